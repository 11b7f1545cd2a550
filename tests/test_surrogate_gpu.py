"""Surrogate kernels on the GPU: the BF16 MFMA forward pass and back-propagation against
the FP64 reference (oracle/surrogate_ref.py), per-track IRT and features, and an
end-to-end check that the trained network replaces IRT on held-out tracks within 3% in
G at a >= 5x lower chemistry cost.

Tolerances (relative L2 norm over a block, |a - b| / |b|):
  * against the reference evaluated with BF16 rounding at the GPU's storage points:
    forward 5e-3, every gradient block 2e-2 (what is left is FP32 accumulation order and
    the BF16 roundings it occasionally flips);
  * against the plain FP64 reference (no BF16 rounding anywhere): every gradient block
    0.15 and cosine >= 0.99.  This gap is not kernel error: BF16 rounding of the forward
    activations flips the ReLU state of units whose pre-activation is within ~0.2% of
    zero, and one flipped sample changes a B-sample gradient sum by about 1/sqrt(B)
    (measured on the CPU with the reference alone: 0.4% at the output layer, 8% at the
    first layer for B = 200).
"""
import numpy as np
import pytest

import surrogate_ref as R
from dna_amd import Surrogate, default_params
from dna_amd import surrogate as S

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def net(ctx):
    s = Surrogate(ctx, 70, 256, 3, 80, seed=11)
    yield s
    s.close()


def test_forward_matches_reference(net):
    rng = np.random.default_rng(5)
    B = 300                                   # partial 64-row tile
    X = rng.normal(0, 1, (B, 70)).astype(np.float32)
    print("surrogate forward B=300", flush=True)
    Y = net.forward(X)
    dims = R.layer_dims(70, 256, 3, 80)
    Yr, _ = R.forward(net.get_params(), dims, X, bf16=True)
    assert Y.shape == (B, 160)
    assert _rel(Y, Yr) < 5e-3, _rel(Y, Yr)
    Yf, _ = R.forward(net.get_params(), dims, X, bf16=False)
    assert _rel(Y, Yf) < 3e-2


def test_gradient_matches_fp64_reference(net):
    rng = np.random.default_rng(6)
    B = 200
    X = rng.normal(0, 1, (B, 70)).astype(np.float32)
    T = rng.normal(0, 1, (B, 80)).astype(np.float32)
    print("surrogate gradient B=200", flush=True)
    loss, g = net.grad(X, T)
    dims = R.layer_dims(70, 256, 3, 80)
    p = net.get_params().astype(np.float64)
    lb, gb = R.loss_and_grad(p, dims, X, T, bf16=True)
    lf, gf = R.loss_and_grad(p, dims, X, T, bf16=False)
    assert abs(loss - lb) <= 2e-3 * abs(lb)
    Wg, bg = R.unflatten(g, dims)
    Wb, bb = R.unflatten(gb, dims)
    Wf, bf = R.unflatten(gf, dims)
    for l in range(len(dims)):
        for a, b, c in ((Wg[l], Wb[l], Wf[l]), (bg[l], bb[l], bf[l])):
            assert _rel(a, b) < 2e-2, (l, _rel(a, b))
            assert _rel(a, c) < 0.15, (l, _rel(a, c))
            cos = float(np.dot(a.ravel(), c.ravel()) / (np.linalg.norm(a) * np.linalg.norm(c)))
            assert cos >= 0.99, (l, cos)
    # the output layer sees no ReLU flips: it must agree with plain FP64 much more closely
    assert _rel(Wg[-1], Wf[-1]) < 1e-2 and _rel(bg[-1], bf[-1]) < 1e-2
    # padded input rows get no gradient
    assert np.all(Wg[0][70:, :] == 0.0)


def test_fused_forward_equals_per_layer_path(ctx, net):
    """The fused forward kernel (DNA_OPT_MLP_FUSED = 1) runs k_gemm's MFMA sequence per
    output tile: raw outputs and gradients equal the one-GEMM-per-layer path bit for bit,
    on a batch that is not a multiple of the 64-row tile."""
    from dna_amd.abi import DNA_OPT_MLP_FUSED
    rng = np.random.default_rng(15)
    X = rng.normal(0, 1, (1000, 70)).astype(np.float32)
    T = rng.normal(0, 1, (1000, 80)).astype(np.float32)
    out = {}
    for fused in (0, 1):
        ctx.set_option(DNA_OPT_MLP_FUSED, fused)
        try:
            print(f"surrogate forward and gradient, fused={fused}", flush=True)
            out[fused] = (net.forward(X),) + net.grad(X, T)
        finally:
            ctx.set_option(DNA_OPT_MLP_FUSED, 1)
    assert np.array_equal(out[0][0], out[1][0])
    # the loss is a sum of FP64 atomics (any order): equal to rounding
    assert abs(out[0][1] - out[1][1]) <= 1e-12 * abs(out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


def test_training_fits_smooth_function(ctx):
    rng = np.random.default_rng(7)
    N = 8192
    X = rng.normal(0, 1, (N, 70)).astype(np.float32)
    a = rng.normal(0, 0.12, (70, 80))
    T = (np.tanh(X @ a) + 0.05 * rng.normal(0, 1, (N, 80))).astype(np.float32)
    s = Surrogate(ctx, 70, 256, 3, 80, seed=2)
    try:
        print("surrogate training on a synthetic function", flush=True)
        l0, _ = s.grad(X[:512], T[:512])
        l1 = s.train(X, T, batch=256, epochs=30, lr=1e-3)
        l2, _ = s.grad(X[:512], T[:512])
        # NLL drops by > 1: the fitted residual variance is e^2 ~ 7x below the target's
        assert l1 < l0 - 1.0 and l2 < l0 - 1.0, (l0, l1, l2)
    finally:
        s.close()


@pytest.fixture(scope="module")
def mixed(ctx):
    E = S.log_uniform_energies(96, 77, 1000.0, 8000.0)
    p = S.transport_params(96, seed=4242)
    print("transport 96 primaries 1-8 keV", flush=True)
    t = ctx.transport(p, energies=E)
    return E, t


def test_irt_tracks_selection_is_exact(ctx, mixed):
    E, t = mixed
    tc = S.checkpoints_16()
    print("irt_tracks full and selected", flush=True)
    o_all, tt_all = ctx.irt_tracks(None, t["rad_n"], 96, 96, 1.0, tc, seed=9)
    o_ref = ctx.irt(None, t["rad_n"], 96, 1.0, tc, seed=9)
    agg = np.array([[o_ref.tally[c][s] for s in range(5)] for c in range(16)], np.uint64)
    assert np.array_equal(tt_all.sum(0, dtype=np.uint64), agg)
    sel = (np.arange(96) % 3 == 1).astype(np.uint8)
    o_sel, tt_sel = ctx.irt_tracks(None, t["rad_n"], 96, 96, 1.0, tc, seed=9, select=sel)
    assert np.array_equal(tt_sel[sel == 1], tt_all[sel == 1])
    assert np.all(tt_sel[sel == 0] == 0)
    agg_sel = np.array([[o_sel.tally[c][s] for s in range(5)] for c in range(16)], np.uint64)
    assert np.array_equal(agg_sel, tt_all[sel == 1].sum(0, dtype=np.uint64))
    assert o_sel.n_tracks == int(sel.sum())


def _features_numpy(rad, E):
    pid = (np.floor(rad[:, 3] + 0.5)).astype(np.int64)
    code, pid = pid % 8, pid // 8
    n = E.size
    F = np.zeros((n, 70))
    lo, hi = np.log(0.5), np.log(20000.0)
    for p in range(n):
        m = pid == p
        c = code[m]
        xyz = rad[m, :3].astype(np.float64)
        part = (c != 4) & (c != 7)
        if part.sum() == 0:
            continue
        cen = xyz[part].mean(0)
        r = np.sqrt(((xyz[part] - cen) ** 2).sum(1))
        b = np.clip(((np.log(np.maximum(r, 0.5)) - lo) * 64 / (hi - lo)).astype(int), 0, 63)
        F[p, :64] = np.bincount(b, minlength=64) / part.sum()
        per100 = 100.0 / E[p]
        F[p, 64] = (c == 0).sum() * per100
        F[p, 65] = ((c == 1) | (c == 5)).sum() * per100
        F[p, 66] = (c == 2).sum() * per100
        F[p, 67] = (c == 3).sum() * per100
        F[p, 68] = (c == 7).sum() * per100
        F[p, 69] = np.log10(E[p] / 1000.0)
    return F


def test_features_match_numpy(net, mixed):
    E, t = mixed
    print("track features resident and host paths", flush=True)
    F = net.features(E)
    rad = t["rad"]
    Fn = _features_numpy(rad, E)
    npart = np.array([((np.floor(rad[:, 3] + 0.5).astype(int) // 8 == p)).sum() for p in range(E.size)])
    assert np.allclose(F[:, 64:], Fn[:, 64:], rtol=1e-5, atol=1e-6)
    # a centroid summed in FP32 may move a particle sitting on a bin edge to the next bin
    moved = np.abs(F[:, :64] - Fn[:, :64]).sum(1) * np.maximum(npart, 1)
    assert np.all(moved <= 4.01), moved.max()
    # the host-buffer path gives the same bits as the resident path
    pid = (np.floor(rad[:, 3] + 0.5).astype(np.int64)) // 8
    off = np.searchsorted(pid, np.arange(E.size + 1)).astype(np.uint64)
    F2 = net.features(E, rad=rad, inst_off=off)
    assert np.array_equal(F, F2)


def test_fused_predict_equals_per_layer_path(ctx, net, mixed):
    """Surrogate chemistry's prediction (mean, sigma, fallback flags and count) from the
    one-launch fused kernel equals the per-layer path with k_predict."""
    from dna_amd.abi import DNA_OPT_MLP_FUSED
    E, t = mixed
    out = {}
    for fused in (0, 1):
        ctx.set_option(DNA_OPT_MLP_FUSED, fused)
        try:
            print(f"surrogate predict, fused={fused}", flush=True)
            mu, sg, fb, nf, _ = net.predict(E, 0.05)
            out[fused] = (mu, sg, fb, nf)
        finally:
            ctx.set_option(DNA_OPT_MLP_FUSED, 1)
    for a, b in zip(out[0][:3], out[1][:3]):
        assert np.array_equal(a, b)
    assert out[0][3] == out[1][3] == int(out[1][2].sum())


def test_surrogate_replaces_irt_on_held_out_tracks(ctx):
    """Train on 48k generated tracks; on a held-out batch of 4096 tracks the surrogate
    chemistry (IRT fallback for the 10% most uncertain tracks) matches plain IRT's
    aggregate G(t) within 3% at every (time, species) with G >= 0.05, costs >= 5x less
    device time, and beats the best input-independent predictor per track.  The uncertain
    tracks are mostly the largest (high-energy) ones, so the fallback costs more than its
    share of tracks: with 20% of them it took 3.2 ms against 16.8 ms of plain IRT."""
    s = Surrogate(ctx, 70, 256, 3, 80, seed=3)
    try:
        print("surrogate: generating 48k training tracks", flush=True)
        d = S.generate(ctx, s, 48 * 1024, seed=S.E10_IRT_SEED, chunk=16384)
        print(f"surrogate: training ({d['t_transport_s']:.2f}s transport, {d['t_irt_s']:.2f}s IRT)", flush=True)
        losses = S.fit(s, d["X"], d["G"], epochs=(16, 8, 4))
        print(f"surrogate: losses {losses}", flush=True)
        sigma_max = S.calibrate_sigma(s, d["X"], 0.1)
        # held-out batch: a different seed stream
        NH = 4096
        E = S.log_uniform_energies(NH, 0xBEEF, 1000.0, 20000.0)
        t = ctx.transport(S.transport_params(NH, seed=0xBEEF), energies=E, copy=0)
        tc = S.checkpoints_16()
        o_irt, tt = ctx.irt_tracks(None, t["rad_n"], NH, NH, 1.0, tc, seed=5)
        G_irt_tr = S.per_track_G(tt, E)
        G_irt = S.aggregate_G(G_irt_tr, E)
        res = S.surrogate_chemistry(ctx, s, t["rad_n"], E, sigma_max, seed=5)
        err = S.g_error(res["G"], G_irt)
        speed = o_irt.t_ms / (res["t_predict_ms"] + res["t_irt_ms"])
        print(f"surrogate: err {err} fallback {res['n_fallback']} irt {o_irt.t_ms:.2f} ms "
              f"surrogate {res['t_predict_ms']:.2f}+{res['t_irt_ms']:.2f} ms speedup {speed:.1f}", flush=True)
        assert err["max_rel"] <= 0.03, err
        assert speed >= 5.0, speed
        # per-track error below that of the training-set mean (the best constant predictor)
        rmse_net = np.sqrt(((res["mu"] - G_irt_tr) ** 2).mean())
        rmse_const = np.sqrt(((d["G"].mean(0) - G_irt_tr) ** 2).mean())
        assert rmse_net < 0.8 * rmse_const, (rmse_net, rmse_const)
        # sigma calibration per track and output (outputs with mean G >= 0.05): the share of
        # |mu - G_IRT| within 1 and 2 sigma, 68.3% and 95.4% for a calibrated Gaussian
        cal = S.sigma_calibration(res["mu"], res["sigma"], G_irt_tr)
        print(f"surrogate: sigma calibration {cal}", flush=True)
        assert 0.55 <= cal["within_1sigma"] <= 0.85 and 0.85 <= cal["within_2sigma"] <= 0.995, cal
    finally:
        s.close()
