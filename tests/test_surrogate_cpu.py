"""CPU checks of the surrogate's FP64 reference (oracle/surrogate_ref.py) and host logic.

The reference back-propagation is pinned by central finite differences of its own loss,
so a dropped term, a wrong sign or a transposed operand in it fails here.
"""
import numpy as np

import surrogate_ref as R
from dna_amd import surrogate as S


def test_bf16_round_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7 + 2 ** -8, -2.5, 3.0e-39,
                  65504.0], np.float32)
    got = R.bf16_round(x)
    # 1+2^-8 is a tie -> even (1.0); 1+3*2^-9 rounds up to 1+2^-7; 1+2^-7+2^-8 tie -> 1+2^-6
    want = np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, None, 65536.0], object)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == want[2] and got[3] == want[3]
    assert got[4] == -2.5 and got[6] == 65536.0
    # a subnormal keeps its leading bits (3e-39 has 6 significant bits left in bfloat16)
    assert abs(got[5] - 3.0e-39) / 3.0e-39 < 2 ** -5


def _tiny(seed=3, n_in=5, H=16, L=2, no=3, B=7):
    rng = np.random.default_rng(seed)
    dims = R.layer_dims(n_in, H, L, no)
    Ws = [rng.normal(0, 0.5, (di, do)) for di, do in dims]
    Ws[0][n_in:, :] = 0.0
    bs = [rng.normal(0, 0.1, do) for _, do in dims]
    flat = R.flatten(Ws, bs)
    X = rng.normal(0, 1, (B, n_in))
    T = rng.normal(0, 1, (B, no))
    return flat, dims, X, T


def test_reference_gradient_matches_finite_differences():
    flat, dims, X, T = _tiny()
    loss, g = R.loss_and_grad(flat, dims, X, T)
    eps = 1e-6
    rng = np.random.default_rng(0)
    pad, n_in = dims[0][0], X.shape[1]
    # every bias and output weight, plus a sample of the hidden weights (not the padded rows)
    idx = []
    off = 0
    for l, (di, do) in enumerate(dims):
        cand = [off + i * do + j for i in range(di if l else n_in) for j in range(do)]
        idx += list(rng.choice(cand, size=min(40, len(cand)), replace=False))
        off += di * do
    idx += list(range(off, flat.size))
    for i in idx:
        p, m = flat.copy(), flat.copy()
        p[i] += eps
        m[i] -= eps
        fd = (R.loss_and_grad(p, dims, X, T)[0] - R.loss_and_grad(m, dims, X, T)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_reference_gradient_zero_on_padded_inputs():
    flat, dims, X, T = _tiny()
    _, g = R.loss_and_grad(flat, dims, X, T)
    W0 = g[:dims[0][0] * dims[0][1]].reshape(dims[0])
    assert np.all(W0[X.shape[1]:, :] == 0.0)


def test_reference_nll_closed_form():
    # one output: mu = 1, s = log 4, t = 3 -> 0.5 (log 4 + 4/4) ; d/dmu = (mu - t)/4 ; d/ds = 0.5 (1 - 1)
    Y = np.array([[1.0, np.log(4.0)]])
    T = np.array([[3.0]])
    loss, dY = R.nll(Y, T)
    assert np.isclose(loss, 0.5 * (np.log(4.0) + 1.0))
    assert np.isclose(dY[0, 0], -0.5) and np.isclose(dY[0, 1], 0.0)
    # outside the clamp the log-variance gradient is zero
    _, dY = R.nll(np.array([[0.0, 20.0]]), np.array([[0.0]]))
    assert dY[0, 1] == 0.0


def test_checkpoint_grid_16_log_spaced():
    t = S.checkpoints_16()
    assert t.size == 16 and np.isclose(t[0], 1e-3) and np.isclose(t[-1], 1e3)
    assert np.allclose(np.diff(np.log10(t)), 6.0 / 15.0)


def test_norm_stats_and_targets():
    rng = np.random.default_rng(1)
    tally = rng.integers(0, 500, (100, 16, 5)).astype(np.uint32)
    E = rng.uniform(1000, 20000, 100).astype(np.float32)
    G = S.per_track_G(tally, E)
    assert G.shape == (100, 80)
    assert np.allclose(G[3].reshape(16, 5), tally[3] / (E[3] / 100.0), rtol=1e-6)
    X = rng.normal(5, 2, (100, 70)).astype(np.float32)
    X[:, 7] = 1.0    # constant column -> std floored
    xm, xs, ym, ys = S.norm_stats(X, G)
    assert np.all(xs > 0) and np.all(ys > 0)
    assert np.isclose(xm[0], X[:, 0].mean(), rtol=1e-5)


def test_aggregate_G_and_error():
    rng = np.random.default_rng(2)
    tally = rng.integers(1, 50, (40, 16, 5)).astype(np.float64)
    E = rng.uniform(1000, 20000, 40)
    G = tally.reshape(40, 80) / (E[:, None] / 100.0)
    agg = S.aggregate_G(G, E)
    assert np.allclose(agg, tally.sum(0).reshape(80) / (E.sum() / 100.0))
    err = S.g_error(agg * 1.02, agg, floor=0.0)
    assert np.isclose(err["max_rel"], 0.02)
