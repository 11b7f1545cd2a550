"""Learned chemistry surrogate: data generation from this library's own transport + IRT,
normalisation, training on the GPU, and surrogate chemistry with IRT fallback.

No counterpart in the reference.  The network (libdna_amd.so, dna_surrogate_*) maps the
70 per-track features of dna_track_features to the track's G(t) for 5 species (OH, e-aq,
H, H2O2, H2) at 16 log-spaced times from 1 ps to 1 us (80 means + 80 log-variances).
A track whose predicted sigma exceeds a threshold is handed to IRT, which stays the
fallback and the reference against which the surrogate's error is measured.

Per-track G uses the reference's normalisation with one thermalised primary:
G = count / (E / 100) (irt-worker.js:670-698 with n_therm = 1); a set of tracks has
G = sum of counts / (sum of E / 100).
"""
import numpy as np

from .geometry import default_params
from .lib import Context, Surrogate
from .abi import DNA_COPY_RESULTS

N_FEAT = 70
N_T = 16
N_SPECIES = 5
N_OUT = N_T * N_SPECIES
E10_IRT_SEED = 0x43484D01     # experiments/lib/seeds.mjs:26 (E10_IRT_G_VALUES)


def checkpoints_16():
    """16 log-spaced times 10^(-3 + 6k/15) ns, k = 0..15 (1 ps .. 1 us)."""
    return 10.0 ** (-3.0 + 6.0 * np.arange(N_T) / 15.0)


def log_uniform_energies(n, seed, e_lo=1000.0, e_hi=20000.0):
    """E = e_lo (e_hi/e_lo)^u, u uniform from a PCG64 stream keyed by seed."""
    u = np.random.Generator(np.random.PCG64(seed)).random(n)
    return (e_lo * (e_hi / e_lo) ** u).astype(np.float32)


def per_track_G(tally, E):
    """[n, 16, 5] species counts -> [n, 80] G values (count per 100 eV of the track)."""
    n = tally.shape[0]
    return (tally.reshape(n, -1).astype(np.float64) / (np.asarray(E, np.float64)[:, None] / 100.0)).astype(np.float32)


def aggregate_G(G, E):
    """G of a set of tracks from per-track G: sum(G_i E_i) / sum(E_i) per output."""
    E = np.asarray(E, np.float64)
    return (np.asarray(G, np.float64) * E[:, None]).sum(0) / E.sum()


def g_error(pred, ref, floor=0.05):
    """Relative error of aggregate G over the outputs whose reference G >= floor."""
    pred, ref = np.asarray(pred, np.float64), np.asarray(ref, np.float64)
    m = ref >= floor
    rel = np.abs(pred[m] - ref[m]) / ref[m]
    return {"max_rel": float(rel.max()) if rel.size else 0.0,
            "mean_rel": float(rel.mean()) if rel.size else 0.0, "n_outputs": int(m.sum())}


def sigma_calibration(mu, sigma, G_ref, floor=0.05):
    """Per-track calibration of the predicted sigma against reference G (IRT): over every
    (track, output) whose mean reference G is at least floor, the share of |mu - G| within
    1 and 2 sigma (68.3% and 95.4% for a calibrated Gaussian), the per-track RMSE and mean
    absolute error of mu, and the mean predicted sigma."""
    mu, sigma, G_ref = (np.asarray(a, np.float64) for a in (mu, sigma, G_ref))
    m = G_ref.mean(0) >= floor
    d = np.abs(mu[:, m] - G_ref[:, m])
    s = sigma[:, m]
    return {"within_1sigma": round(float((d <= s).mean()), 4), "within_2sigma": round(float((d <= 2 * s).mean()), 4),
            "rmse": round(float(np.sqrt((d ** 2).mean())), 5), "mae": round(float(d.mean()), 5),
            "mean_sigma": round(float(s.mean()), 5), "outputs": int(m.sum()), "tracks": int(mu.shape[0])}


def norm_stats(X, G):
    """Per-column mean and std of features and targets (std floored so that constant
    columns normalise to zero rather than dividing by zero)."""
    X = np.asarray(X, np.float64)
    G = np.asarray(G, np.float64)
    xm, xs = X.mean(0), X.std(0)
    ym, ys = G.mean(0), G.std(0)
    xs = np.maximum(xs, 1e-3 * np.maximum(np.abs(xm), 1.0))
    ys = np.maximum(ys, 1e-3 * np.maximum(np.abs(ym), 1.0))
    return [a.astype(np.float32) for a in (xm, xs, ym, ys)]


def transport_params(n, seed, box_nm=15000.0):
    """Primaries at the origin, isotropic, one energy per primary (from `energies`)."""
    p = default_params(10000.0, n, box_nm=box_nm, seed=seed)
    p.dna_enable = 0
    return p


def generate(ctx: Context, surr: Surrogate, n_tracks, seed=E10_IRT_SEED, chunk=16384,
             e_lo=1000.0, e_hi=20000.0, irt_seed=None, log=None):
    """Tracks of one primary each (E log-uniform in [e_lo, e_hi]): transport + prechemistry,
    the track's features and its IRT G(t) at the 16 times.  Returns X [n,70], G [n,80],
    E [n], tally [n,16,5], and the seconds spent in transport and IRT."""
    import time
    tc = checkpoints_16()
    Xs, Gs, Es, Ts = [], [], [], []
    t_tr = t_irt = 0.0
    done, k = 0, 0
    while done < n_tracks:
        m = min(chunk, n_tracks - done)
        E = log_uniform_energies(m, (seed * 1000003 + k) & 0xFFFFFFFFFFFF, e_lo, e_hi)
        p = transport_params(m, seed=(seed + 7919 * k) & 0xFFFFFFFF)
        t0 = time.perf_counter()
        tr = ctx.transport(p, energies=E, copy=DNA_COPY_RESULTS)
        t1 = time.perf_counter()
        F = surr.features(E)
        o, tally = ctx.irt_tracks(None, tr["rad_n"], m, m, float(E.mean()), tc,
                                  seed=(irt_seed if irt_seed is not None else seed) + k)
        t2 = time.perf_counter()
        t_tr += t1 - t0
        t_irt += t2 - t1
        Xs.append(F)
        Gs.append(per_track_G(tally, E))
        Es.append(E)
        Ts.append(tally)
        done += m
        k += 1
        if log:
            log(f"generated {done}/{n_tracks} tracks ({tr['rad_n']} radicals in the last chunk)")
    return {"X": np.concatenate(Xs), "G": np.concatenate(Gs), "E": np.concatenate(Es),
            "tally": np.concatenate(Ts), "t_transport_s": t_tr, "t_irt_s": t_irt}


def fit(surr: Surrogate, X, G, epochs=(24, 12, 6), lrs=(1e-3, 3e-4, 1e-4), batch=512, seed=1):
    """Normalise with the training statistics, then Adam with a stepped learning rate.
    Returns the mean NLL of the last epoch of each stage."""
    surr.set_norm(*norm_stats(X, G))
    losses = []
    for i, (ep, lr) in enumerate(zip(epochs, lrs)):
        losses.append(surr.train(X, G, batch=batch, epochs=ep, lr=lr, seed=seed + i))
    return losses


def predict_host(surr: Surrogate, X, sigma_max):
    """Network on host features: de-normalised mean and sigma [n,80] and fallback flags
    (the arithmetic of dna_surrogate_predict, for tracks that are not resident)."""
    xm, xs, ym, ys = surr.norm
    out = []
    for b0 in range(0, X.shape[0], 65536):
        out.append(surr.forward(X[b0:b0 + 65536]))
    Y = np.concatenate(out)
    no = surr.n_out
    mu = ym + ys * Y[:, :no]
    sg = ys * np.exp(0.5 * np.clip(Y[:, no:], -12.0, 12.0))
    fb = (sg > sigma_max).any(1)
    return mu, sg, fb


def surrogate_chemistry(ctx: Context, surr: Surrogate, rad_n, E, sigma_max, seed=1):
    """Chemistry of the primaries resident from the last transport: the surrogate for
    confident tracks, IRT (same seed as a full run) for the rest.  Returns the aggregate
    G [80], the per-track predictions, the fallback mask and the device times (ms)."""
    import time
    E = np.asarray(E, np.float32)
    n = E.size
    t0 = time.perf_counter()
    mu, sg, fb, nf, t_pred = surr.predict(E, sigma_max)
    t1 = time.perf_counter()
    tc = checkpoints_16()
    tally_fb = np.zeros((n, N_T, N_SPECIES), np.uint32)
    o = None
    if nf:
        o, tally_fb = ctx.irt_tracks(None, rad_n, n, n, float(E.mean()), tc, seed=seed,
                                     select=fb)
    t2 = time.perf_counter()
    fbm = fb.astype(bool)
    Gt = np.where(fbm[:, None], per_track_G(tally_fb, E), mu)
    return {"G": aggregate_G(Gt, E), "mu": mu, "sigma": sg, "fallback": fbm, "n_fallback": int(nf),
            "t_predict_ms": float(t_pred[2]), "t_features_ms": float(t_pred[0]), "t_network_ms": float(t_pred[1]),
            "t_irt_ms": float(o.t_ms) if o is not None else 0.0,
            "wall_predict_s": t1 - t0, "wall_irt_s": t2 - t1}


def calibrate_sigma(surr: Surrogate, X, frac):
    """Threshold on the per-track maximum sigma under which a fraction 1 - frac of the
    given (training) tracks stays with the surrogate."""
    _, sg, _ = predict_host(surr, X, np.inf)
    return float(np.quantile(sg.max(1), 1.0 - frac))
