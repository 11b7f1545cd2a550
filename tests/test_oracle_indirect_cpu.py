"""The indirect-SSB extension in the CPU restatement, pinned independently of its C code.

* Final OH list (oracle_irt_oh): tracks holding one OH each cannot react, so every OH
  survives and its position at 1 us follows in closed form from the counter-based RNG
  (SplitMix64 finaliser, restated here in Python): the worker's displacement (50% sigma =
  0.8/sqrt(3), clf6 draws 1-18, irt-worker.js:475-537) and then one Brownian step with
  sigma = sqrt(2 D_OH 1000 ns) from draws 32-37 (Box-Muller as in the re-pairing).  The
  list size also equals the IRT's own OH tally at 1 us on a real track.
* Indirect scoring (oracle_score_oh): a pure-Python transcription of scoreIndirectSSB
  (ssb-dsb.ts:43-121) on random OH near backbone atoms; identical OH are all scored (no
  duplicate skipping); the LCG is consumed by the indirect pass first and the direct pass
  continues from there (app.ts:176-186); clustering sees the OR of both hit arrays.
"""
import math

import numpy as np
import pytest

import oracle as O
from dna_amd import DNATarget, default_params

M64 = (1 << 64) - 1
G = 0x9E3779B97F4A7C15


def mix64(z):
    z &= M64
    z ^= z >> 30; z = (z * 0xBF58476D1CE4E5B9) & M64
    z ^= z >> 27; z = (z * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def track_key(seed, inst, pid):
    return mix64((mix64((seed + G * (inst + 1)) & M64) + G * (pid + 1)) & M64)


def u01(ps, n):
    return (mix64((ps + G * (n + 1)) & M64) >> 11) * 2.0 ** -53


@pytest.fixture(scope="module")
def dna():
    return DNATarget()


def test_single_oh_tracks_final_positions_closed_form():
    rng = np.random.default_rng(3)
    n = 200
    pos = rng.uniform(-50, 50, (n, 3)).astype(np.float32)
    rad = np.zeros((n, 4), np.float32)
    rad[:, :3] = pos
    rad[:, 3] = np.arange(n) * 8.0                      # pid * 8 + code 0 (OH)
    o, oh = O.irt_oh(rad, n, 1000.0, [1000.0], seed=9, n_threads=4)
    assert oh.shape == (n, 4) and o.n_reacted == 0
    sig_prod = 0.8 / math.sqrt(3.0)
    sig_bm = math.sqrt(2.0 * 2.2 * 1000.0)
    for pid in range(n):
        ps = track_key(9, 0, pid) ^ 0                    # pair state of particle 0 with itself
        dn = 0
        sig = sig_prod if u01(ps, dn) < 0.5 else 0.0
        dn += 1
        p = [float(v) for v in pos[pid]]
        if sig:
            for ax in range(3):
                s = sum(u01(ps, dn + q) for q in range(6)); dn += 6
                p[ax] = p[ax] + sig * ((s - 3) * 1.4142135623730951)
        p = [float(np.float32(v)) for v in p]
        for ax in range(3):
            g = math.sqrt(-2 * math.log(u01(ps, 32 + 2 * ax) + 1e-30)) * math.cos(6.2831853 * u01(ps, 33 + 2 * ax))
            p[ax] = p[ax] + g * sig_bm
        assert np.allclose(oh[pid, :3], np.float32(p), rtol=0, atol=2e-5), pid
        assert oh[pid, 3] == 0.0


def test_final_oh_count_is_the_1us_tally():
    p = default_params(3000.0, 48)
    t = O.transport(p)
    n_therm = int((t["results"]["esc"] == 0).sum())
    o, oh = O.irt_oh(t["rad"], n_therm, 3000.0, [0.001, 1000.0], seed=4)
    assert oh.shape[0] == o.tally[1][0] > 0


def atom(dna, fiber, bp, strand, jitter=(0, 0, 0)):
    rb = (dna.rbb0 if strand == 0 else dna.rbb1).reshape(-1, 2)[bp]
    return [dna.x0 + bp * dna.rise + jitter[0], float(dna.fy[fiber]) + float(rb[0]) + jitter[1],
            float(dna.fz[fiber]) + float(rb[1]) + jitter[2], 0.0]


def _py_indirect(dna, oh, s):
    """Transcription of scoreIndirectSSB for alive-OH lists (independent of score.c)."""
    hits = set()
    x_half = (dna.n_bp_per - 1) * dna.rise * 0.5
    grid_off = -((dna.grid_N - 1) * dna.spacing_nm) * 0.5
    cand = reach = s0 = s1 = 0
    for row in oh:
        x, y, z = (float(v) for v in row[:3])
        if x < -x_half - 0.29 or x > x_half + 0.29:
            continue
        fi = math.floor((y - grid_off) * (1 / dna.spacing_nm) + 0.5)
        fj = math.floor((z - grid_off) * (1 / dna.spacing_nm) + 0.5)
        if fi < 0 or fi >= dna.grid_N or fj < 0 or fj >= dna.grid_N:
            continue
        fib = fi * dna.grid_N + fj
        yr, zr = y - float(dna.fy[fib]), z - float(dna.fz[fib])
        if yr * yr + zr * zr > (dna.r_bb + 0.29) ** 2:
            continue
        cand += 1
        bp = math.floor((x + x_half) * (1 / dna.rise) + 0.5)
        best, bb, bs = math.inf, -1, -1
        for b in range(max(0, bp - 2), min(dna.n_bp_per - 1, bp + 2) + 1):
            dx = x - (dna.x0 + b * dna.rise)
            for st, arr in ((0, dna.rbb0), (1, dna.rbb1)):
                d2 = dx * dx + (yr - float(arr[2 * b])) ** 2 + (zr - float(arr[2 * b + 1])) ** 2
                if d2 < best:
                    best, bb, bs = d2, b, st
        if best < 0.29 * 0.29:
            reach += 1
            s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
            if s / 4294967296.0 < 0.4:
                idx = fib * dna.n_bp_per + bb + bs * dna.n_bp
                if idx not in hits:
                    hits.add(idx)
                    if bs == 0:
                        s0 += 1
                    else:
                        s1 += 1
    return cand, reach, s0, s1, s


def test_indirect_against_transcription(dna):
    rng = np.random.default_rng(8)
    pts = []
    for _ in range(3000):
        f, b, st = int(rng.integers(0, 441)), int(rng.integers(0, dna.n_bp_per)), int(rng.integers(0, 2))
        pts.append(atom(dna, f, b, st, rng.normal(0, 0.2, 3)))
        if rng.random() < 0.2:
            pts.append(list(pts[-1]))                     # identical OH: both scored
    oh = np.array(pts, np.float32)
    o = O.score_oh(dna, np.zeros((0, 4), np.float32), oh, lcg_state=0x12345678)
    c, r, a, b, s = _py_indirect(dna, oh, 0x12345678)
    ind = o["indirect"]
    assert (ind["candidates"], ind["in_reach"], ind["ssb0"], ind["ssb1"], o["lcg_state"]) == (c, r, a, b, s)


def test_indirect_first_then_direct_and_combined_clusters(dna):
    rng = np.random.default_rng(9)
    oh = np.array([atom(dna, 7, 100 + 3 * k, k % 2, rng.normal(0, 0.05, 3)) for k in range(40)], np.float32)
    rad = np.array([atom(dna, 7, 130 + 3 * k, (k + 1) % 2, rng.normal(0, 0.05, 3)) for k in range(40)], np.float32)
    o = O.score_oh(dna, rad, oh, lcg_state=5, want_hits=True)
    c, r, a, b, s_after = _py_indirect(dna, oh, 5)
    d = O.score(dna, rad, lcg_state=s_after, want_hits=True)
    assert o["indirect"]["in_reach"] == r and o["ssb_count"] == d["ssb_count"]
    assert o["lcg_state"] == d["lcg_state"]
    ind_only = O.score_oh(dna, np.zeros((0, 4), np.float32), oh, lcg_state=5, want_hits=True)["hits"]
    combined = ind_only | d["hits"]
    assert np.array_equal(o["hits"], combined)
    from dna_amd.abi import DnaScoreOut
    import ctypes as C
    ref = DnaScoreOut()
    O.lib().oracle_cluster(C.byref(dna.to_c()), O._ptr(combined), C.byref(ref))
    assert (o["dsb"], o["ssb0"], o["ssb1"]) == (ref.dsb, ref.ssb0, ref.ssb1)
