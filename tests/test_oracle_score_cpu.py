"""Scoring restatement (oracle/score.c) against hand-built cases and an independent
pure-Python transcription of src/scoring/ssb-dsb.ts on small inputs; spatial hash
against a numpy evaluation of chemistry.wgsl:76-80."""
import math

import numpy as np
import pytest

import oracle as O
from dna_amd import DNATarget


@pytest.fixture(scope="module")
def dna():
    return DNATarget()


def lcg_draws(s, n):
    out = []
    for _ in range(n):
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        out.append(s / 4294967296.0)
    return out, s


def atom(dna, fiber, bp, strand, jitter=(0, 0, 0)):
    rb = (dna.rbb0 if strand == 0 else dna.rbb1).reshape(-1, 2)[bp]
    return [dna.x0 + bp * dna.rise + jitter[0], float(dna.fy[fiber]) + float(rb[0]) + jitter[1],
            float(dna.fz[fiber]) + float(rb[1]) + jitter[2], 0.0]


def test_geometry_known_answers(dna):
    """buildDNATarget defaults: 441 fibers, floor(3000/0.34) bp per fiber, x0 centring
    (dna-geometry.test.ts:4-39)."""
    assert dna.n_fibers == 441 and dna.n_bp_per == 8823 and dna.n_bp == 441 * 8823
    assert abs(dna.x0 + 8822 * 0.34 / 2) < 1e-12
    assert np.allclose(np.hypot(dna.rbb0[0::2], dna.rbb0[1::2]), 1.0, atol=1e-6)
    assert np.allclose(dna.rbb0, -dna.rbb1, atol=1e-6)


def test_single_in_reach_event_consumes_one_draw(dna):
    s0 = 0x12345678
    rad = np.array([atom(dna, 220, 1000, 0)], np.float32)
    o = O.score(dna, rad, lcg_state=s0, want_hits=True)
    (u,), s1 = lcg_draws(s0, 1)
    assert o["candidates"] == 1 and o["in_reach"] == 1 and o["lcg_state"] == s1
    assert o["ssb_count"] == (1 if u < 0.15 else 0)
    if u < 0.15:
        assert o["hits"][220 * dna.n_bp_per + 1000] == 1


@pytest.mark.parametrize("L", [1, 2, 3, 4, 5])
def test_dedupe_runs(dna, L):
    """ssb-dsb.ts:154-168: a run of L identical entries is visited ceil(L/2) times."""
    rad = np.array([atom(dna, 3, 50, 1)] * L + [atom(dna, 3, 80, 0)], np.float32)
    o = O.score(dna, rad, lcg_state=1)
    assert o["in_reach"] == (L + 1) // 2 + 1


def test_fiber_snap_rounds_half_up(dna):
    """Math.round(x.5) rounds up: a point exactly midway between fibers 9 and 10 in y is
    assigned to fiber 10 and then fails the reach test, an out-of-grid point is skipped."""
    mid_y = float(dna.fy[9 * 21]) + 75.0
    rad = np.array([[0.0, mid_y, float(dna.fz[0]), 0.0],
                    [0.0, float(dna.fy[-1]) + 80.0, 0.0, 0.0]], np.float32)
    o = O.score(dna, rad)
    assert o["candidates"] == 0 and o["in_reach"] == 0


def test_fiber_end_window(dna):
    """bp windows are clipped at 0 and n_bp_per-1 (ssb-dsb.ts:185-186)."""
    rad = np.array([atom(dna, 0, 0, 0), atom(dna, 440, dna.n_bp_per - 1, 1)], np.float32)
    o = O.score(dna, rad, lcg_state=5)
    assert o["in_reach"] == 2


def test_cluster_greedy_known_answer(dna):
    hits = np.zeros(2 * dna.n_bp, np.uint8)
    n = dna.n_bp
    f = 7 * dna.n_bp_per
    for b in (100, 105, 300):
        hits[f + b] = 1
    for b in (95, 108, 200, 311):
        hits[n + f + b] = 1
    o = O.cluster(dna, hits)
    # 100 pairs 95, 105 pairs 108, 300 pairs 311 (|d| = 11 > 10 -> no): dsb = 2
    assert o["dsb"] == 2 and o["ssb0"] == 3 and o["ssb1"] == 4
    p = 21 / dna.n_bp_per
    assert abs(o["dsb_expected"] - (1 - (1 - p) ** 12)) < 1e-15


def _py_direct(dna, rad, s):
    """Transcription of scoreDirectSSB_events for small inputs (independent of score.c)."""
    hits = {}
    x_half = (dna.n_bp_per - 1) * dna.rise * 0.5
    grid_off = -((dna.grid_N - 1) * dna.spacing_nm) * 0.5
    cand = reach = ssb = 0
    i = 0
    n = len(rad)
    while i < n:
        x, y, z = (float(v) for v in rad[i, :3])
        skip = 2 if (i + 1 < n and all(rad[i + 1, k] == rad[i, k] for k in range(3))) else 1
        i += skip
        if x < -x_half - 0.29 or x > x_half + 0.29:
            continue
        fi = math.floor((y - grid_off) * (1 / dna.spacing_nm) + 0.5)
        fj = math.floor((z - grid_off) * (1 / dna.spacing_nm) + 0.5)
        if fi < 0 or fi >= dna.grid_N or fj < 0 or fj >= dna.grid_N:
            continue
        fib = fi * dna.grid_N + fj
        yr, zr = y - float(dna.fy[fib]), z - float(dna.fz[fib])
        if yr * yr + zr * zr > 1.29 ** 2:
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
            if s / 4294967296.0 < 0.15:
                idx = fib * dna.n_bp_per + bb + bs * dna.n_bp
                if idx not in hits:
                    hits[idx] = 1
                    ssb += 1
    return cand, reach, ssb, s


def test_direct_scoring_against_transcription(dna):
    rng = np.random.default_rng(4)
    pts = []
    for _ in range(3000):
        f, b, st = int(rng.integers(0, 441)), int(rng.integers(0, dna.n_bp_per)), int(rng.integers(0, 2))
        pts.append(atom(dna, f, b, st, rng.normal(0, 0.2, 3)))
        if rng.random() < 0.3:
            pts.append(list(pts[-1]))
    rad = np.array(pts, np.float32)
    o = O.score(dna, rad, lcg_state=0x12345678)
    c, r, s, st = _py_direct(dna, rad, 0x12345678)
    assert (o["candidates"], o["in_reach"], o["ssb_count"], o["lcg_state"]) == (c, r, s, st)


def test_hash_against_numpy():
    rng = np.random.default_rng(2)
    pos = rng.normal(0, 30, (50000, 3)).astype(np.float32)
    pos4 = np.concatenate([pos, np.zeros((50000, 1), np.float32)], 1)
    key, start, order = O.hash_cells(pos4, 1.5, 20)
    cell = np.floor(pos / np.float32(1.5)).astype(np.int32)
    with np.errstate(over="ignore"):
        h = ((cell[:, 0].astype(np.uint32) * np.uint32(73856093)) ^
             (cell[:, 1].astype(np.uint32) * np.uint32(19349663)) ^
             (cell[:, 2].astype(np.uint32) * np.uint32(83492791))) & np.uint32((1 << 20) - 1)
    assert np.array_equal(key, h)
    assert np.array_equal(order, np.argsort(h, kind="stable"))
    assert np.array_equal(start, np.concatenate([[0], np.cumsum(np.bincount(h, minlength=1 << 20))]))
