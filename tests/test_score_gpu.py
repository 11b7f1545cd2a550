"""SSB/DSB scoring and spatial hash on the GPU against oracle/score.c (bit-exact)."""
import numpy as np
import pytest

import oracle as O
from dna_amd import DNATarget, default_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dna():
    return DNATarget()


def _synthetic_hits(dna, n, rng):
    """Entries placed on backbone atoms (in reach), with duplicate runs of length 1..4."""
    fib = rng.integers(0, dna.n_fibers, n)
    bp = rng.integers(0, dna.n_bp_per, n)
    strand = rng.integers(0, 2, n)
    rb = np.where(strand[:, None] == 0, dna.rbb0.reshape(-1, 2)[bp], dna.rbb1.reshape(-1, 2)[bp])
    jit = rng.normal(0, 0.12, (n, 3))
    x = dna.x0 + bp * dna.rise + jit[:, 0]
    y = dna.fy[fib] + rb[:, 0] + jit[:, 1]
    z = dna.fz[fib] + rb[:, 1] + jit[:, 2]
    base = np.stack([x, y, z, np.zeros(n)], 1).astype(np.float32)
    reps = rng.integers(1, 5, n)
    return np.repeat(base, reps, axis=0)


def test_score_synthetic_exact(ctx, dna):
    rng = np.random.default_rng(1)
    rad = _synthetic_hits(dna, 20000, rng)
    print("score synthetic", rad.shape, flush=True)
    g = ctx.score(dna, rad, lcg_state=0x12345678, want_hits=True)
    o = O.score(dna, rad, lcg_state=0x12345678, want_hits=True)
    for k in ("ssb_count", "candidates", "in_reach", "dsb", "ssb0", "ssb1", "lcg_state"):
        assert g[k] == o[k], (k, g[k], o[k])
    assert g["dsb_expected"] == o["dsb_expected"]
    assert np.array_equal(g["hits"], o["hits"])
    assert g["dsb"] > 0 and g["in_reach"] > 1000


def test_score_transport_output(ctx, dna):
    p = default_params(10000.0, 64, dna=dna)
    t = ctx.transport(p)
    print("score resident", flush=True)
    g = ctx.score(dna, None, t["rad_n"], lcg_state=0x12345678)
    o = O.score(dna, t["rad"], lcg_state=0x12345678)
    for k in ("ssb_count", "candidates", "in_reach", "dsb", "ssb0", "ssb1", "lcg_state"):
        assert g[k] == o[k], (k, g[k], o[k])


def test_score_empty(ctx, dna):
    g = ctx.score(dna, np.zeros((0, 4), np.float32), lcg_state=7)
    assert g["ssb_count"] == 0 and g["dsb"] == 0 and g["lcg_state"] == 7


def test_hash_cells_exact(ctx):
    rng = np.random.default_rng(3)
    pos = np.concatenate([rng.normal(0, 50, (100000, 3)), rng.uniform(-3000, 3000, (20000, 3))]).astype(np.float32)
    pos4 = np.concatenate([pos, np.zeros((pos.shape[0], 1), np.float32)], 1)
    for bits, cell in ((23, 1.5), (12, 4.0)):
        print("hash", bits, cell, flush=True)
        kg, sg, og = ctx.hash_cells(pos4, cell, bits)
        ko, so, oo = O.hash_cells(pos4, cell, bits)
        assert np.array_equal(kg, ko) and np.array_equal(sg, so) and np.array_equal(og, oo)


def test_hash_cells_empty_and_bad_arguments(ctx):
    """No points: every bucket is empty (cell_start all zero), as the CPU restatement
    gives; out-of-range arguments are refused before any GPU work."""
    empty = np.zeros((0, 4), np.float32)
    kg, sg, og = ctx.hash_cells(empty, 1.5, 10)
    ko, so, oo = O.hash_cells(empty, 1.5, 10)
    assert kg.size == og.size == 0 and np.array_equal(sg, so) and not sg.any()
    with pytest.raises(Exception, match="bits"):
        ctx.hash_cells(empty, 1.5, 0)
    with pytest.raises(Exception, match="cell_nm"):
        ctx.hash_cells(empty, 0.0, 10)
