"""Phase A+B on the GPU against the CPU restatement (oracle/transport.c).

Both run the same FP32 expressions with the same elementary functions, so every
per-primary result, every tally and the canonical radical list must agree bit for bit."""
import numpy as np
import pytest

import oracle as O
from dna_amd import DNATarget, default_params

pytestmark = pytest.mark.gpu


def _cmp(g, o, n):
    assert np.array_equal(g["results"].view(np.uint32).reshape(n, 8), o["results"].view(np.uint32).reshape(n, 8))
    assert np.array_equal(g["counters"], o["counters"])
    assert np.array_equal(g["sec_stats"], o["sec_stats"])
    assert np.array_equal(g["dbg"][:5], o["dbg"][:5])
    assert g["dose_sum"] == o["dose_sum"]
    assert np.array_equal(g["dose"], o["dose"])
    assert g["rad_n"] == o["rad_n"]
    assert np.array_equal(g["rad"].view(np.uint32), o["rad"].view(np.uint32))
    assert g["sec_n"] == o["sec_n"]


@pytest.mark.parametrize("E,n", [(1000.0, 256), (100.0, 128), (10000.0, 32)])
def test_transport_bit_exact(ctx, E, n):
    dna = DNATarget() if E == 10000.0 else None
    p = default_params(E, n, dna=dna)
    print(f"transport E={E} n={n}", flush=True)
    g = ctx.transport(p)
    o = O.transport(p)
    _cmp(g, o, n)


def test_transport_start_cube_and_energies(ctx):
    rng = np.random.default_rng(5)
    n = 96
    E = (100.0 * 300.0 ** rng.random(n)).astype(np.float32)
    p = default_params(5000.0, n, start_half=1000.0, seed=0x54524B01)
    print("transport per-primary energies", flush=True)
    g = ctx.transport(p, energies=E)
    o = O.transport(p, energies=E)
    _cmp(g, o, n)


@pytest.mark.parametrize("lanes", [1, 4, 16, 64])
def test_transport_lanes_per_wavefront(ctx, lanes):
    """The primary kernel's launch shape (active lanes per wavefront, DNA_OPT_PA_LANES)
    leaves every result unchanged: RNG streams are keyed by primary and the radical list is
    put in canonical order."""
    from dna_amd.abi import DNA_OPT_PA_LANES
    p = default_params(3000.0, 200)
    print(f"transport 3 keV x 200 with {lanes} lanes per wavefront", flush=True)
    ctx.set_option(DNA_OPT_PA_LANES, lanes)
    try:
        g = ctx.transport(p)
    finally:
        ctx.set_option(DNA_OPT_PA_LANES, 0)
    _cmp(g, O.transport(p), 200)
    with pytest.raises(Exception, match="DNA_OPT_PA_LANES"):
        ctx.set_option(DNA_OPT_PA_LANES, 3)


@pytest.mark.parametrize("rad_chunk,work_chunk", [(256, 64), (512, 128), (1024, 256), (4096, 1024)])
def test_transport_chunks(ctx, rad_chunk, work_chunk):
    """The radical-slot and secondary-claim chunk sizes (DNA_OPT_RAD_CHUNK, DNA_OPT_WORK_CHUNK)
    change only which wavefront writes which raw slot: the canonical outputs are the
    oracle's."""
    from dna_amd.abi import DNA_OPT_RAD_CHUNK, DNA_OPT_WORK_CHUNK
    p = default_params(3000.0, 200)
    print(f"transport 3 keV x 200 with radical chunk {rad_chunk}, work chunk {work_chunk}", flush=True)
    ctx.set_option(DNA_OPT_RAD_CHUNK, rad_chunk)
    ctx.set_option(DNA_OPT_WORK_CHUNK, work_chunk)
    try:
        g = ctx.transport(p)
    finally:
        ctx.set_option(DNA_OPT_RAD_CHUNK, 0)
        ctx.set_option(DNA_OPT_WORK_CHUNK, 0)
    _cmp(g, O.transport(p), 200)
    for opt, bad in ((DNA_OPT_RAD_CHUNK, 128), (DNA_OPT_RAD_CHUNK, 300), (DNA_OPT_WORK_CHUNK, 32)):
        with pytest.raises(Exception, match="CHUNK"):
            ctx.set_option(opt, bad)


def test_transport_autotune(ctx):
    """The first batch of a class (here 1024 x 2 keV) times the knob candidates and keeps
    the fastest; the run's outputs are the oracle's, and a second run of the class reuses
    the choice."""
    from dna_amd.abi import (DNA_OPT_PA_LANES, DNA_OPT_PA_LANES_LAST, DNA_OPT_RAD_CHUNK,
                             DNA_OPT_RAD_CHUNK_LAST, DNA_OPT_WORK_CHUNK, DNA_OPT_WORK_CHUNK_LAST)
    for o in (DNA_OPT_PA_LANES, DNA_OPT_RAD_CHUNK, DNA_OPT_WORK_CHUNK):
        ctx.set_option(o, 0)
    p = default_params(2000.0, 1024)
    print("transport autotune on 1024 x 2 keV", flush=True)
    g = ctx.transport(p)
    knobs = tuple(ctx.get_option(o) for o in (DNA_OPT_PA_LANES_LAST, DNA_OPT_RAD_CHUNK_LAST, DNA_OPT_WORK_CHUNK_LAST))
    assert knobs[0] in (1, 2, 4, 8, 16, 32, 64) and knobs[1] in (256, 512, 1024) and knobs[2] in (64, 128, 256), knobs
    _cmp(g, O.transport(p, n_threads=8), 1024)
    ctx.transport(p)
    assert tuple(ctx.get_option(o) for o in (DNA_OPT_PA_LANES_LAST, DNA_OPT_RAD_CHUNK_LAST,
                                              DNA_OPT_WORK_CHUNK_LAST)) == knobs


def test_transport_caps(ctx):
    """max_sec / max_rad as caps with drop counts (the reference's MAX_SEC / MAX_RAD,
    dispatch.ts:124-126,190-192): the kept secondaries and radicals are the first ones in
    canonical order, as in the oracle, and an uncapped run afterwards is unaffected."""
    n = 160
    p = default_params(2000.0, n)
    full = O.transport(p)
    p.max_sec = full["sec_n"] // 2
    p.max_rad = full["rad_n"] // 3
    print(f"transport 2 keV x {n} with max_sec={p.max_sec} max_rad={p.max_rad}", flush=True)
    g = ctx.transport(p)
    o = O.transport(p)
    _cmp(g, o, n)
    assert g["sec_n_raw"] == o["sec_n_raw"] == full["sec_n"]
    assert g["sec_dropped"] == o["sec_n_raw"] - o["sec_n"] > 0
    assert g["rad_n_raw"] == o["rad_n_raw"]
    assert g["rad_n"] == p.max_rad and g["rad_dropped"] == o["rad_n_raw"] - p.max_rad > 0
    q = default_params(2000.0, n)
    g2 = ctx.transport(q)
    _cmp(g2, full, n)
    assert g2["sec_dropped"] == 0 and g2["rad_dropped"] == 0
