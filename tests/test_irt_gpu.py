"""Phase C on the GPU against the CPU restatement of public/irt-worker.js (oracle/irt.c).

Both draw from the same counter-based streams and order events by (t, a, b), so the
integer species tallies at every checkpoint and the per-reaction counts must agree
exactly."""
import numpy as np
import pytest

import oracle as O
from dna_amd import default_params

pytestmark = pytest.mark.gpu
CK = [0.001, 0.01, 0.1, 1.0, 10.0, 100.0, 1000.0]


def _tallies(o):
    return (np.array([[o.tally[c][s] for s in range(5)] for c in range(o.n_ckpt)], np.uint64),
            np.array(o.rxn_counts[:], np.uint64), o.n_reacted, o.n_events, o.n_tracks)


def _same(g, o):
    tg, rg, ng, eg, kg = _tallies(g)
    to, ro, no, eo, ko = _tallies(o)
    assert kg == ko
    assert eg == eo, (eg, eo)
    assert ng == no, (ng, no)
    assert np.array_equal(rg, ro), (rg, ro)
    assert np.array_equal(tg, to), (tg, to)


@pytest.fixture(scope="module")
def rad10k(ctx):
    p = default_params(10000.0, 48)
    print("transport 48 x 10 keV for IRT input", flush=True)
    return ctx.transport(p)["rad"]


@pytest.mark.parametrize("seed", [1, 12345])
def test_irt_exact_10kev(ctx, rad10k, seed):
    print(f"irt 10 keV seed={seed}", flush=True)
    g = ctx.irt(rad10k, None, 48, 10000.0, CK, seed=seed)
    o = O.irt(rad10k, 48, 10000.0, CK, seed=seed)
    _same(g, o)


def test_irt_resident_radicals(ctx):
    p = default_params(1000.0, 200)
    t = ctx.transport(p)
    print("irt on resident radicals", flush=True)
    g = ctx.irt(None, t["rad_n"], 200, 1000.0, CK, seed=3)
    o = O.irt(t["rad"], 200, 1000.0, CK, seed=3)
    _same(g, o)


def test_irt_interleaved_input(ctx, rad10k):
    """Entries of different primaries interleaved: grouping must restore input order."""
    rng = np.random.default_rng(0)
    pid = (np.floor(rad10k[:, 3] + 0.5) // 8).astype(int)
    sub = rad10k[pid < 6]
    perm = rng.permutation(sub.shape[0])
    mixed = sub[perm]
    print("irt interleaved", flush=True)
    g = ctx.irt(mixed, None, 6, 10000.0, CK, seed=9)
    o = O.irt(mixed, 6, 10000.0, CK, seed=9)
    _same(g, o)


def test_irt_batch(ctx, rad10k):
    pid = (np.floor(rad10k[:, 3] + 0.5) // 8).astype(int)
    cuts = [0] + [int(np.searchsorted(pid, k)) for k in (5, 11, 30)] + [rad10k.shape[0]]
    off = np.array(cuts, np.uint64)
    nt = np.array([5, 6, 19, 18], np.uint32)
    E = np.full(4, 10000.0, np.float32)
    print("irt batch", flush=True)
    gs = ctx.irt_batch(rad10k, off, nt, E, CK, seed=5)
    os_ = O.irt_batch(rad10k, off, nt, E, CK, seed=5)
    for g, o in zip(gs, os_):
        _same(g, o)


def test_irt_worker_codes_case(ctx):
    """tools/test_worker_codes.cjs synthetic rad_buf: codes 0/1/3/6/7/2 over 4 primaries."""
    def rec(pid, sp, x):
        return [x, 0, 0, pid * 8 + sp]
    rad = np.array([rec(0, 0, 0), rec(0, 1, 2), rec(0, 3, 0), rec(1, 0, 10), rec(1, 6, 10), rec(1, 7, 10),
                    rec(2, 0, 20), rec(2, 0, 20), rec(2, 7, 20), rec(3, 0, 30), rec(3, 2, 30)], np.float32)
    g = ctx.irt(rad, None, 4, 10000.0, CK, seed=2)
    o = O.irt(rad, 4, 10000.0, CK, seed=2)
    _same(g, o)
    # two initial H2 markers counted from 1 ps on (test_worker_codes.cjs:75-85)
    assert g.tally[0][4] >= 2


@pytest.mark.parametrize("block", [64, 128, 256])
def test_irt_workgroup_size_does_not_change_results(ctx, rad10k, block):
    from dna_amd.abi import DNA_OPT_IRT_BLOCK, DNA_OPT_IRT_BLOCK_LAST
    print(f"irt 10 keV with {block}-thread workgroups", flush=True)
    ctx.set_option(DNA_OPT_IRT_BLOCK, block)
    try:
        g = ctx.irt(rad10k, None, 48, 10000.0, CK, seed=77)
        assert ctx.get_option(DNA_OPT_IRT_BLOCK_LAST) == block
    finally:
        ctx.set_option(DNA_OPT_IRT_BLOCK, 0)
    o = O.irt(rad10k, 48, 10000.0, CK, seed=77)
    _same(g, o)


def test_irt_autotune_picks_a_size_and_keeps_results(ctx):
    from dna_amd.abi import DNA_OPT_IRT_BLOCK, DNA_OPT_IRT_BLOCK_LAST
    p = default_params(1000.0, 1536)
    t = ctx.transport(p)
    print("irt autotune on 1536 x 1 keV", flush=True)
    ctx.set_option(DNA_OPT_IRT_BLOCK, 0)
    g = ctx.irt(None, t["rad_n"], 1536, 1000.0, CK, seed=5)
    assert ctx.get_option(DNA_OPT_IRT_BLOCK_LAST) in (64, 128, 256)
    ctx.set_option(DNA_OPT_IRT_BLOCK, 256)
    try:
        r = ctx.irt(None, t["rad_n"], 1536, 1000.0, CK, seed=5)
    finally:
        ctx.set_option(DNA_OPT_IRT_BLOCK, 0)
    _same(g, r)


def test_irt_empty_and_marker_only_inputs(ctx):
    """No radicals, and a primary with only an H2 marker (not a worker track,
    irt-worker.js:387-390, 470), as in the CPU restatement; with DNA_OPT_IRT_FINAL the
    final OH list and particle table are empty and the indirect pass scores nothing."""
    from dna_amd import DNATarget
    from dna_amd.abi import DNA_OPT_IRT_FINAL
    print("IRT on empty and marker-only inputs", flush=True)
    for rad in (np.zeros((0, 4), np.float32), np.array([[0, 0, 0, 7]], np.float32)):
        o = O.irt(rad, 1, 1000.0, CK)
        ctx.set_option(DNA_OPT_IRT_FINAL, 1)
        try:
            g = ctx.irt(rad, rad.shape[0], 1, 1000.0, CK)
        finally:
            ctx.set_option(DNA_OPT_IRT_FINAL, 0)
        assert g.n_tracks == o.n_tracks == 0
        assert [[g.tally[c][s] for s in range(5)] for c in range(7)] == [[o.tally[c][s] for s in range(5)] for c in range(7)]
        assert ctx.irt_oh_final().shape == (0, 4) and ctx.irt_particles().shape == (0, 6)
    dna = DNATarget()
    rad = np.array([[dna.x0, float(dna.fy[3]) + float(dna.rbb0[0]), float(dna.fz[3]) + float(dna.rbb0[1]), 0.0]], np.float32)
    s = ctx.score_oh(dna, rad, 0, None, 0, lcg_state=11)
    d = ctx.score(dna, rad, lcg_state=11)
    assert s["indirect"] == {"ssb0": 0, "ssb1": 0, "candidates": 0, "in_reach": 0}
    assert (s["ssb_count"], s["dsb"], s["lcg_state"]) == (d["ssb_count"], d["dsb"], d["lcg_state"])


def test_irt_interleaved_large_primary(ctx):
    """Entries of three primaries shuffled together, one of them with 9000 entries: the
    grouping's stable sort keeps input order within each primary at any size, that
    primary's species-sorted copy lives in global memory (over 3072 particles) and its
    events outgrow the LDS event list (3504 events against ~2550 LDS slots)."""
    rng = np.random.default_rng(21)

    def blob(pid, n, R):
        r = np.zeros((n, 4), np.float32)
        v = rng.normal(0, 1, (n, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        r[:, :3] = v * R * rng.random(n)[:, None] ** (1 / 3)
        r[:, 3] = pid * 8 + rng.choice([0, 0, 1, 1, 2, 3, 6], n)
        return r
    rad = np.concatenate([blob(7, 9000, 400.0), blob(3, 300, 50.0), blob(11, 500, 80.0)])
    rad = rad[rng.permutation(rad.shape[0])]
    print("irt: 3 interleaved primaries, the largest with 9000 entries", flush=True)
    g = ctx.irt(rad, rad.shape[0], 3, 10000.0, CK, seed=5)
    o = O.irt(rad, 3, 10000.0, CK, seed=5, n_threads=4)
    assert g.n_tracks == o.n_tracks == 3 and o.n_events > 3000
    _same(g, o)


def test_irt_track_beyond_lds_alive_bits(ctx):
    """A 15,000-entry primary beside a small one: its 30,512 particle slots need 3,816 bytes
    of alive bits, more than the loop kernel's LDS budget (2,048), so that launch keeps the
    bits in global memory; tallies still equal the oracle's."""
    rng = np.random.default_rng(22)

    def blob(pid, n, R):
        r = np.zeros((n, 4), np.float32)
        v = rng.normal(0, 1, (n, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        r[:, :3] = v * R * rng.random(n)[:, None] ** (1 / 3)
        r[:, 3] = pid * 8 + rng.choice([0, 0, 1, 1, 2, 3, 6], n)
        return r
    rad = np.concatenate([blob(2, 15000, 600.0), blob(5, 400, 60.0)])
    assert 4 * 2 * ((2 * 15000 + 512 + 63) // 64) > 2048
    print("irt: a 15000-entry primary (alive bits in global memory)", flush=True)
    g = ctx.irt(rad, rad.shape[0], 2, 10000.0, CK, seed=9)
    o = O.irt(rad, 2, 10000.0, CK, seed=9, n_threads=4)
    assert g.n_tracks == o.n_tracks == 2 and o.n_reacted > 2000
    _same(g, o)


@pytest.mark.parametrize("blocks", [1, 3])
def test_irt_few_event_blocks(ctx, rad10k, blocks):
    """DNA_OPT_IRT_EVBLOCKS: with 1 or 3 event blocks per track the blocks double many times
    during the loop (each doubling merges the minima of block pairs); tallies still equal
    the oracle's."""
    from dna_amd.abi import DNA_OPT_IRT_EVBLOCKS
    print(f"irt 10 keV with {blocks} event blocks per track", flush=True)
    ctx.set_option(DNA_OPT_IRT_EVBLOCKS, blocks)
    try:
        assert ctx.get_option(DNA_OPT_IRT_EVBLOCKS) == blocks
        g = ctx.irt(rad10k, None, 48, 10000.0, CK, seed=4)
    finally:
        ctx.set_option(DNA_OPT_IRT_EVBLOCKS, 128)
    o = O.irt(rad10k, 48, 10000.0, CK, seed=4)
    _same(g, o)
    from dna_amd.lib import DnaError
    for bad in (0, 129):
        with pytest.raises(DnaError):
            ctx.set_option(DNA_OPT_IRT_EVBLOCKS, bad)
    assert ctx.get_option(DNA_OPT_IRT_EVBLOCKS) == 128
