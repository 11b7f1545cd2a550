"""The cells of the IRT's initial-pair search (DNA_OPT_IRT_CELL) against the CPU restatement.

The cells only leave out pairs farther apart than R_CUT, and the re-pairing windows built
on them only list slots that the re-pairing's first screen would drop without a draw, so
for every cell width and workgroup size:
  * the tallies, reaction counts and event counts equal the oracle's all-pairs run
    (public/irt-worker.js:541-558 enumerates every pair);
  * the pairs within R_CUT (n_pair_sampled) are the same number as without cells: a pair
    missed or visited twice changes it;
  * the candidates visited (n_pair_cand) and the re-pairing slots scanned are never more
    than without cells;
  * the worker's work counts (all pairs with a channel, pairs within R_CUT, re-pairings)
    equal the oracle's, which counts them on its own all-pairs loops.
Edge cases: a track longer than 128 cells of R_CUT / 4 (the cells widen), every particle
at one point (one cell), a one-particle track, a track along a diagonal."""
import numpy as np
import pytest

import oracle as O
from dna_amd import default_params
from dna_amd.abi import (DNA_OPT_IRT_BLOCK, DNA_OPT_IRT_BLOCK_LAST, DNA_OPT_IRT_CELL,
                         DNA_OPT_IRT_CELL_LAST)

pytestmark = pytest.mark.gpu
CK = [0.001, 0.01, 0.1, 1.0, 10.0, 100.0, 1000.0]


def _same(g, o):
    assert g.n_tracks == o.n_tracks
    assert g.n_events == o.n_events and g.n_reacted == o.n_reacted
    assert list(g.rxn_counts) == list(o.rxn_counts)
    for c in range(len(CK)):
        assert [g.tally[c][s] for s in range(5)] == [o.tally[c][s] for s in range(5)], c
    # the worker's work, counted by the oracle on its all-pairs loops: pairs with a channel,
    # pairs within R_CUT, re-pairings with an alive partner
    assert (g.n_pair_all, g.n_pair_sampled, g.n_repair) == (o.n_pair_all, o.n_pair_sampled, o.n_repair)
    assert g.algo_flops == 9.0 * o.n_pair_all + 6.0 * o.n_pair_sampled + 167.0 * o.n_repair


def _run(ctx, rad, n_therm, E, seed, cell, block):
    ctx.set_option(DNA_OPT_IRT_CELL, cell)
    ctx.set_option(DNA_OPT_IRT_BLOCK, block)
    try:
        g = ctx.irt(rad, None, n_therm, E, CK, seed=seed)
        assert ctx.get_option(DNA_OPT_IRT_CELL_LAST) == cell
        assert ctx.get_option(DNA_OPT_IRT_BLOCK_LAST) == block
    finally:
        ctx.set_option(DNA_OPT_IRT_CELL, 0)
        ctx.set_option(DNA_OPT_IRT_BLOCK, 0)
    return g


@pytest.fixture(scope="module")
def rad20k(ctx):
    print("transport 24 x 20 keV for IRT input", flush=True)
    return ctx.transport(default_params(20000.0, 24))["rad"]


@pytest.mark.parametrize("block", [64, 256])
def test_cells_keep_the_pairs_and_the_results_20kev(ctx, rad20k, block):
    o = O.irt(rad20k, 24, 20000.0, CK, seed=9, n_threads=8)
    base = None
    for cell in (-1, 1, 2, 4):
        print(f"irt 24 x 20 keV, cells {cell}, {block} threads", flush=True)
        g = _run(ctx, rad20k, 24, 20000.0, 9, cell, block)
        _same(g, o)
        if base is None:
            base = g
            continue
        assert g.n_pair_sampled == base.n_pair_sampled, (cell, g.n_pair_sampled, base.n_pair_sampled)
        assert g.n_pair_cand <= base.n_pair_cand, (cell, g.n_pair_cand, base.n_pair_cand)
        assert g.n_repair_scanned <= base.n_repair_scanned
    # 20 keV tracks span several R_CUT: the finest cells visit far fewer candidates, and the
    # re-pairing windows skip list slots
    assert base.n_pair_cand == base.n_pair_all
    assert g.n_pair_cand < 0.8 * base.n_pair_cand
    assert g.n_repair_scanned < 0.9 * base.n_repair_scanned


def _line_track(n, length, rng, direction=(1.0, 0.0, 0.0), pid=0):
    """n radicals (OH, e-aq, H, H3O+, OH-) spread along a segment through the origin."""
    d = np.asarray(direction, np.float64)
    d /= np.linalg.norm(d)
    s = rng.random(n) * length
    pts = s[:, None] * d[None, :] + rng.normal(0.0, 3.0, (n, 3))
    codes = rng.choice(np.array([0, 1, 2, 3, 5, 6]), n)
    rad = np.zeros((n, 4), np.float32)
    rad[:, :3] = pts
    rad[:, 3] = pid * 8 + codes
    return rad


def _cases():
    rng = np.random.default_rng(21)
    long_ = _line_track(1500, 40000.0, rng)                          # > 128 cells of R_CUT / 4
    diag = _line_track(1200, 6000.0, rng, direction=(1.0, -2.0, 0.5), pid=1)
    point = np.zeros((300, 4), np.float32)
    point[:, 3] = 2 * 8 + rng.choice(np.array([0, 1, 2, 3]), 300)     # all at the origin
    one = np.array([[5.0, 5.0, 5.0, 3 * 8 + 1]], np.float32)
    return {"long": long_, "diagonal": diag, "point": point, "single": one,
            "together": np.concatenate([long_, diag, point, one])}


@pytest.mark.parametrize("case", ["long", "diagonal", "point", "single", "together"])
def test_cells_on_synthetic_tracks(ctx, case):
    rad = _cases()[case]
    n_pid = int(np.floor(rad[:, 3].max() / 8)) + 1
    o = O.irt(rad, n_pid, 10000.0, CK, seed=4)
    sampled = None
    for cell in (-1, 1, 4):
        print(f"irt synthetic track '{case}', cells {cell}", flush=True)
        g = _run(ctx, rad, n_pid, 10000.0, 4, cell, 128)
        _same(g, o)
        if sampled is None:
            sampled = g.n_pair_sampled
        assert g.n_pair_sampled == sampled


def test_autotune_picks_block_and_cells(ctx):
    p = default_params(1000.0, 1536)
    t = ctx.transport(p)
    print("irt autotune (block and cells) on 1536 x 1 keV", flush=True)
    ctx.set_option(DNA_OPT_IRT_CELL, 0)
    ctx.set_option(DNA_OPT_IRT_BLOCK, 0)
    g = ctx.irt(None, t["rad_n"], 1536, 1000.0, CK, seed=5)
    assert ctx.get_option(DNA_OPT_IRT_BLOCK_LAST) in (64, 128, 256)
    assert ctx.get_option(DNA_OPT_IRT_CELL_LAST) in (-1, 1, 2, 4)
    ctx.set_option(DNA_OPT_IRT_CELL, -1)
    ctx.set_option(DNA_OPT_IRT_BLOCK, 256)
    try:
        r = ctx.irt(None, t["rad_n"], 1536, 1000.0, CK, seed=5)
    finally:
        ctx.set_option(DNA_OPT_IRT_CELL, 0)
        ctx.set_option(DNA_OPT_IRT_BLOCK, 0)
    _same(g, r)
    assert g.n_pair_sampled == r.n_pair_sampled


def test_cell_option_is_checked(ctx):
    from dna_amd.lib import DnaError
    for bad in (3, -2, 8):
        with pytest.raises(DnaError):
            ctx.set_option(DNA_OPT_IRT_CELL, bad)
    assert ctx.get_option(DNA_OPT_IRT_CELL) == 0


@pytest.mark.parametrize("drain", [0, 1, 2])
def test_drain_placement_keeps_the_results(ctx, rad20k, drain):
    """The initial pairs' SamplePDC drain in the loop kernel (0), the pairs kernel (1) or a
    kernel of its own (2): the same events, so the same results as the oracle."""
    from dna_amd.abi import DNA_OPT_IRT_DRAIN
    o = O.irt(rad20k, 24, 20000.0, CK, seed=13, n_threads=8)
    ctx.set_option(DNA_OPT_IRT_DRAIN, drain)
    try:
        print(f"irt 24 x 20 keV, drain placement {drain}", flush=True)
        g = ctx.irt(rad20k, None, 24, 20000.0, CK, seed=13)
        assert ctx.get_option(DNA_OPT_IRT_DRAIN) == drain
    finally:
        ctx.set_option(DNA_OPT_IRT_DRAIN, 1)
    _same(g, o)


def test_drain_option_is_checked(ctx):
    from dna_amd.abi import DNA_OPT_IRT_DRAIN
    from dna_amd.lib import DnaError
    for bad in (-1, 3):
        with pytest.raises(DnaError):
            ctx.set_option(DNA_OPT_IRT_DRAIN, bad)
    assert ctx.get_option(DNA_OPT_IRT_DRAIN) == 1


@pytest.mark.parametrize("order", [0, 1])
def test_loop_order_keeps_the_results(ctx, rad20k, order):
    """The event-loop workgroups in launch order (0) or heaviest track first (1): each
    track's chain is the same, so the results are the oracle's either way."""
    from dna_amd.abi import DNA_OPT_IRT_ORDER
    o = O.irt(rad20k, 24, 20000.0, CK, seed=17, n_threads=8)
    ctx.set_option(DNA_OPT_IRT_ORDER, order)
    try:
        print(f"irt 24 x 20 keV, loop order {order}", flush=True)
        g = ctx.irt(rad20k, None, 24, 20000.0, CK, seed=17)
        assert ctx.get_option(DNA_OPT_IRT_ORDER) == order
    finally:
        ctx.set_option(DNA_OPT_IRT_ORDER, 1)
    _same(g, o)


def test_loop_order_over_several_chunks(ctx):
    """5000 tracks: two runs of the ordering sort (4096 tracks each at most), the second one
    partial; the results equal the launch-order run's and the oracle's."""
    from dna_amd.abi import DNA_OPT_IRT_ORDER
    print("transport 5000 x 1 keV for IRT input", flush=True)
    rad = ctx.transport(default_params(1000.0, 5000))["rad"]
    o = O.irt(rad, 5000, 1000.0, CK, seed=19, n_threads=8)
    runs = []
    for order in (1, 0):
        ctx.set_option(DNA_OPT_IRT_ORDER, order)
        try:
            print(f"irt 5000 x 1 keV, loop order {order}", flush=True)
            runs.append(ctx.irt(rad, None, 5000, 1000.0, CK, seed=19))
        finally:
            ctx.set_option(DNA_OPT_IRT_ORDER, 1)
    assert runs[0].n_tracks > 4096
    for g in runs:
        _same(g, o)
