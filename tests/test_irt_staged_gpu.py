"""dna_irt_stage / dna_irt_finish: the IRT in two calls gives dna_run_irt's results, also
with another context's IRT between the calls, and refuses to finish what it cannot."""
import numpy as np
import pytest

import oracle as O
from dna_amd import Context, default_params
from dna_amd.lib import DnaError

pytestmark = pytest.mark.gpu
CK = [0.001, 0.01, 0.1, 1.0, 10.0, 100.0, 1000.0]


def _key(o):
    return (np.array([[o.tally[c][s] for s in range(5)] for c in range(o.n_ckpt)], np.uint64).tolist(),
            list(o.rxn_counts[:]), o.n_reacted, o.n_events, o.n_tracks, o.n_pair_cand,
            o.n_pair_sampled, o.n_repair, o.n_repair_scanned, o.n_pair_all)


def test_staged_equals_one_call(ctx):
    p = default_params(1000.0, 200)
    t = ctx.transport(p)
    print("irt in one call, then staged, 200 x 1 keV resident", flush=True)
    one = ctx.irt(None, t["rad_n"], 200, 1000.0, CK, seed=11)
    ctx.irt_stage(None, t["rad_n"], 200, 1000.0, CK)
    two = ctx.irt_finish(11)
    assert _key(two) == _key(one)
    o = O.irt(t["rad"], 200, 1000.0, CK, seed=11)
    assert _key(two)[:5] == (np.array([[o.tally[c][s] for s in range(5)] for c in range(o.n_ckpt)],
                                      np.uint64).tolist(), list(o.rxn_counts[:]), o.n_reacted,
                             o.n_events, o.n_tracks)


def test_staged_host_input_and_two_contexts(ctx):
    """Host radicals; a second context stages and finishes its own IRT between the calls."""
    p = default_params(10000.0, 24)
    rad = ctx.transport(p)["rad"]
    one = ctx.irt(rad, None, 24, 10000.0, CK, seed=5)
    c2 = Context(0)
    try:
        t2 = c2.transport(default_params(1000.0, 64))
        other = c2.irt(None, t2["rad_n"], 64, 1000.0, CK, seed=2)
        print("stage on ctx 1, a whole staged IRT on ctx 2, finish ctx 1", flush=True)
        ctx.irt_stage(rad, None, 24, 10000.0, CK)
        c2.irt_stage(None, t2["rad_n"], 64, 1000.0, CK)
        assert _key(c2.irt_finish(2)) == _key(other)
        assert _key(ctx.irt_finish(5)) == _key(one)
    finally:
        c2.close()


def test_staged_refusals(ctx):
    p = default_params(1000.0, 32)
    t = ctx.transport(p)
    print("finish without a staged IRT", flush=True)
    with pytest.raises(DnaError):
        ctx.irt_finish(1)
    print("transport between stage and finish (resident radicals)", flush=True)
    ctx.irt_stage(None, t["rad_n"], 32, 1000.0, CK)
    t = ctx.transport(p)
    with pytest.raises(DnaError):
        ctx.irt_finish(1)
    print("another IRT call between stage and finish drops the staged one", flush=True)
    ctx.irt_stage(None, t["rad_n"], 32, 1000.0, CK)
    ctx.irt(None, t["rad_n"], 32, 1000.0, CK, seed=1)
    with pytest.raises(DnaError):
        ctx.irt_finish(1)
    print("finish once only", flush=True)
    ctx.irt_stage(None, t["rad_n"], 32, 1000.0, CK)
    ctx.irt_finish(1)
    with pytest.raises(DnaError):
        ctx.irt_finish(1)


def test_staged_empty_input(ctx):
    ctx.irt_stage(np.zeros((0, 4), np.float32), None, 0, 1000.0, CK)
    o = ctx.irt_finish(1)
    assert o.n_tracks == 0 and o.n_events == 0


def test_launch_and_wait_equal_one_call(ctx):
    """dna_irt_launch + dna_irt_wait (event-loop start, then all) give dna_run_irt's results;
    waiting without a launch, or launching twice, is refused."""
    p = default_params(10000.0, 24)
    t = ctx.transport(p)
    one = ctx.irt(None, t["rad_n"], 24, 10000.0, CK, seed=8)
    print("stage, launch, wait for the event loop, wait for all", flush=True)
    ctx.irt_stage(None, t["rad_n"], 24, 10000.0, CK)
    ctx.irt_launch(8)
    with pytest.raises(DnaError):
        ctx.irt_launch(8)
    assert ctx.irt_wait(loop_only=True) is None
    assert _key(ctx.irt_wait()) == _key(one)
    with pytest.raises(DnaError):
        ctx.irt_wait()
