"""Chains in flight: two contexts driven from two host threads at once (what bench.py does)
give the same transport, chemistry and scoring results as one context run sequentially.
Every result is a pure function of its inputs and seeds, so the overlap of the two
contexts' kernels on the GPU must change nothing."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from dna_amd import Context, DNATarget, default_params
from dna_amd.abi import DNA_COPY_RESULTS

pytestmark = pytest.mark.gpu
CK = [0.001, 0.01, 0.1, 1.0, 10.0, 100.0, 1000.0]


def chain(ctx, p, dna):
    t = ctx.transport(p, copy=DNA_COPY_RESULTS)
    n_therm = int((t["results"]["esc"] == 0).sum())
    c = ctx.irt(None, t["rad_n"], n_therm, p.be, CK, seed=3)
    s = ctx.score(dna, None, t["rad_n"], lcg_state=0x12345678)
    tallies = [[c.tally[k][sp] for sp in range(5)] for k in range(len(CK))]
    return (t["results"].tobytes(), t["dose_sum"], t["rad_n"], tallies, list(c.rxn_counts),
            s["ssb_count"], s["dsb"])


def test_two_contexts_in_flight_match_sequential():
    dna = DNATarget()
    params = [default_params(10000.0, 512, dna=dna), default_params(3000.0, 1024, dna=dna)]
    ref = Context(0)
    expect = [chain(ref, p, dna) for p in params]
    ref.close()
    ctxs = [Context(0), Context(0)]
    try:
        print("two contexts, three chain steps each, in flight", flush=True)
        with ThreadPoolExecutor(2) as ex:
            got = list(ex.map(lambda i: [chain(ctxs[i], params[i], dna) for _ in range(3)], range(2)))
        for i in range(2):
            for r in got[i]:
                assert r == expect[i], i
    finally:
        for c in ctxs:
            c.close()
