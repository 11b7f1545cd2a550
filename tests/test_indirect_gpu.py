"""Indirect SSB on the GPU against the CPU restatement, bit for bit: the OH alive at 1 us
(dna_irt_oh_final after an IRT run with DNA_OPT_IRT_FINAL), from a 10 keV transport and
from synthetic one-OH tracks given as host radicals; then dna_score_oh (indirect pass
first, the LCG continuing into the direct pass, combined hits) on the resident and on
host-supplied lists, with every count, the LCG state after the call and the combined
hit array compared."""
import numpy as np
import pytest

import oracle as O
from dna_amd import DNATarget, default_params
from dna_amd.abi import DNA_COPY_RESULTS, DNA_OPT_IRT_FINAL

pytestmark = pytest.mark.gpu
CK = [0.001, 0.1, 10.0, 1000.0]


def run(ctx, p, seed=5):
    t = ctx.transport(p, copy=DNA_COPY_RESULTS | 4)
    n_therm = int((t["results"]["esc"] == 0).sum())
    ctx.set_option(DNA_OPT_IRT_FINAL, 1)
    try:
        c = ctx.irt(None, t["rad_n"], n_therm, p.be, CK, seed=seed)
    finally:
        ctx.set_option(DNA_OPT_IRT_FINAL, 0)
    return t, n_therm, c


def test_oh_final_exact_10kev(ctx):
    p = default_params(10000.0, 64)
    print("OH at 1 us, 10 keV x 64", flush=True)
    t, n_therm, c = run(ctx, p)
    g = ctx.irt_oh_final()
    o, oh = O.irt_oh(t["rad"], n_therm, 10000.0, CK, seed=5, n_threads=8)
    assert g.shape == oh.shape and g.shape[0] == c.tally[3][0]
    assert np.array_equal(g.view(np.uint32), oh.view(np.uint32))


def test_oh_final_exact_host_input(ctx):
    rng = np.random.default_rng(3)
    n = 300
    rad = np.zeros((n, 4), np.float32)
    rad[:, :3] = rng.uniform(-50, 50, (n, 3))
    rad[:, 3] = np.arange(n) * 8.0
    print("OH at 1 us, 300 one-OH tracks", flush=True)
    ctx.set_option(DNA_OPT_IRT_FINAL, 1)
    try:
        ctx.irt(rad, n, n, 1000.0, [1000.0], seed=9)
    finally:
        ctx.set_option(DNA_OPT_IRT_FINAL, 0)
    g = ctx.irt_oh_final()
    _, oh = O.irt_oh(rad, n, 1000.0, [1000.0], seed=9)
    assert np.array_equal(g.view(np.uint32), oh.view(np.uint32))


def test_score_oh_exact(ctx):
    dna = DNATarget()
    p = default_params(10000.0, 512, dna=dna)
    print("indirect + direct scoring, 10 keV x 512", flush=True)
    t, n_therm, c = run(ctx, p, seed=6)
    n_oh = ctx.irt_oh_final(copy=False)
    oh = ctx.irt_oh_final()
    g = ctx.score_oh(dna, None, t["rad_n"], None, n_oh, lcg_state=0x12345678, want_hits=True)
    o = O.score_oh(dna, t["rad"], oh, lcg_state=0x12345678, want_hits=True)
    for k in ("ssb_count", "candidates", "in_reach", "dsb", "ssb0", "ssb1", "lcg_state"):
        assert g[k] == o[k], k
    assert g["indirect"] == o["indirect"]
    assert g["indirect"]["candidates"] > 0
    assert np.array_equal(g["hits"], o["hits"])
    assert g["dsb_expected"] == pytest.approx(o["dsb_expected"], rel=1e-12)
    # host-supplied lists give the same result
    h = ctx.score_oh(dna, t["rad"], 0, oh, 0, lcg_state=0x12345678)
    for k in ("ssb_count", "dsb", "lcg_state"):
        assert h[k] == o[k], k
    assert h["indirect"] == o["indirect"]
    # without OH, the direct scoring alone (dna_score)
    d = ctx.score(dna, None, t["rad_n"], lcg_state=0x12345678)
    od = O.score(dna, t["rad"], lcg_state=0x12345678)
    assert (d["ssb_count"], d["dsb"], d["lcg_state"]) == (od["ssb_count"], od["dsb"], od["lcg_state"])


def test_run_validation_with_indirect_ssb():
    """The host mirror's 10 keV row with the extension on: the indirect pass runs on the
    OH kept by the chemistry and its hits join the direct ones."""
    from dna_amd import host
    print("runValidation at 10 keV with indirect SSB", flush=True)
    out = host.run_validation({"np": 2048, "boxNm": 15000.0, "ceEV": 7.4, "energies": [10000],
                               "indirectSSB": True})
    dmg = out["rows"][0]["damage"]
    base = host.run_validation({"np": 2048, "boxNm": 15000.0, "ceEV": 7.4, "energies": [10000]})
    d0 = base["rows"][0]["damage"]
    print("with indirect:", dmg, "direct only:", d0, flush=True)
    assert d0["ssb_ind"] == 0 and dmg["ssb_ind"] > 0


def test_final_state_survives_a_plan(ctx):
    """Grouping the input of a captured plan leaves the final state kept by
    DNA_OPT_IRT_FINAL untouched: the OH list is the same before and after planning a batch
    of more than 1024 entries."""
    p = default_params(5000.0, 48)
    print("final state, then a plan over the resident radicals", flush=True)
    t, n_therm, c = run(ctx, p, seed=7)
    before = ctx.irt_oh_final()
    assert t["rad_n"] > 1024
    plan = ctx.irt_plan(None, [0, t["rad_n"]], [n_therm], [5000.0], CK)
    try:
        plan.run(3)
    finally:
        plan.close()
    after = ctx.irt_oh_final()
    assert np.array_equal(before.view(np.uint32), after.view(np.uint32))
