"""Host-side mirror of the reference's TypeScript host for the Geant4-DNA chain.

Same names, argument meaning and result layouts as the reference, with the GPU work
routed to libdna_amd.so:
  ensure_pipelines(np, log)                      src/app.ts:65-84
  run_at_energy(device, buffers, pipelines, E_eV, np, box_nm, ce_eV, dna, run_chemistry=None)
                                                 src/gpu/dispatch.ts:85-285 -> EnergyResult
  run_chemistry_worker(rad_buf, rad_n, n_therm, E_eV, log=None)
                                                 src/chemistry/worker.ts:32-87 -> ChemResult
  score_direct_ssb_events / cluster_dsb / combine_hits / make_ssb_rng
                                                 src/scoring/ssb-dsb.ts
  score_damage_at_10kev                          src/app.ts:167-207
  run_validation(cfg)                            src/app.ts:91-161
Errors follow the reference: a missing device gives None plus an 'err' log line
(device.ts:12-21); invalid arguments raise.
"""
import time

import numpy as np

from .abi import DNA_COPY_ALL, DNA_COPY_RESULTS, chem_out_to_dict
from .geometry import (CHECKPOINT_LABELS, CHECKPOINTS_NS, DNA_GRID_N, DNA_LENGTH_NM,
                       DNA_SPACING_NM, KARAMITROS_2011, DNATarget, default_params)
from .lib import Context, DnaError

# src/scoring/estar.ts:8-17
ESTAR = [
    {"E": 100, "csda": 2.4, "sp": 58.0}, {"E": 300, "csda": 9.5, "sp": 35.0},
    {"E": 500, "csda": 16.8, "sp": 23.4}, {"E": 1000, "csda": 48.2, "sp": 12.1},
    {"E": 3000, "csda": 293, "sp": 5.18}, {"E": 5000, "csda": 685, "sp": 3.81},
    {"E": 10000, "csda": 2515, "sp": 2.29}, {"E": 20000, "csda": 8570, "sp": 1.33},
]
DEFAULT_CHEM_BACKEND = "worker"          # src/chemistry/backend.ts:18
RXN_LABELS = ["OH+OH→H2O2 [PDC]", "eaq+OH→OH- [PDC]", "OH+H→H2O [PDC]", "eaq+eaq→2OH-+H2 [TDC]",
              "eaq+H→OH-+H2 [TDC]", "eaq+H3O+→H [PDC]", "H+H→H2 [TDC]", "eaq+H2O2→OH-+OH [PDC]",
              "H3O++OH-→H2O [TDC]"]

_cache = None


def _noop(msg, cls=""):
    pass


def ensure_pipelines(np_, log=_noop):
    """Lazily create the device context (src/app.ts:65-84).  Returns None without a GPU."""
    global _cache
    if _cache is not None and _cache["np"] == np_:
        return _cache
    if _cache is None:
        try:
            ctx = Context(0)
        except (DnaError, FileNotFoundError, OSError) as e:
            log(f"GPU init failed: {e}", "err")
            return None
        # the reference's PipeCache {device, buffers, pipelines, np} (app.ts:65-84): the
        # device is the context; the library sizes its buffers on demand and its kernels are
        # compiled in, so buffers and pipelines carry nothing the calls need
        _cache = {"device": {"ctx": ctx}, "buffers": {"np": np_}, "pipelines": {}, "np": np_, "ctx": ctx}
    else:
        _cache["np"] = np_
        _cache["buffers"] = {"np": np_}
    log(f"Kernels loaded, buffers sized on demand (np={np_})", "data")
    return _cache


def run_at_energy(device, buffers, pipelines, E_eV, np_, box_nm, ce_eV, dna, run_chemistry=None, seed=None):
    """runAtEnergy (dispatch.ts:85-95), with the reference's arguments: Phase A + B on the
    GPU of `device` (from ensure_pipelines), then the host aggregation of
    dispatch.ts:178-226, then the chemistry callback at 10 keV (dispatch.ts:242-246).
    `seed` (not in the reference) overrides its floor(E_eV)."""
    ctx = device["ctx"]
    p = default_params(E_eV, np_, box_nm=box_nm, ce_eV=ce_eV, dna=dna, seed=seed)
    t0 = time.perf_counter()
    t = ctx.transport(p, copy=DNA_COPY_ALL)
    agg = aggregate_energy(E_eV, np_, t["results"], t["counters"], t["dose_sum"], t["sec_n"])
    c = t["counters"]
    rad_buf = t.get("rad")
    chem_result = None
    if E_eV == 10000 and rad_buf is not None and t["rad_n"] > 0 and run_chemistry is not None:
        chem_result = run_chemistry(rad_buf, t["rad_n"], agg["n_therm"], E_eV)
    dt = (time.perf_counter() - t0) * 1e3
    out = {"E": E_eV}
    out.update(agg)
    out.update({
        "dt": dt, "t_prim": t["t_prim_ms"], "t_sec": t["t_sec_ms"],
        "sec_n": t["sec_n"], "sec_dropped": t["sec_dropped"],
        "sec_terminated_cutoff": int(t["sec_stats"][1]), "sec_terminated_bounds": int(t["sec_stats"][2]),
        "sec_steps": int(t["sec_stats"][3]), "sec_tertiary_ions": int(t["sec_stats"][4]),
        "rad_OH": int(c[0]), "rad_eaq": int(c[1]), "rad_H": int(c[2]),
        "rad_n_raw": t["rad_n_raw"], "rad_n_stored": t["rad_n"], "rad_dropped": t["rad_dropped"],
        "kernel_dna_hits": int(c[4]), "chem_result": chem_result, "rad_buf_final": rad_buf,
        "dose_arr": t.get("dose") if E_eV == 10000 else None,
    })
    return out


def aggregate_energy(E_eV, np_, results, counters, dose_sum, sec_n):
    """The host aggregation of runAtEnergy (dispatch.ts:178-226) from the per-primary
    results R[np], the counters, the exact dose sum (100 units per eV) and the secondary
    count: means over the primaries that did not escape, G_init per 100 eV deposited and
    the energy conservation ratio."""
    res = results
    total_deposited_eV = dose_sum / 100.0
    c = counters
    per100 = total_deposited_eV / 100.0
    esc = res["esc"] == 1
    therm = ~esc
    n_therm, n_esc = int(therm.sum()), int(esc.sum())
    path = res["path"][therm].astype(np.float64)
    prod = res["prod"][therm].astype(np.float64)
    fE = res["fE"][therm].astype(np.float64)
    ni = res["ni"][therm].astype(np.float64)
    sp = np.where(path > 0, (E_eV - fE) / np.where(path > 0, path, 1), 0.0)
    mean = (lambda a: float(a.sum() / n_therm) if n_therm > 0 else 0.0)
    expected = n_therm * E_eV
    return {
        "n_therm": n_therm, "n_esc": n_esc,
        "mean_total": mean(path), "mean_prod": mean(prod), "mean_sp": mean(sp), "mean_ions": mean(ni),
        "sec_per_pri": sec_n / np_,
        "total_deposited_eV": total_deposited_eV, "expected_deposited_eV": expected,
        "cons_ratio": total_deposited_eV / expected if expected > 0 else 0.0,
        "G_OH": c[0] / per100 if per100 > 0 else 0.0, "G_eaq": c[1] / per100 if per100 > 0 else 0.0,
        "G_H": c[2] / per100 if per100 > 0 else 0.0,
    }


def table_row_cells(r, estar):
    """The results-table cells of one energy (src/ui/table.ts:49-65, toFixed formatting)
    from an aggregate_energy / run_at_energy dict and its ESTAR entry."""
    return {"n_therm": str(r["n_therm"]), "n_esc": str(r["n_esc"]),
            "mean_total": to_fixed(r["mean_total"], 1),
            "csda_ratio": to_fixed(r["mean_total"] / estar["csda"], 2) + "\u00d7",
            "mean_prod": to_fixed(r["mean_prod"], 1), "mean_sp": to_fixed(r["mean_sp"], 2),
            "sp_ratio": to_fixed(r["mean_sp"] / estar["sp"], 2) + "\u00d7",
            "mean_ions": to_fixed(r["mean_ions"], 1), "sec_per_pri": to_fixed(r["sec_per_pri"], 1),
            "cons_pct": to_fixed(r["cons_ratio"] * 100, 1) + "%",
            "G_OH": to_fixed(r["G_OH"], 2), "G_eaq": to_fixed(r["G_eaq"], 2), "G_H": to_fixed(r["G_H"], 2)}


def to_fixed(x, digits):
    """Number.prototype.toFixed: the exact binary value rounded half away from zero to
    `digits` decimals; zero of either sign prints without a sign."""
    from decimal import ROUND_HALF_UP, Decimal
    x = float(x)
    if x == 0.0:
        x = 0.0
    return str(Decimal(x).quantize(Decimal(1).scaleb(-digits), rounding=ROUND_HALF_UP))


_RXN_TABLE = None


def _rxn_table():
    global _RXN_TABLE
    if _RXN_TABLE is None:
        from .lib import irt_reactions
        _RXN_TABLE = irt_reactions()
    return _RXN_TABLE


def _chem_result(o, t_wall_ms):
    d = chem_out_to_dict(o)
    timeline = []
    for k, (lab, tns) in enumerate(zip(CHECKPOINT_LABELS, CHECKPOINTS_NS)):
        if k >= len(d["G"]):
            break
        g = d["G"][k]
        timeline.append({"label": lab, "t_ns": tns, "G_OH": g[0], "G_eaq": g[1], "G_H": g[2],
                         "G_H2O2": g[3], "G_H2": g[4]})
    tab = _rxn_table()                        # sigma and rc as the worker reports them (irt-worker.js:690-694)
    rxn_info = [{"label": RXN_LABELS[r], "count": d["rxn_counts"][r], "sigma": to_fixed(tab[r, 0], 4),
                 "rc": to_fixed(tab[r, 2], 3)} for r in range(9)]
    return {"chem_n": d["chem_n"], "t_wall": t_wall_ms, "timeline": timeline, "chem_pos_final": None,
            "chem_alive_final": None, "n_reacted": d["n_reacted"], "rxn_info": rxn_info}


def run_chemistry_worker(rad_buf_f32, rad_n, n_therm, E_eV, log=None, max_n=None, seed=1, oh_final=False,
                         options=None):
    """runChemistryWorker (src/chemistry/worker.ts:32-87): the 9-reaction IRT to 1 us on
    the GPU.  rad_buf_f32 is a (rad_n, 4) float32 array or a flat Float32Array-like.
    oh_final: also keep the OH alive at 1 us on the device for the indirect-SSB scoring
    (res["chem_oh_n"]); the worker itself returns no positions (worker.ts:69-70).
    options: {"dump_snapshots": bool, "snapshot_n": int} as runChemistry's
    (schedule.ts:16-27): res["snapshots"] then holds t = 0 and the seven checkpoints of the
    leading snapshot_n particle slots (default 50 000), for the WGDNA4D blob."""
    from .abi import DNA_OPT_IRT_FINAL
    cc = ensure_pipelines(_cache["np"] if _cache else 0)
    if cc is None:
        raise RuntimeError("IRT worker error: no GPU device")
    rad = np.ascontiguousarray(rad_buf_f32, np.float32).reshape(-1, 4)[:rad_n]
    dump = bool(options and options.get("dump_snapshots"))
    keep = oh_final or dump
    t0 = time.perf_counter()
    ctx = cc["ctx"]
    if keep:
        ctx.set_option(DNA_OPT_IRT_FINAL, 1)
    try:
        o = ctx.irt(rad, rad_n, n_therm, float(E_eV), list(CHECKPOINTS_NS), seed=seed)
    finally:
        if keep:
            ctx.set_option(DNA_OPT_IRT_FINAL, 0)
    res = _chem_result(o, (time.perf_counter() - t0) * 1e3)
    if oh_final:
        res["chem_oh_n"] = ctx.irt_oh_final(copy=False)
    if dump:
        from .splat import DEFAULT_SNAPSHOT_N, snapshots_from_particles
        res["snapshots"] = snapshots_from_particles(ctx.irt_particles(),
                                                    options.get("snapshot_n", DEFAULT_SNAPSHOT_N))
    if log:
        for cp in res["timeline"]:
            log(f"  {cp['label']} {cp['G_OH']:.3f} {cp['G_eaq']:.3f} {cp['G_H']:.3f} "
                f"{cp['G_H2O2']:.3f} {cp['G_H2']:.3f}", "data")
        log(f"  IRT worker: {res['n_reacted']} reactions in {res['t_wall'] / 1000:.1f}s "
            f"({res['chem_n']} radicals)", "data")
        for rx in res["rxn_info"]:                                # worker.ts:62-66
            log(f"    rxn {rx['label']}: {rx['count']} (σ={rx['sigma']}nm rc={rx['rc']}nm)", "data")
    return res


def make_ssb_rng(seed=0x12345678):
    """makeSsbRng (ssb-dsb.ts:421-427) as a mutable LCG state shared across the sweep."""
    return {"state": seed & 0xFFFFFFFF}


def score_damage_at_10kev(dna, chem, rad_buf, rad_n, total_deposited_eV, box_nm, rng, log=_noop,
                          kernel_dna_hits=0):
    """scoreDamageAt10keV (src/app.ts:167-207): direct-event SSB + combineHits + clusterDSB
    on the GPU.  Indirect hits are empty on the worker path (app.ts:178-180) unless the
    chemistry kept its OH at 1 us (run_chemistry_worker(oh_final=True)): then
    scoreIndirectSSB runs first on those, as on the reference's GPU-chemistry path."""
    cc = ensure_pipelines(_cache["np"] if _cache else 0, log)
    if cc is None:
        raise DnaError("score_damage_at_10kev", 2, "no device")
    oh_n = chem.get("chem_oh_n") if chem else None
    if oh_n is not None:
        s = cc["ctx"].score_oh(dna, rad_buf, rad_n, None, oh_n, lcg_state=rng["state"])
        ind = s["indirect"]
    else:
        s = cc["ctx"].score(dna, rad_buf, rad_n, lcg_state=rng["state"])
        ind = {"ssb0": 0, "ssb1": 0, "in_reach": 0}
    rng["state"] = s["lcg_state"]
    ssb_dir, ssb_ind, dsb = s["ssb_count"], ind["ssb0"] + ind["ssb1"], s["dsb"]
    dose_box_Gy = (total_deposited_eV * 1.602e-19) / ((2 * box_nm * 1e-9) ** 3 * 1000)
    target_Gbp = dna.n_bp * 1e-9
    dsb_per_gy_gbp = dsb / (dose_box_Gy * target_Gbp) if dose_box_Gy > 0 else 0.0
    ssb_tot = ssb_dir + ssb_ind
    log(f"    DAMAGE: SSB_dir={ssb_dir}  SSB_ind={ssb_ind}  DSB={dsb}  "
        f"(DSB/SSB={(dsb / ssb_tot if ssb_tot else 0):.3f})  box_dose={dose_box_Gy:.3f} Gy  "
        f"target={dna.n_bp / 1e6:.1f} Mbp  reach_dir={s['in_reach']} reach_ind={ind['in_reach']}  "
        f"kernel_hits={kernel_dna_hits}", "data")
    return {"ssb_dir": ssb_dir, "ssb_ind": ssb_ind, "dsb": dsb, "dsb_per_gy_gbp": dsb_per_gy_gbp,
            "dsb_expected": s["dsb_expected"], "in_reach": s["in_reach"]}


def run_validation(cfg):
    """runValidation (src/app.ts:91-161): the 8-energy ESTAR sweep; chemistry and DNA
    scoring at 10 keV.  Returns the rows and a results dict in the layout of
    validation/webgpu-results.json."""
    np_, box_nm, ce_eV = cfg["np"], cfg["boxNm"], cfg["ceEV"]
    log = cfg.get("log", _noop)
    backend = cfg.get("chemBackend", DEFAULT_CHEM_BACKEND)
    energies = cfg.get("energies", [e["E"] for e in ESTAR])
    cc = ensure_pipelines(np_, log)
    if cc is None:
        return None
    log(f"Running {np_} primaries at {len(energies)} energies, box={box_nm}nm, cutoff={ce_eV}eV", "data")
    dna = DNATarget(DNA_LENGTH_NM, DNA_GRID_N, DNA_SPACING_NM)
    ssb_rng = make_ssb_rng()
    chem_cb = None
    indirect = bool(cfg.get("indirectSSB", False))       # extension: score the OH at 1 us too
    if backend != "none":
        def chem_cb(rad, rad_n, n_therm, E):
            return run_chemistry_worker(rad, rad_n, n_therm, E, log, oh_final=indirect and E == 10000)
    rows, out = [], None
    for E in energies:
        ref = next((e for e in ESTAR if e["E"] == E), {"E": E, "csda": float("nan")})
        r = run_at_energy(cc["device"], cc["buffers"], cc["pipelines"], float(E), np_, box_nm, ce_eV, dna, chem_cb)
        damage = None
        if r["chem_result"] and r["rad_buf_final"] is not None and E == 10000:
            damage = score_damage_at_10kev(dna, r["chem_result"], r["rad_buf_final"], r["rad_n_stored"],
                                           r["total_deposited_eV"], box_nm, ssb_rng, log, r["kernel_dna_hits"])
        log(f"  E={E} eV: CSDA={r['mean_total']:.1f}nm ({r['mean_total'] / ref['csda']:.2f}x) "
            f"ions/pri={r['mean_ions']:.1f} sec/pri={r['sec_per_pri']:.1f} "
            f"G_init(OH/eaq/H)={r['G_OH']:.2f}/{r['G_eaq']:.2f}/{r['G_H']:.2f} "
            f"rad_n={r['rad_n_stored']} E_cons={r['cons_ratio'] * 100:.1f}%", "data")
        row = {k: v for k, v in r.items() if k not in ("rad_buf_final", "dose_arr")}
        row["damage"] = damage
        rows.append(row)
        if E == 10000:
            out = results_json(r, damage, np_)
    log("Validation run complete.", "ok")
    return {"rows": rows, "results": out}


def results_json(r, damage, np_):
    """The layout of validation/webgpu-results.json:1-31."""
    cr = r["chem_result"]
    g = cr["timeline"][-1] if cr else {}
    return {
        "primaryEnergyEv": r["E"], "nPrimaries": np_,
        "tracking": {"csdaNm": round(r["mean_total"], 1), "ionsPerPriPrimaryOnly": round(r["mean_ions"], 1),
                     "secPerPri": round(r["sec_per_pri"], 1), "eConservationPct": round(r["cons_ratio"] * 100, 1)},
        "preChemistry": {"G_OH_init": round(r["G_OH"], 2), "G_eaq_init": round(r["G_eaq"], 2),
                         "G_H_init": round(r["G_H"], 2)},
        "chemistry1Us": {k: round(g.get(k, 0.0), 3) for k in ("G_OH", "G_eaq", "G_H", "G_H2O2", "G_H2")},
        "dnaDamage": {"SSB_dir": damage["ssb_dir"] if damage else 0,
                      "SSB_ind": damage["ssb_ind"] if damage else 0,
                      "DSB": damage["dsb"] if damage else 0, "kernelHits": r["kernel_dna_hits"]},
        "karamitros2011": KARAMITROS_2011,
    }
