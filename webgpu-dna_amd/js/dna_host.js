'use strict';
/**
 * dna_host.js — the reference's host API for the Geant4-DNA chain, on the MI355X
 * library through the Node-API addon (lib/dna_amd.node).  CommonJS for Node >= 12.
 *
 * Same names, argument meaning and result layouts as the reference's TypeScript host:
 *   ensurePipelines(np, log)                     src/app.ts:65-84
 *   runAtEnergy(device, buffers, pipelines, E_eV, np, boxNm, ceEV, dna, runChemistry?)
 *                                                src/gpu/dispatch.ts:85-285 -> EnergyResult
 *   runChemistryWorker(radBuf, radN, nTherm, E_eV, log?, maxN?, seed?, options?)
 *                                                src/chemistry/worker.ts:32-87 -> ChemResult
 *                                                (options: {ohFinal} keeps the OH alive at 1 us for
 *                                                the indirect-SSB scoring; {dump_snapshots, snapshot_n}
 *                                                as runChemistry's, src/chemistry/schedule.ts:16-27)
 *   buildDNATarget(L_nm, grid_N, spacing_nm)     src/physics/dna-geometry.ts:16-68
 *   makeSsbRng(seed)                             src/scoring/ssb-dsb.ts:421-427
 *   scoreDamageAt10keV(...)                      src/app.ts:167-207
 *   runValidation(cfg)                           src/app.ts:91-161 (cfg.indirectSSB: the extension
 *                                                that scores the OH at 1 us first, ssb-dsb.ts:43-121)
 * The WGDNA4D snapshot blob is in splat.js.  A missing device gives null plus an 'err' log line (src/gpu/device.ts:12-21).
 * Types: dna_host.d.ts.
 */
const path = require('path');

const addon = require(path.join(__dirname, '..', 'lib', 'dna_amd.node'));

// src/physics/constants.ts
const SOLVATION_THRESHOLD = 7.4;
const VC = 128;
const MAX_SEC_STEPS = 2000;
const DNA_LENGTH_NM = 3000;
const DNA_GRID_N = 21;
const DNA_SPACING_NM = 150;
const KARAMITROS_2011 = { G_OH: 2.50, G_eaq: 2.50, G_H: 0.57, G_H2O2: 0.73, G_H2: 0.42 };
// src/scoring/estar.ts:8-17
const ESTAR = [
  { E: 100, csda: 2.4, sp: 58.0 }, { E: 300, csda: 9.5, sp: 35.0 },
  { E: 500, csda: 16.8, sp: 23.4 }, { E: 1000, csda: 48.2, sp: 12.1 },
  { E: 3000, csda: 293, sp: 5.18 }, { E: 5000, csda: 685, sp: 3.81 },
  { E: 10000, csda: 2515, sp: 2.29 }, { E: 20000, csda: 8570, sp: 1.33 },
];
// public/irt-worker.js:416-417
const CHECKPOINTS_NS = [0.001, 0.01, 0.1, 1.0, 10.0, 100.0, 1000.0];
const CHECKPOINT_LABELS = ['1 ps', '10 ps', '100 ps', '1 ns', '10 ns', '100 ns', '1 us'];
const RXN_LABELS = ['OH+OH→H2O2 [PDC]', 'eaq+OH→OH- [PDC]', 'OH+H→H2O [PDC]', 'eaq+eaq→2OH-+H2 [TDC]',
  'eaq+H→OH-+H2 [TDC]', 'eaq+H3O+→H [PDC]', 'H+H→H2 [TDC]', 'eaq+H2O2→OH-+OH [PDC]',
  'H3O++OH-→H2O [TDC]'];
const COPY_RESULTS = 1, COPY_DOSE = 2, COPY_RAD = 4, COPY_ALL = 7;
// include/dna_amd.h options
const OPT_IRT_BLOCK = 1, OPT_IRT_BLOCK_LAST = 2, OPT_IRT_FINAL = 3, OPT_PA_LANES = 4;

function noop() {}

/** buildDNATarget (dna-geometry.ts:16-68): fibers along X on a (y, z) lattice, 10.5 bp/turn. */
function buildDNATarget(L_nm = DNA_LENGTH_NM, grid_N = DNA_GRID_N, spacing_nm = DNA_SPACING_NM) {
  const rise = 0.34, bpPerTurn = 10.5, r_bb = 1.0;
  const n_bp_per = Math.floor(L_nm / rise);
  const x0 = -(n_bp_per - 1) * rise * 0.5;
  const n_fibers = grid_N * grid_N;
  const off = -((grid_N - 1) * spacing_nm) * 0.5;
  const fy = new Float32Array(n_fibers), fz = new Float32Array(n_fibers);
  for (let i = 0; i < grid_N; i++) {
    for (let j = 0; j < grid_N; j++) {
      fy[i * grid_N + j] = off + i * spacing_nm;
      fz[i * grid_N + j] = off + j * spacing_nm;
    }
  }
  const dPhase = (2 * Math.PI) / bpPerTurn;
  const rbb0 = new Float32Array(2 * n_bp_per), rbb1 = new Float32Array(2 * n_bp_per);
  for (let k = 0; k < n_bp_per; k++) {
    const phi = k * dPhase;
    rbb0[2 * k] = r_bb * Math.cos(phi);
    rbb0[2 * k + 1] = r_bb * Math.sin(phi);
    rbb1[2 * k] = r_bb * Math.cos(phi + Math.PI);
    rbb1[2 * k + 1] = r_bb * Math.sin(phi + Math.PI);
  }
  return {
    rise, r_bb, spacing_nm, x0, L_nm, n_bp_per, n_fibers, n_bp: n_fibers * n_bp_per, grid_N,
    fy, fz, rbb0, rbb1,
  };
}

/** writePrimaryParams (dispatch.ts:22-41) plus the seed (dispatch.ts:98: floor(E)). */
function primaryParams(E_eV, np, boxNm, ceEV, dna, seed) {
  const on = dna != null && E_eV === 10000;
  return {
    n: np, box: boxNm, ce: ceEV, ms: 65536, be: E_eV, maxSec: 0, vc: VC, maxRad: 0, startHalf: 0,
    dnaEnable: on ? 1 : 0, dnaGridN: dna ? dna.grid_N : 0, dirMode: 0,
    dnaRise: dna ? dna.rise : 0, dnaSpacing: dna ? dna.spacing_nm : 0, dnaX0: dna ? dna.x0 : 0,
    dnaRbb: dna ? dna.r_bb : 0, seed: seed == null ? Math.floor(E_eV) : seed, maxSecSteps: MAX_SEC_STEPS,
  };
}

let cache = null;

/** ensurePipelines (app.ts:65-84): lazily create the device context; null without a GPU. */
function ensurePipelines(np, log = noop) {
  if (cache && cache.np === np) return cache;
  if (!cache) {
    let ctx = null;
    try {
      ctx = addon.ctxCreate(0);
    } catch (e) {
      log(`GPU init failed: ${e.message}`, 'err');
      return null;
    }
    if (ctx === null) {
      log(`WebGPU-equivalent device not available: ${addon.lastError()}`, 'err');
      return null;
    }
    // the reference's PipeCache {device, buffers, pipelines, np} (app.ts:65-84): the device
    // is the addon context; buffers are sized by the library on demand and its kernels are
    // compiled in, so buffers and pipelines carry nothing the calls need
    cache = { device: { ctx }, buffers: { np }, pipelines: {}, np, ctx };
  } else {
    cache.np = np;
    cache.buffers = { np };
  }
  log(`Kernels loaded, buffers sized on demand (np=${np})`, 'data');
  return cache;
}

/** runAtEnergy (dispatch.ts:85-285), with the reference's arguments: Phase A + B on the
 * GPU of `device` (the context of ensurePipelines), the host aggregation of
 * dispatch.ts:178-226, then the chemistry callback at 10 keV.  `seed` (not in the
 * reference) overrides the reference's floor(E_eV). */
async function runAtEnergy(device, buffers, pipelines, E_eV, np, boxNm, ceEV, dna, runChemistry, seed) {
  const t0 = Date.now();
  const t = await addon.runTransport(device.ctx, primaryParams(E_eV, np, boxNm, ceEV, dna, seed), null, COPY_ALL);
  const f32 = new Float32Array(t.results), u32 = new Uint32Array(t.results);
  const total_deposited_eV = t.doseSum / 100;
  const per100 = total_deposited_eV / 100;
  const c = t.counters;
  let n_therm = 0, n_esc = 0, sPath = 0, sProd = 0, sSp = 0, sIons = 0;
  for (let i = 0; i < np; i++) {
    const b = i * 8;                       // R = {path, prod, fE, ni, nx, esc, mx, _pad}
    if (u32[b + 5] === 1) { n_esc++; continue; }
    n_therm++;
    const pth = f32[b], fE = f32[b + 2];
    sPath += pth;
    sProd += f32[b + 1];
    sSp += pth > 0 ? (E_eV - fE) / pth : 0;
    sIons += u32[b + 3];
  }
  const mean = (s) => (n_therm > 0 ? s / n_therm : 0);
  const expected = n_therm * E_eV;
  let chem_result = null;
  if (E_eV === 10000 && t.rad && t.radN > 0 && runChemistry) {
    chem_result = await runChemistry(t.rad, t.radN, n_therm, E_eV);
  }
  return {
    E: E_eV, n_therm, n_esc,
    mean_total: mean(sPath), mean_prod: mean(sProd), mean_sp: mean(sSp), mean_ions: mean(sIons),
    dt: Date.now() - t0, t_prim: t.tPrimMs, t_sec: t.tSecMs,
    sec_n: t.secN, sec_dropped: t.secDropped, sec_per_pri: t.secN / np,
    sec_terminated_cutoff: t.secStats[1], sec_terminated_bounds: t.secStats[2],
    sec_steps: t.secStats[3], sec_tertiary_ions: t.secStats[4],
    total_deposited_eV, expected_deposited_eV: expected,
    cons_ratio: expected > 0 ? total_deposited_eV / expected : 0,
    rad_OH: c[0], rad_eaq: c[1], rad_H: c[2],
    G_OH: per100 > 0 ? c[0] / per100 : 0, G_eaq: per100 > 0 ? c[1] / per100 : 0,
    G_H: per100 > 0 ? c[2] / per100 : 0,
    rad_n_raw: t.radNRaw, rad_n_stored: t.radN, rad_dropped: t.radDropped,
    kernel_dna_hits: c[4], chem_result, rad_buf_final: t.rad,
    dose_arr: E_eV === 10000 ? t.dose : null,
  };
}

let rxnTab = null;
/** The IRT's reaction table (dna_irt_reactions): 9 x (sigma, sigma_eff, rc, k_dif, k_act, p). */
function rxnTable() {
  if (!rxnTab) rxnTab = addon.irtReactions();
  return rxnTab;
}

/** runChemistryWorker (worker.ts:32-87): the 9-reaction IRT to 1 us on the GPU. */
async function runChemistryWorker(radBuf, radN, nTherm, E_eV, log, maxN, seed = 1, options = null) {
  const cc = ensurePipelines(cache ? cache.np : 0);
  if (!cc) throw new Error('IRT worker error: no GPU device');
  const t0 = Date.now();
  const rad = radBuf.length > 4 * radN ? radBuf.subarray(0, 4 * radN) : radBuf;
  const ohFinal = !!(options && options.ohFinal), dump = !!(options && options.dump_snapshots);
  if (ohFinal || dump) addon.setOption(cc.ctx, OPT_IRT_FINAL, 1);
  let o;
  try {
    o = await addon.runIrt(cc.ctx, rad, radN, nTherm, E_eV, Float64Array.from(CHECKPOINTS_NS), seed);
  } finally {
    if (ohFinal || dump) addon.setOption(cc.ctx, OPT_IRT_FINAL, 0);
  }
  const timeline = [];
  for (let k = 0; k < o.nCkpt; k++) {
    timeline.push({
      label: CHECKPOINT_LABELS[k], t_ns: CHECKPOINTS_NS[k], G_OH: o.G[k * 5], G_eaq: o.G[k * 5 + 1],
      G_H: o.G[k * 5 + 2], G_H2O2: o.G[k * 5 + 3], G_H2: o.G[k * 5 + 4],
    });
  }
  const res = {
    chem_n: o.chemN, t_wall: Date.now() - t0, timeline, chem_pos_final: null, chem_alive_final: null,
    n_reacted: o.nReacted,
    // sigma and rc as the worker reports them (irt-worker.js:690-694)
    rxn_info: RXN_LABELS.map((label, r) => ({
      label, count: o.rxnCounts[r], sigma: rxnTable()[6 * r].toFixed(4), rc: rxnTable()[6 * r + 2].toFixed(3),
    })),
  };
  if (ohFinal) res.chem_oh_n = (await addon.irtOhFinal(cc.ctx, false)).n;
  if (dump) {
    const splat = require('./splat');
    const p = await addon.irtParticles(cc.ctx);
    res.snapshots = splat.snapshotsFromParticles(p.particles,
      options.snapshot_n != null ? options.snapshot_n : splat.DEFAULT_SNAPSHOT_N);
  }
  if (log) {
    for (const cp of timeline) {
      log(`  ${cp.label} ${cp.G_OH.toFixed(3)} ${cp.G_eaq.toFixed(3)} ${cp.G_H.toFixed(3)} ` +
          `${cp.G_H2O2.toFixed(3)} ${cp.G_H2.toFixed(3)}`, 'data');
    }
    log(`  IRT worker: ${res.n_reacted} reactions in ${(res.t_wall / 1000).toFixed(1)}s (${res.chem_n} radicals)`, 'data');
    for (const rx of res.rxn_info) log(`    rxn ${rx.label}: ${rx.count} (σ=${rx.sigma}nm rc=${rx.rc}nm)`, 'data');
  }
  return res;
}

/** makeSsbRng (ssb-dsb.ts:421-427): the LCG as a state shared across the sweep.  The
 * generator itself runs on the GPU (dna_score advances the state by the draws used). */
function makeSsbRng(seed = 0x12345678) {
  return { state: seed >>> 0 };
}

/** scoreDamageAt10keV (app.ts:167-207): direct-event SSB + combineHits + clusterDSB on
 * the GPU.  The indirect hits are empty on the worker path (app.ts:178-180) unless the
 * chemistry kept its OH at 1 us (runChemistryWorker options.ohFinal): then scoreIndirectSSB
 * runs first on those, as on the reference's GPU-chemistry path. */
async function scoreDamageAt10keV(dna, chem, radBuf, radN, totalDepositedEV, boxNm, rng, log = noop,
  kernelHits = 0) {
  const cc = ensurePipelines(cache ? cache.np : 0, log);
  if (!cc) throw new Error('scoreDamageAt10keV: no device');
  const withOh = chem && chem.chem_oh_n != null;
  const s = withOh ? await addon.scoreOh(cc.ctx, dna, radBuf, radN, null, chem.chem_oh_n, rng.state, false)
    : await addon.score(cc.ctx, dna, radBuf, radN, rng.state, false);
  const ind = withOh ? s.indirect : { ssb0: 0, ssb1: 0, inReach: 0 };
  rng.state = s.lcgState >>> 0;
  const ssb_dir = s.ssbCount, ssb_ind = ind.ssb0 + ind.ssb1, dsb = s.dsb;
  const doseBoxGy = (totalDepositedEV * 1.602e-19) / (Math.pow(2 * boxNm * 1e-9, 3) * 1000);
  const targetGbp = dna.n_bp * 1e-9;
  const ssbTot = ssb_dir + ssb_ind;
  log(`    DAMAGE: SSB_dir=${ssb_dir}  SSB_ind=${ssb_ind}  DSB=${dsb}  ` +
      `(DSB/SSB=${(ssbTot ? dsb / ssbTot : 0).toFixed(3)})  box_dose=${doseBoxGy.toFixed(3)} Gy  ` +
      `target=${(dna.n_bp / 1e6).toFixed(1)} Mbp  reach_dir=${s.inReach} reach_ind=${ind.inReach}  kernel_hits=${kernelHits}`, 'data');
  return {
    ssb_dir, ssb_ind, dsb, dsb_per_gy_gbp: doseBoxGy > 0 ? dsb / (doseBoxGy * targetGbp) : 0,
    dsb_expected: s.dsbExpected, in_reach: s.inReach,
  };
}

const r1 = (x) => Math.round(x * 10) / 10, r2 = (x) => Math.round(x * 100) / 100,
  r3 = (x) => Math.round(x * 1000) / 1000;

/** The layout of validation/webgpu-results.json:1-31. */
function resultsJson(r, damage, np) {
  const g = r.chem_result ? r.chem_result.timeline[r.chem_result.timeline.length - 1] : {};
  return {
    primaryEnergyEv: r.E, nPrimaries: np,
    tracking: {
      csdaNm: r1(r.mean_total), ionsPerPriPrimaryOnly: r1(r.mean_ions), secPerPri: r1(r.sec_per_pri),
      eConservationPct: r1(r.cons_ratio * 100),
    },
    preChemistry: { G_OH_init: r2(r.G_OH), G_eaq_init: r2(r.G_eaq), G_H_init: r2(r.G_H) },
    chemistry1Us: {
      G_OH: r3(g.G_OH || 0), G_eaq: r3(g.G_eaq || 0), G_H: r3(g.G_H || 0), G_H2O2: r3(g.G_H2O2 || 0),
      G_H2: r3(g.G_H2 || 0),
    },
    dnaDamage: {
      SSB_dir: damage ? damage.ssb_dir : 0, SSB_ind: damage ? damage.ssb_ind : 0, DSB: damage ? damage.dsb : 0,
      kernelHits: r.kernel_dna_hits,
    },
    karamitros2011: KARAMITROS_2011,
  };
}

/** runValidation (app.ts:91-161): the ESTAR sweep; chemistry and DNA scoring at 10 keV. */
async function runValidation(cfg) {
  const np = cfg.np, boxNm = cfg.boxNm, ceEV = cfg.ceEV != null ? cfg.ceEV : SOLVATION_THRESHOLD;
  const log = cfg.log || noop;
  const backend = cfg.chemBackend || 'worker';
  const energies = cfg.energies || ESTAR.map((e) => e.E);
  const cc = ensurePipelines(np, log);
  if (!cc) return null;
  const { device, buffers, pipelines } = cc;
  log(`Running ${np} primaries at ${energies.length} energies, box=${boxNm}nm, cutoff=${ceEV}eV`, 'data');
  const dna = buildDNATarget(DNA_LENGTH_NM, DNA_GRID_N, DNA_SPACING_NM);
  const ssbRng = makeSsbRng();
  const indirect = !!cfg.indirectSSB;     // extension: score the OH at 1 us too
  const chemCb = backend === 'none' ? null
    : (rad, radN, nTherm, E) => runChemistryWorker(rad, radN, nTherm, E, log, undefined, 1,
      { ohFinal: indirect && E === 10000 });
  const rows = [];
  let results = null;
  for (const E of energies) {
    const ref = ESTAR.find((e) => e.E === E) || { E, csda: NaN };
    const r = await runAtEnergy(device, buffers, pipelines, E, np, boxNm, ceEV, dna, chemCb);
    let damage = null;
    if (r.chem_result && r.rad_buf_final && E === 10000) {
      damage = await scoreDamageAt10keV(dna, r.chem_result, r.rad_buf_final, r.rad_n_stored,
        r.total_deposited_eV, boxNm, ssbRng, log, r.kernel_dna_hits);
    }
    log(`  E=${E} eV: CSDA=${r.mean_total.toFixed(1)}nm (${(r.mean_total / ref.csda).toFixed(2)}x) ` +
        `ions/pri=${r.mean_ions.toFixed(1)} sec/pri=${r.sec_per_pri.toFixed(1)} ` +
        `G_init(OH/eaq/H)=${r.G_OH.toFixed(2)}/${r.G_eaq.toFixed(2)}/${r.G_H.toFixed(2)} ` +
        `rad_n=${r.rad_n_stored} E_cons=${(r.cons_ratio * 100).toFixed(1)}%`, 'data');
    const row = Object.assign({}, r, { damage });
    delete row.rad_buf_final;
    delete row.dose_arr;
    rows.push(row);
    if (E === 10000) results = resultsJson(r, damage, np);
  }
  log('Validation run complete.', 'ok');
  return { rows, results };
}

module.exports = {
  addon, ESTAR, CHECKPOINTS_NS, CHECKPOINT_LABELS, RXN_LABELS, KARAMITROS_2011, SOLVATION_THRESHOLD,
  COPY_RESULTS, COPY_DOSE, COPY_RAD, COPY_ALL, OPT_IRT_BLOCK, OPT_IRT_BLOCK_LAST, OPT_IRT_FINAL, OPT_PA_LANES,
  buildDNATarget, primaryParams, ensurePipelines, runAtEnergy, runChemistryWorker, makeSsbRng, rxnTable,
  scoreDamageAt10keV, resultsJson, runValidation,
};
