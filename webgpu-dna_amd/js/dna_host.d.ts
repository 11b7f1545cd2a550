// Type declarations of dna_host.js: the shapes of the reference's src/physics/types.ts
// (EnergyResult :86-119, ChemResult :65-74, DNATarget :23-37) and of the addon's results.

export interface DNATarget {
  rise: number; r_bb: number; spacing_nm: number; x0: number; L_nm: number;
  n_bp_per: number; n_fibers: number; n_bp: number; grid_N: number;
  fy: Float32Array; fz: Float32Array; rbb0: Float32Array; rbb1: Float32Array;
}

export interface ChemCheckpoint {
  label: string; t_ns: number; G_OH: number; G_eaq: number; G_H: number; G_H2O2: number; G_H2: number;
}

// src/physics/types.ts:57-63
export interface ChemSnapshot {
  label: string; t_ns: number; n: number; pos: Float32Array; alive: Uint32Array;
}

export interface ChemResult {
  chem_n: number; t_wall: number; timeline: ChemCheckpoint[];
  chem_pos_final: Float32Array | null; chem_alive_final: Uint32Array | null;
  n_reacted: number; rxn_info: { label: string; count: number }[];
  /** with options.ohFinal: the OH alive at 1 us kept on the device for scoreOh */
  chem_oh_n?: number;
  /** with options.dump_snapshots: t = 0 and the seven checkpoints (splat.js) */
  snapshots?: ChemSnapshot[];
}

export interface ChemOptions { ohFinal?: boolean; dump_snapshots?: boolean; snapshot_n?: number; }

export interface EnergyResult {
  E: number; n_therm: number; n_esc: number; mean_total: number; mean_prod: number; mean_sp: number;
  mean_ions: number; dt: number; t_prim: number; t_sec: number; sec_n: number; sec_dropped: number;
  sec_per_pri: number; sec_terminated_cutoff: number; sec_terminated_bounds: number; sec_steps: number;
  sec_tertiary_ions: number; total_deposited_eV: number; expected_deposited_eV: number; cons_ratio: number;
  rad_OH: number; rad_eaq: number; rad_H: number; G_OH: number; G_eaq: number; G_H: number;
  rad_n_raw: number; rad_n_stored: number; rad_dropped: number; kernel_dna_hits: number;
  chem_result: ChemResult | null; rad_buf_final: Float32Array | null; dose_arr: Uint32Array | null;
}

export type LogFn = (msg: string, cls?: string) => void;
export type ChemCallback = (radBuf: Float32Array, radN: number, nTherm: number, E_eV: number) =>
  Promise<ChemResult | null>;
/** the addon context in place of a GPUDevice */
export interface DnaDevice { ctx: object; }
/** app.ts PipeCache {device, buffers, pipelines, np}; ctx = device.ctx */
export interface PipeCache { device: DnaDevice; buffers: { np: number }; pipelines: object; np: number; ctx: object; }
export interface SsbRng { state: number; }
export interface DamageResult {
  ssb_dir: number; ssb_ind: number; dsb: number; dsb_per_gy_gbp: number; dsb_expected: number; in_reach: number;
}
export interface ValidationConfig {
  np: number; boxNm: number; ceEV?: number; log?: LogFn; chemBackend?: 'worker' | 'none';
  energies?: number[];
  /** extension: score the OH alive at 1 us (indirect SSB) before the direct pass */
  indirectSSB?: boolean;
}

export function buildDNATarget(L_nm?: number, grid_N?: number, spacing_nm?: number): DNATarget;
export function ensurePipelines(np: number, log?: LogFn): PipeCache | null;
export function runAtEnergy(device: DnaDevice, buffers: { np: number }, pipelines: object, E_eV: number, np: number,
  boxNm: number, ceEV: number, dna: DNATarget | null, runChemistry?: ChemCallback | null,
  seed?: number): Promise<EnergyResult>;
export function runChemistryWorker(radBuf: Float32Array, radN: number, nTherm: number, E_eV: number,
  log?: LogFn, maxN?: number, seed?: number, options?: ChemOptions | null): Promise<ChemResult>;
export function makeSsbRng(seed?: number): SsbRng;
export function scoreDamageAt10keV(dna: DNATarget, chem: ChemResult | null,
  radBuf: Float32Array, radN: number, totalDepositedEV: number, boxNm: number, rng: SsbRng, log?: LogFn,
  kernelHits?: number): Promise<DamageResult>;
export function runValidation(cfg: ValidationConfig): Promise<{ rows: object[]; results: object | null } | null>;
