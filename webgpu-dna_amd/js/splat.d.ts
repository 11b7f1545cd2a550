// Type declarations of splat.js: the WGDNA4D blob of src/splat/export.ts and main.ts.
import { ChemSnapshot, DNATarget, LogFn } from './dna_host';

export const SNAPSHOT_TIMES_NS: number[];
export const SNAPSHOT_LABELS: string[];
export const DEFAULT_SNAPSHOT_N: number;

export interface ExportConfig { np: number; boxNm: number; ceEV: number; log?: LogFn; snapshot_n?: number; }
export interface SnapshotResult { blob: Buffer; filename: string; sizeMB: string; numCheckpoints: number; }
export interface Blob4D {
  version: number; num_snaps: number; snap_n: number; energy_eV: number;
  dna: { n_fibers: number; n_bp_per: number; fy: Float32Array; fz: Float32Array; rise: number;
    spacing_nm: number; L_nm: number; x0: number; };
  rad0: { rad_n_initial: number; rad_buf: Float32Array };
  snapshots: { label: string; t_ns: number; pos: Float32Array; alive: Uint32Array }[];
}

export function snapshotsFromParticles(particles: Float64Array, snapshotN?: number, tNs?: number[],
  labels?: string[]): ChemSnapshot[];
export function packSnapshotBlob(snapshots: ChemSnapshot[], dna: DNATarget, radBufInitial: Float32Array,
  radNInitial: number, energyEV: number): Buffer;
export function parseSnapshotBlob(buf: Buffer): Blob4D;
export function generateSnapshotAt10keV(cfg: ExportConfig): Promise<SnapshotResult | null>;
export function exportSnapshotAt10keV(cfg: ExportConfig, path?: string): Promise<string | null>;
