'use strict';
/**
 * splat.js — the WGDNA4D snapshot blob for the reference's 4D viewer, fed by the IRT's
 * checkpoints on the MI355X library.  CommonJS for Node >= 12.
 *
 * Same names, argument meaning and layout as the reference:
 *   packSnapshotBlob(snapshots, dna, radBufInitial, radNInitial, energyEV) -> Buffer
 *                                                src/splat/export.ts:46-112 (a Buffer, not a Blob)
 *   parseSnapshotBlob(buf) -> object             src/splat/main.ts:40-96 (parseBlob)
 *   generateSnapshotAt10keV(cfg) -> {blob, filename, sizeMB, numCheckpoints} | null
 *                                                src/splat/export.ts:138-181
 *   exportSnapshotAt10keV(cfg, path?) -> path    src/splat/export.ts:216-225 (a file instead of a
 *                                                browser download)
 * Layout (little-endian): 32-byte header ("WGDNA4D\0", u32 version 1, num_snaps, snap_n,
 * n_fibers, n_bp_per, f32 energy_eV); DNA geometry (u32 n_fibers, f32 fy[], fz[], rise,
 * spacing_nm, L_nm, x0); u32 rad_n_initial + f32 rad_buf[rad_n_initial][4]; per snapshot
 * f32 t_ns, u32 label_len, UTF-8 label, f32 pos[snap_n][4], u32 alive[snap_n].
 *
 * The snapshots come from the IRT's particle table (irtParticles: x, y, z, pid*8+code,
 * t_birth, t_death per slot).  A slot is alive at checkpoint t when t_birth < t <= t_death
 * (the IRT counts a checkpoint before the first event at or after it, so these are the
 * particles of its tallies); at t = 0 the initial particles are.  The IRT moves no particle
 * between reactions, so a slot keeps one position in every snapshot.  pos.w is the rad_buf
 * species code without the pid, as the reference's chem_pos holds it.
 */
const fs = require('fs');
const host = require('./dna_host');

const MAGIC = Buffer.from([0x57, 0x47, 0x44, 0x4e, 0x41, 0x34, 0x44, 0x00]); // export.ts:32
const FORMAT_VERSION = 1;                                                      // export.ts:33
// t = 0 and the seven checkpoints (schedule.ts:33-41 labels, irt-worker.js:416 times)
const SNAPSHOT_TIMES_NS = [0.0, 0.001, 0.01, 0.1, 1.0, 10.0, 100.0, 1000.0];
const SNAPSHOT_LABELS = ['t=0', '1 ps', '10 ps', '100 ps', '1 ns', '10 ns', '100 ns', '1 μs'];
const DEFAULT_SNAPSHOT_N = 50000;                                              // export.ts:140

/** Snapshots {label, t_ns, n, pos: Float32Array(4n), alive: Uint32Array(n)} of the leading
 * min(snapshotN, N) slots of a particle table (Float64Array of N*6). */
function snapshotsFromParticles(particles, snapshotN = DEFAULT_SNAPSHOT_N, tNs = SNAPSHOT_TIMES_NS,
  labels = SNAPSHOT_LABELS) {
  if (tNs.length !== labels.length) throw new Error('one label per snapshot time');
  const n = Math.min(snapshotN, Math.floor(particles.length / 6));
  const pos = new Float32Array(4 * n);
  for (let i = 0; i < n; i++) {
    pos[4 * i] = particles[6 * i];
    pos[4 * i + 1] = particles[6 * i + 1];
    pos[4 * i + 2] = particles[6 * i + 2];
    pos[4 * i + 3] = particles[6 * i + 3] % 8;
  }
  return labels.map((label, k) => {
    const t = tNs[k];
    const alive = new Uint32Array(n);
    for (let i = 0; i < n; i++) {
      const tb = particles[6 * i + 4], td = particles[6 * i + 5];
      const born = tb < t || (t === 0 && tb === 0);
      alive[i] = born && t <= td ? 1 : 0;
    }
    return { label, t_ns: t, n, pos: pos.slice(), alive };
  });
}

/** packSnapshotBlob (export.ts:46-112). */
function packSnapshotBlob(snapshots, dna, radBufInitial, radNInitial, energyEV) {
  if (snapshots.length === 0) throw new Error('no snapshots to pack');
  const snapN = snapshots[0].n;
  const numSnaps = snapshots.length;
  const labelBytes = snapshots.map((s) => Buffer.from(s.label, 'utf8'));
  let total = 32 + 4 + dna.fy.byteLength + dna.fz.byteLength + 16 + 4 + radNInitial * 16;
  for (let i = 0; i < numSnaps; i++) total += 8 + labelBytes[i].length + snapN * 20;
  const buf = Buffer.alloc(total);
  let off = 0;
  MAGIC.copy(buf, off); off += 8;
  buf.writeUInt32LE(FORMAT_VERSION, off); off += 4;
  buf.writeUInt32LE(numSnaps, off); off += 4;
  buf.writeUInt32LE(snapN, off); off += 4;
  buf.writeUInt32LE(dna.n_fibers, off); off += 4;
  buf.writeUInt32LE(dna.n_bp_per, off); off += 4;
  buf.writeFloatLE(energyEV, off); off += 4;
  buf.writeUInt32LE(dna.n_fibers, off); off += 4;
  const raw = (a, bytes) => Buffer.from(a.buffer, a.byteOffset, bytes != null ? bytes : a.byteLength);
  raw(dna.fy).copy(buf, off); off += dna.fy.byteLength;
  raw(dna.fz).copy(buf, off); off += dna.fz.byteLength;
  buf.writeFloatLE(dna.rise, off); off += 4;
  buf.writeFloatLE(dna.spacing_nm, off); off += 4;
  buf.writeFloatLE(dna.L_nm, off); off += 4;
  buf.writeFloatLE(dna.x0, off); off += 4;
  if (radBufInitial.length < 4 * radNInitial) throw new Error('radBufInitial holds fewer than radNInitial entries');
  buf.writeUInt32LE(radNInitial, off); off += 4;
  raw(radBufInitial, radNInitial * 16).copy(buf, off); off += radNInitial * 16;
  for (let i = 0; i < numSnaps; i++) {
    const s = snapshots[i];
    if (s.n !== snapN || s.pos.length !== 4 * snapN || s.alive.length !== snapN) {
      throw new Error('snapshot arrays do not match snap_n');
    }
    buf.writeFloatLE(s.t_ns, off); off += 4;
    buf.writeUInt32LE(labelBytes[i].length, off); off += 4;
    labelBytes[i].copy(buf, off); off += labelBytes[i].length;
    raw(s.pos).copy(buf, off); off += s.pos.byteLength;
    raw(s.alive).copy(buf, off); off += s.alive.byteLength;
  }
  if (off !== total) throw new Error(`blob size mismatch: wrote ${off}, expected ${total}`);
  return buf;
}

/** parseBlob (main.ts:40-96). */
function parseSnapshotBlob(buf) {
  let off = 0;
  if (!buf.subarray(0, 8).equals(MAGIC)) throw new Error(`bad magic: ${JSON.stringify(buf.subarray(0, 8).toString('latin1'))}`);
  off += 8;
  const version = buf.readUInt32LE(off); off += 4;
  if (version !== 1) throw new Error(`unsupported version ${version}`);
  const num_snaps = buf.readUInt32LE(off); off += 4;
  const snap_n = buf.readUInt32LE(off); off += 4;
  const n_fibers = buf.readUInt32LE(off); off += 4;
  const n_bp_per = buf.readUInt32LE(off); off += 4;
  const energy_eV = buf.readFloatLE(off); off += 4;
  if (buf.readUInt32LE(off) !== n_fibers) throw new Error('dna.n_fibers mismatch');
  off += 4;
  const f32 = (n) => { const a = new Float32Array(n); Buffer.from(a.buffer).set(buf.subarray(off, off + 4 * n)); off += 4 * n; return a; };
  const u32 = (n) => { const a = new Uint32Array(n); Buffer.from(a.buffer).set(buf.subarray(off, off + 4 * n)); off += 4 * n; return a; };
  const fy = f32(n_fibers), fz = f32(n_fibers);
  const [rise, spacing_nm, L_nm, x0] = f32(4);
  const rad_n_initial = buf.readUInt32LE(off); off += 4;
  const rad_buf = f32(4 * rad_n_initial);
  const snapshots = [];
  for (let i = 0; i < num_snaps; i++) {
    const t_ns = buf.readFloatLE(off); off += 4;
    const len = buf.readUInt32LE(off); off += 4;
    const label = buf.subarray(off, off + len).toString('utf8'); off += len;
    const pos = f32(4 * snap_n), alive = u32(snap_n);
    snapshots.push({ label, t_ns, pos, alive });
  }
  if (off !== buf.length) throw new Error(`blob size mismatch: read ${off} of ${buf.length} bytes`);
  return {
    version, num_snaps, snap_n, energy_eV,
    dna: { n_fibers, n_bp_per, fy, fz, rise, spacing_nm, L_nm, x0 },
    rad0: { rad_n_initial, rad_buf }, snapshots,
  };
}

/** generateSnapshotAt10keV (export.ts:138-181): one 10 keV pass (transport, then the IRT
 * with the snapshot capture on) and its blob; null (and an 'err' log line) without a device. */
async function generateSnapshotAt10keV(cfg) {
  const { np, boxNm, ceEV } = cfg;
  const log = cfg.log || (() => {});
  const snapshotN = cfg.snapshot_n != null ? cfg.snapshot_n : DEFAULT_SNAPSHOT_N;
  const cc = host.ensurePipelines(np, log);
  if (!cc) {
    log('GPU init failed, cannot generate snapshot.', 'err');
    return null;
  }
  log(`Snapshot capture: ${np} primaries @ 10 keV, snap_n=${snapshotN}`, 'data');
  const dna = host.buildDNATarget();
  const chem = (rad, radN, nTherm, E) => host.runChemistryWorker(rad, radN, nTherm, E, undefined, undefined, 1,
    { dump_snapshots: true, snapshot_n: snapshotN });
  const r = await host.runAtEnergy(cc.device, cc.buffers, cc.pipelines, 10000, np, boxNm, ceEV, dna, chem);
  if (!r.chem_result || !r.chem_result.snapshots || !r.rad_buf_final) {
    log('Snapshot capture: no chemistry result returned.', 'err');
    return null;
  }
  const initialN = Math.min(r.rad_n_stored, snapshotN);
  const blob = packSnapshotBlob(r.chem_result.snapshots, dna, r.rad_buf_final, initialN, 10000);
  const stamp = new Date().toISOString().replace(/[:.]/g, '-').slice(0, 19);
  return {
    blob, filename: `wgdna4d-${np}p-${snapshotN}n-${stamp}.bin`,
    sizeMB: (blob.length / 1024 / 1024).toFixed(2), numCheckpoints: r.chem_result.snapshots.length,
  };
}

/** exportSnapshotAt10keV (export.ts:216-225): generate and write the blob to path (the
 * generated file name in the working directory by default). */
async function exportSnapshotAt10keV(cfg, path) {
  const r = await generateSnapshotAt10keV(cfg);
  if (!r) return null;
  const out = path || r.filename;
  fs.writeFileSync(out, r.blob);
  if (cfg.log) cfg.log(`Snapshot export: wrote ${out} (${r.sizeMB} MB, ${r.numCheckpoints} checkpoints)`, 'ok');
  return out;
}

module.exports = {
  MAGIC, FORMAT_VERSION, SNAPSHOT_TIMES_NS, SNAPSHOT_LABELS, DEFAULT_SNAPSHOT_N,
  snapshotsFromParticles, packSnapshotBlob, parseSnapshotBlob, generateSnapshotAt10keV, exportSnapshotAt10keV,
};
