"""WGDNA4D snapshot blob: the 4D viewer's input, fed by the IRT's checkpoints.

Mirrors the reference's exporter (src/splat/export.ts) and the viewer's parser
(src/splat/main.ts:38-96):
  pack_snapshot_blob(snapshots, dna, rad_buf_initial, rad_n_initial, energy_eV) -> bytes
                                                 export.ts:46-112 (packSnapshotBlob)
  parse_snapshot_blob(buf) -> dict              main.ts:40-96 (parseBlob)
  generate_snapshot_at_10kev(cfg) -> dict       export.ts:138-181 (generateSnapshotAt10keV)
  export_snapshot_at_10kev(cfg, path)           export.ts:216-225 (exportSnapshotAt10keV, a file
                                                 instead of a browser download)

Layout (little-endian throughout):
  header  32 B   "WGDNA4D\\0", u32 version = 1, u32 num_snaps, u32 snap_n, u32 n_fibers,
                 u32 n_bp_per, f32 energy_eV
  dna            u32 n_fibers, f32 fy[n_fibers], f32 fz[n_fibers], f32 rise, spacing_nm, L_nm, x0
  rad0           u32 rad_n_initial, f32 rad_buf[rad_n_initial][4]
  snaps          per snapshot: f32 t_ns, u32 label_len, label (UTF-8), f32 pos[snap_n][4],
                 u32 alive[snap_n]

The reference fills the snapshots from its time-stepped GPU chemistry (schedule.ts:55-246:
t = 0 after the pre-chemistry displacement, then the seven checkpoints from 1 ps to 1 us,
the leading snap_n entries of chem_pos with the species in w and alive 1/0).  Here they come
from the IRT: its per-particle state at the end (dna_irt_particles: position, rad_buf code,
birth and death time) gives, for every checkpoint t, the particles present at t.  A particle
is alive at checkpoint t when it was born before t and had not reacted before t
(t_birth < t <= t_death): the IRT counts a checkpoint before it handles the first event at or
after it, so these are exactly the particles of its species tallies.  At t = 0 the initial
particles (t_birth = 0) are alive.  The IRT moves no particle between its reactions, so a
slot keeps one position (the displaced initial position, or the reaction site for a product)
in every snapshot; products take slots of their own after their track's initial particles.
"""
import time

import numpy as np

MAGIC = b"WGDNA4D\x00"                       # export.ts:32
FORMAT_VERSION = 1                           # export.ts:33
# t = 0 and the seven checkpoints (schedule.ts:33-41 labels, irt-worker.js:416 times)
SNAPSHOT_TIMES_NS = (0.0, 0.001, 0.01, 0.1, 1.0, 10.0, 100.0, 1000.0)
SNAPSHOT_LABELS = ("t=0", "1 ps", "10 ps", "100 ps", "1 ns", "10 ns", "100 ns", "1 μs")
DEFAULT_SNAPSHOT_N = 50_000                  # export.ts:140


def snapshots_from_particles(particles, snapshot_n=DEFAULT_SNAPSHOT_N, t_ns=SNAPSHOT_TIMES_NS,
                             labels=SNAPSHOT_LABELS):
    """Snapshots (label, t_ns, n, pos [n, 4] float32, alive [n] uint32) of the leading
    min(snapshot_n, N) particle slots of an IRT particle table ([N, 6] float64: x, y, z,
    pid*8+code, t_birth, t_death).  pos.w is the rad_buf species code (pid removed), as
    the reference's chem_pos holds it (chemistry.wgsl:90-98)."""
    p = np.asarray(particles, np.float64).reshape(-1, 6)
    if len(t_ns) != len(labels):
        raise ValueError("one label per snapshot time")
    n = min(int(snapshot_n), p.shape[0])
    q = p[:n]
    pos = np.empty((n, 4), np.float32)
    pos[:, :3] = q[:, :3]
    pos[:, 3] = np.mod(q[:, 3], 8.0)
    tb, td = q[:, 4], q[:, 5]
    out = []
    for lab, t in zip(labels, t_ns):
        born = (tb < t) | ((t == 0.0) & (tb == 0.0))
        alive = (born & (t <= td)).astype(np.uint32)
        out.append({"label": lab, "t_ns": float(t), "n": n, "pos": pos.copy(), "alive": alive})
    return out


def pack_snapshot_blob(snapshots, dna, rad_buf_initial, rad_n_initial, energy_eV):
    """packSnapshotBlob (export.ts:46-112): the blob as bytes."""
    if len(snapshots) == 0:
        raise ValueError("no snapshots to pack")
    snap_n = int(snapshots[0]["n"])
    rad0 = np.ascontiguousarray(rad_buf_initial, "<f4").reshape(-1)
    if rad0.size < 4 * rad_n_initial:
        raise ValueError("rad_buf_initial holds fewer than rad_n_initial entries")
    parts = [MAGIC, np.array([FORMAT_VERSION, len(snapshots), snap_n, dna.n_fibers, dna.n_bp_per], "<u4").tobytes(),
             np.array([energy_eV], "<f4").tobytes(),
             np.array([dna.n_fibers], "<u4").tobytes(),
             np.ascontiguousarray(dna.fy, "<f4").tobytes(), np.ascontiguousarray(dna.fz, "<f4").tobytes(),
             np.array([dna.rise, dna.spacing_nm, dna.L_nm, dna.x0], "<f4").tobytes(),
             np.array([rad_n_initial], "<u4").tobytes(), rad0[:4 * rad_n_initial].tobytes()]
    for s in snapshots:
        if int(s["n"]) != snap_n:
            raise ValueError("every snapshot holds snap_n entries")
        lab = s["label"].encode("utf-8")
        pos = np.ascontiguousarray(s["pos"], "<f4").reshape(-1)
        alive = np.ascontiguousarray(s["alive"], "<u4").reshape(-1)
        if pos.size != 4 * snap_n or alive.size != snap_n:
            raise ValueError("snapshot arrays do not match snap_n")
        parts += [np.array([s["t_ns"]], "<f4").tobytes(), np.array([len(lab)], "<u4").tobytes(), lab,
                  pos.tobytes(), alive.tobytes()]
    return b"".join(parts)


def parse_snapshot_blob(buf):
    """parseBlob (main.ts:40-96): the blob's fields as numpy arrays."""
    b = memoryview(bytes(buf))
    if bytes(b[:8]) != MAGIC:
        raise ValueError(f"bad magic: {bytes(b[:8])!r}")
    u32 = lambda o: int(np.frombuffer(b, "<u4", 1, o)[0])
    f32 = lambda o: float(np.frombuffer(b, "<f4", 1, o)[0])
    version = u32(8)
    if version != 1:
        raise ValueError(f"unsupported version {version}")
    num_snaps, snap_n, n_fibers, n_bp_per = u32(12), u32(16), u32(20), u32(24)
    energy_eV = f32(28)
    off = 32
    if u32(off) != n_fibers:
        raise ValueError("dna.n_fibers mismatch")
    off += 4
    fy = np.frombuffer(b, "<f4", n_fibers, off).copy(); off += 4 * n_fibers
    fz = np.frombuffer(b, "<f4", n_fibers, off).copy(); off += 4 * n_fibers
    rise, spacing_nm, L_nm, x0 = (f32(off + 4 * k) for k in range(4)); off += 16
    rad_n_initial = u32(off); off += 4
    rad_buf = np.frombuffer(b, "<f4", 4 * rad_n_initial, off).reshape(-1, 4).copy(); off += 16 * rad_n_initial
    snaps = []
    for _ in range(num_snaps):
        t_ns = f32(off); ll = u32(off + 4); off += 8
        label = bytes(b[off:off + ll]).decode("utf-8"); off += ll
        pos = np.frombuffer(b, "<f4", 4 * snap_n, off).reshape(-1, 4).copy(); off += 16 * snap_n
        alive = np.frombuffer(b, "<u4", snap_n, off).copy(); off += 4 * snap_n
        snaps.append({"label": label, "t_ns": t_ns, "pos": pos, "alive": alive})
    if off != len(b):
        raise ValueError(f"blob size mismatch: read {off} of {len(b)} bytes")
    return {"version": version, "num_snaps": num_snaps, "snap_n": snap_n, "energy_eV": energy_eV,
            "dna": {"n_fibers": n_fibers, "n_bp_per": n_bp_per, "fy": fy, "fz": fz, "rise": rise,
                    "spacing_nm": spacing_nm, "L_nm": L_nm, "x0": x0},
            "rad0": {"rad_n_initial": rad_n_initial, "rad_buf": rad_buf}, "snapshots": snaps}


def generate_snapshot_at_10kev(cfg):
    """generateSnapshotAt10keV (export.ts:138-181): one 10 keV pass (transport, then the
    IRT with the snapshot capture on) and its blob.  cfg: np, boxNm, ceEV, log,
    snapshot_n (default 50 000).  Returns None (and an 'err' log line) without a device."""
    from . import host
    from .geometry import DNA_GRID_N, DNA_LENGTH_NM, DNA_SPACING_NM, DNATarget
    np_, box_nm, ce_eV = cfg["np"], cfg["boxNm"], cfg["ceEV"]
    log = cfg.get("log") or host._noop
    snapshot_n = int(cfg.get("snapshot_n", DEFAULT_SNAPSHOT_N))
    cc = host.ensure_pipelines(np_, log)
    if cc is None:
        log("GPU init failed, cannot generate snapshot.", "err")
        return None
    log(f"Snapshot capture: {np_} primaries @ 10 keV, snap_n={snapshot_n}", "data")
    dna = DNATarget(DNA_LENGTH_NM, DNA_GRID_N, DNA_SPACING_NM)

    def chem(rad, rad_n, n_therm, E_eV):
        return host.run_chemistry_worker(rad, rad_n, n_therm, E_eV,
                                         options={"dump_snapshots": True, "snapshot_n": snapshot_n})

    r = host.run_at_energy(cc["device"], cc["buffers"], cc["pipelines"], 10000, np_, box_nm, ce_eV, dna,
                           run_chemistry=chem)
    chem_r = r["chem_result"]
    if not chem_r or not chem_r.get("snapshots") or r["rad_buf_final"] is None:
        log("Snapshot capture: no chemistry result returned.", "err")
        return None
    initial_n = min(r["rad_n_stored"], snapshot_n)
    blob = pack_snapshot_blob(chem_r["snapshots"], dna, r["rad_buf_final"], initial_n, 10000)
    stamp = time.strftime("%Y-%m-%dT%H-%M-%S", time.gmtime())
    return {"blob": blob, "filename": f"wgdna4d-{np_}p-{snapshot_n}n-{stamp}.bin",
            "sizeMB": f"{len(blob) / 1024 / 1024:.2f}", "numCheckpoints": len(chem_r["snapshots"])}


def export_snapshot_at_10kev(cfg, path=None):
    """exportSnapshotAt10keV (export.ts:216-225): generate and write the blob to path (the
    generated file name in the working directory by default).  Returns the path or None."""
    r = generate_snapshot_at_10kev(cfg)
    if r is None:
        return None
    path = path or r["filename"]
    with open(path, "wb") as f:
        f.write(r["blob"])
    log = cfg.get("log")
    if log:
        log(f"Snapshot export: wrote {path} ({r['sizeMB']} MB, {r['numCheckpoints']} checkpoints)", "ok")
    return path
