"""WGDNA4D snapshot blob (the 4D viewer's input) on the CPU.

* Byte layout: a blob written field by field with struct from the layout of
  packSnapshotBlob (export.ts:36-112: 32-byte header, DNA geometry, initial rad_buf,
  per-snapshot t_ns / label / pos / alive) equals pack_snapshot_blob's bytes, and the
  total size follows export.ts:61-68.
* Round trip: parse_snapshot_blob (parseBlob, main.ts:40-96) gives back every field,
  labels with non-ASCII characters included; bad magic, bad version and a truncated blob
  raise.
* Snapshots from the IRT's particle table (t_birth < t <= t_death, t = 0: the initial
  particles): per-species alive counts at every checkpoint equal the IRT's own tallies
  (OH, e-aq, H, H2O2) on a 3 keV track set of the oracle, and t = 0 counts every initial
  particle.
"""
import struct

import numpy as np
import pytest

import oracle as O
from dna_amd import DNATarget, default_params
from dna_amd.splat import (MAGIC, SNAPSHOT_LABELS, SNAPSHOT_TIMES_NS, pack_snapshot_blob,
                           parse_snapshot_blob, snapshots_from_particles)


class SmallDNA:
    """A 2 x 2 fiber grid (the blob only reads these fields)."""
    n_fibers, n_bp_per = 4, 7
    fy = np.array([-1.5, -1.5, 1.5, 1.5], np.float32)
    fz = np.array([-1.5, 1.5, -1.5, 1.5], np.float32)
    rise, spacing_nm, L_nm, x0 = 0.34, 3.0, 2.04, -1.02


def snaps(n, k, rng):
    labs = ["t=0", "1 μs", "x"][:k]
    return [{"label": labs[i], "t_ns": [0.0, 1000.0, 2.5][i], "n": n,
             "pos": rng.normal(0, 10, (n, 4)).astype(np.float32),
             "alive": rng.integers(0, 2, n).astype(np.uint32)} for i in range(k)]


def test_layout_matches_field_by_field_packing():
    rng = np.random.default_rng(1)
    dna = SmallDNA()
    rad0 = rng.normal(0, 5, (9, 4)).astype(np.float32)
    ss = snaps(5, 3, rng)
    blob = pack_snapshot_blob(ss, dna, rad0, 6, 10000)
    ref = bytearray(b"WGDNA4D\x00")
    ref += struct.pack("<5If", 1, 3, 5, 4, 7, 10000.0)
    assert len(ref) == 32
    ref += struct.pack("<I", 4) + struct.pack("<4f", *dna.fy) + struct.pack("<4f", *dna.fz)
    ref += struct.pack("<4f", dna.rise, dna.spacing_nm, dna.L_nm, dna.x0)
    ref += struct.pack("<I", 6) + struct.pack("<24f", *rad0[:6].ravel())
    for s in ss:
        lab = s["label"].encode("utf-8")
        ref += struct.pack("<fI", s["t_ns"], len(lab)) + lab
        ref += struct.pack("<20f", *s["pos"].ravel()) + struct.pack("<5I", *s["alive"])
    assert blob == bytes(ref)
    labels = sum(len(s["label"].encode("utf-8")) for s in ss)
    assert len(blob) == 32 + 4 + 2 * 4 * 4 + 16 + 4 + 6 * 16 + 3 * (8 + 5 * 20) + labels


def test_round_trip():
    rng = np.random.default_rng(2)
    dna = DNATarget()
    rad0 = rng.normal(0, 5, (100, 4)).astype(np.float32)
    ss = snaps(1000, 3, rng)
    b = parse_snapshot_blob(pack_snapshot_blob(ss, dna, rad0, 100, 10000))
    assert (b["version"], b["num_snaps"], b["snap_n"], b["energy_eV"]) == (1, 3, 1000, 10000.0)
    d = b["dna"]
    assert (d["n_fibers"], d["n_bp_per"]) == (dna.n_fibers, dna.n_bp_per)
    assert np.array_equal(d["fy"], dna.fy) and np.array_equal(d["fz"], dna.fz)
    for k in ("rise", "spacing_nm", "L_nm", "x0"):
        assert d[k] == np.float32(getattr(dna, k))
    assert b["rad0"]["rad_n_initial"] == 100 and np.array_equal(b["rad0"]["rad_buf"], rad0)
    for s, r in zip(ss, b["snapshots"]):
        assert r["label"] == s["label"] and r["t_ns"] == np.float32(s["t_ns"])
        assert np.array_equal(r["pos"], s["pos"]) and np.array_equal(r["alive"], s["alive"])


def test_errors():
    rng = np.random.default_rng(3)
    blob = pack_snapshot_blob(snaps(4, 2, rng), SmallDNA(), np.zeros((2, 4), np.float32), 2, 1000)
    with pytest.raises(ValueError, match="magic"):
        parse_snapshot_blob(b"XGDNA4D\x00" + blob[8:])
    with pytest.raises(ValueError, match="version"):
        parse_snapshot_blob(blob[:8] + struct.pack("<I", 2) + blob[12:])
    with pytest.raises(ValueError):
        parse_snapshot_blob(blob[:-1])
    with pytest.raises(ValueError, match="no snapshots"):
        pack_snapshot_blob([], SmallDNA(), np.zeros((0, 4), np.float32), 0, 1000)
    assert MAGIC == bytes([0x57, 0x47, 0x44, 0x4E, 0x41, 0x34, 0x44, 0x00])


def test_snapshots_reproduce_the_irt_tallies():
    p = default_params(3000.0, 24)
    t = O.transport(p)
    n_therm = int((t["results"]["esc"] == 0).sum())
    ck = list(SNAPSHOT_TIMES_NS[1:])
    o, parts = O.irt_particles(t["rad"], n_therm, 3000.0, ck, seed=4, n_threads=4)
    ss = snapshots_from_particles(parts, snapshot_n=10 ** 9)
    assert [s["label"] for s in ss] == list(SNAPSHOT_LABELS) and ss[0]["n"] == parts.shape[0]
    code = ss[0]["pos"][:, 3]
    n_init = int((parts[:, 4] == 0).sum())
    assert ss[0]["alive"].sum() == n_init > 0
    for k, s in enumerate(ss[1:]):
        got = [int(s["alive"][code == c].sum()) for c in (0, 1, 2, 4)]
        assert got == [int(o.tally[k][j]) for j in range(4)], s["label"]
    # the leading slots only
    few = snapshots_from_particles(parts, snapshot_n=100)
    assert few[3]["n"] == 100 and np.array_equal(few[3]["alive"], ss[3]["alive"][:100])
    assert np.array_equal(few[3]["pos"][:, :3], parts[:100, :3].astype(np.float32))
