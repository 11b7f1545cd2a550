"""The transport step's table lookups evaluated on the GPU (dna_xs_probe: the device code
and LDS tables of k_phase_a/k_phase_b) against the CPU restatement, bit for bit: xs_all,
shell and excitation-level fractions, vibrational total, DEA (helpers.wgsl:15-186) at the
grid energies, the clamp ends, the bin edges of the small tables and random energies, and
the elastic cos(theta) of xs_el_cos (:66-82) at random (E, r)."""
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "webgpu-dna_amd", "data")
MAN = json.load(open(os.path.join(DATA, "cross_sections.json")))
BIN = np.fromfile(os.path.join(DATA, "cross_sections.bin"), dtype="<f4")


def tab(name):
    t = MAN["tables"][name]
    return BIN[t["offset"] // 4: t["offset"] // 4 + t["count"]]


def energies():
    rng = np.random.default_rng(11)
    XE, XVE = tab("XE"), tab("XVE")
    pts = [XE, np.nextafter(XE, np.float32(0)), np.nextafter(XE, np.float32(np.inf)), XVE,
           np.array([0.5, 2.0, 4.0, 4.2, 7.4, 12.99, 13.0, 100.0, 100.01, 1e6], np.float32),
           np.exp(rng.uniform(np.log(0.5), np.log(1e6), 20000)).astype(np.float32)]
    return np.concatenate([np.asarray(p, np.float32) for p in pts])


@pytest.mark.parametrize("which", [0, 1, 2, 3, 4])
def test_lookup_bit_exact(ctx, which):
    E = energies()
    print(f"xs probe which={which} on {E.size} energies", flush=True)
    g = ctx.xs_probe(which, E)
    o = O.xs(which, E)
    assert g.shape == o.shape
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), which


def test_elastic_cos_bit_exact(ctx):
    rng = np.random.default_rng(12)
    E = np.exp(rng.uniform(np.log(5.0), np.log(1e5), 4000)).astype(np.float32)
    r = rng.random(4000, dtype=np.float32)
    r[:4] = [0.0, 1.0, 0.5, np.float32(1.0) - np.float32(2 ** -24)]
    print("xs probe elastic cos", flush=True)
    g = ctx.xs_probe(5, E, r)[:, 0]
    L = O.lib()
    o = np.array([L.oracle_xs_el_cos(float(e), float(x)) for e, x in zip(E, r)], np.float32)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
