"""Host-side physics constants and the DNA fiber target.

Mirrors src/physics/constants.ts and buildDNATarget (src/physics/dna-geometry.ts:16-68).
The fiber grid runs along X on a (y, z) lattice with 10.5 bp per turn, as in the
reference code (its README's "along z, 10 bp/turn" wording is not what the code does).
"""
import math

import numpy as np

# src/physics/constants.ts
NW = 33.4
BIND = (10.79, 13.39, 16.05, 32.30, 539.0)
EXC_E = (8.22, 10.0, 11.24, 12.61, 13.77)
VIB_LEV = (0.01, 0.024, 0.061, 0.092, 0.204, 0.417, 0.460, 0.500, 0.835)
SOLVATION_THRESHOLD = 7.4
MAX_SEC = 5_000_000
MAX_RAD = 16_000_000
VC = 128
MAX_SEC_STEPS = 2000
KARAMITROS_2011 = {"G_OH": 2.50, "G_eaq": 2.50, "G_H": 0.57, "G_H2O2": 0.73, "G_H2": 0.42}
DNA_LENGTH_NM = 3000
DNA_GRID_N = 21
DNA_SPACING_NM = 150
SSB_R_DAMAGE_NM = 0.29
SSB_P_INDIRECT = 0.4
SSB_P_DIRECT = 0.15
DSB_WINDOW_BP = 10

# ESTAR sweep of src/scoring/estar.ts:8-17 (energy eV, CSDA nm, stopping power eV/nm)
CHECKPOINTS_NS = (0.001, 0.01, 0.1, 1.0, 10.0, 100.0, 1000.0)    # irt-worker.js:416
CHECKPOINT_LABELS = ("1 ps", "10 ps", "100 ps", "1 ns", "10 ns", "100 ns", "1 us")


class DNATarget:
    """buildDNATarget(L_nm, grid_N, spacing_nm), dna-geometry.ts:16-68.

    fy/fz/rbb0/rbb1 are float32 arrays computed in double and rounded, as the
    reference fills Float32Arrays from double expressions."""

    def __init__(self, L_nm=3000.0, grid_N=21, spacing_nm=150.0):
        rise, bp_per_turn, r_bb = 0.34, 10.5, 1.0
        n_bp_per = math.floor(L_nm / rise)
        self.rise, self.r_bb, self.n_bp_per = rise, r_bb, n_bp_per
        self.x0 = -(n_bp_per - 1) * rise * 0.5
        self.n_fibers = grid_N * grid_N
        self.n_bp = self.n_fibers * n_bp_per
        self.grid_N, self.spacing_nm, self.L_nm = grid_N, spacing_nm, L_nm
        off = -((grid_N - 1) * spacing_nm) * 0.5
        fi, fj = np.meshgrid(np.arange(grid_N), np.arange(grid_N), indexing="ij")
        self.fy = (off + fi.ravel() * spacing_nm).astype(np.float32)
        self.fz = (off + fj.ravel() * spacing_nm).astype(np.float32)
        d_phase = (2 * math.pi) / bp_per_turn
        phi = np.arange(n_bp_per, dtype=np.float64) * d_phase
        self.rbb0 = np.empty(2 * n_bp_per, np.float32)
        self.rbb1 = np.empty(2 * n_bp_per, np.float32)
        self.rbb0[0::2] = r_bb * np.cos(phi)
        self.rbb0[1::2] = r_bb * np.sin(phi)
        self.rbb1[0::2] = r_bb * np.cos(phi + math.pi)
        self.rbb1[1::2] = r_bb * np.sin(phi + math.pi)

    def to_c(self):
        """dna_dna_target struct (keeps numpy arrays alive on the returned object)."""
        from .abi import DnaTarget
        t = DnaTarget()
        t.rise, t.r_bb, t.spacing_nm, t.x0, t.L_nm = (self.rise, self.r_bb, self.spacing_nm,
                                                       self.x0, self.L_nm)
        t.n_bp_per, t.n_fibers, t.n_bp, t.grid_N = (self.n_bp_per, self.n_fibers, self.n_bp,
                                                     self.grid_N)
        t.fy, t.fz = self.fy.ctypes.data, self.fz.ctypes.data
        t.rbb0, t.rbb1 = self.rbb0.ctypes.data, self.rbb1.ctypes.data
        t._keep = (self.fy, self.fz, self.rbb0, self.rbb1)
        return t


def default_params(E_eV, np_, box_nm=15000.0, ce_eV=7.4, dna=None, seed=None,
                   start_half=0.0, dir_mode=0, max_sec_steps=MAX_SEC_STEPS):
    """writePrimaryParams (src/gpu/dispatch.ts:22-41) as a DnaParams struct."""
    from .abi import DnaParams
    p = DnaParams()
    p.n, p.box, p.ce, p.ms, p.be = np_, box_nm, ce_eV, 65536, E_eV
    p.max_sec, p.vc, p.max_rad = 0, VC, 0
    p.start_half = start_half
    p.dna_enable = 1 if (dna is not None and E_eV == 10000) else 0
    p.dna_grid_n = dna.grid_N if dna is not None else 0
    p.dir_mode = dir_mode
    p.dna_rise = dna.rise if dna is not None else 0.0
    p.dna_spacing = dna.spacing_nm if dna is not None else 0.0
    p.dna_x0 = dna.x0 if dna is not None else 0.0
    p.dna_r_bb = dna.r_bb if dna is not None else 0.0
    p.seed = int(math.floor(E_eV)) if seed is None else int(seed)
    p.max_sec_steps = max_sec_steps
    return p
