"""ctypes mirror of include/dna_amd.h (struct layouts of the C ABI)."""
import ctypes as C

DNA_OK = 0
DNA_MAX_CKPT = 32
DNA_N_RXN = 9
DNA_COPY_RESULTS, DNA_COPY_DOSE, DNA_COPY_RAD, DNA_COPY_ALL = 1, 2, 4, 7
DNA_OPT_IRT_BLOCK, DNA_OPT_IRT_BLOCK_LAST, DNA_OPT_IRT_FINAL, DNA_OPT_PA_LANES = 1, 2, 3, 4
DNA_OPT_IRT_CELL, DNA_OPT_IRT_CELL_LAST = 5, 6
DNA_OPT_RAD_CHUNK, DNA_OPT_WORK_CHUNK = 7, 8
DNA_OPT_PA_LANES_LAST, DNA_OPT_RAD_CHUNK_LAST, DNA_OPT_WORK_CHUNK_LAST = 9, 10, 11
DNA_OPT_MLP_FUSED = 12
DNA_OPT_IRT_DRAIN = 13
DNA_OPT_IRT_ORDER = 14
DNA_OPT_IRT_EVBLOCKS = 15


class DnaParams(C.Structure):
    """Uniform P of primary.wgsl:1-7 (packed at src/gpu/dispatch.ts:22-41) + seed/step cap."""
    _fields_ = [
        ("n", C.c_uint32), ("box", C.c_float), ("ce", C.c_float), ("ms", C.c_uint32),
        ("be", C.c_float), ("max_sec", C.c_uint32), ("vc", C.c_uint32), ("max_rad", C.c_uint32),
        ("start_half", C.c_float), ("dna_enable", C.c_uint32), ("dna_grid_n", C.c_uint32),
        ("dir_mode", C.c_uint32),
        ("dna_rise", C.c_float), ("dna_spacing", C.c_float), ("dna_x0", C.c_float),
        ("dna_r_bb", C.c_float), ("seed", C.c_uint32), ("max_sec_steps", C.c_uint32),
    ]


class DnaResult(C.Structure):
    _fields_ = [("path", C.c_float), ("prod", C.c_float), ("fE", C.c_float),
                ("ni", C.c_uint32), ("nx", C.c_uint32), ("esc", C.c_uint32),
                ("mx", C.c_float), ("_pad", C.c_uint32)]


class DnaTransportOut(C.Structure):
    _fields_ = [
        ("results", C.c_void_p), ("dose", C.c_void_p), ("rad_buf", C.c_void_p),
        ("counters", C.c_uint32 * 8), ("sec_stats", C.c_uint32 * 8), ("dbg", C.c_uint32 * 8),
        ("dose_sum", C.c_uint64),
        ("rad_n_raw", C.c_uint64), ("rad_n_stored", C.c_uint64), ("rad_dropped", C.c_uint64),
        ("sec_n_raw", C.c_uint64), ("sec_n", C.c_uint64), ("sec_dropped", C.c_uint64),
        ("t_prim_ms", C.c_double), ("t_sec_ms", C.c_double), ("t_total_ms", C.c_double),
    ]


class DnaChemOut(C.Structure):
    _fields_ = [
        ("n_ckpt", C.c_int), ("t_ckpt", C.c_double * DNA_MAX_CKPT),
        ("tally", (C.c_uint64 * 5) * DNA_MAX_CKPT), ("G", (C.c_double * 5) * DNA_MAX_CKPT),
        ("rxn_counts", C.c_uint64 * DNA_N_RXN), ("n_reacted", C.c_uint64),
        ("chem_n", C.c_uint64), ("n_tracks", C.c_uint64), ("n_events", C.c_uint64),
        ("t_ms", C.c_double), ("t_kernel_ms", C.c_double),
        ("n_pair_cand", C.c_uint64), ("n_pair_sampled", C.c_uint64), ("n_repair", C.c_uint64),
        ("algo_flops", C.c_double), ("prof", C.c_uint64 * 8),
        ("n_pair_all", C.c_uint64), ("n_repair_scanned", C.c_uint64),
    ]


class DnaTarget(C.Structure):
    _fields_ = [
        ("rise", C.c_double), ("r_bb", C.c_double), ("spacing_nm", C.c_double),
        ("x0", C.c_double), ("L_nm", C.c_double),
        ("n_bp_per", C.c_uint32), ("n_fibers", C.c_uint32), ("n_bp", C.c_uint32),
        ("grid_N", C.c_uint32),
        ("fy", C.c_void_p), ("fz", C.c_void_p), ("rbb0", C.c_void_p), ("rbb1", C.c_void_p),
    ]


class DnaScoreOut(C.Structure):
    _fields_ = [
        ("ssb_count", C.c_uint64), ("candidates", C.c_uint64), ("in_reach", C.c_uint64),
        ("dsb", C.c_uint64), ("ssb0", C.c_uint64), ("ssb1", C.c_uint64),
        ("dsb_expected", C.c_double), ("hits", C.c_void_p), ("t_ms", C.c_double),
    ]


class DnaIndirectOut(C.Structure):
    _fields_ = [("ssb0", C.c_uint64), ("ssb1", C.c_uint64), ("candidates", C.c_uint64),
                ("in_reach", C.c_uint64)]


def chem_out_to_dict(o: DnaChemOut) -> dict:
    nc = o.n_ckpt
    return {
        "t_ckpt": [o.t_ckpt[c] for c in range(nc)],
        "tally": [[int(o.tally[c][s]) for s in range(5)] for c in range(nc)],
        "G": [[o.G[c][s] for s in range(5)] for c in range(nc)],
        "rxn_counts": [int(o.rxn_counts[r]) for r in range(DNA_N_RXN)],
        "n_reacted": int(o.n_reacted), "chem_n": int(o.chem_n),
        "n_tracks": int(o.n_tracks), "n_events": int(o.n_events), "t_ms": o.t_ms,
    }
