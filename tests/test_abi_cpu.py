"""The C ABI library loads here (no GPU) and exports every function include/dna_amd.h
declares; the product package never reaches into oracle/."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "webgpu-dna_amd", "lib", "libdna_amd.so")


def declared_functions():
    text = open(os.path.join(ROOT, "include", "dna_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|void|char\s*\*|const char\s*\*)\s*\**\s*(dna_\w+)\s*\(",
                                 text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("dna_ctx_create", "dna_ctx_destroy", "dna_run_transport", "dna_run_irt",
                 "dna_run_irt_batch", "dna_score", "dna_hash_cells", "dna_last_error"):
        assert must in names, names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "run webgpu-dna_amd/build.py first"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    lib.dna_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.dna_version()


def test_no_gpu_is_reported_not_faked():
    """Without a device, context creation fails with DNA_ERR_NO_DEVICE (2) and a
    message; nothing falls back to CPU compute."""
    import torch
    if torch.cuda.is_available():
        return
    lib = ctypes.CDLL(LIB)
    h = ctypes.c_void_p()
    lib.dna_ctx_create.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]
    rc = lib.dna_ctx_create(0, None, ctypes.byref(h))
    assert rc == 2 and not h.value
    lib.dna_last_error.restype = ctypes.c_char_p
    assert lib.dna_last_error()


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "webgpu-dna_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cc", ".js")):
                src = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "import oracle" not in src and "liboracle" not in src, f
