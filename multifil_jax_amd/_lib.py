"""ctypes binding of the gfx950 engine (include/multifil_amd.h).

Loads ``libmultifil_amd.so`` from this package directory.  There is no CPU
fallback: if the library is missing or fails to load, importing the engine
raises, so a run can never silently use another implementation.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmultifil_amd.so")

MF_ABI_VERSION = 5
MF_N_PARAMS = 44
MF_N_METRICS = 52
MF_MAX_BINS = 256
MF_MAX_POPS = 4
MF_SOLVER_REFERENCE = 0
MF_SOLVER_REFINED = 1
MF_INST_INIT = 1
MF_INST_DYNAMIC = 2
MF_LAUNCH_DYNAMIC = 1
MF_LAUNCH_KINETICS_ONLY = 2
MF_LAUNCH_SOLVE_ONLY = 4
MF_LAUNCH_REBUILD = 8
MF_POP_NONE = 0
MF_POP_MEAN_FIELD = 1
MF_POP_EXPLICIT = 2
MF_POP_XB = 1
MF_POP_TM = 2

EXPORTED = ("mf_abi_version", "mf_last_error", "mf_state_bytes", "mf_table_floats",
            "mf_table_floats_pop", "mf_tables", "mf_run", "mf_run_lds_bytes", "mf_workspace_bytes",
            "mf_run_noalloc")
# bytes of the library's static scratch array (mf_run_noalloc)
MF_NOALLOC_SCRATCH_BYTES = 64 << 20


class MfLattice(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("n_thick", "n_crowns", "n_thin", "n_sites",
                                         "n_xb_per_crown", "n_faces", "width", "n_bins")] + \
               [(n, C.c_int64) for n in ("crown_offsets", "crown_rests", "binding_offsets",
                                         "binding_rests", "thin_rest_cum", "bin_edges",
                                         "bin_centers", "xb_thin", "xb_slot", "xb_valid",
                                         "face_sites", "face_count", "tm_prev", "tm_next")] + \
               [("overlap_lo", C.c_float), ("overlap_hi", C.c_float)]


class MfInstance(C.Structure):
    _fields_ = [("lattice", C.c_int32), ("n_steps", C.c_int32), ("t0", C.c_int32),
                ("flags", C.c_int32), ("seed", C.c_uint64), ("state_off", C.c_int64),
                ("table_off", C.c_int64), ("K_lat", C.c_float), ("nu", C.c_float),
                ("label_off", C.c_int64)]


class MfSolverCfg(C.Structure):
    _fields_ = [("mode", C.c_int32), ("n_newton", C.c_int32), ("n_cg", C.c_int32),
                ("n_refine", C.c_int32), ("tol", C.c_float), ("refine_tol", C.c_float),
                ("dt", C.c_float), ("cg_forcing", C.c_float)]


class MfLaunch(C.Structure):
    _fields_ = [("lattices", C.c_void_p), ("lat_f32", C.c_void_p), ("lat_i32", C.c_void_p),
                ("instances", C.c_void_p), ("params", C.c_void_p), ("drivers", C.c_void_p),
                ("uniforms", C.c_void_p), ("state", C.c_void_p), ("tables", C.c_void_p),
                ("metrics", C.c_void_p), ("force", C.c_void_p),
                ("n_lattices", C.c_int32), ("B", C.c_int32), ("t_stride", C.c_int32),
                ("u_stride", C.c_int32), ("cfg", MfSolverCfg), ("block", C.c_int32),
                ("t_base", C.c_int32), ("launch_flags", C.c_int32), ("pad", C.c_int32),
                ("lattices_host", C.c_void_p), ("workspace", C.c_void_p),
                ("workspace_stride", C.c_int64),
                ("n_pops", C.c_int32), ("pop_mode", C.c_int32), ("pop_flags", C.c_int32),
                ("pad2", C.c_int32), ("pop_params", C.c_void_p), ("pop_fractions", C.c_void_p),
                ("pop_labels", C.c_void_p)]


_lib = None


def load() -> C.CDLL:
    """Load the engine library once; raise if it is absent or mismatched."""
    global _lib
    if _lib is not None:
        return _lib
    # PyTorch ships its own HIP runtime under the same soname
    # (libamdhip64.so.7); loading torch first makes the engine bind to that
    # copy, so device buffers and streams come from a single runtime.
    import torch  # noqa: F401
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
            "g.build()'` (hipcc, gfx950). There is no CPU fallback.")
    lib = C.CDLL(LIB_PATH)
    lib.mf_abi_version.restype = C.c_int
    lib.mf_last_error.restype = C.c_char_p
    lib.mf_state_bytes.restype = C.c_int64
    lib.mf_state_bytes.argtypes = [C.c_int32] * 5
    lib.mf_table_floats.restype = C.c_int64
    lib.mf_table_floats.argtypes = [C.c_int32]
    lib.mf_table_floats_pop.restype = C.c_int64
    lib.mf_table_floats_pop.argtypes = [C.c_int32] * 3
    for fn in ("mf_tables", "mf_run", "mf_run_noalloc"):
        getattr(lib, fn).restype = C.c_int
        getattr(lib, fn).argtypes = [C.POINTER(MfLaunch), C.c_void_p]
    for fn in ("mf_run_lds_bytes", "mf_workspace_bytes"):
        getattr(lib, fn).restype = C.c_int
        getattr(lib, fn).argtypes = [C.POINTER(MfLaunch), C.POINTER(C.c_int64)]
    if lib.mf_abi_version() != MF_ABI_VERSION:
        raise RuntimeError(f"ABI mismatch: library {lib.mf_abi_version()} != {MF_ABI_VERSION}")
    _lib = lib
    return lib


def check(rc: int, what: str):
    if rc != 0:
        msg = load().mf_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed: {msg}")
