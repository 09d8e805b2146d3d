"""ctypes front end of ``cpu_port.c`` (TEST INFRASTRUCTURE ONLY): the C
restatement of the step used as the timed CPU baseline.  Built by
``oracle/Makefile`` into ``oracle/libmfc.so``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import mf_oracle as mo

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "libmfc.so")

_FIELDS = ("thick_k", "thin_k", "titin_a", "titin_b", "titin_rest", "g_k_w", "g_rest_w", "c_k_w",
           "c_rest_w", "g_k_s", "g_rest_s", "c_k_s", "c_rest_s", "solver_tol", "z_line0", "ls0")


class MfcParams(C.Structure):
    _fields_ = [(n, C.c_float) for n in _FIELDS]


_lib = None
_lib_path = None


def build_native(dest_dir: str) -> str:
    """Compile cpu_port.c for the host it runs on (-O3 -march=native, OpenMP)
    into dest_dir; returns the library path.  The in-tree libmfc.so is built
    for a portable x86-64-v2 target instead."""
    import subprocess
    out = os.path.join(dest_dir, "libmfc_native.so")
    cmd = ["gcc", "-O3", "-march=native", "-fopenmp", "-fPIC", "-std=c11", "-shared",
           "-I", os.path.join(os.path.dirname(_HERE), "include"), os.path.join(_HERE, "cpu_port.c"),
           "-o", out, "-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


def load(path: str = None):
    """Load the port (the in-tree build, or ``path``); a later call with
    another path switches libraries."""
    global _lib, _lib_path
    path = path or LIB
    if _lib is None or _lib_path != path:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run `make -C oracle`")
        _lib = C.CDLL(path)
        _lib.mfc_run.restype = C.c_int
        _lib_path = path
    return _lib


def params_struct(p: dict) -> MfcParams:
    """The spring / titin / solver constants the port reads, from a
    DynamicParams-style dict (the XB constants under their reference
    names)."""
    s = MfcParams()
    s.thick_k, s.thin_k = p['thick_k'], p['thin_k']
    s.titin_a, s.titin_b, s.titin_rest = p['titin_a'], p['titin_b'], p['titin_rest']
    gk, gr, ck, cr = mo._xb_springs(p, np.array([False, True]))
    s.g_k_w, s.g_rest_w, s.c_k_w, s.c_rest_w = float(gk[0]), float(gr[0]), float(ck[0]), float(cr[0])
    s.g_k_s, s.g_rest_s, s.c_k_s, s.c_rest_s = float(gk[1]), float(gr[1]), float(ck[1]), float(cr[1])
    s.solver_tol = p['solver_tol']
    s.z_line0, s.ls0 = p['z_line'], p['lattice_spacing']
    return s


def tables(p: dict, pCa: float, ls: float, dt: float, topo):
    """FP32 cumulative rows: TM (27*16) and XB (2 n_bins * 36)."""
    T = mo.make_tables(p, pCa, ls, dt, topo.xb_bin_centers)
    tm = np.cumsum(T.P_tm, axis=2).astype(np.float32).reshape(-1)
    xb = np.cumsum(T.P_xb, axis=2).astype(np.float32).reshape(-1)
    return tm, xb


def prepare(topo, p: dict, drivers, dt=1.0):
    """Per-instance transition tables (from each instance's first pCa and
    spacing) and parameter structs: everything run() needs besides the
    drivers, so a timed run measures the steps alone."""
    from multifil_jax_amd.engine import LatticePack
    drivers = np.asarray(drivers, np.float32)
    B = drivers.shape[0]
    cache = {}
    tm = np.empty((B, 27 * 16), np.float32)
    xb = np.empty((B, 2 * topo.n_xb_bins * 36), np.float32)
    for b in range(B):
        key = (float(drivers[b, 0, 1]), float(drivers[b, 0, 2]))
        if key not in cache:
            cache[key] = tables(p, key[0], key[1], dt, topo)
        tm[b], xb[b] = cache[key]
    P = (MfcParams * B)(*[params_struct(p) for _ in range(B)])
    return dict(pack=LatticePack([topo]), tm=tm, xb=xb, P=P)


def run(topo, p: dict, drivers, seeds, dt=1.0, n_newton=4, n_cg=6, threads=1, prepared=None, lib_path=None):
    """B instances of one lattice with constant pCa / spacing per instance.

    drivers: (B, T, 3) float32 (z_line, pCa, spacing); seeds (B,) uint64.
    Returns force (B, T) float64, residual (B, T), iters (B, T), n_bound (B, T).
    """
    lib = load(lib_path)
    drivers = np.ascontiguousarray(drivers, np.float32)
    B, T, _ = drivers.shape
    pr = prepared or prepare(topo, p, drivers, dt)
    pack, tm, xb, P = pr['pack'], pr['tm'], pr['xb'], pr['P']
    seeds = np.ascontiguousarray(seeds, np.uint64)
    force = np.zeros((B, T), np.float64)
    resid = np.zeros((B, T), np.float32)
    iters = np.zeros((B, T), np.int32)
    nbound = np.zeros((B, T), np.int32)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = lib.mfc_run(C.c_void_p(C.addressof(pack.desc_host)), ptr(pack.f32), ptr(pack.i32), P, ptr(tm), ptr(xb),
                     C.c_int(B), C.c_int(T), ptr(drivers), ptr(seeds), C.c_int(n_newton), C.c_int(n_cg),
                     C.c_int(threads), ptr(force), ptr(resid), ptr(iters), ptr(nbound))
    if rc != 0:
        raise RuntimeError("mfc_run failed")
    return force, resid, iters, nbound
