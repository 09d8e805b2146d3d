"""The C-ABI library loads and exports every symbol include/multifil_amd.h
declares; ctypes mirrors match the C struct layouts; pure-host entry points
answer without a GPU."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "multifil_amd.h")


def _declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mf_[a-z_]+)\s*\(", txt)))


def test_library_exports_all_declared_symbols():
    from multifil_jax_amd import _lib
    lib = _lib.load()
    declared = _declared_functions()
    assert set(declared) == set(_lib.EXPORTED)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.mf_abi_version() == _lib.MF_ABI_VERSION


def test_host_entry_points():
    from multifil_jax_amd import _lib
    from multifil_jax_amd.engine import state_layout
    from multifil_jax_amd.topology import SarcTopology
    lib = _lib.load()
    t = SarcTopology.create(2, 2)
    assert lib.mf_state_bytes(t.n_thick, t.n_crowns, t.n_thin, t.n_sites, t.n_xb_per_crown) \
        == state_layout(t)["bytes"]
    assert lib.mf_table_floats(200) == 16 + 27 * 16 + 400 * 48 + 400 * 4
    # errors are reported, not raised across the boundary
    L = _lib.MfLaunch()
    L.block = 100
    assert lib.mf_run(ctypes.byref(L), None) != 0
    assert b"block" in lib.mf_last_error()


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_struct_layouts_match_c(tmp_path):
    from multifil_jax_amd import _lib
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "multifil_amd.h"\n'
                   'int main(){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(mf_lattice), '
                   'sizeof(mf_instance), sizeof(mf_solver_cfg), sizeof(mf_launch), '
                   'offsetof(mf_launch, cfg), offsetof(mf_launch, lattices_host), '
                   'offsetof(mf_launch, pop_params), offsetof(mf_launch, pop_labels), '
                   'offsetof(mf_instance, label_off));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [ctypes.sizeof(_lib.MfLattice), ctypes.sizeof(_lib.MfInstance),
            ctypes.sizeof(_lib.MfSolverCfg), ctypes.sizeof(_lib.MfLaunch),
            _lib.MfLaunch.cfg.offset, _lib.MfLaunch.lattices_host.offset,
            _lib.MfLaunch.pop_params.offset, _lib.MfLaunch.pop_labels.offset, _lib.MfInstance.label_off.offset]
    assert got == want


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from multifil_jax_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load()


def test_subpopulation_launch_validation():
    """Bad subpopulation descriptors are rejected on the host with a message
    (mf_run_lds_bytes runs the same checks as mf_run, without a GPU)."""
    from multifil_jax_amd import _lib
    from multifil_jax_amd.engine import LatticePack
    from multifil_jax_amd.topology import SarcTopology
    lib = _lib.load()
    pack = LatticePack([SarcTopology.create(2, 2)])
    params = (ctypes.c_float * 8)()
    labels = (ctypes.c_uint8 * 8)()

    def launch(**kw):
        L = _lib.MfLaunch()
        L.n_lattices, L.B, L.block = 1, 1, 768
        L.cfg.mode, L.cfg.n_cg = _lib.MF_SOLVER_REFINED, 6
        L.lattices_host = ctypes.addressof(pack.desc_host)
        L.n_pops = 1
        for k, v in kw.items():
            setattr(L, k, v)
        out = ctypes.c_int64(0)
        rc = lib.mf_run_lds_bytes(ctypes.byref(L), ctypes.byref(out))
        return rc, lib.mf_last_error().decode(), out.value

    assert launch()[0] == 0
    rc, msg, _ = launch(pop_mode=7, n_pops=2)
    assert rc != 0 and "pop_mode" in msg
    rc, msg, _ = launch(pop_mode=_lib.MF_POP_MEAN_FIELD, n_pops=_lib.MF_MAX_POPS + 1)
    assert rc != 0 and "n_pops" in msg
    rc, msg, _ = launch(pop_mode=_lib.MF_POP_MEAN_FIELD, n_pops=2)
    assert rc != 0 and "pop_params" in msg
    rc, msg, _ = launch(pop_mode=_lib.MF_POP_MEAN_FIELD, n_pops=2, pop_params=ctypes.addressof(params))
    assert rc != 0 and "pop_fractions" in msg
    rc, msg, _ = launch(pop_mode=_lib.MF_POP_EXPLICIT, n_pops=2, pop_flags=_lib.MF_POP_XB,
                        pop_params=ctypes.addressof(params))
    assert rc != 0 and "pop_labels" in msg
    # explicit TM populations widen the TM rows in LDS by one 27 x 16 block each
    rc0, _, lds1 = launch()
    rc2, _, lds2 = launch(pop_mode=_lib.MF_POP_EXPLICIT, n_pops=2, pop_flags=_lib.MF_POP_TM,
                          pop_params=ctypes.addressof(params), pop_labels=ctypes.addressof(labels))
    assert rc0 == 0 and rc2 == 0 and lds2 - lds1 >= 27 * 16 * 4
    assert lib.mf_table_floats_pop(200, 1, 1) == lib.mf_table_floats(200)
    assert lib.mf_table_floats_pop(200, 2, 3) == 16 + 2 * 27 * 16 + 3 * (400 * 48 + 400 * 4)
    assert lib.mf_table_floats_pop(200, 0, 1) < 0


def test_noalloc_rejects_launches_beyond_its_scratch(skeletal):
    """mf_run_noalloc checks B x workspace against its static array before
    touching the device: a 6x6 launch of 400 instances (~85 MB of
    workspace) is refused with an error, 280 fit (host-side check only)."""
    from multifil_jax_amd import _lib
    from multifil_jax_amd.engine import LatticePack, SolverConfig
    from multifil_jax_amd.topology import SarcTopology
    s, d = skeletal
    lib = _lib.load()
    pack = LatticePack([SarcTopology.create(6, 6, s, d)])
    L = _lib.MfLaunch()
    L.n_lattices, L.block, L.n_pops = 1, 1024, 1
    L.cfg = SolverConfig(mode="refined").ctype()
    L.lattices_host = ctypes.addressof(pack.desc_host)
    per = ctypes.c_int64(0)
    L.B = 400
    assert lib.mf_workspace_bytes(ctypes.byref(L), ctypes.byref(per)) == 0 and per.value > 0
    stride = (per.value + 255) // 256 * 256
    assert stride * 400 > _lib.MF_NOALLOC_SCRATCH_BYTES >= stride * 280
    assert lib.mf_run_noalloc(ctypes.byref(L), None) != 0
    assert b"scratch" in lib.mf_last_error()
