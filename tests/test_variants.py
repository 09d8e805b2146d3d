"""Host-side variant selection: every lattice the workloads use has a kernel
variant that fits LDS and the register slots at the engine's default block
(no GPU needed: mf_run_lds_bytes is host logic)."""
import ctypes as C

import pytest

from multifil_jax_amd import _lib
from multifil_jax_amd.engine import LatticePack, default_block
from multifil_jax_amd.topology import SarcTopology

LDS_PER_CU = 160 * 1024


@pytest.fixture(scope="module")
def lib():
    try:
        return _lib.load()
    except (RuntimeError, OSError) as e:
        pytest.skip(f"engine library not built: {e}")


def _lds(lib, topos, mode, block, dynamic=False, ws=False):
    pack = LatticePack(topos)
    L = _lib.MfLaunch()
    L.launch_flags = _lib.MF_LAUNCH_DYNAMIC if dynamic else 0
    L.n_lattices = len(topos)
    L.B = 1
    L.block = block
    L.cfg.mode = mode
    L.cfg.n_cg = 6
    L.lattices_host = C.addressof(pack.desc_host)
    out = C.c_int64(0)
    fn = lib.mf_workspace_bytes if ws else lib.mf_run_lds_bytes
    rc = fn(C.byref(L), C.byref(out))
    return rc, out.value


@pytest.mark.parametrize("dynamic", [False, True])
@pytest.mark.parametrize("n", [2, 3, 4, 6])
@pytest.mark.parametrize("mode", [_lib.MF_SOLVER_REFERENCE, _lib.MF_SOLVER_REFINED])
def test_variant_fits(lib, skeletal, n, mode, dynamic):
    s, d = skeletal
    topo = SarcTopology.create(n, n, s, d)
    blk = default_block([topo])
    rc, lds = _lds(lib, [topo], mode, blk, dynamic)
    assert rc == 0, lib.mf_last_error().decode()
    assert 0 < lds <= LDS_PER_CU
    rc, ws = _lds(lib, [topo], mode, blk, dynamic, ws=True)
    assert rc == 0
    # only the 6x6 lattice needs its per-XB / per-site arrays in device memory
    assert (ws > 0) == (n == 6)


@pytest.mark.parametrize("dynamic", [False, True])
@pytest.mark.parametrize("mode", [_lib.MF_SOLVER_REFERENCE, _lib.MF_SOLVER_REFINED])
@pytest.mark.parametrize("preset", ["lethocerus", "drosophila"])
def test_invertebrate_variant_fits(lib, preset, mode, dynamic):
    # 4 heads per crown, 80 candidates per head, 160 sites per strand
    from multifil_jax_amd.params import get_lethocerus_params, get_drosophila_params
    fn, shape = {"lethocerus": (get_lethocerus_params, (2, 2)),
                 "drosophila": (get_drosophila_params, (2, 3))}[preset]
    s, d = fn()
    topo = SarcTopology.create(*shape, s, d)
    assert topo.n_xb_per_crown == 4
    rc, lds = _lds(lib, [topo], mode, default_block([topo]), dynamic)
    assert rc == 0, lib.mf_last_error().decode()
    assert 0 < lds <= LDS_PER_CU


def test_mixed_lattices_share_a_variant(lib, skeletal):
    s, d = skeletal
    topos = [SarcTopology.create(n, n, s, d) for n in (2, 3, 4)]
    rc, lds = _lds(lib, topos, _lib.MF_SOLVER_REFINED, default_block(topos))
    assert rc == 0, lib.mf_last_error().decode()
    assert lds <= LDS_PER_CU


def test_bad_block_is_rejected(lib, skeletal):
    s, d = skeletal
    topo = SarcTopology.create(2, 2, s, d)
    rc, _ = _lds(lib, [topo], _lib.MF_SOLVER_REFINED, 640)
    assert rc != 0


@pytest.mark.parametrize("dynamic", [False, True])
@pytest.mark.parametrize("mode", [_lib.MF_SOLVER_REFERENCE, _lib.MF_SOLVER_REFINED])
@pytest.mark.parametrize("case", ["8x8", "10x10", "lethocerus4x4", "drosophila4x6"])
def test_large_lattice_variant_fits(lib, case, mode, dynamic):
    """The lattices above 6x6 of the reference's dynamic-spacing benchmark
    (examples/benchmarks/benchmark_dynamic_ls.py:136-137: up to 10x10) and
    the larger flight-muscle lattices launch on the level-2 workspace
    variants: positions and per-head state in device memory, only the
    atomically updated site arrays in LDS.  10x10 with dynamic spacing is
    refused (its kernels are switched off)."""
    from multifil_jax_amd.params import get_skeletal_params, get_lethocerus_params, get_drosophila_params
    fn, shape = {"8x8": (get_skeletal_params, (8, 8)), "10x10": (get_skeletal_params, (10, 10)),
                 "lethocerus4x4": (get_lethocerus_params, (4, 4)),
                 "drosophila4x6": (get_drosophila_params, (4, 6))}[case]
    s, d = fn()
    topo = SarcTopology.create(*shape, s, d)
    assert default_block([topo]) == 1024
    rc, lds = _lds(lib, [topo], mode, 1024, dynamic)
    if case == "10x10" and dynamic:
        # the 32-slot dynamic-spacing kernels are switched off: refused on
        # the host with a message that says so
        assert rc != 0 and "dynamic spacing is not available" in lib.mf_last_error().decode()
        return
    assert rc == 0, lib.mf_last_error().decode()
    assert 0 < lds <= LDS_PER_CU
    rc, ws = _lds(lib, [topo], mode, 1024, dynamic, ws=True)
    assert rc == 0 and ws > (8 if mode == _lib.MF_SOLVER_REFINED else 4) * (
        topo.n_thick * topo.n_crowns + topo.n_thin * topo.n_sites)   # positions in the workspace
