"""Structure of the lattice builder (sarc_geometry.py:379-645)."""
import numpy as np
import pytest

from multifil_jax_amd.params import (StaticParams, get_skeletal_params, get_lethocerus_params,
                                     get_drosophila_params, DynamicParams, DYNAMIC_FIELDS)
from multifil_jax_amd.topology import SarcTopology, xb_bin_edges


@pytest.mark.parametrize("n,thick,thin,xbs,dof", [(2, 4, 8, 624, 928), (4, 16, 32, 2496, 3712),
                                                  (6, 36, 72, 5616, 8352)])
def test_counts(skeletal, n, thick, thin, xbs, dof):
    # params.py:106-108 (90 nodes per thin), docs/README.md:52-54,147-148
    s, d = skeletal
    t = SarcTopology.create(n, n, s, d)
    assert (t.n_thick, t.n_thin, t.total_xbs, t.n_dof) == (thick, thin, xbs, dof)
    assert t.n_sites == 90 and t.max_sites_per_face == 30
    assert np.all(t.n_sites_per_face == 30) and t.xb_valid.all()
    assert t.n_titin == 6 * thick


def test_crown_offsets_and_rests(topo22):
    i = np.arange(52, dtype=np.float32)
    assert np.allclose(topo22.crown_offsets, 58.0 + i * np.float32(14.3))
    assert np.all(topo22.crown_rests[:, 0] == 58.0)
    assert np.allclose(topo22.crown_rests[:, 1:], 14.3)


def test_binding_sites(topo22):
    t = topo22
    # offsets strictly decreasing (positions increasing along the thin filament)
    assert np.all(np.diff(t.binding_offsets, axis=1) < 0)
    # rests: consecutive gaps, last one = distance to the Z-line
    gaps = t.binding_offsets[:, :-1] - t.binding_offsets[:, 1:]
    assert np.allclose(t.binding_rests[:, :-1], gaps)
    assert np.allclose(t.binding_rests[:, -1], t.binding_offsets[:, -1])
    # offsets are multiples of the monomer rise 72/26 nm, 1..390 monomers
    k = t.binding_offsets / np.float32(72.0 / 26)
    assert np.allclose(k, np.round(k), atol=1e-3) and k.max() <= 390
    # each face holds two monomers per 26-monomer repeat
    for i in range(t.n_thin):
        for f in range(t.n_faces_per_thin):
            sites = t.face_to_sites[i, f]
            assert len(set(sites.tolist())) == 30
    all_sites = np.sort(t.face_to_sites.reshape(t.n_thin, -1), axis=1)
    assert np.all(all_sites == np.arange(90)[None, :])


def test_tm_neighbours(topo22):
    t = topo22
    own = np.arange(t.n_sites)[None, :]
    prev, nxt = t.tm_prev_neighbor, t.tm_next_neighbor
    rows = np.arange(t.n_thin)[:, None]
    # symmetric: next(prev(i)) == i where prev is real
    real = prev != own
    assert np.all(nxt[rows, prev][real] == own.repeat(t.n_thin, 0)[real])
    # same chain, and exactly one start/end per chain per filament
    assert np.all(t.tm_chains[rows, prev] == t.tm_chains)
    for i in range(t.n_thin):
        for c in (0, 1):
            idx = np.nonzero(t.tm_chains[i] == c)[0]
            assert (prev[i, idx] == idx).sum() == 1 and (nxt[i, idx] == idx).sum() == 1


def test_xb_targets_face_pattern(topo22):
    t = topo22
    # every thick face is used, and each XB's candidates come from its face
    assert np.all(t.thick_to_thin[:, :, 0] >= 0)
    for j in range(0, t.total_xbs, 7):
        row = t.face_to_sites[t.xb_to_thin_id[j], t.xb_to_thin_face[j]]
        assert np.array_equal(t.xb_to_site_indices[j], row)


def test_bin_edges():
    e = xb_bin_edges(-8.0, 35.0, 200)
    assert e.dtype == np.float32 and e.shape == (201,)
    assert e[0] == np.float32(-8.0) and e[-1] == np.float32(35.0)
    assert np.max(np.abs(e.astype(np.float64) - np.linspace(-8, 35, 201))) < 1e-5
    assert np.all(np.diff(e) > 0)


def test_invertebrate_and_errors():
    s, d = get_lethocerus_params()
    t = SarcTopology.create(2, 2, s, d)
    assert t.n_xb_per_crown == 4 and t.n_faces_per_thin == 2 and t.n_crowns == 100
    assert 0.3 < t.xb_valid.mean() < 0.6        # partial coverage at 33.75 deg
    with pytest.raises(ValueError):
        SarcTopology.create(3, 2, s, d)         # invertebrate needs even nrows
    s3, d3 = get_drosophila_params()
    with pytest.raises(ValueError):
        SarcTopology.create(2, 2, s3, d3)        # superlattice needs ncols % 3 == 0
    t3 = SarcTopology.create(2, 3, s3, d3)
    assert len(np.unique(t3.crown_offsets[:, 0])) == 3
    with pytest.raises(ValueError):
        StaticParams(n_superlattice_classes=2)
    with pytest.raises(ValueError):
        StaticParams(legacy_crown_geometry=True, n_xb_per_crown=4)


def test_start_overrides(skeletal):
    s, d = skeletal
    with pytest.raises(ValueError):
        SarcTopology.create(2, 2, s, d, thin_starts=[0, 1])
    t = SarcTopology.create(2, 2, s, d, thin_starts=[0] * 8, thick_starts=[1, 1, 1, 1])
    assert np.all(t.thin_starts == 0)


def test_dynamic_params_api():
    d = DynamicParams()
    assert len(DYNAMIC_FIELDS) == 44
    with pytest.raises(ValueError):
        d.copy(not_a_field=1.0)
    e = d.copy(thick_k=9000.0)
    assert float(e.thick_k) == 9000.0 and float(d.thick_k) == 7500.0
    w = d.with_drivers(5.0, 1000.0, 13.0)
    assert float(w.pCa) == 5.0 and float(w.z_line) == 1000.0
    rows = DynamicParams(thick_k=[7000.0, 8000.0]).rows(2)
    assert rows.shape == (2, 44) and rows[1, 0] == 8000.0


def test_save_load_round_trip(tmp_path, skeletal):
    """The .npz format round-trips every field bit for bit, and a loaded
    lattice packs to the same engine blobs."""
    from multifil_jax_amd.params import get_lethocerus_params
    from multifil_jax_amd.engine import LatticePack
    s, d = skeletal
    ls, ld = get_lethocerus_params()
    for t in (SarcTopology.create(2, 2, s, d), SarcTopology.create(2, 2, ls, ld)):
        path = tmp_path / "lattice.npz"
        t.save(path)
        u = SarcTopology.load(path)
        for k in SarcTopology.INT_FIELDS:
            assert getattr(u, k) == getattr(t, k)
        for k in SarcTopology.ARRAY_FIELDS:
            a, b = np.asarray(getattr(t, k)), np.asarray(getattr(u, k))
            assert a.dtype == b.dtype and np.array_equal(a, b), k
        assert (u.nrows, u.ncols, u.periodic) == (t.nrows, t.ncols, t.periodic)
        p1, p2 = LatticePack([t]), LatticePack([u])
        assert np.array_equal(p1.f32, p2.f32) and np.array_equal(p1.i32, p2.i32)
        assert np.array_equal(p1.desc_bytes, p2.desc_bytes)


def test_load_rejects_bad_files(tmp_path, skeletal):
    s, d = skeletal
    t = SarcTopology.create(2, 2, s, d)
    bad = tmp_path / "other.npz"
    np.savez(bad, x=np.zeros(3))
    with pytest.raises(ValueError):
        SarcTopology.load(bad)
    good = tmp_path / "t.npz"
    t.save(good)
    with np.load(good) as z:
        fields = {k: z[k] for k in z.files}
    fields["format_version"] = np.int64(SarcTopology.FORMAT_VERSION + 1)
    newer = tmp_path / "newer.npz"
    np.savez(newer, **fields)
    with pytest.raises(ValueError):
        SarcTopology.load(newer)
    fields["format_version"] = np.int64(1)
    fields["crown_offsets"] = fields["crown_offsets"][:, :-1]
    broken = tmp_path / "broken.npz"
    np.savez(broken, **fields)
    with pytest.raises(ValueError):
        SarcTopology.load(broken)
