"""Mesh and RWG tables (host code of the library; no GPU needed), checked
against the numpy restatement of get_edge.m / sphere.m and against the
invariants get_edge.m establishes."""
import numpy as np
import pytest

import rwg


def _check_rwg_invariants(t):
    edges = t["edges"].astype(int) - 1
    topol = t["topol"].astype(int) - 1
    trian = t["trian"].astype(int)
    ne = edges.shape[1]
    for e in range(ne):
        tp, tm, vp, vm = edges[:, e]
        assert tp != tm
        # free vertex belongs to its triangle; the other two are the shared edge
        assert vp in topol[:, tp] and vm in topol[:, tm]
        shared_p = set(topol[:, tp]) - {vp}
        shared_m = set(topol[:, tm]) - {vm}
        assert shared_p == shared_m and len(shared_p) == 2
        # trian: +e+1 at the free vertex position of T+, -(e+1) at T-
        assert trian[list(topol[:, tp]).index(vp), tp] == e + 1
        assert trian[list(topol[:, tm]).index(vm), tm] == -(e + 1)
        a, b = sorted(shared_p)
        assert t["ln"][e] == pytest.approx(np.linalg.norm(t["vertex"][:, a] - t["vertex"][:, b]), rel=1e-14)


def test_config1_plates(farcur):
    m = farcur.plates_config(1.0, 20, 4.0)
    nv, nt, ne = m.sizes()
    assert (nv, nt, ne) == (2 * 441, 2 * 800, 2 * 1160)
    t = m.tables()
    _check_rwg_invariants(t)
    # normals +z, areas (1/20)^2 / 2
    assert np.allclose(t["un"], [[0], [0], [1]])
    assert np.allclose(t["ds"], 0.5 / 400)
    # first 1160 RWGs on plate z = 0, the rest on z = 4
    cz = t["cent"][2]
    edges = t["edges"].astype(int) - 1
    assert np.all(cz[edges[0, :1160]] == 0) and np.all(cz[edges[0, 1160:]] == 4)


def test_rwg_matches_get_edge_restatement(farcur):
    for mesh in (farcur.plates_config(1.0, 6, 2.0), farcur.Mesh.octasphere(1.0, 2).build_rwg(),
                 farcur.Mesh.icosphere(1.0, 1.0, 1.0, 2, 0.01, 3).build_rwg()):
        t = mesh.tables()
        ref = rwg.get_edge(t["vertex"], t["topol"])
        np.testing.assert_array_equal(t["edges"], ref["edges"])
        np.testing.assert_array_equal(t["trian"], ref["trian"])
        np.testing.assert_allclose(t["ln"], ref["ln"], rtol=1e-14)
        np.testing.assert_allclose(t["cent"], ref["cent"], rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(t["un"], ref["un"], rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(t["ds"], ref["ds"], rtol=1e-13)


@pytest.mark.parametrize("n", [0, 1, 3, 4])
def test_octasphere_matches_sphere_m(farcur, n):
    m = farcur.Mesh.octasphere(1.0, n).build_rwg()
    t = m.tables()
    V, T = rwg.octasphere(1.0, n)
    np.testing.assert_array_equal(t["topol"], T)
    np.testing.assert_allclose(t["vertex"], V, rtol=0, atol=1e-15)
    assert t["ln"].size == 12 * 4 ** n
    _check_rwg_invariants(t)
    # outward normals: un . cent > 0 (get_edge.m normal convention)
    assert np.all(np.sum(t["un"] * t["cent"], axis=0) > 0)


def test_icosphere_level5_counts(farcur):
    m = farcur.Mesh.icosphere(2.0, 2.0, 2.0, 5, 0.01, 1).build_rwg()
    nv, nt, ne = m.sizes()
    assert (nv, nt, ne) == (10242, 20480, 30720)
    assert nv - ne + nt == 2  # Euler, closed genus-0 surface
    t = m.tables()
    r = np.linalg.norm(t["vertex"], axis=0)
    h = t["ln"].mean()
    # radial jitter within +-1% of the mesh size
    assert np.all(np.abs(r - 2.0) <= 0.0101 * h + 1e-12) and np.std(r) > 0
    assert np.all(np.sum(t["un"] * t["cent"], axis=0) > 0)


def test_ellipsoid_axes(farcur):
    m = farcur.Mesh.icosphere(4.0, 2.0, 1.0, 3, 0.0, 1).build_rwg()
    V = m.tables()["vertex"]
    assert np.allclose(np.sum((V / np.array([[4.0], [2.0], [1.0]])) ** 2, axis=0), 1.0)


def test_partition_sphere(farcur):
    m = farcur.Mesh.icosphere(2.0, 2.0, 2.0, 4, 0.01, 1).build_rwg()
    p = farcur.Partition(m, 0.5, 2.0)
    ne = m.sizes()[2]
    assert np.array_equal(np.sort(p.group_rwg), np.arange(ne))
    ng = p.ngroups
    assert len(p.far_pairs) + len(p.near_pairs) == ng * (ng + 1) // 2
    for a, b in p.far_pairs[:200]:
        d = np.linalg.norm(p.centre[a] - p.centre[b])
        assert d >= 2.0 * max(p.diam[a], p.diam[b])
    for a, b in p.near_pairs[:200]:
        if a != b:
            assert np.linalg.norm(p.centre[a] - p.centre[b]) < 2.0 * max(p.diam[a], p.diam[b])
    ro, ri, co, cix = p.far_blocks()
    assert ro[-1] == ri.size and co[-1] == cix.size


def test_bad_arguments(farcur):
    with pytest.raises(farcur.FarcurError):
        farcur.Mesh.plate(1.0, 1.0, 0, 3)
    with pytest.raises(farcur.FarcurError):
        farcur.Mesh.octasphere(-1.0, 2)
    m = farcur.Mesh.plate(1.0, 1.0, 2, 2)
    with pytest.raises(farcur.FarcurError):
        m.tables()  # RWG tables not built


def test_plane_waves_match_user_compute_Ei(farcur):
    import excitation

    m = farcur.Mesh.icosphere(2.0, 2.0, 2.0, 2, 0.01, 1).build_rwg()
    t = m.tables()
    th = np.linspace(0, np.pi, 5)
    ph = np.array([0.0, 0.3, 1.0, 2.0, 5.0])
    V = m.plane_waves(th, ph, k=2 * np.pi)
    for j in range(5):
        ref = excitation.plane_wave(t["vertex"], t["edges"], t["ln"], t["cent"], 2 * np.pi, th[j], ph[j])
        np.testing.assert_allclose(V[:, j], ref, rtol=1e-12, atol=1e-14)
