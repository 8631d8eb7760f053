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


def _brute_force_groups(m, leaf):
    """Leaf groups, centres and diameters recomputed from the mesh tables in
    numpy: RWG point = midpoint of the shared edge, cells of side `leaf`
    anchored at the bounding-box minimum, centre = mean RWG point, diameter =
    2 x the largest centre-to-vertex distance over the group's triangles."""
    t = m.tables()
    V = t["vertex"].T
    T = t["topol"].T.astype(int) - 1
    E = t["edges"].T.astype(int) - 1
    tp, vp = E[:, 0], E[:, 2]
    pts = np.empty((E.shape[0], 3))
    for e in range(E.shape[0]):
        ab = [v for v in T[tp[e]] if v != vp[e]]
        pts[e] = 0.5 * (V[ab[0]] + V[ab[1]])
    cell = np.floor((pts - pts.min(axis=0)) / leaf).astype(np.int64)
    return pts, cell, V, T, E


def test_partition_brute_force_all_pairs(farcur):
    # Every group against the mesh, and every unordered pair of groups (not a
    # sample) against the admissibility rule |c_a - c_b| >= eta max(d_a, d_b).
    m = farcur.Mesh.icosphere(2.0, 1.5, 1.0, 4, 0.01, 3).build_rwg()
    leaf, eta = 0.5, 2.0
    p = farcur.Partition(m, leaf, eta)
    pts, cell, V, T, E = _brute_force_groups(m, leaf)
    ng = p.ngroups
    for g in range(ng):
        ids = p.group(g)
        assert len({tuple(c) for c in cell[ids]}) == 1  # one leaf cell per group
        c = pts[ids].mean(axis=0)
        np.testing.assert_allclose(p.centre[g], c, rtol=0, atol=1e-12)
        verts = T[np.concatenate([E[ids, 0], E[ids, 1]])].ravel()
        np.testing.assert_allclose(p.diam[g], 2 * np.max(np.linalg.norm(V[verts] - c, axis=1)), rtol=1e-12)
    # distinct groups have distinct cells, and the groups cover every RWG once
    assert len({tuple(cell[p.group(g)[0]]) for g in range(ng)}) == ng
    assert np.array_equal(np.sort(p.group_rwg), np.arange(m.sizes()[2]))
    d = np.linalg.norm(p.centre[:, None, :] - p.centre[None, :, :], axis=2)
    dm = np.maximum(p.diam[:, None], p.diam[None, :])
    far = {(int(a), int(b)) for a, b in p.far_pairs}
    near = {(int(a), int(b)) for a, b in p.near_pairs}
    assert len(far) == len(p.far_pairs) and len(near) == len(p.near_pairs) and not far & near
    for a in range(ng):
        for b in range(a, ng):
            admissible = a != b and d[a, b] >= eta * dm[a, b]
            assert ((a, b) in far) == admissible and ((a, b) in near) == (not admissible), (a, b)


def test_block_list_file_round_trip(farcur, tmp_path):
    m = farcur.Mesh.icosphere(1.0, 1.0, 1.0, 3, 0.01, 1).build_rwg()
    p = farcur.Partition(m, 0.5, 2.0)
    path = tmp_path / "blocks.fcbl"
    p.save(path)
    q = farcur.Partition.load(path)
    assert q.params() == {"leaf_size": 0.5, "eta_adm": 2.0, "nrwg_mesh": m.sizes()[2]}
    for k in ("group_offsets", "group_rwg", "centre", "diam", "far_pairs", "near_pairs"):
        np.testing.assert_array_equal(getattr(q, k), getattr(p, k))
    for x, y in zip(q.far_blocks(), p.far_blocks()):
        np.testing.assert_array_equal(x, y)
    # the reference's Z_comp view: far blocks compressed (comp = 1) with their
    # transposes implied (self = 0); near blocks dense; diagonal blocks self = 1
    zc = q.zcomp()
    assert len(zc) == len(p.far_pairs) + len(p.near_pairs)
    nf = len(p.far_pairs)
    assert all(e["comp"] == 1 and e["self"] == 0 for e in zc[:nf])
    assert all(e["comp"] == 0 for e in zc[nf:])
    assert sum(e["self"] for e in zc) == p.ngroups
    a, b = p.far_pairs[0]
    np.testing.assert_array_equal(zc[0]["m"], p.group(a) + 1)
    np.testing.assert_array_equal(zc[0]["n"], p.group(b) + 1)
    # every entry once: far + near blocks and their transposes tile the matrix
    N = m.sizes()[2]
    cover = np.zeros((N, N), np.int8)
    for e in zc:
        cover[np.ix_(e["m"] - 1, e["n"] - 1)] += 1
        if e["self"] == 0:
            cover[np.ix_(e["n"] - 1, e["m"] - 1)] += 1
    assert cover.min() == 1 and cover.max() == 1


def test_block_list_file_rejects_damage(farcur, tmp_path):
    m = farcur.Mesh.icosphere(1.0, 1.0, 1.0, 2, 0.01, 1).build_rwg()
    p = farcur.Partition(m, 0.5, 2.0)
    path = tmp_path / "blocks.fcbl"
    p.save(path)
    raw = bytearray(path.read_bytes())
    bad = tmp_path / "bad.fcbl"
    for mutate, what in ((lambda b: b.__setitem__(40, b[40] ^ 1), "checksum"),
                         (lambda b: b.__setitem__(0, ord("X")), "not a block-list"),
                         (lambda b: b.__delitem__(slice(len(b) - 20, len(b))), "")):
        x = bytearray(raw)
        mutate(x)
        bad.write_bytes(bytes(x))
        with pytest.raises(farcur.FarcurError, match=what):
            farcur.Partition.load(bad)
    with pytest.raises(farcur.FarcurError, match="cannot open"):
        farcur.Partition.load(tmp_path / "missing.fcbl")
