"""Pins of the CPU oracle (no GPU): every oracle function is checked against
something independent of its own formula — published known-answer vectors,
LAPACK, numpy, closed forms, a high-order quadrature of the defining
integrals, or exact-rank identities."""
import math

import numpy as np
import pytest


# ---------------------------------------------------------------- RNG + sampling
def test_philox_known_answers(oracle):
    # Known-answer vectors of Philox4x32-10 published with the Random123 library
    # (Salmon et al., SC'11).
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in kat:
        assert tuple(int(x) for x in oracle.philox(ctr, key)) == want


@pytest.mark.parametrize("N,r", [(1, 1), (10, 1), (10, 10), (113, 11), (1160, 116), (3072, 614)])
def test_sample_invariants(oracle, N, r):
    for rnd in range(3):
        s = oracle.sample(7, 3, rnd, 1, N, r)
        assert s.size == r
        assert np.all(np.diff(s) > 0), "sorted and distinct"
        assert s.min() >= 0 and s.max() < N


def test_sample_uniform_chi2(oracle):
    # Each index is selected with probability r/N; chi-square over 4000 draws.
    N, r, draws = 40, 7, 4000
    counts = np.zeros(N)
    for b in range(draws):
        counts[oracle.sample(99, b, 0, 2, N, r)] += 1
    expected = draws * r / N
    chi2 = np.sum((counts - expected) ** 2 / expected)
    # dof = 39: P(chi2 > 72) < 1e-3
    assert chi2 < 72.0
    # Pairwise: the first two positions are not systematically chosen together.
    assert abs(counts.mean() - expected) < 1e-9


def test_sample_streams_differ(oracle):
    a = oracle.sample(1, 0, 0, 1, 500, 50)
    assert not np.array_equal(a, oracle.sample(1, 0, 1, 1, 500, 50))
    assert not np.array_equal(a, oracle.sample(1, 1, 0, 1, 500, 50))
    assert not np.array_equal(a, oracle.sample(2, 0, 0, 1, 500, 50))
    assert np.array_equal(a, oracle.sample(1, 0, 0, 1, 500, 50))


def test_vtest_range(oracle):
    v = oracle.vtest(5, 2, 20000)
    assert np.all((v.real >= 0) & (v.real < 1) & (v.imag >= 0) & (v.imag < 1))
    assert abs(v.real.mean() - 0.5) < 0.01 and abs(v.imag.var() - 1 / 12) < 0.005


# ---------------------------------------------------------------- pinv
def _lapack_pinv(W, atol, rtol=0.0):
    u, s, vh = np.linalg.svd(W)
    tau = max(atol, rtol * s[0])
    inv = np.where(s >= tau, 1.0 / np.where(s > 0, s, 1), 0.0)
    return (vh.conj().T * inv) @ u.conj().T, s, int(np.sum(s >= tau))


@pytest.mark.parametrize("r", [1, 2, 5, 17, 44, 64])
def test_pinv_matches_lapack_full_rank(oracle, r):
    rng = np.random.default_rng(r)
    W = rng.standard_normal((r, r)) + 1j * rng.standard_normal((r, r))
    U, s, kept = oracle.pinv(W, 1e-10)
    Uref, sref, kref = _lapack_pinv(W, 1e-10)
    assert kept == kref == r
    np.testing.assert_allclose(np.sort(s)[::-1], sref, rtol=1e-12, atol=1e-13)
    assert np.linalg.norm(U - Uref) <= 1e-10 * np.linalg.norm(Uref)


def test_pinv_threshold_absolute_and_relative(oracle):
    # Singular values 1, 1e-3, ..., 1e-15: the absolute 1e-10 rule keeps four
    # (kernel_Sinv, CU:1435-1444: s < tol -> 0), the relative 1e-6 rule three.
    rng = np.random.default_rng(4)
    r = 6
    q1, _ = np.linalg.qr(rng.standard_normal((r, r)) + 1j * rng.standard_normal((r, r)))
    q2, _ = np.linalg.qr(rng.standard_normal((r, r)) + 1j * rng.standard_normal((r, r)))
    s = np.array([1.0, 1e-3, 1e-6, 1e-9, 1e-12, 1e-15])
    W = (q1 * s) @ q2.conj().T
    for atol, rtol, want in [(1e-10, 0.0, 4), (1e-10, 1e-7, 3), (0.0, 0.0, 6)]:
        U, sig, kept = oracle.pinv(W, atol, rtol)
        assert kept == want
        if want < 6:
            Uref, _, _ = _lapack_pinv(W, atol, rtol)
            # kept part reproduces the truncated pseudo-inverse
            assert np.linalg.norm(U - Uref) <= 1e-6 * np.linalg.norm(Uref)


def test_pinv_zero_matrix(oracle):
    U, s, kept = oracle.pinv(np.zeros((4, 4), complex), 1e-10)
    assert kept == 0 and np.all(U == 0)


def test_pinv_qr_trunc_pins(oracle):
    import scipy.linalg

    rng = np.random.default_rng(3)
    cplx = lambda *shape: rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    # full rank: the inverse
    W = cplx(12, 12)
    U, d, k, perm = oracle.pinv_qr_trunc(W)
    Wi = np.linalg.inv(W)
    assert k == 12
    np.testing.assert_allclose(U, Wi, rtol=0, atol=1e-10 * np.abs(Wi).max())
    # the greedy pivots and |R_jj| of LAPACK's geqp3 (column norms graded so
    # that no two candidates tie)
    for n in (5, 17, 40):
        A = cplx(n, n) * np.logspace(0, -5, n)[rng.permutation(n)]
        _, Rl, pl = scipy.linalg.qr(A, pivoting=True)
        _, d, _, perm = oracle.pinv_qr_trunc(A)
        np.testing.assert_array_equal(perm, pl)
        np.testing.assert_allclose(d, np.abs(np.diag(Rl)), rtol=1e-9)
    # a planted rank under the relative threshold: k pivot columns are
    # interpolated exactly and W U is the orthogonal projector of rank k
    n, k0 = 30, 9
    X = cplx(n, k0) @ cplx(k0, n) + 1e-9 * cplx(n, n)
    U, d, k, perm = oracle.pinv_qr_trunc(X, 1e-10, 1e-6)
    assert k == k0
    UX = U @ X
    np.testing.assert_allclose(UX[:, perm[:k]], np.eye(n)[:, perm[:k]], atol=1e-9)
    P = X @ U
    assert np.abs(P @ P - P).max() <= 1e-8 and np.abs(P - P.conj().T).max() <= 1e-8
    assert np.trace(P).real == pytest.approx(k0, abs=1e-8)
    assert np.all(U[perm[k:], :] == 0)
    # the threshold removes a planted weak direction: the SVD pseudo-inverse's
    # kept count agrees when the gap is wide
    Us, ss, kept = oracle.pinv(X, 1e-10, 1e-6)
    assert kept == k0


# ---------------------------------------------------------------- residual, pivots
def test_residual_matches_numpy(oracle):
    rng = np.random.default_rng(3)
    m, r, n = 37, 9, 23
    C = rng.standard_normal((m, r)) + 1j * rng.standard_normal((m, r))
    U = rng.standard_normal((r, r)) + 1j * rng.standard_normal((r, r))
    R = rng.standard_normal((r, n)) + 1j * rng.standard_normal((r, n))
    v = rng.random(n) + 1j * rng.random(n)
    yo = np.ones(m, complex)
    y, err = oracle.residual(C, U, R, v, yo)
    yref = C @ (U @ (R @ v))
    np.testing.assert_allclose(y, yref, rtol=1e-13, atol=1e-12)
    assert err == pytest.approx(np.linalg.norm(yo - yref) / np.linalg.norm(yref), rel=1e-13)


def test_pivot_rows_first_max(oracle):
    rng = np.random.default_rng(1)
    C = rng.standard_normal((50, 6)) + 1j * rng.standard_normal((50, 6))
    C[10, 2] = 100.0
    C[30, 2] = -100.0  # equal magnitude, later row: first one wins (Julia findmax)
    p = oracle.pivot_rows(C)
    np.testing.assert_array_equal(p, np.argmax(np.abs(C) ** 2, axis=0))
    assert p[2] == 10


# ---------------------------------------------------------------- adaptive loop
@pytest.mark.parametrize("m,n,k", [(120, 90, 5), (300, 310, 12), (8, 9, 3)])
def test_compress_exact_rank(oracle, m, n, k):
    # Skeleton identity: a rank-k matrix is reproduced exactly once the sampled
    # intersection W has rank k (uniform samples, r >= k).
    rng = np.random.default_rng(m)
    A = (rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))) @ (
        rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n)))
    out = oracle.compress_block(None, dense=A, tol=1e-8)
    approx = out["C"] @ out["U"] @ out["R"]
    assert np.linalg.norm(approx - A) <= 1e-9 * np.linalg.norm(A)
    assert out["converged"] and out["kept"] == k
    # C and R are the sampled columns and rows.
    np.testing.assert_array_equal(out["C"], A[:, out["cs"]])
    np.testing.assert_array_equal(out["R"], A[out["rs"], :])


def test_compress_round_sequence(oracle):
    # cuda_cur_compress3 (SW2:889-967): r0 = D/10, then doubling capped at D,
    # a first round whose error is discarded, stop at r == D if tol is never met.
    rng = np.random.default_rng(0)
    m, n = 100, 130
    A = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))  # full rank: never converges early
    out = oracle.compress_block(None, dense=A, tol=1e-14)
    # r: 10 (discarded), 20, 40, 80, 100 -> 5 calls, final r = D
    assert out["r"] == 100 and out["rounds"] == 5 and not out["converged"]
    small = oracle.compress_block(None, dense=A[:7, :9], tol=1e-14)
    # D < 10: r0 = 1, then 2, 4, 7
    assert small["r"] == 7 and small["rounds"] == 4


# ---------------------------------------------------------------- element
def _tables(mesh):
    t = mesh.tables()
    return t


def _tri(t, tri):
    V = t["vertex"]
    ids = t["topol"][:, tri].astype(int) - 1
    return V[:, ids[0]], V[:, ids[1]], V[:, ids[2]]


def _rwg_halves(t, e):
    """[(triangle vertices, sign, free vertex)] for T+ and T- of RWG e (0-based)."""
    out = []
    for h in range(2):
        tri = int(t["edges"][h, e]) - 1
        vf = t["vertex"][:, int(t["edges"][2 + h, e]) - 1]
        out.append((_tri(t, tri), 1.0 if h == 0 else -1.0, vf))
    return out


def _subdivide(p0, p1, p2, level):
    tris = [(p0, p1, p2)]
    for _ in range(level):
        nxt = []
        for a, b, c in tris:
            ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
            nxt += [(a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)]
        tris = nxt
    return tris


def _points_weights(p0, p1, p2, level):
    """Degree-5 seven-point Radau-type rule (Hammer-Marlowe-Stroud) on each of
    4^level sub-triangles; weights sum to the area."""
    a1, b1 = 0.0597158717, 0.4701420641
    a2, b2 = 0.7974269853, 0.1012865073
    bary = [(1 / 3, 1 / 3, 1 / 3), (a1, b1, b1), (b1, a1, b1), (b1, b1, a1), (a2, b2, b2), (b2, a2, b2),
            (b2, b2, a2)]
    w = [0.225, *([0.1323941527] * 3), *([0.1259391805] * 3)]
    pts, wts = [], []
    for a, b, c in _subdivide(p0, p1, p2, level):
        area = 0.5 * np.linalg.norm(np.cross(b - a, c - a))
        for (l0, l1, l2), wi in zip(bary, w):
            pts.append(l0 * a + l1 * b + l2 * c)
            wts.append(wi * area)
    return np.array(pts), np.array(wts)


def _surface_div(f, p0, p1, p2):
    """Surface divergence of the vector field f at the triangle centroid by
    central differences along two in-plane orthonormal directions."""
    t1 = (p1 - p0) / np.linalg.norm(p1 - p0)
    nrm = np.cross(p1 - p0, p2 - p0)
    t2 = np.cross(nrm, t1)
    t2 /= np.linalg.norm(t2)
    c = (p0 + p1 + p2) / 3
    h = 1e-4 * np.linalg.norm(p1 - p0)
    return sum(np.dot(t, (f(c + h * t) - f(c - h * t)) / (2 * h)) for t in (t1, t2))


def _efie_highorder(t, m, n, k, eta, level):
    """Z_mn = -j k eta \\int\\int [f_m.f_n - div f_m div f_n / k^2] e^{-jkR}/(4 pi R)
    with f the RWG function built from its definition (rho/(2A) scaled by l,
    +/- on T+/T-) and the divergence by finite differences."""
    total = 0j
    lm, ln = t["ln"][m], t["ln"][n]
    for (tm, sm, vm) in _rwg_halves(t, m):
        am = 0.5 * np.linalg.norm(np.cross(tm[1] - tm[0], tm[2] - tm[0]))
        fm = lambda r, sm=sm, vm=vm, am=am: sm * lm / (2 * am) * (r - vm)
        dm = _surface_div(fm, *tm)
        pm, wm = _points_weights(*tm, level)
        for (tn, sn, vn) in _rwg_halves(t, n):
            an = 0.5 * np.linalg.norm(np.cross(tn[1] - tn[0], tn[2] - tn[0]))
            fn = lambda r, sn=sn, vn=vn, an=an: sn * ln / (2 * an) * (r - vn)
            dn = _surface_div(fn, *tn)
            pn, wn = _points_weights(*tn, level)
            F1 = np.array([fm(p) for p in pm])
            F2 = np.array([fn(p) for p in pn])
            R = np.linalg.norm(pm[:, None, :] - pn[None, :, :], axis=2)
            G = np.exp(-1j * k * R) / (4 * np.pi * R)
            br = F1 @ F2.T - dm * dn / k ** 2
            total += np.sum(wm[:, None] * wn[None, :] * br * G)
    return -1j * k * eta * total


def test_element_vs_high_order_quadrature(farcur, oracle):
    # Config-1 plates (20x20 squares per 1 x 1 lambda plate, 4 lambda apart):
    # the 3-point rule agrees with a converged high-order quadrature of the
    # defining integral to the rule's discretisation error.
    mesh = farcur.plates_config(1.0, 20, 4.0)
    t = _tables(mesh)
    k, eta = 2 * np.pi, 120 * np.pi
    g = oracle.Geometry(t["vertex"], t["topol"], t["edges"], t["ln"], k, eta)
    ne = t["ln"].size
    rng = np.random.default_rng(11)
    for _ in range(4):
        m = int(rng.integers(0, ne // 2))
        n = int(rng.integers(ne // 2, ne))
        z3 = g.element(m, n)
        zh = _efie_highorder(t, m, n, k, eta, 1)
        zh2 = _efie_highorder(t, m, n, k, eta, 2)
        assert abs(zh2 - zh) <= 1e-6 * abs(zh2)  # reference integral converged
        assert abs(z3 - zh2) <= 2e-3 * abs(zh2)
        assert g.element(n, m) == pytest.approx(z3, rel=1e-13)  # reciprocity


def test_element_near_pairs_same_plate(farcur, oracle):
    # Pairs a few cells apart on one plate: still far from singular for the
    # 3-point rule; agreement with the high-order integral to a few 1e-3.
    mesh = farcur.plates_config(1.0, 20, 4.0)
    t = _tables(mesh)
    k, eta = 2 * np.pi, 120 * np.pi
    g = oracle.Geometry(t["vertex"], t["topol"], t["edges"], t["ln"], k, eta)
    cent = lambda e: t["vertex"][:, (t["edges"][2:, e].astype(int) - 1)].mean(axis=1)
    m = 0
    cands = [e for e in range(1, 1160) if 0.3 < np.linalg.norm(cent(e) - cent(m)) < 0.5]
    n = cands[0]
    z3 = g.element(m, n)
    zh = _efie_highorder(t, m, n, k, eta, 2)
    assert abs(z3 - zh) <= 1e-2 * abs(zh)


def _dipole_field_dot(pm, pn, rm, rn, k, eta):
    """p_m . E(p_n) for a Hertzian current moment p_n at rn observed at rm:
    E = -j k eta G [(1 + 1/(jkR) - 1/(kR)^2) p - (1 + 3/(jkR) - 3/(kR)^2)(p.u)u],
    G = e^{-jkR}/(4 pi R) (the standard closed form of the dipole field)."""
    d = rm - rn
    R = np.linalg.norm(d)
    u = d / R
    G = np.exp(-1j * k * R) / (4 * np.pi * R)
    a = 1 + 1 / (1j * k * R) - 1 / (k * R) ** 2
    b = 1 + 3 / (1j * k * R) - 3 / (k * R) ** 2
    E = -1j * k * eta * G * (a * pn - b * np.dot(pn, u) * u)
    return np.dot(pm, E)


@pytest.mark.parametrize("offset", [(0.3, 0.0, 0.0), (0.0, 0.0, 0.25), (0.2, -0.15, 0.1), (1.3, 0.4, -0.7)])
def test_element_dipole_limit(farcur, oracle, offset):
    # Two tiny RWGs (one-square plates of side 1e-3 lambda): Z_mn -> p_m . E(p_n)
    # with p = l/2 [(c+ - v+) - (c- - v-)]; the near-field terms 1/(kR)^2 test
    # the sign and weight of the scalar-potential (div-div) term.
    a = 1e-3
    mesh = farcur.Mesh.plate(a, a, 1, 1, 0, 0, 0)
    mesh.append(farcur.Mesh.plate(a, a, 1, 1, 0, 0, 0), *offset)
    mesh.build_rwg()
    t = mesh.tables()
    assert t["ln"].size == 2
    k, eta = 2 * np.pi, 120 * np.pi
    g = oracle.Geometry(t["vertex"], t["topol"], t["edges"], t["ln"], k, eta)

    def moment(e):
        l = t["ln"][e]
        (tp, _, vp), (tm, _, vm) = _rwg_halves(t, e)
        cp, cm = sum(tp) / 3, sum(tm) / 3
        return l / 2 * ((cp - vp) - (cm - vm)), (cp + cm) / 2

    pm, rm = moment(0)
    pn, rn = moment(1)
    want = _dipole_field_dot(pm, pn, rm, rn, k, eta)
    got = g.element(0, 1)
    assert abs(got - want) <= 1e-4 * abs(want)


def _cur_like(rng, m, n, r, decay):
    # C U R with a graded spectrum, like a far-field CUR (columns and rows of
    # a numerically low-rank block, U their core's inverse)
    Qa, _ = np.linalg.qr(rng.standard_normal((m, r)) + 1j * rng.standard_normal((m, r)))
    Qb, _ = np.linalg.qr(rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r)))
    s = decay ** np.arange(r)
    Z = (Qa * s) @ Qb.conj().T
    cs, rs = rng.choice(n, r, replace=False), rng.choice(m, r, replace=False)
    C, R = Z[:, cs], Z[rs, :]
    U = np.linalg.pinv(Z[np.ix_(rs, cs)], rcond=1e-13)
    return C, U, R


@pytest.mark.parametrize("thresh", [0.0, 1e-6, 1e-3])
def test_svd_compress_pins(oracle, thresh):
    # svd_compress.m:19-34 pinned by what the mathematics fixes: the kept
    # singular values are those of the m x n product (computed another way),
    # C' and R'^H have orthonormal columns, the residual is the Eckart-Young
    # one, and nothing is dropped at thresh 0.
    rng = np.random.default_rng(3)
    C, U, R = _cur_like(rng, 90, 70, 24, 0.55)
    A = C @ U @ R
    Cp, s, Rp = oracle.svd_compress(C, U, R, thresh)
    sa = np.linalg.svd(A, compute_uv=False)
    k = int(np.sum(sa / sa[0] > thresh)) if thresh > 0 else min(24, int(np.sum(sa > 0)))
    assert s.size == k
    np.testing.assert_allclose(s, sa[:k], rtol=1e-8, atol=1e-13 * sa[0])
    np.testing.assert_allclose(Cp.conj().T @ Cp, np.eye(k), atol=1e-10)
    np.testing.assert_allclose(Rp @ Rp.conj().T, np.eye(k), atol=1e-10)
    resid = np.linalg.norm(Cp @ np.diag(s) @ Rp - A)
    assert abs(resid - np.linalg.norm(sa[k:24])) <= 1e-9 * sa[0]
    # a transposed or unconjugated R' misses the product by O(1)
    assert np.linalg.norm(Cp @ np.diag(s) @ Rp.conj() - A) > 0.1 * np.linalg.norm(A)


def test_svd_compress_planted_rank(oracle):
    rng = np.random.default_rng(4)
    C = rng.standard_normal((40, 5)) + 1j * rng.standard_normal((40, 5))
    Rf = rng.standard_normal((5, 30)) + 1j * rng.standard_normal((5, 30))
    # rank 5 product through an r = 12 CUR-shaped factorisation
    P = rng.standard_normal((5, 12)) + 1j * rng.standard_normal((5, 12))
    Cb = C @ P
    Rb = P.conj().T @ Rf
    Ub = np.linalg.pinv(P) @ np.linalg.pinv(P.conj().T)
    A = Cb @ Ub @ Rb
    Cp, s, Rp = oracle.svd_compress(Cb, Ub, Rb, 1e-8)
    assert s.size == 5
    assert np.linalg.norm(Cp @ np.diag(s) @ Rp - A) <= 1e-10 * np.linalg.norm(A)


def _smooth_block(rng, M, N):
    # a numerically low-rank kernel block between two separated point clouds
    x = rng.standard_normal((M, 3))
    y = rng.standard_normal((N, 3)) + np.array([8.0, 0, 0])
    d = np.linalg.norm(x[:, None, :] - y[None, :, :], axis=2)
    return np.exp(-1j * 2.0 * d) / d


def test_aca_pins(oracle):
    # ACA.m:40-128 pinned by properties the method fixes: the cross
    # approximation reproduces every pivot row and column, the first column
    # pivot is the largest entry of row 0, the error falls with tol, and a
    # block of exact rank 5 is recovered in 5 steps.
    rng = np.random.default_rng(5)
    Z = _smooth_block(rng, 70, 50)
    row, col = (lambda i: Z[i].copy()), (lambda j: Z[:, j].copy())
    U, V, I, J, st = oracle.aca_dense(row, col, 70, 50, 1e-6)
    A = U @ V
    assert st == 1 and I[0] == 0 and J[0] == int(np.argmax(np.abs(Z[0])))
    np.testing.assert_allclose(A[I, :], Z[I, :], atol=1e-10 * np.abs(Z).max())
    np.testing.assert_allclose(A[:, J], Z[:, J], atol=1e-10 * np.abs(Z).max())
    e6 = np.linalg.norm(A - Z) / np.linalg.norm(Z)
    U3, V3, *_ = oracle.aca_dense(row, col, 70, 50, 1e-3)
    e3 = np.linalg.norm(U3 @ V3 - Z) / np.linalg.norm(Z)
    assert e6 < 1e-5 and e3 > e6 and U3.shape[1] < U.shape[1]
    P = (rng.standard_normal((40, 5)) + 1j * rng.standard_normal((40, 5))) @ (
        rng.standard_normal((5, 30)) + 1j * rng.standard_normal((5, 30)))
    U5, V5, I5, J5, _ = oracle.aca_dense(lambda i: P[i].copy(), lambda j: P[:, j].copy(), 40, 30, 1e-12)
    assert U5.shape[1] in (5, 6)
    assert np.linalg.norm(U5 @ V5 - P) <= 1e-10 * np.linalg.norm(P)
    # vectors are returned whole (ACA.m:47)
    Uv, Vv, *_ = oracle.aca_dense(row, col, 1, 50, 1e-6)
    np.testing.assert_array_equal((Uv @ Vv)[0], Z[0])


def test_aca_non_finite_entries_stop_with_status_3(oracle):
    # Coincident row and column triangles make the 3-point element non-finite
    # (R = 0). The cross iteration must then stop with status 3 instead of
    # pivoting on a NaN: a row with no finite entry stops before any cross, a
    # NaN column stops after the first cross. The GPU kernel has the same
    # guards; the library also refuses such blocks when the batch is created.
    rng = np.random.default_rng(6)
    Z = _smooth_block(rng, 30, 20)
    Zn = Z.copy()
    Zn[0, :] = np.nan
    U, V, I, J, st = oracle.aca_dense(lambda i: Zn[i].copy(), lambda j: Zn[:, j].copy(), 30, 20, 1e-6)
    assert st == 3 and U.shape == (30, 0) and V.shape == (0, 20) and I.size == 0
    Zc = Z.copy()
    j0 = int(np.argmax(np.abs(Z[0])))
    Zc[5:, j0] = np.nan  # the first pivot column carries NaN below row 0
    U, V, I, J, st = oracle.aca_dense(lambda i: Zc[i].copy(), lambda j: Zc[:, j].copy(), 30, 20, 1e-6)
    assert st == 3 and U.shape[1] == 1 and J[0] == j0
    # finite entries next to a NaN are still pivots: the NaN is skipped, not chosen
    Zp = Z.copy()
    Zp[0, 3] = np.nan
    U, V, I, J, st = oracle.aca_dense(lambda i: Zp[i].copy(), lambda j: Zp[:, j].copy(), 30, 20, 1e-6)
    assert J[0] == int(np.nanargmax(np.abs(Zp[0])))


# ---------------------------------------------------------------- near field


def _polar_potentials(v, r, n=256):
    """int_T 1/R and int_T (r' - rho)/R by Gauss-Legendre quadrature in polar
    coordinates about the projection rho of r: T splits into the signed
    triangles (rho, v_i, v_i+1), each integrated along rays from rho to the
    edge (v_i, v_i+1) with the Jacobian t, which removes the 1/R singularity.
    Independent of the closed forms under test."""
    v = [np.asarray(x, float) for x in v]
    r = np.asarray(r, float)
    nrm = np.cross(v[1] - v[0], v[2] - v[0])
    nrm /= np.linalg.norm(nrm)
    h = np.dot(r - v[0], nrm)
    rho = r - h * nrm
    x, w = np.polynomial.legendre.leggauss(n)
    Is, Iv = 0.0, np.zeros(3)
    for i in range(3):
        A, B = v[i] - rho, v[(i + 1) % 3] - rho
        sgn = np.sign(np.dot(np.cross(A, B), nrm))
        if sgn == 0:
            continue
        s = 0.5 * (x + 1)  # position along the edge
        P = A[None, :] + s[:, None] * (B - A)[None, :]  # edge points relative to rho
        L = np.linalg.norm(P, axis=1)
        d = P / L[:, None]
        jac = np.linalg.norm(np.cross(A, B)) / L ** 2 * 0.5 * w  # d(angle) = |A x B| / L^2 ds
        t = 0.5 * L[:, None] * (x[None, :] + 1)  # radius along each ray
        f = t / np.sqrt(t * t + h * h) * 0.5 * L[:, None] * w[None, :] * (jac * sgn)[:, None]
        Is += f.sum()
        Iv += ((f * t).sum(axis=1)[:, None] * d).sum(axis=0)
    return Is, Iv


@pytest.mark.parametrize("r", [[0.2, 0.25, 0.0], [0.2, 0.25, 0.3], [0.2, 0.25, -0.05], [1.5, -0.5, 0.0],
                               [2.0, 0.0, 0.0], [-1.0, 0.0, 0.0], [2.0, 0.0, 0.4], [0.0, 0.0, 0.5],
                               [0.5, 0.5, 1e-3], [3.0, 3.0, 1.0]])
def test_potentials_closed_form_vs_polar_quadrature(oracle, r):
    # in-plane inside and outside, above and below, on an edge's extension
    # (R0 = 0 outside the segment), above a vertex, near the hypotenuse, far
    v = ([0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.1, 0.9, 0.0])
    Is, Iv = oracle.potentials(*v, r)
    Js, Jv = _polar_potentials(v, r)
    assert Is == pytest.approx(Js, rel=1e-9, abs=1e-12)
    np.testing.assert_allclose(Iv, Jv, rtol=1e-9, atol=1e-12)


def _plate_geometry(farcur, oracle, nsq):
    mesh = farcur.Mesh.plate(1.0, 1.0, nsq, nsq, 0, 0, 0).build_rwg()
    t = mesh.tables()
    return t, oracle.Geometry(t["vertex"], t["topol"], t["edges"], t["ln"])


def test_near_element_matches_far_rule_on_separated_pairs(farcur, oracle):
    # The singular-extraction path (forced with a huge near_factor) on pairs
    # several triangles apart integrates the same smooth kernel as the 3 x 3
    # rule, with the 7-point rule: both agree to the lower rule's error.
    t, g = _plate_geometry(farcur, oracle, 12)
    rows = np.array([0, 3, 10])
    cols = np.array([int(t["ln"].size) - 1, int(t["ln"].size) - 20])
    Zf = g.fill(rows, cols)
    Zn = g.fill_near(rows, cols, 1e9)
    assert np.abs(Zn - Zf).max() <= 2e-3 * np.abs(Zf).max()
    # and with the default factor these pairs are far: the same rule, equal to rounding
    np.testing.assert_allclose(g.fill_near(rows, cols, 2.0), Zf, rtol=1e-12, atol=1e-14 * np.abs(Zf).max())


def test_near_element_reciprocity_and_finite_self_terms(farcur, oracle):
    t, g = _plate_geometry(farcur, oracle, 6)
    ne = t["ln"].size
    Z = g.fill_near(np.arange(ne), np.arange(ne))
    assert np.all(np.isfinite(Z))
    # the Galerkin form is symmetric; the closed-form inner and 7-point outer
    # integrals break it only at the quadrature error
    assert np.linalg.norm(Z - Z.T) <= 2e-3 * np.linalg.norm(Z)
    # the far rule puts quadrature points on top of each other for coincident triangles
    assert not np.all(np.isfinite(g.fill(np.arange(ne), np.arange(ne))))


def test_mie_series_limits():
    import mie

    k = 2 * np.pi
    a = 1e-3
    # Rayleigh limit of the PEC sphere's backscatter: 9 pi a^2 (ka)^4
    assert mie.rcs(k, a, np.array([np.pi]))[0] == pytest.approx(9 * np.pi * a * a * (k * a) ** 4, rel=1e-4)
    # optical limit: pi a^2
    assert mie.rcs(k, 20.0, np.array([np.pi]))[0] == pytest.approx(np.pi * 400.0, rel=0.05)
    # optical theorem: extinction = (4 pi / k^2) Re S(0), for both amplitudes
    S1, S2 = mie.amplitudes(k * 0.3, np.array([0.0]))
    assert mie.extinction(k, 0.3) == pytest.approx(4 * np.pi / k ** 2 * S1[0].real, rel=1e-10)
    assert S1[0] == pytest.approx(S2[0], rel=1e-12)


@pytest.mark.parametrize("a", [0.25, 0.35])
def test_pec_sphere_rcs_matches_mie(farcur, oracle, a):
    # End to end: the dense near-field matrix of a PEC sphere (1920 RWGs),
    # J = -Z^-1 V for an x-polarised plane wave travelling along -z (the
    # reference's solve, compres_single_block_spheres.jl:3773, and
    # user_compute_Ei.m's excitation), then the bistatic RCS in the E-plane
    # against the Mie series. A wrong sign or weight of the charge term, a
    # lost 1/4 pi or wrong self terms move it by decibels.
    import excitation
    import mie

    mesh = farcur.Mesh.icosphere(a, a, a, 3, 0.0, 1).build_rwg()
    t = mesh.tables()
    k, eta = 2 * np.pi, 120 * np.pi
    g = oracle.Geometry(t["vertex"], t["topol"], t["edges"], t["ln"], k, eta)
    N = t["ln"].size
    Z = g.fill_near(np.arange(N), np.arange(N), 2.0)
    V = excitation.plane_wave(t["vertex"], t["edges"], t["ln"], t["cent"], k, 0.0, 0.0)
    J = np.linalg.solve(Z, -V)
    th = np.linspace(0, np.pi, 37)
    rhat = np.vstack([np.sin(th), np.zeros_like(th), np.cos(th)])
    s_mom = mie.mom_rcs(t["vertex"], t["topol"], t["edges"], t["ln"], J, k, eta, rhat)
    s_mie = mie.rcs(k, a, np.pi - th, "E")
    db = 10 * np.log10(s_mom / s_mie)
    assert np.abs(db).max() <= 0.3, db
