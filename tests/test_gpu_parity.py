"""GPU parity of every kernel against the CPU oracle, through the C ABI.

Tolerances: FP64 fill element-wise 1e-10 relative (to the panel's largest
entry); FP64 pinv 1e-10 relative Frobenius against LAPACK; integer work
(samples, pivots, round counts) bit-exact; complex-FP32 fill 2e-5 relative
Frobenius against FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K, ETA = 2 * np.pi, 120 * np.pi


@pytest.fixture(scope="module")
def plates(farcur, oracle):
    mesh = farcur.plates_config(1.0, 20, 4.0)
    t = mesh.tables()
    ctx = farcur.Context(0)
    ctx.set_mesh(mesh)
    g = oracle.Geometry(t["vertex"], t["topol"], t["edges"], t["ln"], K, ETA)
    yield ctx, g, t
    ctx.close()


def _relmax(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def _relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("nr,nc", [(1, 1), (64, 16), (65, 17), (200, 150), (1160, 40)])
def test_fill_f64_matches_oracle(plates, nr, nc):
    ctx, g, _ = plates
    rng = np.random.default_rng(nr * 1000 + nc)
    rows = np.sort(rng.choice(1160, nr, replace=False))
    cols = 1160 + np.sort(rng.choice(1160, nc, replace=False))
    Z = ctx.fill(rows, cols)
    Zo = g.fill(rows, cols)
    assert _relmax(Z, Zo) <= 1e-10
    nz = np.abs(Zo) > 1e-3 * np.abs(Zo).max()
    assert np.max(np.abs(Z[nz] - Zo[nz]) / np.abs(Zo[nz])) <= 1e-10


def test_fill_same_plate_and_unsorted_ids(plates):
    ctx, g, _ = plates
    rng = np.random.default_rng(5)
    rows = rng.choice(2320, 97)  # repeats and any order are allowed
    cols = rng.choice(2320, 33)
    # keep non-coincident pairs only (no singular self terms in this path)
    Z = ctx.fill(rows, cols)
    Zo = g.fill(rows, cols)
    finite = np.isfinite(Zo)
    assert np.array_equal(finite, np.isfinite(Z))
    assert _relmax(Z[finite], Zo[finite]) <= 1e-10


def test_fill_f32_vs_f64(farcur, plates):
    ctx, g, _ = plates
    rows = np.arange(0, 1160, 3)
    cols = np.arange(1160, 2320, 7)
    Z64 = ctx.fill(rows, cols, farcur.F64)
    Z32 = ctx.fill(rows, cols, farcur.F32)
    print("f32 vs f64: relfro %.3e relmax %.3e" % (_relfro(Z32, Z64), _relmax(Z32, Z64)))
    assert _relfro(Z32, Z64) <= 2e-6
    assert _relmax(Z32, Z64) <= 5e-6


@pytest.mark.parametrize("N,r", [(1, 1), (10, 3), (113, 11), (1160, 116), (3072, 614)])
def test_sample_bitexact(plates, oracle, N, r):
    ctx, _, _ = plates
    for seed, block, rnd, purpose in [(12345, 0, 0, 1), (12345, 0, 1, 2), (7, 911, 3, 1), (2**40 + 3, 5, 0, 2)]:
        np.testing.assert_array_equal(ctx.sample(seed, block, rnd, purpose, N, r),
                                      oracle.sample(seed, block, rnd, purpose, N, r))


def _lapack_pinv(W, atol, rtol=0.0):
    u, s, vh = np.linalg.svd(W)
    tau = max(atol, rtol * s[0])
    inv = np.where(s >= tau, 1.0 / np.where(s > 0, s, 1), 0.0)
    return (vh.conj().T * inv) @ u.conj().T, s, int(np.sum(s >= tau))


@pytest.mark.parametrize("r", [1, 2, 7, 44, 45, 64, 130, 300])
def test_pinv_matches_lapack(plates, r):
    ctx, _, _ = plates
    rng = np.random.default_rng(r)
    W = rng.standard_normal((r, r)) + 1j * rng.standard_normal((r, r))
    U, s, kept = ctx.pinv(W, 1e-10)
    Uref, sref, kref = _lapack_pinv(W, 1e-10)
    assert kept == kref
    np.testing.assert_allclose(s, sref, rtol=1e-11, atol=1e-12 * sref[0])
    assert _relfro(U, Uref) <= 1e-10


@pytest.mark.parametrize("r", [5, 16, 31, 48, 64])
def test_pinv_graded_relative_threshold(plates, r):
    # Singular values spread over 12 decades (the shape of far-field cores),
    # none within a factor 3 of the relative threshold 1e-6: the kept set and
    # the truncated pinv must match LAPACK's. r <= 64 runs the QR-preconditioned
    # LDS path.
    ctx, _, _ = plates
    rng = np.random.default_rng(100 + r)
    qa, _ = np.linalg.qr(rng.standard_normal((r, r)) + 1j * rng.standard_normal((r, r)))
    qb, _ = np.linalg.qr(rng.standard_normal((r, r)) + 1j * rng.standard_normal((r, r)))
    sv = 10.0 ** np.linspace(0, -12, r)
    sv = sv[np.abs(np.log10(sv) + 6) > np.log10(3)]
    sv = np.concatenate([sv, np.full(r - sv.size, 1e-13)])
    W = 3.0 * (qa * sv) @ qb.conj().T
    U, s, kept = ctx.pinv(W, 0.0, 1e-6)
    Uref, sref, kref = _lapack_pinv(W, 0.0, 1e-6)
    assert kept == kref
    np.testing.assert_allclose(s[:kept], sref[:kept], rtol=1e-9)
    assert _relfro(U, Uref) <= 1e-7
    assert _relfro(W @ U @ W, W @ Uref @ W) <= 1e-8


def test_pinv_ill_conditioned_core(plates, oracle):
    # A real intersection matrix of the config-1 block: rank-deficient far
    # field, absolute threshold 1e-10. The truncated pinv is checked where it
    # matters: W U W = W and C U R against the oracle's.
    ctx, g, _ = plates
    rs = np.arange(0, 1160, 29)[:40]
    cs = 1160 + np.arange(3, 1160, 29)[:40]
    W = g.fill(rs, cs)
    U, s, kept = ctx.pinv(W, 1e-10)
    Uo, so, kepto = oracle.pinv(W, 1e-10)
    assert kept == kepto
    np.testing.assert_allclose(np.sort(s)[::-1], np.sort(so)[::-1], rtol=1e-8, atol=1e-13 * s.max())
    # W U W = W up to the dropped singular values (below 1e-10).
    sl = np.linalg.svd(W, compute_uv=False)
    dropped = np.linalg.norm(sl[sl < 1e-10])
    assert np.linalg.norm(W @ U @ W - W) <= 1.5 * dropped + 1e-9 * np.linalg.norm(W)
    Ul, _, _ = _lapack_pinv(W, 1e-10)
    assert _relfro(W @ U @ W, W @ Ul @ W) <= 1e-6


def test_residual_matches_numpy(plates):
    ctx, _, _ = plates
    rng = np.random.default_rng(2)
    for m, r, n in [(1, 1, 1), (113, 22, 97), (1160, 232, 1160), (3072, 300, 64)]:
        C = rng.standard_normal((m, r)) + 1j * rng.standard_normal((m, r))
        U = rng.standard_normal((r, r)) + 1j * rng.standard_normal((r, r))
        R = rng.standard_normal((r, n)) + 1j * rng.standard_normal((r, n))
        v = rng.random(n) + 1j * rng.random(n)
        yo = np.ones(m, complex)
        y, err = ctx.cur_residual(C, U, R, v, yo)
        yref = C @ (U @ (R @ v))
        assert _relfro(y, yref) <= 1e-13
        assert err == pytest.approx(np.linalg.norm(yo - yref) / np.linalg.norm(yref), rel=1e-12)


def test_pivot_rows_bitexact(plates, oracle):
    ctx, g, _ = plates
    C = g.fill(np.arange(1160), 1160 + np.arange(0, 1160, 50))
    np.testing.assert_array_equal(ctx.pivot_rows(C), oracle.pivot_rows(C))
    rng = np.random.default_rng(0)
    X = rng.standard_normal((333, 9)) + 0j
    X[100, 3] = X[200, 3] = 50.0  # tie: first row wins
    p = ctx.pivot_rows(X)
    np.testing.assert_array_equal(p, oracle.pivot_rows(X))
    assert p[3] == 100


def _cfg1_block():
    return np.arange(1160, dtype=np.int32), np.arange(1160, 2320, dtype=np.int32)


@pytest.mark.parametrize("row_mode", [0, 1])
def test_compress_block_matches_oracle(plates, oracle, row_mode):
    # Plain adaptive CUR of the config-1 block (1160 x 1160): same samples,
    # same rounds, panels equal to FP64 rounding, same approximation.
    ctx, g, _ = plates
    rows, cols = _cfg1_block()
    out = ctx.compress_block(rows, cols, tol=1e-4, seed=12345, row_mode=row_mode)
    ref = oracle.compress_block(g, rows, cols, tol=1e-4, seed=12345, row_mode=row_mode)
    assert (out["r"], out["rounds"]) == (ref["r"], ref["rounds"])
    np.testing.assert_array_equal(out["cs"], cols[ref["cs"]])
    np.testing.assert_array_equal(out["rs"], rows[ref["rs"]])
    assert _relmax(out["C"], ref["C"]) <= 1e-10
    assert _relmax(out["R"], ref["R"]) <= 1e-10
    # err ~ 1e-8 is a difference of two ill-conditioned approximations: absolute agreement
    assert out["err"] == pytest.approx(ref["err"], rel=0.05, abs=1e-9)
    A = out["C"] @ out["U"] @ out["R"]
    Ao = ref["C"] @ ref["U"] @ ref["R"]
    Z = g.fill(rows, cols)
    assert _relfro(A, Ao) <= 1e-6  # |U| ~ 1/sigma_min amplifies rounding (absolute 1e-10 threshold)
    assert abs(_relfro(A, Z) - _relfro(Ao, Z)) <= 1e-6
    assert _relfro(A, Z) <= 1e-4  # accuracy bar for the plain method
    assert out["converged"] == 1


def test_compress_block_f32(farcur, plates):
    ctx, g, _ = plates
    rows, cols = _cfg1_block()
    out = ctx.compress_block(rows, cols, tol=1e-4, seed=3, precision=farcur.F32, pinv_rtol=1e-6)
    Z = g.fill(rows, cols)
    assert _relfro(out["C"] @ out["U"] @ out["R"], Z) <= 1e-4


def test_compress_small_blocks(plates, oracle):
    ctx, g, _ = plates
    for m, n in [(1, 1), (3, 5), (9, 30), (12, 10)]:
        rows = np.arange(m, dtype=np.int32)
        cols = 1160 + np.arange(n, dtype=np.int32)
        out = ctx.compress_block(rows, cols, tol=1e-6, seed=1)
        ref = oracle.compress_block(g, rows, cols, tol=1e-6, seed=1)
        assert (out["r"], out["rounds"]) == (ref["r"], ref["rounds"])
        assert _relfro(out["C"] @ out["U"] @ out["R"], ref["C"] @ ref["U"] @ ref["R"]) <= 1e-9


def test_drop_in_round_matches_oracle(farcur, plates, oracle):
    # One call of the 66-argument entry point with caller samples, as
    # pure_cuda_compression2 makes it.
    ctx, g, t = plates
    rng = np.random.default_rng(1)
    r1 = np.arange(1, 1161)
    r2 = np.arange(1161, 2321)
    vt = rng.random(1160) + 1j * rng.random(1160)
    yo = np.ones(1160, complex)
    C, U, R, rs, cs, y, err, ms = farcur.pure_cuda_compression2(vt, yo, 116, 1e-10, r1, r2, t, farcur.EMData(), rng)
    rsi, csi = rs.astype(int) - 1, cs.astype(int) - 1
    Co = g.fill(np.arange(1160), csi)
    Wo = g.fill(rsi, csi)
    Ro = g.fill(rsi, np.arange(1160, 2320))
    assert _relmax(C, Co) <= 1e-10 and _relmax(R, Ro) <= 1e-10
    Uo, _, _ = oracle.pinv(Wo, 1e-10)
    yref, erref = oracle.residual(Co, Uo, Ro, vt, yo)
    assert _relfro(y, yref) <= 1e-8
    assert err == pytest.approx(erref, rel=1e-8)
    assert ms >= 0


def test_reference_adaptive_loop_through_drop_in(farcur, plates):
    ctx, g, t = plates
    rng = np.random.default_rng(2)
    r1 = np.arange(1, 1161)
    r2 = np.arange(1161, 2321)
    C, U, R, rs, cs, err, ms = farcur.cuda_cur_compress3(r1, r2, 1e-4, t, farcur.EMData(), rng)
    Z = g.fill(np.arange(1160), np.arange(1160, 2320))
    assert err <= 1e-4
    assert _relfro(C @ U @ R, Z) <= 1e-4


def test_drop_in_rejects_bad_input(farcur, plates):
    ctx, g, t = plates
    rng = np.random.default_rng(3)
    r1 = np.arange(1, 11)
    r2 = np.array([1161, 1162, 99999])  # out-of-range id
    with pytest.raises(farcur.FarcurError):
        farcur.pure_cuda_compression2(np.ones(3), np.ones(10), 2, 1e-10, r1, r2, t, farcur.EMData(), rng)
    with pytest.raises(farcur.FarcurError):
        ctx.fill([0, 5000], [1])


def test_near_field_blocks_are_refused(farcur, plates):
    # The 3-point far-field quadrature cannot integrate touching or coincident
    # triangle pairs (the reference switches those to singular integration):
    # every entry point that compresses a block refuses one whose row and
    # column triangles share a vertex, and the drop-in also refuses centroid
    # pairs closer than its Rinteg_s / Ranal_s / Rinteg_f radii.
    ctx, g, t = plates
    same_plate = (np.arange(0, 40, dtype=np.int32), np.arange(20, 80, dtype=np.int32))  # shares RWGs
    with pytest.raises(farcur.FarcurError, match="near-field"):
        ctx.compress_block(*same_plate, tol=1e-4, seed=1)
    far = (np.arange(0, 40, dtype=np.int32), np.arange(1160, 1200, dtype=np.int32))
    ro = np.array([0, 40, 80], np.int32)
    ri = np.concatenate([far[0], same_plate[0]])
    co = np.array([0, 40, 100], np.int32)
    ci = np.concatenate([far[1], same_plate[1]])
    with pytest.raises(farcur.FarcurError, match="block 1"):
        farcur.Batch(ctx, ro, ri, co, ci)
    # neighbouring RWGs that only share a vertex or an edge (no common RWG)
    rows = np.array([0], np.int32)
    tri = t["edges"][:2].T.astype(int)  # 1-based T+, T- per RWG
    verts0 = set(t["topol"][:, tri[0] - 1].ravel())
    touch = [j for j in range(1, 1160) if not set(tri[j]) & set(tri[0]) and set(t["topol"][:, tri[j] - 1].ravel()) & verts0]
    assert touch
    with pytest.raises(farcur.FarcurError, match="near-field"):
        ctx.compress_block(rows, np.array(touch, np.int32), tol=1e-4, seed=1)
    # drop-in: the plates are 4 apart, so a 5 m integration radius makes the block near
    rng = np.random.default_rng(4)
    r1, r2 = np.arange(1, 1161), np.arange(1161, 2321)
    em = farcur.EMData(Rint_s=5.0, Ranal_s=5.0, Rint_f=5.0)
    with pytest.raises(farcur.FarcurError, match="integration radius"):
        farcur.pure_cuda_compression2(rng.random(1160) + 0j, np.ones(1160), 20, 1e-10, r1, r2, t, em, rng)
    em = farcur.EMData(Rint_s=1.0, Ranal_s=1.0, Rint_f=1.0)  # the reference's settings (SW2:1092-1106)
    C, U, R, *_ = farcur.pure_cuda_compression2(rng.random(1160) + 0j, np.ones(1160), 20, 1e-10, r1, r2, t, em, rng)
    assert np.all(np.isfinite(C))


def test_learned_mode_needs_heldout_rows(farcur, plates):
    ctx, g, t = plates
    ro = np.array([0, 40], np.int32)
    co = np.array([0, 40], np.int32)
    bt = farcur.Batch(ctx, ro, np.arange(40, dtype=np.int32), co, np.arange(1160, 1200, dtype=np.int32))
    with pytest.raises(farcur.FarcurError, match="heldout"):
        bt.compress(farcur.MODE_LEARNED, tol=1e-4, seed=1, heldout=0)
    bt.close()
