"""Multi-RHS product of the compressed far field (VALU, FP32 MFMA, split
BF16x3 MFMA) against numpy on the downloaded C, U, R panels and against the
dense FP64 far field. The product runs in interpolative form P (R X) with
P = C U formed in FP64, so FP32 arithmetic keeps the tolerances below even
though the core U is ill-conditioned."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def compressed(farcur, oracle):
    mesh = farcur.Mesh.icosphere(1.0, 1.0, 1.0, 3, 0.01, 1).build_rwg()
    part = farcur.Partition(mesh, 0.5, 2.0)
    t = mesh.tables()
    ctx = farcur.Context(0)
    ctx.set_mesh(mesh)
    ro, ri, co, cix = part.far_blocks()
    bt = farcur.Batch(ctx, ro, ri, co, cix)
    bt.compress(farcur.MODE_BATCHED, tol=1e-4, seed=4, precision=farcur.F32, pinv_rtol=1e-6)
    g = oracle.Geometry(t["vertex"], t["topol"], t["edges"], t["ln"])
    N = t["ln"].size
    yield ctx, bt, g, (ro, ri, co, cix), N
    bt.close()
    ctx.close()


def _reference(bt, blocks, V, symmetric):
    ro, ri, co, cix = blocks
    Y = np.zeros_like(V)
    for b in range(bt.nblocks):
        out = bt.block(b)
        A = out["C"] @ out["U"] @ out["R"]
        rows, cols = ri[ro[b]:ro[b + 1]], cix[co[b]:co[b + 1]]
        Y[rows] += A @ V[cols]
        if symmetric:
            Y[cols] += A.T @ V[rows]
    return Y


@pytest.mark.parametrize("symmetric", [False, True])
def test_valu_matches_numpy(farcur, compressed, symmetric):
    ctx, bt, g, blocks, N = compressed
    rng = np.random.default_rng(0)
    V = rng.standard_normal((N, 70)) + 1j * rng.standard_normal((N, 70))  # 70: a partial chunk of 64
    Y, ms = bt.matvec(V, farcur.MV_VALU, symmetric)
    Yr = _reference(bt, blocks, V.astype(np.complex64).astype(np.complex128), symmetric)
    assert np.linalg.norm(Y - Yr) <= 2e-5 * np.linalg.norm(Yr)
    assert ms[0] > 0 and ms[1] > 0
    # P = C U is formed once per compress
    Y2, ms2 = bt.matvec(V, farcur.MV_VALU, symmetric)
    assert ms2[2] == 0.0
    np.testing.assert_array_equal(Y2, Y)


def test_permuted_group_rejected(farcur, compressed):
    # two blocks over the same RWG set in different orders cannot share a target group
    ctx, bt, g, blocks, N = compressed
    ro, ri, co, cix = blocks
    rows = ri[ro[0]:ro[1]]
    cols = cix[co[0]:co[1]]
    r2 = np.concatenate([rows, rows[::-1]]).astype(np.int32)
    c2 = np.concatenate([cols, cols]).astype(np.int32)
    m, n = rows.size, cols.size
    b2 = farcur.Batch(ctx, np.array([0, m, 2 * m], np.int32), r2, np.array([0, n, 2 * n], np.int32), c2)
    b2.compress(farcur.MODE_BATCHED, tol=1e-4, seed=4, precision=farcur.F32)
    with pytest.raises(farcur.FarcurError):
        b2.matvec(np.ones((N, 2)), farcur.MV_VALU, False)
    b2.close()


@pytest.mark.parametrize("mode,tol", [(1, 2e-5), (2, 1e-4)])
def test_mfma_matches_valu(farcur, compressed, mode, tol):
    ctx, bt, g, blocks, N = compressed
    rng = np.random.default_rng(1)
    V = rng.standard_normal((N, 130)) + 1j * rng.standard_normal((N, 130))
    Yv, _ = bt.matvec(V, farcur.MV_VALU, True)
    Ym, _ = bt.matvec(V, mode, True)
    assert np.linalg.norm(Ym - Yv) <= tol * np.linalg.norm(Yv)
    # column by column, no lane-layout permutation
    err_cols = np.linalg.norm(Ym - Yv, axis=0) / np.linalg.norm(Yv, axis=0)
    assert err_cols.max() <= 10 * tol


def test_against_dense_far_field(farcur, compressed):
    ctx, bt, g, blocks, N = compressed
    ro, ri, co, cix = blocks
    rng = np.random.default_rng(2)
    V = rng.standard_normal((N, 8)) + 1j * rng.standard_normal((N, 8))
    Y, _ = bt.matvec(V, farcur.MV_MFMA_F32, True)
    Yd = np.zeros_like(V)
    for b in range(bt.nblocks):
        rows, cols = ri[ro[b]:ro[b + 1]], cix[co[b]:co[b + 1]]
        Z = ctx.fill(rows, cols)
        Yd[rows] += Z @ V[cols]
        Yd[cols] += Z.T @ V[rows]
    assert np.linalg.norm(Y - Yd) <= 1e-4 * np.linalg.norm(Yd)


@pytest.mark.parametrize("symmetric", [False, True])
def test_single_rhs_matches_numpy(farcur, compressed, symmetric):
    # The HBM-bound one-vector kernel (one workgroup per block and
    # orientation, FP32 atomics into y) against numpy on the panels, and
    # against the multi-RHS product's column for the same vector.
    ctx, bt, g, blocks, N = compressed
    rng = np.random.default_rng(5)
    v = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    y, ms = bt.matvec1(v, symmetric)
    yr = _reference(bt, blocks, v.astype(np.complex64).astype(np.complex128)[:, None], symmetric)[:, 0]
    assert np.linalg.norm(y - yr) <= 2e-5 * np.linalg.norm(yr)
    assert ms > 0
    Y, _ = bt.matvec(np.stack([v, v], axis=1), farcur.MV_VALU, symmetric)
    assert np.linalg.norm(y - Y[:, 0]) <= 2e-5 * np.linalg.norm(Y[:, 0])
