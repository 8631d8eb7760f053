"""SVD recompression of compressed blocks (svd_compress.m:19-34) on the GPU
against the oracle's numpy restatement, on the sphere's far blocks and on
ragged hand-made blocks (single rows/columns, thin strips, unsorted ids)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K, ETA = 2 * np.pi, 120 * np.pi


@pytest.fixture(scope="module")
def sphere(farcur):
    mesh = farcur.Mesh.icosphere(1.0, 1.0, 1.0, 3, 0.01, 1).build_rwg()
    part = farcur.Partition(mesh, 0.5, 2.0)
    ctx = farcur.Context(0)
    ctx.set_mesh(mesh)
    yield ctx, part.far_blocks()
    ctx.close()


def _check_block(oracle, out, rc, thresh, prod_tol, orth_tol):
    A = out["C"] @ out["U"] @ out["R"]
    Co, so, Ro = oracle.svd_compress(out["C"], out["U"], out["R"], thresh)
    s = rc["s"]
    if s.size != so.size:
        # only a singular value within rounding of the threshold may flip
        sa = np.linalg.svd(A, compute_uv=False)
        edge = sa[min(s.size, so.size)] / sa[0]
        assert abs(edge - thresh) <= 1e-6 * thresh, (s.size, so.size)
        return
    if s.size == 0:
        return
    # S = R_c U R_r^H is formed with cancellation when |U| is large (the
    # absolute 1e-10 pinv threshold lets |U| reach ~1e10): any two FP64
    # implementations differ by up to ~eps |C| |U| |R|, the backward error of
    # forming the product, on top of the stated relative bars
    fwd = 64 * np.finfo(float).eps * np.prod([np.linalg.norm(x, 2) for x in (out["C"], out["U"], out["R"])])
    np.testing.assert_allclose(s, so, rtol=0, atol=max(1e-9 * so[0], fwd))
    P, Po = rc["C"] @ np.diag(s) @ rc["R"], Co @ np.diag(so) @ Ro
    assert np.linalg.norm(P - Po) <= max(prod_tol * np.linalg.norm(Po), 4 * fwd)
    k = s.size
    assert np.abs(rc["C"].conj().T @ rc["C"] - np.eye(k)).max() <= orth_tol
    assert np.abs(rc["R"] @ rc["R"].conj().T - np.eye(k)).max() <= orth_tol


@pytest.mark.parametrize("precision,thresh", [(0, 1e-5), (0, 0.0), (1, 1e-5), (1, 1e-3)])
def test_recompress_matches_oracle(farcur, oracle, sphere, precision, thresh):
    ctx, (ro, ri, co, cix) = sphere
    bt = farcur.Batch(ctx, ro, ri, co, cix)
    opts = dict(tol=1e-4, seed=7, precision=precision)
    if precision == farcur.F32:
        opts["pinv_rtol"] = 1e-6
    bt.compress(farcur.MODE_BATCHED, **opts)
    ranks, ms = bt.recompress(thresh)
    info = bt.info_arrays()
    assert np.all(ranks <= info["r"]) and np.all(ranks >= 1)
    # FP64 arithmetic on either panel precision; FP32 panels store C', R' in FP32
    prod_tol, orth_tol = (1e-9, 1e-10) if precision == farcur.F64 else (2e-6, 2e-6)
    for b in range(0, bt.nblocks, max(1, bt.nblocks // 15)):
        _check_block(oracle, bt.block(b), bt.recompressed_block(b), thresh, prod_tol, orth_tol)


def test_recompress_ragged_blocks(farcur, oracle):
    mesh = farcur.plates_config(1.0, 20, 4.0)
    ctx = farcur.Context(0)
    ctx.set_mesh(mesh)
    rng = np.random.default_rng(9)
    shapes = [(1, 1), (1, 40), (40, 1), (2, 300), (300, 3), (65, 200), (129, 65)]
    rows = [rng.choice(1160, m, replace=False).astype(np.int32) for m, _ in shapes]
    cols = [(1160 + rng.choice(1160, n, replace=False)).astype(np.int32) for _, n in shapes]
    ro = np.concatenate([[0], np.cumsum([r.size for r in rows])]).astype(np.int32)
    co = np.concatenate([[0], np.cumsum([c.size for c in cols])]).astype(np.int32)
    bt = farcur.Batch(ctx, ro, np.concatenate(rows), co, np.concatenate(cols))
    bt.compress(farcur.MODE_BATCHED, tol=1e-4, seed=3)
    ranks, _ = bt.recompress(1e-6)
    for b, (m, n) in enumerate(shapes):
        rc = bt.recompressed_block(b)
        assert rc["C"].shape == (m, ranks[b]) and rc["R"].shape == (ranks[b], n)
        _check_block(oracle, bt.block(b), rc, 1e-6, 1e-9, 1e-10)
    bt.close()
    ctx.close()


def test_recompress_needs_a_result(farcur, sphere):
    ctx, (ro, ri, co, cix) = sphere
    bt = farcur.Batch(ctx, ro, ri, co, cix)
    with pytest.raises(farcur.FarcurError):
        bt.recompress(1e-5)
    bt.compress(farcur.MODE_BATCHED, tol=1e-4, seed=1)
    with pytest.raises(farcur.FarcurError):
        bt.recompress(-1.0)
    bt.recompress(1e-5)
    # a new compress invalidates the recompression
    bt.compress(farcur.MODE_BATCHED, tol=1e-4, seed=2)
    with pytest.raises(farcur.FarcurError):
        bt.recompressed_block(0)


@pytest.mark.parametrize("precision", [0, 1])
def test_recompress_cores_above_64(farcur, oracle, precision):
    # Cores with r > 64 take the recompression's second launch (streamed
    # Jacobi, register-free): a 400 x 400 plate block started at r = 100 ends
    # its reference loop at r = 200.
    mesh = farcur.plates_config(1.0, 20, 4.0)
    ctx = farcur.Context(0)
    ctx.set_mesh(mesh)
    rng = np.random.default_rng(11)
    rows = np.sort(rng.choice(1160, 400, replace=False)).astype(np.int32)
    cols = (1160 + np.sort(rng.choice(1160, 400, replace=False))).astype(np.int32)
    ro = np.array([0, 400], np.int32)
    bt = farcur.Batch(ctx, ro, rows, ro.copy(), cols)
    opts = dict(tol=1e-4, seed=5, precision=precision, r_start=100)
    if precision == farcur.F32:
        opts["pinv_rtol"] = 1e-6
    bt.compress(farcur.MODE_BATCHED, **opts)
    info = bt.info_arrays()
    assert info["r"][0] > 64
    for thresh in (1e-5, 1e-8):
        ranks, _ = bt.recompress(thresh)
        prod_tol, orth_tol = (1e-9, 1e-10) if precision == farcur.F64 else (2e-6, 2e-6)
        _check_block(oracle, bt.block(0), bt.recompressed_block(0), thresh, prod_tol, orth_tol)
    bt.close()
    ctx.close()
