"""Near-field blocks (singularity extraction for close triangle pairs) and
the compressed operator (near dense + far CUR) with its GMRES solve, on the
GPU: element-wise against the oracle's FP64 near-field element, the operator
product against the oracle's dense matrix, and the solved currents of a PEC
sphere against the Mie series (the end-to-end pin of the element's sign,
normalisation and self terms)."""
import numpy as np
import pytest

import excitation
import mie

pytestmark = pytest.mark.gpu

K, ETA = 2 * np.pi, 120 * np.pi


def _relmax(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.fixture(scope="module")
def small_sphere(farcur, oracle):
    # radius 0.35 lambda, 1920 RWGs, jittered (no exactly tied pair distances)
    mesh = farcur.Mesh.icosphere(0.35, 0.35, 0.35, 3, 0.01, 2).build_rwg()
    t = mesh.tables()
    ctx = farcur.Context(0)
    ctx.set_mesh(mesh)
    g = oracle.Geometry(t["vertex"], t["topol"], t["edges"], t["ln"], K, ETA)
    yield mesh, t, ctx, g
    ctx.close()


def test_fill_near_matches_oracle(small_sphere):
    mesh, t, ctx, g = small_sphere
    rng = np.random.default_rng(1)
    rows = np.concatenate([np.arange(40), rng.choice(1920, 60, replace=False)]).astype(np.int32)
    cols = np.concatenate([np.arange(20, 70), rng.choice(1920, 50, replace=False)]).astype(np.int32)
    for nf in (2.0, 3.5):
        Z = ctx.fill_near(rows, cols, nf)
        Zo = g.fill_near(rows, cols, nf)
        assert np.all(np.isfinite(Z))
        assert _relmax(Z, Zo) <= 1e-10
    # far pairs only (near_factor 0): the far element exactly as the panels fill it
    far_rows, far_cols = np.arange(0, 10, dtype=np.int32), np.arange(1000, 1010, dtype=np.int32)
    np.testing.assert_allclose(ctx.fill_near(far_rows, far_cols, 0.0), ctx.fill(far_rows, far_cols), rtol=1e-12,
                               atol=1e-14 * np.abs(ctx.fill(far_rows, far_cols)).max())


def test_operator_matches_dense_and_solves(farcur, small_sphere, oracle):
    mesh, t, ctx, g = small_sphere
    part = farcur.Partition(mesh, 0.15, 2.0)
    assert len(part.far_pairs) > 50
    near = farcur.NearField.from_partition(ctx, part)
    # every near block against the oracle
    for b in range(0, near.nblocks, max(1, near.nblocks // 25)):
        a_, b_ = part.near_pairs[b]
        assert _relmax(near.block(b), g.fill_near(part.group(a_), part.group(b_))) <= 1e-10
    far = farcur.Batch(ctx, *part.far_blocks())
    far.compress(farcur.MODE_BATCHED, tol=1e-6, seed=3)
    op = farcur.Operator(near, far)
    N = 1920
    Zn = g.fill_near(np.arange(N), np.arange(N))
    # the operator stores one block per pair of groups and applies its
    # transpose for the other (the reference's self == 0); the near element's
    # closed-form inner and 7-point outer integrals are reciprocal only to
    # their quadrature error, so the reference matrix is assembled the same way
    Zd = np.zeros_like(Zn)
    for a, b in np.concatenate([part.near_pairs, part.far_pairs]):
        ra, cb = part.group(a), part.group(b)
        Zd[np.ix_(ra, cb)] = Zn[np.ix_(ra, cb)]
        if a != b:
            Zd[np.ix_(cb, ra)] = Zn[np.ix_(ra, cb)].T
    rng = np.random.default_rng(2)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    y = op.matvec(x)
    yd = Zd @ x
    # the far part is compressed to 1e-6 and applied in FP32
    assert np.linalg.norm(y - yd) <= 2e-5 * np.linalg.norm(yd)
    # GMRES on the device against the dense direct solve
    V = excitation.plane_wave(t["vertex"], t["edges"], t["ln"], t["cent"], K, 0.0, 0.0)
    # the far part runs in FP32: residuals below ~1e-7 are out of reach
    J, it, rr = op.solve(-V, tol=1e-6, restart=100, maxit=1000)
    assert rr[0] <= 1e-6 and 0 < it[0] < 1000
    Jd = np.linalg.solve(Zd, -V)
    assert np.linalg.norm(J - Jd) <= 1e-4 * np.linalg.norm(Jd)
    op.close()


def test_pec_sphere_rcs_through_the_compressed_operator(farcur):
    # A PEC sphere of radius 0.8515 lambda (ka = 5.35, between the interior
    # resonances at 4.97 and 5.76), 7680 RWGs: near blocks dense, far blocks
    # compressed by the learned/batched CUR in complex FP32, GMRES on the
    # device; the bistatic RCS in the E-plane against the Mie series.
    a = 0.8515
    mesh = farcur.Mesh.icosphere(a, a, a, 4, 0.0, 1).build_rwg()
    t = mesh.tables()
    ctx = farcur.Context(0)
    ctx.set_mesh(mesh)
    part = farcur.Partition(mesh, 0.25, 2.0)
    near = farcur.NearField.from_partition(ctx, part)
    far = farcur.Batch(ctx, *part.far_blocks())
    far.compress(farcur.MODE_LEARNED, tol=1e-4, seed=5, precision=farcur.F32, pinv_rtol=1e-6)
    op = farcur.Operator(near, far)
    V = mesh.plane_waves(np.array([0.0]), np.array([0.0]), k=K)[:, 0]
    J, it, rr = op.solve(-V, tol=1e-5, restart=100, maxit=2000)
    assert rr[0] <= 1e-5
    th = np.linspace(0, np.pi, 37)
    rhat = np.vstack([np.sin(th), np.zeros_like(th), np.cos(th)])
    s_mom = farcur.bistatic_rcs(mesh, J, K, rhat)
    s_mie = mie.rcs(K, a, np.pi - th, "E")
    db = 10 * np.log10(s_mom / s_mie)
    assert np.abs(db).max() <= 0.5, db
    # the product path agrees with the oracle's far-field integral
    s_orc = mie.mom_rcs(t["vertex"], t["topol"], t["edges"], t["ln"], J, K, ETA, rhat)
    np.testing.assert_allclose(s_mom, s_orc, rtol=1e-10)
    op.close()
    ctx.close()
