"""FP32 panels of a batch filled through shared triangles in the per-block
tile order (blocks of more than 64 RWGs, where the order is not the list
order), entry by entry against the FP64 oracle element."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K, ETA = 2 * np.pi, 120 * np.pi


@pytest.fixture(scope="module")
def sphere4(farcur, oracle):
    # Radius 1 lambda icosphere, level 4 (7680 RWGs), 0.5 lambda leaves: most
    # far blocks have 100-175 rows and columns.
    mesh = farcur.Mesh.icosphere(1.0, 1.0, 1.0, 4, 0.01, 1).build_rwg()
    part = farcur.Partition(mesh, 0.5, 2.0)
    t = mesh.tables()
    ctx = farcur.Context(0)
    ctx.set_mesh(mesh)
    g = oracle.Geometry(t["vertex"], t["topol"], t["edges"], t["ln"], K, ETA)
    yield ctx, g, part.far_blocks()
    ctx.close()


def _relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _relmax(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def test_batch_f32_panels_match_oracle(farcur, sphere4):
    ctx, g, (ro, ri, co, cix) = sphere4
    m, n = np.diff(ro), np.diff(co)
    assert (m > 64).sum() > 100 and (n > 64).sum() > 100
    bt = farcur.Batch(ctx, ro, ri, co, cix)
    bt.compress(farcur.MODE_BATCHED, tol=1e-4, seed=11, precision=farcur.F32, pinv_rtol=1e-6)
    big = np.flatnonzero((m > 64) & (n > 64))
    checked = 0
    for b in big[:: max(1, big.size // 12)]:
        rows, cols = ri[ro[b]:ro[b + 1]], cix[co[b]:co[b + 1]]
        out = bt.block(b)
        C64 = g.fill(rows, out["cs"])
        R64 = g.fill(out["rs"], cols)
        # same bars as the standalone FP32 fill (test_fill_f32_vs_f64); an entry
        # written to the wrong position would miss by O(1)
        assert _relfro(out["C"], C64) <= 2e-6, b
        assert _relmax(out["C"], C64) <= 5e-6, b
        assert _relfro(out["R"], R64) <= 2e-6, b
        assert _relmax(out["R"], R64) <= 5e-6, b
        checked += 1
    assert checked >= 10


def test_standalone_f32_fill_matches_batch_panel(farcur, sphere4):
    # The standalone fill takes the rows in list order, the batch in tile
    # order: different triangle sets per tile, the same entries to rounding.
    ctx, g, (ro, ri, co, cix) = sphere4
    bt = farcur.Batch(ctx, ro, ri, co, cix)
    bt.compress(farcur.MODE_BATCHED, tol=1e-4, seed=13, precision=farcur.F32, pinv_rtol=1e-6)
    b = int(np.argmax(np.diff(ro)))
    rows = ri[ro[b]:ro[b + 1]]
    out = bt.block(b)
    Cs = ctx.fill(rows, out["cs"], farcur.F32)
    assert _relfro(out["C"], Cs) <= 1e-6
