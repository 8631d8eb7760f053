"""A batch of hand-made ragged blocks between two plates (1160 RWGs each,
4 lambda apart): single rows and columns, thin strips, blocks one past the
64-wide tile edge, unsorted ids and the full 1160 x 1160 block, through every
batch mode, against the oracle's adaptive loop and the dense block."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K, ETA = 2 * np.pi, 120 * np.pi
SHAPES = [(1, 1), (1, 40), (40, 1), (2, 300), (300, 3), (64, 64), (65, 200), (129, 65), (1160, 1160)]


@pytest.fixture(scope="module")
def ragged(farcur, oracle):
    mesh = farcur.plates_config(1.0, 20, 4.0)
    t = mesh.tables()
    ctx = farcur.Context(0)
    ctx.set_mesh(mesh)
    g = oracle.Geometry(t["vertex"], t["topol"], t["edges"], t["ln"], K, ETA)
    rng = np.random.default_rng(7)
    rows, cols = [], []
    for i, (m, n) in enumerate(SHAPES):
        r = rng.choice(1160, m, replace=False).astype(np.int32)
        c = (1160 + rng.choice(1160, n, replace=False)).astype(np.int32)
        if i % 2 == 0:  # every other block in sorted order
            r.sort()
            c.sort()
        rows.append(r)
        cols.append(c)
    ro = np.concatenate([[0], np.cumsum([r.size for r in rows])]).astype(np.int32)
    co = np.concatenate([[0], np.cumsum([c.size for c in cols])]).astype(np.int32)
    blocks = (ro, np.concatenate(rows), co, np.concatenate(cols))
    yield ctx, g, blocks, [g.fill(r, c) for r, c in zip(rows, cols)]
    ctx.close()


def _relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_ragged_batched_matches_oracle(farcur, ragged, oracle):
    ctx, g, (ro, ri, co, cix), dense = ragged
    bt = farcur.Batch(ctx, ro, ri, co, cix)
    bt.compress(farcur.MODE_BATCHED, tol=1e-4, seed=31)
    info = bt.info_arrays()
    for b, (m, n) in enumerate(SHAPES):
        rows, cols = ri[ro[b]:ro[b + 1]], cix[co[b]:co[b + 1]]
        ref = oracle.compress_block(g, rows, cols, tol=1e-4, seed=31, block_id=b)
        out = bt.block(b)
        # integer work (sample counts, rounds, sampled ids) bit-exact
        assert (info["r"][b], info["rounds"][b]) == (ref["r"], ref["rounds"]), (m, n)
        np.testing.assert_array_equal(out["cs"], cols[ref["cs"]])
        np.testing.assert_array_equal(out["rs"], rows[ref["rs"]])
        A = out["C"] @ out["U"] @ out["R"]
        Ao = ref["C"] @ ref["U"] @ ref["R"]
        # same bar as the sphere batch: FP64 rounding of two Jacobi orders,
        # amplified by |U| up to ~1e8 at the absolute 1e-10 threshold
        assert _relfro(A, Ao) <= 1e-6, (m, n)
        assert abs(_relfro(A, dense[b]) - _relfro(Ao, dense[b])) <= 1e-6, (m, n)


def test_ragged_plain_equals_batched(farcur, ragged):
    ctx, g, (ro, ri, co, cix), _ = ragged
    b1 = farcur.Batch(ctx, ro, ri, co, cix)
    b2 = farcur.Batch(ctx, ro, ri, co, cix)
    b1.compress(farcur.MODE_PLAIN, tol=1e-4, seed=33)
    b2.compress(farcur.MODE_BATCHED, tol=1e-4, seed=33)
    i1, i2 = b1.info_arrays(), b2.info_arrays()
    for k in ("r", "rounds", "kept", "converged", "err"):
        np.testing.assert_array_equal(i1[k], i2[k])
    for b in range(len(SHAPES)):
        x, y = b1.block(b), b2.block(b)
        for k in ("C", "U", "R", "rs", "cs"):
            np.testing.assert_array_equal(x[k], y[k])


@pytest.mark.parametrize("mode", [1, 2])
def test_ragged_f32_every_block_accurate(farcur, ragged, mode):
    # FP32 panels with the relative pinv threshold, batched and learned
    # (held-out rows where the block has rows to spare): every block's CUR
    # against the dense FP64 block.
    ctx, g, (ro, ri, co, cix), dense = ragged
    bt = farcur.Batch(ctx, ro, ri, co, cix)
    bt.compress(mode, tol=1e-4, seed=35, precision=farcur.F32, pinv_rtol=1e-6)
    info = bt.info_arrays()
    for b, (m, n) in enumerate(SHAPES):
        out = bt.block(b)
        assert out["C"].shape == (m, info["r"][b]) and out["R"].shape == (info["r"][b], n)
        assert 1 <= info["r"][b] <= min(m, n)
        e = _relfro(out["C"] @ out["U"] @ out["R"], dense[b])
        assert e <= 1e-4, (mode, m, n, e, info["r"][b])
