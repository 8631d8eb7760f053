"""Batched far-field compression (all admissible octree blocks of a sphere)
against the per-block path and the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K, ETA = 2 * np.pi, 120 * np.pi


@pytest.fixture(scope="module")
def sphere(farcur, oracle):
    # Radius 1 lambda icosphere, level 3 (1920 RWGs), 0.5 lambda leaves.
    mesh = farcur.Mesh.icosphere(1.0, 1.0, 1.0, 3, 0.01, 1).build_rwg()
    part = farcur.Partition(mesh, 0.5, 2.0)
    t = mesh.tables()
    ctx = farcur.Context(0)
    ctx.set_mesh(mesh)
    g = oracle.Geometry(t["vertex"], t["topol"], t["edges"], t["ln"], K, ETA)
    blocks = part.far_blocks()
    yield ctx, g, part, blocks
    ctx.close()


def _relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_partition_has_far_blocks(sphere):
    _, _, part, (ro, ri, co, cix) = sphere
    assert len(part.far_pairs) > 20


@pytest.mark.parametrize("row_mode", [0, 1])
def test_batched_equals_plain(farcur, sphere, row_mode):
    ctx, g, part, (ro, ri, co, cix) = sphere
    b1 = farcur.Batch(ctx, ro, ri, co, cix)
    b2 = farcur.Batch(ctx, ro, ri, co, cix)
    b1.compress(farcur.MODE_PLAIN, tol=1e-4, seed=9, row_mode=row_mode)
    b2.compress(farcur.MODE_BATCHED, tol=1e-4, seed=9, row_mode=row_mode)
    i1, i2 = b1.info_arrays(), b2.info_arrays()
    for k in ("r", "rounds", "kept", "converged"):
        np.testing.assert_array_equal(i1[k], i2[k])
    np.testing.assert_array_equal(i1["err"], i2["err"])
    for b in range(0, b1.nblocks, max(1, b1.nblocks // 7)):
        x, y = b1.block(b), b2.block(b)
        for k in ("C", "U", "R", "rs", "cs"):
            np.testing.assert_array_equal(x[k], y[k])


def test_batch_matches_oracle_and_dense(farcur, sphere, oracle):
    ctx, g, part, (ro, ri, co, cix) = sphere
    bt = farcur.Batch(ctx, ro, ri, co, cix)
    bt.compress(farcur.MODE_BATCHED, tol=1e-4, seed=21)
    info = bt.info_arrays()
    worst = 0.0
    for b in range(0, bt.nblocks, max(1, bt.nblocks // 12)):
        rows, cols = ri[ro[b]:ro[b + 1]], cix[co[b]:co[b + 1]]
        ref = oracle.compress_block(g, rows, cols, tol=1e-4, seed=21, block_id=b)
        out = bt.block(b)
        assert (info["r"][b], info["rounds"][b]) == (ref["r"], ref["rounds"])
        np.testing.assert_array_equal(out["cs"], cols[ref["cs"]])
        np.testing.assert_array_equal(out["rs"], rows[ref["rs"]])
        A = out["C"] @ out["U"] @ out["R"]
        Ao = ref["C"] @ ref["U"] @ ref["R"]
        # W keeps singular values down to the absolute 1e-10 threshold, so |U|
        # reaches ~1e8 and FP64 rounding differences of the two Jacobi orders
        # are amplified accordingly; the approximations agree to 1e-6 and have
        # the same error against the dense block.
        assert _relfro(A, Ao) <= 1e-6
        Z = g.fill(rows, cols)
        assert abs(_relfro(A, Z) - _relfro(Ao, Z)) <= 1e-6
        worst = max(worst, _relfro(A, Z))
    assert worst <= 1e-4


def test_batch_f32_accuracy(farcur, sphere):
    ctx, g, part, (ro, ri, co, cix) = sphere
    bt = farcur.Batch(ctx, ro, ri, co, cix)
    bt.compress(farcur.MODE_BATCHED, tol=1e-4, seed=5, precision=farcur.F32, pinv_rtol=1e-6)
    errs = []
    for b in range(0, bt.nblocks, max(1, bt.nblocks // 10)):
        rows, cols = ri[ro[b]:ro[b + 1]], cix[co[b]:co[b + 1]]
        out = bt.block(b)
        errs.append(_relfro(out["C"] @ out["U"] @ out["R"], g.fill(rows, cols)))
    assert max(errs) <= 1e-4


def test_batch_profile_accounting(farcur, sphere):
    ctx, g, part, (ro, ri, co, cix) = sphere
    bt = farcur.Batch(ctx, ro, ri, co, cix)
    bt.profile(True)
    bt.compress(farcur.MODE_BATCHED, tol=1e-4, seed=5)
    kt = bt.kernel_times()
    assert {"sample", "fill", "pinv", "residual", "pinv_r16", "pinv_global"} <= set(kt)
    assert kt["pinv"]["ms"] == pytest.approx(sum(v["ms"] for k, v in kt.items() if k.startswith("pinv_")))
    assert kt["fill"]["ms"] > 0 and kt["fill"]["work"] > 0 and kt["fill"]["launches"] >= 2


@pytest.mark.parametrize("mode,precision", [(1, 1), (1, 0), (2, 1)])
def test_batch_every_block_accurate(farcur, sphere, mode, precision):
    # Every block of the batch (not a sample): a core pinv gone wrong anywhere
    # in the launch shows up as a block whose CUR misses the dense block.
    ctx, g, part, (ro, ri, co, cix) = sphere
    bt = farcur.Batch(ctx, ro, ri, co, cix)
    opts = dict(tol=1e-4, seed=17, precision=precision)  # the library's learned-check defaults
    if precision == farcur.F32:
        opts["pinv_rtol"] = 1e-6
    bt.compress(mode, **opts)
    info = bt.info_arrays()
    if mode == farcur.MODE_LEARNED:
        assert np.all(info["converged"] == 1)
    # (the reference loop ends blocks that reach r = min(m, n) as "stopped at
    # the cap"; their CUR is then exact up to the pinv threshold)
    V = np.random.default_rng(4).standard_normal((1920, 2)) + 0j
    fro, mv, _, _ = bt.verify(V, want_product=False)
    for b in range(bt.nblocks):
        rows, cols = ri[ro[b]:ro[b + 1]], cix[co[b]:co[b + 1]]
        out = bt.block(b)
        Z = g.fill(rows, cols)
        D = out["C"] @ out["U"] @ out["R"] - Z
        e = np.linalg.norm(D) / np.linalg.norm(Z)
        assert e <= 1e-4, (b, e, info["r"][b], info["kept"][b])
        # the device check agrees with this host computation: (C U) R on the
        # device and C (U R) here differ by the rounding of forming the
        # product, ~eps |C| |U| |R| (|U| reaches ~1e8 at the absolute 1e-10
        # pinv threshold)
        fwd = 64 * np.finfo(float).eps * np.prod([np.linalg.norm(out[x], 2) for x in ("C", "U", "R")])
        assert abs(fro[b] - e) <= 1e-6 * e + fwd / np.linalg.norm(Z), (b, fro[b], e)
        v = V[cols, 0]
        emv = np.linalg.norm(D @ v) / np.linalg.norm(Z @ v)
        assert abs(mv[b] - emv) <= 1e-6 * emv + fwd * np.linalg.norm(v) / np.linalg.norm(Z @ v), (b, mv[b], emv)
    assert fro.max() <= 1e-4


def test_verify_dense_product_matches_oracle(farcur, sphere):
    # The device's FP64 dense far-field product (both orientations of every
    # block) against the oracle's dense blocks times the same vectors.
    ctx, g, part, (ro, ri, co, cix) = sphere
    bt = farcur.Batch(ctx, ro, ri, co, cix)
    bt.compress(farcur.MODE_BATCHED, tol=1e-4, seed=2, precision=farcur.F32, pinv_rtol=1e-6)
    rng = np.random.default_rng(8)
    V = rng.standard_normal((1920, 5)) + 1j * rng.standard_normal((1920, 5))
    for symmetric in (False, True):
        _, _, Y, _ = bt.verify(V, symmetric=symmetric)
        Yo = np.zeros_like(V)
        for b in range(bt.nblocks):
            rows, cols = ri[ro[b]:ro[b + 1]], cix[co[b]:co[b + 1]]
            Z = g.fill(rows, cols)
            Yo[rows] += Z @ V[cols]
            if symmetric:
                Yo[cols] += Z.T @ V[rows]
        assert np.linalg.norm(Y - Yo) <= 1e-10 * np.linalg.norm(Yo)


@pytest.mark.parametrize("mode", [1, 2])
def test_pipeline_chunks_do_not_change_results(farcur, sphere, mode):
    # The two-stream chunked rounds run the same per-block work: identical
    # sample counts, residuals and panels with the pipeline on and off.
    ctx, g, part, (ro, ri, co, cix) = sphere
    assert len(ro) - 1 >= 4 * 64  # enough jobs for 4 chunks
    out = []
    for chunks in (1, 4):
        bt = farcur.Batch(ctx, ro, ri, co, cix)
        bt.set_pipeline(chunks)
        bt.compress(mode, tol=1e-4, seed=23, precision=farcur.F32, pinv_rtol=1e-6, heldout=8)
        out.append((bt.info_arrays(), [bt.block(b) for b in range(0, bt.nblocks, 17)]))
    (i1, b1), (i2, b2) = out
    for k in ("r", "rounds", "kept", "converged", "err"):
        np.testing.assert_array_equal(i1[k], i2[k])
    for x, y in zip(b1, b2):
        for k in ("C", "U", "R", "rs", "cs"):
            np.testing.assert_array_equal(x[k], y[k])


def test_round_graphs_replay_matches_eager(farcur, sphere):
    # Rounds captured as hipGraphs and replayed on later compresses give the
    # eager launches' results, every replay, with eager work in between.
    ctx, g, part, (ro, ri, co, cix) = sphere
    opts = dict(tol=1e-4, seed=29, precision=farcur.F32, pinv_rtol=1e-6, heldout=8)
    eager = farcur.Batch(ctx, ro, ri, co, cix)
    eager.set_graphs(False)
    eager.compress(farcur.MODE_LEARNED, **opts)
    ie = eager.info_arrays()
    pick = list(range(0, eager.nblocks, 13))
    be = [eager.block(b) for b in pick]
    assert eager.graph_stats() == (0, 0)
    gb = farcur.Batch(ctx, ro, ri, co, cix)
    rounds = int(ie["rounds"].max())
    for rep in range(3):
        gb.compress(farcur.MODE_LEARNED, **opts)
        ig = gb.info_arrays()
        for k in ("r", "rounds", "kept", "converged", "err", "heldout_err"):
            np.testing.assert_array_equal(ig[k], ie[k])
        for b, x in zip(pick, be):
            y = gb.block(b)
            for k in ("C", "U", "R", "rs", "cs"):
                np.testing.assert_array_equal(x[k], y[k])
        # eager work of the program between replays
        eager.compress(farcur.MODE_BATCHED, tol=1e-4, seed=3)
        ctx.fill(ri[ro[0]:ro[1]], cix[co[0]:co[1]])
    captures, replays = gb.graph_stats()
    assert captures == rounds and replays == 2 * rounds


def test_context_close_releases_its_batches_first(farcur, sphere):
    # A context of its own: closing it releases the batch and model created
    # on it (which hold its device and stream) before the context itself, so
    # a later use of the batch is an error rather than a use of freed state.
    _, _, _, (ro, ri, co, cix) = sphere
    ctx = farcur.Context(0)
    ctx.set_mesh(farcur.Mesh.icosphere(1.0, 1.0, 1.0, 3, 0.01, 1).build_rwg())
    bt = farcur.Batch(ctx, ro, ri, co, cix)
    model = farcur.Model(ctx, seed=3)
    bt.compress(farcur.MODE_BATCHED, tol=1e-4, seed=5, precision=farcur.F32, pinv_rtol=1e-6)
    ctx.close()
    assert not bt._h and not model._h
    with pytest.raises(farcur.FarcurError):
        bt.compress(farcur.MODE_BATCHED, tol=1e-4, seed=5)
    bt.close()
    model.close()


def test_round_graphs_across_changing_plans(farcur, sphere):
    # Calls whose rounds share launch parameters but not plans (the same
    # first round, different fallback rounds at another tolerance) replay
    # cached graphs with their own plans: every call equals an eager run.
    ctx, g, part, (ro, ri, co, cix) = sphere
    base = dict(seed=41, precision=farcur.F32, pinv_rtol=1e-6, heldout=8)
    eager = farcur.Batch(ctx, ro, ri, co, cix)
    eager.set_graphs(False)
    gb = farcur.Batch(ctx, ro, ri, co, cix)
    pick = list(range(0, gb.nblocks, 11))
    for tol in (1e-4, 3e-5, 1e-4, 3e-5):
        eager.compress(farcur.MODE_LEARNED, tol=tol, **base)
        gb.compress(farcur.MODE_LEARNED, tol=tol, **base)
        ie, ig = eager.info_arrays(), gb.info_arrays()
        for k in ("r", "rounds", "kept", "converged", "err", "heldout_err"):
            np.testing.assert_array_equal(ig[k], ie[k])
        for b in pick:
            x, y = eager.block(b), gb.block(b)
            for k in ("C", "U", "R", "rs", "cs"):
                np.testing.assert_array_equal(x[k], y[k])
    captures, replays = gb.graph_stats()
    assert replays > 0
