"""Rank predictor on the GPU: gradients against the CPU oracle, training,
held-out-row sampling parity, and the learned mode against the same
held-out-checked method without the network."""
import importlib.util
import sys

import numpy as np
import pytest

import mlp

pytestmark = pytest.mark.gpu


def _rank_model(farcur):
    if "rank_model" in sys.modules:
        return sys.modules["rank_model"]
    spec = importlib.util.spec_from_file_location("rank_model", farcur.HERE / "rank_model.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rank_model"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def gctx(farcur):
    c = farcur.Context(0)
    yield c
    c.close()


def _data(seed=0, n=300):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((n, 8)) * [0.3, 0.3, 2.0, 0.5, 0.5, 0.4, 1.0, 1.0]
         + [0.8, 0.8, 3.0, 0, 0, 0.6, 4.5, 4.5]).astype(np.float32)
    y = (2.5 + 0.3 * X[:, 0] + 0.3 * X[:, 1] - 0.2 * np.log(X[:, 2] ** 2 + 1) + 0.05 * rng.standard_normal(n))
    return X, y.astype(np.float32)


def test_gradient_matches_oracle(farcur, gctx):
    X, y = _data()
    model = farcur.Model(gctx, seed=3)
    norm = np.concatenate([X.mean(0), X.std(0), [y.mean(), y.std()]]).astype(np.float32)
    model.set(norm=norm)
    p, nrm = model.get()
    loss, g = model.loss_grad(X, y)
    lo, go = mlp.loss_grad(p, nrm, X, y)
    assert loss == pytest.approx(lo, rel=1e-4)
    # FP32 kernels against the FP64 oracle
    assert np.linalg.norm(g - go) <= 1e-4 * np.linalg.norm(go)
    np.testing.assert_allclose(model.predict(X), mlp.forward(p, nrm, X), rtol=1e-5, atol=1e-5)


def test_training_reduces_loss(farcur, gctx):
    X, y = _data(1, 4000)
    model = farcur.Model(gctx, seed=5)
    losses = model.train(X, y, epochs=30, batch=512, lr=3e-3, seed=2)
    assert losses[-1] < 0.25 * losses[0]
    p, nrm = model.get()
    # the trained parameters give the oracle the same loss
    lo, _ = mlp.loss_grad(p, nrm, X, y)
    l_gpu, _ = model.loss_grad(X, y)
    assert l_gpu == pytest.approx(lo, rel=1e-3)


@pytest.fixture(scope="module")
def ellipsoid(farcur, oracle):
    # training geometry and a different deployment geometry
    mesh = farcur.Mesh.icosphere(1.6, 1.2, 1.0, 4, 0.01, 7).build_rwg()
    part = farcur.Partition(mesh, 0.5, 2.0)
    t = mesh.tables()
    g = oracle.Geometry(t["vertex"], t["topol"], t["edges"], t["ln"])
    return mesh, part, g


def test_learned_mode_heldout_sampling_parity(farcur, gctx, ellipsoid, oracle):
    mesh, part, g = ellipsoid
    gctx.set_mesh(mesh)
    ro, ri, co, cix = part.far_blocks()
    bt = farcur.Batch(gctx, ro, ri, co, cix)
    bt.compress(farcur.MODE_LEARNED, tol=1e-4, seed=11, heldout=8)  # no model: reference r0 rule
    info = bt.info_arrays()
    checked = 0
    for b in range(0, bt.nblocks, max(1, bt.nblocks // 25)):
        m, n, r = ro[b + 1] - ro[b], co[b + 1] - co[b], info["r"][b]
        last = info["rounds"][b] - 1  # round index of the final samples
        rows, cols = ri[ro[b]:ro[b + 1]], cix[co[b]:co[b + 1]]
        h = max(0, min(min(max(8, -(-m // 32)), 32), m - r))  # held-out rows: >= 8, one per 32 rows, <= 32
        hp, rsp = oracle.sample_heldout(11, b, last, m, h, r)
        out = bt.block(b)
        np.testing.assert_array_equal(out["rs"], rows[rsp])
        np.testing.assert_array_equal(out["cs"], cols[oracle.sample(11, b, last, 1, n, r)])
        # held-out error as the oracle computes it from the same panels (a
        # sample that took every column is stored exactly, U = R = I)
        if h > 0 and r < min(m, n):
            Zh = g.fill(rows[hp], cols)
            P = out["C"] @ out["U"]
            A = P[hp, :] @ out["R"]
            e = np.linalg.norm(Zh - A) / np.linalg.norm(Zh)
            assert e == pytest.approx(info["heldout_err"][b], rel=1e-3, abs=1e-9)
            # the acceptance estimate: held-out errors scaled by the
            # interpolation norms |P_i|^2 of every unsampled row
            p2 = np.sum(np.abs(P) ** 2, axis=1)
            unsampled = np.ones(m, bool)
            unsampled[rsp] = False
            ztot = np.linalg.norm(out["R"]) ** 2 + (m - r) / h * np.linalg.norm(Zh) ** 2
            est = np.sqrt(np.linalg.norm(Zh - A) ** 2 / p2[hp].sum() * p2[unsampled].sum() / ztot)
            assert est == pytest.approx(info["err"][b], rel=1e-3, abs=1e-9)
            acc = farcur.default_opts().heldout_safety * 1e-4
            assert info["converged"][b] == (est <= acc) or info["r"][b] == min(m, n) or abs(est - acc) < 1e-3 * acc
        checked += 1
    assert checked >= 5


def test_learned_gain_and_accuracy(farcur, gctx, ellipsoid):
    # Train on blocks of two other ellipsoids, deploy on this one: the network
    # must cut the rounds and filled entries of the held-out-checked method,
    # keep the accuracy, and report its miss rate.
    rm = _rank_model(farcur)
    Xs, ys = [], []
    for axes, seed in [((1.4, 1.4, 1.0), 21), ((2.0, 1.0, 1.0), 22), ((1.2, 1.8, 1.5), 23)]:
        X, y = rm.ellipsoid_blocks(farcur, gctx, axes, 4, seed)
        Xs.append(X)
        ys.append(y)
    X, y = np.concatenate(Xs), np.concatenate(ys)
    model, losses = rm.train(farcur, gctx, X, y, epochs=60, batch=1024, lr=3e-3, log=lambda s: None)
    assert losses[-1] < losses[0]

    mesh, part, g = ellipsoid
    gctx.set_mesh(mesh)
    ro, ri, co, cix = part.far_blocks()
    base = farcur.Batch(gctx, ro, ri, co, cix)
    learned = farcur.Batch(gctx, ro, ri, co, cix)
    learned.set_model(model)
    opts = dict(tol=1e-4, seed=3, precision=farcur.F32, pinv_rtol=1e-6)
    base.compress(farcur.MODE_LEARNED, **opts)
    learned.compress(farcur.MODE_LEARNED, **opts)
    ib, il = base.info_arrays(), learned.info_arrays()
    m, n = np.diff(ro), np.diff(co)
    filled = lambda info: np.sum((m + n) * info["r"] + info["r"] ** 2)
    assert il["rounds"].mean() < ib["rounds"].mean()
    assert learned.misses() / learned.nblocks < 0.5
    assert np.all(il["converged"] == 1)
    # every block against its FP64 dense block, on the device
    errs, _, _, _ = learned.verify(np.ones((mesh.sizes()[2], 1), complex), want_product=False)
    assert np.max(errs) <= 1e-4
    # and a sample on the host, with the oracle's dense blocks
    for b in range(0, learned.nblocks, max(1, learned.nblocks // 20)):
        out = learned.block(b)
        Z = g.fill(ri[ro[b]:ro[b + 1]], cix[co[b]:co[b + 1]])
        e = np.linalg.norm(out["C"] @ out["U"] @ out["R"] - Z) / np.linalg.norm(Z)
        assert e <= 1e-4 and abs(e - errs[b]) <= 1e-6 * e + 1e-9
    print(f"rounds {ib['rounds'].mean():.2f} -> {il['rounds'].mean():.2f}, filled {filled(ib):.3e} -> {filled(il):.3e}, "
          f"miss rate {learned.misses() / learned.nblocks:.3f}, max err {np.max(errs):.2e}")


@pytest.mark.parametrize("r,k", [(12, 12), (24, 1), (30, 2), (32, 20), (40, 33), (64, 50), (96, 60)])
def test_pinv_qr_matches_oracle_on_planted_ranks(farcur, gctx, oracle, r, k):
    # Cores of exact rank k plus 1e-13 noise: the relative threshold 1e-8
    # keeps k columns. Covers the LDS classes from 16 to 96 (QR steps in LDS),
    # k = 1 and 2 (the in-place Q formation's first steps), k = r (U = W^-1).
    rng = np.random.default_rng(7000 + 100 * r + k)
    A = rng.standard_normal((r, k)) + 1j * rng.standard_normal((r, k))
    B = rng.standard_normal((k, r)) + 1j * rng.standard_normal((k, r))
    W = A @ B + 1e-13 * (rng.standard_normal((r, r)) + 1j * rng.standard_normal((r, r)))
    U, d, kept = gctx.pinv_qr(W, atol=0.0, rtol=1e-8)
    Uo, do, ko, perm = oracle.pinv_qr_trunc(W, atol=0.0, rtol=1e-8)
    assert kept == ko == k
    # the kept diagonal and pivots must agree; past k the noise columns'
    # order is not fixed by the data
    np.testing.assert_allclose(d[:k], do[:k], rtol=1e-10)
    assert np.all(d[k:] < 1e-8 * d[0])
    assert np.linalg.norm(U - Uo) <= 1e-9 * np.linalg.norm(Uo)
    # interpolation on the pivot columns: (U W)[:, c] = e_c
    UW = U @ W
    for c in perm[:k]:
        e = np.zeros(r)
        e[c] = 1.0
        assert np.linalg.norm(UW[:, c] - e) < 1e-8
    if k == r:
        np.testing.assert_allclose(U, np.linalg.inv(W), rtol=0, atol=1e-10 * np.abs(np.linalg.inv(W)).max())


def test_pinv_qr_rejects_what_its_form_cannot_take(farcur, gctx):
    W = np.eye(8, dtype=np.complex128)
    with pytest.raises(farcur.FarcurError):
        gctx.pinv_qr(W, atol=0.0, rtol=0.0)  # relative thresholds below 1e-8
    with pytest.raises(farcur.FarcurError):
        gctx.pinv_qr(np.eye(97, dtype=np.complex128))  # above the largest LDS class


def test_sampling_parity_in_large_and_small_rounds(farcur, gctx, oracle):
    # Rounds of more than 4,096 jobs sample with a thread per scan, smaller
    # rounds with a wave per scan (a scalar decision chain over 64 lanes):
    # both against the oracle's scans, on the 31,166 far blocks of the
    # 30,720-RWG sphere (without the network, nearly every block takes a
    # second round; the few that need a third are a small round).
    mesh = farcur.Mesh.icosphere(2.0, 2.0, 2.0, 5, 0.01, 1).build_rwg()
    part = farcur.Partition(mesh, 0.5, 2.0)
    gctx.set_mesh(mesh)
    ro, ri, co, cix = part.far_blocks()
    nb = len(ro) - 1
    assert nb > 4096
    bt = farcur.Batch(gctx, ro, ri, co, cix)
    seed = 11
    bt.compress(farcur.MODE_LEARNED, tol=1e-4, seed=seed, precision=farcur.F32, pinv_rtol=1e-6)  # no model: r0 = D/10
    info = bt.info_arrays()
    rounds = info["rounds"]
    heldout = farcur.default_opts().heldout
    # jobs of the round that drew each block's final samples
    jobs_in_round = np.array([np.sum(rounds >= k) for k in range(1, rounds.max() + 1)])
    final_jobs = jobs_in_round[rounds - 1]
    large = np.flatnonzero(final_jobs > 4096)
    small = np.flatnonzero(final_jobs <= 4096)
    assert large.size and small.size, jobs_in_round
    picks = list(large[:: max(1, large.size // 12)][:12]) + list(small[:: max(1, small.size // 12)][:12])
    for b in picks:
        b = int(b)
        m, n, r = ro[b + 1] - ro[b], co[b + 1] - co[b], int(info["r"][b])
        last = int(rounds[b]) - 1
        h = max(0, min(min(max(heldout, -(-m // 32)), 4 * heldout), m - r))
        hp, rsp = oracle.sample_heldout(seed, b, last, m, h, r)
        rows, cols = ri[ro[b]:ro[b + 1]], cix[co[b]:co[b + 1]]
        out = bt.block(b)
        np.testing.assert_array_equal(out["rs"], rows[rsp])
        np.testing.assert_array_equal(out["cs"], cols[oracle.sample(seed, b, last, 1, n, r)])
    bt.close()


def test_learned_core_is_the_truncated_qr(farcur, gctx, ellipsoid, oracle):
    # Learned rounds invert their cores by truncated QR with column pivoting
    # (FARCUR_CORE_QR): each sampled block's U against the oracle's restatement
    # on the same gathered core W = C[rpos, :]; on every sampled block, the
    # pivot columns are interpolated exactly (U W e_c = e_c for k columns).
    mesh, part, g = ellipsoid
    gctx.set_mesh(mesh)
    ro, ri, co, cix = part.far_blocks()
    bt = farcur.Batch(gctx, ro, ri, co, cix)
    bt.compress(farcur.MODE_LEARNED, tol=1e-4, seed=13, precision=farcur.F32, pinv_rtol=1e-6)
    info = bt.info_arrays()
    checked = agree = 0
    for b in range(0, bt.nblocks, max(1, bt.nblocks // 60)):
        m, n, r = ro[b + 1] - ro[b], co[b + 1] - co[b], info["r"][b]
        if r >= min(m, n) or r > 96:
            continue  # stored exactly, or an SVD core above the LDS classes
        out = bt.block(b)
        at = {int(x): i for i, x in enumerate(ri[ro[b]:ro[b + 1]])}
        W = out["C"][[at[int(x)] for x in out["rs"]], :].astype(np.complex128)
        Uo, d, k, perm = oracle.pinv_qr_trunc(W, 1e-10, 1e-6)
        U = out["U"]
        assert info["kept"][b] >= 1
        UW = U @ W
        exact = np.sum(np.linalg.norm(UW - np.eye(r), axis=0) <= 1e-6 * max(1.0, np.linalg.norm(UW, 2)))
        assert exact >= info["kept"][b], (b, exact, info["kept"][b])
        checked += 1
        if k == info["kept"][b] and np.linalg.norm(U - Uo) <= 1e-7 * np.linalg.norm(Uo):
            agree += 1
    assert checked >= 20 and agree >= 0.9 * checked, (checked, agree)


def test_learned_svd_core_option(farcur, gctx, ellipsoid):
    # FARCUR_CORE_SVD keeps the SVD pseudo-inverse in learned rounds: a
    # different U, and every block still within tol of its FP64 dense block.
    mesh, part, g = ellipsoid
    gctx.set_mesh(mesh)
    ro, ri, co, cix = part.far_blocks()
    a = farcur.Batch(gctx, ro, ri, co, cix)
    s = farcur.Batch(gctx, ro, ri, co, cix)
    opts = dict(tol=1e-4, seed=13, precision=farcur.F32, pinv_rtol=1e-6)
    a.compress(farcur.MODE_LEARNED, **opts)
    s.compress(farcur.MODE_LEARNED, learned_core=farcur.CORE_SVD, **opts)
    assert np.all(s.info_arrays()["sweeps"] > 0) or np.any(a.info_arrays()["sweeps"] != s.info_arrays()["sweeps"])
    errs, _, _, _ = s.verify(np.ones((mesh.sizes()[2], 1), complex), want_product=False)
    assert np.max(errs) <= 1e-4
    with pytest.raises(farcur.FarcurError, match="learned_core"):
        s.compress(farcur.MODE_LEARNED, learned_core=7, **opts)
