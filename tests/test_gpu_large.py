"""The large-run regime (config 5's shape at a size the oracle still
reaches): a 4 x 2 x 1 lambda ellipsoid, icosphere level 5 (30,720 RWGs),
1 lambda leaves so that blocks reach hundreds of RWGs, the held-out count
grows past 8 and cores above 64 take the global-memory pinv inside the
learned batch; every block against its FP64 dense block (device check,
itself spot-checked against the oracle) and the 360-plane-wave product
against the FP64 dense far-field product."""
import importlib.util
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K, ETA = 2 * np.pi, 120 * np.pi


def _rank_model(farcur):
    if "rank_model" in sys.modules:
        return sys.modules["rank_model"]
    spec = importlib.util.spec_from_file_location("rank_model", farcur.HERE / "rank_model.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rank_model"] = mod
    spec.loader.exec_module(mod)
    return mod


def test_large_ellipsoid_learned_every_block(farcur, oracle):
    mesh = farcur.Mesh.icosphere(4.0, 2.0, 1.0, 5, 0.01, 1).build_rwg()
    t = mesh.tables()
    N = mesh.sizes()[2]
    part = farcur.Partition(mesh, 1.0, 2.0)
    ro, ri, co, cix = part.far_blocks()
    m, n = np.diff(ro), np.diff(co)
    assert max(m.max(), n.max()) > 256  # held-out rows: one per 32 rows past 256
    ctx = farcur.Context(0)
    ctx.set_mesh(mesh)
    model = _rank_model(farcur).load(farcur, ctx)
    bt = farcur.Batch(ctx, ro, ri, co, cix)
    bt.set_model(model)
    bt.compress(farcur.MODE_LEARNED, tol=1e-4, seed=12345, precision=farcur.F32, pinv_rtol=1e-6)
    info = bt.info_arrays()
    assert info["r"].max() > 64  # some cores took the global-memory pinv
    assert np.all(info["status"] == 1)
    phi = 2 * np.pi * np.arange(360) / 360
    V = mesh.plane_waves(np.full(360, np.pi / 2), phi, k=K)
    sub = V[:, ::23][:, :16]
    fro, mv, Yref, _ = bt.verify(sub, symmetric=True)
    assert fro.max() <= 1e-4, (int(np.argmax(fro)), fro.max())
    # the device check on a few blocks (the largest included) against the oracle's dense blocks
    g = oracle.Geometry(t["vertex"], t["topol"], t["edges"], t["ln"], K, ETA)
    for b in list(np.argsort(m * n)[-3:]) + [0, len(m) // 2]:
        rows, cols = ri[ro[b]:ro[b + 1]], cix[co[b]:co[b + 1]]
        out = bt.block(int(b))
        Z = g.fill(rows, cols)
        e = np.linalg.norm(out["C"] @ out["U"] @ out["R"] - Z) / np.linalg.norm(Z)
        assert e <= 1e-4 and abs(e - fro[b]) <= 1e-6 * e + 1e-9
    # the compressed 360-wave product (FP32 MFMA and split BF16x3) against the FP64 dense far field
    for mode in (farcur.MV_MFMA_F32, farcur.MV_MFMA_BF16X3):
        Y, _ = bt.matvec(V, mode, True)
        Y = Y[:, ::23][:, :16]
        assert np.linalg.norm(Y - Yref) <= 1e-4 * np.linalg.norm(Yref)
    bt.close()
    ctx.close()
