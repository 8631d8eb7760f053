"""Far-field CUR benchmark: all admissible octree blocks of the 30,720-RWG
sphere (radius 2 lambda, icosphere level 5, +-1% radial jitter, 0.5 lambda
leaves, admissible when centre distance >= 2 x max group diameter),
compressed to a successive residual of 1e-4 with complex-FP32 fill and FP64
core pinv. One step = one compression of every far block.

Prints one JSON line (metric: blocks/s). Extra fields carry represented and
filled entries/s, the per-block plain-loop rate, the sampled accuracy against
FP64 dense blocks, the fill kernel's roofline and the CPU oracle baseline.
"""
import argparse
import json
import os
import pathlib
import platform
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
from __graft_entry__ import _load_farcur  # noqa: E402

# MI355X published peaks
PEAK_FP32_VALU_TFLOPS = 157.3
PEAK_FP64_VALU_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# Transcendental ceiling used for the fill (quarter-rate sine/cosine/rsqrt on
# 256 CUs x 64 lanes / 4 at 2.4 GHz).
PEAK_TRANSC_PER_S = 256 * 16 * 2.4e9
FLOPS_PER_EVAL = 22.0  # algorithmic FP32 flops of one Green-function evaluation


TRAFFIC_FILE = ROOT / "profiles" / "traffic_sphere_learned.json"


def measured_traffic(name):
    """HBM bytes per launch of a kernel class from the committed rocprofv3
    --pmc summary (FETCH_SIZE scaled by the plain-copy calibration, plus
    WRITE_SIZE), or None. The numbers come from a separate profiling run of
    this file; see that summary for its command."""
    try:
        d = json.loads(TRAFFIC_FILE.read_text())
        k = d["kernels"][name]
        return {"bytes_per_launch": k["bytes_per_launch"], "fetch_bytes_per_launch": k["fetch_bytes_per_launch"],
                "write_bytes_per_launch": k["write_bytes_per_launch"], "source": str(TRAFFIC_FILE.relative_to(ROOT))}
    except (OSError, KeyError, ValueError):
        return None


def roofline(kt):
    """Roofline of the kernel class with the largest device time in the
    profiled step (pinv counted per r class, each class being its own launch).
    achieved = algorithmic work / device time, both summed over the class's
    launches of the step (= per-launch work / average launch duration)."""
    main = {k: v for k, v in kt.items() if k in ("fill", "residual") or (k.startswith("pinv_") and v["launches"])}
    name = max(main, key=lambda k: main[k]["ms"])
    v = main[name]
    sec = v["ms"] * 1e-3
    common = {"kernel": name, "launches": v["launches"], "avg_launch_ms": v["ms"] / max(v["launches"], 1)}
    traffic = measured_traffic(name)
    if name == "fill":
        achieved = v["work"] * FLOPS_PER_EVAL / sec / 1e12
        return {**common, "bound": "mfma", "compute_unit": "VALU FP32 + transcendental (no MFMA in the fill)",
                "achieved": achieved, "peak": PEAK_FP32_VALU_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_VALU_TFLOPS,
                "work_unit": "Green evaluations x 22 flops (9 x distinct triangle pairs per panel)",
                "evals_per_s": v["work"] / sec,
                "evals_frac_of_transcendental_ceiling": v["work"] / sec / (PEAK_TRANSC_PER_S / 3.0),
                "traffic": traffic["bytes_per_launch"] if traffic else None, "traffic_detail": traffic,
                "note": "compute roof: MI355X publishes 157.3 TFLOP/s for FP32 (vector and matrix); packed FP32 "
                        "is disabled in this build, so the VALU tops out at half of it"}
    if name.startswith("pinv"):
        achieved = v["work"] / sec / 1e12
        return {**common, "bound": "mfma", "compute_unit": "VALU FP64 (no MFMA in the pinv)", "achieved": achieved,
                "peak": PEAK_FP64_VALU_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_VALU_TFLOPS,
                "work_unit": "96 r^3 FP64 flops per r x r complex SVD-pinv",
                "traffic": traffic["bytes_per_launch"] if traffic else None, "traffic_detail": traffic,
                "note": "compute roof: 78.6 TFLOP/s FP64 (vector and matrix)"}
    achieved = v["work"] / sec / 1e9
    return {**common, "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "work_unit": "panel + vector bytes read",
            "traffic": traffic["bytes_per_launch"] if traffic else None, "traffic_detail": traffic}


def build_workload(fc, level=5, radius=2.0, jitter=0.01, seed=1, leaf=0.5, eta_adm=2.0):
    mesh = fc.Mesh.icosphere(radius, radius, radius, level, jitter, seed).build_rwg()
    part = fc.Partition(mesh, leaf, eta_adm)
    return mesh, part


def cpu_model():
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(mesh, blocks, opts, budget_s=12.0, budget_1core_s=8.0, seed=0):
    """Oracle (FP64 restatement of the reference's CPU CUR loop, OpenMP over
    panel columns) on a uniform random sample of the far blocks (seeded), once
    on all host threads and once on one thread, each until its time budget is
    used; blocks/s of each and the sample description."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc

    t = mesh.tables()
    g = orc.Geometry(t["vertex"], t["topol"], t["edges"], t["ln"])
    ro, ri, co, cix = blocks
    nb = len(ro) - 1
    order = np.random.default_rng(seed).permutation(nb)

    def run(budget):
        done, t0 = 0, time.perf_counter()
        while done < nb and time.perf_counter() - t0 < budget:
            b = int(order[done])
            orc.compress_block(g, ri[ro[b]:ro[b + 1]], cix[co[b]:co[b + 1]], tol=opts["tol"], seed=opts["seed"],
                               block_id=b)
            done += 1
        return done, time.perf_counter() - t0

    all_threads = orc.num_threads()
    done, dt = run(budget_s)
    orc.set_threads(1)
    done1, dt1 = run(budget_1core_s)
    orc.set_threads(all_threads)
    m, n = np.diff(ro)[order[:done]], np.diff(co)[order[:done]]
    return {"value": done / dt, "unit": "blocks/s", "cores": all_threads, "kind": "port",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "one_core": {"value": done1 / dt1, "unit": "blocks/s", "cores": 1,
                         "sample": f"first {done1} blocks of the same uniform random order, {dt1:.1f} s"},
            "sample": f"{done} of {nb} far blocks drawn uniformly at random (seed {seed}; sizes {m.min()}-"
                      f"{max(m.max(), n.max())} RWGs, mean {m.mean():.0f} x {n.mean():.0f}), FP64 oracle, "
                      f"{dt:.1f} s on {all_threads} threads"}


def accuracy(bt, v):
    """Every block of the last compress against its FP64 dense block, on the
    device: relative Frobenius error and matvec error for the vector v."""
    fro, mv, _, ms = bt.verify(v[:, None], want_product=False)
    info = bt.info_arrays()
    return {"blocks_checked": int(fro.size), "max_block_relerr": float(fro.max()),
            "median_block_relerr": float(np.median(fro)), "p99_block_relerr": float(np.quantile(fro, 0.99)),
            "blocks_above_1e-4": int(np.sum(fro > 1e-4)), "max_block_matvec_relerr": float(mv.max()),
            "median_block_matvec_relerr": float(np.median(mv)),
            "blocks_unconverged": int(np.sum(info["status"] != 1)), "verify_ms": ms}


def recompress_bench(fc, ctx, batch, blocks, thresh, nsample, log):
    """SVD recompression of every block of the last compress at `thresh`
    (the reference's matlab_postcompress(C, U, R, 0.1 * err)): device time,
    ranks and stored entries before and after, accuracy on sampled blocks."""
    ro, ri, co, cix = blocks
    nb = len(ro) - 1
    r = batch.info_arrays()["r"]
    batch.recompress(thresh)
    runs = [batch.recompress(thresh) for _ in range(3)]
    ranks, ms = runs[-1][0], float(np.mean([x[1] for x in runs]))
    m, n = np.diff(ro), np.diff(co)
    before = float(np.sum((m + n) * r + r * r))
    after = float(np.sum((m + n) * ranks + ranks))
    rng = np.random.default_rng(1)
    errs = []
    for b in rng.choice(nb, min(nsample, nb), replace=False):
        rc = batch.recompressed_block(int(b))
        Z = ctx.fill(ri[ro[b]:ro[b + 1]], cix[co[b]:co[b + 1]], fc.F64)
        errs.append(np.linalg.norm((rc["C"] * rc["s"]) @ rc["R"] - Z) / np.linalg.norm(Z))
    log(f"# recompression at {thresh:g}: {ms:.2f} ms, rank mean {r.mean():.1f} -> {ranks.mean():.1f}, "
        f"stored entries {before:.3e} -> {after:.3e}, accuracy max {max(errs):.2e}")
    return {"thresh": thresh, "ms": ms, "rank_mean_before": float(r.mean()), "rank_mean_after": float(ranks.mean()),
            "stored_entries_before": before, "stored_entries_after": after,
            "max_block_relerr_sampled": float(max(errs)), "median_block_relerr_sampled": float(np.median(errs))}


def aca_bench(fc, ctx, batch, blocks, tol, max_rank, nsample, log):
    """The reference's comparator: partially pivoted ACA of every block with
    the FP64 element at `tol`; device time, ranks and sampled accuracy."""
    ro, ri, co, cix = blocks
    nb = len(ro) - 1
    batch.aca(tol, max_rank)
    runs = [batch.aca(tol, max_rank) for _ in range(2)]
    ranks, ms = runs[-1][0], float(np.mean([x[1] for x in runs]))
    rng = np.random.default_rng(2)
    errs = []
    for b in rng.choice(nb, min(nsample, nb), replace=False):
        out = batch.aca_block(int(b))
        Z = ctx.fill(ri[ro[b]:ro[b + 1]], cix[co[b]:co[b + 1]], fc.F64)
        errs.append(np.linalg.norm(out["U"] @ out["V"] - Z) / np.linalg.norm(Z))
    log(f"# ACA comparator at {tol:g}: {ms:.2f} ms ({nb / (ms * 1e-3):.0f} blocks/s), rank mean {ranks.mean():.1f}, "
        f"accuracy max {max(errs):.2e} median {np.median(errs):.2e}")
    return {"tol": tol, "max_rank": max_rank, "ms": ms, "blocks_per_s": nb / (ms * 1e-3),
            "rank_mean": float(ranks.mean()), "max_block_relerr_sampled": float(max(errs)),
            "median_block_relerr_sampled": float(np.median(errs))}


def rank_model(fc, ctx, log):
    import importlib.util

    spec = importlib.util.spec_from_file_location("rank_model", ROOT / "gpu-randomized-pseudoskeleton_amd" / "rank_model.py")
    rm = importlib.util.module_from_spec(spec)
    sys.modules["rank_model"] = rm
    spec.loader.exec_module(rm)
    if rm.DEFAULT_MODEL.exists():
        log(f"# rank model: {rm.DEFAULT_MODEL.relative_to(ROOT)}")
        return rm.load(fc, ctx), str(rm.DEFAULT_MODEL.relative_to(ROOT))
    log("# rank model: no trained model file, training a small one (40k blocks)")
    X, y = rm.training_set(fc, ctx, 40_000, seed=1, levels=(4,), log=lambda s: None)
    model, _ = rm.train(fc, ctx, X, y, epochs=40, log=lambda s: None)
    return model, "trained in bench (40k blocks, level-4 ellipsoids)"


def product_bench(fc, bt, mesh, blocks, theta, phi, log, reps=3):
    """Compressed far field times the RWG-tested plane waves of (theta, phi),
    symmetric operator (both orientations of every block), in each pass-B
    arithmetic. Flops: 8 per complex multiply-add, pass A + pass B of both
    orientations = 16 s r (m + n) per block."""
    ro, ri, co, cix = blocks
    m, n = np.diff(ro), np.diff(co)
    r = bt.info_arrays()["r"].astype(np.float64)
    V = mesh.plane_waves(theta, phi, k=2 * np.pi)
    s = V.shape[1]
    flops = 16.0 * s * float(np.sum(r * (m + n)))
    out, Yref = {}, None
    for name, mode in (("mfma_f32", fc.MV_MFMA_F32), ("mfma_bf16x3", fc.MV_MFMA_BF16X3), ("valu_f32", fc.MV_VALU)):
        ts = []
        for _ in range(reps):
            Y, t = bt.matvec(V, mode, True)
            ts.append(t)
        t = np.array(ts)
        ta, tb = float(np.median(t[:, 0])), float(np.median(t[:, 1]))
        d = {"ms_pass_a": ta, "ms_pass_b": tb, "tflops": flops / ((ta + tb) * 1e-3) / 1e12,
             "ms_form_p": float(t[0, 2])}
        if Yref is None:
            Yref = Y
        else:
            d["rel_diff_vs_mfma_f32"] = float(np.linalg.norm(Y - Yref) / np.linalg.norm(Yref))
        out[name] = d
        log(f"# product {name}: {s} RHS, pass A {ta:.3f} ms, pass B {tb:.3f} ms, {d['tflops']:.1f} TFLOP/s")
    return {"nrhs": s, "flops": flops, "modes": out}


def matvec1_bench(bt, blocks, v, log, reps=20):
    """The single-RHS CUR matvec over every far block (one orientation, the
    reference's residual product y = C (U (R v)) per block), HBM-bound.
    Algorithmic bytes per block: the complex64 panels P = C U (m x r) and R
    (r x n) read once, v gathered (n), y accumulated (m read + m written)."""
    ro, ri, co, cix = blocks
    m, n = np.diff(ro).astype(np.float64), np.diff(co).astype(np.float64)
    r = bt.info_arrays()["r"].astype(np.float64)
    bt.matvec1(v, False)  # forms P = C U once
    ts = [bt.matvec1(v, False)[1] for _ in range(reps)]
    t = float(np.median(ts))
    nbytes = float(np.sum(8 * r * (m + n) + 8 * n + 16 * m))
    achieved = nbytes / (t * 1e-3) / 1e9
    traffic = measured_traffic("matvec1")
    log(f"# single-RHS CUR matvec: {t * 1e3:.1f} us, {nbytes / 1e6:.1f} MB algorithmic, {achieved:.0f} GB/s")
    return {"kernel": "k_matvec1", "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "ms": t, "bytes_per_launch": nbytes,
            "work_unit": "8 r (m + n) + 8 n + 16 m bytes per block (complex64 panels P = C U and R, vector gather, "
                         "y read-modify-write)",
            "traffic": traffic["bytes_per_launch"] if traffic else None, "traffic_detail": traffic}


def _group_errors(Y, Yref, groups):
    e = [np.linalg.norm(Y[g] - Yref[g]) / np.linalg.norm(Yref[g]) for g in groups if np.linalg.norm(Yref[g]) > 0]
    return float(np.max(e)), float(np.median(e))


def precision_study(fc, ctx, bt32, blocks, groups, V, opts, log, label, model=None):
    """Config 3 on one workload: per-block relative Frobenius and matvec error
    against FP64 dense blocks for the complex-FP32 fill (bt32, already
    compressed) and the FP64 fill, and the multi-RHS product's error per
    arithmetic (FP32 MFMA, split-BF16x3 MFMA, FP32 VALU) against the FP64
    dense far-field product, whole and per target group (the block rows that
    share their RWGs), on 16 of the plane waves."""
    ro, ri, co, cix = blocks
    sub = V[:, ::max(1, V.shape[1] // 16)][:, :16]
    out = {"fill_c64": accuracy(bt32, V[:, 0])}
    bt64 = fc.Batch(ctx, ro, ri, co, cix)
    if model is not None:
        bt64.set_model(model)
    bt64.compress(fc.MODE_LEARNED, **dict(opts, precision=fc.F64, pinv_rtol=0.0))
    out["fill_c128"] = accuracy(bt64, V[:, 0])
    _, _, Yref, vms = bt32.verify(sub, symmetric=True)
    prod = {}
    for name, mode in (("mfma_f32", fc.MV_MFMA_F32), ("mfma_bf16x3", fc.MV_MFMA_BF16X3), ("valu_f32", fc.MV_VALU)):
        Y, _ = bt32.matvec(V, mode, True)
        Y = Y[:, ::max(1, V.shape[1] // 16)][:, :16]
        gmax, gmed = _group_errors(Y, Yref, groups)
        prod[name] = {"relerr_vs_fp64_dense": float(np.linalg.norm(Y - Yref) / np.linalg.norm(Yref)),
                      "max_group_relerr": gmax, "median_group_relerr": gmed}
    Y64, _ = bt64.matvec(V, fc.MV_MFMA_F32, True)
    Y64 = Y64[:, ::max(1, V.shape[1] // 16)][:, :16]
    prod["mfma_f32_on_c128_fill"] = {"relerr_vs_fp64_dense": float(np.linalg.norm(Y64 - Yref) / np.linalg.norm(Yref))}
    out["product"] = prod
    out["dense_reference_ms"] = vms
    bt64.close()
    log(f"# precision study ({label}): fill c64 max {out['fill_c64']['max_block_relerr']:.2e} / c128 max "
        f"{out['fill_c128']['max_block_relerr']:.2e}; product vs FP64 dense: "
        + ", ".join(f"{k} {v['relerr_vs_fp64_dense']:.2e}" for k, v in prod.items()))
    return out


def plates_study(fc, log):
    """Config 1 (two 1 lambda plates, 20 x 20 squares, 4 lambda apart, one
    1160 x 1160 block) through the precision study: learned-mode rounds without
    a rank model (the reference's r0 rule, held-out check), FP32 and FP64 fill,
    256 plane waves (theta_j = j pi / 255, phi = 0)."""
    mesh = fc.plates_config(1.0, 20, 4.0)
    ctx = fc.Context(0)
    ctx.set_mesh(mesh)
    ro = np.array([0, 1160], np.int32)
    blocks = (ro, np.arange(1160, dtype=np.int32), ro.copy(), np.arange(1160, 2320, dtype=np.int32))
    bt = fc.Batch(ctx, *blocks)
    opts = dict(tol=1e-4, seed=12345, precision=fc.F32, pinv_rtol=1e-6)
    bt.compress(fc.MODE_LEARNED, **opts)
    V = mesh.plane_waves(np.arange(256) * np.pi / 255, np.zeros(256), k=2 * np.pi)
    groups = [np.arange(1160), np.arange(1160, 2320)]
    res = precision_study(fc, ctx, bt, blocks, groups, V, opts, log, "plates")
    res["r"] = int(bt.info_arrays()["r"][0])
    bt.close()
    ctx.close()
    return res


def solve_bench(fc, ctx, model, log, a, level, leaf, label, k=2 * np.pi):
    """The full operator on a PEC sphere of radius a: near-field blocks (dense,
    singularity extraction), far-field blocks (learned + batched CUR, complex
    FP32), GMRES(100) on the device to 1e-5 for an x-polarised plane wave
    along -z, and the bistatic RCS in the E-plane against the Mie series."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import mie

    mesh = fc.Mesh.icosphere(a, a, a, level, 0.0, 1).build_rwg()
    N = mesh.sizes()[2]
    part = fc.Partition(mesh, leaf, 2.0)
    ctx.set_em(k)
    ctx.set_mesh(mesh)
    t0 = time.perf_counter()
    near = fc.NearField.from_partition(ctx, part)
    t_near = time.perf_counter() - t0
    near_entries = float(np.sum(np.diff(near.row_off) * np.diff(near.col_off)))
    far = fc.Batch(ctx, *part.far_blocks())
    far.set_model(model)
    t0 = time.perf_counter()
    far.compress(fc.MODE_LEARNED, tol=1e-4, seed=12345, precision=fc.F32, pinv_rtol=1e-6)
    t_far = time.perf_counter() - t0
    op = fc.Operator(near, far)
    V = mesh.plane_waves(np.array([0.0]), np.array([0.0]), k=k)[:, 0]
    t0 = time.perf_counter()
    J, it, rr = op.solve(-V, tol=1e-5, restart=100, maxit=3000)
    t_solve = time.perf_counter() - t0
    th = np.linspace(0, np.pi, 37)
    rhat = np.vstack([np.sin(th), np.zeros_like(th), np.cos(th)])
    db = 10 * np.log10(fc.bistatic_rcs(mesh, J, k, rhat) / mie.rcs(k, a, np.pi - th, "E"))
    res = {"workload": f"PEC sphere a = {a} lambda (ka = {k * a:.2f}), icosphere level {level}, {N} RWGs, "
                       f"{leaf} lambda leaves: {near.nblocks} near blocks ({near_entries:.3e} entries, both "
                       f"orientations stored once), {far.nblocks} far blocks",
           "near_fill_s": t_near, "near_entries_per_s": near_entries / t_near, "far_compress_s": t_far,
           "gmres_iterations": int(it[0]), "relres": float(rr[0]), "solve_s": t_solve,
           "rcs_vs_mie_db_max": float(np.abs(db).max()), "rcs_vs_mie_db_rms": float(np.sqrt(np.mean(db ** 2)))}
    log(f"# solve ({label}): {N} RWGs, near fill {t_near:.2f} s, GMRES {it[0]} products in {t_solve:.2f} s "
        f"(relres {rr[0]:.1e}), RCS vs Mie max {res['rcs_vs_mie_db_max']:.2f} dB")
    op.close()
    near.close()
    far.close()
    return res


def large_run(fc, args, log):
    """Config 5: ellipsoid 4 x 2 x 1 lambda, icosphere level 6 (122,880 RWG),
    learned rank + batched FP32 fill + FP64 core, then the compressed product
    with 360 plane waves (theta = pi/2, phi_j = 2 pi j / 360)."""
    mesh = fc.Mesh.icosphere(4.0, 2.0, 1.0, 6, 0.01, 1).build_rwg()
    part = fc.Partition(mesh, 0.5, 2.0)
    ro, ri, co, cix = blocks = part.far_blocks()
    nb = len(ro) - 1
    m, n = np.diff(ro), np.diff(co)
    N = mesh.sizes()[2]
    ctx = fc.Context(0)
    tuning = ctx.autotune()
    model, model_src = rank_model(fc, ctx, log)
    ctx.set_mesh(mesh)
    log(f"# large workload: {N} RWGs, {part.ngroups} groups, {nb} far blocks (sizes {m.min()}-{max(m.max(), n.max())})")
    opts = dict(tol=1e-4, seed=12345, precision=fc.F32, pinv_rtol=1e-6)
    bt = fc.Batch(ctx, ro, ri, co, cix)
    bt.set_model(model)
    bt.compress(fc.MODE_LEARNED, **opts)
    t_start = time.perf_counter()
    walls = [bt.compress(fc.MODE_LEARNED, **opts) for _ in range(args.steps)]
    ms = (time.perf_counter() - t_start) * 1e3 / args.steps
    info = bt.info_arrays()
    r = info["r"]
    comp_bytes = float(np.sum((m + n) * r * 8 + r * r * 16))
    dense_bytes = 8.0 * N * N
    phi = 2 * np.pi * np.arange(360) / 360
    V360 = mesh.plane_waves(np.full(360, np.pi / 2), phi, k=2 * np.pi)
    acc = accuracy(bt, V360[:, 0])
    log(f"# learned+batched: {ms:.1f} ms/step, {nb / (ms * 1e-3):.0f} blocks/s, miss rate {bt.misses() / nb:.4f}, "
        f"every block vs FP64 dense: max {acc['max_block_relerr']:.2e} median {acc['median_block_relerr']:.2e}, "
        f"{acc['blocks_above_1e-4']} above 1e-4, {acc['blocks_unconverged']} unconverged")
    # plain per-block loop on a sample of blocks
    rng = np.random.default_rng(0)
    sample = rng.choice(nb, min(200, nb), replace=False)
    sro = np.concatenate([[0], np.cumsum(m[sample])]).astype(np.int32)
    sco = np.concatenate([[0], np.cumsum(n[sample])]).astype(np.int32)
    sri = np.concatenate([ri[ro[b]:ro[b + 1]] for b in sample]).astype(np.int32)
    sci = np.concatenate([cix[co[b]:co[b + 1]] for b in sample]).astype(np.int32)
    pb = fc.Batch(ctx, sro, sri, sco, sci)
    pw = pb.compress(fc.MODE_PLAIN, **opts)
    plain_bps = len(sample) / (pw * 1e-3)
    log(f"# plain per-block loop on {len(sample)} sampled blocks: {pw:.0f} ms -> {plain_bps:.1f} blocks/s")
    pb.close()
    # compressed multi-RHS product, 360 plane waves
    prod = product_bench(fc, bt, mesh, blocks, np.full(360, np.pi / 2), phi, log)
    groups = [part.group(g) for g in range(part.ngroups)]
    sub = V360[:, ::23][:, :16]
    _, _, Yref, _ = bt.verify(sub, symmetric=True)
    for name, mode in (("mfma_f32", fc.MV_MFMA_F32), ("mfma_bf16x3", fc.MV_MFMA_BF16X3)):
        Y, _ = bt.matvec(V360, mode, True)
        Y = Y[:, ::23][:, :16]
        gmax, gmed = _group_errors(Y, Yref, groups)
        prod["modes"][name]["relerr_vs_fp64_dense"] = float(np.linalg.norm(Y - Yref) / np.linalg.norm(Yref))
        prod["modes"][name]["max_group_relerr"] = gmax
    res = {
        "metric": "far-field CUR blocks/s (config 5 large run), learned+batched",
        "value": nb / (ms * 1e-3), "unit": "blocks/s", "n_gpus": 1, "steps": args.steps, "ms_per_step": ms,
        "config": {"workload": f"ellipsoid 4x2x1 lambda, {N} RWG, {nb} far blocks, tol 1e-4", "rank_model": model_src},
        "plain_per_block_sampled_blocks_per_s": plain_bps,
        "speedup_vs_plain_per_block": nb / (ms * 1e-3) / plain_bps,
        "miss_rate": bt.misses() / nb,
        "accuracy": acc,
        "hbm_compressed_bytes": comp_bytes, "hbm_dense_complex64_bytes": dense_bytes,
        "compression_ratio": dense_bytes / comp_bytes,
        "product_360": prod,
        "tuning": tuning,
    }
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["sphere", "large"], default="sphere",
                    help="sphere: config 2 (default, the headline); large: config 5")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["f32", "f64"], default="f32")
    ap.add_argument("--level", type=int, default=5)
    ap.add_argument("--no-plain", action="store_true", help="skip the per-block plain-loop comparison")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--check-blocks", type=int, default=64,
                    help="blocks sampled for the host-side checks of the recompression and ACA legs")
    ap.add_argument("--no-study", action="store_true", help="skip the precision study (config 3)")
    ap.add_argument("--no-solve", action="store_true", help="skip the near-field + GMRES solve legs")
    ap.add_argument("--no-product", action="store_true", help="skip the multi-RHS product timing")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched plain-loop comparison")
    ap.add_argument("--tuning", default=None,
                    help="pinv threads per r class, e.g. 64,256,256,256,512 (skips the autotuner's probes, "
                         "for profiler runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 256 MiB device-to-device copy three times (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device copy", "bytes_read": a.numel() * 4, "bytes_written": a.numel() * 4,
                          "copies": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(far-field blocks are independent: a multi-GPU run would split the block list by predicted cost)", file=sys.stderr)
        sys.exit(2)
    log = lambda s: print(s, flush=True)

    fc = _load_farcur()
    if args.workload == "large":
        large_run(fc, args, log)
        return
    mesh, part = build_workload(fc, level=args.level)
    blocks = part.far_blocks()
    ro, ri, co, cix = blocks
    nb = len(ro) - 1
    m = np.diff(ro)
    n = np.diff(co)
    represented = float(np.sum(m.astype(np.float64) * n))
    ctx = fc.Context(0)
    if args.tuning:
        ctx.set_tuning(pinv_threads=[int(x) for x in args.tuning.split(",")])
        tuning = ctx.tuning()
        log(f"# tuning (given): {tuning}")
    else:
        tuning = ctx.autotune()
        log(f"# autotune: {tuning}")
    model, model_src = rank_model(fc, ctx, log)
    ctx.set_mesh(mesh)
    prec = fc.F32 if args.precision == "f32" else fc.F64
    opts = dict(tol=1e-4, seed=12345, precision=prec, pinv_rtol=1e-6 if prec == fc.F32 else 0.0)
    log(f"# workload: {mesh.sizes()[2]} RWGs, {part.ngroups} groups, {nb} far blocks "
        f"(sizes {m.min()}-{max(m.max(), n.max())}), represented {represented:.3e} entries")

    # headline: learned rank + batched rounds + complex-FP32 fill
    learned = fc.Batch(ctx, ro, ri, co, cix)
    learned.set_model(model)
    for _ in range(args.warmup):
        learned.compress(fc.MODE_LEARNED, **opts)
    # K steps between host clock reads; every compress call ends with a stream
    # synchronisation, so the span covers all of their device work
    t_start = time.perf_counter()
    walls = [learned.compress(fc.MODE_LEARNED, **opts) for _ in range(args.steps)]
    ms = (time.perf_counter() - t_start) * 1e3 / args.steps
    il = learned.info_arrays()
    r = il["r"]
    filled = float(np.sum(m * r + r * r + r * n))
    misses = learned.misses()
    log(f"# learned+batched: {ms:.2f} ms/step (steps {', '.join(f'{w:.1f}' for w in walls)}), "
        f"rounds mean {il['rounds'].mean():.3f}, r mean {r.mean():.1f}, miss rate {misses / nb:.4f}")

    learned.profile(True)
    prof_wall = learned.compress(fc.MODE_LEARNED, **opts)
    kt = learned.kernel_times()
    learned.profile(False)
    for k, v in kt.items():
        log(f"# kernel {k:12s} {v['ms']:9.3f} ms  launches {v['launches']:4d}  work {v['work']:.3e}")
    theta256 = np.arange(256) * np.pi / 255
    V256 = mesh.plane_waves(theta256, np.zeros(256), k=2 * np.pi)
    acc = accuracy(learned, V256[:, 0])
    log(f"# learned accuracy, every block vs FP64 dense: max {acc['max_block_relerr']:.2e} median "
        f"{acc['median_block_relerr']:.2e}, {acc['blocks_above_1e-4']} above 1e-4, {acc['blocks_unconverged']} "
        f"unconverged, matvec max {acc['max_block_matvec_relerr']:.2e} ({acc['verify_ms']:.0f} ms on the device)")

    # the single-RHS CUR matvec (HBM roofline)
    mv1 = matvec1_bench(learned, blocks, V256[:, 0], log)
    # config 3: the compressed operator times 256 plane waves (theta_j = j pi / 255, phi = 0)
    prod = None
    if not args.no_product:
        prod = product_bench(fc, learned, mesh, blocks, theta256, np.zeros(256), log)
    study = None
    if not args.no_study:
        groups = [part.group(g) for g in range(part.ngroups)]
        study = {"sphere": precision_study(fc, ctx, learned, blocks, groups, V256, opts, log, "sphere", model),
                 "plates": plates_study(fc, log)}
        ctx.set_mesh(mesh)
    # SVD recompression of the learned result at 0.1 x tol
    recomp = recompress_bench(fc, ctx, learned, blocks, 0.1 * opts["tol"], args.check_blocks, log)
    # the reference's comparator on the same blocks: GPU ACA (FP64 element)
    aca_res = aca_bench(fc, ctx, learned, blocks, opts["tol"], 64, args.check_blocks, log)

    # the learned method without the network: held-out-checked rounds from
    # the reference's first sample count r0 = min(m, n) / 10
    nomodel = fc.Batch(ctx, ro, ri, co, cix)
    nomodel.compress(fc.MODE_LEARNED, **opts)
    t0 = time.perf_counter()
    for _ in range(3):
        nomodel.compress(fc.MODE_LEARNED, **opts)
    nw = (time.perf_counter() - t0) * 1e3 / 3
    inm = nomodel.info_arrays()
    nacc = accuracy(nomodel, V256[:, 0])
    nomodel_res = {"blocks_per_s": nb / (nw * 1e-3), "ms": nw, "rounds_mean": float(inm["rounds"].mean()),
                   "r_mean": float(inm["r"].mean()), "max_block_relerr": nacc["max_block_relerr"],
                   "blocks_above_1e-4": nacc["blocks_above_1e-4"]}
    log(f"# learned mode without the network: {nw:.2f} ms/step, rounds mean {inm['rounds'].mean():.3f}, "
        f"max err {nacc['max_block_relerr']:.2e}")
    nomodel.close()

    # the same without the network: batched reference loop
    batched_res = None
    batched = fc.Batch(ctx, ro, ri, co, cix)
    if not args.no_batched:
        batched.compress(fc.MODE_BATCHED, **opts)
        bw = float(np.mean([batched.compress(fc.MODE_BATCHED, **opts) for _ in range(2)]))
        bacc = accuracy(batched, V256[:, 0])
        log(f"# batched plain loop: {bw:.2f} ms/step, every block: max {bacc['max_block_relerr']:.2e} median "
            f"{bacc['median_block_relerr']:.2e}")
        batched_res = {"blocks_per_s": nb / (bw * 1e-3), "ms": bw, "max_block_relerr": bacc["max_block_relerr"],
                       "blocks_above_1e-4": bacc["blocks_above_1e-4"]}

    plain = None
    if not args.no_plain:
        pw = batched.compress(fc.MODE_PLAIN, **opts)
        plain = {"blocks_per_s": nb / (pw * 1e-3), "ms": pw}
        log(f"# plain per-block loop: {pw:.1f} ms")

    solve = None
    if not args.no_solve:
        solve = {"sphere_0p85": solve_bench(fc, ctx, model, log, 0.8515, 4, 0.25, "a = 0.85 lambda"),
                 "sphere_2": solve_bench(fc, ctx, model, log, 2.0, 5, 0.5, "the 30,720-RWG sphere")}
    cpu = None if args.no_cpu else cpu_baseline(mesh, blocks, opts)

    value = nb / (ms * 1e-3)
    dflt = fc.default_opts()
    res = {
        "metric": "far-field CUR blocks/s (all admissible blocks, relerr<=1e-4), learned+batched",
        "value": value,
        "unit": "blocks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 fill / f64 core" if prec == fc.F32 else "f64",
        "data": "synthetic (seeded jittered icosphere)",
        "config": {"workload": f"icosphere R=2 lambda level {args.level} ({mesh.sizes()[2]} RWG), 0.5 lambda "
                               f"leaves, {nb} far blocks (one block per symmetric pair of admissible groups; "
                               f"Z_ba = Z_ab^T), tol 1e-4, learned rank + batched rounds",
                   "blocks": nb, "represented_entries": represented, "rank_model": model_src,
                   "learned_checks": {k: getattr(dflt, k) for k in ("oversample", "heldout", "heldout_safety",
                                                                     "growth")},
                   "learned_core": "truncated QR" if dflt.learned_core == fc.CORE_QR else "SVD"},
        "represented_entries_per_s": represented / (ms * 1e-3),
        "filled_entries_per_s": filled / (ms * 1e-3),
        "miss_rate": misses / nb,
        "pinv_sweeps_mean": float(np.mean(il["sweeps"])),
        "round_graphs_captured_replayed": learned.graph_stats(),
        "accuracy": acc,
        "learned_without_network": nomodel_res,
        "speedup_from_network": value / nomodel_res["blocks_per_s"],
        "batched_plain": batched_res,
        "plain_per_block": plain,
        "speedup_vs_plain_per_block": value / plain["blocks_per_s"] if plain else None,
        "speedup_vs_batched_plain": value / batched_res["blocks_per_s"] if batched_res else None,
        "profiled_step_ms": prof_wall,
        "product_256": prod,
        "precision_study": study,
        "full_operator_solve": solve,
        "roofline_matvec1": mv1,
        "recompression": recomp,
        "aca_comparator": aca_res,
        "roofline": roofline(kt),
        "kernels": kt,
        "tuning": tuning,
        "cpu_baseline": cpu,
        "host": {"cpu": cpu_model(), "nproc": os.cpu_count()},
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
