"""Python host of libfarcur.so (ctypes), mirroring the reference's Julia driver.

The reference drives its GPU CUR call from Julia
(codes/GPU_implementation/seed/compres_single_block_spheres_for_sweep2.jl):
``pure_cuda_compression2`` (:484-670) samples, converts the geometry tables and
``ccall``s ``cuda_complete_compression``; ``cuda_cur_compress3`` (:878-973)
runs the adaptive loop. Julia is not available here, so this module restates
those two functions over the same C ABI (``cuda_cur_compress3`` and
``pure_cuda_compression2`` below) and wraps the native API (meshes, contexts,
batches). It needs the compiled library and raises if it is missing: there is
no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes
import math
import pathlib
import weakref
from dataclasses import dataclass

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libfarcur.so"

F64, F32 = 0, 1
ROWS_UNIFORM, ROWS_PIVOT = 0, 1
MODE_PLAIN, MODE_BATCHED, MODE_LEARNED = 0, 1, 2
CORE_QR, CORE_SVD = 0, 1  # FARCUR_CORE_*: learned mode's core pseudo-inverse
MV_VALU, MV_MFMA_F32, MV_MFMA_BF16X3 = 0, 1, 2

_lib = None


class FarcurError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FarcurError(f"{LIB_PATH} is not built (run __graft_entry__.build() or make -C {HERE})")
        L = ctypes.CDLL(str(LIB_PATH))
        try:
            _setup(L)
        except AttributeError as e:  # a stale build missing an entry point
            raise FarcurError(f"{LIB_PATH} does not export {e}; rebuild it (make -C {HERE})") from None
        _lib = L  # only a fully configured library is kept
        # Python's exit hooks run before the C runtime's, so every live context
        # (and the batches and models on it) is released while the HIP runtime
        # is still up; objects left to interpreter teardown would otherwise
        # find this module's globals cleared and keep their graphs, streams and
        # pinned buffers until HIP's own teardown.
        atexit.register(_close_all)
    return _lib


_contexts = weakref.WeakSet()


def _close_all():
    for c in list(_contexts):
        c.close()


vp, ci, cd, cu64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_uint64
PINV_CLASSES = 11  # FARCUR_PINV_CLASSES


class Opts(ctypes.Structure):
    _fields_ = [
        ("tol", cd),
        ("pinv_tol", cd),
        ("pinv_rtol", cd),
        ("seed", cu64),
        ("precision", ci),
        ("row_mode", ci),
        ("r_start", ci),
        ("max_rank", ci),
        ("block_id", ci),
        ("heldout", ci),
        ("oversample", ci),
        ("label_eps", cd),
        ("heldout_safety", cd),
        ("growth", cd),
        ("learned_core", ci),
    ]


class BlockInfo(ctypes.Structure):
    _fields_ = [
        ("r", ci),
        ("rounds", ci),
        ("kept", ci),
        ("converged", ci),
        ("err", cd),
        ("heldout_err", cd),
        ("sigma_max", cd),
        ("time_ms", cd),
        ("rank_eps", ci),
        ("status", ci),
        ("sweeps", ci),
        ("pad", ci),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad"}


class Tuning(ctypes.Structure):
    _fields_ = [("pinv_threads", ci * PINV_CLASSES), ("fill_grid", ci), ("resid_threads", ci),
                ("pinv_ms", cd * PINV_CLASSES)]

    def as_dict(self):
        return {"pinv_threads": list(self.pinv_threads), "fill_grid": self.fill_grid,
                "resid_threads": self.resid_threads, "pinv_ms": list(self.pinv_ms)}


def _setup(L: ctypes.CDLL) -> None:
    L.farcur_last_error.restype = ctypes.c_char_p
    L.farcur_version.restype = ctypes.c_char_p
    L.farcur_opts_default.argtypes = [vp]
    L.farcur_mesh_plate.argtypes = [cd, cd, ci, ci, cd, cd, cd, vp]
    L.farcur_mesh_octasphere.argtypes = [cd, ci, vp]
    L.farcur_mesh_icosphere.argtypes = [cd, cd, cd, ci, cd, cu64, vp]
    L.farcur_mesh_append.argtypes = [vp, vp, cd, cd, cd]
    L.farcur_mesh_build_rwg.argtypes = [vp]
    L.farcur_mesh_sizes.argtypes = [vp, vp, vp, vp]
    L.farcur_mesh_tables.argtypes = [vp] * 9
    L.farcur_mesh_destroy.argtypes = [vp]
    L.farcur_mesh_destroy.restype = None
    L.farcur_mesh_plane_waves.argtypes = [vp, cd, vp, vp, ci, cd, vp]
    L.farcur_partition_build.argtypes = [vp, cd, cd, vp]
    L.farcur_partition_sizes.argtypes = [vp, vp, vp, vp, vp]
    L.farcur_partition_tables.argtypes = [vp] * 7
    L.farcur_partition_blocks.argtypes = [vp, vp, vp, vp, vp]
    L.farcur_partition_save.argtypes = [vp, ctypes.c_char_p]
    L.farcur_partition_load.argtypes = [ctypes.c_char_p, vp]
    L.farcur_partition_params.argtypes = [vp, vp, vp, vp]
    L.farcur_partition_destroy.argtypes = [vp]
    L.farcur_partition_destroy.restype = None
    L.farcur_create.argtypes = [ci, vp]
    L.farcur_destroy.argtypes = [vp]
    L.farcur_destroy.restype = None
    L.farcur_set_em.argtypes = [vp, cd, cd]
    L.farcur_set_geometry.argtypes = [vp, vp, ci, vp, ci, vp, ci, vp]
    L.farcur_set_mesh.argtypes = [vp, vp]
    L.farcur_fill.argtypes = [vp, vp, ci, vp, ci, ci, vp]
    L.farcur_pinv.argtypes = [vp, vp, ci, cd, cd, vp, vp, vp]
    L.farcur_pinv_noalloc.argtypes = [vp, vp, ci, cd, cd, vp, vp, vp, vp]
    L.farcur_pinv_qr.argtypes = [vp, vp, ci, cd, cd, vp, vp, vp]
    L.farcur_cur_residual.argtypes = [vp, vp, ci, ci, vp, vp, ci, vp, vp, vp, vp]
    L.farcur_sample.argtypes = [vp, cu64, ci, ci, ci, ci, ci, vp]
    L.farcur_vtest.argtypes = [cu64, ci, ci, vp]
    L.farcur_pivot_rows.argtypes = [vp, vp, ci, ci, vp]
    L.farcur_compress_block.argtypes = [vp, vp, ci, vp, ci, vp, vp]
    L.farcur_block_result.argtypes = [vp] * 7
    L.cuda_complete_compression.restype = cd
    L.cuda_complete_compression.argtypes = (
        [vp, vp, vp, ci, vp, cd, vp, vp, vp, vp, vp] + [ci] * 12 + [vp, ci, cd, cd, cd, cd, cd, ci, ci] + [vp] * 10
        + [ci] * 18 + [ci] * 6
    )
    L.farcur_autotune.argtypes = [vp, vp]
    L.farcur_get_tuning.argtypes = [vp, vp]
    L.farcur_set_tuning.argtypes = [vp, vp]
    L.farcur_pinv_bench.argtypes = [vp, ci, ci, ci, ci, vp, vp]
    if hasattr(L, "farcur_batch_create"):
        L.farcur_batch_create.argtypes = [vp, ci, vp, vp, vp, vp, vp]
        L.farcur_batch_destroy.argtypes = [vp]
        L.farcur_batch_destroy.restype = None
        L.farcur_batch_compress.argtypes = [vp, vp, ci, vp]
        L.farcur_batch_info.argtypes = [vp, vp]
        L.farcur_batch_block.argtypes = [vp, ci, vp, vp, vp, vp, vp]
        L.farcur_batch_recompress.argtypes = [vp, cd, vp, vp]
        L.farcur_batch_recompressed_block.argtypes = [vp, ci, vp, vp, vp, vp]
        L.farcur_batch_aca.argtypes = [vp, cd, ci, vp, vp]
        L.farcur_batch_aca_block.argtypes = [vp, ci, vp, vp, vp, vp, vp, vp]
        L.farcur_batch_profile.argtypes = [vp, ci]
        L.farcur_batch_kernel_times.argtypes = [vp, vp, vp, vp, vp, ci]
        L.farcur_batch_features.argtypes = [vp, vp]
        L.farcur_batch_set_model.argtypes = [vp, vp]
        L.farcur_batch_misses.argtypes = [vp]
        L.farcur_batch_set_pipeline.argtypes = [vp, ci]
        L.farcur_batch_set_graphs.argtypes = [vp, ci]
        L.farcur_batch_graph_stats.argtypes = [vp, vp, vp]
        L.farcur_batch_matvec.argtypes = [vp, vp, ci, vp, ci, ci, vp]
        L.farcur_batch_verify.argtypes = [vp, vp, ci, ci, vp, vp, vp, vp]
        L.farcur_batch_matvec1.argtypes = [vp, vp, vp, ci, vp]
    L.farcur_fill_near.argtypes = [vp, vp, ci, vp, ci, cd, vp]
    L.farcur_near_create.argtypes = [vp, ci, vp, vp, vp, vp, vp, cd, vp]
    L.farcur_near_block.argtypes = [vp, ci, vp]
    L.farcur_near_destroy.argtypes = [vp]
    L.farcur_near_destroy.restype = None
    L.farcur_operator_create.argtypes = [vp, vp, ci, vp]
    L.farcur_operator_destroy.argtypes = [vp]
    L.farcur_operator_destroy.restype = None
    L.farcur_operator_matvec.argtypes = [vp, vp, vp]
    L.farcur_operator_solve.argtypes = [vp, vp, vp, ci, cd, ci, ci, vp, vp]
    L.farcur_model_create.argtypes = [vp, cu64, vp]
    L.farcur_model_destroy.argtypes = [vp]
    L.farcur_model_destroy.restype = None
    L.farcur_model_get.argtypes = [vp, vp, vp]
    L.farcur_model_set.argtypes = [vp, vp, vp]
    L.farcur_model_loss_grad.argtypes = [vp, vp, vp, ci, vp, vp]
    L.farcur_model_train.argtypes = [vp, vp, vp, ci, ci, ci, ctypes.c_float, cu64, vp]
    L.farcur_model_predict.argtypes = [vp, vp, ci, vp]


def check(rc: int) -> None:
    if rc != 0:
        raise FarcurError(f"farcur error {rc}: {lib().farcur_last_error().decode()}")


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().farcur_opts_default(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown option {k}")
        setattr(o, k, v)
    return o


# ---------------------------------------------------------------- geometry
class Mesh:
    """Surface mesh + RWG tables (get_edge.m layout)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def plate(cls, lx, ly, nx, ny, x0=0.0, y0=0.0, z0=0.0):
        h = vp()
        check(lib().farcur_mesh_plate(lx, ly, nx, ny, x0, y0, z0, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def octasphere(cls, radius, n_levels):
        h = vp()
        check(lib().farcur_mesh_octasphere(radius, n_levels, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def icosphere(cls, ax, ay, az, level, jitter=0.0, seed=1):
        h = vp()
        check(lib().farcur_mesh_icosphere(ax, ay, az, level, jitter, seed, ctypes.byref(h)))
        return cls(h)

    def append(self, other: "Mesh", dx=0.0, dy=0.0, dz=0.0) -> "Mesh":
        check(lib().farcur_mesh_append(self._h, other._h, dx, dy, dz))
        return self

    def build_rwg(self) -> "Mesh":
        check(lib().farcur_mesh_build_rwg(self._h))
        return self

    def sizes(self):
        nv, nt, ne = ci(), ci(), ci()
        check(lib().farcur_mesh_sizes(self._h, ctypes.byref(nv), ctypes.byref(nt), ctypes.byref(ne)))
        return nv.value, nt.value, ne.value

    def tables(self) -> dict:
        """Reference layout: 1-based ids as float64, column-major 3 x N etc."""
        nv, nt, ne = self.sizes()
        out = {
            "vertex": np.empty((nv, 3)),
            "topol": np.empty((nt, 3)),
            "trian": np.empty((nt, 3)),
            "edges": np.empty((ne, 4)),
            "un": np.empty((nt, 3)),
            "ds": np.empty(nt),
            "ln": np.empty(ne),
            "cent": np.empty((nt, 3)),
        }
        check(lib().farcur_mesh_tables(self._h, *[_p(out[k]) for k in
                                                   ("vertex", "topol", "trian", "edges", "un", "ds", "ln", "cent")]))
        for k in ("vertex", "topol", "trian", "edges", "un", "cent"):
            out[k] = np.asfortranarray(out[k].T)
        return out

    def plane_waves(self, theta, phi, k=2 * math.pi, rot_eh=math.pi / 2) -> np.ndarray:
        """RWG-tested plane waves, ne x len(theta) (radians; rot_eh = pi/2: theta-polarised)."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        ph = np.ascontiguousarray(phi, dtype=np.float64)
        ne = self.sizes()[2]
        V = np.empty((ne, th.size), dtype=np.complex128, order="F")
        check(lib().farcur_mesh_plane_waves(self._h, k, _p(th), _p(ph), th.size, rot_eh, _p(V)))
        return V

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.farcur_mesh_destroy(self._h)
            self._h = None


def plates_config(side=1.0, nsq=20, zsep=4.0) -> Mesh:
    """Two parallel square plates (normal +z) centred at z = 0 and z = zsep."""
    m = Mesh.plate(side, side, nsq, nsq, 0, 0, 0)
    m.append(Mesh.plate(side, side, nsq, nsq, 0, 0, 0), 0, 0, zsep)
    return m.build_rwg()


def two_octaspheres(radius, n_levels, d) -> Mesh:
    """object_wrapper_spheres_interaction2.m: two sphere.m spheres, the second shifted by d along x."""
    m = Mesh.octasphere(radius, n_levels)
    m.append(Mesh.octasphere(radius, n_levels), d, 0, 0)
    return m.build_rwg()


class Partition:
    """Octree leaf groups and the far/near block list (the reference's Z_comp)."""

    def __init__(self, mesh: Mesh = None, leaf_size=0.5, eta_adm=2.0, _handle=None):
        if _handle is not None:
            self._h = _handle
        else:
            self._h = vp()
            check(lib().farcur_partition_build(mesh._h, leaf_size, eta_adm, ctypes.byref(self._h)))
        self._read()

    @classmethod
    def load(cls, path) -> "Partition":
        """A block-list file written by save()."""
        h = vp()
        check(lib().farcur_partition_load(str(path).encode(), ctypes.byref(h)))
        return cls(_handle=h)

    def save(self, path) -> None:
        check(lib().farcur_partition_save(self._h, str(path).encode()))

    def params(self) -> dict:
        leaf, eta, nr = cd(), cd(), ci()
        check(lib().farcur_partition_params(self._h, ctypes.byref(leaf), ctypes.byref(eta), ctypes.byref(nr)))
        return {"leaf_size": leaf.value, "eta_adm": eta.value, "nrwg_mesh": nr.value}

    def zcomp(self) -> list:
        """The block list as the reference's Z_comp dicts {m, n (1-based RWG ids),
        comp, self} (far blocks first)."""
        nb = len(self.far_pairs) + len(self.near_pairs)
        a, b, comp, slf = (np.empty(nb, np.int32) for _ in range(4))
        check(lib().farcur_partition_blocks(self._h, _p(a), _p(b), _p(comp), _p(slf)))
        return [{"m": self.group(a[i]) + 1, "n": self.group(b[i]) + 1, "comp": int(comp[i]), "self": int(slf[i])}
                for i in range(nb)]

    def _read(self):
        ng, nr, nf, nn = ci(), ci(), ci(), ci()
        check(lib().farcur_partition_sizes(self._h, ctypes.byref(ng), ctypes.byref(nr), ctypes.byref(nf),
                                           ctypes.byref(nn)))
        self.group_offsets = np.empty(ng.value + 1, dtype=np.int32)
        self.group_rwg = np.empty(nr.value, dtype=np.int32)
        self.centre = np.empty((ng.value, 3))
        self.diam = np.empty(ng.value)
        self.far_pairs = np.empty((nf.value, 2), dtype=np.int32)
        self.near_pairs = np.empty((nn.value, 2), dtype=np.int32)
        check(lib().farcur_partition_tables(self._h, _p(self.group_offsets), _p(self.group_rwg), _p(self.centre),
                                            _p(self.diam), _p(self.far_pairs), _p(self.near_pairs)))

    @property
    def ngroups(self):
        return self.diam.size

    def group(self, g):
        return self.group_rwg[self.group_offsets[g]:self.group_offsets[g + 1]]

    def far_blocks(self):
        """(row_off, row_ids, col_off, col_ids) of the admissible pairs (a, b), a < b."""
        rows, cols, ro, co = [], [], [0], [0]
        for a, b in self.far_pairs:
            ra, cb = self.group(a), self.group(b)
            rows.append(ra)
            cols.append(cb)
            ro.append(ro[-1] + ra.size)
            co.append(co[-1] + cb.size)
        cat = lambda xs: np.concatenate(xs).astype(np.int32) if xs else np.zeros(0, np.int32)
        return np.array(ro, np.int32), cat(rows), np.array(co, np.int32), cat(cols)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.farcur_partition_destroy(self._h)
            self._h = None


# ---------------------------------------------------------------- context
def _cvec(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


def _cmat(a) -> np.ndarray:
    return np.asfortranarray(a, dtype=np.complex128)


def _ivec(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


class Context:
    def __init__(self, device=0):
        self._h = vp()
        # batches and models created on this context: they hold its device
        # and stream, so close() releases them first
        self._children = weakref.WeakSet()
        check(lib().farcur_create(device, ctypes.byref(self._h)))
        _contexts.add(self)

    def close(self):
        for ch in list(getattr(self, "_children", ())):
            ch.close()
        if getattr(self, "_h", None):
            lib().farcur_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()

    def set_em(self, k, eta=120 * math.pi):
        check(lib().farcur_set_em(self._h, k, eta))

    def set_mesh(self, mesh: Mesh):
        check(lib().farcur_set_mesh(self._h, mesh._h))

    def set_geometry(self, vertex, topol, edges, ln):
        v = np.asfortranarray(vertex, dtype=np.float64)
        t = np.asfortranarray(topol, dtype=np.float64)
        e = np.asfortranarray(edges, dtype=np.float64)
        l = np.ascontiguousarray(ln, dtype=np.float64)
        check(lib().farcur_set_geometry(self._h, _p(v), v.shape[1], _p(t), t.shape[1], _p(e), e.shape[1], _p(l)))

    def fill(self, rows, cols, precision=F64) -> np.ndarray:
        rows, cols = _ivec(rows), _ivec(cols)
        Z = np.empty((rows.size, cols.size), dtype=np.complex128, order="F")
        check(lib().farcur_fill(self._h, _p(rows), rows.size, _p(cols), cols.size, precision, _p(Z)))
        return Z

    def fill_near(self, rows, cols, near_factor=2.0) -> np.ndarray:
        """Panel with the near-field element (singular extraction for close
        triangle pairs), FP64."""
        rows, cols = _ivec(rows), _ivec(cols)
        Z = np.empty((rows.size, cols.size), dtype=np.complex128, order="F")
        check(lib().farcur_fill_near(self._h, _p(rows), rows.size, _p(cols), cols.size, near_factor, _p(Z)))
        return Z

    def pinv(self, W, atol=1e-10, rtol=0.0):
        W = _cmat(W)
        r = W.shape[0]
        U = np.empty((r, r), dtype=np.complex128, order="F")
        s = np.empty(r)
        kept = ci()
        check(lib().farcur_pinv(self._h, _p(W), r, atol, rtol, _p(U), _p(s), ctypes.byref(kept)))
        return U, s, kept.value

    def pinv_qr(self, W, atol=1e-10, rtol=1e-6):
        """The learned core's truncated QR pseudo-inverse: (U, |R_jj| in pivot
        order, k)."""
        W = _cmat(W)
        r = W.shape[0]
        U = np.empty((r, r), dtype=np.complex128, order="F")
        d = np.empty(r)
        kept = ci()
        check(lib().farcur_pinv_qr(self._h, _p(W), r, atol, rtol, _p(U), _p(d), ctypes.byref(kept)))
        return U, d, kept.value

    def pinv_noalloc(self, W_dev: int, r: int, atol: float, rtol: float, U_dev: int, sigma_dev: int = 0,
                     kept_dev: int = 0, stream: int = 0) -> None:
        """farcur_pinv_noalloc on device addresses (ints): enqueues only, on
        `stream` (0: the context's stream); the caller synchronises."""
        check(lib().farcur_pinv_noalloc(self._h, W_dev, r, atol, rtol, U_dev, sigma_dev or None, kept_dev or None,
                                        stream or None))

    def cur_residual(self, C, U, R, v, y_old):
        C, U, R, v, y_old = _cmat(C), _cmat(U), _cmat(R), _cvec(v), _cvec(y_old)
        m, r = C.shape
        n = R.shape[1]
        y = np.empty(m, dtype=np.complex128)
        err = cd()
        check(lib().farcur_cur_residual(self._h, _p(C), m, r, _p(U), _p(R), n, _p(v), _p(y_old), _p(y),
                                        ctypes.byref(err)))
        return y, err.value

    def sample(self, seed, block, rnd, purpose, N, r) -> np.ndarray:
        out = np.empty(r, dtype=np.int32)
        check(lib().farcur_sample(self._h, seed, block, rnd, purpose, N, r, _p(out)))
        return out

    def pivot_rows(self, C) -> np.ndarray:
        C = _cmat(C)
        m, r = C.shape
        out = np.empty(r, dtype=np.int32)
        check(lib().farcur_pivot_rows(self._h, _p(C), m, r, _p(out)))
        return out

    def autotune(self) -> dict:
        t = Tuning()
        check(lib().farcur_autotune(self._h, ctypes.byref(t)))
        return t.as_dict()

    def tuning(self) -> dict:
        t = Tuning()
        check(lib().farcur_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, pinv_threads=None, fill_grid=None, resid_threads=None):
        t = Tuning()
        check(lib().farcur_get_tuning(self._h, ctypes.byref(t)))
        if pinv_threads is not None:
            for i, v in enumerate(pinv_threads):
                t.pinv_threads[i] = v
        if fill_grid is not None:
            t.fill_grid = fill_grid
        if resid_threads is not None:
            t.resid_threads = resid_threads
        check(lib().farcur_set_tuning(self._h, ctypes.byref(t)))

    def pinv_bench(self, r, nprob, threads, use_lds):
        """(ms, mean sweeps) for nprob random r x r problems."""
        ms, sw = cd(), cd()
        check(lib().farcur_pinv_bench(self._h, r, nprob, threads, 1 if use_lds else 0, ctypes.byref(ms),
                                      ctypes.byref(sw)))
        return ms.value, sw.value

    def compress_block(self, rows, cols, **opts):
        rows, cols = _ivec(rows), _ivec(cols)
        o = default_opts(**opts)
        info = BlockInfo()
        check(lib().farcur_compress_block(self._h, _p(rows), rows.size, _p(cols), cols.size, ctypes.byref(o),
                                          ctypes.byref(info)))
        m, n, r = rows.size, cols.size, info.r
        C = np.empty((m, r), dtype=np.complex128, order="F")
        U = np.empty((r, r), dtype=np.complex128, order="F")
        R = np.empty((r, n), dtype=np.complex128, order="F")
        rs = np.empty(r, dtype=np.int32)
        cs = np.empty(r, dtype=np.int32)
        y = np.empty(m, dtype=np.complex128)
        check(lib().farcur_block_result(self._h, _p(C), _p(U), _p(R), _p(rs), _p(cs), _p(y)))
        return {"C": C, "U": U, "R": R, "rs": rs, "cs": cs, "y": y, **info.as_dict()}


def vtest(seed, block, n) -> np.ndarray:
    v = np.empty(n, dtype=np.complex128)
    check(lib().farcur_vtest(seed, block, n, _p(v)))
    return v


# ---------------------------------------------------------------- batches
class Batch:
    def __init__(self, ctx: Context, row_off, row_ids, col_off, col_ids):
        self.ctx = ctx
        self.row_off, self.row_ids = _ivec(row_off), _ivec(row_ids)
        self.col_off, self.col_ids = _ivec(col_off), _ivec(col_ids)
        self.nblocks = self.row_off.size - 1
        self._h = vp()
        check(lib().farcur_batch_create(ctx._h, self.nblocks, _p(self.row_off), _p(self.row_ids), _p(self.col_off),
                                        _p(self.col_ids), ctypes.byref(self._h)))
        ctx._children.add(self)

    def close(self):
        if getattr(self, "_h", None):
            lib().farcur_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()

    def compress(self, mode=MODE_BATCHED, **opts) -> float:
        o = default_opts(**opts)
        wall = cd()
        check(lib().farcur_batch_compress(self._h, ctypes.byref(o), mode, ctypes.byref(wall)))
        return wall.value

    def info(self) -> list[dict]:
        arr = (BlockInfo * self.nblocks)()
        check(lib().farcur_batch_info(self._h, arr))
        return [a.as_dict() for a in arr]

    def info_arrays(self) -> dict:
        infos = self.info()
        return {k: np.array([d[k] for d in infos]) for k in infos[0]} if infos else {}

    def block(self, b):
        m = int(self.row_off[b + 1] - self.row_off[b])
        n = int(self.col_off[b + 1] - self.col_off[b])
        r = self.info()[b]["r"]
        C = np.empty((m, r), dtype=np.complex128, order="F")
        U = np.empty((r, r), dtype=np.complex128, order="F")
        R = np.empty((r, n), dtype=np.complex128, order="F")
        rs = np.empty(r, dtype=np.int32)
        cs = np.empty(r, dtype=np.int32)
        check(lib().farcur_batch_block(self._h, b, _p(C), _p(U), _p(R), _p(rs), _p(cs)))
        return {"C": C, "U": U, "R": R, "rs": rs, "cs": cs}

    def recompress(self, thresh):
        """SVD recompression of every block of the last compress (svd_compress.m:19-34):
        returns (k per block, device ms)."""
        ranks = np.empty(self.nblocks, dtype=np.int32)
        ms = cd()
        check(lib().farcur_batch_recompress(self._h, float(thresh), _p(ranks), ctypes.byref(ms)))
        return ranks, ms.value

    def recompressed_block(self, b):
        """Block b of the last recompression: C' (m x k), s (k), R' (k x n), with
        C U R ~ C' diag(s) R'."""
        m = int(self.row_off[b + 1] - self.row_off[b])
        n = int(self.col_off[b + 1] - self.col_off[b])
        k = ci()
        check(lib().farcur_batch_recompressed_block(self._h, b, None, None, None, ctypes.byref(k)))
        k = k.value
        C = np.empty((m, k), dtype=np.complex128, order="F")
        s = np.empty(k)
        R = np.empty((k, n), dtype=np.complex128, order="F")
        if k:
            check(lib().farcur_batch_recompressed_block(self._h, b, _p(C), _p(s), _p(R), None))
        return {"C": C, "s": s, "R": R}

    def aca(self, tol, max_rank=0):
        """Partially pivoted ACA of every block (ACA.m:40-128): (k per block, device ms)."""
        ranks = np.empty(self.nblocks, dtype=np.int32)
        ms = cd()
        check(lib().farcur_batch_aca(self._h, float(tol), int(max_rank), _p(ranks), ctypes.byref(ms)))
        return ranks, ms.value

    def aca_block(self, b):
        """Block b of the last ACA: U (m x k), V (k x n), pivots I, J, status."""
        m = int(self.row_off[b + 1] - self.row_off[b])
        n = int(self.col_off[b + 1] - self.col_off[b])
        k, st = ci(), ci()
        check(lib().farcur_batch_aca_block(self._h, b, None, None, None, None, ctypes.byref(k), ctypes.byref(st)))
        k = k.value
        U = np.empty((m, k), dtype=np.complex128, order="F")
        V = np.empty((k, n), dtype=np.complex128, order="F")
        I = np.empty(k, dtype=np.int32)
        J = np.empty(k, dtype=np.int32)
        if k:
            check(lib().farcur_batch_aca_block(self._h, b, _p(U), _p(V), _p(I), _p(J), None, None))
        return {"U": U, "V": V, "I": I, "J": J, "status": st.value}

    def features(self) -> np.ndarray:
        X = np.empty((self.nblocks, 8), dtype=np.float32)
        check(lib().farcur_batch_features(self._h, _p(X)))
        return X

    def set_model(self, model: "Model | None"):
        self._model = model
        check(lib().farcur_batch_set_model(self._h, model._h if model is not None else None))

    def matvec(self, V, mode=1, symmetric=True):
        """Y (N x s) = compressed far field times V (N x s); returns
        (Y, [ms pass A, ms pass B, ms forming P = C U (0 once formed)])."""
        V = np.asfortranarray(V, dtype=np.complex128)
        if V.ndim == 1:
            V = V.reshape(-1, 1, order="F")
        Y = np.empty_like(V, order="F")
        ms = np.zeros(3)
        check(lib().farcur_batch_matvec(self._h, _p(V), V.shape[1], _p(Y), mode, 1 if symmetric else 0, _p(ms)))
        return Y, ms

    def matvec1(self, v, symmetric=True):
        """y (N) = compressed far field times one vector v (N), HBM-bound
        single-RHS kernel; returns (y, device ms of the product kernel)."""
        v = _cvec(v)
        y = np.empty_like(v)
        ms = cd()
        check(lib().farcur_batch_matvec1(self._h, _p(v), _p(y), int(symmetric), ctypes.byref(ms)))
        return y, ms.value

    def verify(self, V, symmetric=True, want_product=True):
        """Accuracy of every block of the last compress against its FP64 dense
        block, on the device: (relative Frobenius error per block, matvec error
        per block for v = V[:, 0], the FP64 dense far-field product Z V (N x s,
        with the transposed blocks when symmetric) or None, device ms).
        V: N x s with 1 <= s <= 16."""
        V = np.asfortranarray(V, dtype=np.complex128)
        if V.ndim == 1:
            V = V[:, None]
        fro = np.empty(self.nblocks)
        mv = np.empty(self.nblocks)
        Y = np.zeros(V.shape, dtype=np.complex128, order="F") if want_product else None
        ms = cd()
        check(lib().farcur_batch_verify(self._h, _p(V), V.shape[1], int(symmetric), _p(fro), _p(mv),
                                        _p(Y) if Y is not None else None, ctypes.byref(ms)))
        return fro, mv, Y, ms.value

    def misses(self) -> int:
        return lib().farcur_batch_misses(self._h)

    def set_pipeline(self, chunks: int):
        """Two-stream pipeline of the batched rounds (chunks <= 1: off)."""
        check(lib().farcur_batch_set_pipeline(self._h, chunks))

    def set_graphs(self, enable: bool):
        """hipGraph capture / replay of the batched rounds (on by default)."""
        check(lib().farcur_batch_set_graphs(self._h, 1 if enable else 0))

    def graph_stats(self):
        """(captures, replays) of round graphs since the batch was created."""
        c, r = ci(), ci()
        check(lib().farcur_batch_graph_stats(self._h, ctypes.byref(c), ctypes.byref(r)))
        return c.value, r.value

    def profile(self, enable=True):
        check(lib().farcur_batch_profile(self._h, 1 if enable else 0))

    def kernel_times(self) -> dict:
        cap = 32  # kernel classes and the host phases
        names = (ctypes.c_char_p * cap)()
        ms = np.zeros(cap)
        launches = np.zeros(cap, dtype=np.int32)
        evals = np.zeros(cap)
        n = lib().farcur_batch_kernel_times(self._h, names, _p(ms), _p(launches), _p(evals), cap)
        return {names[i].decode(): {"ms": float(ms[i]), "launches": int(launches[i]), "work": float(evals[i])}
                for i in range(n)}


# ---------------------------------------------------------------- rank model
N_PARAMS, N_FEATURES, N_NORM = 4801, 8, 18


class Model:
    """MLP 8 -> 64 -> 64 -> 1 (GELU) predicting log(eps-rank); runs on the GPU."""

    def __init__(self, ctx: Context, seed=1):
        self.ctx = ctx
        self._h = vp()
        check(lib().farcur_model_create(ctx._h, seed, ctypes.byref(self._h)))
        ctx._children.add(self)

    def close(self):
        if getattr(self, "_h", None):
            lib().farcur_model_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()

    def get(self):
        p = np.empty(N_PARAMS, dtype=np.float32)
        nrm = np.empty(N_NORM, dtype=np.float32)
        check(lib().farcur_model_get(self._h, _p(p), _p(nrm)))
        return p, nrm

    def set(self, params=None, norm=None):
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        n = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        check(lib().farcur_model_set(self._h, _p(p) if p is not None else None, _p(n) if n is not None else None))

    def loss_grad(self, X, y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.empty(N_PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        check(lib().farcur_model_loss_grad(self._h, _p(X), _p(y), y.size, ctypes.byref(loss), _p(g)))
        return loss.value, g

    def train(self, X, y, epochs=20, batch=4096, lr=2e-3, seed=1):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        losses = np.empty(epochs, dtype=np.float32)
        check(lib().farcur_model_train(self._h, _p(X), _p(y), y.size, epochs, batch, lr, seed, _p(losses)))
        return losses

    def predict(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.empty(X.shape[0], dtype=np.float32)
        check(lib().farcur_model_predict(self._h, _p(X), X.shape[0], _p(y)))
        return y

    def save(self, path):
        import json
        p, n = self.get()
        pathlib.Path(path).write_text(json.dumps({"arch": "mlp-8-64-64-1-gelu", "params": p.tolist(),
                                                  "norm": n.tolist()}))

    def load(self, path):
        import json
        d = json.loads(pathlib.Path(path).read_text())
        if d.get("arch") != "mlp-8-64-64-1-gelu":
            raise FarcurError(f"{path}: unknown model architecture {d.get('arch')}")
        self.set(np.array(d["params"], np.float32), np.array(d["norm"], np.float32))
        return self


class NearField:
    """Dense near-field blocks on the device (the partition's non-admissible
    pairs), filled with the near-field element."""

    def __init__(self, ctx: Context, row_off, row_ids, col_off, col_ids, transpose=None, near_factor=2.0):
        self.row_off, self.col_off = _ivec(row_off), _ivec(col_off)
        row_ids, col_ids = _ivec(row_ids), _ivec(col_ids)
        tr = None if transpose is None else _ivec(transpose)
        self._h = vp()
        self.ctx = ctx
        check(lib().farcur_near_create(ctx._h, self.row_off.size - 1, _p(self.row_off), _p(row_ids),
                                       _p(self.col_off), _p(col_ids), _p(tr) if tr is not None else None,
                                       near_factor, ctypes.byref(self._h)))
        ctx._children.add(self)

    @classmethod
    def from_partition(cls, ctx, part: "Partition", near_factor=2.0):
        rows, cols, tr = [], [], []
        for a, b in part.near_pairs:
            rows.append(part.group(a))
            cols.append(part.group(b))
            tr.append(int(a != b))
        ro = np.concatenate([[0], np.cumsum([r.size for r in rows])])
        co = np.concatenate([[0], np.cumsum([c.size for c in cols])])
        return cls(ctx, ro, np.concatenate(rows), co, np.concatenate(cols), tr, near_factor)

    @property
    def nblocks(self):
        return self.row_off.size - 1

    def block(self, b):
        m = int(self.row_off[b + 1] - self.row_off[b])
        n = int(self.col_off[b + 1] - self.col_off[b])
        Z = np.empty((m, n), dtype=np.complex128, order="F")
        check(lib().farcur_near_block(self._h, b, _p(Z)))
        return Z

    def close(self):
        if getattr(self, "_h", None):
            lib().farcur_near_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()


class Operator:
    """Near dense blocks + far CUR blocks: products and GMRES solves on the device."""

    def __init__(self, near: NearField = None, far: "Batch" = None, far_symmetric=True):
        self._h = vp()
        self._keep = (near, far)
        check(lib().farcur_operator_create(near._h if near else None, far._h if far else None, int(far_symmetric),
                                           ctypes.byref(self._h)))
        (near or far).ctx._children.add(self)

    def matvec(self, x):
        x = _cvec(x)
        y = np.empty_like(x)
        check(lib().farcur_operator_matvec(self._h, _p(x), _p(y)))
        return y

    def solve(self, b, tol=1e-5, restart=100, maxit=2000):
        """x with Z~ x = b (N or N x s), restarted GMRES; returns (x, iters, relres)."""
        b = np.asfortranarray(b, dtype=np.complex128)
        one = b.ndim == 1
        B = b[:, None] if one else b
        X = np.empty_like(B, order="F")
        it = np.empty(B.shape[1], np.int32)
        rr = np.empty(B.shape[1])
        check(lib().farcur_operator_solve(self._h, _p(np.asfortranarray(B)), _p(X), B.shape[1], tol, restart, maxit,
                                          _p(it), _p(rr)))
        return (X[:, 0] if one else X), it, rr

    def close(self):
        if getattr(self, "_h", None):
            lib().farcur_operator_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()


def rwg_far_field(mesh: "Mesh", J, k, rhat):
    """N(rhat) = sum_n J_n int f_n(r') e^{jk rhat.r'} dS' (7-point rule per
    triangle; rhat 3 x nd), and the bistatic radar cross section
    k^2 eta^2 |N_perp|^2 / (4 pi) of the currents for a unit incident field."""
    t = mesh.tables()
    V, T, E, ln = t["vertex"], t["topol"].astype(int) - 1, t["edges"].astype(int) - 1, t["ln"]
    rhat = np.asarray(rhat, float)
    bary = [(1 / 3, 1 / 3, 1 / 3, 0.225)]  # Dunavant's 7-point rule
    for a, b, w in ((0.059715871789770, 0.470142064105115, 0.132394152788506),
                    (0.797426985353087, 0.101286507323456, 0.125939180544827)):
        bary += [(a, b, b, w), (b, a, b, w), (b, b, a, w)]
    N = np.zeros((3, rhat.shape[1]), complex)
    for h, s in ((0, 1.0), (1, -1.0)):
        P = V[:, T[:, E[h]]]
        vf = V[:, E[2 + h]]
        for a, b, c, w in bary:
            q = a * P[:, 0] + b * P[:, 1] + c * P[:, 2]
            N += ((q - vf) * (s * ln * np.asarray(J) / 2 * w)) @ np.exp(1j * k * (rhat.T @ q)).T
    return N


def bistatic_rcs(mesh: "Mesh", J, k, rhat, eta=120 * math.pi):
    N = rwg_far_field(mesh, J, k, rhat)
    Np = N - rhat * np.sum(N * rhat, axis=0)
    return k ** 2 * eta ** 2 * np.sum(np.abs(Np) ** 2, axis=0) / (4 * math.pi)


# ---------------------------------------------------------------- reference driver, restated
@dataclass
class EMData:
    """EM_data_struct (compres_single_block_spheres_for_sweep2.jl:28-39)."""
    k: float = 2 * math.pi
    eta: float = 120 * math.pi
    field: int = 1
    Rint_s: float = 0.0
    Rint_f: float = 0.0
    Ranal_s: float = 0.0
    corr_solid: int = 0
    flag: int = 0


def pure_cuda_compression2(vtest_, Avtest_old, current_size, pinv_tol, r1, r2, obj: dict, EM: EMData, rng):
    """One CUR round through the drop-in symbol, as the reference's
    pure_cuda_compression2 (compres_single_block_spheres_for_sweep2.jl:484-670):
    uniform sorted samples without replacement (StatsBase.sample, :492-495),
    mapped to the block's global 1-based ids (:496), tables as Float64 (:513-545).
    r1, r2 are 1-based RWG ids."""
    r1 = np.asarray(r1)
    r2 = np.asarray(r2)
    MM, NN, cur = r1.size, r2.size, int(current_size)
    col_samples = np.sort(rng.choice(NN, cur, replace=False))
    row_samples = np.sort(rng.choice(MM, cur, replace=False))
    Ccol = np.ascontiguousarray(r2[col_samples], dtype=np.float64)
    Crow = np.ascontiguousarray(r1[row_samples], dtype=np.float64)
    tabs = {k: np.asfortranarray(obj[k], dtype=np.float64) for k in
            ("vertex", "topol", "trian", "edges", "un", "cent")}
    ds = np.ascontiguousarray(obj["ds"], dtype=np.float64)
    ln = np.ascontiguousarray(obj["ln"], dtype=np.float64)
    r1d = np.ascontiguousarray(r1, dtype=np.float64)
    r2d = np.ascontiguousarray(r2, dtype=np.float64)
    C = np.zeros((MM, cur), dtype=np.complex128, order="F")
    U = np.zeros((cur, cur), dtype=np.complex128, order="F")
    R = np.zeros((cur, NN), dtype=np.complex128, order="F")
    v = _cvec(vtest_)
    yo = _cvec(Avtest_old)
    yn = np.zeros(MM, dtype=np.complex128)
    t = np.zeros(1)
    Z = np.zeros(1, dtype=np.complex128)
    d = 8  # divfactor
    V, T, Tr, E, Un, Ce = (tabs[k] for k in ("vertex", "topol", "trian", "edges", "un", "cent"))
    err = lib().cuda_complete_compression(
        _p(C), _p(U), _p(R), cur, _p(t), pinv_tol, _p(yo), _p(yn), _p(v), _p(Ccol), _p(Crow),
        -(-MM // d), -(-cur // d), d, d, -(-cur // d), -(-cur // d), d, d, -(-cur // d), -(-NN // d), d, d,
        _p(Z), EM.field, EM.k, EM.eta, EM.Rint_s, EM.Ranal_s, EM.Rint_f, EM.corr_solid, EM.flag,
        _p(r1d), _p(r2d), _p(V), _p(T), _p(Tr), _p(E), _p(Un), _p(ds), _p(ln), _p(Ce),
        E.shape[1], MM, NN, V.shape[0], V.shape[1], T.shape[0], T.shape[1], T.shape[0], T.shape[1],
        E.shape[0], E.shape[1], Un.shape[0], Un.shape[1], 1, ds.size, ln.size, Ce.shape[0], Ce.shape[1],
        MM, NN, 1, 1, MM, NN)
    if math.isnan(err):
        raise FarcurError(lib().farcur_last_error().decode())
    return C, U, R, Crow, Ccol, yn, err, float(t[0])


def cuda_cur_compress3(r1, r2, tol, obj: dict, EM: EMData, rng, pinv_tol=1e-10):
    """The reference's adaptive loop (compres_single_block_spheres_for_sweep2.jl:878-973)."""
    M, N = len(r1), len(r2)
    vt = rng.random(N) + 1j * rng.random(N)
    D = min(M, N)
    cur = 1 if D < 10 else D // 10
    y_old = np.ones(M, dtype=np.complex128)
    C, U, R, rs, cs, y_new, _, t = pure_cuda_compression2(vt, y_old, cur, pinv_tol, r1, r2, obj, EM, rng)
    time_total = t
    y_old = y_new.copy()
    cur = 2 * cur if 2 * cur < D else D
    err_app, flag_size = math.inf, 0
    while err_app > tol and flag_size == 0:
        C, U, R, rs, cs, y_new, err_app, t = pure_cuda_compression2(vt, y_old, cur, pinv_tol, r1, r2, obj, EM, rng)
        time_total += t
        y_old = y_new.copy()
        if cur > D:
            raise FarcurError("current_size larger than D")
        elif cur == D:
            flag_size = 1
        else:
            cur = 2 * cur if 2 * cur <= D else D
    return C, U, R, rs, cs, err_app, time_total
