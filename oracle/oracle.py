"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline leg). ctypes front end of oracle.c, the CPU FP64 restatement of
the reference's far-field CUR path; see oracle.c's header for the reference
file:line each function follows and the pins that check it.

Geometry is passed as the tables of the reference's object struct
(codes/GPU_implementation/seed/compres_single_block_spheres_for_sweep2.jl:12-26):
``vertex`` 3 x Nv, ``topol`` 3 x Nt and ``edges`` 4 x Ne (1-based ids as
float64, the layout Julia hands to the C library) and ``ln``.
"""
from __future__ import annotations

import ctypes
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liborcfarcur.so"

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        _setup(_lib)
    return _lib


class _Geom(ctypes.Structure):
    _fields_ = [
        ("vertex", ctypes.c_void_p),
        ("topol", ctypes.c_void_p),
        ("edges", ctypes.c_void_p),
        ("ln", ctypes.c_void_p),
        ("ne", ctypes.c_int),
        ("k", ctypes.c_double),
        ("eta", ctypes.c_double),
    ]


class _Opts(ctypes.Structure):
    _fields_ = [
        ("tol", ctypes.c_double),
        ("pinv_tol", ctypes.c_double),
        ("pinv_rtol", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("row_mode", ctypes.c_int),
        ("r_start", ctypes.c_int),
        ("max_rank", ctypes.c_int),
        ("block_id", ctypes.c_int),
    ]


class _Info(ctypes.Structure):
    _fields_ = [
        ("r", ctypes.c_int),
        ("rounds", ctypes.c_int),
        ("kept", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("err", ctypes.c_double),
    ]


class _Block(ctypes.Structure):
    _fields_ = [
        ("g", ctypes.c_void_p),
        ("rows", ctypes.c_void_p),
        ("cols", ctypes.c_void_p),
        ("dense", ctypes.c_void_p),
        ("m", ctypes.c_int),
        ("n", ctypes.c_int),
    ]


def _setup(L: ctypes.CDLL) -> None:
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.orc_element.argtypes = [vp, i, i, vp, vp]
    L.orc_fill.argtypes = [vp, vp, i, vp, i, vp]
    L.orc_block_matvec.argtypes = [vp, vp, i, vp, i, vp, vp]
    L.orc_pinv.argtypes = [vp, i, d, d, vp, vp]
    L.orc_pinv.restype = i
    L.orc_residual.argtypes = [vp, i, i, vp, vp, i, vp, vp, vp]
    L.orc_residual.restype = d
    L.orc_philox.argtypes = [vp, vp, vp]
    L.orc_sample.argtypes = [ctypes.c_uint64, i, i, i, i, i, vp]
    L.orc_vtest.argtypes = [ctypes.c_uint64, i, i, vp]
    L.orc_sample_heldout.argtypes = [ctypes.c_uint64, i, i, i, i, i, vp, vp]
    L.orc_pivot_rows.argtypes = [vp, i, i, vp]
    L.orc_compress_block.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.orc_compress_block.restype = i
    L.orc_potentials.argtypes = [vp, vp, vp, vp, vp, vp]
    L.orc_element_near.argtypes = [vp, i, i, d, vp, vp]
    L.orc_fill_near.argtypes = [vp, vp, i, vp, i, d, vp]
    L.orc_num_threads.restype = i
    L.orc_set_threads.argtypes = [i]


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class Geometry:
    """Reference tables -> oracle geometry (0-based int32 ids internally)."""

    def __init__(self, vertex, topol, edges, ln, k=2 * np.pi, eta=120 * np.pi):
        self.vertex = np.ascontiguousarray(np.asarray(vertex, dtype=np.float64).T)  # nv x 3
        self.topol = np.ascontiguousarray(np.asarray(topol).T.astype(np.int32) - 1)
        self.edges = np.ascontiguousarray(np.asarray(edges).T.astype(np.int32) - 1)
        self.ln = np.ascontiguousarray(np.asarray(ln, dtype=np.float64).ravel())
        self.ne = self.ln.size
        self.k = float(k)
        self.eta = float(eta)
        self._g = _Geom(_ptr(self.vertex), _ptr(self.topol), _ptr(self.edges), _ptr(self.ln), self.ne, self.k, self.eta)

    @property
    def ptr(self):
        return ctypes.addressof(self._g)

    def element(self, m: int, n: int) -> complex:
        re, im = ctypes.c_double(), ctypes.c_double()
        lib().orc_element(self.ptr, int(m), int(n), ctypes.byref(re), ctypes.byref(im))
        return complex(re.value, im.value)

    def fill(self, rows, cols) -> np.ndarray:
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        Z = np.empty((cols.size, rows.size), dtype=np.complex128)  # column-major storage
        lib().orc_fill(self.ptr, _ptr(rows), rows.size, _ptr(cols), cols.size, _ptr(Z))
        return Z.T

    def fill_near(self, rows, cols, near_factor=2.0) -> np.ndarray:
        """Z[rows, cols] with triangle pairs closer than near_factor x the longer
        longest edge integrated by singularity extraction (orc_element_near)."""
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        Z = np.empty((cols.size, rows.size), dtype=np.complex128)
        lib().orc_fill_near(self.ptr, _ptr(rows), rows.size, _ptr(cols), cols.size, float(near_factor), _ptr(Z))
        return Z.T

    def block_matvec(self, rows, cols, v) -> np.ndarray:
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.complex128)
        y = np.empty(rows.size, dtype=np.complex128)
        lib().orc_block_matvec(self.ptr, _ptr(rows), rows.size, _ptr(cols), cols.size, _ptr(v), _ptr(y))
        return y


def _fortran(a: np.ndarray) -> np.ndarray:
    return np.asfortranarray(a, dtype=np.complex128)


def pinv(W: np.ndarray, atol: float = 1e-10, rtol: float = 0.0):
    """Returns (U, sigma (column order), kept)."""
    W = _fortran(W)
    r = W.shape[0]
    U = np.empty((r, r), dtype=np.complex128, order="F")
    sig = np.empty(r, dtype=np.float64)
    kept = lib().orc_pinv(_ptr(W), r, atol, rtol, _ptr(U), _ptr(sig))
    return U, sig, kept


def residual(C, U, R, v, y_old):
    C, U, R = _fortran(C), _fortran(U), _fortran(R)
    v = np.ascontiguousarray(v, dtype=np.complex128)
    y_old = np.ascontiguousarray(y_old, dtype=np.complex128)
    m, r = C.shape
    n = R.shape[1]
    y = np.empty(m, dtype=np.complex128)
    err = lib().orc_residual(_ptr(C), m, r, _ptr(U), _ptr(R), n, _ptr(v), _ptr(y_old), _ptr(y))
    return y, err


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.empty(4, dtype=np.uint32)
    lib().orc_philox(_ptr(c), _ptr(k), _ptr(o))
    return o


def sample(seed: int, block: int, rnd: int, purpose: int, N: int, r: int) -> np.ndarray:
    out = np.empty(r, dtype=np.int32)
    lib().orc_sample(seed, block, rnd, purpose, N, r, _ptr(out))
    return out


def sample_heldout(seed: int, block: int, rnd: int, m: int, h: int, r: int):
    """Returns (held-out positions, sample row positions)."""
    hp = np.empty(max(h, 1), dtype=np.int32)
    rows = np.empty(max(r, 1), dtype=np.int32)
    lib().orc_sample_heldout(seed, block, rnd, m, h, r, _ptr(hp), _ptr(rows))
    return hp[:h], rows[:r]


def vtest(seed: int, block: int, n: int) -> np.ndarray:
    v = np.empty(n, dtype=np.complex128)
    lib().orc_vtest(seed, block, n, _ptr(v))
    return v


def pivot_rows(C) -> np.ndarray:
    C = _fortran(C)
    m, r = C.shape
    out = np.empty(r, dtype=np.int32)
    lib().orc_pivot_rows(_ptr(C), m, r, _ptr(out))
    return out


def compress_block(geom: Geometry | None, rows=None, cols=None, dense=None, *, tol=1e-4, pinv_tol=1e-10,
                   pinv_rtol=0.0, seed=12345, row_mode=0, r_start=0, max_rank=0, block_id=0):
    """cuda_cur_compress3 restated. Returns dict with C, U, R, rs, cs (block positions), y, info."""
    if geom is not None:
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        m, n = rows.size, cols.size
        blk = _Block(geom.ptr, _ptr(rows), _ptr(cols), None, m, n)
    else:
        dense = _fortran(dense)
        m, n = dense.shape
        blk = _Block(None, None, None, _ptr(dense), m, n)
    D = min(m, n)
    cap = max_rank if 0 < max_rank < D else D
    C = np.empty(m * cap, dtype=np.complex128)
    U = np.empty(cap * cap, dtype=np.complex128)
    R = np.empty(cap * n, dtype=np.complex128)
    rs = np.empty(cap, dtype=np.int32)
    cs = np.empty(cap, dtype=np.int32)
    y = np.empty(m, dtype=np.complex128)
    o = _Opts(tol, pinv_tol, pinv_rtol, seed, row_mode, r_start, max_rank, block_id)
    info = _Info()
    lib().orc_compress_block(ctypes.addressof(blk), ctypes.addressof(o), _ptr(C), _ptr(U), _ptr(R), _ptr(rs),
                             _ptr(cs), _ptr(y), ctypes.addressof(info))
    r = info.r
    return {
        "C": C[: m * r].reshape((r, m)).T,
        "U": U[: r * r].reshape((r, r)).T,
        "R": R[: r * n].reshape((n, r)).T,
        "rs": rs[:r].copy(),
        "cs": cs[:r].copy(),
        "y": y,
        "r": r,
        "rounds": info.rounds,
        "kept": info.kept,
        "converged": bool(info.converged),
        "err": info.err,
    }


def potentials(v0, v1, v2, r):
    """(int_T 1/R dS', int_T (r' - rho)/R dS') over the triangle (v0, v1, v2)
    for the point r (Wilton et al. 1984 closed forms, orc_potentials)."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (v0, v1, v2, r)]
    Is = ctypes.c_double()
    Iv = np.zeros(3)
    lib().orc_potentials(*[_ptr(x) for x in a], ctypes.byref(Is), _ptr(Iv))
    return Is.value, Iv


def num_threads() -> int:
    return lib().orc_num_threads()


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def pinv_qr_trunc(W, atol=1e-10, rtol=0.0):
    """Truncated QR pseudo-inverse of the learned rounds' cores (this build's
    learned mode; its SVD alternative is pinv(), CU:1873-1924): Householder QR
    with column pivoting W P = Q R, the pivot at step j being the logical
    column with the largest trailing norm |A[j:, c]| (the first on ties);
    k = the leading |R_jj| >= tau = max(atol, rtol |R_00|); U = P [R_11^{-1}
    Q_k^H; 0]. Returns (U, |R_jj| in pivot order, k, pivots). Pins: inv(W) at
    full rank, LAPACK geqp3's pivots, U W e_c = e_c on the k pivot columns and
    W U an orthogonal projector of rank k (tests/test_oracle.py)."""
    A = np.array(W, dtype=np.complex128)
    r = A.shape[0]
    perm = np.arange(r)
    Q = np.eye(r, dtype=np.complex128)
    for j in range(r):
        norms = np.sum(np.abs(A[j:, perm[j:]]) ** 2, axis=0)
        p = j + int(np.argmax(norms))
        perm[[j, p]] = perm[[p, j]]
        x = A[j:, perm[j]].copy()
        xn = np.linalg.norm(x)
        if xn == 0:
            continue
        ph = x[0] / abs(x[0]) if x[0] != 0 else 1.0
        v = x.copy()
        v[0] += ph * xn
        v /= np.linalg.norm(v)
        cols = perm[j:]
        A[j:, cols] -= 2.0 * np.outer(v, v.conj() @ A[j:, cols])
        Q[:, j:] -= 2.0 * np.outer(Q[:, j:] @ v, v.conj())
    R = A[:, perm]
    d = np.abs(np.diag(R))
    tau = max(atol, rtol * d[0]) if r else 0.0
    k = 0
    while k < r and d[k] >= tau and d[k] > 0:
        k += 1
    U = np.zeros((r, r), dtype=np.complex128)
    if k:
        U[perm[:k], :] = np.linalg.solve(np.triu(R[:k, :k]), Q[:, :k].conj().T)
    return U, d, k, perm


def svd_compress(C, U, R, thresh):
    """SVD recompression of C U R (codes/ACA_vs_CUR/seed/svd_compress.m:19-34,
    called through postcompress_wrapper.m:1-5): QR of C and of R^H (economy),
    SVD of S = R_c U R_r^H, kept s_j / s_1 > thresh (strictly), returns
    (C' = Q_c u_k, s_k, R' = (Q_r v_k)^H). The reference's qr_cbd (absent from
    the tree) is taken as an economy QR: any orthonormal factorisation gives the
    same C' diag(s) R' and s. Pins: the singular values of the m x n product,
    orthonormal C' and R'^H, the Eckart-Young residual, exact reconstruction at
    thresh 0 (tests/test_oracle.py)."""
    C = np.asarray(C, dtype=np.complex128)
    U = np.asarray(U, dtype=np.complex128)
    R = np.asarray(R, dtype=np.complex128)
    qc, rc = np.linalg.qr(C)
    qr_, rr = np.linalg.qr(R.conj().T)
    u, s, vh = np.linalg.svd(rc @ U @ rr.conj().T)
    if s.size == 0 or s[0] == 0:
        k = 0
    else:
        k = int(np.sum(s / s.max() > thresh))
    return qc @ u[:, :k], s[:k], (qr_ @ vh.conj().T[:, :k]).conj().T


def aca(geom: Geometry, rows, cols, tol, max_rank=128):
    """Partially pivoted ACA of Z[rows, cols] (codes/ACA_vs_CUR/seed/ACA.m:40-128),
    rows and columns of the residual filled with the FP64 element. Returns
    (U m x k, V k x n, I, J, status); status 2 when max_rank crosses were taken
    without convergence, 3 when the residual row or column has no finite entry
    left to pivot on or the cross norms are not finite (the element is
    non-finite for coincident triangles). Blocks with one row or column are returned whole
    (ACA.m:47). Pins: exact recovery of a block of rank k in k steps, the
    interpolation property (U V matches Z on every pivot row and column), the
    normZ recursion equals |U V|_F^2 in exact arithmetic bounds (tests/test_oracle.py)."""
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    return aca_dense(lambda i: geom.fill(rows[i:i + 1], cols)[0],
                     lambda j: geom.fill(rows, cols[j:j + 1])[:, 0], rows.size, cols.size, tol, max_rank)


def aca_dense(row, col, M, N, tol, max_rank=128):
    """ACA.m:40-128 over callables row(i) -> Z[i, :] and col(j) -> Z[:, j]."""
    if M == 1 or N == 1:
        if N == 1:
            return col(0)[:, None].copy(), np.ones((1, 1), complex), np.zeros(1, int), np.zeros(1, int), 1
        return np.ones((1, 1), complex), row(0)[None, :].copy(), np.zeros(1, int), np.zeros(1, int), 1
    kend = min(M, N, max_rank)
    used_r = np.zeros(M, bool)
    used_c = np.zeros(N, bool)
    Us, Vs, I, J = [], [], [], []
    Ik, normZ, status = 0, 0.0, 0
    used_r[0] = True
    while True:
        R = row(Ik).astype(complex)
        for u, v in zip(Us, Vs):
            R -= u[Ik] * v
        a = np.abs(R) ** 2
        cand = ~used_c & np.isfinite(a)  # a non-finite entry is never a pivot
        if not cand.any():  # no finite residual entry left: status 3
            status = 3
            break
        Jk = int(np.argmax(np.where(cand, a, -1.0)))  # first index among equal maxima
        piv = R[Jk]
        if piv == 0:
            status = 1
            break
        used_c[Jk] = True
        I.append(Ik)
        J.append(Jk)
        Vk = R / piv
        Uk = col(Jk).astype(complex)
        for u, v in zip(Us, Vs):
            Uk -= u * v[Jk]
        nu, nv = np.vdot(Uk, Uk).real, np.vdot(Vk, Vk).real
        cross = sum(abs(np.vdot(u, Uk)) * abs(np.dot(Vk, v.conj())) for u, v in zip(Us, Vs))
        normZ += 2 * cross + nu * nv
        Us.append(Uk)
        Vs.append(Vk)
        k = len(Us)
        if not (np.isfinite(nu * nv) and np.isfinite(normZ)):  # NaN or Inf in the cross: status 3
            status = 3
            break
        if k > 1 and np.sqrt(nu * nv) <= tol * np.sqrt(normZ):
            status = 1
            break
        if k == min(M, N):
            status = 1
            break
        if k == kend:
            status = 2
            break
        a = np.abs(Uk) ** 2
        cand = ~used_r & np.isfinite(a)
        if not cand.any():
            status = 3
            break
        Ik = int(np.argmax(np.where(cand, a, -1.0)))
        used_r[Ik] = True
    U = np.array(Us).T if Us else np.zeros((M, 0), complex)
    V = np.array(Vs) if Vs else np.zeros((0, N), complex)
    return U, V, np.array(I), np.array(J), status
