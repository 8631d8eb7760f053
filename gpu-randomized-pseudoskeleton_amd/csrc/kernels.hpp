// Work descriptors and launchers of the far-field CUR kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"
#include "element.hpp"

namespace farcur {

// One block in one round of the adaptive CUR. Every pointer is device memory.
// Layouts (column-major, as the reference's panels, CU:1738-1871):
//   C  m x r (ld m): rows = block rows, cols = sampled cols
//   W  r x r (ld r): rows = sampled rows, cols = sampled cols (destroyed by the pinv)
//   R  r x n (ld r): rows = sampled rows, cols = block cols
//   U  r x r (ld r): W^+
struct BlockJob {
    int32_t b;      // block id (selects the RNG streams)
    int32_t m, n, r, round;
    int32_t sample_mode;  // 0: draw cols+rows, 1: samples given, 2: draw cols, rows by argmax |C|,
                          // 3: draw h held-out rows, then cols and rows among the rest
    int32_t h;            // held-out rows (sample_mode 3), 0 otherwise
    int32_t pad1;
    const int32_t* rows;  // m RWG ids
    const int32_t* cols;  // n RWG ids
    int32_t* cs;          // r sampled col ids (RWG ids)
    int32_t* rs;          // r sampled row ids (RWG ids)
    int32_t* rpos;        // r sampled row positions in the block, written by the sampler;
                          // when set, W = C[rpos, :] is gathered from C instead of filled
    void* C;              // c128 or c64 (panel precision)
    void* R;
    c128* W;
    c128* V;  // Jacobi scratch when r exceeds the LDS variant
    c128* U;
    double* sigma;        // r
    const c128* v;        // n test vector
    const c128* y_old;    // m
    c128* y_new;          // m
    int32_t* hs;          // h held-out row ids (RWG ids)
    int32_t* hpos;        // h held-out row positions in the block
    void* H;              // h x n exact rows Z[hs, cols] (ld h), panel precision
    const int32_t* row_perm;  // FP32 fill order of the rows (positions, 64 per tile), or NULL
    const int32_t* col_perm;  // same for the columns
};

// Per-job results written by the kernels.
struct JobOut {
    double err;       // successive residual ||y_old - y|| / ||y||, or the held-out
                      // rows' relative Frobenius error when h > 0
    double sigma_max; // largest singular value of W
    int32_t kept;     // singular values kept by the threshold
    int32_t rank_eps; // singular values >= label_eps * sigma_max (rank label)
    int32_t sweeps;   // Jacobi sweeps run
    int32_t pad;
    double err_est;   // h > 0: estimated relative Frobenius error of the whole block
                      // (ratio estimator over the interpolation norms |C[i, :] U|^2, heldout_error)
};

// One block of an ACA (k_aca): U (m x kmax, ld m), V rows as the columns of
// Vt (n x kmax, ld n), `row` (n) and `used` (m + n) scratch, pivots I, J (kmax).
struct AcaJob {
    int32_t m, n, kmax, pad;
    const int32_t* rows;
    const int32_t* cols;
    c128* U;
    c128* Vt;
    c128* row;
    int32_t* used;
    int32_t* I;
    int32_t* J;
    int32_t* k_out;
    int32_t* status;  // 1 converged (or exhausted min(m, n)), 2 stopped at kmax
};
// One block of an SVD recompression (recompress.hip).
struct RecompressJob {
    int32_t m, n, r, pad;
    const void* C;  // m x r (ld m), panel precision
    const c128* U;  // r x r (ld r)
    const void* R;  // r x n (ld r), panel precision
    c128* ws;       // recompress_ws_elems(m, n, r) elements
    void* Cout;     // m x k (ld m), panel precision
    void* Rout;     // k x n (ld k), panel precision
    double* sigma;  // k kept singular values, descending
    int32_t* k_out; // k
};

// Largest r the recompression takes, and its workspace in c128 elements.
constexpr int kRecompressMaxR = 1024;
size_t recompress_ws_elems(int m, int n, int r);
void launch_recompress(const RecompressJob* jobs, int njobs, double thresh, int f32, hipStream_t s);

struct FillParams {
    const void* rec;     // RWG records (double or float SoA, kRecFields x ne)
    const double* cen;   // RWG centres, 3 x ne SoA (used by the FP32 fill)
    TriGeom tri;         // triangle tables (used by the FP32 fill)
    int32_t ne;
    int32_t f32;      // 1: FP32 fill, panels c64; 0: FP64 fill, panels c128
    double k, eta;
};

void launch_aca(const AcaJob* jobs, int njobs, double tol, const FillParams& fp, hipStream_t s);

// Tiles of the fill for a job of size (m, n, r, h held-out rows) restricted to `panels`.
int fill_tiles(int m, int n, int r, int h, int panels);

void launch_sample(const BlockJob* jobs, int njobs, uint64_t seed, hipStream_t s);
void launch_pivot_rows(const BlockJob* jobs, int njobs, int f32, hipStream_t s);
// panels: bit 0 = C, bit 1 = W, bit 2 = R. tile_prefix: njobs + 1 device ints
// (tiles of job j are [tile_prefix[j], tile_prefix[j+1])).
enum : int { kPanelC = 1, kPanelW = 2, kPanelR = 4, kPanelH = 8 };
void launch_fill(const BlockJob* jobs, int njobs, const int32_t* tile_prefix, int total_tiles, int panels,
                 const FillParams& fp, int grid, hipStream_t s);
// Pinv of W for the listed jobs (indices into jobs). rmax in 1..kPinvLdsMaxR
// selects the LDS variant sized for rmax (every listed job has r <= rmax);
// rmax = 0 selects the global-memory variant.
constexpr int kPinvLdsMaxR = 64;
// The one-buffer LDS form (every core gathered from C, relative threshold
// >= 1e-8: the learned FP32 rounds) fits cores up to this size in 160 KB.
constexpr int kPinvLdsOneBufMaxR = 96;
static_assert(kPinvLdsOneBufMaxR <= 128, "the QR pivot search takes at most two columns per lane");
// r classes (0,16], (16,20], ..., (48,64], (64,96], > 96: the LDS of a launch is
// sized for its class maximum, so finer classes let small cores pack more
// workgroups per CU. The (64,96] class runs in LDS only in the one-buffer form,
// otherwise in global memory.
constexpr int kPinvClasses = 11;
constexpr int kPinvClassMax[kPinvClasses] = {16, 20, 24, 28, 32, 36, 40, 48, 64, 96, 0};
inline int pinv_class(int r) {
    for (int c = 0; c < kPinvClasses - 1; ++c)
        if (r <= kPinvClassMax[c]) return c;
    return kPinvClasses - 1;
}
size_t pinv_lds_bytes(int rmax, bool one_buf);
// Complex elements of the Jacobi scratch V for sample count r (V plus, past
// r = 1024, the r weights).
inline size_t pinv_v_elems(int r) { return (size_t)r * r + (size_t)(r + 1) / 2; }
// Largest sample count any entry point accepts.
constexpr int kMaxSamples = 8192;
// f32: panel precision of C (for jobs whose W is gathered from C). one_buf:
// every listed job gathers W from C (rpos set), so with rtol >= 1e-8 the LDS
// variant may use one r x r buffer (W panel and U as global scratch).
// heldout: the LDS variant also computes the held-out error of jobs with h > 0
// (launch_residual then skips them: skip_heldout_rmax = pinv_lds_limit).
// qr_core: in the one-buffer LDS form, jobs with h > 0 (learned rounds) take the
// truncated QR pseudo-inverse U = P [R_11^{-1} Q_k^H; 0] instead of the SVD one;
// 2: every job of that form (timing probes without held-out rows).
void launch_pinv(const BlockJob* jobs, const int32_t* list, int nlist, double atol, double rtol, double label_eps,
                 int rmax, int threads, int f32, int one_buf, int heldout, int qr_core, JobOut* out, hipStream_t s);
// Largest core the LDS variant takes for a launch with these flags.
inline int pinv_lds_limit(int one_buf, double rtol) {
    return one_buf && rtol >= 1e-8 ? kPinvLdsOneBufMaxR : kPinvLdsMaxR;
}
// skip_heldout_rmax > 0: jobs with h > 0 and r <= skip_heldout_rmax had their
// held-out error computed by the pinv launch and are skipped.
void launch_residual(const BlockJob* jobs, int njobs, int f32, int threads, int skip_heldout_rmax, JobOut* out,
                     hipStream_t s);

// A block whose sample took every row (r = m: R = Z[rows, :] exactly) or
// every column (r = n: C = Z[:, cols] exactly) is stored exactly: C = I, U = I
// (rows, side 0) or U = I, R = I (columns, side 1), so that C U R is the block
// itself instead of a truncated projection of it.
struct ExactJob {
    void* C;  // m x r, panel precision
    c128* U;  // r x r
    void* R;  // r x n, panel precision
    int32_t m, n, r, side;
};
void launch_exact(const ExactJob* jobs, int njobs, int f32, hipStream_t s);

// Plain panel fill Z[i + j*ld] = Z(rows[i], cols[j]) (standalone, one panel).
void launch_fill_panel(void* Z, int ld, const int32_t* rows, int nr, const int32_t* cols, int nc,
                       const FillParams& fp, hipStream_t s);
// v[offsets[b] + j] = test-vector entry j of block b, j < offsets[b+1] - offsets[b].
void launch_vtest(c128* v, const int32_t* offsets, int nblocks, uint64_t seed, hipStream_t s);

}  // namespace farcur
