/*
 * farcur — MI355X (gfx950) randomized pseudoskeleton (CUR) compression of the
 * far-field blocks of the RWG/EFIE impedance matrix. C ABI.
 *
 * Every function returns FARCUR_OK (0) on success or a nonzero status, with a
 * message available from farcur_last_error() (per host thread), except where
 * stated. Host pointers unless stated. Matrices are column-major complex128
 * (interleaved re, im), the layout of Julia's Matrix{ComplexF64}.
 *
 * Reference interfaces replaced (paths under the reference repository):
 *   CU  = codes/GPU_implementation/seed/user_impedance_for_julia_cuda.cu
 *   SW2 = codes/GPU_implementation/seed/compres_single_block_spheres_for_sweep2.jl
 */
#ifndef FARCUR_H
#define FARCUR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double re, im;
} farcur_c128;

enum {
    FARCUR_OK = 0,
    FARCUR_EINVAL = 1, /* bad argument (sizes, ids out of range, ...) */
    FARCUR_EHIP = 2,   /* HIP runtime error or no usable GPU */
    FARCUR_ENOMEM = 3, /* device allocation failed */
    FARCUR_ESTATE = 4  /* call out of order (no geometry, no result, ...) */
};

/* Message of the last failed call on this host thread ("" if none). */
const char* farcur_last_error(void);
const char* farcur_version(void);

/* ------------------------------------------------------------------ drop-in
 * Same symbol, argument list and semantics as the reference's
 * cuda_complete_compression (CU:1564-1629, body CU:1630-2012), as called by
 * pure_cuda_compression2 (SW2:591-660): one CUR round of the block r1 x r2 with
 * the caller's samples. Fills C = Z[r1, col_samples] (n_r1 x cur_order),
 * W = Z[row_samples, col_samples], R = Z[row_samples, r2] (cur_order x n_r2),
 * U = pinv(W) with absolute threshold pinv_tol, Avtest_new = C (U (R vtest)),
 * and returns ||Avtest_old - Avtest_new|| / ||Avtest_new||.
 * col_samples, row_samples, r1, r2 hold 1-based RWG ids stored as double
 * (the reference declares int* but receives Float64, CU:1696-1697, SW2:593).
 * vertex/topol/trian/edges/un/ds/ln/cent are the get_edge.m tables (1-based
 * ids as double). The launch-shape arguments (C/U/R blocks and threads,
 * blocks_x.., limx, limy) and Ztotal are accepted and ignored: this library
 * chooses its own launch shapes. field must be 1 (EFIE); Rinteg_s, Ranal_s,
 * Rinteg_f, cor_solid and flag are ignored (far-field blocks use quadrature
 * only). vec_time_cuda[0] receives the device time in ms of the round.
 * On error returns NaN, sets vec_time_cuda[0] = -1 and farcur_last_error().
 */
double cuda_complete_compression(farcur_c128* Ccompressed, farcur_c128* Ucompressed, farcur_c128* Rcompressed,
                                 int cur_order, double* vec_time_cuda, double pinv_tol, farcur_c128* Avtest_old,
                                 farcur_c128* Avtest_new, farcur_c128* vtest, double* col_samples,
                                 double* row_samples, int C_blocks_x, int C_blocks_y, int C_threads_x,
                                 int C_threads_y, int U_blocks_x, int U_blocks_y, int U_threads_x, int U_threads_y,
                                 int R_blocks_x, int R_blocks_y, int R_threads_x, int R_threads_y,
                                 farcur_c128* Ztotal, int field_arg, double k_arg, double eta_arg,
                                 double Rinteg_s_arg, double Ranal_s_arg, double Rinteg_f_arg, int cor_solid_arg,
                                 int flag_arg, double* r1_arg, double* r2_arg, double* vertex_arg,
                                 double* topol_arg, double* trian_arg, double* edges_arg, double* un_arg,
                                 double* ds_arg, double* ln_arg, double* cent_arg, int N_arg, int n_r1_arg,
                                 int n_r2_arg, int rows_vertex_arg, int cols_vertex_arg, int rows_topol_arg,
                                 int cols_topol_arg, int rows_trian_arg, int cols_trian_arg, int rows_edges_arg,
                                 int cols_edges_arg, int rows_un_arg, int cols_un_arg, int rows_ds_arg,
                                 int cols_ds_arg, int n_ln_arg, int rows_cent_arg, int cols_cent_arg, int blocks_x,
                                 int blocks_y, int threads_x, int threads_y, int limx, int limy);

/* ------------------------------------------------------------------ meshes
 * Host-side geometry producing the get_edge.m tables
 * (codes/GPU_implementation/seed/get_edge.m:16-121), which the reference builds
 * in Matlab (sphere.m, sq_plate.m) and passes through obj_struct (SW2:12-26).
 */
typedef struct farcur_mesh farcur_mesh;

/* Plate lx x ly in the plane z = z0 centred at (x0, y0), nx x ny squares
 * split into right triangles along one diagonal, normal +z. */
int farcur_mesh_plate(double lx, double ly, int nx, int ny, double x0, double y0, double z0, farcur_mesh** out);
/* Octahedral sphere, 12 * 4^n_levels edges, numbering of sphere.m
 * (codes/ACA_vs_CUR/seed/objects/sphere.m:13-53). */
int farcur_mesh_octasphere(double radius, int n_levels, farcur_mesh** out);
/* Icosphere (20 * 4^level triangles) scaled to semi-axes (ax, ay, az), radial
 * vertex jitter U(-jitter, jitter) x mean edge length from `seed`. */
int farcur_mesh_icosphere(double ax, double ay, double az, int level, double jitter, uint64_t seed,
                          farcur_mesh** out);
/* Appends src translated by (dx, dy, dz) to dst (RWG tables must be rebuilt). */
int farcur_mesh_append(farcur_mesh* dst, const farcur_mesh* src, double dx, double dy, double dz);
/* Builds the RWG tables (get_edge.m semantics; junctions are rejected). */
int farcur_mesh_build_rwg(farcur_mesh* mesh);
int farcur_mesh_sizes(const farcur_mesh* mesh, int* nv, int* nt, int* ne);
/* Copies the tables in the reference layout, ids 1-based stored as double:
 * vertex 3 x nv, topol 3 x nt, trian 3 x nt, edges 4 x ne, un 3 x nt, ds nt,
 * ln ne, cent 3 x nt. Any pointer may be NULL. */
int farcur_mesh_tables(const farcur_mesh* mesh, double* vertex, double* topol, double* trian, double* edges,
                       double* un, double* ds, double* ln, double* cent);
void farcur_mesh_destroy(farcur_mesh* mesh);
/* RWG-tested plane waves (centroid rule of user_compute_Ei.m,
 * codes/GPU_implementation/seed/user_compute_Ei.m:40-57): V (ne x nrhs,
 * column-major) for incidence angles theta[s], phi[s] (radians) and the
 * polarisation angle rot_eh (radians; pi/2 = theta-hat). Host computation. */
int farcur_mesh_plane_waves(const farcur_mesh* mesh, double k, const double* theta, const double* phi, int nrhs,
                            double rot_eh, farcur_c128* V);

/* Octree leaf groups of side leaf_size and the admissible (far) pairs
 * |c_a - c_b| >= eta_adm * max(diam_a, diam_b); the Matlab Z_comp partition
 * of the reference multi-block loop (codes/GPU_implementation/seed/
 * compres_single_block_spheres.jl:2539-2696, comp/self flags) is not shipped,
 * this builds the equivalent block list. */
typedef struct farcur_partition farcur_partition;
int farcur_partition_build(const farcur_mesh* mesh, double leaf_size, double eta_adm, farcur_partition** out);
int farcur_partition_sizes(const farcur_partition* p, int* ngroups, int* nrwg, int* nfar, int* nnear);
/* group_offsets ngroups+1, group_rwg nrwg (0-based RWG ids), centre 3 x ngroups,
 * diam ngroups, far_pairs 2 x nfar, near_pairs 2 x nnear (0-based group ids). */
int farcur_partition_tables(const farcur_partition* p, int* group_offsets, int* group_rwg, double* centre,
                            double* diam, int* far_pairs, int* near_pairs);
/* The block list as the reference's Z_comp entries {m, n, comp, self}
 * (codes/GPU_implementation/seed/compres_single_block_spheres.jl:1134-1144,
 * consumed at :2539-2696): nfar + nnear entries, far pairs first. Entry i is
 * group pair (a[i], b[i]); comp[i] = 1 for a far (compressed) block, 0 for a
 * near (dense) one; self[i] = 1 when the entry is the block (a, a) alone, 0
 * when it also stands for the transposed block (b, a) = (a, b)^T (the
 * reference fills transpose(Cs*Us*Rs) for self == 0, :2608-2612). Any
 * pointer may be NULL. */
int farcur_partition_blocks(const farcur_partition* p, int* a, int* b, int* comp, int* self);
/* Block-list file (the partition and its entries; layout in
 * gpu-randomized-pseudoskeleton_amd/csrc/geometry.hpp, with a checksum). load
 * checks the magic, version, checksum, id ranges and the entry flags, and
 * fails with FARCUR_EINVAL on any mismatch. */
int farcur_partition_save(const farcur_partition* p, const char* path);
int farcur_partition_load(const char* path, farcur_partition** out);
/* Build parameters and the RWG count of the mesh it was built on. */
int farcur_partition_params(const farcur_partition* p, double* leaf_size, double* eta_adm, int* nrwg_mesh);
void farcur_partition_destroy(farcur_partition* p);

/* ------------------------------------------------------------------ context
 * A context owns one GPU stream and the device-resident geometry (uploaded
 * once, unlike the per-call upload of CU:1711-1720). Calls on different
 * contexts may run concurrently; calls on one context must not overlap. */
typedef struct farcur_ctx farcur_ctx;

int farcur_create(int device, farcur_ctx** out);
void farcur_destroy(farcur_ctx* ctx);
/* Wavenumber and impedance (EM_data_struct k, eta, SW2:28-39). Default k = 2 pi, eta = 120 pi. */
int farcur_set_em(farcur_ctx* ctx, double k, double eta);
/* Geometry from the reference tables (1-based ids as double, as SW2:513-545
 * converts them): vertex 3 x nv, topol 3 x nt, edges 4 x ne, ln ne. */
int farcur_set_geometry(farcur_ctx* ctx, const double* vertex, int nv, const double* topol, int nt,
                        const double* edges, int ne, const double* ln);
int farcur_set_mesh(farcur_ctx* ctx, const farcur_mesh* mesh);

enum { FARCUR_F64 = 0, FARCUR_F32 = 1 };

/* Panel fill: Z[i + j*nr] = Z(rows[i], cols[j]) for 0-based RWG ids, the role of
 * impedance_matrix_compute_elbyel3 (CU:353-439) / impedance_matrix_cuda_elbyel4
 * (CU:1022-1267). precision FARCUR_F32 computes in complex FP32 and widens. */
int farcur_fill(farcur_ctx* ctx, const int* rows, int nr, const int* cols, int nc, int precision, farcur_c128* Z);

/* Pseudo-inverse of an r x r matrix W with singular values below
 * max(atol, rtol * sigma_max) dropped (CU:1873-1924 with kernel_Sinv CU:1435,
 * my_pinv SW2:1013-1022). sigma (r, descending) may be NULL; *kept receives the
 * number of singular values kept (may be NULL). */
int farcur_pinv(farcur_ctx* ctx, const farcur_c128* W, int r, double atol, double rtol, farcur_c128* U,
                double* sigma, int* kept);

/* The truncated QR pseudo-inverse that the learned rounds apply to their cores
 * (FARCUR_CORE_QR): W P = Q R with column pivoting (the pivot at step j is
 * the column with the largest trailing norm, the first on ties), k = the
 * leading |R_jj| >= max(atol, rtol |R_00|), U = P [R_11^{-1} Q_k^H; 0] (for
 * k = r, W^{-1}). No counterpart in the reference, which inverts every core
 * by SVD (farcur_pinv): for callers and tests of the learned core. Same
 * kernel, launch shapes and LDS form as the learned rounds: r <= 96 and
 * rtol >= 1e-8, else FARCUR_EINVAL. diag (r, may be NULL) receives |R_jj| in
 * pivot order, *kept (may be NULL) the k. */
int farcur_pinv_qr(farcur_ctx* ctx, const farcur_c128* W, int r, double atol, double rtol, farcur_c128* U,
                   double* diag, int* kept);

/* farcur_pinv without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid hipMalloc).
 * No counterpart in the reference. W, U (r x r), sigma (r, descending; may be
 * NULL) and kept (one int; may be NULL) are DEVICE pointers; W is not
 * modified. The call only enqueues its kernels on `stream` (a hipStream_t;
 * NULL: the context's stream) and returns: the results are ready when that
 * stream reaches them. Same kernels, launch shapes and threshold semantics
 * as farcur_pinv, so the outputs match it. Its scratch (W's working copy, the
 * Jacobi workspace, the job record) is a static device array of about 2 MiB,
 * set up once per device by farcur_create and held by every GPU the library
 * runs on until the process ends: r above 256 is rejected (FARCUR_EINVAL).
 * Every call on one GPU shares that array, so calls that overlap in time on
 * one device (other streams or contexts) race; serialise them. */
int farcur_pinv_noalloc(farcur_ctx* ctx, const farcur_c128* W, int r, double atol, double rtol, farcur_c128* U,
                        double* sigma, int* kept, void* stream);

/* y_new = C (U (R v)), returns err = ||y_old - y_new|| / ||y_new|| in *err
 * (CU:1928-1966). C m x r, U r x r, R r x n. */
int farcur_cur_residual(farcur_ctx* ctx, const farcur_c128* C, int m, int r, const farcur_c128* U,
                        const farcur_c128* R, int n, const farcur_c128* v, const farcur_c128* y_old,
                        farcur_c128* y_new, double* err);

/* Uniform sorted sample of r of the indices 0..N-1 for stream
 * (seed, block, round, purpose) (SW2:492-497, StatsBase.sample replace=false
 * then sort); purpose 1 = columns, 2 = rows. */
int farcur_sample(farcur_ctx* ctx, uint64_t seed, int block, int round, int purpose, int N, int r, int* out);

/* Test vector entries j < n of stream (seed, block): U[0,1) + i U[0,1)
 * (rand(N) + im*rand(N), SW2:888). Host computation, no GPU needed. */
int farcur_vtest(uint64_t seed, int block, int n, farcur_c128* v);

/* Row pivots of an m x r panel C: rows[j] = argmax_i |C[i, j]| (first on ties),
 * cur_smart_pinv2 (codes/GPU_implementation/seed/compres_single_block_spheres.jl:2870-2887). */
int farcur_pivot_rows(farcur_ctx* ctx, const farcur_c128* C, int m, int r, int* rows);

/* ------------------------------------------------------------------ adaptive CUR */
enum {
    FARCUR_ROWS_UNIFORM = 0, /* rows sampled like the columns (reference) */
    FARCUR_ROWS_PIVOT = 1    /* rows = argmax |C| per sampled column */
};

/* Core pseudo-inverse of the learned mode's rounds (complex-FP32 panels,
 * relative threshold tau = max(pinv_tol, pinv_rtol * scale), cores up to r = 96):
 *  FARCUR_CORE_QR: truncated QR with column pivoting, W P = Q R, k = the leading
 *    |R_jj| >= tau (scale |R_00|), U = P [R_11^{-1} Q_k^H; 0] (W^{-1} when k = r;
 *    with truncation C U R interpolates the k pivot columns). No SVD.
 *  FARCUR_CORE_SVD: the SVD pseudo-inverse of the plain modes (scale sigma_max).
 * Larger cores and FP64 learned rounds use the SVD pseudo-inverse. */
enum {
    FARCUR_CORE_QR = 0,
    FARCUR_CORE_SVD = 1
};

typedef struct {
    double tol;      /* successive-residual tolerance (reference experiments: 1e-1, SW2:1155) */
    double pinv_tol; /* absolute singular value threshold (reference: 1e-10, SW2:884) */
    double pinv_rtol;/* relative threshold (x sigma_max); 0 = reference behaviour */
    uint64_t seed;   /* RNG key for samples and test vectors */
    int precision;   /* FARCUR_F64 or FARCUR_F32 fill */
    int row_mode;    /* FARCUR_ROWS_UNIFORM or FARCUR_ROWS_PIVOT */
    int r_start;     /* 0: reference rule r0 = D < 10 ? 1 : D / 10 (SW2:889-894) */
    int max_rank;    /* cap on the sample count (0: min(m, n), the reference cap) */
    int block_id;    /* stream id of a single-block call */
    int heldout;     /* held-out rows per block in FARCUR_MODE_LEARNED (default 16; blocks of
                        m > 32 * heldout rows use ceil(m / 32), at most 4 * heldout) */
    int oversample;  /* p added to the predicted rank in FARCUR_MODE_LEARNED (default 18) */
    double label_eps;/* singular values >= label_eps * sigma_max are counted as rank_eps */
    double heldout_safety; /* learned mode accepts a round when its estimated block error <= heldout_safety * tol
                              (0.3); the estimate scales the held-out rows' error by the interpolation norms
                              |C[i, :] U|^2 of every unsampled row (a ratio estimator) */
    double growth;   /* learned mode: sample count factor after a failed check (1.5; the plain loop doubles) */
    int learned_core;/* learned mode's core pseudo-inverse: FARCUR_CORE_QR (default) or FARCUR_CORE_SVD;
                        the plain modes always use the reference's SVD pseudo-inverse */
} farcur_opts;

void farcur_opts_default(farcur_opts* opts);

typedef struct {
    int r;          /* sample count of the final round */
    int rounds;     /* rounds run */
    int kept;       /* singular values kept in the final core */
    int converged;  /* 1 if err <= tol, 0 if stopped at r = D or max_rank */
    double err;     /* final successive residual */
    double heldout_err; /* relative Frobenius error on held-out rows (-1 if not checked) */
    double sigma_max;   /* largest singular value of the final W */
    double time_ms;     /* device time of all rounds */
    int rank_eps;       /* singular values of the final W >= label_eps * sigma_max */
    int status;         /* 1 converged, 2 stopped at the sample cap, 3 NaN residual */
    int sweeps;         /* Jacobi sweeps of the final core pinv */
    int pad;
} farcur_block_info;

/* The adaptive loop of cuda_cur_compress3 (SW2:878-973) on one block, rows x
 * cols (0-based RWG ids), every round on the GPU with device sampling. The
 * result stays on the device until farcur_block_result. */
int farcur_compress_block(farcur_ctx* ctx, const int* rows, int m, const int* cols, int n, const farcur_opts* opts,
                          farcur_block_info* info);
/* Copies the last compress_block result: C m x r, U r x r, R r x n, samples
 * (0-based RWG ids, r each), y (m) = C U R v. Any pointer may be NULL. */
int farcur_block_result(farcur_ctx* ctx, farcur_c128* C, farcur_c128* U, farcur_c128* R, int* row_samples,
                        int* col_samples, farcur_c128* y);

/* ------------------------------------------------------------------ batches
 * Many blocks (the far-field block list) compressed together. Block b has rows
 * row_ids[row_off[b] .. row_off[b+1]) and cols col_ids[col_off[b] .. ). */
typedef struct farcur_batch farcur_batch;

enum {
    FARCUR_MODE_PLAIN = 0,   /* per-block host loop, one block at a time (drop-in semantics) */
    FARCUR_MODE_BATCHED = 1, /* every active block advances one round per batched launch set */
    FARCUR_MODE_LEARNED = 2  /* sample count from the rank model (or, with no model set, the reference's
                                r0 rule), one round checked on held-out rows, doubling fallback */
};

int farcur_batch_create(farcur_ctx* ctx, int nblocks, const int* row_off, const int* row_ids, const int* col_off,
                        const int* col_ids, farcur_batch** out);
void farcur_batch_destroy(farcur_batch* batch);
/* Compresses every block; *wall_ms receives the host wall time of the call
 * (synchronised), may be NULL. */
int farcur_batch_compress(farcur_batch* batch, const farcur_opts* opts, int mode, double* wall_ms);
/* Per-block results (nblocks entries). */
int farcur_batch_info(farcur_batch* batch, farcur_block_info* infos);
/* Copies block b's C, U, R (any may be NULL). */
int farcur_batch_block(farcur_batch* batch, int b, farcur_c128* C, farcur_c128* U, farcur_c128* R, int* row_samples,
                       int* col_samples);

/* SVD recompression of every block of the last compress to its relative rank
 * (svd_compress.m:19-34, reached through postcompress_wrapper.m:1-5; called as
 * matlab_postcompress(C, U, R, 0.1 * err) at
 * compres_single_block_spheres_for_lambda_sweep.jl:568). Per block:
 * C = Q_c R_c and R^H = Q_r R_r (Householder QR), R_c U R_r^H = u diag(s) v^H,
 * the k values with s_j / s_1 > thresh kept (strictly greater, as
 * find(drs > thresh)), C' = Q_c u_k (m x k), U' = diag(s_k), R' = (Q_r v_k)^H
 * (k x n). FP64 arithmetic; C' and R' are stored in the panel precision of the
 * compress. ranks (nblocks, may be NULL) receives k per block, *ms (may be
 * NULL) the device time. Blocks with r above 1024 are rejected. */
int farcur_batch_recompress(farcur_batch* batch, double thresh, int* ranks, double* ms);
/* Block b of the last recompression: C' (m x k), sigma (k, descending),
 * R' (k x n), *k; any pointer may be NULL. U' is diag(sigma). */
int farcur_batch_recompressed_block(farcur_batch* batch, int b, farcur_c128* C, double* sigma, farcur_c128* R,
                                    int* k);

/* Partially pivoted adaptive cross approximation of every block of the batch
 * with the FP64 element (ACA.m:40-128, the reference's comparator; its C twin
 * c_ACA.c:327-466): first row 0; each step takes the residual row, its largest
 * unused column (the first among ties), V_k = row / pivot, U_k = the residual
 * column; normZ += 2 sum_l |(U_l^H U_k)(V_k V_l^H)| + |U_k|^2 |V_k|^2; stops
 * after step k >= 2 when |U_k| |V_k| <= tol sqrt(normZ), on a zero pivot, at
 * min(m, n) or at max_rank crosses (0: 128); the next row is the largest unused
 * entry of U_k. Blocks with one row or column are returned as they are (U V =
 * the block, k = 1). ranks (nblocks, may be NULL) receives k, *ms the device time. */
int farcur_batch_aca(farcur_batch* batch, double tol, int max_rank, int* ranks, double* ms);
/* Block b of the last ACA: U (m x k), V (k x n), the row and column pivots
 * (block positions, k each), k and status (1 converged, 2 stopped at max_rank);
 * any pointer may be NULL. */
int farcur_batch_aca_block(farcur_batch* batch, int b, farcur_c128* U, farcur_c128* V, int* row_pivots,
                           int* col_pivots, int* k, int* status);

/* Eight features per block (nblocks x 8, row-major) for the rank model:
 * diameter of the row group / lambda, diameter of the column group / lambda,
 * centre distance / lambda, cosines between each group's mean normal and the
 * direction to the other group, k x mean RWG edge length, log(rows), log(cols). */
int farcur_batch_features(farcur_batch* batch, float* X);

/* ------------------------------------------------------------------ rank model
 * MLP 8 -> 64 -> 64 -> 1 (GELU) predicting log(eps-rank) of a block from its
 * features, trained and run on the GPU by this library's kernels. No
 * counterpart in the reference, whose sample count always starts at
 * min(m, n) / 10 (SW2:889-894) and doubles. */
typedef struct farcur_model farcur_model;
#define FARCUR_MLP_FEATURES 8
#define FARCUR_MLP_PARAMS 4801
#define FARCUR_MLP_NORM 18 /* feature means[8], stds[8], target mean, std */

int farcur_model_create(farcur_ctx* ctx, uint64_t seed, farcur_model** out);
void farcur_model_destroy(farcur_model* model);
int farcur_model_get(farcur_model* model, float* params, float* norm);
int farcur_model_set(farcur_model* model, const float* params, const float* norm);
/* Mean squared error (normalised target) and its gradient over n samples. */
int farcur_model_loss_grad(farcur_model* model, const float* X, const float* y, int n, float* loss, float* grad);
/* Adam on mini-batches; y = log(rank). The normalisation is taken from (X, y)
 * unless set before. epoch_loss (epochs entries) may be NULL. */
int farcur_model_train(farcur_model* model, const float* X, const float* y, int n, int epochs, int batch, float lr,
                       uint64_t seed, float* epoch_loss);
int farcur_model_predict(farcur_model* model, const float* X, int n, float* y);
/* Model used by FARCUR_MODE_LEARNED on this batch. */
int farcur_batch_set_model(farcur_batch* batch, farcur_model* model);
/* Blocks whose predicted round failed the held-out check in the last learned
 * compress (the miss count; not an error status). */
int farcur_batch_misses(farcur_batch* batch);

/* Two-stream pipeline of the batched rounds (no counterpart in the reference,
 * whose loop is per block): a round's jobs are cut into `chunks` parts of equal
 * estimated cost that alternate between two HIP streams, so that one part's
 * fill overlaps another's core pinv. 0 or 1 turns it off (default). Results
 * do not depend on it. Profiling (farcur_batch_profile) runs one part.
 * Measured on the 30,720-RWG sphere it does not pay (each phase already
 * fills the GPU), so it is off by default. */
int farcur_batch_set_pipeline(farcur_batch* batch, int chunks);

/* hipGraph capture of the batched rounds (no counterpart in the reference):
 * each round (plan upload, the persistent sample / fill / pinv / residual
 * launches, result download) is captured once and replayed as one graph
 * launch whenever a later round has the same plan content and launch
 * parameters (for instance every round of a repeated compress of the same
 * batch). Up to 8 graphs are kept per batch. On by default; profiling runs
 * eagerly. set_graphs(0) drops the cached graphs. graph_stats reports the
 * captures and replays since the batch was created. */
int farcur_batch_set_graphs(farcur_batch* batch, int enable);
int farcur_batch_graph_stats(farcur_batch* batch, int* captures, int* replays);

/* Multi-RHS product of the compressed far field: Y = sum_b C_b U_b R_b V[cols_b]
 * added to rows_b, and with symmetric = 1 also (C_b U_b R_b)^T V[rows_b] added
 * to cols_b (the transposed block of a symmetric operator, which the
 * reference builds with transpose(Cs*Us*Rs) for self == 0,
 * codes/GPU_implementation/seed/compres_single_block_spheres.jl:2608-2612).
 * V and Y are N x nrhs column-major (N = RWGs of the geometry); V is cast to
 * complex64 and Y accumulated in FP32. Rows of Y outside every target set are
 * zero. The blocks' row (and column) id lists must be disjoint or identical
 * (same ids in the same order).
 * mode: FARCUR_MV_VALU, FARCUR_MV_MFMA_F32, FARCUR_MV_MFMA_BF16X3.
 * The product runs in interpolative form, A_b = P_b R_b with P_b = C_b U_b
 * formed in FP64 on the first call after a compress (both grouped GEMMs then
 * have intermediates the size of the result; the core U alone is as
 * ill-conditioned as the kept W).
 * ms (may be NULL) receives 3 values: device ms of pass A, of pass B, and of
 * forming P (0 when already formed). */
enum { FARCUR_MV_VALU = 0, FARCUR_MV_MFMA_F32 = 1, FARCUR_MV_MFMA_BF16X3 = 2 };
int farcur_batch_matvec(farcur_batch* batch, const farcur_c128* V, int nrhs, farcur_c128* Y, int mode, int symmetric,
                        double* ms);

/* Single right-hand side: y = sum_b C_b U_b R_b v[cols_b] added to rows_b (and
 * with symmetric = 1 the transposed blocks), the residual product of CU:1928-1966
 * (three Zgemm with one column) over the whole far field at once. v and y are
 * N complex128 (cast to complex64 on the device, y accumulated in FP32); the
 * panels are read once, one workgroup per block and orientation, with the
 * interpolation matrix P_b = C_b U_b of farcur_batch_matvec (formed on the first
 * product call after a compress). *ms (may be NULL) receives the device time of
 * the product kernel (P not included). Blocks with r above 4096 are rejected. */
int farcur_batch_matvec1(farcur_batch* batch, const farcur_c128* v, farcur_c128* y, int symmetric, double* ms);

/* Accuracy of every block of the last compress against its FP64 dense block,
 * computed on the device (no counterpart kernel in the reference, whose check
 * is norm(C*U*R - Z) / norm(Z) on the host against a dense block,
 * codes/ACA_vs_CUR/seed/compres_single_block_spheres_for_lambda_sweep.jl:515-544).
 * Z_b's entries come from the FP64 element; C U R is formed as (C U) R in
 * FP64 from the stored panels. V: N x nrhs column-major (N = RWGs of the
 * geometry, 1 <= nrhs <= 16). Outputs (any may be NULL except V):
 *   fro_err[b] = ||Z_b - C_b U_b R_b||_F / ||Z_b||_F,
 *   mv_err[b]  = ||(Z_b - C_b U_b R_b) v_b|| / ||Z_b v_b||, v_b = V[cols_b, 0],
 *   Yref (N x nrhs) = sum_b Z_b V[cols_b, :] added to rows_b, and with
 *   symmetric = 1 also Z_b^T V[rows_b, :] added to cols_b: the FP64 dense
 *   far-field product farcur_batch_matvec approximates.
 * *ms (may be NULL) receives the device time. */
int farcur_batch_verify(farcur_batch* batch, const farcur_c128* V, int nrhs, int symmetric, double* fro_err,
                        double* mv_err, farcur_c128* Yref, double* ms);

/* ------------------------------------------------------------------ near field and operator
 * The non-admissible (near) blocks of the partition are dense, as the
 * reference's multi-block loop fills them (comp == 0 branch,
 * codes/GPU_implementation/seed/compres_single_block_spheres.jl:2639-2662,
 * self semantics :2602-2612). Their element integrates triangle pairs whose
 * centroids are closer than near_factor x the longer longest edge by
 * singularity extraction (closed-form 1/R potentials, Wilton et al. 1984;
 * 7-point rules for the remainder and the test triangle), the role of the
 * reference's Rinteg_s / Ranal_s singular rules
 * (compres_single_block_spheres_for_sweep2.jl:1095-1098); other pairs keep the
 * far element's 3 x 3 rule. FP64. */

/* Panel Z[i + j*nr] with the near-field element (host output). */
int farcur_fill_near(farcur_ctx* ctx, const int* rows, int nr, const int* cols, int nc, double near_factor,
                     farcur_c128* Z);

typedef struct farcur_near farcur_near;
/* Fills every block (rows x cols lists as in farcur_batch_create) into device
 * memory (complex128, column-major). transpose[b] = 1 (may be NULL: all 0)
 * when the block also stands for its transpose at (cols, rows), the
 * reference's self == 0. */
int farcur_near_create(farcur_ctx* ctx, int nblocks, const int* row_off, const int* row_ids, const int* col_off,
                       const int* col_ids, const int* transpose, double near_factor, farcur_near** out);
int farcur_near_block(farcur_near* near, int b, farcur_c128* Z);
void farcur_near_destroy(farcur_near* near);

/* The compressed operator: near dense blocks plus the far CUR blocks of a
 * compressed batch (far_symmetric = 1: each far block also applies its
 * transpose, self == 0). Either part may be NULL. */
typedef struct farcur_operator farcur_operator;
int farcur_operator_create(farcur_near* near, farcur_batch* far, int far_symmetric, farcur_operator** out);
void farcur_operator_destroy(farcur_operator* op);
/* y = Z~ x (N complex128 each, host). The near part runs in FP64, the far
 * part in the compressed panels' precision (FP32 accumulation). */
int farcur_operator_matvec(farcur_operator* op, const farcur_c128* x, farcur_c128* y);
/* Z~ x = b for nrhs right-hand sides (N x nrhs column-major) by restarted
 * GMRES(restart) from x = 0, every product on the device: stops at a relative
 * residual <= tol or after maxit products; iters and relres (nrhs each, may
 * be NULL) receive the products used and the final true relative residual.
 * The reference solves the same system directly, J = -Z \ Ei
 * (compres_single_block_spheres.jl:3773). */
int farcur_operator_solve(farcur_operator* op, const farcur_c128* b, farcur_c128* x, int nrhs, double tol,
                          int restart, int maxit, int* iters, double* relres);

/* Device time (ms) and launch count of each kernel class over the last
 * batch_compress call with profiling on: names "sample", "fill", "pinv",
 * "residual", then per pinv class "pinv_r16", "pinv_r20", ... "pinv_r64" (the
 * largest core size of the class) and "pinv_global",
 * then the host wall time of the last compress (measured with or without
 * profiling): "host_predict" (rank inference round trip), "host_plan" (round
 * planning and upload staging), "host_wait" (waiting for the device),
 * "host_advance" (block state updates). Returns the number of entries written
 * (<= cap). */
int farcur_batch_profile(farcur_batch* batch, int enable);
int farcur_batch_kernel_times(farcur_batch* batch, const char** names, double* ms, int* launches,
                              double* evals, int cap);

/* ------------------------------------------------------------------ tuning
 * Launch shapes of the batched kernels. The reference fixes its thread
 * blocks from a `divfactor` sweep (codes/GPU_implementation/run_experiment.sh:8,
 * compres_single_block_spheres_for_sweep2.jl:574-587); here farcur_autotune
 * times candidate shapes on the device and keeps the fastest per context.
 * pinv classes: sample counts (0,16], (16,20], (20,24], (24,28], (28,32],
 * (32,36], (36,40], (40,48], (48,64] (LDS variant sized for the class
 * maximum, so that small cores fit more workgroups per CU), (64,96] (LDS in
 * the one-buffer form of the learned FP32 rounds, else global) and > 96
 * (global). */
#define FARCUR_PINV_CLASSES 11
typedef struct {
    int pinv_threads[FARCUR_PINV_CLASSES];
    int fill_grid;     /* fill workgroups per launch (capped at the launch's tile count) */
    int resid_threads;
    double pinv_ms[FARCUR_PINV_CLASSES]; /* autotune: best time of the class's probe (ms) */
} farcur_tuning;

int farcur_autotune(farcur_ctx* ctx, farcur_tuning* out);
int farcur_get_tuning(farcur_ctx* ctx, farcur_tuning* out);
int farcur_set_tuning(farcur_ctx* ctx, const farcur_tuning* in);
/* Device time (ms, best of 3) of the pinv of nprob random r x r problems with
 * `threads` per workgroup, LDS (use_lds = 1, r <= 64) or global variant;
 * *sweeps (may be NULL) receives the mean Jacobi sweep count. */
int farcur_pinv_bench(farcur_ctx* ctx, int r, int nprob, int threads, int use_lds, double* ms, double* sweeps);

#ifdef __cplusplus
}
#endif

#endif /* FARCUR_H */
