/*
 * skinny.h -- C ABI of the MI355X tall-and-skinny matrix product library.
 *
 * Every product entry point keeps the argument order and meaning of the
 * reference operator boundary
 *
 *   using MatmulFunctionType = std::function<bool(
 *       size_t blockCount, int M, int N, int K, const dtype* A, int lda,
 *       dtype alpha, const dtype* B, int ldb, dtype beta, dtype* C, int ldc)>;
 *                                              (versions.hpp:43-45)
 *
 * with 64-bit sizes, an explicit library handle and an explicit HIP stream
 * appended.  All matrices are caller-owned device pointers; the calls are
 * asynchronous on `stream` and never synchronise or allocate once the
 * handle's workspace is large enough (see skinny_reserve).
 *
 *   TSMTTSM  C = alpha * A^T * B + beta * C
 *            A: K x M row-major (lda >= M), B: K x N row-major (ldb >= N),
 *            C: M x N column-major (ldc >= M).  Complex types use the plain
 *            transpose (no conjugation).  A == B with lda == ldb, M == N is
 *            the Gram (self) case and reads A once.
 *   TSMM     C = alpha * A * B + beta * C
 *            A: K x M row-major (lda >= M), B: M x N column-major (ldb >= M),
 *            C: K x N row-major (ldc >= N).  C may alias A (in place) when
 *            lda == ldc >= max(M, N); every row of A is read before the same
 *            row of C is written.
 *   beta == 0  C is not read (BLAS semantics).
 *
 * blockCount: number of workgroups along K (TSMTTSM: the split-K partial
 * count; TSMM: the row-parallel grid).  0 selects the tuned value.  Results of
 * TSMTTSM are bitwise reproducible for a fixed blockCount.
 *
 * Return values: SKINNY_OK, or SKINNY_UNSUPPORTED where the reference variant
 * would return false (unsupported shape / aliasing), SKINNY_INVALID_ARG for
 * arguments no variant accepts (negative sizes, ld too small), SKINNY_HIP_ERROR
 * and SKINNY_OOM for runtime failures.  skinny_last_error() describes the most
 * recent failure on the calling thread.
 */
#ifndef SKINNY_H
#define SKINNY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct skinny_handle_s* skinny_handle_t;
/* a hipStream_t; 0 is the null stream */
typedef void* skinny_stream_t;

typedef enum {
  SKINNY_OK = 0,
  SKINNY_UNSUPPORTED = 1,
  SKINNY_INVALID_ARG = 2,
  SKINNY_HIP_ERROR = 3,
  SKINNY_OOM = 4
} skinny_status_t;

typedef struct {
  float x, y;
} skinny_float_complex;
typedef struct {
  double x, y;
} skinny_double_complex;
/* a double-double value s + t (the reference's PseudoQuad, PseudoQuad.cuh:5-8) */
typedef struct {
  double s, t;
} skinny_dd;

/* ---- library / handle ------------------------------------------------- */
/* Replaces the reference's namespace-static scratch (d_temp_storage,
 * genv3.cuh:82-98) and static cuBLAS handles (gen_cublas.cuh:121-131) with
 * a per-handle workspace on the device that is current at creation. */
skinny_status_t skinny_create(skinny_handle_t* handle);
skinny_status_t skinny_destroy(skinny_handle_t handle);
/* Each stream a handle is used on gets its own workspace, so calls on
 * different streams may overlap; calls on one stream are ordered by it.
 * skinny_reserve_stream grows the workspace of `stream` to at least `bytes`
 * of scratch now (allocation happens here, not in later calls), as required
 * before capturing calls on that stream into a hipGraph; skinny_reserve does
 * it for the null stream. */
skinny_status_t skinny_reserve(skinny_handle_t handle, size_t bytes);
skinny_status_t skinny_reserve_stream(skinny_handle_t handle, size_t bytes,
                                      skinny_stream_t stream);
/* Workspace bytes a TSMTTSM call of this shape needs (elem = 4, 8 or 16). */
size_t skinny_tsmttsm_workspace(skinny_handle_t handle, size_t blockCount,
                                int64_t M, int64_t N, int64_t K,
                                int elem_bytes);
/* Process-wide handle used by the C++ drop-in shim (created on first use). */
skinny_handle_t skinny_default_handle(void);
const char* skinny_status_string(skinny_status_t s);
const char* skinny_last_error(void);
/* Compute units of the handle's device and the blockCount 0 resolves to. */
int skinny_cu_count(skinny_handle_t handle);
int skinny_tuned_blockcount_tsmttsm(skinny_handle_t handle, int64_t M,
                                    int64_t N, int64_t K, int elem_bytes);
int skinny_tuned_blockcount_tsmm(skinny_handle_t handle, int64_t M, int64_t N,
                                 int64_t K, int elem_bytes);
/* The grid a real TSMM call (elem_bytes 4 or 8) with blockCount 0 launches,
 * for beta == 0 (beta_zero != 0) or not and in place (C == A) or not: the
 * kernel choice and its workgroups per CU depend on both.  The reference
 * fixes its grids per variant in the perf driver (perf.cu:186-204). */
int skinny_tuned_blockcount_tsmm_ex(skinny_handle_t handle, int64_t M,
                                    int64_t N, int64_t K, int elem_bytes,
                                    int beta_zero, int in_place);

/* ---- TSMTTSM: C = alpha A^T B + beta C --------------------------------- */
/* replaces GENV1..8/SPECSMALL::tsmttsm<T,...> and tsmttsm_cublas<T>
 * (tsmttsm/genv3.cuh:85-112, tsmttsm/gen_cublas.cuh:124-160); types per
 * types.hpp:9-57: s = FR, d = DR, c = FC, z = DC, h = HR */
skinny_status_t skinny_dtsmttsm(skinny_handle_t h, size_t blockCount,
                                int64_t M, int64_t N, int64_t K,
                                const double* A, int64_t lda, double alpha,
                                const double* B, int64_t ldb, double beta,
                                double* C, int64_t ldc, skinny_stream_t stream);
/* No counterpart in the reference: the FP64 TSMTTSM above for callers that
 * cannot allocate device memory (code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, plan, kernels and
 * results as skinny_dtsmttsm at the same blockCount; the split-K partials and
 * arrival counters live in a fixed device array inside the library (one per
 * GPU, SKINNY_NOALLOC_PARTIAL_BYTES of partials, held until the process
 * ends) instead of a handle's workspace, so no call allocates, frees or
 * synchronises.  The handle may be NULL (current device, untuned grid; no
 * default handle is created); a given handle is used for its device and
 * tuned grids only.  blockCount 0 takes the largest default split whose
 * partials fit; an explicit blockCount whose partials (blockCount x tiles x
 * tile entries x 8 bytes) exceed the limit returns SKINNY_INVALID_ARG.
 * All calls on one GPU share its array: calls that overlap in time on one
 * GPU (other streams, unordered host threads) race and must not be made. */
#define SKINNY_NOALLOC_PARTIAL_BYTES (16u << 20)
skinny_status_t skinny_dtsmttsm_noalloc(skinny_handle_t h, size_t blockCount,
                                        int64_t M, int64_t N, int64_t K,
                                        const double* A, int64_t lda,
                                        double alpha, const double* B,
                                        int64_t ldb, double beta, double* C,
                                        int64_t ldc, skinny_stream_t stream);
skinny_status_t skinny_stsmttsm(skinny_handle_t h, size_t blockCount,
                                int64_t M, int64_t N, int64_t K,
                                const float* A, int64_t lda, float alpha,
                                const float* B, int64_t ldb, float beta,
                                float* C, int64_t ldc, skinny_stream_t stream);

/* complex (DC, FC): plain transpose; M, N <= 512 */
skinny_status_t skinny_ztsmttsm(skinny_handle_t h, size_t blockCount,
                                int64_t M, int64_t N, int64_t K,
                                const skinny_double_complex* A, int64_t lda,
                                skinny_double_complex alpha,
                                const skinny_double_complex* B, int64_t ldb,
                                skinny_double_complex beta,
                                skinny_double_complex* C, int64_t ldc,
                                skinny_stream_t stream);
skinny_status_t skinny_ctsmttsm(skinny_handle_t h, size_t blockCount,
                                int64_t M, int64_t N, int64_t K,
                                const skinny_float_complex* A, int64_t lda,
                                skinny_float_complex alpha,
                                const skinny_float_complex* B, int64_t ldb,
                                skinny_float_complex beta,
                                skinny_float_complex* C, int64_t ldc,
                                skinny_stream_t stream);

/* Ozaki split on the INT8 matrix cores (no reference counterpart): each
 * FP64 input is rounded to an integer of 8 slices - 2 bits against a running
 * per-column power-of-two scale and written in `slices` (3..5) signed-byte
 * limbs, with exact INT32 accumulation per limb level s + t < L (L = 3 for
 * 3 slices, slices + 1 for 4 and 5) and FP64 reconstruction when a
 * column's scale grows, when a chunk falls 9+ bits below it, or every 16384
 * rows.  Relative error at most about 2^(-(8 slices - 2)) (3 slices: 2^-20)
 * of sum |A||B|, in practice far below (3 slices meet 8 K eps on the
 * cancelling-column data at K = 2^24 and are the fast setting); with
 * blockCount == 1 the computation is one fixed sequence.
 * M, N <= 1024; lda, ldb <= 2^22. */
skinny_status_t skinny_dtsmttsm_ozaki(skinny_handle_t h, size_t blockCount,
                                      int64_t M, int64_t N, int64_t K,
                                      const double* A, int64_t lda,
                                      double alpha, const double* B,
                                      int64_t ldb, double beta, double* C,
                                      int64_t ldc, int slices,
                                      skinny_stream_t stream);
/* Lane layout of the INT8 32x32x32 matrix-core result (diagnostic):
 * host_out[lane * 16 + r] = 32 * row, host_out[1024 + lane * 16 + r] =
 * 32 * column of result element r of lane `lane`.  Synchronous. */
skinny_status_t skinny_probe_i8_layout(int* host_out);

/* HR (types.hpp:49-57): double inputs and output, double-double (PseudoQuad)
 * accumulation through FMA TwoProd and the accurate dd addition; replaces the
 * iT = PseudoQuad instantiations (genv3.cuh:46-79, specsmall.cuh) and their
 * epilogue convert(scale2(renormalize(sum), alpha)) + beta C
 * (genv3.cuh:19-26).  M, N <= 1024. */
skinny_status_t skinny_htsmttsm(skinny_handle_t h, size_t blockCount,
                                int64_t M, int64_t N, int64_t K,
                                const double* A, int64_t lda, double alpha,
                                const double* B, int64_t ldb, double beta,
                                double* C, int64_t ldc, skinny_stream_t stream);
/* the same accumulation with the double-double result kept:
 * C = alpha * sum + beta * C evaluated in dd, C an array of (s, t) pairs */
skinny_status_t skinny_htsmttsm_dd(skinny_handle_t h, size_t blockCount,
                                   int64_t M, int64_t N, int64_t K,
                                   const double* A, int64_t lda, double alpha,
                                   const double* B, int64_t ldb, double beta,
                                   skinny_dd* C, int64_t ldc,
                                   skinny_stream_t stream);

/* ---- TSMM: C = alpha A B + beta C -------------------------------------- */
/* replaces tsmm_fix1..4/fix_fb/fix_blend/var1/varip1..3/varip_blend/
 * fix_ip_ghost/var_ip_ghost and tsmm_cublas<T> (tsmm/fix3.cuh:41-65,
 * tsmm/varip3.cuh, tsmm/cublas.cuh:178-222) */
skinny_status_t skinny_dtsmm(skinny_handle_t h, size_t blockCount, int64_t M,
                             int64_t N, int64_t K, const double* A,
                             int64_t lda, double alpha, const double* B,
                             int64_t ldb, double beta, double* C, int64_t ldc,
                             skinny_stream_t stream);
skinny_status_t skinny_stsmm(skinny_handle_t h, size_t blockCount, int64_t M,
                             int64_t N, int64_t K, const float* A, int64_t lda,
                             float alpha, const float* B, int64_t ldb,
                             float beta, float* C, int64_t ldc,
                             skinny_stream_t stream);

/* complex (DC, FC): M, N <= 512; in place as for the real types.
 * Widths: real M, N <= 1024 (complex 512); above 128 real-view columns the
 * row band of A is staged in LDS and B is packed into the stream's
 * workspace (tsmm/var1.cuh takes any width out of place, varip1..3 in place
 * up to 256 columns). */
skinny_status_t skinny_ztsmm(skinny_handle_t h, size_t blockCount, int64_t M,
                             int64_t N, int64_t K,
                             const skinny_double_complex* A, int64_t lda,
                             skinny_double_complex alpha,
                             const skinny_double_complex* B, int64_t ldb,
                             skinny_double_complex beta,
                             skinny_double_complex* C, int64_t ldc,
                             skinny_stream_t stream);
skinny_status_t skinny_ctsmm(skinny_handle_t h, size_t blockCount, int64_t M,
                             int64_t N, int64_t K,
                             const skinny_float_complex* A, int64_t lda,
                             skinny_float_complex alpha,
                             const skinny_float_complex* B, int64_t ldb,
                             skinny_float_complex beta,
                             skinny_float_complex* C, int64_t ldc,
                             skinny_stream_t stream);

/* ---- TSMM in place, GHOST argument order ------------------------------ */
/* replaces the kernels fix_ip_ghost_kernel / var_ip_ghost_kernel
 * (tsmm/fix_ip_ghost.cuh:4-21, tsmm/var_ip_ghost.cuh:7-27):
 *   x(r, n) <- beta x(r, n) + alpha sum_m x(r, m) w[n stridew + m]
 * for rows r < nrows, n < ncolsout, m < ncolsin; x row-major with row stride
 * stridex >= max(ncolsin, ncolsout).  Same kernels, limits and status codes
 * as skinny_[dszc]tsmm with A = C = x, blockCount 0. */
skinny_status_t skinny_dtsmm_inplace(skinny_handle_t h, double* x,
                                     const double* w, double alpha,
                                     double beta, int64_t nrows,
                                     int64_t ncolsin, int64_t ncolsout,
                                     int64_t stridex, int64_t stridew,
                                     skinny_stream_t stream);
skinny_status_t skinny_stsmm_inplace(skinny_handle_t h, float* x,
                                     const float* w, float alpha, float beta,
                                     int64_t nrows, int64_t ncolsin,
                                     int64_t ncolsout, int64_t stridex,
                                     int64_t stridew, skinny_stream_t stream);
skinny_status_t skinny_ztsmm_inplace(skinny_handle_t h,
                                     skinny_double_complex* x,
                                     const skinny_double_complex* w,
                                     skinny_double_complex alpha,
                                     skinny_double_complex beta, int64_t nrows,
                                     int64_t ncolsin, int64_t ncolsout,
                                     int64_t stridex, int64_t stridew,
                                     skinny_stream_t stream);
skinny_status_t skinny_ctsmm_inplace(skinny_handle_t h,
                                     skinny_float_complex* x,
                                     const skinny_float_complex* w,
                                     skinny_float_complex alpha,
                                     skinny_float_complex beta, int64_t nrows,
                                     int64_t ncolsin, int64_t ncolsout,
                                     int64_t stridex, int64_t stridew,
                                     skinny_stream_t stream);

/* ---- device data generation (counter RNG, identical to the host's) ----- */
/* out[i] = U[-1,1) on a 2^-52 grid from (seed, stream_id, offset + i); the
 * same sequence perf.cu's initKernel role needs (perf.cu:36-42), but random */
skinny_status_t skinny_fill_uniform(double* out, int64_t n, uint64_t seed,
                                    uint64_t stream_id, int64_t offset,
                                    skinny_stream_t stream);
skinny_status_t skinny_fill_uniform_f(float* out, int64_t n, uint64_t seed,
                                      uint64_t stream_id, int64_t offset,
                                      skinny_stream_t stream);
/* integers in [lo, hi] as doubles (the exact-arithmetic test data of
 * test.cu:82-95) */
skinny_status_t skinny_fill_int(double* out, int64_t n, uint64_t seed,
                                uint64_t stream_id, int64_t offset, int lo,
                                int hi, skinny_stream_t stream);

/* ---- autotuner ----------------------------------------------------------- */
/* Time the kernel variants and grid sizes (0.5 .. 8 workgroups per CU) of
 * one shape class (op 0: TSMTTSM, 1: TSMM; dtype 0: double, 1: float,
 * 2: complex double, 3: complex float, 4: HR (TSMTTSM only); M, N <= 1024,
 * complex <= 512, complex widths for complex types; K class =
 * floor(log2 K)) through the public entry points on scratch data, time the
 * three fastest again interleaved, and remember the fastest in the handle:
 * later calls of that type and shape with blockCount == 0 use it.  Replaces
 * the hand-fitted blends (tsmm/fix_blend.cuh:13-28, varip_blend.cuh:13-26)
 * and perf.cu's per-shape grid sweep (perf.cu:186-204).  Synchronous. */
skinny_status_t skinny_autotune(skinny_handle_t h, int op, int dtype,
                                int64_t M, int64_t N, int64_t K, int reps,
                                double* best_ms);
/* The table as text lines "op dtype M N log2K variant blockCount" (TSMTTSM
 * variant bit 0: doubled unroll, bit 1: partials summed by the last arrivals
 * of the same launch instead of a second kernel; real TSMM up to 64 columns,
 * bit 0: the other kernel form, whole rows per thread or matrix cores;
 * complex TSMM: 0); returns the size needed including the terminating NUL. */
size_t skinny_tune_export(skinny_handle_t h, char* buf, size_t len);
skinny_status_t skinny_tune_import(skinny_handle_t h, const char* text);
void skinny_tune_clear(skinny_handle_t h);

/* ---- persistent batched products --------------------------------------- */
/* Replaces multi_dispatch.cuh:3-90 (one launch per compile-time (M, N)) with
 * one persistent launch over a work queue of row ranges of many independent
 * FP64 products of uneven (K, M, N), 1 <= M, N <= 64.  Ops must not read
 * each other's outputs.  TSMM ops may be in place (C == A, lda == ldc). */
typedef struct {
  int op; /* 0: TSMTTSM  C = alpha A^T B + beta C;  1: TSMM  C = alpha A B + beta C */
  int64_t M, N, K;
  const double* A;
  int64_t lda;
  double alpha;
  const double* B;
  int64_t ldb;
  double beta;
  double* C;
  int64_t ldc;
} skinny_batch_op;
typedef struct skinny_batch_s* skinny_batch_t;
/* Plan (allocates and uploads the queue).  Each op is split into items of
 * about item_bytes of row data; item_bytes <= 0 picks about 8 items per
 * workgroup of the persistent grid, between 1 and 8 MiB each.  TSMM items
 * interleave the op's 64-row chunks, so workgroups stream neighbouring rows.
 * TSMTTSM results are reproducible: each op's partials are summed in item
 * order whatever the schedule. */
skinny_status_t skinny_dbatch_create(skinny_handle_t h, int nops,
                                     const skinny_batch_op* ops,
                                     int64_t item_bytes, skinny_batch_t* batch);
/* One launch; no allocation, memset or host synchronisation (graph-safe). */
skinny_status_t skinny_dbatch_run(skinny_batch_t batch, skinny_stream_t stream);
/* Autotune the batch (replaces the per-width ladder of multi_dispatch.cuh:3-90,
 * which fixed one kernel instance per (M, N) at compile time):
 *  1. tile class per op shape: the ops of one (op kind, M, N, Gram or not)
 *     whose class was not set with skinny_dbatch_set_class are run alone
 *     with each candidate tile class (for TSMTTSM with max(M, N) <= 8 the
 *     two densest row packings and the unpacked 16 x 16 tile, for Gram
 *     products the Gram and the general tile, for TSMM with M, N <= 2 the
 *     row-per-thread and the matrix-core form) and take the fastest;
 *  2. work split: the plan is timed with items of the automatic size and of
 *     1, 2, 4, 8 and 16 MiB, on grids of 1 and 2 workgroups per CU, and the
 *     fastest kept.
 * Every timing runs `reps` + 1 times on the null stream (the ops' outputs are
 * overwritten, in-place TSMM ops updated each time).  best_ms, item_bytes
 * (0: automatic) and grid may be NULL.  Synchronous; a captured graph of an
 * earlier plan must be captured again. */
skinny_status_t skinny_dbatch_tune(skinny_batch_t batch, int reps,
                                   double* best_ms, int64_t* item_bytes,
                                   int* grid);
/* Tile class of op `op` (the kernel form the persistent kernel runs it with;
 * the per-width template instances of multi_dispatch.cuh:3-90 made this
 * choice at compile time).  TSMTTSM: 0..5 rows packed P = 16, 8, 5, 4, 3, 2
 * to the matrix-core lanes (P max(M, N) <= 16); 6..14 subtiles
 * (MT, NT) in {1, 2, 4}^2 of 16 x 16 (16 MT >= M, 16 NT >= N); 15..17 Gram
 * products (B == A, ldb == lda, M == N) on 16, 32, 64 square tiles.  TSMM:
 * 0..14 slices S in {1, 2, 4, 8, 16} x NT in {1, 2, 4} column tiles
 * (class 3 iS + iN; 4 S >= M, 16 NT >= N); 15, 16 one thread per row for
 * M <= 1, 2 and N <= 2.  cls = -1 returns to the default rule.  Re-plans the
 * batch; SKINNY_UNSUPPORTED if the class cannot compute the op. */
skinny_status_t skinny_dbatch_set_class(skinny_batch_t batch, int op, int cls);
/* The tile class of every op in the current plan (nops ints). */
skinny_status_t skinny_dbatch_classes(skinny_batch_t batch, int* cls);
skinny_status_t skinny_dbatch_destroy(skinny_batch_t batch);
int64_t skinny_dbatch_items(skinny_batch_t batch);

/* ---- SVQB orthogonalisation (svqb/svqb.cu:72-130) ---------------------- */
/* One block: X is K x M row-major (ldx >= M), M <= 64, updated in place. */
typedef struct {
  int64_t M, K;
  double* X;
  int64_t ldx;
} skinny_svqb_block;
/* An executable hipGraph owning its buffers (see skinny_dsvqb_graph). */
typedef struct skinny_graph_s* skinny_graph_t;
/* `sweeps` SVQB sweeps over `nblocks` independent blocks: per sweep the Gram
 * S = X^T X (TSMTTSM), the FP64 eigensolve of D S D on the device (one
 * workgroup per block) giving T = D U L^-1/2, and X <- X T in place (TSMM).
 * Synchronises `stream` before returning (its per-call buffers are freed). */
skinny_status_t skinny_dsvqb(skinny_handle_t h, int nblocks,
                             const skinny_svqb_block* blocks, int sweeps,
                             skinny_stream_t stream);
/* Capture the same `sweeps` into a hipGraph (no memsets, no host syncs, no
 * allocations inside); replay with skinny_graph_launch.  The graph owns a
 * private handle and buffers, so its device pointers stay valid. */
skinny_status_t skinny_dsvqb_graph(skinny_handle_t h, int nblocks,
                                   const skinny_svqb_block* blocks, int sweeps,
                                   skinny_stream_t stream,
                                   skinny_graph_t* graph);
/* The same sweeps as one cooperative launch of a persistent kernel whose
 * phases (Gram partials of all blocks; per-block sum and eigensolve; X <- X T
 * of all blocks) are separated by grid barriers (the barrier of
 * global_sync/sync.cu:38-53).  Synchronises `stream` before returning; a
 * barrier that waits longer than about 2 s fails the call instead of
 * hanging. */
skinny_status_t skinny_dsvqb_fused(skinny_handle_t h, int nblocks,
                                   const skinny_svqb_block* blocks, int sweeps,
                                   skinny_stream_t stream);
skinny_status_t skinny_graph_launch(skinny_graph_t graph,
                                    skinny_stream_t stream);
skinny_status_t skinny_graph_destroy(skinny_graph_t graph);
/* ||X^T X - I||_F with the double-double Gram (svqb.cu:184-211); synchronous,
 * result on the host. */
skinny_status_t skinny_dorth_error(skinny_handle_t h, int64_t M, int64_t K,
                                   const double* X, int64_t ldx, double* fro,
                                   skinny_stream_t stream);

/* ---- microbenchmarks: this chip's own ceilings ------------------------ */
/* HBM stream over arrays of `bytes_per_array` (allocated internally, far
 * above the Infinity Cache when >= 1 GiB); kind 0 read, 1 copy, 2 triad,
 * 3 write, 4 update (a += s b); best of `reps` and of sixteen launch shapes
 * (1, 2, 4 or 8 workgroups per CU, 1 or 4 accesses in flight per thread,
 * loads with or without the nontemporal hint), in GB/s of array traffic.
 * Kind 5 is a copy over other access patterns: nontemporal stores and/or
 * loads, and per-wave bursts of 16 KiB to 1 MiB (best of 40 variants).
 * Replaces microbenches/membw.cu:19-164 (copy/scale/update/sum/triad). */
skinny_status_t skinny_mb_stream(int kind, size_t bytes_per_array, int reps,
                                 double* gbs, skinny_stream_t stream);
/* kind 0: FP64 vector FMA, 1: FP64 matrix core (16x16x4), 2: INT8 matrix
 * core, 3: FP64 matrix core (4x4x4, four blocks), 4: FP32 matrix core
 * (16x16x4), 5: FP32 matrix core (32x32x2), 6: FP32 matrix core (4x4x1,
 * 16 blocks); TFLOP/s (TOP/s).
 * Replaces microbenches/roofline.cu:19-66. */
skinny_status_t skinny_mb_peak(int kind, int reps, double* tflops,
                               skinny_stream_t stream);
/* Shader clock in MHz while the peak loop of `kind` runs on `stream`, from
 * the shader-clock and real-time counters its first workgroup reads at its
 * start and end (-1: a one-wave spin of about 1 ms, the idle clock).
 * Synchronous.  The reference normalises by a fixed clock
 * (membw.cu:227,231); this measures the one the chip runs at. */
skinny_status_t skinny_mb_clock(int kind, double* mhz, skinny_stream_t stream);
/* Memory hierarchy by working-set size (replaces the LDS-throughput and
 * L1/L2 latency sweeps of microbenches/instruction_rates.cu:30-148):
 * kind 0: ns per dependent load of one lane chasing a chain over the 128-byte
 *         lines of a `bytes` buffer (rounded down to a power of two);
 * kind 1: GB/s of the whole GPU re-reading a `bytes` buffer (each workgroup
 *         the same slice on every pass; from 4 MiB on all workgroups read);
 * kind 2: LDS read GB/s of all CUs (bytes ignored);
 * kind 3: ns per dependent LDS read (bytes ignored).
 * 4 KiB <= bytes <= 16 GiB for kinds 0 and 1. */
skinny_status_t skinny_mb_cache(int kind, size_t bytes, int reps, double* out,
                                skinny_stream_t stream);
/* Microseconds per grid-wide barrier (the barrier of skinny_dsvqb_fused) in a
 * cooperative grid of per_cu workgroups of 256 threads per CU; replaces
 * global_sync/sync.cu:38-53.  SKINNY_UNSUPPORTED if per_cu workgroups do not
 * fit on a CU. */
skinny_status_t skinny_mb_barrier(int per_cu, int reps, double* us,
                                  skinny_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* SKINNY_H */
