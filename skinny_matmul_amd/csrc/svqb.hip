// SVQB orthogonalisation of tall-skinny blocks X (K x M, row-major), built on
// the library's TSMTTSM and TSMM, entirely on the device.
//
// Reference: svqb/svqb.cu:72-130.  One sweep is
//   S = X^T X                                   (svqb.cu:82-93)
//   d_i = 1 / sqrt(S_ii)  (0 if S_ii <= 1e-15)  (svqb.cu:99-105)
//   S <- D S D                                  (svqb.cu:106-110)
//   S = U L U^T  (symmetric eigensolve)         (svqb.cu:113-114, magma_dsyevd)
//   T = D U L^-1/2,  L clamped at 1e-15         (svqb.cu:115-119)
//   X <- X T                                    (svqb.cu:120-129)
// Differences by design: the eigensolve is a cyclic Jacobi method run by one
// workgroup per block on the device (no host round trip, so sweeps can be
// captured in a hipGraph), and the update X <- X T is done in place.
// The orthogonality metric ||X^T X - I||_F of svqb.cu:184-211 uses the
// double-double Gram (as getL2Error does with PseudoQuad).
#include <cmath>
#include <vector>

#include "internal.hpp"
#include "svqb_eig.hpp"

namespace skinny {

struct EigArgs {
  double* const* S;  // per block: M x M Gram (column-major, ld = M)
  double* const* T;  // per block: M x M output transform (column-major, ld M)
  const int* Ms;
  int max_sweeps;
};

__global__ void __launch_bounds__(256) svqb_eig_kernel(EigArgs args) {
  const int blk = blockIdx.x;
  const int M = args.Ms[blk];
  const double* Sg = args.S[blk];
  __shared__ double lds[kEigLdsDoubles];
  __shared__ int ilds[kEigLdsInts];
  for (int e = threadIdx.x; e < M * M; e += 256) {
    const int i = e % M, j = e / M;
    lds[i * kEigLD + j] = Sg[(int64_t)j * M + i];
  }
  svqb_eig_lds(M, args.max_sweeps, lds, ilds, args.T[blk]);
}

struct SvqbPlanBuf {
  std::vector<int> Ms;
  std::vector<double*> S, T;
  int* dMs = nullptr;
  double** dS = nullptr;
  double** dT = nullptr;
  double* mats = nullptr;
  size_t bytes = 0;
};

static skinny_status_t svqb_alloc(const skinny_svqb_block* blocks, int nb,
                                  SvqbPlanBuf& b) {
  size_t doubles = 0;
  for (int i = 0; i < nb; i++) doubles += 2 * (size_t)blocks[i].M * blocks[i].M;
  b.bytes = doubles * sizeof(double) + nb * (2 * sizeof(double*) + sizeof(int)) + 256;
  hipError_t e = hipMalloc(&b.mats, doubles * sizeof(double) + 64);
  if (e == hipSuccess) e = hipMalloc(&b.dS, nb * sizeof(double*));
  if (e == hipSuccess) e = hipMalloc(&b.dT, nb * sizeof(double*));
  if (e == hipSuccess) e = hipMalloc(&b.dMs, nb * sizeof(int));
  if (e != hipSuccess) return hip_fail(e, "svqb alloc");
  size_t off = 0;
  for (int i = 0; i < nb; i++) {
    const size_t mm = (size_t)blocks[i].M * blocks[i].M;
    b.Ms.push_back((int)blocks[i].M);
    b.S.push_back(b.mats + off);
    off += mm;
    b.T.push_back(b.mats + off);
    off += mm;
  }
  e = hipMemcpy(b.dS, b.S.data(), nb * sizeof(double*), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b.dT, b.T.data(), nb * sizeof(double*), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b.dMs, b.Ms.data(), nb * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "svqb upload");
  return SKINNY_OK;
}

static void svqb_free(SvqbPlanBuf& b) {
  (void)hipFree(b.mats);
  (void)hipFree(b.dS);
  (void)hipFree(b.dT);
  (void)hipFree(b.dMs);
  b = SvqbPlanBuf();
}

static skinny_status_t check_blocks(const skinny_svqb_block* blocks, int nb) {
  if (nb < 1 || !blocks) return fail(SKINNY_INVALID_ARG, "svqb: no blocks");
  for (int i = 0; i < nb; i++) {
    const skinny_svqb_block& b = blocks[i];
    if (b.M < 1 || b.K < 0 || !b.X || b.ldx < b.M)
      return fail(SKINNY_INVALID_ARG, "svqb: bad block");
    if (b.M > kSvqbMaxM)
      return fail(SKINNY_UNSUPPORTED, "svqb: M <= 64 only");
  }
  return SKINNY_OK;
}

// Enqueue `sweeps` SVQB sweeps on `st` (no host synchronisation, no
// allocation once the handle's workspace holds the largest Gram's partials).
static skinny_status_t svqb_enqueue(skinny_handle_t h,
                                    const skinny_svqb_block* blocks, int nb,
                                    int sweeps, const SvqbPlanBuf& b,
                                    hipStream_t st) {
  for (int sw = 0; sw < sweeps; sw++) {
    for (int i = 0; i < nb; i++) {
      const skinny_svqb_block& x = blocks[i];
      TsmttsmPlan p = plan_tsmttsm(h, 0, x.M, x.M, x.K, 8, 8);
      skinny_status_t s = launch_tsmttsm_real<double>(
          h, p, x.M, x.M, x.K, x.X, x.ldx, 1.0, x.X, x.ldx, 0.0, b.S[i], x.M, st);
      if (s != SKINNY_OK) return s;
    }
    EigArgs ea;
    ea.S = b.dS;
    ea.T = b.dT;
    ea.Ms = b.dMs;
    ea.max_sweeps = 40;
    hipLaunchKernelGGL(svqb_eig_kernel, dim3(nb), dim3(256), 0, st, ea);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "svqb eig launch");
    for (int i = 0; i < nb; i++) {
      const skinny_svqb_block& x = blocks[i];
      skinny_status_t s = launch_tsmm_real<double>(
          h, 0, x.M, x.M, x.K, x.X, x.ldx, 1.0, b.T[i], x.M, 0.0, x.X, x.ldx, st);
      if (s != SKINNY_OK) return s;
    }
  }
  return SKINNY_OK;
}

static size_t max_gram_scratch(skinny_handle_t h, const skinny_svqb_block* blocks,
                               int nb) {
  size_t m = 0;
  for (int i = 0; i < nb; i++) {
    TsmttsmPlan p = plan_tsmttsm(h, 0, blocks[i].M, blocks[i].M, blocks[i].K, 8, 8);
    if (p.scratch_bytes > m) m = p.scratch_bytes;
  }
  return m;
}

}  // namespace skinny

struct skinny_graph_s {
  skinny_handle_t h = nullptr;  // private handle: its workspace never moves
  skinny::SvqbPlanBuf buf;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

using namespace skinny;

extern "C" {

skinny_status_t skinny_dsvqb(skinny_handle_t h, int nblocks,
                             const skinny_svqb_block* blocks, int sweeps,
                             skinny_stream_t stream) {
  skinny_status_t s = check_blocks(blocks, nblocks);
  if (s != SKINNY_OK) return s;
  if (sweeps < 0) return fail(SKINNY_INVALID_ARG, "svqb: sweeps < 0");
  h = resolve(h);
  if (!h) return no_handle("svqb");
  s = skinny_reserve_stream(h, max_gram_scratch(h, blocks, nblocks), stream);
  if (s != SKINNY_OK) return s;
  SvqbPlanBuf b;
  s = svqb_alloc(blocks, nblocks, b);
  if (s == SKINNY_OK)
    s = svqb_enqueue(h, blocks, nblocks, sweeps, b, (hipStream_t)stream);
  // the per-call buffers must outlive the enqueued kernels
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  svqb_free(b);
  if (s != SKINNY_OK) return s;
  if (e != hipSuccess) return hip_fail(e, "svqb sync");
  return SKINNY_OK;
}

skinny_status_t skinny_dsvqb_graph(skinny_handle_t h_unused, int nblocks,
                                   const skinny_svqb_block* blocks, int sweeps,
                                   skinny_stream_t stream,
                                   skinny_graph_t* graph) {
  (void)h_unused;
  if (!graph) return fail(SKINNY_INVALID_ARG, "svqb graph: null out");
  skinny_status_t s = check_blocks(blocks, nblocks);
  if (s != SKINNY_OK) return s;
  if (sweeps < 1) return fail(SKINNY_INVALID_ARG, "svqb graph: sweeps < 1");
  skinny_graph_t g = new skinny_graph_s();
  s = skinny_create(&g->h);
  if (s == SKINNY_OK) s = svqb_alloc(blocks, nblocks, g->buf);
  if (s != SKINNY_OK) {
    skinny_graph_destroy(g);
    return s;
  }
  hipStream_t st = (hipStream_t)stream;
  hipStream_t cap = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
  // the capture stream's workspace is sized now: nothing allocates inside
  if (e == hipSuccess) {
    s = skinny_reserve_stream(g->h, max_gram_scratch(g->h, blocks, nblocks), cap);
    if (s != SKINNY_OK) {
      (void)hipStreamDestroy(cap);
      skinny_graph_destroy(g);
      return s;
    }
  }
  if (e == hipSuccess) e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    if (cap) (void)hipStreamDestroy(cap);
    skinny_graph_destroy(g);
    return hip_fail(e, "svqb graph: begin capture");
  }
  s = svqb_enqueue(g->h, blocks, nblocks, sweeps, g->buf, cap);
  e = hipStreamEndCapture(cap, &g->graph);
  (void)hipStreamDestroy(cap);
  if (s != SKINNY_OK || e != hipSuccess) {
    skinny_graph_destroy(g);
    return s != SKINNY_OK ? s : hip_fail(e, "svqb graph: end capture");
  }
  e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    skinny_graph_destroy(g);
    return hip_fail(e, "svqb graph: instantiate");
  }
  (void)st;
  *graph = g;
  return SKINNY_OK;
}

skinny_status_t skinny_graph_launch(skinny_graph_t g, skinny_stream_t stream) {
  if (!g || !g->exec) return fail(SKINNY_INVALID_ARG, "graph launch: no graph");
  hipError_t e = hipGraphLaunch(g->exec, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  return SKINNY_OK;
}

skinny_status_t skinny_graph_destroy(skinny_graph_t g) {
  if (!g) return SKINNY_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  (void)hipDeviceSynchronize();
  svqb_free(g->buf);
  if (g->h) skinny_destroy(g->h);
  delete g;
  return SKINNY_OK;
}

skinny_status_t skinny_dorth_error(skinny_handle_t h, int64_t M, int64_t K,
                                   const double* X, int64_t ldx, double* fro,
                                   skinny_stream_t stream) {
  if (!fro || M < 1 || K < 0 || !X || ldx < M)
    return fail(SKINNY_INVALID_ARG, "orth_error: bad arguments");
  h = resolve(h);
  if (!h) return no_handle("orth_error");
  double* G = nullptr;
  hipError_t e = hipMalloc(&G, sizeof(double) * M * M);
  if (e != hipSuccess) return hip_fail(e, "orth_error alloc");
  skinny_status_t s = launch_htsmttsm(h, 0, M, M, K, X, ldx, 1.0, X, ldx, 0.0,
                                      G, false, M, (hipStream_t)stream);
  std::vector<double> hG((size_t)M * M);
  if (s == SKINNY_OK) {
    e = hipMemcpyAsync(hG.data(), G, sizeof(double) * M * M,
                       hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) s = hip_fail(e, "orth_error copy");
  }
  (void)hipFree(G);
  if (s != SKINNY_OK) return s;
  // svqb.cu:201-210
  double l2 = 0.0;
  for (int64_t n = 0; n < M; n++)
    for (int64_t m = 0; m < M; m++) {
      const double err = (m == n ? 1.0 : 0.0) - hG[n * M + m];
      l2 += err * err;
    }
  *fro = std::sqrt(l2);
  return SKINNY_OK;
}

}  // extern "C"
