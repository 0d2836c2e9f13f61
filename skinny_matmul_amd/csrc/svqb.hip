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

namespace skinny {

constexpr int kSvqbMaxM = 64;

struct EigArgs {
  double* const* S;  // per block: M x M Gram (column-major, ld = M); destroyed
  double* const* T;  // per block: M x M output transform (column-major, ld M)
  const int* Ms;
  int max_sweeps;
};

// Round-robin (tournament) ordering: in step r, slots 0..P-1 hold the pairs
// (top[i], bot[i]); index 0 stays, the others rotate.
__device__ __forceinline__ void rr_pair(int i, int r, int n, int* p, int* q) {
  // positions 0..n-1 on a circle of n - 1 movable players plus player 0
  auto player = [&](int pos) -> int {
    if (pos == 0) return 0;
    return 1 + ((pos - 1 + r) % (n - 1));
  };
  int a = player(i), b = player(n - 1 - i);
  *p = a < b ? a : b;
  *q = a < b ? b : a;
}

__global__ void __launch_bounds__(256) svqb_eig_kernel(EigArgs args) {
  const int blk = blockIdx.x;
  const int M = args.Ms[blk];
  const int n = (M + 1) & ~1;  // even size with a dummy index when M is odd
  const int npairs = n / 2;
  double* Sg = args.S[blk];
  double* Tg = args.T[blk];
  const int tid = threadIdx.x;
  constexpr int LD = kSvqbMaxM + 1;
  __shared__ double S[kSvqbMaxM * LD];
  __shared__ double V[kSvqbMaxM * LD];
  __shared__ double dinv[kSvqbMaxM];
  __shared__ double rc[kSvqbMaxM / 2], rs[kSvqbMaxM / 2];
  __shared__ int rp[kSvqbMaxM / 2], rq[kSvqbMaxM / 2];
  __shared__ double red[256];
  __shared__ int done;

  for (int e = tid; e < M * M; e += 256) {
    const int i = e % M, j = e / M;
    S[i * LD + j] = Sg[(int64_t)j * M + i];
    V[i * LD + j] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  if (tid < M) {
    const double sii = S[tid * LD + tid];
    dinv[tid] = (fabs(sii) > 1.0e-15) ? 1.0 / sqrt(sii) : 0.0;
  }
  __syncthreads();
  for (int e = tid; e < M * M; e += 256) {
    const int i = e % M, j = e / M;
    S[i * LD + j] *= dinv[i] * dinv[j];
  }
  __syncthreads();

  for (int sweep = 0; sweep < args.max_sweeps; sweep++) {
    // convergence: off-diagonal vs diagonal mass
    double off = 0.0, dia = 0.0;
    for (int e = tid; e < M * M; e += 256) {
      const int i = e % M, j = e / M;
      const double v = S[i * LD + j];
      if (i == j)
        dia += v * v;
      else
        off += v * v;
    }
    red[tid] = off;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (tid < w) red[tid] += red[tid + w];
      __syncthreads();
    }
    const double offsum = red[0];
    __syncthreads();
    red[tid] = dia;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (tid < w) red[tid] += red[tid + w];
      __syncthreads();
    }
    if (tid == 0) done = (offsum <= 1e-30 * red[0]) ? 1 : 0;
    __syncthreads();
    if (done) break;

    for (int r = 0; r < n - 1; r++) {
      // rotation parameters for the disjoint pairs of this step
      if (tid < npairs) {
        int p, q;
        rr_pair(tid, r, n, &p, &q);
        double c = 1.0, s = 0.0;
        if (q < M) {
          const double apq = S[p * LD + q];
          if (apq != 0.0) {
            const double app = S[p * LD + p], aqq = S[q * LD + q];
            const double tau = (aqq - app) / (2.0 * apq);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) /
                             (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
          }
        }
        rp[tid] = p;
        rq[tid] = q;
        rc[tid] = c;
        rs[tid] = s;
      }
      __syncthreads();
      // rows p, q of S  <-  J^T S
      for (int e = tid; e < npairs * M; e += 256) {
        const int k = e % npairs, j = e / npairs;
        const int p = rp[k], q = rq[k];
        if (q >= M) continue;
        const double c = rc[k], s = rs[k];
        const double a = S[p * LD + j], b = S[q * LD + j];
        S[p * LD + j] = c * a - s * b;
        S[q * LD + j] = s * a + c * b;
      }
      __syncthreads();
      // columns p, q of S and of V  <-  (.) J
      for (int e = tid; e < npairs * M; e += 256) {
        const int k = e % npairs, i = e / npairs;
        const int p = rp[k], q = rq[k];
        if (q >= M) continue;
        const double c = rc[k], s = rs[k];
        const double a = S[i * LD + p], b = S[i * LD + q];
        S[i * LD + p] = c * a - s * b;
        S[i * LD + q] = s * a + c * b;
        const double va = V[i * LD + p], vb = V[i * LD + q];
        V[i * LD + p] = c * va - s * vb;
        V[i * LD + q] = s * va + c * vb;
      }
      __syncthreads();
      if (tid < npairs) {
        const int p = rp[tid], q = rq[tid];
        if (q < M) {
          S[p * LD + q] = 0.0;
          S[q * LD + p] = 0.0;
        }
      }
      __syncthreads();
    }
  }
  // T(m, n) = d_m U(m, n) / sqrt(max(1e-15, L_n))   (svqb.cu:115-119)
  for (int e = tid; e < M * M; e += 256) {
    const int m = e % M, nn = e / M;
    const double lam = S[nn * LD + nn];
    Tg[(int64_t)nn * M + m] = V[m * LD + nn] * dinv[m] * (1.0 / sqrt(fmax(1.0e-15, lam)));
  }
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
