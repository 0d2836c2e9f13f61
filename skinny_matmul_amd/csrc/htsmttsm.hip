// HR TSMTTSM: double inputs, PseudoQuad (double-double) accumulation.
//
// Reference counterparts: the iT = PseudoQuad instantiations of GENV3 /
// SPECSMALL / GENV7 (types.hpp:49-57, genv3.cuh:46-79, specsmall.cuh) and
// the stage-2 epilogue convert(scale2(renormalize(sum), alpha))
// (genv3.cuh:19-26).  Here:
//
//  * Each thread owns a TM x TN block of C entries for a strided set of rows
//    and accumulates every product exactly: TwoProd (FMA) followed by the
//    accurate dd + dd (the reference's slow_axpy2, PseudoQuad.cuh:115-118).
//    The reference's cheap axpy2 (PseudoQuad.cuh:96-108) is not used: it is
//    only double-accurate when the running sum cancels.
//  * Row slots of a workgroup are combined in LDS in a fixed order, then the
//    workgroup partials with the same two-level last-arrival reduction as the
//    FP64 kernel, so the result is reproducible for a fixed blockCount.  The
//    combine stages add dd partials with DdAcc (dd plus the rounding
//    remainders of each addition) and store partials as (s, t, u) triples:
//    partials of very different magnitude (2^124 next to 2^62 in the
//    numeric_tests.cu:23-53 data) would otherwise round the small terms
//    away, and the result would depend on blockCount.
//  * The final step is the reference's: renormalize, scale by alpha (dd * d),
//    then either convert to double and add beta * C, or keep the dd pair.
#include "internal.hpp"
#include "skinny/pseudoquad.hpp"

namespace skinny {

struct HrArgs {
  const double* A;
  const double* B;
  double* C;      // double output, or interleaved (s, t) pairs if dd_out
  int64_t K, lda, ldb, ldc;
  int M, N;
  double alpha, beta;
  int dd_out;
  double* part1;  // [tile][G][PS] (s, t, u) triples
  double* part2;  // [tile][ngroups][PS] triples
  unsigned* cnt;
  int G, GS, ngroups, tilesM, tilesN, tileM, tileN;
};

__device__ __forceinline__ DdAcc ld_tri(const double* p) {
  return DdAcc{PseudoQuad{p[0], p[1]}, p[2]};
}
__device__ __forceinline__ void st_tri(double* p, DdAcc v) {
  p[0] = v.v.s;
  p[1] = v.v.t;
  p[2] = v.u;
}

template <int TM, int TN>
__global__ void __launch_bounds__(256) htsmttsm_kernel(HrArgs a) {
  const int tid = threadIdx.x;
  const int ntiles = a.tilesM * a.tilesN;
  const int tile = blockIdx.x % ntiles;
  const int kp = blockIdx.x / ntiles;
  const int m0 = (tile % a.tilesM) * a.tileM;
  const int n0 = (tile / a.tilesM) * a.tileN;
  const int Mloc = min(a.M - m0, a.tileM);
  const int Nloc = min(a.N - n0, a.tileN);
  const int QM = (Mloc + TM - 1) / TM, QN = (Nloc + TN - 1) / TN;
  const int Q = QM * QN;
  const int slots = 256 / Q;  // >= 1 by construction of the plan
  const int q = tid % Q, slot = tid / Q;
  const bool active = slot < slots;
  const int bm = q % QM, bn = q / QM;

  PseudoQuad acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = PseudoQuad{0.0, 0.0};

  if (active) {
    int mi[TM], nj[TN];
    bool vm[TM], vn[TN];
#pragma unroll
    for (int i = 0; i < TM; i++) {
      mi[i] = m0 + bm * TM + i;
      vm[i] = bm * TM + i < Mloc;
    }
#pragma unroll
    for (int j = 0; j < TN; j++) {
      nj[j] = n0 + bn * TN + j;
      vn[j] = bn * TN + j < Nloc;
    }
    const int64_t stride = (int64_t)a.G * slots;
    for (int64_t r = (int64_t)kp * slots + slot; r < a.K; r += stride) {
      double av[TM], bv[TN];
#pragma unroll
      for (int i = 0; i < TM; i++) av[i] = vm[i] ? a.A[r * a.lda + mi[i]] : 0.0;
#pragma unroll
      for (int j = 0; j < TN; j++) bv[j] = vn[j] ? a.B[r * a.ldb + nj[j]] : 0.0;
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = dd_fma(acc[i][j], av[i], bv[j]);
    }
  }

  // ---- combine the row slots in LDS, fixed order --------------------------
  constexpr int CELLS = TM * TN;
  __shared__ PseudoQuad red[256 * CELLS];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) red[(i * TN + j) * 256 + tid] = acc[i][j];
  __syncthreads();
  const int PS = a.tileM * a.tileN;
  double* mine = a.part1 + ((int64_t)tile * a.G + kp) * PS * 3;
  for (int e = tid; e < PS; e += 256) {
    const int ml = e % a.tileM, nl = e / a.tileM;
    DdAcc s{{0.0, 0.0}, 0.0};
    if (ml < Mloc && nl < Nloc) {
      const int qq = (ml / TM) + (nl / TN) * QM;
      const int cell = (ml % TM) * TN + (nl % TN);
      for (int sl = 0; sl < slots; sl++) s = dda_add(s, red[cell * 256 + sl * Q + qq]);
    }
    st_tri(mine + (int64_t)e * 3, s);
  }

  // ---- two-level last-arrival reduction ------------------------------------
  __shared__ int s_last;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  const int grp = kp / a.GS;
  const int gsz = min(a.GS, a.G - grp * a.GS);
  unsigned* cnt1 = a.cnt + tile * a.ngroups + grp;
  unsigned* cnt2 = a.cnt + ntiles * a.ngroups + tile;
  if (tid == 0) {
    unsigned t = __hip_atomic_fetch_add(cnt1, 1u, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
    s_last = (t == (unsigned)(gsz - 1));
  }
  __syncthreads();
  if (!s_last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");

  const double* src = a.part1 + ((int64_t)tile * a.G + (int64_t)grp * a.GS) * PS * 3;
  int count = gsz;
  if (a.ngroups > 1) {
    double* dst = a.part2 + ((int64_t)tile * a.ngroups + grp) * PS * 3;
    for (int e = tid; e < PS; e += 256) {
      DdAcc s{{0.0, 0.0}, 0.0};
      for (int g = 0; g < gsz; g++) s = dda_add(s, ld_tri(src + ((int64_t)g * PS + e) * 3));
      st_tri(dst + (int64_t)e * 3, s);
    }
    if (tid == 0) *cnt1 = 0u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (tid == 0) {
      unsigned t = __hip_atomic_fetch_add(cnt2, 1u, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
      s_last = (t == (unsigned)(a.ngroups - 1));
    }
    __syncthreads();
    if (!s_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    src = a.part2 + (int64_t)tile * a.ngroups * PS * 3;
    count = a.ngroups;
  }
  for (int e = tid; e < PS; e += 256) {
    const int ml = e % a.tileM, nl = e / a.tileM;
    if (ml >= Mloc || nl >= Nloc) continue;
    DdAcc s{{0.0, 0.0}, 0.0};
    for (int g = 0; g < count; g++) s = dda_add(s, ld_tri(src + ((int64_t)g * PS + e) * 3));
    // genv3.cuh:25-26: beta * C + convert(scale2(renormalize(sum), alpha))
    const PseudoQuad r = dd_scale(dd_renorm(dda_value(s)), a.alpha);
    const int64_t ci = (int64_t)(n0 + nl) * a.ldc + (m0 + ml);
    if (a.dd_out) {
      PseudoQuad o = r;
      if (a.beta != 0.0)
        o = dd_add(dd_scale(PseudoQuad{a.C[2 * ci], a.C[2 * ci + 1]}, a.beta), r);
      a.C[2 * ci] = o.s;
      a.C[2 * ci + 1] = o.t;
    } else {
#pragma clang fp contract(off)
      const double v = dd_to_d(r);
      a.C[ci] = (a.beta == 0.0) ? v : a.beta * a.C[ci] + v;
    }
  }
  if (tid == 0) {
    if (a.ngroups > 1)
      *cnt2 = 0u;
    else
      *cnt1 = 0u;
  }
}

static int isqrt_up(int x) {
  int r = 1;
  while (r * r < x) r++;
  return r;
}

skinny_status_t launch_htsmttsm(skinny_handle_t h, size_t blockCount,
                                int64_t M, int64_t N, int64_t K,
                                const double* A, int64_t lda, double alpha,
                                const double* B, int64_t ldb, double beta,
                                double* C, bool dd_out, int64_t ldc,
                                hipStream_t stream) {
  HrArgs a;
  int TMc;
  if (M * N <= 64) {
    TMc = 1;
    a.tileM = (int)M;
    a.tileN = (int)N;
  } else if (M <= 32 && N <= 32) {
    TMc = 2;
    a.tileM = (int)M;
    a.tileN = (int)N;
  } else {
    TMc = 4;
    a.tileM = (int)std::min<int64_t>(M, 64);
    a.tileN = (int)std::min<int64_t>(N, 64);
  }
  a.tilesM = (int)((M + a.tileM - 1) / a.tileM);
  a.tilesN = (int)((N + a.tileN - 1) / a.tileN);
  const int ntiles = a.tilesM * a.tilesN;
  int64_t G = (int64_t)blockCount;
  TuneEntry te;
  if (G <= 0 && tune_lookup(h, 0, TT_HR, M, N, K, &te)) G = te.G;
  if (G <= 0) {
    int64_t target = (4LL * h->cu_count) / ntiles;
    if (target < 1) target = 1;
    int64_t byrows = K / 1024;
    if (byrows < 1) byrows = 1;
    G = std::min(target, byrows);
  }
  if (G > 65536) G = 65536;
  a.G = (int)G;
  if (a.G <= 16) {
    a.GS = a.G;
    a.ngroups = 1;
  } else {
    a.GS = isqrt_up(a.G);
    a.ngroups = (a.G + a.GS - 1) / a.GS;
  }
  if ((size_t)ntiles * (a.ngroups + 1) > (size_t)kMaxCounters)
    return fail(SKINNY_UNSUPPORTED, "htsmttsm: too many partial groups");
  const size_t PS = (size_t)a.tileM * a.tileN;
  size_t bytes = (size_t)ntiles * a.G * PS * 3 * sizeof(double);
  if (a.ngroups > 1) bytes += (size_t)ntiles * a.ngroups * PS * 3 * sizeof(double);
  Workspace w;
  skinny_status_t st = acquire_workspace(h, stream, bytes, &w);
  if (st != SKINNY_OK) return st;
  a.A = A;
  a.B = B;
  a.C = C;
  a.dd_out = dd_out ? 1 : 0;
  a.K = K;
  a.lda = lda;
  a.ldb = ldb;
  a.ldc = ldc;
  a.M = (int)M;
  a.N = (int)N;
  a.alpha = alpha;
  a.beta = beta;
  a.part1 = static_cast<double*>(ws_scratch(w));
  a.part2 = a.part1 + (size_t)ntiles * a.G * PS * 3;
  a.cnt = ws_counters(w);
  const int blocks = a.G * ntiles;
  switch (TMc) {
    case 1: hipLaunchKernelGGL((htsmttsm_kernel<1, 1>), dim3(blocks), dim3(256), 0, stream, a); break;
    case 2: hipLaunchKernelGGL((htsmttsm_kernel<2, 2>), dim3(blocks), dim3(256), 0, stream, a); break;
    default: hipLaunchKernelGGL((htsmttsm_kernel<4, 4>), dim3(blocks), dim3(256), 0, stream, a); break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "htsmttsm launch");
  return SKINNY_OK;
}

}  // namespace skinny
