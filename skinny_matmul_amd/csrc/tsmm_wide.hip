// TSMM for wide blocks: real-view widths above 128 and up to 1024, out of
// place or in place (C == A), real and complex.
//
// Reference counterparts: the runtime-width kernels tsmm/var1.cuh (any M, N,
// out of place), tsmm/varip1..3.cuh and var_ip_ghost.cuh (in place, N up to
// the block size), which take one thread per output entry and re-read the
// row of A from memory for every entry.  Here:
//
//  * A workgroup owns blocks of 16 R16 consecutive rows (grid-stride).  It
//    first copies the whole row band (16 R16 x M values of A) into LDS, then
//    computes the band's C rows on the matrix cores, then stores them.  Every
//    value of A that a band needs is in LDS before the first store, and no
//    other workgroup touches the band, so C may alias A.
//  * One v_mfma_*_16x16x4 yields a 16-row x 16-column tile of C from four
//    columns of A.  Lane (i, g) = (lane % 16, lane / 16) owns the slice
//    m = g S .. g S + S - 1 of the contraction (S = ceil(M / 4) rounded up to
//    even) and reads it from LDS two values at a time (ds_read_b128 for
//    double); the LDS row stride is 2 mod 32 doubles, so the 16 rows of a
//    lane group fall in distinct banks.
//  * B is packed once per call by a small kernel into the exact operand order
//    of the main loop (column group, step pair, subtile, lane), zero-padded
//    past M and N, so every B operand load of a wave is one contiguous
//    1 KiB access.  For complex products the packed operand is the real
//    2M x 2N matrix [[Re b, Im b], [-Im b, Re b]] of each B(m, n) (the
//    interleaved real view), and the complex alpha and beta are applied in
//    the store after a lane swap of the (re, im) pair.
//  * A wave takes items (row sub-blocks x groups of NTW 16-column subtiles):
//    RW row sub-blocks share each B operand load, NTW subtiles share each A
//    operand read, and RW x NTW independent accumulator chains hide the
//    matrix-core latency.
//
// Wide products are matrix-core bound (2 N flop per 8-byte element of A for
// real data), so the design aims at keeping the FP64 matrix cores busy; the
// HBM traffic is one read of A and one write of C.
#include <algorithm>

#include "internal.hpp"
#include "mfma.hpp"

namespace skinny {

template <typename T>
struct TwArgs {
  const T* A;   // real view, row-major, lda in reals
  const T* Bp;  // packed operand
  T* C;         // real view, row-major, ldc in reals
  int64_t K, lda, ldc;
  int M, N;     // real-view widths
  int S;        // contraction slice per lane group (even, 4 S >= M)
  int LS;       // LDS row stride (elements)
  int R16;      // 16-row sub-blocks per band
  int groups;   // column groups of NTW subtiles
  int nrs;      // row sets per band (R16 / RW)
  T alpha, beta, alpha_i, beta_i;
};

template <typename T>
__device__ __forceinline__ T tw_bval(const T* B, int64_t ldb, int m, int n,
                                     bool cplx) {
  if (!cplx) return B[(int64_t)n * ldb + m];
  const T* b = B + 2 * ((int64_t)(n >> 1) * ldb + (m >> 1));
  const int pa = m & 1, pb = n & 1;
  const T v = (pa ^ pb) ? b[1] : b[0];
  return (pa && !pb) ? -v : v;
}

// Packed operand, in pairs of values: pair q = ((grp S/2 + s2) NTW + t) 64 +
// lane holds Bh(g S + 2 s2 + h, 16 (grp NTW + t) + i) for h = 0, 1, with
// (i, g) = (lane % 16, lane / 16) and zero outside M x N.
template <typename T>
__global__ void __launch_bounds__(256)
    tsmm_wide_pack_kernel(const T* B, int64_t ldb, int M, int N, int S,
                          int ntw, int groups, int cplx, T* Bp) {
  const int S2 = S / 2;
  const int64_t total = (int64_t)groups * S2 * ntw * 64;
  for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * 256) {
    const int lane = (int)(q & 63);
    int64_t r = q >> 6;
    const int t = (int)(r % ntw);
    r /= ntw;
    const int s2 = (int)(r % S2);
    const int grp = (int)(r / S2);
    const int li = lane & 15, lg = lane >> 4;
    const int n = 16 * (grp * ntw + t) + li;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int m = lg * S + 2 * s2 + h;
      Bp[2 * q + h] = (m < M && n < N) ? tw_bval(B, ldb, m, n, cplx != 0) : T(0);
    }
  }
}

template <typename T>
__device__ __forceinline__ T tw_axpby(T s, T c, T alpha, T beta, bool beta0) {
#pragma clang fp contract(off)
  if (beta0) return alpha * s;
  return beta * c + alpha * s;
}

template <typename T, int RW, int NTW, bool BETA0, bool CPLX>
__global__ void __launch_bounds__(256) tsmm_wide_kernel(TwArgs<T> a) {
  using V4 = typename Mfma<T>::v4;
  typedef T T2 __attribute__((ext_vector_type(2)));
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  T* As = reinterpret_cast<T*>(smem_raw);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int M = a.M, N = a.N, S = a.S, LS = a.LS, S2 = a.S / 2;
  const int64_t K = a.K;
  const int ROWS = 16 * a.R16;
  const int64_t nblk = (K + ROWS - 1) / ROWS;
  const int items = a.nrs * a.groups;
  const T2* Bp2 = reinterpret_cast<const T2*>(a.Bp);

  for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int64_t row0 = blk * ROWS;
    // ---- the band's rows of A into LDS: A(row0 + r, m) at As[r LS + m] ----
    // (slice g of a row starts at m = g S); columns M .. 4 S - 1 and rows
    // past K are zero.  Flat element index e = r W4 + m over the band, with
    // SU loads in flight per thread before their LDS writes.
    {
      constexpr int SU = 8;
      const int W4 = 4 * S;
      const int total = ROWS * W4;
      for (int e0 = tid; e0 < total; e0 += 256 * SU) {
        T v[SU];
        int dst[SU];
#pragma unroll
        for (int j = 0; j < SU; j++) {
          const int e = e0 + j * 256;
          const int r = e / W4, m = e - r * W4;
          const int64_t row = row0 + r;
          dst[j] = e < total ? r * LS + m : -1;
          v[j] = (e < total && row < K && m < M) ? ld_stream(a.A + row * a.lda + m) : T(0);
        }
#pragma unroll
        for (int j = 0; j < SU; j++)
          if (dst[j] >= 0) As[dst[j]] = v[j];
      }
    }
    __syncthreads();

    // ---- items: (row set, column group) ----------------------------------
    for (int it = wave; it < items; it += 4) {
      const int rs = it % a.nrs, grp = it / a.nrs;
      V4 acc[RW][NTW];
#pragma unroll
      for (int r = 0; r < RW; r++)
#pragma unroll
        for (int t = 0; t < NTW; t++) acc[r][t] = V4{0, 0, 0, 0};
      const T2* bq = Bp2 + (int64_t)grp * S2 * NTW * 64 + lane;
      const T* arow = As + (rs * RW * 16 + li) * LS + lg * S;
      // operands of step s2 + 1 are loaded while step s2 runs
      T2 bv[NTW], av[RW];
#pragma unroll
      for (int t = 0; t < NTW; t++) bv[t] = bq[(int64_t)t * 64];
#pragma unroll
      for (int r = 0; r < RW; r++)
        av[r] = *reinterpret_cast<const T2*>(arow + r * 16 * LS);
      for (int s2 = 0; s2 < S2; s2++) {
        T2 bn[NTW], an[RW];
        const int sn = s2 + 1 < S2 ? s2 + 1 : s2;
#pragma unroll
        for (int t = 0; t < NTW; t++) bn[t] = bq[(int64_t)(sn * NTW + t) * 64];
#pragma unroll
        for (int r = 0; r < RW; r++)
          an[r] = *reinterpret_cast<const T2*>(arow + r * 16 * LS + 2 * sn);
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
          for (int r = 0; r < RW; r++)
#pragma unroll
            for (int t = 0; t < NTW; t++)
              acc[r][t] = Mfma<T>::mac(av[r][h], bv[t][h], acc[r][t]);
#pragma unroll
        for (int t = 0; t < NTW; t++) bv[t] = bn[t];
#pragma unroll
        for (int r = 0; r < RW; r++) av[r] = an[r];
      }
      // ---- store the C tiles ---------------------------------------------
#pragma unroll
      for (int r = 0; r < RW; r++)
#pragma unroll
        for (int t = 0; t < NTW; t++) {
          const int col = 16 * (grp * NTW + t) + li;
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int64_t row = row0 + (rs * RW + r) * 16 + Mfma<T>::drow(lg, v);
            const T own = acc[r][t][v];
            if (CPLX) {
              // lanes li and li ^ 1 hold the two parts of one complex entry
              const T partner = __shfl_xor(own, 1);
              if (row < K && col < N) {
#pragma clang fp contract(off)
                const bool im_lane = (li & 1) != 0;
                const T re = im_lane ? partner : own;
                const T im = im_lane ? own : partner;
                T out = im_lane ? (a.alpha * im + a.alpha_i * re)
                                : (a.alpha * re - a.alpha_i * im);
                T* c = a.C + row * a.ldc + (col & ~1);
                if (!BETA0) {
                  const T cr = c[0], ci = c[1];
                  const T bc = im_lane ? (a.beta * ci + a.beta_i * cr)
                                       : (a.beta * cr - a.beta_i * ci);
                  out = bc + out;
                }
                c[im_lane ? 1 : 0] = out;
              }
            } else if (row < K && col < N) {
              T* c = a.C + row * a.ldc + col;
              const T old = BETA0 ? T(0) : *c;
              *c = tw_axpby(own, old, a.alpha, a.beta, BETA0);
            }
          }
        }
    }
    __syncthreads();  // the band's LDS is reused by the next band
  }
}

template <typename T, int RW, int NTW, bool CPLX>
static hipError_t launch_wide_rn(const TwArgs<T>& a, int blocks, size_t lds,
                                 bool beta0, hipStream_t st) {
  auto k = beta0 ? tsmm_wide_kernel<T, RW, NTW, true, CPLX>
                 : tsmm_wide_kernel<T, RW, NTW, false, CPLX>;
  hipError_t e = hipFuncSetAttribute((const void*)k,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(blocks), dim3(256), lds, st, a);
  return hipGetLastError();
}

template <typename T, int RW, bool CPLX>
static hipError_t launch_wide_r(const TwArgs<T>& a, int ntw, int blocks,
                                size_t lds, bool beta0, hipStream_t st) {
  switch (ntw) {
    case 1: return launch_wide_rn<T, RW, 1, CPLX>(a, blocks, lds, beta0, st);
    case 2: return launch_wide_rn<T, RW, 2, CPLX>(a, blocks, lds, beta0, st);
    default: return launch_wide_rn<T, RW, 4, CPLX>(a, blocks, lds, beta0, st);
  }
}

template <typename T, bool CPLX>
static hipError_t launch_wide(const TwArgs<T>& a, int rw, int ntw, int blocks,
                              size_t lds, bool beta0, hipStream_t st) {
  switch (rw) {
    case 1: return launch_wide_r<T, 1, CPLX>(a, ntw, blocks, lds, beta0, st);
    case 2: return launch_wide_r<T, 2, CPLX>(a, ntw, blocks, lds, beta0, st);
    default: return launch_wide_r<T, 4, CPLX>(a, ntw, blocks, lds, beta0, st);
  }
}

// Largest LDS a workgroup may take on gfx950.
constexpr size_t kWideLdsMax = 160 * 1024;

template <typename T>
skinny_status_t launch_tsmm_wide(skinny_handle_t h, size_t blockCount,
                                 int64_t M, int64_t N, int64_t K, const T* A,
                                 int64_t lda, T alpha, T alpha_i, const T* B,
                                 int64_t ldb, T beta, T beta_i, T* C,
                                 int64_t ldc, bool cplx, hipStream_t stream) {
  if (M > kTsmmWideMax || N > kTsmmWideMax)
    return fail(SKINNY_UNSUPPORTED, "tsmm: real-view widths above 1024");
  TwArgs<T> a;
  a.A = A;
  a.C = C;
  a.K = K;
  a.lda = lda;
  a.ldc = ldc;
  a.M = (int)M;
  a.N = (int)N;
  a.alpha = alpha;
  a.beta = beta;
  a.alpha_i = alpha_i;
  a.beta_i = beta_i;
  int S = (int)((M + 3) / 4);
  if (S < 1) S = 1;
  S += S & 1;
  a.S = S;
  a.LS = (4 * S + 31) / 32 * 32 + 2;
  const size_t row_bytes = (size_t)a.LS * sizeof(T);
  // rows per band: up to 64, as LDS allows
  int R16 = 4;
  while (R16 > 1 && (size_t)16 * R16 * row_bytes > kWideLdsMax) R16 >>= 1;
  a.R16 = R16;
  const size_t lds = (size_t)16 * R16 * row_bytes;
  // subtiles per item: the fewest matrix-core steps for the busiest wave,
  // ties to the wider group (more reuse of each A operand)
  const int ntiles = (int)((N + 15) / 16);
  int ntw = 4, best = 1 << 30;
  for (int c : {4, 2, 1}) {
    const int g = (ntiles + c - 1) / c;
    const int cost = (g + 3) / 4 * c;
    if (cost < best) {
      best = cost;
      ntw = c;
    }
  }
  a.groups = (ntiles + ntw - 1) / ntw;
  // row sub-blocks per item: all of the band when the column groups alone
  // give every wave work, else one (more items)
  const int rw = a.groups >= 4 ? R16 : 1;
  a.nrs = R16 / rw;

  // pack B into the stream's workspace
  const size_t pack_elems = (size_t)a.groups * (S / 2) * ntw * 64 * 2;
  Workspace w;
  skinny_status_t st = acquire_workspace(h, stream, pack_elems * sizeof(T), &w);
  if (st != SKINNY_OK) return st;
  T* Bp = static_cast<T*>(ws_scratch(w));
  a.Bp = Bp;
  {
    const int64_t pairs = (int64_t)pack_elems / 2;
    int pb = (int)std::min<int64_t>((pairs + 255) / 256, 4 * (int64_t)h->cu_count);
    if (pb < 1) pb = 1;
    hipLaunchKernelGGL(tsmm_wide_pack_kernel<T>, dim3(pb), dim3(256), 0, stream,
                       B, ldb, (int)M, (int)N, S, ntw, a.groups, cplx ? 1 : 0, Bp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "tsmm (wide) pack launch");
  }
  const int64_t nblk = (K + 16 * R16 - 1) / (16 * R16);
  // as many workgroups per CU as the LDS holds (at most 4)
  int64_t per_cu = (int64_t)(kWideLdsMax / lds);
  if (per_cu > 4) per_cu = 4;
  if (per_cu < 1) per_cu = 1;
  int64_t blocks = blockCount > 0 ? (int64_t)blockCount : per_cu * (int64_t)h->cu_count;
  TuneEntry te;
  if (blockCount == 0 &&
      tune_lookup(h, 1, cplx ? (sizeof(T) == 8 ? TT_DC : TT_FC) : (sizeof(T) == 8 ? TT_DR : TT_FR),
                  cplx ? M / 2 : M, cplx ? N / 2 : N, K, &te))
    blocks = te.G;
  if (blocks > nblk) blocks = nblk;
  if (blocks < 1) blocks = 1;
  const bool beta0 = cplx ? (beta == T(0) && beta_i == T(0)) : (beta == T(0));
  hipError_t e = cplx ? launch_wide<T, true>(a, rw, ntw, (int)blocks, lds, beta0, stream)
                      : launch_wide<T, false>(a, rw, ntw, (int)blocks, lds, beta0, stream);
  if (e != hipSuccess) return hip_fail(e, "tsmm (wide) launch");
  return SKINNY_OK;
}

template skinny_status_t launch_tsmm_wide<double>(
    skinny_handle_t, size_t, int64_t, int64_t, int64_t, const double*, int64_t,
    double, double, const double*, int64_t, double, double, double*, int64_t,
    bool, hipStream_t);
template skinny_status_t launch_tsmm_wide<float>(
    skinny_handle_t, size_t, int64_t, int64_t, int64_t, const float*, int64_t,
    float, float, const float*, int64_t, float, float, float*, int64_t, bool,
    hipStream_t);

}  // namespace skinny
