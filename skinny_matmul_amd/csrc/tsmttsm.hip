// TSMTTSM  C = alpha * A^T * B + beta * C  on gfx950 matrix cores.
//
// Reference counterparts: the stage-1 blockProductKernel variants and the
// stage-2 deviceReduce of tsmttsm/genv*.cuh and specsmall.cuh (e.g.
// genv3.cuh:9-27 and 31-80, genv7.cuh:28-111), replaced by one kernel:
//
//  * The contraction index of A^T B is the row index k of the row-major
//    K x M / K x N inputs.  One v_mfma_f64_16x16x4f64 consumes four rows:
//    lane (i, g) = (lane % 16, lane / 16) supplies A(row_g, col_i) and
//    B(row_g, col_i), so each 16-lane group reads one contiguous row segment
//    and the matrix core forms the 16 x 16 outer products.  No LDS and no
//    cross-lane shuffles are needed in the streaming loop.
//  * Narrow blocks (max(M, N) <= 8) pack P = 16 / max(M, N) rows into the 16
//    MFMA lanes (lane i = r * M + m); only the block-diagonal of the MFMA
//    result is kept.  M = N = 1 therefore streams 64 rows per instruction.
//  * Wider blocks use 16-wide subtiles, up to 4 x 4 of them (64 x 64) per wave;
//    larger C is split into 64 x 64 tiles along the grid.
//  * The four waves of a workgroup stream disjoint rows and are summed
//    through LDS in a fixed pairwise order; each workgroup writes one partial.
//  * With G > 1 workgroups per tile the G partials of every entry are summed
//    and the alpha/beta epilogue applied (the two stages of genv3.cuh:9-27
//    and 31-80) either by a second kernel (default) or, in the same launch,
//    by the last arrival of each group and then the last group (grid_finish;
//    an autotuner choice).  The in-launch hand-over writes the partials with
//    agent-scope stores and waits for them before arriving, with no L2
//    write-back (a release at agent scope per workgroup had cost 40-60 us per
//    call at 256-512 workgroups).  Measured without it (FP64, MI355X): equal
//    to the two launches within 1.5% at K = 2^26, 2-3 us slower at K = 2^22.
//    Summation order is fixed, so the result is bitwise reproducible for a
//    fixed blockCount and form.
//  * With G == 1 the one workgroup of a tile applies the epilogue itself.
#include "internal.hpp"
#include "mfma.hpp"
#include "tsmttsm_core.hpp"
#include "tsmttsm_q4.hpp"
#include "team_sum.hpp"

namespace skinny {

template <typename T>
struct TsArgs {
  const T* A;
  const T* B;
  T* C;
  int64_t K, lda, ldb, ldc;
  int M, N;
  T alpha, beta;
  T alpha_i, beta_i;  // imaginary parts (complex view only)
  int cplx;           // 1: A, B, C are complex, viewed as interleaved reals
  T* part1;      // [tile][G][PS]
  T* part2;      // [tile][ngroups][PS]
  unsigned* cnt; // [tile][ngroups] then [tile]
  int G, GS, ngroups, tilesM, tilesN, outM, outN;
  int two_pass;  // 1: partials only, summed by tsmttsm_reduce_kernel
};

constexpr int kWaves = 4;

// FP32 tiles of 3 x 3 subtiles and more run the chunk loop over two register
// sets (TsCore PIPE); measured on MI355X at K = 2^26 against the single set
// (GB/s, general / Gram): M = N = 64 3869 -> 4411 / 3540 -> 3792, 48 4360 ->
// 4708 / 3801 -> 4154.
template <int MT, int NT, int P, int ES>
struct Pipe {
  static constexpr bool value = ES == 4 && P == 1 && MT * NT >= 9;
};

template <int MT, int NT, int P, int ES = 8>
struct Unroll {
  // independent loads kept in flight per lane: about 16
  static constexpr int L = MT + NT;
  static constexpr int base = P > 1 ? 8 : (L >= 8 ? 2 : (L >= 4 ? 4 : 8));
  // an FP32 load carries half the bytes: twice the unroll, except for 16
  // packed rows (M = N = 1) and the pipelined tiles (same measurement, GB/s:
  // M = N = 8 6580 -> 6935, 16 6600 -> 7012, 24 4352 -> 4671; Gram 16 5705 ->
  // 6327, 32 5301 -> 6105; Gram M = N = 1 3762 -> 2942, so not there)
  static constexpr int value =
      (ES == 4 && P != 16 && !Pipe<MT, NT, P, ES>::value) ? 2 * base : base;
};

template <typename T>
__device__ __forceinline__ T epilogue(T s, T alpha, T beta, const T* c) {
#pragma clang fp contract(off)
  T r = alpha * s;
  if (beta != T(0)) r = beta * *c + r;
  return r;
}

// Complex result from the interleaved-real product R = A_r^T B_r:
//   C(m, n) = alpha * [(R00 - R11) + i (R01 + R10)] + beta * C(m, n)
// with Rab = R(2m + a, 2n + b); operation order of cu_complex.h (scale, then
// axpy: beta * C + alpha * s), plain transpose (gen_cublas.cuh:30-33).
template <typename T>
__device__ __forceinline__ void complex_epilogue(const TsArgs<T>& a, T r00,
                                                 T r11, T r01, T r10, T* c) {
#pragma clang fp contract(off)
  T re = r00 - r11;
  T im = r01 + r10;
  T o_re = a.alpha * re - a.alpha_i * im;
  T o_im = a.alpha * im + a.alpha_i * re;
  if (a.beta != T(0) || a.beta_i != T(0)) {
    T c_re = c[0], c_im = c[1];
    T b_re = a.beta * c_re - a.beta_i * c_im;
    T b_im = a.beta * c_im + a.beta_i * c_re;
    o_re = b_re + o_re;
    o_im = b_im + o_im;
  }
  c[0] = o_re;
  c[1] = o_im;
}

// Final epilogue from the summed entries in LDS (real or complex view).
template <typename T>
__device__ void final_epilogue(const TsArgs<T>& a, const T* fin, int m0,
                               int n0, int Mloc, int Nloc) {
  if (a.cplx) {
    const int oMc = a.outM / 2, oNc = a.outN / 2;
    for (int e = threadIdx.x; e < oMc * oNc; e += blockDim.x) {
      const int mc = e % oMc, nc = e / oMc;
      if (2 * mc >= Mloc || 2 * nc >= Nloc) continue;
      const int e00 = (2 * nc) * a.outM + 2 * mc;
      T* c = a.C + 2 * ((int64_t)(n0 / 2 + nc) * a.ldc + (m0 / 2 + mc));
      complex_epilogue(a, fin[e00], fin[e00 + a.outM + 1], fin[e00 + a.outM],
                       fin[e00 + 1], c);
    }
    return;
  }
  const int PS = a.outM * a.outN;
  for (int e = threadIdx.x; e < PS; e += blockDim.x) {
    const int ml = e % a.outM, nl = e / a.outM;
    if (ml < Mloc && nl < Nloc) {
      T* c = a.C + (int64_t)(n0 + nl) * a.ldc + (m0 + ml);
      *c = epilogue(fin[e], a.alpha, a.beta, c);
    }
  }
}

// Grid reduction after a workgroup wrote its partial (single-launch form,
// used with G == 1 and when the autotuner chose one launch): the last
// arrival of each group sums the group, then the last group sums the group
// sums and applies the epilogue.  `red` is the kernel's LDS of LDSN
// elements.
template <typename T, int LDSN>
__device__ __forceinline__ void grid_finish(const TsArgs<T>& a, T* red, int tile, int kp,
                                            int ntiles, int m0, int n0, int Mloc,
                                            int Nloc) {
  const int PS = a.outM * a.outN;
  // ---- grid reduction: last arrival per group, then last group ----------
  // The partials were written with agent-scope stores (write-through); an
  // arrival waits for its stores to complete (published(), no L2 write-back)
  // before it counts itself, and only the last arrival takes the acquire
  // side (L2 invalidate) before it reads the other partials.
  __shared__ int s_last;
  published();
  __syncthreads();
  const int grp = kp / a.GS;
  const int gsz = min(a.GS, a.G - grp * a.GS);
  unsigned* cnt1 = a.cnt + tile * a.ngroups + grp;
  unsigned* cnt2 = a.cnt + ntiles * a.ngroups + tile;
  if (threadIdx.x == 0) {
    unsigned ticket = __hip_atomic_fetch_add(cnt1, 1u, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    s_last = (ticket == (unsigned)(gsz - 1));
  }
  __syncthreads();
  if (!s_last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");

  const T* gsrc = a.part1 + ((int64_t)tile * a.G + (int64_t)grp * a.GS) * PS;
  if (a.ngroups == 1) {
    team_sum<T>(gsrc, gsz, PS, nullptr, red, LDSN, true);
    final_epilogue<T>(a, red + LDSN / 2, m0, n0, Mloc, Nloc);
    if (threadIdx.x == 0) *cnt1 = 0u;
    return;
  }
  T* gdst = a.part2 + ((int64_t)tile * a.ngroups + grp) * PS;
  team_sum<T>(gsrc, gsz, PS, gdst, red, LDSN, false);
  if (threadIdx.x == 0) *cnt1 = 0u;
  published();
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned ticket = __hip_atomic_fetch_add(cnt2, 1u, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    s_last = (ticket == (unsigned)(a.ngroups - 1));
  }
  __syncthreads();
  if (!s_last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  team_sum<T>(a.part2 + (int64_t)tile * a.ngroups * PS, a.ngroups, PS, nullptr,
              red, LDSN, true);
  final_epilogue<T>(a, red + LDSN / 2, m0, n0, Mloc, Nloc);
  if (threadIdx.x == 0) *cnt2 = 0u;
}

template <typename T, int P, int MT, int NT, int U, bool SELF>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu((MT * NT >= 9) ? 2 : 1)))
    tsmttsm_kernel(TsArgs<T> a) {
  using Core = TsCore<T, P, MT, NT, U, SELF, Pipe<MT, NT, P, (int)sizeof(T)>::value>;
  using V4 = typename Core::V4;
  constexpr int DM = Core::DM, DN = Core::DN;
  const int wave = threadIdx.x >> 6;

  const int ntiles = a.tilesM * a.tilesN;
  const int tile = blockIdx.x % ntiles;
  const int kp = blockIdx.x / ntiles;
  const int m0 = (tile % a.tilesM) * DM;
  const int n0 = (tile / a.tilesM) * DN;
  const int Mloc = min(a.M - m0, DM);
  const int Nloc = min(a.N - n0, DN);

  V4 acc[MT][NT];
  Core::zero(acc);
  // rows read by one tile only stream past the caches (nontemporal loads);
  // with several column tiles (only possible with 4 subtiles on a side)
  // neighbouring workgroups read the same rows and share them through L2
  if (MT < 4 && NT < 4)
    Core::template accumulate<true>(a.A, a.lda, a.B, a.ldb, a.M, a.N, m0, n0,
                                    Mloc, Nloc, a.K, (int64_t)kp * kWaves + wave,
                                    (int64_t)a.G * kWaves, acc);
  else if (ntiles == 1)
    Core::template accumulate<true>(a.A, a.lda, a.B, a.ldb, a.M, a.N, m0, n0,
                                    Mloc, Nloc, a.K, (int64_t)kp * kWaves + wave,
                                    (int64_t)a.G * kWaves, acc);
  else
    Core::template accumulate<false>(a.A, a.lda, a.B, a.ldb, a.M, a.N, m0, n0,
                                     Mloc, Nloc, a.K, (int64_t)kp * kWaves + wave,
                                     (int64_t)a.G * kWaves, acc);

  __shared__ T red[Core::RED_ELEMS];
  const int PS = a.outM * a.outN;
  T* mine = a.part1 + ((int64_t)tile * a.G + kp) * PS;
  Core::wg_partial(acc, red, mine, a.M, a.N, a.outM, a.outN);
  if (a.two_pass) return;  // the partials are summed by the next kernel
  grid_finish<T, Core::RED_ELEMS>(a, red, tile, kp, ntiles, m0, n0, Mloc, Nloc);
}

// FP64 64 x 64 tiles on the 4x4x4 matrix-core instruction (TsQ4); the same
// tiles, partials and reductions as tsmttsm_kernel with MT = NT = 4.  Two
// waves per SIMD with U = 1 (64 accumulators and two sets of 16 operands per
// lane), one with U = 2.
template <int U, bool FULL, bool ALIGNED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(U <= 1 ? 2 : 1)))
    tsmttsm_q4_kernel(TsArgs<double> a) {
  using Core = TsQ4<U, FULL, ALIGNED>;
  const int wave = threadIdx.x >> 6;
  const int ntiles = a.tilesM * a.tilesN;
  const int tile = blockIdx.x % ntiles;
  const int kp = blockIdx.x / ntiles;
  const int m0 = (tile % a.tilesM) * 64;
  const int n0 = (tile / a.tilesM) * 64;
  const int Mloc = min(a.M - m0, 64);
  const int Nloc = min(a.N - n0, 64);
  double acc[Core::NACC];
  Core::zero(acc);
  // one column tile: rows are read by one workgroup only (nontemporal
  // loads); several tiles share rows through L2
  if (ntiles == 1)
    Core::template accumulate<true>(a.A, a.lda, a.B, a.ldb, m0, n0, Mloc, Nloc, a.K,
                                    (int64_t)kp * kWaves + wave, (int64_t)a.G * kWaves, acc);
  else
    Core::template accumulate<false>(a.A, a.lda, a.B, a.ldb, m0, n0, Mloc, Nloc, a.K,
                                     (int64_t)kp * kWaves + wave, (int64_t)a.G * kWaves, acc);
  __shared__ double red[Core::RED_ELEMS];
  const int PS = a.outM * a.outN;
  double* mine = a.part1 + ((int64_t)tile * a.G + kp) * PS;
  Core::wg_partial(acc, red, mine, a.outM, a.outN);
  if (a.two_pass) return;
  grid_finish<double, Core::RED_ELEMS>(a, red, tile, kp, ntiles, m0, n0, Mloc, Nloc);
}

// Second pass: sums the G partials of each tile entry in a fixed order and
// applies the epilogue.  An item is one output: one entry of the real view,
// or the four entries (2m, 2n), (2m + 1, 2n), (2m, 2n + 1), (2m + 1, 2n + 1)
// of a complex output.  Workgroup b handles E consecutive items of tile
// b / nsl; thread t takes item t % E and the partials j, j + TPE, ... with
// j = t / E and TPE = 256 / E, summed in increasing order, then the TPE
// sums of an item are added pairwise through LDS.
template <typename T>
__global__ void __launch_bounds__(256) tsmttsm_reduce_kernel(TsArgs<T> a, int E) {
  __shared__ T red[4 * 256];
  const int PS = a.outM * a.outN;
  const int oMi = a.cplx ? a.outM / 2 : a.outM;
  const int nitems = a.cplx ? PS / 4 : PS;
  const int nsl = (nitems + E - 1) / E;
  const int tile = blockIdx.x / nsl;
  const int sl = blockIdx.x % nsl;
  const int TPE = 256 / E;
  const int t = threadIdx.x;
  const int il = t % E, j = t / E;
  const int item = sl * E + il;
  const bool live = item < nitems;
  const int mi = item % oMi, ni = item / oMi;
  const int nent = a.cplx ? 4 : 1;
  int ent[4];
  if (a.cplx) {
    const int e00 = (2 * ni) * a.outM + 2 * mi;
    ent[0] = e00;              // R(2m, 2n)
    ent[1] = e00 + a.outM + 1; // R(2m + 1, 2n + 1)
    ent[2] = e00 + a.outM;     // R(2m, 2n + 1)
    ent[3] = e00 + 1;          // R(2m + 1, 2n)
  } else {
    ent[0] = item;
    ent[1] = ent[2] = ent[3] = item;
  }
  const T* src = a.part1 + (int64_t)tile * a.G * PS;
  for (int k = 0; k < nent; k++) {
    T sum = T(0);
    if (live) {
      const T* p = src + ent[k];
      int q = j;
      for (; q + 3 * TPE < a.G; q += 4 * TPE) {
        const T v0 = p[(int64_t)q * PS];
        const T v1 = p[(int64_t)(q + TPE) * PS];
        const T v2 = p[(int64_t)(q + 2 * TPE) * PS];
        const T v3 = p[(int64_t)(q + 3 * TPE) * PS];
        sum += v0;
        sum += v1;
        sum += v2;
        sum += v3;
      }
      for (; q < a.G; q += TPE) sum += p[(int64_t)q * PS];
    }
    red[k * 256 + t] = sum;
  }
  __syncthreads();
  for (int h = TPE / 2; h >= 1; h >>= 1) {
    if (j < h)
      for (int k = 0; k < nent; k++) red[k * 256 + t] += red[k * 256 + t + h * E];
    __syncthreads();
  }
  if (j != 0 || !live) return;
  const int tm = tile % a.tilesM, tn = tile / a.tilesM;
  const int m0 = tm * a.outM, n0 = tn * a.outN;
  const int Mloc = min(a.M - m0, a.outM), Nloc = min(a.N - n0, a.outN);
  if (a.cplx) {
    if (2 * mi >= Mloc || 2 * ni >= Nloc) return;
    T* c = a.C + 2 * ((int64_t)(n0 / 2 + ni) * a.ldc + (m0 / 2 + mi));
    complex_epilogue(a, red[il], red[256 + il], red[512 + il], red[768 + il], c);
  } else {
    if (mi >= Mloc || ni >= Nloc) return;
    T* c = a.C + (int64_t)(n0 + ni) * a.ldc + (m0 + mi);
    *c = epilogue(red[il], a.alpha, a.beta, c);
  }
}

// ---------------------------------------------------------------------------
// one wide operand (9..16 columns) and one narrow (1..8): paired rows
// ---------------------------------------------------------------------------
// Here a 16-lane matrix-core tile takes one row per step, and the kernel is
// bound by the issue of load instructions (about 13 cycles per 64-lane load
// per CU, whatever the active lanes), not by bytes: M = 9, N = 1 issues as
// many loads as M = N = 16 and runs at 3.7 TB/s.  This kernel halves the
// load instructions per row.  Lane (c, g) with c = 8 h + c' loads columns
// 2c', 2c' + 1 of the wide operand's row 2g + h of an 8-row step with one
// 16-byte load, and column c' of the narrow operand's same row.  Two MFMAs
// take the even and the odd columns: D_p(i, j) = sum_g W(row, 2i' + p)
// Q(row, j') is a valid partial product where i and j lie in the same half
// (the same row), i.e. the 8 x 8 diagonal blocks; C(2c + p, q) = D_p(c, q) +
// D_p(8 + c, 8 + q).  SWAP: B is the wide operand (C = (B^T A)^T).
// ALIGNED: the wide rows allow 16-byte loads (even leading dimension,
// aligned base); otherwise the pair is two 8-byte loads (3 loads per 8 rows
// against 4 for the general kernel).  R = 4 is the same for a wide operand
// of 5..8 columns and a narrow one of at most 4: four rows per 16-lane group
// (lanes 4h + c'), the 4 x 4 diagonal blocks summed, 2 loads per 16 rows
// where the packed general kernel needs 2 per 8 or 12.
template <int U, int R, bool SWAP, bool ALIGNED>
__global__ void __launch_bounds__(256) tsmttsm_pair_kernel(TsArgs<double> a) {
  constexpr int L = 16 / R;  // lanes per row
  using V4 = Mfma<double>::v4;
  typedef double d2v __attribute__((ext_vector_type(2)));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int h = li / L, c = li % L;
  const double* W = SWAP ? a.B : a.A;
  const int64_t ldw = SWAP ? a.ldb : a.lda;
  const int MW = SWAP ? a.N : a.M;
  const double* Q = SWAP ? a.A : a.B;
  const int64_t ldq = SWAP ? a.lda : a.ldb;
  const int NQ = SWAP ? a.M : a.N;
  const int rl = R * lg + h;  // the lane's row within a 4 R-row step
  const bool w2 = 2 * c + 1 < MW, w1 = 2 * c < MW, qok = c < NQ;
  const double* Wb = W + (int64_t)rl * ldw + 2 * c;
  const double* Qb = Q + (int64_t)rl * ldq + c;
  const int64_t K = a.K;
  constexpr int CH = 4 * R * U;
  const int64_t nfull = K / CH;
  const int64_t w0 = (int64_t)blockIdx.x * 4 + wave, ws = (int64_t)gridDim.x * 4;
  V4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
  auto ldw2 = [&](const double* p, bool ok) -> d2v {
    if (ALIGNED) {
      if (ok && w2) return __builtin_nontemporal_load(gmem(reinterpret_cast<const d2v*>(p)));
      return d2v{(ok && w1) ? ld_stream(p) : 0.0, 0.0};
    }
    return d2v{(ok && w1) ? ld_stream(p) : 0.0, (ok && w2) ? ld_stream(p + 1) : 0.0};
  };
  for (int64_t ch = w0; ch < nfull; ch += ws) {
    const double* pw = Wb + ch * CH * ldw;
    const double* pq = Qb + ch * CH * ldq;
    d2v wv[U];
    double qv[U];
#pragma unroll
    for (int s = 0; s < U; s++) {
      wv[s] = ldw2(pw + (int64_t)s * 4 * R * ldw, true);
      qv[s] = qok ? ld_stream(pq + (int64_t)s * 4 * R * ldq) : 0.0;
    }
#pragma unroll
    for (int s = 0; s < U; s++) {
      acc0 = Mfma<double>::mac(wv[s].x, qv[s], acc0);
      acc1 = Mfma<double>::mac(wv[s].y, qv[s], acc1);
    }
  }
  // the one partial chunk, taken by the wave whose turn it is
  if (nfull * CH < K && (nfull % ws) == w0) {
    const int64_t r0 = nfull * CH;
#pragma unroll
    for (int s = 0; s < U; s++) {
      const bool ok = r0 + s * 4 * R + rl < K;
      const d2v wv = ldw2(Wb + (r0 + s * 4 * R) * ldw, ok);
      const double qv = (ok && qok) ? ld_stream(Qb + (r0 + s * 4 * R) * ldq) : 0.0;
      acc0 = Mfma<double>::mac(wv.x, qv, acc0);
      acc1 = Mfma<double>::mac(wv.y, qv, acc1);
    }
  }
  // the four waves' D tiles in LDS: red[wave][p][lane][v]
  __shared__ double red[4 * 2 * 64 * 4];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    red[((wave * 2 + 0) * 64 + lane) * 4 + v] = acc0[v];
    red[((wave * 2 + 1) * 64 + lane) * 4 + v] = acc1[v];
  }
  __syncthreads();
  // D_p(i, j) sits in lane 16 (i % 4) + j, element i / 4
  auto dsum = [&](int p, int i, int j) -> double {
    const int at = (16 * (i & 3) + j) * 4 + (i >> 2);
    const double* r = red + p * 256 + at;
    return (r[0] + r[2 * 512]) + (r[512] + r[3 * 512]);
  };
  const int M = a.M, N = a.N, PS = M * N;
  double* out = a.part1 + (int64_t)blockIdx.x * PS;
  for (int e = threadIdx.x; e < PS; e += 256) {
    const int m = e % M, n = e / M;
    const int wi = SWAP ? n : m, qi = SWAP ? m : n;
    const int p = wi & 1, cw = wi >> 1;
    double v = dsum(p, cw, qi);
#pragma unroll
    for (int hh = 1; hh < R; hh++) v += dsum(p, L * hh + cw, L * hh + qi);
    *gmem(out + e) = v;
  }
}

// ---------------------------------------------------------------------------
// planning
// ---------------------------------------------------------------------------
static int isqrt_ceil(int x) {
  int r = 1;
  while (r * r < x) r++;
  return r;
}

TsmttsmPlan plan_tsmttsm(skinny_handle_t h, size_t blockCount, int64_t M,
                         int64_t N, int64_t K, int elem_bytes,
                         int partial_bytes, int tdtype, int64_t tM, int64_t tN) {
  TsmttsmPlan p;
  int64_t W = M > N ? M : N;
  if (W <= 8) {
    p.P = (int)(16 / (W > 0 ? W : 1));
    p.MT = p.NT = 1;
    p.tilesM = p.tilesN = 1;
    p.tileM = 16;
    p.tileN = 16;
    p.outM = (int)M;
    p.outN = (int)N;
  } else {
    p.P = 1;
    int64_t mt = (M + 15) / 16, nt = (N + 15) / 16;
    p.MT = (int)(mt < 4 ? mt : 4);
    p.NT = (int)(nt < 4 ? nt : 4);
    p.tileM = 16 * p.MT;
    p.tileN = 16 * p.NT;
    p.tilesM = (int)((M + p.tileM - 1) / p.tileM);
    p.tilesN = (int)((N + p.tileN - 1) / p.tileN);
    p.outM = p.tileM;
    p.outN = p.tileN;
  }
  const int ntiles = p.tilesM * p.tilesN;
  int U = p.P > 1 ? 8 : ((p.MT + p.NT) >= 8 ? 2 : ((p.MT + p.NT) >= 4 ? 4 : 8));
  int64_t CH = 4LL * p.P * U;
  int64_t G;
  TuneEntry te;
  if (blockCount > 0) {
    G = (int64_t)blockCount;
    p.fixedG = true;
  } else if ((tdtype >= 0 || elem_bytes == 8 || elem_bytes == 4) &&
             tune_lookup(h, 0, tdtype >= 0 ? tdtype : (elem_bytes == 8 ? TT_DR : TT_FR),
                         tdtype >= 0 ? tM : M, tdtype >= 0 ? tN : N, K, &te)) {
    G = te.G;
    p.uvar = te.uvar & 1;
    p.one_pass = (te.uvar & 2) != 0;
    p.tuned = true;
    p.fixedG = true;
  } else {
    // measured on MI355X (fp64): two workgroups per CU.  The autotuner
    // picked two for most of a 12 x 12 grid of widths 1..64; one per CU
    // left narrow asymmetric products (M = 12, N = 1: 3.1 TB/s) short of
    // loads in flight.
    int64_t per_cu = 2;
    // FP32 with one column on each side: 4-byte lane loads keep too few bytes
    // in flight at two per CU (K = 2^30: 2339 GB/s; sixteen per CU: 4922)
    if (elem_bytes == 4 && M == 1 && N == 1) per_cu = 16;
    int64_t target = (per_cu * h->cu_count) / ntiles;
    if (target < 1) target = 1;
    int64_t byrows = K / (CH * kWaves * 2);
    if (byrows < 1) byrows = 1;
    G = byrows < target ? byrows : target;
  }
  if (G > 65536) G = 65536;
  p.G = (int)G;
  // one level when each thread reads at most 8 partials, else two levels
  if (p.G <= 16 || (int64_t)p.G * p.outM * p.outN <= 256 * 8) {
    p.GS = p.G;
    p.ngroups = 1;
  } else {
    p.GS = isqrt_ceil(p.G);
    p.ngroups = (p.G + p.GS - 1) / p.GS;
  }
  size_t PS = (size_t)p.outM * p.outN;
  p.scratch_bytes = (size_t)ntiles * p.G * PS * partial_bytes;
  if (p.ngroups > 1) p.scratch_bytes += (size_t)ntiles * p.ngroups * PS * partial_bytes;
  p.scratch_bytes = (p.scratch_bytes + 255) & ~(size_t)255;
  (void)elem_bytes;
  return p;
}

template <typename T, int P, int MT, int NT, bool SELF>
static void launch_one(const TsArgs<T>& a, int blocks, int uvar,
                       hipStream_t s) {
  constexpr int U0 = Unroll<MT, NT, P, (int)sizeof(T)>::value;
  if (uvar == 0)
    hipLaunchKernelGGL((tsmttsm_kernel<T, P, MT, NT, U0, SELF>), dim3(blocks),
                       dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((tsmttsm_kernel<T, P, MT, NT, 2 * U0, SELF>),
                       dim3(blocks), dim3(256), 0, s, a);
}

template <bool FULL, bool ALIGNED>
static void launch_q4(const TsArgs<double>& a, int blocks, int uvar, hipStream_t s) {
  if (uvar == 0)
    hipLaunchKernelGGL((tsmttsm_q4_kernel<1, FULL, ALIGNED>), dim3(blocks), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((tsmttsm_q4_kernel<2, FULL, ALIGNED>), dim3(blocks), dim3(256), 0, s, a);
}

// 64 x 64 FP64 tiles of a general product: the 4x4x4 instruction (same box,
// best grid each, K = 2^26: M = N = 64 11.06 -> 10.60 ms; M = 256 at
// K = 2^22: 9.51 -> 9.37 ms, its Gram product 9.30 -> 8.93 ms).  A Gram
// product with one tile stays on 16x16x4 steps, which load its one operand
// once (M = 64: 6.55 against 6.72 ms).
static void dispatch_q4(const TsArgs<double>& a, int blocks, int uvar, hipStream_t s) {
  const bool full = a.M % 64 == 0 && a.N % 64 == 0;
  const bool aligned = a.lda % 2 == 0 && a.ldb % 2 == 0 && (uintptr_t)a.A % 16 == 0 &&
                       (uintptr_t)a.B % 16 == 0;
#define SK_Q4(fu, al) \
  if (full == (fu) && aligned == (al)) return launch_q4<fu, al>(a, blocks, uvar, s);
  SK_Q4(true, true) SK_Q4(true, false) SK_Q4(false, true) SK_Q4(false, false)
#undef SK_Q4
}

template <typename T, bool SELF>
static bool dispatch(const TsmttsmPlan& p, const TsArgs<T>& a, int blocks,
                     hipStream_t s) {
  const int uv = p.uvar;
  if constexpr (sizeof(T) == 8) {
    if (!SELF && p.P == 1 && p.MT == 4 && p.NT == 4) {
      dispatch_q4(a, blocks, uv, s);
      return true;
    }
  }
  if (p.P > 1) {
    switch (p.P) {
      case 16: launch_one<T, 16, 1, 1, SELF>(a, blocks, uv, s); return true;
      case 8: launch_one<T, 8, 1, 1, SELF>(a, blocks, uv, s); return true;
      case 5: launch_one<T, 5, 1, 1, SELF>(a, blocks, uv, s); return true;
      case 4: launch_one<T, 4, 1, 1, SELF>(a, blocks, uv, s); return true;
      case 3: launch_one<T, 3, 1, 1, SELF>(a, blocks, uv, s); return true;
      case 2: launch_one<T, 2, 1, 1, SELF>(a, blocks, uv, s); return true;
      default: return false;
    }
  }
#define SK_CASE(mt, nt)                                   \
  case (mt)*8 + (nt):                                     \
    launch_one<T, 1, mt, nt, SELF>(a, blocks, uv, s);     \
    return true;
  switch (p.MT * 8 + p.NT) {
    SK_CASE(1, 1) SK_CASE(1, 2) SK_CASE(1, 3) SK_CASE(1, 4)
    SK_CASE(2, 1) SK_CASE(2, 2) SK_CASE(2, 3) SK_CASE(2, 4)
    SK_CASE(3, 1) SK_CASE(3, 2) SK_CASE(3, 3) SK_CASE(3, 4)
    SK_CASE(4, 1) SK_CASE(4, 2) SK_CASE(4, 3) SK_CASE(4, 4)
    default: return false;
  }
#undef SK_CASE
}

// FP64 products with one operand 9..16 columns wide and the other at most 8
// (two rows per lane group), or 5..8 wide and at most 4 (four rows) run on
// tsmttsm_pair_kernel; returns the rows per group, 0 for the general kernel.
// A 5-wide operand with odd rows stays on the general kernel: its packing (3
// rows in 15 lanes) needs fewer loads per row than two 8-byte loads per
// pair (measured: M = 5, N = 1 5762 against 5315 GB/s).
static int pair_rows(int64_t M, int64_t N, const void* A, int64_t lda, const void* B,
                     int64_t ldb) {
  const int64_t lo = M < N ? M : N, hi = M < N ? N : M;
  const bool aligned = M < N ? (ldb % 2 == 0 && (uintptr_t)B % 16 == 0)
                             : (lda % 2 == 0 && (uintptr_t)A % 16 == 0);
  if (hi >= 9 && hi <= 16 && lo <= 8) return 2;
  if (hi >= 5 && hi <= 8 && lo <= 4 && (hi > 5 || aligned)) return 4;
  return 0;
}

constexpr int kPairUnroll = 8;

// Default grid of tsmttsm_pair_kernel: two workgroups per CU, no more than
// two chunks of 4 R U rows per wave.  Never above the general kernel's
// default for the same shape, so its workspace size covers this one.
static int64_t pair_grid(skinny_handle_t h, int R, int64_t K) {
  int64_t G = 2 * (int64_t)h->cu_count;
  const int64_t byrows = K / (4 * R * kPairUnroll * 4 * 2);
  if (G > byrows) G = byrows;
  return G < 1 ? 1 : G;
}

int tsmttsm_pair_default_blocks(skinny_handle_t h, int64_t M, int64_t N, int64_t K) {
  // rows of 16-byte aligned, even leading dimension assumed
  const int R = pair_rows(M, N, nullptr, 2, nullptr, 2);
  return R > 0 ? (int)pair_grid(h, R, K) : 0;
}

static skinny_status_t launch_tsmttsm_pair(skinny_handle_t h, const TsmttsmPlan& p,
                                           int64_t M, int64_t N, int64_t K, const double* A,
                                           int64_t lda, double alpha, const double* B,
                                           int64_t ldb, double beta, double* C, int64_t ldc,
                                           hipStream_t stream, const Workspace* fixed) {
  constexpr int U = kPairUnroll;
  const int R = pair_rows(M, N, A, lda, B, ldb);
  const int64_t G = p.fixedG ? p.G : pair_grid(h, R, K);
  const int PS = (int)(M * N);
  const size_t need = (size_t)G * PS * sizeof(double);
  Workspace w;
  if (fixed) {
    if (kCounterBytes + need > fixed->bytes)
      return fail(SKINNY_INVALID_ARG, "tsmttsm: partials exceed the given workspace");
    w = *fixed;
  } else {
    skinny_status_t st = acquire_workspace(h, stream, need, &w);
    if (st != SKINNY_OK) return st;
  }
  TsArgs<double> a{};
  a.A = A;
  a.B = B;
  a.C = C;
  a.K = K;
  a.lda = lda;
  a.ldb = ldb;
  a.ldc = ldc;
  a.M = (int)M;
  a.N = (int)N;
  a.alpha = alpha;
  a.beta = beta;
  a.cplx = 0;
  a.part1 = static_cast<double*>(ws_scratch(w));
  a.part2 = a.part1;
  a.cnt = ws_counters(w);
  a.G = (int)G;
  a.GS = (int)G;
  a.ngroups = 1;
  a.tilesM = a.tilesN = 1;
  a.outM = (int)M;
  a.outN = (int)N;
  a.two_pass = 1;
  const bool swap = M < N;  // B is the wide operand
  const bool aligned = swap ? (ldb % 2 == 0 && (uintptr_t)B % 16 == 0)
                            : (lda % 2 == 0 && (uintptr_t)A % 16 == 0);
#define SK_PAIR(r, sw, al)                                                            \
  if (R == (r) && swap == (sw) && aligned == (al))                                   \
    hipLaunchKernelGGL((tsmttsm_pair_kernel<U, r, sw, al>), dim3(G), dim3(256), 0, stream, a);
  SK_PAIR(2, false, true) SK_PAIR(2, false, false) SK_PAIR(2, true, true)
  SK_PAIR(2, true, false) SK_PAIR(4, false, true) SK_PAIR(4, false, false)
  SK_PAIR(4, true, true) SK_PAIR(4, true, false)
#undef SK_PAIR
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "tsmttsm launch");
  int E = G <= 32 ? 64 : 16;
  while (E > 1 && E / 2 >= PS) E >>= 1;
  const int nsl = (PS + E - 1) / E;
  hipLaunchKernelGGL(tsmttsm_reduce_kernel<double>, dim3(nsl), dim3(256), 0, stream, a, E);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "tsmttsm reduce launch");
  return SKINNY_OK;
}

template <typename T>
skinny_status_t launch_tsmttsm_real(skinny_handle_t h, const TsmttsmPlan& p,
                                    int64_t M, int64_t N, int64_t K,
                                    const T* A, int64_t lda, T alpha,
                                    const T* B, int64_t ldb, T beta, T* C,
                                    int64_t ldc, hipStream_t stream,
                                    const Workspace* fixed) {
  return launch_tsmttsm<T>(h, p, M, N, K, A, lda, alpha, T(0), B, ldb, beta,
                           T(0), C, ldc, false, stream, fixed);
}

template <typename T>
skinny_status_t launch_tsmttsm(skinny_handle_t h, const TsmttsmPlan& p,
                               int64_t M, int64_t N, int64_t K, const T* A,
                               int64_t lda, T alpha, T alpha_i, const T* B,
                               int64_t ldb, T beta, T beta_i, T* C,
                               int64_t ldc, bool cplx, hipStream_t stream,
                               const Workspace* fixed) {
  if (!cplx && sizeof(T) == 8 && pair_rows(M, N, A, lda, B, ldb) > 0)
    return launch_tsmttsm_pair(h, p, M, N, K, (const double*)A, lda, (double)alpha,
                               (const double*)B, ldb, (double)beta, (double*)C, ldc,
                               stream, fixed);
  const int ntiles = p.tilesM * p.tilesN;
  if ((size_t)ntiles * (p.ngroups + 1) > (size_t)kMaxCounters)
    return fail(SKINNY_UNSUPPORTED, "tsmttsm: too many partial groups");
  Workspace w;
  if (fixed) {
    if (kCounterBytes + p.scratch_bytes > fixed->bytes)
      return fail(SKINNY_INVALID_ARG, "tsmttsm: partials exceed the given workspace");
    w = *fixed;
  } else {
    skinny_status_t st = acquire_workspace(h, stream, p.scratch_bytes, &w);
    if (st != SKINNY_OK) return st;
  }
  TsArgs<T> a;
  a.A = A;
  a.B = B;
  a.C = C;
  a.K = K;
  a.lda = lda;
  a.ldb = ldb;
  a.ldc = ldc;
  a.M = (int)M;
  a.N = (int)N;
  a.alpha = alpha;
  a.beta = beta;
  a.alpha_i = alpha_i;
  a.beta_i = beta_i;
  a.cplx = cplx ? 1 : 0;
  a.part1 = static_cast<T*>(ws_scratch(w));
  a.part2 = a.part1 + (size_t)ntiles * p.G * p.outM * p.outN;
  a.cnt = ws_counters(w);
  a.G = p.G;
  a.GS = p.GS;
  a.ngroups = p.ngroups;
  a.tilesM = p.tilesM;
  a.tilesN = p.tilesN;
  a.outM = p.outM;
  a.outN = p.outN;
  // the Gram shortcut (B operand = A operand) holds only when the one tile is
  // on the diagonal; off-diagonal tiles of a large Gram read two column blocks
  const bool self = (A == B) && (lda == ldb) && (M == N) && ntiles == 1;
  const int blocks = p.G * ntiles;
  // G > 1: a second kernel sums the partials, or (one_pass, an autotuner
  // choice) the last arrivals of the same launch do (grid_finish)
  a.two_pass = (p.G > 1 && !p.one_pass) ? 1 : 0;
  // a Gram product loads one operand per step, half the loads in flight of
  // the general product at the same unroll: untuned, it takes the doubled
  // unroll (M = 24 at K = 2^22: 3.4 TB/s with the base unroll)
  TsmttsmPlan q = p;
  if (self && !p.tuned) q.uvar = 1;
  bool ok = self ? dispatch<T, true>(q, a, blocks, stream)
                 : dispatch<T, false>(q, a, blocks, stream);
  if (!ok) return fail(SKINNY_UNSUPPORTED, "tsmttsm: no kernel for this shape");
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "tsmttsm launch");
  if (a.two_pass) {
    // items per workgroup: 16 (a 128-byte line of one partial per 16 lanes,
    // and TPE = 16 threads per item) or 64 when G <= 32 (at most 8 partials
    // per thread); no more than the item count needs
    const int PS = p.outM * p.outN;
    const int nitems = cplx ? PS / 4 : PS;
    int E = p.G <= 32 ? 64 : 16;
    while (E > 1 && E / 2 >= nitems) E >>= 1;
    const int nsl = (nitems + E - 1) / E;
    hipLaunchKernelGGL(tsmttsm_reduce_kernel<T>, dim3(ntiles * nsl), dim3(256), 0,
                       stream, a, E);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "tsmttsm reduce launch");
  }
  return SKINNY_OK;
}

template skinny_status_t launch_tsmttsm<double>(
    skinny_handle_t, const TsmttsmPlan&, int64_t, int64_t, int64_t,
    const double*, int64_t, double, double, const double*, int64_t, double,
    double, double*, int64_t, bool, hipStream_t, const Workspace*);
template skinny_status_t launch_tsmttsm<float>(
    skinny_handle_t, const TsmttsmPlan&, int64_t, int64_t, int64_t,
    const float*, int64_t, float, float, const float*, int64_t, float, float,
    float*, int64_t, bool, hipStream_t, const Workspace*);
template skinny_status_t launch_tsmttsm_real<double>(
    skinny_handle_t, const TsmttsmPlan&, int64_t, int64_t, int64_t,
    const double*, int64_t, double, const double*, int64_t, double, double*,
    int64_t, hipStream_t, const Workspace*);
template skinny_status_t launch_tsmttsm_real<float>(
    skinny_handle_t, const TsmttsmPlan&, int64_t, int64_t, int64_t,
    const float*, int64_t, float, const float*, int64_t, float, float*,
    int64_t, hipStream_t, const Workspace*);

}  // namespace skinny
