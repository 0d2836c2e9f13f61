// TSMM  C = alpha * A * B + beta * C  (out of place, or in place with C == A).
//
// Reference counterparts: tsmm/fix1..4.cuh, fix_fb.cuh, var1.cuh,
// varip1..3.cuh, fix_ip_ghost.cuh, var_ip_ghost.cuh and the blends
// fix_blend.cuh / varip_blend.cuh.  One design replaces them:
//
//  * One thread owns whole rows: it reads row k of A (M values) into
//    registers, forms the N dot products with B and writes row k of C.  A row
//    is read completely before the same row is written, and no other thread
//    touches that row, so C may alias A (in place) with no barrier
//    (the reference's varip kernels need __syncthreads between the reads and
//    the writes of a row, varip1.cuh:18).
//  * B is uniform across the wavefront: every B(m, n) access has the same
//    address in all lanes.  FP32 fetches it with scalar loads into SGPRs,
//    the scalar operand of v_fma (the reference stages B through shared
//    memory, fix3.cuh:10-18); FP64 keeps vector loads, which measured faster
//    there (tm_bload).
//  * Row loads and stores use 16-byte accesses (two doubles or four floats),
//    else two-element ones, when the leading dimensions and base pointers
//    allow it.
//  * The row length is rounded up to a compile-time class MC in
//    {1, 2, 4, 8, 16, 32, 64} and masked with uniform branches; N is a runtime
//    loop unrolled by two, so one instantiation serves a range of widths.
//
// For 3 <= M <= 64 and N <= 64, and for M <= 2 with 4 <= N <= 64, the matrix
// cores take over (tsmm_mfma_kernel):
//  * One v_mfma_f64_16x16x4f64 produces a 16-row x 16-column tile of C from
//    four columns of A: lane (i, g) = (lane % 16, lane / 16) supplies
//    A(row0 + i, m_g) and B(m_g, n0 + i).
//  * Lane (i, g) owns the contiguous slice m = g*S .. g*S + S - 1 of row
//    row0 + i (S = ceil(M / 4)), read two elements per load; MFMA step s uses
//    m_g = g*S + s.  The 16 rows x M columns of a block are one contiguous
//    region when lda == M.
//  * B is the same for every row: each lane keeps its S x NT operand values
//    in registers for the whole kernel.
//  * A wave owns whole 16-row blocks: every A load of a block completes before
//    the MFMAs that consume it, and C of the block is stored after them, so C
//    may alias A.
//
// Complex products with a real beta run on the same matrix-core kernel
// through the interleaved real view: A (K x M complex) is K x 2M real, C is
// K x 2N real, and C = A_r Bh + beta C with the 2M x 2N real matrix
//   Bh(2m, 2n) = Re b',  Bh(2m + 1, 2n) = -Im b',
//   Bh(2m, 2n + 1) = Im b',  Bh(2m + 1, 2n + 1) = Re b',   b' = alpha B(m, n),
// formed from the complex B while the operands are loaded into registers.
#include "internal.hpp"
#include "mfma.hpp"

namespace skinny {

template <typename T>
struct TmArgs {
  const T* A;
  const T* B;
  T* C;
  int64_t K, lda, ldb, ldc;
  int M, N;
  T alpha, beta;
  // complex view (tsmm_mfma_kernel, tsmm_kernel with CPLX): B is complex
  // M/2 x N/2 with ldb in complex units, alpha = 1 and (cre, cim) the
  // complex alpha folded into Bh
  int cplxB;
  T cre, cim;
  // in place with beta != 0 and N <= M (tsmm_mfma_kernel only): the old C
  // columns are A's own columns, so C = A (fa B + fb I) with alpha = 1,
  // beta = 0 and no read of C
  int fold;
  T fa, fb;
};

// Two consecutive elements as one load or store (16 bytes for double, 8 for
// float): half the memory instructions of element-wise access.
template <typename T>
struct Vec2 {
  typedef T type __attribute__((ext_vector_type(2)));
};

// Element (m, n) of the real B operand: B itself, or Bh of the complex view.
template <typename T>
__device__ __forceinline__ T tm_bval(const TmArgs<T>& a, int m, int n) {
  if (!a.cplxB) return a.B[(int64_t)n * a.ldb + m];
  const T* b = a.B + 2 * ((int64_t)(n >> 1) * a.ldb + (m >> 1));
  const T br = b[0], bi = b[1];
  T pr, pi;
  {
#pragma clang fp contract(off)
    pr = a.cre * br - a.cim * bi;
    pi = a.cre * bi + a.cim * br;
  }
  if ((m & 1) == 0) return (n & 1) == 0 ? pr : pi;
  return (n & 1) == 0 ? -pi : pr;
}

// B(m, n) in the whole-row kernel, the same address in every lane: FP32
// reads it with scalar loads into SGPRs (through a constant-memory view; with
// vector loads each step of the column loop waited for them, which shows from
// about 8 columns of C: M = 12, N = 8, beta = 1, best grids, 3280 -> 4887
// GB/s), FP64 with vector loads (scalar ones measured 4-12% slower at M =
// 2..8 with C one or two columns wide, in place)
template <typename T>
__device__ __forceinline__ T tm_bload(const T* p) {
  if constexpr (sizeof(T) == 4)
    return *cmem(p);
  else
    return *p;
}

template <typename T>
__device__ __forceinline__ T fma_t(T a, T b, T c) {
  return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fma_t<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

template <typename T>
__device__ __forceinline__ T axpby_t(T s, T c, T alpha, T beta, bool beta0) {
#pragma clang fp contract(off)
  // tsmm/fix3.cuh:31-35: beta * C + alpha * sum (C unread when beta == 0)
  if (beta0) return alpha * s;
  return beta * c + alpha * s;
}

// rows per thread per iteration (independent rows in flight); FP32 takes two
// rows up to 16 wide (four at 4 wide measured 4-13% slower with C one to
// three columns wide, 1-4% faster at M = 3 with beta = 0)
template <int MC, int EB>
struct TmRows {
  static constexpr int value =
      EB == 8 ? (MC <= 2 ? 4 : (MC <= 8 ? 2 : 1)) : (MC <= 2 ? 4 : (MC <= 16 ? 2 : 1));
};

// VW consecutive elements as one load or store
template <typename T, int VW>
struct VecN {
  typedef T type __attribute__((ext_vector_type(VW)));
};

// VW: elements per row access, when the leading dimensions and the base
// pointers allow it: 1, 2 (16 bytes of double, 8 of float) or 4 (16 bytes of
// float).  CPLX: B is the real view Bh of a complex B (tm_bval), for complex
// products one column wide (real view M = N = 2)
template <typename T, int MC, bool BETA0, int VW, bool CPLX = false>
__global__ void __launch_bounds__(256) tsmm_kernel(TmArgs<T> a) {
  constexpr int R = TmRows<MC, (int)sizeof(T)>::value;
  constexpr int LW = VW < MC ? VW : MC;  // elements per A load
  const int M = a.M, N = a.N;
  const int64_t K = a.K;
  const int64_t nthreads = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const T* __restrict__ B = a.B;

  for (int64_t base = tid; base < K; base += nthreads * R) {
    T av[R][MC];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int64_t row = base + (int64_t)r * nthreads;
      const bool ok = row < K;
      const T* ar = a.A + row * a.lda;
#pragma unroll
      for (int m = 0; m < MC; m += LW) {
        if (m < M) {
          if (LW > 1 && ok && m + LW <= M) {
            const typename VecN<T, LW>::type v =
                *reinterpret_cast<const typename VecN<T, LW>::type*>(ar + m);
#pragma unroll
            for (int j = 0; j < LW; j++) av[r][m + j] = v[j];
          } else {
#pragma unroll
            for (int j = 0; j < LW; j++) av[r][m + j] = (ok && m + j < M) ? ar[m + j] : T(0);
          }
        } else {
#pragma unroll
          for (int j = 0; j < LW; j++) av[r][m + j] = T(0);
        }
      }
    }
    // every A value of these rows is now in registers: in-place safe
    int n = 0;
    if constexpr (VW == 4) {
      // four columns of C per access
      for (; n + 4 <= N; n += 4) {
        T sv[4][R];
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
          for (int r = 0; r < R; r++) sv[j][r] = T(0);
        const T* b0 = B + (int64_t)n * a.ldb;
#pragma unroll
        for (int m = 0; m < MC; m++) {
          if (m < M) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
              const T x = CPLX ? tm_bval(a, m, n + j) : tm_bload(b0 + j * a.ldb + m);
#pragma unroll
              for (int r = 0; r < R; r++) sv[j][r] = fma_t(av[r][m], x, sv[j][r]);
            }
          }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int64_t row = base + (int64_t)r * nthreads;
          if (row < K) {
            typedef typename VecN<T, 4>::type V;
            V* cr = reinterpret_cast<V*>(a.C + row * a.ldc + n);
            V c = BETA0 ? V{} : *cr;
            V o;
#pragma unroll
            for (int j = 0; j < 4; j++) o[j] = axpby_t(sv[j][r], c[j], a.alpha, a.beta, BETA0);
            *cr = o;
          }
        }
      }
    }
    for (; n + 2 <= N; n += 2) {
      T s0[R], s1[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        s0[r] = T(0);
        s1[r] = T(0);
      }
      const T* b0 = B + (int64_t)n * a.ldb;
      const T* b1 = b0 + a.ldb;
#pragma unroll
      for (int m = 0; m < MC; m++) {
        if (m < M) {
          T x0 = CPLX ? tm_bval(a, m, n) : tm_bload(b0 + m);
          T x1 = CPLX ? tm_bval(a, m, n + 1) : tm_bload(b1 + m);
#pragma unroll
          for (int r = 0; r < R; r++) {
            s0[r] = fma_t(av[r][m], x0, s0[r]);
            s1[r] = fma_t(av[r][m], x1, s1[r]);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int64_t row = base + (int64_t)r * nthreads;
        if (row < K) {
          T* cr = a.C + row * a.ldc + n;
          T c0 = BETA0 ? T(0) : cr[0];
          T c1 = BETA0 ? T(0) : cr[1];
          T o0 = axpby_t(s0[r], c0, a.alpha, a.beta, BETA0);
          T o1 = axpby_t(s1[r], c1, a.alpha, a.beta, BETA0);
          if (VW >= 2) {
            *reinterpret_cast<typename VecN<T, 2>::type*>(cr) = typename VecN<T, 2>::type{o0, o1};
          } else {
            cr[0] = o0;
            cr[1] = o1;
          }
        }
      }
    }
    if (n < N) {
      T s0[R];
#pragma unroll
      for (int r = 0; r < R; r++) s0[r] = T(0);
      const T* b0 = B + (int64_t)n * a.ldb;
#pragma unroll
      for (int m = 0; m < MC; m++) {
        if (m < M) {
          T x0 = CPLX ? tm_bval(a, m, n) : tm_bload(b0 + m);
#pragma unroll
          for (int r = 0; r < R; r++) s0[r] = fma_t(av[r][m], x0, s0[r]);
        }
      }
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int64_t row = base + (int64_t)r * nthreads;
        if (row < K) {
          T* cr = a.C + row * a.ldc + n;
          T c0 = BETA0 ? T(0) : cr[0];
          cr[0] = axpby_t(s0[r], c0, a.alpha, a.beta, BETA0);
        }
      }
    }
  }
}

// Outer product C = alpha a b^T + beta C (M = 1, FP64) into a packed C
// (ldc == N, 16-byte aligned).  A wave owns 64 consecutive rows, which are
// 64 N consecutive values of C, and lane l computes the element pairs
// 2 (l + 64 s): every store of the wave is one contiguous 1 KiB, and so is
// every load of old C.  Whole-row threads stored 16 bytes per lane at a
// stride of one row, so that each store of a wave touched 64 cache lines.
// The wave's 64 values of a and the N values of b are read through LDS.
// Each value is fma(a, b, 0) and the same epilogue as the other kernels,
// so results are bitwise theirs.  In place (A == C, lda == N) a wave reads
// its rows' first values before it stores any of them.
template <bool BETA0>
__global__ void __launch_bounds__(256) tsmm_outer_kernel(TmArgs<double> a) {
  __shared__ double sa[4][64];
  __shared__ double sb[64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int N = a.N;
  for (int n = threadIdx.x; n < N; n += 256) sb[n] = a.B[(int64_t)n * a.ldb];
  __syncthreads();
  const int64_t nblk = (a.K + 63) / 64;
  const int64_t nw = (int64_t)gridDim.x * 4;
  // pair p + 64 of a lane is element 2 p + 128: 128 / N rows and 128 % N
  // columns further on
  const int drow = 128 / N, dcol = 128 % N;
  const int row_l = (2 * lane) / N, col_l = (2 * lane) % N;
  for (int64_t blk = (int64_t)blockIdx.x * 4 + w; blk < nblk; blk += nw) {
    const int64_t r0 = blk * 64;
    const int rows = (int)min<int64_t>(64, a.K - r0);
    sa[w][lane] = lane < rows ? a.A[(r0 + lane) * a.lda] : 0.0;
    // the wave's LDS accesses are done in order: every lane's value is
    // written before any lane reads another's
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
    const int nel = rows * N;
    double* cb = a.C + r0 * N;
    int row = row_l, col = col_l;
    for (int e = 2 * lane; e < nel; e += 128) {
      const int row1 = col + 1 == N ? row + 1 : row;
      const int col1 = col + 1 == N ? 0 : col + 1;
      const double v0 = fma_t(sa[w][row], sb[col], 0.0);
      if (e + 1 < nel) {
        const double v1 = fma_t(sa[w][row1], sb[col1], 0.0);
        typedef double d2 __attribute__((ext_vector_type(2)));
        d2* cp = reinterpret_cast<d2*>(cb + e);
        const d2 c = BETA0 ? d2{0.0, 0.0} : *cp;
        *cp = d2{axpby_t(v0, c.x, a.alpha, a.beta, BETA0), axpby_t(v1, c.y, a.alpha, a.beta, BETA0)};
      } else {
        const double c = BETA0 ? 0.0 : cb[e];
        cb[e] = axpby_t(v0, c, a.alpha, a.beta, BETA0);
      }
      col += dcol;
      row += drow;
      if (col >= N) {
        col -= N;
        row++;
      }
    }
    // the next block's a values replace this one's only after every lane
    // has read them
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
  }
}

// ---------------------------------------------------------------------------
// matrix-core TSMM
// ---------------------------------------------------------------------------
template <int S>
struct TmBlocks {
  // 16-row blocks per wave iteration: about 16 loads in flight per lane
  static constexpr int value = S >= 16 ? 1 : (S >= 8 ? 2 : (S >= 4 ? 4 : 8));
};

// TR: the operands swap places and the MFMA produces the transposed tile:
// lane (i, g) then holds C(row0 + i, 16 u + drow(g, v)), so one store
// instruction writes 16 whole rows of up to 4 columns from all 64 lanes,
// where the untransposed tile needs 4 stores with N of every 16 lanes (for
// C at most 8 wide: 1 or 2 stores per 16 rows against 4).
template <typename T, int S, int NT, bool BETA0, bool VEC, bool TR = false>
__global__ void __launch_bounds__(256) tsmm_mfma_kernel(TmArgs<T> a) {
  using V4 = typename Mfma<T>::v4;
  // with beta != 0 the old C tiles of all RB blocks are loaded with A (at
  // most 64 values per lane), so their latency is not paid once per store
  constexpr int RB0 = TmBlocks<S>::value;
  constexpr int RB = BETA0 ? RB0 : (RB0 < 16 / NT ? RB0 : 16 / NT);
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int M = a.M, N = a.N;
  const int64_t K = a.K;
  const int64_t gw = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t nw = (int64_t)gridDim.x * 4;

  // B operand of step s, subtile u: B(g*S + s, 16u + i), B column-major
  T bop[S][NT];
#pragma unroll
  for (int s = 0; s < S; s++)
#pragma unroll
    for (int u = 0; u < NT; u++) {
      const int m = lg * S + s, n = 16 * u + li;
      T b = (m < M && n < N) ? tm_bval(a, m, n) : T(0);
      if (a.fold && m < M && n < N) {
#pragma clang fp contract(off)
        b = a.fa * b + (m == n ? a.fb : T(0));
      }
      bop[s][u] = b;
    }
  const int mbeg = lg * S;

  const int64_t nblk = (K + 15) >> 4;
  for (int64_t b0 = gw * RB; b0 < nblk; b0 += nw * RB) {
    T av[RB][S];
#pragma unroll
    for (int r = 0; r < RB; r++) {
      const int64_t row = (b0 + r) * 16 + li;
      const bool ok = row < K;
      const T* ar = a.A + row * a.lda + mbeg;
      if (VEC) {
        // S even, lda even, A aligned to two elements, M a multiple of 4: no masks
#pragma unroll
        for (int s = 0; s < S; s += 2) {
          if (ok) {
            const typename Vec2<T>::type v =
                *reinterpret_cast<const typename Vec2<T>::type*>(ar + s);
            av[r][s] = v.x;
            av[r][s + 1] = v.y;
          } else {
            av[r][s] = T(0);
            av[r][s + 1] = T(0);
          }
        }
      } else {
#pragma unroll
        for (int s = 0; s < S; s++)
          av[r][s] = (ok && mbeg + s < M) ? ar[s] : T(0);
      }
    }
    T cold[BETA0 ? 1 : RB][BETA0 ? 1 : NT][4];
    if (!BETA0) {
#pragma unroll
      for (int r = 0; r < RB; r++)
#pragma unroll
        for (int u = 0; u < NT; u++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int64_t row = (b0 + r) * 16 + (TR ? li : Mfma<T>::drow(lg, v));
            const int col = 16 * u + (TR ? Mfma<T>::drow(lg, v) : li);
            cold[BETA0 ? 0 : r][BETA0 ? 0 : u][v] =
                (row < K && col < N) ? a.C[row * a.ldc + col] : T(0);
          }
    }
#pragma unroll
    for (int r = 0; r < RB; r++) {
      V4 acc[NT];
#pragma unroll
      for (int u = 0; u < NT; u++) acc[u] = V4{0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s < S; s++)
#pragma unroll
        for (int u = 0; u < NT; u++)
          acc[u] = TR ? Mfma<T>::mac(bop[s][u], av[r][s], acc[u])
                      : Mfma<T>::mac(av[r][s], bop[s][u], acc[u]);
#pragma unroll
      for (int u = 0; u < NT; u++) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int64_t row = (b0 + r) * 16 + (TR ? li : Mfma<T>::drow(lg, v));
          const int col = 16 * u + (TR ? Mfma<T>::drow(lg, v) : li);
          if (row < K && col < N) {
            T* c = a.C + row * a.ldc + col;
            const T old = BETA0 ? T(0) : cold[BETA0 ? 0 : r][BETA0 ? 0 : u][v];
            *c = axpby_t(acc[u][v], old, a.alpha, a.beta, BETA0);
          }
        }
      }
    }
  }
}

// Slice length per lane group (columns of A per MFMA k-lane) rounded up to
// an instantiated class: 1, 2, 4, 6, 8, 12, 16.  6 and 12 keep widths
// 21..24 and 41..48 from padding a third of their matrix-core steps.
static int tm_slice_class(int S) {
  return S <= 1 ? 1 : S <= 2 ? 2 : S <= 4 ? 4 : S <= 6 ? 6 : S <= 8 ? 8 : S <= 12 ? 12 : 16;
}

template <typename T, int S, int NT, bool TR>
static void launch_mfma_s_nt_tr(const TmArgs<T>& a, int blocks, bool beta0,
                                bool vec, hipStream_t st) {
  if (vec && S % 2 == 0) {
    if (beta0)
      hipLaunchKernelGGL((tsmm_mfma_kernel<T, S, NT, true, (S % 2 == 0), TR>),
                         dim3(blocks), dim3(256), 0, st, a);
    else
      hipLaunchKernelGGL((tsmm_mfma_kernel<T, S, NT, false, (S % 2 == 0), TR>),
                         dim3(blocks), dim3(256), 0, st, a);
  } else {
    if (beta0)
      hipLaunchKernelGGL((tsmm_mfma_kernel<T, S, NT, true, false, TR>),
                         dim3(blocks), dim3(256), 0, st, a);
    else
      hipLaunchKernelGGL((tsmm_mfma_kernel<T, S, NT, false, false, TR>),
                         dim3(blocks), dim3(256), 0, st, a);
  }
}

template <typename T, int S, int NT>
static void launch_mfma_s_nt(const TmArgs<T>& a, int blocks, bool beta0,
                             bool vec, hipStream_t st) {
  // measured on MI355X: the transposed tile wins for C at most 4 wide
  // (M = 12, N = 1: 4817 -> 5108 GB/s) and, with beta != 0, up to 8 wide
  // (old C loaded coalesced too: M = 5, N = 5, beta = 1: 3712 -> 5146 GB/s);
  // with beta = 0 and 6..8 columns the untransposed tile is faster
  const bool tr = NT == 1 && (a.N <= 4 || (!beta0 && a.N <= 8));
  if (tr)
    launch_mfma_s_nt_tr<T, S, 1, true>(a, blocks, beta0, vec, st);
  else
    launch_mfma_s_nt_tr<T, S, NT, false>(a, blocks, beta0, vec, st);
}

template <typename T, int S>
static void launch_mfma_s(const TmArgs<T>& a, int blocks, int nt, bool beta0,
                          bool vec, hipStream_t st) {
  switch (nt) {
    case 1: launch_mfma_s_nt<T, S, 1>(a, blocks, beta0, vec, st); break;
    case 2: launch_mfma_s_nt<T, S, 2>(a, blocks, beta0, vec, st); break;
    case 3: launch_mfma_s_nt<T, S, 3>(a, blocks, beta0, vec, st); break;
    default: launch_mfma_s_nt<T, S, 4>(a, blocks, beta0, vec, st); break;
  }
}

// Whole-row threads (tsmm_kernel) or the matrix cores (tsmm_mfma_kernel), for
// M, N <= 64.  Measured on MI355X with K (M + N) = 2^28 doubles: a C at most
// two columns wide uses 2 of the 16 lanes of each matrix-core store, and whole
// rows win up to M = 6 (out of place: M = 5, N = 1 4246 -> 4999 GB/s) or
// M = 8 (in place); three columns up to M = 3.  In place with beta != 0 and C
// wider than A, the matrix-core kernel's old-C tile loads are mostly masked
// lanes, and whole rows win for M <= 3 up to N = 4 (M = 1, N = 2: 1385 ->
// 3793 GB/s).  Out of place with beta != 0, whole rows win up to M = 12 for
// N <= 2 (M = 5, N = 1: 1990 -> 4932 GB/s) and at M <= 4 for N = 4; from
// N = 3 on the matrix cores' transposed tile (coalesced old-C loads) wins
// otherwise (M = 8, N = 3: 3919 -> 5084 GB/s).
//
// FP32 rows take 16-byte accesses and scalar B loads, and measured faster than
// the matrix cores (K = 2^26, the rows at four workgroups per CU) also for C
// three or four columns wide with A up to 12 wide (M = N = 4: 3443 -> 5086
// GB/s; M = 12, N = 4, beta = 1: 3511 -> 5105; in place M = N = 3: 3776 ->
// 4042), except for C three columns wide with A 7..8 wide, or 9..12 wide and
// beta = 0 out of place (0.90-0.94 of the matrix cores), and for C eight
// columns wide with beta != 0 out of place (A 4..12 wide; M = 12, N = 8: 3024
// -> 4954) or in place (A 2..12 wide).
static bool tm_rows_wins(int64_t M, int64_t N, bool inplace, bool beta0, int eb) {
  if (eb == 4 && M <= 12) {
    if (N == 4) return true;
    if (N == 3 && (M <= 6 || (M > 8 && (!beta0 || inplace)))) return true;
    if (N == 8 && (inplace ? M >= 2 : (!beta0 && M >= 4))) return true;
  }
  if (!beta0 && !inplace) {
    if (N <= 2) return M <= 12;
    return N == 4 && M <= 4;
  }
  if (N <= 2) return M <= (inplace ? 8 : 6);
  if (N == 3) return M <= 3;
  if (inplace && !beta0 && N > M) return M <= 3 && N <= 4;
  return false;
}

int tuned_blocks_tsmm(skinny_handle_t h, int64_t M, int64_t N, int64_t K,
                      int elem_bytes, bool real, bool beta0, bool inplace,
                      int tdtype, int64_t tM, int64_t tN) {
  TuneEntry te;
  if (tdtype < 0 && real && (elem_bytes == 8 || elem_bytes == 4)) {
    tdtype = elem_bytes == 8 ? TT_DR : TT_FR;
    tM = M;
    tN = N;
  }
  if (tdtype >= 0 && tune_lookup(h, 1, tdtype, tM, tN, K, &te)) return te.G;
  // measured on MI355X at K = 2^26 (fp64): the one-row-per-thread kernel
  // wants four workgroups per CU for M = 1 or beta != 0 (old C loaded too)
  // and two otherwise; the matrix-core kernel wants eight when A is at most
  // 4 columns wide (few loads per lane) and two otherwise
  int64_t want = (K + 255) / 256;
  int64_t per_cu;
  const bool rows = real ? (M <= 64 && N <= 64 && tm_rows_wins(M, N, inplace, beta0, elem_bytes))
                         : (M <= 2 && N <= 2);
  // FP32 rows (16-byte accesses, up to four rows per thread): four per CU,
  // three from 8 columns of C on (more stores per row), eight for A and C at
  // most 4 and 3 wide with beta != 0 out of place, and for M = N = 1
  // (the complex view's one-column rows keep eight)
  if (rows && elem_bytes == 4 && real)
    per_cu = ((!beta0 && !inplace && M <= 4 && N <= 3) || (M == 1 && N == 1)) ? 8
             : N >= 8                                                           ? 3
                                                                                : 4;
  else if (rows)
    per_cu = elem_bytes == 4 ? 8 : ((M <= 1 || !beta0) ? 4 : 2);
  else
    per_cu = M <= 4 ? 8 : 2;
  int64_t cap = per_cu * h->cu_count;
  if (want > cap) want = cap;
  if (want < 1) want = 1;
  return (int)want;
}

template <typename T, int MC, bool BETA0, int VW, bool CPLX>
static void launch_tm(const TmArgs<T>& a, int blocks, hipStream_t s) {
  hipLaunchKernelGGL((tsmm_kernel<T, MC, BETA0, VW, CPLX>), dim3(blocks), dim3(256), 0, s, a);
}

template <typename T, int MC, bool BETA0, bool CPLX>
static void launch_vw(const TmArgs<T>& a, int blocks, int vw, hipStream_t s) {
  if constexpr (sizeof(T) == 4) {
    if (vw >= 4) return launch_tm<T, MC, BETA0, 4, CPLX>(a, blocks, s);
  }
  if (vw >= 2) return launch_tm<T, MC, BETA0, 2, CPLX>(a, blocks, s);
  launch_tm<T, MC, BETA0, 1, CPLX>(a, blocks, s);
}

// vw: elements per row access (row_vw)
template <typename T, int MC, bool CPLX = false>
static void launch_mc(const TmArgs<T>& a, int blocks, bool beta0, int vw, hipStream_t s) {
  if (beta0)
    launch_vw<T, MC, true, CPLX>(a, blocks, vw, s);
  else
    launch_vw<T, MC, false, CPLX>(a, blocks, vw, s);
}

// widest row access that every row start of A and C allows: 16 bytes (two
// doubles or four floats), else two elements, else one
template <typename T>
static int row_vw(const T* A, int64_t lda, const T* C, int64_t ldc) {
  const uintptr_t pa = (uintptr_t)A, pc = (uintptr_t)C;
  if (sizeof(T) == 4 && lda % 4 == 0 && ldc % 4 == 0 && pa % 16 == 0 && pc % 16 == 0) return 4;
  if (lda % 2 == 0 && ldc % 2 == 0 && pa % (2 * sizeof(T)) == 0 && pc % (2 * sizeof(T)) == 0)
    return 2;
  return 1;
}

template <typename T>
skinny_status_t launch_tsmm_real(skinny_handle_t h, size_t blockCount,
                                 int64_t M, int64_t N, int64_t K, const T* A,
                                 int64_t lda, T alpha, const T* B, int64_t ldb,
                                 T beta, T* C, int64_t ldc, hipStream_t stream) {
  if (M > 64 || N > 64)
    return launch_tsmm_lds<T>(h, blockCount, M, N, K, A, lda, alpha, T(0), B,
                              ldb, beta, T(0), C, ldc, false, stream);
  TmArgs<T> a;
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
  a.cplxB = 0;
  a.cre = T(1);
  a.cim = T(0);
  a.fold = 0;
  a.fa = T(1);
  a.fb = T(0);
  bool beta0 = (beta == T(0));
  const bool inplace = (const void*)A == (const void*)C;
  int blocks = blockCount > 0
                   ? (int)blockCount
                   : tuned_blocks_tsmm(h, M, N, K, (int)sizeof(T), true, beta0, inplace);
  // with blockCount 0, variant bit 0 of a tuned entry: the other kernel form
  bool rows = tm_rows_wins(M, N, inplace, beta0, (int)sizeof(T));
  TuneEntry te;
  if (blockCount == 0 && tune_lookup(h, 1, sizeof(T) == 8 ? TT_DR : TT_FR, M, N, K, &te) &&
      (te.uvar & 1))
    rows = !rows;
  // M = 1 into a packed, 16-byte aligned C: one contiguous block per wave
  // (tsmm_outer_kernel), where it measured faster than the other kernels
  // (K = 2^25, same box): out of place for N = 3..16 with beta = 0 (N = 3:
  // 3033 -> 4945 GB/s, N = 5: 3415 -> 5083, N = 11: 3913 -> 4795) and
  // N = 4..16 otherwise (N = 4: 3537 -> 5300), in place with beta = 0 for
  // N = 2..3 (3424 -> 3995) and with beta != 0 for N = 3..8 (N = 4: 1866 ->
  // 3198).  Outside these ranges the other kernels were as fast or up to 9%
  // faster.
  if constexpr (sizeof(T) == 8) {
    const bool outer_wins = inplace ? (beta0 ? N >= 2 && N <= 3 : N >= 3 && N <= 8)
                                    : (beta0 ? N >= 3 : N >= 4) && N <= 16;
    if (M == 1 && outer_wins && ldc == N && (uintptr_t)C % 16 == 0 &&
        (!inplace || lda == ldc)) {
      const int ob = blockCount > 0 ? (int)blockCount
                                    : (int)std::min<int64_t>(std::max<int64_t>(1, (K + 255) / 256),
                                                             4 * (int64_t)h->cu_count);
      if (beta0)
        hipLaunchKernelGGL((tsmm_outer_kernel<true>), dim3(ob), dim3(256), 0, stream, a);
      else
        hipLaunchKernelGGL((tsmm_outer_kernel<false>), dim3(ob), dim3(256), 0, stream, a);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return hip_fail(e, "tsmm outer launch");
      return SKINNY_OK;
    }
  }
  if (!rows) {
    if (!beta0 && (const void*)A == (const void*)C && lda == ldc && N <= M) {
      a.fold = 1;
      a.fa = alpha;
      a.fb = beta;
      a.alpha = T(1);
      a.beta = T(0);
      beta0 = true;
    }
    const int S = (int)((M + 3) / 4);
    const int Sc = tm_slice_class(S);
    const int nt = (int)((N + 15) / 16);
    // unmasked two-element row slices need every slice start aligned to two
    // elements
    const bool vecm = (M % 4 == 0) && (S % 2 == 0) && (Sc == S) && (lda % 2 == 0) &&
                      ((uintptr_t)A % (2 * sizeof(T)) == 0);
    switch (Sc) {
      case 1: launch_mfma_s<T, 1>(a, blocks, nt, beta0, false, stream); break;
      case 2: launch_mfma_s<T, 2>(a, blocks, nt, beta0, vecm, stream); break;
      case 4: launch_mfma_s<T, 4>(a, blocks, nt, beta0, vecm, stream); break;
      case 6: launch_mfma_s<T, 6>(a, blocks, nt, beta0, vecm, stream); break;
      case 8: launch_mfma_s<T, 8>(a, blocks, nt, beta0, vecm, stream); break;
      case 12: launch_mfma_s<T, 12>(a, blocks, nt, beta0, vecm, stream); break;
      default: launch_mfma_s<T, 16>(a, blocks, nt, beta0, vecm, stream); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "tsmm launch");
    return SKINNY_OK;
  }
  const int vec = row_vw(A, lda, C, ldc);
  int MC = M <= 1 ? 1 : M <= 2 ? 2 : M <= 4 ? 4 : M <= 8 ? 8 : M <= 16 ? 16
           : M <= 32 ? 32 : 64;
  switch (MC) {
    case 1: launch_mc<T, 1>(a, blocks, beta0, 1, stream); break;
    case 2: launch_mc<T, 2>(a, blocks, beta0, vec, stream); break;
    case 4: launch_mc<T, 4>(a, blocks, beta0, vec, stream); break;
    case 8: launch_mc<T, 8>(a, blocks, beta0, vec, stream); break;
    case 16: launch_mc<T, 16>(a, blocks, beta0, vec, stream); break;
    case 32: launch_mc<T, 32>(a, blocks, beta0, vec, stream); break;
    default: launch_mc<T, 64>(a, blocks, beta0, vec, stream); break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "tsmm launch");
  return SKINNY_OK;
}

template <typename T>
skinny_status_t launch_tsmm_cview(skinny_handle_t h, size_t blockCount,
                                  int64_t Mc, int64_t Nc, int64_t K, const T* A,
                                  int64_t lda, T alpha_r, T alpha_i, const T* B,
                                  int64_t ldb, T beta, T* C, int64_t ldc,
                                  hipStream_t stream) {
  const int64_t M = 2 * Mc, N = 2 * Nc;
  if ((Mc < 2 && !(Mc == 1 && Nc == 1)) || M > 64 || N > 64)
    return fail(SKINNY_UNSUPPORTED,
                "tsmm complex view: 2 <= M <= 32, N <= 32, or M = N = 1");
  TmArgs<T> a;
  a.A = A;
  a.B = B;
  a.C = C;
  a.K = K;
  a.lda = 2 * lda;
  a.ldb = ldb;
  a.ldc = 2 * ldc;
  a.M = (int)M;
  a.N = (int)N;
  a.alpha = T(1);
  a.beta = beta;
  a.cplxB = 1;
  a.cre = alpha_r;
  a.cim = alpha_i;
  a.fold = 0;
  a.fa = T(1);
  a.fb = T(0);
  int blocks = blockCount > 0 ? (int)blockCount
                              : tuned_blocks_tsmm(h, M, N, K, (int)sizeof(T), false, true,
                                                  false, sizeof(T) == 8 ? TT_DC : TT_FC, Mc,
                                                  Nc);
  bool beta0 = (beta == T(0));
  if (Mc == 1) {
    // one complex column: the real view is 2 x 2, whole rows per thread
    // (a row is one complex value, one two-element access)
    const int vec = ((uintptr_t)A % (2 * sizeof(T)) == 0) &&
                            ((uintptr_t)C % (2 * sizeof(T)) == 0)
                        ? 2
                        : 1;
    launch_mc<T, 2, true>(a, blocks, beta0, vec, stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "tsmm (complex view) launch");
    return SKINNY_OK;
  }
  // in place, real beta, N <= M: beta I joins the real-view operand (the
  // complex alpha is already in it)
  if (!beta0 && (const void*)A == (const void*)C && lda == ldc && N <= M) {
    a.fold = 1;
    a.fb = beta;
    a.beta = T(0);
    beta0 = true;
  }
  const int S = (int)((M + 3) / 4);
  const int Sc = tm_slice_class(S);
  const int nt = (int)((N + 15) / 16);
  const bool vecm = (M % 4 == 0) && (S % 2 == 0) && (Sc == S) && (a.lda % 2 == 0) &&
                    ((uintptr_t)A % (2 * sizeof(T)) == 0);
  switch (Sc) {
    case 1: launch_mfma_s<T, 1>(a, blocks, nt, beta0, false, stream); break;
    case 2: launch_mfma_s<T, 2>(a, blocks, nt, beta0, vecm, stream); break;
    case 4: launch_mfma_s<T, 4>(a, blocks, nt, beta0, vecm, stream); break;
    case 6: launch_mfma_s<T, 6>(a, blocks, nt, beta0, vecm, stream); break;
    case 8: launch_mfma_s<T, 8>(a, blocks, nt, beta0, vecm, stream); break;
    case 12: launch_mfma_s<T, 12>(a, blocks, nt, beta0, vecm, stream); break;
    default: launch_mfma_s<T, 16>(a, blocks, nt, beta0, vecm, stream); break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "tsmm (complex view) launch");
  return SKINNY_OK;
}

template skinny_status_t launch_tsmm_cview<double>(
    skinny_handle_t, size_t, int64_t, int64_t, int64_t, const double*, int64_t,
    double, double, const double*, int64_t, double, double*, int64_t, hipStream_t);
template skinny_status_t launch_tsmm_cview<float>(
    skinny_handle_t, size_t, int64_t, int64_t, int64_t, const float*, int64_t,
    float, float, const float*, int64_t, float, float*, int64_t, hipStream_t);

template skinny_status_t launch_tsmm_real<double>(
    skinny_handle_t, size_t, int64_t, int64_t, int64_t, const double*, int64_t,
    double, const double*, int64_t, double, double*, int64_t, hipStream_t);
template skinny_status_t launch_tsmm_real<float>(
    skinny_handle_t, size_t, int64_t, int64_t, int64_t, const float*, int64_t,
    float, const float*, int64_t, float, float*, int64_t, hipStream_t);

}  // namespace skinny
