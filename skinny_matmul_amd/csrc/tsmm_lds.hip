// TSMM with B staged in LDS: complex types, and real widths 65..128.
//
// The matrix-core mapping is the one of tsmm_mfma_kernel (16-row x 16-column
// C tiles, lane (i, g) owning the contiguous slice g*S .. g*S + S - 1 of row
// i), but B no longer fits in registers, so each workgroup copies it into LDS
// once (up to 160 KB per workgroup on gfx950) and every MFMA reads its B
// operand from there.  N is a runtime loop over 16-column subtiles taken two
// at a time (two independent accumulator chains).  LDS holds the operand
// zero-padded to 4 S rows and an even number of 16-column subtiles, so an
// operand read is one LDS access with no range checks.
//
// Complex TSMM runs on the interleaved real view: with A_r the K x 2M real
// matrix, C_r = A_r * B^ where B^ (2M x 2N) holds, for each complex B(m, n),
//     [ Re  Im ]
//     [-Im  Re ]   at rows 2m, 2m + 1 and columns 2n, 2n + 1,
// written out in LDS when the workgroup stages B (it was formed per operand
// read from planes Re(B) and Im(B): FC M = N = 64 ran at 1.7 TB/s, bound by
// that VALU work).  Lanes 2j and 2j + 1 hold
// the real and imaginary parts of one C entry; the complex alpha and beta
// are applied after exchanging them with a lane swap (tsmm/fix3.cuh:31-35 and
// cu_complex.h axpby: beta * C + alpha * sum).
#include "internal.hpp"
#include "mfma.hpp"

namespace skinny {

template <typename T>
struct TlArgs {
  const T* A;  // real view, row-major, lda in reals
  const T* B;  // real: column-major M x N; complex: interleaved column-major
  T* C;        // real view, row-major, ldc in reals
  int64_t K, lda, ldb, ldc;  // ldb in B elements (complex units if cplx)
  int M, N;                  // real-view widths
  T alpha, beta, alpha_i, beta_i;
  int vec;  // 16-byte A slice loads allowed
  int ns;   // LDS row stride (elements), at least np
  int np;   // padded columns staged (16 x an even subtile count)
};

template <int S>
struct TlBlocks {
  static constexpr int value = S >= 16 ? 1 : (S >= 8 ? 2 : 4);
};

template <typename T>
__device__ __forceinline__ T axpby_l(T s, T c, T alpha, T beta, bool beta0) {
#pragma clang fp contract(off)
  if (beta0) return alpha * s;
  return beta * c + alpha * s;
}

template <typename T, int S, bool BETA0, bool CPLX>
__global__ void __launch_bounds__(256) tsmm_lds_kernel(TlArgs<T> a) {
  using V4 = typename Mfma<T>::v4;
  constexpr int RB = TlBlocks<S>::value;
  extern __shared__ char smem_raw[];
  T* bs = reinterpret_cast<T*>(smem_raw);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int M = a.M, N = a.N, NS = a.ns, NP = a.np;

  // ---- stage B (or B^) in LDS, [m][n], zero outside M x N ----------------
  for (int idx = tid; idx < 4 * S * NP; idx += 256) {
    const int m = idx / NP, n = idx - m * NP;
    T v = T(0);
    if (m < M && n < N) {
      if (CPLX) {
        const int pa = m & 1, pb = n & 1;
        const T* b = a.B + 2 * ((int64_t)(n >> 1) * a.ldb + (m >> 1));
        v = b[pa ^ pb];  // Re on the diagonal of the 2 x 2 block, Im off it
        if (pa && !pb) v = -v;
      } else {
        v = a.B[(int64_t)n * a.ldb + m];
      }
    }
    bs[m * NS + n] = v;
  }
  __syncthreads();

  const T* bl0 = bs + (lg * S) * NS + li;
  auto bval = [&](int s, int u) -> T { return bl0[s * NS + 16 * u]; };

  const int64_t K = a.K;
  const int64_t gw = (int64_t)blockIdx.x * 4 + (tid >> 6);
  const int64_t nw = (int64_t)gridDim.x * 4;
  const int mbeg = lg * S;
  const int NT = (N + 15) / 16;
  const int64_t nblk = (K + 15) >> 4;

  for (int64_t b0 = gw * RB; b0 < nblk; b0 += nw * RB) {
    T av[RB][S];
#pragma unroll
    for (int r = 0; r < RB; r++) {
      const int64_t row = (b0 + r) * 16 + li;
      const bool ok = row < K;
      const T* ar = a.A + row * a.lda + mbeg;
      if (a.vec) {
        // S even, M even, slice start even: pairs are valid or invalid together
#pragma unroll
        for (int s = 0; s < S; s += 2) {
          if (ok && mbeg + s < M) {
            double2 v = *reinterpret_cast<const double2*>(ar + s);
            av[r][s] = (T)v.x;
            av[r][s + 1] = (T)v.y;
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
#pragma unroll
    for (int r = 0; r < RB; r++) {
      for (int u = 0; u < NT; u += 2) {
        V4 acc0 = V4{0, 0, 0, 0}, acc1 = V4{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < S; s++) {
          const T x = av[r][s];
          acc0 = Mfma<T>::mac(x, bval(s, u), acc0);
          acc1 = Mfma<T>::mac(x, bval(s, u + 1), acc1);
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const V4 acc = h ? acc1 : acc0;
          const int col = 16 * (u + h) + li;
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int64_t row = (b0 + r) * 16 + Mfma<T>::drow(lg, v);
            const T own = acc[v];
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
              *c = axpby_l(own, old, a.alpha, a.beta, BETA0);
            }
          }
        }
      }
    }
  }
}

// ---- complex float with the three-product form ---------------------------
// T1 = Ar Br, T2 = Ai Bi, T3 = (Ar + Ai)(Br + Bi); Re = T1 - T2, Im =
// (T3 - T1) - T2: 3/4 of the real view's matrix-core steps and LDS operand
// reads (the real view's 2Mc x 2Nc operand holds each complex B twice).
// Lane (i, g) owns the complex slice g S2 .. g S2 + S2 - 1 of row i of a
// 16-row block (8-byte loads of (re, im), masked to the row's Mc columns),
// forms Ar + Ai in registers, and reads B_j(g S2 + s, 16 u + i) of the three
// staged matrices Br, Bi, Br + Bi (rows of ns floats, zero outside Mc x Nc).
// Element v of its tiles T_j[u] is entry (4 g + v, 16 u + i) for all three
// j, so the lane combines them, applies the complex alpha and beta and
// stores the complex entry with one 8-byte store.  A wave owns whole
// 16-row blocks (all their loads precede their stores), so C may alias A.
// The sums Ar + Ai and Br + Bi are rounded: the imaginary part's error is
// bounded by eps (|Ar| + |Ai|)(|Br| + |Bi|) terms; integer data is exact.
template <int S2, int NT, bool BETA0>
__global__ void __launch_bounds__(256) ctsmm3_kernel(TlArgs<float> a) {
  using V4 = Mfma<float>::v4;
  // software pipeline (below) where it measured faster: always with beta != 0
  // (C's loads hidden behind the matrix cores: Mc = Nc = 64, 2.0 -> 3.6 TB/s)
  // and at 64 x 64 (one wave per SIMD either way); with beta = 0 and room for
  // two waves per SIMD their interleaving does better (Mc = Nc = 48: 4.1
  // against 3.6 TB/s)
  constexpr bool PF = !BETA0 || (S2 >= 16 && NT >= 4);
  extern __shared__ char smem_raw[];
  float* bs = reinterpret_cast<float*>(smem_raw);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int Mc = a.M, Nc = a.N, NS = a.ns, NP = a.np;  // complex widths here
  const int MR = 4 * S2;                               // staged rows per matrix

  for (int idx = tid; idx < MR * NP; idx += 256) {
    const int m = idx / NP, n = idx - m * NP;
    float br = 0.0f, bi = 0.0f;
    if (m < Mc && n < Nc) {
      const float* b = a.B + 2 * ((int64_t)n * a.ldb + m);
      br = b[0];
      bi = b[1];
    }
    bs[m * NS + n] = br;
    bs[(MR + m) * NS + n] = bi;
    bs[(2 * MR + m) * NS + n] = br + bi;
  }
  __syncthreads();

  const float* bl0 = bs + (lg * S2) * NS + li;
  const int64_t K = a.K;
  const int64_t gw = (int64_t)blockIdx.x * 4 + (tid >> 6);
  const int64_t nw = (int64_t)gridDim.x * 4;
  const int mbeg = lg * S2;
  const int64_t nblk = (K + 15) >> 4;
  typedef float f2 __attribute__((ext_vector_type(2)));

  // A slice of 16-row block blk, zero past column Mc (B's rows there are
  // zero, but a NaN or infinity times zero is not).  With beta != 0 the
  // loads go without branches: rows past K read row K - 1 (they feed only
  // rows of D that are never stored) and columns past Mc read column Mc - 1
  // and are then zeroed; with beta = 0 the extra address registers cost a
  // wave per SIMD and the branches measured faster (Mc = Nc = 48: 4.1
  // against 3.0 TB/s), with beta != 0 slower (3.4 against 3.6 TB/s).
  auto load = [&](int64_t blk, float (&xr)[S2], float (&xi)[S2]) {
    if constexpr (BETA0) {
      const int64_t row = blk * 16 + li;
      const bool ok = row < K;
      const float* ar = a.A + 2 * (row * a.lda + mbeg);
#pragma unroll
      for (int s = 0; s < S2; s++) {
        const bool v = ok && mbeg + s < Mc;
        const f2 e = v ? *gmem(reinterpret_cast<const f2*>(ar + 2 * s)) : f2{0.0f, 0.0f};
        xr[s] = e.x;
        xi[s] = e.y;
      }
    } else {
      const int64_t row = blk * 16 + li < K ? blk * 16 + li : K - 1;
      const float* ar = a.A + 2 * row * a.lda;
#pragma unroll
      for (int s = 0; s < S2; s++) {
        const int col = mbeg + s;
        const f2 e = *gmem(reinterpret_cast<const f2*>(ar + 2 * (col < Mc ? col : Mc - 1)));
        xr[s] = col < Mc ? e.x : 0.0f;
        xi[s] = col < Mc ? e.y : 0.0f;
      }
    }
  };
  // C entries of block blk that this lane owns (beta != 0 only); clamped
  // like A, the entries out of range are never stored
  auto cload = [&](int64_t blk, f2 (&cv)[NT][4]) {
#pragma unroll
    for (int u = 0; u < NT; u++) {
      const int n = 16 * u + li < Nc ? 16 * u + li : Nc - 1;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        int64_t row = blk * 16 + Mfma<float>::drow(lg, v);
        row = row < K ? row : K - 1;
        cv[u][v] = *gmem(reinterpret_cast<const f2*>(a.C + 2 * (row * a.ldc + n)));
      }
    }
  };
  auto work = [&](int64_t blk, const float (&xr)[S2], const float (&xi)[S2],
                  const f2 (&cv)[NT][4]) {
    V4 t1[NT], t2[NT], t3[NT];
#pragma unroll
    for (int u = 0; u < NT; u++) t1[u] = t2[u] = t3[u] = V4{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < S2; s++) {
      const float xs = xr[s] + xi[s];
      const float* bp = bl0 + s * NS;
#pragma unroll
      for (int u = 0; u < NT; u++) {
        t1[u] = Mfma<float>::mac(xr[s], bp[16 * u], t1[u]);
        t2[u] = Mfma<float>::mac(xi[s], bp[MR * NS + 16 * u], t2[u]);
        t3[u] = Mfma<float>::mac(xs, bp[2 * MR * NS + 16 * u], t3[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < NT; u++) {
      const int n = 16 * u + li;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int64_t row = blk * 16 + Mfma<float>::drow(lg, v);
        if (row < K && n < Nc) {
#pragma clang fp contract(off)
          const float re = t1[u][v] - t2[u][v];
          const float im = (t3[u][v] - t1[u][v]) - t2[u][v];
          float o_re = a.alpha * re - a.alpha_i * im;
          float o_im = a.alpha * im + a.alpha_i * re;
          if (!BETA0) {
            o_re = (a.beta * cv[u][v].x - a.beta_i * cv[u][v].y) + o_re;
            o_im = (a.beta * cv[u][v].y + a.beta_i * cv[u][v].x) + o_im;
          }
          *gmem(reinterpret_cast<f2*>(a.C + 2 * (row * a.ldc + n))) = f2{o_re, o_im};
        }
      }
    }
  };

  // A block's A loads all precede its stores, and only its own rows of C are
  // read or written, so C may alias A.
  f2 cv[NT][4];
#pragma unroll
  for (int u = 0; u < NT; u++)
#pragma unroll
    for (int v = 0; v < 4; v++) cv[u][v] = f2{0.0f, 0.0f};
  if constexpr (PF) {
    // two register sets: the A loads of the wave's next block are in flight
    // during this block's matrix-core steps; C of this block (beta != 0) is
    // requested before them, so the epilogue waits for it alone
    float xr0[S2], xi0[S2], xr1[S2], xi1[S2];
    if (gw < nblk) {
      load(gw, xr0, xi0);
      for (int64_t b0 = gw; b0 < nblk; b0 += 2 * nw) {
        if (!BETA0) cload(b0, cv);
        load(b0 + nw, xr1, xi1);
        work(b0, xr0, xi0, cv);
        if (b0 + nw >= nblk) break;
        if (!BETA0) cload(b0 + nw, cv);
        load(b0 + 2 * nw, xr0, xi0);
        work(b0 + nw, xr1, xi1, cv);
      }
    }
  } else {
    // beta = 0 with room for two or more waves per SIMD: RB blocks per pass,
    // their loads issued together
    constexpr int RB = S2 >= 16 ? 1 : 2;
    for (int64_t b0 = gw * RB; b0 < nblk; b0 += nw * RB) {
      float xr[RB][S2], xi[RB][S2];
#pragma unroll
      for (int r = 0; r < RB; r++) load(b0 + r, xr[r], xi[r]);
#pragma unroll
      for (int r = 0; r < RB; r++) work(b0 + r, xr[r], xi[r], cv);
    }
  }
}

template <int S2, int NT>
static hipError_t launch_c3(const TlArgs<float>& a, int blocks, size_t lds, bool beta0,
                            hipStream_t st) {
  auto k = beta0 ? ctsmm3_kernel<S2, NT, true> : ctsmm3_kernel<S2, NT, false>;
  hipError_t e = hipFuncSetAttribute((const void*)k,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(blocks), dim3(256), lds, st, a);
  return hipGetLastError();
}

// Complex float, 17..64 complex columns of A, at most 64 of C: the
// three-product kernel above; complex widths, lda / ldc in complex units.
static skinny_status_t launch_ctsmm3(skinny_handle_t h, size_t blockCount, int Mc, int Nc,
                                     int64_t K, const float* A, int64_t lda, float alpha,
                                     float alpha_i, const float* B, int64_t ldb, float beta,
                                     float beta_i, float* C, int64_t ldc, hipStream_t stream) {
  TlArgs<float> a;
  a.A = A;
  a.B = B;
  a.C = C;
  a.K = K;
  a.lda = lda;
  a.ldb = ldb;
  a.ldc = ldc;
  a.M = Mc;
  a.N = Nc;
  a.alpha = alpha;
  a.beta = beta;
  a.alpha_i = alpha_i;
  a.beta_i = beta_i;
  a.vec = 0;
  const int S2 = Mc <= 32 ? 8 : 16;
  const int NT = (Nc + 15) / 16;
  a.np = 16 * NT;
  // the four lane groups of an operand read sit S2 rows apart: a row stride
  // that puts them on different banks where a small pad can
  a.ns = a.np;
  for (int pad = 0; pad <= 16; pad++) {
    const int64_t words = (int64_t)S2 * (a.np + pad);
    if (words % 64 == 16 || words % 64 == 48) {
      a.ns = a.np + pad;
      break;
    }
  }
  const size_t lds = (size_t)3 * 4 * S2 * a.ns * sizeof(float);
  int blocks = blockCount > 0 ? (int)blockCount
                              : tuned_blocks_tsmm(h, 2 * Mc, 2 * Nc, K, 4, false, true, false,
                                                  TT_FC, Mc, Nc);
  const bool beta0 = beta == 0.0f && beta_i == 0.0f;
  hipError_t e;
#define SK_C3(s2, nt) \
  if (S2 == (s2) && NT == (nt)) e = launch_c3<s2, nt>(a, blocks, lds, beta0, stream); else
  SK_C3(8, 1) SK_C3(8, 2) SK_C3(8, 3) SK_C3(8, 4) SK_C3(16, 1) SK_C3(16, 2) SK_C3(16, 3)
  e = launch_c3<16, 4>(a, blocks, lds, beta0, stream);
#undef SK_C3
  if (e != hipSuccess) return hip_fail(e, "tsmm (complex float, 3M) launch");
  return SKINNY_OK;
}

template <typename T, int S, bool CPLX>
static hipError_t launch_lds_s(const TlArgs<T>& a, int blocks, size_t lds,
                               bool beta0, hipStream_t st) {
  auto k = beta0 ? tsmm_lds_kernel<T, S, true, CPLX>
                 : tsmm_lds_kernel<T, S, false, CPLX>;
  hipError_t e = hipFuncSetAttribute((const void*)k,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(blocks), dim3(256), lds, st, a);
  return hipGetLastError();
}

template <typename T, bool CPLX>
static hipError_t launch_lds(const TlArgs<T>& a, int S, int blocks,
                             size_t lds, bool beta0, hipStream_t st) {
  switch (S) {
    case 1: return launch_lds_s<T, 1, CPLX>(a, blocks, lds, beta0, st);
    case 2: return launch_lds_s<T, 2, CPLX>(a, blocks, lds, beta0, st);
    case 4: return launch_lds_s<T, 4, CPLX>(a, blocks, lds, beta0, st);
    case 8: return launch_lds_s<T, 8, CPLX>(a, blocks, lds, beta0, st);
    case 16: return launch_lds_s<T, 16, CPLX>(a, blocks, lds, beta0, st);
    default: return launch_lds_s<T, 32, CPLX>(a, blocks, lds, beta0, st);
  }
}

// M, N, lda, ldc: real-view values; ldb in B elements (complex units when
// cplx).  Widths above 128 go to the wide kernel (tsmm_wide.hip).
// Complex TSMM for narrow blocks (at most 16 complex columns of A): one
// thread owns whole rows, reads its row of A (M complex values) into
// registers, forms the N complex dot products with B and writes the row of C
// (C may alias A: every A value of a row is loaded before it is written).
// B's address is the same in every lane, so its values come through scalar
// loads.  Per entry: re = sum a.x b.x - a.y b.y, im = sum a.x b.y + a.y b.x
// (fused multiply-adds in m order), then C = alpha s + beta C in complex
// arithmetic (cu_complex.h axpby order: beta C + alpha s).
template <typename T>
struct Cx {
  T x, y;
};

__device__ __forceinline__ double cfma(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}
__device__ __forceinline__ float cfma(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

template <typename T, int MC, bool BETA0>
__global__ void __launch_bounds__(256) ztsmm_rows_kernel(TlArgs<T> a) {
  constexpr int R = 4;  // rows in flight
  const int M = a.M / 2, N = a.N / 2;                   // complex widths
  const int64_t K = a.K;
  const int64_t nthreads = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const Cx<T>* __restrict__ Bc = reinterpret_cast<const Cx<T>*>(a.B);
  const int64_t lda = a.lda / 2, ldc = a.ldc / 2;  // complex units
  const Cx<T>* Ac = reinterpret_cast<const Cx<T>*>(a.A);
  Cx<T>* Cc = reinterpret_cast<Cx<T>*>(a.C);

  for (int64_t base = tid; base < K; base += nthreads * R) {
    Cx<T> av[R][MC];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int64_t row = base + (int64_t)r * nthreads;
      const bool ok = row < K;
#pragma unroll
      for (int m = 0; m < MC; m++)
        av[r][m] = (m < M && ok) ? Ac[row * lda + m] : Cx<T>{T(0), T(0)};
    }
    for (int n = 0; n < N; n++) {
      const Cx<T>* bn = Bc + (int64_t)n * a.ldb;
      T sx[R], sy[R];
#pragma unroll
      for (int r = 0; r < R; r++) sx[r] = sy[r] = T(0);
#pragma unroll
      for (int m = 0; m < MC; m++) {
        if (m < M) {
          const Cx<T> b = bn[m];
#pragma unroll
          for (int r = 0; r < R; r++) {
            sx[r] = cfma(av[r][m].x, b.x, sx[r]);
            sx[r] = cfma(-av[r][m].y, b.y, sx[r]);
            sy[r] = cfma(av[r][m].x, b.y, sy[r]);
            sy[r] = cfma(av[r][m].y, b.x, sy[r]);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int64_t row = base + (int64_t)r * nthreads;
        if (row < K) {
#pragma clang fp contract(off)
          Cx<T>* cr = Cc + row * ldc + n;
          Cx<T> o;
          o.x = a.alpha * sx[r] - a.alpha_i * sy[r];
          o.y = a.alpha * sy[r] + a.alpha_i * sx[r];
          if (!BETA0) {
            const Cx<T> c = *cr;
            o.x = (a.beta * c.x - a.beta_i * c.y) + o.x;
            o.y = (a.beta * c.y + a.beta_i * c.x) + o.y;
          }
          *cr = o;
        }
      }
    }
  }
}

template <typename T, int MC>
static hipError_t launch_rows(const TlArgs<T>& a, int blocks, bool beta0,
                              hipStream_t st) {
  if (beta0)
    hipLaunchKernelGGL((ztsmm_rows_kernel<T, MC, true>), dim3(blocks), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((ztsmm_rows_kernel<T, MC, false>), dim3(blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

template <typename T>
skinny_status_t launch_tsmm_lds(skinny_handle_t h, size_t blockCount,
                                int64_t M, int64_t N, int64_t K, const T* A,
                                int64_t lda, T alpha, T alpha_i, const T* B,
                                int64_t ldb, T beta, T beta_i, T* C,
                                int64_t ldc, bool cplx, hipStream_t stream) {
  if (M > 128 || N > 128)
    return launch_tsmm_wide<T>(h, blockCount, M, N, K, A, lda, alpha, alpha_i,
                               B, ldb, beta, beta_i, C, ldc, cplx, stream);
  if constexpr (sizeof(T) == 8) {
    // FP64 with A at least 29 real columns wide: the 4x4x4 matrix-core form
    const skinny_status_t q = launch_tsmm_q4(h, blockCount, M, N, K, A, lda, alpha, alpha_i,
                                             B, ldb, beta, beta_i, C, ldc, cplx, stream);
    if (q != SKINNY_UNSUPPORTED) return q;
  } else {
    // complex float with 17..64 complex columns of A: the three-product form
    if (cplx && M / 2 >= 17 && M / 2 <= 64 && N / 2 <= 64 && (uintptr_t)A % 8 == 0 &&
        (uintptr_t)C % 8 == 0)
      return launch_ctsmm3(h, blockCount, (int)(M / 2), (int)(N / 2), K, (const float*)A,
                           lda / 2, (float)alpha, (float)alpha_i, (const float*)B, ldb,
                           (float)beta, (float)beta_i, (float*)C, ldc / 2, stream);
  }
  TlArgs<T> a;
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
  const int S0 = (int)((M + 3) / 4);
  const int S = S0 <= 1 ? 1 : S0 <= 2 ? 2 : S0 <= 4 ? 4 : S0 <= 8 ? 8
              : S0 <= 16 ? 16 : 32;
  a.vec = (sizeof(T) == 8 && S % 2 == 0 && M % 2 == 0 && lda % 2 == 0 &&
           (uintptr_t)A % 16 == 0) ? 1 : 0;
  // operand rows 4 S, columns an even number of 16-column subtiles; the row
  // stride puts the four k-lane groups of an operand read (S rows apart) on
  // different banks where a small pad can (float: 16 banks apart)
  const int nt = (int)((N + 15) / 16);
  a.np = 16 * (nt + (nt & 1));
  a.ns = a.np;
  for (int pad = 0; pad <= 16; pad++) {
    const int64_t words = (int64_t)S * (a.np + pad) * (int64_t)sizeof(T) / 4;
    if (sizeof(T) == 4 ? (words % 64 == 16 || words % 64 == 48) : (words % 64 == 32)) {
      a.ns = a.np + pad;
      break;
    }
  }
  size_t lds = (size_t)4 * S * a.ns * sizeof(T);
  if (lds > 160 * 1024) {
    a.ns = a.np;
    lds = (size_t)4 * S * a.ns * sizeof(T);
  }
  int blocks = blockCount > 0
                   ? (int)blockCount
                   : tuned_blocks_tsmm(h, M, N, K, (int)sizeof(T), !cplx, true, false,
                                       cplx ? (sizeof(T) == 8 ? TT_DC : TT_FC) : -1, M / 2,
                                       N / 2);
  const bool beta0 = cplx ? (beta == T(0) && beta_i == T(0)) : (beta == T(0));
  // At most 4 complex columns: whole rows per thread.  Wider rows make each
  // lane's loads span several cache lines and the LDS kernel is faster
  // (measured for 8 and 16 complex columns at K = 2^26).
  if (cplx && M <= 8) {
    const int Mc = (int)(M / 2);
    hipError_t e = Mc <= 1 ? launch_rows<T, 1>(a, blocks, beta0, stream)
                 : Mc <= 2 ? launch_rows<T, 2>(a, blocks, beta0, stream)
                           : launch_rows<T, 4>(a, blocks, beta0, stream);
    if (e != hipSuccess) return hip_fail(e, "tsmm (complex rows) launch");
    return SKINNY_OK;
  }
  hipError_t e = cplx ? launch_lds<T, true>(a, S, blocks, lds, beta0, stream)
                      : launch_lds<T, false>(a, S, blocks, lds, beta0, stream);
  if (e != hipSuccess) return hip_fail(e, "tsmm (LDS) launch");
  return SKINNY_OK;
}

template skinny_status_t launch_tsmm_lds<double>(
    skinny_handle_t, size_t, int64_t, int64_t, int64_t, const double*, int64_t,
    double, double, const double*, int64_t, double, double, double*, int64_t,
    bool, hipStream_t);
template skinny_status_t launch_tsmm_lds<float>(
    skinny_handle_t, size_t, int64_t, int64_t, int64_t, const float*, int64_t,
    float, float, const float*, int64_t, float, float, float*, int64_t, bool,
    hipStream_t);

}  // namespace skinny
