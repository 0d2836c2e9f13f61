// FP64 TSMM  C = alpha A B^ + beta C  on the 4x4x4 matrix-core instruction,
// with the B operand staged in LDS: the products that reach the LDS path
// (real widths 65..128; complex ones above 32 columns or with a complex beta)
// whose real view is at least 29 columns wide.
//
// v_mfma_f64_4x4x4f64 runs four independent 4 x 4 x 4 products per
// instruction; on MI355X it sustains about 70 TFLOP/s against about 47 for
// v_mfma_f64_16x16x4f64, the instruction of tsmm_lds_kernel, so products that
// are matrix-core bound there (complex double at 64 columns: 2 M^2 flop per
// byte of A) gain up to the ratio.
//
// Mapping (lane l = 16 X + 4 b + y; block b takes rows 4 b .. 4 b + 3 of a
// 16-row group):
//  * A operand: lane (X, b, y) supplies A(row0 + 4 b + y, X S + s) at step
//    s, so a lane owns the contiguous slice X S .. X S + S - 1 of one row
//    (S = M / 4 rounded up to a class), read two elements per load.
//  * B operand: lane (X, b, y) supplies B^(X S + s, 4 q + y) for column group
//    q; the four blocks share it, so the 64 lanes read 16 distinct LDS
//    addresses per access (a broadcast), two column groups per 16-byte read.
//  * Result: lane (X, b, y) holds C(row0 + 4 b + X, 4 q + y), one
//    accumulator per column group.
// LDS holds B^ as rows r = (4 s + X) 4 + y of QT + 2 doubles (column group q
// at column q; the two padding doubles put the 16 rows of one access on
// distinct banks).  B^ is B for real products and, for complex ones, the
// real view [[Re, Im], [-Im, Re]] of each complex B(m, n) at rows 2m, 2m + 1
// and columns 2n, 2n + 1, as in tsmm_lds_kernel; alpha and beta are applied
// in the epilogue in the same order as there (complex ones after a lane swap
// between the real and imaginary lanes y, y ^ 1).
//
// A wave owns whole 16-row groups: every A load of a group completes before
// the group's products, and its C values are stored after them (each lane
// reads the old C value it overwrites), so C may alias A.
#include "internal.hpp"
#include "mfma.hpp"

namespace skinny {

struct TqArgs {
  const double* A;  // real view, row-major, lda in doubles
  const double* B;  // real: column-major M x N; complex: interleaved column-major
  double* C;        // real view, row-major, ldc in doubles
  int64_t K, lda, ldb, ldc;  // ldb in B elements (complex units if complex)
  int M, N;                  // real-view widths
  int QT;                    // column groups staged (multiple of QC)
  int vec;                   // unmasked 16-byte A slice loads
  double alpha, beta, alpha_i, beta_i;
};

typedef double tq2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double q4mac(double a, double b, double c) {
  return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

// RB: 16-row groups per wave and pass; each LDS read of B^ feeds 2 RB
// matrix-core steps.
template <int S, int QC, int RB, int NW, bool BETA0, bool CPLX>
__global__ void __launch_bounds__(64 * NW) tsmm_q4_kernel(TqArgs a) {
  extern __shared__ char smem_raw[];
  double* bl = reinterpret_cast<double*>(smem_raw);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int X = lane >> 4, blk = (lane >> 2) & 3, y = lane & 3;
  const int M = a.M, N = a.N, QT = a.QT;
  const int R = QT + 2;  // LDS row stride (doubles)

  // ---- stage B^ in LDS ---------------------------------------------------
  for (int idx = tid; idx < 16 * S * QT; idx += 64 * NW) {
    const int r = idx / QT, q = idx - r * QT;
    const int s = r >> 4, xx = (r >> 2) & 3, yy = r & 3;
    const int m = xx * S + s, n = 4 * q + yy;
    double v = 0.0;
    if (m < M && n < N) {
      if (CPLX) {
        const int pa = m & 1, pb = n & 1;
        const double* b = a.B + 2 * ((int64_t)(n >> 1) * a.ldb + (m >> 1));
        v = b[pa ^ pb];  // Re on the diagonal of the 2 x 2 block, Im off it
        if (pa && !pb) v = -v;
      } else {
        v = a.B[(int64_t)n * a.ldb + m];
      }
    }
    bl[r * R + q] = v;
  }
  for (int idx = tid; idx < 16 * S; idx += 64 * NW) {
    bl[idx * R + QT] = 0.0;
    bl[idx * R + QT + 1] = 0.0;
  }
  __syncthreads();

  const int64_t K = a.K;
  const int64_t gw = (int64_t)blockIdx.x * NW + (tid >> 6);
  const int64_t nw = (int64_t)gridDim.x * NW;
  const int64_t nblk = (K + 15) >> 4;
  const int mbeg = X * S;
  const double* brow = bl + ((4 * 0 + X) * 4 + y) * R;  // row of step 0
  constexpr int RS = 16;  // LDS rows between consecutive steps

  for (int64_t g = gw * RB; g < nblk; g += nw * RB) {
    double av[RB][S];
#pragma unroll
    for (int r = 0; r < RB; r++) {
      const int64_t row = (g + r) * 16 + 4 * blk + y;
      const bool ok = row < K;
      const double* ar = a.A + row * a.lda + mbeg;
      if (a.vec) {
#pragma unroll
        for (int s = 0; s < S; s += 2) {
          if (ok) {
            const double2 v = *reinterpret_cast<const double2*>(ar + s);
            av[r][s] = v.x;
            av[r][s + 1] = v.y;
          } else {
            av[r][s] = 0.0;
            av[r][s + 1] = 0.0;
          }
        }
      } else {
#pragma unroll
        for (int s = 0; s < S; s++) av[r][s] = (ok && mbeg + s < M) ? ar[s] : 0.0;
      }
    }
    for (int q0 = 0; q0 < QT; q0 += QC) {
      double cold[BETA0 ? 1 : RB][BETA0 ? 1 : QC];
      if (!BETA0) {
#pragma unroll
        for (int r = 0; r < RB; r++) {
          const int64_t orow = (g + r) * 16 + 4 * blk + X;
#pragma unroll
          for (int q = 0; q < QC; q++) {
            const int col = 4 * (q0 + q) + y;
            cold[BETA0 ? 0 : r][BETA0 ? 0 : q] =
                (orow < K && col < N) ? a.C[orow * a.ldc + col] : 0.0;
          }
        }
      }
      double acc[RB][QC];
#pragma unroll
      for (int r = 0; r < RB; r++)
#pragma unroll
        for (int q = 0; q < QC; q++) acc[r][q] = 0.0;
#pragma unroll
      for (int s = 0; s < S; s++) {
        const double* bp = brow + s * RS * R + q0;
#pragma unroll
        for (int p = 0; p < QC; p += 2) {
          const tq2 b2 = *lmem(reinterpret_cast<const tq2*>(bp + p));
#pragma unroll
          for (int r = 0; r < RB; r++) {
            acc[r][p] = q4mac(av[r][s], b2.x, acc[r][p]);
            acc[r][p + 1] = q4mac(av[r][s], b2.y, acc[r][p + 1]);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < RB; r++) {
        const int64_t orow = (g + r) * 16 + 4 * blk + X;  // this lane's output row
        const bool rok = orow < K;
        double* crow = a.C + orow * a.ldc;
#pragma unroll
        for (int q = 0; q < QC; q++) {
          const int col = 4 * (q0 + q) + y;
          const double own = acc[r][q];
          if (CPLX) {
            // lanes y and y ^ 1 hold the two parts of one complex entry (N
            // is even, so both are in range or neither is)
            const double partner = __shfl_xor(own, 1);
            const double mine = BETA0 ? 0.0 : cold[BETA0 ? 0 : r][BETA0 ? 0 : q];
            const double other = BETA0 ? 0.0 : __shfl_xor(mine, 1);
            if (rok && col < N) {
#pragma clang fp contract(off)
              const bool im_lane = (y & 1) != 0;
              const double re = im_lane ? partner : own;
              const double im = im_lane ? own : partner;
              double out = im_lane ? (a.alpha * im + a.alpha_i * re)
                                   : (a.alpha * re - a.alpha_i * im);
              if (!BETA0) {
                const double cr = im_lane ? other : mine, ci = im_lane ? mine : other;
                const double bc = im_lane ? (a.beta * ci + a.beta_i * cr)
                                          : (a.beta * cr - a.beta_i * ci);
                out = bc + out;
              }
              crow[col] = out;
            }
          } else if (rok && col < N) {
#pragma clang fp contract(off)
            crow[col] = BETA0 ? a.alpha * own
                              : a.beta * cold[BETA0 ? 0 : r][BETA0 ? 0 : q] + a.alpha * own;
          }
        }
      }
    }
  }
}

// ---- complex double with the three-product form ---------------------------
// C = alpha A B + beta C for complex A (K x Mc), B (Mc x Nc) as three real
// products (the 3M form): T1 = Ar Br, T2 = Ai Bi, T3 = (Ar + Ai)(Br + Bi),
// then Re = T1 - T2, Im = T3 - T1 - T2.  Against the 2Mc x 2Nc real view
// (four real products per complex one) this is 3/4 of the matrix-core steps
// and of the LDS operand reads, which bound the real view from about 40
// complex columns on.  The sums Ar + Ai and Br + Bi are rounded, so the
// imaginary part's error is bounded by eps (|Ar| + |Ai|)(|Br| + |Bi|) terms
// rather than the four-product form's; on integer data both are exact.
//
// Mapping as tsmm_q4_kernel, with complex slices: lane (X, b, y) supplies
// A_j(row0 + 4 b + y, X S2 + s) at step (j, s) (j = 0: Ar, 1: Ai, 2: Ar + Ai;
// the lane owns complex columns X S2 .. X S2 + S2 - 1 of one row, loaded
// 16 bytes at a time), and B_j(X S2 + s, 4 q + y) from LDS (B_0 = Br, B_1 =
// Bi, B_2 = Br + Bi, each laid out as tsmm_q4_kernel's B^ with QT + 2
// doubles per row).  Lane (X, b, y) holds T_j(row0 + 4 b + X, 4 q + y) for
// the three j and combines them itself: one 16-byte store of the complex
// entry per column group.
template <int S2, int QC, int NW, bool BETA0>
__global__ void __launch_bounds__(64 * NW) tsmm_q4c3_kernel(TqArgs a) {
  extern __shared__ char smem_raw[];
  double* bl = reinterpret_cast<double*>(smem_raw);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int X = lane >> 4, blk = (lane >> 2) & 3, y = lane & 3;
  const int Mc = a.M, Nc = a.N, QT = a.QT;  // complex widths here
  const int R = QT + 2;                     // LDS row stride (doubles)
  const int MR = 16 * S2;                   // LDS rows of one B_j

  // ---- stage B_0, B_1, B_2 in LDS -----------------------------------------
  for (int idx = tid; idx < MR * QT; idx += 64 * NW) {
    const int r = idx / QT, q = idx - r * QT;
    const int s = r >> 4, xx = (r >> 2) & 3, yy = r & 3;
    const int m = xx * S2 + s, n = 4 * q + yy;
    double br = 0.0, bi = 0.0;
    if (m < Mc && n < Nc) {
      const double* b = a.B + 2 * ((int64_t)n * a.ldb + m);
      br = b[0];
      bi = b[1];
    }
    bl[r * R + q] = br;
    bl[(MR + r) * R + q] = bi;
    bl[(2 * MR + r) * R + q] = br + bi;
  }
  for (int idx = tid; idx < 3 * MR; idx += 64 * NW) {
    bl[idx * R + QT] = 0.0;
    bl[idx * R + QT + 1] = 0.0;
  }
  __syncthreads();

  const int64_t K = a.K;
  const int64_t gw = (int64_t)blockIdx.x * NW + (tid >> 6);
  const int64_t nw = (int64_t)gridDim.x * NW;
  const int64_t nblk = (K + 15) >> 4;
  const int mbeg = X * S2;  // first complex column of the lane's slice
  const double* brow = bl + (X * 4 + y) * R;  // row of step 0 in B_0
  constexpr int RS = 16;                      // LDS rows between consecutive steps

  for (int64_t g = gw; g < nblk; g += nw) {
    const int64_t row = g * 16 + 4 * blk + y;
    const bool ok = row < K;
    const double* ar = a.A + 2 * (row * a.lda + mbeg);
    // one 16-byte load per complex element (A is 16-byte aligned), masked to
    // the row's Mc columns: a slice may extend past them (Mc < 4 S2), and on
    // the last rows such a read would leave the matrix
    double xr[S2], xi[S2];
#pragma unroll
    for (int s = 0; s < S2; s++) {
      const bool v = ok && mbeg + s < Mc;
      const tq2 e = v ? *gmem(reinterpret_cast<const tq2*>(ar + 2 * s)) : tq2{0.0, 0.0};
      xr[s] = e.x;
      xi[s] = e.y;
    }
    const int64_t orow = g * 16 + 4 * blk + X;  // this lane's output row
    const bool rok = orow < K;
    double* crow = a.C + 2 * (orow * a.ldc);
    for (int q0 = 0; q0 < QT; q0 += QC) {
      tq2 cold[BETA0 ? 1 : QC];
      if (!BETA0) {
#pragma unroll
        for (int q = 0; q < QC; q++) {
          const int n = 4 * (q0 + q) + y;
          cold[BETA0 ? 0 : q] =
              (rok && n < Nc) ? *gmem(reinterpret_cast<const tq2*>(crow + 2 * n)) : tq2{0.0, 0.0};
        }
      }
      double t1[QC], t2[QC], t3[QC];
#pragma unroll
      for (int q = 0; q < QC; q++) t1[q] = t2[q] = t3[q] = 0.0;
#pragma unroll
      for (int s = 0; s < S2; s++) {
        const double* bp = brow + s * RS * R + q0;
        const double xs = xr[s] + xi[s];
#pragma unroll
        for (int p = 0; p < QC; p += 2) {
          const tq2 b0 = *lmem(reinterpret_cast<const tq2*>(bp + p));
          const tq2 b1 = *lmem(reinterpret_cast<const tq2*>(bp + MR * R + p));
          const tq2 b2 = *lmem(reinterpret_cast<const tq2*>(bp + 2 * MR * R + p));
          t1[p] = q4mac(xr[s], b0.x, t1[p]);
          t1[p + 1] = q4mac(xr[s], b0.y, t1[p + 1]);
          t2[p] = q4mac(xi[s], b1.x, t2[p]);
          t2[p + 1] = q4mac(xi[s], b1.y, t2[p + 1]);
          t3[p] = q4mac(xs, b2.x, t3[p]);
          t3[p + 1] = q4mac(xs, b2.y, t3[p + 1]);
        }
      }
#pragma unroll
      for (int q = 0; q < QC; q++) {
        const int n = 4 * (q0 + q) + y;
        if (rok && n < Nc) {
#pragma clang fp contract(off)
          const double re = t1[q] - t2[q];
          const double im = (t3[q] - t1[q]) - t2[q];
          double o_re = a.alpha * re - a.alpha_i * im;
          double o_im = a.alpha * im + a.alpha_i * re;
          if (!BETA0) {
            const tq2 c = cold[BETA0 ? 0 : q];
            o_re = (a.beta * c.x - a.beta_i * c.y) + o_re;
            o_im = (a.beta * c.y + a.beta_i * c.x) + o_im;
          }
          *gmem(reinterpret_cast<tq2*>(crow + 2 * n)) = tq2{o_re, o_im};
        }
      }
    }
  }
}

template <int S2, int QC, int NW>
static hipError_t launch_q4c3_s(const TqArgs& a, int blocks, size_t lds, bool beta0,
                                hipStream_t st) {
  auto k = beta0 ? tsmm_q4c3_kernel<S2, QC, NW, true> : tsmm_q4c3_kernel<S2, QC, NW, false>;
  hipError_t e = hipFuncSetAttribute((const void*)k,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(blocks), dim3(64 * NW), lds, st, a);
  return hipGetLastError();
}

template <int S, int QC, int RB, int NW, bool CPLX>
static hipError_t launch_q4_s(const TqArgs& a, int blocks, size_t lds, bool beta0,
                              hipStream_t st) {
  auto k = beta0 ? tsmm_q4_kernel<S, QC, RB, NW, true, CPLX>
                 : tsmm_q4_kernel<S, QC, RB, NW, false, CPLX>;
  hipError_t e = hipFuncSetAttribute((const void*)k,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(blocks), dim3(64 * NW), lds, st, a);
  return hipGetLastError();
}

// Two 16-row groups per wave and pass with 8 column groups, except complex
// products with S = 32 (64 complex columns), which take one group with 16
// column groups.  Same box, K = 2^25..2^26, against tsmm_lds_kernel: complex
// M = N = 48 42.7 -> 28.5 ms (one group, 16 columns: 34.1), real
// M = N = 128 35.5 -> 21.1 ms (22.2), real 96 17.3 -> 13.9 ms (16.6); complex
// 64 56.4 -> 45.3 ms (two groups: 47.0).
template <int NW, bool CPLX>
static hipError_t launch_q4_nw(const TqArgs& a, int S, int blocks, size_t lds, bool beta0,
                               hipStream_t st) {
  if (CPLX && S == 32) return launch_q4_s<32, 16, 1, NW, CPLX>(a, blocks, lds, beta0, st);
  switch (S) {
    case 8: return launch_q4_s<8, 8, 2, NW, CPLX>(a, blocks, lds, beta0, st);
    case 12: return launch_q4_s<12, 8, 2, NW, CPLX>(a, blocks, lds, beta0, st);
    case 16: return launch_q4_s<16, 8, 2, NW, CPLX>(a, blocks, lds, beta0, st);
    case 24: return launch_q4_s<24, 8, 2, NW, CPLX>(a, blocks, lds, beta0, st);
    default: return launch_q4_s<32, 8, 2, NW, CPLX>(a, blocks, lds, beta0, st);
  }
}

// Returns SKINNY_UNSUPPORTED (and launches nothing) for shapes outside the
// kernel's range: real-view M in 29..128, N in 1..128.
skinny_status_t launch_tsmm_q4(skinny_handle_t h, size_t blockCount, int64_t M, int64_t N,
                               int64_t K, const double* A, int64_t lda, double alpha,
                               double alpha_i, const double* B, int64_t ldb, double beta,
                               double beta_i, double* C, int64_t ldc, bool cplx,
                               hipStream_t stream) {
  if (M < 29 || M > 128 || N < 1 || N > 128) return SKINNY_UNSUPPORTED;
  // complex with 16-byte aligned A and C: the three-product form (same box,
  // K = 2^26, against the real view: M = N = 64 45.5 -> 35.9 ms, 48 28.3 ->
  // 26.9, 40 27.2 -> 25.2; M = 64, N = 8 20.9 -> 17.7; M = 33, N = 64 33.5 ->
  // 27.8)
  if (cplx && (uintptr_t)A % 16 == 0 && (uintptr_t)C % 16 == 0) {
    const int Mc = (int)(M / 2), Nc = (int)(N / 2);
    const int S2 = Mc <= 32 ? 8 : Mc <= 48 ? 12 : 16;
    constexpr int QC3 = 8;
    const int QT3 = ((Nc + 3) / 4 + QC3 - 1) / QC3 * QC3;
    TqArgs c;
    c.A = A;
    c.B = B;
    c.C = C;
    c.K = K;
    c.lda = lda / 2;  // complex units
    c.ldb = ldb;
    c.ldc = ldc / 2;
    c.M = Mc;
    c.N = Nc;
    c.QT = QT3;
    c.vec = 0;  // unused: the kernel masks every load by the column bound
    c.alpha = alpha;
    c.beta = beta;
    c.alpha_i = alpha_i;
    c.beta_i = beta_i;
    const size_t lds = (size_t)3 * 16 * S2 * (QT3 + 2) * sizeof(double);
    const bool big = lds > 80 * 1024;
    int blocks;
    TuneEntry te;
    if (blockCount > 0) {
      blocks = (int)blockCount;
    } else if (tune_lookup(h, 1, TT_DC, Mc, Nc, K, &te)) {
      blocks = te.G;
    } else {
      const int per_cu =
          big ? 1 : (int)std::max<size_t>(1, std::min<size_t>(4, (160 * 1024) / lds));
      const int64_t want = (K + 16 * 4 - 1) / (16 * 4);
      blocks = (int)std::max<int64_t>(1, std::min(want, (int64_t)per_cu * h->cu_count));
    }
    const bool beta0 = beta == 0.0 && beta_i == 0.0;
    hipError_t e;
    switch (S2) {
      case 8: e = big ? launch_q4c3_s<8, QC3, 8>(c, blocks, lds, beta0, stream)
                      : launch_q4c3_s<8, QC3, 4>(c, blocks, lds, beta0, stream); break;
      case 12: e = big ? launch_q4c3_s<12, QC3, 8>(c, blocks, lds, beta0, stream)
                       : launch_q4c3_s<12, QC3, 4>(c, blocks, lds, beta0, stream); break;
      default: e = big ? launch_q4c3_s<16, QC3, 8>(c, blocks, lds, beta0, stream)
                       : launch_q4c3_s<16, QC3, 4>(c, blocks, lds, beta0, stream); break;
    }
    if (e != hipSuccess) return hip_fail(e, "tsmm (complex, 3M) launch");
    return SKINNY_OK;
  }
  const int S0 = (int)((M + 3) / 4);
  const int S = S0 <= 8 ? 8 : S0 <= 12 ? 12 : S0 <= 16 ? 16 : S0 <= 24 ? 24 : 32;
  const int QC = (cplx && S == 32) ? 16 : 8;
  const int Q = (int)((N + 3) / 4);
  const int QT = (Q + QC - 1) / QC * QC;
  TqArgs a;
  a.A = A;
  a.B = B;
  a.C = C;
  a.K = K;
  a.lda = lda;
  a.ldb = ldb;
  a.ldc = ldc;
  a.M = (int)M;
  a.N = (int)N;
  a.QT = QT;
  a.vec = (M == 4 * S && lda % 2 == 0 && (uintptr_t)A % 16 == 0) ? 1 : 0;
  a.alpha = alpha;
  a.beta = beta;
  a.alpha_i = alpha_i;
  a.beta_i = beta_i;
  const size_t lds = (size_t)16 * S * (QT + 2) * sizeof(double);
  // one workgroup holds B^ for all its waves: eight waves when B^ takes more
  // than half the LDS (one workgroup per CU), four otherwise
  const bool big = lds > 80 * 1024;
  int blocks;
  TuneEntry te;
  if (blockCount > 0) {
    blocks = (int)blockCount;
  } else if (tune_lookup(h, 1, cplx ? TT_DC : TT_DR, cplx ? M / 2 : M, cplx ? N / 2 : N, K,
                         &te)) {
    blocks = te.G;
  } else {
    const int per_cu = big ? 1 : (int)std::max<size_t>(1, std::min<size_t>(4, (160 * 1024) / lds));
    int64_t want = (K + 16 * 4 - 1) / (16 * 4);
    const int64_t cap = (int64_t)per_cu * h->cu_count;
    blocks = (int)std::max<int64_t>(1, std::min(want, cap));
  }
  const bool beta0 = beta == 0.0 && beta_i == 0.0;
  hipError_t e;
  if (big)
    e = cplx ? launch_q4_nw<8, true>(a, S, blocks, lds, beta0, stream)
             : launch_q4_nw<8, false>(a, S, blocks, lds, beta0, stream);
  else
    e = cplx ? launch_q4_nw<4, true>(a, S, blocks, lds, beta0, stream)
             : launch_q4_nw<4, false>(a, S, blocks, lds, beta0, stream);
  if (e != hipSuccess) return hip_fail(e, "tsmm (4x4x4) launch");
  return SKINNY_OK;
}

}  // namespace skinny
