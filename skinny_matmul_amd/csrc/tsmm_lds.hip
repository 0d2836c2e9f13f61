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
