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
//    address in all lanes, so it is fetched with scalar loads into SGPRs and
//    broadcast for free as the scalar operand of v_fma (the reference stages B
//    through shared memory, fix3.cuh:10-18).
//  * Row loads and stores use 16-byte vector accesses when the leading
//    dimension and base pointer allow it.
//  * The row length is rounded up to a compile-time class MC in
//    {1, 2, 4, 8, 16, 32, 64} and masked with uniform branches; N is a runtime
//    loop unrolled by four, so one instantiation serves a range of widths.
#include "internal.hpp"

namespace skinny {

template <typename T>
struct TmArgs {
  const T* A;
  const T* B;
  T* C;
  int64_t K, lda, ldb, ldc;
  int M, N;
  T alpha, beta;
};

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

// rows per thread per iteration (independent rows in flight)
template <int MC>
struct TmRows {
  static constexpr int value = MC <= 2 ? 4 : (MC <= 8 ? 2 : 1);
};

template <typename T, int MC, bool BETA0, bool VEC>
__global__ void __launch_bounds__(256) tsmm_kernel(TmArgs<T> a) {
  constexpr int R = TmRows<MC>::value;
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
      if (VEC && (MC % 2 == 0) && sizeof(T) == 8) {
#pragma unroll
        for (int m = 0; m < MC; m += 2) {
          if (m < M) {
            if (ok && m + 1 < M) {
              double2 v = *reinterpret_cast<const double2*>(ar + m);
              av[r][m] = (T)v.x;
              av[r][m + 1] = (T)v.y;
            } else {
              av[r][m] = ok ? ar[m] : T(0);
              av[r][m + 1] = T(0);
            }
          } else {
            av[r][m] = T(0);
            av[r][m + 1] = T(0);
          }
        }
      } else {
#pragma unroll
        for (int m = 0; m < MC; m++) av[r][m] = (m < M && ok) ? ar[m] : T(0);
      }
    }
    // every A value of these rows is now in registers: in-place safe
    int n = 0;
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
          T x0 = b0[m], x1 = b1[m];
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
          if (VEC && sizeof(T) == 8) {
            double2 v;
            v.x = (double)o0;
            v.y = (double)o1;
            *reinterpret_cast<double2*>(cr) = v;
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
          T x0 = b0[m];
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

int tuned_blocks_tsmm(skinny_handle_t h, int64_t M, int64_t N, int64_t K,
                      int elem_bytes) {
  (void)M;
  (void)N;
  (void)elem_bytes;
  int64_t want = (K + 255) / 256;
  int64_t cap = 8LL * h->cu_count;
  if (want > cap) want = cap;
  if (want < 1) want = 1;
  return (int)want;
}

template <typename T, int MC, bool BETA0, bool VEC>
static void launch_tm(const TmArgs<T>& a, int blocks, hipStream_t s) {
  hipLaunchKernelGGL((tsmm_kernel<T, MC, BETA0, VEC>), dim3(blocks), dim3(256),
                     0, s, a);
}

template <typename T, int MC>
static void launch_mc(const TmArgs<T>& a, int blocks, bool beta0, bool vec,
                      hipStream_t s) {
  if (beta0) {
    if (vec)
      launch_tm<T, MC, true, true>(a, blocks, s);
    else
      launch_tm<T, MC, true, false>(a, blocks, s);
  } else {
    if (vec)
      launch_tm<T, MC, false, true>(a, blocks, s);
    else
      launch_tm<T, MC, false, false>(a, blocks, s);
  }
}

template <typename T>
skinny_status_t launch_tsmm_real(skinny_handle_t h, size_t blockCount,
                                 int64_t M, int64_t N, int64_t K, const T* A,
                                 int64_t lda, T alpha, const T* B, int64_t ldb,
                                 T beta, T* C, int64_t ldc, hipStream_t stream) {
  if (M > 64)
    return fail(SKINNY_UNSUPPORTED, "tsmm: M > 64 not supported");
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
  int blocks = blockCount > 0 ? (int)blockCount
                              : tuned_blocks_tsmm(h, M, N, K, (int)sizeof(T));
  const bool beta0 = (beta == T(0));
  // 16-byte row accesses need 16-byte aligned rows for both A and C
  const bool vec = sizeof(T) == 8 && (lda % 2 == 0) && (ldc % 2 == 0) &&
                   ((uintptr_t)A % 16 == 0) && ((uintptr_t)C % 16 == 0);
  int MC = M <= 1 ? 1 : M <= 2 ? 2 : M <= 4 ? 4 : M <= 8 ? 8 : M <= 16 ? 16
           : M <= 32 ? 32 : 64;
  switch (MC) {
    case 1: launch_mc<T, 1>(a, blocks, beta0, false, stream); break;
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

template skinny_status_t launch_tsmm_real<double>(
    skinny_handle_t, size_t, int64_t, int64_t, int64_t, const double*, int64_t,
    double, const double*, int64_t, double, double*, int64_t, hipStream_t);
template skinny_status_t launch_tsmm_real<float>(
    skinny_handle_t, size_t, int64_t, int64_t, int64_t, const float*, int64_t,
    float, const float*, int64_t, float, float*, int64_t, hipStream_t);

}  // namespace skinny
