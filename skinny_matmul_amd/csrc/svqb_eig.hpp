// Device eigensolve of SVQB (svqb/svqb.cu:99-119) for one block by one
// workgroup of 256 threads, shared by the separate-kernel sweep (svqb.hip)
// and the fused grid-barrier sweep (svqb_fused.hip).
//
// Input: the M x M Gram S in LDS (S[i * LD + j], LD = kSvqbMaxM + 1).
//   d_i = 1 / sqrt(S_ii)  (0 if S_ii <= 1e-15)  (svqb.cu:99-105)
//   S <- D S D                                  (svqb.cu:106-110)
//   S = U L U^T by cyclic Jacobi                (svqb.cu:113-114 use magma_dsyevd)
//   T = D U L^-1/2,  L clamped at 1e-15         (svqb.cu:115-119)
// Output: T (column-major, ld M) in global memory.
#pragma once

#include <hip/hip_runtime.h>

namespace skinny {

constexpr int kSvqbMaxM = 64;
constexpr int kEigLD = kSvqbMaxM + 1;
// LDS of one eigensolve: S, V, dinv, rotation cosines and sines, a 256-entry
// reduction buffer (doubles); pair indices and the convergence flag (ints)
constexpr int kEigLdsDoubles = 2 * kSvqbMaxM * kEigLD + kSvqbMaxM + kSvqbMaxM + 256;
constexpr int kEigLdsInts = kSvqbMaxM + 1;

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

// lds: kEigLdsDoubles doubles whose first M x LD entries hold S; ilds:
// kEigLdsInts ints.  Ends after the T stores (no trailing barrier).
__device__ __forceinline__ void svqb_eig_lds(int M, int max_sweeps, double* lds,
                                             int* ilds, double* Tg) {
  const int n = (M + 1) & ~1;  // even size with a dummy index when M is odd
  const int npairs = n / 2;
  const int tid = threadIdx.x;
  constexpr int LD = kEigLD;
  double* S = lds;
  double* V = S + kSvqbMaxM * LD;
  double* dinv = V + kSvqbMaxM * LD;
  double* rc = dinv + kSvqbMaxM;
  double* rs = rc + kSvqbMaxM / 2;
  double* red = rs + kSvqbMaxM / 2;
  int* rp = ilds;
  int* rq = rp + kSvqbMaxM / 2;
  int* done = rq + kSvqbMaxM / 2;

  for (int e = tid; e < M * M; e += 256) {
    const int i = e % M, j = e / M;
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

  for (int sweep = 0; sweep < max_sweeps; sweep++) {
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
    if (tid == 0) *done = (offsum <= 1e-30 * red[0]) ? 1 : 0;
    __syncthreads();
    if (*done) break;

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

}  // namespace skinny
