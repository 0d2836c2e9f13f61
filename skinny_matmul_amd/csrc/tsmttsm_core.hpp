// Workgroup-level building blocks of TSMTTSM on the 16x16x4 matrix cores,
// shared by the one-product kernel (tsmttsm.hip) and the persistent batched
// kernel (batch.hip).
//
//  accumulate()  streams rows [0, K) of A (K x M) and B (K x N), both given
//                already offset to the first row, into per-wave MFMA
//                accumulators.  Wave w of a team of `wstride` waves takes the
//                chunks w, w + wstride, ... of CH = 4 P U rows; the one ragged
//                chunk at the end is masked row by row.
//  wg_partial()  sums the four waves of the workgroup through LDS in the fixed
//                order ((w0 + w2) + (w1 + w3)), lays the D tile out as a matrix
//                and writes the outM x outN partial (block-diagonal sums when
//                P > 1 blocks are packed into the MFMA lanes).
#pragma once

#include "mfma.hpp"

namespace skinny {

// SELF: a Gram product (B == A, one diagonal tile, MT == NT).  Only the
// subtiles t <= u are accumulated; wg_partial mirrors them into the lower
// ones.  Entry (i, j) of subtile (u, t) would be the same products in the
// same order as entry (j, i) of subtile (t, u), so the mirror is bitwise
// what the full computation gives, with MT (MT + 1) / 2 instead of MT^2
// matrix-core steps per row group (10 of 16 at 64 columns).
// PIPE: the chunk loop runs over two register sets (loads of the next chunk
// in flight during this chunk's matrix-core steps).
template <typename T, int P, int MT, int NT, int U, bool SELF, bool PIPE = false>
struct TsCore {
  using V4 = typename Mfma<T>::v4;
  // the dispatch instantiates every (MT, NT) for SELF; only square ones run
  static constexpr bool SYM = SELF && MT == NT;
  // subtile (t, u) is accumulated
  static constexpr bool live(int t, int u) { return !SYM || u >= t; }
  static constexpr int RS = 4 * P;   // rows per MFMA step
  static constexpr int CH = RS * U;  // rows per chunk
  static constexpr int DM = 16 * MT, DN = 16 * NT;
  static constexpr int NACC = MT * NT * 4;
  static constexpr int RED_ELEMS = 2 * NACC * 64;  // LDS elements needed

  static __device__ __forceinline__ void zero(V4 (&acc)[MT][NT]) {
#pragma unroll
    for (int t = 0; t < MT; t++)
#pragma unroll
      for (int u = 0; u < NT; u++) acc[t][u] = V4{0, 0, 0, 0};
  }

  // M, N: widths of the whole product (packing uses them); m0, n0: first
  // column of this tile; Mloc, Nloc: tile columns in range.  STREAM loads the
  // rows with the nontemporal hint: right when every row is read by one
  // workgroup only, wrong when several column tiles of a wide product read
  // the same rows at about the same time and would otherwise share them
  // through the caches.
  template <bool STREAM>
  static __device__ __forceinline__ T ld(const T* p) {
    if (STREAM) return ld_stream(p);
    return *gmem(p);
  }

  template <bool STREAM = true>
  static __device__ __forceinline__ void accumulate(
      const T* A, int64_t lda, const T* B, int64_t ldb, int M, int N, int m0,
      int n0, int Mloc, int Nloc, int64_t K, int64_t w0, int64_t wstride,
      V4 (&acc)[MT][NT]) {
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, lg = lane >> 4;
    // lane -> (row within step, column) of its A and B operands
    int rA = 0, cA = li, rB = 0, cB = li;
    bool vA = true, vB = true;
    if (P > 1) {
      rA = li / M;
      cA = li % M;
      vA = rA < P;
      rB = li / N;
      cB = li % N;
      vB = rB < P;
    }
    const int rowA = rA + P * lg;
    const int rowB = rB + P * lg;
    bool va[MT], vb[NT];
#pragma unroll
    for (int t = 0; t < MT; t++) va[t] = vA && (P > 1 || 16 * t + li < Mloc);
#pragma unroll
    for (int u = 0; u < NT; u++) vb[u] = vB && (P > 1 || 16 * u + li < Nloc);

    const T* Ab = A + (int64_t)rowA * lda + m0 + cA;
    const T* Bb = B + (int64_t)rowB * ldb + n0 + cB;
    const int64_t stepA = (int64_t)RS * lda;
    const int64_t stepB = (int64_t)RS * ldb;
    const int64_t nfull = K / CH;

    auto load = [&](T (&av)[U][MT], T (&bv)[U][NT], int64_t c) {
      const T* pa = Ab + c * CH * lda;
      const T* pb = Bb + c * CH * ldb;
#pragma unroll
      for (int s = 0; s < U; s++) {
#pragma unroll
        for (int t = 0; t < MT; t++)
          av[s][t] = va[t] ? ld<STREAM>(&pa[s * stepA + 16 * t]) : T(0);
        if (!SELF) {
#pragma unroll
          for (int u = 0; u < NT; u++)
            bv[s][u] = vb[u] ? ld<STREAM>(&pb[s * stepB + 16 * u]) : T(0);
        }
      }
    };
    auto mac = [&](const T (&av)[U][MT], const T (&bv)[U][NT]) {
#pragma unroll
      for (int s = 0; s < U; s++)
#pragma unroll
        for (int t = 0; t < MT; t++)
#pragma unroll
          for (int u = 0; u < NT; u++)
            if (live(t, u))
              acc[t][u] = Mfma<T>::mac(av[s][t], SELF ? av[s][u] : bv[s][u],
                                       acc[t][u]);
    };
    if constexpr (PIPE) {
      // two register sets: the loads of the wave's next chunk (index clamped
      // to the last full chunk, so they are issued unconditionally) are in
      // flight while the matrix cores take the current one
      T a0[U][MT], b0[U][NT], a1[U][MT], b1[U][NT];
      if (w0 < nfull) {
        load(a0, b0, w0);
        for (int64_t c = w0; c < nfull; c += 2 * wstride) {
          load(a1, b1, c + wstride < nfull ? c + wstride : nfull - 1);
          mac(a0, b0);
          if (c + wstride >= nfull) break;
          load(a0, b0, c + 2 * wstride < nfull ? c + 2 * wstride : nfull - 1);
          mac(a1, b1);
        }
      }
    } else {
      for (int64_t c = w0; c < nfull; c += wstride) {
        T av[U][MT], bv[U][NT];
        load(av, bv, c);
        mac(av, bv);
      }
    }

    // ragged tail: the one partial chunk, taken by the wave whose turn it is
    if (nfull * CH < K && (nfull % wstride) == w0) {
      const int64_t r0 = nfull * CH;
#pragma unroll
      for (int s = 0; s < U; s++) {
        const bool okA = r0 + s * RS + rowA < K;
        const bool okB = r0 + s * RS + rowB < K;
        T av[MT], bv[NT];
#pragma unroll
        for (int t = 0; t < MT; t++)
          av[t] = (va[t] && okA) ? ld<STREAM>(&Ab[(r0 + s * RS) * lda + 16 * t]) : T(0);
#pragma unroll
        for (int u = 0; u < NT; u++) {
          if (SELF)
            bv[u] = av[u];
          else
            bv[u] = (vb[u] && okB) ? ld<STREAM>(&Bb[(r0 + s * RS) * ldb + 16 * u]) : T(0);
        }
#pragma unroll
        for (int t = 0; t < MT; t++)
#pragma unroll
          for (int u = 0; u < NT; u++)
            if (live(t, u)) acc[t][u] = Mfma<T>::mac(av[t], bv[u], acc[t][u]);
      }
    }
  }

  // Workgroup sum of the four waves' accumulators, written as the outM x outN
  // partial `out` (entry e = (e % outM, e / outM)) with agent-scope stores
  // (st_agent: another workgroup of the same launch may sum it).  `red` is LDS of
  // RED_ELEMS elements.  Ends with a workgroup barrier after the writes.
  static __device__ __forceinline__ void wg_partial(V4 (&acc)[MT][NT], T* red,
                                                    T* out, int M, int N,
                                                    int outM, int outN) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    using LT = __attribute__((address_space(3))) T;
    LT* red0 = lmem(red);
    LT* red1 = red0 + NACC * 64;
    auto slot = [&](LT* base, int t, int u, int v) -> LT& {
      return base[((t * NT + u) * 4 + v) * 64 + lane];
    };
    if (wave >= 2) {
#pragma unroll
      for (int t = 0; t < MT; t++)
#pragma unroll
        for (int u = 0; u < NT; u++)
#pragma unroll
          for (int v = 0; v < 4; v++)
            if (live(t, u)) slot(wave == 2 ? red0 : red1, t, u, v) = acc[t][u][v];
    }
    __syncthreads();
    if (wave < 2) {
#pragma unroll
      for (int t = 0; t < MT; t++)
#pragma unroll
        for (int u = 0; u < NT; u++)
#pragma unroll
          for (int v = 0; v < 4; v++)
            if (live(t, u)) acc[t][u][v] += slot(wave == 0 ? red0 : red1, t, u, v);
    }
    __syncthreads();
    if (wave == 1) {
#pragma unroll
      for (int t = 0; t < MT; t++)
#pragma unroll
        for (int u = 0; u < NT; u++)
#pragma unroll
          for (int v = 0; v < 4; v++)
            if (live(t, u)) slot(red0, t, u, v) = acc[t][u][v];
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
      for (int t = 0; t < MT; t++)
#pragma unroll
        for (int u = 0; u < NT; u++)
#pragma unroll
          for (int v = 0; v < 4; v++)
            if (live(t, u)) acc[t][u][v] += slot(red0, t, u, v);
    }
    __syncthreads();
    // D tile as a DM x DN row-major matrix (MFMA D layout: lane l, element v
    // holds D[drow(l / 16, v)][l % 16])
    LT* Dm = red1;
    if (wave == 0) {
#pragma unroll
      for (int t = 0; t < MT; t++)
#pragma unroll
        for (int u = 0; u < NT; u++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            if (!live(t, u)) continue;
            Dm[(16 * t + Mfma<T>::drow(lg, v)) * DN + 16 * u + li] = acc[t][u][v];
            // SELF: the transposed copy fills subtile (u, t)
            if (SYM && u > t)
              Dm[(16 * u + li) * DN + 16 * t + Mfma<T>::drow(lg, v)] = acc[t][u][v];
          }
    }
    __syncthreads();
    const int PS = outM * outN;
    for (int e = threadIdx.x; e < PS; e += blockDim.x) {
      const int ml = e % outM, nl = e / outM;
      T s;
      if (P > 1) {
        s = T(0);
#pragma unroll
        for (int r = 0; r < P; r++) s += Dm[(r * M + ml) * DN + r * N + nl];
      } else {
        s = Dm[ml * DN + nl];
      }
      st_agent(out + e, s);
    }
    __syncthreads();
  }
};

}  // namespace skinny
