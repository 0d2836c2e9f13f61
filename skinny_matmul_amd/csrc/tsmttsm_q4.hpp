// FP64 TSMTTSM on the 4x4x4 (four-block) matrix-core instruction, for 64 x 64
// tiles of C.
//
// v_mfma_f64_4x4x4f64 computes four independent 4 x 4 x 4 products, one per
// block.  Lane l = 16 X + 4 b + y belongs to block b; it supplies A[y][X] and
// B[X][y] of its block and receives D[X][y] (layout measured on gfx950 by a
// probe that fed powers of two through the operands).  On MI355X the
// instruction sustains about 75 TFLOP/s over independent accumulators, where
// v_mfma_f64_16x16x4f64 sustains about 49 TFLOP/s at the same clock, so the
// tiles whose matrix-core time exceeds their HBM time (64 x 64 and wider) use
// it.
//
// Mapping for C = A^T B (the contraction index is the row k of the row-major
// inputs):
//  * a wave takes rows in groups of four; lane (X, b, y) supplies row X of
//    the group for both operands;
//  * block b = bm + 2 bn splits the tile into halves of 8 columns: A operand
//    register p = 2 s + e holds A(row, 16 s + 8 bm + 2 y + e), B operand
//    register q = 2 t + f holds B(row, 16 t + 8 bn + 2 y + f);
//  * accumulator (p, q) of lane (X, b, y) is C(16 s + 8 bm + 2 X + e,
//    16 t + 8 bn + 2 y + f).
// A register pair (e = 0, 1) is one 16-byte load, and one load instruction
// covers columns 16 s .. 16 s + 15 of each of the four rows: a whole 128-byte
// line per row (each lane's data is needed by the two blocks of its bm, so
// every address is requested by two lanes).  Per group of four rows a lane
// issues 8 loads and 64 matrix-core steps into 64 independent accumulators.
#pragma once

#include "mfma.hpp"

namespace skinny {

typedef double dq2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double mfma_q4(double a, double b, double c) {
  return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

// U: groups of four rows per chunk (loads in flight: 16 U to 32 U per lane).
// FULL: the tile's 64 x 64 columns are all in range (no column masks).
// ALIGNED: even leading dimensions and 16-byte aligned bases (pair loads).
template <int U, bool FULL, bool ALIGNED>
struct TsQ4 {
  static constexpr int CH = 4 * U;  // rows per chunk
  static constexpr int NACC = 64;   // accumulators per lane
  static constexpr int RED_ELEMS = 2 * NACC * 64;
  static constexpr int DM = 64, DN = 64;

  // Two consecutive elements, each loaded only where its flag is set (MASK),
  // or unconditionally (full tiles in the main loop).
  template <bool STREAM, bool MASK = !FULL>
  static __device__ __forceinline__ void ld_pair(const double* p, bool v0, bool v1,
                                                 double& x0, double& x1) {
    if (ALIGNED && (!MASK || (v0 && v1))) {
      dq2 v;
      if (STREAM)
        v = __builtin_nontemporal_load(gmem(reinterpret_cast<const dq2*>(p)));
      else
        v = *gmem(reinterpret_cast<const dq2*>(p));
      x0 = v.x;
      x1 = v.y;
      return;
    }
    x0 = (!MASK || v0) ? (STREAM ? ld_stream(p) : *gmem(p)) : 0.0;
    x1 = (!MASK || v1) ? (STREAM ? ld_stream(p + 1) : *gmem(p + 1)) : 0.0;
  }

  static __device__ __forceinline__ void zero(double (&acc)[NACC]) {
#pragma unroll
    for (int i = 0; i < NACC; i++) acc[i] = 0.0;
  }

  // Rows [0, K) of A and B (given at the first row), columns m0.. and n0.. of
  // this tile (Mloc, Nloc of them in range); wave w0 of a team of wstride
  // takes the chunks w0, w0 + wstride, ... of CH rows, the ragged last chunk
  // masked row by row.
  template <bool STREAM>
  static __device__ __forceinline__ void accumulate(const double* A, int64_t lda,
                                                    const double* B, int64_t ldb,
                                                    int m0, int n0, int Mloc, int Nloc,
                                                    int64_t K, int64_t w0, int64_t wstride,
                                                    double (&acc)[NACC]) {
    const int lane = threadIdx.x & 63;
    const int X = lane >> 4, blk = (lane >> 2) & 3, y = lane & 3;
    const int ca = 8 * (blk & 1) + 2 * y, cb = 8 * (blk >> 1) + 2 * y;
    bool va[8], vb[8];
#pragma unroll
    for (int p = 0; p < 8; p++) {
      va[p] = FULL || (16 * (p >> 1) + ca + (p & 1) < Mloc);
      vb[p] = FULL || (16 * (p >> 1) + cb + (p & 1) < Nloc);
    }
    const double* Ab = A + (int64_t)X * lda + m0 + ca;
    const double* Bb = B + (int64_t)X * ldb + n0 + cb;
    const int64_t nfull = K / CH;

    // software pipeline over two register sets: the loads of a wave's next
    // chunk are in flight while the matrix cores take the current one
    double a0[U][8], b0[U][8], a1[U][8], b1[U][8];
    // chunk offsets are wave-uniform: scalar arithmetic, added to the fixed
    // per-lane bases (no vector temporaries next to loads in flight)
    const int64_t sa = CH * lda, sb = CH * ldb;
    auto load = [&](double (&av)[U][8], double (&bv)[U][8], int64_t c) {
      const double* pa = Ab + c * sa;
      const double* pb = Bb + c * sb;
#pragma unroll
      for (int g = 0; g < U; g++) {
#pragma unroll
        for (int s = 0; s < 4; s++)
          ld_pair<STREAM>(pa + 4 * g * lda + 16 * s, va[2 * s], va[2 * s + 1], av[g][2 * s],
                          av[g][2 * s + 1]);
#pragma unroll
        for (int s = 0; s < 4; s++)
          ld_pair<STREAM>(pb + 4 * g * ldb + 16 * s, vb[2 * s], vb[2 * s + 1], bv[g][2 * s],
                          bv[g][2 * s + 1]);
      }
    };
    auto mac = [&](const double (&av)[U][8], const double (&bv)[U][8]) {
#pragma unroll
      for (int g = 0; g < U; g++)
#pragma unroll
        for (int p = 0; p < 8; p++)
#pragma unroll
          for (int q = 0; q < 8; q++)
            acc[8 * p + q] = mfma_q4(av[g][p], bv[g][q], acc[8 * p + q]);
    };
    // The prefetch is issued on every path (past the last chunk it reloads
    // the last one and its values go unused), so the number of loads in
    // flight is the same everywhere and the wait before each set's
    // matrix-core steps covers only that set; a prefetch under a branch made
    // the compiler wait for both sets.
    // The empty asm with a memory clobber keeps each prefetch ahead of the
    // other set's matrix-core steps: its values are used only past the loop
    // exit test, and without the fence the loads were sunk below the steps.
    int64_t c = __builtin_amdgcn_readfirstlane((int)w0);  // w0 < 2^31
    if (c < nfull) {
      const int64_t last = nfull - 1;
      load(a0, b0, c);
      for (;;) {
        const int64_t c1 = c + wstride;
        load(a1, b1, c1 < nfull ? c1 : last);
        asm volatile("" ::: "memory");
        mac(a0, b0);
        if (c1 >= nfull) break;
        const int64_t c2 = c1 + wstride;
        load(a0, b0, c2 < nfull ? c2 : last);
        asm volatile("" ::: "memory");
        mac(a1, b1);
        if (c2 >= nfull) break;
        c = c2;
      }
    }

    // ragged tail: the one partial chunk, taken by the wave whose turn it is
    if (nfull * CH < K && (nfull % wstride) == w0) {
      const int64_t r0 = nfull * CH;
#pragma unroll
      for (int g = 0; g < U; g++) {
        const bool ok = r0 + 4 * g + X < K;
        double av[8], bv[8];
#pragma unroll
        for (int s = 0; s < 4; s++) {
          ld_pair<STREAM, true>(Ab + (r0 + 4 * g) * lda + 16 * s, ok && va[2 * s],
                                ok && va[2 * s + 1], av[2 * s], av[2 * s + 1]);
          ld_pair<STREAM, true>(Bb + (r0 + 4 * g) * ldb + 16 * s, ok && vb[2 * s],
                                ok && vb[2 * s + 1], bv[2 * s], bv[2 * s + 1]);
        }
#pragma unroll
        for (int p = 0; p < 8; p++)
#pragma unroll
          for (int q = 0; q < 8; q++)
            acc[8 * p + q] = mfma_q4(av[p], bv[q], acc[8 * p + q]);
      }
    }
  }

  // Workgroup sum of the four waves ((w0 + w2) + (w1 + w3), fixed order),
  // written as the outM x outN partial `out` (entry e = (e % outM, e / outM)).
  // `red` is LDS of RED_ELEMS elements.  Ends with a workgroup barrier.
  static __device__ __forceinline__ void wg_partial(double (&acc)[NACC], double* red,
                                                    double* out, int outM, int outN) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    using LT = __attribute__((address_space(3))) double;
    LT* red0 = lmem(red);
    LT* red1 = red0 + NACC * 64;
    if (wave >= 2) {
      LT* dst = wave == 2 ? red0 : red1;
#pragma unroll
      for (int i = 0; i < NACC; i++) dst[i * 64 + lane] = acc[i];
    }
    __syncthreads();
    if (wave < 2) {
      const LT* src = wave == 0 ? red0 : red1;
#pragma unroll
      for (int i = 0; i < NACC; i++) acc[i] += src[i * 64 + lane];
    }
    __syncthreads();
    if (wave == 1) {
#pragma unroll
      for (int i = 0; i < NACC; i++) red0[i * 64 + lane] = acc[i];
    }
    __syncthreads();
    // the tile as a 64 x 64 row-major matrix (entry (m, n) at 64 m + n)
    LT* Dm = red1;
    if (wave == 0) {
      const int X = lane >> 4, blk = (lane >> 2) & 3, y = lane & 3;
      const int mb = 8 * (blk & 1) + 2 * X, nb = 8 * (blk >> 1) + 2 * y;
#pragma unroll
      for (int i = 0; i < NACC; i++) {
        const int p = i >> 3, q = i & 7;
        const int m = 16 * (p >> 1) + mb + (p & 1);
        const int n = 16 * (q >> 1) + nb + (q & 1);
        Dm[m * 64 + n] = acc[i] + red0[i * 64 + lane];
      }
    }
    __syncthreads();
    const int PS = outM * outN;
    for (int e = threadIdx.x; e < PS; e += blockDim.x) {
      const int ml = e % outM, nl = e / outM;
      st_agent(out + e, Dm[ml * 64 + nl]);
    }
    __syncthreads();
  }
};

}  // namespace skinny
