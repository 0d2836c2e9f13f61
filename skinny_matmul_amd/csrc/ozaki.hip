// Ozaki-split FP64 TSMTTSM on the INT8 matrix cores (no reference
// counterpart; an alternative to the FP64 kernel for wide blocks).
//
// C = alpha A^T B + beta C with A (K x M), B (K x N) FP64.  The rows are taken
// in chunks of 32 (one MFMA k-depth):
//  * Each column c has a running scale exponent E_c: every |x| since the last
//    flush is below 2^E_c (after a flush the scale restarts from the next
//    chunk's largest |x|).  An element x is rounded to the integer
//    q = rint(x 2^(8S - 2 - E)), |q| <= 2^(8S - 2), and q is written in S
//    signed bytes (limbs), q = sum_s d_s 2^(8 (S - 1 - s)), the lower ones in
//    [-128, 127] and the top one in [-64, 64]: x = 2^(E - 8S + 2) (q + r),
//    |r| <= 1/2.  This is the error-free splitting of Ozaki et al. (2012)
//    with signed-digit limbs: the sign needs no separate handling, and every
//    limb carries 8 bits.
//  * For each limb-sum level g = s + t (limb 0 the most significant;
//    g < 3 for 3 limbs, g <= S for 4 and 5 limbs, the other pairs dropped as
//    in the truncated Ozaki scheme) the products sum_k d_s(k, m) d'_t(k, n)
//    accumulate exactly in INT32 through v_mfma_i32_32x32x32_i8, one
//    accumulator per level.
//  * The INT32 accumulators are converted to FP64 ("flushed") only when a
//    column's magnitude grows past its scale, when a chunk's largest limb
//    integer falls 9 or more bits below it (the chunk after an outlier), or
//    before INT32 could overflow (every 16384 rows); the flush forms
//      V = sum_{g = L - 1 .. 0} acc_g 2^-8g   (in this order, FP64),
//      C_acc(m, n) += V 2^(E_m + E_n - 12)
//    (x y = 2^(E_m + E_n - 2 (8S - 2)) sum_(s,t) d_s d'_t 2^(8 (2S - 2 - s - t))).
//    For data of steady magnitude the conversion cost is amortised over
//    thousands of rows, so the INT8 units and HBM set the pace.
//  * Workgroups (4 waves, the level accumulators in the matrix-core
//    registers) own a tile of C and a grid-stride set of chunks; their FP64
//    partials are summed by the same two-level last-arrival reduction as the
//    FP64 kernel.  With 3 limbs a tile is 128 x 128 (a 64 x 64 region per
//    wave, 3 levels x 64 registers); with 4 and 5 limbs it is 128 x 64 (a
//    64 x 32 region per wave, so 5 or 6 levels take 160 or 192 registers
//    instead of spilling 320 or 384), which reads A once per 64 columns of
//    B.  With blockCount == 1 the whole computation is one fixed sequence,
//    restated bit for bit on the host by the oracle.
//  * Per chunk the loop is software-pipelined: the next chunks' loads are in
//    flight while this chunk is split (VALU, packed 4 rows to a word) into
//    one LDS buffer and the previous chunk's limbs in the other buffer are
//    multiplied, the two interleaved instruction by instruction.  Two chunks
//    are in flight (prefetch distance 2).
//  * Gram products with 3 limbs (A == B, lda == ldb, M == N) compute only
//    the tiles on and above the diagonal, writing each off-diagonal tile to
//    C twice (C(m, n) and C(n, m)), and read A once: the diagonal tiles run
//    with 64-row chunks whose two 32-row halves take the places of A and B
//    in the staging (one scale per column, shared by both operands).  The
//    loads in flight per workgroup stay the same, so such a tile streams
//    half the bytes of a general one in about half the time.  With 4 and 5
//    limbs a Gram product runs as a general one.
#include <cmath>
#include <utility>

#include "internal.hpp"
#include "team_sum.hpp"

namespace skinny {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef unsigned v2u __attribute__((ext_vector_type(2)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));

// Raw buffer resources: the per-access constant parts of an address go into
// the scalar offset, so unrolled loops over rows and accumulator entries need
// one vector address register instead of one 64-bit pointer per access.
// Word 3 = 0x00020000 is the gfx9 raw-buffer format (32-bit data, no swizzle).
// Loads whose offset (vector offset plus scalar offset) reaches num_records
// return 0 without touching memory (checked on gfx950 for the scalar offset).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t oz_rsrc(const void* p,
                                                          int64_t bytes) {
  const int n = bytes > 0x7fffffff ? 0x7fffffff : (int)bytes;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, n,
                                           0x00020000);
}

__device__ __forceinline__ double oz_ld(__amdgpu_buffer_rsrc_t r, int voff,
                                        int soff) {
  return __builtin_bit_cast(double,
                            __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// Input stream: the same load with the nontemporal hint (cache-policy bit 1,
// `nt` on gfx950), since every row of A and B is read once.
__device__ __forceinline__ double oz_ld_stream(__amdgpu_buffer_rsrc_t r,
                                               int voff, int soff) {
  return __builtin_bit_cast(double,
                            __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 2));
}

__device__ __forceinline__ void oz_st(double v, __amdgpu_buffer_rsrc_t r,
                                      int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r, voff,
                                        soff, 0);
}

// 32x32x32 INT8 MFMA: lane l supplies 16 k-bytes of A row l % 32 (k half
// l / 32) and of B column l % 32; element r of the result in lane l is
// D[i8_drow(l, r)][l % 32].
__device__ __forceinline__ int i8_drow(int lane, int r) {
  return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}

constexpr int kOzTile = 128;
constexpr int kOzRows = 32;  // rows per thread and chunk (rows per chunk: 32, Gram diagonal 64)
constexpr int kOzThreads = 256;
constexpr int kOzFlushRows = 16384;
constexpr int kOzEmin = -4000;

struct OzArgs {
  const double* A;
  const double* B;
  double* C;
  int64_t K, lda, ldb, ldc;
  int M, N;
  double alpha, beta;
  double* part1;
  double* part2;
  int* stage;  // per wave: (S + 1) * 1024 * TT INT32 accumulator staging
  unsigned* cnt;
  int G, GS, ngroups, tilesM, tilesN;
  int ntiles;  // tiles of this launch
  int mode;    // tile set: OZ_ALL, OZ_DIAG or OZ_UPPER
};

// The tile sets a launch covers: all tilesM x tilesN tiles (column-major
// numbering), the diagonal tiles of a Gram product, or its tiles above the
// diagonal (column by column), each also written transposed.
enum { OZ_ALL = 0, OZ_DIAG = 1, OZ_UPPER = 2 };

__device__ __forceinline__ void oz_tile_of(const OzArgs& a, int tile, int& tm,
                                           int& tn) {
  if (a.mode == OZ_ALL) {
    tm = tile % a.tilesM;
    tn = tile / a.tilesM;
  } else if (a.mode == OZ_DIAG) {
    tm = tn = tile;
  } else {  // column tn holds tiles tm = 0 .. tn - 1
    tn = 1;
    while (tile >= tn) {
      tile -= tn;
      tn++;
    }
    tm = tile;
  }
}

// Empty asm that requires the accumulators in the matrix-core (AGPR)
// registers: placed at the top of the chunk loop it keeps the register
// allocator from carrying them around the loop as VGPR copies (which the
// flush path otherwise provokes, spilling them every iteration).
template <int L, int TT>
__device__ __forceinline__ void oz_pin(v16i (&acc)[L][TT]) {
#pragma unroll
  for (int g = 0; g < L; g++)
#pragma unroll
    for (int tt = 0; tt < TT; tt++) asm volatile("" : "+a"(acc[g][tt]));
}

// Convert the INT32 level accumulators to FP64 and add them to this wave's
// region of the workgroup's partial tile in global memory
// (entry (m, n) at n * 128 + m, owned by exactly one lane), then clear them.
// Accumulator tt holds the 32 x 32 block (tt / NSUB, tt % NSUB) of the
// region.  The
// accumulators go straight from the matrix-core registers to the wave's INT32
// staging area and are converted from there one entry at a time, so the
// conversion holds no copy of them.  `part` is the resource of the region's
// first entry, `lofs` the lane's byte offset in it.
template <int L, int TT, int NSUB>
__device__ __forceinline__ void oz_flush(v16i (&acc)[L][TT],
                                         __amdgpu_buffer_rsrc_t stage,
                                         __amdgpu_buffer_rsrc_t part, int lofs,
                                         const int* ea_w, const int* eb_w,
                                         int lane) {
#pragma unroll
  for (int g = 0; g < L; g++)
#pragma unroll
    for (int tt = 0; tt < TT; tt++) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const v4i w = {acc[g][tt][4 * q], acc[g][tt][4 * q + 1],
                       acc[g][tt][4 * q + 2], acc[g][tt][4 * q + 3]};
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u, w), stage,
                                               16 * lane, 1024 * ((g * TT + tt) * 4 + q), 0);
      }
#pragma unroll
      for (int r = 0; r < 16; r++) acc[g][tt][r] = 0;
    }
  // NB entries at a time: their level and partial loads are issued together
  // (one entry at a time, each conversion waited for its own loads)
  constexpr int NB = 8;
#pragma unroll 1
  for (int e0 = 0; e0 < 16 * TT; e0 += NB) {
    int lv[NB][L];
    double pv[NB];
    int so[NB];
    bool live[NB];
#pragma unroll
    for (int i = 0; i < NB; i++) {
      const int e = e0 + i;
      const int tt = e >> 4, r = e & 15;
      const int dm = 32 * (tt / NSUB) + i8_drow(lane, r);
      const int dn = 32 * (tt % NSUB) + (lane & 31);
      live[i] = ea_w[dm] != kOzEmin && eb_w[dn] != kOzEmin;
      const int lo = 16 * lane + 4 * (r & 3);
#pragma unroll
      for (int g = 0; g < L; g++)
        lv[i][g] = (int)__builtin_amdgcn_raw_buffer_load_b32(
            stage, lo, 1024 * ((g * TT + tt) * 4 + (r >> 2)), 0);
      so[i] = 8 * (32 * (tt % NSUB) * kOzTile + 32 * (tt / NSUB) + 8 * (r >> 2) + (r & 3));
      pv[i] = oz_ld(part, lofs, so[i]);
    }
#pragma unroll
    for (int i = 0; i < NB; i++) {
      const int e = e0 + i;
      const int tt = e >> 4, r = e & 15;
      const int dm = 32 * (tt / NSUB) + i8_drow(lane, r);
      const int dn = 32 * (tt % NSUB) + (lane & 31);
      if (live[i]) {
        double v = ldexp((double)lv[i][L - 1], -8 * (L - 1));
#pragma unroll
        for (int g = L - 2; g >= 0; g--) v = v + ldexp((double)lv[i][g], -8 * g);
        oz_st(pv[i] + ldexp(v, ea_w[dm] + eb_w[dn] - 12), part, lofs, so[i]);
      }
    }
  }
}

// Level accumulators per limb count: products d_s d'_t with
// g = s + t < oz_levels<S>() are kept (limb 0 the most significant).  3
// limbs keep 3 levels (6 products per 32 x 32 block and chunk), 4 limbs 5
// and 5 limbs 6.  For 3 limbs the dropped levels 3 and 4 are at most 2^-21
// and 2^-30 of the product of the two column scales, the order of the
// rounding of the two 23-bit limb integers (2^-22).
template <int S>
constexpr int oz_levels() {
  return S == 3 ? 3 : S + 1;
}

// The split of one column against its scale 2^E (every |x| < 2^E).  x is
// rounded to the integer q = rint(x 2^(8S - 2 - E)), |q| <= 2^(8S - 2), and q
// is written in S signed bytes, q = sum_s d_s 2^(8 (S - 1 - s)):
// d_(S-1-p) is byte p of q + sum_(i < p) 2^(8 i + 7) in two's complement, so
// the lower digits lie in [-128, 127] and the top one in [-64, 64].
// q comes from one FMA: r = fma(x, 2^(8S - 2 - E - k), 1.5 2^(52 - k)) rounds
// x 2^(8S - 2 - E) + 1.5 2^52 (scaled by 2^-k) to nearest even once, and for
// |q| < 2^51 the significand of r holds 2^51 + q, whose low 40 bits are q in
// two's complement.  The product is exact inside the FMA, so q is rint of
// the exact t.  k = 128 for E < -900 keeps the scale factor finite (1.5
// 2^(52 - k) stays normal, so the rounding happens at the same bit).
template <int S>
struct OzScale {
  double sc, mg;
  int E;
  __device__ __forceinline__ explicit OzScale(int e) : E(e) {
    const int k = e < -900 ? 128 : 0;
    sc = ldexp(1.0, 8 * S - 2 - e - k);
    mg = ldexp(1.5, 52 - k);
  }
  __device__ __forceinline__ double raw(double x) const {
    return __builtin_fma(x, sc, mg);
  }
  // q of a magnitude a (0 <= a < 2^E, so q <= 2^(8S - 2) < 2^39)
  __device__ __forceinline__ uint64_t q(double a) const {
    const double r = raw(a);
    const uint64_t lo = (unsigned)__double2loint(r);
    if constexpr (S <= 4)
      return lo;
    else
      return lo | ((uint64_t)((unsigned)__double2hiint(r) & 0xFFu) << 32);
  }
  // a row with |x| >= 2^E outgrew the scale
  __device__ __forceinline__ bool grew(double amax) const {
    return amax >= ldexp(1.0, E);
  }
  // the chunk's largest limb integer is nonzero but 9 or more bits below the
  // scale (the chunk after an outlier)
  __device__ __forceinline__ bool shrank(double amax) const {
    const uint64_t qm = q(amax);
    return qm != 0u && qm < ((uint64_t)1 << (8 * S - 11));
  }
};

// max(a, |b|) for finite values: one v_max_f64 with the absolute-value
// modifier.  Written with fmax, IEEE mode adds a canonicalising v_max_f64 for
// every loaded operand (to quiet a signalling NaN), doubling the count.
__device__ __forceinline__ double oz_amax(double a, double b) {
  double r;
  asm("v_max_f64 %0, %1, |%2|" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// Limb words of four consecutive rows of a column: w[s] holds digit s of row
// j in byte j.  Byte p of the four offset integers is gathered with two
// v_perm (selector p picks byte p of the second operand, 4 + p of the first,
// 12 a zero byte).  Raises amax to the rows' largest |x|: the limb integers
// and the growth test of a chunk are monotone in |x|, so its largest one
// decides both.
template <int S>
__device__ __forceinline__ void oz_pack4(const double (&xv)[4],
                                         const OzScale<S>& sc, int (&w)[S],
                                         double& amax) {
  unsigned u[S][4];  // u[p][j]: the dword of row j whose byte p (mod 4) is digit S - 1 - p
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const double r = sc.raw(xv[j]);
    const unsigned lo = (unsigned)__double2loint(r);
#pragma unroll
    for (int p = 0; p < S && p < 4; p++) u[p][j] = lo + (0x80808080u & ((1u << (8 * p)) - 1u));
    if constexpr (S == 5) {  // byte 4: the high dword plus the carry of lo + 0x80808080
      const unsigned l4 = lo + 0x80808080u;
      u[4][j] = (unsigned)__double2hiint(r) + (l4 < lo ? 1u : 0u);
    }
    amax = oz_amax(amax, xv[j]);
  }
#pragma unroll
  for (int p = 0; p < S; p++) {
    const unsigned bp = p & 3;
    w[S - 1 - p] = (int)(__builtin_amdgcn_perm(u[p][1], u[p][0], 0x0C0C0000u | ((4u + bp) << 8) | bp) |
                         __builtin_amdgcn_perm(u[p][3], u[p][2], ((4u + bp) << 24) | (bp << 16) | 0x0C0Cu));
  }
}

// Calls f(integral_constant<int, I>) for I = 0 .. N - 1 in order: unrolled
// at compile time, so indices into register arrays stay constant.
template <class F, int... I>
__device__ __forceinline__ void oz_static_for(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}

// The k-th INT8 product of a chunk (accumulator block tt of TT, level g,
// limb s of A, limb g - s of B), block-major over the pairs with
// s + t < oz_levels<S>().
struct OzMf {
  int tt, g, s;
};
template <int S, int TT>
constexpr int oz_nmfma() {
  int n = 0;
  for (int g = 0; g < oz_levels<S>(); g++)
    for (int s = 0; s < S; s++)
      if (g - s >= 0 && g - s < S) n++;
  return TT * n;
}
template <int S, int TT>
constexpr OzMf oz_mfma_of(int k) {
  int n = 0;
  for (int tt = 0; tt < TT; tt++)
    for (int g = 0; g < oz_levels<S>(); g++)
      for (int s = 0; s < S; s++) {
        const int t = g - s;
        if (t < 0 || t >= S) continue;
        if (n == k) return OzMf{tt, g, s};
        n++;
      }
  return OzMf{0, 0, 0};
}

// SELF: a diagonal tile of a Gram product (64-row chunks, the second index of
// the limb staging is the chunk's 32-row half instead of A/B).
// TN: the tile's width in columns of B (128, or 64 for 4 and 5 limbs, whose
// S + 1 accumulator levels for a 64 x 64 region per wave would not fit the
// 256 matrix-core registers: a 64 x 32 region per wave halves them).
// SHARE: the launch has several tiles, which read the same rows (tiles in
// one row of the tile grid share A's, in one column B's).  The tiles of one
// chunk set kp then run on one XCD (workgroups go to the XCDs round robin,
// so workgroup w sits on XCD w mod 8), and rows are loaded without the
// nontemporal hint, so that the tiles after the first find them in that
// XCD's L2.
// CAT: a general product with M, N <= 64 in one 64 x 64 tile, laid out as the
// diagonal tile of the Gram product of [A B]: wave w streams columns 32 w ..
// 32 w + 31 of [A B] (A's columns 0..63, B's 64..127) over 64-row chunks as
// SELF does, and computes the 32 x 32 block (w & 1, w >> 1) of C.  A chunk
// carries 128 columns that C uses, against 64 + 64 of the 128 + TN that a
// general tile loads at these widths.
// SELF and CAT: a Gram product with M <= 64 in one 64 x 64 tile: columns
// 64..127 are A's columns again, 64 rows further down (128-row chunks, two
// 64-row blocks), and the two copies of a column share one scale.  Wave w
// computes the 32 x 32 block (w & 1, w >> 1) of C from both blocks.
template <int S, bool SELF, int TN, bool SHARE = false, bool CAT = false>
__global__ void __launch_bounds__(kOzThreads)
    __attribute__((amdgpu_waves_per_eu(1, 1))) ozaki_kernel(OzArgs a) {
  static_assert(CAT ? (TN == 64 && !SHARE)
                    : (TN == 128 ? S == 3 : (TN == 64 && !SELF)), "tile width");
  constexpr bool HALVES = SELF || CAT;  // 64-row blocks, two 32-row halves
  constexpr bool GCAT = SELF && CAT;    // two 64-row blocks per chunk
  // operand sets per chunk: (half, block) pairs, each multiplied in turn
  constexpr int NPART = GCAT ? 4 : HALVES ? 2 : 1;
  constexpr int NSUB = CAT ? 1 : TN / 64;  // 32-column blocks per wave
  constexpr int TT = CAT ? 1 : 2 * NSUB;   // 32 x 32 accumulator blocks per wave
  constexpr int NAV = TT / NSUB;           // 32-column blocks of A per wave
  // limbs of two chunks (double buffer): [buffer][A/B or half][limb][column][row];
  // a column's 32 bytes sit in a 48-byte row, so the 16-byte accesses of 16
  // consecutive lanes (32-byte strides otherwise) fall on distinct banks
  __shared__ __attribute__((aligned(16))) signed char sl[2][2][S][kOzTile][kOzRows + 16];
  __shared__ int erun[2][kOzTile];
  __shared__ int s_last;
  __shared__ int vote[3];  // rescale votes, one slot per step in rotation
  __shared__ double tree[2 * kOzThreads];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // this wave's 64 x (TN / 2) region: rows 64 wm.., columns (TN / 2) wn..
  const int wm = wave >> 1, wn = wave & 1;
  const int ntiles = a.ntiles;
  // chunk set kp of tile `tile`; SHARE with G a multiple of 8: workgroups
  // w = 8 j + x take tile j mod ntiles of chunk set x + 8 (j / ntiles)
  int tile, kp;
  if (SHARE && (a.G & 7) == 0) {
    const int x = blockIdx.x & 7, j = blockIdx.x >> 3;
    tile = j % ntiles;
    kp = x + 8 * (j / ntiles);
  } else {
    tile = blockIdx.x % ntiles;
    kp = blockIdx.x / ntiles;
  }
  int tm, tn;
  oz_tile_of(a, tile, tm, tn);
  const int m0 = tm * kOzTile;
  const int n0 = tn * TN;

  // General tiles: waves 0-1 stream the 128 columns of A's tile, waves 2-3
  // the TN columns of B (TN = 64: wave 3 loads nothing); thread `col` holds
  // its column's 32 rows of the current chunk.
  // SELF: wave w streams columns 32 w .. 32 w + 31, lanes 0-31 the chunk's
  // first 32 rows and lanes 32-63 its last 32 (mat = the half), so the two
  // halves of a column meet in one wave.
  constexpr int CR = NPART * kOzRows;  // rows per chunk
  const int mat = HALVES ? lane >> 5 : wave >> 1;
  const int col = HALVES ? 32 * wave + (lane & 31) : tid & 127;
  const bool isA = SELF || (CAT ? wave < 2 : mat == 0);
  const int width = isA ? a.M : a.N;
  const int gcol = CAT ? (GCAT ? col & 63 : isA ? col : col - 64) : (isA ? m0 : n0) + col;
  // a column of the tile (for TN = 64, waves 0-2; stated through the wave
  // index so that the compiler sees a wave-uniform load resource)
  const bool cload = CAT || TN == 128 || isA || (wave & 1) == 0;
  const bool cvalid = cload && gcol < width;
  const double* mbase = isA ? a.A : a.B;
  const int64_t ld = isA ? a.lda : a.ldb;
  const int ld8 = (int)(8 * ld);
  const int voff = 8 * (cvalid ? gcol : 0) + (HALVES ? kOzRows * mat * ld8 : 0) +
                   (GCAT && wave >= 2 ? 2 * kOzRows * ld8 : 0);

  erun[mat][col] = kOzEmin;
  if (tid < 3) vote[tid] = 0;
  constexpr int L = oz_levels<S>();  // accumulator levels g = 0 .. L - 1
  v16i acc[L][TT];
#pragma unroll
  for (int g = 0; g < L; g++)
#pragma unroll
    for (int tt = 0; tt < TT; tt++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[g][tt][r] = 0;
  // this workgroup's FP64 partial tile, accumulated by the (rare) flushes
  constexpr int PS = kOzTile * TN;
  // the wave's region of C: rows (columns of A) from AM, columns (of B)
  // from BN; its first staged column of A (AM) and of B (BS)
  const int AM = CAT ? 32 * (wave & 1) : 64 * wm;
  const int BN = CAT ? 32 * (wave >> 1) : (TN / 2) * wn;
  const int BS = CAT && !SELF ? 64 + BN : BN;
  double* mine = a.part1 + ((int64_t)tile * a.G + kp) * PS;
  const __amdgpu_buffer_rsrc_t part =
      oz_rsrc(mine + BN * kOzTile + AM, 8 * (int64_t)(PS - BN * kOzTile - AM));
  const int lofs = 8 * ((lane & 31) * kOzTile + 4 * (lane >> 5));
#pragma unroll
  for (int tt = 0; tt < TT; tt++)
#pragma unroll
    for (int r = 0; r < 16; r++)
      oz_st(0.0, part, lofs,
            8 * (32 * (tt % NSUB) * kOzTile + 32 * (tt / NSUB) + 8 * (r >> 2) + (r & 3)));
  const __amdgpu_buffer_rsrc_t stage = oz_rsrc(
      a.stage + ((int64_t)blockIdx.x * (kOzThreads / 64) + wave) * (L * TT * 1024),
      (int64_t)L * TT * 1024 * 4);
  const int* ea_w = &erun[0][AM];
  const int* eb_w = &erun[1][BS];
  __syncthreads();

  const int64_t nch = (a.K + CR - 1) / CR;
  // D chunks of this thread's rows in flight while one is split: a chunk's
  // registers are refilled row group by row group, as it is split, with the
  // chunk D = 2 steps ahead (prefetch distance 2)
  constexpr int D = 2;
  double xa[kOzRows], xb[kOzRows];
  // this thread's 32 rows of chunk c in its column: the chunk's resource
  // ends after its last row below K, so rows at or past K read 0 and every
  // chunk issues the same 32 loads back to back.  Padding columns read
  // column 0 and are zeroed through their scale; threads without a column
  // (TN = 64) get an empty resource, so their loads return 0 and touch no
  // memory.
  auto chunk_rsrc = [&](int64_t c) {
    const int64_t r0 = c * CR;
    const int rows = (int)min<int64_t>(CR, a.K - r0);
    return oz_rsrc(mbase + r0 * ld, cload ? 8 * ((int64_t)(rows - 1) * ld + width) : 0);
  };
  auto load_chunk = [&](int64_t c, double (&dst)[kOzRows]) {
    const __amdgpu_buffer_rsrc_t rs = chunk_rsrc(c);
#pragma unroll
    for (int i = 0; i < kOzRows; i++)
      dst[i] = SHARE ? oz_ld(rs, voff, i * ld8) : oz_ld_stream(rs, voff, i * ld8);
  };
  // frexp exponent of the chunk's largest |x| (max |x| < 2^e), kOzEmin for
  // an all-zero chunk or a padding column
  auto chunk_exp = [&](const double (&x)[kOzRows]) {
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < kOzRows; i++) mx = fmax(mx, fabs(x[i]));
    int e = kOzEmin;
    if (cvalid && mx > 0.0) (void)frexp(mx, &e);
    if (HALVES) e = max(e, __shfl_xor(e, 32));  // both halves of the column
    return e;
  };
  // the scale the limbs are taken against: 2^E for a live column; an
  // all-zero column (E = kOzEmin) uses 2^-1074, below which nothing is
  // nonzero, so any nonzero value reads as growth; padding columns use
  // 2^2000, which turns every value into zero limbs
  auto split_scale = [&]() {
    return cvalid ? max(erun[mat][col], -1074) : 2000;
  };
  // split the chunk in x into S limbs and stage them in LDS buffer b
  auto limbs = [&](int b, const double (&x)[kOzRows]) {
    const OzScale<S> Es(split_scale());
#pragma unroll
    for (int half = 0; half < 2; half++) {
      int words[S][4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const double xv[4] = {x[16 * half + 4 * q], x[16 * half + 4 * q + 1],
                              x[16 * half + 4 * q + 2], x[16 * half + 4 * q + 3]};
        int w[S];
        double amax = 0.0;
        oz_pack4<S>(xv, Es, w, amax);
#pragma unroll
        for (int s2 = 0; s2 < S; s2++) words[s2][q] = w[s2];
      }
#pragma unroll
      for (int s2 = 0; s2 < S; s2++) {
        const v4i pk = {words[s2][0], words[s2][1], words[s2][2], words[s2][3]};
        *reinterpret_cast<v4i*>(&sl[b][mat][s2][col][16 * half]) = pk;
      }
    }
  };
  // INT8 products of the chunk staged in buffer b, exact INT32 accumulation
  // per level g = s + t
  const int kh = 16 * (lane >> 5);
  auto multiply = [&](int b) {
#pragma unroll
    for (int p = 0; p < NPART; p++)  // halves (p & 1) and blocks (p >> 1) of the chunk
#pragma unroll
      for (int tt = 0; tt < TT; tt++) {
        const int h = p & 1;
        const int mcol = 64 * (p >> 1) + AM + 32 * (tt / NSUB) + (lane & 31);
        const int ncol = 64 * (p >> 1) + BS + 32 * (tt % NSUB) + (lane & 31);
#pragma unroll
        for (int g = 0; g < L; g++) {
#pragma unroll
          for (int s1 = 0; s1 < S; s1++) {
            const int t1 = g - s1;
            if (t1 < 0 || t1 >= S) continue;
            const v4i av = *reinterpret_cast<const v4i*>(&sl[b][HALVES ? h : 0][s1][mcol][kh]);
            const v4i bv = *reinterpret_cast<const v4i*>(&sl[b][HALVES ? h : 1][t1][ncol][kh]);
            acc[g][tt] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bv, acc[g][tt], 0, 0, 0);
          }
        }
      }
  };
  // Steady-state step: split the chunk in x into LDS buffer bs while the
  // chunk staged in buffer bm is multiplied, and refill x with chunk cnext.
  // The splits of 8 groups of 4 rows, their refill loads and the chunk's
  // products are issued interleaved, in fixed units separated by scheduling
  // fences, so the VALU work runs while the matrix core is busy.  Returns
  // nonzero if the column's scale must restart: a value outgrew it, or the
  // chunk's largest limb integer is nonzero but 9 or more bits below it (the
  // chunk after an outlier).
  // With several operand sets (SELF and CAT: the chunk's two halves; both
  // with two blocks: four sets) set p's products run in units p * 8 / NPART
  // onwards, and set p + 1's operands are read at the end of set p's last
  // unit.
  auto split_multiply = [&](int bm, int bs, double (&x)[kOzRows], int64_t cnext) {
    v4i av[NAV][S], bv[NSUB][S];
    // operand set p: half p & 1, block p >> 1
    auto operands = [&](int p) {
      const int h2 = p & 1, c0 = 64 * (p >> 1);
#pragma unroll
      for (int s2 = 0; s2 < S; s2++) {
#pragma unroll
        for (int h = 0; h < NAV; h++)
          av[h][s2] = *reinterpret_cast<const v4i*>(&sl[bm][HALVES ? h2 : 0][s2][c0 + AM + 32 * h + (lane & 31)][kh]);
#pragma unroll
        for (int h = 0; h < NSUB; h++)
          bv[h][s2] = *reinterpret_cast<const v4i*>(&sl[bm][HALVES ? h2 : 1][s2][c0 + BS + 32 * h + (lane & 31)][kh]);
      }
    };
    operands(0);
    const __amdgpu_buffer_rsrc_t rn = chunk_rsrc(cnext);
    const OzScale<S> Es(split_scale());
    double amax = 0.0;
    int words[S][4];
    oz_static_for(
        [&](auto U_) {
          constexpr int U = decltype(U_)::value;  // rows 4 U .. 4 U + 3
          {
            const double xv[4] = {x[4 * U], x[4 * U + 1], x[4 * U + 2], x[4 * U + 3]};
            int w[S];
            oz_pack4<S>(xv, Es, w, amax);
#pragma unroll
            for (int s2 = 0; s2 < S; s2++) words[s2][U % 4] = w[s2];
          }
          // these 4 rows are split: their registers take chunk cnext
#pragma unroll
          for (int j = 0; j < 4; j++)
            x[4 * U + j] = SHARE ? oz_ld(rn, voff, (4 * U + j) * ld8)
                                 : oz_ld_stream(rn, voff, (4 * U + j) * ld8);
          constexpr int NM = oz_nmfma<S, TT>();
          constexpr int NT = NPART * NM;  // products per chunk
          constexpr int k0 = U * NT / 8, k1 = (U + 1) * NT / 8;
          oz_static_for(
              [&](auto K_) {
                constexpr OzMf m = oz_mfma_of<S, TT>((k0 + decltype(K_)::value) % NM);
                acc[m.g][m.tt] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                    av[m.tt / NSUB][m.s], bv[m.tt % NSUB][m.g - m.s],
                    acc[m.g][m.tt], 0, 0, 0);
              },
              std::make_integer_sequence<int, k1 - k0>{});
          constexpr int UP = 8 / NPART;  // units per operand set
          if constexpr (NPART > 1 && U % UP == UP - 1 && U < 7) operands((U + 1) / UP);
          if constexpr (U % 4 == 3) {
#pragma unroll
            for (int s2 = 0; s2 < S; s2++) {
              const v4i pk = {words[s2][0], words[s2][1], words[s2][2], words[s2][3]};
              *reinterpret_cast<v4i*>(&sl[bs][mat][s2][col][16 * (U / 4)]) = pk;
            }
          }
          __builtin_amdgcn_sched_barrier(0);
        },
        std::make_integer_sequence<int, kOzRows / 4>{});
    if (HALVES) amax = fmax(amax, __shfl_xor(amax, 32));  // both halves share the scale
    return Es.grew(amax) || Es.shrank(amax);
  };

  // Software pipeline: while chunk c is split (VALU) into one LDS buffer,
  // the previous chunk's limbs in the other buffer are multiplied (matrix
  // cores) and chunks c + G and c + 2 G are being loaded.  A flush happens
  // before chunk c when c raises a column's scale or when its rows would take
  // the accumulators past kOzFlushRows rows; `since` counts the rows staged
  // since the last flush.  Growth shows up while c is being split: the
  // previous chunk has been multiplied by then, and c is loaded and split
  // again after the flush.
  // Workgroup-wide any(p) with one barrier: a wave ballot, then one lane of
  // each voting wave sets this call's LDS slot.  The slots rotate over
  // three: call n clears call n + 1's slot before its own barrier (the votes
  // for n + 1 come after it), and call n's slot is cleared again only in
  // call n + 2, after every thread has passed call n + 1's barrier and so
  // has read it.  (__syncthreads_or computes the workgroup size from the
  // dispatch packet and takes three barriers.)
  int vslot = 0;
  auto wg_any = [&](bool p) {
    const int now = vslot;
    vslot = vslot == 2 ? 0 : vslot + 1;
    if (tid == 0) vote[vslot] = 0;
    if (__builtin_amdgcn_ballot_w64(p) != 0 && lane == 0) vote[now] = 1;
    __syncthreads();
    return vote[now] != 0;
  };

  int64_t ch = kp;
  int pb = 0;
  while (ch < nch) {
    // (re)start at chunk ch: accumulators empty, nothing staged, so every
    // column's scale restarts from this chunk (an outlier only coarsens the
    // window it falls in); only this thread touches the entry
    load_chunk(ch, xa);
    erun[mat][col] = chunk_exp(xa);
    __syncthreads();
    if constexpr (GCAT) {  // both blocks of a column share the scale
      const int e2 = max(erun[mat][col], erun[mat][col ^ 64]);
      __syncthreads();
      erun[mat][col] = e2;
      __syncthreads();
    }
    load_chunk(min(ch + a.G, nch - 1), xb);
    limbs(pb, xa);
    load_chunk(min(ch + D * a.G, nch - 1), xa);
    int since = CR;
    __syncthreads();
    bool staged = true;  // buffer pb holds a chunk not yet multiplied
    ch += a.G;
    // one chunk: 0 go on, 1 flush (buffer pb staged), 2 rescale (buffer pb
    // multiplied, chunk ch split again after the flush)
    auto step = [&](double (&x)[kOzRows]) -> int {
      if (ch >= nch) return 1;
      oz_pin<L, TT>(acc);
      if (since + CR > kOzFlushRows) return 1;  // uniform
      const bool rescale = split_multiply(pb, pb ^ 1, x, min(ch + D * a.G, nch - 1));
      if (wg_any(rescale)) return 2;
      since += CR;
      pb ^= 1;
      ch += a.G;
      return 0;
    };
    // xb holds chunk ch and xa chunk ch + G, the two alternate; each step
    // leaves the loop on its own
    int r;
    for (;;) {
      if ((r = step(xb)) != 0) break;
      if ((r = step(xa)) != 0) break;
    }
    staged = r == 1;
    // drain: multiply the staged chunk, convert the accumulators
    if (staged) multiply(pb);
    oz_flush<L, TT, NSUB>(acc, stage, part, lofs, ea_w, eb_w, lane);
    __syncthreads();
  }

  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  const int grp = kp / a.GS;
  const int gsz = min(a.GS, a.G - grp * a.GS);
  unsigned* cnt1 = a.cnt + tile * a.ngroups + grp;
  unsigned* cnt2 = a.cnt + ntiles * a.ngroups + tile;
  if (tid == 0) {
    unsigned t = __hip_atomic_fetch_add(cnt1, 1u, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
    s_last = (t == (unsigned)(gsz - 1));
  }
  __syncthreads();
  if (!s_last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  const int Mloc = min(a.M - m0, kOzTile), Nloc = min(a.N - n0, TN);
  const bool mirror = a.mode == OZ_UPPER;
  auto apply = [&](int e, double s) {
    const int ml = e % kOzTile, nl = e / kOzTile;
    if (ml < Mloc && nl < Nloc) {
#pragma clang fp contract(off)
      double* c = a.C + (int64_t)(n0 + nl) * a.ldc + (m0 + ml);
      double r = a.alpha * s;
      if (a.beta != 0.0) r = a.beta * *c + r;
      *c = r;
      if (mirror) {  // C(n, m) of a Gram product
        double* c2 = a.C + (int64_t)(m0 + ml) * a.ldc + (n0 + nl);
        double r2 = a.alpha * s;
        if (a.beta != 0.0) r2 = a.beta * *c2 + r2;
        *c2 = r2;
      }
    }
  };
  const double* gsrc = a.part1 + ((int64_t)tile * a.G + (int64_t)grp * a.GS) * PS;
  if (a.ngroups == 1) {
    team_reduce<double>(gsrc, gsz, PS, tree, apply);
    if (tid == 0) *cnt1 = 0u;
    return;
  }
  double* gdst = a.part2 + ((int64_t)tile * a.ngroups + grp) * PS;
  team_reduce<double>(gsrc, gsz, PS, tree, [&](int e, double s) { gdst[e] = s; });
  if (tid == 0) *cnt1 = 0u;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (tid == 0) {
    unsigned t = __hip_atomic_fetch_add(cnt2, 1u, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
    s_last = (t == (unsigned)(a.ngroups - 1));
  }
  __syncthreads();
  if (!s_last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  team_reduce<double>(a.part2 + (int64_t)tile * a.ngroups * PS, a.ngroups, PS,
                      tree, apply);
  if (tid == 0) *cnt2 = 0u;
}

// Lane-layout probe of the INT8 MFMA (diagnostic): out[0..1023] = row index
// of element (lane, r) (as 32 * row), out[1024..2047] = column index (32 *
// column), from A = lane row index / B = ones and the reverse.
__global__ void i8_layout_probe(int* out) {
  const int lane = threadIdx.x;
  const int i = lane & 31;
  const int rowv = i | (i << 8) | (i << 16) | (i << 24);
  const int ones = 0x01010101;
  const v4i a1 = {rowv, rowv, rowv, rowv}, b1 = {ones, ones, ones, ones};
  v16i d1 = {};
  d1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, b1, d1, 0, 0, 0);
  v16i d2 = {};
  d2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(b1, a1, d2, 0, 0, 0);
  for (int r = 0; r < 16; r++) {
    out[lane * 16 + r] = d1[r];
    out[1024 + lane * 16 + r] = d2[r];
  }
}

static int isqrt_u(int x) {
  int r = 1;
  while (r * r < x) r++;
  return r;
}

}  // namespace skinny

using namespace skinny;

extern "C" {

skinny_status_t skinny_dtsmttsm_ozaki(skinny_handle_t h, size_t blockCount,
                                      int64_t M, int64_t N, int64_t K,
                                      const double* A, int64_t lda,
                                      double alpha, const double* B,
                                      int64_t ldb, double beta, double* C,
                                      int64_t ldc, int slices,
                                      skinny_stream_t stream) {
  if (M < 0 || N < 0 || K < 0)
    return fail(SKINNY_INVALID_ARG, "ozaki: negative size");
  if (M == 0 || N == 0) return SKINNY_OK;
  if (!C || (K > 0 && (!A || !B)))
    return fail(SKINNY_INVALID_ARG, "ozaki: null matrix");
  if (lda < M || ldb < N || ldc < M)
    return fail(SKINNY_INVALID_ARG, "ozaki: leading dimension too small");
  if (slices < 3 || slices > 5)
    return fail(SKINNY_UNSUPPORTED, "ozaki: 3..5 slices");
  if (M > 1024 || N > 1024)
    return fail(SKINNY_UNSUPPORTED, "ozaki: M, N <= 1024");
  if (lda > (1 << 22) || ldb > (1 << 22))
    return fail(SKINNY_UNSUPPORTED, "ozaki: lda, ldb <= 2^22");
  h = resolve(h);
  if (!h) return no_handle("ozaki");
  // 3 limbs: 128 x 128 tiles.  Gram products skip the tiles below the
  // diagonal and read A once (64-row chunks on the diagonal tiles, whose
  // offsets need lda <= 2^21).  4 and 5 limbs: 128 x 64 tiles, all of them.
  // Other products with M, N <= 64: one tile on [A B] (64-row chunks, so
  // lda, ldb <= 2^21 as well); Gram products with M <= 64 (any limb count):
  // one tile on two 64-row blocks of A (128-row chunks).
  const int levels = slices == 3 ? oz_levels<3>() : slices == 4 ? oz_levels<4>() : oz_levels<5>();
  const bool same = A == B && lda == ldb && M == N && lda <= (1 << 21);
  const bool gcat = same && M <= 64;
  const bool gram = slices == 3 && same && !gcat;
  const bool cat = gcat || (!gram && M <= 64 && N <= 64 && lda <= (1 << 21) && ldb <= (1 << 21));
  const int TN = slices == 3 && !cat ? kOzTile : 64;
  const int tilesM = (int)((M + kOzTile - 1) / kOzTile);
  const int tilesN = (int)((N + TN - 1) / TN);
  struct Set {
    int mode, ntiles;
    bool self;
    OzArgs a;
  };
  Set sets[2];
  int nsets = 0;
  if (gram) {
    sets[nsets++] = Set{OZ_DIAG, tilesM, true, {}};
    if (tilesM > 1) sets[nsets++] = Set{OZ_UPPER, tilesM * (tilesM - 1) / 2, false, {}};
  } else {
    sets[nsets++] = Set{OZ_ALL, tilesM * tilesN, false, {}};
  }
  const size_t PS = (size_t)kOzTile * TN;
  const int TT = cat ? 1 : 2 * (TN / 64);  // accumulator blocks per wave
  size_t bytes = 0;  // the launches run one after the other: the larger one
  for (int i = 0; i < nsets; i++) {
    OzArgs& a = sets[i].a;
    const int ntiles = sets[i].ntiles;
    const int64_t rows = gcat ? 4 * kOzRows : sets[i].self || cat ? 2 * kOzRows : kOzRows;
    int64_t G = (int64_t)blockCount;
    if (G <= 0) {
      G = std::max<int64_t>(1, h->cu_count / ntiles);
      G = std::min<int64_t>(G, std::max<int64_t>(1, K / (rows * 64)));
    }
    if (G > 65536) G = 65536;
    a.G = (int)G;
    if (a.G <= 16) {
      a.GS = a.G;
      a.ngroups = 1;
    } else {
      a.GS = isqrt_u(a.G);
      a.ngroups = (a.G + a.GS - 1) / a.GS;
    }
    if ((size_t)ntiles * (a.ngroups + 1) > (size_t)kMaxCounters)
      return fail(SKINNY_UNSUPPORTED, "ozaki: too many partial groups");
    size_t b = (size_t)ntiles * a.G * PS * sizeof(double);
    if (a.ngroups > 1) b += (size_t)ntiles * a.ngroups * PS * sizeof(double);
    b += (size_t)ntiles * a.G * (kOzThreads / 64) * levels * 1024 * TT * sizeof(int);
    bytes = std::max(bytes, b);
  }
  Workspace w;
  skinny_status_t st = acquire_workspace(h, (hipStream_t)stream, bytes, &w);
  if (st != SKINNY_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < nsets; i++) {
    OzArgs& a = sets[i].a;
    const int ntiles = sets[i].ntiles;
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
    a.tilesM = tilesM;
    a.tilesN = tilesN;
    a.ntiles = ntiles;
    a.mode = sets[i].mode;
    a.part1 = static_cast<double*>(ws_scratch(w));
    a.part2 = a.part1 + (size_t)ntiles * a.G * PS;
    size_t part_bytes = (size_t)ntiles * a.G * PS * sizeof(double);
    if (a.ngroups > 1) part_bytes += (size_t)ntiles * a.ngroups * PS * sizeof(double);
    a.stage = reinterpret_cast<int*>(static_cast<char*>(ws_scratch(w)) + part_bytes);
    a.cnt = ws_counters(w);
    const int blocks = a.G * ntiles;
    const dim3 grid(blocks), wg(kOzThreads);
    if (sets[i].self) {
      hipLaunchKernelGGL((ozaki_kernel<3, true, 128>), grid, wg, 0, s, a);
    } else if (gcat) {
      switch (slices) {
        case 3: hipLaunchKernelGGL((ozaki_kernel<3, true, 64, false, true>), grid, wg, 0, s, a); break;
        case 4: hipLaunchKernelGGL((ozaki_kernel<4, true, 64, false, true>), grid, wg, 0, s, a); break;
        default: hipLaunchKernelGGL((ozaki_kernel<5, true, 64, false, true>), grid, wg, 0, s, a); break;
      }
    } else if (cat) {
      switch (slices) {
        case 3: hipLaunchKernelGGL((ozaki_kernel<3, false, 64, false, true>), grid, wg, 0, s, a); break;
        case 4: hipLaunchKernelGGL((ozaki_kernel<4, false, 64, false, true>), grid, wg, 0, s, a); break;
        default: hipLaunchKernelGGL((ozaki_kernel<5, false, 64, false, true>), grid, wg, 0, s, a); break;
      }
    } else if (ntiles > 1) {  // tiles sharing rows: same XCD, cached loads
      switch (slices) {
        case 3: hipLaunchKernelGGL((ozaki_kernel<3, false, 128, true>), grid, wg, 0, s, a); break;
        case 4: hipLaunchKernelGGL((ozaki_kernel<4, false, 64, true>), grid, wg, 0, s, a); break;
        default: hipLaunchKernelGGL((ozaki_kernel<5, false, 64, true>), grid, wg, 0, s, a); break;
      }
    } else {
      switch (slices) {
        case 3: hipLaunchKernelGGL((ozaki_kernel<3, false, 128>), grid, wg, 0, s, a); break;
        case 4: hipLaunchKernelGGL((ozaki_kernel<4, false, 64>), grid, wg, 0, s, a); break;
        default: hipLaunchKernelGGL((ozaki_kernel<5, false, 64>), grid, wg, 0, s, a); break;
      }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "ozaki launch");
  }
  return SKINNY_OK;
}

skinny_status_t skinny_probe_i8_layout(int* host_out) {
  if (!host_out) return fail(SKINNY_INVALID_ARG, "probe: null out");
  int* d = nullptr;
  hipError_t e = hipMalloc(&d, 2048 * sizeof(int));
  if (e != hipSuccess) return hip_fail(e, "probe alloc");
  hipLaunchKernelGGL(i8_layout_probe, dim3(1), dim3(64), 0, 0, d);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(host_out, d, 2048 * sizeof(int), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "probe");
  return SKINNY_OK;
}

}  // extern "C"
