// Work items of the persistent batched kernels: one (op, row range) of a
// TSMTTSM or TSMM on the matrix cores, dispatched on the op's tile class at
// run time.  Shared by the batched launch (batch.hip) and the fused SVQB
// sweep (svqb_fused.hip).
#pragma once

#include <algorithm>
#include <vector>

#include "internal.hpp"
#include "mfma.hpp"
#include "tsmttsm_core.hpp"

namespace skinny {

struct BatchOpDev {
  int op;  // 0 TSMTTSM, 1 TSMM
  int M, N, cls;
  int64_t K, lda, ldb, ldc;
  const double* A;
  const double* B;
  double* C;
  double alpha, beta;
  int nparts, outM, outN;
  int64_t part_off;  // doubles into the partial buffer
  int cnt;           // completion counter index
  int vec;           // TSMM: A row slices are 16-byte aligned, no masks
};

// TSMTTSM item: rows [r0, r1).  TSMM item: the 64-row chunks r0, r0 + r1,
// r0 + 2 r1, ... of the op (item r0 of r1 interleaved items), so the
// workgroups working on an op at one time stream neighbouring rows.
struct BatchItemDev {
  int op, part;
  int64_t r0, r1;
};

// ---- TSMTTSM tile classes ------------------------------------------------
// 0..5: packed P = 16, 8, 5, 4, 3, 2;  6..14: (MT, NT) in {1, 2, 4}^2;
// 15..17: Gram products (B is A, M == N) of the square classes 6, 10, 14,
// which accumulate the subtiles on and above the diagonal only
static int ts_class(int64_t M, int64_t N) {
  const int64_t W = std::max(M, N);
  if (W <= 8) {
    switch (16 / W) {
      case 16: return 0;
      case 8: return 1;
      case 5: return 2;
      case 4: return 3;
      case 3: return 4;
      default: return 5;
    }
  }
  auto c3 = [](int64_t x) { return x <= 16 ? 0 : (x <= 32 ? 1 : 2); };
  return 6 + 3 * c3(M) + c3(N);
}
static int ts_gram_class(int cls) {
  return cls == 6 ? 15 : cls == 10 ? 16 : cls == 14 ? 17 : cls;
}
static void ts_class_shape(int cls, int64_t M, int64_t N, int* outM, int* outN,
                           int* rows_min) {
  if (cls >= 15) cls = cls == 15 ? 6 : cls == 16 ? 10 : 14;
  if (cls < 6) {
    static const int P[6] = {16, 8, 5, 4, 3, 2};
    *outM = (int)M;
    *outN = (int)N;
    *rows_min = 4 * P[cls] * 8;
    return;
  }
  static const int sz[3] = {1, 2, 4};
  const int mt = sz[(cls - 6) / 3], nt = sz[(cls - 6) % 3];
  *outM = 16 * mt;
  *outN = 16 * nt;
  *rows_min = 4 * 8;
}

// SELF (a Gram class): one operand load per step, so twice the unroll of
// the general product, at most 8 (chunks of at most 32 rows, the item
// granularity)
template <int P, int MT, int NT, bool SELF = false>
__device__ __forceinline__ void ts_item(const BatchOpDev& o,
                                        const BatchItemDev& it, double* red,
                                        double* out) {
  constexpr int U0 = P > 1 ? 8 : ((MT + NT) >= 8 ? 2 : ((MT + NT) >= 4 ? 4 : 8));
  constexpr int U = SELF ? (2 * U0 < 8 ? 2 * U0 : 8) : U0;
  using Core = TsCore<double, P, MT, NT, U, SELF>;
  typename Core::V4 acc[MT][NT];
  Core::zero(acc);
  const int wave = threadIdx.x >> 6;
  Core::accumulate(o.A + it.r0 * o.lda, o.lda, o.B + it.r0 * o.ldb, o.ldb, o.M,
                   o.N, 0, 0, min(o.M, Core::DM), min(o.N, Core::DN),
                   it.r1 - it.r0, wave, 4, acc);
  Core::wg_partial(acc, red, out, o.M, o.N, o.outM, o.outN);
}

__device__ void run_ts_item(const BatchOpDev& o, const BatchItemDev& it,
                            double* red, double* out) {
  switch (o.cls) {
    case 0: ts_item<16, 1, 1>(o, it, red, out); break;
    case 1: ts_item<8, 1, 1>(o, it, red, out); break;
    case 2: ts_item<5, 1, 1>(o, it, red, out); break;
    case 3: ts_item<4, 1, 1>(o, it, red, out); break;
    case 4: ts_item<3, 1, 1>(o, it, red, out); break;
    case 5: ts_item<2, 1, 1>(o, it, red, out); break;
    case 6: ts_item<1, 1, 1>(o, it, red, out); break;
    case 7: ts_item<1, 1, 2>(o, it, red, out); break;
    case 8: ts_item<1, 1, 4>(o, it, red, out); break;
    case 9: ts_item<1, 2, 1>(o, it, red, out); break;
    case 10: ts_item<1, 2, 2>(o, it, red, out); break;
    case 11: ts_item<1, 2, 4>(o, it, red, out); break;
    case 12: ts_item<1, 4, 1>(o, it, red, out); break;
    case 13: ts_item<1, 4, 2>(o, it, red, out); break;
    case 15: ts_item<1, 1, 1, true>(o, it, red, out); break;
    case 16: ts_item<1, 2, 2, true>(o, it, red, out); break;
    case 17: ts_item<1, 4, 4, true>(o, it, red, out); break;
    default: ts_item<1, 4, 4>(o, it, red, out); break;
  }
}

// ---- TSMM: rows [r0, r1) of one op by the four waves --------------------
// classes: S in {1, 2, 4, 8, 16} x NT in {1, 2, 4}  ->  3 * iS + iN;
// 15, 16: M, N <= 2 (M = 1, 2), one thread per row
static int tm_class(int64_t M, int64_t N) {
  if (M <= 2 && N <= 2) return M <= 1 ? 15 : 16;
  const int64_t S = (M + 3) / 4;
  const int iS = S <= 1 ? 0 : S <= 2 ? 1 : S <= 4 ? 2 : S <= 8 ? 3 : 4;
  const int64_t nt = (N + 15) / 16;
  const int iN = nt <= 1 ? 0 : (nt <= 2 ? 1 : 2);
  return 3 * iS + iN;
}

// S of a TSMM matrix-core class (columns of A per lane group), 0 for the
// row classes 15 and 16
static int tm_class_S(int cls) {
  static const int S[5] = {1, 2, 4, 8, 16};
  return cls >= 15 ? 0 : S[cls / 3];
}

// ---- which tile classes can compute an op --------------------------------
// Any of them gives the op's result (on integer data the same bits); they
// differ in loads per row, padding and registers, which is what the batch
// tuner times.  TSMTTSM: packed class P needs P max(M, N) <= 16 (P rows of
// both operands in 16 lanes); subtile class (MT, NT) needs 16 MT >= M and
// 16 NT >= N; the Gram classes a Gram op (B is A, lda == ldb, M == N)
// within their square tile.  TSMM: class (S, NT) needs 4 S >= M and
// 16 NT >= N; the row classes M <= 1 (15) or M <= 2 (16), and N <= 2.
static bool ts_class_valid(int cls, int64_t M, int64_t N, bool gram) {
  if (cls < 0 || cls > 17) return false;
  if (cls < 6) {
    static const int P[6] = {16, 8, 5, 4, 3, 2};
    return (int64_t)P[cls] * std::max(M, N) <= 16;
  }
  if (cls >= 15) {
    const int side = cls == 15 ? 16 : cls == 16 ? 32 : 64;
    return gram && M <= side;
  }
  static const int sz[3] = {1, 2, 4};
  return 16 * sz[(cls - 6) / 3] >= M && 16 * sz[(cls - 6) % 3] >= N;
}
static bool tm_class_valid(int cls, int64_t M, int64_t N) {
  if (cls < 0 || cls > 16) return false;
  if (cls >= 15) return M <= (cls == 15 ? 1 : 2) && N <= 2;
  static const int nt[3] = {1, 2, 4};
  return 4 * tm_class_S(cls) >= M && 16 * nt[cls % 3] >= N;
}

// The classes the batch tuner times for an op: the default rule's class and
// the alternatives that pad nothing or little.  TSMTTSM with max(M, N) <= 8:
// the packed classes of the largest two P that fit and the unpacked 16 x 16
// tile; a Gram op: its Gram class and the general class of the same tile.
// TSMM with M, N <= 2: the row class and the matrix-core class.
static void ts_class_candidates(int64_t M, int64_t N, bool gram, std::vector<int>& out) {
  out.clear();
  const int64_t W = std::max(M, N);
  if (W <= 8) {
    int npacked = 0;
    for (int c = 0; c < 6 && npacked < 2; c++)
      if (ts_class_valid(c, M, N, gram)) {
        out.push_back(c);
        npacked++;
      }
    out.push_back(6);
    if (gram) out.push_back(15);
    return;
  }
  const int base = ts_class(M, N);
  out.push_back(base);
  if (gram && ts_gram_class(base) != base) out.push_back(ts_gram_class(base));
}
static void tm_class_candidates(int64_t M, int64_t N, std::vector<int>& out) {
  out.clear();
  const int base = tm_class(M, N);
  out.push_back(base);
  if (base >= 15) out.push_back(0);  // S = 1, NT = 1: 4 >= M, 16 >= N
}

// BETA0 is a template parameter: with a run-time choice the compiler may
// load C before every store (the store makes the load safe to speculate),
// which adds a read of C to an op that should only write it.
template <int S, int NT, bool BETA0>
__device__ void tm_item(const BatchOpDev& o_, const BatchItemDev& it_) {
  // the records are read once into registers: through the references the
  // compiler would reload their fields after every store to C
  const BatchOpDev o = o_;
  const BatchItemDev it = it_;
  using V4 = Mfma<double>::v4;
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int wave = threadIdx.x >> 6;
  const int M = o.M, N = o.N;
  double bop[S][NT];
#pragma unroll
  for (int s = 0; s < S; s++)
#pragma unroll
    for (int u = 0; u < NT; u++) {
      const int m = lg * S + s, n = 16 * u + li;
      bop[s][u] = (m < M && n < N) ? *gmem(o.B + (int64_t)n * o.ldb + m) : 0.0;
    }
  const int mbeg = lg * S;
  // 64-row chunks r0 + j r1 (j = 0, 1, ...) of the op; wave w owns rows
  // 16 w .. 16 w + 15 of every chunk, so in place needs no barrier, and it
  // keeps the blocks of RB chunks in flight before the first MFMA.  The
  // batched kernels run at two waves per SIMD (fixed by the largest TSMTTSM
  // tile), so a wave keeps about 32 loads per lane in flight.
  // (S = 16 keeps one chunk: its 64-value B operand and two chunks of A
  // would exceed the registers)
  constexpr int RB = S >= 16 ? 1 : (S >= 8 ? 4 : (S >= 4 ? 8 : (S >= 2 ? 16 : 32)));
  const int64_t nch = (o.K + 63) >> 6;
  const int64_t mine = it.r0 < nch ? (nch - it.r0 + it.r1 - 1) / it.r1 : 0;
  for (int64_t j0 = 0; j0 < mine; j0 += RB) {
    double av[RB][S];
#pragma unroll
    for (int r = 0; r < RB; r++) {
      const int64_t row = (it.r0 + (j0 + r) * it.r1) * 64 + wave * 16 + li;
      const bool ok = (j0 + r < mine) && row < o.K;
      const double* ar = o.A + row * o.lda + mbeg;
      bool done = false;
      if constexpr (S % 2 == 0) {
        if (o.vec) {
#pragma unroll
          for (int s = 0; s < S; s += 2) {
            typedef double d2v __attribute__((ext_vector_type(2)));
            const d2v v = ok ? *gmem(reinterpret_cast<const d2v*>(ar + s)) : d2v{0.0, 0.0};
            av[r][s] = v.x;
            av[r][s + 1] = v.y;
          }
          done = true;
        }
      }
      if (!done) {
#pragma unroll
        for (int s = 0; s < S; s++) av[r][s] = (ok && mbeg + s < M) ? *gmem(ar + s) : 0.0;
      }
    }
#pragma unroll
    for (int r = 0; r < RB; r++) {
      if (j0 + r >= mine) break;
      V4 acc[NT];
#pragma unroll
      for (int u = 0; u < NT; u++) acc[u] = V4{0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s < S; s++)
#pragma unroll
        for (int u = 0; u < NT; u++) acc[u] = Mfma<double>::mac(av[r][s], bop[s][u], acc[u]);
      const int64_t rbase = (it.r0 + (j0 + r) * it.r1) * 64 + wave * 16;
#pragma unroll
      for (int u = 0; u < NT; u++) {
        const int col = 16 * u + li;
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int64_t rr = rbase + Mfma<double>::drow(lg, v);
          if (rr < o.K && col < N) {
#pragma clang fp contract(off)
            auto c = gmem(o.C + rr * o.ldc + col);
            const double sres = acc[u][v];
            if (BETA0)
              *c = o.alpha * sres;
            else
              *c = o.beta * *c + o.alpha * sres;
          }
        }
      }
    }
  }
}

// M, N <= 2: a thread owns whole rows (R rows in flight), B in registers;
// every load of a row precedes its stores, so C may alias A
template <int MC, bool BETA0>
__device__ void tm_rows_item(const BatchOpDev& o_, const BatchItemDev& it_) {
  const BatchOpDev o = o_;
  const BatchItemDev it = it_;
  constexpr int R = 16;
  const int M = o.M, N = o.N;
  double b[2][2];
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int n = 0; n < 2; n++)
      b[m][n] = (m < M && n < N) ? *gmem(o.B + (int64_t)n * o.ldb + m) : 0.0;
  // 64-row chunks r0 + j r1 of the op: wave w takes chunks j = 4 q + w,
  // lane l row l of the chunk, R chunks in flight
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t nch = (o.K + 63) >> 6;
  const int64_t mine = it.r0 < nch ? (nch - it.r0 + it.r1 - 1) / it.r1 : 0;
  for (int64_t j0 = wave; j0 < mine; j0 += 4 * R) {
    double av[R][MC];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int64_t j = j0 + 4 * r;
      const int64_t row = (it.r0 + j * it.r1) * 64 + lane;
      const bool ok = j < mine && row < o.K;
      const double* ar = o.A + row * o.lda;
#pragma unroll
      for (int m = 0; m < MC; m++) av[r][m] = (ok && m < M) ? *gmem(ar + m) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int64_t j = j0 + 4 * r;
      const int64_t row = (it.r0 + j * it.r1) * 64 + lane;
      if (j >= mine || row >= o.K) continue;
      auto cr = gmem(o.C + row * o.ldc);
#pragma unroll
      for (int n = 0; n < 2; n++) {
        if (n >= N) continue;
        double sres = 0.0;
#pragma unroll
        for (int m = 0; m < MC; m++) sres = __builtin_fma(av[r][m], b[m][n], sres);
        {
#pragma clang fp contract(off)
          if (BETA0)
            cr[n] = o.alpha * sres;
          else
            cr[n] = o.beta * cr[n] + o.alpha * sres;
        }
      }
    }
  }
}

template <bool BETA0>
__device__ void run_tm_item_b(const BatchOpDev& o, const BatchItemDev& it) {
  switch (o.cls) {
    case 15: tm_rows_item<1, BETA0>(o, it); break;
    case 16: tm_rows_item<2, BETA0>(o, it); break;
    case 0: tm_item<1, 1, BETA0>(o, it); break;
    case 1: tm_item<1, 2, BETA0>(o, it); break;
    case 2: tm_item<1, 4, BETA0>(o, it); break;
    case 3: tm_item<2, 1, BETA0>(o, it); break;
    case 4: tm_item<2, 2, BETA0>(o, it); break;
    case 5: tm_item<2, 4, BETA0>(o, it); break;
    case 6: tm_item<4, 1, BETA0>(o, it); break;
    case 7: tm_item<4, 2, BETA0>(o, it); break;
    case 8: tm_item<4, 4, BETA0>(o, it); break;
    case 9: tm_item<8, 1, BETA0>(o, it); break;
    case 10: tm_item<8, 2, BETA0>(o, it); break;
    case 11: tm_item<8, 4, BETA0>(o, it); break;
    case 12: tm_item<16, 1, BETA0>(o, it); break;
    case 13: tm_item<16, 2, BETA0>(o, it); break;
    default: tm_item<16, 4, BETA0>(o, it); break;
  }
}

__device__ void run_tm_item(const BatchOpDev& o, const BatchItemDev& it) {
  if (o.beta == 0.0)
    run_tm_item_b<true>(o, it);
  else
    run_tm_item_b<false>(o, it);
}

// Host plan of a batch: device op records, the (op, row range) items sorted
// largest first, and the doubles of TSMTTSM partials needed.  An op is split
// into items of about item_bytes of row data (at most 64 partials for a
// TSMTTSM op).  item_bytes <= 0 picks about 8 items per workgroup of a grid
// of `grid` workgroups, within 1..8 MiB per item: a batch of few large ops
// still gives every workgroup work, one of many ops keeps large items.
// cls_force (may be null): per op a tile class to use instead of the default
// rule, or -1; a class that cannot compute the op is refused.
inline bool batch_op_gram(const skinny_batch_op& o) {
  return o.op == 0 && o.A == o.B && o.lda == o.ldb && o.M == o.N;
}
inline skinny_status_t batch_plan(int nops, const skinny_batch_op* ops,
                                  int64_t item_bytes, int grid,
                                  std::vector<BatchOpDev>& dops,
                                  std::vector<BatchItemDev>& items,
                                  size_t& part_doubles,
                                  const int* cls_force = nullptr) {
  if (item_bytes <= 0) {
    double total = 0.0;
    for (int i = 0; i < nops; i++)
      total += 8.0 * (double)(ops[i].M + ops[i].N) * (double)std::max<int64_t>(0, ops[i].K);
    const double per = total / (8.0 * std::max(1, grid));
    item_bytes = (int64_t)std::min(std::max(per, (double)(1 << 20)), (double)(8 << 20));
  }
  dops.assign(nops, BatchOpDev{});
  items.clear();
  part_doubles = 0;
  for (int i = 0; i < nops; i++) {
    const skinny_batch_op& o = ops[i];
    if (o.op != 0 && o.op != 1) return fail(SKINNY_INVALID_ARG, "batch: bad op kind");
    if (o.M < 1 || o.N < 1 || o.K < 0 || o.M > 64 || o.N > 64)
      return fail(SKINNY_UNSUPPORTED, "batch: 1 <= M, N <= 64 only");
    if (!o.A || !o.B || !o.C) return fail(SKINNY_INVALID_ARG, "batch: null matrix");
    BatchOpDev d{};
    d.op = o.op;
    d.M = (int)o.M;
    d.N = (int)o.N;
    d.K = o.K;
    d.lda = o.lda;
    d.ldb = o.ldb;
    d.ldc = o.ldc;
    d.A = o.A;
    d.B = o.B;
    d.C = o.C;
    d.alpha = o.alpha;
    d.beta = o.beta;
    d.cnt = i;
    const int64_t row_bytes = 8 * (o.op == 0 ? (o.M + o.N) : (o.M + o.N));
    if (o.op == 0) {
      if (o.lda < o.M || o.ldb < o.N || o.ldc < o.M)
        return fail(SKINNY_INVALID_ARG, "batch: leading dimension too small");
      int rows_min;
      d.cls = ts_class(o.M, o.N);
      if (batch_op_gram(o)) d.cls = ts_gram_class(d.cls);
      if (cls_force && cls_force[i] >= 0) {
        if (!ts_class_valid(cls_force[i], o.M, o.N, batch_op_gram(o)))
          return fail(SKINNY_UNSUPPORTED, "batch: tile class cannot compute the op");
        d.cls = cls_force[i];
      }
      ts_class_shape(d.cls, o.M, o.N, &d.outM, &d.outN, &rows_min);
      int64_t rows = std::max<int64_t>(item_bytes / row_bytes, rows_min);
      rows = std::max<int64_t>(rows, (o.K + 63) / 64);  // at most 64 partials
      rows = (rows + rows_min - 1) / rows_min * rows_min;
      const int64_t np = std::max<int64_t>(1, (o.K + rows - 1) / rows);
      d.nparts = (int)np;
      d.part_off = (int64_t)part_doubles;
      part_doubles += (size_t)np * d.outM * d.outN;
      for (int64_t p = 0; p < np; p++)
        items.push_back({i, (int)p, p * rows, std::min(o.K, (p + 1) * rows)});
    } else {
      if (o.lda < o.M || o.ldb < o.M || o.ldc < o.N)
        return fail(SKINNY_INVALID_ARG, "batch: leading dimension too small");
      if (o.A == o.C && o.lda != o.ldc)
        return fail(SKINNY_UNSUPPORTED, "batch: in place needs lda == ldc");
      d.cls = tm_class(o.M, o.N);
      if (cls_force && cls_force[i] >= 0) {
        if (!tm_class_valid(cls_force[i], o.M, o.N))
          return fail(SKINNY_UNSUPPORTED, "batch: tile class cannot compute the op");
        d.cls = cls_force[i];
      }
      // 16-byte row slices: the class's slice of S columns per lane group is
      // exactly M / 4 and even (M = 8, 16, 32 or 64 on its own class), even
      // lda, aligned A
      const int S = tm_class_S(d.cls);
      d.vec = (S % 2 == 0 && o.M == 4 * S && o.lda % 2 == 0 &&
               (reinterpret_cast<uintptr_t>(o.A) % 16) == 0) ? 1 : 0;
      // P interleaved items of about item_bytes each
      int64_t rows = std::max<int64_t>(item_bytes / row_bytes, 64);
      const int64_t nch = (o.K + 63) / 64;
      const int64_t P = std::min<int64_t>(nch, std::max<int64_t>(1, (o.K + rows - 1) / rows));
      for (int64_t q = 0; q < P && o.K > 0; q++) items.push_back({i, 0, q, P});
      d.nparts = 0;
    }
    dops[i] = d;
  }
  // largest items first: the tail of the queue is the small remainders
  auto weight = [&](const BatchItemDev& x) -> int64_t {
    const BatchOpDev& a = dops[x.op];
    const int64_t rows = a.op == 0 ? x.r1 - x.r0 : (a.K + x.r1 - 1) / x.r1;
    return rows * (a.M + a.N);
  };
  std::stable_sort(items.begin(), items.end(),
                   [&](const BatchItemDev& x, const BatchItemDev& y) {
                     return weight(x) > weight(y);
                   });
  return SKINNY_OK;
}

}  // namespace skinny
