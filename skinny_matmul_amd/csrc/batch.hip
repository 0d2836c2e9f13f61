// Persistent batched TSMTTSM / TSMM: many independent products of uneven
// (K, M, N) in one launch.
//
// Replaces multi_dispatch.cuh:3-90 (a recursive compile-time ladder mapping a
// runtime (M, N) to one template instance per launch) with a work queue:
//
//  * The host splits every op into items of about 8 MB of row data (at most
//    64 items per op), sorts the items largest first and uploads them once.
//  * A persistent grid of workgroups takes items from a device-wide atomic
//    counter until the queue is empty: the only wait is the queue itself, so
//    every workgroup reaches the exit.  The last workgroup to leave resets
//    the counters, so the batch can be relaunched or replayed from a graph
//    without a memset.
//  * Each op is mapped to a tile class of the matrix-core kernels (packed
//    blocks for widths <= 8, 16-wide subtiles up to 64 x 64); the class is a
//    runtime switch inside the kernel.
//  * A TSMTTSM item writes a partial; the item that completes an op sums the
//    op's partials in item order and applies alpha / beta, so the result does
//    not depend on the schedule.  A TSMM item owns its rows (in place allowed).
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
};

struct BatchItemDev {
  int op, part;
  int64_t r0, r1;
};

struct BatchArgs {
  const BatchOpDev* ops;
  const BatchItemDev* items;
  int nitems;
  double* partials;
  unsigned* cnt;    // per-op completion counters
  unsigned* queue;  // [0] next item, [1] workgroups exited
};

// ---- TSMTTSM tile classes ------------------------------------------------
// 0..5: packed P = 16, 8, 5, 4, 3, 2;  6..14: (MT, NT) in {1, 2, 4}^2
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
static void ts_class_shape(int cls, int64_t M, int64_t N, int* outM, int* outN,
                           int* rows_min) {
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

template <int P, int MT, int NT>
__device__ __forceinline__ void ts_item(const BatchOpDev& o,
                                        const BatchItemDev& it, double* red,
                                        double* out) {
  constexpr int U = P > 1 ? 8 : ((MT + NT) >= 8 ? 2 : ((MT + NT) >= 4 ? 4 : 8));
  using Core = TsCore<double, P, MT, NT, U, false>;
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

template <int S, int NT>
__device__ void tm_item(const BatchOpDev& o, const BatchItemDev& it) {
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
      bop[s][u] = (m < M && n < N) ? o.B[(int64_t)n * o.ldb + m] : 0.0;
    }
  const int mbeg = lg * S;
  const bool beta0 = o.beta == 0.0;
  // rows [r0, r1) in 16-row blocks (r0 is a multiple of 16); a wave takes RB
  // consecutive blocks per step, all their loads in flight before the first
  // MFMA (about 8 loads per lane); wave w owns blocks w RB .. w RB + RB - 1
  // of every group of 4 RB, so in place needs no barrier
  constexpr int RB = S >= 16 ? 1 : (S >= 8 ? 2 : (S >= 4 ? 4 : 8));
  const int64_t nblk = (it.r1 - it.r0 + 15) >> 4;
  for (int64_t b0 = (int64_t)wave * RB; b0 < nblk; b0 += 4 * RB) {
    double av[RB][S];
#pragma unroll
    for (int r = 0; r < RB; r++) {
      const int64_t row = it.r0 + (b0 + r) * 16 + li;
      const bool ok = row < it.r1;
      const double* ar = o.A + row * o.lda + mbeg;
#pragma unroll
      for (int s = 0; s < S; s++) av[r][s] = (ok && mbeg + s < M) ? ar[s] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < RB; r++) {
      V4 acc[NT];
#pragma unroll
      for (int u = 0; u < NT; u++) acc[u] = V4{0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s < S; s++)
#pragma unroll
        for (int u = 0; u < NT; u++) acc[u] = Mfma<double>::mac(av[r][s], bop[s][u], acc[u]);
#pragma unroll
      for (int u = 0; u < NT; u++) {
        const int col = 16 * u + li;
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int64_t rr = it.r0 + (b0 + r) * 16 + Mfma<double>::drow(lg, v);
          if (rr < it.r1 && col < N) {
#pragma clang fp contract(off)
            double* c = o.C + rr * o.ldc + col;
            const double sres = acc[u][v];
            *c = beta0 ? o.alpha * sres : o.beta * *c + o.alpha * sres;
          }
        }
      }
    }
  }
}

// M, N <= 2: a thread owns whole rows (R rows in flight), B in registers;
// every load of a row precedes its stores, so C may alias A
template <int MC>
__device__ void tm_rows_item(const BatchOpDev& o, const BatchItemDev& it) {
  constexpr int R = 4;
  const int M = o.M, N = o.N;
  const bool beta0 = o.beta == 0.0;
  double b[2][2];
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int n = 0; n < 2; n++)
      b[m][n] = (m < M && n < N) ? o.B[(int64_t)n * o.ldb + m] : 0.0;
  const int64_t nrows = it.r1 - it.r0;
  for (int64_t base = threadIdx.x; base < nrows; base += 256 * R) {
    double av[R][MC];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int64_t q = base + 256 * r;
      const double* ar = o.A + (it.r0 + q) * o.lda;
#pragma unroll
      for (int m = 0; m < MC; m++) av[r][m] = (q < nrows && m < M) ? ar[m] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int64_t q = base + 256 * r;
      if (q >= nrows) continue;
      double* cr = o.C + (it.r0 + q) * o.ldc;
#pragma unroll
      for (int n = 0; n < 2; n++) {
        if (n >= N) continue;
        double sres = 0.0;
#pragma unroll
        for (int m = 0; m < MC; m++) sres = __builtin_fma(av[r][m], b[m][n], sres);
        {
#pragma clang fp contract(off)
          cr[n] = beta0 ? o.alpha * sres : o.beta * cr[n] + o.alpha * sres;
        }
      }
    }
  }
}

__device__ void run_tm_item(const BatchOpDev& o, const BatchItemDev& it) {
  switch (o.cls) {
    case 15: tm_rows_item<1>(o, it); break;
    case 16: tm_rows_item<2>(o, it); break;
    case 0: tm_item<1, 1>(o, it); break;
    case 1: tm_item<1, 2>(o, it); break;
    case 2: tm_item<1, 4>(o, it); break;
    case 3: tm_item<2, 1>(o, it); break;
    case 4: tm_item<2, 2>(o, it); break;
    case 5: tm_item<2, 4>(o, it); break;
    case 6: tm_item<4, 1>(o, it); break;
    case 7: tm_item<4, 2>(o, it); break;
    case 8: tm_item<4, 4>(o, it); break;
    case 9: tm_item<8, 1>(o, it); break;
    case 10: tm_item<8, 2>(o, it); break;
    case 11: tm_item<8, 4>(o, it); break;
    case 12: tm_item<16, 1>(o, it); break;
    case 13: tm_item<16, 2>(o, it); break;
    default: tm_item<16, 4>(o, it); break;
  }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
    batch_kernel(BatchArgs a) {
  __shared__ double red[TsCore<double, 1, 4, 4, 2, false>::RED_ELEMS];
  __shared__ int s_item;
  __shared__ int s_last;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0)
      s_item = (int)__hip_atomic_fetch_add(&a.queue[0], 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int idx = s_item;
    __syncthreads();
    if (idx >= a.nitems) break;
    const BatchItemDev it = a.items[idx];
    const BatchOpDev o = a.ops[it.op];
    if (o.op == 1) {
      run_tm_item(o, it);
      continue;
    }
    const int PS = o.outM * o.outN;
    double* part = a.partials + o.part_off;
    run_ts_item(o, it, red, part + (int64_t)it.part * PS);
    // completion: the item that finishes the op sums its partials in order
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (tid == 0) {
      unsigned t = __hip_atomic_fetch_add(&a.cnt[o.cnt], 1u, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
      s_last = (t == (unsigned)(o.nparts - 1));
    }
    __syncthreads();
    if (s_last) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      for (int e = tid; e < PS; e += 256) {
        const int ml = e % o.outM, nl = e / o.outM;
        if (ml >= o.M || nl >= o.N) continue;
        double s = 0.0;
        for (int q = 0; q < o.nparts; q++) s += part[(int64_t)q * PS + e];
        {
#pragma clang fp contract(off)
          double* c = o.C + (int64_t)nl * o.ldc + ml;
          double r = o.alpha * s;
          if (o.beta != 0.0) r = o.beta * *c + r;
          *c = r;
        }
      }
      if (tid == 0) a.cnt[o.cnt] = 0u;
    }
    __syncthreads();
  }
  // leave; the last workgroup out resets the queue for the next launch
  if (tid == 0) {
    unsigned t = __hip_atomic_fetch_add(&a.queue[1], 1u, __ATOMIC_ACQ_REL,
                                        __HIP_MEMORY_SCOPE_AGENT);
    if (t == gridDim.x - 1) {
      __hip_atomic_store(&a.queue[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&a.queue[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace skinny

struct skinny_batch_s {
  int device = 0;
  int grid = 0;
  int nitems = 0;
  int nops = 0;
  size_t partial_bytes = 0;
  skinny::BatchOpDev* d_ops = nullptr;
  skinny::BatchItemDev* d_items = nullptr;
  double* d_part = nullptr;
  unsigned* d_cnt = nullptr;  // nops counters + 2 queue counters
};

using namespace skinny;

extern "C" {

skinny_status_t skinny_dbatch_destroy(skinny_batch_t b) {
  if (!b) return SKINNY_OK;
  (void)hipDeviceSynchronize();
  (void)hipFree(b->d_ops);
  (void)hipFree(b->d_items);
  (void)hipFree(b->d_part);
  (void)hipFree(b->d_cnt);
  delete b;
  return SKINNY_OK;
}

skinny_status_t skinny_dbatch_create(skinny_handle_t h, int nops,
                                     const skinny_batch_op* ops,
                                     int64_t item_bytes,
                                     skinny_batch_t* out) {
  if (!out || nops < 1 || !ops)
    return fail(SKINNY_INVALID_ARG, "batch: no ops");
  h = resolve(h);
  if (!h) return no_handle("batch");
  if (item_bytes <= 0) item_bytes = 8LL << 20;
  std::vector<BatchOpDev> dops(nops);
  std::vector<BatchItemDev> items;
  size_t part_doubles = 0;
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
      int64_t rows = std::max<int64_t>(item_bytes / row_bytes, 64);
      rows = (rows + 63) / 64 * 64;
      for (int64_t r = 0; r < o.K; r += rows)
        items.push_back({i, 0, r, std::min(o.K, r + rows)});
      d.nparts = 0;
    }
    dops[i] = d;
  }
  // largest items first: the tail of the queue is the small remainders
  std::stable_sort(items.begin(), items.end(),
                   [&](const BatchItemDev& x, const BatchItemDev& y) {
                     const BatchOpDev& a = dops[x.op];
                     const BatchOpDev& b = dops[y.op];
                     return (x.r1 - x.r0) * (a.M + a.N) > (y.r1 - y.r0) * (b.M + b.N);
                   });
  skinny_batch_t b = new skinny_batch_s();
  b->device = h->device;
  b->nops = nops;
  b->nitems = (int)items.size();
  b->grid = std::max(1, std::min(b->nitems, 2 * h->cu_count));
  hipError_t e = hipMalloc(&b->d_ops, sizeof(BatchOpDev) * nops);
  if (e == hipSuccess) e = hipMalloc(&b->d_items, sizeof(BatchItemDev) * std::max<size_t>(1, items.size()));
  if (e == hipSuccess) e = hipMalloc(&b->d_part, sizeof(double) * std::max<size_t>(1, part_doubles));
  if (e == hipSuccess) e = hipMalloc(&b->d_cnt, sizeof(unsigned) * (nops + 2));
  if (e == hipSuccess) e = hipMemcpy(b->d_ops, dops.data(), sizeof(BatchOpDev) * nops, hipMemcpyHostToDevice);
  if (e == hipSuccess && !items.empty())
    e = hipMemcpy(b->d_items, items.data(), sizeof(BatchItemDev) * items.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(b->d_cnt, 0, sizeof(unsigned) * (nops + 2));
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    skinny_dbatch_destroy(b);
    return hip_fail(e, "batch create");
  }
  b->partial_bytes = part_doubles * sizeof(double);
  *out = b;
  return SKINNY_OK;
}

skinny_status_t skinny_dbatch_run(skinny_batch_t b, skinny_stream_t stream) {
  if (!b) return fail(SKINNY_INVALID_ARG, "batch run: null batch");
  if (b->nitems == 0) return SKINNY_OK;
  BatchArgs a;
  a.ops = b->d_ops;
  a.items = b->d_items;
  a.nitems = b->nitems;
  a.partials = b->d_part;
  a.cnt = b->d_cnt;
  a.queue = b->d_cnt + b->nops;
  hipLaunchKernelGGL(batch_kernel, dim3(b->grid), dim3(256), 0,
                     (hipStream_t)stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "batch launch");
  return SKINNY_OK;
}

int64_t skinny_dbatch_items(skinny_batch_t b) { return b ? b->nitems : 0; }

}  // extern "C"
