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
#include <map>
#include <tuple>
#include <vector>

#include "batch_items.hpp"

namespace skinny {

struct BatchArgs {
  const BatchOpDev* ops;
  const BatchItemDev* items;
  int nitems;
  double* partials;
  unsigned* cnt;    // per-op completion counters
  unsigned* queue;  // [0] next item, [1] workgroups exited
};

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
    // completion: the item that finishes the op sums its partials in order.
    // The partial was written with agent-scope (write-through) stores, so
    // the arrival waits for them to complete instead of writing the XCD's
    // L2 back (published(), mfma.hpp)
#ifdef BATCH_WBL2
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
#else
    published();
#endif
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
  int cu_count = 256;
  int grid = 0;
  int nitems = 0;
  int nops = 0;
  int64_t item_bytes = 0;  // as planned (0: the automatic size)
  size_t partial_bytes = 0;
  int grid_req = 0;        // grid asked for (grid is at most the item count)
  std::vector<skinny_batch_op> ops;  // host copy, for re-planning
  std::vector<int> cls_force;        // per op: tile class, or -1 (default rule)
  std::vector<char> cls_user;        // per op: the class was set by the caller
  std::vector<int> cls_plan;         // per op: the class of the current plan
  skinny::BatchOpDev* d_ops = nullptr;
  skinny::BatchItemDev* d_items = nullptr;
  double* d_part = nullptr;
  unsigned* d_cnt = nullptr;  // nops counters + 2 queue counters
};

namespace {

void batch_free_device(skinny_batch_t b) {
  (void)hipFree(b->d_ops);
  (void)hipFree(b->d_items);
  (void)hipFree(b->d_part);
  (void)hipFree(b->d_cnt);
  b->d_ops = nullptr;
  b->d_items = nullptr;
  b->d_part = nullptr;
  b->d_cnt = nullptr;
}

// Plan the batch's ops with `item_bytes` per item (0: automatic) for a grid of
// `grid` workgroups and upload the plan (replacing any previous one).
skinny_status_t batch_build(skinny_batch_t b, int64_t item_bytes, int grid) {
  using namespace skinny;
  std::vector<BatchOpDev> dops;
  std::vector<BatchItemDev> items;
  size_t part_doubles = 0;
  skinny_status_t ps = batch_plan(b->nops, b->ops.data(), item_bytes, grid, dops, items,
                                  part_doubles, b->cls_force.data());
  if (ps != SKINNY_OK) return ps;
  b->cls_plan.resize(b->nops);
  for (int i = 0; i < b->nops; i++) b->cls_plan[i] = dops[i].cls;
  b->grid_req = grid;
  batch_free_device(b);
  b->nitems = (int)items.size();
  b->grid = std::max(1, std::min(b->nitems, grid));
  b->item_bytes = item_bytes;
  const int nops = b->nops;
  hipError_t e = hipSetDevice(b->device);
  if (e == hipSuccess) e = hipMalloc(&b->d_ops, sizeof(BatchOpDev) * nops);
  if (e == hipSuccess) e = hipMalloc(&b->d_items, sizeof(BatchItemDev) * std::max<size_t>(1, items.size()));
  if (e == hipSuccess) e = hipMalloc(&b->d_part, sizeof(double) * std::max<size_t>(1, part_doubles));
  if (e == hipSuccess) e = hipMalloc(&b->d_cnt, sizeof(unsigned) * (nops + 2));
  if (e == hipSuccess) e = hipMemcpy(b->d_ops, dops.data(), sizeof(BatchOpDev) * nops, hipMemcpyHostToDevice);
  if (e == hipSuccess && !items.empty())
    e = hipMemcpy(b->d_items, items.data(), sizeof(BatchItemDev) * items.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(b->d_cnt, 0, sizeof(unsigned) * (nops + 2));
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    batch_free_device(b);
    b->nitems = 0;
    return hip_fail(e, "batch plan upload");
  }
  b->partial_bytes = part_doubles * sizeof(double);
  return SKINNY_OK;
}

}  // namespace

using namespace skinny;

extern "C" {

skinny_status_t skinny_dbatch_destroy(skinny_batch_t b) {
  if (!b) return SKINNY_OK;
  (void)hipDeviceSynchronize();
  batch_free_device(b);
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
  skinny_batch_t b = new skinny_batch_s();
  b->device = h->device;
  b->cu_count = h->cu_count;
  b->nops = nops;
  b->ops.assign(ops, ops + nops);
  b->cls_force.assign(nops, -1);
  b->cls_user.assign(nops, 0);
  skinny_status_t s = batch_build(b, item_bytes, 2 * h->cu_count);
  if (s != SKINNY_OK) {
    skinny_dbatch_destroy(b);
    return s;
  }
  *out = b;
  return SKINNY_OK;
}

// Median of `reps` timed runs of the batch's current plan after one warm-up
// run (events on the null stream).
static skinny_status_t batch_time(skinny_batch_t b, int reps, hipEvent_t e0, hipEvent_t e1,
                                  float* ms_out) {
  std::vector<float> t;
  for (int r = 0; r <= reps; r++) {
    (void)hipEventRecord(e0, nullptr);
    skinny_status_t s = skinny_dbatch_run(b, nullptr);
    (void)hipEventRecord(e1, nullptr);
    if (s != SKINNY_OK) return s;
    const hipError_t e = hipEventSynchronize(e1);
    if (e != hipSuccess) return hip_fail(e, "batch tune run");
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (r > 0) t.push_back(ms);  // the first run warms up
  }
  std::sort(t.begin(), t.end());
  *ms_out = t[t.size() / 2];
  return SKINNY_OK;
}

// Tile classes: the ops of one shape (op kind, M, N, Gram or not) whose class
// the caller did not set are timed alone, as a batch of their own, with each
// candidate class (ts_class_candidates / tm_class_candidates), and take the
// fastest.
static skinny_status_t batch_tune_classes(skinny_batch_t b, int reps, hipEvent_t e0,
                                          hipEvent_t e1) {
  std::map<std::tuple<int, int64_t, int64_t, bool>, std::vector<int>> shapes;
  for (int i = 0; i < b->nops; i++) {
    if (b->cls_user[i]) continue;
    const skinny_batch_op& o = b->ops[i];
    shapes[std::make_tuple(o.op, o.M, o.N, batch_op_gram(o))].push_back(i);
  }
  std::vector<int> cand;
  for (const auto& kv : shapes) {
    const int kind = std::get<0>(kv.first);
    const int64_t M = std::get<1>(kv.first), N = std::get<2>(kv.first);
    if (kind == 0)
      ts_class_candidates(M, N, std::get<3>(kv.first), cand);
    else
      tm_class_candidates(M, N, cand);
    if (cand.size() < 2) continue;
    const std::vector<int>& idx = kv.second;
    skinny_batch_s sub;
    sub.device = b->device;
    sub.cu_count = b->cu_count;
    sub.nops = (int)idx.size();
    for (int i : idx) sub.ops.push_back(b->ops[i]);
    sub.cls_user.assign(sub.nops, 0);
    float best = 1e30f;
    int best_c = -1;
    skinny_status_t s = SKINNY_OK;
    for (int c : cand) {
      sub.cls_force.assign(sub.nops, c);
      s = batch_build(&sub, 0, 2 * b->cu_count);
      float ms = 0.0f;
      if (s == SKINNY_OK) s = batch_time(&sub, reps, e0, e1, &ms);
      if (s != SKINNY_OK) break;
      if (ms < best) {
        best = ms;
        best_c = c;
      }
    }
    batch_free_device(&sub);
    if (s != SKINNY_OK) return s;
    for (int i : idx) b->cls_force[i] = best_c;
  }
  return SKINNY_OK;
}

// Tunes the tile class of every op shape (above), then times plans of the
// batch (item sizes: automatic, 1, 2, 4, 8 and 16 MiB; 1 or 2 workgroups per
// CU, the kernel's residency) and keeps the fastest.  TSMTTSM results follow
// the plan's partition: reproducible for a fixed plan.
skinny_status_t skinny_dbatch_tune(skinny_batch_t b, int reps, double* best_ms,
                                   int64_t* item_bytes, int* grid) {
  if (!b) return fail(SKINNY_INVALID_ARG, "batch tune: null batch");
  if (reps < 1) reps = 3;
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) return hip_fail(e, "batch tune events");
  skinny_status_t s = batch_tune_classes(b, reps, e0, e1);
  const int64_t sizes[] = {0, 1 << 20, 2 << 20, 4 << 20, 8 << 20, 16 << 20};
  float best = 1e30f;
  int64_t best_ib = 0;
  int best_g = 2 * b->cu_count;
  for (int per = 1; per <= 2 && s == SKINNY_OK; per++) {
    for (int64_t ib : sizes) {
      s = batch_build(b, ib, per * b->cu_count);
      float ms = 0.0f;
      if (s == SKINNY_OK) s = batch_time(b, reps, e0, e1, &ms);
      if (s != SKINNY_OK) break;
      if (ms < best) {
        best = ms;
        best_ib = ib;
        best_g = per * b->cu_count;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (s == SKINNY_OK) s = batch_build(b, best_ib, best_g);
  if (s != SKINNY_OK) return s;
  if (best_ms) *best_ms = best;
  if (item_bytes) *item_bytes = best_ib;
  if (grid) *grid = b->grid;
  return SKINNY_OK;
}

skinny_status_t skinny_dbatch_set_class(skinny_batch_t b, int op, int cls) {
  if (!b) return fail(SKINNY_INVALID_ARG, "batch: null batch");
  if (op < 0 || op >= b->nops) return fail(SKINNY_INVALID_ARG, "batch: op index out of range");
  const skinny_batch_op& o = b->ops[op];
  if (cls >= 0 && !(o.op == 0 ? ts_class_valid(cls, o.M, o.N, batch_op_gram(o))
                              : tm_class_valid(cls, o.M, o.N)))
    return fail(SKINNY_UNSUPPORTED, "batch: tile class cannot compute the op");
  const int prev = b->cls_force[op];
  const char prev_user = b->cls_user[op];
  b->cls_force[op] = cls < 0 ? -1 : cls;
  b->cls_user[op] = cls < 0 ? 0 : 1;
  skinny_status_t s = batch_build(b, b->item_bytes, b->grid_req);
  if (s != SKINNY_OK) {  // keep the previous plan's choice
    b->cls_force[op] = prev;
    b->cls_user[op] = prev_user;
  }
  return s;
}

skinny_status_t skinny_dbatch_classes(skinny_batch_t b, int* cls) {
  if (!b || !cls) return fail(SKINNY_INVALID_ARG, "batch classes: null argument");
  for (int i = 0; i < b->nops; i++) cls[i] = b->cls_plan[i];
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
