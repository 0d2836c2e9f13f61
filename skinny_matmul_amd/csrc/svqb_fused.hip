// SVQB sweeps in one cooperative launch: the Gram products, the eigensolves
// and the in-place updates of all blocks are phases of a persistent kernel
// separated by grid barriers, in place of one launch per product and per
// eigensolve (svqb.hip, optionally replayed from a hipGraph).
//
// Reference: the algorithm of svqb/svqb.cu:72-130 and the grid-wide
// barrier of global_sync/sync.cu:38-53 (an atomic arrival counter polled by
// every thread).  Per sweep:
//   phase 1  Gram partials: the (block, row range) TSMTTSM items of all
//            blocks (A = B = X), each writing one partial;
//   phase 2  one workgroup per block sums its partials in item order into
//            LDS and runs the device eigensolve, writing T = D U L^-1/2;
//   phase 3  X <- X T in place: the (block, row range) TSMM items.
// Items are taken round robin by workgroup index (largest first), so the
// result does not depend on timing and is bitwise reproducible for a fixed
// grid.
//
// The barrier (grid_barrier.hpp): a monotonic arrival counter with agent-
// scope release/acquire.  The cooperative launch guarantees that all
// workgroups are resident, so every wait ends; a poll limit (about 2 s)
// still bounds it, flags the call as failed and lets the workgroup go on,
// so a broken launch cannot hang the device.  The last workgroup to leave
// resets the counters for the next launch (no memset).
#include <algorithm>
#include <vector>

#include "batch_items.hpp"
#include "grid_barrier.hpp"
#include "svqb_eig.hpp"

namespace skinny {

struct FusedArgs {
  const BatchOpDev* ops;       // [0, nb): Gram ops; [nb, 2 nb): X <- X T
  const BatchItemDev* gitems;  // Gram items
  const BatchItemDev* titems;  // update items
  int ngitems, ntitems, nb, sweeps;
  double* partials;
  unsigned* bar;  // [0] arrivals, [1] exits, [2] poll-limit flag
  unsigned spin_limit;
};

constexpr int kFusedLds =
    kEigLdsDoubles > TsCore<double, 1, 4, 4, 2, false>::RED_ELEMS
        ? kEigLdsDoubles
        : TsCore<double, 1, 4, 4, 2, false>::RED_ELEMS;

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
    svqb_fused_kernel(FusedArgs a) {
  __shared__ double lds[kFusedLds];
  __shared__ int ilds[kEigLdsInts];
  const unsigned G = gridDim.x;
  unsigned nbar = 0;
  for (int sw = 0; sw < a.sweeps; sw++) {
    if (sw > 0) grid_barrier(a.bar, ++nbar * G, a.spin_limit);
    // ---- phase 1: Gram partials ------------------------------------------
    for (int i = blockIdx.x; i < a.ngitems; i += G) {
      const BatchItemDev it = a.gitems[i];
      const BatchOpDev o = a.ops[it.op];
      run_ts_item(o, it, lds,
                  a.partials + o.part_off + (int64_t)it.part * o.outM * o.outN);
    }
    grid_barrier(a.bar, ++nbar * G, a.spin_limit);
    // ---- phase 2: sum of partials, eigensolve, T -------------------------
    for (int b = blockIdx.x; b < a.nb; b += G) {
      const BatchOpDev o = a.ops[b];
      const int M = o.M;
      const int PS = o.outM * o.outN;
      const double* part = a.partials + o.part_off;
      for (int e = threadIdx.x; e < M * M; e += 256) {
        const int i = e % M, j = e / M;
        const double* p = part + i + (int64_t)j * o.outM;
        double s = 0.0;
        for (int q = 0; q < o.nparts; q++) s += p[(int64_t)q * PS];
        lds[i * kEigLD + j] = s;
      }
      svqb_eig_lds(M, 40, lds, ilds, const_cast<double*>(a.ops[a.nb + b].B));
      __syncthreads();  // the LDS is reused by the next block
    }
    grid_barrier(a.bar, ++nbar * G, a.spin_limit);
    // ---- phase 3: X <- X T in place --------------------------------------
    for (int i = blockIdx.x; i < a.ntitems; i += G) {
      const BatchItemDev it = a.titems[i];
      run_tm_item(a.ops[it.op], it);
    }
  }
  grid_exit(a.bar);
}

}  // namespace skinny

using namespace skinny;

extern "C" {

skinny_status_t skinny_dsvqb_fused(skinny_handle_t h, int nblocks,
                                   const skinny_svqb_block* blocks, int sweeps,
                                   skinny_stream_t stream) {
  if (nblocks < 1 || !blocks) return fail(SKINNY_INVALID_ARG, "svqb fused: no blocks");
  if (sweeps < 0) return fail(SKINNY_INVALID_ARG, "svqb fused: sweeps < 0");
  for (int i = 0; i < nblocks; i++) {
    const skinny_svqb_block& b = blocks[i];
    if (b.M < 1 || b.K < 0 || !b.X || b.ldx < b.M)
      return fail(SKINNY_INVALID_ARG, "svqb fused: bad block");
    if (b.M > kSvqbMaxM) return fail(SKINNY_UNSUPPORTED, "svqb fused: M <= 64 only");
  }
  h = resolve(h);
  if (!h) return no_handle("svqb fused");
  if (sweeps == 0) return SKINNY_OK;
  hipStream_t st = (hipStream_t)stream;

  // the transforms T of all blocks (M x M each, column-major)
  std::vector<size_t> toff(nblocks);
  size_t tdoubles = 0;
  for (int i = 0; i < nblocks; i++) {
    toff[i] = tdoubles;
    tdoubles += (size_t)blocks[i].M * blocks[i].M;
  }
  // the plan needs T's address only for the TSMM op records: plan with
  // offsets from a null base, then rebase the records on the workspace
  std::vector<skinny_batch_op> ops(2 * nblocks);
  for (int i = 0; i < nblocks; i++) {
    const skinny_svqb_block& b = blocks[i];
    ops[i] = skinny_batch_op{0, b.M, b.M, b.K, b.X, b.ldx, 1.0, b.X, b.ldx, 0.0, b.X, b.ldx};
    ops[nblocks + i] = skinny_batch_op{1, b.M, b.M, b.K, b.X, b.ldx, 1.0,
                                       (const double*)(uintptr_t)(8 * (toff[i] + 1)), b.M, 0.0,
                                       b.X, b.ldx};
  }
  std::vector<BatchOpDev> dops;
  std::vector<BatchItemDev> items, gitems, titems;
  size_t part_doubles = 0;
  skinny_status_t s = batch_plan(2 * nblocks, ops.data(), 0, 2 * h->cu_count, dops, items,
                                 part_doubles);
  if (s != SKINNY_OK) return s;
  for (const BatchItemDev& it : items) (it.op < nblocks ? gitems : titems).push_back(it);

  // every per-call buffer lives in the stream's workspace (it grows on the
  // first call of a size and is reused afterwards): transforms, op records,
  // items, Gram partials and the barrier counters
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t nit = gitems.size() + titems.size();
  const size_t oT = 0;
  const size_t oOps = oT + up(sizeof(double) * tdoubles);
  const size_t oItems = oOps + up(sizeof(BatchOpDev) * dops.size());
  const size_t oPart = oItems + up(sizeof(BatchItemDev) * std::max<size_t>(1, nit));
  const size_t oBar = oPart + up(sizeof(double) * std::max<size_t>(1, part_doubles));
  const size_t need = oBar + 256;
  Workspace w;
  s = acquire_workspace(h, st, need, &w);
  if (s != SKINNY_OK) return s;
  char* base = static_cast<char*>(ws_scratch(w));
  double* dT = reinterpret_cast<double*>(base + oT);
  BatchOpDev* d_ops = reinterpret_cast<BatchOpDev*>(base + oOps);
  BatchItemDev* d_items = reinterpret_cast<BatchItemDev*>(base + oItems);
  double* d_part = reinterpret_cast<double*>(base + oPart);
  unsigned* d_bar = reinterpret_cast<unsigned*>(base + oBar);
  for (int i = 0; i < nblocks; i++) dops[nblocks + i].B = dT + toff[i];

  hipError_t e = hipMemcpyAsync(d_ops, dops.data(), sizeof(BatchOpDev) * dops.size(),
                                hipMemcpyHostToDevice, st);
  if (e == hipSuccess && !gitems.empty())
    e = hipMemcpyAsync(d_items, gitems.data(), sizeof(BatchItemDev) * gitems.size(),
                       hipMemcpyHostToDevice, st);
  if (e == hipSuccess && !titems.empty())
    e = hipMemcpyAsync(d_items + gitems.size(), titems.data(),
                       sizeof(BatchItemDev) * titems.size(), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(d_bar, 0, sizeof(unsigned) * 4, st);

  int per_cu = 0;
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, svqb_fused_kernel, 256, 0);
  unsigned flag = 0;
  if (e == hipSuccess) {
    const size_t work = std::max<size_t>({gitems.size(), titems.size(), (size_t)nblocks});
    int grid = (int)std::min<size_t>((size_t)std::max(1, per_cu) * h->cu_count, work);
    if (grid < 1) grid = 1;
    FusedArgs fa;
    fa.ops = d_ops;
    fa.gitems = d_items;
    fa.titems = d_items + gitems.size();
    fa.ngitems = (int)gitems.size();
    fa.ntitems = (int)titems.size();
    fa.nb = nblocks;
    fa.sweeps = sweeps;
    fa.partials = d_part;
    fa.bar = d_bar;
    fa.spin_limit = 1u << 25;
    void* args[] = {&fa};
    e = hipLaunchCooperativeKernel((const void*)svqb_fused_kernel, dim3(grid), dim3(256),
                                   args, 0, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(&flag, d_bar + 2, sizeof(unsigned), hipMemcpyDeviceToHost, st);
    // the host copies of the plan and the flag must outlive the stream's use
    const hipError_t e2 = hipStreamSynchronize(st);
    if (e == hipSuccess) e = e2;
  }
  if (e != hipSuccess) return hip_fail(e, "svqb fused launch");
  if (flag) return fail(SKINNY_HIP_ERROR, "svqb fused: a grid barrier timed out");
  return SKINNY_OK;
}

}  // extern "C"
