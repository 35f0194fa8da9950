// lbm_batch.cpp — many lattices stepped together.
//
// Members are grouped by kernel configuration (lattice, storage, boundary mode, collision, MRT
// engine).  One batch step is one persistent launch per group: the group's tiles are laid end to
// end (member m owns tiles [prefix[m], prefix[m+1])) and each workgroup walks one contiguous chunk
// (k_step_batch).  Parameter blocks for both buffer orientations live on the device, so a step
// needs no host->device traffic; two steps are captured in a hipGraph and replayed.  The grid size
// of each group is chosen by timing candidate sizes on the group itself (lbm_batch_tune): a trial
// launch reads every member's current buffer and writes only its other buffer, which the next
// real step overwrites, so tuning leaves the members' states untouched.
//
// While a lattice is a member all of its work goes to the batch's stream, which keeps member
// reads, writes and batch steps in one order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "lbm_handle.hpp"

using namespace lbm;

namespace {

struct Group {
  Config cfg{};
  std::vector<lbm_ctx *> members;
  StepParams *d_params[2] = {nullptr, nullptr};  // [phase][member]
  int32_t *d_prefix = nullptr;
  int32_t total_tiles = 0;
  int grid = 0;
};

bool same_cfg(const Config &a, const Config &b) {
  return a.Q == b.Q && a.storage == b.storage && a.bc == b.bc && a.collision == b.collision &&
         uses_mfma(a) == uses_mfma(b);
}

}  // namespace

struct lbm_batch_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<lbm_ctx *> members;
  std::vector<Group> groups;
  std::vector<int> tuned_grid;  // per group, 0 = default
  bool dirty = true;
  int phase = 0;  // member m reads f[cur_at_build(m) ^ phase]
  hipGraphExec_t graph = nullptr;
  int64_t steps = 0;

  void free_groups() {
    for (auto &g : groups) {
      for (int k = 0; k < 2; ++k)
        if (g.d_params[k]) (void)hipFree(g.d_params[k]);
      if (g.d_prefix) (void)hipFree(g.d_prefix);
    }
    groups.clear();
    if (graph) (void)hipGraphExecDestroy(graph);
    graph = nullptr;
  }
};

namespace {

#define BCHK(call)                                                                                 \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess) return set_error(LBM_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int default_grid(const Group &g, int cus) {
  // enough resident workgroups to cover the chip several times over, never more than tiles
  return std::max(1, std::min(g.total_tiles, 8 * cus));
}

lbm_status rebuild(lbm_batch_ctx *b) {
  // members' own state is current: fold the batch phase into their cur before regrouping
  b->free_groups();
  b->phase = 0;
  for (lbm_ctx *m : b->members) {
    Group *grp = nullptr;
    for (auto &g : b->groups)
      if (same_cfg(g.cfg, m->cfg)) grp = &g;
    if (!grp) {
      b->groups.emplace_back();
      grp = &b->groups.back();
      grp->cfg = m->cfg;
    }
    grp->members.push_back(m);
  }
  const int cus = device_cu_count(b->device);
  b->tuned_grid.resize(b->groups.size(), 0);
  for (size_t gi = 0; gi < b->groups.size(); ++gi) {
    Group &g = b->groups[gi];
    const size_t n = g.members.size();
    std::vector<int32_t> prefix(n + 1, 0);
    std::vector<StepParams> par[2];
    for (int k = 0; k < 2; ++k) par[k].resize(n);
    for (size_t i = 0; i < n; ++i) {
      prefix[i + 1] = prefix[i] + g.members[i]->base.ntiles;
      par[0][i] = g.members[i]->params(g.members[i]->cur);
      par[1][i] = g.members[i]->params(g.members[i]->cur ^ 1);
    }
    g.total_tiles = prefix[n];
    for (int k = 0; k < 2; ++k) {
      BCHK(hipMalloc((void **)&g.d_params[k], n * sizeof(StepParams)));
      BCHK(hipMemcpyAsync(g.d_params[k], par[k].data(), n * sizeof(StepParams), hipMemcpyHostToDevice, b->stream));
    }
    BCHK(hipMalloc((void **)&g.d_prefix, (n + 1) * sizeof(int32_t)));
    BCHK(hipMemcpyAsync(g.d_prefix, prefix.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice, b->stream));
    g.grid = b->tuned_grid[gi] > 0 ? std::min(b->tuned_grid[gi], g.total_tiles) : default_grid(g, cus);
  }
  BCHK(hipStreamSynchronize(b->stream));
  b->dirty = false;
  return LBM_OK;
}

lbm_status launch_phase(lbm_batch_ctx *b, int phase, hipStream_t s) {
  for (auto &g : b->groups) {
    BatchArgs a{g.d_params[phase], g.d_prefix, (int32_t)g.members.size(), g.total_tiles};
    BCHK(launch_step_batch(g.cfg, a, g.grid, s));
  }
  return LBM_OK;
}

lbm_status ensure_graph(lbm_batch_ctx *b) {
  if (b->graph) return LBM_OK;
  hipGraph_t g = nullptr;
  BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  lbm_status s0 = launch_phase(b, 0, b->stream);
  lbm_status s1 = s0 == LBM_OK ? launch_phase(b, 1, b->stream) : s0;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (s1 != LBM_OK || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return set_error(LBM_EDEVICE, "batch graph capture failed");
  }
  e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    b->graph = nullptr;
    return set_error(LBM_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  return LBM_OK;
}

lbm_status enqueue(lbm_batch_ctx *b, int64_t nsteps) {
  if (b->dirty) {
    lbm_status st = rebuild(b);
    if (st != LBM_OK) return st;
  }
  if (b->groups.empty() || nsteps == 0) return LBM_OK;
  int64_t left = nsteps;
  if (left >= 2) {
    lbm_status st = ensure_graph(b);
    if (st != LBM_OK) return st;
  }
  if (b->phase == 1 && left > 0) {
    lbm_status st = launch_phase(b, 1, b->stream);
    if (st != LBM_OK) return st;
    b->phase = 0;
    --left;
  }
  for (; left >= 2; left -= 2) BCHK(hipGraphLaunch(b->graph, b->stream));
  if (left == 1) {
    lbm_status st = launch_phase(b, 0, b->stream);
    if (st != LBM_OK) return st;
    b->phase = 1;
  }
  b->steps += nsteps;
  const int flip = (int)(nsteps & 1);
  for (lbm_ctx *m : b->members) {
    m->cur ^= flip;
    m->steps += nsteps;
  }
  return LBM_OK;
}

}  // namespace

extern "C" {

lbm_status lbm_batch_create(lbm_batch *out) {
  if (!out) return set_error(LBM_EINVAL, "out is NULL");
  *out = nullptr;
  lbm_batch_ctx *b = new (std::nothrow) lbm_batch_ctx();
  if (!b) return set_error(LBM_ENOMEM, "host allocation failed");
  if (hipGetDevice(&b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    lbm_batch_destroy(b);
    return set_error(LBM_EDEVICE, "batch stream/event creation failed");
  }
  *out = b;
  return LBM_OK;
}

lbm_status lbm_batch_add(lbm_batch b, lbm_handle h) {
  if (!b || !h) return set_error(LBM_EINVAL, "NULL argument");
  if (h->batch) return set_error(LBM_EINVAL, "the lattice already belongs to a batch");
  if (h->device != b->device) return set_error(LBM_EINVAL, "the lattice lives on another device");
  BCHK(hipStreamSynchronize(h->stream));
  BCHK(hipStreamSynchronize(b->stream));
  if (b->phase == 1) {  // fold the pending orientation into every current member first
    b->phase = 0;
    b->dirty = true;
  }
  h->batch = b;
  h->borrowed = b->stream;
  b->members.push_back(h);
  b->dirty = true;
  return LBM_OK;
}

lbm_status lbm_batch_remove(lbm_batch b, lbm_handle h) {
  if (!b || !h) return set_error(LBM_EINVAL, "NULL argument");
  auto it = std::find(b->members.begin(), b->members.end(), h);
  if (it == b->members.end()) return set_error(LBM_EINVAL, "the lattice is not a member of this batch");
  BCHK(hipStreamSynchronize(b->stream));
  b->members.erase(it);
  h->batch = nullptr;
  h->borrowed = nullptr;
  b->phase = 0;
  b->dirty = true;
  return LBM_OK;
}

lbm_status lbm_batch_step(lbm_batch b, int64_t nsteps) {
  if (!b) return set_error(LBM_EINVAL, "batch is NULL");
  if (nsteps < 0) return set_error(LBM_EINVAL, "nsteps < 0");
  return enqueue(b, nsteps);
}

lbm_status lbm_batch_step_timed(lbm_batch b, int64_t nsteps, double *ms) {
  if (!b) return set_error(LBM_EINVAL, "batch is NULL");
  if (nsteps < 0) return set_error(LBM_EINVAL, "nsteps < 0");
  if (b->dirty) {
    lbm_status st = rebuild(b);
    if (st != LBM_OK) return st;
  }
  BCHK(hipEventRecord(b->ev0, b->stream));
  lbm_status st = enqueue(b, nsteps);
  if (st != LBM_OK) return st;
  BCHK(hipEventRecord(b->ev1, b->stream));
  BCHK(hipEventSynchronize(b->ev1));
  float t = 0;
  BCHK(hipEventElapsedTime(&t, b->ev0, b->ev1));
  if (ms) *ms = (double)t;
  return LBM_OK;
}

lbm_status lbm_batch_sync(lbm_batch b) {
  if (!b) return set_error(LBM_EINVAL, "batch is NULL");
  BCHK(hipStreamSynchronize(b->stream));
  return LBM_OK;
}

lbm_status lbm_batch_tune(lbm_batch b, int32_t trials) {
  if (!b) return set_error(LBM_EINVAL, "batch is NULL");
  if (b->dirty) {
    lbm_status st = rebuild(b);
    if (st != LBM_OK) return st;
  }
  const int cus = device_cu_count(b->device);
  const int reps = trials > 0 ? trials : 5;
  for (size_t gi = 0; gi < b->groups.size(); ++gi) {
    Group &g = b->groups[gi];
    int best = g.grid;
    float best_ms = 1e30f;
    for (int f : {1, 2, 4, 8, 16, 0}) {  // 0 = one workgroup per tile
      const int grid = f ? std::min(g.total_tiles, f * cus) : g.total_tiles;
      // trial launches read the current buffers and write only the other ones (b->phase)
      BatchArgs a{g.d_params[b->phase], g.d_prefix, (int32_t)g.members.size(), g.total_tiles};
      BCHK(launch_step_batch(g.cfg, a, grid, b->stream));  // warm-up
      BCHK(hipEventRecord(b->ev0, b->stream));
      for (int r = 0; r < reps; ++r) BCHK(launch_step_batch(g.cfg, a, grid, b->stream));
      BCHK(hipEventRecord(b->ev1, b->stream));
      BCHK(hipEventSynchronize(b->ev1));
      float ms = 0;
      BCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
      if (ms < best_ms) {
        best_ms = ms;
        best = grid;
      }
    }
    g.grid = best;
    b->tuned_grid[gi] = best;
  }
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  return LBM_OK;
}

int32_t lbm_batch_size(lbm_batch b) { return b ? (int32_t)b->members.size() : -1; }

lbm_status lbm_batch_destroy(lbm_batch b) {
  if (!b) return LBM_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (lbm_ctx *m : b->members) {
    m->batch = nullptr;
    m->borrowed = nullptr;
  }
  b->free_groups();
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return LBM_OK;
}

}  // extern "C"
