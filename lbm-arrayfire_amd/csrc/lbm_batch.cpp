// lbm_batch.cpp — many lattices stepped together.
//
// Members are grouped by kernel configuration (lattice, storage, boundary mode, collision, MRT
// engine).  One batch step is one launch per group over all of the group's tiles: the members'
// tiles are laid end to end (member m owns tiles [prefix[m], prefix[m+1])) and each workgroup
// walks one contiguous chunk of them (k_step_batch); by default the chunk is one tile and the
// hardware deals workgroups out as compute units free up, and lbm_batch_tune may pick a
// persistent grid of fewer, longer chunks instead, by timing candidate sizes on the group itself
// (a trial launch reads every member's current buffer and writes only its other buffer, which
// the next real step overwrites, so tuning leaves the members' states untouched).
//
// Each group keeps one device table {header, prefix, parameter blocks for both buffer
// orientations} with room for more members than it holds; a membership change rewrites the table
// with one asynchronous copy from pinned staging memory, on the batch's stream.  The kernels take
// only the table's address and read everything else from its header.  Each single-step group's
// two steps are captured in a hipGraph of its own and replayed; the graphs are re-captured when
// the set of groups changes, a table is reallocated (grown) or a launch shape changes.
//
// Groups are cut to the device's last-level cache: members of one configuration are dealt first
// fit into groups whose two population buffers per member total at most the bin size
// (LBM_BATCH_BIN_MIB, default 160 MiB of the 256 MB Infinity Cache; 0 = one group per
// configuration), a member larger than that in a group of its own.  A batch step call runs each
// group's steps back to back, so a group's working set stays in the cache across its steps, as a
// medium lattice stepped alone does; tiny members still share one launch per step.  When all
// groups fit the bin size together, they run concurrently instead: each on a stream of its own,
// forked from and joined back into the batch's stream (LBM_BATCH_CONCURRENT=0 keeps them on the
// batch's stream), and the device overlaps their launches.
//
// While a lattice is a member all of its work goes to the batch's stream, which keeps member
// reads, writes and batch steps in one order.  Joining and leaving are ordered against the
// lattice's own stream with events, not host synchronisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "lbm_handle.hpp"

using namespace lbm;

namespace {


// Device table of a group: BatchTable header | prefix[cap + 1] | prefix_multi[cap + 1] |
// params[0][cap] | params[1][cap].  The header holds the device addresses of the other parts, so
// launches depend only on d_tab.
struct Group {
  Config cfg{};
  std::vector<lbm_ctx *> members;
  int cap = 0;                 // members the table has room for
  size_t prefix_off = 0, pmulti_off = 0, par_off[2] = {0, 0}, bytes = 0;
  char *d_tab = nullptr;       // device table
  char *h_tab = nullptr;       // pinned staging copy
  hipEvent_t uploaded = nullptr;
  int32_t total_tiles = 0;
  int32_t total_multi = 0;     // multi-step tiles (multi_k > 0)
  int multi_k = 0;             // steps per multi-step launch, 0 = single steps only
  int phase = 0;               // multi-step groups: the next launch uses parameter block [phase]
  int grid = 0;
  int tuned = 0;               // grid chosen by lbm_batch_tune, 0 = default
  int nt = 0;                  // single-step launches with non-temporal stores (3D scalar groups too
                               // large for the caches)
  size_t bytes_used = 0;       // both population buffers of every member
  int size_class = 0;          // members' node counts within one factor of 4 (LBM_BATCH_SIZE_CLASSES)
  hipStream_t gs = nullptr;    // the group's own stream (concurrent groups)
  hipEvent_t done = nullptr;   // the group's work of one call, recorded on gs
  hipGraphExec_t graph = nullptr;  // two single steps (phase 0 then 1)

  BatchArgs args(int phase, int launch_grid) const {
    return BatchArgs{(const BatchTable *)d_tab, phase, launch_grid >= (multi_k ? total_multi : total_tiles) ? 1 : 0,
                     multi_k ? 0 : nt};
  }
  void drop_graph() {
    if (graph) (void)hipGraphExecDestroy(graph);
    graph = nullptr;
  }
  void release_table() {
    if (uploaded) (void)hipEventSynchronize(uploaded);
    if (d_tab) (void)hipFree(d_tab);
    if (h_tab) (void)hipHostFree(h_tab);
    if (uploaded) (void)hipEventDestroy(uploaded);
    d_tab = h_tab = nullptr;
    uploaded = nullptr;
    cap = 0;
  }
  void release() {  // the group is gone: its table, graph and stream
    if (gs) (void)hipStreamSynchronize(gs);
    release_table();
    drop_graph();
    if (done) (void)hipEventDestroy(done);
    if (gs) (void)hipStreamDestroy(gs);
    done = nullptr;
    gs = nullptr;
  }
};

// Members whose node counts lie within one factor of 4 of each other share a size class: lattices
// of similar size step well in one launch (8 D3Q27 MRT 64^3 lattices: 1.7x their rate one after
// the other), while a launch over very different sizes ends each step on its largest member's
// tail (8 lattices of sides 40-110 batched: 0.93x one after the other; as 4 concurrent batches:
// 1.14x).
int size_class_of(int64_t nodes) {
  int c = 0;
  while (nodes >= 4) {
    nodes >>= 2;
    ++c;
  }
  return c;
}

bool same_cfg(const Config &a, const Config &b) {
  return a.Q == b.Q && a.storage == b.storage && a.bc == b.bc && a.collision == b.collision &&
         uses_mfma(a) == uses_mfma(b) && (!uses_mfma(a) || a.engine == b.engine);
}

}  // namespace

struct lbm_batch_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, order = nullptr;
  std::vector<lbm_ctx *> members;
  std::vector<Group> groups;
  bool dirty = true;
  hipEvent_t fork = nullptr;  // the batch stream's point where concurrent groups start
  bool use_graph = true;   // LBM_BATCH_GRAPH=0: direct launches instead of graph replays
  bool concurrent = true;  // LBM_BATCH_CONCURRENT=0: every group on the batch's stream
  size_t bin_bytes = 0;  // LBM_BATCH_BIN_MIB: group size limit in MiB, 0 (default) = none
  size_t stream_bytes = (size_t)256 << 20;  // LBM_BATCH_STREAM_MIB: groups above take tiles in stream order
  bool size_classes = true;  // LBM_BATCH_SIZE_CLASSES=0: one group per configuration
  int64_t steps = 0;
  lbm::BatchCounters counters{};

  void drop_graph() {
    for (auto &g : groups) g.drop_graph();
  }
};

namespace {

#define BCHK(call)                                                                                 \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess) return set_error(LBM_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Makes room for n members.  A reallocated table invalidates captured launches (*grown).
lbm_status reserve(Group &g, int n, bool *grown) {
  *grown = false;
  if (n <= g.cap) return LBM_OK;
  *grown = true;
  g.release_table();
  const int cap = std::max({n, 2 * g.cap, 8});
  g.prefix_off = (sizeof(BatchTable) + 63) & ~(size_t)63;
  g.pmulti_off = g.prefix_off + (size_t)(cap + 1) * sizeof(int32_t);
  const size_t pre_end = g.pmulti_off + (size_t)(cap + 1) * sizeof(int32_t);
  g.par_off[0] = (pre_end + 255) & ~(size_t)255;
  g.par_off[1] = g.par_off[0] + (size_t)cap * sizeof(StepParams);
  g.bytes = g.par_off[1] + (size_t)cap * sizeof(StepParams);
  BCHK(hipMalloc((void **)&g.d_tab, g.bytes));
  BCHK(hipHostMalloc((void **)&g.h_tab, g.bytes, hipHostMallocDefault));
  std::memset(g.h_tab, 0, g.bytes);
  BCHK(hipEventCreateWithFlags(&g.uploaded, hipEventDisableTiming));
  g.cap = cap;
  return LBM_OK;
}

// Regroups the members and rewrites the group tables.  Members' cur is current (the batch phase
// is folded in by enqueue), so the tables restart at phase 0.
lbm_status rebuild(lbm_batch_ctx *b) {
  for (auto &g : b->groups) {
    g.members.clear();
    g.bytes_used = 0;
  }
  // first fit, in member order, into groups of one configuration and size class within the bin size
  for (lbm_ctx *m : b->members) {
    const size_t mb = 2 * m->buf_bytes();
    const int sc = b->size_classes ? size_class_of(m->N) : 0;
    Group *grp = nullptr;
    for (auto &g : b->groups)
      if (same_cfg(g.cfg, m->cfg) && g.size_class == sc &&
          (b->bin_bytes == 0 || g.bytes_used == 0 || g.bytes_used + mb <= b->bin_bytes)) {
        grp = &g;
        break;
      }
    if (!grp) {
      b->groups.emplace_back();
      grp = &b->groups.back();
      grp->cfg = m->cfg;
      grp->size_class = sc;
    }
    grp->members.push_back(m);
    grp->bytes_used += mb;
  }
  for (size_t gi = 0; gi < b->groups.size();) {
    if (b->groups[gi].members.empty()) {
      b->groups[gi].release();
      b->groups.erase(b->groups.begin() + (long)gi);
    } else {
      ++gi;
    }
  }
  for (auto &g : b->groups) {
    bool recapture = false;  // this group's captured steps no longer match its table or shape
    const int n = (int)g.members.size();
    bool grown = false;
    lbm_status st = reserve(g, n, &grown);
    if (st != LBM_OK) return st;
    if (grown) recapture = true;
    BCHK(hipEventSynchronize(g.uploaded));  // the previous upload has left the staging copy
    BatchTable *hdr = (BatchTable *)g.h_tab;
    int32_t *prefix = (int32_t *)(g.h_tab + g.prefix_off);
    int32_t *pmulti = (int32_t *)(g.h_tab + g.pmulti_off);
    StepParams *par0 = (StepParams *)(g.h_tab + g.par_off[0]);
    StepParams *par1 = (StepParams *)(g.h_tab + g.par_off[1]);
    // multi-step launches when every member uses the default multi-step shape; a group that
    // switches between multi-step and single-step launches enters or leaves the captured graph
    const int was_multi = g.multi_k;
    g.multi_k = multi_steps_per_launch(g.cfg, MULTI_DEFAULT);
    for (lbm_ctx *m : g.members)
      if (m->multi_variant != MULTI_DEFAULT || m->d.nz != 1) g.multi_k = 0;
    if ((was_multi != 0) != (g.multi_k != 0)) recapture = true;
    prefix[0] = pmulti[0] = 0;
    for (int i = 0; i < n; ++i) {
      const lbm_ctx *m = g.members[(size_t)i];
      prefix[i + 1] = prefix[i] + m->base.ntiles;
      pmulti[i + 1] = pmulti[i] + (g.multi_k ? multi_tiles(MULTI_DEFAULT, m->d.nx, m->d.ny) : 0);
      par0[i] = m->params(m->cur);
      par1[i] = m->params(m->cur ^ 1);
    }
    g.total_tiles = prefix[n];
    g.total_multi = pmulti[n];
    g.phase = 0;
    size_t group_bytes = 0;  // both population buffers of every member
    for (const lbm_ctx *m : g.members) group_bytes += 2 * m->buf_bytes();
    hdr->nmembers = n;
    hdr->total_tiles = g.total_tiles;
    hdr->total_multi = g.total_multi;
    hdr->stream = group_bytes > b->stream_bytes ? 1 : 0;
    // such groups also store non-temporally, as single lattices over the cache size do (3D scalar
    // kernels); the captured launches carry the flag, so a change re-captures them
    int nt = hdr->stream && g.cfg.Q != 9 && !uses_mfma(g.cfg) ? 1 : 0;
    for (const lbm_ctx *m : g.members)
      if (!rows_fill_lines(m->d.nx, m->esize)) nt = 0;  // partial lines stay in the L2 to merge
    if (nt != g.nt) recapture = true;
    g.nt = nt;
    hdr->prefix = (const int32_t *)(g.d_tab + g.prefix_off);
    hdr->prefix_multi = (const int32_t *)(g.d_tab + g.pmulti_off);
    hdr->members[0] = (const StepParams *)(g.d_tab + g.par_off[0]);
    hdr->members[1] = (const StepParams *)(g.d_tab + g.par_off[1]);
    // the whole table (unused rows too) keeps the copy one contiguous transfer
    BCHK(hipMemcpyAsync(g.d_tab, g.h_tab, g.bytes, hipMemcpyHostToDevice, b->stream));
    BCHK(hipEventRecord(g.uploaded, b->stream));
    // default: one workgroup per tile, dispatched by the hardware as they free up (measured
    // faster than statically chunked persistent grids, whose chunks differ in cost: partial
    // tiles, wall nodes, different members); the tuner may pick a persistent grid instead.  The
    // grid keeps room for 1/8 more tiles and shrinks only below half, so members joining and
    // leaving mostly keep the launch shape and the captured graph (spare workgroups exit at once).
    const int need = std::max(g.total_tiles, 1);
    int grid = g.grid;
    if (g.tuned > 0) grid = std::min(g.tuned, need);
    else if (grid < need || grid > 2 * need) grid = need + need / 8 + 1;
    if (grid != g.grid) recapture = true;
    g.grid = grid;
    if (recapture) g.drop_graph();
  }
  b->dirty = false;
  ++b->counters.rebuilds;
  return LBM_OK;
}

// One single step of group g from parameter block [phase], on stream s.
lbm_status launch_single(const Group &g, int phase, hipStream_t s) {
  BCHK(launch_step_batch(g.cfg, g.args(phase, g.grid), g.grid, s));
  return LBM_OK;
}

// The group's two captured single steps (phase 0 then phase 1), captured on stream s.
lbm_status ensure_graph(lbm_batch_ctx *b, Group &g, hipStream_t s) {
  if (g.graph) return LBM_OK;
  hipGraph_t gr = nullptr;
  BCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  lbm_status s0 = launch_single(g, 0, s);
  lbm_status s1 = s0 == LBM_OK ? launch_single(g, 1, s) : s0;
  hipError_t e = hipStreamEndCapture(s, &gr);
  if (s1 != LBM_OK || e != hipSuccess) {
    if (gr) (void)hipGraphDestroy(gr);
    return set_error(LBM_EDEVICE, "batch graph capture failed");
  }
  e = hipGraphInstantiate(&g.graph, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  if (e != hipSuccess) {
    g.graph = nullptr;
    return set_error(LBM_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  ++b->counters.captures;
  return LBM_OK;
}

// nsteps steps of group g on stream s; flips its members' current buffers as the steps do.
lbm_status enqueue_group(lbm_batch_ctx *b, Group &g, int64_t nsteps, hipStream_t s) {
  int64_t flips = 0;
  if (g.multi_k) {  // multi-step groups: K steps per launch, then single steps
    int64_t gl = nsteps;
    for (; gl >= g.multi_k; gl -= g.multi_k, ++flips) {
      BCHK(launch_step_multi_batch(g.cfg, g.args(g.phase, std::max(g.total_multi, 1)), std::max(g.total_multi, 1), s));
      g.phase ^= 1;
    }
    for (; gl > 0; --gl, ++flips) {
      lbm_status st = launch_single(g, g.phase, s);
      if (st != LBM_OK) return st;
      g.phase ^= 1;
    }
  } else {  // single steps: the captured pair replayed from phase 0
    int64_t left = nsteps;
    flips = nsteps;
    if (left > 0 && g.phase == 1) {
      lbm_status st = launch_single(g, 1, s);
      if (st != LBM_OK) return st;
      g.phase = 0;
      --left;
    }
    if (left >= 2 && b->use_graph) {
      lbm_status st = ensure_graph(b, g, s);
      if (st != LBM_OK) return st;
      for (; left >= 2; left -= 2) BCHK(hipGraphLaunch(g.graph, s));
    }
    for (; left >= 2; left -= 2) {
      lbm_status st = launch_single(g, 0, s);
      if (st == LBM_OK) st = launch_single(g, 1, s);
      if (st != LBM_OK) return st;
    }
    if (left == 1) {
      lbm_status st = launch_single(g, 0, s);
      if (st != LBM_OK) return st;
      g.phase = 1;
    }
  }
  for (lbm_ctx *m : g.members) {
    m->cur ^= (int)(flips & 1);
    m->steps += nsteps;
  }
  return LBM_OK;
}

lbm_status enqueue(lbm_batch_ctx *b, int64_t nsteps) {
  if (b->dirty) {
    lbm_status st = rebuild(b);
    if (st != LBM_OK) return st;
  }
  if (b->groups.empty() || nsteps == 0) return LBM_OK;
  b->steps += nsteps;
  size_t total = 0;
  for (const auto &g : b->groups) total += g.bytes_used;
  if (b->groups.size() == 1 || !b->concurrent || (b->bin_bytes && total > b->bin_bytes)) {
    for (auto &g : b->groups) {
      lbm_status st = enqueue_group(b, g, nsteps, b->stream);
      if (st != LBM_OK) return st;
    }
    return LBM_OK;
  }
  // fork: every group's steps on its own stream after the batch stream's work so far (table
  // uploads, member initialisation), then the batch stream waits for all of them
  BCHK(hipEventRecord(b->fork, b->stream));
  for (auto &g : b->groups) {
    if (!g.gs) {
      BCHK(hipStreamCreateWithFlags(&g.gs, hipStreamNonBlocking));
      BCHK(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
    }
    BCHK(hipStreamWaitEvent(g.gs, b->fork, 0));
    lbm_status st = enqueue_group(b, g, nsteps, g.gs);
    if (st != LBM_OK) return st;
    BCHK(hipEventRecord(g.done, g.gs));
    BCHK(hipStreamWaitEvent(b->stream, g.done, 0));
  }
  return LBM_OK;
}

}  // namespace

namespace lbm {
BatchCounters batch_counters(lbm_batch b) { return b ? b->counters : BatchCounters{}; }
hipStream_t batch_stream(lbm_batch b) { return b ? b->stream : nullptr; }
}  // namespace lbm

extern "C" {

lbm_status lbm_batch_create(lbm_batch *out) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!out) return set_error(LBM_EINVAL, "out is NULL");
    *out = nullptr;
    lbm_batch_ctx *b = new (std::nothrow) lbm_batch_ctx();
    if (!b) return set_error(LBM_ENOMEM, "host allocation failed");
    if (hipGetDevice(&b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&b->order, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
      lbm_batch_destroy(b);
      return set_error(LBM_EDEVICE, "batch stream/event creation failed");
    }
    if (const char *e = std::getenv("LBM_BATCH_GRAPH")) b->use_graph = std::atoi(e) != 0;
    if (const char *e = std::getenv("LBM_BATCH_CONCURRENT")) b->concurrent = std::atoi(e) != 0;
    if (const char *e = std::getenv("LBM_BATCH_BIN_MIB")) b->bin_bytes = (size_t)std::max(0, std::atoi(e)) << 20;
    if (const char *e = std::getenv("LBM_BATCH_STREAM_MIB")) b->stream_bytes = (size_t)std::max(0, std::atoi(e)) << 20;
    if (const char *e = std::getenv("LBM_BATCH_SIZE_CLASSES")) b->size_classes = std::atoi(e) != 0;
    *out = b;
    return LBM_OK;
  });
}

lbm_status lbm_batch_add(lbm_batch b, lbm_handle h) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!b || !h) return set_error(LBM_EINVAL, "NULL argument");
    if (h->batch) return set_error(LBM_EINVAL, "the lattice already belongs to a batch");
    if (h->device != b->device) return set_error(LBM_EINVAL, "the lattice lives on another device");
    // the batch's stream runs after everything already queued on the lattice's own stream
    BCHK(hipEventRecord(b->order, h->stream));
    BCHK(hipStreamWaitEvent(b->stream, b->order, 0));
    h->batch = b;
    h->borrowed = b->stream;
    b->members.push_back(h);
    b->dirty = true;
    return LBM_OK;
  });
}

lbm_status lbm_batch_remove(lbm_batch b, lbm_handle h) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!b || !h) return set_error(LBM_EINVAL, "NULL argument");
    auto it = std::find(b->members.begin(), b->members.end(), h);
    if (it == b->members.end()) return set_error(LBM_EINVAL, "the lattice is not a member of this batch");
    // the lattice's own stream runs after the batch work queued so far
    BCHK(hipEventRecord(b->order, b->stream));
    BCHK(hipStreamWaitEvent(h->stream, b->order, 0));
    b->members.erase(it);
    h->batch = nullptr;
    h->borrowed = nullptr;
    b->dirty = true;
    return LBM_OK;
  });
}

lbm_status lbm_batch_step(lbm_batch b, int64_t nsteps) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!b) return set_error(LBM_EINVAL, "batch is NULL");
    if (nsteps < 0) return set_error(LBM_EINVAL, "nsteps < 0");
    return enqueue(b, nsteps);
  });
}

lbm_status lbm_batch_step_timed(lbm_batch b, int64_t nsteps, double *ms) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!b) return set_error(LBM_EINVAL, "batch is NULL");
    if (nsteps < 0) return set_error(LBM_EINVAL, "nsteps < 0");
    if (b->dirty) {
      lbm_status st = rebuild(b);
      if (st != LBM_OK) return st;
    }
    if (nsteps >= 2 && b->use_graph) {  // capture outside the timed region
      size_t total = 0;
      for (const auto &g : b->groups) total += g.bytes_used;
      const bool conc = b->groups.size() > 1 && b->concurrent && !(b->bin_bytes && total > b->bin_bytes);
      for (auto &g : b->groups)
        if (!g.multi_k && g.phase == 0 && (!conc || g.gs)) {
          lbm_status st = ensure_graph(b, g, conc ? g.gs : b->stream);
          if (st != LBM_OK) return st;
        }
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
  });
}

lbm_status lbm_batch_sync(lbm_batch b) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!b) return set_error(LBM_EINVAL, "batch is NULL");
    BCHK(hipStreamSynchronize(b->stream));
    return LBM_OK;
  });
}

lbm_status lbm_batch_tune(lbm_batch b, int32_t trials) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!b) return set_error(LBM_EINVAL, "batch is NULL");
    if (b->dirty) {
      lbm_status st = rebuild(b);
      if (st != LBM_OK) return st;
    }
    const int cus = device_cu_count(b->device);
    const int reps = trials > 0 ? trials : 5;
    for (auto &g : b->groups) {
      int best = g.grid;
      float best_ms = 1e30f;
      for (int f : {1, 2, 4, 8, 16, 0}) {  // 0 = one workgroup per tile
        const int grid = f ? std::min(g.total_tiles, f * cus) : g.total_tiles;
        if (grid < 1) continue;
        // trial launches read the current buffers and write only the other ones
        const BatchArgs a = g.args(g.phase, grid);
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
      g.tuned = best;
      g.grid = best;
    }
    b->drop_graph();
    return LBM_OK;
  });
}

int32_t lbm_batch_size(lbm_batch b) { return b ? (int32_t)b->members.size() : -1; }

lbm_status lbm_batch_destroy(lbm_batch b) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!b) return LBM_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (lbm_ctx *m : b->members) {
      m->batch = nullptr;
      m->borrowed = nullptr;
    }
    b->drop_graph();
    for (auto &g : b->groups) g.release();
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->order) (void)hipEventDestroy(b->order);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return LBM_OK;
  });
}

}  // extern "C"
