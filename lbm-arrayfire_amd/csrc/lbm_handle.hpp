// lbm_handle.hpp — the lattice handle behind the C ABI's lbm_handle (shared by the ABI, batch and
// server implementations).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "lbm_capi.h"
#include "lbm_internal.hpp"

struct lbm_batch_ctx;

namespace lbm {
// The handle's small device scratch: 3 reduction results (monitors, force) in 4 doubles, then the
// two 32-bit step-counter words of the stochastic rounding.
constexpr size_t RED_BYTES = 5 * sizeof(double);
}

namespace lbm {
struct DevPool;
}

struct lbm_ctx {
  lbm_desc d{};
  int device = 0;
  int Q = 9;
  int64_t N = 0;
  int64_t stride = 0;  // elements between population arrays
  int esize = 4;
  hipStream_t stream = nullptr;
  bool own_stream = true;  // false: a stream lent at creation (not destroyed with the handle)
  lbm::DevPool *pool = nullptr;  // device buffers come from and return to this pool (job server)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void *f[2] = {nullptr, nullptr};
  int cur = 0;
  uint8_t *types = nullptr;
  double *partials = nullptr;
  double *red = nullptr;      // 3 reduction results, then (at red + 4) the two step-counter words
  double *fields = nullptr;  // 4N doubles, allocated on first use
  float *mrt_tab = nullptr;  // MFMA MRT per-lane operands
  void *probe = nullptr;     // sensor node list + gathered values
  size_t probe_bytes = 0;
  int64_t steps = 0;
  lbm_batch_ctx *batch = nullptr;     // set while the lattice is a batch member
  hipStream_t borrowed = nullptr;     // the batch's stream while a member: all work goes there
  hipGraphExec_t graph = nullptr;     // two captured steps (cur -> other -> cur)
  int graph_cur = -1;                 // buffer index the graph starts from
  bool use_graph = false;             // direct launches measure faster for single lattices
  int multi_variant = -1;             // multi-step launch shape (D2Q9 BGK 2D), -1 = single steps
  int multi_k = 0;                    // steps per multi-step launch
  lbm::Config cfg{};
  lbm::StepParams base{};

  hipStream_t strm() const { return borrowed ? borrowed : stream; }
  size_t pop_bytes() const { return (size_t)N * (size_t)Q * (size_t)esize; }       // packed
  size_t buf_bytes() const { return (size_t)stride * (size_t)Q * (size_t)esize; }   // allocated

  uint32_t *step_ctr() const { return reinterpret_cast<uint32_t *>(red + 4); }
  lbm::StepParams params(int from) const {
    lbm::StepParams p = base;
    p.fin = f[from];
    p.fout = f[from ^ 1];
    p.step_in = step_ctr() + from;
    p.step_out = step_ctr() + (from ^ 1);
    return p;
  }
};


namespace lbm {
// Sets the thread-local last-error message and returns s.
lbm_status set_error(lbm_status s, const std::string &m);
lbm_status set_error_noexcept(lbm_status s, const char *m) noexcept;
// Device buffers kept for reuse by the job server, so that a finished job's buffers serve the
// next job without a hipFree / hipMalloc pair (each about 30-40 us of host time and a device
// synchronisation).  Blocks are handed out in size classes (2^k times 1, 1.25, 1.5 or 1.75, at
// least 256 B), so lattices of nearby sizes share them; idle blocks stay allocated until trim().
// Used from one thread at a time (the server's).
struct DevPool {
  std::multimap<size_t, void *> idle;
  size_t held = 0;  // bytes in idle blocks
  static size_t size_class(size_t n);
  hipError_t get(void **p, size_t n);
  void put(void *p, size_t n);
  void trim();
  // frees every idle block except one per entry of `keep` (block sizes, each a size class)
  void trim_except(const std::vector<size_t> &keep);
};
// The blocks a lattice takes from a DevPool (sizes rounded to their classes; the transient
// node-type staging block not included), and their sum.
// fields: the lattice's field readout scratch (field_scratch_bytes) too, for jobs that keep their
// fields.
std::vector<size_t> pooled_blocks(const lbm_desc *d, bool fields);
size_t pooled_footprint(const lbm_desc *d, bool fields);
inline size_t field_scratch_bytes(int64_t N) { return 4 * (size_t)N * sizeof(double); }
// lbm_create on a lent stream (nullptr: a stream of its own) with buffers from a pool (nullptr:
// hipMalloc / hipFree), and a batch's stream
lbm_status create_lattice(const lbm_desc *d, const uint8_t *node_type, hipStream_t shared, DevPool *pool,
                          lbm_handle *out);
hipStream_t batch_stream(lbm_batch b);
// The body of a C-ABI entry point: no C++ exception crosses the ABI (a host allocation failure
// returns LBM_ENOMEM, any other exception LBM_EINTERNAL, with the message in lbm_last_error).
template <class F>
inline lbm_status abi_guard(F &&body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc &) {
    return set_error_noexcept(LBM_ENOMEM, "host allocation failed");
  } catch (const std::exception &e) {
    return set_error_noexcept(LBM_EINTERNAL, e.what());
  } catch (...) {
    return set_error_noexcept(LBM_EINTERNAL, "unknown C++ exception");
  }
}
// Number of compute units of the handle's device (cached).
int device_cu_count(int device);
// Work counters of a batch: table rewrites and step-graph captures.
struct BatchCounters {
  int64_t rebuilds = 0, captures = 0;
};
BatchCounters batch_counters(lbm_batch b);
}  // namespace lbm
