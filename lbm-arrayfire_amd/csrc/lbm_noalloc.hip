// lbm_noalloc.hip — lbm_get_fields_at_noalloc: sensor readout that allocates nothing.
//
// lbm_get_fields_at computes the four fields of the whole lattice into a device buffer of 4 N
// doubles (allocated on first use) and gathers the selected nodes from it through a node-list
// buffer it also allocates.  Callers that may not allocate device memory (code under HIP graph
// capture, tight serving loops) use this variant instead: it computes the fields at the listed
// nodes only, with the node list and the results staged in `scratch`, a file-scope device array
// of fixed size (every GPU the library runs on holds one until the process ends).  Values are
// those of lbm_get_fields_at bit for bit (the same per-node code, fields_of_node).
#include <mutex>

#include "lbm_handle.hpp"
#include "lbm_step.hpp"

namespace lbm {

// scratch layout: nodes[LBM_NOALLOC_MAX_NODES] (int64) | out[4][LBM_NOALLOC_MAX_NODES] (double),
// 640 KiB.  Every call writes the slots it reads (the node list before the kernel, the results
// before their download), so nothing has to start at zero.
constexpr size_t kScratchBytes = (size_t)LBM_NOALLOC_MAX_NODES * (sizeof(int64_t) + 4 * sizeof(double));
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once per device; the array's
// name is only a host-side stand-in).
hipError_t scratch_of(int device, void **out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

// One thread per listed node: out[a * count + k] = field a at nodes[k].
template <int Q, int ST, int BC>
__global__ void __launch_bounds__(BLOCK_THREADS) k_fields_at(StepParams p, const int64_t *nodes, int count,
                                                             double *out) {
  const int k = (int)(blockIdx.x * BLOCK_THREADS + threadIdx.x);
  if (k >= count) return;
  const int64_t n = nodes[k];
  NodePos q;
  q.x = (int)(n % p.nx);
  q.y = (int)((n / p.nx) % p.ny);
  q.z = (int)(n / ((int64_t)p.nx * p.ny));
  q.n = (uint32_t)n;
  q.valid = true;
  double r[4];
  fields_of_node<Q, ST, BC>(p, q, r);
  for (int a = 0; a < 4; ++a) out[(size_t)a * count + k] = r[a];
}

hipError_t launch_fields_at(const Config &c, const StepParams &p, const int64_t *nodes, int count, double *out,
                            hipStream_t s) {
  return with_storage(c, [&](auto Qc, auto STc) -> hipError_t {
    constexpr int Q = decltype(Qc)::value, ST = decltype(STc)::value;
    const dim3 grid((unsigned)((count + BLOCK_THREADS - 1) / BLOCK_THREADS));
    if (c.bc == BC_REF)
      hipLaunchKernelGGL((k_fields_at<Q, ST, BC_REF>), grid, dim3(BLOCK_THREADS), 0, s, p, nodes, count, out);
    else
      hipLaunchKernelGGL((k_fields_at<Q, ST, BC_PHYS>), grid, dim3(BLOCK_THREADS), 0, s, p, nodes, count, out);
    return hipGetLastError();
  });
}

}  // namespace
}  // namespace lbm

using namespace lbm;

extern "C" lbm_status lbm_get_fields_at_noalloc(lbm_handle h, const int64_t *nodes, int32_t count, double *rho,
                                                double *ux, double *uy, double *uz) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h || (!nodes && count > 0) || count < 0) return set_error(LBM_EINVAL, "bad argument");
    if (count > LBM_NOALLOC_MAX_NODES)
      return set_error(LBM_EINVAL, "lbm_get_fields_at_noalloc reads at most LBM_NOALLOC_MAX_NODES (" +
                                       std::to_string(LBM_NOALLOC_MAX_NODES) + ") nodes per call; " +
                                       std::to_string(count) + " requested");
    if (count == 0) return LBM_OK;
    for (int32_t k = 0; k < count; ++k)
      if (nodes[k] < 0 || nodes[k] >= h->N) return set_error(LBM_EINVAL, "node index out of range");
    auto chk = [](hipError_t e, const char *what) -> lbm_status {
      return e == hipSuccess ? LBM_OK : set_error(LBM_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
    };
    lbm_status st = chk(hipSetDevice(h->device), "hipSetDevice");
    if (st != LBM_OK) return st;
    void *base = nullptr;
    if ((st = chk(scratch_of(h->device, &base), "hipGetSymbolAddress(scratch)")) != LBM_OK) return st;
    int64_t *dnodes = (int64_t *)base;
    double *dout = (double *)((char *)base + (size_t)LBM_NOALLOC_MAX_NODES * sizeof(int64_t));
    hipStream_t s = h->strm();
    if ((st = chk(hipMemcpyAsync(dnodes, nodes, (size_t)count * sizeof(int64_t), hipMemcpyHostToDevice, s),
                  "node list upload")) != LBM_OK)
      return st;
    if ((st = chk(launch_fields_at(h->cfg, h->params(h->cur), dnodes, count, dout, s), "field kernel")) != LBM_OK)
      return st;
    double *dst[4] = {rho, ux, uy, uz};
    for (int a = 0; a < 4; ++a)
      if (dst[a] && (st = chk(hipMemcpyAsync(dst[a], dout + (size_t)a * count, (size_t)count * sizeof(double),
                                             hipMemcpyDeviceToHost, s),
                              "field download")) != LBM_OK)
        return st;
    return chk(hipStreamSynchronize(s), "hipStreamSynchronize");
  });
}
