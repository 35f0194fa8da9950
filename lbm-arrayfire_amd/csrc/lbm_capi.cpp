// lbm_capi.cpp — host implementation of the C ABI declared in include/lbm_capi.h.
//
// A handle owns two population buffers (A-B pull scheme), the node-type bytes, monitor scratch
// and one HIP stream.  All kernels of a handle are enqueued on its stream; the calls that return
// host data synchronise it (af::sync() in the reference, cpp/d2q9_bgk_lid.cpp:118,182).
#include "lbm_capi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "lbm_handle.hpp"

using namespace lbm;

namespace {

thread_local std::string g_err;
int g_device = 0;

lbm_status fail(lbm_status s, const std::string &m) {
  g_err = m;
  return s;
}

#define HIPCHK(call)                                                                           \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) return fail(LBM_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

double weight(int Q, int i) { return Q == 9 ? Lat<9>::w[i] : (Q == 19 ? Lat<19>::w[i] : Lat<27>::w[i]); }

// Host-side codecs for shifted storage: F16S stores (f - w_i) * 2^15 as IEEE half, BF16S stores
// f - w_i as bfloat16, both rounded to nearest even.
uint16_t enc_half(double v, double w) {
  const float g = (float)(v - (double)(float)w);
  return __builtin_bit_cast(uint16_t, (_Float16)(g * 32768.0f));
}
double dec_half(uint16_t h, double w) {
  return (double)(float)__builtin_bit_cast(_Float16, h) / 32768.0 + (double)(float)w;
}
uint16_t enc_bf16(double v, double w) {
  const float g = (float)(v - (double)(float)w);
  return __builtin_bit_cast(uint16_t, (__bf16)g);
}
double dec_bf16(uint16_t b, double w) {
  return (double)__builtin_bit_cast(float, (uint32_t)b << 16) + (double)(float)w;
}

std::string check_desc(const lbm_desc *d) {
  if (!d) return "descriptor is NULL";
  if (d->lattice != 9 && d->lattice != 19 && d->lattice != 27)
    return "lattice must be 9 (D2Q9), 19 (D3Q19) or 27 (D3Q27)";
  if (d->nx < 1 || d->ny < 1 || d->nz < 1) return "lattice dimensions must be >= 1";
  if (d->lattice == 9 && d->nz != 1) return "D2Q9 requires nz == 1";
  const uint64_t N = (uint64_t)d->nx * (uint64_t)d->ny * (uint64_t)d->nz;
  if (N >= (1ull << 32)) return "lattice has 2^32 or more nodes";
  if (d->collision < 0 || d->collision > 2) return "unknown collision operator";
  if (d->collision != LBM_BGK && d->lattice == 9) return "MRT is implemented for D3Q19 and D3Q27";
  if (d->collision == LBM_MRT_RAW && d->lattice == 19) return "D3Q19 MRT uses the orthogonal basis (LBM_MRT_GS)";
  if (d->storage < 0 || d->storage > 3) return "unknown storage format";
  // 2-byte populations are addressed by 32-bit byte offsets from each population's base
  if ((d->storage == LBM_F16S || d->storage == LBM_BF16S) && N >= (1ull << 31))
    return "F16S / BF16S storage holds at most 2^31 - 1 nodes";
  if (d->bc_mode < 0 || d->bc_mode > 1) return "unknown boundary mode";
  if (d->mrt_engine < 0 || d->mrt_engine > 4) return "unknown MRT engine";
  const bool mfma = d->mrt_engine == LBM_MRT_MFMA_F32 || d->mrt_engine == LBM_MRT_MFMA_FUSED ||
                    d->mrt_engine == LBM_MRT_MFMA_BF16X3;
  if (mfma && d->storage == LBM_F64)
    return "the MFMA engines need FP32 arithmetic (storage F32, F16S or BF16S)";
  if (mfma && d->lattice != 27) return "the MFMA engines are implemented for D3Q27";
  if (d->meq_mode < 0 || d->meq_mode > 1) return "unknown equilibrium-moment mode";
  if (d->inlet_profile < 0 || d->inlet_profile > 1) return "unknown inlet profile";
  if (d->inlet_closure < 0 || d->inlet_closure > 1) return "unknown inlet closure";
  if (!std::isfinite(d->omega)) return "omega is not finite";
  for (int k = 0; k < 27; ++k)
    if (!std::isfinite(d->s[k])) return "MRT rate is not finite";
  if (d->inlet_profile == LBM_INLET_PARABOLIC && (d->ny < 3 || (d->lattice != 9 && d->nz < 3)))
    return "parabolic inlet needs at least 3 nodes across the channel";
  return "";
}

}  // namespace

extern "C" {

int lbm_abi_version(void) { return LBM_ABI_VERSION; }

const char *lbm_last_error(void) { return g_err.c_str(); }

lbm_status lbm_set_device(int device) {
  return lbm::abi_guard([&]() -> lbm_status {
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(LBM_EINVAL, "device index out of range");
    HIPCHK(hipSetDevice(device));
    g_device = device;
    return LBM_OK;
  });
}

size_t lbm_footprint_bytes(const lbm_desc *d) {
  if (!check_desc(d).empty()) return 0;
  const size_t N = (size_t)d->nx * d->ny * d->nz;
  int32_t gx, gy, gz, nt;
  tile_grid(d->nx, d->ny, d->nz, gx, gy, gz, nt);
  return 2 * (size_t)population_stride((int64_t)N, storage_bytes(d->storage)) * (size_t)d->lattice *
             (size_t)storage_bytes(d->storage) + types_block_bytes(d->lattice, d->bc_mode, (int64_t)N) +
         (size_t)nt * 3 * sizeof(double) + RED_BYTES;
}

}  // extern "C"

namespace lbm {
size_t DevPool::size_class(size_t n) {
  if (n <= 256) return 256;
  size_t base = 256;
  while (base * 2 <= n) base *= 2;
  for (size_t m = 4; m <= 7; ++m)
    if (base / 4 * m >= n) return base / 4 * m;
  return base * 2;
}

hipError_t DevPool::get(void **p, size_t n) {
  const size_t c = size_class(n);
  auto it = idle.find(c);
  if (it != idle.end()) {
    *p = it->second;
    idle.erase(it);
    held -= c;
    return hipSuccess;
  }
  return hipMalloc(p, c);
}

void DevPool::put(void *p, size_t n) {
  const size_t c = size_class(n);
  idle.emplace(c, p);
  held += c;
}

void DevPool::trim() {
  for (auto &kv : idle) (void)hipFree(kv.second);
  idle.clear();
  held = 0;
}

void DevPool::trim_except(const std::vector<size_t> &keep) {
  std::map<size_t, int> want;
  for (size_t c : keep) ++want[c];
  for (auto it = idle.begin(); it != idle.end();) {
    auto w = want.find(it->first);
    if (w != want.end() && w->second > 0) {
      --w->second;
      ++it;
      continue;
    }
    (void)hipFree(it->second);
    held -= it->first;
    it = idle.erase(it);
  }
}

std::vector<size_t> pooled_blocks(const lbm_desc *d, bool fields) {
  if (!check_desc(d).empty()) return {};
  const size_t N = (size_t)d->nx * d->ny * d->nz;
  int32_t gx, gy, gz, nt;
  tile_grid(d->nx, d->ny, d->nz, gx, gy, gz, nt);
  const size_t buf = DevPool::size_class((size_t)population_stride((int64_t)N, storage_bytes(d->storage)) *
                                         (size_t)d->lattice * (size_t)storage_bytes(d->storage));
  std::vector<size_t> b = {buf, buf, DevPool::size_class(types_block_bytes(d->lattice, d->bc_mode, (int64_t)N)),
                           DevPool::size_class((size_t)nt * 3 * sizeof(double)),
                           DevPool::size_class(RED_BYTES)};
  if (d->collision != LBM_BGK && d->lattice == 27) b.push_back(DevPool::size_class(64 * MFMA_TAB * sizeof(float)));
  if (fields) b.push_back(DevPool::size_class(field_scratch_bytes((int64_t)N)));
  return b;
}

size_t pooled_footprint(const lbm_desc *d, bool fields) {
  size_t fp = 0;
  for (size_t c : pooled_blocks(d, fields)) fp += c;
  return fp;
}

// lbm_create with the handle's work on the given stream (borrowed, not destroyed with the handle)
// instead of a stream of its own; nullptr creates one.  The job server creates its lattices on
// its batch's stream: a HIP stream costs about 0.6 ms to create and 0.2 ms to destroy, which
// dominated the host time of small jobs.
lbm_status create_lattice(const lbm_desc *d, const uint8_t *node_type, hipStream_t shared, DevPool *pool,
                          lbm_handle *out) {
  if (!out) return fail(LBM_EINVAL, "out is NULL");
  *out = nullptr;
  const std::string why = check_desc(d);
  if (!why.empty()) return fail(LBM_EINVAL, why);
  if (!node_type) return fail(LBM_EINVAL, "node_type is NULL");
  lbm_ctx *h = new (std::nothrow) lbm_ctx();
  if (!h) return fail(LBM_ENOMEM, "host allocation failed");
  h->d = *d;
  h->device = g_device;
  h->Q = d->lattice;
  h->N = (int64_t)d->nx * d->ny * d->nz;
  h->esize = storage_bytes(d->storage);
  h->stride = population_stride(h->N, h->esize);
  int engine = d->mrt_engine;
  if (engine == LBM_MRT_AUTO) engine = d->storage == LBM_F64 ? MRT_ENGINE_SCALAR : MRT_ENGINE_AUTO_FP32;
  h->cfg = Config{d->lattice, d->storage, d->bc_mode, d->collision, engine};

  StepParams &p = h->base;
  p.nx = d->nx;
  p.ny = d->ny;
  p.nz = d->nz;
  tile_grid(d->nx, d->ny, d->nz, p.gx, p.gy, p.gz, p.ntiles);
  p.flat = flat_tiles(d->nx) ? 1 : 0;
  make_fastdiv((uint32_t)d->nx, p.div_nx);
  make_fastdiv((uint32_t)d->ny, p.div_ny);
  p.inlet_profile = d->inlet_profile;
  p.inlet_closure = d->inlet_closure;
  p.meq_fix26 = d->collision == LBM_MRT_RAW && d->meq_mode == LBM_MEQ_REFERENCE ? 1 : 0;
  p.meq_mode = d->meq_mode;
  // Non-temporal population stores for 3D lattices stepped on their own whose buffers exceed the
  // 256 MB Infinity Cache (measured on D3Q27: +2 to +6 % from 192^3 F16S / 160^3 FP32 up, -1.4 %
  // at 160^3 F16S, -18 % at 128^3 F16S), and whose wavefronts' stores fill whole 128-byte lines
  // (rows_fill_lines: flat tiles, or rows of whole lines).  A wavefront over a row that starts
  // inside a line leaves two partial lines per store, which non-temporal writes send to memory as
  // they are (64 x 4 tiles of a 193 x 192 x 192 D3Q27 BGK F16S lattice: 15.2k MLUPS non-temporal
  // against 32.2k with ordinary stores, which the L2 merges).
  // LBM_NT_STORE=0 / 1 forces them off / on.
  h->cfg.nt = h->buf_bytes() > ((size_t)256 << 20) && rows_fill_lines(d->nx, h->esize) ? 1 : 0;
  if (const char *e = std::getenv("LBM_NT_STORE")) h->cfg.nt = std::atoi(e) != 0 ? 1 : 0;
  if (d->lattice == 9) h->cfg.nt = 0;
  p.stream_order = h->cfg.nt;
  p.N = h->N;
  p.stride = h->stride;
  p.omega = d->omega;
  for (int k = 0; k < 3; ++k) p.u_wall[k] = d->u_wall[k];
  p.u_in = d->u_in;
  p.rho_bc = d->rho_bc;
  for (int k = 0; k < 27; ++k) {
    p.s[k] = d->s[k];
    p.s32[k] = (float)d->s[k];
  }

  // Periodic wraps that carry nothing in PHYSICAL mode.  x: every node of column 0 is solid or an
  // inlet and every node of column nx - 1 solid or an outlet (the inlet replaces the populations it
  // pulls across the wrap, the outlet takes all of its populations from its west neighbour); y:
  // rows 0 and ny - 1 are solid (nodes pull nothing from beyond a wall).  Multi-step tiles hold the
  // halo positions across such a wrap instead of updating them.
  if (d->bc_mode == LBM_BC_PHYSICAL && d->nz == 1) {
    bool xcut = d->nx >= 2, ycut = d->ny >= 2;
    for (int y = 0; y < d->ny && xcut; ++y) {
      const uint8_t a = node_type[(size_t)y * d->nx], b = node_type[(size_t)y * d->nx + d->nx - 1];
      xcut = (a & (LBM_NODE_SOLID | LBM_NODE_INLET)) && (b & (LBM_NODE_SOLID | LBM_NODE_OUTLET));
    }
    for (int x = 0; x < d->nx && ycut; ++x)
      ycut = (node_type[x] & LBM_NODE_SOLID) && (node_type[(size_t)(d->ny - 1) * d->nx + x] & LBM_NODE_SOLID);
    p.wrap_cut = (xcut ? 1 : 0) | (ycut ? 2 : 0);
  }

  // Multi-step launches for the lattices that qualify: the default shape (FP64: the default
  // one-copy shape) unless LBM_MULTISTEP names another shape index, or "off" (single steps only).
  if (d->nz == 1 && d->nx < 65536 && d->ny < 65536) {  // the multi-step tiles pack x, y in 16 bits
    int v = d->storage == LBM_F64 ? MULTI_DEFAULT_F64 : MULTI_DEFAULT;
    if (const char *e = std::getenv("LBM_MULTISTEP")) v = std::strcmp(e, "off") == 0 ? -1 : std::atoi(e);
    const int k = v >= 0 ? multi_steps_per_launch(h->cfg, v) : 0;
    if (k > 1) {
      h->multi_variant = v;
      h->multi_k = k;
    }
  }

  h->pool = pool;
  auto dmalloc = [&](void **ptr, size_t n) { return pool ? pool->get(ptr, n) : hipMalloc(ptr, n); };
  auto cleanup = [&](lbm_status s, const std::string &m) {
    lbm_destroy(h);
    return fail(s, m);
  };
  if (hipSetDevice(h->device) != hipSuccess) return cleanup(LBM_EDEVICE, "hipSetDevice failed");
  if (shared) {
    h->own_stream = false;
    hipDevice_t sd = -1;
    if (hipStreamGetDevice(shared, &sd) != hipSuccess || sd != h->device)
      return cleanup(LBM_EINVAL, "the lent stream belongs to another device than the current one");
    h->stream = shared;
  } else if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    return cleanup(LBM_EDEVICE, "hipStreamCreate failed");
  }
  if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
    return cleanup(LBM_EDEVICE, "hipEventCreate failed");
  for (int b = 0; b < 2; ++b)
    if (dmalloc(&h->f[b], h->buf_bytes()) != hipSuccess)
      return cleanup(LBM_ENOMEM, "population buffer allocation failed (" + std::to_string(h->buf_bytes()) + " bytes)");
  if (dmalloc((void **)&h->types, types_block_bytes(h->Q, d->bc_mode, h->N)) != hipSuccess)
    return cleanup(LBM_ENOMEM, "node-type allocation failed");
  if (dmalloc((void **)&h->partials, (size_t)p.ntiles * 3 * sizeof(double)) != hipSuccess ||
      dmalloc((void **)&h->red, RED_BYTES) != hipSuccess)
    return cleanup(LBM_ENOMEM, "monitor scratch allocation failed");
  if (hipMemsetAsync(h->step_ctr(), 0, 2 * sizeof(uint32_t), h->stream) != hipSuccess)
    return cleanup(LBM_EDEVICE, "step counter initialisation failed");
  // Node types: the user's bits as given; PHYSICAL mode then marks, on the device, the non-solid
  // nodes whose pull needs the boundary path (NODE_SLOW).
  {
    uint8_t *raw = nullptr;
    if (dmalloc((void **)&raw, (size_t)h->N) != hipSuccess) return cleanup(LBM_ENOMEM, "node-type staging failed");
    hipError_t e = hipMemcpyAsync(raw, node_type, (size_t)h->N, hipMemcpyHostToDevice, h->stream);
    uint32_t *links = wall_links_kept(h->Q, d->bc_mode) ? (uint32_t *)(h->types + wall_links_offset(h->N)) : nullptr;
    if (e == hipSuccess) e = launch_prepare_types(h->Q, d->bc_mode, d->nx, d->ny, d->nz, raw, h->types, links, h->stream);
    p.links = links;
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (pool)
      pool->put(raw, (size_t)h->N);
    else
      (void)hipFree(raw);
    if (e != hipSuccess) return cleanup(LBM_EDEVICE, std::string("node-type upload failed: ") + hipGetErrorString(e));
  }
  p.type = h->types;
  if (d->collision != LBM_BGK && d->lattice == 27) {
    std::vector<float> tab(64 * MFMA_TAB);
    if (uses_mfma(h->cfg)) build_mfma_table(d->collision, d->s, h->cfg.engine, tab.data());
    if (dmalloc((void **)&h->mrt_tab, tab.size() * sizeof(float)) != hipSuccess)
      return cleanup(LBM_ENOMEM, "MRT table allocation failed");
    if (hipMemcpy(h->mrt_tab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(LBM_EDEVICE, "MRT table upload failed");
    p.mrt_tab = h->mrt_tab;
  }
  *out = h;
  return LBM_OK;
}
}  // namespace lbm

extern "C" {

lbm_status lbm_create(const lbm_desc *d, const uint8_t *node_type, lbm_handle *out) {
  return lbm::abi_guard([&]() -> lbm_status { return lbm::create_lattice(d, node_type, nullptr, nullptr, out); });
}

lbm_status lbm_destroy(lbm_handle h) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h) return LBM_OK;
    if (h->batch) return fail(LBM_EINVAL, "remove the lattice from its batch before destroying it");
    if (h->graph) (void)hipGraphExecDestroy(h->graph);
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    auto dfree = [&](void *ptr, size_t n) {
      if (!ptr) return;
      if (h->pool)
        h->pool->put(ptr, n);
      else
        (void)hipFree(ptr);
    };
    for (int b = 0; b < 2; ++b) dfree(h->f[b], h->buf_bytes());
    dfree(h->types, types_block_bytes(h->Q, h->d.bc_mode, h->N));
    dfree(h->partials, (size_t)h->base.ntiles * 3 * sizeof(double));
    dfree(h->red, RED_BYTES);
    dfree(h->mrt_tab, 64 * MFMA_TAB * sizeof(float));
    dfree(h->fields, field_scratch_bytes(h->N));
    if (h->probe) (void)hipFree(h->probe);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream && h->own_stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return LBM_OK;
  });
}

// Device scratch of the field readout (rho, ux, uy, uz as doubles), allocated on first use; a
// lattice of the job server takes it from the server's pool (and returns it there).
static lbm_status ensure_fields(lbm_handle h) {
  if (!h->fields) {
    const size_t n = field_scratch_bytes(h->N);
    HIPCHK(h->pool ? h->pool->get((void **)&h->fields, n) : hipMalloc((void **)&h->fields, n));
  }
  return LBM_OK;
}

lbm_status lbm_init_fields(lbm_handle h, const double *rho, const double *ux, const double *uy,
                           const double *uz) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h) return fail(LBM_EINVAL, "handle is NULL");
    HIPCHK(hipSetDevice(h->device));
    lbm_status st = ensure_fields(h);
    if (st != LBM_OK) return st;
    const double *src[4] = {rho, ux, uy, uz};
    const double *dev[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int k = 0; k < 4; ++k)
      if (src[k]) {
        double *dst = h->fields + (size_t)k * h->N;
        HIPCHK(hipMemcpyAsync(dst, src[k], (size_t)h->N * sizeof(double), hipMemcpyHostToDevice, h->strm()));
        dev[k] = dst;
      }
    HIPCHK(launch_init_fields(h->cfg, h->params(h->cur), dev[0], dev[1], dev[2], dev[3], h->strm()));
    HIPCHK(hipStreamSynchronize(h->strm()));
    h->steps = 0;
    return LBM_OK;
  });
}

lbm_status lbm_init_equilibrium(lbm_handle h, double rho0, const double u0[3], uint64_t seed,
                                double noise_amp) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h) return fail(LBM_EINVAL, "handle is NULL");
    HIPCHK(hipSetDevice(h->device));
    const double u[3] = {u0 ? u0[0] : 0.0, u0 ? u0[1] : 0.0, u0 ? u0[2] : 0.0};
    HIPCHK(launch_init_equilibrium(h->cfg, h->params(h->cur), rho0, u, seed, noise_amp, h->strm()));
    HIPCHK(hipStreamSynchronize(h->strm()));
    h->steps = 0;
    return LBM_OK;
  });
}

lbm_status lbm_init_reference(lbm_handle h, double rho0) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h) return fail(LBM_EINVAL, "handle is NULL");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(launch_init_reference(h->cfg, h->params(h->cur), rho0, h->strm()));
    HIPCHK(hipStreamSynchronize(h->strm()));
    h->steps = 0;
    return LBM_OK;
  });
}

// Steps are enqueued as replays of a captured two-step graph (cur -> other -> cur); an odd step
// or a graph that starts from the other buffer is launched directly.
static lbm_status enqueue_steps(lbm_handle h, int64_t nsteps) {
  auto direct = [&]() -> lbm_status {
    HIPCHK(launch_step(h->cfg, h->params(h->cur), h->strm()));
    h->cur ^= 1;
    h->steps++;
    return LBM_OK;
  };
  if (nsteps >= 4 && !h->graph && h->use_graph) {
    hipStream_t cs = h->strm();
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    hipError_t e1 = launch_step(h->cfg, h->params(h->cur), cs);
    hipError_t e2 = launch_step(h->cfg, h->params(h->cur ^ 1), cs);
    hipError_t e3 = hipStreamEndCapture(cs, &g);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(LBM_EDEVICE, "step graph capture failed");
    }
    hipError_t e4 = hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e4 != hipSuccess) {
      h->graph = nullptr;
      return fail(LBM_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e4));
    }
    h->graph_cur = h->cur;
  }
  int64_t left = nsteps;
  if (h->multi_k > 1)
    for (; left >= h->multi_k; left -= h->multi_k) {
      HIPCHK(launch_step_multi(h->cfg, h->multi_variant, h->params(h->cur), h->strm()));
      h->cur ^= 1;
      h->steps += h->multi_k;
    }
  if (h->graph && left >= 2 && h->cur != h->graph_cur) {
    lbm_status st = direct();
    if (st != LBM_OK) return st;
    --left;
  }
  if (h->graph)
    for (; left >= 2; left -= 2) {
      HIPCHK(hipGraphLaunch(h->graph, h->strm()));
      h->steps += 2;
    }
  for (; left > 0; --left) {
    lbm_status st = direct();
    if (st != LBM_OK) return st;
  }
  return LBM_OK;
}

lbm_status lbm_step(lbm_handle h, int64_t nsteps) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h) return fail(LBM_EINVAL, "handle is NULL");
    if (nsteps < 0) return fail(LBM_EINVAL, "nsteps < 0");
    if (h->batch) return fail(LBM_EINVAL, "the lattice is a batch member: step the batch");
    HIPCHK(hipSetDevice(h->device));
    return enqueue_steps(h, nsteps);
  });
}

lbm_status lbm_step_timed(lbm_handle h, int64_t nsteps, double *ms) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h) return fail(LBM_EINVAL, "handle is NULL");
    if (nsteps < 0) return fail(LBM_EINVAL, "nsteps < 0");
    if (h->batch) return fail(LBM_EINVAL, "the lattice is a batch member: step the batch");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipEventRecord(h->ev0, h->strm()));
    lbm_status st = enqueue_steps(h, nsteps);
    if (st != LBM_OK) return st;
    HIPCHK(hipEventRecord(h->ev1, h->strm()));
    HIPCHK(hipEventSynchronize(h->ev1));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, h->ev0, h->ev1));
    if (ms) *ms = (double)t;
    return LBM_OK;
  });
}

lbm_status lbm_sync(lbm_handle h) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h) return fail(LBM_EINVAL, "handle is NULL");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->strm()));
    return LBM_OK;
  });
}

int64_t lbm_steps_done(lbm_handle h) { return h ? h->steps : -1; }

lbm_status lbm_check_reciprocal(uint32_t lo, uint32_t hi, uint64_t *mismatches) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!mismatches) return fail(LBM_EINVAL, "mismatches is NULL");
    HIPCHK(hipSetDevice(g_device));
    unsigned long long *d = nullptr;
    HIPCHK(hipMalloc((void **)&d, sizeof(*d)));
    unsigned long long n = 0;
    hipError_t e = hipMemcpy(d, &n, sizeof(n), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_check_reciprocal(lo, hi, d, nullptr);
    if (e == hipSuccess) e = hipMemcpy(&n, d, sizeof(n), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(LBM_EDEVICE, std::string("reciprocal check failed: ") + hipGetErrorString(e));
    *mismatches = (uint64_t)n;
    return LBM_OK;
  });
}

int32_t lbm_steps_per_launch(lbm_handle h) { return h ? (h->multi_k > 1 ? h->multi_k : 1) : -1; }

lbm_status lbm_get_fields(lbm_handle h, double *rho, double *ux, double *uy, double *uz) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h) return fail(LBM_EINVAL, "handle is NULL");
    HIPCHK(hipSetDevice(h->device));
    lbm_status st = ensure_fields(h);
    if (st != LBM_OK) return st;
    double *dst[4] = {rho, ux, uy, uz};
    double *dev[4];
    for (int k = 0; k < 4; ++k) dev[k] = dst[k] ? h->fields + (size_t)k * h->N : nullptr;
    HIPCHK(launch_fields(h->cfg, h->params(h->cur), dev[0], dev[1], dev[2], dev[3], h->strm()));
    for (int k = 0; k < 4; ++k)
      if (dst[k])
        HIPCHK(hipMemcpyAsync(dst[k], dev[k], (size_t)h->N * sizeof(double), hipMemcpyDeviceToHost, h->strm()));
    HIPCHK(hipStreamSynchronize(h->strm()));
    return LBM_OK;
  });
}

lbm_status lbm_get_fields_at(lbm_handle h, const int64_t *nodes, int32_t count, double *rho, double *ux,
                             double *uy, double *uz) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h || (!nodes && count > 0) || count < 0) return fail(LBM_EINVAL, "bad argument");
    if (count == 0) return LBM_OK;
    for (int32_t k = 0; k < count; ++k)
      if (nodes[k] < 0 || nodes[k] >= h->N) return fail(LBM_EINVAL, "node index out of range");
    HIPCHK(hipSetDevice(h->device));
    lbm_status st = ensure_fields(h);
    if (st != LBM_OK) return st;
    const size_t need = (size_t)count * (sizeof(int64_t) + 4 * sizeof(double));
    if (need > h->probe_bytes) {
      if (h->probe) HIPCHK(hipFree(h->probe));
      h->probe = nullptr;
      HIPCHK(hipMalloc(&h->probe, need));
      h->probe_bytes = need;
    }
    int64_t *dnodes = (int64_t *)h->probe;
    double *dout = (double *)((char *)h->probe + (size_t)count * sizeof(int64_t));
    HIPCHK(hipMemcpyAsync(dnodes, nodes, (size_t)count * sizeof(int64_t), hipMemcpyHostToDevice, h->strm()));
    double *f = h->fields;
    HIPCHK(launch_fields(h->cfg, h->params(h->cur), f, f + h->N, f + 2 * h->N, f + 3 * h->N, h->strm()));
    HIPCHK(launch_gather_fields(f, h->N, dnodes, count, dout, h->strm()));
    std::vector<double> host((size_t)4 * count);
    HIPCHK(hipMemcpyAsync(host.data(), dout, host.size() * sizeof(double), hipMemcpyDeviceToHost, h->strm()));
    HIPCHK(hipStreamSynchronize(h->strm()));
    double *dst[4] = {rho, ux, uy, uz};
    for (int a = 0; a < 4; ++a)
      if (dst[a]) std::copy(host.begin() + (long)a * count, host.begin() + (long)(a + 1) * count, dst[a]);
    return LBM_OK;
  });
}

// the state: the populations, then the step count (uint32) and 4 zero bytes
size_t lbm_state_bytes(lbm_handle h) { return h ? h->pop_bytes() + 8 : 0; }

lbm_status lbm_get_state(lbm_handle h, void *dst) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h || !dst) return fail(LBM_EINVAL, "NULL argument");
    HIPCHK(hipSetDevice(h->device));
    const size_t row = (size_t)h->N * h->esize, pitch = (size_t)h->stride * h->esize;
    HIPCHK(hipMemcpy2DAsync(dst, row, h->f[h->cur], pitch, row, (size_t)h->Q, hipMemcpyDeviceToHost, h->strm()));
    uint32_t tail[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(tail, h->step_ctr() + h->cur, sizeof(uint32_t), hipMemcpyDeviceToHost, h->strm()));
    HIPCHK(hipStreamSynchronize(h->strm()));
    std::memcpy((char *)dst + h->pop_bytes(), tail, sizeof(tail));
    return LBM_OK;
  });
}

lbm_status lbm_set_state(lbm_handle h, const void *src) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h || !src) return fail(LBM_EINVAL, "NULL argument");
    HIPCHK(hipSetDevice(h->device));
    // Both buffers: in PHYSICAL mode solid nodes are never rewritten by a step.
    const size_t row = (size_t)h->N * h->esize, pitch = (size_t)h->stride * h->esize;
    for (int b = 0; b < 2; ++b)
      HIPCHK(hipMemcpy2DAsync(h->f[b], pitch, src, row, row, (size_t)h->Q, hipMemcpyHostToDevice, h->strm()));
    uint32_t ctr[2];
    std::memcpy(&ctr[0], (const char *)src + h->pop_bytes(), sizeof(uint32_t));
    ctr[1] = ctr[0];
    HIPCHK(hipMemcpyAsync(h->step_ctr(), ctr, sizeof(ctr), hipMemcpyHostToDevice, h->strm()));
    HIPCHK(hipStreamSynchronize(h->strm()));
    return LBM_OK;
  });
}

lbm_status lbm_get_populations(lbm_handle h, double *f) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h || !f) return fail(LBM_EINVAL, "NULL argument");
    std::vector<unsigned char> raw(h->pop_bytes() + 8);
    lbm_status st = lbm_get_state(h, raw.data());
    if (st != LBM_OK) return st;
    const int64_t N = h->N;
    for (int i = 0; i < h->Q; ++i) {
      const double w = weight(h->Q, i);
      for (int64_t n = 0; n < N; ++n) {
        const size_t k = (size_t)(n + N * i);
        switch (h->d.storage) {
          case LBM_F64: f[k] = reinterpret_cast<const double *>(raw.data())[k]; break;
          case LBM_F32: f[k] = reinterpret_cast<const float *>(raw.data())[k]; break;
          case LBM_F16S: f[k] = dec_half(reinterpret_cast<const uint16_t *>(raw.data())[k], w); break;
          default: f[k] = dec_bf16(reinterpret_cast<const uint16_t *>(raw.data())[k], w); break;
        }
      }
    }
    return LBM_OK;
  });
}

lbm_status lbm_set_populations(lbm_handle h, const double *f) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h || !f) return fail(LBM_EINVAL, "NULL argument");
    std::vector<unsigned char> raw(h->pop_bytes() + 8, 0);
    const int64_t N = h->N;
    for (int i = 0; i < h->Q; ++i) {
      const double w = weight(h->Q, i);
      for (int64_t n = 0; n < N; ++n) {
        const size_t k = (size_t)(n + N * i);
        switch (h->d.storage) {
          case LBM_F64: reinterpret_cast<double *>(raw.data())[k] = f[k]; break;
          case LBM_F32: reinterpret_cast<float *>(raw.data())[k] = (float)f[k]; break;
          case LBM_F16S: reinterpret_cast<uint16_t *>(raw.data())[k] = enc_half(f[k], w); break;
          default: reinterpret_cast<uint16_t *>(raw.data())[k] = enc_bf16(f[k], w); break;
        }
      }
    }
    return lbm_set_state(h, raw.data());  // the zero tail restarts the step count
  });
}

lbm_status lbm_monitors(lbm_handle h, double *mass, double *kinetic_energy, int64_t *nonfinite) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h) return fail(LBM_EINVAL, "handle is NULL");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(launch_monitors(h->cfg, h->params(h->cur), h->partials, h->red, h->strm()));
    double out[3];
    HIPCHK(hipMemcpyAsync(out, h->red, 3 * sizeof(double), hipMemcpyDeviceToHost, h->strm()));
    HIPCHK(hipStreamSynchronize(h->strm()));
    if (mass) *mass = out[0];
    if (kinetic_energy) *kinetic_energy = out[1];
    if (nonfinite) *nonfinite = (int64_t)out[2];
    return out[2] > 0 ? fail(LBM_EUNSTABLE, "non-finite populations") : LBM_OK;
  });
}

lbm_status lbm_force(lbm_handle h, const int32_t box[6], double force[3]) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h || !force) return fail(LBM_EINVAL, "NULL argument");
    int b[6] = {0, h->d.nx, 0, h->d.ny, 0, h->d.nz};
    if (box)
      for (int k = 0; k < 6; ++k) b[k] = box[k];
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(launch_force(h->cfg, h->params(h->cur), b, h->partials, h->red, h->strm()));
    HIPCHK(hipMemcpyAsync(force, h->red, 3 * sizeof(double), hipMemcpyDeviceToHost, h->strm()));
    HIPCHK(hipStreamSynchronize(h->strm()));
    return LBM_OK;
  });
}

}  // extern "C"

namespace lbm {
lbm_status set_error(lbm_status s, const std::string &m) { return fail(s, m); }

lbm_status set_error_noexcept(lbm_status s, const char *m) noexcept {
  try {
    g_err = m;
  } catch (...) {
    g_err.clear();
  }
  return s;
}

int64_t population_stride(int64_t N, int esize) {
  static const bool pad = [] {
    const char *e = std::getenv("LBM_POP_PAD");
    return !(e && e[0] == '0');
  }();
  if (!pad || N * esize < (1 << 20)) return N;  // small arrays live in cache anyway
  const int64_t bytes = ((N * esize + 4095) / 4096) * 4096 + 4096 + 256;
  return bytes / esize;
}

int device_cu_count(int device) {
  static int cached[64] = {0};
  if (device < 0 || device >= 64) return 256;
  if (!cached[device]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0) n = 256;
    cached[device] = n;
  }
  return cached[device];
}
}  // namespace lbm
