// lbm_server.cpp — continuous-batching job server over lbm_batch.
//
// No reference counterpart: each reference program runs one lattice to completion.  Jobs
// (descriptor, geometry, initial state, step count) are queued on submit and admitted while their
// device footprint fits the HBM budget (FIFO with backfill: a smaller job may pass a larger one
// that does not fit yet, unless the head has been passed over too often).  Every tick advances
// all resident jobs together by k = min(steps_per_tick, fewest steps any resident job has left)
// steps in one batched launch sequence, so jobs that arrive while others run join the next tick
// and their work overlaps on the device.  Finished jobs are retired (monitors, optional fields)
// and their memory is returned to the budget.  Every check_every ticks the monitors of all
// resident jobs are read and a job with non-finite populations is retired as LBM_EUNSTABLE
// without touching the others.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "lbm_handle.hpp"

using namespace lbm;

namespace {

struct Job {
  int64_t id = 0;
  lbm_job_spec spec{};
  std::vector<uint8_t> types;
  lbm_handle h = nullptr;
  size_t footprint = 0;
  int64_t done = 0;
  int passed_over = 0;
  double submit_s = 0, admit_s = 0;
};

struct Fields {
  std::vector<double> rho, ux, uy, uz;
};

}  // namespace

struct lbm_server_ctx {
  size_t budget = 0, used = 0;
  int32_t steps_per_tick = 100;
  int32_t check_every = 16;
  int64_t next_id = 1, ticks = 0;
  std::deque<Job *> queue;
  std::vector<Job *> active;
  std::vector<lbm_job_result> finished;
  std::map<int64_t, Fields> kept;
  lbm_batch batch = nullptr;
  lbm::DevPool pool;  // device buffers of retired jobs, reused by later ones
  lbm_server_stats stats{};
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

constexpr int kMaxPassOver = 64;

lbm_status retire(lbm_server_ctx *s, Job *j, bool check_fields) {
  const double t_in = s->now();
  lbm_job_result r{};
  r.job_id = j->id;
  r.submit_s = j->submit_s;
  r.admit_s = j->admit_s;
  r.nodes = (int64_t)j->spec.desc.nx * j->spec.desc.ny * j->spec.desc.nz;
  r.steps = j->done;
  double mass = 0, ke = 0;
  int64_t bad = 0;
  lbm_status st = lbm_monitors(j->h, &mass, &ke, &bad);
  r.status = bad ? LBM_EUNSTABLE : (st == LBM_OK ? LBM_OK : st);
  r.mass = mass;
  r.kinetic_energy = ke;
  if (check_fields && j->spec.want_fields && r.status == LBM_OK) {
    Fields f;
    const size_t N = (size_t)r.nodes;
    f.rho.resize(N);
    f.ux.resize(N);
    f.uy.resize(N);
    f.uz.resize(N);
    // a failed readout is the job's result: no fields are kept for it
    const lbm_status fs = lbm_get_fields(j->h, f.rho.data(), f.ux.data(), f.uy.data(), f.uz.data());
    if (fs == LBM_OK) s->kept[j->id] = std::move(f);
    else r.status = fs;
  }
  lbm_batch_remove(s->batch, j->h);
  lbm_destroy(j->h);
  s->used -= j->footprint;
  r.finish_s = s->now();
  s->finished.push_back(r);
  ++s->stats.retired;
  if (r.status == LBM_EUNSTABLE) ++s->stats.unstable;
  s->stats.retire_s += r.finish_s - t_in;
  delete j;
  return LBM_OK;
}

lbm_status admit(lbm_server_ctx *s) {
  const double t_in = s->now();
  lbm_status st = LBM_OK;
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    Job *j = *it;
    if (s->used + j->footprint > s->budget) {
      if (++j->passed_over > kMaxPassOver) break;  // stop backfilling past a starving head
      ++it;
      continue;
    }
    // on the batch's stream: the job's work runs there anyway once it joins the batch, and a
    // stream of its own would cost more host time to create and destroy than a small job steps
    // idle pooled blocks count against the budget: when the job needs the room, release those it
    // will not reuse, so that once it has taken the rest the device holds used + footprint, within
    // the budget as checked above
    if (s->used + j->footprint + s->pool.held > s->budget)
      s->pool.trim_except(lbm::pooled_blocks(&j->spec.desc, j->spec.want_fields != 0));
    st = lbm::abi_guard([&]() -> lbm_status {
      return lbm::create_lattice(&j->spec.desc, j->types.data(), lbm::batch_stream(s->batch), &s->pool, &j->h);
    });
    if (st == LBM_ENOMEM) {  // the device is fuller than the budget says: wait
      s->pool.trim();
      st = LBM_OK;
      break;
    }
    if (st != LBM_OK) {  // bad job: report and drop it
      lbm_job_result r{};
      r.job_id = j->id;
      r.status = st;
      r.submit_s = j->submit_s;
      r.admit_s = r.finish_s = s->now();
      s->finished.push_back(r);
      delete j;
      it = s->queue.erase(it);
      st = LBM_OK;
      continue;
    }
    st = lbm_init_equilibrium(j->h, j->spec.rho0, j->spec.u0, j->spec.seed, j->spec.noise_amp);
    if (st == LBM_OK) st = lbm_batch_add(s->batch, j->h);
    if (st != LBM_OK) {
      lbm_destroy(j->h);
      j->h = nullptr;
      break;
    }
    s->used += j->footprint;
    j->admit_s = s->now();
    s->active.push_back(j);
    ++s->stats.admitted;
    it = s->queue.erase(it);
  }
  s->stats.admit_s += s->now() - t_in;
  return st;
}

}  // namespace

extern "C" {

lbm_status lbm_server_create(size_t budget_bytes, int32_t steps_per_tick, lbm_server *out) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!out) return set_error(LBM_EINVAL, "out is NULL");
    *out = nullptr;
    if (budget_bytes == 0 || steps_per_tick < 1) return set_error(LBM_EINVAL, "budget and steps_per_tick must be positive");
    lbm_server_ctx *s = new (std::nothrow) lbm_server_ctx();
    if (!s) return set_error(LBM_ENOMEM, "host allocation failed");
    s->budget = budget_bytes;
    s->steps_per_tick = steps_per_tick;
    lbm_status st = lbm_batch_create(&s->batch);
    if (st != LBM_OK) {
      delete s;
      return st;
    }
    *out = s;
    return LBM_OK;
  });
}

lbm_status lbm_server_submit(lbm_server s, const lbm_job_spec *spec, int64_t *job_id) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!s || !spec || !spec->node_type) return set_error(LBM_EINVAL, "NULL argument");
    if (spec->nsteps < 0) return set_error(LBM_EINVAL, "nsteps < 0");
    // buffers come from the server's pool; a job that keeps its fields also holds the field
    // readout scratch (4 doubles per node) when it retires
    const size_t fp = lbm::pooled_footprint(&spec->desc, spec->want_fields != 0);
    if (fp == 0) return set_error(LBM_EINVAL, "invalid job descriptor");
    if (fp > s->budget) return set_error(LBM_ENOMEM, "job footprint exceeds the server's HBM budget");
    Job *j = new (std::nothrow) Job();
    if (!j) return set_error(LBM_ENOMEM, "host allocation failed");
    j->id = s->next_id++;
    j->spec = *spec;
    const size_t N = (size_t)spec->desc.nx * spec->desc.ny * spec->desc.nz;
    j->types.assign(spec->node_type, spec->node_type + N);
    j->spec.node_type = nullptr;
    j->footprint = fp;
    j->submit_s = s->now();
    s->queue.push_back(j);
    if (job_id) *job_id = j->id;
    return LBM_OK;
  });
}

lbm_status lbm_server_tick(lbm_server s, int32_t *active, int32_t *queued) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!s) return set_error(LBM_EINVAL, "server is NULL");
    lbm_status st = admit(s);
    if (st != LBM_OK) return st;
    // zero-step jobs finish at admission
    for (size_t i = 0; i < s->active.size();) {
      Job *j = s->active[i];
      if (j->done >= j->spec.nsteps) {
        s->active.erase(s->active.begin() + (long)i);
        retire(s, j, true);
      } else {
        ++i;
      }
    }
    if (!s->active.empty()) {
      int64_t k = s->steps_per_tick;
      for (Job *j : s->active) k = std::min<int64_t>(k, j->spec.nsteps - j->done);
      const double t_launch = s->now();
      st = lbm_batch_step(s->batch, k);
      if (st != LBM_OK) return st;
      const double t_wait = s->now();
      st = lbm_batch_sync(s->batch);
      if (st != LBM_OK) return st;
      const double t_done = s->now();
      s->stats.launch_s += t_wait - t_launch;
      s->stats.wait_s += t_done - t_wait;
      ++s->stats.ticks;
      s->stats.steps += k;
      ++s->ticks;
      const bool check = s->check_every > 0 && s->ticks % s->check_every == 0;
      for (size_t i = 0; i < s->active.size();) {
        Job *j = s->active[i];
        j->done += k;
        bool bad = false;
        if (check && j->done < j->spec.nsteps) {
          const double t_check = s->now();
          int64_t nf = 0;
          lbm_monitors(j->h, nullptr, nullptr, &nf);
          bad = nf > 0;
          s->stats.check_s += s->now() - t_check;
        }
        if (j->done >= j->spec.nsteps || bad) {
          s->active.erase(s->active.begin() + (long)i);
          retire(s, j, !bad);
        } else {
          ++i;
        }
      }
      st = admit(s);  // memory freed by retired jobs goes to waiting jobs right away
      if (st != LBM_OK) return st;
    }
    if (active) *active = (int32_t)s->active.size();
    if (queued) *queued = (int32_t)s->queue.size();
    return LBM_OK;
  });
}

lbm_status lbm_server_poll(lbm_server s, lbm_job_result *out, int32_t cap, int32_t *n) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!s || (!out && cap > 0)) return set_error(LBM_EINVAL, "NULL argument");
    const int32_t k = (int32_t)std::min<size_t>((size_t)std::max(cap, 0), s->finished.size());
    for (int32_t i = 0; i < k; ++i) out[i] = s->finished[(size_t)i];
    s->finished.erase(s->finished.begin(), s->finished.begin() + k);
    if (n) *n = k;
    return LBM_OK;
  });
}

lbm_status lbm_server_take(lbm_server s, int64_t job_id, double *rho, double *ux, double *uy, double *uz) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!s) return set_error(LBM_EINVAL, "server is NULL");
    auto it = s->kept.find(job_id);
    if (it == s->kept.end()) return set_error(LBM_EINVAL, "no kept fields for this job");
    const Fields &f = it->second;
    if (rho) std::copy(f.rho.begin(), f.rho.end(), rho);
    if (ux) std::copy(f.ux.begin(), f.ux.end(), ux);
    if (uy) std::copy(f.uy.begin(), f.uy.end(), uy);
    if (uz) std::copy(f.uz.begin(), f.uz.end(), uz);
    s->kept.erase(it);
    return LBM_OK;
  });
}

lbm_status lbm_server_cancel(lbm_server s, int64_t job_id) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!s) return set_error(LBM_EINVAL, "server is NULL");
    for (auto it = s->queue.begin(); it != s->queue.end(); ++it)
      if ((*it)->id == job_id) {
        delete *it;
        s->queue.erase(it);
        return LBM_OK;
      }
    for (size_t i = 0; i < s->active.size(); ++i)
      if (s->active[i]->id == job_id) {
        Job *j = s->active[i];
        s->active.erase(s->active.begin() + (long)i);
        lbm_batch_sync(s->batch);
        lbm_batch_remove(s->batch, j->h);
        lbm_destroy(j->h);
        s->used -= j->footprint;
        delete j;
        return LBM_OK;
      }
    return set_error(LBM_EINVAL, "unknown or finished job");
  });
}

lbm_status lbm_server_set_check_every(lbm_server s, int32_t ticks) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!s || ticks < 0) return set_error(LBM_EINVAL, "bad argument");
    s->check_every = ticks;
    return LBM_OK;
  });
}

size_t lbm_server_used_bytes(lbm_server s) { return s ? s->used : 0; }

lbm_status lbm_server_get_stats(lbm_server s, lbm_server_stats *out) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!s || !out) return set_error(LBM_EINVAL, "NULL argument");
    *out = s->stats;
    const BatchCounters c = batch_counters(s->batch);
    out->batch_rebuilds = c.rebuilds;
    out->graph_captures = c.captures;
    return LBM_OK;
  });
}

double lbm_server_clock(lbm_server s) { return s ? s->now() : -1.0; }

lbm_status lbm_server_destroy(lbm_server s) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!s) return LBM_OK;
    if (s->batch) lbm_batch_sync(s->batch);
    for (Job *j : s->active) {
      lbm_batch_remove(s->batch, j->h);
      lbm_destroy(j->h);
      delete j;
    }
    for (Job *j : s->queue) delete j;
    s->pool.trim();
    lbm_batch_destroy(s->batch);
    delete s;
    return LBM_OK;
  });
}

}  // extern "C"
