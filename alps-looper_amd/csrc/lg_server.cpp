// Job server: continuous batching of loop-update jobs on one GPU
// (include/loopgpu_server.h).  A background thread owns the device batch;
// between chunks of sweeps it retires finished jobs and admits queued ones
// whose footprint fits (first come first served, later jobs may backfill),
// then relayouts the batch without moving resident chains.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/loopgpu_server.h"
#include "lg_binning.h"
#include "lg_host.h"

using clock_type = std::chrono::steady_clock;

namespace {

thread_local std::string g_serr;
int serr(int code, const std::string& m) { g_serr = m; return code; }

struct Job {
  uint64_t id = 0;
  lg_job job{};
  int status = LG_JOB_QUEUED;
  clock_type::time_point t_submit, t_admit, t_done;
  bool admitted_once = false;
  uint64_t done = 0;
  lg::Footprint fp;
  looper_amd::Observable obs[LG_JOB_NOBS];
  // preemption: requested for a running job; its configuration (spins,
  // operators, sweep count = Philox counter) is saved at the next chunk
  // boundary and loaded back when it is admitted again
  bool preempt = false;
  bool saved = false;
  lg::Batch::HostConfig cfg;
  uint64_t preemptions = 0;
};

void add_fp(lg::Footprint& a, const lg::Footprint& b, int sgn) {
  a.sites += sgn * b.sites;
  a.bonds += sgn * b.bonds;
  a.ops += sgn * b.ops;
  a.cand += sgn * b.cand;
}

}  // namespace

struct lg_server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::map<uint64_t, Job> jobs;
  std::deque<uint64_t> queue;
  std::vector<uint64_t> active;  // batch order
  lg::Batch batch;
  lg::Capacity cap;
  lg::Footprint used;
  uint32_t max_chunk = 64;
  std::thread th;
  bool stop = false;
  std::string fatal;
  uint64_t next_id = 1, submitted = 0, total_updates = 0, launches = 0, preempted = 0;
  uint64_t graph_captures = 0;    // the batch's, copied under the lock after each chunk
  double graph_capture_s = 0;
  double busy_s = 0;
  std::vector<double> latencies;

  bool fits(const lg::Footprint& f) const {
    return (long long)active_count() < cap.chains && used.sites + f.sites <= cap.sites &&
           used.bonds + f.bonds <= cap.bonds && used.ops + f.ops <= cap.ops && used.cand + f.cand <= cap.cand;
  }
  size_t active_count() const { return admitted_; }
  size_t admitted_ = 0;

  void measure(Job& j, const lg_sweep_record& r) {
    const lg_params& p = j.job.params;
    const double V = (p.lattice == LG_LATTICE_CHAIN) ? p.L : (double)p.L * p.L;
    const double dim = (p.lattice == LG_LATTICE_CHAIN) ? 1 : 2;
    j.obs[0] << r.energy / V;
    if (p.h == 0.0) {  // improved estimators
      j.obs[1] << r.beta * r.umag / V;
      j.obs[2] << r.smag2 / V;
      j.obs[3] << r.beta * r.smag / V;
    } else {           // field: normal estimators (susceptibility.h:304-337)
      j.obs[1] << r.beta * r.umag_a * r.umag_a / V;
      j.obs[2] << r.smag0 * r.smag0 / V;
      j.obs[3] << r.beta * r.smag_a * r.smag_a / V;
    }
    j.obs[4] << r.n_clusters;
    j.obs[5] << r.w2 / (r.beta * dim);
    j.obs[6] << r.n_op;
  }

  void fail_all(const std::string& m) {
    fatal = m;
    for (auto& kv : jobs)
      if (kv.second.status == LG_JOB_RUNNING || kv.second.status == LG_JOB_QUEUED) {
        kv.second.status = LG_JOB_CANCELLED;
        kv.second.t_done = clock_type::now();
      }
    active.clear();
    queue.clear();
    admitted_ = 0;
    cv_done.notify_all();
  }

  void loop() {
    std::vector<lg_sweep_record> rec;
    while (true) {
      std::vector<int> keep;
      std::vector<lg_params> add;
      std::vector<uint64_t> next, added;
      bool changed = false, resumed = false;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || !queue.empty() || !active.empty(); });
        if (stop) return;
        for (size_t i = 0; i < active.size(); ++i) {
          Job& j = jobs[active[i]];
          if (j.status == LG_JOB_RUNNING && j.preempt) {
            // save the configuration the job has reached and queue it again
            // (behind the jobs already waiting); its measurements stay
            int rc = batch.chain_config((int)i, j.cfg.spins, j.cfg.bo, j.cfg.times, j.cfg.types);
            if (rc) {
              fail_all("preempt: " + lg::g_last_error_copy());
              return;
            }
            j.cfg.sweep = batch.host_sweep[i];
            j.saved = true;
            j.preempt = false;
            j.status = LG_JOB_QUEUED;
            ++j.preemptions;
            ++preempted;
            queue.push_back(j.id);
          }
          if (j.status == LG_JOB_RUNNING) {
            keep.push_back((int)i);
            next.push_back(active[i]);
          } else {
            add_fp(used, j.fp, -1);
            --admitted_;
            changed = true;
          }
        }
        for (auto it = queue.begin(); it != queue.end();) {
          Job& j = jobs[*it];
          if (j.status == LG_JOB_CANCELLED) { it = queue.erase(it); continue; }
          if (fits(j.fp)) {
            add_fp(used, j.fp, +1);
            ++admitted_;
            j.status = LG_JOB_RUNNING;
            if (!j.admitted_once) j.t_admit = clock_type::now();
            j.admitted_once = true;
            add.push_back(j.job.params);
            next.push_back(j.id);
            added.push_back(j.id);
            resumed = resumed || j.saved;
            it = queue.erase(it);
            changed = true;
          } else {
            ++it;
          }
        }
        if (resumed) {
          // a job resumes from its saved configuration: every resident chain's
          // configuration goes through the host and the batch is reloaded
          // compactly (a rare event; admissions of fresh jobs stay on the
          // device)
          std::vector<lg::Batch::HostConfig> cfgs(next.size());
          std::vector<lg_params> ps;
          for (size_t k = 0; k < keep.size(); ++k) {
            int rc = batch.chain_config(keep[k], cfgs[k].spins, cfgs[k].bo, cfgs[k].times, cfgs[k].types);
            if (rc) {
              fail_all("resume: " + lg::g_last_error_copy());
              return;
            }
            cfgs[k].sweep = batch.host_sweep[keep[k]];
            ps.push_back(batch.params[keep[k]]);
          }
          for (size_t a = 0; a < added.size(); ++a) {
            Job& j = jobs[added[a]];
            lg::Batch::HostConfig& h = cfgs[keep.size() + a];
            if (j.saved) {
              h = std::move(j.cfg);
              j.cfg = lg::Batch::HostConfig();
              j.saved = false;
            } else {
              const lg::ChainDev d = lg::make_chain_dev(j.job.params);
              h.spins.assign(d.N, 0);  // all up, no operators (path_integral.C:225)
              h.bo.assign(d.Nb + 1, 0);
            }
            ps.push_back(j.job.params);
          }
          int rc = batch.reload(ps, cfgs);
          if (rc) {
            fail_all("resume: " + lg::g_last_error_copy());
            return;
          }
        }
      }
      if (changed && !resumed) {
        int rc = batch.relayout(keep, add);
        if (rc) {
          std::lock_guard<std::mutex> lk(mu);
          fail_all("relayout: " + lg::g_last_error_copy());
          return;
        }
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        active = next;
      }
      if (next.empty()) continue;
      uint32_t chunk = max_chunk;
      {
        std::lock_guard<std::mutex> lk(mu);
        for (uint64_t id : next) {
          const Job& j = jobs[id];
          uint64_t total = (uint64_t)j.job.thermalization + j.job.sweeps;
          uint64_t left = total > j.done ? total - j.done : 1;
          chunk = (uint32_t)std::min<uint64_t>(chunk, left);
        }
      }
      chunk = std::max<uint32_t>(chunk, 1);
      rec.resize((size_t)chunk * next.size());
      auto t0 = clock_type::now();
      int rc = batch.run(chunk, rec.data());
      auto t1 = clock_type::now();
      std::lock_guard<std::mutex> lk(mu);
      if (rc) {
        fail_all("sweep: " + lg::g_last_error_copy());
        return;
      }
      busy_s += std::chrono::duration<double>(t1 - t0).count();
      launches += chunk;
      graph_captures = batch.graph_captures;
      graph_capture_s = batch.graph_capture_s;
      for (size_t k = 0; k < next.size(); ++k) {
        Job& j = jobs[next[k]];
        if (j.status != LG_JOB_RUNNING) continue;  // cancelled meanwhile
        const uint64_t total = (uint64_t)j.job.thermalization + j.job.sweeps;
        for (uint32_t s = 0; s < chunk && j.done < total; ++s) {
          const lg_sweep_record& r = rec[(size_t)s * next.size() + k];
          if (j.done >= j.job.thermalization) measure(j, r);
          total_updates += (uint64_t)r.n_op;
          ++j.done;
        }
        if (j.done >= total) {
          j.status = LG_JOB_DONE;
          j.t_done = t1;
          latencies.push_back(std::chrono::duration<double>(j.t_done - j.t_submit).count());
        }
      }
      cv_done.notify_all();
    }
  }

  void fill(const Job& j, lg_job_result* r) const {
    std::memset(r, 0, sizeof(*r));
    r->status = j.status;
    r->sweeps_done = j.done;
    if (j.status == LG_JOB_DONE || j.status == LG_JOB_CANCELLED)
      r->latency_s = std::chrono::duration<double>(j.t_done - j.t_submit).count();
    if (j.status != LG_JOB_QUEUED)
      r->queue_s = std::chrono::duration<double>(j.t_admit - j.t_submit).count();
    for (int q = 0; q < LG_JOB_NOBS; ++q) {
      r->mean[q] = j.obs[q].mean();
      r->error[q] = j.obs[q].error();
    }
  }
};

extern "C" {

int lg_server_last_error(char* buf, size_t len) {
  if (!buf || !len) return LG_ERR_ARG;
  std::snprintf(buf, len, "%s", g_serr.c_str());
  return LG_OK;
}

int lg_server_start(const lg_server_config* cfg, lg_server** out) {
  if (!cfg || !out) return serr(LG_ERR_ARG, "null pointer");
  *out = nullptr;
  lg_server* s = new lg_server();
  s->cap.chains = cfg->max_jobs > 0 ? cfg->max_jobs : 64;
  s->cap.sites = cfg->max_sites > 0 ? cfg->max_sites : (1 << 20);
  s->cap.bonds = cfg->max_bonds > 0 ? cfg->max_bonds : (1 << 21);
  s->cap.ops = cfg->max_ops > 0 ? cfg->max_ops : (64ll << 20);
  s->cap.cand = cfg->max_cand > 0 ? cfg->max_cand : (64ll << 20);
  s->max_chunk = cfg->max_chunk > 0 ? cfg->max_chunk : 64;
  int rc = s->batch.init_capacity(s->cap);
  if (rc) {
    std::string m = lg::g_last_error_copy();
    delete s;
    return serr(rc, m);
  }
  s->batch.use_graph = true;
  s->th = std::thread([s] { s->loop(); });
  *out = s;
  return LG_OK;
}

int lg_server_submit(lg_server* s, const lg_job* job, uint64_t* id) {
  if (!s || !job) return serr(LG_ERR_ARG, "null pointer");
  int rc = lg::check_params(job->params);
  if (rc) return serr(rc, lg::g_last_error_copy());
  if (job->sweeps == 0) return serr(LG_ERR_ARG, "job needs at least one measurement sweep");
  if (job->params.algorithm != LG_ALGORITHM_PATH_INTEGRAL)
    return serr(LG_ERR_UNSUPPORTED, "the server runs path-integral jobs");
  if (job->params.hx != 0 && job->params.jxy > 0)
    return serr(LG_ERR_UNSUPPORTED, "the server runs sign-free jobs (its results are not sign-weighted)");
  lg::Footprint f = lg::footprint(job->params);
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->fatal.empty()) return serr(LG_ERR_DEVICE, s->fatal);
  if (f.sites > s->cap.sites || f.bonds > s->cap.bonds || f.ops > s->cap.ops || f.cand > s->cap.cand)
    return serr(LG_ERR_CAPACITY, "job footprint exceeds the server capacity");
  Job& j = s->jobs[s->next_id];
  j.id = s->next_id++;
  j.job = *job;
  j.fp = f;
  j.t_submit = clock_type::now();
  s->queue.push_back(j.id);
  ++s->submitted;
  if (id) *id = j.id;
  s->cv_work.notify_all();
  return LG_OK;
}

int lg_server_poll(lg_server* s, uint64_t id, lg_job_result* res) {
  if (!s || !res) return serr(LG_ERR_ARG, "null pointer");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) { res->status = LG_JOB_UNKNOWN; return serr(LG_ERR_ARG, "unknown job id"); }
  s->fill(it->second, res);
  return LG_OK;
}

int lg_server_wait(lg_server* s, uint64_t id, int64_t timeout_ms, lg_job_result* res) {
  if (!s || !res) return serr(LG_ERR_ARG, "null pointer");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) { res->status = LG_JOB_UNKNOWN; return serr(LG_ERR_ARG, "unknown job id"); }
  auto done = [&] { return it->second.status == LG_JOB_DONE || it->second.status == LG_JOB_CANCELLED; };
  if (timeout_ms > 0) s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done);
  s->fill(it->second, res);
  if (!s->fatal.empty() && it->second.status == LG_JOB_CANCELLED) return serr(LG_ERR_DEVICE, s->fatal);
  return LG_OK;
}

int lg_server_cancel(lg_server* s, uint64_t id) {
  if (!s) return serr(LG_ERR_ARG, "null pointer");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return serr(LG_ERR_ARG, "unknown job id");
  Job& j = it->second;
  if (j.status == LG_JOB_QUEUED || j.status == LG_JOB_RUNNING) {
    j.status = LG_JOB_CANCELLED;
    j.t_done = clock_type::now();
    s->cv_work.notify_all();
    s->cv_done.notify_all();
  }
  return LG_OK;
}

int lg_server_preempt(lg_server* s, uint64_t id) {
  if (!s) return serr(LG_ERR_ARG, "null pointer");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return serr(LG_ERR_ARG, "unknown job id");
  Job& j = it->second;
  if (j.status == LG_JOB_RUNNING) {
    j.preempt = true;
    s->cv_work.notify_all();
    return LG_OK;
  }
  if (j.status == LG_JOB_QUEUED) return LG_OK;
  return serr(LG_ERR_STATE, "job is done or cancelled");
}

int lg_server_stats_get(lg_server* s, lg_server_stats* st) {
  if (!s || !st) return serr(LG_ERR_ARG, "null pointer");
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(st, 0, sizeof(*st));
  st->submitted = s->submitted;
  for (auto& kv : s->jobs) {
    if (kv.second.status == LG_JOB_DONE) ++st->done;
    if (kv.second.status == LG_JOB_RUNNING) ++st->running;
    if (kv.second.status == LG_JOB_QUEUED) ++st->queued;
  }
  std::vector<double> l = s->latencies;
  std::sort(l.begin(), l.end());
  if (!l.empty()) {
    auto pct = [&](double q) {
      size_t k = (size_t)std::ceil(q * l.size()) - 1;
      return l[std::min(k, l.size() - 1)];
    };
    st->p50_latency_s = pct(0.50);
    st->p99_latency_s = pct(0.99);
    double m = 0;
    for (double x : l) m += x;
    st->mean_latency_s = m / l.size();
  }
  st->busy_s = s->busy_s;
  st->total_updates = s->total_updates;
  st->updates_per_s = s->busy_s > 0 ? s->total_updates / s->busy_s : 0.0;
  st->sweep_launches = s->launches;
  st->preemptions = s->preempted;
  st->graph_captures = s->graph_captures;
  st->graph_capture_s = s->graph_capture_s;
  return LG_OK;
}

int lg_server_stop(lg_server* s) {
  if (!s) return LG_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  if (s->th.joinable()) s->th.join();
  delete s;
  return LG_OK;
}

}  // extern "C"
