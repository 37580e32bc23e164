// Host-side batch of chains resident in device memory.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/loopgpu.h"
#include "lg_common.h"
#include "lg_kernels.h"

namespace lg {

int set_error(int code, const std::string& msg);
std::string g_last_error_copy();  // message of the calling thread's last failure
int check_params(const lg_params& p);
ChainDev make_chain_dev(const lg_params& p);

// One configuration: tau = 0 spins, operators (bond CSR) and the world-line
// segment arrays built from them.
struct State {
  uint8_t* spins = nullptr;
  int* bo = nullptr;
  double* ot = nullptr;
  int* oi = nullptr;
  int* soff = nullptr;
  double* evt = nullptr;
  uint8_t* segspin = nullptr;
};

// Device memory a chain needs (operators: upper bound of the mean count plus
// 10 standard deviations; candidate slots per bond).
struct Footprint {
  long long sites = 0, bonds = 0, ops = 0, cand = 0;
};
Footprint footprint(const lg_params& p);

struct Capacity {
  int chains = 0;
  long long sites = 0, bonds = 0, ops = 0, cand = 0;
};

class Batch {
 public:
  ~Batch();
  int init(const lg_params* p, int n);
  int init_for_config(const lg_params* p, long long nops);
  int init_capacity(const Capacity& cap);
  int init_exchange(const lg_params& p, const double* betas, int n, uint32_t seed);
  int exchange_state(int32_t* chain_at, uint64_t* accepted, uint64_t* steps);
  int relayout(const std::vector<int>& keep, const std::vector<lg_params>& add);
  int run(uint32_t n, lg_sweep_record* out);
  int tune(uint32_t sweeps_per_trial);
  void get_launch(lg_launch* out) const;
  int set_launch(const lg_launch& l);
  int profile(uint32_t n, double* ms);
  int synchronize();
  int set_config(int c, const int32_t* spins, const int32_t* bond_offsets, const double* times,
                 const int32_t* types, uint64_t sweep);
  int get_config(int c, int32_t* spins, int32_t* bond_offsets, double* times, int32_t* types);
  int num_ops(int c, long long* n);
  // configuration of chain c of any batch (bond offsets relative to the chain)
  int chain_config(int c, std::vector<int>& spins, std::vector<int>& bo, std::vector<double>& times,
                   std::vector<int>& types);
  // whole-batch state from per-chain configurations (a freshly built batch)
  struct HostConfig {
    std::vector<int> spins, bo, types;
    std::vector<double> times;
    uint64_t sweep = 0;
  };
  int load_configs(const std::vector<HostConfig>& cfg);
  // Replace the set of chains by `ps` with configurations `cfg` (chains that
  // start fresh have all spins up and no operators), uploaded compactly to
  // the batch's existing device memory: the path for chains that resume from
  // a saved configuration (serving preemption).
  int reload(const std::vector<lg_params>& ps, const std::vector<HostConfig>& cfg);
  // take over o's device resources and settings (o is left empty)
  void adopt(Batch& o);
  bool is_exchange() const { return nslots > 0; }
  bool has_virtual_sites() const {
    for (const auto& d : chains) if (d.nv > 1) return true;
    return false;
  }
  // some chain has bonds beyond the lattice's (in-site bonds of a
  // single-ion term, site operators of a transverse field), which only the
  // per-operator linking kernel places
  bool is_sse() const { return !chains.empty() && chains[0].sse; }
  bool has_extra_bonds() const {
    for (const auto& d : chains) if (d.Nb > d.nb_lat) return true;
    return false;
  }
  int max_site_lists() const {
    int n = 0;
    for (const auto& d : chains) n = std::max(n, site_lists(d));
    return n;
  }
  // annealing schedule of a single-chain batch: n (start step, temperature)
  // pairs (looper/temperature.h); n == 0 removes it.  Rebuilds the device
  // memory when the schedule's largest beta needs more capacity.
  int set_schedule(int n, const int32_t* steps, const double* temps);
  // inverse temperature of every scheduled chain for its current sweep count,
  // uploaded to the device descriptors
  int sync_schedule();
  int cluster_only(const int32_t* spins, const int32_t* bond_offsets, const double* times,
                   const int32_t* types, uint64_t sweep, int32_t* labels, int32_t* spins_out,
                   int32_t* types_out, int64_t* n_clusters);
  DevBatch dev(int from) const;
  int enqueue_sweep();

  std::vector<lg_params> params;
  std::vector<ChainDev> chains;
  std::vector<uint64_t> host_sweep;
  int nsites = 0, nbonds = 0;
  long long cand_total = 0;
  int op_cap = 0, node_cap = 0;
  int cap_chains = 0, cap_sites = 0, cap_bonds = 0;
  Capacity budget;
  long long min_op_cap = 0;
  // extent of the site / bond / node arrays of st[cur] (previous layout plus
  // chains appended since)
  long long old_ext_sites = 0, old_ext_bonds = 0, old_ext_nodes = 0;
  bool swept_since_layout = false;
  bool tuned = false;
  State st[2];
  int cur = 0;                 // st[cur] holds the latest configuration
  int* sse_str[2] = {nullptr, nullptr};  // SSE batches: operator strings, sse_str[cur] the latest
  LaunchCfg cfg;
  bool use_graph = false;
  uint64_t graph_captures = 0;     // sweep graphs captured and instantiated
  double graph_capture_s = 0;      // host time spent in them
  hipStream_t stream = nullptr;

 private:
  template <class T>
  int alloc(T*& ptr, size_t n);
  int reserve(const Capacity& cap, bool dynamic);
  // new bases (contiguous in the given order), chain tables, sweep counters;
  // check_old: also check the previous bases a relayout leaves the chains with
  int set_layout(std::vector<ChainDev>& ds, const std::vector<uint64_t>& sweeps, bool check_old = false);
  int allocate();
  int reset_state();
  void release();
  void invalidate_graphs();
  // switch launch settings (clears the accumulators when leaving compress mode)
  int apply_cfg(const LaunchCfg& c);
  int capture(int from);
  int check_flags();
  void note_swept();

  std::vector<void*> allocations;
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
  ChainDev* d_chains = nullptr;
  int* bond_chain = nullptr;
  int* site_chain = nullptr;
  int* counts = nullptr;
  int* acc_count = nullptr;
  int* kpre = nullptr;
  double* cand_t = nullptr;
  uint8_t* cand_g = nullptr;
  int* kc_new = nullptr;
  int* l0 = nullptr;
  int* l1 = nullptr;
  int* u0 = nullptr;
  int* u1 = nullptr;
  uint8_t* evflags = nullptr;
  int* chain_node0 = nullptr;
  int* node_site = nullptr;
  int* parent = nullptr;
  uint8_t* fbit = nullptr;
  unsigned long long* fmask = nullptr;
  unsigned long long* rmask = nullptr;
  ClusterAcc* acc = nullptr;
  unsigned long long* scan_status = nullptr;
  int* scan_ctl = nullptr;
  int scan_nparts = 0;
  int cus = 256;                  // compute units of the device
  double* col_part = nullptr;
  int* totals = nullptr;
  int* flags = nullptr;
  uint32_t* d_sweep = nullptr;
  double* rec_acc = nullptr;
  double* rec_ring = nullptr;
  double* rec_host = nullptr;     // pinned host mirror of the record ring (hipHostMalloc)
  int* ring_pos = nullptr;
  int* end_count = nullptr;
  int ring_size = 0;
  // replica exchange (nslots 0: off)
  int nslots = 0;
  int* xslot = nullptr;
  double* xbeta = nullptr;
  double* xscale = nullptr;
  double* xlnr = nullptr;
  unsigned long long* xacc = nullptr;
  uint32_t* xstep = nullptr;
  uint32_t xkey0 = 0, xkey1 = 0;
  // annealing schedules (one chain): table of (start step, temperature) pairs
  std::vector<double> sched_host;
  double* d_sched = nullptr;
  double cap_beta = 0;          // inverse temperature the capacity was sized for
};

}  // namespace lg
