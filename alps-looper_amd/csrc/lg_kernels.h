// Device buffer bundle and launch interface of the sweep kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "lg_common.h"

namespace lg {

enum : int { FLAG_CAND_OVERFLOW = 1, FLAG_OP_OVERFLOW = 2, FLAG_NODE_OVERFLOW = 4, FLAG_RING_OVERFLOW = 8 };

// Pointers of one sweep: "old" = previous configuration (read), "new" =
// configuration being built (written).  Passed by value to every kernel.
struct DevBatch {
  int nchains;
  int nsites, nbonds;          // totals of the new layout
  // descriptors; old_site_base / old_bond_base locate a chain in the previous
  // configuration and differ from site_base / bond_base only in the first
  // sweep after a relayout: k_end_sweep resets them to the new bases
  ChainDev* chains;
  const int* bond_chain;       // [nbonds]
  const int* site_chain;       // [nsites]
  // previous configuration
  const uint8_t* spins_old;    // tau = 0 spins, old site index
  const int* bo_old;           // bond offsets, old bond index
  const double* ot_old;
  const int* oi_old;
  const int* soff_old;         // segment offsets, old site index
  const double* evt_old;       // event time per segment node (EMPTY_TIME for event-free sites)
  const uint8_t* segspin_old;  // post-flip spin per segment node
  // new configuration
  uint8_t* spins_new;
  int* bo_new;
  double* ot_new;
  int* oi_new;
  int* soff_new;
  double* evt_new;
  uint8_t* segspin_new;
  // per-sweep scratch
  int* counts;                 // [nbonds] new operators per bond
  int* acc_count;              // [nbonds] accepted candidates per bond
  int* kpre;                   // [old ops] off-diagonal operators before, within the bond
  double* cand_t;              // candidate slots
  uint8_t* cand_g;
  int* kc_new;                 // [new ops] off-diagonal operators before, within the bond
  int* l0;                     // [new ops] lower-leg node on the source site
  int* l1;                     // [new ops] lower-leg node on the target site
  int* u0;                     // [new ops] upper-leg node on the source site
  int* u1;                     // [new ops] upper-leg node on the target site
  uint8_t* evflags;            // [nodes]
  int* chain_node0;            // [nchains] first segment node of each chain (new configuration)
  int* node_site;              // [nodes] site of each segment (batches of several chains) or of each site's first segment
  int* parent;                 // [nodes]
  uint8_t* fbit;               // [nodes] flip bit of a cluster at its root (compress mode, field path)
  ClusterAcc* acc;             // [nodes] (roots only)
  unsigned long long* scan_status;  // per scan chunk: epoch | state | value
  int* scan_ctl;                    // [0] chunk ticket, [1] launch epoch
  double* col_part;            // [col_chunks(node_cap)][16] collector partial sums
  int* totals;                 // [0] operators, [1] nodes of the new configuration
  int* flags;
  uint32_t* sweep;             // [nchains]
  double* rec_acc;             // [nchains][R_COUNT]
  double* rec_ring;            // [ring][nchains][R_COUNT]
  int* ring_pos;
  int* end_count;              // blocks of end_sweep finished (reset by the last one)
  int ring_size;
  // replica exchange between neighbouring temperature slots (nslots 0: off)
  int any_field;               // some chain has a longitudinal field
  int nslots;
  int* xslot;                  // [nslots] chain holding each temperature
  const double* xbeta;         // [nslots] inverse temperatures
  const double* xscale;        // [nslots] candidate scale 2^52 / (beta W) at each temperature
  const double* xlnr;          // [nslots-1] log(beta[t+1] / beta[t])
  unsigned long long* xacc;    // [nslots-1] accepted swaps per pair
  uint32_t* xstep;             // exchange steps done
  uint32_t xkey0, xkey1;
  int op_cap, node_cap;
  // (fields added later sit here, so that the kernels' argument loads of the
  // fields above keep their layout and the kernels their register counts)
  // annealing schedules: (start step, temperature) pairs of the chains with
  // sched_n > 0, at their sched_off (k_end_sweep sets beta for the next sweep)
  const double* sched;
  unsigned long long* fmask;   // [nodes / 64 + 1] flip bit of every node's cluster, 64 nodes per word
  unsigned long long* rmask;   // [nodes / 64 + 1] 1 for the roots, 64 nodes per word
  int any_split;               // some chain has virtual sites (S=1): segments without wrap, boundary permutation
  int any_extra;               // bonds beyond the lattice's in some chain: 1 single-ion in-site bonds, 2 transverse-field site bonds (bits)
  int any_sse;                 // the chains are SSE chains ("loop; sse"; a batch holds one kind)
  const int* sse_old;          // operator strings of the previous configuration (ChainDev::sse_base)
  int* sse_new;                // operator strings being built
  int* sse_perm;               // [ops] string position -> operator index of the new configuration
};

struct LaunchCfg {
  int group = 64;              // lanes per bond in the diagonal update
  int max_blocks_bond = 1 << 20;
  int blocks_op = 1024;        // grid of the per-operator kernels (grid-stride)
  int blocks_node = 1024;      // grid of the per-node kernels (grid-stride)
  int link_cap = 512;          // linking: 0 one thread per operator, 1 one thread per site, else a group of threads per site staging 128, 512 or 1536 events
  // roots: 1 a compression pass before the estimator (which then reads roots
  // and flip bits in one load), 0 the estimator walks to the roots itself and
  // the collector clears the accumulators it has read
  int compress = 1;
  // estimator: flush of its LDS cluster table to global memory every
  // 256 * est_flush segments (4, 8 or 16; 16 = once per chunk)
  int est_flush = 16;
  // diagonal update, fill pass with a wave per bond: times per list staged
  // in LDS (256, 512 or 1024; the block holds 64 bytes per time)
  int fill_cap = 256;
};

constexpr int NUM_SWEEP_KERNELS = 9;
constexpr int NUM_COLLECTOR_TERMS = 19;   // per-cluster terms summed into the sweep record
extern const char* const kKernelNames[NUM_SWEEP_KERNELS];

void launch_sweep(const DevBatch& B, const LaunchCfg& L, hipStream_t st, hipEvent_t* ev = nullptr);
void launch_cluster_stage(const DevBatch& B, const LaunchCfg& L, hipStream_t st, hipEvent_t* ev = nullptr);
void launch_site_scan(const DevBatch& B, hipStream_t st);
void launch_canon(const DevBatch& B, int chain, int nelem, int* canon, int* labels, int* roots,
                  hipStream_t st);
int scan_blocks(const DevBatch& B);
int col_chunks(int node_cap);

}  // namespace lg
