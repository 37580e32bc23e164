// lg_cluster_config_noalloc (include/loopgpu_noalloc.h): the cluster stage on
// a given configuration -- lg_cluster_config's work -- with every device
// buffer carved out of one static device array, so that no call allocates.
//
// The array `scratch` is a file-scope __device__ object.  Host code never
// passes its name to a kernel or a copy: the name is a host-side stand-in,
// not a device address.  Its address on each device is looked up once, with
// that device current, by hipGetSymbolAddress, and only that address is used.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/loopgpu.h"
#include "../../include/loopgpu_noalloc.h"
#include "lg_common.h"
#include "lg_host.h"
#include "lg_kernels.h"

namespace lg {
namespace {

// capacities of the array's regions (the chain lattice has one bond per
// site, the square lattice two)
constexpr long long NA_SITES = LG_NOALLOC_MAX_SITES;
constexpr long long NA_BONDS = 2 * NA_SITES;
constexpr long long NA_OPS = LG_NOALLOC_MAX_OPS;
constexpr long long NA_NODES = 2 * NA_OPS + NA_SITES;   // segments: max(events, 1) per site
constexpr long long NA_ELEMS = NA_SITES + 4 * NA_OPS;   // label elements
constexpr int NA_RING = 16;                             // record ring (one sweep is used)
constexpr long long NA_SCAN = (NA_BONDS + 2047) / 2048 + (NA_SITES + 2047) / 2048;
constexpr long long NA_COL = (NA_NODES + 1023) / 1024;  // collector chunks of at least 1024 nodes
constexpr int FLAG_INPUT = 1 << 16;                     // invalid configuration (set by k_na_prep)

constexpr size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// byte offsets of the regions in the array
struct Layout {
  size_t chains, bond_chain, site_chain, spins_old, spins_new, counts, bo_new, ot_new, oi_new, kc_new;
  size_t l0, l1, u0, u1, soff_new, evt_new, segspin_new, evflags, node_site, parent, fbit, fmask, rmask, acc;
  size_t chain_node0, scan_status, ctl, col_part, rec_acc, rec_ring, canon, roots, total;
};

constexpr Layout make_layout() {
  Layout L{};
  size_t o = 0;
  auto take = [&o](size_t bytes) { size_t r = o; o += al(bytes); return r; };
  L.chains = take(sizeof(ChainDev));
  L.bond_chain = take(4 * NA_BONDS);
  L.site_chain = take(4 * NA_SITES);
  L.spins_old = take(NA_SITES);
  L.spins_new = take(NA_SITES);
  L.counts = take(4 * NA_BONDS);
  L.bo_new = take(4 * (NA_BONDS + 1));
  L.ot_new = take(8 * NA_OPS);
  L.oi_new = take(4 * NA_OPS);
  L.kc_new = take(4 * NA_OPS);
  L.l0 = take(4 * NA_OPS);
  L.l1 = take(4 * NA_OPS);
  L.u0 = take(4 * NA_OPS);
  L.u1 = take(4 * NA_OPS);
  L.soff_new = take(4 * (NA_SITES + 1));
  L.evt_new = take(8 * NA_NODES);
  L.segspin_new = take(NA_NODES);
  L.evflags = take(NA_NODES);
  L.node_site = take(4 * NA_NODES);
  L.parent = take(4 * NA_NODES);
  L.fbit = take(NA_NODES);
  L.fmask = take(8 * (NA_NODES / 64 + 1));
  L.rmask = take(8 * (NA_NODES / 64 + 1));
  L.acc = take(sizeof(ClusterAcc) * NA_NODES);
  L.chain_node0 = take(4);
  L.scan_status = take(8 * NA_SCAN);
  // ints: scan ticket and epoch, totals (2), flags, sweep, ring position, end count
  L.ctl = take(4 * 8);
  L.col_part = take(8 * NUM_COLLECTOR_TERMS * NA_COL);
  L.rec_acc = take(8 * R_COUNT);
  L.rec_ring = take(8 * R_COUNT * NA_RING);
  L.canon = take(4 * NA_NODES);
  L.roots = take(4 * NA_ELEMS);
  L.total = o;
  return L;
}

constexpr Layout kLayout = make_layout();
static_assert(kLayout.total <= (size_t)17 << 20, "the no-allocation array stays below 17 MiB");

}  // namespace

// The work memory of lg_cluster_config_noalloc, one copy per GPU.
static __device__ __align__(256) unsigned char scratch[kLayout.total];

namespace {

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mutex;                          // serialises calls (and the set-up)
unsigned char* g_na_ptr[NA_MAX_DEVICES] = {};   // device address of `scratch`, per device

// Address of the current device's array, looked up once per device with
// that device current.
int na_scratch(unsigned char** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error(LG_ERR_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= NA_MAX_DEVICES) return set_error(LG_ERR_DEVICE, "device index beyond the supported range");
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(LG_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_na_ptr[dev] = (unsigned char*)p;
  }
  *out = g_na_ptr[dev];
  return LG_OK;
}

// One-chain batch view of the array (previous-configuration arrays are not
// read by the cluster stage; they alias the new ones).
DevBatch na_dev(unsigned char* base, int N, int Nb, bool field) {
  const Layout& L = kLayout;
  int* ctl = (int*)(base + L.ctl);
  DevBatch B;
  std::memset(&B, 0, sizeof(B));
  B.nchains = 1;
  B.nsites = N;
  B.nbonds = Nb;
  B.chains = (ChainDev*)(base + L.chains);
  B.bond_chain = (const int*)(base + L.bond_chain);
  B.site_chain = (const int*)(base + L.site_chain);
  B.spins_old = (const uint8_t*)(base + L.spins_old);
  B.spins_new = (uint8_t*)(base + L.spins_new);
  B.bo_new = (int*)(base + L.bo_new);
  B.ot_new = (double*)(base + L.ot_new);
  B.oi_new = (int*)(base + L.oi_new);
  B.soff_new = (int*)(base + L.soff_new);
  B.evt_new = (double*)(base + L.evt_new);
  B.segspin_new = (uint8_t*)(base + L.segspin_new);
  B.bo_old = B.bo_new;
  B.ot_old = B.ot_new;
  B.oi_old = B.oi_new;
  B.soff_old = B.soff_new;
  B.evt_old = B.evt_new;
  B.segspin_old = B.segspin_new;
  B.counts = (int*)(base + L.counts);
  B.kc_new = (int*)(base + L.kc_new);
  B.l0 = (int*)(base + L.l0);
  B.l1 = (int*)(base + L.l1);
  B.u0 = (int*)(base + L.u0);
  B.u1 = (int*)(base + L.u1);
  B.evflags = (uint8_t*)(base + L.evflags);
  B.chain_node0 = (int*)(base + L.chain_node0);
  B.node_site = (int*)(base + L.node_site);
  B.parent = (int*)(base + L.parent);
  B.fbit = (uint8_t*)(base + L.fbit);
  B.fmask = (unsigned long long*)(base + L.fmask);
  B.rmask = (unsigned long long*)(base + L.rmask);
  B.acc = (ClusterAcc*)(base + L.acc);
  B.scan_status = (unsigned long long*)(base + L.scan_status);
  B.scan_ctl = ctl;
  B.totals = ctl + 2;
  B.flags = ctl + 4;
  B.sweep = (uint32_t*)(ctl + 5);
  B.ring_pos = ctl + 6;
  B.end_count = ctl + 7;
  B.col_part = (double*)(base + L.col_part);
  B.rec_acc = (double*)(base + L.rec_acc);
  B.rec_ring = (double*)(base + L.rec_ring);
  B.ring_size = NA_RING;
  B.any_field = field ? 1 : 0;
  B.any_split = 0;
  B.any_extra = 0;
  B.nslots = 0;
  B.op_cap = (int)NA_OPS;
  B.node_cap = (int)NA_NODES;
  return B;
}

// The configuration into the cluster stage's arrays, validated as
// lg_cluster_config validates it: per bond, its operators' times (in [0, 1),
// ascending), info words (bond << 3 | graph << 1 | offdiagonal) and
// off-diagonal counts before each; per site, the tau = 0 spin.  A bond whose
// offsets leave [0, n_ops] reads nothing.
__global__ void __launch_bounds__(256) k_na_prep(DevBatch B, const int32_t* spins, const int32_t* bo,
                                                 const double* times, const int32_t* types, int nops) {
  const int stride = gridDim.x * blockDim.x;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  int* counts = B.counts;
  int* bond_chain = const_cast<int*>(B.bond_chain);
  int* site_chain = const_cast<int*>(B.site_chain);
  uint8_t* spins_old = const_cast<uint8_t*>(B.spins_old);
  if (tid == 0 && (bo[0] != 0 || bo[B.nbonds] != nops)) atomicOr(B.flags, FLAG_INPUT);
  for (int b = tid; b < B.nbonds; b += stride) {
    bond_chain[b] = 0;
    const int lo = bo[b], hi = bo[b + 1];
    if (lo < 0 || hi < lo || hi > nops) {
      counts[b] = 0;
      atomicOr(B.flags, FLAG_INPUT);
      continue;
    }
    counts[b] = hi - lo;
    int kc = 0;
    double prev = 0.0;
    bool bad = false;
    for (int k = lo; k < hi; ++k) {
      const double t = times[k];
      const int ty = types[k];
      if (!(t >= 0.0 && t < 1.0) || t < prev || ty < 0 || (ty >> 2) > 3) bad = true;
      prev = t;
      B.ot_new[k] = t;
      B.oi_new[k] = (b << 3) | (((ty >> 2) & 3) << 1) | (ty & 1);
      B.kc_new[k] = kc;
      kc += ty & 1;
    }
    if (bad) atomicOr(B.flags, FLAG_INPUT);
  }
  for (int s = tid; s < B.nsites; s += stride) {
    site_chain[s] = 0;
    spins_old[s] = (uint8_t)(spins[s] & 1);
  }
}

// flipped configuration in the ABI's layout
__global__ void __launch_bounds__(256) k_na_out(DevBatch B, int nops, int32_t* spins_out, int32_t* types_out) {
  const int stride = gridDim.x * blockDim.x;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (spins_out)
    for (int s = tid; s < B.nsites; s += stride) spins_out[s] = B.spins_new[s];
  if (types_out)
    for (int k = tid; k < nops; k += stride) {
      const int oi = B.oi_new[k];
      types_out[k] = (oi & 1) | (((oi >> 1) & 3) << 2);
    }
}

int na_flags(const DevBatch& B, int* f) {
  hipError_t e = hipStreamSynchronize(0);
  if (e == hipSuccess) e = hipMemcpy(f, B.flags, sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return set_error(LG_ERR_DEVICE, std::string("cluster stage: ") + hipGetErrorString(e));
  return LG_OK;
}

}  // namespace
}  // namespace lg

extern "C" {

size_t lg_noalloc_scratch_bytes(void) { return lg::kLayout.total; }

int lg_cluster_config_noalloc(const lg_params* p, const int32_t* spins, const int32_t* bond_offsets,
                              const double* times, const int32_t* types, int64_t n_ops, uint64_t sweep,
                              int32_t* labels, int32_t* spins_out, int32_t* types_out, int64_t* n_clusters) {
  using namespace lg;
  if (!p || !spins || !bond_offsets || (n_ops > 0 && (!times || !types)))
    return set_error(LG_ERR_ARG, "null pointer");
  int rc = check_params(*p);
  if (rc) return rc;
  if (p->hx != 0 || p->d != 0 || p->algorithm != LG_ALGORITHM_PATH_INTEGRAL)
    return set_error(LG_ERR_UNSUPPORTED,
                     "the cluster stage of a given configuration takes hx = 0, d = 0 and the path integral");
  ChainDev d = make_chain_dev(*p);
  if (d.N > NA_SITES || n_ops < 0 || n_ops > NA_OPS)
    return set_error(LG_ERR_CAPACITY, "configuration beyond the no-allocation maximum (" +
                                          std::to_string(LG_NOALLOC_MAX_SITES) + " sites, " +
                                          std::to_string(LG_NOALLOC_MAX_OPS) + " operators)");
  const int nops = (int)n_ops;
  std::lock_guard<std::mutex> lk(g_na_mutex);
  unsigned char* base = nullptr;
  if ((rc = na_scratch(&base))) return rc;
  const Layout& L = kLayout;
  DevBatch B = na_dev(base, d.N, d.Nb, p->h != 0.0);
#define NA_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return set_error(LG_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)
  // descriptor, sweep index, and the control words that must start at zero
  // (scan ticket and epoch, totals, flags, ring position, end count; scan
  // status words; record accumulators)
  uint32_t sw = (uint32_t)sweep;
  NA_HIP(hipMemcpy(base + L.chains, &d, sizeof(d), hipMemcpyHostToDevice));
  NA_HIP(hipMemsetAsync(base + L.ctl, 0, 4 * 8, 0));
  NA_HIP(hipStreamSynchronize(0));
  NA_HIP(hipMemcpy(B.sweep, &sw, sizeof(sw), hipMemcpyHostToDevice));
  NA_HIP(hipMemsetAsync(base + L.scan_status, 0, 8 * NA_SCAN, 0));
  NA_HIP(hipMemsetAsync(base + L.rec_acc, 0, 8 * R_COUNT, 0));
  const int pblocks = std::max(1, std::min(1024, (std::max(d.Nb, d.N) + 255) / 256));
  hipLaunchKernelGGL(k_na_prep, dim3(pblocks), dim3(256), 0, 0, B, spins, bond_offsets, times, types, nops);
  NA_HIP(hipGetLastError());
  int f = 0;
  if ((rc = na_flags(B, &f))) return rc;
  if (f & FLAG_INPUT)
    return set_error(LG_ERR_ARG, "invalid configuration (bond offsets, operator times in [0,1) ascending per bond, "
                                 "or operator types)");
  // launch settings from the expected events per site, as a batch chooses
  // them; the compress pass clears the roots' accumulators itself, so the
  // array's accumulators need no clearing between calls
  LaunchCfg c;
  int cus = 256;
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const double per_site = 2.0 * nops / d.N;
  c.link_cap = per_site < 48 ? 0 : (per_site < 96 ? 128 : (per_site <= 400 ? 512 : 1536));
  if (d.nv > 1) {  // virtual sites: the kernels that write separate top segments, and the boundary
    c.link_cap = (per_site < 8 || site_lists(d) > 4) ? 0 : 1;
    B.any_split = 1;
  }
  c.compress = 1;
  c.est_flush = per_site < 48 ? 4 : 16;
  c.blocks_op = c.blocks_node = std::max(64, (int)std::min<double>(cus * 8, std::max(256.0, 2.0 * nops / 256.0)));
  launch_site_scan(B, 0);
  launch_cluster_stage(B, c, 0);
  NA_HIP(hipGetLastError());
  if ((rc = na_flags(B, &f))) return rc;
  if (f) return set_error(LG_ERR_CAPACITY, "device capacity exceeded in the no-allocation cluster stage");
  if (labels) {
    // canonical labels: per root the minimum element index (atomicMin from
    // 0x7f7f7f7f, above every element index), then every element's root's
    const int nnodes = 2 * nops + d.N;
    NA_HIP(hipMemsetAsync(base + L.canon, 0x7f, 4 * (size_t)nnodes, 0));
    launch_canon(B, 0, d.N + 4 * nops, (int*)(base + L.canon), labels, (int*)(base + L.roots), 0);
    NA_HIP(hipGetLastError());
  }
  if (spins_out || (types_out && nops)) {
    const int oblocks = std::max(1, std::min(1024, (std::max(nops, d.N) + 255) / 256));
    hipLaunchKernelGGL(k_na_out, dim3(oblocks), dim3(256), 0, 0, B, nops, spins_out, types_out);
    NA_HIP(hipGetLastError());
  }
  NA_HIP(hipStreamSynchronize(0));
  if (n_clusters) {
    double nc = 0;
    NA_HIP(hipMemcpy(&nc, B.rec_ring + R_NC, sizeof(double), hipMemcpyDeviceToHost));
    *n_clusters = (int64_t)nc;
  }
#undef NA_HIP
  return LG_OK;
}

}  // extern "C"
