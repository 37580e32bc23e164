// gfx950 kernels for one continuous-imaginary-time loop-cluster sweep of a
// batch of S=1/2 XXZ chains (the sweep of path_integral.C:354-864).
//
// One sweep maps the previous configuration (spins at tau=0, operators in
// bond CSR, per-site world-line segments with their post-flip spins) to the
// next one:
//   diag_count   per bond: Poisson candidates from Philox, acceptance against
//                the previous world lines, kept off-diagonal operators counted
//   scan         (single pass) bond offsets of the new operator list and
//                segment offsets of the new world lines (max(events,1) per site)
//   diag_fill    per bond: kept operators (regraphed) merged with accepted
//                candidates in time order
//   link         per site: rank of each event on the site's world line
//                (binary search in the other bond lists, staged in LDS), the
//                segments below and above each leg, segment spins by kink
//                parity
//   uf_union     per operator: lock-free union of segments by the cut rules
//   uf_compress  (compress mode) per node: full compression; per root: clear
//                the accumulator, draw the cluster's flip bit
//   estimate     per node: the root (compress mode: one load; else a walk that
//                also points the node at its root) and the cluster's flip bit
//                (compress mode: the root's; else drawn from the root's Philox
//                counter), improved-estimator terms into the root, flipped
//                segment spins and tau=0 spins
//   flip_collect per operator: type flip; per root: cluster terms -> record
//                (without compress mode: accumulator cleared for the next sweep)
//   end_sweep    per chain: finish and store the sweep record
#include <hip/hip_runtime.h>

#include "lg_common.h"
#include "lg_kernels.h"

namespace lg {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

template <int G>
__device__ __forceinline__ int glane() { return threadIdx.x & (G - 1); }

template <int G>
__device__ __forceinline__ uint64_t gballot(bool p) {
  uint64_t m = __ballot(p);
  if (G == 64) return m;
  int base = lane_id() & ~(G - 1);
  return (m >> base) & ((1ull << G) - 1);
}

template <int G>
__device__ __forceinline__ uint64_t gscan_incl_u64(uint64_t v) {
  int l = glane<G>();
#pragma unroll
  for (int d = 1; d < G; d <<= 1) {
    uint32_t lo = __shfl_up((unsigned int)(v & 0xffffffffu), d, G);
    uint32_t hi = __shfl_up((unsigned int)(v >> 32), d, G);
    uint64_t o = ((uint64_t)hi << 32) | lo;
    if (l >= d) v += o;
  }
  return v;
}

template <int G>
__device__ __forceinline__ uint64_t gbcast_u64(uint64_t v, int src) {
  uint32_t lo = __shfl((unsigned int)(v & 0xffffffffu), src, G);
  uint32_t hi = __shfl((unsigned int)(v >> 32), src, G);
  return ((uint64_t)hi << 32) | lo;
}

template <int G>
__device__ __forceinline__ int gsum_int(int v) {
#pragma unroll
  for (int d = G / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, G);
  return v;
}

// first j in [0,n) with a[j] >= x (n if none)
__device__ __forceinline__ int lower_bound_d(const double* a, int n, double x) {
  int lo = 0, len = n;
  while (len > 0) {
    int h = len >> 1;
    if (a[lo + h] < x) { lo += h + 1; len -= h + 1; } else { len = h; }
  }
  return lo;
}
// first j in [0,n) with a[j * stride] >= x (n if none)
__device__ __forceinline__ int lower_bound_ds(const double* a, long long stride, int n, double x) {
  int lo = 0, len = n;
  while (len > 0) {
    int h = len >> 1;
    if (a[(lo + h) * stride] < x) { lo += h + 1; len -= h + 1; } else { len = h; }
  }
  return lo;
}
// first j in [0,n) with a[j] > x (n if none)
__device__ __forceinline__ int upper_bound_d(const double* a, int n, double x) {
  int lo = 0, len = n;
  while (len > 0) {
    int h = len >> 1;
    if (a[lo + h] <= x) { lo += h + 1; len -= h + 1; } else { len = h; }
  }
  return lo;
}

// Items [lo, hi) of this block.  [0, n) is cut into one contiguous slice per
// XCD (the dispatcher sends workgroup i to XCD i % 8), in proportion to the
// XCD's block count, and each slice into near-equal shares of its blocks, so
// each XCD's L2 serves one contiguous part of the node and operator arrays
// and every XCD gets work also when n is below the block count.
constexpr int NXCD = 8;
__device__ __forceinline__ void xcd_share(int n, int& lo, int& hi) {
  const int nb = gridDim.x;
  const int x = blockIdx.x % NXCD, j = blockIdx.x / NXCD;
  const int q = nb / NXCD, r = nb % NXCD;
  const int nbx = q + (x < r ? 1 : 0);
  const long long s0 = (long long)n * (x * q + min(x, r)) / nb;
  const long long s1 = (long long)n * ((x + 1) * q + min(x + 1, r)) / nb;
  lo = (int)(s0 + (s1 - s0) * j / nbx);
  hi = (int)(s0 + (s1 - s0) * (j + 1) / nbx);
}

__device__ __forceinline__ bool overflowed(const DevBatch& B) {
  return __hip_atomic_load(B.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// chain of a segment node by binary search over the chains' first nodes
// (a few cached loads; no per-node table read)
__device__ __forceinline__ int chain_of_node(const DevBatch& B, int node) {
  int lo = 0, n = B.nchains;
  while (n > 1) {
    const int h = n >> 1;
    if (B.chain_node0[lo + h] <= node) { lo += h; n -= h; } else { n = h; }
  }
  return lo;
}

// chain of `node`, stepping forward from chain c (one at or before it) over
// the few cached chain starts; a binary search after eight steps
__device__ __forceinline__ int chain_from(const DevBatch& B, int c, int node) {
  int t = 0;
  while (t < 8 && c + 1 < B.nchains && B.chain_node0[c + 1] <= node) { ++c; ++t; }
  return t == 8 ? chain_of_node(B, node) : c;
}

// Chain of a segment node.  The site of a segment (node_site) is stored for
// the first segment of each site only, the one the estimator needs: storing
// it for every segment of a batch cost the linking kernels a scattered 4-byte
// store per segment.
__device__ __forceinline__ int node_chain(const DevBatch& B, int node) {
  return B.nchains == 1 ? 0 : chain_of_node(B, node);
}

// site (global index) whose first segment is `node`
__device__ __forceinline__ int first_node_site(const DevBatch& B, int node) {
  return B.node_site[node];
}

// Lower bound of each lane's time in a site's event times t[0..m) (ascending),
// for a lane group whose valid times ascend with the lane and are not below
// the group-uniform cursor `cur` (events before cur are below every time).
// The group reads the line in windows of G events from the cursor, one load
// per lane, and each lane searches the window by shuffles; the cursor moves
// to the last lane's result.  `equal`: the event at the result has the time.
template <int G>
__device__ __forceinline__ int group_lower_bound(const double* t, int m, double tau, int& cur, bool& equal) {
  const int gl = glane<G>();
  const int lane0 = lane_id() & ~(G - 1);
  int res = -1;
  equal = false;
  for (int pos = cur;; pos += G) {
    const double w = (pos + gl < m) ? t[pos + gl] : 3.0;  // past the end: above every time
    int k = 0;
#pragma unroll
    for (int step = G / 2; step >= 1; step >>= 1) {
      const double v = __shfl(w, lane0 + k + step - 1, 64);
      if (v < tau) k += step;
    }
    const double wk = __shfl(w, lane0 + k, 64);
    if (wk < tau) k += 1;  // k in [0, G]: window events below tau
    if (res < 0 && (k < G || pos + G >= m)) {
      res = pos + k;
      equal = (k < G) && (wk == tau);
    }
    if (gballot<G>(res < 0) == 0) break;
  }
  cur = __shfl(res, lane0 + G - 1, 64);
  return res;
}

// spin of the previous world line of a site at each lane's time; -1 where an
// event of the line has exactly that time (such candidates are rejected)
// (groups of fewer than 16 lanes search each lane's time on its own: their
// windows would cover too few events)
template <int G>
__device__ __forceinline__ int group_spin_prev(const DevBatch& B, int base, int m, double tau, int& cur) {
  const double* t = B.evt_old + base;
  int j;
  bool eq;
  if (G >= 16) {
    j = group_lower_bound<G>(t, m, tau, cur, eq);
  } else {
    j = tau < 0 ? 0 : lower_bound_d(t, m, tau);
    eq = j < m && t[j] == tau;
  }
  if (eq) return -1;
  return B.segspin_old[base + (j == m ? 0 : j)];
}

// Both sites' previous world lines staged in LDS (a wave per bond, lines of
// at most DIAG_LDS events): each lane's time is searched in both lines at
// once, one LDS read per line and step, instead of a window of shuffles per
// 64 events.  Same results as group_spin_prev.
constexpr int DIAG_LDS = 384;

__device__ __forceinline__ void lds_spins_prev(const DevBatch& B, const double* t0, int base0, int m0,
                                               const double* t1, int base1, int m1, double tau, int& c0,
                                               int& c1) {
  int lo0 = 0, n0 = m0, lo1 = 0, n1 = m1;
  while (n0 > 0 || n1 > 0) {  // same trip count on every lane (m0, m1 are the bond's)
    const int h0 = n0 >> 1, h1 = n1 >> 1;
    const double v0 = t0[min(lo0 + h0, DIAG_LDS - 1)], v1 = t1[min(lo1 + h1, DIAG_LDS - 1)];
    if (n0 > 0) { if (v0 < tau) { lo0 += h0 + 1; n0 -= h0 + 1; } else { n0 = h0; } }
    if (n1 > 0) { if (v1 < tau) { lo1 += h1 + 1; n1 -= h1 + 1; } else { n1 = h1; } }
  }
  const bool eq0 = lo0 < m0 && t0[min(lo0, DIAG_LDS - 1)] == tau;
  const bool eq1 = lo1 < m1 && t1[min(lo1, DIAG_LDS - 1)] == tau;
  const int s0 = B.segspin_old[base0 + (lo0 == m0 ? 0 : lo0)];
  const int s1 = B.segspin_old[base1 + (lo1 == m1 ? 0 : lo1)];
  c0 = eq0 ? -1 : s0;
  c1 = eq1 ? -1 : s1;
}

// Status word of a scan chunk: launch epoch (30 bits), state (1 sum, 2
// inclusive prefix), value.
__device__ __forceinline__ unsigned long long scan_word(uint32_t epoch, int state, int v) {
  return ((unsigned long long)epoch << 34) | ((unsigned long long)state << 32) | (uint32_t)v;
}

// one thread, in the kernel before a scan: clear the ticket, next epoch
// (never 0, the value of freshly cleared status words)
__device__ __forceinline__ void scan_advance_epoch(const DevBatch& B) {
  B.scan_ctl[0] = 0;
  uint32_t e = ((uint32_t)B.scan_ctl[1] + 1u) & 0x3fffffffu;
  B.scan_ctl[1] = (int)(e == 0 ? 1u : e);
}

// ---------------------------------------------------------------------------
// (a) diagonal update, pass 1: count and store accepted candidates
// ---------------------------------------------------------------------------
// One bond's candidates: kept off-diagonal operators counted, Poisson
// candidates drawn and tested against the previous configuration's spins,
// the accepted ones stored.  CLS: BOND_LATTICE, BOND_INSITE (the
// single-ion term's in-site bonds: their own rate and graphs) or BOND_SITE
// (a site bond of the transverse field, whose candidates come at rate
// beta |hx|/2 with the one site graph g = 0, compatible with either spin,
// graph_impl.h:66-69).  The lattice bonds and the others are separate
// kernels, so that the lattice case keeps its registers.
enum : int { BOND_LATTICE = 0, BOND_INSITE = 1, BOND_SITE = 2 };

template <int G, int CLS>
__device__ __forceinline__ void diag_count_bond(const DevBatch& B, const ChainDev& P, int c, int gb, int gl,
                                                uint64_t below) {
  constexpr bool SITE = CLS == BOND_SITE;
  const int b = gb - P.bond_base;
  const uint32_t sweep = B.sweep[c];
  int s0, s1;
  if constexpr (SITE) s0 = s1 = b - P.nb_real;
  else if constexpr (CLS == BOND_INSITE) dpair_sites(P, b, s0, s1);
  else bond_sites(P, b, s0, s1);
  // kept (off-diagonal) operators of the previous configuration
  const int ob = P.old_bond_base + b;
  const int o0 = B.bo_old[ob], o1 = B.bo_old[ob + 1];
  int kept = 0;
  for (int r0 = o0; r0 < o1; r0 += G) {
    int r = r0 + gl;
    bool off = (r < o1) && (B.oi_old[r] & 1);
    uint64_t m = gballot<G>(off);
    if (r < o1) B.kpre[r] = kept + __popcll(m & below);
    kept += __popcll(m);
  }
  // Poisson candidates on [0,1): exponential gaps on the 2^-52 grid
  int acc = 0;
  if (SITE ? P.site_w > 0 : (CLS == BOND_INSITE ? P.dwsum > 0 : P.has_weight != 0)) {
    const double scale = SITE ? P.site_scale : (CLS == BOND_INSITE ? P.dscale : P.scale);
    // slot j of the bond at j * cand_sstride (a chain holds fewer than 2^31 slots)
    const long long cb = P.cand_base + (long long)b * P.cand_bstride;
    double* const ctb = B.cand_t + cb;
    uint8_t* const cgb = B.cand_g + cb;
    const int base0 = B.soff_old[P.old_site_base + s0], m0 = B.soff_old[P.old_site_base + s0 + 1] - base0;
    const int base1 = B.soff_old[P.old_site_base + s1], m1 = B.soff_old[P.old_site_base + s1 + 1] - base1;
    int cur0 = 0, cur1 = 0;  // world-line cursors (candidate times ascend)
    // a wave per bond: the two lines in LDS when they fit (a wave-uniform
    // choice; the wave's own slice, so no block barrier)
    double* lt0 = nullptr;
    double* lt1 = nullptr;
    if constexpr (G == 64) {
      __shared__ double lines[256 / 64][2][DIAG_LDS];
      if (m0 <= DIAG_LDS && m1 <= DIAG_LDS) {
        lt0 = lines[threadIdx.x >> 6][0];
        lt1 = lines[threadIdx.x >> 6][1];
        __builtin_amdgcn_wave_barrier();  // the previous bond's searches are done
        for (int q = gl; q < m0; q += 64) lt0[q] = B.evt_old[base0 + q];
        for (int q = gl; q < m1; q += 64) lt1[q] = B.evt_old[base1 + q];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
    uint64_t Tb = 0;
    uint32_t kb = 0;
    while (true) {
      U4 x = philox(sweep, STAGE_DIAG, (uint32_t)b, kb + (uint32_t)gl, P.key0, P.key1);
      uint64_t gap = exp_gap(x, scale);
      uint64_t incl = gscan_incl_u64<G>(gap);
      uint64_t T = Tb + incl;
      const bool valid = T < TGRID;
      // lanes past the end of [0,1) search for -1 (found at once)
      const double tau = valid ? (double)T * TGRID_INV : -1.0;
      const int g = SITE ? 0 : (CLS == BOND_INSITE ? choose_graph_d(x.z, P) : choose_graph(x.z, P));
      int c0, c1;
      if (lt0) {
        lds_spins_prev(B, lt0, base0, m0, lt1, base1, m1, tau, c0, c1);
      } else {
        c0 = group_spin_prev<G>(B, base0, m0, tau, cur0);
        c1 = group_spin_prev<G>(B, base1, m1, tau, cur1);
      }
      const bool accept = valid && (c0 >= 0) && (c1 >= 0) && (SITE || ((g & 1) ^ c0 ^ c1) != 0);
      uint64_t m = gballot<G>(accept);
      if (accept) {
        int slot = acc + __popcll(m & below);
        if (slot < P.cand_cap) {
          ctb[slot * P.cand_sstride] = tau;
          cgb[slot * P.cand_sstride] = (uint8_t)g;
        } else {
          atomicOr(B.flags, FLAG_CAND_OVERFLOW);
        }
      }
      acc += __popcll(m);
      Tb += gbcast_u64<G>(incl, G - 1);
      kb += G;
      if (Tb >= TGRID) break;
      // draws beyond 4x the slot count: capacity error (keeps the loop bounded)
      if (kb > 4u * (uint32_t)P.cand_cap + 64u) {
        if (gl == 0) atomicOr(B.flags, FLAG_CAND_OVERFLOW);
        break;
      }
    }
    if (acc > P.cand_cap) acc = P.cand_cap;
  }
  if (gl == 0) {
    B.counts[gb] = kept + acc;
    B.acc_count[gb] = acc;
  }
}

template <int G>
// (at most 96 VGPRs, five waves per SIMD, where that needs no spills)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(G == 32 ? 4 : 5))) k_diag_count(DevBatch B) {
  if (blockIdx.x == 0 && threadIdx.x == 0) scan_advance_epoch(B);  // for k_scan, which follows
  if (overflowed(B)) return;
  const int groups_per_block = blockDim.x / G;
  const int gl = glane<G>();
  const uint64_t below = (gl == 0) ? 0ull : ((1ull << gl) - 1);
  // bonds interleaved over the grid: the work per bond varies with the
  // chain's temperature in a batch
  int gb_first = blockIdx.x * groups_per_block + threadIdx.x / G;
  if (G == 64) gb_first = __builtin_amdgcn_readfirstlane(gb_first);  // one bond per wave: uniform
  for (int gb = gb_first; gb < B.nbonds;
       gb += gridDim.x * groups_per_block) {
    const int c = B.bond_chain[gb];
    const ChainDev P = B.chains[c];
    if (gb - P.bond_base >= P.nb_lat) continue;  // in-site and site bonds: k_diag_count_extra
    diag_count_bond<G, BOND_LATTICE>(B, P, c, gb, gl, below);
  }
}

// the bonds beyond the lattice's, a group of G lanes per (virtual) site:
// CLS BOND_INSITE, the in-site bonds to its real site's higher virtual sites
// (single-ion term); BOND_SITE, its site bond (transverse field).  One kernel
// per class, each with the registers of one instance.
template <int G, int CLS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CLS == BOND_INSITE || G == 32 ? 4 : 5)))
k_diag_count_extra(DevBatch B) {
  if (overflowed(B)) return;
  const int groups_per_block = blockDim.x / G;
  const int gl = glane<G>();
  const uint64_t below = (gl == 0) ? 0ull : ((1ull << gl) - 1);
  int gs_first = blockIdx.x * groups_per_block + threadIdx.x / G;
  if (G == 64) gs_first = __builtin_amdgcn_readfirstlane(gs_first);
  for (int gs = gs_first; gs < B.nsites; gs += gridDim.x * groups_per_block) {
    const int c = B.site_chain[gs];
    const ChainDev P = B.chains[c];
    const int s = gs - P.site_base;
    if constexpr (CLS == BOND_INSITE) {
      if (P.npair > 0) {
        const int r = s / P.nv, a = s - r * P.nv;
        for (int o = a + 1; o < P.nv; ++o)
          diag_count_bond<G, BOND_INSITE>(B, P, c, P.bond_base + P.nb_lat + r * P.npair + pair_index(P.nv, a, o), gl,
                                          below);
      }
    } else {
      if (P.Nb > P.nb_real) diag_count_bond<G, BOND_SITE>(B, P, c, P.bond_base + P.nb_real + s, gl, below);
    }
  }
}

// ---------------------------------------------------------------------------
// scans: bond operator offsets and site segment offsets
// ---------------------------------------------------------------------------
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_CHUNK = SCAN_ITEMS * SCAN_THREADS;

__device__ __forceinline__ int site_events(const DevBatch& B, int gs) {
  const ChainDev& P = B.chains[B.site_chain[gs]];
  int s = gs - P.site_base;
  int m = 0;
  if (P.nv == 1) {
    int bl[4];
    const int nb = site_bonds(P, s, bl);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i < nb) m += B.counts[P.bond_base + bl[i]];
  } else {  // virtual sites: up to eight lists
    const int nb = site_lists(P);
    for (int i = 0; i < nb; ++i) m += B.counts[P.bond_base + site_bond(P, s, i)];
  }
  if (B.any_extra) {
    for (int j = 0, nd = P.npair > 0 ? P.nv - 1 : 0; j < nd; ++j) m += B.counts[P.bond_base + dpair_bond(P, s, j)];  // nv - 1 in-site lists
    if (P.Nb > P.nb_real) m += B.counts[P.bond_base + P.nb_real + s];  // site operators
  }
  return m;
}

__device__ __forceinline__ int scan_value(const DevBatch& B, int which, int i) {
  if (which == 0) return B.counts[i];
  int m = site_events(B, i);
  return m > 0 ? m : 1;
}

__device__ int block_excl_scan(int v, int* total) {
  __shared__ int wsum[SCAN_THREADS / 64];
  int lane = lane_id(), w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int o = __shfl_up(x, d, 64);
    if (lane >= d) x += o;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < SCAN_THREADS / 64; ++i) {
    int s = wsum[i];
    if (i < w) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + x - v;
}

// Single-pass scan with decoupled look-back (Merrill & Garland, "Single-pass
// parallel prefix scan with decoupled look-back", NVIDIA TR 2016).  Blocks
// take chunks in the order they start (a ticket), so a block only waits for
// blocks that are already running.  Chunks [0, nbB) scan the bonds' operator
// counts, the rest the sites' segment counts; each chunk publishes its sum,
// then its inclusive prefix, in one 64-bit status word tagged with the
// launch's epoch (k_diag_count or k_scan_reset advances the epoch and clears
// the ticket before the scan), so the status words never need clearing.
__global__ void k_scan_reset(DevBatch B) { scan_advance_epoch(B); }

__global__ void __launch_bounds__(SCAN_THREADS) k_scan(DevBatch B) {
  if (overflowed(B)) return;
  __shared__ int s_ticket, s_excl;
  if (threadIdx.x == 0) s_ticket = atomicAdd(B.scan_ctl, 1);
  __syncthreads();
  const int t = s_ticket;
  const uint32_t epoch = (uint32_t)B.scan_ctl[1];
  const int nbB = (B.nbonds + SCAN_CHUNK - 1) / SCAN_CHUNK;
  int blk = t, which = 0, n = B.nbonds, first = 0;
  if (blk >= nbB) { blk -= nbB; which = 1; n = B.nsites; first = nbB; }
  const int base = blk * SCAN_CHUNK + threadIdx.x * SCAN_ITEMS;
  int vals[SCAN_ITEMS];
  bool empty[SCAN_ITEMS];
  int sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int v = 0;
    empty[k] = false;
    if (base + k < n) {
      if (which == 0) {
        v = B.counts[base + k];
      } else {
        int m = site_events(B, base + k);
        empty[k] = (m == 0);
        // a virtual site's segments do not wrap: m events, m + 1 segments
        v = empty[k] ? 1 : (B.chains[B.site_chain[base + k]].nv > 1 ? m + 1 : m);
      }
    }
    vals[k] = v;
    sum += v;
  }
  int tot;
  const int loc = block_excl_scan(sum, &tot);
  if (threadIdx.x < 64) {  // the first wave looks back, 64 predecessors at a time
    const int lane = threadIdx.x;
    int excl = 0;
    unsigned long long* st = B.scan_status;
    if (t == first) {
      if (lane == 0) __hip_atomic_store(st + t, scan_word(epoch, 2, tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (lane == 0) __hip_atomic_store(st + t, scan_word(epoch, 1, tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // every predecessor down to `first` holds a lower ticket, so it is
      // running and publishes its sum without waiting; `first` publishes its
      // inclusive prefix, so the window holding it ends the walk
      for (int top = t - 1;; top -= 64) {
        const int p = top - lane;
        const bool valid = p >= first;
        unsigned long long w = 0;
        if (valid) {
          while ((uint32_t)((w = __hip_atomic_load(st + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >> 34) != epoch)
            __builtin_amdgcn_s_sleep(1);
        }
        const unsigned long long incl = __ballot(valid && ((w >> 32) & 3) == 2);
        // nearest inclusive prefix: lanes up to it contribute, else the window
        const int k = incl ? __ffsll((long long)incl) - 1 : 63;
        int v = (valid && lane <= k) ? (int)(uint32_t)w : 0;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
        excl += v;
        if (incl) break;
      }
      if (lane == 0) __hip_atomic_store(st + t, scan_word(epoch, 2, excl + tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0) s_excl = excl;
  }
  __syncthreads();
  const int cap = which ? B.node_cap : B.op_cap;
  int off = loc + s_excl;
  int* out = which ? B.soff_new : B.bo_new;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int i = base + k;
    if (i < n) {
      out[i] = off;
      if (which) {
        const int c = B.site_chain[i];
        const ChainDev& P = B.chains[c];
        if (i == P.site_base) B.chain_node0[c] = off;
        if (empty[k] && off < cap) {
          // event-free site: one node spanning [0,1)
          int node = off;
          B.evt_new[node] = EMPTY_TIME;
          B.segspin_new[node] = B.spins_old[P.old_site_base + (i - P.site_base)];
          B.evflags[node] = (uint8_t)(EV_EMPTY | EV_FIRST | (gauge(P, i - P.site_base) < 0 ? EV_GAUGE : 0) |
                                      (P.nv > 1 ? EV_NOWRAP : 0));
          B.parent[node] = node;
          B.node_site[node] = i;
        }
      }
    }
    off += vals[k];
  }
  if (blk == (n - 1) / SCAN_CHUNK && threadIdx.x == 0) {
    const int total = s_excl + tot;
    B.totals[which] = total;
    out[n] = total;
    if (total > cap) atomicOr(B.flags, which ? FLAG_NODE_OVERFLOW : FLAG_OP_OVERFLOW);
  }
}

// ---------------------------------------------------------------------------
// (a) diagonal update, pass 2: merge kept operators and accepted candidates
// ---------------------------------------------------------------------------
// times per list staged in LDS (a wave per bond): fill_cap, chosen per
// launch from the expected operators per bond; dynamic LDS of
// 256 / 64 waves * 2 lists * fill_cap doubles

template <int G>
__global__ void __launch_bounds__(256) k_diag_fill(DevBatch B, int fill_cap) {
  if (overflowed(B)) return;
  const int groups_per_block = blockDim.x / G;
  const int gl = glane<G>();
  int gb_first = blockIdx.x * groups_per_block + threadIdx.x / G;
  if (G == 64) gb_first = __builtin_amdgcn_readfirstlane(gb_first);  // one bond per wave: uniform
  for (int gb = gb_first; gb < B.nbonds;
       gb += gridDim.x * groups_per_block) {
    const int c = B.bond_chain[gb];
    const ChainDev P = B.chains[c];
    const int b = gb - P.bond_base;
    const uint32_t sweep = B.sweep[c];
    const int ob = P.old_bond_base + b;
    const int o0 = B.bo_old[ob], o1 = B.bo_old[ob + 1];
    const int nb = B.bo_new[gb];
    const int acc = B.acc_count[gb];
    const int kept = B.counts[gb] - acc;
    const long long cb = P.cand_base + (long long)b * P.cand_bstride;  // slot j at cb + j * cs
    const long long cs = P.cand_sstride;
    const double* ct = B.cand_t + cb;
    const int info_b = gb << 3;
    // tc: accepted candidates' times (slot j at tc[j * ts]); to: the old
    // operators' times
    auto place = [&](const double* tc, long long ts, const double* to) {
      // kept operators: regraph (graph_impl.h:323-327) and place
      for (int r = o0 + gl; r < o1; r += G) {
        if (!(B.oi_old[r] & 1)) continue;
        const int kr = B.kpre[r];
        const double tau = B.ot_old[r];
        U4 x = philox(sweep, STAGE_REGRAPH, (uint32_t)b, (uint32_t)kr, P.key0, P.key1);
        // an in-site bond regraphs by its own threshold; a site operator
        // keeps the site graph
        const uint64_t toff = (B.any_extra && b >= P.nb_lat) ? (b >= P.nb_real ? (1ull << 32) : P.dtoff) : P.toff;
        int g = ((uint64_t)x.x < toff) ? 0 : 1;
        int pos = nb + kr + lower_bound_ds(tc, ts, acc, tau);
        B.ot_new[pos] = tau;
        B.oi_new[pos] = info_b | (g << 1) | 1;
        B.kc_new[pos] = kr;
      }
      // accepted candidates: new diagonal operators
      for (int j = gl; j < acc; j += G) {
        const double tau = tc[j * ts];
        const int g = B.cand_g[cb + j * cs];
        int r = lower_bound_d(to, o1 - o0, tau) + o0;
        int kk = (r < o1) ? B.kpre[r] : kept;
        int pos = nb + j + kk;
        B.ot_new[pos] = tau;
        B.oi_new[pos] = info_b | (g << 1);
        B.kc_new[pos] = kk;
      }
    };
    // a wave per bond: both lists searched in LDS when they fit (the wave's
    // own slice, no block barrier)
    if constexpr (G == 64) {
      extern __shared__ double fill_lists[];
      if (acc <= fill_cap && o1 - o0 <= fill_cap) {
        double* lc = fill_lists + (threadIdx.x >> 6) * 2 * fill_cap;
        double* lo = lc + fill_cap;
        __builtin_amdgcn_wave_barrier();  // the previous bond's searches are done
        for (int q = gl; q < acc; q += 64) lc[q] = ct[q * cs];
        for (int q = gl; q < o1 - o0; q += 64) lo[q] = B.ot_old[o0 + q];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        place(lc, 1, lo);
        continue;
      }
    }
    place(ct, cs, B.ot_old + o0);
  }
}

// ---------------------------------------------------------------------------
// (b) world-line linking, a group of threads per site.  The events of a
// site are the operators of its bonds (each bond list sorted by tau); an
// event's rank on the site is its index in its own list plus, in each other
// list, the number of operators before it (ties ordered by bond index), and
// the spin of the segment below it flips with the off-diagonal operators
// counted the same way.  The lists (times and off-diagonal prefix counts) are staged in LDS,
// ranked there, and the site's segment arrays are written in node order
// through the inverse permutation (whole cache lines).  A site with more
// than CAP events searches the lists in global memory and writes directly.
// ---------------------------------------------------------------------------
struct SiteLists {
  int nbd;
  int bond[4];   // chain-local bond index
  int lo[4];     // first operator of the list
  int len[4];
  int off[4];    // staged offset of the list
  int side[4];   // 1 when the site is the bond's target site
};

// off-diagonal operators among the first q of list j (q <= len)
__device__ __forceinline__ int list_pre_global(const DevBatch& B, const SiteLists& S, int j, int q) {
  if (q < S.len[j]) return B.kc_new[S.lo[j] + q];
  if (S.len[j] == 0) return 0;
  const int last = S.lo[j] + S.len[j] - 1;
  return B.kc_new[last] + (B.oi_new[last] & 1);
}

// rank and kink parity of an event of list i (bond bi) at position p; the
// loops over the lists are unrolled so that SiteLists stays in registers
template <bool kLds>
__device__ __forceinline__ void rank_event(const DevBatch& B, const SiteLists& S, int i, int bi, int p,
                                           double tau, const double* lt, const uint16_t* lpre, int& rank,
                                           int& par) {
  rank = p;
  par = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (j >= S.nbd) break;
    if (j == i) {
      par += kLds ? lpre[S.off[j] + j + p] : list_pre_global(B, S, j, p);
      continue;
    }
    const double* t = kLds ? lt + S.off[j] : B.ot_new + S.lo[j];
    const int q = (S.bond[j] < bi) ? upper_bound_d(t, S.len[j], tau) : lower_bound_d(t, S.len[j], tau);
    rank += q;
    par += kLds ? lpre[S.off[j] + j + q] : list_pre_global(B, S, j, q);
  }
}

// union-find on block-local segment indices in LDS (same hooking rule as
// the global one: the larger root goes under the smaller)
__device__ __forceinline__ int lds_find(int* P, int x) {
  int p;
  while ((p = __hip_atomic_load(P + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) != x) x = p;
  return x;
}
__device__ __forceinline__ void lds_unite(int* P, int a, int b) {
  a = lds_find(P, a);
  b = lds_find(P, b);
  while (a != b) {
    if (a < b) { int t = a; a = b; b = t; }
    const int old = atomicCAS(P + a, a, b);
    if (old == a) return;
    a = lds_find(P, old);
    b = lds_find(P, b);
  }
}

// One block step handles SITES = 64 * WAVES / TPS consecutive sites, TPS
// threads per site (several waves per site for long world lines, where one
// wave would rank each list element by element; a quarter wave for short
// ones, where a wave would idle most lanes).  Operators of a bond whose two sites
// are both in the step ("inner" operators: neighbours along the fast lattice
// direction) are united here in LDS, and each segment's parent is written as
// its local root; k_uf_union skips them (u1 is stored complemented).  Roots
// stay cluster minima, so the partition and every result are those of the
// plain union-find.
constexpr int LINK_U = 4;  // staged elements per thread in flight

template <int CAP, int WAVES, int TPS>
__global__ void __launch_bounds__(64 * WAVES) k_link(DevBatch B) {
  if (overflowed(B)) return;
  constexpr int SITES = 64 * WAVES / TPS;
  __shared__ double lt_all[SITES][CAP];           // staged times, list order; then local parents
  __shared__ uint16_t lpre_all[SITES][CAP + 4];   // off-diagonal prefix counts; then element ranks
  __shared__ uint16_t inv_all[SITES][CAP];        // rank -> staged element
  __shared__ uint8_t lfl_all[SITES][CAP];         // staged element -> event flags | spin below << 7
  __shared__ int wn0[SITES], wn1[SITES], wtot[SITES], wnbd[SITES];
  __shared__ int wbond[SITES][4], woff[SITES][4];
  // site slot of the step; with whole waves per site it is wave-uniform (read
  // as such, which keeps the site's list descriptors in scalar registers)
  const int w = TPS >= 64 ? __builtin_amdgcn_readfirstlane(threadIdx.x / TPS) : (int)(threadIdx.x / TPS);
  const int sl = threadIdx.x % TPS;               // thread within the site's group
  double* lt = lt_all[w];
  uint16_t* lpre = lpre_all[w];
  uint16_t* inv = inv_all[w];
  uint8_t* lfl = lfl_all[w];
  int* lpar = reinterpret_cast<int*>(&lt_all[0][0]);  // SITES * CAP * 2 ints
  int slo, shi;
  xcd_share(B.nsites, slo, shi);
  for (int g = slo; g < shi; g += SITES) {  // block-uniform steps
    const int gs = g + w;
    const bool have = gs < shi;
    SiteLists S;
    S.nbd = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) { S.bond[j] = -1; S.off[j] = S.lo[j] = S.len[j] = S.side[j] = 0; }
    int tot = 0, s = 0, n0 = 0, n1 = 0;
    int other[4] = {-1, -1, -1, -1};   // the bond's other site (global index)
    uint8_t spin0 = 0, gflag = 0;
    const ChainDev* Pp = nullptr;
    if (have) {
      Pp = &B.chains[B.site_chain[gs]];
      const ChainDev& P = *Pp;
      s = gs - P.site_base;
      int bl[4];
      S.nbd = site_bonds(P, s, bl);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j >= S.nbd) break;
        const int gb = P.bond_base + bl[j];
        int s0, s1;
        bond_sites(P, bl[j], s0, s1);
        S.bond[j] = bl[j];
        S.lo[j] = B.bo_new[gb];
        S.len[j] = B.bo_new[gb + 1] - S.lo[j];
        S.off[j] = tot;
        S.side[j] = (s0 == s) ? 0 : 1;
        other[j] = P.site_base + ((s0 == s) ? s1 : s0);
        tot += S.len[j];
      }
      n0 = B.soff_new[gs];
      n1 = B.soff_new[gs + 1];
      spin0 = B.spins_old[P.old_site_base + s];
      gflag = gauge(P, s) < 0 ? EV_GAUGE : 0;
      if (tot <= CAP) {
        // LINK_U elements per thread at a time, all loads before any store:
        // time, off-diagonal prefix count and off-diagonal bit (kept in the
        // flags byte until the element's flags replace it)
        for (int e0 = sl; e0 < tot; e0 += LINK_U * TPS) {
          int ek[LINK_U], ej[LINK_U];
          double et[LINK_U];
          int ekc[LINK_U], eoi[LINK_U];
#pragma unroll
          for (int u = 0; u < LINK_U; ++u) {
            const int e = min(e0 + u * TPS, tot - 1);
            int j = 0;
#pragma unroll
            for (int jj = 1; jj < 4; ++jj)
              if (jj < S.nbd && e >= S.off[jj]) j = jj;
            ej[u] = j;
            ek[u] = (j == 0 ? S.lo[0] - S.off[0] : j == 1 ? S.lo[1] - S.off[1] : j == 2 ? S.lo[2] - S.off[2]
                                                                                 : S.lo[3] - S.off[3]) + e;
          }
#pragma unroll
          for (int u = 0; u < LINK_U; ++u) {
            et[u] = B.ot_new[ek[u]];
            ekc[u] = B.kc_new[ek[u]];
            eoi[u] = B.oi_new[ek[u]];
          }
#pragma unroll
          for (int u = 0; u < LINK_U; ++u) {
            const int e = e0 + u * TPS;
            if (e >= tot) break;
            lt[e] = et[u];
            lpre[e + ej[u]] = (uint16_t)ekc[u];
            lfl[e] = (uint8_t)(eoi[u] & 1);
          }
        }
        // each list's total after its last element, one thread per list
        if (sl < S.nbd) {
          const int j = sl;  // selected with constant indices (S stays in registers)
          const int len = j == 0 ? S.len[0] : j == 1 ? S.len[1] : j == 2 ? S.len[2] : S.len[3];
          const int off = j == 0 ? S.off[0] : j == 1 ? S.off[1] : j == 2 ? S.off[2] : S.off[3];
          const int lo = j == 0 ? S.lo[0] : j == 1 ? S.lo[1] : j == 2 ? S.lo[2] : S.lo[3];
          const int last = len > 0 ? lo + len - 1 : 0;  // (operator 0 stands in for an empty list)
          const int v = B.kc_new[last] + (B.oi_new[last] & 1);
          lpre[off + j + len] = (uint16_t)(len > 0 ? v : 0);
        }
      }
      if (sl == 0) {
        wn0[w] = n0;
        wn1[w] = n1;
        wtot[w] = tot;
        wnbd[w] = S.nbd;
#pragma unroll
        for (int j = 0; j < 4; ++j) { wbond[w][j] = S.bond[j]; woff[w][j] = S.off[j]; }
      }
    } else if (sl == 0) {
      wtot[w] = -1;
    }
    __syncthreads();
    // local unions need every site of the step staged and in one chain
    bool inner_ok = true;
    const int c0 = B.site_chain[g];
    for (int v = 0; v < SITES; ++v)
      if (g + v < shi && (wtot[v] > CAP || B.site_chain[g + v] != c0)) inner_ok = false;
    if (have && tot > 0) {
      const bool lds = tot <= CAP;
      for (int e = sl; e < tot; e += TPS) {
        // the element's list, selected with constant indices
        int i = 0, offi = S.off[0], loi = S.lo[0], bi = S.bond[0], side = S.side[0], oth = other[0];
#pragma unroll
        for (int j = 1; j < 4; ++j)
          if (j < S.nbd && e >= S.off[j]) {
            i = j; offi = S.off[j]; loi = S.lo[j]; bi = S.bond[j]; side = S.side[j]; oth = other[j];
          }
        const int p = e - offi;
        const int k = loi + p;
        const double tau = lds ? lt[e] : B.ot_new[k];
        int rank, par;
        if (lds) rank_event<true>(B, S, i, bi, p, tau, lt, lpre, rank, par);
        else rank_event<false>(B, S, i, bi, p, tau, nullptr, nullptr, rank, par);
        const uint8_t fl = (uint8_t)(((lds ? lfl[e] : B.oi_new[k]) & 1) | (side ? 0 : EV_SRC) |
                                     (bond_dir(*Pp, bi) ? EV_DIR1 : 0) | (rank == 0 ? EV_FIRST : 0) | gflag);
        const uint8_t sp = spin0 ^ (uint8_t)(par & 1);
        const int node = n0 + rank;
        const int up = (node + 1 == n1) ? n0 : node + 1;  // segment above the event
        const bool inner = inner_ok && oth >= g && oth < g + SITES && oth < shi;
        if (side) { B.l1[k] = node; B.u1[k] = inner ? ~up : up; } else { B.l0[k] = node; B.u0[k] = up; }
        if (lds) {
          inv[rank] = (uint16_t)e;
          lfl[e] = (uint8_t)(fl | (sp << 7));  // flags use bits 0-5
        } else {
          B.evt_new[node] = tau;
          B.segspin_new[node] = sp;
          B.evflags[node] = fl;
          B.parent[node] = node;
          if (rank == 0) B.node_site[node] = gs;
        }
      }
    }
    __syncthreads();
    // segment arrays in node order; element ranks for the local unions
    if (have && tot > 0 && tot <= CAP) {
      for (int r = sl; r < tot; r += TPS) {
        const int e = inv[r];
        B.evt_new[n0 + r] = lt[e];
        const uint8_t f = lfl[e];
        B.segspin_new[n0 + r] = f >> 7;
        B.evflags[n0 + r] = f & 0x7f;
        if (r == 0) B.node_site[n0 + r] = gs;
        lpre[e] = (uint16_t)r;
      }
    }
    __syncthreads();
    // step's segments [N0, N1): local unions of the inner operators
    const int last = min(g + SITES, shi) - 1 - g;
    const int N0 = wn0[0], N1 = wn1[last];
    if (inner_ok) {
      for (int x = threadIdx.x; x < N1 - N0; x += blockDim.x) lpar[x] = x;
      __syncthreads();
      if (have && tot > 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (j >= S.nbd || S.side[j] != 0) continue;
          const int wt = other[j] - g;
          if (other[j] < g || wt >= SITES || other[j] >= shi) continue;
          int jt = 0;
          while (jt < 3 && wbond[wt][jt] != S.bond[j]) ++jt;
          const uint16_t* rk_t = lpre_all[wt];
          const int t0 = wn0[wt] - N0, t1 = wn1[wt] - N0, m0 = n0 - N0, m1 = n1 - N0;
          for (int p = sl; p < S.len[j]; p += TPS) {
            const int a0 = m0 + lpre[S.off[j] + p];
            const int b0 = (a0 + 1 == m1) ? m0 : a0 + 1;
            const int a1 = t0 + rk_t[woff[wt][jt] + p];
            const int b1 = (a1 + 1 == t1) ? t0 : a1 + 1;
            const int gr = (B.oi_new[S.lo[j] + p] >> 1) & 3;
            if (gr & 2) {            // frozen: all four legs
              lds_unite(lpar, a0, a1);
              lds_unite(lpar, b0, b1);
              lds_unite(lpar, a0, b0);
            } else if (gr == 0) {    // horizontal cut
              lds_unite(lpar, a0, a1);
              lds_unite(lpar, b0, b1);
            } else {                 // cross
              lds_unite(lpar, a0, b1);
              lds_unite(lpar, a1, b0);
            }
          }
        }
      }
      __syncthreads();
      for (int x = threadIdx.x; x < N1 - N0; x += blockDim.x) B.parent[N0 + x] = N0 + lds_find(lpar, x);
    } else if (have && tot > 0 && tot <= CAP) {
      for (int r = sl; r < tot; r += TPS) B.parent[n0 + r] = n0 + r;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// (c) lock-free union-find on segments: hook the larger root under the
// smaller with atomicCAS, so every root is its cluster's minimum node (the
// hooking rule of ECL-CC, Jaiganesh & Burtscher, HPDC'18).  parent[x] <= x
// always holds.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int ld_parent(int* P, int x) {
  return __hip_atomic_load(P + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Plain walk to the root: shortening the path with stores while walking cost
// more write traffic than the longer walks it saves (k_estimate points every
// node at its root once all unions are done).
__device__ __forceinline__ int uf_find(int* P, int v) {
  int cur = v, next;
  while (cur != (next = ld_parent(P, cur))) cur = next;
  return cur;
}

__device__ __forceinline__ void uf_unite(int* P, int a, int b) {
  int ra = uf_find(P, a), rb = uf_find(P, b);
  while (ra != rb) {
    if (ra < rb) { int t = ra; ra = rb; rb = t; }
    // ra > rb: hook ra under rb
    int old = atomicCAS(P + ra, ra, rb);
    if (old == ra) return;
    ra = uf_find(P, old);
    rb = uf_find(P, rb);
  }
}

// every segment its own root, before the unions of k_link_ops
__global__ void __launch_bounds__(256) k_node_init(DevBatch B) {
  if (overflowed(B)) return;
  const int M = B.totals[1];
  int ilo, ihi;
  xcd_share(M, ilo, ihi);
  for (int i = ilo + threadIdx.x; i < ihi; i += blockDim.x) B.parent[i] = i;
}

// Linking, one thread per operator: for each leg, the rank of the event on
// its site by binary search in the site's other bond lists (global memory).
// Used when sites hold few events (a wave per site would idle most lanes).
// The thread then unites its operator's legs by its graph (the k_uf_union
// rules) right away: k_node_init made every segment a root beforehand, and
// the linking itself writes no parent, so the unions of different threads
// only meet in the lock-free union-find.  This saves the separate union
// pass, which re-read the four legs and the type of every operator.
// EXTRA: some chain has bonds beyond the lattice's (in-site bonds of the
// single-ion term: two legs, no bond vector; site operators of the
// transverse field); a separate instance, so that the plain one carries
// none of their tests.
template <bool EXTRA>
__global__ void __launch_bounds__(256) k_link_ops(DevBatch B) {
  if (overflowed(B)) return;
  const int nops = B.totals[0];
  int klo, khi;
  xcd_share(nops, klo, khi);
  for (int k = klo + threadIdx.x; k < khi; k += blockDim.x) {
    const int info = B.oi_new[k];
    const int gb = info >> 3;
    const ChainDev& P = B.chains[B.bond_chain[gb]];
    const int b = gb - P.bond_base;
    const double tau = B.ot_new[k];
    const int r = k - B.bo_new[gb];
    // a site operator (transverse field): one event, on its site, with no
    // bond vector (no source flag), and no union: the world line is cut
    const bool site = EXTRA && b >= P.nb_real;
    const bool lattice = !EXTRA || b < P.nb_lat;  // bonds with a bond vector
    const uint8_t dflag = (lattice && bond_dir(P, b)) ? EV_DIR1 : 0;
    int s0, s1;
    if (EXTRA) op_sites(P, b, s0, s1);
    else bond_sites(P, b, s0, s1);
    int lo_leg[2], up_leg[2];  // lower and upper segment of each side
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      if (side && site) {
        B.l1[k] = lo_leg[0];
        B.u1[k] = up_leg[0];
        break;
      }
      const int s = side ? s1 : s0;
      int rank = r;
      int par = B.kc_new[k];
      // the event's rank among the site's other lists (ties by bond index)
      // and the off-diagonal operators below it there
      auto other_list = [&](int bp) {
        if (bp == b) return;
        const int gbp = P.bond_base + bp;
        const int lo = B.bo_new[gbp], hi = B.bo_new[gbp + 1];
        const int q = (bp < b) ? upper_bound_d(B.ot_new + lo, hi - lo, tau) : lower_bound_d(B.ot_new + lo, hi - lo, tau);
        rank += q;
        if (q < hi - lo) par += B.kc_new[lo + q];
        else if (hi > lo) par += B.kc_new[hi - 1] + (B.oi_new[hi - 1] & 1);
      };
      if (P.nv == 1) {
        int bl[4];
        const int nbd = site_bonds(P, s, bl);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (i >= nbd) break;
          other_list(bl[i]);
        }
      } else {  // virtual sites: up to eight lists
        const int nbd = site_lists(P);
        for (int i = 0; i < nbd; ++i) other_list(site_bond(P, s, i));
      }
      if (EXTRA) {
        for (int j = 0, nd = P.npair > 0 ? P.nv - 1 : 0; j < nd; ++j) other_list(dpair_bond(P, s, j));  // nv - 1 in-site lists
        if (P.Nb > P.nb_real) other_list(P.nb_real + s);                            // the site's own operators
      }
      const int gs = P.site_base + s;
      const int n0 = B.soff_new[gs], n1 = B.soff_new[gs + 1];
      const int node = n0 + rank;
      const bool split = P.nv > 1;
      const int up = (node + 1 == n1) ? n0 : node + 1;  // segment above the event (split: never wraps)
      const uint8_t gflag = gauge(P, s) < 0 ? EV_GAUGE : 0;
      const uint8_t sp = B.spins_old[P.old_site_base + s] ^ (uint8_t)(par & 1);
      B.evt_new[node] = tau;
      B.segspin_new[node] = sp;
      B.evflags[node] = (uint8_t)((info & 1) | ((side || !lattice) ? 0 : EV_SRC) | dflag | (rank == 0 ? EV_FIRST : 0) |
                                  gflag | (split && rank == 0 ? EV_NOWRAP : 0));
      if (rank == 0) B.node_site[node] = gs;
      if (split && up == n1 - 1) {  // the last event: the site's top segment, up to tau = 1
        B.evt_new[up] = 1.0;
        B.segspin_new[up] = sp ^ (uint8_t)(info & 1);
        B.evflags[up] = gflag;
      }
      if (side) { B.l1[k] = node; B.u1[k] = up; } else { B.l0[k] = node; B.u0[k] = up; }
      lo_leg[side] = node;
      up_leg[side] = up;
    }
    if (site) continue;
    const int g = (info >> 1) & 3;
    const int l0 = lo_leg[0], u0 = up_leg[0], l1 = lo_leg[1], u1 = up_leg[1];
    if (g & 2) {            // frozen: all four legs (graph_impl.h:282-283)
      uf_unite(B.parent, l0, l1);
      uf_unite(B.parent, u0, u1);
      uf_unite(B.parent, l0, u0);
    } else if (g == 0) {    // horizontal cut (graph_impl.h:284-288)
      uf_unite(B.parent, l0, l1);
      uf_unite(B.parent, u0, u1);
    } else {                // cross (graph_impl.h:289-293)
      uf_unite(B.parent, l0, u1);
      uf_unite(B.parent, l1, u0);
    }
  }
}

// Linking, one thread per site: the site's bond lists (two or four, each in
// time order) merged in (time, bond) order -- the canonical event order of a
// site -- writing the site's segments in node order and the legs of each
// event; the segment spin flips at every off-diagonal event, so no prefix
// counts or searches are needed.  For sites with a few events each: a
// thread's work is serial in its site's events.  The list heads are kept in
// named registers and selected with constant indices (a dynamically indexed
// register array would live in scratch memory).
__device__ __forceinline__ void link_site_events(const DevBatch& B, int gs) {
  const ChainDev& P = B.chains[B.site_chain[gs]];
  const int s = gs - P.site_base;
  int bl[4];
  const int nbd = site_bonds(P, s, bl);
  int p0, p1, p2 = 0, p3 = 0, e0, e1, e2 = 0, e3 = 0;
  {
    const int b0 = P.bond_base + bl[0], b1 = P.bond_base + bl[1];
    p0 = B.bo_new[b0]; e0 = B.bo_new[b0 + 1];
    p1 = B.bo_new[b1]; e1 = B.bo_new[b1 + 1];
    if (nbd == 4) {
      const int b2 = P.bond_base + bl[2], b3 = P.bond_base + bl[3];
      p2 = B.bo_new[b2]; e2 = B.bo_new[b2 + 1];
      p3 = B.bo_new[b3]; e3 = B.bo_new[b3 + 1];
    }
  }
  if (p0 == e0 && p1 == e1 && p2 == e2 && p3 == e3) return;  // event-free: written by k_scan
  // side (0 source, 1 target) and direction flags of each list's bond
  uint8_t f0, f1, f2 = 0, f3 = 0;
  {
    int a, b;
    bond_sites(P, bl[0], a, b);
    f0 = (uint8_t)((a == s ? EV_SRC : 0) | (bond_dir(P, bl[0]) ? EV_DIR1 : 0));
    bond_sites(P, bl[1], a, b);
    f1 = (uint8_t)((a == s ? EV_SRC : 0) | (bond_dir(P, bl[1]) ? EV_DIR1 : 0));
    if (nbd == 4) {
      bond_sites(P, bl[2], a, b);
      f2 = (uint8_t)((a == s ? EV_SRC : 0) | (bond_dir(P, bl[2]) ? EV_DIR1 : 0));
      bond_sites(P, bl[3], a, b);
      f3 = (uint8_t)((a == s ? EV_SRC : 0) | (bond_dir(P, bl[3]) ? EV_DIR1 : 0));
    }
  }
  double t0 = p0 < e0 ? B.ot_new[p0] : 3.0, t1 = p1 < e1 ? B.ot_new[p1] : 3.0;
  double t2 = p2 < e2 ? B.ot_new[p2] : 3.0, t3 = p3 < e3 ? B.ot_new[p3] : 3.0;
  const int n0 = B.soff_new[gs], n1 = B.soff_new[gs + 1];
  const uint8_t gflag = gauge(P, s) < 0 ? EV_GAUGE : 0;
  const bool split = P.nv > 1;     // virtual site: m + 1 segments, no wrap
  const int ne = split ? n1 - 1 : n1;
  uint8_t spin = B.spins_old[P.old_site_base + s];
  for (int node = n0; node < ne; ++node) {
    // the head with the least (time, bond)
    int i = 0;
    double t = t0;
    int bi = bl[0];
    if (t1 < t || (t1 == t && bl[1] < bi)) { i = 1; t = t1; bi = bl[1]; }
    if (nbd == 4) {
      if (t2 < t || (t2 == t && bl[2] < bi)) { i = 2; t = t2; bi = bl[2]; }
      if (t3 < t || (t3 == t && bl[3] < bi)) { i = 3; t = t3; bi = bl[3]; }
    }
    const int k = i == 0 ? p0 : i == 1 ? p1 : i == 2 ? p2 : p3;
    const uint8_t fl = i == 0 ? f0 : i == 1 ? f1 : i == 2 ? f2 : f3;
    const int off = B.oi_new[k] & 1;
    const int up = (node + 1 == n1) ? n0 : node + 1;  // segment above the event (split: never wraps)
    B.evt_new[node] = t;
    B.segspin_new[node] = spin;
    B.evflags[node] = (uint8_t)(off | fl | (node == n0 ? EV_FIRST | (split ? EV_NOWRAP : 0) : 0) | gflag);
    B.parent[node] = node;
    if (node == n0) B.node_site[node] = gs;
    if (fl & EV_SRC) { B.l0[k] = node; B.u0[k] = up; } else { B.l1[k] = node; B.u1[k] = up; }
    spin ^= (uint8_t)off;
    // advance the chosen list
    if (i == 0) { ++p0; t0 = p0 < e0 ? B.ot_new[p0] : 3.0; }
    else if (i == 1) { ++p1; t1 = p1 < e1 ? B.ot_new[p1] : 3.0; }
    else if (i == 2) { ++p2; t2 = p2 < e2 ? B.ot_new[p2] : 3.0; }
    else { ++p3; t3 = p3 < e3 ? B.ot_new[p3] : 3.0; }
  }
  if (split) {  // the top segment, from the last event up to tau = 1
    B.evt_new[ne] = 1.0;
    B.segspin_new[ne] = spin;
    B.evflags[ne] = gflag;
    B.parent[ne] = ne;
  }
}

// one thread per site, grid-stride over XCD-contiguous shares
__global__ void __launch_bounds__(256) k_link_site(DevBatch B) {
  if (overflowed(B)) return;
  int slo, shi;
  xcd_share(B.nsites, slo, shi);
  for (int gs = slo + threadIdx.x; gs < shi; gs += blockDim.x) link_site_events(B, gs);
}

__global__ void __launch_bounds__(256) k_uf_union(DevBatch B) {
  if (overflowed(B)) return;
  const int nops = B.totals[0];
  int klo, khi;
  xcd_share(nops, klo, khi);
  for (int k = klo + threadIdx.x; k < khi; k += blockDim.x) {
    const int info = B.oi_new[k];
    const int u1 = B.u1[k];
    if (u1 < 0) continue;  // united in LDS by the linking kernel
    const int l0 = B.l0[k], l1 = B.l1[k], u0 = B.u0[k];
    const int g = (info >> 1) & 3;
    if (g & 2) {            // frozen: all four legs (graph_impl.h:282-283)
      uf_unite(B.parent, l0, l1);
      uf_unite(B.parent, u0, u1);
      uf_unite(B.parent, l0, u0);
    } else if (g == 0) {    // horizontal cut (graph_impl.h:284-288)
      uf_unite(B.parent, l0, l1);
      uf_unite(B.parent, u0, u1);
    } else {                // cross (graph_impl.h:289-293)
      uf_unite(B.parent, l0, u1);
      uf_unite(B.parent, l1, u0);
    }
  }
}

// tau boundary of the virtual sites (S > 1/2; path_integral.C:589-615, the
// S > 1/2 branch of the symmetrization): per real site, the tau = 0 ends
// (lowest segments) of its nv virtual sites are joined to their tau = 1 ends
// (top segments) by a permutation that maps equal spins onto each other,
// uniform among those that do (looper::partitioned_random_shuffle,
// looper/permutation.h:88-112).  For each spin value u = 0, 1 in turn, the
// bottoms of spin u in virtual-site order are matched to the tops of spin u
// in virtual-site order permuted by Fisher-Yates: for i = n_u - 1 down to 1,
// j = i - (x mod (i + 1)) with x the next word of the stream (word k is word
// k mod 4 of Philox(sweep, STAGE_BOUNDARY, real site, k / 4)), swap
// positions i and j.  (The spins at tau = 1 are those at tau = 0 in some
// order: every update keeps each real site's magnetization periodic.)  For
// nv = 2 with equal spins this is the exchange iff bit 0 of the first word
// is set.  One thread per real site; nv <= 8, so a permutation lives in one
// register as eight 4-bit fields.
template <typename T>
__device__ __forceinline__ T pick8(int i, T a0, T a1, T a2, T a3, T a4, T a5, T a6, T a7) {
  return i < 4 ? (i == 0 ? a0 : i == 1 ? a1 : i == 2 ? a2 : a3) : (i == 4 ? a4 : i == 5 ? a5 : i == 6 ? a6 : a7);
}

__global__ void __launch_bounds__(256) k_boundary(DevBatch B) {
  if (overflowed(B)) return;
  constexpr int NV = MAX_NV;
  const int stride = gridDim.x * blockDim.x;
  for (int gs = blockIdx.x * blockDim.x + threadIdx.x; gs < B.nsites; gs += stride) {
    const int c = B.site_chain[gs];
    const ChainDev& P = B.chains[c];
    const int nv = P.nv;
    if (nv == 1) continue;
    const int v = gs - P.site_base;
    const int r = v / nv;
    if (v != r * nv) continue;  // one thread per real site: its virtual site 0
    int bot[NV], top[NV];
    uint32_t bs = 0, es = 0;    // bit a: spin of virtual site a's bottom / top segment
#pragma unroll
    for (int a = 0; a < NV; ++a) {
      bot[a] = top[a] = 0;
      if (a < nv) {
        bot[a] = B.soff_new[gs + a];
        top[a] = B.soff_new[gs + a + 1] - 1;
        bs |= (uint32_t)B.segspin_new[bot[a]] << a;
        es |= (uint32_t)B.segspin_new[top[a]] << a;
      }
    }
    const uint32_t all = (1u << nv) - 1;
    // words 0..3, and 4..7 when the two shuffles can take more than four
    // (n_0 - 1 + n_1 - 1 <= nv - 1 words in all)
    const U4 x0 = philox(B.sweep[c], STAGE_BOUNDARY, (uint32_t)r, 0u, P.key0, P.key1);
    U4 x1 = {0u, 0u, 0u, 0u};
    if (nv > 5) x1 = philox(B.sweep[c], STAGE_BOUNDARY, (uint32_t)r, 1u, P.key0, P.key1);
    int k = 0;
    uint32_t perm[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int n = __popc(u ? (bs & all) : (~bs & all));
      uint32_t p = 0x76543210u;
#pragma unroll
      for (int i = NV - 1; i >= 1; --i) {
        if (i < n) {
          const uint32_t w = pick8(k, x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w);
          const int j = i - (int)(w % (uint32_t)(i + 1));
          ++k;
          const uint32_t vi = (p >> (4 * i)) & 15u, vj = (p >> (4 * j)) & 15u;
          p &= ~((15u << (4 * i)) | (15u << (4 * j)));
          p |= (vi << (4 * j)) | (vj << (4 * i));
        }
      }
      perm[u] = p;
    }
#pragma unroll
    for (int a = 0; a < NV; ++a) {
      if (a >= nv) continue;
      const uint32_t u = (bs >> a) & 1u;
      const uint32_t cls_b = u ? bs : ~bs, cls_e = u ? es : ~es;
      const int rank = __popc(cls_b & all & ((1u << a) - 1));
      const int t = (int)(((u ? perm[1] : perm[0]) >> (4 * rank)) & 15u);
      // the top of spin u with rank t among the tops of spin u
      int target = 0, cnt = 0;
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        if (q < nv && ((cls_e >> q) & 1u)) {
          if (cnt == t) target = q;
          ++cnt;
        }
      }
      uf_unite(B.parent, bot[a], pick8(target, top[0], top[1], top[2], top[3], top[4], top[5], top[6], top[7]));
    }
  }
}

// Compress mode: flatten every tree once all unions are done, clear the
// roots' accumulators and draw each cluster's flip bit at its root, keyed by
// the chain-local index of the root (the cluster's minimum segment).
__global__ void __launch_bounds__(256) k_uf_compress(DevBatch B) {
  if (overflowed(B)) return;
  const int M = B.totals[1];
  int ilo, ihi;
  xcd_share(M, ilo, ihi);
  for (int i = ilo + threadIdx.x; i < ihi; i += blockDim.x) {
    int p = B.parent[i];
    if (p != i) {
      int r = p, q;
      while ((q = B.parent[r]) != r) r = q;
      B.parent[i] = r;
    } else {
      B.acc[i] = ClusterAcc{{0, 0, 0, 0}, {0, 0, 0, 0}, 0, 0};
      const int c = node_chain(B, i);
      const ChainDev& P = B.chains[c];
      // with a field the flip waits for the cluster's magnetization (k_field_flip)
      B.fbit[i] = P.h != 0.0 ? (uint8_t)0
                             : (uint8_t)(philox(B.sweep[c], STAGE_FLIP, (uint32_t)(i - B.chain_node0[c]), 0u, P.key0, P.key1).x & 1u);
    }
  }
}


// ---------------------------------------------------------------------------
// (e) improved estimators + (d) flip bits, per segment node.
// WALK (no compress pass): each node walks to its root (the cluster's minimum
// segment) and, when the walk took more than one step, points itself at the
// root, so later kernels find roots in one load.  The cluster's flip bit is
// drawn from the Philox counter of the root's chain-local index by every
// segment of the cluster (the counter k_uf_compress uses), so no kernel has
// to publish it first; with a field the flip waits for the cluster's
// magnetization (k_field_flip) and the bit is 0 here.  Accumulators then
// start at zero because k_flip_collect clears each root's after reading it
// and fresh allocations are cleared.  This saves a pass over the nodes; the
// clearing costs a store per root, so batches of many small clusters keep
// the compress pass.
// Each wave publishes the flip bits and root bits of its 64 consecutive
// segments as one word each (fmask, rmask) for the collector and the
// operator flips.
// A segment adds its length to one class slot of its root (ClusterAcc), its
// tau = 0 boundary (first segment of a site) to one class counter and its
// winding contributions to one packed 64-bit counter.  Each block handles
// chunks of 256 * chunk_items(M) consecutive nodes and merges the segments
// that share a root in an LDS hash table (open addressing, HT slots) before touching global
// memory: the large loops of a cold antiferromagnet (a few clusters holding
// most segments) then cost one global update per chunk instead of one per
// segment.  Segments whose root finds no slot update global memory directly.
// A block of 1024 threads takes a chunk (at most 4 items per thread, 2 in
// flight): every chunk of a large configuration is resident at once, so the
// kernel lasts about one block's serial chain of groups, which more threads
// per chunk shorten; with 16 waves per block only two blocks fit a CU, which
// leaves LDS for a 1024-slot table (53 KB).
// ---------------------------------------------------------------------------
constexpr int EST_THREADS = 1024;
constexpr int EST_U = 2;         // items per thread and group

// Items per thread of the per-node chunked kernels (estimate, collector): 16,
// or 8 or 4 when 16 would give fewer than 1024 chunks.  A small configuration
// otherwise occupies a few blocks, each walking long serial item lists; a
// large one keeps the big chunks that share LDS merges across more segments.
// A function of the segment count only, so every kernel of a sweep agrees.
__device__ __forceinline__ int chunk_items(int M) {
  int it = 16;
  while (it > 4 && (long long)M < 1024LL * 256 * it) it >>= 1;
  return it;
}
constexpr int HT_BITS = 10;
constexpr int HT = 1 << HT_BITS;
constexpr int HT_PROBES = 8;

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Add one contribution (class lengths, tau = 0 boundary counts packed 8 bits
// per class, packed winding) for root r to the block's LDS table, or to the
// root's global accumulator when its probe sequence is full.  The key is
// inserted with a CAS only when a plain read does not already find it, so a
// hot key costs one broadcast read.
__device__ __forceinline__ void ht_add(int* hkey, double (*hL)[HT], unsigned int (*hn)[HT],
                                       unsigned long long* hw, const DevBatch& B, int r,
                                       const double* sL, int n4, long long w) {
  int h = (int)(((uint32_t)r * 2654435761u) >> (32 - HT_BITS));
  int slot = -1;
  for (int p = 0; p < HT_PROBES; ++p) {
    int k = __hip_atomic_load(&hkey[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == -1) k = atomicCAS(&hkey[h], -1, r);
    if (k == -1 || k == r) { slot = h; break; }
    h = (h + 1) & (HT - 1);
  }
  if (slot >= 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (sL[q] != 0.0) atomicAdd(&hL[q][slot], sL[q]);
    if (n4) {
      const unsigned int c0 = n4 & 0xff, c1 = (n4 >> 8) & 0xff, c2 = (n4 >> 16) & 0xff, c3 = (n4 >> 24) & 0xff;
      if (c0 | c1) atomicAdd(&hn[0][slot], c0 | (c1 << 16));
      if (c2 | c3) atomicAdd(&hn[1][slot], c2 | (c3 << 16));
    }
    if (w) atomicAdd(&hw[slot], (unsigned long long)w);
  } else {
    ClusterAcc* a = B.acc + r;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (sL[q] != 0.0) unsafeAtomicAdd(&a->L[q], sL[q]);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if ((n4 >> (8 * q)) & 0xff) atomicAdd(&a->n[q], (n4 >> (8 * q)) & 0xff);
    if (w) atomicAdd((unsigned long long*)&a->w, (unsigned long long)w);
  }
}

template <bool WALK>
// (at most 64 VGPRs: two blocks of 16 waves per CU)
__global__ void __launch_bounds__(EST_THREADS) __attribute__((amdgpu_waves_per_eu(8))) k_estimate(DevBatch B, int flush_items) {
  if (overflowed(B)) return;
  __shared__ int hkey[HT];
  __shared__ double hL[4][HT];
  __shared__ unsigned long long hw[HT];
  __shared__ unsigned int hn[2][HT];   // class counts: [0] = n0 | n1 << 16, [1] = n2 | n3 << 16
  const int tid = threadIdx.x;
  for (int s = tid; s < HT; s += EST_THREADS) {
    hkey[s] = -1;
    hL[0][s] = hL[1][s] = hL[2][s] = hL[3][s] = 0;
    hw[s] = 0;
    hn[0][s] = hn[1][s] = 0;
  }
  __syncthreads();
  const int M = B.totals[1];
  // the chunk of 256 * chunk_items(M) segments over the block's threads
  const int items = chunk_items(M) * 256 / EST_THREADS, echunk = items * EST_THREADS;
  const int flush_every = max(1, flush_items * 256 / EST_THREADS);  // items per thread
  const int nchunks = (M + echunk - 1) / echunk;
  int clo, chi;
  xcd_share(nchunks, clo, chi);
  for (int chunk = clo; chunk < chi; ++chunk) {
    const int cb = chain_of_node(B, min(chunk * echunk, M - 1));  // the chunk's first chain
    // the thread's items in groups of EST_U: every memory step of a group
    // (parents, the walks to the roots, the segments' own data, the events
    // below them) is issued for all its items before any is waited for
    for (int g0 = 0; g0 < items; g0 += EST_U) {
      const int ib = chunk * echunk + g0 * EST_THREADS + tid;
      if (__builtin_amdgcn_readfirstlane(ib) < M) {  // lane 0 holds the wave's lowest node
        // items past the chunk or the end read a valid node and are dropped
        // before any store, so that every load is issued without a branch
        bool ok[EST_U];
        int nd[EST_U], r[EST_U];
        uint8_t fl[EST_U], sp[EST_U], flo[EST_U];
        double thi[EST_U], tlo[EST_U];
#pragma unroll
        for (int u = 0; u < EST_U; ++u) {
          const int i = ib + u * EST_THREADS;
          ok[u] = g0 + u < items && i < M;
          nd[u] = ok[u] ? i : min(ib, M - 1);
        }
#pragma unroll
        for (int u = 0; u < EST_U; ++u) {
          r[u] = B.parent[nd[u]];
          fl[u] = B.evflags[nd[u]];
          sp[u] = B.segspin_new[nd[u]];
          thi[u] = B.evt_new[nd[u]];
          // the segment below: the previous node, except for a site's first
          // segment (replaced below)
          const int pv = max(nd[u] - 1, 0);
          tlo[u] = B.evt_new[pv];
          flo[u] = B.evflags[pv];
        }
        uint32_t f[EST_U];
        if (WALK) {
          // walks to the roots, one hop of every item at a time
          int p0[EST_U];
#pragma unroll
          for (int u = 0; u < EST_U; ++u) p0[u] = r[u];
          for (bool more = true; more;) {
            int q[EST_U];
#pragma unroll
            for (int u = 0; u < EST_U; ++u) q[u] = B.parent[r[u]];
            more = false;
#pragma unroll
            for (int u = 0; u < EST_U; ++u)
              if (q[u] != r[u]) { r[u] = q[u]; more = true; }
          }
#pragma unroll
          for (int u = 0; u < EST_U; ++u) {
            const int c = chain_from(B, cb, nd[u]);
            const ChainDev& P = B.chains[c];
            f[u] = P.h != 0.0 ? 0u
                 : (philox(B.sweep[c], STAGE_FLIP, (uint32_t)(r[u] - B.chain_node0[c]), 0u, P.key0, P.key1).x & 1u);
          }
#pragma unroll
          for (int u = 0; u < EST_U; ++u) {
            const int i = ib + u * EST_THREADS;
            if (ok[u] && r[u] != p0[u]) B.parent[i] = r[u];
          }
        } else {
#pragma unroll
          for (int u = 0; u < EST_U; ++u) f[u] = B.fbit[r[u]];  // drawn for the root by k_uf_compress
        }
        // flip and root bits of the wave's 64 consecutive nodes (the first a
        // multiple of 64), one word each, for the collector
#pragma unroll
        for (int u = 0; u < EST_U; ++u) {
          const unsigned long long fm = __ballot(ok[u] && f[u] != 0);
          const unsigned long long rm = __ballot(ok[u] && r[u] == nd[u]);
          const int w0 = __builtin_amdgcn_readfirstlane(ib) + u * EST_THREADS;
          if (lane_id() == 0 && g0 + u < items && w0 < M) {
            B.fmask[w0 >> 6] = fm;
            B.rmask[w0 >> 6] = rm;
          }
        }
        // a site's first segment wraps through tau = 1: the segment below it is
        // the site's last
        // (one wave-uniform branch; the loads in it are issued for every lane,
        // at node 0 / site 0 for lanes without a first segment)
        int gs[EST_U] = {};
        bool anyfirst = false;
#pragma unroll
        for (int u = 0; u < EST_U; ++u) anyfirst |= (fl[u] & EV_FIRST) != 0;
        if (__ballot(anyfirst)) {
#pragma unroll
          for (int u = 0; u < EST_U; ++u) {
            const int g = first_node_site(B, (fl[u] & EV_FIRST) ? nd[u] : 0);
            gs[u] = (fl[u] & EV_FIRST) ? g : 0;
          }
          int pv[EST_U];
#pragma unroll
          for (int u = 0; u < EST_U; ++u) pv[u] = B.soff_new[gs[u] + 1] - 1;
#pragma unroll
          for (int u = 0; u < EST_U; ++u) {
            const double t = B.evt_new[pv[u]];
            const uint8_t fo = B.evflags[pv[u]];
            if ((fl[u] & (EV_FIRST | EV_EMPTY)) == EV_FIRST) {
              // wrapping through tau = 1: the segment below is the site's last;
              // a virtual site's lowest segment starts at tau = 0, no event below
              const bool nowrap = (fl[u] & EV_NOWRAP) != 0;
              tlo[u] = nowrap ? 0.0 : t;
              flo[u] = nowrap ? (uint8_t)0 : fo;
            }
          }
        }
#pragma unroll
        for (int u = 0; u < EST_U; ++u) {
          const int i = ib + u * EST_THREADS;
          if (!ok[u]) continue;
          const int spin = sp[u];
          const int sig = 1 - 2 * spin;
          const bool first = (fl[u] & EV_FIRST) != 0;
          int dw0 = 0, dw1 = 0;
          double len = 1.0;
          if (!(fl[u] & EV_EMPTY)) {
            len = (first && !(fl[u] & EV_NOWRAP)) ? ((thi[u] + 1.0) - tlo[u]) : (thi[u] - tlo[u]);
            // this segment is the lower leg of the event above it and the upper
            // leg of the event below it (stiffness.h:94-105, source legs only)
            if (fl[u] & EV_SRC) { if (fl[u] & EV_DIR1) dw1 += sig; else dw0 += sig; }
            if (flo[u] & EV_SRC) { if (flo[u] & EV_DIR1) dw1 -= sig; else dw0 -= sig; }
          }
          if (first) B.spins_new[gs[u]] = (uint8_t)(spin ^ f[u]);
          B.segspin_new[i] = (uint8_t)(spin ^ f[u]);
          const int cls = (spin << 1) | ((fl[u] & EV_GAUGE) ? 1 : 0);  // seg_class
          double sL[4] = {0, 0, 0, 0};
          sL[cls] = len;
          ht_add(hkey, hL, hn, hw, B, r[u], sL, first ? (1 << (8 * cls)) : 0, pack_winding(dw0, dw1));
        }
      }
      // the table goes to global memory every 256 * flush_items segments (at
      // the end of a group) and at the chunk's end; a block-uniform
      // condition, so every wave takes the barriers
      if ((g0 + EST_U) % flush_every != 0 && g0 + EST_U < items) continue;
      __syncthreads();
      for (int s = tid; s < HT; s += EST_THREADS) {
        const int r = hkey[s];
        if (r < 0) continue;
        ClusterAcc* a = B.acc + r;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (hL[k][s] != 0.0) { unsafeAtomicAdd(&a->L[k], hL[k][s]); hL[k][s] = 0; }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const unsigned int v = hn[q][s];
          if (v) {
            if (v & 0xffffu) atomicAdd(&a->n[2 * q], (int)(v & 0xffffu));
            if (v >> 16) atomicAdd(&a->n[2 * q + 1], (int)(v >> 16));
            hn[q][s] = 0;
          }
        }
        if (hw[s]) { atomicAdd((unsigned long long*)&a->w, hw[s]); hw[s] = 0; }
        hkey[s] = -1;
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Longitudinal field (the FIELD branch of path_integral.C:798): the cluster
// flips iff 2u - 1 < tanh(beta w_c), w_c = -h umag_c, with u from the same
// Philox counter as the field-free flip; det_tanh keeps host and device
// decisions identical.  Then every segment of a field chain takes its
// root's bit: flipped segment and tau = 0 spins (k_estimate left them
// unflipped: it gives field clusters bit 0).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_field_flip(DevBatch B) {
  if (overflowed(B)) return;
  const int M = B.totals[1];
  int ilo, ihi;
  xcd_share(M, ilo, ihi);
  for (int i = ilo + threadIdx.x; i < ihi; i += blockDim.x) {
    if (B.parent[i] != i) continue;
    const int c = node_chain(B, i);
    const ChainDev& P = B.chains[c];
    if (P.h == 0.0) continue;
    const ClusterAcc a = B.acc[i];
    const double umag = 0.5 * ((a.L[0] + a.L[1]) - (a.L[2] + a.L[3]));
    const double x = P.beta * (-P.h * umag);
    const double u = uniform53(philox(B.sweep[c], STAGE_FLIP, (uint32_t)(i - B.chain_node0[c]), 0u, P.key0, P.key1));
    B.fbit[i] = (uint8_t)((2.0 * u - 1.0) < det_tanh(x) ? 1 : 0);
  }
}

__global__ void __launch_bounds__(256) k_field_apply(DevBatch B) {
  if (overflowed(B)) return;
  const int M = B.totals[1];
  int ilo, ihi;
  xcd_share(M, ilo, ihi);
  for (int i = ilo + threadIdx.x; i < ihi; i += blockDim.x) {
    const int c = node_chain(B, i);
    if (B.chains[c].h == 0.0) continue;
    const uint8_t f = B.fbit[B.parent[i]];
    if (!f) continue;
    atomicOr(&B.fmask[i >> 6], 1ull << (i & 63));
    B.segspin_new[i] ^= 1;
    if (B.evflags[i] & EV_FIRST) B.spins_new[first_node_site(B, i)] ^= 1;
  }
}

// ---------------------------------------------------------------------------
// (d) operator flip, then the improved-estimator collector
// (susceptibility.h:177-198, stiffness.h:117-121) over the roots.
// Collection is chunked: chunk j covers nodes [j*C, (j+1)*C), C = 256 *
// chunk_items(M), and belongs to the chain of its first node; its sum over the roots of that
// chain is written to col_part[j] (no atomics, fixed order), roots of other
// chains in the chunk (chunks straddling a chain boundary) are added to
// rec_acc atomically.  end_sweep adds the partials of each chain in order.
// An operator flips iff the clusters of its source-side legs l0 and u0 flip
// differently; both bits come from the flip bitmap k_estimate wrote (a
// 64-times smaller array than a byte per node, so the gathers stay in L2),
// and the roots come from its root bitmap instead of a parent load per node.
// ---------------------------------------------------------------------------
constexpr int COL_THREADS = 256;
constexpr int COL_MIN_CHUNK = 256 * 4;  // smallest collector chunk (256 * chunk_items)
constexpr int COL_U = 4;                         // items per lane in flight (divides chunk_items)
constexpr int NCOL = NUM_COLLECTOR_TERMS;

// Collector terms of one cluster.  h, flip: the chain's field and the
// cluster's flip; the field weight of the cluster is w = -h umag
// (cluster.h:50-63) and its Zeeman energy after the flip flip ? -w : w
// (path_integral.C:852-858).
__device__ __forceinline__ void cluster_terms(const ClusterAcc& a, double h, int flip, double* v) {
  const double L0 = a.L[0], L1 = a.L[1], L2 = a.L[2], L3 = a.L[3];
  const int n0 = a.n[0], n1 = a.n[1], n2 = a.n[2], n3 = a.n[3];
  int w0, w1;
  unpack_winding(a.w, w0, w1);
  // per-cluster improved estimates (susceptibility.h:96-147 with sigma = 1-2c)
  const double usize0 = 0.5 * (n0 + n1 + n2 + n3), umag0 = 0.5 * (n0 + n1 - n2 - n3);
  const double ssize0 = 0.5 * (n0 - n1 + n2 - n3), smag0 = 0.5 * (n0 - n1 - n2 + n3);
  const double usize = 0.5 * ((L0 + L1) + (L2 + L3)), umag = 0.5 * ((L0 + L1) - (L2 + L3));
  const double ssize = 0.5 * ((L0 - L1) + (L2 - L3)), smag = 0.5 * ((L0 - L1) - (L2 - L3));
  v[0] = 1.0;  // clusters
  v[1] = umag0;
  v[2] = usize0 * usize0;
  v[3] = umag0 * umag0;
  v[4] = v[2] * v[2];
  v[5] = v[3] * v[3];
  v[6] = usize * usize;
  v[7] = umag * umag;
  v[8] = smag0;
  v[9] = ssize0 * ssize0;
  v[10] = smag0 * smag0;
  v[11] = v[9] * v[9];
  v[12] = v[10] * v[10];
  v[13] = ssize * ssize;
  v[14] = smag * smag;
  v[15] = 0.25 * ((double)w0 * w0 + (double)w1 * w1);
  v[16] = umag;   // normal estimator: tau-integrated magnetization
  v[17] = smag;
  const double wf = -h * umag;
  v[18] = flip ? -wf : wf;
}

// record slot of collector term q (R_NC, R_UMAG0 .. R_W2, R_UMAGA, R_SMAGA, R_EFIELD)
__device__ __forceinline__ int term_slot(int q) {
  return q == 0 ? R_NC : (q <= 15 ? R_UMAG0 + (q - 1) : R_UMAGA + (q - 16));
}

// clear: zero each root's accumulator after reading it (no compress pass)
__global__ void __launch_bounds__(COL_THREADS) k_flip_collect(DevBatch B, int clear) {
  if (overflowed(B)) return;
  const int nops = B.totals[0];
  const int M = B.totals[1];
  int klo, khi;
  xcd_share(nops, klo, khi);
  // COL_U operators per lane at a time, every load issued before any store
  // (the last operator of the share stands in for those past its end)
  for (int k0 = klo + threadIdx.x; k0 < khi; k0 += COL_U * COL_THREADS) {
    int kk[COL_U], a[COL_U], b[COL_U], t[COL_U];
    unsigned long long fa[COL_U], fb[COL_U];
#pragma unroll
    for (int u = 0; u < COL_U; ++u) kk[u] = min(k0 + u * COL_THREADS, khi - 1);
#pragma unroll
    for (int u = 0; u < COL_U; ++u) { a[u] = B.l0[kk[u]]; b[u] = B.u0[kk[u]]; t[u] = B.oi_new[kk[u]]; }
#pragma unroll
    for (int u = 0; u < COL_U; ++u) { fa[u] = B.fmask[a[u] >> 6]; fb[u] = B.fmask[b[u] >> 6]; }
#pragma unroll
    for (int u = 0; u < COL_U; ++u)
      if (k0 + u * COL_THREADS < khi && (((fa[u] >> (a[u] & 63)) ^ (fb[u] >> (b[u] & 63))) & 1))
        B.oi_new[kk[u]] = t[u] ^ 1;
  }
  // per-chunk sums by chain: slot k holds chain c_first + k (chains are
  // contiguous in node order, so a chunk meets few of them)
  constexpr int CSLOTS = 8;
  __shared__ double slot[CSLOTS][NCOL];
  // chunks of 256 * chunk_items(M) nodes over the block's threads
  const int cchunk = chunk_items(M) * 256, citems = cchunk / COL_THREADS;
  const int nchunks = (M + cchunk - 1) / cchunk;
  int clo, chi;
  xcd_share(nchunks, clo, chi);
  for (int chunk = clo; chunk < chi; ++chunk) {
    const int base = chunk * cchunk;
    const int c0 = node_chain(B, base);
    const int c0_end = (c0 + 1 < B.nchains) ? B.chain_node0[c0 + 1] : M;  // nodes of chain c0 end here
    for (int s = threadIdx.x; s < CSLOTS * NCOL; s += COL_THREADS) slot[s / NCOL][s % NCOL] = 0;
    __syncthreads();
    // each lane sums the terms of its roots while they belong to one chain
    // (cl) and hands the sum on when the chain changes and at the chunk's end
    double w[NCOL];
#pragma unroll
    for (int q = 0; q < NCOL; ++q) w[q] = 0;
    int cl = -1;
    auto flush = [&](int c) {
      const int k = c - c0;
      if (k < CSLOTS) {
        for (int q = 0; q < NCOL; ++q) atomicAdd(&slot[k][q], w[q]);
      } else {
        double* rec = B.rec_acc + (size_t)c * R_COUNT;
        for (int q = 0; q < NCOL; ++q) unsafeAtomicAdd(rec + term_slot(q), w[q]);
      }
    };
    // COL_U nodes per lane at a time: parents, then the accumulators and
    // flip bits of the roots among them (a non-root reads the chunk's first
    // node's, a cached line, and drops it), then the sums
    for (int it0 = 0; it0 < citems; it0 += COL_U) {
      const int ib = base + it0 * COL_THREADS + threadIdx.x;
      if (__builtin_amdgcn_readfirstlane(ib) >= M) break;  // lane 0 holds the wave's lowest node
      int nd[COL_U];
      unsigned long long rw[COL_U];
      bool ok[COL_U], root[COL_U];
#pragma unroll
      for (int u = 0; u < COL_U; ++u) {
        ok[u] = it0 + u < citems && ib + u * COL_THREADS < M;  // in the chunk and the batch
        nd[u] = ok[u] ? ib + u * COL_THREADS : min(ib, M - 1);
      }
#pragma unroll
      for (int u = 0; u < COL_U; ++u) rw[u] = B.rmask[nd[u] >> 6];
#pragma unroll
      for (int u = 0; u < COL_U; ++u) root[u] = ok[u] && ((rw[u] >> (nd[u] & 63)) & 1);
      ClusterAcc ac[COL_U];
      uint8_t fr[COL_U];
#pragma unroll
      for (int u = 0; u < COL_U; ++u) {
        const int j = root[u] ? nd[u] : base;
        ac[u] = B.acc[j];
        fr[u] = (uint8_t)((B.fmask[j >> 6] >> (j & 63)) & 1);
      }
#pragma unroll
      for (int u = 0; u < COL_U; ++u) {
        if (!root[u]) continue;
        const int i = nd[u];
        // chains are contiguous in node order: from the chunk's first chain
        // forward over the few cached chain starts
        const int c = (i < c0_end) ? c0 : chain_from(B, c0, i);
        if (c != cl) {
          if (cl >= 0) flush(cl);
#pragma unroll
          for (int q = 0; q < NCOL; ++q) w[q] = 0;
          cl = c;
        }
        double v[NCOL];
        cluster_terms(ac[u], B.chains[c].h, fr[u], v);
        if (clear) B.acc[i] = ClusterAcc{{0, 0, 0, 0}, {0, 0, 0, 0}, 0, 0};
#pragma unroll
        for (int q = 0; q < NCOL; ++q) w[q] += v[q];
      }
    }
    int cmax = cl, cmin = (cl < 0) ? 0x7fffffff : cl;
    for (int d = 32; d > 0; d >>= 1) {
      cmax = max(cmax, __shfl_xor(cmax, d, 64));
      cmin = min(cmin, __shfl_xor(cmin, d, 64));
    }
    if (cmax >= 0) {  // some lane of the wave holds a sum
      if (cmin == cmax) {
#pragma unroll
        for (int q = 0; q < NCOL; ++q) w[q] = wave_sum_d(w[q]);
        if (lane_id() == 0) flush(cmax);
      } else if (cl >= 0) {
        flush(cl);
      }
    }
    __syncthreads();
    if (threadIdx.x < NCOL) B.col_part[(size_t)chunk * NCOL + threadIdx.x] = slot[0][threadIdx.x];
    for (int s = NCOL + threadIdx.x; s < CSLOTS * NCOL; s += COL_THREADS) {
      const int k = s / NCOL, q = s % NCOL;
      const double val = slot[k][q];
      if (val != 0.0 && c0 + k < B.nchains)
        unsafeAtomicAdd(B.rec_acc + (size_t)(c0 + k) * R_COUNT + term_slot(q), val);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// end of sweep, one block per chain: add the chain's collector partials in a
// fixed order, finish and store the record, advance the sweep counter.  The
// last block to finish advances the record ring.
// ---------------------------------------------------------------------------
constexpr int END_THREADS = 256;

__global__ void __launch_bounds__(END_THREADS) k_end_sweep(DevBatch B) {
  const int slot = *B.ring_pos;
  const int c = blockIdx.x;
  const int tid = threadIdx.x;
  __shared__ double red[END_THREADS / 64][NCOL];
  __shared__ int is_last;
  // The configuration just built lives at the new bases: the next sweep reads
  // it there.  No kernel after this one in the sweep reads the old bases, and
  // this kernel reads neither.
  if (tid == 0) {
    ChainDev* W = B.chains + c;
    W->old_site_base = W->site_base;
    W->old_bond_base = W->bond_base;
  }
  if (slot >= B.ring_size) {
    if (tid == 0) atomicOr(B.flags, FLAG_RING_OVERFLOW);
    return;
  }
  const bool bad = overflowed(B);
  const ChainDev& P = B.chains[c];
  double v[NCOL];
#pragma unroll
  for (int q = 0; q < NCOL; ++q) v[q] = 0;
  if (!bad) {
    const int start = B.soff_new[P.site_base], end = B.soff_new[P.site_base + P.N];
    const int cchunk = chunk_items(B.totals[1]) * 256;  // as in k_flip_collect
    const int j0 = (start + cchunk - 1) / cchunk, j1 = (end + cchunk - 1) / cchunk;
    for (int j = j0 + tid; j < j1; j += END_THREADS)
      for (int q = 0; q < NCOL; ++q) v[q] += B.col_part[(size_t)j * NCOL + q];
  }
#pragma unroll
  for (int q = 0; q < NCOL; ++q) v[q] = wave_sum_d(v[q]);
  if (lane_id() == 0)
    for (int q = 0; q < NCOL; ++q) red[tid >> 6][q] = v[q];
  __syncthreads();
  double* rec = B.rec_acc + (size_t)c * R_COUNT;
  double* out = B.rec_ring + ((size_t)slot * B.nchains + c) * R_COUNT;
  if (tid < R_COUNT) {
    double t = rec[tid];
    rec[tid] = 0;
    for (int q = 0; q < NCOL; ++q)
      if (term_slot(q) == tid)
        for (int w = 0; w < END_THREADS / 64; ++w) t += red[w][q];
    out[tid] = t;
  }
  __syncthreads();
  // sign problem: the off-diagonal operators of negative weight before the
  // flip (per bond: all operators less the diagonal ones the diagonal update
  // placed), path_integral.C:551-576
  __shared__ int negw[END_THREADS / 64];
  int neg = 0;
  if (P.sgn_lattice | P.sgn_site) {
    for (int b = tid; b < P.Nb; b += END_THREADS) {
      const bool negative = (b < P.nb_lat) ? P.sgn_lattice != 0 : (b >= P.nb_real && P.sgn_site != 0);
      if (negative) neg += B.counts[P.bond_base + b] - B.acc_count[P.bond_base + b];
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) neg += __shfl_xor(neg, d, 64);
    if (lane_id() == 0) negw[tid >> 6] = neg;
    __syncthreads();
    neg = 0;
    for (int w = 0; w < END_THREADS / 64; ++w) neg += negw[w];
  }
  if (tid == 0) {
    const double nop = bad ? 0.0 : (double)(B.bo_new[P.bond_base + P.Nb] - B.bo_new[P.bond_base]);
    out[R_NOP] = nop;
    out[R_SIGN] = (neg & 1) ? -1.0 : 1.0;
    out[R_ENERGY] = P.offset_total - nop / P.beta + out[R_EFIELD];
    out[R_BETA] = P.beta;
    out[R_SWEEP] = (double)B.sweep[c];
    B.sweep[c] += 1;
    // annealing: the next sweep's inverse temperature and candidate scale
    // (looper/temperature.h:93-113 at the next Monte Carlo step)
    if (P.sched_n > 0) {
      ChainDev* W = B.chains + c;
      const double beta = 1.0 / schedule_temperature(B.sched + 2 * P.sched_off, P.sched_n, (int)B.sweep[c]);
      W->beta = beta;
      W->scale = candidate_scale(beta, P.wsum);
      W->site_scale = candidate_scale(beta, P.site_w);
      W->dscale = candidate_scale(beta, P.dwsum);
    }
    __threadfence();
    is_last = (atomicAdd(B.end_count, 1) == (int)gridDim.x - 1);
  }
  __syncthreads();
  if (is_last && tid == 0) {
    *B.end_count = 0;
    *B.ring_pos = slot + 1;
  }
}

// ---------------------------------------------------------------------------
// replica exchange ("loop; path integral; exchange", parapack's exchange
// worker over loop_worker's hooks set_beta / weight_parameter / log_weight,
// path_integral.C:97-109): after a sweep, the chains holding neighbouring
// temperatures t, t+1 (t even on even steps, odd on odd steps) swap
// temperatures with probability min(1, W(a, b_t+1) W(b, b_t) /
// (W(a, b_t) W(b, b_t+1))), W(c, beta) = beta^(operators of c), i.e. when
// log u < (n_a - n_b) log(beta_t+1 / beta_t).  u is Philox(step, 4, t, 0) on
// the exchange key; log is det_log, so the host reproduces every decision.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_exchange(DevBatch B) {
  if (overflowed(B)) return;
  const uint32_t step = *B.xstep;
  for (int t = (int)(step & 1) + 2 * (int)threadIdx.x; t + 1 < B.nslots; t += 2 * (int)blockDim.x) {
    const int a = B.xslot[t], b = B.xslot[t + 1];
    const ChainDev& Pa = B.chains[a];
    const ChainDev& Pb = B.chains[b];
    const int na = B.bo_new[Pa.bond_base + Pa.Nb] - B.bo_new[Pa.bond_base];
    const int nb = B.bo_new[Pb.bond_base + Pb.Nb] - B.bo_new[Pb.bond_base];
    const double x = (double)(na - nb) * B.xlnr[t];
    const bool accept = (x >= 0.0) || (log_uniform(philox(step, STAGE_EXCHANGE, (uint32_t)t, 0u, B.xkey0, B.xkey1)) < x);
    if (accept) {
      B.xslot[t] = b;
      B.xslot[t + 1] = a;
      B.chains[a].beta = B.xbeta[t + 1];
      B.chains[a].scale = B.xscale[t + 1];
      B.chains[b].beta = B.xbeta[t];
      B.chains[b].scale = B.xscale[t];
      B.xacc[t] += 1;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) *B.xstep = step + 1;
}

// ---------------------------------------------------------------------------
// canonical labels of the leg elements of one chain (debug / parity export):
// element s -> boundary of site s, N + 4k + {0,1,2,3} -> legs l0,l1,u0,u1 of
// operator k (chain-local bond-major order); label = min element index of
// its cluster.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int element_node(const DevBatch& B, const ChainDev& P, int e, int op0) {
  if (e < P.N) return B.soff_new[P.site_base + e];
  int k = (e - P.N) >> 2, leg = (e - P.N) & 3;
  int kg = op0 + k;
  if (leg == 0) return B.l0[kg];
  if (leg == 1) return B.l1[kg];
  if (leg == 2) return B.u0[kg];
  const int u1 = B.u1[kg];
  return u1 < 0 ? ~u1 : u1;  // complemented: united in LDS
}

__global__ void k_canon_min(DevBatch B, int chain, int nelem, int* canon) {
  const ChainDev& P = B.chains[chain];
  const int op0 = B.bo_new[P.bond_base];
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nelem; e += gridDim.x * blockDim.x) {
    int node = element_node(B, P, e, op0);
    atomicMin(canon + B.parent[node], e);
  }
}

__global__ void k_canon_label(DevBatch B, int chain, int nelem, const int* canon, int* labels,
                              int* roots) {
  const ChainDev& P = B.chains[chain];
  const int op0 = B.bo_new[P.bond_base];
  const int node0 = B.soff_new[P.site_base];
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nelem; e += gridDim.x * blockDim.x) {
    int node = element_node(B, P, e, op0);
    labels[e] = canon[B.parent[node]];
    roots[e] = B.parent[node] - node0;
  }
}

// ---------------------------------------------------------------------------
// "loop; sse" (sse.C:174-400): the same loop clusters on the stochastic
// series expansion.  A chain's operator string (series order) is walked by
// one wave, all lanes carrying the same state: before each kept operator,
// and again after each insertion, a gap tries an insertion when
// (n + 1) u < beta W_tot (W_tot = Nb sum_g v_g; bond uniform from word 2,
// graph from word 3 by the thresholds, kept if compatible); a diagonal
// operator is removed when beta W_tot u < n (the next one follows with no gap
// between) or else regraphed by choose_diagonal (graph_impl.h:297-312); an
// off-diagonal one by choose_offdiagonal.  W_tot sums every location's
// graphs: lattice bonds, the single-ion term's in-site bonds and the
// transverse field's site operators (the site graph, always compatible; an
// off-diagonal site operator flips its one spin).  Step k of the walk takes the words
// Philox(sweep, 6, 0, k), computed 64 steps at a time, one per lane, and
// staged in LDS with the spins, the per-bond counters and sites and the
// operators in flight, so that a step waits on no global memory.  The new string then becomes bond-major operators at times p / n, on
// which the path-integral cluster stage runs unchanged (lengths in units of
// the string), and the flipped types return to the string.
// ---------------------------------------------------------------------------
constexpr int SSE_THREADS = 64;

__global__ void __launch_bounds__(SSE_THREADS) k_sse_diag(DevBatch B) {
  if (blockIdx.x == 0 && threadIdx.x == 0) scan_advance_epoch(B);  // for k_scan, which follows
  if (overflowed(B)) return;
  const int c = blockIdx.x;
  // (a copy: the string stores below could alias the descriptor, which
  // would make every step reload its fields)
  const ChainDev P = B.chains[c];
  // The walk's working set in LDS, so that a step waits on no global memory:
  // spins, per-bond counters, each bond's two sites, the next 64 old
  // operators and the last (up to 64) emitted ones, stored 64 at a time
  __shared__ uint8_t spc[SSE_MAX_SITES];
  __shared__ int cnt[SSE_MAX_BONDS];
  __shared__ uint32_t ends[SSE_MAX_BONDS];  // source | target << 16
  __shared__ int ochunk[SSE_THREADS], obuf[SSE_THREADS];
  __shared__ U4 xbuf[SSE_THREADS];          // the Philox words of the next 64 steps
  const int lane = threadIdx.x;
  const int Nb = P.Nb;
  for (int s = lane; s < P.N; s += SSE_THREADS) spc[s] = B.spins_old[P.old_site_base + s];
  for (int b = lane; b < Nb; b += SSE_THREADS) {
    int s0, s1;
    op_sites(P, b, s0, s1);
    cnt[b] = 0;
    ends[b] = (uint32_t)s0 | ((uint32_t)s1 << 16);
  }
  __syncthreads();
  const uint32_t sweep = B.sweep[c];
  const int n_old = B.bo_old[P.old_bond_base + Nb] - B.bo_old[P.old_bond_base];
  const int* so = B.sse_old + P.sse_base;
  int* sn = B.sse_new + P.sse_base;
  const double bw = P.beta * P.sse_wtot;
  int nop = n_old, out = 0, i = 0, step = 0;
  int xbase = -SSE_THREADS, obase = -SSE_THREADS;  // first step / old operator held by the lanes
  bool try_gap = true, full = false;
  // Every lane runs the walk on the same values; reading each value loaded
  // from LDS through readfirstlane tells the compiler so, which keeps the
  // walk's state in scalar registers and its branches scalar.
  auto uni = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
  while (try_gap || i < n_old) {
    if (step - xbase >= SSE_THREADS) {
      xbase = step;
      xbuf[lane] = philox(sweep, STAGE_SSE, 0u, (uint32_t)(step + lane), P.key0, P.key1);
    }
    U4 x = xbuf[step - xbase];  // (one wave: its LDS accesses complete in order)
    x.x = (uint32_t)uni((int)x.x);
    x.y = (uint32_t)uni((int)x.y);
    x.z = (uint32_t)uni((int)x.z);
    x.w = (uint32_t)uni((int)x.w);
    ++step;
    const double u = uniform53(x);
    int ne, b;
    if (try_gap) {
      try_gap = false;
      if (!((double)(nop + 1) * u < bw)) continue;
      // the location's class and graph by word 2, the location by word 3
      // (the thresholds ascend, so g counts those at or below z; no array
      // of the descriptor is indexed at run time, which would put the
      // whole copy in scratch memory)
      const uint64_t z = x.z;
      int g = 0;
      if (z < P.sse_cls[0]) {
        b = (int)(((uint64_t)x.w * (uint64_t)P.nb_lat) >> 32);
        g = (int)(z >= P.sse_tl[0]) + (int)(z >= P.sse_tl[1]) + (int)(z >= P.sse_tl[2]);
      } else if (z < P.sse_cls[1]) {
        b = P.nb_lat + (int)(((uint64_t)x.w * (uint64_t)(P.nb_real - P.nb_lat)) >> 32);
        g = (int)(z >= P.sse_td[0]) + (int)(z >= P.sse_td[1]) + (int)(z >= P.sse_td[2]);
      } else {  // a site operator: the site graph, compatible with either spin
        b = P.nb_real + (int)(((uint64_t)x.w * (uint64_t)P.N) >> 32);
      }
      if (b < P.nb_real) {
        const uint32_t e2 = (uint32_t)uni((int)ends[b]);
        if (((g & 1) ^ uni(spc[e2 & 0xffff]) ^ uni(spc[e2 >> 16])) == 0) continue;
      }
      ne = (b << 3) | (g << 1);
      ++nop;
    } else {
      if (i - obase >= SSE_THREADS) {  // the next 64 old operators
        obase = i;
        ochunk[lane] = (i + lane < n_old) ? so[i + lane] : 0;
      }
      const int e = uni(ochunk[i - obase]);
      ++i;
      b = e >> 3;
      const uint32_t e2 = (uint32_t)uni((int)ends[b]);
      const int s0 = (int)(e2 & 0xffff), s1 = (int)(e2 >> 16);
      const int c0 = uni(spc[s0]), c1 = uni(spc[s1]);
      const bool site = b >= P.nb_real, insite = b >= P.nb_lat && !site;
      if ((e & 1) == 0) {
        if (bw * u < (double)nop) {  // removed
          --nop;
          continue;
        }
        const int cc = 1 ^ c0 ^ c1;  // 0 antiparallel, 1 parallel
        const uint64_t pd = insite ? (cc ? P.dpdiag[1] : P.dpdiag[0]) : (cc ? P.pdiag[1] : P.pdiag[0]);
        ne = (b << 3) | (site ? 0 : (((((uint64_t)x.z < pd) ? 0 : 2) ^ cc) << 1));
      } else {
        const uint64_t to = insite ? P.dtoff : P.toff;
        ne = (b << 3) | (site ? 0 : ((((uint64_t)x.z < to) ? 0 : 1) << 1)) | 1;
        if (lane == 0) {
          spc[s0] = (uint8_t)(c0 ^ 1);
          if (!site) spc[s1] = (uint8_t)(c1 ^ 1);  // a site operator flips its one spin
        }
      }
    }
    if (out >= P.sse_cap) {
      full = true;
      break;
    }
    if (lane == 0) {
      obuf[out & (SSE_THREADS - 1)] = ne;
      atomicAdd(&cnt[b], 1);
    }
    ++out;
    if ((out & (SSE_THREADS - 1)) == 0) sn[out - SSE_THREADS + lane] = obuf[lane];  // 64 emitted
    try_gap = true;
  }
  const int rest = out & (SSE_THREADS - 1);
  if (lane < rest) sn[out - rest + lane] = obuf[lane];
  if (full && lane == 0) atomicOr(B.flags, FLAG_OP_OVERFLOW);
  __syncthreads();
  for (int b = lane; b < Nb; b += SSE_THREADS) B.counts[P.bond_base + b] = cnt[b];
}

// the new string into bond-major order (series order within a bond), times
// p / n; perm[p] = the operator index of string position p
__global__ void __launch_bounds__(SSE_THREADS) k_sse_scatter(DevBatch B) {
  if (overflowed(B)) return;
  const int c = blockIdx.x;
  const ChainDev P = B.chains[c];  // (a copy, as in k_sse_diag)
  __shared__ int cur[SSE_MAX_BONDS], koff[SSE_MAX_BONDS];
  const int lane = threadIdx.x, Nb = P.Nb;
  for (int b = lane; b < Nb; b += SSE_THREADS) {
    cur[b] = B.bo_new[P.bond_base + b];
    koff[b] = 0;
  }
  __syncthreads();
  const int n = B.bo_new[P.bond_base + Nb] - B.bo_new[P.bond_base];
  const int* sn = B.sse_new + P.sse_base;
  int* perm = B.sse_perm + P.sse_base;
  const double dn = (double)n;
  for (int p0 = 0; p0 < n; p0 += SSE_THREADS) {
    // lanes past the end hold -1 (bond -1, which no element has)
    const int el = (p0 + lane < n) ? sn[p0 + lane] : -1;
    const int m = min(SSE_THREADS, n - p0);
    // in series order, each element takes its bond's next slot and
    // off-diagonal count: the elements of the same bond before it in these
    // 64 add to the bond's counts so far, and the last of them carries the
    // counts on to the next 64
    int rank = 0, roff = 0;
    bool last = true;
    for (int j = 0; j < m; ++j) {
      const int e = __shfl(el, j);
      if ((e >> 3) == (el >> 3)) {
        if (j < lane) {
          ++rank;
          roff += e & 1;
        } else if (j > lane) {
          last = false;
        }
      }
    }
    if (p0 + lane < n) {
      const int b = el >> 3;
      const int pos = cur[b] + rank, kc = koff[b] + roff;
      if (last) {  // (one wave: these stores follow every lane's loads above)
        cur[b] = pos + 1;
        koff[b] = kc + (el & 1);
      }
      B.ot_new[pos] = (double)(p0 + lane) / dn;
      B.oi_new[pos] = ((P.bond_base + b) << 3) | (el & 7);
      B.kc_new[pos] = kc;
      perm[p0 + lane] = pos;
    }
  }
  // each bond's diagonal operators, as the path-integral diagonal update
  // leaves them (k_end_sweep counts the off-diagonal ones for the sign)
  __syncthreads();
  for (int b = lane; b < Nb; b += SSE_THREADS)
    B.acc_count[P.bond_base + b] = (cur[b] - B.bo_new[P.bond_base + b]) - koff[b];
}

// the flipped operator types back into the string
__global__ void __launch_bounds__(256) k_sse_gather(DevBatch B) {
  if (overflowed(B)) return;
  const int c = blockIdx.x;
  const ChainDev& P = B.chains[c];
  // (k_end_sweep has moved the chain's old bases to the configuration just built)
  const int o0 = B.bo_new[P.bond_base], n = B.bo_new[P.bond_base + P.Nb] - o0;
  int* sn = B.sse_new + P.sse_base;
  const int* perm = B.sse_perm + P.sse_base;
  for (int p = threadIdx.x; p < n; p += blockDim.x) sn[p] = (sn[p] & ~1) | (B.oi_new[perm[p]] & 1);
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
template <int G>
static void launch_diag(const DevBatch& B, const LaunchCfg& L, hipStream_t st, bool fill) {
  int groups_per_block = 256 / G;
  int blocks = (B.nbonds + groups_per_block - 1) / groups_per_block;
  if (blocks > L.max_blocks_bond) blocks = L.max_blocks_bond;
  if (blocks < 1) blocks = 1;
  if (fill) {
    const size_t lds = G == 64 ? (size_t)(256 / 64) * 2 * L.fill_cap * sizeof(double) : 0;
    hipLaunchKernelGGL(k_diag_fill<G>, dim3(blocks), dim3(256), lds, st, B, L.fill_cap);
  }
  else {
    hipLaunchKernelGGL(k_diag_count<G>, dim3(blocks), dim3(256), 0, st, B);
    if (B.any_extra) {
      const int sblocks = max(1, min(L.max_blocks_bond, (B.nsites + groups_per_block - 1) / groups_per_block));
      if (B.any_extra & 1) hipLaunchKernelGGL((k_diag_count_extra<G, BOND_INSITE>), dim3(sblocks), dim3(256), 0, st, B);
      if (B.any_extra & 2) hipLaunchKernelGGL((k_diag_count_extra<G, BOND_SITE>), dim3(sblocks), dim3(256), 0, st, B);
    }
  }
}

static void diag_dispatch(const DevBatch& B, const LaunchCfg& L, hipStream_t st, bool fill) {
  switch (L.group) {
    case 1: launch_diag<1>(B, L, st, fill); break;
    case 2: launch_diag<2>(B, L, st, fill); break;
    case 4: launch_diag<4>(B, L, st, fill); break;
    case 8: launch_diag<8>(B, L, st, fill); break;
    case 16: launch_diag<16>(B, L, st, fill); break;
    case 32: launch_diag<32>(B, L, st, fill); break;
    default: launch_diag<64>(B, L, st, fill); break;
  }
}

int col_chunks(int node_cap) { return (node_cap + COL_MIN_CHUNK - 1) / COL_MIN_CHUNK; }

int scan_blocks(const DevBatch& B) {
  return (B.nbonds + SCAN_CHUNK - 1) / SCAN_CHUNK + (B.nsites + SCAN_CHUNK - 1) / SCAN_CHUNK;
}

const char* const kKernelNames[NUM_SWEEP_KERNELS] = {
    "diag_count", "scan", "diag_fill", "link",
    "uf_union", "uf_compress", "estimate", "flip_collect", "end_sweep"};

// ev (optional): NUM_SWEEP_KERNELS + 1 events recorded around the launches
static inline void mark(hipEvent_t* ev, int i, hipStream_t st) {
  if (ev) (void)hipEventRecord(ev[i], st);
}

static void launch_sweep_sse(const DevBatch& B, const LaunchCfg& L, hipStream_t st, hipEvent_t* ev) {
  mark(ev, 0, st);
  hipLaunchKernelGGL(k_sse_diag, dim3(B.nchains), dim3(SSE_THREADS), 0, st, B);  // also the scan's epoch
  mark(ev, 1, st);
  hipLaunchKernelGGL(k_scan, dim3(scan_blocks(B)), dim3(SCAN_THREADS), 0, st, B);
  mark(ev, 2, st);
  hipLaunchKernelGGL(k_sse_scatter, dim3(B.nchains), dim3(SSE_THREADS), 0, st, B);
  mark(ev, 3, st);
  launch_cluster_stage(B, L, st, ev ? ev + 3 : nullptr);
  hipLaunchKernelGGL(k_sse_gather, dim3(B.nchains), dim3(256), 0, st, B);
}

void launch_sweep(const DevBatch& B, const LaunchCfg& L, hipStream_t st, hipEvent_t* ev) {
  if (B.any_sse) {
    launch_sweep_sse(B, L, st, ev);
    return;
  }
  const int sb = scan_blocks(B);
  mark(ev, 0, st);
  diag_dispatch(B, L, st, false);  // also clears the scan's ticket, next epoch
  mark(ev, 1, st);
  hipLaunchKernelGGL(k_scan, dim3(sb), dim3(SCAN_THREADS), 0, st, B);
  mark(ev, 2, st);
  diag_dispatch(B, L, st, true);
  mark(ev, 3, st);
  launch_cluster_stage(B, L, st, ev ? ev + 3 : nullptr);
}

void launch_cluster_stage(const DevBatch& B, const LaunchCfg& L, hipStream_t st, hipEvent_t* ev) {
  // by the expected events per site: one thread per operator (few events),
  // 16 threads per site staging 128 events with 16 sites per block (29 KB of
  // LDS), a wave per site staging 512 events with 4 sites per block (29 KB),
  // or 1536 events with 2 sites per block and 4 waves per site (43 KB).  The
  // 128- and 512-event kernels take one block step per block (square 128^2,
  // beta=128: 0.119 ms with two steps per block, 0.101 with one); the
  // 1536-event kernel keeps the per-node grid (square 512^2, beta=512: 5.76
  // ms at 8 steps per block, 5.95 at one)
  if (L.link_cap == 0) {
    hipLaunchKernelGGL(k_node_init, dim3(L.blocks_node), dim3(256), 0, st, B);
    if (B.any_extra) hipLaunchKernelGGL(k_link_ops<true>, dim3(L.blocks_op), dim3(256), 0, st, B);
    else hipLaunchKernelGGL(k_link_ops<false>, dim3(L.blocks_op), dim3(256), 0, st, B);
  } else if (L.link_cap == 1) {
    const int blocks = max(1, min(L.blocks_op, (B.nsites + 255) / 256));
    hipLaunchKernelGGL(k_link_site, dim3(blocks), dim3(256), 0, st, B);
  } else if (L.link_cap <= 128) {
    const int blocks = max(1, (B.nsites + 15) / 16);
    hipLaunchKernelGGL((k_link<128, 4, 16>), dim3(blocks), dim3(256), 0, st, B);
  } else if (L.link_cap <= 512) {
    const int blocks = max(1, (B.nsites + 3) / 4);
    hipLaunchKernelGGL((k_link<512, 4, 64>), dim3(blocks), dim3(256), 0, st, B);
  } else {
    const int blocks = min(L.blocks_node, max(1, (B.nsites + 1) / 2));
    hipLaunchKernelGGL((k_link<1536, 8, 256>), dim3(blocks), dim3(512), 0, st, B);
  }
  mark(ev, 1, st);
  if (B.any_split) {
    const int blocks = max(1, min(L.blocks_op, (B.nsites + 255) / 256));
    hipLaunchKernelGGL(k_boundary, dim3(blocks), dim3(256), 0, st, B);
  }
  // (the per-operator linking kernel unites as it links)
  if (L.link_cap != 0) hipLaunchKernelGGL(k_uf_union, dim3(L.blocks_op), dim3(256), 0, st, B);
  mark(ev, 2, st);
  if (L.compress) {
    hipLaunchKernelGGL(k_uf_compress, dim3(L.blocks_node), dim3(256), 0, st, B);
    mark(ev, 3, st);
    hipLaunchKernelGGL(k_estimate<false>, dim3(L.blocks_node), dim3(EST_THREADS), 0, st, B, L.est_flush);
  } else {
    mark(ev, 3, st);
    hipLaunchKernelGGL(k_estimate<true>, dim3(L.blocks_node), dim3(EST_THREADS), 0, st, B, L.est_flush);
  }
  mark(ev, 4, st);
  if (B.any_field) {
    hipLaunchKernelGGL(k_field_flip, dim3(L.blocks_node), dim3(256), 0, st, B);
    hipLaunchKernelGGL(k_field_apply, dim3(L.blocks_node), dim3(256), 0, st, B);
  }
  hipLaunchKernelGGL(k_flip_collect, dim3(L.blocks_node), dim3(COL_THREADS), 0, st, B, L.compress ? 0 : 1);
  mark(ev, 5, st);
  hipLaunchKernelGGL(k_end_sweep, dim3(B.nchains), dim3(END_THREADS), 0, st, B);
  mark(ev, 6, st);
  if (B.nslots > 1) hipLaunchKernelGGL(k_exchange, dim3(1), dim3(256), 0, st, B);
}

void launch_site_scan(const DevBatch& B, hipStream_t st) {
  const int sb = scan_blocks(B);
  hipLaunchKernelGGL(k_scan_reset, dim3(1), dim3(1), 0, st, B);
  hipLaunchKernelGGL(k_scan, dim3(sb), dim3(SCAN_THREADS), 0, st, B);
}

void launch_canon(const DevBatch& B, int chain, int nelem, int* canon, int* labels, int* roots,
                  hipStream_t st) {
  int blocks = (nelem + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_canon_min, dim3(blocks), dim3(256), 0, st, B, chain, nelem, canon);
  hipLaunchKernelGGL(k_canon_label, dim3(blocks), dim3(256), 0, st, B, chain, nelem, canon, labels,
                     roots);
}

}  // namespace lg
