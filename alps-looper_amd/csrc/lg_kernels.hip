// gfx950 kernels for one continuous-imaginary-time loop-cluster sweep of a
// batch of S=1/2 XXZ chains (the sweep of path_integral.C:354-864).
//
// One sweep maps the previous configuration (spins at tau=0, operators in
// bond CSR, per-site world-line segments with their post-flip spins) to the
// next one:
//   diag_count   per bond: Poisson candidates from Philox, acceptance against
//                the previous world lines, kept off-diagonal operators counted
//   scan_*       bond offsets of the new operator list and segment offsets of
//                the new world lines (max(events,1) per site)
//   diag_fill    per bond: kept operators (regraphed) merged with accepted
//                candidates in time order
//   link         per operator: rank of its two legs on the site world lines
//                (merge by binary search), segment spins by kink parity
//   uf_union     per operator: lock-free union of segments by the cut rules
//   uf_compress  per node: full compression, zero the root accumulators
//   estimate     per node: improved-estimator terms into the root, flip bit of
//                the root, flipped segment spins and tau=0 spins
//   flip_collect per operator: type flip; per root: cluster terms -> record
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
// first j in [0,n) with a[j] > x (n if none)
__device__ __forceinline__ int upper_bound_d(const double* a, int n, double x) {
  int lo = 0, len = n;
  while (len > 0) {
    int h = len >> 1;
    if (a[lo + h] <= x) { lo += h + 1; len -= h + 1; } else { len = h; }
  }
  return lo;
}

__device__ __forceinline__ bool overflowed(const DevBatch& B) {
  return __hip_atomic_load(B.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// spin at tau on (chain-local) site s in the previous configuration;
// -1 when tau coincides with an event time there (such candidates are rejected)
__device__ __forceinline__ int spin_prev(const DevBatch& B, const ChainDev& P, int s, double tau) {
  int sg = P.old_site_base + s;
  int base = B.soff_old[sg];
  int m = B.soff_old[sg + 1] - base;
  const double* t = B.evt_old + base;
  int j = lower_bound_d(t, m, tau);
  if (j < m && t[j] == tau) return -1;
  return B.segspin_old[base + (j == m ? 0 : j)];
}

// ---------------------------------------------------------------------------
// (a) diagonal update, pass 1: count and store accepted candidates
// ---------------------------------------------------------------------------
template <int G>
__global__ void __launch_bounds__(256) k_diag_count(DevBatch B) {
  if (overflowed(B)) return;
  const int groups_per_block = blockDim.x / G;
  const int gl = glane<G>();
  const uint64_t below = (gl == 0) ? 0ull : ((1ull << gl) - 1);
  for (int gb = blockIdx.x * groups_per_block + threadIdx.x / G; gb < B.nbonds;
       gb += gridDim.x * groups_per_block) {
    const int c = B.bond_chain[gb];
    const ChainDev P = B.chains[c];
    const int b = gb - P.bond_base;
    const uint32_t sweep = B.sweep[c];
    int s0, s1;
    bond_sites(P, b, s0, s1);
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
    if (P.has_weight) {
      const long long cb = P.cand_base + (long long)b * P.cand_cap;
      uint64_t Tb = 0;
      uint32_t kb = 0;
      while (true) {
        U4 x = philox(sweep, STAGE_DIAG, (uint32_t)b, kb + (uint32_t)gl, P.key0, P.key1);
        uint64_t gap = exp_gap(x, P.scale);
        uint64_t incl = gscan_incl_u64<G>(gap);
        uint64_t T = Tb + incl;
        bool accept = false;
        double tau = 0;
        int g = 0;
        if (T < TGRID) {
          tau = (double)T * TGRID_INV;
          g = choose_graph(x.z, P);
          int c0 = spin_prev(B, P, s0, tau);
          int c1 = spin_prev(B, P, s1, tau);
          accept = (c0 >= 0) && (c1 >= 0) && (((g & 1) ^ c0 ^ c1) != 0);
        }
        uint64_t m = gballot<G>(accept);
        if (accept) {
          int slot = acc + __popcll(m & below);
          if (slot < P.cand_cap) {
            B.cand_t[cb + slot] = tau;
            B.cand_g[cb + slot] = (uint8_t)g;
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
  int bl[4];
  int nb = site_bonds(P, s, bl);
  int m = 0;
  for (int i = 0; i < nb; ++i) m += B.counts[P.bond_base + bl[i]];
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

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(DevBatch B) {
  if (overflowed(B)) return;
  const int nbB = (B.nbonds + SCAN_CHUNK - 1) / SCAN_CHUNK;
  int blk = blockIdx.x, which = 0, n = B.nbonds;
  if (blk >= nbB) { blk -= nbB; which = 1; n = B.nsites; }
  int base = blk * SCAN_CHUNK + threadIdx.x * SCAN_ITEMS;
  int sum = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k)
    if (base + k < n) sum += scan_value(B, which, base + k);
  int tot;
  block_excl_scan(sum, &tot);
  if (threadIdx.x == 0) B.scan_part[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_scan_top(DevBatch B) {
  if (overflowed(B)) return;
  __shared__ int carry[2];
  const int nbB = (B.nbonds + SCAN_CHUNK - 1) / SCAN_CHUNK;
  const int nbS = (B.nsites + SCAN_CHUNK - 1) / SCAN_CHUNK;
  for (int which = 0; which < 2; ++which) {
    int start = which ? nbB : 0, n = which ? nbS : nbB;
    if (threadIdx.x == 0) carry[0] = 0;
    __syncthreads();
    for (int c0 = 0; c0 < n; c0 += 1024) {
      int i = c0 + threadIdx.x;
      int v = (i < n) ? B.scan_part[start + i] : 0;
      // block inclusive scan over 1024 threads
      __shared__ int ws[16];
      int lane = lane_id(), w = threadIdx.x >> 6;
      int x = v;
      for (int d = 1; d < 64; d <<= 1) {
        int o = __shfl_up(x, d, 64);
        if (lane >= d) x += o;
      }
      if (lane == 63) ws[w] = x;
      __syncthreads();
      int off = 0, tot = 0;
      for (int k = 0; k < 16; ++k) { if (k < w) off += ws[k]; tot += ws[k]; }
      int cbase = carry[0];
      if (i < n) B.scan_part[start + i] = cbase + off + x - v;
      __syncthreads();
      if (threadIdx.x == 0) carry[0] = cbase + tot;
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      int total = carry[0];
      B.totals[which] = total;
      int cap = which ? B.node_cap : B.op_cap;
      if (total > cap) atomicOr(B.flags, which ? FLAG_NODE_OVERFLOW : FLAG_OP_OVERFLOW);
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_down(DevBatch B) {
  if (overflowed(B)) return;
  const int nbB = (B.nbonds + SCAN_CHUNK - 1) / SCAN_CHUNK;
  int blk = blockIdx.x, which = 0, n = B.nbonds;
  if (blk >= nbB) { blk -= nbB; which = 1; n = B.nsites; }
  int base = blk * SCAN_CHUNK + threadIdx.x * SCAN_ITEMS;
  int vals[SCAN_ITEMS];
  bool empty[SCAN_ITEMS];
  int sum = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int v = 0;
    empty[k] = false;
    if (base + k < n) {
      if (which == 0) {
        v = B.counts[base + k];
      } else {
        int m = site_events(B, base + k);
        empty[k] = (m == 0);
        v = empty[k] ? 1 : m;
      }
    }
    vals[k] = v;
    sum += v;
  }
  int tot;
  int off = block_excl_scan(sum, &tot) + B.scan_part[blockIdx.x];
  int* out = which ? B.soff_new : B.bo_new;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int i = base + k;
    if (i < n) {
      out[i] = off;
      if (empty[k]) {
        // event-free site: one node spanning [0,1)
        const ChainDev& P = B.chains[B.site_chain[i]];
        int node = off;
        B.evt_new[node] = EMPTY_TIME;
        B.segspin_new[node] = B.spins_old[P.old_site_base + (i - P.site_base)];
        B.evflags[node] = EV_EMPTY;
        B.node_site[node] = i;
        B.parent[node] = node;
      }
    }
    off += vals[k];
  }
  if (blk == (n - 1) / SCAN_CHUNK && threadIdx.x == 0) out[n] = B.totals[which];
}

// ---------------------------------------------------------------------------
// (a) diagonal update, pass 2: merge kept operators and accepted candidates
// ---------------------------------------------------------------------------
template <int G>
__global__ void __launch_bounds__(256) k_diag_fill(DevBatch B) {
  if (overflowed(B)) return;
  const int groups_per_block = blockDim.x / G;
  const int gl = glane<G>();
  for (int gb = blockIdx.x * groups_per_block + threadIdx.x / G; gb < B.nbonds;
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
    const long long cb = P.cand_base + (long long)b * P.cand_cap;
    const double* ct = B.cand_t + cb;
    const int info_b = gb << 3;
    // kept operators: regraph (graph_impl.h:323-327) and place
    for (int r = o0 + gl; r < o1; r += G) {
      if (!(B.oi_old[r] & 1)) continue;
      const int kr = B.kpre[r];
      const double tau = B.ot_old[r];
      U4 x = philox(sweep, STAGE_REGRAPH, (uint32_t)b, (uint32_t)kr, P.key0, P.key1);
      int g = ((uint64_t)x.x < P.toff) ? 0 : 1;
      int pos = nb + kr + lower_bound_d(ct, acc, tau);
      B.ot_new[pos] = tau;
      B.oi_new[pos] = info_b | (g << 1) | 1;
      B.kc_new[pos] = kr;
    }
    // accepted candidates: new diagonal operators
    for (int j = gl; j < acc; j += G) {
      const double tau = ct[j];
      const int g = B.cand_g[cb + j];
      int r = lower_bound_d(B.ot_old + o0, o1 - o0, tau) + o0;
      int kk = (r < o1) ? B.kpre[r] : kept;
      int pos = nb + j + kk;
      B.ot_new[pos] = tau;
      B.oi_new[pos] = info_b | (g << 1);
      B.kc_new[pos] = kk;
    }
  }
}

// ---------------------------------------------------------------------------
// (b) world-line linking
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_link(DevBatch B) {
  if (overflowed(B)) return;
  const int nops = B.totals[0];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nops; k += gridDim.x * blockDim.x) {
    const int info = B.oi_new[k];
    const int gb = info >> 3;
    const int c = B.bond_chain[gb];
    const ChainDev& P = B.chains[c];
    const int b = gb - P.bond_base;
    const double tau = B.ot_new[k];
    const int r = k - B.bo_new[gb];
    const int dir = bond_dir(P, b);
    int s0, s1;
    bond_sites(P, b, s0, s1);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int s = side ? s1 : s0;
      int bl[4];
      const int nbd = site_bonds(P, s, bl);
      int pos = r;
      int par = B.kc_new[k];
      for (int i = 0; i < nbd; ++i) {
        const int bp = bl[i];
        if (bp == b) continue;
        const int gbp = P.bond_base + bp;
        const int lo = B.bo_new[gbp], hi = B.bo_new[gbp + 1];
        const int q = (bp < b) ? upper_bound_d(B.ot_new + lo, hi - lo, tau)
                               : lower_bound_d(B.ot_new + lo, hi - lo, tau);
        pos += q;
        if (q < hi - lo) par += B.kc_new[lo + q];
        else if (hi > lo) par += B.kc_new[hi - 1] + (B.oi_new[hi - 1] & 1);
      }
      const int gs = P.site_base + s;
      const int node = B.soff_new[gs] + pos;
      B.evt_new[node] = tau;
      B.segspin_new[node] = B.spins_old[P.old_site_base + s] ^ (uint8_t)(par & 1);
      B.evflags[node] = (uint8_t)((info & 1) | (side ? 0 : EV_SRC) | (dir ? EV_DIR1 : 0));
      B.node_site[node] = gs;
      B.parent[node] = node;
      if (side) B.l1[k] = node; else B.l0[k] = node;
    }
  }
}

// ---------------------------------------------------------------------------
// (c) lock-free union-find on segments: hook the larger root under the
// smaller with atomicCAS (so every root is its cluster's minimum node) and
// shorten paths while walking (the ECL-CC scheme of Jaiganesh & Burtscher,
// HPDC'18).  parent[x] <= x always holds.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int ld_parent(int* P, int x) {
  return __hip_atomic_load(P + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int uf_find(int* P, int v) {
  int cur = ld_parent(P, v);
  if (cur != v) {
    int prev = v, next;
    while (cur > (next = ld_parent(P, cur))) {
      __hip_atomic_store(P + prev, next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      prev = cur;
      cur = next;
    }
  }
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

__device__ __forceinline__ int next_node(const DevBatch& B, int gs, int node) {
  return (node + 1 == B.soff_new[gs + 1]) ? B.soff_new[gs] : node + 1;
}

__global__ void __launch_bounds__(256) k_uf_union(DevBatch B) {
  if (overflowed(B)) return;
  const int nops = B.totals[0];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nops; k += gridDim.x * blockDim.x) {
    const int info = B.oi_new[k];
    const int gb = info >> 3;
    const ChainDev& P = B.chains[B.bond_chain[gb]];
    int s0, s1;
    bond_sites(P, gb - P.bond_base, s0, s1);
    const int l0 = B.l0[k], l1 = B.l1[k];
    const int u0 = next_node(B, P.site_base + s0, l0);
    const int u1 = next_node(B, P.site_base + s1, l1);
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

__global__ void __launch_bounds__(256) k_uf_compress(DevBatch B) {
  if (overflowed(B)) return;
  const int M = B.totals[1];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
    int p = B.parent[i];
    if (p != i) {
      int r = p, q;
      while ((q = B.parent[r]) != r) r = q;
      B.parent[i] = r;
    } else {
      B.acc[i] = ClusterAcc{{0, 0, 0, 0}, {0, 0, 0, 0}, 0, 0};
    }
  }
}

// ---------------------------------------------------------------------------
// (e) improved estimators + (d) flip bits, per segment node.
// A segment adds its length to one class slot of its root (ClusterAcc), its
// tau = 0 boundary (first segment of a site) to one class counter and its
// winding contributions to one packed 64-bit counter.  Each block handles
// chunks of EST_CHUNK consecutive nodes and merges the segments that share a
// root in an LDS hash table (open addressing, HT slots) before touching global
// memory: the large loops of a cold antiferromagnet (a few clusters holding
// most segments) then cost one global update per chunk instead of one per
// segment.  Segments whose root finds no slot update global memory directly.
// ---------------------------------------------------------------------------
constexpr int EST_THREADS = 256;
constexpr int EST_ITEMS = 16;
constexpr int EST_CHUNK = EST_THREADS * EST_ITEMS;
constexpr int HT_BITS = 9;
constexpr int HT = 1 << HT_BITS;
constexpr int HT_PROBES = 8;

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__global__ void __launch_bounds__(EST_THREADS) k_estimate(DevBatch B) {
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
  const int nchunks = (M + EST_CHUNK - 1) / EST_CHUNK;
  for (int chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    for (int it = 0; it < EST_ITEMS; ++it) {
      const int i = chunk * EST_CHUNK + it * EST_THREADS + tid;
      if (i >= M) break;
      const int gs = B.node_site[i];
      const int c = B.site_chain[gs];
      const ChainDev& P = B.chains[c];
      const int s = gs - P.site_base;
      const int r = B.parent[i];
      const int chain_node0 = B.soff_new[P.site_base];
      const uint32_t f =
          philox(B.sweep[c], STAGE_FLIP, (uint32_t)(r - chain_node0), 0u, P.key0, P.key1).x & 1u;
      B.fbit[i] = (uint8_t)f;
      const uint8_t fl = B.evflags[i];
      const int spin = B.segspin_new[i];
      const int sig = 1 - 2 * spin;
      int dw0 = 0, dw1 = 0;
      double len;
      bool first;
      if (fl & EV_EMPTY) {
        len = 1.0;
        first = true;
      } else {
        const int base = B.soff_new[gs];
        const int last = B.soff_new[gs + 1] - 1;
        first = (i == base);
        const double thi = B.evt_new[i];
        len = first ? ((thi + 1.0) - B.evt_new[last]) : (thi - B.evt_new[i - 1]);
        const uint8_t flo = B.evflags[first ? last : i - 1];
        // this segment is the lower leg of the event above it and the upper
        // leg of the event below it (stiffness.h:94-105, source legs only)
        if (fl & EV_SRC) { if (fl & EV_DIR1) dw1 += sig; else dw0 += sig; }
        if (flo & EV_SRC) { if (flo & EV_DIR1) dw1 -= sig; else dw0 -= sig; }
      }
      if (first) B.spins_new[gs] = (uint8_t)(spin ^ f);
      B.segspin_new[i] = (uint8_t)(spin ^ f);
      const int cls = seg_class(spin, gauge(P, s));
      const long long wp = pack_winding(dw0, dw1);
      int h = (int)(((uint32_t)r * 2654435761u) >> (32 - HT_BITS));
      bool placed = false;
      for (int p = 0; p < HT_PROBES; ++p) {
        const int k = atomicCAS(&hkey[h], -1, r);
        if (k == -1 || k == r) {
          atomicAdd(&hL[cls][h], len);
          if (first) atomicAdd(&hn[cls >> 1][h], 1u << (16 * (cls & 1)));
          if (wp) atomicAdd(&hw[h], (unsigned long long)wp);
          placed = true;
          break;
        }
        h = (h + 1) & (HT - 1);
      }
      if (!placed) {
        ClusterAcc* a = B.acc + r;
        unsafeAtomicAdd(&a->L[cls], len);
        if (first) atomicAdd(&a->n[cls], 1);
        if (wp) atomicAdd((unsigned long long*)&a->w, (unsigned long long)wp);
      }
    }
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

// ---------------------------------------------------------------------------
// (d) operator flip, then the improved-estimator collector
// (susceptibility.h:177-198, stiffness.h:117-121) over the roots.
// Collection is chunked: chunk j covers nodes [j*COL_CHUNK, (j+1)*COL_CHUNK)
// and belongs to the chain of its first node; its sum over the roots of that
// chain is written to col_part[j] (no atomics, fixed order), roots of other
// chains in the chunk (chunks straddling a chain boundary) are added to
// rec_acc atomically.  end_sweep adds the partials of each chain in order.
// ---------------------------------------------------------------------------
constexpr int COL_THREADS = 256;
constexpr int COL_ITEMS = 16;
constexpr int COL_CHUNK = COL_THREADS * COL_ITEMS;
constexpr int NCOL = 16;

__device__ __forceinline__ void cluster_terms(const ClusterAcc& a, double* v) {
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
}

// record slot of collector term q (R_NC, then R_UMAG0 .. R_W2)
__device__ __forceinline__ int term_slot(int q) { return q == 0 ? R_NC : R_UMAG0 + (q - 1); }

__global__ void __launch_bounds__(COL_THREADS) k_flip_collect(DevBatch B) {
  if (overflowed(B)) return;
  const int nops = B.totals[0];
  const int M = B.totals[1];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nops; k += gridDim.x * blockDim.x) {
    const int info = B.oi_new[k];
    const int gb = info >> 3;
    const ChainDev& P = B.chains[B.bond_chain[gb]];
    int s0, s1;
    bond_sites(P, gb - P.bond_base, s0, s1);
    const int l0 = B.l0[k];
    const int u0 = next_node(B, P.site_base + s0, l0);
    if (B.fbit[l0] ^ B.fbit[u0]) B.oi_new[k] = info ^ 1;
  }
  // per-chunk sums by chain: slot k holds chain c_first + k (chains are
  // contiguous in node order, so a chunk meets few of them)
  constexpr int CSLOTS = 8;
  __shared__ double slot[CSLOTS][NCOL];
  const int nchunks = (M + COL_CHUNK - 1) / COL_CHUNK;
  for (int chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const int base = chunk * COL_CHUNK;
    const int c0 = B.site_chain[B.node_site[base]];
    for (int s = threadIdx.x; s < CSLOTS * NCOL; s += COL_THREADS) slot[s / NCOL][s % NCOL] = 0;
    __syncthreads();
    for (int it = 0; it < COL_ITEMS; ++it) {
      const int i = base + it * COL_THREADS + threadIdx.x;
      double w[NCOL];
#pragma unroll
      for (int q = 0; q < NCOL; ++q) w[q] = 0;
      int c = -1;
      if (i < M && B.parent[i] == i) {
        c = B.site_chain[B.node_site[i]];
        cluster_terms(B.acc[i], w);
      }
      int cmax = c, cmin = (c < 0) ? 0x7fffffff : c;
      for (int d = 32; d > 0; d >>= 1) {
        cmax = max(cmax, __shfl_xor(cmax, d, 64));
        cmin = min(cmin, __shfl_xor(cmin, d, 64));
      }
      if (cmax < 0) continue;  // no root in this wave
      if (cmin == cmax) {
#pragma unroll
        for (int q = 0; q < NCOL; ++q) w[q] = wave_sum_d(w[q]);
        if (lane_id() == 0) {
          const int k = cmax - c0;
          if (k < CSLOTS) {
            for (int q = 0; q < NCOL; ++q) atomicAdd(&slot[k][q], w[q]);
          } else {
            double* rec = B.rec_acc + (size_t)cmax * R_COUNT;
            for (int q = 0; q < NCOL; ++q) unsafeAtomicAdd(rec + term_slot(q), w[q]);
          }
        }
      } else if (c >= 0) {
        const int k = c - c0;
        if (k < CSLOTS) {
          for (int q = 0; q < NCOL; ++q) atomicAdd(&slot[k][q], w[q]);
        } else {
          double* rec = B.rec_acc + (size_t)c * R_COUNT;
          for (int q = 0; q < NCOL; ++q) unsafeAtomicAdd(rec + term_slot(q), w[q]);
        }
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
    const int j0 = (start + COL_CHUNK - 1) / COL_CHUNK, j1 = (end + COL_CHUNK - 1) / COL_CHUNK;
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
  if (tid == 0) {
    const double nop = bad ? 0.0 : (double)(B.bo_new[P.bond_base + P.Nb] - B.bo_new[P.bond_base]);
    out[R_NOP] = nop;
    out[R_ENERGY] = P.offset_total - nop / P.beta;
    out[R_BETA] = P.beta;
    out[R_SWEEP] = (double)B.sweep[c];
    B.sweep[c] += 1;
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
// canonical labels of the leg elements of one chain (debug / parity export):
// element s -> boundary of site s, N + 4k + {0,1,2,3} -> legs l0,l1,u0,u1 of
// operator k (chain-local bond-major order); label = min element index of
// its cluster.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int element_node(const DevBatch& B, const ChainDev& P, int e, int op0) {
  if (e < P.N) return B.soff_new[P.site_base + e];
  int k = (e - P.N) >> 2, leg = (e - P.N) & 3;
  int kg = op0 + k;
  int gb = B.oi_new[kg] >> 3;
  int s0, s1;
  bond_sites(P, gb - P.bond_base, s0, s1);
  if (leg == 0) return B.l0[kg];
  if (leg == 1) return B.l1[kg];
  if (leg == 2) return next_node(B, P.site_base + s0, B.l0[kg]);
  return next_node(B, P.site_base + s1, B.l1[kg]);
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
// host-side launchers
// ---------------------------------------------------------------------------
template <int G>
static void launch_diag(const DevBatch& B, const LaunchCfg& L, hipStream_t st, bool fill) {
  int groups_per_block = 256 / G;
  int blocks = (B.nbonds + groups_per_block - 1) / groups_per_block;
  if (blocks > L.max_blocks_bond) blocks = L.max_blocks_bond;
  if (blocks < 1) blocks = 1;
  if (fill) hipLaunchKernelGGL(k_diag_fill<G>, dim3(blocks), dim3(256), 0, st, B);
  else hipLaunchKernelGGL(k_diag_count<G>, dim3(blocks), dim3(256), 0, st, B);
}

static void diag_dispatch(const DevBatch& B, const LaunchCfg& L, hipStream_t st, bool fill) {
  switch (L.group) {
    case 4: launch_diag<4>(B, L, st, fill); break;
    case 8: launch_diag<8>(B, L, st, fill); break;
    case 16: launch_diag<16>(B, L, st, fill); break;
    case 32: launch_diag<32>(B, L, st, fill); break;
    default: launch_diag<64>(B, L, st, fill); break;
  }
}

int col_chunks(int node_cap) { return (node_cap + COL_CHUNK - 1) / COL_CHUNK; }

int scan_blocks(const DevBatch& B) {
  return (B.nbonds + SCAN_CHUNK - 1) / SCAN_CHUNK + (B.nsites + SCAN_CHUNK - 1) / SCAN_CHUNK;
}

const char* const kKernelNames[NUM_SWEEP_KERNELS] = {
    "diag_count", "scan_reduce", "scan_top", "scan_down", "diag_fill", "link",
    "uf_union", "uf_compress", "estimate", "flip_collect", "end_sweep"};

// ev (optional): NUM_SWEEP_KERNELS + 1 events recorded around the launches
static inline void mark(hipEvent_t* ev, int i, hipStream_t st) {
  if (ev) (void)hipEventRecord(ev[i], st);
}

void launch_sweep(const DevBatch& B, const LaunchCfg& L, hipStream_t st, hipEvent_t* ev) {
  const int sb = scan_blocks(B);
  mark(ev, 0, st);
  diag_dispatch(B, L, st, false);
  mark(ev, 1, st);
  hipLaunchKernelGGL(k_scan_reduce, dim3(sb), dim3(SCAN_THREADS), 0, st, B);
  mark(ev, 2, st);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, B);
  mark(ev, 3, st);
  hipLaunchKernelGGL(k_scan_down, dim3(sb), dim3(SCAN_THREADS), 0, st, B);
  mark(ev, 4, st);
  diag_dispatch(B, L, st, true);
  mark(ev, 5, st);
  launch_cluster_stage(B, L, st, ev ? ev + 5 : nullptr);
}

void launch_cluster_stage(const DevBatch& B, const LaunchCfg& L, hipStream_t st, hipEvent_t* ev) {
  hipLaunchKernelGGL(k_link, dim3(L.blocks_op), dim3(256), 0, st, B);
  mark(ev, 1, st);
  hipLaunchKernelGGL(k_uf_union, dim3(L.blocks_op), dim3(256), 0, st, B);
  mark(ev, 2, st);
  hipLaunchKernelGGL(k_uf_compress, dim3(L.blocks_node), dim3(256), 0, st, B);
  mark(ev, 3, st);
  hipLaunchKernelGGL(k_estimate, dim3(L.blocks_node), dim3(EST_THREADS), 0, st, B);
  mark(ev, 4, st);
  hipLaunchKernelGGL(k_flip_collect, dim3(L.blocks_node), dim3(COL_THREADS), 0, st, B);
  mark(ev, 5, st);
  hipLaunchKernelGGL(k_end_sweep, dim3(B.nchains), dim3(END_THREADS), 0, st, B);
  mark(ev, 6, st);
}

void launch_site_scan(const DevBatch& B, hipStream_t st) {
  const int sb = scan_blocks(B);
  hipLaunchKernelGGL(k_scan_reduce, dim3(sb), dim3(SCAN_THREADS), 0, st, B);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, B);
  hipLaunchKernelGGL(k_scan_down, dim3(sb), dim3(SCAN_THREADS), 0, st, B);
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
