// Host side of libloopgpu: batches of chains resident in HBM, sweep launches
// (plain or HIP-graph replay), configuration upload/download and the C-ABI of
// include/loopgpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/loopgpu.h"
#include "lg_common.h"
#include "lg_host.h"
#include "lg_kernels.h"

namespace lg {

thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

std::string g_last_error_copy() { return g_last_error; }

#define LG_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return set_error(LG_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// model setup (host): XXZ weights of weight_impl.h:165-188
// ---------------------------------------------------------------------------
static void xxz_weights(double jxy_in, double jz, double a, double v[4], double* offset) {
  double jxy = std::fabs(jxy_in);
  a = std::min(1.0, std::max(0.0, a));
  auto crop0 = [](double x) { return x > 0 ? x : 0.0; };
  if (jxy + std::fabs(jz) > 1e-10) {
    if (jxy - jz > 2 * a * jxy) {
      v[0] = crop0(std::min(jxy / 2, (jxy + jz) / 4));
      v[1] = crop0(std::min(jxy / 2, (jxy - jz) / 4));
      v[2] = crop0(-(jxy - jz) / 2.0);
      v[3] = crop0(-(jxy + jz) / 2);
    } else {
      v[0] = (1 - a) * jxy / 2;
      v[1] = a * jxy / 2;
      v[2] = -((1 - 2 * a) * jxy - jz) / 2;
      v[3] = 0;
    }
  } else {
    v[0] = v[1] = v[2] = v[3] = 0;
  }
  *offset = (v[0] + v[1] + v[2] + v[3]) / 2;
}

static_assert(LG_MAX_TWO_S == MAX_NV, "the largest 2S is the boundary kernel's virtual-site count");

int check_params(const lg_params& p) {
  if (p.lattice != LG_LATTICE_CHAIN && p.lattice != LG_LATTICE_SQUARE)
    return set_error(LG_ERR_ARG, "lattice must be chain (0) or square (1)");
  if (p.two_s < 0 || p.two_s > LG_MAX_TWO_S)
    return set_error(LG_ERR_ARG, "two_s (2S) must be 1 .. " + std::to_string(LG_MAX_TWO_S) + " (S=1/2 .. " +
                                     std::to_string(LG_MAX_TWO_S / 2) + "; 0 means 1)");
  if (p.L < 2 || (p.L & 1))
    return set_error(LG_ERR_ARG, "L must be even and >= 2 (bipartite, sign-free lattice)");
  {
    // every index of a chain (virtual bonds, in-site and site bonds, sites)
    // stays far inside int32
    const long long nv = p.two_s >= 2 ? p.two_s : 1;
    const long long nr = p.lattice == LG_LATTICE_CHAIN ? p.L : (long long)p.L * p.L;
    long long nb = nv * nv * (p.lattice == LG_LATTICE_CHAIN ? nr : 2 * nr);
    if (p.d != 0) nb += nr * nv * (nv - 1) / 2;  // in-site bonds
    if (p.hx != 0) nb += nr * nv;                // site bonds
    if (nb > (1ll << 27)) return set_error(LG_ERR_ARG, "lattice too large (more than 2^27 bonds)");
  }
  if (!(p.beta > 0) || !std::isfinite(p.beta))
    return set_error(LG_ERR_ARG, "beta must be positive (negative temperature)");
  if (!std::isfinite(p.jxy) || !std::isfinite(p.jz) || !std::isfinite(p.h) || !std::isfinite(p.hx) ||
      !std::isfinite(p.d))
    return set_error(LG_ERR_ARG, "couplings and fields must be finite");
  if (p.algorithm != LG_ALGORITHM_PATH_INTEGRAL && p.algorithm != LG_ALGORITHM_SSE)
    return set_error(LG_ERR_ARG, "algorithm must be 0 (loop; path integral) or 1 (loop; sse)");
  if (p.algorithm == LG_ALGORITHM_SSE) {
    if (p.h != 0)
      return set_error(LG_ERR_ARG, "the SSE worker takes no longitudinal field (h = 0; sse.C:123 rejects it too)");
    const long long nv = p.two_s >= 2 ? p.two_s : 1;
    const long long nr = p.lattice == LG_LATTICE_CHAIN ? p.L : (long long)p.L * p.L;
    long long nb = nv * nv * (p.lattice == LG_LATTICE_CHAIN ? p.L : 2ll * p.L * p.L);
    if (p.d != 0) nb += nr * nv * (nv - 1) / 2;  // in-site bonds
    if (p.hx != 0) nb += nv * nr;                // site bonds
    if (nv * nr > SSE_MAX_SITES || nb > SSE_MAX_BONDS)
      return set_error(LG_ERR_ARG, "the SSE worker takes at most " + std::to_string(SSE_MAX_SITES) +
                                       " (virtual) sites and " + std::to_string(SSE_MAX_BONDS) + " bonds");
  }
  return LG_OK;
}

// graph thresholds of weights v: g = first g with x < thr[g] (2^32 from the
// last graph of nonzero weight on); regraph threshold of an off-diagonal
// operator v0 / (v0 + v1)
static void graph_thresholds(const double* v, uint64_t* thr, uint64_t* toff) {
  const double W = v[0] + v[1] + v[2] + v[3];
  int last = -1;
  for (int g = 0; g < 4; ++g) if (v[g] > 0) last = g;
  double cum = 0;
  for (int g = 0; g < 3; ++g) {
    cum += v[g];
    if (last < 0 || g >= last) thr[g] = 1ull << 32;
    else thr[g] = (uint64_t)std::floor(std::ldexp(cum / W, 32));
  }
  const double s01 = v[0] + v[1];
  *toff = (s01 > 0) ? (uint64_t)std::floor(std::ldexp(v[0] / s01, 32)) : (1ull << 32);
}

// SSE insertion thresholds (ChainDev::sse_cls, sse_tl, sse_td) and the
// in-site bonds' choose_diagonal thresholds; the oracle's philox mode
// evaluates the same expressions
static void sse_selection(ChainDev& d, const double* v, double dw, double dion, double fs) {
  const double wl = d.nb_lat * (v[0] + v[1] + v[2] + v[3]);
  const double wd = (d.nb_real - d.nb_lat) * dw, ws = (d.Nb - d.nb_real) * d.site_w;
  const double wt = wl + wd + ws;
  d.sse_wtot = wt;
  auto t32 = [&](double x) -> uint64_t {
    const double y = std::floor(std::ldexp(x / wt, 32));
    return y >= 4294967296.0 ? (1ull << 32) : (uint64_t)y;
  };
  d.sse_cls[0] = (wd + ws > 0) ? t32(wl) : (1ull << 32);
  d.sse_cls[1] = (ws > 0) ? t32(wl + wd) : (1ull << 32);
  int last = -1;
  for (int g = 0; g < 4; ++g) if (v[g] > 0) last = g;
  double cum = 0;
  for (int g = 0; g < 3; ++g) {
    cum += v[g];
    d.sse_tl[g] = (last < 0 || g >= last) ? d.sse_cls[0] : t32(d.nb_lat * cum);
  }
  double vd[4] = {0, 0, 0, 0}, offd;
  if (dion != 0) xxz_weights(0.0, 2 * dion, fs, vd, &offd);
  last = -1;
  for (int g = 0; g < 4; ++g) if (vd[g] > 0) last = g;
  cum = 0;
  for (int g = 0; g < 3; ++g) {
    cum += vd[g];
    d.sse_td[g] = (last < 0 || g >= last) ? d.sse_cls[1] : t32(wl + (d.nb_real - d.nb_lat) * cum);
  }
  for (int c = 0; c < 2; ++c) {
    const double den = vd[c] + vd[c + 2];
    d.dpdiag[c] = (den > 0) ? (uint64_t)std::floor(std::ldexp(vd[c] / den, 32)) : (1ull << 32);
  }
}

ChainDev make_chain_dev(const lg_params& p) {
  ChainDev d;
  std::memset(&d, 0, sizeof(d));
  d.kind = p.lattice;
  d.L = p.L;
  // S > 1/2: 2S virtual sites per real site, (2S)^2 virtual bonds per real bond
  d.nv = (p.two_s >= 2) ? p.two_s : 1;
  d.N = d.nv * ((p.lattice == LG_LATTICE_CHAIN) ? p.L : p.L * p.L);
  d.Nb = d.nv * d.nv * ((p.lattice == LG_LATTICE_CHAIN) ? p.L : 2 * p.L * p.L);
  d.dim = (p.lattice == LG_LATTICE_CHAIN) ? 1 : 2;
  double v[4], off;
  xxz_weights(p.jxy, p.jz, p.force_scatter, v, &off);
  double W = v[0] + v[1] + v[2] + v[3];
  d.beta = p.beta;
  d.has_weight = W > 0;
  d.wsum = W;
  d.scale = candidate_scale(p.beta, W);
  d.offset_total = d.Nb * off;
  graph_thresholds(v, d.thr, &d.toff);
  // SSE: choose_diagonal (graph_impl.h:297-312): g = 0 w.p. v0 / (v0 + v2)
  // for an antiparallel pair, g = 1 w.p. v1 / (v1 + v3) for a parallel one
  d.sse = p.algorithm == LG_ALGORITHM_SSE ? 1 : 0;
  for (int c = 0; c < 2; ++c) {
    const double den = v[c] + v[c + 2];
    d.pdiag[c] = (den > 0) ? (uint64_t)std::floor(std::ldexp(v[c] / den, 32)) : (1ull << 32);
  }
  d.key0 = p.seed;
  d.key1 = p.chain_id;
  d.h = p.h;
  // single-ion term D (S^z)^2 = D nv/4 + sum_{a<c} 2D s_a s_c per real site
  // (model_parameter.h:592-595, weight_impl.h:194-197, 400-404): in-site
  // bonds after the lattice's, the pairs of each real site's virtual sites,
  // with the weights of an Ising coupling 2D; the constant in the offset
  d.nb_lat = d.Nb;
  double dw = 0;
  if (p.d != 0) {
    double vd[4], offd;
    xxz_weights(0.0, 2 * p.d, p.force_scatter, vd, &offd);
    dw = vd[0] + vd[1] + vd[2] + vd[3];
    d.npair = d.nv * (d.nv - 1) / 2;
    const int nr = d.N / d.nv;
    d.Nb += nr * d.npair;
    d.dwsum = dw;
    d.dscale = candidate_scale(p.beta, dw);
    graph_thresholds(vd, d.dthr, &d.dtoff);
    d.offset_total += nr * d.npair * offd + nr * 0.25 * d.nv * p.d;
  }
  // transverse field: one site graph of weight and offset |hx|/2 per site
  // (weight_impl.h:63-88), its operators on a bond of the site's own after
  // the lattice's (and in-site) bonds
  d.nb_real = d.Nb;
  if (p.hx != 0) {
    d.site_w = std::fabs(p.hx) / 2;
    d.site_scale = candidate_scale(p.beta, d.site_w);
    d.Nb += d.N;
    d.offset_total += d.N * d.site_w;
  }
  double lam = p.beta * std::max(std::max(W, dw), d.site_w);
  d.cand_cap = (int)std::ceil(lam + 12.0 * std::sqrt(lam) + 16.0);
  sse_selection(d, v, dw, d.npair > 0 ? p.d : 0.0, p.force_scatter);
  // an antiferromagnetic XY coupling with a transverse field has a sign
  // problem that no gauge removes (model_parameter.h:617-644 finds the
  // lattice frustrated; the reference takes its SIGN path): every sweep
  // then records the configuration's sign
  if (p.hx != 0 && p.jxy > 0) {
    d.sgn_lattice = 1;
    d.sgn_site = p.hx < 0 ? 1 : 0;
  }
  return d;
}

// upper bound of the mean operator count: beta * Nb * (offset_b + |Jz|/4 + |Jxy|/2)
static long long op_capacity(const lg_params& p, const ChainDev& d) {
  double v[4], off;
  xxz_weights(p.jxy, p.jz, p.force_scatter, v, &off);
  double bound = p.beta * d.nb_lat * (off + std::fabs(p.jz) / 4 + std::fabs(p.jxy) / 2);
  if (d.npair > 0) {  // in-site bonds: an Ising coupling 2D
    double vd[4], offd;
    xxz_weights(0.0, 2 * p.d, p.force_scatter, vd, &offd);
    bound += p.beta * (d.nb_real - d.nb_lat) * (offd + std::fabs(p.d) / 2);
  }
  bound += p.beta * d.N * std::fabs(p.hx);  // site operators: beta |hx| (1/2 + |<S^x>|) per site at most
  return (long long)std::ceil(bound + 10.0 * std::sqrt(bound) + 1024.0);
}

// ---------------------------------------------------------------------------
// Batch
// ---------------------------------------------------------------------------
Batch::~Batch() { release(); }

void Batch::release() {
  for (int i = 0; i < 2; ++i)
    if (graph_exec[i]) { hipGraphExecDestroy(graph_exec[i]); graph_exec[i] = nullptr; }
  for (void* p : allocations) hipFree(p);
  allocations.clear();
  if (rec_host) { hipHostFree(rec_host); rec_host = nullptr; }
  if (stream) { hipStreamDestroy(stream); stream = nullptr; }
}

template <class T>
int Batch::alloc(T*& ptr, size_t n) {
  void* p = nullptr;
  size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    ptr = nullptr;
    return set_error(LG_ERR_DEVICE, "hipMalloc of " + std::to_string(bytes) + " bytes: " +
                                        hipGetErrorString(e));
  }
  allocations.push_back(p);
  ptr = (T*)p;
  return LG_OK;
}

int Batch::init_for_config(const lg_params* p, long long nops) {
  min_op_cap = nops + 1024;
  return init(p, 1);
}

Footprint footprint(const lg_params& p) {
  ChainDev d = make_chain_dev(p);
  Footprint f;
  f.sites = d.N;
  f.bonds = d.Nb;
  f.ops = op_capacity(p, d);
  f.cand = (long long)d.Nb * d.cand_cap;
  return f;
}

static int check_device() {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_error(LG_ERR_DEVICE, "no HIP device available");
  return LG_OK;
}

// Static batch: capacity is exactly what the given chains need.
int Batch::init(const lg_params* p, int n) {
  if (n <= 0) return set_error(LG_ERR_ARG, "need at least one chain");
  for (int c = 0; c < n; ++c) {
    int rc = check_params(p[c]);
    if (rc) return rc;
    if (p[c].algorithm != p[0].algorithm) return set_error(LG_ERR_ARG, "a batch holds chains of one algorithm");
  }
  int rc = check_device();
  if (rc) return rc;
  Capacity cap;
  cap.chains = n;
  for (int c = 0; c < n; ++c) {
    Footprint f = footprint(p[c]);
    cap.sites += f.sites;
    cap.bonds += f.bonds;
    cap.ops += f.ops;
    cap.cand += f.cand;
  }
  cap.ops = std::max(cap.ops, min_op_cap);
  rc = reserve(cap, false);
  if (rc) return rc;
  if (p[0].algorithm == LG_ALGORITHM_SSE)
    for (int i = 0; i < 2; ++i)
      if ((rc = alloc(sse_str[i], op_cap))) return rc;
  std::vector<ChainDev> ds(n);
  for (int c = 0; c < n; ++c) ds[c] = make_chain_dev(p[c]);
  params.assign(p, p + n);
  cap_beta = 0;
  for (int c = 0; c < n; ++c) cap_beta = std::max(cap_beta, p[c].beta);
  rc = set_layout(ds, std::vector<uint64_t>(n, 0));
  if (rc) return rc;
  return reset_state();
}

// Replica-exchange batch: n replicas of model p at inverse temperatures
// betas[t] (replica t keyed by chain_id p.chain_id + t).  Every replica gets
// the candidate slots of the largest beta, since temperatures move between
// chains; the operator capacity is the sum over the temperatures, which the
// exchange permutes but does not change.
int Batch::init_exchange(const lg_params& p, const double* betas, int n, uint32_t seed) {
  if (n < 2) return set_error(LG_ERR_ARG, "exchange needs at least two temperatures");
  if (!betas) return set_error(LG_ERR_ARG, "null pointer");
  if (p.hx != 0 || p.d != 0 || p.algorithm != LG_ALGORITHM_PATH_INTEGRAL)
    return set_error(LG_ERR_UNSUPPORTED,
                     "replica exchange with a transverse field, a single-ion term or SSE chains is not built");
  std::vector<lg_params> ps(n, p);
  for (int t = 0; t < n; ++t) {
    ps[t].beta = betas[t];
    ps[t].chain_id = p.chain_id + (uint32_t)t;
    int rc = check_params(ps[t]);
    if (rc) return rc;
  }
  int rc = check_device();
  if (rc) return rc;
  std::vector<ChainDev> ds(n);
  int cand_cap = 0;
  for (int t = 0; t < n; ++t) {
    ds[t] = make_chain_dev(ps[t]);
    cand_cap = std::max(cand_cap, ds[t].cand_cap);
  }
  Capacity cap;
  cap.chains = n;
  for (int t = 0; t < n; ++t) {
    ds[t].cand_cap = cand_cap;
    Footprint f = footprint(ps[t]);
    cap.sites += f.sites;
    cap.bonds += f.bonds;
    cap.ops += f.ops;
    cap.cand += (long long)ds[t].Nb * cand_cap;
  }
  if ((rc = reserve(cap, false))) return rc;
  params = ps;
  if ((rc = set_layout(ds, std::vector<uint64_t>(n, 0)))) return rc;
  if ((rc = reset_state())) return rc;
  if ((rc = alloc(xslot, n)) || (rc = alloc(xbeta, n)) || (rc = alloc(xscale, n)) ||
      (rc = alloc(xlnr, n - 1)) || (rc = alloc(xacc, n - 1)) || (rc = alloc(xstep, 1)))
    return rc;
  std::vector<int> slot(n);
  std::vector<double> sc(n), lnr(n - 1);
  for (int t = 0; t < n; ++t) { slot[t] = t; sc[t] = ds[t].scale; }
  for (int t = 0; t + 1 < n; ++t) lnr[t] = det_log(betas[t + 1] / betas[t]);
  LG_HIP(hipMemcpy(xslot, slot.data(), n * sizeof(int), hipMemcpyHostToDevice));
  LG_HIP(hipMemcpy(xbeta, betas, n * sizeof(double), hipMemcpyHostToDevice));
  LG_HIP(hipMemcpy(xscale, sc.data(), n * sizeof(double), hipMemcpyHostToDevice));
  LG_HIP(hipMemcpy(xlnr, lnr.data(), (n - 1) * sizeof(double), hipMemcpyHostToDevice));
  LG_HIP(hipMemset(xacc, 0, (n - 1) * sizeof(unsigned long long)));
  LG_HIP(hipMemset(xstep, 0, sizeof(uint32_t)));
  nslots = n;
  xkey0 = seed;
  xkey1 = 0xffffffffu;
  invalidate_graphs();
  return LG_OK;
}

int Batch::exchange_state(int32_t* chain_at, uint64_t* accepted, uint64_t* steps) {
  if (nslots < 2) return set_error(LG_ERR_STATE, "not an exchange batch");
  LG_HIP(hipStreamSynchronize(stream));
  if (chain_at) LG_HIP(hipMemcpy(chain_at, xslot, nslots * sizeof(int), hipMemcpyDeviceToHost));
  if (accepted) LG_HIP(hipMemcpy(accepted, xacc, (nslots - 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
  if (steps) {
    uint32_t s32 = 0;
    LG_HIP(hipMemcpy(&s32, xstep, sizeof(uint32_t), hipMemcpyDeviceToHost));
    *steps = s32;
  }
  return LG_OK;
}

// Server batch: capacity from a budget, no chains yet.  Site/bond arrays get
// room for the previous layout plus newly appended chains (2x the budget).
int Batch::init_capacity(const Capacity& cap) {
  int rc = check_device();
  if (rc) return rc;
  rc = reserve(cap, true);
  if (rc) return rc;
  chains.clear();
  params.clear();
  host_sweep.clear();
  nsites = nbonds = 0;
  old_ext_sites = old_ext_bonds = 0;
  old_ext_nodes = 0;
  cur = 0;
  std::vector<int> z(2, 0);
  LG_HIP(hipMemcpy(st[cur].bo, z.data(), sizeof(int), hipMemcpyHostToDevice));
  LG_HIP(hipMemcpy(st[cur].soff, z.data(), sizeof(int), hipMemcpyHostToDevice));
  return LG_OK;
}

int Batch::reserve(const Capacity& cap, bool dynamic) {
  if (cap.chains < 1) return set_error(LG_ERR_ARG, "capacity needs at least one chain");
  const long long sites = dynamic ? 2 * cap.sites : cap.sites;
  const long long bonds = dynamic ? 2 * cap.bonds : cap.bonds;
  if (sites >= (1ll << 30) || bonds >= (1ll << 28) || cap.ops >= (1ll << 30))
    return set_error(LG_ERR_ARG, "batch too large for 32-bit indexing");
  long long ncap = 2 * cap.ops + sites;
  if (ncap >= (1ll << 31) - 1) return set_error(LG_ERR_ARG, "node capacity exceeds 32-bit indexing");
  budget = cap;
  cap_chains = cap.chains;
  cap_sites = (int)sites;
  cap_bonds = (int)bonds;
  op_cap = (int)cap.ops;
  node_cap = (int)ncap;
  cand_total = cap.cand;
  LG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  return allocate();
}

int Batch::allocate() {
  int rc = 0;
  for (int i = 0; i < 2; ++i) {
    State& s = st[i];
    if ((rc = alloc(s.spins, cap_sites))) return rc;
    if ((rc = alloc(s.bo, cap_bonds + 1))) return rc;
    if ((rc = alloc(s.ot, op_cap))) return rc;
    if ((rc = alloc(s.oi, op_cap))) return rc;
    if ((rc = alloc(s.soff, cap_sites + 1))) return rc;
    if ((rc = alloc(s.evt, node_cap))) return rc;
    if ((rc = alloc(s.segspin, node_cap))) return rc;
  }
  if ((rc = alloc(d_chains, cap_chains))) return rc;
  if ((rc = alloc(bond_chain, cap_bonds))) return rc;
  if ((rc = alloc(site_chain, cap_sites))) return rc;
  if ((rc = alloc(counts, cap_bonds))) return rc;
  if ((rc = alloc(acc_count, cap_bonds))) return rc;
  if ((rc = alloc(kpre, op_cap))) return rc;
  if ((rc = alloc(cand_t, cand_total))) return rc;
  if ((rc = alloc(cand_g, cand_total))) return rc;
  if ((rc = alloc(kc_new, op_cap))) return rc;
  if ((rc = alloc(l0, op_cap))) return rc;
  if ((rc = alloc(l1, op_cap))) return rc;
  if ((rc = alloc(u0, op_cap))) return rc;
  if ((rc = alloc(u1, op_cap))) return rc;
  if ((rc = alloc(evflags, node_cap))) return rc;
  if ((rc = alloc(chain_node0, cap_chains))) return rc;
  // node_site: every segment's site in batches of more than one chain, the
  // first segment's site of each site for one chain
  if ((rc = alloc(node_site, node_cap))) return rc;
  if ((rc = alloc(parent, node_cap))) return rc;
  if ((rc = alloc(fbit, node_cap))) return rc;
  if ((rc = alloc(fmask, node_cap / 64 + 1))) return rc;
  if ((rc = alloc(rmask, node_cap / 64 + 1))) return rc;
  if ((rc = alloc(acc, node_cap))) return rc;
  // the estimator accumulates into zeroed accumulators; the collector
  // clears each root's after reading it
  LG_HIP(hipMemset(acc, 0, (size_t)node_cap * sizeof(ClusterAcc)));
  scan_nparts = (cap_bonds + 2047) / 2048 + (cap_sites + 2047) / 2048;
  if ((rc = alloc(scan_status, scan_nparts))) return rc;
  if ((rc = alloc(scan_ctl, 2))) return rc;
  LG_HIP(hipMemset(scan_status, 0, scan_nparts * sizeof(unsigned long long)));
  LG_HIP(hipMemset(scan_ctl, 0, 2 * sizeof(int)));
  if ((rc = alloc(col_part, (size_t)col_chunks(node_cap) * NUM_COLLECTOR_TERMS))) return rc;
  if ((rc = alloc(totals, 2))) return rc;
  if ((rc = alloc(flags, 1))) return rc;
  if ((rc = alloc(d_sweep, cap_chains))) return rc;
  if ((rc = alloc(rec_acc, (size_t)cap_chains * R_COUNT))) return rc;
  ring_size = (int)std::max<long long>(16, std::min<long long>(4096, (64ll << 20) / ((long long)cap_chains * R_COUNT * 8)));
  if ((rc = alloc(rec_ring, (size_t)ring_size * cap_chains * R_COUNT))) return rc;
  // records reach the host through pinned memory: a copy to pageable memory
  // of a size not seen before costs the runtime milliseconds of staging
  if (rec_host) { hipHostFree(rec_host); rec_host = nullptr; }
  {
    const hipError_t e = hipHostMalloc((void**)&rec_host, (size_t)ring_size * cap_chains * R_COUNT * sizeof(double));
    if (e != hipSuccess) {
      rec_host = nullptr;
      return set_error(LG_ERR_DEVICE, std::string("hipHostMalloc of the record mirror: ") + hipGetErrorString(e));
    }
  }
  if ((rc = alloc(ring_pos, 1))) return rc;
  if ((rc = alloc(end_count, 1))) return rc;
  LG_HIP(hipMemset(end_count, 0, sizeof(int)));
  LG_HIP(hipMemset(rec_acc, 0, (size_t)cap_chains * R_COUNT * sizeof(double)));
  LG_HIP(hipMemset(flags, 0, sizeof(int)));
  LG_HIP(hipMemset(ring_pos, 0, sizeof(int)));
  LG_HIP(hipMemset(totals, 0, 2 * sizeof(int)));
  int dev = 0;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
  cfg.blocks_op = std::max(64, cus * 8);
  cfg.blocks_node = std::max(64, cus * 8);
  return LG_OK;
}

// New bases (contiguous in the given order), chain tables, sweep counters.
int Batch::set_layout(std::vector<ChainDev>& ds, const std::vector<uint64_t>& sweeps, bool check_old) {
  long long sites = 0, bonds = 0, cand = 0, strs = 0;
  for (size_t c = 0; c < ds.size(); ++c) {
    ChainDev& d = ds[c];
    d.site_base = (int)sites;
    d.bond_base = (int)bonds;
    d.cand_base = cand;
    if (d.sse) {  // the chain's operator string: its share of the operator capacity
      d.sse_base = (int)strs;
      d.sse_cap = (int)std::min<long long>(footprint(params[c]).ops, (long long)op_cap - strs);
      strs += d.sse_cap;
    }
    // candidate slots slot-major (slot j of bond b at j * Nb + b) when bonds
    // take few candidates, so that the bonds of a wave write adjacent words
    // (512-chain scan: diag_fill 0.124 -> 0.108 ms); bond-major when they
    // take many, so that a bond's slots share lines (square 128^2, beta=128,
    // about 180 slots per bond: slot-major cost diag_count 0.081 -> 0.084 ms)
    const bool slot_major = d.cand_cap <= 128;
    d.cand_bstride = slot_major ? 1 : d.cand_cap;
    d.cand_sstride = slot_major ? d.Nb : 1;
    sites += d.N;
    bonds += d.Nb;
    cand += (long long)d.Nb * d.cand_cap;
  }
  if ((int)ds.size() > cap_chains || sites > budget.sites || bonds > budget.bonds || cand > cand_total)
    return set_error(LG_ERR_CAPACITY, "chains exceed the batch capacity");
  if (check_old) {
    // every chain's previous configuration lies inside the extent of the
    // arrays it is read from, and no two chains read the same sites or bonds
    std::vector<std::pair<long long, long long>> sr, br;
    for (const auto& d : ds) {
      if (d.old_site_base < 0 || d.old_bond_base < 0 || d.old_site_base + (long long)d.N > old_ext_sites ||
          d.old_bond_base + (long long)d.Nb > old_ext_bonds)
        return set_error(LG_ERR_STATE, "relayout: a chain's previous configuration lies outside the arrays");
      sr.push_back({d.old_site_base, d.old_site_base + d.N});
      br.push_back({d.old_bond_base, d.old_bond_base + d.Nb});
    }
    for (auto* r : {&sr, &br}) {
      std::sort(r->begin(), r->end());
      for (size_t i = 1; i < r->size(); ++i)
        if ((*r)[i].first < (*r)[i - 1].second)
          return set_error(LG_ERR_STATE, "relayout: chains' previous configurations overlap");
    }
    if (sites > cap_sites || bonds > cap_bonds)
      return set_error(LG_ERR_STATE, "relayout: new layout beyond the site or bond arrays");
  }
  chains = ds;
  nsites = (int)sites;
  nbonds = (int)bonds;
  host_sweep = sweeps;
  std::vector<int> bc(std::max(nbonds, 1)), sc(std::max(nsites, 1));
  std::vector<uint32_t> sw(std::max<size_t>(chains.size(), 1));
  for (size_t c = 0; c < chains.size(); ++c) {
    for (int b = 0; b < chains[c].Nb; ++b) bc[chains[c].bond_base + b] = (int)c;
    for (int s = 0; s < chains[c].N; ++s) sc[chains[c].site_base + s] = (int)c;
    sw[c] = (uint32_t)sweeps[c];
  }
  if (!chains.empty()) {
    LG_HIP(hipMemcpy(d_chains, chains.data(), chains.size() * sizeof(ChainDev), hipMemcpyHostToDevice));
    LG_HIP(hipMemcpy(bond_chain, bc.data(), nbonds * sizeof(int), hipMemcpyHostToDevice));
    LG_HIP(hipMemcpy(site_chain, sc.data(), nsites * sizeof(int), hipMemcpyHostToDevice));
    LG_HIP(hipMemcpy(d_sweep, sw.data(), chains.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  double lam_max = 0;
  for (auto& d : chains)
    lam_max = std::max(lam_max, d.beta * std::max(std::max(d.has_weight ? d.wsum : 0.0, d.dwsum), d.site_w));
  int G = 4;
  while (G < 64 && G < lam_max) G *= 2;
  if (!tuned) cfg.group = G;
  // expected events per site (two legs per operator; operators per bond about
  // beta * (offset + half the largest diagonal energy)) picks the staging size
  double ev = 0, ops_bond = 0, nb_total = 0;
  for (size_t c = 0; c < chains.size(); ++c) {
    const ChainDev& d = chains[c];
    // (in-site bonds: counted as lattice bonds; site operators: beta |hx| per site)
    const double per_bond = (d.offset_total - d.N * d.site_w) / d.nb_real +
                            0.5 * (std::fabs(params[c].jz) / 4 + std::fabs(params[c].jxy) / 2);
    ev += 2.0 * d.beta * d.nb_real * per_bond + d.beta * d.N * 2 * d.site_w;
    ops_bond += d.nb_real * d.beta * per_bond + d.beta * d.N * 2 * d.site_w;
    nb_total += d.Nb;
  }
  LaunchCfg c = cfg;
  // the fill pass stages a bond's lists in LDS when they fit: room for the
  // expected operators per bond with a margin, no more (the LDS the block
  // holds limits its occupancy: square 512^2, beta=512, fill 4.05 -> 3.0 ms
  // with 512 times, while 512 would cost square 128^2 0.045 -> 0.061 ms).
  // Not a launch setting: results and the tuner's choices do not depend on it.
  // Sized for the bond-weighted mean, so that a few large jobs in a mixed
  // batch fall back to global memory instead of costing every bond
  // occupancy.
  if (nb_total > 0) ops_bond /= nb_total;
  c.fill_cap = 1.25 * ops_bond <= 256 ? 256 : (1.25 * ops_bond <= 512 ? 512 : 1024);
  bool split = false, many_lists = false;
  for (const auto& d : chains) {
    split = split || d.nv > 1;
    many_lists = many_lists || site_lists(d) > 4;
  }
  if (!tuned) {
    const double per_site = nsites > 0 ? ev / nsites : 0;
    c.link_cap = per_site < 48 ? 0 : (per_site < 96 ? 128 : (per_site <= 400 ? 512 : 1536));
    // virtual sites: only the per-operator and per-site kernels write their
    // separate top segments, and the per-site kernel merges four lists at most
    if (split) c.link_cap = (per_site < 8 || many_lists) ? 0 : 1;
    // the estimator's LDS table: hot configurations (many one- and
    // two-segment clusters, which overflow the table) flush it every 1024
    // segments (512-chain scan: estimate 0.172 -> 0.166 ms); cold ones once
    // per chunk (square 128^2, beta=128: 0.154 ms against 0.19 at 4 or 8)
    c.est_flush = per_site < 48 ? 4 : 16;
    // the estimator finds the roots itself (a pass over the segments less,
    // but the collector then clears every root's accumulator): with the
    // 1024-thread estimator it measured faster than the compress pass for one
    // chain at every size tried (square 128^2, beta=128: 0.660 against 0.678
    // ms; 64^2, beta=64: 0.222 against 0.226) and for the 512-chain scan
    c.compress = 0;
    // grid of the per-operator and per-node kernels: a finer split of large
    // configurations balances the blocks' uneven work (C5: 2048 -> 16384
    // blocks took the sweep from 41.0 to 38.1 ms; C2 is fastest at 2048)
    const int mult = ev < 16.0e6 ? 8 : (ev < 128.0e6 ? 32 : 64);
    // small configurations: about one block per 256 expected events, at
    // least 256 (blocks that find no work still cost their dispatch: chain
    // L=128, beta=100 went 0.090 -> 0.079 ms/sweep, L=64, beta=10 0.077 ->
    // 0.062, square 32^2, beta=8 0.105 -> 0.091)
    const double want = std::max(256.0, ev / 256.0);
    c.blocks_op = c.blocks_node = std::max(64, (int)std::min<double>(cus * mult, want));
  }
  // tuned settings meet chains they were not tuned for when a server admits
  // jobs: the LDS-staged linking kernels need wrapping segments, the
  // per-site one at most four lists per site
  if (split && c.link_cap > 1) c.link_cap = 0;
  if (many_lists && c.link_cap == 1) c.link_cap = 0;
  if (has_extra_bonds()) c.link_cap = 0;  // in-site bonds, site operators: the per-operator kernel
  return apply_cfg(c);
}

int Batch::apply_cfg(const LaunchCfg& c) {
  // the compress pass clears the roots' accumulators before the estimator
  // adds into them; without it they must be zero between sweeps, which the
  // collector keeps once they are
  if (cfg.compress && !c.compress && acc)
    LG_HIP(hipMemsetAsync(acc, 0, (size_t)node_cap * sizeof(ClusterAcc), stream));
  cfg = c;
  invalidate_graphs();
  return LG_OK;
}

// all-up spins, empty operator string (path_integral.C:225) for every chain
int Batch::reset_state() {
  cur = 0;
  State& s = st[cur];
  std::vector<int> zeros(nbonds + 1, 0), soff(nsites + 1);
  for (int i = 0; i <= nsites; ++i) soff[i] = i;
  std::vector<double> evt(nsites, EMPTY_TIME);
  LG_HIP(hipMemset(s.spins, 0, nsites));
  LG_HIP(hipMemcpy(s.bo, zeros.data(), (nbonds + 1) * sizeof(int), hipMemcpyHostToDevice));
  LG_HIP(hipMemcpy(s.soff, soff.data(), (nsites + 1) * sizeof(int), hipMemcpyHostToDevice));
  LG_HIP(hipMemcpy(s.evt, evt.data(), nsites * sizeof(double), hipMemcpyHostToDevice));
  LG_HIP(hipMemset(s.segspin, 0, nsites));
  for (auto& d : chains) { d.old_site_base = d.site_base; d.old_bond_base = d.bond_base; }
  LG_HIP(hipMemcpy(d_chains, chains.data(), chains.size() * sizeof(ChainDev), hipMemcpyHostToDevice));
  std::vector<uint32_t> sw(chains.size(), 0);
  LG_HIP(hipMemcpy(d_sweep, sw.data(), chains.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  host_sweep.assign(chains.size(), 0);
  old_ext_sites = nsites;
  old_ext_bonds = nbonds;
  old_ext_nodes = nsites;
  invalidate_graphs();
  return LG_OK;
}

// Change the set of chains between sweeps: chains keep[i] continue from their
// current configuration (read at their present bases), chains `add` start
// fresh (all-up spins, no operators), appended to the previous configuration
// after its current extent.  Nothing is copied on the device; the next sweep
// reads each chain at its old bases and writes the new layout.
int Batch::relayout(const std::vector<int>& keep, const std::vector<lg_params>& add) {
  if (is_sse()) return set_error(LG_ERR_UNSUPPORTED, "relayout of SSE chains is not built");
  for (const auto& p : add)
    if (p.algorithm != LG_ALGORITHM_PATH_INTEGRAL) return set_error(LG_ERR_UNSUPPORTED, "SSE chains are not served");
  std::vector<ChainDev> ds;
  std::vector<uint64_t> sw;
  std::vector<lg_params> ps;
  for (int i : keep) {
    if (i < 0 || i >= (int)chains.size()) return set_error(LG_ERR_ARG, "relayout: bad chain index");
    ChainDev d = chains[i];
    if (swept_since_layout) { d.old_site_base = d.site_base; d.old_bond_base = d.bond_base; }
    ds.push_back(d);
    sw.push_back(host_sweep[i]);
    ps.push_back(params[i]);
  }
  if (swept_since_layout) {
    old_ext_sites = nsites;
    old_ext_bonds = nbonds;
    int m = 0;
    if (nsites > 0) LG_HIP(hipMemcpy(&m, st[cur].soff + nsites, sizeof(int), hipMemcpyDeviceToHost));
    old_ext_nodes = m;
  }
  if (!add.empty()) {
    State& s = st[cur];
    int b0 = 0;
    LG_HIP(hipMemcpy(&b0, s.bo + old_ext_bonds, sizeof(int), hipMemcpyDeviceToHost));
    long long addN = 0, addNb = 0;
    for (const auto& p : add) {
      int rc = check_params(p);
      if (rc) return rc;
      ChainDev d = make_chain_dev(p);
      addN += d.N;
      addNb += d.Nb;
    }
    if (old_ext_sites + addN > cap_sites || old_ext_bonds + addNb > cap_bonds ||
        old_ext_nodes + addN > node_cap)
      return set_error(LG_ERR_CAPACITY, "relayout: no room to append chains");
    std::vector<int> bo(addNb, b0), soff(addN);
    for (long long k = 0; k < addN; ++k) soff[k] = (int)(old_ext_nodes + k + 1);
    std::vector<double> evt(addN, EMPTY_TIME);
    LG_HIP(hipMemcpy(s.bo + old_ext_bonds + 1, bo.data(), addNb * sizeof(int), hipMemcpyHostToDevice));
    LG_HIP(hipMemcpy(s.soff + old_ext_sites + 1, soff.data(), addN * sizeof(int), hipMemcpyHostToDevice));
    LG_HIP(hipMemcpy(s.evt + old_ext_nodes, evt.data(), addN * sizeof(double), hipMemcpyHostToDevice));
    LG_HIP(hipMemset(s.segspin + old_ext_nodes, 0, addN));
    LG_HIP(hipMemset(s.spins + old_ext_sites, 0, addN));
    for (const auto& p : add) {
      ChainDev d = make_chain_dev(p);
      d.old_site_base = old_ext_sites;
      d.old_bond_base = old_ext_bonds;
      old_ext_sites += d.N;
      old_ext_bonds += d.Nb;
      old_ext_nodes += d.N;
      ds.push_back(d);
      sw.push_back(0);
      ps.push_back(p);
    }
  }
  params = ps;
  swept_since_layout = false;
  return set_layout(ds, sw, true);
}

void Batch::invalidate_graphs() {
  for (int i = 0; i < 2; ++i)
    if (graph_exec[i]) { hipGraphExecDestroy(graph_exec[i]); graph_exec[i] = nullptr; }
}

DevBatch Batch::dev(int from) const {
  const State& o = st[from];
  const State& w = st[1 - from];
  DevBatch B;
  B.nchains = (int)chains.size();
  B.nsites = nsites;
  B.nbonds = nbonds;
  B.chains = d_chains;
  B.bond_chain = bond_chain;
  B.site_chain = site_chain;
  B.spins_old = o.spins; B.bo_old = o.bo; B.ot_old = o.ot; B.oi_old = o.oi;
  B.soff_old = o.soff; B.evt_old = o.evt; B.segspin_old = o.segspin;
  B.spins_new = w.spins; B.bo_new = w.bo; B.ot_new = w.ot; B.oi_new = w.oi;
  B.soff_new = w.soff; B.evt_new = w.evt; B.segspin_new = w.segspin;
  B.counts = counts; B.acc_count = acc_count; B.kpre = kpre;
  B.cand_t = cand_t; B.cand_g = cand_g; B.kc_new = kc_new;
  B.l0 = l0; B.l1 = l1; B.u0 = u0; B.u1 = u1; B.evflags = evflags; B.chain_node0 = chain_node0; B.node_site = node_site;
  B.parent = parent; B.fbit = fbit; B.fmask = fmask; B.rmask = rmask; B.acc = acc;
  B.scan_status = scan_status; B.scan_ctl = scan_ctl; B.col_part = col_part; B.totals = totals; B.flags = flags;
  B.sweep = d_sweep; B.rec_acc = rec_acc; B.rec_ring = rec_ring;
  B.ring_pos = ring_pos; B.end_count = end_count; B.ring_size = ring_size;
  B.any_field = 0;
  B.any_split = 0;
  B.any_extra = 0;
  for (const auto& d : chains) {
    if (d.h != 0.0) B.any_field = 1;
    if (d.nv > 1) B.any_split = 1;
    if (d.nb_real > d.nb_lat) B.any_extra |= 1;
    if (d.Nb > d.nb_real) B.any_extra |= 2;
  }
  B.nslots = nslots; B.xslot = xslot; B.xbeta = xbeta; B.xscale = xscale; B.xlnr = xlnr;
  B.xacc = xacc; B.xstep = xstep; B.xkey0 = xkey0; B.xkey1 = xkey1;
  B.op_cap = op_cap; B.node_cap = node_cap;
  B.any_sse = is_sse() ? 1 : 0;
  B.sse_old = sse_str[from];
  B.sse_new = sse_str[1 - from];
  B.sse_perm = kpre;  // (the path-integral diagonal update's, unused by SSE batches)
  B.sched = d_sched;
  return B;
}

int Batch::capture(int from) {
  const auto t0 = std::chrono::steady_clock::now();
  hipGraph_t g = nullptr;
  LG_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  launch_sweep(dev(from), cfg, stream);
  LG_HIP(hipStreamEndCapture(stream, &g));
  hipError_t e = hipGraphInstantiate(&graph_exec[from], g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return set_error(LG_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  ++graph_captures;
  graph_capture_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return LG_OK;
}

int Batch::check_flags() {
  int f = 0;
  LG_HIP(hipMemcpy(&f, flags, sizeof(int), hipMemcpyDeviceToHost));
  if (f) {
    std::string m = "device capacity exceeded:";
    if (f & FLAG_CAND_OVERFLOW) m += " candidate slots";
    if (f & FLAG_OP_OVERFLOW) m += " operators";
    if (f & FLAG_NODE_OVERFLOW) m += " segments";
    if (f & FLAG_RING_OVERFLOW) m += " record ring";
    return set_error(LG_ERR_CAPACITY, m);
  }
  return LG_OK;
}

int Batch::enqueue_sweep() {
  if (use_graph) {
    if (!graph_exec[cur]) {
      int rc = capture(cur);
      if (rc) return rc;
    }
    LG_HIP(hipGraphLaunch(graph_exec[cur], stream));
  } else {
    launch_sweep(dev(cur), cfg, stream);
    LG_HIP(hipGetLastError());
  }
  cur ^= 1;
  note_swept();
  return LG_OK;
}

// k_end_sweep moves every descriptor's old bases to its new bases; keep the
// host copy the same.
void Batch::note_swept() {
  if (!swept_since_layout)
    for (auto& d : chains) { d.old_site_base = d.site_base; d.old_bond_base = d.bond_base; }
  swept_since_layout = true;
}

int Batch::run(uint32_t n, lg_sweep_record* out) {
  int rc = check_flags();
  if (rc) return rc;
  const size_t nc = chains.size();
  if (nc == 0) return LG_OK;
  uint32_t done = 0;
  while (done < n) {
    uint32_t chunk = std::min<uint32_t>(n - done, (uint32_t)ring_size);
    LG_HIP(hipMemsetAsync(ring_pos, 0, sizeof(int), stream));
    for (uint32_t i = 0; i < chunk; ++i) {
      rc = enqueue_sweep();
      if (rc) return rc;
    }
    const size_t words = (size_t)chunk * nc * R_COUNT;
    if (out) LG_HIP(hipMemcpyAsync(rec_host, rec_ring, words * sizeof(double), hipMemcpyDeviceToHost, stream));
    LG_HIP(hipStreamSynchronize(stream));
    rc = check_flags();
    if (rc) return rc;
    if (out) {
      static_assert(sizeof(lg_sweep_record) == R_COUNT * sizeof(double), "record layout");
      std::memcpy(out + (size_t)done * nc, rec_host, words * sizeof(double));
    }
    for (size_t c = 0; c < nc; ++c) host_sweep[c] += chunk;
    done += chunk;
  }
  return LG_OK;
}

// Autotuner: time candidate lane-group sizes of the diagonal update, grid
// sizes of the per-operator / per-node kernels and the linking kernels on the
// live configuration.
int Batch::tune(uint32_t sweeps_per_trial) {
  if (sweeps_per_trial == 0) sweeps_per_trial = 8;
  int rc = run(sweeps_per_trial, nullptr);  // warm state and code
  if (rc) return rc;
  LaunchCfg best = cfg;
  double best_t = 1e300;
  const int groups[] = {1, 2, 4, 8, 16, 32, 64};
  const int mults[] = {2, 4, 8, 16, 32, 64};
  const int links[] = {0, 1, 128, 512, 1536};
  const int flushes[] = {16, 8, 4};
  // group size first with the current grids, then the grids with the best
  // group, then the linking kernel, then the compress pass, then the
  // estimator's flush interval
  const int trials[] = {7, 6, 5, 2, 3};
  for (int pass = 0; pass < 5; ++pass) {
    for (int i = 0; i < trials[pass]; ++i) {
      LaunchCfg trial = best;
      if (pass == 0) trial.group = groups[i];
      else if (pass == 1) trial.blocks_op = trial.blocks_node = cus * mults[i];
      else if (pass == 2) {
        if (links[i] > 1 && has_virtual_sites()) continue;  // these kernels assume wrapping segments
        if (links[i] == 1 && max_site_lists() > 4) continue;  // it merges four lists at most
        if (links[i] != 0 && has_extra_bonds()) continue;     // in-site bonds, site operators: the per-operator kernel
        trial.link_cap = links[i];
      }
      else if (pass == 3) trial.compress = i;
      else trial.est_flush = flushes[i];
      rc = apply_cfg(trial);
      if (rc) return rc;
      rc = synchronize();
      if (rc) return rc;
      auto t0 = std::chrono::steady_clock::now();
      rc = run(sweeps_per_trial, nullptr);
      if (rc) return rc;
      double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (t < best_t) { best_t = t; best = trial; }
    }
  }
  tuned = true;
  return apply_cfg(best);
}

void Batch::get_launch(lg_launch* out) const {
  out->group = cfg.group;
  out->blocks = cfg.blocks_op;
  out->link = cfg.link_cap;
  out->compress = cfg.compress;
  out->flush = cfg.est_flush;
}

int Batch::set_launch(const lg_launch& l) {
  LaunchCfg c = cfg;
  if (l.group >= 0) {
    if (l.group != 1 && l.group != 2 && l.group != 4 && l.group != 8 && l.group != 16 && l.group != 32 &&
        l.group != 64)
      return set_error(LG_ERR_ARG, "launch group must be 1, 2, 4, 8, 16, 32 or 64");
    c.group = l.group;
  }
  if (l.blocks >= 0) {
    if (l.blocks < 1 || l.blocks > (1 << 20)) return set_error(LG_ERR_ARG, "launch blocks must be in [1, 2^20]");
    c.blocks_op = c.blocks_node = l.blocks;
  }
  if (l.link >= 0) {
    if (l.link != 0 && l.link != 1 && l.link != 128 && l.link != 512 && l.link != 1536)
      return set_error(LG_ERR_ARG, "launch link must be 0, 1, 128, 512 or 1536");
    if (l.link > 1 && has_virtual_sites())
      return set_error(LG_ERR_ARG, "chains with virtual sites (S > 1/2) take linking kernel 0 or 1");
    if (l.link == 1 && max_site_lists() > 4)
      return set_error(LG_ERR_ARG, "sites with more than four bond lists (S > 1 on the chain, S > 1/2 on the square lattice) take linking kernel 0");
    if (l.link != 0 && has_extra_bonds())
      return set_error(LG_ERR_ARG, "chains with a transverse field or a single-ion term take linking kernel 0");
    c.link_cap = l.link;
  }
  if (l.compress >= 0) {
    if (l.compress > 1) return set_error(LG_ERR_ARG, "launch compress must be 0 or 1");
    c.compress = l.compress;
  }
  if (l.flush >= 0) {
    if (l.flush != 4 && l.flush != 8 && l.flush != 16) return set_error(LG_ERR_ARG, "launch flush must be 4, 8 or 16");
    c.est_flush = l.flush;
  }
  tuned = true;
  return apply_cfg(c);
}

// n sweeps with HIP events between the kernel launches on the batch's stream;
// ms[i] = average duration of kernel i per sweep (kKernelNames order).
int Batch::profile(uint32_t n, double* ms) {
  int rc = check_flags();
  if (rc) return rc;
  const int K = NUM_SWEEP_KERNELS;
  std::vector<hipEvent_t> ev((size_t)(K + 1) * std::max<uint32_t>(n, 1));
  for (auto& e : ev) LG_HIP(hipEventCreate(&e));
  for (int i = 0; i < K; ++i) ms[i] = 0;
  uint32_t done = 0;
  while (done < n) {
    uint32_t chunk = std::min<uint32_t>(n - done, (uint32_t)ring_size);
    LG_HIP(hipMemsetAsync(ring_pos, 0, sizeof(int), stream));
    for (uint32_t i = 0; i < chunk; ++i) {
      launch_sweep(dev(cur), cfg, stream, &ev[(size_t)(done + i) * (K + 1)]);
      cur ^= 1;
      note_swept();
    }
    LG_HIP(hipStreamSynchronize(stream));
    for (uint32_t i = 0; i < chunk; ++i)
      for (int k = 0; k < K; ++k) {
        float t = 0;
        LG_HIP(hipEventElapsedTime(&t, ev[(size_t)(done + i) * (K + 1) + k], ev[(size_t)(done + i) * (K + 1) + k + 1]));
        ms[k] += t;
      }
    done += chunk;
    for (size_t c = 0; c < chains.size(); ++c) host_sweep[c] += chunk;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (int i = 0; i < K; ++i) ms[i] /= std::max<uint32_t>(n, 1);
  return check_flags();
}

int Batch::synchronize() {
  LG_HIP(hipStreamSynchronize(stream));
  return LG_OK;
}

// ---- configuration transfer (single-chain helpers) ------------------------
// Builds the per-site world-line arrays of a configuration on the host:
// events ordered by (time, bond, rank), segment offsets max(m,1), event times
// and the spin below each event.
void host_site_arrays(const ChainDev& d, const std::vector<int>& spins, const std::vector<int>& bo,
                      const std::vector<double>& ot, const std::vector<int>& offd,
                      std::vector<int>& soff, std::vector<double>& evt, std::vector<uint8_t>& segspin) {
  const int N = d.N;
  std::vector<std::vector<int>> ev(N);
  for (int b = 0; b < d.Nb; ++b) {
    int s0, s1;
    op_sites(d, b, s0, s1);
    for (int k = bo[b]; k < bo[b + 1]; ++k) {
      ev[s0].push_back(k);
      if (s1 != s0 || b < d.nb_real) ev[s1].push_back(k);  // a site operator is one event
    }
  }
  std::vector<int> bond_of(ot.size());
  for (int b = 0; b < d.Nb; ++b) for (int k = bo[b]; k < bo[b + 1]; ++k) bond_of[k] = b;
  soff.assign(N + 1, 0);
  const bool split = d.nv > 1;  // virtual sites: m events, m + 1 segments, the top one ending at tau = 1
  for (int s = 0; s < N; ++s) {
    std::sort(ev[s].begin(), ev[s].end(), [&](int a, int b) {
      if (ot[a] != ot[b]) return ot[a] < ot[b];
      if (bond_of[a] != bond_of[b]) return bond_of[a] < bond_of[b];
      return a < b;
    });
    const int m = (int)ev[s].size();
    soff[s + 1] = soff[s] + (m == 0 ? 1 : (split ? m + 1 : m));
  }
  evt.assign(soff[N], EMPTY_TIME);
  segspin.assign(soff[N], 0);
  for (int s = 0; s < N; ++s) {
    int c = spins[s];
    if (ev[s].empty()) { segspin[soff[s]] = (uint8_t)c; continue; }
    for (size_t j = 0; j < ev[s].size(); ++j) {
      int k = ev[s][j];
      evt[soff[s] + j] = ot[k];
      segspin[soff[s] + j] = (uint8_t)c;
      c ^= offd[k];
    }
    if (split) {
      evt[soff[s + 1] - 1] = 1.0;
      segspin[soff[s + 1] - 1] = (uint8_t)c;
    }
  }
}

int validate_config(const ChainDev& d, const int32_t* bond_offsets, const double* times,
                    const int32_t* types, long long& nops) {
  if (bond_offsets[0] != 0) return set_error(LG_ERR_ARG, "bond_offsets[0] must be 0");
  for (int b = 0; b < d.Nb; ++b)
    if (bond_offsets[b + 1] < bond_offsets[b]) return set_error(LG_ERR_ARG, "bond_offsets not monotone");
  nops = bond_offsets[d.Nb];
  for (int b = 0; b < d.Nb; ++b)
    for (int k = bond_offsets[b]; k < bond_offsets[b + 1]; ++k) {
      if (!(times[k] >= 0 && times[k] < 1)) return set_error(LG_ERR_ARG, "operator time outside [0,1)");
      if (k > bond_offsets[b] && times[k] < times[k - 1])
        return set_error(LG_ERR_ARG, "operator times not ascending within a bond");
      if (types[k] < 0 || (types[k] >> 2) > 3) return set_error(LG_ERR_ARG, "invalid operator type");
      if (b >= d.nb_real && (types[k] >> 2) != 0)
        return set_error(LG_ERR_ARG, "a site operator (transverse field) takes graph 0");
    }
  return LG_OK;
}

int Batch::set_config(int c, const int32_t* spins_in, const int32_t* bond_offsets, const double* times,
                      const int32_t* types, uint64_t sweep) {
  if (chains.size() != 1) return set_error(LG_ERR_UNSUPPORTED, "set_config needs a single-chain handle");
  const ChainDev& d = chains[c];
  long long nops = 0;
  int rc = validate_config(d, bond_offsets, times, types, nops);
  if (rc) return rc;
  if (nops > op_cap) return set_error(LG_ERR_CAPACITY, "configuration exceeds operator capacity");
  std::vector<int> spins(spins_in, spins_in + d.N), bo(bond_offsets, bond_offsets + d.Nb + 1), offd(nops);
  std::vector<double> ot(times, times + nops);
  std::vector<int> oi(nops);
  for (int b = 0; b < d.Nb; ++b)
    for (int k = bo[b]; k < bo[b + 1]; ++k) {
      offd[k] = types[k] & 1;
      oi[k] = ((d.bond_base + b) << 3) | ((types[k] >> 2) << 1) | (types[k] & 1);
    }
  std::vector<int> soff;
  std::vector<double> evt;
  std::vector<uint8_t> segspin;
  host_site_arrays(d, spins, bo, ot, offd, soff, evt, segspin);
  if ((long long)soff[d.N] > node_cap) return set_error(LG_ERR_CAPACITY, "configuration exceeds segment capacity");
  std::vector<uint8_t> sp(d.N);
  for (int s = 0; s < d.N; ++s) sp[s] = (uint8_t)(spins[s] & 1);
  State& s = st[cur];
  LG_HIP(hipMemcpy(s.spins, sp.data(), d.N, hipMemcpyHostToDevice));
  LG_HIP(hipMemcpy(s.bo, bo.data(), (d.Nb + 1) * sizeof(int), hipMemcpyHostToDevice));
  if (nops) {
    LG_HIP(hipMemcpy(s.ot, ot.data(), nops * sizeof(double), hipMemcpyHostToDevice));
    LG_HIP(hipMemcpy(s.oi, oi.data(), nops * sizeof(int), hipMemcpyHostToDevice));
  }
  LG_HIP(hipMemcpy(s.soff, soff.data(), (d.N + 1) * sizeof(int), hipMemcpyHostToDevice));
  LG_HIP(hipMemcpy(s.evt, evt.data(), evt.size() * sizeof(double), hipMemcpyHostToDevice));
  LG_HIP(hipMemcpy(s.segspin, segspin.data(), segspin.size(), hipMemcpyHostToDevice));
  if (d.sse) {  // SSE: the string in series order, which the times keep (ties by bond)
    std::vector<int> bond_of(nops), order(nops), str(nops);
    for (int b = 0; b < d.Nb; ++b)
      for (int k = bo[b]; k < bo[b + 1]; ++k) bond_of[k] = b;
    for (long long k = 0; k < nops; ++k) order[k] = (int)k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      if (ot[a] != ot[b]) return ot[a] < ot[b];
      return bond_of[a] < bond_of[b];
    });
    if (nops > d.sse_cap) return set_error(LG_ERR_CAPACITY, "configuration exceeds the operator string capacity");
    for (long long p = 0; p < nops; ++p) {
      const int k = order[p];
      str[p] = (bond_of[k] << 3) | ((types[k] >> 2) << 1) | (types[k] & 1);
    }
    if (nops) LG_HIP(hipMemcpy(sse_str[cur] + d.sse_base, str.data(), nops * sizeof(int), hipMemcpyHostToDevice));
  }
  uint32_t sw = (uint32_t)sweep;
  LG_HIP(hipMemcpy(d_sweep, &sw, sizeof(uint32_t), hipMemcpyHostToDevice));
  host_sweep[c] = sweep;
  return sync_schedule();
}

int Batch::get_config(int c, int32_t* spins, int32_t* bond_offsets, double* times, int32_t* types) {
  if (chains.size() != 1) return set_error(LG_ERR_UNSUPPORTED, "get_config needs a single-chain handle");
  const ChainDev& d = chains[c];
  const State& s = st[cur];
  std::vector<uint8_t> sp(d.N);
  std::vector<int> bo(d.Nb + 1);
  LG_HIP(hipMemcpy(sp.data(), s.spins, d.N, hipMemcpyDeviceToHost));
  LG_HIP(hipMemcpy(bo.data(), s.bo, (d.Nb + 1) * sizeof(int), hipMemcpyDeviceToHost));
  int nops = bo[d.Nb] - bo[0];
  std::vector<int> oi(nops);
  if (spins) for (int i = 0; i < d.N; ++i) spins[i] = sp[i];
  if (bond_offsets) for (int b = 0; b <= d.Nb; ++b) bond_offsets[b] = bo[b] - bo[0];
  if (nops && times) LG_HIP(hipMemcpy(times, s.ot + bo[0], nops * sizeof(double), hipMemcpyDeviceToHost));
  if (nops && types) {
    LG_HIP(hipMemcpy(oi.data(), s.oi + bo[0], nops * sizeof(int), hipMemcpyDeviceToHost));
    for (int k = 0; k < nops; ++k) types[k] = (oi[k] & 1) | (((oi[k] >> 1) & 3) << 2);
  }
  return LG_OK;
}

int Batch::num_ops(int c, long long* n) {
  const ChainDev& d = chains[c];
  int a = 0, b = 0;
  LG_HIP(hipMemcpy(&a, st[cur].bo + d.bond_base, sizeof(int), hipMemcpyDeviceToHost));
  LG_HIP(hipMemcpy(&b, st[cur].bo + d.bond_base + d.Nb, sizeof(int), hipMemcpyDeviceToHost));
  *n = b - a;
  return LG_OK;
}

int Batch::chain_config(int c, std::vector<int>& spins, std::vector<int>& bo, std::vector<double>& times,
                        std::vector<int>& types) {
  if (c < 0 || c >= (int)chains.size()) return set_error(LG_ERR_ARG, "chain index out of range");
  if (!swept_since_layout && (chains[c].old_site_base != chains[c].site_base ||
                              chains[c].old_bond_base != chains[c].bond_base))
    return set_error(LG_ERR_STATE, "chain moved since its last sweep");
  LG_HIP(hipStreamSynchronize(stream));
  const ChainDev& d = chains[c];
  const State& s = st[cur];
  std::vector<uint8_t> sp(d.N);
  bo.assign(d.Nb + 1, 0);
  LG_HIP(hipMemcpy(sp.data(), s.spins + d.site_base, d.N, hipMemcpyDeviceToHost));
  LG_HIP(hipMemcpy(bo.data(), s.bo + d.bond_base, (d.Nb + 1) * sizeof(int), hipMemcpyDeviceToHost));
  const int op0 = bo[0], nops = bo[d.Nb] - bo[0];
  for (auto& x : bo) x -= op0;
  spins.assign(sp.begin(), sp.end());
  times.assign(nops, 0.0);
  types.assign(nops, 0);
  std::vector<int> oi(nops);
  if (nops) {
    LG_HIP(hipMemcpy(times.data(), s.ot + op0, nops * sizeof(double), hipMemcpyDeviceToHost));
    LG_HIP(hipMemcpy(oi.data(), s.oi + op0, nops * sizeof(int), hipMemcpyDeviceToHost));
  }
  for (int k = 0; k < nops; ++k) types[k] = (oi[k] & 1) | (((oi[k] >> 1) & 3) << 2);
  return LG_OK;
}

int Batch::load_configs(const std::vector<HostConfig>& cfg) {
  if (cfg.size() != chains.size()) return set_error(LG_ERR_ARG, "one configuration per chain");
  std::vector<uint8_t> sp(std::max(nsites, 1));
  std::vector<int> bo(nbonds + 1), soff(nsites + 1);
  std::vector<double> ot, evt;
  std::vector<int> oi;
  std::vector<uint8_t> segspin;
  long long op_base = 0, node_base = 0;
  for (size_t c = 0; c < chains.size(); ++c) {
    const ChainDev& d = chains[c];
    const HostConfig& h = cfg[c];
    long long nops = 0;
    if ((int)h.spins.size() != d.N || (int)h.bo.size() != d.Nb + 1)
      return set_error(LG_ERR_ARG, "configuration does not match the chain's lattice");
    int rc = validate_config(d, h.bo.data(), h.times.data(), h.types.data(), nops);
    if (rc) return rc;
    if (op_base + nops > op_cap) return set_error(LG_ERR_CAPACITY, "configurations exceed operator capacity");
    std::vector<int> offd(nops);
    for (int b = 0; b < d.Nb; ++b) {
      bo[d.bond_base + b] = (int)(op_base + h.bo[b]);
      for (int k = h.bo[b]; k < h.bo[b + 1]; ++k) {
        offd[k] = h.types[k] & 1;
        oi.push_back(((d.bond_base + b) << 3) | ((h.types[k] >> 2) << 1) | (h.types[k] & 1));
        ot.push_back(h.times[k]);
      }
    }
    std::vector<double> lt(h.times.begin(), h.times.begin() + nops);
    std::vector<int> ls;
    std::vector<double> le;
    std::vector<uint8_t> lsg;
    host_site_arrays(d, h.spins, h.bo, lt, offd, ls, le, lsg);
    for (int i = 0; i < d.N; ++i) {
      soff[d.site_base + i] = (int)(node_base + ls[i]);
      sp[d.site_base + i] = (uint8_t)(h.spins[i] & 1);
    }
    evt.insert(evt.end(), le.begin(), le.end());
    segspin.insert(segspin.end(), lsg.begin(), lsg.end());
    op_base += nops;
    node_base += ls[d.N];
    if (node_base > node_cap) return set_error(LG_ERR_CAPACITY, "configurations exceed segment capacity");
    host_sweep[c] = h.sweep;
  }
  bo[nbonds] = (int)op_base;
  soff[nsites] = (int)node_base;
  State& s = st[cur];
  LG_HIP(hipMemcpy(s.spins, sp.data(), nsites, hipMemcpyHostToDevice));
  LG_HIP(hipMemcpy(s.bo, bo.data(), (nbonds + 1) * sizeof(int), hipMemcpyHostToDevice));
  if (op_base) {
    LG_HIP(hipMemcpy(s.ot, ot.data(), op_base * sizeof(double), hipMemcpyHostToDevice));
    LG_HIP(hipMemcpy(s.oi, oi.data(), op_base * sizeof(int), hipMemcpyHostToDevice));
  }
  LG_HIP(hipMemcpy(s.soff, soff.data(), (nsites + 1) * sizeof(int), hipMemcpyHostToDevice));
  LG_HIP(hipMemcpy(s.evt, evt.data(), node_base * sizeof(double), hipMemcpyHostToDevice));
  LG_HIP(hipMemcpy(s.segspin, segspin.data(), node_base, hipMemcpyHostToDevice));
  std::vector<uint32_t> sw(chains.size());
  for (size_t c = 0; c < chains.size(); ++c) sw[c] = (uint32_t)host_sweep[c];
  LG_HIP(hipMemcpy(d_sweep, sw.data(), sw.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  old_ext_sites = nsites;
  old_ext_bonds = nbonds;
  old_ext_nodes = node_base;
  return sync_schedule();
}

int Batch::reload(const std::vector<lg_params>& ps, const std::vector<HostConfig>& cfg) {
  if (ps.size() != cfg.size()) return set_error(LG_ERR_ARG, "reload: one configuration per chain");
  std::vector<ChainDev> ds;
  std::vector<uint64_t> sw;
  for (size_t c = 0; c < ps.size(); ++c) {
    int rc = check_params(ps[c]);
    if (rc) return rc;
    ds.push_back(make_chain_dev(ps[c]));
    sw.push_back(cfg[c].sweep);
  }
  LG_HIP(hipStreamSynchronize(stream));
  params = ps;
  int rc = set_layout(ds, sw);
  if (rc) return rc;
  // the data is uploaded at the new bases: nothing reads the old ones
  for (auto& d : chains) { d.old_site_base = d.site_base; d.old_bond_base = d.bond_base; }
  if (!chains.empty())
    LG_HIP(hipMemcpy(d_chains, chains.data(), chains.size() * sizeof(ChainDev), hipMemcpyHostToDevice));
  if ((rc = load_configs(cfg))) return rc;
  swept_since_layout = true;
  invalidate_graphs();
  return LG_OK;
}

void Batch::adopt(Batch& o) {
  release();
  const uint64_t captures = graph_captures;
  const double capture_s = graph_capture_s;
  *this = o;  // member-wise: device pointers, host tables, settings
  o.rec_host = nullptr;
  graph_captures += captures;  // the counters cover the handle's lifetime
  graph_capture_s += capture_s;
  o.allocations.clear();
  o.graph_exec[0] = o.graph_exec[1] = nullptr;
  o.stream = nullptr;
}

int Batch::sync_schedule() {
  bool any = false;
  for (size_t c = 0; c < chains.size(); ++c) {
    ChainDev& d = chains[c];
    if (d.sched_n <= 0) continue;
    d.beta = 1.0 / schedule_temperature(sched_host.data() + 2 * d.sched_off, d.sched_n, (int)host_sweep[c]);
    d.scale = candidate_scale(d.beta, d.wsum);
    d.site_scale = candidate_scale(d.beta, d.site_w);
    d.dscale = candidate_scale(d.beta, d.dwsum);
    any = true;
  }
  if (!any) return LG_OK;
  LG_HIP(hipStreamSynchronize(stream));
  LG_HIP(hipMemcpy(d_chains, chains.data(), chains.size() * sizeof(ChainDev), hipMemcpyHostToDevice));
  invalidate_graphs();
  return LG_OK;
}

int Batch::set_schedule(int n, const int32_t* steps, const double* temps) {
  if (is_sse()) return set_error(LG_ERR_UNSUPPORTED, "annealing of SSE chains is not built");
  if (chains.size() != 1) return set_error(LG_ERR_UNSUPPORTED, "an annealing schedule needs a single-chain handle");
  if (n < 0 || (n > 0 && (!steps || !temps))) return set_error(LG_ERR_ARG, "invalid schedule");
  const lg_params p = params[0];
  double bmax = p.beta;
  for (int k = 0; k < n; ++k) {
    if (steps[k] < 1 || (k > 0 && steps[k] <= steps[k - 1]))
      return set_error(LG_ERR_ARG, "schedule steps must be ascending and >= 1");
    if (!(temps[k] > 0) || !std::isfinite(temps[k])) return set_error(LG_ERR_ARG, "schedule temperatures must be positive");
    bmax = std::max(bmax, 1.0 / temps[k]);
  }
  if (n > 0 && 1.0 / temps[n - 1] != p.beta)
    return set_error(LG_ERR_ARG, "the schedule's last temperature must be the chain's (final) temperature");
  if (bmax > cap_beta) {
    // capacity (candidate slots, operators) for the largest beta of the
    // schedule; the configuration and sweep count carry over
    std::vector<HostConfig> cfgs(1);
    int rc = chain_config(0, cfgs[0].spins, cfgs[0].bo, cfgs[0].times, cfgs[0].types);
    if (rc) return rc;
    cfgs[0].sweep = host_sweep[0];
    lg_params q = p;
    q.beta = bmax;
    Batch nb;
    nb.use_graph = use_graph;
    if ((rc = nb.init(&q, 1))) return rc;
    nb.params[0] = p;
    if ((rc = nb.load_configs(cfgs))) return rc;
    adopt(nb);
  }
  sched_host.clear();
  for (int k = 0; k < n; ++k) { sched_host.push_back((double)steps[k]); sched_host.push_back(temps[k]); }
  if (n > 0) {
    if (d_sched) { (void)hipFree(d_sched); allocations.erase(std::find(allocations.begin(), allocations.end(), (void*)d_sched)); d_sched = nullptr; }
    int rc = alloc(d_sched, sched_host.size());
    if (rc) return rc;
    LG_HIP(hipMemcpy(d_sched, sched_host.data(), sched_host.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  chains[0].sched_n = n;
  chains[0].sched_off = 0;
  if (n == 0) {  // fixed temperature again
    chains[0].beta = p.beta;
    chains[0].scale = candidate_scale(p.beta, chains[0].wsum);
    chains[0].site_scale = candidate_scale(p.beta, chains[0].site_w);
    chains[0].dscale = candidate_scale(p.beta, chains[0].dwsum);
    LG_HIP(hipMemcpy(d_chains, chains.data(), sizeof(ChainDev), hipMemcpyHostToDevice));
    invalidate_graphs();
    return LG_OK;
  }
  return sync_schedule();
}

// Cluster stage on a given configuration, written into the "new" buffers as
// if the diagonal update had produced it.
int Batch::cluster_only(const int32_t* spins_in, const int32_t* bond_offsets, const double* times,
                        const int32_t* types, uint64_t sweep, int32_t* labels, int32_t* spins_out,
                        int32_t* types_out, int64_t* n_clusters) {
  const ChainDev& d = chains[0];
  long long nops = 0;
  int rc = validate_config(d, bond_offsets, times, types, nops);
  if (rc) return rc;
  if (nops > op_cap) return set_error(LG_ERR_CAPACITY, "configuration exceeds operator capacity");
  std::vector<int> cnt(d.Nb), oi(nops), kc(nops);
  for (int b = 0; b < d.Nb; ++b) {
    cnt[b] = bond_offsets[b + 1] - bond_offsets[b];
    int k_off = 0;
    for (int k = bond_offsets[b]; k < bond_offsets[b + 1]; ++k) {
      oi[k] = (b << 3) | ((types[k] >> 2) << 1) | (types[k] & 1);
      kc[k] = k_off;
      k_off += types[k] & 1;
    }
  }
  std::vector<uint8_t> sp(d.N);
  for (int s = 0; s < d.N; ++s) sp[s] = (uint8_t)(spins_in[s] & 1);
  const int from = cur;
  State& o = st[from];
  State& w = st[1 - from];
  LG_HIP(hipMemcpy(o.spins, sp.data(), d.N, hipMemcpyHostToDevice));
  LG_HIP(hipMemcpy(counts, cnt.data(), d.Nb * sizeof(int), hipMemcpyHostToDevice));
  if (nops) {
    LG_HIP(hipMemcpy(w.ot, times, nops * sizeof(double), hipMemcpyHostToDevice));
    LG_HIP(hipMemcpy(w.oi, oi.data(), nops * sizeof(int), hipMemcpyHostToDevice));
    LG_HIP(hipMemcpy(kc_new, kc.data(), nops * sizeof(int), hipMemcpyHostToDevice));
  }
  uint32_t sw = (uint32_t)sweep;
  LG_HIP(hipMemcpy(d_sweep, &sw, sizeof(uint32_t), hipMemcpyHostToDevice));
  LG_HIP(hipMemset(ring_pos, 0, sizeof(int)));
  DevBatch B = dev(from);
  launch_site_scan(B, stream);
  launch_cluster_stage(B, cfg, stream);
  LG_HIP(hipGetLastError());
  LG_HIP(hipStreamSynchronize(stream));
  rc = check_flags();
  if (rc) return rc;
  const int nelem = d.N + 4 * (int)nops;
  int* canon = nullptr;
  int* lab = nullptr;
  int* roots = nullptr;
  LG_HIP(hipMalloc(&canon, sizeof(int) * std::max(node_cap, 1)));
  LG_HIP(hipMalloc(&lab, sizeof(int) * std::max(nelem, 1)));
  LG_HIP(hipMalloc(&roots, sizeof(int) * std::max(nelem, 1)));
  std::vector<int> big(node_cap, 0x7fffffff);
  hipError_t e = hipMemcpy(canon, big.data(), sizeof(int) * node_cap, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch_canon(B, 0, nelem, canon, lab, roots, stream);
    e = hipStreamSynchronize(stream);
  }
  std::vector<int> labv(nelem);
  if (e == hipSuccess) e = hipMemcpy(labv.data(), lab, sizeof(int) * nelem, hipMemcpyDeviceToHost);
  hipFree(canon);
  hipFree(lab);
  hipFree(roots);
  if (e != hipSuccess) return set_error(LG_ERR_DEVICE, std::string("canonical labels: ") + hipGetErrorString(e));
  if (labels) std::memcpy(labels, labv.data(), sizeof(int) * nelem);
  std::vector<uint8_t> spo(d.N);
  LG_HIP(hipMemcpy(spo.data(), w.spins, d.N, hipMemcpyDeviceToHost));
  if (spins_out) for (int s = 0; s < d.N; ++s) spins_out[s] = spo[s];
  if (types_out && nops) {
    LG_HIP(hipMemcpy(oi.data(), w.oi, nops * sizeof(int), hipMemcpyDeviceToHost));
    for (long long k = 0; k < nops; ++k) types_out[k] = (oi[k] & 1) | (((oi[k] >> 1) & 3) << 2);
  }
  if (n_clusters) {
    std::vector<double> rec(R_COUNT);
    LG_HIP(hipMemcpy(rec.data(), rec_ring, R_COUNT * sizeof(double), hipMemcpyDeviceToHost));
    *n_clusters = (int64_t)rec[R_NC];
  }
  return LG_OK;
}

}  // namespace lg

// ===========================================================================
// C ABI
// ===========================================================================
using lg::Batch;
using lg::set_error;

struct lg_chain { Batch b; };
struct lg_batch { Batch b; };

extern "C" {

int lg_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return LG_ERR_ARG;
  std::snprintf(buf, len, "%s", lg::g_last_error.c_str());
  return LG_OK;
}

int lg_device_count(int* n) {
  if (!n) return set_error(LG_ERR_ARG, "null pointer");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  return LG_OK;
}

int lg_chain_create(const lg_params* p, lg_chain** out) {
  if (!p || !out) return set_error(LG_ERR_ARG, "null pointer");
  *out = nullptr;
  lg_chain* h = new lg_chain();
  int rc = h->b.init(p, 1);
  if (rc) { delete h; return rc; }
  *out = h;
  return LG_OK;
}

int lg_chain_sweep(lg_chain* h, uint32_t n, lg_sweep_record* rec) {
  if (!h) return set_error(LG_ERR_ARG, "null handle");
  return h->b.run(n, rec);
}

int lg_chain_info(lg_chain* h, int32_t* n_sites, int32_t* n_bonds, int64_t* n_ops, uint64_t* sweep) {
  if (!h) return set_error(LG_ERR_ARG, "null handle");
  if (n_sites) *n_sites = h->b.chains[0].N;
  if (n_bonds) *n_bonds = h->b.chains[0].Nb;
  if (sweep) *sweep = h->b.host_sweep[0];
  if (n_ops) {
    long long n = 0;
    int rc = h->b.num_ops(0, &n);
    if (rc) return rc;
    *n_ops = n;
  }
  return LG_OK;
}

int lg_chain_get_config(lg_chain* h, int32_t* spins, int32_t* bond_offsets, double* times, int32_t* types) {
  if (!h) return set_error(LG_ERR_ARG, "null handle");
  return h->b.get_config(0, spins, bond_offsets, times, types);
}

int lg_chain_set_config(lg_chain* h, const int32_t* spins, const int32_t* bond_offsets, const double* times,
                        const int32_t* types, uint64_t sweep) {
  if (!h || !spins || !bond_offsets) return set_error(LG_ERR_ARG, "null pointer");
  return h->b.set_config(0, spins, bond_offsets, times, types, sweep);
}

int lg_chain_destroy(lg_chain* h) {
  delete h;
  return LG_OK;
}

int lg_chain_set_schedule(lg_chain* h, int32_t n, const int32_t* steps, const double* temperatures) {
  if (!h) return set_error(LG_ERR_ARG, "null handle");
  return h->b.set_schedule(n, steps, temperatures);
}

int lg_cluster_config(const lg_params* p, const int32_t* spins, const int32_t* bond_offsets, const double* times,
                      const int32_t* types, uint64_t sweep, int32_t* labels, int32_t* spins_out,
                      int32_t* types_out, int64_t* n_clusters) {
  if (!p || !spins || !bond_offsets) return set_error(LG_ERR_ARG, "null pointer");
  if (p->hx != 0 || p->d != 0 || p->algorithm != LG_ALGORITHM_PATH_INTEGRAL)
    return set_error(LG_ERR_UNSUPPORTED,
                     "the cluster stage of a given configuration takes hx = 0, d = 0 and the path integral");
  lg_params q = *p;
  // capacity from the configuration itself
  lg::ChainDev d = lg::make_chain_dev(q);
  long long nops = bond_offsets[d.Nb];
  Batch b;
  int rc = b.init_for_config(&q, nops);
  if (rc) return rc;
  return b.cluster_only(spins, bond_offsets, times, types, sweep, labels, spins_out, types_out, n_clusters);
}

int lg_batch_create(const lg_params* p, int32_t n_chains, lg_batch** out) {
  if (!p || !out) return set_error(LG_ERR_ARG, "null pointer");
  *out = nullptr;
  lg_batch* h = new lg_batch();
  int rc = h->b.init(p, n_chains);
  if (rc) { delete h; return rc; }
  *out = h;
  return LG_OK;
}

int lg_batch_run(lg_batch* h, uint32_t n, lg_sweep_record* rec) {
  if (!h) return set_error(LG_ERR_ARG, "null handle");
  return h->b.run(n, rec);
}

int lg_batch_set_graph(lg_batch* h, int enable) {
  if (!h) return set_error(LG_ERR_ARG, "null handle");
  h->b.use_graph = enable != 0;
  return LG_OK;
}

int lg_chain_get_launch(lg_chain* h, lg_launch* out) {
  if (!h || !out) return set_error(LG_ERR_ARG, "null pointer");
  h->b.get_launch(out);
  return LG_OK;
}

int lg_chain_set_launch(lg_chain* h, const lg_launch* l) {
  if (!h || !l) return set_error(LG_ERR_ARG, "null pointer");
  return h->b.set_launch(*l);
}

int lg_batch_get_launch(lg_batch* h, lg_launch* out) {
  if (!h || !out) return set_error(LG_ERR_ARG, "null pointer");
  h->b.get_launch(out);
  return LG_OK;
}

int lg_batch_set_launch(lg_batch* h, const lg_launch* l) {
  if (!h || !l) return set_error(LG_ERR_ARG, "null pointer");
  return h->b.set_launch(*l);
}

int lg_exchange_create(const lg_params* p, const double* betas, int32_t n, uint32_t seed, lg_batch** out) {
  if (!p || !out) return set_error(LG_ERR_ARG, "null pointer");
  *out = nullptr;
  lg_batch* h = new lg_batch();
  int rc = h->b.init_exchange(*p, betas, n, seed);
  if (rc) {
    delete h;
    return rc;
  }
  *out = h;
  return LG_OK;
}

int lg_exchange_state(lg_batch* h, int32_t* chain_at, uint64_t* accepted, uint64_t* steps) {
  if (!h) return set_error(LG_ERR_ARG, "null handle");
  return h->b.exchange_state(chain_at, accepted, steps);
}

int lg_batch_add(lg_batch* h, const lg_params* p, int32_t n) {
  if (!h || !p) return set_error(LG_ERR_ARG, "null pointer");
  if (n <= 0) return set_error(LG_ERR_ARG, "need at least one chain");
  lg::Batch& b = h->b;
  if (b.is_exchange()) return set_error(LG_ERR_UNSUPPORTED, "chains cannot join an exchange batch");
  for (int i = 0; i < n; ++i) {
    int rc = lg::check_params(p[i]);
    if (rc) return rc;
  }
  const size_t n0 = b.chains.size();
  std::vector<lg::Batch::HostConfig> cfg(n0 + n);
  for (size_t c = 0; c < n0; ++c) {
    int rc = b.chain_config((int)c, cfg[c].spins, cfg[c].bo, cfg[c].times, cfg[c].types);
    if (rc) return rc;
    cfg[c].sweep = b.host_sweep[c];
  }
  std::vector<lg_params> all(b.params);
  for (int i = 0; i < n; ++i) {
    const lg::ChainDev d = lg::make_chain_dev(p[i]);
    cfg[n0 + i].spins.assign(d.N, 0);  // all up, empty operator string (path_integral.C:225)
    cfg[n0 + i].bo.assign(d.Nb + 1, 0);
    all.push_back(p[i]);
  }
  lg::Batch nb;
  nb.use_graph = b.use_graph;
  int rc = nb.init(all.data(), (int)all.size());
  if (rc) return rc;
  if ((rc = nb.load_configs(cfg))) return rc;
  b.adopt(nb);
  return LG_OK;
}

// Chain checkpoint blob, in the content order of the reference worker's save
// (path_integral.C:111-124: mcs, spins, operators; an operator is its type,
// location and time, operator.h:108-109,138-139, location_impl.h:43):
//   "LGCHAIN2", sweep (uint64; the Monte Carlo step count and the Philox
//   counter), n_sites (uint64), spins (int32 each, 0 = up), n_ops (uint64),
//   operators in imaginary-time order (ties by bond, then by their order on
//   the bond) as {int32 type = offdiagonal | graph << 2, int32 location =
//   bond << 1 | 1, double time}, then the chain's parameters (lg_params:
//   lattice, couplings, temperature, seed and chain id), which load checks.
static const char kChainMagic[8] = {'L', 'G', 'C', 'H', 'A', 'I', 'N', '2'};
static const size_t kOpRecord = 16;

static size_t chain_blob_size(const lg::ChainDev& d, long long nops) {
  return 8 + 8 + 8 + (size_t)d.N * 4 + 8 + (size_t)nops * kOpRecord + sizeof(lg_params);
}

int lg_chain_state_size(lg_chain* h, size_t* bytes) {
  if (!h || !bytes) return set_error(LG_ERR_ARG, "null pointer");
  long long n = 0;
  int rc = h->b.num_ops(0, &n);
  if (rc) return rc;
  *bytes = chain_blob_size(h->b.chains[0], n);
  return LG_OK;
}

int lg_chain_save(lg_chain* h, void* buf, size_t len) {
  if (!h || !buf) return set_error(LG_ERR_ARG, "null pointer");
  std::vector<int> spins, bo, types;
  std::vector<double> times;
  int rc = h->b.chain_config(0, spins, bo, times, types);
  if (rc) return rc;
  const lg::ChainDev& d = h->b.chains[0];
  const long long nops = (long long)times.size();
  if (len < chain_blob_size(d, nops)) return set_error(LG_ERR_ARG, "checkpoint buffer too small");
  std::vector<int> bond(nops), order(nops);
  for (int b = 0; b < d.Nb; ++b)
    for (int k = bo[b]; k < bo[b + 1]; ++k) bond[k] = b;
  for (long long k = 0; k < nops; ++k) order[k] = (int)k;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (times[a] != times[b]) return times[a] < times[b];
    return bond[a] < bond[b];
  });
  char* o = (char*)buf;
  auto put = [&](const void* p, size_t n) { std::memcpy(o, p, n); o += n; };
  put(kChainMagic, 8);
  const uint64_t sw = h->b.host_sweep[0], ns = (uint64_t)d.N, no = (uint64_t)nops;
  put(&sw, 8);
  put(&ns, 8);
  put(spins.data(), (size_t)d.N * 4);
  put(&no, 8);
  for (long long i = 0; i < nops; ++i) {
    const int k = order[i];
    // looper's location (location_impl.h:37): pos << 1 | is_bond; a site
    // operator (transverse field) sits at its site
    const int32_t type = types[k];
    const int32_t loc = bond[k] >= d.nb_real ? ((bond[k] - d.nb_real) << 1) : ((bond[k] << 1) | 1);
    put(&type, 4);
    put(&loc, 4);
    put(&times[k], 8);
  }
  put(&h->b.params[0], sizeof(lg_params));
  return LG_OK;
}

int lg_chain_load(lg_chain* h, const void* buf, size_t len) {
  if (!h || !buf) return set_error(LG_ERR_ARG, "null pointer");
  const lg::ChainDev& d = h->b.chains[0];
  const char* in = (const char*)buf;
  const char* end = in + len;
  auto get = [&](void* p, size_t n) {
    if ((size_t)(end - in) < n) return false;
    std::memcpy(p, in, n);
    in += n;
    return true;
  };
  char magic[8];
  uint64_t sw = 0, ns = 0, no = 0;
  if (!get(magic, 8) || std::memcmp(magic, kChainMagic, 8) != 0) return set_error(LG_ERR_ARG, "not a chain checkpoint");
  if (!get(&sw, 8) || !get(&ns, 8)) return set_error(LG_ERR_ARG, "checkpoint truncated");
  if (ns != (uint64_t)d.N) return set_error(LG_ERR_ARG, "checkpoint of a different lattice");
  std::vector<int32_t> spins(d.N);
  if (!get(spins.data(), (size_t)d.N * 4) || !get(&no, 8)) return set_error(LG_ERR_ARG, "checkpoint truncated");
  if (no > (uint64_t)(end - in) / kOpRecord) return set_error(LG_ERR_ARG, "checkpoint truncated");
  std::vector<int32_t> type(no), bond(no);
  std::vector<double> time(no);
  for (uint64_t i = 0; i < no; ++i) {
    int32_t loc = 0;
    get(&type[i], 4);
    get(&loc, 4);
    get(&time[i], 8);
    const int pos = loc >> 1;
    const bool ok = (loc & 1) ? (pos >= 0 && pos < d.nb_real) : (d.Nb > d.nb_real && pos >= 0 && pos < d.N);
    if (!ok) return set_error(LG_ERR_ARG, "checkpoint operator location invalid");
    bond[i] = (loc & 1) ? pos : d.nb_real + pos;
  }
  lg_params q;
  if (!get(&q, sizeof(lg_params))) return set_error(LG_ERR_ARG, "checkpoint truncated");
  const lg_params& p = h->b.params[0];
  if (q.lattice != p.lattice || q.L != p.L || q.beta != p.beta || q.jxy != p.jxy || q.jz != p.jz ||
      q.force_scatter != p.force_scatter || q.seed != p.seed || q.chain_id != p.chain_id || q.h != p.h ||
      q.two_s != p.two_s || q.hx != p.hx || q.d != p.d || q.algorithm != p.algorithm)
    return set_error(LG_ERR_ARG, "checkpoint of a different chain (parameters, seed or chain id)");
  // bond-major order for the device: a stable sort by bond keeps each bond's
  // operators in their imaginary-time order
  std::vector<int> order(no);
  for (uint64_t i = 0; i < no; ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return bond[a] < bond[b]; });
  std::vector<int32_t> bo(d.Nb + 1, 0), types(no);
  std::vector<double> times(no);
  for (uint64_t i = 0; i < no; ++i) {
    ++bo[bond[i] + 1];
    types[i] = type[order[i]];
    times[i] = time[order[i]];
  }
  for (int b = 0; b < d.Nb; ++b) bo[b + 1] += bo[b];
  return h->b.set_config(0, spins.data(), bo.data(), times.data(), types.data(), sw);
}

int lg_batch_tune(lg_batch* h, uint32_t sweeps_per_trial) {
  if (!h) return set_error(LG_ERR_ARG, "null handle");
  return h->b.tune(sweeps_per_trial);
}

int lg_batch_chain_info(lg_batch* h, int32_t chain, int64_t* n_ops, uint64_t* sweep) {
  if (!h) return set_error(LG_ERR_ARG, "null handle");
  if (chain < 0 || chain >= (int)h->b.chains.size()) return set_error(LG_ERR_ARG, "chain index out of range");
  if (sweep) *sweep = h->b.host_sweep[chain];
  if (n_ops) {
    long long n = 0;
    int rc = h->b.num_ops(chain, &n);
    if (rc) return rc;
    *n_ops = n;
  }
  return LG_OK;
}

// plain streaming copy used to calibrate HBM counters and measure the
// achievable bandwidth
__global__ void __launch_bounds__(256) k_copy_stream(const double4* __restrict__ a, double4* __restrict__ b,
                                                     size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

int lg_copy_bandwidth(uint64_t bytes, int iters, double* gbs) {
  if (!gbs || bytes < 32 || iters < 1) return set_error(LG_ERR_ARG, "invalid copy size");
  size_t n = bytes / sizeof(double4);
  double4* a = nullptr;
  double4* b = nullptr;
  hipEvent_t e0, e1;
  LG_HIP(hipMalloc(&a, n * sizeof(double4)));
  if (hipMalloc(&b, n * sizeof(double4)) != hipSuccess) { (void)hipFree(a); return set_error(LG_ERR_DEVICE, "hipMalloc"); }
  LG_HIP(hipMemset(a, 1, n * sizeof(double4)));
  LG_HIP(hipEventCreate(&e0));
  LG_HIP(hipEventCreate(&e1));
  int blocks = 256 * 16;
  hipLaunchKernelGGL(k_copy_stream, dim3(blocks), dim3(256), 0, 0, a, b, n);  // warm
  LG_HIP(hipEventRecord(e0, 0));
  for (int i = 0; i < iters; ++i) hipLaunchKernelGGL(k_copy_stream, dim3(blocks), dim3(256), 0, 0, a, b, n);
  LG_HIP(hipEventRecord(e1, 0));
  LG_HIP(hipEventSynchronize(e1));
  float ms = 0;
  LG_HIP(hipEventElapsedTime(&ms, e0, e1));
  *gbs = 2.0 * n * sizeof(double4) * iters / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return LG_OK;
}

int lg_profile_num_kernels(void) { return lg::NUM_SWEEP_KERNELS; }

const char* lg_profile_kernel_name(int i) {
  return (i >= 0 && i < lg::NUM_SWEEP_KERNELS) ? lg::kKernelNames[i] : "";
}

int lg_chain_profile(lg_chain* h, uint32_t n, double* avg_ms) {
  if (!h || !avg_ms) return set_error(LG_ERR_ARG, "null pointer");
  return h->b.profile(n, avg_ms);
}

int lg_batch_profile(lg_batch* h, uint32_t n, double* avg_ms) {
  if (!h || !avg_ms) return set_error(LG_ERR_ARG, "null pointer");
  return h->b.profile(n, avg_ms);
}

int lg_batch_destroy(lg_batch* h) {
  delete h;
  return LG_OK;
}

}  // extern "C"
