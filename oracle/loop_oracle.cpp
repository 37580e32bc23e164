// CPU oracle for the continuous-imaginary-time loop update (S=1/2 XXZ).
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
// this file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it (through ctypes), and only as the checker / the timed CPU
// baseline.
//
// Three restatements live here, all plain serial C++:
//
//  1. "looper mode": the serial path of ALPS/looper's "loop; path integral"
//     worker for bond-only S=1/2 XXZ models without field and sign
//     (path_integral.C:380-864, the #else branches), with looper's union-find
//     semantics (looper/union_find.h:57-82,163-172,242-284,307-350), cut rules
//     and leg accessors (looper/graph_impl.h:255-328,489-494), weights
//     (looper/weight_impl.h:165-188), improved estimators
//     (looper/susceptibility.h:96-256, looper/stiffness.h:82-130,
//     looper/energy.h:55-86).  RNG: std::mt19937 + std distributions and a
//     Walker alias table of our own; the worker's ALPS RNG stream is not
//     reproduced (parity for this mode is statistical).
//
//  2. "standalone mode": standalone/loop.C:39-180 with its own union-find
//     variant (standalone/union_find.h:112-143), consuming std::mt19937(29833)
//     exactly as that program does, so its printed averages can be compared
//     digit for digit with standalone/loop.op and with the reference program
//     built under oracle/_ref.
//
//  3. "philox mode": the counter-based algorithm the GPU kernels implement
//     (per-bond Poisson streams from Philox4x32-10, 2^-52 integer time grid,
//     world-line segments as union-find nodes, min-index roots), written
//     sequentially so GPU and CPU trajectories can be compared bit for bit.
//
// Plus partition oracles on an explicit leg graph: looper's fragment
// algorithm, a breadth-first search, and canonicalisation of labels to the
// minimum element index.
//
// Build: see oracle/Makefile (g++ -O2 -ffp-contract=off; contraction must stay
// off so the deterministic logarithm matches the device's bit for bit).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <random>
#include <stdexcept>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): 10 rounds of a 4x32 multiply-xor network with a Weyl
// key schedule.
// ---------------------------------------------------------------------------
struct u32x4 { uint32_t v[4]; };

inline u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                           uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)M0 * c0;
    uint64_t p1 = (uint64_t)M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += W0; k1 += W1;
  }
  u32x4 out = {{c0, c1, c2, c3}};
  return out;
}

// Counter layout shared with the device: (sweep, stage, element, sub-index),
// key = (seed, chain id).
enum { STAGE_DIAG = 1, STAGE_REGRAPH = 2, STAGE_FLIP = 3, STAGE_BOUNDARY = 5, STAGE_SSE = 6 };

// ---------------------------------------------------------------------------
// Deterministic natural logarithm built only from IEEE-754 +,-,*,/ (no fused
// operations), so that x86-64 and gfx950 produce identical bits.
// log(x) = e*ln2 + 2*atanh(s), s = (m-1)/(m+1), m in (sqrt(1/2), sqrt(2)],
// atanh by its Taylor series s + s^3/3 + s^5/5 + ... (|s| <= 0.1716).
// ---------------------------------------------------------------------------
inline double det_log(double x) {
  uint64_t bits;
  std::memcpy(&bits, &x, 8);
  int e = (int)((bits >> 52) & 0x7ff) - 1023;
  uint64_t mb = (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
  double m;
  std::memcpy(&m, &mb, 8);
  if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  double p = 1.0 / 25.0;
  p = p * z + 1.0 / 23.0;
  p = p * z + 1.0 / 21.0;
  p = p * z + 1.0 / 19.0;
  p = p * z + 1.0 / 17.0;
  p = p * z + 1.0 / 15.0;
  p = p * z + 1.0 / 13.0;
  p = p * z + 1.0 / 11.0;
  p = p * z + 1.0 / 9.0;
  p = p * z + 1.0 / 7.0;
  p = p * z + 1.0 / 5.0;
  p = p * z + 1.0 / 3.0;
  double logm = 2.0 * s + 2.0 * s * (z * p);
  return (double)e * 0.69314718055994530942 + logm;
}

// exp(x) for the field flip, from +,-,*,/ and exact scalings (floor, ldexp):
// Cody-Waite reduction x = k ln2 + r (ln2 split into a 32-bit high part
// and the rest), e^r by its Taylor series to r^14, times 2^k.
inline double det_exp(double x) {
  if (!(x > -745.0)) return 0.0;
  if (x > 709.0) x = 709.0;
  const double k = std::floor(x * 1.4426950408889634 + 0.5);
  const double r = (x - k * 0.6931471803691238) - k * 1.9082149292705877e-10;
  static const double inv_fact[15] = {1.0, 1.0, 0.5, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0, 1.0 / 720.0,
                                      1.0 / 5040.0, 1.0 / 40320.0, 1.0 / 362880.0, 1.0 / 3628800.0,
                                      1.0 / 39916800.0, 1.0 / 479001600.0, 1.0 / 6227020800.0,
                                      1.0 / 87178291200.0};
  double p = inv_fact[14];
  for (int n = 13; n >= 0; --n) p = p * r + inv_fact[n];
  return std::ldexp(p, (int)k);
}

inline double det_tanh(double x) {
  const double a = x < 0 ? -x : x;
  const double e = det_exp(-2.0 * a);
  const double t = (1.0 - e) / (1.0 + e);
  return x < 0 ? -t : t;
}

const uint64_t TGRID = 1ull << 52;  // imaginary time tau = T * 2^-52, T in [0, 2^52)

inline uint64_t exp_gap(const u32x4& x, double scale) {
  uint64_t u53 = ((uint64_t)x.v[0] << 21) | (uint64_t)(x.v[1] >> 11);
  double u = (double)(u53 + 1) * (1.0 / 9007199254740992.0);  // (0, 1]
  double g = -det_log(u) * scale;
  if (!(g < 4503599627370496.0)) return TGRID;
  return (uint64_t)g;
}

// ---------------------------------------------------------------------------
// Lattices with ALPS conventions (pinned by test/spinmodel_helper.op,
// test/virtual_lattice.op): chain bond i = (i, i+1 mod L); square site
// s = x*L + y, bond 2s = (s, s+y^), bond 2s+1 = (s, s+x^); gauge = +1 on the
// sublattice of site 0.
// ---------------------------------------------------------------------------
struct Lattice {
  int kind = 0, L = 0, N = 0, Nb = 0, dim = 1;
  int nv = 1;  // virtual sites per real site (2S)
  int nb_lat = 0;   // the lattice's (virtual) bonds; [nb_lat, nb_real): in-site bonds of the single-ion term
  int nb_real = 0;  // bonds between two sites; bonds nb_real + s are site s's own (transverse field)
  std::vector<int> src, tgt, dir, gauge;
  std::vector<std::vector<int>> site_bonds;  // bonds incident to each site
  // nv > 1 (S = nv/2): looper's virtual lattice for S > 1/2
  // (lattice.h:104-175, virtual_graph): real site r -> virtual sites
  // r*nv + a; real bond rb = (rs, rt) -> the nv^2 virtual bonds
  // rb*nv^2 + a*nv + c = (rs*nv + a, rt*nv + c), in real-bond order and
  // source-major within; every virtual site carries its real site's gauge
  // and every virtual bond its real bond's direction
  void init(int k, int l, int n_virtual = 1) {
    if (n_virtual > 1) {
      Lattice real;
      real.init(k, l, 1);
      kind = k; L = l; nv = n_virtual; dim = real.dim;
      N = real.N * nv; Nb = real.Nb * nv * nv;
      src.resize(Nb); tgt.resize(Nb); dir.resize(Nb);
      for (int rb = 0; rb < real.Nb; ++rb)
        for (int a = 0; a < nv; ++a)
          for (int c = 0; c < nv; ++c) {
            const int b = rb * nv * nv + a * nv + c;
            src[b] = real.src[rb] * nv + a;
            tgt[b] = real.tgt[rb] * nv + c;
            dir[b] = real.dir[rb];
          }
      gauge.resize(N);
      for (int v = 0; v < N; ++v) gauge[v] = real.gauge[v / nv];
      site_bonds.assign(N, std::vector<int>());
      for (int b = 0; b < Nb; ++b) { site_bonds[src[b]].push_back(b); site_bonds[tgt[b]].push_back(b); }
      nb_lat = nb_real = Nb;
      return;
    }
    kind = k; L = l; nv = 1;
    if (kind == 0) {
      N = L; Nb = L; dim = 1;
      src.resize(Nb); tgt.resize(Nb); dir.assign(Nb, 0);
      for (int b = 0; b < Nb; ++b) { src[b] = b; tgt[b] = (b + 1) % L; }
      gauge.resize(N);
      for (int s = 0; s < N; ++s) gauge[s] = (s % 2 == 0) ? 1 : -1;
    } else {
      N = L * L; Nb = 2 * N; dim = 2;
      src.resize(Nb); tgt.resize(Nb); dir.resize(Nb);
      for (int s = 0; s < N; ++s) {
        int x = s / L, y = s % L;
        src[2 * s] = s; tgt[2 * s] = x * L + (y + 1) % L; dir[2 * s] = 1;
        src[2 * s + 1] = s; tgt[2 * s + 1] = ((x + 1) % L) * L + y; dir[2 * s + 1] = 0;
      }
      gauge.resize(N);
      for (int s = 0; s < N; ++s) gauge[s] = ((s / L + s % L) % 2 == 0) ? 1 : -1;
    }
    site_bonds.assign(N, std::vector<int>());
    for (int b = 0; b < Nb; ++b) { site_bonds[src[b]].push_back(b); site_bonds[tgt[b]].push_back(b); }
    nb_lat = nb_real = Nb;
  }
  // Single-ion term D (S^z)^2 of S > 1/2: looper's in-real-site bonds
  // (lattice.h:632-646), after the lattice's bonds, for each real site r the
  // pairs a < c of its virtual sites, a outer: bond (r*nv + a, r*nv + c), an
  // Ising coupling 2D (weight_impl.h:194-197); no bond vector
  void add_d_bonds() {
    const int nr = N / nv;
    for (int r = 0; r < nr; ++r)
      for (int a = 0; a < nv; ++a)
        for (int c = a + 1; c < nv; ++c) {
          const int b = (int)src.size();
          src.push_back(r * nv + a); tgt.push_back(r * nv + c); dir.push_back(-1);
          site_bonds[r * nv + a].push_back(b);
          site_bonds[r * nv + c].push_back(b);
        }
    Nb = (int)src.size();
    nb_real = Nb;
  }
  // Transverse field: site s's operators live on bond nb_real + s, whose
  // source and target are s (a site operator has one leg below and one above
  // on its own site, and no bond vector)
  void add_site_bonds() {
    for (int s = 0; s < N; ++s) {
      src.push_back(s); tgt.push_back(s); dir.push_back(-1);
      site_bonds[s].push_back(Nb + s);
    }
    Nb += N;
  }
  bool is_site_bond(int b) const { return b >= nb_real; }
  bool is_d_bond(int b) const { return b >= nb_lat && b < nb_real; }
  // relative bond vector component d (ALPS bond_vector_relative, unit lattice constant)
  int vr(int b, int d) const { return b >= nb_lat ? 0 : (kind == 0) ? 1 : (dir[b] == d ? 1 : 0); }
};

// ---------------------------------------------------------------------------
// XXZ bond weights, looper/weight_impl.h:165-188 (standard and "ergodic"
// solutions; FORCE_SCATTER = a).
// ---------------------------------------------------------------------------
struct Weights { double v[4]; double offset; int sign; };

inline double crop_0(double x) { return x > 0 ? x : 0; }

Weights xxz_weights(double jxy_in, double jz, double a) {
  Weights w;
  w.sign = (jxy_in <= 0 ? 1 : -1);
  double jxy = std::fabs(jxy_in);
  if (a < 0) a = 0;
  if (a > 1) a = 1;
  if (jxy + std::fabs(jz) > 1e-10) {
    if (jxy - jz > 2 * a * jxy) {
      w.v[0] = crop_0(std::min(jxy / 2, (jxy + jz) / 4));
      w.v[1] = crop_0(std::min(jxy / 2, (jxy - jz) / 4));
      w.v[2] = crop_0(-(jxy - jz) / 2.0);
      w.v[3] = crop_0(-(jxy + jz) / 2);
    } else {
      w.v[0] = (1 - a) * jxy / 2;
      w.v[1] = a * jxy / 2;
      w.v[2] = -((1 - 2 * a) * jxy - jz) / 2;
      w.v[3] = 0;
    }
  } else {
    w.v[0] = w.v[1] = w.v[2] = w.v[3] = 0;
  }
  w.offset = (w.v[0] + w.v[1] + w.v[2] + w.v[3]) / 2;
  return w;
}

// Integer thresholds shared by the philox mode and the device:
// graph g is chosen for 32-bit draw x when thr[g-1] <= x < thr[g].
struct Thresholds { uint64_t thr[4]; uint64_t toff; };

Thresholds make_thresholds(const Weights& w) {
  Thresholds t;
  double W = w.v[0] + w.v[1] + w.v[2] + w.v[3];
  int last = -1;
  for (int g = 0; g < 4; ++g) if (w.v[g] > 0) last = g;
  double cum = 0;
  for (int g = 0; g < 4; ++g) {
    cum += w.v[g];
    if (last < 0 || g >= last) t.thr[g] = 1ull << 32;
    else t.thr[g] = (uint64_t)std::floor(std::ldexp(cum / W, 32));
  }
  double s01 = w.v[0] + w.v[1];
  t.toff = (s01 > 0) ? (uint64_t)std::floor(std::ldexp(w.v[0] / s01, 32)) : (1ull << 32);
  return t;
}

inline int choose_graph(uint32_t x, const Thresholds& t) {
  int g = 0;
  while (g < 3 && (uint64_t)x >= t.thr[g]) ++g;
  return g;
}

// ---------------------------------------------------------------------------
// looper's union-find node and operations, serial branch
// (looper/union_find.h:57-82 node, 163-172 root_index_ph, 242-284 unify,
// 307-350 set_id/copy_id).
// ---------------------------------------------------------------------------
struct UFNode {
  int parent_ = -1;
  int id_ = 0;
  bool is_root() const { return parent_ <= 0; }
  void set_parent(int p) { parent_ = p + 1; }
  int parent() const { return parent_ - 1; }
  void set_weight(int w) { parent_ = -w; }
  int weight() const { return -parent_; }
};

int uf_root_index(const std::vector<UFNode>& v, int g) {
  while (!v[g].is_root()) g = v[g].parent();
  return g;
}

int uf_root_index_ph(std::vector<UFNode>& v, int g) {
  if (v[g].is_root()) return g;
  while (true) {
    int p = v[g].parent();
    if (v[p].is_root()) return p;
    v[g].set_parent(v[p].parent());
    g = p;
  }
}

int uf_unify(std::vector<UFNode>& v, int g0, int g1) {
  int r0 = uf_root_index_ph(v, g0);
  int r1 = uf_root_index_ph(v, g1);
  if (r0 != r1) {
    if (v[r0].weight() < v[r1].weight()) std::swap(r0, r1);
    v[r0].set_weight(v[r0].weight() + v[r1].weight());
    v[r1].set_parent(r0);
  }
  return r0;
}

// Variant with full compression along both paths (looper's unify_compress,
// union_find.h:204-222; same roots and weights as uf_unify).
int uf_unify_compress(std::vector<UFNode>& v, int g0, int g1) {
  int r0 = uf_root_index(v, g0);
  int r1 = uf_root_index(v, g1);
  if (r0 != r1) {
    if (v[r0].weight() < v[r1].weight()) std::swap(r0, r1);
    v[r0].set_weight(v[r0].weight() + v[r1].weight());
    v[r1].set_parent(r0);
  }
  for (int g : {g0, g1}) {
    while (g != r0) { int p = v[g].parent(); v[g].set_parent(r0); g = p; }
  }
  return r0;
}

int uf_set_id(std::vector<UFNode>& v, int start, int n, int nc) {
  for (int i = start; i < start + n; ++i) if (v[i].is_root()) v[i].id_ = nc++;
  return nc;
}
void uf_copy_id(std::vector<UFNode>& v, int start, int n) {
  for (int i = start; i < start + n; ++i) v[i].id_ = v[uf_root_index(v, i)].id_;
}

// Cut rules and reconnect (looper/graph_impl.h:255-295) and leg accessors
// (graph_impl.h:489-494).
inline bool is_compatible(int g, int c0, int c1) { return ((g & 1) ^ c0 ^ c1) != 0; }
inline bool is_frozen(int g) { return (g & 2) == 2; }

struct Reconnect { int fid, curr0, curr1, loop0, loop1; };
Reconnect reconnect(std::vector<UFNode>& frag, int fid, int g, int curr0, int curr1) {
  Reconnect r;
  if ((g & 2) == 2) {
    int u = uf_unify(frag, curr0, curr1);
    r.loop0 = r.loop1 = r.curr0 = r.curr1 = u;
  } else if (g == 0) {
    frag[fid] = UFNode();
    r.loop0 = uf_unify(frag, curr0, curr1);
    r.loop1 = r.curr0 = r.curr1 = fid;
    ++fid;
  } else {
    r.loop0 = curr0; r.loop1 = curr1;
    r.curr0 = curr1; r.curr1 = curr0;
  }
  r.fid = fid;
  return r;
}
inline int loop_l0(int, int l0, int) { return l0; }
inline int loop_l1(int g, int l0, int l1) { return (g == 0 || g == 5) ? l0 : l1; }
inline int loop_u0(int, int, int l1) { return l1; }
inline int loop_u1(int g, int l0, int l1) { return (g == 0 || g == 5) ? l1 : l0; }

// ---------------------------------------------------------------------------
// Sweep record layout (shared with the device library's lg_sweep_record).
// ---------------------------------------------------------------------------
enum {
  R_NOP = 0, R_NC, R_ENERGY,
  R_UMAG0, R_USIZE2, R_UMAG2, R_USIZE4, R_UMAG4, R_USIZE, R_UMAG,
  R_SMAG0, R_SSIZE2, R_SMAG2, R_SSIZE4, R_SMAG4, R_SSIZE, R_SMAG,
  R_W2, R_BETA, R_SWEEP,
  R_UMAGA, R_SMAGA, R_EFIELD,  // normal estimator sums, Zeeman energy (FIELD path)
  R_SIGN,                      // (-1)^(off-diagonal operators of negative weight); 1 if sign-free
  R_COUNT
};

// Sign problem (model_parameter.h:617-644): with Jxy > 0 and a transverse
// field no gauge removes the negative weights; the off-diagonal operators on
// the lattice bonds (antiferromagnetic XY) and, for hx < 0, on the site
// bonds then count against the sign (path_integral.C:551-576)
struct SignRule {
  bool lattice = false, site = false;
  void init(double jxy, double hx) { lattice = hx != 0 && jxy > 0; site = lattice && hx < 0; }
  bool negative(const Lattice& lat, int b) const {
    return b < lat.nb_lat ? lattice : (lat.is_site_bond(b) ? site : false);
  }
};

struct Estimate {   // susceptibility + stiffness improved estimate (per cluster)
  double usize0 = 0, umag0 = 0, usize = 0, umag = 0;
  double ssize0 = 0, smag0 = 0, ssize = 0, smag = 0;
  double winding[3] = {0, 0, 0};
};

inline double p2(double x) { return x * x; }
inline double p4(double x) { return p2(p2(x)); }

void collect(const Estimate& e, double* rec) {
  rec[R_UMAG0] += e.umag0;
  rec[R_USIZE2] += p2(e.usize0);
  rec[R_UMAG2] += p2(e.umag0);
  rec[R_USIZE4] += p4(e.usize0);
  rec[R_UMAG4] += p4(e.umag0);
  rec[R_USIZE] += p2(e.usize);
  rec[R_UMAG] += p2(e.umag);
  rec[R_SMAG0] += e.smag0;
  rec[R_SSIZE2] += p2(e.ssize0);
  rec[R_SMAG2] += p2(e.smag0);
  rec[R_SSIZE4] += p4(e.ssize0);
  rec[R_SMAG4] += p4(e.smag0);
  rec[R_SSIZE] += p2(e.ssize);
  rec[R_SMAG] += p2(e.smag);
  for (int d = 0; d < 3; ++d) rec[R_W2] += p2(0.5 * e.winding[d]);
  rec[R_UMAGA] += e.umag;   // normal estimator (susceptibility.h:258-339)
  rec[R_SMAGA] += e.smag;
}

// end_s / begin_s / start_bottom of susceptibility.h:113-147
inline void est_end(Estimate& e, double t, int gs, int c) {
  e.usize += t * 0.5;
  e.umag += t * (0.5 - c);
  e.ssize += gs * t * 0.5;
  e.smag += gs * t * (0.5 - c);
}
inline void est_begin(Estimate& e, double t, int gs, int c) { est_end(e, -t, gs, c); }
inline void est_start_bottom(Estimate& e, int gs, int c) {
  est_begin(e, 0.0, gs, c);
  e.usize0 += 0.5;
  e.umag0 += (0.5 - c);
  e.ssize0 += gs * 0.5;
  e.smag0 += gs * (0.5 - c);
}

// ---------------------------------------------------------------------------
// Walker alias table (Vose's construction), our own.
// ---------------------------------------------------------------------------
struct Alias {
  std::vector<double> prob;
  std::vector<int> alias;
  void init(const std::vector<double>& w) {
    int n = (int)w.size();
    prob.assign(n, 0); alias.assign(n, 0);
    double sum = 0;
    for (double x : w) sum += x;
    std::vector<double> q(n);
    std::vector<int> small, large;
    for (int i = 0; i < n; ++i) {
      q[i] = w[i] * n / sum;
      (q[i] < 1 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
      int s = small.back(); small.pop_back();
      int l = large.back(); large.pop_back();
      prob[s] = q[s]; alias[s] = l;
      q[l] = (q[l] + q[s]) - 1;
      (q[l] < 1 ? small : large).push_back(l);
    }
    for (int i : large) { prob[i] = 1; alias[i] = i; }
    for (int i : small) { prob[i] = 1; alias[i] = i; }
  }
  template <class RNG>
  int operator()(RNG& u01) const {
    int n = (int)prob.size();
    double x = u01() * n;
    int i = (int)x;
    if (i >= n) i = n - 1;
    return (x - i < prob[i]) ? i : alias[i];
  }
};

// ---------------------------------------------------------------------------
// Annealing schedule, looper/temperature.h:93-113: (start step, temperature)
// pairs with ascending steps, the last one (first step at the final
// temperature, final temperature).  T(s) is the first temperature up to the
// first step, linear in between, the final one after the last step; s < 1
// is taken as 1 (the reference's operator() rejects it).
// ---------------------------------------------------------------------------
struct Schedule {
  std::vector<int> step;
  std::vector<double> temp;
  bool on() const { return !step.empty(); }
  double T(int s) const {
    if (s < 1) s = 1;
    const size_t k = std::lower_bound(step.begin(), step.end(), s) - step.begin();
    if (k == 0) return temp.front();
    if (k == step.size()) return temp.back();
    return temp[k - 1] + (s - step[k - 1]) * (temp[k] - temp[k - 1]) / (step[k] - step[k - 1]);
  }
};

// ---------------------------------------------------------------------------
// Boundary permutation for S > 1/2 (looper/permutation.h): a random
// permutation of a real site's virtual sites under which every tau = 0 spin
// meets an equal tau = 1 spin.
//   shuffle_range:  looper::random_shuffle (permutation.h:36-45), swapping
//                   each position with one drawn from the rest, (int)(n*u);
//   partition_by:   looper::guided_sort_01 (permutation.h:51-80): entries at
//                   positions whose guide is 1 in the front part exchanged
//                   with entries at positions whose guide is 0 in the back,
//                   from both ends inwards;
//   partitioned_shuffle: looper::partitioned_random_shuffle
//                   (permutation.h:88-112): partition the identity by the
//                   tau = 1 spins, shuffle each part, partition by the tau = 0
//                   spins.
// ---------------------------------------------------------------------------
template <class U01>
void shuffle_range(int* first, int n, U01& u01) {
  for (; n > 1; ++first, --n) std::swap(first[0], first[(int)(n * u01())]);
}

void partition_by(int* perm, int n, const int* guide) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    while (lo < hi && guide[lo] == 0) ++lo;
    while (lo < hi && guide[hi] == 1) --hi;
    if (lo >= hi) return;
    std::swap(perm[lo], perm[hi]);
    ++lo;
    --hi;
  }
}

template <class U01>
void partitioned_shuffle(int* perm, int n, const int* guide0, const int* guide1, U01& u01) {
  partition_by(perm, n, guide1);
  int c = 0;
  for (int i = 0; i < n; ++i) c += guide1[i] == 0;
  shuffle_range(perm, c, u01);
  shuffle_range(perm + c, n - c, u01);
  partition_by(perm, n, guide0);
}

// ---------------------------------------------------------------------------
// Looper-mode worker: path_integral.C:380-864, serial, FIELD=SIGN=false,
// IMPROVE=true, S=1/2 (max_virtual_sites == 1).
// ---------------------------------------------------------------------------
struct LooperOp {
  int bond;
  int type;     // bit0 off-diagonal, graph << 2 (operator.h:52-117)
  double time;
  int loop0, loop1;
};

struct LooperChain {
  Lattice lat;
  SignRule sgn;
  Weights w, wd;  // lattice bonds; in-site bonds of the single-ion term
  double beta = 1, offset_total = 0, graph_weight = 0, p_off = 1, p_off_d = 1;
  std::mt19937 eng;
  std::uniform_real_distribution<double> u01{0.0, 1.0};
  Alias alias;
  std::vector<std::pair<int, int>> graph_of;  // alias index -> (bond, g)
  std::vector<int> spins, spins_c, current;
  std::vector<LooperOp> ops, ops_p;
  std::vector<double> times;
  std::vector<UFNode> frag;
  std::vector<Estimate> est;
  std::vector<char> to_flip;
  uint64_t mcs = 0;

  double h = 0;   // longitudinal field, site term -h S^z
  Schedule sched; // annealing: beta = 1 / T(mcs) before each sweep (path_integral.C:325)

  void init(int kind, int L, double b, double jxy, double jz, double fs, uint32_t seed, double field = 0,
            int nv = 1, double hx = 0, double d = 0) {
    lat.init(kind, L, nv);
    sgn.init(jxy, hx);
    w = xxz_weights(jxy, jz, fs);
    beta = b;
    h = field;
    eng.seed(seed);
    std::vector<double> wl;
    for (int bd = 0; bd < lat.Nb; ++bd)
      for (int g = 0; g < 4; ++g)
        if (w.v[g] > 1e-10) { wl.push_back(w.v[g]); graph_of.push_back({bd, g}); graph_weight += w.v[g]; }
    offset_total = lat.Nb * w.offset;
    // single-ion term D (S^z)^2 = D nv/4 + sum_{a<c} 2D s_a s_c per real
    // site: the in-site bonds with the weights of an Ising coupling 2D
    // (weight_impl.h:194-197) and the constant (model_parameter.h:592-595,
    // weight_impl.h:400-404)
    if (d != 0) {
      const int nb = lat.Nb;
      if (nv > 1) lat.add_d_bonds();
      wd = xxz_weights(0.0, 2 * d, fs);
      for (int bd = nb; bd < lat.Nb; ++bd)
        for (int g = 0; g < 4; ++g)
          if (wd.v[g] > 1e-10) { wl.push_back(wd.v[g]); graph_of.push_back({bd, g}); graph_weight += wd.v[g]; }
      offset_total += (lat.Nb - nb) * wd.offset + (lat.N / nv) * 0.25 * nv * d;
      p_off_d = (wd.v[0] + wd.v[1] > 1e-10) ? wd.v[0] / (wd.v[0] + wd.v[1]) : 1;
    }
    // transverse field -hx S^x: the site graph (weight and offset |hx|/2,
    // weight_impl.h:63-88) at every site, drawn with the bond graphs in one
    // table by weight (the model's choose_graph over all locations)
    if (hx != 0) {
      const int nb = lat.Nb;
      lat.add_site_bonds();
      const double vs = std::fabs(hx) / 2;
      for (int s = 0; s < lat.N; ++s) { wl.push_back(vs); graph_of.push_back({nb + s, 0}); graph_weight += vs; }
      offset_total += lat.N * vs;
    }
    if (!wl.empty()) alias.init(wl);
    if (graph_weight < 1e-20) graph_weight = 1e-20;
    p_off = (w.v[0] + w.v[1] > 1e-10) ? w.v[0] / (w.v[0] + w.v[1]) : 1;
    spins.assign(lat.N, 0);
    spins_c.assign(lat.N, 0);
    current.assign(lat.N, 0);
  }

  void sweep(double* rec) {
    const int N = lat.N;
    if (sched.on()) beta = 1.0 / sched.T((int)mcs);
    auto rnd = [&]() { return u01(eng); };
    std::swap(ops, ops_p);
    ops.clear();
    LooperOp sentinel{0, 0, 1.0, 0, 0};
    ops_p.push_back(sentinel);
    spins_c = spins;
    // fill times (path_integral.C:403-425)
    std::exponential_distribution<double> expdist(beta * graph_weight);
    times.clear();
    double t = 0;
    while (t < 1) { t += expdist(eng); times.push_back(t); }
    // fragments (427-453)
    frag.assign(N + ops_p.size() + times.size(), UFNode());
    for (int s = 0; s < N; ++s) current[s] = s;
    int fid = N;
    size_t tmi = 0;
    for (size_t opi = 0; opi < ops_p.size();) {
      if (times[tmi] < ops_p[opi].time) {
        auto bg = graph_of[alias(rnd)];
        int b = bg.first, g = bg.second;
        // a site graph is compatible with either spin (graph_impl.h:66-69)
        if (lat.is_site_bond(b) || is_compatible(g, spins_c[lat.src[b]], spins_c[lat.tgt[b]])) {
          ops.push_back(LooperOp{b, g << 2, times[tmi], 0, 0});
          ++tmi;
        } else {
          ++tmi;
          continue;
        }
      } else {
        if ((ops_p[opi].type & 1) == 0) { ++opi; continue; }
        ops.push_back(ops_p[opi]);
        ++opi;
      }
      LooperOp& oi = ops.back();
      int b = oi.bond, s0 = lat.src[b], s1 = lat.tgt[b];
      if (lat.is_site_bond(b)) {
        // site operator (path_integral.C:555-563): the world line is cut,
        // a new fragment starts above (graph_impl.h:78-84)
        if (oi.type & 1) spins_c[s0] ^= 1;
        frag[fid] = UFNode();
        oi.loop0 = current[s0];
        oi.loop1 = current[s0] = fid;
        ++fid;
        continue;
      }
      if (oi.type & 1) {
        int g = (rnd() < (lat.is_d_bond(b) ? p_off_d : p_off)) ? 0 : 1;    // choose_offdiagonal, graph_impl.h:323-327
        oi.type = (oi.type & 1) | (g << 2);
        spins_c[s0] ^= 1;
        spins_c[s1] ^= 1;
      }
      Reconnect r = reconnect(frag, fid, oi.type >> 2, current[s0], current[s1]);
      fid = r.fid; current[s0] = r.curr0; current[s1] = r.curr1;
      oi.loop0 = r.loop0; oi.loop1 = r.loop1;
    }
    // symmetrize (583-615): S=1/2 closes each world line through tau = 1;
    // virtual sites join a real site's tau = 0 ends to its tau = 1 ends
    // through a random spin-conserving permutation
    if (lat.nv == 1) {
      for (int s = 0; s < N; ++s) uf_unify(frag, s, current[s]);
    } else {
      std::vector<int> perm(lat.nv);
      for (int off = 0; off < N; off += lat.nv) {
        for (int i = 0; i < lat.nv; ++i) perm[i] = i;
        partitioned_shuffle(perm.data(), lat.nv, &spins[off], &spins_c[off], rnd);
        for (int i = 0; i < lat.nv; ++i) uf_unify(frag, off + i, current[off + perm[i]]);
      }
    }
    // cluster ids (626-648)
    int nc = uf_set_id(frag, 0, fid, 0);
    uf_copy_id(frag, 0, fid);
    // accumulate (651-737)
    est.assign(nc, Estimate());
    spins_c = spins;
    for (int s = 0; s < N; ++s) est_start_bottom(est[frag[s].id_], lat.gauge[s], spins_c[s]);
    for (auto& op : ops) {
      int g = op.type >> 2;
      if (lat.is_site_bond(op.bond)) {
        // end_s / begin_s (path_integral.C:714-723): the segment below ends
        // in loop0, the one above starts in loop1; no bond vector
        const int s = lat.src[op.bond];
        est_end(est[frag[op.loop0].id_], op.time, lat.gauge[s], spins_c[s]);
        if (op.type & 1) spins_c[s] ^= 1;
        est_begin(est[frag[op.loop1].id_], op.time, lat.gauge[s], spins_c[s]);
        continue;
      }
      if (is_frozen(g)) continue;
      int b = op.bond, s0 = lat.src[b], s1 = lat.tgt[b];
      double tt = op.time;
      Estimate& el0 = est[frag[loop_l0(g, op.loop0, op.loop1)].id_];
      Estimate& el1 = est[frag[loop_l1(g, op.loop0, op.loop1)].id_];
      est_end(el0, tt, lat.gauge[s0], spins_c[s0]);
      for (int d = 0; d < lat.dim; ++d) el0.winding[d] += (1 - 2 * spins_c[s0]) * lat.vr(b, d);
      est_end(el1, tt, lat.gauge[s1], spins_c[s1]);
      if (op.type & 1) { spins_c[s0] ^= 1; spins_c[s1] ^= 1; }
      Estimate& eu0 = est[frag[loop_u0(g, op.loop0, op.loop1)].id_];
      Estimate& eu1 = est[frag[loop_u1(g, op.loop0, op.loop1)].id_];
      est_begin(eu0, tt, lat.gauge[s0], spins_c[s0]);
      for (int d = 0; d < lat.dim; ++d) eu0.winding[d] -= (1 - 2 * spins_c[s0]) * lat.vr(b, d);
      est_begin(eu1, tt, lat.gauge[s1], spins_c[s1]);
    }
    for (int s = 0; s < N; ++s) est_end(est[frag[current[s]].id_], 1.0, lat.gauge[s], spins_c[s]);
    // collect (740-778)
    for (int i = 0; i < R_COUNT; ++i) rec[i] = 0;
    for (int c = 0; c < nc; ++c) collect(est[c], rec);
    rec[R_NOP] = (double)ops.size();
    rec[R_NC] = nc;
    rec[R_ENERGY] = offset_total - ops.size() / beta;
    rec[R_BETA] = beta;
    rec[R_SWEEP] = (double)mcs;
    int neg = 0;  // the sign (before the flip)
    for (const auto& op : ops) if ((op.type & 1) && sgn.negative(lat, op.bond)) ++neg;
    rec[R_SIGN] = (neg & 1) ? -1.0 : 1.0;
    // determine flip (781-803): (2r-1) < (FIELD ? tanh(beta w_c) : 0), with
    // the field weight w_c = -h umag_c (cluster.h:50-63); Zeeman energy of
    // the flipped cluster added to the energy (852-858)
    to_flip.assign(nc, 0);
    for (int c = 0; c < nc; ++c) {
      const double wc = -h * est[c].umag;
      to_flip[c] = (2 * rnd() - 1) < (h != 0 ? std::tanh(beta * wc) : 0.0);
      rec[R_EFIELD] += to_flip[c] ? -wc : wc;
    }
    rec[R_ENERGY] += rec[R_EFIELD];
    // flip operators & spins (806-825)
    for (auto& op : ops)
      if ((to_flip[frag[op.loop0].id_] ^ to_flip[frag[op.loop1].id_]) & 1) op.type ^= 1;
    for (int s = 0; s < N; ++s) if (to_flip[frag[s].id_]) spins[s] ^= 1;
    ++mcs;
  }

  // "loop; sse" (sse.C:174-400): the same loop clusters on the stochastic
  // series.  The operator string is walked once; before each kept operator
  // (and again after each insertion) a gap tries an insertion with
  // probability beta W / (n + 1), the graph drawn by weight and kept if
  // compatible; a diagonal operator is removed with probability
  // n / (beta W), else its graph is re-chosen (choose_diagonal,
  // graph_impl.h:297-312); an off-diagonal one is regraphed
  // (choose_offdiagonal).  Times are positions in the new string (start at
  // 0, stop at n).  Lattice bonds only (no site operators, no single-ion
  // term, no field).  The record's tau-integrated sums are stored with
  // lengths in units of the string (n > 0): the squared ones (usize, umag,
  // ssize, smag) divided by n^2, the normal estimator's (umag_a, smag_a) by
  // n, so that the susceptibility is beta (n umag + umag2) / (n + 1) / V as
  // in susceptibility.h:213-251.
  void sweep_sse(double* rec) {
    const int N = lat.N;
    if (sched.on()) beta = 1.0 / sched.T((int)mcs);
    auto rnd = [&]() { return u01(eng); };
    int nop = (int)ops.size();
    std::swap(ops, ops_p);
    ops.clear();
    spins_c = spins;
    frag.assign(N, UFNode());
    for (int s = 0; s < N; ++s) current[s] = s;
    int fid = N;
    const double bw = beta * graph_weight;
    const double pd[2] = {(w.v[0] + w.v[2] > 1e-10) ? w.v[0] / (w.v[0] + w.v[2]) : 1,
                          (w.v[1] + w.v[3] > 1e-10) ? w.v[1] / (w.v[1] + w.v[3]) : 1};
    const double pdd[2] = {(wd.v[0] + wd.v[2] > 1e-10) ? wd.v[0] / (wd.v[0] + wd.v[2]) : 1,
                           (wd.v[1] + wd.v[3] > 1e-10) ? wd.v[1] / (wd.v[1] + wd.v[3]) : 1};
    bool try_gap = true;
    size_t opi = 0;
    while (try_gap || opi < ops_p.size()) {
      if (try_gap) {
        if ((nop + 1) * rnd() < bw) {
          auto bg = graph_of[alias(rnd)];
          const int b = bg.first, g = bg.second;
          if (lat.is_site_bond(b) || is_compatible(g, spins_c[lat.src[b]], spins_c[lat.tgt[b]])) {
            ops.push_back(LooperOp{b, g << 2, 0.0, 0, 0});
            ++nop;
          } else {
            try_gap = false;
            continue;
          }
        } else {
          try_gap = false;
          continue;
        }
      } else {
        LooperOp o = ops_p[opi];
        const int c0 = spins_c[lat.src[o.bond]], c1 = spins_c[lat.tgt[o.bond]];
        const bool site = lat.is_site_bond(o.bond), insite = lat.is_d_bond(o.bond);
        if ((o.type & 1) == 0) {
          if (bw * rnd() < nop) {
            --nop;
            ++opi;
            continue;
          }
          const int c = 1 ^ c0 ^ c1;  // 0 antiparallel, 1 parallel
          const double p = insite ? pdd[c] : pd[c];
          o.type = site ? 0 : (((rnd() < p ? 0 : 2) ^ c) << 2);  // the site graph stays
        } else if (!site) {
          o.type = 1 | ((rnd() < (insite ? p_off_d : p_off) ? 0 : 1) << 2);
        }
        ops.push_back(o);
        ++opi;
        try_gap = true;
      }
      LooperOp& oi = ops.back();
      const int b = oi.bond, s0 = lat.src[b], s1 = lat.tgt[b];
      if ((int)frag.size() < fid + 1) frag.resize(fid + 1);
      if (lat.is_site_bond(b)) {  // the site graph: the world line is cut (graph_impl.h:78-84)
        if (oi.type & 1) spins_c[s0] ^= 1;
        frag[fid] = UFNode();
        oi.loop0 = current[s0];
        oi.loop1 = current[s0] = fid;
        ++fid;
        try_gap = true;
        continue;
      }
      if (oi.type & 1) { spins_c[s0] ^= 1; spins_c[s1] ^= 1; }
      Reconnect r = reconnect(frag, fid, oi.type >> 2, current[s0], current[s1]);
      fid = r.fid; current[s0] = r.curr0; current[s1] = r.curr1;
      oi.loop0 = r.loop0; oi.loop1 = r.loop1;
    }
    const int n = (int)ops.size();
    for (int k = 0; k < n; ++k) ops[k].time = k;
    if (lat.nv == 1) {
      for (int s = 0; s < N; ++s) uf_unify(frag, s, current[s]);
    } else {
      std::vector<int> perm(lat.nv);
      for (int off = 0; off < N; off += lat.nv) {
        for (int i = 0; i < lat.nv; ++i) perm[i] = i;
        partitioned_shuffle(perm.data(), lat.nv, &spins[off], &spins_c[off], rnd);
        for (int i = 0; i < lat.nv; ++i) uf_unify(frag, off + i, current[off + perm[i]]);
      }
    }
    int nc = uf_set_id(frag, 0, fid, 0);
    uf_copy_id(frag, 0, fid);
    est.assign(nc, Estimate());
    spins_c = spins;
    for (int s = 0; s < N; ++s) est_start_bottom(est[frag[s].id_], lat.gauge[s], spins_c[s]);
    for (auto& op : ops) {
      const int g = op.type >> 2;
      if (lat.is_site_bond(op.bond)) {  // end_s / begin_s, no bond vector
        const int s = lat.src[op.bond];
        est_end(est[frag[op.loop0].id_], op.time, lat.gauge[s], spins_c[s]);
        if (op.type & 1) spins_c[s] ^= 1;
        est_begin(est[frag[op.loop1].id_], op.time, lat.gauge[s], spins_c[s]);
        continue;
      }
      if (is_frozen(g)) continue;
      const int b = op.bond, s0 = lat.src[b], s1 = lat.tgt[b];
      const double tt = op.time;
      Estimate& el0 = est[frag[loop_l0(g, op.loop0, op.loop1)].id_];
      Estimate& el1 = est[frag[loop_l1(g, op.loop0, op.loop1)].id_];
      est_end(el0, tt, lat.gauge[s0], spins_c[s0]);
      for (int d = 0; d < lat.dim; ++d) el0.winding[d] += (1 - 2 * spins_c[s0]) * lat.vr(b, d);
      est_end(el1, tt, lat.gauge[s1], spins_c[s1]);
      if (op.type & 1) { spins_c[s0] ^= 1; spins_c[s1] ^= 1; }
      Estimate& eu0 = est[frag[loop_u0(g, op.loop0, op.loop1)].id_];
      Estimate& eu1 = est[frag[loop_u1(g, op.loop0, op.loop1)].id_];
      est_begin(eu0, tt, lat.gauge[s0], spins_c[s0]);
      for (int d = 0; d < lat.dim; ++d) eu0.winding[d] -= (1 - 2 * spins_c[s0]) * lat.vr(b, d);
      est_begin(eu1, tt, lat.gauge[s1], spins_c[s1]);
    }
    for (int s = 0; s < N; ++s) est_end(est[frag[current[s]].id_], (double)n, lat.gauge[s], spins_c[s]);
    for (int i = 0; i < R_COUNT; ++i) rec[i] = 0;
    for (int c = 0; c < nc; ++c) collect(est[c], rec);
    if (n > 0) {
      const double n2 = (double)n * n;
      rec[R_USIZE] /= n2; rec[R_UMAG] /= n2; rec[R_SSIZE] /= n2; rec[R_SMAG] /= n2;
      rec[R_UMAGA] /= n; rec[R_SMAGA] /= n;  // the normal estimator's sums
    }
    rec[R_NOP] = n;
    rec[R_NC] = nc;
    rec[R_ENERGY] = offset_total - n / beta;
    rec[R_BETA] = beta;
    rec[R_SWEEP] = (double)mcs;
    int neg = 0;  // the sign (before the flip)
    for (const auto& op : ops) if ((op.type & 1) && sgn.negative(lat, op.bond)) ++neg;
    rec[R_SIGN] = (neg & 1) ? -1.0 : 1.0;
    to_flip.assign(nc, 0);
    for (int c = 0; c < nc; ++c) to_flip[c] = (2 * rnd() - 1) < 0;
    for (auto& op : ops)
      if ((to_flip[frag[op.loop0].id_] ^ to_flip[frag[op.loop1].id_]) & 1) op.type ^= 1;
    for (int s = 0; s < N; ++s) if (to_flip[frag[s].id_]) spins[s] ^= 1;
    ++mcs;
  }
};

// ---------------------------------------------------------------------------
// Standalone mode: standalone/loop.C:39-180 consuming the RNG the same way.
// ---------------------------------------------------------------------------
struct SANode {  // standalone/union_find.h:44-57
  int parent_ = -1;
  int id_ = 0;
  bool is_root() const { return parent_ < 0; }
  int weight() const { return -parent_; }
};
int sa_root(const std::vector<SANode>& v, int g) { while (!v[g].is_root()) g = v[g].parent_; return g; }
int sa_unify(std::vector<SANode>& v, int g0, int g1) {  // unify_weight, :112-126
  int r0 = sa_root(v, g0), r1 = sa_root(v, g1);
  if (r0 != r1) {
    if (v[r0].weight() < v[r1].weight()) std::swap(r0, r1);
    v[r0].parent_ += v[r1].parent_;
    v[r1].parent_ = r0;
  }
  for (int g : {g0, g1}) { while (g != r0) { int p = v[g].parent_; v[g].parent_ = r0; g = p; } }
  return r0;
}

struct RunStat {  // standalone/observable.h
  double sum = 0, esq = 0; unsigned count = 0;
  void add(double x) { sum += x; esq += x * x; ++count; }
  double mean() const { return count ? sum / count : 0; }
  double error() const { return count > 1 ? std::sqrt((esq / count - mean() * mean()) / (count - 1)) : 0; }
};

void standalone_run(unsigned L, double T, unsigned sweeps, unsigned therm, double* mean5, double* err5) {
  const unsigned nsites = L, nbonds = L;
  const double beta = 1 / T;
  std::mt19937 eng(29833);
  std::uniform_real_distribution<> d_uniform;
  std::exponential_distribution<> d_time(beta * nbonds / 2);
  struct Op { int type; unsigned bond; unsigned upper, lower; double time; };
  std::vector<Op> ops, ops_p;
  std::vector<int> spins(nsites, 0), current(nsites);
  std::vector<SANode> frag;
  std::vector<char> to_flip;
  struct Est { double mag = 0, size = 0, length = 0; };
  std::vector<Est> est;
  RunStat o[5];
  auto left = [&](unsigned b) { return (int)b; };
  auto right = [&](unsigned b) { return (b == nbonds - 1) ? 0 : (int)b + 1; };
  for (unsigned mcs = 0; mcs < therm + sweeps; ++mcs) {
    std::swap(ops, ops_p); ops.clear();
    frag.clear(); frag.resize(nsites);
    for (unsigned s = 0; s < nsites; ++s) current[s] = s;
    double t = d_time(eng);
    for (size_t opi = 0; t < 1 || opi != ops_p.size();) {
      if (opi == ops_p.size() || t < ops_p[opi].time) {
        const int b = static_cast<int>(nbonds * d_uniform(eng));
        if (spins[left(b)] != spins[right(b)]) {
          ops.push_back(Op{0, (unsigned)b, 0, 0, t});
          t += d_time(eng);
        } else {
          t += d_time(eng);
          continue;
        }
      } else {
        if (ops_p[opi].type == 0) { ++opi; continue; }
        ops.push_back(ops_p[opi]);
        ++opi;
      }
      Op& oi = ops.back();
      const int s0 = left(oi.bond), s1 = right(oi.bond);
      if (oi.type == 1) { spins[s0] ^= 1; spins[s1] ^= 1; }
      oi.lower = sa_unify(frag, current[s0], current[s1]);
      frag.push_back(SANode());
      oi.upper = current[s0] = current[s1] = (unsigned)frag.size() - 1;
    }
    for (unsigned s = 0; s < nsites; ++s) sa_unify(frag, s, current[s]);
    int nc = 0;
    for (auto& f : frag) if (f.is_root()) f.id_ = nc++;
    for (auto& f : frag) f.id_ = frag[sa_root(frag, (int)(&f - &frag[0]))].id_;
    to_flip.resize(nc);
    est.clear(); est.resize(nc);
    for (auto& op : ops) {
      est[frag[op.lower].id_].length += 2 * op.time;
      est[frag[op.upper].id_].length -= 2 * op.time;
    }
    for (unsigned s = 0; s < nsites; ++s) {
      int id = frag[s].id_;
      est[id].mag += 1 - 2 * spins[s];
      est[id].size += 1;
      est[id].length += 1;
    }
    double usus = 0, smag = 0, ssus = 0;
    for (auto& e : est) { usus += e.mag * e.mag; smag += e.size * e.size; ssus += e.length * e.length; }
    for (int c = 0; c < nc; ++c) to_flip[c] = (d_uniform(eng) < 0.5);
    for (auto& op : ops)
      if (to_flip[frag[op.lower].id_] ^ to_flip[frag[op.upper].id_]) op.type ^= 1;
    for (unsigned s = 0; s < nsites; ++s) if (to_flip[frag[s].id_]) spins[s] ^= 1;
    if (mcs >= therm) {
      o[0].add(nc);
      o[1].add((0.25 * nbonds - ops.size() / beta) / nsites);
      o[2].add(0.25 * beta * usus / nsites);
      o[3].add(0.25 * smag);
      o[4].add(0.25 * beta * ssus / nsites);
    }
  }
  for (int i = 0; i < 5; ++i) { mean5[i] = o[i].mean(); err5[i] = o[i].error(); }
}

// ---------------------------------------------------------------------------
// Canonical configuration and partitions.
//
// A configuration is spins0[N] plus operators in canonical order (bond
// ascending, then time ascending; equal times keep their given order).
// Elements: 0..N-1 the world-line boundary at tau = 0 of each site,
// N + 4k + {0:l0, 1:l1, 2:u0, 3:u1} the legs of operator k; l0/u0 sit on the
// bond's source, l1/u1 on its target.  Events on a site are ordered by
// (time, bond, rank in bond).
// ---------------------------------------------------------------------------
struct SiteEvents {
  std::vector<int> off;           // per site: start in ev (size N+1)
  std::vector<int> ev_op;         // operator index of each event
  std::vector<int> ev_side;       // 0 = source side, 1 = target side
  std::vector<int> pos0, pos1;    // per op: position of its event on source / target site
};

SiteEvents build_site_events(const Lattice& lat, int nops, const int* bond, const double* time) {
  SiteEvents se;
  std::vector<std::vector<std::pair<int, int>>> per(lat.N);  // (op, side)
  for (int k = 0; k < nops; ++k) {
    per[lat.src[bond[k]]].push_back({k, 0});
    if (!lat.is_site_bond(bond[k])) per[lat.tgt[bond[k]]].push_back({k, 1});
  }
  se.off.assign(lat.N + 1, 0);
  se.pos0.assign(nops, -1);
  se.pos1.assign(nops, -1);
  for (int s = 0; s < lat.N; ++s) {
    auto& v = per[s];
    std::stable_sort(v.begin(), v.end(), [&](const std::pair<int, int>& a, const std::pair<int, int>& b) {
      if (time[a.first] != time[b.first]) return time[a.first] < time[b.first];
      if (bond[a.first] != bond[b.first]) return bond[a.first] < bond[b.first];
      return a.first < b.first;
    });
    se.off[s + 1] = se.off[s] + (int)v.size();
    for (size_t j = 0; j < v.size(); ++j) {
      se.ev_op.push_back(v[j].first);
      se.ev_side.push_back(v[j].second);
      if (v[j].second == 0) se.pos0[v[j].first] = (int)j; else se.pos1[v[j].first] = (int)j;
    }
  }
  for (int k = 0; k < nops; ++k)
    if (lat.is_site_bond(bond[k])) se.pos1[k] = se.pos0[k];  // both legs on the one site
  return se;
}

// Leg graph edges (intra-operator cut edges + world-line edges).
std::vector<std::pair<int, int>> leg_graph_edges(const Lattice& lat, int nops, const int* bond,
                                                 const double* time, const int* type) {
  int N = lat.N;
  std::vector<std::pair<int, int>> E;
  for (int k = 0; k < nops; ++k) {
    int g = type[k] >> 2;
    int l0 = N + 4 * k, l1 = l0 + 1, u0 = l0 + 2, u1 = l0 + 3;
    if (g & 2) { E.push_back({l0, l1}); E.push_back({u0, u1}); E.push_back({l0, u0}); }
    else if (g == 0) { E.push_back({l0, l1}); E.push_back({u0, u1}); }
    else { E.push_back({l0, u1}); E.push_back({l1, u0}); }
  }
  SiteEvents se = build_site_events(lat, nops, bond, time);
  for (int s = 0; s < N; ++s) {
    int a = se.off[s], b = se.off[s + 1];
    if (a == b) continue;
    auto lower = [&](int j) { return N + 4 * se.ev_op[j] + (se.ev_side[j] == 0 ? 0 : 1); };
    auto upper = [&](int j) { return N + 4 * se.ev_op[j] + (se.ev_side[j] == 0 ? 2 : 3); };
    E.push_back({s, lower(a)});
    for (int j = a; j + 1 < b; ++j) E.push_back({upper(j), lower(j + 1)});
    E.push_back({upper(b - 1), s});
  }
  return E;
}

void canonicalise(std::vector<int>& comp, int n) {
  // comp[i] = arbitrary component id -> min element index of the component
  std::vector<int> mn(n, INT32_MAX);
  for (int i = 0; i < n; ++i) mn[comp[i]] = std::min(mn[comp[i]], i);
  for (int i = 0; i < n; ++i) comp[i] = mn[comp[i]];
}

void bfs_labels(const Lattice& lat, int nops, const int* bond, const double* time, const int* type,
                int* labels) {
  int n = lat.N + 4 * nops;
  auto E = leg_graph_edges(lat, nops, bond, time, type);
  std::vector<std::vector<int>> adj(n);
  for (auto& e : E) { adj[e.first].push_back(e.second); adj[e.second].push_back(e.first); }
  std::vector<int> comp(n, -1);
  for (int i = 0; i < n; ++i) {
    if (comp[i] >= 0) continue;
    std::deque<int> q{i};
    comp[i] = i;
    while (!q.empty()) {
      int x = q.front(); q.pop_front();
      for (int y : adj[x]) if (comp[y] < 0) { comp[y] = i; q.push_back(y); }
    }
  }
  for (int i = 0; i < n; ++i) labels[i] = comp[i];  // BFS from increasing i -> min index already
}

// looper's fragment algorithm on a given configuration: operators visited in
// tau order ((time, bond, rank) as on a site), reconnect + unify, symmetrize,
// set_id/copy_id; labels mapped to canonical elements through loop0/loop1 and
// the leg accessors.
void looper_labels(const Lattice& lat, int nops, const int* bond, const double* time, const int* type,
                   int* labels, int compress_variant) {
  int N = lat.N;
  std::vector<int> order(nops);
  for (int k = 0; k < nops; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (time[a] != time[b]) return time[a] < time[b];
    if (bond[a] != bond[b]) return bond[a] < bond[b];
    return a < b;
  });
  std::vector<UFNode> frag(N + nops);
  std::vector<int> current(N), loop0(nops), loop1(nops);
  for (int s = 0; s < N; ++s) current[s] = s;
  int fid = N;
  for (int k : order) {
    int g = type[k] >> 2, b = bond[k], s0 = lat.src[b], s1 = lat.tgt[b];
    if (compress_variant && g == 0) {
      frag[fid] = UFNode();
      loop0[k] = uf_unify_compress(frag, current[s0], current[s1]);
      loop1[k] = current[s0] = current[s1] = fid++;
      continue;
    }
    if (compress_variant && (g & 2)) {
      int u = uf_unify_compress(frag, current[s0], current[s1]);
      loop0[k] = loop1[k] = current[s0] = current[s1] = u;
      continue;
    }
    Reconnect r = reconnect(frag, fid, g, current[s0], current[s1]);
    fid = r.fid; current[s0] = r.curr0; current[s1] = r.curr1;
    loop0[k] = r.loop0; loop1[k] = r.loop1;
  }
  for (int s = 0; s < N; ++s) {
    if (compress_variant) uf_unify_compress(frag, s, current[s]);
    else uf_unify(frag, s, current[s]);
  }
  int nc = uf_set_id(frag, 0, fid, 0);
  (void)nc;
  uf_copy_id(frag, 0, fid);
  int n = N + 4 * nops;
  std::vector<int> comp(n);
  for (int s = 0; s < N; ++s) comp[s] = frag[s].id_;
  for (int k = 0; k < nops; ++k) {
    int g = type[k] >> 2;
    comp[N + 4 * k + 0] = frag[loop_l0(g, loop0[k], loop1[k])].id_;
    comp[N + 4 * k + 1] = frag[loop_l1(g, loop0[k], loop1[k])].id_;
    comp[N + 4 * k + 2] = frag[loop_u0(g, loop0[k], loop1[k])].id_;
    comp[N + 4 * k + 3] = frag[loop_u1(g, loop0[k], loop1[k])].id_;
  }
  canonicalise(comp, n);
  for (int i = 0; i < n; ++i) labels[i] = comp[i];
}

// ---------------------------------------------------------------------------
// Philox mode: the device algorithm, sequential.
// ---------------------------------------------------------------------------
struct PhiloxChain {
  Lattice lat;
  SignRule sgn;
  Weights w, wd;     // lattice bonds; in-site bonds of the single-ion term
  Thresholds th, thd;
  double Wd = 0, scale_d = 0;
  double beta = 1, scale = 0, offset_total = 0, W = 0;
  double h = 0;   // longitudinal field
  double vs = 0, site_scale = 0;  // transverse field: site graph weight |hx|/2 and its candidate scale
  Schedule sched; // annealing: beta and candidate scale from T(sweep) before each sweep
  uint32_t key0 = 0, key1 = 0;
  uint32_t sweep = 0;
  std::vector<int> spins;                 // tau = 0 spins
  // operators, bond CSR in canonical order
  std::vector<int> boff;                  // Nb+1
  std::vector<double> otime;
  std::vector<int> otype;                 // offdiag bit | g << 2
  // previous configuration's site events (post-flip segment spins)
  std::vector<int> psoff;                 // N+1, offsets in segment space
  std::vector<int> pm;                    // number of events per site
  std::vector<double> pev_time;           // indexed by segment node (event j's time at psoff[s]+j)
  std::vector<int> pseg_spin;             // post-flip spin of each segment
  // last sweep's cluster data (for inspection)
  std::vector<int> last_root;             // per segment node (chain-local)
  std::vector<int> last_l0, last_l1;      // per op
  int last_nodes = 0;

  void init(int kind, int L, double b, double jxy, double jz, double fs, uint32_t seed, uint32_t chain,
            double field = 0, int nv = 1, double hx = 0, double d = 0) {
    h = field;
    lat.init(kind, L, nv);
    sgn.init(jxy, hx);
    w = xxz_weights(jxy, jz, fs);
    th = make_thresholds(w);
    // choose_diagonal (graph_impl.h:297-312): g = 0 w.p. v0 / (v0 + v2) for
    // an antiparallel pair, g = 1 w.p. v1 / (v1 + v3) for a parallel one
    for (int c = 0; c < 2; ++c) {
      const double den = w.v[c] + w.v[c + 2];
      pdiag[c] = (den > 0) ? (uint64_t)std::floor(std::ldexp(w.v[c] / den, 32)) : (1ull << 32);
    }
    beta = b;
    W = w.v[0] + w.v[1] + w.v[2] + w.v[3];
    scale = (W > 0) ? std::ldexp(1.0, 52) / (beta * W) : 0;
    offset_total = lat.Nb * w.offset;
    // single-ion term: in-site bonds with the weights of an Ising coupling
    // 2D, their own candidate rate and thresholds, and the constant D nv/4
    // per real site
    if (d != 0) {
      const int nb = lat.Nb;
      if (nv > 1) lat.add_d_bonds();
      wd = xxz_weights(0.0, 2 * d, fs);
      thd = make_thresholds(wd);
      Wd = wd.v[0] + wd.v[1] + wd.v[2] + wd.v[3];
      scale_d = (Wd > 0) ? std::ldexp(1.0, 52) / (beta * Wd) : 0;
      offset_total += (lat.Nb - nb) * wd.offset + (lat.N / nv) * 0.25 * nv * d;
    }
    // transverse field -hx S^x: one site graph of weight |hx|/2 and offset
    // |hx|/2 per site (weight_impl.h:63-88), its operators on the site bonds
    if (hx != 0) {
      lat.add_site_bonds();
      vs = std::fabs(hx) / 2;
      site_scale = std::ldexp(1.0, 52) / (beta * vs);
      offset_total += lat.N * vs;
    }
    sse_selection(lat.nb_real > lat.nb_lat ? d : 0.0, fs);
    key0 = seed; key1 = chain;
    spins.assign(lat.N, 0);
    boff.assign(lat.Nb + 1, 0);
    otime.clear(); otype.clear();
    reset_prev_from_ops();
  }

  // Rebuild the "previous configuration" site arrays from the current
  // operators and spins (used at start and after loading a configuration).
  void reset_prev_from_ops() {
    int nops = (int)otime.size();
    std::vector<int> bond(nops);
    for (int b = 0; b < lat.Nb; ++b) for (int k = boff[b]; k < boff[b + 1]; ++k) bond[k] = b;
    SiteEvents se = build_site_events(lat, nops, bond.data(), otime.data());
    psoff.assign(lat.N + 1, 0);
    pm.assign(lat.N, 0);
    for (int s = 0; s < lat.N; ++s) {
      pm[s] = se.off[s + 1] - se.off[s];
      psoff[s + 1] = psoff[s] + segments(pm[s]);
    }
    pev_time.assign(psoff[lat.N], 0);
    pseg_spin.assign(psoff[lat.N], 0);
    for (int s = 0; s < lat.N; ++s) {
      int c = spins[s];
      if (pm[s] == 0) { pseg_spin[psoff[s]] = c; continue; }
      for (int j = 0; j < pm[s]; ++j) {
        int k = se.ev_op[se.off[s] + j];
        pev_time[psoff[s] + j] = otime[k];
        pseg_spin[psoff[s] + j] = c;
        c ^= (otype[k] & 1);
      }
      if (split()) { pev_time[psoff[s] + pm[s]] = 1.0; pseg_spin[psoff[s] + pm[s]] = c; }
    }
  }

  // virtual sites: a site with m events owns m + 1 segments (from tau = 0,
  // between events, up to tau = 1); otherwise the lowest segment wraps
  // through tau = 1 and m segments remain (one for a site without events)
  bool split() const { return lat.nv > 1; }
  int segments(int m) const { return m == 0 ? 1 : (split() ? m + 1 : m); }

  // spin at tau on site s in the previous configuration; returns -1 on a tie
  int spin_at(int s, double tau) const {
    int m = pm[s], base = psoff[s];
    if (m == 0) return pseg_spin[base];
    const double* t = &pev_time[base];
    int j = (int)(std::lower_bound(t, t + m, tau) - t);
    if (j < m && t[j] == tau) return -1;
    return pseg_spin[base + (j == m ? (split() ? m : 0) : j)];
  }

  void sweep_once(double* rec) {
    const int Nb = lat.Nb;
    if (sched.on()) {
      beta = 1.0 / sched.T((int)sweep);
      scale = (W > 0) ? std::ldexp(1.0, 52) / (beta * W) : 0;
      if (vs > 0) site_scale = std::ldexp(1.0, 52) / (beta * vs);
      if (Wd > 0) scale_d = std::ldexp(1.0, 52) / (beta * Wd);
    }
    // (a) diagonal update, per bond
    std::vector<int> nboff(Nb + 1, 0);
    std::vector<double> ntime;
    std::vector<int> ntype;
    ntime.reserve(otime.size() + 16);
    ntype.reserve(otype.size() + 16);
    std::vector<double> acc_t;
    std::vector<int> acc_g;
    for (int b = 0; b < Nb; ++b) {
      int s0 = lat.src[b], s1 = lat.tgt[b];
      // a site bond's candidates: rate beta |hx|/2, the one site graph
      // (g = 0), compatible with either spin (graph_impl.h:66-69)
      const bool site = lat.is_site_bond(b);
      // an in-site bond of the single-ion term: its own rate and graphs
      const bool dbond = lat.is_d_bond(b);
      const Thresholds& tb = dbond ? thd : th;
      acc_t.clear(); acc_g.clear();
      if (site ? vs > 0 : (dbond ? Wd > 0 : W > 0)) {
        uint64_t T = 0;
        for (uint32_t k = 0;; ++k) {
          u32x4 x = philox4x32_10(sweep, STAGE_DIAG, (uint32_t)b, k, key0, key1);
          T += exp_gap(x, site ? site_scale : (dbond ? scale_d : scale));
          if (T >= TGRID) break;
          double tau = (double)T * (1.0 / 4503599627370496.0);
          int g = site ? 0 : choose_graph(x.v[2], tb);
          int c0 = spin_at(s0, tau), c1 = spin_at(s1, tau);
          if (c0 < 0 || c1 < 0) continue;
          if (site || is_compatible(g, c0, c1)) { acc_t.push_back(tau); acc_g.push_back(g); }
        }
      }
      // merge kept (regraphed) off-diagonal ops with accepted candidates
      int kept_idx = 0;
      size_t ai = 0;
      for (int r = boff[b]; r < boff[b + 1]; ++r) {
        if ((otype[r] & 1) == 0) continue;
        u32x4 x = philox4x32_10(sweep, STAGE_REGRAPH, (uint32_t)b, (uint32_t)kept_idx, key0, key1);
        int g = (site || (uint64_t)x.v[0] < tb.toff) ? 0 : 1;
        ++kept_idx;
        while (ai < acc_t.size() && acc_t[ai] < otime[r]) {
          ntime.push_back(acc_t[ai]); ntype.push_back(acc_g[ai] << 2); ++ai;
        }
        ntime.push_back(otime[r]); ntype.push_back(1 | (g << 2));
      }
      while (ai < acc_t.size()) { ntime.push_back(acc_t[ai]); ntype.push_back(acc_g[ai] << 2); ++ai; }
      nboff[b + 1] = (int)ntime.size();
    }
    boff.swap(nboff);
    otime.swap(ntime);
    otype.swap(ntype);
    cluster_stage(rec);
  }

  // "loop; sse" in the device's formulation: the operator string in series
  // order (bond, type), walked once by one thread per chain with one Philox
  // word set per step, Philox(sweep, 6, 0, step): a gap tries an insertion
  // when (n + 1) u < beta W_tot (u from the first 53 bits; the bond
  // uniform from word 2, the graph from word 3 by the thresholds, kept if
  // compatible, and the gap tried again), a kept diagonal operator is
  // removed when beta W_tot u < n (the next operator follows with no gap
  // between) or else regraphed by word 2 against the diagonal threshold of
  // its spin pair, an off-diagonal one by word 2 against the off-diagonal
  // threshold.  The new string becomes bond-major operators with times
  // p / n, the cluster stage runs on them, and the flipped types return
  // to the string.  Lattice bonds only.
  std::vector<int> sbond, stype;          // SSE operator string
  uint64_t pdiag[2] = {0, 0};             // choose_diagonal thresholds: antiparallel, parallel
  uint64_t dpdiag[2] = {0, 0};            // the same for the in-site bonds
  // insertions: the location's class and graph from word 2 (lattice bonds
  // below sse_cls[0], in-site bonds below sse_cls[1], site bonds above; the
  // graph against thresholds scaled into the class's range), the location
  // within the class from word 3
  uint64_t sse_cls[2] = {0, 0}, sse_tl[3] = {0, 0, 0}, sse_td[3] = {0, 0, 0};
  double sse_wtot = 0;

  void sse_selection(double dion, double fs) {
    const double wl = lat.nb_lat * W;
    const double wd = (lat.nb_real - lat.nb_lat) * Wd, ws = (lat.Nb - lat.nb_real) * vs;
    const double wt = wl + wd + ws;
    sse_wtot = wt;
    auto t32 = [&](double x) -> uint64_t {
      const double y = std::floor(std::ldexp(x / wt, 32));
      return y >= 4294967296.0 ? (1ull << 32) : (uint64_t)y;
    };
    sse_cls[0] = (wd + ws > 0) ? t32(wl) : (1ull << 32);
    sse_cls[1] = (ws > 0) ? t32(wl + wd) : (1ull << 32);
    int last = -1;
    for (int g = 0; g < 4; ++g) if (w.v[g] > 0) last = g;
    double cum = 0;
    for (int g = 0; g < 3; ++g) {
      cum += w.v[g];
      sse_tl[g] = (last < 0 || g >= last) ? sse_cls[0] : t32(lat.nb_lat * cum);
    }
    Weights vd = {{0, 0, 0, 0}, 0, 1};
    if (dion != 0) vd = xxz_weights(0.0, 2 * dion, fs);
    last = -1;
    for (int g = 0; g < 4; ++g) if (vd.v[g] > 0) last = g;
    cum = 0;
    for (int g = 0; g < 3; ++g) {
      cum += vd.v[g];
      sse_td[g] = (last < 0 || g >= last) ? sse_cls[1] : t32(wl + (lat.nb_real - lat.nb_lat) * cum);
    }
    for (int c = 0; c < 2; ++c) {
      const double den = vd.v[c] + vd.v[c + 2];
      dpdiag[c] = (den > 0) ? (uint64_t)std::floor(std::ldexp(vd.v[c] / den, 32)) : (1ull << 32);
    }
  }

  void sweep_sse(double* rec) {
    const int Nb = lat.Nb;
    if (sched.on()) beta = 1.0 / sched.T((int)sweep);
    const double bw = beta * sse_wtot;
    std::vector<int> spc(spins), nb, nt;
    int nop = (int)sbond.size();
    uint32_t step = 0;
    bool try_gap = true;
    size_t i = 0;
    while (try_gap || i < sbond.size()) {
      const u32x4 x = philox4x32_10(sweep, STAGE_SSE, 0, step++, key0, key1);
      const uint64_t u53 = ((uint64_t)x.v[0] << 21) | (uint64_t)(x.v[1] >> 11);
      const double u = (double)(u53 + 1) * (1.0 / 9007199254740992.0);
      if (try_gap) {
        try_gap = false;
        if ((double)(nop + 1) * u < bw) {
          const uint64_t z = x.v[2];
          int b, g = 0;
          if (z < sse_cls[0]) {
            b = (int)(((uint64_t)x.v[3] * (uint64_t)lat.nb_lat) >> 32);
            while (g < 3 && z >= sse_tl[g]) ++g;
          } else if (z < sse_cls[1]) {
            b = lat.nb_lat + (int)(((uint64_t)x.v[3] * (uint64_t)(lat.nb_real - lat.nb_lat)) >> 32);
            while (g < 3 && z >= sse_td[g]) ++g;
          } else {
            b = lat.nb_real + (int)(((uint64_t)x.v[3] * (uint64_t)lat.N) >> 32);
          }
          if (lat.is_site_bond(b) || is_compatible(g, spc[lat.src[b]], spc[lat.tgt[b]])) {
            nb.push_back(b); nt.push_back(g << 2);
            ++nop;
            try_gap = true;
          }
        }
        continue;
      }
      const int b = sbond[i], t = stype[i], s0 = lat.src[b], s1 = lat.tgt[b];
      const bool site = lat.is_site_bond(b), insite = lat.is_d_bond(b);
      ++i;
      if ((t & 1) == 0) {
        if (bw * u < (double)nop) { --nop; continue; }
        const int c = 1 ^ spc[s0] ^ spc[s1];
        const uint64_t pd = insite ? dpdiag[c] : pdiag[c];
        nb.push_back(b); nt.push_back(site ? 0 : (((((uint64_t)x.v[2] < pd) ? 0 : 2) ^ c) << 2));
      } else {
        const uint64_t to = insite ? thd.toff : th.toff;
        nb.push_back(b); nt.push_back(1 | (site ? 0 : ((((uint64_t)x.v[2] < to) ? 0 : 1) << 2)));
        spc[s0] ^= 1;
        if (!site) spc[s1] ^= 1;  // a site operator flips its one spin
      }
      try_gap = true;
    }
    const int n = (int)nb.size();
    std::vector<int> cnt(Nb + 1, 0), perm(n);
    for (int p = 0; p < n; ++p) ++cnt[nb[p] + 1];
    boff.assign(Nb + 1, 0);
    for (int b = 0; b < Nb; ++b) boff[b + 1] = boff[b] + cnt[b + 1];
    otime.assign(n, 0);
    otype.assign(n, 0);
    std::vector<int> cur(boff.begin(), boff.end() - 1);
    for (int p = 0; p < n; ++p) {
      const int k = cur[nb[p]]++;
      otime[k] = (double)p / (double)n;
      otype[k] = nt[p];
      perm[p] = k;
    }
    cluster_stage(rec);
    for (int p = 0; p < n; ++p) nt[p] = otype[perm[p]];
    sbond.swap(nb);
    stype.swap(nt);
  }

  // the cluster stage on the operators in boff / otime / otype (linking,
  // union-find, estimators, flips, the record); advances the sweep counter
  void cluster_stage(double* rec) {
    const int N = lat.N, Nb = lat.Nb;
    int nops = (int)otime.size();
    std::vector<int> bond(nops);
    for (int b = 0; b < Nb; ++b) for (int k = boff[b]; k < boff[b + 1]; ++k) bond[k] = b;

    // (b) world-line linking: segments as nodes
    SiteEvents se = build_site_events(lat, nops, bond.data(), otime.data());
    std::vector<int> soff(N + 1, 0), m(N);
    for (int s = 0; s < N; ++s) { m[s] = se.off[s + 1] - se.off[s]; soff[s + 1] = soff[s] + segments(m[s]); }
    int M = soff[N];
    std::vector<int> seg_spin(M), seg_site(M);
    std::vector<double> ev_time(M, 0);
    for (int s = 0; s < N; ++s) {
      int c = spins[s];
      if (m[s] == 0) { seg_spin[soff[s]] = c; seg_site[soff[s]] = s; continue; }
      for (int j = 0; j < m[s]; ++j) {
        int k = se.ev_op[se.off[s] + j];
        ev_time[soff[s] + j] = otime[k];
        seg_spin[soff[s] + j] = c;
        seg_site[soff[s] + j] = s;
        c ^= (otype[k] & 1);
      }
      if (split()) {  // top segment, from the last event up to tau = 1
        ev_time[soff[s] + m[s]] = 1.0;
        seg_spin[soff[s] + m[s]] = c;
        seg_site[soff[s] + m[s]] = s;
      }
    }
    auto node = [&](int s, int j) { return soff[s] + j; };
    auto next = [&](int s, int i) { return (i + 1 == soff[s + 1]) ? soff[s] : i + 1; };
    last_l0.assign(nops, 0); last_l1.assign(nops, 0);
    // (c) union-find with min-index roots (any correct algorithm gives the
    // same partition; roots are the minimum node of each cluster)
    std::vector<int> parent(M);
    for (int i = 0; i < M; ++i) parent[i] = i;
    auto find = [&](int x) { while (parent[x] != x) { parent[x] = parent[parent[x]]; x = parent[x]; } return x; };
    auto unite = [&](int a, int b) {
      a = find(a); b = find(b);
      if (a == b) return;
      if (a < b) parent[b] = a; else parent[a] = b;
    };
    for (int k = 0; k < nops; ++k) {
      int b = bond[k], s0 = lat.src[b], s1 = lat.tgt[b];
      int l0 = node(s0, se.pos0[k]), l1 = node(s1, se.pos1[k]);
      int u0 = next(s0, l0), u1 = next(s1, l1);
      last_l0[k] = l0; last_l1[k] = l1;
      int g = otype[k] >> 2;
      if (g & 2) { unite(l0, l1); unite(u0, u1); unite(l0, u0); }
      else if (g == 0) { unite(l0, l1); unite(u0, u1); }
      else { unite(l0, u1); unite(l1, u0); }
    }
    // virtual sites (S > 1/2): each real site's tau = 0 ends (lowest
    // segments) meet its tau = 1 ends (top segments) through a
    // spin-conserving permutation: for u = 0, 1 in turn, the bottoms of spin
    // u in virtual-site order are matched to the tops of spin u in
    // virtual-site order permuted by Fisher-Yates (i = n_u - 1 .. 1, j = i -
    // (x mod (i + 1)), x the next word of the real site's stream: word k is
    // word k mod 4 of Philox(sweep, STAGE_BOUNDARY, real site, k / 4); swap
    // positions i and j)
    if (split()) {
      const int nv = lat.nv;
      for (int r = 0; r < N / nv; ++r) {
        auto word = [&](int k) {
          return philox4x32_10(sweep, STAGE_BOUNDARY, (uint32_t)r, (uint32_t)(k / 4), key0, key1).v[k % 4];
        };
        int k = 0;
        for (int u = 0; u < 2; ++u) {
          std::vector<int> bots, tops;
          for (int a = 0; a < nv; ++a) {
            const int v = r * nv + a;
            if (seg_spin[soff[v]] == u) bots.push_back(soff[v]);
            if (seg_spin[soff[v + 1] - 1] == u) tops.push_back(soff[v + 1] - 1);
          }
          if (bots.size() != tops.size()) throw std::logic_error("virtual site spins not conserved");
          for (int i = (int)tops.size() - 1; i >= 1; --i) {
            const int j = i - (int)(word(k++) % (uint32_t)(i + 1));
            std::swap(tops[i], tops[j]);
          }
          for (size_t i = 0; i < bots.size(); ++i) unite(bots[i], tops[i]);
        }
      }
    }
    for (int i = 0; i < M; ++i) parent[i] = find(i);
    last_root = parent;
    last_nodes = M;
    // (e) improved estimators per segment
    std::vector<Estimate> acc(M);
    for (int s = 0; s < N; ++s) {
      int gs = lat.gauge[s];
      if (m[s] == 0) {
        int i = soff[s];
        Estimate& e = acc[parent[i]];
        int c = seg_spin[i];
        est_start_bottom(e, gs, c);
        est_end(e, 1.0, gs, c);
        continue;
      }
      const int nseg = segments(m[s]);
      for (int j = 0; j < nseg; ++j) {
        int i = soff[s] + j;
        Estimate& e = acc[parent[i]];
        int c = seg_spin[i];
        // segment j lies between events j-1 and j (wrapping: segment 0 spans
        // the last event to tau = 1 and tau = 0 to the first event; split:
        // segment 0 starts at tau = 0 and segment m ends at tau = 1)
        const double tlo = (j > 0) ? ev_time[i - 1] : (split() ? 0.0 : ev_time[soff[s] + m[s] - 1] - 1.0);
        const double thi = ev_time[i];
        double len = (j == 0 && !split()) ? (thi + 1.0 - ev_time[soff[s] + m[s] - 1]) : (thi - tlo);
        if (j == 0) {
          e.usize0 += 0.5; e.umag0 += 0.5 - c; e.ssize0 += gs * 0.5; e.smag0 += gs * (0.5 - c);
        }
        e.usize += 0.5 * len; e.umag += (0.5 - c) * len;
        e.ssize += gs * 0.5 * len; e.smag += gs * (0.5 - c) * len;
        // winding: this segment is the lower leg of event j (if any) and the
        // upper leg of event j-1 (if any; event m-1 for a wrapping segment 0)
        if (j < m[s]) {
          int ka = se.ev_op[se.off[s] + j], sa = se.ev_side[se.off[s] + j];
          if (sa == 0) for (int d = 0; d < lat.dim; ++d) e.winding[d] += (1 - 2 * c) * lat.vr(bond[ka], d);
        }
        if (j > 0 || !split()) {
          int jb = (j == 0) ? m[s] - 1 : j - 1;
          int kb = se.ev_op[se.off[s] + jb], sb = se.ev_side[se.off[s] + jb];
          if (sb == 0) for (int d = 0; d < lat.dim; ++d) e.winding[d] -= (1 - 2 * c) * lat.vr(bond[kb], d);
        }
      }
    }
    for (int i = 0; i < R_COUNT; ++i) rec[i] = 0;
    int nc = 0;
    for (int i = 0; i < M; ++i) if (parent[i] == i) { collect(acc[i], rec); ++nc; }
    rec[R_NOP] = nops;
    rec[R_NC] = nc;
    rec[R_ENERGY] = offset_total - nops / beta;
    rec[R_BETA] = beta;
    rec[R_SWEEP] = (double)sweep;
    int neg = 0;  // the sign (before the flip)
    for (int k = 0; k < nops; ++k) if ((otype[k] & 1) && sgn.negative(lat, bond[k])) ++neg;
    rec[R_SIGN] = (neg & 1) ? -1.0 : 1.0;
    // (d) flip: one bit per root; with a field the cluster flips iff
    // 2u - 1 < tanh(beta w), w = -h umag, u from the first 53 bits
    std::vector<char> fb(M, 0);
    for (int i = 0; i < M; ++i) {
      if (parent[i] != i) continue;
      u32x4 r = philox4x32_10(sweep, STAGE_FLIP, (uint32_t)i, 0, key0, key1);
      if (h == 0) {
        fb[i] = r.v[0] & 1;
      } else {
        const uint64_t u53 = ((uint64_t)r.v[0] << 21) | (uint64_t)(r.v[1] >> 11);
        const double u = (double)(u53 + 1) * (1.0 / 9007199254740992.0);
        const double wc = -h * acc[i].umag;
        fb[i] = (2.0 * u - 1.0) < det_tanh(beta * wc);
        rec[R_EFIELD] += fb[i] ? -wc : wc;
      }
    }
    rec[R_ENERGY] += rec[R_EFIELD];
    for (int k = 0; k < nops; ++k) {
      int b = bond[k], s0 = lat.src[b];
      int l0 = last_l0[k], u0 = next(s0, l0);
      if (fb[parent[l0]] ^ fb[parent[u0]]) otype[k] ^= 1;
    }
    for (int s = 0; s < N; ++s) if (fb[parent[soff[s]]]) spins[s] ^= 1;
    for (int i = 0; i < M; ++i) seg_spin[i] ^= fb[parent[i]];
    // the flipped configuration becomes "previous" for the next sweep
    psoff = soff; pm = m; pev_time = ev_time; pseg_spin = seg_spin;
    ++sweep;
  }
};

}  // namespace

// ===========================================================================
// C ABI for ctypes
// ===========================================================================
extern "C" {

int or_record_size() { return R_COUNT; }

void or_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  u32x4 x = philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = x.v[i];
}

void or_det_log(const double* x, double* y, int n) { for (int i = 0; i < n; ++i) y[i] = det_log(x[i]); }
void or_det_tanh(const double* x, double* y, int n) { for (int i = 0; i < n; ++i) y[i] = det_tanh(x[i]); }
void or_det_exp(const double* x, double* y, int n) { for (int i = 0; i < n; ++i) y[i] = det_exp(x[i]); }

void or_exp_gap(const uint32_t* x4, double scale, uint64_t* gap) {
  u32x4 x = {{x4[0], x4[1], x4[2], x4[3]}};
  *gap = exp_gap(x, scale);
}

void or_xxz_weights(double jxy, double jz, double force_scatter, double* v4, double* offset, int* sign) {
  Weights w = xxz_weights(jxy, jz, force_scatter);
  for (int g = 0; g < 4; ++g) v4[g] = w.v[g];
  *offset = w.offset;
  *sign = w.sign;
}

void or_thresholds(double jxy, double jz, double force_scatter, uint64_t* thr4, uint64_t* toff) {
  Thresholds t = make_thresholds(xxz_weights(jxy, jz, force_scatter));
  for (int g = 0; g < 4; ++g) thr4[g] = t.thr[g];
  *toff = t.toff;
}

int or_lattice_info(int kind, int L, int nv, int* N, int* Nb, int* dim) {
  if ((kind != 0 && kind != 1) || L < 2 || nv < 1 || nv > 8) return -1;
  Lattice lat; lat.init(kind, L, nv);
  *N = lat.N; *Nb = lat.Nb; *dim = lat.dim;
  return 0;
}
int or_lattice_bonds(int kind, int L, int nv, int* src, int* tgt, int* gauge) {
  Lattice lat; lat.init(kind, L, nv);
  for (int b = 0; b < lat.Nb; ++b) { src[b] = lat.src[b]; tgt[b] = lat.tgt[b]; }
  for (int s = 0; s < lat.N; ++s) gauge[s] = lat.gauge[s];
  return 0;
}

// union-find on an explicit pair list (test/union_find.C semantics):
// returns parent (-1 for roots), root, weight (roots only) per node.
void or_uf_pairs(int n, int npairs, const int* pairs, int variant, int* parent, int* root, int* weight) {
  std::vector<UFNode> v(n);
  for (int i = 0; i < npairs; ++i) {
    if (variant == 0) uf_unify(v, pairs[2 * i], pairs[2 * i + 1]);
    else uf_unify_compress(v, pairs[2 * i], pairs[2 * i + 1]);
  }
  for (int i = 0; i < n; ++i) {
    parent[i] = v[i].is_root() ? -1 : v[i].parent();
    root[i] = uf_root_index(v, i);
    weight[i] = v[i].is_root() ? v[i].weight() : 0;
  }
}

// ---- looper mode ---------------------------------------------------------
void* or_looper_create(int kind, int L, double beta, double jxy, double jz, double fs, uint32_t seed, double h,
                       int nv, double hx, double d) {
  LooperChain* c = new LooperChain();
  c->init(kind, L, beta, jxy, jz, fs, seed, h, nv, hx, d);
  return c;
}
void or_looper_sweep_sse(void* h, int n, double* rec) {
  LooperChain* c = (LooperChain*)h;
  for (int i = 0; i < n; ++i) c->sweep_sse(rec + (size_t)i * R_COUNT);
}
void or_looper_sweep(void* h, int n, double* rec) {
  LooperChain* c = (LooperChain*)h;
  for (int i = 0; i < n; ++i) c->sweep(rec + (size_t)i * R_COUNT);
}
void or_looper_destroy(void* h) { delete (LooperChain*)h; }

// ---- standalone mode -----------------------------------------------------
void or_standalone_run(unsigned L, double T, unsigned sweeps, unsigned therm, double* mean5, double* err5) {
  standalone_run(L, T, sweeps, therm, mean5, err5);
}

// ---- partitions ------------------------------------------------------------
int or_bfs_labels(int kind, int L, int nops, const int* bond, const double* time, const int* type,
                  int* labels) {
  Lattice lat; lat.init(kind, L);
  bfs_labels(lat, nops, bond, time, type, labels);
  return 0;
}
int or_looper_labels(int kind, int L, int nops, const int* bond, const double* time, const int* type,
                     int compress_variant, int* labels) {
  Lattice lat; lat.init(kind, L);
  looper_labels(lat, nops, bond, time, type, labels, compress_variant);
  return 0;
}
// Segment-node index of each element (boundary s -> its segment 0; leg ->
// the world-line segment it touches) for a canonical configuration.
int or_segment_of_elements(int kind, int L, int nops, const int* bond, const double* time, int* seg) {
  Lattice lat; lat.init(kind, L);
  SiteEvents se = build_site_events(lat, nops, bond, time);
  std::vector<int> soff(lat.N + 1, 0);
  for (int s = 0; s < lat.N; ++s) soff[s + 1] = soff[s] + std::max(se.off[s + 1] - se.off[s], 1);
  for (int s = 0; s < lat.N; ++s) seg[s] = soff[s];
  for (int k = 0; k < nops; ++k) {
    int s0 = lat.src[bond[k]], s1 = lat.tgt[bond[k]];
    int l0 = soff[s0] + se.pos0[k], l1 = soff[s1] + se.pos1[k];
    int u0 = (l0 + 1 == soff[s0 + 1]) ? soff[s0] : l0 + 1;
    int u1 = (l1 + 1 == soff[s1 + 1]) ? soff[s1] : l1 + 1;
    seg[lat.N + 4 * k + 0] = l0; seg[lat.N + 4 * k + 1] = l1;
    seg[lat.N + 4 * k + 2] = u0; seg[lat.N + 4 * k + 3] = u1;
  }
  return soff[lat.N];
}

// ---- philox mode -------------------------------------------------------------
void* or_philox_create(int kind, int L, double beta, double jxy, double jz, double fs, double h, uint32_t seed,
                       uint32_t chain, int nv, double hx, double d) {
  PhiloxChain* c = new PhiloxChain();
  c->init(kind, L, beta, jxy, jz, fs, seed, chain, h, nv, hx, d);
  return c;
}
void or_philox_sweep_sse(void* h, int n, double* rec) {
  PhiloxChain* c = (PhiloxChain*)h;
  for (int i = 0; i < n; ++i) c->sweep_sse(rec + (size_t)i * R_COUNT);
}
void or_philox_sweep(void* h, int n, double* rec) {
  PhiloxChain* c = (PhiloxChain*)h;
  for (int i = 0; i < n; ++i) c->sweep_once(rec + (size_t)i * R_COUNT);
}
// new inverse temperature for the following sweeps (replica exchange): the
// candidate scale of a chain created at that beta
void or_philox_set_beta(void* h, double beta) {
  PhiloxChain* c = (PhiloxChain*)h;
  c->beta = beta;
  c->scale = (c->W > 0) ? std::ldexp(1.0, 52) / (beta * c->W) : 0;
  if (c->vs > 0) c->site_scale = std::ldexp(1.0, 52) / (beta * c->vs);
  if (c->Wd > 0) c->scale_d = std::ldexp(1.0, 52) / (beta * c->Wd);
}
int or_philox_num_ops(void* h) { return (int)((PhiloxChain*)h)->otime.size(); }
int or_philox_num_nodes(void* h) { return ((PhiloxChain*)h)->last_nodes; }
// state after the last sweep: spins[N], bond offsets[Nb+1], times[nops], types[nops]
void or_philox_state(void* h, int* spins, int* boff, double* time, int* type) {
  PhiloxChain* c = (PhiloxChain*)h;
  for (int s = 0; s < c->lat.N; ++s) spins[s] = c->spins[s];
  for (int b = 0; b <= c->lat.Nb; ++b) boff[b] = c->boff[b];
  for (size_t k = 0; k < c->otime.size(); ++k) { time[k] = c->otime[k]; type[k] = c->otype[k]; }
}
// roots (chain-local segment nodes) of the last sweep, before the flip
void or_philox_roots(void* h, int* root) {
  PhiloxChain* c = (PhiloxChain*)h;
  for (int i = 0; i < c->last_nodes; ++i) root[i] = c->last_root[i];
}
// load spins + operators (canonical order) and the sweep counter
void or_philox_load(void* h, const int* spins, const int* boff, const double* time, const int* type,
                    uint32_t sweep) {
  PhiloxChain* c = (PhiloxChain*)h;
  for (int s = 0; s < c->lat.N; ++s) c->spins[s] = spins[s];
  for (int b = 0; b <= c->lat.Nb; ++b) c->boff[b] = boff[b];
  int n = boff[c->lat.Nb];
  c->otime.assign(time, time + n);
  c->otype.assign(type, type + n);
  c->sweep = sweep;
  c->reset_prev_from_ops();
}
void or_philox_destroy(void* h) { delete (PhiloxChain*)h; }

// ---- boundary permutation (test/permutation.C's two tests) -----------------
// mt19937(seed), uniforms x / 2^32; test 1: trial1 shuffles of 0..n-1 into
// out1; test 2: per trial, conf0[i] = (int)(2u), conf1 = shuffled conf0, perm =
// the partitioned shuffle of the identity guided by (conf0, conf1)
void or_permutation_test(uint32_t seed, int n, int trial1, int trial2, int* out1, int* conf0, int* conf1,
                         int* perm) {
  std::mt19937 eng(seed);
  auto u01 = [&]() { return (double)eng() / 4294967296.0; };
  for (int t = 0; t < trial1; ++t) {
    int* r = out1 + t * n;
    for (int j = 0; j < n; ++j) r[j] = j;
    shuffle_range(r, n, u01);
  }
  for (int t = 0; t < trial2; ++t) {
    int* c0 = conf0 + t * n;
    int* c1 = conf1 + t * n;
    int* p = perm + t * n;
    for (int i = 0; i < n; ++i) c0[i] = (int)(2 * u01());
    for (int i = 0; i < n; ++i) c1[i] = c0[i];
    shuffle_range(c1, n, u01);
    for (int j = 0; j < n; ++j) p[j] = j;
    partitioned_shuffle(p, n, c0, c1, u01);
  }
}

// ---- annealing -------------------------------------------------------------
static Schedule make_schedule(int n, const int* steps, const double* temps) {
  Schedule s;
  s.step.assign(steps, steps + n);
  s.temp.assign(temps, temps + n);
  return s;
}
double or_schedule_temperature(int n, const int* steps, const double* temps, int step) {
  return make_schedule(n, steps, temps).T(step);
}
void or_philox_set_schedule(void* h, int n, const int* steps, const double* temps) {
  ((PhiloxChain*)h)->sched = make_schedule(n, steps, temps);
}
void or_looper_set_schedule(void* h, int n, const int* steps, const double* temps) {
  ((LooperChain*)h)->sched = make_schedule(n, steps, temps);
}

}  // extern "C"
