// Device/host shared definitions for the gfx950 loop-update kernels.
//
// Layout and conventions (shared by every kernel):
//  * A batch holds independent chains.  Their sites, bonds, operators and
//    world-line segments live in global arrays; chain c owns contiguous
//    ranges [site_base, site_base + N), [bond_base, bond_base + Nb), and the
//    operators / segments that the per-sweep scans place there.
//  * Operators are kept in bond-major order (bond ascending, imaginary time
//    ascending) -- the operator string of path_integral.C:57-195 stored as
//    CSR by bond.  An operator is (time f64, info i32) with
//    info = bond << 3 | graph << 1 | offdiagonal  (cf. operator.h:52-117).
//  * Imaginary time tau lives on the grid k * 2^-52, k in [0, 2^52), and is
//    stored as an exact FP64 value.
//  * A single-ion term adds in-site bonds between the virtual sites of each
//    real site after the lattice's bonds (ChainDev::nb_lat); a transverse
//    field then adds one bond per site (ChainDev::nb_real): its operators
//    are site operators, one event on the site, the world line cut there
//    (looper's site graph, graph_impl.h:66-85).
//  * Union-find nodes are world-line segments: for site s with m events
//    (operator legs touching s, ordered by (tau, bond, rank)), node soff[s]+j
//    is the segment just below event j; node soff[s] is the wrapping segment
//    that also holds the tau = 0 boundary.  A site without events owns one
//    node.  A virtual site (S=1) does not wrap: node soff[s] starts at tau = 0
//    and node soff[s]+m is its top segment up to tau = 1.  Roots are the
//    minimum node index of each cluster.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define LG_HD __host__ __device__ __forceinline__
#else
#define LG_HD inline
#endif

namespace lg {

constexpr uint64_t TGRID = 1ull << 52;
constexpr double TGRID_INV = 1.0 / 4503599627370496.0;
constexpr double EMPTY_TIME = 2.0;  // ev_time of the single node of an event-free site

enum : uint32_t { STAGE_DIAG = 1, STAGE_REGRAPH = 2, STAGE_FLIP = 3, STAGE_EXCHANGE = 4, STAGE_BOUNDARY = 5, STAGE_SSE = 6 };

// ev_flags bits
// EV_FIRST: lowest segment of its site (wraps through tau = 1, unless
// EV_NOWRAP); EV_GAUGE: the site's gauge factor is -1; EV_NOWRAP: the site's
// segments do not wrap (virtual sites: the lowest segment starts at tau = 0
// and the site owns a separate top segment ending at tau = 1, whose time
// entry is 1.0)
enum : uint8_t { EV_OFFDIAG = 1, EV_SRC = 2, EV_DIR1 = 4, EV_EMPTY = 8, EV_FIRST = 16, EV_GAUGE = 32, EV_NOWRAP = 64 };

// Per-chain descriptor (device resident, rebuilt by the host when the set of
// chains in a batch changes).
struct ChainDev {
  int kind;            // 0 chain, 1 square
  int L, N, Nb, dim;   // N, Nb: (virtual) sites and bonds of the lattice the kernels sweep
  int site_base, bond_base;          // layout of the configuration being built
  int old_site_base, old_bond_base;  // layout of the previous configuration
  int cand_cap;                      // candidate slots per bond
  long long cand_base;               // the chain's candidate slots: slot j of bond b at cand_base + b * cand_bstride + j * cand_sstride
  double beta, scale;                // scale = 2^52 / (beta * sum_g v_g)
  double offset_total;               // sum_b offset_b (energy.h / weight_impl.h)
  uint64_t thr[3];                   // graph thresholds: g = first g with x < thr[g], else 3
  uint64_t toff;                     // off-diagonal regraph threshold
  uint32_t key0, key1;               // Philox key = (seed, chain id)
  int has_weight;
  double h;                          // longitudinal field (site term -h S^z)
  double wsum;                       // sum_g v_g of a bond (candidate rate beta * wsum)
  int sched_n, sched_off;            // annealing schedule: entries and offset in the batch's table (0: fixed beta)
  int nv;                            // virtual sites per real site (2S; 1 for S=1/2)
  int cand_bstride, cand_sstride;    // candidate slot strides per bond and per slot
  // transverse field -hx S^x: bonds [nb_real, Nb) are the sites' own,
  // bond nb_real + s holding site s's operators (Nb = nb_real + N); without
  // one nb_real = Nb
  int nb_real;
  double site_w, site_scale;         // site graph weight |hx|/2; candidate scale 2^52 / (beta site_w)
  // single-ion term D (S^z)^2 (S > 1/2): bonds [nb_lat, nb_real) are
  // in-site bonds, npair = nv (nv - 1) / 2 per real site, an Ising coupling
  // 2D between two virtual sites of one real site with graphs of their own;
  // without it nb_lat = nb_real
  int nb_lat, npair;
  double dwsum, dscale;              // in-site bonds: sum_g v_g and candidate scale 2^52 / (beta dwsum)
  uint64_t dthr[3], dtoff;           // in-site bonds: graph and regraph thresholds
  // "loop; sse" (sse.C): the operator string in series order at
  // [sse_base, sse_base + sse_cap) of the batch's string arrays, one int
  // per operator (chain-local bond << 3 | graph << 1 | off-diagonal)
  int sse, sse_base, sse_cap;
  uint64_t pdiag[2];                 // choose_diagonal thresholds (antiparallel, parallel pair)
  uint64_t dpdiag[2];                // the same for the in-site bonds
  // SSE insertions: the location's class from word 2 of the step, lattice
  // bonds below sse_cls[0], in-site bonds below sse_cls[1], site bonds
  // above; the graph by word 2 against sse_tl / sse_td (thresholds scaled
  // into the class's range), the location within the class by word 3
  uint64_t sse_cls[2], sse_tl[3], sse_td[3];
  double sse_wtot;                   // sum of the weights of every (location, graph)
  // sign problem (Jxy > 0 with a transverse field, model_parameter.h:617-644):
  // off-diagonal operators of negative weight on the lattice bonds
  // (antiferromagnetic XY) and on the site bonds (hx < 0); the
  // configuration's sign is (-1)^(their number)
  int sgn_lattice, sgn_site;
};
// SSE chains: their sites' spins and bonds' counters are staged in LDS
constexpr int SSE_MAX_SITES = 4096, SSE_MAX_BONDS = 4096;
// virtual sites per real site (2S) at most: the boundary permutation of a
// real site's virtual world lines is held as eight 4-bit fields
constexpr int MAX_NV = 8;

// Temperature of Monte Carlo step `step` on an annealing schedule of n
// (start step, temperature) pairs stored as doubles, start steps ascending,
// the last pair (first step at the final temperature, final temperature):
// the piecewise-linear interpolation of looper/temperature.h:93-113 (the
// first entry's temperature up to its step, the final one after the last).
// Step 0 takes the first entry's temperature (the reference rejects step 0
// when a schedule is set, temperature.h:96-100, and its worker asks for it
// on its first sweep).  The arithmetic is written out so that host and
// device evaluate it identically.
LG_HD double schedule_temperature(const double* sched, int n, int step) {
#ifdef __HIPCC__
#pragma clang fp contract(off)
#endif
  if (step < 1) step = 1;
  int k = 0;
  while (k < n && sched[2 * k] < (double)step) ++k;
  if (k == 0) return sched[1];
  if (k == n) return sched[2 * n - 1];
  const double s0 = sched[2 * k - 2], t0 = sched[2 * k - 1], s1 = sched[2 * k], t1 = sched[2 * k + 1];
  return t0 + ((double)step - s0) * (t1 - t0) / (s1 - s0);
}

// candidate-time scale 2^52 / (beta * sum_g v_g) of a chain at inverse temperature beta
LG_HD double candidate_scale(double beta, double wsum) {
  return wsum > 0 ? 4503599627370496.0 / (beta * wsum) : 0.0;
}

// Philox4x32-10 (Salmon et al., SC'11): 10 rounds, multipliers D2511F53 /
// CD9E8D57, Weyl increments 9E3779B9 / BB67AE85.
struct U4 { uint32_t x, y, z, w; };

LG_HD U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return U4{c0, c1, c2, c3};
}

// Natural logarithm from IEEE +,-,*,/ only (no contraction), identical on
// host and device: log x = e ln2 + 2 atanh(s), s = (m-1)/(m+1),
// m in (sqrt(1/2), sqrt(2)], atanh by its Taylor series.
LG_HD double det_log(double x) {
#ifdef __HIPCC__
#pragma clang fp contract(off)
#endif
  uint64_t bits = *reinterpret_cast<uint64_t*>(&x);
  int e = (int)((bits >> 52) & 0x7ff) - 1023;
  uint64_t mb = (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
  double m = *reinterpret_cast<double*>(&mb);
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

// exp(x) from IEEE +,-,*,/ and exact scalings only, identical on host and
// device: x = k ln2 + r with the Cody-Waite split of ln2 (high part 32
// significant bits, so k*ln2_hi is exact), e^r by its Taylor series to r^14
// (|r| <= ln2/2), times 2^k.
LG_HD double det_exp(double x) {
#ifdef __HIPCC__
#pragma clang fp contract(off)
#endif
  if (!(x > -745.0)) return 0.0;
  if (x > 709.0) x = 709.0;
  const double k = floor(x * 1.4426950408889634 + 0.5);
  const double r = (x - k * 0.6931471803691238) - k * 1.9082149292705877e-10;
  double p = 1.0 / 87178291200.0;  // 1/14!
  p = p * r + 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)k);
}

// tanh x = sign(x) (1 - e^-2|x|) / (1 + e^-2|x|), from det_exp
LG_HD double det_tanh(double x) {
#ifdef __HIPCC__
#pragma clang fp contract(off)
#endif
  const double a = x < 0 ? -x : x;
  const double e = det_exp(-2.0 * a);
  const double t = (1.0 - e) / (1.0 + e);
  return x < 0 ? -t : t;
}

// uniform on (0,1] from the first 53 bits of a Philox output
LG_HD double uniform53(const U4& r) {
  uint64_t u53 = ((uint64_t)r.x << 21) | (uint64_t)(r.y >> 11);
  return (double)(u53 + 1) * (1.0 / 9007199254740992.0);
}

// log u for u uniform on (0,1] from the first 53 bits of a Philox output
LG_HD double log_uniform(const U4& r) {
#ifdef __HIPCC__
#pragma clang fp contract(off)
#endif
  uint64_t u53 = ((uint64_t)r.x << 21) | (uint64_t)(r.y >> 11);
  double u = (double)(u53 + 1) * (1.0 / 9007199254740992.0);
  return det_log(u);
}

// Exponential gap on the 2^-52 grid for rate lambda (scale = 2^52/lambda),
// from the first 53 bits of a Philox output; saturates at 2^52.
LG_HD uint64_t exp_gap(const U4& r, double scale) {
#ifdef __HIPCC__
#pragma clang fp contract(off)
#endif
  double g = -log_uniform(r) * scale;
  if (!(g < 4503599627370496.0)) return TGRID;
  return (uint64_t)g;
}

// graph g has probability v_g / sum v; thr[g] = floor(2^32 (v_0+..+v_g)/sum v),
// and 2^32 from the last graph of nonzero weight on
LG_HD int choose_graph(uint32_t x, const ChainDev& c) {
  int g = 0;
  while (g < 3 && (uint64_t)x >= c.thr[g]) ++g;
  return g;
}

// the same for the in-site bonds of the single-ion term
LG_HD int choose_graph_d(uint32_t x, const ChainDev& c) {
  int g = 0;
  while (g < 3 && (uint64_t)x >= c.dthr[g]) ++g;
  return g;
}

// ---- lattice geometry (ALPS conventions) ----------------------------------
// chain: bond b = (b, b+1 mod L); site s has bonds {s (source), s-1 (target)}
// square: site s = x*L + y; bond 2s = (s, s+y^) [dir 1], 2s+1 = (s, s+x^) [dir 0];
//         site s has bonds {2s, 2s+1 (source), 2(s-y^), 2(s-x^)+1 (target)}
// Spin S > 1/2 (path_integral.C:589-615, looper's virtual lattice,
// lattice.h:104-175): real site r carries nv = 2S virtual sites v = r*nv + a
// (a < nv); real bond rb = (rs, rt) carries the nv^2 virtual bonds
// rb*nv^2 + a*nv + c = (rs*nv + a, rt*nv + c), source-major as looper's
// virtual_graph orders them.  Virtual site v = r*nv + a has nv*z bond lists
// (z = 2 on the chain, 4 on the square lattice): for the j-th incident real
// bond rb of r (the real order above: source bonds first) and c < nv,
// list j*nv + c is bond rb*nv^2 + a*nv + c (r the source) or
// rb*nv^2 + c*nv + a (r the target).  Every virtual site carries its real
// site's gauge.

// sites of a real bond (real lattice)
LG_HD void real_bond_sites(const ChainDev& c, int b, int& s0, int& s1) {
  if (c.kind == 0) {
    s0 = b;
    s1 = (b + 1 == c.L) ? 0 : b + 1;
  } else {
    int s = b >> 1, x = s / c.L, y = s - x * c.L;
    s0 = s;
    if ((b & 1) == 0) s1 = x * c.L + ((y + 1 == c.L) ? 0 : y + 1);
    else s1 = ((x + 1 == c.L) ? 0 : x + 1) * c.L + y;
  }
}

LG_HD void bond_sites(const ChainDev& c, int b, int& s0, int& s1) {
  if (c.nv == 1) {
    real_bond_sites(c, b, s0, s1);
  } else {
    const int nn = c.nv * c.nv;
    const int rb = b / nn, q = b - rb * nn, a = q / c.nv, d = q - a * c.nv;
    int r0, r1;
    real_bond_sites(c, rb, r0, r1);
    s0 = r0 * c.nv + a;
    s1 = r1 * c.nv + d;
  }
}

// In-site bonds of the single-ion term (looper's in-real-site bonds,
// lattice.h:632-646): real site r's pairs a < c of virtual sites, a outer,
// pair p(a, c) = a nv - a (a + 1) / 2 + (c - a - 1), bond nb_lat + r npair + p.
LG_HD int pair_index(int nv, int a, int c) { return a * nv - a * (a + 1) / 2 + (c - a - 1); }

LG_HD void dpair_sites(const ChainDev& c, int b, int& s0, int& s1) {
  const int q = b - c.nb_lat, r = q / c.npair;
  int rem = q - r * c.npair, a = 0;
  while (rem >= c.nv - 1 - a) { rem -= c.nv - 1 - a; ++a; }
  s0 = r * c.nv + a;
  s1 = r * c.nv + a + 1 + rem;
}

// in-site bond j (j < nv - 1) of virtual site s: with its real site's j-th
// other virtual site
LG_HD int dpair_bond(const ChainDev& c, int s, int j) {
  const int r = s / c.nv, a = s - r * c.nv;
  const int o = (j < a) ? j : j + 1;
  return c.nb_lat + r * c.npair + (a < o ? pair_index(c.nv, a, o) : pair_index(c.nv, o, a));
}

// sites of the operators of bond b: a site bond (transverse field) has its
// one site on both sides
LG_HD void op_sites(const ChainDev& c, int b, int& s0, int& s1) {
  if (b >= c.nb_real) {
    s0 = s1 = b - c.nb_real;
    return;
  }
  if (b >= c.nb_lat) {
    dpair_sites(c, b, s0, s1);
    return;
  }
  bond_sites(c, b, s0, s1);
}

LG_HD int bond_dir(const ChainDev& c, int b) {
  return (c.kind == 0) ? 0 : (((b / (c.nv * c.nv)) & 1) ? 0 : 1);
}

// bond lists of a site: nv * (2 on the chain, 4 on the square lattice)
LG_HD int site_lists(const ChainDev& c) { return c.nv * (c.kind == 0 ? 2 : 4); }

// real bond j of real site r (source bonds first: chain {r, r-1}; square
// {2r, 2r+1, 2(r - y^), 2(r - x^) + 1})
LG_HD int real_site_bond(const ChainDev& c, int r, int j) {
  if (c.kind == 0) return j == 0 ? r : ((r == 0) ? c.L - 1 : r - 1);
  const int x = r / c.L, y = r - x * c.L;
  if (j < 2) return 2 * r + j;
  if (j == 2) return 2 * (x * c.L + ((y == 0) ? c.L - 1 : y - 1));
  return 2 * (((x == 0) ? c.L - 1 : x - 1) * c.L + y) + 1;
}

// bond list i of site s (i < site_lists(c))
LG_HD int site_bond(const ChainDev& c, int s, int i) {
  if (c.nv == 1) return real_site_bond(c, s, i);
  const int r = s / c.nv, a = s - r * c.nv;
  const int j = i / c.nv, q = i - j * c.nv;
  const int rb = real_site_bond(c, r, j);
  const bool source = j < (c.kind == 0 ? 1 : 2);
  return rb * c.nv * c.nv + (source ? a * c.nv + q : q * c.nv + a);
}

// bonds incident to site s when it has at most four lists (S=1/2, and S=1
// on the chain); returns their count.  The four entries are chosen into
// scalars and stored once: stores of different entries on different paths
// get merged into an indexed store, which puts the caller's array in
// scratch memory.
LG_HD int site_bonds(const ChainDev& c, int s, int* bl) {
  int b0, b1, b2, b3, n;
  if (c.nv == 1 && c.kind == 0) {
    b0 = s;
    b1 = (s == 0) ? c.L - 1 : s - 1;
    b2 = b3 = -1;  // unused
    n = 2;
  } else if (c.nv == 1) {
    const int x = s / c.L, y = s - x * c.L;
    const int ym = (y == 0) ? c.L - 1 : y - 1;
    const int xm = (x == 0) ? c.L - 1 : x - 1;
    b0 = 2 * s;
    b1 = 2 * s + 1;
    b2 = 2 * (x * c.L + ym);
    b3 = 2 * (xm * c.L + y) + 1;
    n = 4;
  } else {
    n = site_lists(c);
    b0 = site_bond(c, s, 0);
    b1 = site_bond(c, s, 1);
    b2 = n > 2 ? site_bond(c, s, 2) : -1;
    b3 = n > 3 ? site_bond(c, s, 3) : -1;
  }
  bl[0] = b0;
  bl[1] = b1;
  bl[2] = b2;
  bl[3] = b3;
  return n;
}

LG_HD int gauge(const ChainDev& c, int s) {
  const int r = (c.nv == 1) ? s : s / c.nv;
  if (c.kind == 0) return (r & 1) ? -1 : 1;
  int x = r / c.L, y = r - x * c.L;
  return ((x + y) & 1) ? -1 : 1;
}

// Sweep record layout (identical to lg_sweep_record in include/loopgpu.h).
enum {
  R_NOP = 0, R_NC, R_ENERGY,
  R_UMAG0, R_USIZE2, R_UMAG2, R_USIZE4, R_UMAG4, R_USIZE, R_UMAG,
  R_SMAG0, R_SSIZE2, R_SMAG2, R_SSIZE4, R_SMAG4, R_SSIZE, R_SMAG,
  R_W2, R_BETA, R_SWEEP,
  R_UMAGA, R_SMAGA, R_EFIELD,  // sum_c umag_c, sum_c smag_c (normal estimator), Zeeman energy
  R_SIGN,                      // sign of the configuration's weight (+-1; 1 for sign-free models)
  R_COUNT
};

// Per-cluster accumulator, one 64-byte record per root node.
// Segments are split into four classes k = 2*spin + (gauge < 0): 0 up/+,
// 1 up/-, 2 down/+, 3 down/-, so each segment adds to one slot only.
struct ClusterAcc {
  double L[4];      // world-line length per class
  int n[4];         // tau = 0 boundary sites per class
  long long w;      // winding numbers packed as w_x + 2^32 w_y
  long long pad;
};

LG_HD int seg_class(int spin, int g) { return (spin << 1) | (g < 0 ? 1 : 0); }

LG_HD long long pack_winding(int dw0, int dw1) { return (long long)dw0 + ((long long)dw1 << 32); }

LG_HD void unpack_winding(long long w, int& w0, int& w1) {
  w0 = (int)(uint32_t)(uint64_t)w;      // |w_x| < 2^31: low word sign-extended
  w1 = (int)((w - (long long)w0) >> 32);
}

}  // namespace lg
