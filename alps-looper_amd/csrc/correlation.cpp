// Correlation estimators of the worker (the reference's looper::correlation,
// correlation.h:59-296): on the tau = 0 spins at the end of each sweep
// (after the cluster flip, path_integral.C:822-862),
//
//   Spin Correlations[d]           = 1/4 sum_i s_i s_(i+d) / m_d
//   Staggered Spin Correlations[d] = 1/4 sum_i g_i g_(i+d) s_i s_(i+d) / m_d
//   Spin Structure Factor[k]       = |sum_s (1/2 - spin_s) e^(i k.r_s)|^2 / V
//
// with s = 1 - 2 spin, g the gauge (+1 on the sublattice of site 0), d a
// displacement of the periodic lattice (one distance class per displacement,
// multiplicity m_d = N) and k the lattice momenta 2 pi n / L folded into
// (-pi, pi].  With INITIAL_SITE = o the correlations are 1/4 s_o s_r per site
// r instead (correlation.h:241-252).  Labels follow the reference's output
// (loop.op: "( 0 ) -- ( 1 )", "( 1.5708 )"); on the square lattice the two
// coordinates are written "( x,y )".
//
// The correlations are exact integer pair counts: rows of spins as bit
// strings, the count of disagreeing pairs for a displacement is the popcount
// of a row XOR the displaced row (N L^2 / 64 word operations per sweep on an
// L x L lattice, against the reference's N^2 pair loop).  The structure factor
// is a separable discrete Fourier transform (2 L^3 complex products).
#include "correlation.h"

#include <cmath>
#include <cstdio>
#include <sstream>

namespace looper_amd {

namespace {
std::string num_label(double v) {
  std::ostringstream os;
  os.precision(6);  // ostream default, as the reference's labels
  os << v;
  return os.str();
}
}  // namespace

Correlation::Correlation(int lattice, int L, bool correlations, bool structure_factor, int origin,
                         std::vector<Custom> custom)
    : corr_(correlations), sfac_(structure_factor), origin_(origin), custom_(std::move(custom)) {
  need_corr_ = corr_;
  need_sfac_ = sfac_;
  for (const auto& c : custom_) {
    if (c.kind == 2) need_corr_ = true;
    if (c.kind == 3) need_sfac_ = true;
  }
  nx_ = (lattice == LG_LATTICE_CHAIN) ? 1 : L;
  ny_ = L;
  n_ = nx_ * ny_;
  if (origin_ >= n_) throw std::invalid_argument("INITIAL_SITE out of range");
  auto coord = [&](int s) {
    if (nx_ == 1) return "( " + std::to_string(s) + " )";
    return "( " + std::to_string(s / ny_) + "," + std::to_string(s % ny_) + " )";
  };
  for (int r = 0; r < n_; ++r) site_.push_back(coord(r));
  if (need_corr_) {
    const std::string o = coord(origin_ < 0 ? 0 : origin_);
    for (int r = 0; r < n_; ++r) dist_.push_back(o + " -- " + coord(r));
  }
  if (need_sfac_) {
    for (int a = 0; a < nx_; ++a)
      for (int b = 0; b < ny_; ++b) {
        if (nx_ == 1) mom_.push_back("( " + num_label(momentum(b, ny_)) + " )");
        else mom_.push_back("( " + num_label(momentum(a, nx_)) + "," + num_label(momentum(b, ny_)) + " )");
      }
    for (int m = 0; m < L; ++m) {
      const double ph = 2.0 * M_PI * m / L;
      cos_.push_back(std::cos(ph));
      sin_.push_back(std::sin(ph));
    }
  }
}

double Correlation::momentum(int n, int L) {
  double k = 2.0 * M_PI * n / L;
  if (k > M_PI + 1e-9) k -= 2.0 * M_PI;
  return k;
}

void Correlation::compute(const uint8_t* spins, double* ucorr, double* scorr, double* sfac) const {
  const int nx = nx_, ny = ny_, n = n_;
  if (need_corr_ && origin_ >= 0) {
    // INITIAL_SITE: correlations of the origin with every site
    const int so = 1 - 2 * (spins[origin_] & 1);
    const int go = ((origin_ / ny + origin_ % ny) & 1) ? -1 : 1;
    for (int r = 0; r < n; ++r) {
      const int sr = 1 - 2 * (spins[r] & 1);
      const int gr = ((r / ny + r % ny) & 1) ? -1 : 1;
      ucorr[r] = 0.25 * so * sr;
      scorr[r] = 0.25 * go * gr * so * sr;
    }
  } else if (need_corr_) {
    const int W = (ny + 63) / 64;
    std::vector<uint64_t> row((size_t)nx * W, 0), rot((size_t)nx * W);
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y)
        if (spins[x * ny + y] & 1) row[(size_t)x * W + y / 64] |= 1ull << (y % 64);
    for (int dy = 0; dy < ny; ++dy) {
      // bit y of the displaced row = spin at y + dy
      std::fill(rot.begin(), rot.end(), 0ull);
      for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
          if (spins[x * ny + (y + dy) % ny] & 1) rot[(size_t)x * W + y / 64] |= 1ull << (y % 64);
      for (int dx = 0; dx < nx; ++dx) {
        long long dis = 0;  // pairs (i, i + d) with opposite spins
        for (int x = 0; x < nx; ++x) {
          const uint64_t* a = &row[(size_t)x * W];
          const uint64_t* b = &rot[(size_t)((x + dx) % nx) * W];
          for (int w = 0; w < W; ++w) dis += __builtin_popcountll(a[w] ^ b[w]);
        }
        const int d = dx * ny + dy;
        const double c = 0.25 * (double)(n - 2 * dis) / n;
        ucorr[d] = c;
        scorr[d] = ((dx + dy) & 1) ? -c : c;  // g_i g_(i+d) on the bipartite lattice
      }
    }
  }
  if (need_sfac_) {
    // F(a, y) = sum_x v(x, y) e^(i 2 pi a x / nx), then over y
    std::vector<double> fr((size_t)nx * ny), fi((size_t)nx * ny);
    for (int a = 0; a < nx; ++a)
      for (int y = 0; y < ny; ++y) {
        double re = 0, im = 0;
        for (int x = 0; x < nx; ++x) {
          const double v = 0.5 - (spins[x * ny + y] & 1);
          const int m = (nx == 1) ? 0 : (int)(((long long)a * x) % nx);
          re += v * (nx == 1 ? 1.0 : cos_[m]);
          im += v * (nx == 1 ? 0.0 : sin_[m]);
        }
        fr[(size_t)a * ny + y] = re;
        fi[(size_t)a * ny + y] = im;
      }
    for (int a = 0; a < nx; ++a)
      for (int b = 0; b < ny; ++b) {
        double re = 0, im = 0;
        for (int y = 0; y < ny; ++y) {
          const int m = (int)(((long long)b * y) % ny);
          const double c = cos_[m], s = sin_[m];
          const double xr = fr[(size_t)a * ny + y], xi = fi[(size_t)a * ny + y];
          re += xr * c - xi * s;
          im += xr * s + xi * c;
        }
        sfac[a * ny + b] = (re * re + im * im) / n;
      }
  }
}

void Correlation::measure(const uint8_t* spins, std::map<std::string, Observable>& m) const {
  std::vector<double> u(n_), s(n_), f(n_);
  compute(spins, u.data(), s.data(), f.data());
  if (corr_)
    for (int d = 0; d < n_; ++d) {
      m["Spin Correlations[" + dist_[d] + "]"] << u[d];
      m["Staggered Spin Correlations[" + dist_[d] + "]"] << s[d];
    }
  if (sfac_)
    for (int k = 0; k < n_; ++k) m["Spin Structure Factor[" + mom_[k] + "]"] << f[k];
  // custom measurements of Sz (custom.h:147-250): Sz = 1/2 - spin per site;
  // Sz:Sz correlations and structure factor are the spin correlations and
  // the spin structure factor above
  for (const auto& c : custom_) {
    if (c.kind == 0) {
      double v = 0;
      for (int r = 0; r < n_; ++r) v += 0.5 - (spins[r] & 1);
      m[c.name] << v / n_;
    } else if (c.kind == 1) {
      for (int r = 0; r < n_; ++r) m[c.name + "[" + site_[r] + "]"] << 0.5 - (spins[r] & 1);
    } else if (c.kind == 2) {
      for (int d = 0; d < n_; ++d) m[c.name + "[" + dist_[d] + "]"] << u[d];
    } else {
      for (int k = 0; k < n_; ++k) m[c.name + "[" + mom_[k] + "]"] << f[k];
    }
  }
}

std::vector<std::string> Correlation::names() const {
  std::vector<std::string> v;
  if (corr_)
    for (const auto& d : dist_) {
      v.push_back("Spin Correlations[" + d + "]");
      v.push_back("Staggered Spin Correlations[" + d + "]");
    }
  if (sfac_)
    for (const auto& k : mom_) v.push_back("Spin Structure Factor[" + k + "]");
  for (const auto& c : custom_) {
    if (c.kind == 0) v.push_back(c.name);
    const auto& labels = c.kind == 1 ? site_ : (c.kind == 2 ? dist_ : mom_);
    if (c.kind != 0)
      for (const auto& l : labels) v.push_back(c.name + "[" + l + "]");
  }
  return v;
}

}  // namespace looper_amd

extern "C" int lg_spin_correlations(int32_t lattice, int32_t L, const uint8_t* spins, int32_t origin,
                                    double* ucorr, double* scorr, double* sfac) {
  if (!spins || !ucorr || !scorr || !sfac) return LG_ERR_ARG;
  if ((lattice != LG_LATTICE_CHAIN && lattice != LG_LATTICE_SQUARE) || L < 2 || L % 2 || L > 4096)
    return LG_ERR_ARG;
  try {
    looper_amd::Correlation c(lattice, L, true, true, origin);
    c.compute(spins, ucorr, scorr, sfac);
  } catch (const std::exception&) {
    return LG_ERR_ARG;
  }
  return LG_OK;
}
