// looper_amd::loop_worker -- host C++ mirror of the reference worker
// "loop; path integral" (path_integral.C:57-351) on top of the C-ABI of
// include/loopgpu.h.
//
//  * parameters by the reference's names (alps::Parameters semantics for the
//    subset this path supports): LATTICE, L, W, T / TEMPERATURE / BETA /
//    INVERSE_TEMPERATURE (temperature.h:43-79), J / Jxy / Jz, MODEL, local_S,
//    annealing T_START_i / T_DURATION_i and their TEMPERATURE_, BETA_,
//    INVERSE_TEMPERATURE_ forms (temperature.h:56-81),
//    SWEEPS (default "[65536:]") and THERMALIZATION (default SWEEPS.min/8,
//    montecarlo.h:41-45), SEED, FORCE_SCATTER (weight_impl.h:169), h, Gamma
//    (transverse field, model_parameter.h:67), D (single-ion term),
//    DISABLE_IMPROVED_ESTIMATOR, MEASURE[Correlations] / MEASURE[Structure
//    Factor] / INITIAL_SITE (correlation.h:74-108);
//  * init_observables / run / is_thermalized / progress / save / load with the
//    worker's observable names (path_integral.C:307-321,831-862, energy.h,
//    susceptibility.h, stiffness.h) and its evaluator (Specific Heat,
//    Binder ratios: energy.h:89-102, susceptibility.h:340-374);
//  * binning analysis of every observable (mean, error, integrated
//    autocorrelation time), measurements before thermalization discarded.
//
// Errors are reported like the reference's (std::invalid_argument for bad
// parameters) inside C++, and as status codes across the C-ABI.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/loopgpu.h"
#include "../../include/loopgpu_worker.h"
#include "correlation.h"
#include "lg_binning.h"

namespace looper_amd {

// ---------------------------------------------------------------------------
// parameters: "KEY = value" separated by ';' or newlines, values are numbers,
// quoted strings or arithmetic expressions over other parameters (loop.ip)
// ---------------------------------------------------------------------------
class Parameters {
 public:
  explicit Parameters(const std::string& text) {
    std::string cur;
    bool quoted = false;
    for (char ch : text) {
      if (ch == '"') quoted = !quoted;
      if (!quoted && (ch == ';' || ch == '\n')) { add(cur); cur.clear(); }
      else cur += ch;
    }
    add(cur);
  }
  bool defined(const std::string& k) const { return kv_.count(k) > 0; }
  std::vector<std::string> keys() const {
    std::vector<std::string> v;
    for (const auto& kv : kv_) v.push_back(kv.first);
    return v;
  }
  std::string str(const std::string& k, const std::string& def = "") const {
    auto it = kv_.find(k);
    if (it == kv_.end()) return def;
    std::string v = it->second;
    if (v.size() >= 2 && v.front() == '"' && v.back() == '"') v = v.substr(1, v.size() - 2);
    return v;
  }
  double num(const std::string& k) const {
    if (!defined(k)) throw std::invalid_argument("parameter " + k + " not defined");
    return eval(str(k), 0);
  }
  double num(const std::string& k, double def) const { return defined(k) ? num(k) : def; }

 private:
  static std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
    return (a == std::string::npos) ? "" : s.substr(a, b - a + 1);
  }
  void add(const std::string& stmt) {
    std::string s = trim(stmt);
    if (s.empty() || s[0] == '#' || s == "{" || s == "}") return;
    size_t eq = s.find('=');
    if (eq == std::string::npos) throw std::invalid_argument("malformed parameter: " + s);
    kv_[trim(s.substr(0, eq))] = trim(s.substr(eq + 1));
  }
  // recursive-descent evaluation: expr = term {(+|-) term}, term = factor
  // {(*|/) factor}, factor = number | name | (expr) | -factor
  double eval(const std::string& e, int depth) const {
    if (depth > 32) throw std::invalid_argument("parameter expression nests too deeply");
    size_t pos = 0;
    double v = expr(e, pos, depth);
    skip(e, pos);
    if (pos != e.size()) throw std::invalid_argument("cannot evaluate parameter expression '" + e + "'");
    return v;
  }
  static void skip(const std::string& e, size_t& p) { while (p < e.size() && std::isspace((unsigned char)e[p])) ++p; }
  double expr(const std::string& e, size_t& p, int d) const {
    double v = term(e, p, d);
    for (;;) {
      skip(e, p);
      if (p < e.size() && (e[p] == '+' || e[p] == '-')) { char o = e[p++]; double w = term(e, p, d); v = (o == '+') ? v + w : v - w; }
      else return v;
    }
  }
  double term(const std::string& e, size_t& p, int d) const {
    double v = factor(e, p, d);
    for (;;) {
      skip(e, p);
      if (p < e.size() && (e[p] == '*' || e[p] == '/')) { char o = e[p++]; double w = factor(e, p, d); v = (o == '*') ? v * w : v / w; }
      else return v;
    }
  }
  double factor(const std::string& e, size_t& p, int d) const {
    skip(e, p);
    if (p >= e.size()) throw std::invalid_argument("empty parameter expression '" + e + "'");
    if (e[p] == '-') { ++p; return -factor(e, p, d); }
    if (e[p] == '+') { ++p; return factor(e, p, d); }
    if (e[p] == '(') { ++p; double v = expr(e, p, d); skip(e, p); if (p < e.size() && e[p] == ')') ++p; return v; }
    if (std::isdigit((unsigned char)e[p]) || e[p] == '.') {
      size_t q = p;
      while (q < e.size() && (std::isdigit((unsigned char)e[q]) || e[q] == '.' || e[q] == 'e' || e[q] == 'E' ||
                              ((e[q] == '-' || e[q] == '+') && q > p && (e[q - 1] == 'e' || e[q - 1] == 'E'))))
        ++q;
      double v = std::stod(e.substr(p, q - p));
      p = q;
      return v;
    }
    size_t q = p;
    while (q < e.size() && (std::isalnum((unsigned char)e[q]) || e[q] == '_')) ++q;
    std::string name = e.substr(p, q - p);
    p = q;
    if (name.empty() || !defined(name)) throw std::invalid_argument("unknown name in expression '" + e + "'");
    return eval(str(name), d + 1);
  }
  std::map<std::string, std::string> kv_;
};

struct Result { double mean, error, tau; bool evaluated; };

// jackknife over the common bins of two observables for f(<a>, <b>)
template <class F>
Result jackknife(const Observable& a, const Observable& b, F f) {
  size_t nb = std::min(a.bins().size(), b.bins().size());
  Result r{f(a.mean(), b.mean()), 0.0, 0.0, true};
  if (nb < 2 || a.bin_size() != b.bin_size()) return r;
  double sa = 0, sb = 0;
  for (size_t i = 0; i < nb; ++i) { sa += a.bins()[i]; sb += b.bins()[i]; }
  std::vector<double> jk(nb);
  double m = 0;
  for (size_t i = 0; i < nb; ++i) {
    jk[i] = f((sa - a.bins()[i]) / (nb - 1), (sb - b.bins()[i]) / (nb - 1));
    m += jk[i];
  }
  m /= nb;
  double v = 0;
  for (double x : jk) v += (x - m) * (x - m);
  r.error = std::sqrt(v * (nb - 1) / nb);
  return r;
}

// ---------------------------------------------------------------------------
// the worker
// ---------------------------------------------------------------------------
static int parse_sweeps(const std::string& s, uint64_t& mn, uint64_t& mx) {
  // "[a:b]", "[a:]", or an integer
  std::string t = s;
  t.erase(std::remove(t.begin(), t.end(), ' '), t.end());
  try {
    if (!t.empty() && t.front() == '[') {
      size_t c = t.find(':');
      if (c == std::string::npos || t.back() != ']') return -1;
      std::string a = t.substr(1, c - 1), b = t.substr(c + 1, t.size() - c - 2);
      mn = a.empty() ? 0 : std::stoull(a);
      // "[a:]": no upper bound -- runs until stopped, like parapack's open range
      mx = b.empty() ? UINT64_MAX / 4 : std::stoull(b);
      if (mx < mn) return -1;
      return 0;
    }
    mn = mx = (uint64_t)std::llround(std::stod(t));
  } catch (const std::exception&) {
    return -1;
  }
  return 0;
}

// looper::temperature::init (temperature.h:43-81): the final temperature
// from T, TEMPERATURE, BETA or INVERSE_TEMPERATURE, and the annealing
// schedule T_START_i / T_DURATION_i (or the TEMPERATURE_, BETA_ and
// INVERSE_TEMPERATURE_ forms), i = 0, 1, ... until a pair is missing or not
// positive, each starting where the previous one ended.  steps/temps receive
// the (start step, temperature) pairs, the last one (first step at the final
// temperature, final temperature).
double parse_temperature(const Parameters& p, std::vector<int32_t>& steps, std::vector<double>& temps) {
  double T;
  if (p.defined("T")) T = p.num("T");
  else if (p.defined("TEMPERATURE")) T = p.num("TEMPERATURE");
  else if (p.defined("BETA")) T = 1 / p.num("BETA");
  else if (p.defined("INVERSE_TEMPERATURE")) T = 1 / p.num("INVERSE_TEMPERATURE");
  else throw std::invalid_argument("temperature is not initialized");
  if (!(T > 0)) throw std::invalid_argument("negative temperature");
  steps.clear();
  temps.clear();
  int s_prev = 1;
  for (int i = 0;; ++i) {
    const std::string n = std::to_string(i);
    int td = -1;
    double ts = -1;
    if (p.defined("T_DURATION_" + n) && p.defined("T_START_" + n)) {
      td = (int)p.num("T_DURATION_" + n);
      ts = p.num("T_START_" + n);
    } else if (p.defined("TEMPERATURE_DURATION_" + n) && p.defined("TEMPERATURE_START_" + n)) {
      td = (int)p.num("TEMPERATURE_DURATION_" + n);
      ts = p.num("TEMPERATURE_START_" + n);
    } else if (p.defined("BETA_DURATION_" + n) && p.defined("BETA_START_" + n)) {
      td = (int)p.num("BETA_DURATION_" + n);
      ts = 1 / p.num("BETA_START_" + n);
    } else if (p.defined("INVERSE_TEMPERATURE_DURATION_" + n) && p.defined("INVERSE_TEMPERATURE_START_" + n)) {
      td = (int)p.num("INVERSE_TEMPERATURE_DURATION_" + n);
      ts = 1 / p.num("INVERSE_TEMPERATURE_START_" + n);
    }
    if (td <= 0 || !(ts > 0)) break;
    steps.push_back(s_prev);
    temps.push_back(ts);
    s_prev += td;
  }
  steps.push_back(s_prev);
  temps.push_back(T);
  return T;
}

class loop_worker {
 public:
  explicit loop_worker(const Parameters& p) {
    std::string lat = p.str("LATTICE", "chain lattice");
    if (lat == "chain lattice" || lat == "chain") lp_.lattice = LG_LATTICE_CHAIN;
    else if (lat == "square lattice" || lat == "square") lp_.lattice = LG_LATTICE_SQUARE;
    else throw std::invalid_argument("unsupported LATTICE \"" + lat + "\" (chain lattice, square lattice)");
    lp_.L = (int)std::llround(p.num("L"));
    if (lp_.lattice == LG_LATTICE_SQUARE && p.defined("W") && (int)std::llround(p.num("W")) != lp_.L)
      throw std::invalid_argument("square lattice with W != L is not supported");
    if (p.str("MODEL", "spin") != "spin") throw std::invalid_argument("only MODEL = spin is supported");
    // the worker: "loop; path integral" (path_integral.C) or "loop; sse" (sse.C)
    const std::string alg = p.str("ALGORITHM", "loop; path integral");
    if (alg == "loop; sse") lp_.algorithm = LG_ALGORITHM_SSE;
    else if (alg != "loop; path integral" && alg != "loop")
      throw std::invalid_argument("unsupported ALGORITHM \"" + alg + "\" (loop; path integral, loop; sse)");
    // spin size (model_parameter.h): S > 1/2 through 2S virtual sites per
    // real site (path_integral.C:589-615), up to S = LG_MAX_TWO_S / 2
    lp_.two_s = 1;
    if (p.defined("local_S")) {
      const double S = p.num("local_S");
      const long long t = std::llround(2.0 * S);
      if (std::fabs(2.0 * S - (double)t) > 1e-12 || t < 1 || t > LG_MAX_TWO_S)
        throw std::invalid_argument("local_S must be a multiple of 1/2 from 1/2 to " +
                                    std::to_string(LG_MAX_TWO_S / 2));
      lp_.two_s = (int32_t)t;
    }
    // single-ion term D (S^z)^2 (model_parameter.h:69): in-site bonds of the
    // virtual sites
    lp_.d = p.num("D", 0.0);
    // transverse field -Gamma S^x (model_parameter.h:67, the ALPS spin
    // model's Gamma): site operators; sign-free with Jxy <= 0, signed (the
    // sign measured, normal estimators) with Jxy > 0
    lp_.hx = p.num("Gamma", 0.0);
    lp_.h = p.num("h", 0.0);
    // improved estimators unless a field or DISABLE_IMPROVED_ESTIMATOR (path_integral.C:217)
    improved_ = lp_.h == 0.0 && !p.defined("DISABLE_IMPROVED_ESTIMATOR");
    // temperature and annealing schedule (temperature.h:43-81)
    const double T = parse_temperature(p, sched_steps_, sched_temps_);
    lp_.beta = 1 / T;
    double J = p.num("J", 1.0);
    lp_.jxy = p.num("Jxy", J);
    lp_.jz = p.num("Jz", J);
    // an antiferromagnetic XY coupling with a transverse field: the SIGN
    // path (model_parameter.h:617-644), observables weighted by the sign and
    // measured with the normal estimators.  (The reference keeps its improved
    // estimators there and multiplies them by an "improved" sign that is the
    // parity of the sum of every cluster's negative-operator count,
    // path_integral.C:782-802; each negative operator adds to two cluster
    // ends, so the sum is even, the improved sign is the plain sign, and the
    // improved magnetization estimators, which average over cluster flips
    // that change the sign, are biased.)
    signed_ = lp_.hx != 0 && lp_.jxy > 0;
    if (signed_) improved_ = false;
    // looper scatters by default where it finds the model frustrated, which a
    // positive single-ion term of S > 1/2 makes it (model_parameter.h:646-661,
    // model_impl.h:51)
    const bool frustrated = lp_.d > 0 && lp_.two_s > 1;
    lp_.force_scatter = p.num("FORCE_SCATTER", frustrated ? 0.1 : 0.0);
    lp_.seed = (uint32_t)(int64_t)std::llround(p.num("SEED", 0.0));
    lp_.chain_id = (uint32_t)std::llround(p.num("CHAIN_ID", 0.0));
    if (parse_sweeps(p.str("SWEEPS", "[65536:]"), sweeps_min_, sweeps_max_))
      throw std::invalid_argument("malformed SWEEPS");
    therm_ = p.defined("THERMALIZATION") ? (uint64_t)std::llround(p.num("THERMALIZATION")) : (sweeps_min_ >> 3);
    // path_integral.C:212-213 (annealing_steps = first step at the final temperature - 1)
    if ((uint64_t)(sched_steps_.back() - 1) > therm_)
      throw std::invalid_argument("longer annealing steps than thermalization");
    n_sites_ = (lp_.lattice == LG_LATTICE_CHAIN) ? lp_.L : lp_.L * lp_.L;
    dim_ = (lp_.lattice == LG_LATTICE_CHAIN) ? 1 : 2;
    // correlation.h:74-108; the Green function is disabled there as well
    auto flag = [&](const char* k) {
      if (!p.defined(k)) return false;
      const std::string v = p.str(k);
      return v == "true" || v == "True" || v == "1" || (v != "false" && v != "False" && p.num(k) != 0.0);
    };
    if (flag("MEASURE[Green Function]"))
      std::fprintf(stderr, "WARNING: Green funciton measurement is not implemented yet\n");
    const bool mc = flag("MEASURE[Correlations]"), ms = flag("MEASURE[Structure Factor]");
    // ALPS custom measurements (custom.h:63-250) in their forms for the
    // S=1/2 site operator Sz: MEASURE_AVERAGE / MEASURE_LOCAL [name] = Sz,
    // MEASURE_CORRELATIONS / MEASURE_STRUCTURE_FACTOR [name] = "Sz:Sz"
    std::vector<Correlation::Custom> custom;
    const char* kinds[] = {"MEASURE_AVERAGE[", "MEASURE_LOCAL[", "MEASURE_CORRELATIONS[", "MEASURE_STRUCTURE_FACTOR["};
    for (const auto& k : p.keys()) {
      for (int t = 0; t < 4; ++t) {
        const std::string pre = kinds[t];
        if (k.compare(0, pre.size(), pre) != 0 || k.back() != ']') continue;
        const std::string name = k.substr(pre.size(), k.size() - pre.size() - 1);
        std::string e = p.str(k);
        e.erase(std::remove(e.begin(), e.end(), ' '), e.end());
        if ((t < 2 && e == "Sz") || (t >= 2 && e == "Sz:Sz")) custom.push_back({t, name});
        else
          std::fprintf(stderr, "WARNING: custom measurement %s = %s is not supported (only Sz and Sz:Sz); skipped\n",
                       k.c_str(), e.c_str());
      }
    }
    if (mc || ms || !custom.empty()) {
      if (lp_.two_s > 1) throw std::invalid_argument("correlation measurements are built for S=1/2 only");
      if (lp_.hx != 0) throw std::invalid_argument("correlation measurements are built without a transverse field");
      if (lp_.algorithm != LG_ALGORITHM_PATH_INTEGRAL)
        throw std::invalid_argument("correlation measurements are built for the path-integral worker");
      if (lp_.L < 2 || lp_.L % 2) throw std::invalid_argument("L must be even");
      const int origin = p.defined("INITIAL_SITE") ? (int)std::llround(p.num("INITIAL_SITE")) : -1;
      if (p.defined("INITIAL_SITE") && (origin < 0 || origin >= n_sites_))
        throw std::invalid_argument("INITIAL_SITE out of range");
      corr_.reset(new Correlation(lp_.lattice, lp_.L, mc, ms, origin, custom));
      spins_.resize(n_sites_);
      spins8_.resize(n_sites_);
    }
    int rc = lg_chain_create(&lp_, &h_);
    if (rc == LG_OK && sched_steps_.size() > 1)
      rc = lg_chain_set_schedule(h_, (int32_t)sched_steps_.size(), sched_steps_.data(), sched_temps_.data());
    if (rc != LG_OK) {
      char buf[512];
      lg_last_error(buf, sizeof buf);
      if (h_) { lg_chain_destroy(h_); h_ = nullptr; }
      if (rc == LG_ERR_ARG) throw std::invalid_argument(buf);
      throw std::runtime_error(buf);
    }
  }
  ~loop_worker() { if (h_) lg_chain_destroy(h_); }
  loop_worker(const loop_worker&) = delete;
  loop_worker& operator=(const loop_worker&) = delete;

  // path_integral.C:307-321 + estimator init_observables
  void init_observables(std::map<std::string, Observable>& obs) const {
    for (const auto& n : names(improved_)) obs[n];
    if (signed_)
      for (const char* n : {"Sign", "Weight of Zero-Meron Sector", "Sign in Zero-Meron Sector"}) obs[n];
    if (corr_)
      for (const auto& n : corr_->names()) obs[n];
  }
  bool is_thermalized() const { return mcs_ >= therm_; }
  double progress() const { return (double)mcs_ / (double)(therm_ + sweeps_min_); }
  // montecarlo.h:55-57 (mcs < therm || mcs - therm < SWEEPS.max), stopped
  // once progress() reaches 1 as the scheduler finishes a clone there
  bool can_work() const {
    return (mcs_ < therm_ || mcs_ - therm_ < sweeps_max_) && mcs_ < therm_ + sweeps_min_;
  }
  uint64_t mcs() const { return mcs_; }

  // one Monte Carlo step (path_integral.C:323-351); sweeps are fetched from the
  // device in chunks, so run() usually only consumes a buffered record.  A
  // chunk never goes beyond the `budget` steps the caller still takes, the
  // end of thermalization or the end of the run, so the device's
  // configuration is this step's whenever the caller stops (save() needs it)
  void run(std::map<std::string, Observable>& obs, uint64_t budget = 1) {
    if (next_ == buf_.size()) {
      uint64_t left = (mcs_ < therm_) ? therm_ - mcs_ : (therm_ + sweeps_min_ - mcs_);
      left = std::min<uint64_t>(left, budget);
      // correlations are measured on every sweep's spins: one sweep at a time
      if (corr_ && is_thermalized()) left = 1;
      uint32_t k = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(left, chunk_));
      buf_.resize(k);
      int rc = lg_chain_sweep(h_, k, buf_.data());
      if (rc != LG_OK) {
        char b[512];
        lg_last_error(b, sizeof b);
        throw std::runtime_error(b);
      }
      next_ = 0;
    }
    const lg_sweep_record& r = buf_[next_++];
    if (is_thermalized()) {
      measure(r, obs);
      if (corr_) {
        // tau = 0 spins after this sweep's flip (the chunk held this sweep only)
        int rc = lg_chain_get_config(h_, spins_.data(), nullptr, nullptr, nullptr);
        if (rc != LG_OK) {
          char b[512];
          lg_last_error(b, sizeof b);
          throw std::runtime_error(b);
        }
        for (int i = 0; i < n_sites_; ++i) spins8_[i] = (uint8_t)spins_[i];
        corr_->measure(spins8_.data(), obs);
      }
    }
    ++mcs_;
  }

  // checkpoint: the Monte Carlo counters (mc_steps::save, montecarlo.h:66:
  // mcs, SWEEPS range, thermalization), then the chain's configuration in the
  // reference's content order (path_integral.C:111-124: spins, operators;
  // lg_chain_save), then the measurements so far (the observable set that
  // parapack checkpoints beside the worker), so that a resumed worker ends
  // with the results of an uninterrupted one.
  std::vector<uint8_t> save(const std::map<std::string, Observable>& obs) const {
    if (next_ != buf_.size()) throw std::logic_error("save() with sweeps fetched but not yet run");
    size_t n = 0;
    int rc = lg_chain_state_size(h_, &n);
    std::vector<uint8_t> chain(n);
    if (rc == LG_OK) rc = lg_chain_save(h_, chain.data(), n);
    if (rc != LG_OK) {
      char b[512];
      lg_last_error(b, sizeof b);
      throw std::runtime_error(b);
    }
    std::vector<uint8_t> out;
    auto put = [&](const void* p, size_t k) { const uint8_t* b = (const uint8_t*)p; out.insert(out.end(), b, b + k); };
    const char magic[8] = {'L', 'G', 'W', 'O', 'R', 'K', 'R', '2'};
    put(magic, 8);
    put(&mcs_, 8); put(&sweeps_min_, 8); put(&sweeps_max_, 8); put(&therm_, 8);
    const uint64_t cl = chain.size();
    put(&cl, 8);
    put(chain.data(), chain.size());
    const uint64_t nobs = obs.size();
    put(&nobs, 8);
    for (const auto& kv : obs) {
      const uint64_t nl = kv.first.size();
      put(&nl, 8);
      put(kv.first.data(), nl);
      std::vector<double> v;
      kv.second.save(v);
      const uint64_t nv = v.size();
      put(&nv, 8);
      put(v.data(), 8 * nv);
    }
    return out;
  }
  void load(const uint8_t* p, size_t len, std::map<std::string, Observable>& obs) {
    size_t off = 0;
    auto get = [&](void* d, size_t n) {
      if (n > len - off) throw std::invalid_argument("truncated checkpoint");
      std::memcpy(d, p + off, n);
      off += n;
    };
    char magic[8];
    get(magic, 8);
    if (std::memcmp(magic, "LGWORKR2", 8) != 0) throw std::invalid_argument("not a loopgpu worker checkpoint");
    uint64_t mcs, smin, smax, therm, cl;
    get(&mcs, 8); get(&smin, 8); get(&smax, 8); get(&therm, 8); get(&cl, 8);
    if (cl > len - off) throw std::invalid_argument("truncated checkpoint");
    int rc = lg_chain_load(h_, p + off, cl);
    if (rc != LG_OK) {
      char b[512];
      lg_last_error(b, sizeof b);
      throw std::invalid_argument(b);
    }
    off += cl;
    uint64_t nobs;
    get(&nobs, 8);
    std::map<std::string, Observable> o;
    for (uint64_t i = 0; i < nobs; ++i) {
      uint64_t nl, nv;
      get(&nl, 8);
      if (nl > len - off) throw std::invalid_argument("truncated checkpoint");
      std::string name((const char*)p + off, nl);
      off += nl;
      get(&nv, 8);
      if (nv > (len - off) / 8 || nv < 7) throw std::invalid_argument("truncated checkpoint");
      std::vector<double> v(nv);
      get(v.data(), 8 * nv);
      o[name].load(v.data());
    }
    mcs_ = mcs; sweeps_min_ = smin; sweeps_max_ = smax; therm_ = therm;
    obs.swap(o);
    buf_.clear();
    next_ = 0;
  }
  void set_chunk(uint32_t c) { chunk_ = std::max<uint32_t>(1, c); }

  bool improved() const { return improved_; }

  static std::vector<std::string> names(bool improved = true) {
    static const char* base[] = {
        "Temperature", "Inverse Temperature", "Volume", "Number of Sites", "Number of Clusters",
        "Energy", "Energy Density", "Energy^2", "Stiffness"};
    static const char* mag[] = {
        "Magnetization", "Magnetization Density", "|Magnetization|", "|Magnetization Density|",
        "Magnetization^2", "Magnetization Density^2", "Magnetization^4", "Magnetization Density^4",
        "Susceptibility", "Generalized Magnetization^2", "Generalized Magnetization Density^2",
        "Generalized Magnetization^4", "Generalized Magnetization Density^4", "Generalized Susceptibility"};
    std::vector<std::string> v(base, base + (improved ? 9 : 8));  // Stiffness: improved only
    for (const char* m : mag)
      if (improved || std::string(m).rfind("Generalized ", 0) != 0) v.push_back(m);
    for (const char* m : mag) {
      std::string s(m);
      if (!improved && s.rfind("Generalized ", 0) == 0) continue;
      // insert "Staggered " after a leading "|" or "Generalized "
      if (s[0] == '|') s = "|Staggered " + s.substr(1);
      else if (s.rfind("Generalized ", 0) == 0) s = "Generalized Staggered " + s.substr(12);
      else s = "Staggered " + s;
      v.push_back(s);
    }
    return v;
  }

  int n_sites() const { return n_sites_; }

 private:
  // loop_worker::dispatch measurement (path_integral.C:831-862) and the
  // collector commits (energy.h:70-80, susceptibility.h:200-256,
  // stiffness.h:123-128), sign = 1
  void measure(const lg_sweep_record& r, std::map<std::string, Observable>& m) const {
    const double beta = r.beta, vol = n_sites_, nop = r.n_op;
    auto p2 = [](double x) { return x * x; };
    auto p4 = [](double x) { return x * x * x * x; };
    m["Temperature"] << 1 / beta;
    m["Inverse Temperature"] << beta;
    m["Volume"] << vol;
    m["Number of Sites"] << (double)n_sites_;
    m["Number of Clusters"] << r.n_clusters;
    // sign problem (path_integral.C:836-861): the sign's observables, and
    // every observable below weighted by the configuration's sign (the
    // evaluator divides by the mean sign)
    const double sg = signed_ ? r.sign : 1.0;
    if (signed_) {
      m["Sign"] << r.sign;
      m["Weight of Zero-Meron Sector"] << 1.0;  // (the cluster signs always sum to an even number)
      m["Sign in Zero-Meron Sector"] << r.sign;
    }
    auto put = [&](const std::string& k, double v) { m[k] << sg * v; };
    put("Energy", r.energy);
    put("Energy Density", r.energy / vol);
    put("Energy^2", p2(r.energy) - nop / p2(beta));
    // SSE: the tau-integrated sums come in units of the string;
    // beta (n S + S0) / (n + 1) (susceptibility.h:213-251, 326-333)
    const bool sse = lp_.algorithm == LG_ALGORITHM_SSE;
    auto chi = [&](double s, double s0) { return sse ? beta * (nop * s + s0) / (nop + 1) / vol : beta * s / vol; };
    if (!improved_) {
      // normal estimator (susceptibility.h:304-337), the configuration before the flip
      const double mags[2][2] = {{r.umag0, r.umag_a}, {r.smag0, r.smag_a}};
      for (int k = 0; k < 2; ++k) {
        const std::string P = k ? "Staggered " : "", bar = "|" + P;
        const double mg = mags[k][0], ma = mags[k][1];
        put(P + "Magnetization", mg);
        put(P + "Magnetization Density", mg / vol);
        put(bar + "Magnetization|", std::fabs(mg));
        put(bar + "Magnetization Density|", std::fabs(mg) / vol);
        put(P + "Magnetization^2", p2(mg));
        put(P + "Magnetization Density^2", p2(mg / vol));
        put(P + "Magnetization^4", p4(mg));
        put(P + "Magnetization Density^4", p4(mg / vol));
        put(P + "Susceptibility", chi(p2(ma), p2(mg)));
      }
      return;
    }
    put("Stiffness", r.w2 / (beta * dim_));
    struct Set { const char* pre; double mag0, size2, mag2, size4, mag4, sus, gsus; };
    const Set sets[2] = {
        {"", r.umag0, r.usize2, r.umag2, r.usize4, r.umag4, r.umag, r.usize},
        {"Staggered ", r.smag0, r.ssize2, r.smag2, r.ssize4, r.smag4, r.smag, r.ssize}};
    for (const Set& s : sets) {
      std::string P(s.pre);
      std::string bar = "|" + P;
      std::string gen = "Generalized " + P;
      put(P + "Magnetization", 0.0);
      put(P + "Magnetization Density", 0.0);
      put(bar + "Magnetization|", std::fabs(s.mag0));
      put(bar + "Magnetization Density|", std::fabs(s.mag0) / vol);
      put(P + "Magnetization^2", s.mag2);
      put(P + "Magnetization Density^2", s.mag2 / p2(vol));
      put(P + "Magnetization^4", 3 * p2(s.mag2) - 2 * s.mag4);
      put(P + "Magnetization Density^4", (3 * p2(s.mag2) - 2 * s.mag4) / p4(vol));
      put(P + "Susceptibility", chi(s.sus, s.mag2));
      put(gen + "Magnetization^2", s.size2);
      put(gen + "Magnetization Density^2", s.size2 / p2(vol));
      put(gen + "Magnetization^4", 3 * p2(s.size2) - 2 * s.size4);
      put(gen + "Magnetization Density^4", (3 * p2(s.size2) - 2 * s.size4) / p4(vol));
      put(gen + "Susceptibility", chi(s.gsus, s.size2));
    }
  }

  lg_params lp_{};
  std::vector<int32_t> sched_steps_;   // annealing schedule (start step, temperature); last = final
  std::vector<double> sched_temps_;
  bool improved_ = true;
  bool signed_ = false;          // sign problem: Jxy > 0 with a transverse field
  lg_chain* h_ = nullptr;
  uint64_t mcs_ = 0, therm_ = 0, sweeps_min_ = 0, sweeps_max_ = 0;
  int n_sites_ = 0, dim_ = 1;
  std::vector<lg_sweep_record> buf_;
  size_t next_ = 0;
  uint32_t chunk_ = 1024;
  std::unique_ptr<Correlation> corr_;
  std::vector<int32_t> spins_;
  std::vector<uint8_t> spins8_;
};

// evaluator (energy.h:89-102, susceptibility.h:340-374)
// jackknife over the common bins of three observables for f(<a>, <b>, <c>)
template <class F>
Result jackknife3(const Observable& a, const Observable& b, const Observable& c, F f) {
  size_t nb = std::min(std::min(a.bins().size(), b.bins().size()), c.bins().size());
  Result r{f(a.mean(), b.mean(), c.mean()), 0.0, 0.0, true};
  if (nb < 2 || a.bin_size() != b.bin_size() || a.bin_size() != c.bin_size()) return r;
  double sa = 0, sb = 0, sc = 0;
  for (size_t i = 0; i < nb; ++i) { sa += a.bins()[i]; sb += b.bins()[i]; sc += c.bins()[i]; }
  std::vector<double> jk(nb);
  double m = 0;
  for (size_t i = 0; i < nb; ++i) {
    jk[i] = f((sa - a.bins()[i]) / (nb - 1), (sb - b.bins()[i]) / (nb - 1), (sc - c.bins()[i]) / (nb - 1));
    m += jk[i];
  }
  m /= nb;
  double v = 0;
  for (double x : jk) v += (x - m) * (x - m);
  r.error = std::sqrt(v * (nb - 1) / nb);
  return r;
}

std::map<std::string, Result> evaluate(const std::map<std::string, Observable>& obs) {
  std::map<std::string, Result> out;
  for (const auto& kv : obs)
    if (kv.second.count()) out[kv.first] = Result{kv.second.mean(), kv.second.error(), kv.second.tau(), false};
  auto has = [&](const char* k) { return obs.count(k) && obs.at(k).count() > 0; };
  if (has("Sign")) {
    // sign problem: every sign-weighted observable as <s O> / <s> (ALPS's
    // signed observables); the unweighted ones and the sign's own stay
    const Observable& sg = obs.at("Sign");
    static const char* plain[] = {"Temperature", "Inverse Temperature", "Volume", "Number of Sites",
                                  "Number of Clusters", "Sign", "Weight of Zero-Meron Sector",
                                  "Sign in Zero-Meron Sector"};
    for (const auto& kv : obs) {
      if (!kv.second.count() || std::find(std::begin(plain), std::end(plain), kv.first) != std::end(plain)) continue;
      Result r = jackknife(kv.second, sg, [](double x, double s) { return x / s; });
      r.tau = kv.second.tau();
      r.evaluated = false;
      out[kv.first] = r;
    }
    if (has("Energy") && has("Energy^2") && has("Inverse Temperature") && has("Volume")) {
      double b = obs.at("Inverse Temperature").mean(), v = obs.at("Volume").mean();
      out["Specific Heat"] = jackknife3(obs.at("Energy^2"), obs.at("Energy"), sg, [&](double e2, double e, double s) {
        return b * b * (e2 / s - (e / s) * (e / s)) / v;
      });
    }
  } else if (has("Energy") && has("Energy^2") && has("Inverse Temperature") && has("Volume")) {
    double b = obs.at("Inverse Temperature").mean(), v = obs.at("Volume").mean();
    out["Specific Heat"] = jackknife(obs.at("Energy^2"), obs.at("Energy"),
                                     [&](double e2, double e) { return b * b * (e2 - e * e) / v; });
  }
  const char* binder[][3] = {
      {"Magnetization^2", "Magnetization^4", "Binder Ratio of Magnetization"},
      {"Staggered Magnetization^2", "Staggered Magnetization^4", "Binder Ratio of Staggered Magnetization"},
      {"Generalized Magnetization^2", "Generalized Magnetization^4", "Binder Ratio of Generalized Magnetization"},
      {"Generalized Staggered Magnetization^2", "Generalized Staggered Magnetization^4",
       "Binder Ratio of Generalized Staggered Magnetization"}};
  for (auto& b : binder)
    if (has(b[0]) && has(b[1])) {
      if (has("Sign"))  // (m2 / s)^2 / (m4 / s)
        out[b[2]] = jackknife3(obs.at(b[0]), obs.at(b[1]), obs.at("Sign"),
                               [](double m2, double m4, double s) { return m2 * m2 / (m4 * s); });
      else
        out[b[2]] = jackknife(obs.at(b[0]), obs.at(b[1]), [](double m2, double m4) { return m2 * m2 / m4; });
    }
  return out;
}

std::string report(const std::map<std::string, Result>& res) {
  std::ostringstream os;
  os.precision(6);
  os << "[results]\n";
  for (const auto& kv : res) {
    os << kv.first << ": " << kv.second.mean << " +/- " << kv.second.error;
    if (!kv.second.evaluated) os << "; tau = " << std::round(kv.second.tau * 100) / 100;
    os << "\n";
  }
  return os.str();
}

}  // namespace looper_amd

// ===========================================================================
// C-ABI of the worker (include/loopgpu_worker.h)
// ===========================================================================
struct lg_worker {
  std::unique_ptr<looper_amd::loop_worker> w;
  std::map<std::string, looper_amd::Observable> obs;
};

namespace {
thread_local std::string g_werr;
int werr(int code, const std::string& m) { g_werr = m; return code; }
}  // namespace

extern "C" {

int lg_worker_last_error(char* buf, size_t len) {
  if (!buf || !len) return LG_ERR_ARG;
  std::snprintf(buf, len, "%s", g_werr.c_str());
  return LG_OK;
}

int lg_worker_create(const char* params, lg_worker** out) {
  if (!params || !out) return werr(LG_ERR_ARG, "null pointer");
  *out = nullptr;
  try {
    looper_amd::Parameters p(params);
    lg_worker* h = new lg_worker();
    try {
      h->w.reset(new looper_amd::loop_worker(p));
    } catch (...) {
      delete h;
      throw;
    }
    h->w->init_observables(h->obs);
    *out = h;
    return LG_OK;
  } catch (const std::invalid_argument& e) {
    return werr(LG_ERR_ARG, e.what());
  } catch (const std::exception& e) {
    return werr(LG_ERR_DEVICE, e.what());
  }
}

int lg_worker_run(lg_worker* h, uint64_t max_steps, uint64_t* done) {
  if (!h) return werr(LG_ERR_ARG, "null handle");
  uint64_t n = 0;
  try {
    while (h->w->can_work() && n < max_steps) { h->w->run(h->obs, max_steps - n); ++n; }
  } catch (const std::exception& e) {
    if (done) *done = n;
    return werr(LG_ERR_DEVICE, e.what());
  }
  if (done) *done = n;
  return LG_OK;
}

int lg_worker_status(lg_worker* h, uint64_t* mcs, int* thermalized, double* progress, int* can_work) {
  if (!h) return werr(LG_ERR_ARG, "null handle");
  if (mcs) *mcs = h->w->mcs();
  if (thermalized) *thermalized = h->w->is_thermalized();
  if (progress) *progress = h->w->progress();
  if (can_work) *can_work = h->w->can_work();
  return LG_OK;
}

int lg_worker_result(lg_worker* h, const char* name, double* mean, double* error, double* tau) {
  if (!h || !name) return werr(LG_ERR_ARG, "null pointer");
  auto res = looper_amd::evaluate(h->obs);
  auto it = res.find(name);
  if (it == res.end()) return werr(LG_ERR_ARG, std::string("no observable named ") + name);
  if (mean) *mean = it->second.mean;
  if (error) *error = it->second.error;
  if (tau) *tau = it->second.tau;
  return LG_OK;
}

int lg_worker_report(lg_worker* h, char* buf, size_t len, size_t* needed) {
  if (!h) return werr(LG_ERR_ARG, "null handle");
  std::string s = looper_amd::report(looper_amd::evaluate(h->obs));
  if (needed) *needed = s.size() + 1;
  if (buf && len) std::snprintf(buf, len, "%s", s.c_str());
  return LG_OK;
}

int lg_worker_save(lg_worker* h, void* buf, size_t* len) {
  if (!h || !len) return werr(LG_ERR_ARG, "null pointer");
  try {
    std::vector<uint8_t> b = h->w->save(h->obs);
    if (buf && *len >= b.size()) std::memcpy(buf, b.data(), b.size());
    bool fits = buf && *len >= b.size();
    *len = b.size();
    return (buf && !fits) ? werr(LG_ERR_ARG, "buffer too small") : LG_OK;
  } catch (const std::exception& e) {
    return werr(LG_ERR_STATE, e.what());
  }
}

int lg_worker_load(lg_worker* h, const void* buf, size_t len) {
  if (!h || !buf) return werr(LG_ERR_ARG, "null pointer");
  try {
    h->w->load((const uint8_t*)buf, len, h->obs);
    return LG_OK;
  } catch (const std::invalid_argument& e) {
    return werr(LG_ERR_ARG, e.what());
  } catch (const std::exception& e) {
    return werr(LG_ERR_DEVICE, e.what());
  }
}

int lg_worker_set_chunk(lg_worker* h, uint32_t sweeps) {
  if (!h) return werr(LG_ERR_ARG, "null handle");
  h->w->set_chunk(sweeps);
  return LG_OK;
}

int lg_worker_destroy(lg_worker* h) {
  delete h;
  return LG_OK;
}

int lg_temperature_schedule(const char* params, int32_t max_n, int32_t* n, int32_t* steps, double* temperatures) {
  if (!params || !n) return werr(LG_ERR_ARG, "null pointer");
  try {
    looper_amd::Parameters p(params);
    std::vector<int32_t> st;
    std::vector<double> tp;
    looper_amd::parse_temperature(p, st, tp);
    *n = (int32_t)st.size();
    if ((int32_t)st.size() > max_n) return werr(LG_ERR_ARG, "schedule longer than max_n");
    for (size_t k = 0; k < st.size(); ++k) {
      if (steps) steps[k] = st[k];
      if (temperatures) temperatures[k] = tp[k];
    }
    return LG_OK;
  } catch (const std::exception& e) {
    return werr(LG_ERR_ARG, e.what());
  }
}

}  // extern "C"
