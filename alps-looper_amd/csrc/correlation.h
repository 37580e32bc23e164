// Correlation estimators of the worker (looper::correlation,
// correlation.h:59-296 of the reference): spin correlations, staggered spin
// correlations and the spin structure factor of the tau = 0 spins.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/loopgpu.h"
#include "lg_binning.h"

namespace looper_amd {

class Correlation {
 public:
  // custom measurement of Sz: kind 0 average, 1 local, 2 correlations (Sz:Sz),
  // 3 structure factor (Sz:Sz); name as given in MEASURE_...[name]
  struct Custom { int kind; std::string name; };
  // origin < 0: every pair, by displacement; else the INITIAL_SITE's pairs
  Correlation(int lattice, int L, bool correlations, bool structure_factor, int origin,
              std::vector<Custom> custom = {});
  bool enabled() const { return corr_ || sfac_ || !custom_.empty(); }
  // per-site / per-displacement / per-momentum values (N each)
  void compute(const uint8_t* spins, double* ucorr, double* scorr, double* sfac) const;
  void measure(const uint8_t* spins, std::map<std::string, Observable>& m) const;
  std::vector<std::string> names() const;
  static double momentum(int n, int L);

 private:
  bool corr_, sfac_;
  bool need_corr_, need_sfac_;  // also for custom measurements
  int origin_;
  std::vector<Custom> custom_;
  std::vector<std::string> site_;
  int nx_ = 1, ny_ = 1, n_ = 1;
  std::vector<std::string> dist_, mom_;
  std::vector<double> cos_, sin_;
};

}  // namespace looper_amd
