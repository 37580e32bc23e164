// Binning accumulator shared by the worker and the job server.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace looper_amd {

// ---------------------------------------------------------------------------
// binning: level-0 moments plus a bounded array of bin means (bin size
// doubles whenever 2*NBINS bins are full); error from the bins, tau from the
// ratio of binned and naive variances, functions of means by jackknife.
// ---------------------------------------------------------------------------
class Observable {
 public:
  static const size_t NBINS = 512;
  void operator<<(double x) {
    ++n_;
    s1_ += x;
    s2_ += x * x;
    cur_ += x;
    if (++cur_n_ == bin_size_) {
      bins_.push_back(cur_ / bin_size_);
      cur_ = 0;
      cur_n_ = 0;
      if (bins_.size() == 2 * NBINS) {
        for (size_t i = 0; i < NBINS; ++i) bins_[i] = 0.5 * (bins_[2 * i] + bins_[2 * i + 1]);
        bins_.resize(NBINS);
        bin_size_ *= 2;
      }
    }
  }
  uint64_t count() const { return n_; }
  double mean() const { return n_ ? s1_ / n_ : 0.0; }
  double naive_error() const {
    if (n_ < 2) return 0;
    double m = mean(), v = s2_ / n_ - m * m;
    return v > 0 ? std::sqrt(v / (n_ - 1)) : 0.0;
  }
  double error() const {
    size_t nb = bins_.size();
    if (nb < 2) return naive_error();
    double m = 0, v = 0;
    for (double b : bins_) m += b;
    m /= nb;
    for (double b : bins_) v += (b - m) * (b - m);
    return std::max(naive_error(), std::sqrt(v / (nb - 1) / nb));
  }
  double tau() const {
    double e0 = naive_error(), e = error();
    return (e0 > 0) ? 0.5 * (e * e / (e0 * e0) - 1) : 0.0;
  }
  const std::vector<double>& bins() const { return bins_; }
  uint64_t bin_size() const { return bin_size_; }
  void save(std::vector<double>& out) const {
    out.push_back((double)n_); out.push_back(s1_); out.push_back(s2_); out.push_back(cur_);
    out.push_back((double)cur_n_); out.push_back((double)bin_size_); out.push_back((double)bins_.size());
    out.insert(out.end(), bins_.begin(), bins_.end());
  }
  size_t load(const double* p) {
    n_ = (uint64_t)p[0]; s1_ = p[1]; s2_ = p[2]; cur_ = p[3]; cur_n_ = (uint64_t)p[4];
    bin_size_ = (uint64_t)p[5];
    size_t nb = (size_t)p[6];
    bins_.assign(p + 7, p + 7 + nb);
    return 7 + nb;
  }

 private:
  uint64_t n_ = 0, cur_n_ = 0, bin_size_ = 1;
  double s1_ = 0, s2_ = 0, cur_ = 0;
  std::vector<double> bins_;
};

}  // namespace looper_amd
