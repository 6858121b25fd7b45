// gcm_cache_manager.hpp -- header-only C++ facade over the C ABI in
// include/gcm.h with the shape of the reference's Cache::Manager
// (src/CacheManager/include/CacheManager.h:56-410): Build, Fill returning a
// deferred std::future<bool>, PropagateParameters, CopyHistogramContents,
// CopyEventWeights, ValidateHistogramContents, GetResidentMemory, the timers,
// ParameterIndex and the toggles (per object here: there is no singleton).  GUNDAM's
// Propagator can hold one of these instead of the singleton; the tables it
// passes are the EventDialCache content laid out as gcm_tables.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <future>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gcm.h"

namespace gcm {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc, const char* where) {
  if (rc != GCM_OK)
    throw Error(rc, std::string(where) + ": " + gcm_last_error());
}

// Running timer over the last N intervals, the shape of the reference's
// GenericToolbox::Time::AveragedTimer<10> behind GetCacheFillTimer and
// GetPullFromDeviceTimer (CacheManager.h:128-136).
template <int N>
class AveragedTimer {
 public:
  void add(double ms) {
    last_ = ms;
    ring_[count_ % N] = ms;
    ++count_;
  }
  double LastMs() const { return last_; }
  double AverageMs() const {
    const long n = std::min<long>(count_, N);
    double s = 0;
    for (long i = 0; i < n; ++i) s += ring_[i];
    return n ? s / n : 0.0;
  }
  long Count() const { return count_; }

 private:
  double ring_[N] = {};
  double last_ = 0;
  long count_ = 0;
};

class CacheManager {
 public:
  explicit CacheManager(int device = 0, int precision = GCM_PRECISION_FP64) {
    check(gcm_create(device, &ctx_), "gcm_create");
    check(gcm_set_precision(ctx_, precision), "gcm_set_precision");
  }
  ~CacheManager() { gcm_destroy(ctx_); }
  CacheManager(const CacheManager&) = delete;
  CacheManager& operator=(const CacheManager&) = delete;

  // Cache::Manager::HasGPU (CacheManager.h:75-78)
  static bool HasGPU() {
    int n = 0;
    return gcm_device_count(&n) == GCM_OK && n > 0;
  }
  // Cache::Manager::HasCUDA (CacheManager.h:70-73): this build is HIP on
  // gfx950, with no CUDA.
  static bool HasCUDA() { return false; }

  // SetIsHistContentCopyEnabled (CacheManager.h:109-115): whether the future of
  // Fill copies the histograms back; returns the previous setting.
  bool SetIsHistContentCopyEnabled(bool enable) {
    const bool old = hist_copy_;
    hist_copy_ = enable;
    return old;
  }
  // SetEnableDebugPrintouts (CacheManager.h:125-126): one line per completed
  // Fill (enqueue and wait times, total -2 ln L of point 0) on stdout.
  void SetEnableDebugPrintouts(bool enable) { debug_ = enable; }

  // GetCacheFillTimer / GetPullFromDeviceTimer (CacheManager.h:128-136): the
  // time Fill spends uploading and enqueueing, and the time its future's get()
  // waits for the device and copies the results back.
  const AveragedTimer<10>& GetCacheFillTimer() const { return fill_timer_; }
  const AveragedTimer<10>& GetPullFromDeviceTimer() const { return pull_timer_; }

  // ParameterIndex (CacheManager.h:159-163): parameters are addressed by
  // their position in the theta vector the caller passes to Fill, so the
  // index of parameter p is p itself (negative if it is not one of them).
  int ParameterIndex(int p) const { return (p >= 0 && p < n_par_) ? p : -1; }

  void SetLaunchConfig(const gcm_launch_config& cfg) {
    check(gcm_set_launch_config(ctx_, &cfg), "gcm_set_launch_config");
  }

  // Cache::Manager::Build + Update (CacheManager.cpp:198-740)
  bool Build(const gcm_tables& tables, int max_points = 1) {
    check(gcm_build(ctx_, &tables, max_points), "gcm_build");
    max_points_ = max_points;
    n_bins_ = tables.n_bins;
    n_events_ = tables.n_events;
    n_par_ = tables.n_parameters;
    table_len_ = tables.n_tables ? tables.table_begin[tables.n_tables] : 0;
    built_ = true;
    return true;
  }
  bool IsBuilt() const { return built_; }

  // SetIsEventWeightCopyEnabled (CacheManager.h:117-123)
  bool SetIsEventWeightCopyEnabled(bool enable) {
    const bool old = weight_copy_;
    check(gcm_enable_weights(ctx_, enable ? 1 : 0), "gcm_enable_weights");
    weight_copy_ = enable;
    return old;
  }

  void SetLikelihood(const gcm_llh_config& cfg, const std::vector<double>& data,
                     const std::vector<double>& prior_mean,
                     const std::vector<double>& prior_sigma) {
    check(gcm_set_likelihood(ctx_, &cfg, data.data(),
                             prior_mean.empty() ? nullptr : prior_mean.data(),
                             prior_sigma.empty() ? nullptr : prior_sigma.data()),
          "gcm_set_likelihood");
  }

  // The prior penalty of each of the last k_points points.
  std::vector<double> Penalty(int k_points = 1) {
    std::vector<double> out(k_points);
    check(gcm_get_llh(ctx_, nullptr, nullptr, out.data()), "gcm_get_llh");
    return out;
  }

  // Correlated prior penalties per parameter set (LikelihoodInterface.cpp:
  // 186-213), added after the diagonal term in set order.
  void SetPenaltySets(const std::vector<gcm_penalty_set>& sets) {
    check(gcm_set_penalty_sets(ctx_, (int32_t)sets.size(), sets.empty() ? nullptr : sets.data()),
          "gcm_set_penalty_sets");
  }

  // Device-side table refresh: the update kernel `kernel` of the code object
  // at `code_object` (include/gcm_table_update.h) fills the jobs' tables at
  // every Fill, in place of table_values (TabulatedDialFactory.cpp:133-141).
  void AddTableUpdate(const std::string& code_object, const std::string& kernel,
                      const std::vector<gcm_table_job>& jobs) {
    check(gcm_add_table_update(ctx_, code_object.c_str(), kernel.c_str(), (int32_t)jobs.size(),
                               jobs.data()),
          "gcm_add_table_update");
  }

  // Cache::Manager::Fill (CacheManager.cpp:750-810): upload the parameter
  // point(s) and the lookup tables, enqueue the evaluation, and return a
  // deferred future whose get() synchronises and copies the histograms when
  // the copy is enabled.
  std::future<bool> Fill(const std::vector<double>& theta, int k_points = 1,
                         const std::vector<double>* table_values = nullptr) {
    const auto t0 = Clock::now();
    check(gcm_set_parameters(ctx_, theta.data(), n_par_, k_points), "gcm_set_parameters");
    if (table_values && table_len_)
      check(gcm_set_table_values(ctx_, table_values->data(), table_len_, k_points),
            "gcm_set_table_values");
    check(gcm_fill(ctx_, nullptr), "gcm_fill");
    fill_timer_.add(ms_since(t0));
    return std::async(std::launch::deferred, [this] {
      const auto t1 = Clock::now();
      check(gcm_sync(ctx_), "gcm_sync");
      if (hist_copy_) CopyHistogramContents();
      if (weight_copy_) CopyEventWeights();
      pull_timer_.add(ms_since(t1));
      if (debug_) {
        double llh = 0;
        check(gcm_get_llh(ctx_, &llh, nullptr, nullptr), "gcm_get_llh");
        std::printf("[gcm] fill %ld: enqueue %.3f ms, wait+copy %.3f ms, -2lnL %.10g\n",
                    pull_timer_.Count(), fill_timer_.LastMs(), pull_timer_.LastMs(), llh);
      }
      return true;
    });
  }

  bool PropagateParameters(const std::vector<double>& theta,
                           const std::vector<double>* table_values = nullptr) {
    return Fill(theta, 1, table_values).get();
  }

  // CopyHistogramContents -> CacheSampleHistFiller::copyHistogram
  // (CacheSampleHistFiller.cpp:9-19): sumWeights and sqrtSumSqWeights.
  bool CopyHistogramContents(int k = 0) {
    sum_w_.resize(n_bins_);
    std::vector<double> s2(n_bins_);
    check(gcm_get_hist(ctx_, k, sum_w_.data(), s2.data()), "gcm_get_hist");
    sqrt_sum_w2_.resize(n_bins_);
    for (int b = 0; b < n_bins_; ++b) sqrt_sum_w2_[b] = std::sqrt(s2[b]);
    return true;
  }
  bool CopyEventWeights() {
    weights_.assign((size_t)n_events_, std::nan(""));
    check(gcm_get_weights(ctx_, weights_.data()), "gcm_get_weights");
    return true;
  }

  // Cache::Manager::ValidateHistogramContents against CPU-filled histograms
  // (sumWeights, sqrtSumSqWeights per global bin).
  bool ValidateHistogramContents(const std::vector<double>& cpu_sums,
                                 const std::vector<double>& cpu_sqrt_sum_w2, bool verbose = false,
                                 int k = 0) {
    int32_t bad = 0;
    check(gcm_validate_hist(ctx_, k, cpu_sums.data(), cpu_sqrt_sum_w2.data(), 0.3,
                            verbose ? 1 : 0, &bad),
          "gcm_validate_hist");
    return bad == 0;
  }

  std::vector<double> Likelihood(int k_points = 1) {
    std::vector<double> tot(k_points);
    check(gcm_get_llh(ctx_, tot.data(), nullptr, nullptr), "gcm_get_llh");
    return tot;
  }

  // Batched central-difference gradient of the total -2 ln L for Minuit:
  // the points theta and theta +- h e_j for every j in `params` go out in
  // launches of up to max_points points (the reference's RootMinimizer asks
  // Minuit for a numerical gradient, one evalFit per point,
  // RootMinimizer.cpp:125-126).  table_fn(points, k, values) fills the
  // lookup tables of k points (k x n_parameters, row-major) when the tables
  // depend on the parameters.  Returns -2 ln L at theta; grad[i] is the
  // derivative along params[i].
  using TableFn = std::function<void(const std::vector<double>&, int, std::vector<double>&)>;
  double Gradient(const std::vector<double>& theta, double h, const std::vector<int>& params,
                  std::vector<double>* grad, const TableFn& table_fn = {}) {
    const int m = (int)params.size();
    const int n = 2 * m + 1;
    std::vector<double> pts((size_t)n * n_par_), llh(n), tab;
    for (int i = 0; i < n; ++i) std::copy(theta.begin(), theta.end(), pts.begin() + (size_t)i * n_par_);
    for (int j = 0; j < m; ++j) {
      pts[(size_t)(1 + 2 * j) * n_par_ + params[j]] += h;
      pts[(size_t)(2 + 2 * j) * n_par_ + params[j]] -= h;
    }
    for (int i0 = 0; i0 < n; i0 += max_points_) {
      const int k = std::min(max_points_, n - i0);
      std::vector<double> chunk(pts.begin() + (size_t)i0 * n_par_,
                                pts.begin() + (size_t)(i0 + k) * n_par_);
      const std::vector<double>* tv = nullptr;
      if (table_fn && table_len_) {
        table_fn(chunk, k, tab);
        tv = &tab;
      }
      const bool hc = hist_copy_, wc = weight_copy_;
      hist_copy_ = weight_copy_ = false;
      Fill(chunk, k, tv).get();
      hist_copy_ = hc;
      weight_copy_ = wc;
      check(gcm_get_llh(ctx_, llh.data() + i0, nullptr, nullptr), "gcm_get_llh");
    }
    if (grad) {
      grad->resize(m);
      for (int j = 0; j < m; ++j) (*grad)[j] = (llh[1 + 2 * j] - llh[2 + 2 * j]) / (2 * h);
    }
    return llh[0];
  }

  std::size_t GetResidentMemory() const {
    int64_t b = 0;
    gcm_resident_bytes(ctx_, &b);
    return (std::size_t)b;
  }

  const std::vector<double>& SumWeights() const { return sum_w_; }
  const std::vector<double>& SqrtSumSqWeights() const { return sqrt_sum_w2_; }
  const std::vector<double>& EventWeights() const { return weights_; }
  gcm_context* context() { return ctx_; }

 private:
  gcm_context* ctx_ = nullptr;
  bool built_ = false;
  using Clock = std::chrono::steady_clock;
  static double ms_since(Clock::time_point t) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
  }
  bool hist_copy_ = true;  // fIsHistContentCopyEnabled default true
  bool weight_copy_ = false;
  bool debug_ = false;
  AveragedTimer<10> fill_timer_, pull_timer_;
  int max_points_ = 1;
  int n_bins_ = 0, n_par_ = 0;
  int64_t n_events_ = 0, table_len_ = 0;
  std::vector<double> sum_w_, sqrt_sum_w2_, weights_;
};

}  // namespace gcm
