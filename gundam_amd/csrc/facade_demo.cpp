// facade_demo.cpp -- C++ host program over the Cache::Manager-shaped facade:
// builds the 100k-event monotonic-spline workload, bins it on the device,
// takes the Asimov histogram at nominal as data, and evaluates the Poisson
// LLH at a few parameter points, the way GUNDAM's Propagator +
// LikelihoodInterface would drive the cache.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "gcm_cache_manager.hpp"
#include "synth.h"

int main() {
  if (!gcm::CacheManager::HasGPU()) {
    std::fprintf(stderr, "no HIP device\n");
    return 2;
  }
  void* wl = gcm_synth_create(1, 0, 0, 8);
  gcm_synth_view v;
  gcm_synth_view_get(wl, &v);
  gcm::check(gcm_bin_events(0, v.n_events, v.n_vars, v.vars, v.n_bins, v.edges, v.bin),
             "gcm_bin_events");
  gcm_tables t{};
  t.n_events = v.n_events;
  t.base_weight = v.base_weight;
  t.bin = v.bin;
  t.dial_begin = v.dial_begin;
  t.n_dials = v.n_dials;
  t.dial_type = v.dial_type;
  t.dial_input = v.dial_input;
  t.data_begin = v.data_begin;
  t.dial_data = v.dial_data;
  t.n_inputs = v.n_inputs;
  t.inputs = v.inputs;
  t.n_tables = v.n_tables;
  t.table_begin = v.table_begin;
  t.n_parameters = v.n_parameters;
  t.n_bins = v.n_bins;
  t.reweight_cap = v.reweight_cap;
  t.dial_input2 = v.dial_input2;

  gcm::CacheManager cm(0);
  cm.Build(t, 16);  // up to 16 points per launch (the gradient below)
  std::vector<double> nominal(v.prior_mean, v.prior_mean + v.n_parameters);
  cm.PropagateParameters(nominal);
  std::vector<double> data = cm.SumWeights();  // Asimov data
  // Barlow-Beeston OA2021: finite for bins empty at the prior (the Poisson
  // LLH of the reference returns +inf for a bin predicted <= 0)
  gcm_llh_config llh{GCM_LLH_BARLOW_OA2021, 0, 1e150, -1.0};
  cm.SetLikelihood(llh, data, nominal,
                   std::vector<double>(v.prior_sigma, v.prior_sigma + v.n_parameters));
  std::mt19937_64 rng(7);
  std::normal_distribution<double> g(0.0, 1.0);
  double total = 0;
  for (int i = 0; i < 4; ++i) {
    std::vector<double> th(nominal);
    for (int p = 0; p < v.n_parameters; ++p) th[p] += 0.3 * v.prior_sigma[p] * g(rng);
    auto fut = cm.Fill(th);
    fut.get();
    const double l = cm.Likelihood()[0];
    total += l;
    std::printf("point %d: LLH %.10g\n", i, l);
  }
  // a Minuit-style gradient over the first five parameters in one batched
  // launch of 11 points, checked against single-point evaluations
  std::vector<int> params{0, 1, 2, 3, 4};
  std::vector<double> grad;
  std::vector<double> th0(nominal);
  for (int p = 0; p < v.n_parameters; ++p) th0[p] += 0.2 * v.prior_sigma[p];
  const double h = 1e-3;
  const double l0 = cm.Gradient(th0, h, params, &grad);
  double worst = 0;
  for (int j = 0; j < 5; ++j) {
    std::vector<double> tp(th0), tm(th0);
    tp[params[j]] += h;
    tm[params[j]] -= h;
    cm.Fill(tp).get();
    const double lp = cm.Likelihood()[0];
    cm.Fill(tm).get();
    const double lm = cm.Likelihood()[0];
    worst = std::max(worst, std::fabs((lp - lm) / (2 * h) - grad[j]));
  }
  std::printf("gradient at LLH %.10g: d/dp0 %.6g ... max |batched - single| %.3g\n",
              l0, grad[0], worst);
  if (worst != 0.0) {
    std::printf("batched gradient differs from single-point evaluations\n");
    return 1;
  }
  // a correlated prior on parameters 0 and 1 (LikelihoodInterface's per-set
  // penalty): the device's penalty must grow by exactly d^T (C^-1 d), formed
  // in the same order on the host
  {
    const int32_t pars[2] = {0, 1};
    const double prior[2] = {nominal[0], nominal[1]};
    const double inv[4] = {2.0, -0.5, -0.5, 1.5};
    gcm_penalty_set ps{2, pars, prior, inv, nullptr, nullptr};
    cm.Fill(th0).get();
    const double pen0 = cm.Penalty()[0];
    cm.SetPenaltySets({ps});
    cm.Fill(th0).get();
    const double pen1 = cm.Penalty()[0];
    const double d0 = th0[0] - prior[0], d1 = th0[1] - prior[1];
    double y0 = 0, y1 = 0;
    y0 += d0 * inv[0];
    y0 += d1 * inv[1];
    y1 += d0 * inv[2];
    y1 += d1 * inv[3];
    double q = 0;
    q += d0 * (0.0 + y0);
    q += d1 * (0.0 + y1);
    std::printf("penalty %.10g -> %.10g with the correlated set (expected +%.10g)\n", pen0,
                pen1, q);
    if (pen1 != pen0 + q) {
      std::printf("correlated penalty differs\n");
      return 1;
    }
  }
  // the toggles, timers and parameter indices of the reference's interface
  if (gcm::CacheManager::HasCUDA()) {
    std::printf("HasCUDA must be false\n");
    return 1;
  }
  if (cm.ParameterIndex(0) != 0 || cm.ParameterIndex(v.n_parameters) >= 0 ||
      cm.ParameterIndex(-1) >= 0) {
    std::printf("ParameterIndex mismatch\n");
    return 1;
  }
  {
    // with the histogram copy off, the future leaves the last copy untouched
    cm.Fill(nominal).get();
    const std::vector<double> before = cm.SumWeights();
    const bool was = cm.SetIsHistContentCopyEnabled(false);
    cm.SetEnableDebugPrintouts(true);
    cm.Fill(th0).get();
    cm.SetEnableDebugPrintouts(false);
    const bool unchanged = cm.SumWeights() == before;
    cm.SetIsHistContentCopyEnabled(was);
    cm.Fill(th0).get();
    const bool refreshed = cm.SumWeights() != before;
    const auto& ft = cm.GetCacheFillTimer();
    const auto& pt = cm.GetPullFromDeviceTimer();
    std::printf("fill timer: %ld fills, last %.3f ms, mean of last 10 %.3f ms; pull %.3f ms\n",
                ft.Count(), ft.LastMs(), ft.AverageMs(), pt.AverageMs());
    if (!was || !unchanged || !refreshed || ft.Count() != pt.Count() || ft.Count() < 10 ||
        !(ft.AverageMs() > 0)) {
      std::printf("toggle/timer check failed\n");
      return 1;
    }
  }
  std::printf("resident %.1f MiB, nominal LLH check %.3g\n",
              cm.GetResidentMemory() / 1048576.0, total);
  gcm_synth_destroy(wl);
  return 0;
}
