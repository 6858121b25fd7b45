// oracle.cpp -- TEST INFRASTRUCTURE (the oracle), never the product.
//
// A faithful CPU restatement of the reference's CPU reweight-and-fill path,
// with the reference's own data structures, so that it serves both as the
// parity oracle for the HIP path and as the timed CPU baseline:
//
//   Propagator::reweightEvents / refillHistograms  Propagator.cpp:360-393
//     EventDialCache::reweightEntry                EventDialCache.cpp:156-171
//       DialResponseCache::getResponse             EventDialCache.h:50-59
//         DialInterface::evalResponse              DialInterface.h:57
//           DialResponseSupervisor::process        DialResponseSupervisor.cpp:10-18
//           <Dial>::evalResponse (virtual)         DialDefinitions/src/*.cpp
//       DialInputBuffer::update (mirror, fmod)     DialInputBuffer.cpp:44-89
//     Histogram::refillHistogram                   Histogram.cpp:107-135
//   LLH: PoissonLogLikelihood::eval                PoissonLogLikelihood.h:19-53
//        BarlowBeestonBanff2022::eval              BarlowBeestonBanff2022.h:92-294
//        BarlowBeeston::eval (classic)             BarlowBeeston.h:25-62
//        BarlowBeestonBanff2020::eval              BarlowBeestonBanff2020.h:18-85
//        BarlowBeestonBanff2022Sfgd::eval          BarlowBeestonBanff2022Sfgd.h:20-134
//        ChiSquared::eval, LeastSquares::eval      ChiSquared.h:19-29, LeastSquares.h:45-52
//        zero-prediction-at-prior bin mask         JointProbabilityBase.h:56-77
//        LikelihoodInterface penalty + total       LikelihoodInterface.cpp:165-215
//        per-set covariance penalty, bounds        LikelihoodInterface.cpp:186-213,
//                                                  ParameterSet.cpp:441-463
//   Event -> bin: LoaderUtils::fillBinIndex        LoaderUtils.cpp:32-40
//                 Bin::isBetweenEdges              Bin.cpp:167-176
//
// How each part is pinned (the tests named): the dial math bit-exact against
// golden vectors from the reference's own Calculate*.h headers compiled in
// place (tests/test_oracle_golden.py); event assembly against an independent
// Python restatement (tests/test_oracle_path.py); every LLH type against the
// numerical profile minimum or its closed form and the reference unit test's
// sweeps (tests/test_llh_variants.py); the penalty sets against exact
// rational arithmetic and an independent solve (tests/test_penalty_oracle.py).
// No function is left unpinned.
//
// Threading follows the reference partitioning: contiguous event ranges per
// thread for the reweight, contiguous bin ranges per thread for the refill,
// and a single-threaded LLH.  Only tests/, the smoke check and the bench's
// CPU-baseline leg load this library.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../include/gcm.h"
#include "oracle_calc.hpp"

namespace oracle {

// ---------------------------------------------------------------------------
// Dial inputs (DialInputBuffer) and response supervisors.
// ---------------------------------------------------------------------------
struct DialInputBuffer {
  int parameter{-1};
  double mirror_lower{std::nan("")};
  double mirror_range{std::nan("")};
  double buffer{std::nan("unset")};
  bool update_requested{true};

  // DialInputBuffer.cpp:44-89.
  void update(const double* theta) {
    update_requested = false;
    double v = theta[parameter];
    if (!std::isnan(mirror_lower)) {
      v = std::abs(std::fmod(v - mirror_lower, 2 * mirror_range));
      if (v > mirror_range) {
        v -= 2 * mirror_range;
        v = -v;
      }
      v += mirror_lower;
    }
    if (buffer != v) {
      update_requested = true;
      buffer = v;
    }
  }
};

struct DialResponseSupervisor {
  double min_response{std::nan("unset")};
  double max_response{std::nan("unset")};
  // DialResponseSupervisor.cpp:10-18.
  double process(double r) const {
    if (!std::isnan(min_response) && r < min_response) return min_response;
    else if (!std::isnan(max_response) && r > max_response) return max_response;
    return r;
  }
};

// ---------------------------------------------------------------------------
// Dials (DialBase hierarchy), evaluated through virtual dispatch as the
// reference does.
// ---------------------------------------------------------------------------
struct DialBase {
  virtual ~DialBase() = default;
  virtual double evalResponse(const DialInputBuffer& in) const = 0;
};

struct NormDial : DialBase {  // Norm.h:21
  double evalResponse(const DialInputBuffer& in) const override {
    return in.buffer;
  }
};

struct ShiftDial : DialBase {  // Shift.h:21
  double shift{1};
  double evalResponse(const DialInputBuffer&) const override { return shift; }
};

// Shared by the uniformly spaced splines: the record and the bounds used when
// extrapolation is forbidden (MonotonicSpline.cpp:78-81).
struct UniformKnotDial : DialBase {
  std::vector<double> data;
  bool allow_extrapolation{true};
  double bound_min{0}, bound_max{0};
  double input(const DialInputBuffer& in) const {
    double x = in.buffer;
    if (!allow_extrapolation) {
      if (x <= bound_min) x = bound_min;
      else if (x >= bound_max) x = bound_max;
    }
    return x;
  }
};

struct MonotonicDial : UniformKnotDial {  // MonotonicSpline.cpp:71-84
  double evalResponse(const DialInputBuffer& in) const override {
    return calc_monotonic(input(in), -1E20, 1E20, data.data(),
                          int(data.size() - 2));
  }
};

struct CompactDial : UniformKnotDial {  // CompactSpline.cpp:71-84
  double evalResponse(const DialInputBuffer& in) const override {
    return calc_compact(input(in), -1E20, 1E20, data.data(),
                        int(data.size() - 2));
  }
};

struct UniformDial : UniformKnotDial {  // UniformSpline.cpp:47-60
  double evalResponse(const DialInputBuffer& in) const override {
    return calc_uniform(input(in), -1E20, 1E20, data.data(), int(data.size()));
  }
};

struct GeneralDial : UniformKnotDial {  // GeneralSpline.cpp:33-42
  double evalResponse(const DialInputBuffer& in) const override {
    return calc_general(input(in), -1E20, 1E20, data.data(), int(data.size()));
  }
};

struct GraphDial : DialBase {  // Graph.cpp:41-54
  std::vector<double> data;
  bool allow_extrapolation{true};
  double evalResponse(const DialInputBuffer& in) const override {
    double x = in.buffer;
    if (!allow_extrapolation) {
      if (x <= data[1]) return data[0];
      else if (x >= data.back()) return data[data.size() - 2];
    }
    return calc_graph(x, -1E20, 1E20, data.data(), int(data.size()));
  }
};

struct TabulatedDial : DialBase {  // Tabulated.h:16-28
  const std::vector<double>* table{nullptr};
  double fraction{0};  // holds a float-rounded value
  int index{0};
  double evalResponse(const DialInputBuffer&) const override {
    return fraction * (*table)[index] + (1.0 - fraction) * (*table)[index + 1];
  }
};

struct KrigedDial : DialBase {  // Kriged.cpp:14-21
  const std::vector<double>* table{nullptr};
  std::vector<std::pair<int, float>> weights;
  double evalResponse(const DialInputBuffer&) const override {
    const std::vector<double>& tbl = *table;
    double v{0.0};
    for (const std::pair<int, float>& w : weights) v += w.second * tbl[w.first];
    return v;
  }
};

// Bilinear.cpp:40-64 and Bicubic.cpp:40-64: two inputs (DialInputBuffer
// slots 0 and 1), clamped to the knot range unless extrapolation is allowed.
struct Dial2D : DialBase {
  std::vector<double> data;  // nx, ny, x knots, y knots, values
  const DialInputBuffer* second{nullptr};
  bool allow_extrapolation{false};
  bool bicubic{false};
  double evalResponse(const DialInputBuffer& in) const override {
    double input0 = in.buffer;
    double input1 = second->buffer;
    const int nx = int(data[0]), ny = int(data[1]);
    const double* xx = data.data() + 2;
    const double* yy = xx + nx;
    if (!allow_extrapolation) {
      if (input0 < xx[0]) input0 = xx[0];
      if (input0 > xx[nx - 1]) input0 = xx[nx - 1];
      if (input1 < yy[0]) input1 = yy[0];
      if (input1 > yy[ny - 1]) input1 = yy[ny - 1];
    }
    return bicubic ? calc_bicubic(input0, input1, -1E20, 1E20, data.data())
                   : calc_bilinear(input0, input1, -1E20, 1E20, data.data());
  }
};

// DialInterface.h:57.
struct DialInterface {
  std::unique_ptr<DialBase> dial;
  DialInputBuffer* input{nullptr};
  const DialResponseSupervisor* supervisor{nullptr};
  double evalResponse() const {
    return supervisor->process(dial->evalResponse(*input));
  }
};

struct Event {
  double base{1};
  double current{1};
  int bin{-1};
};

// EventDialCache.h:38-60 and EventDialCache.cpp:156-171.
struct DialResponseCache {
  DialInterface* iface;
  double response{std::nan("unset")};
  bool* update_requested;
  double getResponse() {
    if (*update_requested) response = iface->evalResponse();
    return response;
  }
};

struct CacheEntry {
  Event* event;
  std::vector<DialResponseCache> dials;
};

struct BinContent {
  double sumWeights{0};
  double sqrtSumSqWeights{0};
};

struct Model {
  int n_threads{1};
  int n_parameters{0};
  int n_bins{0};
  bool cap_enabled{false};
  double cap{0};
  std::vector<DialInputBuffer> inputs;
  std::vector<DialResponseSupervisor> supervisors;
  std::vector<std::vector<double>> tables;
  std::vector<Event> events;
  std::vector<DialInterface> interfaces;
  std::vector<CacheEntry> cache;
  std::vector<std::vector<Event*>> bin_events;  // Histogram BinContext lists
  bool always{true};  // 2-D dials re-evaluate on every propagation
  std::vector<BinContent> hist;

  void reweightEntry(CacheEntry& entry) {
    double temp{1};
    for (auto& d : entry.dials) temp *= d.getResponse();
    if (cap_enabled && temp > cap) temp = cap;
    entry.event->current = entry.event->base;
    entry.event->current *= temp;
  }

  // Histogram.cpp:107-135 for the bins [b0, b1).
  void refill(int b0, int b1) {
    double w;
    for (int b = b0; b < b1; ++b) {
      BinContent& c = hist[b];
      c.sumWeights = 0;
      c.sqrtSumSqWeights = 0;
      for (Event* e : bin_events[b]) {
        w = e->current;
        c.sumWeights += w;
        c.sqrtSumSqWeights += w * w;
      }
      c.sqrtSumSqWeights = std::sqrt(c.sqrtSumSqWeights);
    }
  }
};

// Contiguous [begin, end) share of n items for thread i of n_threads.
static void thread_bounds(int i, int n_threads, int64_t n, int64_t* b,
                          int64_t* e) {
  int64_t chunk = n / n_threads;
  int64_t rest = n % n_threads;
  *b = i * chunk + std::min<int64_t>(i, rest);
  *e = *b + chunk + (i < rest ? 1 : 0);
}

template <typename F>
static void parallel(int n_threads, F&& f) {
  if (n_threads <= 1) {
    f(0);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(n_threads);
  for (int i = 0; i < n_threads; ++i) pool.emplace_back([&f, i] { f(i); });
  for (auto& t : pool) t.join();
}

thread_local std::string g_error;

}  // namespace oracle

using namespace oracle;

extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

// Build the reference object graph from the same tables the device consumes.
void* oracle_create(const gcm_tables* t, int n_threads) {
  try {
    auto m = std::make_unique<Model>();
    m->n_threads = std::max(1, n_threads);
    m->n_parameters = t->n_parameters;
    m->n_bins = t->n_bins;
    m->cap_enabled = std::isfinite(t->reweight_cap);
    m->cap = t->reweight_cap;
    m->inputs.resize(t->n_inputs);
    m->supervisors.resize(t->n_inputs);
    for (int i = 0; i < t->n_inputs; ++i) {
      const gcm_input& in = t->inputs[i];
      if (in.parameter < 0 || in.parameter >= t->n_parameters) {
        g_error = "input parameter out of range";
        return nullptr;
      }
      m->inputs[i].parameter = in.parameter;
      m->inputs[i].mirror_lower = in.mirror_lower;
      m->inputs[i].mirror_range = in.mirror_range;
      m->supervisors[i].min_response = in.min_response;
      m->supervisors[i].max_response = in.max_response;
    }
    m->tables.resize(t->n_tables);
    for (int i = 0; i < t->n_tables; ++i)
      m->tables[i].assign(t->table_begin[i + 1] - t->table_begin[i], 1.0);

    m->events.resize(t->n_events);
    m->interfaces.resize(t->n_dials);
    for (int64_t d = 0; d < t->n_dials; ++d) {
      const double* p = t->dial_data + t->data_begin[d];
      const int64_t len = t->data_begin[d + 1] - t->data_begin[d];
      const int in = t->dial_input[d];
      if (in < 0 || in >= t->n_inputs) {
        g_error = "dial input out of range";
        return nullptr;
      }
      const bool extrap = !(t->inputs[in].flags & GCM_INPUT_NO_EXTRAPOLATION);
      std::unique_ptr<DialBase> dial;
      switch (t->dial_type[d]) {
        case GCM_DIAL_NORM: dial = std::make_unique<NormDial>(); break;
        case GCM_DIAL_SHIFT: {
          auto s = std::make_unique<ShiftDial>();
          s->shift = p[0];
          dial = std::move(s);
          break;
        }
        case GCM_DIAL_MONOTONIC:
        case GCM_DIAL_COMPACT:
        case GCM_DIAL_UNIFORM:
        case GCM_DIAL_GENERAL: {
          std::unique_ptr<UniformKnotDial> s;
          int knots;
          switch (t->dial_type[d]) {
            case GCM_DIAL_MONOTONIC: s = std::make_unique<MonotonicDial>(); knots = int(len - 2); break;
            case GCM_DIAL_COMPACT: s = std::make_unique<CompactDial>(); knots = int(len - 2); break;
            case GCM_DIAL_UNIFORM: s = std::make_unique<UniformDial>(); knots = int((len - 2) / 2); break;
            default: s = std::make_unique<GeneralDial>(); knots = int((len - 2) / 3); break;
          }
          s->data.assign(p, p + len);
          s->allow_extrapolation = extrap;
          s->bound_min = p[0];
          if (t->dial_type[d] == GCM_DIAL_GENERAL)
            s->bound_max = p[2 + 3 * (knots - 1) + 2];
          else
            s->bound_max = p[0] + (knots - 1) * p[1];
          dial = std::move(s);
          break;
        }
        case GCM_DIAL_GRAPH: {
          auto g = std::make_unique<GraphDial>();
          g->data.assign(p, p + len);
          g->allow_extrapolation = extrap;
          dial = std::move(g);
          break;
        }
        case GCM_DIAL_TABULATED: {
          auto tb = std::make_unique<TabulatedDial>();
          tb->table = &m->tables.at(int(p[0]));
          tb->index = int(p[1]);
          tb->fraction = (float)p[2];
          dial = std::move(tb);
          break;
        }
        case GCM_DIAL_KRIGED: {
          auto kr = std::make_unique<KrigedDial>();
          kr->table = &m->tables.at(int(p[0]));
          for (int64_t k = 1; k + 1 < len; k += 2)
            kr->weights.emplace_back(int(p[k]), float(p[k + 1]));
          dial = std::move(kr);
          break;
        }
        case GCM_DIAL_BILINEAR:
        case GCM_DIAL_BICUBIC: {
          const int in2 = t->dial_input2 ? t->dial_input2[d] : -1;
          if (in2 < 0 || in2 >= t->n_inputs) {
            g_error = "2-D dial second input out of range";
            return nullptr;
          }
          const int nx = len >= 2 ? int(p[0]) : 0, ny = len >= 2 ? int(p[1]) : 0;
          if (nx < 2 || ny < 2 || len != 2 + nx + ny + (int64_t)nx * ny) {
            g_error = "2-D dial needs nx, ny >= 2 and 2 + nx + ny + nx*ny values";
            return nullptr;
          }
          auto g2 = std::make_unique<Dial2D>();
          g2->data.assign(p, p + len);
          g2->second = &m->inputs[in2];
          g2->allow_extrapolation = extrap;
          g2->bicubic = t->dial_type[d] == GCM_DIAL_BICUBIC;
          dial = std::move(g2);
          break;
        }
        default:
          g_error = "unsupported dial type";
          return nullptr;
      }
      m->interfaces[d].dial = std::move(dial);
      m->interfaces[d].input = &m->inputs[in];
      m->interfaces[d].supervisor = &m->supervisors[in];
    }
    m->cache.resize(t->n_events);
    m->bin_events.resize(t->n_bins);
    for (int64_t e = 0; e < t->n_events; ++e) {
      Event& ev = m->events[e];
      ev.base = t->base_weight[e];
      ev.current = ev.base;
      ev.bin = t->bin[e];
      CacheEntry& ce = m->cache[e];
      ce.event = &ev;
      for (int64_t d = t->dial_begin[e]; d < t->dial_begin[e + 1]; ++d) {
        DialInterface* di = &m->interfaces[d];
        const bool two_d = t->dial_type[d] == GCM_DIAL_BILINEAR ||
                           t->dial_type[d] == GCM_DIAL_BICUBIC;
        ce.dials.push_back(DialResponseCache{di, std::nan("unset"),
                                             two_d ? &m->always : &di->input->update_requested});
      }
      if (ev.bin >= 0) {
        if (ev.bin >= t->n_bins) {
          g_error = "bin out of range";
          return nullptr;
        }
        m->bin_events[ev.bin].push_back(&ev);
      }
    }
    m->hist.resize(t->n_bins);
    return m.release();
  } catch (const std::exception& ex) {
    g_error = ex.what();
    return nullptr;
  }
}

void oracle_destroy(void* h) { delete static_cast<Model*>(h); }

// Tabulated/Kriged tables for the next evaluation (the factory callback
// output, TabulatedDialFactory.cpp:133-141).
int oracle_set_table_values(void* h, const double* values, int64_t len) {
  Model* m = static_cast<Model*>(h);
  int64_t off = 0;
  for (auto& t : m->tables) {
    if (off + (int64_t)t.size() > len) return 1;
    std::copy(values + off, values + off + t.size(), t.begin());
    off += t.size();
  }
  return 0;
}

// One propagation: input-buffer update, reweight (event ranges per thread),
// refill (bin ranges per thread).  Outputs are optional.
int oracle_propagate(void* h, const double* theta, double* weights,
                     double* sums, double* sqrt_sum_sq) {
  Model* m = static_cast<Model*>(h);
  for (auto& in : m->inputs) in.update(theta);
  const int64_t n = (int64_t)m->cache.size();
  parallel(m->n_threads, [&](int i) {
    int64_t b, e;
    thread_bounds(i, m->n_threads, n, &b, &e);
    for (int64_t k = b; k < e; ++k) m->reweightEntry(m->cache[k]);
  });
  parallel(m->n_threads, [&](int i) {
    int64_t b, e;
    thread_bounds(i, m->n_threads, m->n_bins, &b, &e);
    m->refill(int(b), int(e));
  });
  if (weights)
    for (int64_t k = 0; k < n; ++k) weights[k] = m->events[k].current;
  for (int b = 0; b < m->n_bins; ++b) {
    if (sums) sums[b] = m->hist[b].sumWeights;
    if (sqrt_sum_sq) sqrt_sum_sq[b] = m->hist[b].sqrtSumSqWeights;
  }
  return 0;
}

// PoissonLogLikelihood.h:19-53 (err unused).
double oracle_llh_poisson_bin(double data, double pred) {
  double llh{0};
  if (pred <= 0) llh = std::numeric_limits<double>::infinity();
  else if (data <= 0) llh = 2.0 * pred;
  else llh = 2.0 * (pred - data + data * std::log(data / pred));
  return llh;
}

// BarlowBeestonBanff2022.h:150-294 with BBNoUpdateWeights = false.
// err is sqrtSumSqWeights; the reference squares it (:125-126).
double oracle_llh_barlow_bin(double data, double pred, double err,
                             const gcm_llh_config* cfg) {
  double mcuncert = err;
  mcuncert *= mcuncert;
  if (!std::isfinite(mcuncert) || mcuncert < 0.0)
    return std::numeric_limits<double>::infinity();
  if (pred <= cfg->expected_value_minimum) {
    if (pred <= 0.0 && data > 0.0) return std::numeric_limits<double>::infinity();
    return 0.0;
  }
  double newmc = pred;
  double penalty = 0.0;
  if (mcuncert > 0.0) {
    double fractional =
        std::min(std::sqrt(mcuncert) / pred, cfg->fractional_error_limit);
    double temp = pred * fractional * fractional - 1;
    double temp2 = temp * temp + 4 * data * fractional * fractional;
    if (temp2 < 0) return std::numeric_limits<double>::infinity();
    double beta = (-1 * temp + std::sqrt(temp2)) / 2.;
    newmc = std::max(pred * beta, std::numeric_limits<double>::min());
    penalty = (beta - 1) * (beta - 1) / (2 * fractional * fractional);
  }
  double stat = 0;
  if (data <= std::numeric_limits<double>::epsilon()) {
    stat = newmc;
  } else if (newmc > std::numeric_limits<double>::min()) {
    stat = newmc - data + data * (std::log(data) - std::log(newmc));
  } else {
    const double mc = std::numeric_limits<double>::min();
    stat = mc - data + data * (std::log(data) - std::log(mc));
  }
  double chisq = 2.0 * stat;
  if (!cfg->use_poisson_likelihood) chisq += 2.0 * penalty;
  return chisq;
}

// BarlowBeeston.h:33-62 ("classic"; err is sqrtSumSqWeights as the base
// class passes it, used as a variance).
double oracle_llh_barlow_classic_bin(double data, double pred, double err) {
  const double dataVal = std::max(data, 0.0);
  const double predVal = std::max(pred, std::sqrt(std::numeric_limits<double>::min()));
  const double mcUncert = std::max(err, std::numeric_limits<double>::min());
  const double rel_var = mcUncert / (predVal * predVal);
  const double b = (predVal * rel_var) - 1;
  const double c = 4 * dataVal * rel_var;
  const double beta = (-b + std::sqrt(b * b + c)) / 2.0;
  const double mc_hat = predVal * beta;
  double chi2 = 0.0;
  if (dataVal <= 0.0) {
    chi2 = 2 * mc_hat;
    chi2 += (beta - 1) * (beta - 1) / rel_var;
  } else {
    chi2 = 2 * (mc_hat - dataVal);
    if (dataVal > 0.0) chi2 += 2 * dataVal * (std::log(dataVal) - std::log(mc_hat));
    chi2 += (beta - 1) * (beta - 1) / rel_var;
  }
  return chi2;
}

// BarlowBeestonBanff2020.h:18-85 (fractional = sqrt(err) / pred, err as the
// base class passes it) and BarlowBeestonBanff2022Sfgd.h:20-134 (fractional
// = (err + pred x detector fraction) / pred).  The reference throws where
// this returns NaN.
double oracle_llh_barlow_banff_bin(int type, double data, double pred, double err,
                                   double extra_fraction) {
  double fractional;
  if (type == GCM_LLH_BARLOW_OA2020) {
    fractional = std::sqrt(err) / pred;
  } else {
    const double sfgUncert = err + pred * extra_fraction;
    fractional = sfgUncert / pred;
  }
  const double temp = pred * fractional * fractional - 1;
  const double temp2 = temp * temp + 4 * data * fractional * fractional;
  if (temp2 < 0) return std::numeric_limits<double>::quiet_NaN();
  const double beta = (-1 * temp + std::sqrt(temp2)) / 2.;
  const double newmc = pred * beta;
  const double penalty = (beta - 1) * (beta - 1) / (2 * fractional * fractional);
  double stat = 0;
  if (data == 0) stat = newmc;
  else if (newmc > 0) stat = newmc - data + data * std::log(data / newmc);
  if (type == GCM_LLH_BARLOW_OA2020 && std::isnan(penalty))
    return std::numeric_limits<double>::infinity();
  double chisq = 0.;
  if ((pred > 0.0) && (data > 0.0)) chisq += 2.0 * (stat + penalty);
  else if (pred > 0.0) chisq += 2.0 * (stat + penalty);
  if (type == GCM_LLH_BARLOW_OA2020 && std::isnan(chisq))
    return std::numeric_limits<double>::quiet_NaN();
  return chisq;
}

// ChiSquared.h:19-29.
double oracle_llh_chi2_bin(double data, double pred) {
  if (pred == 0) return std::numeric_limits<double>::infinity();
  return (pred - data) * (pred - data) / pred;
}

// LeastSquares.h:45-52.
double oracle_llh_least_squares_bin(double data, double pred, int poissonian) {
  double v = data - pred;
  v = v * v;
  if (poissonian && data > 1.0) v /= 0.5 * data;
  return v;
}

// One bin of any type: JointProbabilityBase::eval(SamplePair, bin)
// (JointProbabilityBase.h:37-43) with the bin mask of :74-77 and the OA2021
// BBNoUpdateWeights switch (BarlowBeestonBanff2022.h:109-123).
double oracle_llh_bin(const gcm_llh_config* cfg, int b, double data, double pred,
                      double err) {
  if (cfg->ignore_bins && cfg->ignore_bins[b]) return 0.0;
  switch (cfg->type) {
    case GCM_LLH_POISSON: return oracle_llh_poisson_bin(data, pred);
    case GCM_LLH_BARLOW_OA2021:
      return oracle_llh_barlow_bin(
          data, pred, (cfg->bb_no_update_weights && cfg->nominal_err) ? cfg->nominal_err[b] : err,
          cfg);
    case GCM_LLH_BARLOW_BEESTON: return oracle_llh_barlow_classic_bin(data, pred, err);
    case GCM_LLH_BARLOW_OA2020:
    case GCM_LLH_BARLOW_OA2021_SFGD:
      return oracle_llh_barlow_banff_bin(cfg->type, data, pred, err,
                                         cfg->extra_fraction ? cfg->extra_fraction[b] : 0.0);
    case GCM_LLH_CHI_SQUARED: return oracle_llh_chi2_bin(data, pred);
    case GCM_LLH_LEAST_SQUARES:
      return oracle_llh_least_squares_bin(data, pred, cfg->lsq_poissonian_approximation);
    default: return std::numeric_limits<double>::quiet_NaN();
  }
}

// JointProbabilityBase::eval(SamplePair) summed over bins in order, plus the
// diagonal-prior penalty of LikelihoodInterface.cpp:179-215; total =
// stat + penalty.
int oracle_llh(const gcm_llh_config* cfg, int n_bins, const double* data,
               const double* sums, const double* sqrt_sum_sq, int n_par,
               const double* theta, const double* prior_mean,
               const double* prior_sigma, double* total, double* stat_out,
               double* penalty_out) {
  double stat = 0;
  for (int b = 0; b < n_bins; ++b)
    stat += oracle_llh_bin(cfg, b, data[b], sums[b], sqrt_sum_sq[b]);
  double pen = 0;
  for (int p = 0; p < n_par; ++p) {
    if (!prior_sigma || !(prior_sigma[p] > 0)) continue;
    double z = (theta[p] - prior_mean[p]) / prior_sigma[p];
    pen += z * z;
  }
  if (stat_out) *stat_out = stat;
  if (penalty_out) *penalty_out = pen;
  if (total) *total = stat + pen;
  return 0;
}

// LikelihoodInterface::evalPenaltyLikelihood(const ParameterSet&)
// (LikelihoodInterface.cpp:186-213) for a set with a prior covariance and no
// eigen decomposition: the delta vector of the valid correlated parameters
// (ParameterSet::updateDeltaVector, ParameterSet.cpp:456-463), the inverse
// covariance times it (ROOT's TMatrixD * TVectorD: per row a sum from 0 over
// the columns in order, added to the zeroed target), then the dot product
// from 0 in index order.  With bounds (applyInfinitePenaltyOnOutOfBoundPar,
// ParameterSet.cpp:441-453: a NaN value or one outside [lower, upper] is out
// of bounds; a NaN bound is no bound) an out-of-bounds set gives +inf.
// inv_cov NULL: no quadratic term (a set without a covariance matrix).
double oracle_penalty_set(int n, const double* theta, const int32_t* parameters,
                          const double* prior, const double* inv_cov,
                          const double* lower, const double* upper) {
  std::vector<double> delta(n), y(n, 0.0);
  for (int i = 0; i < n; ++i) delta[i] = theta[parameters[i]] - prior[i];
  double buffer = 0;
  if (inv_cov) {
    for (int i = 0; i < n; ++i) {
      double sum = 0;
      for (int j = 0; j < n; ++j) sum += delta[j] * inv_cov[(int64_t)i * n + j];
      y[i] += sum;
    }
    double dot = 0;
    for (int i = 0; i < n; ++i) dot += delta[i] * y[i];
    buffer = dot;
  }
  if (lower && upper) {
    bool out = false;
    for (int i = 0; i < n; ++i) {
      const double v = theta[parameters[i]];
      if (std::isnan(v) || v < lower[i] || v > upper[i]) out = true;
    }
    if (out) return std::numeric_limits<double>::infinity();
  }
  return buffer;
}

// LoaderUtils.cpp:32-40 with Bin::isBetweenEdges (Bin.cpp:167-176), default
// [min, max) edges.
int oracle_bin_events(int64_t n_events, int n_vars, const double* values,
                      int n_bins, const double* edges, int32_t* out) {
  for (int64_t e = 0; e < n_events; ++e) {
    const double* v = values + e * n_vars;
    int found = -1;
    for (int b = 0; b < n_bins && found < 0; ++b) {
      bool in = true;
      for (int k = 0; k < n_vars && in; ++k) {
        const double lo = edges[(int64_t(b) * n_vars + k) * 2];
        const double hi = edges[(int64_t(b) * n_vars + k) * 2 + 1];
        if (lo == hi) in = (lo == v[k]);
        else if (lo > v[k]) in = false;
        else if (hi <= v[k]) in = false;
      }
      if (in) found = b;
    }
    out[e] = found;
  }
  return 0;
}

// Direct access to the restated Calculate* functions for the golden checks.
// One 2-D dial evaluation (the golden-vector check): data is the payload
// layout [nx, ny, x knots, y knots, values].
double oracle_calc2d(int type, double x, double y, double lo, double hi,
                     const double* data) {
  if (type == GCM_DIAL_BILINEAR) return calc_bilinear(x, y, lo, hi, data);
  if (type == GCM_DIAL_BICUBIC) return calc_bicubic(x, y, lo, hi, data);
  return std::nan("");
}

double oracle_calc(int type, double x, double lo, double hi, const double* data,
                   int dim) {
  switch (type) {
    case GCM_DIAL_MONOTONIC: return calc_monotonic(x, lo, hi, data, dim);
    case GCM_DIAL_COMPACT: return calc_compact(x, lo, hi, data, dim);
    case GCM_DIAL_UNIFORM: return calc_uniform(x, lo, hi, data, dim);
    case GCM_DIAL_GENERAL: return calc_general(x, lo, hi, data, dim);
    case GCM_DIAL_GRAPH: return calc_graph(x, lo, hi, data, dim);
    default: return std::nan("");
  }
}

}  // extern "C"
