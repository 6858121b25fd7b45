// synth.cpp -- seeded synthetic event/dial tables for the benchmark and test
// workloads (host only).  Produces exactly the arrays that gcm_build takes,
// so the device path and the CPU restatement consume identical bytes.
//
// Randomness: SplitMix64 seeds one xoshiro256** stream per chunk of 4096
// events, so the output does not depend on the number of generator threads.
// Exp(mu) = -mu*log1p(-u) with u a 53-bit uniform; N(0,1) by Box-Muller.
//
// Workloads (event kinematics p ~ Exp(0.8) clipped at 5, cos(theta) ~
// 1 - Exp(0.3) clipped at -1; base weight 1):
//   1: 100k events, 1 sample, 20x20 bins, 20 parameters; per event 4
//      monotonic splines (7 knots on [-3,3]) on distinct parameters 0..15 and
//      one norm on parameter 16 + p quartile.
//   2: 5M events, samples 2M/1.5M/1M/0.5M of 30x50 bins, 300 parameters
//      (30 norm), 12 dials per event from the seven-type mix.
//   3: the workload-1 events re-split into samples of 50k/30k/15k/5k.
//   4: 20M events, 6 samples of 60x50 bins, 500 parameters (50 norm), 15
//      dials per event from the same mix.
//   0: a small feature workload: every dial type including shifts, mirrored
//      inputs, response caps, a parameter read by two collections with
//      different clamps, no-extrapolation inputs and a reweight cap.
// n_events > 0 rescales the sample sizes proportionally.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gcm.h"
#include "synth.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
};

struct Xoshiro256ss {
  uint64_t s[4];
  explicit Xoshiro256ss(uint64_t seed) {
    SplitMix64 sm{seed};
    for (auto& v : s) v = sm.next();
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return double(next() >> 11) * 0x1.0p-53; }
  double exponential(double mu) { return -mu * std::log1p(-uniform()); }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 <= 0) u1 = 0x1.0p-53;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  int integer(int n) { return int(uniform() * n); }  // [0, n)
};

enum ParType : int { P_MONO = 0, P_NORM, P_UNIFORM, P_GENERAL, P_COMPACT,
                     P_GRAPH, P_TABULATED, P_KRIGED, P_NTYPES };

struct Spec {
  int id = 2;
  std::vector<int64_t> sample_events;
  int p_bins = 30, c_bins = 50;
  int n_par = 300, n_norm = 30;
  int dials_per_event = 12;
  bool mixed = true;  // seven-type mix; false: the monotonic+norm workload
  bool features = false;
  uint64_t seed = 20261203;
};

struct Chunk {
  std::vector<double> vars;  // [n][3]: sample, p, cos(theta)
  std::vector<double> base;
  std::vector<int64_t> n_dials;
  std::vector<uint8_t> type;
  std::vector<int32_t> input;
  std::vector<int32_t> input2;  // second input of 2-D dials, else -1
  std::vector<int64_t> data_len;
  std::vector<double> data;
};

struct Synth {
  Spec spec;
  int64_t n_events = 0;
  int n_bins = 0;
  // parameters
  std::vector<int> par_type;
  std::vector<double> prior_mean, prior_sigma;
  std::vector<std::vector<int>> pool;  // parameters per type
  // inputs
  std::vector<gcm_input> inputs;
  std::vector<int> par_input;        // default input of each parameter
  std::vector<int> par_input_alt;    // second collection (features only)
  // tables
  std::vector<int> par_table;        // table id of a tabulated/kriged par
  std::vector<int64_t> table_begin;
  std::vector<int> table_par;
  std::vector<int> table_kind;       // 0: tabulated (64), 1: kriged (256)
  // events and dials
  std::vector<double> vars;
  std::vector<double> base;
  std::vector<int32_t> bin;
  std::vector<int64_t> dial_begin;
  std::vector<uint8_t> type;
  std::vector<int32_t> input;
  std::vector<int32_t> input2;
  std::vector<int64_t> data_begin;
  std::vector<double> data;
  std::vector<double> edges;         // [n_bins][3][2]
  double cap = std::nan("");
};

double knot_y(Xoshiro256ss& r, double x) {
  return std::max(1.0 + 0.1 * r.normal() * x, 1e-3);
}

// DialUtils.cpp:95-125 Catmull-Rom slopes with getSlope = dy/dx
// (DialUtils.cpp:304-306).
void catmull_rom(const std::vector<double>& x, const std::vector<double>& y,
                 std::vector<double>& s) {
  const size_t n = x.size();
  s.assign(n, 0.0);
  if (n < 2) return;
  s[0] = (y[1] - y[0]) / (x[1] - x[0]);
  s[n - 1] = (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]);
  for (size_t i = 1; i + 1 < n; ++i)
    s[i] = (y[i + 1] - y[i - 1]) / (x[i + 1] - x[i - 1]);
}

// Sorted abscissae on [-3, 3] with fixed ends and gaps >= 0.05.
void random_knots(Xoshiro256ss& r, int n, std::vector<double>& x) {
  for (;;) {
    x.assign(n, 0.0);
    x[0] = -3.0;
    x[n - 1] = 3.0;
    for (int i = 1; i < n - 1; ++i) x[i] = -3.0 + 6.0 * r.uniform();
    std::sort(x.begin(), x.end());
    bool ok = true;
    for (int i = 1; i < n; ++i) ok = ok && (x[i] - x[i - 1] >= 0.05);
    if (ok) return;
  }
}

void emit_dial(Synth& S, Chunk& c, Xoshiro256ss& r, int kind, int par,
               int in) {
  std::vector<double> x, y, s;
  int64_t start = (int64_t)c.data.size();
  uint8_t t = 0;
  switch (kind) {
    case P_NORM: t = GCM_DIAL_NORM; break;
    case P_MONO:
    case P_COMPACT: {
      t = (kind == P_MONO) ? GCM_DIAL_MONOTONIC : GCM_DIAL_COMPACT;
      const int n = 7;
      c.data.push_back(-3.0);
      c.data.push_back((3.0 - -3.0) / (n - 1.0));
      for (int k = 0; k < n; ++k) c.data.push_back(knot_y(r, -3.0 + k));
      break;
    }
    case P_UNIFORM: {
      t = GCM_DIAL_UNIFORM;
      const int n = 9;
      for (int k = 0; k < n; ++k) x.push_back(-3.0 + k * 0.75);
      for (int k = 0; k < n; ++k) y.push_back(knot_y(r, x[k]));
      catmull_rom(x, y, s);
      c.data.push_back(x.front());
      c.data.push_back((x.back() - x.front()) / (n - 1.0));
      for (int k = 0; k < n; ++k) {
        c.data.push_back(y[k]);
        c.data.push_back(s[k]);
      }
      break;
    }
    case P_GENERAL: {
      t = GCM_DIAL_GENERAL;
      const int n = 5 + r.integer(11);
      random_knots(r, n, x);
      for (int k = 0; k < n; ++k) y.push_back(knot_y(r, x[k]));
      catmull_rom(x, y, s);
      c.data.push_back(x.front());
      c.data.push_back((x.back() - x.front()) / (n - 1.0));
      for (int k = 0; k < n; ++k) {
        c.data.push_back(y[k]);
        c.data.push_back(s[k]);
        c.data.push_back(x[k]);
      }
      break;
    }
    case P_GRAPH: {
      t = GCM_DIAL_GRAPH;
      const int n = 3 + r.integer(5);
      random_knots(r, n, x);
      for (int k = 0; k < n; ++k) {
        c.data.push_back(knot_y(r, x[k]));
        c.data.push_back(x[k]);
      }
      break;
    }
    case P_TABULATED: {
      t = GCM_DIAL_TABULATED;
      const double u = 63.0 * r.uniform();
      const int i = int(u);
      c.data.push_back(S.par_table[par]);
      c.data.push_back(i);
      c.data.push_back((double)(float)(u - i));
      break;
    }
    case P_KRIGED: {
      t = GCM_DIAL_KRIGED;
      c.data.push_back(S.par_table[par]);
      double e[8], tot = 0;
      int idx[8];
      for (int k = 0; k < 8; ++k) {
        idx[k] = r.integer(256);
        e[k] = -std::log1p(-r.uniform());
        tot += e[k];
      }
      for (int k = 0; k < 8; ++k) {
        c.data.push_back(idx[k]);
        c.data.push_back((double)(float)(e[k] / tot));
      }
      break;
    }
  }
  c.type.push_back(t);
  c.input.push_back(in);
  c.input2.push_back(-1);
  c.data_len.push_back((int64_t)c.data.size() - start);
}

// A 2-D dial (Bilinear/Bicubic) on the parameters behind inputs in (x) and
// in2 (y): nx in [2, 18] (past the 16 knots of the reference's index search),
// ny in [2, 6], sorted knots, smooth positive values.
void emit_dial2d(Chunk& c, Xoshiro256ss& r, int in, int in2) {
  const uint8_t t = r.uniform() < 0.5 ? GCM_DIAL_BILINEAR : GCM_DIAL_BICUBIC;
  const int nx = 2 + r.integer(17), ny = 2 + r.integer(5);
  std::vector<double> xk, yk;
  random_knots(r, nx, xk);
  random_knots(r, ny, yk);
  for (double& v : yk) v *= 0.8;
  const int64_t start = (int64_t)c.data.size();
  c.data.push_back(nx);
  c.data.push_back(ny);
  for (double v : xk) c.data.push_back(v);
  for (double v : yk) c.data.push_back(v);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      c.data.push_back(std::max(1.0 + 0.1 * r.normal() * (xk[i] + 0.5 * yk[j]), 0.05));
  c.type.push_back(t);
  c.input.push_back(in);
  c.input2.push_back(in2);
  c.data_len.push_back((int64_t)c.data.size() - start);
}

void make_chunk(Synth& S, int64_t e0, int64_t e1, Chunk& c) {
  const Spec& sp = S.spec;
  Xoshiro256ss r(sp.seed * 0x100000001B3ull + (uint64_t)(e0 / 4096) + 1);
  // sample of each event: the samples are contiguous event ranges
  std::vector<int64_t> sample_end(sp.sample_events.size());
  int64_t acc = 0;
  for (size_t s = 0; s < sp.sample_events.size(); ++s)
    sample_end[s] = (acc += sp.sample_events[s]);
  std::vector<int> used;
  for (int64_t e = e0; e < e1; ++e) {
    int smp = 0;
    while (e >= sample_end[smp]) ++smp;
    const double p = std::min(r.exponential(0.8), 5.0);
    const double ct = std::max(1.0 - r.exponential(0.3), -1.0);
    c.vars.push_back(smp);
    c.vars.push_back(p);
    c.vars.push_back(ct);
    c.base.push_back(1.0);
    int64_t before = (int64_t)c.type.size();
    used.clear();
    if (!sp.mixed) {
      for (int k = 0; k < 4; ++k) {
        int par;
        do par = S.pool[P_MONO][r.integer((int)S.pool[P_MONO].size())];
        while (std::find(used.begin(), used.end(), par) != used.end());
        used.push_back(par);
        emit_dial(S, c, r, P_MONO, par, S.par_input[par]);
      }
      const int q = std::min(int(p / 1.25), 3);
      const int par = S.pool[P_NORM][q];
      emit_dial(S, c, r, P_NORM, par, S.par_input[par]);
    } else {
      static const int kinds[7] = {P_UNIFORM, P_GENERAL, P_COMPACT, P_GRAPH,
                                   P_NORM, P_TABULATED, P_KRIGED};
      const int nkinds = sp.features ? 8 : 7;
      for (int k = 0; k < sp.dials_per_event; ++k) {
        int kind, par;
        for (;;) {
          const int ki = r.integer(nkinds);
          kind = (ki == 7) ? P_MONO : kinds[ki];
          par = S.pool[kind][r.integer((int)S.pool[kind].size())];
          if (std::find(used.begin(), used.end(), par) == used.end()) break;
        }
        used.push_back(par);
        int in = S.par_input[par];
        if (sp.features && S.par_input_alt[par] >= 0 && r.uniform() < 0.5)
          in = S.par_input_alt[par];
        emit_dial(S, c, r, kind, par, in);
      }
      if (sp.features && r.uniform() < 0.2) {
        // a Shift dial (CacheManager.cpp:638-642 folds these into the base
        // weight; the CPU path multiplies them in dial-list order)
        c.type.push_back(GCM_DIAL_SHIFT);
        c.input.push_back(S.par_input[S.pool[P_NORM][0]]);
        c.input2.push_back(-1);
        c.data.push_back(0.5 + r.uniform());
        c.data_len.push_back(1);
      }
      if (sp.features && r.uniform() < 0.3) {
        // a 2-D dial on two spline parameters (Bilinear / Bicubic)
        const auto& pool = S.pool[P_MONO];
        const int a = pool[r.integer((int)pool.size())];
        int b = pool[r.integer((int)pool.size())];
        if (b == a) b = pool[(std::find(pool.begin(), pool.end(), a) - pool.begin() + 1) % pool.size()];
        emit_dial2d(c, r, S.par_input[a], S.par_input[b]);
      }
    }
    c.n_dials.push_back((int64_t)c.type.size() - before);
  }
}

void setup_parameters(Synth& S) {
  const Spec& sp = S.spec;
  S.par_type.assign(sp.n_par, P_MONO);
  S.pool.assign(P_NTYPES, {});
  if (!sp.mixed) {
    for (int i = 0; i < sp.n_par - sp.n_norm; ++i) S.par_type[i] = P_MONO;
    for (int i = sp.n_par - sp.n_norm; i < sp.n_par; ++i) S.par_type[i] = P_NORM;
  } else {
    static const int shape[7] = {P_UNIFORM, P_GENERAL, P_COMPACT, P_GRAPH,
                                 P_TABULATED, P_KRIGED, P_MONO};
    const int nshape = sp.features ? 7 : 6;
    for (int i = 0; i < sp.n_norm; ++i) S.par_type[i] = P_NORM;
    for (int i = sp.n_norm; i < sp.n_par; ++i)
      S.par_type[i] = shape[(i - sp.n_norm) % nshape];
  }
  S.prior_mean.resize(sp.n_par);
  S.prior_sigma.resize(sp.n_par);
  S.par_table.assign(sp.n_par, -1);
  S.table_begin.assign(1, 0);
  for (int i = 0; i < sp.n_par; ++i) {
    const int t = S.par_type[i];
    S.pool[t].push_back(i);
    S.prior_mean[i] = (t == P_NORM) ? 1.0 : 0.0;
    S.prior_sigma[i] = (t == P_NORM) ? 0.1 : 1.0;
    if (t == P_TABULATED || t == P_KRIGED) {
      S.par_table[i] = (int)S.table_par.size();
      S.table_par.push_back(i);
      S.table_kind.push_back(t == P_TABULATED ? 0 : 1);
      S.table_begin.push_back(S.table_begin.back() +
                              (t == P_TABULATED ? 64 : 256));
    }
  }
  // One dial collection per parameter: minDialResponse defaults to 0 in code
  // (DialCollection.cpp:453), maxDialResponse unset, no mirroring.
  const double nan = std::nan("");
  S.par_input.resize(sp.n_par);
  S.par_input_alt.assign(sp.n_par, -1);
  for (int i = 0; i < sp.n_par; ++i) {
    gcm_input in{i, 0, nan, nan, 0.0, nan};
    if (sp.features) {
      if (i % 5 == 1) { in.mirror_lower = -1.5; in.mirror_range = 3.0; }
      if (i % 7 == 2) in.max_response = 1.05;
      if (i % 6 == 3) in.flags |= GCM_INPUT_NO_EXTRAPOLATION;
    }
    S.par_input[i] = (int)S.inputs.size();
    S.inputs.push_back(in);
  }
  if (sp.features) {
    // a second collection on every fourth parameter, with other clamps
    for (int i = 0; i < sp.n_par; i += 4) {
      gcm_input in{i, 0, nan, nan, 0.97, 1.02};
      S.par_input_alt[i] = (int)S.inputs.size();
      S.inputs.push_back(in);
    }
    S.cap = 1.6;
  }
}

void setup_bins(Synth& S) {
  const Spec& sp = S.spec;
  const int ns = (int)sp.sample_events.size();
  S.n_bins = ns * sp.p_bins * sp.c_bins;
  S.edges.resize((size_t)S.n_bins * 6);
  size_t k = 0;
  for (int s = 0; s < ns; ++s)
    for (int ip = 0; ip < sp.p_bins; ++ip)
      for (int ic = 0; ic < sp.c_bins; ++ic) {
        S.edges[k++] = s;  // sample: equality "condition variable"
        S.edges[k++] = s;
        S.edges[k++] = ip * 5.0 / sp.p_bins;
        S.edges[k++] = (ip + 1) * 5.0 / sp.p_bins;
        S.edges[k++] = (ic - sp.c_bins / 2) / double(sp.c_bins / 2);
        S.edges[k++] = (ic + 1 - sp.c_bins / 2) / double(sp.c_bins / 2);
      }
}

Spec make_spec(int id, int64_t n_events, uint64_t seed) {
  Spec sp;
  sp.id = id;
  sp.seed = seed ? seed : 20261203ull + (uint64_t)id;
  switch (id) {
    case 1:
      sp.sample_events = {100000};
      sp.p_bins = 20; sp.c_bins = 20;
      sp.n_par = 20; sp.n_norm = 4; sp.mixed = false;
      break;
    case 3:
      sp.sample_events = {50000, 30000, 15000, 5000};
      sp.p_bins = 20; sp.c_bins = 20;
      sp.n_par = 20; sp.n_norm = 4; sp.mixed = false;
      break;
    case 4:
      sp.sample_events = {8000000, 5000000, 3000000, 2000000, 1000000, 1000000};
      sp.p_bins = 60; sp.c_bins = 50;
      sp.n_par = 500; sp.n_norm = 50; sp.dials_per_event = 15;
      break;
    case 0:
      sp.sample_events = {3000, 2000};
      sp.p_bins = 10; sp.c_bins = 10;
      sp.n_par = 40; sp.n_norm = 5; sp.dials_per_event = 9;
      sp.features = true;
      break;
    default:
      sp.id = 2;
      sp.sample_events = {2000000, 1500000, 1000000, 500000};
      sp.p_bins = 30; sp.c_bins = 50;
      sp.n_par = 300; sp.n_norm = 30; sp.dials_per_event = 12;
      break;
  }
  if (n_events > 0) {
    int64_t tot = 0;
    for (auto v : sp.sample_events) tot += v;
    int64_t acc = 0;
    for (size_t s = 0; s < sp.sample_events.size(); ++s) {
      int64_t v = (s + 1 == sp.sample_events.size())
                      ? n_events - acc
                      : std::max<int64_t>(1, sp.sample_events[s] * n_events / tot);
      sp.sample_events[s] = v;
      acc += v;
    }
  }
  return sp;
}

thread_local std::string g_err;

}  // namespace

extern "C" {


const char* gcm_synth_last_error(void) { return g_err.c_str(); }

void* gcm_synth_create(int workload, int64_t n_events, uint64_t seed,
                       int n_threads) {
  try {
    auto S = std::make_unique<Synth>();
    S->spec = make_spec(workload, n_events, seed);
    for (auto v : S->spec.sample_events) S->n_events += v;
    setup_parameters(*S);
    setup_bins(*S);
    const int64_t N = S->n_events;
    const int64_t CH = 4096;
    const int64_t nch = (N + CH - 1) / CH;
    std::vector<Chunk> chunks(nch);
    n_threads = std::max(1, std::min<int>(n_threads, 64));
    std::vector<std::thread> pool;
    for (int t = 0; t < n_threads; ++t)
      pool.emplace_back([&, t] {
        for (int64_t c = t; c < nch; c += n_threads)
          make_chunk(*S, c * CH, std::min(N, (c + 1) * CH), chunks[c]);
      });
    for (auto& th : pool) th.join();
    int64_t nd = 0, ndata = 0;
    for (auto& c : chunks) { nd += c.type.size(); ndata += c.data.size(); }
    S->vars.reserve(N * 3);
    S->base.reserve(N);
    S->dial_begin.reserve(N + 1);
    S->type.reserve(nd);
    S->input.reserve(nd);
    S->input2.reserve(nd);
    S->data_begin.reserve(nd + 1);
    S->data.reserve(ndata);
    S->dial_begin.push_back(0);
    S->data_begin.push_back(0);
    for (auto& c : chunks) {
      S->vars.insert(S->vars.end(), c.vars.begin(), c.vars.end());
      S->base.insert(S->base.end(), c.base.begin(), c.base.end());
      for (auto n : c.n_dials) S->dial_begin.push_back(S->dial_begin.back() + n);
      S->type.insert(S->type.end(), c.type.begin(), c.type.end());
      S->input.insert(S->input.end(), c.input.begin(), c.input.end());
      S->input2.insert(S->input2.end(), c.input2.begin(), c.input2.end());
      for (auto n : c.data_len) S->data_begin.push_back(S->data_begin.back() + n);
      S->data.insert(S->data.end(), c.data.begin(), c.data.end());
      c = Chunk();
    }
    S->bin.assign(N, -1);
    return S.release();
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return nullptr;
  }
}

void gcm_synth_destroy(void* h) { delete static_cast<Synth*>(h); }

int gcm_synth_view_get(void* h, gcm_synth_view* v) {
  Synth* S = static_cast<Synth*>(h);
  v->n_events = S->n_events;
  v->n_vars = 3;
  v->n_bins = S->n_bins;
  v->n_parameters = S->spec.n_par;
  v->n_inputs = (int32_t)S->inputs.size();
  v->n_tables = (int32_t)S->table_par.size();
  v->n_samples = (int32_t)S->spec.sample_events.size();
  v->n_dials = (int64_t)S->type.size();
  v->n_data = (int64_t)S->data.size();
  v->table_len = S->table_begin.back();
  v->reweight_cap = S->cap;
  v->vars = S->vars.data();
  v->edges = S->edges.data();
  v->base_weight = S->base.data();
  v->bin = S->bin.data();
  v->dial_begin = S->dial_begin.data();
  v->dial_type = S->type.data();
  v->dial_input = S->input.data();
  v->data_begin = S->data_begin.data();
  v->dial_data = S->data.data();
  v->inputs = S->inputs.data();
  v->table_begin = S->table_begin.data();
  v->prior_mean = S->prior_mean.data();
  v->prior_sigma = S->prior_sigma.data();
  v->par_type = S->par_type.data();
  v->dial_input2 = S->input2.data();
  return 0;
}

// The Tabulated/Kriged factory callback (TabulatedDialFactory.cpp:133-141
// analogue): T[j] = max(1 + 0.05 theta cos(pi j / (len-1)), 0) for the table
// of each tabulated (64 entries) or kriged (256 entries) parameter.
int gcm_synth_table_values(void* h, const double* theta, double* out) {
  Synth* S = static_cast<Synth*>(h);
  for (size_t t = 0; t < S->table_par.size(); ++t) {
    const double th = theta[S->table_par[t]];
    const int64_t b = S->table_begin[t], n = S->table_begin[t + 1] - b;
    for (int64_t j = 0; j < n; ++j)
      out[b + j] = std::max(1.0 + 0.05 * th * std::cos(M_PI * j / double(n - 1)), 0.0);
  }
  return 0;
}

}  // extern "C"
