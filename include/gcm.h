/*
 * gcm.h -- C ABI of the MI355X event-reweighting cache ("gcm" = GUNDAM cache
 * manager).  This is the drop-in boundary that replaces the reference's
 * Cache::Manager (src/CacheManager/include/CacheManager.h:56-410): parameter
 * upload, dial-table build, weight and histogram fetch, plus the binned
 * likelihood that the reference evaluates on the host after the copy-back
 * (src/StatisticalInference/JointProbability/include/JointProbabilityBase.h:46-54).
 *
 * Conventions
 *   - Every function returns an int status (GCM_OK == 0); on failure the
 *     message is available from gcm_last_error() (thread-local).  Nothing is
 *     thrown across the ABI and nothing calls exit().
 *   - Plain pointers and sizes only.  Host pointers unless a name ends in
 *     _device.  All floating point is IEEE binary64 at the boundary.
 *   - One context owns one set of device-resident tables on one GPU.  A
 *     context (and a gcm_mlp handle) is driven by one host thread at a time;
 *     different contexts share no mutable state and may be used from
 *     different threads at once.
 *
 * Dial payload layouts (one record per dial, in gcm_tables.dial_data).  They
 * are the reference's own record layouts so the facade can pass
 * DialBase::getDialData() through unchanged:
 *   NORM       : (empty)                      Norm.h:21 -- response = input value
 *   MONOTONIC  : x0, step, y[0..n-1]          MonotonicSpline.cpp:12-69
 *   UNIFORM    : x0, step, (y,slope)[0..n-1]  UniformSpline.cpp:14-45
 *   GENERAL    : x0, step, (y,slope,x)[0..n-1] GeneralSpline.cpp:12-31
 *   COMPACT    : x0, step, y[0..n-1]          CompactSpline.cpp:11-69
 *   GRAPH      : (y,x)[0..n-1]                Graph.cpp:32-39
 *   TABULATED  : table, index, fraction       Tabulated.h:16-28 (fraction is
 *                rounded to float as the reference does)
 *   KRIGED     : table, (index, weight)[0..k-1]  Kriged.cpp:3-21 (weights are
 *                rounded to float as the reference does)
 *   SHIFT      : value                        Shift.h:21
 *   BILINEAR,  : nx, ny, x[0..nx-1], y[0..ny-1], v[0..nx*ny-1] (x-major)
 *   BICUBIC      Bilinear.cpp:5-32, Bicubic.cpp:5-32; nx, ny >= 2.  The
 *                dial reads dial_input (x) and dial_input2 (y); the response
 *                supervisor is dial_input's; both coordinates are clamped to
 *                the knot range when dial_input has GCM_INPUT_NO_EXTRAPOLATION
 *                (the reference's default for these dials).
 * "table" is an index into gcm_tables.table_begin; "index" is relative to the
 * start of that table.
 */
#ifndef GCM_H
#define GCM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define GCM_OK 0
#define GCM_ERR_ARG 1         /* invalid argument or size                   */
#define GCM_ERR_STATE 2       /* call order violated (e.g. fill before build) */
#define GCM_ERR_HIP 3         /* HIP runtime error                          */
#define GCM_ERR_UNSUPPORTED 4 /* dial type or feature not supported          */
#define GCM_ERR_NUMERIC 5     /* NaN weight or non-finite likelihood         */
#define GCM_ERR_NOMEM 6       /* device or host allocation failed            */

/* ---- dial types (DialBase subclasses with device support,
 *      CacheManager.cpp:570-667) ------------------------------------------ */
#define GCM_DIAL_NORM 0
#define GCM_DIAL_MONOTONIC 1
#define GCM_DIAL_UNIFORM 2
#define GCM_DIAL_GENERAL 3
#define GCM_DIAL_COMPACT 4
#define GCM_DIAL_GRAPH 5
#define GCM_DIAL_TABULATED 6
#define GCM_DIAL_KRIGED 7
#define GCM_DIAL_SHIFT 8
#define GCM_DIAL_BILINEAR 9  /* 2-D: Bilinear.cpp, CalculateBilinearInterpolation.h */
#define GCM_DIAL_BICUBIC 10  /* 2-D: Bicubic.cpp, CalculateBicubicSpline.h        */
#define GCM_DIAL_NTYPES 11

/* ---- input flags --------------------------------------------------------- */
#define GCM_INPUT_NO_EXTRAPOLATION 1 /* DialBase::setAllowExtrapolation(false) */

/* ---- precision modes ----------------------------------------------------- */
#define GCM_PRECISION_FP64 0 /* FP64 records and dial evaluation (reference)  */
#define GCM_PRECISION_FP32 1 /* FP32 records and dial evaluation; FP64 clamp,
                                compensated FP64 bin sums and FP64 LLH       */

/* ---- likelihood types (JointProbability.cpp:61-94) ---------------------- */
#define GCM_LLH_POISSON 0           /* PoissonLogLikelihood.h:19-53          */
#define GCM_LLH_BARLOW_OA2021 1     /* BarlowBeestonBanff2022.h:92-294      */
#define GCM_LLH_BARLOW_BEESTON 2    /* classic, BarlowBeeston.h:25-62       */
#define GCM_LLH_BARLOW_OA2020 3     /* BarlowBeestonBanff2020.h:18-85       */
#define GCM_LLH_BARLOW_OA2021_SFGD 4 /* BarlowBeestonBanff2022Sfgd.h:20-134 */
#define GCM_LLH_CHI_SQUARED 5       /* ChiSquared.h:19-29                    */
#define GCM_LLH_LEAST_SQUARES 6     /* LeastSquares.h:45-52                  */
#define GCM_LLH_NTYPES 7

/* One dial input: the parameter reference of a DialInputBuffer with its
 * mirror edges (DialInputBuffer.cpp:44-89), and the DialResponseSupervisor of
 * the dial collection (DialResponseSupervisor.cpp:10-18).  Dials of different
 * collections that read the same parameter use different inputs, which keeps
 * the CPU path's per-collection clamp semantics. */
typedef struct gcm_input {
  int32_t parameter;    /* index into the parameter vector                 */
  int32_t flags;        /* GCM_INPUT_*                                      */
  double mirror_lower;  /* NaN: no mirroring                                */
  double mirror_range;  /* mirror range (upper - lower)                     */
  double min_response;  /* NaN: no lower clamp (the reference's collections
                           default to 0, DialCollection.cpp:453: pass 0)   */
  double max_response;  /* NaN: no upper clamp                              */
} gcm_input;

/* Dial tables in the reference's event-major EventDialCache order
 * (EventDialCache.h:64-67): event e owns dials [dial_begin[e], dial_begin[e+1])
 * in dial-list order, which is the order of the response product
 * (EventDialCache.cpp:156-171). */
typedef struct gcm_tables {
  int64_t n_events;
  const double* base_weight;  /* [n_events] Event::Weights::base          */
  const int32_t* bin;         /* [n_events] global bin (sample offset
                                 included), -1 = unbinned (dropped,
                                 CacheManager.cpp:510)                     */
  const int64_t* dial_begin;  /* [n_events+1]                               */
  int64_t n_dials;
  const uint8_t* dial_type;   /* [n_dials] GCM_DIAL_*                       */
  const int32_t* dial_input;  /* [n_dials] index into inputs                */
  const int64_t* data_begin;  /* [n_dials+1] offsets into dial_data         */
  const double* dial_data;    /* payloads, layouts above                    */
  int32_t n_inputs;
  const gcm_input* inputs;    /* [n_inputs]                                 */
  int32_t n_tables;
  const int64_t* table_begin; /* [n_tables+1] offsets into the table vector */
  int32_t n_parameters;
  int32_t n_bins;
  double reweight_cap;        /* GlobalEventReweightCap (EventDialCache.h:26-35);
                                 NaN or +inf: disabled                     */
  const int32_t* dial_input2; /* [n_dials] second input of 2-D dials (the
                                 y axis; DialInputBuffer slot 1), ignored
                                 for 1-D dials; may be NULL without 2-D
                                 dials                                      */
} gcm_tables;

/* Likelihood configuration: the options of the reference's joint
 * probabilities that change arithmetic.  The per-bin arrays are [n_bins] in
 * global bin order, copied by gcm_set_likelihood; NULL means "not used".
 * Every type receives, per bin, (data, sum w, sqrt(sum w^2)) as
 * JointProbabilityBase::eval(SamplePair, bin) passes them
 * (JointProbabilityBase.h:37-43); the OA2021 types square the uncertainty
 * themselves (BarlowBeestonBanff2022.h:125-126), the classic and OA2020 ones
 * use it as given. */
typedef struct gcm_llh_config {
  int32_t type;                   /* GCM_LLH_*                              */
  int32_t use_poisson_likelihood; /* OA2021: drop the beta penalty          */
  double fractional_error_limit;  /* OA2021, default 1e150 (capped 1e152)   */
  double expected_value_minimum;  /* OA2021, default -1                     */
  int32_t lsq_poissonian_approximation; /* LeastSquares.h:28, 50          */
  int32_t bb_no_update_weights;   /* OA2021 BBNoUpdateWeights (:36, 109-123):
                                     uncertainty from nominal_err          */
  const double* nominal_err;      /* sqrt(sum w^2) of the nominal MC        */
  const double* extra_fraction;   /* OA2021_SFGD: the sample's detector
                                     fraction, added as pred x fraction to
                                     sqrt(sum w^2) (Sfgd.h:28-69; give the one
                                     nonzero SFGD or WAGASCI value)         */
  const uint8_t* ignore_bins;     /* nonzero: the bin contributes 0
                                     (ignoreBinsWithZeroPredictionAtPrior,
                                     JointProbabilityBase.h:56-77)          */
} gcm_llh_config;

/* Kernel launch shape.  Zero fields mean "use the built-in default".  The
 * tile record capacity changes the device layout, so it is fixed at build. */
typedef struct gcm_launch_config {
  int32_t block_threads;    /* threads per block of the fill kernel         */
  int32_t tile_words;       /* record words (8 B) staged in LDS per tile    */
  int32_t blocks_per_cu;    /* resident blocks per CU for the grid size     */
  int32_t ring_buffers;     /* LDS tile buffers (2 or 3); producers prefetch
                               ring_buffers - 1 tiles ahead                 */
} gcm_launch_config;

typedef struct gcm_context gcm_context;

/* ---- library ------------------------------------------------------------ */
/* Last error message of the calling thread ("" if none). */
const char* gcm_last_error(void);
/* ABI version (major*10000 + minor*100 + patch). */
int gcm_version(void);
/* Replaces Cache::Manager::HasGPU (CacheManager.h:73-78): number of visible
 * devices. */
int gcm_device_count(int* count);

/* ---- lifetime (Cache::Manager ctor/Build, CacheManager.cpp:46-188,198-475) */
int gcm_create(int device, gcm_context** out);
int gcm_destroy(gcm_context* ctx);
/* Precision and launch shape; call before gcm_build. */
int gcm_set_precision(gcm_context* ctx, int precision);
int gcm_set_launch_config(gcm_context* ctx, const gcm_launch_config* cfg);
int gcm_get_launch_config(gcm_context* ctx, gcm_launch_config* cfg);
/* Replaces Cache::Manager::Build + Update (CacheManager.cpp:198-475,477-740):
 * validates the tables, lays them out for the device and uploads them.  The
 * tables are copied; the caller's arrays may be freed afterwards.
 * max_points is the largest K that gcm_set_parameters will be given. */
int gcm_build(gcm_context* ctx, const gcm_tables* tables, int32_t max_points);
/* Device-resident bytes (Cache::Manager::GetResidentMemory,
 * CacheManager.h:192). */
int gcm_resident_bytes(gcm_context* ctx, int64_t* bytes);
/* Events kept after dropping unbinned ones; dials kept. */
int gcm_counts(gcm_context* ctx, int64_t* events, int64_t* dials, int64_t* tiles);

/* ---- per-evaluation inputs ----------------------------------------------- */
/* Replaces Cache::Parameters::SetParameter (CacheParameters.cpp:141-154) for K
 * parameter points at once: theta is [k_points][n_parameters], row-major. */
int gcm_set_parameters(gcm_context* ctx, const double* theta,
                       int32_t n_parameters, int32_t k_points);
/* Tabulated/Kriged lookup tables for every point: [k_points][table_len]
 * (the per-Fill host upload of WeightTabulated.cpp:164-171 and
 * WeightKriged.cpp:216-223). */
int gcm_set_table_values(gcm_context* ctx, const double* values,
                         int64_t table_len, int32_t k_points);
/* Likelihood: data histogram [n_bins], Gaussian priors [n_parameters]
 * (sigma <= 0 or NaN: no penalty term; LikelihoodInterface.cpp:179-215 with a
 * diagonal covariance). */
int gcm_set_likelihood(gcm_context* ctx, const gcm_llh_config* cfg,
                       const double* data_hist, const double* prior_mean,
                       const double* prior_sigma);
/* Correlated prior penalty of one parameter set (LikelihoodInterface.cpp:
 * 186-213, ParameterSet::updateDeltaVector ParameterSet.cpp:456-463): with
 * d_i = theta[parameters[i]] - prior[i], the set adds
 *   sum_i d_i * (sum_j inv_cov[i][j] * d_j)
 * summed from 0 in index order (TMatrixD x TVectorD, then the dot product),
 * or +inf when lower/upper are given and some theta[parameters[i]] is NaN or
 * outside [lower[i], upper[i]] (NaN bound: none; the reference's
 * applyInfinitePenaltyOnOutOfBoundPar, ParameterSet.cpp:441-453).  inv_cov
 * NULL: bounds only.  The caller lists the set's valid correlated parameters
 * (ParameterSet.cpp:225-233) and passes the inverse of their covariance. */
typedef struct gcm_penalty_set {
  int32_t n;
  const int32_t* parameters;  /* [n] indices into the parameter vector      */
  const double* prior;        /* [n]                                        */
  const double* inv_cov;      /* [n][n] row-major, or NULL                  */
  const double* lower;        /* [n] or NULL                                */
  const double* upper;        /* [n] or NULL                                */
} gcm_penalty_set;
/* Replaces the penalty sets (n_sets = 0 clears them).  The penalty of a point
 * is (the diagonal prior term of gcm_set_likelihood) + set 0 + set 1 + ...,
 * in that order; with no diagonal term this is the reference's sum over
 * parameter sets.  The arrays are copied. */
int gcm_set_penalty_sets(gcm_context* ctx, int32_t n_sets,
                         const gcm_penalty_set* sets);

/* Device-side Tabulated/Kriged table refresh, in place of the per-Fill host
 * callback and upload (TabulatedDialFactory.cpp:133-141 calls a user function
 * found with dlsym; WeightTabulated.cpp:164-171 and WeightKriged.cpp:216-223
 * copy the result to the device).  The user function is a gfx950 kernel in a
 * code object (hipcc --genco), written with include/gcm_table_update.h; each
 * gcm_fill launches it once, one block per (job, point), after the input
 * mirroring and before the dial evaluation, so graph replays refresh too.
 * A job fills table `table` of every point from the mirrored values of its
 * inputs (gcm_input indices: DialInputBuffer semantics), in the order given.
 * Tables that no job covers keep the values of gcm_set_table_values (or 0).
 * A table may be covered by at most one job over all calls. */
typedef struct gcm_table_job {
  int32_t table;          /* index into gcm_tables.table_begin              */
  int32_t n_inputs;
  const int32_t* inputs;  /* [n_inputs] indices into gcm_tables.inputs      */
  int32_t user_id;        /* passed to the function (the reference passes
                             the factory name)                             */
} gcm_table_job;
/* Load `kernel_name` from the code object file and add its jobs. */
int gcm_add_table_update(gcm_context* ctx, const char* code_object_path,
                         const char* kernel_name, int32_t n_jobs,
                         const gcm_table_job* jobs);
/* Remove every device table update (and unload the code objects). */
int gcm_clear_table_updates(gcm_context* ctx);
/* The lookup tables point k was last evaluated with: [table_len]. */
int gcm_get_table_values(gcm_context* ctx, int32_t k, double* values);

/* Keep per-event weights on the device so gcm_get_weights can fetch them
 * (Cache::Manager::SetIsEventWeightCopyEnabled, CacheManager.h:104-109). */
int gcm_enable_weights(gcm_context* ctx, int enable);

/* ---- evaluation (Cache::Manager::Fill, CacheManager.cpp:750-810) --------- */
/* Enqueue mirror/clamp, dial evaluation, weight product, histogram and
 * likelihood for the K points set last, on `stream` (a hipStream_t, NULL for
 * the context's own stream).  Asynchronous. */
int gcm_fill(gcm_context* ctx, void* stream);
/* Wait for the last fill; reports NaN weights (CacheWeights.cpp:83-86) as
 * GCM_ERR_NUMERIC. */
int gcm_sync(gcm_context* ctx);
/* hipGraph mode (no reference counterpart; the reference launches every
 * kernel eagerly, CacheWeights.cpp:153-183): the first gcm_fill for a point
 * count K captures the uploads and the kernels into a graph, later fills
 * replay it.  Timing (gcm_enable_timing) forces eager launches. */
int gcm_enable_graph(gcm_context* ctx, int enable);

/* ---- fetch (CopyHistogramContents / CopyEventWeights,
 *      CacheManager.cpp:832-866) ------------------------------------------- */
/* Sum of weights and sum of squared weights per bin for point k. */
int gcm_get_hist(gcm_context* ctx, int32_t k, double* sums, double* sums2);
/* Replaces Cache::Manager::ValidateHistogramContents (CacheManager.cpp:868-887,
 * CacheSampleHistFiller.cpp:21-50): compares point k's bin sums and
 * sqrt(sum w^2) with the caller's (CPU-filled) ones, [n_bins] each, with the
 * reference's GundamUtils::almostEqual rule (GundamAlmostEqual.h:24-58: a
 * relative tolerance eps^(1 - fraction_of_bits_lost), eps the FP64 epsilon;
 * non-finite values never match; the reference's default fraction is 0.3).
 * *n_bad receives the number of mismatching bins; verbose != 0 prints each
 * to stderr. */
int gcm_validate_hist(gcm_context* ctx, int32_t k, const double* sums,
                      const double* sqrt_sum_w2, double fraction_of_bits_lost,
                      int verbose, int32_t* n_bad);
/* Per-event weights of point 0 in the caller's event order; unbinned events
 * are left untouched. */
int gcm_get_weights(gcm_context* ctx, double* weights);
/* Total likelihood (stat + penalty) per point, and the two parts; any pointer
 * may be NULL. */
int gcm_get_llh(gcm_context* ctx, double* total, double* stat,
                double* penalty);

/* ---- kernel timing ------------------------------------------------------ */
/* Record HIP events around the fill and finalize kernels of every gcm_fill on
 * the stream they run on (replaces the AveragedTimer cacheFillTimer,
 * CacheManager.h:128-135).  Enabling resets the totals. */
int gcm_enable_timing(gcm_context* ctx, int enable);
/* Totals over the completed fills since timing was enabled. */
int gcm_get_timing(gcm_context* ctx, double* fill_ms, double* finalize_ms,
                   int64_t* count);

/* ---- event -> bin indexing (LoaderUtils.cpp:32-40, Bin.cpp:167-176) ------ */
/* First bin whose [min, max) box holds the event on every variable (min ==
 * max means equality); -1 if none.  values is [n_events][n_vars]; edges is
 * [n_bins][n_vars][2].  Runs on `device`. */
int gcm_bin_events(int device, int64_t n_events, int32_t n_vars,
                   const double* values, int32_t n_bins, const double* edges,
                   int32_t* out_bin);

/* ---- learned surrogate (no reference counterpart: an MLP that predicts the bin contents) --
 * An MLP n_in -> n_hidden -> n_hidden -> n_out with GELU, BF16 weights and
 * FP32 accumulation on MFMA, trained with Adam on the mean squared error.
 * Parameter layout (FP32 master copy): W1[h][n_in] b1[h] W2[h][h] b2[h]
 * W3[n_out][h] b3[n_out].  Inputs and targets are normalised by the caller. */
typedef struct gcm_mlp gcm_mlp;
int gcm_mlp_create(int device, int n_in, int n_hidden, int n_out, uint64_t seed,
                   gcm_mlp** out);
int gcm_mlp_destroy(gcm_mlp* m);
int gcm_mlp_param_count(gcm_mlp* m, int64_t* count);
int gcm_mlp_get_params(gcm_mlp* m, float* out);
int gcm_mlp_set_params(gcm_mlp* m, const float* in);
/* y[rows][n_out] = MLP(x[rows][n_in]) (host arrays). */
int gcm_mlp_predict(gcm_mlp* m, const float* x, int64_t rows, float* y);
/* Mean squared error and its gradient w.r.t. every parameter, no update. */
int gcm_mlp_gradients(gcm_mlp* m, const float* x, const float* t, int rows,
                      float* grad, double* loss);
/* Resident training set and Adam epochs over it (shuffled minibatches);
 * loss_out[epoch] (optional) is the mean of sampled minibatch losses. */
int gcm_mlp_set_data(gcm_mlp* m, const float* x, const float* t, int64_t rows);
int gcm_mlp_train(gcm_mlp* m, int epochs, int batch, float lr, uint64_t seed,
                  double* loss_out);
/* Likelihood head: for output b, pred = exp(mean[b] + scale[b] * y[b])
 * (mean[b] = -inf: no events, pred = 0) with MC variance pred * var_ratio[b];
 * gcm_mlp_predict_llh returns each row's Barlow-Beeston OA2021 chi2 against
 * data[] (the same per-bin formula as GCM_LLH_BARLOW_OA2021), summed over the
 * n_out bins in FP64 on the device. */
int gcm_mlp_set_head(gcm_mlp* m, const double* mean, const double* scale,
                     const double* data, const double* var_ratio);
int gcm_mlp_predict_llh(gcm_mlp* m, const float* x, int64_t rows, double* chi2);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph stream capture, tight serving and MCMC loops):
 * gcm_mlp_predict_llh on device arrays -- x_device [rows][n_in] floats,
 * chi2_device [rows] doubles -- enqueued on `stream` (a hipStream_t; NULL:
 * the handle's stream) without synchronising, and without allocating: every
 * intermediate (activations, split-K partials) lives in one fixed 16 MiB
 * device array per GPU, held until the process ends.  The rows must fit in
 * it: 4 x (4 n_hidden + n_out) bytes per row (each part rounded up to
 * 256 B) plus the split-K partials of long reductions; larger calls return
 * GCM_ERR_ARG (about 500 rows for 300 -> 512 -> 512 -> 6000).  Results equal
 * gcm_mlp_predict_llh's for the same rows.  All calls on one GPU share the
 * array: calls that overlap in time on one GPU (other streams, other
 * handles, other threads) race; serialise them. */
int gcm_mlp_predict_llh_noalloc(gcm_mlp* m, const float* x_device, int64_t rows,
                                double* chi2_device, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GCM_H */
