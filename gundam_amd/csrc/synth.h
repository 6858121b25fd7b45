// synth.h -- C interface of the seeded synthetic workload generator
// (synth.cpp): the event and dial tables of the benchmark and test workloads
// in the layout gcm_build takes.
#ifndef GCM_SYNTH_H
#define GCM_SYNTH_H

#include <stdint.h>

#include "../../include/gcm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gcm_synth_view {
  int64_t n_events;
  int32_t n_vars;
  int32_t n_bins;
  int32_t n_parameters;
  int32_t n_inputs;
  int32_t n_tables;
  int32_t n_samples;
  int64_t n_dials;
  int64_t n_data;
  int64_t table_len;
  double reweight_cap;
  const double* vars;         /* [n_events][n_vars]                       */
  const double* edges;        /* [n_bins][n_vars][2]                      */
  const double* base_weight;
  int32_t* bin;               /* filled by the caller (binning)           */
  const int64_t* dial_begin;
  const uint8_t* dial_type;
  const int32_t* dial_input;
  const int64_t* data_begin;
  const double* dial_data;
  const gcm_input* inputs;
  const int64_t* table_begin;
  const double* prior_mean;
  const double* prior_sigma;
  const int32_t* par_type;    /* generator's parameter kind               */
  const int32_t* dial_input2; /* [n_dials] second input of 2-D dials, -1   */
} gcm_synth_view;

const char* gcm_synth_last_error(void);
void* gcm_synth_create(int workload, int64_t n_events, uint64_t seed, int n_threads);
void gcm_synth_destroy(void* h);
int gcm_synth_view_get(void* h, gcm_synth_view* v);
int gcm_synth_table_values(void* h, const double* theta, double* out);

#ifdef __cplusplus
}
#endif

#endif /* GCM_SYNTH_H */
