/*
 * gcm_table_update.h -- writing a device-side lookup-table update function
 * for gcm_add_table_update (include/gcm.h).
 *
 * The reference refreshes Tabulated/Kriged tables on the host with a user
 * function `update(name, table, size, inputs, n_inputs)` loaded from a
 * shared library (TabulatedDialFactory.cpp:133-141).  Here the same function
 * body runs on the GPU: GCM_TABLE_UPDATE(name) { ... } defines the kernel
 * `name`, and the body sees
 *     double* table          the table of one point, `size` entries
 *     int32_t size
 *     const double* inputs   the job's mirrored input values
 *     int32_t n_inputs
 *     int32_t user_id        the job's user_id
 * Every thread of the block runs the body, so it loops over the entries:
 *
 *     #include "gcm_table_update.h"
 *     GCM_TABLE_UPDATE(my_update) {
 *       for (int j = threadIdx.x; j < size; j += blockDim.x)
 *         table[j] = 1.0 + 0.1 * inputs[0] * j;
 *     }
 *
 * Build: hipcc --genco --offload-arch=gfx950 -O3 -I<repo>/include f.hip -o f.hsaco
 */
#ifndef GCM_TABLE_UPDATE_H
#define GCM_TABLE_UPDATE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

/* One job as the library hands it to the kernel (device memory). */
typedef struct gcm_table_job_dev {
  int64_t table_offset;  /* first entry of the table in a point's tables    */
  int64_t input_offset;  /* first input of the job in a point's input values */
  int32_t size;          /* entries in the table                            */
  int32_t n_inputs;
  int32_t user_id;
  int32_t reserved;
} gcm_table_job_dev;

/* Kernel arguments: jobs[gridDim.x], tables [points][table_stride],
 * inputs [points][input_stride]; block (job, point) = (blockIdx.x,
 * blockIdx.y). */
#define GCM_TABLE_UPDATE(name)                                                      \
  static __device__ void name##_body(double* table, int32_t size,                   \
                                     const double* inputs, int32_t n_inputs,        \
                                     int32_t user_id);                              \
  extern "C" __global__ void name(const gcm_table_job_dev* jobs, double* tables,     \
                                  int64_t table_stride, const double* inputs,       \
                                  int64_t input_stride) {                           \
    const gcm_table_job_dev jb = jobs[blockIdx.x];                                  \
    name##_body(tables + (int64_t)blockIdx.y * table_stride + jb.table_offset,      \
                jb.size, inputs + (int64_t)blockIdx.y * input_stride + jb.input_offset, \
                jb.n_inputs, jb.user_id);                                           \
  }                                                                                 \
  static __device__ void name##_body(double* table, int32_t size,                   \
                                     const double* inputs, int32_t n_inputs,        \
                                     int32_t user_id)

#endif /* GCM_TABLE_UPDATE_H */
