// The synthetic workloads' Tabulated/Kriged table function as a device table
// update (gcm_add_table_update): T[j] = max(1 + 0.05 x cos(pi j / (size - 1)), 0)
// with x the job's first (mirrored) input.  It is the device form of the host
// callback gcm_synth_table_values (synth.cpp), evaluated in the same operation
// order; cos is the device library's, so entries may differ from the host's
// in the last bit.  Built as a code object: gundam_amd/table_update_cos.hsaco.
#include <math.h>

#include "../../include/gcm_table_update.h"

GCM_TABLE_UPDATE(gcm_synth_cos_table) {
  const double x = inputs[0];
  for (int j = threadIdx.x; j < size; j += blockDim.x)
    table[j] = fmax(1.0 + 0.05 * x * cos(M_PI * j / double(size - 1)), 0.0);
}
