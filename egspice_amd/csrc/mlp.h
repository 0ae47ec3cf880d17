// Internal seam between the lane engine and the warm-start network
// (include/espice_mlp.h): the engine owns the lanes' start point and
// stream, the network writes its prediction there.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/espice_mlp.h"

namespace esp {

// Records `name` as the calling thread's last error (espice_lanes_error_message).
void set_error(const char *name);

// The network's prediction for params[count][n_in] (host) written into
// x[i * N + l], i < n, l < count (device, lane-major), on `stream`.
espice_status mlp_warm_start(espice_mlp *m, const double *params, uint32_t count, double *x, uint32_t n, uint32_t N,
                             hipStream_t stream);

}  // namespace esp
