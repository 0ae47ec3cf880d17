/* The learned warm start of the lane engine: a small MLP that maps each
 * lane's sweep parameters to a Newton starting point, trained on the GPU
 * on operating points the engine itself solved, with BF16 MFMA kernels
 * (FP32 accumulation, FP32 master weights, Adam).
 *
 * The reference has no counterpart: its lanes start cold (MODEINITJCT) or
 * from the nominal operating point (src/analysis/dc/op.zig:588,671). The
 * network only moves the start; Newton convergence under the unchanged
 * tolerances decides every result, and a lane that does not converge from
 * its predicted start is rerun from the nominal start by the caller.
 *
 * Shapes: n_in inputs -> `depth` hidden layers of `hidden` tanh units ->
 * n_out linear outputs. Parameters are packed layer by layer as W (n_out
 * rows of n_in, row-major, out x in) then b. Inputs are standardized with
 * the caller's (mean, std) before the first layer, outputs are
 * de-standardized after the last; a zero std marks a constant column
 * (input 0, output exactly its mean). */
#ifndef ESPICE_MLP_H
#define ESPICE_MLP_H

#include <stddef.h>
#include <stdint.h>

#include "espice_lanes.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct espice_mlp espice_mlp;

typedef struct {
    uint32_t n_in, hidden, depth, n_out;
    uint64_t seed; /* Xavier-uniform initial weights from splitmix64(seed) */
} espice_mlp_shape;

typedef struct {
    uint32_t epochs, batch;
    float lr, beta1, beta2, eps;
    uint64_t seed; /* the per-epoch shuffles */
} espice_mlp_train_opts;

espice_status espice_mlp_create(const espice_mlp_shape *shape, espice_mlp **out);
espice_status espice_mlp_destroy(espice_mlp *m);
/* Number of packed parameters. */
espice_status espice_mlp_param_count(const espice_mlp *m, size_t *count);
/* in_mean/in_std[n_in], out_mean/out_std[n_out]. */
espice_status espice_mlp_set_scaling(espice_mlp *m, const double *in_mean, const double *in_std, const double *out_mean,
    const double *out_std);
/* FP32 master weights in packed order (param_count floats). */
espice_status espice_mlp_get_params(espice_mlp *m, float *params);
espice_status espice_mlp_set_params(espice_mlp *m, const float *params);
/* Adam on the mean squared error of the standardized outputs over
 * x[count][n_in] -> y[count][n_out]; loss[epochs] receives each epoch's
 * mean batch loss (may be NULL). */
espice_status espice_mlp_train(espice_mlp *m, const double *x, const double *y, uint32_t count,
    const espice_mlp_train_opts *opts, float *loss);
/* One forward and backward pass over the whole of x/y (count <= 4096) with
 * no update: grads[param_count] in packed order, *loss the batch loss. */
espice_status espice_mlp_gradients(espice_mlp *m, const double *x, const double *y, uint32_t count, float *grads,
    float *loss);
/* y[count][n_out], de-standardized. */
espice_status espice_mlp_predict(espice_mlp *m, const double *x, uint32_t count, double *y);
/* Writes the network's prediction for params[n_lanes][n_in] into the lanes'
 * start point (n_out must equal the template's n), on the lanes' stream;
 * the next espice_lanes_newton_op with init_fix 0 starts there.
 * kernel_ms (optional): GPU time of the inference and the write. */
espice_status espice_lanes_warm_start(espice_lanes *l, espice_mlp *m, const double *params, double *kernel_ms);
/* espice_lanes_warm_start, then espice_lanes_newton_op's kernel from that
 * start (opts->init_fix as given: 0 for a learned start), in one call with
 * one host synchronisation: out[n_lanes] as espice_lanes_newton_op gives it
 * and, when x is not NULL, the solution x[n * n_lanes] as
 * espice_lanes_download_x gives it. warm_ms and newton_ms (optional): GPU
 * time of the inference and of the Newton kernel. No counterpart in the
 * reference, which starts its lanes cold or from the nominal point
 * (src/analysis/dc/op.zig:588,671). */
espice_status espice_lanes_newton_learned(espice_lanes *l, espice_mlp *m, const double *params,
    const espice_newton_opts *opts, espice_lane_result *out, double *x, double *warm_ms, double *newton_ms);

#ifdef __cplusplus
}
#endif

#endif
