"""Learned warm start of operating-point lanes (include/espice_mlp.h).

A sweep's lanes differ only in a few parameters, so their operating points
are a smooth function of those parameters almost everywhere. `train_warm_start`
solves a training set of cold operating points with the engine itself,
fits the network to them on the GPU, and `learned_op` then starts every lane
of a sweep at the network's prediction: Newton from there under the same
tolerances, and a lane that does not converge is rerun from the nominal
operating point (the reference's warm start, src/analysis/dc/op.zig:588,671)
and then cold.

The reference starts every point cold or from the nominal point; the start
only changes where Newton begins, and each lane's result is still the
converged Newton iterate under the unchanged gates.
"""
import time

import numpy as np

from . import Lanes, Mlp

CONVERGED = 1


def cold_ops(template, set_lane_params, params):
    """Cold operating points of len(params) lanes: set_lane_params(lanes,
    params) writes each lane's parameters. Returns (x[N][n], status)."""
    N = len(params)
    lanes = Lanes(template, N)
    set_lane_params(lanes, params)
    lanes.prepare()
    x, st, _, _ = lanes.newton_op()
    return x.T.copy(), st


# Adam's step size falls in three stages (epochs, learning rate): the
# sweep's steep region (an inverter chain's switching point) is learned in
# the later, finer stages.
SCHEDULE = ((600, 1e-3), (300, 2e-4), (300, 5e-5))


def train_warm_start(template, set_lane_params, params, hidden=256, depth=3, schedule=SCHEDULE, batch=256, seed=1):
    """Network trained on the converged cold operating points of `params`
    (rows of sweep parameters). Returns (mlp, info)."""
    t0 = time.perf_counter()
    params = np.asarray(params, dtype=np.float64).reshape(len(params), -1)
    x, st = cold_ops(template, set_lane_params, params)
    ok = (st & CONVERGED) != 0
    t1 = time.perf_counter()
    mlp = Mlp(params.shape[1], template.n, hidden=hidden, depth=depth, seed=seed)
    mlp.fit_scaling(params[ok], x[ok])
    losses = [mlp.train(params[ok], x[ok], epochs=ep, batch=batch, lr=lr, seed=seed + 1 + k)
              for k, (ep, lr) in enumerate(schedule)]
    t2 = time.perf_counter()
    return mlp, {"samples": int(ok.sum()), "solve_s": t1 - t0, "train_s": t2 - t1,
                 "epochs": int(sum(ep for ep, _ in schedule)), "loss_first": float(losses[0][0]),
                 "loss_last": float(losses[-1][-1])}


def learned_op(lanes, mlp, params, nominal_x=None, timing=None, **newton_kw):
    """Every lane from the network's start, then the fallbacks for lanes
    that did not converge: the nominal point (when given), then cold.
    Returns (x[n][N], status, iterations summed over the attempts, GPU ms);
    `timing` (a dict) receives the GPU ms of the inference ("warm_ms") and
    of the Newton kernels ("newton_ms") and the fallback lane counts."""
    kw = dict(newton_kw)
    kw["init_fix"] = False
    x, st, it, w_ms, k_ms = lanes.newton_learned(mlp, params, **kw)
    ms = w_ms + k_ms
    n_ms = k_ms
    fell = {}
    iters = it.astype(np.int64)
    for fallback in ("nominal", "cold"):
        bad = (st & CONVERGED) == 0
        if not bad.any() or (fallback == "nominal" and nominal_x is None):
            continue
        x0 = x.copy()
        if fallback == "nominal":
            x0[:, bad] = np.asarray(nominal_x, dtype=np.float64).reshape(-1, 1)
            lanes.upload_x(x0.reshape(-1))
            x2, st2, it2, k_ms = lanes.newton(**kw)
        else:
            x0[:, bad] = 0.0
            lanes.upload_x(x0.reshape(-1))
            x2, st2, it2, k_ms = lanes.newton(**dict(newton_kw, init_fix=True))
        ms += k_ms
        n_ms += k_ms
        fell[fallback] = int(bad.sum())
        x[:, bad], st[bad] = x2[:, bad], st2[bad]
        iters[bad] += it2[bad]
    if timing is not None:
        timing.update(warm_ms=w_ms, newton_ms=n_ms, fallbacks=fell)
    return x, st, iters, ms
