"""Benchmark of the MI355X lane engine on its sweep workloads.

One step = one full lane solve of the workload on the GPU: reset every
lane's solver, place the start points, run the static Newton kernel to
convergence, read back the per-lane outcomes. Prints one JSON line.

Workload: the 64-point DC sweep of a 5-stage mos1 CMOS inverter chain, each
sweep point an independent operating point on its own lane, in two arms at
the same Newton tolerances:
  plain    every lane starts cold (MODEINITJCT seeds, INITFIX);
  learned  every lane starts at the warm-start network's prediction
           (egspice_amd/warm.py), with the nominal point and then a cold
           start as fallbacks; the network is trained beforehand, untimed,
           on 4096 cold operating points of other input voltages that the
           engine solves itself.
`value` is the learned arm's converged lanes/s; the plain arm's is in
`config`.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged sweep lanes/sec at equal Newton tol: learned+batched vs plain GPU lanes"
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)

# Static FP64 operation counts of one evaluation per device kind (value +
# 4-wide forward-mode derivatives, the limiter, the stamps), counted from
# egspice_amd/csrc/devices.h and engine.h; transcendentals count 20 each.
EVAL_FLOPS = {0: 8, 1: 2, 2: 10, 3: 0, 4: 90, 5: 330}
LIMIT_FLOPS = {4: 12, 5: 40}


def lu_flops(col_ptr, row_idx, vals, n):
    """Refactor + solve flops of one lane tick on the lead factor's
    structure: 2 sum over U entries of |L column| + |L| divides (the
    tape's count, sparse_lu.zig:1186), solve 2(|L| + |U|) + n."""
    import egspice_amd as E
    tab, idx = E.tab_build(n, col_ptr, row_idx, vals)
    lp = idx[tab.lp:tab.lp + n + 1]
    ui = idx[tab.ui:tab.ui + tab.nnz_u]
    refactor = 2 * int(sum(int(lp[i + 1] - lp[i]) for i in ui)) + int(tab.nnz_l)
    solve = 2 * (int(tab.nnz_l) + int(tab.nnz_u)) + n
    return refactor + solve, tab


def host_threads():
    """Host cores this process may use, capped at the 16 a one-GPU share
    of the machine gets."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return max(1, min(16, cores))


def cpu_baseline(n_lanes, budget_s=10.0):
    """The oracle's restatement of the reference CPU path (converger.newton
    with a direct.Solver per lane on the template's one column order) on the
    same workload, lanes dealt over the host's cores, repeated for about
    budget_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from egspice_amd import workloads as W
    lib = O.load()
    c = O.Circuit(lib)
    W.inverter_chain(c.add)
    c.freeze()
    vin = W.dc_sweep_points(n_lanes)
    L = O.Lanes(c, n_lanes)
    L.set_param(O.VSRC, 1, O.VSRC_DC, vin)
    L.prepare()
    threads = host_threads()
    reps, t0 = 0, time.perf_counter()
    converged = 0
    while True:
        _, st, _ = L.newton_op(mode=1, threads=threads)
        converged += int((st & 1).sum())
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    return {"value": converged / el, "unit": "lanes/s", "cores": threads, "kind": "port",
            "sample": f"{reps} repetitions of the {n_lanes}-lane sweep ({reps * n_lanes} lane solves, {el:.1f} s), "
                      f"oracle CPU path on {threads} threads, {platform.processor() or platform.machine()}"}


def profiled_traffic(name):
    """HBM bytes per launch measured for this kernel and workload by the
    FETCH_SIZE/WRITE_SIZE counter passes, copy-calibrated, as recorded in
    profiles/<name>.md (`traffic_bytes_per_launch:`); None when absent."""
    path = os.path.join(ROOT, "profiles", name + ".md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def copy_calibration(mib):
    """Five device-to-device copies of `mib` MiB (hipMemcpy D2D, the
    runtime's copy kernel): a kernel whose HBM traffic is known exactly
    (mib MiB read and mib MiB written per dispatch), against which the
    profiler's FETCH_SIZE and WRITE_SIZE counters are scaled."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    assert bool((dst[-4:] == src[-4:]).all())
    print(json.dumps({"copy_calibration_mib": mib, "dispatches": 5, "bytes_read_per_dispatch": mib << 20,
                      "bytes_written_per_dispatch": mib << 20}), flush=True)


def tran_lanes(N, seed=11):
    """The measure.sp RC deck on N lanes: R ~ U(500, 2000) ohm and C ~
    U(0.5, 2) nF per lane (lane 0 nominal), its ten .meas lines as passes."""
    import egspice_amd as E
    from egspice_amd import workloads as W
    t = E.Template()
    W.rc_pulse(t.add)
    t.freeze()
    rng = np.random.default_rng(seed)
    r = rng.uniform(500.0, 2000.0, N)
    c = rng.uniform(0.5e-9, 2e-9, N)
    r[0], c[0] = 1e3, 1e-9
    vin, vout = t.node("in"), t.node("out")
    passes = [E.meas_cross(vout, 0.1, rise=1), E.meas_cross(vout, 0.9, rise=1), E.meas_cross(vin, 0.5, rise=1),
              E.meas_cross(vout, 0.5, rise=1), E.meas_at(vout, 3e-6), E.meas_cross(vout, 0.5, fall=1),
              E.meas_extremum(vout, 1), E.meas_extremum(vout, 2, frm=0.0, to=10e-6),
              E.meas_integ(vout, rms=True, frm=0.0, to=10e-6), E.meas_at(vout, 2e-6, deriv=True),
              E.meas_integ(2, frm=1e-6, to=6e-6)]
    return t, r, c, passes


def bench_tran(args):
    """Transient lanes/s: one step = every lane's whole `.tran 10n 10u`
    (operating point, ~1029 accepted points, the .meas passes) in one launch."""
    import egspice_amd as E
    N = args.lanes
    t, r, c, passes = tran_lanes(N)
    lanes = E.Lanes(t, N)
    lanes.set_param(E.RESISTOR, 0, 0, r)
    lanes.set_param(E.CAPACITOR, 0, 0, c)
    lanes.prepare()
    opts = E.tran_opts(10e-6, 10e-9, max_steps=200_000)

    def step():
        return lanes.tran(opts, passes)

    for _ in range(args.warmup):
        step()
    kms, done, t0 = [], 0, time.perf_counter()
    for _ in range(args.steps):
        res, meas, _, _, ms = step()
        kms.append(ms)
        done += int(((res["status"] == 0) & np.all(np.isfinite(meas[:, :, 0]), axis=1)).sum())
    wall = time.perf_counter() - t0
    attempts = float(res["attempts"].mean())
    # FP64 model of one lane: per Newton tick the evaluation (V 8, R 10,
    # C 3 with its charge), the companion rows and a = g + ag0 c (5n + 2 nnz),
    # refactor + solve on the lane's tape, update and gates (12n); per
    # accepted step the LTE bound (30 per charge site), the predictor (3n),
    # advanceCurrent (4n) and the .meas passes (10 each)
    lu, _ = lu_flops(t.col_ptr, t.row_idx, np.ones(t.nnz), t.n)
    per_tick = 8 + 10 + 3 + 5 * t.n + 2 * t.nnz + lu + 12 * t.n
    per_step = 30 + 7 * t.n + 10 * len(passes)
    flops = float(res["nr_iters"].sum()) * per_tick + float(res["steps"].sum()) * per_step
    kms = float(np.mean(kms))
    achieved = flops / (kms * 1e-3) / 1e12
    line = {
        "metric": "converged transient lanes/sec (every .meas resolved)", "value": done / wall, "unit": "lanes/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "docs/examples/measure.sp RC, .tran 10n 10u, 11 .meas passes, R and C per lane",
                   "lanes": N, "unknowns": t.n, "points_per_lane": float(res["points"].mean()),
                   "attempts_per_lane": attempts, "newton_ticks_per_lane": float(res["nr_iters"].mean()),
                   "kernel_ms": kms},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": profiled_traffic("tran_wave_kernel_rc_measure_4096") if N == 4096 else None,
                     "kernel": "tran_wave_kernel", "flops_per_launch": flops},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_tran(N, args.cpu_budget)
    print(json.dumps(line), flush=True)


def cpu_baseline_tran(n_lanes, budget_s=10.0):
    """The oracle's transient (tran.Sim restated) over the same lanes, one
    lane per call on one thread, for about budget_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from egspice_amd import workloads as W
    lib = O.load()
    c = O.Circuit(lib)
    W.rc_pulse(c.add)
    c.freeze()
    _, r, cc, _ = tran_lanes(n_lanes)
    L = O.Lanes(c, n_lanes)
    L.set_param(O.RES, 0, 0, r)
    L.set_param(O.CAP, 0, 0, cc)
    L.prepare()
    opts = O.tran_opts(10e-6, 10e-9, max_steps=200_000)
    done, t0 = 0, time.perf_counter()
    while True:
        _, _, st = L.tran(done % n_lanes, opts)
        done += 1 if st.status == 0 else 0
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    return {"value": done / el, "unit": "lanes/s", "cores": 1, "kind": "port",
            "sample": f"{done} lane transients in {el:.1f} s (waveform recorded, .meas not included), "
                      f"oracle CPU path on 1 thread, {platform.processor() or platform.machine()}"}


OPAMP_STEP = (1.0, 10e-9, 1e-9, 500e-9)  # VINP PULSE(0.9 1.0 10n 1n 1n 500n 1)


def opamp_template(step=OPAMP_STEP):
    import egspice_amd as E
    from egspice_amd import workloads as W
    t = E.Template()
    inst = W.miller_opamp(t.add, lambda nm: E.param_index(E.PSP103, nm), step=step)
    t.freeze()
    return t, inst


def opamp_lanes(t, inst, inputs):
    import egspice_amd as E
    from egspice_amd import workloads as W
    lanes = E.Lanes(t, len(inputs))
    for kind, name, slot, vals in W.opamp_lane_params(inputs, lambda nm: E.param_index(E.PSP103, nm)):
        lanes.set_param(kind, inst[name], slot, vals)
    lanes.prepare()
    return lanes


def opamp_nominal(t, inst):
    """The nominal operating point by the operating-point ladder from a
    cold start (the plain rung, then dynamic gmin stepping), one lane."""
    lanes = opamp_lanes(t, inst, np.array([[1, 1, 1, 1, 1.8, 27.0, 1, 1e-12]]))
    xs, st, _, rung, _ = lanes.op_ladder()
    if st[0] != 1:
        raise RuntimeError("the nominal operating point did not converge")
    return xs[:, 0].copy()


def mixed_summary(op_m, op_n, node_rows):
    """The mixed-precision operating points (FP32 factor and solves, FP64
    refinement) against the FP64 lanes from the same start: convergence,
    ticks, FP32 solves and FP64 fallbacks per lane, and the error per lane
    (the largest node-voltage difference; its maximum and median)."""
    x_m, st_m, it_m, solves, fallbacks, ms = op_m
    x_n, st_n, it_n = op_n[0], op_n[1], op_n[2]
    both = ((st_m & 1) != 0) & ((st_n & 1) != 0)
    err = np.abs(x_m - x_n)[node_rows][:, both].max(axis=0) if both.any() else np.zeros(0)
    rel = (np.abs(x_m - x_n)[node_rows][:, both] / np.maximum(np.abs(x_n[node_rows][:, both]), 1e-30)).max(axis=0) \
        if both.any() else np.zeros(0)
    n = len(st_m)
    return {"newton_ms": ms, "converged_lanes": int((st_m & 1).sum()), "ticks_per_lane": float(it_m.sum()) / n,
            "same_ticks_as_fp64_lanes": int((it_m == it_n).sum()),
            "fp32_solves_per_lane": float(solves.sum()) / n, "fp32_solves_per_tick": float(solves.sum()) /
            max(float(it_m.sum()), 1.0), "fp64_fallback_ticks": int(fallbacks.sum()),
            "max_node_err_vs_fp64_V": float(err.max()) if err.size else None,
            "median_node_err_vs_fp64_V": float(np.median(err)) if err.size else None,
            "max_node_rel_err_vs_fp64": float(rel.max()) if rel.size else None}


def opamp_passes(t):
    """The op-amp's .meas passes: .meas tran vmax MAX v(out); .meas tran
    tpd WHEN v(out)=0.95 RISE=1 (the 50% point of the 100 mV step)."""
    import egspice_amd as E
    out = t.node("out")
    return [E.meas_extremum(out, 1), E.meas_cross(out, 0.95, rise=1)]


def bench_opamp(args):
    """Config 5: the psp103 Miller op-amp's Monte Carlo lanes, each an
    operating point and a 1 us transient (`.tran 10n 1u`, a 100 mV input
    pulse, two .meas passes), learned warm start of the operating point
    against the nominal warm start at the same Newton tolerances. One step
    is one launch of every lane's whole analysis."""
    import egspice_amd as E
    from egspice_amd import workloads as W
    N = args.lanes
    t, inst = opamp_template()
    nominal = opamp_nominal(t, inst)
    inputs = W.opamp_inputs(N, seed=1)
    lanes = opamp_lanes(t, inst, inputs)
    passes = opamp_passes(t)
    x_nom = np.repeat(nominal[:, None], N, axis=1).reshape(-1)
    topts = E.tran_opts(1e-6, 10e-9, op_warm=True, max_steps=200_000)

    def lane_ok(res, meas):
        return (res["status"] == 0) & np.all(np.isfinite(meas[:, :, 0]), axis=1)

    def nominal_step():
        lanes.upload_x(x_nom)
        res, meas, _, _, ms = lanes.tran(topts, passes)
        return res, meas, ms, 0.0

    def timed(step, steps):
        # one untimed launch warms every kernel (further warm-up launches of
        # a multi-second step only repeat it)
        for _ in range(min(args.warmup, 1)):
            step()
        kms, wms, ticks, conv = [], [], [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            res, meas, ms, wm = step()
            kms.append(ms)
            wms.append(wm)
            ticks.append(float(res["nr_iters"].sum()))
            conv.append(int(lane_ok(res, meas).sum()))
        return res, meas, kms, wms, ticks, conv, time.perf_counter() - t0

    # the nominal arm is the comparison, reported in config: a few steps
    nominal_steps = min(args.steps, 5)
    res_n, meas_n, kms_n, _, ticks_n, conv_n, wall_n = timed(nominal_step, nominal_steps)
    nominal_value = sum(conv_n) / wall_n

    # operating points alone, from the nominal point (the training set's
    # solver and the OP-only comparison)
    def op_nominal():
        lanes.upload_x(x_nom)
        return lanes.newton(init_fix=False)

    op_n = [op_nominal() for _ in range(2)][-1]

    # the same operating points with FP32 factors and FP64 refinement: the
    # error against the FP64 lanes (bitwise the oracle's) per lane
    def op_mixed():
        lanes.upload_x(x_nom)
        return lanes.newton_mixed(init_fix=False)

    op_m = [op_mixed() for _ in range(2)][-1]

    # training set: other Monte Carlo draws, solved by the engine from the
    # nominal point (untimed; reported with the amortised rate)
    t_train = time.perf_counter()
    train_in = W.opamp_inputs(args.train_lanes, seed=2)
    tl = opamp_lanes(t, inst, train_in)
    tl.upload_x(np.repeat(nominal[:, None], len(train_in), axis=1).reshape(-1))
    xt, stt, _, _ = tl.newton(init_fix=False)
    ok = (stt & 1) != 0
    del tl
    solve_s = time.perf_counter() - t_train
    mlp = E.Mlp(8, t.n, hidden=256, depth=3, seed=1)
    mlp.fit_scaling(train_in[ok], xt[:, ok].T)
    t1 = time.perf_counter()
    losses = [mlp.train(train_in[ok], xt[:, ok].T, epochs=ep, batch=512, lr=lr, seed=3 + k)
              for k, (ep, lr) in enumerate(OPAMP_SCHEDULE)]
    train_s = time.perf_counter() - t1

    def learned_step():
        wm = lanes.warm_start(mlp, inputs)
        res, meas, _, _, ms = lanes.tran(topts, passes)
        return res, meas, ms, wm

    res_l, meas_l, kms_l, wms_l, ticks_l, conv_l, wall_l = timed(learned_step, args.steps)
    value = sum(conv_l) / wall_l

    # both arms once more with the operating point solved to reltol 1e-9
    # (the steps keep 1e-3): the two starts then reach nearly the same
    # operating point; how many lanes' .meas values then agree to the
    # north star's 1e-9
    tight = E.tran_opts(1e-6, 10e-9, op_warm=True, max_steps=200_000, op_reltol=1e-9)
    lanes.warm_start(mlp, inputs)
    res_tl, meas_tl, _, _, ms_tl = lanes.tran(tight, passes)
    lanes.upload_x(x_nom)
    res_tn, meas_tn, _, _, ms_tn = lanes.tran(tight, passes)
    ok_t = lane_ok(res_tl, meas_tl) & lane_ok(res_tn, meas_tn)
    tight_dev = np.abs(meas_tl[ok_t, :, 0] - meas_tn[ok_t, :, 0]) / np.maximum(np.abs(meas_tn[ok_t, :, 0]), 1e-30)
    tight_op = {"op_reltol": 1e-9, "lanes_both_ok": int(ok_t.sum()),
                "meas_max_rel_dev_learned_vs_nominal": float(tight_dev.max()) if tight_dev.size else None,
                "meas_median_rel_dev_learned_vs_nominal": float(np.median(tight_dev)) if tight_dev.size else None,
                "lanes_meas_within_1e-9": int((tight_dev.max(axis=1) <= 1e-9).sum()) if tight_dev.size else 0,
                "learned_tran_kernel_ms": ms_tl, "nominal_tran_kernel_ms": ms_tn,
                "learned_newton_ticks_per_lane": float(res_tl["nr_iters"].sum()) / N,
                "nominal_newton_ticks_per_lane": float(res_tn["nr_iters"].sum()) / N}

    # operating points alone from the network's start
    op_l = [lanes.newton_learned(mlp, inputs, init_fix=False) for _ in range(2)][-1]
    x_l, st_l = op_l[0], op_l[1]
    x_n, st_n = op_n[0], op_n[1]
    node_rows = [t.node(nm) for nm in ("vdd", "inp", "nbias", "tail", "n1", "n2", "out")]
    both = ((st_l & 1) != 0) & ((st_n & 1) != 0)
    dev = np.abs(x_l - x_n)[node_rows][:, both]
    ok_both = lane_ok(res_l, meas_l) & lane_ok(res_n, meas_n)
    meas_dev = np.abs(meas_l[ok_both, :, 0] - meas_n[ok_both, :, 0]) / np.maximum(np.abs(meas_n[ok_both, :, 0]), 1e-30)
    amortised = (sum(conv_l) / args.steps) / (wall_l / args.steps + (solve_s + train_s))
    line = {
        "metric": METRIC, "value": value, "unit": "lanes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall_l * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "psp103 two-stage Miller op-amp (unity-gain buffer, 34 unknowns) Monte Carlo: "
                               "operating point + .tran 10n 1u with a 100 mV input pulse and 2 .meas passes per lane; "
                               "learned warm start vs nominal warm start",
                   "lanes": N, "unknowns": t.n, "pattern_nnz": t.nnz,
                   "inputs": "4 width scales U(0.9,1.1), VDD U(1.6,2.0), T U(-40,125) degC, Ibias U(0.8,1.2)x, "
                             "CL U(0.1,2) pF, seeded",
                   "arm": "learned warm start (BF16 MFMA MLP 8->3x256->34) of the operating point + FP64 GPU "
                          "transient lanes",
                   "newton": "reltol 1e-3 abstol 1e-12 vntol 1e-6 residual_tol 1e-9 itl1 100",
                   "converged_lanes": conv_l[-1],
                   "newton_ticks_per_lane": ticks_l[-1] / N, "tran_kernel_ms": float(np.mean(kms_l)),
                   "warm_start_ms": float(np.mean(wms_l)),
                   "points_per_lane": float(res_l["points"].mean()),
                   "nominal_lanes_per_s": nominal_value, "nominal_ms_per_step": wall_n * 1e3 / nominal_steps,
                   "nominal_steps": nominal_steps, "warmup_launches_per_arm": min(args.warmup, 1),
                   "nominal_converged_lanes": conv_n[-1], "nominal_newton_ticks_per_lane": ticks_n[-1] / N,
                   "nominal_tran_kernel_ms": float(np.mean(kms_n)),
                   "speedup_vs_nominal": value / nominal_value,
                   "meas_max_rel_dev_learned_vs_nominal": float(meas_dev.max()) if meas_dev.size else None,
                   "tight_op": tight_op,
                   "op_only": {"learned_newton_ms": op_l[4], "learned_inference_ms": op_l[3],
                               "learned_ticks_per_lane": float(op_l[2].sum()) / N,
                               "nominal_newton_ms": op_n[3], "nominal_ticks_per_lane": float(op_n[2].sum()) / N,
                               "max_node_dev_learned_vs_nominal_V": float(dev.max()) if dev.size else None,
                               "mixed": mixed_summary(op_m, op_n, node_rows)},
                   "training": {"samples": int(ok.sum()), "epochs": int(sum(e for e, _ in OPAMP_SCHEDULE)),
                                "batch": 512, "schedule": [list(e) for e in OPAMP_SCHEDULE],
                                "solve_s": round(solve_s, 3), "train_s": round(train_s, 3),
                                "loss_first": float(losses[0][0]), "loss_last": float(losses[-1][-1]),
                                "amortised_lanes_per_s": amortised}},
    }
    flops = opamp_tran_flops(t, res_l, len(passes))
    achieved = flops / (float(np.mean(kms_l)) * 1e-3) / 1e12
    line["roofline"] = {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP64_PEAK_TFLOPS, "traffic": profiled_traffic(f"tran_wave_kernel_opamp_{N}"),
                        "kernel": "tran_wave_kernel", "flops_per_launch": flops}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_opamp(t, inst, nominal, args.cpu_budget)
    print(json.dumps(line), flush=True)


# the learning rate falls in stages (epochs, rate)
OPAMP_SCHEDULE = ((60, 1e-3), (30, 2e-4), (30, 5e-5))


def opamp_tran_flops(t, res, n_passes):
    """Algorithmic FP64 work of one transient launch: per Newton tick the
    eight PSP103 evaluations with charges, the other devices, the companion
    rows and a = g + ag0 c, refactor + solve, update and gates; per accepted
    step the charges re-read at the solution, the LTE bound over the charge
    sites, the predictor and currents and the passes."""
    vals = np.random.default_rng(0).uniform(0.1, 1.0, t.nnz)
    vals[t.diag] = 2.0 * t.n
    lu, _ = lu_flops(t.col_ptr, t.row_idx, vals, t.n)
    sites = 8 * PSP103_SITES + 2
    per_tick = 8 * PSP103_TRAN_EVAL_FLOPS + 2 * EVAL_FLOPS[0] + EVAL_FLOPS[1] + 5 * t.n + 2 * t.nnz + lu + 12 * t.n
    per_step = 8 * PSP103_CHARGE_FLOPS + 30 * sites + 7 * t.n + 10 * n_passes
    return float(res["nr_iters"].sum()) * per_tick + float(res["steps"].sum()) * per_step


# FP64 operations of one PSP103 evaluation (value and derivative lanes, the
# stamps; the charges alone: values only) along the paths the op-amp's
# transistors take at its nominal operating point, mean of the eight:
# `python -m egspice_amd.vagen.flops egspice_amd/csrc/gen/psp103.h psp103`
# (FMA 2, exp and log 30, pow 62)
PSP103_EVAL_FLOPS = 5724
PSP103_TRAN_EVAL_FLOPS = 6529
PSP103_CHARGE_FLOPS = 1616
PSP103_SITES = 11
# the same for BSIM4 at an inverter's transistors over its input swing:
# `python -m egspice_amd.vagen.flops egspice_amd/csrc/gen/bsim4va.h bsim4va`
BSIM4_EVAL_FLOPS = 2315
BSIM4_TRAN_EVAL_FLOPS = 3633
BSIM4_CHARGE_FLOPS = 1630
BSIM4_SITES = 5


def ring_tran_flops(t, res, n_passes, stages=31):
    """Algorithmic FP64 work of one ring transient launch: per Newton tick
    the 62 BSIM4 evaluations with charges and the limiter's passes, the
    loads, the source, the companion rows, a = g + ag0 c, refactor + solve,
    update and gates; per step the charges re-read at the solution, the LTE
    bound, the predictor and the passes."""
    vals = np.random.default_rng(0).uniform(0.1, 1.0, t.nnz)
    vals[t.diag] = 2.0 * t.n
    lu, _ = lu_flops(t.col_ptr, t.row_idx, vals, t.n)
    sites = 2 * stages * BSIM4_SITES + stages
    per_tick = (2 * stages * (BSIM4_TRAN_EVAL_FLOPS + 60) + stages * 8 + EVAL_FLOPS[0] + 5 * t.n + 2 * t.nnz + lu
                + 12 * t.n)
    per_step = 2 * stages * BSIM4_CHARGE_FLOPS + 30 * sites + 7 * t.n + 10 * n_passes
    return float(res["nr_iters"].sum()) * per_tick + float(res["steps"].sum()) * per_step


def bench_ring(args):
    """Config 2: the 31-stage BSIM4 ring oscillator's Monte Carlo lanes
    (every transistor's VTH0 ~ N(+-0.3 V, 15 mV), TOXE ~ N(3 nm, 2%)), each
    a cold operating point (MODEINITJCT seeds, limited Newton) and a 5 ns
    trapezoidal transient (`.tran 10p 5n`, a 20 ps kick into n1) with the
    period's two crossings as .meas passes (v(n1) rising through VDD/2,
    the first two after 0.2 ns; frequency = 1/period on the host). A lane counts when its transient reached t_stop and both
    crossings resolved. One step is one launch of every lane's analysis."""
    import egspice_amd as E
    from egspice_amd import workloads as W
    N = args.lanes
    t = E.Template()
    idx = lambda nm: E.param_index(E.BSIM4, nm)
    inst = W.ring_oscillator(t.add, idx)
    t.freeze()
    lanes = E.Lanes(t, N)
    for kind, name, slot, vals in W.ring_lane_params(N, idx, seed=1):
        lanes.set_param(kind, inst[name], slot, vals)
    lanes.prepare()
    n1 = t.node("n1")
    m = W.RING_MEAS
    passes = [E.meas_cross(n1, m["val"], rise=1, td=m["td"]), E.meas_cross(n1, m["val"], rise=2, td=m["td"])]
    topts = E.tran_opts(5e-9, 10e-12, max_steps=200_000)

    def step():
        res, meas, _, _, ms = lanes.tran(topts, passes)
        ok = (res["status"] == 0) & np.all(np.isfinite(meas[:, :, 0]), axis=1)
        return res, meas, ms, int(ok.sum())

    for _ in range(min(args.warmup, 1)):
        step()
    kms, conv = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, meas, ms, c = step()
        kms.append(ms)
        conv.append(c)
    wall = time.perf_counter() - t0
    value = sum(conv) / wall
    ok = (res["status"] == 0) & np.all(np.isfinite(meas[:, :, 0]), axis=1)
    period = meas[ok, 1, 0] - meas[ok, 0, 0]
    line = {
        "metric": "converged Monte Carlo lanes/sec (ring oscillator transient)", "value": value, "unit": "lanes/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "31-stage BSIM4 (bsim4va) ring oscillator Monte Carlo, 1 fF per stage: cold operating "
                               "point + .tran 10p 5n trapezoidal, period by 2 crossing passes",
                   "lanes": N, "unknowns": t.n, "pattern_nnz": t.nnz,
                   "inputs": "per transistor VTH0 N(+-0.3 V, 15 mV), TOXE N(3 nm, 2%), seeded",
                   "completed_lanes": int((res["status"] == 0).sum()), "lanes_with_period": conv[-1],
                   "period_ns_mean": float(period.mean() * 1e9) if period.size else None,
                   "period_ns_std": float(period.std() * 1e9) if period.size else None,
                   "frequency_ghz_mean": float(np.mean(1.0 / period) * 1e-9) if period.size else None,
                   "newton_ticks_per_lane": float(res["nr_iters"].mean()), "steps_per_lane": float(res["steps"].mean()),
                   "tran_kernel_ms": float(np.mean(kms))},
    }
    flops = ring_tran_flops(t, res, len(passes))
    achieved = flops / (float(np.mean(kms)) * 1e-3) / 1e12
    line["roofline"] = {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP64_PEAK_TFLOPS, "traffic": profiled_traffic(f"tran_wave_kernel_ring_{N}"),
                        "kernel": "tran_wave_kernel", "flops_per_launch": flops}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_ring(args.cpu_budget)
    print(json.dumps(line), flush=True)


def cpu_baseline_ring(budget_s=10.0, n_lanes=256):
    """The oracle's CPU path on the same ring lanes (cold operating point,
    `.tran 10p 5n`, both crossing passes), dealt over the host's cores for
    about budget_s seconds."""
    import concurrent.futures as cf
    import ctypes as C
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from egspice_amd import workloads as W
    lib = O.load()
    c = O.Circuit(lib)
    idx = lambda nm: lib.or_va_param_index(O.BSIM4, nm.encode())
    oinst = W.ring_oscillator(c.add, idx)
    c.freeze()
    L = O.Lanes(c, n_lanes)
    for kind, name, slot, vals in W.ring_lane_params(n_lanes, idx, seed=1):
        L.set_param(kind, oinst[name], slot, vals)
    L.prepare()
    opts = O.tran_opts(5e-9, 10e-12, max_steps=200_000)
    n1 = c.node("n1")
    m = W.RING_MEAS
    clauses = [O.clause(val=m["val"], rise=1, td=m["td"]), O.clause(val=m["val"], rise=2, td=m["td"])]

    def lane(l):
        tt, xx, st = L.tran(l, opts)
        if st.status != 0:
            return 0
        y = np.ascontiguousarray(xx[:, n1])
        got = [lib.or_meas_cross(O.f64(tt)[1], O.f64(y)[1], len(tt), C.byref(cl)) for cl in clauses]
        return 1 if all(np.isfinite(g) for g in got) else 0

    threads = host_threads()
    done, t0, l = 0, time.perf_counter(), 0
    with cf.ThreadPoolExecutor(threads) as ex:
        while True:
            done += sum(ex.map(lane, [(l + k) % n_lanes for k in range(threads * 2)]))
            l += threads * 2
            el = time.perf_counter() - t0
            if el >= budget_s:
                break
    return {"value": done / el, "unit": "lanes/s", "cores": threads, "kind": "port",
            "sample": f"{l} Monte Carlo ring lanes (cold operating point, .tran 10p 5n, 2 crossing passes) in "
                      f"{el:.1f} s, oracle CPU path on {threads} threads, {platform.processor() or platform.machine()}"}


def cpu_baseline_opamp(t, inst, nominal, budget_s=10.0, n_lanes=512):
    """The oracle's CPU path on the same Monte Carlo lanes: each lane's
    operating point warm from the nominal point, its `.tran 10n 1u`
    (tran.Sim restated) and the two .meas passes over the waveform, lanes
    dealt over the host's cores (the C oracle releases the interpreter lock),
    for about budget_s seconds."""
    import concurrent.futures as cf
    import ctypes as C
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from egspice_amd import workloads as W
    lib = O.load()
    inputs = W.opamp_inputs(n_lanes, seed=1)
    c = O.Circuit(lib)
    idx = lambda nm: lib.or_va_param_index(O.PSP103, nm.encode())
    oinst = W.miller_opamp(c.add, idx, step=OPAMP_STEP)
    c.freeze()
    L = O.Lanes(c, n_lanes)
    for kind, name, slot, vals in W.opamp_lane_params(inputs, idx):
        L.set_param(kind, oinst[name], slot, vals)
    L.prepare()
    opts = O.tran_opts(1e-6, 10e-9, max_steps=200_000)
    out = c.node("out")
    clause = O.clause(val=0.95, rise=1)

    def lane(l):
        tt, xx, st = L.tran(l, opts, x0=nominal)
        if st.status != 0:
            return 0
        y = np.ascontiguousarray(xx[:, out])
        vmax, at = C.c_double(), C.c_double()
        lib.or_meas_extremum(O.f64(tt)[1], O.f64(y)[1], len(tt), 0.0, 0.0, 1, C.byref(vmax), C.byref(at))
        tpd = lib.or_meas_cross(O.f64(tt)[1], O.f64(y)[1], len(tt), C.byref(clause))
        return 1 if np.isfinite(vmax.value) and np.isfinite(tpd) else 0

    threads = host_threads()
    done, t0, l = 0, time.perf_counter(), 0
    with cf.ThreadPoolExecutor(threads) as ex:
        while True:
            done += sum(ex.map(lane, [(l + k) % n_lanes for k in range(threads * 4)]))
            l += threads * 4
            el = time.perf_counter() - t0
            if el >= budget_s:
                break
    return {"value": done / el, "unit": "lanes/s", "cores": threads, "kind": "port",
            "sample": f"{l} Monte Carlo lanes (operating point from the nominal point, .tran 10n 1u, 2 .meas passes) "
                      f"in {el:.1f} s, oracle CPU path on {threads} threads, {platform.processor() or platform.machine()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--lanes", type=int, default=None,
                    help="lanes (default: 65536 for opamp-mc, 4096 for ring-bsim4 and tran-rc-measure, 64 for "
                         "dc-inverter-chain)")
    ap.add_argument("--train-lanes", type=int, default=65536)
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("opamp-mc", "ring-bsim4", "dc-inverter-chain", "tran-rc-measure"),
                    default="opamp-mc",
                    help="opamp-mc: the psp103 op-amp's Monte Carlo lanes, learned vs nominal warm start "
                         "(default); ring-bsim4: the 31-stage BSIM4 ring oscillator's Monte Carlo transient; "
                         "dc-inverter-chain: the mos1 chain's 64-point DC sweep; tran-rc-measure: the RC transient "
                         "of docs/examples/measure.sp with its .meas passes")
    ap.add_argument("--arm", choices=("both", "learned", "plain"), default="both",
                    help="dc-inverter-chain: time both arms (default), or only one (for profiling one arm's "
                         "kernels); with one arm, the other's fields are null")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (the reference kernel for "
                         "scaling profiler byte counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(lanes shard across GPUs by contiguous lane ranges, a design not built here)", file=sys.stderr)
        sys.exit(2)

    if args.lanes is None:
        args.lanes = {"opamp-mc": 65536, "ring-bsim4": 4096, "dc-inverter-chain": 64,
                      "tran-rc-measure": 4096}[args.workload]
    if args.workload == "ring-bsim4":
        bench_ring(args)
        return
    if args.workload == "tran-rc-measure":
        bench_tran(args)
        return
    if args.workload == "opamp-mc":
        bench_opamp(args)
        return

    import egspice_amd as E
    from egspice_amd import workloads as W

    t = E.Template()
    W.inverter_chain(t.add)
    t.freeze()
    N = args.lanes
    vin = W.dc_sweep_points(N)
    lanes = E.Lanes(t, N)
    lanes.set_param(E.VSOURCE, 1, E.VSOURCE_DC, vin)
    lanes.prepare()
    x0 = np.zeros(t.n * N)

    def plain_step():
        lanes.reset()
        return lanes.newton_op(x0)

    def timed(step):
        for _ in range(args.warmup):
            step()
        kernel_ms, iters_all, conv_all = [], [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            x, st, it, ms = step()
            kernel_ms.append(ms)
            iters_all.append(it.copy())
            conv_all.append(int((st & 1).sum()))
        wall = time.perf_counter() - t0
        return x, kernel_ms, iters_all, conv_all, wall

    if args.arm != "learned":
        x_cold, kernel_ms, iters_all, conv_all, wall = timed(plain_step)
        plain_value = sum(conv_all) / wall
        plain_ms = wall * 1e3 / args.steps
        plain_ticks = float(np.mean([it.sum() for it in iters_all]))
    else:
        lanes.reset()
        x_cold = lanes.newton_op(x0)[0]
        plain_value = plain_ms = plain_ticks = None

    # learned arm: train (untimed) on cold operating points of other inputs
    # (also for --arm plain, so both runs share the setup)
    from egspice_amd import warm

    def set_vin(ln, p):
        ln.set_param(E.VSOURCE, 1, E.VSOURCE_DC, np.asarray(p, dtype=np.float64)[:, 0])

    train_in = np.random.default_rng(2024).uniform(0.0, 1.8, (4096, 1))
    mlp, info = warm.train_warm_start(t, set_vin, train_in, batch=256)
    nominal, _ = warm.cold_ops(t, set_vin, np.zeros((1, 1)))
    sweep = vin.reshape(-1, 1)

    timings = []

    def learned_step():
        lanes.reset()
        tm = {}
        out = warm.learned_op(lanes, mlp, sweep, nominal_x=nominal[0], timing=tm)
        timings.append(tm)
        return out

    if args.arm == "plain":
        x_l, l_kernel_ms, l_iters, l_conv, l_wall = x_cold, kernel_ms, iters_all, conv_all, wall
        timings = [{"newton_ms": k, "warm_ms": 0.0, "fallbacks": {}} for k in kernel_ms]
    else:
        x_l, l_kernel_ms, l_iters, l_conv, l_wall = timed(learned_step)
    timings = timings[-args.steps:]
    newton_ms = float(np.mean([tm["newton_ms"] for tm in timings]))
    warm_ms = float(np.mean([tm["warm_ms"] for tm in timings]))
    fallbacks = sum(sum(tm["fallbacks"].values()) for tm in timings)
    value = sum(l_conv) / l_wall
    ms_per_step = l_wall * 1e3 / args.steps
    converged = l_conv[-1]
    node_rows = [t.node(nm) for nm in ["in", "vdd"] + [f"n{k}" for k in range(1, 6)]]
    max_dev = float(np.abs(x_l - x_cold)[node_rows].max())

    # roofline of the Newton kernel: algorithmic FP64 flops / kernel time
    g_eval = sum(EVAL_FLOPS[k] * t_count for k, t_count in [(0, 2), (5, 2 * 5)])
    g_lim = LIMIT_FLOPS[5] * 10
    # the lead structure at lane 0's seeded first matrix is the typical lane's
    vals = np.ones(t.nnz)
    lu, _ = lu_flops(t.col_ptr, t.row_idx, vals, t.n)
    per_tick = g_eval + g_lim + lu + 12 * t.n
    ticks = float(np.mean([it.sum() for it in l_iters]))
    flops = per_tick * ticks
    kms = float(np.mean(l_kernel_ms))
    achieved = flops / (newton_ms * 1e-3) / 1e12
    line = {
        "metric": METRIC, "value": value, "unit": "lanes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "mos1 5-stage inverter chain, 64-point DC sweep as independent operating points",
                   "lanes": N, "unknowns": t.n, "pattern_nnz": t.nnz,
                   "arm": ("plain FP64 GPU lanes (cold starts)" if args.arm == "plain" else
                           "learned warm start (BF16 MFMA MLP 1->3x256->n) + FP64 GPU lanes"),
                   "newton": "reltol 1e-3 abstol 1e-12 vntol 1e-6 residual_tol 1e-9 itl1 100",
                   "converged_lanes": converged, "newton_ticks_per_solve": ticks,
                   "kernel_ms_per_solve": kms, "newton_kernel_ms": newton_ms, "warm_start_ms": warm_ms,
                   "fallback_lanes": fallbacks,
                   "plain_lanes_per_s": plain_value, "plain_ms_per_step": plain_ms,
                   "plain_newton_ticks_per_solve": plain_ticks,
                   "speedup_vs_plain": value / plain_value if plain_value else None,
                   "max_node_dev_vs_cold_V": max_dev,
                   "training": {"samples": info["samples"], "epochs": info["epochs"], "batch": 256,
                                "schedule": [list(e) for e in warm.SCHEDULE],
                                "solve_s": round(info["solve_s"], 3), "train_s": round(info["train_s"], 3),
                                "loss_first": info["loss_first"], "loss_last": info["loss_last"]}},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": (profiled_traffic("newton_wave_kernel_inverter_chain_64" if args.arm == "plain" else
                                                  "newton_wave_kernel_inverter_chain_64_learned") if N == 64 else None),
                     "kernel": "newton_wave_kernel", "flops_per_launch": flops},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(N, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
