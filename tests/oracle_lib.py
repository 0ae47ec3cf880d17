"""ctypes binding of the CPU oracle (oracle/liboracle.so).

The oracle is the checker of the HIP lane engine. Only tests, the smoke
check and bench.py's CPU-baseline leg load it; the engine never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

VSRC, ISRC, RES, CAP, DIODE, MOS1, PSP103, BSIM4 = range(8)
UK = {VSRC: 71, ISRC: 71, RES: 1, CAP: 1, DIODE: 19, MOS1: 36}
COMPILED = (PSP103, BSIM4)

# user parameter slots (oracle/oracle_circuit.h)
VSRC_DC = 0
DIO = dict(IS=0, N=1, CJO=2, VJ=3, M=4, FC=5, TT=6, GMIN=7, AREA=8, TNOM=9, TEMP=10, DTEMP=11, EG=12, XTI=13)
M1 = {name: i for i, name in enumerate(
    "TYPE VTO KP GAMMA PHI LAMBDA TOX LD U0 NSUB NSS TPG IS JS CBD CBS CJ CJSW PB MJ MJSW FC CGSO CGDO CGBO "
    "TNOM GMIN TEMP DTEMP L W AD AS PD PS M".split())}


class LaneTab(C.Structure):
    _fields_ = [(f, C.c_uint32) for f in ("n", "nnz", "nnz_l", "nnz_u")] + \
               [("growth", C.c_double), ("wide", C.c_double)] + \
               [(f, C.c_uint32) for f in ("col_ptr", "prow", "q", "pinv", "lp", "li", "up", "ui", "flags",
                                          "void_slots", "n_void_slots", "_pad")]


class NewtonOpts(C.Structure):
    _fields_ = [("reltol", C.c_double), ("abstol", C.c_double), ("vntol", C.c_double),
                ("residual_tol", C.c_double), ("gmin", C.c_double),
                ("max_iter", C.c_uint32), ("init_fix", C.c_uint32)]


class TranOpts(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("t_stop", "dt_init", "dt_max", "dt_min", "xmu", "reltol", "abstol", "vntol",
                                          "chgtol", "trtol", "residual_tol_op")] + \
               [(f, C.c_uint32) for f in ("method", "max_iter", "max_steps", "pad")]


class TranStats(C.Structure):
    _fields_ = [(f, C.c_uint32) for f in ("points", "steps", "attempts", "nr_iters", "rej_newton", "rej_lte",
                                          "bp_landings")] + [("status", C.c_int32)]


class MeasClause(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("frm", "to", "val", "td")] + \
               [(f, C.c_int32) for f in ("rise", "fall", "cross", "pad")]


def clause(val=1e99, rise=-1, fall=-1, cross=-1, td=0.0, frm=0.0, to=0.0):
    return MeasClause(frm, to, val, td, rise, fall, cross, 0)


def tran_opts(t_stop, dt_init, dt_max=None, method=1, xmu=0.5, dt_min=1e-18, reltol=1e-3, abstol=1e-12, vntol=1e-6,
              chgtol=1e-14, trtol=7.0, max_iter=100, max_steps=1_000_000_000):
    """`.tran dt_init t_stop` with the deck's ceiling min(dt_init, t_stop/50)
    (src/frontend/analyses.zig:565-566, src/core/query.zig:673-680)."""
    if dt_max is None:
        dt_max = min(dt_init, t_stop / 50)
    return TranOpts(t_stop, dt_init, dt_max, dt_min, xmu, reltol, abstol, vntol, chgtol, trtol, 1e-9, method, max_iter,
                    max_steps, 0)


class Outcome(C.Structure):
    _fields_ = [("status", C.c_uint32), ("iterations", C.c_uint32)]


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    u32p, f64p = C.POINTER(C.c_uint32), C.POINTER(C.c_double)
    sig = {
        "om_exp": (C.c_double, [C.c_double]), "om_log": (C.c_double, [C.c_double]), "om_sin": (C.c_double, [C.c_double]),
        "om_pow": (C.c_double, [C.c_double, C.c_double]),
        "or_order": (C.c_int, [C.c_uint32, u32p, u32p, u32p]),
        "or_amd": (C.c_int, [C.c_uint32, u32p, u32p, u32p]),
        "or_blocks": (C.c_int, [C.c_uint32, u32p, u32p, u32p, u32p]),
        "or_chol_count": (C.c_uint64, [C.c_uint32, u32p, u32p, u32p, C.c_uint64]),
        "or_lu_new": (P, [C.c_uint32, u32p, u32p]), "or_lu_free": (None, [P]),
        "or_lu_factor": (C.c_int, [P, u32p, u32p, f64p, C.c_double]),
        "or_lu_refactor": (C.c_int, [P, u32p, f64p, C.c_double]),
        "or_lu_solve": (None, [P, f64p, f64p]), "or_lu_solve_t": (None, [P, f64p, f64p]),
        "or_lu_set_caps": (None, [P, C.c_double, C.c_double, C.c_uint64]),
        "or_lu_nnz_l": (C.c_uint32, [P]), "or_lu_nnz_u": (C.c_uint32, [P]),
        "or_lu_pinv": (u32p, [P]), "or_lu_lx": (f64p, [P]), "or_lu_ux": (f64p, [P]),
        "or_lu_udiag": (f64p, [P]), "or_lu_w": (f64p, [P]), "or_lu_void_col": (C.POINTER(C.c_uint8), [P]),
        "or_lu_scaled_pivot": (C.POINTER(C.c_uint8), [P]), "or_lu_sn_reached": (C.c_int, [P]),
        "or_solver_new": (P, [C.c_uint32, u32p, u32p]), "or_solver_free": (None, [P]),
        "or_solver_factor": (C.c_int, [P, f64p]), "or_solver_solve_neg": (None, [P, f64p, f64p]),
        "or_solver_lu": (P, [P]), "or_solver_q": (u32p, [P]),
        "or_is_tridiag": (C.c_int, [C.c_uint32, u32p, u32p]),
        "or_build_tab": (u32p, [P, u32p, u32p, C.c_double, C.c_double, C.POINTER(LaneTab), C.POINTER(C.c_size_t)]),
        "or_lane_refactor": (C.c_int, [C.POINTER(LaneTab), u32p, f64p, f64p, f64p, f64p, f64p, C.c_uint32, C.c_uint32]),
        "or_lane_solve_neg": (None, [C.POINTER(LaneTab), u32p, f64p, f64p, f64p, f64p, f64p, f64p, C.c_uint32, C.c_uint32]),
        "or_dense_solve_neg": (C.c_int, [C.c_uint32, f64p, f64p]),
        "or_ckt_new": (P, []), "or_ckt_free": (None, [P]),
        "or_ckt_node": (C.c_uint32, [P, C.c_char_p]),
        "or_ckt_add": (C.c_int32, [P, C.c_int, u32p, f64p]),
        "or_ckt_freeze": (C.c_int, [P]), "or_ckt_n": (C.c_uint32, [P]), "or_ckt_nnz": (C.c_uint32, [P]),
        "or_ckt_col_ptr": (u32p, [P]), "or_ckt_row_idx": (u32p, [P]), "or_ckt_diag": (u32p, [P]),
        "or_ckt_cur_row": (C.POINTER(C.c_uint8), [P]),
        "or_lanes_new": (P, [P, C.c_uint32]), "or_lanes_free": (None, [P]),
        "or_lanes_set_param": (None, [P, C.c_int, C.c_uint32, C.c_uint32, f64p]),
        "or_lanes_prepare": (None, [P]),
        "or_newton_op": (C.c_int, [P, C.c_int, C.POINTER(NewtonOpts), f64p, C.POINTER(Outcome)]),
        "or_newton_op_mt": (C.c_int, [P, C.c_int, C.POINTER(NewtonOpts), f64p, C.POINTER(Outcome), C.c_uint32]),
        "or_lead_tab": (C.c_int, [P, f64p, C.c_int, C.POINTER(LaneTab), C.POINTER(u32p), C.POINTER(C.c_size_t)]),
        "or_eval_lane": (None, [P, C.c_uint32, f64p, f64p, f64p]),
        "or_dc_sweep": (C.c_int, [P, C.c_uint32, C.c_double, C.c_double, C.c_uint32, C.POINTER(NewtonOpts), f64p, u32p]),
        "or_newton_hom": (C.c_int, [P, C.c_uint32, C.POINTER(NewtonOpts), f64p, C.POINTER(Outcome), u32p, C.c_double]),
        "or_lanes_set_model_gmin": (None, [P, C.c_uint32, C.c_double]),
        "or_gmin_stamps": (None, [P, C.c_uint32, u32p]),
        "or_gmin_factor": (C.c_double, [C.c_double, C.c_uint32, C.c_uint32]),
        "or_gmin_stalled": (C.c_int, [f64p, C.POINTER(C.c_int), C.c_uint32, C.c_double, C.c_double]),
        "or_gmin_replay": (C.c_uint32, [C.POINTER(C.c_int), u32p, C.c_uint32, C.c_double, C.c_uint32, f64p,
                                        C.POINTER(C.c_int)]),
        "or_op_ladder": (C.c_int, [P, C.POINTER(NewtonOpts), C.c_double, C.c_uint32, f64p, C.POINTER(Outcome), u32p]),
        "or_eval_lane_gc": (None, [P, C.c_uint32, f64p, f64p, f64p]),
        "or_dc_sweep_lanes": (C.c_int, [P, C.c_int, C.c_uint32, f64p, C.c_uint32, C.POINTER(NewtonOpts), C.c_double,
                                        C.c_uint32, f64p, u32p]),
        "or_fetlim": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "or_tran_coeffs": (None, [C.c_int, C.c_double, C.c_double, C.c_double, f64p]),
        "or_tran_run": (C.c_int, [P, C.c_uint32, C.POINTER(TranOpts), C.POINTER(f64p), C.POINTER(f64p),
                                  C.POINTER(TranStats)]),
        "or_tran_run_from": (C.c_int, [P, C.c_uint32, C.POINTER(TranOpts), f64p, C.POINTER(f64p), C.POINTER(f64p),
                                       C.POINTER(TranStats)]),
        "or_meas_cross": (C.c_double, [f64p, f64p, C.c_uint32, C.POINTER(MeasClause)]),
        "or_meas_at": (C.c_double, [f64p, f64p, C.c_uint32, C.c_double, C.c_int]),
        "or_meas_find_when": (C.c_double, [f64p, f64p, f64p, C.c_uint32, C.POINTER(MeasClause), C.c_int]),
        "or_meas_extremum": (None, [f64p, f64p, C.c_uint32, C.c_double, C.c_double, C.c_int, f64p, f64p]),
        "or_meas_integ": (None, [f64p, f64p, C.c_uint32, C.c_double, C.c_double, C.c_int, f64p, f64p, f64p]),
        "or_limvds": (C.c_double, [C.c_double, C.c_double]),
        "or_site_key": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "or_mc_uniform": (C.c_double, [C.c_uint64, C.c_uint64, C.c_uint32]),
        "or_normal_quantile": (C.c_double, [C.c_double]),
        "or_trial_draw": (C.c_double, [C.c_uint64, C.c_uint32, C.c_uint64, C.c_int, f64p, C.c_uint32]),
        "or_wyhash": (C.c_uint64, [C.c_uint64, C.c_char_p, C.c_size_t]),
        "or_param_site": (C.c_uint64, [C.c_char_p, C.c_uint32]),
        "or_va_nparam": (C.c_int, [C.c_int]), "or_va_nu": (C.c_int, [C.c_int]),
        "or_va_nport": (C.c_int, [C.c_int]), "or_va_nprep": (C.c_int, [C.c_int]),
        "or_va_nsite": (C.c_int, [C.c_int]),
        "or_va_gmask": (C.c_uint, [C.c_int, C.c_int]), "or_va_cmask": (C.c_uint, [C.c_int, C.c_int]),
        "or_va_param_index": (C.c_int, [C.c_int, C.c_char_p]),
        "or_va_setup": (C.c_int, [C.c_int, f64p, f64p]),
        "or_va_eval": (None, [C.c_int, f64p, f64p, f64p, f64p]),
        "or_va_eval_tran": (None, [C.c_int, f64p, f64p, f64p, f64p, f64p, f64p, f64p]),
        "or_va_eval_charge": (None, [C.c_int, f64p, f64p, f64p, f64p]),
        "or_va_limw": (C.c_uint, [C.c_int]),
        "or_va_limit": (C.c_int, [C.c_int, f64p, f64p, f64p, f64p]),
        "or_va_seed": (None, [C.c_int, f64p, f64p, f64p]),
        "or_va_limit_args": (None, [C.c_int, f64p, f64p, f64p]),
        "or_limit_mos1": (C.c_int, [f64p, f64p, f64p, f64p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib.free_ = C.CDLL(None).free
    lib.free_.argtypes = [C.c_void_p]
    return lib


def u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint32))


def f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def newton_opts(init_fix=True, max_iter=100, reltol=1e-3, abstol=1e-12, vntol=1e-6, residual_tol=1e-9, gmin=0.0):
    """converger.Options defaults (src/core/numerics.zig:378-399)."""
    return NewtonOpts(reltol, abstol, vntol, residual_tol, gmin, max_iter, 1 if init_fix else 0)


class Circuit:
    """Oracle circuit template built card by card."""

    def __init__(self, lib):
        self.lib = lib
        self.h = lib.or_ckt_new()
        self.frozen = False
        self.counts = [0] * 8

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_ckt_free(self.h)
            self.h = None

    def node(self, name):
        return self.lib.or_ckt_node(self.h, name.encode())

    def add(self, kind, ports, params=None):
        rows = [self.node(p) if isinstance(p, str) else p for p in ports]
        if kind in COMPILED and kind not in UK:
            UK[kind] = self.lib.or_va_nparam(kind)
        pr = np.full(UK[kind], np.nan)
        if params:
            for k, v in params.items():
                pr[k] = v
        _, pp = u32(rows)
        a, ap = f64(pr)
        r = self.lib.or_ckt_add(self.h, kind, pp, ap)
        assert r >= 0
        self.counts[kind] += 1
        return r

    def freeze(self):
        assert self.lib.or_ckt_freeze(self.h) == 0
        self.frozen = True
        self.n = self.lib.or_ckt_n(self.h)
        self.nnz = self.lib.or_ckt_nnz(self.h)
        self.col_ptr = np.ctypeslib.as_array(self.lib.or_ckt_col_ptr(self.h), (self.n + 1,)).copy()
        self.row_idx = np.ctypeslib.as_array(self.lib.or_ckt_row_idx(self.h), (self.nnz,)).copy()
        self.diag = np.ctypeslib.as_array(self.lib.or_ckt_diag(self.h), (self.n,)).copy()
        self.cur_row = np.ctypeslib.as_array(self.lib.or_ckt_cur_row(self.h), (self.n,)).copy()
        return self


class Lanes:
    def __init__(self, ckt, n_lanes):
        self.ckt = ckt
        self.lib = ckt.lib
        self.N = n_lanes
        self.h = self.lib.or_lanes_new(ckt.h, n_lanes)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_lanes_free(self.h)
            self.h = None

    def set_param(self, kind, inst, k, values):
        v, vp = f64(np.broadcast_to(np.asarray(values, dtype=np.float64), (self.N,)))
        self.lib.or_lanes_set_param(self.h, kind, inst, k, vp)

    def prepare(self):
        self.lib.or_lanes_prepare(self.h)

    def newton_op(self, mode=0, x0=None, threads=1, **kw):
        n, N = self.ckt.n, self.N
        x = np.zeros(n * N) if x0 is None else np.array(x0, dtype=np.float64).reshape(-1)
        xa, xp = f64(x)
        out = (Outcome * N)()
        o = newton_opts(**kw)
        rc = self.lib.or_newton_op_mt(self.h, mode, C.byref(o), xp, out, threads)
        if rc != 0:
            raise RuntimeError("no lane tape")
        status = np.array([out[i].status for i in range(N)], dtype=np.uint32)
        iters = np.array([out[i].iterations for i in range(N)], dtype=np.uint32)
        return xa.reshape(n, N), status, iters

    def lead_tab(self, x0=None, init_fix=True):
        n, N = self.ckt.n, self.N
        x = np.zeros(n * N) if x0 is None else np.asarray(x0, dtype=np.float64).reshape(-1)
        xa, xp = f64(x)
        tab = LaneTab()
        idx = C.POINTER(C.c_uint32)()
        ln = C.c_size_t()
        rc = self.lib.or_lead_tab(self.h, xp, 1 if init_fix else 0, C.byref(tab), C.byref(idx), C.byref(ln))
        if rc != 0:
            return None, None
        arr = np.ctypeslib.as_array(idx, (ln.value,)).copy()
        self.lib.free_(C.cast(idx, C.c_void_p))
        return tab, arr

    def eval_lane(self, l, x):
        n, nnz = self.ckt.n, self.ckt.nnz
        xa, xp = f64(x)
        rhs, rp = f64(np.zeros(n))
        vals, vp = f64(np.zeros(nnz))
        self.lib.or_eval_lane(self.h, l, xp, rp, vp)
        return rhs, vals

    def tran(self, l, opts, x0=None):
        """Lane l's transient: (times, rows (points, n), stats); x0 (n
        values) starts its operating point warm."""
        tp, rp = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
        st = TranStats()
        if x0 is None:
            k = self.lib.or_tran_run(self.h, l, C.byref(opts), C.byref(tp), C.byref(rp), C.byref(st))
        else:
            xa, xp = f64(np.asarray(x0, dtype=np.float64))
            k = self.lib.or_tran_run_from(self.h, l, C.byref(opts), xp, C.byref(tp), C.byref(rp), C.byref(st))
        n = self.ckt.n
        t = np.ctypeslib.as_array(tp, (k,)).copy() if k else np.zeros(0)
        x = np.ctypeslib.as_array(rp, (k * n,)).copy().reshape(k, n) if k else np.zeros((0, n))
        self.lib.free_(C.cast(tp, C.c_void_p))
        self.lib.free_(C.cast(rp, C.c_void_p))
        return t, x, st

    def eval_gc(self, l, x):
        """Lane l's G and C planes at x, as dense n x n matrices."""
        n, nnz = self.ckt.n, self.ckt.nnz
        xa, xp = f64(np.asarray(x, dtype=np.float64))
        g, gp = f64(np.zeros(nnz))
        cc, cp = f64(np.zeros(nnz))
        self.lib.or_eval_lane_gc(self.h, l, xp, gp, cp)
        G, Cm = np.zeros((n, n)), np.zeros((n, n))
        for j in range(n):
            for p in range(self.ckt.col_ptr[j], self.ckt.col_ptr[j + 1]):
                G[self.ckt.row_idx[p], j] += g[p]
                Cm[self.ckt.row_idx[p], j] += cc[p]
        return G, Cm

    def gmin_stamps(self, l=0):
        n = self.ckt.n
        st, sp = u32(np.zeros(2 * n))
        self.lib.or_gmin_stamps(self.h, l, sp)
        return st[:n].copy(), st[n:].copy()

    def newton_hom(self, l, x, gshunt, srcfact, stamps=True, gmodel=None, **kw):
        """Lane l's Newton under a homotopy from x (n values); returns
        (x, status, iterations)."""
        n = self.ckt.n
        if gmodel is not None:
            self.lib.or_lanes_set_model_gmin(self.h, l, gmodel)
        xa, xp = f64(np.array(x, dtype=np.float64))
        sp = None
        if stamps:
            st, sp = u32(np.zeros(2 * n))
            self.lib.or_gmin_stamps(self.h, l, sp)
        o = newton_opts(gmin=gshunt, **kw)
        out = Outcome()
        assert self.lib.or_newton_hom(self.h, l, C.byref(o), xp, C.byref(out), sp, srcfact) == 0
        return xa, out.status, out.iterations

    def op_ladder(self, x0=None, gmin_start=1e-2, ladder_max_iter=100, **kw):
        """The plain rung then the stepping rungs per failing lane; returns
        (x (n, N), status, iterations, rung)."""
        n, N = self.ckt.n, self.N
        x = np.zeros(n * N) if x0 is None else np.array(x0, dtype=np.float64).reshape(-1)
        xa, xp = f64(x)
        out = (Outcome * N)()
        rung, rp = u32(np.zeros(N))
        o = newton_opts(**kw)
        assert self.lib.or_op_ladder(self.h, C.byref(o), gmin_start, ladder_max_iter, xp, out, rp) == 0
        status = np.array([out[i].status for i in range(N)], dtype=np.uint32)
        iters = np.array([out[i].iterations for i in range(N)], dtype=np.uint32)
        return xa.reshape(n, N), status, iters, rung

    def dc_sweep_lanes(self, kind, inst, values, gmin_start=1e-2, ladder_max_iter=100, **kw):
        """The `.dc` sweep over every lane: (x (points, n, N), converged
        (points, N))."""
        n, N = self.ckt.n, self.N
        v, vp = f64(np.asarray(values, dtype=np.float64))
        xo, xp = f64(np.zeros(v.size * n * N))
        cv, cp = u32(np.zeros(v.size * N))
        o = newton_opts(**kw)
        assert self.lib.or_dc_sweep_lanes(self.h, kind, inst, vp, v.size, C.byref(o), gmin_start, ladder_max_iter,
                                          xp, cp) == 0
        return xo.reshape(v.size, n, N), cv.reshape(v.size, N)

    def dc_sweep(self, vsrc, start, step, npts, **kw):
        n = self.ckt.n
        xo, xp = f64(np.zeros(npts * n))
        conv, cp = u32(np.zeros(npts))
        o = newton_opts(**kw)
        assert self.lib.or_dc_sweep(self.h, vsrc, start, step, npts, C.byref(o), xp, cp) == 0
        return xo.reshape(npts, n), conv


def va_params(lib, kind, **values):
    """{slot: value} of named parameters of a compiled kind."""
    out = {}
    for k, v in values.items():
        i = lib.or_va_param_index(kind, k.encode())
        assert i >= 0, k
        out[i] = v
    return out
