"""MI355X lane engine for ESPice sweeps: Python binding of the C ABI.

The engine is the native library ``libespice_lanes.so`` (HIP kernels for
gfx950 plus a C++ host), declared in ``include/espice_lanes.h``. This module
only marshals arguments through ctypes; every solve runs on the GPU, and
importing fails loudly when the library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libespice_lanes.so")

VSOURCE, ISOURCE, RESISTOR, CAPACITOR, DIODE, MOS1, PSP103, BSIM4 = range(8)
# user parameter counts per kind (include/espice_lanes.h kinds); a compiled
# kind's count comes from the library (param_count)
PARAM_COUNT = {VSOURCE: 71, ISOURCE: 71, RESISTOR: 1, CAPACITOR: 1, DIODE: 19, MOS1: 36}
COMPILED_KINDS = (PSP103, BSIM4)
VSOURCE_DC = 0
DIODE_PARAMS = dict(IS=0, N=1, CJO=2, VJ=3, M=4, FC=5, TT=6, GMIN=7, AREA=8, TNOM=9, TEMP=10, DTEMP=11, EG=12, XTI=13,
                    RS=14, TRS=15, TRS2=16, TM1=17, TM2=18)
MOS1_PARAMS = {name: i for i, name in enumerate(
    "TYPE VTO KP GAMMA PHI LAMBDA TOX LD U0 NSUB NSS TPG IS JS CBD CBS CJ CJSW PB MJ MJSW FC CGSO CGDO CGBO "
    "TNOM GMIN TEMP DTEMP L W AD AS PD PS M".split())}

OK = 0
EXPORTS = (
    "espice_lanes_error_message", "espice_lanes_template_create", "espice_lanes_template_destroy",
    "espice_lanes_template_node", "espice_lanes_template_add", "espice_lanes_template_freeze",
    "espice_lanes_template_info", "espice_lanes_template_pattern", "espice_lanes_create", "espice_lanes_destroy",
    "espice_lanes_reset", "espice_lanes_set_param", "espice_lanes_prepare", "espice_lanes_upload_x", "espice_lanes_download_x",
    "espice_lanes_newton_op", "espice_lanes_tran", "espice_lanes_tab_build", "espice_lanes_tab_refactor_solve",
    "espice_lanes_mc_draw", "espice_mlp_create", "espice_mlp_destroy", "espice_mlp_param_count",
    "espice_mlp_set_scaling", "espice_mlp_get_params", "espice_mlp_set_params", "espice_mlp_train",
    "espice_mlp_gradients", "espice_mlp_predict", "espice_lanes_warm_start", "espice_lanes_newton_learned",
    "espice_lanes_first_factors", "espice_lanes_param_count", "espice_lanes_param_index",
    "espice_lanes_tab_refactor_solve_device", "espice_lanes_template_gmin_stamps", "espice_lanes_shape",
    "espice_lanes_newton_homotopy", "espice_lanes_op_ladder", "espice_lanes_set_source", "espice_lanes_dc_sweep",
    "espice_lanes_newton_mixed", "espice_lanes_template_ac_drive", "espice_lanes_ac",
)


class EngineError(RuntimeError):
    pass


class NewtonOpts(C.Structure):
    _fields_ = [("reltol", C.c_double), ("abstol", C.c_double), ("vntol", C.c_double),
                ("residual_tol", C.c_double), ("gmin", C.c_double),
                ("max_iter", C.c_uint32), ("init_fix", C.c_uint32)]


class LaneResult(C.Structure):
    _fields_ = [("status", C.c_uint32), ("iterations", C.c_uint32)]


class LaneHomotopy(C.Structure):
    _fields_ = [("gshunt", C.c_double), ("gmodel", C.c_double), ("srcfact", C.c_double),
                ("flags", C.c_uint32), ("pad", C.c_uint32)]


class LadderOpts(C.Structure):
    _fields_ = [("gmin_start", C.c_double), ("max_iter", C.c_uint32), ("pad", C.c_uint32)]


HOM_ACTIVE, HOM_COLD = 1, 2


class TranOpts(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("t_stop", "dt_init", "dt_max", "dt_min", "xmu", "reltol", "abstol", "vntol",
                                          "chgtol", "trtol")] + \
               [(f, C.c_uint32) for f in ("method", "max_iter", "max_steps", "record_capacity", "op_warm",
                                          "mixed_refine", "early_stop", "pad2")] + \
               [("abort_after", C.c_double), ("abort_mask", C.c_uint64), ("op_reltol", C.c_double)]


class MeasPass(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("row", C.c_uint32), ("rise", C.c_int32), ("fall", C.c_int32),
                ("cross", C.c_int32), ("op", C.c_uint32)] + \
               [(f, C.c_double) for f in ("val", "td", "frm", "to", "at")]


class MeasResult(C.Structure):
    _fields_ = [("value", C.c_double), ("aux0", C.c_double), ("aux1", C.c_double)]


class TranResult(C.Structure):
    _fields_ = [(f, C.c_uint32) for f in ("status", "points", "steps", "attempts", "rej_newton", "rej_lte",
                                          "bp_landings", "nr_iters", "fp32_solves", "fp64_fallbacks")]


MEAS_CROSS, MEAS_AT, MEAS_EXTREMUM, MEAS_INTEG, MEAS_FIND = range(5)
LAST = -2


class LaneTab(C.Structure):
    _fields_ = [(f, C.c_uint32) for f in ("n", "nnz", "nnz_l", "nnz_u")] + \
               [("growth", C.c_double), ("wide", C.c_double)] + \
               [(f, C.c_uint32) for f in ("col_ptr", "prow", "q", "pinv", "lp", "li", "up", "ui", "flags",
                                          "void_slots", "n_void_slots", "pad")]


class MlpShape(C.Structure):
    _fields_ = [(f, C.c_uint32) for f in ("n_in", "hidden", "depth", "n_out")] + [("seed", C.c_uint64)]


class MlpTrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("batch", C.c_uint32)] + \
               [(f, C.c_float) for f in ("lr", "beta1", "beta2", "eps")] + [("seed", C.c_uint64)]


_lib = None


def library():
    """The loaded engine; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EngineError(f"lane engine not built: {LIB_PATH} is missing (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    P, u32p, f64p = C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_double)
    sig = {
        "espice_lanes_error_message": [C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)],
        "espice_lanes_template_create": [C.POINTER(P)],
        "espice_lanes_template_destroy": [P],
        "espice_lanes_template_node": [P, C.c_char_p, u32p],
        "espice_lanes_template_add": [P, C.c_uint32, u32p, f64p, C.c_size_t, u32p],
        "espice_lanes_template_freeze": [P],
        "espice_lanes_template_info": [P, u32p, u32p, u32p],
        "espice_lanes_template_pattern": [P, u32p, u32p, u32p, C.POINTER(C.c_uint8), u32p],
        "espice_lanes_create": [P, C.c_uint32, C.POINTER(P)],
        "espice_lanes_destroy": [P],
        "espice_lanes_reset": [P],
        "espice_lanes_set_param": [P, C.c_uint32, C.c_uint32, C.c_uint32, f64p],
        "espice_lanes_param_count": [C.c_uint32, C.POINTER(C.c_uint32)],
        "espice_lanes_param_index": [C.c_uint32, C.c_char_p, C.POINTER(C.c_uint32)],
        "espice_lanes_tab_refactor_solve_device": [C.POINTER(LaneTab), u32p, C.c_size_t, C.c_uint32, P, P, P, P, P],
        "espice_lanes_prepare": [P],
        "espice_lanes_upload_x": [P, f64p],
        "espice_lanes_download_x": [P, f64p],
        "espice_lanes_newton_op": [P, C.POINTER(NewtonOpts), C.POINTER(LaneResult), f64p],
        "espice_lanes_tran": [P, C.POINTER(TranOpts), C.POINTER(MeasPass), C.c_uint32, C.POINTER(TranResult),
                              C.POINTER(MeasResult), f64p, f64p, f64p],
        "espice_lanes_mc_draw": [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, f64p, C.c_uint32, f64p],
        "espice_lanes_tab_build": [C.c_uint32, u32p, u32p, f64p, C.POINTER(LaneTab), u32p, C.c_size_t,
                                   C.POINTER(C.c_size_t)],
        "espice_lanes_tab_refactor_solve": [C.POINTER(LaneTab), u32p, C.c_size_t, C.c_uint32, f64p, f64p, f64p,
                                            C.POINTER(C.c_uint8), f64p],
        "espice_mlp_create": [C.POINTER(MlpShape), C.POINTER(P)],
        "espice_mlp_destroy": [P],
        "espice_mlp_param_count": [P, C.POINTER(C.c_size_t)],
        "espice_mlp_set_scaling": [P, f64p, f64p, f64p, f64p],
        "espice_mlp_get_params": [P, C.POINTER(C.c_float)],
        "espice_mlp_set_params": [P, C.POINTER(C.c_float)],
        "espice_mlp_train": [P, f64p, f64p, C.c_uint32, C.POINTER(MlpTrainOpts), C.POINTER(C.c_float)],
        "espice_mlp_gradients": [P, f64p, f64p, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)],
        "espice_mlp_predict": [P, f64p, C.c_uint32, f64p],
        "espice_lanes_warm_start": [P, P, f64p, f64p],
        "espice_lanes_newton_learned": [P, P, f64p, C.POINTER(NewtonOpts), C.POINTER(LaneResult), f64p, f64p, f64p],
        "espice_lanes_first_factors": [P, u32p, u32p],
        "espice_lanes_template_gmin_stamps": [P, u32p, u32p],
        "espice_lanes_shape": [P, u32p, u32p],
        "espice_lanes_newton_homotopy": [P, C.POINTER(NewtonOpts), C.POINTER(LaneHomotopy), f64p,
                                         C.POINTER(LaneResult), f64p],
        "espice_lanes_op_ladder": [P, C.POINTER(NewtonOpts), C.POINTER(LadderOpts), C.POINTER(LaneResult), u32p,
                                   f64p],
        "espice_lanes_set_source": [P, C.c_uint32, C.c_uint32, C.c_double],
        "espice_lanes_newton_mixed": [P, C.POINTER(NewtonOpts), C.c_uint32, C.POINTER(LaneResult), u32p, f64p],
        "espice_lanes_template_ac_drive": [P, C.c_uint32, C.c_uint32, C.c_double, C.c_double, f64p],
        "espice_lanes_ac": [P, f64p, C.c_uint32, f64p, f64p, f64p],
        "espice_lanes_dc_sweep": [P, C.c_uint32, C.c_uint32, f64p, C.c_uint32, C.POINTER(NewtonOpts),
                                  C.POINTER(LadderOpts), f64p, u32p, f64p],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.restype = C.c_uint32
        f.argtypes = args
    lib.espice_lanes_site_key.restype = C.c_uint64
    lib.espice_lanes_site_key.argtypes = [C.c_uint64, C.c_uint64]
    _lib = lib
    return lib


def _check(status):
    if status != OK:
        buf = C.create_string_buffer(256)
        req = C.c_size_t()
        library().espice_lanes_error_message(buf, 256, C.byref(req))
        raise EngineError(f"espice status {status}: {buf.value.decode()}")


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint32))


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def newton_opts(init_fix=True, max_iter=100, reltol=1e-3, abstol=1e-12, vntol=1e-6, residual_tol=1e-9, gmin=0.0):
    """converger.Options defaults (src/core/numerics.zig:378-399)."""
    return NewtonOpts(reltol, abstol, vntol, residual_tol, gmin, max_iter, 1 if init_fix else 0)


def tran_opts(t_stop, dt_init, dt_max=None, method=1, xmu=0.5, dt_min=1e-18, reltol=1e-3, abstol=1e-12, vntol=1e-6,
              chgtol=1e-14, trtol=7.0, max_iter=100, max_steps=1_000_000_000, record_capacity=0, op_warm=False,
              early_stop=False, abort_after=float("nan"), abort_mask=(1 << 64) - 1, mixed_refine=0, op_reltol=0.0):
    """`.tran dt_init t_stop`; a deck's step ceiling is min(dt_init, t_stop/50)
    (src/frontend/analyses.zig:565-566, src/core/query.zig:673-680), the
    tolerances are converger.Options' (src/core/numerics.zig:378-399).
    early_stop / abort_after / abort_mask: the points options' measure
    handling (a lane ends once its passes stop; past abort_after, a running
    pass of the mask aborts it with status 4). mixed_refine k > 0: every
    Newton tick's linear step in FP32 with FP64 refinement, at most k FP32
    solves per tick."""
    if dt_max is None:
        dt_max = min(dt_init, t_stop / 50)
    return TranOpts(t_stop, dt_init, dt_max, dt_min, xmu, reltol, abstol, vntol, chgtol, trtol, method, max_iter,
                    max_steps, record_capacity, 1 if op_warm else 0, mixed_refine, 1 if early_stop else 0, 0,
                    abort_after, abort_mask, op_reltol)


def meas_cross(row, val, rise=-1, fall=-1, cross=-1, td=0.0, frm=0.0, to=0.0):
    """WHEN / TRIG / TARG on unknown `row` (src/output/scan.zig Cross)."""
    return MeasPass(MEAS_CROSS, row, rise, fall, cross, 0, val, td, frm, to, 0.0)


def meas_at(row, at, deriv=False):
    """FIND / DERIV ... AT= (src/output/scan.zig At)."""
    return MeasPass(MEAS_AT, row, -1, -1, -1, 1 if deriv else 0, 0.0, 0.0, 0.0, 0.0, at)


def meas_find(row, deriv=False):
    """FIND / DERIV ... WHEN on unknown `row`: must follow its WHEN's
    meas_cross pass (src/output/scan.zig Find)."""
    return MeasPass(MEAS_FIND, row, -1, -1, -1, 1 if deriv else 0, 0.0, 0.0, 0.0, 0.0, 0.0)


def meas_extremum(row, op, frm=0.0, to=0.0):
    """MIN (op 0), MAX (1), AVG (2) (src/output/scan.zig Extremum)."""
    return MeasPass(MEAS_EXTREMUM, row, -1, -1, -1, op, 0.0, 0.0, frm, to, 0.0)


def meas_integ(row, rms=False, frm=0.0, to=0.0):
    """INTEG / RMS (src/output/scan.zig Integ)."""
    return MeasPass(MEAS_INTEG, row, -1, -1, -1, 1 if rms else 0, 0.0, 0.0, frm, to, 0.0)


def param_count(kind):
    """User parameters of a kind (a compiled kind: its module's)."""
    if kind in PARAM_COUNT:
        return PARAM_COUNT[kind]
    c = C.c_uint32()
    _check(library().espice_lanes_param_count(kind, C.byref(c)))
    PARAM_COUNT[kind] = c.value
    return c.value


def param_index(kind, name):
    """The slot of a named parameter of a compiled kind (aliases resolved)."""
    i = C.c_uint32()
    _check(library().espice_lanes_param_index(kind, name.encode(), C.byref(i)))
    return i.value


def named_params(kind, **values):
    """{slot: value} of named parameters of a compiled kind."""
    return {param_index(kind, k): v for k, v in values.items()}


class Template:
    """A circuit frozen into one MNA pattern, shared by every lane."""

    def __init__(self):
        self.lib = library()
        self.h = C.c_void_p()
        _check(self.lib.espice_lanes_template_create(C.byref(self.h)))
        self.rows = {}

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.espice_lanes_template_destroy(self.h)
            self.h = None

    def node(self, name):
        """The row of node `name`, numbered on first touch; a frozen
        template answers only for names it already holds."""
        if name in self.rows:
            return self.rows[name]
        row = C.c_uint32()
        _check(self.lib.espice_lanes_template_node(self.h, name.encode(), C.byref(row)))
        self.rows[name] = row.value
        return row.value

    def add(self, kind, ports, params=None):
        rows = [self.node(p) if isinstance(p, str) else p for p in ports]
        if kind not in PARAM_COUNT and kind not in COMPILED_KINDS:
            raise EngineError(f"unknown device kind {kind}")
        pr = np.full(param_count(kind), np.nan)
        for k, v in (params or {}).items():
            if not 0 <= k < len(pr):
                raise EngineError(f"parameter {k} out of range for kind {kind} ({len(pr)} parameters)")
            pr[k] = v
        _, rp = _u32(rows)
        a, ap = _f64(pr)
        inst = C.c_uint32()
        _check(self.lib.espice_lanes_template_add(self.h, kind, rp, ap, len(pr), C.byref(inst)))
        return inst.value

    def freeze(self):
        _check(self.lib.espice_lanes_template_freeze(self.h))
        n, nnz, tri = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(self.lib.espice_lanes_template_info(self.h, C.byref(n), C.byref(nnz), C.byref(tri)))
        self.n, self.nnz, self.tridiagonal = n.value, nnz.value, bool(tri.value)
        cp, cpp = _u32(np.zeros(self.n + 1))
        ri, rip = _u32(np.zeros(self.nnz))
        dg, dgp = _u32(np.zeros(self.n))
        cr = np.zeros(self.n, dtype=np.uint8)
        q, qp = _u32(np.zeros(self.n))
        _check(self.lib.espice_lanes_template_pattern(self.h, cpp, rip, dgp, cr.ctypes.data_as(C.POINTER(C.c_uint8)), qp))
        self.col_ptr, self.row_idx, self.diag, self.cur_row, self.q = cp, ri, dg, cr, q
        return self

    def ac_drive(self, sources):
        """The AC excitation (2n: real parts, then imaginary) of sources =
        [(kind, instance, mag, phase_deg)]."""
        d, dp = _f64(np.zeros(2 * self.n))
        for kind, inst, mag, phase in sources:
            _check(self.lib.espice_lanes_template_ac_drive(self.h, kind, inst, mag, phase, dp))
        return d

    def gmin_stamps(self):
        """(col, slot) per row: where the operating point's stepping rungs
        load their diagonal gmin (NONE where a row takes none)."""
        col, colp = _u32(np.zeros(self.n))
        slot, slotp = _u32(np.zeros(self.n))
        _check(self.lib.espice_lanes_template_gmin_stamps(self.h, colp, slotp))
        return col, slot


class Lanes:
    """N lanes of one template, resident on the GPU."""

    def __init__(self, template, n_lanes):
        self.t = template
        self.lib = template.lib
        self.N = n_lanes
        self.h = C.c_void_p()
        _check(self.lib.espice_lanes_create(template.h, n_lanes, C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.espice_lanes_destroy(self.h)
            self.h = None

    def set_param(self, kind, inst, k, values):
        v, vp = _f64(np.broadcast_to(np.asarray(values, dtype=np.float64), (self.N,)))
        _check(self.lib.espice_lanes_set_param(self.h, kind, inst, k, vp))

    def prepare(self):
        _check(self.lib.espice_lanes_prepare(self.h))

    def reset(self):
        _check(self.lib.espice_lanes_reset(self.h))

    def upload_x(self, x):
        a, ap = _f64(np.asarray(x, dtype=np.float64).reshape(-1))
        assert a.size == self.t.n * self.N
        _check(self.lib.espice_lanes_upload_x(self.h, ap))

    def download_x(self):
        a, ap = _f64(np.zeros(self.t.n * self.N))
        _check(self.lib.espice_lanes_download_x(self.h, ap))
        return a.reshape(self.t.n, self.N)

    def newton_op(self, x0=None, **kw):
        """Static Newton on every lane; returns (x, status, iterations, kernel_ms)."""
        self.upload_x(np.zeros(self.t.n * self.N) if x0 is None else x0)
        return self.newton(**kw)

    def warm_start(self, mlp, params):
        """The network's start point for params[N][n_in] written into the
        lanes' x on the GPU; returns the GPU milliseconds it took."""
        a, ap = _f64(np.asarray(params, dtype=np.float64).reshape(self.N, -1))
        ms = C.c_double()
        _check(self.lib.espice_lanes_warm_start(self.h, mlp.h, ap, C.byref(ms)))
        return ms.value

    def newton(self, **kw):
        """Static Newton from the lanes' current x (no upload)."""
        o = newton_opts(**kw)
        res = np.zeros((self.N, 2), dtype=np.uint32)  # espice_lane_result: status, iterations
        ms = C.c_double()
        _check(self.lib.espice_lanes_newton_op(self.h, C.byref(o), res.ctypes.data_as(C.POINTER(LaneResult)),
                                               C.byref(ms)))
        return self.download_x(), res[:, 0].copy(), res[:, 1].copy(), ms.value

    def newton_homotopy(self, gshunt, gmodel, srcfact, flags, **kw):
        """Static Newton of the active lanes under per-lane homotopy controls
        (arrays or scalars) from the lanes' current x; returns
        (x, status, iterations, kernel_ms)."""
        h = np.zeros(self.N, dtype=[("gshunt", "f8"), ("gmodel", "f8"), ("srcfact", "f8"), ("flags", "u4"),
                                    ("pad", "u4")])
        h["gshunt"], h["gmodel"], h["srcfact"], h["flags"] = gshunt, gmodel, srcfact, flags
        o = newton_opts(**kw)
        res = np.zeros((self.N, 2), dtype=np.uint32)
        ms = C.c_double()
        _check(self.lib.espice_lanes_newton_homotopy(self.h, C.byref(o), h.ctypes.data_as(C.POINTER(LaneHomotopy)),
                                                     None, res.ctypes.data_as(C.POINTER(LaneResult)), C.byref(ms)))
        return self.download_x(), res[:, 0].copy(), res[:, 1].copy(), ms.value

    def op_ladder(self, x0=None, gmin_start=1e-2, ladder_max_iter=100, **kw):
        """Every lane's operating point by the plain rung and, where it
        fails, the stepping rungs; returns (x, status, iterations, rung,
        kernel_ms)."""
        self.upload_x(np.zeros(self.t.n * self.N) if x0 is None else x0)
        o = newton_opts(**kw)
        lo = LadderOpts(gmin_start, ladder_max_iter, 0)
        res = np.zeros((self.N, 2), dtype=np.uint32)
        rung, rp = _u32(np.zeros(self.N))
        ms = C.c_double()
        _check(self.lib.espice_lanes_op_ladder(self.h, C.byref(o), C.byref(lo),
                                               res.ctypes.data_as(C.POINTER(LaneResult)), rp, C.byref(ms)))
        return self.download_x(), res[:, 0].copy(), res[:, 1].copy(), rung, ms.value

    def ac(self, omega, drive):
        """Every lane's small-signal response at its x: complex array
        (frequencies, n, N) and the kernel's milliseconds."""
        w, wp = _f64(np.asarray(omega, dtype=np.float64))
        d, dp = _f64(np.asarray(drive, dtype=np.float64))
        assert d.size == 2 * self.t.n
        out, op = _f64(np.zeros(w.size * 2 * self.t.n * self.N))
        ms = C.c_double()
        _check(self.lib.espice_lanes_ac(self.h, wp, w.size, dp, op, C.byref(ms)))
        o = out.reshape(w.size, 2, self.t.n, self.N)
        return o[:, 0] + 1j * o[:, 1], ms.value

    def newton_mixed(self, max_refine=10, **kw):
        """Static Newton from the lanes' current x with FP32 factors and
        FP64 refinement; returns (x, status, iterations, solves, fallbacks,
        kernel_ms)."""
        o = newton_opts(**kw)
        res = np.zeros((self.N, 2), dtype=np.uint32)
        ref, rp = _u32(np.zeros(2 * self.N))
        ms = C.c_double()
        _check(self.lib.espice_lanes_newton_mixed(self.h, C.byref(o), max_refine,
                                                  res.ctypes.data_as(C.POINTER(LaneResult)), rp, C.byref(ms)))
        return self.download_x(), res[:, 0].copy(), res[:, 1].copy(), ref[0::2].copy(), ref[1::2].copy(), ms.value

    def dc_sweep(self, kind, inst, values, gmin_start=1e-2, ladder_max_iter=100, **kw):
        """A `.dc` sweep of one source's dc value over every lane; returns
        (x (points, n, N), converged (points, N), kernel_ms)."""
        v, vp = _f64(np.asarray(values, dtype=np.float64))
        npts = v.size
        o = newton_opts(**kw)
        lo = LadderOpts(gmin_start, ladder_max_iter, 0)
        xo, xp = _f64(np.zeros(npts * self.t.n * self.N))
        cv, cp = _u32(np.zeros(npts * self.N))
        ms = C.c_double()
        _check(self.lib.espice_lanes_dc_sweep(self.h, kind, inst, vp, npts, C.byref(o), C.byref(lo), xp, cp,
                                              C.byref(ms)))
        return xo.reshape(npts, self.t.n, self.N), cv.reshape(npts, self.N), ms.value

    def first_factors(self):
        """(predicted, serial): how the last Newton launch factored each
        lane's first matrix."""
        p, q = C.c_uint32(), C.c_uint32()
        _check(self.lib.espice_lanes_first_factors(self.h, C.byref(p), C.byref(q)))
        return p.value, q.value

    def newton_learned(self, mlp, params, **kw):
        """The network's start for params[N][n_in], then static Newton from
        it, in one call. Returns (x, status, iterations, warm_ms, newton_ms).
        The argument buffers and their ctypes pointers are kept per lanes
        object, so a repeated solve marshals nothing new."""
        p = np.asarray(params, dtype=np.float64).reshape(self.N, -1)
        key = (p.shape, tuple(sorted(kw.items())))
        buf = getattr(self, "_learned", None)
        if buf is None or buf["key"] != key:
            o = newton_opts(**kw)
            pa = np.zeros(p.shape)
            res = np.zeros((self.N, 2), dtype=np.uint32)
            x = np.zeros(self.t.n * self.N)
            wm, nm = C.c_double(), C.c_double()
            buf = self._learned = {
                "key": key, "p": pa, "res": res, "x": x, "wm": wm, "nm": nm,
                "args": (pa.ctypes.data_as(C.POINTER(C.c_double)), C.byref(o),
                         res.ctypes.data_as(C.POINTER(LaneResult)), x.ctypes.data_as(C.POINTER(C.c_double)),
                         C.byref(wm), C.byref(nm)),
                "opts": o,
            }
        np.copyto(buf["p"], p)
        _check(self.lib.espice_lanes_newton_learned(self.h, mlp.h, *buf["args"]))
        res = buf["res"]
        return (buf["x"].reshape(self.t.n, self.N).copy(), res[:, 0].copy(), res[:, 1].copy(), buf["wm"].value,
                buf["nm"].value)


    def tran(self, opts, passes=(), record_capacity=None):
        """Every lane's transient (its operating point cold, or warm at the
        lanes' x with opts.op_warm).

        Returns (results (N,) TranResult fields as a dict of arrays, meas
        (N, P, 3), times (N, cap), rows (N, cap, n), kernel_ms); times/rows
        are None when nothing is recorded."""
        if record_capacity is not None:
            opts.record_capacity = record_capacity
        cap, n, N, P = opts.record_capacity, self.t.n, self.N, len(passes)
        pa = (MeasPass * max(P, 1))(*passes)
        out = (TranResult * N)()
        meas, mp = _f64(np.zeros(max(N * P * 3, 1)))
        times, tp = _f64(np.zeros(max(N * cap, 1)))
        rows, rp = _f64(np.zeros(max(N * cap * n, 1)))
        ms = C.c_double()
        _check(self.lib.espice_lanes_tran(self.h, C.byref(opts), pa, P, out,
                                          C.cast(mp, C.POINTER(MeasResult)), tp, rp, C.byref(ms)))
        res = {f: np.array([getattr(out[i], f) for i in range(N)], dtype=np.uint32) for f, _ in TranResult._fields_}
        m = meas[:N * P * 3].reshape(N, P, 3)
        if not cap:
            return res, m, None, None, ms.value
        return res, m, times.reshape(N, cap), rows.reshape(N, cap, n), ms.value


# ---- the points API (include/espice.h)

ESPICE_FILE, ESPICE_BYTES = 0, 1
ESPICE_NGSPICE, ESPICE_HSPICE, ESPICE_SPECTRE = 0, 1, 2
ESPICE_CPU, ESPICE_AUTO, ESPICE_CUDA, ESPICE_HIP = 0, 1, 2, 3
REASONS = {0: "OK", 1: "NO_CONVERGENCE", 2: "TIMESTEP_TOO_SMALL", 3: "SINGULAR_MATRIX", 4: "MEASURE_NOT_FOUND",
           5: "MEASURE_UNRESOLVED", 6: "ABORTED", 7: "MODEL_BIN_NOT_FOUND", 8: "TOPOLOGY_CHANGED",
           9: "INVALID_PARAMETER", 10: "DEVICE_REFUSED", 255: "OTHER"}


class EspiceBytes(C.Structure):
    _fields_ = [("data", C.c_char_p), ("len", C.c_size_t)]


def _bytes(b):
    b = b.encode() if isinstance(b, str) else b
    return EspiceBytes(b, len(b))


class CreateOptions(C.Structure):
    _fields_ = [(f, C.c_uint32) for f in ("abi_version", "struct_size", "source_kind", "dialect", "backend",
                                          "explicit_gpu", "output_format", "max_parallel")] + \
               [(f, EspiceBytes) for f in ("source", "origin", "output_path", "zig_path", "share_path")] + \
               [("live_params", C.POINTER(EspiceBytes)), ("live_param_count", C.c_size_t),
                ("display", C.c_void_p), ("display_user", C.c_void_p)]


class PointsOptions(C.Structure):
    _fields_ = [("struct_size", C.c_size_t), ("abort_measure", EspiceBytes), ("abort_after", C.c_double),
                ("measures_only", C.c_uint8), ("early_stop", C.c_uint8), ("warm_start", C.c_uint8),
                ("tangent_step", C.c_double)]


def points_options(abort_measure=None, abort_after=float("nan"), measures_only=False, early_stop=False,
                   warm_start=False, tangent_step=0.0):
    """espice_points_options (include/espice.h) for Problem.append_points."""
    o = PointsOptions(C.sizeof(PointsOptions), _bytes(abort_measure or ""), abort_after, int(measures_only),
                      int(early_stop), int(warm_start), tangent_step)
    o._keep = abort_measure
    return o


def _points_lib():
    lib = library()
    if getattr(lib, "_points_bound", False):
        return lib
    P = C.c_void_p
    sig = {
        "espice_abi_version": (C.c_uint32, []),
        "espice_default_options": (None, [C.POINTER(CreateOptions)]),
        "espice_create": (C.c_uint32, [C.POINTER(CreateOptions), C.POINTER(P), C.c_char_p, C.c_size_t]),
        "espice_destroy": (C.c_uint32, [P]),
        "espice_query_count": (C.c_uint32, [P, C.POINTER(C.c_uint32)]),
        "espice_run_all": (C.c_uint32, [P]),
        "espice_error_message": (C.c_uint32, [P, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "espice_append_points": (C.c_uint32, [P, C.POINTER(C.c_double), C.c_size_t, C.POINTER(PointsOptions),
                                              C.POINTER(C.c_uint32)]),
        "espice_copy_point_measures": (C.c_uint32, [P, C.c_uint32, C.c_size_t, C.POINTER(EspiceBytes), C.c_size_t,
                                                    C.POINTER(C.c_double), C.POINTER(C.c_uint32)]),
        "espice_release_points": (C.c_uint32, [P, C.c_uint32, C.c_size_t]),
        "espice_lane_width": (C.c_uint32, [P, C.POINTER(C.c_uint32)]),
        "espice_cli_report": (C.c_uint32, [P, C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "espice_spice_number": (C.c_int, [C.c_char_p, C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib._points_bound = True
    return lib


def spice_number(text):
    """ngspice's INPevaluate of one number with its scale suffix, or None."""
    v = C.c_double()
    return v.value if _points_lib().espice_spice_number(text.encode(), C.byref(v)) else None


class PointsError(EngineError):
    def __init__(self, status, name):
        super().__init__(f"espice status {status}: {name}")
        self.status = status
        self.name = name


def create_options(text=None, path=None, live=(), backend=ESPICE_HIP, dialect=ESPICE_NGSPICE):
    o = CreateOptions()
    _points_lib().espice_default_options(C.byref(o))
    o.backend = backend
    o.explicit_gpu = 1 if backend == ESPICE_HIP else 0
    o.dialect = dialect
    if path is not None:
        o.source_kind, o.source = ESPICE_FILE, _bytes(path)
    else:
        o.source_kind, o.source = ESPICE_BYTES, _bytes(text)
    names = (EspiceBytes * max(len(live), 1))(*[_bytes(n) for n in live])
    o.live_params = C.cast(names, C.POINTER(EspiceBytes))
    o.live_param_count = len(live)
    o._keep = (names, text, path)
    return o


class Problem:
    """One deck on the lane engine through the points API (include/espice.h)."""

    def __init__(self, text=None, path=None, live=(), backend=ESPICE_HIP, options=None):
        self.lib = _points_lib()
        self.h = C.c_void_p()
        o = options or create_options(text, path, live, backend)
        diag = C.create_string_buffer(256)
        st = self.lib.espice_create(C.byref(o), C.byref(self.h), diag, 256)
        if st != OK:
            self.h = None
            raise PointsError(st, diag.value.decode())
        self.live = list(live)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.espice_destroy(self.h)
            self.h = None

    def _check(self, st):
        if st != OK:
            buf = C.create_string_buffer(256)
            req = C.c_size_t()
            self.lib.espice_error_message(self.h, buf, 256, C.byref(req))
            raise PointsError(st, buf.value.decode())

    def query_count(self):
        n = C.c_uint32()
        self._check(self.lib.espice_query_count(self.h, C.byref(n)))
        return n.value

    def append_points(self, rows, options=None):
        rows = np.ascontiguousarray(np.atleast_2d(np.asarray(rows, dtype=np.float64)))
        first = C.c_uint32()
        self._check(self.lib.espice_append_points(self.h, rows.ctypes.data_as(C.POINTER(C.c_double)), rows.shape[0],
                                                  C.byref(options) if options is not None else None,
                                                  C.byref(first)))
        return first.value

    def run_all(self):
        self._check(self.lib.espice_run_all(self.h))

    def measures(self, first_id, count, names):
        out = np.zeros((count, len(names)))
        reasons = np.zeros(count, dtype=np.uint32)
        nb = (EspiceBytes * max(len(names), 1))(*[_bytes(n) for n in names])
        self._check(self.lib.espice_copy_point_measures(self.h, first_id, count, nb, len(names),
                                                        out.ctypes.data_as(C.POINTER(C.c_double)),
                                                        reasons.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out, reasons

    def release(self, first_id, count):
        self._check(self.lib.espice_release_points(self.h, first_id, count))

    def lane_width(self):
        n = C.c_uint32()
        self._check(self.lib.espice_lane_width(self.h, C.byref(n)))
        return n.value

    def report(self, json=False, tables=False, plot_format=None):
        """Point 0's .meas lines as the CLI prints them; json: its measures
        as JSON; tables: the run's plots in SPICE3 print format (point 0's
        AC response, a `.dc monte` deck's Monte Carlo table); plot_format
        "csv", "ascii" or "binary": those plots in that format (bytes for
        binary)."""
        mode = {"csv": 3, "ascii": 4, "binary": 5}[plot_format] if plot_format else 2 if tables else 1 if json else 0
        req = C.c_size_t()
        self.lib.espice_cli_report(self.h, mode, None, 0, C.byref(req))
        buf = C.create_string_buffer(req.value)
        self._check(self.lib.espice_cli_report(self.h, mode, buf, req.value, C.byref(req)))
        raw = buf.raw[:req.value - 1]
        return raw if plot_format == "binary" else raw.decode()


MC_AGAUSS, MC_GAUSS, MC_AUNIF, MC_UNIF, MC_LIMIT = range(5)


def site_key(owner, salt):
    """netlist.siteKey: owner mixed with salt (splitmix64's finalizer)."""
    return library().espice_lanes_site_key(owner & (2**64 - 1), salt & (2**64 - 1))


def mc_draw(seed, site, first_trial, count, fn, *args):
    """Trials first_trial .. first_trial + count - 1 of distribution fn at
    site: one value per lane (agauss(nom, abs, sigmas[, mult]) and kin)."""
    a, ap = _f64(np.asarray(args, dtype=np.float64))
    out, op = _f64(np.zeros(count))
    _check(library().espice_lanes_mc_draw(seed & (2**64 - 1), site & (2**64 - 1), first_trial, count, fn, ap,
                                          len(a), op))
    return out


def tab_build(n, col_ptr, row_idx, vals):
    """LaneTab and index words of one host factor (ownTape/buildTab)."""
    lib = library()
    cp, cpp = _u32(col_ptr)
    ri, rip = _u32(row_idx)
    v, vp = _f64(vals)
    tab = LaneTab()
    ln = C.c_size_t()
    lib.espice_lanes_tab_build(n, cpp, rip, vp, C.byref(tab), None, 0, C.byref(ln))
    idx, ip = _u32(np.zeros(ln.value))
    _check(lib.espice_lanes_tab_build(n, cpp, rip, vp, C.byref(tab), ip, ln.value, C.byref(ln)))
    return tab, idx


def tab_refactor_solve_device(tab, idx, n_lanes, a_ptr, rhs_ptr, dx_ptr, fail_ptr, stream=None):
    """The allocation-free replay: a, rhs, dx and fail are device addresses
    (ints); enqueued on `stream` (a hipStream_t address or None); the caller
    synchronizes before reading dx and fail."""
    lib = library()
    ii, ip = _u32(idx)
    V = C.c_void_p
    _check(lib.espice_lanes_tab_refactor_solve_device(C.byref(tab), ip, len(ii), n_lanes, V(a_ptr), V(rhs_ptr),
                                                      V(dx_ptr), V(fail_ptr), V(stream)))


def tab_refactor_solve(tab, idx, a, rhs):
    """Lead-tape replay on the GPU: a (nnz, N) and rhs (n, N) lane-major."""
    lib = library()
    a = np.ascontiguousarray(a, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    N = a.shape[1]
    ii, ip = _u32(idx)
    aa, ap = _f64(a)
    rr, rp = _f64(rhs)
    dx, dxp = _f64(np.zeros_like(rhs))
    fail = np.zeros(N, dtype=np.uint8)
    ms = C.c_double()
    _check(lib.espice_lanes_tab_refactor_solve(C.byref(tab), ip, len(ii), N, ap, rp, dxp,
                                               fail.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(ms)))
    return dx.reshape(rhs.shape), fail, ms.value


class Mlp:
    """The warm-start network (include/espice_mlp.h): n_in -> depth x hidden
    tanh -> n_out, BF16 MFMA kernels with FP32 master weights."""

    def __init__(self, n_in, n_out, hidden=256, depth=3, seed=1):
        self.lib = library()
        self.n_in, self.n_out, self.hidden, self.depth = n_in, n_out, hidden, depth
        h = C.c_void_p()
        _check(self.lib.espice_mlp_create(C.byref(MlpShape(n_in, hidden, depth, n_out, seed)), C.byref(h)))
        self.h = h
        n = C.c_size_t()
        _check(self.lib.espice_mlp_param_count(self.h, C.byref(n)))
        self.param_count = n.value

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.espice_mlp_destroy(self.h)
            self.h = None

    def set_scaling(self, in_mean, in_std, out_mean, out_std):
        arrs = [_f64(np.asarray(a, dtype=np.float64).reshape(-1)) for a in (in_mean, in_std, out_mean, out_std)]
        assert arrs[0][0].size == arrs[1][0].size == self.n_in and arrs[2][0].size == arrs[3][0].size == self.n_out
        _check(self.lib.espice_mlp_set_scaling(self.h, *[p for _, p in arrs]))

    def fit_scaling(self, x, y):
        """Mean and standard deviation of the training set's columns."""
        x, y = np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64)
        self.set_scaling(x.mean(0), x.std(0), y.mean(0), y.std(0))

    def params(self):
        out = np.zeros(self.param_count, dtype=np.float32)
        _check(self.lib.espice_mlp_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.param_count
        _check(self.lib.espice_mlp_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def train(self, x, y, epochs=100, batch=256, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, seed=7):
        xa, xp = _f64(np.asarray(x, dtype=np.float64).reshape(-1, self.n_in))
        ya, yp = _f64(np.asarray(y, dtype=np.float64).reshape(-1, self.n_out))
        assert xa.shape[0] == ya.shape[0]
        loss = np.zeros(epochs, dtype=np.float32)
        o = MlpTrainOpts(epochs, batch, lr, beta1, beta2, eps, seed)
        _check(self.lib.espice_mlp_train(self.h, xp, yp, xa.shape[0], C.byref(o),
                                         loss.ctypes.data_as(C.POINTER(C.c_float))))
        return loss

    def gradients(self, x, y):
        """(grads in packed order, loss) of one pass over x, y."""
        xa, xp = _f64(np.asarray(x, dtype=np.float64).reshape(-1, self.n_in))
        ya, yp = _f64(np.asarray(y, dtype=np.float64).reshape(-1, self.n_out))
        g = np.zeros(self.param_count, dtype=np.float32)
        loss = C.c_float()
        _check(self.lib.espice_mlp_gradients(self.h, xp, yp, xa.shape[0], g.ctypes.data_as(C.POINTER(C.c_float)),
                                             C.byref(loss)))
        return g, loss.value

    def predict(self, x):
        xa, xp = _f64(np.asarray(x, dtype=np.float64).reshape(-1, self.n_in))
        y, yp = _f64(np.zeros((xa.shape[0], self.n_out)))
        _check(self.lib.espice_mlp_predict(self.h, xp, xa.shape[0], yp))
        return y
