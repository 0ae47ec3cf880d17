"""ctypes binding of tests/libhostlane.so: the engine's per-lane bodies run
on the host (test infrastructure)."""
import ctypes as C

import numpy as np

from oracle_lib import NewtonOpts, f64, u32


def load(path):
    hl = C.CDLL(path)
    P, u32p, f64p = C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_double)
    for name, res, args in [
        ("hl_template_new", P, []), ("hl_template_free", None, [P]), ("hl_node", C.c_uint32, [P, C.c_char_p]),
        ("hl_add", C.c_int, [P, C.c_int, u32p, f64p]), ("hl_freeze", C.c_int, [P]),
        ("hl_n", C.c_uint32, [P]), ("hl_nnz", C.c_uint32, [P]), ("hl_pattern", None, [P, u32p, u32p, u32p, u32p]),
        ("hl_lanes_new", P, [P, C.c_uint32]), ("hl_lanes_free", None, [P]),
        ("hl_set_param", None, [P, C.c_int, C.c_uint32, C.c_uint32, f64p]), ("hl_prepare", None, [P]),
        ("hl_prm", None, [P, C.c_int, f64p]),
        ("hl_newton_op", None, [P, C.POINTER(NewtonOpts), f64p, u32p, u32p]),
        ("hl_newton_wave", None, [P, C.POINTER(NewtonOpts), f64p, u32p, u32p]),
        ("hl_column_order", C.c_int, [C.c_uint32, u32p, u32p, u32p]),
        ("hl_amd", C.c_int, [C.c_uint32, u32p, u32p, u32p]),
        ("hl_prep", None, [C.c_int, f64p, f64p]),
        ("hl_exp", C.c_double, [C.c_double]), ("hl_log", C.c_double, [C.c_double]), ("hl_sin", C.c_double, [C.c_double]),
        ("hl_pow", C.c_double, [C.c_double, C.c_double]),
    ]:
        f = getattr(hl, name)
        f.restype = res
        f.argtypes = args
    return hl


class HostTemplate:
    def __init__(self, hl):
        self.hl = hl
        self.h = hl.hl_template_new()

    def __del__(self):
        if getattr(self, "h", None):
            self.hl.hl_template_free(self.h)
            self.h = None

    def node(self, name):
        return self.hl.hl_node(self.h, name.encode())

    def add(self, kind, ports, params=None):
        import oracle_lib
        rows = [self.node(p) if isinstance(p, str) else p for p in ports]
        pr = np.full(oracle_lib.UK[kind], np.nan)
        for k, v in (params or {}).items():
            pr[k] = v
        _, rp = u32(rows)
        a, ap = f64(pr)
        r = self.hl.hl_add(self.h, kind, rp, ap)
        assert r >= 0
        return r

    def freeze(self):
        assert self.hl.hl_freeze(self.h) == 0
        self.n, self.nnz = self.hl.hl_n(self.h), self.hl.hl_nnz(self.h)
        cp, cpp = u32(np.zeros(self.n + 1))
        ri, rip = u32(np.zeros(self.nnz))
        dg, dgp = u32(np.zeros(self.n))
        q, qp = u32(np.zeros(self.n))
        self.hl.hl_pattern(self.h, cpp, rip, dgp, qp)
        self.col_ptr, self.row_idx, self.diag, self.q = cp, ri, dg, q
        return self


class HostLanes:
    def __init__(self, t, N):
        self.t, self.hl, self.N = t, t.hl, N
        self.h = self.hl.hl_lanes_new(t.h, N)

    def __del__(self):
        if getattr(self, "h", None):
            self.hl.hl_lanes_free(self.h)
            self.h = None

    def set_param(self, kind, inst, k, values):
        v, vp = f64(np.broadcast_to(np.asarray(values, dtype=np.float64), (self.N,)))
        self.hl.hl_set_param(self.h, kind, inst, k, vp)

    def prepare(self):
        self.hl.hl_prepare(self.h)

    def newton_op(self, x0=None, wave=False, **kw):
        import oracle_lib
        n, N = self.t.n, self.N
        x = np.zeros(n * N) if x0 is None else np.array(x0, dtype=np.float64).reshape(-1)
        xa, xp = f64(x)
        st, sp = u32(np.zeros(N))
        it, ip = u32(np.zeros(N))
        o = oracle_lib.newton_opts(**kw)
        (self.hl.hl_newton_wave if wave else self.hl.hl_newton_op)(self.h, C.byref(o), xp, sp, ip)
        return xa.reshape(n, N), st, it
