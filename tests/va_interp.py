"""A direct evaluation of the Verilog-A compiler's normalised IR (test-only).

It walks the analog block statement by statement in Python floats, as the
Verilog-A semantics read, with none of the compiler's later stages: no
setup/evaluation split, no relevance pruning, no derivative lanes, no C.
Its residual and charge rows therefore check what those stages emit
(values only; the derivatives are checked by finite differences)."""
import math

from egspice_amd.vagen import ir as IR
from egspice_amd.vagen import parse as PA
from egspice_amd.vagen import pp as PP
from egspice_amd.vagen.compile import split_ddt


def load(path):
    return IR.Model(PA.parse(PP.preprocess(path)))


class Interp:
    def __init__(self, model, user, temp=300.15, gmin=1e-12):
        """user: {param name: value}; the rest take their defaults."""
        self.m = model
        self.temp, self.gmin = temp, gmin
        self.given = set()
        self.p = {}
        ptype = {p["name"]: p["type"] for p in model.params}
        # parameter names are case-insensitive, as instance and model cards
        # read them
        user = {k.lower(): v for k, v in user.items()}
        for p in model.params:
            nm = p["name"]
            if nm.lower() in user:
                v = user[nm.lower()]
                self.given.add(nm)
            else:
                v = self.expr(model.expr(p["default"], [{}], []), None)
            if ptype[nm] == "integer":
                v = int(self.rtoi(v))
            self.p[nm] = v
        self.vtype = model.vartype

    @staticmethod
    def rtoi(v):
        if isinstance(v, int):
            return v
        return int(math.floor(v + 0.5)) if v >= 0 else -int(math.floor(-v + 0.5))

    def run(self, x, tran=True):
        """Residual rows, charge rows at the unknowns' values x."""
        nu = len(self.m.unknowns)
        self.x = x
        self.v = {}
        self.res = [0.0] * nu
        self.q = [0.0] * nu
        self.tran = tran
        self.stmts(self.m.body)
        return self.res, self.q

    # ------------------------------------------------------------ values
    def var(self, name):
        if name in self.v:
            return self.v[name]
        return 0 if self.vtype.get(name) == "integer" else 0.0

    def expr(self, e, x):
        k = e[0]
        if k == "num":
            return e[1]
        if k == "par":
            return self.p[e[1]]
        if k == "var":
            return self.var(e[1])
        if k == "probe":
            a, b = e[1], e[2]
            va = x[a] if a >= 0 else 0.0
            vb = x[b] if b >= 0 else 0.0
            return va - vb
        if k == "zero":
            return 0.0
        if k == "ddx":
            # only operating-point outputs read ddx(); none reaches a row
            return math.nan
        if k == "temp":
            return self.temp
        if k == "gmin":
            return self.gmin
        if k == "mfactor":
            return 1.0
        if k == "given":
            return 1 if e[1] in self.given else 0
        if k == "analysis":
            names = set(e[1])
            return 1 if (names & {"dc", "static", "op"}) or ("tran" in names and self.tran) else 0
        if k == "un":
            a = self.expr(e[2], x)
            if e[1] == "-":
                return -a
            if e[1] == "!":
                return 1 if a == 0 else 0
        if k == "bin":
            op = e[1]
            a, b = self.expr(e[2], x), self.expr(e[3], x)
            if op == "+":
                return a + b
            if op == "-":
                return a - b
            if op == "*":
                return a * b
            if op == "/":
                if isinstance(a, int) and isinstance(b, int):
                    if b == 0:
                        return 0
                    q = abs(a) // abs(b)
                    return q if (a >= 0) == (b >= 0) else -q
                return a / b if b != 0 else (math.copysign(math.inf, a) * math.copysign(1, b) if a else math.nan)
            if op == "%":
                # integer remainder, truncated as C's (0 for a zero divisor)
                if b == 0:
                    return 0
                r = abs(a) % abs(b)
                return r if a >= 0 else -r
            if op == "<":
                return int(a < b)
            if op == "<=":
                return int(a <= b)
            if op == ">":
                return int(a > b)
            if op == ">=":
                return int(a >= b)
            if op == "==":
                return int(a == b)
            if op == "!=":
                return int(a != b)
            if op == "&&":
                return int(a != 0 and b != 0)
            if op == "||":
                return int(a != 0 or b != 0)
            raise ValueError(op)
        if k == "tern":
            return self.expr(e[2], x) if self.expr(e[1], x) != 0 else self.expr(e[3], x)
        if k == "call":
            args = [self.expr(a, x) for a in e[2]]
            fn = e[1]
            if fn == "exp":
                return math.exp(args[0]) if args[0] < 709.78 else math.inf
            if fn == "ln":
                return math.log(args[0]) if args[0] > 0 else (-math.inf if args[0] == 0 else math.nan)
            if fn == "log":
                return math.log10(args[0])
            if fn == "sqrt":
                return math.sqrt(args[0]) if args[0] >= 0 else math.nan
            if fn == "pow":
                try:
                    return math.pow(args[0], args[1])
                except (ValueError, OverflowError):
                    return math.nan
            if fn == "abs":
                return abs(args[0])
            if fn == "min":
                return args[1] if args[1] < args[0] else args[0]
            if fn == "max":
                return args[1] if args[1] > args[0] else args[0]
            if fn == "floor":
                return float(math.floor(args[0]))
            if fn == "ceil":
                return float(math.ceil(args[0]))
            if fn == "tanh":
                return math.tanh(args[0])
            if fn == "atan":
                return math.atan(args[0])
            raise ValueError(fn)
        raise ValueError(k)

    # -------------------------------------------------------- statements
    def stmts(self, sts):
        for s in sts:
            k = s[0]
            if k == "assign":
                v = self.expr(s[2], self.x)
                if self.vtype.get(s[1]) == "integer":
                    v = self.rtoi(v)
                self.v[s[1]] = v
            elif k == "if":
                self.stmts(s[2] if self.expr(s[1], self.x) != 0 else s[3])
            elif k == "while":
                it = 0
                while self.expr(s[1], self.x) != 0 and it < 100000:
                    self.stmts(s[2])
                    it += 1
            elif k == "contrib":
                a, b = s[1], s[2]
                r, qs = split_ddt(s[3])
                if r is not None:
                    val = self.expr(r, self.x)
                    if a >= 0:
                        self.res[a] += val
                    if b >= 0:
                        self.res[b] -= val
                if self.tran:
                    for sign, q, _ in qs:
                        val = sign * self.expr(q, self.x)
                        if a >= 0:
                            self.q[a] += val
                        if b >= 0:
                            self.q[b] -= val
            elif k == "limit":
                # evaluated at x itself (a limited image is an x too)
                self.v[s[1]] = self.expr(s[3], self.x)
            elif k == "collapse":
                pass
            else:
                raise ValueError(k)
