"""C emission of a compiled model (compile.Compiler).

The generated header is plain C that compiles as C (the CPU oracle, gcc)
and as HIP C++ (the lane engine, host and gfx950): the including file
defines VA_FN (function qualifiers) and the elementary functions
VA_EXP, VA_LOG, VA_LOG10, VA_POW, VA_SQRT, VA_FABS, VA_FLOOR, VA_CEIL,
VA_FMA, VA_TANH, VA_ATAN, VA_SINH, VA_COSH, VA_SIN, VA_COS.
"""
from .compile import CMP, LOGIC, expr_reads, split_ddt
from .parse import ParseError
import re

LN10_INV = "0.43429448190325182765"


def cnum(v, is_int):
    if is_int:
        return str(int(v))
    if v != v:
        return "(0.0/0.0)"
    if v == float("inf"):
        return "VA_INF"
    if v == float("-inf"):
        return "(-VA_INF)"
    r = repr(float(v))
    if "e" not in r and "." not in r and "n" not in r:
        r += ".0"
    return r


class Val:
    __slots__ = ("v", "d", "ty")

    def __init__(self, v, d=None, ty="real"):
        self.v = v
        self.d = d or {}
        self.ty = ty


class Emitter:
    def __init__(self, c):
        self.c = c
        self.m = c.m
        self.P = c.prefix
        self.lines = []
        self.ind = 1
        self.tmp = 0
        self.noad = False  # values only: no derivative lanes

    # ------------------------------------------------------------ output
    def out(self, s):
        self.lines.append("    " * self.ind + s)

    def t(self):
        self.tmp += 1
        return f"t{self.tmp}"

    def const(self, expr, ty="real"):
        name = self.t()
        self.out(f"const {'double' if ty == 'real' else 'int'} {name} = {expr};")
        return name

    # ------------------------------------------------- symbol references
    def ref_var(self, name):
        c = self.c
        ty = c.vtype[name]
        if self.mode == "setup":
            return Val(f"v_{name}", ty=ty)
        if name in c.dyn:
            lanes = () if self.noad or ty != "real" else sorted(c.vl.get(name, ()))
            return Val(f"v_{name}", {l: f"v_{name}__{l}" for l in lanes}, ty)
        self.used_static.add(("v", name))
        return Val(f"@v:{name}@", ty=ty)

    def ref_par(self, name):
        ty = self.c.ptype[name]
        if self.mode == "setup":
            return Val(f"p_{name}", ty=ty)
        self.used_static.add(("p", name))
        return Val(f"@p:{name}@", ty=ty)

    # -------------------------------------------------------- expressions
    def as_real(self, a):
        if a.ty == "int":
            return Val(f"((double){a.v})")
        return a

    def ex(self, e):
        k = e[0]
        if k == "num":
            return Val(cnum(e[1], e[2]), ty="int" if e[2] else "real")
        if k == "str":
            raise ParseError("a string in an expression")
        if k == "par":
            return self.ref_par(e[1])
        if k == "var":
            return self.ref_var(e[1])
        if k == "probe":
            a, b = e[1], e[2]
            if a < 0 and b < 0:
                return Val("0.0")
            if self.mode == "setup":
                raise ParseError("a probe in static code")
            lane, sign = self.c.probe_lane(a, b)
            d = {} if self.noad else {lane: "1.0" if sign > 0 else "-1.0"}
            if b < 0:
                return Val(f"x[{a}]", d)
            if a < 0:
                return Val(self.const(f"-(x[{b}])"), d)
            return Val(self.const(f"x[{a}] - x[{b}]"), d)
        if k == "zero":
            return Val("0.0")
        if k == "temp":
            return Val("va_temp" if self.mode == "setup" else "s_va_temp")
        if k == "gmin":
            return Val("va_gmin" if self.mode == "setup" else "s_va_gmin")
        if k == "mfactor":
            return Val("1.0")
        if k == "given":
            if self.mode != "setup":
                self.used_static.add(("g", e[1]))
                return Val(f"@g:{e[1]}@", ty="int")
            return Val(f"g_{e[1]}", ty="int")
        if k == "analysis":
            names = set(e[1])
            # the large-signal analyses: the operating point and the transient
            hit_static = bool(names & {"dc", "static", "op"})
            hit_tran = bool(names & {"tran"})
            if hit_static and hit_tran:
                return Val("1", ty="int")
            if not hit_static and not hit_tran:
                return Val("0", ty="int")
            if self.mode == "setup":
                raise ParseError("analysis() in static code")
            return Val(("VA_TRAN" if hit_tran else "(!VA_TRAN)"), ty="int")
        if k == "abstime":
            raise ParseError("$abstime is not supported")
        if k == "ddt":
            raise ParseError("ddt() outside a contribution")
        if k == "ddx":
            # the node partial from the branch-voltage lanes: a lane's pair
            # (a, b) moves with node a (+) and against node b (-); a value
            # without derivative lanes of its own
            if self.mode == "setup":
                raise ParseError("ddx() in static code")
            x = self.as_real(self.ex(e[1]))
            node = e[2]
            terms = []
            for l in sorted(x.d):
                pa, pb = self.c.probe_pair[l]
                if pa == node:
                    terms.append(x.d[l])
                elif pb == node:
                    terms.append(f"-({x.d[l]})")
            if not terms:
                return Val("0.0")
            acc = self.const(terms[0])
            for t in terms[1:]:
                acc = self.const(f"{acc} + {t}")
            return Val(acc)
        if k == "un":
            a = self.ex(e[2])
            if e[1] == "-":
                if a.ty == "int":
                    return Val(self.const(f"-({a.v})", "int"), ty="int")
                return Val(self.const(f"-({a.v})"), {l: self.const(f"-({d})") for l, d in a.d.items()})
            if e[1] == "!":
                return Val(self.const(f"({a.v} == 0)", "int"), ty="int")
            raise ParseError(f"unary {e[1]}")
        if k == "bin":
            return self.binop(e[1], self.ex(e[2]), self.ex(e[3]))
        if k == "tern":
            return self.tern(e)
        if k == "call":
            return self.call(e[1], e[2])
        raise ParseError(f"cannot emit {k}")

    def binop(self, op, a, b):
        if op in CMP:
            return Val(self.const(f"({a.v} {op} {b.v})", "int"), ty="int")
        if op in LOGIC:
            return Val(self.const(f"(({a.v}) != 0 {op} ({b.v}) != 0)", "int"), ty="int")
        if a.ty == "int" and b.ty == "int":
            if op in "+-*":
                return Val(self.const(f"{a.v} {op} {b.v}", "int"), ty="int")
            if op == "/":
                return Val(self.const(f"VA_IDIV({a.v}, {b.v})", "int"), ty="int")
            if op == "%":
                return Val(self.const(f"VA_IMOD({a.v}, {b.v})", "int"), ty="int")
            raise ParseError(f"integer operator {op}")
        a, b = self.as_real(a), self.as_real(b)
        if op in "+-":
            v = self.const(f"{a.v} {op} {b.v}")
            d = {}
            for l in sorted(set(a.d) | set(b.d)):
                if l in a.d and l in b.d:
                    d[l] = self.const(f"{a.d[l]} {op} {b.d[l]}")
                elif l in a.d:
                    d[l] = a.d[l]
                else:
                    d[l] = b.d[l] if op == "+" else self.const(f"-({b.d[l]})")
            return Val(v, d)
        if op == "*":
            v = self.const(f"{a.v} * {b.v}")
            d = {}
            for l in sorted(set(a.d) | set(b.d)):
                if l in a.d and l in b.d:
                    d[l] = self.const(f"VA_FMA({b.d[l]}, {a.v}, {a.d[l]} * {b.v})")
                elif l in a.d:
                    d[l] = self.const(f"{a.d[l]} * {b.v}")
                else:
                    d[l] = self.const(f"{b.d[l]} * {a.v}")
            return Val(v, d)
        if op == "/":
            q = self.const(f"{a.v} / {b.v}")
            d = {}
            if a.d or b.d:
                inv = self.const(f"1.0 / {b.v}")
                for l in sorted(set(a.d) | set(b.d)):
                    if l in b.d:
                        num = a.d.get(l, "0.0")
                        d[l] = self.const(f"VA_FMA({b.d[l]}, -{q}, {num}) * {inv}")
                    else:
                        d[l] = self.const(f"{a.d[l]} * {inv}")
            return Val(q, d)
        raise ParseError(f"real operator {op}")

    def tern(self, e):
        cnd = self.ex(e[1])
        lanes = sorted(self.c.elanes(e[2]) | self.c.elanes(e[3])) if self.mode != "setup" and not self.noad else []
        ty = "int" if self.c.etype(e[2]) == self.c.etype(e[3]) == "int" else "real"
        r = self.t()
        ctype = "double" if ty == "real" else "int"
        self.out(f"{ctype} {r}" + "".join(f", {r}__{l}" for l in lanes) + ";")
        self.out(f"if ({cnd.v}) {{")
        self.ind += 1
        a = self.ex(e[2])
        if ty == "real":
            a = self.as_real(a)
        self.out(f"{r} = {a.v};" + "".join(f" {r}__{l} = {a.d.get(l, '0.0')};" for l in lanes))
        self.ind -= 1
        self.out("} else {")
        self.ind += 1
        b = self.ex(e[3])
        if ty == "real":
            b = self.as_real(b)
        self.out(f"{r} = {b.v};" + "".join(f" {r}__{l} = {b.d.get(l, '0.0')};" for l in lanes))
        self.ind -= 1
        self.out("}")
        return Val(r, {l: f"{r}__{l}" for l in lanes}, ty)

    def map1(self, a, value, slope):
        """f(a) with value `value` and derivative factor `slope` (a C
        expression, evaluated only when a carries lanes)."""
        v = self.const(value)
        d = {}
        if a.d:
            s = self.const(slope.replace("$v", v))
            for l, x in a.d.items():
                d[l] = self.const(f"{x} * {s}")
        return Val(v, d)

    def call(self, fn, args):
        vals = [self.ex(a) for a in args]
        if fn in ("min", "max"):
            a, b = vals
            if a.ty == "int" and b.ty == "int":
                op = "<" if fn == "min" else ">"
                return Val(self.const(f"({b.v} {op} {a.v}) ? {b.v} : {a.v}", "int"), ty="int")
            a, b = self.as_real(a), self.as_real(b)
            op = "<" if fn == "min" else ">"
            pick = self.const(f"({b.v} {op} {a.v})", "int")
            v = self.const(f"{pick} ? {b.v} : {a.v}")
            d = {}
            for l in sorted(set(a.d) | set(b.d)):
                d[l] = self.const(f"{pick} ? {b.d.get(l, '0.0')} : {a.d.get(l, '0.0')}")
            return Val(v, d)
        if fn == "abs":
            a = vals[0]
            if a.ty == "int":
                return Val(self.const(f"({a.v} < 0 ? -{a.v} : {a.v})", "int"), ty="int")
            v = self.const(f"VA_FABS({a.v})")
            d = {}
            if a.d:
                neg = self.const(f"({a.v} < 0.0)", "int")
                for l, x in a.d.items():
                    d[l] = self.const(f"{neg} ? -({x}) : {x}")
            return Val(v, d)
        vals = [self.as_real(a) for a in vals]
        if fn == "pow":
            a, b = vals
            p = self.const(f"VA_POW({a.v}, {b.v})")
            d = {}
            if a.d:
                sa = self.const(f"{a.v} != 0.0 ? {b.v} * {p} / {a.v} : {b.v} * VA_POW({a.v}, {b.v} - 1.0)")
                sa = self.const(f"VA_FABS({sa}) < VA_INF ? {sa} : 0.0")
            if b.d:
                sb = self.const(f"{a.v} > 0.0 ? {p} * VA_LOG({a.v}) : 0.0")
            for l in sorted(set(a.d) | set(b.d)):
                if l in a.d and l in b.d:
                    d[l] = self.const(f"{a.d[l]} * {sa} + {b.d[l]} * {sb}")
                elif l in a.d:
                    d[l] = self.const(f"{a.d[l]} * {sa}")
                else:
                    d[l] = self.const(f"{b.d[l]} * {sb}")
            return Val(p, d)
        a = vals[0]
        if fn == "exp":
            return self.map1(a, f"VA_EXP({a.v})", "$v")
        if fn == "ln":
            return self.map1(a, f"VA_LOG({a.v})", f"1.0 / {a.v}")
        if fn == "log":
            return self.map1(a, f"VA_LOG10({a.v})", f"{LN10_INV} / {a.v}")
        if fn == "sqrt":
            return self.map1(a, f"VA_SQRT({a.v})", "$v > 0.0 ? 0.5 / $v : 0.0")
        if fn in ("floor", "ceil"):
            return Val(self.const(f"VA_{fn.upper()}({a.v})"))
        if fn == "tanh":
            return self.map1(a, f"VA_TANH({a.v})", "1.0 - $v * $v")
        if fn == "atan":
            return self.map1(a, f"VA_ATAN({a.v})", f"1.0 / (1.0 + {a.v} * {a.v})")
        if fn == "sinh":
            return self.map1(a, f"VA_SINH({a.v})", f"VA_COSH({a.v})")
        if fn == "cosh":
            return self.map1(a, f"VA_COSH({a.v})", f"VA_SINH({a.v})")
        if fn == "sin":
            return self.map1(a, f"VA_SIN({a.v})", f"VA_COS({a.v})")
        if fn == "cos":
            return self.map1(a, f"VA_COS({a.v})", f"-VA_SIN({a.v})")
        raise ParseError(f"function {fn} is not supported by the emitter")

    # --------------------------------------------------------- statements
    def keep(self, st):
        """Does leaf statement st belong to the function being emitted?"""
        c = self.c
        k = st.kind
        if self.mode == "setup":
            if k == "collapse":
                return True
            if k == "assign":
                v = st.a[1]
                return v not in c.dyn and any(v in c.rel[m] for m in ("dc", "tran", "limitargs"))
            return False
        if self.mode == "limitargs":
            if k == "limit":
                return True
            if k == "assign":
                v = st.a[1]
                return v in c.dyn and v in c.rel["limitargs"]
            return False
        if k == "contrib":
            r, qs = split_ddt(st.a[3])
            if self.mode == "charge":
                return bool(qs)
            return r is not None or (self.mode == "tran" and qs)
        if k in ("assign", "limit"):
            v = st.a[1]
            return v in c.dyn and v in c.rel[self.mode]
        return False

    def any_kept(self, sts):
        for st in sts:
            if st.kind in ("if", "while"):
                if any(self.any_kept(k) for k in st.kids):
                    return True
            elif self.keep(st):
                return True
        return False

    def stmts(self, sts):
        for st in sts:
            k = st.kind
            if k == "if":
                if not (self.any_kept(st.kids[0]) or self.any_kept(st.kids[1])):
                    continue
                self.out("{")
                self.ind += 1
                cnd = self.ex(st.a[1])
                self.out(f"if ({cnd.v}) {{")
                self.ind += 1
                self.stmts(st.kids[0])
                self.ind -= 1
                if self.any_kept(st.kids[1]):
                    self.out("} else {")
                    self.ind += 1
                    self.stmts(st.kids[1])
                    self.ind -= 1
                self.out("}")
                self.ind -= 1
                self.out("}")
            elif k == "while":
                if not self.any_kept(st.kids[0]):
                    continue
                self.out("for (int va_it = 0;; va_it++) {")
                self.ind += 1
                self.out("if (va_it >= 100000) break;")
                cnd = self.ex(st.a[1])
                self.out(f"if (!({cnd.v})) break;")
                self.stmts(st.kids[0])
                self.ind -= 1
                self.out("}")
            elif self.keep(st):
                self.leaf(st)

    def leaf(self, st):
        k = st.kind
        c = self.c
        if k == "limit" and self.mode == "limitargs":
            # the limiter's arguments where the call stands in program order
            base = self.limit_base[id(st)]
            self.out("{")
            self.ind += 1
            for i, a in enumerate(st.a[4]):
                val = self.as_real(self.ex(a))
                self.out(f"la[{base + i}] = {val.v};")
            self.ind -= 1
            self.out("}")
            return
        if k == "limit":
            # the evaluation runs at the limited image of the unknowns, so
            # $limit(V(a, b), ...) reads V(a, b) there
            st = type(st)("assign", ("assign", st.a[1], st.a[3]), st.pos, st.conds, st.loops)
            k = "assign"
        if k == "assign":
            v = st.a[1]
            self.out("{")
            self.ind += 1
            val = self.ex(st.a[2])
            ty = c.vtype[v]
            if ty == "int":
                if val.ty == "int":
                    self.out(f"v_{v} = {val.v};")
                else:
                    self.out(f"v_{v} = VA_RTOI({val.v});")
            else:
                val = self.as_real(val)
                line = f"v_{v} = {val.v};"
                if self.mode != "setup" and not self.noad:
                    for l in sorted(c.vl.get(v, ())):
                        line += f" v_{v}__{l} = {val.d.get(l, '0.0')};"
                self.out(line)
            self.ind -= 1
            self.out("}")
        elif k == "collapse":
            self.out(f"prep[{self.coll_base + self.coll_index[id(st)]}] = 1.0;")
        elif k == "contrib":
            a, b = st.a[1], st.a[2]
            r, qs = split_ddt(st.a[3])
            self.out("{")
            self.ind += 1
            if r is not None and self.mode != "charge":
                val = self.as_real(self.ex(r))
                self.stamp(a, b, val, "res", "jac", self.jidx)
            if self.mode in ("tran", "charge"):
                for i, (sign, q, _) in enumerate(qs):
                    val = self.as_real(self.ex(q))
                    if sign < 0:
                        val = Val(self.const(f"-({val.v})"), {l: self.const(f"-({x})") for l, x in val.d.items()})
                    self.stamp(a, b, val, "q", "cjac", self.cidx)
                    self.out(f"site[{self.site_index[(id(st), i)]}] = {val.v};")
            self.ind -= 1
            self.out("}")

    def stamp(self, a, b, val, rv, jv, jidx):
        """Rows a (+) and b (-); each lane's partial onto its probe's node
        columns by the chain rule (+ on the pair's first node, - on its
        second)."""
        for node, sgn in ((a, 1), (b, -1)):
            if node < 0:
                continue
            self.out(f"{rv}[{node}] {'+' if sgn > 0 else '-'}= {val.v};")
            if self.noad or self.mode == "charge":
                continue
            for l in sorted(val.d):
                pa, pb = self.c.probe_pair[l]
                for col, cs in ((pa, 1), (pb, -1)):
                    if col < 0:
                        continue
                    op = "+" if sgn * cs > 0 else "-"
                    self.out(f"{jv}[{jidx[(node, col)]}] {op}= {val.d[l]};")


def structure(c):
    """The static G and C patterns (row, col) of the model and its charge
    sites, from every contribution's lanes."""
    g, cc = set(), set()
    sites = []
    for st, r, qs in c.contribs:
        a, b = st.a[1], st.a[2]
        rows = [x for x in (a, b) if x >= 0]
        def cols(lanes):
            out = set()
            for l in lanes:
                out |= {x for x in c.probe_pair[l] if x >= 0}
            return out

        if r is not None:
            for row in rows:
                for col in cols(c.elanes(r)):
                    g.add((row, col))
        for i, (sign, q, _) in enumerate(qs):
            sites.append((id(st), i))
            for row in rows:
                for col in cols(c.elanes(q)):
                    cc.add((row, col))
    return sorted(g), sorted(cc), sites


def generate(c):
    m = c.m
    P = c.prefix
    nu = c.nu
    gpat, cpat, sites = structure(c)
    jidx = {rc: i for i, rc in enumerate(gpat)}
    cidx = {rc: i for i, rc in enumerate(cpat)}
    site_index = {s: i for i, s in enumerate(sites)}
    collapses = [st for st in c.flat if st.kind == "collapse"]
    params = [p for p in m.params if not p.get("local")]
    locals_ = [p for p in m.params if p.get("local")]
    pindex = {p["name"]: i for i, p in enumerate(params)}

    body = []
    # the limiter's arguments, in call order
    limit_base, nla = {}, 0
    for st in c.limits:
        limit_base[id(st)] = nla
        nla += len(st.a[4])
    modes = ("dc", "tran", "charge") + (("limitargs",) if c.limits else ())
    # ---------------------------------------------------------- eval
    evals = {}
    used_static = set()
    for mode in modes:
        e = Emitter(c)
        e.mode = mode
        e.limit_base = limit_base
        # the charge values alone carry no derivative lanes, unless ddx()
        # reads lanes on their path; the limiter's arguments are values
        e.noad = (mode == "charge" and not c.ddx_in["charge"]) or (mode == "limitargs" and not c.ddx_in["limitargs"])
        e.used_static = set()
        e.jidx, e.cidx, e.site_index = jidx, cidx, site_index
        e.coll_base, e.coll_index = 0, {}
        e.stmts(c.tree)
        evals[mode] = (e.lines, e.tmp)
        used_static |= e.used_static
    statics = sorted(used_static)
    sidx = {s: i for i, s in enumerate(statics)}
    nprep = len(statics) + 2 + len(collapses)
    coll_base = len(statics) + 2

    # ---------------------------------------------------------- setup
    e = Emitter(c)
    e.mode = "setup"
    e.used_static = set()
    e.coll_base = coll_base
    e.coll_index = {id(st): i for i, st in enumerate(collapses)}
    e.stmts(c.tree)
    setup_lines = e.lines

    H = []
    H.append(f"/* {P}: generated from the Verilog-A module {m.name} by egspice_amd.vagen. Do not edit. */")
    H.append(f"#define {P.upper()}_NU {nu}")
    H.append(f"#define {P.upper()}_NPORT {m.nport}")
    H.append(f"#define {P.upper()}_NPARAM {len(params)}")
    H.append(f"#define {P.upper()}_NPREP {nprep}")
    H.append(f"#define {P.upper()}_NJ {len(gpat)}")
    H.append(f"#define {P.upper()}_NCJ {len(cpat)}")
    H.append(f"#define {P.upper()}_NSITE {len(sites)}")
    H.append(f"#define {P.upper()}_NCOLL {len(collapses)}")
    H.append(f"#define {P.upper()}_COLL_BASE {coll_base}")
    gm = [0] * nu
    cm = [0] * nu
    for r, l in gpat:
        gm[r] |= 1 << l
    for r, l in cpat:
        cm[r] |= 1 << l
    H.append(f"VA_TABLE unsigned {P}_gmask[{nu}] = {{{', '.join(hex(x) + 'u' for x in gm)}}};")
    H.append(f"VA_TABLE unsigned {P}_cmask[{nu}] = {{{', '.join(hex(x) + 'u' for x in cm)}}};")
    cl = ", ".join(f"{st.a[1]}, {st.a[2]}" for st in collapses) or "0"
    H.append(f"VA_TABLE int {P}_collapse[{max(1, 2 * len(collapses))}] = {{{cl}}};")
    H.append(f"VA_HOST_TABLE char *const {P}_param_names[{len(params)}] = {{")
    for i in range(0, len(params), 8):
        H.append("    " + " ".join(f'"{p["name"]}",' for p in params[i:i + 8]))
    H.append("};")
    H.append(f"VA_HOST_TABLE unsigned char {P}_param_instance[{len(params)}] = {{")
    for i in range(0, len(params), 32):
        H.append("    " + " ".join(f'{1 if p["instance"] else 0},' for p in params[i:i + 32]))
    H.append("};")
    aliases = [(a, pindex[t]) for a, t in sorted(m.aliases.items()) if t in pindex]
    H.append(f"#define {P.upper()}_NALIAS {len(aliases)}")
    H.append(f"VA_HOST_TABLE char *const {P}_alias_names[{max(1, len(aliases))}] = {{{', '.join(chr(34) + a + chr(34) for a, _ in aliases) or '0'}}};")
    H.append(f"VA_HOST_TABLE int {P}_alias_target[{max(1, len(aliases))}] = {{{', '.join(str(t) for _, t in aliases) or '0'}}};")

    # setup function
    H.append("")
    H.append(f"/* u[{len(params)}]: parameter values in declaration order, NaN = not given. Returns 0, or")
    H.append(" * 1 + the index of a parameter outside its declared range. */")
    H.append(f"VA_FN int {P}_setup(const double *u, double va_temp, double va_gmin, double *prep)")
    H.append("{")
    # declarations
    for p in params:
        ct = "double" if c.ptype[p["name"]] == "real" else "int"
        H.append(f"    {ct} p_{p['name']};")
        H.append(f"    const int g_{p['name']} = u[{pindex[p['name']]}] == u[{pindex[p['name']]}];")
    for p in locals_:
        ct = "double" if c.ptype[p["name"]] == "real" else "int"
        H.append(f"    {ct} p_{p['name']};")
    for v, ty in sorted(c.vtype.items()):
        if v in c.dyn:
            continue
        H.append(f"    {'double' if ty == 'real' else 'int'} v_{v} = 0;")
    H.append(f"    for (int k = 0; k < {nprep}; k++) prep[k] = 0.0;")
    # parameter values: given, or the default expression; then the range
    pe = Emitter(c)
    pe.mode = "setup"
    pe.used_static = set()
    for p in params + locals_:
        nm = p["name"]
        ty = c.ptype[nm]
        pe.lines = []
        pe.ind = 1
        pe.out("{")
        pe.ind += 1
        dv = pe.ex(m_expr(m, p["default"]))
        dvs = dv.v if ty == "int" or dv.ty == "real" else f"((double){dv.v})"
        if ty == "int" and dv.ty == "real":
            dvs = f"VA_RTOI({dv.v})"
        if p.get("local"):
            pe.out(f"p_{nm} = {dvs};")
        else:
            k = pindex[nm]
            if ty == "int":
                pe.out(f"p_{nm} = g_{nm} ? VA_RTOI(u[{k}]) : {dvs};")
            else:
                pe.out(f"p_{nm} = g_{nm} ? u[{k}] : {dvs};")
            for kind, lo, hi, lo_open, hi_open in p["ranges"]:
                lv = pe.ex(m_expr(m, lo))
                if hi is None:
                    cond = f"(p_{nm} == {lv.v})"
                else:
                    hv = pe.ex(m_expr(m, hi))
                    cond = (f"(p_{nm} {'>' if lo_open else '>='} {lv.v} && p_{nm} {'<' if hi_open else '<='} {hv.v})")
                if kind == "from":
                    pe.out(f"if (!{cond}) return {k + 1};")
                else:
                    pe.out(f"if ({cond}) return {k + 1};")
        pe.ind -= 1
        pe.out("}")
        H.extend(pe.lines)
    H.extend(setup_lines)
    for (kind, nm), i in sidx.items():
        src = {"v": f"v_{nm}", "p": f"p_{nm}", "g": f"g_{nm}"}[kind]
        H.append(f"    prep[{i}] = (double){src};")
    H.append(f"    prep[{len(statics)}] = va_temp;")
    H.append(f"    prep[{len(statics) + 1}] = va_gmin;")
    H.append("    return 0;")
    H.append("}")

    # eval functions: the static evaluation, the transient one, and the
    # transient's charge rows and sites alone, values only (the same
    # operations on the value path, so the same bits as eval_tran's)
    for mode in modes:
        lines, _ = evals[mode]
        fname = {"dc": f"{P}_eval", "tran": f"{P}_eval_tran", "charge": f"{P}_eval_charge",
                 "limitargs": f"{P}_limit_args"}[mode]
        args = "const double *VA_RESTRICT prep, const double *VA_RESTRICT x"
        outs = []
        if mode == "limitargs":
            args += ", double *VA_RESTRICT la"
        if mode in ("dc", "tran"):
            args += ", double *VA_RESTRICT res_out, double *VA_RESTRICT jac_out"
            outs += [("res", nu), ("jac", len(gpat))]
        if mode == "tran":
            args += ", double *VA_RESTRICT q_out, double *VA_RESTRICT cjac_out, double *VA_RESTRICT site_out"
            outs += [("q", nu), ("cjac", len(cpat)), ("site", len(sites))]
        if mode == "charge":
            args += ", double *VA_RESTRICT q_out, double *VA_RESTRICT site_out"
            outs += [("q", nu), ("site", len(sites))]
        H.append("")
        H.append(f"VA_FN void {fname}({args})")
        H.append("{")
        H.append(f"#define VA_TRAN {0 if mode in ('dc', 'limitargs') else 1}")
        if mode == "limitargs":
            H.append(f"    for (int k = 0; k < {nla}; k++) la[k] = 0.0;")
        for nm, cnt in outs:
            H.append(f"    double {nm}[{max(1, cnt)}];")
            H.append(f"    for (int k = 0; k < {max(1, cnt)}; k++) {nm}[k] = 0.0;")
        H.append(f"    const double s_va_temp = prep[{len(statics)}];")
        H.append(f"    const double s_va_gmin = prep[{len(statics) + 1}];")
        H.append("    (void)s_va_temp; (void)s_va_gmin;")
        def sub(mm):
            kind, nm = mm.group(1), mm.group(2)
            i = sidx[(kind, nm)]
            ty = c.ptype[nm] if kind == "p" else ("int" if kind == "g" else c.vtype[nm])
            return f"prep[{i}]" if ty == "real" else f"((int)prep[{i}])"

        lines = [re.sub(r"@([vpg]):([^@]+)@", sub, ln) for ln in lines]
        for v in sorted(c.dyn):
            if v not in c.rel[mode]:
                continue
            ty = c.vtype[v]
            if ty == "int":
                H.append(f"    int v_{v} = 0;")
            else:
                noad = mode in ("charge", "limitargs") and not c.ddx_in[mode]
                ls = sorted(c.vl.get(v, ())) if not noad else []
                H.append(f"    double v_{v} = 0.0" + "".join(f", v_{v}__{l} = 0.0" for l in ls) + ";")
        H.extend(lines)
        for nm, cnt in outs:
            if cnt:
                H.append(f"    for (int k = 0; k < {cnt}; k++) {nm}_out[k] = {nm}[k];")
        H.append("#undef VA_TRAN")
        H.append("}")
    H.extend(limit_functions(c, P, nu, limit_base, nla, [(st.a[1], st.a[2]) for st in collapses], coll_base))
    return "\n".join(H) + "\n", dict(nprep=nprep, nj=len(gpat), ncj=len(cpat), nsite=len(sites),
                                     ncoll=len(collapses), nparam=len(params), gpat=gpat, cpat=cpat)


def m_expr(m, e):
    """A parameter default or range bound normalised in the module scope."""
    return m.expr(e, [{}], [])


def limit_functions(c, P, nu, limit_base, nla, pairs, coll_base):
    """The limiter of a model's $limit calls, as the reference's VerA
    groups them for a MOSFET (models/mos1.va:286-308): "fetlimds" on both
    gate legs and "limvds" on the channel limit the leg that controls the
    channel in the sign of the old vds, then derive the other; "pnjlimds"
    on both bulk legs limits one junction, picked by the limited vds, and
    moves only the bulk. The argument after the frame sign `type` is the
    MODEINITJCT seed. P_limit writes the limited image of the unknowns
    (node space: the source is kept, the gate, drain and bulk move) and
    returns pnjlim's check; P_seed writes the seeds' image. An unknown
    collapsed onto a limited one (the same node of the instance) takes its
    image, so the evaluation reads one potential per node: P_LIMW names
    the limited unknowns and every unknown a collapse may join to one. The
    includer defines VA_FETLIM, VA_LIMVDS, VA_PNJLIM and VA_SNAP (the 16-ulp
    snap of the image back onto the unknowns)."""
    U = P.upper()
    if not c.limits:
        return [f"#define {U}_LIMW 0u"]
    legs = {}
    for st in c.limits:
        fn = st.a[2]
        a, b = st.a[3][1], st.a[3][2]
        legs.setdefault(fn, []).append((a, b, limit_base[id(st)], len(st.a[4])))
    if sorted(legs) != ["fetlimds", "limvds", "pnjlimds"] or len(legs["fetlimds"]) != 2 or \
            len(legs["limvds"]) != 1 or len(legs["pnjlimds"]) != 2:
        raise ParseError("unsupported $limit group: " + ", ".join(f"{k} x{len(v)}" for k, v in sorted(legs.items())))
    (g1, s_, f1, n1), (g2, d_, f2, n2) = legs["fetlimds"]
    (dv, sv, lv, nl), = legs["limvds"]
    (b1, sb, p1, np1), (b2, db, p2, np2) = legs["pnjlimds"]
    if not (g1 == g2 and dv == d_ and sv == s_ and sb == s_ and db == d_ and b1 == b2) or min(g1, s_, d_, b1) < 0:
        raise ParseError("a $limit group whose legs do not share gate, drain, source and bulk")
    if n1 < 3 or n2 < 2 or nl < 2 or np1 < 4 or np2 < 3:
        raise ParseError("a $limit leg without its frame sign or seed")
    g, d, src, b = g1, d_, s_, b1
    writes = (1 << g) | (1 << d) | (1 << b)
    reads = writes | (1 << src)
    # unknowns some collapse can join to a limited one
    comp = list(range(nu))

    def root(u):
        while comp[u] != u:
            u = comp[u]
        return u

    for a_, b_ in pairs:
        if a_ >= 0 and b_ >= 0:
            ra, rb = root(a_), root(b_)
            if ra != rb:
                comp[max(ra, rb)] = min(ra, rb)
    partners = [u for u in range(nu) if not (writes >> u & 1) and any(root(u) == root(w) for w in (g, d, b))]
    if any(root(src) == root(w) for w in (g, d, b)):
        raise ParseError("a $limit group whose source may collapse onto a limited node")
    ext = writes | sum(1 << u for u in partners)

    def partner_lines():
        out = []
        if not partners:
            return out
        out.append("    {")
        out.append(f"        int rep[{nu}];")
        out.append(f"        for (int u = 0; u < {nu}; u++) rep[u] = u;")
        for k, (a_, b_) in enumerate(pairs):
            if a_ < 0 or b_ < 0:
                continue
            out.append(f"        if (prep[{coll_base + k}] != 0.0) {{")
            out.append(f"            int ra = {a_}, rb = {b_};")
            out.append("            while (rep[ra] != ra) ra = rep[ra];")
            out.append("            while (rep[rb] != rb) rb = rep[rb];")
            out.append("            if (ra < rb) rep[rb] = ra; else if (rb < ra) rep[ra] = rb;")
            out.append("        }")
        out.append(f"        for (int u = 0; u < {nu}; u++)")
        out.append("            while (rep[rep[u]] != rep[u]) rep[u] = rep[rep[u]];")
        for u in partners:
            for w in (g, d, b):
                if root(u) == root(w):
                    out.append(f"        if (rep[{u}] == rep[{w}]) lim[{u}] = lim[{w}];")
        out.append("    }")
        return out

    L = []
    L.append(f"#define {U}_LIMW {hex(ext)}u")
    L.append("")
    L.append(f"VA_FN int {P}_limit(const double *VA_RESTRICT prep, const double *VA_RESTRICT cur, const double *VA_RESTRICT old,")
    L.append("                    double *VA_RESTRICT lim)")
    L.append("{")
    L.append(f"    double la[{max(1, nla)}];")
    L.append(f"    {P}_limit_args(prep, old, la);")
    L.append(f"    const double type = la[{f1 + 1}];")
    L.append(f"    const double vgs_o = type * (old[{g}] - old[{src}]), vds_o = type * (old[{d}] - old[{src}]);")
    L.append(f"    const double vbs_o = type * (old[{b}] - old[{src}]), vbd_o = type * (old[{b}] - old[{d}]);")
    L.append("    const double vgd_o = vgs_o - vds_o;")
    L.append(f"    double vgs = type * (cur[{g}] - cur[{src}]), vds = type * (cur[{d}] - cur[{src}]);")
    L.append(f"    double vbs = type * (cur[{b}] - cur[{src}]), vbd = type * (cur[{b}] - cur[{d}]);")
    L.append("    double vgd = vgs - vds;")
    L.append("    if (vds_o >= 0.0) {")
    L.append(f"        vgs = VA_FETLIM(vgs, vgs_o, la[{f1}]);")
    L.append("        vds = vgs - vgd;")
    L.append("        vds = VA_LIMVDS(vds, vds_o);")
    L.append("    } else {")
    L.append(f"        vgd = VA_FETLIM(vgd, vgd_o, la[{f2}]);")
    L.append("        vds = vgs - vgd;")
    L.append("        vds = -VA_LIMVDS(-vds, -vds_o);")
    L.append("        vgs = vgd + vds;")
    L.append("    }")
    L.append("    int check = 0;")
    L.append("    if (vds >= 0.0) {")
    L.append(f"        vbs = VA_PNJLIM(vbs, vbs_o, la[{p1}], la[{p1 + 1}], &check);")
    L.append("    } else {")
    L.append(f"        vbd = VA_PNJLIM(vbd, vbd_o, la[{p2}], la[{p2 + 1}], &check);")
    L.append("        vbs = vbd + vds;")
    L.append("    }")
    L.append(f"    for (int u = 0; u < {nu}; u++) lim[u] = cur[u];")
    L.append(f"    const double s0 = cur[{src}];")
    L.append(f"    lim[{d}] = s0 + type * vds;")
    L.append(f"    lim[{g}] = s0 + type * vgs;")
    L.append(f"    lim[{b}] = s0 + type * vbs;")
    L.append(f"    VA_SNAP({nu}, {hex(reads)}u, {hex(writes)}u, lim, cur);")
    L.extend(partner_lines())
    L.append("    return check;")
    L.append("}")
    L.append("")
    L.append(f"VA_FN void {P}_seed(const double *VA_RESTRICT prep, const double *VA_RESTRICT cur, double *VA_RESTRICT lim)")
    L.append("{")
    L.append(f"    double la[{max(1, nla)}];")
    L.append(f"    {P}_limit_args(prep, cur, la);")
    L.append(f"    const double type = la[{f1 + 1}];")
    L.append(f"    for (int u = 0; u < {nu}; u++) lim[u] = cur[u];")
    L.append(f"    const double s0 = cur[{src}];")
    L.append(f"    lim[{d}] = s0 + type * la[{lv + 1}];")
    L.append(f"    lim[{g}] = s0 + type * la[{f1 + 2}];")
    L.append(f"    lim[{b}] = s0 + type * la[{p1 + 3}];")
    L.extend(partner_lines())
    L.append("}")
    return L
