"""Normalisation of a parsed Verilog-A module into the compiler's IR.

Names are resolved (parameters through aliasparam, block-local variables
renamed apart, nodes and branches to local unknown indices), `case` and
`for` become if-chains and while loops, analog functions are inlined at
their call sites, and the system functions the lane engine gives meaning
to are turned into IR nodes:

  expressions  ('num', v, is_int) ('par', name) ('var', name)
               ('probe', a, b)        V(a, b); -1 is ground
               ('zero',)              a noise source or a flow probe of a
                                      branch that only carries noise
               ('temp',) ('gmin',) ('mfactor',) ('abstime',)
               ('given', param)       $param_given
               ('analysis', tran?)    analysis("...") in the large-signal run
               ('call', fn, [args]) ('un', op, e) ('bin', op, a, b)
               ('tern', c, a, b) ('ddt', e)
               ('ddx', e, n)          the partial of e in node n's potential
  statements   ('assign', var, e) ('if', c, [then], [else])
               ('while', c, [body]) ('contrib', a, b, e) ('collapse', a, b)
               ('limit', var, kind, e_probe, [args])   (bsim4's $limit)
"""
from .parse import ParseError

MATH1 = {"exp", "ln", "log", "sqrt", "abs", "floor", "ceil", "tanh", "atan", "sinh", "cosh", "sin", "cos", "tan",
         "asinh", "acosh", "atanh", "asin", "acos", "limexp"}
MATH2 = {"pow", "min", "max", "atan2", "hypot"}
NOISE = {"white_noise", "flicker_noise", "noise_table", "noise_table_log"}
IGNORED_TASKS = {"$strobe", "$display", "$write", "$monitor", "$warning", "$error", "$fatal", "$finish", "$stop",
                 "$debug", "$info", "$bound_step", "$discontinuity"}


class Model:
    """A normalised module."""

    def __init__(self, mod):
        self.name = mod.name
        self.mod = mod
        thermal = getattr(mod, "thermal", set())
        ports = [p for p in mod.ports if p not in thermal]
        internal = [n for n in mod.nodes if n not in ports and n not in thermal]
        self.unknowns = ports + internal
        self.nport = len(ports)
        self.index = {n: i for i, n in enumerate(self.unknowns)}
        self.branches = mod.branches
        self.params = [p for p in mod.params]
        self.param_index = {p["name"]: p for p in self.params}
        self.aliases = dict(mod.aliases)
        self.vartype = {}
        self.counter = 0
        self.noise_branches = set()
        self.inline_depth = 0
        for nm, ty in mod.vars.items():
            self.vartype[nm] = ty
        scope = [{nm: nm for nm in mod.vars}]
        self._find_noise_branches(mod.analog)
        self.body = self.stmt(mod.analog, scope)

    # ---------------------------------------------------------- helpers
    def _find_noise_branches(self, st):
        """Branches whose only flow contributions are noise sources; a flow
        probe of one reads 0 in the large-signal analyses."""
        contrib = {}

        def walk(s):
            if s is None:
                return
            k = s[0]
            if k == "block":
                for x in s[3]:
                    walk(x)
            elif k == "if":
                walk(s[2])
                walk(s[3])
            elif k in ("while",):
                walk(s[2])
            elif k == "for":
                walk(s[4])
            elif k == "case":
                for _, x in s[2]:
                    walk(x)
            elif k == "event":
                walk(s[2])
            elif k == "contrib" and s[1] == "I" and len(s[2]) == 1 and s[2][0] in self.branches:
                e = s[3]
                only_noise = e[0] == "call" and e[1] in NOISE
                contrib.setdefault(s[2][0], []).append(only_noise)

        walk(st)
        self.noise_branches = {b for b, v in contrib.items() if all(v)}

    def fresh(self, base, ty):
        self.counter += 1
        nm = f"{base}__{self.counter}"
        self.vartype[nm] = ty
        return nm

    def resolve_param(self, name):
        name = self.aliases.get(name, name)
        return name if name in self.param_index else None

    def node_pair(self, args, what):
        names = []
        for a in args:
            if a[0] != "id":
                raise ParseError(f"{what}: node argument expected")
            names.append(a[1])
        if len(names) == 1 and names[0] in self.branches:
            a, b = self.branches[names[0]]
            names = [a] if b is None else [a, b]
        idx = []
        for n in names:
            if n in self.index:
                idx.append(self.index[n])
            elif n in ("gnd", "0"):
                idx.append(-1)
            else:
                raise ParseError(f"{what}: unknown node {n}")
        if len(idx) == 1:
            idx.append(-1)
        return idx[0], idx[1]

    # ------------------------------------------------------- expressions
    def expr(self, e, scope, pre):
        k = e[0]
        if k == "num":
            return e
        if k == "str":
            return e
        if k == "id":
            nm = e[1]
            if nm.startswith("$"):
                return self.expr(("call", nm, []), scope, pre)
            for sc in reversed(scope):
                if nm in sc:
                    return ("var", sc[nm])
            p = self.resolve_param(nm)
            if p is not None:
                return ("par", p)
            raise ParseError(f"unknown identifier {nm}")
        if k == "un":
            return ("un", e[1], self.expr(e[2], scope, pre))
        if k == "bin":
            op = e[1]
            a, b = self.expr(e[2], scope, pre), self.expr(e[3], scope, pre)
            if op == "**":
                return ("call", "pow", [a, b])
            return ("bin", op, a, b)
        if k == "tern":
            return ("tern", self.expr(e[1], scope, pre), self.expr(e[2], scope, pre), self.expr(e[3], scope, pre))
        if k == "call":
            fn, args = e[1], e[2]
            if fn == "V":
                a, b = self.node_pair(args, "V()")
                return ("probe", a, b)
            if fn == "I":
                if len(args) == 1 and args[0][0] == "id" and args[0][1] in self.noise_branches:
                    return ("zero",)
                raise ParseError("flow probes are supported only on noise-only branches")
            if fn in NOISE:
                return ("zero",)
            if fn == "ddt":
                return ("ddt", self.expr(args[0], scope, pre))
            if fn == "ddx":
                # ddx(x, V(n)): the partial of x in the potential of node n
                if len(args) != 2 or args[1][0] != "call" or args[1][1] != "V" or len(args[1][2]) != 1:
                    raise ParseError("ddx() takes an expression and V(node)")
                n, _ = self.node_pair(args[1][2], "ddx")
                return ("ddx", self.expr(args[0], scope, pre), n)
            if fn == "$temperature":
                return ("temp",)
            if fn == "$vt":
                t = ("temp",) if not args else self.expr(args[0], scope, pre)
                return ("bin", "/", ("bin", "*", ("num", 1.3806503e-23, False), t), ("num", 1.602176462e-19, False))
            if fn == "$simparam":
                what = args[0][1] if args and args[0][0] == "str" else None
                if what == "gmin":
                    return ("gmin",)
                if what == "tnom":
                    # the circuit's nominal temperature, degC: 27 unless a
                    # deck's .options tnom sets it (frontend/analyses.zig:397)
                    return ("num", 27.0, False)
                if len(args) > 1:
                    return self.expr(args[1], scope, pre)
                raise ParseError(f"$simparam({what}) without a default")
            if fn == "$mfactor":
                return ("mfactor",)
            if fn == "$abstime":
                return ("abstime",)
            if fn == "$param_given":
                p = self.resolve_param(args[0][1])
                if p is None:
                    raise ParseError(f"$param_given of unknown {args[0][1]}")
                return ("given", p)
            if fn == "$port_connected":
                return ("num", 1, True)
            if fn == "analysis":
                names = {a[1] for a in args if a[0] == "str"}
                return ("analysis", tuple(sorted(names)))
            if fn in ("$given",):
                raise ParseError(f"{fn} is not supported")
            if fn in MATH1 or fn in MATH2:
                return ("call", fn, [self.expr(a, scope, pre) for a in args])
            if fn in self.mod.functions:
                return self.inline(self.mod.functions[fn], args, scope, pre)
            if fn == "$limit":
                # lifted into a limit statement on a fresh variable, read here
                a, b = self.node_pair(args[0][2], "$limit") if args[0][0] == "call" else (None, None)
                lf = args[1][1] if len(args) > 1 and args[1][0] == "str" else None
                extra = [self.expr(x, scope, pre) for x in args[2:]]
                tmp = self.fresh("limit", "real")
                pre.append(("limit", tmp, lf, ("probe", a, b), extra))
                return ("var", tmp)
            raise ParseError(f"unsupported function {fn}")
        raise ParseError(f"unsupported expression {k}")

    def inline(self, f, args, scope, pre):
        """An analog function call: its body is emitted into `pre` (the
        statements before the one that holds the call) with its arguments
        and locals renamed apart; the call reads the result variable."""
        self.inline_depth += 1
        if self.inline_depth > 16:
            raise ParseError("analog functions nest too deep")
        local = {}
        for nm, ty in f["types"].items():
            local[nm] = self.fresh(f["name"] + "_" + nm, ty)
        local[f["name"]] = self.fresh(f["name"] + "_ret", f["rtype"])
        outs = []
        for a, actual in zip(f["args"], args):
            d = f["dirs"][a]
            if a not in local:
                local[a] = self.fresh(f["name"] + "_" + a, "real")
            if d in ("input", "inout"):
                pre.append(("assign", local[a], self.expr(actual, scope, pre)))
            if d in ("output", "inout"):
                if actual[0] != "id":
                    raise ParseError("an output argument must be a variable")
                outs.append((local[a], actual))
        body = self.stmt(f["body"], [local])
        pre.extend(body)
        for src, actual in outs:
            tgt = self.expr(actual, scope, pre)
            pre.append(("assign", tgt[1], ("var", src)))
        self.inline_depth -= 1
        return ("var", local[f["name"]])

    # -------------------------------------------------------- statements
    def stmt(self, s, scope):
        """Returns a list of IR statements."""
        if s is None:
            return []
        k = s[0]
        if k == "nop":
            return []
        if k == "block":
            sc = {}
            for nm, ty in s[2].items():
                sc[nm] = self.fresh(nm, ty)
            out = []
            for x in s[3]:
                out.extend(self.stmt(x, scope + [sc]))
            return out
        if k == "if":
            pre = []
            c = self.expr(s[1], scope, pre)
            return pre + [("if", c, self.stmt(s[2], scope), self.stmt(s[3], scope))]
        if k == "case":
            pre = []
            e = self.expr(s[1], scope, pre)
            chain = []
            default = []
            for labels, body in s[2]:
                if labels is None:
                    default = self.stmt(body, scope)
                    continue
                cond = None
                for lab in labels:
                    c = ("bin", "==", e, self.expr(lab, scope, pre))
                    cond = c if cond is None else ("bin", "||", cond, c)
                chain.append((cond, self.stmt(body, scope)))
            out = default
            for cond, body in reversed(chain):
                out = [("if", cond, body, out)]
            return pre + out
        if k == "for":
            init = self.stmt(s[1], scope)
            pre = []
            c = self.expr(s[2], scope, pre)
            if pre:
                raise ParseError("function calls in a loop condition are not supported")
            body = self.stmt(s[4], scope) + self.stmt(s[3], scope)
            return init + [("while", c, body)]
        if k == "while":
            pre = []
            c = self.expr(s[1], scope, pre)
            if pre:
                raise ParseError("function calls in a loop condition are not supported")
            return [("while", c, self.stmt(s[2], scope))]
        if k == "event":
            # @(initial_step) and the like: the large-signal run executes the
            # guarded statements at its first evaluation; their results only
            # ever set up constants
            return self.stmt(s[2], scope)
        if k == "assign":
            nm = s[1]
            tgt = None
            for sc in reversed(scope):
                if nm in sc:
                    tgt = sc[nm]
                    break
            if tgt is None:
                raise ParseError(f"assignment to unknown variable {nm}")
            pre = []
            e = s[2]
            if e[0] == "call" and e[1] == "$limit":
                a, b = self.node_pair(e[2][0][2], "$limit") if e[2][0][0] == "call" else (None, None)
                fn = e[2][1][1] if e[2][1][0] == "str" else None
                extra = [self.expr(x, scope, pre) for x in e[2][2:]]
                return pre + [("limit", tgt, fn, ("probe", a, b), extra)]
            return pre + [("assign", tgt, self.expr(e, scope, pre))]
        if k == "contrib":
            kind, nodes, e = s[1], s[2], s[3]
            a, b = self.node_pair([("id", n) for n in nodes], "contribution")
            if kind == "V":
                if e[0] == "num" and e[1] == 0:
                    return [("collapse", a, b)]
                raise ParseError("only V(a, b) <+ 0 (a collapse) is supported as a potential contribution")
            if kind != "I":
                raise ParseError(f"contribution of {kind} is not supported")
            if len(nodes) == 1 and nodes[0] in self.noise_branches:
                return []
            pre = []
            ex = self.expr(e, scope, pre)
            return pre + [("contrib", a, b, ex)]
        if k == "task":
            if s[1] in IGNORED_TASKS:
                return []
            raise ParseError(f"unsupported task {s[1]}")
        raise ParseError(f"unsupported statement {k}")
