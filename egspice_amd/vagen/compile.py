"""Analyses and C emission of a normalised Verilog-A model (ir.Model).

1. Types: integer or real for every variable and parameter, and for every
   expression by the LRM's rules (integer op integer stays integer).
2. Bias dependence: a variable is *dynamic* when any assignment to it
   reads a probe, $abstime or a dynamic variable, or sits under a dynamic
   condition; otherwise *static* (a function of the parameters and the
   temperature only). Static statements move to the setup function, which
   runs once per instance and lane; a static value read by the evaluation
   is stored in the instance's prepared constants. A static variable read
   by a dynamic statement and assigned again after it becomes dynamic, so
   moving the static statements never changes a value.
3. Relevance: only statements whose results reach a contribution (or a
   collapse) are kept, separately for the static evaluation (currents), the
   transient one (currents and charges) and the charge values alone.
4. Derivative lanes: each dynamic real carries the partial derivatives
   with respect to the branch voltages it depends on, and only those
   (sparse forward-mode AD): every distinct probed pair of nodes {a, b} is
   one lane (V(a, b) and V(b, a) are the same lane with opposite signs), a
   variable's lanes the union over its assignments; a contribution's
   Jacobian on node columns follows by the chain rule at the stamp
   (d/dx_a = +d_p, d/dx_b = -d_p). A compact model's core reads three or
   four branch voltages where it would carry four or more node lanes.
5. Emission: three-address C with one temporary per operation; the
   derivative rules follow the reference's dual numbers
   (src/device/eval.zig:107-350): product rule and quotient numerator
   fused (fma), quotient lanes scaled by the reciprocal, sqrt and pow slopes
   with their zero and non-finite guards, map(a, f(a.v), f'(a.v)) for the
   elementary functions.
"""
from .ir import Model
from .parse import ParseError

CMP = {"<", "<=", ">", ">=", "==", "!="}
LOGIC = {"&&", "||"}


def walk_expr(e, fn):
    fn(e)
    k = e[0]
    if k == "un":
        walk_expr(e[2], fn)
    elif k == "bin":
        walk_expr(e[2], fn)
        walk_expr(e[3], fn)
    elif k == "tern":
        walk_expr(e[1], fn)
        walk_expr(e[2], fn)
        walk_expr(e[3], fn)
    elif k == "call":
        for a in e[2]:
            walk_expr(a, fn)
    elif k in ("ddt", "ddx"):
        walk_expr(e[1], fn)


def expr_reads(e):
    vs, ps = set(), set()

    def f(x):
        if x[0] == "var":
            vs.add(x[1])
        elif x[0] in ("par", "given"):
            ps.add(x[1])

    walk_expr(e, f)
    return vs, ps


def has_ddt(e):
    found = []
    walk_expr(e, lambda x: found.append(1) if x[0] == "ddt" else None)
    return bool(found)


def split_ddt(e, sign=1):
    """A contribution's expression as (resistive part or None, [(sign, q,
    factor)]): q is the charge of one ddt() term (times its constant factor,
    or factor None)."""
    k = e[0]
    if k == "ddt":
        if has_ddt(e[1]):
            raise ParseError("nested ddt")
        return None, [(sign, e[1], None)]
    if not has_ddt(e):
        return (e if sign > 0 else ("un", "-", e)), []
    if k == "un" and e[1] == "-":
        return split_ddt(e[2], -sign)
    if k == "bin" and e[1] in ("+", "-"):
        ra, qa = split_ddt(e[2], sign)
        rb, qb = split_ddt(e[3], sign if e[1] == "+" else -sign)
        if ra is None:
            r = rb
        elif rb is None:
            r = ra
        else:
            r = ("bin", "+", ra, rb)
        return r, qa + qb
    if k == "bin" and e[1] == "*" and (e[2][0] == "ddt") != (e[3][0] == "ddt"):
        # k * ddt(q) = ddt(k * q) for a time-invariant k (checked by the
        # compiler: no probe, no state)
        kf, d = (e[3], e[2]) if e[2][0] == "ddt" else (e[2], e[3])
        if has_ddt(kf) or has_ddt(d[1]):
            raise ParseError("nested ddt")
        return None, [(sign, ("bin", "*", kf, d[1]), kf)]
    raise ParseError("ddt() must be an added term of a contribution")


def map_expr(e, fn):
    """e with every ('var', name) node replaced by fn(name)."""
    k = e[0]
    if k == "var":
        return fn(e[1])
    if k == "un":
        return (k, e[1], map_expr(e[2], fn))
    if k == "bin":
        return (k, e[1], map_expr(e[2], fn), map_expr(e[3], fn))
    if k == "tern":
        return (k, map_expr(e[1], fn), map_expr(e[2], fn), map_expr(e[3], fn))
    if k == "call":
        return (k, e[1], [map_expr(a, fn) for a in e[2]])
    if k == "ddt":
        return (k, map_expr(e[1], fn))
    if k == "ddx":
        return (k, map_expr(e[1], fn), e[2])
    return e


def split_webs(body, vartype):
    """Renames every variable apart into its webs (the definitions that
    reach a common use, closed under sharing a use), so that a temporary
    reused for parameter arithmetic and later for bias-dependent values is
    two variables, and the bias-dependence split sees each on its own.
    Reaching definitions over the structured body (if: both branches
    merge; while: to a fixed point over the back edge); a variable read
    before any assignment reaches from its initial value. Returns the new
    body; vartype gains the new names."""
    parent = {}

    def find(d):
        while parent[d] != d:
            parent[d] = parent[parent[d]]
            d = parent[d]
        return d

    def union(a, b):
        a, b = find(a), find(b)
        if a != b:
            parent[max(a, b)] = min(a, b)

    defs = {}  # id(statement) -> def id
    counter = [0]

    def new_def(key):
        if key not in defs:
            defs[key] = counter[0]
            parent[counter[0]] = counter[0]
            counter[0] += 1
        return defs[key]

    def init_def(v):
        return new_def(("init", v))

    def reads(e):
        vs, _ = expr_reads(e)
        return vs

    def use(v, reach):
        ds = reach.get(v)
        if ds is None:
            ds = {init_def(v)}
        it = iter(ds)
        first = next(it)
        for d in it:
            union(first, d)
        return first

    def merge(a, b):
        out = dict(a)
        for v, ds in b.items():
            out[v] = out.get(v, set()) | ds if v in out else set(ds)
        for v in a:
            if v not in b:
                out[v] = set(a[v]) | {init_def(v)}
        for v in b:
            if v not in a:
                out[v] = out[v] | {init_def(v)}
        return out

    def walk(stmts, reach, rename):
        """One pass; with rename, returns the rewritten statements too."""
        out = []
        for s in stmts:
            k = s[0]

            def rx(e):
                for v in reads(e):
                    use(v, reach)
                if rename:
                    return map_expr(e, lambda v: ("var", name_of(v, find(use(v, reach)))))
                return e

            if k == "assign":
                e = rx(s[2])
                d = new_def(("d", id(s)))
                reach[s[1]] = {d}
                out.append(("assign", name_of(s[1], find(d)) if rename else s[1], e))
            elif k == "limit":
                pr = rx(s[3])
                args = [rx(a) for a in s[4]]
                d = new_def(("d", id(s)))
                reach[s[1]] = {d}
                out.append(("limit", name_of(s[1], find(d)) if rename else s[1], s[2], pr, args))
            elif k == "contrib":
                out.append(("contrib", s[1], s[2], rx(s[3])))
            elif k == "collapse":
                out.append(s)
            elif k == "if":
                c = rx(s[1])
                r1 = dict((v, set(ds)) for v, ds in reach.items())
                r2 = dict((v, set(ds)) for v, ds in reach.items())
                t = walk(s[2], r1, rename)
                f = walk(s[3], r2, rename)
                reach.clear()
                reach.update(merge(r1, r2))
                out.append(("if", c, t, f))
            elif k == "while":
                head = dict((v, set(ds)) for v, ds in reach.items())
                while True:
                    body_in = dict((v, set(ds)) for v, ds in head.items())
                    for v in reads(s[1]):
                        use(v, body_in)
                    walk(s[2], body_in, False)
                    nxt = merge(head, body_in)
                    if all(nxt.get(v) == head.get(v) for v in set(nxt) | set(head)):
                        break
                    head = nxt
                reach.clear()
                reach.update(head)
                if rename:
                    c = map_expr(s[1], lambda v: ("var", name_of(v, find(use(v, dict(head))))))
                    inner = dict((v, set(ds)) for v, ds in head.items())
                    b = walk(s[2], inner, True)
                    out.append(("while", c, b))
                else:
                    out.append(s)
            else:
                raise ParseError(f"statement {k}")
        return out

    names = {}

    def name_of(v, web):
        key = (v, web)
        if key not in names:
            taken = sum(1 for (w, _) in names if w == v)
            names[key] = v if taken == 0 else f"{v}__w{taken}"
            vartype[names[key]] = vartype[v]
        return names[key]

    walk(body, {}, False)
    return walk(body, {}, True)


class Stmt:
    """A statement with its place in program order."""
    __slots__ = ("kind", "a", "pos", "end", "conds", "loops", "kids")

    def __init__(self, kind, a, pos, conds, loops):
        self.kind, self.a, self.pos, self.conds, self.loops = kind, a, pos, conds, loops
        self.kids = []
        self.end = pos


class Compiler:
    def __init__(self, model: Model, prefix):
        self.m = model
        self.prefix = prefix
        self.nu = len(model.unknowns)
        self.flat = []  # every leaf statement (assign/contrib/collapse) as Stmt
        self.pos = 0
        self.body = split_webs(model.body, model.vartype)
        self.tree = self.index(self.body, (), ())
        self.types()
        self.probes()
        self.classify()
        self.relevance()
        self.lanes()

    # ------------------------------------------------------------ index
    def index(self, stmts, conds, loops):
        out = []
        for s in stmts:
            k = s[0]
            self.pos += 1
            st = Stmt(k, s, self.pos, conds, loops)
            if k == "if":
                st.kids = [self.index(s[2], conds + (s[1],), loops), self.index(s[3], conds + (s[1],), loops)]
            elif k == "while":
                st.kids = [self.index(s[2], conds + (s[1],), loops + (st,))]
            else:
                self.flat.append(st)
            st.end = self.pos
            out.append(st)
        return out

    # ------------------------------------------------------------ types
    def types(self):
        self.ptype = {p["name"]: ("int" if p["type"] == "integer" else "real") for p in self.m.params}
        self.vtype = {v: ("int" if t == "integer" else "real") for v, t in self.m.vartype.items()}

    def etype(self, e):
        k = e[0]
        if k == "num":
            return "int" if e[2] else "real"
        if k == "par":
            return self.ptype[e[1]]
        if k == "var":
            return self.vtype[e[1]]
        if k in ("given", "analysis"):
            return "int"
        if k == "un":
            return "int" if e[1] == "!" else self.etype(e[2])
        if k == "bin":
            if e[1] in CMP or e[1] in LOGIC:
                return "int"
            ta, tb = self.etype(e[2]), self.etype(e[3])
            return "int" if ta == tb == "int" else "real"
        if k == "tern":
            ta, tb = self.etype(e[2]), self.etype(e[3])
            return "int" if ta == tb == "int" else "real"
        if k == "call":
            if e[1] in ("min", "max", "abs"):
                return "int" if all(self.etype(a) == "int" for a in e[2]) else "real"
            return "real"
        return "real"

    # ---------------------------------------------------- bias dependence
    def expr_dyn(self, e):
        hit = []

        def f(x):
            k = x[0]
            if k in ("probe", "abstime", "ddt", "ddx"):
                hit.append(1)
            elif k == "var" and x[1] in self.dyn:
                hit.append(1)

        walk_expr(e, f)
        return bool(hit)

    def assigned(self, st):
        k = st.kind
        if k == "assign":
            return st.a[1]
        if k == "limit":
            return st.a[1]
        return None

    def stmt_reads(self, st):
        """Expressions a leaf statement reads (conditions included)."""
        k = st.kind
        es = list(st.conds)
        if k == "assign":
            es.append(st.a[2])
        elif k == "contrib":
            es.append(st.a[3])
        elif k == "limit":
            es.append(st.a[3])
        return es

    def classify(self):
        self.dyn = set()
        changed = True
        while changed:
            changed = False
            for st in self.flat:
                v = self.assigned(st)
                if v is None or v in self.dyn:
                    continue
                if st.kind == "limit" or any(self.expr_dyn(e) for e in self.stmt_reads(st)):
                    self.dyn.add(v)
                    changed = True
            if changed:
                continue
            # hazard: a static variable read by a dynamic statement and
            # assigned (statically) later, or inside a loop around the read
            assigns = {}
            for st in self.flat:
                v = self.assigned(st)
                if v is not None and v not in self.dyn:
                    assigns.setdefault(v, []).append(st)
            for st in self.flat:
                if not self.is_dynamic_stmt(st):
                    continue
                for e in self.stmt_reads(st) + (list(st.a[4]) if st.kind == "limit" else []):
                    vs, _ = expr_reads(e)
                    for v in vs:
                        if v in self.dyn or v not in assigns:
                            continue
                        for a in assigns[v]:
                            if a.pos > st.pos or any(lp.pos <= a.pos <= lp.end for lp in st.loops):
                                self.dyn.add(v)
                                changed = True
                                break

    def is_dynamic_stmt(self, st):
        if st.kind in ("contrib",):
            return True
        if st.kind == "collapse":
            return False
        v = self.assigned(st)
        return v in self.dyn

    # ------------------------------------------------------------ probes
    def probes(self):
        """One derivative lane per distinct probed node pair, canonical
        orientation (smaller index first; ground is -1)."""
        self.probe_id = {}
        self.probe_pair = []

        def f(e):
            if e[0] == "probe" and not (e[1] < 0 and e[2] < 0):
                key = (min(e[1], e[2]), max(e[1], e[2]))
                if key not in self.probe_id:
                    self.probe_id[key] = len(self.probe_pair)
                    self.probe_pair.append(key)

        for st in self.flat:
            for e in self.stmt_reads(st):
                walk_expr(e, f)

    def probe_lane(self, a, b):
        """(lane, sign) of V(a, b)."""
        key = (min(a, b), max(a, b))
        return self.probe_id[key], (1.0 if (a, b) == key else -1.0)

    # ---------------------------------------------------------- relevance
    def relevance(self):
        self.contribs = []
        for st in self.flat:
            if st.kind == "contrib":
                r, qs = split_ddt(st.a[3])
                for _, _, kf in qs:
                    if kf is not None and self.expr_dyn(kf):
                        raise ParseError("a bias-dependent factor of ddt()")
                self.contribs.append((st, r, qs))
        self.rel = {}
        self.limits = [st for st in self.flat if st.kind == "limit"]
        for mode in ("dc", "tran", "charge", "limitargs"):
            need = set()
            seeds = []
            if mode == "limitargs":
                # the limiter's arguments (evaluated at the limiter's own
                # iterate, values only)
                for st in self.limits:
                    seeds.extend(st.conds)
                    seeds.extend(st.a[4])
            for st, r, qs in self.contribs:
                if mode == "limitargs" or (mode == "charge" and not qs):
                    continue
                seeds.extend(st.conds)
                if r is not None and mode != "charge":
                    seeds.append(r)
                if mode != "dc":
                    seeds.extend(q for _, q, _ in qs)
            for st in self.flat:
                if st.kind == "collapse" and mode != "limitargs":
                    seeds.extend(st.conds)
            work = []
            for e in seeds:
                vs, _ = expr_reads(e)
                work.extend(vs)
            by_var = {}
            for st in self.flat:
                v = self.assigned(st)
                if v is not None:
                    by_var.setdefault(v, []).append(st)
            while work:
                v = work.pop()
                if v in need:
                    continue
                need.add(v)
                for st in by_var.get(v, []):
                    for e in self.stmt_reads(st):
                        vs, _ = expr_reads(e)
                        work.extend(vs - need)
            self.rel[mode] = need
        # does a ddx() sit on a statement a mode keeps?
        self.ddx_in = {}
        for mode in self.rel:
            hit = False
            for st in self.flat:
                v = self.assigned(st)
                if v is None or v not in self.rel[mode]:
                    continue
                for e in self.stmt_reads(st):
                    walk_expr(e, lambda x: None)
                    found = []
                    walk_expr(e, lambda x: found.append(1) if x[0] == "ddx" else None)
                    hit = hit or bool(found)
            self.ddx_in[mode] = hit

    # -------------------------------------------------------------- lanes
    def lanes(self):
        self.vl = {v: frozenset() for v in self.dyn}
        changed = True
        while changed:
            changed = False
            for st in self.flat:
                v = self.assigned(st)
                if v is None or v not in self.dyn or self.vtype.get(v) == "int":
                    continue
                if st.kind == "limit":
                    l = self.elanes(st.a[3])
                else:
                    l = self.elanes(st.a[2])
                if not l <= self.vl[v]:
                    self.vl[v] = self.vl[v] | l
                    changed = True

    def elanes(self, e):
        k = e[0]
        if k == "probe":
            if e[1] < 0 and e[2] < 0:
                return frozenset()
            return frozenset([self.probe_lane(e[1], e[2])[0]])
        if k == "var":
            if self.vtype.get(e[1]) == "int":
                return frozenset()
            return self.vl.get(e[1], frozenset())
        if k == "un":
            return frozenset() if e[1] == "!" else self.elanes(e[2])
        if k == "bin":
            if e[1] in CMP or e[1] in LOGIC:
                return frozenset()
            return self.elanes(e[2]) | self.elanes(e[3])
        if k == "tern":
            return self.elanes(e[2]) | self.elanes(e[3])
        if k == "call":
            if e[1] in ("floor", "ceil"):
                return frozenset()
            out = frozenset()
            for a in e[2]:
                out |= self.elanes(a)
            return out
        if k == "ddt":
            return self.elanes(e[1])
        # ddx() is a value: the derivative lanes stop there
        return frozenset()
