"""Lexer and parser of the Verilog-A subset compact models are written in
(Verilog-AMS LRM 2.4: modules with electrical ports and internal nodes,
branches, parameters with ranges, aliasparam, real/integer variables,
analog functions, and one analog block of begin/end blocks, if/else,
case, for, while, assignments, contributions and system tasks).

The tree is made of tuples:
  expressions  ('num', value, is_int) ('str', s) ('id', name)
               ('call', name, [args]) ('un', op, e) ('bin', op, a, b)
               ('tern', c, a, b)
  statements   ('block', name, decls, [stmts]) ('if', c, then, else)
               ('case', e, [(labels or None, stmt)]) ('for', init, c, step, body)
               ('while', c, body) ('assign', name, e) ('contrib', kind, nodes, e)
               ('task', name, [args]) ('nop',)
"""
import re

KEYWORDS = {
    "module", "endmodule", "macromodule", "analog", "begin", "end", "if", "else", "case", "endcase", "default",
    "for", "while", "repeat", "parameter", "localparam", "real", "integer", "string", "from", "exclude", "inf",
    "branch", "inout", "input", "output", "electrical", "thermal", "ground", "aliasparam", "function",
    "endfunction", "genvar", "or",
}

_tok = re.compile(r"""
    (?P<ws>\s+)
  | (?P<attr>\(\*.*?\*\))
  | (?P<num>(\d+\.\d*|\.\d+|\d+)([eE][+-]?\d+)?[TGMKkmunpfa]?(?![A-Za-z_0-9]))
  | (?P<str>"(\\.|[^"\\])*")
  | (?P<id>\$?[A-Za-z_][A-Za-z0-9_$]*)
  | (?P<op><\+|<=|>=|==|!=|&&|\|\||\*\*|<<|>>|[-+*/%<>!~&|^?:;,.()\[\]{}=#@])
""", re.S | re.X)

SI = {"T": 1e12, "G": 1e9, "M": 1e6, "K": 1e3, "k": 1e3, "m": 1e-3, "u": 1e-6, "n": 1e-9, "p": 1e-12,
      "f": 1e-15, "a": 1e-18}


class ParseError(Exception):
    pass


def lex(text):
    toks = []
    pos, line = 0, 1
    n = len(text)
    while pos < n:
        m = _tok.match(text, pos)
        if not m:
            raise ParseError(f"line {line}: cannot read {text[pos:pos + 30]!r}")
        k = m.lastgroup
        s = m.group(0)
        if k == "num":
            body = s
            scale = None
            if body[-1] in SI and not body[-1].isdigit():
                scale = SI[body[-1]]
                body = body[:-1]
            is_int = scale is None and re.fullmatch(r"\d+", body) is not None
            if is_int:
                toks.append(("num", (int(body), True), line))
            else:
                v = float(body)
                if scale is not None:
                    # the scale factor multiplies the mantissa: 1.5n = 1.5e-9
                    v = float(body + "e" + str(int(round(__import__("math").log10(scale)))))
                toks.append(("num", (v, False), line))
        elif k == "str":
            toks.append(("str", s[1:-1], line))
        elif k == "id":
            toks.append(("kw" if s in KEYWORDS else "id", s, line))
        elif k == "op":
            toks.append(("op", s, line))
        elif k == "attr":
            toks.append(("attr", s, line))
        line += s.count("\n")
        pos = m.end()
    toks.append(("eof", None, line))
    return toks


# binary operator precedence, lowest first (LRM 4.2)
BINPREC = [
    ["||"], ["&&"], ["|"], ["^"], ["&"], ["==", "!="], ["<", "<=", ">", ">="], ["<<", ">>"], ["+", "-"],
    ["*", "/", "%"], ["**"],
]


class Module:
    def __init__(self, name):
        self.name = name
        self.ports = []
        self.nodes = []          # every electrical node, declaration order (ports too)
        self.branches = {}       # name -> (a, b or None)
        self.params = []         # dicts: name, type, default, ranges, instance
        self.aliases = {}        # alias -> param
        self.vars = {}           # module-level variables: name -> type
        self.functions = {}      # analog functions
        self.analog = None


class Parser:
    def __init__(self, toks):
        self.t = toks
        self.i = 0

    # -------------------------------------------------------------- tokens
    def peek(self, k=0):
        return self.t[self.i + k]

    def next(self):
        tok = self.t[self.i]
        self.i += 1
        return tok

    def at(self, val, k=0):
        tok = self.t[self.i + k]
        return tok[1] == val and tok[0] in ("op", "kw")

    def accept(self, val):
        if self.at(val):
            self.i += 1
            return True
        return False

    def expect(self, val):
        tok = self.next()
        if tok[1] != val:
            raise ParseError(f"line {tok[2]}: expected {val!r}, got {tok[1]!r}")
        return tok

    def ident(self):
        tok = self.next()
        if tok[0] != "id":
            raise ParseError(f"line {tok[2]}: expected an identifier, got {tok[1]!r}")
        return tok[1]

    # -------------------------------------------------------------- module
    def parse(self):
        mods = []
        while self.peek()[0] != "eof":
            if self.at("module") or self.at("macromodule"):
                mods.append(self.module())
            elif self.peek()[0] == "kw" and self.peek()[1] in ("electrical",):
                self.next()
            else:
                # discipline/nature declarations outside modules are skipped
                self.next()
        return mods

    def module(self):
        self.next()
        m = Module(self.ident())
        if self.accept("("):
            if not self.at(")"):
                m.ports.append(self.ident())
                while self.accept(","):
                    m.ports.append(self.ident())
            self.expect(")")
        self.expect(";")
        while not self.accept("endmodule"):
            self.module_item(m)
        return m

    def module_item(self, m):
        tok = self.peek()
        v = tok[1]
        if v in ("inout", "input", "output"):
            self.next()
            if self.peek()[1] in ("electrical", "thermal"):
                self.next()
            self.ident_list()
            self.expect(";")
        elif v in ("electrical", "thermal"):
            self.next()
            for nm in self.ident_list():
                if nm not in m.nodes:
                    m.nodes.append(nm)
                if v == "thermal":
                    m.thermal = getattr(m, "thermal", set()) | {nm}
            self.expect(";")
        elif v == "ground":
            self.next()
            self.ident_list()
            self.expect(";")
        elif v == "branch":
            self.next()
            self.expect("(")
            a = self.ident()
            b = None
            if self.accept(","):
                b = self.ident()
            self.expect(")")
            for nm in self.ident_list():
                m.branches[nm] = (a, b)
            self.expect(";")
        elif v in ("parameter", "localparam"):
            self.next()
            self.param_decl(m, local=(v == "localparam"))
        elif v == "aliasparam":
            self.next()
            a = self.ident()
            self.expect("=")
            p = self.ident()
            self.expect(";")
            m.aliases[a] = p
        elif v in ("real", "integer", "string"):
            self.next()
            for nm, init in self.var_list():
                m.vars[nm] = v
            self.expect(";")
        elif v == "genvar":
            self.next()
            self.ident_list()
            self.expect(";")
        elif v == "analog":
            self.next()
            if self.at("function"):
                self.function(m)
            else:
                if m.analog is not None:
                    raise ParseError("more than one analog block")
                m.analog = self.statement()
        elif tok[0] == "attr":
            self.next()
            if re.search(r'type\s*=\s*"instance"', v):
                self.instance_attr = True
            return
        else:
            raise ParseError(f"line {tok[2]}: unexpected {v!r} in module")
        self.instance_attr = False

    def ident_list(self):
        names = [self.ident()]
        while self.accept(","):
            names.append(self.ident())
        return names

    def var_list(self):
        out = []
        while True:
            nm = self.ident()
            if self.accept("["):
                raise ParseError("array variables are not supported")
            init = None
            if self.accept("="):
                init = self.expr()
            out.append((nm, init))
            if not self.accept(","):
                return out

    def param_decl(self, m, local=False):
        ptype = "real"
        if self.peek()[1] in ("real", "integer", "string"):
            ptype = self.next()[1]
        while True:
            name = self.ident()
            self.expect("=")
            default = self.expr()
            ranges = []
            while self.at("from") or self.at("exclude"):
                kind = self.next()[1]
                if self.at("[") or self.at("("):
                    lo_open = self.next()[1] == "("
                    lo = self.expr()
                    self.expect(":")
                    hi = self.expr()
                    hi_tok = self.next()[1]
                    if hi_tok not in ("]", ")"):
                        raise ParseError(f"bad range end {hi_tok!r}")
                    ranges.append((kind, lo, hi, lo_open, hi_tok == ")"))
                elif self.at("'"):
                    raise ParseError("set ranges are not supported")
                else:
                    ranges.append((kind, self.expr(), None, False, False))
            m.params.append(dict(name=name, type=ptype, default=default, ranges=ranges,
                                 instance=getattr(self, "instance_attr", False), local=local))
            if not self.accept(","):
                break
        self.expect(";")

    def function(self, m):
        self.expect("function")
        rtype = "real"
        if self.peek()[1] in ("real", "integer"):
            rtype = self.next()[1]
        name = self.ident()
        self.expect(";")
        args, dirs, types = [], {}, {}
        while not (self.at("begin") or self.peek()[0] == "id" or self.at("if") or self.at("case")):
            d = self.next()[1]
            if d in ("input", "output", "inout"):
                for a in self.ident_list():
                    args.append(a)
                    dirs[a] = d
                self.expect(";")
            elif d in ("real", "integer"):
                for a, _ in self.var_list():
                    types[a] = d
                self.expect(";")
            else:
                raise ParseError(f"function {name}: unexpected {d!r}")
        body = self.statement()
        self.expect("endfunction")
        m.functions[name] = dict(name=name, rtype=rtype, args=args, dirs=dirs, types=types, body=body)

    # ---------------------------------------------------------- statements
    def statement(self):
        tok = self.peek()
        v = tok[1]
        if tok[0] == "attr":
            self.next()
            return self.statement()
        if self.accept(";"):
            return ("nop",)
        if self.at("begin"):
            self.next()
            name = None
            if self.accept(":"):
                name = self.ident()
            decls = {}
            body = []
            while not self.accept("end"):
                if self.peek()[1] in ("real", "integer", "string") and self.peek()[0] == "kw":
                    ty = self.next()[1]
                    for nm, init in self.var_list():
                        decls[nm] = ty
                        if init is not None:
                            body.append(("assign", nm, init))
                    self.expect(";")
                elif self.peek()[0] == "attr":
                    self.next()
                else:
                    body.append(self.statement())
            return ("block", name, decls, body)
        if self.at("if"):
            self.next()
            self.expect("(")
            c = self.expr()
            self.expect(")")
            th = self.statement()
            el = None
            if self.accept("else"):
                el = self.statement()
            return ("if", c, th, el)
        if self.at("case"):
            self.next()
            self.expect("(")
            e = self.expr()
            self.expect(")")
            items = []
            while not self.accept("endcase"):
                if self.accept("default"):
                    self.accept(":")
                    items.append((None, self.statement()))
                else:
                    labels = [self.expr()]
                    while self.accept(","):
                        labels.append(self.expr())
                    self.expect(":")
                    items.append((labels, self.statement()))
            return ("case", e, items)
        if self.at("for"):
            self.next()
            self.expect("(")
            init = self.simple_assign()
            self.expect(";")
            c = self.expr()
            self.expect(";")
            step = self.simple_assign()
            self.expect(")")
            return ("for", init, c, step, self.statement())
        if self.at("while"):
            self.next()
            self.expect("(")
            c = self.expr()
            self.expect(")")
            return ("while", c, self.statement())
        if self.at("repeat"):
            raise ParseError("repeat is not supported")
        if self.at("@"):
            # event control: @(initial_step[("tran")]) stmt
            self.next()
            self.expect("(")
            ev = self.expr()
            self.expect(")")
            return ("event", ev, self.statement())
        if tok[0] == "id" and tok[1].startswith("$"):
            name = self.next()[1]
            args = []
            if self.accept("("):
                if not self.at(")"):
                    args.append(self.expr())
                    while self.accept(","):
                        args.append(self.expr())
                self.expect(")")
            self.expect(";")
            return ("task", name, args)
        if tok[0] == "id" and self.at("(", 1):
            # a contribution: I(a, b) <+ e; V(br) <+ e;
            kind = self.next()[1]
            self.expect("(")
            nodes = [self.ident()]
            while self.accept(","):
                nodes.append(self.ident())
            self.expect(")")
            self.expect("<+")
            e = self.expr()
            self.expect(";")
            return ("contrib", kind, nodes, e)
        if tok[0] == "id":
            s = self.simple_assign()
            self.expect(";")
            return s
        raise ParseError(f"line {tok[2]}: unexpected {v!r}")

    def simple_assign(self):
        name = self.ident()
        self.expect("=")
        return ("assign", name, self.expr())

    # --------------------------------------------------------- expressions
    def expr(self):
        c = self.binary(0)
        if self.accept("?"):
            a = self.expr()
            self.expect(":")
            b = self.expr()
            return ("tern", c, a, b)
        return c

    def binary(self, level):
        if level == len(BINPREC):
            return self.unary()
        left = self.binary(level + 1)
        ops = BINPREC[level]
        while self.peek()[0] == "op" and self.peek()[1] in ops:
            op = self.next()[1]
            if op == "**":
                right = self.binary(level)  # right associative
            else:
                right = self.binary(level + 1)
            left = ("bin", op, left, right)
        return left

    def unary(self):
        if self.peek()[0] == "op" and self.peek()[1] in ("-", "+", "!", "~"):
            op = self.next()[1]
            e = self.unary()
            if op == "+":
                return e
            if op == "-" and e[0] == "num":
                return ("num", -e[1], e[2])
            return ("un", op, e)
        return self.primary()

    def primary(self):
        tok = self.next()
        if tok[0] == "num":
            return ("num", tok[1][0], tok[1][1])
        if tok[0] == "str":
            return ("str", tok[1])
        if tok[0] == "op" and tok[1] == "(":
            e = self.expr()
            self.expect(")")
            return e
        if tok[0] == "kw" and tok[1] == "inf":
            return ("num", float("inf"), False)
        if tok[0] == "id":
            if self.accept("("):
                args = []
                if not self.at(")"):
                    args.append(self.expr())
                    while self.accept(","):
                        args.append(self.expr())
                self.expect(")")
                return ("call", tok[1], args)
            return ("id", tok[1])
        raise ParseError(f"line {tok[2]}: unexpected {tok[1]!r} in an expression")


def parse(text):
    mods = Parser(lex(text)).parse()
    if len(mods) != 1:
        raise ParseError(f"expected one module, found {len(mods)}")
    return mods[0]
