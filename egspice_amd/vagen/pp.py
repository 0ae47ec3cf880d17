"""Verilog-A preprocessor: comments, `define/`undef (object-like and with
arguments), `ifdef/`ifndef/`elsif/`else/`endif and `include of the
standard headers (constants.vams, disciplines.vams; any other include must
be found beside the source). Macro bodies are expanded where they are used,
and the result is scanned again, as the Verilog-AMS LRM (2.4, chapter 10)
prescribes."""
import os
import re

# Verilog-AMS LRM 2.4, Annex D (constants.vams): the mathematical and
# physical constants a model may name as `M_* or `P_*.
STANDARD_MACROS = {
    "M_E": "2.7182818284590452354",
    "M_LOG2E": "1.4426950408889634074",
    "M_LOG10E": "0.43429448190325182765",
    "M_LN2": "0.69314718055994530942",
    "M_LN10": "2.30258509299404568402",
    "M_PI": "3.14159265358979323846",
    "M_TWO_PI": "6.28318530717958647652",
    "M_PI_2": "1.57079632679489661923",
    "M_PI_4": "0.78539816339744830962",
    "M_1_PI": "0.31830988618379067154",
    "M_2_PI": "0.63661977236758134308",
    "M_2_SQRTPI": "1.12837916709551257390",
    "M_SQRT2": "1.41421356237309504880",
    "M_SQRT1_2": "0.70710678118654752440",
    "P_Q": "1.602176462e-19",
    "P_C": "2.99792458e8",
    "P_K": "1.3806503e-23",
    "P_H": "6.62606876e-34",
    "P_EPS0": "8.854187817e-12",
    "P_U0": "(4.0e-7 * 3.14159265358979323846)",
    "P_CELSIUS0": "273.15",
}
STANDARD_INCLUDES = {"constants.vams", "disciplines.vams", "constants.h", "discipline.h", "disciplines.h"}


class PPError(Exception):
    pass


def strip_comments(text):
    """Removes // and /* */ comments outside string literals, keeping every
    newline so that line numbers survive."""
    out = []
    i, n = 0, len(text)
    while i < n:
        c = text[i]
        if c == '"':
            j = i + 1
            while j < n and text[j] != '"':
                j += 2 if text[j] == "\\" else 1
            out.append(text[i:j + 1])
            i = j + 1
        elif text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            if j < 0:
                raise PPError("unterminated block comment")
            out.append("\n" * text.count("\n", i, j))
            i = j + 2
        else:
            out.append(c)
            i += 1
    return "".join(out)


_ident = re.compile(r"[A-Za-z_][A-Za-z0-9_$]*")


def _split_args(text, i):
    """Reads a parenthesised, comma-separated argument list starting at
    text[i] == '('. Returns (args, index after ')')."""
    assert text[i] == "("
    depth, start, args = 0, i + 1, []
    j = i
    while j < len(text):
        c = text[j]
        if c == '"':
            j = text.index('"', j + 1)
        elif c in "([{":
            depth += 1
        elif c in ")]}":
            depth -= 1
            if depth == 0:
                args.append(text[start:j])
                return args, j + 1
        elif c == "," and depth == 1:
            args.append(text[start:j])
            start = j + 1
        j += 1
    raise PPError("unterminated macro argument list")


class Preprocessor:
    def __init__(self, defines=None, include_dirs=()):
        self.macros = {}  # name -> (params or None, body)
        for k, v in STANDARD_MACROS.items():
            self.macros[k] = (None, v)
        for k, v in (defines or {}).items():
            self.macros[k] = (None, str(v))
        self.include_dirs = list(include_dirs)

    def expand(self, text, depth=0):
        """Expands every macro use in text (no directives inside)."""
        if depth > 64:
            raise PPError("macro expansion too deep")
        out = []
        i, n = 0, len(text)
        while i < n:
            c = text[i]
            if c == '"':
                j = text.index('"', i + 1)
                out.append(text[i:j + 1])
                i = j + 1
                continue
            if c != "`":
                j = i
                while j < n and text[j] not in '`"':
                    j += 1
                out.append(text[i:j])
                i = j
                continue
            m = _ident.match(text, i + 1)
            if not m:
                raise PPError(f"stray backtick near {text[i:i + 30]!r}")
            name = m.group(0)
            j = m.end()
            if name not in self.macros:
                raise PPError(f"undefined macro `{name}")
            params, body = self.macros[name]
            if params is None:
                out.append(" " + self.expand(body, depth + 1) + " ")
                i = j
                continue
            k = j
            while k < n and text[k] in " \t\n":
                k += 1
            if k >= n or text[k] != "(":
                raise PPError(f"macro `{name} needs arguments")
            args, end = _split_args(text, k)
            if len(args) != len(params):
                raise PPError(f"macro `{name} takes {len(params)} arguments, got {len(args)}")
            args = [self.expand(a, depth + 1) for a in args]
            sub = dict(zip(params, args))
            body2 = _ident.sub(lambda mm: sub.get(mm.group(0), mm.group(0)) if not _preceded_by_tick(body, mm) else mm.group(0), body)
            out.append(" " + self.expand(body2, depth + 1) + " ")
            i = end
        return "".join(out)

    def process(self, text, path="<input>"):
        text = strip_comments(text)
        # join continued lines, keeping the line count with trailing newlines
        lines = text.split("\n")
        logical = []
        buf, nbuf = "", 0
        for ln in lines:
            if ln.rstrip().endswith("\\"):
                buf += ln.rstrip()[:-1] + " "
                nbuf += 1
                continue
            logical.append(buf + ln)
            logical.extend([""] * nbuf)
            buf, nbuf = "", 0
        if buf:
            logical.append(buf)
        out = []
        stack = []  # (active_now, taken_any, parent_active)
        active = True
        for ln in logical:
            s = ln.strip()
            m = re.match(r"`(define|undef|ifdef|ifndef|elsif|else|endif|include|timescale|resetall|default_nettype)\b(.*)", s)
            if not m:
                out.append(self.expand(ln) if active and "`" in ln else (ln if active else ""))
                continue
            d, rest = m.group(1), m.group(2).strip()
            if d in ("ifdef", "ifndef"):
                name = rest.split()[0]
                cond = (name in self.macros) == (d == "ifdef")
                stack.append((active, cond))
                active = active and cond
            elif d == "elsif":
                parent, taken = stack[-1]
                name = rest.split()[0]
                cond = (not taken) and name in self.macros
                stack[-1] = (parent, taken or cond)
                active = parent and cond
            elif d == "else":
                parent, taken = stack[-1]
                stack[-1] = (parent, True)
                active = parent and not taken
            elif d == "endif":
                active = stack.pop()[0]
            elif not active:
                pass
            elif d == "define":
                mm = _ident.match(rest)
                name = mm.group(0)
                tail = rest[mm.end():]
                if tail.startswith("("):
                    args, end = _split_args(tail, 0)
                    params = [a.strip() for a in args]
                    self.macros[name] = (params, tail[end:].strip())
                else:
                    self.macros[name] = (None, tail.strip())
            elif d == "undef":
                self.macros.pop(rest.split()[0], None)
            elif d == "include":
                fname = rest.strip().strip('"')
                if fname in STANDARD_INCLUDES:
                    pass
                else:
                    found = None
                    for dd in [os.path.dirname(path)] + self.include_dirs:
                        p = os.path.join(dd, fname)
                        if os.path.exists(p):
                            found = p
                            break
                    if not found:
                        raise PPError(f"include {fname} not found")
                    with open(found) as f:
                        out.append(self.process(f.read(), found))
            out.append("")
        if stack:
            raise PPError("unterminated `ifdef")
        return "\n".join(out)


def _preceded_by_tick(body, m):
    return m.start() > 0 and body[m.start() - 1] == "`"


def preprocess(path, defines=None):
    with open(path) as f:
        text = f.read()
    return Preprocessor(defines).process(text, path)
