// The deck subset reader and its builder onto a lane template (deck.h).
#include "deck.h"
#include "sampler.h"

#include <cctype>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/espice_lanes.h"

namespace {
// The engine's PWL table (devices.h PWL_CAP; models/vsource.va holds 64).
constexpr size_t PWL_POINTS = 16;
}  // namespace

namespace esp {

// ------------------------------------------------------------- numbers

// 10^e as glibc's pow(10.0, e) gives it: correctly rounded for every e but
// 23 and 210, which it rounds up; 0 below 1e-330, inf past 1e308
// (src/frontend/lines.zig:250-263).
static double pow10_glibc(int e)
{
    static double table[308 + 330 + 1];
    static bool init = false;
    if (!init) {
        char buf[32];
        for (int k = -330; k <= 308; k++) {
            std::snprintf(buf, sizeof buf, "1e%d", k);
            table[k + 330] = std::strtod(buf, nullptr);
        }
        table[23 + 330] = 1.0000000000000001e23;
        table[210 + 330] = 1.0000000000000001e210;
        init = true;
    }
    if (e < -330) return 0.0;
    if (e > 308) return HUGE_VAL;
    return table[e + 330];
}

// Length of the leading decimal literal (numberEnd, lines.zig:150-178,
// with `d` as an exponent letter); 0 when there is none.
static size_t number_end(const std::string &t, size_t at)
{
    size_t end = at;
    bool digit = false, dot = false, exp = false;
    while (end < t.size()) {
        const char c = t[end];
        if (c >= '0' && c <= '9') {
            digit = true;
        } else if (c == '.' && !dot && !exp) {
            dot = true;
        } else if ((c == 'e' || c == 'd') && !exp && digit && end + 1 < t.size() &&
                   (std::isdigit((unsigned char)t[end + 1]) || t[end + 1] == '-' || t[end + 1] == '+')) {
            exp = true;
            end++;
            if (!std::isdigit((unsigned char)t[end])) {
                if (end + 1 >= t.size() || !std::isdigit((unsigned char)t[end + 1])) return 0;
            }
        } else if ((c == '+' || c == '-') && end == at) {
        } else {
            break;
        }
        end++;
    }
    return digit ? end - at : 0;
}

// inpEvaluate (lines.zig:222-248): one f64 mantissa of every digit, then
// sign * mantissa * 10^exponent.
static double inp_evaluate(const std::string &lit, int scale_exp, bool mil)
{
    size_t i = 0;
    double sign = 1;
    if (lit[0] == '+' || lit[0] == '-') {
        if (lit[0] == '-') sign = -1;
        i = 1;
    }
    double mantis = 0;
    int e = scale_exp;
    while (i < lit.size() && std::isdigit((unsigned char)lit[i])) mantis = 10 * mantis + (double)(lit[i++] - '0');
    if (i < lit.size() && lit[i] == '.') {
        i++;
        while (i < lit.size() && std::isdigit((unsigned char)lit[i])) {
            mantis = 10 * mantis + (double)(lit[i++] - '0');
            e -= 1;
        }
    }
    if (i < lit.size()) {
        i++;
        int esign = 1;
        if (lit[i] == '+' || lit[i] == '-') {
            if (lit[i] == '-') esign = -1;
            i++;
        }
        int x = 0;
        while (i < lit.size()) {
            x = 10 * x + (lit[i] - '0');
            if (x > 100000) x = 100000;
            i++;
        }
        e += esign * x;
    }
    if (mil) mantis *= 25.4;
    if (mantis == 0) return sign * 0.0;
    return sign * mantis * pow10_glibc(e);
}

// parseSpiceNum (lines.zig:182-213) from position `at`; `len` receives the
// characters read (the literal and its suffix letters).
static bool spice_number_at(const std::string &t, size_t at, double &out, size_t &len)
{
    const size_t n = number_end(t, at);
    if (n == 0) return false;
    size_t s = at + n;
    int scale = 0;
    bool mil = false;
    if (t.compare(s, 3, "meg") == 0) scale = 6;
    else if (t.compare(s, 3, "mil") == 0) scale = -6, mil = true;
    else if (s < t.size()) {
        switch (t[s]) {
        case 't': scale = 12; break;
        case 'g': scale = 9; break;
        case 'k': scale = 3; break;
        case 'm': scale = -3; break;
        case 'u': scale = -6; break;
        case 'n': scale = -9; break;
        case 'p': scale = -12; break;
        case 'f': scale = -15; break;
        case 'a': scale = -18; break;
        default: break;
        }
    }
    out = inp_evaluate(t.substr(at, n), scale, mil);
    size_t e = s;
    while (e < t.size() && (std::isalpha((unsigned char)t[e]) || t[e] == '_')) e++;
    len = e - at;
    return true;
}

bool spice_number(const std::string &text, double &out)
{
    size_t len = 0;
    return spice_number_at(text, 0, out, len) && len == text.size();
}

// ---------------------------------------------------------- expressions

struct Expr {
    enum Op { NUM, NAME, NEG, ADD, SUB, MUL, DIV, POW, CALL } op = NUM;
    double num = 0;
    std::string name;
    std::vector<ExprP> args;
};

namespace {

struct ExprParser {
    const std::string &s;
    size_t i = 0;
    std::string err;

    void ws()
    {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    ExprP make(Expr::Op op, std::vector<ExprP> args)
    {
        auto e = std::make_shared<Expr>();
        e->op = op;
        e->args = std::move(args);
        return e;
    }
    ExprP expr()
    {
        ExprP l = term();
        for (;;) {
            ws();
            if (!l || i >= s.size() || (s[i] != '+' && s[i] != '-')) return l;
            const char c = s[i++];
            ExprP r = term();
            if (!r) return nullptr;
            l = make(c == '+' ? Expr::ADD : Expr::SUB, {l, r});
        }
    }
    ExprP term()
    {
        ExprP l = unary();
        for (;;) {
            ws();
            if (!l || i >= s.size()) return l;
            if (s[i] == '*' && !(i + 1 < s.size() && s[i + 1] == '*')) {
                i++;
                ExprP r = unary();
                if (!r) return nullptr;
                l = make(Expr::MUL, {l, r});
            } else if (s[i] == '/') {
                i++;
                ExprP r = unary();
                if (!r) return nullptr;
                l = make(Expr::DIV, {l, r});
            } else {
                return l;
            }
        }
    }
    ExprP unary()
    {
        ws();
        if (i < s.size() && (s[i] == '-' || s[i] == '+')) {
            const char c = s[i++];
            ExprP a = unary();
            if (!a) return nullptr;
            return c == '-' ? make(Expr::NEG, {a}) : a;
        }
        return power();
    }
    ExprP power()
    {
        ExprP b = primary();
        ws();
        if (!b) return nullptr;
        if (i < s.size() && (s[i] == '^' || (s[i] == '*' && i + 1 < s.size() && s[i + 1] == '*'))) {
            i += s[i] == '^' ? 1 : 2;
            ExprP x = unary();
            if (!x) return nullptr;
            return make(Expr::POW, {b, x});
        }
        return b;
    }
    ExprP primary()
    {
        ws();
        if (i >= s.size()) {
            err = "ExpressionEnd";
            return nullptr;
        }
        if (s[i] == '(') {
            i++;
            ExprP e = expr();
            ws();
            if (!e || i >= s.size() || s[i] != ')') {
                if (err.empty()) err = "UnbalancedParenthesis";
                return nullptr;
            }
            i++;
            return e;
        }
        double v;
        size_t len;
        if ((std::isdigit((unsigned char)s[i]) || s[i] == '.') && spice_number_at(s, i, v, len)) {
            i += len;
            auto e = std::make_shared<Expr>();
            e->op = Expr::NUM;
            e->num = v;
            return e;
        }
        if (std::isalpha((unsigned char)s[i]) || s[i] == '_') {
            size_t b = i;
            while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == '.')) i++;
            auto e = std::make_shared<Expr>();
            e->name = s.substr(b, i - b);
            ws();
            if (i < s.size() && s[i] == '(') {
                i++;
                e->op = Expr::CALL;
                for (;;) {
                    ExprP a = expr();
                    if (!a) return nullptr;
                    e->args.push_back(a);
                    ws();
                    if (i < s.size() && s[i] == ',') {
                        i++;
                        continue;
                    }
                    if (i < s.size() && s[i] == ')') {
                        i++;
                        break;
                    }
                    err = "UnbalancedParenthesis";
                    return nullptr;
                }
            } else {
                e->op = Expr::NAME;
            }
            return e;
        }
        err = "BadExpression";
        return nullptr;
    }
};

}  // namespace

ExprP parse_expr(const std::string &text, std::string &err)
{
    ExprParser p{text};
    ExprP e = p.expr();
    p.ws();
    if (e && p.i != text.size()) {
        err = "BadExpression";
        return nullptr;
    }
    if (!e) err = p.err.empty() ? "BadExpression" : p.err;
    return e;
}

double eval_expr(const Expr &e, const std::map<std::string, double> &env, std::string &err, Draw *draw)
{
    auto arg = [&](size_t k) { return eval_expr(*e.args[k], env, err, draw); };
    switch (e.op) {
    case Expr::NUM: return e.num;
    case Expr::NAME: {
        auto it = env.find(e.name);
        if (it == env.end()) {
            err = "UnknownParameter " + e.name;
            return NAN;
        }
        return it->second;
    }
    case Expr::NEG: return -arg(0);
    case Expr::ADD: return arg(0) + arg(1);
    case Expr::SUB: return arg(0) - arg(1);
    case Expr::MUL: return arg(0) * arg(1);
    case Expr::DIV: return arg(0) / arg(1);
    case Expr::POW: return std::pow(arg(0), arg(1));
    case Expr::CALL: {
        const std::string &f = e.name;
        const size_t n = e.args.size();
        const int dist = f == "agauss" ? MC_AGAUSS : f == "gauss" ? MC_GAUSS : f == "aunif" ? MC_AUNIF
                       : f == "unif"   ? MC_UNIF   : f == "limit" ? MC_LIMIT : -1;
        if (dist >= 0) {
            if (n < 1 || n > 4) {
                err = "BadDistribution " + f;
                return NAN;
            }
            double a[4];
            for (size_t k = 0; k < n; k++) a[k] = arg(k);
            if (!draw || draw->trial == 0) return a[0];
            const uint64_t site = site_key(draw->owner, draw->ordinal++);
            return trial_draw(draw->seed, draw->trial, site, dist, a, (uint32_t)n);
        }
        if (n == 1) {
            const double x = arg(0);
            if (f == "sqrt") return std::sqrt(x);
            if (f == "exp") return std::exp(x);
            if (f == "log" || f == "ln") return std::log(x);
            if (f == "log10") return std::log10(x);
            if (f == "abs") return std::fabs(x);
            if (f == "sin") return std::sin(x);
            if (f == "cos") return std::cos(x);
            if (f == "tan") return std::tan(x);
            if (f == "atan") return std::atan(x);
            if (f == "sinh") return std::sinh(x);
            if (f == "cosh") return std::cosh(x);
            if (f == "tanh") return std::tanh(x);
            if (f == "floor") return std::floor(x);
            if (f == "ceil") return std::ceil(x);
        } else if (n == 2) {
            const double x = arg(0), y = arg(1);
            if (f == "min") return x < y ? x : y;
            if (f == "max") return x > y ? x : y;
            if (f == "pow" || f == "pwr") return std::pow(x, y);
        }
        err = "UnknownFunction " + f;
        return NAN;
    }
    }
    return NAN;
}

void expr_names(const Expr &e, std::vector<std::string> &out)
{
    if (e.op == Expr::NAME) out.push_back(e.name);
    for (auto &a : e.args) expr_names(*a, out);
}

// ------------------------------------------------------------ the cards

namespace {

// A card's tokens: words, and `(`, `)`, `=` on their own; {...} and '...'
// stay whole (their inner text, marked by the leading brace or quote).
std::vector<std::string> tokens(const std::string &line)
{
    std::vector<std::string> out;
    size_t i = 0;
    while (i < line.size()) {
        const char c = line[i];
        if (std::isspace((unsigned char)c) || c == ',') {
            i++;
        } else if (c == '(' || c == ')' || c == '=') {
            out.emplace_back(1, c);
            i++;
        } else if (c == '{' || c == '\'' || c == '"') {
            const char close = c == '{' ? '}' : c;
            size_t j = line.find(close, i + 1);
            if (j == std::string::npos) j = line.size();
            out.push_back(std::string(1, '{') + line.substr(i + 1, j - i - 1));
            i = j + 1;
        } else {
            size_t j = i;
            while (j < line.size() && !std::isspace((unsigned char)line[j]) && line[j] != '(' && line[j] != ')' &&
                   line[j] != '=' && line[j] != ',' && line[j] != '{' && line[j] != '\'')
                j++;
            out.push_back(line.substr(i, j - i));
            i = j;
        }
    }
    return out;
}

bool to_value(const std::string &tok, Value &v, std::string &err)
{
    v.given = true;
    if (!tok.empty() && tok[0] == '{') {
        v.expr = parse_expr(tok.substr(1), err);
        return v.expr != nullptr;
    }
    if (spice_number(tok, v.num)) return true;
    v.expr = parse_expr(tok, err);
    return v.expr != nullptr;
}

struct Cursor {
    const std::vector<std::string> &t;
    size_t i;
    bool more() const { return i < t.size(); }
    const std::string &peek() const { return t[i]; }
    bool next_is(const char *s) const { return i < t.size() && t[i] == s; }
};

// k = v pairs to the end of the card (parentheses skipped).
bool kv_pairs(Cursor &c, std::map<std::string, Value> &kv, std::string &err)
{
    while (c.more()) {
        if (c.peek() == "(" || c.peek() == ")") {
            c.i++;
            continue;
        }
        const std::string key = c.peek();
        if (c.i + 2 >= c.t.size() + 1 || !(c.i + 1 < c.t.size() && c.t[c.i + 1] == "=") || c.i + 2 >= c.t.size()) {
            err = "ExpectedKeyValue " + key;
            return false;
        }
        Value v;
        if (!to_value(c.t[c.i + 2], v, err)) return false;
        kv[key] = v;
        c.i += 3;
    }
    return true;
}

// v(node) or i(name) at the cursor.
bool vector_at(Cursor &c, std::string &vec, std::string &err)
{
    if (c.i + 3 < c.t.size() + 0 && (c.peek() == "v" || c.peek() == "i") && c.t[c.i + 1] == "(" && c.t[c.i + 3] == ")") {
        vec = c.peek() + "(" + c.t[c.i + 2] + ")";
        c.i += 4;
        return true;
    }
    err = "UnsupportedVector";
    return false;
}

int count_of(const std::string &tok)
{
    if (tok == "last") return -2;
    double v;
    if (!spice_number(tok, v) || v < 1 || v != std::floor(v)) return -3;
    return (int)v;
}

// TRIG/TARG/WHEN side conditions: VAL=, TD=, RISE=, FALL=, CROSS=, FROM=, TO=.
bool side_conditions(Cursor &c, Meas::Side &s, std::string &err, const char *stop)
{
    while (c.more() && !(stop && c.peek() == stop)) {
        const std::string key = c.peek();
        if (key == "when") {
            err = "BadMeasure when";
            return false;
        }
        if (!(c.i + 2 < c.t.size() && c.t[c.i + 1] == "=")) {
            err = "BadMeasure " + key;
            return false;
        }
        const std::string &val = c.t[c.i + 2];
        c.i += 3;
        if (key == "val") {
            if (!to_value(val, s.val, err)) return false;
        } else if (key == "td") {
            if (!to_value(val, s.td, err)) return false;
        } else if (key == "from") {
            if (!to_value(val, s.from, err)) return false;
        } else if (key == "to") {
            if (!to_value(val, s.to, err)) return false;
        } else if (key == "at") {
            if (!to_value(val, s.at, err)) return false;
        } else if (key == "rise" || key == "fall" || key == "cross") {
            const int k = count_of(val);
            if (k == -3) {
                err = "BadMeasureCount";
                return false;
            }
            (key == "rise" ? s.rise : key == "fall" ? s.fall : s.cross) = k;
        } else {
            err = "UnsupportedMeasureOption " + key;
            return false;
        }
    }
    return true;
}

bool read_meas(const std::vector<std::string> &t, Meas &m, std::string &err)
{
    if (t.size() < 4) {
        err = "BadMeasure";
        return false;
    }
    m.analysis = t[1];
    m.name = t[2];
    if (m.analysis != "tran" && m.analysis != "op") {
        err = "UnsupportedMeasureAnalysis " + m.analysis;
        return false;
    }
    Cursor c{t, 3};
    const std::string f = c.peek();
    c.i++;
    if (f == "trig") {
        m.kind = "trig_targ";
        if (!vector_at(c, m.a.vec, err) || !side_conditions(c, m.a, err, "targ")) return false;
        if (!c.next_is("targ")) {
            err = "MeasureWithoutTarg";
            return false;
        }
        c.i++;
        return vector_at(c, m.b.vec, err) && side_conditions(c, m.b, err, nullptr);
    }
    if (f == "when") {
        m.kind = "when";
        if (!vector_at(c, m.a.vec, err)) return false;
        if (!(c.next_is("=") && c.i + 1 < t.size())) {
            err = "BadWhen";
            return false;
        }
        if (!to_value(t[c.i + 1], m.a.val, err)) return false;
        c.i += 2;
        return side_conditions(c, m.a, err, nullptr);
    }
    if (f == "find" || f == "deriv" || f == "derivative") {
        m.kind = f == "find" ? "find" : "deriv";
        if (!vector_at(c, m.a.vec, err) || !side_conditions(c, m.a, err, "when")) return false;
        if (c.next_is("when")) {
            // FIND / DERIV sig WHEN sig2=val [RISE= FALL= CROSS= ...]: the
            // WHEN's Cross, then the value of sig at its event
            c.i++;
            m.kind += "_when";
            if (!vector_at(c, m.b.vec, err)) return false;
            if (!(c.next_is("=") && c.i + 1 < t.size())) {
                err = "BadWhen";
                return false;
            }
            if (!to_value(t[c.i + 1], m.b.val, err)) return false;
            c.i += 2;
            if (!side_conditions(c, m.b, err, nullptr)) return false;
            if (m.analysis != "tran" || m.b.td.given || m.b.from.given || m.b.rise == -2 || m.b.fall == -2 ||
                m.b.cross == -2) {
                // measured as it goes only without TD=, FROM= or a LAST count
                err = "UnsupportedMeasure " + f + " WHEN";
                return false;
            }
            return true;
        }
        if (m.analysis == "tran" && !m.a.at.given) {
            err = "UnsupportedMeasure " + f;
            return false;
        }
        return true;
    }
    if (f == "min" || f == "max" || f == "avg" || f == "rms" || f == "integ" || f == "integral") {
        m.kind = f == "integral" ? "integ" : f;
        return vector_at(c, m.a.vec, err) && side_conditions(c, m.a, err, nullptr);
    }
    if (f == "param") {
        m.kind = "param";
        if (!(c.next_is("=") && c.i + 1 < t.size())) {
            err = "BadParamMeasure";
            return false;
        }
        const std::string &x = t[c.i + 1];
        m.param = parse_expr(x[0] == '{' ? x.substr(1) : x, err);
        return m.param != nullptr;
    }
    err = "UnsupportedMeasure " + f;
    return false;
}

bool read_element(const std::vector<std::string> &t, Element &e, std::string &err)
{
    e.name = t[0];
    e.letter = t[0][0];
    size_t ports = 0;
    switch (e.letter) {
    case 'r':
    case 'c':
    case 'v':
    case 'i':
    case 'd': ports = 2; break;
    case 'm': ports = 4; break;
    default: err = "UnsupportedElement " + e.name; return false;
    }
    if (t.size() < 1 + ports) {
        err = "MissingNodes " + e.name;
        return false;
    }
    for (size_t k = 0; k < ports; k++) e.nodes.push_back(t[1 + k]);
    Cursor c{t, 1 + ports};
    if (e.letter == 'r' || e.letter == 'c') {
        if (!c.more() || !to_value(c.peek(), e.value, err)) {
            if (err.empty()) err = "MissingValue " + e.name;
            return false;
        }
        c.i++;
        if (c.more()) {
            err = "UnsupportedOption " + e.name;
            return false;
        }
        return true;
    }
    if (e.letter == 'd' || e.letter == 'm') {
        if (!c.more()) {
            err = "MissingModel " + e.name;
            return false;
        }
        e.model = c.peek();
        c.i++;
        if (e.letter == 'd' && c.more() && !(c.i + 1 < t.size() && t[c.i + 1] == "=")) {
            Value area;
            if (!to_value(c.peek(), area, err)) return false;
            e.kv["area"] = area;
            c.i++;
        }
        if (!kv_pairs(c, e.kv, err)) return false;
        return true;
    }
    // V and I: [dc] value, ac ..., pulse(...)
    bool have_value = false;
    while (c.more()) {
        const std::string w = c.peek();
        if (w == "dc") {
            if (c.i + 1 >= t.size() || !to_value(t[c.i + 1], e.value, err)) {
                if (err.empty()) err = "MissingValue " + e.name;
                return false;
            }
            have_value = true;
            c.i += 2;
        } else if (w == "ac") {
            c.i++;
            double x;
            e.ac = true;
            for (int k = 0; k < 2 && c.more() && spice_number(c.peek(), x); k++) {
                (k == 0 ? e.ac_mag : e.ac_phase) = x;
                c.i++;
            }
        } else if ((w == "pulse" || w == "sin" || w == "exp" || w == "am" || (w == "sffm" && e.letter == 'v')) &&
                   (e.letter == 'v' || e.letter == 'i') && e.wave == 0) {
            // positional arguments fill the waveform's slots in order; the
            // ones not written keep the model defaults
            // (frontend/builder.zig:2841-2871, 2985-2990)
            // PULSE's eighth argument is ngspice's pulse count NP and AM
            // reads ngspice's eight (VO VMO VMA FM FC TD PHASEM PHASEC); the
            // build maps both (frontend/builder.zig:1478-1562)
            const int most = w == "pulse" ? 8 : w == "sffm" ? 8 : w == "am" ? 8 : 6;
            const std::string bad = (w == "pulse" ? "BadPulse " : w == "sin" ? "BadSin " : w == "exp" ? "BadExp "
                                     : w == "sffm" ? "BadSffm " : "BadAm ") + e.name;
            c.i++;
            if (!c.next_is("(")) {
                err = bad;
                return false;
            }
            c.i++;
            int k = 0;
            while (c.more() && !c.next_is(")")) {
                if (k >= most || !to_value(c.peek(), e.wave_args[k], err)) {
                    if (err.empty()) err = bad;
                    return false;
                }
                k++;
                c.i++;
            }
            // trailing arguments may be left off (their slots keep the
            // model defaults); AM needs two to be read in ngspice's order
            if (!c.next_is(")") || k < (w == "am" ? 2 : 1)) {
                err = bad;
                return false;
            }
            c.i++;
            e.wave = w == "pulse" ? 1 : w == "sin" ? 2 : w == "exp" ? 3 : w == "sffm" ? 5 : 6;
            e.n_wave_args = k;
        } else if (w == "pwl" && (e.letter == 'v' || e.letter == 'i') && e.wave == 0) {
            // (time, value) pairs (frontend/builder.zig:2848-2859), then
            // `r=` and `td=` on the card
            c.i++;
            if (!c.next_is("(")) {
                err = "BadPwl " + e.name;
                return false;
            }
            c.i++;
            while (c.more() && !c.next_is(")")) {
                Value v;
                if (!to_value(c.peek(), v, err)) {
                    if (err.empty()) err = "BadPwl " + e.name;
                    return false;
                }
                e.pwl.push_back(v);
                c.i++;
            }
            if (!c.next_is(")") || e.pwl.empty() || e.pwl.size() % 2) {
                err = "BadPwl " + e.name;
                return false;
            }
            if (e.pwl.size() / 2 > PWL_POINTS) {
                err = "PwlTooLong " + e.name;
                return false;
            }
            c.i++;
            e.wave = 4;
            while (c.i + 2 < t.size() && (c.peek() == "r" || c.peek() == "td") && t[c.i + 1] == "=") {
                Value &dst = c.peek() == "r" ? e.pwl_r : e.pwl_td;
                if (!to_value(t[c.i + 2], dst, err)) return false;
                c.i += 3;
            }
            // a numeric r= must be one of the table's times before its last
            // (frontend/builder.zig:1580-1594, ngspice vsrcpar.c:160-176)
            if (e.pwl_r.given && !e.pwl_r.expr) {
                bool on_point = false;
                for (size_t k = 0; k + 1 < e.pwl.size(); k += 2)
                    on_point = on_point || (!e.pwl[k].expr && e.pwl[k].num == e.pwl_r.num);
                const Value &last = e.pwl[e.pwl.size() - 2];
                const double t_last = last.expr ? e.pwl_r.num : last.num;
                if (!on_point || e.pwl_r.num >= t_last) {
                    err = "InvalidParameterValue " + e.name + " PWL r= is not a time point before the last";
                    return false;
                }
            }
        } else if (w == "sin" || w == "pwl" || w == "exp" || w == "sffm" || w == "am" || w == "pulse") {
            err = "UnsupportedWaveform " + w;
            return false;
        } else if (!have_value) {
            if (!to_value(w, e.value, err)) return false;
            have_value = true;
            c.i++;
        } else {
            err = "UnsupportedOption " + e.name;
            return false;
        }
    }
    return true;
}

}  // namespace

// `.param name = value ...` from the card's text: a value is a quoted or
// braced expression, or bare text running to the next top-level `name =`
// (so a function call keeps its parentheses and commas).
static bool read_params(const std::string &line, Deck &d, std::string &err)
{
    auto word = [](char c) { return std::isalnum((unsigned char)c) || c == '_' || c == '.'; };
    auto space = [](char c) { return std::isspace((unsigned char)c) || c == ','; };
    size_t i = line.find_first_of(" \t");
    if (i == std::string::npos) return true;
    for (;;) {
        while (i < line.size() && space(line[i])) i++;
        if (i >= line.size()) return true;
        const size_t b = i;
        while (i < line.size() && word(line[i])) i++;
        const std::string name = line.substr(b, i - b);
        while (i < line.size() && std::isspace((unsigned char)line[i])) i++;
        if (name.empty() || i >= line.size() || line[i] != '=') {
            err = "BadParam";
            return false;
        }
        i++;
        while (i < line.size() && std::isspace((unsigned char)line[i])) i++;
        std::string text;
        bool braced = false;
        if (i < line.size() && (line[i] == '{' || line[i] == '\'' || line[i] == '"')) {
            const char close = line[i] == '{' ? '}' : line[i];
            const size_t j = line.find(close, i + 1);
            if (j == std::string::npos) {
                err = "BadParam";
                return false;
            }
            text = line.substr(i + 1, j - i - 1);
            braced = true;
            i = j + 1;
        } else {
            const size_t v0 = i;
            int depth = 0;
            while (i < line.size()) {
                const char c = line[i];
                if (c == '(') depth++;
                if (c == ')') depth--;
                if (depth == 0 && space(c)) {
                    size_t k = i;
                    while (k < line.size() && space(line[k])) k++;
                    size_t k2 = k;
                    while (k2 < line.size() && word(line[k2])) k2++;
                    size_t k3 = k2;
                    while (k3 < line.size() && std::isspace((unsigned char)line[k3])) k3++;
                    if (k >= line.size() || (k2 > k && k3 < line.size() && line[k3] == '=')) break;
                }
                i++;
            }
            text = line.substr(v0, i - v0);
        }
        Value v;
        v.given = true;
        double x;
        size_t len;
        if (!braced && spice_number_at(text, 0, x, len) && len == text.size()) {
            v.num = x;
        } else {
            v.expr = parse_expr(text, err);
            if (!v.expr) return false;
        }
        d.params.emplace_back(name, v);
    }
}

bool read_deck(const std::string &text, Deck &d, std::string &err)
{
    // fold case, split lines, join '+' continuations, drop comments
    std::string low(text);
    for (char &ch : low) ch = (char)std::tolower((unsigned char)ch);
    std::vector<std::pair<int, std::string>> lines;
    size_t at = 0;
    int no = 0;
    bool first = true;
    while (at <= low.size()) {
        size_t e = low.find('\n', at);
        if (e == std::string::npos) e = low.size();
        std::string ln = low.substr(at, e - at);
        at = e + 1;
        no++;
        if (!ln.empty() && ln.back() == '\r') ln.pop_back();
        if (first) {
            d.title = text.substr(0, text.find('\n') == std::string::npos ? text.size() : text.find('\n'));
            if (!d.title.empty() && d.title.back() == '\r') d.title.pop_back();
            first = false;
            continue;
        }
        const size_t semi = ln.find(';');
        if (semi != std::string::npos) ln.resize(semi);
        for (size_t k = 1; k < ln.size(); k++)
            if (ln[k] == '$' && std::isspace((unsigned char)ln[k - 1])) {
                ln.resize(k);
                break;
            }
        size_t b = ln.find_first_not_of(" \t");
        if (b == std::string::npos || ln[b] == '*') continue;
        ln = ln.substr(b);
        if (ln[0] == '+') {
            if (lines.empty()) {
                err = "ContinuationFirst";
                return false;
            }
            lines.back().second += " " + ln.substr(1);
            continue;
        }
        lines.emplace_back(no, ln);
        if (e == low.size()) break;
    }
    for (auto &card : lines) {
        const int line_no = card.first;
        const std::vector<std::string> t = tokens(card.second);
        if (t.empty()) continue;
        const std::string &w = t[0];
        auto fail_at = [&](const std::string &e) {
            err = e + " (line " + std::to_string(line_no) + ")";
            return false;
        };
        if (w[0] != '.') {
            Element e;
            e.line = line_no;
            if (!read_element(t, e, err)) return fail_at(err);
            d.elements.push_back(std::move(e));
        } else if (w == ".end") {
            break;
        } else if (w == ".param") {
            if (!read_params(card.second, d, err)) return fail_at(err);
        } else if (w == ".model") {
            if (t.size() < 3) return fail_at("BadModel");
            Model m;
            m.name = t[1];
            m.type = t[2];
            m.line = line_no;
            Cursor c{t, 3};
            if (!kv_pairs(c, m.kv, err)) return fail_at(err);
            if (m.type != "nmos" && m.type != "pmos" && m.type != "d") return fail_at("UnsupportedModel " + m.type);
            if (m.type != "d") {
                auto lv = m.kv.find("level");
                if (lv != m.kv.end() && !(lv->second.expr == nullptr && lv->second.num == 1.0))
                    return fail_at("UnsupportedModelLevel");
                m.kv.erase("level");
            }
            d.models[m.name] = m;
        } else if (w == ".op") {
            d.op = true;
        } else if (w == ".option" || w == ".options") {
            // the sampler's options (variants.zig:1694-1705); any other
            // option would change the solve and is refused
            Cursor c{t, 1};
            while (c.more()) {
                if (!(c.i + 2 < t.size() && t[c.i + 1] == "=")) return fail_at("UnsupportedCard " + w);
                const std::string &k = t[c.i], &v = t[c.i + 2];
                double x;
                if (k == "seed" && spice_number(v, x)) {
                    d.seed = x != x ? 0 : std::fabs(x) >= 18446744073709551616.0 ? UINT64_MAX : (uint64_t)std::fabs(x);
                    d.has_seed = true;
                } else if (!(k == "sampling_method" && v == "srs")) {
                    return fail_at("UnsupportedCard " + w + " " + k);
                }
                c.i += 3;
            }
        } else if (w == ".dc" && t.size() == 4 && t[1] == "monte" && t[2] == "=") {
            double x;
            if (!spice_number(t[3], x) || !(x >= 1) || x > 1e6) return fail_at("BadMonte");
            d.monte_dc = (uint32_t)x;
        } else if (w == ".tran") {
            Cursor c{t, 1};
            Value *dst[4] = {&d.tstep, &d.tstop, &d.tstart, &d.tmax};
            int k = 0;
            while (c.more()) {
                if (c.peek() == "uic") return fail_at("UicUnsupported");
                if (c.peek() == "sweep") {
                    double x;
                    if (!(c.i + 4 == t.size() && t[c.i + 1] == "monte" && t[c.i + 2] == "=" &&
                          spice_number(t[c.i + 3], x) && x >= 1 && x <= 1e6))
                        return fail_at("UnsupportedSweep");
                    d.monte_tran = (uint32_t)x;
                    break;
                }
                if (k >= 4 || !to_value(c.peek(), *dst[k], err)) return fail_at(err.empty() ? "BadTran" : err);
                k++;
                c.i++;
            }
            if (k < 2) return fail_at("BadTran");
            if (d.tstart.given && (d.tstart.expr || d.tstart.num != 0)) return fail_at("TranStartUnsupported");
            d.tran = true;
        } else if (w == ".meas" || w == ".measure") {
            Meas m;
            m.line = line_no;
            if (!read_meas(t, m, err)) return fail_at(err);
            d.meas.push_back(std::move(m));
        } else if (w == ".ac") {
            // .ac dec|oct|lin N fstart fstop
            double np = 0, f0 = 0, f1 = 0;
            if (t.size() != 5 || (t[1] != "dec" && t[1] != "oct" && t[1] != "lin") || !spice_number(t[2], np) ||
                !spice_number(t[3], f0) || !spice_number(t[4], f1) || !(np >= 1) || np > 1e7)
                return fail_at("BadAc");
            d.ac_kind = t[1] == "dec" ? 0 : t[1] == "oct" ? 1 : 2;
            d.ac_points = (uint32_t)std::floor(np + 0.5);
            d.ac_fstart = f0;
            d.ac_fstop = f1;
        } else if (w == ".dc" && t.size() == 5 && t[1] != "monte" && (t[1][0] == 'v' || t[1][0] == 'i')) {
            // .dc NAME start stop step, one source
            double a = 0, b = 0, st = 0;
            if (!spice_number(t[2], a) || !spice_number(t[3], b) || !spice_number(t[4], st))
                return fail_at("BadDc");
            d.dc_src = t[1];
            d.dc_start = a;
            d.dc_stop = b;
            d.dc_step = st;
        } else if (w == ".dc" && t.size() > 5 && t[1] != "monte") {
            return fail_at("UnsupportedCard .dc (a nested sweep)");
        } else if (w == ".dc" || w == ".noise" || w == ".print" || w == ".plot" || w == ".save" ||
                   w == ".probe" || w == ".width" || w == ".title") {
            d.ignored.push_back(w);
        } else {
            return fail_at("UnsupportedCard " + w);
        }
    }
    return true;
}

bool param_env(const Deck &d, const std::map<std::string, double> &live, std::map<std::string, double> &env,
               std::string &err, uint32_t trial)
{
    env.clear();
    for (auto &[name, v] : d.params) {
        auto it = live.find(name);
        if (it != live.end()) {
            env[name] = it->second;
            continue;
        }
        Draw dr;
        dr.seed = d.seed;
        dr.trial = trial;
        dr.owner = top_owner(name.data(), name.size());
        env[name] = v.expr ? eval_expr(*v.expr, env, err, &dr) : v.num;
        if (!err.empty()) return false;
    }
    for (auto &[name, v] : live)
        if (!env.count(name)) {
            err = "NotAParameter " + name;
            return false;
        }
    return true;
}

// ------------------------------------------------------------ the build

namespace {

// Engine parameter indices of the deck's keys (include/espice_lanes.h).
const char *const MOS1_KEYS[] = {"type", "vto", "kp", "gamma", "phi", "lambda", "tox", "ld", "u0", "nsub", "nss",
                                 "tpg", "is", "js", "cbd", "cbs", "cj", "cjsw", "pb", "mj", "mjsw", "fc", "cgso",
                                 "cgdo", "cgbo", "tnom", "gmin", "temp", "dtemp", "l", "w", "ad", "as", "pd", "ps", "m"};
const char *const DIODE_KEYS[] = {"is", "n", "cjo", "vj", "m", "fc", "tt", "gmin", "area", "tnom", "temp", "dtemp",
                                  "eg", "xti", "rs", "trs", "trs2", "tm1", "tm2"};

int key_index(const char *const *keys, size_t n, const std::string &k)
{
    for (size_t i = 0; i < n; i++)
        if (k == keys[i]) return (int)i;
    return -1;
}

Value lit(double x)
{
    Value v;
    v.given = true;
    v.num = x;
    return v;
}

}  // namespace

// PULSE with ngspice's pulse count NP (the eighth argument,
// frontend/builder.zig:1505-1545 pulseTrain): false, and the card keeps its
// first seven arguments, when the count never ends the train before the
// .tran stop; else `prm` gets a PWL of NP pulses' corners and then V1 (true),
// or err is set (a table past the engine's PWL capacity, arguments that are
// not numbers).
bool pulse_train(const Deck &d, const Element &e, std::vector<std::pair<uint32_t, Value>> &prm, std::string &err)
{
    double a[8];
    for (int q = 0; q < 8; q++) {
        if (e.wave_args[q].expr || !d.tran || d.tstop.expr || d.tstep.expr) {
            err = "UnresolvedParameter " + e.name + " a PULSE count needs numbers and a numeric .tran";
            return false;
        }
        a[q] = e.wave_args[q].num;
    }
    const double tstep = d.tstep.num, tstop = d.tstop.num;
    const double v1 = a[0], v2 = a[1], td = a[2];
    const double tr = a[3] != 0 ? a[3] : tstep, tf = a[4] != 0 ? a[4] : tstep;
    const double pw = a[5] != 0 ? a[5] : tstop;
    const double per = std::max(a[6] != 0 ? a[6] : tstop, tr + pw + tf);
    const double end = td + a[7] * per;
    if (!(a[7] > 0) || end >= tstop) return false;
    const double pulses = std::ceil(a[7]);
    std::vector<double> pts{td, v1};
    double t_prev = td, v_prev = v1;
    bool cut = false;
    for (double k = 0; k < pulses && !cut && pts.size() <= 2 * 64; k += 1) {
        const double base = td + k * per;
        const double corners[4][2] = {{base + tr, v2}, {base + tr + pw, v2}, {base + tr + pw + tf, v1}, {base + per, v1}};
        for (auto &c : corners) {
            if (c[0] == t_prev) continue;
            if (c[0] > end) {
                const double v_end = v_prev + (c[1] - v_prev) * (end - t_prev) / (c[0] - t_prev);
                pts.insert(pts.end(), {end, v_end, end, v1});
                cut = true;
                break;
            }
            pts.insert(pts.end(), {c[0], c[1]});
            t_prev = c[0];
            v_prev = c[1];
        }
    }
    const size_t np = pts.size() / 2;
    if (np > PWL_POINTS) {
        err = "PwlTooLong " + e.name + " a PULSE count past the engine's PWL table";
        return false;
    }
    prm.emplace_back(1, lit(4.0));
    prm.emplace_back(15, lit((double)np));
    for (size_t k = 0; k < np; k++) {
        prm.emplace_back((uint32_t)(18 + k), lit(pts[2 * k]));
        prm.emplace_back((uint32_t)(34 + k), lit(pts[2 * k + 1]));
    }
    if (e.value.given) prm.emplace_back(0, e.value);
    return true;
}

// The `.dc` card's swept values (src/analysis/dc/dc.zig:657-665 sweepCount,
// 383-390): floor((stop - start) / step + 1e-6) + 1 points, one for a zero
// step or one pointing away from stop, each value the previous plus step
// (accumulated, as ngspice does).
std::vector<double> dc_values(const Deck &d)
{
    std::vector<double> v;
    if (d.dc_src.empty()) return v;
    const double a = d.dc_start, b = d.dc_stop, st = d.dc_step;
    const double sgn = st > 0 ? 1.0 : st < 0 ? -1.0 : 0.0;
    size_t count = 1;
    if (!(st == 0 || (b - a) * sgn < 0)) count = (size_t)std::floor((b - a) / st + 1e-6) + 1;
    double x = a;
    for (size_t k = 0; k < count; k++) {
        if (k) x += st;
        v.push_back(x);
    }
    return v;
}

std::vector<double> ac_axis(const Deck &d)
{
    // FreqSweep.count and .at (src/core/numerics.zig:286-327) with the .ac
    // card's land_on_stop: a dec grid spanning a decade or more stretches
    // its ratio so the last point is fstop; the others step by the base's
    // 1/points power from fstart, keeping every point up to fstop plus
    // ngspice's stop tolerance; a lin grid spaces points evenly.
    std::vector<double> f;
    const double f0 = d.ac_fstart, f1 = d.ac_fstop;
    const uint32_t pts = d.ac_points;
    if (d.ac_kind < 0) return f;
    if (d.ac_kind == 2) {
        const uint32_t n = f1 != f0 ? std::max<uint32_t>(pts, 1) : 1;
        for (uint32_t k = 0; k < n; k++)
            f.push_back(n <= 1 ? f0 : f0 + ((double)k / (double)(n - 1)) * (f1 - f0));
        return f;
    }
    const double base = d.ac_kind == 1 ? 2.0 : 10.0;
    if (!(f0 > 0) || !(f1 >= f0) || pts == 0) return {f0};
    double stretched = -1.0;
    if (d.ac_kind == 0 && f1 / 10.0 >= f0) {
        const double st = std::floor(std::log10(f1 / f0) * (double)pts);
        if (st >= 1 && std::isfinite(st)) stretched = st;
    }
    const double ln_step = std::log(base) / (double)pts;
    const double ratio = f1 / f0;
    const double ln_span = std::isfinite(ratio) ? std::log(ratio) : std::log(f1) - std::log(f0);
    const double steps = stretched >= 0 ? stretched : (ln_span + std::log1p(std::exp(ln_step) * 1e-3)) / ln_step;
    if (!std::isfinite(steps) || steps < 0) return {f0};
    const uint32_t n = (uint32_t)std::min(std::floor(steps), 4294967294.0) + 1;
    for (uint32_t k = 0; k < n; k++) {
        const double kf = (double)k;
        f.push_back(stretched >= 0 ? f0 * std::pow(f1 / f0, kf / stretched) : f0 * std::pow(base, kf / (double)pts));
    }
    return f;
}

bool build_deck(BuiltDeck &bd, ::espice_lanes_template *t, std::string &err)
{
    Deck &d = bd.deck;
    // NetBuilder.build's bucket order over the subset's letters
    const char *order = "vlifhwkabcdegjmnopqrstuxyz";
    bd.elem_kind.assign(d.elements.size(), -1);
    bd.elem_inst.assign(d.elements.size(), 0);
    for (const char *o = order; *o; o++)
        for (size_t k = 0; k < d.elements.size(); k++)
            if (d.elements[k].letter == *o) bd.add_order.push_back((uint32_t)k);
    uint32_t next_row = 1;  // rows are numbered in order of allocation
    // the template holds the nominal circuit: expressions at the deck's own
    // .param values (points overwrite them per lane)
    std::map<std::string, double> nominal;
    std::string nerr;
    param_env(d, {}, nominal, nerr);
    for (uint32_t k : bd.add_order) {
        Element &e = d.elements[k];
        uint32_t rows[4];
        for (size_t p = 0; p < e.nodes.size(); p++) {
            const std::string &nn = e.nodes[p];
            if (espice_lanes_template_node(t, nn.c_str(), &rows[p]) != ESPICE_OK) {
                err = "BadNode " + nn;
                return false;
            }
            if (nn != "0" && nn != "gnd") bd.rows[nn] = rows[p];
            if (rows[p] + 1 > next_row) next_row = rows[p] + 1;
        }
        int kind;
        std::vector<std::pair<uint32_t, Value>> prm;
        switch (e.letter) {
        case 'v':
        case 'i': {
            // a current source carries the voltage source's slots
            // (models/isource.va:63-130)
            kind = e.letter == 'v' ? ESPICE_LANES_VSOURCE : ESPICE_LANES_ISOURCE;
            if (e.wave == 4) {
                // PWL: point count (slot 15), r= (16), td= (17), times from
                // slot 18 and values from slot 34
                prm.emplace_back(1, lit(4.0));
                const size_t np = e.pwl.size() / 2;
                prm.emplace_back(15, lit((double)np));
                if (e.pwl_r.given) prm.emplace_back(16, e.pwl_r);
                if (e.pwl_td.given) prm.emplace_back(17, e.pwl_td);
                for (size_t k = 0; k < np; k++) {
                    prm.emplace_back((uint32_t)(18 + k), e.pwl[2 * k]);
                    prm.emplace_back((uint32_t)(34 + k), e.pwl[2 * k + 1]);
                }
                if (e.value.given) prm.emplace_back(0, e.value);
            } else if (e.wave == 6) {
                // ngspice's AM(VO VMO VMA FM FC TD PHASEM PHASEC), value
                // VO + (VMO + VMA sin(2 pi FM t' + PHASEM)) sin(2 pi FC t' +
                // PHASEC) on t' = t - TD, onto the models' AM(VA VO MF FC TD
                // PHASEC PHASES) (frontend/builder.zig:1547-1562 amArgs): the
                // delay moves into the phases; VMA = 0 holds the envelope at
                // sin(90 deg); FM and FC default to 5/TSTOP and 500/TSTOP; a
                // nonzero VO has no slot and is refused
                double a[8] = {0, 0, 1, 0, 0, 0, 0, 0};
                const bool tstop_num = d.tran && d.tstop.given && !d.tstop.expr;
                const double tstop = tstop_num ? d.tstop.num : 0.0;
                a[3] = 5.0 / tstop;
                a[4] = 500.0 / tstop;
                for (int q = 0; q < 8 && q < e.n_wave_args; q++) {
                    if (!e.wave_args[q].given) continue;
                    if (e.wave_args[q].expr) {
                        err = "UnresolvedParameter " + e.name + " AM arguments must be numbers";
                        return false;
                    }
                    a[q] = e.wave_args[q].num;
                }
                if (!tstop_num && (e.n_wave_args < 5)) {
                    err = "UnresolvedParameter " + e.name + " AM's FM and FC defaults need a numeric .tran stop";
                    return false;
                }
                const double vo = a[0], vmo = a[1], vma = a[2], fm = a[3], fc = a[4], td = a[5], phm = a[6], phc = a[7];
                if (vo != 0.0) {
                    err = "UnsupportedCard " + e.name + " AM with a nonzero VO offset";
                    return false;
                }
                const double ours[7] = {vma != 0.0 ? vma : vmo, vma != 0.0 ? vmo / vma : 0.0, vma != 0.0 ? fm : 0.0, fc,
                                        td, phc - 360.0 * fc * td, vma != 0.0 ? phm - 360.0 * fm * td : 90.0};
                prm.emplace_back(1, lit(6.0));
                for (int q = 0; q < 7; q++) prm.emplace_back((uint32_t)(64 + q), lit(ours[q]));
                if (e.value.given) prm.emplace_back(0, e.value);
            } else if (e.wave == 1 && e.n_wave_args == 8 && pulse_train(d, e, prm, err)) {
                // NP pulses within the run: a PWL of their corners
            } else if (e.wave) {
                if (!err.empty()) return false;
                // PULSE args go to slots 2..8, SIN args to 9..14, EXP args
                // to 50..55, SFFM's to 56..63, AM's to 64..70; a dc left unwritten is filled from the
                // waveform's t = 0 value when the lane constants are
                // prepared. EXP's time constants start at 0 on a card, so
                // the .tran step fills the omitted ones
                // (frontend/builder.zig:2860-2866)
                const uint32_t base = e.wave == 1 ? 2 : e.wave == 2 ? 9 : e.wave == 3 ? 50 : e.wave == 5 ? 56 : 64;
                prm.emplace_back(1, lit(double(e.wave)));
                if (e.wave == 3) {
                    if (!e.wave_args[3].given) prm.emplace_back(53, lit(0.0));
                    if (!e.wave_args[5].given) prm.emplace_back(55, lit(0.0));
                }
                for (int a = 0; a < (e.wave == 1 ? 7 : 8); a++)
                    if (e.wave_args[a].given) prm.emplace_back(base + a, e.wave_args[a]);
                if (e.value.given) prm.emplace_back(0, e.value);
            } else {
                prm.emplace_back(0, e.value.given ? e.value : lit(0.0));
            }
            break;
        }
        case 'r':
            kind = ESPICE_LANES_RESISTOR;
            prm.emplace_back(0, e.value);
            break;
        case 'c':
            kind = ESPICE_LANES_CAPACITOR;
            prm.emplace_back(0, e.value);
            break;
        case 'd':
        case 'm': {
            auto mi = d.models.find(e.model);
            if (mi == d.models.end()) {
                err = "UnknownModel " + e.model;
                return false;
            }
            const Model &m = mi->second;
            const bool mos = e.letter == 'm';
            if (mos != (m.type != "d")) {
                err = "ModelTypeMismatch " + e.name;
                return false;
            }
            kind = mos ? ESPICE_LANES_MOS1 : ESPICE_LANES_DIODE;
            const char *const *keys = mos ? MOS1_KEYS : DIODE_KEYS;
            const size_t nk = mos ? 36 : 19;
            if (mos) prm.emplace_back(0, lit(m.type == "nmos" ? 1.0 : -1.0));
            for (auto &[k, v] : m.kv) {
                const int idx = key_index(keys, nk, k);
                if (idx < 0 || (mos && idx == 0)) {
                    err = "UnsupportedModelParameter " + k;
                    return false;
                }
                prm.emplace_back((uint32_t)idx, v);
            }
            for (auto &[k, v] : e.kv) {
                const int idx = key_index(keys, nk, k);
                if (idx < 0 || (mos && idx == 0)) {
                    err = "UnsupportedInstanceParameter " + k;
                    return false;
                }
                prm.emplace_back((uint32_t)idx, v);
            }
            break;
        }
        default: err = "UnsupportedElement " + e.name; return false;
        }
        // template constants: the literals now; expressions are point values
        const int npk[] = {ESPICE_LANES_VSOURCE_PARAMS,   ESPICE_LANES_ISOURCE_PARAMS, ESPICE_LANES_RESISTOR_PARAMS,
                           ESPICE_LANES_CAPACITOR_PARAMS, ESPICE_LANES_DIODE_PARAMS,   ESPICE_LANES_MOS1_PARAMS};
        std::vector<double> pv(npk[kind]);
        for (int a = 0; a < npk[kind]; a++) pv[a] = NAN;
        for (auto &[idx, v] : prm) {
            if (idx >= pv.size()) {
                err = "InvalidParam " + e.name;
                return false;
            }
            std::string xerr;
            pv[idx] = v.expr ? eval_expr(*v.expr, nominal, xerr) : v.num;
            if (!xerr.empty()) pv[idx] = NAN;
        }
        uint32_t inst = 0;
        if (espice_lanes_template_add(t, (uint32_t)kind, rows, pv.data(), (size_t)npk[kind], &inst) != ESPICE_OK) {
            size_t req;
            char buf[128];
            espice_lanes_error_message(buf, sizeof buf, &req);
            err = std::string(buf) + " " + e.name;
            return false;
        }
        bd.elem_kind[k] = kind;
        bd.elem_inst[k] = inst;
        for (auto &[idx, v] : prm) bd.slots.push_back(ParamSlot{(uint32_t)kind, inst, idx, v, e.name});
        // a V source's branch current takes the next row when it is added,
        // and so does a diode's internal anode when its RS is positive
        if (e.letter == 'v') bd.vsrc_branch[e.name] = next_row++;
        if (e.letter == 'd' && pv[ESPICE_LANES_DIODE_RS] > 0.0) next_row++;
    }
    if (espice_lanes_template_freeze(t) != ESPICE_OK) {
        size_t req;
        char buf[128];
        espice_lanes_error_message(buf, sizeof buf, &req);
        err = buf;
        return false;
    }
    uint32_t n = 0, nnz = 0, tri = 0;
    espice_lanes_template_info(t, &n, &nnz, &tri);
    bd.n = n;
    if (n != next_row) {
        err = "RowCount";
        return false;
    }
    return true;
}

double slot_value(const Deck &d, const ParamSlot &sl, const std::map<std::string, double> &env, std::string &err,
                  uint32_t trial)
{
    if (!sl.value.expr) return sl.value.num;
    Draw dr;
    dr.seed = d.seed;
    dr.trial = trial;
    dr.owner = top_owner(sl.card.data(), sl.card.size());
    return eval_expr(*sl.value.expr, env, err, &dr);
}

}  // namespace esp
