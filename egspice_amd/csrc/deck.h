// A SPICE deck subset read into a lane template: the cards the points API
// and the CLI run on the lane engine (include/espice.h lists them).
//
// Reading follows the reference front end where results depend on it:
// numbers are ngspice's INPevaluate (src/frontend/lines.zig:150-262: one
// f64 mantissa of every digit, times a correctly rounded power of ten);
// text is case-folded; elements are added bucket by bucket in the letter
// order of NetBuilder.build (src/frontend/builder.zig:886), each bucket in
// card order, and nodes are numbered on first touch (rowOf :559-565), so
// patterns, column orders and therefore every rounding match a deck built
// by the reference; PULSE fields not given are resolved from .tran as
// resolvePulseDefaults does (:845-877).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

struct espice_lanes_template;

namespace esp {

// ngspice's INPevaluate of a number with its scale suffix (`10meg`, `2.5u`,
// `1mil`); false when `text` does not start with a number.
bool spice_number(const std::string &text, double &out);

// An expression over numbers, .param names and a few functions, parsed once
// and evaluated per point.
struct Expr;
using ExprP = std::shared_ptr<const Expr>;
ExprP parse_expr(const std::string &text, std::string &err);

// A Monte Carlo trial's draws (src/frontend/variants.zig:1716-1720,
// 1862-1895): trial 0 answers every distribution call (agauss, gauss, aunif,
// unif, limit) with its nominal; trial k >= 1 draws the call at site
// site_key(owner, ordinal), the ordinal counting the owner's calls in
// evaluation order (src/frontend/netlist.zig:3595-3601).
struct Draw {
    uint64_t seed = 1;
    uint32_t trial = 0;
    uint64_t owner = 0;
    uint32_t ordinal = 0;
};
double eval_expr(const Expr &e, const std::map<std::string, double> &env, std::string &err, Draw *draw = nullptr);
// Names an expression reads.
void expr_names(const Expr &e, std::vector<std::string> &out);

// A value as written on a card: a literal, or an expression.
struct Value {
    bool given = false;
    double num = 0;
    ExprP expr;  // null for a literal
};

struct Element {
    char letter = 0;
    std::string name;                 // full card name, folded
    std::vector<std::string> nodes;   // port nets
    std::string model;                // D and M cards
    Value value;                      // R, C, I value; V DC value
    int wave = 0;                     // V or I waveform: 0 none, 1 PULSE(v1 v2 td tr tf pw per),
                                      // 2 SIN(vo va freq td theta phase),
                                      // 3 EXP(v1 v2 td1 tau1 td2 tau2), 4 PWL(t1 v1 ...),
                                      // 5 SFFM(vo va fm mdi fc td phasem phasec), V only,
                                      // 6 AM(va vo mf fc td phasec phases)
    Value wave_args[8];
    int n_wave_args = 0;              // arguments the card gave
    std::vector<Value> pwl;           // PWL times and values, interleaved
    Value pwl_r, pwl_td;              // PWL r= and td=
    std::map<std::string, Value> kv;  // W=, L=, AD=, M=, area
    // V and I: the card's AC spec (frontend/builder.zig:3124-3151 sourceAc:
    // AC [mag [phase]], 1 and 0 when left off)
    bool ac = false;
    double ac_mag = 1.0, ac_phase = 0.0;
    int line = 0;
};

struct Model {
    std::string name, type;  // type: nmos, pmos, d
    std::map<std::string, Value> kv;
    int line = 0;
};

// One .meas card. kind: trig_targ, when, find, deriv, min, max, avg, rms,
// integ, param. A vector is v(node) or i(vsource).
struct Meas {
    std::string analysis, name, kind;
    struct Side {
        std::string vec;
        Value val, td, at, from, to;
        int rise = -1, fall = -1, cross = -1;  // -2 LAST
    } a, b;
    ExprP param;
    int line = 0;
};

struct Deck {
    std::string title;
    std::vector<std::pair<std::string, Value>> params;  // .param in card order
    std::vector<Element> elements;                       // card order
    std::map<std::string, Model> models;
    bool op = false, tran = false;
    Value tstep, tstop, tstart, tmax;
    std::vector<Meas> meas;
    std::vector<std::string> ignored;  // cards read and not run (.dc, .noise, .print, ...)
    // `.ac dec|oct|lin N fstart fstop` (core/numerics.zig:263-372
    // FreqSweep, the .ac card landing a dec grid on fstop): kind 0 dec,
    // 1 oct, 2 lin; -1 none
    // `.dc NAME start stop step` over one V or I source (dc.zig Options):
    // dc_src empty when the deck has none
    std::string dc_src;
    double dc_start = 0.0, dc_stop = 0.0, dc_step = 0.0;
    int ac_kind = -1;
    uint32_t ac_points = 0;
    double ac_fstart = 0.0, ac_fstop = 0.0;
    // Monte Carlo: `.option seed=` (the reference's default 1), `.dc
    // monte=N` (N operating-point trials, each warm from the nominal
    // operating point) and `.tran ... sweep monte=N`
    uint64_t seed = 1;
    bool has_seed = false;
    uint32_t monte_dc = 0, monte_tran = 0;
};

// Reads deck text; false with a named error (and its line) on anything
// outside the subset.
bool read_deck(const std::string &text, Deck &d, std::string &err);

// ---- the deck built as a lane template

// Device parameter slots of the lane engine that a point sets.
struct ParamSlot {
    uint32_t kind, inst, index;
    Value value;
    std::string card;  // the element's card name: its distribution calls' site owner
};

struct BuiltDeck {
    Deck deck;
    // card order of addition and, per element, its engine kind and instance
    std::vector<uint32_t> add_order;
    std::vector<int> elem_kind;
    std::vector<uint32_t> elem_inst;
    std::vector<ParamSlot> slots;        // every parameter given by the deck
    std::map<std::string, uint32_t> rows;  // node name -> row
    std::map<std::string, uint32_t> vsrc_branch;  // V card name -> branch row
    uint32_t n = 0;
};

// The `.dc` card's swept values (src/analysis/dc/dc.zig:657-665 sweepCount,
// accumulated by += step as ngspice steps them).
std::vector<double> dc_values(const Deck &d);

// The `.ac` card's frequencies in Hz (FreqSweep.count/at).
std::vector<double> ac_axis(const Deck &d);

// Builds `bd.deck` into template `t` (empty, not frozen) and freezes it.
bool build_deck(BuiltDeck &bd, struct ::espice_lanes_template *t, std::string &err);

// The .param environment of one point: every .param in card order, with
// `live` values taking the place of their cards' expressions; Monte Carlo
// trial `trial` (0: nominal) draws each parameter's distribution calls under
// the parameter's own site owner.
bool param_env(const Deck &d, const std::map<std::string, double> &live, std::map<std::string, double> &env,
               std::string &err, uint32_t trial = 0);

// A card's expression value for trial `trial` (its distribution calls owned
// by the card).
double slot_value(const Deck &d, const ParamSlot &s, const std::map<std::string, double> &env, std::string &err,
                  uint32_t trial);

}  // namespace esp
