// Compact models compiled from Verilog-A by egspice_amd.vagen (the lane
// engine's own compiler; the reference compiles the same sources with
// VerA). Each generated header gives, per model, its unknowns (ports then
// internal nodes), the static G and C patterns as row masks, the collapse
// pairs (V(a, b) <+ 0 under a parameter condition), the setup function
// (parameters with defaults and range checks, every bias-independent
// statement, into per-instance prepared constants) and the evaluations:
// currents and their Jacobian, and for the transient also the charges,
// their Jacobian and one value per ddt() site. The header is C; here it is
// compiled as ESP_HD code with the engine's own elementary functions, so a
// lane's device evaluation is bitwise its host evaluation.
//
//   psp103   models/psp103.va (module PSP103VA; D G S B, internal NOI GP SI
//            DI BP BI BS BD; PSP 103.8.2 with JUNCAP 200.6.2)
//   bsim4va  models/bsim4va.va (BSIM 4.8; D G S B, internal DI GI SI BI GM
//            SBULK DBULK, most collapsed by rgateMod/rbodyMod/rdsMod); its
//            $limit group limits the gate, drain and bulk legs
#pragma once

#include "emath.h"
#include "limiting.h"

namespace esp {

#define VA_FN ESP_HD
#define VA_TABLE constexpr
#define VA_HOST_TABLE static const
#define VA_RESTRICT __restrict__
#define VA_EXP e_exp
#define VA_LOG e_log
#define VA_LOG10(x) (e_log(x) * 0.43429448190325182765)
#define VA_POW e_pow
#define VA_SQRT __builtin_sqrt
#define VA_FABS fabs_
#define VA_FLOOR __builtin_floor
#define VA_CEIL __builtin_ceil
#define VA_FMA __builtin_fma
// parameter arithmetic only (a model's setup, which the lane engine runs on
// the host): the C library's
#define VA_ATAN atan
#define VA_INF __builtin_huge_val()
#define VA_IDIV(a, b) ((b) ? (a) / (b) : 0)
#define VA_IMOD(a, b) ((b) ? (a) % (b) : 0)
// real to integer: nearest, ties away from zero (LRM 4.2.1.1)
#define VA_RTOI(x) ((int)__builtin_round(x))

#define VA_FETLIM fetlim
#define VA_LIMVDS limvds
#define VA_PNJLIM pnjlim
#define VA_SNAP snap_limited

#include "gen/psp103.h"
#include "gen/bsim4va.h"

#undef VA_FN
#undef VA_TABLE
#undef VA_HOST_TABLE
#undef VA_RESTRICT
#undef VA_EXP
#undef VA_LOG
#undef VA_LOG10
#undef VA_POW
#undef VA_SQRT
#undef VA_FABS
#undef VA_FLOOR
#undef VA_CEIL
#undef VA_FMA
#undef VA_ATAN
#undef VA_INF
#undef VA_IDIV
#undef VA_IMOD
#undef VA_RTOI
#undef VA_FETLIM
#undef VA_LIMVDS
#undef VA_PNJLIM
#undef VA_SNAP

// $simparam("gmin") the models read: the deck's gmin option, 1e-12 S by
// default (src/core/numerics.zig:386, src/device/abi.zig:408).
constexpr double VA_SIMPARAM_GMIN = 1e-12;

}  // namespace esp
