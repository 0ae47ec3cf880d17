/* espice.h — the points subset of ESPice's public C API, served by the MI355X
 * lane engine (egspice_amd/libespice_lanes.so).
 *
 * One deck is a template; each point is a row of its live .param values;
 * every point runs the deck's .op and .tran on GPU lanes (one workgroup per
 * point) and its .meas cards become device passes. The declarations, struct
 * layouts, status and reason codes are those of the reference header
 * (include/espice.h), restricted to the calls below; a caller written for
 * the reference's points API links against this library unchanged.
 *
 * Deck subset (ngspice dialect): V and I (DC, PULSE, SIN, EXP, SFFM, AM,
 * PWL; I's SFFM refused), R, C, D, M (LEVEL=1) elements; .model NMOS/PMOS/D;
 * .param with {expr} and 'expr';
 * .op; .tran tstep tstop [tstart [tmax]]; .ac dec|oct|lin N fstart fstop
 * with the sources' AC mag [phase]; .dc NAME start stop step over one V or I
 * source (a nested sweep refused); .dc monte=N trials; .meas tran TRIG/TARG,
 * WHEN, FIND/DERIV ... AT=, MIN, MAX, AVG, RMS, INTEG [FROM= TO=], PARAM=;
 * .meas op FIND. Any other card that would change an analysis result fails
 * create with a named error. */
#ifndef ESPICE_H
#define ESPICE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/espice.h:14 */
#define ESPICE_ABI_VERSION 1u

typedef struct espice_problem espice_problem;
typedef uint32_t espice_status;
/* include/espice.h:33-38 */
enum {
    ESPICE_OK = 0, ESPICE_INVALID_ARGUMENT = 1, ESPICE_BUFFER_TOO_SMALL = 2,
    ESPICE_OUT_OF_MEMORY = 3, ESPICE_INVALID_QUERY = 4,
    ESPICE_RESULT_UNAVAILABLE = 5, ESPICE_FAILED = 6, ESPICE_ABI_MISMATCH = 7,
    ESPICE_INVALID_HANDLE = 8
};
/* include/espice.h:41-47 */
enum { ESPICE_FILE = 0, ESPICE_BYTES = 1 };
enum { ESPICE_NGSPICE = 0, ESPICE_HSPICE = 1, ESPICE_SPECTRE = 2 };
enum { ESPICE_CPU = 0, ESPICE_AUTO = 1, ESPICE_CUDA = 2, ESPICE_HIP = 3 };

/* include/espice.h:91-109: the same layout. Only ESPICE_HIP and
 * ESPICE_AUTO run (both on the GPU lanes; ESPICE_AUTO without a GPU fails
 * like an explicit HIP, since this library has no CPU path). output_format,
 * output_path, zig_path, share_path and the display callback are accepted
 * and unused: points deliver their results through the copy calls. */
typedef struct { const char *data; size_t len; } espice_bytes;
typedef struct {
    uint32_t abi_version, struct_size;
    uint32_t source_kind, dialect, backend, explicit_gpu;
    uint32_t output_format, max_parallel;
    espice_bytes source, origin, output_path;
    espice_bytes zig_path, share_path;
    const espice_bytes *live_params;
    size_t live_param_count;
    void (*display)(const char *text, size_t len, void *user);
    void *display_user;
} espice_create_options;

/* include/espice.h:260-284 */
typedef enum {
    ESPICE_REASON_OK = 0,
    ESPICE_REASON_NO_CONVERGENCE = 1,
    ESPICE_REASON_TIMESTEP_TOO_SMALL = 2,
    ESPICE_REASON_SINGULAR_MATRIX = 3,
    ESPICE_REASON_MEASURE_NOT_FOUND = 4,
    ESPICE_REASON_MEASURE_UNRESOLVED = 5,
    ESPICE_REASON_ABORTED = 6,
    ESPICE_REASON_MODEL_BIN_NOT_FOUND = 7,
    ESPICE_REASON_TOPOLOGY_CHANGED = 8,
    ESPICE_REASON_INVALID_PARAMETER = 9,
    ESPICE_REASON_DEVICE_REFUSED = 10,
    ESPICE_REASON_OTHER = 255
} espice_reason;

/* include/espice.h:286-308: the same 48-byte layout and meanings. The
 * lanes measure every .tran card as they run and keep no waveform, so
 * measures_only changes nothing; early_stop ends a point's transient once
 * its measures resolve (the values are those of the whole run);
 * abort_after (NaN off) fails a point with ESPICE_REASON_ABORTED once t
 * passes it while abort_measure's card (or, unnamed, any card) is still
 * unresolved (src/analysis/tran/tran.zig:1059-1064); warm_start starts
 * every point's operating point from the template's nominal one
 * (src/analysis/dc/op.zig:588, 671). tangent_step must be 0
 * (INVALID_ARGUMENT "TangentsUnavailable"). Points of one append share
 * their options; points of appends with different options run in
 * separate lane batches. */
typedef struct {
    size_t struct_size;
    espice_bytes abort_measure;
    double abort_after;
    uint8_t measures_only;
    uint8_t early_stop;
    uint8_t warm_start;
    double tangent_step;
} espice_points_options;

/* include/espice.h:157 */
uint32_t espice_abi_version(void);
/* include/espice.h:160 */
void espice_default_options(espice_create_options *options);
/* include/espice.h:168-169: parses the deck and builds its lane template;
 * nothing runs yet. diagnostic receives the error name on failure. */
espice_status espice_create(const espice_create_options *options, espice_problem **out, char *diagnostic,
    size_t diagnostic_capacity);
/* include/espice.h:172 */
espice_status espice_destroy(espice_problem *problem);
/* include/espice.h:173 */
espice_status espice_query_count(espice_problem *, uint32_t *out);
/* include/espice.h:202: runs every appended point not yet run, in lane
 * batches of at most espice_lane_width points. */
espice_status espice_run_all(espice_problem *);
/* include/espice.h:239 */
espice_status espice_error_message(espice_problem *, char *, size_t, size_t *required);
/* include/espice.h:325-326 */
espice_status espice_append_points(espice_problem *, const double *values, size_t count,
    const espice_points_options *options, uint32_t *first_id);
/* include/espice.h:338-339 */
espice_status espice_copy_point_measures(espice_problem *, uint32_t first_id, size_t count,
    const espice_bytes *names, size_t measure_count, double *out, uint32_t *reasons);
/* include/espice.h:357 */
espice_status espice_release_points(espice_problem *, uint32_t first_id, size_t count);
/* include/espice.h:362: the GPU lane cap, 16384. */
espice_status espice_lane_width(espice_problem *, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif
