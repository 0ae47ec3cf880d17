// Library entry points for the command line front end and its tests; not
// part of the reference's API (include/espice.h).
#pragma once

#include "../../include/espice.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Point 0's .meas lines as the reference prints them (mode 0), its
 * measures as a JSON object (mode 1), or the run's plots (point 0's AC
 * response, a `.dc monte` deck's Monte Carlo table) as the reference's
 * writers write them: mode 2 SPICE3 print, 3 CSV, 4 ASCII raw, 5 binary
 * raw (binary bytes: *required - 1 of them count); buffer protocol as
 * espice_copy_*. */
espice_status espice_cli_report(espice_problem *problem, int mode, char *buf, size_t cap, size_t *required);
/* ngspice's INPevaluate of one number with its scale suffix; 1 when `text`
 * is a number, else 0. */
int espice_spice_number(const char *text, double *out);

#ifdef __cplusplus
}
#endif
