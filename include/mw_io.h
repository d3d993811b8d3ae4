/*
 * mw_io.h -- host-side readers for the reference's input files and the
 * evaluation-state set-up, exported by libmw_separation_hip.so.
 *
 * These are the calls the reference client makes before it evaluates
 * (src/milkyway_client_app.c:218-254).  They run on the host only and never
 * touch the GPU.  They accept the reference's file formats; where the
 * reference's fscanf readers silently leave fields unset on malformed input,
 * these return an error instead.
 */
#ifndef MW_IO_H
#define MW_IO_H

#include "mw_separation_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces read_astronomy_parameters (milkyway_separation/cpu/parameters.c:60-89,
 * format fread_astronomy_parameters, parameters.c:107-228; arrays
 * util/io_util.c:55-141).  Fills *ap, allocating its arrays.  Returns 0, or 1
 * with a message on stderr when the file is missing or malformed. */
int mw_read_astronomy_parameters(const char* filename, ASTRONOMY_PARAMETERS* ap);

/* Replaces write_astronomy_parameters (parameters.c:91-105) with the format of
 * fwrite_astronomy_parameters (parameters.c:230-305; arrays util/io_util.c:30-50).
 * Returns 0, or 1 with a message on stderr. */
int mw_write_astronomy_parameters(const char* filename, const ASTRONOMY_PARAMETERS* ap);

/* Replaces free_parameters (parameters.c:27-58). */
void mw_free_astronomy_parameters(ASTRONOMY_PARAMETERS* ap);

/* Replaces read_star_points (milkyway_separation/cpu/star_points.c:26-81):
 * a count line, then one "l b r" triple per line.  Returns 0, or 1 with a
 * message on stderr. */
int mw_read_star_points(const char* filename, STAR_POINTS* sp);

/* Replaces write_star_points (star_points.c:53-67, format :83-93). */
int mw_write_star_points(const char* filename, const STAR_POINTS* sp);

/* Replaces free_star_points (star_points.c:96-104). */
void mw_free_star_points(STAR_POINTS* sp);

/* Replaces initialize_state (milkyway_separation/cpu/evaluation_state.c:146-168)
 * with initialize_integral_area (evaluation_state.c:97-124).  Returns 0, or 1
 * when an allocation fails. */
int mw_initialize_state(const ASTRONOMY_PARAMETERS* ap, const STAR_POINTS* sp,
                        EVALUATION_STATE* es);

/* Replaces free_state (evaluation_state.c). */
void mw_free_state(EVALUATION_STATE* es);

/* Replaces write_checkpoint (milkyway_separation/cpu/evaluation_state.c:214-257,
 * areas :33-66): the state's totals, star-pass position and every area's
 * geometry, calculation range and partial integrals, to filename (the
 * reference resolves "astronomy_checkpoint" through BOINC).  Returns 0, or 1
 * with a message on stderr. */
int mw_write_checkpoint(const char* filename, const EVALUATION_STATE* es);

/* Replaces read_checkpoint (evaluation_state.c:259-290, areas :68-95) into a
 * state set up by mw_initialize_state for the same model.  A missing file is
 * not an error (returns 0 and leaves es unchanged), as in the reference; a
 * file that does not parse or does not match es's stream and region counts
 * returns 1. */
int mw_read_checkpoint(const char* filename, EVALUATION_STATE* es);

/* Replaces calculate_progress (evaluation_optimized.c:301-333): the fraction
 * of integral cells and stars done, from current_integral, each area's
 * current_calculation and current_star_point. */
double mw_calculate_progress(const EVALUATION_STATE* es);

/* Replaces get_optimized_parameter_count (parameters.c:307-330). */
int mw_get_optimized_parameter_count(const ASTRONOMY_PARAMETERS* ap);

/* Replaces set_astronomy_parameters (parameters.c:332-363): writes the
 * optimised-parameter vector into ap, in the reference's order. */
void mw_set_astronomy_parameters(ASTRONOMY_PARAMETERS* ap, const double* parameters);

#ifdef __cplusplus
}
#endif

#endif /* MW_IO_H */
