/*
 * sepmisc.h -- potentials, wrap macro and small setters (seplib C99 API).
 */
#ifndef SEP_MI355X_SEPMISC_H
#define SEP_MI355X_SEPMISC_H

#include <stdio.h>

#include "sepstrct.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Square and single-image minimum-image wrap, as the reference's macros
 * (include/sepmisc.h:69, 81-85): at most one box length is added or
 * subtracted, with strict comparisons against +-L/2. */
#define sep_Sq(x) ((x) * (x))
#define sep_Wrap(x, y)                          \
	{                                       \
		if ((x) > 0.5 * (y)) (x) -= (y);        \
		else if ((x) < -0.5 * (y)) (x) += (y);  \
	}

/* Plain LJ (eps = sigma = 1): 'f' returns 48 r^-14 - 24 r^-8 (force / r),
 * 'u' returns 4 (r^-12 - r^-6).  Replaces include/sepmisc.h:115
 * (source/sepmisc.c:115-130).  Its address selects the device kernel in
 * sep_force_pairs. */
double sep_lj(double r2, char opt);

/* LJ shifted by SEP_LJCF2 (zero energy at rc = 2.5).  Replaces
 * include/sepmisc.h:122 (source/sepmisc.c:132-147). */
double sep_lj_shift(double r2, char opt);

/* WCA: LJ shifted by +1 (use with rc = 2^(1/6)).  Replaces
 * include/sepmisc.h:129 (source/sepmisc.c:149-164). */
double sep_wca(double r2, char opt);

/* f = 0 for every atom and sys->max_dist2 = 0.  Replaces
 * include/sepmisc.h:245 (source/sepmisc.c:393-400). */
void sep_reset_force(seppart *ptr, sepsys *sys);

/* Flushes stdout (include/sepmisc.h:39). */
#define SEP_FLUSH fflush(stdout)

/* Sets sys->omp_flag and sys->nthreads.  The device path ignores the thread
 * count.  Replaces include/sepmisc.h:303 (source/sepmisc.c:1116-1123). */
void sep_set_omp(unsigned nthreads, sepsys *sys);

/* Sets the list skin; like the reference it does not recompute the cell
 * grid.  Replaces include/sepmisc.h:311 (source/sepmisc.c:1125-1129). */
void sep_set_skin(sepsys *sys, double value);

/* Writes the atoms whose type is one of the characters of partnames in the
 * sep_init_xyz format ("N\nLx Ly Lz\n", then "type x y z vx vy vz m q"
 * per atom, %.15f), fopen mode `mode`.  A resident system is synchronised
 * first.  Replaces include/sepmisc.h:164 (source/sepmisc.c:536-573). */
void sep_save_xyz(seppart *ptr, const char *partnames, const char *file, char *mode,
		  sepsys *sys);

/* Raw binary dump of npart seppart records; a resident system is
 * synchronised first.  Replaces include/sepmisc.h:144
 * (source/sepmisc.c:319-330). */
void sep_save(seppart *ptr, size_t npart, const char *file);

/* Reads npart raw seppart records and allocates each atom a new neighbour
 * array of nneighb ints (the pointers read from the file are replaced); a
 * resident system's device state is replaced by the loaded atoms and its
 * list is rebuilt at the next pair call.
 * Replaces include/sepmisc.h:154 (source/sepmisc.c:333-354). */
void sep_load(seppart *ptr, int nneighb, size_t npart, const char *file);

/* Diagnostics printed as "sep-error -> ..." / "sep-warning -> ..." on
 * stdout.  sep_error does not exit (the reference's -DOCTAVE behaviour,
 * source/sepmisc.c:59-65); it records SEP_GPU_EHOST for sep_gpu_last_error.
 * Replaces include/sepmisc.h:101,108. */
void sep_error(char *str, ...);
void sep_warning(char *str, ...);

#ifdef __cplusplus
}
#endif
#endif
