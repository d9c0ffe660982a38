/*
 * sepmol.h -- molecular topology and bonded forces (seplib C99 API).
 */
#ifndef SEP_MI355X_SEPMOL_H
#define SEP_MI355X_SEPMOL_H

#include "sepstrct.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reads the [ bonds ], [ angles ] and [ dihedrals ] sections of a .top
 * file into sys->molptr and the per-atom partner arrays.
 * Replaces include/sepmol.h:18 (source/sepmol.c:22-369). */
void sep_read_topology_file(sepatom *aptr, const char *file, sepsys *sys,
			    char opt);

/* Harmonic bonds of one type: U = ks/2 (d - lbond)^2.  Adds forces, energy,
 * ret->pot_P and ret->pot_P_bond; stores each bond length in
 * sys->molptr->blengths.
 * Replaces include/sepmol.h:31 (source/sepmol.c:372-414). */
void sep_stretch_harmonic(sepatom *aptr, int type, const double lbond,
			  const double ks, sepsys *sys, sepret *ret);

/* Cosine-squared angles: U = k/2 (cos theta' - cos(pi - angle0))^2.
 * Replaces include/sepmol.h:57 (source/sepmol.c:418-467). */
void sep_angle_cossq(sepatom *ptr, int type, const double angle0,
		     const double k, sepsys *sys, sepret *ret);

/* Harmonic angles, U = k/2 (theta - angle0)^2, with the reference's force
 * expression (which omits the 1/sin(theta) factor of the gradient).
 * Replaces include/sepmol.h:70 (source/sepmol.c:469-516). */
void sep_angle_harmonic(sepatom *ptr, int type, const double angle0,
			const double k, sepsys *sys, sepret *ret);

/* Ryckaert-Bellemans torsion, U = sum_n g[n] cos^n(psi).
 * Replaces include/sepmol.h:83 (source/sepmol.c:520-587). */
void sep_torsion_Ryckaert(sepatom *ptr, int type, const double g[6],
			  sepsys *sys, sepret *ret);

#ifdef __cplusplus
}
#endif
#endif
