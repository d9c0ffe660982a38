/*
 * sepprfrc.h -- short-range pair forces on the neighbour-list path
 * (seplib C99 API).
 */
#ifndef SEP_MI355X_SEPPRFRC_H
#define SEP_MI355X_SEPPRFRC_H

#include "sepstrct.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Pair forces of a potential given as fun(r2,'f'|'u') between the two type
 * labels in types[0..1].  Rebuilds the neighbour list first when
 * sys->neighb_flag is set (cell list + Verlet list on the device, with the
 * exclusion rule opt), then adds the pair forces to f, ASSIGNS the pair
 * energy to ret->epot and adds the configurational stress to ret->pot_P.
 * fun must be one of sep_lj, sep_lj_shift, sep_wca (dispatch is by address);
 * other potentials are rejected with SEP_GPU_EUNSUPPORTED.
 * Replaces include/sepprfrc.h:49 (source/sepprfrc.c:226-274, 94-224). */
int sep_force_pairs(seppart *ptr, const char *types, double cf,
		    double (*fun)(double, char), sepsys *sys,
		    sepret *retval, const unsigned opt);

/* Lennard-Jones pair forces with param = {rc, epsilon, sigma, aw}; the
 * energy is shifted to zero at rc and ADDED to ret->epot.
 * Replaces include/sepprfrc.h:69 (source/sepprfrc.c:743-780, 782-923). */
void sep_force_lj(seppart *ptr, const char *types,
		  const double *param, sepsys *sys,
		  sepret *retval, const unsigned opt);

#ifdef __cplusplus
}
#endif
#endif
