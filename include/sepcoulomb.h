/*
 * sepcoulomb.h -- shifted-force Coulomb (seplib C99 API).
 */
#ifndef SEP_MI355X_SEPCOULOMB_H
#define SEP_MI355X_SEPCOULOMB_H

#include "sepstrct.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Shifted-force Coulomb between the charges z over the current neighbour
 * list (list path) or all pairs (SEP_BRUTE).  Adds forces, adds the energy
 * to both ret->epot and ret->ecoul, adds the stress to ret->pot_P.  On the
 * list path the list is never rebuilt here and opt is ignored, as in the
 * reference's serial path.
 * Replaces include/sepcoulomb.h:39 (source/sepcoulomb.c:5-18, 96-160). */
void sep_coulomb_sf(seppart *ptr, double cf, sepsys *sys, sepret *retval,
		    const unsigned opt);

#ifdef __cplusplus
}
#endif
#endif
