/*
 * sepinit.h -- allocation and system setup (seplib C99 API).
 */
#ifndef SEP_MI355X_SEPINIT_H
#define SEP_MI355X_SEPINIT_H

#include <stddef.h>
#include "sepstrct.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Allocates npart atoms, each with an nneighb-entry half neighbour list, and
 * sets the reference defaults (type 'A', m = 1, molindex -1, bonded partner
 * arrays -1).  Unlike the reference, xn/v/f are zero-filled rather than left
 * uninitialised.
 * Replaces include/sepinit.h:29 (source/sepinit.c:16-56). */
seppart *sep_init(size_t npart, size_t nneighb);

/* Frees the atoms, their lists and any device mirror the library holds for
 * them.  Replaces include/sepinit.h:36 (source/sepinit.c:58-65). */
void sep_close(seppart *ptr, size_t npart);

/* Reads "N / Lx Ly Lz / type x y z vx vy vz m q" text and returns a fresh
 * atom array.  Replaces include/sepinit.h:48 (source/sepinit.c:69-109). */
seppart *sep_init_xyz(double *lbox, int *npart, const char *file, char verbose);

/* Builds the system record: cell grid nsubbox = (int)(L/(maxcf+0.25)),
 * skin 0.25, ndof 3N-3, neighb_flag 1.  Replaces include/sepinit.h:77
 * (source/sepinit.c:246-305). */
sepsys sep_sys_setup(double lengthx, double lengthy, double lengthz,
		     double maxsyscf, double dt, size_t npart, size_t update);

/* Frees the topology lists held by sys->molptr.  Replaces
 * include/sepinit.h:84 (source/sepinit.c:308-315). */
void sep_free_sys(sepsys *ptr);

#ifdef __cplusplus
}
#endif
#endif
