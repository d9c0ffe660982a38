/*
 * sepomp.h -- the OpenMP Model II entry points of the seplib C99 API
 * (reference include/sepomp.h): one force term at a time, its forces added
 * to a caller-owned N x 3 array (a sep_matrix) instead of the atoms' f, so
 * that a program can evaluate independent terms in separate sections and
 * sum them itself (reference prgs/prg5.c:55-77).
 *
 * Each term runs on the device like its sep_* counterpart and returns when
 * its forces are in ftot.  Energies and stress are not computed, as in the
 * reference.  Synchronised mode only: on a resident system the calls fail
 * with SEP_GPU_ESTATE (sep_gpu_run with SEP_GPU_RUN_CONCURRENT is the
 * resident form of the same structure).
 */
#ifndef SEPOMP_H
#define SEPOMP_H

#include <stddef.h>

#include "sepstrct.h"

/* reference include/sepomp.h:32, source/sepomp.c:16-33: builds the
 * neighbour list for opt (SEP_ALL, SEP_NEIGHB_EXCL_BONDED,
 * SEP_NEIGHB_EXCL_SAME_MOL) if sys->neighb_flag is set, then clears it */
void sep_omp_make_neighb(seppart *ptr, double cf, sepsys *sys, const unsigned opt);

/* reference include/sepomp.h:46, source/sepomp.c:35-79: the forces of
 * fun (sep_lj, sep_lj_shift or sep_wca) between types[0] and types[1] */
void sep_omp_pairs(double **ftot, const seppart *ptr, const char *types, double cf,
		   double (*fun)(double, char), const sepsys *sys);

/* reference include/sepomp.h:62, source/sepomp.c:81-133: LJ with
 * param = {cf, sigma, epsilon} in the reference's form
 * ft = 48/epsilon (sigma/r^2)^3 ((sigma/r^2)^3 - 1/2) sigma/r^2 */
void sep_omp_pairs_lj(double **ftot, const seppart *ptr, const char *types,
		      const double *param, const sepsys *sys);

/* reference include/sepomp.h:75, source/sepomp.c:136-176: shifted-force
 * Coulomb forces over the current list */
void sep_omp_coulomb(double **ftot, seppart *ptr, double cf, sepsys *sys);

/* reference include/sepomp.h:87, source/sepomp.c:179-212: harmonic bonds */
void sep_omp_bond(double **ftot, seppart *aptr, int type, const double lbond,
		  const double ks, sepsys *sys);

/* reference include/sepomp.h:100, source/sepomp.c:215-261: the
 * cosine-squared angle force, f = -k (cos - cos(pi - angle0)) */
void sep_omp_angle(double **ftot, seppart *ptr, int type, const double angle0,
		   const double k, sepsys *sys);

/* reference include/sepomp.h:112, source/sepomp.c:264-328: Ryckaert-
 * Bellemans torsions */
void sep_omp_torsion(double **ftot, seppart *ptr, int type, const double g[6],
		     sepsys *sys);

/* the N x 3 arrays the Model II terms add into: sep_matrix, sep_matrix_set
 * and sep_free_matrix, declared where the reference declares them */
#include "separray.h"

#endif
