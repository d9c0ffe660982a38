/*
 * seporacle.h -- CPU oracle for the seplib per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain serial FP64 C99 restatement of
 * the reference algorithm, used by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg as the checker.  The product library
 * (libsep.so) never links, loads or calls it.
 *
 * Every function works on the same host structures as the API (seppart AoS,
 * sepsys, sepret from include/sepstrct.h) and cites the reference lines it
 * restates.  The names carry an orc_ prefix so the oracle can be loaded in
 * the same process as libsep.so and the compiled reference.
 *
 * Pinning (how each function is checked independently of itself) is listed
 * in the header of seporacle.c.
 */
#ifndef SEP_ORACLE_H
#define SEP_ORACLE_H

#include "sepstrct.h"

typedef double (*orc_pot_fn)(double r2, char opt);

/* allocation (source/sepinit.c:16-56 defaults, zero-filled) */
seppart *orc_alloc(long npart, int nneighb);
void orc_free(seppart *p, long npart);
sepsys orc_sys_setup(double lx, double ly, double lz, double cf, double dt,
		     long npart, unsigned update);
void orc_free_sys(sepsys *s);

/* potentials (source/sepmisc.c:115-164) */
double orc_lj(double r2, char opt);
double orc_lj_shift(double r2, char opt);
double orc_wca(double r2, char opt);

/* return values (source/sepret.c:19-82) */
void orc_reset_retval(sepret *r);
void orc_pressure_tensor(sepret *r, const sepsys *s);
void orc_reset_force(seppart *p, sepsys *s);

/* neighbour lists: returns 0, or -1 on overflow / cell index error */
int orc_build_list(seppart *p, sepsys *s, unsigned opt, int nneighb);

/* forces; return 0 or -1 (cutoff above sys->cf, list error) */
int orc_force_pairs(seppart *p, const char *types, double cf, orc_pot_fn fun,
		    sepsys *s, sepret *r, unsigned opt, int nneighb);
int orc_force_lj(seppart *p, const char *types, const double *param,
		 sepsys *s, sepret *r, unsigned opt, int nneighb);
void orc_coulomb_sf(seppart *p, double cf, sepsys *s, sepret *r, unsigned opt);

/* bonded forces (source/sepmol.c:372-587) */
void orc_stretch_harmonic(seppart *p, int type, double lbond, double ks,
			  sepsys *s, sepret *r);
void orc_angle_harmonic(seppart *p, int type, double angle0, double k,
			sepsys *s, sepret *r);
void orc_angle_cossq(seppart *p, int type, double angle0, double k,
		     sepsys *s, sepret *r);
void orc_torsion_ryckaert(seppart *p, int type, const double *g, sepsys *s,
			  sepret *r);
/* OpenMP Model II terms (source/sepomp.c:35-328): forces added to ftot
 * (npart row pointers of 3 doubles) only -- no energy, stress or per-term
 * geometry -- the pair terms over the half lists already in p[i].neighb */
int orc_omp_make_neighb(seppart *p, double cf, sepsys *s, unsigned opt, int nneighb);
void orc_omp_pairs(double **ftot, const seppart *p, const char *types, double cf,
		   orc_pot_fn fun, const sepsys *s);
void orc_omp_pairs_lj(double **ftot, const seppart *p, const char *types,
		      const double *param, const sepsys *s);
void orc_omp_coulomb(double **ftot, const seppart *p, double cf, const sepsys *s);
void orc_omp_bond(double **ftot, const seppart *p, int type, double lbond, double ks,
		  const sepsys *s);
void orc_omp_angle(double **ftot, const seppart *p, int type, double angle0, double k,
		   const sepsys *s);
void orc_omp_torsion(double **ftot, const seppart *p, int type, const double *g,
		     const sepsys *s);

/* .top reader (source/sepmol.c:22-369); returns 0 or -1 */
int orc_read_topology(seppart *p, const char *file, sepsys *s);

/* integrators (source/sepintgr.c:18-88, 149-168) */
void orc_nosehoover(seppart *p, double Td, double *alpha, double tau,
		    sepsys *s);
void orc_leapfrog(seppart *p, sepsys *s, sepret *r);

/* half-list helpers for tests: number of entries of atom i's list */
int orc_list_len(const seppart *p, long i);
long orc_half_pairs(const seppart *p, long npart, int *out, long maxpairs);

#endif
