/*
 * sepstrct.h -- host-visible state of the seplib C99 API.
 *
 * The three structures below are part of the drop-in ABI: callers read and
 * write their fields directly (atoms[i].f[k], ret.epot, sys.nupdate_neighb),
 * so member order, types and therefore every offset must equal the
 * reference's.  tests/test_abi_layout.py checks each offset against values
 * recorded from the reference headers.
 *
 * Reference: include/sepstrct.h:23-61 (seppart), :70-96 (sepmolinfo),
 * :103-135 (sepsys), :173-204 (sepret).
 *
 * Fields that the MI355X hot path reads or writes are marked [hot]; the rest
 * are carried for layout compatibility only.
 */
#ifndef SEP_MI355X_SEPSTRCT_H
#define SEP_MI355X_SEPSTRCT_H

#include "sepdef.h"
#include <stdbool.h>

/* One atom, array-of-structures.  568 bytes on LP64. */
typedef struct {
	double x[3];            /* [hot] position, wrapped into [0, L)          */
	double v[3];            /* [hot] velocity (leapfrog: half step ahead)   */
	double f[3];            /* [hot] total force                            */
	double a[3];            /* [hot] acceleration of the last leapfrog      */
	double m;               /* [hot] mass                                   */
	char type;              /* [hot] one-character type label               */
	double z;               /* [hot] point charge                           */

	int *neighb;            /* [hot] half neighbour list, -1 terminated     */

	int cross_neighb[3];    /* [hot] box crossings since the last rebuild   */
	int crossings[3];       /* [hot] box crossings since the start          */

	int molindex;           /* [hot] molecule index, -1 if none             */
	int bond[SEP_BOND];     /* [hot] bonded partners, -1 terminated         */
	int angle[SEP_ANGLE];   /* [hot] angle partners, -1 terminated          */
	int dihed[SEP_DIHED];   /* [hot] dihedral partners, -1 terminated       */

	double sigma;
	int *collid;
	double *colltime;

	double ldiff;

	double xtrue[3];

	double x0[3];
	double xn[3];           /* [hot] position at the last list rebuild      */

	double xp[3];

	double px[3];
	double pv[3];
	double pa[3];

	double randn[3];
	double prevf[3];
} seppart;

typedef seppart sepatom;

/* Intramolecular topology read from a .top file. */
typedef struct {
	unsigned num_mols;
	unsigned max_nuau;

	int flag_bonds, flag_angles, flag_dihedrals;

	unsigned num_bonds;     /* [hot] */
	unsigned *blist;        /* [hot] (a, b, type) per bond                  */
	unsigned num_btypes;

	unsigned num_angles;    /* [hot] */
	unsigned *alist;        /* [hot] (a, b, c, type) per angle              */
	unsigned num_atypes;

	unsigned num_dihedrals; /* [hot] */
	unsigned *dlist;        /* [hot] (a, b, c, d, type) per dihedral        */
	unsigned num_dtypes;

	double *blengths;       /* [hot] bond length per bond (output)          */
	double *angles;         /* [hot] angle per angle (output)               */
	double *dihedrals;      /* [hot] dihedral per dihedral (output)         */

	unsigned flag_Fij;
	float ***Fij;
	float ***Fiajb;
} sepmolinfo;

/* System description. */
typedef struct {
	long int npart;          /* [hot] */
	double length[3];        /* [hot] periodic box                           */
	double volume;           /* [hot] */

	int intgr_type;
	double dt;               /* [hot] */
	double tnow;             /* [hot] */
	unsigned ndof;           /* [hot] */
	double max_dist2;        /* [hot] largest squared displacement since xn  */

	double cf;               /* [hot] maximum interaction cutoff             */
	double lsubbox[3];       /* [hot] cell edge per axis                     */
	int nsubbox[3];          /* [hot] cells per axis                         */
	double skin;             /* [hot] */
	unsigned neighb_update;  /* [hot] SEP_BRUTE / SEP_NEIGHBLIST / SEP_LLIST_NEIGHBLIST */
	unsigned neighb_flag;    /* [hot] 1: rebuild at the next pair call       */
	unsigned nupdate_neighb; /* [hot] rebuild count                          */

	bool omp_flag;
	unsigned int nthreads;

	int fun_cstate;

	sepmolinfo *molptr;      /* [hot] */
} sep3D;

typedef sep3D sepsys;

/* Molecule record (only its size and layout matter here). */
typedef struct {
	double m;
	double x[3], xtrue[3];
	double v[3];
	unsigned nuau;
	int *index;
	double ete[3];
	double re2;
	double rg;
	double S[3];
	double s[3];
	double inertia[3][3];
	double w[3];
	int method_w;
	double pel[3];
	char type;
	unsigned nbonds;
	double *blength;
	int shake_flag;
} sepmol;

/* Per-step return values. */
typedef struct {
	double etot;
	double ekin;            /* [hot] */
	double epot;            /* [hot] */
	double ecoul;           /* [hot] */
	double sumv2;

	double P[3][3];         /* [hot] */
	double kin_P[3][3];     /* [hot] */
	double pot_P[3][3];     /* [hot] */
	double p;               /* [hot] */

	double P_mol[3][3];
	double kin_P_mol[3][3];
	double pot_P_mol[3][3];
	double p_mol;

	double pot_P_conservative[3][3];
	double pot_P_random[3][3];
	double pot_P_dissipative[3][3];
	double pot_P_bond[3][3];   /* [hot] */

	double pot_T_mol[3][3];
	double kin_T_mol[3][3];
	double T_mol[3][3];
	double t_mol;
} sepret;

#endif
