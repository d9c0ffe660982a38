/*
 * sep_forces.c -- the per-timestep functions of the seplib C99 API, served
 * by the device module: pair forces (list and brute-force paths),
 * shifted-force Coulomb, bonded terms, Nose-Hoover and leapfrog.
 *
 * Synchronised mode follows each reference function's contract on the
 * AoS: upload what it reads, run, write back what it changes.  Resident
 * mode only enqueues work on the device.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sephost.h"

/* Runs fn with a synchronised scratch slot and merges the slot into ret. */
static int sync_ret_begin(seph_ctx *c)
{
	return seph_dev_fail(sepdev_ret_zero(c->dev, c->ret_scratch));
}

static int sync_ret_end(seph_ctx *c, sepret *ret, int epot_assign, int bond_stress)
{
	double w[SEPDEV_RET_WORDS];
	int rc = seph_dev_fail(sepdev_ret_get(c->dev, c->ret_scratch, w));
	if (rc == 0)
		seph_ret_merge(ret, w, epot_assign, bond_stress);
	return rc;
}

static int ret_slot_for(seph_ctx *c, sepret *ret)
{
	return c->resident ? seph_ret_slot(c, ret) : c->ret_scratch;
}

/* ---- sep_reset_force: source/sepmisc.c:393-400 ------------------------ */

void sep_reset_force(seppart *ptr, sepsys *sys)
{
	seph_ctx *c = seph_find(ptr);
	if (c && c->resident) {
		seph_dev_fail(sepdev_zero_forces(c->dev));
		seph_dev_fail(sepdev_reset_max_dist(c->dev));
		return;
	}
	for (long i = 0; i < sys->npart; i++)
		ptr[i].f[0] = ptr[i].f[1] = ptr[i].f[2] = 0.0;
	sys->max_dist2 = 0.0;
}

/* ---- shared pair-call driver ------------------------------------------ */

typedef struct {
	const char *types;
	double cf;          /* interaction cutoff                    */
	unsigned opt;       /* exclusion rule                        */
	sepdev_lj_args lj;  /* constants; cf2/types filled here      */
	int epot_assign_list;
	const sepdev_ljtab_args *table; /* multi-type table instead of lj */
} pair_call;

/* True when the type filter of source/sepprfrc.c:116-121 keeps every pair:
 * both labels equal and every atom carries that label.  ("AB" never keeps
 * an A-A pair, so it always needs the filter.)  Cached per context until
 * the types are uploaded again. */
static int types_cover_all(seph_ctx *c, const char *types)
{
	if (types[0] != types[1])
		return 0;
	if (c->uniform_type_valid)
		return c->uniform_type == types[0];
	c->uniform_type = c->atoms[0].type;
	for (long i = 1; i < c->n; i++)
		if (c->atoms[i].type != c->uniform_type) {
			c->uniform_type = 0;
			break;
		}
	c->uniform_type_valid = c->resident;
	return c->uniform_type == types[0];
}

static int needs_partners(unsigned opt) { return opt == SEP_EXCL_BONDED; }

/* Brute-force and list paths of source/sepprfrc.c:226-274 and :743-780. */
static int run_pairs(seppart *ptr, sepsys *sys, sepret *ret, pair_call *pc)
{
	seph_ctx *c = seph_ctx_for(ptr, sys);
	if (!c)
		return sep_gpu_last_error();
	int rc;
	const int brute = sys->neighb_update == SEP_BRUTE;
	if (brute)
		sys->omp_flag = false;
	if (!c->resident) {
		if ((rc = seph_put_x(c)) || (rc = seph_put_f(c)) || (rc = seph_put_scalars(c)))
			return rc;
		if (needs_partners(pc->opt) && (rc = seph_put_partners(c)))
			return rc;
		if ((rc = sync_ret_begin(c)))
			return rc;
	} else if (needs_partners(pc->opt) && !c->partners_on_device) {
		if ((rc = seph_put_partners(c)))
			return rc;
	}

	sepdev_grid g;
	if (brute) {
		/* every call rebuilds at the interaction cutoff, bonds-only
		 * exclusion test (source/sepprfrc.c:35-42) */
		seph_grid_own(sys, pc->cf, pc->opt, 0, &g);
		rc = c->resident ? seph_build_resident(c, &g, 0)
				 : seph_build(c, sys, &g, 0);
		if (rc)
			return rc;
	} else if (c->resident) {
		if (sys->neighb_update == SEP_NEIGHBLIST) {
			seph_grid_own(sys, sys->cf + sys->skin, pc->opt, 0, &g);
			const double rl = sys->cf + sys->skin;
			g.rlist2 = rl * rl;
		} else {
			seph_grid_reference(sys, pc->opt, &g);
		}
		if ((rc = seph_build_resident(c, &g, 1)))
			return rc;
	} else if (sys->neighb_flag == 1) {
		if (sys->neighb_update == SEP_NEIGHBLIST) {
			seph_grid_own(sys, sys->cf + sys->skin, pc->opt, 0, &g);
			const double rl = sys->cf + sys->skin;
			g.rlist2 = rl * rl;
		} else {
			seph_grid_reference(sys, pc->opt, &g);
		}
		if ((rc = seph_build(c, sys, &g, 0)))
			return rc;
		sys->neighb_flag = 0;
	}

	const int slot = ret_slot_for(c, ret);
	if (slot < 0)
		return sep_gpu_last_error();
	int assign = 0;
	if (pc->table) {
		sepdev_ljtab_args ta = *pc->table;
		for (int k = 0; k < 3; k++)
			ta.L[k] = sys->length[k];
		if ((rc = seph_dev_fail(sepdev_lj_table(c->dev, &ta, slot))))
			return rc;
	} else {
		pc->lj.cf2 = pc->cf * pc->cf;
		for (int k = 0; k < 3; k++)
			pc->lj.L[k] = sys->length[k];
		pc->lj.t0 = pc->types[0];
		pc->lj.t1 = pc->types[1];
		pc->lj.typed = !types_cover_all(c, pc->types);
		pc->lj.epot_assign = brute ? 0 : pc->epot_assign_list;
		pc->lj.mixed = c->precision == SEP_GPU_MIXED;
		assign = pc->lj.epot_assign;
		if ((rc = seph_dev_fail(sepdev_lj(c->dev, &pc->lj, slot))))
			return rc;
	}
	if (c->resident)
		return 0;
	if ((rc = seph_get_f(c)))
		return rc;
	return sync_ret_end(c, ret, assign, 0);
}

int sep_force_pairs(seppart *ptr, const char *types, double cf,
		    double (*fun)(double, char), sepsys *sys, sepret *retval,
		    const unsigned opt)
{
	if (cf > sys->cf) {
		sep_warning("cutoff for an interaction cannot be larger than maximum cutoff");
		seph_fail(SEP_GPU_EARG, "cutoff %f above the system cutoff %f", cf, sys->cf);
		return SEP_FAILURE;
	}
	const int pot = seph_identify_potential(fun);
	if (pot == SEPH_POT_NONE) {
		sep_warning("sep_force_pairs: the device path evaluates sep_lj, sep_lj_shift and sep_wca only");
		seph_fail(SEP_GPU_EUNSUPPORTED, "potential function not available on the device");
		return SEP_FAILURE;
	}
	pair_call pc;
	memset(&pc, 0, sizeof pc);
	pc.types = types;
	pc.cf = cf;
	pc.opt = opt;
	/* sep_lj / sep_lj_shift / sep_wca are LJ with eps = sigma = aw = 1 and
	 * an energy offset (source/sepmisc.c:115-164) */
	pc.lj.eps48 = 48.0;
	pc.lj.eps4 = 4.0;
	pc.lj.awh = 0.5;
	pc.lj.aw = 1.0;
	pc.lj.sig2 = 1.0;
	pc.lj.ushift = pot == SEPH_POT_LJ_SHIFT ? SEP_LJCF2 : pot == SEPH_POT_WCA ? 1.0 : 0.0;
	pc.epot_assign_list = 1; /* source/sepprfrc.c:222 */
	int rc = run_pairs(ptr, sys, retval, &pc);
	if (rc)
		sep_warning("sep_force_pairs: %s", (char *)sep_gpu_last_error_message());
	return SEP_SUCCESS;
}

void sep_force_lj(seppart *ptr, const char *types, const double *param,
		  sepsys *sys, sepret *retval, const unsigned opt)
{
	const double cf = param[0], eps = param[1], sigma = param[2], aw = param[3];
	if (cf > sys->cf)
		sep_error("cutoff for an interaction cannot be larger than maximum cutoff");
	pair_call pc;
	memset(&pc, 0, sizeof pc);
	pc.types = types;
	pc.cf = cf;
	pc.opt = opt;
	/* source/sepprfrc.c:786-799; the shift uses the C library's pow */
	const double shift = 4.0 * eps * (pow(sigma / cf, 12.) - aw * pow(sigma / cf, 6.));
	pc.lj.eps48 = 48.0 * eps;
	pc.lj.eps4 = 4.0 * eps;
	pc.lj.awh = 0.5 * aw;
	pc.lj.aw = aw;
	pc.lj.sig2 = sigma * sigma;
	pc.lj.ushift = -shift;
	pc.epot_assign_list = 0; /* source/sepprfrc.c:922 */
	if (run_pairs(ptr, sys, retval, &pc))
		sep_warning("sep_force_lj: %s", (char *)sep_gpu_last_error_message());
}

/* ---- shifted-force Coulomb: source/sepcoulomb.c:5-160 ----------------- */

void sep_coulomb_sf(seppart *ptr, double cf, sepsys *sys, sepret *retval,
		    const unsigned opt)
{
	seph_ctx *c = seph_ctx_for(ptr, sys);
	int rc = 0;
	if (!c) {
		rc = sep_gpu_last_error();
		goto out;
	}
	if (!c->resident) {
		if ((rc = seph_put_x(c)) || (rc = seph_put_f(c)) || (rc = seph_put_scalars(c)))
			goto out;
		if ((rc = sync_ret_begin(c)))
			goto out;
	}
	if (sys->neighb_update == SEP_BRUTE) {
		/* all pairs within cf; the bonded exclusion of the brute path
		 * drops a pair only when exactly one partnership holds */
		if (opt == SEP_EXCL_BONDED && (!c->resident || !c->partners_on_device) &&
		    (rc = seph_put_partners(c)))
			goto out;
		sepdev_grid g;
		seph_grid_own(sys, cf, opt, 2, &g);
		rc = c->resident ? seph_build_resident(c, &g, 0)
				 : seph_build(c, sys, &g, 0);
		if (rc)
			goto out;
	}
	sepdev_coul_args a;
	a.cf = cf;
	a.cf2 = cf * cf;
	a.icf2 = 1.0 / a.cf2;
	a.icf = 1.0 / cf;
	for (int k = 0; k < 3; k++)
		a.L[k] = sys->length[k];
	a.mixed = c->precision == SEP_GPU_MIXED;
	const int slot = ret_slot_for(c, retval);
	if (slot < 0) {
		rc = sep_gpu_last_error();
		goto out;
	}
	if ((rc = seph_dev_fail(sepdev_coulomb(c->dev, &a, slot))))
		goto out;
	if (!c->resident) {
		if ((rc = seph_get_f(c)))
			goto out;
		rc = sync_ret_end(c, retval, 0, 0);
	}
out:
	if (rc)
		sep_warning("sep_coulomb_sf: %s", (char *)sep_gpu_last_error_message());
}

/* ---- bonded: source/sepmol.c:372-587 ---------------------------------- */

enum { BONDED_BOND, BONDED_ANGLE, BONDED_COSSQ, BONDED_TORSION };

static void bonded_call(seppart *ptr, int kind, int type, const double *prm,
			sepsys *sys, sepret *ret, int values)
{
	seph_ctx *c = seph_ctx_for(ptr, sys);
	int rc = 0;
	if (!c) {
		rc = sep_gpu_last_error();
		goto out;
	}
	if ((rc = seph_put_topology(c, sys)))
		goto out;
	if (!c->resident) {
		if ((rc = seph_put_x(c)) || (rc = seph_put_f(c)) || (rc = sync_ret_begin(c)))
			goto out;
	}
	const int slot = ret_slot_for(c, ret);
	if (slot < 0) {
		rc = sep_gpu_last_error();
		goto out;
	}
	switch (kind) {
	case BONDED_BOND:
		rc = sepdev_bond(c->dev, type, prm[0], prm[1], sys->length, slot);
		break;
	case BONDED_ANGLE:
		rc = sepdev_angle(c->dev, type, prm[0], prm[1], 0, sys->length, slot);
		break;
	case BONDED_COSSQ:
		rc = sepdev_angle(c->dev, type, prm[0], prm[1], 1, sys->length, slot);
		break;
	default:
		rc = sepdev_torsion(c->dev, type, prm, sys->length, slot);
		break;
	}
	if ((rc = seph_dev_fail(rc)))
		goto out;
	if (c->resident)
		return;
	if ((rc = seph_get_f(c)))
		goto out;
	if ((rc = sync_ret_end(c, ret, 0, kind == BONDED_BOND)))
		goto out;
	/* per-term geometry is written only for terms of this type (the Model
	 * II terms leave it alone, source/sepomp.c:179-328) */
	if (values) {
		sepmolinfo *m = sys->molptr;
		const unsigned nterm = kind == BONDED_BOND ? c->topo_nb
				       : kind == BONDED_TORSION ? c->topo_nd : c->topo_na;
		if (nterm == 0)
			return;
		double *tmp = malloc(sizeof(double) * nterm);
		rc = seph_dev_fail(sepdev_get_bonded_values(
			c->dev, kind == BONDED_BOND ? tmp : NULL,
			kind == BONDED_ANGLE || kind == BONDED_COSSQ ? tmp : NULL,
			kind == BONDED_TORSION ? tmp : NULL));
		if (rc == 0)
			for (unsigned n = 0; n < nterm; n++) {
				if (kind == BONDED_BOND && (int)m->blist[3 * n + 2] == type)
					m->blengths[n] = tmp[n];
				else if (kind == BONDED_TORSION && (int)m->dlist[5 * n + 4] == type)
					m->dihedrals[n] = tmp[n];
				else if ((kind == BONDED_ANGLE || kind == BONDED_COSSQ) &&
					 (int)m->alist[4 * n + 3] == type)
					m->angles[n] = tmp[n];
			}
		free(tmp);
	}
out:
	if (rc)
		sep_warning("bonded forces: %s", (char *)sep_gpu_last_error_message());
}

void sep_stretch_harmonic(sepatom *aptr, int type, const double lbond,
			  const double ks, sepsys *sys, sepret *ret)
{
	const double p[2] = {lbond, ks};
	bonded_call(aptr, BONDED_BOND, type, p, sys, ret, 1);
}

void sep_angle_harmonic(sepatom *ptr, int type, const double angle0,
			const double k, sepsys *sys, sepret *ret)
{
	const double p[2] = {angle0, k};
	bonded_call(ptr, BONDED_ANGLE, type, p, sys, ret, 1);
}

void sep_angle_cossq(sepatom *ptr, int type, const double angle0,
		     const double k, sepsys *sys, sepret *ret)
{
	const double p[2] = {angle0, k};
	bonded_call(ptr, BONDED_COSSQ, type, p, sys, ret, 1);
}

void sep_torsion_Ryckaert(sepatom *ptr, int type, const double g[6],
			  sepsys *sys, sepret *ret)
{
	bonded_call(ptr, BONDED_TORSION, type, g, sys, ret, 1);
}

/* ---- Nose-Hoover: source/sepintgr.c:149-168 --------------------------- */

void sep_nosehoover(seppart *ptr, double Td, double *alpha, const double Q,
		    sepsys *sys)
{
	seph_ctx *c = seph_ctx_for(ptr, sys);
	int rc = 0;
	if (!c) {
		rc = sep_gpu_last_error();
		goto out;
	}
	int slot;
	if (c->resident) {
		slot = seph_alpha_slot(c, alpha);
		if (slot < 0) {
			rc = sep_gpu_last_error();
			goto out;
		}
	} else {
		slot = c->alpha_scratch;
		if ((rc = seph_put_v(c)) || (rc = seph_put_f(c)) || (rc = seph_put_scalars(c)) ||
		    (rc = seph_dev_fail(sepdev_scalar_put(c->dev, slot, *alpha))))
			goto out;
	}
	if ((rc = seph_dev_fail(sepdev_nosehoover(c->dev, Td, Q, sys->dt, slot))))
		goto out;
	if (!c->resident) {
		if ((rc = seph_get_f(c)))
			goto out;
		rc = seph_dev_fail(sepdev_scalar_get(c->dev, slot, alpha));
	}
out:
	if (rc)
		sep_warning("sep_nosehoover: %s", (char *)sep_gpu_last_error_message());
}

/* ---- leapfrog: source/sepintgr.c:18-88 -------------------------------- */

void sep_leapfrog(seppart *ptr, sepsys *sys, sepret *retval)
{
	seph_ctx *c = seph_ctx_for(ptr, sys);
	int rc = 0;
	if (!c) {
		rc = sep_gpu_last_error();
		goto out;
	}
	if (c->resident) {
		const int slot = seph_ret_slot(c, retval);
		if (slot < 0) {
			rc = sep_gpu_last_error();
			goto out;
		}
		rc = seph_dev_fail(sepdev_leapfrog(c->dev, sys->dt, sys->skin, sys->length, slot));
		goto out;
	}
	{
		sepdev_sys ds;
		memset(&ds, 0, sizeof ds);
		ds.max_dist2 = sys->max_dist2;
		ds.tnow = sys->tnow;
		ds.nupdate_neighb = sys->nupdate_neighb;
		ds.neighb_flag = 0;
		if ((rc = seph_put_x(c)) || (rc = seph_put_v(c)) || (rc = seph_put_f(c)) ||
		    (rc = seph_put_xn_cross(c)) || (rc = seph_put_scalars(c)) ||
		    (rc = seph_dev_fail(sepdev_put_sys(c->dev, &ds))) || (rc = sync_ret_begin(c)))
			goto out;
		if ((rc = seph_dev_fail(sepdev_leapfrog(c->dev, sys->dt, sys->skin, sys->length,
							 c->ret_scratch))))
			goto out;
		if ((rc = seph_get_state(c)) || (rc = seph_dev_fail(sepdev_get_sys(c->dev, &ds))))
			goto out;
		if ((rc = sync_ret_end(c, retval, 0, 0)))
			goto out;
		sys->max_dist2 = ds.max_dist2;
		if (ds.neighb_flag) {
			sys->neighb_flag = 1;
			sys->nupdate_neighb++;
			for (long i = 0; i < sys->npart; i++)
				for (int k = 0; k < 3; k++) {
					ptr[i].xn[k] = ptr[i].x[k];
					ptr[i].cross_neighb[k] = 0;
				}
		}
		sys->tnow += sys->dt;
	}
out:
	if (rc)
		sep_warning("sep_leapfrog: %s", (char *)sep_gpu_last_error_message());
}

/* ---- multi-type LJ table and its force-matching adjoint (extension) ---- */

int sep_gpu_force_lj_table(seppart *ptr, int ntypes, const char *types,
			   const double *params, sepsys *sys, sepret *retval,
			   const unsigned opt)
{
	double cmax = 0.0;
	for (int i = 0; i < ntypes * ntypes; i++)
		if (params[4 * i] > cmax)
			cmax = params[4 * i];
	if (cmax > sys->cf)
		return seph_fail(SEP_GPU_EARG, "a table cutoff exceeds the maximum cutoff %g", sys->cf);
	sepdev_ljtab_args ta;
	memset(&ta, 0, sizeof ta);
	ta.ntypes = ntypes;
	ta.types = types;
	ta.params = params;
	pair_call pc;
	memset(&pc, 0, sizeof pc);
	pc.types = "\0\0";
	pc.cf = cmax;
	pc.opt = opt;
	pc.table = &ta;
	return run_pairs(ptr, sys, retval, &pc);
}

int sep_gpu_lj_adjoint(seppart *ptr, sepsys *sys, int ntypes, const char *types,
		       const double *params, const double *dLdF, double *dLdx, double *dLdtheta)
{
	seph_ctx *c = seph_ctx_for(ptr, sys);
	if (!c)
		return sep_gpu_last_error();
	int rc;
	if (!c->resident && ((rc = seph_put_x(c)) || (rc = seph_put_scalars(c))))
		return rc;
	sepdev_ljtab_args ta;
	memset(&ta, 0, sizeof ta);
	ta.ntypes = ntypes;
	ta.types = types;
	ta.params = params;
	for (int k = 0; k < 3; k++)
		ta.L[k] = sys->length[k];
	return seph_dev_fail(sepdev_lj_adjoint(c->dev, &ta, dLdF, dLdx, dLdtheta));
}

/* The same adjoint without any allocation: the device context must exist
 * (a force call made it), and the device buffers are the library's fixed
 * scratch array instead of memory allocated on the first call. */
int sep_gpu_lj_adjoint_noalloc(seppart *ptr, sepsys *sys, int ntypes, const char *types,
			       const double *params, const double *dLdF, double *dLdx,
			       double *dLdtheta)
{
	seph_ctx *c = seph_find(ptr);
	if (!c || !c->dev)
		return seph_fail(SEP_GPU_ESTATE,
				 "sep_gpu_lj_adjoint_noalloc: no device state for these atoms "
				 "(a force call creates it)");
	if (sys->npart != c->n)
		return seph_fail(SEP_GPU_EARG, "sys->npart (%ld) differs from the %ld atoms",
				 sys->npart, c->n);
	int rc;
	if (!c->resident && ((rc = seph_put_x(c)) || (rc = seph_put_scalars(c))))
		return rc;
	sepdev_ljtab_args ta;
	memset(&ta, 0, sizeof ta);
	ta.ntypes = ntypes;
	ta.types = types;
	ta.params = params;
	for (int k = 0; k < 3; k++)
		ta.L[k] = sys->length[k];
	return seph_dev_fail(sepdev_lj_adjoint_noalloc(c->dev, &ta, dLdF, dLdx, dLdtheta));
}

/* ---- OpenMP Model II: source/sepomp.c:16-328 ---------------------------- */

/* A Model II term runs through the synchronised path of its sep_*
 * counterpart with the atoms' forces parked: the path then leaves exactly
 * the term's forces in f, which go to ftot before the parked forces come
 * back.  Energies land in a scratch sepret and are dropped, as the
 * reference computes none. */
typedef struct {
	seppart *ptr;
	long n;
	double *saved;
	int err_before;
} omp_term;

static int omp_begin(omp_term *t, seppart *ptr, sepsys *sys, const char *who)
{
	seph_ctx *c = seph_ctx_for(ptr, sys);
	if (!c)
		return sep_gpu_last_error();
	if (c->resident) {
		sep_warning("%s: a resident system runs concurrent terms through sep_gpu_run", who);
		return seph_fail(SEP_GPU_ESTATE, "%s on a resident system", who);
	}
	t->ptr = ptr;
	t->n = sys->npart;
	t->err_before = sep_gpu_last_error();
	t->saved = malloc(sizeof(double) * 3 * (size_t)t->n);
	if (!t->saved)
		return seph_fail(SEP_GPU_EHOST, "%s: out of host memory", who);
	for (long i = 0; i < t->n; i++)
		for (int k = 0; k < 3; k++) {
			t->saved[3 * i + k] = ptr[i].f[k];
			ptr[i].f[k] = 0.0;
		}
	return 0;
}

/* rc: the term's own status; a void sep_* call reports through the error
 * state instead */
static void omp_end(omp_term *t, double **ftot, int rc)
{
	const int ok = rc == 0 && sep_gpu_last_error() == t->err_before;
	for (long i = 0; i < t->n; i++)
		for (int k = 0; k < 3; k++) {
			if (ok)
				ftot[i][k] += t->ptr[i].f[k];
			t->ptr[i].f[k] = t->saved[3 * i + k];
		}
	free(t->saved);
}

void sep_omp_make_neighb(seppart *ptr, double cf, sepsys *sys, const unsigned opt)
{
	if (cf > sys->cf) {
		sep_warning("cutoff for an interaction cannot be larger than maximum cutoff");
		seph_fail(SEP_GPU_EARG, "cutoff %f above the system cutoff %f", cf, sys->cf);
		return;
	}
	if (sys->neighb_flag != 1)
		return;
	seph_ctx *c = seph_ctx_for(ptr, sys);
	int rc = c ? 0 : sep_gpu_last_error();
	if (!rc && c->resident)
		rc = seph_fail(SEP_GPU_ESTATE, "sep_omp_make_neighb on a resident system");
	if (!rc)
		rc = seph_put_x(c);
	if (!rc)
		rc = seph_put_scalars(c);
	if (!rc && needs_partners(opt))
		rc = seph_put_partners(c);
	if (!rc) {
		sepdev_grid g;
		if (sys->neighb_update == SEP_NEIGHBLIST) {
			const double rl = sys->cf + sys->skin;
			seph_grid_own(sys, rl, opt, 0, &g);
			g.rlist2 = rl * rl;
		} else {
			seph_grid_reference(sys, opt, &g);
		}
		rc = seph_build(c, sys, &g, 0);
	}
	if (rc) {
		sep_warning("sep_omp_make_neighb: %s", (char *)sep_gpu_last_error_message());
		return;
	}
	sys->neighb_flag = 0;
}

/* The pair terms use the list as it stands, as the reference's do: a set
 * rebuild flag is held off for the call. */
static void omp_pairs(double **ftot, seppart *ptr, sepsys *sys, pair_call *pc,
		      const char *who)
{
	omp_term t;
	if (omp_begin(&t, ptr, sys, who))
		return;
	sepret r;
	sep_reset_retval(&r);
	const unsigned flag = sys->neighb_flag;
	sys->neighb_flag = 0;
	const int rc = run_pairs(ptr, sys, &r, pc);
	sys->neighb_flag = flag;
	if (rc)
		sep_warning("%s: %s", who, (char *)sep_gpu_last_error_message());
	omp_end(&t, ftot, rc);
}

void sep_omp_pairs(double **ftot, const seppart *ptr, const char *types, double cf,
		   double (*fun)(double, char), const sepsys *sys)
{
	const int pot = seph_identify_potential(fun);
	if (pot == SEPH_POT_NONE) {
		sep_warning("sep_omp_pairs: the device path evaluates sep_lj, sep_lj_shift and sep_wca only");
		seph_fail(SEP_GPU_EUNSUPPORTED, "potential function not available on the device");
		return;
	}
	pair_call pc;
	memset(&pc, 0, sizeof pc);
	pc.types = types;
	pc.cf = cf;
	pc.opt = SEP_ALL;
	pc.lj.eps48 = 48.0;
	pc.lj.eps4 = 4.0;
	pc.lj.awh = 0.5;
	pc.lj.aw = 1.0;
	pc.lj.sig2 = 1.0;
	omp_pairs(ftot, (seppart *)ptr, (sepsys *)sys, &pc, "sep_omp_pairs");
}

void sep_omp_pairs_lj(double **ftot, const seppart *ptr, const char *types,
		      const double *param, const sepsys *sys)
{
	/* source/sepomp.c:112-118: rri = sigma/r^2, ft = 48/eps rri^3 (rri^3 -
	 * 1/2) rri -- the pair kernel's form with sig2 = sigma, eps48 = 48/eps */
	pair_call pc;
	memset(&pc, 0, sizeof pc);
	pc.types = types;
	pc.cf = param[0];
	pc.opt = SEP_ALL;
	pc.lj.eps48 = 48.0 / param[2];
	pc.lj.eps4 = 4.0 / param[2];
	pc.lj.awh = 0.5;
	pc.lj.aw = 1.0;
	pc.lj.sig2 = param[1];
	omp_pairs(ftot, (seppart *)ptr, (sepsys *)sys, &pc, "sep_omp_pairs_lj");
}

void sep_omp_coulomb(double **ftot, seppart *ptr, double cf, sepsys *sys)
{
	omp_term t;
	if (omp_begin(&t, ptr, sys, "sep_omp_coulomb"))
		return;
	sepret r;
	sep_reset_retval(&r);
	sep_coulomb_sf(ptr, cf, sys, &r, SEP_ALL);
	omp_end(&t, ftot, 0);
}

static void omp_bonded(double **ftot, seppart *ptr, int kind, int type, const double *prm,
		       sepsys *sys, const char *who)
{
	omp_term t;
	if (omp_begin(&t, ptr, sys, who))
		return;
	sepret r;
	sep_reset_retval(&r);
	bonded_call(ptr, kind, type, prm, sys, &r, 0);
	omp_end(&t, ftot, 0);
}

void sep_omp_bond(double **ftot, seppart *aptr, int type, const double lbond,
		  const double ks, sepsys *sys)
{
	const double p[2] = {lbond, ks};
	omp_bonded(ftot, aptr, BONDED_BOND, type, p, sys, "sep_omp_bond");
}

void sep_omp_angle(double **ftot, seppart *ptr, int type, const double angle0,
		   const double k, sepsys *sys)
{
	/* source/sepomp.c:217-250: the cosine-squared force */
	const double p[2] = {angle0, k};
	omp_bonded(ftot, ptr, BONDED_COSSQ, type, p, sys, "sep_omp_angle");
}

void sep_omp_torsion(double **ftot, seppart *ptr, int type, const double g[6], sepsys *sys)
{
	omp_bonded(ftot, ptr, BONDED_TORSION, type, g, sys, "sep_omp_torsion");
}
