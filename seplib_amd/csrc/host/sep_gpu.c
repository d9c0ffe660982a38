/*
 * sep_gpu.c -- the MI355X extension of the API (include/sepgpu.h):
 * precision, residency, fused multi-step runs, diagnostics.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sephost.h"

int sep_gpu_set_precision(seppart *atoms, sepsys *sys, int mode)
{
	if (mode != SEP_GPU_FP64 && mode != SEP_GPU_MIXED)
		return seph_fail(SEP_GPU_EARG, "precision mode %d", mode);
	seph_ctx *c = seph_ctx_for(atoms, sys);
	if (!c)
		return sep_gpu_last_error();
	c->precision = mode;
	return 0;
}

int sep_gpu_resident(seppart *atoms, sepsys *sys, int on)
{
	seph_ctx *c = seph_ctx_for(atoms, sys);
	if (!c)
		return sep_gpu_last_error();
	if (on && !c->resident) {
		int rc = seph_put_all(c, sys);
		if (rc)
			return rc;
		c->resident = 1;
		return 0;
	}
	if (!on && c->resident) {
		int rc = seph_get_all(c, sys);
		c->resident = 0;
		c->nrets = 0;
		c->nalphas = 0;
		return rc;
	}
	return 0;
}

int sep_gpu_sync(seppart *atoms, sepsys *sys)
{
	seph_ctx *c = seph_ctx_for(atoms, sys);
	if (!c)
		return sep_gpu_last_error();
	if (!c->resident)
		return 0;
	return seph_get_all(c, sys);
}

int sep_gpu_upload(seppart *atoms, sepsys *sys)
{
	seph_ctx *c = seph_ctx_for(atoms, sys);
	if (!c)
		return sep_gpu_last_error();
	if (!c->resident)
		return 0;
	/* bound sepret / alpha values are re-read from the host as well */
	int rc = seph_put_all(c, sys);
	if (rc)
		return rc;
	c->nrets = 0;
	c->nalphas = 0;
	return 0;
}

/* ---- fused run -------------------------------------------------------- */

/* A step of one pair call, an optional sep_nosehoover and sep_leapfrog: on
 * a resident system the device runs it as one fused kernel. */
static int fusable(const sep_gpu_op *ops, int nops)
{
	if (nops < 2 || nops > 3)
		return 0;
	if (ops[0].op != SEP_GPU_OP_PAIRS && ops[0].op != SEP_GPU_OP_LJ)
		return 0;
	if (ops[nops - 1].op != SEP_GPU_OP_LEAPFROG)
		return 0;
	return nops == 2 || ops[1].op == SEP_GPU_OP_NOSEHOOVER;
}

/* A step whose bonded terms can overlap its pair terms: a pair call
 * (which fixes the atom order for the step) before a bonded call. */
static int overlappable(const sep_gpu_op *ops, int nops)
{
	int pair = 0;
	for (int i = 0; i < nops; i++) {
		const int op = ops[i].op;
		if (op == SEP_GPU_OP_PAIRS || op == SEP_GPU_OP_LJ || op == SEP_GPU_OP_COULOMB)
			pair = 1;
		else if (pair && op >= SEP_GPU_OP_BOND && op <= SEP_GPU_OP_TORSION)
			return 1;
	}
	return 0;
}

/* A pair call followed directly by a Coulomb call: one list walk serves
 * both. */
static int pair_then_coulomb(const sep_gpu_op *ops, int nops)
{
	for (int i = 0; i + 1 < nops; i++)
		if ((ops[i].op == SEP_GPU_OP_PAIRS || ops[i].op == SEP_GPU_OP_LJ) &&
		    ops[i + 1].op == SEP_GPU_OP_COULOMB)
			return 1;
	return 0;
}

static int run_ops(seppart *atoms, sepsys *sys, sepret *ret, const sep_gpu_op *ops,
		   int nops);

static int run_one_step(seppart *atoms, sepsys *sys, sepret *ret,
			const sep_gpu_op *ops, int nops, int flags)
{
	seph_ctx *c = seph_ctx_for(atoms, sys);
	if (!c)
		return sep_gpu_last_error();
	const int fuse = c->resident && fusable(ops, nops);
	const int conc = c->resident && (flags & SEP_GPU_RUN_CONCURRENT) && overlappable(ops, nops);
	if (fuse)
		sepdev_fuse_step(c->dev, 1);
	const int pcf = c->resident && pair_then_coulomb(ops, nops);
	int rc = conc ? seph_dev_fail(sepdev_concurrent_terms(c->dev, 1)) : 0;
	if (!rc && pcf)
		rc = seph_dev_fail(sepdev_fuse_pair_coulomb(c->dev, 1));
	if (!rc)
		rc = run_ops(atoms, sys, ret, ops, nops);
	if (pcf) {
		const int rc2 = seph_dev_fail(sepdev_fuse_pair_coulomb(c->dev, 0));
		if (!rc)
			rc = rc2;
	}
	if (fuse) {
		const int rc2 = seph_dev_fail(sepdev_fuse_step(c->dev, 0));
		if (!rc)
			rc = rc2;
	}
	if (conc) {
		const int rc2 = seph_dev_fail(sepdev_concurrent_terms(c->dev, 0));
		if (!rc)
			rc = rc2;
	}
	return rc;
}

static int run_ops(seppart *atoms, sepsys *sys, sepret *ret, const sep_gpu_op *ops,
		   int nops)
{
	sep_reset_retval(ret);
	sep_reset_force(atoms, sys);
	for (int i = 0; i < nops; i++) {
		const sep_gpu_op *o = &ops[i];
		switch (o->op) {
		case SEP_GPU_OP_PAIRS:
			sep_force_pairs(atoms, o->types, o->p[0], o->fun, sys, ret, o->opt);
			break;
		case SEP_GPU_OP_LJ:
			sep_force_lj(atoms, o->types, o->p, sys, ret, o->opt);
			break;
		case SEP_GPU_OP_COULOMB:
			sep_coulomb_sf(atoms, o->p[0], sys, ret, o->opt);
			break;
		case SEP_GPU_OP_BOND:
			sep_stretch_harmonic(atoms, o->type, o->p[0], o->p[1], sys, ret);
			break;
		case SEP_GPU_OP_ANGLE:
			sep_angle_harmonic(atoms, o->type, o->p[0], o->p[1], sys, ret);
			break;
		case SEP_GPU_OP_COSSQ:
			sep_angle_cossq(atoms, o->type, o->p[0], o->p[1], sys, ret);
			break;
		case SEP_GPU_OP_TORSION:
			sep_torsion_Ryckaert(atoms, o->type, o->p, sys, ret);
			break;
		case SEP_GPU_OP_NOSEHOOVER:
			sep_nosehoover(atoms, o->p[0], o->alpha, o->p[1], sys);
			break;
		case SEP_GPU_OP_LEAPFROG:
			sep_leapfrog(atoms, sys, ret);
			break;
		default:
			return seph_fail(SEP_GPU_EARG, "unknown op %d", o->op);
		}
	}
	return sep_gpu_last_error();
}

/* The device error word after a run: a failed step and the ones after it
 * left the state alone; report the failure, growing an overflowed list
 * (within any bound) so that the next run can continue. */
static int run_check(seph_ctx *c)
{
	sepdev_sys ds;
	int rc = seph_dev_fail(sepdev_get_sys(c->dev, &ds));
	if (rc || !ds.error)
		return rc;
	if (ds.error == SEP_GPU_ENEIGHB) {
		const int max_full = ds.max_full;
		if ((rc = seph_grow_list(c, &ds)))
			return rc;
		return seph_fail(SEP_GPU_ENEIGHB,
				 "neighbour list overflow in a resident run (longest list %d): "
				 "the system stays at the step before it; the list was grown",
				 max_full);
	}
	return seph_check_device_error(c);
}

#define SEPH_GRAPH_STEPS 16 /* steps a replayed graph holds in long runs (even) */

static int run_steps(seppart *atoms, sepsys *sys, sepret *ret, const sep_gpu_op *ops,
		     int nops, long nsteps, int flags);

int sep_gpu_run(seppart *atoms, sepsys *sys, sepret *ret, const sep_gpu_op *ops,
		int nops, long nsteps, int flags)
{
	seph_ctx *c = seph_ctx_for(atoms, sys);
	if (!c)
		return sep_gpu_last_error();
	if (!c->resident)
		return seph_fail(SEP_GPU_ESTATE, "sep_gpu_run needs a resident system");
	if (nsteps <= 0)
		return 0;
	sep_gpu_clear_error();
	int rc = run_steps(atoms, sys, ret, ops, nops, nsteps, flags);
	const int rc2 = run_check(c);
	return rc ? rc : rc2;
}

int sep_gpu_list_limit(seppart *atoms, sepsys *sys, int entries)
{
	if (entries < 0)
		return seph_fail(SEP_GPU_EARG, "list limit %d", entries);
	seph_ctx *c = seph_ctx_for(atoms, sys);
	if (!c)
		return sep_gpu_last_error();
	int applied = 0;
	int rc = seph_dev_fail(sepdev_list_limit(c->dev, entries, &applied));
	if (!rc)
		c->list_limit = applied;
	return rc;
}

static int run_steps(seppart *atoms, sepsys *sys, sepret *ret, const sep_gpu_op *ops,
		     int nops, long nsteps, int flags)
{
	seph_ctx *c = seph_ctx_for(atoms, sys);
	/* the first step runs eagerly: it binds slots and sizes every buffer */
	int rc = run_one_step(atoms, sys, ret, ops, nops, flags);
	if (rc || nsteps == 1)
		return rc;
	if (!(flags & SEP_GPU_RUN_GRAPH)) {
		for (long s = 1; s < nsteps; s++)
			if ((rc = run_one_step(atoms, sys, ret, ops, nops, flags)))
				return rc;
		return 0;
	}
	/* An even number of steps per graph: a fused step writes the new
	 * positions into the other of two buffers, so consecutive steps
	 * alternate between them and a replayed graph returns to the captured
	 * assignment.  Each replay costs a few microseconds on the device
	 * before its first kernel, so runs of SEPH_GRAPH_STEPS steps or more
	 * replay that many a launch (slot 0) and the pairs left over two a
	 * launch (slot 1); both are captured before either is replayed. */
	const long rest = nsteps - 1;
	const long big = rest >= SEPH_GRAPH_STEPS ? rest / SEPH_GRAPH_STEPS : 0;
	const long pairs = (rest - big * SEPH_GRAPH_STEPS) / 2;
	const long sizes[2] = {big ? SEPH_GRAPH_STEPS : 2, 2};
	const long times[2] = {big ? big : pairs, big ? pairs : 0};
	for (int g = 0; g < 2; g++) {
		if (!times[g])
			continue;
		if ((rc = seph_dev_fail(sepdev_graph_begin(c->dev, g))))
			return rc;
		for (long s = 0; !rc && s < sizes[g]; s++)
			rc = run_one_step(atoms, sys, ret, ops, nops, flags);
		int rc2 = seph_dev_fail(sepdev_graph_end(c->dev, g));
		if (rc || rc2)
			return rc ? rc : rc2;
	}
	/* the captured steps were not executed */
	for (int g = 0; g < 2; g++)
		if (times[g] && (rc = seph_dev_fail(sepdev_graph_launch(c->dev, g, times[g]))))
			return rc;
	if (rest % 2)
		rc = run_one_step(atoms, sys, ret, ops, nops, flags);
	return rc;
}

/* ---- device checkpoint ------------------------------------------------ */

#define SEPH_CKPT_TAG "SEPGPUH1"

int sep_gpu_checkpoint(seppart *atoms, sepsys *sys, const sepret *ret, const double *alpha,
		       const char *file)
{
	seph_ctx *c = seph_ctx_for(atoms, sys);
	if (!c)
		return sep_gpu_last_error();
	if (!c->resident)
		return seph_fail(SEP_GPU_ESTATE, "sep_gpu_checkpoint needs a resident system");
	/* the values the run left in its sepret and thermostat variable */
	sepret r;
	memset(&r, 0, sizeof r);
	double al = 0.0;
	int rc = 0;
	for (int b = 0; !rc && b < c->nrets; b++)
		if (c->rets[b].host == ret) {
			double w[SEPDEV_RET_WORDS];
			rc = seph_dev_fail(sepdev_ret_get(c->dev, c->rets[b].slot, w));
			seph_ret_unpack(&r, w);
		}
	if (!rc && ret && c->nrets == 0)
		r = *ret;
	int have_alpha = 0;
	for (int b = 0; !rc && b < c->nalphas; b++)
		if (c->alphas[b].host == alpha) {
			rc = seph_dev_fail(sepdev_scalar_get(c->dev, c->alphas[b].slot, &al));
			have_alpha = 1;
		}
	if (!rc && alpha && !have_alpha)
		al = *alpha;
	if (rc)
		return rc;
	FILE *f = fopen(file, "wb");
	if (!f)
		return seph_fail(SEP_GPU_EARG, "sep_gpu_checkpoint: cannot open %s", file);
	int has[2] = {ret != NULL, alpha != NULL};
	if (fwrite(SEPH_CKPT_TAG, 1, 8, f) != 8 || fwrite(has, sizeof has, 1, f) != 1 ||
	    fwrite(&r, sizeof r, 1, f) != 1 || fwrite(&al, sizeof al, 1, f) != 1)
		rc = seph_fail(SEP_GPU_EHOST, "sep_gpu_checkpoint: write failed");
	if (!rc)
		rc = seph_dev_fail(sepdev_checkpoint_write(c->dev, f));
	if (fclose(f) != 0 && !rc)
		rc = seph_fail(SEP_GPU_EHOST, "sep_gpu_checkpoint: write failed");
	return rc;
}

int sep_gpu_restore(seppart *atoms, sepsys *sys, sepret *ret, double *alpha, const char *file)
{
	seph_ctx *c = seph_ctx_for(atoms, sys);
	if (!c)
		return sep_gpu_last_error();
	if (!c->resident)
		return seph_fail(SEP_GPU_ESTATE, "sep_gpu_restore needs a resident system");
	FILE *f = fopen(file, "rb");
	if (!f)
		return seph_fail(SEP_GPU_EARG, "sep_gpu_restore: cannot open %s", file);
	char tag[8];
	int has[2];
	sepret r;
	double al;
	int rc = 0;
	if (fread(tag, 1, 8, f) != 8 || memcmp(tag, SEPH_CKPT_TAG, 8) != 0 ||
	    fread(has, sizeof has, 1, f) != 1 || fread(&r, sizeof r, 1, f) != 1 ||
	    fread(&al, sizeof al, 1, f) != 1)
		rc = seph_fail(SEP_GPU_EARG, "sep_gpu_restore: %s is not a device checkpoint", file);
	if (!rc)
		rc = seph_dev_fail(sepdev_checkpoint_read(c->dev, f));
	fclose(f);
	if (rc)
		return rc;
	sepdev_sys ds;
	if ((rc = seph_dev_fail(sepdev_get_sys(c->dev, &ds))))
		return rc;
	sys->tnow = ds.tnow;
	sys->neighb_flag = ds.neighb_flag;
	sys->nupdate_neighb = ds.nupdate_neighb;
	sys->max_dist2 = ds.max_dist2;
	if (ret && has[0])
		*ret = r;
	if (alpha && has[1])
		*alpha = al;
	/* bound values are taken from the host again at their next use */
	c->nrets = 0;
	c->nalphas = 0;
	return 0;
}

/* ---- diagnostics ------------------------------------------------------ */

int sep_gpu_neighbour_pairs(seppart *atoms, sepsys *sys, int *pairs, long cap,
			    long *npairs)
{
	seph_ctx *c = seph_ctx_for(atoms, sys);
	if (!c)
		return sep_gpu_last_error();
	const int stride = sepdev_list_capacity(c->dev);
	if (stride <= 0)
		return seph_fail(SEP_GPU_ESTATE, "no neighbour list on the device");
	int *counts = malloc(sizeof(int) * (size_t)c->n);
	int *half = malloc(sizeof(int) * (size_t)c->n * (size_t)stride);
	int maxc = 0;
	int rc = seph_dev_fail(sepdev_get_half_list(c->dev, counts, half, stride, &maxc));
	if (rc == 0) {
		long tot = 0;
		for (long i = 0; i < c->n; i++)
			tot += counts[i];
		*npairs = tot;
		if (pairs && tot <= cap) {
			long w = 0;
			for (long i = 0; i < c->n; i++)
				for (int e = 0; e < counts[i]; e++) {
					pairs[2 * w] = (int)i;
					pairs[2 * w + 1] = half[(size_t)i * stride + e];
					w++;
				}
		}
	}
	free(counts);
	free(half);
	return rc;
}

int sep_gpu_timing_enable(seppart *atoms, int on)
{
	seph_ctx *c = seph_find(atoms);
	if (!c)
		return seph_fail(SEP_GPU_ESTATE, "no device context for these atoms");
	return seph_dev_fail(sepdev_timing(c->dev, on));
}

int sep_gpu_timing_reset(seppart *atoms)
{
	seph_ctx *c = seph_find(atoms);
	if (!c)
		return seph_fail(SEP_GPU_ESTATE, "no device context for these atoms");
	return seph_dev_fail(sepdev_timing_reset(c->dev));
}

int sep_gpu_timing_get(seppart *atoms, const char *kernel, double *total_ms, long *count)
{
	seph_ctx *c = seph_find(atoms);
	if (!c)
		return seph_fail(SEP_GPU_ESTATE, "no device context for these atoms");
	return seph_dev_fail(sepdev_timing_get(c->dev, kernel, total_ms, count));
}

int sep_gpu_time_pair_kernel(seppart *atoms, int reps, double *ms_per_launch)
{
	seph_ctx *c = seph_find(atoms);
	if (!c)
		return seph_fail(SEP_GPU_ESTATE, "no device context for these atoms");
	return seph_dev_fail(sepdev_time_pair(c->dev, reps, ms_per_launch));
}

int sep_gpu_list_stats(seppart *atoms, double *mean_full, int *max_full)
{
	seph_ctx *c = seph_find(atoms);
	if (!c)
		return seph_fail(SEP_GPU_ESTATE, "no device context for these atoms");
	return seph_dev_fail(sepdev_list_stats(c->dev, mean_full, max_full));
}

/* ---- per-step record ---------------------------------------------------- */

int sep_gpu_record_steps(seppart *atoms, long capacity)
{
	seph_ctx *c = seph_find(atoms);
	if (!c)
		return seph_fail(SEP_GPU_ESTATE, "no device context for these atoms");
	return seph_dev_fail(sepdev_ring_alloc(c->dev, capacity));
}

int sep_gpu_read_steps(seppart *atoms, sepret *rets, double *tnow, long max, long *n,
		       long *lost)
{
	if (!n || !lost)
		return seph_fail(SEP_GPU_EARG, "read_steps: null count");
	*n = 0;
	*lost = 0;
	seph_ctx *c = seph_find(atoms);
	if (!c)
		return seph_fail(SEP_GPU_ESTATE, "no device context for these atoms");
	if (max <= 0)
		return seph_fail(SEP_GPU_EARG, "read_steps: max %ld", max);
	double *w = malloc((size_t)max * SEPDEV_RING_WORDS * sizeof(double));
	if (!w)
		return seph_fail(SEP_GPU_EHOST, "read_steps: out of host memory");
	int rc = seph_dev_fail(sepdev_ring_read(c->dev, w, max, n, lost));
	for (long k = 0; !rc && k < *n; k++) {
		const double *rec = w + (size_t)k * SEPDEV_RING_WORDS;
		if (rets) {
			memset(&rets[k], 0, sizeof(sepret));
			seph_ret_unpack(&rets[k], rec);
		}
		if (tnow)
			tnow[k] = rec[SEPDEV_RET_WORDS];
	}
	free(w);
	return rc;
}

int sep_gpu_synchronize(seppart *atoms)
{
	seph_ctx *c = seph_find(atoms);
	if (!c)
		return seph_fail(SEP_GPU_ESTATE, "no device context for these atoms");
	return seph_dev_fail(sepdev_synchronize(c->dev));
}
