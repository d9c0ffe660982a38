/*
 * sep_batch.c -- batches of independent systems (include/sepgpu.h): many
 * seppart/sepsys pairs, each set up as for the reference's step loop
 * (prgs/prg1.c:55-65), advanced together on the device, one persistent
 * launch per step; and the autotuner that picks the lanes per atom of each
 * size bucket and the persistent grid.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sephost.h"

struct sep_gpu_batch {
	sepdev_batch *dev;
	int nsys;
	seppart **atoms;
	sepsys **sys;
	int precision;
};

sep_gpu_batch *sep_gpu_batch_create(int nsys, seppart **atoms, sepsys **sys,
				    const char *types, double cf, const double *Td,
				    const double *tau, const double *alpha, int precision,
				    int cap)
{
	if (nsys <= 0 || !atoms || !sys || !types || !Td || !tau) {
		seph_fail(SEP_GPU_EARG, "batch: %d systems or a null argument", nsys);
		return NULL;
	}
	if (precision != SEP_GPU_FP64 && precision != SEP_GPU_MIXED) {
		seph_fail(SEP_GPU_EARG, "batch: precision mode %d", precision);
		return NULL;
	}
	if (types[0] == '\0' || types[1] != types[0]) {
		seph_fail(SEP_GPU_EUNSUPPORTED, "batch: one atom type per system (\"AA\")");
		return NULL;
	}
	sepdev_batch_sys *in = calloc(nsys, sizeof *in);
	int bad = 0;
	for (int s = 0; s < nsys && !bad; s++) {
		const sepsys *y = sys[s];
		const seppart *p = atoms[s];
		if (!y || !p || y->npart <= 0) {
			bad = seph_fail(SEP_GPU_EARG, "batch: system %d is empty", s);
			break;
		}
		if (cf > y->cf) {
			bad = seph_fail(SEP_GPU_EARG, "batch: system %d: cf %g above sys.cf %g", s, cf,
					y->cf);
			break;
		}
		const long n = y->npart;
		double *x = malloc(3 * n * sizeof(double));
		double *v = malloc(3 * n * sizeof(double));
		double *m = malloc(n * sizeof(double));
		int *cr = malloc(3 * n * sizeof(int));
		for (long i = 0; i < n; i++) {
			if (p[i].type != types[0] && !bad)
				bad = seph_fail(SEP_GPU_EUNSUPPORTED,
						"batch: system %d atom %ld has type '%c'", s, i,
						p[i].type);
			for (int k = 0; k < 3; k++) {
				x[3 * i + k] = p[i].x[k];
				v[3 * i + k] = p[i].v[k];
				cr[3 * i + k] = p[i].crossings[k];
			}
			m[i] = p[i].m;
		}
		in[s].n = n;
		for (int k = 0; k < 3; k++)
			in[s].L[k] = y->length[k];
		const double rl = y->cf + y->skin;
		in[s].rlist2 = rl * rl;
		in[s].dt = y->dt;
		in[s].Td = Td[s];
		in[s].tau = tau[s];
		in[s].skin = y->skin;
		in[s].alpha = alpha ? alpha[s] : 0.0;
		in[s].tnow = y->tnow;
		in[s].x = x;
		in[s].v = v;
		in[s].m = m;
		in[s].cr = cr;
	}
	sep_gpu_batch *b = NULL;
	if (!bad) {
		int err = 0;
		sepdev_batch *d = sepdev_batch_create(nsys, in, cf, cap, &err);
		if (!d) {
			seph_dev_fail(err ? err : SEP_GPU_EDEVICE);
		} else {
			b = calloc(1, sizeof *b);
			b->dev = d;
			b->nsys = nsys;
			b->precision = precision;
			b->atoms = malloc(nsys * sizeof(seppart *));
			b->sys = malloc(nsys * sizeof(sepsys *));
			memcpy(b->atoms, atoms, nsys * sizeof(seppart *));
			memcpy(b->sys, sys, nsys * sizeof(sepsys *));
		}
	}
	for (int s = 0; s < nsys; s++) {
		free((void *)in[s].x);
		free((void *)in[s].v);
		free((void *)in[s].m);
		free((void *)in[s].cr);
	}
	free(in);
	return b;
}

void sep_gpu_batch_free(sep_gpu_batch *b)
{
	if (!b)
		return;
	sepdev_batch_destroy(b->dev);
	free(b->atoms);
	free(b->sys);
	free(b);
}

int sep_gpu_batch_run(sep_gpu_batch *b, long nsteps)
{
	if (!b)
		return seph_fail(SEP_GPU_EARG, "batch: null");
	return seph_dev_fail(sepdev_batch_run(b->dev, nsteps, b->precision == SEP_GPU_MIXED, 1));
}

int sep_gpu_batch_run_eager(sep_gpu_batch *b, long nsteps)
{
	if (!b)
		return seph_fail(SEP_GPU_EARG, "batch: null");
	return seph_dev_fail(sepdev_batch_run(b->dev, nsteps, b->precision == SEP_GPU_MIXED, 0));
}

int sep_gpu_batch_eval(sep_gpu_batch *b, int precision)
{
	if (!b)
		return seph_fail(SEP_GPU_EARG, "batch: null");
	if (precision != SEP_GPU_FP64 && precision != SEP_GPU_MIXED)
		return seph_fail(SEP_GPU_EARG, "batch: precision mode %d", precision);
	return seph_dev_fail(sepdev_batch_eval(b->dev, precision == SEP_GPU_MIXED));
}

int sep_gpu_batch_synchronize(sep_gpu_batch *b)
{
	if (!b)
		return seph_fail(SEP_GPU_EARG, "batch: null");
	return seph_dev_fail(sepdev_batch_synchronize(b->dev));
}

int sep_gpu_batch_sync(sep_gpu_batch *b, sepret *rets, double *alpha)
{
	if (!b)
		return seph_fail(SEP_GPU_EARG, "batch: null");
	int rc = 0;
	for (int s = 0; s < b->nsys && !rc; s++) {
		sepsys *y = b->sys[s];
		seppart *p = b->atoms[s];
		const long n = y->npart;
		double *x = malloc(3 * n * sizeof(double));
		double *v = malloc(3 * n * sizeof(double));
		double *f = malloc(3 * n * sizeof(double));
		double *a = malloc(3 * n * sizeof(double));
		int *cr = malloc(3 * n * sizeof(int));
		double w[SEPDEV_RET_WORDS], scal[8];
		rc = seph_dev_fail(sepdev_batch_get(b->dev, s, x, v, f, a, cr, w, scal));
		if (!rc) {
			for (long i = 0; i < n; i++)
				for (int k = 0; k < 3; k++) {
					p[i].x[k] = x[3 * i + k];
					p[i].v[k] = v[3 * i + k];
					p[i].f[k] = f[3 * i + k];
					p[i].a[k] = a[3 * i + k];
					p[i].crossings[k] = cr[3 * i + k];
				}
			y->max_dist2 = scal[1];
			y->tnow = scal[2];
			y->nupdate_neighb = (int)scal[3];
			y->neighb_flag = (int)scal[4];
			if (rets) {
				sepret *r = &rets[s];
				r->epot = w[SEPDEV_RET_EPOT];
				r->ecoul = w[SEPDEV_RET_ECOUL];
				r->ekin = w[SEPDEV_RET_EKIN];
				for (int i = 0; i < 3; i++)
					for (int k = 0; k < 3; k++) {
						r->pot_P[i][k] = w[SEPDEV_RET_POTP + 3 * i + k];
						r->kin_P[i][k] = w[SEPDEV_RET_KINP + 3 * i + k];
					}
			}
			if (alpha)
				alpha[s] = scal[0];
		}
		free(x);
		free(v);
		free(f);
		free(a);
		free(cr);
	}
	return rc;
}

int sep_gpu_batch_set_shape(sep_gpu_batch *b, const int *lanes, int grid)
{
	if (!b)
		return seph_fail(SEP_GPU_EARG, "batch: null");
	return seph_dev_fail(sepdev_batch_set_shape(b->dev, lanes, grid));
}

int sep_gpu_batch_get_shape(sep_gpu_batch *b, int *lanes, int *grid)
{
	if (!b || !lanes || !grid)
		return seph_fail(SEP_GPU_EARG, "batch: null");
	return seph_dev_fail(sepdev_batch_get_shape(b->dev, lanes, grid));
}

int sep_gpu_batch_time_eval(sep_gpu_batch *b, const int *mask, int reps, double *ms)
{
	if (!b || !ms)
		return seph_fail(SEP_GPU_EARG, "batch: null");
	return seph_dev_fail(sepdev_batch_time_eval(b->dev, mask, b->precision == SEP_GPU_MIXED,
						    reps, ms));
}

/* Size buckets of the autotuner: N < 2048, N < 8192, larger. */
static int bucket_of(long n)
{
	return n < 2048 ? 0 : (n < 8192 ? 1 : 2);
}

int sep_gpu_batch_tune(sep_gpu_batch *b, int reps, int *lanes_out, int *grid_out)
{
	static const int lane_choices[] = {1, 2, 4, 8, 16};
	static const int grid_choices[] = {256, 512, 1024, 2048, 4096};
	if (!b)
		return seph_fail(SEP_GPU_EARG, "batch: null");
	if (reps <= 0)
		reps = 20;
	const int ns = b->nsys;
	int *lanes = malloc(ns * sizeof(int));
	int *mask = malloc(ns * sizeof(int));
	int grid = 0, rc = seph_dev_fail(sepdev_batch_get_shape(b->dev, lanes, &grid));
	for (int bk = 0; bk < 3 && !rc; bk++) {
		int any = 0;
		for (int s = 0; s < ns; s++) {
			mask[s] = bucket_of(b->sys[s]->npart) == bk;
			any |= mask[s];
		}
		if (!any)
			continue;
		double best = INFINITY;
		int pick = lanes[0];
		for (size_t c = 0; c < sizeof lane_choices / sizeof lane_choices[0] && !rc; c++) {
			for (int s = 0; s < ns; s++)
				if (mask[s])
					lanes[s] = lane_choices[c];
			double ms = 0.0;
			rc = seph_dev_fail(sepdev_batch_set_shape(b->dev, lanes, 0));
			if (!rc)
				rc = seph_dev_fail(sepdev_batch_time_eval(
					b->dev, mask, b->precision == SEP_GPU_MIXED, reps, &ms));
			if (!rc && ms < best) {
				best = ms;
				pick = lane_choices[c];
			}
		}
		for (int s = 0; s < ns; s++)
			if (mask[s])
				lanes[s] = pick;
	}
	if (!rc)
		rc = seph_dev_fail(sepdev_batch_set_shape(b->dev, lanes, 0));
	double best = INFINITY;
	int gpick = grid;
	for (size_t c = 0; c < sizeof grid_choices / sizeof grid_choices[0] && !rc; c++) {
		double ms = 0.0;
		rc = seph_dev_fail(sepdev_batch_set_shape(b->dev, NULL, grid_choices[c]));
		if (!rc)
			rc = seph_dev_fail(sepdev_batch_time_eval(
				b->dev, NULL, b->precision == SEP_GPU_MIXED, reps, &ms));
		if (!rc && ms < best) {
			best = ms;
			gpick = grid_choices[c];
		}
	}
	if (!rc)
		rc = seph_dev_fail(sepdev_batch_set_shape(b->dev, lanes, gpick));
	if (!rc && lanes_out)
		memcpy(lanes_out, lanes, ns * sizeof(int));
	if (!rc && grid_out)
		*grid_out = gpick;
	free(lanes);
	free(mask);
	return rc;
}
