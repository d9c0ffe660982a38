// sepdev_bonded.hip -- harmonic bonds, harmonic / cosine-squared angles and
// Ryckaert-Bellemans torsions.
//
// One thread per bonded term writes the force on each of its atoms into a
// per-term slot (zero for terms of another type); a gather kernel then adds,
// for every atom, its slots in term order.  Each atom therefore receives
// its bonded forces in the same order as the reference's sequential loop
// adds them, and no atomics are involved.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sepdev_impl.h"

#define SEPK_PI 3.14159265358979 /* the reference's SEP_PI */

// ---- topology upload -------------------------------------------------------

// CSR from atom to the slots (term*width + role) that act on it, in term
// order; atoms are in original numbering.
static void build_csr(const unsigned *list, int nterm, int stride, int width,
		      long natom, int **off_out, int **idx_out)
{
	int *off = (int *)calloc((size_t)natom + 1, sizeof(int));
	for (int n = 0; n < nterm; n++)
		for (int r = 0; r < width; r++)
			off[list[(size_t)n * stride + r] + 1]++;
	for (long o = 0; o < natom; o++)
		off[o + 1] += off[o];
	int *fill = (int *)calloc((size_t)natom, sizeof(int));
	int *idx = (int *)malloc(sizeof(int) * (size_t)(nterm * width > 0 ? nterm * width : 1));
	for (int n = 0; n < nterm; n++)
		for (int r = 0; r < width; r++) {
			const unsigned o = list[(size_t)n * stride + r];
			idx[off[o] + fill[o]++] = n * width + r;
		}
	free(fill);
	*off_out = off;
	*idx_out = idx;
}

static int upload(void **dptr, const void *src, size_t bytes, hipStream_t st)
{
	if (*dptr)
		(void)hipFree(*dptr);
	*dptr = nullptr;
	SEPK_CHECK(hipMalloc(dptr, bytes ? bytes : 8));
	if (bytes)
		SEPK_CHECK(hipMemcpyAsync(*dptr, src, bytes, hipMemcpyHostToDevice, st));
	return 0;
}

static int upload_family(sepdev *d, const unsigned *list, int nterm, int stride,
			 int width, unsigned **dlist, double **contrib, int **doff,
			 int **didx, double **dval)
{
	for (int n = 0; n < nterm; n++)
		for (int r = 0; r < width; r++)
			if (list[(size_t)n * stride + r] >= (unsigned)d->n)
				return sepk_fail(SEP_GPU_EARG, "topology atom index %u >= %ld",
						 list[(size_t)n * stride + r], d->n);
	int *off, *idx;
	build_csr(list, nterm, stride, width, d->n, &off, &idx);
	int rc = 0;
	rc |= upload((void **)dlist, list, sizeof(unsigned) * (size_t)nterm * stride, d->st);
	rc |= upload((void **)doff, off, sizeof(int) * ((size_t)d->n + 1), d->st);
	rc |= upload((void **)didx, idx, sizeof(int) * (size_t)nterm * width, d->st);
	if (*contrib)
		(void)hipFree(*contrib);
	*contrib = nullptr;
	SEPK_CHECK(hipMalloc(contrib, sizeof(double) * 3 * ((size_t)nterm * width + 1)));
	if (*dval)
		(void)hipFree(*dval);
	*dval = nullptr;
	SEPK_CHECK(hipMalloc(dval, sizeof(double) * ((size_t)nterm + 1)));
	SEPK_CHECK(hipStreamSynchronize(d->st));
	free(off);
	free(idx);
	return rc;
}

extern "C" int sepdev_put_topology(sepdev *d, const unsigned *blist, int nb,
				   const unsigned *alist, int na, const unsigned *dlist,
				   int nd)
{
	if (d->capturing)
		return sepk_fail(SEP_GPU_ESTATE, "topology upload during capture");
	// one partial row per 256 terms must fit a reduction region
	const long maxterms = (d->region_rows - 8) * (long)SEPK_BLOCK;
	if (nb > maxterms || na > maxterms || nd > maxterms)
		return sepk_fail(SEP_GPU_EARG, "too many bonded terms for %ld atoms", d->n);
	int rc = 0;
	rc |= upload_family(d, blist, nb, 3, 2, &d->blist, &d->bcontrib, &d->bcsr_off,
			    &d->bcsr_idx, &d->blen);
	rc |= upload_family(d, alist, na, 4, 3, &d->alist, &d->acontrib, &d->acsr_off,
			    &d->acsr_idx, &d->aval);
	rc |= upload_family(d, dlist, nd, 5, 4, &d->dlist, &d->dcontrib, &d->dcsr_off,
			    &d->dcsr_idx, &d->dval);
	d->nb = nb;
	d->na = na;
	d->nd = nd;
	return rc;
}

// ---- gather ------------------------------------------------------------------

// The owed gathers of sepdev::gpend in order: f (or 0 when the first
// starts from f = 0) plus each family's contributions of the atom, in CSR
// order, so the sums are those of one gather launch per family.
__global__ void k_bonded_gather(sepk_state s, long n, long npad, sepk_gpend p)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n || *s.err)
		return;
	const int o = s.orig[t];
	double f[3];
	for (int k = 0; k < 3; k++)
		f[k] = p.over ? 0.0 : s.f[k * npad + t];
	for (int q = 0; q < p.n; q++)
		for (int e = p.g[q].off[o]; e < p.g[q].off[o + 1]; e++) {
			const long sl = p.g[q].idx[e];
			for (int k = 0; k < 3; k++)
				f[k] += p.g[q].contrib[3 * sl + k];
		}
	for (int k = 0; k < 3; k++)
		s.f[k * npad + t] = f[k];
}

int sepk_flush_gathers(sepdev *d)
{
	if (d->ngpend == 0)
		return 0;
	sepk_gpend p;
	p.n = d->ngpend;
	p.over = d->gpend_over;
	for (int q = 0; q < p.n; q++)
		p.g[q] = d->gpend[q];
	d->ngpend = 0;
	d->gpend_over = 0;
	k_bonded_gather<<<sepk_blocks(d->n), SEPK_BLOCK, 0, d->st>>>(d->s[0], d->n, d->npad, p);
	SEPK_LAUNCH_CHECK();
	return 0;
}

// A family's gather is owed, not launched: the next reader of f applies
// it (the leapfrog inside its own kernel), so a step's bonded families
// cost no launch of their own.
static int gather(sepdev *d, const double *contrib, const int *off, const int *idx)
{
	if (d->ngpend == SEPK_MAXGATHER) {
		int rc = sepk_flush_gathers(d);
		if (rc)
			return rc;
	}
	if (d->ngpend == 0)
		d->gpend_over = d->f_zero_pending != 0;
	d->gpend[d->ngpend].contrib = contrib;
	d->gpend[d->ngpend].off = off;
	d->gpend[d->ngpend].idx = idx;
	d->ngpend++;
	d->f_zero_pending = 0;
	return 0;
}

// A second call of a family in one step rewrites the family's per-term
// slots: a gather still owed from the first goes to f before that.
static int settle(sepdev *d, const double *contrib)
{
	for (int q = 0; q < d->ngpend; q++)
		if (d->gpend[q].contrib == contrib)
			return sepk_flush_gathers(d);
	return 0;
}

static int bonded_finish(sepdev *d, unsigned nb, int ret_slot, int with_stress, int region)
{
	sepk_task f;
	memset(&f, 0, sizeof f);
	f.mode = SEPK_FIN_BONDED;
	f.region = region;
	f.nblk = (int)nb;
	f.nv = 7;
	f.with_stress = with_stress;
	sepk_task_ret(d, f, ret_slot);
	return sepk_queue_task(d, f);
}

__device__ __forceinline__ sepk_pos sepk_at(const sepk_state &s, const int *iperm, unsigned o)
{
	return s.pos[iperm[o]];
}

// ---- harmonic bond (source/sepmol.c:372-414) ---------------------------------

__global__ __launch_bounds__(SEPK_BLOCK) void k_bond(
	sepk_state s, const int *iperm, const unsigned *bl, int nb, int type,
	double lbond, double ks, double L0, double L1, double L2, double *contrib,
	double *blen, double *partial, long rows)
{
	if (*s.err)
		return; // the whole block: a step after a device error changes nothing
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	if (n < nb) {
		double f[3] = {0.0, 0.0, 0.0};
		if ((int)bl[3 * n + 2] == type) {
			const sepk_pos pa = sepk_at(s, iperm, bl[3 * n]);
			const sepk_pos pb = sepk_at(s, iperm, bl[3 * n + 1]);
			const double r[3] = {sepk_wrap(pa.x - pb.x, L0), sepk_wrap(pa.y - pb.y, L1),
					     sepk_wrap(pa.z - pb.z, L2)};
			const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
			const double dist = sqrt(r2);
			const double ft = -ks * (dist - lbond) / dist;
			for (int k = 0; k < 3; k++)
				f[k] = ft * r[k];
			acc[0] = 0.5 * ks * ((dist - lbond) * (dist - lbond));
			acc[1] = f[0] * r[0];
			acc[2] = f[0] * r[1];
			acc[3] = f[0] * r[2];
			acc[4] = f[1] * r[1];
			acc[5] = f[1] * r[2];
			acc[6] = f[2] * r[2];
			blen[n] = dist;
		}
		for (int k = 0; k < 3; k++) {
			contrib[3 * (2 * n) + k] = f[k];
			contrib[3 * (2 * n + 1) + k] = -f[k];
		}
	}
	sepk_block_partial<7>(acc, 0u, partial, blockIdx.x, rows);
}

extern "C" int sepdev_bond(sepdev *d, int type, double lbond, double ks,
			   const double *L, int ret_slot)
{
	if (d->nb == 0)
		return 0;
	int region;
	hipStream_t st;
	int rc = sepk_flush_nh(d);
	if (rc || (rc = settle(d, d->bcontrib)) || (rc = sepk_term_region(d, &region)) ||
	    (rc = sepk_term_stream(d, 1, &st)))
		return rc;
	const unsigned nb = sepk_blocks(d->nb);
	k_bond<<<nb, SEPK_BLOCK, 0, st>>>(d->s[0], d->iperm, d->blist, d->nb, type, lbond, ks,
					   L[0], L[1], L[2], d->bcontrib, d->blen,
					   sepk_region(d, region), d->region_rows);
	SEPK_LAUNCH_CHECK();
	if ((rc = sepk_term_join(d, st)) || (rc = bonded_finish(d, nb, ret_slot, 1, region)))
		return rc;
	return gather(d, d->bcontrib, d->bcsr_off, d->bcsr_idx);
}

// ---- angles (source/sepmol.c:418-516) ----------------------------------------

__device__ __forceinline__ double sepk_dot3(const double *a, const double *b)
{
	return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__global__ __launch_bounds__(SEPK_BLOCK) void k_angle(
	sepk_state s, const int *iperm, const unsigned *al, int na, int type,
	double angle0, double kth, double ccon, int cossq, double L0, double L1,
	double L2, double *contrib, double *aval, double *partial, long rows)
{
	if (*s.err)
		return; // the whole block: a step after a device error changes nothing
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	if (n < na) {
		double fa[3] = {0.0, 0.0, 0.0}, fb[3] = {0.0, 0.0, 0.0}, fc[3] = {0.0, 0.0, 0.0};
		if ((int)al[4 * n + 3] == type) {
			const sepk_pos pa = sepk_at(s, iperm, al[4 * n]);
			const sepk_pos pb = sepk_at(s, iperm, al[4 * n + 1]);
			const sepk_pos pc = sepk_at(s, iperm, al[4 * n + 2]);
			const double d1[3] = {sepk_wrap(pb.x - pa.x, L0), sepk_wrap(pb.y - pa.y, L1),
					      sepk_wrap(pb.z - pa.z, L2)};
			const double d2[3] = {sepk_wrap(pc.x - pb.x, L0), sepk_wrap(pc.y - pb.y, L1),
					      sepk_wrap(pc.z - pb.z, L2)};
			const double c11 = sepk_dot3(d1, d1), c12 = sepk_dot3(d1, d2),
				     c22 = sepk_dot3(d2, d2);
			const double cD = sqrt(c11 * c22);
			double f;
			if (cossq) {
				const double cc = c12 / cD;
				f = -kth * (cc - ccon);
				acc[0] = 0.5 * kth * ((cc - ccon) * (cc - ccon));
				aval[n] = SEPK_PI - acos(cc);
			} else {
				const double theta = SEPK_PI - acos(c12 / cD);
				f = -kth * (theta - angle0);
				acc[0] = 0.5 * kth * ((theta - angle0) * (theta - angle0));
				aval[n] = theta;
			}
			for (int k = 0; k < 3; k++) {
				const double f1 = f * ((c12 / c11) * d1[k] - d2[k]) / cD;
				const double f2 = f * (d1[k] - (c12 / c22) * d2[k]) / cD;
				fa[k] = f1;
				fb[k] = -f1 - f2;
				fc[k] = f2;
			}
		}
		for (int k = 0; k < 3; k++) {
			contrib[3 * (3 * n) + k] = fa[k];
			contrib[3 * (3 * n + 1) + k] = fb[k];
			contrib[3 * (3 * n + 2) + k] = fc[k];
		}
	}
	sepk_block_partial<7>(acc, 0u, partial, blockIdx.x, rows);
}

extern "C" int sepdev_angle(sepdev *d, int type, double angle0, double k, int cossq,
			    const double *L, int ret_slot)
{
	if (d->na == 0)
		return 0;
	// cos(pi - angle0) on the host with the C library, as the reference
	const double ccon = cos(SEPK_PI - angle0);
	int region;
	hipStream_t st;
	int rc = sepk_flush_nh(d);
	if (rc || (rc = settle(d, d->acontrib)) || (rc = sepk_term_region(d, &region)) ||
	    (rc = sepk_term_stream(d, 2, &st)))
		return rc;
	const unsigned nb = sepk_blocks(d->na);
	k_angle<<<nb, SEPK_BLOCK, 0, st>>>(d->s[0], d->iperm, d->alist, d->na, type, angle0, k,
					    ccon, cossq, L[0], L[1], L[2], d->acontrib, d->aval,
					    sepk_region(d, region), d->region_rows);
	SEPK_LAUNCH_CHECK();
	if ((rc = sepk_term_join(d, st)) || (rc = bonded_finish(d, nb, ret_slot, 0, region)))
		return rc;
	return gather(d, d->acontrib, d->acsr_off, d->acsr_idx);
}

// ---- Ryckaert-Bellemans torsion (source/sepmol.c:520-587) --------------------

struct sepk_rb {
	double g[6];
};

__global__ __launch_bounds__(SEPK_BLOCK) void k_torsion(
	sepk_state s, const int *iperm, const unsigned *dl, int nd, int type,
	sepk_rb g, double L0, double L1, double L2, double *contrib, double *dval,
	double *partial, long rows)
{
	if (*s.err)
		return; // the whole block: a step after a device error changes nothing
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	if (n < nd) {
		double F[4][3] = {{0.0}};
		if ((int)dl[5 * n + 4] == type) {
			const sepk_pos pa = sepk_at(s, iperm, dl[5 * n]);
			const sepk_pos pb = sepk_at(s, iperm, dl[5 * n + 1]);
			const sepk_pos pc = sepk_at(s, iperm, dl[5 * n + 2]);
			const sepk_pos pd = sepk_at(s, iperm, dl[5 * n + 3]);
			const double d1[3] = {sepk_wrap(pb.x - pa.x, L0), sepk_wrap(pb.y - pa.y, L1),
					      sepk_wrap(pb.z - pa.z, L2)};
			const double d2[3] = {sepk_wrap(pc.x - pb.x, L0), sepk_wrap(pc.y - pb.y, L1),
					      sepk_wrap(pc.z - pb.z, L2)};
			const double d3[3] = {sepk_wrap(pd.x - pc.x, L0), sepk_wrap(pd.y - pc.y, L1),
					      sepk_wrap(pd.z - pc.z, L2)};
			const double c11 = sepk_dot3(d1, d1), c12 = sepk_dot3(d1, d2),
				     c13 = sepk_dot3(d1, d3), c22 = sepk_dot3(d2, d2),
				     c23 = sepk_dot3(d2, d3), c33 = sepk_dot3(d3, d3);
			const double cA = c13 * c22 - c12 * c23;
			const double cB1 = c11 * c22 - c12 * c12;
			const double cB2 = c22 * c33 - c23 * c23;
			const double cD = sqrt(cB1 * cB2);
			const double cc = cA / cD;
			const double *gg = g.g;
			const double f = -(gg[1] + (2. * gg[2] + (3. * gg[3] + (4. * gg[4] + 5. * gg[5] * cc) * cc) * cc) * cc);
			const double t1 = cA, t2 = c11 * c23 - c12 * c13, t3 = -cB1;
			const double t4 = cB2, t5 = c13 * c23 - c12 * c33, t6 = -cA;
			const double cR1 = c12 / c22, cR2 = c23 / c22;
			for (int k = 0; k < 3; k++) {
				const double f1 = f * c22 * (t1 * d1[k] + t2 * d2[k] + t3 * d3[k]) / (cD * cB1);
				const double f2 = f * c22 * (t4 * d1[k] + t5 * d2[k] + t6 * d3[k]) / (cD * cB2);
				F[0][k] = f1;
				F[1][k] = -(1.0 + cR1) * f1 + cR2 * f2;
				F[2][k] = cR1 * f1 - (1.0 + cR2) * f2;
				F[3][k] = f2;
			}
			acc[0] = gg[0] + (gg[1] + (gg[2] + (gg[3] + (gg[4] + gg[5] * cc) * cc) * cc) * cc) * cc;
			dval[n] = SEPK_PI - acos(cc);
		}
		for (int r = 0; r < 4; r++)
			for (int k = 0; k < 3; k++)
				contrib[3 * (4 * n + r) + k] = F[r][k];
	}
	sepk_block_partial<7>(acc, 0u, partial, blockIdx.x, rows);
}

extern "C" int sepdev_torsion(sepdev *d, int type, const double *g, const double *L,
			      int ret_slot)
{
	if (d->nd == 0)
		return 0;
	sepk_rb gg;
	for (int i = 0; i < 6; i++)
		gg.g[i] = g[i];
	int region;
	hipStream_t st;
	int rc = sepk_flush_nh(d);
	if (rc || (rc = settle(d, d->dcontrib)) || (rc = sepk_term_region(d, &region)) ||
	    (rc = sepk_term_stream(d, 4, &st)))
		return rc;
	const unsigned nb = sepk_blocks(d->nd);
	k_torsion<<<nb, SEPK_BLOCK, 0, st>>>(d->s[0], d->iperm, d->dlist, d->nd, type, gg, L[0],
					      L[1], L[2], d->dcontrib, d->dval,
					      sepk_region(d, region), d->region_rows);
	SEPK_LAUNCH_CHECK();
	if ((rc = sepk_term_join(d, st)) || (rc = bonded_finish(d, nb, ret_slot, 0, region)))
		return rc;
	return gather(d, d->dcontrib, d->dcsr_off, d->dcsr_idx);
}

extern "C" int sepdev_get_bonded_values(sepdev *d, double *blengths, double *angles,
					double *dihedrals)
{
	int rc = sepk_flush_tasks(d);
	if (rc)
		return rc;
	if (blengths && d->nb)
		SEPK_CHECK(hipMemcpyAsync(blengths, d->blen, sizeof(double) * d->nb, hipMemcpyDeviceToHost, d->st));
	if (angles && d->na)
		SEPK_CHECK(hipMemcpyAsync(angles, d->aval, sizeof(double) * d->na, hipMemcpyDeviceToHost, d->st));
	if (dihedrals && d->nd)
		SEPK_CHECK(hipMemcpyAsync(dihedrals, d->dval, sizeof(double) * d->nd, hipMemcpyDeviceToHost, d->st));
	SEPK_CHECK(hipStreamSynchronize(d->st));
	return 0;
}
