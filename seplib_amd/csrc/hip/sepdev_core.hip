// sepdev_core.hip -- device state lifetime, host<->device transfers in the
// caller's atom order, return/scalar slots, timing and graph capture.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sepdev_impl.h"

// ---- errors ---------------------------------------------------------------

static char g_msg[512];
static int g_code;

int sepk_fail(int code, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_msg, sizeof g_msg, fmt, ap);
	va_end(ap);
	g_code = code;
	return code;
}

extern "C" const char *sepdev_error_message(void) { return g_msg; }
extern "C" int sepdev_last_code(void) { return g_code; }
extern "C" void sepdev_clear_error(void)
{
	g_code = 0;
	g_msg[0] = '\0';
}

// ---- allocation -------------------------------------------------------------

template <class T>
static int dalloc(T **p, size_t count)
{
	*p = nullptr;
	if (count == 0)
		count = 1;
	SEPK_CHECK(hipMalloc((void **)p, count * sizeof(T)));
	return 0;
}

static int alloc_state(sepk_state *s, long npad)
{
	memset(s, 0, sizeof *s);
	int rc = 0;
	rc |= dalloc(&s->pos, npad);
	rc |= dalloc(&s->qpos, npad);
	rc |= dalloc(&s->v, 3 * npad);
	rc |= dalloc(&s->f, 3 * npad);
	rc |= dalloc(&s->a, 3 * npad);
	rc |= dalloc(&s->xn, 3 * npad);
	rc |= dalloc(&s->cn, 3 * npad);
	rc |= dalloc(&s->cr, 3 * npad);
	rc |= dalloc(&s->m, npad);
	rc |= dalloc(&s->type, npad);
	rc |= dalloc(&s->mol, npad);
	rc |= dalloc(&s->orig, npad);
	return rc;
}

static void free_state(sepk_state *s)
{
	void *ptrs[] = {s->pos, s->qpos, s->v, s->f, s->a, s->xn, s->cn, s->cr,
			s->m, s->type, s->mol, s->orig, s->partners};
	for (void *p : ptrs)
		if (p)
			(void)hipFree(p);
	memset(s, 0, sizeof *s);
}

__global__ void k_init_identity(sepk_state s, int *iperm, long n, long npad)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= npad)
		return;
	s.orig[t] = t < n ? (int)t : -1;
	if (t < n)
		iperm[t] = (int)t;
	// slots past the last atom are phantoms: lists are padded with the
	// first of them, and a NaN position fails every cutoff test
	const double nan = __builtin_nan("");
	s.pos[t] = t < n ? sepk_pos{0.0, 0.0, 0.0, 0.0} : sepk_pos{nan, nan, nan, 0.0};
	s.qpos[t] = make_int4(0, 0, 0, t < n ? 0 : SEPK_QPHANTOM);
	s.m[t] = 1.0;
	s.type[t] = t < n ? 'A' : 0;
	s.mol[t] = -1;
	for (int c = 0; c < 3; c++) {
		s.v[c * npad + t] = 0.0;
		s.f[c * npad + t] = 0.0;
		s.a[c * npad + t] = 0.0;
		s.xn[c * npad + t] = 0.0;
		s.cn[c * npad + t] = 0;
		s.cr[c * npad + t] = 0;
	}
}

// The fused step's second position buffer: phantoms as in k_init_identity.
__global__ void k_init_alt(sepk_pos *pos, int4 *qpos, long n, long npad)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= npad)
		return;
	const double nan = __builtin_nan("");
	pos[t] = t < n ? sepk_pos{0.0, 0.0, 0.0, 0.0} : sepk_pos{nan, nan, nan, 0.0};
	qpos[t] = make_int4(0, 0, 0, t < n ? 0 : SEPK_QPHANTOM);
}

extern "C" sepdev *sepdev_create(long npart, int *err)
{
	*err = 0;
	if (npart <= 0 || npart > SEPDEV_MAXN) {
		*err = sepk_fail(SEP_GPU_EARG, "sepdev_create: npart = %ld (1 to %ld)", npart,
				 (long)SEPDEV_MAXN);
		return nullptr;
	}
	sepdev *d = (sepdev *)calloc(1, sizeof(sepdev));
	d->n = npart;
	if (hipGetDevice(&d->device) != hipSuccess)
		d->device = 0;
	// at least one slot past the last atom: the phantom that pads lists
	d->npad = (npart + 1 + 63) / 64 * 64;
	d->qpos_stale = 1;
	int rc = 0;
	if (hipStreamCreateWithFlags(&d->st, hipStreamNonBlocking) != hipSuccess) {
		*err = sepk_fail(SEP_GPU_EDEVICE, "hipStreamCreate failed");
		free(d);
		return nullptr;
	}
	rc |= alloc_state(&d->s[0], d->npad);
	rc |= alloc_state(&d->s[1], d->npad);
	rc |= dalloc(&d->pos_alt, d->npad);
	rc |= dalloc(&d->qpos_alt, d->npad);
	rc |= dalloc(&d->iperm, d->npad);
	rc |= dalloc(&d->cellid_old, d->npad);
	rc |= dalloc(&d->order_tmp, d->npad);
	rc |= dalloc(&d->cell_slot, d->npad);
	rc |= dalloc(&d->nbr_count, d->npad);
	{
		// rows per region: the widest launch is a list kernel (SEPK_SUB
		// lanes per atom, XCD-padded) or a bonded kernel over up to 20
		// terms per atom
		long rows = d->npad * SEPK_SUB / SEPK_BLOCK + 2 * SEPK_NXCD;
		const long brows = d->npad * 20 / SEPK_BLOCK + 8;
		if (brows > rows)
			rows = brows;
		d->region_rows = rows;
		d->region_words = rows * SEPK_MAXNV;
		d->partials_words = d->region_words * SEPK_NREGION;
	}
	rc |= dalloc(&d->partials, d->partials_words);
	rc |= dalloc(&d->fin_l2, (size_t)SEPK_MAXTASK * SEPK_MAXNV * SEPK_FIN_BLOCKS);
	rc |= dalloc(&d->sys, 1);
	d->ret_cap = 16;
	rc |= dalloc(&d->ret, (size_t)d->ret_cap * SEPDEV_RET_WORDS);
	d->ret_zero = (unsigned char *)calloc(d->ret_cap, 1);
	d->scal_cap = 16;
	rc |= dalloc(&d->scal, d->scal_cap);
	d->stage_bytes = (size_t)d->npad * 40 * sizeof(int);
	if (d->stage_bytes < (size_t)d->npad * 3 * sizeof(double))
		d->stage_bytes = (size_t)d->npad * 3 * sizeof(double);
	if (hipHostMalloc(&d->hstage, d->stage_bytes, hipHostMallocDefault) != hipSuccess)
		rc |= sepk_fail(SEP_GPU_EDEVICE, "hipHostMalloc staging failed");
	rc |= dalloc((char **)&d->dstage, d->stage_bytes);
	if (rc) {
		*err = rc;
		sepdev_destroy(d);
		return nullptr;
	}
	d->s[0].err = d->s[1].err = &d->sys->error;
	// list lengths start at zero so that a pair kernel that runs before any
	// build reads no list entries
	(void)hipMemsetAsync(d->nbr_count, 0, d->npad * sizeof(int), d->st);
	(void)hipMemsetAsync(d->sys, 0, sizeof(sepdev_sys), d->st);
	(void)hipMemsetAsync(d->ret, 0, (size_t)d->ret_cap * SEPDEV_RET_WORDS * sizeof(double), d->st);
	for (int b = 0; b < 2; b++) {
		k_init_identity<<<sepk_blocks(d->npad), SEPK_BLOCK, 0, d->st>>>(d->s[b], d->iperm, d->n, d->npad);
	}
	k_init_alt<<<sepk_blocks(d->npad), SEPK_BLOCK, 0, d->st>>>(d->pos_alt, d->qpos_alt, d->n,
								    d->npad);
	if (hipStreamSynchronize(d->st) != hipSuccess) {
		*err = sepk_fail(SEP_GPU_EDEVICE, "device init failed");
		sepdev_destroy(d);
		return nullptr;
	}
	return d;
}

extern "C" void sepdev_destroy(sepdev *d)
{
	if (!d)
		return;
	if (d->st)
		(void)hipStreamSynchronize(d->st);
	sepdev_graph_drop(d);
	free_state(&d->s[0]);
	free_state(&d->s[1]);
	void *ptrs[] = {d->iperm, d->cell_count, d->cell_start,
			d->cellid_old, d->order_tmp, d->cell_slot,
			d->nbr, d->nbr_count, d->partials, d->fin_l2, d->ring, d->ring_n, d->ret,
			d->scal, d->sys, d->dstage, d->blist, d->alist, d->dlist,
			d->bcontrib, d->acontrib, d->dcontrib, d->bcsr_off,
			d->bcsr_idx, d->acsr_off, d->acsr_idx, d->dcsr_off,
			d->dcsr_idx, d->blen, d->aval, d->dval};
	for (void *p : ptrs)
		if (p)
			(void)hipFree(p);
	if (d->hstage)
		(void)hipHostFree(d->hstage);
	free(d->ret_zero);
	free(d->last_lj);
	(void)hipFree(d->pos_alt);
	(void)hipFree(d->qpos_alt);
	(void)hipFree(d->tscratch);
	void *adj[] = {d->adj_g, d->adj_x, d->adj_out, d->adj_part};
	for (void *p : adj)
		if (p)
			(void)hipFree(p);
	if (d->timing) {
		for (int i = 0; i < 64; i++) {
			(void)hipEventDestroy(d->t_pair.beg[i]);
			(void)hipEventDestroy(d->t_pair.end[i]);
		}
	}
	if (d->side) {
		(void)hipStreamSynchronize(d->side);
		(void)hipEventDestroy(d->fork_ev);
		(void)hipEventDestroy(d->join_ev);
		(void)hipStreamDestroy(d->side);
	}
	if (d->st)
		(void)hipStreamDestroy(d->st);
	free(d);
}

// ---- transfers ----------------------------------------------------------------

// dst block (3 x npad, device order) <- src (3n interleaved, original order)
__global__ void k_scatter_vec(double *dst, const double *src, const int *iperm,
			      long n, long npad)
{
	long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= n)
		return;
	const long t = iperm[o];
	dst[t] = src[3 * o];
	dst[npad + t] = src[3 * o + 1];
	dst[2 * npad + t] = src[3 * o + 2];
}

__global__ void k_scatter_pos(sepk_pos *pos, const double *src, const int *iperm,
			      long n)
{
	long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= n)
		return;
	const long t = iperm[o];
	sepk_pos p = pos[t];
	p.x = src[3 * o];
	p.y = src[3 * o + 1];
	p.z = src[3 * o + 2];
	pos[t] = p;
}

__global__ void k_gather_vec(double *dst, const double *src, const int *iperm,
			     long n, long npad)
{
	long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= n)
		return;
	const long t = iperm[o];
	dst[3 * o] = src[t];
	dst[3 * o + 1] = src[npad + t];
	dst[3 * o + 2] = src[2 * npad + t];
}

__global__ void k_gather_pos(double *dst, const sepk_pos *pos, const int *iperm,
			     long n)
{
	long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= n)
		return;
	const sepk_pos p = pos[iperm[o]];
	dst[3 * o] = p.x;
	dst[3 * o + 1] = p.y;
	dst[3 * o + 2] = p.z;
}

__global__ void k_scatter_ivec(int *dst, const int *src, const int *iperm,
			       long n, long npad)
{
	long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= n)
		return;
	const long t = iperm[o];
	dst[t] = src[3 * o];
	dst[npad + t] = src[3 * o + 1];
	dst[2 * npad + t] = src[3 * o + 2];
}

__global__ void k_gather_ivec(int *dst, const int *src, const int *iperm,
			      long n, long npad)
{
	long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= n)
		return;
	const long t = iperm[o];
	dst[3 * o] = src[t];
	dst[3 * o + 1] = src[npad + t];
	dst[3 * o + 2] = src[2 * npad + t];
}

__global__ void k_zero_blocks(double *p, long count)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < count)
		p[t] = 0.0;
}

static int stage_in(sepdev *d, const void *src, size_t bytes)
{
	if (bytes > d->stage_bytes)
		return sepk_fail(SEP_GPU_EARG, "staging overflow (%zu bytes)", bytes);
	// the pinned buffer may still feed an earlier copy
	SEPK_CHECK(hipStreamSynchronize(d->st));
	memcpy(d->hstage, src, bytes);
	SEPK_CHECK(hipMemcpyAsync(d->dstage, d->hstage, bytes, hipMemcpyHostToDevice, d->st));
	return 0;
}

static int stage_out(sepdev *d, void *dst, size_t bytes)
{
	SEPK_CHECK(hipMemcpyAsync(d->hstage, d->dstage, bytes, hipMemcpyDeviceToHost, d->st));
	SEPK_CHECK(hipStreamSynchronize(d->st));
	memcpy(dst, d->hstage, bytes);
	return 0;
}

static double *vec_field(sepdev *d, int which)
{
	sepk_state &s = d->s[0];
	switch (which) {
	case SEPDEV_V: return s.v;
	case SEPDEV_F: return s.f;
	case SEPDEV_A: return s.a;
	case SEPDEV_XN: return s.xn;
	}
	return nullptr;
}

extern "C" int sepdev_put_vec(sepdev *d, int which, const double *src3n)
{
	if (d->capturing)
		return sepk_fail(SEP_GPU_ESTATE, "transfer during graph capture");
	// owed bonded gathers belong to the forces before this upload
	int rc = sepk_flush_nh(d);
	if (rc || (rc = sepk_flush_gathers(d)))
		return rc;
	if (which == SEPDEV_V)
		d->sumv2_valid = 0;
	if ((rc = stage_in(d, src3n, (size_t)d->n * 3 * sizeof(double))))
		return rc;
	const unsigned nb = sepk_blocks(d->n);
	if (which == SEPDEV_X) {
		k_scatter_pos<<<nb, SEPK_BLOCK, 0, d->st>>>(d->s[0].pos, (const double *)d->dstage, d->iperm, d->n);
		d->qpos_stale = 1; // requantized by the next mixed pair call
	} else {
		double *dst = vec_field(d, which);
		if (!dst)
			return sepk_fail(SEP_GPU_EARG, "put_vec: field %d", which);
		k_scatter_vec<<<nb, SEPK_BLOCK, 0, d->st>>>(dst, (const double *)d->dstage, d->iperm, d->n, d->npad);
		if (which == SEPDEV_F)
			d->f_zero_pending = 0;
	}
	SEPK_LAUNCH_CHECK();
	return 0;
}

extern "C" int sepdev_get_vec(sepdev *d, int which, double *dst3n)
{
	if (d->capturing)
		return sepk_fail(SEP_GPU_ESTATE, "transfer during graph capture");
	const unsigned nb = sepk_blocks(d->n);
	int rc = sepk_flush_nh(d);
	if (rc)
		return rc;
	if (which == SEPDEV_F && (rc = sepk_materialize_forces(d)))
		return rc;
	if (which == SEPDEV_X) {
		k_gather_pos<<<nb, SEPK_BLOCK, 0, d->st>>>((double *)d->dstage, d->s[0].pos, d->iperm, d->n);
	} else {
		double *src = vec_field(d, which);
		if (!src)
			return sepk_fail(SEP_GPU_EARG, "get_vec: field %d", which);
		k_gather_vec<<<nb, SEPK_BLOCK, 0, d->st>>>((double *)d->dstage, src, d->iperm, d->n, d->npad);
	}
	SEPK_LAUNCH_CHECK();
	return stage_out(d, dst3n, (size_t)d->n * 3 * sizeof(double));
}

// A host 3n array in original order into / out of any 3 x npad device block
// in device order (adjoint inputs and outputs).
int sepk_put_vec_to(sepdev *d, double *dst, const double *src3n)
{
	int rc = stage_in(d, src3n, (size_t)d->n * 3 * sizeof(double));
	if (rc)
		return rc;
	k_scatter_vec<<<sepk_blocks(d->n), SEPK_BLOCK, 0, d->st>>>(dst, (const double *)d->dstage, d->iperm, d->n, d->npad);
	SEPK_LAUNCH_CHECK();
	return 0;
}

int sepk_get_vec_from(sepdev *d, double *dst3n, const double *src)
{
	k_gather_vec<<<sepk_blocks(d->n), SEPK_BLOCK, 0, d->st>>>((double *)d->dstage, src, d->iperm, d->n, d->npad);
	SEPK_LAUNCH_CHECK();
	return stage_out(d, dst3n, (size_t)d->n * 3 * sizeof(double));
}

extern "C" int sepdev_put_ivec(sepdev *d, int which, const int *src3n)
{
	int rc = stage_in(d, src3n, (size_t)d->n * 3 * sizeof(int));
	if (rc)
		return rc;
	int *dst = which == SEPDEV_CROSS_NB ? d->s[0].cn : d->s[0].cr;
	k_scatter_ivec<<<sepk_blocks(d->n), SEPK_BLOCK, 0, d->st>>>(dst, (const int *)d->dstage, d->iperm, d->n, d->npad);
	SEPK_LAUNCH_CHECK();
	return 0;
}

extern "C" int sepdev_get_ivec(sepdev *d, int which, int *dst3n)
{
	const int *src = which == SEPDEV_CROSS_NB ? d->s[0].cn : d->s[0].cr;
	k_gather_ivec<<<sepk_blocks(d->n), SEPK_BLOCK, 0, d->st>>>((int *)d->dstage, src, d->iperm, d->n, d->npad);
	SEPK_LAUNCH_CHECK();
	return stage_out(d, dst3n, (size_t)d->n * 3 * sizeof(int));
}

__global__ void k_scatter_scalars(sepk_state s, const double *mass, const double *q,
				  const signed char *type, const int *mol,
				  const int *iperm, long n)
{
	long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= n)
		return;
	const long t = iperm[o];
	if (mass)
		s.m[t] = mass[o];
	if (q)
		s.pos[t].q = q[o];
	if (type)
		s.type[t] = type[o];
	if (mol)
		s.mol[t] = mol[o];
}

extern "C" int sepdev_put_scalars(sepdev *d, const double *mass, const double *charge,
				  const signed char *type, const int *molindex)
{
	// pack into one staging upload: mass | charge | mol | type
	const long n = d->n;
	if (mass)
		d->sumv2_valid = 0;
	if (charge)
		d->qpos_stale = 1; // the quantized records carry the charges
	const size_t bytes = (size_t)n * (2 * sizeof(double) + sizeof(int) + 1);
	if (bytes > d->stage_bytes)
		return sepk_fail(SEP_GPU_EARG, "staging overflow");
	SEPK_CHECK(hipStreamSynchronize(d->st));
	char *h = (char *)d->hstage;
	double *hm = (double *)h;
	double *hq = hm + n;
	int *hmol = (int *)(hq + n);
	signed char *ht = (signed char *)(hmol + n);
	if (mass) memcpy(hm, mass, n * sizeof(double));
	if (charge) memcpy(hq, charge, n * sizeof(double));
	if (molindex) memcpy(hmol, molindex, n * sizeof(int));
	if (type) memcpy(ht, type, n);
	SEPK_CHECK(hipMemcpyAsync(d->dstage, d->hstage, bytes, hipMemcpyHostToDevice, d->st));
	char *g = (char *)d->dstage;
	double *gm = (double *)g;
	double *gq = gm + n;
	int *gmol = (int *)(gq + n);
	signed char *gt = (signed char *)(gmol + n);
	k_scatter_scalars<<<sepk_blocks(n), SEPK_BLOCK, 0, d->st>>>(
		d->s[0], mass ? gm : nullptr, charge ? gq : nullptr,
		type ? gt : nullptr, molindex ? gmol : nullptr, d->iperm, n);
	SEPK_LAUNCH_CHECK();
	return 0;
}

extern "C" int sepdev_put_charges(sepdev *d, const double *charge)
{
	return sepdev_put_scalars(d, nullptr, charge, nullptr, nullptr);
}

extern "C" int sepdev_put_types(sepdev *d, const signed char *type)
{
	return sepdev_put_scalars(d, nullptr, nullptr, type, nullptr);
}

__global__ void k_scatter_partners(int *dst, const int *src, const int *iperm,
				   long n, long npad)
{
	long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= n)
		return;
	const long t = iperm[o];
	for (int k = 0; k < 40; k++)
		dst[k * npad + t] = src[40 * o + k];
}

extern "C" int sepdev_put_partners(sepdev *d, const int *partners40n)
{
	for (int b = 0; b < 2; b++)
		if (!d->s[b].partners) {
			int rc = dalloc(&d->s[b].partners, (size_t)40 * d->npad);
			if (rc)
				return rc;
		}
	int rc = stage_in(d, partners40n, (size_t)d->n * 40 * sizeof(int));
	if (rc)
		return rc;
	k_scatter_partners<<<sepk_blocks(d->n), SEPK_BLOCK, 0, d->st>>>(
		d->s[0].partners, (const int *)d->dstage, d->iperm, d->n, d->npad);
	SEPK_LAUNCH_CHECK();
	return 0;
}

extern "C" int sepdev_put_sys(sepdev *d, const sepdev_sys *s)
{
	// the host record replaces everything, including the cached sum m v^2
	d->max_reset_pending = 0;
	d->sumv2_valid = 0;
	SEPK_CHECK(hipStreamSynchronize(d->st));
	memcpy(d->hstage, s, sizeof *s);
	SEPK_CHECK(hipMemcpyAsync(d->sys, d->hstage, sizeof *s, hipMemcpyHostToDevice, d->st));
	return 0;
}

extern "C" int sepdev_get_sys(sepdev *d, sepdev_sys *s)
{
	int rc = sepk_flush_tasks(d);
	if (rc)
		return rc;
	SEPK_CHECK(hipMemcpyAsync(d->hstage, d->sys, sizeof *s, hipMemcpyDeviceToHost, d->st));
	SEPK_CHECK(hipStreamSynchronize(d->st));
	memcpy(s, d->hstage, sizeof *s);
	if (d->max_reset_pending)
		s->max_dist2 = 0.0;
	return 0;
}

// ---- slots --------------------------------------------------------------------

extern "C" int sepdev_ret_alloc(sepdev *d)
{
	if (d->nret == d->ret_cap) {
		double *nr;
		int rc = dalloc(&nr, (size_t)2 * d->ret_cap * SEPDEV_RET_WORDS);
		if (rc)
			return -1;
		SEPK_CHECK(hipMemsetAsync(nr, 0, (size_t)2 * d->ret_cap * SEPDEV_RET_WORDS * sizeof(double), d->st));
		SEPK_CHECK(hipMemcpyAsync(nr, d->ret, (size_t)d->ret_cap * SEPDEV_RET_WORDS * sizeof(double), hipMemcpyDeviceToDevice, d->st));
		SEPK_CHECK(hipStreamSynchronize(d->st));
		(void)hipFree(d->ret);
		d->ret = nr;
		d->ret_zero = (unsigned char *)realloc(d->ret_zero, 2 * d->ret_cap);
		memset(d->ret_zero + d->ret_cap, 0, d->ret_cap);
		d->ret_cap *= 2;
	}
	return d->nret++;
}

extern "C" int sepdev_ret_zero(sepdev *d, int slot)
{
	if (slot < 0 || slot >= d->nret)
		return sepk_fail(SEP_GPU_EARG, "ret slot %d", slot);
	// owed to the next kernel that writes the slot (it zeroes it first)
	d->ret_zero[slot] = 1;
	return 0;
}

extern "C" int sepdev_ret_get(sepdev *d, int slot, double *words)
{
	if (slot < 0 || slot >= d->nret)
		return sepk_fail(SEP_GPU_EARG, "ret slot %d", slot);
	int rc = sepk_flush_tasks(d);
	if (rc)
		return rc;
	if (d->ret_zero[slot]) {
		// a fill kernel rather than a memset node, so it may sit in a graph
		k_zero_blocks<<<1, 64, 0, d->st>>>(d->ret + (size_t)slot * SEPDEV_RET_WORDS,
						   SEPDEV_RET_WORDS);
		SEPK_LAUNCH_CHECK();
		d->ret_zero[slot] = 0;
	}
	SEPK_CHECK(hipMemcpyAsync(d->hstage, d->ret + (size_t)slot * SEPDEV_RET_WORDS,
				  SEPDEV_RET_WORDS * sizeof(double), hipMemcpyDeviceToHost, d->st));
	SEPK_CHECK(hipStreamSynchronize(d->st));
	memcpy(words, d->hstage, SEPDEV_RET_WORDS * sizeof(double));
	return 0;
}

extern "C" int sepdev_ret_put(sepdev *d, int slot, const double *words)
{
	if (slot < 0 || slot >= d->nret)
		return sepk_fail(SEP_GPU_EARG, "ret slot %d", slot);
	SEPK_CHECK(hipStreamSynchronize(d->st));
	d->ret_zero[slot] = 0;
	memcpy(d->hstage, words, SEPDEV_RET_WORDS * sizeof(double));
	SEPK_CHECK(hipMemcpyAsync(d->ret + (size_t)slot * SEPDEV_RET_WORDS, d->hstage,
				  SEPDEV_RET_WORDS * sizeof(double), hipMemcpyHostToDevice, d->st));
	SEPK_CHECK(hipStreamSynchronize(d->st));
	return 0;
}

extern "C" int sepdev_scalar_alloc(sepdev *d, double init)
{
	if (d->nscal == d->scal_cap) {
		double *ns;
		if (dalloc(&ns, (size_t)2 * d->scal_cap))
			return -1;
		SEPK_CHECK(hipMemcpyAsync(ns, d->scal, (size_t)d->scal_cap * sizeof(double), hipMemcpyDeviceToDevice, d->st));
		SEPK_CHECK(hipStreamSynchronize(d->st));
		(void)hipFree(d->scal);
		d->scal = ns;
		d->scal_cap *= 2;
	}
	int slot = d->nscal++;
	if (sepdev_scalar_put(d, slot, init))
		return -1;
	return slot;
}

extern "C" int sepdev_scalar_get(sepdev *d, int slot, double *v)
{
	int rc = sepk_flush_nh(d);
	if (rc || (rc = sepk_flush_tasks(d)))
		return rc;
	SEPK_CHECK(hipMemcpyAsync(d->hstage, d->scal + slot, sizeof(double), hipMemcpyDeviceToHost, d->st));
	SEPK_CHECK(hipStreamSynchronize(d->st));
	memcpy(v, d->hstage, sizeof(double));
	return 0;
}

extern "C" int sepdev_scalar_put(sepdev *d, int slot, double v)
{
	SEPK_CHECK(hipStreamSynchronize(d->st));
	memcpy(d->hstage, &v, sizeof(double));
	SEPK_CHECK(hipMemcpyAsync(d->scal + slot, d->hstage, sizeof(double), hipMemcpyHostToDevice, d->st));
	SEPK_CHECK(hipStreamSynchronize(d->st));
	return 0;
}

// ---- misc -----------------------------------------------------------------------

extern "C" int sepdev_synchronize(sepdev *d)
{
	int rc = sepk_flush_tasks(d);
	if (rc)
		return rc;
	SEPK_CHECK(hipStreamSynchronize(d->st));
	return 0;
}

extern "C" int sepdev_timing(sepdev *d, int on)
{
	if (on && !d->timing) {
		for (int i = 0; i < 64; i++) {
			SEPK_CHECK(hipEventCreate(&d->t_pair.beg[i]));
			SEPK_CHECK(hipEventCreate(&d->t_pair.end[i]));
		}
		d->t_pair.used = 0;
		d->t_pair.ms = 0.0;
		d->t_pair.count = 0;
		d->timing = 1;
	} else if (!on && d->timing) {
		SEPK_CHECK(hipStreamSynchronize(d->st));
		for (int i = 0; i < 64; i++) {
			(void)hipEventDestroy(d->t_pair.beg[i]);
			(void)hipEventDestroy(d->t_pair.end[i]);
		}
		d->timing = 0;
	}
	return 0;
}

// Folds recorded event pairs into the running total.
static int timer_flush(sepk_timer *t)
{
	for (int i = 0; i < t->used; i++) {
		SEPK_CHECK(hipEventSynchronize(t->end[i]));
		float ms = 0.0f;
		SEPK_CHECK(hipEventElapsedTime(&ms, t->beg[i], t->end[i]));
		t->ms += ms;
		t->count++;
	}
	t->used = 0;
	return 0;
}

int sepk_timer_begin(sepdev *d, sepk_timer *t)
{
	if (!d->timing || d->capturing)
		return 0;
	if (t->used == 64) {
		int rc = timer_flush(t);
		if (rc)
			return rc;
	}
	SEPK_CHECK(hipEventRecord(t->beg[t->used], d->st));
	return 0;
}

int sepk_timer_end(sepdev *d, sepk_timer *t)
{
	if (!d->timing || d->capturing)
		return 0;
	SEPK_CHECK(hipEventRecord(t->end[t->used], d->st));
	t->used++;
	return 0;
}

extern "C" int sepdev_timing_reset(sepdev *d)
{
	if (!d->timing)
		return 0;
	int rc = timer_flush(&d->t_pair);
	d->t_pair.ms = 0.0;
	d->t_pair.count = 0;
	return rc;
}

extern "C" int sepdev_timing_get(sepdev *d, const char *name, double *ms, long *count)
{
	if (!d->timing)
		return sepk_fail(SEP_GPU_ESTATE, "timing not enabled");
	if (strcmp(name, "pair") != 0)
		return sepk_fail(SEP_GPU_EARG, "unknown timer '%s'", name);
	int rc = timer_flush(&d->t_pair);
	*ms = d->t_pair.ms;
	*count = d->t_pair.count;
	return rc;
}

extern "C" int sepdev_graph_begin(sepdev *d, int slot)
{
	if (slot < 0 || slot > 1)
		return sepk_fail(SEP_GPU_EARG, "graph slot %d", slot);
	// gathers owed before the capture are applied now, not in every replay
	int rc = sepk_flush_gathers(d);
	if (rc)
		return rc;
	if (slot == 0)
		sepdev_graph_drop(d);
	else if (d->exec[1] || d->graph[1])
		return sepk_fail(SEP_GPU_ESTATE, "graph slot 1 in use");
	SEPK_CHECK(hipStreamBeginCapture(d->st, hipStreamCaptureModeRelaxed));
	d->capturing = 1;
	return 0;
}

extern "C" int sepdev_capturing(sepdev *d) { return d->capturing; }

extern "C" int sepdev_graph_end(sepdev *d, int slot)
{
	// the captured step must leave nothing queued for the host side
	int rc = sepk_flush_gathers(d);
	if (!rc)
		rc = sepk_flush_tasks(d);
	d->capturing = 0;
	if (rc) {
		hipGraph_t g = nullptr;
		if (hipStreamEndCapture(d->st, &g) == hipSuccess && g)
			(void)hipGraphDestroy(g);
		return rc;
	}
	SEPK_CHECK(hipStreamEndCapture(d->st, &d->graph[slot & 1]));
	SEPK_CHECK(hipGraphInstantiate(&d->exec[slot & 1], d->graph[slot & 1], nullptr, nullptr, 0));
	return 0;
}

extern "C" int sepdev_graph_launch(sepdev *d, int slot, long times)
{
	if (slot < 0 || slot > 1 || !d->exec[slot])
		return sepk_fail(SEP_GPU_ESTATE, "no graph in slot %d", slot);
	for (long i = 0; i < times; i++)
		SEPK_CHECK(hipGraphLaunch(d->exec[slot], d->st));
	return 0;
}

extern "C" int sepdev_graph_drop(sepdev *d)
{
	// the replays of the last run may still be queued on the stream: the
	// executable graphs go only after they are done
	if (d->exec[0] || d->exec[1])
		(void)hipStreamSynchronize(d->st);
	for (int k = 0; k < 2; k++) {
		if (d->exec[k])
			(void)hipGraphExecDestroy(d->exec[k]);
		if (d->graph[k])
			(void)hipGraphDestroy(d->graph[k]);
		d->exec[k] = nullptr;
		d->graph[k] = nullptr;
	}
	return 0;
}

// ---- device checkpoint ---------------------------------------------------
//
// The state a resident run carries from one step to the next, as it stands
// between steps: every per-atom array in device order (with the order
// itself), the neighbour list, the system scalars and the host-side facts
// the next step's kernels depend on (list geometry, whether sum m v^2 of
// the current velocities is known, the box of the quantized positions).
// Read back into a device of the same size, the run continues bit for bit.

#define SEPK_CKPT_MAGIC "SEPGPUC1"

struct sepk_ckpt_head {
	char magic[8];
	long n, npad;
	int lcap, cap, cap_limit, list_built;
	int has_partners, sumv2_valid, qpos_stale, pad_;
	double qL[3];
	sepdev_grid grid;
	sepdev_sys sys;
};

static int ckpt_io(FILE *f, void *host, size_t bytes, int writing)
{
	const size_t done = writing ? fwrite(host, 1, bytes, f) : fread(host, 1, bytes, f);
	if (done != bytes)
		return sepk_fail(SEP_GPU_EARG, "checkpoint: short %s (%zu of %zu bytes)",
				 writing ? "write" : "read", done, bytes);
	return 0;
}

// One device array through a host buffer, in either direction.
static int ckpt_array(sepdev *d, FILE *f, void *dev, size_t bytes, void *buf, int writing)
{
	int rc;
	if (writing) {
		SEPK_CHECK(hipMemcpy(buf, dev, bytes, hipMemcpyDeviceToHost));
		return ckpt_io(f, buf, bytes, 1);
	}
	if ((rc = ckpt_io(f, buf, bytes, 0)))
		return rc;
	SEPK_CHECK(hipMemcpy(dev, buf, bytes, hipMemcpyHostToDevice));
	return 0;
}

static int ckpt_arrays(sepdev *d, FILE *f, const sepk_ckpt_head &h, int writing)
{
	const size_t np = (size_t)d->npad;
	const size_t list_ints = (size_t)h.lcap * SEPK_SUB * np;
	size_t most = np * 40 * sizeof(int);
	if (h.list_built && list_ints * sizeof(int) > most)
		most = list_ints * sizeof(int);
	void *buf = malloc(most);
	if (!buf)
		return sepk_fail(SEP_GPU_EHOST, "checkpoint: out of host memory");
	sepk_state &s = d->s[0];
	struct {
		void *p;
		size_t bytes;
	} parts[] = {
		{s.pos, np * sizeof(sepk_pos)}, {s.qpos, np * sizeof(int4)},
		{s.v, 3 * np * sizeof(double)}, {s.f, 3 * np * sizeof(double)},
		{s.a, 3 * np * sizeof(double)}, {s.xn, 3 * np * sizeof(double)},
		{s.cn, 3 * np * sizeof(int)},   {s.cr, 3 * np * sizeof(int)},
		{s.m, np * sizeof(double)},     {s.type, np},
		{s.mol, np * sizeof(int)},      {s.orig, np * sizeof(int)},
		{d->iperm, np * sizeof(int)},   {d->nbr_count, np * sizeof(int)},
	};
	int rc = 0;
	for (auto &p : parts)
		if ((rc = ckpt_array(d, f, p.p, p.bytes, buf, writing)))
			break;
	if (!rc && h.has_partners)
		rc = ckpt_array(d, f, s.partners, np * 40 * sizeof(int), buf, writing);
	if (!rc && h.list_built)
		rc = ckpt_array(d, f, d->nbr, list_ints * sizeof(int), buf, writing);
	free(buf);
	return rc;
}

extern "C" int sepdev_checkpoint_write(sepdev *d, void *file)
{
	FILE *f = (FILE *)file;
	if (d->capturing)
		return sepk_fail(SEP_GPU_ESTATE, "checkpoint during a graph capture");
	// the step's deferred work is done: the recorded calls, the finisher
	int rc = sepk_flush_pair(d);
	if (rc || (rc = sepk_flush_nh(d)) || (rc = sepk_materialize_forces(d)) ||
	    (rc = sepk_flush_tasks(d)))
		return rc;
	sepk_ckpt_head h;
	memset(&h, 0, sizeof h);
	memcpy(h.magic, SEPK_CKPT_MAGIC, 8);
	h.n = d->n;
	h.npad = d->npad;
	h.list_built = d->list_built && d->nbr;
	h.lcap = h.list_built ? d->lcap : 0;
	h.cap = h.list_built ? d->cap : 0;
	h.cap_limit = d->cap_limit;
	h.has_partners = d->s[0].partners != nullptr;
	h.sumv2_valid = d->sumv2_valid;
	h.qpos_stale = d->qpos_stale;
	for (int k = 0; k < 3; k++)
		h.qL[k] = d->qL[k];
	h.grid = d->grid;
	if ((rc = sepdev_get_sys(d, &h.sys)))
		return rc;
	SEPK_CHECK(hipStreamSynchronize(d->st));
	if ((rc = ckpt_io(f, &h, sizeof h, 1)))
		return rc;
	return ckpt_arrays(d, f, h, 1);
}

extern "C" int sepdev_checkpoint_read(sepdev *d, void *file)
{
	FILE *f = (FILE *)file;
	if (d->capturing)
		return sepk_fail(SEP_GPU_ESTATE, "restore during a graph capture");
	sepk_ckpt_head h;
	int rc = ckpt_io(f, &h, sizeof h, 0);
	if (rc)
		return rc;
	if (memcmp(h.magic, SEPK_CKPT_MAGIC, 8) != 0)
		return sepk_fail(SEP_GPU_EARG, "not a device checkpoint of this library");
	if (h.n != d->n || h.npad != d->npad)
		return sepk_fail(SEP_GPU_EARG, "checkpoint of %ld atoms, system of %ld", h.n, d->n);
	if ((rc = sepk_flush_pair(d)) || (rc = sepk_flush_nh(d)) || (rc = sepk_flush_tasks(d)))
		return rc;
	SEPK_CHECK(hipStreamSynchronize(d->st));
	if (h.list_built && (rc = sepdev_reserve_list(d, h.cap)))
		return rc;
	if (h.has_partners && !d->s[0].partners) {
		for (int b = 0; b < 2; b++)
			if ((rc = dalloc(&d->s[b].partners, (size_t)40 * d->npad)))
				return rc;
	}
	if ((rc = ckpt_arrays(d, f, h, 0)))
		return rc;
	if ((rc = sepdev_put_sys(d, &h.sys)))
		return rc;
	SEPK_CHECK(hipStreamSynchronize(d->st));
	// the host side of the device state
	d->list_built = h.list_built;
	if (h.list_built) {
		d->cap = h.cap;
		d->grid = h.grid;
	}
	d->cap_limit = h.cap_limit;
	d->sumv2_valid = h.sumv2_valid;
	d->qpos_stale = h.qpos_stale;
	for (int k = 0; k < 3; k++)
		d->qL[k] = h.qL[k];
	d->f_zero_pending = 0;
	d->ngpend = 0; // owed gathers were for the forces just replaced
	d->gpend_over = 0;
	d->max_reset_pending = 0;
	d->nh_pending = 0;
	d->fork_ready = 0;
	return 0;
}
