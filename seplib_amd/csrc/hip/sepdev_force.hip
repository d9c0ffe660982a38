// sepdev_force.hip -- pair forces over the full neighbour list, the
// Nose-Hoover and leapfrog updates, and list extraction for the host.
//
// Pair kernels: one thread per atom (device order), the full list read in
// its interleaved layout, forces kept in registers and written once per
// atom (no atomics).  Energy and stress are summed per atom, halved (each
// pair is seen from both ends), reduced per block in a fixed order and
// finished by the last block, so repeated runs give identical bits.
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sepdev_impl.h"

int sepk_timer_begin(sepdev *d, sepk_timer *t);
int sepk_timer_end(sepdev *d, sepk_timer *t);

// ---- lazy force reset ------------------------------------------------------

__global__ void k_zero_forces(double *f, long count)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < count)
		f[t] = 0.0;
}

int sepk_materialize_forces(sepdev *d)
{
	if (!d->f_zero_pending)
		return 0;
	const long cnt = 3 * d->npad;
	k_zero_forces<<<sepk_blocks(cnt), SEPK_BLOCK, 0, d->st>>>(d->s[0].f, cnt);
	SEPK_LAUNCH_CHECK();
	d->f_zero_pending = 0;
	return 0;
}

extern "C" int sepdev_zero_forces(sepdev *d)
{
	// The next force kernel overwrites instead of accumulating; anything
	// else that reads f first materialises the zeros.
	d->f_zero_pending = 1;
	return 0;
}

// ---- LJ family ---------------------------------------------------------------

struct sepk_lj {
	double cf2, eps48, eps4, awh, aw, sig2, ushift;
	double L0, L1, L2;
	float eps48f, eps4f, awhf, awf, sig2f, ushiftf;
	signed char t0, t1;
};

// Shared epilogue: write or add the per-atom force, reduce energy + stress.
// acc = {e, wxx, wxy, wxz, wyy, wyz, wzz}, each summed over the full list.
template <bool OVERWRITE>
__device__ __forceinline__ void sepk_pair_epilogue(
	long t, long n, long npad, bool act, double fx, double fy, double fz,
	double (&acc)[7], double *f, double *partial, unsigned *counter,
	double *ret, int epot_assign, int also_ecoul)
{
	if (t < n) {
		if (OVERWRITE) {
			f[t] = fx;
			f[npad + t] = fy;
			f[2 * npad + t] = fz;
		} else if (act) {
			f[t] += fx;
			f[npad + t] += fy;
			f[2 * npad + t] += fz;
		}
	}
	sepk_block_sum<7>(acc, partial);
	if (sepk_last_block(counter)) {
		__shared__ double tot[7];
		if (threadIdx.x < 7)
			tot[threadIdx.x] = sepk_sum_partials(partial, gridDim.x, 7, threadIdx.x);
		__syncthreads();
		if (threadIdx.x == 0) {
			const double e = 0.5 * tot[0];
			if (epot_assign)
				ret[SEPDEV_RET_EPOT] = e;
			else
				ret[SEPDEV_RET_EPOT] += e;
			if (also_ecoul)
				ret[SEPDEV_RET_ECOUL] += e;
			sepk_add_sym9(ret + SEPDEV_RET_POTP, tot + 1, 0.5);
			*counter = 0;
		}
	}
}

template <bool TYPED, bool MIXED, bool OVERWRITE>
__global__ __launch_bounds__(SEPK_BLOCK) void k_lj(
	sepk_state s, const int *__restrict__ nbr, const int *__restrict__ cnt,
	long n, long npad, sepk_lj c, double *partial, unsigned *counter,
	double *ret, int epot_assign)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double fx = 0.0, fy = 0.0, fz = 0.0;
	double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	bool act = t < n;
	signed char ti = 0;
	if (act && TYPED) {
		ti = s.type[t];
		act = ti == c.t0 || ti == c.t1;
	}
	if (act) {
		const sepk_pos pi = s.pos[t];
		const int nc = cnt[t];
		for (int k = 0; k < nc; k++) {
			const int j = nbr[(long)k * npad + t];
			if (TYPED) {
				const signed char tj = s.type[j];
				if (!((ti == c.t0 && tj == c.t1) || (ti == c.t1 && tj == c.t0)))
					continue;
			}
			const sepk_pos pj = s.pos[j];
			const double dx = sepk_wrap(pi.x - pj.x, c.L0);
			const double dy = sepk_wrap(pi.y - pj.y, c.L1);
			const double dz = sepk_wrap(pi.z - pj.z, c.L2);
			const double r2 = dx * dx + dy * dy + dz * dz;
			if (!(r2 < c.cf2))
				continue;
			if (MIXED) {
				const float q = c.sig2f * __builtin_amdgcn_rcpf((float)r2);
				const float q3 = q * q * q;
				const double ft = (double)(c.eps48f * q3 * (q3 - c.awhf) * q);
				const double gx = ft * dx, gy = ft * dy, gz = ft * dz;
				fx += gx;
				fy += gy;
				fz += gz;
				acc[0] += (double)(c.eps4f * q3 * (q3 - c.awf) + c.ushiftf);
				acc[1] += gx * dx;
				acc[2] += gx * dy;
				acc[3] += gx * dz;
				acc[4] += gy * dy;
				acc[5] += gy * dz;
				acc[6] += gz * dz;
			} else {
				const double q = c.sig2 / r2;
				const double q3 = q * q * q;
				const double ft = c.eps48 * q3 * (q3 - c.awh) * q;
				const double gx = ft * dx, gy = ft * dy, gz = ft * dz;
				fx += gx;
				fy += gy;
				fz += gz;
				acc[0] += c.eps4 * q3 * (q3 - c.aw) + c.ushift;
				acc[1] += gx * dx;
				acc[2] += gx * dy;
				acc[3] += gx * dz;
				acc[4] += gy * dy;
				acc[5] += gy * dz;
				acc[6] += gz * dz;
			}
		}
	}
	sepk_pair_epilogue<OVERWRITE>(t, n, npad, act, fx, fy, fz, acc, s.f, partial,
				      counter, ret, epot_assign, 0);
}

typedef void (*sepk_lj_kernel)(sepk_state, const int *, const int *, long, long,
			       sepk_lj, double *, unsigned *, double *, int);

template <bool T, bool M>
static sepk_lj_kernel pick_lj(bool over)
{
	return over ? k_lj<T, M, true> : k_lj<T, M, false>;
}

extern "C" int sepdev_lj(sepdev *d, const sepdev_lj_args *a, int ret_slot)
{
	if (!d->list_built)
		return sepk_fail(SEP_GPU_ESTATE, "pair forces before a neighbour list was built");
	if (ret_slot < 0 || ret_slot >= d->nret)
		return sepk_fail(SEP_GPU_EARG, "ret slot %d", ret_slot);
	sepk_lj c;
	c.cf2 = a->cf2;
	c.eps48 = a->eps48;
	c.eps4 = a->eps4;
	c.awh = a->awh;
	c.aw = a->aw;
	c.sig2 = a->sig2;
	c.ushift = a->ushift;
	c.L0 = a->L[0];
	c.L1 = a->L[1];
	c.L2 = a->L[2];
	c.eps48f = (float)a->eps48;
	c.eps4f = (float)a->eps4;
	c.awhf = (float)a->awh;
	c.awf = (float)a->aw;
	c.sig2f = (float)a->sig2;
	c.ushiftf = (float)a->ushift;
	c.t0 = a->t0;
	c.t1 = a->t1;
	const bool over = d->f_zero_pending != 0;
	sepk_lj_kernel k;
	if (a->typed)
		k = a->mixed ? pick_lj<true, true>(over) : pick_lj<true, false>(over);
	else
		k = a->mixed ? pick_lj<false, true>(over) : pick_lj<false, false>(over);
	int rc = sepk_timer_begin(d, &d->t_pair);
	if (rc)
		return rc;
	k<<<sepk_blocks(d->n), SEPK_BLOCK, 0, d->st>>>(
		d->s[0], d->nbr, d->nbr_count, d->n, d->npad, c, d->partials,
		d->counters + SEPK_CTR_PAIR, d->ret + (size_t)ret_slot * SEPDEV_RET_WORDS,
		a->epot_assign);
	SEPK_LAUNCH_CHECK();
	d->f_zero_pending = 0;
	return sepk_timer_end(d, &d->t_pair);
}

// ---- shifted-force Coulomb (source/sepcoulomb.c:96-160) ---------------------

struct sepk_coul {
	double cf, cf2, icf, icf2;
	double L0, L1, L2;
};

template <bool OVERWRITE>
__global__ __launch_bounds__(SEPK_BLOCK) void k_coulomb(
	sepk_state s, const int *__restrict__ nbr, const int *__restrict__ cnt,
	long n, long npad, sepk_coul c, double *partial, unsigned *counter,
	double *ret)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double fx = 0.0, fy = 0.0, fz = 0.0;
	double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	bool act = t < n;
	sepk_pos pi = {0.0, 0.0, 0.0, 0.0};
	if (act) {
		pi = s.pos[t];
		act = !(fabs(pi.q) < DBL_EPSILON);
	}
	if (act) {
		const int nc = cnt[t];
		for (int k = 0; k < nc; k++) {
			const int j = nbr[(long)k * npad + t];
			const sepk_pos pj = s.pos[j];
			if (fabs(pj.q) < DBL_EPSILON)
				continue;
			const double dx = sepk_wrap(pi.x - pj.x, c.L0);
			const double dy = sepk_wrap(pi.y - pj.y, c.L1);
			const double dz = sepk_wrap(pi.z - pj.z, c.L2);
			const double r2 = dx * dx + dy * dy + dz * dz;
			if (!(r2 < c.cf2))
				continue;
			const double zz = pj.q * pi.q;
			const double r = sqrt(r2);
			const double ft = zz * (1.0 / r2 - c.icf2) / r;
			const double gx = ft * dx, gy = ft * dy, gz = ft * dz;
			fx += gx;
			fy += gy;
			fz += gz;
			acc[0] += zz * (1.0 / r + (r - c.cf) * c.icf2 - c.icf);
			acc[1] += gx * dx;
			acc[2] += gx * dy;
			acc[3] += gx * dz;
			acc[4] += gy * dy;
			acc[5] += gy * dz;
			acc[6] += gz * dz;
		}
	}
	sepk_pair_epilogue<OVERWRITE>(t, n, npad, act, fx, fy, fz, acc, s.f, partial,
				      counter, ret, 0, 1);
}

extern "C" int sepdev_coulomb(sepdev *d, const sepdev_coul_args *a, int ret_slot)
{
	if (!d->list_built)
		return sepk_fail(SEP_GPU_ESTATE, "Coulomb forces before a neighbour list was built");
	if (ret_slot < 0 || ret_slot >= d->nret)
		return sepk_fail(SEP_GPU_EARG, "ret slot %d", ret_slot);
	sepk_coul c = {a->cf, a->cf2, a->icf, a->icf2, a->L[0], a->L[1], a->L[2]};
	const bool over = d->f_zero_pending != 0;
	auto k = over ? k_coulomb<true> : k_coulomb<false>;
	k<<<sepk_blocks(d->n), SEPK_BLOCK, 0, d->st>>>(
		d->s[0], d->nbr, d->nbr_count, d->n, d->npad, c, d->partials,
		d->counters + SEPK_CTR_COUL, d->ret + (size_t)ret_slot * SEPDEV_RET_WORDS);
	SEPK_LAUNCH_CHECK();
	d->f_zero_pending = 0;
	return 0;
}

// ---- Nose-Hoover (source/sepintgr.c:149-168) ---------------------------------

__global__ __launch_bounds__(SEPK_BLOCK) void k_nh_sum(
	sepk_state s, long n, long npad, double npart, double Td,
	double dt_tau2, double *alpha, double *partial, unsigned *counter,
	sepdev_sys *sys)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double acc[1] = {0.0};
	if (t < n) {
		const double m = s.m[t];
		for (int k = 0; k < 3; k++) {
			const double v = s.v[k * npad + t];
			acc[0] += v * v * m;
		}
	}
	sepk_block_sum<1>(acc, partial);
	if (sepk_last_block(counter)) {
		if (threadIdx.x == 0) {
			const double sum = sepk_sum_partials(partial, gridDim.x, 1, 0);
			sys->sumv2 = sum;
			// ekin = 0.5*sum/npart; T = 0.666667*ekin
			const double ekin = 0.5 * sum / npart;
			const double temp = 0.666667 * ekin;
			*alpha = *alpha + dt_tau2 * (temp / Td - 1.0);
			*counter = 0;
		}
	}
}

__global__ void k_nh_apply(sepk_state s, long n, long npad, const double *alpha)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	const double al = *alpha;
	const double m = s.m[t];
	for (int k = 0; k < 3; k++)
		s.f[k * npad + t] -= al * m * s.v[k * npad + t];
}

extern "C" int sepdev_nosehoover(sepdev *d, double Td, double tau, double dt,
				 int alpha_slot)
{
	if (alpha_slot < 0 || alpha_slot >= d->nscal)
		return sepk_fail(SEP_GPU_EARG, "alpha slot %d", alpha_slot);
	int rc = sepk_materialize_forces(d);
	if (rc)
		return rc;
	double *alpha = d->scal + alpha_slot;
	k_nh_sum<<<sepk_blocks(d->n), SEPK_BLOCK, 0, d->st>>>(
		d->s[0], d->n, d->npad, (double)d->n, Td, dt / (tau * tau), alpha,
		d->partials, d->counters + SEPK_CTR_NH, d->sys);
	SEPK_LAUNCH_CHECK();
	k_nh_apply<<<sepk_blocks(d->n), SEPK_BLOCK, 0, d->st>>>(d->s[0], d->n, d->npad, alpha);
	SEPK_LAUNCH_CHECK();
	return 0;
}

// ---- leapfrog (source/sepintgr.c:18-88) ---------------------------------------

__global__ __launch_bounds__(SEPK_BLOCK) void k_leapfrog(
	sepk_state s, long n, long npad, double dt, double L0, double L1, double L2,
	double skin, double *partial, unsigned *counter, sepdev_sys *sys,
	double *ret)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	double d2 = 0.0;
	const int pending = sys->xn_pending;
	if (t < n) {
		const double L[3] = {L0, L1, L2};
		sepk_pos p = s.pos[t];
		double x[3] = {p.x, p.y, p.z};
		const double m = s.m[t];
		double vh[3];
		for (int k = 0; k < 3; k++) {
			const long i = k * npad + t;
			double xn = s.xn[i];
			int cn = s.cn[i];
			if (pending) {
				xn = x[k];
				cn = 0;
				s.xn[i] = xn;
			}
			const double a = s.f[i] / m;
			double v = s.v[i] + a * dt;
			x[k] += v * dt;
			vh[k] = v - 0.5 * a * dt;
			acc[0] += vh[k] * vh[k] * m;
			int cr = s.cr[i];
			if (x[k] > L[k]) {
				x[k] -= L[k];
				cn++;
				cr++;
			} else if (x[k] < 0.0) {
				x[k] += L[k];
				cn--;
				cr--;
			}
			const double dr = (x[k] + cn * L[k]) - xn;
			d2 += dr * dr;
			s.a[i] = a;
			s.v[i] = v;
			s.cn[i] = cn;
			s.cr[i] = cr;
		}
		p.x = x[0];
		p.y = x[1];
		p.z = x[2];
		s.pos[t] = p;
		acc[1] = vh[0] * vh[0] * m;
		acc[2] = vh[0] * vh[1] * m;
		acc[3] = vh[0] * vh[2] * m;
		acc[4] = vh[1] * vh[1] * m;
		acc[5] = vh[1] * vh[2] * m;
		acc[6] = vh[2] * vh[2] * m;
	}
	const long nblk = gridDim.x;
	sepk_block_sum<7>(acc, partial);
	sepk_block_max(d2, partial + nblk * 7 + blockIdx.x);
	if (sepk_last_block(counter)) {
		__shared__ double tot[7];
		if (threadIdx.x < 7)
			tot[threadIdx.x] = sepk_sum_partials(partial, nblk, 7, threadIdx.x);
		__syncthreads();
		if (threadIdx.x == 0) {
			double mx = sys->max_dist2;
			for (long b = 0; b < nblk; b++) {
				const double v = sepk_ld(partial + nblk * 7 + b);
				if (v > mx)
					mx = v;
			}
			sys->max_dist2 = mx;
			ret[SEPDEV_RET_EKIN] += 0.5 * tot[0];
			sepk_add_sym9(ret + SEPDEV_RET_KINP, tot + 1, 1.0);
			if (pending)
				sys->xn_pending = 0;
			if (sqrt(mx) > skin * 0.5) {
				sys->neighb_flag = 1;
				sys->nupdate_neighb++;
				sys->xn_pending = 1;
			}
			sys->tnow += dt;
			*counter = 0;
		}
	}
}

extern "C" int sepdev_leapfrog(sepdev *d, double dt, double skin, const double *L,
			       int ret_slot)
{
	if (ret_slot < 0 || ret_slot >= d->nret)
		return sepk_fail(SEP_GPU_EARG, "ret slot %d", ret_slot);
	int rc = sepk_materialize_forces(d);
	if (rc)
		return rc;
	k_leapfrog<<<sepk_blocks(d->n), SEPK_BLOCK, 0, d->st>>>(
		d->s[0], d->n, d->npad, dt, L[0], L[1], L[2], skin, d->partials,
		d->counters + SEPK_CTR_LF, d->sys, d->ret + (size_t)ret_slot * SEPDEV_RET_WORDS);
	SEPK_LAUNCH_CHECK();
	return 0;
}

// ---- list extraction ------------------------------------------------------------

__global__ void k_half_list(const int *orig, const int *nbr, const int *cnt, long n,
			    long npad, int *counts, int *half, int stride)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	const int o = orig[t];
	int c = 0;
	for (int k = 0; k < cnt[t]; k++) {
		const int oj = orig[nbr[(long)k * npad + t]];
		if (oj > o) {
			if (c < stride)
				half[(long)o * stride + c] = oj;
			c++;
		}
	}
	counts[o] = c;
}

extern "C" int sepdev_get_half_list(sepdev *d, int *counts, int *half, int stride,
				    int *maxc)
{
	if (!d->list_built)
		return sepk_fail(SEP_GPU_ESTATE, "no neighbour list");
	int *dc = nullptr, *dh = nullptr;
	SEPK_CHECK(hipMalloc(&dc, d->n * sizeof(int)));
	SEPK_CHECK(hipMalloc(&dh, (size_t)d->n * stride * sizeof(int)));
	k_half_list<<<sepk_blocks(d->n), SEPK_BLOCK, 0, d->st>>>(d->s[0].orig, d->nbr, d->nbr_count,
								 d->n, d->npad, dc, dh, stride);
	SEPK_LAUNCH_CHECK();
	SEPK_CHECK(hipMemcpyAsync(counts, dc, d->n * sizeof(int), hipMemcpyDeviceToHost, d->st));
	SEPK_CHECK(hipMemcpyAsync(half, dh, (size_t)d->n * stride * sizeof(int), hipMemcpyDeviceToHost, d->st));
	SEPK_CHECK(hipStreamSynchronize(d->st));
	(void)hipFree(dc);
	(void)hipFree(dh);
	int mc = 0;
	for (long o = 0; o < d->n; o++)
		if (counts[o] > mc)
			mc = counts[o];
	*maxc = mc;
	return 0;
}

extern "C" int sepdev_list_stats(sepdev *d, double *mean_full, int *max_full)
{
	if (!d->list_built)
		return sepk_fail(SEP_GPU_ESTATE, "no neighbour list");
	int *h = (int *)malloc(d->n * sizeof(int));
	SEPK_CHECK(hipMemcpyAsync(h, d->nbr_count, d->n * sizeof(int), hipMemcpyDeviceToHost, d->st));
	SEPK_CHECK(hipStreamSynchronize(d->st));
	double s = 0.0;
	int mx = 0;
	for (long i = 0; i < d->n; i++) {
		s += h[i];
		if (h[i] > mx)
			mx = h[i];
	}
	free(h);
	*mean_full = s / (double)d->n;
	*max_full = mx;
	return 0;
}
