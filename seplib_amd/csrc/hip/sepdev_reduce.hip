// sepdev_reduce.hip -- the one-block finisher that turns per-block partial
// rows into the energies, stress tensors, thermostat state and rebuild
// bookkeeping of a step, and the deferred-work helpers around it.
//
// Kernels leave one row per block in a region of the partials buffer and
// queue a task; queued tasks run together in one launch of k_finish at the
// next flush (each leapfrog, each host read, the end of a graph capture).
// Every reduction is fixed-order -- each thread sums a strided set of rows
// in row order, then a shared-memory tree of fixed shape -- so repeated runs
// give identical bits.
#include <string.h>

#include "sepdev_impl.h"

#define FIN_THREADS 1024

__device__ static void fin_apply(const sepk_task &k, const double *tot, sepdev_sys *sys)
{
	double *ret = k.ret;
	if (ret && k.zero_first)
		for (int w = 0; w < SEPDEV_RET_WORDS; w++)
			ret[w] = 0.0;
	switch (k.mode) {
	case SEPK_FIN_PAIR: {
		// each pair was seen from both ends: halve energy and stress
		const double e = 0.5 * tot[0];
		if (k.epot_assign)
			ret[SEPDEV_RET_EPOT] = e;
		else
			ret[SEPDEV_RET_EPOT] += e;
		if (k.also_ecoul)
			ret[SEPDEV_RET_ECOUL] += e;
		sepk_add_sym9(ret + SEPDEV_RET_POTP, tot + 1, 0.5);
		break;
	}
	case SEPK_FIN_BONDED:
		ret[SEPDEV_RET_EPOT] += tot[0];
		if (k.with_stress) {
			sepk_add_sym9(ret + SEPDEV_RET_POTP, tot + 1, 1.0);
			sepk_add_sym9(ret + SEPDEV_RET_BONDP, tot + 1, 1.0);
		}
		break;
	case SEPK_FIN_NH: {
		// nblk == 0: the sum of m v^2 is already in sys->sumv2
		const double sum = k.nblk > 0 ? tot[0] : sys->sumv2;
		sys->sumv2 = sum;
		*k.alpha = sepk_nh_alpha(*k.alpha, sum, k.npart, k.Td, k.dt_tau2);
		break;
	}
	case SEPK_FIN_LF: {
		if (k.nh_fused)
			*k.alpha = sepk_nh_alpha(*k.alpha, sys->sumv2, k.npart, k.Td, k.dt_tau2);
		ret[SEPDEV_RET_EKIN] += 0.5 * tot[0];
		sepk_add_sym9(ret + SEPDEV_RET_KINP, tot + 1, 1.0);
		// source/sepintgr.c:61,72-84: running maximum since sep_reset_force
		double mx = k.reset_max ? 0.0 : sys->max_dist2;
		if (tot[7] > mx)
			mx = tot[7];
		sys->max_dist2 = mx;
		sys->sumv2 = tot[8];
		sys->xn_pending = 0; // applied by the leapfrog kernel
		if (sqrt(mx) > k.skin * 0.5) {
			sys->neighb_flag = 1;
			sys->nupdate_neighb++;
			sys->xn_pending = 1;
		}
		sys->tnow += k.dt;
		break;
	}
	case SEPK_FIN_BUILD:
		sys->max_full = (int)tot[0];
		sys->neighb_flag = 0;
		sys->xn_pending = 0; // applied by the permutation
		sys->unclean = 0;    // judged afresh by the next binning
		break;
	}
}

// One workgroup.  Per task: each thread sums rows tid, tid + 1024, ... of
// every column (column-major rows: consecutive threads read consecutive
// words), then a wave butterfly and a combine over the 16 waves, all in a
// fixed order.
__global__ __launch_bounds__(FIN_THREADS) void k_finish(const double *partials, sepk_fin f)
{
	__shared__ double wred[SEPK_MAXNV][FIN_THREADS / 64];
	__shared__ double tot[SEPK_MAXNV];
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	sepdev_sys *sys = f.sys;
	for (int ti = 0; ti < f.ntask; ti++) {
		const sepk_task &k = f.t[ti];
		// a conditional build that did not run leaves nothing to finish
		if (k.mode == SEPK_FIN_BUILD && k.conditional && sys->neighb_flag == 0)
			continue;
		const double *p = partials + (size_t)k.region * f.region_words;
		// every load of the task is issued before any is used (addresses are
		// clamped to a valid row instead of branched around): one memory
		// round trip per task instead of one per column
		double a[SEPK_MAXNV];
		if (k.nblk > 0) {
			double v[SEPK_MAXNV][4];
#pragma unroll
			for (int i = 0; i < SEPK_MAXNV; i++) {
				const double *col = p + (size_t)(i < k.nv ? i : 0) * f.region_rows;
#pragma unroll
				for (int r = 0; r < 4; r++) {
					const int b = threadIdx.x + r * FIN_THREADS;
					v[i][r] = col[b < k.nblk ? b : 0];
				}
			}
#pragma unroll
			for (int i = 0; i < SEPK_MAXNV; i++) {
				double sum = 0.0, mx = 0.0;
#pragma unroll
				for (int r = 0; r < 4; r++) {
					const double x = (threadIdx.x + r * FIN_THREADS < k.nblk) ? v[i][r] : 0.0;
					sum += x;
					mx = fmax(mx, x);
				}
				a[i] = ((k.maxmask >> i) & 1u) ? mx : sum;
			}
		} else {
#pragma unroll
			for (int i = 0; i < SEPK_MAXNV; i++)
				a[i] = 0.0;
		}
		for (int b = threadIdx.x + 4 * FIN_THREADS; b < k.nblk; b += FIN_THREADS)
			for (int i = 0; i < k.nv; i++) {
				const double v = p[(size_t)i * f.region_rows + b];
				a[i] = ((k.maxmask >> i) & 1u) ? fmax(a[i], v) : a[i] + v;
			}
#pragma unroll
		for (int i = 0; i < SEPK_MAXNV; i++) {
			const bool mx = (k.maxmask >> i) & 1u;
			double x = a[i];
			for (int o = 32; o > 0; o >>= 1) {
				const double y = __shfl_xor(x, o, 64);
				x = mx ? fmax(x, y) : x + y;
			}
			if (lane == 0 && i < k.nv)
				wred[i][wid] = x;
		}
		__syncthreads();
		if (threadIdx.x == 0) {
			for (int i = 0; i < k.nv; i++) {
				const bool mx = (k.maxmask >> i) & 1u;
				double a = wred[i][0];
				for (int w = 1; w < FIN_THREADS / 64; w++)
					a = mx ? fmax(a, wred[i][w]) : a + wred[i][w];
				tot[i] = a;
			}
			fin_apply(k, tot, sys);
		}
		__syncthreads();
	}
}

double *sepk_region(sepdev *d, int region)
{
	return d->partials + (size_t)region * d->region_words;
}

int sepk_flush_tasks(sepdev *d)
{
	if (d->npending == 0)
		return 0;
	sepk_fin f;
	memset(&f, 0, sizeof f);
	for (int i = 0; i < d->npending; i++)
		f.t[i] = d->pending[i];
	f.ntask = d->npending;
	f.region_words = d->region_words;
	f.region_rows = d->region_rows;
	f.sys = d->sys;
	d->npending = 0;
	k_finish<<<1, FIN_THREADS, 0, d->st>>>(d->partials, f);
	SEPK_LAUNCH_CHECK();
	return 0;
}

int sepk_claim_region(sepdev *d, int region)
{
	for (int i = 0; i < d->npending; i++)
		if (d->pending[i].region == region)
			return sepk_flush_tasks(d);
	return 0;
}

int sepk_queue_task(sepdev *d, const sepk_task &t)
{
	if (t.nv > SEPK_MAXNV)
		return sepk_fail(SEP_GPU_EARG, "finisher: %d columns", t.nv);
	if (t.nblk > d->region_rows)
		return sepk_fail(SEP_GPU_EARG, "finisher: %d rows overflow the region", t.nblk);
	// a region can hold one pending result; finish the queue first if the
	// new task's region is taken or the queue is full
	bool clash = d->npending == SEPK_MAXTASK;
	for (int i = 0; i < d->npending; i++)
		clash = clash || d->pending[i].region == t.region;
	if (clash) {
		int rc = sepk_flush_tasks(d);
		if (rc)
			return rc;
	}
	d->pending[d->npending++] = t;
	return 0;
}

// Points a task at a return slot and consumes the slot's pending reset.
void sepk_task_ret(sepdev *d, sepk_task &t, int slot)
{
	t.ret = d->ret + (size_t)slot * SEPDEV_RET_WORDS;
	t.zero_first = d->ret_zero[slot];
	d->ret_zero[slot] = 0;
}

// ---- deferred Nose-Hoover ----------------------------------------------------

__global__ void k_nh_apply2(sepk_state s, long n, long npad, const double *alpha)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	const double al = *alpha;
	const double m = s.m[t];
	for (int k = 0; k < 3; k++)
		s.f[k * npad + t] -= al * m * s.v[k * npad + t];
}

// Applies a Nose-Hoover call that was waiting to be fused into a leapfrog,
// because something else is about to read or change f or v.
int sepk_flush_nh(sepdev *d)
{
	if (!d->nh_pending)
		return 0;
	d->nh_pending = 0;
	int rc = sepk_materialize_forces(d);
	if (rc)
		return rc;
	sepk_task k;
	memset(&k, 0, sizeof k);
	k.mode = SEPK_FIN_NH;
	k.region = SEPK_R_MISC;
	k.alpha = d->scal + d->nh_slot;
	k.Td = d->nh_Td;
	k.dt_tau2 = d->nh_dt_tau2;
	k.npart = (double)d->n;
	if ((rc = sepk_queue_task(d, k)) || (rc = sepk_flush_tasks(d)))
		return rc;
	k_nh_apply2<<<sepk_blocks(d->n), SEPK_BLOCK, 0, d->st>>>(d->s[0], d->n, d->npad,
								   d->scal + d->nh_slot);
	SEPK_LAUNCH_CHECK();
	return 0;
}
