// sepdev_reduce.hip -- the finisher that turns per-block partial
// rows into the energies, stress tensors, thermostat state and rebuild
// bookkeeping of a step, and the deferred-work helpers around it.
//
// Kernels leave one row per block in a region of the partials buffer and
// queue a task; queued tasks run together in the finisher's launches at the
// next flush (each leapfrog, each host read, the end of a graph capture).
// Every reduction is fixed-order, so repeated runs give identical bits.
#include <string.h>

#include "sepdev_impl.h"

#define FIN_THREADS (64 * SEPK_MAXTASK)

// ret, alpha and sys point at the finisher's LDS copies of the task's
// return slot, thermostat variable and the system scalars.
// Row-major entry m of a symmetric 3x3 tensor stored as its six distinct
// entries (xx, xy, xz, yy, yz, zz): the index of that entry.
__device__ static int fin_sym(int m)
{
	return (int)((0x542431210ull >> (4 * m)) & 15ull);
}

// One word w of a task's return slot: what the task adds to (or sets in)
// that word, given the word's value v before the task.
__device__ static double fin_word(const sepk_task &k, const double *tot, int w, double v)
{
	if (k.zero_first)
		v = 0.0;
	const int mp = w - SEPDEV_RET_POTP, mk = w - SEPDEV_RET_KINP, mb = w - SEPDEV_RET_BONDP;
	switch (k.mode) {
	case SEPK_FIN_PAIR: {
		// each pair was seen from both ends: halve energy and stress
		const double e = 0.5 * tot[0];
		if (w == SEPDEV_RET_EPOT)
			v = k.epot_assign ? e : v + e;
		if (w == SEPDEV_RET_ECOUL && k.also_ecoul)
			v += e;
		if (mp >= 0 && mp < 9)
			v += 0.5 * tot[1 + fin_sym(mp)];
		break;
	}
	case SEPK_FIN_BONDED:
		if (w == SEPDEV_RET_EPOT)
			v += tot[0];
		if (k.with_stress && mp >= 0 && mp < 9)
			v += 1.0 * tot[1 + fin_sym(mp)];
		if (k.with_stress && mb >= 0 && mb < 9)
			v += 1.0 * tot[1 + fin_sym(mb)];
		break;
	case SEPK_FIN_LF:
		if (w == SEPDEV_RET_EKIN)
			v += 0.5 * tot[0];
		if (mk >= 0 && mk < 9)
			v += 1.0 * tot[1 + fin_sym(mk)];
		break;
	default:
		break;
	}
	return v;
}

// A task's effect on the system scalars and its thermostat variable.
__device__ static void fin_sys(const sepk_task &k, const double *tot, double *alpha,
			       sepdev_sys *sys)
{
	switch (k.mode) {
	case SEPK_FIN_NH: {
		// nblk == 0: the sum of m v^2 is already in sys->sumv2
		const double sum = k.nblk > 0 ? tot[0] : sys->sumv2;
		sys->sumv2 = sum;
		*alpha = sepk_nh_alpha(*alpha, sum, k.npart, k.Td, k.dt_tau2);
		break;
	}
	case SEPK_FIN_LF: {
		if (k.nh_fused)
			*alpha = sepk_nh_alpha(*alpha, sys->sumv2, k.npart, k.Td, k.dt_tau2);
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
		// sys->max_full was set by the build's atomic maximum
		sys->neighb_flag = 0;
		sys->xn_pending = 0; // applied by the permutation
		sys->unclean = 0;    // judged afresh by the next binning
		break;
	default:
		break;
	}
}

// Two launches.  k_fin_rows: SEPK_FIN_BLOCKS workgroups of one wave per
// queued task (at least two); wave w of every workgroup takes task w: a
// contiguous chunk of its rows, one lane per row (SEPK_MAXNV columns loaded
// at once with clamped addresses, all tasks' loads in flight together),
// reduced in a fixed order -- lane-strided rows, then a wave butterfly --
// to one value per column.  k_fin_apply, one workgroup, combines the
// SEPK_FIN_BLOCKS values of each column in block order (wave w again for
// task w) and applies the tasks in queue order: lane w of the first wave
// takes word w of every return slot, lane 0 of the second wave the system
// scalars and thermostat variables (no task's words depend on them).  The kernel boundary makes
// the first launch's values visible to the second, where a last-workgroup
// handshake in one launch would need a device-scope fence in every
// workgroup (a write-back of its XCD's L2).  Neither step depends on
// timing, so repeated runs give identical bits.
__device__ static double fin_comb(bool mx, double a, double b)
{
	return mx ? fmax(a, b) : a + b;
}

static_assert(FIN_THREADS == 64 * SEPK_MAXTASK, "one finisher wave per task");
static_assert(sizeof(sepdev_sys) <= 64 * sizeof(int), "the second wave copies the scalars");

__global__ __launch_bounds__(FIN_THREADS) void k_fin_rows(const double *partials, sepk_fin f)
{
	const int lane = threadIdx.x & 63;
	const int ti = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); // this wave's task
	const int G = gridDim.x;
	if (ti < f.ntask) {
		const sepk_task &k = f.t[ti];
		const double *p = partials + (size_t)k.region * f.region_words;
		const long chunk = (k.nblk + G - 1) / G;
		const long r0 = (long)blockIdx.x * chunk;
		const long r1 = min(r0 + chunk, (long)k.nblk);
		double a[SEPK_MAXNV];
#pragma unroll
		for (int i = 0; i < SEPK_MAXNV; i++)
			a[i] = 0.0;
		for (long r = r0 + lane; r < r1; r += 64) {
			double v[SEPK_MAXNV];
#pragma unroll
			for (int i = 0; i < SEPK_MAXNV; i++)
				v[i] = p[(size_t)(i < k.nv ? i : 0) * f.region_rows + r];
#pragma unroll
			for (int i = 0; i < SEPK_MAXNV; i++)
				a[i] = fin_comb((k.maxmask >> i) & 1u, a[i], v[i]);
		}
		// butterfly step by step over all columns at once (one wait a step)
#pragma unroll
		for (int o = 32; o > 0; o >>= 1) {
			double y[SEPK_MAXNV];
#pragma unroll
			for (int i = 0; i < SEPK_MAXNV; i++)
				y[i] = __shfl_xor(a[i], o, 64);
#pragma unroll
			for (int i = 0; i < SEPK_MAXNV; i++)
				a[i] = fin_comb((k.maxmask >> i) & 1u, a[i], y[i]);
		}
#pragma unroll
		for (int i = 0; i < SEPK_MAXNV; i++)
			if (lane == 0 && i < k.nv)
				f.l2[((size_t)ti * SEPK_MAXNV + i) * G + blockIdx.x] = a[i];
	}
}

__global__ __launch_bounds__(FIN_THREADS) void k_fin_apply(sepk_fin f, int G)
{
	__shared__ double tot[SEPK_MAXTASK][SEPK_MAXNV];
	__shared__ double sret[SEPK_MAXTASK][SEPDEV_RET_WORDS];
	__shared__ double salpha[SEPK_MAXTASK];
	__shared__ sepdev_sys ssys;
	const int lane = threadIdx.x & 63;
	const int ti = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); // this wave's task
	// The tasks read and update the system scalars, their return slots and
	// thermostat variables: one parallel load of each into LDS (issued with
	// the loads of the row results), the tasks applied there in queue order,
	// one parallel store back.
	const signed char *rmap = f.rmap, *amap = f.amap; // first task sharing the pointer
	const int nsys = (int)(sizeof(sepdev_sys) / sizeof(int));
	for (int w = threadIdx.x; w < f.ntask * SEPDEV_RET_WORDS; w += blockDim.x) {
		const int u = w / SEPDEV_RET_WORDS, word = w % SEPDEV_RET_WORDS;
		if (rmap[u] == u)
			sret[u][word] = f.t[u].ret[word];
	}
	if (threadIdx.x < f.ntask && amap[threadIdx.x] == (int)threadIdx.x)
		salpha[threadIdx.x] = *f.t[threadIdx.x].alpha;
	if (threadIdx.x >= 64 && threadIdx.x < 64 + nsys)
		((int *)&ssys)[threadIdx.x - 64] = ((const int *)f.sys)[threadIdx.x - 64];
	const unsigned long long ring_at =
		f.ring_task >= 0 && ti == 0 ? *f.t[f.ring_task].ring_n : 0ull;
	if (ti < f.ntask) {
		const sepk_task &k = f.t[ti];
		double x[SEPK_MAXNV];
#pragma unroll
		for (int i = 0; i < SEPK_MAXNV; i++) {
			const double *col = f.l2 + ((size_t)ti * SEPK_MAXNV + (i < k.nv ? i : 0)) * G;
			x[i] = lane < G ? col[lane] : 0.0;
		}
#pragma unroll
		for (int o = 32; o > 0; o >>= 1) {
			double y[SEPK_MAXNV];
#pragma unroll
			for (int i = 0; i < SEPK_MAXNV; i++)
				y[i] = __shfl_xor(x[i], o, 64);
#pragma unroll
			for (int i = 0; i < SEPK_MAXNV; i++)
				x[i] = fin_comb((k.maxmask >> i) & 1u, x[i], y[i]);
		}
#pragma unroll
		for (int i = 0; i < SEPK_MAXNV; i++)
			if (lane == 0)
				tot[ti][i] = x[i];
	}
	__syncthreads();
	// after a device error the step changed nothing, and neither does this
	const bool frozen = ssys.error != 0;
	if (threadIdx.x < SEPDEV_RET_WORDS && !frozen) {
		for (int u = 0; u < f.ntask; u++)
			if (rmap[u] >= 0)
				sret[rmap[u]][threadIdx.x] =
					fin_word(f.t[u], tot[u], threadIdx.x, sret[rmap[u]][threadIdx.x]);
	} else if (threadIdx.x == 64 && !frozen) {
		for (int u = 0; u < f.ntask; u++) {
			const sepk_task &k = f.t[u];
			// a conditional build that did not run leaves nothing to finish
			if (k.mode == SEPK_FIN_BUILD && k.conditional && ssys.neighb_flag == 0)
				continue;
			fin_sys(k, tot[u], amap[u] >= 0 ? &salpha[amap[u]] : nullptr, &ssys);
		}
	}
	__syncthreads();
	for (int w = threadIdx.x; w < f.ntask * SEPDEV_RET_WORDS; w += blockDim.x) {
		const int u = w / SEPDEV_RET_WORDS, word = w % SEPDEV_RET_WORDS;
		if (rmap[u] == u)
			f.t[u].ret[word] = sret[u][word];
	}
	// the per-step record: the leapfrog's return slot and the new tnow; the
	// count was read at the start by this same wave, which alone stores it
	if (f.ring_task >= 0 && !frozen && threadIdx.x <= SEPDEV_RET_WORDS) {
		const sepk_task &k = f.t[f.ring_task];
		double *rec = k.ring + (size_t)(ring_at % (unsigned long long)k.ring_cap) * SEPDEV_RING_WORDS;
		rec[threadIdx.x] = threadIdx.x < SEPDEV_RET_WORDS ? sret[rmap[f.ring_task]][threadIdx.x]
								  : ssys.tnow;
		if (threadIdx.x == 0)
			*k.ring_n = ring_at + 1;
	}
	if (threadIdx.x < f.ntask && amap[threadIdx.x] == (int)threadIdx.x)
		*f.t[threadIdx.x].alpha = salpha[threadIdx.x];
	if (threadIdx.x >= 64 && threadIdx.x < 64 + nsys)
		((int *)f.sys)[threadIdx.x - 64] = ((const int *)&ssys)[threadIdx.x - 64];
}

double *sepk_region(sepdev *d, int region)
{
	return d->partials + (size_t)region * d->region_words;
}

int sepk_flush_tasks(sepdev *d)
{
	// a recorded pair call queues its task when it is launched
	int rc = sepk_flush_pair(d);
	if (rc)
		return rc;
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
	f.l2 = d->fin_l2;
	for (int u = 0; u < f.ntask; u++) {
		f.rmap[u] = f.amap[u] = -1;
		for (int v = u; v >= 0; v--) {
			if (f.t[u].ret && f.t[v].ret == f.t[u].ret)
				f.rmap[u] = (signed char)v;
			if (f.t[u].alpha && f.t[v].alpha == f.t[u].alpha)
				f.amap[u] = (signed char)v;
		}
	}
	// the leapfrog's per-step record (a leapfrog task always ends a flush,
	// so there is at most one)
	f.ring_task = -1;
	for (int u = 0; u < f.ntask; u++)
		if (f.t[u].mode == SEPK_FIN_LF && f.t[u].ring && f.rmap[u] >= 0)
			f.ring_task = u;
	d->npending = 0;
	// one wave per task, at least two (the system scalars are copied by the
	// second wave)
	const unsigned threads = 64u * (unsigned)(f.ntask > 2 ? f.ntask : 2);
	k_fin_rows<<<SEPK_FIN_BLOCKS, threads, 0, d->st>>>(d->partials, f);
	k_fin_apply<<<1, threads, 0, d->st>>>(f, SEPK_FIN_BLOCKS);
	SEPK_LAUNCH_CHECK();
	return 0;
}

int sepk_term_region(sepdev *d, int *region)
{
	int rc = sepk_flush_pair(d);
	if (rc)
		return rc;
	for (int r = SEPK_R_PAIR; r < SEPK_R_PAIR + SEPK_NTERM; r++) {
		bool held = false;
		for (int i = 0; i < d->npending; i++)
			held = held || d->pending[i].region == r;
		if (!held) {
			*region = r;
			return 0;
		}
	}
	// the finisher reads the term regions: the side stream waits for it
	d->fork_ready = 0;
	*region = SEPK_R_PAIR;
	return sepk_flush_tasks(d);
}

int sepk_fork_point(sepdev *d)
{
	if (!d->fork_armed || d->fork_ready)
		return 0;
	SEPK_CHECK(hipEventRecord(d->fork_ev, d->st));
	d->fork_ready = 1;
	d->fork_used = 0;
	return 0;
}

int sepk_term_stream(sepdev *d, unsigned family, hipStream_t *ks)
{
	*ks = d->st;
	if (!d->fork_armed || !d->fork_ready)
		return 0;
	// a second call of the family since the fork rewrites the family's
	// per-term slots: it waits for the gather of the first
	if (d->fork_used & family) {
		SEPK_CHECK(hipEventRecord(d->fork_ev, d->st));
		d->fork_used = 0;
	}
	d->fork_used |= family;
	SEPK_CHECK(hipStreamWaitEvent(d->side, d->fork_ev, 0));
	*ks = d->side;
	return 0;
}

int sepk_term_join(sepdev *d, hipStream_t ks)
{
	if (ks == d->st)
		return 0;
	SEPK_CHECK(hipEventRecord(d->join_ev, ks));
	SEPK_CHECK(hipStreamWaitEvent(d->st, d->join_ev, 0));
	return 0;
}

extern "C" int sepdev_concurrent_terms(sepdev *d, int on)
{
	if (on && !d->side) {
		SEPK_CHECK(hipStreamCreateWithFlags(&d->side, hipStreamNonBlocking));
		SEPK_CHECK(hipEventCreateWithFlags(&d->fork_ev, hipEventDisableTiming));
		SEPK_CHECK(hipEventCreateWithFlags(&d->join_ev, hipEventDisableTiming));
	}
	d->fork_armed = on != 0;
	d->fork_ready = 0;
	return 0;
}

int sepk_claim_region(sepdev *d, int region)
{
	int rc = sepk_flush_pair(d);
	if (rc)
		return rc;
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
	if (t >= n || *s.err)
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

// ---- per-step record -----------------------------------------------------------

extern "C" int sepdev_ring_alloc(sepdev *d, long capacity)
{
	if (d->capturing)
		return sepk_fail(SEP_GPU_ESTATE, "step record changed during graph capture");
	if (capacity < 0)
		return sepk_fail(SEP_GPU_EARG, "step record capacity %ld", capacity);
	// finish queued work that may still write the old record, and drop a
	// captured graph that holds its address
	int rc = sepk_flush_tasks(d);
	if (rc)
		return rc;
	SEPK_CHECK(hipStreamSynchronize(d->st));
	sepdev_graph_drop(d);
	(void)hipFree(d->ring);
	(void)hipFree(d->ring_n);
	d->ring = nullptr;
	d->ring_n = nullptr;
	d->ring_cap = 0;
	if (capacity == 0)
		return 0;
	SEPK_CHECK(hipMalloc(&d->ring, (size_t)capacity * SEPDEV_RING_WORDS * sizeof(double)));
	SEPK_CHECK(hipMalloc(&d->ring_n, sizeof(unsigned long long)));
	SEPK_CHECK(hipMemsetAsync(d->ring_n, 0, sizeof(unsigned long long), d->st));
	d->ring_cap = capacity;
	return 0;
}

extern "C" int sepdev_ring_read(sepdev *d, double *words, long max, long *n, long *lost)
{
	*n = 0;
	*lost = 0;
	if (!d->ring)
		return sepk_fail(SEP_GPU_ESTATE, "no step record (sep_gpu_record_steps)");
	if (max < d->ring_cap)
		return sepk_fail(SEP_GPU_EARG, "step record read of %ld < capacity %ld", max,
				 d->ring_cap);
	int rc = sepk_flush_tasks(d);
	if (rc)
		return rc;
	unsigned long long cnt = 0;
	SEPK_CHECK(hipMemcpyAsync(&cnt, d->ring_n, sizeof cnt, hipMemcpyDeviceToHost, d->st));
	SEPK_CHECK(hipStreamSynchronize(d->st));
	const unsigned long long cap = (unsigned long long)d->ring_cap;
	const unsigned long long held = cnt < cap ? cnt : cap;
	const size_t rec = SEPDEV_RING_WORDS * sizeof(double);
	// oldest first: entries (cnt - held) .. cnt - 1, modulo the capacity
	const unsigned long long first = (cnt - held) % cap;
	const unsigned long long tail = held < cap - first ? held : cap - first;
	if (tail)
		SEPK_CHECK(hipMemcpyAsync(words, d->ring + first * SEPDEV_RING_WORDS, tail * rec,
					  hipMemcpyDeviceToHost, d->st));
	if (held > tail)
		SEPK_CHECK(hipMemcpyAsync(words + tail * SEPDEV_RING_WORDS, d->ring,
					  (held - tail) * rec, hipMemcpyDeviceToHost, d->st));
	SEPK_CHECK(hipMemsetAsync(d->ring_n, 0, sizeof(unsigned long long), d->st));
	SEPK_CHECK(hipStreamSynchronize(d->st));
	*n = (long)held;
	*lost = (long)(cnt - held);
	return 0;
}
