// sepdev_list.hip -- cell binning and Verlet neighbour-list build.
//
// A rebuild is five kernels on the library stream:
//   k_bin        cell index per atom (reference formula or floor+wrap),
//                per-cell counts by atomics
//   k_scan       one workgroup: exclusive scan of the counts -> cell_start
//   k_scatter    atom -> slot inside its cell segment (atomic fill)
//   k_cell_sort  one wave per cell: order the segment by original atom
//                index, so the device order and every list are independent
//                of atomic timing
//   k_permute    move all per-atom state into the new (cell-sorted) order
//                in a spare copy; k_copy_back copies it home
//   k_build      one thread per atom: scan the stencil cells' contiguous
//                segments and append neighbours within the list cutoff
// With conditional = 1 each kernel returns at once unless the device
// rebuild flag is set, so the whole sequence can sit in a captured graph.
//
// The pair multiset equals the reference's half list of
// source/sepprfrc.c:419-513 (and its exclusion variants :517-700) taken in
// both directions: every half-shell offset s contributes j in cell(i)+s
// ("forward") and cell(i)-s ("backward"), with the reference's
// "m1 != m2 || j2 > j1" rule applied on original indices when the offset
// wraps back onto the home cell.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sepdev_impl.h"

__device__ __forceinline__ void sepk_set_error(sepdev_sys *sys, int code)
{
	atomicCAS(&sys->error, 0, code);
}

__device__ __forceinline__ int sepk_mod(int a, int n)
{
	int r = a % n;
	return r < 0 ? r + n : r;
}

// Linear cell index of a position, or -1 if it falls outside the grid.
__device__ __forceinline__ long sepk_cell_of(const sepk_pos &p, const sepdev_grid &g)
{
	const double q[3] = {p.x / g.l[0], p.y / g.l[1], p.z / g.l[2]};
	int c[3];
	for (int k = 0; k < 3; k++) {
		if (!(fabs(q[k]) < 1.0e9))
			return -1;
		if (g.ref_index)
			c[k] = (int)q[k]; // truncation, as the reference
		else
			c[k] = sepk_mod((int)floor(q[k]), g.n[k]);
	}
	const long lin = (long)c[0] + (long)c[1] * g.n[0] + (long)c[2] * g.n[0] * g.n[1];
	const long ncell = (long)g.n[0] * g.n[1] * g.n[2];
	if (lin < 0 || lin >= ncell)
		return -1;
	return lin;
}

__global__ void k_bin(const sepk_pos *pos, long n, sepdev_grid g, int *cellid,
		      int *count, sepdev_sys *sys, int conditional)
{
	if (conditional && sys->neighb_flag == 0)
		return;
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	long c = sepk_cell_of(pos[t], g);
	if (c < 0) {
		sepk_set_error(sys, SEP_GPU_ECELL);
		c = 0;
	}
	cellid[t] = (int)c;
	atomicAdd(&count[c], 1);
}

// Exclusive scan of count[0..ncell) into start[0..ncell]; also zeroes count
// and fill for the next rebuild.  One workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_scan(int *count, int *start, int *fill,
					       long ncell, sepdev_sys *sys,
					       int conditional)
{
	if (conditional && sys->neighb_flag == 0)
		return;
	__shared__ int wsum[16];
	__shared__ int carry;
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	if (threadIdx.x == 0)
		carry = 0;
	__syncthreads();
	for (long base = 0; base < ncell; base += 1024) {
		const long c = base + threadIdx.x;
		const int v = c < ncell ? count[c] : 0;
		int x = v; // inclusive wave scan
		for (int o = 1; o < 64; o <<= 1) {
			int y = __shfl_up(x, o, 64);
			if (lane >= o)
				x += y;
		}
		if (lane == 63)
			wsum[wid] = x;
		__syncthreads();
		int off = carry;
		for (int w = 0; w < wid; w++)
			off += wsum[w];
		if (c < ncell) {
			start[c] = off + x - v;
			count[c] = 0;
			fill[c] = 0;
		}
		__syncthreads();
		if (threadIdx.x == 1023) {
			int tot = carry;
			for (int w = 0; w < 16; w++)
				tot += wsum[w];
			carry = tot;
		}
		__syncthreads();
	}
	if (threadIdx.x == 0)
		start[ncell] = carry;
}

__global__ void k_scatter(const int *cellid, const int *start, int *fill,
			  int *order, long n, sepdev_sys *sys, int conditional)
{
	if (conditional && sys->neighb_flag == 0)
		return;
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	const int c = cellid[t];
	const int slot = start[c] + atomicAdd(&fill[c], 1);
	order[slot] = (int)t;
}

// One wave per cell: rank each member by its original index.
__global__ void k_cell_sort(const int *order, const int *start, const int *orig,
			    int *neworder, int *cellid_new, long ncell,
			    sepdev_sys *sys, int conditional)
{
	if (conditional && sys->neighb_flag == 0)
		return;
	const long c = ((long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
	const int lane = threadIdx.x & 63;
	if (c >= ncell)
		return;
	const int s = start[c], e = start[c + 1];
	for (int i = s + lane; i < e; i += 64) {
		const int mine = order[i];
		const int key = orig[mine];
		int rank = 0;
		for (int j = s; j < e; j++)
			rank += orig[order[j]] < key;
		neworder[s + rank] = mine;
		cellid_new[s + rank] = (int)c;
	}
}

__global__ void k_permute(sepk_state src, sepk_state dst, const int *neworder,
			  int *iperm, long n, long npad, sepdev_sys *sys,
			  int conditional, int has_partners)
{
	if (conditional && sys->neighb_flag == 0)
		return;
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	const long o = neworder[t];
	const int pending = sys->xn_pending;
	const sepk_pos p = src.pos[o];
	dst.pos[t] = p;
	const double xyz[3] = {p.x, p.y, p.z};
	for (int k = 0; k < 3; k++) {
		dst.v[k * npad + t] = src.v[k * npad + o];
		dst.f[k * npad + t] = src.f[k * npad + o];
		dst.a[k * npad + t] = src.a[k * npad + o];
		dst.xn[k * npad + t] = pending ? xyz[k] : src.xn[k * npad + o];
		dst.cn[k * npad + t] = pending ? 0 : src.cn[k * npad + o];
		dst.cr[k * npad + t] = src.cr[k * npad + o];
	}
	dst.m[t] = src.m[o];
	dst.type[t] = src.type[o];
	dst.mol[t] = src.mol[o];
	const int og = src.orig[o];
	dst.orig[t] = og;
	iperm[og] = (int)t;
	if (has_partners)
		for (int k = 0; k < 40; k++)
			dst.partners[k * npad + t] = src.partners[k * npad + o];
}

// Copies the permuted spare state back into the current buffers, so that
// the buffers every other kernel uses never change address.
__global__ void k_copy_back(sepk_state src, sepk_state dst, long n, long npad,
			    sepdev_sys *sys, int conditional, int has_partners)
{
	if (conditional && sys->neighb_flag == 0)
		return;
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	dst.pos[t] = src.pos[t];
	for (int k = 0; k < 3; k++) {
		dst.v[k * npad + t] = src.v[k * npad + t];
		dst.f[k * npad + t] = src.f[k * npad + t];
		dst.a[k * npad + t] = src.a[k * npad + t];
		dst.xn[k * npad + t] = src.xn[k * npad + t];
		dst.cn[k * npad + t] = src.cn[k * npad + t];
		dst.cr[k * npad + t] = src.cr[k * npad + t];
	}
	dst.m[t] = src.m[t];
	dst.type[t] = src.type[t];
	dst.mol[t] = src.mol[t];
	dst.orig[t] = src.orig[t];
	if (has_partners)
		for (int k = 0; k < 40; k++)
			dst.partners[k * npad + t] = src.partners[k * npad + t];
}

// Partner scan of source/sepprfrc.c:703-731: stops at the first -1 in
// either atom's array.  Partner values are original indices.
__device__ __forceinline__ int sepk_share(const int *P, long npad, long ti, long tj,
					  int base, int cap, int oi, int oj)
{
	for (int k = 0; k < cap; k++) {
		const int a = P[(base + k) * npad + ti], b = P[(base + k) * npad + tj];
		if (a == -1 || b == -1)
			return 0;
		if (a == oj || b == oi)
			return 1;
	}
	return 0;
}

__global__ __launch_bounds__(SEPK_BLOCK) void k_build(
	sepk_state s, const int *cellid, const int *start, sepdev_grid g,
	long n, long npad, int cap, int *nbr, int *cnt, double *partial,
	unsigned *counter, sepdev_sys *sys, int conditional)
{
	if (conditional && sys->neighb_flag == 0)
		return;
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double mx = 0.0;
	if (t < n) {
		const int c = cellid[t];
		const int cx = c % g.n[0], cy = (c / g.n[0]) % g.n[1], cz = c / (g.n[0] * g.n[1]);
		const sepk_pos pi = s.pos[t];
		const int oi = s.orig[t];
		const int mi = s.mol[t];
		int m = 0;
		for (int o = 0; o < g.noff; o++) {
			const int tx = sepk_mod(cx + g.off[o][0], g.n[0]);
			const int ty = sepk_mod(cy + g.off[o][1], g.n[1]);
			const int tz = sepk_mod(cz + g.off[o][2], g.n[2]);
			const int c2 = tx + ty * g.n[0] + tz * g.n[0] * g.n[1];
			const int rule = g.off[o][3];
			const bool same = c2 == c;
			const int js = start[c2], je = start[c2 + 1];
			for (int j = js; j < je; j++) {
				if (same) {
					if (rule == SEPDEV_RULE_NE) {
						if (j == t)
							continue;
					} else {
						const int oj = s.orig[j];
						if (rule == SEPDEV_RULE_GT ? oj <= oi : oj >= oi)
							continue;
					}
				}
				if (g.opt == 3 /* SEP_EXCL_SAME_MOL */ && mi != -1 && mi == s.mol[j])
					continue;
				const sepk_pos pj = s.pos[j];
				const double dx = sepk_wrap(pi.x - pj.x, g.L[0]);
				const double dy = sepk_wrap(pi.y - pj.y, g.L[1]);
				const double dz = sepk_wrap(pi.z - pj.z, g.L[2]);
				const double r2 = dx * dx + dy * dy + dz * dz;
				if (!(r2 < g.rlist2))
					continue;
				if (g.opt == 2 /* SEP_EXCL_BONDED */) {
					const int oj = s.orig[j];
					int sh = sepk_share(s.partners, npad, t, j, 0, 10, oi, oj);
					if (g.bonded_all)
						sh += sepk_share(s.partners, npad, t, j, 10, 10, oi, oj) +
						      sepk_share(s.partners, npad, t, j, 20, 20, oi, oj);
					// rule 2 is the Coulomb brute-force test of
					// source/sepcoulomb.c:44, which excludes only
					// when exactly one kind of partnership holds
					if (g.bonded_all == 2 ? sh == 1 : sh != 0)
						continue;
				}
				if (m < cap)
					nbr[(long)m * npad + t] = j;
				m++;
			}
		}
		if (m > cap)
			sepk_set_error(sys, SEP_GPU_ENEIGHB);
		cnt[t] = m < cap ? m : cap;
		mx = (double)m;
	}
	double mxb = sepk_block_max(mx, partial + blockIdx.x);
	(void)mxb;
	if (sepk_last_block(counter)) {
		if (threadIdx.x == 0) {
			double best = 0.0;
			for (unsigned b = 0; b < gridDim.x; b++)
				best = fmax(best, sepk_ld(partial + b));
			sys->max_full = (int)best;
			sys->neighb_flag = 0;
			sys->xn_pending = 0;
			*counter = 0;
		}
	}
}

// ---- host side --------------------------------------------------------------

extern "C" int sepdev_reserve_list(sepdev *d, int cap)
{
	if (d->capturing)
		return sepk_fail(SEP_GPU_ESTATE, "list resize during graph capture");
	cap = (cap + 31) / 32 * 32;
	if (d->nbr && cap <= d->cap)
		return 0;
	SEPK_CHECK(hipStreamSynchronize(d->st));
	if (d->nbr)
		(void)hipFree(d->nbr);
	d->nbr = nullptr;
	SEPK_CHECK(hipMalloc(&d->nbr, (size_t)cap * d->npad * sizeof(int)));
	SEPK_CHECK(hipMemsetAsync(d->nbr_count, 0, d->npad * sizeof(int), d->st));
	d->cap = cap;
	d->list_built = 0;
	return 0;
}

__global__ void k_sys_reset_max(sepdev_sys *sys) { sys->max_dist2 = 0.0; }

extern "C" int sepdev_reset_max_dist(sepdev *d)
{
	k_sys_reset_max<<<1, 1, 0, d->st>>>(d->sys);
	SEPK_LAUNCH_CHECK();
	return 0;
}

extern "C" int sepdev_list_capacity(sepdev *d) { return d->nbr ? d->cap : 0; }

extern "C" int sepdev_build_list(sepdev *d, const sepdev_grid *gp, int conditional)
{
	const sepdev_grid &g = *gp;
	if (g.noff <= 0 || g.noff > SEPDEV_MAX_OFF)
		return sepk_fail(SEP_GPU_EARG, "build_list: %d offsets", g.noff);
	for (int k = 0; k < 3; k++)
		if (g.n[k] < 1)
			return sepk_fail(SEP_GPU_EARG, "build_list: %d cells on axis %d", g.n[k], k);
	const long ncell = (long)g.n[0] * g.n[1] * g.n[2];
	if (ncell > (1L << 28))
		return sepk_fail(SEP_GPU_EARG, "build_list: %ld cells", ncell);
	if (g.opt == 2 && !d->s[0].partners)
		return sepk_fail(SEP_GPU_ESTATE, "SEP_EXCL_BONDED needs bonded partners");

	// size buffers on the host side of any capture
	if (ncell > d->ncell_cap) {
		if (d->capturing)
			return sepk_fail(SEP_GPU_ESTATE, "cell grid grew during graph capture");
		(void)hipFree(d->cell_count);
		(void)hipFree(d->cell_start);
		(void)hipFree(d->cell_fill);
		SEPK_CHECK(hipMalloc(&d->cell_count, ncell * sizeof(int)));
		SEPK_CHECK(hipMalloc(&d->cell_start, (ncell + 1) * sizeof(int)));
		SEPK_CHECK(hipMalloc(&d->cell_fill, ncell * sizeof(int)));
		SEPK_CHECK(hipMemsetAsync(d->cell_count, 0, ncell * sizeof(int), d->st));
		SEPK_CHECK(hipMemsetAsync(d->cell_fill, 0, ncell * sizeof(int), d->st));
		d->ncell_cap = ncell;
	}
	if (!d->nbr) {
		// capacity from the mean density with a generous margin
		const double V = g.L[0] * g.L[1] * g.L[2];
		const double rl = sqrt(g.rlist2);
		double expect = 4.18879020478639 * rl * rl * rl * (double)d->n / V;
		if (!(expect < 1e7))
			expect = 1e7;
		long cap = (long)(2.0 * expect) + 64;
		if (cap > d->n * 4 + 64)
			cap = d->n * 4 + 64;
		cap = (cap + 31) / 32 * 32;
		if (d->capturing)
			return sepk_fail(SEP_GPU_ESTATE, "list allocated during graph capture");
		SEPK_CHECK(hipMalloc(&d->nbr, (size_t)cap * d->npad * sizeof(int)));
		d->cap = (int)cap;
	}

	const unsigned nb = sepk_blocks(d->n);
	sepk_state &src = d->s[0]; // current state, fixed addresses
	sepk_state &dst = d->s[1]; // spare: receives the permuted order
	k_bin<<<nb, SEPK_BLOCK, 0, d->st>>>(src.pos, d->n, g, d->cellid_old, d->cell_count, d->sys, conditional);
	k_scan<<<1, 1024, 0, d->st>>>(d->cell_count, d->cell_start, d->cell_fill, ncell, d->sys, conditional);
	k_scatter<<<nb, SEPK_BLOCK, 0, d->st>>>(d->cellid_old, d->cell_start, d->cell_fill, d->order_tmp, d->n, d->sys, conditional);
	k_cell_sort<<<(unsigned)((ncell * 64 + SEPK_BLOCK - 1) / SEPK_BLOCK), SEPK_BLOCK, 0, d->st>>>(
		d->order_tmp, d->cell_start, src.orig, d->neworder, d->cellid_new, ncell, d->sys, conditional);
	k_permute<<<nb, SEPK_BLOCK, 0, d->st>>>(src, dst, d->neworder, d->iperm, d->n, d->npad, d->sys,
						 conditional, src.partners != nullptr);
	k_copy_back<<<nb, SEPK_BLOCK, 0, d->st>>>(dst, src, d->n, d->npad, d->sys, conditional,
						   src.partners != nullptr);
	SEPK_LAUNCH_CHECK();
	k_build<<<nb, SEPK_BLOCK, 0, d->st>>>(src, d->cellid_new, d->cell_start, g, d->n, d->npad, d->cap,
					      d->nbr, d->nbr_count, d->partials, d->counters + SEPK_CTR_MISC,
					      d->sys, conditional);
	SEPK_LAUNCH_CHECK();
	d->grid = g;
	d->list_built = 1;
	return 0;
}
