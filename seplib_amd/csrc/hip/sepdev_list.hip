// sepdev_list.hip -- cell binning and Verlet neighbour-list build.
//
// A rebuild is a short kernel sequence on the library stream:
//   k_bin          cell index per atom (reference formula or floor+wrap),
//                  per-cell counts by atomics
//   k_scan         one workgroup: exclusive scan of the counts -> cell_start
//   k_scatter      atom -> slot inside its cell segment (atomic fill)
//   k_sort_permute one wave per cell: order the segment by original atom
//                  index (so the device order and every list are independent
//                  of atomic timing) and move each member's state to its new
//                  cell-sorted index in the spare copy
//   k_build        SEPK_SUB lanes per atom: copy the atom's state home and
//                  scan the stencil cells' contiguous segments, appending
//                  neighbours within the list cutoff to the lane's own slots
// and a BUILD task for the step's finisher.  With conditional = 1 every
// kernel returns at once unless the device rebuild flag is set, so the
// sequence can sit in a captured graph.
//
// The pair multiset equals the reference's half list of
// source/sepprfrc.c:419-513 (and its exclusion variants :517-700) taken in
// both directions: every half-shell offset s contributes j in cell(i)+s
// ("forward") and cell(i)-s ("backward"), with the reference's
// "m1 != m2 || j2 > j1" rule applied on original indices when the offset
// wraps back onto the home cell.  r2 is formed as the reference does
// (difference, single-image wrap, x2 + y2 + z2 in that order, each
// operation rounded on its own), so the set is identical bit for bit.
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
			c[k] = (int)q[k]; // truncation, as source/sepprfrc.c:403-405
		else
			c[k] = sepk_mod((int)floor(q[k]), g.n[k]);
	}
	const long lin = (long)c[0] + (long)c[1] * g.n[0] + (long)c[2] * g.n[0] * g.n[1];
	const long ncell = (long)g.n[0] * g.n[1] * g.n[2];
	if (lin < 0 || lin >= ncell)
		return -1;
	return lin;
}

#define SEPK_GRID_CAP 2048 // grid-stride launches: at most this many blocks

__global__ void k_bin(const sepk_pos *pos, long n, sepdev_grid g, int *cellid,
		      int *count, sepdev_sys *sys, int conditional)
{
	if (conditional && sys->neighb_flag == 0)
		return;
	for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
	     t += (long)gridDim.x * blockDim.x) {
		const sepk_pos p = pos[t];
		long c = sepk_cell_of(p, g);
		if (c < 0) {
			sepk_set_error(sys, SEP_GPU_ECELL);
			c = 0;
		}
		// an atom outside [0, n*l) on some axis (allowed: it is binned by
		// truncation or wrapping) rules out the per-cell image shifts
		const double q[3] = {p.x / g.l[0], p.y / g.l[1], p.z / g.l[2]};
		if (!(q[0] >= 0.0 && q[0] < g.n[0] && q[1] >= 0.0 && q[1] < g.n[1] &&
		      q[2] >= 0.0 && q[2] < g.n[2]))
			sys->unclean = 1;
		cellid[t] = (int)c;
		atomicAdd(&count[c], 1);
	}
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
	for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
	     t += (long)gridDim.x * blockDim.x) {
		const int c = cellid[t];
		const int slot = start[c] + atomicAdd(&fill[c], 1);
		order[slot] = (int)t;
	}
}

// One wave per cell (grid-stride over cells): rank each member by its
// original index and move its state to start + rank in the spare copy.
__global__ void k_sort_permute(const int *order, const int *start, sepk_state src,
			       sepk_state dst, int *iperm, int *cellid_new, long ncell,
			       long npad, sepdev_sys *sys, int conditional,
			       int has_partners)
{
	if (conditional && sys->neighb_flag == 0)
		return;
	const int pending = sys->xn_pending;
	const int lane = threadIdx.x & 63;
	const long waves = (long)gridDim.x * (blockDim.x / 64);
	for (long c = ((long)blockIdx.x * blockDim.x + threadIdx.x) / 64; c < ncell; c += waves) {
		const int s = start[c], e = start[c + 1];
		for (int i = s + lane; i < e; i += 64) {
			const long o = order[i];
			const int key = src.orig[o];
			int rank = 0;
			for (int j = s; j < e; j++)
				rank += src.orig[order[j]] < key;
			const long t = s + rank;
			cellid_new[t] = (int)c;
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
			dst.orig[t] = key;
			iperm[key] = (int)t;
			if (has_partners)
				for (int k = 0; k < 40; k++)
					dst.partners[k * npad + t] = src.partners[k * npad + o];
		}
	}
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

// Copies atom t of the permuted spare state back to the current buffers,
// whose addresses every other kernel (and a captured graph) holds.
__device__ __forceinline__ void sepk_copy_home(const sepk_state &src, const sepk_state &dst,
					       long t, long npad, int has_partners)
{
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

// SEPK_SUB lanes per atom scan each stencil cell's contiguous segment
// together, SEPK_SUB consecutive candidates per round; a wave ballot gives
// every accepting lane its position, so each atom's list comes out in
// stencil-cell order and, within a cell, in device order -- the same order
// whatever the lane count.  Entry k goes to lane k % SEPK_SUB's slots.
//
// SHIFT: with at least 5 cells on every axis and every atom inside its own
// cell, |x_i - x_j| < 2l <= 0.4 L for a target cell reached without
// crossing the box and > L - 2l >= 0.6 L for one reached across it, so the
// reference's single-image wrap (subtract or add L exactly when |d| > L/2)
// is the constant shift of the target cell; d + shift rounds exactly like
// its d - L / d + L.  Otherwise the wrap is evaluated per candidate.
template <bool SHIFT>
__device__ __forceinline__ void sepk_build_atom(
	const sepk_state &s, long t, int sub, int gshift, const int *cellid,
	const int *start, const sepdev_grid &g, long npad, int cap, int *nbr, int &m_out)
{
	const unsigned long long below = (1ull << sub) - 1ull;
	const int c = cellid[t];
	const int cx = c % g.n[0], cy = (c / g.n[0]) % g.n[1], cz = c / (g.n[0] * g.n[1]);
	const sepk_pos pi = s.pos[t];
	const int oi = s.orig[t];
	const int mi = s.mol[t];
	int m = 0; // entries so far; equal in the atom's lanes
	for (int o = 0; o < g.noff; o++) {
		const int ux = cx + g.off[o][0], uy = cy + g.off[o][1], uz = cz + g.off[o][2];
		const int tx = sepk_mod(ux, g.n[0]);
		const int ty = sepk_mod(uy, g.n[1]);
		const int tz = sepk_mod(uz, g.n[2]);
		const double sx = ux >= g.n[0] ? -g.L[0] : ux < 0 ? g.L[0] : 0.0;
		const double sy = uy >= g.n[1] ? -g.L[1] : uy < 0 ? g.L[1] : 0.0;
		const double sz = uz >= g.n[2] ? -g.L[2] : uz < 0 ? g.L[2] : 0.0;
		const int c2 = tx + ty * g.n[0] + tz * g.n[0] * g.n[1];
		const int rule = g.off[o][3];
		const bool same = c2 == c;
		const int js = start[c2], je = start[c2 + 1];
		for (int base = js; base < je; base += SEPK_SUB) {
			const int j = base + sub;
			bool take = j < je;
			if (same && take) {
				if (rule == SEPDEV_RULE_NE) {
					take = j != t;
				} else {
					const int oj = s.orig[j];
					take = rule == SEPDEV_RULE_GT ? oj > oi : oj < oi;
				}
			}
			if (take && g.opt == 3 /* SEP_EXCL_SAME_MOL */)
				take = mi == -1 || mi != s.mol[j];
			if (take) {
				const sepk_pos pj = s.pos[j];
				double dx, dy, dz;
				if (SHIFT) {
					dx = (pi.x - pj.x) + sx;
					dy = (pi.y - pj.y) + sy;
					dz = (pi.z - pj.z) + sz;
				} else {
					dx = sepk_wrap(pi.x - pj.x, g.L[0]);
					dy = sepk_wrap(pi.y - pj.y, g.L[1]);
					dz = sepk_wrap(pi.z - pj.z, g.L[2]);
				}
				const double r2 = dx * dx + dy * dy + dz * dz;
				take = r2 < g.rlist2;
			}
			if (take && g.opt == 2 /* SEP_EXCL_BONDED */) {
				const int oj = s.orig[j];
				int sh = sepk_share(s.partners, npad, t, j, 0, 10, oi, oj);
				if (g.bonded_all)
					sh += sepk_share(s.partners, npad, t, j, 10, 10, oi, oj) +
					      sepk_share(s.partners, npad, t, j, 20, 20, oi, oj);
				// rule 2 is the Coulomb brute-force test of
				// source/sepcoulomb.c:44, which excludes only when
				// exactly one kind of partnership holds
				take = g.bonded_all == 2 ? sh != 1 : sh == 0;
			}
			const unsigned long long vote = __ballot(take);
			const unsigned long long mine = (vote >> gshift) & ((1ull << SEPK_SUB) - 1ull);
			if (take) {
				const int k = m + __popcll(mine & below);
				if (k < cap)
					nbr[sepk_slot(k / SEPK_SUB, t, npad, k % SEPK_SUB)] = j;
			}
			m += __popcll(mine);
		}
	}
	m_out = m;
}

// One group of SEPK_SUB lanes per atom; lane 0 also copies the atom's
// permuted state home.
__global__ __launch_bounds__(SEPK_BLOCK) void k_build(
	sepk_state s, sepk_state home, const int *cellid, const int *start, sepdev_grid g,
	long n, long npad, int cap, int *nbr, int *cnt, double *partial, long rows,
	sepdev_sys *sys, int conditional, int has_partners)
{
	if (conditional && sys->neighb_flag == 0)
		return;
	const long lb = sepk_xcd_block();
	const long t = lb * (SEPK_BLOCK / SEPK_SUB) + threadIdx.x / SEPK_SUB;
	const int sub = threadIdx.x % SEPK_SUB;
	const int gshift = (threadIdx.x & (SEPK_WAVE - 1)) & ~(SEPK_SUB - 1);
	const bool shift = g.shift_ok && sys->unclean == 0;
	double mx = 0.0;
	if (t < n) {
		if (sub == 0)
			sepk_copy_home(s, home, t, npad, has_partners);
		int m;
		if (shift)
			sepk_build_atom<true>(s, t, sub, gshift, cellid, start, g, npad, cap, nbr, m);
		else
			sepk_build_atom<false>(s, t, sub, gshift, cellid, start, g, npad, cap, nbr, m);
		if (m > cap)
			sepk_set_error(sys, SEP_GPU_ENEIGHB);
		if (sub == 0)
			cnt[t] = m < cap ? m : cap;
		mx = (double)m;
	}
	double row[1] = {mx};
	sepk_block_partial<1>(row, 1u, partial, lb, rows);
}

// ---- host side --------------------------------------------------------------

extern "C" int sepdev_reserve_list(sepdev *d, int cap)
{
	if (d->capturing)
		return sepk_fail(SEP_GPU_ESTATE, "list resize during graph capture");
	int lcap = (cap + SEPK_SUB - 1) / SEPK_SUB;
	lcap = (lcap + 7) / 8 * 8;
	if (d->nbr && lcap <= d->lcap)
		return 0;
	SEPK_CHECK(hipStreamSynchronize(d->st));
	if (d->nbr)
		(void)hipFree(d->nbr);
	d->nbr = nullptr;
	SEPK_CHECK(hipMalloc(&d->nbr, (size_t)lcap * SEPK_SUB * d->npad * sizeof(int)));
	SEPK_CHECK(hipMemsetAsync(d->nbr_count, 0, d->npad * sizeof(int), d->st));
	d->lcap = lcap;
	d->cap = lcap * SEPK_SUB;
	d->list_built = 0;
	return 0;
}

extern "C" int sepdev_reset_max_dist(sepdev *d)
{
	// owed to the next leapfrog, which then starts its maximum at zero
	d->max_reset_pending = 1;
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
	int rc = sepk_flush_nh(d);
	if (rc)
		return rc;

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
		if (d->capturing)
			return sepk_fail(SEP_GPU_ESTATE, "list allocated during graph capture");
		if ((rc = sepdev_reserve_list(d, (int)cap)))
			return rc;
	}
	if ((rc = sepk_claim_region(d, SEPK_R_BUILD)))
		return rc;

	sepk_state &cur = d->s[0];   // current state, fixed addresses
	sepk_state &spare = d->s[1]; // receives the permuted order
	const int hp = cur.partners != nullptr;
	unsigned nb = sepk_blocks(d->n);
	if (nb > SEPK_GRID_CAP)
		nb = SEPK_GRID_CAP;
	long sb = (ncell * 64 + SEPK_BLOCK - 1) / SEPK_BLOCK;
	if (sb > SEPK_GRID_CAP)
		sb = SEPK_GRID_CAP;
	k_bin<<<nb, SEPK_BLOCK, 0, d->st>>>(cur.pos, d->n, g, d->cellid_old, d->cell_count, d->sys,
					    conditional);
	k_scan<<<1, 1024, 0, d->st>>>(d->cell_count, d->cell_start, d->cell_fill, ncell, d->sys,
				      conditional);
	k_scatter<<<nb, SEPK_BLOCK, 0, d->st>>>(d->cellid_old, d->cell_start, d->cell_fill,
						d->order_tmp, d->n, d->sys, conditional);
	k_sort_permute<<<(unsigned)sb, SEPK_BLOCK, 0, d->st>>>(
		d->order_tmp, d->cell_start, cur, spare, d->iperm, d->cellid_new, ncell, d->npad,
		d->sys, conditional, hp);
	SEPK_LAUNCH_CHECK();
	const unsigned bb = sepk_blocks_xcd(d->n * SEPK_SUB);
	k_build<<<bb, SEPK_BLOCK, 0, d->st>>>(spare, cur, d->cellid_new, d->cell_start, g, d->n,
					      d->npad, d->cap, d->nbr, d->nbr_count,
					      sepk_region(d, SEPK_R_BUILD), d->region_rows, d->sys,
					      conditional, hp);
	SEPK_LAUNCH_CHECK();
	// largest list; clears the rebuild flag and the owed xn reset
	sepk_task f;
	memset(&f, 0, sizeof f);
	f.mode = SEPK_FIN_BUILD;
	f.region = SEPK_R_BUILD;
	f.nblk = (int)bb;
	f.nv = 1;
	f.maxmask = 1u;
	f.conditional = conditional;
	if ((rc = sepk_queue_task(d, f)))
		return rc;
	// an unconditional build is finished at once, so the host can check it
	if (!conditional && (rc = sepk_flush_tasks(d)))
		return rc;
	d->grid = g;
	d->list_built = 1;
	return 0;
}
