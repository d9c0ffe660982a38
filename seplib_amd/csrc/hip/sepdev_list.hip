// sepdev_list.hip -- cell binning and Verlet neighbour-list build.
//
// A rebuild is a short kernel sequence on the library stream:
//   k_bin          cell index per atom (reference formula or floor+wrap),
//                  per-cell counts and slots by wave-aggregated atomics
//   k_scan         one workgroup: exclusive scan of the counts -> cell_start
//   k_scatter      atom -> start of its cell + its slot
//                  (k_scan_scatter: both in one launch, every workgroup
//                  scanning the counts itself, for grids of up to 16352
//                  cells)
//   k_sort_permute a half-wave per cell (a wave for cells of more than 32
//                  atoms on average): order the segment by original atom
//                  index (so the device order and every list are independent
//                  of atomic timing) and move each member's state to its new
//                  cell-sorted index in the spare copy
//   k_build        one workgroup per cell: stage the stencil cells' segments
//                  in LDS and append the candidates within the list cutoff
//                  (each wave every fourth home atom, 64 candidates a round),
//                  then copy the cells' permuted state home
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

// Cell index of every atom and its slot among the cell's members.  The
// lanes of a wave that share a cell take their slots with one atomic add per
// distinct cell (atoms arrive in the last build's cell order, so a wave holds
// a handful of runs), which keeps the device-wide atomics to a few per wave.
__global__ void k_bin(const sepk_pos *pos, long n, sepdev_grid g, int *cellid, int *slot,
		      int *count, sepdev_sys *sys, int conditional)
{
	if ((conditional && sys->neighb_flag == 0) || sys->error)
		return;
	// the build's atomic maximum starts here
	if (blockIdx.x == 0 && threadIdx.x == 0)
		sys->max_full = 0;
	const int lane = threadIdx.x & 63;
	const unsigned long long below = (1ull << lane) - 1ull;
	const long stride = (long)gridDim.x * blockDim.x;
	// whole waves iterate together (the bound is rounded up to a wave)
	const long nw = (n + 63) / 64 * 64;
	for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < nw; t += stride) {
		const bool act = t < n;
		long c = 0;
		if (act) {
			const sepk_pos p = pos[t];
			c = sepk_cell_of(p, g);
			if (c < 0) {
				sepk_set_error(sys, SEP_GPU_ECELL);
				c = 0;
			}
			// an atom outside [0, n*l) on some axis (allowed: it is
			// binned by truncation or wrapping) rules out the per-cell
			// image shifts
			const double q[3] = {p.x / g.l[0], p.y / g.l[1], p.z / g.l[2]};
			if (!(q[0] >= 0.0 && q[0] < g.n[0] && q[1] >= 0.0 && q[1] < g.n[1] &&
			      q[2] >= 0.0 && q[2] < g.n[2]))
				sys->unclean = 1;
			cellid[t] = (int)c;
		}
		unsigned long long todo = __ballot(act);
		int my = 0;
		while (todo) {
			const int lead = __ffsll((long long)todo) - 1;
			const int cl = __shfl((int)c, lead, 64);
			const unsigned long long same = __ballot(act && (int)c == cl);
			int base = 0;
			if (lane == lead)
				base = atomicAdd(&count[cl], __popcll(same));
			base = __shfl(base, lead, 64);
			if (act && (int)c == cl)
				my = base + __popcll(same & below);
			todo &= ~same;
		}
		if (act)
			slot[t] = my;
	}
}

// Exclusive scan of count[0..ncell) into start[0..ncell], zeroing count for
// the next rebuild.  One workgroup; each thread scans SEPK_SCAN_ITEMS
// consecutive cells of a pass in registers (one memory round trip per pass).
// It also fixes whether the rest of the rebuild runs: sys->frozen = an error
// is set (by an earlier step, or by this binning), read by the scatter,
// permutation and build kernels, which must all run or all skip (the build
// may itself raise an error while it runs).  The counts are zeroed either
// way, so a later rebuild starts from zero.
#define SEPK_SCAN_ITEMS 16
__global__ __launch_bounds__(1024) void k_scan(int *count, int *start, long ncell,
					       sepdev_sys *sys, int conditional)
{
	if (conditional && sys->neighb_flag == 0)
		return;
	if (threadIdx.x == 0)
		sys->frozen = sys->error != 0;
	__shared__ int wsum[16];
	__shared__ int carry_s;
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	int carry = 0;
	const long pass = 1024L * SEPK_SCAN_ITEMS;
	for (long base = 0; base < ncell; base += pass) {
		const long c0 = base + (long)threadIdx.x * SEPK_SCAN_ITEMS;
		int v[SEPK_SCAN_ITEMS];
		if (c0 + SEPK_SCAN_ITEMS <= ncell) {
			// 16-byte aligned (count comes from hipMalloc, c0 % 16 == 0)
			const int4 *q = (const int4 *)(count + c0);
#pragma unroll
			for (int k = 0; k < SEPK_SCAN_ITEMS / 4; k++) {
				const int4 w = q[k];
				v[4 * k] = w.x;
				v[4 * k + 1] = w.y;
				v[4 * k + 2] = w.z;
				v[4 * k + 3] = w.w;
			}
		} else {
#pragma unroll
			for (int k = 0; k < SEPK_SCAN_ITEMS; k++)
				v[k] = c0 + k < ncell ? count[c0 + k] : 0;
		}
		int tot = 0;
#pragma unroll
		for (int k = 0; k < SEPK_SCAN_ITEMS; k++)
			tot += v[k];
		int x = tot; // inclusive wave scan of the thread totals
		for (int o = 1; o < 64; o <<= 1) {
			const int y = __shfl_up(x, o, 64);
			if (lane >= o)
				x += y;
		}
		if (lane == 63)
			wsum[wid] = x;
		__syncthreads();
		int off = carry;
		for (int w = 0; w < wid; w++)
			off += wsum[w];
		off += x - tot;
#pragma unroll
		for (int k = 0; k < SEPK_SCAN_ITEMS; k++) {
			if (c0 + k < ncell) {
				start[c0 + k] = off;
				count[c0 + k] = 0;
			}
			off += v[k];
		}
		if (threadIdx.x == 1023)
			carry_s = off;
		__syncthreads();
		carry = carry_s;
		__syncthreads();
	}
	if (threadIdx.x == 0)
		start[ncell] = carry;
}

// The scan and the scatter in one launch, for grids of at most
// SEPK_PLACE_CELLS cells: every workgroup scans the cell counts into LDS
// (one pass of SEPK_SCAN_ITEMS cells a thread, as k_scan), the first also
// writes them to start[], and each then places its share of the atoms.  No
// workgroup waits for another.  It fixes sys->frozen as k_scan does; the
// counts, which every workgroup reads, are zeroed by k_sort_permute.
#define SEPK_PLACE_CELLS (1024 * SEPK_SCAN_ITEMS - 32) // 64 KB of LDS with the wave sums
__global__ __launch_bounds__(1024) void k_scan_scatter(const int *count, int *start,
						       const int *cellid, const int *slot, int *order,
						       long ncell, long n, sepdev_sys *sys,
						       int conditional)
{
	if (conditional && sys->neighb_flag == 0)
		return;
	const bool frozen = sys->error != 0;
	if (blockIdx.x == 0 && threadIdx.x == 0)
		sys->frozen = frozen;
	if (frozen)
		return; // the permutation and the build skip as well
	__shared__ int sstart[SEPK_PLACE_CELLS + 1];
	__shared__ int wsum[16];
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	const long c0 = (long)threadIdx.x * SEPK_SCAN_ITEMS;
	int v[SEPK_SCAN_ITEMS];
	if (c0 + SEPK_SCAN_ITEMS <= ncell) {
		// 16-byte aligned (count comes from hipMalloc, c0 % 16 == 0)
		const int4 *q = (const int4 *)(count + c0);
#pragma unroll
		for (int k = 0; k < SEPK_SCAN_ITEMS / 4; k++) {
			const int4 w = q[k];
			v[4 * k] = w.x;
			v[4 * k + 1] = w.y;
			v[4 * k + 2] = w.z;
			v[4 * k + 3] = w.w;
		}
	} else {
#pragma unroll
		for (int k = 0; k < SEPK_SCAN_ITEMS; k++)
			v[k] = c0 + k < ncell ? count[c0 + k] : 0;
	}
	int tot = 0;
#pragma unroll
	for (int k = 0; k < SEPK_SCAN_ITEMS; k++)
		tot += v[k];
	int x = tot; // inclusive wave scan of the thread totals
	for (int o = 1; o < 64; o <<= 1) {
		const int y = __shfl_up(x, o, 64);
		if (lane >= o)
			x += y;
	}
	if (lane == 63)
		wsum[wid] = x;
	__syncthreads();
	int off = 0;
	for (int w = 0; w < wid; w++)
		off += wsum[w];
	off += x - tot;
#pragma unroll
	for (int k = 0; k < SEPK_SCAN_ITEMS; k++) {
		if (c0 + k < ncell)
			sstart[c0 + k] = off;
		off += v[k];
		if (c0 + k == ncell - 1)
			sstart[ncell] = off;
	}
	__syncthreads();
	if (blockIdx.x == 0)
		for (long c = threadIdx.x; c <= ncell; c += blockDim.x)
			start[c] = sstart[c];
	for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
	     t += (long)gridDim.x * blockDim.x)
		order[sstart[cellid[t]] + slot[t]] = (int)t;
}

// Atom t -> its place in the cell-sorted order (slot from k_bin).
__global__ void k_scatter(const int *cellid, const int *slot, const int *start, int *order,
			  long n, sepdev_sys *sys, int conditional)
{
	if ((conditional && sys->neighb_flag == 0) || sys->frozen)
		return;
	for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
	     t += (long)gridDim.x * blockDim.x)
		order[start[cellid[t]] + slot[t]] = (int)t;
}

// Copies the partner table of atom o in src to slot t of dst, eight
// entries per round trip.
__device__ __forceinline__ void sepk_copy_partners(const sepk_state &src, const sepk_state &dst,
						   long o, long t, long npad)
{
	for (int k0 = 0; k0 < 40; k0 += 8) {
		int p[8];
#pragma unroll
		for (int k = 0; k < 8; k++)
			p[k] = sepk_at(src.partners + (size_t)(k0 + k) * npad, (unsigned)o);
#pragma unroll
		for (int k = 0; k < 8; k++)
			sepk_at(dst.partners + (size_t)(k0 + k) * npad, (unsigned)t) = p[k];
	}
}

// Moves atom o of src to slot t of dst.  Each half of the state is loaded
// whole before any of it is stored, so the move pays two memory round
// trips instead of one per field (source and destination may alias as far
// as the compiler knows).  pending: xn = x and cross_neighb = 0 are owed.
__device__ __forceinline__ void sepk_move_atom(const sepk_state &src, long o,
					       const sepk_state &dst, long t, long npad,
					       bool pending)
{
	const unsigned ou = (unsigned)o, tu = (unsigned)t;
	const sepk_pos p = sepk_at(src.pos, ou);
	{
		const int4 q = sepk_at(src.qpos, ou);
		const double v0 = sepk_at(src.v, ou), v1 = sepk_at(src.v + npad, ou);
		const double v2 = sepk_at(src.v + 2 * npad, ou);
		const double f0 = sepk_at(src.f, ou), f1 = sepk_at(src.f + npad, ou);
		const double f2 = sepk_at(src.f + 2 * npad, ou);
		const double a0 = sepk_at(src.a, ou), a1 = sepk_at(src.a + npad, ou);
		const double a2 = sepk_at(src.a + 2 * npad, ou);
		sepk_at(dst.pos, tu) = p;
		sepk_at(dst.qpos, tu) = q;
		sepk_at(dst.v, tu) = v0;
		sepk_at(dst.v + npad, tu) = v1;
		sepk_at(dst.v + 2 * npad, tu) = v2;
		sepk_at(dst.f, tu) = f0;
		sepk_at(dst.f + npad, tu) = f1;
		sepk_at(dst.f + 2 * npad, tu) = f2;
		sepk_at(dst.a, tu) = a0;
		sepk_at(dst.a + npad, tu) = a1;
		sepk_at(dst.a + 2 * npad, tu) = a2;
	}
	double x0 = sepk_at(src.xn, ou), x1 = sepk_at(src.xn + npad, ou);
	double x2 = sepk_at(src.xn + 2 * npad, ou);
	int n0 = sepk_at(src.cn, ou), n1 = sepk_at(src.cn + npad, ou), n2 = sepk_at(src.cn + 2 * npad, ou);
	const int r0 = sepk_at(src.cr, ou), r1 = sepk_at(src.cr + npad, ou);
	const int r2 = sepk_at(src.cr + 2 * npad, ou);
	const double m = sepk_at(src.m, ou);
	const int mol = sepk_at(src.mol, ou), orig = sepk_at(src.orig, ou);
	const signed char ty = sepk_at(src.type, ou);
	if (pending) {
		x0 = p.x;
		x1 = p.y;
		x2 = p.z;
		n0 = n1 = n2 = 0;
	}
	sepk_at(dst.xn, tu) = x0;
	sepk_at(dst.xn + npad, tu) = x1;
	sepk_at(dst.xn + 2 * npad, tu) = x2;
	sepk_at(dst.cn, tu) = n0;
	sepk_at(dst.cn + npad, tu) = n1;
	sepk_at(dst.cn + 2 * npad, tu) = n2;
	sepk_at(dst.cr, tu) = r0;
	sepk_at(dst.cr + npad, tu) = r1;
	sepk_at(dst.cr + 2 * npad, tu) = r2;
	sepk_at(dst.m, tu) = m;
	sepk_at(dst.mol, tu) = mol;
	sepk_at(dst.orig, tu) = orig;
	sepk_at(dst.type, tu) = ty;
}

// HW lanes per cell (grid-stride over cells; HW = 32: the two halves of a
// wave take two cells, for cells of up to about 32 atoms; HW = 64: a wave
// per cell): rank each member by its original index -- the keys of the
// cell are held one per lane and compared through shuffles within the
// group -- and move its state to start + rank in the spare copy.
template <int HW>
__global__ void k_sort_permute(const int *order, const int *start, sepk_state src,
			       sepk_state dst, int *iperm, long ncell,
			       long npad, sepdev_sys *sys, int conditional,
			       int has_partners, int *count)
{
	if (conditional && sys->neighb_flag == 0)
		return;
	// the counts start from zero at the next rebuild, whether this one runs
	// on or stops here
	for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
	     c += (long)gridDim.x * blockDim.x)
		count[c] = 0;
	if (sys->frozen)
		return;
	const int pending = sys->xn_pending;
	static_assert(HW == 32 || HW == 64, "half or whole waves");
	const int hl = threadIdx.x % HW;
	const long halves = (long)gridDim.x * (blockDim.x / HW);
	for (long c = ((long)blockIdx.x * blockDim.x + threadIdx.x) / HW; c < ncell; c += halves) {
		const int s = start[c], e = start[c + 1];
		for (int i0 = s; i0 < e; i0 += HW) {
			const int i = i0 + hl;
			const long o = i < e ? sepk_at(order, (unsigned)i) : 0;
			const int key = i < e ? sepk_at(src.orig, (unsigned)o) : 0;
			int rank = 0;
			for (int b = s; b < e; b += HW) {
				// keys of members b .. b + HW - 1, one per lane of the group
				const unsigned ob = b + hl < e ? (unsigned)sepk_at(order, (unsigned)(b + hl)) : 0u;
				const int kj = b == i0 ? key : (b + hl < e ? sepk_at(src.orig, ob) : 0);
				const int m = min(HW, e - b);
				for (int jj = 0; jj < m; jj++)
					rank += __shfl(kj, jj, HW) < key;
			}
			if (i >= e)
				continue;
			const long t = s + rank;
			sepk_move_atom(src, o, dst, t, npad, pending);
			sepk_at(iperm, (unsigned)key) = (int)t;
			if (has_partners)
				sepk_copy_partners(src, dst, o, t, npad);
		}
	}
}

// The same for cells of more than 64 atoms on average (dense molecular
// boxes): a workgroup per cell (grid-stride over cells).  The cell's keys
// go to LDS once, SEPK_SORT_KEYS at a time, and each thread ranks its
// members against them with broadcast reads, so a cell costs one round
// trip for its keys and one for its moves where a wave per cell pays one
// per 64 keys, and four waves work on it instead of one.  Same ranks, so
// the same permutation as k_sort_permute.
#define SEPK_SORT_KEYS 1024
__global__ __launch_bounds__(SEPK_BLOCK) void k_sort_permute_wg(
	const int *order, const int *start, sepk_state src, sepk_state dst, int *iperm,
	long ncell, long npad, sepdev_sys *sys, int conditional, int has_partners, int *count)
{
	if (conditional && sys->neighb_flag == 0)
		return;
	for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
	     c += (long)gridDim.x * blockDim.x)
		count[c] = 0;
	if (sys->frozen)
		return;
	const int pending = sys->xn_pending;
	__shared__ int keys[SEPK_SORT_KEYS];
	for (long c = blockIdx.x; c < ncell; c += gridDim.x) {
		const int s = start[c], e = start[c + 1];
		for (int i0 = s; i0 < e; i0 += SEPK_BLOCK) {
			const int i = i0 + (int)threadIdx.x;
			const long o = i < e ? sepk_at(order, (unsigned)i) : 0;
			const int key = i < e ? sepk_at(src.orig, (unsigned)o) : 0;
			int rank = 0;
			for (int b = s; b < e; b += SEPK_SORT_KEYS) {
				const int m = min(SEPK_SORT_KEYS, e - b);
				__syncthreads(); // the previous keys are no longer read
				for (int q = threadIdx.x; q < m; q += SEPK_BLOCK) {
					const int x = b + q; // a thread's own member: its key
					keys[q] = x == i ? key
							 : sepk_at(src.orig, (unsigned)sepk_at(order, (unsigned)x));
				}
				__syncthreads();
#pragma unroll 4
				for (int jj = 0; jj < m; jj++)
					rank += keys[jj] < key;
			}
			if (i >= e)
				continue;
			const long t = s + rank;
			sepk_move_atom(src, o, dst, t, npad, pending);
			sepk_at(iperm, (unsigned)key) = (int)t;
			if (has_partners)
				sepk_copy_partners(src, dst, o, t, npad);
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
	sepk_move_atom(src, t, dst, t, npad, false);
	if (has_partners)
		sepk_copy_partners(src, dst, t, t, npad);
}

// One workgroup per home cell at a time (cells split into one contiguous
// range per XCD, walked with a stride of the range's workgroup count).
// The stencil cells' segments are staged in LDS as SoA chunks of
// SEPK_BCHUNK candidates (position, device index, stencil offset, molecule,
// quantized position) and the cell's members ("home" atoms, SEPK_BHOME at a
// time) beside them.  Each wave owns every fourth home atom: a round takes
// 64 staged candidates into registers (one a lane) and tests them against
// each of the wave's atoms, whose position is an LDS broadcast; the lanes
// that pass store their candidate to the atom's list at once, each at the
// atom's length so far plus its rank among the passing lanes (mbcnt), and
// the length (lane kk of a register for the wave's kk-th atom) grows by
// their number.  Lists therefore come out in stencil-offset order and,
// within a cell, in device order.  Entry k of atom t goes to slot
// sepk_entry(k, t, npad).
//
// SHIFT: with at least 5 cells on every axis and every atom inside its own
// cell, |x_i - x_j| < 2l <= 0.4 L for a target cell reached without
// crossing the box and > L - 2l >= 0.6 L for one reached across it, so the
// reference's single-image wrap (subtract or add L exactly when |d| > L/2)
// is the constant shift of the target cell; d + shift rounds exactly like
// its d - L / d + L, and d + 0 is d.  Otherwise the wrap is evaluated per
// candidate.
//
// Quantized test (SHIFT and a cubic box): every coordinate is also held as
// floor(x * 65536 / L) mod 2^16, x and y in one word and z in the low half
// of another.  A packed 16-bit subtraction gives the minimum-image
// difference in units of u = L / 65536, each component within one unit of
// the exact one (|x_i - x_j| < 0.4 L keeps it away from the wrap), and two
// saturating packed dot products its square r2q.  With R = r_list / u the
// distance is below r_list - 0.26 u when r2q < floor((R - 2)^2) and above
// r_list + 0.26 u when r2q > ceil((R + 2)^2) (sqrt(3) < 1.74), far outside
// the rounding of the FP64 r2; only candidates between the two bounds (a
// shell about 4 u thick: a few in ten thousand) take the FP64 test, the
// reference's own, so the accepted set is the FP64 test's exactly.
#define SEPK_BT 256      // threads of a build workgroup (4 waves)
#define SEPK_BCHUNK 384  // candidates staged at a time
#define SEPK_BHOME 64    // home atoms staged at a time
#define SEPK_BGRID 1792  // build workgroups at most (a multiple of 8)
static_assert(SEPK_BCHUNK % 64 == 0, "candidates are tested 64 at a time");
static_assert(SEPK_BHOME <= 64 * (SEPK_BT / 64), "a wave's home atoms fit its lanes");

struct sepk_build_lds {
	double cx[SEPK_BCHUNK], cy[SEPK_BCHUNK], cz[SEPK_BCHUNK];
	uint2 cq[SEPK_BCHUNK];
	int cj[SEPK_BCHUNK], cm[SEPK_BCHUNK];
	unsigned char co[SEPK_BCHUNK];
	double hx[SEPK_BHOME], hy[SEPK_BHOME], hz[SEPK_BHOME];
	uint2 hq[SEPK_BHOME];
	int horig[SEPK_BHOME], hmol[SEPK_BHOME];
	int segjs[SEPDEV_MAX_OFF], segpre[SEPDEV_MAX_OFF + 1];
	int segrule[SEPDEV_MAX_OFF]; // rule when the target is the home cell, else -1
	double segsh[SEPDEV_MAX_OFF][3];
	int hs, h, anyshift;
};

// Lists of fewer than 2^20 padded atoms and under 4 GiB: a list slot is a
// 32-bit byte offset from its atom's first slot, formed with a 24-bit
// multiply (otherwise sepk_entry's 64-bit arithmetic).
__device__ __forceinline__ bool sepk_np32(long npad, int cap)
{
	return npad < (1l << 20) &&
	       (long)((cap + SEPK_PAD - 1) / SEPK_PAD) * npad * SEPK_PAD * 4 < (1l << 32);
}

// Bounds of the quantized test; on = the test applies (SHIFT, cubic box).
struct sepk_qtest {
	double scale; // 65536 / L
	int lo2, hi2;
	bool on;
};

__device__ __forceinline__ sepk_qtest sepk_qtest_of(const sepdev_grid &g, bool shift)
{
	sepk_qtest q;
	q.scale = 65536.0 / g.L[0];
	const double R = sqrt(g.rlist2) * q.scale;
	// SHIFT has L >= 5 r_list, so R <= 13108 and both bounds fit 31 bits
	q.on = shift && g.L[0] == g.L[1] && g.L[1] == g.L[2] && R < 30000.0;
	q.lo2 = R > 2.0 ? (int)floor((R - 2.0) * (R - 2.0)) : 0;
	q.hi2 = q.on ? (int)ceil((R + 2.0) * (R + 2.0)) : 0;
	return q;
}

// x and y in the low and high halves of .x, z in the low half of .y
__device__ __forceinline__ uint2 sepk_quantize(const sepk_pos &p, double scale)
{
	const unsigned qx = (unsigned)(int)floor(p.x * scale) & 0xffffu;
	const unsigned qy = (unsigned)(int)floor(p.y * scale) & 0xffffu;
	const unsigned qz = (unsigned)(int)floor(p.z * scale) & 0xffffu;
	return make_uint2(qx | qy << 16, qz);
}

typedef short sepk_short2 __attribute__((ext_vector_type(2)));

// Squared quantized minimum-image distance, saturated at 2^31 - 1.
__device__ __forceinline__ int sepk_r2q(uint2 a, uint2 b)
{
	const sepk_short2 dxy =
		__builtin_bit_cast(sepk_short2, a.x) - __builtin_bit_cast(sepk_short2, b.x);
	const sepk_short2 dz =
		__builtin_bit_cast(sepk_short2, a.y) - __builtin_bit_cast(sepk_short2, b.y);
	return __builtin_amdgcn_sdot2(dz, dz, __builtin_amdgcn_sdot2(dxy, dxy, 0, true), true);
}

// Lane `lane` of v (a wave-uniform lane number) becomes val, whatever the
// execution mask.
__device__ __forceinline__ void sepk_writelane(int &v, int val, int lane)
{
	asm volatile("v_writelane_b32 %0, %1, %2" : "+v"(v) : "s"(val), "{m0}"(lane));
}

// One round: the 64 staged candidates [q0, q0 + 64) of a chunk of kn
// against the wave's home atoms i = wid, wid + 4, ... < hn, storing the
// accepted ones to their lists.  Lane kk of cnt holds the list length of
// the wave's kk-th home atom; entries at or beyond cap are counted but not
// stored.
// QNT: the quantized test (SHIFT holds).  SH: some stencil target of the
// cell lies across the box (shift adds; the quantized test needs none).
// EXCL: 0 = SEP_ALL, 2 = SEP_EXCL_BONDED, 3 = SEP_EXCL_SAME_MOL.
template <bool SHIFT, bool SH, int EXCL, bool QNT, bool NP32>
__device__ __forceinline__ void sepk_build_round(
	sepk_build_lds &L, const sepk_state &s, const sepdev_grid &g, const sepk_qtest &qt,
	int q0, int kn, long tbase, int hn, long npad, int cap, int *nbr, int &cnt)
{
	const int lane = threadIdx.x & 63;
	const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); // wave-uniform
	const int q = min(q0 + lane, kn - 1);
	// lanes past the staged candidates take bounds that no distance meets
	const bool valid = lane < kn - q0;
	// NP32 (sepk_np32): 32-bit byte offsets with a full-rate 24-bit multiply
	const unsigned npad16 = (unsigned)npad * SEPK_PAD;
	const int j = L.cj[q];
	const int o = L.co[q];
	const int rule = L.segrule[o];
	const int mj = EXCL == 3 ? L.cm[q] : -1;
	// lanes whose candidate is in the home cell (SHIFT: through offset 0,
	// whose rule is NE)
	const unsigned long long home = __ballot(rule >= 0);
	// original index of a home-cell candidate, for the GT / LT rules
	const int oj = !SHIFT && rule >= SEPDEV_RULE_GT ? s.orig[j] : 0;
	uint2 cq = make_uint2(0u, 0u);
	double px = 0.0, py = 0.0, pz = 0.0;
	if (QNT) {
		cq = L.cq[q];
	} else {
		px = L.cx[q];
		py = L.cy[q];
		pz = L.cz[q];
	}
	int lo2 = valid ? qt.lo2 : INT_MIN, hi2 = valid ? qt.hi2 : -1;
	double rl2 = valid ? g.rlist2 : -1.0;
	// the home atom i's own candidate is the one with j - tbase == i
	int jrel = j - (int)tbase;
	// held in registers as they are: the compiler would otherwise fold the
	// selects back into every test of the loop below
	asm volatile("" : "+v"(lo2), "+v"(hi2), "+v"(rl2), "+v"(jrel));
	double sx = 0.0, sy = 0.0, sz = 0.0;
	if (SH && !QNT) {
		sx = L.segsh[o][0];
		sy = L.segsh[o][1];
		sz = L.segsh[o][2];
	}
	int kk = 0;
	for (int i = wid; i < hn; i += SEPK_BT / 64, kk++) {
		const long t = tbase + i;
		const int t32 = (int)t;
		unsigned long long take;
		if (QNT) {
			const int r2q = sepk_r2q(L.hq[i], cq);
			take = __ballot(r2q < lo2);
			const unsigned long long near = __ballot(r2q <= hi2) & ~take;
			if (near) {
				// the reference's FP64 test, with the target cell's shift
				const double dx = L.hx[i] - L.cx[q] + L.segsh[o][0];
				const double dy = L.hy[i] - L.cy[q] + L.segsh[o][1];
				const double dz = L.hz[i] - L.cz[q] + L.segsh[o][2];
				const double r2 = dx * dx + dy * dy + dz * dz;
				take |= __ballot(r2 < g.rlist2) & near;
			}
		} else {
			double dx, dy, dz;
			if (SHIFT) {
				dx = L.hx[i] - px;
				dy = L.hy[i] - py;
				dz = L.hz[i] - pz;
				if (SH) {
					dx += sx;
					dy += sy;
					dz += sz;
				}
			} else {
				dx = sepk_wrap(L.hx[i] - px, g.L[0]);
				dy = sepk_wrap(L.hy[i] - py, g.L[1]);
				dz = sepk_wrap(L.hz[i] - pz, g.L[2]);
			}
			const double r2 = dx * dx + dy * dy + dz * dz;
			take = __ballot(r2 < rl2);
		}
		if (SHIFT) {
			// the 27-cell stencil reaches the home cell only through
			// offset 0: drop the atom itself
			take &= ~__ballot(jrel == i);
		} else if (take & home) {
			const int oi = L.horig[i];
			const bool drop = rule == SEPDEV_RULE_NE ? j == t32
					  : rule == SEPDEV_RULE_GT ? oj <= oi
								   : oj >= oi;
			take &= ~(__ballot(drop) & home);
		}
		if (EXCL == 3 /* SEP_EXCL_SAME_MOL */) {
			const int mi = L.hmol[i];
			if (mi != -1)
				take &= ~__ballot(mj == mi);
		}
		if (EXCL == 2 /* SEP_EXCL_BONDED */ && take) {
			bool ok = false;
			if (__builtin_amdgcn_inverse_ballot_w64(take)) {
				const int oi = L.horig[i];
				const int oj2 = s.orig[j];
				int sh = sepk_share(s.partners, npad, t, j, 0, 10, oi, oj2);
				if (g.bonded_all)
					sh += sepk_share(s.partners, npad, t, j, 10, 10, oi, oj2) +
					      sepk_share(s.partners, npad, t, j, 20, 20, oi, oj2);
				// rule 2 is the Coulomb brute-force test of
				// source/sepcoulomb.c:44, which excludes only when
				// exactly one kind of partnership holds
				ok = g.bonded_all == 2 ? sh != 1 : sh == 0;
			}
			take = __ballot(ok);
		}
		if (take) {
			const int e0 = __builtin_amdgcn_readlane(cnt, kk);
			const int e = (int)__builtin_amdgcn_mbcnt_hi(
				(unsigned)(take >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)take, (unsigned)e0));
			if (__builtin_amdgcn_inverse_ballot_w64(take) && e < cap) {
				if (NP32) {
					// the atom's first slot as a scalar base, a 32-bit byte
					// offset (sepk_entry)
					const unsigned off = __umul24((unsigned)e / SEPK_PAD, npad16) +
							     ((unsigned)e % SEPK_SUB) * SEPK_BATCH +
							     ((unsigned)e / SEPK_SUB) % SEPK_BATCH;
					*(int *)((char *)(nbr + (size_t)t * SEPK_PAD) + (size_t)(off * 4u)) = j;
				} else {
					nbr[sepk_entry(e, t, npad)] = j;
				}
			}
			sepk_writelane(cnt, e0 + __popcll(take), kk);
		}
	}
}

// Slice `slice` of `nslice` of cell c: its home atoms
// [h * slice / nslice, h * (slice + 1) / nslice), SEPK_BHOME at a time.
template <bool SHIFT, int EXCL>
__device__ __forceinline__ void sepk_build_cell(
	sepk_build_lds &L, const sepk_state &s, long c, int slice, int nslice,
	const int *start, const sepdev_grid &g, const sepk_qtest &qt, long npad, int cap,
	int phantom, int *nbr, int *cnt, sepdev_sys *sys, int &mymax)
{
	const int lane = threadIdx.x & 63;
	const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); // wave-uniform
	const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
	__syncthreads(); // the previous cell's readers are done with the LDS
	if (wid == 0) {
		// stencil segments, their prefix sums and whether any is shifted
		int len = 0;
		bool shifted = false;
		if (lane < g.noff) {
			const int o = lane;
			const int cxi = (int)(c % n0), cyi = (int)((c / n0) % n1);
			const int czi = (int)(c / ((long)n0 * n1));
			const int ux = cxi + g.off[o][0], uy = cyi + g.off[o][1], uz = czi + g.off[o][2];
			const int tx = sepk_mod(ux, n0), ty = sepk_mod(uy, n1), tz = sepk_mod(uz, n2);
			const long c2 = tx + (long)ty * n0 + (long)tz * n0 * n1;
			const double sx = ux >= n0 ? -g.L[0] : ux < 0 ? g.L[0] : 0.0;
			const double sy = uy >= n1 ? -g.L[1] : uy < 0 ? g.L[1] : 0.0;
			const double sz = uz >= n2 ? -g.L[2] : uz < 0 ? g.L[2] : 0.0;
			L.segsh[o][0] = sx;
			L.segsh[o][1] = sy;
			L.segsh[o][2] = sz;
			shifted = sx != 0.0 || sy != 0.0 || sz != 0.0;
			const int js = start[c2];
			L.segjs[o] = js;
			len = start[c2 + 1] - js;
			L.segrule[o] = c2 == c ? g.off[o][3] : -1;
		}
		int x = len; // inclusive scan over the segments
		for (int d = 1; d < 64; d <<= 1) {
			const int y = __shfl_up(x, d, 64);
			if (lane >= d)
				x += y;
		}
		if (lane == 0)
			L.segpre[0] = 0;
		if (lane < g.noff)
			L.segpre[lane + 1] = x;
		const unsigned long long sv = __ballot(shifted);
		if (lane == 0) {
			L.anyshift = sv != 0ull;
			L.hs = start[c];
			L.h = start[c + 1] - L.hs;
		}
	}
	__syncthreads();
	const int hs = L.hs, h = L.h, T = L.segpre[g.noff];
	const bool anyshift = L.anyshift;
	// h < 2^24 atoms a cell and nslice <= 64: 32-bit products
	const int hbeg = nslice == 1 ? 0 : (int)((unsigned)(h * slice) / (unsigned)nslice);
	const int hend = nslice == 1 ? h : (int)((unsigned)(h * (slice + 1)) / (unsigned)nslice);
	const bool qnt = SHIFT && qt.on;
	const bool np32 = sepk_np32(npad, cap);
	for (int hb = hbeg; hb < hend; hb += SEPK_BHOME) {
		const int hn = min(hend - hb, SEPK_BHOME);
		int len = 0; // lane kk: list length of the wave's kk-th home atom
		for (int k0 = 0; k0 < T; k0 += SEPK_BCHUNK) {
			const int kn = min(T - k0, SEPK_BCHUNK);
			if (k0 > 0)
				__syncthreads(); // the previous chunk is no longer read
			// stage [k0, k0 + kn): thread x takes candidates x, x + SEPK_BT,
			// ...; each finds its segment by a binary search of the prefix
			// sums and all of its loads are issued before any LDS store
			constexpr int PER = (SEPK_BCHUNK + SEPK_BT - 1) / SEPK_BT;
			int jq[PER], oq[PER];
#pragma unroll
			for (int r = 0; r < PER; r++) {
				const int q = threadIdx.x + r * SEPK_BT;
				const int k = k0 + q;
				int o = 0; // the last segment starting at or before k
#pragma unroll
				for (int step = 32; step > 0; step >>= 1)
					if (o + step < g.noff && L.segpre[o + step] <= k)
						o += step;
				oq[r] = q < kn && q < SEPK_BCHUNK ? o : -1;
				jq[r] = L.segjs[o] + (k - L.segpre[o]);
			}
			sepk_pos pq[PER];
			int mq[PER];
#pragma unroll
			for (int r = 0; r < PER; r++) {
				if (oq[r] >= 0) {
					pq[r] = sepk_at(s.pos, (unsigned)jq[r]);
					if (EXCL == 3)
						mq[r] = sepk_at(s.mol, (unsigned)jq[r]);
				}
			}
#pragma unroll
			for (int r = 0; r < PER; r++) {
				if (oq[r] >= 0) {
					const int q = threadIdx.x + r * SEPK_BT;
					L.cx[q] = pq[r].x;
					L.cy[q] = pq[r].y;
					L.cz[q] = pq[r].z;
					if (qnt)
						L.cq[q] = sepk_quantize(pq[r], qt.scale);
					L.cj[q] = jq[r];
					L.co[q] = (unsigned char)oq[r];
					if (EXCL == 3)
						L.cm[q] = mq[r];
				}
			}
			if (k0 == 0)
				for (int i = threadIdx.x; i < hn; i += SEPK_BT) {
					const unsigned t = (unsigned)(hs + hb + i);
					const sepk_pos p = sepk_at(s.pos, t);
					L.hx[i] = p.x;
					L.hy[i] = p.y;
					L.hz[i] = p.z;
					if (qnt)
						L.hq[i] = sepk_quantize(p, qt.scale);
					L.horig[i] = sepk_at(s.orig, t);
					L.hmol[i] = EXCL == 3 ? sepk_at(s.mol, t) : -1;
				}
			__syncthreads();
			const long tb = hs + hb;
			for (int q0 = 0; q0 < kn; q0 += 64) {
				if (qnt && np32)
					sepk_build_round<true, false, EXCL, true, true>(L, s, g, qt, q0, kn, tb,
											hn, npad, cap, nbr, len);
				else if (qnt)
					sepk_build_round<true, false, EXCL, true, false>(L, s, g, qt, q0, kn, tb,
											 hn, npad, cap, nbr, len);
				else if (SHIFT && anyshift)
					sepk_build_round<SHIFT, true, EXCL, false, false>(L, s, g, qt, q0, kn, tb,
											  hn, npad, cap, nbr, len);
				else
					sepk_build_round<SHIFT, false, EXCL, false, false>(L, s, g, qt, q0, kn, tb,
											   hn, npad, cap, nbr, len);
			}
		}
		// lane kk of each wave: its home atom's count and length bound
		const int i = wid + (SEPK_BT / 64) * lane;
		if (lane < SEPK_BHOME / (SEPK_BT / 64) && i < hn) {
			const long t = hs + hb + i;
			cnt[t] = len < cap ? len : cap;
			if (len > cap)
				sepk_set_error(sys, SEP_GPU_ENEIGHB);
			mymax = max(mymax, len);
		}
		// padding with the phantom slot to a multiple of SEPK_PAD, four
		// atoms a store: lane 16 a + s takes entry s of the last block of
		// the a-th atom of the group when that entry lies in [len, pad)
		const int nk = wid < hn ? (hn - wid + (SEPK_BT / 64) - 1) / (SEPK_BT / 64) : 0;
		for (int g0 = 0; g0 < nk; g0 += 64 / SEPK_PAD) {
			const int kk = g0 + lane / SEPK_PAD;
			const int lk = __shfl(len, kk, 64);
			const int pad = min((lk + SEPK_PAD - 1) / SEPK_PAD * SEPK_PAD, cap);
			const int e = lk / SEPK_PAD * SEPK_PAD + lane % SEPK_PAD;
			if (kk < nk && e >= lk && e < pad)
				nbr[sepk_entry(e, hs + hb + wid + (SEPK_BT / 64) * kk, npad)] = phantom;
		}
		if (hb + SEPK_BHOME < hend)
			__syncthreads(); // the next batch restages the candidates
	}
}

// Persistent over (cell, slice) items, nslice per cell; the largest list
// length goes to sys->max_full by an atomic maximum (order-independent, so
// the result is deterministic).  The permuted state of the workgroup's cells
// goes home at the end, in one batch, so its memory round trip is paid once.
template <int EXCL>
__global__ __launch_bounds__(SEPK_BT) __attribute__((amdgpu_waves_per_eu(EXCL == 2 ? 4 : 7))) void k_build(
	sepk_state s, sepk_state home, const int *start, sepdev_grid g, long ncell, int nslice,
	long n, long npad, int cap, int *nbr, int *cnt, sepdev_sys *sys, int conditional,
	int has_partners)
{
	if ((conditional && sys->neighb_flag == 0) || sys->frozen)
		return;
	__shared__ sepk_build_lds L;
	const bool shift = g.shift_ok && sys->unclean == 0;
	const sepk_qtest qt = sepk_qtest_of(g, shift);
	const long per = gridDim.x / SEPK_NXCD;
	const long xcd = blockIdx.x % SEPK_NXCD, k = blockIdx.x / SEPK_NXCD;
	const long nitem = ncell * nslice;
	const long iper = (nitem + SEPK_NXCD - 1) / SEPK_NXCD;
	const long ilo = xcd * iper, ihi = min(ilo + iper, nitem);
	int mymax = 0;
	for (long it = ilo + k; it < ihi; it += per) {
		// items < 2^31 (the host bounds nslice by the atoms a cell)
		const long c = nslice == 1 ? it : (long)((unsigned)it / (unsigned)nslice);
		const int sl = (int)(it - c * nslice);
		if (shift)
			sepk_build_cell<true, EXCL>(L, s, c, sl, nslice, start, g, qt, npad, cap,
						    (int)n, nbr, cnt, sys, mymax);
		else
			sepk_build_cell<false, EXCL>(L, s, c, sl, nslice, start, g, qt, npad, cap,
						     (int)n, nbr, cnt, sys, mymax);
	}
	// one wave per cell, SEPK_BT / 64 cells at a time
	const long cper = (ncell + SEPK_NXCD - 1) / SEPK_NXCD;
	const long lo = xcd * cper, hi = min(lo + cper, ncell);
	const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
	for (long c = lo + k + wv * per; c < hi; c += (SEPK_BT / 64) * per) {
		const int a = start[c], b = start[c + 1];
		for (int t = a + ln; t < b; t += 64)
			sepk_copy_home(s, home, t, npad, has_partners);
	}
	for (int o = 32; o > 0; o >>= 1)
		mymax = max(mymax, __shfl_xor(mymax, o, 64));
	__shared__ int bmax;
	if (threadIdx.x == 0)
		bmax = 0;
	__syncthreads();
	if ((threadIdx.x & 63) == 0)
		atomicMax(&bmax, mymax);
	__syncthreads();
	if (threadIdx.x == 0 && bmax > 0)
		atomicMax(&sys->max_full, bmax);
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

extern "C" int sepdev_list_limit(sepdev *d, int cap, int *applied)
{
	if (cap < 0)
		return sepk_fail(SEP_GPU_EARG, "list limit %d", cap);
	// whole batches: a truncated list then ends on a batch boundary, and
	// the pair kernels read no entry the build did not write
	if (cap > 0) {
		cap = cap / SEPK_PAD * SEPK_PAD;
		if (cap < SEPK_PAD)
			cap = SEPK_PAD;
	}
	d->cap_limit = cap;
	if (cap > 0 && (!d->nbr || cap > d->cap)) {
		int rc = sepdev_reserve_list(d, cap);
		if (rc)
			return rc;
	}
	if (applied)
		*applied = cap;
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
	// a recorded pair call reads the current list: launch it before the
	// list can be resized or rebuilt
	int rc = sepk_flush_pair(d);
	if (rc)
		return rc;
	rc = sepk_flush_nh(d);
	if (rc)
		return rc;
	// a rebuild permutes the atoms: a fork point recorded before it (also
	// by the flushed pair call above) must not serve the terms after it
	d->fork_ready = 0;

	// size buffers on the host side of any capture
	if (ncell > d->ncell_cap) {
		if (d->capturing)
			return sepk_fail(SEP_GPU_ESTATE, "cell grid grew during graph capture");
		(void)hipFree(d->cell_count);
		(void)hipFree(d->cell_start);
		SEPK_CHECK(hipMalloc(&d->cell_count, ncell * sizeof(int)));
		SEPK_CHECK(hipMalloc(&d->cell_start, (ncell + 1) * sizeof(int)));
		SEPK_CHECK(hipMemsetAsync(d->cell_count, 0, ncell * sizeof(int), d->st));
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
	// half a wave per cell when the cells hold 32 atoms or fewer on
	// average, a wave up to 64, a workgroup beyond
	const bool half = d->n <= 32 * ncell, wg = d->n > 64 * ncell;
	long sb = wg ? ncell : (ncell * (half ? 32 : 64) + SEPK_BLOCK - 1) / SEPK_BLOCK;
	if (sb > SEPK_GRID_CAP)
		sb = SEPK_GRID_CAP;
	k_bin<<<nb, SEPK_BLOCK, 0, d->st>>>(cur.pos, d->n, g, d->cellid_old, d->cell_slot,
					    d->cell_count, d->sys,
					    conditional);
	if (ncell <= SEPK_PLACE_CELLS) {
		// one launch: each workgroup scans the counts itself
		long pb = (d->n + 1023) / 1024;
		if (pb > 256)
			pb = 256;
		k_scan_scatter<<<(unsigned)pb, 1024, 0, d->st>>>(
			d->cell_count, d->cell_start, d->cellid_old, d->cell_slot, d->order_tmp, ncell,
			d->n, d->sys, conditional);
	} else {
		k_scan<<<1, 1024, 0, d->st>>>(d->cell_count, d->cell_start, ncell, d->sys,
					      conditional);
		k_scatter<<<nb, SEPK_BLOCK, 0, d->st>>>(d->cellid_old, d->cell_slot,
							d->cell_start, d->order_tmp, d->n, d->sys,
							conditional);
	}
	(wg ? k_sort_permute_wg : half ? k_sort_permute<32> : k_sort_permute<64>)<<<(unsigned)sb, SEPK_BLOCK, 0, d->st>>>(
		d->order_tmp, d->cell_start, cur, spare, d->iperm, ncell, d->npad,
		d->sys, conditional, hp, d->cell_count);
	SEPK_LAUNCH_CHECK();
	// Home atoms of a cell split over nslice workgroups when the cells are
	// few and full (small boxes, dense molecular systems): at least 16 home
	// atoms a slice, at most SEPK_BHOME, about 1024 items where cells alone
	// are fewer.  The lists do not depend on the split.
	const long avg = (d->n + ncell - 1) / ncell;
	long nslice = (avg + 15) / 16;
	const long fill = (1024 + ncell - 1) / ncell;
	if (nslice > fill)
		nslice = fill;
	if (nslice < (avg + SEPK_BHOME - 1) / SEPK_BHOME)
		nslice = (avg + SEPK_BHOME - 1) / SEPK_BHOME;
	if (nslice < 1)
		nslice = 1;
	if (nslice > 64)
		nslice = 64;
	long bb = (ncell * nslice + SEPK_NXCD - 1) / SEPK_NXCD * SEPK_NXCD;
	if (bb > SEPK_BGRID)
		bb = SEPK_BGRID;
	auto kb = g.opt == 2 ? k_build<2> : g.opt == 3 ? k_build<3> : k_build<0>;
	const int capb = d->cap_limit > 0 && d->cap_limit < d->cap ? d->cap_limit : d->cap;
	kb<<<(unsigned)bb, SEPK_BT, 0, d->st>>>(spare, cur, d->cell_start, g, ncell, (int)nslice,
						      d->n, d->npad, capb, d->nbr, d->nbr_count,
						      d->sys, conditional, hp);
	SEPK_LAUNCH_CHECK();
	// clears the rebuild flag and the owed xn reset
	sepk_task f;
	memset(&f, 0, sizeof f);
	f.mode = SEPK_FIN_BUILD;
	f.region = SEPK_R_BUILD;
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
