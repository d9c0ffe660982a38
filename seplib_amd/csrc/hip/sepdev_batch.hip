// sepdev_batch.hip -- many independent Lennard-Jones systems advanced
// together: one persistent launch per step for all of them.
//
// Each system is the reference's prg1 loop (prgs/prg1.c:55-65) on its own
// box: sep_reset_retval, sep_reset_force, sep_force_pairs("AA", cf,
// sep_lj_shift) with the list rebuilt when flagged, sep_nosehoover and
// sep_leapfrog.  All systems' atoms are concatenated on the device (system s
// owns slots off[s] .. off[s] + npad[s], npad = a multiple of 64 above n,
// slot n being a phantom with a NaN position that pads lists).  Per step:
//
//   k_bcount      every atom of a flagged system: its cell and its slot
//                 among the cell's members (global counters)
//   k_bscan       one workgroup per flagged system: scan of the counts
//                 into cell starts
//   k_bscatter    the members of every cell in slot order
//                 (batches of more than SEPB_FMAP systems: k_bbin, one
//                 workgroup per flagged system with LDS counters, instead
//                 of these three)
//   k_bsort       one wave per cell of a flagged system: rank the members
//                 by original index (the new order, independent of the
//                 atomic order that listed them) and move their state to it
//                 in the spare copy
//   k_bbuild      one workgroup per cell of a flagged system: the
//                 candidates of the distinct neighbour cells staged in LDS,
//                 tested 64 at a time against each home atom (ballots kept
//                 in LDS), then written to the lists in candidate order
//   k_bstep       persistent: the grid walks a table of tiles (a system and
//                 a run of its atoms, lanes per atom chosen per system);
//                 pair forces, Nose-Hoover and leapfrog for the tile's
//                 atoms and one row of block sums per tile
//   k_bfinish     one workgroup per system: its rows reduced in tile order,
//                 then energies, stress, thermostat, the state copy and
//                 position slot in use, and the rebuild flag.
//
// The rebuild launches return at once for systems whose flag is clear, so
// the launches of a step replay from one captured graph.  Workgroups never wait for
// each other inside a launch: on this chip a device-scope fence writes back
// and invalidates an XCD's L2, so every combination across workgroups
// happens at a kernel boundary instead.  Positions are double
// buffered inside a state copy (a tile reads slot p and writes slot 1-p;
// other atoms' tiles still read slot p), every other per-atom field is
// written only by its own atom.  Unlike the reference's cell-list builder,
// a system with fewer than three cells on an axis visits each distinct
// neighbour cell once, so every pair is listed once from each end.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sepdev_impl.h"

#define SEPB_BLOCK 256
#define SEPB_NV 16
#define SEPB_BATCH 4

// One system on the device.  Read by every kernel; written only by the
// finishing tile of k_bstep, the last cell of k_bbuild and the host.
struct sepb_sys {
	long off;         // first slot
	long list_off;    // first list slot (cap * off)
	int n, npad;
	int sub;          // lanes per atom of its tiles and its list layout
	int ntile, row0;  // its tiles' partial rows: row0 .. row0 + ntile - 1
	int ncell[3], cell_off;  // cell_start slots: cell_off .. + ncells
	double lcell[3], L[3], iL[3];
	double ilcell[3]; // 1 / lcell
	double rlist2;
	double dt, Td, dt_tau2, skin;
	int cur, p;       // state copy in use, position slot in use
	unsigned neighb_flag, nupdate;
	int error, max_full;
	int unclean;      // the last binning saw an atom outside [0, L)
	double alpha, sumv2, max_dist2, tnow;
	double ret[SEPDEV_RET_WORDS];
};

// One copy of every system's per-atom state, concatenated.
struct sepb_state {
	sepk_pos *pos[2];
	double *v, *f, *a, *xn; // x, y, z blocks of `slots` doubles
	int *cn, *cr;           // x, y, z blocks of `slots` ints
	double *m;
	int *orig;              // original index within the system
};

struct sepb_lj {
	double cf2, A, nB, C, nD, ushift;
	float Af, nBf, Cf, nDf, ushiftf;
};

struct sepb_view {
	sepb_sys *sys;
	const sepb_state *st; // two copies, device memory
	long slots;
	int nsys;
	int *cellid, *cslot, *otmp;
	int *gcount;          // per-cell member counts of the binning (0 between)
	int *outside;         // per system: an atom outside the box (0 between)
	int *cell_start;
	const int *cell_sys;  // system of every global cell index
	long ncell_tot;
	int *nbr, *nbr_count;
	int cap;
	double *partial;
	long prow;
};

struct sepdev_batch {
	int nsys;
	long slots, ncell_tot;
	hipStream_t st;
	sepb_state S[2];
	sepb_state *dS;
	sepb_sys *dsys;
	sepb_sys *hsys; // host copy of the descriptors as last uploaded
	int *cellid, *cslot, *otmp, *cell_start, *cell_sys;
	int *gcount, *outside;
	int *nbr, *nbr_count;
	int cap;
	double *partial;
	long prow;
	int4 *tiles; // full table
	int ntiles;
	int4 *sub_tiles; // a subset (autotuner)
	int grid;
	sepb_lj lj;
	hipGraph_t graph;
	hipGraphExec_t exec;
	int graph_mixed, graph_steps;
};

__device__ __forceinline__ int sepb_mod(int a, int n)
{
	const int r = a % n;
	return r < 0 ? r + n : r;
}

// ---- rebuild ------------------------------------------------------------------

#define SEPB_BIN_THREADS 1024
#define SEPB_MAX_CELLS 8192 // cells per system (LDS counters)
#define SEPB_BIN_UNROLL 8

// Cell of position q in system y (floor of x / l, by the reciprocal edge il,
// and periodic wrap); an atom outside [0, L) on some axis (x == L after a
// wrap, or a starting position) sets `outside`, which rules out the
// per-segment image shifts of the build.  Rounding of x il can put an atom
// within an ulp of a cell face into the next cell; its partners across the
// face are then at most one cell away, and only a pair of two such atoms
// at nearly the list cutoff (beyond the force cutoff) could be missed.
__device__ __forceinline__ int sepb_cell(sepb_sys *y, const double (&il)[3], const sepk_pos &q,
					 int &outside)
{
	const double xs[3] = {q.x, q.y, q.z};
	int ci[3];
	for (int k = 0; k < 3; k++) {
		const double w = xs[k] * il[k];
		if (!(fabs(w) < 1.0e9)) {
			atomicCAS(&y->error, 0, SEP_GPU_ECELL);
			ci[k] = 0;
		} else {
			const int f = (int)floor(w);
			if (f < 0 || f >= y->ncell[k])
				outside = 1;
			ci[k] = sepb_mod(f, y->ncell[k]);
		}
	}
	return ci[0] + y->ncell[0] * (ci[1] + y->ncell[1] * ci[2]);
}

// One workgroup per flagged system: the cell of every atom, its slot among
// the cell's members (LDS atomics), the exclusive scan of the counts into
// cell_start (one extra entry: n), and otmp[off + start + slot] = atom.
__global__ __launch_bounds__(SEPB_BIN_THREADS) void k_bbin(sepb_view b)
{
	__shared__ int cnt[SEPB_MAX_CELLS];
	__shared__ int wsum[SEPB_BIN_THREADS / 64];
	sepb_sys *y = b.sys + blockIdx.x;
	if (!y->neighb_flag)
		return;
	const int nc = y->ncell[0] * y->ncell[1] * y->ncell[2];
	const long off = y->off;
	const int n = y->n;
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	__shared__ int outside;
	for (int c = threadIdx.x; c < nc; c += SEPB_BIN_THREADS)
		cnt[c] = 0;
	if (threadIdx.x == 0)
		outside = 0;
	__syncthreads();
	const sepk_pos *pos = b.st[y->cur].pos[y->p] + off;
	const double il[3] = {y->ilcell[0], y->ilcell[1], y->ilcell[2]};
	// SEPB_BIN_UNROLL positions per thread are loaded before any is
	// binned, so a large system pays a few memory round trips, not one
	// per SEPB_BIN_THREADS atoms
	for (int t0 = threadIdx.x; t0 < n; t0 += SEPB_BIN_THREADS * SEPB_BIN_UNROLL) {
		sepk_pos qq[SEPB_BIN_UNROLL];
#pragma unroll
		for (int u = 0; u < SEPB_BIN_UNROLL; u++) {
			const int t = t0 + u * SEPB_BIN_THREADS;
			if (t < n)
				qq[u] = pos[t];
		}
#pragma unroll
		for (int u = 0; u < SEPB_BIN_UNROLL; u++) {
			const int t = t0 + u * SEPB_BIN_THREADS;
			if (t < n) {
				const int c = sepb_cell(y, il, qq[u], outside);
				b.cellid[off + t] = c;
				b.cslot[off + t] = atomicAdd(&cnt[c], 1);
			}
		}
	}
	__syncthreads();
	// exclusive scan of cnt in place, SEPB_BIN_THREADS cells at a time
	int *st = b.cell_start + y->cell_off;
	int carry = 0;
	for (int base = 0; base < nc; base += SEPB_BIN_THREADS) {
		const int c = base + threadIdx.x;
		const int v = c < nc ? cnt[c] : 0;
		int incl = v;
		for (int o = 1; o < 64; o <<= 1) {
			const int u = __shfl_up(incl, o, 64);
			if (lane >= o)
				incl += u;
		}
		if (lane == 63)
			wsum[wid] = incl;
		__syncthreads();
		int before = carry, tot = 0;
		for (int w = 0; w < SEPB_BIN_THREADS / 64; w++) {
			if (w < wid)
				before += wsum[w];
			tot += wsum[w];
		}
		if (c < nc) {
			cnt[c] = before + incl - v;
			st[c] = before + incl - v;
		}
		carry += tot;
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		st[nc] = carry;
		y->unclean = outside;
	}
	for (int t0 = threadIdx.x; t0 < n; t0 += SEPB_BIN_THREADS * SEPB_BIN_UNROLL) {
		int cc[SEPB_BIN_UNROLL], ss[SEPB_BIN_UNROLL];
#pragma unroll
		for (int u = 0; u < SEPB_BIN_UNROLL; u++) {
			const int t = t0 + u * SEPB_BIN_THREADS;
			cc[u] = t < n ? b.cellid[off + t] : 0;
			ss[u] = t < n ? b.cslot[off + t] : 0;
		}
#pragma unroll
		for (int u = 0; u < SEPB_BIN_UNROLL; u++) {
			const int t = t0 + u * SEPB_BIN_THREADS;
			if (t < n)
				b.otmp[off + cnt[cc[u]] + ss[u]] = t;
		}
	}
}


// Rank of `key` among the original indices of the m members seg[0..m) of a
// cell (all distinct), by the whole wave: the members' new order within the
// cell, independent of the atomic order that listed them.
__device__ __forceinline__ int sepb_rank(const int *seg, const int *orig, int m, int key)
{
	const int lane = threadIdx.x & 63;
	int r = 0;
	for (int cb = 0; cb < m; cb += 64) {
		const int kk = cb + lane < m ? orig[seg[cb + lane]] : 0x7fffffff;
		const int lim = min(m - cb, 64);
		for (int j = 0; j < lim; j++)
			r += __shfl(kk, j, 64) < key;
	}
	return r;
}

// The distinct neighbour-cell offsets of an axis with `nc` cells.
__device__ __forceinline__ int sepb_offsets(int nc, int *o)
{
	if (nc >= 3) {
		o[0] = -1;
		o[1] = 0;
		o[2] = 1;
		return 3;
	}
	o[0] = 0;
	o[1] = 1;
	return nc;
}

// The cells of the flagged systems, numbered in system order, are the work
// of k_bsort and k_bbuild: fpre[s] is the number of such cells before
// system s (fpre[nsys]: all of them), scanned by each workgroup into LDS.
// Batches of more than SEPB_FMAP systems walk every cell instead.
#define SEPB_FMAP 1024
__device__ __forceinline__ void sepb_flag_prefix(const sepb_view &b, int *fpre,
						 bool atoms = false)
{
	__shared__ int wtot[SEPB_BLOCK / 64];
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	int carry = 0;
	for (int base = 0; base < b.nsys; base += SEPB_BLOCK) {
		const int s = base + threadIdx.x;
		int v = 0;
		if (s < b.nsys) {
			const sepb_sys *y = b.sys + s;
			v = !y->neighb_flag ? 0 : atoms ? y->n : y->ncell[0] * y->ncell[1] * y->ncell[2];
		}
		int incl = v;
		for (int o = 1; o < 64; o <<= 1) {
			const int u = __shfl_up(incl, o, 64);
			if (lane >= o)
				incl += u;
		}
		if (lane == 63)
			wtot[wid] = incl;
		__syncthreads();
		int before = carry, tot = 0;
		for (int w = 0; w < SEPB_BLOCK / 64; w++) {
			if (w < wid)
				before += wtot[w];
			tot += wtot[w];
		}
		if (s < b.nsys)
			fpre[s] = before + incl - v;
		carry += tot;
		__syncthreads();
	}
	if (threadIdx.x == 0)
		fpre[b.nsys] = carry;
	__syncthreads();
}

// The system of flagged item w (cell or atom, by the prefix fpre): the last
// system whose items start at or before w.
__device__ __forceinline__ int sepb_flag_system(const int *fpre, int nsys, long w)
{
	int lo = 0, hi = nsys;
	while (hi - lo > 1) {
		const int mid = (lo + hi) >> 1;
		if (fpre[mid] <= w)
			lo = mid;
		else
			hi = mid;
	}
	return lo;
}

// Work item w of a rebuild kernel -> (system, cell); false if the cell's
// system is not flagged (only when every cell is walked).
__device__ __forceinline__ bool sepb_work_cell(const sepb_view &b, const int *fpre, bool mapped,
					       long w, int &s, int &c)
{
	if (mapped) {
		s = sepb_flag_system(fpre, b.nsys, w);
		c = (int)(w - fpre[s]);
		return true;
	}
	s = b.cell_sys[w];
	if (s < 0 || !b.sys[s].neighb_flag)
		return false; // the scan's extra entry of a system, or no rebuild
	c = (int)(w - b.sys[s].cell_off);
	return true;
}

// Binning of the flagged systems' atoms across the whole grid, for batches
// of up to SEPB_FMAP systems (k_bbin, one workgroup per system, serves
// larger ones): k_bcount gives every atom its cell and its slot among the
// cell's members (global counters, one atomic add per distinct cell of a
// wave), k_bscan turns each flagged system's counts into cell starts (one
// workgroup per system) and k_bscatter lists the members of every cell.

// Flagged atom w -> (system, atom); the atoms are numbered in system order.
__global__ __launch_bounds__(SEPB_BLOCK) void k_bcount(sepb_view b)
{
	__shared__ int apre[SEPB_FMAP + 1];
	sepb_flag_prefix(b, apre, true);
	const long total = apre[b.nsys];
	const int lane = threadIdx.x & 63;
	const unsigned long long below = (1ull << lane) - 1ull;
	const long stride = (long)gridDim.x * blockDim.x;
	// whole waves iterate together (the bound is rounded up to a wave)
	const long tw = (total + 63) / 64 * 64;
	for (long w = (long)blockIdx.x * blockDim.x + threadIdx.x; w < tw; w += stride) {
		const bool act = w < total;
		int gc = 0, s = 0;
		long at = 0;
		if (act) {
			s = sepb_flag_system(apre, b.nsys, w);
			sepb_sys *y = b.sys + s;
			const int t = (int)(w - apre[s]);
			at = y->off + t;
			const double il[3] = {y->ilcell[0], y->ilcell[1], y->ilcell[2]};
			int outside = 0;
			const int c = sepb_cell(y, il, b.st[y->cur].pos[y->p][at], outside);
			if (outside)
				atomicOr(&b.outside[s], 1);
			b.cellid[at] = c;
			gc = y->cell_off + c;
		}
		unsigned long long todo = __ballot(act);
		int my = 0;
		while (todo) {
			const int lead = __ffsll((long long)todo) - 1;
			const int cl = __shfl(gc, lead, 64);
			const unsigned long long same = __ballot(act && gc == cl);
			int base = 0;
			if (lane == lead)
				base = atomicAdd(&b.gcount[cl], __popcll(same));
			base = __shfl(base, lead, 64);
			if (act && gc == cl)
				my = base + __popcll(same & below);
			todo &= ~same;
		}
		if (act)
			b.cslot[at] = my;
	}
}

// One workgroup per flagged system: the exclusive scan of its cells' counts
// into cell_start (one extra entry: n), the counts and the outside flag
// zeroed for the next rebuild, y->unclean set from the flag.
__global__ __launch_bounds__(SEPB_BIN_THREADS) void k_bscan(sepb_view b)
{
	__shared__ int wsum[SEPB_BIN_THREADS / 64];
	sepb_sys *y = b.sys + blockIdx.x;
	if (!y->neighb_flag)
		return;
	const int nc = y->ncell[0] * y->ncell[1] * y->ncell[2];
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	int *cnt = b.gcount + y->cell_off;
	int *st = b.cell_start + y->cell_off;
	int carry = 0;
	for (int base = 0; base < nc; base += SEPB_BIN_THREADS) {
		const int c = base + threadIdx.x;
		const int v = c < nc ? cnt[c] : 0;
		int incl = v;
		for (int o = 1; o < 64; o <<= 1) {
			const int u = __shfl_up(incl, o, 64);
			if (lane >= o)
				incl += u;
		}
		if (lane == 63)
			wsum[wid] = incl;
		__syncthreads();
		int before = carry, tot = 0;
		for (int w = 0; w < SEPB_BIN_THREADS / 64; w++) {
			if (w < wid)
				before += wsum[w];
			tot += wsum[w];
		}
		if (c < nc) {
			st[c] = before + incl - v;
			cnt[c] = 0;
		}
		carry += tot;
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		st[nc] = carry;
		y->unclean = b.outside[blockIdx.x];
		b.outside[blockIdx.x] = 0;
	}
}

// Every flagged atom into its cell's member list: otmp[off + start + slot].
__global__ __launch_bounds__(SEPB_BLOCK) void k_bscatter(sepb_view b)
{
	__shared__ int apre[SEPB_FMAP + 1];
	sepb_flag_prefix(b, apre, true);
	const long total = apre[b.nsys];
	const long stride = (long)gridDim.x * blockDim.x;
	for (long w = (long)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += stride) {
		const int s = sepb_flag_system(apre, b.nsys, w);
		const sepb_sys *y = b.sys + s;
		const int t = (int)(w - apre[s]);
		const long at = y->off + t;
		b.otmp[y->off + b.cell_start[y->cell_off + b.cellid[at]] + b.cslot[at]] = t;
	}
}

// One wave per cell (grid-stride): the new order of the cell's members and
// their state moved to it.
__global__ __launch_bounds__(SEPB_BLOCK) void k_bsort(sepb_view b)
{
	__shared__ int fpre[SEPB_FMAP + 1];
	const bool mapped = b.nsys <= SEPB_FMAP;
	long total = b.ncell_tot;
	if (mapped) {
		sepb_flag_prefix(b, fpre);
		total = fpre[b.nsys];
	}
	const int lane = threadIdx.x & 63;
	const long nw = ((long)gridDim.x * blockDim.x) >> 6;
	for (long w = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < total; w += nw) {
		int s, c;
		if (!sepb_work_cell(b, fpre, mapped, w, s, c))
			continue;
		const sepb_sys *y = b.sys + s;
		const long off = y->off, slots = b.slots;
		const int *st = b.cell_start + y->cell_off;
		const int h0 = st[c], m = st[c + 1] - h0;
		const sepb_state &So = b.st[y->cur];
		const sepb_state &Sn = b.st[1 - y->cur];
		const int *seg = b.otmp + off + h0;
		const int *korig = So.orig + off;
		for (int hb = 0; hb < m; hb += 64) {
			const bool v = hb + lane < m;
			const int o = v ? seg[hb + lane] : 0;
			const int r = sepb_rank(seg, korig, m, v ? korig[o] : 0x7fffffff);
			if (!v)
				continue;
			// xn = x and cross_neighb = 0 go with a rebuild
			// (source/sepintgr.c:72-84)
			// every field is loaded before any is stored (one memory
			// round trip: the copies may alias as far as the compiler
			// knows)
			const long go = off + o, gn = off + h0 + r;
			const sepk_pos q = So.pos[y->p][go];
			double vv[3], ff[3], aa[3];
			int cr[3];
#pragma unroll
			for (int k = 0; k < 3; k++) {
				vv[k] = So.v[k * slots + go];
				ff[k] = So.f[k * slots + go];
				aa[k] = So.a[k * slots + go];
				cr[k] = So.cr[k * slots + go];
			}
			const double mm = So.m[go];
			const int og = So.orig[go];
			Sn.pos[0][gn] = q;
#pragma unroll
			for (int k = 0; k < 3; k++) {
				Sn.v[k * slots + gn] = vv[k];
				Sn.f[k * slots + gn] = ff[k];
				Sn.a[k * slots + gn] = aa[k];
				Sn.cr[k * slots + gn] = cr[k];
				Sn.cn[k * slots + gn] = 0;
			}
			Sn.xn[gn] = q.x;
			Sn.xn[slots + gn] = q.y;
			Sn.xn[2 * slots + gn] = q.z;
			Sn.m[gn] = mm;
			Sn.orig[gn] = og;
		}
	}
}

// One workgroup per cell of a flagged system (the flagged cells' work
// list).  Wave 0 lays out the candidate segments (one per distinct
// neighbour cell: first new index, prefix of lengths, and the image shift
// that brings the cell next to the home cell); every thread then stages
// candidates -- position with the shift added, new index -- in LDS, up to
// SEPB_BCHUNK at a time, and the home atoms beside them.  Test pass: wave w
// owns home atoms w, w + 4, ... and tests 64 register-held candidates per
// round against each, keeping the ballot in LDS.  Emit pass: one thread per
// (home atom, 64-candidate word) finds the word's first entry number from
// the atom's running length and the earlier words' popcounts and writes
// the set bits in order, so every list is in candidate order.
#define SEPB_BCHUNK 384
#define SEPB_BWORDS (SEPB_BCHUNK / 64)
#define SEPB_BHOME 64
#define SEPB_BGRID 1792 // build workgroups (7 per CU)
static_assert(SEPB_BCHUNK % 64 == 0, "candidates are tested 64 at a time");

struct sepb_build_lds {
	double cx[SEPB_BCHUNK], cy[SEPB_BCHUNK], cz[SEPB_BCHUNK];
	int cj[SEPB_BCHUNK];
	double hx[SEPB_BHOME], hy[SEPB_BHOME], hz[SEPB_BHOME];
	unsigned long long vote[SEPB_BHOME][SEPB_BWORDS];
	int hcnt[2][SEPB_BHOME]; // list lengths before / after the chunk
	int segjs[27], segpre[28];
	double segsh[27][3];
	int wmax[SEPB_BLOCK / 64];
};

// Tests the staged candidates [q0, q0 + 64) of a chunk of kn against the
// wave's home atoms (hbase + i for i = wid, wid + 4, ... < hn) and keeps
// each ballot in L.vote[i][q0 / 64].  SHIFT: the staged positions carry
// their segment's image shift and r^2 needs no wrap; otherwise the
// single-image wrap of every difference.
template <bool SHIFT>
__device__ __forceinline__ void sepb_build_test(sepb_build_lds &L, int q0, int kn, int hbase,
						int hn, double L0, double L1, double L2,
						double rl2)
{
	const int lane = threadIdx.x & 63;
	const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); // wave-uniform
	const int q = min(q0 + lane, kn - 1);
	const bool valid = q0 + lane < kn;
	const double cx = L.cx[q], cy = L.cy[q], cz = L.cz[q];
	const int j = L.cj[q];
	unsigned vlo = 0u, vhi = 0u; // lane kk keeps the ballot of the kk-th atom
	int kk = 0;
	for (int i = wid; i < hn; i += SEPB_BLOCK / 64, kk++) {
		double ddx = L.hx[i] - cx;
		double ddy = L.hy[i] - cy;
		double ddz = L.hz[i] - cz;
		if (!SHIFT) {
			ddx = sepk_wrap(ddx, L0);
			ddy = sepk_wrap(ddy, L1);
			ddz = sepk_wrap(ddz, L2);
		}
		const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
		const bool hit = valid & (r2 < rl2) & (j != hbase + i);
		const unsigned long long vote = __ballot(hit);
		if (lane == kk) {
			vlo = (unsigned)vote;
			vhi = (unsigned)(vote >> 32);
		}
	}
	const int i = wid + (SEPB_BLOCK / 64) * lane;
	if (lane < kk && i < hn)
		L.vote[i][q0 / 64] = ((unsigned long long)vhi << 32) | vlo;
}

// Writes the accepted candidates of a chunk of kn to the lists of home
// atoms hbase .. hbase + hn - 1 (list slot of entry e of atom h:
// (e >> lg) * lstride + h * sub + (e & (sub - 1))).  Lengths before the
// chunk are in L.hcnt[par]; the last word of each atom leaves the length
// after it in L.hcnt[par ^ 1].  Entries at or beyond cap are counted only.
__device__ __forceinline__ void sepb_build_emit(sepb_build_lds &L, int kn, int hn, int par,
						int hbase, int sub, int lg, int lstride, int cap,
						int *lst)
{
	const int nw = (kn + 63) / 64;
	for (int it = threadIdx.x; it < hn * nw; it += SEPB_BLOCK) {
		const int i = it / nw, w = it - i * nw;
		int e = L.hcnt[par][i];
		for (int u = 0; u < w; u++)
			e += __popcll(L.vote[i][u]);
		unsigned long long m = L.vote[i][w];
		if (w == nw - 1)
			L.hcnt[par ^ 1][i] = e + __popcll(m);
		const int h = hbase + i;
		const int *cj = L.cj + w * 64;
		for (; m; m &= m - 1, e++)
			if (e < cap)
				lst[(e >> lg) * lstride + h * sub + (e & (sub - 1))] =
					cj[__builtin_ctzll(m)];
	}
}

__global__ __launch_bounds__(SEPB_BLOCK) __attribute__((amdgpu_waves_per_eu(7))) void k_bbuild(sepb_view b)
{
	__shared__ sepb_build_lds L;
	__shared__ int fpre[SEPB_FMAP + 1];
	const bool mapped = b.nsys <= SEPB_FMAP;
	long total = b.ncell_tot;
	if (mapped) {
		sepb_flag_prefix(b, fpre);
		total = fpre[b.nsys];
	}
	const int lane = threadIdx.x & 63;
	const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); // wave-uniform
	for (long w = blockIdx.x; w < total; w += gridDim.x) {
		int s, c;
		if (!sepb_work_cell(b, fpre, mapped, w, s, c))
			continue; // the same for the whole workgroup
		sepb_sys *y = b.sys + s;
		const int nx = y->ncell[0], ny = y->ncell[1], nz = y->ncell[2];
		const long off = y->off;
		const int n = y->n, npad = y->npad, sub = y->sub, cap = b.cap;
		const int lg = __ffs(sub) - 1;
		const int *st = b.cell_start + y->cell_off;
		// the new order, written by k_bsort
		const sepk_pos *pn = b.st[1 - y->cur].pos[0] + off;
		const double L0 = y->L[0], L1 = y->L[1], L2 = y->L[2];
		const double rl2 = y->rlist2;
		const int h0 = st[c], m = st[c + 1] - h0;
		int *lst = b.nbr + y->list_off;
		// a system's list region holds cap * npad < 2^31 entries
		const int lstride = npad * sub;
		const bool shift = nx >= 3 && ny >= 3 && nz >= 3 && !y->unclean;
		// the distinct neighbour cells, in a fixed order
		int ox[3], oy[3], oz[3];
		const int nox = sepb_offsets(nx, ox), noy = sepb_offsets(ny, oy),
			  noz = sepb_offsets(nz, oz);
		const int nseg = nox * noy * noz;
		__syncthreads(); // the previous cell's readers are done with the LDS
		if (wid == 0) {
			int len = 0;
			if (lane < nseg) {
				const int cxi = c % nx, cyi = (c / nx) % ny, czi = c / (nx * ny);
				const int ux = cxi + ox[lane % nox], uy = cyi + oy[(lane / nox) % noy],
					  uz = czi + oz[lane / (nox * noy)];
				const int c2 =
					sepb_mod(ux, nx) + nx * (sepb_mod(uy, ny) + ny * sepb_mod(uz, nz));
				L.segjs[lane] = st[c2];
				L.segsh[lane][0] = ux >= nx ? L0 : (ux < 0 ? -L0 : 0.0);
				L.segsh[lane][1] = uy >= ny ? L1 : (uy < 0 ? -L1 : 0.0);
				L.segsh[lane][2] = uz >= nz ? L2 : (uz < 0 ? -L2 : 0.0);
				len = st[c2 + 1] - st[c2];
			}
			int incl = len;
			for (int o = 1; o < 64; o <<= 1) {
				const int u = __shfl_up(incl, o, 64);
				if (lane >= o)
					incl += u;
			}
			if (lane < nseg)
				L.segpre[lane + 1] = incl;
			if (lane == 0)
				L.segpre[0] = 0;
		}
		__syncthreads();
		const int T = L.segpre[nseg];
		int full = 0;
		for (int hb = 0; hb < m; hb += SEPB_BHOME) {
			const int hn = min(m - hb, SEPB_BHOME);
			if (threadIdx.x < SEPB_BHOME)
				L.hcnt[0][threadIdx.x] = 0;
			int par = 0;
			for (int k0 = 0; k0 < T; k0 += SEPB_BCHUNK, par ^= 1) {
				const int kn = min(T - k0, SEPB_BCHUNK);
				if (k0 > 0)
					__syncthreads(); // the previous chunk is no longer read
				// stage [k0, k0 + kn): thread x takes x, x + SEPB_BLOCK, ...;
				// its loads are issued before any LDS store
				constexpr int PER = (SEPB_BCHUNK + SEPB_BLOCK - 1) / SEPB_BLOCK;
				int jq[PER], rq[PER];
				sepk_pos pq[PER];
#pragma unroll
				for (int r = 0; r < PER; r++) {
					const int q = threadIdx.x + r * SEPB_BLOCK;
					const int k = k0 + q;
					int o = 0; // the last segment starting at or before k
#pragma unroll
					for (int step = 16; step > 0; step >>= 1)
						if (o + step < nseg && L.segpre[o + step] <= k)
							o += step;
					rq[r] = q < kn ? o : -1;
					jq[r] = L.segjs[o] + (k - L.segpre[o]);
				}
#pragma unroll
				for (int r = 0; r < PER; r++)
					if (rq[r] >= 0)
						pq[r] = pn[jq[r]];
#pragma unroll
				for (int r = 0; r < PER; r++) {
					if (rq[r] < 0)
						continue;
					const int q = threadIdx.x + r * SEPB_BLOCK;
					double px = pq[r].x, py = pq[r].y, pz = pq[r].z;
					if (shift) {
						px += L.segsh[rq[r]][0];
						py += L.segsh[rq[r]][1];
						pz += L.segsh[rq[r]][2];
					}
					L.cx[q] = px;
					L.cy[q] = py;
					L.cz[q] = pz;
					L.cj[q] = jq[r];
				}
				if (k0 == 0 && threadIdx.x < hn) {
					const sepk_pos q = pn[h0 + hb + threadIdx.x];
					L.hx[threadIdx.x] = q.x;
					L.hy[threadIdx.x] = q.y;
					L.hz[threadIdx.x] = q.z;
				}
				__syncthreads();
				for (int q0 = 0; q0 < kn; q0 += 64) {
					if (shift)
						sepb_build_test<true>(L, q0, kn, h0 + hb, hn, L0, L1, L2, rl2);
					else
						sepb_build_test<false>(L, q0, kn, h0 + hb, hn, L0, L1, L2, rl2);
				}
				__syncthreads();
				sepb_build_emit(L, kn, hn, par, h0 + hb, sub, lg, lstride, cap, lst);
			}
			__syncthreads();
			if (threadIdx.x < hn) {
				const int h = h0 + hb + threadIdx.x;
				int cnt = L.hcnt[par][threadIdx.x];
				full = max(full, cnt);
				if (cnt > cap) {
					atomicCAS(&y->error, 0, SEP_GPU_ENEIGHB);
					cnt = cap;
				}
				// pad to whole batches of every lane with the phantom
				const int qq = sub * SEPB_BATCH;
				const int pad = min((cnt + qq - 1) / qq * qq, cap);
				for (int e = cnt; e < pad; e++)
					lst[(e >> lg) * lstride + h * sub + (e & (sub - 1))] = n;
				b.nbr_count[off + h] = cnt;
			}
			if (hb + SEPB_BHOME < m)
				__syncthreads(); // the next batch restages the candidates
		}
		// one atomic per cell: every home atom of a system hitting one
		// address would serialise in L2
		for (int o = 32; o > 0; o >>= 1)
			full = max(full, __shfl_xor(full, o, 64));
		if (lane == 0)
			L.wmax[wid] = full;
		__syncthreads();
		if (threadIdx.x == 0) {
			int mx = 0;
			for (int v = 0; v < SEPB_BLOCK / 64; v++)
				mx = max(mx, L.wmax[v]);
			if (mx > 0)
				atomicMax(&y->max_full, mx);
		}
	}
}

// ---- the step ------------------------------------------------------------------

// Columns of a tile's row: 0 pair energy, 1-6 pair stress (full-list
// sums), 7 sum m vh^2, 8-13 kinetic stress, 14 sum m v^2, 15 max d^2.
template <bool MIXED, bool INTEGRATE>
__global__ __launch_bounds__(SEPB_BLOCK) void k_bstep(sepb_view b, sepb_lj c, const int4 *tiles,
						      int ntiles)
{
	for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
		const int4 td = tiles[tile]; // system, first atom, row
		sepb_sys *y = b.sys + td.x;
		const long off = y->off;
		const int n = y->n, npad = y->npad, sub = y->sub;
		// a system rebuilt this step (flag still set) is in the spare
		// copy, at position slot 0
		const bool rebuilt = y->neighb_flag != 0u;
		const sepb_state &S = b.st[rebuilt ? 1 - y->cur : y->cur];
		const int p = rebuilt ? 0 : y->p;
		const long slots = b.slots;
		const double L0 = y->L[0], L1 = y->L[1], L2 = y->L[2];
		const int t = td.y + (int)threadIdx.x / sub;
		const int l = (int)threadIdx.x % sub;
		const bool act = t < n;
		double fx = 0.0, fy = 0.0, fz = 0.0;
		// pair energy and the crossing part of the pair stress
		double pacc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
		sepk_pos pi = {0.0, 0.0, 0.0, 0.0};
		const sepk_pos *pos = S.pos[p] + off;
		if (act) {
			const double iL0 = y->iL[0], iL1 = y->iL[1], iL2 = y->iL[2];
			pi = pos[t];
			const int q = sub * SEPB_BATCH;
			const int nbat = (b.nbr_count[off + t] + q - 1) / q;
			const long stride = (long)npad * sub;
			const int *lp = b.nbr + y->list_off + (long)t * sub + l;
			for (int bb = 0; bb < nbat; bb++, lp += SEPB_BATCH * stride) {
				int j[SEPB_BATCH];
#pragma unroll
				for (int u = 0; u < SEPB_BATCH; u++)
					j[u] = lp[u * stride];
				double px[SEPB_BATCH], py[SEPB_BATCH], pz[SEPB_BATCH];
#pragma unroll
				for (int u = 0; u < SEPB_BATCH; u++) {
					const sepk_pos &r = pos[j[u]];
					px[u] = r.x;
					py[u] = r.y;
					pz[u] = r.z;
				}
#pragma unroll
				for (int u = 0; u < SEPB_BATCH; u++) {
					// minimum image d - L k, k = rint(d / L): the
					// single-image rule for every listed pair, whose
					// separation stays below 1.5 L
					const double kx = rint((pi.x - px[u]) * iL0);
					const double ky = rint((pi.y - py[u]) * iL1);
					const double kz = rint((pi.z - pz[u]) * iL2);
					const double wx = fma(-L0, kx, pi.x - px[u]);
					const double wy = fma(-L1, ky, pi.y - py[u]);
					const double wz = fma(-L2, kz, pi.z - pz[u]);
					const double r2 = wx * wx + wy * wy + wz * wz;
					if (!(r2 < c.cf2))
						continue;
					double ft, e;
					if (MIXED) {
						const float q1 = __builtin_amdgcn_rcpf((float)r2);
						const float q3 = q1 * q1 * q1;
						ft = (double)(fmaf(c.Af, q3, c.nBf) * q3 * q1);
						e = (double)fmaf(fmaf(c.Cf, q3, c.nDf), q3, c.ushiftf);
					} else {
						const double q1 = 1.0 / r2;
						const double q3 = q1 * q1 * q1;
						ft = fma(c.A, q3, c.nB) * q3 * q1;
						e = fma(fma(c.C, q3, c.nD), q3, c.ushift);
					}
					fx = fma(ft, wx, fx);
					fy = fma(ft, wy, fy);
					fz = fma(ft, wz, fz);
					pacc[0] += e;
					if (kx != 0.0 || ky != 0.0 || kz != 0.0) {
						// the crossing term of the stress: -f (x) L k
						const double sx = kx * L0, sy = ky * L1, sz = kz * L2;
						const double gx = -ft * wx, gy = -ft * wy, gz = -ft * wz;
						pacc[1] = fma(gx, sx, pacc[1]);
						pacc[2] = fma(gx, sy, pacc[2]);
						pacc[3] = fma(gx, sz, pacc[3]);
						pacc[4] = fma(gy, sy, pacc[4]);
						pacc[5] = fma(gy, sz, pacc[5]);
						pacc[6] = fma(gz, sz, pacc[6]);
					}
				}
			}
		}
		double acc[SEPB_NV];
#pragma unroll
		for (int k = 0; k < 7; k++)
			acc[k] = pacc[k];
#pragma unroll
		for (int k = 7; k < SEPB_NV; k++)
			acc[k] = 0.0;
		for (int o = 1; o < sub; o <<= 1) {
			fx += __shfl_xor(fx, o, SEPK_WAVE);
			fy += __shfl_xor(fy, o, SEPK_WAVE);
			fz += __shfl_xor(fz, o, SEPK_WAVE);
		}
		double al = 0.0;
		if (INTEGRATE)
			al = sepk_nh_alpha(y->alpha, y->sumv2, (double)n, y->Td, y->dt_tau2);
		if (act && l == 0) {
			// stress: 2 F_i (x) x_i per atom (the full-list identity)
			const double gx = 2.0 * fx, gy = 2.0 * fy, gz = 2.0 * fz;
			acc[1] = fma(gx, pi.x, acc[1]);
			acc[2] = fma(gx, pi.y, acc[2]);
			acc[3] = fma(gx, pi.z, acc[3]);
			acc[4] = fma(gy, pi.y, acc[4]);
			acc[5] = fma(gy, pi.z, acc[5]);
			acc[6] = fma(gz, pi.z, acc[6]);
			const long g = off + t;
			if (!INTEGRATE) {
				S.f[g] = fx;
				S.f[slots + g] = fy;
				S.f[2 * slots + g] = fz;
			} else {
				// sep_nosehoover then sep_leapfrog
				// (source/sepintgr.c:149-168, :42-88)
				const double m = S.m[g];
				const double Lk[3] = {L0, L1, L2};
				const double F[3] = {fx, fy, fz};
				double x[3] = {pi.x, pi.y, pi.z};
				double vh[3];
				double d2 = 0.0;
				for (int k = 0; k < 3; k++) {
					const long i = k * slots + g;
					double v = S.v[i];
					const double f = F[k] - al * m * v;
					const double a = f / m;
					v += a * y->dt;
					x[k] += v * y->dt;
					vh[k] = v - 0.5 * a * y->dt;
					acc[7] += vh[k] * vh[k] * m;
					acc[14] += v * v * m;
					int cn = S.cn[i], cr = S.cr[i];
					if (x[k] > Lk[k]) {
						x[k] -= Lk[k];
						cn++;
						cr++;
					} else if (x[k] < 0.0) {
						x[k] += Lk[k];
						cn--;
						cr--;
					}
					const double dr = (x[k] + cn * Lk[k]) - S.xn[i];
					d2 += dr * dr;
					S.f[i] = f;
					S.a[i] = a;
					S.v[i] = v;
					S.cn[i] = cn;
					S.cr[i] = cr;
				}
				sepk_pos np = {x[0], x[1], x[2], pi.q};
				S.pos[1 - p][g] = np;
				acc[8] = vh[0] * vh[0] * m;
				acc[9] = vh[0] * vh[1] * m;
				acc[10] = vh[0] * vh[2] * m;
				acc[11] = vh[1] * vh[1] * m;
				acc[12] = vh[1] * vh[2] * m;
				acc[13] = vh[2] * vh[2] * m;
				acc[15] = d2;
			}
		}
		sepk_block_partial<SEPB_NV>(acc, 1u << 15, b.partial, td.z, b.prow);
		__syncthreads(); // sepk_block_partial's LDS is reused by the next tile
	}
}

// One workgroup per system, after k_bstep: the system's rows reduced in
// tile order (16 threads per column over strided rows, then the 16 part
// sums in order: a fixed order, so deterministic) and the finishing step.
template <bool INTEGRATE>
__global__ __launch_bounds__(SEPB_BLOCK) void k_bfinish(sepb_view b)
{
	__shared__ double red[16][SEPB_NV];
	__shared__ double tot[SEPB_NV];
	sepb_sys *y = b.sys + blockIdx.x;
	{
		const int col = threadIdx.x & 15, part = threadIdx.x >> 4;
		const double *r = b.partial + (size_t)col * b.prow + y->row0;
		const int nt = y->ntile;
		double sacc = 0.0;
		for (int k = part; k < nt; k += 16)
			sacc = col == 15 ? fmax(sacc, r[k]) : sacc + r[k];
		red[part][col] = sacc;
	}
	__syncthreads();
	if (threadIdx.x < SEPB_NV) {
		const int col = threadIdx.x;
		double sacc = red[0][col];
		for (int k = 1; k < 16; k++)
			sacc = col == 15 ? fmax(sacc, red[k][col]) : sacc + red[k][col];
		tot[col] = sacc;
	}
	__syncthreads();
	if (threadIdx.x != 0)
		return;
	double *ret = y->ret;
	for (int k = 0; k < SEPDEV_RET_WORDS; k++)
		ret[k] = 0.0;
	// each pair was seen from both ends
	ret[SEPDEV_RET_EPOT] = 0.5 * tot[0];
	sepk_add_sym9(ret + SEPDEV_RET_POTP, tot + 1, 0.5);
	const bool rebuilt = y->neighb_flag != 0u;
	int p = y->p;
	if (rebuilt) {
		// k_bsort moved the state to the spare copy, slot 0
		y->cur = 1 - y->cur;
		p = 0;
		y->neighb_flag = 0u;
	}
	if (INTEGRATE) {
		ret[SEPDEV_RET_EKIN] = 0.5 * tot[7];
		sepk_add_sym9(ret + SEPDEV_RET_KINP, tot + 8, 1.0);
		y->alpha = sepk_nh_alpha(y->alpha, y->sumv2, (double)y->n, y->Td, y->dt_tau2);
		y->sumv2 = tot[14];
		y->max_dist2 = tot[15];
		if (sqrt(tot[15]) > 0.5 * y->skin) {
			y->neighb_flag = 1u;
			y->nupdate++;
		}
		y->tnow += y->dt;
		p = 1 - p;
	}
	y->p = p;
}

// ---- host side ---------------------------------------------------------------------

template <typename T>
static int balloc(T **p, size_t count)
{
	*p = nullptr;
	if (count == 0)
		count = 1;
	if (hipMalloc((void **)p, count * sizeof(T)) != hipSuccess)
		return sepk_fail(SEP_GPU_EDEVICE, "batch: hipMalloc of %zu bytes failed",
				 count * sizeof(T));
	return 0;
}

static void free_state(sepb_state &s)
{
	for (int k = 0; k < 2; k++)
		(void)hipFree(s.pos[k]);
	(void)hipFree(s.v);
	(void)hipFree(s.f);
	(void)hipFree(s.a);
	(void)hipFree(s.xn);
	(void)hipFree(s.cn);
	(void)hipFree(s.cr);
	(void)hipFree(s.m);
	(void)hipFree(s.orig);
}

static int alloc_state(sepb_state &s, long slots)
{
	memset(&s, 0, sizeof s);
	int rc = 0;
	rc |= balloc(&s.pos[0], slots);
	rc |= balloc(&s.pos[1], slots);
	rc |= balloc(&s.v, 3 * slots);
	rc |= balloc(&s.f, 3 * slots);
	rc |= balloc(&s.a, 3 * slots);
	rc |= balloc(&s.xn, 3 * slots);
	rc |= balloc(&s.cn, 3 * slots);
	rc |= balloc(&s.cr, 3 * slots);
	rc |= balloc(&s.m, slots);
	rc |= balloc(&s.orig, slots);
	return rc;
}

extern "C" void sepdev_batch_destroy(sepdev_batch *b)
{
	if (!b)
		return;
	if (b->st)
		(void)hipStreamSynchronize(b->st);
	if (b->exec)
		(void)hipGraphExecDestroy(b->exec);
	if (b->graph)
		(void)hipGraphDestroy(b->graph);
	free_state(b->S[0]);
	free_state(b->S[1]);
	(void)hipFree(b->dS);
	(void)hipFree(b->dsys);
	(void)hipFree(b->cellid);
	(void)hipFree(b->cslot);
	(void)hipFree(b->otmp);
	(void)hipFree(b->gcount);
	(void)hipFree(b->outside);
	(void)hipFree(b->cell_start);
	(void)hipFree(b->cell_sys);
	(void)hipFree(b->nbr);
	(void)hipFree(b->nbr_count);
	(void)hipFree(b->partial);
	(void)hipFree(b->tiles);
	(void)hipFree(b->sub_tiles);
	if (b->st)
		(void)hipStreamDestroy(b->st);
	free(b->hsys);
	free(b);
}

static sepb_view view_of(const sepdev_batch *b)
{
	sepb_view v;
	memset(&v, 0, sizeof v);
	v.sys = b->dsys;
	v.st = b->dS;
	v.slots = b->slots;
	v.nsys = b->nsys;
	v.cellid = b->cellid;
	v.cslot = b->cslot;
	v.otmp = b->otmp;
	v.gcount = b->gcount;
	v.outside = b->outside;
	v.cell_start = b->cell_start;
	v.cell_sys = b->cell_sys;
	v.ncell_tot = b->ncell_tot;
	v.nbr = b->nbr;
	v.nbr_count = b->nbr_count;
	v.cap = b->cap;
	v.partial = b->partial;
	v.prow = b->prow;
	return v;
}

// Tiles of every system at its lanes per atom; rows numbered in table order.
static int make_tiles(sepdev_batch *b)
{
	long nt = 0;
	for (int s = 0; s < b->nsys; s++) {
		const int apt = SEPB_BLOCK / b->hsys[s].sub;
		nt += (b->hsys[s].n + apt - 1) / apt;
	}
	if (nt > b->prow)
		return sepk_fail(SEP_GPU_EARG, "batch: %ld tiles exceed %ld rows", nt, b->prow);
	int4 *h = (int4 *)malloc(nt * sizeof(int4));
	long k = 0;
	for (int s = 0; s < b->nsys; s++) {
		sepb_sys &y = b->hsys[s];
		const int apt = SEPB_BLOCK / y.sub;
		y.row0 = (int)k;
		y.ntile = (y.n + apt - 1) / apt;
		for (int i = 0; i < y.ntile; i++, k++)
			h[k] = make_int4(s, i * apt, (int)k, 0);
	}
	b->ntiles = (int)nt;
	hipError_t e = hipMemcpy(b->tiles, h, nt * sizeof(int4), hipMemcpyHostToDevice);
	free(h);
	if (e != hipSuccess)
		return sepk_fail(SEP_GPU_EDEVICE, "batch: tile upload: %s", hipGetErrorString(e));
	return 0;
}

static int get_sys(sepdev_batch *b)
{
	SEPK_CHECK(hipStreamSynchronize(b->st));
	SEPK_CHECK(hipMemcpy(b->hsys, b->dsys, b->nsys * sizeof(sepb_sys), hipMemcpyDeviceToHost));
	return 0;
}

static int put_sys(sepdev_batch *b)
{
	SEPK_CHECK(hipStreamSynchronize(b->st));
	SEPK_CHECK(hipMemcpy(b->dsys, b->hsys, b->nsys * sizeof(sepb_sys), hipMemcpyHostToDevice));
	return 0;
}

static void drop_graph(sepdev_batch *b)
{
	// replays of the old graph may still be queued on the stream
	if (b->exec) {
		(void)hipStreamSynchronize(b->st);
		(void)hipGraphExecDestroy(b->exec);
	}
	if (b->graph)
		(void)hipGraphDestroy(b->graph);
	b->exec = nullptr;
	b->graph = nullptr;
}

extern "C" sepdev_batch *sepdev_batch_create(int nsys, const sepdev_batch_sys *in, double cf,
					     int cap, int *err)
{
	*err = 0;
	if (nsys <= 0 || !in) {
		*err = sepk_fail(SEP_GPU_EARG, "batch: %d systems", nsys);
		return nullptr;
	}
	if (cap <= 0)
		cap = 256;
	cap = (cap + 63) / 64 * 64;
	sepdev_batch *b = (sepdev_batch *)calloc(1, sizeof(sepdev_batch));
	b->nsys = nsys;
	b->cap = cap;
	b->hsys = (sepb_sys *)calloc(nsys, sizeof(sepb_sys));
	long slots = 0, cells = 0;
	for (int s = 0; s < nsys; s++) {
		const sepdev_batch_sys &a = in[s];
		sepb_sys &y = b->hsys[s];
		if (a.n <= 0 || a.n > (1L << 26)) {
			*err = sepk_fail(SEP_GPU_EARG, "batch: system %d has %ld atoms", s, a.n);
			sepdev_batch_destroy(b);
			return nullptr;
		}
		const double rl = sqrt(a.rlist2);
		y.off = slots;
		y.n = (int)a.n;
		y.npad = (int)((a.n + 1 + 63) / 64 * 64);
		y.sub = SEPDEV_LANES;
		for (int k = 0; k < 3; k++) {
			// single image: the list cutoff stays below half the box
			if (!(0.5 * a.L[k] >= rl) || !(cf * cf <= a.rlist2)) {
				*err = sepk_fail(SEP_GPU_EARG,
						 "batch: system %d: box %g too small for list cutoff %g",
						 s, a.L[k], rl);
				sepdev_batch_destroy(b);
				return nullptr;
			}
			int nc = (int)(a.L[k] / rl);
			if (nc < 1)
				nc = 1;
			y.ncell[k] = nc;
			y.lcell[k] = a.L[k] / nc;
			y.ilcell[k] = 1.0 / y.lcell[k];
			y.L[k] = a.L[k];
			y.iL[k] = 1.0 / a.L[k];
		}
		if ((long)y.ncell[0] * y.ncell[1] * y.ncell[2] > SEPB_MAX_CELLS) {
			*err = sepk_fail(SEP_GPU_EARG, "batch: system %d needs more than %d cells", s,
					 SEPB_MAX_CELLS);
			sepdev_batch_destroy(b);
			return nullptr;
		}
		y.cell_off = (int)cells;
		cells += (long)y.ncell[0] * y.ncell[1] * y.ncell[2] + 1;
		y.rlist2 = a.rlist2;
		y.dt = a.dt;
		y.Td = a.Td;
		y.dt_tau2 = a.dt / (a.tau * a.tau);
		y.skin = a.skin;
		y.alpha = a.alpha;
		y.tnow = a.tnow;
		y.neighb_flag = 1u; // the first step builds
		slots += y.npad;
	}
	b->slots = slots;
	b->ncell_tot = cells;
	b->prow = slots / 16 + nsys + 64; // tiles hold at least 16 atoms
	int rc = 0;
	if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess)
		rc = sepk_fail(SEP_GPU_EDEVICE, "batch: hipStreamCreate failed");
	rc |= alloc_state(b->S[0], slots);
	rc |= alloc_state(b->S[1], slots);
	rc |= balloc(&b->dS, 2);
	rc |= balloc(&b->dsys, nsys);
	rc |= balloc(&b->cellid, slots);
	rc |= balloc(&b->cslot, slots);
	rc |= balloc(&b->otmp, slots);
	rc |= balloc(&b->gcount, cells);
	rc |= balloc(&b->outside, nsys);
	rc |= balloc(&b->cell_start, cells);
	rc |= balloc(&b->cell_sys, cells);
	rc |= balloc(&b->nbr, (size_t)slots * cap);
	rc |= balloc(&b->nbr_count, slots);
	rc |= balloc(&b->partial, (size_t)b->prow * SEPB_NV);
	rc |= balloc(&b->tiles, b->prow);
	rc |= balloc(&b->sub_tiles, b->prow);
	if (rc) {
		*err = rc;
		sepdev_batch_destroy(b);
		return nullptr;
	}
	for (int s = 0; s < nsys; s++)
		b->hsys[s].list_off = b->hsys[s].off * (long)cap;
	for (int s = 0; s < nsys; s++)
		if ((long)b->hsys[s].npad * cap >= (1L << 31)) {
			*err = sepk_fail(SEP_GPU_EARG, "batch: system %d: list of %d x %d entries", s,
					 b->hsys[s].npad, cap);
			sepdev_batch_destroy(b);
			return nullptr;
		}
	// LJ constants of sep_lj_shift (source/sepmisc.c:132-147)
	b->lj.cf2 = cf * cf;
	b->lj.A = 48.0;
	b->lj.nB = -24.0;
	b->lj.C = 4.0;
	b->lj.nD = -4.0;
	b->lj.ushift = 0.016316891136;
	b->lj.Af = 48.0f;
	b->lj.nBf = -24.0f;
	b->lj.Cf = 4.0f;
	b->lj.nDf = -4.0f;
	b->lj.ushiftf = (float)0.016316891136;
	b->grid = 1024;

	// host images of the per-slot and per-cell tables, then the state
	int *csys = (int *)malloc(cells * sizeof(int));
	for (int s = 0; s < nsys; s++) {
		const sepb_sys &y = b->hsys[s];
		const int nc = y.ncell[0] * y.ncell[1] * y.ncell[2];
		for (int k = 0; k < nc; k++)
			csys[y.cell_off + k] = s;
		csys[y.cell_off + nc] = -1;
	}
	const size_t dbytes = (size_t)slots * sizeof(double);
	sepk_pos *hp = (sepk_pos *)malloc(slots * sizeof(sepk_pos));
	double *hv = (double *)calloc(3 * slots, sizeof(double));
	double *hm = (double *)calloc(slots, sizeof(double));
	int *ho = (int *)calloc(slots, sizeof(int));
	int *hz = (int *)calloc(3 * slots, sizeof(int));
	const double nan = __builtin_nan("");
	for (long k = 0; k < slots; k++)
		hp[k] = sepk_pos{nan, nan, nan, 0.0};
	for (int s = 0; s < nsys; s++) {
		const sepdev_batch_sys &a = in[s];
		sepb_sys &y = b->hsys[s];
		double sv2 = 0.0;
		for (long t = 0; t < a.n; t++) {
			const long g = y.off + t;
			hp[g] = sepk_pos{a.x[3 * t], a.x[3 * t + 1], a.x[3 * t + 2], 0.0};
			for (int k = 0; k < 3; k++) {
				hv[k * slots + g] = a.v[3 * t + k];
				sv2 += a.v[3 * t + k] * a.v[3 * t + k] * a.m[t];
			}
			hm[g] = a.m[t];
			ho[g] = (int)t;
			if (a.cr)
				for (int k = 0; k < 3; k++)
					hz[k * slots + g] = a.cr[3 * t + k];
		}
		// sum m v^2 of the starting velocities, for the first thermostat
		// update
		y.sumv2 = sv2;
	}
	hipError_t e = hipSuccess;
	for (int c = 0; c < 2 && e == hipSuccess; c++) {
		for (int k = 0; k < 2 && e == hipSuccess; k++)
			e = hipMemcpy(b->S[c].pos[k], hp, slots * sizeof(sepk_pos), hipMemcpyHostToDevice);
		if (e == hipSuccess)
			e = hipMemcpy(b->S[c].v, hv, 3 * dbytes, hipMemcpyHostToDevice);
		if (e == hipSuccess)
			e = hipMemcpy(b->S[c].m, hm, dbytes, hipMemcpyHostToDevice);
		if (e == hipSuccess)
			e = hipMemcpy(b->S[c].orig, ho, slots * sizeof(int), hipMemcpyHostToDevice);
		if (e == hipSuccess)
			e = hipMemcpy(b->S[c].cr, hz, 3 * slots * sizeof(int), hipMemcpyHostToDevice);
		if (e == hipSuccess)
			e = hipMemset(b->S[c].cn, 0, 3 * slots * sizeof(int));
		if (e == hipSuccess)
			e = hipMemset(b->S[c].f, 0, 3 * dbytes);
		if (e == hipSuccess)
			e = hipMemset(b->S[c].a, 0, 3 * dbytes);
		if (e == hipSuccess)
			e = hipMemset(b->S[c].xn, 0, 3 * dbytes);
	}
	if (e == hipSuccess)
		e = hipMemcpy(b->dS, b->S, 2 * sizeof(sepb_state), hipMemcpyHostToDevice);
	if (e == hipSuccess)
		e = hipMemcpy(b->cell_sys, csys, cells * sizeof(int), hipMemcpyHostToDevice);
	if (e == hipSuccess)
		e = hipMemset(b->nbr_count, 0, slots * sizeof(int));
	// the binning's counters start (and stay, between rebuilds) at zero
	if (e == hipSuccess)
		e = hipMemset(b->gcount, 0, cells * sizeof(int));
	if (e == hipSuccess)
		e = hipMemset(b->outside, 0, nsys * sizeof(int));
	free(csys);
	free(hp);
	free(hv);
	free(hm);
	free(ho);
	free(hz);
	if (e != hipSuccess) {
		*err = sepk_fail(SEP_GPU_EDEVICE, "batch: upload: %s", hipGetErrorString(e));
		sepdev_batch_destroy(b);
		return nullptr;
	}
	if ((rc = make_tiles(b)) || (rc = put_sys(b))) {
		*err = rc;
		sepdev_batch_destroy(b);
		return nullptr;
	}
	return b;
}

static int launch_rebuild(sepdev_batch *b)
{
	const sepb_view v = view_of(b);
	if (b->nsys <= SEPB_FMAP) {
		k_bcount<<<SEPB_BGRID, SEPB_BLOCK, 0, b->st>>>(v);
		k_bscan<<<b->nsys, SEPB_BIN_THREADS, 0, b->st>>>(v);
		k_bscatter<<<SEPB_BGRID, SEPB_BLOCK, 0, b->st>>>(v);
	} else {
		k_bbin<<<b->nsys, SEPB_BIN_THREADS, 0, b->st>>>(v);
	}
	SEPK_LAUNCH_CHECK();
	k_bsort<<<512, SEPB_BLOCK, 0, b->st>>>(v);
	SEPK_LAUNCH_CHECK();
	k_bbuild<<<SEPB_BGRID, SEPB_BLOCK, 0, b->st>>>(v);
	SEPK_LAUNCH_CHECK();
	return 0;
}

// The step (or a force evaluation) over `tiles`; `finish` adds the
// per-system finishing launch, which needs every tile of the batch.
static int launch_step(sepdev_batch *b, int mixed, int integrate, const int4 *tiles, int ntiles,
		       int finish)
{
	const sepb_view v = view_of(b);
	const int grid = ntiles < b->grid ? ntiles : b->grid;
	if (grid > 0) {
		if (mixed) {
			if (integrate)
				k_bstep<true, true><<<grid, SEPB_BLOCK, 0, b->st>>>(v, b->lj, tiles, ntiles);
			else
				k_bstep<true, false><<<grid, SEPB_BLOCK, 0, b->st>>>(v, b->lj, tiles, ntiles);
		} else {
			if (integrate)
				k_bstep<false, true><<<grid, SEPB_BLOCK, 0, b->st>>>(v, b->lj, tiles, ntiles);
			else
				k_bstep<false, false><<<grid, SEPB_BLOCK, 0, b->st>>>(v, b->lj, tiles, ntiles);
		}
		SEPK_LAUNCH_CHECK();
	}
	if (finish) {
		if (integrate)
			k_bfinish<true><<<b->nsys, SEPB_BLOCK, 0, b->st>>>(v);
		else
			k_bfinish<false><<<b->nsys, SEPB_BLOCK, 0, b->st>>>(v);
		SEPK_LAUNCH_CHECK();
	}
	return 0;
}

#define SEPB_GRAPH_STEPS 8

extern "C" int sepdev_batch_run(sepdev_batch *b, long nsteps, int mixed, int graph)
{
	if (nsteps <= 0)
		return 0;
	int rc;
	if (!graph) {
		for (long i = 0; i < nsteps; i++)
			if ((rc = launch_rebuild(b)) || (rc = launch_step(b, mixed, 1, b->tiles, b->ntiles, 1)))
				return rc;
		return 0;
	}
	// SEPB_GRAPH_STEPS steps a replayed graph in long runs (each replay
	// costs a few microseconds on the device before its first kernel), the
	// few left over eagerly; one a replay in short runs
	const int per = nsteps >= SEPB_GRAPH_STEPS ? SEPB_GRAPH_STEPS : 1;
	if (!b->exec || b->graph_mixed != mixed || b->graph_steps != per) {
		drop_graph(b);
		SEPK_CHECK(hipStreamBeginCapture(b->st, hipStreamCaptureModeRelaxed));
		rc = 0;
		for (int k = 0; !rc && k < per; k++)
			if (!(rc = launch_rebuild(b)))
				rc = launch_step(b, mixed, 1, b->tiles, b->ntiles, 1);
		hipGraph_t g = nullptr;
		const hipError_t e = hipStreamEndCapture(b->st, &g);
		if (rc) {
			if (g)
				(void)hipGraphDestroy(g);
			return rc;
		}
		if (e != hipSuccess)
			return sepk_fail(SEP_GPU_EDEVICE, "batch: capture: %s", hipGetErrorString(e));
		b->graph = g;
		SEPK_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
		b->graph_mixed = mixed;
		b->graph_steps = per;
	}
	for (long i = 0; i < nsteps / per; i++)
		SEPK_CHECK(hipGraphLaunch(b->exec, b->st));
	for (long i = nsteps % per; i > 0; i--)
		if ((rc = launch_rebuild(b)) || (rc = launch_step(b, mixed, 1, b->tiles, b->ntiles, 1)))
			return rc;
	return 0;
}

// Forces, energy and stress at the current positions (rebuilding flagged
// lists first); nothing moves.
extern "C" int sepdev_batch_eval(sepdev_batch *b, int mixed)
{
	int rc = launch_rebuild(b);
	if (!rc)
		rc = launch_step(b, mixed, 0, b->tiles, b->ntiles, 1);
	return rc;
}

extern "C" int sepdev_batch_synchronize(sepdev_batch *b)
{
	SEPK_CHECK(hipStreamSynchronize(b->st));
	return 0;
}

// State of system s in its original atom order (3n interleaved vectors;
// any pointer may be null), its return slot (SEPDEV_RET_WORDS) and scalars
// {alpha, max_dist2, tnow, nupdate, neighb_flag, max_full, error, sumv2}.
extern "C" int sepdev_batch_get(sepdev_batch *b, int s, double *x, double *v, double *f,
				double *a, int *cr, double *ret, double *scal)
{
	if (s < 0 || s >= b->nsys)
		return sepk_fail(SEP_GPU_EARG, "batch: system %d", s);
	int rc = get_sys(b);
	if (rc)
		return rc;
	const sepb_sys &y = b->hsys[s];
	const sepb_state &S = b->S[y.cur];
	const long n = y.n, slots = b->slots;
	sepk_pos *hp = (sepk_pos *)malloc(n * sizeof(sepk_pos));
	double *hv = (double *)malloc(3 * n * sizeof(double));
	int *ho = (int *)malloc(n * sizeof(int));
	int *hi = (int *)malloc(3 * n * sizeof(int));
	hipError_t e = hipMemcpy(ho, S.orig + y.off, n * sizeof(int), hipMemcpyDeviceToHost);
	if (e == hipSuccess && x) {
		e = hipMemcpy(hp, S.pos[y.p] + y.off, n * sizeof(sepk_pos), hipMemcpyDeviceToHost);
		for (long t = 0; t < n && e == hipSuccess; t++) {
			x[3 * ho[t]] = hp[t].x;
			x[3 * ho[t] + 1] = hp[t].y;
			x[3 * ho[t] + 2] = hp[t].z;
		}
	}
	double *fields[3] = {v, f, a};
	double *src[3] = {S.v, S.f, S.a};
	for (int q = 0; q < 3 && e == hipSuccess; q++) {
		if (!fields[q])
			continue;
		for (int k = 0; k < 3 && e == hipSuccess; k++)
			e = hipMemcpy(hv + k * n, src[q] + k * slots + y.off, n * sizeof(double),
				      hipMemcpyDeviceToHost);
		for (long t = 0; t < n && e == hipSuccess; t++)
			for (int k = 0; k < 3; k++)
				fields[q][3 * ho[t] + k] = hv[k * n + t];
	}
	if (cr && e == hipSuccess) {
		for (int k = 0; k < 3 && e == hipSuccess; k++)
			e = hipMemcpy(hi + k * n, S.cr + k * slots + y.off, n * sizeof(int),
				      hipMemcpyDeviceToHost);
		for (long t = 0; t < n && e == hipSuccess; t++)
			for (int k = 0; k < 3; k++)
				cr[3 * ho[t] + k] = hi[k * n + t];
	}
	free(hp);
	free(hv);
	free(ho);
	free(hi);
	if (e != hipSuccess)
		return sepk_fail(SEP_GPU_EDEVICE, "batch: download: %s", hipGetErrorString(e));
	if (ret)
		memcpy(ret, y.ret, SEPDEV_RET_WORDS * sizeof(double));
	if (scal) {
		scal[0] = y.alpha;
		scal[1] = y.max_dist2;
		scal[2] = y.tnow;
		scal[3] = y.nupdate;
		scal[4] = y.neighb_flag;
		scal[5] = y.max_full;
		scal[6] = y.error;
		scal[7] = y.sumv2;
	}
	if (y.error)
		return sepk_fail(y.error, "batch: system %d: device error %d%s", s, y.error,
				 y.error == SEP_GPU_ENEIGHB ? " (neighbour list overflow)" : "");
	return 0;
}

// Lanes per atom of each system (1, 2, 4, 8 or 16) and the persistent
// grid; systems whose lanes change are rebuilt before the next step.
extern "C" int sepdev_batch_set_shape(sepdev_batch *b, const int *sub, int grid)
{
	int rc = get_sys(b);
	if (rc)
		return rc;
	if (sub)
		for (int s = 0; s < b->nsys; s++) {
			const int u = sub[s];
			if (u != 1 && u != 2 && u != 4 && u != 8 && u != 16)
				return sepk_fail(SEP_GPU_EARG, "batch: %d lanes per atom", u);
			if (u * SEPB_BATCH > b->cap)
				return sepk_fail(SEP_GPU_EARG, "batch: %d lanes exceed the list capacity", u);
			if (u != b->hsys[s].sub) {
				b->hsys[s].sub = u;
				b->hsys[s].neighb_flag = 1u;
			}
		}
	if (grid > 0)
		b->grid = grid;
	if ((rc = make_tiles(b)) || (rc = put_sys(b)))
		return rc;
	drop_graph(b);
	return 0;
}

extern "C" int sepdev_batch_get_shape(sepdev_batch *b, int *sub, int *grid)
{
	for (int s = 0; s < b->nsys; s++)
		sub[s] = b->hsys[s].sub;
	*grid = b->grid;
	return 0;
}

// Mean duration of one force evaluation (no integration) over the tiles of
// the systems with mask[s] != 0 (null: all), reps launches back to back.
extern "C" int sepdev_batch_time_eval(sepdev_batch *b, const int *mask, int mixed, int reps,
				      double *ms)
{
	// a complete evaluation first (rebuilds pending lists and finishes
	// them), then the tiles of the chosen systems alone
	int rc = sepdev_batch_eval(b, mixed);
	if (rc || (rc = get_sys(b)))
		return rc;
	int4 *h = (int4 *)malloc(b->prow * sizeof(int4));
	int nt = 0;
	for (int s = 0; s < b->nsys; s++) {
		if (mask && !mask[s])
			continue;
		const sepb_sys &y = b->hsys[s];
		const int apt = SEPB_BLOCK / y.sub;
		for (int i = 0; i < y.ntile; i++)
			h[nt++] = make_int4(s, i * apt, y.row0 + i, 0);
	}
	hipError_t e = hipMemcpy(b->sub_tiles, h, nt * sizeof(int4), hipMemcpyHostToDevice);
	free(h);
	if (e != hipSuccess)
		return sepk_fail(SEP_GPU_EDEVICE, "batch: tiles: %s", hipGetErrorString(e));
	hipEvent_t e0, e1;
	SEPK_CHECK(hipEventCreate(&e0));
	SEPK_CHECK(hipEventCreate(&e1));
	rc = launch_step(b, mixed, 0, b->sub_tiles, nt, 0);
	if (!rc)
		(void)hipEventRecord(e0, b->st);
	for (int r = 0; r < reps && !rc; r++)
		rc = launch_step(b, mixed, 0, b->sub_tiles, nt, 0);
	if (!rc) {
		(void)hipEventRecord(e1, b->st);
		e = hipEventSynchronize(e1);
		float f = 0.0f;
		if (e == hipSuccess)
			e = hipEventElapsedTime(&f, e0, e1);
		if (e != hipSuccess)
			rc = sepk_fail(SEP_GPU_EDEVICE, "batch: timing: %s", hipGetErrorString(e));
		*ms = reps > 0 ? f / reps : 0.0;
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	return rc;
}
