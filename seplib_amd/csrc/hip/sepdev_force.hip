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

__global__ void k_zero_forces(double *f, long count, const int *err)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < count && !*err)
		f[t] = 0.0;
}

int sepk_materialize_forces(sepdev *d)
{
	int rc = sepk_flush_pair(d);
	if (rc || (rc = sepk_flush_gathers(d)))
		return rc;
	if (!d->f_zero_pending)
		return 0;
	const long cnt = 3 * d->npad;
	k_zero_forces<<<sepk_blocks(cnt), SEPK_BLOCK, 0, d->st>>>(d->s[0].f, cnt, d->s[0].err);
	SEPK_LAUNCH_CHECK();
	d->f_zero_pending = 0;
	return 0;
}

extern "C" int sepdev_zero_forces(sepdev *d)
{
	// The next force kernel overwrites instead of accumulating; anything
	// else that reads f first materialises the zeros.  Owed gathers would
	// be zeroed with the rest: they are dropped.
	d->f_zero_pending = 1;
	d->ngpend = 0;
	d->gpend_over = 0;
	return 0;
}

// ---- LJ family ---------------------------------------------------------------

// ft = eps48 q^3 (q^3 - aw/2) q and u = eps4 q^3 (q^3 - aw) + ushift with
// q = sig2 / r^2, evaluated as ft = (A q^3 + nB) q^3 q, u = (C q^3 + nD) q^3 + ushift
// (nB = -eps48 aw/2, nD = -eps4 aw)

// Shared epilogue of the list kernels: combine the SEPK_SUB lanes of each
// atom (fixed xor order), write or add its force from lane 0, and leave the
// block's energy + stress row in the pair region.  acc = {e, wxx, wxy, wxz,
// wyy, wyz, wzz}, each summed over the lane's share of the full list.
// The stress of a pair term, sum over listed pairs of f (x) r, is not
// accumulated pair by pair.  With r_ij = x_i - x_j - s_ij (s_ij = L k_ij,
// the periodic shift taken off) and f_ji = -f_ij, the full-list sum is
//     sum_{i,j} f_ij (x) r_ij = 2 sum_i F_i (x) x_i - sum_{i,j} f_ij (x) s_ij,
// so a pair adds to the stress only when it crosses the box (k != 0), and
// each atom adds 2 F_i (x) x_i once, here, from its summed force.  The six
// components are {xx, xy, xz, yy, yz, zz} of f (x) r; the finisher halves
// the full-list sums.  frozen: the device error word is set, and f is left
// as it is.
template <bool OVERWRITE>
__device__ __forceinline__ void sepk_pair_epilogue(
	long t, int sub, long n, long npad, bool act, double fx, double fy, double fz,
	double px, double py, double pz, double (&acc)[7], double *f, double *partial,
	long row, long rows, bool frozen)
{
#pragma unroll
	for (int o = 1; o < SEPK_SUB; o <<= 1) {
		fx += __shfl_xor(fx, o, SEPK_WAVE);
		fy += __shfl_xor(fy, o, SEPK_WAVE);
		fz += __shfl_xor(fz, o, SEPK_WAVE);
	}
	if (t < n && sub == 0 && !frozen) {
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
	if (act && sub == 0) {
		const double gx = 2.0 * fx, gy = 2.0 * fy, gz = 2.0 * fz;
		acc[1] = fma(gx, px, acc[1]);
		acc[2] = fma(gx, py, acc[2]);
		acc[3] = fma(gx, pz, acc[3]);
		acc[4] = fma(gy, py, acc[4]);
		acc[5] = fma(gy, pz, acc[5]);
		acc[6] = fma(gz, pz, acc[6]);
	}
	sepk_block_partial<7>(acc, 0u, partial, row, rows);
}

// Minimum image of d, returning d - L k with the image count k in
// {-1, 0, 1} (the reference's single-image rule, or its FAST form).
template <bool FAST>
__device__ __forceinline__ double sepk_image(double d, double L, double iL, double &k)
{
	if (FAST) {
		k = rint(d * iL);
		return fma(-L, k, d);
	}
	k = 0.0;
	if (d > 0.5 * L) {
		k = 1.0;
		d -= L;
	} else if (d < -0.5 * L) {
		k = -1.0;
		d += L;
	}
	return d;
}

// Adds one pair's force ft r to the atom and its energy; the stress gets
// only the crossing term -f (x) s (see sepk_pair_epilogue).
__device__ __forceinline__ void sepk_pair_cross(
	double ft, double dx, double dy, double dz, double kx, double ky, double kz, double L0,
	double L1, double L2, double (&acc)[7])
{
	const double sx = kx * L0, sy = ky * L1, sz = kz * L2;
	const double gx = -ft * dx, gy = -ft * dy, gz = -ft * dz;
	acc[1] = fma(gx, sx, acc[1]);
	acc[2] = fma(gx, sy, acc[2]);
	acc[3] = fma(gx, sz, acc[3]);
	acc[4] = fma(gy, sy, acc[4]);
	acc[5] = fma(gy, sz, acc[5]);
	acc[6] = fma(gz, sz, acc[6]);
}

__device__ __forceinline__ void sepk_pair_accumulate(
	double ft, double e, double dx, double dy, double dz, double kx, double ky, double kz,
	double L0, double L1, double L2, double &fx, double &fy, double &fz, double (&acc)[7])
{
	fx = fma(ft, dx, fx);
	fy = fma(ft, dy, fy);
	fz = fma(ft, dz, fz);
	acc[0] += e;
	if (kx != 0.0 || ky != 0.0 || kz != 0.0)
		sepk_pair_cross(ft, dx, dy, dz, kx, ky, kz, L0, L1, L2, acc);
}

// Entries of one lane are taken SEPK_BATCH at a time: every index and
// every neighbour position of the batch is loaded before any is used, so a
// wave keeps SEPK_BATCH gathers in flight instead of one.  Lists are padded
// to whole batches with the phantom slot (NaN position, type 0, charge 0),
// which fails every test, so the loads need neither clamps nor branches; the
// accumulation order is the list order.
//
// Addressing: lane `sub` of atom t reads the SEPK_BATCH indices of its
// batch b as one int4 (sepk_entry, sepk_lane_list); neighbour records are
// fetched with 32-bit byte offsets from the uniform base.
__device__ __forceinline__ const sepk_pos &sepk_rec(const char *base, int j)
{
	return *(const sepk_pos *)(base + (unsigned)j * (unsigned)sizeof(sepk_pos));
}

// 1 / r2 in FP64: hardware reciprocal and two Newton steps (within an ulp of
// the division; the error bar of the FP64 path is 1e-12)
__device__ __forceinline__ double sepk_recip(double r2)
{
	double y = __builtin_amdgcn_rcp(r2);
	double e = fma(-r2, y, 1.0);
	y = fma(y, e, y);
	e = fma(-r2, y, 1.0);
	return fma(y, e, y);
}

// The cutoff test of the mixed LJ loops.  Each quantized coordinate is
// within half a quantization step of the FP64 one, so the quantized r^2 is
// within c.band of the r^2 the plain kernel forms; a pair that close to the
// cutoff is decided on the FP64 records, as the plain kernel decides it.
// (The force of sep_lj and sep_lj_shift jumps by F(rc) = -0.039 at rc = 2.5:
// a pair taken by one form and dropped by the other would put that whole
// jump into the force error of both atoms.)  About one pair in 10^8 lies in
// the band.
template <bool FAST>
__device__ __forceinline__ bool sepk_inside_mixed(const sepk_state &s, long t, unsigned j,
						  double r2, const sepk_lj &c)
{
	if (!(fabs(r2 - c.cf2) < c.band))
		return r2 < c.cf2;
	const sepk_pos pi = s.pos[t], pj = s.pos[j];
	double kx, ky, kz;
	const double dx = sepk_image<FAST>(pi.x - pj.x, c.L0, c.iL0, kx);
	const double dy = sepk_image<FAST>(pi.y - pj.y, c.L1, c.iL1, ky);
	const double dz = sepk_image<FAST>(pi.z - pj.z, c.L2, c.iL2, kz);
	return dx * dx + dy * dy + dz * dz < c.cf2;
}

// One pair of the mixed LJ loop: the FP32 force and energy terms from the
// quantized squared distance r2 (dx, dy, dz in FP64), summed in FP64, and
// the crossing stress when the integer difference wrapped.
__device__ __forceinline__ void sepk_lj_mixed_add(
	const sepk_lj &c, unsigned ux, unsigned vx, unsigned uy, unsigned vy, unsigned uz,
	unsigned vz, int ix, int iy, int iz, double dx, double dy, double dz, double r2,
	double &fx, double &fy, double &fz, double (&acc)[7])
{
	const float q = c.sig2f * __builtin_amdgcn_rcpf((float)r2);
	const float q3 = q * q * q;
	const double ft = (double)(fmaf(c.Af, q3, c.nBf) * q3 * q);
	const double e = (double)fmaf(fmaf(c.Cf, q3, c.nDf), q3, c.ushiftf);
	fx = fma(ft, dx, fx);
	fy = fma(ft, dy, fy);
	fz = fma(ft, dz, fz);
	acc[0] += e;
	const bool bx = (ux < vx) != (ix < 0), by = (uy < vy) != (iy < 0),
		   bz = (uz < vz) != (iz < 0);
	if (bx | by | bz) {
		// k = +1 when the unwrapped difference is >= 2^31 steps (wrapped
		// negative), -1 when it is below -2^31 (wrapped non-negative)
		const double kx = bx ? (ix < 0 ? 1.0 : -1.0) : 0.0;
		const double ky = by ? (iy < 0 ? 1.0 : -1.0) : 0.0;
		const double kz = bz ? (iz < 0 ? 1.0 : -1.0) : 0.0;
		sepk_pair_cross(ft, dx, dy, dz, kx, ky, kz, c.L0, c.L1, c.L2, acc);
	}
}

// One entry of a mixed LJ walk (untyped): the pair of records qi, qj,
// added when inside c.cf2lo; one nearer the cutoff than that sets band and
// is left to sepk_lj_mixed_band.
__device__ __forceinline__ void sepk_lj_mixed_entry(const sepk_lj &c, const int4 &qi,
						    const int4 &qj, double &fx, double &fy,
						    double &fz, double (&acc)[7], bool &band)
{
	if (qj.w == SEPK_QPHANTOM)
		return;
	const unsigned ux = (unsigned)qi.x, vx = (unsigned)qj.x;
	const unsigned uy = (unsigned)qi.y, vy = (unsigned)qj.y;
	const unsigned uz = (unsigned)qi.z, vz = (unsigned)qj.z;
	const int ix = (int)(ux - vx), iy = (int)(uy - vy), iz = (int)(uz - vz);
	const double dx = (double)ix * c.qd0;
	const double dy = (double)iy * c.qd1;
	const double dz = (double)iz * c.qd2;
	const double r2 = dx * dx + dy * dy + dz * dz;
	if (!(r2 < c.cf2lo)) {
		// near the cutoff: decided after the walk
		band |= r2 < c.cf2hi;
		return;
	}
	sepk_lj_mixed_add(c, ux, vx, uy, vy, uz, vz, ix, iy, iz, dx, dy, dz, r2, fx, fy, fz, acc);
}

// The pairs of the mixed loop that lay within c.band of the cutoff (the
// loop left them out and flagged the lane): decided on the FP64 records as
// the plain kernel decides them (sepk_inside_mixed), and added when inside.
// A lane gets here for about one pair in 10^8.
template <bool TYPED, bool FAST>
__device__ __forceinline__ void sepk_lj_mixed_band(
	const sepk_state &s, const int *__restrict__ nbr, const int *__restrict__ cnt, long npad,
	const sepk_lj &c, long t, int sub, signed char ti, int4 qi, double &fx, double &fy,
	double &fz, double (&acc)[7])
{
	const int n = cnt[t];
	for (int e = sub; e < n; e += SEPK_SUB) {
		const unsigned j = (unsigned)nbr[sepk_entry(e, t, npad)];
		const int4 qj = sepk_at(s.qpos, j);
		if (TYPED) {
			const signed char tj = s.type[j];
			if (!((ti == c.t0 && tj == c.t1) || (ti == c.t1 && tj == c.t0)))
				continue;
		}
		if (qj.w == SEPK_QPHANTOM)
			continue;
		const unsigned ux = (unsigned)qi.x, vx = (unsigned)qj.x;
		const unsigned uy = (unsigned)qi.y, vy = (unsigned)qj.y;
		const unsigned uz = (unsigned)qi.z, vz = (unsigned)qj.z;
		const int ix = (int)(ux - vx), iy = (int)(uy - vy), iz = (int)(uz - vz);
		const double dx = (double)ix * c.qd0;
		const double dy = (double)iy * c.qd1;
		const double dz = (double)iz * c.qd2;
		const double r2 = dx * dx + dy * dy + dz * dz;
		if (r2 < c.cf2lo || !(r2 < c.cf2hi) || !sepk_inside_mixed<FAST>(s, t, j, r2, c))
			continue;
		sepk_lj_mixed_add(c, ux, vx, uy, vy, uz, vz, ix, iy, iz, dx, dy, dz, r2, fx, fy, fz,
				  acc);
	}
}

// The pair loop of atom t, lane `sub`: the lane's share of the atom's force
// and of the block's energy and crossing stress, and the position the
// stress identity pairs the summed force with (px, py, pz).
template <bool TYPED, bool MIXED, bool FAST>
__device__ __forceinline__ void sepk_lj_pairs(
	const sepk_state &s, const int *__restrict__ nbr, const int *__restrict__ cnt, long npad,
	const sepk_lj &c, long t, int sub, bool act, signed char ti, double &fx, double &fy,
	double &fz, double (&acc)[7], double &px, double &py, double &pz)
{
	if (MIXED) {
		// positions quantized to the box (sepk_quant): the integer
		// difference is the minimum image, one 16-byte gather per entry;
		// the stress identity uses the quantized positions x~ = u L / 2^32
		// and the image count k of the pair, nonzero when the unwrapped
		// difference borrows differently from the wrapped one
		int4 qi = make_int4(0, 0, 0, SEPK_QPHANTOM);
		if (act) {
			qi = s.qpos[t];
			bool band = false;
			const int nbat = (cnt[t] + SEPK_PAD - 1) / SEPK_PAD;
			const size_t stride = (size_t)npad * SEPK_SUB;
			const int4 *lp = sepk_wave_list(nbr, t, sub);
			const unsigned ln = threadIdx.x & 63;
			// batch 0 goes out with the count (every list allocation
			// holds it), and batch b + 1 while batch b is computed
			int4 jb = sepk_at(lp, ln);
			for (int b = 0; b < nbat; b++) {
				const int j[SEPK_BATCH] = {jb.x, jb.y, jb.z, jb.w};
				int4 qj[SEPK_BATCH];
				signed char tj[SEPK_BATCH];
#pragma unroll
				for (int u = 0; u < SEPK_BATCH; u++) {
					qj[u] = sepk_at(s.qpos, (unsigned)j[u]);
					if (TYPED)
						tj[u] = s.type[(unsigned)j[u]];
				}
				// the wave's last batch reloads itself (a cache hit)
				if (__ballot(b + 1 < nbat))
					lp += stride;
				jb = sepk_at(lp, ln);
#pragma unroll
				for (int u = 0; u < SEPK_BATCH; u++) {
					if (TYPED && !((ti == c.t0 && tj[u] == c.t1) ||
						       (ti == c.t1 && tj[u] == c.t0)))
						continue;
					sepk_lj_mixed_entry(c, qi, qj[u], fx, fy, fz, acc, band);
				}
			}
			if (band)
				sepk_lj_mixed_band<TYPED, FAST>(s, nbr, cnt, npad, c, t, sub, ti, qi, fx,
								fy, fz, acc);
		}
		px = act ? (double)(unsigned)qi.x * c.qd0 : 0.0;
		py = act ? (double)(unsigned)qi.y * c.qd1 : 0.0;
		pz = act ? (double)(unsigned)qi.z * c.qd2 : 0.0;
		return;
	}
	sepk_pos pi = {0.0, 0.0, 0.0, 0.0};
	if (act) {
		pi = s.pos[t];
		const int nbat = (cnt[t] + SEPK_PAD - 1) / SEPK_PAD;
		const size_t stride = (size_t)npad * SEPK_SUB;
		const int4 *lp = sepk_wave_list(nbr, t, sub);
		const unsigned ln = threadIdx.x & 63;
		const char *pb = (const char *)s.pos;
		// batch b + 1 in flight while batch b is computed, as above
		int4 jb = sepk_at(lp, ln);
		for (int b = 0; b < nbat; b++) {
			const int j[SEPK_BATCH] = {jb.x, jb.y, jb.z, jb.w};
			double px[SEPK_BATCH], py[SEPK_BATCH], pz[SEPK_BATCH];
			signed char tj[SEPK_BATCH];
#pragma unroll
			for (int u = 0; u < SEPK_BATCH; u++) {
				const sepk_pos &r = sepk_rec(pb, j[u]);
				px[u] = r.x;
				py[u] = r.y;
				pz[u] = r.z;
				if (TYPED)
					tj[u] = s.type[(unsigned)j[u]];
			}
			// the wave's last batch reloads itself (a cache hit)
			if (__ballot(b + 1 < nbat))
				lp += stride;
			jb = sepk_at(lp, ln);
#pragma unroll
			for (int u = 0; u < SEPK_BATCH; u++) {
				if (TYPED && !((ti == c.t0 && tj[u] == c.t1) ||
					       (ti == c.t1 && tj[u] == c.t0)))
					continue;
				double kx, ky, kz;
				const double dx = sepk_image<FAST>(pi.x - px[u], c.L0, c.iL0, kx);
				const double dy = sepk_image<FAST>(pi.y - py[u], c.L1, c.iL1, ky);
				const double dz = sepk_image<FAST>(pi.z - pz[u], c.L2, c.iL2, kz);
				// r^2 in the reference's order
				const double r2 = dx * dx + dy * dy + dz * dz;
				if (!(r2 < c.cf2))
					continue;
				const double q = c.sig2 * sepk_recip(r2);
				const double q3 = q * q * q;
				const double ft = fma(c.A, q3, c.nB) * q3 * q;
				const double e = fma(fma(c.C, q3, c.nD), q3, c.ushift);
				sepk_pair_accumulate(ft, e, dx, dy, dz, kx, ky, kz, c.L0, c.L1, c.L2,
						     fx, fy, fz, acc);
			}
		}
	}
	px = pi.x;
	py = pi.y;
	pz = pi.z;
}


template <bool TYPED, bool MIXED, bool OVERWRITE, bool FAST>
__global__ __launch_bounds__(SEPK_BLOCK) void k_lj(
	sepk_state s, const int *__restrict__ nbr, const int *__restrict__ cnt,
	long n, long npad, sepk_lj c, double *partial, long rows)
{
	const long lb = sepk_xcd_block();
	const long t = lb * (SEPK_BLOCK / SEPK_SUB) + threadIdx.x / SEPK_SUB;
	const int sub = threadIdx.x % SEPK_SUB;
	const bool frozen = *s.err != 0;
	double fx = 0.0, fy = 0.0, fz = 0.0;
	double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	bool act = t < n;
	signed char ti = 0;
	if (act && TYPED) {
		ti = s.type[t];
		act = ti == c.t0 || ti == c.t1;
	}
	double px = 0.0, py = 0.0, pz = 0.0;
	sepk_lj_pairs<TYPED, MIXED, FAST>(s, nbr, cnt, npad, c, t, sub, act, ti, fx, fy, fz, acc,
					   px, py, pz);
	sepk_pair_epilogue<OVERWRITE>(t, sub, n, npad, act, fx, fy, fz, px, py, pz, acc, s.f,
				      partial, lb, rows, frozen);
}

typedef void (*sepk_lj_kernel)(sepk_state, const int *, const int *, long, long,
			       sepk_lj, double *, long);

// qpos from pos for every slot (sepk_quant), after a position or charge
// upload or a change of box.
__global__ void k_quantize(sepk_state s, long npad, double q0, double q1, double q2)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < npad)
		s.qpos[t] = sepk_quant(s.pos[t], q0, q1, q2);
}

static int sepk_fresh_qpos(sepdev *d, const double *L)
{
	if (!d->qpos_stale && d->qL[0] == L[0] && d->qL[1] == L[1] && d->qL[2] == L[2])
		return 0;
	k_quantize<<<sepk_blocks(d->npad), SEPK_BLOCK, 0, d->st>>>(
		d->s[0], d->npad, SEPK_QSCALE / L[0], SEPK_QSCALE / L[1], SEPK_QSCALE / L[2]);
	SEPK_LAUNCH_CHECK();
	for (int k = 0; k < 3; k++)
		d->qL[k] = L[k];
	d->qpos_stale = 0;
	return 0;
}

template <bool T, bool M, bool O>
static sepk_lj_kernel pick_lj3(bool fast)
{
	return fast ? k_lj<T, M, O, true> : k_lj<T, M, O, false>;
}

template <bool T, bool M>
static sepk_lj_kernel pick_lj(bool over, bool fast)
{
	return over ? pick_lj3<T, M, true>(fast) : pick_lj3<T, M, false>(fast);
}

// The fast image is exact for listed pairs when half the box exceeds the
// list cutoff plus the skin with room to spare (see sepk_wrap_fast and
// sepk_image).
static bool fast_image_ok(const sepdev *d, const double *L)
{
	if (!d->list_built)
		return false;
	// listed pairs start within the list cutoff and each atom moves at
	// most half the skin before a rebuild; twice the cutoff bounds both
	const double reach = 2.0 * sqrt(d->grid.rlist2);
	for (int k = 0; k < 3; k++)
		if (!(0.5 * L[k] > 1.05 * reach))
			return false;
	return true;
}

// One pair-force launch and its finishing task.
static int launch_lj(sepdev *d, const sepk_lj &c, int typed, int mixed, bool fast, int ret_slot,
		     int epot_assign)
{
	// first: claiming a region may launch a recorded pair call, which
	// consumes the pending f = 0
	int rc, region;
	if ((rc = sepk_term_region(d, &region)) || (rc = sepk_fork_point(d)))
		return rc;
	// owed bonded gathers go to f first (the sums keep their order)
	if ((rc = sepk_flush_gathers(d)))
		return rc;
	const bool over = d->f_zero_pending != 0;
	sepk_lj_kernel k;
	if (typed)
		k = mixed ? pick_lj<true, true>(over, fast) : pick_lj<true, false>(over, fast);
	else
		k = mixed ? pick_lj<false, true>(over, fast) : pick_lj<false, false>(over, fast);
	if ((rc = sepk_timer_begin(d, &d->t_pair)))
		return rc;
	const unsigned nb = sepk_blocks_xcd(d->n * SEPK_SUB);
	double *part = sepk_region(d, region);
	k<<<nb, SEPK_BLOCK, 0, d->st>>>(d->s[0], d->nbr, d->nbr_count, d->n, d->npad, c, part,
					  d->region_rows);
	SEPK_LAUNCH_CHECK();
	d->f_zero_pending = 0;
	if ((rc = sepk_timer_end(d, &d->t_pair)))
		return rc;
	if (!d->last_lj)
		d->last_lj = malloc(sizeof(sepk_lj));
	memcpy(d->last_lj, &c, sizeof c);
	d->last_lj_kernel = (const void *)(typed ? (mixed ? pick_lj<true, true>(true, fast)
							  : pick_lj<true, false>(true, fast))
					       : (mixed ? pick_lj<false, true>(true, fast)
							: pick_lj<false, false>(true, fast)));
	d->last_fused = 0;
	sepk_task f;
	memset(&f, 0, sizeof f);
	f.mode = SEPK_FIN_PAIR;
	f.region = region;
	f.nblk = (int)nb;
	f.nv = 7;
	f.epot_assign = epot_assign;
	sepk_task_ret(d, f, ret_slot);
	return sepk_queue_task(d, f);
}

extern "C" int sepdev_lj(sepdev *d, const sepdev_lj_args *a, int ret_slot)
{
	if (!d->list_built)
		return sepk_fail(SEP_GPU_ESTATE, "pair forces before a neighbour list was built");
	if (ret_slot < 0 || ret_slot >= d->nret)
		return sepk_fail(SEP_GPU_EARG, "ret slot %d", ret_slot);
	sepk_lj c;
	c.cf2 = a->cf2;
	c.A = a->eps48;
	c.nB = -(a->eps48 * a->awh);
	c.C = a->eps4;
	c.nD = -(a->eps4 * a->aw);
	c.sig2 = a->sig2;
	c.ushift = a->ushift;
	c.L0 = a->L[0];
	c.L1 = a->L[1];
	c.L2 = a->L[2];
	c.iL0 = 1.0 / a->L[0];
	c.iL1 = 1.0 / a->L[1];
	c.iL2 = 1.0 / a->L[2];
	c.qd0 = a->L[0] / SEPK_QSCALE;
	c.qd1 = a->L[1] / SEPK_QSCALE;
	c.qd2 = a->L[2] / SEPK_QSCALE;
	{
		// |dx~ - dx| <= one step per axis, so |r2~ - r2| <= q (2 sqrt(3) rc + 3 q)
		const double q = fmax(c.qd0, fmax(c.qd1, c.qd2));
		c.band = q * (4.0 * sqrt(a->cf2) + 4.0 * q) + 1e-12 * a->cf2;
		c.cf2lo = a->cf2 - c.band;
		c.cf2hi = a->cf2 + c.band;
	}
	c.Af = (float)c.A;
	c.nBf = (float)c.nB;
	c.Cf = (float)c.C;
	c.nDf = (float)c.nD;
	c.sig2f = (float)a->sig2;
	c.ushiftf = (float)a->ushift;
	c.t0 = a->t0;
	c.t1 = a->t1;
	int rc = sepk_flush_nh(d);
	if (rc || (a->mixed && (rc = sepk_fresh_qpos(d, a->L))))
		return rc;
	const bool over = d->f_zero_pending != 0;
	const bool fast = fast_image_ok(d, a->L);
	if (d->fuse_armed && !a->typed && over && !d->fused_pending) {
		// recorded: the step's leapfrog launches it fused with itself
		d->fused_pending = 1;
		d->fused_lj = c;
		d->fused_mixed = a->mixed;
		d->fused_fast = fast;
		d->fused_ret = ret_slot;
		d->fused_epot_assign = a->epot_assign;
		return 0;
	}
	if (d->pc_armed && !d->pc_pending && !d->fused_pending) {
		// recorded: a Coulomb call right after launches both terms
		d->pc_pending = 1;
		d->pc_lj = c;
		d->pc_mixed = a->mixed;
		d->pc_typed = a->typed;
		d->pc_fast = fast;
		d->pc_ret = ret_slot;
		d->pc_epot_assign = a->epot_assign;
		return 0;
	}
	return launch_lj(d, c, a->typed, a->mixed, fast, ret_slot, a->epot_assign);
}

// Launches a recorded pair call on its own (something other than the
// step's leapfrog needs its forces or its rows).
int sepk_flush_pair(sepdev *d)
{
	if (d->pc_pending) {
		d->pc_pending = 0;
		return launch_lj(d, d->pc_lj, d->pc_typed, d->pc_mixed, d->pc_fast != 0, d->pc_ret,
				 d->pc_epot_assign);
	}
	if (!d->fused_pending)
		return 0;
	d->fused_pending = 0;
	return launch_lj(d, d->fused_lj, 0, d->fused_mixed, d->fused_fast, d->fused_ret,
			 d->fused_epot_assign);
}

extern "C" int sepdev_fuse_step(sepdev *d, int on)
{
	d->fuse_armed = on != 0;
	return on ? 0 : sepk_flush_pair(d);
}

extern "C" int sepdev_fuse_pair_coulomb(sepdev *d, int on)
{
	d->pc_armed = on != 0;
	return on ? 0 : sepk_flush_pair(d);
}


// ---- shifted-force Coulomb (source/sepcoulomb.c:96-160) ---------------------

struct sepk_coul {
	double cf, cf2, icf, icf2;
	double L0, L1, L2, iL0, iL1, iL2;
	double qd0, qd1, qd2; // quantization step per axis (mixed form)
	float cff, icff, icf2f;
};

// The mixed form of the pair loop below: positions and charges from the
// 16-byte quantized records (integer minimum image, charge as FP32 in w),
// r^2 and the cutoff test in FP64 from the exact integer differences, the
// pair force and energy in FP32 (hardware reciprocal square root, factored
// forms of the shifted-force terms), sums in FP64.  The charge test is the reference's |q| < DBL_EPSILON; the phantom's
// NaN charge fails it too.
__device__ __forceinline__ void sepk_coul_pairs_mixed(
	const sepk_state &s, const int *__restrict__ nbr, const int *__restrict__ cnt,
	long npad, const sepk_coul &c, long t, int sub, int4 qi, double &fx, double &fy,
	double &fz, double (&acc)[7])
{
	const float qif = __int_as_float(qi.w);
	const float epsf = (float)DBL_EPSILON;
	const int nbat = (cnt[t] + SEPK_PAD - 1) / SEPK_PAD;
	const size_t stride = (size_t)npad * SEPK_SUB;
	const int4 *lp = sepk_lane_list(nbr, t, sub);
	for (int b = 0; b < nbat; b++, lp += stride) {
		const int4 jb = *lp;
		const int j[SEPK_BATCH] = {jb.x, jb.y, jb.z, jb.w};
		int4 qj[SEPK_BATCH];
#pragma unroll
		for (int u = 0; u < SEPK_BATCH; u++)
			qj[u] = sepk_at(s.qpos, (unsigned)j[u]);
#pragma unroll
		for (int u = 0; u < SEPK_BATCH; u++) {
			const float qjf = __int_as_float(qj[u].w);
			if (!(fabsf(qjf) >= epsf))
				continue;
			const unsigned ux = (unsigned)qi.x, vx = (unsigned)qj[u].x;
			const unsigned uy = (unsigned)qi.y, vy = (unsigned)qj[u].y;
			const unsigned uz = (unsigned)qi.z, vz = (unsigned)qj[u].z;
			const int ix = (int)(ux - vx), iy = (int)(uy - vy), iz = (int)(uz - vz);
			const double dx = (double)ix * c.qd0;
			const double dy = (double)iy * c.qd1;
			const double dz = (double)iz * c.qd2;
			const double r2 = dx * dx + dy * dy + dz * dz;
			if (!(r2 < c.cf2))
				continue;
			// 1/r^2 - 1/rc^2 = (rc - r)(rc + r) / (r^2 rc^2) and
			// 1/r - 1/rc + (r - rc)/rc^2 = (rc - r)^2 / (r rc^2): the
			// reference's expressions without their cancellation near
			// the cutoff, which FP32 would carry into the energy sum
			const float r2f = (float)r2;
			const float ri = __builtin_amdgcn_rsqf(r2f);
			const float rf = r2f * ri;
			const float qq = qjf * qif;
			const float dc = c.cff - rf;
			const float qs = qq * c.icf2f * ri;
			const double ft = (double)(qs * (dc * (c.cff + rf)) * (ri * ri));
			const double e = (double)(qs * (dc * dc));
			fx = fma(ft, dx, fx);
			fy = fma(ft, dy, fy);
			fz = fma(ft, dz, fz);
			acc[0] += e;
			const bool bx = (ux < vx) != (ix < 0), by = (uy < vy) != (iy < 0),
				   bz = (uz < vz) != (iz < 0);
			if (bx | by | bz) {
				const double kx = bx ? (ix < 0 ? 1.0 : -1.0) : 0.0;
				const double ky = by ? (iy < 0 ? 1.0 : -1.0) : 0.0;
				const double kz = bz ? (iz < 0 ? 1.0 : -1.0) : 0.0;
				sepk_pair_cross(ft, dx, dy, dz, kx, ky, kz, c.L0, c.L1, c.L2, acc);
			}
		}
	}
}

template <bool OVERWRITE, bool FAST, bool MIXED>
__global__ __launch_bounds__(SEPK_BLOCK) void k_coulomb(
	sepk_state s, const int *__restrict__ nbr, const int *__restrict__ cnt,
	long n, long npad, sepk_coul c, double *partial, long rows)
{
	const long lb = sepk_xcd_block();
	const long t = lb * (SEPK_BLOCK / SEPK_SUB) + threadIdx.x / SEPK_SUB;
	const int sub = threadIdx.x % SEPK_SUB;
	const bool frozen = *s.err != 0;
	double fx = 0.0, fy = 0.0, fz = 0.0;
	double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	bool act = t < n;
	sepk_pos pi = {0.0, 0.0, 0.0, 0.0};
	if (act) {
		pi = s.pos[t];
		act = !(fabs(pi.q) < DBL_EPSILON);
	}
	if (MIXED) {
		int4 qi = make_int4(0, 0, 0, SEPK_QPHANTOM);
		if (act) {
			qi = s.qpos[t];
			sepk_coul_pairs_mixed(s, nbr, cnt, npad, c, t, sub, qi, fx, fy, fz, acc);
		}
		// the stress identity on the quantized positions
		const double px = act ? (double)(unsigned)qi.x * c.qd0 : 0.0;
		const double py = act ? (double)(unsigned)qi.y * c.qd1 : 0.0;
		const double pz = act ? (double)(unsigned)qi.z * c.qd2 : 0.0;
		sepk_pair_epilogue<OVERWRITE>(t, sub, n, npad, act, fx, fy, fz, px, py, pz, acc,
					      s.f, partial, lb, rows, frozen);
		return;
	}
	if (act) {
		const int nbat = (cnt[t] + SEPK_PAD - 1) / SEPK_PAD;
		const size_t stride = (size_t)npad * SEPK_SUB;
		const int4 *lp = sepk_lane_list(nbr, t, sub);
		const char *pb = (const char *)s.pos;
		for (int b = 0; b < nbat; b++, lp += stride) {
			const int4 jb = *lp;
			const int j[SEPK_BATCH] = {jb.x, jb.y, jb.z, jb.w};
			sepk_pos pj[SEPK_BATCH];
#pragma unroll
			for (int u = 0; u < SEPK_BATCH; u++)
				pj[u] = sepk_rec(pb, j[u]);
#pragma unroll
			for (int u = 0; u < SEPK_BATCH; u++) {
				// the phantom (charge 0) is skipped here
				if (fabs(pj[u].q) < DBL_EPSILON)
					continue;
				double kx, ky, kz;
				const double dx = sepk_image<FAST>(pi.x - pj[u].x, c.L0, c.iL0, kx);
				const double dy = sepk_image<FAST>(pi.y - pj[u].y, c.L1, c.iL1, ky);
				const double dz = sepk_image<FAST>(pi.z - pj[u].z, c.L2, c.iL2, kz);
				const double r2 = dx * dx + dy * dy + dz * dz;
				if (!(r2 < c.cf2))
					continue;
				const double qq = pj[u].q * pi.q;
				const double r = sqrt(r2);
				const double ft = qq * (1.0 / r2 - c.icf2) / r;
				const double e = qq * (1.0 / r + (r - c.cf) * c.icf2 - c.icf);
				sepk_pair_accumulate(ft, e, dx, dy, dz, kx, ky, kz, c.L0, c.L1, c.L2,
						     fx, fy, fz, acc);
			}
		}
	}
	sepk_pair_epilogue<OVERWRITE>(t, sub, n, npad, act, fx, fy, fz, pi.x, pi.y, pi.z, acc,
				      s.f, partial, lb, rows, frozen);
}

// A pair call (LJ, typed or not) and the shifted-force Coulomb call right
// after it, in either form, in one walk of the list: one gather of each
// neighbour's record serves both terms.  Each term keeps its own force sum,
// energy, crossing stress and partial rows, accumulated entry by entry in
// the order of its own kernel (k_lj, k_coulomb), and the atom's force is
// written as the two calls would leave it: the LJ sum (over or added to f),
// then the Coulomb sum added.  The results are those of the two launches,
// bit for bit.
template <bool TYPED, bool OVERWRITE, bool MIXED, bool FAST>
__global__ __launch_bounds__(SEPK_BLOCK) void k_lj_coul(
	sepk_state s, const int *__restrict__ nbr, const int *__restrict__ cnt, long n,
	long npad, sepk_lj c, sepk_coul cc, double *part_lj, double *part_c, long rows)
{
	const long lb = sepk_xcd_block();
	const long t = lb * (SEPK_BLOCK / SEPK_SUB) + threadIdx.x / SEPK_SUB;
	const int sub = threadIdx.x % SEPK_SUB;
	const bool frozen = *s.err != 0;
	double fl[3] = {0.0, 0.0, 0.0}, fc[3] = {0.0, 0.0, 0.0};
	double al[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	double ac[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	bool act_l = t < n, act_c = t < n;
	signed char ti = 0;
	if (act_l && TYPED) {
		ti = s.type[t];
		act_l = ti == c.t0 || ti == c.t1;
	}
	sepk_pos pi = {0.0, 0.0, 0.0, 0.0};
	if (act_l || act_c)
		pi = s.pos[t];
	if (act_c)
		act_c = !(fabs(pi.q) < DBL_EPSILON);
	int4 qi = make_int4(0, 0, 0, SEPK_QPHANTOM);
	if (!MIXED && (act_l || act_c)) {
		// the FP64 loops of sepk_lj_pairs and k_coulomb on 32-byte records
		const int nbat = (cnt[t] + SEPK_PAD - 1) / SEPK_PAD;
		const size_t stride = (size_t)npad * SEPK_SUB;
		const int4 *lp = sepk_wave_list(nbr, t, sub);
		const unsigned ln = threadIdx.x & 63;
		const char *pb = (const char *)s.pos;
		// batch b + 1 in flight while batch b is computed, as above
		int4 jb = sepk_at(lp, ln);
		for (int b = 0; b < nbat; b++) {
			const int j[SEPK_BATCH] = {jb.x, jb.y, jb.z, jb.w};
			sepk_pos pj[SEPK_BATCH];
			signed char tj[SEPK_BATCH];
#pragma unroll
			for (int u = 0; u < SEPK_BATCH; u++) {
				pj[u] = sepk_rec(pb, j[u]);
				if (TYPED)
					tj[u] = s.type[(unsigned)j[u]];
			}
			// the wave's last batch reloads itself (a cache hit)
			if (__ballot(b + 1 < nbat))
				lp += stride;
			jb = sepk_at(lp, ln);
#pragma unroll
			for (int u = 0; u < SEPK_BATCH; u++) {
				double kx, ky, kz;
				const double dx = sepk_image<FAST>(pi.x - pj[u].x, c.L0, c.iL0, kx);
				const double dy = sepk_image<FAST>(pi.y - pj[u].y, c.L1, c.iL1, ky);
				const double dz = sepk_image<FAST>(pi.z - pj[u].z, c.L2, c.iL2, kz);
				const double r2 = dx * dx + dy * dy + dz * dz;
				const bool tl = !TYPED || (ti == c.t0 && tj[u] == c.t1) ||
						(ti == c.t1 && tj[u] == c.t0);
				if (act_l && tl && r2 < c.cf2) {
					const double q = c.sig2 * sepk_recip(r2);
					const double q3 = q * q * q;
					const double ft = fma(c.A, q3, c.nB) * q3 * q;
					const double e = fma(fma(c.C, q3, c.nD), q3, c.ushift);
					sepk_pair_accumulate(ft, e, dx, dy, dz, kx, ky, kz, c.L0, c.L1, c.L2,
							     fl[0], fl[1], fl[2], al);
				}
				if (act_c && !(fabs(pj[u].q) < DBL_EPSILON) && r2 < cc.cf2) {
					const double qq = pj[u].q * pi.q;
					const double r = sqrt(r2);
					const double ft = qq * (1.0 / r2 - cc.icf2) / r;
					const double e = qq * (1.0 / r + (r - cc.cf) * cc.icf2 - cc.icf);
					sepk_pair_accumulate(ft, e, dx, dy, dz, kx, ky, kz, cc.L0, cc.L1,
							     cc.L2, fc[0], fc[1], fc[2], ac);
				}
			}
		}
	}
	if (MIXED && (act_l || act_c)) {
		qi = s.qpos[t];
		const float qif = __int_as_float(qi.w);
		const float epsf = (float)DBL_EPSILON;
		const int nbat = (cnt[t] + SEPK_PAD - 1) / SEPK_PAD;
		const size_t stride = (size_t)npad * SEPK_SUB;
		const int4 *lp = sepk_wave_list(nbr, t, sub);
		const unsigned ln = threadIdx.x & 63;
		bool band = false;
		int4 jb = sepk_at(lp, ln); // batch b + 1 in flight, as above
		for (int b = 0; b < nbat; b++) {
			const int j[SEPK_BATCH] = {jb.x, jb.y, jb.z, jb.w};
			int4 qj[SEPK_BATCH];
			signed char tj[SEPK_BATCH];
#pragma unroll
			for (int u = 0; u < SEPK_BATCH; u++) {
				qj[u] = sepk_at(s.qpos, (unsigned)j[u]);
				if (TYPED)
					tj[u] = s.type[(unsigned)j[u]];
			}
			// the wave's last batch reloads itself (a cache hit)
			if (__ballot(b + 1 < nbat))
				lp += stride;
			jb = sepk_at(lp, ln);
#pragma unroll
			for (int u = 0; u < SEPK_BATCH; u++) {
				const unsigned ux = (unsigned)qi.x, vx = (unsigned)qj[u].x;
				const unsigned uy = (unsigned)qi.y, vy = (unsigned)qj[u].y;
				const unsigned uz = (unsigned)qi.z, vz = (unsigned)qj[u].z;
				const int ix = (int)(ux - vx), iy = (int)(uy - vy), iz = (int)(uz - vz);
				const double dx = (double)ix * c.qd0;
				const double dy = (double)iy * c.qd1;
				const double dz = (double)iz * c.qd2;
				const double r2 = dx * dx + dy * dy + dz * dz;
				const bool bx = (ux < vx) != (ix < 0), by = (uy < vy) != (iy < 0),
					   bz = (uz < vz) != (iz < 0);
				const double kx = bx ? (ix < 0 ? 1.0 : -1.0) : 0.0;
				const double ky = by ? (iy < 0 ? 1.0 : -1.0) : 0.0;
				const double kz = bz ? (iz < 0 ? 1.0 : -1.0) : 0.0;
				// the LJ term as in sepk_lj_pairs, pairs near the cutoff
				// left to sepk_lj_mixed_band after the loop
				const bool tl = !TYPED || (ti == c.t0 && tj[u] == c.t1) ||
						(ti == c.t1 && tj[u] == c.t0);
				const bool lj = act_l && tl && qj[u].w != SEPK_QPHANTOM;
				if (lj && !(r2 < c.cf2lo))
					band |= r2 < c.cf2hi;
				if (lj && r2 < c.cf2lo) {
					const float q = c.sig2f * __builtin_amdgcn_rcpf((float)r2);
					const float q3 = q * q * q;
					const double ft = (double)(fmaf(c.Af, q3, c.nBf) * q3 * q);
					const double e = (double)fmaf(fmaf(c.Cf, q3, c.nDf), q3, c.ushiftf);
					fl[0] = fma(ft, dx, fl[0]);
					fl[1] = fma(ft, dy, fl[1]);
					fl[2] = fma(ft, dz, fl[2]);
					al[0] += e;
					if (bx | by | bz)
						sepk_pair_cross(ft, dx, dy, dz, kx, ky, kz, c.L0, c.L1, c.L2, al);
				}
				// the Coulomb term as in sepk_coul_pairs_mixed
				const float qjf = __int_as_float(qj[u].w);
				if (act_c && fabsf(qjf) >= epsf && r2 < cc.cf2) {
					const float r2f = (float)r2;
					const float ri = __builtin_amdgcn_rsqf(r2f);
					const float rf = r2f * ri;
					const float qq = qjf * qif;
					const float dc = cc.cff - rf;
					const float qs = qq * cc.icf2f * ri;
					const double ft = (double)(qs * (dc * (cc.cff + rf)) * (ri * ri));
					const double e = (double)(qs * (dc * dc));
					fc[0] = fma(ft, dx, fc[0]);
					fc[1] = fma(ft, dy, fc[1]);
					fc[2] = fma(ft, dz, fc[2]);
					ac[0] += e;
					if (bx | by | bz)
						sepk_pair_cross(ft, dx, dy, dz, kx, ky, kz, cc.L0, cc.L1, cc.L2, ac);
				}
			}
		}
		if (band)
			sepk_lj_mixed_band<TYPED, FAST>(s, nbr, cnt, npad, c, t, sub, ti, qi, fl[0],
							fl[1], fl[2], al);
	}
	// the two epilogues of sepk_pair_epilogue, the second adding to the first
#pragma unroll
	for (int o = 1; o < SEPK_SUB; o <<= 1)
#pragma unroll
		for (int k = 0; k < 3; k++) {
			fl[k] += __shfl_xor(fl[k], o, SEPK_WAVE);
			fc[k] += __shfl_xor(fc[k], o, SEPK_WAVE);
		}
	if (t < n && sub == 0 && !frozen) {
		for (int k = 0; k < 3; k++) {
			double v = OVERWRITE ? fl[k] : s.f[k * npad + t];
			if (!OVERWRITE && act_l)
				v += fl[k];
			if (act_c)
				v += fc[k];
			s.f[k * npad + t] = v;
		}
	}
	// the stress identity's positions: quantized (mixed) or FP64
	const double px = MIXED ? (double)(unsigned)qi.x * c.qd0 : pi.x;
	const double py = MIXED ? (double)(unsigned)qi.y * c.qd1 : pi.y;
	const double pz = MIXED ? (double)(unsigned)qi.z * c.qd2 : pi.z;
	if (act_l && sub == 0) {
		const double gx = 2.0 * fl[0], gy = 2.0 * fl[1], gz = 2.0 * fl[2];
		al[1] = fma(gx, px, al[1]);
		al[2] = fma(gx, py, al[2]);
		al[3] = fma(gx, pz, al[3]);
		al[4] = fma(gy, py, al[4]);
		al[5] = fma(gy, pz, al[5]);
		al[6] = fma(gz, pz, al[6]);
	}
	if (act_c && sub == 0) {
		const double gx = 2.0 * fc[0], gy = 2.0 * fc[1], gz = 2.0 * fc[2];
		ac[1] = fma(gx, px, ac[1]);
		ac[2] = fma(gx, py, ac[2]);
		ac[3] = fma(gx, pz, ac[3]);
		ac[4] = fma(gy, py, ac[4]);
		ac[5] = fma(gy, pz, ac[5]);
		ac[6] = fma(gz, pz, ac[6]);
	}
	sepk_block_partial<7>(al, 0u, part_lj, lb, rows);
	__syncthreads(); // the partial helper's LDS is reused
	sepk_block_partial<7>(ac, 0u, part_c, lb, rows);
}

// The recorded pair call and this Coulomb call as one k_lj_coul launch,
// each with its own term region and finishing task, queued in call order.
static int launch_lj_coul(sepdev *d, const sepk_coul &cc, int ret_slot)
{
	d->pc_pending = 0;
	int rc;
	// two distinct free term regions (the queue is finished first if not)
	int r1 = -1, r2 = -1;
	for (int r = SEPK_R_PAIR; r < SEPK_R_PAIR + SEPK_NTERM && r2 < 0; r++) {
		bool held = false;
		for (int i = 0; i < d->npending; i++)
			held = held || d->pending[i].region == r;
		if (!held) {
			if (r1 < 0)
				r1 = r;
			else
				r2 = r;
		}
	}
	if (r2 < 0) {
		d->fork_ready = 0;
		if ((rc = sepk_flush_tasks(d)))
			return rc;
		r1 = SEPK_R_PAIR;
		r2 = SEPK_R_PAIR + 1;
	}
	if ((rc = sepk_fork_point(d)) || (rc = sepk_flush_gathers(d)))
		return rc;
	const bool over = d->f_zero_pending != 0;
	typedef void (*kfn)(sepk_state, const int *, const int *, long, long, sepk_lj, sepk_coul,
			    double *, double *, long);
	static const kfn table[2][2][2][2] = {
		{{{k_lj_coul<false, false, false, false>, k_lj_coul<false, false, false, true>},
		  {k_lj_coul<false, false, true, false>, k_lj_coul<false, false, true, true>}},
		 {{k_lj_coul<false, true, false, false>, k_lj_coul<false, true, false, true>},
		  {k_lj_coul<false, true, true, false>, k_lj_coul<false, true, true, true>}}},
		{{{k_lj_coul<true, false, false, false>, k_lj_coul<true, false, false, true>},
		  {k_lj_coul<true, false, true, false>, k_lj_coul<true, false, true, true>}},
		 {{k_lj_coul<true, true, false, false>, k_lj_coul<true, true, false, true>},
		  {k_lj_coul<true, true, true, false>, k_lj_coul<true, true, true, true>}}}};
	const kfn k = table[d->pc_typed ? 1 : 0][over ? 1 : 0][d->pc_mixed ? 1 : 0][d->pc_fast ? 1 : 0];
	const unsigned nb = sepk_blocks_xcd(d->n * SEPK_SUB);
	k<<<nb, SEPK_BLOCK, 0, d->st>>>(d->s[0], d->nbr, d->nbr_count, d->n, d->npad, d->pc_lj, cc,
					  sepk_region(d, r1), sepk_region(d, r2), d->region_rows);
	SEPK_LAUNCH_CHECK();
	d->f_zero_pending = 0;
	d->last_fused = 0;
	sepk_task f;
	memset(&f, 0, sizeof f);
	f.mode = SEPK_FIN_PAIR;
	f.region = r1;
	f.nblk = (int)nb;
	f.nv = 7;
	f.epot_assign = d->pc_epot_assign;
	sepk_task_ret(d, f, d->pc_ret);
	if ((rc = sepk_queue_task(d, f)))
		return rc;
	memset(&f, 0, sizeof f);
	f.mode = SEPK_FIN_PAIR;
	f.region = r2;
	f.nblk = (int)nb;
	f.nv = 7;
	f.also_ecoul = 1;
	sepk_task_ret(d, f, ret_slot);
	return sepk_queue_task(d, f);
}

extern "C" int sepdev_coulomb(sepdev *d, const sepdev_coul_args *a, int ret_slot)
{
	if (!d->list_built)
		return sepk_fail(SEP_GPU_ESTATE, "Coulomb forces before a neighbour list was built");
	if (ret_slot < 0 || ret_slot >= d->nret)
		return sepk_fail(SEP_GPU_EARG, "ret slot %d", ret_slot);
	sepk_coul c = {a->cf, a->cf2, a->icf, a->icf2, a->L[0], a->L[1], a->L[2],
		       1.0 / a->L[0], 1.0 / a->L[1], 1.0 / a->L[2],
		       a->L[0] / SEPK_QSCALE, a->L[1] / SEPK_QSCALE, a->L[2] / SEPK_QSCALE,
		       (float)a->cf, (float)a->icf, (float)a->icf2};
	int rc = sepk_flush_nh(d);
	if (rc || (a->mixed && (rc = sepk_fresh_qpos(d, a->L))))
		return rc;
	// the recorded pair call joins this one when both read the same
	// records with the same image rule
	if (d->pc_pending && (a->mixed != 0) == (d->pc_mixed != 0) &&
	    (a->mixed || (fast_image_ok(d, a->L) ? 1 : 0) == (d->pc_fast ? 1 : 0)))
		return launch_lj_coul(d, c, ret_slot);
	int region;
	if ((rc = sepk_term_region(d, &region)) || (rc = sepk_fork_point(d)) ||
	    (rc = sepk_flush_gathers(d)))
		return rc;
	const bool over = d->f_zero_pending != 0;
	const bool fast = fast_image_ok(d, a->L);
	auto k = a->mixed ? (over ? k_coulomb<true, false, true> : k_coulomb<false, false, true>)
		 : over   ? (fast ? k_coulomb<true, true, false> : k_coulomb<true, false, false>)
			  : (fast ? k_coulomb<false, true, false> : k_coulomb<false, false, false>);
	const unsigned nb = sepk_blocks_xcd(d->n * SEPK_SUB);
	k<<<nb, SEPK_BLOCK, 0, d->st>>>(d->s[0], d->nbr, d->nbr_count, d->n, d->npad, c,
					  sepk_region(d, region), d->region_rows);
	SEPK_LAUNCH_CHECK();
	d->f_zero_pending = 0;
	sepk_task f;
	memset(&f, 0, sizeof f);
	f.mode = SEPK_FIN_PAIR;
	f.region = region;
	f.nblk = (int)nb;
	f.nv = 7;
	f.also_ecoul = 1;
	sepk_task_ret(d, f, ret_slot);
	return sepk_queue_task(d, f);
}

// ---- Nose-Hoover (source/sepintgr.c:149-168) ---------------------------------
//
// When the sum of m v^2 of the current velocities is known (the previous
// leapfrog computed it) the call is only recorded and the next leapfrog
// applies f -= alpha m v itself; otherwise: sum kernel, finisher, apply.

// Sum of m v^2 over the atoms, with the leapfrog's blocks (SEPK_SUB lanes
// per atom, lane 0 working, XCD-remapped block index) and its order of
// terms, so that it is the same bits as the sum a leapfrog leaves in
// sys->sumv2 for the same velocities.
__global__ __launch_bounds__(SEPK_BLOCK) void k_nh_sum(sepk_state s, long n, long npad,
							double *partial, long rows)
{
	if (*s.err)
		return; // the finisher applies nothing either
	const long lb = sepk_xcd_block();
	const long t = lb * (SEPK_BLOCK / SEPK_SUB) + threadIdx.x / SEPK_SUB;
	double acc[1] = {0.0};
	if (t < n && threadIdx.x % SEPK_SUB == 0) {
		const double m = s.m[t];
		const double vx = s.v[t], vy = s.v[npad + t], vz = s.v[2 * npad + t];
		acc[0] = (vx * vx * m + vy * vy * m) + vz * vz * m;
	}
	sepk_block_partial<1>(acc, 0u, partial, lb, rows);
}

__global__ void k_nh_apply(sepk_state s, long n, long npad, const double *alpha)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n || *s.err)
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
	int rc = sepk_flush_nh(d);
	if (rc)
		return rc;
	if (d->sumv2_valid) {
		d->nh_pending = 1;
		d->nh_slot = alpha_slot;
		d->nh_Td = Td;
		d->nh_dt_tau2 = dt / (tau * tau);
		return 0;
	}
	if ((rc = sepk_materialize_forces(d)))
		return rc;
	const unsigned nb = sepk_blocks_xcd(d->n * SEPK_SUB);
	if ((rc = sepk_flush_tasks(d)))
		return rc;
	k_nh_sum<<<nb, SEPK_BLOCK, 0, d->st>>>(d->s[0], d->n, d->npad,
						sepk_region(d, SEPK_R_MISC), d->region_rows);
	SEPK_LAUNCH_CHECK();
	sepk_task f;
	memset(&f, 0, sizeof f);
	f.mode = SEPK_FIN_NH;
	f.region = SEPK_R_MISC;
	f.nblk = (int)nb;
	f.nv = 1;
	f.alpha = d->scal + alpha_slot;
	f.Td = Td;
	f.dt_tau2 = dt / (tau * tau);
	f.npart = (double)d->n;
	if ((rc = sepk_queue_task(d, f)) || (rc = sepk_flush_tasks(d)))
		return rc;
	d->sumv2_valid = 1;
	k_nh_apply<<<sepk_blocks(d->n), SEPK_BLOCK, 0, d->st>>>(d->s[0], d->n, d->npad,
								 d->scal + alpha_slot);
	SEPK_LAUNCH_CHECK();
	return 0;
}

// ---- leapfrog (source/sepintgr.c:18-88) ---------------------------------------
//
// Per atom: optional fused Nose-Hoover force term, a = f/m, v += a dt,
// x += v dt, wrap once per axis counting crossings, squared displacement
// from xn.  Row: {ekin sum, kin_P (6), max d2, sum m v_new^2}.

// Output arrays of the per-atom integrator step: the state's own arrays,
// except the positions of a fused step (the spare buffer, while other atoms
// still gather the old ones) and a timing run's scratch.
struct sepk_lfout {
	double *v, *f, *a, *xn;
	int *cn, *cr;
	sepk_pos *pos;
	int4 *qpos;
};

// sep_nosehoover's force term and sep_leapfrog for atom t
// (source/sepintgr.c:149-168, :42-88) from its force before the thermostat
// term, spread over lanes 0..2 of the atom's SEPK_SUB lanes, one axis each
// (lane k is given component k of the force as Fk).  Lane 0 gathers the
// other axes' values by shuffles and forms the sums in axis order.  acc:
// m vh^2 summed, the six m vh (x) vh components, d^2 and m v^2.  Every lane
// of the wave calls it.
__device__ __forceinline__ void sepk_lf_atom_lanes(const sepk_state &s, const sepk_lfout &o,
						   long t, long npad, int sub, bool act, double F0,
						   double F1, double F2, double dt, double L0,
						   double L1, double L2, int nh, double al,
						   int pending, double qs0, double qs1, double qs2,
						   double (&acc)[9])
{
	double x = 0.0, v = 0.0, vh = 0.0, dr = 0.0, m = 0.0, q = 0.0;
	if (act && sub < 3) {
		const int k = sub;
		const sepk_pos p = s.pos[t];
		q = p.q;
		x = k == 0 ? p.x : (k == 1 ? p.y : p.z);
		const double L = k == 0 ? L0 : (k == 1 ? L1 : L2);
		const double Fk = k == 0 ? F0 : (k == 1 ? F1 : F2);
		m = s.m[t];
		const long i = k * npad + t;
		double xn = s.xn[i];
		int cn = s.cn[i];
		if (pending) {
			xn = x;
			cn = 0;
			o.xn[i] = xn;
		}
		v = s.v[i];
		double f = Fk;
		if (nh)
			f -= al * m * v;
		const double a = f / m;
		v += a * dt;
		x += v * dt;
		vh = v - 0.5 * a * dt;
		int cr = s.cr[i];
		if (x > L) {
			x -= L;
			cn++;
			cr++;
		} else if (x < 0.0) {
			x += L;
			cn--;
			cr--;
		}
		dr = (x + cn * L) - xn;
		o.f[i] = f;
		o.a[i] = a;
		o.v[i] = v;
		o.cn[i] = cn;
		o.cr[i] = cr;
	}
	const int base = (int)(threadIdx.x & (SEPK_WAVE - 1)) - sub;
	const double x1 = __shfl(x, base + 1, SEPK_WAVE), x2 = __shfl(x, base + 2, SEPK_WAVE);
	const double v1 = __shfl(v, base + 1, SEPK_WAVE), v2 = __shfl(v, base + 2, SEPK_WAVE);
	const double h1 = __shfl(vh, base + 1, SEPK_WAVE), h2 = __shfl(vh, base + 2, SEPK_WAVE);
	const double r1 = __shfl(dr, base + 1, SEPK_WAVE), r2 = __shfl(dr, base + 2, SEPK_WAVE);
	if (act && sub == 0) {
		acc[0] = (vh * vh * m + h1 * h1 * m) + h2 * h2 * m;
		acc[8] = (v * v * m + v1 * v1 * m) + v2 * v2 * m;
		acc[1] = vh * vh * m;
		acc[2] = vh * h1 * m;
		acc[3] = vh * h2 * m;
		acc[4] = h1 * h1 * m;
		acc[5] = h1 * h2 * m;
		acc[6] = h2 * h2 * m;
		acc[7] = (dr * dr + r1 * r1) + r2 * r2;
		const sepk_pos p = {x, x1, x2, q};
		o.pos[t] = p;
		o.qpos[t] = sepk_quant(p, qs0, qs1, qs2);
	}
}

// The leapfrog on its own.  Its atoms map onto blocks exactly as the pair
// kernels' (SEPK_SUB lanes per atom, lane 0 working, XCD-remapped block
// index), so that its rows -- and hence every reduced sum -- are the same
// bits as those of the fused step kernel.
__global__ __launch_bounds__(SEPK_BLOCK) void k_leapfrog(
	sepk_state s, long n, long npad, double dt, double L0, double L1, double L2,
	const sepdev_sys *sys, int nh, const double *alpha, double Td, double dt_tau2,
	double *partial, long rows, double qs0, double qs1, double qs2, sepk_gpend gp)
{
	if (sys->error)
		return; // positions are updated in place: nothing to copy
	const long lb = sepk_xcd_block();
	const long t = lb * (SEPK_BLOCK / SEPK_SUB) + threadIdx.x / SEPK_SUB;
	const int sub = threadIdx.x % SEPK_SUB;
	double acc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	const int pending = sys->xn_pending;
	double al = 0.0;
	if (nh)
		al = sepk_nh_alpha(*alpha, sys->sumv2, (double)n, Td, dt_tau2);
	// one axis a lane, as in the fused step kernel (sepk_lf_atom_lanes), so
	// the two give the same bits; lane k needs only force component k, to
	// which it adds the owed bonded gathers as k_bonded_gather would
	const bool act = t < n;
	double Fk = 0.0;
	if (act && sub < 3) {
		Fk = gp.over ? 0.0 : s.f[sub * npad + t];
		if (gp.n) {
			const int o = s.orig[t];
			for (int q = 0; q < gp.n; q++)
				for (int e = gp.g[q].off[o]; e < gp.g[q].off[o + 1]; e++)
					Fk += gp.g[q].contrib[3 * (long)gp.g[q].idx[e] + sub];
		}
	}
	const sepk_lfout o = {s.v, s.f, s.a, s.xn, s.cn, s.cr, s.pos, s.qpos};
	sepk_lf_atom_lanes(s, o, t, npad, sub, act, Fk, Fk, Fk, dt, L0, L1, L2, nh, al, pending,
			   qs0, qs1, qs2, acc);
	sepk_block_partial<9>(acc, 1u << 7, partial, lb, rows);
}

// Pair forces, Nose-Hoover and leapfrog in one launch (a step whose only
// calls are one overwriting pair call, sep_nosehoover and sep_leapfrog):
// the pair loop and its rows as in k_lj, then the atom's integrator step
// from its summed force, with the new positions to o.pos / o.qpos.
template <bool MIXED, bool FAST>
__global__ __launch_bounds__(SEPK_BLOCK) __attribute__((amdgpu_waves_per_eu(MIXED ? 7 : FAST ? 6 : 5))) void k_lj_step(
	sepk_state s, const int *__restrict__ nbr, const int *__restrict__ cnt, long n,
	long npad, sepk_lj c, const sepdev_sys *sys, int nh, const double *alpha, double Td,
	double dt_tau2, double dt, double qs0, double qs1, double qs2, sepk_lfout o,
	double *ppart, double *lpart, long rows)
{
	const long lb = sepk_xcd_block();
	const long t = lb * (SEPK_BLOCK / SEPK_SUB) + threadIdx.x / SEPK_SUB;
	const int sub = threadIdx.x % SEPK_SUB;
	const bool act = t < n;
	if (sys->error) {
		// a step after a device error changes nothing: the positions go
		// through to the other buffer unchanged (the whole grid returns)
		if (act && sub == 0) {
			o.pos[t] = s.pos[t];
			o.qpos[t] = s.qpos[t];
		}
		return;
	}
	double fx = 0.0, fy = 0.0, fz = 0.0;
	double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	double px = 0.0, py = 0.0, pz = 0.0;
	sepk_lj_pairs<false, MIXED, FAST>(s, nbr, cnt, npad, c, t, sub, act, 0, fx, fy, fz, acc,
					   px, py, pz);
#pragma unroll
	for (int m = 1; m < SEPK_SUB; m <<= 1) {
		fx += __shfl_xor(fx, m, SEPK_WAVE);
		fy += __shfl_xor(fy, m, SEPK_WAVE);
		fz += __shfl_xor(fz, m, SEPK_WAVE);
	}
	double lf[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	const int pending = sys->xn_pending;
	double al = 0.0;
	if (nh)
		al = sepk_nh_alpha(*alpha, sys->sumv2, (double)n, Td, dt_tau2);
	if (act && sub == 0) {
		// the stress identity's per-atom term, as sepk_pair_epilogue adds it
		const double gx = 2.0 * fx, gy = 2.0 * fy, gz = 2.0 * fz;
		acc[1] = fma(gx, px, acc[1]);
		acc[2] = fma(gx, py, acc[2]);
		acc[3] = fma(gx, pz, acc[3]);
		acc[4] = fma(gy, py, acc[4]);
		acc[5] = fma(gy, pz, acc[5]);
		acc[6] = fma(gz, pz, acc[6]);
	}
	sepk_lf_atom_lanes(s, o, t, npad, sub, act, fx, fy, fz, dt, c.L0, c.L1, c.L2, nh, al,
			   pending, qs0, qs1, qs2, lf);
	sepk_block_partial<7>(acc, 0u, ppart, lb, rows);
	sepk_block_partial<9>(lf, 1u << 7, lpart, lb, rows);
}

// The leapfrog's finishing task (after the step's pair task, if fused).
static int queue_lf_task(sepdev *d, int nh, double *alpha, unsigned nb, double dt, double skin,
			 int ret_slot)
{
	sepk_task f;
	memset(&f, 0, sizeof f);
	f.mode = SEPK_FIN_LF;
	f.region = SEPK_R_LF;
	f.nblk = (int)nb;
	f.nv = 9;
	f.maxmask = 1u << 7;
	f.nh_fused = nh;
	f.alpha = alpha;
	f.Td = d->nh_Td;
	f.dt_tau2 = d->nh_dt_tau2;
	f.npart = (double)d->n;
	f.reset_max = d->max_reset_pending;
	d->max_reset_pending = 0;
	f.skin = skin;
	f.dt = dt;
	sepk_task_ret(d, f, ret_slot);
	if (d->ring) {
		f.ring = d->ring;
		f.ring_cap = d->ring_cap;
		f.ring_n = d->ring_n;
	}
	// one finisher launch for the whole step: any queued build and pair
	// results, then the leapfrog's
	int rc;
	if ((rc = sepk_queue_task(d, f)) || (rc = sepk_flush_tasks(d)))
		return rc;
	d->sumv2_valid = 1;
	return 0;
}

static void launch_step_kernel(sepdev *d, const sepk_lj &c, int mixed, bool fast, int nh,
			       const double *alpha, double dt, const double *L,
			       const sepk_lfout &o, unsigned nb)
{
	auto k = mixed ? (fast ? k_lj_step<true, true> : k_lj_step<true, false>)
		       : (fast ? k_lj_step<false, true> : k_lj_step<false, false>);
	k<<<nb, SEPK_BLOCK, 0, d->st>>>(d->s[0], d->nbr, d->nbr_count, d->n, d->npad, c, d->sys,
					  nh, alpha, d->nh_Td, d->nh_dt_tau2, dt,
					  SEPK_QSCALE / L[0], SEPK_QSCALE / L[1], SEPK_QSCALE / L[2], o,
					  sepk_region(d, SEPK_R_PAIR), sepk_region(d, SEPK_R_LF),
					  d->region_rows);
}

// The recorded pair call, the pending Nose-Hoover call and this leapfrog
// in one launch.
static int sepdev_fused_step(sepdev *d, double dt, double skin, const double *L, int ret_slot)
{
	d->fused_pending = 0;
	int rc;
	if ((rc = sepk_claim_region(d, SEPK_R_PAIR)) || (rc = sepk_claim_region(d, SEPK_R_LF)))
		return rc;
	const int nh = d->nh_pending;
	d->nh_pending = 0;
	double *alpha = nh ? d->scal + d->nh_slot : d->scal;
	const unsigned nb = sepk_blocks_xcd(d->n * SEPK_SUB);
	const sepk_state &s = d->s[0];
	const sepk_lfout o = {s.v, s.f, s.a, s.xn, s.cn, s.cr, d->pos_alt, d->qpos_alt};
	launch_step_kernel(d, d->fused_lj, d->fused_mixed, d->fused_fast != 0, nh, alpha, dt, L, o,
			   nb);
	SEPK_LAUNCH_CHECK();
	d->f_zero_pending = 0;
	// the new positions are current from here on
	sepk_pos *p = d->s[0].pos;
	d->s[0].pos = d->pos_alt;
	d->pos_alt = p;
	int4 *q = d->s[0].qpos;
	d->s[0].qpos = d->qpos_alt;
	d->qpos_alt = q;
	for (int k = 0; k < 3; k++) {
		d->qL[k] = L[k];
		d->fused_L[k] = L[k];
	}
	d->qpos_stale = 0;
	d->last_fused = 1;
	d->fused_nh = nh;
	d->fused_alpha = alpha;
	d->fused_dt = dt;
	sepk_task f;
	memset(&f, 0, sizeof f);
	f.mode = SEPK_FIN_PAIR;
	f.region = SEPK_R_PAIR;
	f.nblk = (int)nb;
	f.nv = 7;
	f.epot_assign = d->fused_epot_assign;
	sepk_task_ret(d, f, d->fused_ret);
	if ((rc = sepk_queue_task(d, f)))
		return rc;
	return queue_lf_task(d, nh, alpha, nb, dt, skin, ret_slot);
}

extern "C" int sepdev_leapfrog(sepdev *d, double dt, double skin, const double *L,
			       int ret_slot)
{
	if (ret_slot < 0 || ret_slot >= d->nret)
		return sepk_fail(SEP_GPU_EARG, "ret slot %d", ret_slot);
	d->fork_ready = 0; // the atoms move
	if (d->fused_pending)
		return sepdev_fused_step(d, dt, skin, L, ret_slot);
	// a recorded pair call goes first (it applies any owed gathers before
	// itself); the gathers owed after it are folded into the kernel
	int rc = sepk_flush_pair(d);
	if (rc)
		return rc;
	sepk_gpend gp;
	gp.n = d->ngpend;
	gp.over = d->gpend_over;
	for (int q = 0; q < gp.n; q++)
		gp.g[q] = d->gpend[q];
	d->ngpend = 0;
	d->gpend_over = 0;
	if ((rc = sepk_materialize_forces(d)))
		return rc;
	if ((rc = sepk_claim_region(d, SEPK_R_LF)))
		return rc;
	const int nh = d->nh_pending;
	d->nh_pending = 0;
	const unsigned nb = sepk_blocks_xcd(d->n * SEPK_SUB);
	double *alpha = nh ? d->scal + d->nh_slot : d->scal;
	k_leapfrog<<<nb, SEPK_BLOCK, 0, d->st>>>(d->s[0], d->n, d->npad, dt, L[0], L[1], L[2],
						  d->sys, nh, alpha, d->nh_Td, d->nh_dt_tau2,
						  sepk_region(d, SEPK_R_LF), d->region_rows,
						  SEPK_QSCALE / L[0], SEPK_QSCALE / L[1],
						  SEPK_QSCALE / L[2], gp);
	SEPK_LAUNCH_CHECK();
	for (int k = 0; k < 3; k++)
		d->qL[k] = L[k];
	d->qpos_stale = 0;
	return queue_lf_task(d, nh, alpha, nb, dt, skin, ret_slot);
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
		const int oj = orig[nbr[sepk_entry(k, t, npad)]];
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
	int rc0 = sepk_flush_tasks(d);
	if (rc0)
		return rc0;
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

extern "C" int sepdev_time_pair(sepdev *d, int reps, double *ms_per_launch)
{
	if ((!d->last_lj && !d->last_fused) || !d->list_built)
		return sepk_fail(SEP_GPU_ESTATE, "no pair-force launch to time");
	if (reps < 1 || d->capturing)
		return sepk_fail(SEP_GPU_EARG, "time_pair: reps %d%s", reps,
				 d->capturing ? " during capture" : "");
	// owed bonded gathers go to f before the timed launches replace it
	int rc = sepk_flush_nh(d);
	if (rc || (rc = sepk_flush_gathers(d)) || (rc = sepk_flush_tasks(d)))
		return rc;
	sepk_lj c;
	memcpy(&c, d->last_fused ? (const void *)&d->fused_lj : d->last_lj, sizeof c);
	const sepk_lj_kernel k = (sepk_lj_kernel)d->last_lj_kernel;
	const unsigned nb = sepk_blocks_xcd(d->n * SEPK_SUB);
	double *part = sepk_region(d, SEPK_R_PAIR);
	// a fused step kernel is timed with its integrator outputs sent to
	// scratch, so that back-to-back launches see the same state
	sepk_lfout dry;
	memset(&dry, 0, sizeof dry);
	if (d->last_fused) {
		const size_t np = (size_t)d->npad;
		if (!d->tscratch)
			SEPK_CHECK(hipMalloc(&d->tscratch, 21 * np * sizeof(double)));
		double *b = d->tscratch;
		dry.v = b;
		dry.f = b + 3 * np;
		dry.a = b + 6 * np;
		dry.xn = b + 9 * np;
		dry.cn = (int *)(b + 12 * np);
		dry.cr = (int *)(b + 12 * np) + 3 * np;
		dry.pos = (sepk_pos *)(b + 15 * np);
		dry.qpos = (int4 *)(b + 19 * np);
	}
	hipEvent_t e0, e1;
	SEPK_CHECK(hipEventCreate(&e0));
	SEPK_CHECK(hipEventCreate(&e1));
	hipError_t err = hipEventRecord(e0, d->st);
	for (int r = 0; r < reps && err == hipSuccess; r++) {
		if (d->last_fused)
			launch_step_kernel(d, c, d->fused_mixed, d->fused_fast != 0, d->fused_nh,
					   d->fused_alpha, d->fused_dt, d->fused_L, dry, nb);
		else
			k<<<nb, SEPK_BLOCK, 0, d->st>>>(d->s[0], d->nbr, d->nbr_count, d->n,
							  d->npad, c, part, d->region_rows);
		err = hipGetLastError();
	}
	if (err == hipSuccess)
		err = hipEventRecord(e1, d->st);
	if (err == hipSuccess)
		err = hipEventSynchronize(e1);
	float ms = 0.0f;
	if (err == hipSuccess)
		err = hipEventElapsedTime(&ms, e0, e1);
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (err != hipSuccess)
		return sepk_fail(SEP_GPU_EDEVICE, "time_pair: %s", hipGetErrorString(err));
	d->f_zero_pending = 0;
	*ms_per_launch = ms / reps;
	return 0;
}
