// sepdev_adjoint.hip -- multi-type Lennard-Jones from a type-pair table in
// one list pass, and the adjoint of a force-matching loss.
//
// Pair force (the definition of the reference's sep_force_lj,
// source/sepprfrc.c:785-840, per type pair ab with its own eps, sigma, rc,
// aw):  f_ij = ft(s) r_ij,  s = |r_ij|^2,  q = sigma^2 / s,
//     ft = 48 eps (q^7 - aw/2 q^4),  u = 4 eps (q^6 - aw q^3) - shift_ab.
// For a loss L(F) with g_i = dL/dF_i, the adjoint over the full list is
//     dL/d eps_ab   = sum_{i, j in list(i), types ab} (ft / eps)        (g_i . r_ij)
//     dL/d sigma_ab = sum_{i, j in list(i), types ab} (48 eps / sigma)
//                       (14 q^7 - 4 aw q^4)                            (g_i . r_ij)
//     dL/dx_i = sum_{j in list(i)} J(r_ij) (g_i - g_j),
//               J(r) = ft I + 2 (dft/ds) r r^T,
//               dft/ds = -(48 eps / s) (7 q^7 - 2 aw q^4),
// each unordered pair appearing once from each end.  Cutoffs are held
// fixed (not differentiated).  One writer per atom, no atomics; the
// parameter gradient is reduced per block and then by one workgroup in a
// fixed order.
#include <math.h>
#include <string.h>

#include "sepdev_impl.h"

#define SEPK_NPAIR (SEPDEV_MAXT * (SEPDEV_MAXT + 1) / 2)

// Per unordered type pair, in LDS for the kernels.
struct sepk_tpair {
	double A, nB, C, nD, sig2, cf2, ushift, eps, sigma, aw;
};

struct sepk_ljtab {
	sepk_tpair p[SEPK_NPAIR];
	signed char tmap[128]; // type label -> type index, -1 when not in the table
	int nt;
	double L0, L1, L2, iL0, iL1, iL2;
};

__device__ __forceinline__ int sepk_pair_index(int a, int b, int nt)
{
	const int lo = min(a, b), hi = max(a, b);
	return lo * nt - lo * (lo - 1) / 2 + (hi - lo);
}

// the reference's single-image rule (exact for any box these lists serve)
__device__ __forceinline__ double sepk_img(double d, double L, double iL)
{
	(void)iL;
	return sepk_wrap(d, L);
}

template <bool OVERWRITE>
__global__ __launch_bounds__(SEPK_BLOCK) void k_lj_table(
	sepk_state s, const int *__restrict__ nbr, const int *__restrict__ cnt, long n,
	long npad, sepk_ljtab c, double *partial, long rows)
{
	__shared__ sepk_tpair P[SEPK_NPAIR];
	for (int i = threadIdx.x; i < SEPK_NPAIR; i += blockDim.x)
		P[i] = c.p[i];
	__syncthreads();
	const long lb = sepk_xcd_block();
	const long t = lb * (SEPK_BLOCK / SEPK_SUB) + threadIdx.x / SEPK_SUB;
	const int sub = threadIdx.x % SEPK_SUB;
	double fx = 0.0, fy = 0.0, fz = 0.0;
	double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	int ti = -1;
	if (t < n)
		ti = c.tmap[s.type[t] & 127];
	if (ti >= 0) {
		const sepk_pos pi = s.pos[t];
		const int nbat = (cnt[t] + SEPK_PAD - 1) / SEPK_PAD;
		const size_t stride = (size_t)npad * SEPK_SUB;
		const int4 *lp = sepk_lane_list(nbr, t, sub);
		for (int b = 0; b < nbat; b++, lp += stride) {
			const int4 jb = *lp;
			const int j[SEPK_BATCH] = {jb.x, jb.y, jb.z, jb.w};
			sepk_pos pj[SEPK_BATCH];
			int tj[SEPK_BATCH];
#pragma unroll
			for (int u = 0; u < SEPK_BATCH; u++) {
				pj[u] = s.pos[(unsigned)j[u]];
				tj[u] = c.tmap[s.type[(unsigned)j[u]] & 127];
			}
#pragma unroll
			for (int u = 0; u < SEPK_BATCH; u++) {
				if (tj[u] < 0)
					continue;
				const sepk_tpair &q = P[sepk_pair_index(ti, tj[u], c.nt)];
				const double dx = sepk_img(pi.x - pj[u].x, c.L0, c.iL0);
				const double dy = sepk_img(pi.y - pj[u].y, c.L1, c.iL1);
				const double dz = sepk_img(pi.z - pj[u].z, c.L2, c.iL2);
				const double xx = dx * dx, yy = dy * dy, zz = dz * dz;
				const double r2 = xx + yy + zz;
				if (!(r2 < q.cf2))
					continue;
				const double qq = q.sig2 / r2;
				const double q3 = qq * qq * qq;
				const double ft = fma(q.A, q3, q.nB) * q3 * qq;
				const double e = fma(fma(q.C, q3, q.nD), q3, q.ushift);
				fx = fma(ft, dx, fx);
				fy = fma(ft, dy, fy);
				fz = fma(ft, dz, fz);
				acc[0] += e;
				acc[1] = fma(ft, xx, acc[1]);
				acc[2] = fma(ft, dx * dy, acc[2]);
				acc[3] = fma(ft, dx * dz, acc[3]);
				acc[4] = fma(ft, yy, acc[4]);
				acc[5] = fma(ft, dy * dz, acc[5]);
				acc[6] = fma(ft, zz, acc[6]);
			}
		}
	}
#pragma unroll
	for (int o = 1; o < SEPK_SUB; o <<= 1) {
		fx += __shfl_xor(fx, o, SEPK_WAVE);
		fy += __shfl_xor(fy, o, SEPK_WAVE);
		fz += __shfl_xor(fz, o, SEPK_WAVE);
	}
	// f stays as it is while the device error word is set
	if (t < n && sub == 0 && !*s.err) {
		if (OVERWRITE) {
			s.f[t] = fx;
			s.f[npad + t] = fy;
			s.f[2 * npad + t] = fz;
		} else if (ti >= 0) {
			s.f[t] += fx;
			s.f[npad + t] += fy;
			s.f[2 * npad + t] += fz;
		}
	}
	sepk_block_partial<7>(acc, 0u, partial, lb, rows);
}

// The adjoint pass: g (3 x npad, device order) in, dL/dx (3 x npad) out, and
// per block a row of the 2 * SEPK_NPAIR parameter-gradient sums.
__global__ __launch_bounds__(SEPK_BLOCK) void k_lj_adjoint(
	sepk_state s, const int *__restrict__ nbr, const int *__restrict__ cnt, long n,
	long npad, sepk_ljtab c, const double *__restrict__ g, double *dldx, double *partial,
	long rows)
{
	__shared__ sepk_tpair P[SEPK_NPAIR];
	for (int i = threadIdx.x; i < SEPK_NPAIR; i += blockDim.x)
		P[i] = c.p[i];
	__syncthreads();
	const long lb = sepk_xcd_block();
	const long t = lb * (SEPK_BLOCK / SEPK_SUB) + threadIdx.x / SEPK_SUB;
	const int sub = threadIdx.x % SEPK_SUB;
	double hx = 0.0, hy = 0.0, hz = 0.0;
	double gp[2 * SEPK_NPAIR];
#pragma unroll
	for (int k = 0; k < 2 * SEPK_NPAIR; k++)
		gp[k] = 0.0;
	int ti = -1;
	if (t < n)
		ti = c.tmap[s.type[t] & 127];
	if (ti >= 0) {
		const sepk_pos pi = s.pos[t];
		const double gix = g[t], giy = g[npad + t], giz = g[2 * npad + t];
		const int nbat = (cnt[t] + SEPK_PAD - 1) / SEPK_PAD;
		const size_t stride = (size_t)npad * SEPK_SUB;
		const int4 *lp = sepk_lane_list(nbr, t, sub);
		for (int b = 0; b < nbat; b++, lp += stride) {
			const int4 jb = *lp;
			const int j[SEPK_BATCH] = {jb.x, jb.y, jb.z, jb.w};
#pragma unroll
			for (int u = 0; u < SEPK_BATCH; u++) {
				const unsigned jj = (unsigned)j[u];
				const int tj = c.tmap[s.type[jj] & 127];
				if (tj < 0)
					continue;
				const int pidx = sepk_pair_index(ti, tj, c.nt);
				const sepk_tpair &q = P[pidx];
				const sepk_pos pj = s.pos[jj];
				const double dx = sepk_img(pi.x - pj.x, c.L0, c.iL0);
				const double dy = sepk_img(pi.y - pj.y, c.L1, c.iL1);
				const double dz = sepk_img(pi.z - pj.z, c.L2, c.iL2);
				const double r2 = dx * dx + dy * dy + dz * dz;
				if (!(r2 < q.cf2))
					continue;
				const double qq = q.sig2 / r2;
				const double q3 = qq * qq * qq;
				const double q4 = q3 * qq, q7 = q4 * q3;
				const double e48 = 48.0 * q.eps;
				const double ft = e48 * (q7 - 0.5 * q.aw * q4);
				const double fts = -(e48 / r2) * (7.0 * q7 - 2.0 * q.aw * q4);
				// parameter gradient: (dft/dtheta) (g_i . r)
				const double gr = gix * dx + giy * dy + giz * dz;
				const double deps = (ft / q.eps) * gr;
				const double dsig = (e48 / q.sigma) * (14.0 * q7 - 4.0 * q.aw * q4) * gr;
#pragma unroll
				for (int k = 0; k < SEPK_NPAIR; k++)
					if (k == pidx) {
						gp[2 * k] += deps;
						gp[2 * k + 1] += dsig;
					}
				// position gradient: J(r) (g_i - g_j)
				const double ux = gix - g[jj], uy = giy - g[npad + jj],
					     uz = giz - g[2 * npad + jj];
				const double ru = 2.0 * fts * (dx * ux + dy * uy + dz * uz);
				hx += ft * ux + ru * dx;
				hy += ft * uy + ru * dy;
				hz += ft * uz + ru * dz;
			}
		}
	}
#pragma unroll
	for (int o = 1; o < SEPK_SUB; o <<= 1) {
		hx += __shfl_xor(hx, o, SEPK_WAVE);
		hy += __shfl_xor(hy, o, SEPK_WAVE);
		hz += __shfl_xor(hz, o, SEPK_WAVE);
	}
	if (t < n && sub == 0) {
		dldx[t] = hx;
		dldx[npad + t] = hy;
		dldx[2 * npad + t] = hz;
	}
	sepk_block_partial<2 * SEPK_NPAIR>(gp, 0u, partial, lb, rows);
}

// One workgroup: column sums of the adjoint's partial rows, fixed order.
__global__ __launch_bounds__(256) void k_colsum(const double *partial, long rows, int ncol,
						int nblk, double *out)
{
	__shared__ double red[4];
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	for (int col = 0; col < ncol; col++) {
		double a = 0.0;
		for (int r = threadIdx.x; r < nblk; r += 256)
			a += partial[(size_t)col * rows + r];
		for (int o = 32; o > 0; o >>= 1)
			a += __shfl_xor(a, o, 64);
		if (lane == 0)
			red[wid] = a;
		__syncthreads();
		if (threadIdx.x == 0)
			out[col] = ((red[0] + red[1]) + red[2]) + red[3];
		__syncthreads();
	}
}

static int fill_table(const sepdev_ljtab_args *a, sepk_ljtab &c)
{
	memset(&c, 0, sizeof c);
	if (a->ntypes < 1 || a->ntypes > SEPDEV_MAXT)
		return sepk_fail(SEP_GPU_EARG, "LJ table: %d types (1..%d)", a->ntypes, SEPDEV_MAXT);
	c.nt = a->ntypes;
	for (int k = 0; k < 128; k++)
		c.tmap[k] = -1;
	for (int i = 0; i < a->ntypes; i++) {
		const int ch = (unsigned char)a->types[i];
		if (ch == 0 || ch >= 128 || c.tmap[ch] >= 0)
			return sepk_fail(SEP_GPU_EARG, "LJ table: bad or repeated type label");
		c.tmap[ch] = (signed char)i;
	}
	for (int i = 0; i < a->ntypes; i++)
		for (int k = i; k < a->ntypes; k++) {
			// params[(i * ntypes + k) * 4 + {0,1,2,3}] = rc, eps, sigma, aw
			const double *p = a->params + (size_t)(i * a->ntypes + k) * 4;
			const double rc = p[0], eps = p[1], sig = p[2], aw = p[3];
			const int idx = i * c.nt - i * (i - 1) / 2 + (k - i);
			sepk_tpair &q = c.p[idx];
			q.eps = eps;
			q.sigma = sig;
			q.aw = aw;
			q.A = 48.0 * eps;
			q.nB = -(48.0 * eps * 0.5 * aw);
			q.C = 4.0 * eps;
			q.nD = -(4.0 * eps * aw);
			q.sig2 = sig * sig;
			q.cf2 = rc * rc;
			// source/sepprfrc.c:786: shift = 4 eps ((s/rc)^12 - aw (s/rc)^6)
			q.ushift = -4.0 * eps * (pow(sig / rc, 12.0) - aw * pow(sig / rc, 6.0));
		}
	c.L0 = a->L[0];
	c.L1 = a->L[1];
	c.L2 = a->L[2];
	c.iL0 = 1.0 / a->L[0];
	c.iL1 = 1.0 / a->L[1];
	c.iL2 = 1.0 / a->L[2];
	return 0;
}

extern "C" int sepdev_lj_table(sepdev *d, const sepdev_ljtab_args *a, int ret_slot)
{
	if (!d->list_built)
		return sepk_fail(SEP_GPU_ESTATE, "pair forces before a neighbour list was built");
	if (ret_slot < 0 || ret_slot >= d->nret)
		return sepk_fail(SEP_GPU_EARG, "ret slot %d", ret_slot);
	sepk_ljtab c;
	int rc = fill_table(a, c);
	int region;
	if (rc || (rc = sepk_flush_nh(d)) || (rc = sepk_term_region(d, &region)) ||
	    (rc = sepk_fork_point(d)) || (rc = sepk_flush_gathers(d)))
		return rc;
	const bool over = d->f_zero_pending != 0;
	const unsigned nb = sepk_blocks_xcd(d->n * SEPK_SUB);
	auto k = over ? k_lj_table<true> : k_lj_table<false>;
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
	f.epot_assign = 0; // sep_force_lj adds to epot
	sepk_task_ret(d, f, ret_slot);
	return sepk_queue_task(d, f);
}

// dldtheta[(i * nt + k) * 2 + {0, 1}] = dL/d eps, dL/d sigma of pair ik from
// the pair sums, written for both orders of a pair
static void unpack_theta(const sepdev_ljtab_args *a, const double *full, double *dldtheta)
{
	for (int i = 0; i < a->ntypes; i++)
		for (int k = 0; k < a->ntypes; k++) {
			const int lo = i < k ? i : k, hi = i < k ? k : i;
			const int idx = lo * a->ntypes - lo * (lo - 1) / 2 + (hi - lo);
			dldtheta[(i * a->ntypes + k) * 2] = full[2 * idx];
			dldtheta[(i * a->ntypes + k) * 2 + 1] = full[2 * idx + 1];
		}
}

extern "C" int sepdev_lj_adjoint(sepdev *d, const sepdev_ljtab_args *a, const double *g3n,
				 double *dldx3n, double *dldtheta)
{
	if (!d->list_built)
		return sepk_fail(SEP_GPU_ESTATE, "adjoint before a neighbour list was built");
	sepk_ljtab c;
	int rc = fill_table(a, c);
	if (rc || (rc = sepk_flush_nh(d)) || (rc = sepk_flush_tasks(d)))
		return rc;
	const size_t bytes = (size_t)3 * d->npad * sizeof(double);
	if (!d->adj_g) {
		if (d->capturing)
			return sepk_fail(SEP_GPU_ESTATE, "adjoint buffers allocated during capture");
		SEPK_CHECK(hipMalloc(&d->adj_g, bytes));
		SEPK_CHECK(hipMalloc(&d->adj_x, bytes));
		SEPK_CHECK(hipMalloc(&d->adj_out, 2 * SEPK_NPAIR * sizeof(double)));
		SEPK_CHECK(hipMalloc(&d->adj_part,
				     (size_t)d->region_rows * 2 * SEPK_NPAIR * sizeof(double)));
	}
	// g arrives in original order
	if ((rc = sepk_put_vec_to(d, d->adj_g, g3n)))
		return rc;
	const unsigned nb = sepk_blocks_xcd(d->n * SEPK_SUB);
	if ((long)nb > d->region_rows)
		return sepk_fail(SEP_GPU_EARG, "adjoint: %u blocks overflow the partial rows", nb);
	double *part = d->adj_part;
	k_lj_adjoint<<<nb, SEPK_BLOCK, 0, d->st>>>(d->s[0], d->nbr, d->nbr_count, d->n, d->npad,
						   c, d->adj_g, d->adj_x, part, d->region_rows);
	SEPK_LAUNCH_CHECK();
	k_colsum<<<1, 256, 0, d->st>>>(part, d->region_rows, 2 * SEPK_NPAIR, (int)nb, d->adj_out);
	SEPK_LAUNCH_CHECK();
	double full[2 * SEPK_NPAIR];
	SEPK_CHECK(hipMemcpyAsync(full, d->adj_out, sizeof full, hipMemcpyDeviceToHost, d->st));
	if ((rc = sepk_get_vec_from(d, dldx3n, d->adj_x)))
		return rc;
	SEPK_CHECK(hipStreamSynchronize(d->st));
	unpack_theta(a, full, dldtheta);
	return 0;
}

// ---- the adjoint without allocation ----------------------------------------
//
// The device buffers of sepdev_lj_adjoint live in this array instead of
// memory allocated on first use.  Each GPU the library runs on holds its own
// copy (SEPDEV_SCRATCH_BYTES) until the process ends.  Host code never uses
// the array's name, only the device address hipGetSymbolAddress returns for
// it on each device (scratch_dev).  Every call on a GPU uses the same array:
// nothing in it is read before the call has written it, and calls that
// overlap on one GPU race.
static __device__ __align__(256) double scratch[SEPDEV_SCRATCH_BYTES / sizeof(double)];

#define SEPK_MAX_DEVICES 64
static double *scratch_dev[SEPK_MAX_DEVICES];

// The device address of scratch on the device of d, looked up once per
// device with that device current.
static int scratch_for(sepdev *d, double **out)
{
	if (d->device < 0 || d->device >= SEPK_MAX_DEVICES)
		return sepk_fail(SEP_GPU_EARG, "device %d beyond the scratch table", d->device);
	if (!scratch_dev[d->device]) {
		int cur = 0;
		SEPK_CHECK(hipGetDevice(&cur));
		if (cur != d->device)
			SEPK_CHECK(hipSetDevice(d->device));
		void *p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (cur != d->device)
			(void)hipSetDevice(cur);
		if (e != hipSuccess || !p)
			return sepk_fail(SEP_GPU_EDEVICE, "hipGetSymbolAddress(scratch): %s",
					 hipGetErrorString(e));
		scratch_dev[d->device] = (double *)p;
	}
	*out = scratch_dev[d->device];
	return 0;
}

extern "C" int sepdev_lj_adjoint_noalloc(sepdev *d, const sepdev_ljtab_args *a,
					 const double *g3n, double *dldx3n, double *dldtheta)
{
	if (!d->list_built)
		return sepk_fail(SEP_GPU_ESTATE, "adjoint before a neighbour list was built");
	if (d->capturing)
		return sepk_fail(SEP_GPU_ESTATE, "adjoint during a graph capture");
	sepk_ljtab c;
	int rc = fill_table(a, c);
	if (rc)
		return rc;
	// layout (doubles): g (3 npad), dL/dx (3 npad), the pair sums (2 NPAIR,
	// padded to 32), the partial rows (one per block, 2 NPAIR columns)
	const unsigned nb = sepk_blocks_xcd(d->n * SEPK_SUB);
	const size_t vec = (size_t)3 * d->npad, sums = 32;
	const size_t need = 2 * vec + sums + (size_t)nb * 2 * SEPK_NPAIR;
	const size_t have = SEPDEV_SCRATCH_BYTES / sizeof(double);
	if (need > have)
		return sepk_fail(SEP_GPU_EARG,
				 "the no-allocation adjoint needs %zu bytes for %ld atoms; its scratch "
				 "holds %ld bytes (about %ld atoms)",
				 need * sizeof(double), d->n, (long)SEPDEV_SCRATCH_BYTES,
				 (long)(have / (6.0 + 2.0 * SEPK_NPAIR * SEPK_SUB / SEPK_BLOCK)) - 1024);
	double *base = nullptr;
	if ((rc = scratch_for(d, &base)) || (rc = sepk_flush_nh(d)) || (rc = sepk_flush_tasks(d)))
		return rc;
	double *g = base, *dldx = base + vec, *out = base + 2 * vec, *part = out + sums;
	if ((rc = sepk_put_vec_to(d, g, g3n)))
		return rc;
	k_lj_adjoint<<<nb, SEPK_BLOCK, 0, d->st>>>(d->s[0], d->nbr, d->nbr_count, d->n, d->npad,
						   c, g, dldx, part, (long)nb);
	SEPK_LAUNCH_CHECK();
	k_colsum<<<1, 256, 0, d->st>>>(part, (long)nb, 2 * SEPK_NPAIR, (int)nb, out);
	SEPK_LAUNCH_CHECK();
	// the sums go out through the pinned staging buffer after dL/dx
	if ((rc = sepk_get_vec_from(d, dldx3n, dldx)))
		return rc;
	double full[2 * SEPK_NPAIR];
	SEPK_CHECK(hipMemcpyAsync(d->hstage, out, sizeof full, hipMemcpyDeviceToHost, d->st));
	SEPK_CHECK(hipStreamSynchronize(d->st));
	memcpy(full, d->hstage, sizeof full);
	unpack_theta(a, full, dldtheta);
	return 0;
}
