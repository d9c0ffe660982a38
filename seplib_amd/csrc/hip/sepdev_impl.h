// sepdev_impl.h -- device-side state and helpers shared by the HIP sources
// of libsep.so.  Internal; not installed.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../sepdev.h"
#include "sepgpu.h"

#define SEPK_BLOCK 256
#define SEPK_WAVE 64

// 32-byte aligned position record: x, y, z and the point charge.  One
// record per atom so that a neighbour gather is a single 32-byte access.
struct __attribute__((aligned(32))) sepk_pos {
	double x, y, z, q;
};

// One copy of the per-atom state, in device (cell-sorted) order.  Vector
// fields are stored as three blocks of npad doubles (x block, y block,
// z block) so that per-atom streaming is coalesced.
struct sepk_state {
	sepk_pos *pos;
	double *v, *f, *a, *xn;
	int *cn, *cr;
	double *m;
	signed char *type;
	int *mol;
	int *orig;
	int *partners; // 40 blocks of npad, allocated with the topology
};

struct sepk_timer {
	hipEvent_t beg[64], end[64];
	int used;
	double ms;
	long count;
};

struct sepdev {
	long n, npad;
	hipStream_t st;
	sepk_state s[2]; // s[0] current (fixed addresses), s[1] rebuild spare

	int *iperm; // original index -> device index

	// cell binning
	long ncell_cap;
	int *cell_count, *cell_start, *cell_fill;
	int *cellid_old, *cellid_new, *order_tmp, *neworder;

	// full neighbour list, entry k of atom t at nbr[k*npad + t]
	int cap;
	int *nbr, *nbr_count;
	int list_built;
	sepdev_grid grid; // grid of the last build

	// reductions: per-block partial sums and last-block counters
	double *partials;
	long partials_words;
	unsigned *counters;

	// return and scalar slots
	double *ret;
	int nret, ret_cap;
	double *scal;
	int nscal, scal_cap;
	sepdev_sys *sys;

	// staging
	void *hstage, *dstage;
	size_t stage_bytes;

	// forces are known to be zero on the device (lazy sep_reset_force)
	int f_zero_pending;

	// topology (original atom numbering)
	int nb, na, nd;
	unsigned *blist, *alist, *dlist;
	double *bcontrib, *acontrib, *dcontrib;    // per-term force slots
	int *bcsr_off, *bcsr_idx, *acsr_off, *acsr_idx, *dcsr_off, *dcsr_idx;
	double *blen, *aval, *dval;

	// timing
	int timing;
	sepk_timer t_pair;

	// graph
	hipGraph_t graph;
	hipGraphExec_t exec;
	int capturing;
};

// error recording (sepdev.hip)
int sepk_fail(int code, const char *fmt, ...);
#define SEPK_CHECK(call)                                                      \
	do {                                                                  \
		hipError_t e_ = (call);                                       \
		if (e_ != hipSuccess)                                         \
			return sepk_fail(SEP_GPU_EDEVICE, "%s: %s (%s:%d)", #call, \
					 hipGetErrorString(e_), __FILE__, __LINE__); \
	} while (0)
#define SEPK_LAUNCH_CHECK()                                                   \
	do {                                                                  \
		hipError_t e_ = hipGetLastError();                            \
		if (e_ != hipSuccess)                                         \
			return sepk_fail(SEP_GPU_EDEVICE, "launch: %s (%s:%d)",   \
					 hipGetErrorString(e_), __FILE__, __LINE__); \
	} while (0)

static inline unsigned sepk_blocks(long n)
{
	return (unsigned)((n + SEPK_BLOCK - 1) / SEPK_BLOCK);
}

// Counter indices for the last-block reductions.
enum { SEPK_CTR_PAIR = 0, SEPK_CTR_LF = 1, SEPK_CTR_NH = 2, SEPK_CTR_BOND = 3,
       SEPK_CTR_COUL = 4, SEPK_CTR_MISC = 5, SEPK_NCTR = 8 };

// ---- device helpers ----------------------------------------------------

// Single-image minimum image with strict comparisons against +-L/2, the
// rule of the reference's sep_Wrap macro (include/sepmisc.h:81-85).
__device__ __forceinline__ double sepk_wrap(double d, double L)
{
	if (d > 0.5 * L)
		d -= L;
	else if (d < -0.5 * L)
		d += L;
	return d;
}

// Block-wide sum of NV doubles per thread into partial[blockIdx.x*NV + i],
// in a fixed order (wave butterfly, then waves in index order).
template <int NV>
__device__ __forceinline__ void sepk_block_sum(double (&v)[NV], double *partial)
{
	__shared__ double sh[SEPK_BLOCK / SEPK_WAVE][NV];
	const int lane = threadIdx.x & (SEPK_WAVE - 1), wid = threadIdx.x / SEPK_WAVE;
#pragma unroll
	for (int i = 0; i < NV; i++) {
		double x = v[i];
#pragma unroll
		for (int o = SEPK_WAVE / 2; o > 0; o >>= 1)
			x += __shfl_xor(x, o, SEPK_WAVE);
		v[i] = x;
	}
	if (lane == 0)
#pragma unroll
		for (int i = 0; i < NV; i++)
			sh[wid][i] = v[i];
	__syncthreads();
	if (threadIdx.x < NV) {
		double s = 0.0;
		for (int w = 0; w < SEPK_BLOCK / SEPK_WAVE; w++)
			s += sh[w][threadIdx.x];
		partial[(size_t)blockIdx.x * NV + threadIdx.x] = s;
	}
}

__device__ __forceinline__ double sepk_block_max(double v, double *partial_slot)
{
	__shared__ double shm[SEPK_BLOCK / SEPK_WAVE];
	const int lane = threadIdx.x & (SEPK_WAVE - 1), wid = threadIdx.x / SEPK_WAVE;
#pragma unroll
	for (int o = SEPK_WAVE / 2; o > 0; o >>= 1)
		v = fmax(v, __shfl_xor(v, o, SEPK_WAVE));
	if (lane == 0)
		shm[wid] = v;
	__syncthreads();
	if (threadIdx.x == 0) {
		double m = shm[0];
		for (int w = 1; w < SEPK_BLOCK / SEPK_WAVE; w++)
			m = fmax(m, shm[w]);
		*partial_slot = m;
	}
	return v;
}

// Returns true in every thread of the block that finished last.  Each block
// has written its partials before calling; the fences make them visible to
// the last block (device scope, across XCDs).
__device__ __forceinline__ bool sepk_last_block(unsigned *counter)
{
	__shared__ int last;
	__threadfence();
	__syncthreads();
	if (threadIdx.x == 0) {
		unsigned t = atomicAdd(counter, 1u);
		last = (t == gridDim.x - 1);
	}
	__syncthreads();
	if (last)
		__threadfence();
	return last;
}

__device__ __forceinline__ double sepk_ld(const double *p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of column i over nblk partial rows (NV words each), in block order,
// done by the last block with one thread per column.
__device__ __forceinline__ double sepk_sum_partials(const double *partial,
						    int nblk, int nv, int i)
{
	double s = 0.0;
	for (int b = 0; b < nblk; b++)
		s += sepk_ld(partial + (size_t)b * nv + i);
	return s;
}

// Writes the 6 unique stress components (xx xy xz yy yz zz) into a
// row-major 3x3 block.
__device__ __forceinline__ void sepk_add_sym9(double *dst9, const double *w6,
					      double scale)
{
	dst9[0] += scale * w6[0];
	dst9[1] += scale * w6[1];
	dst9[2] += scale * w6[2];
	dst9[3] += scale * w6[1];
	dst9[4] += scale * w6[3];
	dst9[5] += scale * w6[4];
	dst9[6] += scale * w6[2];
	dst9[7] += scale * w6[4];
	dst9[8] += scale * w6[5];
}

// internal launch helpers implemented in sepdev.hip
int sepk_materialize_forces(sepdev *d);
