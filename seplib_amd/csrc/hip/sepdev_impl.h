// sepdev_impl.h -- device-side state and helpers shared by the HIP sources
// of libsep.so.  Internal; not installed.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../sepdev.h"
#include "sepgpu.h"

#define SEPK_BLOCK 256
#define SEPK_WAVE 64
// Lanes per atom in the list kernels.  Entry k of atom t's list is read by
// lane k%SEPK_SUB of the atom (slot: sepk_entry): a wave covers
// 64/SEPK_SUB consecutive atoms, each of its list reads is 1 KB of
// contiguous memory, and since every list is in cell order the lanes of one
// atom gather neighbours from the same cache lines.
#define SEPK_SUB SEPDEV_LANES
// entries a lane of a pair kernel takes at a time (one int4 of indices);
// every atom's list is padded with the phantom slot n to a multiple of
// SEPK_PAD entries, so a lane's entries come in full batches
#define SEPK_BATCH 4
#define SEPK_PAD (SEPK_SUB * SEPK_BATCH)
// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2); list kernels remap so that each XCD gets one contiguous range of
// cell-sorted atoms, whose neighbours are mostly in the same range.
#define SEPK_NXCD 8

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
	// the mixed-form pair kernels' copy of pos: each coordinate mod L as an
	// unsigned 32-bit fraction of the box (sepk_quant), w the FP32 charge
	// bits (SEPK_QPHANTOM on phantoms)
	int4 *qpos;
	double *v, *f, *a, *xn;
	int *cn, *cr;
	double *m;
	signed char *type;
	int *mol;
	int *orig;
	int *partners; // 40 blocks of npad, allocated with the topology
	// the device error word (sepdev_sys::error): a step's kernels leave the
	// state unchanged while it is set
	const int *err;
};

// Constants of one LJ pair-force launch (sep_force_pairs / sep_force_lj).
struct sepk_lj {
	double cf2, A, nB, C, nD, sig2, ushift;
	double L0, L1, L2, iL0, iL1, iL2;
	double qd0, qd1, qd2; // L / 2^32: one step of the quantized positions
	// the quantized r^2 is within band of the FP64 one (mixed kernels);
	// cf2lo = cf2 - band, cf2hi = cf2 + band
	double band, cf2lo, cf2hi;
	float Af, nBf, Cf, nDf, sig2f, ushiftf;
	signed char t0, t1;
};

// What the finisher does with the combined rows.  Partial rows live in
// separate regions so that a build, the force terms and a leapfrog of one
// step can all be finished by a single launch after the leapfrog.  Force
// terms take the first of the SEPK_NTERM term regions that no queued task
// holds (SEPK_R_PAIR is the first; the fused step always uses it).
enum { SEPK_FIN_PAIR = 0, SEPK_FIN_BONDED = 1, SEPK_FIN_NH = 2, SEPK_FIN_LF = 3,
       SEPK_FIN_BUILD = 4 };
#define SEPK_NTERM 4
enum { SEPK_R_BUILD = 0, SEPK_R_LF = 1, SEPK_R_MISC = 2, SEPK_R_PAIR = 3,
       SEPK_NREGION = SEPK_R_PAIR + SEPK_NTERM };
#define SEPK_MAXNV 10
#define SEPK_MAXTASK 8
// workgroups of the finisher; the last one to finish combines their rows
#define SEPK_FIN_BLOCKS 64

struct sepk_task {
	int mode, nblk, nv, region;
	unsigned maxmask;
	double *ret;      // return slot (32 words) or null
	int zero_first;   // the slot was reset since its last write
	int epot_assign;  // PAIR: epot = e instead of +=
	int also_ecoul;   // PAIR: Coulomb energy also goes to ecoul
	int with_stress;  // BONDED: add the stress to pot_P and pot_P_bond
	double *alpha;    // NH / LF with a fused thermostat
	double Td, dt_tau2, npart;
	int nh_fused;     // LF: the leapfrog applied a pending Nose-Hoover
	int reset_max;    // LF: max_dist2 was reset since the last leapfrog
	double skin, dt;
	int conditional;  // BUILD: only when the rebuild flag is set
	// LF: the per-step record (null: none); entry (*ring_n % ring_cap)
	// receives the return slot's words and tnow, then *ring_n counts it
	double *ring;
	long ring_cap;
	unsigned long long *ring_n;
};

// One bonded term family's per-atom gather: atom o (original numbering)
// adds contrib[3 idx[e] + k] for e in [off[o], off[o + 1]) to its force.
#define SEPK_MAXGATHER 3
struct sepk_gsrc {
	const double *contrib;
	const int *off, *idx;
};
struct sepk_gpend {
	sepk_gsrc g[SEPK_MAXGATHER];
	int n, over;
};

struct sepk_fin {
	sepk_task t[SEPK_MAXTASK];
	int ntask;
	long region_words, region_rows;
	sepdev_sys *sys;
	double *l2;       // SEPK_MAXTASK x SEPK_MAXNV x SEPK_FIN_BLOCKS block results
	// per task: the first task with the same return slot / thermostat
	// variable (-1: none), so that shared ones are loaded and stored once
	signed char rmap[SEPK_MAXTASK], amap[SEPK_MAXTASK];
	int ring_task;    // the leapfrog task writing the per-step record, or -1
};

struct sepk_timer {
	hipEvent_t beg[64], end[64]; // eager launches, flushed in batches
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
	int *cell_count, *cell_start;
	int *cellid_old, *cell_slot, *order_tmp;

	// full neighbour list: entry k of atom t at slot
	// ((k/SEPK_SUB)*npad + t)*SEPK_SUB + k%SEPK_SUB, nbr_count[t] entries,
	// at most cap = lcap*SEPK_SUB
	int cap, lcap;
	int cap_limit; // entries a build may write per atom (0: cap)
	int *nbr, *nbr_count;
	int list_built;
	sepdev_grid grid; // grid of the last build

	// reductions: one row of partial results per block, read by k_fin_rows;
	// SEPK_NREGION regions of region_words = SEPK_MAXNV * region_rows
	// doubles, each column-major
	double *partials;
	long partials_words, region_words, region_rows;
	sepk_task pending[SEPK_MAXTASK];
	int npending;
	double *fin_l2;
	double *ring;                // per-step record (SEPDEV_RING_WORDS each)
	long ring_cap;
	unsigned long long *ring_n;  // steps recorded since the last read

	// return and scalar slots
	double *ret;
	int nret, ret_cap;
	double *scal;
	int nscal, scal_cap;
	sepdev_sys *sys;

	// staging
	void *hstage, *dstage;
	size_t stage_bytes;

	// Deferred work, decided on the host so that a captured step replays
	// the same kernels every time:
	int f_zero_pending;          // f = 0 owed (the next force kernel overwrites)
	// bonded per-atom gathers owed to f, applied in order by the next reader
	// of f (the leapfrog folds them in; sepk_flush_gathers otherwise)
	sepk_gsrc gpend[SEPK_MAXGATHER];
	int ngpend, gpend_over;      // gpend_over: the first starts from f = 0
	int max_reset_pending;       // sys->max_dist2 = 0 owed (the next leapfrog starts at 0)
	unsigned char *ret_zero;     // per return slot: zero owed before the next write
	int nh_pending;              // a Nose-Hoover call waits to be fused into leapfrog
	int nh_slot;
	double nh_Td, nh_dt_tau2;
	int sumv2_valid;             // sys->sumv2 holds sum m v^2 of the current v

	// topology (original atom numbering)
	int nb, na, nd;
	unsigned *blist, *alist, *dlist;
	double *bcontrib, *acontrib, *dcontrib;    // per-term force slots
	int *bcsr_off, *bcsr_idx, *acsr_off, *acsr_idx, *dcsr_off, *dcsr_idx;
	double *blen, *aval, *dval;

	// timing
	int timing;
	sepk_timer t_pair;
	// the last pair-force launch in its overwriting form, for
	// sepdev_time_pair (kernel pointer and its sepk_lj argument block)
	const void *last_lj_kernel;
	void *last_lj;

	// force-matching adjoint: g and dL/dx (3 x npad), parameter-gradient
	// partial rows and sums (allocated on first use)
	double *adj_g, *adj_x, *adj_out, *adj_part;

	// qpos holds pos quantized with box qL, unless qpos_stale (set by
	// position uploads; the leapfrog and the rebuild keep qpos current)
	int qpos_stale;
	double qL[3];

	// Fused step (pair forces, Nose-Hoover and leapfrog in one kernel),
	// armed by the host around a step whose only calls are one pair call,
	// sep_nosehoover and sep_leapfrog: the pair call is recorded instead of
	// launched and the leapfrog launches both.  The fused kernel writes the
	// new positions into pos_alt / qpos_alt (other atoms still gather the
	// old ones) and the host swaps them with s[0].pos / s[0].qpos, so that
	// s[0] always names the current positions.
	sepk_pos *pos_alt;
	int4 *qpos_alt;
	int fuse_armed, fused_pending;
	sepk_lj fused_lj;
	int fused_mixed, fused_fast, fused_ret, fused_epot_assign;
	// the last fused launch, for sepdev_time_pair (dry: outputs to tscratch)
	int last_fused, fused_nh;
	double *fused_alpha;
	double fused_dt, fused_L[3];
	double *tscratch;

	// Concurrent force terms (source/sepomp.c: the reference's OpenMP
	// sections over independent terms), armed by the host around a step.
	// The per-term kernel of a bonded call runs on `side`, which waits for
	// fork_ev; st waits for join_ev before that call's gather, so each atom
	// still receives its forces in call order.  fork_ev is recorded on st
	// just before the step's first pair kernel (after any list rebuild);
	// fork_ready is cleared by anything that moves atoms or by a finisher
	// launch that reads a term region the side stream may write next.
	// Pair call and Coulomb call in one list walk (k_lj_coul), armed by
	// the host around a step that makes the two calls in a row: the pair
	// call is recorded, and the Coulomb call launches both
	// (anything else first launches the pair call on its own).
	int pc_armed, pc_pending;
	sepk_lj pc_lj;
	int pc_mixed, pc_typed, pc_fast, pc_ret, pc_epot_assign;

	hipStream_t side;
	hipEvent_t fork_ev, join_ev;
	int fork_armed, fork_ready;
	unsigned fork_used; // bonded families launched on side since fork_ev

	// graph
	hipGraph_t graph[2];     // captured step sequences (sepdev_graph_begin slots)
	hipGraphExec_t exec[2];
	int capturing;

	int device; // the HIP device current at sepdev_create
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

// ---- device helpers ----------------------------------------------------

static inline unsigned sepk_blocks_xcd(long threads)
{
	const unsigned nb = sepk_blocks(threads);
	return (nb + SEPK_NXCD - 1) / SEPK_NXCD * SEPK_NXCD;
}

// Logical block of an XCD-remapped launch (gridDim.x a multiple of 8).
__device__ __forceinline__ long sepk_xcd_block()
{
	const unsigned per = gridDim.x / SEPK_NXCD;
	return (long)(blockIdx.x % SEPK_NXCD) * per + blockIdx.x / SEPK_NXCD;
}

// Slot of entry e of atom t's list.  Entries come in batches of SEPK_PAD;
// within batch b, lane s of the atom (s = e % SEPK_SUB) takes entries
// s, s + SEPK_SUB, ... and finds them side by side, so it reads its
// SEPK_BATCH indices of a batch with one 16-byte load, and a wave's load
// is 1 KB of consecutive memory.
__device__ __forceinline__ size_t sepk_entry(long e, long t, long npad)
{
	return (((size_t)(e / SEPK_PAD) * npad + t) * SEPK_SUB + e % SEPK_SUB) * SEPK_BATCH +
	       (e / SEPK_SUB) % SEPK_BATCH;
}

// Element j of a per-atom array (or of one component block of npad
// entries) addressed by a 32-bit byte offset from the wave-uniform base:
// one VGPR per address where an index would take two.  No block holds more
// than SEPDEV_MAXN + 64 records of at most 32 bytes (sepdev_create), so the
// offset fits.
template <class T>
__device__ __forceinline__ const T &sepk_at(const T *base, unsigned j)
{
	return *(const T *)((const char *)base + j * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ T &sepk_at(T *base, unsigned j)
{
	return *(T *)((char *)base + j * (unsigned)sizeof(T));
}

// Batch 0 of lane `sub` of atom t; batch b is at [b * npad * SEPK_SUB].
__device__ __forceinline__ const int4 *sepk_lane_list(const int *nbr, long t, int sub)
{
	return (const int4 *)nbr + (size_t)t * SEPK_SUB + sub;
}

// The same for a list kernel's wave, whose lanes take atoms t and lanes sub
// with t * SEPK_SUB + sub = (first of the wave) + (lane in the wave): batch 0
// of the wave's first lane as a wave-uniform base, which lane l reads at
// element l (sepk_at); batch b is at [b * npad * SEPK_SUB].
__device__ __forceinline__ const int4 *sepk_wave_list(const int *nbr, long t, int sub)
{
	const long w = (long)t * SEPK_SUB + sub - (long)(threadIdx.x & 63);
	const long wu = (long)(((unsigned long)(unsigned)__builtin_amdgcn_readfirstlane((int)(w >> 32)) << 32) |
			       (unsigned)__builtin_amdgcn_readfirstlane((int)w));
	return (const int4 *)nbr + wu;
}

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

// The same image with three instructions: d - L*rint(d/L).  Equal to
// sepk_wrap bit for bit whenever |d| < 1.5 L and |d mod L| is not within an
// ulp of L/2 -- true for every listed pair once L/2 exceeds the list
// cutoff plus the skin, which the host checks before choosing it.
__device__ __forceinline__ double sepk_wrap_fast(double d, double L, double invL)
{
	return fma(-L, rint(d * invL), d);
}

// A position as the mixed pair kernels read it: x mod L as the unsigned
// 32-bit fraction rint(x 2^32 / L) mod 2^32 (resolution L / 2^32, 1.6e-8 at
// L = 67), so that (int)(qi - qj) (L / 2^32) is the minimum image of
// xi - xj to within that resolution, with no wrap arithmetic; one 16-byte
// record per neighbour gather instead of the 32-byte FP64 one.  w holds the
// charge as FP32 bits (the mixed Coulomb kernel's charges); the phantom
// (NaN position) gets SEPK_QPHANTOM, a NaN pattern no FP32 conversion of a
// double produces.
#define SEPK_QSCALE 4294967296.0 // 2^32
#define SEPK_QPHANTOM (-1)

__device__ __forceinline__ int4 sepk_quant(const sepk_pos &p, double s0, double s1, double s2)
{
	if (!(p.x == p.x))
		return make_int4(0, 0, 0, SEPK_QPHANTOM);
	return make_int4((int)(unsigned)(long long)rint(p.x * s0),
			 (int)(unsigned)(long long)rint(p.y * s1),
			 (int)(unsigned)(long long)rint(p.z * s2), __float_as_int((float)p.q));
}

// Per-block reduction of NV values per thread into column-major partial
// rows, partial[i*rows + row]:
// columns whose bit is set in maxmask take the maximum, the others the sum,
// in a fixed order (wave butterfly, then the block's waves in index order).
// The finisher kernels (k_fin_rows, k_fin_apply) combine the rows
// afterwards; the kernel boundary makes every row visible to them, so no
// fences are needed.
template <int NV>
__device__ __forceinline__ void sepk_block_partial(double (&v)[NV], unsigned maxmask,
						   double *partial, long row, long rows)
{
	__shared__ double sh[SEPK_BLOCK / SEPK_WAVE][NV];
	const int lane = threadIdx.x & (SEPK_WAVE - 1), wid = threadIdx.x / SEPK_WAVE;
	// butterfly step by step over all columns at once, so that each step's
	// shuffles are in flight together (one wait a step, not one a column)
#pragma unroll
	for (int o = SEPK_WAVE / 2; o > 0; o >>= 1) {
		double y[NV];
#pragma unroll
		for (int i = 0; i < NV; i++)
			y[i] = __shfl_xor(v[i], o, SEPK_WAVE);
#pragma unroll
		for (int i = 0; i < NV; i++)
			v[i] = (maxmask >> i) & 1u ? fmax(v[i], y[i]) : v[i] + y[i];
	}
	if (lane == 0) {
#pragma unroll
		for (int i = 0; i < NV; i++)
			sh[wid][i] = v[i];
	}
	__syncthreads();
	if (threadIdx.x < NV) {
		const int i = threadIdx.x;
		const bool mx = (maxmask >> i) & 1u;
		double s = sh[0][i];
		for (int w = 1; w < SEPK_BLOCK / SEPK_WAVE; w++)
			s = mx ? fmax(s, sh[w][i]) : s + sh[w][i];
		partial[(size_t)i * rows + row] = s;
	}
}

// Nose-Hoover friction update of source/sepintgr.c:151-162, from the sum of
// m v^2 over all atoms: ekin = 0.5 sum / N, T = 0.666667 ekin,
// alpha += dt/tau^2 (T/Td - 1).
__device__ __forceinline__ double sepk_nh_alpha(double alpha, double sumv2, double npart,
						double Td, double dt_tau2)
{
	const double ekin = 0.5 * sumv2 / npart;
	const double temp = 0.666667 * ekin;
	return alpha + dt_tau2 * (temp / Td - 1.0);
}

// Queues a finishing task; it runs (with any queued before it) in one
// launch at the next sepk_flush_tasks, which every leapfrog, every host
// read of device results and the end of a graph capture perform.
int sepk_queue_task(sepdev *d, const sepk_task &t);
int sepk_flush_tasks(sepdev *d);
double *sepk_region(sepdev *d, int region);
// Finishes queued work first if a queued task still reads `region`, which
// the caller's next kernel is about to overwrite.
int sepk_claim_region(sepdev *d, int region);
// A term region for a force call's partial rows: the first one that no
// queued task holds (finishing the queue first if all are held).
int sepk_term_region(sepdev *d, int *region);
// Concurrent terms: records fork_ev on st if the step is armed and it is
// not yet recorded; the stream a bonded call's per-term kernel goes to
// (side once fork_ev is recorded, else st), and the join after it.
int sepk_fork_point(sepdev *d);
// family: 1 bonds, 2 angles, 4 torsions (each has one set of per-term slots)
int sepk_term_stream(sepdev *d, unsigned family, hipStream_t *ks);
int sepk_term_join(sepdev *d, hipStream_t ks);

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

// deferred-work helpers (sepdev_force.hip, sepdev_reduce.hip)
int sepk_materialize_forces(sepdev *d);
// Launches the owed bonded gathers (one kernel for all of them).
int sepk_flush_gathers(sepdev *d);
// launches a pair call recorded for a fused step, if any (sepdev_force.hip)
int sepk_flush_pair(sepdev *d);
int sepk_flush_nh(sepdev *d);
int sepk_put_vec_to(sepdev *d, double *dst, const double *src3n);
int sepk_get_vec_from(sepdev *d, double *dst3n, const double *src);
void sepk_task_ret(sepdev *d, sepk_task &t, int slot);
