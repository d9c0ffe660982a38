/*
 * sepgpu.h -- MI355X extension of the seplib C99 API.
 *
 * Nothing here has a counterpart in the reference; the reference API in
 * sepinit.h ... sepcoulomb.h is served unchanged by libsep.so and runs its
 * hot path on the GPU.  This header adds what a device library needs on
 * top of that API:
 *
 *   - error reporting (the reference's functions are void and its
 *     SEP_SUCCESS == SEP_FAILURE, so failures are reported here);
 *   - the precision mode of the pair kernels;
 *   - device residency: with sep_gpu_resident(atoms, sys, 1) the reference
 *     API calls on those atoms work on device state only and return without
 *     synchronising; sep_gpu_sync() brings atoms, sys and every sepret /
 *     Nose-Hoover alpha used since then back to the host;
 *   - a fused multi-step run of a recorded call sequence, replayed from a
 *     HIP graph;
 *   - diagnostics used by tests and benchmarks (neighbour pairs, timings).
 *
 * Every function returns SEP_GPU_OK (0) or an error code and never exits.
 * A system on the device holds at most 2^27 - 64 atoms (every per-atom
 * record is reached by a 32-bit byte offset); a larger one is refused with
 * SEP_GPU_EARG when its device state is created.
 */
#ifndef SEP_MI355X_SEPGPU_H
#define SEP_MI355X_SEPGPU_H

#include <stddef.h>
#include "sepstrct.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- errors ---------------------------------------------------------- */
#define SEP_GPU_OK 0
#define SEP_GPU_EDEVICE 1      /* HIP runtime failure (message has details)   */
#define SEP_GPU_EUNSUPPORTED 2 /* potential / option not on the device path   */
#define SEP_GPU_ENEIGHB 3      /* per-atom neighbour capacity exceeded        */
#define SEP_GPU_ECELL 4        /* atom maps outside the cell grid (x>=L, NaN) */
#define SEP_GPU_EARG 5         /* invalid argument                            */
#define SEP_GPU_EHOST 6        /* sep_error() was called                      */
#define SEP_GPU_ESTATE 7       /* call not valid in the current mode          */

/* Code of the most recent failure since the last clear (0 if none). */
int sep_gpu_last_error(void);
/* Text of the most recent failure ("" if none). */
const char *sep_gpu_last_error_message(void);
void sep_gpu_clear_error(void);

/* ---- precision ------------------------------------------------------- */
#define SEP_GPU_FP64 0  /* all pair arithmetic in FP64 (parity mode)          */
#define SEP_GPU_MIXED 1 /* FP32 pair arithmetic on FP64-formed dx, FP64 sums,
                           FP64 positions and integrator                      */
int sep_gpu_set_precision(seppart *atoms, sepsys *sys, int mode);

/* ---- residency ------------------------------------------------------- */
/* on = 1: upload atoms/sys and keep the device copy authoritative; on = 0:
 * synchronise back to the host and return to per-call synchronisation. */
int sep_gpu_resident(seppart *atoms, sepsys *sys, int on);
/* Copy device state of a resident system back: x, v, f, a, xn, crossings,
 * cross_neighb, sys->{neighb_flag, nupdate_neighb, max_dist2}, and every
 * sepret and alpha that device calls have used. */
int sep_gpu_sync(seppart *atoms, sepsys *sys);
/* Push host edits of a resident system (positions, velocities, masses,
 * charges, types, box) to the device. */
int sep_gpu_upload(seppart *atoms, sepsys *sys);

/* ---- fused run ------------------------------------------------------- */
/* One entry of a per-step call sequence; the fields used depend on op and
 * mirror the arguments of the reference function named in the comment. */
enum {
	SEP_GPU_OP_PAIRS = 1,     /* sep_force_pairs(types, cf=p[0], fun, opt)          */
	SEP_GPU_OP_LJ = 2,        /* sep_force_lj(types, p[0..3], opt)                  */
	SEP_GPU_OP_COULOMB = 3,   /* sep_coulomb_sf(cf=p[0], opt)                        */
	SEP_GPU_OP_BOND = 4,      /* sep_stretch_harmonic(type, lbond=p[0], ks=p[1])    */
	SEP_GPU_OP_ANGLE = 5,     /* sep_angle_harmonic(type, angle0=p[0], k=p[1])      */
	SEP_GPU_OP_COSSQ = 6,     /* sep_angle_cossq(type, angle0=p[0], k=p[1])         */
	SEP_GPU_OP_TORSION = 7,   /* sep_torsion_Ryckaert(type, g=p[0..5])              */
	SEP_GPU_OP_NOSEHOOVER = 8,/* sep_nosehoover(Td=p[0], alpha, Q=p[1])             */
	SEP_GPU_OP_LEAPFROG = 9   /* sep_leapfrog()                                      */
};
typedef struct {
	int op;
	char types[2];
	int type;
	unsigned opt;
	double p[6];
	double (*fun)(double, char);
	double *alpha;
} sep_gpu_op;

#define SEP_GPU_RUN_GRAPH 1 /* replay the step from a captured HIP graph */
/* Run the per-term kernels of the bonded calls on a second stream, beside
 * the pair kernels of the same step (the GPU form of the reference's
 * OpenMP sections over independent force terms, source/sepomp.c and
 * prgs/prg5.c:58-77).  Forces, energies and stress are bit-identical to the
 * single-stream step: each atom still adds its terms in call order. */
#define SEP_GPU_RUN_CONCURRENT 2

/* Runs nsteps of { sep_reset_retval(ret); sep_reset_force(); ops[0..nops) }
 * on a resident system without host synchronisation; ret (and each alpha)
 * hold the last step's values after sep_gpu_sync.  At the end it reads the
 * device error word once (one small copy): a step that fails on the device
 * (SEP_GPU_ENEIGHB: a list build overflowed; SEP_GPU_ECELL: an atom outside
 * the cell grid) and every step after it change nothing, so the system stays
 * at the step before the failure (sys->tnow tells which), and the code is
 * returned.  After SEP_GPU_ENEIGHB the list has been grown (unless a
 * sep_gpu_list_limit forbids it), so calling again continues the run. */
int sep_gpu_run(seppart *atoms, sepsys *sys, sepret *ret,
		const sep_gpu_op *ops, int nops, long nsteps, int flags);

/* Bound on the device neighbour list: at most `entries` neighbours per atom
 * in the full list (both directions; rounded down to a multiple of 16, at
 * least 16), 0 for none.  The reference sizes each atom's half list once,
 * with the nneighb of sep_init, and stops at "Too many neighbours"
 * (source/sepprfrc.c:499-501); this library grows its list instead, unless
 * a bound is set: then a build that needs more raises SEP_GPU_ENEIGHB.  The
 * list is grown to the bound if it is smaller; the current list stays in
 * use until the next build.  No counterpart in the reference. */
int sep_gpu_list_limit(seppart *atoms, sepsys *sys, int entries);

/* ---- device checkpoint ------------------------------------------------- */
/* Writes the device state of a resident system to `file`: positions (also in
 * their quantized form), velocities, forces, accelerations, xn,
 * cross_neighb and crossings, masses, types, molecules and bonded partners,
 * all in the device's own atom order together with that order; the current
 * neighbour list; the sepsys counters (tnow, neighb_flag, nupdate_neighb,
 * max_dist2) and the thermostat's running sum; and, if given, the sepret
 * and Nose-Hoover alpha the run uses.  No host copy of the atoms is made.
 * sep_gpu_restore reads it into a resident system of the same size (which
 * may be another system, in another process) and the run continues exactly
 * as the one that wrote it, bit for bit; it also sets the sepsys counters,
 * *ret and *alpha on the host.  The reference's own checkpoint is sep_save,
 * the raw AoS records (source/sepmisc.c:320-330); this one has no
 * counterpart there. */
int sep_gpu_checkpoint(seppart *atoms, sepsys *sys, const sepret *ret, const double *alpha,
		       const char *file);
int sep_gpu_restore(seppart *atoms, sepsys *sys, sepret *ret, double *alpha,
		    const char *file);

/* ---- multi-type LJ and force matching --------------------------------- */
/* The pair forces of sep_force_lj (include/sepprfrc.h:69, source/
 * sepprfrc.c:743-923) for every pair of the ntypes labels in `types` in one
 * list pass: params[(i * ntypes + k) * 4 + {0,1,2,3}] = {rc, eps, sigma, aw}
 * of type pair (types[i], types[k]) (symmetric), at most 3 types.  Adds to
 * f, ret.epot and ret.pot_P as the reference's three calls would. */
int sep_gpu_force_lj_table(seppart *atoms, int ntypes, const char *types,
			   const double *params, sepsys *sys, sepret *ret,
			   const unsigned opt);

/* Adjoint of a force-matching loss L(F) with F the forces above, on the
 * current positions and neighbour list (call sep_gpu_force_lj_table on the
 * same positions first): dLdF (3n, original order, dL/dF_i) in; out dL/dx
 * (3n) and dLdtheta[(i * ntypes + k) * 2 + {0,1}] = dL/d eps, dL/d sigma of
 * pair (i, k).  Cutoffs are held fixed.  No reference counterpart. */
int sep_gpu_lj_adjoint(seppart *atoms, sepsys *sys, int ntypes, const char *types,
		       const double *params, const double *dLdF, double *dLdx,
		       double *dLdtheta);

/* sep_gpu_lj_adjoint without any device or host allocation, for callers that
 * cannot allocate (a fitting loop in a process that forbids allocation after
 * its set-up, code that must not reach hipMalloc).  No counterpart in the
 * reference.  Same arguments and results, bit for bit.  Its device buffers
 * (g, dL/dx, the parameter-gradient rows and sums: about 50 bytes per atom)
 * are a fixed 32 MiB array that each GPU the library runs on holds until the
 * process ends, so it serves systems of up to about 670000 atoms and refuses
 * larger ones with SEP_GPU_EARG.  The atoms must already have device state
 * (any force call on them creates it; otherwise SEP_GPU_ESTATE).  All calls
 * on one GPU share the array: calls that overlap on one GPU (from several
 * host threads) race. */
int sep_gpu_lj_adjoint_noalloc(seppart *atoms, sepsys *sys, int ntypes, const char *types,
			       const double *params, const double *dLdF, double *dLdx,
			       double *dLdtheta);

/* ---- per-step record --------------------------------------------------- */
/* No counterpart in the reference, whose samplers read sepret on the host
 * after every step (the stress autocorrelation of source/sepsampler.c:
 * 280-351 reads pot_P and kin_P each step).  With a record of `capacity`
 * steps, every leapfrog of these atoms appends its sepret's epot, ecoul,
 * ekin, pot_P, kin_P and pot_P_bond (the values the reference's
 * sep_leapfrog call would leave in that sepret) and sys->tnow to a device
 * ring, inside captured graphs too, so a resident run needs no host sync
 * per step.  capacity 0 stops recording.  Allocates: call it outside
 * sep_gpu_run. */
int sep_gpu_record_steps(seppart *atoms, long capacity);
/* Copies the recorded steps still held, oldest first, into rets[0..*n)
 * (only the fields listed above; others are zero) and tnow[0..*n) (either
 * may be NULL), and empties the record.  max (the room in each array) must
 * be at least the capacity; *lost receives the number of steps overwritten
 * since the last read. */
int sep_gpu_read_steps(seppart *atoms, sepret *rets, double *tnow, long max, long *n,
		       long *lost);

/* ---- diagnostics ----------------------------------------------------- */
/* Unordered neighbour pairs (i < j, original atom indices) of the current
 * device list, written as pairs[2k], pairs[2k+1]; *npairs receives the
 * count (also when cap is too small, then nothing is written). */
int sep_gpu_neighbour_pairs(seppart *atoms, sepsys *sys, int *pairs,
			    long cap, long *npairs);

/* Kernel timing with HIP events on the library's stream around each eager
 * launch of the pair-force kernel (kernel name "pair"; launches inside a
 * captured graph are not timed).  sep_gpu_timing_reset starts a new window. */
int sep_gpu_timing_enable(seppart *atoms, int on);
int sep_gpu_timing_reset(seppart *atoms);
int sep_gpu_timing_get(seppart *atoms, const char *kernel, double *total_ms,
		       long *count);

/* Mean duration of the pair-force kernel: the last configuration that a
 * pair call launched, re-launched reps times back to back on the library
 * stream between one pair of HIP events.  Afterwards f holds the pair forces
 * of the current positions (as after sep_reset_force + that call).  When
 * the last launch was a fused step (pair forces, Nose-Hoover and leapfrog
 * of a sep_gpu_run step), that kernel is timed instead, with its integrator
 * outputs sent to scratch: the state is left as it was. */
int sep_gpu_time_pair_kernel(seppart *atoms, int reps, double *ms_per_launch);

/* Mean full-list length of the current device list (entries per atom). */
int sep_gpu_list_stats(seppart *atoms, double *mean_full, int *max_full);

/* Device synchronisation barrier for timing. */
int sep_gpu_synchronize(seppart *atoms);

/* ---- batches of independent systems ---------------------------------- */
/* No counterpart in the reference, whose API advances one system per call
 * sequence.  A batch holds nsys systems, each the caller's own seppart
 * array and sepsys as set up for the reference's step loop
 * (prgs/prg1.c:55-65): sep_reset_retval, sep_reset_force,
 * sep_force_pairs(atoms, types, cf, sep_lj_shift, sys, ret, SEP_ALL),
 * sep_nosehoover(atoms, Td[s], &alpha[s], tau[s], sys), sep_leapfrog.
 * Every atom must carry the type types[0] (== types[1]); sys->cf and
 * sys->skin give each system's list cutoff, which must stay below half its
 * box; cells per axis = floor(L / (cf + skin)) and, unlike the reference's
 * builder, a box with fewer than three cells on an axis still lists every
 * pair once from each end.  The systems are advanced together, one
 * persistent launch per step (replayed from a HIP graph) after the list
 * rebuild launches, which return at once for systems whose rebuild flag is
 * clear.  precision: SEP_GPU_FP64 or SEP_GPU_MIXED; cap: per-atom full-list
 * capacity (0: 256).  Returns NULL on failure (sep_gpu_last_error()). */
typedef struct sep_gpu_batch sep_gpu_batch;
sep_gpu_batch *sep_gpu_batch_create(int nsys, seppart **atoms, sepsys **sys,
				    const char *types, double cf, const double *Td,
				    const double *tau, const double *alpha, int precision,
				    int cap);
void sep_gpu_batch_free(sep_gpu_batch *b);
/* nsteps steps of every system, replayed from a captured graph
 * (run_eager: the same launches without the graph). */
int sep_gpu_batch_run(sep_gpu_batch *b, long nsteps);
int sep_gpu_batch_run_eager(sep_gpu_batch *b, long nsteps);
/* Pair forces, energy and stress at the current positions in the given
 * precision (nothing moves; f holds the pair forces afterwards). */
int sep_gpu_batch_eval(sep_gpu_batch *b, int precision);
/* Copy every system's x, v, f, a, crossings and sys->{max_dist2, tnow,
 * nupdate_neighb, neighb_flag} back; rets[s] gets epot, ekin, pot_P and
 * kin_P of the last step (or eval) and alpha[s] the thermostat state
 * (either may be NULL). */
int sep_gpu_batch_sync(sep_gpu_batch *b, sepret *rets, double *alpha);
int sep_gpu_batch_synchronize(sep_gpu_batch *b);
/* Tile shape: lanes per atom of each system (1, 2, 4, 8 or 16; NULL keeps
 * them) and the persistent grid in workgroups (<= 0 keeps it). */
int sep_gpu_batch_set_shape(sep_gpu_batch *b, const int *lanes, int grid);
int sep_gpu_batch_get_shape(sep_gpu_batch *b, int *lanes, int *grid);
/* Mean duration (ms) of a force evaluation over the systems with
 * mask[s] != 0 (NULL: all), over reps back-to-back launches. */
int sep_gpu_batch_time_eval(sep_gpu_batch *b, const int *mask, int reps, double *ms);
/* Autotuner: for each size bucket (N < 2048, N < 8192, larger) the lanes
 * per atom with the fastest force evaluation of that bucket's systems, then
 * the fastest persistent grid; reps launches per candidate.  The chosen
 * shape stays set and is returned in lanes_out[nsys] and grid_out (either
 * may be NULL). */
int sep_gpu_batch_tune(sep_gpu_batch *b, int reps, int *lanes_out, int *grid_out);

#ifdef __cplusplus
}
#endif
#endif
