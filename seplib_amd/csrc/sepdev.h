/*
 * sepdev.h -- internal C ABI between the C99 host layer (csrc/host) and
 * the HIP device module (csrc/hip) of libsep.so.
 *
 * Plain pointers and sizes only.  Host arrays passed in or out are in the
 * caller's ORIGINAL atom order; the device keeps its own cell-sorted order
 * and converts.  Every function returns SEPDEV_OK or a SEP_GPU_* code from
 * sepgpu.h and records a message retrievable with sepdev_error_message().
 */
#ifndef SEPDEV_H
#define SEPDEV_H

#ifdef __cplusplus
extern "C" {
#endif

#define SEPDEV_OK 0
/* lanes per atom in the list kernels; list capacities are multiples */
#define SEPDEV_LANES 4

typedef struct sepdev sepdev;

/* ---- return-value slots: a device mirror of the sepret fields the hot
 * path writes.  Layout of one slot (doubles): */
enum {
	SEPDEV_RET_EPOT = 0,
	SEPDEV_RET_ECOUL = 1,
	SEPDEV_RET_EKIN = 2,
	SEPDEV_RET_POTP = 3,   /* 9 doubles, row-major */
	SEPDEV_RET_KINP = 12,  /* 9 */
	SEPDEV_RET_BONDP = 21, /* 9 */
	SEPDEV_RET_WORDS = 32
};

/* ---- system scalars mirrored on the device */
typedef struct {
	double max_dist2;
	double tnow;
	unsigned neighb_flag;
	unsigned nupdate_neighb;
	unsigned xn_pending; /* xn = x and cross_neighb = 0 still owed */
	int error;           /* sticky device error code (SEP_GPU_*)   */
	int max_full;        /* largest full-list length of the last build */
	int unclean;         /* the last binning saw an atom outside [0, n*l) */
	int frozen;          /* error was set when this rebuild's binning ended */
	double sumv2;        /* last sum m v^2 (Nose-Hoover)           */
} sepdev_sys;
/* While error is nonzero every kernel of a step leaves the state alone (a
 * fused step copies its positions through unchanged) and the finisher
 * applies nothing, so a resident run that fails stays at the step before
 * the failure; the host reads and clears error. */

/* ---- cell grid + stencil of one neighbour-list build */
#define SEPDEV_MAX_OFF 28
enum { SEPDEV_RULE_NE = 0, SEPDEV_RULE_GT = 1, SEPDEV_RULE_LT = 2 };
typedef struct {
	int n[3];        /* cells per axis                                  */
	int ref_index;   /* 1: reference cell formula (truncation, linear
			    index aliasing); 0: floor + periodic wrap        */
	double inv[3];   /* unused by ref_index (which divides by l)        */
	double l[3];     /* cell edge per axis                              */
	double L[3];     /* box                                             */
	double rlist2;   /* list cutoff squared                             */
	int noff;
	int opt;         /* SEP_ALL / SEP_EXCL_BONDED / SEP_EXCL_SAME_MOL    */
	int bonded_all;  /* EXCL_BONDED: 0 bonds only, 1 bonds+angles+dihedrals,
			    2 exclude when exactly one kind is shared          */
	int shift_ok;    /* >= 5 cells on every axis: per-cell image shifts  */
	int off[SEPDEV_MAX_OFF][4]; /* dx, dy, dz, rule                     */
} sepdev_grid;

/* ---- lifetime.  At most SEPDEV_MAXN atoms a device (2^27 - 64, so that
 * every per-atom record is reached by a 32-bit byte offset). */
#define SEPDEV_MAXN ((1L << 27) - 64)
sepdev *sepdev_create(long npart, int *err);
void sepdev_destroy(sepdev *d);
const char *sepdev_error_message(void);
int sepdev_last_code(void);
void sepdev_clear_error(void);

/* ---- host <-> device transfers (original order, AoS-interleaved 3-vectors
 * x0 y0 z0 x1 y1 z1 ...) */
enum {
	SEPDEV_X = 0, SEPDEV_V = 1, SEPDEV_F = 2, SEPDEV_A = 3, SEPDEV_XN = 4
};
enum { SEPDEV_CROSS_NB = 0, SEPDEV_CROSS = 1 };
int sepdev_put_vec(sepdev *d, int which, const double *src3n);
int sepdev_get_vec(sepdev *d, int which, double *dst3n);
int sepdev_put_ivec(sepdev *d, int which, const int *src3n);
int sepdev_get_ivec(sepdev *d, int which, int *dst3n);
int sepdev_put_scalars(sepdev *d, const double *mass, const double *charge,
		       const signed char *type, const int *molindex);
int sepdev_put_charges(sepdev *d, const double *charge);
int sepdev_put_types(sepdev *d, const signed char *type);
/* bonded partners, 40 ints per atom (bond[10], angle[10], dihed[20]) */
int sepdev_put_partners(sepdev *d, const int *partners40n);
int sepdev_put_sys(sepdev *d, const sepdev_sys *s);
int sepdev_get_sys(sepdev *d, sepdev_sys *s);

/* ---- return slots */
int sepdev_ret_alloc(sepdev *d);                 /* returns slot id >= 0 */
int sepdev_ret_zero(sepdev *d, int slot);
int sepdev_ret_get(sepdev *d, int slot, double *words);
int sepdev_ret_put(sepdev *d, int slot, const double *words);
/* scalar slots (Nose-Hoover alpha) */
/* ---- per-step record: at every leapfrog, the leapfrog's return slot
 * (SEPDEV_RET_WORDS) and then tnow are appended to a device ring of
 * `capacity` steps.  capacity 0 frees the ring.  Reading copies the steps
 * still held, oldest first (*n of them; *lost: steps overwritten since the
 * last read), and empties the ring; max must be at least the capacity. */
#define SEPDEV_RING_WORDS (SEPDEV_RET_WORDS + 1)
int sepdev_ring_alloc(sepdev *d, long capacity);
int sepdev_ring_read(sepdev *d, double *words, long max, long *n, long *lost);

int sepdev_scalar_alloc(sepdev *d, double init);
int sepdev_scalar_get(sepdev *d, int slot, double *v);
int sepdev_scalar_put(sepdev *d, int slot, double v);

/* ---- neighbour list.  conditional = 1: do nothing unless the device
 * neighb_flag is set (resident mode); the flag is cleared afterwards. */
int sepdev_build_list(sepdev *d, const sepdev_grid *g, int conditional);
/* grow the per-atom full-list capacity (drops the current list) */
int sepdev_reserve_list(sepdev *d, int cap);
/* upper bound on the entries a build writes per atom (full list), rounded
 * down to a multiple of SEPDEV_LANES * 4; 0: the allocation alone.  The
 * list is grown to the bound if it is smaller; a longer list raises
 * SEP_GPU_ENEIGHB.  Returns the bound in effect in *applied. */
int sepdev_list_limit(sepdev *d, int cap, int *applied);
int sepdev_list_capacity(sepdev *d);
/* sys->max_dist2 = 0 on the device (resident sep_reset_force) */
int sepdev_reset_max_dist(sepdev *d);
/* half list in original numbering: counts[o] entries j > o for atom o,
 * written to half[o*stride + e]; returns the largest count in *maxc */
int sepdev_get_half_list(sepdev *d, int *counts, int *half, int stride,
			 int *maxc);
int sepdev_list_stats(sepdev *d, double *mean_full, int *max_full);

/* ---- forces */
typedef struct {
	double cf2;             /* interaction cutoff squared               */
	double eps48, eps4, awh, aw, sig2, ushift;  /* LJ constants         */
	double L[3];
	char t0, t1;            /* type filter                              */
	int typed;              /* 0: every atom has a matching type        */
	int epot_assign;        /* 1: ret.epot = e (sep_force_pairs list)   */
	int mixed;              /* FP32 pair arithmetic                      */
} sepdev_lj_args;
int sepdev_lj(sepdev *d, const sepdev_lj_args *a, int ret_slot);

typedef struct {
	double cf, cf2, icf, icf2;
	double L[3];
	int mixed;
} sepdev_coul_args;
int sepdev_coulomb(sepdev *d, const sepdev_coul_args *a, int ret_slot);

/* multi-type LJ from a type-pair table (one list pass) and the adjoint of a
 * force-matching loss.  params[(i * ntypes + k) * 4 + {0,1,2,3}] =
 * {rc, eps, sigma, aw} of type pair (types[i], types[k]), symmetric. */
#define SEPDEV_MAXT 3
typedef struct {
	int ntypes;
	const char *types;
	const double *params;
	double L[3];
} sepdev_ljtab_args;
int sepdev_lj_table(sepdev *d, const sepdev_ljtab_args *a, int ret_slot);
/* g3n: dL/dF per atom (original order, 3n); out: dL/dx (3n) and
 * dldtheta[(i * ntypes + k) * 2 + {0,1}] = dL/d eps, dL/d sigma */
int sepdev_lj_adjoint(sepdev *d, const sepdev_ljtab_args *a, const double *g3n,
		      double *dldx3n, double *dldtheta);
/* The same with its device buffers (g, dL/dx, the parameter-gradient rows
 * and sums) in a fixed device array of SEPDEV_SCRATCH_BYTES per GPU instead
 * of memory allocated on first use: no allocation at all.  Calls that
 * overlap on one GPU share the array and race. */
#define SEPDEV_SCRATCH_BYTES (32L << 20)
int sepdev_lj_adjoint_noalloc(sepdev *d, const sepdev_ljtab_args *a, const double *g3n,
			      double *dldx3n, double *dldtheta);

/* bonded terms; lists hold ORIGINAL atom indices as in sepmolinfo */
int sepdev_put_topology(sepdev *d, const unsigned *blist, int nb,
			const unsigned *alist, int na, const unsigned *dlist,
			int nd);
int sepdev_bond(sepdev *d, int type, double lbond, double ks,
		const double *L, int ret_slot);
int sepdev_angle(sepdev *d, int type, double angle0, double k, int cossq,
		 const double *L, int ret_slot);
int sepdev_torsion(sepdev *d, int type, const double *g, const double *L,
		   int ret_slot);
int sepdev_get_bonded_values(sepdev *d, double *blengths, double *angles,
			     double *dihedrals);

/* ---- integrators */
int sepdev_zero_forces(sepdev *d);
int sepdev_nosehoover(sepdev *d, double Td, double tau, double dt,
		      int alpha_slot);
int sepdev_leapfrog(sepdev *d, double dt, double skin, const double *L,
		    int ret_slot);
/* on = 1: until the matching off call, a non-typed pair call that starts the
 * step's forces is recorded and launched fused with the next leapfrog (any
 * other use of the forces launches it first); off launches a recorded call
 * that is still waiting */
int sepdev_fuse_step(sepdev *d, int on);
/* Arms (on = 1) or disarms concurrent force terms for the calls of one step:
 * bonded per-term kernels overlap the pair kernels on a second stream. */
int sepdev_concurrent_terms(sepdev *d, int on);
/* Arms (on = 1) or disarms, for the calls of one step, the single list
 * walk of a pair call followed by a Coulomb call. */
int sepdev_fuse_pair_coulomb(sepdev *d, int on);

/* ---- device checkpoint: the device state between steps (per-atom arrays
 * in device order, the order, the neighbour list, the system scalars and
 * the host-side facts the next step depends on) to or from an open FILE;
 * reading needs a device of the same size and replaces its state */
int sepdev_checkpoint_write(sepdev *d, void *file);
int sepdev_checkpoint_read(sepdev *d, void *file);

/* ---- misc */
int sepdev_synchronize(sepdev *d);
int sepdev_timing(sepdev *d, int on);
int sepdev_timing_reset(sepdev *d);
int sepdev_timing_get(sepdev *d, const char *name, double *ms, long *count);
/* mean duration of the last pair-force configuration over reps back-to-back
 * launches (replaces f with the pair forces; queues no reduction) */
int sepdev_time_pair(sepdev *d, int reps, double *ms_per_launch);
/* Two graph slots: capturing into slot 0 drops both graphs first (after
 * their queued replays), capturing into slot 1 keeps slot 0. */
int sepdev_graph_begin(sepdev *d, int slot);
/* 1 between sepdev_graph_begin and sepdev_graph_end */
int sepdev_capturing(sepdev *d);
int sepdev_graph_end(sepdev *d, int slot);
int sepdev_graph_launch(sepdev *d, int slot, long times);
int sepdev_graph_drop(sepdev *d);

/* ---- batches: many independent LJ systems (sep_force_pairs "AA" with
 * sep_lj_shift, sep_nosehoover, sep_leapfrog each), one persistent launch
 * per step.  Host arrays in original order, 3n interleaved. */
typedef struct sepdev_batch sepdev_batch;
typedef struct {
	long n;
	double L[3];
	double rlist2;                 /* (sys.cf + skin)^2                 */
	double dt, Td, tau, skin, alpha;
	double tnow;
	const double *x, *v, *m;
	const int *cr;                 /* crossings (may be null)           */
} sepdev_batch_sys;
sepdev_batch *sepdev_batch_create(int nsys, const sepdev_batch_sys *sys, double cf,
				  int cap, int *err);
void sepdev_batch_destroy(sepdev_batch *b);
int sepdev_batch_run(sepdev_batch *b, long nsteps, int mixed, int graph);
int sepdev_batch_eval(sepdev_batch *b, int mixed);
int sepdev_batch_synchronize(sepdev_batch *b);
/* scal: alpha, max_dist2, tnow, nupdate, neighb_flag, max_full, error, sumv2 */
int sepdev_batch_get(sepdev_batch *b, int s, double *x, double *v, double *f,
		     double *a, int *cr, double *ret, double *scal);
int sepdev_batch_set_shape(sepdev_batch *b, const int *sub, int grid);
int sepdev_batch_get_shape(sepdev_batch *b, int *sub, int *grid);
int sepdev_batch_time_eval(sepdev_batch *b, const int *mask, int mixed, int reps,
			   double *ms);

#ifdef __cplusplus
}
#endif
#endif
