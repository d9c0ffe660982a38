/*
 * sephost.h -- internal declarations of the C99 host layer of libsep.so.
 *
 * The host layer keeps the reference's AoS structures as the caller-visible
 * state and drives the device module (sepdev.h).  Each atom array returned
 * by sep_init is registered here; its device context is created at the
 * first call that needs the GPU.
 *
 * Two modes per atom array:
 *   synchronised (default): every API call uploads what it reads from the
 *     AoS, runs its kernels and writes back what it changes before
 *     returning, so the caller sees exactly the reference's post-call state;
 *   resident (sep_gpu_resident): the device copy is authoritative, calls
 *     only enqueue kernels, and sep_gpu_sync() writes the state back.
 */
#ifndef SEPHOST_H
#define SEPHOST_H

#include "sep.h"
#include "../sepdev.h"

#define SEPH_MAX_BIND 32

typedef struct {
	const void *host; /* sepret* or double* (Nose-Hoover alpha) */
	int slot;
} seph_bind;

#define SEPH_CHECKED 4
typedef struct {
	int n[3], opt;
	double rlist2;
} seph_grid_key;

typedef struct {
	seppart *atoms;
	long n;
	int nneighb;          /* per-atom half-list capacity from sep_init */
	sepdev *dev;
	int resident;
	int precision;        /* SEP_GPU_FP64 / SEP_GPU_MIXED */
	int ret_scratch;      /* slot for synchronised calls */
	int alpha_scratch;
	seph_bind rets[SEPH_MAX_BIND];
	int nrets;
	seph_bind alphas[SEPH_MAX_BIND];
	int nalphas;
	/* topology last uploaded */
	const unsigned *topo_b, *topo_a, *topo_d;
	unsigned topo_nb, topo_na, topo_nd;
	int partners_on_device;
	int timing;
	/* the single type label shared by all atoms (0 if mixed); trusted only
	 * while resident, where types change only through sep_gpu_upload */
	char uniform_type;
	int uniform_type_valid;
	/* grids whose list capacity was confirmed on the host (resident mode);
	 * one step may alternate between grids, e.g. brute-force pair and
	 * Coulomb calls at different cutoffs */
	seph_grid_key checked[SEPH_CHECKED];
	int nchecked;
	/* sep_gpu_list_limit: entries per atom a build may write (0: none) */
	int list_limit;
	/* host staging */
	double *v3;
	int *i3;
	double *s1, *s2;
	signed char *c1;
	int *i1;
} seph_ctx;

/* registry */
void seph_register(seppart *atoms, long n, int nneighb);
void seph_unregister(seppart *atoms);
/* The per-atom neighbour arrays were reallocated with nneighb ints. */
void seph_set_nneighb(seppart *atoms, int nneighb);
seph_ctx *seph_find(const seppart *atoms);
/* context with a device; creates the device state on first use */
seph_ctx *seph_ctx_for(seppart *atoms, const sepsys *sys);

/* errors */
int seph_fail(int code, const char *fmt, ...);
int seph_dev_fail(int rc); /* copies the device module's message */

/* transfers between the AoS and the device (original order) */
int seph_put_x(seph_ctx *c);
int seph_put_v(seph_ctx *c);
int seph_put_f(seph_ctx *c);
int seph_put_xn_cross(seph_ctx *c);
int seph_put_scalars(seph_ctx *c);
int seph_put_partners(seph_ctx *c);
int seph_put_topology(seph_ctx *c, const sepsys *sys);
int seph_get_f(seph_ctx *c);
int seph_get_state(seph_ctx *c); /* x, v, a, cross counters */
int seph_put_all(seph_ctx *c, const sepsys *sys);
int seph_get_all(seph_ctx *c, sepsys *sys);
/* The per-atom state only: device -> AoS (a resident system), and AoS ->
 * device (x, v, f, a, xn and crossings, mass, charge, type, molecule). */
int seph_get_atoms(seph_ctx *c);
int seph_put_atoms(seph_ctx *c);
/* The sepret fields of a device return slot (SEPDEV_RET_WORDS) into r;
 * the fields the hot path does not write are left as they are. */
void seph_ret_unpack(sepret *r, const double *w);
int seph_check_device_error(seph_ctx *c);
/* after a list overflow: grow the list (clearing the error) or report it */
int seph_grow_list(seph_ctx *c, sepdev_sys *ds);

/* return/alpha slots */
int seph_ret_slot(seph_ctx *c, sepret *ret);   /* resident: bound slot */
int seph_alpha_slot(seph_ctx *c, double *alpha);
void seph_ret_merge(sepret *ret, const double *w, int epot_assign, int bond_stress);

/* grids */
void seph_grid_reference(const sepsys *sys, unsigned opt, sepdev_grid *g);
void seph_grid_own(const sepsys *sys, double rlist, unsigned opt, int bonded_rule,
		   sepdev_grid *g);

/* list build + AoS half list; used by the pair functions */
int seph_build(seph_ctx *c, sepsys *sys, const sepdev_grid *g, int conditional);
/* device-only build (resident mode); the first build of a grid is checked
 * for list overflow on the host and grown once */
int seph_build_resident(seph_ctx *c, const sepdev_grid *g, int conditional);

/* potentials served on the device */
enum { SEPH_POT_LJ = 0, SEPH_POT_LJ_SHIFT = 1, SEPH_POT_WCA = 2, SEPH_POT_NONE = -1 };
int seph_identify_potential(double (*fun)(double, char));

#endif
