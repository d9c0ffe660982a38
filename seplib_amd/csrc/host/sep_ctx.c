/*
 * sep_ctx.c -- registry of atom arrays, their device contexts, error state,
 * AoS <-> device marshalling, return-value slots and cell grids.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sephost.h"

/* ---- errors ----------------------------------------------------------- */

static int g_code;
static char g_msg[512];

int seph_fail(int code, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_msg, sizeof g_msg, fmt, ap);
	va_end(ap);
	g_code = code;
	return code;
}

int seph_dev_fail(int rc)
{
	if (rc == 0)
		return 0;
	g_code = rc;
	snprintf(g_msg, sizeof g_msg, "%s", sepdev_error_message());
	return rc;
}

int sep_gpu_last_error(void) { return g_code; }
const char *sep_gpu_last_error_message(void) { return g_msg; }
void sep_gpu_clear_error(void)
{
	g_code = 0;
	g_msg[0] = '\0';
	sepdev_clear_error();
}

/* ---- registry --------------------------------------------------------- */

typedef struct {
	seppart *atoms;
	long n;
	int nneighb;
	seph_ctx *ctx;
} seph_entry;

static seph_entry *g_reg;
static int g_nreg, g_capreg;

static seph_entry *find_entry(const seppart *atoms)
{
	for (int i = 0; i < g_nreg; i++)
		if (g_reg[i].atoms == atoms)
			return &g_reg[i];
	return NULL;
}

void seph_register(seppart *atoms, long n, int nneighb)
{
	seph_entry *e = find_entry(atoms);
	if (!e) {
		if (g_nreg == g_capreg) {
			g_capreg = g_capreg ? 2 * g_capreg : 16;
			g_reg = realloc(g_reg, sizeof(seph_entry) * (size_t)g_capreg);
		}
		e = &g_reg[g_nreg++];
		e->ctx = NULL;
	}
	e->atoms = atoms;
	e->n = n;
	e->nneighb = nneighb;
}

static void free_ctx(seph_ctx *c)
{
	if (!c)
		return;
	sepdev_destroy(c->dev);
	free(c->v3);
	free(c->i3);
	free(c->s1);
	free(c->s2);
	free(c->c1);
	free(c->i1);
	free(c);
}

void seph_set_nneighb(seppart *atoms, int nneighb)
{
	seph_entry *e = find_entry(atoms);
	if (!e)
		return;
	e->nneighb = nneighb;
	if (e->ctx)
		e->ctx->nneighb = nneighb;
}

void seph_unregister(seppart *atoms)
{
	for (int i = 0; i < g_nreg; i++)
		if (g_reg[i].atoms == atoms) {
			free_ctx(g_reg[i].ctx);
			g_reg[i] = g_reg[--g_nreg];
			return;
		}
}

seph_ctx *seph_find(const seppart *atoms)
{
	seph_entry *e = find_entry(atoms);
	return e ? e->ctx : NULL;
}

/* Visits every context, for calls that only see a sepret pointer. */
int seph_each_ctx(int (*fn)(seph_ctx *, void *), void *arg)
{
	for (int i = 0; i < g_nreg; i++)
		if (g_reg[i].ctx) {
			int rc = fn(g_reg[i].ctx, arg);
			if (rc)
				return rc;
		}
	return 0;
}

seph_ctx *seph_ctx_for(seppart *atoms, const sepsys *sys)
{
	if (!atoms || !sys) {
		seph_fail(SEP_GPU_EARG, "null atoms or system pointer");
		return NULL;
	}
	seph_entry *e = find_entry(atoms);
	if (!e) {
		/* an array the library did not allocate: no list capacity known */
		seph_register(atoms, sys->npart, 0);
		e = find_entry(atoms);
	}
	if (e->n != sys->npart) {
		seph_fail(SEP_GPU_EARG, "sys->npart (%ld) differs from the %ld atoms allocated",
			  sys->npart, e->n);
		return NULL;
	}
	if (e->ctx)
		return e->ctx;
	seph_ctx *c = calloc(1, sizeof(seph_ctx));
	int err = 0;
	c->atoms = atoms;
	c->n = e->n;
	c->nneighb = e->nneighb;
	c->dev = sepdev_create(c->n, &err);
	if (!c->dev) {
		seph_dev_fail(err ? err : SEP_GPU_EDEVICE);
		free(c);
		return NULL;
	}
	c->v3 = malloc(sizeof(double) * 3 * (size_t)c->n);
	c->i3 = malloc(sizeof(int) * 40 * (size_t)c->n);
	c->s1 = malloc(sizeof(double) * (size_t)c->n);
	c->s2 = malloc(sizeof(double) * (size_t)c->n);
	c->c1 = malloc((size_t)c->n);
	c->i1 = malloc(sizeof(int) * (size_t)c->n);
	c->ret_scratch = sepdev_ret_alloc(c->dev);
	c->alpha_scratch = sepdev_scalar_alloc(c->dev, 0.0);
	if (c->ret_scratch < 0 || c->alpha_scratch < 0) {
		seph_dev_fail(SEP_GPU_EDEVICE);
		free_ctx(c);
		return NULL;
	}
	e->ctx = c;
	return c;
}

/* ---- marshalling ------------------------------------------------------ */

#define PACK3(field)                                                    \
	for (long i = 0; i < c->n; i++)                                  \
		for (int k = 0; k < 3; k++)                              \
			c->v3[3 * i + k] = c->atoms[i].field[k]
#define UNPACK3(field)                                                  \
	for (long i = 0; i < c->n; i++)                                  \
		for (int k = 0; k < 3; k++)                              \
			c->atoms[i].field[k] = c->v3[3 * i + k]

int seph_put_x(seph_ctx *c)
{
	PACK3(x);
	return seph_dev_fail(sepdev_put_vec(c->dev, SEPDEV_X, c->v3));
}

int seph_put_v(seph_ctx *c)
{
	PACK3(v);
	return seph_dev_fail(sepdev_put_vec(c->dev, SEPDEV_V, c->v3));
}

int seph_put_f(seph_ctx *c)
{
	PACK3(f);
	return seph_dev_fail(sepdev_put_vec(c->dev, SEPDEV_F, c->v3));
}

static int put_a(seph_ctx *c)
{
	PACK3(a);
	return seph_dev_fail(sepdev_put_vec(c->dev, SEPDEV_A, c->v3));
}

int seph_put_xn_cross(seph_ctx *c)
{
	PACK3(xn);
	int rc = seph_dev_fail(sepdev_put_vec(c->dev, SEPDEV_XN, c->v3));
	if (rc)
		return rc;
	for (long i = 0; i < c->n; i++)
		for (int k = 0; k < 3; k++)
			c->i3[3 * i + k] = c->atoms[i].cross_neighb[k];
	rc = seph_dev_fail(sepdev_put_ivec(c->dev, SEPDEV_CROSS_NB, c->i3));
	if (rc)
		return rc;
	for (long i = 0; i < c->n; i++)
		for (int k = 0; k < 3; k++)
			c->i3[3 * i + k] = c->atoms[i].crossings[k];
	return seph_dev_fail(sepdev_put_ivec(c->dev, SEPDEV_CROSS, c->i3));
}

int seph_put_scalars(seph_ctx *c)
{
	c->uniform_type_valid = 0;
	for (long i = 0; i < c->n; i++) {
		c->s1[i] = c->atoms[i].m;
		c->s2[i] = c->atoms[i].z;
		c->c1[i] = (signed char)c->atoms[i].type;
		c->i1[i] = c->atoms[i].molindex;
	}
	return seph_dev_fail(sepdev_put_scalars(c->dev, c->s1, c->s2, c->c1, c->i1));
}

int seph_put_partners(seph_ctx *c)
{
	for (long i = 0; i < c->n; i++) {
		int *p = c->i3 + 40 * i;
		memcpy(p, c->atoms[i].bond, sizeof(int) * SEP_BOND);
		memcpy(p + 10, c->atoms[i].angle, sizeof(int) * SEP_ANGLE);
		memcpy(p + 20, c->atoms[i].dihed, sizeof(int) * SEP_DIHED);
	}
	int rc = seph_dev_fail(sepdev_put_partners(c->dev, c->i3));
	if (rc == 0)
		c->partners_on_device = 1;
	return rc;
}

int seph_put_topology(seph_ctx *c, const sepsys *sys)
{
	const sepmolinfo *m = sys->molptr;
	if (!m)
		return 0;
	const unsigned nb = m->flag_bonds == 1 ? m->num_bonds : 0;
	const unsigned na = m->flag_angles == 1 ? m->num_angles : 0;
	const unsigned nd = m->flag_dihedrals == 1 ? m->num_dihedrals : 0;
	if (c->topo_b == m->blist && c->topo_a == m->alist && c->topo_d == m->dlist &&
	    c->topo_nb == nb && c->topo_na == na && c->topo_nd == nd)
		return 0;
	int rc = seph_dev_fail(sepdev_put_topology(c->dev, m->blist, (int)nb, m->alist, (int)na,
						   m->dlist, (int)nd));
	if (rc)
		return rc;
	c->topo_b = m->blist;
	c->topo_a = m->alist;
	c->topo_d = m->dlist;
	c->topo_nb = nb;
	c->topo_na = na;
	c->topo_nd = nd;
	return 0;
}

int seph_get_f(seph_ctx *c)
{
	int rc = seph_dev_fail(sepdev_get_vec(c->dev, SEPDEV_F, c->v3));
	if (rc)
		return rc;
	UNPACK3(f);
	return 0;
}

int seph_get_state(seph_ctx *c)
{
	int rc = seph_dev_fail(sepdev_get_vec(c->dev, SEPDEV_X, c->v3));
	if (rc)
		return rc;
	UNPACK3(x);
	if ((rc = seph_dev_fail(sepdev_get_vec(c->dev, SEPDEV_V, c->v3))))
		return rc;
	UNPACK3(v);
	if ((rc = seph_dev_fail(sepdev_get_vec(c->dev, SEPDEV_A, c->v3))))
		return rc;
	UNPACK3(a);
	if ((rc = seph_dev_fail(sepdev_get_ivec(c->dev, SEPDEV_CROSS_NB, c->i3))))
		return rc;
	for (long i = 0; i < c->n; i++)
		for (int k = 0; k < 3; k++)
			c->atoms[i].cross_neighb[k] = c->i3[3 * i + k];
	if ((rc = seph_dev_fail(sepdev_get_ivec(c->dev, SEPDEV_CROSS, c->i3))))
		return rc;
	for (long i = 0; i < c->n; i++)
		for (int k = 0; k < 3; k++)
			c->atoms[i].crossings[k] = c->i3[3 * i + k];
	return 0;
}

static int any_partners(const seph_ctx *c)
{
	for (long i = 0; i < c->n; i++)
		if (c->atoms[i].bond[0] != -1 || c->atoms[i].angle[0] != -1 ||
		    c->atoms[i].dihed[0] != -1)
			return 1;
	return 0;
}

static void sys_to_dev(const sepsys *sys, sepdev_sys *ds)
{
	memset(ds, 0, sizeof *ds);
	ds->max_dist2 = sys->max_dist2;
	ds->tnow = sys->tnow;
	ds->neighb_flag = sys->neighb_flag;
	ds->nupdate_neighb = sys->nupdate_neighb;
}

int seph_put_all(seph_ctx *c, const sepsys *sys)
{
	int rc;
	if ((rc = seph_put_x(c)) || (rc = seph_put_v(c)) || (rc = seph_put_f(c)) ||
	    (rc = put_a(c)) || (rc = seph_put_xn_cross(c)) || (rc = seph_put_scalars(c)))
		return rc;
	if (any_partners(c) && (rc = seph_put_partners(c)))
		return rc;
	if ((rc = seph_put_topology(c, sys)))
		return rc;
	sepdev_sys ds;
	sys_to_dev(sys, &ds);
	return seph_dev_fail(sepdev_put_sys(c->dev, &ds));
}

static void ret_pack(const sepret *r, double *w)
{
	memset(w, 0, sizeof(double) * SEPDEV_RET_WORDS);
	w[SEPDEV_RET_EPOT] = r->epot;
	w[SEPDEV_RET_ECOUL] = r->ecoul;
	w[SEPDEV_RET_EKIN] = r->ekin;
	for (int a = 0; a < 3; a++)
		for (int b = 0; b < 3; b++) {
			w[SEPDEV_RET_POTP + 3 * a + b] = r->pot_P[a][b];
			w[SEPDEV_RET_KINP + 3 * a + b] = r->kin_P[a][b];
			w[SEPDEV_RET_BONDP + 3 * a + b] = r->pot_P_bond[a][b];
		}
}

void seph_ret_unpack(sepret *r, const double *w)
{
	r->epot = w[SEPDEV_RET_EPOT];
	r->ecoul = w[SEPDEV_RET_ECOUL];
	r->ekin = w[SEPDEV_RET_EKIN];
	for (int a = 0; a < 3; a++)
		for (int b = 0; b < 3; b++) {
			r->pot_P[a][b] = w[SEPDEV_RET_POTP + 3 * a + b];
			r->kin_P[a][b] = w[SEPDEV_RET_KINP + 3 * a + b];
			r->pot_P_bond[a][b] = w[SEPDEV_RET_BONDP + 3 * a + b];
		}
}

/* The per-atom state of a resident system back into the AoS; ds gets the
 * device's system scalars. */
static int get_atoms(seph_ctx *c, sepdev_sys *ds)
{
	int rc;
	if ((rc = seph_get_state(c)) || (rc = seph_get_f(c)))
		return rc;
	if ((rc = seph_dev_fail(sepdev_get_vec(c->dev, SEPDEV_XN, c->v3))))
		return rc;
	UNPACK3(xn);
	if ((rc = seph_dev_fail(sepdev_get_sys(c->dev, ds))))
		return rc;
	if (ds->xn_pending) {
		/* the reference resets xn and cross_neighb right after the
		 * leapfrog that raised the flag (source/sepintgr.c:72-84) */
		for (long i = 0; i < c->n; i++)
			for (int k = 0; k < 3; k++) {
				c->atoms[i].xn[k] = c->atoms[i].x[k];
				c->atoms[i].cross_neighb[k] = 0;
			}
	}
	return 0;
}

int seph_get_atoms(seph_ctx *c)
{
	sepdev_sys ds;
	return get_atoms(c, &ds);
}

int seph_put_atoms(seph_ctx *c)
{
	int rc;
	if ((rc = seph_put_x(c)) || (rc = seph_put_v(c)) || (rc = seph_put_f(c)) ||
	    (rc = put_a(c)) || (rc = seph_put_xn_cross(c)) || (rc = seph_put_scalars(c)))
		return rc;
	return 0;
}

int seph_get_all(seph_ctx *c, sepsys *sys)
{
	int rc;
	sepdev_sys ds;
	if ((rc = get_atoms(c, &ds)))
		return rc;
	sys->max_dist2 = ds.max_dist2;
	sys->tnow = ds.tnow;
	sys->neighb_flag = ds.neighb_flag;
	sys->nupdate_neighb = ds.nupdate_neighb;
	for (int b = 0; b < c->nrets; b++) {
		double w[SEPDEV_RET_WORDS];
		if ((rc = seph_dev_fail(sepdev_ret_get(c->dev, c->rets[b].slot, w))))
			return rc;
		seph_ret_unpack((sepret *)c->rets[b].host, w);
	}
	for (int b = 0; b < c->nalphas; b++)
		if ((rc = seph_dev_fail(sepdev_scalar_get(c->dev, c->alphas[b].slot,
							  (double *)c->alphas[b].host))))
			return rc;
	if (sys->molptr && c->topo_nb + c->topo_na + c->topo_nd > 0) {
		sepmolinfo *m = sys->molptr;
		rc = seph_dev_fail(sepdev_get_bonded_values(
			c->dev, c->topo_nb ? m->blengths : NULL, c->topo_na ? m->angles : NULL,
			c->topo_nd ? m->dihedrals : NULL));
		if (rc)
			return rc;
	}
	return seph_check_device_error(c);
}

int seph_check_device_error(seph_ctx *c)
{
	sepdev_sys ds;
	int rc = seph_dev_fail(sepdev_get_sys(c->dev, &ds));
	if (rc)
		return rc;
	if (ds.error) {
		const int code = ds.error;
		ds.error = 0;
		sepdev_put_sys(c->dev, &ds);
		if (code == SEP_GPU_ENEIGHB)
			return seph_fail(code, "neighbour list capacity exceeded on the device (largest list %d)",
					 ds.max_full);
		if (code == SEP_GPU_ECELL)
			return seph_fail(code, "an atom position maps outside the cell grid (x < 0, x >= L or not finite)");
		return seph_fail(code, "device error %d", code);
	}
	return 0;
}

/* ---- slots ------------------------------------------------------------ */

int seph_ret_slot(seph_ctx *c, sepret *ret)
{
	for (int b = 0; b < c->nrets; b++)
		if (c->rets[b].host == ret)
			return c->rets[b].slot;
	if (c->nrets == SEPH_MAX_BIND) {
		seph_fail(SEP_GPU_EARG, "too many sepret structures bound to one system");
		return -1;
	}
	const int slot = sepdev_ret_alloc(c->dev);
	if (slot < 0) {
		seph_dev_fail(SEP_GPU_EDEVICE);
		return -1;
	}
	double w[SEPDEV_RET_WORDS];
	ret_pack(ret, w);
	if (seph_dev_fail(sepdev_ret_put(c->dev, slot, w)))
		return -1;
	c->rets[c->nrets].host = ret;
	c->rets[c->nrets].slot = slot;
	c->nrets++;
	return slot;
}

int seph_alpha_slot(seph_ctx *c, double *alpha)
{
	for (int b = 0; b < c->nalphas; b++)
		if (c->alphas[b].host == alpha)
			return c->alphas[b].slot;
	if (c->nalphas == SEPH_MAX_BIND) {
		seph_fail(SEP_GPU_EARG, "too many thermostat states bound to one system");
		return -1;
	}
	const int slot = sepdev_scalar_alloc(c->dev, *alpha);
	if (slot < 0) {
		seph_dev_fail(SEP_GPU_EDEVICE);
		return -1;
	}
	c->alphas[c->nalphas].host = alpha;
	c->alphas[c->nalphas].slot = slot;
	c->nalphas++;
	return slot;
}

void seph_ret_merge(sepret *ret, const double *w, int epot_assign, int bond_stress)
{
	if (epot_assign)
		ret->epot = w[SEPDEV_RET_EPOT];
	else
		ret->epot += w[SEPDEV_RET_EPOT];
	ret->ecoul += w[SEPDEV_RET_ECOUL];
	ret->ekin += w[SEPDEV_RET_EKIN];
	for (int a = 0; a < 3; a++)
		for (int b = 0; b < 3; b++) {
			ret->pot_P[a][b] += w[SEPDEV_RET_POTP + 3 * a + b];
			ret->kin_P[a][b] += w[SEPDEV_RET_KINP + 3 * a + b];
			if (bond_stress)
				ret->pot_P_bond[a][b] += w[SEPDEV_RET_BONDP + 3 * a + b];
		}
}

/* ---- grids ------------------------------------------------------------ */

/* The reference's 14-offset half-shell stencil (source/sepprfrc.c:424-426),
 * one (x, y, z) row per offset, in the reference's order. */
static const int HALF_SHELL[14][3] = {
	{0, 0, 0},  {1, 0, 0},  {1, 1, 0},  {0, 1, 0},  {-1, 1, 0},
	{0, 0, 1},  {1, 0, 1},  {1, 1, 1},  {0, 1, 1},  {-1, 1, 1},
	{-1, 0, 1}, {-1, -1, 1}, {0, -1, 1}, {1, -1, 1}};

void seph_grid_reference(const sepsys *sys, unsigned opt, sepdev_grid *g)
{
	memset(g, 0, sizeof *g);
	for (int k = 0; k < 3; k++) {
		g->n[k] = sys->nsubbox[k];
		g->l[k] = sys->lsubbox[k];
		g->L[k] = sys->length[k];
	}
	g->ref_index = 1;
	const double rl = sys->cf + sys->skin;
	g->rlist2 = rl * rl;
	g->opt = (int)opt;
	g->bonded_all = 1;
	int o = 0;
	if (g->n[0] >= 3 && g->n[1] >= 3 && g->n[2] >= 3) {
		/* No half-shell offset wraps onto its own cell and every
		 * (offset, -offset) pair reaches two distinct cells, so the
		 * reference's pairs seen from both ends are exactly the 27-cell
		 * stencil with j != i in the home cell.  Visit it in ascending
		 * (z, y, x) order: lists then come out in cell order. */
		for (int dz = -1; dz <= 1; dz++)
			for (int dy = -1; dy <= 1; dy++)
				for (int dx = -1; dx <= 1; dx++) {
					g->off[o][0] = dx;
					g->off[o][1] = dy;
					g->off[o][2] = dz;
					g->off[o++][3] = SEPDEV_RULE_NE;
				}
		g->noff = o;
		g->shift_ok = g->n[0] >= 5 && g->n[1] >= 5 && g->n[2] >= 5;
		return;
	}
	/* Fewer than three cells on some axis: offsets can alias.  Offset 0 in
	 * both directions = every other atom of the home cell; each further
	 * half-shell offset forward (j > i when it wraps onto the home cell)
	 * and backward (j < i), reproducing the reference's duplicates. */
	g->off[o][0] = g->off[o][1] = g->off[o][2] = 0;
	g->off[o++][3] = SEPDEV_RULE_NE;
	for (int s = 1; s < 14; s++) {
		for (int k = 0; k < 3; k++)
			g->off[o][k] = HALF_SHELL[s][k];
		g->off[o++][3] = SEPDEV_RULE_GT;
		for (int k = 0; k < 3; k++)
			g->off[o][k] = -HALF_SHELL[s][k];
		g->off[o++][3] = SEPDEV_RULE_LT;
	}
	g->noff = o;
}

void seph_grid_own(const sepsys *sys, double rlist, unsigned opt, int bonded_rule,
		   sepdev_grid *g)
{
	memset(g, 0, sizeof *g);
	int span[3];
	for (int k = 0; k < 3; k++) {
		int n = (int)(sys->length[k] / (rlist * (1.0 + 1e-9)));
		if (n < 3)
			n = 1;
		g->n[k] = n;
		g->l[k] = sys->length[k] / n;
		g->L[k] = sys->length[k];
		span[k] = n >= 3 ? 1 : 0;
	}
	g->ref_index = 0;
	g->rlist2 = rlist * rlist;
	g->opt = (int)opt;
	g->bonded_all = bonded_rule;
	g->shift_ok = g->n[0] >= 5 && g->n[1] >= 5 && g->n[2] >= 5;
	int o = 0;
	for (int dz = -span[2]; dz <= span[2]; dz++)
		for (int dy = -span[1]; dy <= span[1]; dy++)
			for (int dx = -span[0]; dx <= span[0]; dx++) {
				g->off[o][0] = dx;
				g->off[o][1] = dy;
				g->off[o][2] = dz;
				g->off[o++][3] = SEPDEV_RULE_NE;
			}
	g->noff = o;
}

/* ---- list build ------------------------------------------------------- */

/* Writes the device list into atoms[i].neighb as a -1 terminated half list
 * (entries j > i, original numbering). */
static int write_half_lists(seph_ctx *c)
{
	if (c->nneighb <= 0)
		return 0;
	const int stride = sepdev_list_capacity(c->dev);
	int *counts = malloc(sizeof(int) * (size_t)c->n);
	int *half = malloc(sizeof(int) * (size_t)c->n * (size_t)stride);
	int maxc = 0;
	int rc = seph_dev_fail(sepdev_get_half_list(c->dev, counts, half, stride, &maxc));
	if (rc == 0 && maxc >= c->nneighb)
		rc = seph_fail(SEP_GPU_ENEIGHB, "Too many neighbours (%d >= capacity %d)", maxc,
			       c->nneighb);
	if (rc == 0)
		for (long i = 0; i < c->n; i++) {
			memcpy(c->atoms[i].neighb, half + (size_t)i * stride,
			       sizeof(int) * (size_t)counts[i]);
			c->atoms[i].neighb[counts[i]] = -1;
		}
	free(counts);
	free(half);
	return rc;
}

/* After a build that overflowed the list (SEP_GPU_ENEIGHB, longest list
 * max_full): grow the list and clear the error, or, where the longest list
 * exceeds the bound set by sep_gpu_list_limit, report the overflow. */
int seph_grow_list(seph_ctx *c, sepdev_sys *ds)
{
	if (c->list_limit > 0 && ds->max_full > c->list_limit) {
		ds->error = 0;
		sepdev_put_sys(c->dev, ds);
		return seph_fail(SEP_GPU_ENEIGHB,
				 "Too many neighbours: a list of %d entries exceeds the limit of %d",
				 ds->max_full, c->list_limit);
	}
	ds->error = 0;
	int rc = seph_dev_fail(sepdev_put_sys(c->dev, ds));
	if (!rc)
		rc = seph_dev_fail(sepdev_reserve_list(c->dev, ds->max_full + ds->max_full / 4 + 32));
	return rc;
}

int seph_build(seph_ctx *c, sepsys *sys, const sepdev_grid *g, int conditional)
{
	(void)sys;
	int rc = seph_dev_fail(sepdev_build_list(c->dev, g, conditional));
	if (rc || conditional)
		return rc;
	/* synchronised: check overflow and grow once */
	sepdev_sys ds;
	if ((rc = seph_dev_fail(sepdev_get_sys(c->dev, &ds))))
		return rc;
	if (ds.error == SEP_GPU_ENEIGHB) {
		if ((rc = seph_grow_list(c, &ds)))
			return rc;
		if ((rc = seph_dev_fail(sepdev_build_list(c->dev, g, 0))))
			return rc;
	}
	if ((rc = seph_check_device_error(c)))
		return rc;
	return write_half_lists(c);
}

static int grid_checked(const seph_ctx *c, const sepdev_grid *g)
{
	for (int i = 0; i < c->nchecked && i < SEPH_CHECKED; i++) {
		const seph_grid_key *k = &c->checked[i];
		if (k->n[0] == g->n[0] && k->n[1] == g->n[1] && k->n[2] == g->n[2] &&
		    k->opt == g->opt && k->rlist2 == g->rlist2)
			return 1;
	}
	return 0;
}

int seph_build_resident(seph_ctx *c, const sepdev_grid *g, int conditional)
{
	int rc = seph_dev_fail(sepdev_build_list(c->dev, g, conditional));
	/* a capture cannot hold the host check below (a stream
	 * synchronisation); an overflow there freezes the run, and sep_gpu_run
	 * reports it after the replay */
	if (rc || grid_checked(c, g) || sepdev_capturing(c->dev))
		return rc;
	/* first build of this grid: the initial capacity comes from the mean
	 * density, which grids of fewer than 3 cells per axis (whose stencil
	 * lists pairs more than once) or dense clusters can exceed; confirm
	 * it here, once, instead of letting SEP_GPU_ENEIGHB surface later */
	sepdev_sys ds;
	if ((rc = seph_dev_fail(sepdev_get_sys(c->dev, &ds))))
		return rc;
	if (ds.error == SEP_GPU_ENEIGHB) {
		if ((rc = seph_grow_list(c, &ds)) ||
		    (rc = seph_dev_fail(sepdev_build_list(c->dev, g, 0))))
			return rc;
	}
	if ((rc = seph_check_device_error(c)))
		return rc;
	seph_grid_key *k = &c->checked[c->nchecked % SEPH_CHECKED];
	for (int a = 0; a < 3; a++)
		k->n[a] = g->n[a];
	k->opt = g->opt;
	k->rlist2 = g->rlist2;
	c->nchecked++;
	return 0;
}

int seph_identify_potential(double (*fun)(double, char))
{
	if (fun == sep_lj)
		return SEPH_POT_LJ;
	if (fun == sep_lj_shift)
		return SEPH_POT_LJ_SHIFT;
	if (fun == sep_wca)
		return SEPH_POT_WCA;
	return SEPH_POT_NONE;
}
