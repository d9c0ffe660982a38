/*
 * seporacle.c -- CPU oracle for the seplib per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY: a serial FP64 restatement of the reference
 * algorithm.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.
 *
 * Built with -O2 -ffp-contract=off so that every product and sum is rounded
 * on its own, in the operation order of the reference source.
 *
 * How each part is pinned independently of this file (tests/
 * test_golden_oracle.py against vectors written by the compiled reference,
 * test_reference_live.py against that reference run live,
 * test_oracle_math.py against the mathematics); every function below is
 * covered, none is "parity unpinned":
 *   neighbour lists     -> golden pair-set hashes of 1000 steps (bitwise),
 *                          O(N^2) enumeration with the same r^2 op order
 *   LJ / WCA forces     -> golden vectors (bitwise), two-atom closed forms
 *                          (F(2^(1/6)) = 0, U(rc) = 0, the sigma^2 factor of
 *                          sep_force_lj), O(N^2) numpy sums, Newton's third law
 *   shifted-force Coulomb -> golden vectors (bitwise), two-charge closed
 *                          form, F(rc) = U(rc) = 0, O(N^2) numpy sum
 *   bond / cos^2 angle / RB torsion -> golden vectors (bitwise), central
 *                          finite differences of their own energy
 *   harmonic angle      -> golden vectors (bitwise); F = sin(theta) (-grad U)
 *                          against central differences
 *   Nose-Hoover, leapfrog -> golden step sequences (bitwise), free-particle
 *                          recurrences, wrap/crossing cases, skin/2 trigger
 *   topology reader     -> the reference's own .top inputs (golden vectors)
 *   Model II terms (orc_omp_*) -> the compiled reference's sep_omp_* run
 *                          live (bitwise), and the same forces as the
 *                          sep_* term of each (the cos^2 form for the angle)
 */
#include "seporacle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* allocation                                                          */

/* Defaults of source/sepinit.c:33-51; everything else zero. */
seppart *orc_alloc(long npart, int nneighb)
{
	seppart *p = calloc((size_t)npart, sizeof(seppart));
	if (!p)
		return NULL;
	for (long i = 0; i < npart; i++) {
		p[i].neighb = malloc((size_t)(nneighb > 0 ? nneighb : 1) * sizeof(int));
		if (!p[i].neighb)
			return NULL;
		for (int k = 0; k < nneighb; k++)
			p[i].neighb[k] = -1;
		p[i].type = 'A';
		p[i].m = 1.0;
		p[i].molindex = -1;
		p[i].ldiff = 1.0;
		for (int k = 0; k < SEP_BOND; k++) p[i].bond[k] = -1;
		for (int k = 0; k < SEP_ANGLE; k++) p[i].angle[k] = -1;
		for (int k = 0; k < SEP_DIHED; k++) p[i].dihed[k] = -1;
	}
	return p;
}

void orc_free(seppart *p, long npart)
{
	if (!p)
		return;
	for (long i = 0; i < npart; i++)
		free(p[i].neighb);
	free(p);
}

/* source/sepinit.c:246-305 with source/sepmisc.c:454-462 (cell count). */
sepsys orc_sys_setup(double lx, double ly, double lz, double cf, double dt,
		     long npart, unsigned update)
{
	sepsys s;
	memset(&s, 0, sizeof s);
	const double skin0 = 0.25;
	const double L[3] = {lx, ly, lz};
	for (int k = 0; k < 3; k++) {
		s.length[k] = L[k];
		s.nsubbox[k] = (int)(L[k] / (cf + skin0));
		s.lsubbox[k] = L[k] / s.nsubbox[k];
	}
	s.volume = lx * ly * lz;
	s.npart = npart;
	s.cf = cf;
	s.dt = dt;
	s.ndof = (unsigned)(3 * npart - 3);
	s.neighb_update = update;
	s.neighb_flag = 1;
	s.skin = skin0;
	s.molptr = calloc(1, sizeof(sepmolinfo));
	return s;
}

void orc_free_sys(sepsys *s)
{
	sepmolinfo *m = s->molptr;
	if (!m)
		return;
	free(m->blist);
	free(m->alist);
	free(m->dlist);
	free(m->blengths);
	free(m->angles);
	free(m->dihedrals);
	free(m);
	s->molptr = NULL;
}

/* ------------------------------------------------------------------ */
/* potentials: source/sepmisc.c:115-164                                */

double orc_lj(double r2, char opt)
{
	double inv = 1.0 / r2, inv3 = inv * inv * inv;
	if (opt == 'f')
		return 48. * inv3 * (inv3 - 0.5) * inv;
	if (opt == 'u')
		return 4.0 * inv3 * (inv3 - 1.0);
	return 0.0;
}

double orc_lj_shift(double r2, char opt)
{
	double inv = 1.0 / r2, inv3 = inv * inv * inv;
	if (opt == 'f')
		return 48.0 * inv3 * (inv3 - 0.5) * inv;
	if (opt == 'u')
		return 4.0 * inv3 * (inv3 - 1.0) + SEP_LJCF2;
	return 0.0;
}

double orc_wca(double r2, char opt)
{
	double inv = 1.0 / r2, inv3 = inv * inv * inv;
	if (opt == 'f')
		return 48.0 * inv3 * (inv3 - 0.5) * inv;
	if (opt == 'u')
		return 4.0 * inv3 * (inv3 - 1.0) + 1.0;
	return 0.0;
}

/* ------------------------------------------------------------------ */
/* return values: source/sepret.c:19-82, source/sepmisc.c:393-400      */

void orc_reset_retval(sepret *r)
{
	/* every field the reference zeroes is a scalar energy or a 3x3
	 * tensor; etot, p, p_mol and t_mol are left alone as in
	 * source/sepret.c:19-47 */
	double keep_etot = r->etot, keep_p = r->p, keep_pmol = r->p_mol,
	       keep_tmol = r->t_mol;
	memset(r, 0, sizeof *r);
	r->etot = keep_etot;
	r->p = keep_p;
	r->p_mol = keep_pmol;
	r->t_mol = keep_tmol;
}

void orc_pressure_tensor(sepret *r, const sepsys *s)
{
	const double ivol = 1.0 / s->volume;
	for (int a = 0; a < 3; a++)
		for (int b = 0; b < 3; b++)
			r->P[a][b] = (r->kin_P[a][b] + r->pot_P[a][b]) * ivol;
	double tr = 0.0;
	for (int a = 0; a < 3; a++)
		tr += r->P[a][a];
	r->p = tr / 3.0;
}

void orc_reset_force(seppart *p, sepsys *s)
{
	for (long i = 0; i < s->npart; i++)
		p[i].f[0] = p[i].f[1] = p[i].f[2] = 0.0;
	s->max_dist2 = 0.0;
}

/* ------------------------------------------------------------------ */
/* geometry                                                            */

/* Single-image minimum image of include/sepmisc.h:81-85. */
static inline double min_image(double d, double L)
{
	if (d > 0.5 * L)
		d -= L;
	else if (d < -0.5 * L)
		d += L;
	return d;
}

/* Separation x_a - x_b and its square, summed in axis order as every
 * reference pair loop does (e.g. source/sepprfrc.c:486-490). */
static inline double separation(const seppart *p, long a, long b,
				const double *L, double d[3])
{
	double r2 = 0.0;
	for (int k = 0; k < 3; k++) {
		d[k] = min_image(p[a].x[k] - p[b].x[k], L[k]);
		r2 += d[k] * d[k];
	}
	return r2;
}

static inline double dot3(const double *a, const double *b)
{
	double s = 0.0; /* source/seputil.c:393-403 */
	for (int k = 0; k < 3; k++)
		s += a[k] * b[k];
	return s;
}

/* ------------------------------------------------------------------ */
/* exclusion predicates: source/sepprfrc.c:703-740                     */

/* The partner scan stops at the first -1 in EITHER atom's array. */
static int share_in(const int *pa, const int *pb, int cap, long a, long b)
{
	for (int k = 0; k < cap; k++) {
		if (pa[k] == -1 || pb[k] == -1)
			return 0;
		if (pa[k] == b || pb[k] == a)
			return 1;
	}
	return 0;
}

static int bond_share(const seppart *p, long a, long b)
{
	return share_in(p[a].bond, p[b].bond, SEP_BOND, a, b);
}

static int bonded_count(const seppart *p, long a, long b)
{
	return share_in(p[a].bond, p[b].bond, SEP_BOND, a, b) +
	       share_in(p[a].angle, p[b].angle, SEP_ANGLE, a, b) +
	       share_in(p[a].dihed, p[b].dihed, SEP_DIHED, a, b);
}

/* ------------------------------------------------------------------ */
/* neighbour lists                                                     */

/* Half-shell stencil of source/sepprfrc.c:424-426, one (x, y, z) row per
 * offset, in the reference's order. */
static const int HALF_SHELL[14][3] = {
	{0, 0, 0},  {1, 0, 0},  {1, 1, 0},  {0, 1, 0},  {-1, 1, 0},
	{0, 0, 1},  {1, 0, 1},  {1, 1, 1},  {0, 1, 1},  {-1, 1, 1},
	{-1, 0, 1}, {-1, -1, 1}, {0, -1, 1}, {1, -1, 1}};

static inline int wrap_cell(int c, int n)
{
	return ((c % n) + n) % n;
}

/* Linked cell list of source/sepprfrc.c:394-415: head[c] is the last atom
 * inserted into cell c, next[i] the one inserted before atom i. */
static int make_cells(const seppart *p, const sepsys *s, int *head, int *next)
{
	const int nx = s->nsubbox[0], ny = s->nsubbox[1], nz = s->nsubbox[2];
	const long ncell = (long)nx * ny * nz;
	for (long c = 0; c < ncell; c++)
		head[c] = -1;
	for (long i = 0; i < s->npart; i++) {
		long c = (long)(int)(p[i].x[0] / s->lsubbox[0]) +
			 (long)(int)(p[i].x[1] / s->lsubbox[1]) * nx +
			 (long)(int)(p[i].x[2] / s->lsubbox[2]) * nx * ny;
		if (c < 0 || c >= ncell)
			return -1;
		next[i] = head[c];
		head[c] = (int)i;
	}
	return 0;
}

static int pair_allowed(const seppart *p, long a, long b, unsigned opt,
			int bonded_rule)
{
	if (opt == SEP_EXCL_SAME_MOL)
		return p[a].molindex == -1 || p[a].molindex != p[b].molindex;
	if (opt == SEP_EXCL_BONDED)
		return bonded_rule ? bonded_count(p, a, b) == 0
				   : bond_share(p, a, b) == 0;
	return 1;
}

/* source/sepprfrc.c:419-513 (SEP_ALL), :517-603 (SEP_EXCL_BONDED, which
 * uses bonds, angles and dihedrals), :606-700 (SEP_EXCL_SAME_MOL). */
static int list_from_cells(seppart *p, sepsys *s, unsigned opt, int nneighb)
{
	const int n[3] = {s->nsubbox[0], s->nsubbox[1], s->nsubbox[2]};
	const long ncell = (long)n[0] * n[1] * n[2];
	const double rl = s->cf + s->skin, rl2 = rl * rl;
	int *head = malloc(sizeof(int) * (size_t)ncell);
	int *next = malloc(sizeof(int) * (size_t)s->npart);
	int *len = calloc((size_t)s->npart, sizeof(int));
	int rc = 0;

	for (long i = 0; i < s->npart; i++)
		for (int k = 0; k < nneighb; k++)
			p[i].neighb[k] = -1;

	if (make_cells(p, s, head, next) != 0) {
		rc = -1;
		goto done;
	}

	for (int cz = 0; cz < n[2]; cz++)
	for (int cy = 0; cy < n[1]; cy++)
	for (int cx = 0; cx < n[0]; cx++) {
		const long home = cx + (long)cy * n[0] + (long)cz * n[0] * n[1];
		for (int a = head[home]; a != -1; a = next[a]) {
			for (int o = 0; o < 14; o++) {
				const int tx = wrap_cell(cx + HALF_SHELL[o][0], n[0]);
				const int ty = wrap_cell(cy + HALF_SHELL[o][1], n[1]);
				const int tz = wrap_cell(cz + HALF_SHELL[o][2], n[2]);
				const long other = tx + (long)ty * n[0] + (long)tz * n[0] * n[1];
				for (int b = head[other]; b != -1; b = next[b]) {
					if (home == other && b <= a)
						continue;
					if (opt == SEP_EXCL_SAME_MOL &&
					    !pair_allowed(p, a, b, opt, 1))
						continue;
					double d[3];
					const double r2 = separation(p, a, b, s->length, d);
					int take = r2 < rl2;
					if (take && opt == SEP_EXCL_BONDED)
						take = pair_allowed(p, a, b, opt, 1);
					if (take)
						p[a].neighb[len[a]++] = b;
					if (len[a] == nneighb) {
						rc = -1;
						goto done;
					}
				}
			}
		}
	}
done:
	free(head);
	free(next);
	free(len);
	return rc;
}

/* O(N^2) half list of source/sepprfrc.c:306-343; its SEP_EXCL_BONDED test
 * looks at bonds only. */
static int list_all_pairs(seppart *p, sepsys *s, unsigned opt, int nneighb)
{
	const double rl = s->cf + s->skin, rl2 = rl * rl;
	for (long i = 0; i < s->npart; i++)
		for (int k = 0; k < nneighb; k++)
			p[i].neighb[k] = -1;
	for (long a = 0; a < s->npart - 1; a++) {
		int len = 0;
		for (long b = a + 1; b < s->npart; b++) {
			if (!pair_allowed(p, a, b, opt, 0))
				continue;
			double d[3];
			if (separation(p, a, b, s->length, d) < rl2) {
				if (len + 1 >= nneighb)
					return -1;
				p[a].neighb[len++] = (int)b;
			}
		}
	}
	return 0;
}

int orc_build_list(seppart *p, sepsys *s, unsigned opt, int nneighb)
{
	if (s->neighb_update == SEP_NEIGHBLIST)
		return list_all_pairs(p, s, opt, nneighb);
	return list_from_cells(p, s, opt, nneighb);
}

/* Unordered pairs (min, max) of the half lists p[i].neighb, written as
 * 2 ints per pair to out (room for maxpairs); returns the pair count, or -1
 * if out is too small.  Reads any AoS with the seppart layout. */
long orc_half_pairs(const seppart *p, long npart, int *out, long maxpairs)
{
	long c = 0;
	for (long i = 0; i < npart; i++) {
		for (int k = 0; p[i].neighb[k] != -1; k++) {
			const int j = p[i].neighb[k];
			if (c >= maxpairs)
				return -1;
			out[2 * c] = (int)(i < j ? i : j);
			out[2 * c + 1] = (int)(i < j ? j : i);
			c++;
		}
	}
	return c;
}

int orc_list_len(const seppart *p, long i)
{
	int n = 0;
	while (p[i].neighb[n] != -1)
		n++;
	return n;
}

/* ------------------------------------------------------------------ */
/* pair forces                                                         */

static int types_match(const seppart *p, long a, long b, const char *t)
{
	return (p[a].type == t[0] && p[b].type == t[1]) ||
	       (p[a].type == t[1] && p[b].type == t[0]);
}

static void add_stress(sepret *r, const double *f, const double *d)
{
	for (int a = 0; a < 3; a++)
		for (int b = 0; b < 3; b++)
			r->pot_P[a][b] += f[a] * d[b];
}

/* Rebuild step shared by source/sepprfrc.c:249-267 and :757-774. */
static int rebuild_if_flagged(seppart *p, sepsys *s, unsigned opt, int nneighb)
{
	if (s->neighb_flag != 1)
		return 0;
	int rc = orc_build_list(p, s, opt, nneighb);
	s->neighb_flag = 0;
	return rc;
}

/* source/sepprfrc.c:19-91 */
static void pairs_brute(seppart *p, const char *types, double cf,
			orc_pot_fn fun, sepsys *s, sepret *r, unsigned opt)
{
	const double cf2 = cf * cf;
	for (long a = 0; a < s->npart - 1; a++) {
		for (long b = a + 1; b < s->npart; b++) {
			if (opt == SEP_EXCL_BONDED && bond_share(p, a, b) == 1)
				continue;
			if (opt == SEP_EXCL_SAME_MOL && p[a].molindex == p[b].molindex &&
			    p[a].molindex != -1)
				continue;
			if (!types_match(p, a, b, types))
				continue;
			double d[3], f[3];
			const double r2 = separation(p, a, b, s->length, d);
			if (r2 >= cf2)
				continue;
			const double ft = fun(r2, 'f');
			for (int k = 0; k < 3; k++) {
				f[k] = ft * d[k];
				p[a].f[k] += f[k];
				p[b].f[k] -= f[k];
			}
			r->epot += fun(r2, 'u');
			add_stress(r, f, d);
		}
	}
}

/* Serial list branch of source/sepprfrc.c:94-224: forces go through a
 * zeroed 3N scratch array that is added to f at the end, and the pair
 * energy is ASSIGNED to r->epot (:222). */
static void pairs_list(seppart *p, const char *types, double cf,
		       orc_pot_fn fun, sepsys *s, sepret *r)
{
	const double cf2 = cf * cf;
	const long n = s->npart;
	double *acc = calloc((size_t)(3 * n), sizeof(double));
	double epot = 0.0;
	for (long a = 0; a < n; a++) {
		if (p[a].type != types[0] && p[a].type != types[1])
			continue;
		for (int e = 0; p[a].neighb[e] != -1; e++) {
			const long b = p[a].neighb[e];
			if (!types_match(p, a, b, types))
				continue;
			double d[3], f[3];
			const double r2 = separation(p, a, b, s->length, d);
			if (r2 >= cf2)
				continue;
			const double ft = fun(r2, 'f');
			for (int k = 0; k < 3; k++) {
				f[k] = ft * d[k];
				acc[3 * a + k] += f[k];
				acc[3 * b + k] += -f[k];
			}
			epot += fun(r2, 'u');
			add_stress(r, f, d);
		}
	}
	for (long a = 0; a < n; a++)
		for (int k = 0; k < 3; k++)
			p[a].f[k] += acc[3 * a + k];
	free(acc);
	r->epot = epot;
}

/* source/sepprfrc.c:226-274 (shipped -DOCTAVE build: an oversized cutoff
 * warns and returns). */
int orc_force_pairs(seppart *p, const char *types, double cf, orc_pot_fn fun,
		    sepsys *s, sepret *r, unsigned opt, int nneighb)
{
	if (cf > s->cf)
		return -1;
	if (s->neighb_update == SEP_BRUTE) {
		s->omp_flag = false;
		pairs_brute(p, types, cf, fun, s, r, opt);
		return 0;
	}
	if (rebuild_if_flagged(p, s, opt, nneighb) != 0)
		return -1;
	pairs_list(p, types, cf, fun, s, r);
	return 0;
}

/* Parameter block of source/sepprfrc.c:785-799: p = {rc, eps, sigma, aw}. */
typedef struct {
	double cf2, shift, eps48, eps4, awh, aw, sig2;
} lj_consts;

static lj_consts lj_setup(const double *prm)
{
	lj_consts c;
	const double cf = prm[0], eps = prm[1], sigma = prm[2], aw = prm[3];
	c.shift = 4.0 * eps * (pow(sigma / cf, 12.) - aw * pow(sigma / cf, 6.));
	c.cf2 = cf * cf;
	c.eps48 = 48.0 * eps;
	c.eps4 = 4.0 * eps;
	c.awh = 0.5 * aw;
	c.aw = aw;
	c.sig2 = sigma * sigma;
	return c;
}

/* Serial branch of source/sepprfrc.c:782-923: forces straight into f and
 * the energy ADDED to r->epot (:922). */
static void lj_list(seppart *p, const char *types, const lj_consts *c,
		    sepsys *s, sepret *r)
{
	double epot = 0.0;
	for (long a = 0; a < s->npart; a++) {
		if (p[a].type != types[0] && p[a].type != types[1])
			continue;
		for (int e = 0; p[a].neighb[e] != -1; e++) {
			const long b = p[a].neighb[e];
			if (!types_match(p, a, b, types))
				continue;
			double d[3], f[3];
			const double r2 = separation(p, a, b, s->length, d);
			if (r2 >= c->cf2)
				continue;
			const double q = c->sig2 / r2, q3 = q * q * q;
			const double ft = c->eps48 * q3 * (q3 - c->awh) * q;
			for (int k = 0; k < 3; k++) {
				f[k] = ft * d[k];
				p[a].f[k] += f[k];
				p[b].f[k] += -f[k];
			}
			epot += c->eps4 * q3 * (q3 - c->aw) - c->shift;
			add_stress(r, f, d);
		}
	}
	r->epot += epot;
}

/* source/sepprfrc.c:925-1000 */
static void lj_brute(seppart *p, const char *types, const double *prm,
		     const lj_consts *c, sepsys *s, sepret *r, unsigned opt)
{
	const double eps = prm[1], sigma = prm[2], aw = prm[3];
	for (long a = 0; a < s->npart - 1; a++) {
		for (long b = a + 1; b < s->npart; b++) {
			if (opt == SEP_EXCL_BONDED && bond_share(p, a, b) == 1)
				continue;
			if (opt == SEP_EXCL_SAME_MOL && p[a].molindex == p[b].molindex &&
			    p[a].molindex != -1)
				continue;
			if (!types_match(p, a, b, types))
				continue;
			double d[3], f[3];
			const double r2 = separation(p, a, b, s->length, d);
			if (r2 >= c->cf2)
				continue;
			const double q = sigma * sigma / r2, q3 = q * q * q;
			const double ft = 48.0 * eps * q3 * (q3 - 0.5 * aw) * q;
			for (int k = 0; k < 3; k++) {
				f[k] = ft * d[k];
				p[a].f[k] += f[k];
				p[b].f[k] -= f[k];
			}
			r->epot += 4.0 * eps * q3 * (q3 - aw) - c->shift;
			add_stress(r, f, d);
		}
	}
}

/* source/sepprfrc.c:743-780.  The shipped -DOCTAVE build only prints when
 * the cutoff is above sys->cf and carries on; so does the oracle. */
int orc_force_lj(seppart *p, const char *types, const double *prm,
		 sepsys *s, sepret *r, unsigned opt, int nneighb)
{
	const lj_consts c = lj_setup(prm);
	int rc = prm[0] > s->cf ? -1 : 0;
	if (s->neighb_update == SEP_BRUTE) {
		s->omp_flag = false;
		lj_brute(p, types, prm, &c, s, r, opt);
		return rc;
	}
	if (rebuild_if_flagged(p, s, opt, nneighb) != 0)
		return -1;
	lj_list(p, types, &c, s, r);
	return rc;
}

/* ------------------------------------------------------------------ */
/* shifted-force Coulomb: source/sepcoulomb.c:5-160                    */

static void coulomb_pair(seppart *p, long a, long b, double r2,
			 const double *d, double cf, double icf, double icf2,
			 sepret *r)
{
	const double zz = p[b].z * p[a].z;
	const double rr = sqrt(r2);
	const double ft = zz * (1.0 / r2 - icf2) / rr;
	double f[3];
	for (int k = 0; k < 3; k++) {
		f[k] = ft * d[k];
		p[a].f[k] += f[k];
		p[b].f[k] -= f[k];
	}
	add_stress(r, f, d);
	const double e = zz * (1.0 / rr + (rr - cf) * icf2 - icf);
	r->epot += e;
	r->ecoul += e;
}

void orc_coulomb_sf(seppart *p, double cf, sepsys *s, sepret *r, unsigned opt)
{
	const double cf2 = cf * cf, icf2 = 1.0 / cf2, icf = 1.0 / cf;
	if (s->neighb_update == SEP_BRUTE) {
		/* source/sepcoulomb.c:20-93; note the bonded test is "== 1" */
		for (long a = 0; a < s->npart - 1; a++) {
			if (fabs(p[a].z) < DBL_EPSILON)
				continue;
			for (long b = a + 1; b < s->npart; b++) {
				if (opt == SEP_EXCL_BONDED && bonded_count(p, a, b) == 1)
					continue;
				if (opt == SEP_EXCL_SAME_MOL &&
				    p[a].molindex == p[b].molindex && p[a].molindex != -1)
					continue;
				double d[3];
				const double r2 = separation(p, a, b, s->length, d);
				if (r2 < cf2)
					coulomb_pair(p, a, b, r2, d, cf, icf, icf2, r);
			}
		}
		return;
	}
	/* serial list path, source/sepcoulomb.c:96-160 */
	for (long a = 0; a < s->npart; a++) {
		if (fabs(p[a].z) < DBL_EPSILON)
			continue;
		for (int e = 0; p[a].neighb[e] != -1; e++) {
			const long b = p[a].neighb[e];
			double d[3];
			const double r2 = separation(p, a, b, s->length, d);
			if (r2 < cf2)
				coulomb_pair(p, a, b, r2, d, cf, icf, icf2, r);
		}
	}
}

/* ------------------------------------------------------------------ */
/* bonded forces: source/sepmol.c:372-587                              */

void orc_stretch_harmonic(seppart *p, int type, double lbond, double ks,
			  sepsys *s, sepret *r)
{
	sepmolinfo *m = s->molptr;
	for (unsigned n = 0; n < m->num_bonds; n++) {
		if ((int)m->blist[3 * n + 2] != type)
			continue;
		const long a = m->blist[3 * n], b = m->blist[3 * n + 1];
		double d[3], f[3];
		const double r2 = separation(p, a, b, s->length, d);
		const double dist = sqrt(r2);
		const double ft = -ks * (dist - lbond) / dist;
		for (int k = 0; k < 3; k++) {
			f[k] = ft * d[k];
			p[a].f[k] += f[k];
			p[b].f[k] -= f[k];
		}
		for (int i = 0; i < 3; i++)
			for (int j = 0; j < 3; j++) {
				r->pot_P[i][j] += f[i] * d[j];
				r->pot_P_bond[i][j] += f[i] * d[j];
			}
		r->epot += 0.5 * ks * ((dist - lbond) * (dist - lbond));
		m->blengths[n] = dist;
	}
}

/* Bond vectors b - a and c - b of an angle, each minimum-imaged. */
static void angle_vectors(const seppart *p, long a, long b, long c,
			  const double *L, double d1[3], double d2[3])
{
	for (int k = 0; k < 3; k++) {
		d1[k] = min_image(p[b].x[k] - p[a].x[k], L[k]);
		d2[k] = min_image(p[c].x[k] - p[b].x[k], L[k]);
	}
}

/* Force distribution shared by both angle forms: a gets f1, c gets f2, b
 * gets -(f1 + f2). */
static void angle_apply(seppart *p, long a, long b, long c, double f,
			const double *d1, const double *d2, double c11,
			double c12, double c22, double cD)
{
	for (int k = 0; k < 3; k++) {
		const double f1 = f * ((c12 / c11) * d1[k] - d2[k]) / cD;
		const double f2 = f * (d1[k] - (c12 / c22) * d2[k]) / cD;
		p[a].f[k] += f1;
		p[b].f[k] += (-f1 - f2);
		p[c].f[k] += f2;
	}
}

void orc_angle_cossq(seppart *p, int type, double angle0, double k,
		     sepsys *s, sepret *r)
{
	sepmolinfo *m = s->molptr;
	const double c0 = cos(SEP_PI - angle0);
	for (unsigned n = 0; n < m->num_angles; n++) {
		if ((int)m->alist[4 * n + 3] != type)
			continue;
		const long a = m->alist[4 * n], b = m->alist[4 * n + 1],
			   c = m->alist[4 * n + 2];
		double d1[3], d2[3];
		angle_vectors(p, a, b, c, s->length, d1, d2);
		const double c11 = dot3(d1, d1), c12 = dot3(d1, d2), c22 = dot3(d2, d2);
		const double cD = sqrt(c11 * c22);
		const double cc = c12 / cD;
		angle_apply(p, a, b, c, -k * (cc - c0), d1, d2, c11, c12, c22, cD);
		r->epot += 0.5 * k * ((cc - c0) * (cc - c0));
		m->angles[n] = SEP_PI - acos(cc);
	}
}

void orc_angle_harmonic(seppart *p, int type, double angle0, double k,
			sepsys *s, sepret *r)
{
	sepmolinfo *m = s->molptr;
	for (unsigned n = 0; n < m->num_angles; n++) {
		if ((int)m->alist[4 * n + 3] != type)
			continue;
		const long a = m->alist[4 * n], b = m->alist[4 * n + 1],
			   c = m->alist[4 * n + 2];
		double d1[3], d2[3];
		angle_vectors(p, a, b, c, s->length, d1, d2);
		const double c11 = dot3(d1, d1), c12 = dot3(d1, d2), c22 = dot3(d2, d2);
		const double cD = sqrt(c11 * c22);
		const double theta = SEP_PI - acos(c12 / cD);
		angle_apply(p, a, b, c, -k * (theta - angle0), d1, d2, c11, c12, c22, cD);
		r->epot += 0.5 * k * ((theta - angle0) * (theta - angle0));
		m->angles[n] = theta;
	}
}

void orc_torsion_ryckaert(seppart *p, int type, const double *g, sepsys *s,
			  sepret *r)
{
	sepmolinfo *m = s->molptr;
	for (unsigned n = 0; n < m->num_dihedrals; n++) {
		if ((int)m->dlist[5 * n + 4] != type)
			continue;
		const long a = m->dlist[5 * n], b = m->dlist[5 * n + 1],
			   c = m->dlist[5 * n + 2], e = m->dlist[5 * n + 3];
		double d1[3], d2[3], d3[3];
		for (int k = 0; k < 3; k++) {
			d1[k] = min_image(p[b].x[k] - p[a].x[k], s->length[k]);
			d2[k] = min_image(p[c].x[k] - p[b].x[k], s->length[k]);
			d3[k] = min_image(p[e].x[k] - p[c].x[k], s->length[k]);
		}
		const double c11 = dot3(d1, d1), c12 = dot3(d1, d2), c13 = dot3(d1, d3);
		const double c22 = dot3(d2, d2), c23 = dot3(d2, d3), c33 = dot3(d3, d3);
		const double cA = c13 * c22 - c12 * c23;
		const double cB1 = c11 * c22 - c12 * c12;
		const double cB2 = c22 * c33 - c23 * c23;
		const double cD = sqrt(cB1 * cB2);
		const double cc = cA / cD;
		/* -dU/dcos, Horner form */
		const double f = -(g[1] + (2. * g[2] + (3. * g[3] + (4. * g[4] + 5. * g[5] * cc) * cc) * cc) * cc);
		const double t1 = cA, t2 = c11 * c23 - c12 * c13, t3 = -cB1;
		const double t4 = cB2, t5 = c13 * c23 - c12 * c33, t6 = -cA;
		const double cR1 = c12 / c22, cR2 = c23 / c22;
		for (int k = 0; k < 3; k++) {
			const double f1 = f * c22 * (t1 * d1[k] + t2 * d2[k] + t3 * d3[k]) / (cD * cB1);
			const double f2 = f * c22 * (t4 * d1[k] + t5 * d2[k] + t6 * d3[k]) / (cD * cB2);
			p[a].f[k] += f1;
			p[b].f[k] += (-(1.0 + cR1) * f1 + cR2 * f2);
			p[c].f[k] += (cR1 * f1 - (1.0 + cR2) * f2);
			p[e].f[k] += f2;
		}
		r->epot += g[0] + (g[1] + (g[2] + (g[3] + (g[4] + g[5] * cc) * cc) * cc) * cc) * cc;
		m->dihedrals[n] = SEP_PI - acos(cc);
	}
}

/* ------------------------------------------------------------------ */
/* OpenMP Model II terms (source/sepomp.c:35-328)                      */

/* source/sepomp.c:16-33: the list for opt when the rebuild flag is set;
 * returns -1 for a cutoff above sys->cf or a list error */
int orc_omp_make_neighb(seppart *p, double cf, sepsys *s, unsigned opt, int nneighb)
{
	if (cf > s->cf)
		return -1;
	if (s->neighb_flag == 1) {
		if (orc_build_list(p, s, opt, nneighb))
			return -1;
		s->neighb_flag = 0;
	}
	return 0;
}

/* Pair terms over the half lists: atom i1 of the types, each listed i2 of
 * the other type, force ft r added to i1 and taken from i2, in list order
 * (source/sepomp.c:35-79, 81-133, 136-176). */
static int omp_types_ok(const seppart *p, long a, long b, const char *t)
{
	return (p[a].type == t[0] && p[b].type == t[1]) ||
	       (p[a].type == t[1] && p[b].type == t[0]);
}

static void omp_add_pair(double **ftot, long a, long b, double ft, const double *r)
{
	double f[3];
	for (int k = 0; k < 3; k++)
		f[k] = ft * r[k];
	for (int k = 0; k < 3; k++) {
		ftot[a][k] += f[k];
		ftot[b][k] -= f[k];
	}
}

void orc_omp_pairs(double **ftot, const seppart *p, const char *types, double cf,
		   orc_pot_fn fun, const sepsys *s)
{
	const double cf2 = cf * cf;
	for (long i1 = 0; i1 < s->npart; i1++) {
		if (p[i1].type != types[0] && p[i1].type != types[1])
			continue;
		for (int n = 0; p[i1].neighb[n] != -1; n++) {
			const long i2 = p[i1].neighb[n];
			if (!omp_types_ok(p, i1, i2, types))
				continue;
			double r[3];
			const double r2 = separation(p, i1, i2, s->length, r);
			if (r2 < cf2)
				omp_add_pair(ftot, i1, i2, fun(r2, 'f'), r);
		}
	}
}

void orc_omp_pairs_lj(double **ftot, const seppart *p, const char *types,
		      const double *param, const sepsys *s)
{
	const double cf = param[0], sigma = param[1], epsilon = param[2];
	const double cf2 = cf * cf;
	for (long i1 = 0; i1 < s->npart; i1++) {
		if (p[i1].type != types[0] && p[i1].type != types[1])
			continue;
		for (int n = 0; p[i1].neighb[n] != -1; n++) {
			const long i2 = p[i1].neighb[n];
			if (!omp_types_ok(p, i1, i2, types))
				continue;
			double r[3];
			const double r2 = separation(p, i1, i2, s->length, r);
			if (r2 < cf2) {
				/* source/sepomp.c:112-115, left to right */
				const double rri = sigma / r2;
				const double rri3 = rri * rri * rri;
				const double ft = 48.0 / epsilon * rri3 * (rri3 - 0.5) * rri;
				omp_add_pair(ftot, i1, i2, ft, r);
			}
		}
	}
}

void orc_omp_coulomb(double **ftot, const seppart *p, double cf, const sepsys *s)
{
	const double cf2 = cf * cf;
	const double icf2 = 1.0 / cf2;
	for (long i1 = 0; i1 < s->npart; i1++)
		for (int n = 0; p[i1].neighb[n] != -1; n++) {
			const long i2 = p[i1].neighb[n];
			double r[3];
			const double r2 = separation(p, i1, i2, s->length, r);
			if (r2 < cf2) {
				const double zizj = p[i2].z * p[i1].z;
				const double rr = sqrt(r2);
				omp_add_pair(ftot, i1, i2, zizj * (1.0 / r2 - icf2) / rr, r);
			}
		}
}

/* source/sepomp.c:179-212 */
void orc_omp_bond(double **ftot, const seppart *p, int type, double lbond, double ks,
		  const sepsys *s)
{
	const sepmolinfo *m = s->molptr;
	for (unsigned n = 0; n < m->num_bonds; n++) {
		if ((int)m->blist[3 * n + 2] != type)
			continue;
		const long a = m->blist[3 * n], b = m->blist[3 * n + 1];
		double r[3];
		const double dist = sqrt(separation(p, a, b, s->length, r));
		const double ft = -ks * (dist - lbond) / dist;
		for (int k = 0; k < 3; k++) {
			const double f = ft * r[k];
			ftot[a][k] += f;
			ftot[b][k] -= f;
		}
	}
}

/* source/sepomp.c:215-261: the cosine-squared force */
void orc_omp_angle(double **ftot, const seppart *p, int type, double angle0, double k,
		   const sepsys *s)
{
	const double ccon = cos(SEP_PI - angle0);
	const sepmolinfo *m = s->molptr;
	for (unsigned n = 0; n < m->num_angles; n++) {
		if ((int)m->alist[4 * n + 3] != type)
			continue;
		const long a = m->alist[4 * n], b = m->alist[4 * n + 1], c = m->alist[4 * n + 2];
		double d1[3], d2[3];
		angle_vectors(p, a, b, c, s->length, d1, d2);
		const double c11 = dot3(d1, d1), c12 = dot3(d1, d2), c22 = dot3(d2, d2);
		const double cd = sqrt(c11 * c22);
		const double f = -k * (c12 / cd - ccon);
		for (int q = 0; q < 3; q++) {
			const double f1 = f * ((c12 / c11) * d1[q] - d2[q]) / cd;
			const double f2 = f * (d1[q] - (c12 / c22) * d2[q]) / cd;
			ftot[a][q] += f1;
			ftot[b][q] += (-f1 - f2);
			ftot[c][q] += f2;
		}
	}
}

/* source/sepomp.c:264-328 */
void orc_omp_torsion(double **ftot, const seppart *p, int type, const double *g,
		     const sepsys *s)
{
	const sepmolinfo *m = s->molptr;
	for (unsigned n = 0; n < m->num_dihedrals; n++) {
		if ((int)m->dlist[5 * n + 4] != type)
			continue;
		const long a = m->dlist[5 * n], b = m->dlist[5 * n + 1];
		const long c = m->dlist[5 * n + 2], d = m->dlist[5 * n + 3];
		double d1[3], d2[3], d3[3];
		for (int k = 0; k < 3; k++) {
			d1[k] = min_image(p[b].x[k] - p[a].x[k], s->length[k]);
			d2[k] = min_image(p[c].x[k] - p[b].x[k], s->length[k]);
			d3[k] = min_image(p[d].x[k] - p[c].x[k], s->length[k]);
		}
		const double c11 = dot3(d1, d1), c12 = dot3(d1, d2), c13 = dot3(d1, d3);
		const double c22 = dot3(d2, d2), c23 = dot3(d2, d3), c33 = dot3(d3, d3);
		const double ca = c13 * c22 - c12 * c23;
		const double cb1 = c11 * c22 - c12 * c12;
		const double cb2 = c22 * c33 - c23 * c23;
		const double cd = sqrt(cb1 * cb2);
		const double cc = ca / cd;
		const double f = -(g[1] + (2. * g[2] + (3. * g[3] + (4. * g[4] + 5. * g[5] * cc) * cc) * cc) * cc);
		const double t1 = ca, t2 = c11 * c23 - c12 * c13, t3 = -cb1;
		const double t4 = cb2, t5 = c13 * c23 - c12 * c33, t6 = -ca;
		const double cr1 = c12 / c22, cr2 = c23 / c22;
		for (int k = 0; k < 3; k++) {
			const double f1 = f * c22 * (t1 * d1[k] + t2 * d2[k] + t3 * d3[k]) / (cd * cb1);
			const double f2 = f * c22 * (t4 * d1[k] + t5 * d2[k] + t6 * d3[k]) / (cd * cb2);
			ftot[a][k] += f1;
			ftot[b][k] += (-(1.0 + cr1) * f1 + cr2 * f2);
			ftot[c][k] += (cr1 * f1 - (1.0 + cr2) * f2);
			ftot[d][k] += f2;
		}
	}
}

/* ------------------------------------------------------------------ */
/* topology reader: source/sepmol.c:4-369                              */

/* Positions f just after the line following the exact header line, or
 * returns 0 if the header is absent (source/sepmol.c:4-20). */
static int seek_section(FILE *f, const char *header)
{
	char line[256];
	rewind(f);
	while (fgets(line, sizeof line, f)) {
		if (strcmp(line, header) == 0)
			return fgets(line, sizeof line, f) != NULL;
	}
	return 0;
}

/* Reads rows of `width` unsigned ints until a '[' line or EOF; returns the
 * number of rows, or -1 on a malformed row.  Rows are appended to *rows. */
static long read_rows(FILE *f, int width, unsigned **rows)
{
	char line[256];
	long n = 0, cap = 0;
	*rows = NULL;
	while (fgets(line, sizeof line, f)) {
		if (line[0] == '[')
			break;
		unsigned v[6];
		int got;
		char *s = line;
		while (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\n')
			s++;
		if (*s == '\0')
			continue;
		if (width == 4)
			got = sscanf(s, "%u%u%u%u", &v[0], &v[1], &v[2], &v[3]);
		else if (width == 5)
			got = sscanf(s, "%u%u%u%u%u", &v[0], &v[1], &v[2], &v[3], &v[4]);
		else
			got = sscanf(s, "%u%u%u%u%u%u", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5]);
		if (got != width)
			return -1;
		if (n == cap) {
			cap = cap ? 2 * cap : 1024;
			*rows = realloc(*rows, sizeof(unsigned) * (size_t)(cap * width));
		}
		memcpy(*rows + n * width, v, sizeof(unsigned) * (size_t)width);
		n++;
	}
	return n;
}

/* Appends `who` to the first free slot (-1) of a partner array. */
static int push_partner(int *arr, int cap, int who)
{
	for (int k = 0; k < cap; k++)
		if (arr[k] == -1) {
			arr[k] = who;
			return 0;
		}
	return -1;
}

int orc_read_topology(seppart *p, const char *file, sepsys *s)
{
	sepmolinfo *m = s->molptr;
	FILE *f = fopen(file, "r");
	if (!f)
		return -1;
	unsigned *rows;
	long n;

	/* bonds: mol a b type */
	m->flag_bonds = 1;
	m->num_bonds = m->num_btypes = m->num_mols = 0;
	if (seek_section(f, "[ bonds ]\n")) {
		if ((n = read_rows(f, 4, &rows)) < 0)
			goto bad;
		m->num_bonds = (unsigned)n;
		m->blist = malloc(sizeof(unsigned) * 3 * (size_t)(n ? n : 1));
		for (long i = 0; i < n; i++) {
			const unsigned *v = rows + 4 * i;
			m->blist[3 * i] = v[1];
			m->blist[3 * i + 1] = v[2];
			m->blist[3 * i + 2] = v[3];
			p[v[1]].molindex = (int)v[0];
			p[v[2]].molindex = (int)v[0];
			if (push_partner(p[v[1]].bond, SEP_BOND, (int)v[2]) ||
			    push_partner(p[v[2]].bond, SEP_BOND, (int)v[1]))
				goto bad;
			if (v[3] > m->num_btypes) m->num_btypes = v[3];
			if (v[0] > m->num_mols) m->num_mols = v[0];
		}
		free(rows);
		m->num_btypes++;
		m->num_mols++;
	}
	m->blengths = malloc(sizeof(double) * (m->num_bonds ? m->num_bonds : 1));

	/* angles: mol a b c type */
	m->flag_angles = 1;
	m->num_angles = m->num_atypes = 0;
	if (seek_section(f, "[ angles ]\n")) {
		if ((n = read_rows(f, 5, &rows)) < 0)
			goto bad;
		m->num_angles = (unsigned)n;
		m->alist = malloc(sizeof(unsigned) * 4 * (size_t)(n ? n : 1));
		for (long i = 0; i < n; i++) {
			const unsigned *v = rows + 5 * i;
			for (int k = 0; k < 4; k++)
				m->alist[4 * i + k] = v[1 + k];
			const int a = (int)v[1], b = (int)v[2], c = (int)v[3];
			if (push_partner(p[a].angle, SEP_ANGLE, b) || push_partner(p[a].angle, SEP_ANGLE, c) ||
			    push_partner(p[b].angle, SEP_ANGLE, a) || push_partner(p[b].angle, SEP_ANGLE, c) ||
			    push_partner(p[c].angle, SEP_ANGLE, a) || push_partner(p[c].angle, SEP_ANGLE, b))
				goto bad;
			if (v[4] > m->num_atypes) m->num_atypes = v[4];
		}
		free(rows);
		m->num_atypes++;
	}
	m->angles = malloc(sizeof(double) * (m->num_angles ? m->num_angles : 1));

	/* dihedrals: mol a b c d type */
	m->flag_dihedrals = 1;
	m->num_dihedrals = m->num_dtypes = 0;
	if (seek_section(f, "[ dihedrals ]\n")) {
		if ((n = read_rows(f, 6, &rows)) < 0)
			goto bad;
		m->num_dihedrals = (unsigned)n;
		m->dlist = malloc(sizeof(unsigned) * 5 * (size_t)(n ? n : 1));
		for (long i = 0; i < n; i++) {
			const unsigned *v = rows + 6 * i;
			for (int k = 0; k < 5; k++)
				m->dlist[5 * i + k] = v[1 + k];
			const int q[4] = {(int)v[1], (int)v[2], (int)v[3], (int)v[4]};
			for (int x = 0; x < 4; x++)
				for (int y = 0; y < 4; y++)
					if (x != y && push_partner(p[q[x]].dihed, SEP_DIHED, q[y]))
						goto bad;
			if (v[5] > m->num_dtypes) m->num_dtypes = v[5];
		}
		free(rows);
		m->num_dtypes++;
	}
	m->dihedrals = malloc(sizeof(double) * (m->num_dihedrals ? m->num_dihedrals : 1));
	fclose(f);
	return 0;
bad:
	fclose(f);
	return -1;
}

/* ------------------------------------------------------------------ */
/* integrators: source/sepintgr.c:18-88, 149-168                       */

void orc_nosehoover(seppart *p, double Td, double *alpha, double tau,
		    sepsys *s)
{
	double sum = 0.0;
	for (long i = 0; i < s->npart; i++)
		for (int k = 0; k < 3; k++)
			sum += (p[i].v[k] * p[i].v[k]) * p[i].m;
	const double ekin = 0.5 * sum / s->npart;
	const double temp = 0.666667 * ekin;
	*alpha = *alpha + s->dt / (tau * tau) * (temp / Td - 1.0);
	for (long i = 0; i < s->npart; i++)
		for (int k = 0; k < 3; k++)
			p[i].f[k] -= *alpha * p[i].m * p[i].v[k];
}

/* Wrap once per axis, count crossings, and return the squared distance
 * from the last-rebuild position (source/sepintgr.c:18-40). */
static double wrap_and_displacement(seppart *p, long i, const sepsys *s)
{
	double d2 = 0.0;
	for (int k = 0; k < 3; k++) {
		const double L = s->length[k];
		if (p[i].x[k] > L) {
			p[i].x[k] -= L;
			p[i].cross_neighb[k]++;
			p[i].crossings[k]++;
		} else if (p[i].x[k] < 0.0) {
			p[i].x[k] += L;
			p[i].cross_neighb[k]--;
			p[i].crossings[k]--;
		}
		const double dr = (p[i].x[k] + p[i].cross_neighb[k] * L) - p[i].xn[k];
		d2 += dr * dr;
	}
	return d2;
}

void orc_leapfrog(seppart *p, sepsys *s, sepret *r)
{
	const double dt = s->dt;
	double sum = 0.0;
	for (long i = 0; i < s->npart; i++) {
		double vh[3];
		for (int k = 0; k < 3; k++) {
			p[i].a[k] = p[i].f[k] / p[i].m;
			p[i].v[k] += p[i].a[k] * dt;
			p[i].x[k] += p[i].v[k] * dt;
			vh[k] = p[i].v[k] - 0.5 * p[i].a[k] * dt;
			sum += vh[k] * vh[k] * p[i].m;
		}
		const double d2 = wrap_and_displacement(p, i, s);
		if (d2 > s->max_dist2)
			s->max_dist2 = d2;
		for (int a = 0; a < 3; a++)
			for (int b = 0; b < 3; b++)
				r->kin_P[a][b] += vh[a] * vh[b] * p[i].m;
	}
	if (sqrt(s->max_dist2) > s->skin * 0.5) {
		s->neighb_flag = 1;
		s->nupdate_neighb++;
		for (long i = 0; i < s->npart; i++)
			for (int k = 0; k < 3; k++) {
				p[i].xn[k] = p[i].x[k];
				p[i].cross_neighb[k] = 0;
			}
	}
	s->tnow += dt;
	r->ekin += 0.5 * sum;
}
