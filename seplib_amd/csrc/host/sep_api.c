/*
 * sep_api.c -- allocation, system setup, return values, potentials and
 * diagnostics of the seplib C99 API (host side; no device work).
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sephost.h"

int seph_each_ctx(int (*fn)(seph_ctx *, void *), void *arg);

/* ---- diagnostics: include/sepmisc.h:101,108 --------------------------- */

/* Expands the %d %i %f %c %s subset the reference's printer handles. */
static void print_tagged(const char *tag, const char *str, va_list ap)
{
	printf("%s", tag);
	for (const char *p = str; *p; p++) {
		if (*p != '%') {
			putchar(*p);
			continue;
		}
		switch (*++p) {
		case 'd':
		case 'i': printf("%d", va_arg(ap, int)); break;
		case 'f': printf("%f", va_arg(ap, double)); break;
		case 'c': putchar(va_arg(ap, int)); break;
		case 's': fputs(va_arg(ap, char *), stdout); break;
		case '\0': p--; break;
		default: putchar(*p); break;
		}
	}
	printf("\n");
	fflush(stdout);
}

void sep_error(char *str, ...)
{
	va_list ap;
	va_start(ap, str);
	print_tagged("sep-error -> ", str, ap);
	va_end(ap);
	seph_fail(SEP_GPU_EHOST, "sep_error: %s", str);
}

void sep_warning(char *str, ...)
{
	va_list ap;
	va_start(ap, str);
	print_tagged("sep-warning -> ", str, ap);
	va_end(ap);
}

/* ---- allocation: source/sepinit.c:16-65 ------------------------------- */

seppart *sep_init(size_t npart, size_t nneighb)
{
	seppart *p = calloc(npart ? npart : 1, sizeof(seppart));
	if (!p) {
		sep_error("%s: Couldn't allocate memory", __func__);
		return NULL;
	}
	for (size_t i = 0; i < npart; i++) {
		p[i].neighb = malloc((nneighb ? nneighb : 1) * sizeof(int));
		if (!p[i].neighb) {
			sep_error("%s: Couldn't allocate memory", __func__);
			return NULL;
		}
		if (nneighb)
			p[i].neighb[0] = -1;
		p[i].type = 'A';
		p[i].m = 1.0;
		p[i].molindex = -1;
		p[i].ldiff = 1.0;
		for (int k = 0; k < SEP_BOND; k++) p[i].bond[k] = -1;
		for (int k = 0; k < SEP_ANGLE; k++) p[i].angle[k] = -1;
		for (int k = 0; k < SEP_DIHED; k++) p[i].dihed[k] = -1;
	}
	seph_register(p, (long)npart, (int)nneighb);
	return p;
}

void sep_close(seppart *ptr, size_t npart)
{
	if (!ptr)
		return;
	seph_unregister(ptr);
	for (size_t i = 0; i < npart; i++)
		free(ptr[i].neighb);
	free(ptr);
}

/* Binary dump of the AoS, as the reference: npart raw seppart records
 * (source/sepmisc.c:319-330).  A resident system is synchronised first. */
void sep_save(seppart *ptr, size_t npart, const char *file)
{
	seph_ctx *c = seph_find(ptr);
	if (c && c->resident && seph_get_atoms(c) != 0) {
		sep_warning("sep_save: %s", (char *)sep_gpu_last_error_message());
		return;
	}
	FILE *fout = fopen(file, "wb");
	if (fout == NULL) {
		sep_error("%s at line %d: Couldn't open the file: %s\n", __func__, __LINE__, file);
		return;
	}
	fwrite(ptr, sizeof(seppart), npart, fout);
	fclose(fout);
}

/* Reads npart raw seppart records into ptr and gives every atom a fresh
 * neighbour array of nneighb ints (source/sepmisc.c:333-354); the pointer
 * fields read from the file are overwritten.  A resident system's device
 * state is replaced by the loaded atoms and its list is rebuilt at the next
 * pair call (the reference leaves that to the caller's sys->neighb_flag). */
void sep_load(seppart *ptr, int nneighb, size_t npart, const char *file)
{
	for (size_t n = 0; n < npart; n++)
		free(ptr[n].neighb);
	FILE *fin = fopen(file, "rb");
	size_t got = 0;
	if (fin == NULL)
		sep_error("%s at line %d: Couldn't open the file: %s\n", __func__, __LINE__, file);
	else {
		got = fread(ptr, sizeof(seppart), npart, fin);
		if (got == 0)
			sep_error("%s at line %d: Error reading file %s\n", __func__, __LINE__, file);
		fclose(fin);
	}
	for (size_t n = 0; n < npart; n++) {
		ptr[n].neighb = malloc((nneighb > 0 ? nneighb : 1) * sizeof(int));
		if (ptr[n].neighb == NULL) {
			sep_error("%s at line %d: Couldn't allocate memory\n", __func__, __LINE__);
			return;
		}
		ptr[n].neighb[0] = -1;
	}
	// half lists written back to these atoms must fit the new arrays
	seph_set_nneighb(ptr, nneighb > 0 ? nneighb : 1);
	// a resident system takes the loaded atoms, and its next pair call
	// rebuilds the list for them
	seph_ctx *c = seph_find(ptr);
	if (got && c && c->resident) {
		sepdev_sys ds;
		/* the file's bond, angle and dihedral partners replace the
		 * device's at the next call that excludes bonded pairs */
		c->partners_on_device = 0;
		int rc = seph_put_atoms(c);
		if (!rc)
			rc = seph_dev_fail(sepdev_get_sys(c->dev, &ds));
		if (!rc) {
			ds.neighb_flag = 1;
			rc = seph_dev_fail(sepdev_put_sys(c->dev, &ds));
		}
		if (rc)
			sep_warning("sep_load: %s", (char *)sep_gpu_last_error_message());
	}
}

/* Text format "N\nLx Ly Lz\n" then "type x y z vx vy vz m q" per atom
 * (source/sepinit.c:69-109). */
seppart *sep_init_xyz(double *lbox, int *npart, const char *file, char verbose)
{
	if (verbose == 'v')
		fprintf(stdout, "Opening %s\n", file);
	FILE *fin = fopen(file, "r");
	if (!fin) {
		sep_error("%s: Couldn't open file", __func__);
		return NULL;
	}
	if (fscanf(fin, "%d", npart) != 1 || *npart <= 0) {
		fclose(fin);
		sep_error("%s: Error reading xyz file", __func__);
		return NULL;
	}
	seppart *p = sep_init((size_t)*npart, SEP_NUM_NEIGHB);
	if (!p) {
		fclose(fin);
		return NULL;
	}
	if (fscanf(fin, "%lf%lf%lf\n", &lbox[0], &lbox[1], &lbox[2]) != 3) {
		fclose(fin);
		sep_error("%s: Error reading xyz file", __func__);
		return p;
	}
	for (int i = 0; i < *npart; i++) {
		if (fscanf(fin, "%c%lf%lf%lf%lf%lf%lf%lf%lf\n", &p[i].type, &p[i].x[0],
			   &p[i].x[1], &p[i].x[2], &p[i].v[0], &p[i].v[1], &p[i].v[2],
			   &p[i].m, &p[i].z) != 9) {
			sep_error("%s: Error reading xyz file", __func__);
			break;
		}
	}
	if (verbose == 'v') {
		fprintf(stdout, "Closing file\n");
		fprintf(stdout, "Number of particles: %d\n", *npart);
		fflush(stdout);
	}
	fclose(fin);
	return p;
}

/* ---- system: source/sepinit.c:246-315 --------------------------------- */

sepsys sep_sys_setup(double lengthx, double lengthy, double lengthz,
		     double maxsyscf, double dt, size_t npart, size_t update)
{
	sepsys s;
	memset(&s, 0, sizeof s);
	const double skin0 = 0.25;
	const double L[3] = {lengthx, lengthy, lengthz};
	const char axis[3] = {'x', 'y', 'z'};
	for (int k = 0; k < 3; k++) {
		s.length[k] = L[k];
		s.nsubbox[k] = (int)(L[k] / (maxsyscf + skin0));
		if (update > 1 && s.nsubbox[k] < 3)
			printf("sep-warning -> %s: Number of subboxes in %c direction are less than three\n",
			       __func__, axis[k]);
		s.lsubbox[k] = L[k] / s.nsubbox[k];
	}
	s.volume = lengthx * lengthy * lengthz;
	s.npart = (long)npart;
	s.cf = maxsyscf;
	s.dt = dt;
	s.ndof = (unsigned)(3 * npart - 3);
	s.tnow = 0.0;
	s.nupdate_neighb = 0;
	s.neighb_update = (unsigned)update;
	s.neighb_flag = 1;
	s.skin = skin0;
	s.molptr = calloc(1, sizeof(sepmolinfo));
	s.omp_flag = false;
	s.fun_cstate = 0;
	return s;
}

void sep_free_sys(sepsys *ptr)
{
	sepmolinfo *m = ptr->molptr;
	if (!m)
		return;
	if (m->flag_bonds == 1) {
		free(m->blist);
		free(m->blengths);
		m->blist = NULL;
		m->blengths = NULL;
	}
	if (m->flag_angles == 1) {
		free(m->alist);
		free(m->angles);
		m->alist = NULL;
		m->angles = NULL;
	}
	if (m->flag_dihedrals == 1) {
		free(m->dlist);
		free(m->dihedrals);
		m->dlist = NULL;
		m->dihedrals = NULL;
	}
}

void sep_set_omp(unsigned nthreads, sepsys *sys)
{
	sys->omp_flag = true;
	sys->nthreads = nthreads;
}

void sep_set_skin(sepsys *sys, double value) { sys->skin = value; }

/* ---- return values: source/sepret.c:19-82 ----------------------------- */

static int zero_bound_ret(seph_ctx *c, void *arg)
{
	if (!c->resident)
		return 0;
	for (int b = 0; b < c->nrets; b++)
		if (c->rets[b].host == arg)
			return seph_dev_fail(sepdev_ret_zero(c->dev, c->rets[b].slot));
	return 0;
}

void sep_reset_retval(sepret *r)
{
	r->epot = 0.0;
	r->ecoul = 0.0;
	r->ekin = 0.0;
	r->sumv2 = 0.0;
	memset(r->P, 0, sizeof r->P);
	memset(r->kin_P, 0, sizeof r->kin_P);
	memset(r->pot_P, 0, sizeof r->pot_P);
	memset(r->P_mol, 0, sizeof r->P_mol);
	memset(r->kin_P_mol, 0, sizeof r->kin_P_mol);
	memset(r->pot_P_mol, 0, sizeof r->pot_P_mol);
	memset(r->pot_P_conservative, 0, sizeof r->pot_P_conservative);
	memset(r->pot_P_random, 0, sizeof r->pot_P_random);
	memset(r->pot_P_dissipative, 0, sizeof r->pot_P_dissipative);
	memset(r->pot_P_bond, 0, sizeof r->pot_P_bond);
	memset(r->pot_T_mol, 0, sizeof r->pot_T_mol);
	memset(r->kin_T_mol, 0, sizeof r->kin_T_mol);
	memset(r->T_mol, 0, sizeof r->T_mol);
	seph_each_ctx(zero_bound_ret, r);
}

void sep_pressure_tensor(sepret *r, sepsys *sys)
{
	const double ivol = 1.0 / sys->volume;
	for (int a = 0; a < 3; a++)
		for (int b = 0; b < 3; b++)
			r->P[a][b] = (r->kin_P[a][b] + r->pot_P[a][b]) * ivol;
	double tr = 0.0;
	for (int a = 0; a < 3; a++)
		tr += r->P[a][a];
	r->p = tr / 3.0;
}

double sep_get_pressure(sepret *r, sepsys *sys)
{
	sep_pressure_tensor(r, sys);
	return r->p;
}

double sep_get_temperature(sepret *r, sepsys *sys)
{
	return 2.0 / (3.0 * sys->ndof) * r->ekin;
}

/* ---- potentials: source/sepmisc.c:115-164 ----------------------------- */
/* Host evaluations for callers; sep_force_pairs recognises these three by
 * address and evaluates the same expressions on the device. */

double sep_lj(double r2, char opt)
{
	const double inv = 1.0 / r2, inv3 = inv * inv * inv;
	if (opt == 'f')
		return 48. * inv3 * (inv3 - 0.5) * inv;
	if (opt == 'u')
		return 4.0 * inv3 * (inv3 - 1.0);
	return 0.0;
}

double sep_lj_shift(double r2, char opt)
{
	const double inv = 1.0 / r2, inv3 = inv * inv * inv;
	if (opt == 'f')
		return 48.0 * inv3 * (inv3 - 0.5) * inv;
	if (opt == 'u')
		return 4.0 * inv3 * (inv3 - 1.0) + SEP_LJCF2;
	return 0.0;
}

double sep_wca(double r2, char opt)
{
	const double inv = 1.0 / r2, inv3 = inv * inv * inv;
	if (opt == 'f')
		return 48.0 * inv3 * (inv3 - 0.5) * inv;
	if (opt == 'u')
		return 4.0 * inv3 * (inv3 - 1.0) + 1.0;
	return 0.0;
}

/* ---- configuration output: source/sepmisc.c:536-573 -------------------- */

void sep_save_xyz(seppart *ptr, const char *partnames, const char *file, char *mode,
		  sepsys *sys)
{
	seph_ctx *c = seph_find(ptr);
	if (c && c->resident && seph_get_all(c, sys) != 0) {
		sep_warning("sep_save_xyz: %s", (char *)sep_gpu_last_error_message());
		return;
	}
	FILE *fout = fopen(file, mode);
	if (fout == NULL) {
		sep_error("%s: I couldn't open file %s", __func__, file);
		return;
	}
	const size_t ntype = strlen(partnames);
	unsigned long ntotal = 0;
	for (long n = 0; n < sys->npart; n++)
		for (size_t k = 0; k < ntype; k++)
			if (ptr[n].type == partnames[k])
				ntotal++;
	fprintf(fout, "%lu\n%f %f %f\n", ntotal, sys->length[0], sys->length[1], sys->length[2]);
	for (long n = 0; n < sys->npart; n++)
		for (size_t k = 0; k < ntype; k++)
			if (ptr[n].type == partnames[k])
				fprintf(fout, "%c %.15f %.15f %.15f %.15f %.15f %.15f %.15f %.15f\n",
					ptr[n].type, ptr[n].x[0], ptr[n].x[1], ptr[n].x[2], ptr[n].v[0],
					ptr[n].v[1], ptr[n].v[2], ptr[n].m, ptr[n].z);
	fclose(fout);
}
