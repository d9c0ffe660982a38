/*
 * sep_topology.c -- .top reader of the seplib C99 API
 * (source/sepmol.c:4-369).
 *
 * A section starts at the exact line "[ bonds ]" (or "[ angles ]",
 * "[ dihedrals ]"); the line after it is a comment and is skipped; rows
 * follow until a line starting with '[' or the end of the file.  Blank
 * lines between rows are skipped.  Rows:
 *   bonds      mol a b type
 *   angles     mol a b c type
 *   dihedrals  mol a b c d type
 * Bonds set molindex of both atoms; every kind appends the other atoms of
 * the term to each atom's partner array (bond[], angle[], dihed[]) in row
 * order.  A missing section leaves its list empty.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sephost.h"

static int find_section(FILE *f, const char *header)
{
	char line[256];
	rewind(f);
	while (fgets(line, sizeof line, f))
		if (strcmp(line, header) == 0)
			return fgets(line, sizeof line, f) != NULL;
	return 0;
}

/* Returns the number of rows read into a growing array, -1 on bad rows. */
static long read_section(FILE *f, int width, unsigned **out)
{
	char line[256];
	long n = 0, cap = 0;
	unsigned *rows = NULL;
	while (fgets(line, sizeof line, f)) {
		if (line[0] == '[')
			break;
		char *s = line + strspn(line, " \t\r\n");
		if (*s == '\0')
			continue;
		unsigned v[6];
		int got = 0;
		char *end = s;
		for (int k = 0; k < width; k++) {
			char *nx;
			unsigned long x = strtoul(end, &nx, 10);
			if (nx == end)
				break;
			v[k] = (unsigned)x;
			end = nx;
			got++;
		}
		if (got != width) {
			free(rows);
			return -1;
		}
		if (n == cap) {
			cap = cap ? 2 * cap : 1024;
			rows = realloc(rows, sizeof(unsigned) * (size_t)(cap * width));
		}
		memcpy(rows + n * width, v, sizeof(unsigned) * (size_t)width);
		n++;
	}
	*out = rows;
	return n;
}

static int add_partner(int *arr, int cap, int who)
{
	for (int k = 0; k < cap; k++)
		if (arr[k] == -1) {
			arr[k] = who;
			return 0;
		}
	return -1;
}

static void report(const char *what, const char *file)
{
	sep_error("sep_read_topology_file: %s in %s", what, file);
}

void sep_read_topology_file(sepatom *aptr, const char *file, sepsys *sys, char opt)
{
	sepmolinfo *m = sys->molptr;
	FILE *f = fopen(file, "r");
	if (!f) {
		report("Couldn't open file", file);
		return;
	}
	unsigned *rows = NULL;
	long n;

	m->flag_bonds = 1;
	m->num_bonds = m->num_btypes = m->num_mols = 0;
	m->blist = NULL;
	if (find_section(f, "[ bonds ]\n")) {
		if ((n = read_section(f, 4, &rows)) < 0) {
			report("Format in top file not correct", file);
			goto done;
		}
		m->num_bonds = (unsigned)n;
		m->blist = malloc(sizeof(unsigned) * 3 * (size_t)(n ? n : 1));
		for (long i = 0; i < n; i++) {
			const unsigned *v = rows + 4 * i;
			if (v[1] >= (unsigned)sys->npart || v[2] >= (unsigned)sys->npart) {
				report("atom index out of range", file);
				goto done;
			}
			m->blist[3 * i] = v[1];
			m->blist[3 * i + 1] = v[2];
			m->blist[3 * i + 2] = v[3];
			aptr[v[1]].molindex = (int)v[0];
			aptr[v[2]].molindex = (int)v[0];
			if (add_partner(aptr[v[1]].bond, SEP_BOND, (int)v[2]) ||
			    add_partner(aptr[v[2]].bond, SEP_BOND, (int)v[1])) {
				report("Index exceeds the allowed number of bonds", file);
				goto done;
			}
			if (v[3] > m->num_btypes) m->num_btypes = v[3];
			if (v[0] > m->num_mols) m->num_mols = v[0];
		}
		free(rows);
		rows = NULL;
		m->num_btypes++;
		m->num_mols++;
		if (opt == 'v')
			printf("Succesfully read 'bond' section in file %s -> Found %d molecules, %d bond(s) and %d bond type(s).\n",
			       file, m->num_mols, m->num_bonds, m->num_btypes);
	}
	m->blengths = calloc(m->num_bonds ? m->num_bonds : 1, sizeof(double));

	m->flag_angles = 1;
	m->num_angles = m->num_atypes = 0;
	m->alist = NULL;
	if (find_section(f, "[ angles ]\n")) {
		if ((n = read_section(f, 5, &rows)) < 0) {
			report("Format in top file not correct", file);
			goto done;
		}
		m->num_angles = (unsigned)n;
		m->alist = malloc(sizeof(unsigned) * 4 * (size_t)(n ? n : 1));
		for (long i = 0; i < n; i++) {
			const unsigned *v = rows + 5 * i;
			for (int k = 0; k < 4; k++)
				m->alist[4 * i + k] = v[1 + k];
			const int q[3] = {(int)v[1], (int)v[2], (int)v[3]};
			for (int x = 0; x < 3; x++) {
				if (q[x] >= sys->npart) {
					report("atom index out of range", file);
					goto done;
				}
				for (int y = 0; y < 3; y++)
					if (x != y && add_partner(aptr[q[x]].angle, SEP_ANGLE, q[y])) {
						report("Index exceeds SEP_ANGLE", file);
						goto done;
					}
			}
			if (v[4] > m->num_atypes) m->num_atypes = v[4];
		}
		free(rows);
		rows = NULL;
		m->num_atypes++;
		if (opt == 'v')
			printf("Succesfully read 'angles' section in file %s -> Found %d angles(s) and %d bond angles(s).\n",
			       file, m->num_angles, m->num_atypes);
	}
	m->angles = calloc(m->num_angles ? m->num_angles : 1, sizeof(double));

	m->flag_dihedrals = 1;
	m->num_dihedrals = m->num_dtypes = 0;
	m->dlist = NULL;
	if (find_section(f, "[ dihedrals ]\n")) {
		if ((n = read_section(f, 6, &rows)) < 0) {
			report("Format in top file not correct", file);
			goto done;
		}
		m->num_dihedrals = (unsigned)n;
		m->dlist = malloc(sizeof(unsigned) * 5 * (size_t)(n ? n : 1));
		for (long i = 0; i < n; i++) {
			const unsigned *v = rows + 6 * i;
			for (int k = 0; k < 5; k++)
				m->dlist[5 * i + k] = v[1 + k];
			const int q[4] = {(int)v[1], (int)v[2], (int)v[3], (int)v[4]};
			for (int x = 0; x < 4; x++) {
				if (q[x] >= sys->npart) {
					report("atom index out of range", file);
					goto done;
				}
				for (int y = 0; y < 4; y++)
					if (x != y && add_partner(aptr[q[x]].dihed, SEP_DIHED, q[y])) {
						report("Index exceeds SEP_DIHED", file);
						goto done;
					}
			}
			if (v[5] > m->num_dtypes) m->num_dtypes = v[5];
		}
		free(rows);
		rows = NULL;
		m->num_dtypes++;
		if (opt == 'v')
			printf("Succesfully read 'dihedrals' section in file %s -> Found %d dihedrals(s) and %d dihedral types(s).\n",
			       file, m->num_dihedrals, m->num_dtypes);
	}
	m->dihedrals = calloc(m->num_dihedrals ? m->num_dihedrals : 1, sizeof(double));
done:
	free(rows);
	fclose(f);
}
