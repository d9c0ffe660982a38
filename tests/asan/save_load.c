/* sep_save in one process, sep_load and sep_close in another (the
 * reference's binary checkpoint, source/sepmisc.c:319-354).  Built with
 * AddressSanitizer over the library's host sources by
 * tests/test_asan_host.py; any invalid free or out-of-bounds access in the
 * load / close path aborts the run.
 *
 *   save_load save FILE N        writes N atoms with known fields
 *   save_load load FILE N NNEIGHB reads them into a fresh system, checks
 *                                 them, fills each new neighbour array and
 *                                 closes the system; prints "ok" */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sep.h"

static double field(long i, int k) { return 0.25 * (double)i + 0.125 * (double)k; }

int main(int argc, char **argv)
{
	if (argc < 4)
		return 2;
	const long n = atol(argv[3]);
	if (strcmp(argv[1], "save") == 0) {
		seppart *a = sep_init((size_t)n, 8);
		for (long i = 0; i < n; i++) {
			for (int k = 0; k < 3; k++) {
				a[i].x[k] = field(i, k);
				a[i].v[k] = -field(i, k);
			}
			a[i].m = 1.0 + 0.5 * (double)(i % 3);
			a[i].type = (char)('A' + i % 2);
			a[i].neighb[0] = (int)((i + 1) % n);
			a[i].neighb[1] = -1;
		}
		sep_save(a, (size_t)n, argv[2]);
		sep_close(a, (size_t)n);
		return 0;
	}
	if (strcmp(argv[1], "load") == 0 && argc >= 5) {
		const int nn = atoi(argv[4]);
		seppart *a = sep_init((size_t)n, 4);
		sep_load(a, nn, (size_t)n, argv[2]);
		int bad = 0;
		for (long i = 0; i < n; i++) {
			for (int k = 0; k < 3; k++)
				bad += a[i].x[k] != field(i, k) || a[i].v[k] != -field(i, k);
			bad += a[i].m != 1.0 + 0.5 * (double)(i % 3);
			bad += a[i].type != (char)('A' + i % 2);
			bad += a[i].neighb[0] != -1;
			/* the new neighbour arrays hold nn ints each */
			for (int k = 0; k < nn; k++)
				a[i].neighb[k] = k;
		}
		sep_close(a, (size_t)n);
		if (bad) {
			printf("mismatch %d\n", bad);
			return 1;
		}
		printf("ok\n");
		return 0;
	}
	return 2;
}
