/*
 * sep_array.c -- the N x 3 arrays of the OpenMP Model II terms (reference
 * include/separray.h:39,90,206; source/separray.c:31-52,215-222,419-425):
 * one zero-filled row of ncol doubles per row pointer.
 */
#include <stdlib.h>

#include "sephost.h"

double **sep_matrix(size_t nrow, size_t ncol)
{
	double **a = malloc((nrow ? nrow : 1) * sizeof(double *));
	if (!a) {
		sep_error("sep_matrix: couldn't allocate memory");
		return NULL;
	}
	for (size_t n = 0; n < nrow; n++) {
		a[n] = calloc(ncol ? ncol : 1, sizeof(double));
		if (!a[n]) {
			sep_free_matrix(a, n);
			sep_error("sep_matrix: couldn't allocate memory");
			return NULL;
		}
	}
	return a;
}

void sep_matrix_set(double **a, size_t nrow, size_t ncol, double value)
{
	for (size_t n = 0; n < nrow; n++)
		for (size_t m = 0; m < ncol; m++)
			a[n][m] = value;
}

void sep_free_matrix(double **a, size_t nrow)
{
	if (!a)
		return;
	for (size_t n = 0; n < nrow; n++)
		free(a[n]);
	free(a);
}
