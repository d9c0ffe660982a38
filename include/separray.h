/*
 * separray.h -- the array helpers of the reference's include/separray.h that
 * the per-timestep path uses: the N x 3 force arrays the OpenMP Model II
 * terms (sepomp.h) add into.  Same names and signatures as the reference
 * (include/separray.h:39, :90, :206); the reference's other vector, tensor
 * and integer-array helpers are not part of this library.
 */
#ifndef SEP_MI355X_SEPARRAY_H
#define SEP_MI355X_SEPARRAY_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reference include/separray.h:39 (source/separray.c:31-52): nrow zero-filled
 * rows of ncol doubles, one row pointer each */
double **sep_matrix(size_t nrow, size_t ncol);

/* reference include/separray.h:206 (source/separray.c:215-222) */
void sep_matrix_set(double **a, size_t nrow, size_t ncol, double value);

/* reference include/separray.h:90 (source/separray.c:419-425) */
void sep_free_matrix(double **ptr, size_t nrow);

#ifdef __cplusplus
}
#endif
#endif
