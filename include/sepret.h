/*
 * sepret.h -- return values: energies, stress, pressure (seplib C99 API).
 */
#ifndef SEP_MI355X_SEPRET_H
#define SEP_MI355X_SEPRET_H

#include "sepstrct.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Zeroes the scalar energies and every 3x3 tensor of ret.  When ret is
 * bound to a device-resident system (sep_gpu_resident) its device copy is
 * zeroed too.  Replaces include/sepret.h:36 (source/sepret.c:19-47). */
void sep_reset_retval(sepret *retval);

/* P = (kin_P + pot_P)/V and p = tr(P)/3; returns p.
 * Replaces include/sepret.h:44 (source/sepret.c:50-55). */
double sep_get_pressure(sepret *retval, sepsys *sys);

/* T = 2/(3 ndof) ekin.  Replaces include/sepret.h:52 (source/sepret.c:58-62). */
double sep_get_temperature(sepret *retval, sepsys *sys);

/* P = (kin_P + pot_P)/V, p = tr(P)/3.
 * Replaces include/sepret.h:61 (source/sepret.c:65-82). */
void sep_pressure_tensor(sepret *retval, sepsys *sys);

#ifdef __cplusplus
}
#endif
#endif
