/*
 * sepintgr.h -- integrators on the hot path (seplib C99 API).
 */
#ifndef SEP_MI355X_SEPINTGR_H
#define SEP_MI355X_SEPINTGR_H

#include "sepstrct.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Leapfrog step: a = f/m, v += a dt, x += v dt, wrap into the box counting
 * crossings, accumulate ekin and kin_P from v - a dt/2, track the largest
 * displacement since the last rebuild and raise sys->neighb_flag (resetting
 * xn and cross_neighb) when it exceeds skin/2.
 * Replaces include/sepintgr.h:35 (source/sepintgr.c:42-88, 18-40). */
void sep_leapfrog(seppart *ptr, sepsys *sys, sepret *retval);

/* Nose-Hoover: T = 0.666667 * (sum m v^2)/(2N), alpha += dt/Q^2 (T/Td - 1),
 * f -= alpha m v.  Only alpha[0] is used.
 * Replaces include/sepintgr.h:58 (source/sepintgr.c:149-168). */
void sep_nosehoover(seppart *ptr, double Td, double *alpha, const double Q,
		    sepsys *sys);

#ifdef __cplusplus
}
#endif
#endif
