/*
 * sepdef.h -- constants of the seplib C99 API as served by the MI355X
 * library (libsep.so).
 *
 * Values must equal the reference's so that existing callers compiled
 * against the reference headers behave the same when linked against this
 * library.  Reference: include/sepdef.h:14-64.
 */
#ifndef SEP_MI355X_SEPDEF_H
#define SEP_MI355X_SEPDEF_H

#define SEP_FALSE 0
#define SEP_TRUE 1

/* Per-atom capacities of the bonded-partner arrays (include/sepdef.h:20-22). */
#define SEP_BOND 10
#define SEP_ANGLE 10
#define SEP_DIHED 20

/* Per-atom neighbour capacity used by sep_init(N, SEP_NEIGHB)
 * (include/sepdef.h:24-25). */
#define SEP_NEIGHB 3000
#define SEP_NUM_NEIGHB 3000
#define SEP_NO_NEIGHB 0

/* Exclusion options for the pair functions (include/sepdef.h:29-39). */
#define SEP_NEIGHB_ALL 1
#define SEP_ALL 1
#define SEP_NEIGHB_EXCL_BONDED 2
#define SEP_EXCL_BONDED 2
#define SEP_NEIGHB_EXCL_SAME_MOL 3
#define SEP_EXCL_SAME_MOL 3

#define SEP_MAX_NUM_MOL 10000

/* The reference's truncated pi and its LJ constants (include/sepdef.h:41-45).
 * SEP_LJCF2 is -4(rc^-12 - rc^-6) at rc = 2.5 rounded to 12 digits; the
 * shifted-LJ energy adds it verbatim, so it must stay bit-identical. */
#define SEP_PI 3.14159265358979
#define SEP_WCACF 1.12246204830937
#define SEP_LJCF2 0.016316891136

/* Force-evaluation algorithms selected in sep_sys_setup
 * (include/sepdef.h:53-55). */
#define SEP_BRUTE 0
#define SEP_NEIGHBLIST 1
#define SEP_LLIST_NEIGHBLIST 2

#define SEP_LEAPFROG 0
#define SEP_NOSEHOOVER 1

/* The reference defines both as 1 (include/sepdef.h:63-64); kept so that
 * callers comparing against them compile and behave identically. */
#define SEP_SUCCESS 1
#define SEP_FAILURE 1

#endif
