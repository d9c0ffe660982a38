/*
 * sep.h -- umbrella header of the MI355X seplib hot-path library (libsep.so).
 * Same role as the reference's include/sep.h:19-31, restricted to the
 * per-timestep path plus the device extension in sepgpu.h.
 */
#ifndef SEP_MI355X_SEP_H
#define SEP_MI355X_SEP_H

#include <stdio.h>
#include <stdlib.h>
#include <math.h>

#include "sepstrct.h"
#include "sepinit.h"
#include "sepprfrc.h"
#include "sepintgr.h"
#include "sepmol.h"
#include "sepmisc.h"
#include "sepret.h"
#include "sepcoulomb.h"
#include "separray.h"
#include "sepomp.h"
#include "sepgpu.h"

#endif
