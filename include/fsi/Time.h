/*
 * fsi/Time.h -- a header name the reference's progs (and the progs/common helpers
 * they include) use; the declarations are those of biblbm_plb.h, the
 * device-backed C++ layer of biblbm_amd.
 */
#ifndef BIBLBM_SHIM_FSI_TIME_H
#define BIBLBM_SHIM_FSI_TIME_H
#include "../biblbm_plb.h"
#endif
