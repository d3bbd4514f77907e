/*
 * palabos3D.h -- the header the reference's progs include first
 * (progs/capsule_in_shear/capsule_in_shear.cpp:4): here the device-backed
 * C++ layer of biblbm_amd (biblbm_plb.h), which holds the Palabos and fsi
 * names the progs use on the coupled lattice / immersed-boundary path.
 */
#ifndef BIBLBM_PALABOS3D_H
#define BIBLBM_PALABOS3D_H
#include "biblbm_plb.h"
#endif
