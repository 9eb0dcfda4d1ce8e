/*
 * Single sphere in a periodic box, the reference's sphere example in its
 * current API: 32^3 lattice, sphere radius 4.5 at the box centre
 * (15.5, 15.5, 15.5), uniform body force 1e-5 along x on fluid nodes,
 * Stokes flow, default tau 0.75, 2000 steps, then measure().
 *
 *   gcc -O2 -Iinclude examples/config1.c -Lsunlightlb_amd -lsunlightlb -lm
 *
 * Set SLB_PRECISION=32 for the FP32 shifted-population mode.
 */
#include <stdio.h>
#include <stdlib.h>
#include "sunlightlb.h"

int main(int argc, char **argv) {
  int nsteps = argc > 1 ? atoi(argv[1]) : 2000;
  LBSys lb = newlb(32, 32, 32, "null");
  lb->stokes_flow = 1;
  LBObj sph = insertsph(lb, 4.5, 15.5, 15.5, 15.5);
  zerobf(lb);
  addbf(lb, 1e-5, 0.0, 0.0);
  for (int i = 0; i < nsteps; i++) lbstep(lb);
  measure(lb);
  printf("drag %.17g vmean %.17g discrepancy %.6f\n", sph->frcx, lb->vmeanx, discrf(lb));
  return 0;
}
