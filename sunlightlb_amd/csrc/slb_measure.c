/*
 * slb_measure.c -- flow and force measurement of the sunlightlb.h API on the
 * GPU: measure, accvmean, vzero, vgrab, vsave, zeroforce, normforce,
 * distribf, wrmeas, discrf (reference src/lbmeasure.c:30-332).
 *
 * Node velocities and densities are reduced in FP64 on the device
 * (fixed-order two-stage sums); per-node momentum-exchange forces are FP64
 * and per-object force/torque sums are FP64.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include "slb_internal.h"

/* measure(): two force-collecting steps with velocity weights 1/4, 1/2, 1/4
 * and the node force halved (src/lbmeasure.c:30-49). */
void measure(LBSys lb) {
  if (lb->status != POST_COLLIDE) {
    printf("Measurements apparently out of position\n");
    wrstatus(lb, stdout);
  }
  zeroforce(lb);
  lb->vmeanx = lb->vmeany = lb->vmeanz = 0.0;
  accvmean(lb, 0.25);
  slb_forcestep(lb);
  accvmean(lb, 0.5);
  slb_forcestep(lb);
  accvmean(lb, 0.25);
  normforce(lb, 0.5);
  distribf(lb);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : measure done\n", lb->step);
}

/* Velocity correction factor from the cycle position (src/lbmeasure.c:71-77). */
static double vel_fac(LBSys lb, const char *who) {
  double fac = 0.0;
  if (lb->vel_corrected) switch (lb->status) {
    case POST_PROPAGATE: fac = 0.5; break;
    case POST_COLLIDE: fac = -0.5; break;
    default:
      printf("%s apparently of position\n", who);
      wrstatus(lb, stdout);
      break;
  }
  return fac;
}

void accvmean(LBSys lb, LBReal prefac) {
  slbd_params P;
  double s[5];
  double fac = vel_fac(lb, "Accvmean");
  slb_params(lb, &P);
  slb_dev_geom(lb);
  slb_dev_pops(lb);
  slbd_vmean(PRIV(lb)->dev, &P, fac, lb->psi, s);
  /* sum of weights renormalised to the fluid count (src/lbmeasure.c:96-98) */
  double wfac = s[4] / s[3];
  lb->vmeanx += prefac * (s[0] * wfac) / (lb->nnodes);
  lb->vmeany += prefac * (s[1] * wfac) / (lb->nnodes);
  lb->vmeanz += prefac * (s[2] * wfac) / (lb->nnodes);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : accvmean done\n", lb->step);
}

void vzero(LBSys lb) {
  slb_dev_geom(lb);
  slb_dev_u(lb);
  slbd_vzero(PRIV(lb)->dev);
  PRIV(lb)->u_loc = LOC_DEV;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : vzero done\n", lb->step);
}

void vgrab(LBSys lb, LBReal prefac) {
  slbd_params P;
  double fac = vel_fac(lb, "vgrab");
  slb_params(lb, &P);
  slb_dev_geom(lb);
  slb_dev_pops(lb);
  slb_dev_u(lb);
  slbd_vgrab(PRIV(lb)->dev, &P, fac, prefac);
  PRIV(lb)->u_loc = LOC_DEV;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : vgrab done\n", lb->step);
}

/* Text dump of the measured velocity field (src/lbmeasure.c:173-203). */
void vsave(LBSys lb, char *savefile) {
  FILE *fp = fopen(savefile, "w");
  if (fp == NULL) {
    printf("savefile: %s could not be opened\n", savefile);
    return;
  }
  slb_host_u(lb);
  double fac = 12.0 / lb->rho0;
  for (int i = 0; i < lb->esi; i++)
    for (int j = 0; j < lb->esj; j++)
      for (int k = 0; k < lb->esk; k++) {
        long n = (long)lb->esk * ((long)lb->esj * i + j) + k;
        LBReal ux, uy, uz;
        if (lb->obj[n] != NULL) {
          if (lb->skip_solid) ux = uy = uz = 0.0;
          else { ux = fac * lb->ux[n]; uy = fac * lb->uy[n]; uz = fac * lb->uz[n]; }
        } else {
          ux = lb->ux[n]; uy = lb->uy[n]; uz = lb->uz[n];
        }
        fprintf(fp, "%g %g %g %i %i %i\n", ux, uy, uz, i, j, k);
      }
  fclose(fp);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : lbsavevel done\n", lb->step);
}

void zeroforce(LBSys lb) {
  slbd_zero_f(PRIV(lb)->dev);
  PRIV(lb)->f_loc = LOC_DEV;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : zeroforce done\n", lb->step);
}

void normforce(LBSys lb, LBReal fac) {
  slb_dev_f(lb);
  slbd_scale_f(PRIV(lb)->dev, fac);
  PRIV(lb)->f_loc = LOC_DEV;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : normforce(%f) done\n", lb->step, fac);
}

/* Per-object force and torque (src/lbmeasure.c:236-263): zero every
 * object, then sum the node forces of each object's nodes, torque about the
 * object centre without minimum image. */
void distribf(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  for (LBObj p = lb->obj_head; p != NULL; p = p->next_obj) {
    p->frcx = p->frcy = p->frcz = 0.0;
    p->trqx = p->trqy = p->trqz = 0.0;
  }
  slb_dev_geom(lb);
  slb_dev_f(lb);
  slb_objtab(lb);
  if (S->nobj > 0) {
    double *frc = (double *)slb_xmalloc(sizeof(double) * 3 * S->nobj);
    double *trq = (double *)slb_xmalloc(sizeof(double) * 3 * S->nobj);
    slbd_distribf(S->dev, S->nobj, S->objtab, frc, trq);
    for (int o = 0; o < S->nobj; o++) {
      LBObj p = S->objlist[o];
      p->frcx = frc[3 * o]; p->frcy = frc[3 * o + 1]; p->frcz = frc[3 * o + 2];
      p->trqx = trq[3 * o]; p->trqy = trq[3 * o + 1]; p->trqz = trq[3 * o + 2];
    }
    free(frc); free(trq);
  }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : distribf done\n", lb->step);
}

/* Expected total force: the body force summed over fluid nodes, including
 * the live uniform part dpdx/dpdy/dpdz. */
static void expected_force(LBSys lb, double *fx, double *fy, double *fz) {
  double sx = 0.0, sy = 0.0, sz = 0.0;
  long nf = 0, N = lb->nnodes;
  for (long n = 0; n < N; n++)
    if (lb->obj[n] == NULL) {
      sx += lb->bx[n]; sy += lb->by[n]; sz += lb->bz[n];
      nf++;
    }
  *fx = sx + nf * (double)lb->dpdx;
  *fy = sy + nf * (double)lb->dpdy;
  *fz = sz + nf * (double)lb->dpdz;
}

void wrmeas(LBSys lb) {
  double fxtot = 0.0, fytot = 0.0, fztot = 0.0, fxsum, fysum, fzsum;
  printf("\n\nProperties at lb steps %i to %i\n", ((lb->step) - 2), (lb->step));
  printf(" mean flow rate = %0.6f %0.6f %0.6f\n", lb->vmeanx, lb->vmeany, lb->vmeanz);
  for (LBObj p = lb->obj_head; p != NULL; p = p->next_obj) {
    if (slb_verbosity & FORCE_DETAILS) {
      printf(" object id %3i (base id %3i), force = %8.3f %8.3f %8.3f\n", p->id,
             p->base_obj->id, p->frcx, p->frcy, p->frcz);
      printf("                             torque = %8.3f %8.3f %8.3f\n", p->trqx, p->trqy,
             p->trqz);
    }
    fxtot += p->frcx; fytot += p->frcy; fztot += p->frcz;
  }
  printf("            total force = %8.3f %8.3f %8.3f\n", fxtot, fytot, fztot);
  expected_force(lb, &fxsum, &fysum, &fzsum);
  printf("  total force should be = %8.3f %8.3f %8.3f\n", fxsum, fysum, fzsum);
  double dx = fxtot - fxsum, dy = fytot - fysum, dz = fztot - fzsum;
  double discr = sqrt(dx * dx + dy * dy + dz * dz);
  double norm = sqrt(fxsum * fxsum + fysum * fysum + fzsum * fzsum);
  if (norm != 0.0) discr = discr / norm;
  printf(" force discrepancy = %0.4f (norm = %0.4f)\n", discr, norm);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : wrforce done\n", lb->step);
}

LBReal discrf(LBSys lb) {
  double fxtot = 0.0, fytot = 0.0, fztot = 0.0, fxsum, fysum, fzsum;
  for (LBObj p = lb->obj_head; p != NULL; p = p->next_obj) {
    fxtot += p->frcx; fytot += p->frcy; fztot += p->frcz;
  }
  expected_force(lb, &fxsum, &fysum, &fzsum);
  double dx = fxtot - fxsum, dy = fytot - fysum, dz = fztot - fzsum;
  LBReal discr = sqrt(dx * dx + dy * dy + dz * dz);
  double norm = sqrt(fxsum * fxsum + fysum * fysum + fzsum * fzsum);
  if (norm != 0.0) discr = discr / norm;
  if (slb_verbosity & TRACE_ROUTINE_EXITS)
    printf("LB %i : wrforce done, discr = %0.4f\n", lb->step, discr);
  return discr;
}
