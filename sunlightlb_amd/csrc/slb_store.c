/*
 * slb_store.c -- checkpoint and field I/O of the sunlightlb.h API
 * (reference src/lbstore.c:36-678).  File formats are byte-compatible with
 * the reference: binary moment files hold 10 LBStoreReal per node (drho, j,
 * s_xx, s_yy, s_zz, s_xy, s_xz, s_yz), k fastest; velocity files hold
 * (rho, u) per node as LBReal (binary) or "%g %g %g %g %i %i %i" text.
 *
 * The host mirrors are synchronised around each call; the bbl / propagate /
 * collide phases these routines take run on the GPU.
 */
#include <stdio.h>
#include <stdlib.h>
#include "slb_internal.h"

/* Moments of one node's shifted populations: drho, j, and the stress
 * s = Pi - drho/3 I (same moment set as the collision). */
static void node_moments(const LBReal *a, LBReal *drho, LBReal j[3], LBReal s[6]) {
  LBReal d = 0, jx = 0, jy = 0, jz = 0, pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
  static const int c[15][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
                               {0, 0, 1},  {0, 0, -1}, {1, 1, 1},  {-1, 1, 1},  {1, -1, 1},
                               {-1, -1, 1}, {1, 1, -1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};
  for (int l = 0; l < 15; l++) {
    LBReal v = a[l];
    d += v;
    jx += c[l][0] * v; jy += c[l][1] * v; jz += c[l][2] * v;
    pxx += c[l][0] * c[l][0] * v; pyy += c[l][1] * c[l][1] * v; pzz += c[l][2] * c[l][2] * v;
    pxy += c[l][0] * c[l][1] * v; pxz += c[l][0] * c[l][2] * v; pyz += c[l][1] * c[l][2] * v;
  }
  *drho = d;
  j[0] = jx; j[1] = jy; j[2] = jz;
  s[0] = pxx - d / 3.0; s[1] = pyy - d / 3.0; s[2] = pzz - d / 3.0;
  s[3] = pxy; s[4] = pxz; s[5] = pyz;
}

/* Populations from (drho, j, s): a_l = (w_l/72)[drho + 3 c.j + 4.5(cc:s - tr/3)]
 * (the rebuild of src/lbstore.c:145-197, :375-399). */
static void node_rebuild(LBReal *a, LBReal drho, const LBReal j[3], const LBReal s[6],
                         LBReal weight, int add) {
  static const int c[15][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
                               {0, 0, 1},  {0, 0, -1}, {1, 1, 1},  {-1, 1, 1},  {1, -1, 1},
                               {-1, -1, 1}, {1, 1, -1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};
  static const int w[15] = {16, 8, 8, 8, 8, 8, 8, 1, 1, 1, 1, 1, 1, 1, 1};
  LBReal tr3 = (s[0] + s[1] + s[2]) / 3.0;
  for (int l = 0; l < 15; l++) {
    LBReal cc = c[l][0] * c[l][0] * s[0] + c[l][1] * c[l][1] * s[1] + c[l][2] * c[l][2] * s[2] +
                2.0 * (c[l][0] * c[l][1] * s[3] + c[l][0] * c[l][2] * s[4] + c[l][1] * c[l][2] * s[5]);
    LBReal v = w[l] / 72.0 * (drho + 3.0 * (c[l][0] * j[0] + c[l][1] * j[1] + c[l][2] * j[2]) +
                              4.5 * (cc - tr3));
    if (add) a[l] += weight * v; else a[l] = v;
  }
}

/* The .dms record of one post-propagate node (src/lbstore.c:56-84): the
 * moments from partial sums in LBReal, each stored as LBStoreReal, with the
 * diagonal stresses built on the density already rounded to LBStoreReal
 * (the format's writer keeps drho in an LBStoreReal variable before it
 * enters s_aa = 2/3 drho - (a_0 + axis pairs) + (pair along a)).  Links 7..10
 * have c_z = +1, 11..14 c_z = -1; "up"/"dn" are their sums weighted by 1,
 * c_x, c_y and c_x c_y. */
static void dms_moments(const LBReal *a, LBStoreReal v[10]) {
  const LBReal px = a[1] + a[2], py = a[3] + a[4], pz = a[5] + a[6];
  const LBReal up = a[7] + a[8] + a[9] + a[10], dn = a[11] + a[12] + a[13] + a[14];
  const LBReal up_x = a[7] - a[8] + a[9] - a[10], dn_x = a[11] - a[12] + a[13] - a[14];
  const LBReal up_y = a[7] + a[8] - a[9] - a[10], dn_y = a[11] + a[12] - a[13] - a[14];
  const LBReal up_xy = a[7] - a[8] - a[9] + a[10], dn_xy = a[11] - a[12] - a[13] + a[14];
  const LBReal rest = a[0] + px + py + pz;
  const LBStoreReal drho = (LBStoreReal)(rest + (up + dn));
  const LBReal shift = (2.0 / 3.0) * drho - rest;
  v[0] = drho;
  v[1] = (LBStoreReal)(a[1] - a[2] + up_x + dn_x);
  v[2] = (LBStoreReal)(a[3] - a[4] + up_y + dn_y);
  v[3] = (LBStoreReal)(a[5] - a[6] + up - dn);
  v[4] = (LBStoreReal)(shift + px);
  v[5] = (LBStoreReal)(shift + py);
  v[6] = (LBStoreReal)(shift + pz);
  v[7] = (LBStoreReal)(up_xy + dn_xy);
  v[8] = (LBStoreReal)(up_x - dn_x);
  v[9] = (LBStoreReal)(up_y - dn_y);
}

/* One node rebuilt from its .dms record as the format's reader builds it
 * (src/lbstore.c:138-197): a_l = (w_l/72)[drho + 3 c.j + 9/2 (cc:s - tr/3)],
 * where the trace and the momentum combinations of the diagonal links are
 * summed in LBStoreReal (the reader adds its LBStoreReal variables before
 * promoting them), the rest in LBReal; the momenta were multiplied by
 * rescale into LBStoreReal first.  weight/add: load_average. */
static void dms_rebuild(LBReal *a, const LBStoreReal v[10], LBReal weight, int add) {
  const LBStoreReal jx = v[1], jy = v[2], jz = v[3];
  const LBStoreReal sxx = v[4], syy = v[5], szz = v[6], sxy = v[7], sxz = v[8], syz = v[9];
  const LBReal r1 = v[0] / 72.0, r8 = 8.0 * r1;
  const LBReal third = 8.0 * (1.0 / 24.0), d24 = 1.0 / 24.0;
  const LBReal tr3 = (LBStoreReal)(sxx + syy + szz) / 3.0;
  LBReal o[15];
  o[0] = 2.0 * r8 - tr3;
  LBReal g = third * jx, h = 0.5 * (sxx - tr3);
  o[1] = r8 + g + h; o[2] = r8 - g + h;
  g = third * jy; h = 0.5 * (syy - tr3);
  o[3] = r8 + g + h; o[4] = r8 - g + h;
  g = third * jz; h = 0.5 * (szz - tr3);
  o[5] = r8 + g + h; o[6] = r8 - g + h;
  g = d24 * (LBStoreReal)(jx + jy + jz); h = 0.125 * (tr3 + sxy + sxz + syz);
  o[7] = r1 + g + h; o[14] = r1 - g + h;
  g = d24 * (LBStoreReal)(jx - jy - jz); h = 0.125 * (tr3 - sxy - sxz + syz);
  o[8] = r1 - g + h; o[13] = r1 + g + h;
  g = d24 * (LBStoreReal)(jx - jy + jz); h = 0.125 * (tr3 - sxy + sxz - syz);
  o[9] = r1 + g + h; o[12] = r1 - g + h;
  g = d24 * (LBStoreReal)(jx + jy - jz); h = 0.125 * (tr3 + sxy - sxz - syz);
  o[10] = r1 - g + h; o[11] = r1 + g + h;
  for (int l = 0; l < 15; l++) a[l] = add ? a[l] + weight * o[l] : o[l];
}

/* lbsave: bbl + propagate, write the 10 moments, collide (src/lbstore.c:36-91). */
void lbsave(LBSys lb, char *savefile) {
  FILE *fp = fopen(savefile, "wb");
  if (fp == NULL) {
    printf("savefile: %s could not be opened\n", savefile);
    return;
  }
  printf("\nSaving density, momentum, stress to %s\n", savefile);
  bbl(lb);
  propagate(lb);
  slb_host_pops(lb);
  for (long n = 0; n < lb->nnodes; n++) {
    LBStoreReal v[10];
    dms_moments(lb->a + 15 * n, v);
    fwrite(v, sizeof(LBStoreReal), 10, fp);
  }
  fclose(fp);
  collide(lb);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : lbsave done\n", lb->step);
}

/* lbload: rebuild populations from the 10 moments (or average them in with
 * load_average), j scaled by rescale, then collide (src/lbstore.c:102-218). */
void lbload(LBSys lb, char *loadfile) {
  if (lb->changed_objects) syncobjs(lb);
  FILE *fp = fopen(loadfile, "r");
  if (fp == NULL) {
    printf("lbload: %s could not be opened\n", loadfile);
    exit(EXIT_FAILURE);
  }
  if (lb->load_average) printf("\nAveraging density, momentum, stress with %s\n", loadfile);
  else printf("\nLoading density, momentum, stress from %s\n", loadfile);
  if (lb->rescale != 1.0) {
    printf("Multiplying velocity field (only) by %f\n", lb->rescale);
    printf("*** this factor will be reset to unity at end of loading ***\n");
  }
  slb_host_pops_w(lb);
  for (long n = 0; n < lb->nnodes; n++) {
    LBStoreReal v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (fread(v, sizeof(LBStoreReal), 10, fp) != 10) { /* short file: zeros, as fread leaves */ }
    for (int q = 1; q <= 3; q++) v[q] = (LBStoreReal)(lb->rescale * v[q]);
    dms_rebuild(lb->a + 15 * n, v, 0.5, lb->load_average);
  }
  fclose(fp);
  if (lb->skip_solid)
    for (long n = 0; n < lb->nnodes; n++)
      if (lb->obj[n] != NULL)
        for (int l = 0; l < 15; l++) lb->a[15 * n + l] = 0.0;
  lb->status = POST_PROPAGATE;
  collide(lb);
  lb->rescale = 1.0;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : lbload done\n", lb->step);
}

/* lbsavevel: (rho, u) per node with the force correction taken, as in the
 * reference, from the node force field fx (src/lbstore.c:227-296). */
void lbsavevel(LBSys lb, char *savefile, char *type) {
  int bin = (type[0] == 'b' || type[0] == 'B');
  FILE *fp = fopen(savefile, "w");
  if (fp == NULL) {
    printf("savefile: %s could not be opened\n", savefile);
    return;
  }
  double fac = 0.0;
  if (lb->vel_corrected) {
    printf("\nSaving corrected velocity field in %s to %s", bin ? "binary" : "ascii", savefile);
    switch (lb->status) {
      case POST_BBL: printf(" [post-bounceback - not a good idea]\n"); break;
      case POST_PROPAGATE: fac = 0.5; printf("\n"); break;
      case POST_COLLIDE: fac = -0.5; printf("\n"); break;
    }
  } else {
    printf("\nSaving ");
    switch (lb->status) {
      case POST_BBL: printf("post-bounceback [not a good idea]"); break;
      case POST_PROPAGATE: printf("post-propagation"); break;
      case POST_COLLIDE: printf("post-collision"); break;
    }
    printf(" uncorrected velocity field in %s to %s\n", bin ? "binary" : "ascii", savefile);
  }
  slb_host_pops(lb);
  slb_host_f(lb);
  LBReal rho = 0.0;
  for (int i = 0; i < lb->esi; i++)
    for (int j = 0; j < lb->esj; j++)
      for (int k = 0; k < lb->esk; k++) {
        long n = (long)lb->esk * ((long)lb->esj * i + j) + k;
        LBReal ux, uy, uz;
        if (lb->obj[n] != NULL && lb->skip_solid) {
          ux = uy = uz = 0.0;
        } else {
          LBReal d, jj[3], s[6];
          node_moments(lb->a + 15 * n, &d, jj, s);
          rho = lb->rho0 + d;
          ux = (jj[0] + fac * lb->fx[n]) / rho;
          uy = (jj[1] + fac * lb->fy[n]) / rho;
          uz = (jj[2] + fac * lb->fz[n]) / rho;
        }
        if (bin) {
          fwrite(&rho, sizeof(LBReal), 1, fp);
          fwrite(&ux, sizeof(LBReal), 1, fp);
          fwrite(&uy, sizeof(LBReal), 1, fp);
          fwrite(&uz, sizeof(LBReal), 1, fp);
        } else {
          fprintf(fp, "%g %g %g %g %i %i %i\n", rho, ux, uy, uz, i, j, k);
        }
      }
  fclose(fp);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : lbsavevel done\n", lb->step);
}

/* lbloadvel: equilibrium populations from (rho, u); stress zero for Stokes
 * flow, u j otherwise; then collide (src/lbstore.c:305-404).  This is the
 * warm-start injection point. */
void lbloadvel(LBSys lb, char *loadfile, char *type) {
  int bin = (type[0] == 'b' || type[0] == 'B');
  if (lb->changed_objects) syncobjs(lb);
  if (myscmp("null", loadfile)) {
    printf("\nNo load file\n");
    return;
  }
  printf("\nLoading velocity field from %s file %s\n", bin ? "binary" : "ascii", loadfile);
  printf(" (approximate LB initialisation)\n");
  FILE *fp = fopen(loadfile, "r");
  if (fp == NULL) {
    printf("lbloadvel: %s could not be opened\n", loadfile);
    exit(EXIT_FAILURE);
  }
  if (lb->rescale != 1.0) {
    printf("Multiplying velocity field by %f\n", lb->rescale);
    printf("*** this factor will be reset to unity at end of loading ***\n");
  }
  slb_host_pops_w(lb);
  for (long n = 0; n < lb->nnodes; n++) {
    LBReal rho = 0, ux = 0, uy = 0, uz = 0;
    if (bin) {
      LBReal v[4] = {0, 0, 0, 0};
      if (fread(v, sizeof(LBReal), 4, fp) == 4) { rho = v[0]; ux = v[1]; uy = v[2]; uz = v[3]; }
    } else {
      float tr = 0, tx = 0, ty = 0, tz = 0;
      int i, j, k;
      if (fscanf(fp, "%f%f%f%f%i%i%i", &tr, &tx, &ty, &tz, &i, &j, &k) != 7) {}
      rho = tr; ux = tx; uy = ty; uz = tz;
    }
    ux *= lb->rescale; uy *= lb->rescale; uz *= lb->rescale;
    LBReal jj[3] = {rho * ux, rho * uy, rho * uz}, s[6] = {0, 0, 0, 0, 0, 0};
    if (!lb->stokes_flow) {
      s[0] = ux * jj[0]; s[1] = uy * jj[1]; s[2] = uz * jj[2];
      s[3] = ux * jj[1]; s[4] = ux * jj[2]; s[5] = uy * jj[2];
    }
    node_rebuild(lb->a + 15 * n, rho - lb->rho0, jj, s, 1.0, 0);
  }
  fclose(fp);
  if (lb->skip_solid)
    for (long n = 0; n < lb->nnodes; n++)
      if (lb->obj[n] != NULL)
        for (int l = 0; l < 15; l++) lb->a[15 * n + l] = 0.0;
  lb->status = POST_PROPAGATE;
  collide(lb);
  lb->rescale = 1.0;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : lbloadvel done\n", lb->step);
}

/* lbsavestr: stress averaged over the post-propagate and post-collide states
 * of one step, plus a void-space flag (src/lbstore.c:412-515).  The binary
 * records hold the stresses converted to LBStoreReal; the reference writes
 * the first sizeof(LBStoreReal) bytes of each LBReal instead (:88-93), which
 * with float storage of double stresses is not the value. */
void lbsavestr(LBSys lb, char *savefile, char *type) {
  long N = lb->nnodes;
  LBReal *acc = (LBReal *)malloc(sizeof(LBReal) * 6 * (size_t)N);
  if (acc == NULL) {
    printf("lbsavestr: ran out of memory for temporary storage\n");
    return;
  }
  bbl(lb);
  propagate(lb);
  slb_host_pops(lb);
  for (long n = 0; n < N; n++) {
    LBReal d, j[3], s[6];
    node_moments(lb->a + 15 * n, &d, j, s);
    for (int q = 0; q < 6; q++) acc[6 * n + q] = s[q];
  }
  collide(lb);
  slb_host_pops(lb);
  for (long n = 0; n < N; n++) {
    LBReal d, j[3], s[6];
    node_moments(lb->a + 15 * n, &d, j, s);
    for (int q = 0; q < 6; q++) acc[6 * n + q] += s[q];
  }
  int bin = (type[0] == 'b' || type[0] == 'B');
  FILE *fp = fopen(savefile, "wb");
  if (fp == NULL) {
    printf("savefile: %s could not be opened\n", savefile);
    free(acc);
    return;
  }
  printf("\nSaving stress field in %s to %s\n", bin ? "binary" : "ascii", savefile);
  for (long n = 0; n < N; n++) {
    LBReal t[6];
    for (int q = 0; q < 6; q++)
      t[q] = (lb->obj[n] != NULL && lb->skip_solid) ? 0.0 : 0.5 * acc[6 * n + q];
    int void_space = lb->obj[n] == NULL;
    if (bin) {
      for (int q = 0; q < 6; q++) {
        LBStoreReal v = (LBStoreReal)t[q];
        fwrite(&v, sizeof(LBStoreReal), 1, fp);
      }
      fwrite(&void_space, sizeof(int), 1, fp);
    } else {
      fprintf(fp, "%f\t%f\t%f\t%f\t%f\t%f\t%i\n", t[0], t[1], t[2], t[3], t[4], t[5], void_space);
    }
  }
  fclose(fp);
  free(acc);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : lbsavestr done\n", lb->step);
}

/* Object state as text records POS/VEL/ANG/FRC/TRQ (src/lbstore.c:523-546). */
void lbsaveobj(LBSys lb, char *savefile) {
  int more = 0, opened = 0;
  FILE *fp;
  if (myscmp("stdout", savefile)) { fp = stdout; more = 1; }
  else if (myscmp("stderr", savefile)) { fp = stderr; more = 1; }
  else if ((fp = fopen(savefile, "w")) == NULL) {
    printf("lbsaveobj: %s could not be opened\n", savefile);
    return;
  } else {
    opened = 1;
  }
  printf("\nSaving object data to %s\n", savefile);
  if (more) fprintf(fp, "START of particle data: TYPE, id, base id, x, y, z\n\n");
  for (LBObj p = lb->obj_head; p != NULL; p = p->next_obj) {
    int id = p->id, bid = p->base_obj->id;
    fprintf(fp, "POS %3i %3i %f %f %f\n", id, bid, p->posx, p->posy, p->posz);
    fprintf(fp, "VEL %3i %3i %f %f %f\n", id, bid, p->velx, p->vely, p->velz);
    fprintf(fp, "ANG %3i %3i %f %f %f\n", id, bid, p->angx, p->angy, p->angz);
    fprintf(fp, "FRC %3i %3i %f %f %f\n", id, bid, p->frcx, p->frcy, p->frcz);
    fprintf(fp, "TRQ %3i %3i %f %f %f\n", id, bid, p->trqx, p->trqy, p->trqz);
    if (more) fprintf(fp, "\n");
  }
  if (more) fprintf(fp, "END of particle data\n");
  if (opened) fclose(fp);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : lbsaveobj done\n", lb->step);
}

/* Read POS/VEL/ANG back (FRC/TRQ skipped); velocities scaled by rescale
 * (src/lbstore.c:554-586). */
void lbloadobj(LBSys lb, char *loadfile) {
  FILE *fp = fopen(loadfile, "r");
  if (fp == NULL) {
    printf("lbloadobj: %s could not be opened\n", loadfile);
    return;
  }
  printf("Loading object data from %s\n", loadfile);
  if (lb->rescale != 1.0) {
    printf("Multiplying velocities and angular velocities by %f\n", lb->rescale);
    printf("*** this factor will be reset to unity at end of loading ***\n");
  }
  int irc = 0, i1, i2;
  float t1, t2, t3;
  while (fscanf(fp, "%*s%i%i%f%f%f", &i1, &i2, &t1, &t2, &t3) == 5) {
    LBObj p1 = lookupobj(lb, i1);
    if (p1 == NULL) { irc = 1; break; }
    if (i2 != i1) {
      if (lookupobj(lb, i2) == NULL) { irc = 2; break; }
      p1->base_obj->id = i2;
    }
    p1->posx = t1; p1->posy = t2; p1->posz = t3;
    if (fscanf(fp, "%*s%i%i%f%f%f", &i1, &i2, &t1, &t2, &t3) != 5) { irc = 3; break; }
    p1->velx = lb->rescale * t1; p1->vely = lb->rescale * t2; p1->velz = lb->rescale * t3;
    if (fscanf(fp, "%*s%i%i%f%f%f", &i1, &i2, &t1, &t2, &t3) != 5) { irc = 4; break; }
    p1->angx = lb->rescale * t1; p1->angy = lb->rescale * t2; p1->angz = lb->rescale * t3;
    if (fscanf(fp, "%*s%i%i%f%f%f", &i1, &i2, &t1, &t2, &t3) != 5) { irc = 5; break; }
    if (fscanf(fp, "%*s%i%i%f%f%f", &i1, &i2, &t1, &t2, &t3) != 5) { irc = 6; break; }
  }
  if (irc == 1 || irc == 2) printf("UFO found (unidentified floating object)\n");
  if (irc > 3) printf("Ran out of data apparently\n");
  fclose(fp);
  lb->rescale = 1.0;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : lbloadobj done\n", lb->step);
}

/* Node forces: binary / ascii per node, or a list of boundary nodes with
 * their object (src/lbstore.c:595-640). */
void lbsavefrc(LBSys lb, char *frcfile, char *type) {
  int bin = (type[0] == 'b' || type[0] == 'B'), list = (type[0] == 'l' || type[0] == 'L');
  FILE *fp = fopen(frcfile, "w");
  if (fp == NULL) {
    printf("lbsavefrc: %s could not be opened\n", frcfile);
    exit(EXIT_FAILURE);
  }
  slb_host_f(lb);
  slb_host_mask(lb);
  if (list) {
    printf("\nSaving force data as a list to %s\n", frcfile);
    for (int i = 0; i < lb->esi; i++)
      for (int j = 0; j < lb->esj; j++)
        for (int k = 0; k < lb->esk; k++) {
          long n = (long)lb->esk * ((long)lb->esj * i + j) + k;
          if (lb->mask[n] > 0) {
            LBObj p = lb->obj[n];
            fprintf(fp, "%i %i %i %f %f %f %i %f %f %f\n", i, j, k, lb->fx[n], lb->fy[n],
                    lb->fz[n], p->id, p->posx, p->posy, p->posz);
          }
        }
  } else {
    printf("\nSaving %s force data to %s\n", bin ? "binary" : "ascii", frcfile);
    for (long n = 0; n < lb->nnodes; n++) {
      LBStoreReal f[3] = {(LBStoreReal)lb->fx[n], (LBStoreReal)lb->fy[n], (LBStoreReal)lb->fz[n]};
      if (bin) fwrite(f, sizeof(LBStoreReal), 3, fp);
      else fprintf(fp, "%f %f %f\n", f[0], f[1], f[2]);
    }
  }
  fclose(fp);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : lbsavefrc done\n", lb->step);
}

/* Solid map as a node list or a byte per node, 0 fluid / 255 solid
 * (src/lbstore.c:647-678). */
void lbsavemap(LBSys lb, char *mapfile, char *type) {
  FILE *fp = fopen(mapfile, "w");
  if (fp == NULL) {
    printf("lbsavemap: %s could not be opened\n", mapfile);
    exit(EXIT_FAILURE);
  }
  if (type[0] == 'l' || type[0] == 'L') {
    printf("\nSaving list of solid nodes to %s\n", mapfile);
    for (int i = 0; i < lb->esi; i++)
      for (int j = 0; j < lb->esj; j++)
        for (int k = 0; k < lb->esk; k++)
          if (lb->obj[(long)lb->esk * ((long)lb->esj * i + j) + k] != NULL)
            fprintf(fp, "%i %i %i\n", i, j, k);
  } else {
    printf("\nSaving object map as character array to %s\n", mapfile);
    for (long n = 0; n < lb->nnodes; n++) fputc(lb->obj[n] == NULL ? 0 : 255, fp);
  }
  fclose(fp);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : lbsavemap done\n", lb->step);
}
