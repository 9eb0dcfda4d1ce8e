/*
 * slb_mp.c -- moment propagation (passive scalar) of the sunlightlb.h API
 * on the GPU, plus its setup, diagnostics and adsorption-site helpers
 * (reference src/mpkernel.c:33-43, src/mpmisc.c:38-499, src/mpobjs.c:28-261,
 * src/mpads.c:37-186).
 *
 * The tracer field P of each MP system lives on the device (ping-pong pair);
 * mp->p is its host mirror.  mp->newp is kept allocated for layout
 * compatibility; the gather kernel does not need it.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "slb_internal.h"

static time_t mp_tstart;

void slb_host_p(MPSys mp) {
  slb_mppriv *M = MPRIV(mp);
  if (M->p_loc == LOC_DEV) {
    slbd_mp_get(PRIV(mp->lb)->dev, M->dev, 0, mp->p);
    M->p_loc = LOC_SYNC;
  }
}
void slb_host_p_w(MPSys mp) { slb_host_p(mp); MPRIV(mp)->p_loc = LOC_HOST; }
void slb_dev_p(MPSys mp) {
  slb_mppriv *M = MPRIV(mp);
  if (M->p_loc == LOC_HOST) {
    slbd_mp_put(PRIV(mp->lb)->dev, M->dev, 0, mp->p);
    M->p_loc = LOC_SYNC;
  }
}

/* newmp: tracer and density fields, density reset to rho0 (src/mpmisc.c:38-59).
 * alpha starts at 0 here (the reference leaves it unset). */
MPSys newmp(LBSys lb, char *name) {
  slb_mpwrap *W = (slb_mpwrap *)calloc(1, sizeof(slb_mpwrap));
  if (W == NULL) error("No room for mpsys structure\n");
  MPSys mp = &W->pub;
  long N = lb->nnodes;
  mp->lb = lb;
  mp->p = slb_ralloc(N);
  mp->newp = slb_ralloc(N);
  if (lb->rho != NULL) slb_big_free(lb->rho);
  lb->rho = slb_ralloc(N);
  for (long n = 0; n < N; n++) lb->rho[n] = lb->rho0;
  PRIV(lb)->rho_loc = LOC_HOST;
  mp->objects_moving = lb->objects_moving;
  mp->to_be_initialised = 1;
  mp->alpha = 0.0;
  if (myscmp("null", name)) mp->name = NULL; else mp->name = mycopy(name);
  W->priv.nmom = 1;
  W->priv.axis = 0;
  W->priv.pads = 0.0;
  W->priv.dev = slbd_mp_create(PRIV(lb)->dev, 1);
  W->priv.p_loc = LOC_SYNC;
  W->priv.next = PRIV(lb)->mps;
  PRIV(lb)->mps = W;
  printf("MP '%s' initialised for %i * %i * %i lattice\n", mp->name, lb->esi, lb->esj, lb->esk);
  return mp;
}

/* Frozen density: rho0 + sum a on fluid nodes, rho0 on solid
 * (src/mpmisc.c:64-79). */
void mprhoinit(MPSys mp) {
  LBSys lb = mp->lb;
  slb_dev_geom(lb);
  slb_dev_pops(lb);
  slbd_mp_rhoinit(PRIV(lb)->dev, lb->rho0);
  PRIV(lb)->rho_loc = LOC_DEV;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP '%s' %i : mprhoinit done\n", mp->name, mp->step);
}

/* mpinit: density, tracer cleared, step reset (src/mpmisc.c:83-99). */
void mpinit(MPSys mp) {
  slb_mppriv *M = MPRIV(mp);
  LBSys lb = mp->lb;
  mprhoinit(mp);
  memset(mp->p, 0, sizeof(LBReal) * (size_t)lb->nnodes);
  for (int f = 0; f < M->nmom; f++) slbd_mp_zero(PRIV(lb)->dev, M->dev, f);
  M->p_loc = LOC_SYNC;
  mp->step = 0;
  mp->to_be_initialised = 0;
  mp_tstart = time(NULL);
  printf("MP '%s' initialised and ready to go.\n", mp->name);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP '%s' %i : mpinit done\n", mp->name, mp->step);
}

void mpwrpars(MPSys mp) {
  LBReal d = (1.0 - mp->alpha) / 6.0;
  printf("\nMP '%s' parameter list\n\n", mp->name);
  printf(" | LB size  = %i * %i * %i = %i\n", mp->lb->esi, mp->lb->esj, mp->lb->esk,
         mp->lb->nnodes);
  printf(" | alpha, D = %0.4f %0.6f\n", mp->alpha, d);
  if (mp->objects_moving) printf(" | Objects allowed to move\n");
  else printf(" | All objects are stationary\n");
  if (MPRIV(mp)->pads > 0.0) printf(" | Partial adsorption probability %g\n", MPRIV(mp)->pads);
  printf("\n");
}

static void mp_steps(MPSys mp, int nsteps) {
  slb_mppriv *M = MPRIV(mp);
  LBSys lb = mp->lb;
  slb_lbpriv *S = PRIV(lb);
  slb_dev_geom(lb);
  slb_dev_pops(lb);
  slb_dev_rho(lb);
  slb_dev_p(mp);
  if (mp->objects_moving) slb_objtab(lb);
  slbd_mp_step(S->dev, M->dev, lb->rho0, mp->alpha, mp->objects_moving, S->nobj, S->objtab,
               M->pads, M->nmom, M->axis, nsteps);
  M->p_loc = LOC_DEV;
}

/* One MP step on the post-collide populations (src/mpkernel.c:33-43). */
void mpstep(MPSys mp) {
  if (mp->to_be_initialised) mpinit(mp);
  if (mp->lb->status != POST_COLLIDE) {
    printf("Moment propagation apparently out of position\n");
    wrstatus(mp->lb, stdout);
  }
  mp_steps(mp, 1);
  if (slb_verbosity & TRACE_ROUTINE_EXITS)
    printf("MP '%s' %i : %s done\n", mp->name, mp->step, mp->objects_moving ? "mpmove" : "mpstat");
#ifdef SIGNAL
  if (slb_signal_flag) sighandle(mp->lb);
#endif
  mp->step += 1;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP '%s' %i : mpdo done\n", mp->name, mp->step);
}

void slb_mp_run(MPSys mp, int nsteps) {
  if (nsteps <= 0) return;
  if (mp->to_be_initialised) mpinit(mp);
  if (mp->lb->status != POST_COLLIDE) {
    printf("Moment propagation apparently out of position\n");
    wrstatus(mp->lb, stdout);
  }
  mp_steps(mp, nsteps);
#ifdef SIGNAL
  if (slb_signal_flag) sighandle(mp->lb);
#endif
  mp->step += nsteps;
}

double slb_mp_time_steps(MPSys mp, int nsteps) {
  slb_mppriv *M = MPRIV(mp);
  LBSys lb = mp->lb;
  if (mp->to_be_initialised) mpinit(mp);
  slb_dev_geom(lb);
  slb_dev_pops(lb);
  slb_dev_rho(lb);
  slb_dev_p(mp);
  double ms = slbd_time_mp(PRIV(lb)->dev, M->dev, lb->rho0, mp->alpha, M->pads, M->nmom, nsteps);
  M->p_loc = LOC_DEV;
  mp->step += nsteps;
  return ms;
}

void slb_mp_sums(MPSys mp, double out[6]) {
  slb_dev_geom(mp->lb);
  slb_dev_p(mp);
  slbd_mp_sums(PRIV(mp->lb)->dev, MPRIV(mp)->dev, out);
}

void slb_mp_set_adsorption(MPSys mp, double pads) {
  if (pads < 0.0) pads = 0.0;
  if (pads > 1.0) pads = 1.0;
  MPRIV(mp)->pads = pads;
}

void slb_mp_track_moments(MPSys mp, int axis) {
  slb_mppriv *M = MPRIV(mp);
  LBSys lb = mp->lb;
  slb_host_p(mp);
  if (M->nmom < 3) {
    slbd_mp_destroy(M->dev);
    M->dev = slbd_mp_create(PRIV(lb)->dev, 3);
    M->nmom = 3;
    M->p_loc = LOC_HOST;
  } else {
    slbd_mp_zero(PRIV(lb)->dev, M->dev, 1);
    slbd_mp_zero(PRIV(lb)->dev, M->dev, 2);
  }
  M->axis = (axis >= 0 && axis < 3) ? axis : 0;
}

void slb_mp_sync_host(MPSys mp) { slb_host_p(mp); slb_host_rho(mp->lb); }
void slb_mp_sync_device(MPSys mp) { MPRIV(mp)->p_loc = LOC_HOST; }

/* ------------------------------------------------------------------ */
/* Diagnostics (src/mpmisc.c:121-246, :466-499).                        */

void mpcheck(MPSys mp) {
  double s[6];
  slb_mp_sums(mp, s);
  printf("MP '%s' %i : fluid + solid scalar = %10.6f + %10.6f = %10.6f\n", mp->name, mp->step,
         s[0], s[1], s[0] + s[1]);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP '%s' %i : mpcheck done\n", mp->name, mp->step);
}

LBReal mpmean(MPSys mp) {
  double s[6];
  slb_mp_sums(mp, s);
  LBReal pmean = s[0] / (LBReal)(countfluid(mp->lb));
  if (slb_verbosity & TRACE_ROUTINE_EXITS)
    printf("MP '%s' %i : mpnorm done, ptot = %f, pmean = %f\n", mp->name, mp->step, s[0], pmean);
  return pmean;
}

static LBReal mp_extreme(MPSys mp, int want_max) {
  slb_host_p(mp);
  LBSys lb = mp->lb;
  int found = 0;
  LBReal e = 0.0;
  for (long n = 0; n < lb->nnodes; n++) {
    if (lb->obj[n] != NULL) continue;
    LBReal v = mp->p[n];
    if (!found) { e = v; found = 1; }
    else if (want_max ? (e < v) : (e > v)) e = v;
  }
  return e;
}

LBReal mpmax(MPSys mp) {
  LBReal v = mp_extreme(mp, 1);
  if (slb_verbosity & TRACE_ROUTINE_EXITS)
    printf("MP '%s' %i : mpnorm done, pmax = %f\n", mp->name, mp->step, v);
  return v;
}

LBReal mpmin(MPSys mp) {
  LBReal v = mp_extreme(mp, 0);
  if (slb_verbosity & TRACE_ROUTINE_EXITS)
    printf("MP '%s' %i : mpnorm done, pmin = %f\n", mp->name, mp->step, v);
  return v;
}

void mprescale(MPSys mp, LBReal fac) {
  slb_host_p_w(mp);
  for (long n = 0; n < mp->lb->nnodes; n++)
    if (mp->lb->obj[n] == NULL) mp->p[n] *= fac;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP '%s' %i : mprescale done\n", mp->name, mp->step);
}

/* Adds the fluid part of mp's tracer into mp_dest (src/mpmisc.c:233-246). */
void mpcopy(MPSys mp, MPSys mp_dest) {
  slb_host_p(mp);
  slb_host_p_w(mp_dest);
  for (long n = 0; n < mp->lb->nnodes; n++)
    if (mp->lb->obj[n] == NULL) mp_dest->p[n] += mp->p[n];
  if (slb_verbosity & TRACE_ROUTINE_EXITS)
    printf("MP '%s' %i : mpcopy done, into MP '%s'\n", mp->name, mp->step, mp_dest->name);
}

void mpsave(MPSys mp, char *file, char *type) {
  int bin = (type[0] == 'b' || type[0] == 'B');
  FILE *fp = fopen(file, "w");
  if (fp == NULL) {
    printf("mpsave: %s could not be opened\n", file);
    return;
  }
  printf("MP '%s' %6i: Saving mp data to %s\n", mp->name, mp->step, file);
  slb_host_p(mp);
  for (long n = 0; n < mp->lb->nnodes; n++) {
    LBStoreReal v = mp->p[n];
    if (bin) fwrite(&v, sizeof(LBStoreReal), 1, fp);
    else fprintf(fp, "%f\n", v);
  }
  fclose(fp);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP '%s' %i : mpsave done\n", mp->name, mp->step);
}

void mpload(MPSys mp, char *file, char *type) {
  int bin = (type[0] == 'b' || type[0] == 'B');
  FILE *fp = fopen(file, "r");
  if (fp == NULL) {
    printf("mpload: %s could not be opened\n", file);
    exit(1);
  }
  printf("MP '%s' %6i: Loading mp data from %s\n", mp->name, mp->step, file);
  slb_host_p_w(mp);
  for (long n = 0; n < mp->lb->nnodes; n++) {
    LBStoreReal v = 0;
    if (bin) { if (fread(&v, sizeof(LBStoreReal), 1, fp) != 1) {} }
    else { if (fscanf(fp, "%f", &v) != 1) {} }
    mp->p[n] = v;
  }
  fclose(fp);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP '%s' %i : mpload done\n", mp->name, mp->step);
}

/* Tracer field writer with unique/append naming and ascii/binary/byte
 * encodings (src/mpmisc.c:329-398). */
void mpconwr(MPSys mp, char *mpconfile, char *method) {
  char file[512];
  char meth[8] = "w";
  char filetype = 'a';
  switch (method[0]) {
    case 'u':
      snprintf(file, sizeof file, mpconfile, mp->step);
      filetype = method[1];
      break;
    case 'n':
      snprintf(file, sizeof file, "%s", mpconfile);
      if (method[1] != '\0') {
        if (method[1] == 'a') strcpy(meth, "a");
        filetype = method[2];
      } else filetype = 'a';
      break;
    default:
      error("Unrecognised method code (1st letter) in mpconwr");
  }
  if (filetype != 'b' && filetype != 'c') filetype = 'a';
  LBReal pmin = mp->p2bytemin, pmax = mp->p2bytemax;
  if (pmin >= pmax) {
    printf("Switching to p2bytemin = 0.0 and p2bytemax = 1.0\n");
    pmin = 0.0; pmax = 1.0;
  }
  FILE *fp = fopen(file, meth);
  if (fp == NULL) {
    printf("mpconfile: %s could not be opened\n", file);
    return;
  }
  printf("MP '%s' %6i: ", mp->name, mp->step);
  if (meth[0] == 'w') printf("Writing");
  if (meth[0] == 'a') printf("Appending");
  switch (filetype) {
    case 'a': printf(" ascii"); break;
    case 'b': printf(" binary"); break;
    case 'c': printf(" byte-converted"); break;
  }
  printf(" scalar density to %s\n", file);
  slb_host_p(mp);
  for (long n = 0; n < mp->lb->nnodes; n++) {
    LBStoreReal val = (mp->lb->obj[n] == NULL) ? (LBStoreReal)mp->p[n] : 0.0f;
    if (filetype == 'c') {
      int iv = (int)(256.0 * (val - pmin) / (pmax - pmin));
      if (iv > 255) iv = 255;
      if (iv < 0) iv = 0;
      char cv = (char)iv;
      fwrite(&cv, 1, 1, fp);
    } else if (filetype == 'b') {
      fwrite(&val, sizeof(LBStoreReal), 1, fp);
    } else {
      fprintf(fp, "%f\n", val);
    }
  }
  fclose(fp);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP '%s' %i : mpconwr done\n", mp->name, mp->step);
}

MPIso newmpiso(MPSys mp, LBReal val) {
  MPIso iso = (MPIso)malloc(sizeof(MPIso_type));
  if (iso == NULL) {
    fprintf(stderr, "No room at the inn for another isosurface\n");
    return NULL;
  }
  iso->mp = mp;
  iso->val = val;
  iso->count = slb_ialloc(mp->lb->nnodes);
  return iso;
}

void mpiso(MPIso iso) {
  slb_host_p(iso->mp);
  LBReal *p = iso->mp->p;
  for (long n = 0; n < iso->mp->lb->nnodes; n++)
    if (p[n] > iso->val) iso->count[n]++;
  if (slb_verbosity & TRACE_ROUTINE_EXITS)
    printf("MP '%s' %i : mpiso done\n", iso->mp->name, iso->mp->step);
}

void mpisowr(MPIso iso, char *isofile, char *type) {
  int bin = (type[0] == 'b' || type[0] == 'B');
  FILE *fp = fopen(isofile, "w");
  if (fp == NULL) {
    printf("mpisowr: %s could not be opened\n", isofile);
    return;
  }
  printf("MP '%s' %6i: Saving cumulative isosurface data to %s\n", iso->mp->name, iso->mp->step,
         isofile);
  for (long n = 0; n < iso->mp->lb->nnodes; n++) {
    int v = iso->count[n];
    if (bin) fwrite(&v, sizeof(int), 1, fp);
    else fprintf(fp, "%i\n", v);
  }
  fclose(fp);
  if (slb_verbosity & TRACE_ROUTINE_EXITS)
    printf("MP '%s' %i : mpisowr done\n", iso->mp->name, iso->mp->step);
}

/* Variance of the tracer position along dirn over fluid nodes, raw lattice
 * coordinates (src/mpmisc.c:466-499). */
LBReal mpspread(MPSys mp, char *dirn) {
  int fx = 0, fy = 0, fz = 0;
  switch (dirn[0]) {
    case 'x': case 'i': fx = 1; break;
    case 'y': case 'j': fy = 1; break;
    case 'z': case 'k': fz = 1; break;
    default: printf("Unrecognised direction for measuring spread\n"); return -1.0;
  }
  slb_host_p(mp);
  LBSys lb = mp->lb;
  LBReal tot = 0.0, mean = 0.0, var = 0.0;
  for (int i = 0; i < lb->esi; i++)
    for (int j = 0; j < lb->esj; j++)
      for (int k = 0; k < lb->esk; k++) {
        long n = (long)lb->esk * ((long)lb->esj * i + j) + k;
        if (lb->obj[n] != NULL) continue;
        LBReal pv = mp->p[n];
        int pos = fx * i + fy * j + fz * k;
        tot += pv;
        mean += pv * pos;
        var += pv * pos * pos;
      }
  if (tot > 0.0) {
    mean = mean / tot;
    var = var / tot - mean * mean;
  }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP '%s' %i : mpspread done\n", mp->name, mp->step);
  return var;
}

/* ------------------------------------------------------------------ */
/* Tracer sources (src/mpobjs.c:28-261).                                */

static inline long mnode(LBSys lb, int i, int j, int k) {
  return (long)lb->esk * ((long)lb->esj * i + j) + k;
}

void mpall(MPSys mp, LBReal pval) {
  LBSys lb = mp->lb;
  slb_host_p_w(mp);
  int ns = 0;
  for (long n = 0; n < lb->nnodes; n++)
    if (lb->obj[n] == NULL) { mp->p[n] = pval; ns++; }
  if (slb_verbosity & MP_OBJECT_DETAILS) {
    printf("Diffusing material slab inserted everywhere, ");
    printf("%i sites set\n", ns);
  }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP %i : mpall done\n", mp->step);
}

void mpslab(MPSys mp, LBReal x1, LBReal x2, LBReal y1, LBReal y2, LBReal z1, LBReal z2, LBReal pval) {
  LBSys lb = mp->lb;
  slb_host_p_w(mp);
  int ns = 0;
  for (int i = 0; i < lb->esi; i++)
    for (int j = 0; j < lb->esj; j++)
      for (int k = 0; k < lb->esk; k++)
        if (i > x1 && i < x2 && j > y1 && j < y2 && k > z1 && k < z2) {
          long n = mnode(lb, i, j, k);
          if (lb->obj[n] == NULL) { mp->p[n] = pval; ns++; }
        }
  if (slb_verbosity & MP_OBJECT_DETAILS) {
    printf("Diffusing material slab from %5.2f %5.2f %5.2f ", x1, y1, z1);
    printf("to %5.2f %5.2f %5.2f, ", x2, y2, z2);
    printf("%i sites set\n", ns);
  }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP %i : mpslab done\n", mp->step);
}

void mpwall(MPSys mp, char *type, int wallpos, LBReal pval) {
  LBSys lb = mp->lb;
  int lim;
  switch (type[0]) {
    case 'x': case 'i': lim = lb->esi; break;
    case 'y': case 'j': lim = lb->esj; break;
    case 'z': case 'k': lim = lb->esk; break;
    default: printf("Unrecognised direction for walls\n"); return;
  }
  if (wallpos < 0 || wallpos > lim - 1) { printf("wallpos out of range\n"); return; }
  slb_host_p_w(mp);
  int ns = 0;
  for (int i = 0; i < lb->esi; i++)
    for (int j = 0; j < lb->esj; j++)
      for (int k = 0; k < lb->esk; k++) {
        int q = (type[0] == 'x' || type[0] == 'i') ? i : (type[0] == 'y' || type[0] == 'j') ? j : k;
        if (q != wallpos) continue;
        long n = mnode(lb, i, j, k);
        if (lb->obj[n] == NULL) { mp->p[n] = pval; ns++; }
      }
  if (slb_verbosity & MP_OBJECT_DETAILS) {
    printf("Diffusing material wall at %c = %i, ", type[0], wallpos);
    printf("%i sites set\n", ns);
  }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP %i : mpwall done\n", mp->step);
}

void mpsph(MPSys mp, LBReal rad, LBReal ox, LBReal oy, LBReal oz, LBReal pval) {
  LBSys lb = mp->lb;
  slb_host_p_w(mp);
  int ns = 0;
  LBReal radsq = rad * rad;
  for (int i = 0; i < lb->esi; i++)
    for (int j = 0; j < lb->esj; j++)
      for (int k = 0; k < lb->esk; k++) {
        volatile LBReal dx = i - ox, dy = j - oy, dz = k - oz;
        LBReal rsq = dx * dx + dy * dy + dz * dz;
        if (rsq < radsq) {
          long n = mnode(lb, i, j, k);
          if (lb->obj[n] == NULL) { mp->p[n] = pval; ns++; }
        }
      }
  if (slb_verbosity & MP_OBJECT_DETAILS) {
    printf("Diffusing material sphere, radius %5.2f at %5.2f %5.2f %5.2f, ", rad, ox, oy, oz);
    printf("%i sites set\n", ns);
  }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP %i : mpsph done\n", mp->step);
}

/* Radial predicate shared by the cylinder helpers: type[0] names the axis,
 * a second letter inverts (hole). */
static int in_cyl(char *type, int i, int j, int k, LBReal radsq, LBReal a, LBReal b) {
  volatile LBReal u, v;
  switch (type[0]) {
    case 'x': case 'i': u = j - a; v = k - b; break;
    case 'y': case 'j': u = i - a; v = k - b; break;
    case 'z': case 'k': u = i - a; v = j - b; break;
    default: error("Unrecognised cylinder direction"); return 0;
  }
  LBReal rsq = u * u + v * v;
  int within = (rsq < radsq);
  if (type[1] != '\0') within = !within;
  return within;
}

void mpcyl(MPSys mp, char *type, LBReal rad, LBReal a, LBReal b, LBReal pval) {
  LBSys lb = mp->lb;
  slb_host_p_w(mp);
  int ns = 0;
  LBReal radsq = rad * rad;
  for (int i = 0; i < lb->esi; i++)
    for (int j = 0; j < lb->esj; j++)
      for (int k = 0; k < lb->esk; k++)
        if (in_cyl(type, i, j, k, radsq, a, b)) {
          long n = mnode(lb, i, j, k);
          if (lb->obj[n] == NULL) { mp->p[n] = pval; ns++; }
        }
  if (slb_verbosity & MP_OBJECT_DETAILS) {
    if (type[1] == '\0') printf("Diffusing material cylinder, \n");
    else printf("Diffusing material cylindrical hole, \n");
    printf("radius %5.2f, axis %5.2f %5.2f along %c, ", rad, a, b, type[0]);
    printf("%i sites set\n", ns);
  }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP %i : mpcyl done\n", mp->step);
}

void mpsubs(MPSys mp, char *subsfile, int ilo, int ihi, int jlo, int jhi, int klo, int khi, LBReal pval) {
  FILE *fp = fopen(subsfile, "r");
  if (fp == NULL) {
    printf("mpsubs: %s could not be opened\n", subsfile);
    return;
  }
  if (slb_verbosity & MP_OBJECT_DETAILS) {
    printf("\nInserting mp substrate data from %s\n", subsfile);
    printf("Substrate region is (%i-%i,%i-%i,%i-%i) inclusive\n", ilo, ihi, jlo, jhi, klo, khi);
  }
  LBSys lb = mp->lb;
  slb_host_p_w(mp);
  int ns = 0;
  LBReal ptot = 0.0;
  for (int i = ilo; i <= ihi; i++)
    for (int j = jlo; j <= jhi; j++)
      for (int k = klo; k <= khi; k++) {
        unsigned char byte;
        if (fscanf(fp, "%c", &byte) != 1) {
          printf("mpsubs: ran out of data in %s\n", subsfile);
          goto done;
        }
        long n = mnode(lb, i, j, k);
        if (lb->obj[n] == NULL) {
          LBReal val = pval * (double)byte / 255.0;
          mp->p[n] = val; ns++; ptot += val;
        }
      }
done:
  fclose(fp);
  if (slb_verbosity & MP_OBJECT_DETAILS) {
    printf("Inserted mp substrate field, %i sites set\n", ns);
    printf("Total amount of added material is %f\n", ptot);
  }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP %i : mpsubs done\n", mp->step);
}

void mplist(MPSys mp, char *subsfile, LBReal pval) {
  FILE *fp = fopen(subsfile, "r");
  if (fp == NULL) {
    printf("mpsubs: %s could not be opened\n", subsfile);
    return;
  }
  if (slb_verbosity & MP_OBJECT_DETAILS) printf("\nInserting mp list data from %s\n", subsfile);
  LBSys lb = mp->lb;
  slb_host_p_w(mp);
  int ns = 0, i, j, k;
  LBReal ptot = 0.0;
  while (fscanf(fp, "%i%i%i", &i, &j, &k) == 3) {
    if (i < 0 || i >= lb->esi || j < 0 || j >= lb->esj || k < 0 || k >= lb->esk) continue;
    long n = mnode(lb, i, j, k);
    if (lb->obj[n] == NULL) { mp->p[n] = pval; ns++; ptot += pval; }
  }
  fclose(fp);
  if (slb_verbosity & MP_OBJECT_DETAILS) {
    printf("Inserted mp substrate field, %i sites set\n", ns);
    printf("Total amount of added material is %f\n", ptot);
  }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP %i : mpsubs done\n", mp->step);
}

/* ------------------------------------------------------------------ */
/* Adsorption sites: bit 0x4000 of the mask, only on nodes whose mask is
 * non-zero (src/mpads.c:37-186). */

void adall(MPSys mp) {
  LBSys lb = mp->lb;
  slb_host_mask_w(lb);
  int ns = 0;
  for (long n = 0; n < lb->nnodes; n++)
    if (lb->mask[n] != 0) { lb->mask[n] |= 0x4000; ns++; }
  printf("All nodes are adsorption sites, %i sites set\n", ns);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP %i : adall done\n", mp->step);
}

void adnone(MPSys mp) {
  LBSys lb = mp->lb;
  slb_host_mask_w(lb);
  int ns = 0;
  for (long n = 0; n < lb->nnodes; n++)
    if (lb->mask[n] & 0x4000) { lb->mask[n] &= ~0x4000; ns++; }
  printf("Reset %i adsorption sites to non-adsorbing sites\n", ns);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP %i : adnone done\n", mp->step);
}

void adslab(MPSys mp, LBReal x1, LBReal x2, LBReal y1, LBReal y2, LBReal z1, LBReal z2) {
  LBSys lb = mp->lb;
  slb_host_mask_w(lb);
  int ns = 0;
  for (int i = 0; i < lb->esi; i++)
    for (int j = 0; j < lb->esj; j++)
      for (int k = 0; k < lb->esk; k++)
        if (i > x1 && i < x2 && j > y1 && j < y2 && k > z1 && k < z2) {
          long n = mnode(lb, i, j, k);
          if (lb->mask[n] != 0) { lb->mask[n] |= 0x4000; ns++; }
        }
  printf("Adsorption site slab from %5.2f %5.2f %5.2f ", x1, y1, z1);
  printf("to %5.2f %5.2f %5.2f, ", x2, y2, z2);
  printf("%i sites set\n", ns);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP %i : adslab done\n", mp->step);
}

void adsph(MPSys mp, LBReal rad, LBReal ox, LBReal oy, LBReal oz) {
  LBSys lb = mp->lb;
  slb_host_mask_w(lb);
  int ns = 0;
  LBReal radsq = rad * rad;
  for (int i = 0; i < lb->esi; i++)
    for (int j = 0; j < lb->esj; j++)
      for (int k = 0; k < lb->esk; k++) {
        volatile LBReal dx = i - ox, dy = j - oy, dz = k - oz;
        LBReal rsq = dx * dx + dy * dy + dz * dz;
        if (rsq < radsq) {
          long n = mnode(lb, i, j, k);
          if (lb->mask[n] != 0) { lb->mask[n] |= 0x4000; ns++; }
        }
      }
  printf("Adsorption site sphere, radius %5.2f at %5.2f %5.2f %5.2f, ", rad, ox, oy, oz);
  printf("%i sites set\n", ns);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP %i : adsph done\n", mp->step);
}

void adcyl(MPSys mp, char *type, LBReal rad, LBReal a, LBReal b) {
  LBSys lb = mp->lb;
  slb_host_mask_w(lb);
  int ns = 0;
  LBReal radsq = rad * rad;
  for (int i = 0; i < lb->esi; i++)
    for (int j = 0; j < lb->esj; j++)
      for (int k = 0; k < lb->esk; k++)
        if (in_cyl(type, i, j, k, radsq, a, b)) {
          long n = mnode(lb, i, j, k);
          if (lb->mask[n] != 0) { lb->mask[n] |= 0x4000; ns++; }
        }
  if (type[1] == '\0') printf("Adsorbtion site cylinder, \n");
  else printf("Adsorbtion site cylindrical hole, \n");
  printf("radius %5.2f, axis %5.2f %5.2f along %c, ", rad, a, b, type[0]);
  printf("%i sites set\n", ns);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP %i : adcyl done\n", mp->step);
}

/* Total adsorbed amount; appends "step*imult total" to adtfile unless it is
 * "null" (src/mpads.c:149-172). */
LBReal adtotal(MPSys mp, char *adtfile, int imult) {
  static int nadt = 0;
  LBSys lb = mp->lb;
  slb_host_p(mp);
  slb_host_mask(lb);
  LBReal ptot = 0.0;
  for (long n = 0; n < lb->nnodes; n++)
    if (lb->mask[n] & 0x4000) ptot += mp->p[n];
  if (!myscmp("null", adtfile)) {
    FILE *fp = fopen(adtfile, nadt == 0 ? "w" : "a");
    if (fp == NULL) printf("adtotal: %s could not be opened to write/append data\n", adtfile);
    else {
      fprintf(fp, "%8i %f\n", nadt * imult, ptot);
      fclose(fp);
    }
  }
  nadt++;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP %i : adtotal done\n", mp->step);
  return ptot;
}

void adzero(MPSys mp) {
  LBSys lb = mp->lb;
  slb_host_mask(lb);
  slb_host_p_w(mp);
  for (long n = 0; n < lb->nnodes; n++)
    if (lb->mask[n] & 0x4000) mp->p[n] = 0.0;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP %i : adzero done\n", mp->step);
}
