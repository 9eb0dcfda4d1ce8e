/*
 * slb_core.c -- lattice construction, host/device coherence and the LB step
 * entry points of the sunlightlb.h API, driving the gfx950 kernels through
 * slb_dev.h.
 *
 * Reference interfaces replaced here: newlb (src/lbmisc.c:81-152), lbstep,
 * bbl, bblforce, propagate, collide (src/lbkernel.c:35-51, :623-877),
 * makemask, syncobjs, countfluid, countmask, vspread, zerobf, addbf, vset
 * (src/lbobjs.c:526-714).
 */
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include "slb_internal.h"

typedef char slb_lbreal_is_double[(sizeof(LBReal) == sizeof(double)) ? 1 : -1];

/* Zero-filled arrays; release with slb_big_free. */
LBReal *slb_ralloc(long n) {
  LBReal *p = (LBReal *)slb_big_calloc(n, sizeof(LBReal));
  if (p == NULL) error("No room for LBReal array");
  return p;
}

int *slb_ialloc(long n) {
  int *p = (int *)slb_big_calloc(n, sizeof(int));
  if (p == NULL) error("No room for int array");
  return p;
}

/* D3Q15 velocity set and integer weights (layout of src/lbmisc.c:89-104:
 * c[3l+d], w[l]). */
static const int d3q15_c[45] = {0, 0, 0,  1, 0, 0,  -1, 0, 0,  0, 1, 0,  0, -1, 0,
                                0, 0, 1,  0, 0, -1,  1, 1, 1,  -1, 1, 1,  1, -1, 1,
                                -1, -1, 1,  1, 1, -1,  -1, 1, -1,  1, -1, -1,  -1, -1, -1};
static const int d3q15_w[15] = {16, 8, 8, 8, 8, 8, 8, 1, 1, 1, 1, 1, 1, 1, 1};

static int default_precision(void) {
  const char *e = getenv("SLB_PRECISION");
  if (e && (atoi(e) == 32 || atoi(e) == 64)) return atoi(e);
  return 64;
}

LBSys newlb(int reqesi, int reqesj, int reqesk, char *name) {
  /* nnodes is an int in the struct layout drivers link against */
  if (reqesi < 1 || reqesj < 1 || reqesk < 1 ||
      (long)reqesi * reqesj * reqesk > (long)INT_MAX)
    error("newlb: lattice dimensions must be positive, with at most 2^31 - 1 nodes");
  slb_lbwrap *W = (slb_lbwrap *)calloc(1, sizeof(slb_lbwrap));
  if (W == NULL) error("No room for lbsys structure\n");
  LBSys lb = &W->pub;
  slb_lbpriv *S = &W->priv;
  lb->c = slb_ialloc(45);
  lb->w = slb_ialloc(15);
  memcpy(lb->c, d3q15_c, sizeof d3q15_c);
  memcpy(lb->w, d3q15_w, sizeof d3q15_w);
  lb->esi = reqesi; lb->esj = reqesj; lb->esk = reqesk;
  lb->nnodes = reqesi * reqesj * reqesk;
  long N = (long)lb->nnodes;
  lb->a = slb_ralloc(15 * N);
  lb->mask = slb_ialloc(N);
  lb->obj = (LBObj *)slb_big_calloc(N, sizeof(LBObj));
  if (lb->obj == NULL) error("No room for LBObj array\n");
  lb->ux = slb_ralloc(N); lb->uy = slb_ralloc(N); lb->uz = slb_ralloc(N);
  lb->fx = slb_ralloc(N); lb->fy = slb_ralloc(N); lb->fz = slb_ralloc(N);
  lb->bx = slb_ralloc(N); lb->by = slb_ralloc(N); lb->bz = slb_ralloc(N);
  lb->tau = 0.75;
  lb->rho0 = 0.0;
  for (int l = 0; l < 15; l++) lb->rho0 += lb->w[l];
  lb->objects_moving = 0;
  lb->skip_solid = 0;
  lb->stokes_flow = 0;
  lb->vel_corrected = 1;
  lb->step = 0;
  lb->next_id = 0;
  lb->obj_head = NULL;
  lb->status = POST_COLLIDE;
  lb->changed_objects = 0;
  lb->rescale = 1.0;
  lb->load_average = 0;
  lb->psi = NULL;
  lb->rho = NULL;
  lb->dpdx = lb->dpdy = lb->dpdz = 0.0;
  slb_verbosity = 0;
  slb_clock_reset();
  if (myscmp("null", name)) lb->name = NULL; else lb->name = mycopy(name);
  lb->monitoring = (lb->name == NULL) ? 0 : 1;
  /* device state: every mirror starts in sync (all zero, all fluid) */
  S->prec = default_precision();
  S->dev = slbd_create(reqesi, reqesj, reqesk, S->prec);
  S->pops_loc = S->u_loc = S->f_loc = S->mask_loc = S->rho_loc = LOC_SYNC;
  S->objidx = (int32_t *)slb_xmalloc(sizeof(int32_t) * (size_t)(N > 0 ? N : 1));
  for (long n = 0; n < N; n++) S->objidx[n] = -1;
  S->objidx_dirty = 0;
  S->bf_dirty = 0;
  S->bf_uniform = 1;
  S->nfluid = lb->nnodes;
  S->flags_stale = 0;
  printf("LB system '%s' initialised for %i * %i * %i lattice\n", lb->name, lb->esi, lb->esj,
         lb->esk);
#ifdef SIGNAL
  siginit();
  printf("Signalling implemented, process id = %i\n", (int)getpid());
#else
  printf("Signalling not implemented\n");
#endif
  return lb;
}

/* ------------------------------------------------------------------ */
/* Mirror coherence.                                                    */

void slb_host_pops(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  if (S->pops_loc == LOC_DEV) { slbd_get_pops(S->dev, lb->a); S->pops_loc = LOC_SYNC; }
}
void slb_host_pops_w(LBSys lb) { slb_host_pops(lb); PRIV(lb)->pops_loc = LOC_HOST; }
void slb_dev_pops(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  if (S->pops_loc == LOC_HOST) { slbd_put_pops(S->dev, lb->a); S->pops_loc = LOC_SYNC; }
}

void slb_host_u(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  if (S->u_loc == LOC_DEV) { slbd_get_u(S->dev, lb->ux, lb->uy, lb->uz); S->u_loc = LOC_SYNC; }
}
void slb_host_u_w(LBSys lb) { slb_host_u(lb); PRIV(lb)->u_loc = LOC_HOST; }
void slb_dev_u(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  if (S->u_loc == LOC_HOST) { slbd_put_u(S->dev, lb->ux, lb->uy, lb->uz); S->u_loc = LOC_SYNC; }
}

void slb_host_f(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  if (S->f_loc == LOC_DEV) { slbd_get_f(S->dev, lb->fx, lb->fy, lb->fz); S->f_loc = LOC_SYNC; }
}
void slb_host_f_w(LBSys lb) { slb_host_f(lb); PRIV(lb)->f_loc = LOC_HOST; }
void slb_dev_f(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  if (S->f_loc == LOC_HOST) { slbd_put_f(S->dev, lb->fx, lb->fy, lb->fz); S->f_loc = LOC_SYNC; }
}

void slb_host_rho(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  if (S->rho_loc == LOC_DEV && lb->rho) { slbd_mp_get_rho(S->dev, lb->rho); S->rho_loc = LOC_SYNC; }
}
void slb_host_rho_w(LBSys lb) { slb_host_rho(lb); PRIV(lb)->rho_loc = LOC_HOST; }
void slb_dev_rho(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  if (S->rho_loc == LOC_HOST && lb->rho) { slbd_mp_put_rho(S->dev, lb->rho); S->rho_loc = LOC_SYNC; }
}

void slb_host_mask(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  if (S->mask_loc == LOC_DEV) { slbd_get_mask(S->dev, lb->mask); S->mask_loc = LOC_SYNC; }
}
void slb_host_mask_w(LBSys lb) {
  slb_host_mask(lb);
  PRIV(lb)->mask_loc = LOC_HOST;
  PRIV(lb)->flags_stale = 1;
}

/* Dense object index: list order from obj_head.  A small open-addressing
 * table maps object pointers to indices while the per-node index is built. */
static int objtab_grow(slb_lbpriv *S, int n) {
  if (n <= S->objcap) return 0;
  int cap = S->objcap ? S->objcap : 16;
  while (cap < n) cap *= 2;
  S->objlist = (LBObj *)realloc(S->objlist, sizeof(LBObj) * cap);
  S->objtab = (double *)realloc(S->objtab, sizeof(double) * SLBD_OBJ_STRIDE * cap);
  if (!S->objlist || !S->objtab) error("No room for object table");
  S->objcap = cap;
  return 0;
}

void slb_objtab(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  int n = 0;
  for (LBObj p = lb->obj_head; p != NULL; p = p->next_obj) n++;
  objtab_grow(S, n);
  n = 0;
  for (LBObj p = lb->obj_head; p != NULL; p = p->next_obj) {
    S->objlist[n] = p;
    double *t = S->objtab + (long)SLBD_OBJ_STRIDE * n;
    t[0] = p->posx; t[1] = p->posy; t[2] = p->posz;
    t[3] = p->velx; t[4] = p->vely; t[5] = p->velz;
    t[6] = p->angx; t[7] = p->angy; t[8] = p->angz;
    n++;
  }
  S->nobj = n;
}

/* Refresh only the numeric rows (pos/vel/ang) for the current dense order. */
static void objtab_values(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  for (int o = 0; o < S->nobj; o++) {
    LBObj p = S->objlist[o];
    double *t = S->objtab + (long)SLBD_OBJ_STRIDE * o;
    t[0] = p->posx; t[1] = p->posy; t[2] = p->posz;
    t[3] = p->velx; t[4] = p->vely; t[5] = p->velz;
    t[6] = p->angx; t[7] = p->angy; t[8] = p->angz;
  }
}

static uint64_t ptr_hash(const void *p) {
  uint64_t x = (uint64_t)(uintptr_t)p;
  x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33;
  return x;
}

typedef struct {
  LBObj *obj;
  int32_t *objidx;
  LBObj *keys;
  int *vals;
  int cap;
  long nfluid[SLB_PAR_MAX];
  int bad[SLB_PAR_MAX];
} objidx_pass;

static void objidx_range(void *c, long lo, long hi, int tid) {
  objidx_pass *J = (objidx_pass *)c;
  LBObj last = NULL;
  int lastv = -1;
  long nf = 0;
  J->bad[tid] = 0;
  for (long n = lo; n < hi; n++) {
    LBObj p = J->obj[n];
    if (p == NULL) { J->objidx[n] = -1; nf++; continue; }
    if (p != last) {
      uint64_t h = ptr_hash(p) & (J->cap - 1);
      while (J->keys[h] != NULL && J->keys[h] != p) h = (h + 1) & (J->cap - 1);
      if (J->keys[h] != p) { J->bad[tid] = 1; break; }
      last = p; lastv = J->vals[h];
    }
    J->objidx[n] = lastv;
  }
  J->nfluid[tid] = nf;
}

static void build_objidx(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  slb_objtab(lb);
  int cap = 16;
  while (cap < 2 * S->nobj + 2) cap *= 2;
  LBObj *keys = (LBObj *)slb_xcalloc(cap, sizeof(LBObj));
  int *vals = (int *)slb_xmalloc(sizeof(int) * cap);
  for (int o = 0; o < S->nobj; o++) {
    uint64_t h = ptr_hash(S->objlist[o]) & (cap - 1);
    while (keys[h] != NULL) h = (h + 1) & (cap - 1);
    keys[h] = S->objlist[o];
    vals[h] = o;
  }
  objidx_pass J;
  J.obj = lb->obj; J.objidx = S->objidx; J.keys = keys; J.vals = vals; J.cap = cap;
  const int nr = slb_par_for(lb->nnodes, objidx_range, &J);
  free(keys); free(vals);
  long nfluid = 0;
  for (int t = 0; t < nr; t++) {
    if (J.bad[t]) error("Node refers to an object that is not in the object list");
    nfluid += J.nfluid[t];
  }
  S->nfluid = (int)nfluid;
  slbd_put_objidx(S->dev, S->objidx);
  S->objidx_dirty = 0;
  S->flags_stale = 1;
}

int slb_objindex(LBSys lb, LBObj p) {
  slb_lbpriv *S = PRIV(lb);
  for (int o = 0; o < S->nobj; o++) if (S->objlist[o] == p) return o;
  return -1;
}

void slb_obj_changed(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  S->objidx_dirty = 1;
  S->bf_dirty = 1;
}

void slb_bf_changed(LBSys lb) { PRIV(lb)->bf_dirty = 1; }

/* Device objidx, mask and pull flags current for the next kernel. */
void slb_dev_geom(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  if (S->objidx_dirty) build_objidx(lb);
  if (S->mask_loc == LOC_HOST) {
    slbd_put_mask(S->dev, lb->mask);
    S->mask_loc = LOC_SYNC;
    S->flags_stale = 1;
  }
  if (S->flags_stale) { slbd_build_flags(S->dev); S->flags_stale = 0; }
}

/* Body force analysis: the force on a fluid node is b[n] + (dpdx,dpdy,dpdz).
 * When b is the same on every fluid node it travels as a kernel argument,
 * otherwise the per-node arrays are uploaded. */
typedef struct {
  LBSys lb;
  int have[SLB_PAR_MAX], uniform[SLB_PAR_MAX];
  double u[SLB_PAR_MAX][3];
} bf_pass;

/* per range: its first fluid node's force and whether every later fluid
 * node of the range equals it */
static void bf_range(void *c, long lo, long hi, int tid) {
  bf_pass *J = (bf_pass *)c;
  LBSys lb = J->lb;
  int have = 0, uniform = 1;
  double u0 = 0, u1 = 0, u2 = 0;
  for (long n = lo; n < hi; n++) {
    if (lb->obj[n] != NULL) continue;
    if (!have) { u0 = lb->bx[n]; u1 = lb->by[n]; u2 = lb->bz[n]; have = 1; continue; }
    if (lb->bx[n] != u0 || lb->by[n] != u1 || lb->bz[n] != u2) { uniform = 0; break; }
  }
  J->have[tid] = have; J->uniform[tid] = uniform;
  J->u[tid][0] = u0; J->u[tid][1] = u1; J->u[tid][2] = u2;
}

static void analyse_bf(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  bf_pass J;
  J.lb = lb;
  const int nr = slb_par_for(lb->nnodes, bf_range, &J);
  /* the same verdict and value as one scan in node order: the first fluid
   * node's force, uniform when every range is and starts with that force */
  int have = 0, uniform = 1;
  double u0 = 0, u1 = 0, u2 = 0;
  for (int t = 0; t < nr && uniform; t++) {
    if (!J.have[t]) continue;
    if (!J.uniform[t]) { uniform = 0; break; }
    if (!have) { u0 = J.u[t][0]; u1 = J.u[t][1]; u2 = J.u[t][2]; have = 1; continue; }
    if (J.u[t][0] != u0 || J.u[t][1] != u1 || J.u[t][2] != u2) uniform = 0;
  }
  S->bf_uniform = uniform;
  if (uniform) {
    S->bfu[0] = u0; S->bfu[1] = u1; S->bfu[2] = u2;
  } else {
    S->bfu[0] = S->bfu[1] = S->bfu[2] = 0.0;
    slbd_put_bf(S->dev, lb->bx, lb->by, lb->bz);
  }
  S->bf_dirty = 0;
}

void slb_params(LBSys lb, slbd_params *P) {
  slb_lbpriv *S = PRIV(lb);
  if (S->objidx_dirty) build_objidx(lb);
  if (S->bf_dirty) analyse_bf(lb);
  P->tau = lb->tau;
  P->rho0 = lb->rho0;
  P->stokes_flow = lb->stokes_flow;
  P->skip_solid = lb->skip_solid;
  P->moving = lb->objects_moving;
  P->bf[0] = S->bfu[0] + lb->dpdx;
  P->bf[1] = S->bfu[1] + lb->dpdy;
  P->bf[2] = S->bfu[2] + lb->dpdz;
  if (!S->bf_uniform) P->bf_mode = 2;
  else P->bf_mode = (P->bf[0] != 0.0 || P->bf[1] != 0.0 || P->bf[2] != 0.0) ? 1 : 0;
}

void slb_after_collide(LBSys lb) {
  lb->step += 1;
  lb->status = POST_COLLIDE;
#ifdef SIGNAL
  if (slb_signal_flag) sighandle(lb);
#endif
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : collide done\n", lb->step);
}

/* ------------------------------------------------------------------ */
/* Step entry points.                                                   */

static void prepare_step(LBSys lb, int force, slbd_params *P) {
  slb_lbpriv *S = PRIV(lb);
  slb_params(lb, P);
  slb_dev_geom(lb);
  slb_dev_pops(lb);
  if (lb->objects_moving) slb_dev_u(lb);
  if (force) slb_dev_f(lb);
  (void)S;
}

static void fused_steps(LBSys lb, int nsteps, int force) {
  slbd_params P;
  prepare_step(lb, force, &P);
  slbd_step(PRIV(lb)->dev, &P, nsteps, force);
  PRIV(lb)->pops_loc = LOC_DEV;
  if (force) PRIV(lb)->f_loc = LOC_DEV;
}

static void trace_bbl(LBSys lb, int force) {
  if (slb_verbosity & TRACE_ROUTINE_EXITS)
    printf("LB %i : %s (%s) done\n", lb->step, force ? "bblforce" : "bbl",
           lb->objects_moving ? "MOV" : "STAT");
}

/* bbl (+force) -> propagate -> collide as one fused kernel.  The
 * out-of-order warnings and trace lines are those of the three phases. */
static void one_step(LBSys lb, int force) {
  if (lb->changed_objects) syncobjs(lb);
  if (lb->status != POST_COLLIDE) {
    printf("Bounceback apparently out of position\n");
    wrstatus(lb, stdout);
  }
  fused_steps(lb, 1, force);
  trace_bbl(lb, force);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : propagate done\n", lb->step);
  slb_after_collide(lb);
}

void lbstep(LBSys lb) { one_step(lb, 0); }

/* Internal: measure()'s force-collecting step. */
void slb_forcestep(LBSys lb) { one_step(lb, 1); }

static void bbl_phase(LBSys lb, int force) {
  if (lb->changed_objects) syncobjs(lb);
  if (lb->status != POST_COLLIDE) {
    printf("Bounceback apparently out of position\n");
    wrstatus(lb, stdout);
  }
  slbd_params P;
  prepare_step(lb, force, &P);
  slbd_bbl(PRIV(lb)->dev, &P, force);
  PRIV(lb)->pops_loc = LOC_DEV;
  if (force) PRIV(lb)->f_loc = LOC_DEV;
  lb->status = POST_BBL;
  trace_bbl(lb, force);
}

void bbl(LBSys lb) { bbl_phase(lb, 0); }
void bblforce(LBSys lb) { bbl_phase(lb, 1); }

void propagate(LBSys lb) {
  if (lb->status != POST_BBL) {
    printf("Propagation apparently out of position\n");
    wrstatus(lb, stdout);
  }
  slb_dev_pops(lb);
  slbd_propagate(PRIV(lb)->dev);
  PRIV(lb)->pops_loc = LOC_DEV;
  lb->status = POST_PROPAGATE;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : propagate done\n", lb->step);
}

void collide(LBSys lb) {
  if (lb->status != POST_PROPAGATE) {
    printf("Collision apparently out of position\n");
    wrstatus(lb, stdout);
  }
  slbd_params P;
  slb_params(lb, &P);
  slb_dev_geom(lb);
  slb_dev_pops(lb);
  slbd_collide(PRIV(lb)->dev, &P);
  PRIV(lb)->pops_loc = LOC_DEV;
  slb_after_collide(lb);
}

/* ------------------------------------------------------------------ */
/* Tables: src/lbobjs.c:526-560 (makemask), :648-663, :671-714.         */

void makemask(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  build_objidx(lb);
  slbd_makemask(S->dev);
  S->mask_loc = LOC_DEV;      /* the host copy follows when a host reader asks */
  slbd_build_flags(S->dev);
  S->flags_stale = 0;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : makemask done\n", lb->step);
}

int countfluid(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  /* the fluid count of obj[] as the dense index was built from it */
  if (S->objidx_dirty) build_objidx(lb);
  const int count = S->nfluid;
  if (slb_verbosity & TRACE_ROUTINE_EXITS)
    printf("LB %i : countfluid done, = %i\n", lb->step, count);
  return count;
}

void syncobjs(LBSys lb) {
  makemask(lb);
  countfluid(lb);
  lb->changed_objects = 0;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : syncobjs done.\n", lb->step);
}

int countmask(LBSys lb, char *type) {
  slb_lbpriv *S = PRIV(lb);
  long N = lb->nnodes;
  int count = 0;
  /* the device copy of the mask is current unless the host wrote it */
  if (S->mask_loc != LOC_HOST && (type[0] == 'a' || type[0] == 'l' || type[0] == 'f' ||
                                  type[0] == 's')) {
    const int faces = (type[0] == 'f' || type[0] == 's');
    count = (int)slbd_countmask(S->dev, faces, lb->w);
    if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : countmask done\n", lb->step);
    return count;
  }
  switch (type[0]) {
  case 'a': case 'l':
    for (long n = 0; n < N; n++) {
      int m = lb->mask[n];
      if (m <= 0) continue;
      for (int l = 1; l < 15; l++) if (m & (1 << (l - 1))) count += lb->w[l];
    }
    break;
  case 'f': case 's':
    for (long n = 0; n < N; n++) {
      int m = lb->mask[n];
      if (m <= 0) continue;
      for (int l = 1; l <= 6; l++) if (m & (1 << (l - 1))) count++;
    }
    break;
  default:
    printf("countmask: expected 'links' or 'faces'\n");
  }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : countmask done\n", lb->step);
  return count;
}

/* Wall velocities from the object state, src/lbobjs.c:565-589. */
void vspread(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  slb_dev_geom(lb);
  objtab_values(lb);
  slbd_vspread(S->dev, S->nobj, S->objtab, lb->rho0);
  S->u_loc = LOC_DEV;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : vspread done\n", lb->step);
}

/* Body force on fluid nodes, src/lbobjs.c:593-619. */
/* A store of +0.0 only where the value differs in its bits: the result is
 * the same, and pages of the force arrays that were never written (newlb's
 * zeroed allocation) are not touched. */
static inline void set_zero(LBReal *p) {
  static const LBReal z = 0.0;
  if (memcmp(p, &z, sizeof z) != 0) *p = 0.0;
}
static void zerobf_range(void *c, long lo, long hi, int tid) {
  LBSys lb = (LBSys)c;
  (void)tid;
  for (long n = lo; n < hi; n++)
    if (lb->obj[n] == NULL) { set_zero(&lb->bx[n]); set_zero(&lb->by[n]); set_zero(&lb->bz[n]); }
}

void zerobf(LBSys lb) {
  slb_par_for(lb->nnodes, zerobf_range, lb);
  slb_bf_changed(lb);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : zerobf done\n", lb->step);
}

typedef struct {
  LBSys lb;
  LBReal d[3];
} addbf_pass;
static void addbf_range(void *c, long lo, long hi, int tid) {
  addbf_pass *J = (addbf_pass *)c;
  LBSys lb = J->lb;
  (void)tid;
  for (long n = lo; n < hi; n++)
    if (lb->obj[n] == NULL) { lb->bx[n] += J->d[0]; lb->by[n] += J->d[1]; lb->bz[n] += J->d[2]; }
}

void addbf(LBSys lb, double dpdx, double dpdy, double dpdz) {
  addbf_pass J = {lb, {(LBReal)dpdx, (LBReal)dpdy, (LBReal)dpdz}};
  slb_par_for(lb->nnodes, addbf_range, &J);
  slb_bf_changed(lb);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : zerobf done\n", lb->step);
}

/* One node of the wall-velocity field, src/lbobjs.c:623-643. */
void vset(LBSys lb, int i, int j, int k, LBReal vx, LBReal vy, LBReal vz) {
  if (i >= 0 && i < lb->esi && j >= 0 && j < lb->esj && k >= 0 && k < lb->esk) {
    long n = (long)lb->esk * ((long)lb->esj * i + j) + k;
    if (lb->obj[n] != NULL) {
      slb_host_u_w(lb);
      lb->ux[n] = lb->rho0 * vx / 12.0;
      lb->uy[n] = lb->rho0 * vy / 12.0;
      lb->uz[n] = lb->rho0 * vz / 12.0;
    }
  }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : vset done\n", lb->step);
}

/* ------------------------------------------------------------------ */
/* Extensions (sunlightlb_gpu.h).                                       */

int slb_get_precision(LBSys lb) { return PRIV(lb)->prec; }

int slb_set_precision(LBSys lb, int bits) {
  slb_lbpriv *S = PRIV(lb);
  if (bits != 32 && bits != 64) return -1;
  if (bits == S->prec) return 0;
  /* bring every device-resident field to the host, switch, re-upload */
  slb_sync_host(lb);
  for (slb_mpwrap *m = S->mps; m; m = m->priv.next) {
    slbd_mp_destroy(m->priv.dev);
    m->priv.dev = NULL;
  }
  slbd_set_precision(S->dev, bits, lb->a);
  S->prec = bits;
  S->pops_loc = LOC_SYNC;
  if (lb->rho) S->rho_loc = LOC_HOST;
  S->bf_dirty = 1;
  for (slb_mpwrap *m = S->mps; m; m = m->priv.next) {
    m->priv.dev = slbd_mp_create(S->dev, m->priv.nmom);
    m->priv.p_loc = LOC_HOST;
  }
  return 0;
}

void slb_sync_host(LBSys lb) {
  slb_host_pops(lb);
  slb_host_u(lb);
  slb_host_f(lb);
  slb_host_mask(lb);
  slb_host_rho(lb);
  for (slb_mpwrap *m = PRIV(lb)->mps; m; m = m->priv.next) slb_host_p(&m->pub);
}

void slb_sync_device(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  S->pops_loc = LOC_HOST;
  S->u_loc = LOC_HOST;
  S->f_loc = LOC_HOST;
  S->mask_loc = LOC_HOST;
  if (lb->rho) S->rho_loc = LOC_HOST;
  S->objidx_dirty = 1;
  S->bf_dirty = 1;
  S->flags_stale = 1;
}

void slb_run(LBSys lb, int nsteps) {
  if (nsteps <= 0) return;
  if (lb->changed_objects) syncobjs(lb);
  if (lb->status != POST_COLLIDE) {
    printf("Bounceback apparently out of position\n");
    wrstatus(lb, stdout);
  }
  int trace = slb_verbosity & TRACE_ROUTINE_EXITS;
#ifdef SIGNAL
  int sig = 1;
#else
  int sig = 0;
#endif
  if (trace || sig) {
    /* chunks between host checks of the trace / signal state */
    int done = 0;
    while (done < nsteps) {
      int chunk = trace ? 1 : (nsteps - done < 256 ? nsteps - done : 256);
      fused_steps(lb, chunk, 0);
      if (trace) {
        trace_bbl(lb, 0);
        printf("LB %i : propagate done\n", lb->step);
      }
      lb->step += chunk - 1;
      slb_after_collide(lb);
      done += chunk;
    }
  } else {
    fused_steps(lb, nsteps, 0);
    lb->step += nsteps - 1;
    slb_after_collide(lb);
  }
}

double slb_residual(LBSys lb) {
  slbd_params P;
  slb_params(lb, &P);
  slb_dev_geom(lb);
  slb_dev_pops(lb);
  return slbd_residual(PRIV(lb)->dev, &P);
}

void slb_residual_reset(LBSys lb) { slbd_residual_reset(PRIV(lb)->dev); }

int slb_run_to_residual(LBSys lb, double tol, int interval, int maxsteps, double *res) {
  if (interval < 1) interval = 1;
  int steps = 0;
  double r = -1.0;
  slb_residual_reset(lb);
  slb_residual(lb);
  while (steps < maxsteps) {
    int n = interval;
    if (steps + n > maxsteps) n = maxsteps - steps;
    slb_run(lb, n);
    steps += n;
    r = slb_residual(lb);
    if (r >= 0.0 && r < tol) break;
  }
  if (res) *res = r;
  return steps;
}

/* Solve of a small dense system (k <= SLBD_VEC_MAX) by Gaussian elimination
 * with partial pivoting; returns -1 if singular. */
static int small_solve(int k, double *A, double *b) {
  for (int c = 0; c < k; c++) {
    int piv = c;
    for (int r = c + 1; r < k; r++)
      if (fabs(A[r * k + c]) > fabs(A[piv * k + c])) piv = r;
    if (!(fabs(A[piv * k + c]) > 0.0)) return -1;
    if (piv != c) {
      for (int q = 0; q < k; q++) { double t = A[c * k + q]; A[c * k + q] = A[piv * k + q]; A[piv * k + q] = t; }
      double t = b[c]; b[c] = b[piv]; b[piv] = t;
    }
    for (int r = c + 1; r < k; r++) {
      const double f = A[r * k + c] / A[c * k + c];
      for (int q = c; q < k; q++) A[r * k + q] -= f * A[c * k + q];
      b[r] -= f * b[c];
    }
  }
  for (int r = k - 1; r >= 0; r--) {
    double t = b[r];
    for (int q = r + 1; q < k; q++) t -= A[r * k + q] * b[q];
    b[r] = t / A[r * k + r];
  }
  return 0;
}

/* Run to the velocity residual with Anderson mixing of the steady-state
 * iteration.  One iteration is G = `interval` ordinary steps from state x_i;
 * the residual of G(x_i) against x_i is the residual slb_run_to_residual
 * tests, and the solve ends, in the state G(x_i), the first time it is below
 * tol.  Otherwise the next state is
 *   x_{i+1} = G(x_i) - sum_j gamma_j dG_j,
 * gamma minimising |f_i - sum_j gamma_j dF_j| over the last `window`
 * differences dF_j = f_{j+1} - f_j, dG_j = G(x_{j+1}) - G(x_j) of the
 * fixed-point residuals f = G(x) - x (Anderson's type II mixing, equivalent
 * to GMRES for the linear Stokes iteration).  The normal equations are
 * solved in FP64 with a 1e-10 relative ridge; the history restarts when the
 * residual grows tenfold over the best one.  Memory: 2 window + 3 population
 * arrays on the device. */
int slb_run_to_residual_aa(LBSys lb, double tol, int interval, int maxsteps, int window,
                           double *res) {
  if (window < 1) return slb_run_to_residual(lb, tol, interval, maxsteps, res);
  if (window > SLBD_VEC_MAX) window = SLBD_VEC_MAX;
  if (interval < 1) interval = 1;
  slb_lbpriv *S = PRIV(lb);
  slbd_lattice *D = S->dev;
  slb_dev_geom(lb);
  slb_dev_pops(lb);
  const int m = window;
  void *base = slbd_vec_alloc(D, 3 + 2 * m);
  void *F = slbd_vec_at(D, base, 0), *Fp = slbd_vec_at(D, base, 1);
  void *Gp = slbd_vec_at(D, base, 2);
  void *dF[SLBD_VEC_MAX], *dG[SLBD_VEC_MAX];
  for (int j = 0; j < m; j++) {
    dF[j] = slbd_vec_at(D, base, 3 + j);
    dG[j] = slbd_vec_at(D, base, 3 + m + j);
  }
  double gram[SLBD_VEC_MAX * SLBD_VEC_MAX], A[SLBD_VEC_MAX * SLBD_VEC_MAX];
  double rhs[SLBD_VEC_MAX], row[SLBD_VEC_MAX];
  int k = 0, head = 0, have_prev = 0, steps = 0;
  double r = -1.0, best = HUGE_VAL;
  slb_residual_reset(lb);
  slb_residual(lb);
  while (steps < maxsteps) {
    int n = interval;
    if (steps + n > maxsteps) n = maxsteps - steps;
    slbd_vec_copy(D, F, slbd_vec_cur(D));      /* x_i */
    slb_run(lb, n);
    steps += n;
    r = slb_residual(lb);
    if (r >= 0.0 && r < tol) break;
    if (steps >= maxsteps) break;              /* end on an iterate G(x_i), not a mix */
    if (r > 10.0 * best) { k = 0; head = 0; have_prev = 0; }   /* restart the history */
    if (r >= 0.0 && r < best) best = r;
    void *g = slbd_vec_cur(D);
    slbd_vec_sub(D, F, g, F);                  /* f_i = G(x_i) - x_i */
    if (have_prev) {
      const int slot = head;
      slbd_vec_sub2(D, dF[slot], F, Fp, dG[slot], g, Gp);
      if (k < m) k++;
      head = (head + 1) % m;
      /* new Gram row and right-hand side in one pass over the dF columns */
      slbd_vec_dots2(D, dF[slot], F, dF, k, row, rhs);
      for (int j = 0; j < k; j++) { gram[slot * m + j] = row[j]; gram[j * m + slot] = row[j]; }
    }
    { void *t = Fp; Fp = F; F = t; }          /* keep f_i; F is free for x_{i+1} */
    slbd_vec_copy(D, Gp, g);
    have_prev = 1;
    if (k > 0) {
      double dmax = 0.0;
      for (int j = 0; j < k; j++) dmax = gram[j * m + j] > dmax ? gram[j * m + j] : dmax;
      for (int a = 0; a < k; a++)
        for (int b = 0; b < k; b++) A[a * k + b] = gram[a * m + b] + (a == b ? 1e-10 * dmax : 0.0);
      if (small_solve(k, A, rhs) == 0) {
        slbd_vec_mix(D, g, g, dG, rhs, k);     /* x_{i+1}, in place */
        S->pops_loc = LOC_DEV;
      } else {
        k = 0; head = 0;
      }
    }
    slb_residual_reset(lb);
    slb_residual(lb);
  }
  slbd_vec_free(base);
  if (res) *res = r;
  return steps;
}

const char *slb_step_kernel(LBSys lb) { return slbd_step_kernel(PRIV(lb)->dev); }

double slb_time_steps(LBSys lb, int nsteps) {
  if (lb->changed_objects) syncobjs(lb);
  slbd_params P;
  prepare_step(lb, 0, &P);
  double ms = slbd_time_steps(PRIV(lb)->dev, &P, nsteps);
  PRIV(lb)->pops_loc = LOC_DEV;
  lb->step += nsteps;
  lb->status = POST_COLLIDE;
  return ms;
}

void slb_scale_pops(LBSys lb, double s) {
  slb_dev_pops(lb);
  slbd_scale_pops(PRIV(lb)->dev, s);
  PRIV(lb)->pops_loc = LOC_DEV;
}

void slb_sum_pops(LBSys lb, double out[15]) {
  slb_dev_geom(lb);
  slb_dev_pops(lb);
  slbd_sum_pops(PRIV(lb)->dev, out);
}

size_t slb_device_bytes(LBSys lb) { return slbd_bytes(PRIV(lb)->dev); }
const char *slb_device_arch(void) { return slbd_arch(); }

double slb_hbm_copy_test(size_t bytes, int reps) {
  if (slbd_device_count() < 1) error("no HIP device is visible");
  return slbd_copy_test(bytes, reps);
}

/* Release a lattice, its objects, its MP systems and all device memory.
 * The handle (and every MPSys, MPIso and object pointer of it) is invalid
 * afterwards. */
void slb_destroy(LBSys lb) {
  if (lb == NULL) return;
  slb_lbpriv *S = PRIV(lb);
  for (slb_mpwrap *m = S->mps; m;) {
    slb_mpwrap *next = m->priv.next;
    slbd_mp_destroy(m->priv.dev);
    slb_big_free(m->pub.p);
    slb_big_free(m->pub.newp);
    free(m->pub.name);
    free(m);
    m = next;
  }
  slbd_destroy(S->dev);
  for (LBObj p = lb->obj_head; p != NULL;) {
    LBObj q = p->next_obj;
    free(p);
    p = q;
  }
  slb_big_free(lb->c); slb_big_free(lb->w); slb_big_free(lb->a); slb_big_free(lb->mask);
  slb_big_free(lb->obj);
  slb_big_free(lb->ux); slb_big_free(lb->uy); slb_big_free(lb->uz);
  slb_big_free(lb->fx); slb_big_free(lb->fy); slb_big_free(lb->fz);
  slb_big_free(lb->bx); slb_big_free(lb->by); slb_big_free(lb->bz);
  slb_big_free(lb->rho); free(lb->name);
  free(S->objlist); free(S->objtab); free(S->objidx);
  free((slb_lbwrap *)lb);
}
