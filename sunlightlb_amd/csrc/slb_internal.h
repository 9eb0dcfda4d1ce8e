/*
 * slb_internal.h -- private state of the C host layer.
 *
 * Every LBSys handed out by newlb() is the first member of a larger block
 * (slb_lbwrap) whose second member holds the GPU lattice and the coherence
 * state of the host mirrors, so the public struct keeps the reference layout
 * and no lookup table is needed.  MPSys is wrapped the same way.
 */
#ifndef SLB_INTERNAL_H
#define SLB_INTERNAL_H

#include <stddef.h>
#include <stdint.h>
#include "sunlightlb.h"
#include "sunlightlb_gpu.h"
#include "slb_dev.h"

/* error() prints and exits (src/lbmisc.c:39-42); said here so that the
 * compiler and its analyser know that no call returns */
void error(char *s1) __attribute__((noreturn));

/* host allocations that end the program with error() when memory runs out */
#include <stdlib.h>
static inline void *slb_xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (p == NULL) error("out of host memory");
  return p;
}
static inline void *slb_xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (p == NULL) error("out of host memory");
  return p;
}
static inline void *slb_xrealloc(void *q, size_t n) {
  void *p = realloc(q, n ? n : 1);
  if (p == NULL) error("out of host memory");
  return p;
}

/* O(N) host passes split over host threads (slb_par.c): fn(ctx, lo, hi, tid)
 * runs on contiguous ranges in order of tid; slb_par_for returns how many
 * ranges it used, so the caller combines per-range slots 0 .. that - 1. */
#define SLB_PAR_MAX 64
typedef void (*slb_range_fn)(void *ctx, long lo, long hi, int tid);
int slb_host_threads(void);
int slb_par_for(long n, slb_range_fn fn, void *ctx);
/* zero-filled per-node host arrays, large ones on 2 MiB pages (slb_par.c);
 * release with slb_big_free */
void *slb_big_calloc(long n, size_t size);
void slb_big_free(void *p);

/* Where the newest copy of a mirrored array lives. */
enum { LOC_SYNC = 0, LOC_HOST = 1, LOC_DEV = 2 };

typedef struct {
  slbd_lattice *dev;
  int prec;
  int pops_loc, u_loc, f_loc, mask_loc, rho_loc;
  int objidx_dirty;     /* host obj[] changed since objidx was uploaded */
  int bf_dirty;         /* host bx/by/bz (or the fluid set) changed */
  int bf_uniform;       /* analysis result: 1 uniform on fluid nodes */
  double bfu[3];        /* the uniform value when bf_uniform */
  int bf_nonzero;       /* any fluid node carries a nonzero body force */
  int bf_uploaded;      /* per-node arrays are on the device */
  int nobj;             /* objects in the dense table */
  LBObj *objlist;       /* dense index -> object, list order */
  int objcap;
  double *objtab;       /* SLBD_OBJ_STRIDE doubles per object */
  int32_t *objidx;      /* host copy of the dense per-node object index */
  int nfluid;
  struct slb_mpwrap_s *mps;   /* MP systems attached to this lattice */
  int flags_stale;      /* device flags need rebuilding */
} slb_lbpriv;

typedef struct {
  LBSys_type pub;
  slb_lbpriv priv;
} slb_lbwrap;

typedef struct {
  slbd_mp *dev;         /* device fields of this MP system */
  int p_loc;            /* P field mirror state */
  int nmom;             /* 1, or 3 when displacement moments are tracked */
  int axis;
  double pads;          /* partial adsorption probability */
  struct slb_mpwrap_s *next;
} slb_mppriv;

typedef struct slb_mpwrap_s {
  MPSys_type pub;
  slb_mppriv priv;
} slb_mpwrap;

static inline slb_lbpriv *PRIV(LBSys lb) { return &((slb_lbwrap *)lb)->priv; }
static inline slb_mppriv *MPRIV(MPSys mp) { return &((slb_mpwrap *)mp)->priv; }

/* mirror coherence (slb_core.c) */
void slb_host_pops(LBSys lb);          /* host a[] readable */
void slb_host_pops_w(LBSys lb);        /* host a[] about to be written */
void slb_dev_pops(LBSys lb);           /* device populations current */
void slb_host_u(LBSys lb);
void slb_host_u_w(LBSys lb);
void slb_dev_u(LBSys lb);
void slb_host_f(LBSys lb);
void slb_host_f_w(LBSys lb);
void slb_dev_f(LBSys lb);
void slb_host_mask(LBSys lb);
void slb_host_mask_w(LBSys lb);
void slb_dev_geom(LBSys lb);           /* objidx, mask and flags current */
void slb_host_rho(LBSys lb);
void slb_host_rho_w(LBSys lb);
void slb_dev_rho(LBSys lb);
void slb_params(LBSys lb, slbd_params *P);
void slb_objtab(LBSys lb);             /* refresh the dense object table */
void slb_obj_changed(LBSys lb);        /* host obj[] was modified */
void slb_bf_changed(LBSys lb);         /* host bx/by/bz were modified */
int  slb_objindex(LBSys lb, LBObj p);  /* dense index of an object */
void slb_after_collide(LBSys lb);      /* step++, status, signal */

/* MP (slb_mp.c) */
void slb_host_p(MPSys mp);
void slb_host_p_w(MPSys mp);
void slb_dev_p(MPSys mp);

void slb_forcestep(LBSys lb);          /* bblforce+propagate+collide */
void slb_clock_reset(void);            /* newlb() resets the run clock */

LBReal *slb_ralloc(long n);
int *slb_ialloc(long n);

#endif
