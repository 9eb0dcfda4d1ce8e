/*
 * slb_batchapi.c -- batched resistance problems (sunlightlb_gpu.h): many
 * small periodic boxes, each with several right-hand sides (object velocity
 * sets), stepped together by the batched kernel of slb_batch.hip.  A box is
 * described by an ordinary lattice built with the reference API (newlb,
 * insertsph, ...), so its solid map and boundary-link mask are the ones the
 * single-lattice path would use.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "slb_internal.h"

struct slb_batch_s {
  slbd_batch *d;
  int prec;
  int stepped;
};

SLBBatch slb_batch_new(int prec) {
  SLBBatch b = (SLBBatch)calloc(1, sizeof(struct slb_batch_s));
  if (b == NULL) error("slb_batch_new: out of host memory");
  b->prec = (prec == 32) ? 32 : 64;
  b->d = slbd_batch_create(b->prec);
  return b;
}

int slb_batch_add(SLBBatch b, LBSys lb, int nrhs, const double *objvel) {
  slb_lbpriv *S = PRIV(lb);
  if (lb->changed_objects) syncobjs(lb);
  slb_dev_geom(lb);
  slb_host_mask(lb);
  slb_objtab(lb);
  /* body force: uniform over fluid nodes, plus the live dpdx trio */
  double bf[3] = {lb->dpdx, lb->dpdy, lb->dpdz};
  int have = 0;
  for (long n = 0; n < lb->nnodes; n++) {
    if (lb->obj[n] != NULL) continue;
    if (!have) { bf[0] += lb->bx[n]; bf[1] += lb->by[n]; bf[2] += lb->bz[n]; have = 1; continue; }
    if (lb->bx[n] + lb->dpdx != bf[0] || lb->by[n] + lb->dpdy != bf[1] ||
        lb->bz[n] + lb->dpdz != bf[2]) {
      fprintf(stderr, "slb_batch_add: non-uniform body force is not supported in a batch\n");
      return -1;
    }
  }
  const int nobj = S->nobj;
  double *tabs = (double *)slb_xmalloc(sizeof(double) * SLBD_OBJ_STRIDE * (nobj > 0 ? nobj : 1) * nrhs);
  for (int r = 0; r < nrhs; r++)
    for (int o = 0; o < nobj; o++) {
      double *t = tabs + ((long)r * nobj + o) * SLBD_OBJ_STRIDE;
      memcpy(t, S->objtab + (long)o * SLBD_OBJ_STRIDE, sizeof(double) * SLBD_OBJ_STRIDE);
      if (objvel) memcpy(t + 3, objvel + ((long)r * nobj + o) * 6, sizeof(double) * 6);
    }
  int moving = lb->objects_moving || objvel != NULL;
  int box = slbd_batch_add_box(b->d, lb->esi, lb->esj, lb->esk, S->objidx, lb->mask, nobj, nrhs,
                               tabs, lb->tau, lb->rho0, lb->stokes_flow, lb->skip_solid, moving,
                               bf);
  free(tabs);
  return box;
}

int slb_batch_commit(SLBBatch b) { return slbd_batch_commit(b->d); }

void slb_batch_set_config(SLBBatch b, int block, int nodes_per_thread, int grid) {
  slbd_batch_config(b->d, block, nodes_per_thread, grid);
}

void slb_batch_get_config(SLBBatch b, int cfg[4]) { slbd_batch_get_config(b->d, cfg); }

/* Shape class of a batch: precision, right-hand sides per problem bucket,
 * and a histogram of box edge lengths in bins of 8 with counts rounded to a
 * power of two.  Batches of one class share a tuned launch shape. */
static void shape_key(SLBBatch b, char *key, size_t len) {
  int hist[64] = {0};
  int nbox = slbd_batch_nbox(b->d), nprob = slbd_batch_nprob(b->d);
  for (int q = 0; q < nbox; q++) {
    int dims[3];
    slbd_batch_box_dims(b->d, q, dims);
    int e = dims[0] > dims[1] ? dims[0] : dims[1];
    e = e > dims[2] ? e : dims[2];
    int bin = e / 8;
    if (bin > 63) bin = 63;
    hist[bin]++;
  }
  int off = snprintf(key, len, "fp%d-p%d", b->prec, nprob > 0 ? 1 << (31 - __builtin_clz(nprob)) : 0);
  for (int q = 0; q < 64 && off < (int)len - 16; q++)
    if (hist[q]) off += snprintf(key + off, len - off, "-e%dx%d", 8 * q, 1 << (31 - __builtin_clz(hist[q])));
}

int slb_batch_autotune(SLBBatch b, const char *cache_path) {
  if (b->stepped) {
    fprintf(stderr, "slb_batch_autotune: call before the first step (tuning steps the state)\n");
    return -1;
  }
  char key[512];
  shape_key(b, key, sizeof key);
  if (cache_path) {
    FILE *f = fopen(cache_path, "r");
    if (f) {
      char k[512];
      int blk, npt, grid;
      while (fscanf(f, "%511s %d %d %d", k, &blk, &npt, &grid) == 4)
        if (strcmp(k, key) == 0) {
          fclose(f);
          slbd_batch_config(b->d, blk, npt, grid);
          return 1;
        }
      fclose(f);
    }
  }
  static const int blocks[] = {64, 128, 256, 512};
  static const int npts[] = {1, 2, 4};
  const int cu = slbd_batch_cu_count();
  const int grids[] = {0, 4 * cu, 8 * cu, 16 * cu};
  double best = 1e300;
  int bb = 256, bn = 1, bg = 0;
  for (size_t i = 0; i < sizeof blocks / sizeof blocks[0]; i++)
    for (size_t j = 0; j < sizeof npts / sizeof npts[0]; j++)
      for (size_t g = 0; g < sizeof grids / sizeof grids[0]; g++) {
        slbd_batch_config(b->d, blocks[i], npts[j], grids[g]);
        slbd_batch_run(b->d, 2, 0);
        double ms = slbd_batch_time(b->d, 4, 0);
        if (ms < best) { best = ms; bb = blocks[i]; bn = npts[j]; bg = grids[g]; }
      }
  slbd_batch_config(b->d, bb, bn, bg);
  /* the trial steps moved the state: start again from rest */
  slbd_batch_reset(b->d);
  if (cache_path) {
    FILE *f = fopen(cache_path, "a");
    if (f) { fprintf(f, "%s %d %d %d\n", key, bb, bn, bg); fclose(f); }
  }
  return 0;
}

void slb_batch_run(SLBBatch b, int nsteps) {
  b->stepped = 1;
  slbd_batch_run(b->d, nsteps, 1);
}

double slb_batch_time_steps(SLBBatch b, int nsteps) {
  b->stepped = 1;
  return slbd_batch_time(b->d, nsteps, 1);
}

void slb_batch_measure(SLBBatch b) {
  b->stepped = 1;
  slbd_batch_measure(b->d);
}

int slb_batch_forces(SLBBatch b, int box, int rhs, double *frc, double *trq) {
  return slbd_batch_forces(b->d, box, rhs, frc, trq);
}

int slb_batch_get_pops(SLBBatch b, int box, int rhs, double *aos) {
  return slbd_batch_get_pops(b->d, box, rhs, aos);
}

long slb_batch_nodes(SLBBatch b) { return slbd_batch_nodes(b->d); }

void slb_batch_free(SLBBatch b) {
  if (!b) return;
  slbd_batch_destroy(b->d);
  free(b);
}
