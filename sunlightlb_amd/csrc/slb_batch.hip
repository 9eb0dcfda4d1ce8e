/*
 * slb_batch.hip -- many independent small lattices ("problems") stepped
 * together: one kernel launch per time step covers every node of every
 * problem, the step loop is captured in a HIP graph and replayed, and the
 * launch shape (block size, nodes per thread, persistent grid size) is chosen
 * by a timing autotuner.
 *
 * Each problem is one periodic box with its own size, solid map, relaxation
 * time and object velocities (one right-hand side of a resistance problem).
 * Problems of the same box share the node tables (flags, objidx).  The step
 * is the single-lattice fused pull step (bbl + propagate + collide, see
 * slb_dev.hip) with the moving-wall velocity of a solid node computed from
 * its object row (src/lbobjs.c:565-589) instead of a stored field; the force
 * variant accumulates the momentum exchange of src/lbkernel.c:464-618
 * directly into per-object force and torque (the distribf sums of
 * src/lbmeasure.c:236-263) in FP64.
 *
 * Work decomposition: a tile is a contiguous run of blockDim*NPT nodes of one
 * problem; the tile table lists (problem, first node) for all tiles.  A
 * persistent grid of G blocks walks the tile table with stride G.
 */
#include <string.h>
#include "slb_dev.h"
#include "slb_kern.h"

struct BProb {
  int esi, esj, esk, plane;
  long N;
  long pop_off;   /* first population of this problem in the arena (tiled
                     layout pidx, whole tiles per problem) */
  long node_off;  /* first node of this problem's box in the node tables */
  int tab_off;    /* first object row (9 doubles each) */
  int nobj;
  int acc_off;    /* first object accumulator (6 doubles each) */
  int moving, stokes, skip_solid;
  double omega, comega, rho0;
  double bf[3];
};

#define BATCH_LDS_OBJ 64

/* Wall rows: per object row, in the working precision and scaled by
 * rho0/12, w0 = v - omega x pos (the velocity of the lattice origin in the
 * object's rigid motion) and omega, so that the wall velocity at solid node s
 * is U = w0 + omega x s and only c.U = c.w0 + c.(omega x s) enters a link. */
#define WROW 8

template <typename T, bool FORCE, int NPT>
__global__ void __launch_bounds__(512)
k_bstep(const T *__restrict__ src, T *__restrict__ dst,
        const uint16_t *__restrict__ flags, const int32_t *__restrict__ objidx,
        const double *__restrict__ tab, const T *__restrict__ wtab,
        const BProb *__restrict__ probs,
        const int2 *__restrict__ tiles, int ntiles, int xcd_chunk, double *__restrict__ acc) {
  __shared__ double sacc[FORCE ? BATCH_LDS_OBJ * 6 : 1];
  /* Work items v = blockIdx.x, + gridDim.x, ...  Workgroups are dispatched
   * round robin over the 8 XCDs (each with its own L2), so block b runs on
   * XCD b % 8.  With xcd_chunk > 0 (and gridDim.x a multiple of 8) item v is
   * tile (v % 8) * xcd_chunk + v / 8: each XCD walks one contiguous eighth of
   * the tile table in order, and the rows a tile pulls from (j and i
   * neighbours, a few tiles away in a small box) were fetched into the same
   * L2 moments before instead of by another XCD. */
  const int nv = xcd_chunk > 0 ? 8 * xcd_chunk : ntiles;
  for (int v = blockIdx.x; v < nv; v += gridDim.x) {
    const int tile = xcd_chunk > 0 ? (v & 7) * xcd_chunk + (v >> 3) : v;
    if (tile >= ntiles) continue;
    const int2 td = tiles[tile];
    const BProb P = probs[td.x];
    const bool lds = FORCE && P.nobj <= BATCH_LDS_OBJ;
    if (lds) {
      for (int q = threadIdx.x; q < P.nobj * 6; q += blockDim.x) sacc[q] = 0.0;
      __syncthreads();
    }
    const T *s = src + P.pop_off;
    T *d = dst + P.pop_off;
    const uint16_t *fl = flags + P.node_off;
    const int32_t *oid = objidx + P.node_off;
    PhysArgs A;
    A.omega = P.omega; A.comega = P.comega; A.rho0 = P.rho0;
    A.stokes = P.stokes; A.skip_solid = P.skip_solid;
    A.bfx = P.bf[0]; A.bfy = P.bf[1]; A.bfz = P.bf[2];
    const bool bfon = (P.bf[0] != 0.0 || P.bf[1] != 0.0 || P.bf[2] != 0.0);
#pragma unroll
    for (int q = 0; q < NPT; q++) {
      /* 15 pop_nodes(N) < 2^31 per problem (checked at commit): 32-bit
       * node and population indices */
      const int x = td.y + q * (int)blockDim.x + (int)threadIdx.x;
      if (x >= P.N) continue;
      const int i = x / P.plane;
      const int rem = x - i * P.plane;
      const int j = rem / P.esk, k = rem - j * P.esk;
      const AxisOff32 oi = axis_off32(i, P.esi, P.plane);
      const AxisOff32 oj = axis_off32(j, P.esj, P.esk);
      const AxisOff32 ok = axis_off32(k, P.esk, 1);
      const unsigned f = fl[x];
      const bool solid = (f & FLAG_SOLID) != 0;
      T v[15];
      /* the node's loads issue at raised wave priority (8.57 -> 8.47 ms per
       * step on the 512-box batch, FP64 15.83 -> 15.20 ms) */
      __builtin_amdgcn_s_setprio(3);
      v[0] = s[pidx32(0, x)];
      Links<1, 15>::run([&](auto LL) {
        constexpr int l = decltype(LL)::value;
        const bool bb = (f >> (l - 1)) & 1u;
        const int idx = bb ? pidx32(OPP(l), x) : pidx32(l, x + nb_off<l, -1>(oi, oj, ok));
        v[l] = s[idx];
      });
      __builtin_amdgcn_s_setprio(0);
      if ((P.moving || FORCE) && (f & FLAG_LINKS)) {
        double Fx = 0.0, Fy = 0.0, Fz = 0.0;
        Links<1, 15>::run([&](auto LL) {
          constexpr int l = decltype(LL)::value;
          if ((f >> (l - 1)) & 1u) {
            const int xs = x + nb_off<l, -1>(oi, oj, ok);
            T cu = T(0);
            if (P.moving) {
              /* wall velocity U = (rho0/12)(v + w x r) of the solid end of
               * the link (vspread, src/lbobjs.c:565-589, r = s - pos without
               * minimum image) = w0 + w x s with the scaled wall row; only
               * its component along c_l enters */
              int sn = x;
              int si = i, sj = j, sk = k;
              if (!solid) {
                sn = xs;
                si = i - CX(l); si += (si < 0) ? P.esi : 0; si -= (si >= P.esi) ? P.esi : 0;
                sj = j - CY(l); sj += (sj < 0) ? P.esj : 0; sj -= (sj >= P.esj) ? P.esj : 0;
                sk = k - CZ(l); sk += (sk < 0) ? P.esk : 0; sk -= (sk >= P.esk) ? P.esk : 0;
              }
              const T *wr = wtab + ((long)P.tab_off + oid[sn]) * WROW;
              const T sx = T(si), sy = T(sj), sz = T(sk);
              T cv = T(0);
              if (CX(l)) cv += T(CX(l)) * (wr[0] + wr[4] * sz - wr[5] * sy);
              if (CY(l)) cv += T(CY(l)) * (wr[1] + wr[5] * sx - wr[3] * sz);
              if (CZ(l)) cv += T(CZ(l)) * (wr[2] + wr[3] * sy - wr[4] * sx);
              cu = T(WL(l)) * cv;
            }
            if (FORCE && solid) {
              const double a = (double)v[l];
              const double b = (double)s[pidx32(l, xs)];
              const double g = -2.0 * (a - b + (double)cu);
              if (CX(l)) Fx += CX(l) * g;
              if (CY(l)) Fy += CY(l) * g;
              if (CZ(l)) Fz += CZ(l) * g;
            }
            if (P.moving) v[l] += cu;
          }
        });
        if (FORCE && solid && (Fx != 0.0 || Fy != 0.0 || Fz != 0.0)) {
          const int o = oid[x];
          const double *po = tab + ((long)P.tab_off + o) * SLBD_OBJ_STRIDE;
          const double rx = i - po[0], ry = j - po[1], rz = k - po[2];
          const double tq[6] = {Fx, Fy, Fz, ry * Fz - rz * Fy, rz * Fx - rx * Fz, rx * Fy - ry * Fx};
          double *dstp = lds ? sacc + 6 * o : acc + 6 * ((long)P.acc_off + o);
#pragma unroll
          for (int c = 0; c < 6; c++) atomicAdd(dstp + c, tq[c]);
        }
      }
      if (!(P.skip_solid && solid)) {
        T jbx = T(0), jby = T(0), jbz = T(0);
        if (bfon && !solid) { jbx = T(P.bf[0]); jby = T(P.bf[1]); jbz = T(P.bf[2]); }
        collide_node<T>(v, jbx, jby, jbz, A);
      }
#pragma unroll
      /* non-temporal: a batch step streams tens of GB (8.95 -> 8.59 ms at 512 boxes) */
      for (int l = 0; l < 15; l++) __builtin_nontemporal_store(v[l], &d[pidx32(l, x)]);
    }
    if (lds) {
      __syncthreads();
      for (int q = threadIdx.x; q < P.nobj * 6; q += blockDim.x)
        if (sacc[q] != 0.0) atomicAdd(acc + 6 * (long)P.acc_off + q, sacc[q]);
      __syncthreads();
    }
  }
}

template <typename T>
__global__ void k_bscale(double *__restrict__ a, long n, double s) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) a[x] *= s;
}

/* Pull flags for one box from its object index and mask (as k_flags). */
__global__ void __launch_bounds__(256)
k_bflags(const int32_t *__restrict__ objidx, const int32_t *__restrict__ mask,
         uint16_t *__restrict__ flags, int esi, int esj, int esk) {
  const int plane = esj * esk;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= plane) return;
  const int i = blockIdx.y, j = t / esk, k = t - j * esk;
  const long x = (long)i * plane + t;
  const AxisOff oi = axis_off(i, esi, plane);
  const AxisOff oj = axis_off(j, esj, esk);
  const AxisOff ok = axis_off(k, esk, 1);
  const int mx = mask[x];
  unsigned f = (objidx[x] >= 0) ? FLAG_SOLID : 0u;
  Links<1, 15>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    const int ms = mask[x + nb_off<l, -1>(oi, oj, ok)];
    if (((ms >> (l - 1)) & 1) || ((mx >> (OPP(l) - 1)) & 1)) f |= 1u << (l - 1);
  });
  flags[x] = (uint16_t)f;
}

/* ---------------------------------------------------------------------- */

struct BBox {
  int esi, esj, esk, nobj, nrhs, moving, stokes, skip_solid;
  double tau, rho0, bf[3];
  int32_t *objidx, *mask;   /* host copies until commit */
  double *tab;              /* nrhs * nobj rows */
};

struct slbd_batch {
  int prec;
  int nbox, cap;
  BBox *box;
  int committed;
  /* device arenas */
  void *pops[2];
  int cur;
  uint16_t *flags;
  int32_t *objidx;
  double *tab, *acc;
  void *wtab;               /* wall rows in the working precision */
  BProb *probs;
  int2 *tiles;
  int nprob, ntiles, nacc;
  long nnodes_total, npop_total, nnode_tab;
  BProb *hprobs;
  int *box_first_prob;
  /* launch configuration */
  int block, npt, grid;
  int tiled_block, tiled_npt;   /* tile table built for this shape */
  int xcd_order;                /* XCD-contiguous tile order (k_bstep) */
  hipStream_t stream;
  /* captured step loops, one per live buffer: a graph replays the pointers
   * it was captured with, so the loop captured from buffer 0 must not run
   * when buffer 1 is live (after an odd number of plain steps) */
  hipGraphExec_t gexec[2];
  int graph_valid[2];
  hipEvent_t ev0, ev1;
};

/* Drop the captured step loops (launch shape or tile table changed). */
static void invalidate_graphs(slbd_batch *B) {
  for (int c = 0; c < 2; c++)
    if (B->graph_valid[c]) {
      HIPCHECK(hipGraphExecDestroy(B->gexec[c]));
      B->graph_valid[c] = 0;
    }
}

static size_t bts(int prec) { return prec == 32 ? 4 : 8; }

extern "C" slbd_batch *slbd_batch_create(int prec) {
  if (slbd_device_count() < 1) {
    fprintf(stderr, "sunlightlb: no HIP device is visible; the GPU library cannot run\n");
    exit(EXIT_FAILURE);
  }
  slbd_batch *B = (slbd_batch *)calloc(1, sizeof(slbd_batch));
  B->prec = (prec == 32) ? 32 : 64;
  B->block = 256; B->npt = 1; B->grid = 0;
  /* XCD-contiguous tile order: measured slower on the 512-box batch (8.59 ->
   * 8.76 ms per step), off unless SLB_BATCH_XCD=1 */
  B->xcd_order = getenv("SLB_BATCH_XCD") ? atoi(getenv("SLB_BATCH_XCD")) : 0;
  HIPCHECK(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
  HIPCHECK(hipEventCreate(&B->ev0));
  HIPCHECK(hipEventCreate(&B->ev1));
  return B;
}

extern "C" int slbd_batch_add_box(slbd_batch *B, int esi, int esj, int esk,
                                  const int32_t *objidx, const int32_t *mask, int nobj,
                                  int nrhs, const double *tabs, double tau, double rho0,
                                  int stokes, int skip_solid, int moving, const double *bf) {
  if (B->committed || nrhs < 1) return -1;
  if (15 * pop_nodes((long)esi * esj * esk) > 0x7fffffffL) {
    fprintf(stderr, "slb_batch_add: a %d x %d x %d box is too large for a batch (at most "
            "2^31 populations per box)\n", esi, esj, esk);
    return -1;
  }
  if (B->nbox == B->cap) {
    B->cap = B->cap ? 2 * B->cap : 64;
    B->box = (BBox *)realloc(B->box, sizeof(BBox) * B->cap);
  }
  BBox *b = &B->box[B->nbox];
  const long N = (long)esi * esj * esk;
  b->esi = esi; b->esj = esj; b->esk = esk; b->nobj = nobj; b->nrhs = nrhs;
  b->moving = moving; b->stokes = stokes; b->skip_solid = skip_solid;
  b->tau = tau; b->rho0 = rho0;
  for (int c = 0; c < 3; c++) b->bf[c] = bf ? bf[c] : 0.0;
  b->objidx = (int32_t *)malloc(sizeof(int32_t) * N);
  b->mask = (int32_t *)malloc(sizeof(int32_t) * N);
  memcpy(b->objidx, objidx, sizeof(int32_t) * N);
  memcpy(b->mask, mask, sizeof(int32_t) * N);
  const size_t tb = sizeof(double) * SLBD_OBJ_STRIDE * (size_t)(nobj > 0 ? nobj : 1) * nrhs;
  b->tab = (double *)malloc(tb);
  if (nobj > 0) memcpy(b->tab, tabs, sizeof(double) * SLBD_OBJ_STRIDE * nobj * nrhs);
  return B->nbox++;
}

static void build_tiles(slbd_batch *B) {
  const long per = (long)B->block * B->npt;
  long nt = 0;
  for (int p = 0; p < B->nprob; p++) nt += (B->hprobs[p].N + per - 1) / per;
  int2 *h = (int2 *)malloc(sizeof(int2) * (nt > 0 ? nt : 1));
  long t = 0;
  for (int p = 0; p < B->nprob; p++)
    for (long s = 0; s < B->hprobs[p].N; s += per) { h[t].x = p; h[t].y = (int)s; t++; }
  if (B->tiles) HIPCHECK(hipFree(B->tiles));
  HIPCHECK(hipMalloc(&B->tiles, sizeof(int2) * (nt > 0 ? nt : 1)));
  HIPCHECK(hipMemcpy(B->tiles, h, sizeof(int2) * nt, hipMemcpyHostToDevice));
  free(h);
  B->ntiles = (int)nt;
  B->tiled_block = B->block;
  B->tiled_npt = B->npt;
  invalidate_graphs(B);
}

extern "C" int slbd_batch_commit(slbd_batch *B) {
  if (B->committed || B->nbox == 0) return -1;
  int nprob = 0, nacc = 0;
  long nodes = 0, pops = 0, ntab = 0;
  for (int b = 0; b < B->nbox; b++) {
    const long N = (long)B->box[b].esi * B->box[b].esj * B->box[b].esk;
    nprob += B->box[b].nrhs;
    nodes += N;
    pops += 15 * pop_nodes(N) * B->box[b].nrhs;
    ntab += (long)B->box[b].nobj * B->box[b].nrhs;
    nacc += B->box[b].nobj * B->box[b].nrhs;
  }
  B->nprob = nprob;
  B->nacc = nacc;
  B->nnode_tab = nodes;
  B->npop_total = pops;
  B->hprobs = (BProb *)calloc(nprob, sizeof(BProb));
  B->box_first_prob = (int *)malloc(sizeof(int) * B->nbox);
  HIPCHECK(hipMalloc(&B->pops[0], bts(B->prec) * pops));
  HIPCHECK(hipMalloc(&B->pops[1], bts(B->prec) * pops));
  HIPCHECK(hipMemsetAsync(B->pops[0], 0, bts(B->prec) * pops, B->stream));
  HIPCHECK(hipMemsetAsync(B->pops[1], 0, bts(B->prec) * pops, B->stream));
  HIPCHECK(hipMalloc(&B->flags, sizeof(uint16_t) * nodes));
  HIPCHECK(hipMalloc(&B->objidx, sizeof(int32_t) * nodes));
  int32_t *dmask;
  HIPCHECK(hipMalloc(&dmask, sizeof(int32_t) * nodes));
  HIPCHECK(hipMalloc(&B->tab, sizeof(double) * SLBD_OBJ_STRIDE * (ntab > 0 ? ntab : 1)));
  HIPCHECK(hipMalloc(&B->wtab, bts(B->prec) * WROW * (ntab > 0 ? ntab : 1)));
  double *hw = (double *)calloc((size_t)WROW * (ntab > 0 ? ntab : 1), sizeof(double));
  HIPCHECK(hipMalloc(&B->acc, sizeof(double) * 6 * (nacc > 0 ? nacc : 1)));
  HIPCHECK(hipMemsetAsync(B->acc, 0, sizeof(double) * 6 * (nacc > 0 ? nacc : 1), B->stream));
  long node_off = 0, pop_off = 0;
  int tab_off = 0, p = 0;
  for (int b = 0; b < B->nbox; b++) {
    BBox *x = &B->box[b];
    const long N = (long)x->esi * x->esj * x->esk;
    HIPCHECK(hipMemcpy(B->objidx + node_off, x->objidx, sizeof(int32_t) * N, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dmask + node_off, x->mask, sizeof(int32_t) * N, hipMemcpyHostToDevice));
    if (x->nobj > 0)
      HIPCHECK(hipMemcpy(B->tab + (long)tab_off * SLBD_OBJ_STRIDE, x->tab,
                         sizeof(double) * SLBD_OBJ_STRIDE * x->nobj * x->nrhs,
                         hipMemcpyHostToDevice));
    for (long o = 0; o < (long)x->nobj * x->nrhs; o++) {
      const double *po = x->tab + o * SLBD_OBJ_STRIDE;   /* pos, vel, ang */
      double *w = hw + (tab_off + o) * WROW;
      const double us = x->rho0 / 12.0;
      w[0] = us * (po[3] - (po[7] * po[2] - po[8] * po[1]));
      w[1] = us * (po[4] - (po[8] * po[0] - po[6] * po[2]));
      w[2] = us * (po[5] - (po[6] * po[1] - po[7] * po[0]));
      w[3] = us * po[6]; w[4] = us * po[7]; w[5] = us * po[8];
    }
    hipLaunchKernelGGL(k_bflags, dim3((x->esj * x->esk + 255) / 256, x->esi), dim3(256), 0,
                       B->stream, B->objidx + node_off, dmask + node_off, B->flags + node_off,
                       x->esi, x->esj, x->esk);
    LAUNCHCHECK();
    B->box_first_prob[b] = p;
    for (int r = 0; r < x->nrhs; r++, p++) {
      BProb *q = &B->hprobs[p];
      q->esi = x->esi; q->esj = x->esj; q->esk = x->esk; q->plane = x->esj * x->esk;
      q->N = N;
      q->pop_off = pop_off;
      q->node_off = node_off;
      q->tab_off = tab_off + r * x->nobj;
      q->nobj = x->nobj;
      q->acc_off = tab_off + r * x->nobj;
      q->moving = x->moving; q->stokes = x->stokes; q->skip_solid = x->skip_solid;
      q->omega = 1.0 / x->tau; q->comega = 1.0 - q->omega; q->rho0 = x->rho0;
      for (int c = 0; c < 3; c++) q->bf[c] = x->bf[c];
      pop_off += 15 * pop_nodes(N);
    }
    node_off += N;
    tab_off += x->nobj * x->nrhs;
    free(x->objidx); free(x->mask); free(x->tab);
    x->objidx = x->mask = NULL; x->tab = NULL;
  }
  if (B->prec == 32) {
    float *hf = (float *)malloc(sizeof(float) * WROW * (ntab > 0 ? ntab : 1));
    for (long q = 0; q < (long)WROW * (ntab > 0 ? ntab : 1); q++) hf[q] = (float)hw[q];
    HIPCHECK(hipMemcpy(B->wtab, hf, sizeof(float) * WROW * (ntab > 0 ? ntab : 1),
                       hipMemcpyHostToDevice));
    free(hf);
  } else {
    HIPCHECK(hipMemcpy(B->wtab, hw, sizeof(double) * WROW * (ntab > 0 ? ntab : 1),
                       hipMemcpyHostToDevice));
  }
  free(hw);
  HIPCHECK(hipStreamSynchronize(B->stream));
  HIPCHECK(hipFree(dmask));
  HIPCHECK(hipMalloc(&B->probs, sizeof(BProb) * nprob));
  HIPCHECK(hipMemcpy(B->probs, B->hprobs, sizeof(BProb) * nprob, hipMemcpyHostToDevice));
  long tot = 0;
  for (int q = 0; q < nprob; q++) tot += B->hprobs[q].N;
  B->nnodes_total = tot;
  build_tiles(B);
  B->committed = 1;
  return 0;
}

template <typename T, bool FORCE, int NPT>
static void blaunch_t(slbd_batch *B) {
  int grid = (B->grid > 0 && B->grid < B->ntiles) ? B->grid : B->ntiles;
  int chunk = 0;
  if (B->xcd_order && B->ntiles >= 64) {
    chunk = (B->ntiles + 7) / 8;
    grid = (grid + 7) & ~7;   /* every block stays on one XCD's range */
  }
  hipLaunchKernelGGL((k_bstep<T, FORCE, NPT>), dim3(grid), dim3(B->block), 0, B->stream,
                     (const T *)B->pops[B->cur], (T *)B->pops[B->cur ^ 1], B->flags, B->objidx,
                     B->tab, (const T *)B->wtab, B->probs, B->tiles, B->ntiles, chunk, B->acc);
  LAUNCHCHECK();
  B->cur ^= 1;
}

template <typename T>
static void blaunch(slbd_batch *B, bool force) {
  if (force) {
    if (B->npt == 2) blaunch_t<T, true, 2>(B); else if (B->npt == 4) blaunch_t<T, true, 4>(B);
    else blaunch_t<T, true, 1>(B);
  } else {
    if (B->npt == 2) blaunch_t<T, false, 2>(B); else if (B->npt == 4) blaunch_t<T, false, 4>(B);
    else blaunch_t<T, false, 1>(B);
  }
}

static void bstep(slbd_batch *B, bool force) {
  if (B->tiled_block != B->block || B->tiled_npt != B->npt) build_tiles(B);
  if (B->prec == 32) blaunch<float>(B, force); else blaunch<double>(B, force);
}

extern "C" void slbd_batch_config(slbd_batch *B, int block, int npt, int grid) {
  if (block != 64 && block != 128 && block != 256 && block != 512) block = 256;
  if (npt != 1 && npt != 2 && npt != 4) npt = 1;
  if (block != B->block || npt != B->npt || grid != B->grid) invalidate_graphs(B);
  B->block = block; B->npt = npt; B->grid = grid;
}

#define BATCH_GRAPH_STEPS 32

/* Capture BATCH_GRAPH_STEPS steps starting from the live buffer B->cur. */
static void build_graph(slbd_batch *B) {
  if (B->tiled_block != B->block || B->tiled_npt != B->npt) build_tiles(B);
  const int cur0 = B->cur;
  if (B->graph_valid[cur0]) return;
  hipGraph_t g;
  HIPCHECK(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
  for (int s = 0; s < BATCH_GRAPH_STEPS; s++) bstep(B, false);
  HIPCHECK(hipStreamEndCapture(B->stream, &g));
  HIPCHECK(hipGraphInstantiate(&B->gexec[cur0], g, NULL, NULL, 0));
  HIPCHECK(hipGraphDestroy(g));
  B->cur = cur0;   /* an even number of steps returns to the same buffer */
  B->graph_valid[cur0] = 1;
}

extern "C" void slbd_batch_run(slbd_batch *B, int nsteps, int use_graph) {
  if (!B->committed) return;
  int s = 0;
  if (use_graph && nsteps >= BATCH_GRAPH_STEPS) {
    /* the loop captured from the live buffer (a shape change since the
     * last capture rebuilds the tile table and drops both loops) */
    if (B->tiled_block != B->block || B->tiled_npt != B->npt) build_tiles(B);
    build_graph(B);
    for (; s + BATCH_GRAPH_STEPS <= nsteps; s += BATCH_GRAPH_STEPS)
      HIPCHECK(hipGraphLaunch(B->gexec[B->cur], B->stream));
  }
  for (; s < nsteps; s++) bstep(B, false);
}

extern "C" double slbd_batch_time(slbd_batch *B, int nsteps, int use_graph) {
  if (use_graph && B->committed && nsteps >= BATCH_GRAPH_STEPS) build_graph(B);
  HIPCHECK(hipStreamSynchronize(B->stream));
  HIPCHECK(hipEventRecord(B->ev0, B->stream));
  slbd_batch_run(B, nsteps, use_graph);
  HIPCHECK(hipEventRecord(B->ev1, B->stream));
  HIPCHECK(hipEventSynchronize(B->ev1));
  float ms = 0.0f;
  HIPCHECK(hipEventElapsedTime(&ms, B->ev0, B->ev1));
  return (double)ms;
}

/* measure(): force accumulators zeroed, two force-collecting steps, halved
 * (src/lbmeasure.c:30-49). */
extern "C" void slbd_batch_measure(slbd_batch *B) {
  if (!B->committed) return;
  const long n = 6L * (B->nacc > 0 ? B->nacc : 1);
  HIPCHECK(hipMemsetAsync(B->acc, 0, sizeof(double) * n, B->stream));
  bstep(B, true);
  bstep(B, true);
  hipLaunchKernelGGL(k_bscale<double>, dim3((n + 255) / 256), dim3(256), 0, B->stream, B->acc, n, 0.5);
  LAUNCHCHECK();
  HIPCHECK(hipStreamSynchronize(B->stream));
}

extern "C" int slbd_batch_forces(slbd_batch *B, int box, int rhs, double *frc, double *trq) {
  if (!B->committed || box < 0 || box >= B->nbox || rhs < 0 || rhs >= B->box[box].nrhs) return -1;
  const BProb *q = &B->hprobs[B->box_first_prob[box] + rhs];
  if (q->nobj == 0) return 0;
  double *h = (double *)malloc(sizeof(double) * 6 * q->nobj);
  HIPCHECK(hipMemcpy(h, B->acc + 6L * q->acc_off, sizeof(double) * 6 * q->nobj, hipMemcpyDeviceToHost));
  for (int o = 0; o < q->nobj; o++)
    for (int c = 0; c < 3; c++) {
      frc[3 * o + c] = h[6 * o + c];
      trq[3 * o + c] = h[6 * o + 3 + c];
    }
  free(h);
  return q->nobj;
}

extern "C" int slbd_batch_get_pops(slbd_batch *B, int box, int rhs, double *aos) {
  if (!B->committed || box < 0 || box >= B->nbox || rhs < 0 || rhs >= B->box[box].nrhs) return -1;
  HIPCHECK(hipStreamSynchronize(B->stream));
  const BProb *q = &B->hprobs[B->box_first_prob[box] + rhs];
  const size_t ts = bts(B->prec);
  const long np = 15 * pop_nodes(q->N);
  char *tmp = (char *)malloc(np * ts);
  HIPCHECK(hipMemcpy(tmp, (char *)B->pops[B->cur] + q->pop_off * ts, np * ts,
                     hipMemcpyDeviceToHost));
  for (int l = 0; l < 15; l++)
    for (long n = 0; n < q->N; n++)
      aos[15 * n + l] = (B->prec == 32) ? (double)((float *)tmp)[pidx(l, n)]
                                        : ((double *)tmp)[pidx(l, n)];
  free(tmp);
  return 0;
}

extern "C" long slbd_batch_nodes(const slbd_batch *B) { return B->nnodes_total; }
extern "C" int slbd_batch_nprob(const slbd_batch *B) { return B->nprob; }
extern "C" int slbd_batch_nbox(const slbd_batch *B) { return B->nbox; }
extern "C" int slbd_batch_box_dims(const slbd_batch *B, int box, int *dims) {
  if (box < 0 || box >= B->nbox) return -1;
  dims[0] = B->box[box].esi; dims[1] = B->box[box].esj; dims[2] = B->box[box].esk;
  return B->box[box].nrhs;
}
extern "C" void slbd_batch_get_config(const slbd_batch *B, int *cfg) {
  /* tile count of the configured shape (the table itself is rebuilt lazily
   * at the next run, so B->ntiles may still describe an earlier shape) */
  const long per = (long)B->block * B->npt;
  long nt = 0;
  for (int p = 0; p < B->nprob; p++) nt += (B->hprobs[p].N + per - 1) / per;
  cfg[0] = B->block; cfg[1] = B->npt; cfg[2] = B->grid; cfg[3] = (int)nt;
}
extern "C" int slbd_batch_cu_count(void) {
  hipDeviceProp_t prop;
  HIPCHECK(hipGetDeviceProperties(&prop, 0));
  return prop.multiProcessorCount;
}

extern "C" void slbd_batch_destroy(slbd_batch *B) {
  if (!B) return;
  HIPCHECK(hipStreamSynchronize(B->stream));
  invalidate_graphs(B);
  void *ps[] = {B->pops[0], B->pops[1], B->flags, B->objidx, B->tab, B->wtab, B->acc, B->probs,
                B->tiles};
  for (size_t q = 0; q < sizeof ps / sizeof ps[0]; q++)
    if (ps[q]) HIPCHECK(hipFree(ps[q]));
  for (int b = 0; b < B->nbox; b++) { free(B->box[b].objidx); free(B->box[b].mask); free(B->box[b].tab); }
  free(B->box); free(B->hprobs); free(B->box_first_prob);
  HIPCHECK(hipEventDestroy(B->ev0));
  HIPCHECK(hipEventDestroy(B->ev1));
  HIPCHECK(hipStreamDestroy(B->stream));
  free(B);
}

/* Every problem back to rest (populations and force sums zero). */
extern "C" void slbd_batch_reset(slbd_batch *B) {
  if (!B->committed) return;
  HIPCHECK(hipMemsetAsync(B->pops[0], 0, bts(B->prec) * B->npop_total, B->stream));
  HIPCHECK(hipMemsetAsync(B->pops[1], 0, bts(B->prec) * B->npop_total, B->stream));
  HIPCHECK(hipMemsetAsync(B->acc, 0, sizeof(double) * 6 * (B->nacc > 0 ? B->nacc : 1), B->stream));
  HIPCHECK(hipStreamSynchronize(B->stream));
  B->cur = 0;
}
