/*
 * slb_vacf.c -- tagged-tracer velocity autocorrelation by moment
 * propagation (the entry points of src/mpvacf.c).
 *
 * Two MP systems carry the computation: P (mp) holds the tracer density
 * tagged with its initial velocity component, P2 (mp2) the untagged tracer
 * density.  Both are advanced by the ordinary mpstep (GPU).  mpvacf()
 * removes the projection of P on the averaging weight psi and forms
 *   C(t) = sum_r psi P(r) u_b(r) / sum_r psi P2(r)
 * over fluid nodes with FP64 reductions on the GPU; mpvacfwr() integrates
 * D = C(0)/2 + sum_{t>0} C(t).
 *
 * The initialisations are one-off O(N) passes on the host mirrors, as in the
 * reference.  Four reference defects are not reproduced:
 *   - the rest-population contribution to P2(r,1) is added at the source
 *     node itself (src/mpvacf.c:120 adds it one node along +x, where it is
 *     lost when that node is solid);
 *   - the boundary link along c_10 carries the c_11 velocity factor like the
 *     other 13 links (src/mpvacf.c:207 omits it);
 *   - C(0) is normalised by N_tracers = sum_r psi^2 rho, as the reference's
 *     own comment states (src/mpvacf.c:245-246); its code also adds
 *     sum_r psi P2(r,1) (:267), counting every tracer twice, so that a fluid at
 *     rest gives D = (1 - alpha)/12 instead of the moment-propagation
 *     diffusion coefficient (1 - alpha)/6 (src/mpmisc.c:106);
 *   - mpvacfwr() with the file name "null" does not close a null stream
 *     (src/mpvacf.c:421), and nvals larger than the number of values writes
 *     every value instead of dividing by zero.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "slb_internal.h"

/* one autocorrelation run at a time, as the reference's file statics */
static int vacf_l1, vacf_l2, vacf_n, vacf_max;
static LBReal *vacf_val;

static int axis_of(char ch, const char *what) {
  switch (ch) {
    case 'x': case 'i': return 0;
    case 'y': case 'j': return 1;
    case 'z': case 'k': return 2;
    case '\0': {
      char msg[96];
      snprintf(msg, sizeof msg, "mpvacfinit: please supply %s direction", what);
      error(msg);
    }
  }
  {
    char msg[96];
    snprintf(msg, sizeof msg, "mpvacfinit: unrecognised %s direction", what);
    error(msg);
  }
  return 0;
}

static long node_of(const LBSys lb, int i, int j, int k) {
  i = (i % lb->esi + lb->esi) % lb->esi;
  j = (j % lb->esj + lb->esj) % lb->esj;
  k = (k % lb->esk + lb->esk) % lb->esk;
  return (long)lb->esk * ((long)lb->esj * i + j) + k;
}

static int opposite(const int *c, int l) {
  for (int m = 0; m < 15; m++)
    if (c[3 * m] == -c[3 * l] && c[3 * m + 1] == -c[3 * l + 1] && c[3 * m + 2] == -c[3 * l + 2])
      return m;
  return l;
}

void mpvacfinit(MPSys mp, MPSys mp2, char *vacfdir, int reqnvacf) {
  LBSys lb = mp->lb;
  free(vacf_val);
  vacf_val = (LBReal *)malloc(sizeof(LBReal) * (size_t)(reqnvacf + 1));
  if (vacf_val == NULL) {
    fprintf(stderr, "Not enough memory for %i vacf values\n", reqnvacf + 1);
    return;
  }
  vacf_max = reqnvacf;
  vacf_l1 = axis_of(vacfdir[0], "initial");
  vacf_l2 = axis_of(vacfdir[1], "final");
  if (lb->status != POST_COLLIDE) {
    fprintf(stderr, "vacf initialisation apparently out of position\n");
    wrstatus(lb, stdout);
  }
  if (lb->changed_objects) syncobjs(lb);
  slb_host_pops(lb);
  slb_host_rho(lb);
  slb_host_mask(lb);
  slb_host_p_w(mp);
  slb_host_p_w(mp2);
  const int *c = lb->c, *w = lb->w;
  const LBReal rho0 = lb->rho0, alpha = mp->alpha, css = 1.0 / 3.0;
  LBReal *p = mp->p, *p2 = mp2->p, *rho = lb->rho, *psi = lb->psi, *a = lb->a;
  int opp[15];
  for (int l = 0; l < 15; l++) opp[l] = opposite(c, l);
  /* One propagation of the tagged and untagged densities, P(r,1) and P2(r,1):
   * each fluid node y sends q_l = a_l(y) + (w_l/72)(rho0 - alpha rho(y))
   * (+ alpha rho(y) on the rest link) along c_l, weighted by psi(y), the
   * tagged copy times c_l[l1]; a link into a non-adsorbing solid node comes
   * back to y (bounce-back) with the opposite direction's velocity. */
  for (int i = 0; i < lb->esi; i++)
    for (int j = 0; j < lb->esj; j++)
      for (int k = 0; k < lb->esk; k++) {
        const long y = node_of(lb, i, j, k);
        if (lb->obj[y] == NULL) {
          const LBReal wf = psi ? psi[y] : 1.0;
          const LBReal base = (rho0 - alpha * rho[y]) / 72.0;
          for (int l = 0; l < 15; l++) {
            LBReal q = a[15 * y + l] + w[l] * base;
            if (l == 0) q += alpha * rho[y];
            const long x = node_of(lb, i + c[3 * l], j + c[3 * l + 1], k + c[3 * l + 2]);
            p[x] += wf * c[3 * l + vacf_l1] * q;
            p2[x] += wf * q;
          }
        } else {
          const int m = lb->mask[y];
          if (m == 0 || (m & 0x4000)) continue;
          for (int l = 1; l < 15; l++) {
            if (!((m >> (l - 1)) & 1)) continue;
            /* fluid node f = y + c_l sent its lbar population into y */
            const long f = node_of(lb, i + c[3 * l], j + c[3 * l + 1], k + c[3 * l + 2]);
            const int lb_ = opp[l];
            const LBReal wf = psi ? psi[f] : 1.0;
            const LBReal q = a[15 * f + lb_] + w[lb_] * (rho0 - alpha * rho[f]) / 72.0;
            p[f] += wf * c[3 * lb_ + vacf_l1] * q;
            p2[f] += wf * q;
          }
        }
      }
  for (long n = 0; n < lb->nnodes; n++)
    if (lb->obj[n] != NULL) { p[n] = 0.0; p2[n] = 0.0; }
  /* C(0) = sum_r psi^2 [rho css (1 - alpha) delta_ab + S_ab] / N_tracers,
   * N_tracers = sum_r psi^2 rho, S = Pi - (drho/3) I (fluid nodes) */
  LBReal tot = 0.0, ntr = 0.0;
  for (long n = 0; n < lb->nnodes; n++) {
    if (lb->obj[n] != NULL) continue;
    LBReal d = 0.0, s = 0.0;
    for (int l = 0; l < 15; l++) {
      d += a[15 * n + l];
      s += c[3 * l + vacf_l1] * c[3 * l + vacf_l2] * a[15 * n + l];
    }
    if (vacf_l1 == vacf_l2) s += rho[n] * css * (1.0 - alpha) - d / 3.0;
    const LBReal wf = psi ? psi[n] : 1.0;
    ntr += wf * wf * rho[n];
    tot += wf * wf * s;
  }
  vacf_val[0] = tot / ntr;
  vacf_n = 1;
  printf("MP '%s' %i : Initialised for vacf, init drn %c, final drn %c\n", mp->name, mp->step,
         vacfdir[0], vacfdir[1]);
  printf("MP '%s' %i : Initialised storage for %i vacf values\n", mp->name, mp->step, vacf_max);
  printf("MP '%s' %i : ntracers = %f\n", mp->name, mp->step, ntr);
  printf("MP '%s' %i : vacf(0) = %f\n", mp->name, mp->step, vacf_val[0]);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP '%s' %i : mpvacfinit done\n", mp->name, mp->step);
}

void mpvacfinit2(MPSys mp, MPSys mp2) {
  LBSys lb = mp->lb;
  if (lb->changed_objects) syncobjs(lb);
  slb_host_pops(lb);
  slb_host_rho(lb);
  slb_host_p_w(mp);
  slb_host_p_w(mp2);
  const int *c = lb->c;
  /* P(r,0) = psi u_a(r), P2(r,0) = psi on fluid nodes, u = j/rho */
  for (long n = 0; n < lb->nnodes; n++) {
    if (lb->obj[n] != NULL) continue;
    LBReal j = 0.0;
    for (int l = 1; l < 15; l++) j += c[3 * l + vacf_l1] * lb->a[15 * n + l];
    const LBReal wf = lb->psi ? lb->psi[n] : 1.0;
    mp->p[n] = wf * j / lb->rho[n];
    mp2->p[n] = wf;
  }
  vacf_n = 0;
  printf("MP '%s' %i : Re-initialised vacf calculation\n", mp->name, mp->step);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP '%s' %i : mpvacfinit2 done\n", mp->name, mp->step);
}

void mpvacf(MPSys mp, MPSys mp2) {
  LBSys lb = mp->lb;
  if (vacf_val == NULL) {
    fprintf(stderr, "mpvacf: mpvacfinit has not been called\n");
    return;
  }
  if (vacf_n > vacf_max) {
    fprintf(stderr, "Ran out of storage after %i vacf values\n", vacf_max);
    return;
  }
  if (lb->changed_objects) syncobjs(lb);
  slb_dev_geom(lb);
  slb_dev_pops(lb);
  slb_dev_rho(lb);
  slb_dev_p(mp);
  slb_dev_p(mp2);
  double out[2];
  slbd_vacf(PRIV(lb)->dev, MPRIV(mp)->dev, MPRIV(mp2)->dev, lb->psi, vacf_l2, out);
  MPRIV(mp)->p_loc = LOC_DEV;   /* P had its psi projection removed */
  vacf_val[vacf_n++] = out[0] / out[1];
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP '%s' %i : mpvacf done\n", mp->name, mp->step);
}

LBReal mpvacfwr(MPSys mp, MPSys mp2, char *vacffile, int nvals) {
  (void)mp2;
  FILE *fp = NULL;
  if (!myscmp("null", vacffile)) {
    if ((fp = fopen(vacffile, "w")) == NULL) {
      fprintf(stderr, "%s could not be opened\n", vacffile);
      return 0.0;
    }
    printf("MP '%s' %i : Writing vacf data to %s\n", mp->name, mp->step, vacffile);
  } else {
    printf("MP '%s' %i : Analyzing vacf data\n", mp->name, mp->step);
  }
  if (vacf_val == NULL || vacf_n == 0) {
    if (fp) fclose(fp);
    return 0.0;
  }
  int every = nvals > 0 ? vacf_n / nvals : 1;
  if (every < 1) every = 1;
  LBReal d = 0.5 * vacf_val[0];
  for (int t = 1; t < vacf_n; t++) {
    d += vacf_val[t];
    if (fp != NULL && t % every == 0) fprintf(fp, "%6i %g %g\n", t, vacf_val[t], d);
    if (t > vacf_n - 4)
      printf("MP '%s' %i : step, d = %8i %g %g\n", mp->name, mp->step, t, vacf_val[t], d);
  }
  if (fp != NULL) fclose(fp);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("MP '%s' %i : mpvacfwr done\n", mp->name, mp->step);
  return d;
}
