/*
 * CPU oracle for the SunlightLB D3Q15 lattice-Boltzmann step.
 *
 * TEST INFRASTRUCTURE ONLY (see slb_oracle.h).  A clean-room restatement of
 * the reference algorithm written from its mathematics: table-driven loops
 * over the 15 links instead of the reference's unrolled code.  Every
 * function names the reference file:line it restates.  The functions are
 * pinned against the reference library itself (golden fixtures written by
 * tests/golden/make_golden.py) and against closed forms in tests/.
 *
 * The partial-adsorption probability and the displacement-moment fields of
 * orc_mpstep() have no reference counterpart; they are this project's own
 * definitions, pinned by special cases (pads = 0 and pads = 1 reduce to the
 * reference's reflecting and adsorbing rules) and by random-walk closed forms.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "slb_oracle.h"

/* Velocity set and integer weights: src/lbmisc.c:89-104.  b = sum w = 72
 * = rho0 (src/lbmisc.c:120). */
const int orc_c[15][3] = {
  { 0, 0, 0},
  { 1, 0, 0}, {-1, 0, 0}, { 0, 1, 0}, { 0,-1, 0}, { 0, 0, 1}, { 0, 0,-1},
  { 1, 1, 1}, {-1, 1, 1}, { 1,-1, 1}, {-1,-1, 1},
  { 1, 1,-1}, {-1, 1,-1}, { 1,-1,-1}, {-1,-1,-1}
};
const int orc_w[15] = {16, 8, 8, 8, 8, 8, 8, 1, 1, 1, 1, 1, 1, 1, 1};
const int orc_opp[15] = {0, 2, 1, 4, 3, 6, 5, 14, 13, 12, 11, 10, 9, 8, 7};

static inline int wrap(int v, int n) { v %= n; return v < 0 ? v + n : v; }

static inline long nb(int esi, int esj, int esk, int i, int j, int k, int l, int sgn) {
  int ii = wrap(i + sgn * orc_c[l][0], esi);
  int jj = wrap(j + sgn * orc_c[l][1], esj);
  int kk = wrap(k + sgn * orc_c[l][2], esk);
  return (long)esk * ((long)esj * ii + jj) + kk;
}

/* ------------------------------------------------------------------ */
/* Geometry: src/lbobjs.c:204-232 (insertsph), :526-560 (makemask),
 * :648-655 (countfluid), :671-714 (countmask). */

int orc_insertsph(int esi, int esj, int esk, int32_t *objid, int id,
                  double rad, double ox, double oy, double oz, int *overlap) {
  /* A node is inside when (i-ox)^2+(j-oy)^2+(k-oz)^2 < rad^2, evaluated in
   * FP64 left to right; no periodic wrap; the last insert wins. */
  volatile double radsq = rad * rad;
  int ns = 0, ov = 0;
  for (int i = 0; i < esi; i++)
    for (int j = 0; j < esj; j++)
      for (int k = 0; k < esk; k++) {
        volatile double dx = i - ox, dy = j - oy, dz = k - oz;
        volatile double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
        volatile double s = dx2 + dy2;
        volatile double rsq = s + dz2;
        if (rsq < radsq) {
          long n = (long)esk * ((long)esj * i + j) + k;
          if (objid[n] >= 0) ov = 1;
          objid[n] = id; ns++;
        }
      }
  if (overlap) *overlap = ov;
  return ns;
}

void orc_makemask(int esi, int esj, int esk, const int32_t *objid, int32_t *mask) {
  /* Bit (l-1) is set on a solid node iff its neighbour at +c_l is fluid. */
  for (int i = 0; i < esi; i++)
    for (int j = 0; j < esj; j++)
      for (int k = 0; k < esk; k++) {
        long n = (long)esk * ((long)esj * i + j) + k;
        int32_t m = 0;
        if (objid[n] >= 0)
          for (int l = 1; l < 15; l++)
            if (objid[nb(esi, esj, esk, i, j, k, l, +1)] < 0) m |= 1 << (l - 1);
        mask[n] = m;
      }
}

int orc_countfluid(long nnodes, const int32_t *objid) {
  int c = 0;
  for (long n = 0; n < nnodes; n++) c += objid[n] < 0;
  return c;
}

int orc_countmask(long nnodes, const int32_t *mask, int links) {
  /* links != 0: weighted link count (sum of w_l); else count of axis faces. */
  int c = 0;
  for (long n = 0; n < nnodes; n++) {
    int32_t m = mask[n];
    if (m <= 0) continue;
    for (int l = 1; l < 15; l++) {
      if (!(m & (1 << (l - 1)))) continue;
      if (links) c += orc_w[l];
      else if (l <= 6) c++;
    }
  }
  return c;
}

/* ------------------------------------------------------------------ */
/* Bounce-back: src/lbkernel.c:58-170 (stationary), :179-321 (stationary,
 * with force), :330-454 (moving), :464-618 (moving, with force).
 * For every solid node s and set mask bit (l-1), with f = s + c_l:
 *   a = a_l(s), b = a_lbar(f), fac = w_l (c_l . U(s))   (0 if stationary)
 *   a_l(s) <- b + fac,  a_lbar(f) <- a - fac,
 *   force(s) += 2 (a - b - fac) c_l                      (force variants)
 * U is the precomputed node velocity from vspread (rho0/12 scaling). */

void orc_bbl(int esi, int esj, int esk, double *a, const int32_t *mask,
             const double *ux, const double *uy, const double *uz,
             double *fx, double *fy, double *fz) {
  for (int i = 0; i < esi; i++)
    for (int j = 0; j < esj; j++)
      for (int k = 0; k < esk; k++) {
        long s = (long)esk * ((long)esj * i + j) + k;
        int32_t m = mask[s];
        if (m <= 0) continue;
        for (int l = 1; l < 15; l++) {
          if (!(m & (1 << (l - 1)))) continue;
          long f = nb(esi, esj, esk, i, j, k, l, +1);
          int lb = orc_opp[l];
          double av = a[15 * s + l], bv = a[15 * f + lb], fac = 0.0;
          if (ux)
            fac = orc_w[l] * (orc_c[l][0] * ux[s] + orc_c[l][1] * uy[s] + orc_c[l][2] * uz[s]);
          a[15 * s + l] = bv + fac;
          a[15 * f + lb] = av - fac;
          if (fx) {
            double g = 2.0 * (av - bv - fac);
            fx[s] += g * orc_c[l][0]; fy[s] += g * orc_c[l][1]; fz[s] += g * orc_c[l][2];
          }
        }
      }
}

/* Streaming: src/lbkernel.c:623-762.  a_l(x) <- a_l(x - c_l), periodic.
 * The reference does three in-place axis sweeps; a pure permutation, so an
 * out-of-place gather gives identical values. */
void orc_propagate(int esi, int esj, int esk, double *a) {
  long N = (long)esi * esj * esk;
  double *t = (double *)malloc(sizeof(double) * 15 * N);
  memcpy(t, a, sizeof(double) * 15 * N);
  for (int i = 0; i < esi; i++)
    for (int j = 0; j < esj; j++)
      for (int k = 0; k < esk; k++) {
        long x = (long)esk * ((long)esj * i + j) + k;
        for (int l = 1; l < 15; l++) a[15 * x + l] = t[15 * nb(esi, esj, esk, i, j, k, l, -1) + l];
      }
  free(t);
}

/* Moments of one node: drho, j (3), Pi (6: xx yy zz xy xz yz). */
static void node_moments(const double *al, double *drho, double j[3], double pi[6]) {
  double d = 0, jx = 0, jy = 0, jz = 0, p[6] = {0, 0, 0, 0, 0, 0};
  for (int l = 0; l < 15; l++) {
    double v = al[l];
    int cx = orc_c[l][0], cy = orc_c[l][1], cz = orc_c[l][2];
    d += v; jx += cx * v; jy += cy * v; jz += cz * v;
    p[0] += cx * cx * v; p[1] += cy * cy * v; p[2] += cz * cz * v;
    p[3] += cx * cy * v; p[4] += cx * cz * v; p[5] += cy * cz * v;
  }
  *drho = d; j[0] = jx; j[1] = jy; j[2] = jz;
  for (int q = 0; q < 6; q++) pi[q] = p[q];
}

/* Rebuild the 15 shifted populations from (drho, j, s) where s is the
 * traceless-shifted stress:
 *   a_l = (w_l/72) [ drho + 3 c_l.j + (9/2)(c_l c_l : s - tr(s)/3) ]
 * (src/lbkernel.c:837-865, src/lbstore.c:375-399). */
static void node_rebuild(double *al, double drho, const double j[3], const double s[6]) {
  double tr = s[0] + s[1] + s[2];
  for (int l = 0; l < 15; l++) {
    int cx = orc_c[l][0], cy = orc_c[l][1], cz = orc_c[l][2];
    double cc = cx * cx * s[0] + cy * cy * s[1] + cz * cz * s[2]
              + 2.0 * (cx * cy * s[3] + cx * cz * s[4] + cy * cz * s[5]);
    al[l] = orc_w[l] / 72.0 * (drho + 3.0 * (cx * j[0] + cy * j[1] + cz * j[2])
                               + 4.5 * (cc - tr / 3.0));
  }
}

/* Collision: src/lbkernel.c:769-877.  Project on (drho, j, s = Pi - drho/3 I),
 * add the body force to j on fluid nodes (:816-819), relax s with
 * omega = 1/tau (Stokes: s' = (1-omega)s; Navier-Stokes adds
 * omega j j/(rho0+drho)), drop the 5 ghost moments, rebuild. */
void orc_collide(long nnodes, double *a, const int32_t *objid,
                 const double *bx, const double *by, const double *bz,
                 double tau, double rho0, int stokes_flow, int skip_solid) {
  double omega = 1.0 / tau, comega = 1.0 - omega;
  for (long n = 0; n < nnodes; n++) {
    if (skip_solid && objid[n] >= 0) continue;
    double drho, j[3], pi[6], s[6];
    node_moments(a + 15 * n, &drho, j, pi);
    if (objid[n] < 0) { j[0] += bx[n]; j[1] += by[n]; j[2] += bz[n]; }
    for (int q = 0; q < 6; q++) s[q] = pi[q] - (q < 3 ? drho / 3.0 : 0.0);
    if (stokes_flow) {
      for (int q = 0; q < 6; q++) s[q] *= comega;
    } else {
      double o = omega / (rho0 + drho);
      s[0] = comega * s[0] + o * j[0] * j[0];
      s[1] = comega * s[1] + o * j[1] * j[1];
      s[2] = comega * s[2] + o * j[2] * j[2];
      s[3] = comega * s[3] + o * j[0] * j[1];
      s[4] = comega * s[4] + o * j[0] * j[2];
      s[5] = comega * s[5] + o * j[1] * j[2];
    }
    node_rebuild(a + 15 * n, drho, j, s);
  }
}

/* One full step as lbstep: bbl, propagate, collide (src/lbkernel.c:49-51). */
void orc_lbstep(int esi, int esj, int esk, double *a, const int32_t *mask,
                const int32_t *objid, const double *bx, const double *by,
                const double *bz, const double *ux, const double *uy,
                const double *uz, double tau, double rho0, int stokes_flow,
                int skip_solid, int moving, long nsteps) {
  long N = (long)esi * esj * esk;
  for (long t = 0; t < nsteps; t++) {
    orc_bbl(esi, esj, esk, a, mask, moving ? ux : NULL, uy, uz, NULL, NULL, NULL);
    orc_propagate(esi, esj, esk, a);
    orc_collide(N, a, objid, bx, by, bz, tau, rho0, stokes_flow, skip_solid);
  }
}

/* ------------------------------------------------------------------ */
/* Node velocity of solid nodes: src/lbobjs.c:565-589.
 * U = rho0 (v + omega x r)/12, r = node - object centre (no minimum image);
 * fluid nodes get 0. */
void orc_vspread(int esi, int esj, int esk, const int32_t *objid,
                 const double *objpos, const double *objvel, const double *objang,
                 double rho0, double *ux, double *uy, double *uz) {
  for (int i = 0; i < esi; i++)
    for (int j = 0; j < esj; j++)
      for (int k = 0; k < esk; k++) {
        long n = (long)esk * ((long)esj * i + j) + k;
        int o = objid[n];
        if (o < 0) { ux[n] = uy[n] = uz[n] = 0.0; continue; }
        const double *P = objpos + 3 * o, *V = objvel + 3 * o, *W = objang + 3 * o;
        double rx = i - P[0], ry = j - P[1], rz = k - P[2];
        ux[n] = rho0 * (V[0] + W[1] * rz - W[2] * ry) / 12.0;
        uy[n] = rho0 * (V[1] + W[2] * rx - W[0] * rz) / 12.0;
        uz[n] = rho0 * (V[2] + W[0] * ry - W[1] * rx) / 12.0;
      }
}

/* Mean superficial velocity contribution: src/lbmeasure.c:58-106.
 * Sum over fluid nodes of (j + fac b)/rho, rho = rho0 + sum a, optionally
 * weighted by psi^2 and renormalised to the fluid count, divided by nnodes.
 * fac = +1/2 post-propagate, -1/2 post-collide, 0 if uncorrected. */
void orc_accvmean(long nnodes, const double *a, const int32_t *objid,
                  const double *bx, const double *by, const double *bz,
                  const double *psi, double rho0, double fac, double out[3]) {
  double tx = 0, ty = 0, tz = 0, wn = 0;
  long nf = 0;
  for (long n = 0; n < nnodes; n++) {
    if (objid[n] >= 0) continue;
    double d, j[3], pi[6];
    node_moments(a + 15 * n, &d, j, pi);
    double rho = rho0 + d, wf = 1.0;
    if (psi) wf = psi[n] * psi[n];
    tx += wf * (j[0] + fac * bx[n]) / rho;
    ty += wf * (j[1] + fac * by[n]) / rho;
    tz += wf * (j[2] + fac * bz[n]) / rho;
    wn += wf; nf++;
  }
  double r = (wn != 0.0) ? nf / wn : 0.0;
  out[0] = tx * r / nnodes; out[1] = ty * r / nnodes; out[2] = tz * r / nnodes;
}

/* Velocity field on fluid nodes (src/lbmeasure.c:128-166 with prefac 1,
 * into zeroed arrays): u = (j + fac b)/rho; solid nodes untouched. */
void orc_vfield(long nnodes, const double *a, const int32_t *objid,
                const double *bx, const double *by, const double *bz,
                double rho0, double fac, double *ux, double *uy, double *uz) {
  for (long n = 0; n < nnodes; n++) {
    if (objid[n] >= 0) continue;
    double d, j[3], pi[6];
    node_moments(a + 15 * n, &d, j, pi);
    double rho = rho0 + d;
    ux[n] = (j[0] + fac * bx[n]) / rho;
    uy[n] = (j[1] + fac * by[n]) / rho;
    uz[n] = (j[2] + fac * bz[n]) / rho;
  }
}

/* Per-object force and torque: src/lbmeasure.c:236-263.
 * F = sum f, T = sum (r - R) x f over the object's nodes, no minimum image. */
void orc_distribf(int esi, int esj, int esk, const int32_t *objid, int nobj,
                  const double *objpos, const double *fx, const double *fy,
                  const double *fz, double *frc, double *trq) {
  memset(frc, 0, sizeof(double) * 3 * nobj);
  memset(trq, 0, sizeof(double) * 3 * nobj);
  for (int i = 0; i < esi; i++)
    for (int j = 0; j < esj; j++)
      for (int k = 0; k < esk; k++) {
        long n = (long)esk * ((long)esj * i + j) + k;
        int o = objid[n];
        if (o < 0) continue;
        double rx = i - objpos[3 * o], ry = j - objpos[3 * o + 1], rz = k - objpos[3 * o + 2];
        frc[3 * o] += fx[n]; frc[3 * o + 1] += fy[n]; frc[3 * o + 2] += fz[n];
        trq[3 * o]     += ry * fz[n] - rz * fy[n];
        trq[3 * o + 1] += rz * fx[n] - rx * fz[n];
        trq[3 * o + 2] += rx * fy[n] - ry * fx[n];
      }
}

/* measure(): src/lbmeasure.c:30-49.  Two force-measuring steps with
 * velocity weights 1/4, 1/2, 1/4, node force halved, then distributed. */
void orc_measure(int esi, int esj, int esk, double *a, const int32_t *mask,
                 const int32_t *objid, const double *bx, const double *by,
                 const double *bz, const double *ux, const double *uy,
                 const double *uz, double tau, double rho0, int stokes_flow,
                 int skip_solid, int moving, int vel_corrected, int nobj,
                 const double *objpos, double *fx, double *fy, double *fz,
                 double vmean[3], double *frc, double *trq) {
  long N = (long)esi * esj * esk;
  double fac = vel_corrected ? -0.5 : 0.0, v[3];
  static const double wgt[3] = {0.25, 0.5, 0.25};
  memset(fx, 0, sizeof(double) * N); memset(fy, 0, sizeof(double) * N);
  memset(fz, 0, sizeof(double) * N);
  vmean[0] = vmean[1] = vmean[2] = 0.0;
  for (int stage = 0; stage < 3; stage++) {
    if (stage > 0) {
      orc_bbl(esi, esj, esk, a, mask, moving ? ux : NULL, uy, uz, fx, fy, fz);
      orc_propagate(esi, esj, esk, a);
      orc_collide(N, a, objid, bx, by, bz, tau, rho0, stokes_flow, skip_solid);
    }
    orc_accvmean(N, a, objid, bx, by, bz, NULL, rho0, fac, v);
    for (int d = 0; d < 3; d++) vmean[d] += wgt[stage] * v[d];
  }
  for (long n = 0; n < N; n++) { fx[n] *= 0.5; fy[n] *= 0.5; fz[n] *= 0.5; }
  orc_distribf(esi, esj, esk, objid, nobj, objpos, fx, fy, fz, frc, trq);
}

/* Equilibrium populations from (rho, u): src/lbstore.c:346-399 (lbloadvel).
 * drho = rho - rho0, j = rho u, s = 0 (Stokes) or u j (Navier-Stokes). */
void orc_equilibrium(long nnodes, double *a, const double *rho,
                     const double *ux, const double *uy, const double *uz,
                     double rho0, int stokes_flow) {
  for (long n = 0; n < nnodes; n++) {
    double r = rho[n], j[3] = {r * ux[n], r * uy[n], r * uz[n]}, s[6] = {0, 0, 0, 0, 0, 0};
    if (!stokes_flow) {
      s[0] = ux[n] * j[0]; s[1] = uy[n] * j[1]; s[2] = uz[n] * j[2];
      s[3] = ux[n] * j[1]; s[4] = ux[n] * j[2]; s[5] = uy[n] * j[2];
    }
    node_rebuild(a + 15 * n, r - rho0, j, s);
  }
}

/* The 10 stored moments per node of lbsave (src/lbstore.c:36-91):
 * drho, jx, jy, jz, sxx, syy, szz, sxy, sxz, syz. */
void orc_moments(long nnodes, const double *a, double *m10) {
  for (long n = 0; n < nnodes; n++) {
    double d, j[3], pi[6];
    node_moments(a + 15 * n, &d, j, pi);
    double *o = m10 + 10 * n;
    o[0] = d; o[1] = j[0]; o[2] = j[1]; o[3] = j[2];
    o[4] = pi[0] - d / 3.0; o[5] = pi[1] - d / 3.0; o[6] = pi[2] - d / 3.0;
    o[7] = pi[3]; o[8] = pi[4]; o[9] = pi[5];
  }
}

/* ------------------------------------------------------------------ */
/* Frozen density for moment propagation: src/mpmisc.c:64-79.
 * Fluid: rho0 + sum a; solid: rho0. */
void orc_mprhoinit(long nnodes, const double *a, const int32_t *objid,
                   double rho0, double *rho) {
  for (long n = 0; n < nnodes; n++) {
    double r = rho0;
    if (objid[n] < 0) for (int l = 0; l < 15; l++) r += a[15 * n + l];
    rho[n] = r;
  }
}

/* Moment-propagation step, scatter form: src/mpkernel.c:47-201 (mpstat)
 * and :206-368 (mpmove).  A walker on fluid node y moves along link l with
 *   pi_l(y) = a_l(y)/rho(y) + (w_l/72)(rho0/rho(y) - alpha)  (+ alpha, l = 0).
 * A non-adsorbing solid node s returns, along each of its boundary links l
 * (f = s + c_l fluid), what f sent towards it, pi_lbar(f) P(f), plus the
 * moving-wall term w_l (P(f)/rho(f)) c_l.U(s), U = rho0 (v + w x r)/12 from
 * the object state.  Final update (:190-197): fluid p = new; adsorbing solid
 * (mask bit 0x4000) p += new; other solid p = 0.
 *
 * Extensions (this project's definitions, no reference counterpart):
 *   pads   partial adsorption probability on boundary solid nodes without
 *          bit 0x4000: the reflected amount is scaled by (1 - pads) and the
 *          solid node keeps p += pads * new.  pads = 0 is the reference's
 *          reflecting rule, pads = 1 its all-adsorbing rule.
 *   nmom=3 also carries displacement moments m1 = sum P d, m2 = sum P d^2
 *          along `axis`: a hop along c_l adds d = c_l[axis]; a reflection
 *          adds 0.  Adsorbed walkers keep their moments on the solid node. */
void orc_mpstep(int esi, int esj, int esk, const double *a, const int32_t *mask,
                const int32_t *objid, const double *rho, double rho0,
                double alpha, int moving, const double *objpos,
                const double *objvel, const double *objang, double pads,
                int nmom, double *p, double *m1, double *m2, int axis) {
  long N = (long)esi * esj * esk;
  double *np = (double *)calloc(N, sizeof(double));
  double *n1 = nmom > 1 ? (double *)calloc(N, sizeof(double)) : NULL;
  double *n2 = nmom > 1 ? (double *)calloc(N, sizeof(double)) : NULL;
  for (int i = 0; i < esi; i++)
    for (int j = 0; j < esj; j++)
      for (int k = 0; k < esk; k++) {
        long y = (long)esk * ((long)esj * i + j) + k;
        if (objid[y] < 0) {
          double P = p[y], r = rho[y];
          for (int l = 0; l < 15; l++) {
            double pr = a[15 * y + l] / r + orc_w[l] / 72.0 * (rho0 / r - alpha);
            if (l == 0) pr += alpha;
            long x = nb(esi, esj, esk, i, j, k, l, +1);
            np[x] += pr * P;
            if (nmom > 1) {
              double d = orc_c[l][axis];
              n1[x] += pr * (m1[y] + d * P);
              n2[x] += pr * (m2[y] + 2.0 * d * m1[y] + d * d * P);
            }
          }
          continue;
        }
        int32_t m = mask[y];
        if (!m || (m & 0x4000)) continue;
        double keep = 1.0 - pads;
        double U[3] = {0, 0, 0};
        if (moving) {
          int o = objid[y];
          const double *P0 = objpos + 3 * o, *V = objvel + 3 * o, *W = objang + 3 * o;
          double rx = i - P0[0], ry = j - P0[1], rz = k - P0[2];
          U[0] = rho0 * (V[0] + W[1] * rz - W[2] * ry) / 12.0;
          U[1] = rho0 * (V[1] + W[2] * rx - W[0] * rz) / 12.0;
          U[2] = rho0 * (V[2] + W[0] * ry - W[1] * rx) / 12.0;
        }
        for (int l = 1; l < 15; l++) {
          if (!(m & (1 << (l - 1)))) continue;
          long f = nb(esi, esj, esk, i, j, k, l, +1);
          int lb = orc_opp[l];
          double r = rho[f], P = p[f];
          double pr = a[15 * f + lb] / r + orc_w[l] / 72.0 * (rho0 / r - alpha);
          if (moving)
            pr += orc_w[l] * (orc_c[l][0] * U[0] + orc_c[l][1] * U[1] + orc_c[l][2] * U[2]) / r;
          pr *= keep;
          np[f] += pr * P;
          if (nmom > 1) { n1[f] += pr * m1[f]; n2[f] += pr * m2[f]; }
        }
      }
  for (long n = 0; n < N; n++) {
    int32_t m = mask[n];
    double frac;
    if (objid[n] < 0) frac = -1.0;                 /* replace */
    else if (m & 0x4000) frac = 1.0;               /* adsorbing site */
    else if (m != 0 && pads > 0.0) frac = pads;    /* partial adsorption */
    else frac = 0.0;                               /* cleared */
    if (frac < 0.0) {
      p[n] = np[n];
      if (nmom > 1) { m1[n] = n1[n]; m2[n] = n2[n]; }
    } else if (frac > 0.0) {
      p[n] += frac * np[n];
      if (nmom > 1) { m1[n] += frac * n1[n]; m2[n] += frac * n2[n]; }
    } else {
      p[n] = 0.0;
      if (nmom > 1) { m1[n] = 0.0; m2[n] = 0.0; }
    }
  }
  free(np); free(n1); free(n2);
}
