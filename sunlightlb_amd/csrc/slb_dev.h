/*
 * slb_dev.h -- internal C ABI between the C host layer (slb_*.c) and the
 * gfx950 HIP kernels (slb_dev.hip).  Plain pointers and sizes only.
 *
 * Device data layout for one lattice of N = esi*esj*esk nodes
 * (node index n = esk*(esj*i+j)+k, k fastest):
 *   pops[2]   structure-of-arrays populations, pops[b][l*N + n], T = float or
 *             double (the run-time precision); ping-pong pair for the pull
 *             step, `cur` names the live one (post-collide between steps)
 *   flags     uint16 per node: bit l-1 (l=1..14) = the population arriving
 *             along c_l is bounced back (link with exactly one solid end),
 *             bit 14 = node is solid
 *   objidx    int32 per node, -1 fluid, else dense object index
 *   mask      int32 per node, the reference boundary-link mask (bits 0..13)
 *             plus the adsorption bit 0x4000
 *   ux,uy,uz  double per node: wall velocity (rho0/12 scaled) on solid nodes,
 *             vgrab accumulator on fluid nodes
 *   fx,fy,fz  double per node, momentum-exchange force (allocated lazily)
 *   bx,by,bz  T per node, body force (only when it is not uniform)
 */
#ifndef SLB_DEV_H
#define SLB_DEV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct slbd_lattice slbd_lattice;

/* Per-call physics parameters, read from the host struct at every call. */
typedef struct {
  double tau, rho0;
  int stokes_flow, skip_solid, moving;
  int bf_mode;          /* 0: no body force, 1: uniform bf[], 2: per-node arrays + bf[] */
  double bf[3];         /* uniform part of the body force on fluid nodes */
} slbd_params;

/* Object table row: pos[3], vel[3], ang[3]. */
#define SLBD_OBJ_STRIDE 9

int  slbd_device_count(void);
slbd_lattice *slbd_create(int esi, int esj, int esk, int prec);
void slbd_destroy(slbd_lattice *L);
int  slbd_precision(const slbd_lattice *L);
int  slbd_set_precision(slbd_lattice *L, int prec, double *scratch_aos);
void slbd_sync(slbd_lattice *L);
const char *slbd_arch(void);

/* host <-> device transfers (host populations are AoS double a[15n+l]) */
void slbd_put_pops(slbd_lattice *L, const double *aos);
void slbd_get_pops(slbd_lattice *L, double *aos);
void slbd_zero_pops(slbd_lattice *L);
void slbd_put_objidx(slbd_lattice *L, const int32_t *objidx);
void slbd_get_objidx(slbd_lattice *L, int32_t *objidx);
void slbd_makemask(slbd_lattice *L);                 /* mask from objidx (bit-exact) */
long slbd_countmask(slbd_lattice *L, int faces, const int *w); /* countmask on the device mask */
void slbd_put_mask(slbd_lattice *L, const int32_t *mask);
void slbd_get_mask(slbd_lattice *L, int32_t *mask);
void slbd_build_flags(slbd_lattice *L);              /* pull flags from mask + objidx */
void slbd_put_u(slbd_lattice *L, const double *ux, const double *uy, const double *uz);
void slbd_get_u(slbd_lattice *L, double *ux, double *uy, double *uz);
void slbd_put_bf(slbd_lattice *L, const double *bx, const double *by, const double *bz);
void slbd_get_f(slbd_lattice *L, double *fx, double *fy, double *fz);
void slbd_put_f(slbd_lattice *L, const double *fx, const double *fy, const double *fz);
void slbd_zero_f(slbd_lattice *L);
void slbd_scale_f(slbd_lattice *L, double s);
void slbd_scale_pops(slbd_lattice *L, double s);

/* LB step: nsteps fused bbl+propagate+collide; force != 0 accumulates the
 * bblforce node force (fx,fy,fz) during the bounce-back. */
void slbd_step(slbd_lattice *L, const slbd_params *P, int nsteps, int force);
/* Random sequential addition, candidate pre-filter on the device (see
 * k_rsa_filter in slb_dev.hip): a cell grid of the accepted spheres (host
 * layout: heads per cell, next per sphere, centres), and for trials t0 ..
 * t0+m-1 of the splitmix64 stream from seed, pass[q] = 0 when the trial's
 * centre clearly overlaps an uploaded sphere (squared minimum-image
 * distance below d2min). */
typedef struct slbd_rsa slbd_rsa;
slbd_rsa *slbd_rsa_create(const double L[3], const double cs[3], const int nc[3]);
void slbd_rsa_update(slbd_rsa *R, const int *head, const int *next, const double *xyz, int n);
void slbd_rsa_filter(slbd_rsa *R, uint64_t seed, long t0, int m, double d2min, uint8_t *pass);
void slbd_rsa_destroy(slbd_rsa *R);
/* Sphere images (n rows of centre {ox, oy, oz} and clipped bounding box
 * {ilo, jlo, klo, ihi, jhi, khi}) against insertsph's FP64 predicate with
 * r^2 = radsq: node count of each image, or the byte map of all of them
 * (N bytes, copied to the host). */
void slbd_sph_count(slbd_lattice *L, const double *o, const int *box, int n, double radsq,
                    int *counts);
void slbd_sph_map(slbd_lattice *L, const double *o, const int *box, int n, double radsq,
                  uint8_t *map);

/* Launch plan of the two-step sweep over esi planes in chunks of P (see
 * sweep2 in slb_dev.hip): rows of {phase, i0, np, sring, dring}, phase 1 =
 * step t from the live buffer into the plane window of the other buffer,
 * phase 2 = step t+1 from the window back into the live buffer.  Returns
 * the launch count (rows written up to maxrows), 0 when the sweep does not
 * apply.  Host-only. */
int slbd_sweep_plan(int esi, int P, int *rows, int maxrows);
/* Separate phases on the live buffer. */
void slbd_bbl(slbd_lattice *L, const slbd_params *P, int force);
void slbd_propagate(slbd_lattice *L);
void slbd_collide(slbd_lattice *L, const slbd_params *P);

/* Reductions (FP64). out[0..2] = sum over fluid of w (j + fac b)/rho,
 * out[3] = sum w, out[4] = fluid count.  psi may be NULL (weights 1). */
void slbd_vmean(slbd_lattice *L, const slbd_params *P, double fac,
                const double *psi_host, double out[5]);
/* ux += prefac (j + fac b)/rho on fluid nodes (vgrab); vzero zeroes fluid. */
void slbd_vgrab(slbd_lattice *L, const slbd_params *P, double fac, double prefac);
void slbd_vzero(slbd_lattice *L);
void slbd_vspread(slbd_lattice *L, int nobj, const double *objtab, double rho0);
void slbd_distribf(slbd_lattice *L, int nobj, const double *objtab,
                   double *frc, double *trq);
/* Relative velocity residual ||u - u_prev|| / ||u|| over fluid nodes, using
 * the corrected post-collide velocity; stores u as the next u_prev.  Returns
 * -1 on the first call (no previous field). */
double slbd_residual(slbd_lattice *L, const slbd_params *P);
void slbd_residual_reset(slbd_lattice *L);
/* Sum over fluid nodes of the 10 moments (checks / conservation tests). */
void slbd_sum_pops(slbd_lattice *L, double out[15]);

/* Moment propagation.  The frozen density and the MP link flags belong to
 * the lattice; each MP system owns its fields (P, and with nmom = 3 the
 * displacement moments m1, m2), ping-pong pairs in the lattice precision. */
typedef struct slbd_mp slbd_mp;
void slbd_mp_rhoinit(slbd_lattice *L, double rho0);
void slbd_mp_put_rho(slbd_lattice *L, const double *rho);
void slbd_mp_get_rho(slbd_lattice *L, double *rho);
slbd_mp *slbd_mp_create(slbd_lattice *L, int nmom);
void slbd_mp_destroy(slbd_mp *M);
void slbd_mp_put(slbd_lattice *L, slbd_mp *M, int field, const double *v);
void slbd_mp_get(slbd_lattice *L, slbd_mp *M, int field, double *v);
void slbd_mp_zero(slbd_lattice *L, slbd_mp *M, int field);
void slbd_mp_step(slbd_lattice *L, slbd_mp *M, double rho0, double alpha, int moving,
                  int nobj, const double *objtab, double pads, int nmom,
                  int axis, int nsteps);
/* out: [0] sum fluid P, [1] sum solid P, [2] sum adsorbing-site P,
 *      [3] sum fluid m1, [4] sum fluid m2, [5] fluid count */
void slbd_mp_sums(slbd_lattice *L, slbd_mp *M, double out[6]);
/* mpvacf on the device: P -= psi * (sum psi P / sum psi^2) on fluid nodes of
 * M, then out[0] = sum psi P u_comp (u = j/rho, frozen MP rho), out[1] =
 * sum psi P2 (P2 of M2), fluid nodes, FP64; psi_host may be NULL */
void slbd_vacf(slbd_lattice *L, slbd_mp *M, slbd_mp *M2, const double *psi_host, int comp,
               double out[2]);

/* Timing support: run nsteps fused steps between HIP events on the
 * lattice stream; returns the elapsed milliseconds. */
double slbd_time_steps(slbd_lattice *L, const slbd_params *P, int nsteps);
const char *slbd_step_kernel(const slbd_lattice *L);
double slbd_time_mp(slbd_lattice *L, slbd_mp *M, double rho0, double alpha,
                    double pads, int nmom, int nsteps);
size_t slbd_bytes(const slbd_lattice *L);
/* Plain 16-byte-per-lane copy of `bytes` between two device buffers, reps
 * times; mean milliseconds per copy (HBM roofline and counter calibration). */
double slbd_copy_test(size_t bytes, int reps);

/* ---- warm start (features, targets, injection) ---- */
long slbd_nodes(const slbd_lattice *L);
void slbd_dims(const slbd_lattice *L, int *d);
/* channel 0 solid flag, channel 1 signed distance / ell (periodic jump
 * flooding) with ell the mean distance from a fluid node to the nearest
 * solid node (returned); C-channel BF16 device tensor */
double slbd_features(slbd_lattice *L, uint16_t *X, int C);
/* The same features with the seed buffers in the library's static device
 * scratch (no device allocation): lattices of at most
 * SLBD_FEATURES_MAX_NODES nodes; returns -1 (and prints why) beyond that.
 * slbd_scratch_init resolves the scratch's device address on the current
 * device (called at lattice creation). */
#define SLBD_FEATURES_MAX_NODES (1L << 21)
/* Project the velocity channels 0..2 of the FP32 device tensor U [n][CU]
 * towards the discretely mass-conserving fields of the fluid nodes: passes
 * times (face divergence, iters conjugate-gradient iterations on the
 * fluid-graph Laplacian, node correction); returns the last pass's final
 * relative CG residual. */
double slbd_project(slbd_lattice *L, float *U, int CU, int iters, int passes);
double slbd_features_noalloc(slbd_lattice *L, uint16_t *X, int C);
void slbd_scratch_init(void);
/* the 10 moments of the live state on fluid nodes (0 on solid), scaled per
 * group: 0..2 (j - b/2)/rho * sc[0], 3 drho * sc[1], 4..9 stress
 * (xx,yy,zz,xy,xz,yz) * sc[2]; BF16 [n][C] */
void slbd_targets(slbd_lattice *L, const slbd_params *P, const float *sc, uint16_t *Y, int C);
/* the same in FP32 */
void slbd_moments(slbd_lattice *L, const slbd_params *P, const float *sc, float *Y, int C);
/* populations from predicted moments (the layout above) times inv_sc:
 * mode 0 equilibrium at rho0 and u before a collide (lbloadvel), mode 1 the
 * post-collide state with the predicted drho, u and stress */
void slbd_inject(slbd_lattice *L, const slbd_params *P, const float *U, int CU,
                 const double *inv_sc, int mode);

/* ---- population vectors for Anderson mixing (whole population arrays in
 * the lattice precision, on the lattice's stream) ---- */
#define SLBD_VEC_MAX 16
void *slbd_vec_cur(slbd_lattice *L);                 /* current populations */
void *slbd_vec_alloc(slbd_lattice *L, int count);    /* count vectors */
void *slbd_vec_at(slbd_lattice *L, void *base, int k);
void  slbd_vec_free(void *p);
void  slbd_vec_copy(slbd_lattice *L, void *dst, const void *src);
void  slbd_vec_sub(slbd_lattice *L, void *z, const void *x, const void *y);   /* z = x - y */
/* out[j] = a . B[j], FP64, j < m <= SLBD_VEC_MAX, fixed summation order */
void  slbd_vec_dots(slbd_lattice *L, const void *a, void *const *B, int m, double *out);
/* out[j] = a . B[j] and out2[j] = a2 . B[j] in one pass over B */
void  slbd_vec_dots2(slbd_lattice *L, const void *a, const void *a2, void *const *B, int m,
                     double *out, double *out2);
/* z1 = x1 - y1 and z2 = x2 - y2 in one pass */
void  slbd_vec_sub2(slbd_lattice *L, void *z1, const void *x1, const void *y1, void *z2,
                    const void *x2, const void *y2);
/* z = g - sum_j gamma[j] B[j] */
void  slbd_vec_mix(slbd_lattice *L, void *z, const void *g, void *const *B,
                   const double *gamma, int m);

/* ---- batched independent lattices (slb_batch.hip) ---- */
typedef struct slbd_batch slbd_batch;
slbd_batch *slbd_batch_create(int prec);
int  slbd_batch_add_box(slbd_batch *B, int esi, int esj, int esk, const int32_t *objidx,
                        const int32_t *mask, int nobj, int nrhs, const double *tabs,
                        double tau, double rho0, int stokes, int skip_solid, int moving,
                        const double *bf);
int  slbd_batch_commit(slbd_batch *B);
void slbd_batch_config(slbd_batch *B, int block, int npt, int grid);
void slbd_batch_get_config(const slbd_batch *B, int *cfg);
void slbd_batch_run(slbd_batch *B, int nsteps, int use_graph);
double slbd_batch_time(slbd_batch *B, int nsteps, int use_graph);
void slbd_batch_measure(slbd_batch *B);
int  slbd_batch_forces(slbd_batch *B, int box, int rhs, double *frc, double *trq);
int  slbd_batch_get_pops(slbd_batch *B, int box, int rhs, double *aos);
long slbd_batch_nodes(const slbd_batch *B);
int  slbd_batch_nprob(const slbd_batch *B);
int  slbd_batch_nbox(const slbd_batch *B);
int  slbd_batch_box_dims(const slbd_batch *B, int box, int *dims);
int  slbd_batch_cu_count(void);
void slbd_batch_reset(slbd_batch *B);
void slbd_batch_destroy(slbd_batch *B);

#ifdef __cplusplus
}
#endif
#endif
