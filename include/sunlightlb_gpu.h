/*
 * sunlightlb_gpu.h -- extensions of the MI355X SunlightLB library that have
 * no counterpart in the reference interface src/sunlightlb.h.  The reference
 * header (sunlightlb.h here) stays a strict superset-compatible drop-in; a
 * driver that never includes this file sees only reference behaviour.
 *
 * All entry points take the reference handles (LBSys, MPSys) and plain
 * pointers and sizes.
 */
#ifndef SUNLIGHTLB_GPU_H_INC
#define SUNLIGHTLB_GPU_H_INC

#include "sunlightlb.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- precision and host/device coherence ---- */

/* GPU arithmetic and storage precision of the populations: 64 (default,
 * matches the reference's double LBReal) or 32 (production: shifted
 * populations a = f - w rho0/72 held in float, reductions, forces and
 * torques in double).  Switching converts the live state.  Returns 0 on
 * success, -1 for an unsupported value.  The environment variable
 * SLB_PRECISION=32|64 sets the default for newlb(). */
int  slb_set_precision(LBSys lb, int bits);
int  slb_get_precision(LBSys lb);

/* Bring every host mirror (a, mask, ux.., fx.., rho, and the P field of
 * every MP system) up to date with the device. */
void slb_sync_host(LBSys lb);
/* Declare that the driver wrote host mirrors directly through the struct
 * pointers (a, ux.., bx.., mask, obj): they are uploaded before the next
 * GPU operation. */
void slb_sync_device(LBSys lb);
void slb_mp_sync_host(MPSys mp);
void slb_mp_sync_device(MPSys mp);

/* Release a lattice created by newlb(): its objects, its MP systems and all
 * device memory (the reference API has no destructor).  Every handle
 * derived from lb is invalid afterwards. */
void slb_destroy(LBSys lb);

/* ---- stepping ---- */

/* nsteps lbstep() calls without host round trips (the step counter, status
 * and signal handling advance as for lbstep). */
void slb_run(LBSys lb, int nsteps);

/* Relative velocity residual ||u(t) - u(t_prev)||_2 / ||u(t)||_2 over fluid
 * nodes, with u = (j - b/2)/rho the corrected post-collide velocity and
 * u(t_prev) the field stored by the previous call (FP64 sums).  The first
 * call after slb_residual_reset() stores the field and returns -1. */
double slb_residual(LBSys lb);
void   slb_residual_reset(LBSys lb);

/* Run in chunks of `interval` steps until slb_residual() < tol or maxsteps
 * steps were taken.  Returns the steps taken; *res receives the last
 * residual. */
int slb_run_to_residual(LBSys lb, double tol, int interval, int maxsteps, double *res);

/* Time nsteps fused steps on the device between HIP events (the state
 * advances).  Returns milliseconds. */
double slb_time_steps(LBSys lb, int nsteps);

/* Multiply every population by s (a Stokes-flow solution scales with the
 * body force). */
void slb_scale_pops(LBSys lb, double s);

/* Sum over fluid nodes of each population a_l, l = 0..14 (FP64). */
void slb_sum_pops(LBSys lb, double out[15]);

/* ---- geometry ---- */

/* Insert n spheres, rows (radius, x, y, z), each exactly as insertsph()
 * would, in order.  Returns the number of spheres that set at least one
 * node. */
int slb_insert_spheres(LBSys lb, int n, const double *rxyz);

/* Random sequential addition of non-overlapping spheres of radius R into the
 * periodic box (minimum-image overlap test, splitmix64 stream from `seed`),
 * until the lattice solid fraction (solid nodes / nnodes, counting the 27
 * periodic images that the insertion writes) reaches phi, or maxtrials
 * trials were made.  Each accepted sphere is inserted with insertsph() at
 * every periodic image whose bounding box meets the box.  Writes up to
 * maxcentres centres (x,y,z rows) to centres if non-NULL.  Returns the
 * number of spheres placed; *phi_out receives the achieved lattice solid
 * fraction. */
int slb_rsa_pack(LBSys lb, double R, double phi, unsigned long long seed,
                 long maxtrials, double *centres, int maxcentres, double *phi_out);

/* ---- moment propagation extensions ---- */

/* Partial adsorption: each boundary solid node without the 0x4000 bit keeps
 * a fraction pads of what arrives from the fluid and reflects (1 - pads).
 * pads = 0 is the reference rule (reflect); pads = 1 equals adall(). */
void slb_mp_set_adsorption(MPSys mp, double pads);

/* Also propagate displacement moments along axis (0 x, 1 y, 2 z): m1 = sum
 * P d, m2 = sum P d^2 with d the accumulated lattice displacement of the
 * walkers (unwrapped).  Starts the moments at zero. */
void slb_mp_track_moments(MPSys mp, int axis);

/* Moment sums: out[0] = sum fluid P, [1] sum solid P, [2] sum over
 * adsorbing sites, [3] sum fluid m1, [4] sum fluid m2, [5] fluid nodes. */
void slb_mp_sums(MPSys mp, double out[6]);

/* Run nsteps mpstep() calls without host round trips. */
void slb_mp_run(MPSys mp, int nsteps);

/* Time nsteps MP steps on the device between HIP events (the state
 * advances).  Returns milliseconds. */
double slb_mp_time_steps(MPSys mp, int nsteps);

/* ---- introspection ---- */

/* Plain device copy of `bytes` (16 bytes per lane) repeated reps times:
 * mean milliseconds per copy.  The achievable HBM copy rate of the device,
 * and the known-size kernel against which memory counters are scaled. */
double slb_hbm_copy_test(size_t bytes, int reps);

/* Device memory held by the lattice, bytes. */
size_t slb_device_bytes(LBSys lb);
/* GPU architecture name, e.g. "gfx950". */
const char *slb_device_arch(void);

#ifdef __cplusplus
}
#endif
#endif
