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

/* ---- batched independent boxes (resistance problems) ---- */

/* A batch of independent periodic boxes, each with nrhs right-hand sides
 * (object velocity sets), stepped together: one kernel launch per step over
 * all problems, the step loop replayed from a captured HIP graph. */
typedef struct slb_batch_s *SLBBatch;

/* New empty batch; prec 32 or 64 (populations; force sums are FP64). */
SLBBatch slb_batch_new(int prec);

/* Add lb's box: its solid map and boundary-link mask (lb is synchronised),
 * tau, stokes_flow, skip_solid and uniform body force, as nrhs problems
 * starting from rest.  objvel holds nrhs x nobj rows of (vx, vy, vz, wx, wy,
 * wz) for lb's objects in list order (NULL: every rhs uses the objects' own
 * velx..angz); wall velocities follow vspread() (rho0 (v + w x r)/12, r from
 * the object centre).  lb itself is not changed and may be destroyed after
 * the call.  Returns the box index, or -1. */
int  slb_batch_add(SLBBatch b, LBSys lb, int nrhs, const double *objvel);

/* Allocate the device arenas and the tile schedule; no more boxes after. */
int  slb_batch_commit(SLBBatch b);

/* Choose block size, nodes per thread and persistent grid size by timing
 * candidate shapes on this batch (before its first step; the state is reset
 * to rest afterwards).  With cache_path, a shape class already in the file is
 * reused without timing, and new results are appended.  Returns 1 on a cache
 * hit, 0 after timing, -1 on error. */
int  slb_batch_autotune(SLBBatch b, const char *cache_path);
void slb_batch_set_config(SLBBatch b, int block, int nodes_per_thread, int grid);
/* cfg = {block, nodes per thread, grid (0 = one block per tile), tiles} */
void slb_batch_get_config(SLBBatch b, int cfg[4]);

void   slb_batch_run(SLBBatch b, int nsteps);
double slb_batch_time_steps(SLBBatch b, int nsteps);
/* measure(): two force-collecting steps, force and torque per object halved
 * (FP64 sums, torque about each object's own centre, no minimum image). */
void slb_batch_measure(SLBBatch b);
/* Force and torque (3 each per object, list order) of one problem after
 * slb_batch_measure(); returns the object count or -1. */
int  slb_batch_forces(SLBBatch b, int box, int rhs, double *frc, double *trq);
/* Populations of one problem, AoS a[15n+l] as lb->a. */
int  slb_batch_get_pops(SLBBatch b, int box, int rhs, double *aos);
long slb_batch_nodes(SLBBatch b);
void slb_batch_free(SLBBatch b);

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
