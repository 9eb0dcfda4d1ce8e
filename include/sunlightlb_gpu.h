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

#include <stdint.h>
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
/* The same solve with Anderson mixing over the last `window` (1..16)
 * iterations of `interval` steps each: the next state is the combination of
 * recent iterates that minimises the fixed-point residual.  Ends, as above,
 * in a state whose plain `interval`-step residual is below tol.  Needs
 * 2 window + 3 extra population arrays on the device.  window 0: plain. */
int slb_run_to_residual_aa(LBSys lb, double tol, int interval, int maxsteps, int window,
                           double *res);

/* Time nsteps fused steps on the device between HIP events (the state
 * advances).  Returns milliseconds. */
double slb_time_steps(LBSys lb, int nsteps);

/* Name of the kernel that the lattice's fused steps launch, fixed when the
 * lattice is made from its shape and precision: "k_step_lds" (rows of 64,
 * 128, 256 or 512 nodes, k pulls staged in LDS), "k_step_ldsr" (other rows
 * of up to 512 nodes, where staging measured faster) or "k_step".  All three
 * give bit-identical results; SLB_STEP_LDS=0 selects k_step. */
const char *slb_step_kernel(LBSys lb);

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

/* slb_rsa_pack followed by flattenobj() in one call, with the node counts
 * and the rasterisation of the sphere images done on the GPU (the same FP64
 * predicate): the same spheres, centres, solid fraction, node map and single
 * object (id 0 at the centroid of every solid node, objects already present
 * included).  Entries of centres beyond the returned count are
 * unspecified. */
int slb_rsa_pack_flat(LBSys lb, double R, double phi, unsigned long long seed,
                      long maxtrials, double *centres, int maxcentres, double *phi_out);

/* The solid map (solid[n] = 1, n = esk*(esj*i+j)+k) of a periodic sphere
 * pack of radius R to lattice solid fraction phi, without a lattice:
 * overlap = 0 is the random sequential addition of slb_rsa_pack (same
 * seed, same spheres, same map as slb_rsa_pack followed by flattenobj);
 * overlap = 1 places spheres at random without an overlap test (a Boolean
 * model, for fractions RSA cannot reach).  Thread-safe.  Returns the spheres
 * placed (-1 on allocation failure); *phi_out the achieved fraction. */
int slb_pack_map(int esi, int esj, int esk, double R, double phi, unsigned long long seed,
                 int overlap, long maxtrials, unsigned char *solid, double *phi_out);
/* Insert the nodes with solid[n] != 0 as one object (NULL if none). */
LBObj slb_insert_map(LBSys lb, const unsigned char *solid);

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

/* Run nsteps mpstep() calls without host round trips.  A run of 32 steps or
 * more (SLB_MP_WEIGHTS=n sets the minimum, 0 never) computes each node's
 * link factors from the frozen flow once and reuses them: the tracer a
 * partially adsorbing node keeps is then (sum of its link factors) x P
 * instead of the sum of the products, a rounding-level difference. */
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

/* ---- learned warm start (no reference counterpart) ----
 * A 3D convolutional encoder-decoder (4 levels, 16-64 channels, 3x3x3 taps,
 * periodic padding; BF16 weights and activations, FP32 accumulation, all on
 * the GPU) maps the solid flag and signed distance of a lattice to the 10
 * hydrodynamic moments of its steady Stokes state per unit body force F:
 * m[0..2] corrected velocity u, m[3] density deviation drho, m[4..9] stress
 * s = Pi - drho/3 I (xx, yy, zz, xy, xz, yz), zero on solid nodes.  Needs a
 * uniform non-zero body force; lattice edges must be multiples of 8 (>= 16).
 * Initialisations (slb_warm_set_mode):
 *   SLB_WARM_EQUILIBRIUM  equilibrium at rho0 and the predicted velocity,
 *                         then one collide: what lbloadvel builds;
 *   SLB_WARM_MOMENTS      (default) the post-collide state rebuilt from all
 *                         10 predicted moments. */
#define SLB_WARM_EQUILIBRIUM 0
#define SLB_WARM_MOMENTS 1
typedef struct slb_warmnet slb_warmnet;
/* New network with seeded He-initialised weights. */
slb_warmnet *slb_warm_new(unsigned long long seed);
void   slb_warm_free(slb_warmnet *net);
/* Weights file (magic, layer table, fallback threshold, FP32 weights).
 * save returns 0 / -1; load returns NULL on any mismatch. */
int    slb_warm_save(const slb_warmnet *net, const char *path);
slb_warmnet *slb_warm_load(const char *path);
/* Add the lattice's current state (a converged solve) as a training sample
 * held in device memory; returns the sample count.  sample_stats gives the
 * mean square of the scaled velocity, density and stress targets and the
 * sample's normalising length ell. */
int    slb_warm_add_sample(slb_warmnet *net, LBSys lb);
int    slb_warm_nsamples(const slb_warmnet *net);
void   slb_warm_sample_stats(const slb_warmnet *net, int i, double out[4]);
/* The network's inputs and targets for a lattice, written to a device BF16
 * tensor [n][channels] (zeros beyond the used channels).  Features: channel
 * 0 solid flag, 1 signed distance to the interface / ell (periodic jump
 * flooding: + on fluid, - on solid), ell = the mean distance from a fluid
 * node to the solid, returned.  Targets: the 10 moments above scaled by
 * eta/(F ell^2) (velocity), 1/(8 F ell) (density), 1/(F ell) (stress).
 * moments writes the same targets in FP32; inject builds populations from
 * scaled moments m (device FP32 [n][channels >= 10]) in the given mode. */
double slb_warm_features(LBSys lb, void *x_dev, int channels);
/* slb_warm_features without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * HIP graph capture around it, tight serving loops, harnesses that forbid
 * hipMalloc).  The jump-flooding seed buffers live in a static device
 * scratch of the library (24 MiB + 256 B per GPU, held until the process
 * ends), so lattices are limited to 2^21 nodes (128^3): a larger lattice is
 * refused with a message on stderr and -1 is returned, x_dev untouched.
 * Same result as slb_warm_features, bit for bit.  Every call on one GPU
 * shares the scratch: calls that overlap on one GPU (several host threads)
 * race and must be serialised by the caller. */
double slb_warm_features_noalloc(LBSys lb, void *x_dev, int channels);
void   slb_warm_targets(LBSys lb, void *t_dev, int channels, double ell);
void   slb_warm_moments(LBSys lb, float *m_dev, int channels, double ell);
void   slb_warm_inject(LBSys lb, const float *m_dev, int channels, int mode, double ell);
/* iters Adam steps (cosine learning-rate decay from lr to lr/20) over
 * seeded epochs of the samples; loss = mean over the three target groups of
 * the relative mean-square error.  history (iters doubles) may be NULL.
 * Returns the mean loss over the last min(iters, samples) steps. */
double slb_warm_train(slb_warmnet *net, int iters, double lr, unsigned long long seed,
                      double *history);
/* Relative mean-square error of the prediction against the lattice's
 * current state (group errors to grp[3] when non-NULL). */
double slb_warm_eval(slb_warmnet *net, LBSys lb, double *grp);
/* Predicted moments m[10n + c] (lattice units) of the lattice's solid map
 * and body force. */
void   slb_warm_predict(slb_warmnet *net, LBSys lb, float *m);
void   slb_warm_set_mode(slb_warmnet *net, int mode);
/* Divergence projection of the predicted velocity before injection
 * (iters > 0 and passes > 0; iters = 0: off, the default).  The face
 * fluxes (u(x) + u(x+e_a))/2 between fluid neighbours (none through walls,
 * periodic wrap) lose their divergence: A phi = -div F by iters
 * conjugate-gradient iterations, A the Laplacian of the graph of fluid
 * nodes and open faces; then each fluid node's velocity component a loses
 * the mean of phi's differences across its open faces along a.  That node
 * correction removes the face divergence only approximately, so the whole
 * step is repeated `passes` times.  slb_warm_project applies the same to
 * the channels 0..2 of a device FP32 tensor m [n][channels] on lb's solid
 * map and returns the last pass's final relative CG residual. */
void   slb_warm_set_projection(slb_warmnet *net, int iters, int passes);
void   slb_warm_get_projection(const slb_warmnet *net, int *iters, int *passes);
double slb_warm_project(LBSys lb, float *m_dev, int channels, int iters, int passes);
int    slb_warm_get_mode(const slb_warmnet *net);
/* Initialise the populations from the prediction (the mode above). */
void   slb_warm_start(slb_warmnet *net, LBSys lb);
/* Cold-start fallback.  note_cold records a cold solve's residual history
 * (res[k] after (k+1)*interval steps, n entries); the network keeps the
 * median curve over the recorded solves.  A warm solve with the same
 * interval falls back to a cold start (populations reset to rest) at the
 * first check, from the second on, where the steps projected from its last
 * residual decay ratio exceed the steps the median cold curve takes to the
 * tolerance.  set_fallback replaces the curve by a single threshold
 * probe_res checked from step `probe` on (probe 0: never fall back). */
void   slb_warm_set_fallback(slb_warmnet *net, int probe, double probe_res);
void   slb_warm_get_fallback(const slb_warmnet *net, int *probe, double *probe_res);
void   slb_warm_note_cold(slb_warmnet *net, const double *res, int n, int interval);
/* Start policy.  note_warm records the residual history of a calibration
 * warm solve (slb_warm_start, then plain steps; same form as note_cold); the
 * network keeps the median curve.  slb_warm_policy returns 1 when a solve to
 * tol should start warm: always, unless both median curves exist for this
 * interval and the warm one falls below tol no sooner than the cold one.
 * slb_warm_solve starts cold (from rest) when it returns 0.  set_fallback
 * clears both curves. */
void   slb_warm_note_warm(slb_warmnet *net, const double *res, int n, int interval);
int    slb_warm_policy(const slb_warmnet *net, double tol, int interval);
/* Warm start (or a cold one, by the start policy), then run to
 * slb_residual() < tol as slb_run_to_residual does, with the fallback.
 * Returns the total steps; *used_warm is 0 when the solve started cold or
 * fell back to a cold start. */
int    slb_warm_solve(slb_warmnet *net, LBSys lb, double tol, int interval, int maxsteps,
                      double *res, int *used_warm);
/* Parameter access for tests: the layer table, the parameter arena
 * (W[co][Kp] then 64 bias slots per layer), and one forward+backward on
 * device tensors x (dims d, 16 channels) and t (10 channels) with host loss
 * weights wc[10], returning the weighted squared error, the FP32 output
 * (host, [n][16]) and the gradient (host, nparams floats). */
long   slb_warm_nparams(const slb_warmnet *net);
void   slb_warm_layout(int *nlayer, int *cin, int *cout, long *wofs, long *bofs);
void   slb_warm_get_params(const slb_warmnet *net, float *w);
void   slb_warm_set_params(slb_warmnet *net, const float *w);
double slb_warm_fwd_bwd(slb_warmnet *net, const int *d, const uint16_t *x, const uint16_t *t,
                        const float *wc, float *outf, float *grad);

#ifdef __cplusplus
}
#endif
#endif
