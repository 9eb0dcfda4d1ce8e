/*
 * CPU oracle for the SunlightLB D3Q15 lattice-Boltzmann step: a clean-room
 * restatement of the reference algorithm on plain arrays.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or as the timed CPU baseline).  The product library never links
 * or calls it.
 *
 * Conventions (all functions):
 *   - node index  inode = esk*(esj*i + j) + k   (k fastest), i<->x, j<->y, k<->z
 *   - populations AoS a[15*inode + l], shifted: a_l = f_l - w_l*rho0/72
 *   - objid[inode] = -1 for fluid, >= 0 object index for solid
 *   - FP64 throughout
 */
#ifndef SLB_ORACLE_H
#define SLB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* D3Q15 velocity set, integer weights (sum 72) and opposite links. */
extern const int orc_c[15][3];
extern const int orc_w[15];
extern const int orc_opp[15];

/* ---- geometry tables ---- */
int  orc_insertsph(int esi, int esj, int esk, int32_t *objid, int id,
                   double rad, double ox, double oy, double oz, int *overlap);
void orc_makemask(int esi, int esj, int esk, const int32_t *objid, int32_t *mask);
int  orc_countfluid(long nnodes, const int32_t *objid);
int  orc_countmask(long nnodes, const int32_t *mask, int links);

/* ---- LB step phases ---- */
void orc_bbl(int esi, int esj, int esk, double *a, const int32_t *mask,
             const double *ux, const double *uy, const double *uz,
             double *fx, double *fy, double *fz);
void orc_propagate(int esi, int esj, int esk, double *a);
void orc_collide(long nnodes, double *a, const int32_t *objid,
                 const double *bx, const double *by, const double *bz,
                 double tau, double rho0, int stokes_flow, int skip_solid);
void orc_lbstep(int esi, int esj, int esk, double *a, const int32_t *mask,
                const int32_t *objid, const double *bx, const double *by,
                const double *bz, const double *ux, const double *uy,
                const double *uz, double tau, double rho0, int stokes_flow,
                int skip_solid, int moving, long nsteps);

/* ---- measurement ---- */
void orc_vspread(int esi, int esj, int esk, const int32_t *objid,
                 const double *objpos, const double *objvel, const double *objang,
                 double rho0, double *ux, double *uy, double *uz);
void orc_accvmean(long nnodes, const double *a, const int32_t *objid,
                  const double *bx, const double *by, const double *bz,
                  const double *psi, double rho0, double fac, double out[3]);
void orc_distribf(int esi, int esj, int esk, const int32_t *objid, int nobj,
                  const double *objpos, const double *fx, const double *fy,
                  const double *fz, double *frc, double *trq);
void orc_measure(int esi, int esj, int esk, double *a, const int32_t *mask,
                 const int32_t *objid, const double *bx, const double *by,
                 const double *bz, const double *ux, const double *uy,
                 const double *uz, double tau, double rho0, int stokes_flow,
                 int skip_solid, int moving, int vel_corrected, int nobj,
                 const double *objpos, double *fx, double *fy, double *fz,
                 double vmean[3], double *frc, double *trq);
void orc_vfield(long nnodes, const double *a, const int32_t *objid,
                const double *bx, const double *by, const double *bz,
                double rho0, double fac, double *ux, double *uy, double *uz);
void orc_equilibrium(long nnodes, double *a, const double *rho,
                     const double *ux, const double *uy, const double *uz,
                     double rho0, int stokes_flow);
void orc_moments(long nnodes, const double *a, double *m10);

/* ---- moment propagation ---- */
void orc_mprhoinit(long nnodes, const double *a, const int32_t *objid,
                   double rho0, double *rho);
void orc_mpstep(int esi, int esj, int esk, const double *a, const int32_t *mask,
                const int32_t *objid, const double *rho, double rho0,
                double alpha, int moving, const double *objpos,
                const double *objvel, const double *objang, double pads,
                int nmom, double *p, double *m1, double *m2, int axis);

#ifdef __cplusplus
}
#endif
#endif
