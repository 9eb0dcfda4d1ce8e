/*
 * sunlightlb.h -- public C API of the MI355X (gfx950) SunlightLB library.
 *
 * Drop-in for the reference interface src/sunlightlb.h: every macro, struct
 * layout (same field order and types) and prototype of src/sunlightlb.h:21-316
 * is declared here, so a driver written for the reference (for example
 * examples/sphere.c) compiles and links against this library unchanged.
 *
 * The one addition to the reference layout: LBSys_type gains the fields
 * dpdx, dpdy, dpdz at its END.  They are a live uniform body force applied to
 * fluid nodes on top of the per-node bx/by/bz field (default 0), read at every
 * step; examples/sphere.c:50-52 and the reference's Perl/Python examples set
 * them.  Because they are appended, the offsets of all reference fields are
 * unchanged.
 *
 * Storage model.  The lattice state lives on the GPU.  The host arrays named
 * in the structs (lb->a, lb->mask, lb->ux.., lb->fx.., mp->p, lb->rho) are
 * mirrors.  Every function below that reads or writes one of them on the host
 * first brings it up to date, and every GPU step first uploads any host-side
 * change made through this API.  A driver that writes a mirror directly
 * through the pointer calls slb_sync_device() (sunlightlb_gpu.h) afterwards;
 * a driver that reads one directly calls slb_sync_host() first.
 *
 * GPU-specific extensions (precision, run-to-residual, batched solves,
 * packs) are declared in sunlightlb_gpu.h.
 */
#ifndef SUNLIGHTLB_H_INC
#define SUNLIGHTLB_H_INC

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Host-side real type of the API (src/sunlightlb.h:25-29).  The GPU
 * arithmetic precision is chosen separately at run time
 * (slb_set_precision in sunlightlb_gpu.h). */
#ifdef _SINGLE
typedef float LBReal;
#else
typedef double LBReal;
#endif

/* Real type of the binary checkpoint files (src/sunlightlb.h:31-35). */
#ifdef _STORE_DOUBLE
typedef double LBStoreReal;
#else
typedef float LBStoreReal;
#endif

/* Verbosity bits (src/sunlightlb.h:39-43). */
#define TRACE_ROUTINE_EXITS   0x02
#define TRACING_ROUTINE_EXITS 0x02
#define OBJECT_DETAILS        0x04
#define MP_OBJECT_DETAILS     0x04
#define FORCE_DETAILS         0x08

/* Position in the bbl -> propagate -> collide cycle (src/sunlightlb.h:47-49). */
#define POST_BBL       42
#define POST_PROPAGATE 47
#define POST_COLLIDE   49

/* A rigid object embedded in the lattice (src/sunlightlb.h:53-62). */
typedef struct LBObj_struct {
  int id;
  LBReal posx, posy, posz;
  LBReal velx, vely, velz;
  LBReal angx, angy, angz;
  LBReal frcx, frcy, frcz;
  LBReal trqx, trqy, trqz;
  struct LBObj_struct *next_obj;
  struct LBObj_struct *base_obj;
} LBObj_type;

typedef LBObj_type *LBObj;

/* The lattice system (src/sunlightlb.h:68-100), plus dpdx/dpdy/dpdz. */
typedef struct LBSys_struct {
  int *w, *c;                          /* D3Q15 weights w[15], velocities c[45] */
  LBReal *a;                           /* 15 shifted populations per node (mirror) */
  LBReal *ux, *uy, *uz;                /* node velocity: walls (solid) / vgrab (fluid) */
  LBReal *fx, *fy, *fz;                /* per-node force from bblforce */
  LBReal *bx, *by, *bz;                /* per-node body force */
  int esi, esj, esk, nnodes, step, status, changed_objects;
  int objects_moving, skip_solid, stokes_flow, vel_corrected, monitoring;
  LBReal tau, rho0;
  LBReal vmeanx, vmeany, vmeanz;       /* written by measure()/accvmean() */
  LBReal *psi;                         /* optional averaging weight */
  LBReal *rho;                         /* frozen density for moment propagation */
  char *name;
  LBReal rescale;
  int load_average;
  LBObj *obj;                          /* per-node object, NULL = fluid */
  LBObj obj_head;
  int *mask;                           /* boundary-link bits 0..13, adsorb bit 14 */
  int next_id;
  LBReal dpdx, dpdy, dpdz;             /* appended: live uniform body force */
} LBSys_type;

typedef LBSys_type *LBSys;

/* Moment propagation (passive scalar) system (src/sunlightlb.h:106-119). */
typedef struct MPSys_struct {
  LBSys lb;
  LBReal *p, *newp;
  int step;
  int objects_moving;
  int to_be_initialised;
  LBReal alpha, isoslice, poffset;
  LBReal p2bytemin, p2bytemax;
  char *name;
} MPSys_type;

typedef MPSys_type *MPSys;

/* Isosurface counter (src/sunlightlb.h:123-129). */
typedef struct MPIso_struct {
  MPSys mp;
  LBReal val;
  int *count;
} MPIso_type;

typedef MPIso_type *MPIso;

/* Colours, colour maps and images (src/sunlightlb.h:133-161). */
typedef struct LBColour_struct {
  int red, green, blue;
} LBColour_type;

typedef LBColour_type *LBColour;

typedef struct LBColourMap_struct {
  int red_lo, green_lo, blue_lo;
  int red_hi, green_hi, blue_hi;
  LBReal v_lo, v_hi;
} LBColourMap_type;

typedef LBColourMap_type *LBColourMap;

typedef struct LBImage_struct {
  int width, height;
  int *red, *green, *blue;
  int back_red, back_green, back_blue;
  int object_red, object_green, object_blue;
} LBImage_type;

typedef LBImage_type *LBImage;

/* Global verbosity word (src/sunlightlb.h:165-171); defined in the library. */
extern int slb_verbosity;

/* ---- basic lattice Boltzmann (src/sunlightlb.h:176-238) ---- */
LBSys newlb(int, int, int, char *);
void set_verbosity(int);
void unset_verbosity(int);
void report_verbosity();
void lbstep(LBSys);
void bbl(LBSys);
void bblforce(LBSys);
void collide(LBSys);
void propagate(LBSys);
LBObj lookupobj(LBSys, int);
LBObj addobj(LBSys);
void zeroobj(LBObj);
LBObj insertsubs(LBSys, char *, int, int, int, int, int, int);
LBObj insertlist(LBSys, char *);
LBObj insertslab(LBSys, LBReal, LBReal, LBReal, LBReal, LBReal, LBReal);
LBObj insertsph(LBSys, LBReal, LBReal, LBReal, LBReal);
LBObj insertcyl(LBSys, char *, LBReal, LBReal, LBReal);
LBObj insertwall(LBSys, char *, int);
LBObj flattenobj(LBSys);
LBObj invertobj(LBSys);
void removeobjs(LBSys);
void wrobjs(LBSys);
void wrmap(LBSys, char *);
void makemask(LBSys);
void vspread(LBSys);
void zerobf(LBSys);
void addbf(LBSys, double, double, double);
void vset(LBSys, int, int, int, LBReal, LBReal, LBReal);
int countmask(LBSys, char *);
int countfluid(LBSys);
void syncobjs(LBSys);
void lbsave(LBSys, char *);
void lbload(LBSys, char *);
void lbsavevel(LBSys, char *, char *);
void lbloadvel(LBSys, char *, char *);
void lbsavestr(LBSys, char *, char *);
void lbsaveobj(LBSys, char *);
void lbloadobj(LBSys, char *);
void lbsavefrc(LBSys, char *, char *);
void lbsavemap(LBSys, char *, char *);
void error(char *);
void about();
void wrpars(LBSys);
void wrstatus(LBSys, FILE *);
void wrlinks(LBSys);
void wrstats(LBSys);
void fwrstats(FILE *, int, int);
void startclock();
char *mycat(char *, char *);
char *mycopy(char *);
int myscmp(char *, char *);
void accvmean(LBSys, LBReal);
void vzero(LBSys);
void vgrab(LBSys, LBReal);
void vsave(LBSys, char *);
void measure(LBSys);
void zeroforce(LBSys);
void normforce(LBSys, LBReal);
void distribf(LBSys);
void wrmeas(LBSys);
LBReal discrf(LBSys);
void wrmon(LBSys, char *);

/* ---- moment propagation (src/sunlightlb.h:242-277) ---- */
void mpstep(MPSys);
MPSys newmp(LBSys, char *);
MPIso newmpiso(MPSys, LBReal);
void mpiso(MPIso);
void mpisowr(MPIso, char *, char *);
void mpinit(MPSys);
void mpwrpars(MPSys);
void mprhoinit(MPSys);
void mpcheck(MPSys);
LBReal mpmean(MPSys);
LBReal mpmax(MPSys);
LBReal mpmin(MPSys);
void mprescale(MPSys, LBReal);
void mpcopy(MPSys, MPSys);
void mpsave(MPSys, char *, char *);
void mpload(MPSys, char *, char *);
void mpconwr(MPSys, char *, char *);
LBReal mpspread(MPSys, char *);
void mpall(MPSys, LBReal);
void mpslab(MPSys, LBReal, LBReal, LBReal, LBReal, LBReal, LBReal, LBReal);
void mpwall(MPSys, char *, int, LBReal);
void mpsph(MPSys, LBReal, LBReal, LBReal, LBReal, LBReal);
void mpcyl(MPSys, char *, LBReal, LBReal, LBReal, LBReal);
void mpsubs(MPSys, char *, int, int, int, int, int, int, LBReal);
void mplist(MPSys, char *, LBReal);
void mpvacfinit(MPSys, MPSys, char *, int);
void mpvacfinit2(MPSys, MPSys);
void mpvacf(MPSys, MPSys);
LBReal mpvacfwr(MPSys, MPSys, char *, int);
void adall(MPSys);
void adnone(MPSys);
void adslab(MPSys, LBReal, LBReal, LBReal, LBReal, LBReal, LBReal);
void adsph(MPSys, LBReal, LBReal, LBReal, LBReal);
void adcyl(MPSys, char *, LBReal, LBReal, LBReal);
LBReal adtotal(MPSys, char *, int);
void adzero(MPSys);

/* ---- images (src/sunlightlb.h:281-294) ---- */
LBColour newlbcolour(int, int, int);
void lbcolourset(LBColour, int, int, int);
void lbcolourcopy(LBColour, LBColour);
LBColourMap newlbcolourmap(LBReal, LBReal, LBColour, LBColour);
void lbcolourmapset(LBColourMap, LBReal, LBReal, LBColour, LBColour);
LBImage newlbimage(int, int, LBColour, LBColour);
void lbimageclear(LBImage);
void lbimagemap(LBImage, LBSys, char *, int);
void lbimagevel(LBImage, LBColourMap, LBSys, char *, char *, int);
void lbcolourget(LBImage, LBColour, int, int);
int lbgetred(LBImage, int, int);
int lbgetgreen(LBImage, int, int);
int lbgetblue(LBImage, int, int);
void lbcolourclamp(int *, int *, int *);
void mpimageset(LBImage, MPSys, LBColourMap, char *, char *, int);

/* ---- signals (src/sunlightlb.h:298-314) ---- */
#ifndef _NO_SIGNALLING
#define SIGNAL
extern int slb_signal_flag;
void sighandle(LBSys);
void sighook(int);
void siginit();
#endif

#ifdef __cplusplus
}
#endif

#endif /* SUNLIGHTLB_H_INC */
