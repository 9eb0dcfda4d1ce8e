/*
 * slb_objs.c -- object list and solid-node rasterisation of the sunlightlb.h
 * API (reference src/lbobjs.c:33-516).
 *
 * The node map obj[] (one object pointer per node) is host data, as in the
 * reference; the GPU receives a dense per-node object index when a kernel
 * next needs it.  Rasterisation follows the reference predicates exactly
 * (same FP64 expressions, same last-insert-wins order), so the solid map and
 * the boundary-link mask derived from it are bit-identical.  insertsph only
 * visits the bounding box of the sphere padded by one node: a node further
 * away has |i - ox| >= rad + 1, so its FP64 squared distance cannot fall
 * below rad^2 and the reference predicate is false there too.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "slb_internal.h"

static LBObj newobj(void) {
  LBObj p = (LBObj)malloc(sizeof(LBObj_type));
  if (p == NULL) error("Ran out of memory for object creation\n");
  return p;
}

/* New object at the head of the list, next id (src/lbobjs.c:45-52). */
LBObj addobj(LBSys lb) {
  LBObj p = newobj();
  p->base_obj = p;
  p->id = lb->next_id;
  lb->next_id++;
  p->next_obj = lb->obj_head;
  lb->obj_head = p;
  slb_obj_changed(lb);  /* dense object indices shift with a new head */
  return p;
}

void zeroobj(LBObj p) {
  p->posx = p->posy = p->posz = 0.0;
  p->velx = p->vely = p->velz = 0.0;
  p->angx = p->angy = p->angz = 0.0;
  p->frcx = p->frcy = p->frcz = 0.0;
  p->trqx = p->trqy = p->trqz = 0.0;
}

LBObj lookupobj(LBSys lb, int tid) {
  for (LBObj p = lb->obj_head; p != NULL; p = p->next_obj)
    if (tid == p->id) return p;
  return NULL;
}

/* Drop the just-added head object after an empty insertion. */
static void drop_head(LBSys lb, LBObj p) {
  lb->obj_head = p->next_obj;
  free(p);
  slb_obj_changed(lb);
}

static inline long nidx(LBSys lb, int i, int j, int k) {
  return (long)lb->esk * ((long)lb->esj * i + j) + k;
}

/* Substrate byte file, bit 0 = solid (src/lbobjs.c:77-123). */
LBObj insertsubs(LBSys lb, char *subsfile, int ilo, int ihi, int jlo, int jhi, int klo, int khi) {
  FILE *fp = fopen(subsfile, "r");
  if (fp == NULL) {
    printf("insertsubs: %s could not be opened\n", subsfile);
    return NULL;
  }
  if (slb_verbosity & OBJECT_DETAILS) {
    printf("\nInserting substrate data from %s\n", subsfile);
    printf("Substrate region is (%i-%i, %i-%i, %i-%i) inclusive\n", ilo, ihi, jlo, jhi, klo, khi);
  }
  LBObj p = addobj(lb);
  zeroobj(p);
  int ns = 0, overlap = 0, tot = 0;
  for (int i = ilo; i <= ihi; i++)
    for (int j = jlo; j <= jhi; j++)
      for (int k = klo; k <= khi; k++) {
        unsigned char byte;
        if (fscanf(fp, "%c", &byte) != 1) {
          if (slb_verbosity & OBJECT_DETAILS) printf("Ran out of data in %s\n", subsfile);
          break;
        }
        long n = nidx(lb, i, j, k);
        if (byte & 1) {
          if (lb->obj[n] != NULL) overlap = 1;
          lb->obj[n] = p;
          ns++;
        }
        if (byte & 2) tot++;
      }
  if (slb_verbosity & OBJECT_DETAILS) {
    printf("Inserted substrate data, id no = %i, # nodes = %i\n", p->id, ns);
    if (overlap) printf("Overlaps occurred\n");
  }
  if (ns == 0) {
    if (slb_verbosity & OBJECT_DETAILS) printf("No sites set in object\n");
    drop_head(lb, p);
    return NULL;
  }
  lb->changed_objects = 1;
  slb_obj_changed(lb);
  fclose(fp);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : insertsubs done\n", lb->step);
  return p;
}

/* Node list file "i j k" per line (src/lbobjs.c:128-165). */
LBObj insertlist(LBSys lb, char *listfile) {
  FILE *fp = fopen(listfile, "r");
  if (fp == NULL) {
    printf("insertlist: %s could not be opened\n", listfile);
    return NULL;
  }
  if (slb_verbosity & OBJECT_DETAILS) printf("\nInserting substrate node list data from %s\n", listfile);
  LBObj p = addobj(lb);
  zeroobj(p);
  int ns = 0, overlap = 0, i, j, k;
  while (fscanf(fp, "%i%i%i", &i, &j, &k) == 3) {
    if (i >= 0 && i < lb->esi && j >= 0 && j < lb->esj && k >= 0 && k < lb->esk) {
      long n = nidx(lb, i, j, k);
      if (lb->obj[n] != NULL) overlap = 1;
      lb->obj[n] = p;
      ns++;
    } else if (slb_verbosity & OBJECT_DETAILS) {
      printf("insertlist: ignoring node (%i,%i,%i), it is out of range\n", i, j, k);
    }
  }
  if (slb_verbosity & OBJECT_DETAILS) {
    printf("Inserted node list data, id no = %i, # nodes = %i\n", p->id, ns);
    if (overlap) printf("Overlaps occurred\n");
  }
  if (ns == 0) {
    if (slb_verbosity & OBJECT_DETAILS) printf("No sites set in object\n");
    drop_head(lb, p);
    return NULL;
  }
  lb->changed_objects = 1;
  slb_obj_changed(lb);
  fclose(fp);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : insertlist done\n", lb->step);
  return p;
}

/* One object's full-lattice insertion pass on host threads: each range of
 * nodes tests the shape's predicate, takes the nodes for object p and counts
 * them and the overlaps (one object per pass, so the order of the ranges
 * does not matter). */
typedef struct {
  LBSys lb;
  LBObj p;
  int shape;              /* 0 slab, 1 cylinder */
  double b[6];            /* slab x1 x2 y1 y2 z1 z2; cylinder rad^2, a, b */
  char axis, hole;        /* cylinder axis letter, hole when nonzero */
  long ns[SLB_PAR_MAX];
  int ov[SLB_PAR_MAX];
} shape_pass;

static void shape_range(void *c, long lo, long hi, int tid) {
  shape_pass *J = (shape_pass *)c;
  LBSys lb = J->lb;
  const long esk = lb->esk, plane = (long)lb->esj * esk;
  long ns = 0;
  int ov = 0;
  for (long n = lo; n < hi; n++) {
    const int i = (int)(n / plane), r = (int)(n - (long)i * plane);
    const int j = (int)(r / esk), k = (int)(r - (long)j * esk);
    int within;
    if (J->shape == 0) {
      within = i > J->b[0] && i < J->b[1] && j > J->b[2] && j < J->b[3] && k > J->b[4] &&
               k < J->b[5];
    } else {
      double u, v;
      switch (J->axis) {
        case 'x': case 'i': u = j - J->b[1]; v = k - J->b[2]; break;
        case 'y': case 'j': u = i - J->b[1]; v = k - J->b[2]; break;
        default: u = i - J->b[1]; v = j - J->b[2]; break;
      }
      const double rsq = u * u + v * v;
      within = rsq < J->b[0];
      if (J->hole) within = !within;
    }
    if (within) {
      if (lb->obj[n] != NULL) ov = 1;
      lb->obj[n] = J->p;
      ns++;
    }
  }
  J->ns[tid] = ns;
  J->ov[tid] = ov;
}

static long shape_insert(shape_pass *J, int *overlap) {
  const int nr = slb_par_for(J->lb->nnodes, shape_range, J);
  long ns = 0;
  for (int t = 0; t < nr; t++) { ns += J->ns[t]; if (J->ov[t]) *overlap = 1; }
  return ns;
}

/* Open slab x1 < i < x2, y1 < j < y2, z1 < k < z2 (src/lbobjs.c:170-199). */
LBObj insertslab(LBSys lb, LBReal x1, LBReal x2, LBReal y1, LBReal y2, LBReal z1, LBReal z2) {
  LBObj p = addobj(lb);
  zeroobj(p);
  int overlap = 0;
  shape_pass J;
  J.lb = lb; J.p = p; J.shape = 0;
  J.b[0] = x1; J.b[1] = x2; J.b[2] = y1; J.b[3] = y2; J.b[4] = z1; J.b[5] = z2;
  J.axis = 0; J.hole = 0;
  int ns = (int)shape_insert(&J, &overlap);
  if (slb_verbosity & OBJECT_DETAILS) {
    printf("Inserted slab ");
    printf("%5.2f < x < %5.2f, ", x1, x2);
    printf("%5.2f < y < %5.2f, ", y1, y2);
    printf("%5.2f < z < %5.2f, ", z1, z2);
    printf("   id no = %i, # nodes = %i\n", p->id, ns);
    if (overlap) printf("Overlaps occurred\n");
  }
  if (ns == 0) {
    if (slb_verbosity & OBJECT_DETAILS) printf("No sites set in object\n");
    drop_head(lb, p);
    return NULL;
  }
  lb->changed_objects = 1;
  slb_obj_changed(lb);
  p->posx = 0.5 * (x1 + x2 - 1); p->posy = 0.5 * (y1 + y2 - 1); p->posz = 0.5 * (z1 + z2 - 1);
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : insertslab done\n", lb->step);
  return p;
}

/* Index range [lo, hi] of an axis of length n that can hold nodes within
 * distance r of centre c, padded by one node (empty when lo > hi). */
static void axis_range(double c, double r, int n, int *lo, int *hi) {
  double a = floor(c - r) - 1.0, b = ceil(c + r) + 1.0;
  if (!(a == a) || !(b == b)) { *lo = 0; *hi = -1; return; }
  if (a < 0.0) a = 0.0;
  if (b > (double)(n - 1)) b = (double)(n - 1);
  if (a > b) { *lo = 0; *hi = -1; return; }
  *lo = (int)a; *hi = (int)b;
}

/* Solid-node predicate of insertsph (src/lbobjs.c:212-218): FP64, the three
 * squares summed left to right, (dx^2 + dy^2) + dz^2 (this file is built
 * with -ffp-contract=off, so no product is fused into a sum).  A row (i, j)
 * whose dx^2 + dy^2 already reaches rad^2 holds no node of the sphere:
 * adding dz^2 >= 0 cannot round the sum below it. */
static inline double row_part(int i, int j, double ox, double oy) {
  const double dx = i - ox, dy = j - oy;
  return dx * dx + dy * dy;
}
static inline int in_sphere_row(double part, int k, double oz, double radsq) {
  const double dz = k - oz;
  return part + dz * dz < radsq;
}

/* Rasterise one sphere into obj[] with object p; returns the node count. */
static int raster_sphere(LBSys lb, LBObj p, double rad, double ox, double oy, double oz,
                         int *overlap) {
  int ilo, ihi, jlo, jhi, klo, khi, ns = 0;
  double radsq = rad * rad, ra = fabs(rad);
  axis_range(ox, ra, lb->esi, &ilo, &ihi);
  axis_range(oy, ra, lb->esj, &jlo, &jhi);
  axis_range(oz, ra, lb->esk, &klo, &khi);
  for (int i = ilo; i <= ihi; i++)
    for (int j = jlo; j <= jhi; j++) {
      const double part = row_part(i, j, ox, oy);
      if (!(part < radsq)) continue;
      for (int k = klo; k <= khi; k++)
        if (in_sphere_row(part, k, oz, radsq)) {
          long n = nidx(lb, i, j, k);
          if (lb->obj[n] != NULL) *overlap = 1;
          lb->obj[n] = p;
          ns++;
        }
    }
  return ns;
}

/* Sphere of radius rad at (ox,oy,oz), no periodic wrap (src/lbobjs.c:204-232). */
LBObj insertsph(LBSys lb, LBReal rad, LBReal ox, LBReal oy, LBReal oz) {
  LBObj p = addobj(lb);
  zeroobj(p);
  int overlap = 0;
  int ns = raster_sphere(lb, p, rad, ox, oy, oz, &overlap);
  if (slb_verbosity & OBJECT_DETAILS) {
    printf("Sphere radius %5.2f at %6.2f %6.2f %6.2f ", rad, ox, oy, oz);
    printf("id = %5i, # nodes = %i", p->id, ns);
    if (overlap) printf(" (overlaps)\n"); else printf("\n");
  }
  if (ns == 0) {
    if (slb_verbosity & OBJECT_DETAILS) printf("No sites set in object\n");
    drop_head(lb, p);
    lb->next_id--;
    return NULL;
  }
  lb->changed_objects = 1;
  slb_obj_changed(lb);
  p->posx = ox; p->posy = oy; p->posz = oz;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : insertsph done\n", lb->step);
  return p;
}

/* Cylinder (or, with a second letter in type, a cylindrical hole) along the
 * axis named by type[0] (src/lbobjs.c:241-287). */
LBObj insertcyl(LBSys lb, char *type, LBReal rad, LBReal a, LBReal b) {
  switch (type[0]) {
    case 'x': case 'i': case 'y': case 'j': case 'z': case 'k': break;
    default: error("Unrecognised cylinder direction"); return NULL;
  }
  LBObj p = addobj(lb);
  zeroobj(p);
  int overlap = 0;
  /* FP64 predicate (src/lbobjs.c:255-266): u^2 + v^2 < rad^2, this file
   * built with -ffp-contract=off */
  shape_pass J;
  J.lb = lb; J.p = p; J.shape = 1;
  J.b[0] = rad * rad; J.b[1] = a; J.b[2] = b;
  J.axis = type[0]; J.hole = type[1] != '\0';
  int ns = (int)shape_insert(&J, &overlap);
  if (slb_verbosity & OBJECT_DETAILS) {
    if (type[1] == '\0') printf("Cylinder "); else printf("Cylindrical hole ");
    printf("radius %5.2f, axis %5.2f %5.2f along %c ", rad, a, b, type[0]);
    printf("id = %i, # nodes = %i", p->id, ns);
    if (ns == 0) printf(" (no sites set)");
    if (overlap) printf(" (overlaps)\n"); else printf("\n");
  }
  if (ns == 0) {
    drop_head(lb, p);
    lb->next_id--;
    return NULL;
  }
  lb->changed_objects = 1;
  slb_obj_changed(lb);
  switch (type[0]) {
    case 'x': case 'i': p->posx = 0.5 * (lb->esi - 1); p->posy = a; p->posz = b; break;
    case 'y': case 'j': p->posx = a; p->posy = 0.5 * (lb->esj - 1); p->posz = b; break;
    default: p->posx = a; p->posy = b; p->posz = 0.5 * (lb->esk - 1); break;
  }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : insertcyl done\n", lb->step);
  return p;
}

/* One-node-thick wall at wallpos normal to type[0] (src/lbobjs.c:294-357).
 * As in the reference, a bad wallpos or direction returns NULL and leaves
 * the new empty object in the list. */
LBObj insertwall(LBSys lb, char *type, int wallpos) {
  LBObj p = addobj(lb);
  zeroobj(p);
  int ns = 0, overlap = 0;
  int esi = lb->esi, esj = lb->esj, esk = lb->esk;
  switch (type[0]) {
    case 'x': case 'i':
      if (wallpos < 0 || wallpos > esi - 1) { printf("insertwall: wallpos out of range\n"); return NULL; }
      for (int j = 0; j < esj; j++)
        for (int k = 0; k < esk; k++) {
          long n = nidx(lb, wallpos, j, k);
          if (lb->obj[n] != NULL) overlap = 1;
          lb->obj[n] = p; ns++;
        }
      break;
    case 'y': case 'j':
      if (wallpos < 0 || wallpos > esj - 1) { printf("insertwall: wallpos out of range\n"); return NULL; }
      for (int i = 0; i < esi; i++)
        for (int k = 0; k < esk; k++) {
          long n = nidx(lb, i, wallpos, k);
          if (lb->obj[n] != NULL) overlap = 1;
          lb->obj[n] = p; ns++;
        }
      break;
    case 'z': case 'k':
      if (wallpos < 0 || wallpos > esk - 1) { printf("insertwall: wallpos out of range\n"); return NULL; }
      for (int i = 0; i < esi; i++)
        for (int j = 0; j < esj; j++) {
          long n = nidx(lb, i, j, wallpos);
          if (lb->obj[n] != NULL) overlap = 1;
          lb->obj[n] = p; ns++;
        }
      break;
    default:
      printf("insertwall: unrecognised direction for walls\n");
      return NULL;
  }
  switch (type[0]) {
    case 'x': case 'i': p->posx = wallpos; p->posy = 0.5 * (esj - 1); p->posz = 0.5 * (esk - 1); break;
    case 'y': case 'j': p->posx = 0.5 * (esi - 1); p->posy = wallpos; p->posz = 0.5 * (esk - 1); break;
    default: p->posx = 0.5 * (esi - 1); p->posy = 0.5 * (esj - 1); p->posz = wallpos; break;
  }
  if (ns == 0) {
    if (slb_verbosity & OBJECT_DETAILS) printf("No sites set in object\n");
    drop_head(lb, p);
    return NULL;
  }
  lb->changed_objects = 1;
  slb_obj_changed(lb);
  if (slb_verbosity & OBJECT_DETAILS) {
    printf("Inserted wall at %c = %i, ", type[0], wallpos);
    printf("id no = %i, # nodes = %i\n", p->id, ns);
    if (overlap) printf("Overlaps occurred\n");
  }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : insertwall done\n", lb->step);
  return p;
}

static void free_list(LBSys lb) {
  LBObj p = lb->obj_head;
  while (p != NULL) {
    LBObj q = p;
    p = p->next_obj;
    free(q);
  }
  lb->next_id = 0;
  lb->obj_head = NULL;
}

/* flattenobj / invertobj pass: the nodes that end up in the new object p
 * (the solid ones, or for invert the fluid ones, whose solid neighbours turn
 * fluid), with per-range node counts and index sums.  The reference sums the
 * indices as doubles in node order; the integer sums here, converted once,
 * are the same numbers (every partial sum is an integer below 2^53). */
typedef struct {
  LBSys lb;
  LBObj p;
  int invert;
  const unsigned char *map;   /* extra solid nodes (flatten only), or NULL */
  long ns[SLB_PAR_MAX];
  long long si[SLB_PAR_MAX], sj[SLB_PAR_MAX], sk[SLB_PAR_MAX];
} merge_pass;

static void merge_range(void *c, long lo, long hi, int tid) {
  merge_pass *J = (merge_pass *)c;
  LBSys lb = J->lb;
  const long esk = lb->esk, plane = (long)lb->esj * esk;
  long ns = 0;
  long long si = 0, sj = 0, sk = 0;
  for (long n = lo; n < hi; n++) {
    const int take = J->invert ? (lb->obj[n] == NULL)
                               : (lb->obj[n] != NULL || (J->map && J->map[n]));
    if (J->invert && !take) { lb->obj[n] = NULL; continue; }
    if (!take) continue;
    const long i = n / plane, r = n - i * plane, j = r / esk, k = r - j * esk;
    si += i; sj += j; sk += k; ns++;
    lb->obj[n] = J->p;
  }
  J->ns[tid] = ns; J->si[tid] = si; J->sj[tid] = sj; J->sk[tid] = sk;
}

static long merge_nodes(LBSys lb, LBObj p, int invert, const unsigned char *map, LBReal *ox,
                        LBReal *oy, LBReal *oz) {
  merge_pass J;
  J.lb = lb; J.p = p; J.invert = invert; J.map = map;
  const int nr = slb_par_for(lb->nnodes, merge_range, &J);
  long ns = 0;
  long long si = 0, sj = 0, sk = 0;
  for (int t = 0; t < nr; t++) { ns += J.ns[t]; si += J.si[t]; sj += J.sj[t]; sk += J.sk[t]; }
  *ox = (LBReal)si; *oy = (LBReal)sj; *oz = (LBReal)sk;
  return ns;
}

/* All solid nodes become one object at their centroid (src/lbobjs.c:362-392). */
LBObj flattenobj(LBSys lb) {
  free_list(lb);
  LBObj p = addobj(lb);
  zeroobj(p);
  LBReal ox = 0.0, oy = 0.0, oz = 0.0;
  int ns = (int)merge_nodes(lb, p, 0, NULL, &ox, &oy, &oz);
  if (ns == 0) {
    if (slb_verbosity & OBJECT_DETAILS) printf("No sites set in object\n");
    drop_head(lb, p);
    slb_obj_changed(lb);
    return NULL;
  }
  lb->changed_objects = 1;
  slb_obj_changed(lb);
  if (slb_verbosity & OBJECT_DETAILS) printf("Object space flattened to id no %i, # nodes = %i\n", p->id, ns);
  p->posx = ox / ns; p->posy = oy / ns; p->posz = oz / ns;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : flattenobj done\n", lb->step);
  return p;
}

/* Swap solid and fluid; the new solid is one object (src/lbobjs.c:397-428). */
LBObj invertobj(LBSys lb) {
  free_list(lb);
  LBObj p = addobj(lb);
  zeroobj(p);
  LBReal ox = 0.0, oy = 0.0, oz = 0.0;
  int ns = (int)merge_nodes(lb, p, 1, NULL, &ox, &oy, &oz);
  slb_obj_changed(lb);
  if (ns == 0) {
    if (slb_verbosity & OBJECT_DETAILS) printf("No sites set in object\n");
    drop_head(lb, p);
    return NULL;
  }
  lb->changed_objects = 1;
  if (slb_verbosity & OBJECT_DETAILS) printf("Object space inverted to id no %i, # nodes = %i\n", p->id, ns);
  p->posx = ox / ns; p->posy = oy / ns; p->posz = oz / ns;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : invertobj done\n", lb->step);
  return p;
}

/* Remove every object and reset the populations, node velocities and node
 * forces to zero (src/lbobjs.c:433-453). */
void removeobjs(LBSys lb) {
  slb_lbpriv *S = PRIV(lb);
  free_list(lb);
  long N = lb->nnodes;
  memset(lb->obj, 0, sizeof(LBObj) * (size_t)N);
  memset(lb->ux, 0, sizeof(LBReal) * N); memset(lb->uy, 0, sizeof(LBReal) * N);
  memset(lb->uz, 0, sizeof(LBReal) * N);
  memset(lb->fx, 0, sizeof(LBReal) * N); memset(lb->fy, 0, sizeof(LBReal) * N);
  memset(lb->fz, 0, sizeof(LBReal) * N);
  memset(lb->a, 0, sizeof(LBReal) * 15 * N);
  slbd_zero_pops(S->dev);
  S->pops_loc = LOC_SYNC;
  S->u_loc = LOC_HOST;
  S->f_loc = LOC_HOST;
  slb_obj_changed(lb);
  if (slb_verbosity & OBJECT_DETAILS) printf("All objects removed\n");
  lb->changed_objects = 1;
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("LB %i : removeobjs done\n", lb->step);
}

void wrobjs(LBSys lb) {
  for (LBObj p = lb->obj_head; p != NULL; p = p->next_obj) {
    printf("Object id (base) = %i (%i)\n", p->id, p->base_obj->id);
    printf(" posx, y, z = %f %f %f\n", p->posx, p->posy, p->posz);
    printf(" velx, y, z = %f %f %f\n", p->velx, p->vely, p->velz);
    printf(" angx, y, z = %f %f %f\n", p->angx, p->angy, p->angz);
    printf(" frcx, y, z = %f %f %f\n", p->frcx, p->frcy, p->frcz);
    printf(" trqx, y, z = %f %f %f\n", p->trqx, p->trqy, p->trqz);
  }
}

/* Human-readable map of the solid nodes, one k slice at a time, objects
 * drawn with the glyph of their id (src/lbobjs.c:472-516). */
void wrmap(LBSys lb, char *name) {
  static const char glyphs[] = "ox+*#acemnrsuvwxz?";
  const int ng = (int)strlen(glyphs);
  FILE *fp = fopen(name, "w");
  if (fp == NULL) {
    fprintf(stderr, "Couldn't open %s\n", name);
    return;
  }
  int *count = (int *)calloc(ng, sizeof(int));
  if (count == NULL) error("wrmap: out of host memory");
  for (int k = 0; k < lb->esk; k++) {
    fprintf(fp, "\nSlice k = %3i\n", k);
    fprintf(fp, "            ");
    for (int i = 0; i < lb->esi; i++) fprintf(fp, "%c", i < 10 ? ' ' : ('0' + (i / 10) % 10));
    fprintf(fp, "\n");
    fprintf(fp, "Col i =     ");
    for (int i = 0; i < lb->esi; i++) fprintf(fp, "%c", ('0' + (i % 10)));
    fprintf(fp, "\n");
    for (int j = 0; j < lb->esj; j++) {
      fprintf(fp, "Row j = %3i ", j);
      for (int i = 0; i < lb->esi; i++) {
        LBObj p = lb->obj[nidx(lb, i, j, k)];
        char c = '.';
        if (p != NULL) {
          int id = p->id;
          if (id >= ng) id = ng - 1;
          c = glyphs[id];
          count[id]++;
        }
        fprintf(fp, "%c", c);
      }
      fprintf(fp, "\n");
    }
  }
  fclose(fp);
  printf("\n id shape count\n");
  for (int g = 0; g < ng; g++)
    if (count[g] > 0) printf("%3i   %c %6i\n", g, glyphs[g], count[g]);
  printf("\n");
  free(count);
  printf("Map of object space written to %s\n", name);
}

typedef struct {
  LBSys lb;
  long n[SLB_PAR_MAX];
} count_pass;
static void solid_range(void *c, long lo, long hi, int tid) {
  count_pass *J = (count_pass *)c;
  long m = 0;
  for (long n = lo; n < hi; n++) m += J->lb->obj[n] != NULL;
  J->n[tid] = m;
}
/* solid nodes of obj[] */
static long slb_count_solid(LBSys lb) {
  count_pass J;
  J.lb = lb;
  const int nr = slb_par_for(lb->nnodes, solid_range, &J);
  long m = 0;
  for (int t = 0; t < nr; t++) m += J.n[t];
  return m;
}

/* ------------------------------------------------------------------ */
/* Extensions: batched spheres and random sequential addition packs.   */

int slb_insert_spheres(LBSys lb, int n, const double *rxyz) {
  int placed = 0;
  for (int s = 0; s < n; s++)
    if (insertsph(lb, rxyz[4 * s], rxyz[4 * s + 1], rxyz[4 * s + 2], rxyz[4 * s + 3])) placed++;
  return placed;
}

/* splitmix64 (Steele, Lea, Flood 2014) */
static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static double unit01(uint64_t *s) { return (splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

/* Centre of trial t: the stream's draws 3t+1, 3t+2, 3t+3 times the box edges,
 * x then y then z (the state after k draws is seed + k times the splitmix64
 * increment, so trial t needs no earlier draws). */
static void rsa_centre(uint64_t seed, long t, const double *L, double *c) {
  uint64_t st = seed + 3ULL * (uint64_t)t * 0x9E3779B97F4A7C15ULL;
  const double x = unit01(&st) * L[0];
  const double y = unit01(&st) * L[1];
  const double z = unit01(&st) * L[2];
  c[0] = x; c[1] = y; c[2] = z;
}

/* Cell grid for the minimum-image overlap test. */
typedef struct {
  int nc[3];
  double cs[3], L[3];
  int *head, *next;
  double *xyz;
  int n, cap;
} rsa_grid;

static int rsa_overlaps(const rsa_grid *g, const double *c, double d2min) {
  int ci[3];
  for (int d = 0; d < 3; d++) {
    ci[d] = (int)floor(c[d] / g->cs[d]);
    if (ci[d] >= g->nc[d]) ci[d] = g->nc[d] - 1;
    if (ci[d] < 0) ci[d] = 0;
  }
  for (int a = -1; a <= 1; a++)
    for (int b = -1; b <= 1; b++)
      for (int e = -1; e <= 1; e++) {
        int x = (ci[0] + a + g->nc[0]) % g->nc[0];
        int y = (ci[1] + b + g->nc[1]) % g->nc[1];
        int z = (ci[2] + e + g->nc[2]) % g->nc[2];
        for (int q = g->head[(x * g->nc[1] + y) * g->nc[2] + z]; q >= 0; q = g->next[q]) {
          double r2 = 0.0;
          for (int d = 0; d < 3; d++) {
            double dx = c[d] - g->xyz[3 * q + d];
            dx -= g->L[d] * floor(dx / g->L[d] + 0.5);
            r2 += dx * dx;
          }
          if (r2 < d2min) return 1;
        }
      }
  return 0;
}

static void rsa_add(rsa_grid *g, const double *c) {
  if (g->n == g->cap) {
    g->cap = g->cap ? 2 * g->cap : 1024;
    g->xyz = (double *)slb_xrealloc(g->xyz, sizeof(double) * 3 * g->cap);
    g->next = (int *)slb_xrealloc(g->next, sizeof(int) * g->cap);
  }
  int ci[3];
  for (int d = 0; d < 3; d++) {
    ci[d] = (int)floor(c[d] / g->cs[d]);
    if (ci[d] >= g->nc[d]) ci[d] = g->nc[d] - 1;
    if (ci[d] < 0) ci[d] = 0;
  }
  int cell = (ci[0] * g->nc[1] + ci[1]) * g->nc[2] + ci[2];
  memcpy(g->xyz + 3 * g->n, c, sizeof(double) * 3);
  g->next[g->n] = g->head[cell];
  g->head[cell] = g->n;
  g->n++;
}

int slb_rsa_pack(LBSys lb, double R, double phi, unsigned long long seed, long maxtrials,
                 double *centres, int maxcentres, double *phi_out) {
  rsa_grid g;
  memset(&g, 0, sizeof g);
  g.L[0] = lb->esi; g.L[1] = lb->esj; g.L[2] = lb->esk;
  int ncell = 1;
  for (int d = 0; d < 3; d++) {
    g.nc[d] = (int)floor(g.L[d] / (2.0 * R));
    if (g.nc[d] < 1) g.nc[d] = 1;
    if (g.nc[d] < 3) g.nc[d] = 1;  /* small boxes: one cell, direct scan */
    g.cs[d] = g.L[d] / g.nc[d];
    ncell *= g.nc[d];
  }
  g.head = (int *)slb_xmalloc(sizeof(int) * ncell);
  for (int c = 0; c < ncell; c++) g.head[c] = -1;
  long target = (long)ceil(phi * (double)lb->nnodes);
  long solid = slb_count_solid(lb);
  int placed = 0;
  int saved = slb_verbosity;
  slb_verbosity &= ~OBJECT_DETAILS;
  /* Trials in batches: the device drops the candidates that clearly overlap
   * a sphere accepted before the batch (late in a pack, nearly all of them);
   * the rest get the exact sequential test here, in trial order. */
  slbd_rsa *D = slbd_rsa_create(g.L, g.cs, g.nc);
  int m = 1024, uploaded = -1;
  uint8_t *pass = (uint8_t *)slb_xmalloc(1 << 20);
  long t = 0;
  while (t < maxtrials && solid < target) {
    const int mm = (maxtrials - t < m) ? (int)(maxtrials - t) : m;
    if (uploaded != g.n) { slbd_rsa_update(D, g.head, g.next, g.xyz, g.n); uploaded = g.n; }
    slbd_rsa_filter(D, (uint64_t)seed, t, mm, 4.0 * R * R, pass);
    int q, npass = 0;
    for (q = 0; q < mm && solid < target; q++) {
      if (!pass[q]) continue;
      npass++;
      double c[3];
      rsa_centre((uint64_t)seed, t + q, g.L, c);
      if (rsa_overlaps(&g, c, 4.0 * R * R)) continue;
      rsa_add(&g, c);
      for (int a = -1; a <= 1; a++)
        for (int b = -1; b <= 1; b++)
          for (int e = -1; e <= 1; e++) {
            double o[3] = {c[0] + a * g.L[0], c[1] + b * g.L[1], c[2] + e * g.L[2]};
            if (o[0] + R < 0.0 || o[0] - R > g.L[0] || o[1] + R < 0.0 || o[1] - R > g.L[1] ||
                o[2] + R < 0.0 || o[2] - R > g.L[2])
              continue;
            LBObj p = addobj(lb);
            zeroobj(p);
            int ov = 0;
            int ns = raster_sphere(lb, p, R, o[0], o[1], o[2], &ov);
            if (ns == 0) { drop_head(lb, p); lb->next_id--; continue; }
            p->posx = o[0]; p->posy = o[1]; p->posz = o[2];
            solid += ns;
            lb->changed_objects = 1;
          }
      if (centres && placed < maxcentres) memcpy(centres + 3 * placed, c, sizeof(double) * 3);
      placed++;
    }
    t += q;
    /* batch size: a few hundred survivors per batch at most */
    if (npass < 64 && m < (1 << 20)) m *= 2;
    else if (npass > 512 && m > 256) m /= 2;
  }
  free(pass);
  slbd_rsa_destroy(D);
  slb_verbosity = saved;
  slb_obj_changed(lb);
  if (phi_out) *phi_out = (double)solid / (double)lb->nnodes;
  free(g.head); free(g.next); free(g.xyz);
  return placed;
}

/* slb_rsa_pack followed by flattenobj, with the sphere images counted and
 * rasterised on the GPU.  The acceptance of candidates is the same
 * sequential test in trial order; the spheres accepted from one batch of
 * trials have their images' node counts taken on the device together, and
 * the pack ends after the first sphere whose images bring the solid count to
 * the target, as slb_rsa_pack ends.  The images of the placed spheres are
 * then marked in a byte map on the device, and every solid node (those of
 * objects already present and the map's) becomes the one object of
 * flattenobj.  Same centres, solid fraction, object and node map. */
typedef struct {
  double *o;
  int *box, *sph;
  int n, cap;
} img_list;

static void img_push(img_list *L, const double *o, const int *box, int sph) {
  if (L->n == L->cap) {
    L->cap = L->cap ? 2 * L->cap : 1024;
    L->o = (double *)slb_xrealloc(L->o, sizeof(double) * 3 * L->cap);
    L->box = (int *)slb_xrealloc(L->box, sizeof(int) * 6 * L->cap);
    L->sph = (int *)slb_xrealloc(L->sph, sizeof(int) * L->cap);
  }
  memcpy(L->o + 3 * L->n, o, sizeof(double) * 3);
  memcpy(L->box + 6 * L->n, box, sizeof(int) * 6);
  L->sph[L->n] = sph;
  L->n++;
}

int slb_rsa_pack_flat(LBSys lb, double R, double phi, unsigned long long seed, long maxtrials,
                      double *centres, int maxcentres, double *phi_out) {
  slb_lbpriv *S = PRIV(lb);
  rsa_grid g;
  memset(&g, 0, sizeof g);
  g.L[0] = lb->esi; g.L[1] = lb->esj; g.L[2] = lb->esk;
  int ncell = 1;
  for (int d = 0; d < 3; d++) {
    g.nc[d] = (int)floor(g.L[d] / (2.0 * R));
    if (g.nc[d] < 1) g.nc[d] = 1;
    if (g.nc[d] < 3) g.nc[d] = 1;
    g.cs[d] = g.L[d] / g.nc[d];
    ncell *= g.nc[d];
  }
  g.head = (int *)slb_xmalloc(sizeof(int) * ncell);
  for (int c = 0; c < ncell; c++) g.head[c] = -1;
  const long target = (long)ceil(phi * (double)lb->nnodes);
  long solid = slb_count_solid(lb);
  const double radsq = R * R, ra = fabs(R);
  const int dims[3] = {lb->esi, lb->esj, lb->esk};
  img_list imgs = {NULL, NULL, NULL, 0, 0};
  int *counts = NULL, ccap = 0;
  int placed = 0, done = 0;
  slbd_rsa *D = slbd_rsa_create(g.L, g.cs, g.nc);
  int m = 1024, uploaded = -1;
  uint8_t *pass = (uint8_t *)slb_xmalloc(1 << 20);
  long t = 0;
  while (t < maxtrials && solid < target && !done) {
    const int mm = (maxtrials - t < m) ? (int)(maxtrials - t) : m;
    if (uploaded != g.n) { slbd_rsa_update(D, g.head, g.next, g.xyz, g.n); uploaded = g.n; }
    slbd_rsa_filter(D, (uint64_t)seed, t, mm, 4.0 * R * R, pass);
    const int first_img = imgs.n;
    int npass = 0;
    for (int q = 0; q < mm; q++) {
      if (!pass[q]) continue;
      npass++;
      double c[3];
      rsa_centre((uint64_t)seed, t + q, g.L, c);
      if (rsa_overlaps(&g, c, 4.0 * R * R)) continue;
      rsa_add(&g, c);
      for (int a = -1; a <= 1; a++)
        for (int b = -1; b <= 1; b++)
          for (int e = -1; e <= 1; e++) {
            double o[3] = {c[0] + a * g.L[0], c[1] + b * g.L[1], c[2] + e * g.L[2]};
            if (o[0] + R < 0.0 || o[0] - R > g.L[0] || o[1] + R < 0.0 || o[1] - R > g.L[1] ||
                o[2] + R < 0.0 || o[2] - R > g.L[2])
              continue;
            int box[6];
            for (int d = 0; d < 3; d++) axis_range(o[d], ra, dims[d], &box[d], &box[3 + d]);
            img_push(&imgs, o, box, placed);
          }
      if (centres && placed < maxcentres) memcpy(centres + 3 * placed, c, sizeof(double) * 3);
      placed++;
    }
    t += mm;
    /* node counts of this batch's images, then the stopping sphere */
    const int nb = imgs.n - first_img;
    if (nb > ccap) { ccap = nb; counts = (int *)slb_xrealloc(counts, sizeof(int) * ccap); }
    slbd_sph_count(S->dev, imgs.o + 3 * first_img, imgs.box + 6 * first_img, nb, radsq, counts);
    for (int q = 0; q < nb; q++) {
      solid += counts[q];
      const int sp = imgs.sph[first_img + q];
      const int last = (q + 1 == nb) || imgs.sph[first_img + q + 1] != sp;
      if (last && solid >= target) {
        placed = sp + 1;
        imgs.n = first_img + q + 1;
        done = 1;
        break;
      }
    }
    if (npass < 64 && m < (1 << 20)) m *= 2;
    else if (npass > 512 && m > 256) m /= 2;
  }
  free(pass);
  free(counts);
  slbd_rsa_destroy(D);
  free(g.head); free(g.next); free(g.xyz);
  /* the map of the placed spheres, then flattenobj over it and the objects
   * already present */
  unsigned char *map = (unsigned char *)slb_big_calloc(lb->nnodes, 1);
  if (map == NULL) error("out of host memory");
  slbd_sph_map(S->dev, imgs.o, imgs.box, imgs.n, radsq, map);
  free(imgs.o); free(imgs.box); free(imgs.sph);
  free_list(lb);
  LBObj p = addobj(lb);
  zeroobj(p);
  LBReal ox = 0.0, oy = 0.0, oz = 0.0;
  int ns = (int)merge_nodes(lb, p, 0, map, &ox, &oy, &oz);
  slb_big_free(map);
  if (ns == 0) {
    if (slb_verbosity & OBJECT_DETAILS) printf("No sites set in object\n");
    drop_head(lb, p);
  } else {
    lb->changed_objects = 1;
    if (slb_verbosity & OBJECT_DETAILS)
      printf("Object space flattened to id no %i, # nodes = %i\n", p->id, ns);
    p->posx = ox / ns; p->posy = oy / ns; p->posz = oz / ns;
  }
  slb_obj_changed(lb);
  if (phi_out) *phi_out = (double)solid / (double)lb->nnodes;
  return placed;
}

/* ------------------------------------------------------------------ */
/* Solid maps of sphere packs without a lattice (thread-safe: no globals,
 * no objects), and their insertion as one object.                       */

/* Mark the nodes of one sphere image; returns the nodes that were fluid. */
static long raster_map(unsigned char *solid, const int *L, double R, const double *o) {
  int lo[3], hi[3];
  for (int d = 0; d < 3; d++) axis_range(o[d], R, L[d], &lo[d], &hi[d]);
  const double radsq = R * R;
  long added = 0;
  for (int i = lo[0]; i <= hi[0]; i++)
    for (int j = lo[1]; j <= hi[1]; j++) {
      const double part = row_part(i, j, o[0], o[1]);
      if (!(part < radsq)) continue;
      for (int k = lo[2]; k <= hi[2]; k++)
        if (in_sphere_row(part, k, o[2], radsq)) {
          long n = (long)L[2] * ((long)L[1] * i + j) + k;
          if (!solid[n]) { solid[n] = 1; added++; }
        }
    }
  return added;
}

int slb_pack_map(int esi, int esj, int esk, double R, double phi, unsigned long long seed,
                 int overlap, long maxtrials, unsigned char *solid, double *phi_out) {
  const int L[3] = {esi, esj, esk};
  const long N = (long)esi * esj * esk;
  rsa_grid g;
  memset(&g, 0, sizeof g);
  int ncell = 1;
  for (int d = 0; d < 3; d++) {
    g.L[d] = L[d];
    g.nc[d] = (int)floor(g.L[d] / (2.0 * R));
    if (g.nc[d] < 3) g.nc[d] = 1;
    g.cs[d] = g.L[d] / g.nc[d];
    ncell *= g.nc[d];
  }
  if (!overlap) {
    g.head = (int *)malloc(sizeof(int) * ncell);
    if (!g.head) return -1;
    for (int c = 0; c < ncell; c++) g.head[c] = -1;
  }
  memset(solid, 0, (size_t)N);
  const long target = (long)ceil(phi * (double)N);
  long nsolid = 0;
  int placed = 0;
  for (long t = 0; t < maxtrials && nsolid < target; t++) {
    double c[3];
    rsa_centre((uint64_t)seed, t, g.L, c);
    if (!overlap) {
      if (rsa_overlaps(&g, c, 4.0 * R * R)) continue;
      rsa_add(&g, c);
    }
    for (int a = -1; a <= 1; a++)
      for (int b = -1; b <= 1; b++)
        for (int e = -1; e <= 1; e++) {
          double o[3] = {c[0] + a * g.L[0], c[1] + b * g.L[1], c[2] + e * g.L[2]};
          if (o[0] + R < 0.0 || o[0] - R > g.L[0] || o[1] + R < 0.0 || o[1] - R > g.L[1] ||
              o[2] + R < 0.0 || o[2] - R > g.L[2])
            continue;
          nsolid += raster_map(solid, L, R, o);
        }
    placed++;
  }
  free(g.head); free(g.next); free(g.xyz);
  if (phi_out) *phi_out = (double)nsolid / (double)N;
  return placed;
}

typedef struct {
  LBSys lb;
  LBObj p;
  const unsigned char *solid;
  long ns[SLB_PAR_MAX];
} map_pass;
static void map_range(void *c, long lo, long hi, int tid) {
  map_pass *J = (map_pass *)c;
  long ns = 0;
  for (long n = lo; n < hi; n++)
    if (J->solid[n]) { J->lb->obj[n] = J->p; ns++; }
  J->ns[tid] = ns;
}

LBObj slb_insert_map(LBSys lb, const unsigned char *solid) {
  LBObj p = addobj(lb);
  zeroobj(p);
  map_pass J;
  J.lb = lb; J.p = p; J.solid = solid;
  const int nr = slb_par_for(lb->nnodes, map_range, &J);
  long ns = 0;
  for (int t = 0; t < nr; t++) ns += J.ns[t];
  if (ns == 0) {
    drop_head(lb, p);
    lb->next_id--;
    return NULL;
  }
  lb->changed_objects = 1;
  slb_obj_changed(lb);
  return p;
}
