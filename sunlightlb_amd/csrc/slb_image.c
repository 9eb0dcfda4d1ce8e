/*
 * slb_image.c -- colour images of lattice slices (the entry points of
 * src/mpimage.c): colours, colour maps, images; the solid map, a velocity
 * component or magnitude, or an MP concentration field on one lattice plane.
 *
 * Images are host arrays of width x height pixels, pixel (a, b) at
 * height * a + b.  A slice normal to x, y or z maps its two in-plane lattice
 * axes (j, k), (i, k) or (i, j) to (a, b), centred in the image; lattice
 * nodes that fall outside the image are not drawn.  The fields are read from
 * the host mirrors, which are brought up to date first (one transfer of the
 * populations or of the MP field per call; images are diagnostics).
 *
 * Deliberate differences from src/mpimage.c, whose slice and colour code
 * have these defects:
 *   - slices normal to x and y there write pixel (slice, j) and (i, slice):
 *     the plane is collapsed onto one column or row; here (j, k) and (i, k);
 *   - the centring offsets there are (lattice - image)/2 added to the lattice
 *     coordinate, which moves a smaller lattice off the image; here
 *     (image - lattice)/2, with pixels outside the image skipped;
 *   - lbimagevel there starts the blue channel from green_lo; here blue_lo;
 *   - the velocity correction there uses the object force field fx, fy, fz;
 *     here the body force (b plus dpdx, dpdy, dpdz), as accvmean and vgrab do
 *     (src/lbmeasure.c:71-93);
 *   - the pixel getters there accept i == width or j == height (one past the
 *     end); here they are out of range.
 * On a slice normal to z of a lattice the size of the image, with
 * green_lo == blue_lo and no body force, the images equal the reference's.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include "slb_internal.h"

static int clamp255(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }

void lbcolourclamp(int *r, int *g, int *b) {
  *r = clamp255(*r);
  *g = clamp255(*g);
  *b = clamp255(*b);
}

void lbcolourset(LBColour c, int r, int g, int b) {
  c->red = r; c->green = g; c->blue = b;
  lbcolourclamp(&c->red, &c->green, &c->blue);
}

LBColour newlbcolour(int r, int g, int b) {
  LBColour c = (LBColour)malloc(sizeof(LBColour_type));
  if (c == NULL) error("No room for colour\n");
  lbcolourset(c, r, g, b);
  return c;
}

void lbcolourcopy(LBColour d, LBColour s) { *d = *s; }

void lbcolourmapset(LBColourMap m, LBReal lo, LBReal hi, LBColour a, LBColour b) {
  m->v_lo = lo; m->v_hi = hi;
  m->red_lo = a->red; m->green_lo = a->green; m->blue_lo = a->blue;
  m->red_hi = b->red; m->green_hi = b->green; m->blue_hi = b->blue;
  printf("Colour Map low  end p = %f, colour = (%03i,%03i,%03i)\n", lo, m->red_lo, m->green_lo,
         m->blue_lo);
  printf("Colour Map high end p = %f, colour = (%03i,%03i,%03i)\n", hi, m->red_hi, m->green_hi,
         m->blue_hi);
}

LBColourMap newlbcolourmap(LBReal lo, LBReal hi, LBColour a, LBColour b) {
  LBColourMap m = (LBColourMap)malloc(sizeof(LBColourMap_type));
  if (m == NULL) error("No room for colour map\n");
  lbcolourmapset(m, lo, hi, a, b);
  return m;
}

LBImage newlbimage(int w, int h, LBColour back, LBColour objc) {
  LBImage im = (LBImage)malloc(sizeof(LBImage_type));
  if (im == NULL || w < 1 || h < 1) error("No room for image\n");
  const size_t area = (size_t)w * h;
  im->width = w; im->height = h;
  im->red = (int *)malloc(area * sizeof(int));
  im->green = (int *)malloc(area * sizeof(int));
  im->blue = (int *)malloc(area * sizeof(int));
  if (!im->red || !im->green || !im->blue) error("No room for image\n");
  im->back_red = back->red; im->back_green = back->green; im->back_blue = back->blue;
  im->object_red = objc->red; im->object_green = objc->green; im->object_blue = objc->blue;
  printf("Image initialised to size %i * %i = %i\n", h, w, (int)area);
  printf("Image background = (%03i,%03i,%03i), ", im->back_red, im->back_green, im->back_blue);
  printf("objects = (%03i,%03i,%03i)\n", im->object_red, im->object_green, im->object_blue);
  return im;
}

void lbimageclear(LBImage im) {
  const long area = (long)im->width * im->height;
  for (long p = 0; p < area; p++) {
    im->red[p] = im->back_red;
    im->green[p] = im->back_green;
    im->blue[p] = im->back_blue;
  }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("Image cleared to background\n");
}

/* One lattice plane mapped onto an image: node index and pixel index of
 * in-plane position (a, b), or -1 when the pixel is outside the image. */
typedef struct {
  int na, nb;           /* in-plane extents */
  long base, sa, sb;    /* node = base + a*sa + b*sb */
  int hoff, voff, w, h;
} Slice;

static int slice_of(const LBSys lb, const LBImage im, const char *dirn, int slice,
                    const char *who, Slice *S) {
  const long pj = lb->esk, pi = (long)lb->esj * lb->esk;
  int n;
  switch (dirn ? dirn[0] : '\0') {
    case 'x': case 'i':
      n = lb->esi; S->na = lb->esj; S->nb = lb->esk; S->sa = pj; S->sb = 1; S->base = slice * pi;
      break;
    case 'y': case 'j':
      n = lb->esj; S->na = lb->esi; S->nb = lb->esk; S->sa = pi; S->sb = 1; S->base = slice * pj;
      break;
    case 'z': case 'k':
      n = lb->esk; S->na = lb->esi; S->nb = lb->esj; S->sa = pi; S->sb = pj; S->base = slice;
      break;
    default:
      printf("Unrecognised direction %s in %s\n", dirn ? dirn : "(null)", who);
      return 0;
  }
  if (slice < 0 || slice >= n) {
    printf("%s: slice %i out of range 0 -- %i\n", who, slice, n - 1);
    return 0;
  }
  S->w = im->width; S->h = im->height;
  S->hoff = (im->width - S->na) / 2;
  S->voff = (im->height - S->nb) / 2;
  return 1;
}

static long pixel_of(const Slice *S, int a, int b) {
  const int pa = a + S->hoff, pb = b + S->voff;
  if (pa < 0 || pa >= S->w || pb < 0 || pb >= S->h) return -1;
  return (long)S->h * pa + pb;
}

static void put(LBImage im, long p, int r, int g, int b) {
  im->red[p] = r; im->green[p] = g; im->blue[p] = b;
}

void lbimagemap(LBImage im, LBSys lb, char *dirn, int slice) {
  Slice S;
  if (!slice_of(lb, im, dirn, slice, "lbimagemap", &S)) return;
  if (lb->changed_objects) syncobjs(lb);
  for (int a = 0; a < S.na; a++)
    for (int b = 0; b < S.nb; b++) {
      const long p = pixel_of(&S, a, b);
      if (p < 0) continue;
      const long n = S.base + a * S.sa + b * S.sb;
      if (lb->obj[n] == NULL) put(im, p, im->back_red, im->back_green, im->back_blue);
      else put(im, p, im->object_red, im->object_green, im->object_blue);
    }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("Image cleared to background plus objects\n");
}

/* colour of value v on map m: lo + (v - v_lo) * (1/(v_hi - v_lo)) * (hi - lo)
 * per channel, truncated toward zero, then clamped */
static void map_colour(const LBColourMap m, double v, int *r, int *g, int *b) {
  const double sf = 1.0 / (m->v_hi - m->v_lo);
  const double t = sf * (v - m->v_lo);
  *r = (int)(m->red_lo + t * (double)(m->red_hi - m->red_lo));
  *g = (int)(m->green_lo + t * (double)(m->green_hi - m->green_lo));
  *b = (int)(m->blue_lo + t * (double)(m->blue_hi - m->blue_lo));
  lbcolourclamp(r, g, b);
}

void lbimagevel(LBImage im, LBColourMap m, LBSys lb, char *what, char *dirn, int slice) {
  Slice S;
  if (!slice_of(lb, im, dirn, slice, "lbimagevel", &S)) return;
  int comp = 3;   /* magnitude */
  switch (what ? what[0] : '\0') {
    case 'x': case 'i': comp = 0; break;
    case 'y': case 'j': comp = 1; break;
    case 'z': case 'k': comp = 2; break;
  }
  if (lb->changed_objects) syncobjs(lb);
  slb_host_pops(lb);
  double fac = 0.0;
  if (lb->vel_corrected) {
    if (lb->status == POST_PROPAGATE) fac = 0.5;
    else if (lb->status == POST_COLLIDE) fac = -0.5;
  }
  const int *c = lb->c;
  for (int a = 0; a < S.na; a++)
    for (int b = 0; b < S.nb; b++) {
      const long p = pixel_of(&S, a, b);
      if (p < 0) continue;
      const long n = S.base + a * S.sa + b * S.sb;
      if (lb->obj[n] != NULL) {
        put(im, p, im->object_red, im->object_green, im->object_blue);
        continue;
      }
      const LBReal *an = lb->a + 15 * n;
      double rho = lb->rho0, j[3] = {0.0, 0.0, 0.0};
      for (int l = 0; l < 15; l++) {
        rho += an[l];
        for (int q = 0; q < 3; q++)
          if (c[3 * l + q]) j[q] += c[3 * l + q] * an[l];
      }
      const double bf[3] = {lb->bx[n] + lb->dpdx, lb->by[n] + lb->dpdy, lb->bz[n] + lb->dpdz};
      double u[3];
      for (int q = 0; q < 3; q++) u[q] = (j[q] + fac * bf[q]) / rho;
      const double v = comp < 3 ? u[comp] : sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
      int r, g, bl;
      map_colour(m, v, &r, &g, &bl);
      put(im, p, r, g, bl);
    }
  if (slb_verbosity & TRACE_ROUTINE_EXITS) printf("velocity field %s imaged\n", what);
}

void mpimageset(LBImage im, MPSys mp, LBColourMap m, char *what, char *dirn, int slice) {
  int add;
  switch (what ? what[0] : '\0') {
    case 'a': add = 1; break;
    case 's': add = 0; break;
    default:
      printf("Unrecognised add/subtract colour command %s\n", what ? what : "(null)");
      return;
  }
  LBSys lb = mp->lb;
  Slice S;
  if (!slice_of(lb, im, dirn, slice, "mpimage", &S)) return;
  if (lb->changed_objects) syncobjs(lb);
  slb_host_p(mp);
  const double sf = 1.0 / (m->v_hi - m->v_lo);
  const double rr = m->red_hi - m->red_lo, gr = m->green_hi - m->green_lo,
               br = m->blue_hi - m->blue_lo;
  for (int a = 0; a < S.na; a++)
    for (int b = 0; b < S.nb; b++) {
      const long p = pixel_of(&S, a, b);
      if (p < 0) continue;
      const long n = S.base + a * S.sa + b * S.sb;
      if (lb->obj[n] != NULL) {
        put(im, p, im->object_red, im->object_green, im->object_blue);
        continue;
      }
      const double t = sf * (mp->p[n] - m->v_lo);
      const int sg = add ? 1 : -1;
      int r = im->red[p] + sg * (int)(t * rr);
      int g = im->green[p] + sg * (int)(t * gr);
      int bl = im->blue[p] + sg * (int)(t * br);
      lbcolourclamp(&r, &g, &bl);
      put(im, p, r, g, bl);
    }
  if (slb_verbosity & TRACE_ROUTINE_EXITS)
    printf("MP Concentration field colour %s image\n", add ? "added to" : "subtracted from");
}

static int in_image(const LBImage im, int i, int j, const char *who) {
  if (i < 0 || i >= im->width || j < 0 || j >= im->height) {
    printf("%s: out of range (i, j) = (%i, %i)\n", who, i, j);
    printf("%s: range = (0, 0) -- (%i, %i)\n", who, im->width - 1, im->height - 1);
    return 0;
  }
  return 1;
}

void lbcolourget(LBImage im, LBColour c, int i, int j) {
  if (!in_image(im, i, j, "lbcolourget")) return;
  const long p = (long)im->height * i + j;
  c->red = im->red[p]; c->green = im->green[p]; c->blue = im->blue[p];
}

int lbgetred(LBImage im, int i, int j) {
  return in_image(im, i, j, "lbgetred") ? im->red[(long)im->height * i + j] : 0;
}

int lbgetgreen(LBImage im, int i, int j) {
  return in_image(im, i, j, "lbgetgreen") ? im->green[(long)im->height * i + j] : 0;
}

int lbgetblue(LBImage im, int i, int j) {
  return in_image(im, i, j, "lbgetblue") ? im->blue[(long)im->height * i + j] : 0;
}
