/*
 * slb_extra.c -- image entry points of the sunlightlb.h API (reference
 * src/mpimage.c).  These visualisation routines are outside the GPU hot
 * path; here they report that they are not provided by this library and
 * return neutral values.  (The velocity autocorrelation is in slb_vacf.c.)
 */
#include <stdio.h>
#include <stdlib.h>
#include "slb_internal.h"

static void unsupported(const char *what) {
  fprintf(stderr, "sunlightlb: %s is not provided by the MI355X library\n", what);
}


LBColour newlbcolour(int r, int g, int b) {
  LBColour c = (LBColour)malloc(sizeof(LBColour_type));
  c->red = r; c->green = g; c->blue = b;
  return c;
}
void lbcolourset(LBColour c, int r, int g, int b) { c->red = r; c->green = g; c->blue = b; }
void lbcolourcopy(LBColour d, LBColour s) { *d = *s; }
LBColourMap newlbcolourmap(LBReal lo, LBReal hi, LBColour a, LBColour b) {
  LBColourMap m = (LBColourMap)malloc(sizeof(LBColourMap_type));
  lbcolourmapset(m, lo, hi, a, b);
  return m;
}
void lbcolourmapset(LBColourMap m, LBReal lo, LBReal hi, LBColour a, LBColour b) {
  m->v_lo = lo; m->v_hi = hi;
  m->red_lo = a->red; m->green_lo = a->green; m->blue_lo = a->blue;
  m->red_hi = b->red; m->green_hi = b->green; m->blue_hi = b->blue;
}
LBImage newlbimage(int w, int h, LBColour back, LBColour objc) { (void)w; (void)h; (void)back; (void)objc; unsupported("newlbimage"); return NULL; }
void lbimageclear(LBImage im) { (void)im; unsupported("lbimageclear"); }
void lbimagemap(LBImage im, LBSys lb, char *d, int s) { (void)im; (void)lb; (void)d; (void)s; unsupported("lbimagemap"); }
void lbimagevel(LBImage im, LBColourMap m, LBSys lb, char *a, char *b, int s) { (void)im; (void)m; (void)lb; (void)a; (void)b; (void)s; unsupported("lbimagevel"); }
void lbcolourget(LBImage im, LBColour c, int i, int j) { (void)im; (void)c; (void)i; (void)j; unsupported("lbcolourget"); }
int lbgetred(LBImage im, int i, int j) { (void)im; (void)i; (void)j; return 0; }
int lbgetgreen(LBImage im, int i, int j) { (void)im; (void)i; (void)j; return 0; }
int lbgetblue(LBImage im, int i, int j) { (void)im; (void)i; (void)j; return 0; }
void lbcolourclamp(int *r, int *g, int *b) {
  if (*r < 0) *r = 0; if (*r > 255) *r = 255;
  if (*g < 0) *g = 0; if (*g > 255) *g = 255;
  if (*b < 0) *b = 0; if (*b > 255) *b = 255;
}
void mpimageset(LBImage im, MPSys mp, LBColourMap m, char *a, char *b, int s) { (void)im; (void)mp; (void)m; (void)a; (void)b; (void)s; unsupported("mpimageset"); }
