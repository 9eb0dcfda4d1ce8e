/*
 * slb_warm.c -- learned warm start for steady Stokes flow through a periodic
 * solid map.
 *
 * A 4-level 3D convolutional encoder-decoder (16-64 channels, 3x3x3 taps,
 * periodic padding, ReLU, 2x2x2 average pooling, nearest upsampling with
 * additive skips) maps two features of the lattice -- the solid flag and
 * the signed distance to the solid/fluid interface -- to the 10
 * hydrodynamic moments of the steady state: corrected velocity, density
 * deviation and stress, each per unit body force F (a Stokes solution is
 * linear in F) and the velocity times the viscosity eta = rho0 (tau-1/2)/3.
 * The prediction only initialises the populations; the LB then runs to the
 * requested residual.  Two initialisations:
 *   WARM_EQUILIBRIUM -- equilibrium at rho0 and the predicted velocity,
 *                       then one collide, exactly as lbloadvel builds its
 *                       state (src/lbstore.c:346-401);
 *   WARM_MOMENTS     -- the post-collide state rebuilt from all 10
 *                       predicted moments (density, momentum, stress).
 * Two safeguards keep a warm solve from costing more than a cold one: a
 * start policy, which starts cold when the median residual curve of
 * calibration warm solves reaches the tolerance no sooner than the median
 * cold curve, and a fallback, which resets the populations to rest when the
 * residual of a warm solve does not fall fast enough.
 *
 * Everything runs on the GPU: features (jump flooding), forward and
 * backward passes (BF16 implicit-GEMM MFMA convolutions, FP32
 * accumulation), the loss and Adam.  The training set lives in device
 * memory as 2-channel inputs and 3-channel targets per sample.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sunlightlb.h"
#include "sunlightlb_gpu.h"
#include "slb_internal.h"
#include "slb_dev.h"
#include "sunlightlb_nn.h"

enum { E1A, E1B, E2A, E2B, E3A, E3B, E4A, E4B, D3, D2, D1, OUT, NLAYER };
static const int LCIN[NLAYER] = {16, 16, 16, 32, 32, 64, 64, 64, 64, 32, 16, 16};
static const int LCOUT[NLAYER] = {16, 16, 32, 32, 64, 64, 64, 64, 32, 16, 16, 16};
static const int LLEVEL[NLAYER] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 1, 0, 0};
#define NLEVEL 4
static const int LEVEL_CMAX[NLEVEL] = {16, 32, 64, 64};

#define WN_C 16          /* channels of the input and output tensors */
#define WN_FEAT 2        /* used input channels: solid flag, signed distance */
#define WN_TARG 10       /* used output channels: u (3), drho, stress (6) */
/* Targets per group, with ell the pack's normalising length (mean distance
 * from a fluid node to the solid) -- a Stokes solution at fixed F and eta
 * scales as ell^2 (velocity), ell (density deviation, stress):
 *   u eta / (F ell^2) * WN_SU,  drho / (F ell) * WN_SR,  s / (F ell) * WN_SS */
#define WN_SU 1.0
#define WN_SR 0.125
#define WN_SS 1.0
#define WN_MAGIC "SLBWNET4"
#define WN_CURVE 256     /* checks of the cold reference curve */
static const int WN_GROUP[WN_TARG] = {0, 0, 0, 1, 2, 2, 2, 2, 2, 2};

typedef struct {
  int dims[3];
  long nodes;
  uint16_t *x;   /* [n][WN_FEAT] */
  uint16_t *t;   /* [n][WN_TARG] */
  float *wc;     /* device: per-channel loss weights (WN_TARG) */
  double ms[3];  /* mean square of each target group */
  double ell;    /* normalising length */
} wn_sample;

struct slb_warmnet {
  long wofs[NLAYER], bofs[NLAYER], nparam;
  float *w, *g, *m, *v; /* FP32 master weights, gradient, Adam moments */
  uint16_t *wb;         /* BF16 copy of w */
  uint16_t *wt[NLAYER]; /* transposed taps (input gradient) */
  int adam_step;
  /* activations for the current dims */
  int dims[3];
  long nl[NLEVEL];
  uint16_t *x0, *act[NLAYER], *pool[3], *up[3];
  float *outf;
  uint16_t *gA[NLEVEL], *gB[NLEVEL], *gP[NLEVEL], *dout;
  uint16_t *tbuf; /* targets of an evaluated lattice */
  float *loss;
  float *wc_eval;
  int mode;
  int proj_iters;       /* divergence projection of the prediction (0: off): */
  int proj_passes;      /* CG iterations per pass, passes */
  /* training set */
  int nsamp, cap;
  wn_sample *samp;
  /* cold-start fallback */
  int probe;
  double probe_res;
  int nnote, note_interval;
  double *notes;
  int ncurve, curve_interval;
  double curve[WN_CURVE];
  /* start policy: median residual curve of calibration warm solves */
  int nwnote, wnote_interval;
  double *wnotes;
  int nwcurve, wcurve_interval;
  double wcurve[WN_CURVE];
  uint64_t rng;
};

/* ------------------------------------------------------------------ */

static uint64_t splitmix(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static double urand(uint64_t *s) { return (splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }
static double nrand(uint64_t *s) {
  double u1 = urand(s), u2 = urand(s);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static void die(const char *msg) {
  fprintf(stderr, "sunlightlb warm start: %s\n", msg);
  exit(EXIT_FAILURE);
}

static void layout(slb_warmnet *net) {
  long o = 0;
  for (int L = 0; L < NLAYER; L++) {
    net->wofs[L] = o;
    o += (long)LCOUT[L] * slbn_kpad(LCIN[L]);
    net->bofs[L] = o;
    o += 64;
  }
  net->nparam = o;
}

static void refresh_wt(slb_warmnet *net) {
  for (int L = 0; L < NLAYER; L++)
    slbn_wtrans(net->wb + net->wofs[L], LCOUT[L], LCIN[L], net->wt[L]);
}

static void alloc_params(slb_warmnet *net) {
  layout(net);
  const size_t n = (size_t)net->nparam;
  net->w = slbn_malloc(n * 4);
  net->g = slbn_malloc(n * 4);
  net->m = slbn_malloc(n * 4);
  net->v = slbn_malloc(n * 4);
  net->wb = slbn_malloc(n * 2);
  slbn_zero(net->m, n * 4);
  slbn_zero(net->v, n * 4);
  for (int L = 0; L < NLAYER; L++)
    net->wt[L] = slbn_malloc((size_t)LCIN[L] * slbn_kpad(LCOUT[L]) * 2);
  net->loss = slbn_malloc(4 * 4096);
  net->wc_eval = slbn_malloc(4 * WN_TARG);
  net->mode = SLB_WARM_MOMENTS;
  net->probe = 200;
  net->probe_res = HUGE_VAL;
}

static void upload_params(slb_warmnet *net, const float *hw) {
  slbn_h2d(net->w, hw, (size_t)net->nparam * 4);
  slbn_f2bf(net->w, net->wb, net->nparam);
  refresh_wt(net);
  slbn_sync();
}

slb_warmnet *slb_warm_new(unsigned long long seed) {
  slb_warmnet *net = calloc(1, sizeof(*net));
  if (!net) die("out of host memory");
  net->rng = seed;
  alloc_params(net);
  float *hw = calloc((size_t)net->nparam, sizeof(float));
  if (!hw) die("out of host memory");
  uint64_t s = seed ^ 0x5EEDull;
  for (int L = 0; L < NLAYER; L++) {
    const int cin = LCIN[L], cout = LCOUT[L], Kp = slbn_kpad(cin);
    const double sd = (L == OUT ? 1.0 : sqrt(2.0)) / sqrt(27.0 * cin);
    for (int co = 0; co < cout; co++)
      for (int k = 0; k < 27 * cin; k++) hw[net->wofs[L] + (long)co * Kp + k] = (float)(sd * nrand(&s));
  }
  upload_params(net, hw);
  free(hw);
  return net;
}

static void free_acts(slb_warmnet *net) {
  slbn_free(net->x0);
  for (int L = 0; L < NLAYER; L++) slbn_free(net->act[L]);
  for (int i = 0; i < 3; i++) { slbn_free(net->pool[i]); slbn_free(net->up[i]); }
  slbn_free(net->outf);
  for (int l = 0; l < NLEVEL; l++) { slbn_free(net->gA[l]); slbn_free(net->gB[l]); slbn_free(net->gP[l]); }
  slbn_free(net->dout);
  slbn_free(net->tbuf);
  net->x0 = NULL;
  memset(net->act, 0, sizeof(net->act));
  memset(net->pool, 0, sizeof(net->pool));
  memset(net->up, 0, sizeof(net->up));
  net->outf = NULL;
  memset(net->gA, 0, sizeof(net->gA));
  memset(net->gB, 0, sizeof(net->gB));
  memset(net->gP, 0, sizeof(net->gP));
  net->dout = NULL;
  net->tbuf = NULL;
  net->dims[0] = net->dims[1] = net->dims[2] = 0;
}

static void check_dims(const int *d) {
  for (int a = 0; a < 3; a++)
    if (d[a] < 16 || d[a] % 8) die("lattice edges must be multiples of 8 and at least 16");
  if ((long)d[0] * d[1] * d[2] * WN_C >= (1L << 31)) die("lattice too large for the network");
}

static void ensure_acts(slb_warmnet *net, const int *d) {
  if (net->dims[0] == d[0] && net->dims[1] == d[1] && net->dims[2] == d[2]) return;
  free_acts(net);
  check_dims(d);
  memcpy(net->dims, d, sizeof(net->dims));
  const long N0 = (long)d[0] * d[1] * d[2];
  for (int l = 0; l < NLEVEL; l++) net->nl[l] = N0 >> (3 * l);
  const size_t b = 2;
  net->x0 = slbn_malloc(N0 * WN_C * b);
  for (int L = 0; L < NLAYER; L++)
    if (L != OUT) net->act[L] = slbn_malloc(net->nl[LLEVEL[L]] * LCOUT[L] * b);
  /* pool[l] at level l+1 with the channels of level l; up[l] at level l */
  const int pc[3] = {16, 32, 64};
  for (int l = 0; l < 3; l++) {
    net->pool[l] = slbn_malloc(net->nl[l + 1] * pc[l] * b);
    net->up[l] = slbn_malloc(net->nl[l] * pc[l] * b);
  }
  net->outf = slbn_malloc(N0 * WN_C * 4);
  for (int l = 0; l < NLEVEL; l++) {
    net->gA[l] = slbn_malloc(net->nl[l] * LEVEL_CMAX[l] * b);
    net->gB[l] = slbn_malloc(net->nl[l] * LEVEL_CMAX[l] * b);
    net->gP[l] = slbn_malloc(net->nl[l] * LEVEL_CMAX[l] * b);
  }
  net->dout = slbn_malloc(N0 * WN_C * b);
  net->tbuf = slbn_malloc(N0 * WN_TARG * b);
}

void slb_warm_free(slb_warmnet *net) {
  if (!net) return;
  free_acts(net);
  slbn_free(net->w); slbn_free(net->g); slbn_free(net->m); slbn_free(net->v); slbn_free(net->wb);
  for (int L = 0; L < NLAYER; L++) slbn_free(net->wt[L]);
  slbn_free(net->loss);
  slbn_free(net->wc_eval);
  for (int i = 0; i < net->nsamp; i++) {
    slbn_free(net->samp[i].x);
    slbn_free(net->samp[i].t);
    slbn_free(net->samp[i].wc);
  }
  free(net->samp);
  free(net->notes);
  free(net->wnotes);
  free(net);
}

/* ------------------------------------------------------------------ */
/* Forward and backward passes over net->x0 (dims net->dims).            */

static const int *lvl_dims(const slb_warmnet *net, int l, int *out) {
  for (int a = 0; a < 3; a++) out[a] = net->dims[a] >> l;
  return out;
}

static uint16_t *layer_in(slb_warmnet *net, int L) {
  switch (L) {
    case E1A: return net->x0;
    case E1B: return net->act[E1A];
    case E2A: return net->pool[0];
    case E2B: return net->act[E2A];
    case E3A: return net->pool[1];
    case E3B: return net->act[E3A];
    case E4A: return net->pool[2];
    case E4B: return net->act[E4A];
    case D3: return net->up[2];
    case D2: return net->up[1];
    case D1: return net->up[0];
    default: return net->act[D1];
  }
}

static void conv_fwd(slb_warmnet *net, int L) {
  int d[3];
  lvl_dims(net, LLEVEL[L], d);
  slbn_conv(layer_in(net, L), LCIN[L], net->wb + net->wofs[L], net->w + net->bofs[L],
            L == OUT ? NULL : net->act[L], L == OUT ? net->outf : NULL, LCOUT[L], L != OUT, 0,
            d[0], d[1], d[2]);
}

static void forward(slb_warmnet *net) {
  int d[3];
  conv_fwd(net, E1A);
  conv_fwd(net, E1B);
  lvl_dims(net, 0, d);
  slbn_pool(net->act[E1B], net->pool[0], 16, d[0], d[1], d[2]);
  conv_fwd(net, E2A);
  conv_fwd(net, E2B);
  lvl_dims(net, 1, d);
  slbn_pool(net->act[E2B], net->pool[1], 32, d[0], d[1], d[2]);
  conv_fwd(net, E3A);
  conv_fwd(net, E3B);
  lvl_dims(net, 2, d);
  slbn_pool(net->act[E3B], net->pool[2], 64, d[0], d[1], d[2]);
  conv_fwd(net, E4A);
  conv_fwd(net, E4B);
  slbn_up_add(net->act[E4B], net->act[E3B], net->up[2], 64, d[0], d[1], d[2]);
  conv_fwd(net, D3);
  lvl_dims(net, 1, d);
  slbn_up_add(net->act[D3], net->act[E2B], net->up[1], 32, d[0], d[1], d[2]);
  conv_fwd(net, D2);
  lvl_dims(net, 0, d);
  slbn_up_add(net->act[D2], net->act[E1B], net->up[0], 16, d[0], d[1], d[2]);
  conv_fwd(net, D1);
  conv_fwd(net, OUT);
}

/* Gradient of layer L given the gradient of its output gy (pre-activation
 * gradient formed in gp when the layer has a ReLU): accumulates dW, db and,
 * when gx != NULL, writes the gradient of the layer input into gx. */
static void conv_bwd(slb_warmnet *net, int L, const uint16_t *gy, uint16_t *gp, uint16_t *gx) {
  int d[3];
  lvl_dims(net, LLEVEL[L], d);
  const long n = net->nl[LLEVEL[L]];
  const uint16_t *dp = gy;
  if (L != OUT) {
    /* ReLU backward and the bias gradient in one pass */
    slbn_relu_bwd_bias(net->act[L], gy, gp, LCOUT[L], (int)n, net->g + net->bofs[L]);
    dp = gp;
  } else {
    slbn_bias_grad(dp, LCOUT[L], (int)n, net->g + net->bofs[L]);
  }
  slbn_conv_wgrad(layer_in(net, L), LCIN[L], dp, LCOUT[L], net->g + net->wofs[L], d[0], d[1], d[2]);
  if (gx)
    slbn_conv(dp, LCOUT[L], net->wt[L], NULL, gx, NULL, LCIN[L], 0, 1, d[0], d[1], d[2]);
}

static void backward(slb_warmnet *net) {
  int d[3];
  uint16_t **A = net->gA, **B = net->gB, **P = net->gP;
  /* level 0, decoder side */
  conv_bwd(net, OUT, net->dout, NULL, A[0]);          /* A0 = d act[D1] */
  conv_bwd(net, D1, A[0], P[0], B[0]);                /* B0 = d up[0] (kept: skip of E1B) */
  lvl_dims(net, 0, d);
  slbn_up_bwd(B[0], A[1], 16, d[0], d[1], d[2]);      /* A1 = d act[D2] */
  conv_bwd(net, D2, A[1], P[1], B[1]);                /* B1 = d up[1] (skip of E2B) */
  lvl_dims(net, 1, d);
  slbn_up_bwd(B[1], A[2], 32, d[0], d[1], d[2]);      /* A2 = d act[D3] */
  conv_bwd(net, D3, A[2], P[2], B[2]);                /* B2 = d up[2] (skip of E3B) */
  lvl_dims(net, 2, d);
  slbn_up_bwd(B[2], A[3], 64, d[0], d[1], d[2]);      /* A3 = d act[E4B] */
  /* bottleneck */
  conv_bwd(net, E4B, A[3], P[3], B[3]);               /* B3 = d act[E4A] */
  conv_bwd(net, E4A, B[3], P[3], A[3]);               /* A3 = d pool[2] */
  /* level 2 encoder */
  slbn_pool_bwd(A[3], A[2], 64, d[0], d[1], d[2]);    /* A2 = d act[E3B] via pool */
  slbn_add_into(A[2], B[2], net->nl[2] * 64);
  conv_bwd(net, E3B, A[2], P[2], B[2]);               /* B2 = d act[E3A] */
  conv_bwd(net, E3A, B[2], P[2], A[2]);               /* A2 = d pool[1] (32 ch) */
  lvl_dims(net, 1, d);
  slbn_pool_bwd(A[2], A[1], 32, d[0], d[1], d[2]);
  slbn_add_into(A[1], B[1], net->nl[1] * 32);
  conv_bwd(net, E2B, A[1], P[1], B[1]);
  conv_bwd(net, E2A, B[1], P[1], A[1]);               /* A1 = d pool[0] (16 ch) */
  lvl_dims(net, 0, d);
  slbn_pool_bwd(A[1], A[0], 16, d[0], d[1], d[2]);
  slbn_add_into(A[0], B[0], net->nl[0] * 16);
  conv_bwd(net, E1B, A[0], P[0], B[0]);
  conv_bwd(net, E1A, B[0], P[0], NULL);
}

/* ------------------------------------------------------------------ */
/* Lattice side.                                                         */

/* group scales of the targets for this lattice (per unit force and
 * normalising length ell) */
static void target_scales(LBSys lb, const slbd_params *P, double ell, float sc[3]) {
  const double F = sqrt(P->bf[0] * P->bf[0] + P->bf[1] * P->bf[1] + P->bf[2] * P->bf[2]);
  const double eta = lb->rho0 * (lb->tau - 0.5) / 3.0;
  if (!(F > 0.0) || P->bf_mode == 2) die("the warm start needs a uniform, non-zero body force");
  sc[0] = (float)(eta / (F * ell * ell) * WN_SU);
  sc[1] = (float)(WN_SR / (F * ell));
  sc[2] = (float)(WN_SS / (F * ell));
}

/* features into net->x0; returns ell */
static double lattice_features(slb_warmnet *net, LBSys lb, slbd_params *P) {
  slb_params(lb, P);
  slb_dev_geom(lb);
  int d[3] = {lb->esi, lb->esj, lb->esk};
  ensure_acts(net, d);
  slbn_sync();
  return slbd_features(PRIV(lb)->dev, net->x0, WN_C);
}

/* Mean square of each target group (host pass over a device tensor) and
 * the loss weights that make every group's relative mean-square error
 * count equally: wc[c] = 1 / (3 * nodes * |group| * ms[group]). */
static void target_weights(const uint16_t *t_dev, long N, double ms[3], float wc[WN_TARG]) {
  uint16_t *h = malloc((size_t)N * WN_TARG * 2);
  if (!h) die("out of host memory");
  slbn_d2h(h, t_dev, (size_t)N * WN_TARG * 2);
  double ss[WN_TARG] = {0};
  for (long n = 0; n < N; n++)
    for (int c = 0; c < WN_TARG; c++) {
      union { uint32_t u; float f; } cv;
      cv.u = (uint32_t)h[n * WN_TARG + c] << 16;
      ss[c] += (double)cv.f * cv.f;
    }
  free(h);
  int cnt[3] = {0, 0, 0};
  ms[0] = ms[1] = ms[2] = 0.0;
  for (int c = 0; c < WN_TARG; c++) { ms[WN_GROUP[c]] += ss[c]; cnt[WN_GROUP[c]]++; }
  for (int g = 0; g < 3; g++) ms[g] /= (double)N * cnt[g];
  for (int c = 0; c < WN_TARG; c++) {
    const int g = WN_GROUP[c];
    wc[c] = (float)(1.0 / (3.0 * (double)N * cnt[g] * (ms[g] > 0.0 ? ms[g] : 1.0)));
  }
}

int slb_warm_add_sample(slb_warmnet *net, LBSys lb) {
  slbd_params P;
  slb_params(lb, &P);
  slb_dev_geom(lb);
  slb_dev_pops(lb);
  int d[3] = {lb->esi, lb->esj, lb->esk};
  check_dims(d);
  const long N = (long)d[0] * d[1] * d[2];
  if (net->nsamp == net->cap) {
    net->cap = net->cap ? 2 * net->cap : 64;
    net->samp = realloc(net->samp, (size_t)net->cap * sizeof(wn_sample));
    if (!net->samp) die("out of host memory");
  }
  wn_sample *s = &net->samp[net->nsamp];
  memcpy(s->dims, d, sizeof(d));
  s->nodes = N;
  s->x = slbn_malloc(N * WN_FEAT * 2);
  s->t = slbn_malloc(N * WN_TARG * 2);
  s->wc = slbn_malloc(4 * WN_TARG);
  float sc[3], wc[WN_TARG];
  slbn_sync();
  s->ell = slbd_features(PRIV(lb)->dev, s->x, WN_FEAT);
  target_scales(lb, &P, s->ell, sc);
  slbd_targets(PRIV(lb)->dev, &P, sc, s->t, WN_TARG);
  target_weights(s->t, N, s->ms, wc);
  slbn_h2d(s->wc, wc, sizeof(wc));
  return ++net->nsamp;
}

void slb_warm_sample_stats(const slb_warmnet *net, int i, double out[4]) {
  if (i < 0 || i >= net->nsamp) die("sample index out of range");
  memcpy(out, net->samp[i].ms, sizeof(double) * 3);
  out[3] = net->samp[i].ell;
}

int slb_warm_nsamples(const slb_warmnet *net) { return net->nsamp; }

/* populations from scaled moments m[n][channels] (device FP32) */
static void inject(LBSys lb, const slbd_params *P, const float sc[3], const float *m, int channels,
                   int mode) {
  slb_lbpriv *S = PRIV(lb);
  const double inv[3] = {1.0 / sc[0], 1.0 / sc[1], 1.0 / sc[2]};
  slbd_inject(S->dev, P, m, channels, inv, mode == SLB_WARM_EQUILIBRIUM ? 0 : 1);
  S->pops_loc = LOC_DEV;
  if (mode == SLB_WARM_EQUILIBRIUM) {
    /* lbloadvel: the rebuilt state is post-propagate, then collide */
    slbd_collide(S->dev, P);
    slb_after_collide(lb);
  } else {
    lb->status = POST_COLLIDE;
  }
}

double slb_warm_features(LBSys lb, void *x, int channels) {
  slbd_params P;
  slb_params(lb, &P);
  slb_dev_geom(lb);
  slbn_sync();
  return slbd_features(PRIV(lb)->dev, (uint16_t *)x, channels);
}

double slb_warm_features_noalloc(LBSys lb, void *x, int channels) {
  slbd_params P;
  slb_params(lb, &P);
  slb_dev_geom(lb);
  slbn_sync();
  return slbd_features_noalloc(PRIV(lb)->dev, (uint16_t *)x, channels);
}

void slb_warm_moments(LBSys lb, float *m, int channels, double ell) {
  slbd_params P;
  float sc[3];
  slb_params(lb, &P);
  slb_dev_geom(lb);
  slb_dev_pops(lb);
  target_scales(lb, &P, ell, sc);
  slbn_sync();
  slbd_moments(PRIV(lb)->dev, &P, sc, m, channels);
}

void slb_warm_inject(LBSys lb, const float *m, int channels, int mode, double ell) {
  slbd_params P;
  float sc[3];
  if (channels < WN_TARG) die("slb_warm_inject needs at least 10 channels");
  if (mode != SLB_WARM_EQUILIBRIUM && mode != SLB_WARM_MOMENTS) die("unknown warm-start mode");
  slb_params(lb, &P);
  slb_dev_geom(lb);
  target_scales(lb, &P, ell, sc);
  slbn_sync();
  inject(lb, &P, sc, m, channels, mode);
}

void slb_warm_targets(LBSys lb, void *t, int channels, double ell) {
  slbd_params P;
  float sc[3];
  slb_params(lb, &P);
  slb_dev_geom(lb);
  slb_dev_pops(lb);
  target_scales(lb, &P, ell, sc);
  slbn_sync();
  slbd_targets(PRIV(lb)->dev, &P, sc, (uint16_t *)t, channels);
}

/* One optimisation step on sample s; the squared error goes to loss[slot]. */
static void train_step(slb_warmnet *net, const wn_sample *s, double lr, int slot) {
  ensure_acts(net, s->dims);
  slbn_zero(net->g, (size_t)net->nparam * 4);
  slbn_expand(s->x, WN_FEAT, net->x0, WN_C, s->nodes);
  forward(net);
  slbn_mse(net->outf, WN_C, s->t, WN_TARG, WN_TARG, (int)s->nodes, s->wc, net->dout,
           net->loss + slot);
  backward(net);
  net->adam_step++;
  slbn_adam(net->w, net->wb, net->g, net->m, net->v, net->nparam, (float)lr, 0.9f, 0.999f, 1e-8f,
            net->adam_step);
  refresh_wt(net);
}

double slb_warm_train(slb_warmnet *net, int iters, double lr, unsigned long long seed,
                      double *history) {
  if (net->nsamp == 0) die("no training samples");
  if (iters <= 0) return 0.0;
  uint64_t s = seed;
  int *perm = malloc(sizeof(int) * net->nsamp);
  float *hl = malloc(sizeof(float) * 4096);
  double *rel = malloc(sizeof(double) * iters);
  if (!perm || !hl || !rel) die("out of host memory");
  int pos = net->nsamp;
  for (int it0 = 0; it0 < iters; it0 += 4096) {
    const int nb = (iters - it0 < 4096) ? iters - it0 : 4096;
    slbn_zero(net->loss, 4 * 4096);
    for (int b = 0; b < nb; b++) {
      if (pos == net->nsamp) {
        for (int i = 0; i < net->nsamp; i++) perm[i] = i;
        for (int i = net->nsamp - 1; i > 0; i--) {
          const int j = (int)(splitmix(&s) % (uint64_t)(i + 1));
          const int t = perm[i]; perm[i] = perm[j]; perm[j] = t;
        }
        pos = 0;
      }
      const int it = it0 + b;
      /* linear warm-up over the first 5% of the steps, then cosine decay
       * to lr/20 */
      const double wu = 0.05 * iters;
      double f = 0.05 + 0.95 * 0.5 * (1.0 + cos(3.141592653589793 * it / (double)iters));
      if (it < wu) f *= (it + 1) / wu;
      train_step(net, &net->samp[perm[pos++]], lr * f, b);
    }
    slbn_d2h(hl, net->loss, 4 * nb);
    for (int b = 0; b < nb; b++) rel[it0 + b] = hl[b];
  }
  if (history)
    for (int i = 0; i < iters; i++) history[i] = rel[i];
  const int tail = iters < net->nsamp ? iters : net->nsamp;
  double m = 0.0;
  for (int i = iters - tail; i < iters; i++) m += rel[i];
  free(perm); free(hl); free(rel);
  return m / tail;
}

/* Relative mean-square error of the prediction against the lattice's
 * current state (the mean over the three target groups of each group's
 * squared error over its mean square); group errors to grp when non-NULL. */
double slb_warm_eval(slb_warmnet *net, LBSys lb, double *grp) {
  slbd_params P;
  float sc[3], wc[WN_TARG];
  double ms[3];
  slb_dev_pops(lb);
  const double ell = lattice_features(net, lb, &P);
  target_scales(lb, &P, ell, sc);
  forward(net);
  slbn_sync();
  const long N = net->nl[0];
  slbd_targets(PRIV(lb)->dev, &P, sc, net->tbuf, WN_TARG);
  target_weights(net->tbuf, N, ms, wc);
  double tot = 0.0;
  for (int g = 0; g < 3; g++) {
    float wg[WN_TARG];
    for (int c = 0; c < WN_TARG; c++) wg[c] = WN_GROUP[c] == g ? 3.f * wc[c] : 0.f;
    slbn_h2d(net->wc_eval, wg, sizeof(wg));
    slbn_zero(net->loss, 4);
    slbn_mse(net->outf, WN_C, net->tbuf, WN_TARG, WN_TARG, (int)N, net->wc_eval, net->dout,
             net->loss);
    float se;
    slbn_d2h(&se, net->loss, 4);
    if (grp) grp[g] = se;
    tot += se / 3.0;
  }
  return tot;
}

void slb_warm_set_mode(slb_warmnet *net, int mode) {
  if (mode != SLB_WARM_EQUILIBRIUM && mode != SLB_WARM_MOMENTS) die("unknown warm-start mode");
  net->mode = mode;
}
int slb_warm_get_mode(const slb_warmnet *net) { return net->mode; }

void slb_warm_start(slb_warmnet *net, LBSys lb) {
  slbd_params P;
  float sc[3];
  const double ell = lattice_features(net, lb, &P);
  target_scales(lb, &P, ell, sc);
  forward(net);
  slbn_sync();
  if (net->proj_iters > 0 && net->proj_passes > 0)
    slbd_project(PRIV(lb)->dev, net->outf, WN_C, net->proj_iters, net->proj_passes);
  inject(lb, &P, sc, net->outf, WN_C, net->mode);
}

void slb_warm_set_projection(slb_warmnet *net, int iters, int passes) {
  net->proj_iters = iters > 0 ? iters : 0;
  net->proj_passes = passes > 0 ? passes : 0;
}
void slb_warm_get_projection(const slb_warmnet *net, int *iters, int *passes) {
  if (iters) *iters = net->proj_iters;
  if (passes) *passes = net->proj_passes;
}

double slb_warm_project(LBSys lb, float *m, int channels, int iters, int passes) {
  slb_dev_geom(lb);
  slbn_sync();
  return slbd_project(PRIV(lb)->dev, m, channels, iters, passes);
}

void slb_warm_predict(slb_warmnet *net, LBSys lb, float *m) {
  slbd_params P;
  float sc[3];
  const double ell = lattice_features(net, lb, &P);
  target_scales(lb, &P, ell, sc);
  forward(net);
  const long N = net->nl[0];
  float *h = malloc(N * WN_C * 4);
  if (!h) die("out of host memory");
  slbn_d2h(h, net->outf, N * WN_C * 4);
  for (long n = 0; n < N; n++)
    for (int c = 0; c < WN_TARG; c++)
      m[WN_TARG * n + c] = lb->obj[n] != NULL ? 0.f : h[n * WN_C + c] / sc[WN_GROUP[c]];
  free(h);
}

/* ------------------------------------------------------------------ */
/* Solve with the warm start and the cold-start fallback.                */

void slb_warm_set_fallback(slb_warmnet *net, int probe, double probe_res) {
  net->probe = probe;
  net->probe_res = probe_res;
  net->ncurve = 0;
  net->nwcurve = 0;
}

void slb_warm_get_fallback(const slb_warmnet *net, int *probe, double *probe_res) {
  if (probe) *probe = net->probe;
  if (probe_res) *probe_res = net->probe_res;
}

/* Cold reference curve: the median over the recorded cold solves of the
 * residual after k * interval steps (a finished solve counts as its last
 * residual from then on). */
static double median(double *v, int n) {
  for (int i = 1; i < n; i++)
    for (int j = i; j > 0 && v[j - 1] > v[j]; j--) { const double x = v[j]; v[j] = v[j - 1]; v[j - 1] = x; }
  return v[n / 2];
}

/* Append one residual history (res[k] after (k+1)*interval steps, n
 * entries) to a set of runs and recompute the set's median curve; a set
 * recorded with another interval starts again. */
static void add_run(int *count, int *set_interval, double **runs, double *curve,
                    const double *res, int n, int interval) {
  if (*count == 0 || interval != *set_interval) {
    *count = 0;
    *set_interval = interval;
  }
  double *nn = realloc(*runs, sizeof(double) * (size_t)(*count + 1) * WN_CURVE);
  if (!nn) die("out of host memory");
  *runs = nn;
  double *row = nn + (size_t)*count * WN_CURVE;
  for (int k = 0; k < WN_CURVE; k++) row[k] = res[k < n ? k : n - 1];
  (*count)++;
  double *t = malloc(sizeof(double) * *count);
  if (!t) die("out of host memory");
  for (int k = 0; k < WN_CURVE; k++) {
    for (int i = 0; i < *count; i++) t[i] = nn[(size_t)i * WN_CURVE + k];
    curve[k] = median(t, *count);
  }
  free(t);
}

void slb_warm_note_cold(slb_warmnet *net, const double *res, int n, int interval) {
  if (n < 1) return;
  add_run(&net->nnote, &net->note_interval, &net->notes, net->curve, res, n, interval);
  net->ncurve = WN_CURVE;
  net->curve_interval = interval;
  net->probe = 2 * interval;
  net->probe_res = net->curve[1];
}

void slb_warm_note_warm(slb_warmnet *net, const double *res, int n, int interval) {
  if (n < 1) return;
  add_run(&net->nwnote, &net->wnote_interval, &net->wnotes, net->wcurve, res, n, interval);
  net->nwcurve = WN_CURVE;
  net->wcurve_interval = interval;
}

/* Steps a median curve takes to fall below tol (HUGE_VAL if it never does). */
static double curve_steps(const double *curve, int ncurve, int interval, double tol) {
  for (int k = 0; k < ncurve; k++)
    if (curve[k] < tol) return (double)(k + 1) * interval;
  return HUGE_VAL;
}

int slb_warm_policy(const slb_warmnet *net, double tol, int interval) {
  if (net->nwcurve == 0 || net->ncurve == 0 || interval != net->wcurve_interval ||
      interval != net->curve_interval)
    return 1;
  const double w = curve_steps(net->wcurve, net->nwcurve, interval, tol);
  const double c = curve_steps(net->curve, net->ncurve, interval, tol);
  return w < c;
}

int slb_warm_solve(slb_warmnet *net, LBSys lb, double tol, int interval, int maxsteps, double *res,
                   int *used_warm) {
  if (interval < 1) interval = 1;
  double r = -1.0;
  int steps = 0, warm = slb_warm_policy(net, tol, interval), done = 0;
  if (warm) {
    slb_warm_start(net, lb);
  } else {
    /* the calibration says a warm start does not reach tol sooner: start
     * from rest, as the fallback does */
    slb_lbpriv *S = PRIV(lb);
    slbd_zero_pops(S->dev);
    S->pops_loc = LOC_DEV;
  }
  slb_residual_reset(lb);
  slb_residual(lb);
  /* cold steps to tol on the reference curve */
  const int use_curve = net->ncurve > 0 && interval == net->curve_interval;
  double cold_steps = HUGE_VAL;
  if (use_curve) {
    int k = 0;
    while (k < net->ncurve && !(net->curve[k] < tol)) k++;
    cold_steps = (double)(k + 1) * interval;
  }
  double rprev = -1.0;
  while (steps < maxsteps) {
    int n = interval;
    if (steps + n > maxsteps) n = maxsteps - steps;
    slb_run(lb, n);
    steps += n;
    r = slb_residual(lb);
    if (r >= 0.0 && r < tol) { done = 1; break; }
    if (!warm) continue;
    /* The warm residual must fall faster than a cold start: with the
     * recorded curve, from the second check on, the steps projected from
     * the last decay ratio must not exceed the cold steps to tol; without
     * it, the residual must be below probe_res from step `probe` on. */
    int fall_back = 0;
    if (use_curve) {
      if (rprev > 0.0 && steps >= 2 * interval) {
        const double q = r / rprev;
        const double proj = q < 1.0 ? steps + interval * log(tol / r) / log(q) : HUGE_VAL;
        fall_back = proj > cold_steps;
      }
    } else if (net->probe > 0 && steps >= net->probe) {
      fall_back = !(r >= 0.0 && r < net->probe_res);
    }
    rprev = r;
    if (fall_back) {
      /* not falling faster than a cold start: restart from rest */
      warm = 0;
      slb_lbpriv *S = PRIV(lb);
      slbd_zero_pops(S->dev);
      S->pops_loc = LOC_DEV;
      slb_residual_reset(lb);
      slb_residual(lb);
    }
  }
  (void)done;
  if (res) *res = r;
  if (used_warm) *used_warm = warm;
  return steps;
}

/* ------------------------------------------------------------------ */
/* Weights file: magic, layer table, fallback, FP32 weights.             */

int slb_warm_save(const slb_warmnet *net, const char *path) {
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  float *hw = malloc(sizeof(float) * net->nparam);
  if (!hw) { fclose(f); return -1; }
  slbn_d2h(hw, net->w, sizeof(float) * net->nparam);
  int hdr[2 * NLAYER + 2];
  hdr[0] = NLAYER;
  for (int L = 0; L < NLAYER; L++) { hdr[1 + 2 * L] = LCIN[L]; hdr[2 + 2 * L] = LCOUT[L]; }
  hdr[2 * NLAYER + 1] = net->probe;
  int ok = fwrite(WN_MAGIC, 1, 8, f) == 8 && fwrite(hdr, sizeof(hdr), 1, f) == 1 &&
           fwrite(&net->probe_res, sizeof(double), 1, f) == 1 &&
           fwrite(&net->ncurve, sizeof(int), 1, f) == 1 &&
           fwrite(&net->curve_interval, sizeof(int), 1, f) == 1 &&
           fwrite(net->curve, sizeof(double), WN_CURVE, f) == WN_CURVE &&
           fwrite(&net->nwcurve, sizeof(int), 1, f) == 1 &&
           fwrite(&net->wcurve_interval, sizeof(int), 1, f) == 1 &&
           fwrite(net->wcurve, sizeof(double), WN_CURVE, f) == WN_CURVE &&
           fwrite(&net->nparam, sizeof(long), 1, f) == 1 &&
           fwrite(hw, sizeof(float), net->nparam, f) == (size_t)net->nparam;
  free(hw);
  if (fclose(f) != 0) ok = 0;
  return ok ? 0 : -1;
}

slb_warmnet *slb_warm_load(const char *path) {
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  char magic[8];
  int hdr[2 * NLAYER + 2];
  double pr;
  long np;
  slb_warmnet tmp;
  memset(&tmp, 0, sizeof(tmp));
  layout(&tmp);
  int ncurve = 0, cint = 0, nwcurve = 0, wcint = 0;
  double curve[WN_CURVE], wcurve[WN_CURVE];
  int ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, WN_MAGIC, 8) == 0 &&
           fread(hdr, sizeof(hdr), 1, f) == 1 && hdr[0] == NLAYER &&
           fread(&pr, sizeof(double), 1, f) == 1 && fread(&ncurve, sizeof(int), 1, f) == 1 &&
           fread(&cint, sizeof(int), 1, f) == 1 &&
           fread(curve, sizeof(double), WN_CURVE, f) == WN_CURVE &&
           fread(&nwcurve, sizeof(int), 1, f) == 1 && fread(&wcint, sizeof(int), 1, f) == 1 &&
           fread(wcurve, sizeof(double), WN_CURVE, f) == WN_CURVE &&
           fread(&np, sizeof(long), 1, f) == 1 && np == tmp.nparam;
  for (int L = 0; ok && L < NLAYER; L++)
    if (hdr[1 + 2 * L] != LCIN[L] || hdr[2 + 2 * L] != LCOUT[L]) ok = 0;
  /* curve lengths index the fixed-size curve tables; intervals are step
   * counts */
  if (ok && (ncurve < 0 || ncurve > WN_CURVE || nwcurve < 0 || nwcurve > WN_CURVE ||
             cint < 0 || wcint < 0 || hdr[2 * NLAYER + 1] < 0))
    ok = 0;
  float *hw = ok ? malloc(sizeof(float) * np) : NULL;
  if (ok && (!hw || fread(hw, sizeof(float), np, f) != (size_t)np)) ok = 0;
  fclose(f);
  if (!ok) { free(hw); return NULL; }
  slb_warmnet *net = calloc(1, sizeof(*net));
  if (!net) die("out of host memory");
  alloc_params(net);
  upload_params(net, hw);
  free(hw);
  net->probe = hdr[2 * NLAYER + 1];
  net->probe_res = pr;
  net->ncurve = ncurve;
  net->curve_interval = cint;
  memcpy(net->curve, curve, sizeof(curve));
  net->nwcurve = nwcurve;
  net->wcurve_interval = wcint;
  memcpy(net->wcurve, wcurve, sizeof(wcurve));
  return net;
}

/* ------------------------------------------------------------------ */
/* Parameter access (tests: gradients against a CPU network).             */

long slb_warm_nparams(const slb_warmnet *net) { return net->nparam; }

void slb_warm_layout(int *nlayer, int *cin, int *cout, long *wofs, long *bofs) {
  slb_warmnet tmp;
  memset(&tmp, 0, sizeof(tmp));
  layout(&tmp);
  if (nlayer) *nlayer = NLAYER;
  for (int L = 0; L < NLAYER; L++) {
    if (cin) cin[L] = LCIN[L];
    if (cout) cout[L] = LCOUT[L];
    if (wofs) wofs[L] = tmp.wofs[L];
    if (bofs) bofs[L] = tmp.bofs[L];
  }
}

void slb_warm_get_params(const slb_warmnet *net, float *w) {
  slbn_d2h(w, net->w, sizeof(float) * net->nparam);
}
void slb_warm_set_params(slb_warmnet *net, const float *w) { upload_params(net, w); }

/* Forward and backward on a caller's device input x (dims d, WN_C
 * channels, BF16) and target t (WN_TARG channels, BF16) with host loss
 * weights wc[WN_TARG]: writes the FP32 output (host, [n][WN_C]) to outf when non-NULL,
 * the parameter gradient (host, nparam floats) to grad when non-NULL, and
 * returns the summed squared error. */
double slb_warm_fwd_bwd(slb_warmnet *net, const int *d, const uint16_t *x, const uint16_t *t,
                        const float *wc, float *outf, float *grad) {
  ensure_acts(net, d);
  const long N = net->nl[0];
  slbn_sync();
  slbn_zero(net->g, (size_t)net->nparam * 4);
  slbn_zero(net->loss, 4);
  /* device-to-device copy of the input through the expand kernel */
  slbn_expand(x, WN_C, net->x0, WN_C, N);
  forward(net);
  slbn_h2d(net->wc_eval, wc, sizeof(float) * WN_TARG);
  slbn_mse(net->outf, WN_C, t, WN_TARG, WN_TARG, (int)N, net->wc_eval, net->dout, net->loss);
  backward(net);
  float se;
  slbn_d2h(&se, net->loss, 4);
  if (outf) slbn_d2h(outf, net->outf, N * WN_C * 4);
  if (grad) slbn_d2h(grad, net->g, sizeof(float) * net->nparam);
  return se;
}
