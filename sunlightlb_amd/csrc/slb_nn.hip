/*
 * slb_nn.hip -- gfx950 kernels of the learned warm start: a 3D convolutional
 * encoder-decoder over the lattice, trained and evaluated on the GPU.
 *
 * Tensors are channels-last BF16, X[n][c] with n the lattice node index
 * (esk*(esj*i+j)+k) and C a multiple of 8, so the 8 channels one MFMA lane
 * needs are one 16-byte load.  All convolutions are 3x3x3 with periodic
 * padding (the lattice is periodic) and are implicit GEMMs on the BF16 MFMA
 * v_mfma_f32_16x16x32_bf16 with FP32 accumulation:
 *
 *   forward      M = Cout, N = 16 consecutive nodes, K = 27*Cin (k = tap*Cin + ci)
 *   grad input   M = Cin,  N = 16 consecutive nodes, K = 27*Cout (transposed taps)
 *   grad weight  M = Cout, N = Cin (one tap), K = nodes, staged through LDS
 *
 * MFMA operand layout (checked on gfx950): lane l holds A[l%16][(l/16)*8+e]
 * and B[(l/16)*8+e][l%16] (e = 0..7) and receives D[(l/16)*4+r][l%16]
 * (r = 0..3).  Weights are W[co][Kp] with Kp = 27*Cin rounded up to 32 and
 * zero-filled.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "sunlightlb_nn.h"

#define HIPCHECK(call)                                                        \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      fprintf(stderr, "sunlightlb: HIP error '%s' (%d) at %s:%d in %s\n",     \
              hipGetErrorString(e_), (int)e_, __FILE__, __LINE__, #call);     \
      fflush(stderr);                                                         \
      exit(EXIT_FAILURE);                                                     \
    }                                                                         \
  } while (0)
#define LAUNCHCHECK() HIPCHECK(hipGetLastError())

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;
__device__ __forceinline__ bf16x8 as_bf(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ float bf2f(uint16_t h) {
  return __uint_as_float(((uint32_t)h) << 16);
}
/* round to nearest even */
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  /* NaN */
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

struct NGeom {
  int D, H, Wd;     /* i, j, k extents */
  int plane;        /* H*Wd */
  int N;            /* nodes */
};

/* node n shifted by tap t (dz,dy,dx) = (t/9-1, (t/3)%3-1, t%3-1), times sgn,
 * periodic */
__device__ __forceinline__ int tap_node(const NGeom &G, int z, int y, int x, int t, int sgn) {
  int dz = (t / 9 - 1) * sgn, dy = ((t / 3) % 3 - 1) * sgn, dx = (t % 3 - 1) * sgn;
  int zz = z + dz, yy = y + dy, xx = x + dx;
  zz += (zz < 0) ? G.D : 0; zz -= (zz >= G.D) ? G.D : 0;
  yy += (yy < 0) ? G.H : 0; yy -= (yy >= G.H) ? G.H : 0;
  xx += (xx < 0) ? G.Wd : 0; xx -= (xx >= G.Wd) ? G.Wd : 0;
  return (zz * G.H + yy) * G.Wd + xx;
}

/* ---------------------------------------------------------------------- */
/* Forward / grad-input convolution (same kernel, tap direction sgn).       */
/* Y[n][co] = act( bias[co] + sum_k W[co][k] * X[n + sgn*off(tap(k))][ci(k)] ) */
/* Block: 256 threads = 4 waves; wave w handles NTW tiles of 16 consecutive   */
/* nodes and all output channels (MT = Cout/16 tiles); each K chunk's         */
/* weights are loaded once and applied to all NTW tiles, whose NTW input     */
/* loads are in flight together.                                            */

template <int MT, int NTW>
__global__ void __launch_bounds__(256)
k_conv(const uint16_t *__restrict__ X, int Cin, const uint16_t *__restrict__ W, int Kp,
       const float *__restrict__ bias, uint16_t *__restrict__ Y, float *__restrict__ Yf,
       int Cout, int relu, int sgn, NGeom G) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n0 = (blockIdx.x * 4 + wv) * 16 * NTW;
  if (n0 >= G.N) return;
  const int grp = lane >> 4;
  int zc[NTW], yc[NTW], xc[NTW];
#pragma unroll
  for (int u = 0; u < NTW; u++) {
    int nn = n0 + 16 * u + (lane & 15);
    if (nn >= G.N) nn = n0;
    zc[u] = nn / G.plane;
    const int rem = nn - zc[u] * G.plane;
    yc[u] = rem / G.Wd;
    xc[u] = rem - yc[u] * G.Wd;
  }
  const int K = 27 * Cin;
  f32x4 acc[NTW][MT];
#pragma unroll
  for (int u = 0; u < NTW; u++)
#pragma unroll
    for (int m = 0; m < MT; m++) acc[u][m] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < Kp; k0 += 32) {
    const int kk = k0 + 8 * grp;
    /* the weights of this K chunk, loaded once for all NTW node tiles */
    bf16x8 a[MT];
#pragma unroll
    for (int m = 0; m < MT; m++) {
      const int co = 16 * m + (lane & 15);
      a[m] = (bf16x8){0, 0, 0, 0, 0, 0, 0, 0};
      if (co < Cout) a[m] = *(const bf16x8 *)(W + (long)co * Kp + kk);
    }
    bf16x8 b[NTW];
    if (kk < K) {
      const int t = kk / Cin, ci = kk - t * Cin;
#pragma unroll
      for (int u = 0; u < NTW; u++) {
        const int src = tap_node(G, zc[u], yc[u], xc[u], t, sgn);
        b[u] = *(const bf16x8 *)(X + (long)src * Cin + ci);
      }
    } else {
#pragma unroll
      for (int u = 0; u < NTW; u++) b[u] = (bf16x8){0, 0, 0, 0, 0, 0, 0, 0};
    }
#pragma unroll
    for (int u = 0; u < NTW; u++)
#pragma unroll
      for (int m = 0; m < MT; m++)
        acc[u][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[u], acc[u][m], 0, 0, 0);
  }
  /* lane holds D[co = 16m + 4*grp + r][node n0 + 16u + lane%16] */
#pragma unroll
  for (int u = 0; u < NTW; u++) {
    const int n = n0 + 16 * u + (lane & 15);
    if (n >= G.N) continue;
#pragma unroll
    for (int m = 0; m < MT; m++) {
      const int cb = 16 * m + 4 * grp;
      if (cb >= Cout) continue;
      float v[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        v[r] = acc[u][m][r] + (bias ? bias[cb + r] : 0.f);
        if (relu) v[r] = v[r] > 0.f ? v[r] : 0.f;
      }
      if (Y) {
        uint2 pk;
        pk.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
        pk.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
        *(uint2 *)(Y + (long)n * Cout + cb) = pk;
      }
      if (Yf) {
#pragma unroll
        for (int r = 0; r < 4; r++) Yf[(long)n * Cout + cb + r] = v[r];
      }
    }
  }
}

/* The same convolution for CIN = 16, 32 or 64 input channels with the input
 * brick staged in LDS: a block computes XT = 1024/CIN x-nodes x 4 rows of
 * ZT planes (wave w: row w, XT/16 tiles of 16 nodes per plane) and first
 * loads the (XT+2) x 6 x (ZT+2) halo brick (38-41 KB for one plane, 51-55 KB
 * for two; periodic wrap applied at load) once,
 * so the 27 taps read the input from LDS instead of 27 times through L2.
 * The K loop, MFMA order and epilogue are those of k_conv: results are
 * bit-identical.  Needs Wd % XT == 0 and H % 4 == 0. */
#define CL_TY 6
/* ZT = 2: the block computes planes z0 and z0 + 1 from one (XT+2) x 6 x 4 brick
 * (the input of the four planes z0 - 1 .. z0 + 2 staged once for both, 3x the
 * output nodes instead of 4.5x), each chunk's weights applied to both planes;
 * every output tile keeps k_conv's K order and MFMA sequence.  Needs D even. */
template <int CIN, int MT, int ZT>
__global__ void __launch_bounds__(256)
k_conv_lds(const uint16_t *__restrict__ X, const uint16_t *__restrict__ W, int Kp,
           const float *__restrict__ bias, uint16_t *__restrict__ Y, float *__restrict__ Yf,
           int Cout, int relu, int sgn, NGeom G) {
  constexpr int XT = 1024 / CIN, NTW = XT / 16, TX = XT + 2, CH = CIN / 8, TZ = ZT + 2;
  __shared__ __attribute__((aligned(16))) uint16_t sX[TZ * CL_TY * TX * CIN];
  const int nbx = G.Wd / XT, nby = G.H >> 2;
  const int bx = blockIdx.x % nbx, rest = blockIdx.x / nbx;
  const int by = rest % nby, z = (rest / nby) * ZT;
  const int x0 = bx * XT, y0 = by * 4;
  for (int e = threadIdx.x; e < TZ * CL_TY * TX * CH; e += blockDim.x) {
    const int part = e % CH, node = e / CH;
    const int lx = node % TX, r = node / TX, ly = r % CL_TY, lz = r / CL_TY;
    int gx = x0 + lx - 1, gy = y0 + ly - 1, gz = z + lz - 1;
    gx += (gx < 0) ? G.Wd : 0; gx -= (gx >= G.Wd) ? G.Wd : 0;
    gy += (gy < 0) ? G.H : 0; gy -= (gy >= G.H) ? G.H : 0;
    gz += (gz < 0) ? G.D : 0; gz -= (gz >= G.D) ? G.D : 0;
    const long src = ((long)(gz * G.H + gy) * G.Wd + gx) * CIN + part * 8;
    *(u32x4 *)(sX + node * CIN + part * 8) = *(const u32x4 *)(X + src);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, grp = lane >> 4, nl = lane & 15;
  constexpr int K = 27 * CIN;
  f32x4 acc[ZT][NTW][MT];
#pragma unroll
  for (int zp = 0; zp < ZT; zp++)
#pragma unroll
    for (int u = 0; u < NTW; u++)
#pragma unroll
      for (int m = 0; m < MT; m++) acc[zp][u][m] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < Kp; k0 += 32) {
    const int kk = k0 + 8 * grp;
    bf16x8 a[MT];
#pragma unroll
    for (int m = 0; m < MT; m++) {
      const int co = 16 * m + nl;
      a[m] = (bf16x8){0, 0, 0, 0, 0, 0, 0, 0};
      if (co < Cout) a[m] = *(const bf16x8 *)(W + (long)co * Kp + kk);
    }
    bf16x8 b[ZT][NTW];
    if (kk < K) {
      const int t = kk / CIN, ci = kk % CIN;
      const int dz = (t / 9 - 1) * sgn, dy = ((t / 3) % 3 - 1) * sgn, dx = (t % 3 - 1) * sgn;
#pragma unroll
      for (int zp = 0; zp < ZT; zp++) {
        const int base = ((1 + dz + zp) * CL_TY + (wv + 1 + dy)) * TX + 1 + dx + nl;
#pragma unroll
        for (int u = 0; u < NTW; u++)
          b[zp][u] = as_bf(*(const u32x4 *)(sX + (base + 16 * u) * CIN + ci));
      }
    } else {
#pragma unroll
      for (int zp = 0; zp < ZT; zp++)
#pragma unroll
        for (int u = 0; u < NTW; u++) b[zp][u] = (bf16x8){0, 0, 0, 0, 0, 0, 0, 0};
    }
#pragma unroll
    for (int zp = 0; zp < ZT; zp++)
#pragma unroll
      for (int u = 0; u < NTW; u++)
#pragma unroll
        for (int m = 0; m < MT; m++)
          acc[zp][u][m] =
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[zp][u], acc[zp][u][m], 0, 0, 0);
  }
#pragma unroll
  for (int zp = 0; zp < ZT; zp++)
#pragma unroll
  for (int u = 0; u < NTW; u++) {
    const int n = ((z + zp) * G.H + y0 + wv) * G.Wd + x0 + 16 * u + nl;
#pragma unroll
    for (int m = 0; m < MT; m++) {
      const int cb = 16 * m + 4 * grp;
      if (cb >= Cout) continue;
      float v[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        v[r] = acc[zp][u][m][r] + (bias ? bias[cb + r] : 0.f);
        if (relu) v[r] = v[r] > 0.f ? v[r] : 0.f;
      }
      if (Y) {
        uint2 pk;
        pk.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
        pk.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
        *(uint2 *)(Y + (long)n * Cout + cb) = pk;
      }
      if (Yf) {
#pragma unroll
        for (int r = 0; r < 4; r++) Yf[(long)n * Cout + cb + r] = v[r];
      }
    }
  }
}

/* ---------------------------------------------------------------------- */
/* Weight gradient for one tap t: dW[co][t*Cin+ci] += sum_n dY[n][co] *     */
/* X[n + off(t)][ci].  Block: 256 threads; wave w computes co tile w (if    */
/* w < Cout/16) against all Cin/16 ci tiles; K = nodes in chunks of 32,      */
/* staged through LDS transposed ([c][node]).                               */

#define WG_CHUNK 32
template <int NT>
__global__ void __launch_bounds__(256)
k_conv_wgrad(const uint16_t *__restrict__ X, int Cin, const uint16_t *__restrict__ dY, int Cout,
             float *__restrict__ dW, int Kp, int t, int nodes_per_block, NGeom G) {
  __shared__ uint16_t sD[64 * WG_CHUNK];   /* [co][node] */
  __shared__ uint16_t sX[64 * WG_CHUNK];   /* [ci][node] */
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int grp = lane >> 4;
  const int nbeg = blockIdx.x * nodes_per_block;
  const int nend = min(G.N, nbeg + nodes_per_block);
  f32x4 acc[NT];
#pragma unroll
  for (int q = 0; q < NT; q++) acc[q] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const bool active = (wv * 16) < Cout;
  for (int c0 = nbeg; c0 < nend; c0 += WG_CHUNK) {
    __syncthreads();
    /* stage dY[c0..c0+31][0..Cout) and X[shift(c0..)][0..Cin) transposed */
    for (int e = threadIdx.x; e < WG_CHUNK * (Cout / 8); e += blockDim.x) {
      const int p = e / (Cout / 8), cg = (e - p * (Cout / 8)) * 8;
      const int n = c0 + p;
      uint16_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (n < nend) *(uint4 *)v = *(const uint4 *)(dY + (long)n * Cout + cg);
#pragma unroll
      for (int q = 0; q < 8; q++) sD[(cg + q) * WG_CHUNK + p] = v[q];
    }
    for (int e = threadIdx.x; e < WG_CHUNK * (Cin / 8); e += blockDim.x) {
      const int p = e / (Cin / 8), cg = (e - p * (Cin / 8)) * 8;
      const int n = c0 + p;
      uint16_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (n < nend) {
        const int z = n / G.plane, rem = n - z * G.plane, y = rem / G.Wd, x = rem - y * G.Wd;
        const int src = tap_node(G, z, y, x, t, 1);
        *(uint4 *)v = *(const uint4 *)(X + (long)src * Cin + cg);
      }
#pragma unroll
      for (int q = 0; q < 8; q++) sX[(cg + q) * WG_CHUNK + p] = v[q];
    }
    __syncthreads();
    if (active) {
      const int co = wv * 16 + (lane & 15);
      const bf16x8 a = *(const bf16x8 *)(sD + co * WG_CHUNK + 8 * grp);
#pragma unroll
      for (int q = 0; q < NT; q++) {
        const int ci = q * 16 + (lane & 15);
        const bf16x8 b = *(const bf16x8 *)(sX + ci * WG_CHUNK + 8 * grp);
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[q], 0, 0, 0);
      }
    }
  }
  if (!active) return;
  /* lane holds D[co = wv*16 + 4*grp + r][ci = q*16 + lane%16] */
#pragma unroll
  for (int q = 0; q < NT; q++) {
    const int ci = q * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int co = wv * 16 + 4 * grp + r;
      atomicAdd(dW + (long)co * Kp + t * Cin + ci, acc[q][r]);
    }
  }
}

/* ---------------------------------------------------------------------- */
/* Weight gradient as a GEMM over the node-blocked layout                    */
/*   Xb[((n >> 3) * C + c) * 8 + (n & 7)]                                     */
/* (8 consecutive nodes of one channel in 16 bytes, x-rows of a multiple of  */
/* 8 nodes).  Block (group range, tap t): M = Cout, N = Cin, K = nodes; the   */
/* 4 waves take interleaved K-steps of 32 nodes (4 groups) and are reduced in */
/* LDS; partial sums per block go to part[t][block][co][ci] and are added in  */
/* a fixed order by k_wgrad_reduce (deterministic).  The tap's x shift of one */
/* node is a one-element funnel of two 8-node groups of the same x-row.        */

__global__ void k_block(const uint16_t *__restrict__ X, int C, uint16_t *__restrict__ Xb, long N) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (N >> 3) * C) return;
  const long g = e / C;
  const int c = (int)(e - g * C);
  uint32_t w[4];
#pragma unroll
  for (int k = 0; k < 4; k++)
    w[k] = (uint32_t)X[(g * 8 + 2 * k) * C + c] | ((uint32_t)X[(g * 8 + 2 * k + 1) * C + c] << 16);
  *(uint4 *)(Xb + e * 8) = make_uint4(w[0], w[1], w[2], w[3]);
}



template <int MT, int NT, int TX>
__global__ void __launch_bounds__(256)
k_wgrad_b(const uint16_t *__restrict__ Xb, const uint16_t *__restrict__ Db,
          float *__restrict__ part, int gpb, NGeom G) {
  /* TX = 1: blockIdx.y is the tap; TX = 3: blockIdx.y is the (dz, dy) pair
   * and the block computes the three taps dx = -1, 0, +1 from one load of
   * the output gradient and of the input groups either side */
  constexpr int Cin = 16 * NT, Cout = 16 * MT;
  __shared__ float red[3][TX * MT * NT * 4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, grp = lane >> 4, m = lane & 15;
  const int tq = blockIdx.y;
  const int dz = TX == 3 ? tq / 3 - 1 : tq / 9 - 1;
  const int dy = TX == 3 ? tq % 3 - 1 : (tq / 3) % 3 - 1;
  const int dx1 = TX == 3 ? 0 : tq % 3 - 1;   /* the single dx of TX = 1 */
  const int gW = G.Wd >> 3;
  const long ngr = (long)G.N >> 3;
  const long gb = (long)blockIdx.x * gpb;
  const long ge = min(ngr, gb + gpb);
  f32x4 acc[TX][MT][NT];
#pragma unroll
  for (int d = 0; d < TX; d++)
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
      for (int j = 0; j < NT; j++) acc[d][i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (long g0 = gb + 4 * wv; g0 < ge; g0 += 16) {
    const long g = g0 + grp;
    const bool ok = g < ge;
    const long gg = ok ? g : gb;
    bf16x8 a[MT];
#pragma unroll
    for (int i = 0; i < MT; i++) {
      u32x4 v = *(const u32x4 *)(Db + (gg * Cout + 16 * i + m) * 8);
      if (!ok) v = (u32x4){0u, 0u, 0u, 0u};
      a[i] = as_bf(v);
    }
    const int row = (int)(gg / gW), xg = (int)(gg - (long)row * gW);
    const int z = row / G.H, y = row - z * G.H;
    int zz = z + dz, yy = y + dy;
    zz += (zz < 0) ? G.D : 0; zz -= (zz >= G.D) ? G.D : 0;
    yy += (yy < 0) ? G.H : 0; yy -= (yy >= G.H) ? G.H : 0;
    const long rb = ((long)zz * G.H + yy) * gW;
    const long s0 = rb + xg;
    const long sm = rb + (xg > 0 ? xg - 1 : gW - 1);     /* the group at x - 8 */
    const long sp = rb + (xg < gW - 1 ? xg + 1 : 0);     /* the group at x + 8 */
#pragma unroll
    for (int j = 0; j < NT; j++) {
      const int ci = 16 * j + m;
      const u32x4 v0 = *(const u32x4 *)(Xb + (s0 * Cin + ci) * 8);
      u32x4 vm = v0, vp = v0;
      if (TX == 3 || dx1 < 0) vm = *(const u32x4 *)(Xb + (sm * Cin + ci) * 8);
      if (TX == 3 || dx1 > 0) vp = *(const u32x4 *)(Xb + (sp * Cin + ci) * 8);
#pragma unroll
      for (int d = 0; d < TX; d++) {
        const int dx = TX == 3 ? d - 1 : dx1;
        u32x4 v;
        if (dx == 0) {
          v = v0;
        } else if (dx > 0) {
          /* nodes x0+1 .. x0+8: elements 1..7 of v0, element 0 of vp */
          v.x = (v0.x >> 16) | (v0.y << 16);
          v.y = (v0.y >> 16) | (v0.z << 16);
          v.z = (v0.z >> 16) | (v0.w << 16);
          v.w = (v0.w >> 16) | (vp.x << 16);
        } else {
          /* nodes x0-1 .. x0+6: element 7 of vm, elements 0..6 of v0 */
          v.x = (vm.w >> 16) | (v0.x << 16);
          v.y = (v0.x >> 16) | (v0.y << 16);
          v.z = (v0.y >> 16) | (v0.z << 16);
          v.w = (v0.z >> 16) | (v0.w << 16);
        }
        if (!ok) v = (u32x4){0u, 0u, 0u, 0u};
        const bf16x8 bb = as_bf(v);
#pragma unroll
        for (int i = 0; i < MT; i++)
          acc[d][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bb, acc[d][i][j], 0, 0, 0);
      }
    }
  }
  /* reduce the 4 waves' partial tiles: waves 1..3 write, wave 0 adds */
  if (wv > 0) {
#pragma unroll
    for (int d = 0; d < TX; d++)
#pragma unroll
      for (int i = 0; i < MT; i++)
#pragma unroll
        for (int j = 0; j < NT; j++)
#pragma unroll
          for (int r = 0; r < 4; r++)
            red[wv - 1][((d * MT + i) * NT + j) * 4 + r][lane] = acc[d][i][j][r];
  }
  __syncthreads();
  if (wv > 0) return;
#pragma unroll
  for (int d = 0; d < TX; d++) {
    const int t = TX == 3 ? 3 * tq + d : tq;
    float *out = part + ((long)t * gridDim.x + blockIdx.x) * Cout * Cin;
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
      for (int j = 0; j < NT; j++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int q = ((d * MT + i) * NT + j) * 4 + r;
          const float v = acc[d][i][j][r] + red[0][q][lane] + red[1][q][lane] + red[2][q][lane];
          /* lane holds D[co = 16i + 4 grp + r][ci = 16j + m] */
          out[(16 * i + 4 * grp + r) * Cin + 16 * j + m] = v;
        }
  }
}

/* Weight gradient of the 16-to-16 full-resolution layers, staged in LDS.
 * A brick is two x-rows (y0, y0 + 1) of plane z.  The block holds the 12 input
 * rows (z - 1 .. z + 1) x (y0 - 1 .. y0 + 2), 6 of them carried over from the
 * previous brick of the same plane, and stages the 2 output-gradient rows
 * transposed to [row][channel][x] (channel rows padded by 8 elements, so the
 * 16 channels of one read fall in different banks), and all 27 taps read them
 * there: the MFMA operands are 8 consecutive nodes of one channel (aligned
 * 16-byte reads), the x shifts of +-1 a funnel with one dword of the
 * neighbouring group, wrapping periodically within the row.  Wave w takes
 * 32-node chunks w, w + 4, ... of each brick and keeps all 27 taps'
 * accumulators; a block walks a contiguous range of bricks (blocks on one XCD
 * take adjacent ranges), then reduces its 4 waves in LDS and writes
 * part[t][block][co][ci] for k_wgrad_reduce.  Requires Wd % 8 == 0,
 * Wd <= 128 and H even. */
#define WG16_SW(Wd) ((Wd) + 8)
__global__ void __launch_bounds__(256)
k_wgrad16_lds(const uint16_t *__restrict__ X, const uint16_t *__restrict__ dY,
              float *__restrict__ part, int per, NGeom G) {
  extern __shared__ __align__(16) uint16_t lds[];
  const int Wd = G.Wd, SW = WG16_SW(G.Wd), hb = G.H >> 1;
  const int nbricks = G.D * hb;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, grp = lane >> 4, m = lane & 15;
  /* blocks b, b + 8, ... run on one XCD: give them adjacent brick ranges */
  int b = blockIdx.x;
  if ((gridDim.x & 7) == 0) b = (b & 7) * (gridDim.x >> 3) + (b >> 3);
  const int kb = b * per, ke = min(nbricks, kb + per);
  uint16_t *Xs = lds;                       /* [12][16][SW] */
  uint16_t *Ds = lds + 12 * 16 * SW;        /* [2][16][SW] */
  const int gpr = Wd >> 3;                  /* 8-node groups per row */
  const int cpr = (gpr + 3) >> 2;           /* 32-node chunks per row */
  f32x4 acc[27];
#pragma unroll
  for (int t = 0; t < 27; t++) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
  int zprev = -1;
  for (int k = kb; k < ke; k++) {
    const int z = k / hb, y0 = 2 * (k - z * hb);
    /* input rows y0 - 1 .. y0 + 2 of each plane sit in ring slots (y + 4) & 3
     * (y unwrapped within the plane): the next brick of the same plane keeps
     * rows y0 + 1, y0 + 2 and loads y0 + 3, y0 + 4 into the two freed slots */
    const bool step = (z == zprev);
    zprev = z;
    const int nx = step ? 6 : 12;
    __syncthreads();                        /* the previous brick is consumed */
    const int pairs = Wd >> 1;
    for (int p = threadIdx.x; p < (nx + 2) * pairs; p += 256) {
      const int r = p / pairs, x = 2 * (p - r * pairs);
      int zz, yy;
      const uint16_t *src;
      uint16_t *dst;
      if (r < nx) {
        const int pl = step ? r >> 1 : r >> 2;
        const int yu = step ? y0 + 1 + (r & 1) : y0 - 1 + (r & 3);
        zz = z + pl - 1;
        yy = yu;
        src = X;
        dst = Xs + (pl * 4 + ((yu + 4) & 3)) * 16 * SW;
      } else {
        zz = z;
        yy = y0 + (r - nx);
        src = dY;
        dst = Ds + (r - nx) * 16 * SW;
      }
      zz += (zz < 0) ? G.D : 0; zz -= (zz >= G.D) ? G.D : 0;
      yy += (yy < 0) ? G.H : 0; yy -= (yy >= G.H) ? G.H : 0;
      const long n = ((long)zz * G.H + yy) * Wd + x;
      const u32x4 a0 = *(const u32x4 *)(src + n * 16);
      const u32x4 a1 = *(const u32x4 *)(src + n * 16 + 8);
      const u32x4 b0 = *(const u32x4 *)(src + (n + 1) * 16);
      const u32x4 b1 = *(const u32x4 *)(src + (n + 1) * 16 + 8);
      const uint32_t av[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
      const uint32_t bv[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
      for (int q = 0; q < 8; q++) {
        /* channel 2q: low halves of nodes x, x + 1; channel 2q + 1: high halves */
        *(uint32_t *)(dst + (2 * q) * SW + x) = (av[q] & 0xffffu) | (bv[q] << 16);
        *(uint32_t *)(dst + (2 * q + 1) * SW + x) = (av[q] >> 16) | (bv[q] & 0xffff0000u);
      }
    }
    __syncthreads();
    for (int c = wv; c < 2 * cpr; c += 4) {
      const int q = c / cpr, gx = 4 * (c - q * cpr) + grp;
      const bool ok = gx < gpr;
      const int x0 = ok ? 8 * gx : 0;
      const int xp = (x0 == 0) ? Wd - 2 : x0 - 2;      /* dword holding node x0 - 1 */
      const int xn = (x0 + 8 == Wd) ? 0 : x0 + 8;      /* dword holding node x0 + 8 */
      u32x4 av = *(const u32x4 *)(Ds + (q * 16 + m) * SW + x0);
      if (!ok) av = (u32x4){0u, 0u, 0u, 0u};
      const bf16x8 A = as_bf(av);
#pragma unroll
      for (int pr = 0; pr < 9; pr++) {
        const int dz = pr / 3, dy = pr % 3;             /* offsets + 1 */
        const int slot = (y0 + q + dy - 1 + 4) & 3;     /* ring slot of row y0 + q + dy - 1 */
        const uint16_t *row = Xs + ((dz * 4 + slot) * 16 + m) * SW;
        const u32x4 v0 = *(const u32x4 *)(row + x0);
        const uint32_t vm = *(const uint32_t *)(row + xp);
        const uint32_t vp = *(const uint32_t *)(row + xn);
        u32x4 bm, bp;
        /* nodes x0 - 1 .. x0 + 6 */
        bm.x = (vm >> 16) | (v0.x << 16);
        bm.y = (v0.x >> 16) | (v0.y << 16);
        bm.z = (v0.y >> 16) | (v0.z << 16);
        bm.w = (v0.z >> 16) | (v0.w << 16);
        /* nodes x0 + 1 .. x0 + 8 */
        bp.x = (v0.x >> 16) | (v0.y << 16);
        bp.y = (v0.y >> 16) | (v0.z << 16);
        bp.z = (v0.z >> 16) | (v0.w << 16);
        bp.w = (v0.w >> 16) | (vp << 16);
        const int t = pr * 3;
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, as_bf(bm), acc[t], 0, 0, 0);
        acc[t + 1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, as_bf(v0), acc[t + 1], 0, 0, 0);
        acc[t + 2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, as_bf(bp), acc[t + 2], 0, 0, 0);
      }
    }
  }
  /* reduce the 4 waves in three rounds of 9 taps: waves 1..3 write, wave 0 adds */
  float *red = (float *)lds;                /* [3][9 * 4][64] */
#pragma unroll
  for (int rd = 0; rd < 3; rd++) {
    __syncthreads();
    if (wv > 0) {
#pragma unroll
      for (int u = 0; u < 9; u++)
#pragma unroll
        for (int r = 0; r < 4; r++) red[((wv - 1) * 36 + u * 4 + r) * 64 + lane] = acc[9 * rd + u][r];
    }
    __syncthreads();
    if (wv == 0) {
#pragma unroll
      for (int u = 0; u < 9; u++) {
        const int t = 9 * rd + u;
        float *out = part + ((long)t * gridDim.x + blockIdx.x) * 256;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int qq = u * 4 + r;
          const float v = acc[t][r] + red[qq * 64 + lane] + red[(36 + qq) * 64 + lane] +
                          red[(72 + qq) * 64 + lane];
          /* lane holds D[co = 4 grp + r][ci = m] */
          out[(4 * grp + r) * 16 + m] = v;
        }
      }
    }
  }
}

static size_t wg16_lds_bytes(int Wd) {
  const size_t stage = (size_t)14 * 16 * WG16_SW(Wd) * sizeof(uint16_t);
  const size_t red = (size_t)3 * 36 * 64 * sizeof(float);
  return stage > red ? stage : red;
}

/* dW[co][t*Cin + ci] += sum over blocks b (in order) of part[t][b][co][ci] */
/* Block of 256 threads = 32 outputs x 8 strands: strand s sums the partials
 * b = s, s + 8, ... in order, then the 8 strand sums are added in order
 * (a fixed order: deterministic). */
__global__ void __launch_bounds__(256)
k_wgrad_reduce(const float *__restrict__ part, int nb, int Cout, int Cin, int Kp,
               float *__restrict__ dW) {
  __shared__ float red[8][32];
  const int o = threadIdx.x & 31, sd = threadIdx.x >> 5;
  const int e = blockIdx.x * 32 + o;
  const bool ok = e < 27 * Cout * Cin;
  float acc = 0.f;
  if (ok) {
    const int t = e / (Cout * Cin), rem = e - t * Cout * Cin;
    const float *p = part + (long)t * nb * Cout * Cin + rem;
    for (int b = sd; b < nb; b += 8) acc += p[(long)b * Cout * Cin];
  }
  red[sd][o] = acc;
  __syncthreads();
  if (sd == 0 && ok) {
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < 8; q++) s += red[q][o];
    const int t = e / (Cout * Cin), rem = e - t * Cout * Cin, co = rem / Cin, ci = rem - co * Cin;
    dW[(long)co * Kp + t * Cin + ci] += s;
  }
}

/* ---------------------------------------------------------------------- */
/* Elementwise and reduction helpers.                                        */

/* WT[ci][t*Cout + co] = W[co][t*Cin + ci]  (taps of the transposed conv) */
__global__ void k_wtrans(const uint16_t *__restrict__ W, int Cout, int Cin, int Kp,
                         uint16_t *__restrict__ WT, int KpT) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Cin * KpT) return;
  const int ci = e / KpT, k = e - ci * KpT;
  uint16_t v = 0;
  if (k < 27 * Cout) {
    const int t = k / Cout, co = k - t * Cout;
    v = W[(long)co * Kp + t * Cin + ci];
  }
  WT[e] = v;
}

/* dPre = dY * (Y > 0) (ReLU derivative from the stored output) */
__global__ void k_relu_bwd(const uint16_t *__restrict__ Y, const uint16_t *__restrict__ dY,
                           uint16_t *__restrict__ dP, long n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  dP[e] = (bf2f(Y[e]) > 0.f) ? dY[e] : (uint16_t)0;
}

/* db[c] += sum_n dP[n][c] for C dividing 256: thread t sums channel t % C
 * over rows t / C, t / C + 256/C, ... of its block's grid-stride share
 * (coalesced rows), the block reduces in LDS and adds one value per channel */
__global__ void __launch_bounds__(256)
k_bias_grad(const uint16_t *__restrict__ dP, int C, int N, float *__restrict__ db) {
  __shared__ float sm[256];
  const int R = 256 / C;
  const int c = threadIdx.x % C, r = threadIdx.x / C;
  float s = 0.f;
  for (long n = (long)blockIdx.x * R + r; n < N; n += (long)gridDim.x * R)
    s += bf2f(dP[n * C + c]);
  sm[threadIdx.x] = s;
  __syncthreads();
  if ((int)threadIdx.x < C) {
    float t = 0.f;
    for (int k = 0; k < R; k++) t += sm[k * C + threadIdx.x];
    atomicAdd(db + threadIdx.x, t);
  }
}

/* 2x2x2 average pooling (channels-last) and its gradient */
__global__ void k_pool(const uint16_t *__restrict__ X, uint16_t *__restrict__ Y, int C, NGeom Gi,
                       NGeom Go) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)Go.N * C) return;
  const int n = (int)(e / C), c = (int)(e - (long)n * C);
  const int z = n / Go.plane, rem = n - z * Go.plane, y = rem / Go.Wd, x = rem - y * Go.Wd;
  float s = 0.f;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int d = 0; d < 2; d++)
        s += bf2f(X[(long)(((2 * z + a) * Gi.H + 2 * y + b) * Gi.Wd + 2 * x + d) * C + c]);
  Y[e] = f2bf(0.125f * s);
}
__global__ void k_pool_bwd(const uint16_t *__restrict__ dY, uint16_t *__restrict__ dX, int C,
                           NGeom Gi, NGeom Go) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)Gi.N * C) return;
  const int n = (int)(e / C), c = (int)(e - (long)n * C);
  const int z = n / Gi.plane, rem = n - z * Gi.plane, y = rem / Gi.Wd, x = rem - y * Gi.Wd;
  const int m = ((z / 2) * Go.H + y / 2) * Go.Wd + x / 2;
  dX[e] = f2bf(0.125f * bf2f(dY[(long)m * C + c]));
}

/* nearest x2 upsampling added to a skip tensor: Y = up(Xc) + S; the
 * gradient of Xc is the sum over the 8 children */
__global__ void k_up_add(const uint16_t *__restrict__ Xc, const uint16_t *__restrict__ S,
                         uint16_t *__restrict__ Y, int C, NGeom Gf, NGeom Gc) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)Gf.N * C) return;
  const int n = (int)(e / C), c = (int)(e - (long)n * C);
  const int z = n / Gf.plane, rem = n - z * Gf.plane, y = rem / Gf.Wd, x = rem - y * Gf.Wd;
  const int m = ((z / 2) * Gc.H + y / 2) * Gc.Wd + x / 2;
  Y[e] = f2bf(bf2f(Xc[(long)m * C + c]) + (S ? bf2f(S[e]) : 0.f));
}
__global__ void k_up_bwd(const uint16_t *__restrict__ dY, uint16_t *__restrict__ dXc, int C,
                         NGeom Gf, NGeom Gc) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)Gc.N * C) return;
  const int n = (int)(e / C), c = (int)(e - (long)n * C);
  const int z = n / Gc.plane, rem = n - z * Gc.plane, y = rem / Gc.Wd, x = rem - y * Gc.Wd;
  float s = 0.f;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int d = 0; d < 2; d++)
        s += bf2f(dY[(long)(((2 * z + a) * Gf.H + 2 * y + b) * Gf.Wd + 2 * x + d) * C + c]);
  dXc[e] = f2bf(s);
}

/* dA += dB elementwise (skip gradients), BF16 storage, FP32 add */
__global__ void k_add_into(uint16_t *__restrict__ A, const uint16_t *__restrict__ B, long n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  A[e] = f2bf(bf2f(A[e]) + bf2f(B[e]));
}

/* Weighted squared error of an FP32 prediction against a BF16 target over
 * the first Cv channels, sum_n sum_c wc[c] (Yf - T)^2 into loss[0] (atomics),
 * and its gradient 2 wc[c] (Yf - T) into dY (BF16, C channels, zero beyond
 * Cv).  wc: Cv FP32 weights in device memory. */
__global__ void __launch_bounds__(256)
k_mse(const float *__restrict__ Yf, int C, const uint16_t *__restrict__ T, int Ct, int Cv,
      int N, const float *__restrict__ wc, uint16_t *__restrict__ dY, float *__restrict__ loss) {
  __shared__ float sm[4];
  float s = 0.f;
  for (long n = (long)blockIdx.x * blockDim.x + threadIdx.x; n < N;
       n += (long)gridDim.x * blockDim.x) {
    for (int c = 0; c < C; c++) {
      float g = 0.f;
      if (c < Cv) {
        const float d = Yf[n * C + c] - bf2f(T[n * Ct + c]);
        s += wc[c] * d * d;
        g = 2.f * d * wc[c];
      }
      dY[n * C + c] = f2bf(g);
    }
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, sm[0] + sm[1] + sm[2] + sm[3]);
}

/* The same loss for C = 16 prediction channels and an even Ct <= 16: one
 * node per thread, its 16 FP32 predictions read as four 16-byte vectors, the
 * Ct BF16 targets as Ct/2 words and its 16 gradients written as two 16-byte
 * vectors (every value computed as by k_mse). */
__global__ void __launch_bounds__(256)
k_mse16(const float *__restrict__ Yf, const uint16_t *__restrict__ T, int Ct, int Cv, int N,
        const float *__restrict__ wc, uint16_t *__restrict__ dY, float *__restrict__ loss) {
  __shared__ float sm[4];
  float w[16];
#pragma unroll
  for (int c = 0; c < 16; c++) w[c] = c < Cv ? wc[c] : 0.f;
  float s = 0.f;
  for (long n = (long)blockIdx.x * blockDim.x + threadIdx.x; n < N;
       n += (long)gridDim.x * blockDim.x) {
    float y[16], t[16];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const f32x4 v = ((const f32x4 *)(Yf + n * 16))[q];
      y[4 * q] = v[0]; y[4 * q + 1] = v[1]; y[4 * q + 2] = v[2]; y[4 * q + 3] = v[3];
    }
    const uint32_t *tw = (const uint32_t *)(T + n * Ct);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const uint32_t u = (2 * q < Ct) ? tw[q] : 0u;
      t[2 * q] = __uint_as_float(u << 16);
      t[2 * q + 1] = __uint_as_float(u & 0xffff0000u);
    }
    float g[16];
#pragma unroll
    for (int c = 0; c < 16; c++) {
      g[c] = 0.f;
      if (c < Cv) {
        const float d = y[c] - t[c];
        s += w[c] * d * d;
        g[c] = 2.f * d * w[c];
      }
    }
    uint32_t o[8];
#pragma unroll
    for (int q = 0; q < 8; q++) o[q] = (uint32_t)f2bf(g[2 * q]) | ((uint32_t)f2bf(g[2 * q + 1]) << 16);
    u32x4 *dst = (u32x4 *)(dY + n * 16);
    dst[0] = (u32x4){o[0], o[1], o[2], o[3]};
    dst[1] = (u32x4){o[4], o[5], o[6], o[7]};
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, sm[0] + sm[1] + sm[2] + sm[3]);
}

/* Adam on FP32 master weights, refreshing the BF16 copy (zero-padded K
 * columns stay zero because their gradient is zero). */
__global__ void k_adam(float *__restrict__ w, uint16_t *__restrict__ wb, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gg = g[e];
  const float mm = b1 * m[e] + (1.f - b1) * gg;
  const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  const float ww = w[e] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  w[e] = ww;
  if (wb) wb[e] = f2bf(ww);
}

__global__ void k_f2bf(const float *__restrict__ a, uint16_t *__restrict__ b, long n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) b[e] = f2bf(a[e]);
}
__global__ void k_bf2f(const uint16_t *__restrict__ a, float *__restrict__ b, long n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) b[e] = bf2f(a[e]);
}

/* 16-byte forms of the element-wise kernels: 8 BF16 values per thread, for
 * tensors whose element count (or channel count) is a multiple of 8.  Each
 * value is computed exactly as by the scalar kernel above. */
__device__ __forceinline__ void unpack8(u32x4 v, float f[8]) {
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int q = 0; q < 4; q++) {
    f[2 * q] = __uint_as_float(w[q] << 16);
    f[2 * q + 1] = __uint_as_float(w[q] & 0xffff0000u);
  }
}
__device__ __forceinline__ u32x4 pack8(const float f[8]) {
  uint32_t w[4];
#pragma unroll
  for (int q = 0; q < 4; q++) w[q] = (uint32_t)f2bf(f[2 * q]) | ((uint32_t)f2bf(f[2 * q + 1]) << 16);
  return (u32x4){w[0], w[1], w[2], w[3]};
}

__global__ void k_relu_bwd8(const uint16_t *__restrict__ Y, const uint16_t *__restrict__ dY,
                            uint16_t *__restrict__ dP, long n8) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n8) return;
  const u32x4 y = ((const u32x4 *)Y)[e], d = ((const u32x4 *)dY)[e];
  const uint32_t yw[4] = {y.x, y.y, y.z, y.w}, dw[4] = {d.x, d.y, d.z, d.w};
  uint32_t o[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const bool lo = __uint_as_float(yw[q] << 16) > 0.f;
    const bool hi = __uint_as_float(yw[q] & 0xffff0000u) > 0.f;
    o[q] = (lo ? (dw[q] & 0xffffu) : 0u) | (hi ? (dw[q] & 0xffff0000u) : 0u);
  }
  ((u32x4 *)dP)[e] = (u32x4){o[0], o[1], o[2], o[3]};
}

__global__ void k_add_into8(uint16_t *__restrict__ A, const uint16_t *__restrict__ B, long n8) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n8) return;
  float a[8], b[8];
  unpack8(((const u32x4 *)A)[e], a);
  unpack8(((const u32x4 *)B)[e], b);
#pragma unroll
  for (int q = 0; q < 8; q++) a[q] += b[q];
  ((u32x4 *)A)[e] = pack8(a);
}

/* element e8 = (node, channel group of 8) */
__global__ void k_up_add8(const uint16_t *__restrict__ Xc, const uint16_t *__restrict__ S,
                          uint16_t *__restrict__ Y, int C, NGeom Gf, NGeom Gc) {
  const int G8 = C >> 3;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)Gf.N * G8) return;
  const int n = (int)(e / G8), g = (int)(e - (long)n * G8);
  const int z = n / Gf.plane, rem = n - z * Gf.plane, y = rem / Gf.Wd, x = rem - y * Gf.Wd;
  const int m = ((z / 2) * Gc.H + y / 2) * Gc.Wd + x / 2;
  float a[8], b[8];
  unpack8(*(const u32x4 *)(Xc + (long)m * C + 8 * g), a);
  if (S) {
    unpack8(((const u32x4 *)S)[e], b);
#pragma unroll
    for (int q = 0; q < 8; q++) a[q] += b[q];
  } else {
#pragma unroll
    for (int q = 0; q < 8; q++) a[q] += 0.f;
  }
  ((u32x4 *)Y)[e] = pack8(a);
}

__global__ void k_pool_bwd8(const uint16_t *__restrict__ dY, uint16_t *__restrict__ dX, int C,
                            NGeom Gi, NGeom Go) {
  const int G8 = C >> 3;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)Gi.N * G8) return;
  const int n = (int)(e / G8), g = (int)(e - (long)n * G8);
  const int z = n / Gi.plane, rem = n - z * Gi.plane, y = rem / Gi.Wd, x = rem - y * Gi.Wd;
  const int m = ((z / 2) * Go.H + y / 2) * Go.Wd + x / 2;
  float a[8];
  unpack8(*(const u32x4 *)(dY + (long)m * C + 8 * g), a);
#pragma unroll
  for (int q = 0; q < 8; q++) a[q] = 0.125f * a[q];
  ((u32x4 *)dX)[e] = pack8(a);
}

/* db[c] += sum_n dP[n][c]: thread t sums channel group t % (C/8) over rows
 * t / (C/8), + 256/(C/8), ... of its block's grid-stride share, the block
 * reduces in LDS in a fixed order and adds one value per channel */
__global__ void __launch_bounds__(256)
k_bias_grad8(const uint16_t *__restrict__ dP, int C, int N, float *__restrict__ db) {
  __shared__ float sm[256 * 8];
  const int G8 = C >> 3, R = 256 / G8;
  const int g = threadIdx.x % G8, r = threadIdx.x / G8;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (long n = (long)blockIdx.x * R + r; n < N; n += (long)gridDim.x * R) {
    float f[8];
    unpack8(*(const u32x4 *)(dP + n * C + 8 * g), f);
#pragma unroll
    for (int q = 0; q < 8; q++) acc[q] += f[q];
  }
#pragma unroll
  for (int q = 0; q < 8; q++) sm[threadIdx.x * 8 + q] = acc[q];
  __syncthreads();
  if ((int)threadIdx.x < C) {
    const int gg = threadIdx.x >> 3, q = threadIdx.x & 7;
    float t = 0.f;
    for (int k = 0; k < R; k++) t += sm[(k * G8 + gg) * 8 + q];
    atomicAdd(db + threadIdx.x, t);
  }
}

/* dP = dY * (Y > 0) and db[c] += sum_n dP[n][c] in one pass: the ReLU
 * backward of k_relu_bwd8 fused into k_bias_grad8's loop (same per-thread
 * order and block reduction, one atomic per channel and block), so dP is
 * written once and not read back. */
__global__ void __launch_bounds__(256)
k_relu_bwd_bias8(const uint16_t *__restrict__ Y, const uint16_t *__restrict__ dY,
                 uint16_t *__restrict__ dP, int C, int N, float *__restrict__ db) {
  __shared__ float sm[256 * 8];
  const int G8 = C >> 3, R = 256 / G8;
  const int g = threadIdx.x % G8, r = threadIdx.x / G8;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (long n = (long)blockIdx.x * R + r; n < N; n += (long)gridDim.x * R) {
    const long e = n * G8 + g;    /* 16-byte group of channels 8g .. 8g + 7 */
    const u32x4 y = ((const u32x4 *)Y)[e], d = ((const u32x4 *)dY)[e];
    const uint32_t yw[4] = {y.x, y.y, y.z, y.w}, dw[4] = {d.x, d.y, d.z, d.w};
    uint32_t o[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const bool lo = __uint_as_float(yw[q] << 16) > 0.f;
      const bool hi = __uint_as_float(yw[q] & 0xffff0000u) > 0.f;
      o[q] = (lo ? (dw[q] & 0xffffu) : 0u) | (hi ? (dw[q] & 0xffff0000u) : 0u);
    }
    const u32x4 ov = (u32x4){o[0], o[1], o[2], o[3]};
    ((u32x4 *)dP)[e] = ov;
    float f[8];
    unpack8(ov, f);
#pragma unroll
    for (int q = 0; q < 8; q++) acc[q] += f[q];
  }
#pragma unroll
  for (int q = 0; q < 8; q++) sm[threadIdx.x * 8 + q] = acc[q];
  __syncthreads();
  if ((int)threadIdx.x < C) {
    const int gg = threadIdx.x >> 3, q = threadIdx.x & 7;
    float t = 0.f;
    for (int k = 0; k < R; k++) t += sm[(k * G8 + gg) * 8 + q];
    atomicAdd(db + threadIdx.x, t);
  }
}

/* ---------------------------------------------------------------------- */
/* Host ABI.                                                                */

static NGeom ngeom(int D, int H, int Wd) {
  NGeom g;
  g.D = D; g.H = H; g.Wd = Wd; g.plane = H * Wd; g.N = D * H * Wd;
  return g;
}

/* node tiles of 16 per wave in the convolution (weights of a K chunk are
 * loaded once for all of them) */
#ifndef SLBN_CONV_NTW
#define SLBN_CONV_NTW 4
#endif

static hipStream_t nn_stream = 0;
extern "C" void slbn_set_stream(void *s) { nn_stream = (hipStream_t)s; }

static unsigned blocks_for(long n, int b) { return (unsigned)((n + b - 1) / b); }

extern "C" int slbn_kpad(int cin) { return ((27 * cin + 31) / 32) * 32; }

extern "C" void slbn_conv(const uint16_t *X, int cin, const uint16_t *W, const float *bias,
                          uint16_t *Y, float *Yf, int cout, int relu, int transposed, int D, int H,
                          int Wd) {
  const NGeom G = ngeom(D, H, Wd);
  const int Kp = slbn_kpad(cin);
  /* 4 waves per block, each NTW node tiles of 16 */
  const unsigned g = blocks_for(G.N, 64 * SLBN_CONV_NTW);
  const int sgn = transposed ? -1 : 1;
  if (cin < 8 || cin % 8 || cout < 1 || cout % 4 || (long)G.N * (cin > cout ? cin : cout) >= (1L << 31)) {
    fprintf(stderr, "slbn_conv: cin must be a multiple of 8, cout of 4, tensors < 2^31 elements\n");
    exit(EXIT_FAILURE);
  }
  const int mt = (cout + 15) / 16;
  /* 16, 32 or 64 input channels on rows of a multiple of 1024/cin nodes:
   * the LDS-staged kernel (SLBN_CONV_LDS=0 selects the general one) */
  const char *lds_env = getenv("SLBN_CONV_LDS");
  const int xt = cin >= 16 ? 1024 / cin : 0;
  if ((cin == 16 || cin == 32 || cin == 64) && Wd % xt == 0 && H % 4 == 0 && mt <= 4 &&
      !(lds_env && atoi(lds_env) == 0)) {
    /* two planes per block where D is even and it measured faster (128^3
     * network, us per call): 16 inputs with up to 32 outputs (96.6 -> 83.4),
     * 32 -> 32 (43.0 -> 41.0), 64 inputs (41.6 -> 38.8, 32.2 -> 28.3); not
     * 32 -> 16 (32.1 -> 32.9) or 32 -> 64 (24.7 -> 30.7) */
    const int zt = (D % 2 == 0 && ((cin == 16 && mt <= 2) || (cin == 32 && mt == 2) ||
                                   cin == 64)) ? 2 : 1;
    const dim3 gl((unsigned)((Wd / xt) * (H / 4) * (D / zt)));
#define SLBN_CL(C, MT, ZT)                                                                  \
  hipLaunchKernelGGL((k_conv_lds<C, MT, ZT>), gl, dim3(256), 0, nn_stream, X, W, Kp, bias, Y, \
                     Yf, cout, relu, sgn, G)
#define SLBN_CLM(C, ZT)                                                                     \
  if (mt == 1) SLBN_CL(C, 1, ZT); else if (mt == 2) SLBN_CL(C, 2, ZT);                      \
  else if (mt == 3) SLBN_CL(C, 3, ZT); else SLBN_CL(C, 4, ZT);
    if (cin == 16 && zt == 2) {
      if (mt == 1) SLBN_CL(16, 1, 2); else SLBN_CL(16, 2, 2);
    } else if (cin == 16) { SLBN_CLM(16, 1) }
    else if (cin == 32 && zt == 2) { SLBN_CL(32, 2, 2); }
    else if (cin == 32) { SLBN_CLM(32, 1) }
    else if (zt == 2) { SLBN_CLM(64, 2) }
    else { SLBN_CLM(64, 1) }
#undef SLBN_CLM
#undef SLBN_CL
    LAUNCHCHECK();
    return;
  }
#define SLBN_CONV(MT)                                                                      \
  hipLaunchKernelGGL((k_conv<MT, SLBN_CONV_NTW>), dim3(g), dim3(256), 0, nn_stream, X, cin, W, \
                     Kp, bias, Y, Yf, \
                     cout, relu, sgn, G)
  if (mt == 1) SLBN_CONV(1);
  else if (mt == 2) SLBN_CONV(2);
  else if (mt == 3) SLBN_CONV(3);
  else if (mt == 4) SLBN_CONV(4);
  else { fprintf(stderr, "slbn_conv: at most 64 output channels\n"); exit(EXIT_FAILURE); }
#undef SLBN_CONV
  LAUNCHCHECK();
}

/* workspace of the blocked weight gradient (grown on demand, kept) */
static void *wg_ws = nullptr;
static size_t wg_ws_bytes = 0;
static void *wg_workspace(size_t bytes) {
  if (bytes > wg_ws_bytes) {
    HIPCHECK(hipStreamSynchronize(nn_stream));
    if (wg_ws) HIPCHECK(hipFree(wg_ws));
    HIPCHECK(hipMalloc(&wg_ws, bytes));
    wg_ws_bytes = bytes;
  }
  return wg_ws;
}

/* the original per-tap LDS-staged kernel: any extents (used when the
 * x-rows are not a multiple of 8 nodes) */
static void wgrad_staged(const uint16_t *X, int cin, const uint16_t *dY, int cout, float *dW,
                         const NGeom &G) {
  const int Kp = slbn_kpad(cin);
  int npb = WG_CHUNK * 64;
  if ((long)npb * 64 < G.N) npb = (int)(((G.N + 63) / 64 + WG_CHUNK - 1) / WG_CHUNK) * WG_CHUNK;
  const unsigned g = blocks_for(G.N, npb);
  for (int t = 0; t < 27; t++) {
#define SLBN_WG(NT)                                                                       \
  hipLaunchKernelGGL(k_conv_wgrad<NT>, dim3(g), dim3(256), 0, nn_stream, X, cin, dY, cout, dW, \
                     Kp, t, npb, G)
    const int nt = cin / 16;
    if (nt == 1) SLBN_WG(1);
    else if (nt == 2) SLBN_WG(2);
    else if (nt == 3) SLBN_WG(3);
    else SLBN_WG(4);
#undef SLBN_WG
    LAUNCHCHECK();
  }
}

/* the LDS-staged 16 -> 16 weight gradient applies (SLBN_WGRAD_LDS=0 turns it off) */
extern "C" int slbn_wgrad_lds16_ok(int cin, int cout, int D, int H, int Wd) {
  const char *env = getenv("SLBN_WGRAD_LDS");   /* read per call, as SLBN_CONV_LDS */
  return !(env && atoi(env) == 0) && cin == 16 && cout == 16 && D >= 1 && H >= 2 && H % 2 == 0 &&
         Wd >= 8 && Wd % 8 == 0 && Wd <= 128;
}

extern "C" int slbn_wgrad_blocked_ok(int cin, int cout, int D, int H, int Wd) {
  (void)D; (void)H;
  return Wd % 8 == 0 && (cin == 16 || cin == 32 || cin == 64) &&
         (cout == 16 || cout == 32 || cout == 64);
}

extern "C" void slbn_conv_wgrad(const uint16_t *X, int cin, const uint16_t *dY, int cout,
                                float *dW, int D, int H, int Wd) {
  const NGeom G = ngeom(D, H, Wd);
  const int Kp = slbn_kpad(cin);
  if (cin > 64 || cout > 64 || (cin % 16) || (cout % 16)) {
    fprintf(stderr, "slbn_conv_wgrad: channels must be 16..64 in steps of 16\n");
    exit(EXIT_FAILURE);
  }
  if (!slbn_wgrad_blocked_ok(cin, cout, D, H, Wd)) {
    wgrad_staged(X, cin, dY, cout, dW, G);
    return;
  }
  if (slbn_wgrad_lds16_ok(cin, cout, D, H, Wd)) {
    /* 16 -> 16 channels: LDS-staged bricks, about two blocks per CU */
    const int nbricks = D * (H / 2);
    int nb = nbricks < 512 ? nbricks : 512;
    const int per = (nbricks + nb - 1) / nb;
    nb = (nbricks + per - 1) / per;
    float *part = (float *)wg_workspace((size_t)27 * nb * 256 * sizeof(float));
    hipLaunchKernelGGL(k_wgrad16_lds, dim3((unsigned)nb), dim3(256), wg16_lds_bytes(Wd), nn_stream,
                       X, dY, part, per, G);
    LAUNCHCHECK();
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks_for(27L * 256, 32)), dim3(256), 0, nn_stream,
                       part, nb, 16, 16, Kp, dW);
    LAUNCHCHECK();
    return;
  }
  const long ngr = (long)G.N >> 3;
  /* enough blocks per tap to fill the chip with 27 taps, each block a
   * contiguous group range in steps of 16 groups */
  const int mt = cout / 16, nt = cin / 16;
  /* three taps per block where the accumulators fit (MT NT <= 2) */
  const int tx = (mt * nt <= 2) ? 3 : 1;
  const long nbmax = tx == 3 ? 192 : 64;
  long nb = (ngr + 255) / 256;
  if (nb > nbmax) nb = nbmax;
  if (nb < 1) nb = 1;
  int gpb = (int)(((ngr + nb - 1) / nb + 15) / 16 * 16);
  nb = (ngr + gpb - 1) / gpb;
  const size_t xb = (size_t)G.N * cin * 2, db = (size_t)G.N * cout * 2;
  const size_t pb = (size_t)27 * nb * cout * cin * 4;
  char *ws = (char *)wg_workspace(xb + db + pb);
  uint16_t *Xb = (uint16_t *)ws, *Db = (uint16_t *)(ws + xb);
  float *part = (float *)(ws + xb + db);
  hipLaunchKernelGGL(k_block, dim3(blocks_for(ngr * cin, 256)), dim3(256), 0, nn_stream, X, cin,
                     Xb, (long)G.N);
  LAUNCHCHECK();
  hipLaunchKernelGGL(k_block, dim3(blocks_for(ngr * cout, 256)), dim3(256), 0, nn_stream, dY, cout,
                     Db, (long)G.N);
  LAUNCHCHECK();
  const dim3 grid((unsigned)nb, 27 / tx);
#define SLBN_WGB(MT, NT, TX)                                                             \
  hipLaunchKernelGGL((k_wgrad_b<MT, NT, TX>), grid, dim3(256), 0, nn_stream, Xb, Db, part, \
                     gpb, G)
  if (mt == 1 && nt == 1) SLBN_WGB(1, 1, 3);
  else if (mt == 1 && nt == 2) SLBN_WGB(1, 2, 3);
  else if (mt == 2 && nt == 1) SLBN_WGB(2, 1, 3);
  else if (mt == 1 && nt == 4) SLBN_WGB(1, 4, 1);
  else if (mt == 2 && nt == 2) SLBN_WGB(2, 2, 1);
  else if (mt == 2 && nt == 4) SLBN_WGB(2, 4, 1);
  else if (mt == 4 && nt == 1) SLBN_WGB(4, 1, 1);
  else if (mt == 4 && nt == 2) SLBN_WGB(4, 2, 1);
  else SLBN_WGB(4, 4, 1);
#undef SLBN_WGB
  LAUNCHCHECK();
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks_for(27L * cout * cin, 32)), dim3(256), 0,
                     nn_stream, part, (int)nb, cout, cin, Kp, dW);
  LAUNCHCHECK();
}

extern "C" void slbn_wtrans(const uint16_t *W, int cout, int cin, uint16_t *WT) {
  const int Kp = slbn_kpad(cin), KpT = slbn_kpad(cout);
  hipLaunchKernelGGL(k_wtrans, dim3(blocks_for((long)cin * KpT, 256)), dim3(256), 0, nn_stream,
                     W, cout, cin, Kp, WT, KpT);
  LAUNCHCHECK();
}

extern "C" void slbn_relu_bwd(const uint16_t *Y, const uint16_t *dY, uint16_t *dP, long n) {
  if (n % 8 == 0)
    hipLaunchKernelGGL(k_relu_bwd8, dim3(blocks_for(n / 8, 256)), dim3(256), 0, nn_stream, Y, dY,
                       dP, n / 8);
  else
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(n, 256)), dim3(256), 0, nn_stream, Y, dY, dP,
                       n);
  LAUNCHCHECK();
}

extern "C" void slbn_bias_grad(const uint16_t *dP, int C, int N, float *db) {
  if (C < 1 || C > 256 || 256 % C) {
    fprintf(stderr, "slbn_bias_grad: the channel count must divide 256\n");
    exit(EXIT_FAILURE);
  }
  if (C % 8 == 0) {
    const long rows = 256 / (C / 8);
    long nb = ((long)N + rows - 1) / rows;
    if (nb > 1024) nb = 1024;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_bias_grad8, dim3((unsigned)nb), dim3(256), 0, nn_stream, dP, C, N, db);
    LAUNCHCHECK();
    return;
  }
  const long rows = 256 / C;
  long nb = ((long)N + rows - 1) / rows;
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_bias_grad, dim3((unsigned)nb), dim3(256), 0, nn_stream, dP, C, N, db);
  LAUNCHCHECK();
}

extern "C" void slbn_relu_bwd_bias(const uint16_t *Y, const uint16_t *dY, uint16_t *dP, int C,
                                   int N, float *db) {
  if (C < 8 || C % 8 || 256 % C) {
    slbn_relu_bwd(Y, dY, dP, (long)N * C);
    slbn_bias_grad(dP, C, N, db);
    return;
  }
  /* the grid of slbn_bias_grad (4096 blocks: 23 -> 38 us per call, the
   * per-block atomics on C addresses contend) */
  const long rows = 256 / (C / 8);
  long nb = ((long)N + rows - 1) / rows;
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_relu_bwd_bias8, dim3((unsigned)nb), dim3(256), 0, nn_stream, Y, dY, dP, C, N,
                     db);
  LAUNCHCHECK();
}

extern "C" void slbn_pool(const uint16_t *X, uint16_t *Y, int C, int D, int H, int Wd) {
  const NGeom Gi = ngeom(D, H, Wd), Go = ngeom(D / 2, H / 2, Wd / 2);
  hipLaunchKernelGGL(k_pool, dim3(blocks_for((long)Go.N * C, 256)), dim3(256), 0, nn_stream, X, Y,
                     C, Gi, Go);
  LAUNCHCHECK();
}
extern "C" void slbn_pool_bwd(const uint16_t *dY, uint16_t *dX, int C, int D, int H, int Wd) {
  const NGeom Gi = ngeom(D, H, Wd), Go = ngeom(D / 2, H / 2, Wd / 2);
  if (C % 8 == 0)
    hipLaunchKernelGGL(k_pool_bwd8, dim3(blocks_for((long)Gi.N * (C / 8), 256)), dim3(256), 0,
                       nn_stream, dY, dX, C, Gi, Go);
  else
    hipLaunchKernelGGL(k_pool_bwd, dim3(blocks_for((long)Gi.N * C, 256)), dim3(256), 0,
                       nn_stream, dY, dX, C, Gi, Go);
  LAUNCHCHECK();
}
extern "C" void slbn_up_add(const uint16_t *Xc, const uint16_t *S, uint16_t *Y, int C, int D, int H,
                            int Wd) {
  const NGeom Gf = ngeom(D, H, Wd), Gc = ngeom(D / 2, H / 2, Wd / 2);
  if (C % 8 == 0)
    hipLaunchKernelGGL(k_up_add8, dim3(blocks_for((long)Gf.N * (C / 8), 256)), dim3(256), 0,
                       nn_stream, Xc, S, Y, C, Gf, Gc);
  else
    hipLaunchKernelGGL(k_up_add, dim3(blocks_for((long)Gf.N * C, 256)), dim3(256), 0, nn_stream,
                       Xc, S, Y, C, Gf, Gc);
  LAUNCHCHECK();
}
extern "C" void slbn_up_bwd(const uint16_t *dY, uint16_t *dXc, int C, int D, int H, int Wd) {
  const NGeom Gf = ngeom(D, H, Wd), Gc = ngeom(D / 2, H / 2, Wd / 2);
  hipLaunchKernelGGL(k_up_bwd, dim3(blocks_for((long)Gc.N * C, 256)), dim3(256), 0, nn_stream, dY,
                     dXc, C, Gf, Gc);
  LAUNCHCHECK();
}
extern "C" void slbn_add_into(uint16_t *A, const uint16_t *B, long n) {
  if (n % 8 == 0)
    hipLaunchKernelGGL(k_add_into8, dim3(blocks_for(n / 8, 256)), dim3(256), 0, nn_stream, A, B,
                       n / 8);
  else
    hipLaunchKernelGGL(k_add_into, dim3(blocks_for(n, 256)), dim3(256), 0, nn_stream, A, B, n);
  LAUNCHCHECK();
}
extern "C" void slbn_mse(const float *Yf, int C, const uint16_t *T, int Ct, int Cv, int N,
                         const float *wc, uint16_t *dY, float *loss) {
  unsigned nb = blocks_for(N, 256);
  if (nb > 2048) nb = 2048;
  if (C == 16 && Ct % 2 == 0 && Ct <= 16 && Cv <= Ct)
    hipLaunchKernelGGL(k_mse16, dim3(nb), dim3(256), 0, nn_stream, Yf, T, Ct, Cv, N, wc, dY,
                       loss);
  else
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, nn_stream, Yf, C, T, Ct, Cv, N, wc, dY,
                       loss);
  LAUNCHCHECK();
}
extern "C" void slbn_adam(float *w, uint16_t *wb, const float *g, float *m, float *v, long n,
                          float lr, float b1, float b2, float eps, int step) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n, 256)), dim3(256), 0, nn_stream, w, wb, g, m, v, n,
                     lr, b1, b2, eps, c1, c2);
  LAUNCHCHECK();
}
extern "C" void slbn_f2bf(const float *a, uint16_t *b, long n) {
  hipLaunchKernelGGL(k_f2bf, dim3(blocks_for(n, 256)), dim3(256), 0, nn_stream, a, b, n);
  LAUNCHCHECK();
}
extern "C" void slbn_bf2f(const uint16_t *a, float *b, long n) {
  hipLaunchKernelGGL(k_bf2f, dim3(blocks_for(n, 256)), dim3(256), 0, nn_stream, a, b, n);
  LAUNCHCHECK();
}

/* Weight gradient for small channel counts: one thread per (co, t, ci)
 * element and node chunk, FP32, atomics into dW. */
__global__ void __launch_bounds__(256)
k_conv_wgrad_small(const uint16_t *__restrict__ X, int Cin, const uint16_t *__restrict__ dY,
                   int Cout, float *__restrict__ dW, int Kp, int chunk, NGeom G) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Cout * 27 * Cin) return;
  const int co = e / (27 * Cin), rem = e - co * 27 * Cin, t = rem / Cin, ci = rem - t * Cin;
  const int nb = blockIdx.y * chunk, ne = min(G.N, nb + chunk);
  float s = 0.f;
  for (int n = nb; n < ne; n++) {
    const float g = bf2f(dY[(long)n * Cout + co]);
    if (g == 0.f) continue;
    const int z = n / G.plane, r2 = n - z * G.plane, y = r2 / G.Wd, x = r2 - y * G.Wd;
    s += g * bf2f(X[(long)tap_node(G, z, y, x, t, 1) * Cin + ci]);
  }
  if (s != 0.f) atomicAdd(dW + (long)co * Kp + t * Cin + ci, s);
}

extern "C" void slbn_conv_wgrad_small(const uint16_t *X, int cin, const uint16_t *dY, int cout,
                                      float *dW, int D, int H, int Wd) {
  const NGeom G = ngeom(D, H, Wd);
  const int Kp = slbn_kpad(cin);
  const int chunk = 4096;
  dim3 g(blocks_for((long)cout * 27 * cin, 256), blocks_for(G.N, chunk));
  hipLaunchKernelGGL(k_conv_wgrad_small, g, dim3(256), 0, nn_stream, X, cin, dY, cout, dW, Kp,
                     chunk, G);
  LAUNCHCHECK();
}

__global__ void k_expand(const uint16_t *__restrict__ S, int Cs, uint16_t *__restrict__ X, int C,
                         long N) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N * C) return;
  const long n = e / C;
  const int c = (int)(e - n * C);
  X[e] = (c < Cs) ? S[n * Cs + c] : (uint16_t)0;
}

extern "C" void slbn_expand(const uint16_t *S, int Cs, uint16_t *X, int C, long N) {
  hipLaunchKernelGGL(k_expand, dim3(blocks_for(N * C, 256)), dim3(256), 0, nn_stream, S, Cs, X, C, N);
  LAUNCHCHECK();
}

/* ---------------------------------------------------------------------- */
/* Device memory helpers for the host-side network code (plain C).          */

extern "C" void *slbn_malloc(size_t bytes) {
  void *p = nullptr;
  if (bytes == 0) bytes = 256;
  HIPCHECK(hipMalloc(&p, bytes));
  return p;
}
extern "C" void slbn_free(void *p) {
  if (p) HIPCHECK(hipFree(p));
}
extern "C" void slbn_zero(void *p, size_t bytes) {
  HIPCHECK(hipMemsetAsync(p, 0, bytes, nn_stream));
}
extern "C" void slbn_h2d(void *dst, const void *src, size_t bytes) {
  HIPCHECK(hipStreamSynchronize(nn_stream));
  HIPCHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}
extern "C" void slbn_d2h(void *dst, const void *src, size_t bytes) {
  HIPCHECK(hipStreamSynchronize(nn_stream));
  HIPCHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}
extern "C" void slbn_sync(void) { HIPCHECK(hipStreamSynchronize(nn_stream)); }
