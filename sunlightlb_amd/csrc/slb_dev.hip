/*
 * slb_dev.hip -- gfx950 kernels for the SunlightLB D3Q15 lattice-Boltzmann
 * step and the thin C ABI (slb_dev.h) the C host layer calls.
 *
 * The hot path is one fused pull kernel per time step (k_step): every node
 * gathers the 15 post-collide populations that stream into it, applies link
 * bounce-back where the link has exactly one solid end, collides, and writes
 * its 15 new populations to the other buffer of a ping-pong pair.  This is
 * bbl -> propagate -> collide of src/lbkernel.c:49-51 in one pass over HBM:
 * each population is read once and written once.
 *
 * Equivalence of the pull form with the reference's swap-then-shift
 * (src/lbkernel.c:58-618 then :623-762): after the reference's bbl and
 * propagate, the value at x along l is
 *     a_lbar(x) + w_l c_l.U(s)   if exactly one of x, x - c_l is solid
 *                                (s = the solid one; U from vspread),
 *     a_l(x - c_l)               otherwise,
 * with all a post-collide.  The per-node `flags` word holds the first case
 * as one bit per link, derived from the reference mask (bit l-1 of the solid
 * node x - c_l, or bit lbar-1 of x itself).
 *
 * Momentum-exchange force (bblforce, src/lbkernel.c:179-321, :464-618): for a
 * solid node s with boundary link m (f = s + c_m fluid) the reference adds
 * 2 (a_m(s) - a_mbar(f) - w_m c_m.U(s)) c_m to the force of s, using values
 * before the swap.  In the pull kernel the solid node s sees that link as an
 * incoming bounce along l = mbar; it already loads a_m(s) and loads a_mbar(f)
 * = a_l(s - c_l) as well, so each solid node accumulates its own force with
 * no atomics.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <type_traits>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "slb_dev.h"

#include "slb_kern.h"

/* ---------------------------------------------------------------------- */
/* Fused step kernel.                                                      */

/* Plane addressing of one step launch.  A launch updates the planes
 * i0 .. i0 + gridDim.y - 1 (mod esi).  Source and destination populations
 * of plane i live at plane slot sring/dring(i) of their buffer: ring 0 is
 * the plain layout (slot i); ring R > 0 keeps planes 0 and 1 in slots 0
 * and 1 and plane i >= 2 in slot 2 + (i - 2) mod R, so a buffer of R + 2
 * planes holds a sliding window of the lattice (the intermediate step of
 * the two-step sweep).  Flags, wall velocities, body and node forces are
 * always indexed by the global node. */
struct SweepMap {
  int i0, sring, dring;
};

__device__ __forceinline__ int plane_slot(int i, int ring) {
  return (ring == 0 || i < 2) ? i : 2 + (i - 2) % ring;
}

/* I: index type, int when the lattice holds fewer than 2^31 populations
 * (half the address arithmetic), long otherwise. */
template <typename T, bool MOVING, bool FORCE, int BF, typename I>
__global__ void __launch_bounds__(512)
k_step(const T *__restrict__ src, T *__restrict__ dst,
       const uint16_t *__restrict__ flags,
       const double *__restrict__ ux, const double *__restrict__ uy,
       const double *__restrict__ uz,
       const T *__restrict__ bx, const T *__restrict__ by, const T *__restrict__ bz,
       double *__restrict__ fx, double *__restrict__ fy, double *__restrict__ fz,
       Geom G, PhysArgs P, SweepMap M) {
  int blk_x, blk_y;
  step_block(G, blk_x, blk_y);
  const int t = blk_x * blockDim.x + threadIdx.x;
  if (t >= G.plane) return;
  int i = blk_y + M.i0;
  if (i >= G.esi) i -= G.esi;
  const int j = t / G.esk;
  const int k = t - j * G.esk;
  const I x = (I)i * (I)G.plane + (I)t;
  const AxisOffT<I> oi = axis_off_t<I>(i, G.esi, (I)G.plane);
  const AxisOffT<I> oj = axis_off_t<I>(j, G.esj, (I)G.esk);
  const AxisOffT<I> ok = axis_off_t<I>(k, G.esk, (I)1);
  /* source node and plane offsets in the source buffer's slots */
  I xs = x;
  AxisOffT<I> os = oi;
  if (M.sring) {
    const int si = plane_slot(i, M.sring);
    const int im = i > 0 ? i - 1 : G.esi - 1, ip = i < G.esi - 1 ? i + 1 : 0;
    xs = (I)si * (I)G.plane + (I)t;
    os.m = (I)(plane_slot(im, M.sring) - si) * (I)G.plane;
    os.p = (I)(plane_slot(ip, M.sring) - si) * (I)G.plane;
  }
  const I xd = M.dring ? (I)plane_slot(i, M.dring) * (I)G.plane + (I)t : x;
  const unsigned fl = flags[x];
  const bool solid = (fl & FLAG_SOLID) != 0;
  T v[15];
  if (G.nt_load) {
    v[0] = __builtin_nontemporal_load(&src[pidx_t<I>(0, xs)]);
    Links<1, 15>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      const bool bb = (fl >> (l - 1)) & 1u;
      const I idx = bb ? pidx_t<I>(OPP(l), xs) : pidx_t<I>(l, xs + nb_off<l, -1>(os, oj, ok));
      v[l] = __builtin_nontemporal_load(&src[idx]);
    });
  } else {
    v[0] = src[pidx_t<I>(0, xs)];
    Links<1, 15>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      const bool bb = (fl >> (l - 1)) & 1u;
      const I idx = bb ? pidx_t<I>(OPP(l), xs) : pidx_t<I>(l, xs + nb_off<l, -1>(os, oj, ok));
      v[l] = src[idx];
    });
  }
  if ((MOVING || FORCE) && (fl & FLAG_LINKS)) {
    double Fx = 0.0, Fy = 0.0, Fz = 0.0;
    Links<1, 15>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      if ((fl >> (l - 1)) & 1u) {
        double cu = 0.0;
        if (MOVING) {
          const I s = solid ? x : x + nb_off<l, -1>(oi, oj, ok);
          cu = (double)WL(l) * ((CX(l) ? CX(l) * ux[s] : 0.0) + (CY(l) ? CY(l) * uy[s] : 0.0) +
                                (CZ(l) ? CZ(l) * uz[s] : 0.0));
        }
        if (FORCE && solid) {
          /* link m = lbar of solid x: a = a_m(x), b = a_l(x - c_l) */
          const double a = (double)v[l];
          const double b = (double)src[pidx_t<I>(l, xs + nb_off<l, -1>(os, oj, ok))];
          const double g = -2.0 * (a - b + cu);
          if (CX(l)) Fx += CX(l) * g;
          if (CY(l)) Fy += CY(l) * g;
          if (CZ(l)) Fz += CZ(l) * g;
        }
        if (MOVING) v[l] = (T)((double)v[l] + cu);
      }
    });
    if (FORCE && solid) {
      fx[x] += Fx; fy[x] += Fy; fz[x] += Fz;
    }
  }
  if (!(P.skip_solid && solid)) {
    T jbx = T(0), jby = T(0), jbz = T(0);
    if (BF != 0 && !solid) {
      jbx = T(P.bfx); jby = T(P.bfy); jbz = T(P.bfz);
      if (BF == 2) { jbx += bx[x]; jby += by[x]; jbz += bz[x]; }
    }
    collide_node<T>(v, jbx, jby, jbz, P);
  }
  if (G.nt_store) {
#pragma unroll
    for (int l = 0; l < 15; l++) __builtin_nontemporal_store(v[l], &dst[pidx_t<I>(l, xd)]);
  } else {
#pragma unroll
    for (int l = 0; l < 15; l++) dst[pidx_t<I>(l, xd)] = v[l];
  }
}

/* The step with the k pulls staged in LDS.  For rows whose length esk
 * divides the 512-thread block (64, 128, 256 or 512 nodes), a block holds
 * whole rows, and the population a_l pulled from x - c_l lies in row
 * (i - cx, j - cy) at k - cz (periodic within the row).  The 10 links with
 * cz != 0 are loaded aligned -- each thread loads the source row's value at
 * its own k -- into LDS, and each thread reads its value at k - cz there, so
 * every global load of a wavefront covers whole 64-value stretches instead
 * of one shifted by an element (two 128-byte lines per wave and population
 * instead of three).  Links with cz = 0 and bounced links (a node's own
 * opposite population) load from global memory as in k_step.  Same values,
 * same arithmetic: bit-identical to k_step.  (512-node segments of longer
 * rows, with the two values beyond a segment's ends staged by its end
 * threads, measured no faster than k_step at 1024^3 and slower at rows of
 * 2048: 13.7 against 12.8 ms at 512 x 512 x 2048.) */
template <typename T, bool MOVING, bool FORCE, int BF, typename I>
__global__ void __launch_bounds__(512)
k_step_lds(const T *__restrict__ src, T *__restrict__ dst,
           const uint16_t *__restrict__ flags,
           const double *__restrict__ ux, const double *__restrict__ uy,
           const double *__restrict__ uz,
           const T *__restrict__ bx, const T *__restrict__ by, const T *__restrict__ bz,
           double *__restrict__ fx, double *__restrict__ fy, double *__restrict__ fz,
           Geom G, PhysArgs P) {
  __shared__ T row[10][512];
  int blk_x, blk_y;
  step_block(G, blk_x, blk_y);
  const int t = blk_x * 512 + threadIdx.x;          /* plane % 512 == 0: all in range */
  const int i = blk_y;
  const int kmask = G.esk - 1;
  const int j = t / G.esk;
  const int k = t & kmask;
  const I x = (I)i * (I)G.plane + (I)t;
  const AxisOffT<I> oi = axis_off_t<I>(i, G.esi, (I)G.plane);
  const AxisOffT<I> oj = axis_off_t<I>(j, G.esj, (I)G.esk);
  const AxisOffT<I> ok = axis_off_t<I>(k, G.esk, (I)1);
  const AxisOffT<I> o0 = {0, 0};
  const unsigned fl = flags[x];
  const bool solid = (fl & FLAG_SOLID) != 0;
  T v[15];
  /* FP32: the wave's loads issue at raised priority until the barrier
   * (357.6 against 359.6 us per step at 256^3; FP64 measured 0.8% slower
   * with it, so it stays at the default there) */
  if (sizeof(T) == 4) __builtin_amdgcn_s_setprio(3);
  /* stage: the source row of each cz != 0 link, aligned (no k shift) */
  Links<5, 15>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    row[l - 5][threadIdx.x] = src[pidx_t<I>(l, x + nb_off<l, -1>(oi, oj, o0))];
  });
  v[0] = src[pidx_t<I>(0, x)];
  Links<1, 5>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    const bool bb = (fl >> (l - 1)) & 1u;
    v[l] = src[bb ? pidx_t<I>(OPP(l), x) : pidx_t<I>(l, x + nb_off<l, -1>(oi, oj, ok))];
  });
  T own[15];
  Links<5, 15>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    own[l] = ((fl >> (l - 1)) & 1u) ? src[pidx_t<I>(OPP(l), x)] : T(0);
  });
  __syncthreads();
  if (sizeof(T) == 4) __builtin_amdgcn_s_setprio(0);
  const int rbase = threadIdx.x - k;      /* first thread of this row in the block */
  Links<5, 15>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    const T pulled = row[l - 5][rbase + ((k - CZ(l)) & kmask)];
    v[l] = ((fl >> (l - 1)) & 1u) ? own[l] : pulled;
    if (FORCE) own[l] = pulled;   /* the force needs a_l(x - c_l) as well */
  });
  if ((MOVING || FORCE) && (fl & FLAG_LINKS)) {
    double Fx = 0.0, Fy = 0.0, Fz = 0.0;
    Links<1, 15>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      if ((fl >> (l - 1)) & 1u) {
        double cu = 0.0;
        if (MOVING) {
          const I s = solid ? x : x + nb_off<l, -1>(oi, oj, ok);
          cu = (double)WL(l) * ((CX(l) ? CX(l) * ux[s] : 0.0) + (CY(l) ? CY(l) * uy[s] : 0.0) +
                                (CZ(l) ? CZ(l) * uz[s] : 0.0));
        }
        if (FORCE && solid) {
          const double a = (double)v[l];
          const double b = CZ(l) ? (double)own[l]
                                 : (double)src[pidx_t<I>(l, x + nb_off<l, -1>(oi, oj, ok))];
          const double g = -2.0 * (a - b + cu);
          if (CX(l)) Fx += CX(l) * g;
          if (CY(l)) Fy += CY(l) * g;
          if (CZ(l)) Fz += CZ(l) * g;
        }
        if (MOVING) v[l] = (T)((double)v[l] + cu);
      }
    });
    if (FORCE && solid) {
      fx[x] += Fx; fy[x] += Fy; fz[x] += Fz;
    }
  }
  if (!(P.skip_solid && solid)) {
    T jbx = T(0), jby = T(0), jbz = T(0);
    if (BF != 0 && !solid) {
      jbx = T(P.bfx); jby = T(P.bfy); jbz = T(P.bfz);
      if (BF == 2) { jbx += bx[x]; jby += by[x]; jbz += bz[x]; }
    }
    collide_node<T>(v, jbx, jby, jbz, P);
  }
  if (G.nt_store) {
#pragma unroll
    for (int l = 0; l < 15; l++) __builtin_nontemporal_store(v[l], &dst[pidx_t<I>(l, x)]);
  } else {
#pragma unroll
    for (int l = 0; l < 15; l++) dst[pidx_t<I>(l, x)] = v[l];
  }
}

/* The staged step for rows of any length esk <= 512: a block holds rpb
 * whole rows (rpb divides esj, rpb * esk <= 512) and as many threads,
 * rounded up to whole wavefronts; threads past rpb * esk only take part in
 * the barrier.  The staging and the wrap of k - cz within the row are those
 * of k_step_lds with the row length a runtime value (compares instead of a
 * mask), the LDS sized to the block (10 values per thread).  Same values,
 * same arithmetic: bit-identical to k_step. */
template <typename T, bool MOVING, bool FORCE, int BF, typename I>
__global__ void __launch_bounds__(512)
k_step_ldsr(const T *__restrict__ src, T *__restrict__ dst,
            const uint16_t *__restrict__ flags,
            const double *__restrict__ ux, const double *__restrict__ uy,
            const double *__restrict__ uz,
            const T *__restrict__ bx, const T *__restrict__ by, const T *__restrict__ bz,
            double *__restrict__ fx, double *__restrict__ fy, double *__restrict__ fz,
            Geom G, PhysArgs P, int rpb) {
  extern __shared__ __attribute__((aligned(16))) unsigned char ldsr_raw[];
  T *row = (T *)ldsr_raw;                 /* [10][blockDim.x] */
  const int bs = blockDim.x;
  int blk_x, blk_y;
  step_block(G, blk_x, blk_y);
  const int esk = G.esk;
  const int r = threadIdx.x / esk;        /* row within the block */
  const int k = threadIdx.x - r * esk;
  const int j = blk_x * rpb + r;
  const bool act = r < rpb;               /* rpb divides esj: j < esj */
  const int i = blk_y;
  const I x = (I)i * (I)G.plane + (I)j * (I)esk + (I)k;
  const AxisOffT<I> oi = axis_off_t<I>(i, G.esi, (I)G.plane);
  const AxisOffT<I> oj = axis_off_t<I>(act ? j : 0, G.esj, (I)esk);
  const AxisOffT<I> ok = axis_off_t<I>(k, esk, (I)1);
  const AxisOffT<I> o0 = {0, 0};
  const unsigned fl = act ? flags[x] : 0u;
  const bool solid = (fl & FLAG_SOLID) != 0;
  T v[15];
  T own[15];
  if (act) {
    Links<5, 15>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      row[(l - 5) * bs + threadIdx.x] = src[pidx_t<I>(l, x + nb_off<l, -1>(oi, oj, o0))];
    });
    v[0] = src[pidx_t<I>(0, x)];
    Links<1, 5>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      const bool bb = (fl >> (l - 1)) & 1u;
      v[l] = src[bb ? pidx_t<I>(OPP(l), x) : pidx_t<I>(l, x + nb_off<l, -1>(oi, oj, ok))];
    });
    Links<5, 15>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      own[l] = ((fl >> (l - 1)) & 1u) ? src[pidx_t<I>(OPP(l), x)] : T(0);
    });
  }
  __syncthreads();
  if (!act) return;
  const int rbase = threadIdx.x - k;
  Links<5, 15>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    int kw = k - CZ(l);
    kw += (kw < 0) ? esk : 0;
    kw -= (kw >= esk) ? esk : 0;
    const T pulled = row[(l - 5) * bs + rbase + kw];
    v[l] = ((fl >> (l - 1)) & 1u) ? own[l] : pulled;
    if (FORCE) own[l] = pulled;
  });
  if ((MOVING || FORCE) && (fl & FLAG_LINKS)) {
    double Fx = 0.0, Fy = 0.0, Fz = 0.0;
    Links<1, 15>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      if ((fl >> (l - 1)) & 1u) {
        double cu = 0.0;
        if (MOVING) {
          const I s = solid ? x : x + nb_off<l, -1>(oi, oj, ok);
          cu = (double)WL(l) * ((CX(l) ? CX(l) * ux[s] : 0.0) + (CY(l) ? CY(l) * uy[s] : 0.0) +
                                (CZ(l) ? CZ(l) * uz[s] : 0.0));
        }
        if (FORCE && solid) {
          const double a = (double)v[l];
          const double b = CZ(l) ? (double)own[l]
                                 : (double)src[pidx_t<I>(l, x + nb_off<l, -1>(oi, oj, ok))];
          const double g = -2.0 * (a - b + cu);
          if (CX(l)) Fx += CX(l) * g;
          if (CY(l)) Fy += CY(l) * g;
          if (CZ(l)) Fz += CZ(l) * g;
        }
        if (MOVING) v[l] = (T)((double)v[l] + cu);
      }
    });
    if (FORCE && solid) {
      fx[x] += Fx; fy[x] += Fy; fz[x] += Fz;
    }
  }
  if (!(P.skip_solid && solid)) {
    T jbx = T(0), jby = T(0), jbz = T(0);
    if (BF != 0 && !solid) {
      jbx = T(P.bfx); jby = T(P.bfy); jbz = T(P.bfz);
      if (BF == 2) { jbx += bx[x]; jby += by[x]; jbz += bz[x]; }
    }
    collide_node<T>(v, jbx, jby, jbz, P);
  }
  if (G.nt_store) {
#pragma unroll
    for (int l = 0; l < 15; l++) __builtin_nontemporal_store(v[l], &dst[pidx_t<I>(l, x)]);
  } else {
#pragma unroll
    for (int l = 0; l < 15; l++) dst[pidx_t<I>(l, x)] = v[l];
  }
}

/* Moment storage.  After a collision a node's 15 populations are the
 * rebuild of its 10 post-collision moments (rebuild_link), so a run of
 * steps can keep the moments instead: 10 values per node in tiles of 1024
 * nodes x 10 runs, read and written instead of 15 populations.  The step
 * rebuilds each pulled a_l(x - c_l) from the source node's moments (links
 * grouped by the source node: own, x -+ 1, y -+ 1 and the four diagonal
 * sources), the k-shifted links through LDS as in k_step_lds, and bounced
 * links from the node's own moments; the collision is collide_node's.  The
 * first step of a run reads populations (IN_POPS), the last writes them
 * (OUT_POPS).  Every pulled value equals the population the plain step
 * would have stored and loaded, so the populations after the run are
 * bit-identical to k_step's.  Needs rows of 64-512 nodes dividing 512, no
 * skip_solid (a skipped solid node keeps 15 independent populations) and no
 * force accumulation. */
template <typename I>
__host__ __device__ __forceinline__ I midx_t(int m, I x) {
  return ((x >> POP_SHIFT) * I(10) + I(m)) * I(POP_TILE) + (x & I(POP_TILE - 1));
}

template <typename T, typename I>
__device__ __forceinline__ Mom<T> load_mom(const T *__restrict__ M, I y) {
  Mom<T> m;
  m.drho = M[midx_t<I>(0, y)];
  m.jx = M[midx_t<I>(1, y)];
  m.jy = M[midx_t<I>(2, y)];
  m.jz = M[midx_t<I>(3, y)];
  m.sxx = M[midx_t<I>(4, y)];
  m.syy = M[midx_t<I>(5, y)];
  m.szz = M[midx_t<I>(6, y)];
  m.sxy = M[midx_t<I>(7, y)];
  m.sxz = M[midx_t<I>(8, y)];
  m.syz = M[midx_t<I>(9, y)];
  return m;
}

/* the moments an axis link's rebuild reads (drho, the trace, one j) */
template <typename T, typename I, int AX>
__device__ __forceinline__ Mom<T> load_mom_axis(const T *__restrict__ M, I y) {
  Mom<T> m;
  m.drho = M[midx_t<I>(0, y)];
  m.jx = (AX == 0) ? M[midx_t<I>(1, y)] : T(0);
  m.jy = (AX == 1) ? M[midx_t<I>(2, y)] : T(0);
  m.jz = (AX == 2) ? M[midx_t<I>(3, y)] : T(0);
  m.sxx = M[midx_t<I>(4, y)];
  m.syy = M[midx_t<I>(5, y)];
  m.szz = M[midx_t<I>(6, y)];
  m.sxy = m.sxz = m.syz = T(0);
  return m;
}

template <typename T, bool MOVING, int BF, typename I, bool IN_POPS, bool OUT_POPS>
__global__ void __launch_bounds__(512)
k_step_mom(const T *__restrict__ srcP, const T *__restrict__ srcM, T *__restrict__ dstP,
           T *__restrict__ dstM, const uint16_t *__restrict__ flags,
           const double *__restrict__ ux, const double *__restrict__ uy,
           const double *__restrict__ uz,
           const T *__restrict__ bx, const T *__restrict__ by, const T *__restrict__ bz,
           Geom G, PhysArgs P) {
  __shared__ T row[10][512];
  int blk_x, blk_y;
  step_block(G, blk_x, blk_y);
  const int t = blk_x * 512 + threadIdx.x;          /* plane % 512 == 0: all in range */
  const int i = blk_y;
  const int kmask = G.esk - 1;
  const int j = t / G.esk;
  const int k = t & kmask;
  const I x = (I)i * (I)G.plane + (I)t;
  const AxisOffT<I> oi = axis_off_t<I>(i, G.esi, (I)G.plane);
  const AxisOffT<I> oj = axis_off_t<I>(j, G.esj, (I)G.esk);
  const AxisOffT<I> ok = axis_off_t<I>(k, G.esk, (I)1);
  const AxisOffT<I> o0 = {0, 0};
  const unsigned fl = flags[x];
  const bool solid = (fl & FLAG_SOLID) != 0;
  T v[15];
  T own[15];
  if constexpr (IN_POPS) {
    Links<5, 15>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      row[l - 5][threadIdx.x] = srcP[pidx_t<I>(l, x + nb_off<l, -1>(oi, oj, o0))];
    });
    v[0] = srcP[pidx_t<I>(0, x)];
    Links<1, 5>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      const bool bb = (fl >> (l - 1)) & 1u;
      v[l] = srcP[bb ? pidx_t<I>(OPP(l), x) : pidx_t<I>(l, x + nb_off<l, -1>(oi, oj, ok))];
    });
    Links<5, 15>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      own[l] = ((fl >> (l - 1)) & 1u) ? srcP[pidx_t<I>(OPP(l), x)] : T(0);
    });
  } else {
    const Mom<T> mo = load_mom<T, I>(srcM, x);
    /* diagonal sources (cx, cy) = (+1,+1): links 7, 11; (-1,+1): 8, 12;
     * (+1,-1): 9, 13; (-1,-1): 10, 14 -- at the node's own k */
    {
      const Mom<T> m = load_mom<T, I>(srcM, x + nb_off<7, -1>(oi, oj, o0));
      row[7 - 5][threadIdx.x] = rebuild_link<7>(m);
      row[11 - 5][threadIdx.x] = rebuild_link<11>(m);
    }
    {
      const Mom<T> m = load_mom<T, I>(srcM, x + nb_off<8, -1>(oi, oj, o0));
      row[8 - 5][threadIdx.x] = rebuild_link<8>(m);
      row[12 - 5][threadIdx.x] = rebuild_link<12>(m);
    }
    {
      const Mom<T> m = load_mom<T, I>(srcM, x + nb_off<9, -1>(oi, oj, o0));
      row[9 - 5][threadIdx.x] = rebuild_link<9>(m);
      row[13 - 5][threadIdx.x] = rebuild_link<13>(m);
    }
    {
      const Mom<T> m = load_mom<T, I>(srcM, x + nb_off<10, -1>(oi, oj, o0));
      row[10 - 5][threadIdx.x] = rebuild_link<10>(m);
      row[14 - 5][threadIdx.x] = rebuild_link<14>(m);
    }
    row[5 - 5][threadIdx.x] = rebuild_link<5>(mo);
    row[6 - 5][threadIdx.x] = rebuild_link<6>(mo);
    v[0] = rebuild_link<0>(mo);
    {
      const bool b1 = fl & 1u, b2 = (fl >> 1) & 1u, b3 = (fl >> 2) & 1u, b4 = (fl >> 3) & 1u;
      const Mom<T> m1 = load_mom_axis<T, I, 0>(srcM, x + nb_off<1, -1>(oi, oj, ok));
      const Mom<T> m2 = load_mom_axis<T, I, 0>(srcM, x + nb_off<2, -1>(oi, oj, ok));
      const Mom<T> m3 = load_mom_axis<T, I, 1>(srcM, x + nb_off<3, -1>(oi, oj, ok));
      const Mom<T> m4 = load_mom_axis<T, I, 1>(srcM, x + nb_off<4, -1>(oi, oj, ok));
      v[1] = b1 ? rebuild_link<2>(mo) : rebuild_link<1>(m1);
      v[2] = b2 ? rebuild_link<1>(mo) : rebuild_link<2>(m2);
      v[3] = b3 ? rebuild_link<4>(mo) : rebuild_link<3>(m3);
      v[4] = b4 ? rebuild_link<3>(mo) : rebuild_link<4>(m4);
    }
    Links<5, 15>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      own[l] = ((fl >> (l - 1)) & 1u) ? rebuild_link<OPP(l)>(mo) : T(0);
    });
  }
  __syncthreads();
  const int rbase = threadIdx.x - k;
  Links<5, 15>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    const T pulled = row[l - 5][rbase + ((k - CZ(l)) & kmask)];
    v[l] = ((fl >> (l - 1)) & 1u) ? own[l] : pulled;
  });
  if (MOVING && (fl & FLAG_LINKS)) {
    Links<1, 15>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      if ((fl >> (l - 1)) & 1u) {
        const I s = solid ? x : x + nb_off<l, -1>(oi, oj, ok);
        const double cu =
            (double)WL(l) * ((CX(l) ? CX(l) * ux[s] : 0.0) + (CY(l) ? CY(l) * uy[s] : 0.0) +
                             (CZ(l) ? CZ(l) * uz[s] : 0.0));
        v[l] = (T)((double)v[l] + cu);
      }
    });
  }
  T jbx = T(0), jby = T(0), jbz = T(0);
  if (BF != 0 && !solid) {
    jbx = T(P.bfx); jby = T(P.bfy); jbz = T(P.bfz);
    if (BF == 2) { jbx += bx[x]; jby += by[x]; jbz += bz[x]; }
  }
  const Mom<T> mn = collide_moments<T>(v, jbx, jby, jbz, P);
  if constexpr (OUT_POPS) {
    rebuild_all<T>(mn, v);
#pragma unroll
    for (int l = 0; l < 15; l++) __builtin_nontemporal_store(v[l], &dstP[pidx_t<I>(l, x)]);
  } else {
    __builtin_nontemporal_store(mn.drho, &dstM[midx_t<I>(0, x)]);
    __builtin_nontemporal_store(mn.jx, &dstM[midx_t<I>(1, x)]);
    __builtin_nontemporal_store(mn.jy, &dstM[midx_t<I>(2, x)]);
    __builtin_nontemporal_store(mn.jz, &dstM[midx_t<I>(3, x)]);
    __builtin_nontemporal_store(mn.sxx, &dstM[midx_t<I>(4, x)]);
    __builtin_nontemporal_store(mn.syy, &dstM[midx_t<I>(5, x)]);
    __builtin_nontemporal_store(mn.szz, &dstM[midx_t<I>(6, x)]);
    __builtin_nontemporal_store(mn.sxy, &dstM[midx_t<I>(7, x)]);
    __builtin_nontemporal_store(mn.sxz, &dstM[midx_t<I>(8, x)]);
    __builtin_nontemporal_store(mn.syz, &dstM[midx_t<I>(9, x)]);
  }
}

/* ---------------------------------------------------------------------- */
/* Separate phases (bbl, propagate, collide) on the live buffer.           */

/* bbl / bblforce in place: src/lbkernel.c:58-618 restated per solid node. */
template <typename T, bool MOVING, bool FORCE>
__global__ void __launch_bounds__(256)
k_bbl(T *__restrict__ a, const int32_t *__restrict__ mask,
      const double *__restrict__ ux, const double *__restrict__ uy,
      const double *__restrict__ uz, double *__restrict__ fx,
      double *__restrict__ fy, double *__restrict__ fz, Geom G) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.plane) return;
  const int i = blockIdx.y, j = t / G.esk, k = t - j * G.esk;
  const long x = (long)i * G.plane + t;
  const int m = mask[x];
  if (m <= 0) return;
  const AxisOff oi = axis_off(i, G.esi, G.plane);
  const AxisOff oj = axis_off(j, G.esj, G.esk);
  const AxisOff ok = axis_off(k, G.esk, 1);
  double Fx = 0.0, Fy = 0.0, Fz = 0.0;
  Links<1, 15>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    if ((m >> (l - 1)) & 1) {
      const long f = x + nb_off<l, +1>(oi, oj, ok);
      const double av = (double)a[pidx(l, x)];
      const double bv = (double)a[pidx(OPP(l), f)];
      double fac = 0.0;
      if (MOVING)
        fac = (double)WL(l) * ((CX(l) ? CX(l) * ux[x] : 0.0) + (CY(l) ? CY(l) * uy[x] : 0.0) +
                               (CZ(l) ? CZ(l) * uz[x] : 0.0));
      a[pidx(l, x)] = (T)(bv + fac);
      a[pidx(OPP(l), f)] = (T)(av - fac);
      if (FORCE) {
        const double g = 2.0 * (av - bv - fac);
        if (CX(l)) Fx += CX(l) * g;
        if (CY(l)) Fy += CY(l) * g;
        if (CZ(l)) Fz += CZ(l) * g;
      }
    }
  });
  if (FORCE) { fx[x] += Fx; fy[x] += Fy; fz[x] += Fz; }
}

/* propagate: dst_l(x) = src_l(x - c_l), src/lbkernel.c:623-762. */
template <typename T>
__global__ void __launch_bounds__(256)
k_propagate(const T *__restrict__ src, T *__restrict__ dst, Geom G) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.plane) return;
  const int i = blockIdx.y, j = t / G.esk, k = t - j * G.esk;
  const long x = (long)i * G.plane + t;
  const AxisOff oi = axis_off(i, G.esi, G.plane);
  const AxisOff oj = axis_off(j, G.esj, G.esk);
  const AxisOff ok = axis_off(k, G.esk, 1);
  dst[pidx(0, x)] = src[pidx(0, x)];
  Links<1, 15>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    dst[pidx(l, x)] = src[pidx(l, x + nb_off<l, -1>(oi, oj, ok))];
  });
}

/* collide in place, src/lbkernel.c:769-877. */
template <typename T, int BF>
__global__ void __launch_bounds__(256)
k_collide(T *__restrict__ a, const uint16_t *__restrict__ flags,
          const T *__restrict__ bx, const T *__restrict__ by, const T *__restrict__ bz,
          Geom G, PhysArgs P) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= G.N) return;
  const bool solid = (flags[x] & FLAG_SOLID) != 0;
  if (P.skip_solid && solid) return;
  T v[15];
#pragma unroll
  for (int l = 0; l < 15; l++) v[l] = a[pidx(l, x)];
  T jbx = T(0), jby = T(0), jbz = T(0);
  if (BF != 0 && !solid) {
    jbx = T(P.bfx); jby = T(P.bfy); jbz = T(P.bfz);
    if (BF == 2) { jbx += bx[x]; jby += by[x]; jbz += bz[x]; }
  }
  collide_node<T>(v, jbx, jby, jbz, P);
#pragma unroll
  for (int l = 0; l < 15; l++) a[pidx(l, x)] = v[l];
}

/* ---------------------------------------------------------------------- */
/* Geometry tables.                                                        */

/* makemask, src/lbobjs.c:526-560: bit l-1 on a solid node iff the node at
 * +c_l is fluid.  Bit-exact integer work. */
__global__ void __launch_bounds__(256)
k_makemask(const int32_t *__restrict__ objidx, int32_t *__restrict__ mask, Geom G) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.plane) return;
  const int i = blockIdx.y, j = t / G.esk, k = t - j * G.esk;
  const long x = (long)i * G.plane + t;
  int m = 0;
  if (objidx[x] >= 0) {
    const AxisOff oi = axis_off(i, G.esi, G.plane);
    const AxisOff oj = axis_off(j, G.esj, G.esk);
    const AxisOff ok = axis_off(k, G.esk, 1);
    Links<1, 15>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      if (objidx[x + nb_off<l, +1>(oi, oj, ok)] < 0) m |= 1 << (l - 1);
    });
  }
  mask[x] = m;
}

/* countmask (src/lbobjs.c:671-714) over the device mask: per node with
 * mask > 0, the weights w_l of its link bits (links) or the number of its
 * face bits l = 1..6 (faces); exact integer sums. */
struct CountW {
  int w[15];
};
__global__ void __launch_bounds__(256)
k_countmask(const int32_t *__restrict__ mask, long N, int faces, CountW W,
            unsigned long long *__restrict__ out) {
  __shared__ unsigned red[256];
  unsigned v = 0;
  for (long x = (long)blockIdx.x * 256 + threadIdx.x; x < N; x += (long)gridDim.x * 256) {
    const int m = mask[x];
    if (m > 0) {
      if (faces) {
        v += (unsigned)__popc((unsigned)m & 0x3fu);
      } else {
        for (int l = 1; l < 15; l++)
          if (m & (1 << (l - 1))) v += (unsigned)W.w[l];
      }
    }
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0] != 0) atomicAdd(out, (unsigned long long)red[0]);
}

/* Pull flags: bit l-1 = bounce along l at x, i.e. bit l-1 of mask(x - c_l)
 * or bit lbar-1 of mask(x); bit 14 = solid. */
__global__ void __launch_bounds__(256)
k_flags(const int32_t *__restrict__ objidx, const int32_t *__restrict__ mask,
        uint16_t *__restrict__ flags, Geom G) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.plane) return;
  const int i = blockIdx.y, j = t / G.esk, k = t - j * G.esk;
  const long x = (long)i * G.plane + t;
  const AxisOff oi = axis_off(i, G.esi, G.plane);
  const AxisOff oj = axis_off(j, G.esj, G.esk);
  const AxisOff ok = axis_off(k, G.esk, 1);
  const int mx = mask[x];
  unsigned f = (objidx[x] >= 0) ? FLAG_SOLID : 0u;
  Links<1, 15>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    const int ms = mask[x + nb_off<l, -1>(oi, oj, ok)];
    if (((ms >> (l - 1)) & 1) || ((mx >> (OPP(l) - 1)) & 1)) f |= 1u << (l - 1);
  });
  flags[x] = (uint16_t)f;
}

__global__ void __launch_bounds__(256)
k_vspread(const int32_t *__restrict__ objidx, const double *__restrict__ tab,
          double rho0, double *__restrict__ ux, double *__restrict__ uy,
          double *__restrict__ uz, Geom G) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.plane) return;
  const int i = blockIdx.y, j = t / G.esk, k = t - j * G.esk;
  const long x = (long)i * G.plane + t;
  const int o = objidx[x];
  double U0 = 0.0, U1 = 0.0, U2 = 0.0;
  if (o >= 0) obj_velocity(tab, o, i, j, k, rho0, U0, U1, U2);
  ux[x] = U0; uy[x] = U1; uz[x] = U2;
}

/* ---------------------------------------------------------------------- */
/* Reductions.                                                             */

#define RED_BLOCKS 1024

template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&s)[NV], double *__restrict__ part) {
  __shared__ double red[NV][256 / 64];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double v = s[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    s[q] = v;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) red[q][wv] = s[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) {
      double v = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) v += red[q][w];
      part[(long)blockIdx.x * NV + q] = v;
    }
  }
}

/* Final pass: one block sums nparts rows of NV doubles in a fixed order. */
template <int NV>
__global__ void __launch_bounds__(256)
k_final(const double *__restrict__ part, int nparts, double *__restrict__ out) {
  double s[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) s[q] = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x)
#pragma unroll
    for (int q = 0; q < NV; q++) s[q] += part[(long)b * NV + q];
  __shared__ double red[NV][256 / 64];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double v = s[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    s[q] = v;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; q++) red[q][wv] = s[q];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NV; q++) {
      double v = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) v += red[q][w];
      out[q] = v;
    }
}

/* Node momentum and density for measurement (FP64 from T populations). */
template <typename T>
__device__ __forceinline__ void node_rho_j(const T *__restrict__ a, long x, long N,
                                           double &d, double &jx, double &jy, double &jz) {
  double v[15];
#pragma unroll
  for (int l = 0; l < 15; l++) v[l] = (double)a[pidx(l, x)];
  d = 0.0; jx = 0.0; jy = 0.0; jz = 0.0;
#pragma unroll
  for (int l = 0; l < 15; l++) {
    d += v[l];
    if (CX(l)) jx += CX(l) * v[l];
    if (CY(l)) jy += CY(l) * v[l];
    if (CZ(l)) jz += CZ(l) * v[l];
  }
}

/* accvmean sums, src/lbmeasure.c:58-106. */
template <typename T, int BF>
__global__ void __launch_bounds__(256)
k_vmean(const T *__restrict__ a, const uint16_t *__restrict__ flags,
        const T *__restrict__ bx, const T *__restrict__ by, const T *__restrict__ bz,
        const double *__restrict__ psi, double fac, Geom G, PhysArgs P,
        double *__restrict__ part) {
  double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const long N = G.N;
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < N;
       x += (long)gridDim.x * blockDim.x) {
    if (flags[x] & FLAG_SOLID) continue;
    double d, jx, jy, jz;
    node_rho_j<T>(a, x, N, d, jx, jy, jz);
    double b0 = 0.0, b1 = 0.0, b2 = 0.0;
    if (BF != 0) {
      b0 = P.bfx; b1 = P.bfy; b2 = P.bfz;
      if (BF == 2) { b0 += (double)bx[x]; b1 += (double)by[x]; b2 += (double)bz[x]; }
    }
    const double rho = P.rho0 + d;
    double w = 1.0;
    if (psi) { w = psi[x]; w *= w; }
    s[0] += w * ((jx + fac * b0) / rho);
    s[1] += w * ((jy + fac * b1) / rho);
    s[2] += w * ((jz + fac * b2) / rho);
    s[3] += w;
    s[4] += 1.0;
  }
  block_reduce_store<5>(s, part);
}

/* vgrab, src/lbmeasure.c:128-166; vzero :110-121. */
template <typename T, int BF>
__global__ void __launch_bounds__(256)
k_vgrab(const T *__restrict__ a, const uint16_t *__restrict__ flags,
        const T *__restrict__ bx, const T *__restrict__ by, const T *__restrict__ bz,
        double fac, double prefac, Geom G, PhysArgs P, double *__restrict__ ux,
        double *__restrict__ uy, double *__restrict__ uz, int zero) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= G.N) return;
  if (flags[x] & FLAG_SOLID) return;
  if (zero) { ux[x] = 0.0; uy[x] = 0.0; uz[x] = 0.0; return; }
  double d, jx, jy, jz;
  node_rho_j<T>(a, x, G.N, d, jx, jy, jz);
  double b0 = 0.0, b1 = 0.0, b2 = 0.0;
  if (BF != 0) {
    b0 = P.bfx; b1 = P.bfy; b2 = P.bfz;
    if (BF == 2) { b0 += (double)bx[x]; b1 += (double)by[x]; b2 += (double)bz[x]; }
  }
  const double rho = P.rho0 + d;
  ux[x] += prefac * (jx + fac * b0) / rho;
  uy[x] += prefac * (jy + fac * b1) / rho;
  uz[x] += prefac * (jz + fac * b2) / rho;
}

/* Velocity residual: sums of |u - u_prev|^2 and |u|^2 over fluid nodes,
 * u = (j - b/2)/rho post-collide; u_prev <- u. */
template <typename T, int BF>
__global__ void __launch_bounds__(256)
k_residual(const T *__restrict__ a, const uint16_t *__restrict__ flags,
           const T *__restrict__ bx, const T *__restrict__ by, const T *__restrict__ bz,
           Geom G, PhysArgs P, double *__restrict__ up, double *__restrict__ part) {
  double s[2] = {0.0, 0.0};
  const long N = G.N;
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < N;
       x += (long)gridDim.x * blockDim.x) {
    if (flags[x] & FLAG_SOLID) continue;
    double d, jx, jy, jz;
    node_rho_j<T>(a, x, N, d, jx, jy, jz);
    double b0 = 0.0, b1 = 0.0, b2 = 0.0;
    if (BF != 0) {
      b0 = P.bfx; b1 = P.bfy; b2 = P.bfz;
      if (BF == 2) { b0 += (double)bx[x]; b1 += (double)by[x]; b2 += (double)bz[x]; }
    }
    const double rho = P.rho0 + d;
    const double u0 = (jx - 0.5 * b0) / rho, u1 = (jy - 0.5 * b1) / rho, u2 = (jz - 0.5 * b2) / rho;
    const double e0 = u0 - up[x], e1 = u1 - up[N + x], e2 = u2 - up[2 * N + x];
    s[0] += e0 * e0 + e1 * e1 + e2 * e2;
    s[1] += u0 * u0 + u1 * u1 + u2 * u2;
    up[x] = u0; up[N + x] = u1; up[2 * N + x] = u2;
  }
  block_reduce_store<2>(s, part);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_sum_pops(const T *__restrict__ a, const uint16_t *__restrict__ flags, Geom G,
           double *__restrict__ part) {
  double s[15];
#pragma unroll
  for (int l = 0; l < 15; l++) s[l] = 0.0;
  const long N = G.N;
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < N;
       x += (long)gridDim.x * blockDim.x) {
    if (flags[x] & FLAG_SOLID) continue;
#pragma unroll
    for (int l = 0; l < 15; l++) s[l] += (double)a[pidx(l, x)];
  }
  block_reduce_store<15>(s, part);
}

/* distribf, src/lbmeasure.c:236-263: per object F = sum f, T = sum r x f.
 * Small object counts accumulate in LDS first. */
#define DF_LDS_OBJ 96
__global__ void __launch_bounds__(256)
k_distribf(const int32_t *__restrict__ objidx, const double *__restrict__ fx,
           const double *__restrict__ fy, const double *__restrict__ fz,
           const double *__restrict__ tab, int nobj, Geom G,
           double *__restrict__ acc /* 6*nobj */) {
  __shared__ double sacc[DF_LDS_OBJ * 6];
  const bool use_lds = nobj <= DF_LDS_OBJ;
  if (use_lds) {
    for (int q = threadIdx.x; q < nobj * 6; q += blockDim.x) sacc[q] = 0.0;
    __syncthreads();
  }
  const long N = G.N;
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < N;
       x += (long)gridDim.x * blockDim.x) {
    const int o = objidx[x];
    if (o < 0) continue;
    const double f0 = fx[x], f1 = fy[x], f2 = fz[x];
    if (f0 == 0.0 && f1 == 0.0 && f2 == 0.0) continue;
    const int i = (int)(x / G.plane);
    const int rem = (int)(x - (long)i * G.plane);
    const int j = rem / G.esk, k = rem - j * G.esk;
    const double *p = tab + (long)o * SLBD_OBJ_STRIDE;
    const double rx = i - p[0], ry = j - p[1], rz = k - p[2];
    const double t0 = ry * f2 - rz * f1, t1 = rz * f0 - rx * f2, t2 = rx * f1 - ry * f0;
    double *dstp = use_lds ? sacc + 6 * o : acc + 6 * (long)o;
    atomicAdd(dstp + 0, f0); atomicAdd(dstp + 1, f1); atomicAdd(dstp + 2, f2);
    atomicAdd(dstp + 3, t0); atomicAdd(dstp + 4, t1); atomicAdd(dstp + 5, t2);
  }
  if (use_lds) {
    __syncthreads();
    for (int q = threadIdx.x; q < nobj * 6; q += blockDim.x)
      if (sacc[q] != 0.0) atomicAdd(acc + q, sacc[q]);
  }
}

/* ---------------------------------------------------------------------- */
/* Moment propagation (src/mpkernel.c), gather form.                      */

/* mprhoinit, src/mpmisc.c:64-79 */
template <typename T>
__global__ void __launch_bounds__(256)
k_mprho(const T *__restrict__ a, const uint16_t *__restrict__ flags, double rho0,
        T *__restrict__ rho, Geom G) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= G.N) return;
  double r = rho0;
  if (!(flags[x] & FLAG_SOLID)) {
#pragma unroll
    for (int l = 0; l < 15; l++) r += (double)a[pidx(l, x)];
  }
  rho[x] = (T)r;
}

/* MP flags: bit l (0..14) = node x - c_l is fluid; bit 16+l (1..14) = node
 * s = x - c_l is a solid with boundary link l that reflects (mask bit l-1
 * set, adsorption bit 0x4000 clear). */
__global__ void __launch_bounds__(256)
k_mpflags(const int32_t *__restrict__ objidx, const int32_t *__restrict__ mask,
          uint32_t *__restrict__ mpf, Geom G) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.plane) return;
  const int i = blockIdx.y, j = t / G.esk, k = t - j * G.esk;
  const long x = (long)i * G.plane + t;
  const AxisOff oi = axis_off(i, G.esi, G.plane);
  const AxisOff oj = axis_off(j, G.esj, G.esk);
  const AxisOff ok = axis_off(k, G.esk, 1);
  uint32_t f = (objidx[x] < 0) ? 1u : 0u;
  Links<1, 15>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    const long s = x + nb_off<l, -1>(oi, oj, ok);
    if (objidx[s] < 0) {
      f |= 1u << l;
    } else {
      const int m = mask[s];
      if (m && !(m & 0x4000) && ((m >> (l - 1)) & 1)) f |= 1u << (16 + l);
    }
  });
  mpf[x] = f;
}

struct MPArgs {
  double rho0, alpha, pads;
  int axis;
};

/* MP fields live in node records of W = 2 (P, 1/rho) or W = 4 (P, m1, m2,
 * 1/rho) values, so that the gather below reads one vector per incoming link
 * instead of four scalars; 1/rho (the lattice's frozen density) is kept in
 * each record and refreshed whenever the density changes.
 *
 * One moment-propagation step for every node x:
 *   new(x) = sum_l [x - c_l fluid] pi_l(x - c_l) P(x - c_l)
 *          + sum_l [x - c_l reflecting solid] (1 - pads)(pi_lbar(x)
 *                 + w_l c_l.U(x - c_l)/rho(x)) P(x)
 *   pi_l(y) = a_l(y)/rho(y) + (w_l/72)(rho0/rho(y) - alpha) (+ alpha, l = 0)
 * then the update rule of src/mpkernel.c:190-197 (fluid replace, adsorbing
 * site accumulate, other solid clear; partial adsorption keeps pads*new). */
/* FP64 records are 32 bytes: capped at 4 waves per SIMD, the compiler keeps the
 * whole gather in flight instead of interleaving loads and waits to hold
 * eight waves' worth of registers (the FP32 kernel does both) */
template <typename T, int NMOM, bool MOVING, typename I>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(1, sizeof(T) == 8 ? 4 : 8)))
k_mpstep(const T *__restrict__ a, const T *__restrict__ qin, T *__restrict__ qout,
         const uint32_t *__restrict__ mpf, const int32_t *__restrict__ mask,
         const int32_t *__restrict__ objidx, const double *__restrict__ tab,
         Geom G, MPArgs A) {
  constexpr int W = NMOM > 1 ? 4 : 2;
  typedef T V __attribute__((ext_vector_type(W)));
  const V *__restrict__ Q = (const V *)qin;
  /* mp_xcd 1: each XCD walks a contiguous range of planes, so the records of
   * planes i - 1 and i + 1 that a plane gathers were loaded into the same
   * XCD's L2 by the neighbouring planes */
  Geom Gb = G;
  Gb.xcd_map = G.mp_xcd;
  int blk_x, blk_y;
  step_block(Gb, blk_x, blk_y);
  const int t = blk_x * blockDim.x + threadIdx.x;
  if (t >= G.plane) return;
  const int i = blk_y, j = t / G.esk, k = t - j * G.esk;
  const I x = (I)i * (I)G.plane + (I)t;
  const AxisOffT<I> oi = axis_off_t<I>(i, G.esi, (I)G.plane);
  const AxisOffT<I> oj = axis_off_t<I>(j, G.esj, (I)G.esk);
  const AxisOffT<I> ok = axis_off_t<I>(k, G.esk, (I)1);
  const uint32_t f = mpf[x];
  const T r0 = T(A.rho0), al = T(A.alpha);
  const V qx = Q[x];
  const T rix = qx[W - 1];
  T np = T(0), n1 = T(0), n2 = T(0);
  /* rest link: x itself */
  if (f & 1u) {
    const T pr = a[pidx_t<I>(0, x)] * rix + T(16.0 / 72.0) * (r0 * rix - al) + al;
    np += pr * qx[0];
    if (NMOM > 1) { n1 += pr * qx[1]; n2 += pr * qx[2]; }
  }
  /* Branch-free gather: every source y = x - c_l is an in-range node
   * (periodic wrap), so its record and population are loaded whether or not
   * the link carries tracer, all 28 loads are issued before any is used, and
   * a select drops the links from solid sources. */
  V qv[15];
  T av[15];
  Links<1, 15>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    const I y = x + nb_off<l, -1>(oi, oj, ok);
    qv[l] = Q[y];
    av[l] = a[pidx_t<I>(l, y)];
  });
  Links<1, 15>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    const V qy = qv[l];
    const T ri = qy[W - 1], P = qy[0];
    const T pr0 = av[l] * ri + T(WL(l) / 72.0) * (r0 * ri - al);
    const T pr = (f & (1u << l)) ? pr0 : T(0);
    np += pr * P;
    if (NMOM > 1) {
      const int dd = (A.axis == 0) ? CX(l) : (A.axis == 1) ? CY(l) : CZ(l);
      const T q1 = qy[1];
      n1 += pr * (q1 + T(dd) * P);
      n2 += pr * (qy[2] + T(2 * dd) * q1 + T(dd * dd) * P);
    }
  });
  if (f >> 17) {
    const T P = qx[0];
    const T keep = T(1.0 - A.pads);
    Links<1, 15>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      av[l] = a[pidx_t<I>(OPP(l), x)];
    });
    Links<1, 15>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      if (f & (1u << (16 + l))) {
        T pr = av[l] * rix + T(WL(l) / 72.0) * (r0 * rix - al);
        if (MOVING) {
          const I s = x + nb_off<l, -1>(oi, oj, ok);
          const int si = (i - CX(l) + G.esi) % G.esi;
          const int sj = (j - CY(l) + G.esj) % G.esj;
          const int sk = (k - CZ(l) + G.esk) % G.esk;
          double Ux, Uy, Uz;
          obj_velocity(tab, objidx[s], si, sj, sk, A.rho0, Ux, Uy, Uz);
          const double cu = (double)WL(l) * (CX(l) * Ux + CY(l) * Uy + CZ(l) * Uz);
          pr += (T)(cu * (double)rix);
        }
        pr *= keep;
        np += pr * P;
        if (NMOM > 1) { n1 += pr * qx[1]; n2 += pr * qx[2]; }
      }
    });
  }
  V o;
  o[W - 1] = rix;
  o[0] = T(0);
  if (NMOM > 1) { o[1] = T(0); o[2] = T(0); }
  if (f & 1u) {
    o[0] = np;
    if (NMOM > 1) { o[1] = n1; o[2] = n2; }
  } else {
    const int m = mask[x];
    T frac = T(0);
    if (m & 0x4000) frac = T(1);
    else if (m != 0 && A.pads > 0.0) frac = T(A.pads);
    if (frac != T(0)) {
      o[0] = qx[0] + frac * np;
      if (NMOM > 1) { o[1] = qx[1] + frac * n1; o[2] = qx[2] + frac * n2; }
    }
  }
  /* a plain store: the next step gathers these records from 15 neighbours,
   * and at MP sizes they stay in the caches (non-temporal stores: 38 -> 45 us
   * per step at 128^3) */
  ((V *)qout)[x] = o;
}

/* Transition weights of the MP gather for a frozen flow.  Over the steps of
 * one run the populations, the density, alpha, the adsorption probability
 * and the object velocities do not change, so the factor that a link l
 * carries from its source y = x - c_l,
 *   pr_l(x) = a_l(y) / rho(y) + (w_l / 72) (rho0 / rho(y) - alpha)
 * (zero when y is solid; the rest link adds alpha), is the same at every
 * step, and so is the factor of everything a node keeps of its own tracer:
 * the rest link's and the sum over its partially adsorbing links.
 * k_mpweights stores the 15 factors of each node (slot 0 the own one);
 * k_mpstep_w then reads them as four vector loads instead of 15 pulled
 * population loads (and the own-node loads of partially adsorbing links),
 * and gathers only the records.  The one difference in arithmetic: the
 * kept own tracer is (sum of the own factors) x P rather than the sum of
 * the products. */
/* weights in tiles of 1024 nodes: runs of 1024 four-value vectors (weights
 * 0-3, 4-7, 8-11) and one of 1024 three-value vectors (12-14), so a wave's
 * load of one run is contiguous */
template <typename I>
__device__ __forceinline__ I wt_idx(int c, I x) {
  const I tile = (x >> POP_SHIFT) * I(15 * POP_TILE), n = x & I(POP_TILE - 1);
  return c < 3 ? tile + (I(c) * I(POP_TILE) + n) * I(4) : tile + I(12 * POP_TILE) + n * I(3);
}
template <typename T>
struct __attribute__((aligned(sizeof(T)))) Wt3 {
  T a, b, c;
};

template <typename T, bool MOVING, typename I>
__global__ void __launch_bounds__(256)
k_mpweights(const T *__restrict__ a, const T *__restrict__ qin, T *__restrict__ wt,
            const uint32_t *__restrict__ mpf, const int32_t *__restrict__ objidx,
            const double *__restrict__ tab, int W, Geom G, MPArgs A) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.plane) return;
  const int i = blockIdx.y, j = t / G.esk, k = t - j * G.esk;
  const I x = (I)i * (I)G.plane + (I)t;
  const AxisOffT<I> oi = axis_off_t<I>(i, G.esi, (I)G.plane);
  const AxisOffT<I> oj = axis_off_t<I>(j, G.esj, (I)G.esk);
  const AxisOffT<I> ok = axis_off_t<I>(k, G.esk, (I)1);
  const uint32_t f = mpf[x];
  const T r0 = T(A.rho0), al = T(A.alpha);
  T w[16];
  const T rix = qin[(long)x * W + W - 1];
  w[0] = (f & 1u) ? a[pidx_t<I>(0, x)] * rix + T(16.0 / 72.0) * (r0 * rix - al) + al : T(0);
  Links<1, 15>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    const I y = x + nb_off<l, -1>(oi, oj, ok);
    const T ri = qin[(long)y * W + W - 1];
    const T pr0 = a[pidx_t<I>(l, y)] * ri + T(WL(l) / 72.0) * (r0 * ri - al);
    w[l] = (f & (1u << l)) ? pr0 : T(0);
  });
  T R = T(0);
  if (f >> 17) {
    const T keep = T(1.0 - A.pads);
    Links<1, 15>::run([&](auto LL) {
      constexpr int l = decltype(LL)::value;
      if (f & (1u << (16 + l))) {
        T pr = a[pidx_t<I>(OPP(l), x)] * rix + T(WL(l) / 72.0) * (r0 * rix - al);
        if (MOVING) {
          const I s = x + nb_off<l, -1>(oi, oj, ok);
          const int si = (i - CX(l) + G.esi) % G.esi;
          const int sj = (j - CY(l) + G.esj) % G.esj;
          const int sk = (k - CZ(l) + G.esk) % G.esk;
          double Ux, Uy, Uz;
          obj_velocity(tab, objidx[s], si, sj, sk, A.rho0, Ux, Uy, Uz);
          const double cu = (double)WL(l) * (CX(l) * Ux + CY(l) * Uy + CZ(l) * Uz);
          pr += (T)(cu * (double)rix);
        }
        R += pr * keep;
      }
    });
  }
  w[0] += R;      /* zero where the node keeps nothing of its own */
  typedef T V4 __attribute__((ext_vector_type(4)));
#pragma unroll
  for (int c = 0; c < 3; c++)
    *(V4 *)(wt + wt_idx<I>(c, x)) = (V4){w[4 * c], w[4 * c + 1], w[4 * c + 2], w[4 * c + 3]};
  *(Wt3<T> *)(wt + wt_idx<I>(3, x)) = Wt3<T>{w[12], w[13], w[14]};
}

/* Compact node records of a run on link factors: {P} or {P, m1, m2} without
 * the 1/rho slot, which only the factors use (12 bytes per node FP32 with
 * moments instead of 16). */
template <typename T, int NM> struct CRec;
template <typename T> struct __attribute__((aligned(sizeof(T)))) CRec<T, 3> { T p, m1, m2; };
template <typename T> struct __attribute__((aligned(sizeof(T)))) CRec<T, 1> { T p; };

template <typename T, int NM>
__global__ void k_rec_pack(const T *__restrict__ q, int W, CRec<T, NM> *__restrict__ c, long N) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  CRec<T, NM> r;
  r.p = q[x * W];
  if constexpr (NM > 1) { r.m1 = q[x * W + 1]; r.m2 = q[x * W + 2]; }
  c[x] = r;
}
template <typename T, int NM>
__global__ void k_rec_unpack(const CRec<T, NM> *__restrict__ c, T *__restrict__ q, int W, long N) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  const CRec<T, NM> r = c[x];
  q[x * W] = r.p;
  if constexpr (NM > 1) { q[x * W + 1] = r.m1; q[x * W + 2] = r.m2; }
}

/* k_mpstep with the gather weights precomputed by k_mpweights: the same
 * sums over the same links in the same order, on compact records. */
template <typename T, int NMOM, typename I>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(1, sizeof(T) == 8 ? 4 : 8)))
k_mpstep_w(const T *__restrict__ wt, const CRec<T, NMOM> *__restrict__ Q,
           CRec<T, NMOM> *__restrict__ qout, const uint32_t *__restrict__ mpf,
           const int32_t *__restrict__ mask, Geom G, MPArgs A) {
  typedef T V4 __attribute__((ext_vector_type(4)));
  typedef CRec<T, NMOM> R;
  Geom Gb = G;
  Gb.xcd_map = G.mp_xcd;
  int blk_x, blk_y;
  step_block(Gb, blk_x, blk_y);
  const int t = blk_x * blockDim.x + threadIdx.x;
  if (t >= G.plane) return;
  const int i = blk_y, j = t / G.esk, k = t - j * G.esk;
  const I x = (I)i * (I)G.plane + (I)t;
  const AxisOffT<I> oi = axis_off_t<I>(i, G.esi, (I)G.plane);
  const AxisOffT<I> oj = axis_off_t<I>(j, G.esj, (I)G.esk);
  const AxisOffT<I> ok = axis_off_t<I>(k, G.esk, (I)1);
  const uint32_t f = mpf[x];
  const R qx = Q[x];
  T wv[15];
  {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const V4 v = *(const V4 *)(wt + wt_idx<I>(c, x));
      wv[4 * c] = v[0]; wv[4 * c + 1] = v[1]; wv[4 * c + 2] = v[2]; wv[4 * c + 3] = v[3];
    }
    const Wt3<T> v = *(const Wt3<T> *)(wt + wt_idx<I>(3, x));
    wv[12] = v.a; wv[13] = v.b; wv[14] = v.c;
  }
  R qv[15];
  Links<1, 15>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    qv[l] = Q[x + nb_off<l, -1>(oi, oj, ok)];
  });
  T np = T(0), n1 = T(0), n2 = T(0);
  if ((f & 1u) || (f >> 17)) {
    const T pr = wv[0];
    np += pr * qx.p;
    if constexpr (NMOM > 1) { n1 += pr * qx.m1; n2 += pr * qx.m2; }
  }
  Links<1, 15>::run([&](auto LL) {
    constexpr int l = decltype(LL)::value;
    const R qy = qv[l];
    const T P = qy.p;
    const T pr = wv[l];
    np += pr * P;
    if constexpr (NMOM > 1) {
      const int dd = (A.axis == 0) ? CX(l) : (A.axis == 1) ? CY(l) : CZ(l);
      const T q1 = qy.m1;
      n1 += pr * (q1 + T(dd) * P);
      n2 += pr * (qy.m2 + T(2 * dd) * q1 + T(dd * dd) * P);
    }
  });
  R o;
  o.p = T(0);
  if constexpr (NMOM > 1) { o.m1 = T(0); o.m2 = T(0); }
  if (f & 1u) {
    o.p = np;
    if constexpr (NMOM > 1) { o.m1 = n1; o.m2 = n2; }
  } else {
    const int m = mask[x];
    T frac = T(0);
    if (m & 0x4000) frac = T(1);
    else if (m != 0 && A.pads > 0.0) frac = T(A.pads);
    if (frac != T(0)) {
      o.p = qx.p + frac * np;
      if constexpr (NMOM > 1) { o.m1 = qx.m1 + frac * n1; o.m2 = qx.m2 + frac * n2; }
    }
  }
  qout[x] = o;
}

/* record component access: q[x * W + c] */
template <typename T>
__global__ void k_rec_put(T *__restrict__ q, int W, int c, const double *__restrict__ v, long N) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < N) q[x * W + c] = (T)v[x];
}
template <typename T>
__global__ void k_rec_get(const T *__restrict__ q, int W, int c, double *__restrict__ v, long N) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < N) v[x] = (double)q[x * W + c];
}
template <typename T>
__global__ void k_rec_zero(T *__restrict__ q, int W, int c, long N) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < N) q[x * W + c] = T(0);
}
/* 1/rho into the last slot of both record buffers */
template <typename T>
__global__ void k_rec_rinv(T *__restrict__ q0, T *__restrict__ q1, int W,
                           const T *__restrict__ rho, long N) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  const T r = T(1) / rho[x];
  q0[x * W + W - 1] = r;
  q1[x * W + W - 1] = r;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_mpsums(const T *__restrict__ q, int W, const uint16_t *__restrict__ flags,
         const int32_t *__restrict__ mask, int nmom, Geom G, double *__restrict__ part) {
  double s[6] = {0, 0, 0, 0, 0, 0};
  const long N = G.N;
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < N;
       x += (long)gridDim.x * blockDim.x) {
    const double v = (double)q[x * W];
    if (flags[x] & FLAG_SOLID) {
      s[1] += v;
    } else {
      s[0] += v;
      s[5] += 1.0;
      if (nmom > 1) { s[3] += (double)q[x * W + 1]; s[4] += (double)q[x * W + 2]; }
    }
    if (mask[x] & 0x4000) s[2] += v;
  }
  block_reduce_store<6>(s, part);
}

/* ---------------------------------------------------------------------- */
/* Elementwise helpers.                                                    */

template <typename S, typename D>
__global__ void k_convert(const S *__restrict__ s, D *__restrict__ d, long n) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) d[x] = (D)s[x];
}
template <typename T>
__global__ void k_scale(T *__restrict__ d, long n, double s) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) d[x] = (T)((double)d[x] * s);
}

/* ---------------------------------------------------------------------- */
/* Host side.                                                              */

struct slbd_lattice {
  int esi, esj, esk, prec;
  int idx32;   /* population indices fit 32 bits (and SLB_INDEX64 does not force 64) */
  long rho_version;   /* bumped whenever rho changes (MP 1/rho slots) */
  long N;
  Geom G;
  hipStream_t stream;
  void *pops[2];
  int cur;
  uint16_t *flags;
  int32_t *objidx, *mask;
  double *ux, *uy, *uz;
  double *fx, *fy, *fz;
  void *bx, *by, *bz;
  double *part, *dout;      /* reduction partials and result */
  double *psi;              /* device copy of psi when used */
  double *tab;              /* object table */
  int tab_cap;
  double *acc;              /* distribf accumulators */
  int acc_cap;
  double *uprev;            /* residual previous velocity */
  int uprev_valid;
  /* moment propagation: frozen density and link flags (per lattice) */
  void *rho;
  uint32_t *mpf;
  hipEvent_t ev0, ev1;
  int sweep_nt_mid;         /* two-step sweep: non-temporal window stores */
  int step_lds;             /* step with the k pulls staged in LDS (k_step_lds) */
  int ldsr_rpb, ldsr_bs;    /* k_step_ldsr: rows per block, threads per block (0: off) */
  void *mom[2];             /* moment storage of a run (k_step_mom), or NULL */
  void *sph_buf;            /* sphere images of a pack being placed (grow-only) */
  size_t sph_cap;
};

struct slbd_mp {
  void *q[2];      /* node records {P, (m1, m2,) 1/rho}, ping-pong pair */
  int cur, nmom, prec, W;
  long N;
  long rho_version;  /* lattice density version the 1/rho slots hold */
  void *wt;        /* gather weights of the current run (15 per node, tiled), or NULL */
  void *cq[2];     /* compact records of a run on the weights, or NULL */
};

static size_t tsize(int prec) { return prec == 32 ? sizeof(float) : sizeof(double); }

static void *dalloc(size_t bytes) {
  void *p = NULL;
  if (bytes == 0) bytes = 8;
  HIPCHECK(hipMalloc(&p, bytes));
  return p;
}

static dim3 plane_grid(const slbd_lattice *L) {
  return dim3((unsigned)((L->G.plane + 255) / 256), (unsigned)L->esi, 1);
}
static dim3 lin_grid(long n) { return dim3((unsigned)((n + 255) / 256), 1, 1); }

/* Block order of the step kernel (see step_block): plain dispatch order by
 * default; SLB_XCD_MAP selects 1 (contiguous range per XCD) or T (T x T
 * tiles per XCD, when T x T tiles divide the (row block, plane) grid into a
 * multiple of 8 tiles). */
static int step_order(const Geom &G) {
  const long nbx = (G.plane + G.step_tpb - 1) / G.step_tpb, nby = G.esi, nb = nbx * nby;
  auto fits = [&](int T) {
    return T >= 2 && nbx % T == 0 && nby % T == 0 && nb % (8L * T * T) == 0;
  };
  const char *e = getenv("SLB_XCD_MAP");
  if (e) {
    const int m = atoi(e);
    if (m == 0 || m == 1) return (m == 1 && nb % 8 == 0) ? 1 : 0;
    return fits(m) ? m : 0;
  }
  /* measured on the tiled population layout: the plain order is fastest at
   * 256^3; from planes of 2^19 nodes on, the pulls from planes i - 1 and
   * i + 1 lie 30 MB and more apart, and each XCD walking a contiguous range
   * of blocks keeps the span of its loads compact: 27.0 -> 24.0-25.2 ms per
   * FP32 step at 1024^3, 14.9 -> 12.8 ms at 512 x 512 x 2048 */
  return (G.plane >= (1 << 19) && nb % 8 == 0) ? 1 : 0;
}

extern "C" int slbd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

extern "C" const char *slbd_arch(void) {
  static char name[128];
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return "none";
  snprintf(name, sizeof name, "%s", prop.gcnArchName);
  return name;
}

/* The step kernel of a lattice, from its shape and precision (made again
 * when the precision changes).  k_step_lds needs whole rows per 512-thread
 * block. */
static void step_kernel_policy(slbd_lattice *L) {
  const int esj = L->esj, esk = L->esk, prec = L->prec;
  {
    /* measured: 362.5 -> 359.1 us per FP32 step at 256^3 (729 against 733 us
     * FP64), bit-identical */
    const int want = getenv("SLB_STEP_LDS") ? atoi(getenv("SLB_STEP_LDS")) : 1;
    const int e = esk;
    L->step_lds = want && (e == 64 || e == 128 || e == 256 || e == 512) &&
                  (L->G.plane % 512) == 0;
    /* other rows of up to 512 nodes: k_step_ldsr with the most whole rows
     * that divide the plane's rows and fit 512 threads, where at least 85%
     * of the block's threads hold a node, on lattices of 2^18 nodes and
     * more, and where either every block starts on a 128-byte line or both
     * population buffers fit the last-level cache.  Measured FP32 against
     * k_step: 80^3 13.2/15.0 us, 96^3 20.1/23.5, 100^3 26.4/27.2, 120^3
     * 42.3/43.6, 160^3 94.0/99.0, 192^3 154.6/154.3, 224^3 241.4/259.2 (FP64
     * 96^3 38.6/43.2, 160^3 185/188); but
     * rows whose blocks start off the line boundary, streamed from HBM, are
     * slower (200^3 212/181, 300^3 724/646 us), as are lattices too small
     * to fill the GPU (32^3 4.7/4.5, 48^3 7.0/6.1 us).  SLB_STEP_LDS=2 uses
     * it wherever rows fit (for tests). */
    L->ldsr_rpb = L->ldsr_bs = 0;
    if (want && !L->step_lds && e <= 512) {
      int rpb = 1;
      for (int q = 512 / e; q >= 1; q--)
        if (esj % q == 0) { rpb = q; break; }
      const int n = rpb * e, bs = (n + 63) & ~63;
      const long ts = (long)tsize(prec);
      const bool aligned = ((long)n * ts) % 128 == 0 && ((long)L->G.plane * ts) % 128 == 0;
      const bool resident = 30.0 * (double)L->N * (double)ts <= 192e6;
      if (want == 2 || (bs >= 128 && 100 * n >= 85 * bs && L->N >= (1L << 18) &&
                        (aligned || resident))) {
        L->ldsr_rpb = rpb;
        L->ldsr_bs = bs;
      }
    }
  }
}

extern "C" slbd_lattice *slbd_create(int esi, int esj, int esk, int prec) {
  if (slbd_device_count() < 1) {
    fprintf(stderr, "sunlightlb: no HIP device is visible; the GPU library cannot run\n");
    exit(EXIT_FAILURE);
  }
  if (prec != 32 && prec != 64) prec = 64;
  slbd_scratch_init();
  slbd_lattice *L = (slbd_lattice *)calloc(1, sizeof(slbd_lattice));
  L->esi = esi; L->esj = esj; L->esk = esk; L->prec = prec;
  L->N = (long)esi * esj * esk;
  L->idx32 = 15 * pop_nodes(L->N) <= 0x7fffffffL &&
             !(getenv("SLB_INDEX64") && atoi(getenv("SLB_INDEX64")) != 0);
  L->G.esi = esi; L->G.esj = esj; L->G.esk = esk; L->G.plane = esj * esk; L->G.N = L->N;
  L->G.nt_store = getenv("SLB_NT_STORE") ? atoi(getenv("SLB_NT_STORE")) : 1;
  L->G.nt_load = getenv("SLB_NT_LOAD") ? atoi(getenv("SLB_NT_LOAD")) : 0;
  /* threads per block of the step kernel: 512 measured fastest at 256^3 */
  L->G.step_tpb = 512;
  if (getenv("SLB_STEP_BLOCK")) {
    const int b = atoi(getenv("SLB_STEP_BLOCK"));
    if (b == 256 || b == 512) L->G.step_tpb = b;
  }
  L->G.xcd_map = step_order(L->G);
  L->sweep_nt_mid = getenv("SLB_SWEEP_NT") ? atoi(getenv("SLB_SWEEP_NT")) : 0;
  step_kernel_policy(L);
  {
    /* MP kernel: 256-thread blocks over (row block, plane) */
    const long nb = (long)((L->G.plane + 255) / 256) * esi;
    const int want = getenv("SLB_MP_XCD_MAP") ? atoi(getenv("SLB_MP_XCD_MAP")) : 1;
    L->G.mp_xcd = (want == 1 && nb % 8 == 0) ? 1 : 0;
  }
  HIPCHECK(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
  HIPCHECK(hipEventCreate(&L->ev0));
  HIPCHECK(hipEventCreate(&L->ev1));
  const size_t pb = 15 * (size_t)pop_nodes(L->N) * tsize(prec);
  L->pops[0] = dalloc(pb);
  L->pops[1] = dalloc(pb);
  HIPCHECK(hipMemsetAsync(L->pops[0], 0, pb, L->stream));
  HIPCHECK(hipMemsetAsync(L->pops[1], 0, pb, L->stream));
  L->flags = (uint16_t *)dalloc(sizeof(uint16_t) * L->N);
  L->objidx = (int32_t *)dalloc(sizeof(int32_t) * L->N);
  L->mask = (int32_t *)dalloc(sizeof(int32_t) * L->N);
  HIPCHECK(hipMemsetAsync(L->flags, 0, sizeof(uint16_t) * L->N, L->stream));
  HIPCHECK(hipMemsetAsync(L->objidx, 0xff, sizeof(int32_t) * L->N, L->stream));
  HIPCHECK(hipMemsetAsync(L->mask, 0, sizeof(int32_t) * L->N, L->stream));
  L->ux = (double *)dalloc(sizeof(double) * L->N);
  L->uy = (double *)dalloc(sizeof(double) * L->N);
  L->uz = (double *)dalloc(sizeof(double) * L->N);
  HIPCHECK(hipMemsetAsync(L->ux, 0, sizeof(double) * L->N, L->stream));
  HIPCHECK(hipMemsetAsync(L->uy, 0, sizeof(double) * L->N, L->stream));
  HIPCHECK(hipMemsetAsync(L->uz, 0, sizeof(double) * L->N, L->stream));
  L->part = (double *)dalloc(sizeof(double) * RED_BLOCKS * 16);
  L->dout = (double *)dalloc(sizeof(double) * 16);
  HIPCHECK(hipStreamSynchronize(L->stream));
  return L;
}

static void free_mp(slbd_lattice *L) {
  if (L->rho) { HIPCHECK(hipFree(L->rho)); L->rho = NULL; }
  if (L->mpf) { HIPCHECK(hipFree(L->mpf)); L->mpf = NULL; }
  L->rho_version++;
}

extern "C" void slbd_destroy(slbd_lattice *L) {
  if (!L) return;
  HIPCHECK(hipStreamSynchronize(L->stream));
  void *ps[] = {L->pops[0], L->pops[1], L->flags, L->objidx, L->mask, L->ux, L->uy, L->uz,
                L->fx, L->fy, L->fz, L->bx, L->by, L->bz, L->part, L->dout, L->psi,
                L->tab, L->acc, L->uprev, L->sph_buf, L->mom[0], L->mom[1]};
  for (size_t q = 0; q < sizeof ps / sizeof ps[0]; q++)
    if (ps[q]) HIPCHECK(hipFree(ps[q]));
  free_mp(L);
  HIPCHECK(hipEventDestroy(L->ev0));
  HIPCHECK(hipEventDestroy(L->ev1));
  HIPCHECK(hipStreamDestroy(L->stream));
  free(L);
}

extern "C" int slbd_precision(const slbd_lattice *L) { return L->prec; }
extern "C" void slbd_sync(slbd_lattice *L) { HIPCHECK(hipStreamSynchronize(L->stream)); }

extern "C" size_t slbd_bytes(const slbd_lattice *L) {
  size_t b = 2 * 15 * (size_t)pop_nodes(L->N) * tsize(L->prec);
  b += (size_t)L->N * (2 + 4 + 4 + 24);
  if (L->fx) b += (size_t)L->N * 24;
  if (L->bx) b += (size_t)L->N * 3 * tsize(L->prec);
  return b;
}

/* Populations: host AoS double a[15n + l] <-> device tiled layout (pidx),
 * rearranged and converted on the host, one copy each way. */
extern "C" void slbd_put_pops(slbd_lattice *L, const double *aos) {
  const long N = L->N, Np = pop_nodes(N);
  const size_t ts = tsize(L->prec);
  void *tmp = calloc((size_t)15 * Np, ts);
  if (!tmp) { fprintf(stderr, "sunlightlb: out of host memory\n"); exit(EXIT_FAILURE); }
  for (long n = 0; n < N; n++)
    for (int l = 0; l < 15; l++) {
      if (L->prec == 32) ((float *)tmp)[pidx(l, n)] = (float)aos[15 * n + l];
      else ((double *)tmp)[pidx(l, n)] = aos[15 * n + l];
    }
  HIPCHECK(hipMemcpyAsync(L->pops[L->cur], tmp, (size_t)15 * Np * ts, hipMemcpyHostToDevice,
                          L->stream));
  HIPCHECK(hipStreamSynchronize(L->stream));
  free(tmp);
}

extern "C" void slbd_get_pops(slbd_lattice *L, double *aos) {
  const long N = L->N, Np = pop_nodes(N);
  const size_t ts = tsize(L->prec);
  void *tmp = malloc((size_t)15 * Np * ts);
  if (!tmp) { fprintf(stderr, "sunlightlb: out of host memory\n"); exit(EXIT_FAILURE); }
  HIPCHECK(hipMemcpyAsync(tmp, L->pops[L->cur], (size_t)15 * Np * ts, hipMemcpyDeviceToHost,
                          L->stream));
  HIPCHECK(hipStreamSynchronize(L->stream));
  for (long n = 0; n < N; n++)
    for (int l = 0; l < 15; l++)
      aos[15 * n + l] = L->prec == 32 ? (double)((const float *)tmp)[pidx(l, n)]
                                      : ((const double *)tmp)[pidx(l, n)];
  free(tmp);
}

extern "C" void slbd_zero_pops(slbd_lattice *L) {
  HIPCHECK(hipMemsetAsync(L->pops[L->cur], 0, 15 * (size_t)pop_nodes(L->N) * tsize(L->prec), L->stream));
  HIPCHECK(hipStreamSynchronize(L->stream));
}

extern "C" int slbd_set_precision(slbd_lattice *L, int prec, double *scratch_aos) {
  if (prec != 32 && prec != 64) return -1;
  if (prec == L->prec) return 0;
  slbd_get_pops(L, scratch_aos);
  HIPCHECK(hipStreamSynchronize(L->stream));
  HIPCHECK(hipFree(L->pops[0]));
  HIPCHECK(hipFree(L->pops[1]));
  if (L->bx) { HIPCHECK(hipFree(L->bx)); HIPCHECK(hipFree(L->by)); HIPCHECK(hipFree(L->bz)); }
  L->bx = L->by = L->bz = NULL;
  free_mp(L);
  for (int q = 0; q < 2; q++)
    if (L->mom[q]) { HIPCHECK(hipFree(L->mom[q])); L->mom[q] = NULL; }
  L->prec = prec;
  step_kernel_policy(L);
  const size_t pb = 15 * (size_t)pop_nodes(L->N) * tsize(prec);
  L->pops[0] = dalloc(pb);
  L->pops[1] = dalloc(pb);
  L->cur = 0;
  slbd_put_pops(L, scratch_aos);
  return 1;
}

static void copy_in(void *dev, const void *host, size_t bytes, hipStream_t s) {
  HIPCHECK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s));
  HIPCHECK(hipStreamSynchronize(s));
}
static void copy_out(void *host, const void *dev, size_t bytes, hipStream_t s) {
  HIPCHECK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
}

extern "C" void slbd_put_objidx(slbd_lattice *L, const int32_t *objidx) {
  copy_in(L->objidx, objidx, sizeof(int32_t) * L->N, L->stream);
}
extern "C" void slbd_get_objidx(slbd_lattice *L, int32_t *objidx) {
  copy_out(objidx, L->objidx, sizeof(int32_t) * L->N, L->stream);
}
extern "C" void slbd_makemask(slbd_lattice *L) {
  hipLaunchKernelGGL(k_makemask, plane_grid(L), dim3(256), 0, L->stream, L->objidx, L->mask, L->G);
  LAUNCHCHECK();
}
extern "C" long slbd_countmask(slbd_lattice *L, int faces, const int *w) {
  CountW W;
  for (int l = 0; l < 15; l++) W.w[l] = w[l];
  /* the reduction result slot serves as the counter */
  unsigned long long *cnt = (unsigned long long *)L->dout, h = 0;
  HIPCHECK(hipMemsetAsync(cnt, 0, sizeof h, L->stream));
  /* a grid-stride pass over at most 2048 blocks: few atomics on the counter */
  const long nb = (L->N + 255) / 256;
  hipLaunchKernelGGL(k_countmask, dim3((unsigned)(nb < 2048 ? nb : 2048)), dim3(256), 0,
                     L->stream, L->mask, L->N, faces, W, cnt);
  LAUNCHCHECK();
  HIPCHECK(hipMemcpyAsync(&h, cnt, sizeof h, hipMemcpyDeviceToHost, L->stream));
  HIPCHECK(hipStreamSynchronize(L->stream));
  return (long)h;
}
extern "C" void slbd_put_mask(slbd_lattice *L, const int32_t *mask) {
  copy_in(L->mask, mask, sizeof(int32_t) * L->N, L->stream);
}
extern "C" void slbd_get_mask(slbd_lattice *L, int32_t *mask) {
  copy_out(mask, L->mask, sizeof(int32_t) * L->N, L->stream);
}
extern "C" void slbd_build_flags(slbd_lattice *L) {
  hipLaunchKernelGGL(k_flags, plane_grid(L), dim3(256), 0, L->stream, L->objidx, L->mask,
                     L->flags, L->G);
  LAUNCHCHECK();
  if (L->mpf) {
    hipLaunchKernelGGL(k_mpflags, plane_grid(L), dim3(256), 0, L->stream, L->objidx, L->mask,
                       L->mpf, L->G);
    LAUNCHCHECK();
  }
  HIPCHECK(hipStreamSynchronize(L->stream));
}

extern "C" void slbd_put_u(slbd_lattice *L, const double *ux, const double *uy, const double *uz) {
  copy_in(L->ux, ux, sizeof(double) * L->N, L->stream);
  copy_in(L->uy, uy, sizeof(double) * L->N, L->stream);
  copy_in(L->uz, uz, sizeof(double) * L->N, L->stream);
}
extern "C" void slbd_get_u(slbd_lattice *L, double *ux, double *uy, double *uz) {
  copy_out(ux, L->ux, sizeof(double) * L->N, L->stream);
  copy_out(uy, L->uy, sizeof(double) * L->N, L->stream);
  copy_out(uz, L->uz, sizeof(double) * L->N, L->stream);
}

extern "C" void slbd_put_bf(slbd_lattice *L, const double *bx, const double *by, const double *bz) {
  const size_t ts = tsize(L->prec);
  if (!L->bx) {
    L->bx = dalloc(L->N * ts); L->by = dalloc(L->N * ts); L->bz = dalloc(L->N * ts);
  }
  const double *src[3] = {bx, by, bz};
  void *dst[3] = {L->bx, L->by, L->bz};
  double *stage = (double *)dalloc(sizeof(double) * L->N);
  for (int q = 0; q < 3; q++) {
    copy_in(stage, src[q], sizeof(double) * L->N, L->stream);
    if (L->prec == 32)
      hipLaunchKernelGGL((k_convert<double, float>), lin_grid(L->N), dim3(256), 0, L->stream,
                         stage, (float *)dst[q], L->N);
    else
      hipLaunchKernelGGL((k_convert<double, double>), lin_grid(L->N), dim3(256), 0, L->stream,
                         stage, (double *)dst[q], L->N);
    LAUNCHCHECK();
  }
  HIPCHECK(hipStreamSynchronize(L->stream));
  HIPCHECK(hipFree(stage));
}

static void ensure_f(slbd_lattice *L) {
  if (L->fx) return;
  L->fx = (double *)dalloc(sizeof(double) * L->N);
  L->fy = (double *)dalloc(sizeof(double) * L->N);
  L->fz = (double *)dalloc(sizeof(double) * L->N);
  HIPCHECK(hipMemsetAsync(L->fx, 0, sizeof(double) * L->N, L->stream));
  HIPCHECK(hipMemsetAsync(L->fy, 0, sizeof(double) * L->N, L->stream));
  HIPCHECK(hipMemsetAsync(L->fz, 0, sizeof(double) * L->N, L->stream));
}
extern "C" void slbd_get_f(slbd_lattice *L, double *fx, double *fy, double *fz) {
  if (!L->fx) {
    memset(fx, 0, sizeof(double) * L->N); memset(fy, 0, sizeof(double) * L->N);
    memset(fz, 0, sizeof(double) * L->N);
    return;
  }
  copy_out(fx, L->fx, sizeof(double) * L->N, L->stream);
  copy_out(fy, L->fy, sizeof(double) * L->N, L->stream);
  copy_out(fz, L->fz, sizeof(double) * L->N, L->stream);
}
extern "C" void slbd_put_f(slbd_lattice *L, const double *fx, const double *fy, const double *fz) {
  ensure_f(L);
  copy_in(L->fx, fx, sizeof(double) * L->N, L->stream);
  copy_in(L->fy, fy, sizeof(double) * L->N, L->stream);
  copy_in(L->fz, fz, sizeof(double) * L->N, L->stream);
}
extern "C" void slbd_zero_f(slbd_lattice *L) {
  ensure_f(L);
  HIPCHECK(hipMemsetAsync(L->fx, 0, sizeof(double) * L->N, L->stream));
  HIPCHECK(hipMemsetAsync(L->fy, 0, sizeof(double) * L->N, L->stream));
  HIPCHECK(hipMemsetAsync(L->fz, 0, sizeof(double) * L->N, L->stream));
}
extern "C" void slbd_scale_f(slbd_lattice *L, double s) {
  ensure_f(L);
  double *fs[3] = {L->fx, L->fy, L->fz};
  for (int q = 0; q < 3; q++) {
    hipLaunchKernelGGL(k_scale<double>, lin_grid(L->N), dim3(256), 0, L->stream, fs[q], L->N, s);
    LAUNCHCHECK();
  }
}
extern "C" void slbd_scale_pops(slbd_lattice *L, double s) {
  const long n = 15 * pop_nodes(L->N);
  if (L->prec == 32)
    hipLaunchKernelGGL(k_scale<float>, lin_grid(n), dim3(256), 0, L->stream,
                       (float *)L->pops[L->cur], n, s);
  else
    hipLaunchKernelGGL(k_scale<double>, lin_grid(n), dim3(256), 0, L->stream,
                       (double *)L->pops[L->cur], n, s);
  LAUNCHCHECK();
}

/* ---------------------------------------------------------------------- */
/* Population vectors (whole arrays of 15 pop_nodes(N) values in the        */
/* lattice precision) for Anderson mixing of the steady-state iteration.   */

struct VecSet {
  const void *p[SLBD_VEC_MAX];
  double g[SLBD_VEC_MAX];
  int m;
};

/* All population vectors hold a multiple of 4 values (15 pop_nodes(N), with
 * pop_nodes a multiple of 1024) at 16-byte aligned addresses: the kernels
 * below move 4 values per lane per access. */
template <typename T>
using V4 = T __attribute__((ext_vector_type(4)));

template <typename T>
__global__ void __launch_bounds__(256)
k_vsub(T *__restrict__ z, const T *__restrict__ x, const T *__restrict__ y, long n4) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (long)gridDim.x * blockDim.x)
    ((V4<T> *)z)[i] = ((const V4<T> *)x)[i] - ((const V4<T> *)y)[i];
}

/* z1 = x1 - y1, z2 = x2 - y2 in one pass */
template <typename T>
__global__ void __launch_bounds__(256)
k_vsub2(T *__restrict__ z1, const T *__restrict__ x1, const T *__restrict__ y1,
        T *__restrict__ z2, const T *__restrict__ x2, const T *__restrict__ y2, long n4) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (long)gridDim.x * blockDim.x) {
    ((V4<T> *)z1)[i] = ((const V4<T> *)x1)[i] - ((const V4<T> *)y1)[i];
    ((V4<T> *)z2)[i] = ((const V4<T> *)x2)[i] - ((const V4<T> *)y2)[i];
  }
}

/* part[b * SLBD_VEC_MAX + j] = block b's share of a . B_j (FP64); with TWO,
 * part[(gridDim + b) * SLBD_VEC_MAX + j] = block b's share of a2 . B_j, each
 * B_j read once for both.  M = number of vectors B_j (compile time). */
template <typename T, int M, bool TWO>
__global__ void __launch_bounds__(256)
k_vdots(const T *__restrict__ a, const T *__restrict__ a2, VecSet B, long n4,
        double *__restrict__ part) {
  __shared__ double red[256];
  double acc[M], acc2[TWO ? M : 1];
#pragma unroll
  for (int j = 0; j < M; j++) acc[j] = 0.0;
#pragma unroll
  for (int j = 0; j < (TWO ? M : 1); j++) acc2[j] = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (long)gridDim.x * blockDim.x) {
    const V4<T> av = ((const V4<T> *)a)[i];
    V4<T> bv = av;
    if (TWO) bv = ((const V4<T> *)a2)[i];
#pragma unroll
    for (int j = 0; j < M; j++) {
      const V4<T> v = ((const V4<T> *)B.p[j])[i];
      acc[j] += (double)av.x * (double)v.x + (double)av.y * (double)v.y +
                (double)av.z * (double)v.z + (double)av.w * (double)v.w;
      if (TWO)
        acc2[TWO ? j : 0] += (double)bv.x * (double)v.x + (double)bv.y * (double)v.y +
                             (double)bv.z * (double)v.z + (double)bv.w * (double)v.w;
    }
  }
  for (int q = 0; q < (TWO ? 2 : 1); q++)
#pragma unroll
    for (int j = 0; j < M; j++) {
      red[threadIdx.x] = q ? acc2[TWO ? j : 0] : acc[j];
      __syncthreads();
      for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
      }
      if (threadIdx.x == 0) part[((long)q * gridDim.x + blockIdx.x) * SLBD_VEC_MAX + j] = red[0];
      __syncthreads();
    }
}

/* z = g - sum_j gamma_j B_j (z may be g) */
template <typename T, int M>
__global__ void __launch_bounds__(256)
k_vmix(T *z, const T *g, VecSet B, long n4) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (long)gridDim.x * blockDim.x) {
    const V4<T> gv = ((const V4<T> *)g)[i];
    double v0 = gv.x, v1 = gv.y, v2 = gv.z, v3 = gv.w;
#pragma unroll
    for (int j = 0; j < M; j++) {
      const V4<T> b = ((const V4<T> *)B.p[j])[i];
      const double c = B.g[j];
      v0 -= c * (double)b.x; v1 -= c * (double)b.y; v2 -= c * (double)b.z; v3 -= c * (double)b.w;
    }
    V4<T> o;
    o.x = (T)v0; o.y = (T)v1; o.z = (T)v2; o.w = (T)v3;
    ((V4<T> *)z)[i] = o;
  }
}

static long vec_len(const slbd_lattice *L) { return 15 * pop_nodes(L->N); }
static dim3 vec_grid(long n4) {
  const long b = (n4 + 255) / 256;
  return dim3((unsigned)(b < 4096 ? b : 4096));
}

extern "C" void *slbd_vec_cur(slbd_lattice *L) { return L->pops[L->cur]; }

extern "C" void *slbd_vec_alloc(slbd_lattice *L, int count) {
  return dalloc((size_t)count * vec_len(L) * tsize(L->prec));
}

extern "C" void *slbd_vec_at(slbd_lattice *L, void *base, int k) {
  return (char *)base + (size_t)k * vec_len(L) * tsize(L->prec);
}

extern "C" void slbd_vec_free(void *p) {
  if (p) HIPCHECK(hipFree(p));
}

extern "C" void slbd_vec_copy(slbd_lattice *L, void *dst, const void *src) {
  HIPCHECK(hipMemcpyAsync(dst, src, (size_t)vec_len(L) * tsize(L->prec),
                          hipMemcpyDeviceToDevice, L->stream));
}

extern "C" void slbd_vec_sub(slbd_lattice *L, void *z, const void *x, const void *y) {
  const long n4 = vec_len(L) / 4;
  if (L->prec == 32)
    hipLaunchKernelGGL(k_vsub<float>, vec_grid(n4), dim3(256), 0, L->stream, (float *)z,
                       (const float *)x, (const float *)y, n4);
  else
    hipLaunchKernelGGL(k_vsub<double>, vec_grid(n4), dim3(256), 0, L->stream, (double *)z,
                       (const double *)x, (const double *)y, n4);
  LAUNCHCHECK();
}

template <typename T, int M>
static void vdots_launch(slbd_lattice *L, dim3 grid, const void *a, const void *a2,
                         const VecSet &S, long n4) {
  if (a2)
    hipLaunchKernelGGL((k_vdots<T, M, true>), grid, dim3(256), 0, L->stream, (const T *)a,
                       (const T *)a2, S, n4, L->part);
  else
    hipLaunchKernelGGL((k_vdots<T, M, false>), grid, dim3(256), 0, L->stream, (const T *)a,
                       (const T *)nullptr, S, n4, L->part);
}

template <typename T>
static void vdots_t(slbd_lattice *L, dim3 grid, const void *a, const void *a2, const VecSet &S,
                    long n4) {
  switch (S.m) {
#define SLBD_VD(M) case M: vdots_launch<T, M>(L, grid, a, a2, S, n4); break;
    SLBD_VD(1) SLBD_VD(2) SLBD_VD(3) SLBD_VD(4) SLBD_VD(5) SLBD_VD(6) SLBD_VD(7) SLBD_VD(8)
    SLBD_VD(9) SLBD_VD(10) SLBD_VD(11) SLBD_VD(12) SLBD_VD(13) SLBD_VD(14) SLBD_VD(15)
    SLBD_VD(16)
#undef SLBD_VD
  }
}

extern "C" void slbd_vec_dots2(slbd_lattice *L, const void *a, const void *a2, void *const *B,
                               int m, double *out, double *out2) {
  if (m < 1 || m > SLBD_VEC_MAX) return;
  VecSet S;
  memset(&S, 0, sizeof S);
  S.m = m;
  for (int j = 0; j < m; j++) S.p[j] = B[j];
  const long n4 = vec_len(L) / 4;
  /* two partial sets must fit L->part (RED_BLOCKS x 16 doubles) */
  dim3 grid((unsigned)((n4 + 255) / 256));
  const unsigned cap = a2 ? RED_BLOCKS / 2 : RED_BLOCKS;
  if (grid.x > cap) grid.x = cap;
  if (L->prec == 32) vdots_t<float>(L, grid, a, a2, S, n4);
  else vdots_t<double>(L, grid, a, a2, S, n4);
  LAUNCHCHECK();
  static double host[RED_BLOCKS * SLBD_VEC_MAX];
  const int sets = a2 ? 2 : 1;
  HIPCHECK(hipMemcpyAsync(host, L->part, sizeof(double) * sets * grid.x * SLBD_VEC_MAX,
                          hipMemcpyDeviceToHost, L->stream));
  HIPCHECK(hipStreamSynchronize(L->stream));
  for (int q = 0; q < sets; q++)
    for (int j = 0; j < m; j++) {
      double t = 0.0;
      for (unsigned b = 0; b < grid.x; b++) t += host[((long)q * grid.x + b) * SLBD_VEC_MAX + j];
      (q ? out2 : out)[j] = t;
    }
}

extern "C" void slbd_vec_dots(slbd_lattice *L, const void *a, void *const *B, int m,
                              double *out) {
  slbd_vec_dots2(L, a, nullptr, B, m, out, nullptr);
}

extern "C" void slbd_vec_sub2(slbd_lattice *L, void *z1, const void *x1, const void *y1,
                              void *z2, const void *x2, const void *y2) {
  const long n4 = vec_len(L) / 4;
  if (L->prec == 32)
    hipLaunchKernelGGL(k_vsub2<float>, vec_grid(n4), dim3(256), 0, L->stream, (float *)z1,
                       (const float *)x1, (const float *)y1, (float *)z2, (const float *)x2,
                       (const float *)y2, n4);
  else
    hipLaunchKernelGGL(k_vsub2<double>, vec_grid(n4), dim3(256), 0, L->stream, (double *)z1,
                       (const double *)x1, (const double *)y1, (double *)z2,
                       (const double *)x2, (const double *)y2, n4);
  LAUNCHCHECK();
}

template <typename T>
static void vmix_t(slbd_lattice *L, void *z, const void *g, const VecSet &S, long n4) {
  const dim3 grid = vec_grid(n4);
  switch (S.m) {
#define SLBD_VM(M)                                                                       \
  case M:                                                                                \
    hipLaunchKernelGGL((k_vmix<T, M>), grid, dim3(256), 0, L->stream, (T *)z, (const T *)g, \
                       S, n4);                                                           \
    break;
    SLBD_VM(1) SLBD_VM(2) SLBD_VM(3) SLBD_VM(4) SLBD_VM(5) SLBD_VM(6) SLBD_VM(7) SLBD_VM(8)
    SLBD_VM(9) SLBD_VM(10) SLBD_VM(11) SLBD_VM(12) SLBD_VM(13) SLBD_VM(14) SLBD_VM(15)
    SLBD_VM(16)
#undef SLBD_VM
  }
}

extern "C" void slbd_vec_mix(slbd_lattice *L, void *z, const void *g, void *const *B,
                             const double *gamma, int m) {
  if (m < 1 || m > SLBD_VEC_MAX) return;
  VecSet S;
  memset(&S, 0, sizeof S);
  S.m = m;
  for (int j = 0; j < m; j++) { S.p[j] = B[j]; S.g[j] = gamma[j]; }
  const long n4 = vec_len(L) / 4;
  if (L->prec == 32) vmix_t<float>(L, z, g, S, n4);
  else vmix_t<double>(L, z, g, S, n4);
  LAUNCHCHECK();
}

static PhysArgs phys(const slbd_params *P) {
  PhysArgs a;
  a.omega = 1.0 / P->tau;
  a.comega = 1.0 - a.omega;
  a.rho0 = P->rho0;
  a.stokes = P->stokes_flow;
  a.skip_solid = P->skip_solid;
  a.bfx = P->bf[0]; a.bfy = P->bf[1]; a.bfz = P->bf[2];
  return a;
}

static int bf_mode(const slbd_lattice *L, const slbd_params *P) {
  if (P->bf_mode == 2 && L->bx) return 2;
  if (P->bf_mode >= 1) return 1;
  return 0;
}

/* One launch of the step kernel over planes i0 .. i0 + np - 1 (mod esi). */
template <typename T, bool MOVING, bool FORCE, int BF>
static void launch_step_t(slbd_lattice *L, const PhysArgs &A, const void *srcv, void *dstv,
                          int i0, int np, int sring, int dring, const Geom &G) {
  const T *src = (const T *)srcv;
  T *dst = (T *)dstv;
  const SweepMap M = {i0, sring, dring};
  if (L->step_lds && i0 == 0 && np == L->esi && sring == 0 && dring == 0) {
    const dim3 g2((unsigned)(G.plane / 512), (unsigned)np);
    if (L->idx32)
      hipLaunchKernelGGL((k_step_lds<T, MOVING, FORCE, BF, int>), g2, dim3(512), 0, L->stream,
                         src, dst, L->flags, L->ux, L->uy, L->uz, (const T *)L->bx,
                         (const T *)L->by, (const T *)L->bz, L->fx, L->fy, L->fz, G, A);
    else
      hipLaunchKernelGGL((k_step_lds<T, MOVING, FORCE, BF, long>), g2, dim3(512), 0, L->stream,
                         src, dst, L->flags, L->ux, L->uy, L->uz, (const T *)L->bx,
                         (const T *)L->by, (const T *)L->bz, L->fx, L->fy, L->fz, G, A);
    LAUNCHCHECK();
    return;
  }
  if (L->ldsr_rpb && i0 == 0 && np == L->esi && sring == 0 && dring == 0) {
    /* the tile orders (xcd_map >= 2) were chosen for k_step's grid */
    Geom Gr = G;
    if (Gr.xcd_map != 1) Gr.xcd_map = 0;
    const dim3 g2((unsigned)(G.esj / L->ldsr_rpb), (unsigned)np);
    const size_t sh = (size_t)10 * (size_t)L->ldsr_bs * sizeof(T);
    if (L->idx32)
      hipLaunchKernelGGL((k_step_ldsr<T, MOVING, FORCE, BF, int>), g2, dim3(L->ldsr_bs), sh,
                         L->stream, src, dst, L->flags, L->ux, L->uy, L->uz, (const T *)L->bx,
                         (const T *)L->by, (const T *)L->bz, L->fx, L->fy, L->fz, Gr, A,
                         L->ldsr_rpb);
    else
      hipLaunchKernelGGL((k_step_ldsr<T, MOVING, FORCE, BF, long>), g2, dim3(L->ldsr_bs), sh,
                         L->stream, src, dst, L->flags, L->ux, L->uy, L->uz, (const T *)L->bx,
                         (const T *)L->by, (const T *)L->bz, L->fx, L->fy, L->fz, Gr, A,
                         L->ldsr_rpb);
    LAUNCHCHECK();
    return;
  }
  const dim3 grid((unsigned)((G.plane + G.step_tpb - 1) / G.step_tpb), (unsigned)np);
  if (L->idx32)
    hipLaunchKernelGGL((k_step<T, MOVING, FORCE, BF, int>), grid, dim3(G.step_tpb), 0,
                       L->stream, src, dst, L->flags, L->ux, L->uy, L->uz, (const T *)L->bx,
                       (const T *)L->by, (const T *)L->bz, L->fx, L->fy, L->fz, G, A, M);
  else
    hipLaunchKernelGGL((k_step<T, MOVING, FORCE, BF, long>), grid, dim3(G.step_tpb), 0,
                       L->stream, src, dst, L->flags, L->ux, L->uy, L->uz, (const T *)L->bx,
                       (const T *)L->by, (const T *)L->bz, L->fx, L->fy, L->fz, G, A, M);
  LAUNCHCHECK();
}

template <typename T>
static void launch_step(slbd_lattice *L, const PhysArgs &A, bool moving, bool force, int bf,
                        const void *src, void *dst, int i0, int np, int sring, int dring,
                        const Geom &G) {
#define SLB_STEP_CASE(M, F, B)                                                           \
  if (moving == M && force == F && bf == B) {                                            \
    launch_step_t<T, M, F, B>(L, A, src, dst, i0, np, sring, dring, G);                 \
    return;                                                                              \
  }
  SLB_STEP_CASE(false, false, 0) SLB_STEP_CASE(false, false, 1) SLB_STEP_CASE(false, false, 2)
  SLB_STEP_CASE(false, true, 0) SLB_STEP_CASE(false, true, 1) SLB_STEP_CASE(false, true, 2)
  SLB_STEP_CASE(true, false, 0) SLB_STEP_CASE(true, false, 1) SLB_STEP_CASE(true, false, 2)
  SLB_STEP_CASE(true, true, 0) SLB_STEP_CASE(true, true, 1) SLB_STEP_CASE(true, true, 2)
#undef SLB_STEP_CASE
}

/* Default bytes of one chunk's planes in the two-step sweep; 0: the sweep
 * is off unless SLB_SWEEP names a chunk size. */
#ifndef SLB_SWEEP_BYTES
#define SLB_SWEEP_BYTES 0L
#endif

/* Planes per chunk of the two-step sweep (0: off), from SLB_SWEEP (planes)
 * or the default byte budget of a chunk's intermediate planes. */
static int sweep_planes(const slbd_lattice *L) {
  const long pbytes = 15L * L->G.plane * (long)tsize(L->prec);
  int P;
  if (getenv("SLB_SWEEP")) {
    P = atoi(getenv("SLB_SWEEP"));
  } else {
    const long budget = SLB_SWEEP_BYTES;
    P = (int)(budget / pbytes);
  }
  if (P < 4) return 0;
  /* the sliding window of P + 4 planes must fit the second buffer, and the
   * ring must start each plane on a whole population tile */
  if (L->esi < 2 * P + 2 || L->G.plane % POP_TILE != 0) return 0;
  /* at most 256 launches per sweep */
  if ((L->esi + P - 1) / P * 2 + 1 > 256) return 0;
  return P;
}

/* Two steps as one sweep over the planes (temporal blocking through the
 * device's last-level cache).  Step t+1 of plane i needs step t of planes
 * i-1..i+1, which needs step t-1 of planes i-2..i+2.  The planes are taken
 * in chunks of P: launch one computes step t of the chunk's planes from the
 * live buffer A into a sliding window of P + 2 planes (+ planes 0 and 1) in
 * the second buffer, launch two computes step t+1 of the planes one behind
 * from that window back into A, in place: the step t-1 values it overwrites
 * have no reader left, since the planes that read them (i-1..i+1 of step t)
 * are done.  Planes esi-1 and 0 of step t+1 close the sweep, when step t of
 * planes esi-2..1 exists.  The window of the latest chunk (a few tens of MB)
 * is written and read back while it is still in the 256 MB last-level
 * cache, so HBM sees the live buffer read once and written once per two
 * steps.  Every node computes exactly what the one-step kernel computes:
 * the results are bit-identical. */
extern "C" int slbd_sweep_plan(int esi, int P, int *rows, int maxrows) {
  if (P < 4 || esi < 2 * P + 2) return 0;
  const int R = P + 2;
  int n = 0;
  auto put = [&](int phase, int i0, int np, int sring, int dring) {
    if (rows && n < maxrows) {
      int *r = rows + 5 * n;
      r[0] = phase; r[1] = i0; r[2] = np; r[3] = sring; r[4] = dring;
    }
    n++;
  };
  for (int c0 = 0; c0 < esi; c0 += P) {
    const int np1 = c0 + P <= esi ? P : esi - c0;
    put(1, c0, np1, 0, R);
    const int lo = c0 == 0 ? 1 : c0 - 1;
    const int hi = c0 + np1 - 1 < esi - 1 ? c0 + np1 - 1 : esi - 1;
    if (hi > lo) put(2, lo, hi - lo, R, 0);
  }
  put(2, esi - 1, 2, R, 0);
  return n;
}

template <typename T>
static void sweep2(slbd_lattice *L, const PhysArgs &A, bool moving, int bf, int P) {
  void *Abuf = L->pops[L->cur], *Bbuf = L->pops[L->cur ^ 1];
  Geom g1 = L->G, g2 = L->G;
  g1.xcd_map = 0; g2.xcd_map = 0;
  g1.nt_store = L->sweep_nt_mid;   /* window stores: cacheable by default */
  g2.nt_load = 0;
  int rows[5 * 256];
  const int n = slbd_sweep_plan(L->esi, P, rows, 256);
  if (n < 1 || n > 256) {
    fprintf(stderr, "sunlightlb: two-step sweep plan of %d launches\n", n);
    exit(EXIT_FAILURE);
  }
  for (int q = 0; q < n; q++) {
    const int *r = rows + 5 * q;
    if (r[0] == 1)
      launch_step<T>(L, A, moving, false, bf, Abuf, Bbuf, r[1], r[2], r[3], r[4], g1);
    else
      launch_step<T>(L, A, moving, false, bf, Bbuf, Abuf, r[1], r[2], r[3], r[4], g2);
  }
}

/* Runs of steps on moment storage (k_step_mom): SLB_STEP_MOM=1, read per
 * call; runs of 2 steps or more without force accumulation, on lattices the
 * staged step covers, with 32-bit indices and skip_solid off. */
static bool mom_run_ok(const slbd_lattice *L, const slbd_params *P, int nsteps, int force) {
  const char *e = getenv("SLB_STEP_MOM");
  if (!(e && atoi(e) == 1)) return false;
  return !force && nsteps >= 2 && L->step_lds && L->idx32 && !P->skip_solid;
}

template <typename T, bool IN_POPS, bool OUT_POPS>
static void launch_mom(slbd_lattice *L, const PhysArgs &A, bool moving, int bf,
                       const T *srcP, const T *srcM, T *dstP, T *dstM) {
  const dim3 g2((unsigned)(L->G.plane / 512), (unsigned)L->esi);
  const T *bx = (const T *)L->bx, *by = (const T *)L->by, *bz = (const T *)L->bz;
#define SLB_MOM_CASE(M, B)                                                                    \
  if (moving == M && bf == B) {                                                               \
    hipLaunchKernelGGL((k_step_mom<T, M, B, int, IN_POPS, OUT_POPS>), g2, dim3(512), 0,      \
                       L->stream, srcP, srcM, dstP, dstM, L->flags, L->ux, L->uy, L->uz, bx, \
                       by, bz, L->G, A);                                                      \
    LAUNCHCHECK();                                                                            \
    return;                                                                                   \
  }
  SLB_MOM_CASE(false, 0) SLB_MOM_CASE(false, 1) SLB_MOM_CASE(false, 2)
  SLB_MOM_CASE(true, 0) SLB_MOM_CASE(true, 1) SLB_MOM_CASE(true, 2)
#undef SLB_MOM_CASE
}

template <typename T>
static void mom_run(slbd_lattice *L, const PhysArgs &A, bool moving, int bf, int nsteps) {
  const size_t mb = 10 * (size_t)pop_nodes(L->N) * sizeof(T);
  for (int q = 0; q < 2; q++)
    if (!L->mom[q]) L->mom[q] = dalloc(mb);
  T *M[2] = {(T *)L->mom[0], (T *)L->mom[1]};
  const T *P0 = (const T *)L->pops[L->cur];
  T *P1 = (T *)L->pops[L->cur ^ 1];
  launch_mom<T, true, false>(L, A, moving, bf, P0, nullptr, nullptr, M[0]);
  int a = 0;
  for (int s = 1; s < nsteps - 1; s++) {
    launch_mom<T, false, false>(L, A, moving, bf, nullptr, M[a], nullptr, M[a ^ 1]);
    a ^= 1;
  }
  launch_mom<T, false, true>(L, A, moving, bf, nullptr, M[a], P1, nullptr);
  L->cur ^= 1;
}

extern "C" void slbd_step(slbd_lattice *L, const slbd_params *P, int nsteps, int force) {
  const PhysArgs A = phys(P);
  const int bf = bf_mode(L, P);
  if (force) ensure_f(L);
  const int sw = force ? 0 : sweep_planes(L);
  if (sw == 0 && mom_run_ok(L, P, nsteps, force)) {
    if (L->prec == 32) mom_run<float>(L, A, P->moving != 0, bf, nsteps);
    else mom_run<double>(L, A, P->moving != 0, bf, nsteps);
    return;
  }
  int s = 0;
  if (sw > 0) {
    for (; s + 2 <= nsteps; s += 2) {
      if (L->prec == 32) sweep2<float>(L, A, P->moving != 0, bf, sw);
      else sweep2<double>(L, A, P->moving != 0, bf, sw);
    }
  }
  for (; s < nsteps; s++) {
    void *src = L->pops[L->cur], *dst = L->pops[L->cur ^ 1];
    if (L->prec == 32)
      launch_step<float>(L, A, P->moving != 0, force != 0, bf, src, dst, 0, L->esi, 0, 0, L->G);
    else
      launch_step<double>(L, A, P->moving != 0, force != 0, bf, src, dst, 0, L->esi, 0, 0, L->G);
    L->cur ^= 1;
  }
}

/* Name of the kernel that the lattice's plain steps launch. */
extern "C" const char *slbd_step_kernel(const slbd_lattice *L) {
  if (L->step_lds) return "k_step_lds";
  if (L->ldsr_rpb) return "k_step_ldsr";
  return "k_step";
}

extern "C" double slbd_time_steps(slbd_lattice *L, const slbd_params *P, int nsteps) {
  HIPCHECK(hipStreamSynchronize(L->stream));
  HIPCHECK(hipEventRecord(L->ev0, L->stream));
  slbd_step(L, P, nsteps, 0);
  HIPCHECK(hipEventRecord(L->ev1, L->stream));
  HIPCHECK(hipEventSynchronize(L->ev1));
  float ms = 0.0f;
  HIPCHECK(hipEventElapsedTime(&ms, L->ev0, L->ev1));
  return (double)ms;
}

template <typename T>
static void launch_bbl(slbd_lattice *L, bool moving, bool force) {
  T *a = (T *)L->pops[L->cur];
  if (moving && force)
    hipLaunchKernelGGL((k_bbl<T, true, true>), plane_grid(L), dim3(256), 0, L->stream, a, L->mask,
                       L->ux, L->uy, L->uz, L->fx, L->fy, L->fz, L->G);
  else if (moving)
    hipLaunchKernelGGL((k_bbl<T, true, false>), plane_grid(L), dim3(256), 0, L->stream, a, L->mask,
                       L->ux, L->uy, L->uz, L->fx, L->fy, L->fz, L->G);
  else if (force)
    hipLaunchKernelGGL((k_bbl<T, false, true>), plane_grid(L), dim3(256), 0, L->stream, a, L->mask,
                       L->ux, L->uy, L->uz, L->fx, L->fy, L->fz, L->G);
  else
    hipLaunchKernelGGL((k_bbl<T, false, false>), plane_grid(L), dim3(256), 0, L->stream, a, L->mask,
                       L->ux, L->uy, L->uz, L->fx, L->fy, L->fz, L->G);
  LAUNCHCHECK();
}

extern "C" void slbd_bbl(slbd_lattice *L, const slbd_params *P, int force) {
  if (force) ensure_f(L);
  if (L->prec == 32) launch_bbl<float>(L, P->moving != 0, force != 0);
  else launch_bbl<double>(L, P->moving != 0, force != 0);
}

extern "C" void slbd_propagate(slbd_lattice *L) {
  if (L->prec == 32)
    hipLaunchKernelGGL(k_propagate<float>, plane_grid(L), dim3(256), 0, L->stream,
                       (const float *)L->pops[L->cur], (float *)L->pops[L->cur ^ 1], L->G);
  else
    hipLaunchKernelGGL(k_propagate<double>, plane_grid(L), dim3(256), 0, L->stream,
                       (const double *)L->pops[L->cur], (double *)L->pops[L->cur ^ 1], L->G);
  LAUNCHCHECK();
  L->cur ^= 1;
}

template <typename T>
static void launch_collide(slbd_lattice *L, const PhysArgs &A, int bf) {
  T *a = (T *)L->pops[L->cur];
  const T *bx = (const T *)L->bx, *by = (const T *)L->by, *bz = (const T *)L->bz;
  if (bf == 2)
    hipLaunchKernelGGL((k_collide<T, 2>), lin_grid(L->N), dim3(256), 0, L->stream, a, L->flags, bx, by, bz, L->G, A);
  else if (bf == 1)
    hipLaunchKernelGGL((k_collide<T, 1>), lin_grid(L->N), dim3(256), 0, L->stream, a, L->flags, bx, by, bz, L->G, A);
  else
    hipLaunchKernelGGL((k_collide<T, 0>), lin_grid(L->N), dim3(256), 0, L->stream, a, L->flags, bx, by, bz, L->G, A);
  LAUNCHCHECK();
}

extern "C" void slbd_collide(slbd_lattice *L, const slbd_params *P) {
  const PhysArgs A = phys(P);
  const int bf = bf_mode(L, P);
  if (L->prec == 32) launch_collide<float>(L, A, bf);
  else launch_collide<double>(L, A, bf);
}

static int red_blocks(long N) {
  long b = (N + 255) / 256;
  return (int)(b < RED_BLOCKS ? b : RED_BLOCKS);
}

template <typename T>
static void launch_vmean(slbd_lattice *L, const PhysArgs &A, int bf, double fac, const double *psi) {
  const int nb = red_blocks(L->N);
  const T *a = (const T *)L->pops[L->cur];
  const T *bx = (const T *)L->bx, *by = (const T *)L->by, *bz = (const T *)L->bz;
  if (bf == 2)
    hipLaunchKernelGGL((k_vmean<T, 2>), dim3(nb), dim3(256), 0, L->stream, a, L->flags, bx, by, bz, psi, fac, L->G, A, L->part);
  else if (bf == 1)
    hipLaunchKernelGGL((k_vmean<T, 1>), dim3(nb), dim3(256), 0, L->stream, a, L->flags, bx, by, bz, psi, fac, L->G, A, L->part);
  else
    hipLaunchKernelGGL((k_vmean<T, 0>), dim3(nb), dim3(256), 0, L->stream, a, L->flags, bx, by, bz, psi, fac, L->G, A, L->part);
  LAUNCHCHECK();
  hipLaunchKernelGGL(k_final<5>, dim3(1), dim3(256), 0, L->stream, L->part, nb, L->dout);
  LAUNCHCHECK();
}

extern "C" void slbd_vmean(slbd_lattice *L, const slbd_params *P, double fac,
                           const double *psi_host, double out[5]) {
  const PhysArgs A = phys(P);
  const int bf = bf_mode(L, P);
  const double *psi = NULL;
  if (psi_host) {
    if (!L->psi) L->psi = (double *)dalloc(sizeof(double) * L->N);
    copy_in(L->psi, psi_host, sizeof(double) * L->N, L->stream);
    psi = L->psi;
  }
  if (L->prec == 32) launch_vmean<float>(L, A, bf, fac, psi);
  else launch_vmean<double>(L, A, bf, fac, psi);
  copy_out(out, L->dout, sizeof(double) * 5, L->stream);
}

template <typename T>
static void launch_vgrab(slbd_lattice *L, const PhysArgs &A, int bf, double fac, double prefac, int zero) {
  const T *a = (const T *)L->pops[L->cur];
  const T *bx = (const T *)L->bx, *by = (const T *)L->by, *bz = (const T *)L->bz;
  if (bf == 2)
    hipLaunchKernelGGL((k_vgrab<T, 2>), lin_grid(L->N), dim3(256), 0, L->stream, a, L->flags, bx, by, bz, fac, prefac, L->G, A, L->ux, L->uy, L->uz, zero);
  else if (bf == 1)
    hipLaunchKernelGGL((k_vgrab<T, 1>), lin_grid(L->N), dim3(256), 0, L->stream, a, L->flags, bx, by, bz, fac, prefac, L->G, A, L->ux, L->uy, L->uz, zero);
  else
    hipLaunchKernelGGL((k_vgrab<T, 0>), lin_grid(L->N), dim3(256), 0, L->stream, a, L->flags, bx, by, bz, fac, prefac, L->G, A, L->ux, L->uy, L->uz, zero);
  LAUNCHCHECK();
}

extern "C" void slbd_vgrab(slbd_lattice *L, const slbd_params *P, double fac, double prefac) {
  const PhysArgs A = phys(P);
  const int bf = bf_mode(L, P);
  if (L->prec == 32) launch_vgrab<float>(L, A, bf, fac, prefac, 0);
  else launch_vgrab<double>(L, A, bf, fac, prefac, 0);
}

extern "C" void slbd_vzero(slbd_lattice *L) {
  slbd_params P;
  memset(&P, 0, sizeof P);
  P.tau = 1.0;
  const PhysArgs A = phys(&P);
  if (L->prec == 32) launch_vgrab<float>(L, A, 0, 0.0, 0.0, 1);
  else launch_vgrab<double>(L, A, 0, 0.0, 0.0, 1);
}

static void put_tab(slbd_lattice *L, int nobj, const double *objtab) {
  if (nobj > L->tab_cap) {
    if (L->tab) HIPCHECK(hipFree(L->tab));
    L->tab_cap = nobj + 16;
    L->tab = (double *)dalloc(sizeof(double) * SLBD_OBJ_STRIDE * L->tab_cap);
  }
  if (nobj > 0) copy_in(L->tab, objtab, sizeof(double) * SLBD_OBJ_STRIDE * nobj, L->stream);
}

extern "C" void slbd_vspread(slbd_lattice *L, int nobj, const double *objtab, double rho0) {
  put_tab(L, nobj, objtab);
  hipLaunchKernelGGL(k_vspread, plane_grid(L), dim3(256), 0, L->stream, L->objidx, L->tab, rho0,
                     L->ux, L->uy, L->uz, L->G);
  LAUNCHCHECK();
}

extern "C" void slbd_distribf(slbd_lattice *L, int nobj, const double *objtab, double *frc,
                              double *trq) {
  if (nobj <= 0) return;
  ensure_f(L);
  put_tab(L, nobj, objtab);
  if (nobj > L->acc_cap) {
    if (L->acc) HIPCHECK(hipFree(L->acc));
    L->acc_cap = nobj + 16;
    L->acc = (double *)dalloc(sizeof(double) * 6 * L->acc_cap);
  }
  HIPCHECK(hipMemsetAsync(L->acc, 0, sizeof(double) * 6 * nobj, L->stream));
  hipLaunchKernelGGL(k_distribf, dim3(red_blocks(L->N)), dim3(256), 0, L->stream, L->objidx,
                     L->fx, L->fy, L->fz, L->tab, nobj, L->G, L->acc);
  LAUNCHCHECK();
  double *h = (double *)malloc(sizeof(double) * 6 * nobj);
  copy_out(h, L->acc, sizeof(double) * 6 * nobj, L->stream);
  for (int o = 0; o < nobj; o++) {
    frc[3 * o] = h[6 * o]; frc[3 * o + 1] = h[6 * o + 1]; frc[3 * o + 2] = h[6 * o + 2];
    trq[3 * o] = h[6 * o + 3]; trq[3 * o + 1] = h[6 * o + 4]; trq[3 * o + 2] = h[6 * o + 5];
  }
  free(h);
}

template <typename T>
static void launch_residual(slbd_lattice *L, const PhysArgs &A, int bf, int nb) {
  const T *a = (const T *)L->pops[L->cur];
  const T *bx = (const T *)L->bx, *by = (const T *)L->by, *bz = (const T *)L->bz;
  if (bf == 2)
    hipLaunchKernelGGL((k_residual<T, 2>), dim3(nb), dim3(256), 0, L->stream, a, L->flags, bx, by, bz, L->G, A, L->uprev, L->part);
  else if (bf == 1)
    hipLaunchKernelGGL((k_residual<T, 1>), dim3(nb), dim3(256), 0, L->stream, a, L->flags, bx, by, bz, L->G, A, L->uprev, L->part);
  else
    hipLaunchKernelGGL((k_residual<T, 0>), dim3(nb), dim3(256), 0, L->stream, a, L->flags, bx, by, bz, L->G, A, L->uprev, L->part);
  LAUNCHCHECK();
}

extern "C" void slbd_residual_reset(slbd_lattice *L) { L->uprev_valid = 0; }

extern "C" double slbd_residual(slbd_lattice *L, const slbd_params *P) {
  if (!L->uprev) {
    L->uprev = (double *)dalloc(sizeof(double) * 3 * L->N);
    L->uprev_valid = 0;
  }
  if (!L->uprev_valid) HIPCHECK(hipMemsetAsync(L->uprev, 0, sizeof(double) * 3 * L->N, L->stream));
  const PhysArgs A = phys(P);
  const int bf = bf_mode(L, P);
  const int nb = red_blocks(L->N);
  if (L->prec == 32) launch_residual<float>(L, A, bf, nb);
  else launch_residual<double>(L, A, bf, nb);
  hipLaunchKernelGGL(k_final<2>, dim3(1), dim3(256), 0, L->stream, L->part, nb, L->dout);
  LAUNCHCHECK();
  double o[2];
  copy_out(o, L->dout, sizeof o, L->stream);
  const int had = L->uprev_valid;
  L->uprev_valid = 1;
  if (!had) return -1.0;
  return o[1] > 0.0 ? sqrt(o[0] / o[1]) : 0.0;
}

extern "C" void slbd_sum_pops(slbd_lattice *L, double out[15]) {
  const int nb = red_blocks(L->N);
  if (L->prec == 32)
    hipLaunchKernelGGL(k_sum_pops<float>, dim3(nb), dim3(256), 0, L->stream,
                       (const float *)L->pops[L->cur], L->flags, L->G, L->part);
  else
    hipLaunchKernelGGL(k_sum_pops<double>, dim3(nb), dim3(256), 0, L->stream,
                       (const double *)L->pops[L->cur], L->flags, L->G, L->part);
  LAUNCHCHECK();
  hipLaunchKernelGGL(k_final<15>, dim3(1), dim3(256), 0, L->stream, L->part, nb, L->dout);
  LAUNCHCHECK();
  copy_out(out, L->dout, sizeof(double) * 15, L->stream);
}

/* ---- moment propagation ---- */

static void ensure_mp_lattice(slbd_lattice *L) {
  const size_t ts = tsize(L->prec);
  if (!L->rho) {
    L->rho = dalloc(L->N * ts);
    HIPCHECK(hipMemsetAsync(L->rho, 0, L->N * ts, L->stream));
  }
  if (!L->mpf) {
    L->mpf = (uint32_t *)dalloc(sizeof(uint32_t) * L->N);
    hipLaunchKernelGGL(k_mpflags, plane_grid(L), dim3(256), 0, L->stream, L->objidx, L->mask,
                       L->mpf, L->G);
    LAUNCHCHECK();
  }
}

extern "C" slbd_mp *slbd_mp_create(slbd_lattice *L, int nmom) {
  ensure_mp_lattice(L);
  slbd_mp *M = (slbd_mp *)calloc(1, sizeof(slbd_mp));
  M->nmom = nmom > 1 ? 3 : 1;
  M->W = M->nmom > 1 ? 4 : 2;
  M->prec = L->prec;
  M->N = L->N;
  M->rho_version = -1;
  const size_t bytes = (size_t)L->N * M->W * tsize(L->prec);
  for (int b = 0; b < 2; b++) {
    M->q[b] = dalloc(bytes);
    HIPCHECK(hipMemsetAsync(M->q[b], 0, bytes, L->stream));
  }
  HIPCHECK(hipStreamSynchronize(L->stream));
  return M;
}

extern "C" void slbd_mp_destroy(slbd_mp *M) {
  if (!M) return;
  HIPCHECK(hipDeviceSynchronize());
  for (int b = 0; b < 2; b++)
    if (M->q[b]) HIPCHECK(hipFree(M->q[b]));
  if (M->wt) HIPCHECK(hipFree(M->wt));
  for (int b = 0; b < 2; b++)
    if (M->cq[b]) HIPCHECK(hipFree(M->cq[b]));
  free(M);
}

static void check_mp(const slbd_lattice *L, const slbd_mp *M) {
  if (M->prec != L->prec || M->N != L->N) {
    fprintf(stderr, "sunlightlb: MP fields do not match the lattice (precision %d vs %d)\n",
            M->prec, L->prec);
    exit(EXIT_FAILURE);
  }
}

/* refresh the 1/rho slots when the lattice density changed */
static void mp_sync_rho(slbd_lattice *L, slbd_mp *M) {
  if (M->rho_version == L->rho_version) return;
  if (L->prec == 32)
    hipLaunchKernelGGL(k_rec_rinv<float>, lin_grid(L->N), dim3(256), 0, L->stream,
                       (float *)M->q[0], (float *)M->q[1], M->W, (const float *)L->rho, L->N);
  else
    hipLaunchKernelGGL(k_rec_rinv<double>, lin_grid(L->N), dim3(256), 0, L->stream,
                       (double *)M->q[0], (double *)M->q[1], M->W, (const double *)L->rho, L->N);
  LAUNCHCHECK();
  M->rho_version = L->rho_version;
}

extern "C" void slbd_mp_rhoinit(slbd_lattice *L, double rho0) {
  ensure_mp_lattice(L);
  if (L->prec == 32)
    hipLaunchKernelGGL(k_mprho<float>, lin_grid(L->N), dim3(256), 0, L->stream,
                       (const float *)L->pops[L->cur], L->flags, rho0, (float *)L->rho, L->G);
  else
    hipLaunchKernelGGL(k_mprho<double>, lin_grid(L->N), dim3(256), 0, L->stream,
                       (const double *)L->pops[L->cur], L->flags, rho0, (double *)L->rho, L->G);
  LAUNCHCHECK();
  L->rho_version++;
}

static void put_field_T(slbd_lattice *L, void *dev, const double *host) {
  if (L->prec == 64) { copy_in(dev, host, sizeof(double) * L->N, L->stream); return; }
  float *t = (float *)malloc(sizeof(float) * L->N);
  for (long n = 0; n < L->N; n++) t[n] = (float)host[n];
  copy_in(dev, t, sizeof(float) * L->N, L->stream);
  free(t);
}
static void get_field_T(slbd_lattice *L, double *host, const void *dev) {
  if (L->prec == 64) { copy_out(host, dev, sizeof(double) * L->N, L->stream); return; }
  float *t = (float *)malloc(sizeof(float) * L->N);
  copy_out(t, dev, sizeof(float) * L->N, L->stream);
  for (long n = 0; n < L->N; n++) host[n] = (double)t[n];
  free(t);
}

extern "C" void slbd_mp_put_rho(slbd_lattice *L, const double *rho) {
  ensure_mp_lattice(L);
  put_field_T(L, L->rho, rho);
  L->rho_version++;
}
extern "C" void slbd_mp_get_rho(slbd_lattice *L, double *rho) {
  ensure_mp_lattice(L);
  get_field_T(L, rho, L->rho);
}
extern "C" void slbd_mp_put(slbd_lattice *L, slbd_mp *M, int field, const double *v) {
  check_mp(L, M);
  if (field >= M->nmom) return;
  double *stage = (double *)dalloc(sizeof(double) * L->N);
  copy_in(stage, v, sizeof(double) * L->N, L->stream);
  if (L->prec == 32)
    hipLaunchKernelGGL(k_rec_put<float>, lin_grid(L->N), dim3(256), 0, L->stream,
                       (float *)M->q[M->cur], M->W, field, stage, L->N);
  else
    hipLaunchKernelGGL(k_rec_put<double>, lin_grid(L->N), dim3(256), 0, L->stream,
                       (double *)M->q[M->cur], M->W, field, stage, L->N);
  LAUNCHCHECK();
  HIPCHECK(hipStreamSynchronize(L->stream));
  HIPCHECK(hipFree(stage));
}
extern "C" void slbd_mp_get(slbd_lattice *L, slbd_mp *M, int field, double *v) {
  check_mp(L, M);
  if (field >= M->nmom) { memset(v, 0, sizeof(double) * L->N); return; }
  double *stage = (double *)dalloc(sizeof(double) * L->N);
  if (L->prec == 32)
    hipLaunchKernelGGL(k_rec_get<float>, lin_grid(L->N), dim3(256), 0, L->stream,
                       (const float *)M->q[M->cur], M->W, field, stage, L->N);
  else
    hipLaunchKernelGGL(k_rec_get<double>, lin_grid(L->N), dim3(256), 0, L->stream,
                       (const double *)M->q[M->cur], M->W, field, stage, L->N);
  LAUNCHCHECK();
  copy_out(v, stage, sizeof(double) * L->N, L->stream);
  HIPCHECK(hipFree(stage));
}
extern "C" void slbd_mp_zero(slbd_lattice *L, slbd_mp *M, int field) {
  check_mp(L, M);
  if (field >= M->nmom) return;
  if (L->prec == 32)
    hipLaunchKernelGGL(k_rec_zero<float>, lin_grid(L->N), dim3(256), 0, L->stream,
                       (float *)M->q[M->cur], M->W, field, L->N);
  else
    hipLaunchKernelGGL(k_rec_zero<double>, lin_grid(L->N), dim3(256), 0, L->stream,
                       (double *)M->q[M->cur], M->W, field, L->N);
  LAUNCHCHECK();
}

/* MP steps of one run with precomputed gather weights: runs of at least
 * this many steps (SLB_MP_WEIGHTS=0 turns the weights off) */
static int mp_weights_min(void) {
  /* the factor pass and the record packing cost about four steps; read per
   * call so that a caller can change it (0: never) */
  const char *e = getenv("SLB_MP_WEIGHTS");
  return e ? atoi(e) : 32;
}

template <typename T, int NMOM, bool MOVING>
static void launch_mp_w(slbd_lattice *L, slbd_mp *M, const MPArgs &A, int nsteps) {
  if (!M->wt) M->wt = dalloc((size_t)pop_nodes(L->N) * 15 * sizeof(T));
  if (L->idx32)
    hipLaunchKernelGGL((k_mpweights<T, MOVING, int>), plane_grid(L), dim3(256), 0, L->stream,
                       (const T *)L->pops[L->cur], (const T *)M->q[M->cur], (T *)M->wt, L->mpf,
                       L->objidx, L->tab, M->W, L->G, A);
  else
    hipLaunchKernelGGL((k_mpweights<T, MOVING, long>), plane_grid(L), dim3(256), 0, L->stream,
                       (const T *)L->pops[L->cur], (const T *)M->q[M->cur], (T *)M->wt, L->mpf,
                       L->objidx, L->tab, M->W, L->G, A);
  LAUNCHCHECK();
  /* the run on compact records: packed from the current records, unpacked
   * back into them at the end (their 1/rho slots stay as they are) */
  typedef CRec<T, NMOM> R;
  for (int b = 0; b < 2; b++)
    if (!M->cq[b]) M->cq[b] = dalloc((size_t)L->N * sizeof(R));
  hipLaunchKernelGGL((k_rec_pack<T, NMOM>), lin_grid(L->N), dim3(256), 0, L->stream,
                     (const T *)M->q[M->cur], M->W, (R *)M->cq[0], L->N);
  LAUNCHCHECK();
  int c = 0;
  for (int s = 0; s < nsteps; s++) {
    if (L->idx32)
      hipLaunchKernelGGL((k_mpstep_w<T, NMOM, int>), plane_grid(L), dim3(256), 0,
                         L->stream, (const T *)M->wt, (const R *)M->cq[c], (R *)M->cq[c ^ 1],
                         L->mpf, L->mask, L->G, A);
    else
      hipLaunchKernelGGL((k_mpstep_w<T, NMOM, long>), plane_grid(L), dim3(256), 0,
                         L->stream, (const T *)M->wt, (const R *)M->cq[c], (R *)M->cq[c ^ 1],
                         L->mpf, L->mask, L->G, A);
    LAUNCHCHECK();
    c ^= 1;
  }
  hipLaunchKernelGGL((k_rec_unpack<T, NMOM>), lin_grid(L->N), dim3(256), 0, L->stream,
                     (const R *)M->cq[c], (T *)M->q[M->cur], M->W, L->N);
  LAUNCHCHECK();
}

template <typename T, int NMOM, bool MOVING>
static void launch_mp_t(slbd_lattice *L, slbd_mp *M, const MPArgs &A) {
  const int c = M->cur;
  /* 32-bit indices pay off (fewer VGPRs) only without the wall term */
  if (!MOVING && L->idx32)
    hipLaunchKernelGGL((k_mpstep<T, NMOM, MOVING, int>), plane_grid(L), dim3(256), 0, L->stream,
                       (const T *)L->pops[L->cur], (const T *)M->q[c], (T *)M->q[c ^ 1], L->mpf,
                       L->mask, L->objidx, L->tab, L->G, A);
  else
    hipLaunchKernelGGL((k_mpstep<T, NMOM, MOVING, long>), plane_grid(L), dim3(256), 0,
                       L->stream, (const T *)L->pops[L->cur], (const T *)M->q[c],
                       (T *)M->q[c ^ 1], L->mpf, L->mask, L->objidx, L->tab, L->G, A);
  LAUNCHCHECK();
  M->cur ^= 1;
}

template <typename T>
static void launch_mp(slbd_lattice *L, slbd_mp *M, const MPArgs &A, bool moving, int nsteps) {
  const int wmin = mp_weights_min();
  if (wmin > 0 && nsteps >= wmin) {
    if (M->nmom > 1) {
      if (moving) launch_mp_w<T, 3, true>(L, M, A, nsteps);
      else launch_mp_w<T, 3, false>(L, M, A, nsteps);
    } else {
      if (moving) launch_mp_w<T, 1, true>(L, M, A, nsteps);
      else launch_mp_w<T, 1, false>(L, M, A, nsteps);
    }
    return;
  }
  for (int s = 0; s < nsteps; s++) {
    if (M->nmom > 1) {
      if (moving) launch_mp_t<T, 3, true>(L, M, A); else launch_mp_t<T, 3, false>(L, M, A);
    } else {
      if (moving) launch_mp_t<T, 1, true>(L, M, A); else launch_mp_t<T, 1, false>(L, M, A);
    }
  }
}

extern "C" void slbd_mp_step(slbd_lattice *L, slbd_mp *M, double rho0, double alpha,
                             int moving, int nobj, const double *objtab, double pads,
                             int nmom, int axis, int nsteps) {
  (void)nmom;
  check_mp(L, M);
  ensure_mp_lattice(L);
  mp_sync_rho(L, M);
  if (moving) put_tab(L, nobj, objtab);
  MPArgs A;
  A.rho0 = rho0; A.alpha = alpha; A.pads = pads; A.axis = axis;
  if (L->prec == 32) launch_mp<float>(L, M, A, moving != 0, nsteps);
  else launch_mp<double>(L, M, A, moving != 0, nsteps);
}

extern "C" double slbd_time_mp(slbd_lattice *L, slbd_mp *M, double rho0, double alpha,
                               double pads, int nmom, int nsteps) {
  check_mp(L, M);
  ensure_mp_lattice(L);
  mp_sync_rho(L, M);
  HIPCHECK(hipStreamSynchronize(L->stream));
  HIPCHECK(hipEventRecord(L->ev0, L->stream));
  slbd_mp_step(L, M, rho0, alpha, 0, 0, NULL, pads, nmom, 0, nsteps);
  HIPCHECK(hipEventRecord(L->ev1, L->stream));
  HIPCHECK(hipEventSynchronize(L->ev1));
  float ms = 0.0f;
  HIPCHECK(hipEventElapsedTime(&ms, L->ev0, L->ev1));
  return (double)ms;
}

extern "C" void slbd_mp_sums(slbd_lattice *L, slbd_mp *M, double out[6]) {
  check_mp(L, M);
  const int nb = red_blocks(L->N);
  const int c = M->cur;
  if (L->prec == 32)
    hipLaunchKernelGGL(k_mpsums<float>, dim3(nb), dim3(256), 0, L->stream,
                       (const float *)M->q[c], M->W, L->flags, L->mask, M->nmom, L->G, L->part);
  else
    hipLaunchKernelGGL(k_mpsums<double>, dim3(nb), dim3(256), 0, L->stream,
                       (const double *)M->q[c], M->W, L->flags, L->mask, M->nmom, L->G, L->part);
  LAUNCHCHECK();
  hipLaunchKernelGGL(k_final<6>, dim3(1), dim3(256), 0, L->stream, L->part, nb, L->dout);
  LAUNCHCHECK();
  copy_out(out, L->dout, sizeof(double) * 6, L->stream);
}

/* ---- velocity autocorrelation (src/mpvacf.c mpvacf) ---- */

/* pass 1: sum over fluid nodes of psi P and psi^2 */
template <typename T>
__global__ void __launch_bounds__(256)
k_vacf1(const T *__restrict__ q, int W, const uint16_t *__restrict__ flags,
        const double *__restrict__ psi, Geom G, double *__restrict__ part) {
  double s[2] = {0.0, 0.0};
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < G.N;
       x += (long)gridDim.x * blockDim.x) {
    if (flags[x] & FLAG_SOLID) continue;
    const double w = psi ? psi[x] : 1.0;
    s[0] += w * (double)q[x * W];
    s[1] += w * w;
  }
  block_reduce_store<2>(s, part);
}

/* pass 2: P -= psi pmean on fluid nodes, then the sums of psi P u_b and
 * psi P2 (u = j/rho with the frozen MP density) */
template <typename T>
__global__ void __launch_bounds__(256)
k_vacf2(const T *__restrict__ a, T *__restrict__ q, int W, const T *__restrict__ q2, int W2,
        const T *__restrict__ rho, const uint16_t *__restrict__ flags,
        const double *__restrict__ psi, double pmean, int comp, Geom G,
        double *__restrict__ part) {
  double s[2] = {0.0, 0.0};
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < G.N;
       x += (long)gridDim.x * blockDim.x) {
    if (flags[x] & FLAG_SOLID) continue;
    const double w = psi ? psi[x] : 1.0;
    const T pv = (T)((double)q[x * W] - w * pmean);
    q[x * W] = pv;
    double j = 0.0;
#pragma unroll
    for (int l = 1; l < 15; l++) {
      const int c = comp == 0 ? CX(l) : (comp == 1 ? CY(l) : CZ(l));
      if (c) j += c * (double)a[pidx(l, x)];
    }
    const double u = j / (double)rho[x];
    s[0] += w * (double)pv * u;
    s[1] += w * (double)q2[x * W2];
  }
  block_reduce_store<2>(s, part);
}

template <typename T>
static void launch_vacf(slbd_lattice *L, slbd_mp *M, slbd_mp *M2, const double *psi, int comp,
                        double out[2]) {
  const int nb = red_blocks(L->N);
  hipLaunchKernelGGL(k_vacf1<T>, dim3(nb), dim3(256), 0, L->stream, (const T *)M->q[M->cur], M->W,
                     L->flags, psi, L->G, L->part);
  LAUNCHCHECK();
  hipLaunchKernelGGL(k_final<2>, dim3(1), dim3(256), 0, L->stream, L->part, nb, L->dout);
  LAUNCHCHECK();
  double h[2];
  copy_out(h, L->dout, sizeof(double) * 2, L->stream);
  const double pmean = h[0] / h[1];
  hipLaunchKernelGGL(k_vacf2<T>, dim3(nb), dim3(256), 0, L->stream,
                     (const T *)L->pops[L->cur], (T *)M->q[M->cur], M->W,
                     (const T *)M2->q[M2->cur], M2->W, (const T *)L->rho, L->flags, psi, pmean,
                     comp, L->G, L->part);
  LAUNCHCHECK();
  hipLaunchKernelGGL(k_final<2>, dim3(1), dim3(256), 0, L->stream, L->part, nb, L->dout);
  LAUNCHCHECK();
  copy_out(out, L->dout, sizeof(double) * 2, L->stream);
}

extern "C" void slbd_vacf(slbd_lattice *L, slbd_mp *M, slbd_mp *M2, const double *psi_host,
                          int comp, double out[2]) {
  check_mp(L, M);
  check_mp(L, M2);
  ensure_mp_lattice(L);
  const double *psi = NULL;
  if (psi_host) {
    if (!L->psi) L->psi = (double *)dalloc(sizeof(double) * L->N);
    copy_in(L->psi, psi_host, sizeof(double) * L->N, L->stream);
    psi = L->psi;
  }
  if (L->prec == 32) launch_vacf<float>(L, M, M2, psi, comp, out);
  else launch_vacf<double>(L, M, M2, psi, comp, out);
}

/* ---- device copy bandwidth (roofline and counter calibration) ---- */

__global__ void __launch_bounds__(256)
k_copy16(const uint4 *__restrict__ s, uint4 *__restrict__ d, long n) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) d[x] = s[x];
}

/* Copy `bytes` (rounded down to 16) between two fresh device buffers reps
 * times; returns the mean milliseconds per copy. */
extern "C" double slbd_copy_test(size_t bytes, int reps) {
  const long n = (long)(bytes / 16);
  if (n <= 0 || reps <= 0) return 0.0;
  uint4 *a = (uint4 *)dalloc((size_t)n * 16), *b = (uint4 *)dalloc((size_t)n * 16);
  HIPCHECK(hipMemset(a, 0, (size_t)n * 16));
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  const unsigned g = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_copy16, dim3(g), dim3(256), 0, 0, a, b, n);
  HIPCHECK(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy16, dim3(g), dim3(256), 0, 0, a, b, n);
  HIPCHECK(hipEventRecord(e1, 0));
  HIPCHECK(hipEventSynchronize(e1));
  LAUNCHCHECK();
  float ms = 0.0f;
  HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
  HIPCHECK(hipEventDestroy(e0));
  HIPCHECK(hipEventDestroy(e1));
  HIPCHECK(hipFree(a));
  HIPCHECK(hipFree(b));
  return (double)ms / reps;
}

/* ---------------------------------------------------------------------- */
/* Warm start: network features and targets from a lattice, and equilibrium */
/* populations from a predicted velocity field.                             */

__device__ __forceinline__ void node_ijk(const Geom &G, long n, int &i, int &j, int &k) {
  i = (int)(n / G.plane);
  const int r = (int)(n - (long)i * G.plane);
  j = r / G.esk;
  k = r - j * G.esk;
}

/* periodic minimum-image squared distance between two nodes */
__device__ __forceinline__ int pdist2(const Geom &G, long a, long b) {
  int ai, aj, ak, bi, bj, bk;
  node_ijk(G, a, ai, aj, ak);
  node_ijk(G, b, bi, bj, bk);
  int di = abs(ai - bi), dj = abs(aj - bj), dk = abs(ak - bk);
  di = min(di, G.esi - di); dj = min(dj, G.esj - dj); dk = min(dk, G.esk - dk);
  return di * di + dj * dj + dk * dk;
}

__global__ void k_jfa_init(const uint16_t *__restrict__ flags, int seed_solid,
                           int *__restrict__ nearest, Geom G) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.N) return;
  const int solid = (flags[n] & FLAG_SOLID) ? 1 : 0;
  nearest[n] = (solid == seed_solid) ? (int)n : -1;
}

/* one jump-flooding pass with stride s (periodic): keep the closest seed
 * among the 27 nodes at offsets s*(dz,dy,dx) */
__global__ void k_jfa_pass(const int *__restrict__ in, int *__restrict__ out, int s, Geom G) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.N) return;
  int i, j, k;
  node_ijk(G, n, i, j, k);
  int best = in[n];
  int bd = best >= 0 ? pdist2(G, n, best) : 0x7fffffff;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        int ii = ((i + dz * s) % G.esi + G.esi) % G.esi;
        int jj = ((j + dy * s) % G.esj + G.esj) % G.esj;
        int kk = ((k + dx * s) % G.esk + G.esk) % G.esk;
        const int c = in[((long)ii * G.esj + jj) * G.esk + kk];
        if (c < 0) continue;
        const int d = pdist2(G, n, c);
        if (d < bd) { bd = d; best = c; }
      }
  out[n] = best;
}

/* sum over fluid nodes of the distance to the nearest solid node (FP64,
 * atomics), and the fluid node count */
__global__ void __launch_bounds__(256)
k_fluid_dist(const uint16_t *__restrict__ flags, const int *__restrict__ ns, double *__restrict__ out,
             Geom G) {
  __shared__ double sm[2][4];
  double d = 0.0, c = 0.0;
  for (long n = (long)blockIdx.x * blockDim.x + threadIdx.x; n < G.N;
       n += (long)gridDim.x * blockDim.x) {
    if (flags[n] & FLAG_SOLID) continue;
    d += (ns[n] >= 0) ? (double)sqrtf((float)pdist2(G, n, ns[n])) : 64.0;
    c += 1.0;
  }
  for (int o = 32; o > 0; o >>= 1) {
    d += __shfl_xor(d, o, 64);
    c += __shfl_xor(c, o, 64);
  }
  if ((threadIdx.x & 63) == 0) { sm[0][threadIdx.x >> 6] = d; sm[1][threadIdx.x >> 6] = c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(out, sm[0][0] + sm[0][1] + sm[0][2] + sm[0][3]);
    atomicAdd(out + 1, sm[1][0] + sm[1][1] + sm[1][2] + sm[1][3]);
  }
}

/* channels: 0 solid (1/0), 1 signed distance * scale (+ to the nearest
 * solid node on fluid nodes, - to the nearest fluid node on solid nodes) */
__global__ void k_features(const uint16_t *__restrict__ flags, const int *__restrict__ ns,
                           const int *__restrict__ nf, uint16_t *__restrict__ X, int C, float scale,
                           Geom G) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.N) return;
  const bool solid = (flags[n] & FLAG_SOLID) != 0;
  float d;
  if (solid) d = (nf[n] >= 0) ? -sqrtf((float)pdist2(G, n, nf[n])) : -64.f;
  else d = (ns[n] >= 0) ? sqrtf((float)pdist2(G, n, ns[n])) : 64.f;
  for (int c = 0; c < C; c++) {
    float v = 0.f;
    if (c == 0) v = solid ? 1.f : 0.f;
    else if (c == 1) v = d * scale;
    const uint32_t u = __float_as_uint(v) + 0x7fffu + ((__float_as_uint(v) >> 16) & 1u);
    X[n * C + c] = (uint16_t)(u >> 16);
  }
}

/* targets: the 10 hydrodynamic moments of the live (post-collide) state,
 * scaled per group -- channels 0..2 corrected velocity (j - b/2)/rho *
 * sc[0], channel 3 drho * sc[1], channels 4..9 the stress s = Pi - drho/3 I
 * (xx, yy, zz, xy, xz, yz) * sc[2] -- on fluid nodes, 0 on solid nodes;
 * channels from 10 to C-1 are zero (C < 10 keeps the first C). */
template <typename T, typename O>
__global__ void k_targets(const T *__restrict__ a, const uint16_t *__restrict__ flags, double rho0,
                          double bx, double by, double bz, float su, float sr, float ss,
                          O *__restrict__ Y, int C, Geom G) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.N) return;
  float m[10];
#pragma unroll
  for (int q = 0; q < 10; q++) m[q] = 0.f;
  if (!(flags[n] & FLAG_SOLID)) {
    double d = 0, j[3] = {0, 0, 0}, P[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int l = 0; l < 15; l++) {
      const double v = (double)a[pidx(l, n)];
      d += v;
      j[0] += CX(l) * v; j[1] += CY(l) * v; j[2] += CZ(l) * v;
      P[0] += CX(l) * CX(l) * v; P[1] += CY(l) * CY(l) * v; P[2] += CZ(l) * CZ(l) * v;
      P[3] += CX(l) * CY(l) * v; P[4] += CX(l) * CZ(l) * v; P[5] += CY(l) * CZ(l) * v;
    }
    const double rho = rho0 + d;
    m[0] = (float)((j[0] - 0.5 * bx) / rho * su);
    m[1] = (float)((j[1] - 0.5 * by) / rho * su);
    m[2] = (float)((j[2] - 0.5 * bz) / rho * su);
    m[3] = (float)(d * sr);
    m[4] = (float)((P[0] - d / 3.0) * ss);
    m[5] = (float)((P[1] - d / 3.0) * ss);
    m[6] = (float)((P[2] - d / 3.0) * ss);
    m[7] = (float)(P[3] * ss);
    m[8] = (float)(P[4] * ss);
    m[9] = (float)(P[5] * ss);
  }
  for (int c = 0; c < C; c++) {
    const float v = c < 10 ? m[c] : 0.f;
    if constexpr (sizeof(O) == 4) {
      Y[n * C + c] = v;
    } else {
      const uint32_t w = __float_as_uint(v) + 0x7fffu + ((__float_as_uint(v) >> 16) & 1u);
      Y[n * C + c] = (uint16_t)(w >> 16);
    }
  }
}

/* Populations from predicted moments U[n][0..9] (the layout of k_targets)
 * divided by the group scales (iu, ir, is = 1/sc):
 *   mode 0 -- the lbloadvel rebuild (src/lbstore.c:346-399): drho = 0,
 *             j = rho0 u - b/2 on fluid nodes (so that the collide that
 *             follows, adding b, leaves the corrected velocity u), stress 0
 *             (Stokes) or u j; rest on solid nodes;
 *   mode 1 -- the post-collide state itself: drho, j = (rho0 + drho) u + b/2
 *             and the predicted stress on fluid nodes, rest on solid nodes.
 * Both use the node algebra of collide: a_l = (w_l/72)(drho + 3 c.j +
 * 9/2 (cc:s - tr s / 3)). */
template <typename T>
__global__ void k_inject(T *__restrict__ a, const uint16_t *__restrict__ flags,
                         const float *__restrict__ U, int CU, double iu, double ir, double is,
                         double rho0, double bx, double by, double bz, int stokes, int mode,
                         Geom G) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.N) return;
  double d = 0.0, j[3] = {0.0, 0.0, 0.0};
  double s[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  if (!(flags[n] & FLAG_SOLID)) {
    const float *u = U + n * CU;
    if (mode == 0) {
      j[0] = rho0 * (u[0] * iu) - 0.5 * bx;
      j[1] = rho0 * (u[1] * iu) - 0.5 * by;
      j[2] = rho0 * (u[2] * iu) - 0.5 * bz;
      if (!stokes)
        for (int p = 0; p < 3; p++)
          for (int q = 0; q < 3; q++) s[p][q] = j[p] * j[q] / rho0;
    } else {
      d = u[3] * ir;
      const double rho = rho0 + d;
      j[0] = rho * (u[0] * iu) + 0.5 * bx;
      j[1] = rho * (u[1] * iu) + 0.5 * by;
      j[2] = rho * (u[2] * iu) + 0.5 * bz;
      s[0][0] = u[4] * is; s[1][1] = u[5] * is; s[2][2] = u[6] * is;
      s[0][1] = s[1][0] = u[7] * is;
      s[0][2] = s[2][0] = u[8] * is;
      s[1][2] = s[2][1] = u[9] * is;
    }
  }
  const double tr = s[0][0] + s[1][1] + s[2][2];
#pragma unroll
  for (int l = 0; l < 15; l++) {
    const int c[3] = {CX(l), CY(l), CZ(l)};
    double cs = 0.0;
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++) cs += c[p] * c[q] * s[p][q];
    const double cj = c[0] * j[0] + c[1] * j[1] + c[2] * j[2];
    a[pidx(l, n)] = (T)((WL(l) / 72.0) * (d + 3.0 * cj + 4.5 * (cs - tr / 3.0)));
  }
}

extern "C" long slbd_nodes(const slbd_lattice *L) { return L->N; }
extern "C" void slbd_dims(const slbd_lattice *L, int *d) { d[0] = L->esi; d[1] = L->esj; d[2] = L->esk; }

__global__ void k_zero_f64(double *__restrict__ p, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = 0.0;
}

/* Features of the lattice's solid map into a C-channel BF16 device tensor,
 * with the jump-flooding seed buffers buf[0..2] (N ints each) and the two
 * accumulators acc[0..1] supplied by the caller.  Returns ell. */
static double features_run(slbd_lattice *L, uint16_t *X, int C, int *const buf[3],
                           double *acc) {
  int maxd = L->esi > L->esj ? L->esi : L->esj;
  maxd = maxd > L->esk ? maxd : L->esk;
  int s0 = 1;
  while (2 * s0 < maxd) s0 *= 2;
  /* phase 0 seeds on solid nodes (distance to solid), phase 1 on fluid */
  int *ping = buf[0], *pong = buf[1], *res[2];
  for (int phase = 0; phase < 2; phase++) {
    if (phase == 1) pong = buf[2];
    hipLaunchKernelGGL(k_jfa_init, lin_grid(L->N), dim3(256), 0, L->stream, L->flags,
                       phase == 0 ? 1 : 0, ping, L->G);
    LAUNCHCHECK();
    for (int s = s0; s >= 1; s /= 2) {
      hipLaunchKernelGGL(k_jfa_pass, lin_grid(L->N), dim3(256), 0, L->stream, ping, pong, s, L->G);
      LAUNCHCHECK();
      int *t = ping; ping = pong; pong = t;
    }
    /* one more stride-1 pass (JFA+1) corrects most of the residual errors */
    hipLaunchKernelGGL(k_jfa_pass, lin_grid(L->N), dim3(256), 0, L->stream, ping, pong, 1, L->G);
    LAUNCHCHECK();
    res[phase] = pong;   /* ping is free for the next phase */
  }
  /* normalising length: mean distance from a fluid node to the solid */
  double h[2];
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, L->stream, acc, 2);
  LAUNCHCHECK();
  hipLaunchKernelGGL(k_fluid_dist, dim3((unsigned)red_blocks(L->N)), dim3(256), 0, L->stream,
                     L->flags, res[0], acc, L->G);
  LAUNCHCHECK();
  HIPCHECK(hipMemcpyAsync(h, acc, 2 * sizeof(double), hipMemcpyDeviceToHost, L->stream));
  HIPCHECK(hipStreamSynchronize(L->stream));
  const double ell = (h[1] > 0.0 && h[0] > 0.0) ? h[0] / h[1] : 1.0;
  hipLaunchKernelGGL(k_features, lin_grid(L->N), dim3(256), 0, L->stream, L->flags, res[0], res[1],
                     X, C, (float)(1.0 / ell), L->G);
  LAUNCHCHECK();
  HIPCHECK(hipStreamSynchronize(L->stream));
  return ell;
}

extern "C" double slbd_features(slbd_lattice *L, uint16_t *X, int C) {
  int *buf[3];
  double *acc;
  for (int b = 0; b < 3; b++) HIPCHECK(hipMalloc(&buf[b], sizeof(int) * L->N));
  HIPCHECK(hipMalloc(&acc, 2 * sizeof(double)));
  const double ell = features_run(L, X, C, buf, acc);
  for (int b = 0; b < 3; b++) HIPCHECK(hipFree(buf[b]));
  HIPCHECK(hipFree(acc));
  return ell;
}

/* Allocation-free features: the seed buffers and accumulators live in one
 * file-scope device array, `scratch`, held by every GPU the library runs on
 * until the process ends: 256 bytes of accumulators, then three seed
 * buffers of SLBD_FEATURES_MAX_NODES ints (24 MiB + 256 B).  In host code
 * `scratch` names a host-side stand-in, not device memory: its device
 * address is resolved once per device with hipGetSymbolAddress (at lattice
 * creation, or the first call on a device) and only that address is passed
 * to kernels.  Every call on one GPU shares the array; calls that overlap
 * on one GPU (two host threads, or two streams in flight) race. */
#define SCRATCH_ACC_BYTES 256
static __device__ __align__(256) unsigned char
    scratch[SCRATCH_ACC_BYTES + 3 * sizeof(int) * (size_t)SLBD_FEATURES_MAX_NODES];
#define SCRATCH_MAX_DEVICES 64
static unsigned char *scratch_dev[SCRATCH_MAX_DEVICES];

/* Device address of `scratch` on the current device (resolved on first use
 * there); NULL when it cannot be resolved. */
static unsigned char *scratch_ptr(void) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SCRATCH_MAX_DEVICES) return NULL;
  if (!scratch_dev[dev]) {
    void *p = NULL;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      fprintf(stderr, "sunlightlb: hipGetSymbolAddress(scratch) failed on device %d: %s\n",
              dev, hipGetErrorString(e));
      return NULL;
    }
    scratch_dev[dev] = (unsigned char *)p;
  }
  return scratch_dev[dev];
}

extern "C" void slbd_scratch_init(void) { (void)scratch_ptr(); }

extern "C" double slbd_features_noalloc(slbd_lattice *L, uint16_t *X, int C) {
  if (L->N > (long)SLBD_FEATURES_MAX_NODES) {
    fprintf(stderr, "slb_warm_features_noalloc: a lattice of %ld nodes exceeds the static "
            "scratch of %ld nodes\n", L->N, (long)SLBD_FEATURES_MAX_NODES);
    return -1.0;
  }
  unsigned char *base = scratch_ptr();
  if (!base) return -1.0;
  int *buf[3];
  for (int b = 0; b < 3; b++)
    buf[b] = (int *)(base + SCRATCH_ACC_BYTES + (size_t)b * sizeof(int) * SLBD_FEATURES_MAX_NODES);
  return features_run(L, X, C, buf, (double *)base);
}

template <typename O>
static void launch_targets(slbd_lattice *L, const slbd_params *P, const float *sc, O *Y, int C) {
  const double bx = P->bf[0], by = P->bf[1], bz = P->bf[2];
  if (L->prec == 32)
    hipLaunchKernelGGL((k_targets<float, O>), lin_grid(L->N), dim3(256), 0, L->stream,
                       (const float *)L->pops[L->cur], L->flags, P->rho0, bx, by, bz, sc[0], sc[1],
                       sc[2], Y, C, L->G);
  else
    hipLaunchKernelGGL((k_targets<double, O>), lin_grid(L->N), dim3(256), 0, L->stream,
                       (const double *)L->pops[L->cur], L->flags, P->rho0, bx, by, bz, sc[0], sc[1],
                       sc[2], Y, C, L->G);
  LAUNCHCHECK();
  HIPCHECK(hipStreamSynchronize(L->stream));
}

extern "C" void slbd_targets(slbd_lattice *L, const slbd_params *P, const float *sc, uint16_t *Y,
                             int C) {
  launch_targets<uint16_t>(L, P, sc, Y, C);
}

extern "C" void slbd_moments(slbd_lattice *L, const slbd_params *P, const float *sc, float *Y,
                             int C) {
  launch_targets<float>(L, P, sc, Y, C);
}

extern "C" void slbd_inject(slbd_lattice *L, const slbd_params *P, const float *U, int CU,
                            const double *inv_sc, int mode) {
  const double rho0 = P->rho0, bx = P->bf[0], by = P->bf[1], bz = P->bf[2];
  const int st = P->stokes_flow;
  if (L->prec == 32)
    hipLaunchKernelGGL(k_inject<float>, lin_grid(L->N), dim3(256), 0, L->stream,
                       (float *)L->pops[L->cur], L->flags, U, CU, inv_sc[0], inv_sc[1], inv_sc[2],
                       rho0, bx, by, bz, st, mode, L->G);
  else
    hipLaunchKernelGGL(k_inject<double>, lin_grid(L->N), dim3(256), 0, L->stream,
                       (double *)L->pops[L->cur], L->flags, U, CU, inv_sc[0], inv_sc[1], inv_sc[2],
                       rho0, bx, by, bz, st, mode, L->G);
  LAUNCHCHECK();
  HIPCHECK(hipStreamSynchronize(L->stream));
}

/* ---------------------------------------------------------------------- */
/* Warm start: projection of a velocity field onto the discretely         */
/* mass-conserving fields of the lattice's fluid nodes.                   */
/*                                                                        */
/* Face fluxes F(x, x+e_a) = open(x, x+e_a) (u_a(x) + u_a(x+e_a)) / 2,     */
/* open = both nodes fluid (no flux through walls), periodic wrap.  The   */
/* potential phi solves A phi = -div F with the fluid-graph Laplacian     */
/* A phi(x) = sum over open faces (phi(x) - phi(neighbour)) (conjugate    */
/* gradients, FP64, scalars kept on the device), so that the fluxes       */
/* F - (phi(x+e_a) - phi(x)) on open faces carry no divergence.  Each      */
/* fluid node's velocity is then corrected by the mean of its open face   */
/* differences along each axis.                                           */

struct ProjNb {
  long xm[3], xp[3];   /* neighbour indices along x, y, z */
};

__device__ __forceinline__ ProjNb proj_nb(long n, const Geom &G) {
  const int i = (int)(n / G.plane), t = (int)(n - (long)i * G.plane);
  const int j = t / G.esk, k = t - j * G.esk;
  const AxisOff oi = axis_off(i, G.esi, G.plane);
  const AxisOff oj = axis_off(j, G.esj, G.esk);
  const AxisOff ok = axis_off(k, G.esk, 1);
  ProjNb b;
  b.xm[0] = n + oi.m; b.xp[0] = n + oi.p;
  b.xm[1] = n + oj.m; b.xp[1] = n + oj.p;
  b.xm[2] = n + ok.m; b.xp[2] = n + ok.p;
  return b;
}

__device__ __forceinline__ bool proj_fluid(const uint16_t *__restrict__ flags, long n) {
  return !(flags[n] & FLAG_SOLID);
}

/* rhs = -div F on fluid nodes (0 on solid); x = 0; r = p = rhs; partial r.r */
__global__ void __launch_bounds__(256)
k_proj_init(const uint16_t *__restrict__ flags, const float *__restrict__ U, int CU,
            double *__restrict__ x, double *__restrict__ r, double *__restrict__ p, Geom G,
            double *__restrict__ part) {
  double s[1] = {0.0};
  for (long n = (long)blockIdx.x * blockDim.x + threadIdx.x; n < G.N;
       n += (long)gridDim.x * blockDim.x) {
    double rhs = 0.0;
    if (proj_fluid(flags, n)) {
      const ProjNb b = proj_nb(n, G);
      double div = 0.0;
      for (int a = 0; a < 3; a++) {
        const double u0 = U[n * CU + a];
        if (proj_fluid(flags, b.xp[a])) div += 0.5 * (u0 + (double)U[b.xp[a] * CU + a]);
        if (proj_fluid(flags, b.xm[a])) div -= 0.5 * ((double)U[b.xm[a] * CU + a] + u0);
      }
      rhs = -div;
    }
    x[n] = 0.0;
    r[n] = rhs;
    p[n] = rhs;
    s[0] += rhs * rhs;
  }
  block_reduce_store<1>(s, part);
}

/* q = A p; partial p.q */
__global__ void __launch_bounds__(256)
k_proj_ap(const uint16_t *__restrict__ flags, const double *__restrict__ p,
          double *__restrict__ q, Geom G, double *__restrict__ part) {
  double s[1] = {0.0};
  for (long n = (long)blockIdx.x * blockDim.x + threadIdx.x; n < G.N;
       n += (long)gridDim.x * blockDim.x) {
    double v = 0.0;
    if (proj_fluid(flags, n)) {
      const ProjNb b = proj_nb(n, G);
      const double p0 = p[n];
      for (int a = 0; a < 3; a++) {
        if (proj_fluid(flags, b.xp[a])) v += p0 - p[b.xp[a]];
        if (proj_fluid(flags, b.xm[a])) v += p0 - p[b.xm[a]];
      }
      s[0] += p0 * v;
    }
    q[n] = v;
  }
  block_reduce_store<1>(s, part);
}

/* Converged: r.r at most PROJ_TOL^2 of its start.  Iterating on past that
 * point lets p.Ap, a ratio of roundoff-level sums, push the iterate away
 * again, so a converged iteration changes nothing. */
#define PROJ_TOL 1e-10
__device__ __forceinline__ bool proj_done(const double *__restrict__ scal, int rr_slot) {
  return !(scal[rr_slot] > PROJ_TOL * PROJ_TOL * scal[3]);
}

/* alpha = rr / pq; x += alpha p; r -= alpha q; partial r.r (new) */
__global__ void __launch_bounds__(256)
k_proj_xr(const double *__restrict__ p, const double *__restrict__ q, double *__restrict__ x,
          double *__restrict__ r, const double *__restrict__ scal, int rr_slot, long N,
          double *__restrict__ part) {
  const double rr = scal[rr_slot], pq = scal[2];
  const double alpha = (pq > 0.0 && !proj_done(scal, rr_slot)) ? rr / pq : 0.0;
  double s[1] = {0.0};
  for (long n = (long)blockIdx.x * blockDim.x + threadIdx.x; n < N;
       n += (long)gridDim.x * blockDim.x) {
    x[n] += alpha * p[n];
    const double rn = r[n] - alpha * q[n];
    r[n] = rn;
    s[0] += rn * rn;
  }
  block_reduce_store<1>(s, part);
}

/* beta = rr_new / rr_old; p = r + beta p */
__global__ void __launch_bounds__(256)
k_proj_p(const double *__restrict__ r, double *__restrict__ p, const double *__restrict__ scal,
         int old_slot, int new_slot, long N) {
  const double ro = scal[old_slot], rn = scal[new_slot];
  const double beta = (ro > 0.0 && !proj_done(scal, old_slot)) ? rn / ro : 0.0;
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) p[n] = r[n] + beta * p[n];
}

/* u_a(x) -= mean over the open faces along a of the potential difference */
__global__ void __launch_bounds__(256)
k_proj_correct(const uint16_t *__restrict__ flags, const double *__restrict__ phi,
               float *__restrict__ U, int CU, Geom G) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.N || !proj_fluid(flags, n)) return;
  const ProjNb b = proj_nb(n, G);
  const double f0 = phi[n];
  for (int a = 0; a < 3; a++) {
    double d = 0.0;
    int c = 0;
    if (proj_fluid(flags, b.xp[a])) { d += phi[b.xp[a]] - f0; c++; }
    if (proj_fluid(flags, b.xm[a])) { d += f0 - phi[b.xm[a]]; c++; }
    if (c > 0) U[n * CU + a] = (float)((double)U[n * CU + a] - d / c);
  }
}

extern "C" double slbd_project(slbd_lattice *L, float *U, int CU, int iters, int passes) {
  if (iters < 1 || passes < 1 || CU < 3) return 0.0;
  const long N = L->N;
  double *w = NULL;
  HIPCHECK(hipMalloc(&w, sizeof(double) * (4 * (size_t)N + 8)));
  double *x = w, *r = w + N, *p = w + 2 * N, *q = w + 3 * N, *scal = w + 4 * N;
  const int nb = red_blocks(N);
  double h[4] = {0.0, 0.0, 0.0, 0.0};
  /* each pass: divergence of the current field, CG, correction.  The node
   * correction (mean of the face differences) removes the face divergence
   * only approximately, so later passes remove what the earlier ones left.
   * scal[0], scal[1]: r.r of alternate iterations; scal[2]: p.Ap; scal[3]:
   * r.r at the start of the pass */
  for (int pass = 0; pass < passes; pass++) {
    hipLaunchKernelGGL(k_proj_init, dim3(nb), dim3(256), 0, L->stream, L->flags, U, CU, x, r, p,
                       L->G, L->part);
    LAUNCHCHECK();
    hipLaunchKernelGGL(k_final<1>, dim3(1), dim3(256), 0, L->stream, L->part, nb, scal + 0);
    hipLaunchKernelGGL(k_final<1>, dim3(1), dim3(256), 0, L->stream, L->part, nb, scal + 3);
    LAUNCHCHECK();
    int last = 0;   /* slot of the latest r.r */
    for (int it = 0; it < iters; it++) {
      const int cur = it & 1, nxt = cur ^ 1;
      if (it > 0 && (it & 31) == 0) {
        /* stop early once converged (a converged iteration is a no-op) */
        HIPCHECK(hipMemcpyAsync(h, scal, sizeof h, hipMemcpyDeviceToHost, L->stream));
        HIPCHECK(hipStreamSynchronize(L->stream));
        if (!(h[cur] > PROJ_TOL * PROJ_TOL * h[3])) break;
      }
      last = nxt;
      hipLaunchKernelGGL(k_proj_ap, dim3(nb), dim3(256), 0, L->stream, L->flags, p, q, L->G,
                         L->part);
      hipLaunchKernelGGL(k_final<1>, dim3(1), dim3(256), 0, L->stream, L->part, nb, scal + 2);
      hipLaunchKernelGGL(k_proj_xr, dim3(nb), dim3(256), 0, L->stream, p, q, x, r, scal, cur, N,
                         L->part);
      hipLaunchKernelGGL(k_final<1>, dim3(1), dim3(256), 0, L->stream, L->part, nb, scal + nxt);
      hipLaunchKernelGGL(k_proj_p, lin_grid(N), dim3(256), 0, L->stream, r, p, scal, cur, nxt, N);
      LAUNCHCHECK();
    }
    hipLaunchKernelGGL(k_proj_correct, lin_grid(N), dim3(256), 0, L->stream, L->flags, x, U, CU,
                       L->G);
    LAUNCHCHECK();
    HIPCHECK(hipMemcpyAsync(h, scal, sizeof h, hipMemcpyDeviceToHost, L->stream));
    HIPCHECK(hipStreamSynchronize(L->stream));
    h[2] = h[last];
  }
  HIPCHECK(hipFree(w));
  const double r0 = h[3], rl = h[2];
  return r0 > 0.0 ? sqrt(rl / r0) : 0.0;
}

/* ---------------------------------------------------------------------- */
/* Random sequential addition: candidate pre-filter.                       */
/*                                                                        */
/* Trial t of the splitmix64 stream from `seed` draws its centre from the  */
/* stream's outputs 3t+1, 3t+2, 3t+3 (x, y, z), and output k is            */
/* mix(seed + k * 0x9E3779B97F4A7C15): any trial's centre is computed      */
/* directly.  A batch of candidates is tested in parallel against the      */
/* spheres accepted before the batch (cell grid, minimum image); a         */
/* candidate that overlaps one of them is rejected by the sequential       */
/* process too, since that sphere was accepted first.  The host then runs  */
/* its exact sequential test on the survivors only, in trial order, so the */
/* accepted spheres are the sequential process's.  The device test drops   */
/* only clear overlaps (squared distance below (1 - 1e-9) of the limit):   */
/* its rounding never decides a case the host would decide otherwise.      */

__host__ __device__ inline uint64_t rsa_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256)
k_rsa_filter(uint64_t seed, long t0, int m, const int *__restrict__ head,
             const int *__restrict__ next, const double *__restrict__ xyz, int nc0, int nc1,
             int nc2, double cs0, double cs1, double cs2, double L0, double L1, double L2,
             double d2lim, uint8_t *__restrict__ pass) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const uint64_t g = 0x9E3779B97F4A7C15ULL;
  const uint64_t k0 = 3ULL * (uint64_t)(t0 + q);
  const double s53 = 1.0 / 9007199254740992.0;
  const double L[3] = {L0, L1, L2}, cs[3] = {cs0, cs1, cs2};
  const int nc[3] = {nc0, nc1, nc2};
  double c[3];
  int ci[3];
  for (int d = 0; d < 3; d++) {
    c[d] = (double)(rsa_mix(seed + (k0 + 1 + d) * g) >> 11) * s53 * L[d];
    int v = (int)floor(c[d] / cs[d]);
    ci[d] = v >= nc[d] ? nc[d] - 1 : (v < 0 ? 0 : v);
  }
  uint8_t ok = 1;
  for (int a = -1; a <= 1 && ok; a++)
    for (int b = -1; b <= 1 && ok; b++)
      for (int e = -1; e <= 1 && ok; e++) {
        const int x = (ci[0] + a + nc0) % nc0, y = (ci[1] + b + nc1) % nc1,
                  z = (ci[2] + e + nc2) % nc2;
        for (int s = head[(x * nc1 + y) * nc2 + z]; s >= 0; s = next[s]) {
          double r2 = 0.0;
          for (int d = 0; d < 3; d++) {
            double dx = c[d] - xyz[3 * s + d];
            dx -= L[d] * floor(dx / L[d] + 0.5);
            r2 += dx * dx;
          }
          if (r2 < d2lim) { ok = 0; break; }
        }
      }
  pass[q] = ok;
}

struct slbd_rsa {
  int nc[3], ncell, cap, n;
  double cs[3], L[3];
  int *head, *next;
  double *xyz;
  uint8_t *pass;
  int pass_cap;
  hipStream_t stream;
};

extern "C" slbd_rsa *slbd_rsa_create(const double L[3], const double cs[3], const int nc[3]) {
  slbd_rsa *R = (slbd_rsa *)calloc(1, sizeof(slbd_rsa));
  for (int d = 0; d < 3; d++) { R->nc[d] = nc[d]; R->cs[d] = cs[d]; R->L[d] = L[d]; }
  R->ncell = nc[0] * nc[1] * nc[2];
  R->cap = 4096;
  HIPCHECK(hipStreamCreateWithFlags(&R->stream, hipStreamNonBlocking));
  HIPCHECK(hipMalloc(&R->head, sizeof(int) * R->ncell));
  HIPCHECK(hipMalloc(&R->next, sizeof(int) * R->cap));
  HIPCHECK(hipMalloc(&R->xyz, sizeof(double) * 3 * R->cap));
  HIPCHECK(hipMemsetAsync(R->head, 0xff, sizeof(int) * R->ncell, R->stream));
  return R;
}

/* The accepted spheres so far: the host's cell heads, and spheres n0.. with
 * their list links and centres. */
extern "C" void slbd_rsa_update(slbd_rsa *R, const int *head, const int *next,
                                const double *xyz, int n) {
  if (n > R->cap) {
    int cap = R->cap;
    while (cap < n) cap *= 2;
    int *nx;
    double *xy;
    HIPCHECK(hipMalloc(&nx, sizeof(int) * cap));
    HIPCHECK(hipMalloc(&xy, sizeof(double) * 3 * cap));
    HIPCHECK(hipStreamSynchronize(R->stream));
    HIPCHECK(hipFree(R->next));
    HIPCHECK(hipFree(R->xyz));
    R->next = nx; R->xyz = xy; R->cap = cap;
    R->n = 0;   /* everything is uploaded again below */
  }
  if (n > R->n) {
    HIPCHECK(hipMemcpyAsync(R->next + R->n, next + R->n, sizeof(int) * (n - R->n),
                            hipMemcpyHostToDevice, R->stream));
    HIPCHECK(hipMemcpyAsync(R->xyz + 3L * R->n, xyz + 3L * R->n, sizeof(double) * 3 * (n - R->n),
                            hipMemcpyHostToDevice, R->stream));
  }
  HIPCHECK(hipMemcpyAsync(R->head, head, sizeof(int) * R->ncell, hipMemcpyHostToDevice, R->stream));
  HIPCHECK(hipStreamSynchronize(R->stream));
  R->n = n;
}

/* pass[q] = 0 when trial t0 + q clearly overlaps a sphere uploaded so far. */
extern "C" void slbd_rsa_filter(slbd_rsa *R, uint64_t seed, long t0, int m, double d2min,
                                uint8_t *pass) {
  if (m > R->pass_cap) {
    if (R->pass) HIPCHECK(hipFree(R->pass));
    HIPCHECK(hipMalloc(&R->pass, m));
    R->pass_cap = m;
  }
  hipLaunchKernelGGL(k_rsa_filter, dim3((m + 255) / 256), dim3(256), 0, R->stream, seed, t0, m,
                     R->head, R->next, R->xyz, R->nc[0], R->nc[1], R->nc[2], R->cs[0], R->cs[1],
                     R->cs[2], R->L[0], R->L[1], R->L[2], d2min * (1.0 - 1e-9), R->pass);
  LAUNCHCHECK();
  HIPCHECK(hipMemcpyAsync(pass, R->pass, m, hipMemcpyDeviceToHost, R->stream));
  HIPCHECK(hipStreamSynchronize(R->stream));
}

/* Sphere images of a pack on the device: for each image, the nodes of its
 * bounding box (clipped to the lattice by the caller, one node of padding)
 * inside the sphere by insertsph's predicate (src/lbobjs.c:212-218),
 * (i - ox)^2 + (j - oy)^2 + (k - oz)^2 < r^2 in FP64 with every product and
 * sum rounded on its own (contraction into fused multiply-adds is off in
 * the kernel, as in the host code built with -ffp-contract=off) and the
 * squares added left to right.  mode 0 counts each image's nodes, mode 1
 * marks them in a byte map. */
struct SphImg {
  double o[3];
  int lo[3], hi[3];
};

__global__ void __launch_bounds__(256)
k_sph_images(const SphImg *__restrict__ img, double radsq, int esj, int esk, int mode,
             int *__restrict__ counts, uint8_t *__restrict__ map) {
#pragma clang fp contract(off)
  __shared__ int red[256];
  const SphImg S = img[blockIdx.x];
  const int ni = S.hi[0] - S.lo[0] + 1, nj = S.hi[1] - S.lo[1] + 1, nk = S.hi[2] - S.lo[2] + 1;
  const long tot = (ni > 0 && nj > 0 && nk > 0) ? (long)ni * nj * nk : 0;
  int c = 0;
  for (long q = threadIdx.x; q < tot; q += blockDim.x) {
    const int a = (int)(q / ((long)nj * nk));
    const int r = (int)(q - (long)a * nj * nk);
    const int b = r / nk, e = r - b * nk;
    const int i = S.lo[0] + a, j = S.lo[1] + b, k = S.lo[2] + e;
    const double dx = (double)i - S.o[0], dy = (double)j - S.o[1], dz = (double)k - S.o[2];
    const double part = dx * dx + dy * dy;
    if (part + dz * dz < radsq) {
      if (mode) map[((long)i * esj + j) * esk + k] = 1;
      else c++;
    }
  }
  if (mode) return;
  red[threadIdx.x] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[blockIdx.x] = red[0];
}

/* images: n rows of {ox, oy, oz} (o) and {ilo, jlo, klo, ihi, jhi, khi} (box),
 * staged in a grow-only buffer of the lattice (images, then n counts) */
static SphImg *sph_upload(slbd_lattice *L, const double *o, const int *box, int n) {
  const size_t need = (sizeof(SphImg) + sizeof(int)) * (size_t)n;
  if (need > L->sph_cap) {
    if (L->sph_buf) HIPCHECK(hipFree(L->sph_buf));
    L->sph_cap = need + need / 2;
    L->sph_buf = dalloc(L->sph_cap);
  }
  SphImg *h = (SphImg *)malloc(sizeof(SphImg) * (size_t)n);
  if (!h) { fprintf(stderr, "sunlightlb: out of host memory\n"); exit(EXIT_FAILURE); }
  for (int q = 0; q < n; q++)
    for (int d = 0; d < 3; d++) {
      h[q].o[d] = o[3 * q + d];
      h[q].lo[d] = box[6 * q + d];
      h[q].hi[d] = box[6 * q + 3 + d];
    }
  SphImg *dv = (SphImg *)L->sph_buf;
  copy_in(dv, h, sizeof(SphImg) * (size_t)n, L->stream);
  free(h);
  return dv;
}

extern "C" void slbd_sph_count(slbd_lattice *L, const double *o, const int *box, int n,
                               double radsq, int *counts) {
  if (n <= 0) return;
  SphImg *dv = sph_upload(L, o, box, n);
  int *dc = (int *)(dv + n);
  hipLaunchKernelGGL(k_sph_images, dim3((unsigned)n), dim3(256), 0, L->stream, dv, radsq,
                     L->esj, L->esk, 0, dc, (uint8_t *)nullptr);
  LAUNCHCHECK();
  copy_out(counts, dc, sizeof(int) * (size_t)n, L->stream);
}

extern "C" void slbd_sph_map(slbd_lattice *L, const double *o, const int *box, int n,
                             double radsq, uint8_t *map) {
  uint8_t *dm = (uint8_t *)dalloc((size_t)L->N);
  HIPCHECK(hipMemsetAsync(dm, 0, (size_t)L->N, L->stream));
  if (n > 0) {
    SphImg *dv = sph_upload(L, o, box, n);
    hipLaunchKernelGGL(k_sph_images, dim3((unsigned)n), dim3(256), 0, L->stream, dv, radsq,
                       L->esj, L->esk, 1, (int *)nullptr, dm);
    LAUNCHCHECK();
  }
  copy_out(map, dm, (size_t)L->N, L->stream);
  HIPCHECK(hipFree(dm));
}

extern "C" void slbd_rsa_destroy(slbd_rsa *R) {
  if (!R) return;
  HIPCHECK(hipStreamSynchronize(R->stream));
  HIPCHECK(hipFree(R->head));
  HIPCHECK(hipFree(R->next));
  HIPCHECK(hipFree(R->xyz));
  if (R->pass) HIPCHECK(hipFree(R->pass));
  HIPCHECK(hipStreamDestroy(R->stream));
  free(R);
}
