/*
 * slb_kern.h -- device-side building blocks shared by the gfx950 kernels of
 * the single-lattice path (slb_dev.hip) and the batched path (slb_batch.hip):
 * the D3Q15 tables, periodic neighbour offsets, the collision of one node and
 * the wall velocity of an object node.
 */
#ifndef SLB_KERN_H
#define SLB_KERN_H

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <type_traits>

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

/* ---------------------------------------------------------------------- */
/* D3Q15 tables (the reference's src/lbmisc.c:89-104 ordering).            */

__host__ __device__ constexpr int CX(int l) {
  return (l == 1 || l == 7 || l == 9 || l == 11 || l == 13) ? 1
       : (l == 2 || l == 8 || l == 10 || l == 12 || l == 14) ? -1 : 0;
}
__host__ __device__ constexpr int CY(int l) {
  return (l == 3 || l == 7 || l == 8 || l == 11 || l == 12) ? 1
       : (l == 4 || l == 9 || l == 10 || l == 13 || l == 14) ? -1 : 0;
}
__host__ __device__ constexpr int CZ(int l) {
  return (l == 5 || (l >= 7 && l <= 10)) ? 1 : (l == 6 || l >= 11) ? -1 : 0;
}
__host__ __device__ constexpr int WL(int l) { return l == 0 ? 16 : (l <= 6 ? 8 : 1); }
__host__ __device__ constexpr int OPP(int l) {
  return l == 0 ? 0 : (l <= 6 ? ((l & 1) ? l + 1 : l - 1) : 21 - l);
}

/* Population layout in HBM: tiles of POP_TILE consecutive nodes, each tile
 * holding the 15 populations as 15 contiguous runs of POP_TILE values
 * (structure of arrays within a tile).  A block's own populations and
 * those of its row neighbours then lie in a few contiguous 60-120 KB
 * stretches rather than in 15 arrays N apart, so the bounce-back loads of
 * a node's own opposite populations stay within the pages and lines the
 * pull is already touching.  Allocations hold pop_nodes(N) nodes (N rounded
 * up to a whole tile). */
#ifndef POP_SHIFT
#define POP_SHIFT 10
#endif
#define POP_TILE (1L << POP_SHIFT)
template <typename I>
__host__ __device__ __forceinline__ I pidx_t(int l, I x) {
  return ((x >> POP_SHIFT) * I(15) + I(l)) * I(POP_TILE) + (x & I(POP_TILE - 1));
}
__host__ __device__ __forceinline__ long pidx(int l, long x) { return pidx_t<long>(l, x); }
__host__ __device__ __forceinline__ long pop_nodes(long N) {
  return (N + POP_TILE - 1) & ~(POP_TILE - 1);
}

#define FLAG_SOLID 0x4000u
#define FLAG_LINKS 0x3fffu

/* Neighbour offsets along one axis for the node at coordinate q of an axis of
 * length n and stride st: off[0] for component -1 (source q+1), off[2] for
 * component +1 (source q-1); used with sign = -1 for pull (x - c) or +1
 * for push (x + c). */
template <typename I>
struct AxisOffT {
  I m, p;  /* offset of q-1 and q+1 with periodic wrap */
};
typedef AxisOffT<long> AxisOff;
template <typename I>
__device__ __forceinline__ AxisOffT<I> axis_off_t(int q, int n, I st) {
  AxisOffT<I> o;
  o.m = (q > 0) ? -st : I(n - 1) * st;
  o.p = (q < n - 1) ? st : -I(n - 1) * st;
  return o;
}
__device__ __forceinline__ AxisOff axis_off(int q, int n, long st) {
  return axis_off_t<long>(q, n, st);
}
/* offset of x + s*c_l (s = +1 push, -1 pull) */
template <int L, int S, typename I>
__device__ __forceinline__ I nb_off(const AxisOffT<I> &oi, const AxisOffT<I> &oj,
                                    const AxisOffT<I> &ok) {
  constexpr int cx = S * CX(L), cy = S * CY(L), cz = S * CZ(L);
  I o = 0;
  if (cx > 0) o += oi.p; else if (cx < 0) o += oi.m;
  if (cy > 0) o += oj.p; else if (cy < 0) o += oj.m;
  if (cz > 0) o += ok.p; else if (cz < 0) o += ok.m;
  return o;
}

/* 32-bit forms of the above, for lattices of fewer than 2^31 populations:
 * half the address registers and arithmetic. */
__host__ __device__ __forceinline__ int pidx32(int l, int x) { return pidx_t<int>(l, x); }
typedef AxisOffT<int> AxisOff32;
__device__ __forceinline__ AxisOff32 axis_off32(int q, int n, int st) {
  return axis_off_t<int>(q, n, st);
}

struct Geom {
  int esi, esj, esk, plane;
  long N;
  int xcd_map;   /* step kernel block order: 0 plain, 1 contiguous range per
                    XCD, T >= 2 tiles of T x T rows per XCD */
  int nt_store;  /* step kernel: non-temporal population stores */
  int nt_load;   /* step kernel: non-temporal population loads */
  int step_tpb;  /* step kernel: threads per block (256 or 512) */
  int mp_xcd;    /* moment-propagation kernel block order (as xcd_map: 0 or 1) */
};

/* Block coordinates of the step kernel (blockIdx.x: block of rows within a
 * plane, blockIdx.y: plane i).  Workgroups are dispatched round robin over
 * the 8 XCDs, each with its own L2.  A bounce-back link makes a node read
 * its own opposite population, a cache line that the row on the other side
 * of the link also reads; when the two rows run on different XCDs, or far
 * apart in time, the line is fetched from HBM twice.
 *   xcd_map == 1: XCD b % 8 walks a contiguous 1/8 of the rows.
 *   xcd_map == T: the rows are cut in T x T tiles of (plane, row block);
 *   XCD x takes tiles x, x + 8, x + 16, ... in dispatch order and walks each
 *   tile row block by row block, so both neighbours of a row are within T
 *   blocks on the same XCD, while the 8 XCDs stay in neighbouring tiles.
 * The caller sets xcd_map only when the grid divides (slbd_step_order). */
__device__ __forceinline__ void step_block(const Geom &G, int &bx, int &by) {
  bx = blockIdx.x;
  by = blockIdx.y;
  if (G.xcd_map == 1) {
    const int nbx = gridDim.x, nb = gridDim.x * gridDim.y;
    if ((nb & 7) == 0) {
      const int lin = by * nbx + bx;
      const int lin2 = (lin & 7) * (nb >> 3) + (lin >> 3);
      bx = lin2 % nbx;
      by = lin2 / nbx;
    }
  } else if (G.xcd_map >= 2) {
    const int T = G.xcd_map, TT = T * T, nbx = gridDim.x;
    const int lin = by * nbx + bx;
    const int k = lin >> 3;
    const int tile = (k / TT) * 8 + (lin & 7), w = k - (k / TT) * TT;
    const int ntx = nbx / T;
    const int ti = tile / ntx, tj = tile - ti * ntx;
    by = ti * T + w / T;
    bx = tj * T + (w - (w / T) * T);
  }
}

/* Compile-time unrolled loop over links 1..14 */
template <int L, int END>
struct Links {
  template <typename F>
  __device__ __forceinline__ static void run(F &&f) {
    f(std::integral_constant<int, L>{});
    Links<L + 1, END>::run(f);
  }
};
template <int END>
struct Links<END, END> {
  template <typename F>
  __device__ __forceinline__ static void run(F &&) {}
};

/* ---------------------------------------------------------------------- */
/* Collision of one node (src/lbkernel.c:797-866).                        */

struct PhysArgs {
  double omega, comega, rho0;
  int stokes, skip_solid;
  double bfx, bfy, bfz;
};

/* Post-collision moments of a node: density deviation, momentum (with the
 * body force) and the relaxed traceless-shifted stress.  The 15
 * post-collision populations are rebuilt from these 10 numbers alone. */
template <typename T>
struct Mom {
  T drho, jx, jy, jz, sxx, syy, szz, sxy, sxz, syz;
};

template <typename T>
__device__ __forceinline__ Mom<T> collide_moments(const T v[15], T jbx, T jby, T jbz,
                                                  const PhysArgs &P) {
  const T ax = v[1] + v[2], dx = v[1] - v[2];
  const T ay = v[3] + v[4], dy = v[3] - v[4];
  const T az = v[5] + v[6], dz = v[5] - v[6];
  /* diagonal links 7..14, c = (sx,sy,sz): sums weighted by 1, sx, sy, sz,
   * sx sy, sx sz, sy sz */
  const T p7 = v[7] + v[14], m7 = v[7] - v[14];     /* (+,+,+) */
  const T p8 = v[8] + v[13], m8 = v[8] - v[13];     /* (-,+,+) */
  const T p9 = v[9] + v[12], m9 = v[9] - v[12];     /* (+,-,+) */
  const T p10 = v[10] + v[11], m10 = v[10] - v[11]; /* (-,-,+) */
  const T D = (p7 + p8) + (p9 + p10);
  const T Dx = (m7 - m8) + (m9 - m10);
  const T Dy = (m7 + m8) - (m9 + m10);
  const T Dz = (m7 + m8) + (m9 + m10);
  const T Dxy = (p7 - p8) - (p9 - p10);
  const T Dxz = (p7 - p8) + (p9 - p10);
  const T Dyz = (p7 + p8) - (p9 + p10);
  const T a0a = v[0] + ax + ay + az;
  Mom<T> m;
  m.drho = a0a + D;
  m.jx = dx + Dx + jbx;
  m.jy = dy + Dy + jby;
  m.jz = dz + Dz + jbz;
  /* traceless-shifted stress s = Pi - drho/3 I (src/lbkernel.c:821-823) */
  const T sh = T(2.0 / 3.0) * m.drho - a0a;
  T sxx = sh + ax, syy = sh + ay, szz = sh + az;
  T sxy = Dxy, sxz = Dxz, syz = Dyz;
  const T cw = T(P.comega);
  if (P.stokes) {
    sxx *= cw; syy *= cw; szz *= cw; sxy *= cw; sxz *= cw; syz *= cw;
  } else {
    const T o = T(P.omega) / (T(P.rho0) + m.drho);
    sxx = cw * sxx + o * m.jx * m.jx;
    syy = cw * syy + o * m.jy * m.jy;
    szz = cw * szz + o * m.jz * m.jz;
    sxy = cw * sxy + o * m.jx * m.jy;
    sxz = cw * sxz + o * m.jx * m.jz;
    syz = cw * syz + o * m.jy * m.jz;
  }
  m.sxx = sxx; m.syy = syy; m.szz = szz; m.sxy = sxy; m.sxz = sxz; m.syz = syz;
  return m;
}

/* Post-collision population of link l rebuilt from the moments,
 * a_l = (w_l/72)[drho + 3 c.j + 4.5 (cc:s - tr s/3)] (src/lbkernel.c:837-865),
 * with no contraction into fused multiply-adds, so that the rebuild gives the
 * same value wherever it is evaluated (the step that stores populations and
 * the step that pulls them from stored moments). */
template <int l, typename T>
__device__ __forceinline__ T rebuild_link(const Mom<T> &m) {
#pragma clang fp contract(off)
  const T r72 = m.drho * T(1.0 / 72.0);
  const T t3 = (m.sxx + m.syy + m.szz) * T(1.0 / 3.0);
  if constexpr (l == 0) {
    const T r9 = T(8.0) * r72;
    return T(2.0) * r9 - t3;
  } else if constexpr (l <= 6) {
    const T r9 = T(8.0) * r72;
    const T sd = (l <= 2) ? m.sxx : (l <= 4) ? m.syy : m.szz;
    const T jd = (l <= 2) ? m.jx : (l <= 4) ? m.jy : m.jz;
    const T q = r9 + T(0.5) * (sd - t3);
    const T g = T(1.0 / 3.0) * jd;
    return (l & 1) ? q + g : q - g;
  } else if constexpr (l == 7 || l == 14) {
    const T q = r72 + T(0.125) * (t3 + m.sxy + m.sxz + m.syz);
    const T g = T(1.0 / 24.0) * (m.jx + m.jy + m.jz);
    return (l == 7) ? q + g : q - g;
  } else if constexpr (l == 8 || l == 13) {
    const T q = r72 + T(0.125) * (t3 - m.sxy - m.sxz + m.syz);
    const T g = T(1.0 / 24.0) * (-m.jx + m.jy + m.jz);
    return (l == 8) ? q + g : q - g;
  } else if constexpr (l == 9 || l == 12) {
    const T q = r72 + T(0.125) * (t3 - m.sxy + m.sxz - m.syz);
    const T g = T(1.0 / 24.0) * (m.jx - m.jy + m.jz);
    return (l == 9) ? q + g : q - g;
  } else {
    const T q = r72 + T(0.125) * (t3 + m.sxy - m.sxz - m.syz);
    const T g = T(1.0 / 24.0) * (-m.jx - m.jy + m.jz);
    return (l == 10) ? q + g : q - g;
  }
}

template <typename T>
__device__ __forceinline__ void rebuild_all(const Mom<T> &m, T v[15]) {
  v[0] = rebuild_link<0>(m);
  v[1] = rebuild_link<1>(m); v[2] = rebuild_link<2>(m);
  v[3] = rebuild_link<3>(m); v[4] = rebuild_link<4>(m);
  v[5] = rebuild_link<5>(m); v[6] = rebuild_link<6>(m);
  v[7] = rebuild_link<7>(m); v[8] = rebuild_link<8>(m);
  v[9] = rebuild_link<9>(m); v[10] = rebuild_link<10>(m);
  v[11] = rebuild_link<11>(m); v[12] = rebuild_link<12>(m);
  v[13] = rebuild_link<13>(m); v[14] = rebuild_link<14>(m);
}

/* Collision of one node (src/lbkernel.c:797-866): moments, relaxation,
 * rebuild of the 15 populations in place. */
template <typename T>
__device__ __forceinline__ void collide_node(T v[15], T jbx, T jby, T jbz,
                                             const PhysArgs &P) {
  const Mom<T> m = collide_moments<T>(v, jbx, jby, jbz, P);
  rebuild_all<T>(m, v);
}

/* Wall velocity of solid nodes from the object table, src/lbobjs.c:565-589:
 * U = rho0 (v + w x r)/12, r = node - centre (no minimum image). */
__device__ __forceinline__ void obj_velocity(const double *__restrict__ tab, int o,
                                             int i, int j, int k, double rho0,
                                             double &Ux, double &Uy, double &Uz) {
  const double *p = tab + (long)o * SLBD_OBJ_STRIDE;
  const double rx = i - p[0], ry = j - p[1], rz = k - p[2];
  Ux = rho0 * (p[3] + p[7] * rz - p[8] * ry) / 12.0;
  Uy = rho0 * (p[4] + p[8] * rx - p[6] * rz) / 12.0;
  Uz = rho0 * (p[5] + p[6] * ry - p[7] * rx) / 12.0;
}

#endif
