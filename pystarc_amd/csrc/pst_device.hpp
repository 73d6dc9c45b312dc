// Device helpers for the batched BD step on gfx950: counter-based RNG,
// quaternion algebra, trilinear grid gather and the Yukawa multipole
// far-field.  Every physics formula cites the PySTARC reference it restates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include "pst_types.hpp"

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11, "Parallel random numbers: as easy as
// 1, 2, 3"): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments
// 0x9E3779B9 / 0xBB67AE85.  Counter = (traj lo, traj hi, step, stream).
// ---------------------------------------------------------------------------
struct PhiloxOut { uint32_t v[4]; };

__host__ __device__ inline PhiloxOut philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                   uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  PhiloxOut o;
  o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
  return o;
}

enum { RNG_INIT = 1, RNG_ESC = 2, RNG_RETQ = 3, RNG_NOISE = 4, RNG_BRIDGE = 5, RNG_OUTER = 6 };

// Force-kernel geometry shared by the direct and the binned gathers.
constexpr int kForceBlock = 256;
constexpr int kLdsAtomsMax = 1536;  // ligand atoms staged in LDS (<= 48 KB in FP64)

// Whether receptor run k (PST_REC_CHUNK consecutive atoms, bounding sphere
// rec_chunk[k]) can hold an atom within `reach` of (cx, cy, cz).  False only
// when every atom of the run is farther than reach (the sphere radius and the
// reaches it is used with carry margins above rounding).
__device__ inline bool rec_chunk_near(const ForceSys& S, int k, double cx, double cy, double cz,
                                      double reach) {
  const double* s = S.rec_chunk + 4 * k;
  const double dx = s[0] - cx, dy = s[1] - cy, dz = s[2] - cz;
  const double lim = reach + s[3];
  return dx * dx + dy * dy + dz * dz <= lim * lim;
}

// Sum over the G lanes of a trajectory group (butterfly, FP64).
template <int G>
__device__ inline double group_sum(double v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

// Two uniforms in [0,1) with 53 random bits each.
__device__ inline void philox_uniform2(uint64_t seed, uint64_t traj, uint32_t step,
                                       uint32_t stream, uint32_t sub, double& u0, double& u1) {
  PhiloxOut o = philox4x32_10((uint32_t)traj, (uint32_t)(traj >> 32), step,
                              stream | (sub << 8), (uint32_t)seed, (uint32_t)(seed >> 32));
  uint64_t a = ((uint64_t)o.v[0] << 21) ^ (uint64_t)(o.v[1] >> 11);
  uint64_t b = ((uint64_t)o.v[2] << 21) ^ (uint64_t)(o.v[3] >> 11);
  a &= ((1ull << 53) - 1);
  b &= ((1ull << 53) - 1);
  const double s = 1.0 / 9007199254740992.0;  // 2^-53
  u0 = (double)a * s;
  u1 = (double)b * s;
}

// sin and cos of 2 pi u.  Through sincospi the argument needs no reduction
// (2u is reduced exactly), so the general sin / cos and their large-argument
// reduction (dozens of registers and a scratch frame in every kernel that
// draws a normal or a rotation) stay out of the step kernels.  The values
// differ from sin(2 pi u) by about one ulp.
__device__ inline void sincos_2pi(double u, double* s, double* c) { sincospi(2.0 * u, s, c); }

// Two standard normals by Box-Muller from one Philox block.
__device__ inline void philox_normal2(uint64_t seed, uint64_t traj, uint32_t step,
                                      uint32_t stream, uint32_t sub, double& z0, double& z1) {
  double u0, u1;
  philox_uniform2(seed, traj, step, stream, sub, u0, u1);
  const double one_minus = 1.0 - u0;  // in (0, 1]
  double r = sqrt(-2.0 * log(one_minus));
  double s, c;
  sincos_2pi(u1, &s, &c);
  z0 = r * c;
  z1 = r * s;
}

// ---------------------------------------------------------------------------
// Quaternions, (w, x, y, z) order.
// ---------------------------------------------------------------------------
// R(q), gpu_batch_simulator.py:1611-1629.
__device__ inline void quat_to_rot(const double* q, double* R) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - w * z); R[2] = 2 * (x * z + w * y);
  R[3] = 2 * (x * y + w * z); R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - w * x);
  R[6] = 2 * (x * z - w * y); R[7] = 2 * (y * z + w * x); R[8] = 1 - 2 * (x * x + y * y);
}

// Shoemake uniform rotation from three uniforms, gpu_batch_simulator.py:1597-1609.
__device__ inline void shoemake(double u0, double u1, double u2, double* q) {
  const double a = sqrt(1.0 - u0), b = sqrt(u0);
  double s1, c1, s2, c2;
  sincos_2pi(u1, &s1, &c1);
  sincos_2pi(u2, &s2, &c2);
  q[0] = a * s1;
  q[1] = a * c1;
  q[2] = b * s2;
  q[3] = b * c2;
}

// q <- normalize(q (x) dq(omega)), gpu_batch_simulator.py:1631-1660.
__device__ inline void apply_rotation(double* q, double ox, double oy, double oz) {
  const double ang = sqrt(ox * ox + oy * oy + oz * oz);
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (ang > 1e-10) { ax = ox / ang; ay = oy / ang; az = oz / ang; }
  // sin and cos of ang / 2 as sincospi of ang / (2 pi) (see sincos_2pi)
  double s, c;
  sincospi(ang * (0.5 * M_1_PI), &s, &c);
  const double w2 = c, x2 = ax * s, y2 = ay * s, z2 = az * s;
  const double w1 = q[0], x1 = q[1], y1 = q[2], z1 = q[3];
  double nw = w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2;
  double nx = w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2;
  double ny = w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2;
  double nz = w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2;
  const double n = sqrt(nw * nw + nx * nx + ny * ny + nz * nz);
  q[0] = nw / n; q[1] = nx / n; q[2] = ny / n; q[3] = nz / n;
}

__device__ inline double norm3(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }

// Cube-map cell of a direction d (the order of the split evaluation's
// reverse-Born pass: d = R^T (-c), the receptor centre seen from the ligand):
// face = dominant axis and its sign (6), then the two other components over
// the dominant one, each in kDirSide cells; face-major, row-major, so
// neighbouring cells are neighbouring directions.
constexpr int kDirSide = 6;
constexpr int kDirCells = 6 * kDirSide * kDirSide;   // 216
__device__ inline int dir_cell(double dx, double dy, double dz) {
  const double d[3] = {dx, dy, dz};
  const double ax = fabs(dx), ay = fabs(dy), az = fabs(dz);
  const int dom = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
  const double m = fmax(fabs(d[dom]), 1e-300);
  const int u = dom == 0 ? 1 : 0, v = dom == 2 ? 1 : 2;
  int iu = (int)((d[u] / m + 1.0) * (0.5 * kDirSide));
  int iv = (int)((d[v] / m + 1.0) * (0.5 * kDirSide));
  iu = iu < 0 ? 0 : (iu >= kDirSide ? kDirSide - 1 : iu);
  iv = iv < 0 ? 0 : (iv >= kDirSide ? kDirSide - 1 : iv);
  const int face = 2 * dom + (d[dom] > 0 ? 1 : 0);
  return (face * kDirSide + iu) * kDirSide + iv;
}

// ---------------------------------------------------------------------------
// Trilinear gather, batch_force_kernel (gpu_batch_engine.py:95-110): cell
// index floor((p - o) / sp).  The reference's value 0 for a cell outside
// [0, n-2] never applies here: a gather is evaluated only strictly inside
// the 3-spacing valid box, where the whole stencil is inside the grid.
// Index arithmetic is FP64; weights and sums are in TC (float or double).
// The value type V is TC for one grid, or Pair<TC> when the receptor
// potential and Born grids share a geometry and are stored interleaved
// (float2 per node): one stencil then fetches both fields from the same lines.
// ---------------------------------------------------------------------------
template <typename T> struct Pair { T a, b; };
template <typename T> __device__ inline Pair<T> operator+(Pair<T> x, Pair<T> y) { return {x.a + y.a, x.b + y.b}; }
template <typename T> __device__ inline Pair<T> operator-(Pair<T> x, Pair<T> y) { return {x.a - y.a, x.b - y.b}; }
template <typename T> __device__ inline Pair<T> operator*(Pair<T> x, T s) { return {x.a * s, x.b * s}; }
template <typename T> __device__ inline Pair<T> operator*(T s, Pair<T> x) { return {x.a * s, x.b * s}; }
template <typename T> __device__ inline Pair<T> operator/(Pair<T> x, T s) { return {x.a / s, x.b / s}; }

template <typename V, typename TG> __device__ inline V gld(const TG* d, long i);
template <> __device__ inline double gld<double, double>(const double* d, long i) { return d[i]; }
template <> __device__ inline double gld<double, float>(const float* d, long i) { return (double)d[i]; }
template <> __device__ inline float gld<float, float>(const float* d, long i) { return d[i]; }
template <> __device__ inline float gld<float, double>(const double* d, long i) { return (float)d[i]; }
template <> __device__ inline Pair<float> gld<Pair<float>, float2>(const float2* d, long i) {
  const float2 v = d[i];
  return {v.x, v.y};
}
template <> __device__ inline Pair<double> gld<Pair<double>, float2>(const float2* d, long i) {
  const float2 v = d[i];
  return {(double)v.x, (double)v.y};
}

// Node addressing.  The interpolation arithmetic always works in the grid's
// own index space (cell i0 = floor((p - o) * inv)); an accessor turns node
// (i, j, k) into an offset ox(i) + oy(j) + oz(k) into the array it reads: the
// whole grid in HBM, or a tile of it staged in LDS whose first node is
// (ti, tj, tk).
//
// In HBM every grid is stored in bricks of 4 x 4 x 4 nodes: bricks in C
// order, nodes z-fastest inside a brick (one x-slice of a brick is 16 nodes,
// 64 B of FP32 or 128 B of float2).  The 20 nodes of one atom's stencil then
// sit in ~3-5 cache lines instead of the ~8-9 of a C-order array, where
// every (i, j) column of the stencil is a line of its own.
// ZP: fetch the z nodes of every stencil column as aligned pairs (the direct
// kernel's "z-pair" launch variant, zcol3 below).
template <bool ZP>
struct GlobalAccT {
  static constexpr bool kZPairs = ZP;
  long bx, by;  // nodes per brick step along x and along y
  __device__ long ox(int i) const { return (long)(i >> 2) * bx + ((i & 3) << 4); }
  __device__ long oy(int j) const { return (long)(j >> 2) * by + ((j & 3) << 2); }
  __device__ long oz(int k) const { return ((long)(k >> 2) << 6) + (k & 3); }
  __device__ long at(int i, int j, int k) const { return ox(i) + oy(j) + oz(k); }
};
using GlobalAcc = GlobalAccT<false>;
struct TileAcc {
  static constexpr bool kZPairs = false;
  long sx, sy;
  int ti, tj, tk;
  __device__ long ox(int i) const { return (long)(i - ti) * sx; }
  __device__ long oy(int j) const { return (long)(j - tj) * sy; }
  __device__ long oz(int k) const { return (long)(k - tk); }
  __device__ long at(int i, int j, int k) const { return ox(i) + oy(j) + oz(k); }
};
// Two z-adjacent nodes of one (i, j) column in one 8- or 16-byte load (the
// values are those of two single loads).
typedef float pst_f2a4 __attribute__((ext_vector_type(2), aligned(4)));
typedef float pst_f4a8 __attribute__((ext_vector_type(4), aligned(8)));
typedef double pst_d2a8 __attribute__((ext_vector_type(2), aligned(8)));
template <typename V, typename TG> __device__ inline void gld_pair2(const TG* d, long o, V& lo, V& hi);
template <> __device__ inline void gld_pair2<float, float>(const float* d, long o, float& lo, float& hi) {
  const pst_f2a4 v = *(const pst_f2a4*)(d + o);
  lo = v.x; hi = v.y;
}
template <> __device__ inline void gld_pair2<double, float>(const float* d, long o, double& lo, double& hi) {
  const pst_f2a4 v = *(const pst_f2a4*)(d + o);
  lo = v.x; hi = v.y;
}
template <> __device__ inline void gld_pair2<double, double>(const double* d, long o, double& lo, double& hi) {
  const pst_d2a8 v = *(const pst_d2a8*)(d + o);
  lo = v.x; hi = v.y;
}
template <> __device__ inline void gld_pair2<float, double>(const double* d, long o, float& lo, float& hi) {
  const pst_d2a8 v = *(const pst_d2a8*)(d + o);
  lo = (float)v.x; hi = (float)v.y;
}
template <> __device__ inline void gld_pair2<Pair<float>, float2>(const float2* d, long o, Pair<float>& lo,
                                                                 Pair<float>& hi) {
  const pst_f4a8 v = *(const pst_f4a8*)(d + o);
  lo = {v.x, v.y}; hi = {v.z, v.w};
}
template <> __device__ inline void gld_pair2<Pair<double>, float2>(const float2* d, long o,
                                                                  Pair<double>& lo, Pair<double>& hi) {
  const pst_f4a8 v = *(const pst_f4a8*)(d + o);
  lo = {(double)v.x, (double)v.y}; hi = {(double)v.z, (double)v.w};
}
// Nodes kmin, kmin + 1, kmin + 2 of one column of a bricked grid from its
// aligned pairs (e, e + 1) at offset oa and (e + 2, e + 3) at ob, e = kmin
// rounded down to even (odd: kmin = e + 1).  An even node starts a pair
// inside its brick's z-row, so neither load crosses a brick and every lane
// takes the same two loads, wherever its stencil sits; a wave whose pairs
// were fetched only where they fit in a brick would issue both the pair and
// the single loads whenever its lanes disagree.
template <typename V, typename TG>
__device__ inline void zcol3(const TG* d, long oa, long ob, bool odd, V& m0, V& m1, V& m2) {
  V q0, q1, q2, q3;
  gld_pair2<V, TG>(d, oa, q0, q1);
  gld_pair2<V, TG>(d, ob, q2, q3);
  m0 = odd ? q1 : q0;
  m1 = odd ? q2 : q1;
  m2 = odd ? q3 : q2;
}

template <bool ZP = false>
__device__ inline GlobalAccT<ZP> global_acc(const GridDev& g) {
  const long nbz = (g.nz + 3) >> 2, nby = (g.ny + 3) >> 2;
  return {nby * nbz * 64, nbz * 64};
}


// Strictly inside the valid box (gpu_batch_engine.py:317-319).
__device__ inline bool grid_inside(const GridDev& g, double x, double y, double z) {
  return x > g.lo[0] && y > g.lo[1] && z > g.lo[2] && x < g.hi[0] && y < g.hi[1] && z < g.hi[2];
}

// phi and grad phi by central differences at +-sp/2 of trilinear
// interpolations (batch_force_kernel, gpu_batch_engine.py:95-115: cell index
// floor((p - o) / sp), seven interpolations of 8 nodes each), evaluated from
// the 20 distinct stencil nodes: the probes at p -+ sp/2 along an axis sit in neighbouring
// cells with one common fraction f, so
//   phi(p + h) - phi(p - h) = sum_yz w_yz [(1-f)(v[m+1]-v[m]) + f(v[m+2]-v[m+1])],
// m = floor(i - 1/2).  Neighbouring grid values are subtracted before they
// are weighted, which keeps FP32 free of the cancellation of two nearly
// equal interpolated values, and every node is loaded once.  Callers
// guarantee the whole stencil is inside the grid (the 3-spacing valid box).
//
// The evaluation comes in three parts so that the loads can be issued by
// other lanes than the one that evaluates (stencil_team_load): the geometry
// of the point (stencil_geo), the 20 node values (per lane: stencil_load),
// and the arithmetic on them (stencil_eval).

// Centre cell (i0, j0, k0) and fractions f; the half-shifted cell's fractions
// e and whether its lower plane is the centre cell's lower plane (a = 0) or
// the one below it (a = -1), per axis.
template <typename TC>
struct StencilGeo {
  int i0, j0, k0;
  int ai, aj, ak;
  TC fx, fy, fz;
  TC ex, ey, ez;
};

template <typename TC>
__device__ inline StencilGeo<TC> stencil_geo(const GridDev& g, double px, double py, double pz) {
  StencilGeo<TC> s;
  const double ix = (px - g.o[0]) * g.inv[0];
  const double iy = (py - g.o[1]) * g.inv[1];
  const double iz = (pz - g.o[2]) * g.inv[2];
  const double fi = floor(ix), fj = floor(iy), fk = floor(iz);
  s.i0 = (int)fi; s.j0 = (int)fj; s.k0 = (int)fk;
  s.fx = (TC)(ix - fi); s.fy = (TC)(iy - fj); s.fz = (TC)(iz - fk);
  const double mi = floor(ix - 0.5), mj = floor(iy - 0.5), mk = floor(iz - 0.5);
  s.ex = (TC)(ix - 0.5 - mi); s.ey = (TC)(iy - 0.5 - mj); s.ez = (TC)(iz - 0.5 - mk);
  s.ai = (int)mi - s.i0; s.aj = (int)mj - s.j0; s.ak = (int)mk - s.k0;  // -1 or 0
  return s;
}

// The 20 nodes: centre cell (c), the x- and y-extra planes (x, y) at z nodes
// (k0, k0 + 1), and the z-extra plane (z) at node k0 - 1 (ak = -1) or
// k0 + 2 (ak = 0).
template <typename V>
struct StencilNodes {
  V c000, c001, c010, c011, c100, c101, c110, c111;
  V x00, x01, x10, x11, y00, y01, y10, y11;
  V z00, z01, z10, z11;
};

// The stencil as 8 (i, j) columns of 3 z nodes kmin .. kmin + 2
// (kmin = k0 + ak): column c = 0..3 the centre cell's (i0 + (c >> 1),
// j0 + (c & 1)), 4..5 the x-extra plane at j0 + (c & 1), 6..7 the y-extra
// plane at i0 + (c & 1).  Offsets of column c's (i, j) through accessor a.
template <typename A, typename TC>
__device__ inline long stencil_col_xy(const A& a, const StencilGeo<TC>& s, int c) {
  const int xe = s.ai < 0 ? s.i0 - 1 : s.i0 + 2, ye = s.aj < 0 ? s.j0 - 1 : s.j0 + 2;
  const int i = c < 4 ? s.i0 + (c >> 1) : (c < 6 ? xe : s.i0 + (c & 1));
  const int j = c < 4 ? s.j0 + (c & 1) : (c < 6 ? s.j0 + (c & 1) : ye);
  return a.ox(i) + a.oy(j);
}

// Column c's nodes kmin .. kmin + 2 (m0, m1, m2) -> the stencil's nodes.
template <typename V, typename TC>
__device__ inline void stencil_put_col(const StencilGeo<TC>& s, int c, const V& m0, const V& m1,
                                       const V& m2, StencilNodes<V>& n) {
  // the z-extra node is m0 (ak = -1) or m2; the centre cell's nodes the others
  const bool low = s.ak < 0;
  const V lo = low ? m1 : m0, hi = low ? m2 : m1, ze = low ? m0 : m2;
  switch (c) {
    case 0: n.c000 = lo; n.c001 = hi; n.z00 = ze; break;
    case 1: n.c010 = lo; n.c011 = hi; n.z01 = ze; break;
    case 2: n.c100 = lo; n.c101 = hi; n.z10 = ze; break;
    case 3: n.c110 = lo; n.c111 = hi; n.z11 = ze; break;
    case 4: n.x00 = lo; n.x01 = hi; break;
    case 5: n.x10 = lo; n.x11 = hi; break;
    case 6: n.y00 = lo; n.y01 = hi; break;
    default: n.y10 = lo; n.y11 = hi; break;
  }
}

// Per-lane loads of the 20 nodes through accessor a (z-pair fetches when the
// accessor asks for them).
template <typename TG, typename V, typename A, typename TC>
__device__ inline void stencil_load(const TG* __restrict__ d, const A& a, const StencilGeo<TC>& s,
                                    StencilNodes<V>& n) {
  const int i0 = s.i0, j0 = s.j0, k0 = s.k0, ai = s.ai, aj = s.aj, ak = s.ak;
  // per-axis offsets of the centre cell's two planes and of the extra plane
  // (index -1 when a = -1, +2 when a = 0)
  const long X0 = a.ox(i0), X1 = a.ox(i0 + 1), Xe = a.ox(ai < 0 ? i0 - 1 : i0 + 2);
  const long Y0 = a.oy(j0), Y1 = a.oy(j0 + 1), Ye = a.oy(aj < 0 ? j0 - 1 : j0 + 2);
  if constexpr (A::kZPairs && std::is_same<TG, float>::value && std::is_same<V, float>::value) {
    // a single FP32 field: every column's nodes kmin .. kmin + 2 from the
    // 16-byte z-row (4 nodes of one brick) that holds kmin, and the next
    // brick's row only where the three cross it (kmin mod 4 >= 2, half the
    // points): 8-16 loads per stencil, and the lanes that need no second row
    // cost the L1 no lookup in the other 8
    const int kmin = k0 + ak, kq = kmin & 3, rb = kmin - kq;
    const bool two = kq >= 2;
    const long Zr0 = a.oz(rb), Zr1 = a.oz(rb + 4);
    const long xy[8] = {X0 + Y0, X0 + Y1, X1 + Y0, X1 + Y1, Xe + Y0, Xe + Y1, X0 + Ye, X1 + Ye};
    float4 r0[8];
    float2 r1[8];   // of the next row only its first two nodes can be needed
#pragma unroll
    for (int c = 0; c < 8; ++c) r0[c] = *(const float4*)(d + xy[c] + Zr0);
    if (two) {
#pragma unroll
      for (int c = 0; c < 8; ++c) r1[c] = *(const float2*)(d + xy[c] + Zr1);
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const float4 u = r0[c];
      const float2 w = two ? r1[c] : make_float2(u.z, u.w);
      // nodes kq, kq + 1, kq + 2 of the 8 nodes (u, w)
      const float m0 = kq == 0 ? u.x : kq == 1 ? u.y : kq == 2 ? u.z : u.w;
      const float m1 = kq == 0 ? u.y : kq == 1 ? u.z : kq == 2 ? u.w : w.x;
      const float m2 = kq == 0 ? u.z : kq == 1 ? u.w : kq == 2 ? w.x : w.y;
      stencil_put_col(s, c, m0, m1, m2, n);
    }
  } else if constexpr (A::kZPairs) {
    // every column's nodes kmin .. kmin + 2 (kmin = k0 + ak) in two aligned
    // pair loads: 16 loads per stencil, the same for every lane
    const int kmin = k0 + ak, e = kmin & ~1;
    const bool odd = (kmin & 1) != 0;
    const long Za = a.oz(e), Zb = a.oz(e + 2);
    const long xy[8] = {X0 + Y0, X0 + Y1, X1 + Y0, X1 + Y1, Xe + Y0, Xe + Y1, X0 + Ye, X1 + Ye};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      V m0, m1, m2;
      zcol3<V, TG>(d, xy[c] + Za, xy[c] + Zb, odd, m0, m1, m2);
      stencil_put_col(s, c, m0, m1, m2, n);
    }
  } else {
    const long Z0 = a.oz(k0), Z1 = a.oz(k0 + 1), Ze = a.oz(ak < 0 ? k0 - 1 : k0 + 2);
    n.c000 = gld<V>(d, X0 + Y0 + Z0); n.c001 = gld<V>(d, X0 + Y0 + Z1);
    n.c010 = gld<V>(d, X0 + Y1 + Z0); n.c011 = gld<V>(d, X0 + Y1 + Z1);
    n.c100 = gld<V>(d, X1 + Y0 + Z0); n.c101 = gld<V>(d, X1 + Y0 + Z1);
    n.c110 = gld<V>(d, X1 + Y1 + Z0); n.c111 = gld<V>(d, X1 + Y1 + Z1);
    n.x00 = gld<V>(d, Xe + Y0 + Z0); n.x01 = gld<V>(d, Xe + Y0 + Z1);
    n.x10 = gld<V>(d, Xe + Y1 + Z0); n.x11 = gld<V>(d, Xe + Y1 + Z1);
    n.y00 = gld<V>(d, X0 + Ye + Z0); n.y01 = gld<V>(d, X0 + Ye + Z1);
    n.y10 = gld<V>(d, X1 + Ye + Z0); n.y11 = gld<V>(d, X1 + Ye + Z1);
    n.z00 = gld<V>(d, X0 + Y0 + Ze); n.z01 = gld<V>(d, X0 + Y1 + Ze);
    n.z10 = gld<V>(d, X1 + Y0 + Ze); n.z11 = gld<V>(d, X1 + Y1 + Ze);
  }
}

// phi and the gradient from the 20 nodes.
template <typename TC, typename V>
__device__ inline void stencil_eval(const GridDev& g, const StencilGeo<TC>& s,
                                    const StencilNodes<V>& n, V& phi, V& gx, V& gy, V& gz) {
  const TC fx = s.fx, fy = s.fy, fz = s.fz, ex = s.ex, ey = s.ey, ez = s.ez;
  const int ai = s.ai, aj = s.aj, ak = s.ak;
  const V c000 = n.c000, c001 = n.c001, c010 = n.c010, c011 = n.c011;
  const V c100 = n.c100, c101 = n.c101, c110 = n.c110, c111 = n.c111;
  const V x00 = n.x00, x01 = n.x01, x10 = n.x10, x11 = n.x11;
  const V y00 = n.y00, y01 = n.y01, y10 = n.y10, y11 = n.y11;
  const V z00 = n.z00, z01 = n.z01, z10 = n.z10, z11 = n.z11;
  const TC one = (TC)1;
  const TC gxw = one - fx, gyw = one - fy, gzw = one - fz;
  phi = c000 * gxw * gyw * gzw + c100 * fx * gyw * gzw + c010 * gxw * fy * gzw +
        c001 * gxw * gyw * fz + c110 * fx * fy * gzw + c101 * fx * gyw * fz +
        c011 * gxw * fy * fz + c111 * fx * fy * fz;
  {
    // x: nodes at (m, m+1, m+2) for the four (y,z) corners of the centre cell
    const V n0_00 = ai < 0 ? x00 : c000, n1_00 = ai < 0 ? c000 : c100, n2_00 = ai < 0 ? c100 : x00;
    const V n0_01 = ai < 0 ? x01 : c001, n1_01 = ai < 0 ? c001 : c101, n2_01 = ai < 0 ? c101 : x01;
    const V n0_10 = ai < 0 ? x10 : c010, n1_10 = ai < 0 ? c010 : c110, n2_10 = ai < 0 ? c110 : x10;
    const V n0_11 = ai < 0 ? x11 : c011, n1_11 = ai < 0 ? c011 : c111, n2_11 = ai < 0 ? c111 : x11;
    const TC e1 = one - ex;
    const V d00 = e1 * (n1_00 - n0_00) + ex * (n2_00 - n1_00);
    const V d01 = e1 * (n1_01 - n0_01) + ex * (n2_01 - n1_01);
    const V d10 = e1 * (n1_10 - n0_10) + ex * (n2_10 - n1_10);
    const V d11 = e1 * (n1_11 - n0_11) + ex * (n2_11 - n1_11);
    gx = (d00 * gyw * gzw + d10 * fy * gzw + d01 * gyw * fz + d11 * fy * fz) * (TC)g.inv[0];
  }
  {
    // along y for the four (x,z) corners; y-extra loads are ordered (x0 z0, x0 z1, x1 z0, x1 z1)
    const V n0_00 = aj < 0 ? y00 : c000, n1_00 = aj < 0 ? c000 : c010, n2_00 = aj < 0 ? c010 : y00;
    const V n0_01 = aj < 0 ? y01 : c001, n1_01 = aj < 0 ? c001 : c011, n2_01 = aj < 0 ? c011 : y01;
    const V n0_10 = aj < 0 ? y10 : c100, n1_10 = aj < 0 ? c100 : c110, n2_10 = aj < 0 ? c110 : y10;
    const V n0_11 = aj < 0 ? y11 : c101, n1_11 = aj < 0 ? c101 : c111, n2_11 = aj < 0 ? c111 : y11;
    const TC e1 = one - ey;
    const V d00 = e1 * (n1_00 - n0_00) + ey * (n2_00 - n1_00);
    const V d01 = e1 * (n1_01 - n0_01) + ey * (n2_01 - n1_01);
    const V d10 = e1 * (n1_10 - n0_10) + ey * (n2_10 - n1_10);
    const V d11 = e1 * (n1_11 - n0_11) + ey * (n2_11 - n1_11);
    gy = (d00 * gxw * gzw + d10 * fx * gzw + d01 * gxw * fz + d11 * fx * fz) * (TC)g.inv[1];
  }
  {
    // along z for the four (x,y) corners; z-extra loads are ordered (x0 y0, x0 y1, x1 y0, x1 y1)
    const V n0_00 = ak < 0 ? z00 : c000, n1_00 = ak < 0 ? c000 : c001, n2_00 = ak < 0 ? c001 : z00;
    const V n0_01 = ak < 0 ? z01 : c010, n1_01 = ak < 0 ? c010 : c011, n2_01 = ak < 0 ? c011 : z01;
    const V n0_10 = ak < 0 ? z10 : c100, n1_10 = ak < 0 ? c100 : c101, n2_10 = ak < 0 ? c101 : z10;
    const V n0_11 = ak < 0 ? z11 : c110, n1_11 = ak < 0 ? c110 : c111, n2_11 = ak < 0 ? c111 : z11;
    const TC e1 = one - ez;
    const V d00 = e1 * (n1_00 - n0_00) + ez * (n2_00 - n1_00);
    const V d01 = e1 * (n1_01 - n0_01) + ez * (n2_01 - n1_01);
    const V d10 = e1 * (n1_10 - n0_10) + ez * (n2_10 - n1_10);
    const V d11 = e1 * (n1_11 - n0_11) + ez * (n2_11 - n1_11);
    gz = (d00 * gxw * gyw + d10 * fx * gyw + d01 * gxw * fy + d11 * fx * fy) * (TC)g.inv[2];
  }
}

// Cooperative stencil loads of one wave (64 lanes, one point each, the
// z-pair accessor).  A lane's 16 loads touch ~6 cache lines, but the 64 lanes
// of a load instruction touch ~53 (every lane a different point), and the
// L1 looks up one line per clock: that lookup rate, not HBM or L2, bounds a
// gather over scattered points.  Here the lanes form 8 teams of 8; in round
// r (0..7) team t loads the stencil of lane 8t + r, lane `rank` of the team
// its column `rank` (two aligned z-pair loads, as stencil_load does), so one
// instruction touches the few lines of 8 points.  Each lane stores its
// column's 3 nodes in the owner's record in LDS (`buf`, kTeamWords values of
// V per wave: node k of column c of lane L at k * 520 + c * 65 + L, which
// keeps both the stores of a round and the owner's reads free of bank
// conflicts), and every lane then reads its own 20 nodes back.  The nodes
// are those stencil_load reads, so stencil_eval gives the same bits.  All 64
// lanes must call it together (lanes without a stencil pass need = false).
constexpr int kTeamWords = 3 * 520;

template <typename TG, typename V, typename A, typename TC>
__device__ inline void stencil_team_load(const TG* __restrict__ d, const A& a,
                                         const StencilGeo<TC>& s, bool need, V* buf,
                                         StencilNodes<V>& n) {
  static_assert(A::kZPairs, "team loads use the z-pair accessor");
  const int lane = __lane_id();
  const int team = lane >> 3, rank = lane & 7;
  const int w0 = (s.i0 & 0xFFFF) | (s.j0 << 16);
  const int w1 = (s.k0 & 0xFFFF) | ((s.ai < 0) << 16) | ((s.aj < 0) << 17) | ((s.ak < 0) << 18) |
                 ((need ? 1 : 0) << 19);
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int owner = (team << 3) + r;
    const int o0 = __shfl(w0, owner), o1 = __shfl(w1, owner);
    if (o1 & (1 << 19)) {
      StencilGeo<TC> t;
      t.i0 = o0 & 0xFFFF; t.j0 = o0 >> 16; t.k0 = o1 & 0xFFFF;
      t.ai = (o1 >> 16) & 1 ? -1 : 0; t.aj = (o1 >> 17) & 1 ? -1 : 0; t.ak = (o1 >> 18) & 1 ? -1 : 0;
      const long xy = stencil_col_xy(a, t, rank);
      const int kmin = t.k0 + t.ak, e = kmin & ~1;
      V m0, m1, m2;
      zcol3<V, TG>(d, xy + a.oz(e), xy + a.oz(e + 2), (kmin & 1) != 0, m0, m1, m2);
      V* rec = buf + rank * 65 + owner;
      rec[0] = m0;
      rec[520] = m1;
      rec[1040] = m2;
    }
  }
  // the stores of the other lanes of this wave are visible after this point
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (need) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const V* rec = buf + c * 65 + lane;
      stencil_put_col(s, c, rec[0], rec[520], rec[1040], n);
    }
  }
  // the records may be overwritten by the next call only after every lane read its own
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename TG, typename TC, typename V, typename A>
__device__ inline void grid_phi_grad_stencil_a(const GridDev& g, const TG* __restrict__ d,
                                               const A& a, double px, double py, double pz,
                                               V& phi, V& gx, V& gy, V& gz) {
  const StencilGeo<TC> s = stencil_geo<TC>(g, px, py, pz);
  StencilNodes<V> n;
  stencil_load<TG, V>(d, a, s, n);
  stencil_eval<TC, V>(g, s, n, phi, gx, gy, gz);
}

template <typename TG, typename TC, typename V, bool ZP = false>
__device__ inline void grid_phi_grad_stencil(const GridDev& g, double px, double py, double pz,
                                             V& phi, V& gx, V& gy, V& gz) {
  grid_phi_grad_stencil_a<TG, TC, V>(g, (const TG*)g.data, global_acc<ZP>(g), px, py, pz, phi, gx,
                                     gy, gz);
}

// Both precisions evaluate the node-difference stencil (20 nodes); the
// reference's seven trilinear interpolations (gpu_batch_engine.py:95-115)
// give the same phi and gradient algebraically wherever the whole stencil is
// inside the grid, which every caller guarantees (the 3-spacing valid box),
// so FP64 results equal theirs to rounding (~1e-16 relative) with 20 node
// loads instead of 56.
template <typename TC, bool ZP = false>
__device__ inline void grid_phi_grad_any(const GridDev& g, double px, double py, double pz,
                                         TC& phi, TC& gx, TC& gy, TC& gz) {
  if (g.dtype == 1)
    grid_phi_grad_stencil<double, TC, TC, ZP>(g, px, py, pz, phi, gx, gy, gz);
  else
    grid_phi_grad_stencil<float, TC, TC, ZP>(g, px, py, pz, phi, gx, gy, gz);
}

// Interleaved potential + Born grid (float2 per node): both fields at once.
template <typename TC, bool ZP = false>
__device__ inline void grid_phi_grad_pair(const GridDev& g, double px, double py, double pz,
                                          Pair<TC>& phi, Pair<TC>& gx, Pair<TC>& gy, Pair<TC>& gz) {
  grid_phi_grad_stencil<float2, TC, Pair<TC>, ZP>(g, px, py, pz, phi, gx, gy, gz);
}


// The same two gathers through an accessor (a grid tile staged in LDS):
// the float2 potential + Born pair, and a single FP32 field.
template <typename TC, typename A>
__device__ inline void grid_phi_grad_pair_a(const GridDev& g, const float2* d, const A& a,
                                            double px, double py, double pz, Pair<TC>& phi,
                                            Pair<TC>& gx, Pair<TC>& gy, Pair<TC>& gz) {
  grid_phi_grad_stencil_a<float2, TC, Pair<TC>>(g, d, a, px, py, pz, phi, gx, gy, gz);
}
template <typename TC, typename A>
__device__ inline void grid_phi_grad_f32_a(const GridDev& g, const float* d, const A& a,
                                           double px, double py, double pz, TC& phi, TC& gx,
                                           TC& gy, TC& gz) {
  grid_phi_grad_stencil_a<float, TC, TC>(g, d, a, px, py, pz, phi, gx, gy, gz);
}

template <typename TC> __device__ inline TC pst_exp(TC x);
template <> __device__ inline double pst_exp<double>(double x) { return exp(x); }
template <> __device__ inline float pst_exp<float>(float x) { return __expf(x); }

// Yukawa monopole + dipole + quadrupole with the reference's radial-only
// gradient (gpu_batch_engine.py:401-465).  Accumulates F += -q dphi/dr r_hat,
// E += q phi for an atom at lab position p (receptor at the origin).
template <typename TC> __device__ inline TC pst_rsqrt(TC x);
template <> __device__ inline double pst_rsqrt<double>(double x) { return rsqrt(x); }
template <> __device__ inline float pst_rsqrt<float>(float x) { return rsqrtf(x); }

// Yukawa monopole + dipole + quadrupole with the reference's radial-only
// gradient (gpu_batch_engine.py:401-465), r clamped to >= 1 A (:404):
//   phi  = e^{-r/l} [V/r + (p.u)/(fpe r^2)(1 + r/l) + (uQu)/(fpe r^3)(1 + r/l + r^2/(3l^2))]
//   dphi = e^{-r/l} [V(-1/r^2 - 1/(rl)) + (p.u)/fpe(-2/r^3 - 2/(r^2 l) - 1/(r l^2))
//          + (uQu)/fpe(-3/r^4 - 3/(r^3 l) - 1/(r^2 l^2) - 1/(3 r l^3))]
// with u = p/r.  The quadrupole derivative keeps the reference's
// -1/(r^2 l^2) term.  One reciprocal square root per atom; every other
// constant is pre-divided on the host.  Accumulates F += -q dphi u, E += q phi.
template <typename TC>
__device__ inline void yukawa_atom(const ForceSys& S, TC px, TC py, TC pz, TC q, TC& fx, TC& fy,
                                   TC& fz, TC& e) {
  const TC r2 = px * px + py * py + pz * pz;
  TC inv, sr;
  if (r2 > (TC)1) {
    inv = pst_rsqrt<TC>(r2);
    sr = r2 * inv;
  } else {
    inv = (TC)1;
    sr = (TC)1;
  }
  const TC il = (TC)S.inv_lam;
  const TC ex = pst_exp<TC>(-sr * il);
  const TC rx = px * inv, ry = py * inv, rz = pz * inv;
  const TC V = (TC)S.V_fac;
  const TC i2 = inv * inv;
  TC phi = V * inv;
  TC dphi = -V * (i2 + inv * il);
  if (S.has_dip) {
    const TC pdr = rx * (TC)S.dip_s[0] + ry * (TC)S.dip_s[1] + rz * (TC)S.dip_s[2];
    phi += pdr * i2 * ((TC)1 + sr * il);
    dphi -= pdr * ((TC)2 * i2 * inv + (TC)2 * i2 * il + inv * (TC)S.inv_lam2);
  }
  if (S.has_quad) {
    const TC qx = (TC)S.quad_s[0] * rx + (TC)S.quad_s[1] * ry + (TC)S.quad_s[2] * rz;
    const TC qy = (TC)S.quad_s[3] * rx + (TC)S.quad_s[4] * ry + (TC)S.quad_s[5] * rz;
    const TC qz = (TC)S.quad_s[6] * rx + (TC)S.quad_s[7] * ry + (TC)S.quad_s[8] * rz;
    const TC rQr = rx * qx + ry * qy + rz * qz;
    phi += rQr * i2 * inv * ((TC)1 + sr * il + sr * sr * (TC)S.c3);
    dphi -= rQr * ((TC)3 * i2 * i2 + (TC)3 * i2 * inv * il + i2 * (TC)S.inv_lam2 + inv * (TC)S.c3l);
  }
  const TC qd = q * dphi * ex;
  fx -= qd * rx;
  fy -= qd * ry;
  fz -= qd * rz;
  e += q * phi * ex;
}
