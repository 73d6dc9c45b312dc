// Two-body Rotne-Prager-Yamakawa grand mobility (Zuk, Wajnryb, Mizerski &
// Szymczak, J. Fluid Mech. 741 R5, 2014) and its FP64 Cholesky noise.
//
// The reference keeps only the translational coupling block
// (pystarc/hydrodynamics/rotne_prager.py:101-123) and reduces it to the scalar
// D_parallel of gpu_batch_simulator.py:1074-1084.  Here both bodies move:
//
//   X = [x_rec, x_lig, theta_rec, theta_lig]  (lab frame, receptor at origin)
//   dX = Df dt M f + sqrt(2 Df dt) G xi,   G G^T = M,   Df = kT / eta,
//   f = [F_rec, F_lig, T_rec, T_lig].
//
// In the frame whose third axis is u = r/|r| the 12x12 M splits into one 4x4
// block A over (t0x, t1x, r0y, r1y), its twin S A S (S = diag(1,1,-1,-1)) over
// (t0y, t1y, r0x, r1x), and two 2x2 blocks over (t0z, t1z) and (r0z, r1z):
// one 4x4 and two 2x2 Cholesky factorisations in registers per trajectory
// instead of a 12x12 one.  Noise index layout: xi[0:4] -> A, xi[4:8] -> twin,
// xi[8:10] -> (t0z, t1z), xi[10:12] -> (r0z, r1z).
#pragma once
#include <hip/hip_runtime.h>
#include "pst_device.hpp"

namespace pst {

struct RpyCoef {
  double m0, m1, p0, p1;        // self: 1/(6 pi a), 1/(8 pi a^3)
  double al, be, ga, de;        // tt_01 = al I + be uu, rr_01 = ga I + de uu
  double x01, x10;              // rt_ij = x_ij eps.u_ij
};

__host__ __device__ inline RpyCoef rpy_coef(double r, double a0, double a1) {
  RpyCoef c;
  const double pi = M_PI;
  c.m0 = 1.0 / (6.0 * pi * a0);
  c.m1 = 1.0 / (6.0 * pi * a1);
  c.p0 = 1.0 / (8.0 * pi * a0 * a0 * a0);
  c.p1 = 1.0 / (8.0 * pi * a1 * a1 * a1);
  const double s = a0 + a1, d = fabs(a0 - a1);
  if (r > s) {
    const double sq = a0 * a0 + a1 * a1, r2 = r * r, r3 = r2 * r;
    c.al = (1.0 + sq / (3.0 * r2)) / (8.0 * pi * r);
    c.be = (1.0 - sq / r2) / (8.0 * pi * r);
    c.ga = -1.0 / (16.0 * pi * r3);
    c.de = 3.0 / (16.0 * pi * r3);
    c.x01 = c.x10 = 1.0 / (8.0 * pi * r2);
  } else if (r > d) {
    const double r2 = r * r, r3 = r2 * r, dd = (a0 - a1) * (a0 - a1);
    const double t = dd + 3.0 * r2, w = dd - r2;
    c.al = (16.0 * r3 * s - t * t) / (32.0 * r3) / (6.0 * pi * a0 * a1);
    c.be = 3.0 * w * w / (32.0 * r3) / (6.0 * pi * a0 * a1);
    const double q = a0 * a0 + 4.0 * a0 * a1 + a1 * a1;
    const double dsq = a0 * a0 - a1 * a1;
    const double A = (5.0 * r3 * r3 - 27.0 * r2 * r2 * (a0 * a0 + a1 * a1) +
                      32.0 * r3 * (a0 * a0 * a0 + a1 * a1 * a1) - 9.0 * r2 * dsq * dsq -
                      dd * dd * q) / (64.0 * r3);
    const double B = 3.0 * w * w * (q - r2) / (64.0 * r3);
    const double v = 8.0 * pi * a0 * a0 * a0 * a1 * a1 * a1;
    c.ga = A / v;
    c.de = B / v;
    const double e01 = a0 - a1 + r, e10 = a1 - a0 + r;
    c.x01 = e01 * e01 * (a1 * a1 + 2.0 * a1 * (a0 + r) - 3.0 * (a0 - r) * (a0 - r)) /
            (128.0 * pi * a0 * a0 * a0 * a1 * r2);
    c.x10 = e10 * e10 * (a0 * a0 + 2.0 * a0 * (a1 + r) - 3.0 * (a1 - r) * (a1 - r)) /
            (128.0 * pi * a1 * a1 * a1 * a0 * r2);
  } else {
    const double big = fmax(a0, a1);
    c.al = 1.0 / (6.0 * pi * big);
    c.be = 0.0;
    c.ga = 1.0 / (8.0 * pi * big * big * big);
    c.de = 0.0;
    c.x01 = a0 > a1 ? r / (8.0 * pi * a0 * a0 * a0) : 0.0;
    c.x10 = a1 > a0 ? r / (8.0 * pi * a1 * a1 * a1) : 0.0;
  }
  return c;
}

// Orthonormal e1, e2 with e1 x e2 = u (Duff et al., JCGT 6(1), 2017).
__device__ inline void rpy_basis(const double* u, double* e1, double* e2) {
  const double sgn = copysign(1.0, u[2]);
  const double a = -1.0 / (sgn + u[2]);
  const double b = u[0] * u[1] * a;
  e1[0] = 1.0 + sgn * u[0] * u[0] * a; e1[1] = sgn * b; e1[2] = -sgn * u[0];
  e2[0] = b; e2[1] = sgn + u[1] * u[1] * a; e2[2] = -u[1];
}

// Lower Cholesky factors: L4 of A (row-major 4x4), LC and LD of the 2x2 blocks.
struct RpyFactor { double L4[4][4]; double LC[3]; double LD[3]; };

__device__ inline RpyFactor rpy_factor(const RpyCoef& c) {
  const double A[4][4] = {{c.m0, c.al, 0.0, -c.x10},
                          {c.al, c.m1, c.x01, 0.0},
                          {0.0, c.x01, c.p0, c.ga},
                          {-c.x10, 0.0, c.ga, c.p1}};
  RpyFactor f;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) f.L4[i][j] = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    double s = A[j][j];
#pragma unroll
    for (int k = 0; k < j; ++k) s -= f.L4[j][k] * f.L4[j][k];
    const double d = sqrt(s);
    f.L4[j][j] = d;
#pragma unroll
    for (int i = j + 1; i < 4; ++i) {
      double t = A[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) t -= f.L4[i][k] * f.L4[j][k];
      f.L4[i][j] = t / d;
    }
  }
  // 2x2 [[a, b], [b, c]] -> [l00, l10, l11]
  const double tc = c.al + c.be, rc = c.ga + c.de;
  f.LC[0] = sqrt(c.m0); f.LC[1] = tc / f.LC[0]; f.LC[2] = sqrt(c.m1 - f.LC[1] * f.LC[1]);
  f.LD[0] = sqrt(c.p0); f.LD[1] = rc / f.LD[0]; f.LD[2] = sqrt(c.p1 - f.LD[1] * f.LD[1]);
  return f;
}

// Aligned-frame noise G_aligned xi -> t0, t1, r0, r1 (aligned components).
__device__ inline void rpy_noise_aligned(const RpyFactor& f, const double* xi, double* t0,
                                         double* t1, double* r0, double* r1) {
  double yA[4], yB[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int k = 0; k <= i; ++k) {
      a += f.L4[i][k] * xi[k];
      b += f.L4[i][k] * (k < 2 ? xi[4 + k] : -xi[4 + k]);
    }
    yA[i] = a;
    yB[i] = i < 2 ? b : -b;
  }
  const double c0 = f.LC[0] * xi[8], c1 = f.LC[1] * xi[8] + f.LC[2] * xi[9];
  const double d0 = f.LD[0] * xi[10], d1 = f.LD[1] * xi[10] + f.LD[2] * xi[11];
  t0[0] = yA[0]; t0[1] = yB[0]; t0[2] = c0;
  t1[0] = yA[1]; t1[1] = yB[1]; t1[2] = c1;
  r0[0] = yB[2]; r0[1] = yA[2]; r0[2] = d0;
  r1[0] = yB[3]; r1[1] = yA[3]; r1[2] = d1;
}

__device__ inline void rpy_to_lab(const double* e1, const double* e2, const double* u,
                                  const double* v, double* out) {
#pragma unroll
  for (int k = 0; k < 3; ++k) out[k] = e1[k] * v[0] + e2[k] * v[1] + u[k] * v[2];
}

__device__ inline void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// Force and torque on the receptor from the pair: Newton's third law; torque
// zero in the reference's zero-torque mode (gpu_batch_engine.py:597-599), else
// the angular-momentum balance about the origin T_rec = -T_lig - pos x F_lig.
__device__ inline void rpy_receptor_load(int torque_mode, const double* pos, const double* FL,
                                         const double* TL, double* FR, double* TR) {
  FR[0] = -FL[0]; FR[1] = -FL[1]; FR[2] = -FL[2];
  if (torque_mode == 0) {
    TR[0] = TR[1] = TR[2] = 0.0;
  } else {
    double c[3];
    cross3(pos, FL, c);
    TR[0] = -TL[0] - c[0]; TR[1] = -TL[1] - c[1]; TR[2] = -TL[2] - c[2];
  }
}

// dX (12, lab frame) for the pose `pos` (ligand centre, receptor at origin).
__device__ inline void rpy_displacement(double a0, double a1, double Df, double dt,
                                        const double* pos, const double* FR, const double* FL,
                                        const double* TR, const double* TL, const double* xi,
                                        double* dX) {
  const double r = norm3(pos[0], pos[1], pos[2]);
  const double u[3] = {pos[0] / r, pos[1] / r, pos[2] / r};
  const RpyCoef c = rpy_coef(r, a0, a1);
  // drift  M f  (tensor forms; eps.u v = v x u)
  const double uFL = u[0] * FL[0] + u[1] * FL[1] + u[2] * FL[2];
  const double uFR = u[0] * FR[0] + u[1] * FR[1] + u[2] * FR[2];
  const double uTL = u[0] * TL[0] + u[1] * TL[1] + u[2] * TL[2];
  const double uTR = u[0] * TR[0] + u[1] * TR[1] + u[2] * TR[2];
  double uxTL[3], TRxu[3], uxFL[3], FRxu[3];
  cross3(u, TL, uxTL);
  cross3(TR, u, TRxu);
  cross3(u, FL, uxFL);
  cross3(FR, u, FRxu);
  const double h = Df * dt;
  double m[12];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    m[k] = c.m0 * FR[k] + c.al * FL[k] + c.be * u[k] * uFL + c.x10 * uxTL[k];
    m[3 + k] = c.al * FR[k] + c.be * u[k] * uFR + c.m1 * FL[k] + c.x01 * TRxu[k];
    m[6 + k] = c.p0 * TR[k] + c.ga * TL[k] + c.de * u[k] * uTL + c.x01 * uxFL[k];
    m[9 + k] = c.ga * TR[k] + c.de * u[k] * uTR + c.p1 * TL[k] + c.x10 * FRxu[k];
  }
  // noise  G xi
  const RpyFactor f = rpy_factor(c);
  double e1[3], e2[3], t0[3], t1[3], r0[3], r1[3], g[12];
  rpy_basis(u, e1, e2);
  rpy_noise_aligned(f, xi, t0, t1, r0, r1);
  rpy_to_lab(e1, e2, u, t0, g);
  rpy_to_lab(e1, e2, u, t1, g + 3);
  rpy_to_lab(e1, e2, u, r0, g + 6);
  rpy_to_lab(e1, e2, u, r1, g + 9);
  const double sg = sqrt(2.0 * Df * dt);
#pragma unroll
  for (int k = 0; k < 12; ++k) dX[k] = h * m[k] + sg * g[k];
}

__device__ inline void rotvec_quat(const double* w, double* q) {
  const double ang = norm3(w[0], w[1], w[2]);
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (ang > 1e-10) { ax = w[0] / ang; ay = w[1] / ang; az = w[2] / ang; }
  const double s = sin(0.5 * ang), c = cos(0.5 * ang);
  q[0] = c; q[1] = ax * s; q[2] = ay * s; q[3] = az * s;
}

__device__ inline void qmul(const double* a, const double* b, double* o) {
  o[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  o[1] = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
  o[2] = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
  o[3] = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
}

// New pose in the turned receptor frame:
//   pos' = R(theta_rec)^T (pos + dx_lig - dx_rec)
//   q'   = conj(dq(theta_rec)) dq(theta_lig) q   (normalised)
__device__ inline void rpy_apply(const double* dX, double* pos, double* q) {
  double dr[4], dl[4], R[9], t[4], o[4];
  rotvec_quat(dX + 6, dr);
  rotvec_quat(dX + 9, dl);
  const double p0 = pos[0] + dX[3] - dX[0];
  const double p1 = pos[1] + dX[4] - dX[1];
  const double p2 = pos[2] + dX[5] - dX[2];
  quat_to_rot(dr, R);
  pos[0] = R[0] * p0 + R[3] * p1 + R[6] * p2;
  pos[1] = R[1] * p0 + R[4] * p1 + R[7] * p2;
  pos[2] = R[2] * p0 + R[5] * p1 + R[8] * p2;
  qmul(dl, q, t);
  const double cr[4] = {dr[0], -dr[1], -dr[2], -dr[3]};
  qmul(cr, t, o);
  const double n = sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2] + o[3] * o[3]);
  q[0] = o[0] / n; q[1] = o[1] / n; q[2] = o[2] / n; q[3] = o[3] / n;
}

}  // namespace pst
