// LMZ outer propagator on the device: one trajectory per thread is carried
// through the outer region b < r < q (q = 20 x the larger molecular radius)
// until it reaches the b-sphere or leaves through the q-sphere, where it
// returns with the LMZ probability rate(b)/rate(q)
// (pystarc/simulation/outer_propagator.py:259-401).  Near either sphere the
// Lamm-Schulten step (step_near_surface.py:31-130) decides absorption; in
// between, radial drift + diffusion steps with the step-size rule of
// outer_propagator.py:340-377.  On return the ligand orientation takes the
// rotational diffusion of both bodies over the elapsed time
// (diffusional_rotation.py:87-126; for 0.25 < D_r t < 4 the rotation angle
// is drawn by inverting the SO(3) heat-kernel CDF).  Random numbers: one
// Philox call per draw keyed (trajectory, step, RNG_OUTER, draw counter).
#pragma once
#include "pst_device.hpp"

namespace pst {

struct OuterDraws {
  uint64_t seed, key;
  uint32_t step, c;
  __device__ double u() {
    double a, b;
    philox_uniform2(seed, key, step, RNG_OUTER, c++, a, b);
    return a;
  }
  __device__ double n() {
    double a, b;
    philox_normal2(seed, key, step, RNG_OUTER, c++, a, b);
    return a;
  }
};

constexpr uint32_t kOuterMaxDraws = 16000000u;  // draw counters stay below 2^24

__device__ inline double outer_D(const StepParams& P, double r) {
  return P.lmz_hi ? P.lmz_dpre * (P.lmz_ainv - 3.0 / r + 2.0 * P.lmz_a2 / (r * r * r)) : P.lmz_Dt;
}

__device__ inline double outer_force(const StepParams& P, double r) {
  return P.lmz_V * exp(-r / P.lmz_L) / r * (1.0 / r + 1.0 / P.lmz_L);
}

// Lamm-Schulten step at distance x0 from an absorbing wall; returns whether
// the particle survives, its new distance and the elapsed time.
__device__ inline bool outer_step_near(OuterDraws& dr, double x0, double F, double D,
                                       double* new_x, double* time) {
  const double bb = -F;
  const double tau = x0 * x0 / D;
  const double st = sqrt(tau), st2 = 2.0 * st, bt = bb * tau;
  const double erfmt = erf((x0 - bt) / st2), erfpt = erf((x0 + bt) / st2);
  const double psurv =
      fmin(1.0, fmax(0.0, 0.5 * (exp(bb * x0) * (erfpt - 1.0) + erfmt + 1.0)));
  if (dr.u() < psurv) {
    const double E = erfmt;
    double x = 0.0;
    bool found = false;
    for (int a = 0; a < 10000 && !found; ++a) {
      const double iearg = fmin(1.0 - 1e-12, fmax(-1.0 + 1e-12, dr.u() * (E + 1.0) - E));
      x = 2.0 * st * erfinv(iearg) - bt + x0;
      if (x < 0.0) continue;
      const double t4 = 4.0 * tau;
      const double p0 = exp(-((x - x0 + bt) * (x - x0 + bt)) / t4);
      const double p1 = exp(bb * x0 - ((x + x0 + bt) * (x + x0 + bt)) / t4);
      const double p2 = -2.0 * exp(-((x + bt) * (x + bt) + x0 * (x0 + 2.0 * (x - bt))) / t4);
      const double p = p0 + p1 + p2;
      const double den = 0.5 * (erfmt + 1.0);
      const double pu = den > 1e-30 ? p0 / den : 0.0;
      if (p < 0 || pu <= 0) continue;
      found = dr.u() < p / pu;
    }
    if (!found) x = fmax(x0, 0.001);
    *new_x = fmax(0.0, x);
    *time = tau / D;
    return true;
  }
  const double x02 = x0 * x0, b2 = bb * bb;
  const double tmax = fabs(bb * x0) < 0.5 ? x02 * (1.0 / 6.0 - x02 * b2 / 216.0)
                                          : (sqrt(b2 * x02 + 9.0) - 3.0) / b2;
  auto pt = [&](double t) {
    return t <= 0 ? 0.0
                  : x0 * exp(-((x0 - bb * t) * (x0 - bb * t)) / (4.0 * t)) /
                        (2.0 * sqrt(M_PI * t * t * t));
  };
  double pmax = pt(tmax);
  if (pmax <= 0) pmax = 1e-30;
  double ts = 0.0;
  for (int a = 0; a < 10000; ++a) {
    ts = tau * dr.u();
    if (pmax * dr.u() < pt(ts)) break;
  }
  *new_x = 0.0;
  *time = ts / D;
  return false;
}

// CDF of the rotation angle after rotational diffusion for tau = D_r t: the
// SO(3) heat kernel sum_l (2l+1) e^{-l(l+1) tau} chi_l(theta) on the class
// measure (1 - cos theta)/pi, integrated term by term.
__device__ inline double outer_angle_cdf(double th, double tau, int lmax) {
  double s = th - sin(th);
  for (int l = 1; l <= lmax; ++l)
    s += (2 * l + 1) * exp(-(double)l * (l + 1) * tau) * (sin(l * th) / l - sin((l + 1) * th) / (l + 1));
  return s / M_PI;
}

__device__ inline void outer_rotation(OuterDraws& dr, double tau, double* q) {
  q[0] = 1.0; q[1] = q[2] = q[3] = 0.0;
  if (tau <= 0.0) return;
  if (tau <= 0.25) {
    const double s = sqrt(2.0 * tau);
    const double wx = s * dr.n(), wy = s * dr.n(), wz = s * dr.n();
    const double a = sqrt(wx * wx + wy * wy + wz * wz);
    if (a < 1e-12) return;
    const double h = sin(0.5 * a) / a;
    q[0] = cos(0.5 * a); q[1] = wx * h; q[2] = wy * h; q[3] = wz * h;
    return;
  }
  if (tau < 4.0) {
    const double u = dr.u();
    const int lmax = min(64, (int)ceil(sqrt(40.0 / tau)) + 2);
    double lo = 0.0, hi = M_PI;
    for (int i = 0; i < 55; ++i) {
      const double m = 0.5 * (lo + hi);
      if (outer_angle_cdf(m, tau, lmax) < u) lo = m; else hi = m;
    }
    const double th = 0.5 * (lo + hi);
    const double ax = dr.n(), ay = dr.n(), az = dr.n();
    const double nrm = sqrt(ax * ax + ay * ay + az * az);
    if (nrm < 1e-14) return;
    const double h = sin(0.5 * th) / nrm;
    q[0] = cos(0.5 * th); q[1] = ax * h; q[2] = ay * h; q[3] = az * h;
    return;
  }
  const double a = dr.n(), b = dr.n(), c = dr.n(), d = dr.n();
  const double nrm = sqrt(a * a + b * b + c * c + d * d);
  q[0] = a / nrm; q[1] = b / nrm; q[2] = c / nrm; q[3] = d / nrm;
}

__device__ inline void outer_qmul(const double* a, const double* b, double* o) {
  o[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  o[1] = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
  o[2] = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
  o[3] = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
}

// One excursion from pos (|pos| >= lmz_trigger).  Returns true when the
// trajectory is back on the b-sphere (pos, q updated), false when it escaped.
__device__ inline bool outer_propagate(const StepParams& P, uint64_t key, uint32_t step,
                                       double* pos, double* q) {
  OuterDraws dr{P.seed, key, step, 0u};
  const double LARGE = 1.0e30;
  double t = 0.0;
  bool back = false, out = false;
  while (!(back || out)) {
    if (dr.c > kOuterMaxDraws) { out = true; break; }
    const double r = norm3(pos[0], pos[1], pos[2]);
    if (r <= P.lmz_b) {
      const double s = P.lmz_b / r;
      pos[0] *= s; pos[1] *= s; pos[2] *= s;
      back = true;
    } else if (r >= P.lmz_q) {
      if (dr.u() < P.lmz_return_q) {
        const double s = P.lmz_b / r;
        pos[0] *= s; pos[1] *= s; pos[2] *= s;
        back = true;
      } else {
        out = true;
      }
      t = LARGE;
    } else if (r < P.lmz_b_cover) {
      double nx, dt;
      const bool surv = outer_step_near(dr, r - P.lmz_b, outer_force(P, r), outer_D(P, r), &nx, &dt);
      back = !surv;
      t += dt;
      const double s = (P.lmz_b + nx) / r;
      pos[0] *= s; pos[1] *= s; pos[2] *= s;
    } else if (r > P.lmz_q_cover) {
      double nx, dt;
      const bool surv = outer_step_near(dr, P.lmz_q - r, -outer_force(P, r), outer_D(P, r), &nx, &dt);
      if (!surv) {
        if (dr.u() < P.lmz_return_q) {
          const double s = P.lmz_b / r;
          pos[0] *= s; pos[1] *= s; pos[2] *= s;
          back = true;
        } else {
          out = true;
        }
        t = LARGE;
      } else {
        const double s = (P.lmz_q - nx) / r;
        pos[0] *= s; pos[1] *= s; pos[2] *= s;
      }
    } else {
      const double Fr0 = outer_force(P, r);
      const double L = P.lmz_L;
      const double D0 = outer_D(P, r);
      const double g = 1.0 / r + 1.0 / L;
      const double Fr1 = -(P.lmz_V * exp(-r / L) / r) * g * g - 2.0 * Fr0 / r;
      const double alpha = 0.01;
      const double Dfac = 6.0 * M_PI * P.lmz_dpre;
      double dtf;
      if (P.lmz_hi) {
        const double rm1 = 1.0 / r;
        const double D1 = -3.0 * D0 * rm1 - Dfac * rm1 * rm1 / M_PI;
        const double D2 = -4.0 * D1 * rm1 + Dfac * rm1 * rm1 * rm1 / M_PI;
        const double D3 = -5.0 * D2 * rm1 - 2.0 * Dfac * rm1 * rm1 * rm1 * rm1 / M_PI;
        if (fabs(Fr0) > 0 && r < 3.0 * L) {
          const double num = D1 + D0 * Fr0;
          const double den = D0 * D3 + (D1 + 2 * Fr0 * D0) * D2 + Fr0 * D1 * D1 +
                             (3 * Fr1 + Fr0 * Fr0) * D0 * D1 + Fr0 * Fr1 * D0 * D0;
          dtf = fabs(den) > 0 ? alpha * fabs(num / den) : LARGE;
        } else {
          const double den = D0 * D3 + D1 * D2;
          dtf = fabs(den) > 0 ? alpha * fabs(D1 / den) : LARGE;
        }
      } else {
        dtf = (fabs(Fr0) > 0 && r < 3.0 * L) ? (fabs(Fr1) > 0 ? alpha / fabs(D0 * Fr1) : LARGE)
                                              : LARGE;
      }
      const double edge = fmin(P.lmz_q - r, r - P.lmz_b);
      const double dt = fmin(edge * edge / (18.0 * D0), dtf);
      t += dt;
      const double ux = pos[0] / r, uy = pos[1] / r, uz = pos[2] / r;
      const double drift = D0 * Fr0 * dt;
      pos[0] += drift * ux; pos[1] += drift * uy; pos[2] += drift * uz;
      const double sr = sqrt(2.0 * D0 * dt);
      if (P.lmz_hi) {
        const double rm1 = 1.0 / r;
        const double Dp = Dfac * (P.lmz_ainv / (6.0 * M_PI) -
                                  2.0 * (rm1 - 2.0 * P.lmz_a2 * rm1 * rm1 * rm1) / (8.0 * M_PI));
        const double sp = sqrt(2.0 * Dp * dt);
        const double rho = sqrt(ux * ux + uy * uy);
        double tx, ty, tz, px, py, pz;
        if (rho == 0.0) {
          tx = 1.0; ty = 0.0; tz = 0.0; px = 0.0; py = 1.0; pz = 0.0;
        } else {
          tx = uz * ux / rho; ty = uz * uy / rho; tz = -rho;
          px = -uy / rho; py = ux / rho; pz = 0.0;
        }
        const double a = sr * dr.n();
        pos[0] += a * ux; pos[1] += a * uy; pos[2] += a * uz;
        const double b = sp * dr.n();
        pos[0] += b * tx; pos[1] += b * ty; pos[2] += b * tz;
        const double c = sp * dr.n();
        pos[0] += c * px; pos[1] += c * py; pos[2] += c * pz;
      } else {
        pos[0] += sr * dr.n(); pos[1] += sr * dr.n(); pos[2] += sr * dr.n();
      }
    }
  }
  if (back) {
    double dq0[4], dq1[4], a[4], b[4];
    outer_rotation(dr, t * P.lmz_drot0, dq0);
    outer_rotation(dr, t * P.lmz_drot1, dq1);
    outer_qmul(dq1, q, a);
    const double c0[4] = {dq0[0], -dq0[1], -dq0[2], -dq0[3]};
    outer_qmul(c0, a, b);
    const double n = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2] + b[3] * b[3]);
    q[0] = b[0] / n; q[1] = b[1] / n; q[2] = b[2] / n; q[3] = b[3] / n;
  }
  return back;
}

}  // namespace pst
