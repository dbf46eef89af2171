// fdtd_setup.cpp -- host-side set-up of the FDTD library: CPML grading, material coefficients,
// oblique plane-wave projection.  No device calls; compiled with -ffp-contract=off so each
// product is rounded on its own, matching the reference's float/double expression types.
#include <cmath>

#include "fdtd_internal.hpp"

namespace fdtd {

// Polynomially graded CPML, src/simulation.h:730-831.  The reference stores sigma, alpha and kappa
// as float (`real`) and evaluates exp/pow in double; the expressions below keep those types.  E
// profile points sit at integer depths q = 0..P-1 (graded by (P-1-q)/(P-1)); H points at q + 1/2
// for q = 0..P-2.  The high-side slab entries [P, 2P) are the mirror image of the low side.
CpmlProfiles compute_cpml(const CpmlParams &p, float delX, float delT) {
  CpmlProfiles pr;
  const size_t P = (size_t)p.P;
  const float m = p.poly_m, am = p.alpha_m, amax = p.alpha_max, kmax = p.kappa_max;
  pr.sig_max = (float)(p.sig_factor * (m + 1.0) / (std::sqrt(kMu0 / kEps0) * delX));
  const float smax = pr.sig_max;
  pr.bE.assign(2 * P, 0.0f);
  pr.cE.assign(2 * P, 0.0f);
  pr.bH.assign(2 * P, 0.0f);
  pr.cH.assign(2 * P, 0.0f);
  pr.kapE.assign(P, 0.0f);
  pr.kapH.assign(P, 0.0f);
  for (size_t q = 0; q < P; ++q) {
    const double depth = (P - q - 1) / (P - 1.0);
    const float sg = (float)(smax * std::pow(depth, m));
    const float al = (float)(amax * std::pow((q + 1 - 1.0) / (P - 1.0), am));
    const float ka = (float)(1.0 + (kmax - 1.0) * std::pow(depth, m));
    const float b = (float)std::exp(-(sg / ka + al) * delT / kEps0);
    float c;
    if (sg == 0.0 && al == 0.0 && q + 1 == P - 1)
      c = 0.0f;
    else
      c = (float)(sg / ka / (sg + ka * al) * (b - 1.0));
    pr.kapE[q] = ka;
    pr.bE[q] = b;
    pr.cE[q] = c;
    pr.bE[2 * P - 1 - q] = b;  // high side, mirrored
    pr.cE[2 * P - 1 - q] = c;
  }
  for (size_t q = 0; q + 1 < P; ++q) {
    const double depth = (P - q - 1 - 0.5) / (P - 1.0);
    const float sg = (float)(smax * std::pow(depth, m));
    const float al = (float)(amax * std::pow((q + 1 - 0.5) / (P - 1.0), am));
    const float ka = (float)(1.0 + (kmax - 1.0) * std::pow(depth, m));
    const float b = (float)std::exp(-(sg / ka + al) * delT / kEps0);
    const float c = (float)(sg / ka / (sg + ka * al) * (b - 1.0));
    pr.kapH[q] = ka;
    pr.bH[q] = b;
    pr.cH[q] = c;
    pr.bH[2 * P - 2 - q] = b;  // high side, mirrored about the last H point
    pr.cH[2 * P - 2 - q] = c;
  }
  return pr;
}

// 1/kappa along an axis of Q cells (1 outside the slabs; exactly 1 everywhere when kappa_max = 1).
std::vector<float> inv_kappa_E(const CpmlProfiles &pr, int P, size_t Q) {
  std::vector<float> v(Q, 1.0f);
  const size_t PP = (size_t)P;
  for (size_t q = 0; q < Q; ++q) {
    if (q < PP) v[q] = (float)(1.0 / pr.kapE[q]);
    else if (q + PP >= Q) v[q] = (float)(1.0 / pr.kapE[Q - 1 - q]);
  }
  return v;
}

std::vector<float> inv_kappa_H(const CpmlProfiles &pr, int P, size_t Q) {
  std::vector<float> v(Q, 1.0f);
  const size_t PP = (size_t)P;
  for (size_t q = 0; q < Q; ++q) {
    if (q + 1 < PP) v[q] = (float)(1.0 / pr.kapH[q]);
    else if (q + PP >= Q && q + 1 < Q) v[q] = (float)(1.0 / pr.kapH[Q - 2 - q]);
  }
  return v;
}

void material_coeffs(float eps, float sig, float delT, float den, float *Ce, float *Ch) {
  const double twoEpsE = 2 * kEps0 * eps;
  const float sdt = sig * delT;
  *Ce = (float)((twoEpsE - sdt) / (twoEpsE + sdt));
  *Ch = (float)(2 * delT / (twoEpsE + sdt) * den);
}

// Ratio dx/ds for which a 1D line with the grid's dt has the 3D grid's numerical wavenumber along
// (kx, ky, kz) at frequency f0 (Taflove & Hagness, 3rd ed., sec. 5.10).  Bisection on both
// dispersion relations; returns 1 when f0 is not positive or not resolvable on the grid.
double matched_ratio(double kx, double ky, double kz, double dx, double dt, double f0) {
  if (!(f0 > 0)) return 1.0;
  const double A = std::sin(2 * kPi * f0 * dt / 2) / (kC0 * dt);
  auto grid = [&](double k) {
    const double a = std::sin(k * kx * dx / 2), b = std::sin(k * ky * dx / 2), c = std::sin(k * kz * dx / 2);
    return (a * a + b * b + c * c) / (dx * dx) - A * A;
  };
  double lo = 0, hi = kPi / dx;
  if (grid(hi) < 0) return 1.0;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (grid(mid) < 0) lo = mid; else hi = mid;
  }
  const double k = 0.5 * (lo + hi);
  auto line = [&](double ds) { return std::sin(k * ds / 2) / ds - A; };
  double a = dx / 4, b = dx;
  if (line(b) >= 0) return 1.0;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (a + b);
    if (line(mid) > 0) a = mid; else b = mid;
  }
  return dx / (0.5 * (a + b));
}

// Oblique incidence: propagation k = (sin t cos p, sin t sin p, cos t), E along
// cos(psi) theta_hat + sin(psi) phi_hat, H along -(k x e) times the line's H (the line's H equals
// -E/eta for a wave travelling towards +s).  The line coordinate of a Yee point r is
// s = sc + ratio * k.(r - corner) with the corner the first box corner the wave reaches; the
// per-axis tables hold the three terms so a kernel adds them in a fixed order, (x + y) + z.
ObliqueSetup oblique_setup(double th, double ph, double ps, const int lo[3], const int hi[3],
                           const int Q[3], int P, float delX, float delT, float fctr) {
  ObliqueSetup o;
  const double k[3] = {std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)};
  const double ut[3] = {std::cos(th) * std::cos(ph), std::cos(th) * std::sin(ph), -std::sin(th)};
  const double up[3] = {-std::sin(ph), std::cos(ph), 0.0};
  double e[3], h[3];
  for (int u = 0; u < 3; ++u) e[u] = std::cos(ps) * ut[u] + std::sin(ps) * up[u];
  h[0] = k[1] * e[2] - k[2] * e[1];
  h[1] = k[2] * e[0] - k[0] * e[2];
  h[2] = k[0] * e[1] - k[1] * e[0];
  const double r = matched_ratio(std::fabs(k[0]), std::fabs(k[1]), std::fabs(k[2]), (double)delX, (double)delT, (double)fctr);
  o.ratio = r;
  double corner[3], span = 0;
  for (int u = 0; u < 3; ++u) {
    corner[u] = k[u] >= 0 ? (double)lo[u] - 1.0 : (double)hi[u] + 1.0;
    span += std::fabs(k[u]) * ((double)(hi[u] - lo[u]) + 2.0);
  }
  o.src = P + 3;
  const double sc = (double)o.src + 3.0;
  o.N1 = (int)std::ceil(sc + r * span) + 4 + P;
  for (int u = 0; u < 3; ++u) {
    const double baseE = u == 0 ? sc : 0.0, baseH = u == 0 ? sc - 0.5 : 0.0;
    for (int t = 0; t < 2; ++t) {
      o.tabE[u][t].assign((size_t)Q[u], 0.0f);
      o.tabH[u][t].assign((size_t)Q[u], 0.0f);
    }
    for (int q = 0; q < Q[u]; ++q) {
      const double di = (double)q - corner[u], dh = (double)q + 0.5 - corner[u];
      o.tabE[u][0][q] = (float)(baseE + r * k[u] * di);
      o.tabE[u][1][q] = (float)(baseE + r * k[u] * dh);
      o.tabH[u][0][q] = (float)(baseH + r * k[u] * di);
      o.tabH[u][1][q] = (float)(baseH + r * k[u] * dh);
    }
    o.polE[u] = (float)e[u];
    o.polH[u] = (float)(-h[u]);
  }
  o.ds = (float)((double)delX / r);
  o.Ch1 = (float)(delT / kMu0 / o.ds);
  o.Cb1 = (float)(delT / kEps0 / o.ds);
  return o;
}

}  // namespace fdtd
