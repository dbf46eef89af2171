// fdtd_internal.hpp -- shared definitions of the MI355X FDTD library (host set-up, kernels, C ABI).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace fdtd {

// Physical constants as src/simulation.h:12-21 (pi = the header's fallback value, used when
// <math.h> does not define M_PI).
constexpr double kPi = 3.1415926535897;
constexpr double kC0 = 299792458;
constexpr double kMu0 = 4 * kPi * 1e-7;
constexpr double kEps0 = 1 / kC0 / kC0 / kMu0;

enum { EX = 0, EY, EZ, HX, HY, HZ };
// psi array ids, reference order (src/simulation.h:1318-1332)
enum { P_EXY = 0, P_EXZ, P_EYX, P_EYZ, P_EZX, P_EZY, P_HXY, P_HXZ, P_HYX, P_HYZ, P_HZX, P_HZY };

struct CpmlParams {
  int P = 10;
  float poly_m = 3.0f, kappa_max = 1.0f, alpha_max = 0.08f, alpha_m = 1.0f;
  double sig_factor = 0.8;
};

// CPML profile set computed on the host (src/simulation.h:730-831).
struct CpmlProfiles {
  float sig_max = 0;
  std::vector<float> bE, cE, bH, cH;  // [2P]
  std::vector<float> kapE, kapH;      // [P]
};
CpmlProfiles compute_cpml(const CpmlParams &p, float delX, float delT);
std::vector<float> inv_kappa_E(const CpmlProfiles &pr, int P, size_t Q);
std::vector<float> inv_kappa_H(const CpmlProfiles &pr, int P, size_t Q);

// Coefficients of one material (src/simulation.h:924-938): Ce = (2 eps0 eps - sig dt)/(2 eps0 eps
// + sig dt); Ch = 2 dt/(2 eps0 eps + sig dt) * den with den = 1/dx (0 on the last x plane).
void material_coeffs(float eps, float sig, float delT, float den, float *Ce, float *Ch);

// Oblique plane-wave set-up (extension): projection tables and polarisation of the 1D line.
struct ObliqueSetup {
  double ratio = 1.0;  // dx / ds
  int N1 = 0, src = 0;
  float ds = 0, Ch1 = 0, Cb1 = 0;
  std::vector<float> tabE[3][2], tabH[3][2];  // [axis][0 integer | 1 half position][q]
  float polE[3] = {0, 0, 0}, polH[3] = {0, 0, 0};
};
double matched_ratio(double kx, double ky, double kz, double dx, double dt, double f0);
ObliqueSetup oblique_setup(double theta, double phi, double psi, const int lo[3], const int hi[3],
                           const int Q[3], int P, float delX, float delT, float fctr);

// Device-side parameter block of the TF/SF corrections (generic face formulation).
struct PWDev {
  int active_h = 0, active_e = 0;  // bit (3*face_axis + component)
  int N1 = 0;
  const float *Eh = nullptr, *Hh = nullptr;  // history rows [T][N1]
  const float *tabE[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
  const float *tabH[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
  float polE[3] = {0, 0, 0}, polH[3] = {0, 0, 0};
};

}  // namespace fdtd
