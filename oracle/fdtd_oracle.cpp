/*
 * fdtd_oracle.cpp -- CPU restatement of the enricoballini/maxwell_solver 3D Yee FDTD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: it may be built, loaded and
 * called only by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and only
 * to check results against (or as the timed CPU baseline), never as a product code path.
 *
 * The reference has no CPU solver (src/cpucode/FunLib.cpp only generates a waveform), so this is
 * a restatement of the reference's device arithmetic (src/gpu.h) and host setup
 * (src/simulation.h), templated on the real type R.  With R = float it follows the reference's
 * float/double mixing literally (the reference declares its scalars `real` = float and evaluates
 * constants such as eps0 in double); with R = double it is "the reference with real = double".
 * Build with -ffp-contract=off so that no multiply-add is fused.
 *
 * Every function cites the reference file:line it restates.  Extensions that the reference lacks
 * (runtime CPML parameters with per-axis 1/kappa, a TF/SF inset, oblique and layered plane waves,
 * probes, snapshots, a graded mesh) reduce to the reference arithmetic when set to the reference preset
 * (P=10, m=3, kappa_max=1, alpha_max=0.08, alpha_m=1, inset 15, vacuum 1D line).
 *
 * Parity pins (checked in tests/test_oracle_*.py): closed-form CPML profiles and coefficients,
 * a vectorised numpy statement of whole steps on a tiny non-cubic grid, discrete divergence of H, Yee energy in a
 * lossless closed cavity, SF-region leakage of the axis-aligned TF/SF (exact up to rounding), 1D
 * line vs the textbook recurrence, CPML reflection, float vs double convergence, Fresnel
 * reflection of a lossy half-space.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include <omp.h>

namespace orc {

// src/simulation.h:12-21.  pi is the fallback value simulation.h uses when <math.h> does not
// define M_PI (MSVC without _USE_MATH_DEFINES, the reference's evident build host).
constexpr double kPi = 3.1415926535897;
constexpr double kC0 = 299792458;
constexpr double kMu0 = 4 * kPi * 1e-7;
constexpr double kEps0 = 1 / kC0 / kC0 / kMu0;

enum { EX = 0, EY, EZ, HX, HY, HZ };
// psi arrays, names as in src/simulation.h:1318-1332
enum { P_EXY = 0, P_EXZ, P_EYX, P_EYZ, P_EZX, P_EZY, P_HXY, P_HXZ, P_HYX, P_HYZ, P_HZX, P_HZY };

static thread_local std::string g_err;

struct Base {
  virtual ~Base() {}
};

template <typename R>
struct Sim : Base {
  // ---- state flags (src/simulation.h:1305-1311) ----
  bool init_const = false, init_source = false, init_antenna = false, init_ready = false;
  bool have_1d = false;

  // ---- setConst (src/simulation.h:79-131) ----
  R fctr = 0, delX = 0, delT = 0;
  size_t T = 0, I = 0, J = 0, K = 0, nAnt = 0, IJ = 0, IJK = 0;
  // CPML parameters (src/simulation.h:85-90; runtime here)
  size_t P = 10;
  R poly_m = 3, alpha_m = 1.0, alpha_max = 0.08, kappa_max = 1.0;
  double sig_factor = 0.8;
  R sig_max = 0;
  size_t width = 15;  // TF/SF inset (src/simulation.h:102)
  size_t i0 = 0, i1 = 0, j0 = 0, j1 = 0, k0 = 0, k1 = 0;

  std::vector<R> eps[3], sig[3];  // eps_inf{x,y,z}, sigma_s{x,y,z}

  // ---- readySimulation (src/simulation.h:640-979) ----
  std::vector<R> bE, cE, bH, cH;          // [2P], shared by all axes and the 1D line
  std::vector<R> kapE, kapH;              // kappa profiles [P] (profile index)
  std::vector<R> ikE[3], ikH[3];          // per-axis 1/kappa along x, y, z (times dx / spacing on a graded mesh)
  // graded mesh (extension): cell sizes along x, y, z as given (empty: uniform dx along that axis)
  std::vector<double> mesh[3];
  std::vector<R> Ce[3], Ch[3];            // Cexe/Ceye/Ceze, Cexh/Ceyh/Cezh per cell
  std::vector<R> den_ex;                  // [I] (den_ex depends on i only)
  R Chh = 1, Che = 0;
  // per-cell mu (createMu, src/simulation.h:28-68): relative permeabilities as given, and the
  // per-cell H coefficients Chxe/Chye/Chze of K2 (empty until both mu and den_hx exist)
  std::vector<float> mu_in[3];
  std::vector<R> Chmu[3];
  std::vector<R> F[6];                    // Ex Ey Ez Hx Hy Hz
  std::vector<R> psi[12];

  // ---- source / probes ----
  std::vector<R> source;
  std::vector<size_t> coax_ind;           // src/simulation.h:1047-1067
  std::vector<int> probe_comps = {EZ};    // recorded components (0..5)
  std::vector<R> port;                    // [ncomp][T][nAnt]
  size_t rec_step = 0;

  // ---- 1D incident line (src/simulation.h:134-254) ----
  size_t N1 = 0;
  std::vector<R> E1h, H1h, Einvh, Hinvh;  // histories [T*N1]
  // plane-wave extension
  int pw_mode = 0;            // 0: reference (forward + inverse vacuum lines); 1..4 with layered; 5 oblique
  int pw_layered = 0;
  double pw_theta = 0, pw_phi = 0, pw_psi = 0;
  double pw_ratio = 1;        // dx / ds of the oblique line
  size_t pw_src = 0;
  std::vector<R> tabE_i[3], tabE_h[3], tabH_i[3], tabH_h[3];  // oblique projection tables
  R polE[3] = {0, 0, 0}, polH[3] = {0, 0, 0};

  // =====================================================================================
  // Graded (tensor-product) mesh, the build's use of the reference's non-uniform-mesh
  // scaffolding (FDTD_UNIFORM_MESH and den_ey/den_ez/den_hy/den_hz, src/simulation.h:2, 679-685,
  // 718-724, allocated and zeroed but never read there): d[a][n] is the distance between the
  // grid planes n and n+1 along axis a (the last entry, a CPML cell, equals dx).  The H update differences
  // E across the primal cell n (spacing d[n]) and the E update differences H across the dual
  // cell n (spacing (d[n-1] + d[n]) / 2); both enter as factors dx / spacing on the per-axis
  // tables (the reference's den_h* = 1/(dx kappa_H), den_e* = 1/(dx kappa_E), :861-912), so
  // Che and Ch keep their 1/dx.
  int set_mesh(const float *const d[3]) {
    if (!init_const) { g_err = "simFDTD:InvalidInput|SetMesh before SetConst"; return 2; }
    if (init_ready) { g_err = "simFDTD:InvalidInput|SetMesh after ReadySimulation"; return 2; }
    const size_t Q[3] = {I, J, K};
    for (int a = 0; a < 3; ++a) {
      mesh[a].clear();
      if (!d[a]) continue;
      for (size_t n = 0; n < Q[a]; ++n) {
        if (!std::isfinite(d[a][n]) || !(d[a][n] > 0)) { g_err = "simFDTD:InvalidInput|SetMesh: spacings must be positive"; return 1; }
        mesh[a].push_back((double)d[a][n]);
      }
    }
    return 0;
  }
  bool graded() const { return !mesh[0].empty() || !mesh[1].empty() || !mesh[2].empty(); }
  // The CPML layers keep the uniform spacing dx (cells n <= P and n >= Q-P-2), so that the psi
  // updates, which difference over dx, and every CPML table are those of the uniform grid.
  int grade_tables() {
    const size_t Q[3] = {I, J, K};
    const double dx = (double)delX;
    // Courant limit of the finest cells: dt <= 1 / (c sqrt(sum_a 1 / min_n d[a][n]^2))
    double inv2 = 0;
    for (int a = 0; a < 3; ++a) {
      double m = dx;
      for (double v : mesh[a]) m = std::min(m, v);
      inv2 += 1.0 / (m * m);
    }
    if ((double)delT > 1.0 / (kC0 * std::sqrt(inv2))) {
      g_err = "simFDTD:InvalidInput|SetMesh: dt exceeds the Courant limit of the finest cells";
      return 1;
    }
    for (int a = 0; a < 3; ++a) {
      const std::vector<double> &d = mesh[a];
      if (d.empty()) continue;
      for (size_t n = 0; n < Q[a]; ++n)
        if ((n <= P || n + P + 2 >= Q[a]) && d[n] != dx) {
          g_err = "simFDTD:InvalidInput|SetMesh: the CPML cells (P+1 at each end) must keep the spacing dx";
          return 1;
        }
      for (size_t n = 0; n + 1 < Q[a]; ++n) ikH[a][n] = (R)((double)ikH[a][n] * (dx / d[n]));
      for (size_t n = 1; n < Q[a]; ++n) ikE[a][n] = (R)((double)ikE[a][n] * (dx / (0.5 * (d[n - 1] + d[n]))));
    }
    return 0;
  }

  int set_const(const float *freqRes, const int32_t *timeDim) {
    if (init_const) { g_err = "simFDTD:AlreadyInitialized|SetConst: already called."; return 3; }
    fctr = freqRes[0];
    delX = freqRes[1];
    delT = freqRes[2];
    poly_m = 3;
    alpha_m = 1.0;
    alpha_max = 0.08;
    kappa_max = 1.0;
    if (timeDim[0] < 1 || timeDim[1] < 1 || timeDim[2] < 1 || timeDim[3] < 1 || timeDim[4] < 0) {
      g_err = "simFDTD:InvalidInput|SetConst: bad TimeDim";
      return 1;
    }
    T = (size_t)timeDim[0];
    I = (size_t)timeDim[1];
    J = (size_t)timeDim[2];
    K = (size_t)timeDim[3];
    nAnt = (size_t)timeDim[4];
    IJ = I * J;
    IJK = I * J * K;
    for (int c = 0; c < 3; ++c) {
      eps[c].assign(IJK, (R)1.0);
      sig[c].assign(IJK, (R)0.0);
    }
    init_const = true;
    return 0;
  }

  int set_cpml(const double *p) {
    if (!init_const || init_ready) { g_err = "simFDTD:InvalidInput|SetCPML: call between SetConst and ReadySimulation"; return 2; }
    if (p[0] < 2) { g_err = "simFDTD:InvalidInput|SetCPML: P >= 2 required"; return 1; }
    P = (size_t)p[0];
    poly_m = (R)p[1];
    kappa_max = (R)p[2];
    alpha_max = (R)p[3];
    alpha_m = (R)p[4];
    sig_factor = p[5];
    return 0;
  }

  void box() {
    // src/simulation.h:102-108 (width generalised)
    i0 = width; i1 = I - width - 1;
    j0 = width; j1 = J - width - 1;
    k0 = width; k1 = K - width - 1;
  }

  // src/simulation.h:288-328
  int update_material(int flag, const float *v) {
    if (!init_const) { g_err = "simFDTD:InvalidInput|UpdateMaterial before SetConst"; return 2; }
    std::vector<R> *dst = nullptr;
    if (flag >= 1 && flag <= 3) dst = &sig[flag - 1];
    else if (flag >= 4 && flag <= 6) dst = &eps[flag - 4];
    else return 0;  // the reference silently ignores other flags
    for (size_t i = 0; i < IJK; ++i) (*dst)[i] = v[i];
    return 0;
  }

  // CreateMu, src/simulation.h:28-68, with its initialisation order fixed: the reference sets
  // init_mu only when ReadySimulation has not run (:30-32) but reads den_hx, which ReadySimulation
  // fills (:861-885), so neither order works there.  Here mu may be given any time after SetConst;
  // the coefficients are computed once den_hx exists (now, or at ReadySimulation).
  int create_mu(const float *mx, const float *my, const float *mz) {
    if (!init_const) { g_err = "simFDTD:InvalidInput|CreateMu before SetConst"; return 2; }
    const float *m[3] = {mx, my, mz};
    for (int c = 0; c < 3; ++c) mu_in[c].assign(m[c], m[c] + IJK);
    if (init_ready) mu_coefficients();
    return 0;
  }

  // src/simulation.h:43-61: C = delT / mu0 / mu / (1 + sigma_m delT / 2 / mu0 / mu) * den_hx with
  // the reference's operand types (delT, mu, sigma_m, den_hx float; mu0 double), stored as float.
  // den_hx (:861-885) is 1/dx on i < I-1 and 0 on the last x plane; its kappa factor is applied
  // per axis through ikH as for den_ex (equal at kappa_max = 1).
  void mu_coefficients() {
    const R sigma_m = 0;  // :914
    for (int c = 0; c < 3; ++c) {
      Chmu[c].assign(IJK, 0);
      for (size_t id = 0; id < IJK; ++id) {
        const R mu = (R)mu_in[c][id];
        const R den = den_ex[id % I];
        Chmu[c][id] = (R)(delT / kMu0 / mu / (1 + sigma_m * delT / 2 / kMu0 / mu) * den);
      }
    }
  }

  // src/simulation.h:1068-1095
  int set_source(const float *s, int update) {
    if (!init_const) { g_err = "simFDTD:InvalidInput|SetSource before SetConst"; return 2; }
    if (update) {
      if (!init_source) return 0;
    } else if (init_source) {
      return 0;
    }
    source.assign(T, 0);
    for (size_t i = 0; i < T; ++i) source[i] = s[i];
    init_source = true;
    return 0;
  }

  // src/simulation.h:1047-1067 (cell = (P + x, P + y, P + z)); resolved against the current P
  std::vector<int32_t> ant_loc;
  int set_antenna(const int32_t *loc) {
    if (!init_const) { g_err = "simFDTD:InvalidInput|SetAntennaLocations before SetConst"; return 2; }
    ant_loc.assign(loc, loc + 3 * nAnt);
    init_antenna = true;
    return resolve_antenna();
  }
  int resolve_antenna() {
    coax_ind.assign(nAnt, 0);
    for (size_t a = 0; a < nAnt; ++a) {
      long x = (long)P + ant_loc[a], y = (long)P + ant_loc[a + nAnt], z = (long)P + ant_loc[a + 2 * nAnt];
      if (x < 0 || y < 0 || z < 0 || x >= (long)I || y >= (long)J || z >= (long)K) {
        g_err = "simFDTD:InvalidInput|antenna outside the grid";
        init_antenna = false;
        return 1;
      }
      coax_ind[a] = (size_t)x + (size_t)y * I + (size_t)z * IJ;
    }
    port.assign(probe_comps.size() * T * nAnt, 0);
    return 0;
  }

  int set_probe_comps(const int32_t *c, int n) {
    if (n < 1 || n > 6) { g_err = "simFDTD:InvalidInput|1..6 probe components"; return 1; }
    probe_comps.clear();
    for (int q = 0; q < n; ++q) {
      if (c[q] < 1 || c[q] > 6) { g_err = "simFDTD:InvalidInput|probe component must be 1..6"; return 1; }
      probe_comps.push_back(c[q] - 1);
    }
    if (init_antenna) port.assign(probe_comps.size() * T * nAnt, 0);
    return 0;
  }

  // ---- CPML profiles, src/simulation.h:730-831 ----
  void cpml_profiles() {
    sig_max = sig_factor * (poly_m + 1.0) / (sqrt(kMu0 / kEps0) * delX);  // :87
    std::vector<R> sigE(P, 0), alE(P, 0), sigH(P, 0), alH(P, 0);
    kapE.assign(P, 0);
    kapH.assign(P, 0);
    bE.assign(2 * P, 0); cE.assign(2 * P, 0); bH.assign(2 * P, 0); cH.assign(2 * P, 0);
    std::vector<R> bE2(P, 0), cE2(P, 0), bH2(P, 0), cH2(P, 0);
    for (size_t i = 0; i < P; i++) {  // :731-747 (E, low side) and :769-789 (E, high side)
      sigE[i] = sig_max * pow(((P - i - 1) / (P - 1.0)), poly_m);
      alE[i] = alpha_max * pow(((i + 1 - 1.0) / (P - 1.0)), alpha_m);
      kapE[i] = 1.0 + (kappa_max - 1.0) * pow(((P - i - 1) / (P - 1.0)), poly_m);
      bE[i] = exp(-(sigE[i] / kapE[i] + alE[i]) * delT / kEps0);
      if ((sigE[i] == 0.0) && (alE[i] == 0.0) && ((i + 1) == (P - 1)))
        cE[i] = 0.0;
      else
        cE[i] = sigE[i] / kapE[i] / (sigE[i] + kapE[i] * alE[i]) * (bE[i] - 1.0);
      bE2[P - i - 1] = bE[i];
      cE2[P - i - 1] = cE[i];
    }
    for (size_t i = 0; i < P - 1; i++) {  // :750-765 and :792-806
      sigH[i] = sig_max * pow(((P - i - 1 - 0.5) / (P - 1.0)), poly_m);
      alH[i] = alpha_max * pow(((i + 1 - 0.5) / (P - 1.0)), alpha_m);
      kapH[i] = 1.0 + (kappa_max - 1.0) * pow(((P - i - 1 - 0.5) / (P - 1.0)), poly_m);
      bH[i] = exp(-(sigH[i] / kapH[i] + alH[i]) * delT / kEps0);
      cH[i] = sigH[i] / kapH[i] / (sigH[i] + kapH[i] * alH[i]) * (bH[i] - 1.0);
      bH2[P - i - 1] = bH[i];
      cH2[P - i - 1] = cH[i];
    }
    for (size_t i = 0; i < P; i++) {  // :807-811
      bE[P + i] = bE2[i];
      cE[P + i] = cE2[i];
    }
    for (size_t i = 0; i < P - 1; i++) {  // :813-817
      bH[P + i] = bH2[i + 1];
      cH[P + i] = cH2[i + 1];
    }
  }

  // 1/kappa along an axis of length Q.  The reference applies kappa only through den_ex (x only,
  // src/simulation.h:887-912); this restatement applies 1/kappa per axis (standard CPML).  With
  // kappa_max = 1 every entry is exactly 1 and the arithmetic is the reference's.
  std::vector<R> inv_kappa_E(size_t Q) const {
    std::vector<R> v(Q, (R)1.0);
    for (size_t q = 0; q < Q; ++q) {
      if (q < P) v[q] = (R)(1.0 / kapE[q]);
      else if (q + P >= Q) v[q] = (R)(1.0 / kapE[Q - 1 - q]);
    }
    return v;
  }
  std::vector<R> inv_kappa_H(size_t Q) const {
    std::vector<R> v(Q, (R)1.0);
    for (size_t q = 0; q < Q; ++q) {
      if (q + 1 < P) v[q] = (R)(1.0 / kapH[q]);
      else if (q + P >= Q && q + 1 < Q) v[q] = (R)(1.0 / kapH[Q - 2 - q]);
    }
    return v;
  }

  // src/simulation.h:924-938 and :1099-1113 (identical formulas)
  void coefficients() {
    for (int c = 0; c < 3; ++c) {
      Ce[c].assign(IJK, 0);
      Ch[c].assign(IJK, 0);
    }
    for (size_t id = 0; id < IJK; id++) {
      R den = den_ex[id % I];
      for (int c = 0; c < 3; ++c) {
        Ce[c][id] = (2 * kEps0 * eps[c][id] - sig[c][id] * delT) / (2 * kEps0 * eps[c][id] + sig[c][id] * delT);
        Ch[c][id] = 2 * delT / (2 * kEps0 * eps[c][id] + sig[c][id] * delT) * den;
      }
    }
  }

  int ready() {
    if (!init_const) { g_err = "simFDTD:InvalidInput|ReadySimulation before SetConst"; return 2; }
    if (init_ready) { g_err = "simFDTD:AlreadyInitialized|ReadySimulation: already called."; return 3; }
    if (I < 2 * P + 2 || J < 2 * P + 2 || K < 2 * P + 2) { g_err = "simFDTD:InvalidInput|grid smaller than 2P+2"; return 1; }
    if (2 * width + 3 > I || 2 * width + 3 > J || 2 * width + 3 > K || width < 1) {
      g_err = "simFDTD:InvalidInput|TF/SF inset too large"; return 1;
    }
    box();
    cpml_profiles();
    // den_ex, src/simulation.h:887-912 with kappa moved into ik*: 1/dx for i < I-1, 0 at I-1
    den_ex.assign(I, 0);
    for (size_t i = 0; i + 1 < I; i++) den_ex[i] = (R)(1.0 / (delX * (R)1.0));
    ikE[0] = inv_kappa_E(I); ikE[1] = inv_kappa_E(J); ikE[2] = inv_kappa_E(K);
    ikH[0] = inv_kappa_H(I); ikH[1] = inv_kappa_H(J); ikH[2] = inv_kappa_H(K);
    if (graded())
      if (int rc = grade_tables()) return rc;
    // src/simulation.h:914-922 (sigma_m = 0)
    R sigma_m = 0;
    Chh = (1 - sigma_m * delT / 2 / kMu0) / (1 + sigma_m * delT / 2 / kMu0);
    Che = delT / kMu0 / (1 + sigma_m * delT / 2 / kMu0) / delX;
    coefficients();
    if (!mu_in[0].empty()) mu_coefficients();
    // initField, src/simulation.h:980-1045
    for (int c = 0; c < 6; ++c) F[c].assign(IJK, 0);
    size_t sx = 2 * P * J * K, sy = 2 * I * P * K, sz = 2 * I * J * P;
    size_t sizes[12] = {sy, sz, sx, sz, sx, sy, sy, sz, sx, sz, sx, sy};
    for (int q = 0; q < 12; ++q) psi[q].assign(sizes[q], 0);
    init_ready = true;
    if (init_antenna && resolve_antenna()) return 1;
    return 0;
  }

  int update_coeffs() {  // UpdateCuda, src/simulation.h:1097-1127
    if (!init_ready) { g_err = "simFDTD:InvalidInput|UpdateCuda before ReadySimulation"; return 2; }
    coefficients();
    return 0;
  }

  int reset() {  // src/simulation.h:1128-1151
    if (!init_ready) { g_err = "simFDTD:InvalidInput|ResetFieldAll before ReadySimulation"; return 2; }
    for (int c = 0; c < 6; ++c) std::fill(F[c].begin(), F[c].end(), (R)0);
    for (int q = 0; q < 12; ++q) std::fill(psi[q].begin(), psi[q].end(), (R)0);
    std::fill(port.begin(), port.end(), (R)0);
    rec_step = 0;
    return 0;
  }

  // =====================================================================================
  // 1D incident line.  src/simulation.h:134-254 with src/gpu.h:40-53, 84-149.
  struct Line {
    size_t N;
    R Ch, ds;
    std::vector<R> Ce, Cb, ikE1, ikH1;
  };

  void run_line(const Line &L, size_t src, const std::vector<R> &pw, std::vector<R> &Eh, std::vector<R> &Hh) {
    const size_t N = L.N;
    std::vector<R> E(N, 0), H(N, 0), pH(2 * P, 0), pE(2 * P, 0);
    Eh.assign(T * N, 0);
    Hh.assign(T * N, 0);
    for (size_t n = 0; n < T; n++) {
      size_t time = n * N;
      for (size_t i = 0; i < N; ++i) Eh[time + i] = E[i];           // copy1d  gpu.h:40-46
      for (size_t i = 0; i + 1 < N; ++i)                            // GPU_updateH_1D gpu.h:134-140
        H[i] = H[i] + L.Ch * ((E[i + 1] - E[i]) * L.ikH1[i]);
      for (size_t t = 0; t < 2 * P; ++t) {                          // GPU_updateHPml_1D gpu.h:84-107
        bool c1 = t < P, c2 = (t < 2 * P - 1) && (t >= P);
        if (!(c1 || c2)) continue;
        size_t idx = c1 ? t : t + N - 2 * P;
        pH[t] = bH[t] * pH[t] + cH[t] * (E[idx + 1] - E[idx]) / L.ds;
        H[idx] = H[idx] + L.Ch * pH[t] * L.ds;
      }
      for (size_t i = 1; i < N; ++i)                                // GPU_updateE_1D gpu.h:142-149
        E[i] = L.Ce[i] * E[i] + L.Cb[i] * ((H[i] - H[i - 1]) * L.ikE1[i]);
      for (size_t t = 0; t < 2 * P; ++t) {                          // GPU_updateEPml_1D gpu.h:109-132
        bool c1 = (t > 0) && (t < P), c2 = (t < 2 * P) && (t >= P);
        if (!(c1 || c2)) continue;
        size_t idx = c1 ? t : t + N - 2 * P;
        pE[t] = bE[t] * pE[t] + cE[t] * (H[idx] - H[idx - 1]) / L.ds;
        E[idx] = E[idx] + L.Cb[idx] * pE[t] * L.ds;
      }
      E[src] = E[src] + pw[n];                                      // AddSource_GPU_1D gpu.h:48-53
      for (size_t i = 0; i < N; ++i) Hh[time + i] = H[i];
    }
  }

  Line vacuum_line(size_t N, R ds, R Ch1, R Cb1) const {
    Line L;
    L.N = N; L.ds = ds; L.Ch = Ch1;
    L.Ce.assign(N, (R)1.0);
    L.Cb.assign(N, Cb1);
    L.ikE1 = inv_kappa_E(N);
    L.ikH1 = inv_kappa_H(N);
    return L;
  }

  int set_plane_wave(int mode, const double *prm) {
    if (!init_const) { g_err = "simFDTD:InvalidInput|SetPlaneWave before SetConst"; return 2; }
    if (mode < 1 || mode > 7) { g_err = "simFDTD:InvalidInput|plane-wave mode must be 1..7"; return 1; }
    pw_mode = mode;
    if (mode == 5) {
      pw_theta = prm[0]; pw_phi = prm[1]; pw_psi = prm[2]; pw_layered = 0;
    } else {
      pw_layered = prm[0] != 0.0;
    }
    return 0;
  }

  // Dispersion-matched ds for an oblique line: the 1D grid with the same dt has the 3D grid's
  // numerical wavenumber along k at omega = 2 pi fctr (Taflove & Hagness 3rd ed. sec. 5.10).
  static double matched_ratio(double kx, double ky, double kz, double dx, double dt, double f0) {
    if (!(f0 > 0)) return 1.0;
    double A = sin(2 * kPi * f0 * dt / 2) / (kC0 * dt);
    auto F = [&](double k) {
      double sx = sin(k * kx * dx / 2), sy = sin(k * ky * dx / 2), sz = sin(k * kz * dx / 2);
      return (sx * sx + sy * sy + sz * sz) / (dx * dx) - A * A;
    };
    double lo = 0, hi = kPi / dx;
    if (F(hi) < 0) return 1.0;
    for (int it = 0; it < 200; ++it) {
      double mid = 0.5 * (lo + hi);
      if (F(mid) < 0) lo = mid; else hi = mid;
    }
    double k = 0.5 * (lo + hi);
    auto G = [&](double ds) { return sin(k * ds / 2) / ds - A; };
    double a = dx / 4, b = dx;
    if (G(b) >= 0) return 1.0;
    for (int it = 0; it < 200; ++it) {
      double mid = 0.5 * (a + b);
      if (G(mid) > 0) a = mid; else b = mid;
    }
    return dx / (0.5 * (a + b));
  }

  // src/simulation.h:134-254 (modes 1-4, vacuum), generalised: layered (modes 1-4) and oblique (5)
  int init_1d(float rate, const float *pwsrc) {
    if (!init_ready) { g_err = "simFDTD:InvalidInput|Init1D before ReadySimulation"; return 2; }
    std::vector<R> pw(T);
    R source_rate = rate;
    for (size_t i = 0; i < T; i++) pw[i] = pwsrc[i] * source_rate;  // :188-192
    if (pw_mode == 5) return init_oblique(pw);
    if (pw_layered) return init_layered(pw);
    N1 = std::max(std::max(I, J), K);                                   // :137
    R Ceh_1D = delT / kEps0 / delX;                                      // :144
    Line L = vacuum_line(N1, delX, Che, Ceh_1D);
    run_line(L, N1 - 13, pw, E1h, H1h);                                  // :193-215
    run_line(L, 13, pw, Einvh, Hinvh);                                   // :216-244
    have_1d = true;
    return 0;
  }

  // Layered line for the axis modes: exact restriction of the grid's centre column along the
  // propagation axis (length Q, so index q maps to q), polarised-component coefficients per node.
  int init_layered(const std::vector<R> &pw) {
    int m = pw_mode;
    int ax = (m <= 2) ? 2 : (m <= 4) ? 0 : 1;     // modes 1,2 along z; 3,4 along x; 6,7 along y
    int pol = (m == 1) ? EY : (m == 2) ? EX : (m == 3) ? EZ : (m == 4) ? EY : (m == 6) ? EX : EZ;
    size_t Q = ax == 2 ? K : ax == 0 ? I : J;
    size_t ic = I / 2, jc = J / 2, kc = K / 2;
    N1 = Q;
    Line L;
    L.N = Q; L.ds = delX; L.Ch = Che;
    L.Ce.assign(Q, 1); L.Cb.assign(Q, 0);
    for (size_t q = 0; q < Q; ++q) {
      size_t id = ax == 2 ? (ic + jc * I + q * IJ) : ax == 0 ? (q + jc * I + kc * IJ) : (ic + q * I + kc * IJ);
      L.Ce[q] = Ce[pol][id];
      L.Cb[q] = Ch[pol][id];
    }
    L.ikE1 = inv_kappa_E(Q);
    L.ikH1 = inv_kappa_H(Q);
    if (m == 4) run_line(L, 13, pw, Einvh, Hinvh);
    else run_line(L, Q - 13, pw, E1h, H1h);
    have_1d = true;
    return 0;
  }

  int init_oblique(const std::vector<R> &pw) {
    double th = pw_theta, ph = pw_phi, ps = pw_psi;
    double k[3] = {sin(th) * cos(ph), sin(th) * sin(ph), cos(th)};
    double ut[3] = {cos(th) * cos(ph), cos(th) * sin(ph), -sin(th)};
    double up[3] = {-sin(ph), cos(ph), 0.0};
    double e[3], h[3];
    for (int u = 0; u < 3; ++u) e[u] = cos(ps) * ut[u] + sin(ps) * up[u];
    h[0] = k[1] * e[2] - k[2] * e[1];
    h[1] = k[2] * e[0] - k[0] * e[2];
    h[2] = k[0] * e[1] - k[1] * e[0];
    double r = matched_ratio(fabs(k[0]), fabs(k[1]), fabs(k[2]), (double)delX, (double)delT, (double)fctr);
    pw_ratio = r;
    size_t lo[3] = {i0, j0, k0}, hi[3] = {i1, j1, k1}, Q[3] = {I, J, K};
    double corner[3], span = 0;
    for (int u = 0; u < 3; ++u) {
      corner[u] = k[u] >= 0 ? (double)lo[u] - 1.0 : (double)hi[u] + 1.0;
      span += fabs(k[u]) * ((double)(hi[u] - lo[u]) + 2.0);
    }
    pw_src = P + 3;
    double sc = (double)pw_src + 3.0;
    N1 = (size_t)ceil(sc + r * span) + 4 + P;
    for (int u = 0; u < 3; ++u) {
      double base = u == 0 ? sc : 0.0, baseH = u == 0 ? sc - 0.5 : 0.0;
      tabE_i[u].assign(Q[u], 0); tabE_h[u].assign(Q[u], 0);
      tabH_i[u].assign(Q[u], 0); tabH_h[u].assign(Q[u], 0);
      for (size_t q = 0; q < Q[u]; ++q) {
        tabE_i[u][q] = (R)(base + r * k[u] * ((double)q - corner[u]));
        tabE_h[u][q] = (R)(base + r * k[u] * ((double)q + 0.5 - corner[u]));
        tabH_i[u][q] = (R)(baseH + r * k[u] * ((double)q - corner[u]));
        tabH_h[u][q] = (R)(baseH + r * k[u] * ((double)q + 0.5 - corner[u]));
      }
      polE[u] = (R)e[u];
      polH[u] = (R)(-h[u]);
    }
    R ds = (R)((double)delX / r);
    R Ch1 = delT / kMu0 / ds;
    R Cb1 = delT / kEps0 / ds;
    Line L = vacuum_line(N1, ds, Ch1, Cb1);
    run_line(L, pw_src, pw, E1h, H1h);
    have_1d = true;
    return 0;
  }

  // =====================================================================================
  // K1 GPU_updateH, src/gpu.h:575-602 (Chxh = Chh, Chxe = Che; 1/kappa factors are 1 at the preset)
  void update_h_all() {
    R *Ex = F[EX].data(), *Ey = F[EY].data(), *Ez = F[EZ].data();
    R *Hx = F[HX].data(), *Hy = F[HY].data(), *Hz = F[HZ].data();
    const R Chxh = Chh, Chxe = Che;
    const R *ikx = ikH[0].data(), *iky = ikH[1].data(), *ikz = ikH[2].data();
    if (!Chmu[0].empty()) {
      // K2 GPU_updateH_Full, src/gpu.h:4-30: per-cell Chxe/Chye/Chze, no Chxh factor
      const R *Cx = Chmu[0].data(), *Cy = Chmu[1].data(), *Cz = Chmu[2].data();
#pragma omp parallel for schedule(static)
      for (long kk = 0; kk < (long)K - 1; ++kk) {
        size_t k = (size_t)kk;
        for (size_t j = 0; j + 1 < J; ++j)
          for (size_t i = 0; i + 1 < I; ++i) {
            size_t id = i + j * I + k * IJ;
            R ez = Ez[id], ey = Ey[id], ex = Ex[id];
            Hx[id] = Hx[id] + Cx[id] * (-(Ez[id + I] - ez) * iky[j] + (Ey[id + IJ] - ey) * ikz[k]);
            Hy[id] = Hy[id] + Cy[id] * (-(Ex[id + IJ] - ex) * ikz[k] + (Ez[id + 1] - ez) * ikx[i]);
            Hz[id] = Hz[id] + Cz[id] * (-(Ey[id + 1] - ey) * ikx[i] + (Ex[id + I] - ex) * iky[j]);
          }
      }
      return;
    }
#pragma omp parallel for schedule(static)
    for (long kk = 0; kk < (long)K - 1; ++kk) {
      size_t k = (size_t)kk;
      for (size_t j = 0; j + 1 < J; ++j)
        for (size_t i = 0; i + 1 < I; ++i) {
          size_t id = i + j * I + k * IJ;
          R ez = Ez[id], ey = Ey[id], ex = Ex[id];
          Hx[id] = Chxh * Hx[id] + Chxe * (-(Ez[id + I] - ez) * iky[j] + (Ey[id + IJ] - ey) * ikz[k]);
          Hy[id] = Chxh * Hy[id] + Chxe * (-(Ex[id + IJ] - ex) * ikz[k] + (Ez[id + 1] - ez) * ikx[i]);
          Hz[id] = Chxh * Hz[id] + Chxe * (-(Ey[id + 1] - ey) * ikx[i] + (Ex[id + I] - ex) * iky[j]);
        }
    }
  }

  // K3 GPU_updateE, src/gpu.h:727-763
  void update_e_all() {
    R *Ex = F[EX].data(), *Ey = F[EY].data(), *Ez = F[EZ].data();
    R *Hx = F[HX].data(), *Hy = F[HY].data(), *Hz = F[HZ].data();
    const R *Cexe = Ce[0].data(), *Ceye = Ce[1].data(), *Ceze = Ce[2].data();
    const R *Cexh = Ch[0].data(), *Ceyh = Ch[1].data(), *Cezh = Ch[2].data();
    const R *ikx = ikE[0].data(), *iky = ikE[1].data(), *ikz = ikE[2].data();
#pragma omp parallel for schedule(static)
    for (long kk = 0; kk < (long)K; ++kk) {
      size_t k = (size_t)kk;
      for (size_t j = 0; j < J; ++j)
        for (size_t i = 0; i < I; ++i) {
          size_t id = i + j * I + k * IJ;
          R hx = Hx[id], hy = Hy[id], hz = Hz[id];
          if (j != 0 && k != 0)
            Ex[id] = Cexe[id] * Ex[id] + Cexh[id] * ((hz - Hz[id - I]) * iky[j] - (hy - Hy[id - IJ]) * ikz[k]);
          if (i != 0 && k != 0)
            Ey[id] = Ceye[id] * Ey[id] + Ceyh[id] * ((hx - Hx[id - IJ]) * ikz[k] - (hz - Hz[id - 1]) * ikx[i]);
          if (i != 0 && j != 0)
            Ez[id] = Ceze[id] * Ez[id] + Cezh[id] * ((hy - Hy[id - 1]) * ikx[i] - (hx - Hx[id - I]) * iky[j]);
        }
    }
  }

  // updateH_PML, src/simulation.h:392-403: K5 (Y), K4 (X), K6 (Z) in this order
  void update_h_pml() {
    R *Ex = F[EX].data(), *Ey = F[EY].data(), *Ez = F[EZ].data();
    R *Hx = F[HX].data(), *Hy = F[HY].data(), *Hz = F[HZ].data();
    const R dx = delX;
    const long PP = (long)P;
    // K5 GPU_updateHPmlY, src/gpu.h:644-685
    {
      R *pxy = psi[P_HXY].data(), *pzy = psi[P_HZY].data();
#pragma omp parallel for schedule(static)
      for (long kk = 0; kk < (long)K - 1; ++kk)
        for (long s = 0; s < 2 * PP - 1; ++s)
          for (size_t i = 0; i + 1 < I; ++i) {
            size_t k = (size_t)kk;
            size_t t = i + I * (size_t)s + I * 2 * P * k;
            size_t jj = s < PP ? (size_t)s : (size_t)s + J - 2 * P;
            size_t id = i + jj * I + k * IJ;
            pxy[t] = bH[s] * pxy[t] + cH[s] * (Ez[id + I] - Ez[id]) / dx;
            Hx[id] = Hx[id] - Che * pxy[t] * dx;
            pzy[t] = bH[s] * pzy[t] + cH[s] * (Ex[id + I] - Ex[id]) / dx;
            Hz[id] = Hz[id] + Che * pzy[t] * dx;
          }
    }
    // K4 GPU_updateHPmlX, src/gpu.h:604-642
    {
      R *pyx = psi[P_HYX].data(), *pzx = psi[P_HZX].data();
#pragma omp parallel for schedule(static)
      for (long kk = 0; kk < (long)K - 1; ++kk)
        for (size_t j = 0; j + 1 < J; ++j)
          for (long s = 0; s < 2 * PP - 1; ++s) {
            size_t k = (size_t)kk;
            size_t t = (size_t)s + 2 * P * j + 2 * P * J * k;
            size_t ii = s < PP ? (size_t)s : I - 2 * P + (size_t)s;
            size_t id = ii + j * I + k * IJ;
            pyx[t] = bH[s] * pyx[t] + cH[s] * (Ez[id + 1] - Ez[id]) / dx;
            Hy[id] = Hy[id] + Che * pyx[t] * dx;
            pzx[t] = bH[s] * pzx[t] + cH[s] * (Ey[id + 1] - Ey[id]) / dx;
            Hz[id] = Hz[id] - Che * pzx[t] * dx;
          }
    }
    // K6 GPU_updateHPmlZ, src/gpu.h:687-725
    {
      R *pxz = psi[P_HXZ].data(), *pyz = psi[P_HYZ].data();
#pragma omp parallel for schedule(static)
      for (long s = 0; s < 2 * PP - 1; ++s)
        for (size_t j = 0; j + 1 < J; ++j)
          for (size_t i = 0; i + 1 < I; ++i) {
            size_t t = i + I * j + IJ * (size_t)s;
            size_t kk = s < PP ? (size_t)s : (size_t)s + K - 2 * P;
            size_t id = i + j * I + kk * IJ;
            pxz[t] = bH[s] * pxz[t] + cH[s] * (Ey[id + IJ] - Ey[id]) / dx;
            Hx[id] = Hx[id] + Che * pxz[t] * dx;
            pyz[t] = bH[s] * pyz[t] + cH[s] * (Ex[id + IJ] - Ex[id]) / dx;
            Hy[id] = Hy[id] - Che * pyz[t] * dx;
          }
    }
  }

  // updateE_PML, src/simulation.h:404-418: K8 (Y), K9 (Z), K7 (X) in this order
  void update_e_pml() {
    R *Ex = F[EX].data(), *Ey = F[EY].data(), *Ez = F[EZ].data();
    R *Hx = F[HX].data(), *Hy = F[HY].data(), *Hz = F[HZ].data();
    const R *Cexh = Ch[0].data(), *Ceyh = Ch[1].data(), *Cezh = Ch[2].data();
    const R dx = delX;
    const long PP = (long)P;
    // K8 GPU_updateEPmlY, src/gpu.h:790-827
    {
      R *pxy = psi[P_EXY].data(), *pzy = psi[P_EZY].data();
#pragma omp parallel for schedule(static)
      for (long kk = 1; kk < (long)K - 1; ++kk)
        for (long s = 1; s < 2 * PP; ++s)
          for (size_t i = 1; i + 1 < I; ++i) {
            size_t k = (size_t)kk;
            size_t t = i + I * (size_t)s + I * 2 * P * k;
            size_t jj = s < PP ? (size_t)s : (size_t)s + J - 2 * P;
            size_t id = i + jj * I + k * IJ;
            pxy[t] = bE[s] * pxy[t] + cE[s] * (Hz[id] - Hz[id - I]) / dx;
            Ex[id] = Ex[id] + Cexh[id] * pxy[t] * dx;
            pzy[t] = bE[s] * pzy[t] + cE[s] * (Hx[id] - Hx[id - I]) / dx;
            Ez[id] = Ez[id] - Cezh[id] * pzy[t] * dx;
          }
    }
    // K9 GPU_updateEPmlZ, src/gpu.h:829-868
    {
      R *pxz = psi[P_EXZ].data(), *pyz = psi[P_EYZ].data();
#pragma omp parallel for schedule(static)
      for (long s = 1; s < 2 * PP; ++s)
        for (size_t j = 1; j + 1 < J; ++j)
          for (size_t i = 1; i + 1 < I; ++i) {
            size_t t = i + I * j + IJ * (size_t)s;
            size_t kk = s < PP ? (size_t)s : (size_t)s + K - 2 * P;
            size_t id = i + j * I + kk * IJ;
            pxz[t] = bE[s] * pxz[t] + cE[s] * (Hy[id] - Hy[id - IJ]) / dx;
            Ex[id] = Ex[id] - Cexh[id] * pxz[t] * dx;
            pyz[t] = bE[s] * pyz[t] + cE[s] * (Hx[id] - Hx[id - IJ]) / dx;
            Ey[id] = Ey[id] + Ceyh[id] * pyz[t] * dx;
          }
    }
    // K7 GPU_updateEPmlX, src/gpu.h:765-788
    {
      R *pyx = psi[P_EYX].data(), *pzx = psi[P_EZX].data();
#pragma omp parallel for schedule(static)
      for (long kk = 1; kk < (long)K - 1; ++kk)
        for (size_t j = 1; j + 1 < J; ++j)
          for (long s = 1; s < 2 * PP; ++s) {
            size_t k = (size_t)kk;
            size_t t = (size_t)s + 2 * P * j + 2 * P * J * k;
            size_t ii = s < PP ? (size_t)s : (size_t)s + I - 2 * P;
            size_t id = ii + j * I + k * IJ;
            pyx[t] = bE[s] * pyx[t] + cE[s] * (Hz[id] - Hz[id - 1]) / dx;
            Ey[id] = Ey[id] - Ceyh[id] * pyx[t] * dx;
            pzx[t] = bE[s] * pzx[t] + cE[s] * (Hy[id] - Hy[id - 1]) / dx;
            Ez[id] = Ez[id] + Cezh[id] * pzx[t] * dx;
          }
    }
    record();
  }

  void record() {
    if (init_antenna && rec_step < T) {
      for (size_t q = 0; q < probe_comps.size(); ++q)
        for (size_t a = 0; a < nAnt; ++a)
          port[(q * T + rec_step) * nAnt + a] = F[probe_comps[q]][coax_ind[a]];
    }
    rec_step++;
  }

  // =====================================================================================
  // Plane wave, src/simulation.h:330-369 and :421-570 with the kernels of src/gpu.h:152-573.
  size_t IDX(size_t i, size_t j, size_t k) const { return i + j * I + k * IJ; }

  int update_h_pw(int n, int mode) {
    if (mode < 1 || mode > 7) return 0;
    if (graded()) { g_err = "simFDTD:InvalidInput|a graded mesh takes no plane wave (the 1D incident line is uniform)"; return 1; }
    if (!have_1d) { g_err = "simFDTD:InvalidInput|plane wave before Init1D"; return 2; }
    if (n < 0 || (size_t)n >= T) { g_err = "simFDTD:InvalidInput|time index outside [0,T)"; return 1; }
    if (mode == 5) return oblique_h((size_t)n);
    const bool inv = (mode == 4);
    const std::vector<R> &Eh = inv ? Einvh : E1h;
    if (Eh.empty()) { g_err = "simFDTD:InvalidInput|the 1D line for this mode was not initialised"; return 2; }
    R *Hx = F[HX].data(), *Hy = F[HY].data(), *Hz = F[HZ].data();
    size_t nstep = (size_t)n * N1;
    const R c = 1, Chxe = Che, Chye = Che, Chze = Che;
    if (mode == 1) {  // updateH_PW_ZY: K30 gpu.h:424-449, K31 gpu.h:451-477
      const R *E = Eh.data() + nstep + N1 - K;
      for (size_t y = j0; y <= j1 - 1; ++y)
        for (size_t x = i0; x <= i1; ++x) {
          size_t a = IDX(x, y, k0 - 1), b = IDX(x, y, k1);
          Hx[a] = Hx[a] - Chxe * E[k0] * c;
          Hx[b] = Hx[b] + Chxe * E[k1] * c;
        }
      for (size_t z = k0; z <= k1; ++z)
        for (size_t y = j0; y <= j1 - 1; ++y) {
          size_t a = IDX(i0 - 1, y, z), b = IDX(i1, y, z);
          Hz[a] = Hz[a] + Chze * E[z] * c;
          Hz[b] = Hz[b] - Chze * E[z] * c;
        }
    } else if (mode == 2) {  // updateH_PW_ZX: K27 gpu.h:528-542, K26 gpu.h:512-527
      const R *E = Eh.data() + nstep + N1 - K;
      for (size_t y = j0; y <= j1; ++y)
        for (size_t x = i0; x <= i1 - 1; ++x) {
          size_t a = IDX(x, y, k0 - 1), b = IDX(x, y, k1);
          Hy[a] = Hy[a] + Chye * E[k0] * c;
          Hy[b] = Hy[b] - Chye * E[k1] * c;
        }
      for (size_t z = k0; z <= k1; ++z)
        for (size_t x = i0; x <= i1 - 1; ++x) {
          size_t a = IDX(x, j0 - 1, z), b = IDX(x, j1, z);
          Hz[a] = Hz[a] - Chze * E[z] * c;
          Hz[b] = Hz[b] + Chze * E[z] * c;
        }
    } else if (mode == 3) {  // updateH_PW_XZ: K18 gpu.h:152-166, K19 gpu.h:168-182
      const R *E = Eh.data() + nstep + N1 - I;
      for (size_t z = k0; z <= k1 - 1; ++z)
        for (size_t x = i0; x <= i1; ++x) {
          size_t a = IDX(x, j0 - 1, z), b = IDX(x, j1, z);
          Hx[a] = Hx[a] + Chxe * E[x] * c;
          Hx[b] = Hx[b] - Chxe * E[x] * c;
        }
      for (size_t z = k0; z <= k1 - 1; ++z)
        for (size_t y = j0; y <= j1; ++y) {
          size_t a = IDX(i0 - 1, y, z), b = IDX(i1, y, z);
          Hy[a] = Hy[a] - Chye * E[i0] * c;
          Hy[b] = Hy[b] + Chye * E[i1] * c;
        }
    } else if (mode == 4) {  // updateH_PW_XY: K22 gpu.h:217-231, K23 gpu.h:233-247
      const R *E = Eh.data() + nstep + N1 - I;
      for (size_t y = j0; y <= j1 - 1; ++y)
        for (size_t x = i0; x <= i1; ++x) {
          size_t a = IDX(x, y, k0 - 1), b = IDX(x, y, k1);
          Hx[a] = Hx[a] - Chxe * E[x] * c;
          Hx[b] = Hx[b] + Chxe * E[x] * c;
        }
      for (size_t z = k0; z <= k1; ++z)
        for (size_t y = j0; y <= j1 - 1; ++y) {
          size_t a = IDX(i0 - 1, y, z), b = IDX(i1, y, z);
          Hz[a] = Hz[a] + Chze * E[i0] * c;
          Hz[b] = Hz[b] - Chze * E[i1] * c;
        }
    } else if (mode == 6) {  // updateH_PW_YX (src/simulation.h:446-456): gpu.h:294-324
      const R *E = Eh.data() + nstep + N1 - J;  // the y modes' offset by analogy with :330-369
      for (size_t y = j0; y <= j1; ++y)
        for (size_t x = i0; x <= i1 - 1; ++x) {
          size_t a = IDX(x, y, k0 - 1), b = IDX(x, y, k1);
          Hy[a] = Hy[a] + Chye * E[y] * c;
          Hy[b] = Hy[b] - Chye * E[y] * c;
        }
      for (size_t z = k0; z <= k1; ++z)
        for (size_t x = i0; x <= i1 - 1; ++x) {
          size_t a = IDX(x, j0 - 1, z), b = IDX(x, j1, z);
          Hz[a] = Hz[a] - Chze * E[j0] * c;
          Hz[b] = Hz[b] + Chze * E[j1] * c;
        }
    } else {  // mode 7, updateH_PW_YZ (src/simulation.h:457-467): gpu.h:359-390
      const R *E = Eh.data() + nstep + N1 - J;
      for (size_t z = k0; z <= k1 - 1; ++z)
        for (size_t x = i0; x <= i1; ++x) {
          size_t a = IDX(x, j0 - 1, z), b = IDX(x, j1, z);
          Hx[a] = Hx[a] + Chxe * E[j0] * c;
          Hx[b] = Hx[b] - Chxe * E[j1] * c;
        }
      for (size_t z = k0; z <= k1 - 1; ++z)
        for (size_t y = j0; y <= j1; ++y) {
          size_t a = IDX(i0 - 1, y, z), b = IDX(i1, y, z);
          Hy[a] = Hy[a] - Chye * E[y] * c;
          Hy[b] = Hy[b] + Chye * E[y] * c;
        }
    }
    return 0;
  }

  int update_e_pw(int n, int mode) {
    if (mode < 1 || mode > 7) return 0;
    if (graded()) { g_err = "simFDTD:InvalidInput|a graded mesh takes no plane wave (the 1D incident line is uniform)"; return 1; }
    if (!have_1d) { g_err = "simFDTD:InvalidInput|plane wave before Init1D"; return 2; }
    if (n < 0 || (size_t)n >= T) { g_err = "simFDTD:InvalidInput|time index outside [0,T)"; return 1; }
    if (mode == 5) return oblique_e((size_t)n);
    const bool inv = (mode == 4);
    const std::vector<R> &Hh = inv ? Hinvh : H1h;
    if (Hh.empty()) { g_err = "simFDTD:InvalidInput|the 1D line for this mode was not initialised"; return 2; }
    R *Ex = F[EX].data(), *Ey = F[EY].data(), *Ez = F[EZ].data();
    const R *Cexh = Ch[0].data(), *Ceyh = Ch[1].data(), *Cezh = Ch[2].data();
    size_t nstep = (size_t)n * N1;
    if (mode == 1) {  // updateE_PW_ZY (c = +1): K32 gpu.h:479-493, K33 gpu.h:495-509
      const R c = 1;
      const R *H = Hh.data() + nstep + N1 - K;
      for (size_t y = j0; y <= j1 - 1; ++y)
        for (size_t x = i0; x <= i1; ++x) {
          size_t a = IDX(x, y, k0), b = IDX(x, y, k1);
          Ey[a] = Ey[a] - Ceyh[a] * H[k0 - 1] * c;
          Ey[b] = Ey[b] + Ceyh[b] * H[k1] * c;
        }
      for (size_t z = k0; z <= k1 - 1; ++z)
        for (size_t x = i0; x <= i1; ++x) {
          size_t a = IDX(x, j0, z), b = IDX(x, j1, z);
          Ez[a] = Ez[a] + Cezh[a] * H[z] * c;
          Ez[b] = Ez[b] - Cezh[b] * H[z] * c;
        }
    } else if (mode == 2) {  // updateE_PW_ZX (kernels get -c): K28 gpu.h:544-558, K29 gpu.h:560-573
      const R c = -1;
      const R *H = Hh.data() + nstep + N1 - K;
      for (size_t y = j0; y <= j1; ++y)
        for (size_t x = i0; x <= i1 - 1; ++x) {
          size_t a = IDX(x, y, k0), b = IDX(x, y, k1);
          Ex[a] = Ex[a] + Cexh[a] * H[k0 - 1] * c;
          Ex[b] = Ex[b] - Cexh[b] * H[k1] * c;
        }
      for (size_t z = k0; z <= k1 - 1; ++z)
        for (size_t y = j0; y <= j1; ++y) {
          size_t a = IDX(i0, y, z), b = IDX(i1, y, z);
          Ez[a] = Ez[a] - Cezh[a] * H[z] * c;
          Ez[b] = Ez[b] + Cezh[b] * H[z] * c;
        }
    } else if (mode == 3) {  // updateE_PW_XZ (c = +1): K20 gpu.h:184-198, K21 gpu.h:200-214
      const R c = 1;
      const R *H = Hh.data() + nstep + N1 - I;
      for (size_t y = j0; y <= j1; ++y)
        for (size_t x = i0; x <= i1 - 1; ++x) {
          size_t a = IDX(x, y, k0), b = IDX(x, y, k1);
          Ex[a] = Ex[a] + Cexh[a] * H[x] * c;
          Ex[b] = Ex[b] - Cexh[b] * H[x] * c;
        }
      for (size_t z = k0; z <= k1 - 1; ++z)
        for (size_t y = j0; y <= j1; ++y) {
          size_t a = IDX(i0, y, z), b = IDX(i1, y, z);
          Ez[a] = Ez[a] - Cezh[a] * H[i0 - 1] * c;
          Ez[b] = Ez[b] + Cezh[b] * H[i1] * c;
        }
    } else if (mode == 4) {  // updateE_PW_XY (kernels get -c): K24 gpu.h:249-275, K25 gpu.h:277-291
      const R c = -1;
      const R *H = Hh.data() + nstep + N1 - I;
      for (size_t z = k0; z <= k1; ++z)
        for (size_t x = i0; x <= i1 - 1; ++x) {
          size_t a = IDX(x, j0, z), b = IDX(x, j1, z);
          Ex[a] = Ex[a] - Cexh[a] * H[x] * c;
          Ex[b] = Ex[b] + Cexh[b] * H[x] * c;
        }
      for (size_t z = k0; z <= k1; ++z)
        for (size_t y = j0; y <= j1 - 1; ++y) {
          size_t a = IDX(i0, y, z), b = IDX(i1, y, z);
          Ey[a] = Ey[a] + Ceyh[a] * H[i0 - 1] * c;
          Ey[b] = Ey[b] - Ceyh[b] * H[i1] * c;
        }
    } else if (mode == 6) {  // updateE_PW_YX (c = +1, src/simulation.h:519-531): gpu.h:326-356
      const R c = 1;
      const R *H = Hh.data() + nstep + N1 - J;
      for (size_t z = k0; z <= k1; ++z)
        for (size_t x = i0; x <= i1 - 1; ++x) {
          size_t a = IDX(x, j0, z), b = IDX(x, j1, z);
          Ex[a] = Ex[a] - Cexh[a] * H[j0 - 1] * c;
          Ex[b] = Ex[b] + Cexh[b] * H[j1] * c;
        }
      for (size_t z = k0; z <= k1; ++z)
        for (size_t y = j0; y <= j1 - 1; ++y) {
          size_t a = IDX(i0, y, z), b = IDX(i1, y, z);
          Ey[a] = Ey[a] + Ceyh[a] * H[y] * c;
          Ey[b] = Ey[b] - Ceyh[b] * H[y] * c;
        }
    } else {  // mode 7, updateE_PW_YZ (kernels get -c, src/simulation.h:532-544): gpu.h:391-422
      const R c = -1;
      const R *H = Hh.data() + nstep + N1 - J;
      for (size_t y = j0; y <= j1 - 1; ++y)
        for (size_t x = i0; x <= i1; ++x) {
          size_t a = IDX(x, y, k0), b = IDX(x, y, k1);
          Ey[a] = Ey[a] - Ceyh[a] * H[y] * c;
          Ey[b] = Ey[b] + Ceyh[b] * H[y] * c;
        }
      for (size_t z = k0; z <= k1 - 1; ++z)
        for (size_t x = i0; x <= i1; ++x) {
          size_t a = IDX(x, j0, z), b = IDX(x, j1, z);
          Ez[a] = Ez[a] + Cezh[a] * H[j0 - 1] * c;
          Ez[b] = Ez[b] - Cezh[b] * H[j1] * c;
        }
    }
    return 0;
  }

  // ---- oblique TF/SF (extension; no reference counterpart) ----
  // Incident value of the 1D line row at fractional index s: linear interpolation.
  static R lerp_line(const R *row, R s) {
    long m = (long)std::floor(s);
    R w = s - (R)m;
    if (w == (R)0) return row[m];
    return row[m] + w * (row[m + 1] - row[m]);
  }
  // E component c at the Yee position of cell (x,y,z); staggering: Ex (h,i,i) Ey (i,h,i) Ez (i,i,h)
  R incE(int c, size_t x, size_t y, size_t z, const R *row) const {
    R sx = c == 0 ? tabE_h[0][x] : tabE_i[0][x];
    R sy = c == 1 ? tabE_h[1][y] : tabE_i[1][y];
    R sz = c == 2 ? tabE_h[2][z] : tabE_i[2][z];
    return polE[c] * lerp_line(row, (sx + sy) + sz);
  }
  // H component c at its Yee position: Hx (i,h,h) Hy (h,i,h) Hz (h,h,i)
  R incH(int c, size_t x, size_t y, size_t z, const R *row) const {
    R sx = c == 0 ? tabH_i[0][x] : tabH_h[0][x];
    R sy = c == 1 ? tabH_i[1][y] : tabH_h[1][y];
    R sz = c == 2 ? tabH_i[2][z] : tabH_h[2][z];
    return polH[c] * lerp_line(row, (sx + sy) + sz);
  }

  // H corrections: faces normal to x, then y, then z; each face pair low then high.
  int oblique_h(size_t n) {
    R *Hx = F[HX].data(), *Hy = F[HY].data(), *Hz = F[HZ].data();
    const R *row = E1h.data() + n * N1;
    // x faces: Hy (Ez extent) and Hz (Ey extent)
    for (size_t z = k0; z <= k1 - 1; ++z)
      for (size_t y = j0; y <= j1; ++y) {
        size_t a = IDX(i0 - 1, y, z), b = IDX(i1, y, z);
        Hy[a] = Hy[a] - Che * incE(2, i0, y, z, row);
        Hy[b] = Hy[b] + Che * incE(2, i1, y, z, row);
      }
    for (size_t z = k0; z <= k1; ++z)
      for (size_t y = j0; y <= j1 - 1; ++y) {
        size_t a = IDX(i0 - 1, y, z), b = IDX(i1, y, z);
        Hz[a] = Hz[a] + Che * incE(1, i0, y, z, row);
        Hz[b] = Hz[b] - Che * incE(1, i1, y, z, row);
      }
    // y faces: Hx (Ez extent) and Hz (Ex extent)
    for (size_t z = k0; z <= k1 - 1; ++z)
      for (size_t x = i0; x <= i1; ++x) {
        size_t a = IDX(x, j0 - 1, z), b = IDX(x, j1, z);
        Hx[a] = Hx[a] + Che * incE(2, x, j0, z, row);
        Hx[b] = Hx[b] - Che * incE(2, x, j1, z, row);
      }
    for (size_t z = k0; z <= k1; ++z)
      for (size_t x = i0; x <= i1 - 1; ++x) {
        size_t a = IDX(x, j0 - 1, z), b = IDX(x, j1, z);
        Hz[a] = Hz[a] - Che * incE(0, x, j0, z, row);
        Hz[b] = Hz[b] + Che * incE(0, x, j1, z, row);
      }
    // z faces: Hx (Ey extent) and Hy (Ex extent)
    for (size_t y = j0; y <= j1 - 1; ++y)
      for (size_t x = i0; x <= i1; ++x) {
        size_t a = IDX(x, y, k0 - 1), b = IDX(x, y, k1);
        Hx[a] = Hx[a] - Che * incE(1, x, y, k0, row);
        Hx[b] = Hx[b] + Che * incE(1, x, y, k1, row);
      }
    for (size_t y = j0; y <= j1; ++y)
      for (size_t x = i0; x <= i1 - 1; ++x) {
        size_t a = IDX(x, y, k0 - 1), b = IDX(x, y, k1);
        Hy[a] = Hy[a] + Che * incE(0, x, y, k0, row);
        Hy[b] = Hy[b] - Che * incE(0, x, y, k1, row);
      }
    return 0;
  }

  int oblique_e(size_t n) {
    R *Ex = F[EX].data(), *Ey = F[EY].data(), *Ez = F[EZ].data();
    const R *Cexh = Ch[0].data(), *Ceyh = Ch[1].data(), *Cezh = Ch[2].data();
    const R *row = H1h.data() + n * N1;
    // x faces: Ez uses Hy at i0-1/2 and i1+1/2; Ey uses Hz
    for (size_t z = k0; z <= k1 - 1; ++z)
      for (size_t y = j0; y <= j1; ++y) {
        size_t a = IDX(i0, y, z), b = IDX(i1, y, z);
        Ez[a] = Ez[a] - Cezh[a] * incH(1, i0 - 1, y, z, row);
        Ez[b] = Ez[b] + Cezh[b] * incH(1, i1, y, z, row);
      }
    for (size_t z = k0; z <= k1; ++z)
      for (size_t y = j0; y <= j1 - 1; ++y) {
        size_t a = IDX(i0, y, z), b = IDX(i1, y, z);
        Ey[a] = Ey[a] + Ceyh[a] * incH(2, i0 - 1, y, z, row);
        Ey[b] = Ey[b] - Ceyh[b] * incH(2, i1, y, z, row);
      }
    // y faces: Ex uses Hz, Ez uses Hx
    for (size_t z = k0; z <= k1; ++z)
      for (size_t x = i0; x <= i1 - 1; ++x) {
        size_t a = IDX(x, j0, z), b = IDX(x, j1, z);
        Ex[a] = Ex[a] - Cexh[a] * incH(2, x, j0 - 1, z, row);
        Ex[b] = Ex[b] + Cexh[b] * incH(2, x, j1, z, row);
      }
    for (size_t z = k0; z <= k1 - 1; ++z)
      for (size_t x = i0; x <= i1; ++x) {
        size_t a = IDX(x, j0, z), b = IDX(x, j1, z);
        Ez[a] = Ez[a] + Cezh[a] * incH(0, x, j0 - 1, z, row);
        Ez[b] = Ez[b] - Cezh[b] * incH(0, x, j1, z, row);
      }
    // z faces: Ex uses Hy, Ey uses Hx
    for (size_t y = j0; y <= j1; ++y)
      for (size_t x = i0; x <= i1 - 1; ++x) {
        size_t a = IDX(x, y, k0), b = IDX(x, y, k1);
        Ex[a] = Ex[a] + Cexh[a] * incH(1, x, y, k0 - 1, row);
        Ex[b] = Ex[b] - Cexh[b] * incH(1, x, y, k1, row);
      }
    for (size_t y = j0; y <= j1 - 1; ++y)
      for (size_t x = i0; x <= i1; ++x) {
        size_t a = IDX(x, y, k0), b = IDX(x, y, k1);
        Ey[a] = Ey[a] - Ceyh[a] * incH(0, x, y, k0 - 1, row);
        Ey[b] = Ey[b] + Ceyh[b] * incH(0, x, y, k1, row);
      }
    return 0;
  }

  // K17 AddSource_GPU_3D_1DThread, src/gpu.h:55-82 via addCurrents src/simulation.h:256-286
  int add_currents(const int32_t *time, const int32_t *ind) {
    if (!init_ready) { g_err = "simFDTD:InvalidInput|AddCurrents before ReadySimulation"; return 2; }
    if (!init_source) { g_err = "simFDTD:InvalidInput|AddCurrents before SetSource"; return 2; }
    int n = time[0], mode = time[1];
    if (n < 0 || (size_t)n >= T) { g_err = "simFDTD:InvalidInput|time index outside [0,T)"; return 1; }
    if (mode < 1 || mode > 3) return 0;
    long a0 = ind[0], a1 = ind[1], b0 = ind[2], b1 = ind[3], c0 = ind[4], c1 = ind[5];
    if (a1 < a0 || b1 < b0 || c1 < c0) return 0;
    // The reference takes the box as int (simulation.h:260-266) and hands it to the kernel as
    // size_t (gpu.h:56-58): a negative start becomes a huge unsigned value, so either i1 < i0
    // holds and the kernel returns (gpu.h:60-61), or every index fails the i < Imax test
    // (gpu.h:75).  A negative start on any axis therefore adds nothing.
    if (a0 < 0 || b0 < 0 || c0 < 0) return 0;
    R *E = F[mode - 1].data();
    const R *Cxh = Ch[mode - 1].data();
    const R s = source[(size_t)n];
    // the upper ends are clipped per cell (gpu.h:75)
    for (long z = c0; z <= c1 && z < (long)K; ++z)
      for (long y = b0; y <= b1 && y < (long)J; ++y)
        for (long x = a0; x <= a1 && x < (long)I; ++x) {
          size_t id = IDX((size_t)x, (size_t)y, (size_t)z);
          E[id] = E[id] - Cxh[id] * s * delX / delX / delX;
        }
    return 0;
  }

  // One full step in the canonical order of highfrequency.m:86-91.
  int step(int n, int mode) {
    update_h_all();
    int rc = 0;
    if (mode >= 1 && mode <= 7) rc = update_h_pw(n, mode);
    if (rc) return rc;
    update_h_pml();
    update_e_all();
    if (mode >= 1 && mode <= 7) rc = update_e_pw(n, mode);
    if (rc) return rc;
    update_e_pml();
    return 0;
  }
};

}  // namespace orc

using namespace orc;

template <typename R>
static Sim<R> *S(void *h) { return static_cast<Sim<R> *>(static_cast<Base *>(h)); }

struct Handle {
  int dbl;
  Base *p;
};

#define DISPATCH(h, expr)                                              \
  do {                                                                 \
    Handle *H_ = (Handle *)(h);                                        \
    if (!H_) { g_err = "simFDTD:InvalidInput|null handle"; return 1; } \
    if (H_->dbl) { auto *s = S<double>(H_->p); return (expr); }         \
    else { auto *s = S<float>(H_->p); return (expr); }                 \
  } while (0)

template <typename R>
static void widen_out(const std::vector<R> &v, double *out) {
  for (size_t i = 0; i < v.size(); ++i) out[i] = (double)v[i];
}

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

// OpenMP threads of the parallel loops (the CPU baseline times 1 thread and all of them).
int orc_set_threads(int n) {
  if (n < 1) return 1;
  omp_set_num_threads(n);
  return 0;
}
int orc_max_threads(void) { return omp_get_max_threads(); }

void *orc_create(int dbl) {
  Handle *h = new Handle;
  h->dbl = dbl;
  if (dbl) h->p = new Sim<double>();
  else h->p = new Sim<float>();
  return h;
}
void orc_destroy(void *h) {
  Handle *H_ = (Handle *)h;
  if (!H_) return;
  delete H_->p;
  delete H_;
}
int orc_set_const(void *h, const float *fr, const int32_t *td) { DISPATCH(h, s->set_const(fr, td)); }
int orc_set_cpml(void *h, const double *p) { DISPATCH(h, s->set_cpml(p)); }
int orc_set_tfsf(void *h, int32_t width) {
  DISPATCH(h, (s->init_ready ? 2 : (s->width = (size_t)width, 0)));
}
int orc_set_source(void *h, const float *src, int update) { DISPATCH(h, s->set_source(src, update)); }
int orc_set_mesh(void *h, const float *dx, const float *dy, const float *dz) {
  const float *const d[3] = {dx, dy, dz};
  DISPATCH(h, s->set_mesh(d));
}
int orc_set_antenna(void *h, const int32_t *loc) { DISPATCH(h, s->set_antenna(loc)); }
int orc_set_probe_comps(void *h, const int32_t *c, int n) { DISPATCH(h, s->set_probe_comps(c, n)); }
int orc_update_material(void *h, int32_t flag, const float *v) { DISPATCH(h, s->update_material(flag, v)); }
int orc_create_mu(void *h, const float *mx, const float *my, const float *mz) { DISPATCH(h, s->create_mu(mx, my, mz)); }
int orc_ready(void *h) { DISPATCH(h, s->ready()); }
int orc_set_plane_wave(void *h, int32_t mode, const double *prm) { DISPATCH(h, s->set_plane_wave(mode, prm)); }
int orc_init_1d(void *h, float rate, const float *pw) { DISPATCH(h, s->init_1d(rate, pw)); }
int orc_update_coeffs(void *h) { DISPATCH(h, s->update_coeffs()); }
int orc_reset(void *h) { DISPATCH(h, s->reset()); }
int orc_update_h_all(void *h) { DISPATCH(h, (s->update_h_all(), 0)); }
int orc_update_e_all(void *h) { DISPATCH(h, (s->update_e_all(), 0)); }
int orc_update_h_pml(void *h) { DISPATCH(h, (s->update_h_pml(), 0)); }
int orc_update_e_pml(void *h) { DISPATCH(h, (s->update_e_pml(), 0)); }
int orc_update_h_pw(void *h, const int32_t *t) { DISPATCH(h, s->update_h_pw(t[0], t[1])); }
int orc_update_e_pw(void *h, const int32_t *t) { DISPATCH(h, s->update_e_pw(t[0], t[1])); }
int orc_add_currents(void *h, const int32_t *t, const int32_t *box) { DISPATCH(h, s->add_currents(t, box)); }

// Steps n0 .. n0+count-1 in the canonical order; optional snapshots of component `snap_comp`
// (0..5) after every `snap_every`-th completed step into snap_out (double, IJK per snapshot).
int orc_run(void *h, int32_t n0, int32_t count, int32_t mode, int32_t snap_every, int32_t snap_comp,
            double *snap_out) {
  Handle *H_ = (Handle *)h;
  auto body = [&](auto *s) -> int {
    if (!s->init_ready) { g_err = "simFDTD:InvalidInput|Run before ReadySimulation"; return 2; }
    size_t nsnap = 0;
    for (int32_t q = 0; q < count; ++q) {
      int rc = s->step(n0 + q, mode);
      if (rc) return rc;
      if (snap_every > 0 && snap_out && (s->rec_step % (size_t)snap_every) == 0) {
        const auto &f = s->F[snap_comp];
        double *dst = snap_out + nsnap * s->IJK;
        for (size_t i = 0; i < s->IJK; ++i) dst[i] = (double)f[i];
        nsnap++;
      }
    }
    return 0;
  };
  if (!H_) return 1;
  if (H_->dbl) return body(S<double>(H_->p));
  return body(S<float>(H_->p));
}

// ---- readback (all values widened to double; float -> double is exact) ----

// what: 0..5 fields, 6..11 Ce{x,y,z}/Ch{x,y,z}, 12..17 eps x,y,z / sigma x,y,z, 20..31 psi
int orc_get_array(void *h, int32_t what, double *out) {
  Handle *H_ = (Handle *)h;
  auto body = [&](auto *s) -> int {
    if (what >= 0 && what < 6) widen_out(s->F[what], out);
    else if (what >= 6 && what < 9) widen_out(s->Ce[what - 6], out);
    else if (what >= 9 && what < 12) widen_out(s->Ch[what - 9], out);
    else if (what >= 12 && what < 15) widen_out(s->eps[what - 12], out);
    else if (what >= 15 && what < 18) widen_out(s->sig[what - 15], out);
    else if (what >= 20 && what < 32) widen_out(s->psi[what - 20], out);
    else { g_err = "simFDTD:InvalidInput|unknown array"; return 1; }
    return 0;
  };
  if (!H_) return 1;
  if (H_->dbl) return body(S<double>(H_->p));
  return body(S<float>(H_->p));
}

int64_t orc_array_len(void *h, int32_t what) {
  Handle *H_ = (Handle *)h;
  auto body = [&](auto *s) -> int64_t {
    if (what >= 0 && what < 18) return (int64_t)s->IJK;
    if (what >= 20 && what < 32) return (int64_t)s->psi[what - 20].size();
    return -1;
  };
  if (!H_) return -1;
  if (H_->dbl) return body(S<double>(H_->p));
  return body(S<float>(H_->p));
}

// what: 0 bE, 1 cE, 2 bH, 3 cH (2P each), 4 kapE, 5 kapH (P), 6..8 ikE x/y/z, 9..11 ikH x/y/z,
// 12 scalars [Chh, Che, sig_max, N1, i0, i1, j0, j1, k0, k1, P, ratio, src]
int orc_get_setup(void *h, int32_t what, double *out) {
  Handle *H_ = (Handle *)h;
  auto body = [&](auto *s) -> int {
    switch (what) {
      case 0: widen_out(s->bE, out); break;
      case 1: widen_out(s->cE, out); break;
      case 2: widen_out(s->bH, out); break;
      case 3: widen_out(s->cH, out); break;
      case 4: widen_out(s->kapE, out); break;
      case 5: widen_out(s->kapH, out); break;
      case 6: case 7: case 8: widen_out(s->ikE[what - 6], out); break;
      case 9: case 10: case 11: widen_out(s->ikH[what - 9], out); break;
      case 12: {
        double v[13] = {(double)s->Chh, (double)s->Che, (double)s->sig_max, (double)s->N1,
                        (double)s->i0, (double)s->i1, (double)s->j0, (double)s->j1,
                        (double)s->k0, (double)s->k1, (double)s->P, s->pw_ratio, (double)s->pw_src};
        std::memcpy(out, v, sizeof(v));
        break;
      }
      case 32: {
        for (int c = 0; c < 3; ++c) { out[c] = (double)s->polE[c]; out[3 + c] = (double)s->polH[c]; }
        break;
      }
      default:
        if (what >= 20 && what < 32) {  // oblique tables: 20 + 4*axis + {E int, E half, H int, H half}
          const int u = (what - 20) / 4, q = (what - 20) % 4;
          widen_out(q == 0 ? s->tabE_i[u] : q == 1 ? s->tabE_h[u] : q == 2 ? s->tabH_i[u] : s->tabH_h[u], out);
          break;
        }
        g_err = "simFDTD:InvalidInput|unknown setup array";
        return 1;
    }
    return 0;
  };
  if (!H_) return 1;
  if (H_->dbl) return body(S<double>(H_->p));
  return body(S<float>(H_->p));
}

// 1D histories: which 0 E fwd, 1 H fwd, 2 E inv, 3 H inv -> T*N1 values
int orc_get_line(void *h, int32_t which, double *out) {
  Handle *H_ = (Handle *)h;
  auto body = [&](auto *s) -> int {
    const auto &v = which == 0 ? s->E1h : which == 1 ? s->H1h : which == 2 ? s->Einvh : s->Hinvh;
    if (v.empty()) { g_err = "simFDTD:InvalidInput|line not initialised"; return 2; }
    widen_out(v, out);
    return 0;
  };
  if (!H_) return 1;
  if (H_->dbl) return body(S<double>(H_->p));
  return body(S<float>(H_->p));
}

// probe record [ncomp][T][nAnt]
int orc_get_probes(void *h, double *out) {
  Handle *H_ = (Handle *)h;
  auto body = [&](auto *s) -> int {
    if (!s->init_antenna) { g_err = "simFDTD:InvalidInput|Antenna data is not initialized."; return 2; }
    widen_out(s->port, out);
    return 0;
  };
  if (!H_) return 1;
  if (H_->dbl) return body(S<double>(H_->p));
  return body(S<float>(H_->p));
}

// Overwrite a field (0..5) from double values (converted to R) -- test set-up of initial states.
int orc_set_field(void *h, int32_t comp, const double *v) {
  Handle *H_ = (Handle *)h;
  auto body = [&](auto *s) -> int {
    if (comp < 0 || comp > 5 || !s->init_ready) return 1;
    using RR = typename std::remove_reference<decltype(s->F[0][0])>::type;
    for (size_t i = 0; i < s->IJK; ++i) s->F[comp][i] = (RR)v[i];
    return 0;
  };
  if (!H_) return 1;
  if (H_->dbl) return body(S<double>(H_->p));
  return body(S<float>(H_->p));
}

}  // extern "C"
