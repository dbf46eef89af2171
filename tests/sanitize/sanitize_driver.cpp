// Host-only driver run under AddressSanitizer and UndefinedBehaviorSanitizer (built with g++ by
// tests/test_sanitizers.py): the library's host set-up (CPML grading, 1/kappa tables, material
// coefficients, the oblique plane-wave tables) and the CPU oracle's ops on small grids in every
// plane-wave mode, with probes, AddCurrents boxes past the grid and with negative starts, snapshots
// and every readback.  Any out-of-bounds access, leak or undefined operation aborts the run.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "fdtd_internal.hpp"

extern "C" {
void *orc_create(int dbl);
void orc_destroy(void *h);
int orc_set_const(void *h, const float *fr, const int32_t *td);
int orc_set_cpml(void *h, const double *p);
int orc_set_tfsf(void *h, int32_t width);
int orc_set_source(void *h, const float *src, int update);
int orc_set_antenna(void *h, const int32_t *loc);
int orc_set_probe_comps(void *h, const int32_t *c, int n);
int orc_update_material(void *h, int32_t flag, const float *v);
int orc_create_mu(void *h, const float *mx, const float *my, const float *mz);
int orc_ready(void *h);
int orc_set_plane_wave(void *h, int32_t mode, const double *prm);
int orc_init_1d(void *h, float rate, const float *pw);
int orc_update_coeffs(void *h);
int orc_reset(void *h);
int orc_update_h_all(void *h);
int orc_update_e_all(void *h);
int orc_update_h_pml(void *h);
int orc_update_e_pml(void *h);
int orc_update_h_pw(void *h, const int32_t *t);
int orc_update_e_pw(void *h, const int32_t *t);
int orc_add_currents(void *h, const int32_t *t, const int32_t *box);
int orc_run(void *h, int32_t n0, int32_t count, int32_t mode, int32_t snap_every, int32_t snap_comp, double *snap_out);
int orc_get_array(void *h, int32_t what, double *out);
int64_t orc_array_len(void *h, int32_t what);
int orc_get_setup(void *h, int32_t what, double *out);
int orc_get_line(void *h, int32_t which, double *out);
int orc_get_probes(void *h, double *out);
const char *orc_last_error(void);
}

static int fails = 0;
#define CHECK(x)                                                                            \
  do {                                                                                      \
    if ((x) != 0) {                                                                         \
      std::printf("FAIL %s:%d %s -> %s\n", __FILE__, __LINE__, #x, orc_last_error());      \
      ++fails;                                                                              \
    }                                                                                       \
  } while (0)

static void host_setup() {
  using namespace fdtd;
  const CpmlParams sets[3] = {CpmlParams(), {8, 4.0f, 2.5f, 0.2f, 1.5f, 1.1}, {12, 3.0f, 5.0f, 0.0f, 1.0f, 0.8}};
  for (const CpmlParams &p : sets) {
    CpmlProfiles pr = compute_cpml(p, 0.01f, 1.9e-11f);
    for (size_t Q : {(size_t)2 * p.P + 2, (size_t)64}) {
      std::vector<float> a = inv_kappa_E(pr, p.P, Q), b = inv_kappa_H(pr, p.P, Q);
      if (a.size() != Q || b.size() != Q) ++fails;
    }
  }
  float ce = 0, ch = 0;
  material_coeffs(4.0f, 0.02f, 1.9e-11f, 100.0f, &ce, &ch);
  if (!(ce > 0 && ch > 0)) ++fails;
  const int lo[3] = {12, 13, 14}, hi[3] = {40, 45, 50}, Q[3] = {54, 60, 66};
  for (double th : {0.0, 0.4, 1.2, 3.0})
    for (double ph : {0.0, 2.0, 5.5}) {
      ObliqueSetup o = oblique_setup(th, ph, 0.7, lo, hi, Q, 8, 0.01f, 1.9e-11f, 3e9f);
      for (int a = 0; a < 3; ++a)
        for (int h = 0; h < 2; ++h)
          if (o.tabE[a][h].size() != (size_t)Q[a] || o.tabH[a][h].size() != (size_t)Q[a]) ++fails;
    }
  (void)matched_ratio(0.3, 0.4, 0.866, 0.01, 1.9e-11, 3e9);
}

static void oracle_run(int dbl, int mode) {
  const int T = 30, I = 30, J = 32, K = 34, nAnt = 3;
  void *h = orc_create(dbl);
  const float fr[3] = {3e9f, 0.01f, 1.9e-11f};
  const int32_t td[5] = {T, I, J, K, nAnt};
  CHECK(orc_set_const(h, fr, td));
  const double cp[6] = {8, 3, 5, 0.05, 1, 0.8};
  CHECK(orc_set_cpml(h, cp));
  CHECK(orc_set_tfsf(h, 11));
  std::vector<float> src(T);
  for (int n = 0; n < T; ++n) src[n] = std::exp(-((n - 8.0f) * (n - 8.0f)) / 9.0f);
  CHECK(orc_set_source(h, src.data(), 0));
  const int32_t comps[3] = {1, 3, 5};
  CHECK(orc_set_probe_comps(h, comps, 3));
  const int32_t loc[9] = {0, 5, I - 17, 0, 6, J - 17, 0, 7, K - 17};
  CHECK(orc_set_antenna(h, loc));
  std::vector<float> eps((size_t)I * J * K, 1.0f), sig((size_t)I * J * K, 0.0f);
  for (size_t q = 0; q < eps.size(); q += 7) { eps[q] = 3.0f; sig[q] = 0.01f; }
  for (int f = 1; f <= 3; ++f) CHECK(orc_update_material(h, f, sig.data()));
  for (int f = 4; f <= 6; ++f) CHECK(orc_update_material(h, f, eps.data()));
  if (mode == 6) {
    std::vector<float> mu(eps.size(), 2.0f);
    CHECK(orc_create_mu(h, mu.data(), mu.data(), mu.data()));
  }
  if (mode == 5) {
    const double prm[4] = {0.7, 2.1, 0.4, 0};
    CHECK(orc_set_plane_wave(h, 5, prm));
  } else if (mode == 2) {
    const double prm[4] = {1, 0, 0, 0};  // layered line
    CHECK(orc_set_plane_wave(h, 2, prm));
  }
  CHECK(orc_ready(h));
  if (mode) CHECK(orc_init_1d(h, 1.0f, src.data()));
  CHECK(orc_update_coeffs(h));
  CHECK(orc_reset(h));
  for (int n = 0; n < 6; ++n) {
    const int32_t t[2] = {n, mode};
    CHECK(orc_update_h_all(h));
    CHECK(orc_update_h_pw(h, t));
    CHECK(orc_update_h_pml(h));
    CHECK(orc_update_e_all(h));
    CHECK(orc_update_e_pw(h, t));
    CHECK(orc_update_e_pml(h));
    const int32_t tc[2] = {n, 1 + n % 3};
    const int32_t boxes[3][6] = {{2, 40, 3, 3, 4, 50}, {-1, 5, 2, 3, 2, 3}, {3, 2, 2, 3, 2, 3}};
    CHECK(orc_add_currents(h, tc, boxes[n % 3]));
  }
  std::vector<double> snaps((size_t)(24 / 5 + 1) * I * J * K);
  CHECK(orc_run(h, 6, 24, mode, 5, 2, snaps.data()));
  for (int what : {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31}) {
    const int64_t n = orc_array_len(h, what);
    if (n <= 0) continue;
    std::vector<double> out((size_t)n);
    CHECK(orc_get_array(h, what, out.data()));
  }
  for (int what : {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12}) {
    std::vector<double> out(4096);
    CHECK(orc_get_setup(h, what, out.data()));
  }
  if (mode && mode != 2 && mode != 5) {
    std::vector<double> line((size_t)T * (K > I ? K : I) * 2);
    CHECK(orc_get_line(h, 0, line.data()));
  }
  std::vector<double> pr((size_t)3 * T * nAnt);
  CHECK(orc_get_probes(h, pr.data()));
  orc_destroy(h);
}

int main() {
  host_setup();
  for (int dbl : {0, 1})
    for (int mode : {0, 1, 2, 3, 4, 5, 6, 7}) oracle_run(dbl, mode);
  if (fails) {
    std::printf("SANITIZE FAILED (%d)\n", fails);
    return 1;
  }
  std::printf("SANITIZE OK\n");
  return 0;
}
