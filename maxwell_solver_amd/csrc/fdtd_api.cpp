// fdtd_api.cpp -- simulation state and C ABI of libfdtd_hip.so (declared in include/fdtd.h).
//
// One fdtd_sim per simulation, as the reference's `class simulation` (src/simulation.h:24).  Host
// set-up (CPML grading, coefficients, plane-wave tables) runs in fdtd_prepare without touching the
// device; ReadySimulation then allocates and uploads.  The fields and CPML psi arrays live in HBM
// for the life of the simulation; the fused step ping-pongs between two copies of them.
#include <dlfcn.h>
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <deque>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <map>
#include <mutex>
#include <set>
#include <tuple>
#include <vector>

#include "../../include/fdtd.h"
#include "fdtd_internal.hpp"
#include "fdtd_kernels.hpp"

using namespace fdtd;

namespace {

thread_local std::string g_err;
thread_local std::string g_err_id;

int fail(int code, const std::string &msg) {
  g_err = msg;
  switch (code) {
    case FDTD_ERR_ALREADY: g_err_id = "simFDTD:AlreadyInitialized"; break;
    case FDTD_ERR_DEVICE: g_err_id = "simFDTD:DeviceError"; break;
    case FDTD_ERR_NOMEM: g_err_id = "simFDTD:OutOfMemory"; break;
    case FDTD_ERR_UNKNOWN_OP: g_err_id = "simFDTD:UnknownOperation"; break;
    default: g_err_id = "simFDTD:InvalidInput"; break;
  }
  return code;
}

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fail(FDTD_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

// Row pitch of the per-cell device arrays.  A row of I floats that is not a whole number of
// 128-byte lines starts off a line boundary in every row after the first, so each wave's load and
// store of 64 consecutive columns straddles three lines, two of them partly written by another
// wave.  Where the state is larger than the 256 MB Infinity Cache the step's stores stream to HBM
// and such lines cost a read-modify-write there (252 x 256 x 256: 60.3 Gcell/s, 68.7 with rows
// padded to 256; 256^3: 74.4); where it fits, the partial lines meet in the cache and padding
// gains nothing (250 x 280 x 50: 54.2 against 53.5 padded) while adding bytes.  So rows are padded
// to a multiple of 32 floats when the state streams; a multiple of 64 columns is kept as it is
// (the split row layout needs I itself).  FDTD_ROW_ALIGN = 1 (no padding), 16, 32 or 64 sets the
// multiple whatever the state size.
static int row_pitch(int I, bool streams) {
  int a = streams ? 32 : 1;
  if (const char *e = std::getenv("FDTD_ROW_ALIGN")) {
    const int v = std::atoi(e);
    if (v == 1 || v == 16 || v == 32 || v == 64) a = v;
  }
  if (a == 1 || I % 64 == 0) return I;
  return (I + a - 1) / a * a;
}

struct fdtd_sim {
  // ---- flags (src/simulation.h:1305-1311) ----
  bool init_const = false, init_source = false, init_antenna = false, init_ready = false;
  bool prepared = false, have_1d = false;

  // ---- SetConst ----
  float fctr = 0, delX = 0, delT = 0;
  int T = 0, I = 0, J = 0, K = 0, nAnt = 0;
  size_t IJ = 0, IJK = 0;
  // per-cell device arrays (fields, materials, per-cell coefficients): rows of Ip >= I elements
  // (row_pitch), planes of IJp = Ip*J, IJKp = Ip*J*K elements; host arrays stay I*J*K
  int Ip = 0;
  size_t IJp = 0, IJKp = 0;
  CpmlParams cp;
  int width = 15;

  // ---- materials ----
  bool have_arrays = false;  // six host arrays (UpdateMaterial layout)
  std::vector<float> eps[3], sig[3];
  bool have_index = false;   // SetMaterialIndex form
  std::vector<uint8_t> midx;
  std::vector<float> mtab;   // [n][6] sigma x,y,z, eps x,y,z
  int nmat = 0;

  // ---- prepared host set-up ----
  CpmlProfiles prof;
  std::vector<float> ikE[3], ikH[3];
  float Chh = 1, Che = 0, den = 0;
  // graded mesh (fdtd_set_mesh): cell sizes along x, y, z (empty: uniform dx along that axis)
  std::vector<float> mesh[3];
  bool graded() const { return !mesh[0].empty() || !mesh[1].empty() || !mesh[2].empty(); }
  GridDev g{};
  bool index_mode = true;
  std::vector<uint8_t> cell_mat;   // index mode: material id per cell
  std::vector<float> coef_tab;     // index mode: [n][6] Ce x,y,z, Ch x,y,z
  std::vector<float> cell_Ce[3], cell_Ch[3];  // full mode
  int pw_mode = 0, pw_layered = 0;
  double pw_prm[3] = {0, 0, 0};
  ObliqueSetup obl;
  int N1 = 0, line_src = 0;

  // ---- source / probes ----
  std::vector<float> source;
  std::vector<int32_t> ant_loc;
  std::vector<int> coax;
  std::vector<int> probe_comps = {EZ};
  int rec_step = 0;

  // ---- device ----
  hipStream_t stream = nullptr;
  int device = 0;  // the HIP device of ReadySimulation
  float *dFb[2] = {}, *dPsiB[2] = {};  // packed field / psi sets
  float *dF[2][6] = {};
  float *dPsi[2][12] = {};
  size_t psi_len[12] = {};
  size_t psi_total = 0;
  size_t cs = 0;                                   // field component stride
  static constexpr size_t kCompSkew = 2112;        // 8 KiB + 256 B
  static constexpr size_t kSetSkew = 3136;         // 12 KiB + 256 B
  int cur = 0;
  bool second = false;
  uint8_t *dMat = nullptr;
  float *dTab = nullptr;
  float *dCe[3] = {}, *dCh[3] = {};
  float *dIkE[3] = {}, *dIkH[3] = {};
  float *dbE = nullptr, *dcE = nullptr, *dbH = nullptr, *dcH = nullptr;
  float *dE1h = nullptr, *dH1h = nullptr, *dEinv = nullptr, *dHinv = nullptr;
  float *dPrm = nullptr;      // bE cE bH cH, 1/kappa E x,y,z, 1/kappa H x,y,z
  float *dCoefFull = nullptr; // full mode: Ce x,y,z, Ch x,y,z per cell
  float *dPwTab[8] = {};      // per plane-wave mode (1..7): E tables x,y,z (int, half), then H tables
  float *dXs[2] = {nullptr, nullptr};
  std::vector<float> mu_in[3];               // CreateMu: relative permeabilities as given
  float *dChmu[3] = {nullptr, nullptr, nullptr};  // their per-cell H coefficients (K2)  // face sheets of the TF/SF box, one buffer per state set (fused step)
  int *dProbeIdx = nullptr, *dProbeComps = nullptr;
  float *dPort = nullptr;
  size_t port_len = 0;
  std::vector<float *> dSnaps;
  int nsnap = 0;
  PWDev pwdev[8];
  StepTuning tune;
  bool addr_global = false;  // fdtd_set_addressing(1): the row kernel's 64-bit global path
  bool tune_user = false;  // fdtd_set_tuning chose the tuning: the size-class table leaves it
  int tune_policy = 1;     // fdtd_set_autotune: 0 heuristics, 1 size-class table, 2 table + autotune on a miss
  // device argument blocks of the fused kernel per (state parity, plane-wave mode); rebuilt when
  // args_gen moves (any buffer or tuning change)
  struct ArgSlot {
    int gen = -1, nwg = 0;
    void *dptr = nullptr;
    std::vector<char> host;
  };
  ArgSlot argslots[2][8];
  int args_gen = 0;
  // fdtd_run_noalloc: while set, the fused step takes its TF/SF face sheets and its row-kernel
  // argument blocks from the GPU's device scratch (scratch_region) instead of the simulation's own
  // buffers, and builds the argument blocks in na_host; nothing is allocated.
  struct NoAlloc {
    float *xs[2];        // face-sheet buffers of state sets 0 and 1
    char *args[2];       // argument blocks of state sets 0 and 1 (device)
    bool uploaded[2];    // argument block of the set built and copied in this call
    int nwg[2];
  };
  NoAlloc *na = nullptr;
  alignas(16) char na_host[2][kScratchArgBytes];
  int timing = 0;  // HIP events around every timing-th step of fdtd_run (0: none)
  std::vector<hipEvent_t> ev;
  double kernel_ms = 0, rows_ms = 0;  // summed step / row-kernel event times of the last timed run
  int64_t kernel_launches = 0;
  std::map<void *, size_t> allocs;  // device allocations and their bytes
  size_t dev_bytes = 0;  // device bytes held (snapshots excluded)
  std::vector<void *> line_allocs;  // buffers of the last Init1D (line histories, its uploads, tables)
  // Step currents: AddCurrents(n, pol, box) after every step of fdtd_step / fdtd_run, the loop of
  // GICsimulation.m:119-124 (pol 0: none).
  int step_pol = 0, step_box[6] = {0, 0, 0, 0, 0, 0};
  int currents_gen = 0, source_gen = 0;
  // Graph replay of runs (fdtd_set_graph_steps): graph_steps steps recorded once into a HIP graph
  // and replayed; the step currents of recorded launches read their sample from dSrc[*dStepN + q].
  int graph_steps = 0;
  float *dSrc = nullptr;
  int dsrc_gen = -1;
  int *dStepN = nullptr;
  hipGraphExec_t gexec = nullptr;
  int gx_cur = -1, gx_args = -1, gx_steps = 0, gx_currents = -1, gx_mode = -1;
  bool gx_probes = false;

  ~fdtd_sim() {
    if (stream) hipStreamSynchronize(stream);
    if (gexec) hipGraphExecDestroy(gexec);
    for (auto &kv : allocs) hipFree(kv.first);
    for (float *p : dSnaps) hipFree(p);
    for (hipEvent_t e : ev) hipEventDestroy(e);
    if (stream) hipStreamDestroy(stream);
  }

  template <typename Tp>
  int dalloc(Tp **p, size_t n) {
    void *q = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&q, n * sizeof(Tp));
    if (e != hipSuccess) return fail(FDTD_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs[q] = n * sizeof(Tp);
    dev_bytes += n * sizeof(Tp);
    *p = (Tp *)q;
    return 0;
  }
  // Frees a buffer of dalloc once the simulation's queued work is done, and sets *p to null.
  template <typename Tp>
  void dfree(Tp **p) {
    if (!*p) return;
    auto it = allocs.find((void *)*p);
    if (it != allocs.end()) {
      if (stream) hipStreamSynchronize(stream);
      hipFree(it->first);
      dev_bytes -= it->second;
      allocs.erase(it);
    }
    *p = nullptr;
  }
  // The line buffers of the previous Init1D (before a new one replaces them).
  void free_line_buffers() {
    for (void *q : line_allocs) dfree(&q);
    line_allocs.clear();
    for (int m = 0; m < 8; ++m) dfree(&dPwTab[m]);
    dE1h = dH1h = dEinv = dHinv = nullptr;
  }

  // Upper bound of the device bytes the simulation holds once ready, with its 1D line and after
  // fused steps (two field and psi sets, materials, tables, line histories, face sheets, probe
  // record), from the set-up alone (no device call): the request server admits against it.
  size_t device_bytes_bound(int mode) const {
    const size_t P = (size_t)cp.P;
    const size_t sx = 2 * P * J * K, sy = 2 * (size_t)I * P * K, sz = 2 * (size_t)I * J * P;
    const size_t IJKb = (size_t)((I + 63) / 64 * 64) * J * K;  // per-cell arrays at the widest row pitch
    size_t b = 2 * ((6 * (IJKb + kCompSkew) + kSetSkew) + 4 * (sx + sy + sz)) * sizeof(float);
    b += have_index ? IJKb + 256 * 6 * sizeof(float) : 6 * IJKb * sizeof(float);
    b += (8 * P + 2 * ((size_t)I + J + K)) * sizeof(float);
    if (mode) {
      const size_t N = (size_t)I + J + K + 64;                // line nodes, oblique included
      b += 2 * (2 * (size_t)T * N + 4 * N + T) * sizeof(float);  // two lines: histories + uploads
      b += 6 * 4 * ((size_t)I + J + K) * sizeof(float);       // plane-wave tables
      const size_t ni = I, nj = J, nk = K;
      b += 2 * 8 * (nj * nk + ni * nk + ni * nj) * sizeof(float);  // face sheets, two buffers
    }
    if (init_antenna) b += (probe_comps.size() * (size_t)T * nAnt + nAnt + probe_comps.size()) * sizeof(float);
    if (!mu_in[0].empty()) b += 3 * IJKb * sizeof(float);  // CreateMu coefficients
    return b + ((size_t)1 << 20);  // argument blocks and small tables
  }
  template <typename Tp>
  int upload(Tp **p, const std::vector<Tp> &v) {
    if (int rc = dalloc(p, v.size())) return rc;
    HIPCHK(hipMemcpyAsync(*p, v.data(), v.size() * sizeof(Tp), hipMemcpyHostToDevice, stream));
    return 0;
  }
  // A host array of I*J*K cells to a per-cell device array (rows of Ip) and back.  Padded rows go
  // through a host copy of the padded layout (zero padding columns) and one linear transfer: a
  // strided transfer from pageable memory moves a row at a time (a 320^3 material upload took
  // longer than the solve).  The padded case completes before returning.
  template <typename Tp>
  hipError_t put_cells(Tp *dev, const Tp *host) const {
    if (Ip == I) return hipMemcpyAsync(dev, host, IJK * sizeof(Tp), hipMemcpyHostToDevice, stream);
    std::vector<Tp> pad(IJKp, Tp(0));
    for (size_t r = 0; r < (size_t)J * K; ++r) std::memcpy(&pad[r * Ip], host + r * I, (size_t)I * sizeof(Tp));
    hipError_t e = hipMemcpyAsync(dev, pad.data(), IJKp * sizeof(Tp), hipMemcpyHostToDevice, stream);
    return e != hipSuccess ? e : hipStreamSynchronize(stream);
  }
  template <typename Tp>
  hipError_t get_cells(Tp *host, const Tp *dev) const {
    if (Ip == I) return hipMemcpyAsync(host, dev, IJK * sizeof(Tp), hipMemcpyDeviceToHost, stream);
    std::vector<Tp> pad(IJKp);
    hipError_t e = hipMemcpyAsync(pad.data(), dev, IJKp * sizeof(Tp), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return e;
    for (size_t r = 0; r < (size_t)J * K; ++r) std::memcpy(host + r * I, &pad[r * Ip], (size_t)I * sizeof(Tp));
    return hipSuccess;
  }

  FieldSet fields(int which) const {
    FieldSet F;
    for (int c = 0; c < 6; ++c) F.f[c] = dF[which][c];
    return F;
  }
  PsiSet psis(int which) const {
    PsiSet P;
    for (int q = 0; q < 12; ++q) P.p[q] = dPsi[which][q];
    return P;
  }
  CoefDev coefs() const {
    CoefDev co;
    co.mat = index_mode ? dMat : nullptr;
    co.tab = dTab;
    co.ntab = index_mode ? (int)(coef_tab.size() / 6) : 0;
    for (int c = 0; c < 3; ++c) {
      co.Ce[c] = dCe[c];
      co.Ch[c] = dCh[c];
      co.ikE[c] = dIkE[c];
      co.ikH[c] = dIkH[c];
    }
    co.bE = dbE; co.cE = dcE; co.bH = dbH; co.cH = dcH;
    co.Chh = Chh; co.Che = Che; co.dx = delX;
    for (int c = 0; c < 3; ++c) co.Chmu[c] = dChmu[c];
    return co;
  }

  // CreateMu coefficients, src/simulation.h:43-61: C = delT / mu0 / mu / (1 + sigma_m delT / 2 /
  // mu0 / mu) * den_hx with the reference's operand types (float delT, mu, sigma_m, den_hx; double
  // mu0), stored as float.  den_hx (:861-885) is 1/dx on i < I-1 and 0 on the last x plane; its
  // kappa factor is applied per axis through the 1/kappa tables, as for den_ex.
  int upload_mu() {
    const float sigma_m = 0;
    std::vector<float> C(IJK);
    if (!dChmu[0]) {
      float *blk = nullptr;
      if (int rc = dalloc(&blk, 3 * IJKp)) return rc;
      HIPCHK(hipMemsetAsync(blk, 0, 3 * IJKp * sizeof(float), stream));  // the padding columns
      for (int c = 0; c < 3; ++c) dChmu[c] = blk + c * IJKp;
      args_gen++;
    }
    for (int c = 0; c < 3; ++c) {
      for (size_t id = 0; id < IJK; ++id) {
        const float mu = mu_in[c][id];
        const float dh = ((int)(id % (size_t)I) == I - 1) ? 0.0f : den;
        C[id] = (float)(delT / kMu0 / mu / (1 + sigma_m * delT / 2 / kMu0 / mu) * dh);
      }
      HIPCHK(put_cells(dChmu[c], C.data()));
      HIPCHK(hipStreamSynchronize(stream));  // C is rebuilt for the next component
    }
    return 0;
  }

  void ensure_arrays() {
    if (have_arrays) return;
    for (int c = 0; c < 3; ++c) {
      eps[c].assign(IJK, 1.0f);
      sig[c].assign(IJK, 0.0f);
    }
    if (have_index) {  // expand the index form into the six arrays
      for (size_t id = 0; id < IJK; ++id) {
        const float *r = &mtab[6 * midx[id]];
        for (int c = 0; c < 3; ++c) { sig[c][id] = r[c]; eps[c][id] = r[3 + c]; }
      }
      have_index = false;
    }
    have_arrays = true;
  }

  // Material classification: the uint8 table form when the grid has at most 256 distinct
  // (sigma, eps) tuples, per-cell coefficient arrays otherwise.
  void build_coefficients() {
    std::vector<float> rows;  // [n][6] sigma, eps
    cell_mat.clear();
    index_mode = true;
    if (have_index) {
      rows = mtab;
      cell_mat = midx;
    } else if (have_arrays) {
      struct Key {
        float v[6];
        bool operator==(const Key &o) const { return std::memcmp(v, o.v, sizeof(v)) == 0; }
      };
      struct KH {
        size_t operator()(const Key &k) const {
          uint32_t u[6];
          std::memcpy(u, k.v, sizeof(u));
          size_t h = 1469598103934665603ull;
          for (int q = 0; q < 6; ++q) h = (h ^ u[q]) * 1099511628211ull;
          return h;
        }
      };
      std::unordered_map<Key, int, KH> ids;
      cell_mat.assign(IJK, 0);
      Key last{};
      int last_id = -1;
      for (size_t id = 0; id < IJK && index_mode; ++id) {
        Key kk;
        for (int c = 0; c < 3; ++c) { kk.v[c] = sig[c][id]; kk.v[3 + c] = eps[c][id]; }
        if (last_id >= 0 && kk == last) { cell_mat[id] = (uint8_t)last_id; continue; }
        auto it = ids.find(kk);
        int m;
        if (it == ids.end()) {
          m = (int)ids.size();
          if (m >= 256) { index_mode = false; break; }
          ids.emplace(kk, m);
          rows.insert(rows.end(), kk.v, kk.v + 6);
        } else {
          m = it->second;
        }
        cell_mat[id] = (uint8_t)m;
        last = kk;
        last_id = m;
      }
    } else {
      rows = {0, 0, 0, 1, 1, 1};
      cell_mat.assign(IJK, 0);
    }
    coef_tab.clear();
    for (int c = 0; c < 3; ++c) { cell_Ce[c].clear(); cell_Ch[c].clear(); }
    if (index_mode) {
      const size_t n = rows.size() / 6;
      coef_tab.assign(n * 6, 0.0f);
      for (size_t m = 0; m < n; ++m)
        for (int c = 0; c < 3; ++c)
          material_coeffs(rows[6 * m + 3 + c], rows[6 * m + c], delT, den, &coef_tab[6 * m + c], &coef_tab[6 * m + 3 + c]);
    } else {
      cell_mat.clear();
      for (int c = 0; c < 3; ++c) {
        cell_Ce[c].assign(IJK, 0.0f);
        cell_Ch[c].assign(IJK, 0.0f);
      }
      for (size_t id = 0; id < IJK; ++id) {
        const float d = ((int)(id % (size_t)I) == I - 1) ? 0.0f : den;
        for (int c = 0; c < 3; ++c) material_coeffs(eps[c][id], sig[c][id], delT, d, &cell_Ce[c][id], &cell_Ch[c][id]);
      }
    }
  }

  int prepare() {
    if (!init_const) return fail(FDTD_ERR_STATE, "ReadySimulation before SetConst");
    const int P = cp.P;
    if (I < 2 * P + 2 || J < 2 * P + 2 || K < 2 * P + 2) return fail(FDTD_ERR_INVALID, "grid smaller than 2P+2 cells along an axis");
    if (width < 1 || 2 * width + 3 > I || 2 * width + 3 > J || 2 * width + 3 > K)
      return fail(FDTD_ERR_INVALID, "TF/SF inset too large for the grid");
    if (!init_ready) {  // the pitch of the device arrays, fixed once they exist
      const size_t sx = 2 * (size_t)P * J * K, sy = 2 * (size_t)I * P * K, sz = 2 * (size_t)I * J * P;
      const size_t state = 2 * 24 * (IJK + kCompSkew) + 2 * 16 * (sx + sy + sz) + IJK;
      Ip = row_pitch(I, state > ((size_t)256 << 20));
      IJp = (size_t)Ip * J;
      IJKp = IJp * K;
    }
    g.I = I; g.J = J; g.K = K; g.P = P;
    g.Ip = Ip;
    g.i0 = width; g.i1 = I - width - 1;
    g.j0 = width; g.j1 = J - width - 1;
    g.k0 = width; g.k1 = K - width - 1;
    prof = compute_cpml(cp, delX, delT);
    ikE[0] = inv_kappa_E(prof, P, I); ikE[1] = inv_kappa_E(prof, P, J); ikE[2] = inv_kappa_E(prof, P, K);
    ikH[0] = inv_kappa_H(prof, P, I); ikH[1] = inv_kappa_H(prof, P, J); ikH[2] = inv_kappa_H(prof, P, K);
    // graded mesh: the H update differences E across the primal cell n (spacing d[n]), the E
    // update H across the dual cell n (spacing (d[n-1] + d[n]) / 2); both enter the per-axis tables
    // as dx / spacing (the reference's den_h* and den_e*, src/simulation.h:861-912, 679-685), the
    // CPML cells keeping dx so that the psi updates and the CPML tables are the uniform grid's
    if (graded()) {
      const size_t Q[3] = {(size_t)I, (size_t)J, (size_t)K};
      const double dx = (double)delX;
      // the time step of SetConst must stay within the Courant limit of the finest cells,
      // dt <= 1 / (c sqrt(sum over the axes of 1 / min spacing^2))
      double inv2 = 0;
      for (int a = 0; a < 3; ++a) {
        double m = dx;
        for (float v : mesh[a]) m = std::min(m, (double)v);
        inv2 += 1.0 / (m * m);
      }
      if ((double)delT > 1.0 / (kC0 * std::sqrt(inv2)))
        return fail(FDTD_ERR_INVALID, "SetMesh: dt exceeds the Courant limit of the finest cells");
      for (int a = 0; a < 3; ++a) {
        const std::vector<float> &d = mesh[a];
        if (d.empty()) continue;
        for (size_t n = 0; n < Q[a]; ++n)
          if ((n <= (size_t)P || n + (size_t)P + 2 >= Q[a]) && (double)d[n] != dx)
            return fail(FDTD_ERR_INVALID, "SetMesh: the CPML cells (P+1 at each end) must keep the spacing dx");
        for (size_t n = 0; n + 1 < Q[a]; ++n) ikH[a][n] = (float)((double)ikH[a][n] * (dx / (double)d[n]));
        for (size_t n = 1; n < Q[a]; ++n)
          ikE[a][n] = (float)((double)ikE[a][n] * (dx / (0.5 * ((double)d[n - 1] + (double)d[n]))));
      }
    }
    // src/simulation.h:914-922 with sigma_m = 0; den_ex = 1/dx (kappa is in the 1/kappa tables)
    const float sigma_m = 0;
    Chh = (float)((1 - sigma_m * delT / 2 / kMu0) / (1 + sigma_m * delT / 2 / kMu0));
    Che = (float)(delT / kMu0 / (1 + sigma_m * delT / 2 / kMu0) / delX);
    den = (float)(1.0 / (delX * 1.0f));
    build_coefficients();
    if (pw_mode == 5) {
      const int lo[3] = {g.i0, g.j0, g.k0}, hi[3] = {g.i1, g.j1, g.k1}, Q[3] = {I, J, K};
      obl = oblique_setup(pw_prm[0], pw_prm[1], pw_prm[2], lo, hi, Q, P, delX, delT, fctr);
    }
    prepared = true;
    return 0;
  }

  int resolve_antenna() {
    coax.assign(nAnt, 0);
    for (int a = 0; a < nAnt; ++a) {
      const long x = cp.P + (long)ant_loc[a], y = cp.P + (long)ant_loc[a + nAnt], z = cp.P + (long)ant_loc[a + 2 * nAnt];
      if (x < 0 || y < 0 || z < 0 || x >= I || y >= J || z >= K) {
        init_antenna = false;
        return fail(FDTD_ERR_INVALID, "antenna location outside the grid");
      }
      coax[a] = (int)(x + y * (long)Ip + z * (long)IJp);
    }
    return 0;
  }

  int alloc_probes() {
    if (!init_antenna || !init_ready) return 0;
    if (int rc = resolve_antenna()) return rc;
    port_len = probe_comps.size() * (size_t)T * nAnt;
    dfree(&dPort);
    dfree(&dProbeIdx);
    dfree(&dProbeComps);
    if (int rc = dalloc(&dPort, port_len)) return rc;
    HIPCHK(hipMemsetAsync(dPort, 0, port_len * sizeof(float), stream));
    if (int rc = upload(&dProbeIdx, coax)) return rc;
    if (int rc = upload(&dProbeComps, probe_comps)) return rc;
    return 0;
  }

  // Materials to the device (ReadySimulation, UpdateCuda).  The buffers of the mode in use are
  // allocated once and copied into afterwards, as the reference copies into its arrays
  // (src/simulation.h:1116-1127); switching between the index and the per-cell form frees the
  // other form's buffers.
  int upload_coefficients() {
    args_gen++;
    if (index_mode) {
      dfree(&dCoefFull);
      for (int c = 0; c < 3; ++c) dCe[c] = dCh[c] = nullptr;
      if (!dMat) {
        if (dalloc(&dMat, IJKp)) return FDTD_ERR_NOMEM;
        HIPCHK(hipMemsetAsync(dMat, 0, IJKp, stream));  // the padding columns: material 0
      }
      if (!dTab && dalloc(&dTab, (size_t)256 * 6)) return FDTD_ERR_NOMEM;
      HIPCHK(put_cells(dMat, cell_mat.data()));
      HIPCHK(hipMemcpyAsync(dTab, coef_tab.data(), coef_tab.size() * sizeof(float), hipMemcpyHostToDevice, stream));
    } else {
      dfree(&dMat);
      dfree(&dTab);
      if (!dCoefFull) {
        if (dalloc(&dCoefFull, 6 * IJKp)) return FDTD_ERR_NOMEM;
        HIPCHK(hipMemsetAsync(dCoefFull, 0, 6 * IJKp * sizeof(float), stream));  // the padding columns
      }
      for (int c = 0; c < 3; ++c) {
        dCe[c] = dCoefFull + c * IJKp;
        dCh[c] = dCoefFull + (3 + c) * IJKp;
        HIPCHK(put_cells(dCe[c], cell_Ce[c].data()));
        HIPCHK(put_cells(dCh[c], cell_Ch[c].data()));
      }
    }
    return 0;
  }

  int ready() {
    if (init_ready) return fail(FDTD_ERR_ALREADY, "ReadySimulation: already called.");
    if (int rc = prepare()) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FDTD_ERR_DEVICE, "no HIP device");
    HIPCHK(hipStreamCreate(&stream));
    HIPCHK(hipGetDevice(&device));
    if (int rc = alloc_set(0)) return rc;
    // both state sets of the fused step's ping-pong at set-up, so that no step allocates
    if (int rc = ensure_second()) return rc;
    if (int rc = upload_coefficients()) return rc;
    // parameter table: bE cE bH cH [2P each], 1/kappa E x,y,z, 1/kappa H x,y,z
    std::vector<float> prm;
    for (const auto *v : {&prof.bE, &prof.cE, &prof.bH, &prof.cH, &ikE[0], &ikE[1], &ikE[2], &ikH[0], &ikH[1], &ikH[2]})
      prm.insert(prm.end(), v->begin(), v->end());
    if (int rc = upload(&dPrm, prm)) return rc;
    const size_t P2 = 2 * (size_t)cp.P;
    dbE = dPrm; dcE = dPrm + P2; dbH = dPrm + 2 * P2; dcH = dPrm + 3 * P2;
    dIkE[0] = dPrm + 4 * P2; dIkE[1] = dIkE[0] + I; dIkE[2] = dIkE[1] + J;
    dIkH[0] = dIkE[2] + K; dIkH[1] = dIkH[0] + I; dIkH[2] = dIkH[1] + J;
    init_ready = true;
    if (!mu_in[0].empty())
      if (int rc = upload_mu()) return rc;
    if (int rc = alloc_probes()) return rc;
    return zero_state();
  }

  int zero_state() {
    HIPCHK(hipMemsetAsync(dFb[cur], 0, 6 * cs * sizeof(float), stream));
    HIPCHK(hipMemsetAsync(dPsiB[cur], 0, psi_total * sizeof(float), stream));
    if (dPort) HIPCHK(hipMemsetAsync(dPort, 0, port_len * sizeof(float), stream));
    rec_step = 0;
    return 0;
  }

  // One field set (six components at stride Ip*J*K plus a skew) and one psi set, packed by slab orientation:
  // x slabs [Eyx Ezx Hyx Hzx], y slabs [Exy Ezy Hxy Hzy], z slabs [Exz Eyz Hxz Hyz].
  int alloc_set(int which) {
    const size_t P = (size_t)cp.P;
    const size_t sx = 2 * P * J * K, sy = 2 * (size_t)I * P * K, sz = 2 * (size_t)I * J * P;
    // components cs = I*J*K + skew apart, and the two sets offset differently, so that the twelve
    // streams a step touches at one cell do not sit at large power-of-two distances (HBM channel
    // interleaving would otherwise send them to the same channels)
    cs = IJKp + kCompSkew;
    float *blk = nullptr;
    if (int rc = dalloc(&blk, 6 * cs + kSetSkew)) return rc;
    dFb[which] = blk + (which ? kSetSkew : 0);
    for (int c = 0; c < 6; ++c) dF[which][c] = dFb[which] + c * cs;
    if (int rc = dalloc(&dPsiB[which], 4 * (sx + sy + sz))) return rc;
    float *px = dPsiB[which], *py = px + 4 * sx, *pz = py + 4 * sy;
    // three groups of four psi arrays, the members of a group interleaved per element (element t
    // of member m at 4 t + m): the fused step reads and writes a cell's four values at once
    float **q = dPsi[which];
    q[P_EYX] = px; q[P_EZX] = px + 1; q[P_HYX] = px + 2; q[P_HZX] = px + 3;
    q[P_EXY] = py; q[P_EZY] = py + 1; q[P_HXY] = py + 2; q[P_HZY] = py + 3;
    q[P_EXZ] = pz; q[P_EYZ] = pz + 1; q[P_HXZ] = pz + 2; q[P_HYZ] = pz + 3;
    const size_t len[12] = {sy, sz, sx, sz, sx, sy, sy, sz, sx, sz, sx, sy};
    for (int k = 0; k < 12; ++k) psi_len[k] = len[k];
    psi_total = 4 * (sx + sy + sz);
    return 0;
  }

  int ensure_second() {
    if (second) return 0;
    if (int rc = alloc_set(1)) return rc;
    second = true;
    args_gen++;
    return 0;
  }

  // ---- 1D line and plane-wave parameter blocks ----
  int run_line(int N, float ds, float Ch1, const std::vector<float> &Ce, const std::vector<float> &Cb, int src,
               const std::vector<float> &pw, float **dEh, float **dHh) {
    std::vector<float> ikE1 = inv_kappa_E(prof, cp.P, (size_t)N), ikH1 = inv_kappa_H(prof, cp.P, (size_t)N);
    float *dCe1, *dCb1, *dikE1, *dikH1, *dpw;
    if (int rc = upload(&dCe1, Ce)) return rc;
    if (int rc = upload(&dCb1, Cb)) return rc;
    if (int rc = upload(&dikE1, ikE1)) return rc;
    if (int rc = upload(&dikH1, ikH1)) return rc;
    if (int rc = upload(&dpw, pw)) return rc;
    if (int rc = dalloc(dEh, (size_t)T * N)) return rc;
    if (int rc = dalloc(dHh, (size_t)T * N)) return rc;
    for (void *q : {(void *)dCe1, (void *)dCb1, (void *)dikE1, (void *)dikH1, (void *)dpw, (void *)*dEh, (void *)*dHh})
      line_allocs.push_back(q);
    LineDev L;
    L.N = N; L.T = T; L.P = cp.P; L.src = src; L.Ch = Ch1; L.ds = ds;
    L.Ce = dCe1; L.Cb = dCb1; L.ikE = dikE1; L.ikH = dikH1;
    L.bE = dbE; L.cE = dcE; L.bH = dbH; L.cH = dcH; L.pw = dpw; L.Eh = *dEh; L.Hh = *dHh;
    HIPCHK(launch_line(stream, L));
    return 0;
  }

  int init_1d(float rate, const float *pwsrc) {
    if (!init_ready) return fail(FDTD_ERR_STATE, "Init1D before ReadySimulation");
    std::vector<float> pw((size_t)T);
    for (int i = 0; i < T; ++i) pw[i] = pwsrc[i] * rate;  // src/simulation.h:188-192
    have_1d = false;
    free_line_buffers();  // a repeated Init1D replaces the previous line
    if (pw_mode == 5) {
      N1 = obl.N1;
      line_src = obl.src;
      if (int rc = run_line(N1, obl.ds, obl.Ch1, std::vector<float>(N1, 1.0f), std::vector<float>(N1, obl.Cb1), obl.src, pw, &dE1h, &dH1h)) return rc;
    } else if (pw_layered && pw_mode >= 1 && pw_mode != 5 && pw_mode <= 7) {
      // exact restriction of the centre column along the propagation axis
      const int ax = pw_mode <= 2 ? 2 : pw_mode <= 4 ? 0 : 1;
      const int pol = pw_mode == 1 ? EY : pw_mode == 2 ? EX : pw_mode == 3 ? EZ : pw_mode == 4 ? EY : pw_mode == 6 ? EX : EZ;
      const int Q = ax == 2 ? K : ax == 0 ? I : J;
      const size_t ic = I / 2, jc = J / 2, kc = K / 2;
      std::vector<float> Ce(Q), Cb(Q);
      for (int q = 0; q < Q; ++q) {
        const size_t id = ax == 2 ? (ic + jc * I + (size_t)q * IJ)
                          : ax == 0 ? ((size_t)q + jc * I + kc * IJ) : (ic + (size_t)q * I + kc * IJ);
        cell_coeff(id, pol, &Ce[q], &Cb[q]);
      }
      N1 = Q;
      if (pw_mode == 4) {
        line_src = 13;
        if (int rc = run_line(Q, delX, Che, Ce, Cb, 13, pw, &dEinv, &dHinv)) return rc;
      } else {
        line_src = Q - 13;
        if (int rc = run_line(Q, delX, Che, Ce, Cb, Q - 13, pw, &dE1h, &dH1h)) return rc;
      }
    } else {
      // src/simulation.h:134-254: vacuum line of max(I,J,K) nodes, forward and inverse passes
      N1 = std::max(std::max(I, J), K);
      const float Ceh_1D = (float)(delT / kEps0 / delX);
      std::vector<float> Ce(N1, 1.0f), Cb(N1, Ceh_1D);
      line_src = N1 - 13;
      if (int rc = run_line(N1, delX, Che, Ce, Cb, N1 - 13, pw, &dE1h, &dH1h)) return rc;
      if (int rc = run_line(N1, delX, Che, Ce, Cb, 13, pw, &dEinv, &dHinv)) return rc;
    }
    if (int rc = build_pw()) return rc;
    have_1d = true;
    return 0;
  }

  void cell_coeff(size_t id, int comp, float *ce, float *ch) const {
    if (index_mode) {
      const float *r = &coef_tab[6 * cell_mat[id]];
      *ce = r[comp];
      *ch = ((int)(id % (size_t)I) == I - 1) ? 0.0f : r[3 + comp];
    } else {
      *ce = cell_Ce[comp][id];
      *ch = cell_Ch[comp][id];
    }
  }

  // Per-mode incident-field tables, one buffer each: E tables (x int, x half, y int, y half, z int,
  // z half) then the same for H.  Axis modes map the propagation coordinate q to line node
  // q + N1 - Q (src/simulation.h:330-369) and hold zeros on the other axes: 1, 2 along z, 3, 4
  // along x, 6 (YX) and 7 (YZ) along y, the y-propagating modes the reference defines but never
  // dispatches (src/simulation.h:446-467, 519-544), with the y extent J in place of K or I.
  int build_pw() {
    args_gen++;
    const size_t tl = 2 * ((size_t)I + J + K);
    for (int m = 0; m < 8; ++m) pwdev[m] = PWDev();
    for (int m = 1; m <= 7; ++m) {
      if (m == 5 && pw_mode != 5) continue;
      std::vector<float> t(2 * tl, 0.0f);
      const int Q[3] = {I, J, K};
      size_t off[3] = {0, 2 * (size_t)I, 2 * (size_t)I + 2 * (size_t)J};
      for (int eh = 0; eh < 2; ++eh)
        for (int u = 0; u < 3; ++u)
          for (int h = 0; h < 2; ++h)
            for (int q = 0; q < Q[u]; ++q) {
              float v = 0.0f;
              if (m == 5) v = eh == 0 ? obl.tabE[u][h][q] : obl.tabH[u][h][q];
              else if ((m <= 2 && u == 2) || ((m == 3 || m == 4) && u == 0) || (m >= 6 && u == 1))
                v = (float)(q + N1 - Q[u]);
              t[eh * tl + off[u] + (size_t)h * Q[u] + q] = v;
            }
      if (int rc = upload(&dPwTab[m], t)) return rc;
      PWDev &p = pwdev[m];
      for (int u = 0; u < 3; ++u)
        for (int h = 0; h < 2; ++h) {
          p.tabE[u][h] = dPwTab[m] + off[u] + (size_t)h * Q[u];
          p.tabH[u][h] = dPwTab[m] + tl + off[u] + (size_t)h * Q[u];
        }
      p.N1 = N1;
      p.Eh = m == 4 ? dEinv : dE1h;
      p.Hh = m == 4 ? dHinv : dH1h;
      if (!p.Eh || !p.Hh) continue;  // the line of this mode was not computed
      const int all = (1 << 1) | (1 << 2) | (1 << 3) | (1 << 5) | (1 << 6) | (1 << 7);
      if (m == 1) { p.polE[1] = 1; p.polH[0] = 1; p.active_h = (1 << 6) | (1 << 2); p.active_e = (1 << 7) | (1 << 5); }
      if (m == 2) { p.polE[0] = 1; p.polH[1] = -1; p.active_h = (1 << 7) | (1 << 5); p.active_e = (1 << 6) | (1 << 2); }
      if (m == 3) { p.polE[2] = 1; p.polH[1] = 1; p.active_h = (1 << 3) | (1 << 1); p.active_e = (1 << 6) | (1 << 2); }
      if (m == 4) { p.polE[1] = 1; p.polH[2] = -1; p.active_h = (1 << 6) | (1 << 2); p.active_e = (1 << 3) | (1 << 1); }
      // YX (src/gpu.h:294-356): Hy on the z faces, Hz on the y faces; Ex on the y faces, Ey on
      // the x faces; the E kernels get +c
      if (m == 6) { p.polE[0] = 1; p.polH[2] = 1; p.active_h = (1 << 7) | (1 << 5); p.active_e = (1 << 3) | (1 << 1); }
      // YZ (src/gpu.h:359-422): Hx on the y faces, Hy on the x faces; Ey on the z faces, Ez on
      // the y faces; the E kernels get -c (updateE_PW_YZ, src/simulation.h:532-544)
      if (m == 7) { p.polE[2] = 1; p.polH[0] = -1; p.active_h = (1 << 3) | (1 << 1); p.active_e = (1 << 7) | (1 << 5); }
      if (m == 5) {
        for (int c = 0; c < 3; ++c) { p.polE[c] = obl.polE[c]; p.polH[c] = obl.polH[c]; }
        p.active_h = all;
        p.active_e = all;
      }
    }
    return 0;
  }

  int check_pw(int n, int mode, const PWDev **out) {
    *out = nullptr;
    if (mode < 1 || mode > 7) return 0;  // the reference ignores other modes
    if (graded()) return fail(FDTD_ERR_INVALID, "a graded mesh takes no plane wave (the 1D incident line is uniform)");
    if (!have_1d) return fail(FDTD_ERR_STATE, "plane wave before Init1D");
    if (n < 0 || n >= T) return fail(FDTD_ERR_INVALID, "time index outside [0, T)");
    if (!pwdev[mode].active_h) return fail(FDTD_ERR_STATE, "the 1D line for this plane-wave mode was not initialised");
    *out = &pwdev[mode];
    return 0;
  }

  int record_probes(int which) {
    if (dPort && rec_step < T)
      HIPCHK(launch_probes(stream, fields(which), dProbeIdx, nAnt, dProbeComps, (int)probe_comps.size(), dPort, rec_step, T));
    rec_step++;
    return 0;
  }

  // The fused step's description and row-kernel argument block for current set `which` and
  // plane wave pw (null: none), built and uploaded when anything it depends on changed.
  // Floats of one face-sheet buffer of the TF/SF box (0 when the box is empty).
  size_t sheet_floats() const {
    if (g.i1 < g.i0 || g.j1 < g.j0 || g.k1 < g.k0) return 0;
    const size_t ni = g.i1 - g.i0 + 1, nj = g.j1 - g.j0 + 1, nk = g.k1 - g.k0 + 1;
    return 8 * (nj * nk + ni * nk + ni * nj);
  }

  // The row-kernel argument block of set `which` in an allocation-free run: built into na_host and
  // copied to the device scratch once per call (the scratch is shared by every simulation of the
  // GPU, so a previous call may have left another simulation's block there).
  int noalloc_args(int which, const FusedDev &f, const void **dargs, const void **hargs, int *nwg) {
    if (!na->uploaded[which]) {
      std::memset(na_host[which], 0, sizeof(na_host[which]));
      na->nwg[which] = make_step_args(f, tune, na_host[which]);
      if (na->nwg[which] < 0 && (tune.cpl || tune.tile_rows))
        return fail(FDTD_ERR_INVALID, "launch tuning not supported for this grid");
      if (na->nwg[which] >= 0)
        HIPCHK(hipMemcpyAsync(na->args[which], na_host[which], step_args_size(), hipMemcpyHostToDevice, stream));
      na->uploaded[which] = true;
    }
    *dargs = na->args[which];
    *hargs = na_host[which];
    *nwg = na->nwg[which];
    return 0;
  }

  int fused_slot(int which, int mode, const PWDev *pw, FusedDev &f, ArgSlot **out) {
    if (int rc = ensure_second()) return rc;
    const int slot_mode = pw ? mode : 0;
    if (na) {  // allocation-free run: the face sheets are the device scratch's
      fill_fused(which, mode, pw, pw ? na->xs[which] : nullptr, f);
      *out = nullptr;
      return 0;
    }
    if (pw && !dXs[0] && g.i1 >= g.i0 && g.j1 >= g.j0 && g.k1 >= g.k0) {
      const size_t ni = g.i1 - g.i0 + 1, nj = g.j1 - g.j0 + 1, nk = g.k1 - g.k0 + 1;
      const size_t nxs = 8 * (nj * nk + ni * nk + ni * nj);
      if (dalloc(&dXs[0], nxs) || dalloc(&dXs[1], nxs)) return FDTD_ERR_NOMEM;
      args_gen++;
    }
    ArgSlot &slot = argslots[which][slot_mode];
    fill_fused(which, mode, pw, dXs[which], f);
    if (slot.gen != args_gen) {
      slot.host.assign(step_args_size(), 0);
      slot.nwg = make_step_args(f, tune, slot.host.data());
      if (slot.nwg < 0 && (tune.cpl || tune.tile_rows))
        return fail(FDTD_ERR_INVALID, "launch tuning not supported for this grid");
      if (slot.nwg >= 0) {
        if (!slot.dptr && dalloc((char **)&slot.dptr, slot.host.size())) return FDTD_ERR_NOMEM;
        HIPCHK(hipMemcpyAsync(slot.dptr, slot.host.data(), slot.host.size(), hipMemcpyHostToDevice, stream));
      }
      slot.gen = args_gen;
    }
    *out = &slot;
    return 0;
  }

  // The fused step's description for current set `which`, plane wave pw (null: none) and face
  // sheets xs.
  void fill_fused(int which, int mode, const PWDev *pw, float *xs, FusedDev &f) const {
    f = FusedDev();
    f.g = g;
    f.force_global = addr_global;
    f.Fc = dFb[which]; f.Fn = dFb[1 - which]; f.Pc = dPsiB[which]; f.Pn = dPsiB[1 - which];
    f.cs = cs;
    f.mat = index_mode ? dMat : nullptr;
    f.coef = index_mode ? dTab : dCoefFull;
    f.ntab = index_mode ? (int)(coef_tab.size() / 6) : 0;
    f.prm = dPrm;
    if (pw) {
      f.pwtab = dPwTab[mode]; f.Eh = pw->Eh; f.Hh = pw->Hh; f.N1 = pw->N1; f.xs = xs;
      f.act_h = pw->active_h; f.act_e = pw->active_e;
      for (int c = 0; c < 3; ++c) { f.polE[c] = pw->polE[c]; f.polH[c] = pw->polH[c]; }
    }
    f.Chh = Chh; f.Che = Che; f.dx = delX;
    f.Chmu = dChmu[0];  // the three arrays are one block
    f.gr = graded();
  }

  // State carried from step to step by a run of consecutive steps (fdtd_run): which face sheets
  // the previous step filled in advance, and the probe record row still to be gathered from the
  // current state (the slab part's launch of the next step, or flush_probes, gathers it).
  struct RunFold {
    int sheets_n = -1, sheets_set = -1;
    int probe_row = -1;
  };

  int flush_probes(RunFold &rf) {
    if (rf.probe_row >= 0)
      HIPCHK(launch_probes(stream, fields(cur), dProbeIdx, nAnt, dProbeComps, (int)probe_comps.size(), dPort, rf.probe_row, T));
    rf.probe_row = -1;
    return 0;
  }

  int step_currents(int n) {
    if (!step_pol) return 0;
    if (n < 0 || n >= T) return fail(FDTD_ERR_INVALID, "time index outside [0, T)");
    HIPCHK(launch_add_currents(stream, g, fields(cur), coefs(), step_pol - 1, step_box, source[n], delX));
    return 0;
  }

  // One fused time step.  dry: launch the kernels (they read the current set and write the next)
  // but keep the current set and the probe recorder: the state is unchanged (used for timing).
  // rf (a run of steps): fold the next step's face sheets and this state's probe gather into the
  // step's launches.
  int step(int n, int mode, bool timed, int *ev_slot, bool dry = false, RunFold *rf = nullptr) {
    if (!init_ready) return fail(FDTD_ERR_STATE, "Step before ReadySimulation");
    const PWDev *pw = nullptr;
    if (int rc = check_pw(n, mode, &pw)) return rc;
    FusedDev f;
    ArgSlot *slotp = nullptr;
    if (int rc = fused_slot(cur, mode, pw, f, &slotp)) return rc;
    const void *dargs = nullptr, *hargs = nullptr;
    int nwg = -1;
    if (na) {
      if (int rc = noalloc_args(cur, f, &dargs, &hargs, &nwg)) return rc;
    } else {
      dargs = slotp->dptr;
      hargs = slotp->host.data();
      nwg = slotp->nwg;
    }
    if (nwg < 0) {
      if (dry) return fail(FDTD_ERR_INVALID, "grid outside the fused step's layouts");
      if (rf) {
        if (int rc = flush_probes(*rf)) return rc;
        rf->sheets_n = -1;
      }
      // a grid outside the fused step's layouts (rows over 1024 columns, P over 64, ...): the
      // per-op kernels in the reference op order, in place on the current set
      if (timed) {
        for (int q = 0; q < 3; ++q) HIPCHK(hipEventRecord(ev[(*ev_slot)++], stream));
      }
      HIPCHK(launch_update_h_all(stream, g, fields(cur), coefs()));
      if (pw) HIPCHK(launch_pw_h(stream, g, fields(cur), coefs(), *pw, n));
      HIPCHK(launch_update_h_pml(stream, g, fields(cur), psis(cur), coefs()));
      HIPCHK(launch_update_e_all(stream, g, fields(cur), coefs()));
      if (pw) HIPCHK(launch_pw_e(stream, g, fields(cur), coefs(), *pw, n));
      HIPCHK(launch_update_e_pml(stream, g, fields(cur), psis(cur), coefs()));
      if (timed) HIPCHK(hipEventRecord(ev[(*ev_slot)++], stream));
      if (int rc = record_probes(cur)) return rc;
      return step_currents(n);
    }
    // timed: events before the step, around the row kernel, after the step
    hipEvent_t *e = timed ? &ev[*ev_slot] : nullptr;
    if (timed) {
      HIPCHK(hipEventRecord(e[0], stream));
      *ev_slot += 4;
    }
    if (!rf || dry) {
      HIPCHK(launch_step_fused(stream, f, dargs, hargs, nwg, n, e ? e + 1 : nullptr));
      if (timed) HIPCHK(hipEventRecord(e[3], stream));
      if (dry) return 0;
      cur = 1 - cur;
      if (int rc = record_probes(cur)) return rc;
      return step_currents(n);
    }
    StepFold fo;
    fo.T = T;
    fo.sheets_ready = rf->sheets_n == n && rf->sheets_set == cur;
    fo.xs_next = n + 1 < T ? (na ? na->xs[1 - cur] : dXs[1 - cur]) : nullptr;
    if (rf->probe_row >= 0) {
      fo.F = fields(cur);
      fo.idx = dProbeIdx; fo.comps = dProbeComps; fo.port = dPort;
      fo.nAnt = nAnt; fo.ncomp = (int)probe_comps.size(); fo.row = rf->probe_row; fo.T = T;
    }
    if (step_pol && !dPort) {  // with probes the currents follow the step's probe gather
      if (n < 0 || n >= T) return fail(FDTD_ERR_INVALID, "time index outside [0, T)");
      fo.cur_comp = step_pol - 1;
      for (int q = 0; q < 6; ++q) fo.cur_box[q] = step_box[q];
      fo.cur_s = source[n];
      fo.cur_dx = delX;
    }
    HIPCHK(launch_step_fused(stream, f, dargs, hargs, nwg, n, e ? e + 1 : nullptr, &fo));
    if (timed) HIPCHK(hipEventRecord(e[3], stream));
    rf->sheets_n = fo.next_filled ? n + 1 : -1;
    rf->sheets_set = 1 - cur;
    cur = 1 - cur;
    rf->probe_row = dPort && rec_step < T ? rec_step : -1;
    rec_step++;
    if (step_pol) {  // the probes sample the step's E before its currents (as UpdateE_PML does)
      if (int rc = flush_probes(*rf)) return rc;
      if (!fo.cur_done) return step_currents(n);
    }
    return 0;
  }

  // Average device time of `reps` steps at time index n in the launch form of a run (the row
  // kernel's two launches, the slab part filling the next step's face sheets), with the tuning in
  // force, run dry: the kernels read the current set and write the other, the current set and the
  // probe record stay as they are.  One untimed step first (it also fills the sheets of step n).
  int time_variant(int n, int mode, int reps, hipEvent_t e0, hipEvent_t e1, double *ms) {
    const PWDev *pw = nullptr;
    if (int rc = check_pw(n, mode, &pw)) return rc;
    FusedDev f;
    ArgSlot *slot = nullptr;
    if (int rc = fused_slot(cur, mode, pw, f, &slot)) return rc;
    if (slot->nwg < 0) return fail(FDTD_ERR_INVALID, "grid outside the fused step's layouts");
    HIPCHK(launch_step_fused(stream, f, slot->dptr, slot->host.data(), slot->nwg, n));
    HIPCHK(hipEventRecord(e0, stream));
    for (int q = 0; q < reps; ++q) {
      StepFold fo;
      fo.T = T;
      fo.sheets_ready = true;
      fo.xs_next = pw ? dXs[1 - cur] : nullptr;
      HIPCHK(launch_step_fused(stream, f, slot->dptr, slot->host.data(), slot->nwg, n, nullptr, &fo));
    }
    HIPCHK(hipEventRecord(e1, stream));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    *ms = (double)t / reps;
    return 0;
  }

  // Steps [n0, n0 + count) as replays of one recorded graph of graph_steps steps (an even count, so
  // a replay ends on the set it starts from), where every launch of a step is the same for all
  // steps: no plane wave, probes, snapshots or timing, a grid the fused step covers.  Returns the
  // steps run (a multiple of graph_steps), 0 when the run does not qualify.
  int run_graphs(int n0, int count, int mode, int *done, RunFold &rf) {
    *done = 0;
    const int G = graph_steps;
    if (G < 2 || count < G || timing) return 0;
    if (dPort && step_pol) return 0;  // probes then sample between a step and its currents: eager
    if ((step_pol || mode) && (n0 < 0 || n0 + count > T)) return 0;  // the step loop reports the bad index
    const PWDev *pw = nullptr;
    if (int rc = check_pw(n0, mode, &pw)) return rc;
    const int gmode = pw ? mode : 0;
    FusedDev f[2];
    ArgSlot *slot[2];
    for (int w = 0; w < 2; ++w) {  // argument blocks (and face-sheet buffers) before the recording
      if (int rc = fused_slot(w, mode, pw, f[w], &slot[w])) return rc;
      if (slot[w]->nwg < 0) return 0;
    }
    if (step_pol && dsrc_gen != source_gen) {
      if (!dSrc && dalloc(&dSrc, (size_t)T)) return FDTD_ERR_NOMEM;
      HIPCHK(hipMemcpy(dSrc, source.data(), (size_t)T * sizeof(float), hipMemcpyHostToDevice));
      dsrc_gen = source_gen;
    }
    if (!dStepN && dalloc(&dStepN, 3)) return FDTD_ERR_NOMEM;
    const bool probes = dPort != nullptr;
    if (!gexec || gx_cur != cur || gx_args != args_gen || gx_steps != G || gx_currents != currents_gen ||
        gx_mode != gmode || gx_probes != probes) {
      if (gexec) {
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(hipGraphExecDestroy(gexec));
        gexec = nullptr;
      }
      HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      int c = cur;
      hipError_t e = hipSuccess;
      for (int q = 0; q < G && e == hipSuccess; ++q) {
        // step q of a replay: its time index and record row are dStepN[0] + q and dStepN[1] + q;
        // its face sheets were filled by the step before (the previous replay's last step, or the
        // k_sheets launch before the first replay), it fills those of the next time index, and it
        // gathers the probes of the state it reads (the previous step's record)
        StepFold fo;
        fo.sheets_ready = true;
        fo.xs_next = pw ? dXs[1 - c] : nullptr;
        fo.gidx = dStepN;
        fo.gq = q;
        fo.T = T;
        if (probes) {
          fo.probe_rel = true;
          fo.F = fields(c);
          fo.idx = dProbeIdx; fo.comps = dProbeComps; fo.port = dPort;
          fo.nAnt = nAnt; fo.ncomp = (int)probe_comps.size();
        }
        if (step_pol) {  // the step currents ride in the slab part's launch where they can
          fo.cur_comp = step_pol - 1;
          for (int b = 0; b < 6; ++b) fo.cur_box[b] = step_box[b];
          fo.cur_dx = delX; fo.cur_src = dSrc; fo.cur_dn = dStepN; fo.cur_q = q;
        }
        e = launch_step_fused(stream, f[c], slot[c]->dptr, slot[c]->host.data(), slot[c]->nwg, n0 + q, nullptr, &fo);
        c = 1 - c;
        if (e == hipSuccess && step_pol && !fo.cur_done)
          e = launch_add_currents(stream, g, fields(c), coefs(), step_pol - 1, step_box, 0.0f, delX, dSrc, dStepN, q);
      }
      if (e == hipSuccess) e = launch_advance(stream, dStepN, G);
      hipGraph_t graph = nullptr;
      const hipError_t e2 = hipStreamEndCapture(stream, &graph);
      if (e != hipSuccess || e2 != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        return fail(FDTD_ERR_DEVICE, std::string("graph recording failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
      }
      const hipError_t e3 = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (e3 != hipSuccess) {
        gexec = nullptr;
        return fail(FDTD_ERR_DEVICE, std::string("graph instantiation failed: ") + hipGetErrorString(e3));
      }
      gx_cur = cur; gx_args = args_gen; gx_steps = G; gx_currents = currents_gen; gx_mode = gmode; gx_probes = probes;
    }
    // the first step's sheets (unless the run before filled them), then the replays
    if (pw && !(rf.sheets_n == n0 && rf.sheets_set == cur)) HIPCHK(launch_sheets(stream, f[cur], n0));
    if (int rc = flush_probes(rf)) return rc;  // a record still pending from before
    const int idx[3] = {n0, rec_step, rec_step};
    HIPCHK(hipMemcpy(dStepN, idx, sizeof(idx), hipMemcpyHostToDevice));
    int q = 0;
    for (; q + G <= count; q += G) HIPCHK(hipGraphLaunch(gexec, stream));
    rec_step += q;
    *done = q;
    // what the last replayed step left for the steps after it: the next step's sheets (filled in
    // the current set's buffer) and its own record, still to be gathered
    rf.sheets_n = pw && n0 + q < T ? n0 + q : -1;
    rf.sheets_set = cur;
    rf.probe_row = probes && rec_step - 1 < T ? rec_step - 1 : -1;
    return 0;
  }
};


// ------------------------------------------------------------------------------------------
// Batched engine: many independent grids advanced together.  Every batch step launches each
// kernel of the fused step once for all the grids still running (face sheets, the row kernel's
// interior and slab parts per row-kernel variant, probes), so small grids fill the
// chip together and the launch count per step does not grow with the number of grids.  Grids
// whose layout the fused step does not cover run on their own afterwards.
// ------------------------------------------------------------------------------------------
struct fdtd_batch {
  struct Item {
    fdtd_sim *s;
    int n0, count, mode;
  };
  std::vector<Item> items;
  int strategy = 0;  // 0 automatic (merged), 1 merged launches, 2 one stream per grid, 3 device work queue
  int *dqueue = nullptr;  // per sub-batch stream: two slots of 8 (one per XCD) ticket counters
  // sub-batches run on their own streams (the GPU has four hardware queues per process), so the
  // kernels of one sub-batch overlap those of another
  static constexpr int kStreams = 4;
  hipStream_t streams[kStreams] = {};
  void *dtab = nullptr;
  size_t dtab_bytes = 0;
  std::vector<void *> dslab;  // per queued grid: slab argument blocks of sets 0 and 1
  ~fdtd_batch() {
    for (hipStream_t st : streams)
      if (st) {
        hipStreamSynchronize(st);
        hipStreamDestroy(st);
      }
    if (dtab) hipFree(dtab);
    if (dqueue) hipFree(dqueue);
    for (void *p : dslab) hipFree(p);
  }
};

namespace {

struct BatchPlan {
  // one epoch: the batch steps [t0, t1) with a fixed set of running grids
  struct Launch {
    int kind;  // 0 TF/SF face sheets, 1 row kernel part 0 (interior), 2 row kernel part 1 (slabs)
    RowsKey key;
    bool idx = false;
    size_t tab = 0;  // element offset of the table in the device buffer
    int ng = 0, nwg = 0;
  };
  struct Epoch {
    int t0 = 0, t1 = 0;
    std::vector<Launch> launches;
    size_t probes = 0;
    int nprobe = 0, max_ant = 0;
  };
  std::vector<Epoch> epochs;
};

}  // namespace

static int apply_tuning_class(fdtd_sim *s, int n, int mode);

// One stream per grid: every grid's steps are queued at once on its own stream and the GPU
// overlaps the grids as its hardware queues allow (the caller waits for the streams).
static int batch_queue_streams(const std::vector<fdtd_batch::Item> &items) {
  for (const auto &it : items) {
    if (!it.s->init_ready) return fail(FDTD_ERR_STATE, "batch: ReadySimulation has not been called");
    if (int rc = apply_tuning_class(it.s, it.n0, it.mode)) return rc;
    int slot = 0;
    for (int q = 0; q < it.count; ++q)
      if (int rc = it.s->step(it.n0 + q, it.mode, false, &slot)) return rc;
  }
  return 0;
}

static int batch_run_merged(fdtd_batch *b, const std::vector<fdtd_batch::Item> &items);

static int batch_run(fdtd_batch *b) {
  // strategy 2 runs every grid on its own stream; 0 and 1 merge the launches of every grid the
  // fused step covers.  The automatic choice is the merged form: on a mix of 32..268-cell grids it
  // beat both one stream per grid and a mix (grids of 2^18 cells and more on their own streams
  // beside merged small grids, which was slower than either, the own streams contending with
  // the sub-batch streams for the four hardware queues)
  std::vector<fdtd_batch::Item> big, small;
  for (const auto &it : b->items) (b->strategy == 2 ? big : small).push_back(it);
  if (int rc = batch_queue_streams(big)) return rc;
  const int rc = small.empty() ? 0 : batch_run_merged(b, small);
  for (const auto &it : big) HIPCHK(hipStreamSynchronize(it.s->stream));
  return rc;
}

static int batch_run_merged(fdtd_batch *b, const std::vector<fdtd_batch::Item> &items) {
  std::vector<fdtd_batch::Item> fused, alone;
  struct Prep {
    fdtd_batch::Item it;
    const void *rows_arg[2];
    int rows_nwg, rows_nwg1;
    RowsKey key;
    const void *slab_arg[2];
    unsigned nb_sheets;
    bool idx;
  };
  std::vector<Prep> preps;
  for (void *p : b->dslab) hipFree(p);
  b->dslab.clear();
  const size_t sab = slab_args_size();
  for (const auto &it : items) {
    fdtd_sim *s = it.s;
    if (!s->init_ready) return fail(FDTD_ERR_STATE, "batch: ReadySimulation has not been called");
    if (it.count <= 0) continue;
    const PWDev *pw = nullptr;
    if (int rc = s->check_pw(it.n0, it.mode, &pw)) return rc;
    if (pw && it.n0 + it.count > s->T) return fail(FDTD_ERR_INVALID, "batch: steps beyond the source length T");
    if (int rc = apply_tuning_class(s, it.n0, it.mode)) return rc;
    Prep pr;
    pr.it = it;
    bool ok = true;
    std::vector<char> h0(sab), h1(sab);
    for (int w = 0; w < 2 && ok; ++w) {
      FusedDev f;
      fdtd_sim::ArgSlot *slot = nullptr;
      if (int rc = s->fused_slot(w, it.mode, pw, f, &slot)) return rc;
      if (slot->nwg < 0 || s->step_pol || s->dChmu[0] || s->graded()) {  // per-op layouts, step currents, mu, graded meshes run alone
        ok = false;
        break;
      }
      pr.rows_arg[w] = slot->dptr;
      pr.rows_nwg = rows_part_nwg(slot->host.data(), 0);
      pr.rows_nwg1 = rows_part_nwg(slot->host.data(), 1);
      pr.key = rows_key(slot->host.data());
      make_slab_args(f, w ? h1.data() : h0.data(), &pr.nb_sheets);
      pr.idx = f.mat != nullptr;
    }
    if (!ok) {
      alone.push_back(it);
      continue;
    }
    void *d = nullptr;
    if (hipMalloc(&d, 2 * sab) != hipSuccess) return fail(FDTD_ERR_NOMEM, "batch: argument blocks");
    b->dslab.push_back(d);
    HIPCHK(hipMemcpyAsync(d, h0.data(), sab, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync((char *)d + sab, h1.data(), sab, hipMemcpyHostToDevice, s->stream));
    pr.slab_arg[0] = d;
    pr.slab_arg[1] = (char *)d + sab;
    preps.push_back(pr);
  }
  // the batch streams run after everything already queued on the grids' own streams
  for (int q = 0; q < fdtd_batch::kStreams; ++q)
    if (!b->streams[q]) HIPCHK(hipStreamCreateWithFlags(&b->streams[q], hipStreamNonBlocking));
  for (const auto &pr : preps) HIPCHK(hipStreamSynchronize(pr.it.s->stream));

  // sub-batches balanced by cell updates (largest first onto the least loaded)
  const int nsub = std::min<int>(fdtd_batch::kStreams, (int)preps.size());
  std::vector<std::vector<const Prep *>> subs(nsub);
  {
    std::vector<const Prep *> order;
    for (const auto &pr : preps) order.push_back(&pr);
    std::sort(order.begin(), order.end(), [](const Prep *a, const Prep *c) {
      const fdtd_sim *x = a->it.s, *y = c->it.s;
      return (double)x->IJK * a->it.count > (double)y->IJK * c->it.count;
    });
    std::vector<double> load(nsub, 0.0);
    for (const Prep *pr : order) {
      const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      subs[q].push_back(pr);
      load[q] += (double)pr->it.s->IJK * pr->it.count;
    }
  }
  std::vector<BatchGrid> tabs;
  std::vector<ProbeJob> jobs;
  std::vector<BatchPlan> plans(nsub);
  for (int q = 0; q < nsub; ++q) {
    // epochs: the running set changes only where a grid's step count ends
    std::vector<int> ends;
    for (const Prep *pr : subs[q]) ends.push_back(pr->it.count);
    std::sort(ends.begin(), ends.end());
    ends.erase(std::unique(ends.begin(), ends.end()), ends.end());
    int t0 = 0;
    for (int t1 : ends) {
      BatchPlan::Epoch ep;
      ep.t0 = t0;
      ep.t1 = t1;
      std::vector<const Prep *> run;
      for (const Prep *pr : subs[q])
        if (pr->it.count > t0) run.push_back(pr);
      auto add_table = [&](int kind, const std::vector<const Prep *> &grids, RowsKey key, bool idx) {
        BatchPlan::Launch L;
        L.kind = kind;
        L.key = key;
        L.idx = idx;
        L.tab = tabs.size();
        L.ng = (int)grids.size();
        int off = 0;
        for (const Prep *pr : grids) {
          BatchGrid bg;
          bg.off = off;
          bg.cur0 = pr->it.s->cur ^ (t0 & 1);
          bg.n0 = pr->it.n0 + t0;
          for (int w = 0; w < 2; ++w) bg.arg[w] = kind == 0 ? pr->slab_arg[w] : pr->rows_arg[w];
          tabs.push_back(bg);
          off += kind == 0 ? (int)pr->nb_sheets : kind == 1 ? pr->rows_nwg : pr->rows_nwg1;
        }
        BatchGrid end;
        end.off = off;
        tabs.push_back(end);
        L.nwg = off;
        ep.launches.push_back(L);
      };
      add_table(0, run, RowsKey(), false);
      std::map<RowsKey, std::vector<const Prep *>> by_key;
      for (const Prep *pr : run) by_key[pr->key].push_back(pr);
      for (auto &kv : by_key) add_table(1, kv.second, kv.first, false);
      for (auto &kv : by_key) add_table(2, kv.second, kv.first, false);
      ep.probes = jobs.size();
      for (const Prep *pr : run) {
        fdtd_sim *s = pr->it.s;
        if (!s->dPort) continue;
        ProbeJob jb;
        jb.F[0] = s->fields(0);
        jb.F[1] = s->fields(1);
        jb.idx = s->dProbeIdx;
        jb.comps = s->dProbeComps;
        jb.port = s->dPort;
        jb.nAnt = s->nAnt;
        jb.ncomp = (int)s->probe_comps.size();
        jb.row0 = s->rec_step + t0;
        jb.T = s->T;
        jb.cur0 = s->cur ^ (t0 & 1);
        jobs.push_back(jb);
        ep.max_ant = std::max(ep.max_ant, s->nAnt);
      }
      ep.nprobe = (int)(jobs.size() - ep.probes);
      plans[q].epochs.push_back(ep);
      t0 = t1;
    }
  }
  // one device buffer: all tables, then all probe jobs (uploaded on stream 0; the others wait)
  const size_t tab_bytes = tabs.size() * sizeof(BatchGrid), job_bytes = jobs.size() * sizeof(ProbeJob);
  if (tab_bytes + job_bytes > b->dtab_bytes) {
    if (b->dtab) {
      for (hipStream_t st : b->streams) HIPCHK(hipStreamSynchronize(st));
      hipFree(b->dtab);
    }
    b->dtab = nullptr;
    b->dtab_bytes = 0;
    if (hipMalloc(&b->dtab, tab_bytes + job_bytes) != hipSuccess) return fail(FDTD_ERR_NOMEM, "batch: tables");
    b->dtab_bytes = tab_bytes + job_bytes;
  }
  hipStream_t s0 = b->streams[0];
  const bool queue = b->strategy == 3;
  if (queue) {
    if (!b->dqueue && hipMalloc(&b->dqueue, 16 * fdtd_batch::kStreams * sizeof(int)) != hipSuccess)
      return fail(FDTD_ERR_NOMEM, "batch: work queues");
    HIPCHK(hipMemsetAsync(b->dqueue, 0, 16 * fdtd_batch::kStreams * sizeof(int), s0));
  }
  if (tab_bytes) HIPCHK(hipMemcpyAsync(b->dtab, tabs.data(), tab_bytes, hipMemcpyHostToDevice, s0));
  if (job_bytes) HIPCHK(hipMemcpyAsync((char *)b->dtab + tab_bytes, jobs.data(), job_bytes, hipMemcpyHostToDevice, s0));
  HIPCHK(hipStreamSynchronize(s0));
  const BatchGrid *dt = (const BatchGrid *)b->dtab;
  const ProbeJob *dj = (const ProbeJob *)((char *)b->dtab + tab_bytes);
  // interleave the sub-batches step by step so every stream stays fed
  std::vector<size_t> epi(nsub, 0);
  std::vector<int> qslot(nsub, 0);  // queue launches so far per stream (their slots alternate)
  int tmax = 0;
  for (const auto &pr : preps) tmax = std::max(tmax, pr.it.count);
  for (int t = 0; t < tmax; ++t) {
    for (int q = 0; q < nsub; ++q) {
      auto &eps = plans[q].epochs;
      while (epi[q] < eps.size() && t >= eps[epi[q]].t1) ++epi[q];
      if (epi[q] >= eps.size()) continue;
      const BatchPlan::Epoch &ep = eps[epi[q]];
      const int step = t - ep.t0;
      hipStream_t st = b->streams[q];
      for (const auto &L : ep.launches)
        if (L.kind == 0) HIPCHK(launch_sheets_batch(st, dt + L.tab, L.ng, L.nwg, step));
      for (const auto &L : ep.launches) {
        if (L.kind < 1) continue;
        if (queue) {
          if (L.nwg <= 0) continue;
          HIPCHK(launch_rows_queue(st, L.key, L.kind - 1, dt + L.tab, L.ng, L.nwg, step, b->dqueue + 16 * q, qslot[q] & 1,
                                   nsub));
          qslot[q]++;
        } else {
          HIPCHK(launch_rows_batch(st, L.key, L.kind - 1, dt + L.tab, L.ng, L.nwg, step));
        }
      }
      HIPCHK(launch_probes_batch(st, dj + ep.probes, ep.nprobe, ep.max_ant, step));
    }
  }
  for (int q = 0; q < nsub; ++q) HIPCHK(hipStreamSynchronize(b->streams[q]));
  for (const auto &pr : preps) {
    fdtd_sim *s = pr.it.s;
    s->cur ^= pr.it.count & 1;
    s->rec_step += pr.it.count;
  }
  // grids outside the fused layouts: their own steps
  for (const auto &it : alone) {
    int slot = 0;
    for (int q = 0; q < it.count; ++q)
      if (int rc = it.s->step(it.n0 + q, it.mode, false, &slot)) return rc;
  }
  return 0;
}

// ------------------------------------------------------------------------------------------
// Request server: solves arrive one at a time and are advanced together (continuous batching).
// A round admits queued requests in arrival order while the active slots and the HBM budget allow
// (ReadySimulation and Init1D run at admission, so a queued request holds no device memory),
// advances every active request by up to steps_per_round steps through the batched engine, and
// retires the requests that reached their last step.  Requests admitted in one round join the
// running ones at the next round boundary; each request's result is that of fdtd_run alone.
// ------------------------------------------------------------------------------------------
static void destroy_sim(fdtd_sim *s);

struct fdtd_server {
  struct Req {
    int64_t id = 0;
    fdtd_sim *s = nullptr;
    float rate = 0;
    std::vector<float> pw;
    int count = 0, mode = 0, done = 0;
    size_t bytes = 0;
    double t_arrive = 0, t_admit = -1, t_done = -1;
    int err = 0;          // FDTD_ERR_* of a failed admission (0: none)
    std::string err_msg;  // its message
  };
  size_t budget = 0, in_use = 0, peak = 0;
  int max_active = 0, steps_per_round = 0;
  int64_t next_id = 1;
  std::deque<Req *> pending;        // arrival order
  std::vector<Req *> active;
  std::deque<Req *> finished;       // completed, not yet collected
  std::map<int64_t, Req *> live;    // every request not yet released
  fdtd_batch batch;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  int64_t cell_updates = 0;

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

  ~fdtd_server() {
    for (auto &kv : live) {
      destroy_sim(kv.second->s);
      delete kv.second;
    }
  }

  int admit(Req *r) {
    fdtd_sim *s = r->s;
    if (int rc = s->ready()) return rc;
    if (r->mode)
      if (int rc = s->init_1d(r->rate, r->pw.data())) return rc;
    r->pw.clear();
    r->pw.shrink_to_fit();
    in_use += r->bytes;
    peak = std::max(peak, in_use);
    r->t_admit = now();
    active.push_back(r);
    return 0;
  }

  int round() {
    while (!pending.empty() && (int)active.size() < max_active && in_use + pending.front()->bytes <= budget) {
      Req *r = pending.front();
      pending.pop_front();
      if (int rc = admit(r)) {
        // a request whose set-up fails at admission (ReadySimulation or Init1D: a TF/SF inset too
        // large for its grid, device memory, ...) completes at once as failed: it is charged
        // nothing, collect returns it (admission time -1), fdtd_server_request_error reports the
        // error, and release frees what its set-up allocated
        r->err = rc;
        r->err_msg = fdtd_last_error();
        r->t_done = now();
        r->pw.clear();
        r->pw.shrink_to_fit();
        finished.push_back(r);
      }
    }
    if (active.empty()) return 0;
    batch.items.clear();
    for (Req *r : active) {
      const int k = std::min(steps_per_round, r->count - r->done);
      batch.items.push_back({r->s, r->done, k, r->mode});
    }
    if (int rc = batch_run(&batch)) return rc;
    const double t = now();
    std::vector<Req *> still;
    for (size_t q = 0; q < active.size(); ++q) {
      Req *r = active[q];
      const int k = batch.items[q].count;
      r->done += k;
      cell_updates += (int64_t)r->s->IJK * k;
      if (r->done >= r->count) {
        r->t_done = t;
        finished.push_back(r);
      } else {
        still.push_back(r);
      }
    }
    active.swap(still);
    batch.items.clear();
    return 0;
  }
};

// ------------------------------------------------------------------------------------------
// Live simulation handles.  The reference's handle is a raw pointer in a uint64 and a deleted one
// is undefined behaviour (src/simFDTD.cu:32-52); here every entry point checks that the handle
// was created by fdtd_create and not yet destroyed, and rejects it with simFDTD:InvalidInput
// otherwise.
// ------------------------------------------------------------------------------------------
namespace {
std::mutex g_live_mu;
std::set<const fdtd_sim *> g_live;

void register_sim(const fdtd_sim *s) {
  std::lock_guard<std::mutex> lk(g_live_mu);
  g_live.insert(s);
}
bool sim_live(const fdtd_sim *s) {
  std::lock_guard<std::mutex> lk(g_live_mu);
  return g_live.count(s) != 0;
}
}  // namespace

// Destroys a simulation and forgets its handle (fdtd_destroy, and the request server's release).
static void destroy_sim(fdtd_sim *s) {
  {
    std::lock_guard<std::mutex> lk(g_live_mu);
    g_live.erase(s);
  }
  delete s;
}

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------
#define NEED(s)                                                                                     \
  do {                                                                                              \
    if (!(s)) return fail(FDTD_ERR_INVALID, "null simulation handle");                              \
    if (!sim_live(s)) return fail(FDTD_ERR_INVALID, "invalid simulation handle (deleted or never created)"); \
  } while (0)
#define NEED_READY(s)                                                          \
  do {                                                                         \
    NEED(s);                                                                   \
    if (!(s)->init_ready) return fail(FDTD_ERR_STATE, "ReadySimulation has not been called"); \
  } while (0)

extern "C" {

const char *fdtd_last_error(void) { return g_err.c_str(); }
const char *fdtd_last_error_id(void) { return g_err_id.c_str(); }
const char *fdtd_version(void) { return "maxwell-mi355x fdtd 0.1 (gfx950)"; }

int fdtd_create(fdtd_sim **out) {
  if (!out) return fail(FDTD_ERR_INVALID, "null output pointer");
  *out = new fdtd_sim();
  register_sim(*out);
  return 0;
}

int fdtd_destroy(fdtd_sim *s) {
  NEED(s);
  destroy_sim(s);
  return 0;
}

int fdtd_set_const(fdtd_sim *s, const float freqRes[3], const int32_t timeDim[5]) {
  NEED(s);
  if (s->init_const) return fail(FDTD_ERR_ALREADY, "SetConst: already called.");
  if (!freqRes || !timeDim) return fail(FDTD_ERR_INVALID, "freqRes and TimeDim required.");
  if (timeDim[0] < 1 || timeDim[1] < 1 || timeDim[2] < 1 || timeDim[3] < 1 || timeDim[4] < 0)
    return fail(FDTD_ERR_INVALID, "TimeDim must be [T>0, I>0, J>0, K>0, nAnt>=0]");
  s->fctr = freqRes[0];
  s->delX = freqRes[1];
  s->delT = freqRes[2];
  s->T = timeDim[0];
  s->I = timeDim[1];
  s->J = timeDim[2];
  s->K = timeDim[3];
  s->nAnt = timeDim[4];
  s->IJ = (size_t)s->I * s->J;
  s->IJK = s->IJ * s->K;
  s->Ip = s->I;  // fixed by the set-up before the device arrays are allocated (prepare)
  s->IJp = s->IJ;
  s->IJKp = s->IJK;
  s->init_const = true;
  return 0;
}

int fdtd_set_cpml(fdtd_sim *s, const double p[6]) {
  NEED(s);
  if (!s->init_const || s->init_ready) return fail(FDTD_ERR_STATE, "SetCPML: call between SetConst and ReadySimulation");
  if (!p || p[0] < 2 || p[0] > 64) return fail(FDTD_ERR_INVALID, "SetCPML: 2 <= P <= 64");
  s->cp.P = (int)p[0];
  s->cp.poly_m = (float)p[1];
  s->cp.kappa_max = (float)p[2];
  s->cp.alpha_max = (float)p[3];
  s->cp.alpha_m = (float)p[4];
  s->cp.sig_factor = p[5];
  return 0;
}

int fdtd_set_mesh(fdtd_sim *s, const float *dx, const float *dy, const float *dz) {
  NEED(s);
  if (!s->init_const) return fail(FDTD_ERR_STATE, "SetMesh before SetConst");
  if (s->init_ready) return fail(FDTD_ERR_STATE, "SetMesh after ReadySimulation");
  const float *d[3] = {dx, dy, dz};
  const int Q[3] = {s->I, s->J, s->K};
  for (int a = 0; a < 3; ++a)
    if (d[a])
      for (int n = 0; n < Q[a]; ++n)
        if (!std::isfinite(d[a][n]) || !(d[a][n] > 0)) return fail(FDTD_ERR_INVALID, "SetMesh: spacings must be positive");
  for (int a = 0; a < 3; ++a) {
    if (d[a]) s->mesh[a].assign(d[a], d[a] + Q[a]);
    else s->mesh[a].clear();
  }
  s->args_gen++;
  return 0;
}

int fdtd_set_tfsf(fdtd_sim *s, int32_t inset) {
  NEED(s);
  if (s->init_ready) return fail(FDTD_ERR_STATE, "SetTFSF after ReadySimulation");
  if (inset < 1) return fail(FDTD_ERR_INVALID, "TF/SF inset must be >= 1");
  s->width = inset;
  return 0;
}

int fdtd_set_source(fdtd_sim *s, const float *src, int32_t update) {
  NEED(s);
  if (!s->init_const) return fail(FDTD_ERR_STATE, "SetSource before SetConst");
  if (!src) return fail(FDTD_ERR_INVALID, "SourceIn required.");
  if (update ? !s->init_source : s->init_source) return 0;  // src/simulation.h:1070-1074, :1085-1089
  s->source.assign(src, src + s->T);
  s->init_source = true;
  s->source_gen++;
  return 0;
}

int fdtd_set_step_currents(fdtd_sim *s, int32_t pol, const int32_t box[6]) {
  NEED(s);
  if (pol < 0 || pol > 3) return fail(FDTD_ERR_INVALID, "step currents: pol must be 0..3");
  if (pol && !box) return fail(FDTD_ERR_INVALID, "step currents: box required");
  if (pol && !s->init_source) return fail(FDTD_ERR_STATE, "step currents before SetSource");
  s->step_pol = pol;
  for (int q = 0; q < 6; ++q) s->step_box[q] = pol ? box[q] : 0;
  // a box that writes nothing (empty, or a negative start: src/gpu.h:55-82) turns them off
  if (pol && (box[1] < box[0] || box[3] < box[2] || box[5] < box[4] || box[0] < 0 || box[2] < 0 || box[4] < 0))
    s->step_pol = 0;
  s->currents_gen++;
  return 0;
}

int fdtd_set_graph_steps(fdtd_sim *s, int32_t steps) {
  NEED(s);
  if (steps < 0 || steps % 2) return fail(FDTD_ERR_INVALID, "graph steps must be an even count (0: off)");
  s->graph_steps = steps;
  return 0;
}

int fdtd_set_antenna_locations(fdtd_sim *s, const int32_t *loc) {
  NEED(s);
  if (!s->init_const) return fail(FDTD_ERR_STATE, "SetAntennaLocations before SetConst");
  if (!loc && s->nAnt > 0) return fail(FDTD_ERR_INVALID, "Locations required.");
  s->ant_loc.assign(loc, loc + 3 * (size_t)s->nAnt);
  s->init_antenna = s->nAnt > 0;
  if (s->init_ready) return s->alloc_probes();
  return 0;
}

int fdtd_set_probe_components(fdtd_sim *s, const int32_t *comps, int32_t n) {
  NEED(s);
  if (!comps || n < 1 || n > 6) return fail(FDTD_ERR_INVALID, "1..6 probe components");
  std::vector<int> v;
  for (int q = 0; q < n; ++q) {
    if (comps[q] < 1 || comps[q] > 6) return fail(FDTD_ERR_INVALID, "probe component must be 1..6");
    v.push_back(comps[q] - 1);
  }
  s->probe_comps = v;
  if (s->init_ready) return s->alloc_probes();
  return 0;
}

int fdtd_update_material(fdtd_sim *s, int32_t flag, const float *values) {
  NEED(s);
  if (!s->init_const) return fail(FDTD_ERR_STATE, "UpdateMaterial before SetConst");
  if (!values) return fail(FDTD_ERR_INVALID, "flag, values required.");
  if (flag < 1 || flag > 6) return 0;  // src/simulation.h:288-328 ignores other flags
  s->ensure_arrays();
  std::vector<float> &dst = flag <= 3 ? s->sig[flag - 1] : s->eps[flag - 4];
  std::memcpy(dst.data(), values, s->IJK * sizeof(float));
  return 0;
}

int fdtd_create_mu(fdtd_sim *s, const float *mu_x, const float *mu_y, const float *mu_z) {
  NEED(s);
  if (!s->init_const) return fail(FDTD_ERR_STATE, "CreateMu before SetConst");
  if (!mu_x || !mu_y || !mu_z) return fail(FDTD_ERR_INVALID, "mu_x, mu_y, mu_z required.");
  const float *m[3] = {mu_x, mu_y, mu_z};
  for (int c = 0; c < 3; ++c) {
    for (size_t id = 0; id < s->IJK; ++id)
      if (!(m[c][id] > 0.0f) || !std::isfinite(m[c][id])) return fail(FDTD_ERR_INVALID, "CreateMu: mu must be positive and finite");
  }
  for (int c = 0; c < 3; ++c) s->mu_in[c].assign(m[c], m[c] + s->IJK);
  if (s->init_ready) return s->upload_mu();
  return 0;
}

int fdtd_set_material_index(fdtd_sim *s, const uint8_t *index, const float *table, int32_t n) {
  NEED(s);
  if (!s->init_const) return fail(FDTD_ERR_STATE, "SetMaterialIndex before SetConst");
  if (!index || !table || n < 1 || n > 256) return fail(FDTD_ERR_INVALID, "SetMaterialIndex: index, table and 1 <= n <= 256 required");
  for (size_t id = 0; id < s->IJK; ++id)
    if (index[id] >= n) return fail(FDTD_ERR_INVALID, "material index beyond the table");
  s->midx.assign(index, index + s->IJK);
  s->mtab.assign(table, table + 6 * (size_t)n);
  s->nmat = n;
  s->have_index = true;
  s->have_arrays = false;
  for (int c = 0; c < 3; ++c) { s->eps[c].clear(); s->sig[c].clear(); }
  return 0;
}

int fdtd_set_plane_wave(fdtd_sim *s, int32_t mode, const double params[4]) {
  NEED(s);
  if (!s->init_const) return fail(FDTD_ERR_STATE, "SetPlaneWave before SetConst");
  if (mode < 1 || mode > 7 || !params) return fail(FDTD_ERR_INVALID, "plane-wave mode must be 1..7");
  s->pw_mode = mode;
  if (mode == 5) {
    s->pw_prm[0] = params[0]; s->pw_prm[1] = params[1]; s->pw_prm[2] = params[2];
    s->pw_layered = 0;
    if (s->prepared) {
      const int lo[3] = {s->g.i0, s->g.j0, s->g.k0}, hi[3] = {s->g.i1, s->g.j1, s->g.k1}, Q[3] = {s->I, s->J, s->K};
      s->obl = oblique_setup(params[0], params[1], params[2], lo, hi, Q, s->cp.P, s->delX, s->delT, s->fctr);
    }
  } else {
    s->pw_layered = params[0] != 0.0;
  }
  return 0;
}

int fdtd_prepare(fdtd_sim *s) {
  NEED(s);
  return s->prepare();
}

int fdtd_ready(fdtd_sim *s) {
  NEED(s);
  return s->ready();
}

int fdtd_init_1d(fdtd_sim *s, float rate, const float *pw) {
  NEED(s);
  if (!pw) return fail(FDTD_ERR_INVALID, "rate and PW_source required.");
  return s->init_1d(rate, pw);
}

int fdtd_update_coeffs(fdtd_sim *s) {
  NEED_READY(s);
  s->build_coefficients();
  return s->upload_coefficients();
}

int fdtd_reset_fields(fdtd_sim *s) {
  NEED_READY(s);
  return s->zero_state();
}

int fdtd_update_h_all(fdtd_sim *s) {
  NEED_READY(s);
  HIPCHK(launch_update_h_all(s->stream, s->g, s->fields(s->cur), s->coefs()));
  return 0;
}

int fdtd_update_e_all(fdtd_sim *s) {
  NEED_READY(s);
  HIPCHK(launch_update_e_all(s->stream, s->g, s->fields(s->cur), s->coefs()));
  return 0;
}

int fdtd_update_h_pml(fdtd_sim *s) {
  NEED_READY(s);
  HIPCHK(launch_update_h_pml(s->stream, s->g, s->fields(s->cur), s->psis(s->cur), s->coefs()));
  return 0;
}

int fdtd_update_e_pml(fdtd_sim *s) {
  NEED_READY(s);
  HIPCHK(launch_update_e_pml(s->stream, s->g, s->fields(s->cur), s->psis(s->cur), s->coefs()));
  return s->record_probes(s->cur);
}

int fdtd_update_h_planewave(fdtd_sim *s, const int32_t time[2]) {
  NEED_READY(s);
  if (!time) return fail(FDTD_ERR_INVALID, "time required.");
  const PWDev *pw = nullptr;
  if (int rc = s->check_pw(time[0], time[1], &pw)) return rc;
  if (pw) HIPCHK(launch_pw_h(s->stream, s->g, s->fields(s->cur), s->coefs(), *pw, time[0]));
  return 0;
}

int fdtd_update_e_planewave(fdtd_sim *s, const int32_t time[2]) {
  NEED_READY(s);
  if (!time) return fail(FDTD_ERR_INVALID, "time required.");
  const PWDev *pw = nullptr;
  if (int rc = s->check_pw(time[0], time[1], &pw)) return rc;
  if (pw) HIPCHK(launch_pw_e(s->stream, s->g, s->fields(s->cur), s->coefs(), *pw, time[0]));
  return 0;
}

int fdtd_add_currents(fdtd_sim *s, const int32_t time[2], const int32_t box[6]) {
  NEED_READY(s);
  if (!time || !box) return fail(FDTD_ERR_INVALID, "time and indices required.");
  if (!s->init_source) return fail(FDTD_ERR_STATE, "AddCurrents before SetSource");
  const int n = time[0], pol = time[1];
  if (n < 0 || n >= s->T) return fail(FDTD_ERR_INVALID, "time index outside [0, T)");
  if (pol < 1 || pol > 3) return 0;
  // an empty box or a negative start writes nothing (src/gpu.h:55-82; see current_box)
  if (box[1] < box[0] || box[3] < box[2] || box[5] < box[4] || box[0] < 0 || box[2] < 0 || box[4] < 0) return 0;
  const int b[6] = {box[0], box[1], box[2], box[3], box[4], box[5]};
  HIPCHK(launch_add_currents(s->stream, s->g, s->fields(s->cur), s->coefs(), pol - 1, b, s->source[n], s->delX));
  return 0;
}

int fdtd_step(fdtd_sim *s, int32_t n, int32_t mode) {
  NEED(s);
  int slot = 0;
  return s->step(n, mode, false, &slot);
}

int fdtd_batch_create(fdtd_batch **out) {
  if (!out) return fail(FDTD_ERR_INVALID, "null output pointer");
  *out = new fdtd_batch();
  return 0;
}

int fdtd_batch_destroy(fdtd_batch *b) {
  delete b;
  return 0;
}

int fdtd_batch_add(fdtd_batch *b, fdtd_sim *s, int32_t n0, int32_t count, int32_t mode) {
  if (!b) return fail(FDTD_ERR_INVALID, "null batch handle");
  NEED_READY(s);
  if (count < 0 || n0 < 0) return fail(FDTD_ERR_INVALID, "negative step index or count");
  for (const auto &it : b->items)
    if (it.s == s) return fail(FDTD_ERR_INVALID, "simulation already queued in this batch");
  b->items.push_back({s, n0, count, mode});
  return 0;
}

int fdtd_batch_clear(fdtd_batch *b) {
  if (!b) return fail(FDTD_ERR_INVALID, "null batch handle");
  b->items.clear();
  return 0;
}

int fdtd_batch_set_strategy(fdtd_batch *b, int32_t strategy) {
  if (!b) return fail(FDTD_ERR_INVALID, "null batch handle");
  if (strategy < 0 || strategy > 3) return fail(FDTD_ERR_INVALID, "strategy must be 0, 1, 2 or 3");
  b->strategy = strategy;
  return 0;
}

int fdtd_batch_run(fdtd_batch *b) {
  if (!b) return fail(FDTD_ERR_INVALID, "null batch handle");
  return batch_run(b);
}

int fdtd_server_create(fdtd_server **out, uint64_t hbm_budget, int32_t max_active, int32_t steps_per_round) {
  if (!out) return fail(FDTD_ERR_INVALID, "null output pointer");
  if (max_active < 1 || steps_per_round < 1 || hbm_budget == 0)
    return fail(FDTD_ERR_INVALID, "server: budget, slots and steps per round must be positive");
  fdtd_server *srv = new fdtd_server();
  srv->budget = (size_t)hbm_budget;
  srv->max_active = max_active;
  srv->steps_per_round = steps_per_round;
  *out = srv;
  return 0;
}

int fdtd_server_destroy(fdtd_server *srv) {
  delete srv;
  return 0;
}

int fdtd_server_now(fdtd_server *srv, double *t) {
  if (!srv || !t) return fail(FDTD_ERR_INVALID, "null server handle or output");
  *t = srv->now();
  return 0;
}

int fdtd_server_submit(fdtd_server *srv, fdtd_sim *s, float rate, const float *pw_source, int32_t count, int32_t mode,
                       double t_arrive, int64_t *id) {
  if (!srv || !id) return fail(FDTD_ERR_INVALID, "null server handle or output");
  NEED(s);
  if (s->init_ready) return fail(FDTD_ERR_STATE, "server: the simulation must not be ready yet (admission readies it)");
  if (!s->init_const) return fail(FDTD_ERR_STATE, "server: SetConst has not been called");
  if (count < 1) return fail(FDTD_ERR_INVALID, "server: step count must be positive");
  if (mode < 0 || mode > 5) return fail(FDTD_ERR_INVALID, "server: plane-wave mode must be 0..5");
  if (mode && !pw_source) return fail(FDTD_ERR_INVALID, "server: a plane-wave mode needs its source samples");
  if (mode && count > s->T) return fail(FDTD_ERR_INVALID, "server: steps beyond the source length T");
  if (mode == 5 && s->pw_mode != 5) return fail(FDTD_ERR_STATE, "server: mode 5 needs the oblique plane-wave set-up");
  for (const auto &kv : srv->live)
    if (kv.second->s == s) return fail(FDTD_ERR_INVALID, "server: simulation already submitted");
  auto *r = new fdtd_server::Req();
  r->id = srv->next_id++;
  r->s = s;
  r->rate = rate;
  if (mode) r->pw.assign(pw_source, pw_source + s->T);
  r->count = count;
  r->mode = mode;
  r->bytes = s->device_bytes_bound(mode);
  r->t_arrive = t_arrive >= 0 ? t_arrive : srv->now();
  if (r->bytes > srv->budget) {
    delete r;
    return fail(FDTD_ERR_NOMEM, "server: the request alone exceeds the HBM budget");
  }
  srv->pending.push_back(r);
  srv->live[r->id] = r;
  *id = r->id;
  return 0;
}

int fdtd_server_round(fdtd_server *srv, int32_t *active, int32_t *queued) {
  if (!srv) return fail(FDTD_ERR_INVALID, "null server handle");
  const int rc = srv->round();
  if (active) *active = (int32_t)srv->active.size();
  if (queued) *queued = (int32_t)srv->pending.size();
  return rc;
}

int fdtd_server_collect(fdtd_server *srv, int64_t *ids, double *times, int32_t max, int32_t *n) {
  if (!srv || !n || (max > 0 && (!ids || !times))) return fail(FDTD_ERR_INVALID, "null server handle or output");
  int k = 0;
  while (k < max && !srv->finished.empty()) {
    fdtd_server::Req *r = srv->finished.front();
    srv->finished.pop_front();
    ids[k] = r->id;
    times[3 * k] = r->t_arrive;
    times[3 * k + 1] = r->t_admit;
    times[3 * k + 2] = r->t_done;
    ++k;
  }
  *n = k;
  return 0;
}

int fdtd_server_release(fdtd_server *srv, int64_t id) {
  if (!srv) return fail(FDTD_ERR_INVALID, "null server handle");
  auto it = srv->live.find(id);
  if (it == srv->live.end()) return fail(FDTD_ERR_INVALID, "server: unknown request id");
  fdtd_server::Req *r = it->second;
  if (r->t_done < 0) return fail(FDTD_ERR_STATE, "server: request not completed");
  for (const auto &f : srv->finished)
    if (f == r) return fail(FDTD_ERR_STATE, "server: request not collected");
  destroy_sim(r->s);
  if (!r->err) srv->in_use -= r->bytes;  // a failed admission was never charged
  srv->live.erase(it);
  delete r;
  return 0;
}

int fdtd_server_request_error(fdtd_server *srv, int64_t id, int32_t *code, char *msg, int32_t msg_len) {
  if (!srv || !code) return fail(FDTD_ERR_INVALID, "null server handle or output");
  auto it = srv->live.find(id);
  if (it == srv->live.end()) return fail(FDTD_ERR_INVALID, "server: unknown request id");
  const fdtd_server::Req *r = it->second;
  *code = r->err;
  if (msg && msg_len > 0) {
    const size_t n = std::min(r->err_msg.size(), (size_t)msg_len - 1);
    std::memcpy(msg, r->err_msg.data(), n);
    msg[n] = 0;
  }
  return 0;
}

int fdtd_server_request_bytes(fdtd_server *srv, int64_t id, uint64_t *bytes) {
  if (!srv || !bytes) return fail(FDTD_ERR_INVALID, "null server handle or output");
  auto it = srv->live.find(id);
  if (it == srv->live.end()) return fail(FDTD_ERR_INVALID, "server: unknown request id");
  *bytes = it->second->bytes;
  return 0;
}

int fdtd_server_stats(fdtd_server *srv, uint64_t *bytes_in_use, uint64_t *bytes_peak, int64_t *cell_updates) {
  if (!srv) return fail(FDTD_ERR_INVALID, "null server handle");
  if (bytes_in_use) *bytes_in_use = srv->in_use;
  if (bytes_peak) *bytes_peak = srv->peak;
  if (cell_updates) *cell_updates = srv->cell_updates;
  return 0;
}

// ------------------------------------------------------------------------------------------
// Size-class tuning table: the autotuner's choices per class of grid, shared by the process's
// simulations and kept in a JSON file when FDTD_TUNE_CACHE names one.  A class is (I exactly: it
// fixes the row layouts; floor(log2 J) and floor(log2 K): the parallelism the z-chunk depth
// trades against; P; index-form materials; plane wave; per-cell mu).
// ------------------------------------------------------------------------------------------
namespace {
struct TuneClass {
  int I = 0, Jl = 0, Kl = 0, P = 0, idx = 0, pw = 0, mu = 0;
  bool operator<(const TuneClass &o) const {
    return std::tie(I, Jl, Kl, P, idx, pw, mu) < std::tie(o.I, o.Jl, o.Kl, o.P, o.idx, o.pw, o.mu);
  }
};
std::mutex g_tune_mu;
std::map<TuneClass, StepTuning> g_tune;
std::string g_tune_loaded = "\x01";  // the FDTD_TUNE_CACHE path last read (none yet)

int ilog2(int v) {
  int r = 0;
  while (v > 1) { v >>= 1; ++r; }
  return r;
}

TuneClass tune_class(const fdtd_sim *s, int mode) {
  TuneClass c;
  c.I = s->I; c.Jl = ilog2(s->J); c.Kl = ilog2(s->K); c.P = s->cp.P;
  c.idx = s->index_mode ? 1 : 0;
  c.pw = mode >= 1 && mode <= 7 ? 1 : 0;
  c.mu = s->dChmu[0] ? 1 : 0;
  return c;
}

std::string tune_path() {
  const char *p = std::getenv("FDTD_TUNE_CACHE");
  return p ? std::string(p) : std::string();
}

// The table shipped beside the library (tuning_gfx950.json in its directory), read when
// FDTD_TUNE_CACHE names no file of the caller's; never written.
std::string packaged_tune_path() {
  Dl_info info;
  if (!dladdr((const void *)&packaged_tune_path, &info) || !info.dli_fname) return std::string();
  std::string lib = info.dli_fname;
  const size_t slash = lib.find_last_of('/');
  return (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/tuning_gfx950.json";
}

// The file: {"classes": [{"I": 256, "log2J": 8, "log2K": 8, "P": 10, "index": 1, "plane_wave": 1,
// "mu": 0, "rows": 0, "z_chunk": 8, "cols_per_lane": 1, "slab_chunk": 0}, ...]} (integers only;
// an entry without "slab_chunk" keeps the slab part's automatic chunks).
void tune_load_locked() {
  std::string path = tune_path();
  if (path.empty()) path = packaged_tune_path();
  if (path == g_tune_loaded) return;  // read once per path (its entries join the table)
  g_tune_loaded = path;
  if (path.empty()) return;
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) return;
  std::string txt;
  char buf[4096];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) txt.append(buf, n);
  std::fclose(f);
  size_t pos = txt.find("\"classes\"");
  while (pos != std::string::npos) {
    const size_t a = txt.find('{', pos), b = a == std::string::npos ? a : txt.find('}', a);
    if (a == std::string::npos || b == std::string::npos) break;
    std::map<std::string, int> kv;
    size_t q = a;
    while (true) {
      const size_t k0 = txt.find('"', q + 1);
      if (k0 == std::string::npos || k0 > b) break;
      const size_t k1 = txt.find('"', k0 + 1);
      const size_t colon = txt.find(':', k1);
      if (k1 == std::string::npos || colon == std::string::npos || colon > b) break;
      kv[txt.substr(k0 + 1, k1 - k0 - 1)] = (int)std::strtol(txt.c_str() + colon + 1, nullptr, 10);
      q = txt.find_first_of(",}", colon);
      if (q == std::string::npos || q >= b) break;
    }
    TuneClass c;
    c.I = kv["I"]; c.Jl = kv["log2J"]; c.Kl = kv["log2K"]; c.P = kv["P"];
    c.idx = kv["index"]; c.pw = kv["plane_wave"]; c.mu = kv["mu"];
    StepTuning t;
    t.tile_rows = kv["rows"]; t.z_chunk = kv["z_chunk"]; t.slab_chunk = kv["slab_chunk"];
    const int cpl = kv["cols_per_lane"];
    t.cpl = cpl < 0 ? 1 : cpl;
    t.no_split = cpl < 0;
    if (c.I > 0) g_tune[c] = t;
    pos = b;
  }
}

void tune_save_locked() {
  const std::string path = tune_path();
  if (path.empty()) return;
  const std::string tmp = path + ".tmp";
  FILE *f = std::fopen(tmp.c_str(), "w");
  if (!f) return;
  std::fprintf(f, "{\"classes\": [");
  bool first = true;
  for (const auto &kv : g_tune) {
    const TuneClass &c = kv.first;
    const StepTuning &t = kv.second;
    std::fprintf(f, "%s\n  {\"I\": %d, \"log2J\": %d, \"log2K\": %d, \"P\": %d, \"index\": %d, \"plane_wave\": %d, "
                 "\"mu\": %d, \"rows\": %d, \"z_chunk\": %d, \"cols_per_lane\": %d, \"slab_chunk\": %d}",
                 first ? "" : ",", c.I, c.Jl, c.Kl, c.P, c.idx, c.pw, c.mu, t.tile_rows, t.z_chunk,
                 t.no_split ? -1 : t.cpl, t.slab_chunk);
    first = false;
  }
  std::fprintf(f, "\n]}\n");
  std::fclose(f);
  std::rename(tmp.c_str(), path.c_str());
}

// The autotuner: every columns-per-lane and z-chunk choice and two tile heights, each timed over
// `reps` dry steps in a run's launch form, then the slab part's chunk depth with the best of
// those; a choice replaces the automatic one only when it is faster by more than 1% (timing noise
// would otherwise pick between equals).
int autotune_sim(fdtd_sim *s, int n, int mode, int reps, double *best_ms) {
  const StepTuning saved = s->tune;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  double base = -1, best_t = -1;
  StepTuning best;
  int rc = 0;
  for (int rows : {0, 4})
    for (int cpl : {0, 1, 2, 4})
      for (int kc : {0, 4, 8, 16}) {
        s->tune = StepTuning();
        s->tune.tile_rows = rows;
        s->tune.cpl = cpl;
        s->tune.z_chunk = kc;
        s->args_gen++;
        double ms = 0;
        const int r = s->time_variant(n, mode, reps, e0, e1, &ms);
        if (r == FDTD_ERR_DEVICE) { rc = r; break; }
        if (r) continue;  // a layout this grid does not take
        if (rows == 0 && cpl == 0 && kc == 0) base = ms;
        if (best_t < 0 || ms < best_t) { best_t = ms; best = s->tune; }
      }
  if (base > 0 && best_t > 0.99 * base) { best = StepTuning(); best_t = base; }
  // the slab part's chunk depth (y-band chunks of the interior planes, z slabs cut to at most as
  // many planes) with the layout chosen above
  if (!rc && best_t > 0) {
    const StepTuning lay = best;
    const double lay_t = best_t;
    for (int sc : {3, 4, 6, 8, 12, 16}) {
      s->tune = lay;
      s->tune.slab_chunk = sc;
      s->args_gen++;
      double ms = 0;
      const int r = s->time_variant(n, mode, reps, e0, e1, &ms);
      if (r == FDTD_ERR_DEVICE) { rc = r; break; }
      if (r) continue;
      if (ms < 0.99 * lay_t && ms < best_t) { best_t = ms; best = s->tune; }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) { s->tune = saved; s->args_gen++; return rc; }
  if (best_t < 0) { s->tune = saved; s->args_gen++; return fail(FDTD_ERR_INVALID, "grid outside the fused step's layouts"); }
  s->tune = best;
  s->args_gen++;
  g_err.clear();
  if (best_ms) *best_ms = best_t;
  return 0;
}
}  // namespace

// The size-class table's tuning for a run in plane-wave mode `mode` (fdtd_run, the batched
// engine): applied when the simulation's tuning is automatic and not chosen by fdtd_set_tuning;
// with policy 2 a class not in the table is autotuned first (at time index n) and added.
static int apply_tuning_class(fdtd_sim *s, int n, int mode) {
  if (s->tune_policy == 0 || s->tune_user || !s->init_ready) return 0;
  const TuneClass c = tune_class(s, mode);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_load_locked();
  auto it = g_tune.find(c);
  if (it != g_tune.end()) {
    const StepTuning &t = it->second;
    if (t.tile_rows != s->tune.tile_rows || t.z_chunk != s->tune.z_chunk || t.cpl != s->tune.cpl ||
        t.no_split != s->tune.no_split || t.slab_chunk != s->tune.slab_chunk) {
      s->tune = t;
      s->args_gen++;
    }
    return 0;
  }
  if (s->tune_policy < 2) return 0;
  const PWDev *pw = nullptr;
  if (s->check_pw(n, mode, &pw)) return 0;  // the run reports a bad time index itself
  double ms = 0;
  if (int rc = autotune_sim(s, n, mode, 10, &ms)) return rc == FDTD_ERR_DEVICE ? rc : 0;
  g_tune[c] = s->tune;
  tune_save_locked();
  return 0;
}

int fdtd_autotune(fdtd_sim *s, int32_t n, int32_t mode, int32_t reps, double *best_ms) {
  NEED_READY(s);
  if (reps < 1) return fail(FDTD_ERR_INVALID, "reps must be >= 1");
  if (int rc = autotune_sim(s, n, mode, reps, best_ms)) return rc;
  s->tune_user = true;  // the caller asked for this simulation's own choice
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_load_locked();
  g_tune[tune_class(s, mode)] = s->tune;
  tune_save_locked();
  return 0;
}

int fdtd_set_autotune(fdtd_sim *s, int32_t policy) {
  NEED(s);
  if (policy < 0 || policy > 2) return fail(FDTD_ERR_INVALID, "autotune policy must be 0, 1 or 2");
  s->tune_policy = policy;
  return 0;
}

int fdtd_run(fdtd_sim *s, int32_t n0, int32_t count, int32_t mode, int32_t snap_every, int32_t snap_comp) {
  NEED_READY(s);
  if (count < 0) return fail(FDTD_ERR_INVALID, "negative step count");
  if (snap_every > 0 && (snap_comp < 1 || snap_comp > 6)) return fail(FDTD_ERR_INVALID, "snapshot component must be 1..6");
  if (int rc = s->ensure_second()) return rc;
  if (count > 0)
    if (int rc = apply_tuning_class(s, n0, mode)) return rc;
  // snapshot buffers are allocated before the loop so the loop itself allocates nothing
  s->nsnap = 0;
  if (snap_every > 0) {
    const int need = (s->rec_step + count) / snap_every - s->rec_step / snap_every;
    while ((int)s->dSnaps.size() < need) {
      void *p = nullptr;
      hipError_t e = hipMalloc(&p, s->IJKp * sizeof(float));
      if (e != hipSuccess) return fail(FDTD_ERR_NOMEM, "snapshot buffer allocation failed");
      s->dSnaps.push_back((float *)p);
    }
  }
  const int every = s->timing;  // time steps q with q % every == 0 (0: none)
  const int ntimed = every > 0 ? (count + every - 1) / every : 0;
  if (every > 0) {
    while ((int)s->ev.size() < 4 * ntimed) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      s->ev.push_back(e);
    }
  }
  int slot = 0, q0 = 0;
  fdtd_sim::RunFold rf;
  if (snap_every <= 0)
    if (int rc = s->run_graphs(n0, count, mode, &q0, rf)) return rc;
  for (int q = q0; q < count; ++q) {
    if (int rc = s->step(n0 + q, mode, every > 0 && q % every == 0, &slot, false, &rf)) {
      s->flush_probes(rf);
      return rc;
    }
    if (snap_every > 0 && s->rec_step % snap_every == 0) {
      HIPCHK(hipMemcpyAsync(s->dSnaps[s->nsnap], s->dF[s->cur][snap_comp - 1], s->IJKp * sizeof(float),
                            hipMemcpyDeviceToDevice, s->stream));
      s->nsnap++;
    }
  }
  if (int rc = s->flush_probes(rf)) return rc;
  if (ntimed > 0) {
    HIPCHK(hipStreamSynchronize(s->stream));
    double tot = 0, rows = 0;
    for (int q = 0; q < ntimed; ++q) {
      float ms = 0, rms = 0;
      HIPCHK(hipEventElapsedTime(&ms, s->ev[4 * q], s->ev[4 * q + 3]));
      HIPCHK(hipEventElapsedTime(&rms, s->ev[4 * q + 1], s->ev[4 * q + 2]));
      tot += ms;
      rows += rms;
    }
    s->kernel_ms = tot;
    s->rows_ms = rows;
    s->kernel_launches = ntimed;
  }
  return 0;
}

int fdtd_run_noalloc(fdtd_sim *s, int32_t n0, int32_t count, int32_t mode) {
  NEED_READY(s);
  if (count < 0) return fail(FDTD_ERR_INVALID, "negative step count");
  if (s->na) return fail(FDTD_ERR_STATE, "an allocation-free run of this simulation is in progress");
  const PWDev *pw = nullptr;
  if (count > 0)
    if (int rc = s->check_pw(n0, mode, &pw)) return rc;
  if (pw && s->sheet_floats() * sizeof(float) > kScratchSheetBytes)
    return fail(FDTD_ERR_INVALID, "run_noalloc: the TF/SF face sheets of this grid (" +
                                      std::to_string(s->sheet_floats() * sizeof(float)) +
                                      " bytes) exceed the scratch's " + std::to_string(kScratchSheetBytes) +
                                      " bytes per buffer");
  static_assert(kScratchArgBytes >= 1024, "argument block scratch too small");
  if (step_args_size() > kScratchArgBytes) return fail(FDTD_ERR_INVALID, "run_noalloc: argument block too large");
  int prev = 0;
  HIPCHK(hipGetDevice(&prev));
  if (prev != s->device) HIPCHK(hipSetDevice(s->device));
  void *base = nullptr;
  const hipError_t e = scratch_region(&base);
  if (e != hipSuccess) {
    if (prev != s->device) hipSetDevice(prev);
    return fail(FDTD_ERR_DEVICE, std::string("scratch_region: ") + hipGetErrorString(e));
  }
  char *b = (char *)base;
  fdtd_sim::NoAlloc na{{(float *)b, (float *)(b + kScratchSheetBytes)},
                       {b + 2 * kScratchSheetBytes, b + 2 * kScratchSheetBytes + kScratchArgBytes},
                       {false, false},
                       {-1, -1}};
  s->na = &na;
  fdtd_sim::RunFold rf;
  int slot = 0, rc = 0;
  for (int q = 0; q < count && !rc; ++q) rc = s->step(n0 + q, mode, false, &slot, false, &rf);
  const int rc2 = s->flush_probes(rf);
  s->na = nullptr;
  if (prev != s->device) hipSetDevice(prev);
  return rc ? rc : rc2;
}

int fdtd_snapshot_count(fdtd_sim *s, int32_t *count) {
  NEED(s);
  if (!count) return fail(FDTD_ERR_INVALID, "null count");
  *count = s->nsnap;
  return 0;
}

int fdtd_output_snapshots(fdtd_sim *s, float *out, uint64_t size) {
  NEED_READY(s);
  if (!out || size < (uint64_t)s->nsnap * s->IJK) return fail(FDTD_ERR_INVALID, "snapshot output too small");
  for (int q = 0; q < s->nsnap; ++q)
    HIPCHK(s->get_cells(out + (size_t)q * s->IJK, s->dSnaps[q]));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int fdtd_stream_ceiling(fdtd_sim *s, int32_t reps, double *ms) {
  NEED_READY(s);
  if (reps < 1 || !ms) return fail(FDTD_ERR_INVALID, "stream ceiling: reps >= 1 and an output required");
  if (!s->dMat) return fail(FDTD_ERR_STATE, "stream ceiling: needs the uint8 material index form");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  hipError_t e = launch_stream(s->stream, s->dFb[s->cur], s->dFb[1 - s->cur], s->dMat, s->cs, s->IJKp);  // warm
  if (e == hipSuccess) e = hipEventRecord(e0, s->stream);
  for (int q = 0; q < reps && e == hipSuccess; ++q)
    e = launch_stream(s->stream, s->dFb[(s->cur + q) & 1], s->dFb[(s->cur + q + 1) & 1], s->dMat, s->cs, s->IJKp);
  if (e == hipSuccess) e = hipEventRecord(e1, s->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (e != hipSuccess) return fail(FDTD_ERR_DEVICE, std::string("stream ceiling: ") + hipGetErrorString(e));
  *ms = t / reps;
  return 0;
}

int fdtd_device_bytes(fdtd_sim *s, uint64_t *bytes) {
  NEED(s);
  if (!bytes) return fail(FDTD_ERR_INVALID, "null output");
  *bytes = s->dev_bytes + s->dSnaps.size() * s->IJKp * sizeof(float);
  return 0;
}

int fdtd_synchronize(fdtd_sim *s) {
  NEED(s);
  if (s->stream) HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int fdtd_output_ant(fdtd_sim *s, float *out, uint32_t size) {
  NEED(s);
  if (!out) return fail(FDTD_ERR_INVALID, "OutAntSize required.");
  if (!s->init_ready || !s->dPort) {
    // src/simulation.h:634-637: without antennas the reference prints this and leaves the output
    // as it is (the gateway's freshly created, zero-filled array)
    std::fprintf(stderr, "Antenna data is not initialized.\n");
    return 0;
  }
  const size_t n = std::min((size_t)size, s->port_len);
  HIPCHK(hipMemcpyAsync(out, s->dPort, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int fdtd_output_vis(fdtd_sim *s, float *out, uint32_t size, int32_t flag) {
  NEED_READY(s);
  if (!out) return fail(FDTD_ERR_INVALID, "OutAntSize and flag required.");
  if (size < s->IJK) return fail(FDTD_ERR_INVALID, "OutputVis: size smaller than I*J*K");
  if (flag >= 1 && flag <= 6) {
    HIPCHK(s->get_cells(out, s->dF[s->cur][flag - 1]));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
  }
  if (flag <= 0 && flag >= -2) {  // host sigma arrays, src/simulation.h:574-588
    const int c = -flag;
    if (s->have_arrays) std::memcpy(out, s->sig[c].data(), s->IJK * sizeof(float));
    else if (s->have_index)
      for (size_t id = 0; id < s->IJK; ++id) out[id] = s->mtab[6 * s->midx[id] + c];
    else std::memset(out, 0, s->IJK * sizeof(float));
    return 0;
  }
  return 0;  // other flags leave the output untouched, as the reference
}

int fdtd_read_array(fdtd_sim *s, int32_t what, float *out, uint64_t *count) {
  NEED_READY(s);
  const float *src = nullptr;
  size_t n = 0;
  std::vector<float> tmp;
  bool host = false;
  if (what >= 0 && what < 6) { src = s->dF[s->cur][what]; n = s->IJK; }
  else if (what >= 6 && what < 12) {
    const int c = (what - 6) % 3;
    const bool isCh = what >= 9;
    tmp.resize(s->IJK);
    for (size_t id = 0; id < s->IJK; ++id) {
      float ce, ch;
      s->cell_coeff(id, c, &ce, &ch);
      tmp[id] = isCh ? ch : ce;
    }
    host = true; n = s->IJK;
  } else if (what >= 20 && what < 32) { src = s->dPsi[s->cur][what - 20]; n = s->psi_len[what - 20]; }
  else if (what >= 40 && what < 44) {
    src = what == 40 ? s->dE1h : what == 41 ? s->dH1h : what == 42 ? s->dEinv : s->dHinv;
    if (!src) return fail(FDTD_ERR_STATE, "line history not initialised");
    n = (size_t)s->T * s->N1;
  } else {
    return fail(FDTD_ERR_INVALID, "unknown array id");
  }
  if (count) *count = n;
  if (!out) return 0;
  if (host) { std::memcpy(out, tmp.data(), n * sizeof(float)); return 0; }
  if (what >= 0 && what < 6) {  // a field component: rows of Ip on the device
    HIPCHK(s->get_cells(out, src));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
  }
  if (what >= 20 && what < 32) {  // psi: every 4th float of its interleaved group (one linear copy)
    std::vector<float> grp(4 * n);
    HIPCHK(hipMemcpyAsync(grp.data(), src, (4 * n - 3) * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (size_t t = 0; t < n; ++t) out[t] = grp[4 * t];
    return 0;
  }
  HIPCHK(hipMemcpyAsync(out, src, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int fdtd_read_setup(fdtd_sim *s, int32_t what, float *out, uint64_t *count) {
  NEED(s);
  if (!s->prepared) return fail(FDTD_ERR_STATE, "set-up not prepared");
  std::vector<float> v;
  if (what == 0) v = s->prof.bE;
  else if (what == 1) v = s->prof.cE;
  else if (what == 2) v = s->prof.bH;
  else if (what == 3) v = s->prof.cH;
  else if (what == 4) v = s->prof.kapE;
  else if (what == 5) v = s->prof.kapH;
  else if (what >= 6 && what < 9) v = s->ikE[what - 6];
  else if (what >= 9 && what < 12) v = s->ikH[what - 9];
  else if (what == 12) {
    const int n1 = s->pw_mode == 5 ? s->obl.N1 : s->N1;
    v = {s->Chh, s->Che, s->prof.sig_max, (float)n1, (float)s->g.i0, (float)s->g.i1, (float)s->g.j0,
         (float)s->g.j1, (float)s->g.k0, (float)s->g.k1, (float)s->cp.P, (float)s->obl.ratio, (float)s->obl.src};
  } else if (what >= 13 && what < 19) {
    const int c = (what - 13) % 3;
    const bool isCh = what >= 16;
    v.resize(s->IJK);
    for (size_t id = 0; id < s->IJK; ++id) {
      float ce, ch;
      s->cell_coeff(id, c, &ce, &ch);
      v[id] = isCh ? ch : ce;
    }
  } else if (what == 19) {
    v.resize(s->IJK);
    for (size_t id = 0; id < s->IJK; ++id) v[id] = s->index_mode ? (float)s->cell_mat[id] : -1.0f;
  } else if (what >= 20 && what < 32) {
    const int u = (what - 20) / 4, q = (what - 20) % 4;
    v = q < 2 ? s->obl.tabE[u][q] : s->obl.tabH[u][q - 2];
  } else if (what == 32) {
    v = {s->obl.polE[0], s->obl.polE[1], s->obl.polE[2], s->obl.polH[0], s->obl.polH[1], s->obl.polH[2]};
  } else {
    return fail(FDTD_ERR_INVALID, "unknown set-up array id");
  }
  if (count) *count = v.size();
  if (out) std::memcpy(out, v.data(), v.size() * sizeof(float));
  return 0;
}

int fdtd_set_tuning(fdtd_sim *s, int32_t rows_per_group, int32_t z_chunk, int32_t cols_per_lane) {
  NEED(s);
  if (rows_per_group != 0 && rows_per_group != 1 && rows_per_group != 2 && rows_per_group != 4 && rows_per_group != 8 &&
      rows_per_group != 16)
    return fail(FDTD_ERR_INVALID, "rows_per_group must be 0, 1, 2, 4, 8 or 16");
  if (z_chunk < 0) return fail(FDTD_ERR_INVALID, "z_chunk must be >= 0");
  if (cols_per_lane != 0 && cols_per_lane != 1 && cols_per_lane != 2 && cols_per_lane != 4 && cols_per_lane != -1)
    return fail(FDTD_ERR_INVALID, "cols_per_lane must be 0, 1, 2, 4 or -1");
  s->tune.tile_rows = rows_per_group;
  s->tune.z_chunk = z_chunk;
  s->tune_user = rows_per_group != 0 || z_chunk != 0 || cols_per_lane != 0 || s->tune.slab_chunk != 0;
  s->tune.cpl = cols_per_lane < 0 ? 1 : cols_per_lane;
  s->tune.no_split = cols_per_lane < 0;
  s->args_gen++;
  return 0;
}

int fdtd_get_tuning(fdtd_sim *s, int32_t out[4]) {
  NEED(s);
  if (!out) return fail(FDTD_ERR_INVALID, "null output");
  out[0] = s->tune.tile_rows;
  out[1] = s->tune.z_chunk;
  out[2] = s->tune.no_split ? -1 : s->tune.cpl;
  out[3] = s->tune.slab_chunk;
  return 0;
}

int fdtd_set_addressing(fdtd_sim *s, int32_t mode) {
  NEED(s);
  if (mode != 0 && mode != 1) return fail(FDTD_ERR_INVALID, "addressing mode must be 0 (automatic) or 1 (global)");
  s->addr_global = mode == 1;
  s->args_gen++;
  return 0;
}

int fdtd_set_slab_chunk(fdtd_sim *s, int32_t planes) {
  NEED(s);
  if (planes < 0) return fail(FDTD_ERR_INVALID, "slab chunk must be >= 0");
  s->tune.slab_chunk = planes;
  if (planes) s->tune_user = true;
  s->args_gen++;
  return 0;
}

int fdtd_set_timing(fdtd_sim *s, int32_t every) {
  NEED(s);
  if (every < 0) return fail(FDTD_ERR_INVALID, "negative timing interval");
  s->timing = every;
  return 0;
}

int fdtd_kernel_time(fdtd_sim *s, double *avg_ms, int64_t *launches) {
  NEED(s);
  if (avg_ms) *avg_ms = s->kernel_launches ? s->kernel_ms / (double)s->kernel_launches : 0.0;
  if (launches) *launches = s->kernel_launches;
  return 0;
}

int fdtd_row_kernel_time(fdtd_sim *s, double *avg_ms) {
  NEED(s);
  if (avg_ms) *avg_ms = s->kernel_launches ? s->rows_ms / (double)s->kernel_launches : 0.0;
  return 0;
}

}  // extern "C"
