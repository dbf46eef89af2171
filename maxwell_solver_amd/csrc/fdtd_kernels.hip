// fdtd_kernels.hip -- gfx950 kernels of the 3D Yee FDTD library.
//
// Two families:
//  * per-op kernels, one per op of the reference MEX gateway, updating the state in place with
//    the reference's masks and operation order (src/gpu.h);
//  * the fused time step over a ping-ponged state (read the current set, write the next):
//    k_sheets (the TF/SF face sheets of the step), then the row kernel, H and E of every cell in
//    one pass: one wave per aligned row segment marching up a chunk of planes, all CPML slabs and
//    TF/SF corrections inline, in two launches (the interior rows and planes; the slab rows and
//    planes).  Each field is read once and written once per step outside the recomputed halo.
//    In a run of steps the slab launch also fills the next step's sheets and gathers the probes,
//    so a step is two launches.
//
// All arithmetic is plain IEEE float with no contraction (-ffp-contract=off); expressions keep the
// reference's association order so results equal the CPU oracle bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "fdtd_kernels.hpp"

namespace fdtd {

// ------------------------------------------------------------------------------------------
// shared device helpers
// ------------------------------------------------------------------------------------------

// Linear interpolation of a 1D-line history row at fractional node s (exact lookup when s is an
// integer, as for the axis-aligned modes).
__device__ __forceinline__ float line_at(const float *row, float s) {
  const float fl = floorf(s);
  const int m = (int)fl;
  const float w = s - fl;
  if (w == 0.0f) return row[m];
  return row[m] + w * (row[m + 1] - row[m]);
}

// Incident E component c at the Yee position of cell (x,y,z): Ex (h,i,i), Ey (i,h,i), Ez (i,i,h).
__device__ __forceinline__ float inc_e(const PWDev &pw, const float *row, int c, int x, int y, int z) {
  const float sx = pw.tabE[0][c == 0][x];
  const float sy = pw.tabE[1][c == 1][y];
  const float sz = pw.tabE[2][c == 2][z];
  return pw.polE[c] * line_at(row, (sx + sy) + sz);
}
// Incident H component c: Hx (i,h,h), Hy (h,i,h), Hz (h,h,i).
__device__ __forceinline__ float inc_h(const PWDev &pw, const float *row, int c, int x, int y, int z) {
  const float sx = pw.tabH[0][c != 0][x];
  const float sy = pw.tabH[1][c != 1][y];
  const float sz = pw.tabH[2][c != 2][z];
  return pw.polH[c] * line_at(row, (sx + sy) + sz);
}

// TF/SF corrections of an H cell on the SF side of face axis FA (generic form of the reference's
// mode kernels src/gpu.h:152-573).  lo: the cell is at (lo face - 1), else at the high face.
template <int FA>
__device__ __forceinline__ void pw_h_face(const GridDev &g, float Che, const PWDev &pw, const float *row, int i, int j,
                                          int k, bool lo, float &hx, float &hy, float &hz) {
  const int m = pw.active_h;
  if (FA == 0) {
    const int xe = lo ? g.i0 : g.i1;
    if ((m & (1 << 1)) && j >= g.j0 && j <= g.j1 && k >= g.k0 && k <= g.k1 - 1) {
      const float v = Che * inc_e(pw, row, 2, xe, j, k);
      hy = lo ? hy - v : hy + v;
    }
    if ((m & (1 << 2)) && j >= g.j0 && j <= g.j1 - 1 && k >= g.k0 && k <= g.k1) {
      const float v = Che * inc_e(pw, row, 1, xe, j, k);
      hz = lo ? hz + v : hz - v;
    }
  } else if (FA == 1) {
    const int ye = lo ? g.j0 : g.j1;
    if ((m & (1 << 3)) && i >= g.i0 && i <= g.i1 && k >= g.k0 && k <= g.k1 - 1) {
      const float v = Che * inc_e(pw, row, 2, i, ye, k);
      hx = lo ? hx + v : hx - v;
    }
    if ((m & (1 << 5)) && i >= g.i0 && i <= g.i1 - 1 && k >= g.k0 && k <= g.k1) {
      const float v = Che * inc_e(pw, row, 0, i, ye, k);
      hz = lo ? hz - v : hz + v;
    }
  } else {
    const int ze = lo ? g.k0 : g.k1;
    if ((m & (1 << 6)) && i >= g.i0 && i <= g.i1 && j >= g.j0 && j <= g.j1 - 1) {
      const float v = Che * inc_e(pw, row, 1, i, j, ze);
      hx = lo ? hx - v : hx + v;
    }
    if ((m & (1 << 7)) && i >= g.i0 && i <= g.i1 - 1 && j >= g.j0 && j <= g.j1) {
      const float v = Che * inc_e(pw, row, 0, i, j, ze);
      hy = lo ? hy + v : hy - v;
    }
  }
}

// TF/SF corrections of an E cell on face axis FA (lo: on the low face, else the high face).
template <int FA>
__device__ __forceinline__ void pw_e_face(const GridDev &g, const float ch[3], const PWDev &pw, const float *row, int i,
                                          int j, int k, bool lo, float &ex, float &ey, float &ez) {
  const int m = pw.active_e;
  if (FA == 0) {
    const int xh = lo ? g.i0 - 1 : g.i1;
    if ((m & (1 << 2)) && j >= g.j0 && j <= g.j1 && k >= g.k0 && k <= g.k1 - 1) {
      const float v = ch[2] * inc_h(pw, row, 1, xh, j, k);
      ez = lo ? ez - v : ez + v;
    }
    if ((m & (1 << 1)) && j >= g.j0 && j <= g.j1 - 1 && k >= g.k0 && k <= g.k1) {
      const float v = ch[1] * inc_h(pw, row, 2, xh, j, k);
      ey = lo ? ey + v : ey - v;
    }
  } else if (FA == 1) {
    const int yh = lo ? g.j0 - 1 : g.j1;
    if ((m & (1 << 3)) && i >= g.i0 && i <= g.i1 - 1 && k >= g.k0 && k <= g.k1) {
      const float v = ch[0] * inc_h(pw, row, 2, i, yh, k);
      ex = lo ? ex - v : ex + v;
    }
    if ((m & (1 << 5)) && i >= g.i0 && i <= g.i1 && k >= g.k0 && k <= g.k1 - 1) {
      const float v = ch[2] * inc_h(pw, row, 0, i, yh, k);
      ez = lo ? ez + v : ez - v;
    }
  } else {
    const int zh = lo ? g.k0 - 1 : g.k1;
    if ((m & (1 << 6)) && i >= g.i0 && i <= g.i1 - 1 && j >= g.j0 && j <= g.j1) {
      const float v = ch[0] * inc_h(pw, row, 1, i, j, zh);
      ex = lo ? ex + v : ex - v;
    }
    if ((m & (1 << 7)) && i >= g.i0 && i <= g.i1 && j >= g.j0 && j <= g.j1 - 1) {
      const float v = ch[1] * inc_h(pw, row, 0, i, j, zh);
      ey = lo ? ey - v : ey + v;
    }
  }
}

// Material coefficients of cell id (column i).  Index mode: table row of the cell's material,
// with Ch = 0 on the last x plane (den_ex = 0 there, src/simulation.h:887-912).
// row pitch of the per-cell arrays (GridDev::Ip; 0 means I)
__host__ __device__ __forceinline__ size_t gpitch(const GridDev &g) { return (size_t)(g.Ip > 0 ? g.Ip : g.I); }

template <bool IDX>
__device__ __forceinline__ void cell_coef(const CoefDev &co, const float *tab, size_t id, int i, int I, float ce[3],
                                          float ch[3]) {
  if (IDX) {
    const float *r = tab + 6 * (int)co.mat[id];
    ce[0] = r[0]; ce[1] = r[1]; ce[2] = r[2];
    const bool last = (i == I - 1);
    ch[0] = last ? 0.0f : r[3];
    ch[1] = last ? 0.0f : r[4];
    ch[2] = last ? 0.0f : r[5];
  } else {
    ce[0] = co.Ce[0][id]; ce[1] = co.Ce[1][id]; ce[2] = co.Ce[2][id];
    ch[0] = co.Ch[0][id]; ch[1] = co.Ch[1][id]; ch[2] = co.Ch[2][id];
  }
}

// ------------------------------------------------------------------------------------------
// per-op kernels (in place)
// ------------------------------------------------------------------------------------------

// K1 GPU_updateH, src/gpu.h:575-602; with per-cell mu (co.Chmu set) K2 GPU_updateH_Full,
// src/gpu.h:4-30: per-cell coefficients and no Chxh factor
__global__ void __launch_bounds__(256) k_h_all(GridDev g, FieldSet F, CoefDev co) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= g.I - 1 || j >= g.J - 1 || k >= g.K - 1) return;
  const size_t I = gpitch(g), IJ = I * g.J;  // strides of the per-cell arrays
  const size_t id = (size_t)i + (size_t)j * I + (size_t)k * IJ;
  const float *Ex = F.f[EX], *Ey = F.f[EY], *Ez = F.f[EZ];
  float *Hx = F.f[HX], *Hy = F.f[HY], *Hz = F.f[HZ];
  const float ez = Ez[id], ey = Ey[id], ex = Ex[id];
  const float kx = co.ikH[0][i], ky = co.ikH[1][j], kz = co.ikH[2][k];
  if (co.Chmu[0]) {
    Hx[id] = Hx[id] + co.Chmu[0][id] * (-(Ez[id + I] - ez) * ky + (Ey[id + IJ] - ey) * kz);
    Hy[id] = Hy[id] + co.Chmu[1][id] * (-(Ex[id + IJ] - ex) * kz + (Ez[id + 1] - ez) * kx);
    Hz[id] = Hz[id] + co.Chmu[2][id] * (-(Ey[id + 1] - ey) * kx + (Ex[id + I] - ex) * ky);
    return;
  }
  Hx[id] = co.Chh * Hx[id] + co.Che * (-(Ez[id + I] - ez) * ky + (Ey[id + IJ] - ey) * kz);
  Hy[id] = co.Chh * Hy[id] + co.Che * (-(Ex[id + IJ] - ex) * kz + (Ez[id + 1] - ez) * kx);
  Hz[id] = co.Chh * Hz[id] + co.Che * (-(Ey[id + 1] - ey) * kx + (Ex[id + I] - ex) * ky);
}

// K3 GPU_updateE, src/gpu.h:727-763
template <bool IDX>
__global__ void __launch_bounds__(256) k_e_all(GridDev g, FieldSet F, CoefDev co) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= g.I || j >= g.J) return;
  const size_t I = gpitch(g), IJ = I * g.J;  // strides of the per-cell arrays
  const size_t id = (size_t)i + (size_t)j * I + (size_t)k * IJ;
  float *Ex = F.f[EX], *Ey = F.f[EY], *Ez = F.f[EZ];
  const float *Hx = F.f[HX], *Hy = F.f[HY], *Hz = F.f[HZ];
  float ce[3], ch[3];
  cell_coef<IDX>(co, co.tab, id, i, g.I, ce, ch);
  const float hx = Hx[id], hy = Hy[id], hz = Hz[id];
  const float kx = co.ikE[0][i], ky = co.ikE[1][j], kz = co.ikE[2][k];
  if (j != 0 && k != 0) Ex[id] = ce[0] * Ex[id] + ch[0] * ((hz - Hz[id - I]) * ky - (hy - Hy[id - IJ]) * kz);
  if (i != 0 && k != 0) Ey[id] = ce[1] * Ey[id] + ch[1] * ((hx - Hx[id - IJ]) * kz - (hz - Hz[id - 1]) * kx);
  if (i != 0 && j != 0) Ez[id] = ce[2] * Ez[id] + ch[2] * ((hy - Hy[id - 1]) * kx - (hx - Hx[id - I]) * ky);
}

// K5 GPU_updateHPmlY, src/gpu.h:644-685: i < I-1, s in [0, 2P-1), k < K-1
__global__ void __launch_bounds__(256) k_h_pml_y(GridDev g, FieldSet F, PsiSet ps, CoefDev co) {
  const int i = blockIdx.x * 64 + threadIdx.x, s = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  const int I = g.I, J = g.J, P = g.P;
  if (i >= I - 1 || s >= 2 * P - 1) return;
  const size_t Ip = gpitch(g), IJp = Ip * J;
  const size_t t = (size_t)i + (size_t)I * s + (size_t)I * 2 * P * k;
  const int jj = s < P ? s : s + J - 2 * P;
  const size_t id = (size_t)i + (size_t)jj * Ip + (size_t)k * IJp;
  const float dx = co.dx;
  float *pxy = ps.p[P_HXY], *pzy = ps.p[P_HZY];
  pxy[4 * t] = co.bH[s] * pxy[4 * t] + co.cH[s] * (F.f[EZ][id + Ip] - F.f[EZ][id]) / dx;
  F.f[HX][id] = F.f[HX][id] - co.Che * pxy[4 * t] * dx;
  pzy[4 * t] = co.bH[s] * pzy[4 * t] + co.cH[s] * (F.f[EX][id + Ip] - F.f[EX][id]) / dx;
  F.f[HZ][id] = F.f[HZ][id] + co.Che * pzy[4 * t] * dx;
}

// K4 GPU_updateHPmlX, src/gpu.h:604-642: s in [0, 2P-1), j < J-1, k < K-1
__global__ void __launch_bounds__(256) k_h_pml_x(GridDev g, FieldSet F, PsiSet ps, CoefDev co) {
  const int I = g.I, J = g.J, P = g.P;
  const int q = blockIdx.x * 256 + threadIdx.x, k = blockIdx.y;
  const int ns = 2 * P - 1;
  if (q >= ns * (J - 1)) return;
  const int s = q % ns, j = q / ns;
  const size_t Ip = gpitch(g), IJp = Ip * J;
  const size_t t = (size_t)s + (size_t)2 * P * j + (size_t)2 * P * J * k;
  const int ii = s < P ? s : I - 2 * P + s;
  const size_t id = (size_t)ii + (size_t)j * Ip + (size_t)k * IJp;
  const float dx = co.dx;
  float *pyx = ps.p[P_HYX], *pzx = ps.p[P_HZX];
  pyx[4 * t] = co.bH[s] * pyx[4 * t] + co.cH[s] * (F.f[EZ][id + 1] - F.f[EZ][id]) / dx;
  F.f[HY][id] = F.f[HY][id] + co.Che * pyx[4 * t] * dx;
  pzx[4 * t] = co.bH[s] * pzx[4 * t] + co.cH[s] * (F.f[EY][id + 1] - F.f[EY][id]) / dx;
  F.f[HZ][id] = F.f[HZ][id] - co.Che * pzx[4 * t] * dx;
}

// K6 GPU_updateHPmlZ, src/gpu.h:687-725: i < I-1, j < J-1, s in [0, 2P-1)
__global__ void __launch_bounds__(256) k_h_pml_z(GridDev g, FieldSet F, PsiSet ps, CoefDev co) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, s = blockIdx.z;
  const int I = g.I, J = g.J, K = g.K, P = g.P;
  if (i >= I - 1 || j >= J - 1) return;
  const size_t IJ = (size_t)I * J, Ip = gpitch(g), IJp = Ip * J;
  const size_t t = (size_t)i + (size_t)I * j + IJ * s;
  const int kk = s < P ? s : s + K - 2 * P;
  const size_t id = (size_t)i + (size_t)j * Ip + (size_t)kk * IJp;
  const float dx = co.dx;
  float *pxz = ps.p[P_HXZ], *pyz = ps.p[P_HYZ];
  pxz[4 * t] = co.bH[s] * pxz[4 * t] + co.cH[s] * (F.f[EY][id + IJp] - F.f[EY][id]) / dx;
  F.f[HX][id] = F.f[HX][id] + co.Che * pxz[4 * t] * dx;
  pyz[4 * t] = co.bH[s] * pyz[4 * t] + co.cH[s] * (F.f[EX][id + IJp] - F.f[EX][id]) / dx;
  F.f[HY][id] = F.f[HY][id] - co.Che * pyz[4 * t] * dx;
}

// K8 GPU_updateEPmlY, src/gpu.h:790-827: 0 < i < I-1, s in [1, 2P), 1 <= k < K-1
template <bool IDX>
__global__ void __launch_bounds__(256) k_e_pml_y(GridDev g, FieldSet F, PsiSet ps, CoefDev co) {
  const int i = blockIdx.x * 64 + threadIdx.x, s = blockIdx.y * 4 + threadIdx.y + 1, k = blockIdx.z + 1;
  const int I = g.I, J = g.J, P = g.P;
  if (i < 1 || i >= I - 1 || s >= 2 * P) return;
  const size_t Ip = gpitch(g), IJp = Ip * J;
  const size_t t = (size_t)i + (size_t)I * s + (size_t)I * 2 * P * k;
  const int jj = s < P ? s : s + J - 2 * P;
  const size_t id = (size_t)i + (size_t)jj * Ip + (size_t)k * IJp;
  float ce[3], ch[3];
  cell_coef<IDX>(co, co.tab, id, i, I, ce, ch);
  const float dx = co.dx;
  float *pxy = ps.p[P_EXY], *pzy = ps.p[P_EZY];
  pxy[4 * t] = co.bE[s] * pxy[4 * t] + co.cE[s] * (F.f[HZ][id] - F.f[HZ][id - Ip]) / dx;
  F.f[EX][id] = F.f[EX][id] + ch[0] * pxy[4 * t] * dx;
  pzy[4 * t] = co.bE[s] * pzy[4 * t] + co.cE[s] * (F.f[HX][id] - F.f[HX][id - Ip]) / dx;
  F.f[EZ][id] = F.f[EZ][id] - ch[2] * pzy[4 * t] * dx;
}

// K9 GPU_updateEPmlZ, src/gpu.h:829-868: 0 < i < I-1, 0 < j < J-1, s in [1, 2P)
template <bool IDX>
__global__ void __launch_bounds__(256) k_e_pml_z(GridDev g, FieldSet F, PsiSet ps, CoefDev co) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, s = blockIdx.z + 1;
  const int I = g.I, J = g.J, K = g.K, P = g.P;
  if (i < 1 || i >= I - 1 || j < 1 || j >= J - 1) return;
  const size_t IJ = (size_t)I * J, Ip = gpitch(g), IJp = Ip * J;
  const size_t t = (size_t)i + (size_t)I * j + IJ * s;
  const int kk = s < P ? s : s + K - 2 * P;
  const size_t id = (size_t)i + (size_t)j * Ip + (size_t)kk * IJp;
  float ce[3], ch[3];
  cell_coef<IDX>(co, co.tab, id, i, I, ce, ch);
  const float dx = co.dx;
  float *pxz = ps.p[P_EXZ], *pyz = ps.p[P_EYZ];
  pxz[4 * t] = co.bE[s] * pxz[4 * t] + co.cE[s] * (F.f[HY][id] - F.f[HY][id - IJp]) / dx;
  F.f[EX][id] = F.f[EX][id] - ch[0] * pxz[4 * t] * dx;
  pyz[4 * t] = co.bE[s] * pyz[4 * t] + co.cE[s] * (F.f[HX][id] - F.f[HX][id - IJp]) / dx;
  F.f[EY][id] = F.f[EY][id] + ch[1] * pyz[4 * t] * dx;
}

// K7 GPU_updateEPmlX, src/gpu.h:765-788: s in [1, 2P), 0 < j < J-1, 1 <= k < K-1
template <bool IDX>
__global__ void __launch_bounds__(256) k_e_pml_x(GridDev g, FieldSet F, PsiSet ps, CoefDev co) {
  const int I = g.I, J = g.J, P = g.P;
  const int q = blockIdx.x * 256 + threadIdx.x, k = blockIdx.y + 1;
  const int ns = 2 * P - 1;  // s = 1 .. 2P-1
  if (q >= ns * (J - 2)) return;
  const int s = q % ns + 1, j = q / ns + 1;
  const size_t Ip = gpitch(g), IJp = Ip * J;
  const size_t t = (size_t)s + (size_t)2 * P * j + (size_t)2 * P * J * k;
  const int ii = s < P ? s : s + I - 2 * P;
  const size_t id = (size_t)ii + (size_t)j * Ip + (size_t)k * IJp;
  float ce[3], ch[3];
  cell_coef<IDX>(co, co.tab, id, ii, I, ce, ch);
  const float dx = co.dx;
  float *pyx = ps.p[P_EYX], *pzx = ps.p[P_EZX];
  pyx[4 * t] = co.bE[s] * pyx[4 * t] + co.cE[s] * (F.f[HZ][id] - F.f[HZ][id - 1]) / dx;
  F.f[EY][id] = F.f[EY][id] - ch[1] * pyx[4 * t] * dx;
  pzx[4 * t] = co.bE[s] * pzx[4 * t] + co.cE[s] * (F.f[HY][id] - F.f[HY][id - 1]) / dx;
  F.f[EZ][id] = F.f[EZ][id] + ch[2] * pzx[4 * t] * dx;
}

// TF/SF corrections of face axis FA, both faces (one kernel per face axis, launched x, y, z).
// Threads cover the face's tangential box range [lo, hi] of the two other axes.
template <int FA>
__global__ void __launch_bounds__(256) k_pw_h(GridDev g, FieldSet F, CoefDev co, PWDev pw, int n) {
  const int u = blockIdx.x * 16 + threadIdx.x, v = blockIdx.y * 16 + threadIdx.y;
  int i, j, k;
  const float *row = pw.Eh + (size_t)n * pw.N1;
  const size_t I = gpitch(g), IJ = I * g.J;  // strides of the per-cell arrays
  for (int side = 0; side < 2; ++side) {
    const bool lo = side == 0;
    if (FA == 0) { i = lo ? g.i0 - 1 : g.i1; j = g.j0 + u; k = g.k0 + v; if (j > g.j1 || k > g.k1) return; }
    else if (FA == 1) { j = lo ? g.j0 - 1 : g.j1; i = g.i0 + u; k = g.k0 + v; if (i > g.i1 || k > g.k1) return; }
    else { k = lo ? g.k0 - 1 : g.k1; i = g.i0 + u; j = g.j0 + v; if (i > g.i1 || j > g.j1) return; }
    const size_t id = (size_t)i + (size_t)j * I + (size_t)k * IJ;
    float hx = F.f[HX][id], hy = F.f[HY][id], hz = F.f[HZ][id];
    pw_h_face<FA>(g, co.Che, pw, row, i, j, k, lo, hx, hy, hz);
    F.f[HX][id] = hx; F.f[HY][id] = hy; F.f[HZ][id] = hz;
  }
}

template <int FA, bool IDX>
__global__ void __launch_bounds__(256) k_pw_e(GridDev g, FieldSet F, CoefDev co, PWDev pw, int n) {
  const int u = blockIdx.x * 16 + threadIdx.x, v = blockIdx.y * 16 + threadIdx.y;
  int i, j, k;
  const float *row = pw.Hh + (size_t)n * pw.N1;
  const size_t I = gpitch(g), IJ = I * g.J;  // strides of the per-cell arrays
  for (int side = 0; side < 2; ++side) {
    const bool lo = side == 0;
    if (FA == 0) { i = lo ? g.i0 : g.i1; j = g.j0 + u; k = g.k0 + v; if (j > g.j1 || k > g.k1) return; }
    else if (FA == 1) { j = lo ? g.j0 : g.j1; i = g.i0 + u; k = g.k0 + v; if (i > g.i1 || k > g.k1) return; }
    else { k = lo ? g.k0 : g.k1; i = g.i0 + u; j = g.j0 + v; if (i > g.i1 || j > g.j1) return; }
    const size_t id = (size_t)i + (size_t)j * I + (size_t)k * IJ;
    float ce[3], ch[3];
    cell_coef<IDX>(co, co.tab, id, i, g.I, ce, ch);
    float ex = F.f[EX][id], ey = F.f[EY][id], ez = F.f[EZ][id];
    pw_e_face<FA>(g, ch, pw, row, i, j, k, lo, ex, ey, ez);
    F.f[EX][id] = ex; F.f[EY][id] = ey; F.f[EZ][id] = ez;
  }
}

// K17 AddSource_GPU_3D_1DThread, src/gpu.h:55-82 (box already clipped to the grid)
template <bool IDX>
// The source sample is s, or, for launches recorded into a graph that is replayed for later
// steps, src[*dn + q]: the step index of the replay's first step (*dn) plus this launch's offset.
__global__ void __launch_bounds__(256) k_add_currents(GridDev g, FieldSet F, CoefDev co, int comp, int a0, int b0,
                                                      int c0, int ni, int nj, int nk, float s, float dx,
                                                      const float *src, const int *dn, int q) {
  const long tid = (long)blockIdx.x * 256 + threadIdx.x;
  if (tid >= (long)ni * nj * nk) return;
  if (src) s = src[*dn + q];
  const int i = a0 + (int)(tid % ni), j = b0 + (int)((tid / ni) % nj), k = c0 + (int)(tid / ((long)ni * nj));
  const size_t id = (size_t)i + (size_t)j * gpitch(g) + (size_t)k * gpitch(g) * g.J;
  float ce[3], ch[3];
  cell_coef<IDX>(co, co.tab, id, i, g.I, ce, ch);
  float *E = F.f[comp];
  E[id] = E[id] - ch[comp] * s * dx / dx / dx;
}

// Advances the indices read by graph-recorded launches by the steps of one replay: dn[0] the time
// index, dn[1] the probe record row of the replay's first step (dn[2], the first record row of the
// run, stays).
__global__ void k_advance(int *dn, int steps) {
  if (threadIdx.x == 0) {
    dn[0] = dn[0] + steps;
    dn[1] = dn[1] + steps;
  }
}

// 1D incident line for all T steps in one workgroup (src/simulation.h:198-244 with the kernels
// src/gpu.h:40-53, 84-149): history rows E^n (before the step) and H^{n+1/2} (after it).
__global__ void __launch_bounds__(1024) k_line(LineDev L) {
  extern __shared__ float sm[];
  float *E = sm, *H = sm + L.N;
  __shared__ float pE[128], pH[128];
  const int t = threadIdx.x, N = L.N, P = L.P;
  for (int i = t; i < N; i += 1024) { E[i] = 0.0f; H[i] = 0.0f; }
  if (t < 2 * P) { pE[t] = 0.0f; pH[t] = 0.0f; }
  __syncthreads();
  for (int n = 0; n < L.T; ++n) {
    float *Eh = L.Eh + (size_t)n * N, *Hh = L.Hh + (size_t)n * N;
    for (int i = t; i < N; i += 1024) Eh[i] = E[i];
    for (int i = t; i < N - 1; i += 1024) H[i] = H[i] + L.Ch * ((E[i + 1] - E[i]) * L.ikH[i]);
    __syncthreads();
    if (t < 2 * P) {
      const bool c1 = t < P, c2 = (t < 2 * P - 1) && (t >= P);
      if (c1 || c2) {
        const int idx = c1 ? t : t + N - 2 * P;
        pH[t] = L.bH[t] * pH[t] + L.cH[t] * (E[idx + 1] - E[idx]) / L.ds;
        H[idx] = H[idx] + L.Ch * pH[t] * L.ds;
      }
    }
    __syncthreads();
    for (int i = t; i < N; i += 1024)
      if (i > 0) E[i] = L.Ce[i] * E[i] + L.Cb[i] * ((H[i] - H[i - 1]) * L.ikE[i]);
    __syncthreads();
    if (t < 2 * P) {
      const bool c1 = (t > 0) && (t < P), c2 = (t < 2 * P) && (t >= P);
      if (c1 || c2) {
        const int idx = c1 ? t : t + N - 2 * P;
        pE[t] = L.bE[t] * pE[t] + L.cE[t] * (H[idx] - H[idx - 1]) / L.ds;
        E[idx] = E[idx] + L.Cb[idx] * pE[t] * L.ds;
      }
    }
    __syncthreads();
    if (t == 0) E[L.src] = E[L.src] + L.pw[n];
    __syncthreads();
    for (int i = t; i < N; i += 1024) Hh[i] = H[i];
  }
}

__global__ void k_probes(FieldSet F, const int *idx, int nAnt, const int *comps, int ncomp, float *port, int row, int T) {
  const int a = blockIdx.x * 64 + threadIdx.x, q = blockIdx.y;
  if (a >= nAnt || q >= ncomp) return;
  port[((size_t)q * T + row) * nAnt + a] = F.f[comps[q]][idx[a]];
}

// ------------------------------------------------------------------------------------------
// fused time step, reading the cur state and writing the nxt state
//
//  k_sheets   the incident values the six faces of the TF/SF box need at step n (tfsf_sheet).
//  k_rows     every cell: one wave per aligned row segment of 64 CPL columns (CPL consecutive
//             columns per lane), marching k up a chunk of planes; fused H and E with the TF/SF
//             corrections and the CPML of all three axes inline.  Part 0 (k_rows_c1 for one
//             column per lane): the rows j in [P+1, J-P) of the planes k in [P, K-P); part 1: the
//             other rows and planes (y-slab rows and the row above the low slab, z-slab planes).
// Per cell both apply exactly the operations of the per-op kernels, in the same order.
// ------------------------------------------------------------------------------------------
struct StepArgs {
  int I, J, K, P;
  int i0, i1, j0, j1, k0, k1;
  int S, R, CPL, nTy;  // row segments per row, rows per workgroup (slab part), columns per lane, row tiles
  int R0, nTmid0;      // rows per workgroup and row tiles of the interior part
  int KC, ntab, act_h, act_e;
  int KCS, nz1, nz2;  // z chunks: nz1 of KCS planes over [0, P), nz2 of KC over [P, K-P), nz1 over [K-P, K)
  int KCY, nzy;       // the slab part's y-band rows over [P, K-P): nzy chunks of KCY planes
  int Wrow;           // LDS row stride of the row exchange: 64 CPL S columns
  int nTlo, nThi, nTmid;    // row tiles of the bands [0, P], [J-P, J) and [P+1, J-P)
  int nwg0, nwg1;           // workgroups of the two parts (host information)
  int xni, xnj, xnk;  // TF/SF box extents (i1-i0+1, j1-j0+1, k1-k0+1)
  int spl;            // split edge segment (1 column per lane, Ip a multiple of 64; see rows_body)
  int ntf;            // no TF/SF work in the slab part (the box lies clear of its rows and planes)
  int glob;           // 64-bit global addressing although every array fits a buffer resource
  int gr;             // graded mesh: every row and plane applies the per-axis table factors
  int Ip;             // row pitch of the per-cell arrays (>= I)
  size_t IJK, sx, sy, sz;  // IJK: Ip*J*K, the extent of a per-cell array
  size_t cs;  // distance between the six field components of a set (Ip*J*K plus a skew)
  const float *Fc;
  float *Fn;
  const float *Pc;
  float *Pn;
  const uint8_t *mat;
  const float *coef;   // index mode: [ntab][6]; full mode: [6][IJK]
  const float *prm;    // bE cE bH cH [2P each], 1/kappa E x,y,z, 1/kappa H x,y,z
  const float *xs;      // TF/SF face sheets (tfsf_sheet)
  float Chh, Che, dx;
  const float *mu;      // per-cell H coefficients of CreateMu [3][IJK] (K2), or null
};

// Pointers taken from the argument block are generic; casting them to the global address space
// lets the compiler emit global_load/store (counted by vmcnt only) instead of flat accesses,
// which also count against lgkmcnt.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gptr(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}

// Workgroup barrier ordering LDS only: global loads in flight are not waited for, unlike
// __syncthreads(), whose fence also covers global memory.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Workgroups are dealt to the 8 XCDs round-robin; remap so that each XCD walks a contiguous range
// of tiles (neighbouring row tiles of one z chunk share their halo rows in that XCD's L2).
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int per = nb / 8, rem = nb % 8;
  const int x = b % 8, q = b / 8;
  return x < rem ? x * (per + 1) + q : rem * (per + 1) + (x - rem) * per + q;
}

// Addressing of the row kernel.  Every access is (array, uniform offset u, lane offset v), both in
// bytes.  Buffer mode (each array under 4 GiB): a buffer resource per array and 32-bit offsets,
// u in a scalar register (soffset) and v in one vector register (voffset), so that the many
// row / plane / component neighbours of a cell cost no 64-bit address registers; out-of-range
// buffer accesses read 0 and drop stores.  Global mode: 64-bit pointer arithmetic.
#define AS1 __attribute__((address_space(1)))
#define AS4 __attribute__((address_space(4)))
typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
typedef unsigned int u2v __attribute__((ext_vector_type(2)));

// third dword of a gfx9-family buffer resource: 32-bit data format, raw (untyped) access
constexpr int kBufWord3 = 0x00020000;

template <bool BUF>
struct Acc;

template <>
struct Acc<true> {
  using O = uint32_t;
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ Acc(const void *base, size_t bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)(uint32_t)bytes, kBufWord3)) {}
  template <int CPL>
  __device__ __forceinline__ void ld(O v, O u, float (&x)[CPL]) const {
    if constexpr (CPL == 4) {
      const f4v t = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(r, (int)v, (int)u, 0));
      x[0] = t.x; x[1] = t.y; x[2] = t.z; x[3] = t.w;
    } else if constexpr (CPL == 2) {
      const f2v t = __builtin_bit_cast(f2v, __builtin_amdgcn_raw_buffer_load_b64(r, (int)v, (int)u, 0));
      x[0] = t.x; x[1] = t.y;
    } else {
      x[0] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)v, (int)u, 0));
    }
  }
  __device__ __forceinline__ float ld1(O v, O u) const {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)v, (int)u, 0));
  }
  // Store cache policy: NT (non-temporal) or write-back.  Nothing of the new state is read again
  // within a step; when the grid's state is larger than the 256 MB Infinity Cache, streaming it
  // past the L2 keeps the halo rows the neighbouring tiles re-read resident (256^3: interior part
  // 177.0 -> 172.9 us, slab part 67.6 -> 64.4 us); when the state fits, write-back lets the next
  // step read it from the Infinity Cache (250x280x50: 87.1 -> 71.3 us per step).
  template <int CPL, bool NT = true>
  __device__ __forceinline__ void st(O v, O u, const float (&x)[CPL]) const {
    constexpr int aux = NT ? 2 : 0;
    if constexpr (CPL == 4) {
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4v, f4v{x[0], x[1], x[2], x[3]}), r, (int)v, (int)u, aux);
    } else if constexpr (CPL == 2) {
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2v, f2v{x[0], x[1]}), r, (int)v, (int)u, aux);
    } else {
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x[0]), r, (int)v, (int)u, aux);
    }
  }
  __device__ __forceinline__ void st1(O v, O u, float x) const {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, (int)v, (int)u, 2);
  }
  // CPL consecutive bytes at byte offset v + u
  template <int CPL>
  __device__ __forceinline__ void ldm(O v, O u, int (&m)[CPL]) const {
    if constexpr (CPL == 4) {
      const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(r, (int)v, (int)u, 0);
      m[0] = w & 255; m[1] = (w >> 8) & 255; m[2] = (w >> 16) & 255; m[3] = w >> 24;
    } else if constexpr (CPL == 2) {
      const uint32_t w = __builtin_amdgcn_raw_buffer_load_b16(r, (int)v, (int)u, 0);
      m[0] = w & 255; m[1] = w >> 8;
    } else {
      m[0] = __builtin_amdgcn_raw_buffer_load_b8(r, (int)v, (int)u, 0);
    }
  }
};

template <>
struct Acc<false> {
  using O = uint64_t;
  AS1 char *p;
  __device__ __forceinline__ Acc(const void *base, size_t) : p((AS1 char *)const_cast<void *>(base)) {}
  template <int CPL>
  __device__ __forceinline__ void ld(O v, O u, float (&x)[CPL]) const {
    const AS1 char *q = p + u + v;
    if constexpr (CPL == 4) {
      const f4v t = *(const AS1 f4v *)q;
      x[0] = t.x; x[1] = t.y; x[2] = t.z; x[3] = t.w;
    } else if constexpr (CPL == 2) {
      const f2v t = *(const AS1 f2v *)q;
      x[0] = t.x; x[1] = t.y;
    } else {
      x[0] = *(const AS1 float *)q;
    }
  }
  __device__ __forceinline__ float ld1(O v, O u) const { return *(const AS1 float *)(p + u + v); }
  // non-temporal stores (arrays over 4 GiB: the state is far larger than the Infinity Cache)
  template <int CPL, bool NT = true>
  __device__ __forceinline__ void st(O v, O u, const float (&x)[CPL]) const {
    AS1 char *q = p + u + v;
    if constexpr (CPL == 4) {
      __builtin_nontemporal_store(f4v{x[0], x[1], x[2], x[3]}, (AS1 f4v *)q);
    } else if constexpr (CPL == 2) {
      __builtin_nontemporal_store(f2v{x[0], x[1]}, (AS1 f2v *)q);
    } else {
      __builtin_nontemporal_store(x[0], (AS1 float *)q);
    }
  }
  __device__ __forceinline__ void st1(O v, O u, float x) const { __builtin_nontemporal_store(x, (AS1 float *)(p + u + v)); }
  template <int CPL>
  __device__ __forceinline__ void ldm(O v, O u, int (&m)[CPL]) const {
    const AS1 uint8_t *q = (const AS1 uint8_t *)(p + u + v);
    if constexpr (CPL == 4) {
      const uint32_t w = *(const AS1 uint32_t *)q;
      m[0] = w & 255; m[1] = (w >> 8) & 255; m[2] = (w >> 16) & 255; m[3] = w >> 24;
    } else if constexpr (CPL == 2) {
      const uint32_t w = *(const AS1 uint16_t *)q;
      m[0] = w & 255; m[1] = w >> 8;
    } else {
      m[0] = *q;
    }
  }
};

template <typename O>
__device__ __forceinline__ float ldo(const float *base, O off) {
  return *(const AS1 float *)((const AS1 char *)base + off);
}

// Values the whole wave shares, moved to scalar registers.  The compiler cannot always prove a
// value wave-uniform (anything derived from threadIdx.x >> 6, or selected under per-lane control
// flow); a buffer access with such an operand would otherwise be wrapped in a loop over its
// distinct values.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint32_t uni(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
__device__ __forceinline__ uint64_t uni(uint64_t x) {
  return (uint64_t)uni((uint32_t)x) | ((uint64_t)uni((uint32_t)(x >> 32)) << 32);
}
template <typename T>
__device__ __forceinline__ T *uni(T *p) {
  return (T *)uni((uint64_t)(uintptr_t)p);
}
// The same pointer, opaque to the optimiser: loads through it are not hoisted out of the branch
// that makes them, so rarely used arguments do not stay live in scalar registers across a loop.
template <typename T>
__device__ __forceinline__ const T *opaque(const T *p) {
  asm volatile("" : "+s"(p));
  return p;
}

template <typename T>
__device__ __forceinline__ const AS4 T *opaque_c(const AS4 T *p) {
  asm volatile("" : "+s"(p));
  return p;
}

// TF/SF face sheets: the incident values the six faces of the TF/SF box need at time step n,
// computed once per step by the slab-H launch, so that the row kernel applies every TF/SF
// correction with one load.  Sheet q = 4 eh + 2 hi + u of face axis a (x, y, z): update eh (0: H,
// from the incident E; 1: E, from the incident H), face hi (0: low, 1: high) and the face's two
// tangential components u, over the face's other two axes within the box (x faces (j, k), y faces
// (i, k), z faces (i, j), first index fastest).  The x sheets come first, then y, then z.
// Component, range and face position per (a, eh, u), those of the TF/SF corrections of
// src/gpu.h:152-573 (range: the box, one short along the axis named):
//   a  H u=0           H u=1           E u=0            E u=1
//   x  E_z  k<k1 ->Hy  E_y  j<j1 ->Hz  H_y  k<k1 ->Ez   H_z  j<j1 ->Ey
//   y  E_z  k<k1 ->Hx  E_x  i<i1 ->Hz  H_z  i<i1 ->Ex   H_x  k<k1 ->Ez
//   z  E_y  j<j1 ->Hx  E_x  i<i1 ->Hy  H_y  i<i1 ->Ex   H_x  j<j1 ->Ey
// H sheets sit on the face planes i0|i1 (j0|j1, k0|k1), E sheets on i0-1|i1 (j0-1|j1, k0-1|k1).
// Entries outside their range are not written (and not read).
// (32-bit index arithmetic: the sheets of a grid hold far fewer than 2^32 entries.)
__device__ __forceinline__ void tfsf_sheet(const GridDev &g, const PWDev &pw, const float *rowE, const float *rowH,
                                           float *xs, int ni, int nj, int nk, uint32_t t) {
  const uint32_t Sx = (uint32_t)nj * nk, Sy = (uint32_t)ni * nk, Sz = (uint32_t)ni * nj;
  int a;
  uint32_t r;
  if (t < 8 * Sx) { a = 0; r = t; }
  else if (t < 8 * (Sx + Sy)) { a = 1; r = t - 8 * Sx; }
  else { a = 2; r = t - 8 * (Sx + Sy); }
  const uint32_t Sa = a == 0 ? Sx : (a == 1 ? Sy : Sz);
  const int q = (int)(r / Sa);
  const uint32_t e = r - (uint32_t)q * Sa;
  const int eh = q >> 2, hi = (q >> 1) & 1, u = q & 1;
  const uint32_t nf = a == 0 ? nj : ni;
  const int fa = (int)(e % nf), sa = (int)(e / nf);
  // bit, incident component and the axis whose range stops one short, per [eh][a][u]
  constexpr int kBit[2][3][2] = {{{1, 2}, {3, 5}, {6, 7}}, {{2, 1}, {3, 5}, {6, 7}}};
  constexpr int kComp[2][3][2] = {{{2, 1}, {2, 0}, {1, 0}}, {{1, 2}, {2, 0}, {1, 0}}};
  constexpr int kShort[2][3][2] = {{{2, 1}, {2, 0}, {1, 0}}, {{2, 1}, {0, 2}, {0, 1}}};
  const int bit = kBit[eh][a][u], comp = kComp[eh][a][u], sh = kShort[eh][a][u];
  if (!((eh == 0 ? pw.active_h : pw.active_e) & (1 << bit))) return;
  int i, j, k;
  const int lo0[3] = {g.i0, g.j0, g.k0}, hi0[3] = {g.i1, g.j1, g.k1};
  const int pos = eh == 0 ? (hi ? hi0[a] : lo0[a]) : (hi ? hi0[a] : lo0[a] - 1);
  if (a == 0) { i = pos; j = g.j0 + fa; k = g.k0 + sa; }
  else if (a == 1) { i = g.i0 + fa; j = pos; k = g.k0 + sa; }
  else { i = g.i0 + fa; j = g.j0 + sa; k = pos; }
  const int c3[3] = {i, j, k};
  if (c3[sh] > hi0[sh] - 1) return;
  const float v = eh == 0 ? inc_e(pw, rowE, comp, i, j, k) : inc_h(pw, rowH, comp, i, j, k);
  const uint32_t off = a == 0 ? 0 : (a == 1 ? 8 * Sx : 8 * (Sx + Sy));
  xs[off + r] = v;
}

// Row kernel (B).  Waves are S row segments x R rows per workgroup; lane l of segment g holds the
// columns i = 64 CPL g + CPL l + c, c < CPL, of row j and marches k through its chunk of planes.
// Per plane it computes H(i, j, k) at n+1/2 and Hx, Hz at (i, j-1, k) (the row below belongs to
// another wave, so those two components are recomputed from the old state; on the first row the
// row below is a y-slab row, whose new H the slab-H kernel has written), then E(i, j, k) at n+1.
// H at (i, j, k-1) comes from the previous plane in registers; a chunk that does not start on
// plane P first recomputes H on the plane below it.  Column neighbours come from the lane's own
// registers, from the next / previous lane through cross-lane shuffles and, across segments,
// through LDS with one barrier per plane.  The lane offset o addresses (first column, j-1, k);
// rows j and j+1, plane k+1 and the field component are uniform offsets.  1/kappa_y and 1/kappa_z
// are exactly 1 on these rows and planes and their factors are omitted; 1/kappa_x is applied in
// the x-slab columns (it is 1, an exact no-op, elsewhere).
//
// Two instantiations of the march share the kernel: plain waves (every lane in the grid, no x-slab
// column, no column on an x face of the TF/SF box, a row no y face of the box touches) carry no
// per-lane branches; edge waves hold their per-lane constants (x-slab profile values, 1/kappa_x,
// predicate bits, face columns) in registers across the march.  (Issuing the next plane's loads
// before computing the current one was measured slower at 256^3: the second plane buffer costs
// occupancy, and the plain march already streams within 5% of a pure 7-in / 6-out copy of the
// same arrays.)

// Loads of one plane of the march (see rows_body).
template <int CPL, bool MU>
struct RowPlaneMu {};  // per-cell H coefficients of a plane (mu runs only; empty otherwise)
template <int CPL>
struct RowPlaneMu<CPL, true> {
  float mux[CPL], muy[CPL], muz[CPL], muxm[CPL], muzm[CPL];
};
template <int CPL, bool MU = false>
struct RowPlane : RowPlaneMu<CPL, MU> {
  float nx[CPL], ny[CPL], nz[CPL], nym[CPL];     // E(k+1) of row j, Ey(k+1) of row j-1
  float exm[CPL], ezm[CPL], exq[CPL], ezq[CPL];  // Ex, Ez of rows j-1 and j+1 at plane k
  float hx[CPL], hy[CPL], hz[CPL];               // H of row j at plane k (old state)
  float hxm[CPL], hzm[CPL];                      // Hx, Hz of row j-1 (old, or new if not recomputed)
  int mt[CPL];                                   // material index (index mode)
  float pHy[CPL], pHz[CPL], pHzm[CPL], pEy[CPL], pEz[CPL];  // x-slab psi (edge waves)
  float vhy, vhz, vhzm, vez, vey;                // x-face incident values (edge waves)
  float qHxy[CPL], qHzy[CPL], qExy[CPL], qEzy[CPL], qHxym[CPL], qHzym[CPL];  // y-slab psi (slab waves)
  float qHxz[CPL], qHyz[CPL], qExz[CPL], qEyz[CPL], qHxzm[CPL];              // z-slab psi
};

// Workgroup size bound of the row kernel per columns per lane: wider lanes hold more registers.
constexpr int kRowsMaxThreads[5] = {0, 1024, 512, 0, 256};

// MU: per-cell H coefficients (CreateMu): H = Chh H + C(i,j,k) curl E, which is K2's H + C curl E
// bit for bit since Chh is exactly 1.
// WB: write-back stores (a state that fits in the Infinity Cache), else non-temporal.
// NTF: the slab part of a grid whose TF/SF box lies clear of the slab rows and planes (StepArgs
// ntf): the march carries no TF/SF code (256^3 slab part 74.1 -> 70.0 us).
// GR: a graded mesh.  The per-axis tables carry dx / spacing on every index (1/kappa in the CPML
// cells, where the spacing is dx), so every march applies the y and z factors per row and plane
// and the x factor of each lane's column, which the uniform grid's marches omit where they are 1.
template <bool IDX, bool BUF, int CPL, int PART, bool MU = false, bool WB = false, bool NTF = false, bool GR = false>
__device__ __forceinline__ void rows_body(const StepArgs *__restrict__ pa_in, int L) {
  // The argument block does not change during a launch: read through the constant address space
  // it is loaded with scalar loads into scalar registers (also where the block's address was
  // itself read from memory, as in the batched launch).
  const AS4 StepArgs *pa = (const AS4 StepArgs *)pa_in;
  using O = typename Acc<BUF>::O;
  constexpr int W = 64 * CPL;
  const int I = pa->I, J = pa->J, K = pa->K, P = pa->P;
  const int bi0 = pa->i0, bi1 = pa->i1, bj0 = pa->j0, bj1 = pa->j1, bk0 = pa->k0, bk1 = pa->k1;
  const int S = pa->S, R = PART == 0 ? pa->R0 : pa->R, nTy = pa->nTy, KC = pa->KC, ntab = pa->ntab;
  const int act_h = PART == 1 && NTF ? 0 : uni(pa->act_h), act_e = PART == 1 && NTF ? 0 : uni(pa->act_e);
  // IJ: plane of the psi z-slab records (unpadded); IJp: plane of the per-cell arrays
  const size_t cs = pa->cs, IJK = pa->IJK, IJ = (size_t)I * J, Ip = pa->Ip, IJp = Ip * J;
  const float Chh = pa->Chh, Che = pa->Che, dx = pa->dx;
  const int P2 = 2 * P;

  __shared__ float sCoef[IDX ? 256 * 6 : 6];
  __shared__ float sSlab[6][128];   // per x-slab index s: bE cE bH cH, 1/kappa_x of E and of H
  __shared__ float sX[16][5];       // per wave: E ey ez eym (lane e_pub), new H hy hz (lane h_pub)
  extern __shared__ float sRow[];   // per row of the tile: Ex, Ez (plane k, old), Hx, Hz (new), Wrow each
  {
    const auto prm = gptr(pa->prm);
    if (IDX) {
      const auto coef = gptr(pa->coef);
      for (int q = threadIdx.x; q < ntab * 6; q += blockDim.x) sCoef[q] = coef[q];
    }
    for (int q = threadIdx.x; q < P2; q += blockDim.x) {
      const int i = q < P ? q : I - P2 + q;
      sSlab[0][q] = prm[q];
      sSlab[1][q] = prm[P2 + q];
      sSlab[2][q] = prm[2 * P2 + q];
      sSlab[3][q] = prm[3 * P2 + q];
      sSlab[4][q] = prm[4 * P2 + i];
      sSlab[5][q] = prm[4 * P2 + I + J + K + i];
    }
  }
  __syncthreads();

  // the wave index is uniform across the wave; readfirstlane lets the compiler keep the row, its
  // offsets and flags in scalar registers (and branch on them) instead of per-lane values
  // Waves of a workgroup take their segments in an order rotated per row and per workgroup (a hash
  // of L), so that the edge segments of the workgroups resident on a CU do not all fall on the same
  // SIMD (256^3 interior part 175.0 -> 173.7 us).
  const int lane = threadIdx.x & 63, wp = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = wp / S, rot = (int)(((uint32_t)L * 0x9E3779B1u) >> 30);
  const int seg = uni((wp % S + r + rot) % S), wv = uni(r * S + seg);
  // Part 0: the interior band of rows [P+1, J-P) over the interior chunks of planes [P, K-P).
  // Part 1: the rest -- the bands [0, P] and [J-P, J) over every chunk, and the interior band over
  // the z-slab chunks.  The two parts read only the old state and write disjoint cells.
  int row0, row_end, tz;
  {
    const int nTlo = pa->nTlo, nThi = pa->nThi, nTmid = pa->nTmid, n1 = pa->nz1, n2 = pa->nz2;
    if (PART == 0) {
      const int nT0 = pa->nTmid0;
      row0 = P + 1 + R * (L % nT0); row_end = J - P; tz = n1 + L / nT0;
    } else {
      const int ny = pa->nzy, nb = nTlo + nThi, nA = nb * (2 * n1 + ny);
      if (L < nA) {
        const int t = L % nb;
        tz = L / nb;
        if (t < nTlo) { row0 = R * t; row_end = P + 1; }
        else { row0 = J - P + R * (t - nTlo); row_end = J; }
      } else {
        const int q = L - nA, zz = q / nTmid;
        row0 = P + 1 + R * (q % nTmid); row_end = J - P; tz = zz < n1 ? zz : n1 + ny + (zz - n1);
      }
      (void)n2;
    }
    (void)nTy;
  }
  const int j = uni(row0 + r);
  // z chunk of the wave; chunks do not straddle the z-slab boundaries P and K-P (the interior
  // planes [P, K-P) in chunks of KC planes in the interior part, of KCY in the slab part)
  int kb, ke;
  {
    const int n1 = pa->nz1, n2 = PART == 0 ? pa->nz2 : pa->nzy, KCS = pa->KCS, KCm = PART == 0 ? KC : pa->KCY;
    if (tz < n1) { kb = tz * KCS; ke = min(kb + KCS, P); }
    else if (tz < n1 + n2) { kb = P + (tz - n1) * KCm; ke = min(kb + KCm, K - P); }
    else { kb = K - P + (tz - n1 - n2) * KCS; ke = min(kb + KCS, K); }
    kb = uni(kb); ke = uni(ke);
  }
  const bool row_ok = j < row_end;
  if (S == 1 && R == 1 && !row_ok) return;  // a one-wave workgroup: no barrier follows
  // Rows of a tile exchange through LDS: row j reads the old Ex, Ez of row j+1 and the new Hx, Hz of
  // row j-1 there.  The tile's first row recomputes Hx, Hz of the row below it from the old state,
  // and its last row (or the band's) reads row j+1 from memory.
  const bool firstr = uni(r) == 0, lastq = uni((int)(r == R - 1 || j + 1 >= row_end)) != 0;
  const bool xsync = S > 1 || R > 1;
  const int Wrow = pa->Wrow;
  // Columns of the lane.  Plain layout: segment g holds the columns [64 CPL g, 64 CPL (g+1)).
  // Split layout (1 column per lane, the row pitch Ip a multiple of 64): segment 0 holds both ends
  // of the row, lanes 0-31 the columns [0, 32) and lanes 32-63 the columns [Ip-32, Ip) (those from
  // I on are padding, off), and segment g >= 1 the columns [64 g - 32, 64 g + 32); the x CPML slabs
  // and (usually) the x faces of the TF/SF box then fall in one wave per row instead of two, and
  // every other wave runs the plain march (padded 252 x 256 x 256: 68.5 -> 69.8 Gcell/s).
  const bool spl = CPL == 1 && pa->spl;
  const int ib = spl ? (seg == 0 ? (lane < 32 ? lane : (int)Ip - 64 + lane) : 64 * seg - 32 + lane) : seg * W + CPL * lane;
  // Lanes and waves of the exchange with the neighbouring segments: lane e_pub's E (its column's
  // left neighbour is in wave w_prev) and lane h_pub's new H (its right neighbour is in wave
  // w_next) go through LDS to lane e_rd of wave w_prev and lane h_rd of wave w_next.
  const int wrow = wv - seg;
  const int w_next = wrow + (spl ? (seg + 1) % S : seg + 1), w_prev = wrow + (spl ? (seg + S - 1) % S : seg - 1);
  const bool has_next = spl || seg + 1 < S, has_prev = spl || seg > 0;
  const int e_pub = spl && seg == 0 ? 32 : 0, h_pub = spl && seg == 0 ? 31 : 63;
  const int e_rd = h_pub, h_rd = e_pub;
  const bool on = row_ok && ib < I;     // I is a multiple of CPL: all CPL columns are in the grid
  const bool has_jm = j >= 1;           // row j-1 exists (its Hx, Hz are recomputed from the old state)

  // arrays and uniform offsets (bytes): component c at c*cs; the lane offset o0 addresses row
  // max(j-1, 0), row j is at r1 and row j+1 at r2 from it (on the first row, row j-1 is row 0 and
  // on the last, row j+1 is row j: values read there are not used); plane k+1 at IJp
  const Acc<BUF> Fc(pa->Fc, 6 * cs * 4), Fn(pa->Fn, 6 * cs * 4);
  const Acc<BUF> Mt(pa->mat, IDX ? IJK : 0), Cf(pa->coef, IDX ? 0 : 24 * IJK);
  const Acc<BUF> Cm(MU ? pa->mu : nullptr, MU ? 12 * IJK : 0);
  const O cm1 = (O)(IJK * 4), cm2 = 2 * cm1;
  const O c1 = (O)(cs * 4), c2 = 2 * c1, c3 = 3 * c1, c4 = 4 * c1, c5 = 5 * c1;
  const O r1 = uni(has_jm ? (O)(Ip * 4) : (O)0), r2 = uni(r1 + (j + 1 < J ? (O)(Ip * 4) : (O)0)), pz = (O)(IJp * 4);

  const int k_first = uni(kb > 0 ? kb - 1 : kb), n_planes = uni(ke - k_first);
  const size_t col = (on ? (size_t)ib : 0) + (size_t)(has_jm ? j - 1 : 0) * Ip;
  const O o0 = (O)((col + (size_t)k_first * IJp) * 4);

  // plain wave (see above); the z faces of the box are applied on their planes under uniform
  // branches in both marches
  // slab wave: a row in a y CPML slab or next to one (its row j-1 is recomputed with the slab's psi),
  // or a chunk of z-slab planes
  const bool slabw = j <= P || j >= J - P || kb < P || ke > K - P;
  const int lo = spl ? 64 * seg - 32 : seg * W, hi = lo + W;
  bool plain = row_ok && !slabw && !(spl && seg == 0) && lo >= P && hi <= I - P;
  if (act_h | act_e) {
    plain = plain && !(bi0 - 1 >= lo && bi0 - 1 < hi) && !(bi0 >= lo && bi0 < hi) && !(bi1 >= lo && bi1 < hi);
    plain = plain && j != bj0 - 1 && j != bj0 && j != bj1 && j != bj1 + 1;
  }
  const bool zrow = j >= bj0 && j <= bj1, zrow_m = j - 1 >= bj0 && j - 1 <= bj1 - 1;

  // y-slab rows: slab index and 1/kappa_y of rows j and j-1 (1 off the slabs)
  const AS4 float *prmc = (const AS4 float *)pa->prm;
  const bool ys0 = j < P || j >= J - P, ysm0 = has_jm && (j - 1 < P || j - 1 >= J - P);
  const int sy = j < P ? j : j - (J - P2), sym = j - 1 < P ? j - 1 : j - 1 - (J - P2);
  const int ikE0 = 4 * P2, ikH0 = 4 * P2 + I + J + K;  // 1/kappa tables of E and H: x, y, z
  // Waves run one of four marches: 0 plain, 1 edge (x-slab columns, TF/SF faces, partial
  // segments), 2 slab (y-slab rows and the rows above them, z-slab chunks; also the edge work),
  // 3 slab without the y CPML (the z-slab chunks of the rows between the y slabs: 1/kappa_y is
  // exactly 1 there and its factors are omitted).
  auto march = [&](auto kind_tag) {
    constexpr int KIND = decltype(kind_tag)::value;
    constexpr bool EDGE = KIND >= 1, SLAB = KIND >= 2, YC = KIND != 3;
    // factors of the per-axis tables: applied where 1/kappa may differ from 1 (x per lane in edge
    // waves, y per row and z per plane in slab waves) and, on a graded mesh, everywhere
    constexpr bool XF = EDGE || GR, YF = (SLAB && YC) || GR;
    const bool ys = YC && ys0, ysm = YC && ysm0;
    const bool lon = !EDGE || on;
    const float kyH = YF ? prmc[ikH0 + I + j] : 1.0f, kyE = YF ? prmc[ikE0 + I + j] : 1.0f;
    const float kyHm = YF && has_jm ? prmc[ikH0 + I + j - 1] : 1.0f;
    const int syc = ys ? sy : 0, symc = ysm ? sym : 0;
    const float bHy = sSlab[2][syc], cHy = sSlab[3][syc], bEy = sSlab[0][syc], cEy = sSlab[1][syc];
    const float bHym = sSlab[2][symc], cHym = sSlab[3][symc];

    // ---- per-lane constants of the march ----
    // Predicates are bits of one register (FS per column c, then lane bits), re-read through an
    // opaque copy in every plane: the compiler would otherwise hoist each predicate out of the
    // march as a 64-bit lane mask in scalar registers and spill them.  A column is an x-slab
    // column when it updates H psi (s < 2P-1) or E psi (s >= 1).
    constexpr int FS = 7, F_H = 0, F_E = 1, F_LAST = 2, F_SLH = 3, F_SLE = 4, F_XH = 5, F_XE = 6;
    constexpr int F_SLAB = 28, F_ON = 29, F_XHL = 30, F_XEL = 31;  // lane: x slab, in grid, x faces
    static_assert(FS * CPL <= F_SLAB, "lane flags overlap the column flags");
    int sl[CPL];          // x-slab index of each column (-1 outside the x slabs)
    float bH[CPL], cH[CPL], bE[CPL], cE[CPL], kH[CPL], kE[CPL];
    unsigned lf = on ? 1u << F_ON : 0u;
    int xh_c = -1, xe_c = -1;  // x-face TF/SF columns (the faces are at least CPL columns apart)
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const int i = ib + c;
      sl[c] = EDGE && on ? (i < P ? i : (i >= I - P ? i - (I - P2) : -1)) : -1;
      const int s = sl[c] < 0 ? 0 : sl[c];
      bE[c] = sSlab[0][s]; cE[c] = sSlab[1][s]; bH[c] = sSlab[2][s]; cH[c] = sSlab[3][s];
      kE[c] = sl[c] >= 0 ? sSlab[4][s] : (GR && on ? prmc[ikE0 + i] : 1.0f);
      kH[c] = sl[c] >= 0 ? sSlab[5][s] : (GR && on ? prmc[ikH0 + i] : 1.0f);
      const unsigned b = (unsigned)FS * c;
      if (sl[c] >= 0) lf |= 1u << F_SLAB;
      if (sl[c] >= 0 && sl[c] < P2 - 1) lf |= 1u << (b + F_SLH);
      if (sl[c] >= 1) lf |= 1u << (b + F_SLE);
      if (on && i < I - 1) lf |= 1u << (b + F_H);
      if (on && i != 0) lf |= 1u << (b + F_E);
      if (i == I - 1) lf |= 1u << (b + F_LAST);  // den_ex = 0 on the last x plane
      if (EDGE && on && act_h && (i == bi0 - 1 || i == bi1)) { lf |= (1u << (b + F_XH)) | (1u << F_XHL); xh_c = c; }
      if (EDGE && on && act_e && (i == bi0 || i == bi1)) { lf |= (1u << (b + F_XE)) | (1u << F_XEL); xe_c = c; }
    }
    const bool xh_lo = xh_c >= 0 && ib + xh_c == bi0 - 1, xe_lo = xe_c >= 0 && ib + xe_c == bi0;
    // sign of the x-face corrections: a - v == a + (-v) exactly, so the face side is a factor
    const float sxh = xh_lo ? -1.0f : 1.0f, sxe = xe_lo ? -1.0f : 1.0f;
    auto bit = [](unsigned f, int b) { return ((f >> b) & 1u) != 0; };
    auto slab_col = [](unsigned f, int c) { return ((f >> (FS * c + F_SLH)) & 3u) != 0; };
    const int xnj = pa->xnj;
    const size_t xS = (size_t)xnj * pa->xnk;
    const uint32_t xh_sel = xh_lo ? 0u : (uint32_t)(2 * xS), xe_sel = xe_lo ? (uint32_t)(4 * xS) : (uint32_t)(6 * xS);
    // row predicates of the march (bits): j in [j0, j1], j in [j0, j1-1], j-1 in [j0, j1-1],
    // and the rows the z faces of the box cover (j, and j-1 for the recomputed Hx)
    const int jfl = uni((j >= bj0 && j <= bj1 ? 1 : 0) | (j >= bj0 && j <= bj1 - 1 ? 2 : 0) |
                        (j - 1 >= bj0 && j - 1 <= bj1 - 1 ? 4 : 0) | (zrow ? 8 : 0) | (zrow_m ? 16 : 0));
    // psi groups, the four members of a group interleaved per element (16 bytes): x [Eyx Ezx Hyx
    // Hzx], 2P J K elements, element s + 2P j + 2P J k; y [Exy Ezy Hxy Hzy], 2P I K elements,
    // i + I s + 2P I k; z [Exz Eyz Hxz Hyz], 2P I J elements, i + I j + I J s.  A lane reads and
    // writes the four values of a column with one 16-byte access.
    const size_t sx = pa->sx, syy = pa->sy, szz = pa->sz;
    const Acc<BUF> Pc(pa->Pc, 16 * (sx + syy + szz));
    const Acc<BUF> Pn(pa->Pn, 16 * (sx + syy + szz));
    const O gY = (O)(16 * sx), gZ = gY + (O)(16 * syy);  // byte offsets of the y and z groups
    const O oc = (O)((on ? (size_t)ib : 0) * 16);  // lane offset of the y- and z-slab psi (column i)
    O so[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) so[c] = (O)((sl[c] < 0 ? 0 : sl[c]) * 16);

    // ---- carried state: E of row j at plane k, Ey of row j-1 at plane k, H(i, j, k-1) ----
    float ex[CPL], ey[CPL], ez[CPL], eym[CPL], hpx[CPL], hpy[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) ex[c] = ey[c] = ez[c] = eym[c] = hpx[c] = hpy[c] = 0.0f;
    if (lon) {
      Fc.ld(o0 + r1, 0, ex); Fc.ld(o0 + r1, c1, ey); Fc.ld(o0 + r1, c2, ez);
      Fc.ld(o0, c1, eym);
    }
    float *const sRq = sRow + (size_t)(4 * r) * Wrow + (on ? ib : 0);  // this row's LDS slots, column ib

    // ---- loads of plane index it ----
    auto load = [&](RowPlane<CPL, MU> &Lp, int it, auto own_tag) {
      constexpr bool own = decltype(own_tag)::value;
      const int k = uni(k_first + it);
      const O o = o0 + (O)it * pz;
      const O o1 = o + r1, o2 = o + r2;
      // plane k+1 (the last plane of the grid reads itself: its H is not updated)
      const O pn1 = SLAB ? uni(k + 1 < K ? pz : (O)0) : pz;
      unsigned f = lf;
      int ah = act_h, ae = act_e, jf = jfl;
      if constexpr (PART == 1 && NTF) asm volatile("" : "+v"(f), "+s"(jf));
      else if (EDGE) asm volatile("" : "+v"(f), "+s"(ah), "+s"(ae), "+s"(jf));
      const bool xj0 = jf & 1, xj1 = jf & 2, xjm1 = jf & 4;
      if (!EDGE || bit(f, F_ON)) {
        Fc.ld(o1, pn1, Lp.nx); Fc.ld(o1, c1 + pn1, Lp.ny); Fc.ld(o1, c2 + pn1, Lp.nz);
        Fc.ld(o1, c3, Lp.hx); Fc.ld(o1, c4, Lp.hy); Fc.ld(o1, c5, Lp.hz);
        if (firstr) {  // row j-1, recomputed by the tile's first row
          Fc.ld(o, c1 + pn1, Lp.nym);
          Fc.ld(o, 0, Lp.exm); Fc.ld(o, c2, Lp.ezm);
          Fc.ld(o, c3, Lp.hxm);
          Fc.ld(o, c5, Lp.hzm);
        }
        if (lastq) { Fc.ld(o2, 0, Lp.exq); Fc.ld(o2, c2, Lp.ezq); }
        if constexpr (MU) {
          Cm.ld(o1, 0, Lp.mux); Cm.ld(o1, cm1, Lp.muy); Cm.ld(o1, cm2, Lp.muz);
          if (own && firstr) { Cm.ld(o, 0, Lp.muxm); Cm.ld(o, cm2, Lp.muzm); }
        }
        if (IDX && own) Mt.ldm(o1 >> 2, 0, Lp.mt);
        // y-slab psi of rows j and j-1
        if (SLAB && ys) {
          const O u = uni(gY + (O)((size_t)I * ((size_t)sy + (size_t)P2 * k) * 16));
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            float v[4];
            Pc.template ld<4>(oc + (O)(16 * c), u, v);
            Lp.qExy[c] = v[0]; Lp.qEzy[c] = v[1]; Lp.qHxy[c] = v[2]; Lp.qHzy[c] = v[3];
          }
        }
        if (SLAB && own && ysm && firstr) {
          const O u = uni(gY + (O)((size_t)I * ((size_t)sym + (size_t)P2 * k) * 16));
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            float v[4];
            Pc.template ld<4>(oc + (O)(16 * c), u, v);
            Lp.qHxym[c] = v[2]; Lp.qHzym[c] = v[3];
          }
        }
        // z-slab psi of plane k (also on the plane below an interior chunk, when that is a slab plane)
        if ((SLAB || !own) && (k < P || k >= K - P)) {
          const int sz = k < P ? k : k - (K - P2);
          const O u = uni(gZ + (O)(((size_t)I * j + IJ * (size_t)sz) * 16));
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            float v[4];
            Pc.template ld<4>(oc + (O)(16 * c), u, v);
            Lp.qExz[c] = v[0]; Lp.qEyz[c] = v[1]; Lp.qHxz[c] = v[2]; Lp.qHyz[c] = v[3];
            if (own && has_jm && firstr) Lp.qHxzm[c] = Pc.ld1(oc + (O)(16 * c), u - (O)(I * 16) + 8);
          }
        }
        if (EDGE && bit(f, F_SLAB)) {
          const O pq = uni((O)((size_t)P2 * ((size_t)j + (size_t)J * k) * 16));
#pragma unroll
          for (int c = 0; c < CPL; ++c)
            if (slab_col(f, c)) {
              float v[4];
              Pc.template ld<4>(so[c], pq, v);
              Lp.pEy[c] = v[0]; Lp.pEz[c] = v[1]; Lp.pHy[c] = v[2]; Lp.pHz[c] = v[3];
              if (has_jm && firstr) Lp.pHzm[c] = Pc.ld1(so[c], pq - (O)(P2 * 16) + 12);
            }
        }
      }
      if (EDGE) {
        Lp.vhy = Lp.vhz = Lp.vhzm = Lp.vez = Lp.vey = 0.0f;
        const uint32_t xo = (uint32_t)((k - bk0) * xnj + (j - bj0));
        const bool xk0 = k >= bk0 && k <= bk1 - 1, xk1 = k >= bk0 && k <= bk1;
        const float *xs = opaque_c(pa)->xs;
        if (bit(f, F_XHL)) {
          if ((ah & 2) && xj0 && xk0) Lp.vhy = ldo(xs, (xh_sel + xo) * 4u);
          if ((ah & 4) && xj1 && xk1) Lp.vhz = ldo(xs + xS, (xh_sel + xo) * 4u);
          if (has_jm && firstr && (ah & 4) && xjm1 && xk1) Lp.vhzm = ldo(xs + xS, (xh_sel + xo - 1u) * 4u);
        }
        if (bit(f, F_XEL)) {
          if ((ae & 4) && xj0 && xk0) Lp.vez = ldo(xs, (xe_sel + xo) * 4u);
          if ((ae & 2) && xj1 && xk1) Lp.vey = ldo(xs + xS, (xe_sel + xo) * 4u);
        }
      }
    };

    // ---- plane index it: H(k) of row j, Hx, Hz(k) of row j-1, then E(k) of row j ----
    auto compute = [&](RowPlane<CPL, MU> &Lp, int it, auto own_tag) {
      const int k = uni(k_first + it);
      constexpr bool own = decltype(own_tag)::value;  // a plane of the chunk (the plane below only feeds H(k-1))
      const O o1 = o0 + (O)it * pz + r1;
      const bool xk0 = k >= bk0 && k <= bk1 - 1, xk1 = k >= bk0 && k <= bk1;
      unsigned f = lf;
      int ah = act_h, ae = act_e, jf = jfl;
      if constexpr (PART == 1 && NTF) asm volatile("" : "+v"(f), "+s"(jf));
      else if (EDGE) asm volatile("" : "+v"(f), "+s"(ah), "+s"(ae), "+s"(jf));
      const bool xj0 = jf & 1, xj1 = jf & 2, xjm1 = jf & 4, zrow = jf & 8, zrow_m = jf & 16;
      const bool lon = !EDGE || bit(f, F_ON);

      // exchange 1 (old E of plane k): Ex, Ez of each row for the row below it, and Ey, Ez (and Ey
      // of row j-1) at the first column of each segment for the segment on its left
      float ey_n = __shfl_down(ey[0], 1), ez_n = __shfl_down(ez[0], 1), eym_n = __shfl_down(eym[0], 1);
      if (xsync) {
        if (R > 1 && lon) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) { sRq[c] = ex[c]; sRq[Wrow + c] = ez[c]; }
        }
        if (S > 1 && lane == e_pub) { sX[wv][0] = ey[0]; sX[wv][1] = ez[0]; sX[wv][2] = eym[0]; }
        lds_barrier();
        if (!lastq && lon) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) { Lp.exq[c] = sRq[4 * Wrow + c]; Lp.ezq[c] = sRq[5 * Wrow + c]; }
        }
        if (S > 1 && lane == e_rd && has_next) { ey_n = sX[w_next][0]; ez_n = sX[w_next][1]; eym_n = sX[w_next][2]; }
      }

      // z-slab plane: slab index, profile and 1/kappa_z (the plane below an interior chunk may be
      // a slab plane too); off the slabs 1/kappa_z is exactly 1 and its factor is omitted
      constexpr bool ZC = SLAB || !own;
      const bool zs = ZC && (k < P || k >= K - P);
      const int sz = zs ? (k < P ? k : k - (K - P2)) : 0;
      const float kzH = ZC || GR ? prmc[ikH0 + I + J + k] : 1.0f, kzE = SLAB || GR ? prmc[ikE0 + I + J + k] : 1.0f;
      const float bHz = sSlab[2][sz], cHz = sSlab[3][sz], bEz = sSlab[0][sz], cEz = sSlab[1][sz];
      // uniform parts of the update masks (slab waves reach the grid's first and last rows / planes)
      const bool h_jk = !SLAB || (j < J - 1 && k < K - 1);

      // H(i, j, k): K1 GPU_updateH (src/gpu.h:592-601), TF/SF, CPML Y-X-Z (src/gpu.h:604-725)
      float hx[CPL], hy[CPL], hz[CPL], hxm[CPL], hzm[CPL];
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        const float ey1 = c + 1 < CPL ? ey[c + 1 < CPL ? c + 1 : 0] : ey_n;
        const float ez1 = c + 1 < CPL ? ez[c + 1 < CPL ? c + 1 : 0] : ez_n;
        hx[c] = Lp.hx[c]; hy[c] = Lp.hy[c]; hz[c] = Lp.hz[c];
        if (h_jk && (!EDGE || bit(f, FS * c + F_H))) {
          const float dyz = XF ? (ez1 - ez[c]) * kH[c] : ez1 - ez[c];
          const float dzy = XF ? (ey1 - ey[c]) * kH[c] : ey1 - ey[c];
          float cx = Che, cy = Che, cz = Che;
          if constexpr (MU) { cx = Lp.mux[c]; cy = Lp.muy[c]; cz = Lp.muz[c]; }
          hx[c] = Chh * hx[c] + cx * (-(Lp.ezq[c] - ez[c]) * kyH + (Lp.ny[c] - ey[c]) * kzH);
          hy[c] = Chh * hy[c] + cy * (-(Lp.nx[c] - ex[c]) * kzH + dyz);
          hz[c] = Chh * hz[c] + cz * (-dzy + (Lp.exq[c] - ex[c]) * kyH);
        }
        if (EDGE && bit(f, FS * c + F_XH)) {
          if ((ah & 2) && xj0 && xk0) hy[c] = hy[c] + (Che * Lp.vhy) * sxh;
          if ((ah & 4) && xj1 && xk1) hz[c] = hz[c] - (Che * Lp.vhz) * sxh;
        }
      }
      // TF/SF y faces of H (rows j0-1, j1): edge waves only (plain waves never sit on them)
      if (EDGE && ah && (j == bj0 - 1 || j == bj1) && k >= bk0 && k <= bk1) {
        const AS4 StepArgs *q = opaque_c(pa);
        const int ni = q->xni, nj = q->xnj, nk = q->xnk;
        const size_t yS = (size_t)ni * nk;
        const bool lo = j == bj0 - 1;
        const float *yr = q->xs + 8 * (size_t)nj * nk + (lo ? 0 : 2 * yS) + ((ptrdiff_t)(k - bk0) * ni - bi0);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const int i = ib + c;
          // both sheet values are loaded before either is used: one wait for the pair
          const bool in = bit(f, F_ON) && i >= bi0 && i <= bi1;
          const bool ua = in && (ah & (1 << 3)) && k <= bk1 - 1, ub = in && (ah & (1 << 5)) && i <= bi1 - 1;
          float va = 0.0f, vb = 0.0f;
          if (ua) va = ldo(yr, (uint32_t)i * 4u);
          if (ub) vb = ldo(yr + yS, (uint32_t)i * 4u);
          if (ua) {
            const float v = Che * va;
            hx[c] = lo ? hx[c] + v : hx[c] - v;
          }
          if (ub) {
            const float v = Che * vb;
            hz[c] = lo ? hz[c] - v : hz[c] + v;
          }
        }
      }
      // TF/SF z faces of H (planes k0-1, k1)
      const bool zf_h = ah && (k == bk0 - 1 || k == bk1);
      if (zf_h && zrow) {
        const AS4 StepArgs *q = opaque_c(pa);
        const int ni = q->xni, nj = q->xnj, nk = q->xnk;
        const size_t zS = (size_t)ni * nj;
        const bool lo = k == bk0 - 1;
        const float *zr = q->xs + 8 * ((size_t)nj * nk + (size_t)ni * nk) + (lo ? 0 : 2 * zS) +
                          ((ptrdiff_t)(j - bj0) * ni - bi0);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const int i = ib + c;
          const bool in = lon && i >= bi0 && i <= bi1;
          const bool ua = in && (ah & (1 << 6)) && j <= bj1 - 1, ub = in && (ah & (1 << 7)) && i <= bi1 - 1;
          float va = 0.0f, vb = 0.0f;
          if (ua) va = ldo(zr, (uint32_t)i * 4u);
          if (ub) vb = ldo(zr + zS, (uint32_t)i * 4u);
          if (ua) {
            const float v = Che * va;
            hx[c] = lo ? hx[c] - v : hx[c] + v;
          }
          if (ub) {
            const float v = Che * vb;
            hy[c] = lo ? hy[c] + v : hy[c] - v;
          }
        }
      }
      // CPML Y (K5: rows of the y slabs, s < 2P-1, i < I-1, k < K-1)
      if (SLAB && ys && sy < P2 - 1 && k < K - 1) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          if (bit(f, FS * c + F_H)) {
            const float pxy = bHy * Lp.qHxy[c] + cHy * (Lp.ezq[c] - ez[c]) / dx;
            hx[c] = hx[c] - Che * pxy * dx;
            const float pzy = bHy * Lp.qHzy[c] + cHy * (Lp.exq[c] - ex[c]) / dx;
            hz[c] = hz[c] + Che * pzy * dx;
            Lp.qHxy[c] = pxy;
            Lp.qHzy[c] = pzy;
          }
        }
      }
      // CPML X (K4: x-slab columns, s < 2P-1, j < J-1, k < K-1)
      if (EDGE && h_jk && bit(f, F_SLAB)) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          if (bit(f, FS * c + F_SLH)) {
            const float ey1 = c + 1 < CPL ? ey[c + 1 < CPL ? c + 1 : 0] : ey_n;
            const float ez1 = c + 1 < CPL ? ez[c + 1 < CPL ? c + 1 : 0] : ez_n;
            const float pyx = bH[c] * Lp.pHy[c] + cH[c] * (ez1 - ez[c]) / dx;
            hy[c] = hy[c] + Che * pyx * dx;
            const float pzx = bH[c] * Lp.pHz[c] + cH[c] * (ey1 - ey[c]) / dx;
            hz[c] = hz[c] - Che * pzx * dx;
            Lp.pHy[c] = pyx;
            Lp.pHz[c] = pzx;
          }
        }
      }
      // CPML Z (K6: z-slab planes, s < 2P-1, i < I-1, j < J-1)
      if (zs && sz < P2 - 1 && (!SLAB || j < J - 1)) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          if (!EDGE || bit(f, FS * c + F_H)) {
            const float pxz = bHz * Lp.qHxz[c] + cHz * (Lp.ny[c] - ey[c]) / dx;
            hx[c] = hx[c] + Che * pxz * dx;
            const float pyz = bHz * Lp.qHyz[c] + cHz * (Lp.nx[c] - ex[c]) / dx;
            hy[c] = hy[c] - Che * pyz * dx;
            Lp.qHxz[c] = pxz;
            Lp.qHyz[c] = pyz;
          }
        }
      }

      // Hx, Hz at (i, j-1, k), recomputed the same way (not needed on the plane below the chunk)
#pragma unroll
      for (int c = 0; c < CPL; ++c) { hxm[c] = Lp.hxm[c]; hzm[c] = Lp.hzm[c]; }
      if (own && has_jm && firstr) {
        const int jm = j - 1;
        const bool hm_k = !SLAB || k < K - 1;
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const float eym1 = c + 1 < CPL ? eym[c + 1 < CPL ? c + 1 : 0] : eym_n;
          if (hm_k && (!EDGE || bit(f, FS * c + F_H))) {
            const float dzy = XF ? (eym1 - eym[c]) * kH[c] : eym1 - eym[c];
            float cx = Che, cz = Che;
            if constexpr (MU) { cx = Lp.muxm[c]; cz = Lp.muzm[c]; }
            hxm[c] = Chh * hxm[c] + cx * (-(ez[c] - Lp.ezm[c]) * kyHm + (Lp.nym[c] - eym[c]) * kzH);
            hzm[c] = Chh * hzm[c] + cz * (-dzy + (ex[c] - Lp.exm[c]) * kyHm);
          }
          if (EDGE && bit(f, FS * c + F_XH) && (ah & 4) && xjm1 && xk1) hzm[c] = hzm[c] - (Che * Lp.vhzm) * sxh;
        }
        if (EDGE && ah && (jm == bj0 - 1 || jm == bj1) && k >= bk0 && k <= bk1) {
          const AS4 StepArgs *q = opaque_c(pa);
          const int ni = q->xni, nj = q->xnj, nk = q->xnk;
          const size_t yS = (size_t)ni * nk;
          const bool lo = jm == bj0 - 1;
          const float *yr = q->xs + 8 * (size_t)nj * nk + (lo ? 0 : 2 * yS) + ((ptrdiff_t)(k - bk0) * ni - bi0);
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            const int i = ib + c;
            const bool in = bit(f, F_ON) && i >= bi0 && i <= bi1;
            const bool ua = in && (ah & (1 << 3)) && k <= bk1 - 1, ub = in && (ah & (1 << 5)) && i <= bi1 - 1;
            float va = 0.0f, vb = 0.0f;
            if (ua) va = ldo(yr, (uint32_t)i * 4u);
            if (ub) vb = ldo(yr + yS, (uint32_t)i * 4u);
            if (ua) {
              const float v = Che * va;
              hxm[c] = lo ? hxm[c] + v : hxm[c] - v;
            }
            if (ub) {
              const float v = Che * vb;
              hzm[c] = lo ? hzm[c] - v : hzm[c] + v;
            }
          }
        }
        if (zf_h && (ah & (1 << 6)) && zrow_m) {
          const AS4 StepArgs *q = opaque_c(pa);
          const int ni = q->xni, nj = q->xnj, nk = q->xnk;
          const size_t zS = (size_t)ni * nj;
          const bool lo = k == bk0 - 1;
          const float *zr = q->xs + 8 * ((size_t)nj * nk + (size_t)ni * nk) + (lo ? 0 : 2 * zS) +
                            ((ptrdiff_t)(jm - bj0) * ni - bi0);
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            const int i = ib + c;
            if (lon && i >= bi0 && i <= bi1) {
              const float v = Che * ldo(zr, (uint32_t)i * 4u);
              hxm[c] = lo ? hxm[c] - v : hxm[c] + v;
            }
          }
        }
        // CPML Y of row j-1 (a y-slab row)
        if (SLAB && ysm && sym < P2 - 1 && k < K - 1) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            if (bit(f, FS * c + F_H)) {
              const float pxy = bHym * Lp.qHxym[c] + cHym * (ez[c] - Lp.ezm[c]) / dx;
              hxm[c] = hxm[c] - Che * pxy * dx;
              const float pzy = bHym * Lp.qHzym[c] + cHym * (ex[c] - Lp.exm[c]) / dx;
              hzm[c] = hzm[c] + Che * pzy * dx;
            }
          }
        }
        if (EDGE && hm_k && bit(f, F_SLAB)) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            if (bit(f, FS * c + F_SLH)) {
              const float eym1 = c + 1 < CPL ? eym[c + 1 < CPL ? c + 1 : 0] : eym_n;
              const float pzx = bH[c] * Lp.pHzm[c] + cH[c] * (eym1 - eym[c]) / dx;
              hzm[c] = hzm[c] - Che * pzx * dx;
            }
          }
        }
        // CPML Z of row j-1 (Hx only)
        if (SLAB && zs && sz < P2 - 1) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            if (bit(f, FS * c + F_H)) {
              const float pxz = bHz * Lp.qHxzm[c] + cHz * (Lp.nym[c] - eym[c]) / dx;
              hxm[c] = hxm[c] + Che * pxz * dx;
            }
          }
        }
      }

      // exchange 2 (new H of plane k): Hx, Hz of each row for the row above it, and Hy, Hz at the last
      // column of each segment for the segment on its right (the plane below the chunk only
      // synchronises, so that its reads of exchange 1 end before the next plane's writes)
      float hy_p = __shfl_up(hy[CPL - 1], 1), hz_p = __shfl_up(hz[CPL - 1], 1);
      if (xsync) {
        if (own && R > 1 && lon) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) { sRq[2 * Wrow + c] = hx[c]; sRq[3 * Wrow + c] = hz[c]; }
        }
        if (own && S > 1 && lane == h_pub) { sX[wv][3] = hy[CPL - 1]; sX[wv][4] = hz[CPL - 1]; }
        lds_barrier();
        if (own && !firstr && lon) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) { hxm[c] = sRq[-2 * Wrow + c]; hzm[c] = sRq[-Wrow + c]; }
        }
        if (own && S > 1 && lane == h_rd && has_prev) { hy_p = sX[w_prev][3]; hz_p = sX[w_prev][4]; }
      }

      if (own) {
        if (lon) {
          Fn.template st<CPL, !WB>(o1, c3, hx);
          Fn.template st<CPL, !WB>(o1, c4, hy);
          Fn.template st<CPL, !WB>(o1, c5, hz);
        }
        // E(i, j, k): K3 GPU_updateE (src/gpu.h:745-762), TF/SF, CPML Y-Z-X (src/gpu.h:765-868)
        float ce[3][CPL], ch[3][CPL];
        if (IDX) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            const float *rw = sCoef + 6 * Lp.mt[c];
            ce[0][c] = rw[0]; ce[1][c] = rw[1]; ce[2][c] = rw[2];
            ch[0][c] = rw[3]; ch[1][c] = rw[4]; ch[2][c] = rw[5];
            if (EDGE && bit(f, FS * c + F_LAST)) ch[0][c] = ch[1][c] = ch[2][c] = 0.0f;
          }
        } else {
#pragma unroll
          for (int c = 0; c < CPL; ++c) ce[0][c] = ce[1][c] = ce[2][c] = ch[0][c] = ch[1][c] = ch[2][c] = 0.0f;
          if (lon) {
            const O cq = (O)(IJK * 4);
#pragma unroll
            for (int qq = 0; qq < 3; ++qq) {
              Cf.ld(o1, (O)qq * cq, ce[qq]);
              Cf.ld(o1, (O)(3 + qq) * cq, ch[qq]);
            }
          }
        }
        float ux[CPL], uy[CPL], uz[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const float hym1 = c > 0 ? hy[c > 0 ? c - 1 : 0] : hy_p;
          const float hzm1 = c > 0 ? hz[c > 0 ? c - 1 : 0] : hz_p;
          ux[c] = ex[c];
          uy[c] = ey[c];
          uz[c] = ez[c];
          if (!SLAB || (j != 0 && k != 0))
            ux[c] = ce[0][c] * ex[c] + ch[0][c] * ((hz[c] - hzm[c]) * kyE - (hy[c] - hpy[c]) * kzE);
          if (!EDGE || bit(f, FS * c + F_E)) {
            const float dzx = XF ? (hz[c] - hzm1) * kE[c] : hz[c] - hzm1;
            const float dyx = XF ? (hy[c] - hym1) * kE[c] : hy[c] - hym1;
            if (!SLAB || k != 0) uy[c] = ce[1][c] * ey[c] + ch[1][c] * ((hx[c] - hpx[c]) * kzE - dzx);
            if (!SLAB || j != 0) uz[c] = ce[2][c] * ez[c] + ch[2][c] * (dyx - (hx[c] - hxm[c]) * kyE);
          }
          if (EDGE && bit(f, FS * c + F_XE)) {
            if ((ae & 4) && xj0 && xk0) uz[c] = uz[c] + (ch[2][c] * Lp.vez) * sxe;
            if ((ae & 2) && xj1 && xk1) uy[c] = uy[c] - (ch[1][c] * Lp.vey) * sxe;
          }
        }
        // TF/SF y faces of E (rows j0, j1): edge waves only
        if (EDGE && ae && (j == bj0 || j == bj1) && k >= bk0 && k <= bk1) {
          const AS4 StepArgs *q = opaque_c(pa);
          const int ni = q->xni, nj = q->xnj, nk = q->xnk;
          const size_t yS = (size_t)ni * nk;
          const bool lo = j == bj0;
          const float *yr = q->xs + 8 * (size_t)nj * nk + (lo ? 4 : 6) * yS + ((ptrdiff_t)(k - bk0) * ni - bi0);
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            const int i = ib + c;
            const bool in = i >= bi0 && i <= bi1;
            const bool ua = in && (ae & (1 << 3)) && i <= bi1 - 1, ub = in && (ae & (1 << 5)) && k <= bk1 - 1;
            float va = 0.0f, vb = 0.0f;
            if (ua) va = ldo(yr, (uint32_t)i * 4u);
            if (ub) vb = ldo(yr + yS, (uint32_t)i * 4u);
            if (ua) {
              const float v = ch[0][c] * va;
              ux[c] = lo ? ux[c] - v : ux[c] + v;
            }
            if (ub) {
              const float v = ch[2][c] * vb;
              uz[c] = lo ? uz[c] + v : uz[c] - v;
            }
          }
        }
        // TF/SF z faces of E (planes k0, k1)
        if (ae && (k == bk0 || k == bk1) && zrow) {
          const AS4 StepArgs *q = opaque_c(pa);
          const int ni = q->xni, nj = q->xnj, nk = q->xnk;
          const size_t zS = (size_t)ni * nj;
          const bool lo = k == bk0;
          const float *zr = q->xs + 8 * ((size_t)nj * nk + (size_t)ni * nk) + (lo ? 4 : 6) * zS +
                            ((ptrdiff_t)(j - bj0) * ni - bi0);
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            const int i = ib + c;
            const bool in = i >= bi0 && i <= bi1;
            const bool ua = in && (ae & (1 << 6)) && i <= bi1 - 1, ub = in && (ae & (1 << 7)) && j <= bj1 - 1;
            float va = 0.0f, vb = 0.0f;
            if (ua) va = ldo(zr, (uint32_t)i * 4u);
            if (ub) vb = ldo(zr + zS, (uint32_t)i * 4u);
            if (ua) {
              const float v = ch[0][c] * va;
              ux[c] = lo ? ux[c] + v : ux[c] - v;
            }
            if (ub) {
              const float v = ch[1][c] * vb;
              uy[c] = lo ? uy[c] - v : uy[c] + v;
            }
          }
        }
        // CPML Y (K8: rows of the y slabs, s >= 1, 0 < i < I-1, 0 < k < K-1)
        if (SLAB && ys && sy >= 1 && k >= 1 && k < K - 1) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            if (bit(f, FS * c + F_E) && bit(f, FS * c + F_H)) {
              const float pxy = bEy * Lp.qExy[c] + cEy * (hz[c] - hzm[c]) / dx;
              ux[c] = ux[c] + ch[0][c] * pxy * dx;
              const float pzy = bEy * Lp.qEzy[c] + cEy * (hx[c] - hxm[c]) / dx;
              uz[c] = uz[c] - ch[2][c] * pzy * dx;
              Lp.qExy[c] = pxy;
              Lp.qEzy[c] = pzy;
            }
          }
        }
        // CPML Z (K9: z-slab planes, s >= 1, 0 < i < I-1, 0 < j < J-1)
        if (SLAB && zs && sz >= 1 && j >= 1 && j < J - 1) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            if (bit(f, FS * c + F_E) && bit(f, FS * c + F_H)) {
              const float pxz = bEz * Lp.qExz[c] + cEz * (hy[c] - hpy[c]) / dx;
              ux[c] = ux[c] - ch[0][c] * pxz * dx;
              const float pyz = bEz * Lp.qEyz[c] + cEz * (hx[c] - hpx[c]) / dx;
              uy[c] = uy[c] + ch[1][c] * pyz * dx;
              Lp.qExz[c] = pxz;
              Lp.qEyz[c] = pyz;
            }
          }
        }
        // CPML X (K7: x-slab columns, s >= 1, 0 < j < J-1, 0 < k < K-1)
        if (EDGE && (!SLAB || (j >= 1 && j < J - 1 && k >= 1 && k < K - 1)) && bit(f, F_SLAB)) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            if (bit(f, FS * c + F_SLE)) {
              const float hym1 = c > 0 ? hy[c > 0 ? c - 1 : 0] : hy_p;
              const float hzm1 = c > 0 ? hz[c > 0 ? c - 1 : 0] : hz_p;
              const float pyx = bE[c] * Lp.pEy[c] + cE[c] * (hz[c] - hzm1) / dx;
              uy[c] = uy[c] - ch[1][c] * pyx * dx;
              const float pzx = bE[c] * Lp.pEz[c] + cE[c] * (hy[c] - hym1) / dx;
              uz[c] = uz[c] + ch[2][c] * pzx * dx;
              Lp.pEy[c] = pyx;
              Lp.pEz[c] = pzx;
            }
          }
        }
        if (lon) {
          Fn.template st<CPL, !WB>(o1, 0, ux);
          Fn.template st<CPL, !WB>(o1, c1, uy);
          Fn.template st<CPL, !WB>(o1, c2, uz);
          // psi: the four members of each slab column in one 16-byte store; every element is written
          // each step (updated, or copied outside the masks)
          if (EDGE && bit(f, F_SLAB)) {
            const O pq = uni((O)((size_t)P2 * ((size_t)j + (size_t)J * k) * 16));
#pragma unroll
            for (int c = 0; c < CPL; ++c)
              if (slab_col(f, c)) {
                const float v[4] = {Lp.pEy[c], Lp.pEz[c], Lp.pHy[c], Lp.pHz[c]};
                // write-back: a row's block of 2P elements (320 bytes at P = 10) is not a whole
                // number of 128-byte lines; the neighbouring rows' stores complete its lines in
                // the L2 (256^3 interior part -0.5 us against non-temporal)
                Pn.template st<4, false>(so[c], pq, v);
              }
          }
          if (SLAB && ys) {
            const O u = uni(gY + (O)((size_t)I * ((size_t)sy + (size_t)P2 * k) * 16));
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
              const float v[4] = {Lp.qExy[c], Lp.qEzy[c], Lp.qHxy[c], Lp.qHzy[c]};
              Pn.template st<4, !WB>(oc + (O)(16 * c), u, v);
            }
          }
          if (SLAB && zs) {
            const O u = uni(gZ + (O)(((size_t)I * j + IJ * (size_t)sz) * 16));
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
              const float v[4] = {Lp.qExz[c], Lp.qEyz[c], Lp.qHxz[c], Lp.qHyz[c]};
              Pn.template st<4, !WB>(oc + (O)(16 * c), u, v);
            }
          }
        }
      }
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        hpx[c] = hx[c];
        hpy[c] = hy[c];
        ex[c] = Lp.nx[c]; ey[c] = Lp.ny[c]; ez[c] = Lp.nz[c]; eym[c] = Lp.nym[c];
      }
    };

    RowPlane<CPL, MU> A;
    // lanes outside the grid (edge waves) never load: their plane values stay zero
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      A.nx[c] = A.ny[c] = A.nz[c] = A.nym[c] = A.exm[c] = A.ezm[c] = A.exq[c] = A.ezq[c] = 0.0f;
      A.hx[c] = A.hy[c] = A.hz[c] = A.hxm[c] = A.hzm[c] = 0.0f;
      A.pHy[c] = A.pHz[c] = A.pHzm[c] = A.pEy[c] = A.pEz[c] = 0.0f;
      A.qHxy[c] = A.qHzy[c] = A.qExy[c] = A.qEzy[c] = A.qHxym[c] = A.qHzym[c] = 0.0f;
      A.qHxz[c] = A.qHyz[c] = A.qExz[c] = A.qEyz[c] = A.qHxzm[c] = 0.0f;
      A.mt[c] = 0;
      if constexpr (MU) A.mux[c] = A.muy[c] = A.muz[c] = A.muxm[c] = A.muzm[c] = 0.0f;
    }
    // The plane below the chunk (H only) is peeled off the march, so that every load of a march
    // iteration is consumed within it: a load left in flight across the back edge (the material
    // byte of a plane whose E is not updated) made the compiler wait for all memory operations,
    // the previous plane's stores included, before the next plane's loads.
    int it0 = 0;
    if (k_first < kb) {
      load(A, 0, std::false_type{});
      compute(A, 0, std::false_type{});
      it0 = 1;
    }
    __builtin_amdgcn_s_waitcnt(0);
    for (int it = it0; it < n_planes; ++it) {
      load(A, it, std::true_type{});
      compute(A, it, std::true_type{});
    }
  };

  if constexpr (PART == 0) {
    if (uni((int)plain)) {
      march(std::integral_constant<int, 0>{});
    } else {
      // edge waves carry the x-slab psi and set the tile's critical path: they issue first
      // (256^3 interior part 173.9 -> 173.0 us on one box)
      __builtin_amdgcn_s_setprio(2);
      march(std::integral_constant<int, 1>{});
    }
  } else {
    (void)slabw;
    if (!ys0 && !ysm0) march(std::integral_constant<int, 3>{});
    else march(std::integral_constant<int, 2>{});
  }
}

// Work the slab part's launch carries after its rows (fdtd_run): the TF/SF face sheets of the next
// time step, into the sheet buffer of the other state set, and the probe gather of the state the
// step reads (the result of the previous step).  Neither touches what the rows read or write, so
// these workgroups fill the slots the slab rows leave free instead of taking launches of their own.
struct TailArgs {
  GridDev g;
  PWDev pw;
  float *xs;                 // sheets of time index nsh (nb_sheets workgroups; 0: none)
  int xni, xnj, xnk, nsh;
  unsigned nb_sheets, nb_probes;
  FieldSet F;                // probes (nb_probes workgroups): record row prow of port
  const int *pidx, *pcomps;
  float *port;
  int nAnt, ncomp, prow, T;
  // graph-recorded launches: nsh and prow are offsets from idx[0] (time index) and idx[1] (record
  // row); sheets of time indices >= T and record rows < idx[2] (before the run) are skipped
  const int *idx;
  // step currents (nb_cur workgroups) on the new state: the box lies in the interior part's cells,
  // which that launch has written; sample s, or src[*dn + q] in a recorded graph
  FieldSet Fn;
  CoefDev co;
  int comp, a0, b0, c0, ni, nj, nk, q;
  float s, dx;
  const float *src;
  const int *dn;
  unsigned nb_cur;
};

template <bool IDX>
__device__ __forceinline__ void tail_body(const TailArgs &ta, unsigned blk) {
  const unsigned nthr = blockDim.x;
  if (blk < ta.nb_sheets) {
    const int nsh = ta.idx ? ta.idx[0] + ta.nsh : ta.nsh;
    if (nsh >= ta.T) return;  // past the line history: a step after the last has no sheets
    const uint32_t ns = 8u * ((uint32_t)ta.xnj * ta.xnk + (uint32_t)ta.xni * ta.xnk + (uint32_t)ta.xni * ta.xnj);
    const float *rowE = ta.pw.active_h ? ta.pw.Eh + (size_t)nsh * ta.pw.N1 : nullptr;
    const float *rowH = ta.pw.active_e ? ta.pw.Hh + (size_t)nsh * ta.pw.N1 : nullptr;
    for (uint32_t t = blk * nthr + threadIdx.x; t < ns; t += ta.nb_sheets * nthr)
      tfsf_sheet(ta.g, ta.pw, rowE, rowH, ta.xs, ta.xni, ta.xnj, ta.xnk, t);
    return;
  }
  blk -= ta.nb_sheets;
  if (blk < ta.nb_probes) {
    const int prow = ta.idx ? ta.idx[1] + ta.prow : ta.prow;
    if (prow < 0 || prow >= ta.T || (ta.idx && prow < ta.idx[2])) return;
    const uint32_t np = (uint32_t)ta.nAnt * ta.ncomp;
    for (uint32_t t = blk * nthr + threadIdx.x; t < np; t += ta.nb_probes * nthr) {
      const int q = t / ta.nAnt, a = t % ta.nAnt;
      ta.port[((size_t)q * ta.T + prow) * ta.nAnt + a] = ta.F.f[ta.pcomps[q]][ta.pidx[a]];
    }
    return;
  }
  blk -= ta.nb_probes;
  // K17 AddSource_GPU_3D_1DThread (src/gpu.h:55-82), as k_add_currents
  const float s = ta.src ? ta.src[*ta.dn + ta.q] : ta.s;
  const long nc = (long)ta.ni * ta.nj * ta.nk;
  for (long tid = (long)blk * nthr + threadIdx.x; tid < nc; tid += (long)ta.nb_cur * nthr) {
    const int i = ta.a0 + (int)(tid % ta.ni), j = ta.b0 + (int)((tid / ta.ni) % ta.nj), k = ta.c0 + (int)(tid / ((long)ta.ni * ta.nj));
    const size_t id = (size_t)i + (size_t)j * gpitch(ta.g) + (size_t)k * gpitch(ta.g) * ta.g.J;
    float ce[3], ch[3];
    cell_coef<IDX>(ta.co, ta.co.tab, id, i, ta.g.I, ce, ch);
    float *E = ta.Fn.f[ta.comp];
    E[id] = E[id] - ch[ta.comp] * s * ta.dx / ta.dx / ta.dx;
  }
}

// The row kernel's workgroups [0, nwg); the slab part's launch may carry tail workgroups after them.
template <bool IDX, bool BUF, int CPL, int PART, bool MU, bool WB = false, bool NTF = false>
__global__ void __launch_bounds__(kRowsMaxThreads[CPL]) k_rows(const StepArgs *__restrict__ pa, int n, int nwg,
                                                               const TailArgs ta) {
  const int b = blockIdx.x;
  if (PART == 1 && b >= nwg) {
    tail_body<IDX>(ta, b - nwg);
    return;
  }
  rows_body<IDX, BUF, CPL, PART, MU, WB, NTF>(pa, xcd_remap(b, nwg));
}
// Graded mesh: both parts with one column per lane, non-temporal stores, no TF/SF code (a graded
// mesh takes no plane wave); the slab part's launch carries the tail workgroups as k_rows's does.
template <bool IDX, bool BUF, int PART>
__global__ void __launch_bounds__(kRowsMaxThreads[1]) k_rows_gr(const StepArgs *__restrict__ pa, int n, int nwg,
                                                                const TailArgs ta) {
  const int b = blockIdx.x;
  if (PART == 1 && b >= nwg) {
    tail_body<IDX>(ta, b - nwg);
    return;
  }
  rows_body<IDX, BUF, 1, PART, false, false, true, true>(pa, xcd_remap(b, nwg));
}
// The interior part with one column per lane, held to 64 registers so that 8 waves share each SIMD
// (measured 4% faster at 256^3 than the 68 the compiler picks freely).
template <bool IDX, bool BUF, bool WB = false>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_rows_c1(const StepArgs *__restrict__ pa, int n) {
  rows_body<IDX, BUF, 1, 0, false, WB>(pa, xcd_remap(blockIdx.x, gridDim.x));
}

// Batched launches cover the workgroups of many grids: tab[g].off is grid g's first workgroup
// (tab[ng].off the total), arg[p] its argument block while its current state set is p.  At batch
// step `step` grid g is at time index n0 + step with current set cur0 ^ (step & 1).
__device__ __forceinline__ int find_grid(const BatchGrid *__restrict__ tab, int ng, int b) {
  int lo = 0, hi = ng;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tab[mid].off <= b) lo = mid;
    else hi = mid;
  }
  return uni(lo);
}

// Blocks of one grid whose local index agrees mod 8 run on the same XCD (blocks are dealt to the
// XCDs round-robin), so the per-grid remap is xcd_remap of the local index.
template <bool IDX, bool BUF, int CPL, int PART, bool NTF = false>
__global__ void __launch_bounds__(kRowsMaxThreads[CPL]) k_rows_batch(const BatchGrid *__restrict__ tab, int ng, int step) {
  const int b = blockIdx.x, g = find_grid(tab, ng, b);
  const int par = (tab[g].cur0 + step) & 1;
  const int off = tab[g].off, nb = tab[g + 1].off - off;
  rows_body<IDX, BUF, CPL, PART, false, false, NTF>((const StepArgs *)tab[g].arg[par], xcd_remap(b - off, nb));
}


// The same tiles pulled from device work queues by resident workgroups.  Workgroups are dealt to
// the 8 XCDs round-robin, and the XCD of workgroup w (w mod 8) keeps a queue of its own: the
// work items b = 8 t + x for tickets t = 0, 1, ... of counter queue[8 slot + x], the block
// indices a merged launch would run on that XCD, so a tile's halo rows stay in its XCD's L2.  A
// workgroup takes tickets until they pass nitems.  Workgroup 0 clears the counters of slot ^ 1
// for the launch after this one (launches on one stream alternate the slots; the previous launch
// has finished with that slot).
template <bool IDX, bool BUF, int CPL, int PART>
__global__ void __launch_bounds__(kRowsMaxThreads[CPL]) k_rows_queue(const BatchGrid *__restrict__ tab, int ng, int step,
                                                                     int *queue, int slot, int nitems) {
  __shared__ int s_item;
  if (blockIdx.x == 0 && threadIdx.x < 8) queue[8 * (slot ^ 1) + threadIdx.x] = 0;
  const int x = blockIdx.x & 7;
  int *q = queue + 8 * slot + x;
  for (;;) {
    if (threadIdx.x == 0) s_item = 8 * atomicAdd(q, 1) + x;
    __syncthreads();
    const int b = uni(s_item);
    if (b >= nitems) return;  // the same ticket in every wave: the whole workgroup leaves
    const int g = find_grid(tab, ng, b);
    const int par = (tab[g].cur0 + step) & 1;
    const int off = tab[g].off, nb = tab[g + 1].off - off;
    rows_body<IDX, BUF, CPL, PART, false>((const StepArgs *)tab[g].arg[par], xcd_remap(b - off, nb));
    __syncthreads();  // every wave is done with the tile's LDS (and has read s_item) before the next ticket
  }
}

// Argument block of the TF/SF face-sheet fill (k_sheets), one per (state parity, plane-wave mode).
struct SlabArgs {
  GridDev g;
  FieldSet cur, nxt;
  PsiSet pcur, pnxt;
  CoefDev co;
  PWDev pw;
  int n;
  float *xs;
  int xni, xnj, xnk;
};

// The argument block of a batched workgroup, copied to LDS once (it is read field by field).
__device__ __forceinline__ void load_slab_args(SlabArgs &dst, const void *src) {
  const uint32_t *s = (const uint32_t *)src;
  for (int q = threadIdx.x; q < (int)(sizeof(SlabArgs) / 4); q += blockDim.x) ((uint32_t *)&dst)[q] = s[q];
  __syncthreads();
}

// The TF/SF face sheets of time step n (one thread per sheet entry).
__device__ __forceinline__ void sheets_body(const SlabArgs &a, unsigned blk, int n) {
  const uint32_t t = blk * 256u + threadIdx.x;
  if (t < 8u * ((uint32_t)a.xnj * a.xnk + (uint32_t)a.xni * a.xnk + (uint32_t)a.xni * a.xnj)) {
    const float *rowE = a.pw.active_h ? a.pw.Eh + (size_t)n * a.pw.N1 : nullptr;
    const float *rowH = a.pw.active_e ? a.pw.Hh + (size_t)n * a.pw.N1 : nullptr;
    tfsf_sheet(a.g, a.pw, rowE, rowH, a.xs, a.xni, a.xnj, a.xnk, t);
  }
}

__global__ void __launch_bounds__(256) k_sheets(const SlabArgs a) { sheets_body(a, blockIdx.x, a.n); }

__global__ void __launch_bounds__(256) k_sheets_batch(const BatchGrid *__restrict__ tab, int ng, int step) {
  __shared__ __align__(16) char raw[sizeof(SlabArgs)];
  SlabArgs &sa = *reinterpret_cast<SlabArgs *>(raw);
  const int b = blockIdx.x, g = find_grid(tab, ng, b);
  const int par = (tab[g].cur0 + step) & 1;
  load_slab_args(sa, tab[g].arg[par]);
  sheets_body(sa, b - tab[g].off, tab[g].n0 + step);
}

// Probe gather of many grids after a batch step: job blockIdx.y, probes blockIdx.x * 64 + lane,
// sampled from the set that is current after the step.
__global__ void __launch_bounds__(64) k_probes_batch(const ProbeJob *__restrict__ jobs, int step) {
  const ProbeJob &jb = jobs[blockIdx.y];
  const int a = blockIdx.x * 64 + threadIdx.x, row = jb.row0 + step;
  if (a >= jb.nAnt || row >= jb.T) return;
  const int cur = (jb.cur0 + step + 1) & 1;
  for (int q = 0; q < jb.ncomp; ++q)
    jb.port[((size_t)q * jb.T + row) * jb.nAnt + a] = jb.F[cur].f[jb.comps[q]][jb.idx[a]];
}

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------
static inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// Cells K17 AddSource_GPU_3D_1DThread writes for an AddCurrents box [i0,i1,j0,j1,k0,k1]
// (src/gpu.h:55-82 via src/simulation.h:256-286): the box, with its upper ends clipped to the
// grid (the kernel's i < Imax test, gpu.h:75).  The reference hands the int box to the kernel as
// size_t, so a negative start either makes i1 < i0 (the kernel returns, gpu.h:60-61) or puts
// every index past the grid: a box with a negative start, or an empty box, writes nothing.
bool current_box(const GridDev &g, const int box[6], int &a0, int &a1, int &b0, int &b1, int &c0, int &c1) {
  if (box[1] < box[0] || box[3] < box[2] || box[5] < box[4]) return false;
  if (box[0] < 0 || box[2] < 0 || box[4] < 0) return false;
  a0 = box[0]; b0 = box[2]; c0 = box[4];
  a1 = box[1] >= g.I ? g.I - 1 : box[1];
  b1 = box[3] >= g.J ? g.J - 1 : box[3];
  c1 = box[5] >= g.K ? g.K - 1 : box[5];
  return a0 <= a1 && b0 <= b1 && c0 <= c1;
}

hipError_t launch_update_h_all(hipStream_t st, const GridDev &g, const FieldSet &F, const CoefDev &co) {
  if (g.I < 2 || g.J < 2 || g.K < 2) return hipSuccess;
  dim3 grid(cdiv(g.I - 1, 64), cdiv(g.J - 1, 4), g.K - 1), blk(64, 4, 1);
  hipLaunchKernelGGL(k_h_all, grid, blk, 0, st, g, F, co);
  return hipGetLastError();
}

hipError_t launch_update_e_all(hipStream_t st, const GridDev &g, const FieldSet &F, const CoefDev &co) {
  dim3 grid(cdiv(g.I, 64), cdiv(g.J, 4), g.K), blk(64, 4, 1);
  if (co.mat) hipLaunchKernelGGL(k_e_all<true>, grid, blk, 0, st, g, F, co);
  else hipLaunchKernelGGL(k_e_all<false>, grid, blk, 0, st, g, F, co);
  return hipGetLastError();
}

hipError_t launch_update_h_pml(hipStream_t st, const GridDev &g, const FieldSet &F, const PsiSet &ps, const CoefDev &co) {
  const int P = g.P;
  hipLaunchKernelGGL(k_h_pml_y, dim3(cdiv(g.I - 1, 64), cdiv(2 * P - 1, 4), g.K - 1), dim3(64, 4, 1), 0, st, g, F, ps, co);
  hipLaunchKernelGGL(k_h_pml_x, dim3(cdiv((long)(2 * P - 1) * (g.J - 1), 256), g.K - 1), dim3(256), 0, st, g, F, ps, co);
  hipLaunchKernelGGL(k_h_pml_z, dim3(cdiv(g.I - 1, 64), cdiv(g.J - 1, 4), 2 * P - 1), dim3(64, 4, 1), 0, st, g, F, ps, co);
  return hipGetLastError();
}

hipError_t launch_update_e_pml(hipStream_t st, const GridDev &g, const FieldSet &F, const PsiSet &ps, const CoefDev &co) {
  const int P = g.P;
  dim3 gy(cdiv(g.I, 64), cdiv(2 * P - 1, 4), g.K - 2), gz(cdiv(g.I, 64), cdiv(g.J, 4), 2 * P - 1);
  dim3 gx(cdiv((long)(2 * P - 1) * (g.J - 2), 256), g.K - 2);
  if (co.mat) {
    hipLaunchKernelGGL(k_e_pml_y<true>, gy, dim3(64, 4, 1), 0, st, g, F, ps, co);
    hipLaunchKernelGGL(k_e_pml_z<true>, gz, dim3(64, 4, 1), 0, st, g, F, ps, co);
    hipLaunchKernelGGL(k_e_pml_x<true>, gx, dim3(256), 0, st, g, F, ps, co);
  } else {
    hipLaunchKernelGGL(k_e_pml_y<false>, gy, dim3(64, 4, 1), 0, st, g, F, ps, co);
    hipLaunchKernelGGL(k_e_pml_z<false>, gz, dim3(64, 4, 1), 0, st, g, F, ps, co);
    hipLaunchKernelGGL(k_e_pml_x<false>, gx, dim3(256), 0, st, g, F, ps, co);
  }
  return hipGetLastError();
}

hipError_t launch_pw_h(hipStream_t st, const GridDev &g, const FieldSet &F, const CoefDev &co, const PWDev &pw, int n) {
  const int ni = g.i1 - g.i0 + 1, nj = g.j1 - g.j0 + 1, nk = g.k1 - g.k0 + 1;
  dim3 blk(16, 16, 1);
  hipLaunchKernelGGL(k_pw_h<0>, dim3(cdiv(nj, 16), cdiv(nk, 16)), blk, 0, st, g, F, co, pw, n);
  hipLaunchKernelGGL(k_pw_h<1>, dim3(cdiv(ni, 16), cdiv(nk, 16)), blk, 0, st, g, F, co, pw, n);
  hipLaunchKernelGGL(k_pw_h<2>, dim3(cdiv(ni, 16), cdiv(nj, 16)), blk, 0, st, g, F, co, pw, n);
  return hipGetLastError();
}

hipError_t launch_pw_e(hipStream_t st, const GridDev &g, const FieldSet &F, const CoefDev &co, const PWDev &pw, int n) {
  const int ni = g.i1 - g.i0 + 1, nj = g.j1 - g.j0 + 1, nk = g.k1 - g.k0 + 1;
  dim3 blk(16, 16, 1);
  if (co.mat) {
    hipLaunchKernelGGL((k_pw_e<0, true>), dim3(cdiv(nj, 16), cdiv(nk, 16)), blk, 0, st, g, F, co, pw, n);
    hipLaunchKernelGGL((k_pw_e<1, true>), dim3(cdiv(ni, 16), cdiv(nk, 16)), blk, 0, st, g, F, co, pw, n);
    hipLaunchKernelGGL((k_pw_e<2, true>), dim3(cdiv(ni, 16), cdiv(nj, 16)), blk, 0, st, g, F, co, pw, n);
  } else {
    hipLaunchKernelGGL((k_pw_e<0, false>), dim3(cdiv(nj, 16), cdiv(nk, 16)), blk, 0, st, g, F, co, pw, n);
    hipLaunchKernelGGL((k_pw_e<1, false>), dim3(cdiv(ni, 16), cdiv(nk, 16)), blk, 0, st, g, F, co, pw, n);
    hipLaunchKernelGGL((k_pw_e<2, false>), dim3(cdiv(ni, 16), cdiv(nj, 16)), blk, 0, st, g, F, co, pw, n);
  }
  return hipGetLastError();
}

hipError_t launch_add_currents(hipStream_t st, const GridDev &g, const FieldSet &F, const CoefDev &co, int comp,
                               const int box[6], float s, float dx, const float *src, const int *dn, int q) {
  int a0, a1, b0, b1, c0, c1;
  if (!current_box(g, box, a0, a1, b0, b1, c0, c1)) return hipSuccess;
  const int ni = a1 - a0 + 1, nj = b1 - b0 + 1, nk = c1 - c0 + 1;
  const unsigned nb = cdiv((long)ni * nj * nk, 256);
  if (co.mat) hipLaunchKernelGGL(k_add_currents<true>, dim3(nb), dim3(256), 0, st, g, F, co, comp, a0, b0, c0, ni, nj, nk, s, dx, src, dn, q);
  else hipLaunchKernelGGL(k_add_currents<false>, dim3(nb), dim3(256), 0, st, g, F, co, comp, a0, b0, c0, ni, nj, nk, s, dx, src, dn, q);
  return hipGetLastError();
}

hipError_t launch_advance(hipStream_t st, int *dn, int steps) {
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, dn, steps);
  return hipGetLastError();
}

hipError_t launch_line(hipStream_t st, const LineDev &L) {
  if (2 * L.P > 128) return hipErrorInvalidValue;
  const size_t shm = (size_t)2 * L.N * sizeof(float);
  if (shm > 64 * 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_line, dim3(1), dim3(1024), shm, st, L);
  return hipGetLastError();
}

hipError_t launch_probes(hipStream_t st, const FieldSet &F, const int *idx, int nAnt, const int *comps, int ncomp,
                         float *port, int row, int T) {
  if (nAnt <= 0 || ncomp <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_probes, dim3(cdiv(nAnt, 64), ncomp), dim3(64), 0, st, F, idx, nAnt, comps, ncomp, port, row, T);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Stream ceiling: the fused step's compulsory traffic with no stencil -- per cell the six field
// components and the material byte read from one state set and the six components written to the
// other (49 bytes), one pass, the row kernel's layout (components cs floats apart) and store
// policy.  Its rate bounds what any single-pass step can reach on this GPU.
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_stream(const float *__restrict__ Fc, float *__restrict__ Fn,
                                               const uint8_t *__restrict__ mat, size_t cs, size_t n) {
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t id = (size_t)blockIdx.x * 256 + threadIdx.x; id < n; id += stride) {
    const float m = (float)mat[id];
    const float ex = Fc[id], ey = Fc[cs + id], ez = Fc[2 * cs + id];
    const float hx = Fc[3 * cs + id], hy = Fc[4 * cs + id], hz = Fc[5 * cs + id];
    __builtin_nontemporal_store(ex + m, Fn + id);
    __builtin_nontemporal_store(ey + m, Fn + cs + id);
    __builtin_nontemporal_store(ez + m, Fn + 2 * cs + id);
    __builtin_nontemporal_store(hx + m, Fn + 3 * cs + id);
    __builtin_nontemporal_store(hy + m, Fn + 4 * cs + id);
    __builtin_nontemporal_store(hz + m, Fn + 5 * cs + id);
  }
}

hipError_t launch_stream(hipStream_t st, const float *Fc, float *Fn, const uint8_t *mat, size_t cs, size_t n) {
  const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, (size_t)256 * 64);
  hipLaunchKernelGGL(k_stream, dim3(nb), dim3(256), 0, st, Fc, Fn, mat, cs, n);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// The device scratch of fdtd_run_noalloc: one file-scope array per GPU, held until the process
// ends.  Host code never passes `scratch` itself (in host code the name stands for a host-side
// shadow, not a device address): scratch_region() looks the array's device address up once per
// device with hipGetSymbolAddress and every user takes that address.
// ------------------------------------------------------------------------------------------
static __device__ __align__(256) char scratch[kScratchBytes];

hipError_t scratch_region(void **ptr) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void *addr[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = p;
  }
  *ptr = addr[dev];
  return hipSuccess;
}

bool fused_supported(const GridDev &g) {
  return g.P >= 1 && g.P <= 64 && g.I >= 2 * g.P + 2 && g.J >= 2 * g.P + 2 && g.K >= 2 * g.P + 2;
}

// Row layout of the row kernel: CPL columns per lane (I a multiple of CPL, for aligned vector
// accesses) and S segments of 64 CPL columns per row, with the fewest idle lanes, then the fewest
// columns per lane (more waves in flight: measured faster at 256^3); at most 8 segments with one
// column per lane (rows up to 512 columns; 2 rows x 8 waves fill the 1024-thread workgroup) and 4
// with two or four (rows up to 1024 columns).  force_cpl != 0 restricts CPL.
static bool row_layout(int I, int force_cpl, int &cpl, int &S) {
  int best_waste = -1;
  for (int c : {1, 2, 4}) {
    if (I % c || (force_cpl && c != force_cpl)) continue;
    const int s = (I + 64 * c - 1) / (64 * c);
    if (s > (c == 1 ? 8 : 4)) continue;
    const int waste = s * 64 * c - I;
    if (best_waste < 0 || waste < best_waste) {
      best_waste = waste;
      cpl = c;
      S = s;
    }
  }
  return best_waste >= 0;
}

// LDS bytes of the row exchange: four rows of Wrow floats per row of the tile (none for one row)
static unsigned rows_lds(const StepArgs &a, int part) {
  const int R = part == 0 ? a.R0 : a.R;
  return R > 1 ? 16u * R * a.Wrow : 0u;
}

template <bool IDX, bool BUF, int PART, bool MU>
static void launch_rows_m(hipStream_t st, int cpl, int nwg, unsigned ntail, dim3 blk, unsigned lds, const StepArgs *d,
                          int n, const TailArgs &ta, bool wb, bool ntf, bool gr) {
  const dim3 grid(nwg + ntail);
  if (grid.x == 0) return;
  if constexpr (!MU) {
    if (gr) {  // make_step_args gives a graded mesh one column per lane and no per-cell mu
      hipLaunchKernelGGL((k_rows_gr<IDX, BUF, PART>), grid, blk, lds, st, d, n, nwg, ta);
      return;
    }
  }
  if constexpr (BUF && !MU) {  // write-back variants: one column per lane (grids that fit the cache)
    if (wb && cpl == 1) {
      if constexpr (PART == 0) hipLaunchKernelGGL((k_rows_c1<IDX, BUF, true>), grid, blk, lds, st, d, n);
      else if (ntf) hipLaunchKernelGGL((k_rows<IDX, BUF, 1, PART, false, true, true>), grid, blk, lds, st, d, n, nwg, ta);
      else hipLaunchKernelGGL((k_rows<IDX, BUF, 1, PART, false, true>), grid, blk, lds, st, d, n, nwg, ta);
      return;
    }
    // the slab part without TF/SF code (one column per lane, buffer addressing)
    if (PART == 1 && ntf && cpl == 1) {
      hipLaunchKernelGGL((k_rows<IDX, BUF, 1, PART, false, false, true>), grid, blk, lds, st, d, n, nwg, ta);
      return;
    }
  }
  if (cpl == 4) hipLaunchKernelGGL((k_rows<IDX, BUF, 4, PART, MU>), grid, blk, lds, st, d, n, nwg, ta);
  else if (cpl == 2) hipLaunchKernelGGL((k_rows<IDX, BUF, 2, PART, MU>), grid, blk, lds, st, d, n, nwg, ta);
  else if constexpr (PART == 0 && !MU) hipLaunchKernelGGL((k_rows_c1<IDX, BUF>), grid, blk, lds, st, d, n);
  // (with mu the interior part takes the register-unbounded kernel: held to 64 registers its
  // march would spill)
  else hipLaunchKernelGGL((k_rows<IDX, BUF, 1, PART, MU>), grid, blk, lds, st, d, n, nwg, ta);
}

// mu: the per-cell H coefficients of CreateMu are in use
template <bool IDX, bool BUF, int PART>
static void launch_rows_w(hipStream_t st, int cpl, int nwg, unsigned ntail, dim3 blk, unsigned lds, const StepArgs *d,
                          int n, const TailArgs &ta, bool mu, bool wb, bool ntf, bool gr) {
  if (mu) launch_rows_m<IDX, BUF, PART, true>(st, cpl, nwg, ntail, blk, lds, d, n, ta, wb, ntf, false);
  else launch_rows_m<IDX, BUF, PART, false>(st, cpl, nwg, ntail, blk, lds, d, n, ta, wb, ntf, gr);
}

// part: 0 the interior part, 1 the slab part (with the tail workgroups of ta, if any), -1 both
// (one after the other)
static hipError_t launch_rows(hipStream_t st, const StepArgs &a, const StepArgs *dargs, int part, int n,
                              const TailArgs *tail = nullptr) {
  // buffer addressing when every array the kernel touches (a state set, a psi set, the material
  // bytes or the per-cell coefficients) is below 4 GiB
  const size_t lim = ((size_t)1 << 32) - 1;
  const bool buf = !a.glob && 24 * a.cs < lim && 16 * (a.sx + a.sy + a.sz) < lim && (a.mat ? a.IJK : 24 * a.IJK) < lim &&
                   (!a.mu || 12 * a.IJK < lim);
  const bool mu = a.mu != nullptr;
  // the state (two field sets, two psi sets, materials) fits in the 256 MB Infinity Cache
  const size_t state = 2 * 24 * a.cs + 2 * 16 * (a.sx + a.sy + a.sz) + (a.mat ? a.IJK : 24 * a.IJK);
  const bool wb = state <= ((size_t)256 << 20);
  const bool ntf = a.ntf != 0, gr = a.gr != 0;
  const int n0 = part == 1 ? 0 : a.nwg0, n1 = part == 0 ? 0 : a.nwg1;
  TailArgs ta{};
  if (tail && part != 0) ta = *tail;
  const unsigned nt = ta.nb_sheets + ta.nb_probes + ta.nb_cur;
  const dim3 b0(64 * a.S * a.R0), b1(64 * a.S * a.R);
  const unsigned L0 = rows_lds(a, 0), L1 = rows_lds(a, 1);
  if (a.mat) {
    if (buf) { launch_rows_w<true, true, 0>(st, a.CPL, n0, 0, b0, L0, dargs, n, ta, mu, wb, ntf, gr); launch_rows_w<true, true, 1>(st, a.CPL, n1, nt, b1, L1, dargs, n, ta, mu, wb, ntf, gr); }
    else { launch_rows_w<true, false, 0>(st, a.CPL, n0, 0, b0, L0, dargs, n, ta, mu, wb, ntf, gr); launch_rows_w<true, false, 1>(st, a.CPL, n1, nt, b1, L1, dargs, n, ta, mu, wb, ntf, gr); }
  } else {
    if (buf) { launch_rows_w<false, true, 0>(st, a.CPL, n0, 0, b0, L0, dargs, n, ta, mu, wb, ntf, gr); launch_rows_w<false, true, 1>(st, a.CPL, n1, nt, b1, L1, dargs, n, ta, mu, wb, ntf, gr); }
    else { launch_rows_w<false, false, 0>(st, a.CPL, n0, 0, b0, L0, dargs, n, ta, mu, wb, ntf, gr); launch_rows_w<false, false, 1>(st, a.CPL, n1, nt, b1, L1, dargs, n, ta, mu, wb, ntf, gr); }
  }
  return hipGetLastError();
}

size_t step_args_size() { return sizeof(StepArgs); }

// Argument block of the row kernel (host copy, uploaded once per state parity / plane-wave mode);
// returns the number of row-kernel workgroups (0 if there are no such rows), -1 if the grid is
// not supported by the fused step.
int make_step_args(const FusedDev &f, const StepTuning &tune, void *out) {
  const GridDev &g = f.g;
  if (!fused_supported(g)) return -1;
  if (f.mat && f.ntab > 256) return -1;
  int cpl = 1, S = 1;
  // a graded mesh: one column per lane (its x factors are per lane) and no per-cell mu (the
  // per-op operations step such a grid)
  if (f.gr && f.Chmu) return -1;
  if (!row_layout(g.I, f.gr ? 1 : tune.cpl, cpl, S)) return -1;
  // the two x faces of the TF/SF box must not share a lane (at most one face column per lane)
  if ((f.act_h | f.act_e) && g.i1 - g.i0 < cpl) {
    if (tune.cpl || !row_layout(g.I, 1, cpl, S)) return -1;
  }
  // rows per tile (a workgroup of R S waves, at most the row kernel's bound).  Interior part: two
  // rows (measured at 256^3: 180 us against 190 with one; the rows share their neighbour rows
  // through LDS; for the one- and two-segment rows of a batch of small grids two rows beat four
  // by 3%).  Slab part: one row where a row takes several segments (its heavier waves keep fewer
  // workgroups resident; two rows cost it 35%), else two.
  const int max_waves = kRowsMaxThreads[cpl] / 64;
  int R0 = 2, R = S == 1 ? 2 : 1;
  while (R0 > 1 && S * R0 > max_waves) R0 /= 2;
  while (R > 1 && S * R > max_waves) R /= 2;
  if (tune.tile_rows > 0) R = R0 = tune.tile_rows;
  if (S * R > max_waves || S * R0 > max_waves) return -1;
  // every row and plane of the grid; z chunks do not straddle the z-slab boundaries
  const int cJ = g.J, cK = g.K - 2 * g.P;
  const int nTy = (cJ + R - 1) / R;
  const int nTlo = (g.P + 1 + R - 1) / R, nThi = (g.P + R - 1) / R, nTmid = (g.J - 2 * g.P - 1 + R - 1) / R;
  const int nTmid0 = (g.J - 2 * g.P - 1 + R0 - 1) / R0;
  int kc = tune.z_chunk;
  if (kc <= 0) {
    // 8 planes per chunk (one recomputed plane below each chunk: 1/8 extra reads), fewer when
    // the grid would not fill the chip with about two waves per wave slot (256 CUs x 32)
    const long layer = (long)nTy * R * S;
    kc = 8;
    while (kc > 2 && layer * ((cK + kc - 1) / kc) < 16384) kc /= 2;
    if (kc > cK) kc = cK;
  }
  // The slab part: its y-band rows take the interior planes in chunks of kcy planes and every row
  // its z slabs in chunks of kcs (the whole slab when it is thin); tune.slab_chunk sets both (the
  // z slabs cut into chunks of at most that many planes).  By default a grid with full-depth
  // interior chunks gives its y-band rows chunks as deep as a whole z slab: the slab part's waves
  // are then all of one length, fewer and with less recompute (256^3, P = 10: slab part 68.0 ->
  // 65.5 us with 10-plane y-band chunks against 8).
  const int sc = tune.slab_chunk;
  const int kcs = sc > 0 ? std::min(sc, g.P) : (g.P <= 2 * kc ? g.P : kc);
  const int kcy = sc > 0 ? std::min(sc, cK) : (kc >= 8 && kcs == g.P ? std::min(std::max(kc, g.P), cK) : kc);
  const int nz1 = (g.P + kcs - 1) / kcs, nz2 = (cK + kc - 1) / kc, nzy = (cK + kcy - 1) / kcy;
  const int nKc = 2 * nz1 + nzy;
  StepArgs a;
  std::memset(&a, 0, sizeof(a));
  a.I = g.I; a.J = g.J; a.K = g.K; a.P = g.P;
  a.i0 = g.i0; a.i1 = g.i1; a.j0 = g.j0; a.j1 = g.j1; a.k0 = g.k0; a.k1 = g.k1;
  a.S = S; a.R = R; a.CPL = cpl; a.nTy = nTy;
  a.spl = (cpl == 1 && gpitch(g) % 64 == 0 && !tune.no_split) ? 1 : 0;
  a.KC = kc; a.ntab = f.ntab; a.act_h = f.act_h; a.act_e = f.act_e;
  // The slab part meets the TF/SF corrections only where the box reaches its rows (j <= P,
  // j >= J-P; a row's H of row j-1 included) or planes (k < P, k >= K-P, and the plane K-P-1 below
  // the upper z-slab chunks): faces on the rows j0-1..j1+1 and the planes k0-1..k1+1.
  a.glob = f.force_global ? 1 : 0;
  a.gr = f.gr ? 1 : 0;
  a.ntf = !(f.act_h | f.act_e) ||
          (g.j0 >= g.P + 2 && g.j1 <= g.J - g.P - 2 && g.k0 >= g.P + 2 && g.k1 <= g.K - g.P - 2);
  a.KCS = kcs; a.nz1 = nz1; a.nz2 = nz2;
  a.KCY = kcy; a.nzy = nzy;
  a.Wrow = 64 * cpl * S;
  a.nTlo = nTlo; a.nThi = nThi; a.nTmid = nTmid;
  a.R0 = R0; a.nTmid0 = nTmid0;
  a.nwg0 = nTmid0 * nz2;
  a.nwg1 = (nTlo + nThi) * nKc + nTmid * 2 * nz1;
  a.xni = g.i1 - g.i0 + 1;
  a.xnj = g.j1 - g.j0 + 1;
  a.xnk = g.k1 - g.k0 + 1;
  a.Ip = (int)gpitch(g);
  a.IJK = gpitch(g) * g.J * g.K;
  a.cs = f.cs;
  a.sx = (size_t)2 * g.P * g.J * g.K; a.sy = (size_t)2 * g.I * g.P * g.K; a.sz = (size_t)2 * g.I * g.J * g.P;
  a.Fc = f.Fc; a.Fn = f.Fn; a.Pc = f.Pc; a.Pn = f.Pn;
  a.mat = f.mat; a.coef = f.coef; a.prm = f.prm;
  a.xs = f.xs;
  a.mu = f.Chmu;
  a.Chh = f.Chh; a.Che = f.Che; a.dx = f.dx;
  std::memcpy(out, &a, sizeof(a));
  return a.nwg0 + a.nwg1;
}

static SlabArgs slab_args(const FusedDev &f, int n) {
  SlabArgs s;
  const GridDev &g = f.g;
  const size_t IJK = gpitch(g) * g.J * g.K;  // extent of a per-cell array
  const size_t sx = (size_t)2 * g.P * g.J * g.K, sy = (size_t)2 * g.I * g.P * g.K, sz = (size_t)2 * g.I * g.J * g.P;
  s.g = g;
  for (int c = 0; c < 6; ++c) {
    s.cur.f[c] = const_cast<float *>(f.Fc) + c * f.cs;
    s.nxt.f[c] = f.Fn + c * f.cs;
  }
  auto psi = [&](float *base, PsiSet &ps) {  // interleaved groups: element t of member m at 4 t + m
    float *px = base, *py = px + 4 * sx, *pz = py + 4 * sy;
    ps.p[P_EYX] = px; ps.p[P_EZX] = px + 1; ps.p[P_HYX] = px + 2; ps.p[P_HZX] = px + 3;
    ps.p[P_EXY] = py; ps.p[P_EZY] = py + 1; ps.p[P_HXY] = py + 2; ps.p[P_HZY] = py + 3;
    ps.p[P_EXZ] = pz; ps.p[P_EYZ] = pz + 1; ps.p[P_HXZ] = pz + 2; ps.p[P_HYZ] = pz + 3;
  };
  psi(const_cast<float *>(f.Pc), s.pcur);
  psi(f.Pn, s.pnxt);
  CoefDev &co = s.co;
  co.mat = f.mat;
  co.tab = f.mat ? f.coef : nullptr;
  co.ntab = f.ntab;
  for (int c = 0; c < 3; ++c) {
    co.Ce[c] = f.mat ? nullptr : f.coef + c * IJK;
    co.Ch[c] = f.mat ? nullptr : f.coef + (3 + c) * IJK;
  }
  const size_t P2 = 2 * (size_t)g.P;
  co.bE = f.prm; co.cE = f.prm + P2; co.bH = f.prm + 2 * P2; co.cH = f.prm + 3 * P2;
  co.ikE[0] = f.prm + 4 * P2; co.ikE[1] = co.ikE[0] + g.I; co.ikE[2] = co.ikE[1] + g.J;
  co.ikH[0] = co.ikE[2] + g.K; co.ikH[1] = co.ikH[0] + g.I; co.ikH[2] = co.ikH[1] + g.J;
  co.Chh = f.Chh; co.Che = f.Che; co.dx = f.dx;
  PWDev &pw = s.pw;
  pw.active_h = f.act_h; pw.active_e = f.act_e; pw.N1 = f.N1; pw.Eh = f.Eh; pw.Hh = f.Hh;
  if (f.pwtab) {
    const size_t tl = 2 * ((size_t)g.I + g.J + g.K);
    const size_t off[3] = {0, 2 * (size_t)g.I, 2 * (size_t)g.I + 2 * (size_t)g.J};
    const int Q[3] = {g.I, g.J, g.K};
    for (int u = 0; u < 3; ++u)
      for (int h = 0; h < 2; ++h) {
        pw.tabE[u][h] = f.pwtab + off[u] + (size_t)h * Q[u];
        pw.tabH[u][h] = f.pwtab + tl + off[u] + (size_t)h * Q[u];
      }
  }
  for (int c = 0; c < 3; ++c) { pw.polE[c] = f.polE[c]; pw.polH[c] = f.polH[c]; }
  s.n = n;
  s.xs = f.xs;
  s.xni = g.i1 - g.i0 + 1;
  s.xnj = g.j1 - g.j0 + 1;
  s.xnk = g.k1 - g.k0 + 1;
  return s;
}

RowsKey rows_key(const void *hargs) {
  const StepArgs &a = *(const StepArgs *)hargs;
  const size_t lim = ((size_t)1 << 32) - 1;
  RowsKey k;
  k.idx = a.mat != nullptr;
  k.buf = !a.glob && 24 * a.cs < lim && 16 * (a.sx + a.sy + a.sz) < lim && (a.mat ? a.IJK : 24 * a.IJK) < lim;
  k.cpl = a.CPL;
  k.threads = 64 * a.S * a.R;
  k.threads0 = 64 * a.S * a.R0;
  k.lds = rows_lds(a, 1);
  k.lds0 = rows_lds(a, 0);
  k.ntf = a.ntf != 0;
  return k;
}

size_t slab_args_size() { return sizeof(SlabArgs); }

void make_slab_args(const FusedDev &f, void *out, unsigned *nb_sheets) {
  const SlabArgs sa = slab_args(f, 0);
  const bool sheets = f.xs && (f.act_h | f.act_e) && sa.xni > 0 && sa.xnj > 0 && sa.xnk > 0;
  const long nsheet = 8L * ((long)sa.xnj * sa.xnk + (long)sa.xni * sa.xnk + (long)sa.xni * sa.xnj);
  std::memcpy(out, &sa, sizeof(sa));
  *nb_sheets = sheets ? cdiv(nsheet, 256) : 0;
}

template <bool IDX, bool BUF, int PART>
static void launch_rows_batch_w(hipStream_t st, int cpl, dim3 grid, dim3 blk, unsigned lds, const BatchGrid *tab, int ng,
                                int step, bool ntf) {
  if constexpr (BUF && PART == 1) {  // every grid of the launch has its TF/SF box clear of the slab part
    if (ntf && cpl == 1) {
      hipLaunchKernelGGL((k_rows_batch<IDX, BUF, 1, PART, true>), grid, blk, lds, st, tab, ng, step);
      return;
    }
  }
  if (cpl == 4) hipLaunchKernelGGL((k_rows_batch<IDX, BUF, 4, PART>), grid, blk, lds, st, tab, ng, step);
  else if (cpl == 2) hipLaunchKernelGGL((k_rows_batch<IDX, BUF, 2, PART>), grid, blk, lds, st, tab, ng, step);
  else hipLaunchKernelGGL((k_rows_batch<IDX, BUF, 1, PART>), grid, blk, lds, st, tab, ng, step);
}

template <int PART>
static void launch_rows_batch_p(hipStream_t st, const RowsKey &key, dim3 grid, dim3 blk, const BatchGrid *tab, int ng,
                                int step) {
  if (key.idx) {
    if (key.buf) launch_rows_batch_w<true, true, PART>(st, key.cpl, grid, blk, PART == 0 ? key.lds0 : key.lds, tab, ng, step, key.ntf);
    else launch_rows_batch_w<true, false, PART>(st, key.cpl, grid, blk, PART == 0 ? key.lds0 : key.lds, tab, ng, step, key.ntf);
  } else {
    if (key.buf) launch_rows_batch_w<false, true, PART>(st, key.cpl, grid, blk, PART == 0 ? key.lds0 : key.lds, tab, ng, step, key.ntf);
    else launch_rows_batch_w<false, false, PART>(st, key.cpl, grid, blk, PART == 0 ? key.lds0 : key.lds, tab, ng, step, key.ntf);
  }
}

// Resident workgroups of a queue launch: the occupancy of the instantiation at this workgroup size
// times the CU count (looked up once per instantiation and size).
template <typename F>
static int queue_capacity(F kernel, int threads, unsigned lds) {
  static std::mutex mu;
  static std::map<std::tuple<const void *, int, unsigned>, int> cache;
  const auto key = std::make_tuple((const void *)kernel, threads, lds);
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, lds) != hipSuccess || per < 1)
    return 0;
  return cache[key] = cus * per;
}

template <bool IDX, bool BUF, int CPL, int PART>
static hipError_t launch_rows_queue_k(hipStream_t st, int threads, unsigned lds, const BatchGrid *tab, int ng, int nwg,
                                      int step, int *queue, int slot, int share) {
  auto k = k_rows_queue<IDX, BUF, CPL, PART>;
  const int cap = queue_capacity(k, threads, lds);
  if (cap <= 0) return hipErrorInvalidConfiguration;
  // this launch's share of the resident workgroups (launches of other streams run beside it), a
  // multiple of 8 so that every XCD queue has workgroups
  const int want = std::max(8, cap / std::max(share, 1) / 8 * 8);
  const dim3 grid(std::min((nwg + 7) / 8 * 8, want)), blk(threads);
  hipLaunchKernelGGL(k, grid, blk, lds, st, tab, ng, step, queue, slot, nwg);
  return hipGetLastError();
}

template <bool IDX, bool BUF, int PART>
static hipError_t launch_rows_queue_c(hipStream_t st, int cpl, int threads, unsigned lds, const BatchGrid *tab, int ng,
                                      int nwg, int step, int *queue, int slot, int share) {
  if (cpl == 4) return launch_rows_queue_k<IDX, BUF, 4, PART>(st, threads, lds, tab, ng, nwg, step, queue, slot, share);
  if (cpl == 2) return launch_rows_queue_k<IDX, BUF, 2, PART>(st, threads, lds, tab, ng, nwg, step, queue, slot, share);
  return launch_rows_queue_k<IDX, BUF, 1, PART>(st, threads, lds, tab, ng, nwg, step, queue, slot, share);
}

hipError_t launch_rows_queue(hipStream_t st, const RowsKey &key, int part, const BatchGrid *tab, int ng, int nwg,
                             int step, int *queue, int slot, int share) {
  if (nwg <= 0) return hipSuccess;
  const int threads = part == 0 ? key.threads0 : key.threads;
  const unsigned lds = part == 0 ? key.lds0 : key.lds;
  if (key.idx) {
    if (key.buf) return part == 0 ? launch_rows_queue_c<true, true, 0>(st, key.cpl, threads, lds, tab, ng, nwg, step, queue, slot, share)
                                  : launch_rows_queue_c<true, true, 1>(st, key.cpl, threads, lds, tab, ng, nwg, step, queue, slot, share);
    return part == 0 ? launch_rows_queue_c<true, false, 0>(st, key.cpl, threads, lds, tab, ng, nwg, step, queue, slot, share)
                     : launch_rows_queue_c<true, false, 1>(st, key.cpl, threads, lds, tab, ng, nwg, step, queue, slot, share);
  }
  if (key.buf) return part == 0 ? launch_rows_queue_c<false, true, 0>(st, key.cpl, threads, lds, tab, ng, nwg, step, queue, slot, share)
                                : launch_rows_queue_c<false, true, 1>(st, key.cpl, threads, lds, tab, ng, nwg, step, queue, slot, share);
  return part == 0 ? launch_rows_queue_c<false, false, 0>(st, key.cpl, threads, lds, tab, ng, nwg, step, queue, slot, share)
                   : launch_rows_queue_c<false, false, 1>(st, key.cpl, threads, lds, tab, ng, nwg, step, queue, slot, share);
}

hipError_t launch_rows_batch(hipStream_t st, const RowsKey &key, int part, const BatchGrid *tab, int ng, int nwg,
                             int step) {
  if (nwg <= 0) return hipSuccess;
  const dim3 grid(nwg), blk(part == 0 ? key.threads0 : key.threads);
  if (part == 0) launch_rows_batch_p<0>(st, key, grid, blk, tab, ng, step);
  else launch_rows_batch_p<1>(st, key, grid, blk, tab, ng, step);
  return hipGetLastError();
}

int rows_part_nwg(const void *hargs, int part) {
  const StepArgs &a = *(const StepArgs *)hargs;
  return part == 0 ? a.nwg0 : a.nwg1;
}

hipError_t launch_sheets_batch(hipStream_t st, const BatchGrid *tab, int ng, int nwg, int step) {
  if (nwg <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sheets_batch, dim3(nwg), dim3(256), 0, st, tab, ng, step);
  return hipGetLastError();
}

hipError_t launch_probes_batch(hipStream_t st, const ProbeJob *jobs, int njobs, int max_ant, int step) {
  if (njobs <= 0 || max_ant <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_probes_batch, dim3(cdiv(max_ant, 64), njobs), dim3(64), 0, st, jobs, step);
  return hipGetLastError();
}

hipError_t launch_sheets(hipStream_t st, const FusedDev &f, int n) {
  const SlabArgs sa = slab_args(f, n);
  const bool sheets = f.xs && (f.act_h | f.act_e) && sa.xni > 0 && sa.xnj > 0 && sa.xnk > 0;
  if (!sheets) return hipSuccess;
  const long nsheet = 8L * ((long)sa.xnj * sa.xnk + (long)sa.xni * sa.xnk + (long)sa.xni * sa.xnj);
  hipLaunchKernelGGL(k_sheets, dim3(cdiv(nsheet, 256)), dim3(256), 0, st, sa);
  return hipGetLastError();
}

// One time step: the TF/SF face sheets of step n (when a plane wave is active), then the row
// kernel over every cell: its interior part, then its slab part.  ev_rows (optional): two events
// recorded around the interior part (the step's dominant launch).  fold (optional, fdtd_run): the
// sheets of step n may be filled already, and the slab part's launch carries the sheets of step
// n + 1 (into fold->xs_next) and the probe gather of the state the step reads (record row
// fold->row) as tail workgroups.
hipError_t launch_step_fused(hipStream_t st, const FusedDev &f, const void *dargs, const void *hargs, int nwg, int n,
                             hipEvent_t *ev_rows, StepFold *fold) {
  const SlabArgs sa = slab_args(f, n);
  const bool sheets = f.xs && (f.act_h | f.act_e) && sa.xni > 0 && sa.xnj > 0 && sa.xnk > 0;
  const long nsheet = 8L * ((long)sa.xnj * sa.xnk + (long)sa.xni * sa.xnk + (long)sa.xni * sa.xnj);
  const unsigned nbx = sheets ? cdiv(nsheet, 256) : 0;
  if (nbx && !(fold && fold->sheets_ready)) hipLaunchKernelGGL(k_sheets, dim3(nbx), dim3(256), 0, st, sa);
  if (ev_rows) {
    const hipError_t e = hipEventRecord(ev_rows[0], st);
    if (e != hipSuccess) return e;
  }
  const StepArgs &a = *(const StepArgs *)hargs;
  if (nwg > 0) {
    const hipError_t e = launch_rows(st, a, (const StepArgs *)dargs, 0, n);
    if (e != hipSuccess) return e;
  }
  if (ev_rows) {
    const hipError_t e = hipEventRecord(ev_rows[1], st);
    if (e != hipSuccess) return e;
  }
  TailArgs ta{};
  if (fold) {
    fold->next_filled = false;
    fold->cur_done = false;
    const unsigned nthr = 64u * a.S * a.R;
    if (nwg <= 0) {  // no row launch to carry the tail
      if (fold->row >= 0)
        return launch_probes(st, fold->F, fold->idx, fold->nAnt, fold->comps, fold->ncomp, fold->port, fold->row, fold->T);
      return hipGetLastError();
    }
    ta.idx = fold->gidx;
    ta.T = fold->T;
    if (sheets && fold->xs_next) {
      ta.g = sa.g; ta.pw = sa.pw;
      ta.xs = fold->xs_next; ta.xni = sa.xni; ta.xnj = sa.xnj; ta.xnk = sa.xnk;
      ta.nsh = fold->gidx ? fold->gq + 1 : n + 1;
      ta.nb_sheets = std::min(cdiv(nsheet, nthr), 1024u);
      fold->next_filled = true;
    }
    const bool rec = fold->gidx ? fold->probe_rel : (fold->row >= 0 && fold->row < fold->T);
    if (rec && fold->nAnt > 0 && fold->ncomp > 0) {
      ta.F = fold->F; ta.pidx = fold->idx; ta.pcomps = fold->comps; ta.port = fold->port;
      ta.nAnt = fold->nAnt; ta.ncomp = fold->ncomp; ta.prow = fold->gidx ? fold->gq - 1 : fold->row;
      ta.nb_probes = std::min(cdiv((long)fold->nAnt * fold->ncomp, nthr), 64u);
    }
    if (fold->cur_comp >= 0) {
      const GridDev &g = f.g;
      int a0, a1, b0, b1, c0, c1;
      if (!current_box(g, fold->cur_box, a0, a1, b0, b1, c0, c1)) {
        fold->cur_done = true;  // an empty box adds nothing
      } else if (b0 >= g.P + 1 && b1 <= g.J - g.P - 1 && c0 >= g.P && c1 <= g.K - g.P - 1) {
        ta.g = g; ta.Fn = sa.nxt; ta.co = sa.co;
        ta.comp = fold->cur_comp; ta.a0 = a0; ta.b0 = b0; ta.c0 = c0;
        ta.ni = a1 - a0 + 1; ta.nj = b1 - b0 + 1; ta.nk = c1 - c0 + 1;
        ta.s = fold->cur_s; ta.dx = fold->cur_dx; ta.src = fold->cur_src; ta.dn = fold->cur_dn; ta.q = fold->cur_q;
        ta.nb_cur = std::min(cdiv((long)ta.ni * ta.nj * ta.nk, nthr), 64u);
        fold->cur_done = true;
      }
    }
  }
  if (nwg > 0) {
    const hipError_t e = launch_rows(st, a, (const StepArgs *)dargs, 1, n, &ta);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

}  // namespace fdtd
