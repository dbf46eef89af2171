// fdtd_kernels.hip -- gfx950 kernels of the 3D Yee FDTD library.
//
// Two families:
//  * per-op kernels, one per op of the reference MEX gateway, updating the state in place with
//    the reference's masks and operation order (src/gpu.h);
//  * k_step_fused: one launch per time step.  Workgroups own (row tile, z chunk) blocks, one
//    wave per grid row (x fastest, 64 lanes x NC chunks, coalesced), and march up the z chunk.
//    For each plane k the workgroup computes H(k) at n+1/2 (main update, TF/SF correction, CPML)
//    from E(k), E(k+1) and H(k), stages it in LDS, then E(k) at n+1 from H(k), H(k-1) (kept in
//    registers) and E(k).  The state is ping-ponged (read cur, write nxt), so a workgroup can
//    recompute the H values of its halo row / halo plane from old data without racing its
//    neighbours.  Each field is read once and written once per step.
//
// All arithmetic is plain IEEE float with no contraction (-ffp-contract=off); expressions keep the
// reference's association order so results equal the CPU oracle bit for bit.
#include <hip/hip_runtime.h>

#include <cstring>

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

// All TF/SF corrections of one H cell, faces in axis order x, y, z.
__device__ __forceinline__ void pw_h_cell(const GridDev &g, float Che, const PWDev &pw, const float *row, int i, int j,
                                          int k, float &hx, float &hy, float &hz) {
  if (i < g.i0 - 1 || i > g.i1 || j < g.j0 - 1 || j > g.j1 || k < g.k0 - 1 || k > g.k1) return;
  if (i == g.i0 - 1 || i == g.i1) pw_h_face<0>(g, Che, pw, row, i, j, k, i == g.i0 - 1, hx, hy, hz);
  if (j == g.j0 - 1 || j == g.j1) pw_h_face<1>(g, Che, pw, row, i, j, k, j == g.j0 - 1, hx, hy, hz);
  if (k == g.k0 - 1 || k == g.k1) pw_h_face<2>(g, Che, pw, row, i, j, k, k == g.k0 - 1, hx, hy, hz);
}

__device__ __forceinline__ void pw_e_cell(const GridDev &g, const float ch[3], const PWDev &pw, const float *row, int i,
                                          int j, int k, float &ex, float &ey, float &ez) {
  if (i < g.i0 || i > g.i1 || j < g.j0 || j > g.j1 || k < g.k0 || k > g.k1) return;
  if (i == g.i0 || i == g.i1) pw_e_face<0>(g, ch, pw, row, i, j, k, i == g.i0, ex, ey, ez);
  if (j == g.j0 || j == g.j1) pw_e_face<1>(g, ch, pw, row, i, j, k, j == g.j0, ex, ey, ez);
  if (k == g.k0 || k == g.k1) pw_e_face<2>(g, ch, pw, row, i, j, k, k == g.k0, ex, ey, ez);
}

// Material coefficients of cell id (column i).  Index mode: table row of the cell's material,
// with Ch = 0 on the last x plane (den_ex = 0 there, src/simulation.h:887-912).
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
// H-cell and E-cell updates shared by the fused kernel (main update, TF/SF, CPML in the
// reference's order).  Neighbour values are passed in; psi is read from pc and, when `own`, the
// new psi (or an unchanged copy for slab cells outside the update masks) is written to pn.
// ------------------------------------------------------------------------------------------
struct HNb {
  float ex, ey, ez;         // E(i,j,k)
  float ex_j1, ez_j1;       // E(i,j+1,k)
  float ey_i1, ez_i1;       // E(i+1,j,k)
  float ex_k1, ey_k1;       // E(i,j,k+1)
};
struct ENb {
  float hx, hy, hz;         // H(i,j,k) at n+1/2
  float hx_jm, hz_jm;       // H(i,j-1,k)
  float hy_im, hz_im;       // H(i-1,j,k)
  float hx_km, hy_km;       // H(i,j,k-1)
};

__device__ __forceinline__ void h_cell(const GridDev &g, const CoefDev &co, const PWDev &pw, const float *rowE,
                                       const PsiSet &pc, const PsiSet &pn, bool own, int i, int j, int k, size_t IJ,
                                       const HNb &e, float &hx, float &hy, float &hz) {
  const int I = g.I, J = g.J, K = g.K, P = g.P;
  const float Che = co.Che, dx = co.dx;
  // K1 GPU_updateH, src/gpu.h:592-601
  if (i < I - 1 && j < J - 1 && k < K - 1) {
    const float kx = co.ikH[0][i], ky = co.ikH[1][j], kz = co.ikH[2][k];
    hx = co.Chh * hx + Che * (-(e.ez_j1 - e.ez) * ky + (e.ey_k1 - e.ey) * kz);
    hy = co.Chh * hy + Che * (-(e.ex_k1 - e.ex) * kz + (e.ez_i1 - e.ez) * kx);
    hz = co.Chh * hz + Che * (-(e.ey_i1 - e.ey) * kx + (e.ex_j1 - e.ex) * ky);
  }
  // updateH_PlaneWave (src/simulation.h:330-349)
  if (pw.active_h) pw_h_cell(g, Che, pw, rowE, i, j, k, hx, hy, hz);
  // updateH_PML: Y (K5), X (K4), Z (K6), src/simulation.h:392-403
  if (j < P || j >= J - P) {
    const int s = j < P ? j : j - (J - 2 * P);
    const size_t t = (size_t)i + (size_t)I * s + (size_t)I * 2 * P * k;
    if (s < 2 * P - 1 && i < I - 1 && k < K - 1) {
      const float pxy = co.bH[s] * pc.p[P_HXY][t] + co.cH[s] * (e.ez_j1 - e.ez) / dx;
      hx = hx - Che * pxy * dx;
      const float pzy = co.bH[s] * pc.p[P_HZY][t] + co.cH[s] * (e.ex_j1 - e.ex) / dx;
      hz = hz + Che * pzy * dx;
      if (own) { pn.p[P_HXY][t] = pxy; pn.p[P_HZY][t] = pzy; }
    } else if (own) {
      pn.p[P_HXY][t] = pc.p[P_HXY][t];
      pn.p[P_HZY][t] = pc.p[P_HZY][t];
    }
  }
  if (i < P || i >= I - P) {
    const int s = i < P ? i : i - (I - 2 * P);
    const size_t t = (size_t)s + (size_t)2 * P * j + (size_t)2 * P * J * k;
    if (s < 2 * P - 1 && j < J - 1 && k < K - 1) {
      const float pyx = co.bH[s] * pc.p[P_HYX][t] + co.cH[s] * (e.ez_i1 - e.ez) / dx;
      hy = hy + Che * pyx * dx;
      const float pzx = co.bH[s] * pc.p[P_HZX][t] + co.cH[s] * (e.ey_i1 - e.ey) / dx;
      hz = hz - Che * pzx * dx;
      if (own) { pn.p[P_HYX][t] = pyx; pn.p[P_HZX][t] = pzx; }
    } else if (own) {
      pn.p[P_HYX][t] = pc.p[P_HYX][t];
      pn.p[P_HZX][t] = pc.p[P_HZX][t];
    }
  }
  if (k < P || k >= K - P) {
    const int s = k < P ? k : k - (K - 2 * P);
    const size_t t = (size_t)i + (size_t)I * j + IJ * s;
    if (s < 2 * P - 1 && i < I - 1 && j < J - 1) {
      const float pxz = co.bH[s] * pc.p[P_HXZ][t] + co.cH[s] * (e.ey_k1 - e.ey) / dx;
      hx = hx + Che * pxz * dx;
      const float pyz = co.bH[s] * pc.p[P_HYZ][t] + co.cH[s] * (e.ex_k1 - e.ex) / dx;
      hy = hy - Che * pyz * dx;
      if (own) { pn.p[P_HXZ][t] = pxz; pn.p[P_HYZ][t] = pyz; }
    } else if (own) {
      pn.p[P_HXZ][t] = pc.p[P_HXZ][t];
      pn.p[P_HYZ][t] = pc.p[P_HYZ][t];
    }
  }
}

__device__ __forceinline__ void e_cell(const GridDev &g, const CoefDev &co, const PWDev &pw, const float *rowH,
                                       const PsiSet &pc, const PsiSet &pn, int i, int j, int k, size_t IJ,
                                       const float ce[3], const float ch[3], const ENb &h, float &ex, float &ey,
                                       float &ez) {
  const int I = g.I, J = g.J, K = g.K, P = g.P;
  const float dx = co.dx;
  // K3 GPU_updateE, src/gpu.h:745-762
  {
    const float kx = co.ikE[0][i], ky = co.ikE[1][j], kz = co.ikE[2][k];
    if (j != 0 && k != 0) ex = ce[0] * ex + ch[0] * ((h.hz - h.hz_jm) * ky - (h.hy - h.hy_km) * kz);
    if (i != 0 && k != 0) ey = ce[1] * ey + ch[1] * ((h.hx - h.hx_km) * kz - (h.hz - h.hz_im) * kx);
    if (i != 0 && j != 0) ez = ce[2] * ez + ch[2] * ((h.hy - h.hy_im) * kx - (h.hx - h.hx_jm) * ky);
  }
  // updateE_PlaneWave (src/simulation.h:350-369)
  if (pw.active_e) pw_e_cell(g, ch, pw, rowH, i, j, k, ex, ey, ez);
  // updateE_PML: Y (K8), Z (K9), X (K7), src/simulation.h:404-418
  if (j < P || j >= J - P) {
    const int s = j < P ? j : j - (J - 2 * P);
    const size_t t = (size_t)i + (size_t)I * s + (size_t)I * 2 * P * k;
    if (s >= 1 && i >= 1 && i < I - 1 && k >= 1 && k < K - 1) {
      const float pxy = co.bE[s] * pc.p[P_EXY][t] + co.cE[s] * (h.hz - h.hz_jm) / dx;
      ex = ex + ch[0] * pxy * dx;
      const float pzy = co.bE[s] * pc.p[P_EZY][t] + co.cE[s] * (h.hx - h.hx_jm) / dx;
      ez = ez - ch[2] * pzy * dx;
      pn.p[P_EXY][t] = pxy;
      pn.p[P_EZY][t] = pzy;
    } else {
      pn.p[P_EXY][t] = pc.p[P_EXY][t];
      pn.p[P_EZY][t] = pc.p[P_EZY][t];
    }
  }
  if (k < P || k >= K - P) {
    const int s = k < P ? k : k - (K - 2 * P);
    const size_t t = (size_t)i + (size_t)I * j + IJ * s;
    if (s >= 1 && i >= 1 && i < I - 1 && j >= 1 && j < J - 1) {
      const float pxz = co.bE[s] * pc.p[P_EXZ][t] + co.cE[s] * (h.hy - h.hy_km) / dx;
      ex = ex - ch[0] * pxz * dx;
      const float pyz = co.bE[s] * pc.p[P_EYZ][t] + co.cE[s] * (h.hx - h.hx_km) / dx;
      ey = ey + ch[1] * pyz * dx;
      pn.p[P_EXZ][t] = pxz;
      pn.p[P_EYZ][t] = pyz;
    } else {
      pn.p[P_EXZ][t] = pc.p[P_EXZ][t];
      pn.p[P_EYZ][t] = pc.p[P_EYZ][t];
    }
  }
  if (i < P || i >= I - P) {
    const int s = i < P ? i : i - (I - 2 * P);
    const size_t t = (size_t)s + (size_t)2 * P * j + (size_t)2 * P * J * k;
    if (s >= 1 && j >= 1 && j < J - 1 && k >= 1 && k < K - 1) {
      const float pyx = co.bE[s] * pc.p[P_EYX][t] + co.cE[s] * (h.hz - h.hz_im) / dx;
      ey = ey - ch[1] * pyx * dx;
      const float pzx = co.bE[s] * pc.p[P_EZX][t] + co.cE[s] * (h.hy - h.hy_im) / dx;
      ez = ez + ch[2] * pzx * dx;
      pn.p[P_EYX][t] = pyx;
      pn.p[P_EZX][t] = pzx;
    } else {
      pn.p[P_EYX][t] = pc.p[P_EYX][t];
      pn.p[P_EZX][t] = pc.p[P_EZX][t];
    }
  }
}

// ------------------------------------------------------------------------------------------
// per-op kernels (in place)
// ------------------------------------------------------------------------------------------

// K1 GPU_updateH, src/gpu.h:575-602
__global__ void __launch_bounds__(256) k_h_all(GridDev g, FieldSet F, CoefDev co) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= g.I - 1 || j >= g.J - 1 || k >= g.K - 1) return;
  const size_t I = g.I, IJ = (size_t)g.I * g.J;
  const size_t id = (size_t)i + (size_t)j * I + (size_t)k * IJ;
  const float *Ex = F.f[EX], *Ey = F.f[EY], *Ez = F.f[EZ];
  float *Hx = F.f[HX], *Hy = F.f[HY], *Hz = F.f[HZ];
  const float ez = Ez[id], ey = Ey[id], ex = Ex[id];
  const float kx = co.ikH[0][i], ky = co.ikH[1][j], kz = co.ikH[2][k];
  Hx[id] = co.Chh * Hx[id] + co.Che * (-(Ez[id + I] - ez) * ky + (Ey[id + IJ] - ey) * kz);
  Hy[id] = co.Chh * Hy[id] + co.Che * (-(Ex[id + IJ] - ex) * kz + (Ez[id + 1] - ez) * kx);
  Hz[id] = co.Chh * Hz[id] + co.Che * (-(Ey[id + 1] - ey) * kx + (Ex[id + I] - ex) * ky);
}

// K3 GPU_updateE, src/gpu.h:727-763
template <bool IDX>
__global__ void __launch_bounds__(256) k_e_all(GridDev g, FieldSet F, CoefDev co) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= g.I || j >= g.J) return;
  const size_t I = g.I, IJ = (size_t)g.I * g.J;
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
  const size_t IJ = (size_t)I * J;
  const size_t t = (size_t)i + (size_t)I * s + (size_t)I * 2 * P * k;
  const int jj = s < P ? s : s + J - 2 * P;
  const size_t id = (size_t)i + (size_t)jj * I + (size_t)k * IJ;
  const float dx = co.dx;
  float *pxy = ps.p[P_HXY], *pzy = ps.p[P_HZY];
  pxy[t] = co.bH[s] * pxy[t] + co.cH[s] * (F.f[EZ][id + I] - F.f[EZ][id]) / dx;
  F.f[HX][id] = F.f[HX][id] - co.Che * pxy[t] * dx;
  pzy[t] = co.bH[s] * pzy[t] + co.cH[s] * (F.f[EX][id + I] - F.f[EX][id]) / dx;
  F.f[HZ][id] = F.f[HZ][id] + co.Che * pzy[t] * dx;
}

// K4 GPU_updateHPmlX, src/gpu.h:604-642: s in [0, 2P-1), j < J-1, k < K-1
__global__ void __launch_bounds__(256) k_h_pml_x(GridDev g, FieldSet F, PsiSet ps, CoefDev co) {
  const int I = g.I, J = g.J, P = g.P;
  const int q = blockIdx.x * 256 + threadIdx.x, k = blockIdx.y;
  const int ns = 2 * P - 1;
  if (q >= ns * (J - 1)) return;
  const int s = q % ns, j = q / ns;
  const size_t IJ = (size_t)I * J;
  const size_t t = (size_t)s + (size_t)2 * P * j + (size_t)2 * P * J * k;
  const int ii = s < P ? s : I - 2 * P + s;
  const size_t id = (size_t)ii + (size_t)j * I + (size_t)k * IJ;
  const float dx = co.dx;
  float *pyx = ps.p[P_HYX], *pzx = ps.p[P_HZX];
  pyx[t] = co.bH[s] * pyx[t] + co.cH[s] * (F.f[EZ][id + 1] - F.f[EZ][id]) / dx;
  F.f[HY][id] = F.f[HY][id] + co.Che * pyx[t] * dx;
  pzx[t] = co.bH[s] * pzx[t] + co.cH[s] * (F.f[EY][id + 1] - F.f[EY][id]) / dx;
  F.f[HZ][id] = F.f[HZ][id] - co.Che * pzx[t] * dx;
}

// K6 GPU_updateHPmlZ, src/gpu.h:687-725: i < I-1, j < J-1, s in [0, 2P-1)
__global__ void __launch_bounds__(256) k_h_pml_z(GridDev g, FieldSet F, PsiSet ps, CoefDev co) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, s = blockIdx.z;
  const int I = g.I, J = g.J, K = g.K, P = g.P;
  if (i >= I - 1 || j >= J - 1) return;
  const size_t IJ = (size_t)I * J;
  const size_t t = (size_t)i + (size_t)I * j + IJ * s;
  const int kk = s < P ? s : s + K - 2 * P;
  const size_t id = (size_t)i + (size_t)j * I + (size_t)kk * IJ;
  const float dx = co.dx;
  float *pxz = ps.p[P_HXZ], *pyz = ps.p[P_HYZ];
  pxz[t] = co.bH[s] * pxz[t] + co.cH[s] * (F.f[EY][id + IJ] - F.f[EY][id]) / dx;
  F.f[HX][id] = F.f[HX][id] + co.Che * pxz[t] * dx;
  pyz[t] = co.bH[s] * pyz[t] + co.cH[s] * (F.f[EX][id + IJ] - F.f[EX][id]) / dx;
  F.f[HY][id] = F.f[HY][id] - co.Che * pyz[t] * dx;
}

// K8 GPU_updateEPmlY, src/gpu.h:790-827: 0 < i < I-1, s in [1, 2P), 1 <= k < K-1
template <bool IDX>
__global__ void __launch_bounds__(256) k_e_pml_y(GridDev g, FieldSet F, PsiSet ps, CoefDev co) {
  const int i = blockIdx.x * 64 + threadIdx.x, s = blockIdx.y * 4 + threadIdx.y + 1, k = blockIdx.z + 1;
  const int I = g.I, J = g.J, P = g.P;
  if (i < 1 || i >= I - 1 || s >= 2 * P) return;
  const size_t IJ = (size_t)I * J;
  const size_t t = (size_t)i + (size_t)I * s + (size_t)I * 2 * P * k;
  const int jj = s < P ? s : s + J - 2 * P;
  const size_t id = (size_t)i + (size_t)jj * I + (size_t)k * IJ;
  float ce[3], ch[3];
  cell_coef<IDX>(co, co.tab, id, i, I, ce, ch);
  const float dx = co.dx;
  float *pxy = ps.p[P_EXY], *pzy = ps.p[P_EZY];
  pxy[t] = co.bE[s] * pxy[t] + co.cE[s] * (F.f[HZ][id] - F.f[HZ][id - I]) / dx;
  F.f[EX][id] = F.f[EX][id] + ch[0] * pxy[t] * dx;
  pzy[t] = co.bE[s] * pzy[t] + co.cE[s] * (F.f[HX][id] - F.f[HX][id - I]) / dx;
  F.f[EZ][id] = F.f[EZ][id] - ch[2] * pzy[t] * dx;
}

// K9 GPU_updateEPmlZ, src/gpu.h:829-868: 0 < i < I-1, 0 < j < J-1, s in [1, 2P)
template <bool IDX>
__global__ void __launch_bounds__(256) k_e_pml_z(GridDev g, FieldSet F, PsiSet ps, CoefDev co) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, s = blockIdx.z + 1;
  const int I = g.I, J = g.J, K = g.K, P = g.P;
  if (i < 1 || i >= I - 1 || j < 1 || j >= J - 1) return;
  const size_t IJ = (size_t)I * J;
  const size_t t = (size_t)i + (size_t)I * j + IJ * s;
  const int kk = s < P ? s : s + K - 2 * P;
  const size_t id = (size_t)i + (size_t)j * I + (size_t)kk * IJ;
  float ce[3], ch[3];
  cell_coef<IDX>(co, co.tab, id, i, I, ce, ch);
  const float dx = co.dx;
  float *pxz = ps.p[P_EXZ], *pyz = ps.p[P_EYZ];
  pxz[t] = co.bE[s] * pxz[t] + co.cE[s] * (F.f[HY][id] - F.f[HY][id - IJ]) / dx;
  F.f[EX][id] = F.f[EX][id] - ch[0] * pxz[t] * dx;
  pyz[t] = co.bE[s] * pyz[t] + co.cE[s] * (F.f[HX][id] - F.f[HX][id - IJ]) / dx;
  F.f[EY][id] = F.f[EY][id] + ch[1] * pyz[t] * dx;
}

// K7 GPU_updateEPmlX, src/gpu.h:765-788: s in [1, 2P), 0 < j < J-1, 1 <= k < K-1
template <bool IDX>
__global__ void __launch_bounds__(256) k_e_pml_x(GridDev g, FieldSet F, PsiSet ps, CoefDev co) {
  const int I = g.I, J = g.J, P = g.P;
  const int q = blockIdx.x * 256 + threadIdx.x, k = blockIdx.y + 1;
  const int ns = 2 * P - 1;  // s = 1 .. 2P-1
  if (q >= ns * (J - 2)) return;
  const int s = q % ns + 1, j = q / ns + 1;
  const size_t IJ = (size_t)I * J;
  const size_t t = (size_t)s + (size_t)2 * P * j + (size_t)2 * P * J * k;
  const int ii = s < P ? s : s + I - 2 * P;
  const size_t id = (size_t)ii + (size_t)j * I + (size_t)k * IJ;
  float ce[3], ch[3];
  cell_coef<IDX>(co, co.tab, id, ii, I, ce, ch);
  const float dx = co.dx;
  float *pyx = ps.p[P_EYX], *pzx = ps.p[P_EZX];
  pyx[t] = co.bE[s] * pyx[t] + co.cE[s] * (F.f[HZ][id] - F.f[HZ][id - 1]) / dx;
  F.f[EY][id] = F.f[EY][id] - ch[1] * pyx[t] * dx;
  pzx[t] = co.bE[s] * pzx[t] + co.cE[s] * (F.f[HY][id] - F.f[HY][id - 1]) / dx;
  F.f[EZ][id] = F.f[EZ][id] + ch[2] * pzx[t] * dx;
}

// TF/SF corrections of face axis FA, both faces (one kernel per face axis, launched x, y, z).
// Threads cover the face's tangential box range [lo, hi] of the two other axes.
template <int FA>
__global__ void __launch_bounds__(256) k_pw_h(GridDev g, FieldSet F, CoefDev co, PWDev pw, int n) {
  const int u = blockIdx.x * 16 + threadIdx.x, v = blockIdx.y * 16 + threadIdx.y;
  int i, j, k;
  const float *row = pw.Eh + (size_t)n * pw.N1;
  const size_t I = g.I, IJ = (size_t)g.I * g.J;
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
  const size_t I = g.I, IJ = (size_t)g.I * g.J;
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
__global__ void __launch_bounds__(256) k_add_currents(GridDev g, FieldSet F, CoefDev co, int comp, int a0, int b0,
                                                      int c0, int ni, int nj, int nk, float s, float dx) {
  const long tid = (long)blockIdx.x * 256 + threadIdx.x;
  if (tid >= (long)ni * nj * nk) return;
  const int i = a0 + (int)(tid % ni), j = b0 + (int)((tid / ni) % nj), k = c0 + (int)(tid / ((long)ni * nj));
  const size_t id = (size_t)i + (size_t)j * g.I + (size_t)k * g.I * g.J;
  float ce[3], ch[3];
  cell_coef<IDX>(co, co.tab, id, i, g.I, ce, ch);
  float *E = F.f[comp];
  E[id] = E[id] - ch[comp] * s * dx / dx / dx;
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
// fused time step: three launches per step, reading the cur state and writing the nxt state
//
//  A  k_slab_h   one thread per H cell of the CPML slabs S (the cells within P of a face): main
//                update, TF/SF, CPML Y-X-Z with psi, in the reference order.
//  B  k_interior the interior box C = [P, I-P) x [P, J-P) x [P, K-P), fused H and E.  Tiles of
//                64 columns x TY rows x KC planes; waves 0..TY+1 own one row each (rows jb-1 and
//                jb+TY are halo rows), wave TY+2 covers the halo columns ib-1 and ib+64.  Each
//                thread stands for one cell (i, j) and marches k up the chunk: E(k) staged in LDS,
//                barrier, H(k) at n+1/2 for rows jb-1..jb+TY-1 and columns ib-1..ib+63 into LDS
//                (own cells also to HBM), barrier, E(k) at n+1 for own cells.  H of halo cells in C
//                is recomputed from the old state; halo cells in S take the new H written by A.
//                No CPML code in this kernel: its loop only streams fields, the material byte and
//                the next plane's prefetch.
//  C  k_slab_e   one thread per E cell of S: main update, TF/SF, CPML Y-Z-X with psi, reading the
//                new H of A and B.
// All three apply per cell exactly the operations of the per-op kernels, in the same order.
// ------------------------------------------------------------------------------------------
struct StepArgs {
  int I, J, K, P;
  int i0, i1, j0, j1, k0, k1;
  int nTx, nTy, KC, ntab;
  int act_h, act_e, N1, pad;
  size_t IJK, sx, sy, sz;
  size_t cs;  // distance between the six field components of a set (I*J*K plus a skew)
  const float *Fc;
  float *Fn;
  const float *Pc;
  float *Pn;
  const uint8_t *mat;
  const float *coef;   // index mode: [ntab][6]; full mode: [6][IJK]
  const float *prm;    // bE cE bH cH [2P each], 1/kappa E x,y,z, 1/kappa H x,y,z
  const float *pwtab;  // E tables x(int,half) y(int,half) z(int,half), then the same for H
  const float *Eh, *Hh;  // 1D line histories [T][N1]
  float Chh, Che, dx;
  float polE[3], polH[3];
};

// Pointers taken from the argument block are generic; casting them to the global address space
// lets the compiler emit global_load/store (counted by vmcnt only) instead of flat accesses,
// which also count against lgkmcnt and would drain the prefetched loads at every LDS wait.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gptr(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}

// Workgroup barrier ordering LDS only: global loads in flight (the next plane's prefetch) are not
// waited for, unlike __syncthreads(), whose fence also covers global memory.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// psi groups of one set: x slabs [Eyx Ezx Hyx Hzx], y slabs [Exy Ezy Hxy Hzy], z slabs [Exz Eyz Hxz Hyz]
struct PsiView {
  __attribute__((address_space(1))) const float *c;
  __attribute__((address_space(1))) float *n;
  size_t sx, sy, sz;
  __device__ __forceinline__ size_t off_x(int q) const { return (size_t)q * sx; }
  __device__ __forceinline__ size_t off_y(int q) const { return 4 * sx + (size_t)q * sy; }
  __device__ __forceinline__ size_t off_z(int q) const { return 4 * sx + 4 * sy + (size_t)q * sz; }
};

__device__ __forceinline__ float line_at_g(const float *rowp, float s) {
  const auto row = gptr(rowp);
  const float fl = floorf(s);
  const int m = (int)fl;
  const float w = s - fl;
  if (w == 0.0f) return row[m];
  return row[m] + w * (row[m + 1] - row[m]);
}

__device__ __forceinline__ float tab_e(const StepArgs &a, const float *row, int c, int x, int y, int z) {
  const auto t = gptr(a.pwtab);
  const float sx = t[(c == 0 ? a.I : 0) + x];
  const float sy = t[2 * a.I + (c == 1 ? a.J : 0) + y];
  const float sz = t[2 * a.I + 2 * a.J + (c == 2 ? a.K : 0) + z];
  return a.polE[c] * line_at_g(row, (sx + sy) + sz);
}
__device__ __forceinline__ float tab_h(const StepArgs &a, const float *row, int c, int x, int y, int z) {
  const auto t = gptr(a.pwtab + 2 * (a.I + a.J + a.K));
  const float sx = t[(c != 0 ? a.I : 0) + x];
  const float sy = t[2 * a.I + (c != 1 ? a.J : 0) + y];
  const float sz = t[2 * a.I + 2 * a.J + (c != 2 ? a.K : 0) + z];
  return a.polH[c] * line_at_g(row, (sx + sy) + sz);
}

struct F3 {
  float x, y, z;
};

// TF/SF corrections of an H cell (faces x, y, z in order), generic form of src/gpu.h:152-573.
// Out of line: only TF/SF surface cells take this path.
__device__ __noinline__ F3 fused_pw_h(const StepArgs *__restrict__ pa, int n, int i, int j, int k, F3 hv) {
  const StepArgs &a = *pa;
  const float *row = a.Eh + (size_t)n * a.N1;
  float hx = hv.x, hy = hv.y, hz = hv.z;
  const int m = a.act_h;
  const float Che = a.Che;
  if (i == a.i0 - 1 || i == a.i1) {
    const bool lo = i == a.i0 - 1;
    const int xe = lo ? a.i0 : a.i1;
    if ((m & (1 << 1)) && j >= a.j0 && j <= a.j1 && k >= a.k0 && k <= a.k1 - 1) {
      const float v = Che * tab_e(a, row, 2, xe, j, k);
      hy = lo ? hy - v : hy + v;
    }
    if ((m & (1 << 2)) && j >= a.j0 && j <= a.j1 - 1 && k >= a.k0 && k <= a.k1) {
      const float v = Che * tab_e(a, row, 1, xe, j, k);
      hz = lo ? hz + v : hz - v;
    }
  }
  if (j == a.j0 - 1 || j == a.j1) {
    const bool lo = j == a.j0 - 1;
    const int ye = lo ? a.j0 : a.j1;
    if ((m & (1 << 3)) && i >= a.i0 && i <= a.i1 && k >= a.k0 && k <= a.k1 - 1) {
      const float v = Che * tab_e(a, row, 2, i, ye, k);
      hx = lo ? hx + v : hx - v;
    }
    if ((m & (1 << 5)) && i >= a.i0 && i <= a.i1 - 1 && k >= a.k0 && k <= a.k1) {
      const float v = Che * tab_e(a, row, 0, i, ye, k);
      hz = lo ? hz - v : hz + v;
    }
  }
  if (k == a.k0 - 1 || k == a.k1) {
    const bool lo = k == a.k0 - 1;
    const int ze = lo ? a.k0 : a.k1;
    if ((m & (1 << 6)) && i >= a.i0 && i <= a.i1 && j >= a.j0 && j <= a.j1 - 1) {
      const float v = Che * tab_e(a, row, 1, i, j, ze);
      hx = lo ? hx - v : hx + v;
    }
    if ((m & (1 << 7)) && i >= a.i0 && i <= a.i1 - 1 && j >= a.j0 && j <= a.j1) {
      const float v = Che * tab_e(a, row, 0, i, j, ze);
      hy = lo ? hy + v : hy - v;
    }
  }
  return F3{hx, hy, hz};
}

__device__ __noinline__ F3 fused_pw_e(const StepArgs *__restrict__ pa, int n, F3 chv, int i, int j, int k, F3 ev) {
  const StepArgs &a = *pa;
  const float *row = a.Hh + (size_t)n * a.N1;
  const float ch[3] = {chv.x, chv.y, chv.z};
  float ex = ev.x, ey = ev.y, ez = ev.z;
  const int m = a.act_e;
  if (i == a.i0 || i == a.i1) {
    const bool lo = i == a.i0;
    const int xh = lo ? a.i0 - 1 : a.i1;
    if ((m & (1 << 2)) && j >= a.j0 && j <= a.j1 && k >= a.k0 && k <= a.k1 - 1) {
      const float v = ch[2] * tab_h(a, row, 1, xh, j, k);
      ez = lo ? ez - v : ez + v;
    }
    if ((m & (1 << 1)) && j >= a.j0 && j <= a.j1 - 1 && k >= a.k0 && k <= a.k1) {
      const float v = ch[1] * tab_h(a, row, 2, xh, j, k);
      ey = lo ? ey + v : ey - v;
    }
  }
  if (j == a.j0 || j == a.j1) {
    const bool lo = j == a.j0;
    const int yh = lo ? a.j0 - 1 : a.j1;
    if ((m & (1 << 3)) && i >= a.i0 && i <= a.i1 - 1 && k >= a.k0 && k <= a.k1) {
      const float v = ch[0] * tab_h(a, row, 2, i, yh, k);
      ex = lo ? ex - v : ex + v;
    }
    if ((m & (1 << 5)) && i >= a.i0 && i <= a.i1 && k >= a.k0 && k <= a.k1 - 1) {
      const float v = ch[2] * tab_h(a, row, 0, i, yh, k);
      ez = lo ? ez + v : ez - v;
    }
  }
  if (k == a.k0 || k == a.k1) {
    const bool lo = k == a.k0;
    const int zh = lo ? a.k0 - 1 : a.k1;
    if ((m & (1 << 6)) && i >= a.i0 && i <= a.i1 - 1 && j >= a.j0 && j <= a.j1) {
      const float v = ch[0] * tab_h(a, row, 1, i, j, zh);
      ex = lo ? ex + v : ex - v;
    }
    if ((m & (1 << 7)) && i >= a.i0 && i <= a.i1 && j >= a.j0 && j <= a.j1 - 1) {
      const float v = ch[1] * tab_h(a, row, 0, i, j, zh);
      ey = lo ? ey - v : ey + v;
    }
  }
  return F3{ex, ey, ez};
}

// Workgroups are dealt to the 8 XCDs round-robin; remap so that each XCD walks a contiguous range
// of tiles (neighbouring tiles of one z chunk share halo rows and columns in that XCD's L2).
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int per = nb / 8, rem = nb % 8;
  const int x = b % 8, q = b / 8;
  return x < rem ? x * (per + 1) + q : rem * (per + 1) + (x - rem) * per + q;
}


// Interior kernel: one wave per grid row segment, barrier-free.  Lane l of a wave stands for column
// i = P - 1 + 62 tx + l of row j; lanes 1..62 own their cells, lane 0 supplies H at i-1 and lane 63
// supplies E at i+1 through cross-lane shuffles.  Each lane marches k up the chunk and per plane
// computes H at (i, j) and Hx, Hz at (i, j-1) (the row below belongs to another wave, so those two
// components are recomputed from the old state), then E at (i, j).  H at (i, j, k-1) comes from
// the previous plane in registers.  H of cells in the CPML slabs (outside C) is not recomputed: it
// is read from the new state written by the slab-H kernel.
constexpr int kOwnCols = 62;

template <bool IDX>
__global__ void __launch_bounds__(256) k_interior(const StepArgs *__restrict__ pa, int n) {
  const int I = pa->I, J = pa->J, K = pa->K, P = pa->P;
  const int bi0 = pa->i0, bi1 = pa->i1, bj0 = pa->j0, bj1 = pa->j1, bk0 = pa->k0, bk1 = pa->k1;
  const int nTx = pa->nTx, nTy = pa->nTy, KC = pa->KC, ntab = pa->ntab, act_h = pa->act_h, act_e = pa->act_e;
  const size_t IJK = pa->IJK, cs = pa->cs;
  const auto Fc = gptr(pa->Fc);
  const auto Fn = gptr(pa->Fn);
  const auto mat = gptr(pa->mat);
  const auto coef = gptr(pa->coef);
  const float Chh = pa->Chh, Che = pa->Che;
  __shared__ float sCoef[IDX ? 256 * 6 : 6];
  if (IDX)
    for (int q = threadIdx.x; q < ntab * 6; q += 256) sCoef[q] = coef[q];
  __syncthreads();

  const size_t IJ = (size_t)I * J;
  const int cx1 = I - P, cy1 = J - P, cz1 = K - P;  // C = [P, cx1) x [P, cy1) x [P, cz1)
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = L % nTx, t2 = L / nTx;
  const int ty = t2 % nTy, tz = t2 / nTy;
  const int i = P - 1 + kOwnCols * tx + lane;
  const int j = P + 4 * ty + w;
  const int kb = P + tz * KC, ke = min(kb + KC, cz1);
  if (j >= cy1) return;  // whole wave: no barrier follows
  const bool inI = i < I;
  const bool iC = (i >= P) && (i < cx1);          // column inside C
  const bool own = iC && lane >= 1 && lane <= kOwnCols;
  const bool hlane = lane <= kOwnCols && inI;      // lanes whose H feeds an own E
  const bool hC = hlane && iC;                     // H recomputed here
  const bool jm_C = (j - 1) >= P;                  // row j-1 inside C

  const auto Ex = Fc, Ey = Fc + cs, Ez = Fc + 2 * cs;
  const auto Hx = Fc + 3 * cs, Hy = Fc + 4 * cs, Hz = Fc + 5 * cs;
  const auto Nx = Fn + 3 * cs, Ny = Fn + 4 * cs, Nz = Fn + 5 * cs;
  const size_t col = inI ? (size_t)i + (size_t)j * I : (size_t)j * I;

  int k = kb > P ? kb - 1 : kb;
  // E of row j at plane k, and Ey of row j-1 at plane k (carried from plane to plane)
  float ex = 0, ey = 0, ez = 0, eym = 0;
  float hpx = 0, hpy = 0;  // H(i, j, k-1) of own cells
  {
    const size_t id = col + (size_t)k * IJ;
    if (inI) {
      ex = Ex[id]; ey = Ey[id]; ez = Ez[id];
      eym = Ey[id - I];
    }
    if (own && kb == P) {  // plane P-1 is a slab plane: its new H is already written
      hpx = Nx[id - IJ];
      hpy = Ny[id - IJ];
    }
  }
  for (; k < ke; ++k) {
    const size_t id = col + (size_t)k * IJ;
    const bool own_plane = k >= kb;
    // ---- loads of this plane (independent, issued together) ----
    float nx = 0, ny = 0, nz = 0, nym = 0;       // E(k+1) row j, Ey(k+1) row j-1
    float exm = 0, ezm = 0, exp_ = 0, ezp = 0;   // Ex, Ez of rows j-1 and j+1
    float hx = 0, hy = 0, hz = 0, hxm = 0, hzm = 0;
    int mt = 0;
    if (inI) {
      nx = Ex[id + IJ]; ny = Ey[id + IJ]; nz = Ez[id + IJ];
      nym = Ey[id + IJ - I];
      exm = Ex[id - I]; ezm = Ez[id - I];
      exp_ = Ex[id + I]; ezp = Ez[id + I];
    }
    if (hC) { hx = Hx[id]; hy = Hy[id]; hz = Hz[id]; }
    else if (hlane) { hx = Nx[id]; hy = Ny[id]; hz = Nz[id]; }
    if (hC && jm_C) { hxm = Hx[id - I]; hzm = Hz[id - I]; }
    else if (hlane) { hxm = Nx[id - I]; hzm = Nz[id - I]; }
    if (IDX && own && own_plane) mt = mat[id];
    // ---- x neighbours through the wave ----
    const float ey_i1 = __shfl_down(ey, 1), ez_i1 = __shfl_down(ez, 1), eym_i1 = __shfl_down(eym, 1);
    // ---- H(i, j, k) and Hx, Hz at (i, j-1, k): K1 GPU_updateH, src/gpu.h:592-601 ----
    // (C cells satisfy its mask; 1/kappa is exactly 1 in C, so the x * 1.0f factors are omitted)
    if (hC) {
      hx = Chh * hx + Che * (-(ezp - ez) + (ny - ey));
      hy = Chh * hy + Che * (-(nx - ex) + (ez_i1 - ez));
      hz = Chh * hz + Che * (-(ey_i1 - ey) + (exp_ - ex));
      if (act_h && i >= bi0 - 1 && i <= bi1 && j >= bj0 - 1 && j <= bj1 && k >= bk0 - 1 && k <= bk1 &&
          (i == bi0 - 1 || i == bi1 || j == bj0 - 1 || j == bj1 || k == bk0 - 1 || k == bk1)) {
        const F3 o = fused_pw_h(pa, n, i, j, k, F3{hx, hy, hz});
        hx = o.x; hy = o.y; hz = o.z;
      }
      if (jm_C) {
        const int jm = j - 1;
        hxm = Chh * hxm + Che * (-(ez - ezm) + (nym - eym));
        hzm = Chh * hzm + Che * (-(eym_i1 - eym) + (ex - exm));
        if (act_h && i >= bi0 - 1 && i <= bi1 && jm >= bj0 - 1 && jm <= bj1 && k >= bk0 - 1 && k <= bk1 &&
            (i == bi0 - 1 || i == bi1 || jm == bj0 - 1 || jm == bj1 || k == bk0 - 1 || k == bk1)) {
          const F3 o = fused_pw_h(pa, n, i, jm, k, F3{hxm, 0.0f, hzm});
          hxm = o.x; hzm = o.z;
        }
      }
    }
    const float hy_im = __shfl_up(hy, 1), hz_im = __shfl_up(hz, 1);
    if (own && own_plane) {
      Nx[id] = hx;
      Ny[id] = hy;
      Nz[id] = hz;
      float ce[3], ch[3];
      if (IDX) {
        const float *row = sCoef + 6 * mt;
        ce[0] = row[0]; ce[1] = row[1]; ce[2] = row[2];
        ch[0] = row[3]; ch[1] = row[4]; ch[2] = row[5];  // C cells are never on the last x plane
      } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) { ce[c] = coef[c * IJK + id]; ch[c] = coef[(3 + c) * IJK + id]; }
      }
      // K3 GPU_updateE, src/gpu.h:745-762 (C cells satisfy its masks; 1/kappa = 1 in C)
      float ux = ce[0] * ex + ch[0] * ((hz - hzm) - (hy - hpy));
      float uy = ce[1] * ey + ch[1] * ((hx - hpx) - (hz - hz_im));
      float uz = ce[2] * ez + ch[2] * ((hy - hy_im) - (hx - hxm));
      if (act_e && i >= bi0 && i <= bi1 && j >= bj0 && j <= bj1 && k >= bk0 && k <= bk1 &&
          (i == bi0 || i == bi1 || j == bj0 || j == bj1 || k == bk0 || k == bk1)) {
        const F3 o = fused_pw_e(pa, n, F3{ch[0], ch[1], ch[2]}, i, j, k, F3{ux, uy, uz});
        ux = o.x; uy = o.y; uz = o.z;
      }
      Fn[id] = ux;
      Fn[cs + id] = uy;
      Fn[2 * cs + id] = uz;
    }
    hpx = hx;
    hpy = hy;
    ex = nx; ey = ny; ez = nz; eym = nym;
  }
}

// Decode thread t of the slab kernels into a cell of S (z slabs, then y slabs of the remaining
// planes, then x slabs of the remaining rows).
__device__ __forceinline__ bool slab_cell(long t, int I, int J, int K, int P, int &i, int &j, int &k) {
  const long nz = 2L * P * I * J, ny = 2L * P * I * (K - 2 * P), nx = 2L * P * (J - 2 * P) * (K - 2 * P);
  if (t < nz) {
    const long pl = t / ((long)I * J), rem = t % ((long)I * J);
    k = pl < P ? (int)pl : K - 2 * P + (int)pl;
    j = (int)(rem / I);
    i = (int)(rem % I);
    return true;
  }
  t -= nz;
  if (t < ny) {
    const long q = t / I;
    i = (int)(t % I);
    const int js = (int)(q % (2 * P));
    k = P + (int)(q / (2 * P));
    j = js < P ? js : J - 2 * P + js;
    return true;
  }
  t -= ny;
  if (t < nx) {
    const long q = t / (2 * P);
    const int is = (int)(t % (2 * P));
    i = is < P ? is : I - 2 * P + is;
    j = P + (int)(q % (J - 2 * P));
    k = P + (int)(q / (J - 2 * P));
    return true;
  }
  return false;
}

struct SlabArgs {
  GridDev g;
  FieldSet cur, nxt;
  PsiSet pcur, pnxt;
  CoefDev co;
  PWDev pw;
  int n;
};

// A: H cells of the CPML slabs (src/gpu.h:575-602 + plane wave + src/gpu.h:604-725)
__global__ void __launch_bounds__(256) k_slab_h(const SlabArgs a) {
  int i, j, k;
  if (!slab_cell((long)blockIdx.x * 256 + threadIdx.x, a.g.I, a.g.J, a.g.K, a.g.P, i, j, k)) return;
  const int I = a.g.I, J = a.g.J, K = a.g.K;
  const size_t IJ = (size_t)I * J;
  const size_t id = (size_t)i + (size_t)j * I + (size_t)k * IJ;
  const float *Ex = a.cur.f[EX], *Ey = a.cur.f[EY], *Ez = a.cur.f[EZ];
  HNb e;
  e.ex = Ex[id]; e.ey = Ey[id]; e.ez = Ez[id];
  const bool in = i + 1 < I, jn = j + 1 < J, kn = k + 1 < K;
  e.ey_i1 = in ? Ey[id + 1] : 0.0f;
  e.ez_i1 = in ? Ez[id + 1] : 0.0f;
  e.ex_j1 = jn ? Ex[id + I] : 0.0f;
  e.ez_j1 = jn ? Ez[id + I] : 0.0f;
  e.ex_k1 = kn ? Ex[id + IJ] : 0.0f;
  e.ey_k1 = kn ? Ey[id + IJ] : 0.0f;
  float hx = a.cur.f[HX][id], hy = a.cur.f[HY][id], hz = a.cur.f[HZ][id];
  const float *rowE = a.pw.active_h ? a.pw.Eh + (size_t)a.n * a.pw.N1 : nullptr;
  h_cell(a.g, a.co, a.pw, rowE, a.pcur, a.pnxt, true, i, j, k, IJ, e, hx, hy, hz);
  a.nxt.f[HX][id] = hx;
  a.nxt.f[HY][id] = hy;
  a.nxt.f[HZ][id] = hz;
}

// C: E cells of the CPML slabs (src/gpu.h:727-763 + plane wave + src/gpu.h:765-868), reading the
// new H of kernels A and B.
template <bool IDX>
__global__ void __launch_bounds__(256) k_slab_e(const SlabArgs a) {
  int i, j, k;
  if (!slab_cell((long)blockIdx.x * 256 + threadIdx.x, a.g.I, a.g.J, a.g.K, a.g.P, i, j, k)) return;
  const int I = a.g.I, J = a.g.J;
  const size_t IJ = (size_t)I * J;
  const size_t id = (size_t)i + (size_t)j * I + (size_t)k * IJ;
  const float *Hx = a.nxt.f[HX], *Hy = a.nxt.f[HY], *Hz = a.nxt.f[HZ];
  ENb h;
  h.hx = Hx[id]; h.hy = Hy[id]; h.hz = Hz[id];
  h.hy_im = i > 0 ? Hy[id - 1] : 0.0f;
  h.hz_im = i > 0 ? Hz[id - 1] : 0.0f;
  h.hx_jm = j > 0 ? Hx[id - I] : 0.0f;
  h.hz_jm = j > 0 ? Hz[id - I] : 0.0f;
  h.hx_km = k > 0 ? Hx[id - IJ] : 0.0f;
  h.hy_km = k > 0 ? Hy[id - IJ] : 0.0f;
  float ce[3], ch[3];
  cell_coef<IDX>(a.co, a.co.tab, id, i, I, ce, ch);
  float ex = a.cur.f[EX][id], ey = a.cur.f[EY][id], ez = a.cur.f[EZ][id];
  const float *rowH = a.pw.active_e ? a.pw.Hh + (size_t)a.n * a.pw.N1 : nullptr;
  e_cell(a.g, a.co, a.pw, rowH, a.pcur, a.pnxt, i, j, k, IJ, ce, ch, h, ex, ey, ez);
  a.nxt.f[EX][id] = ex;
  a.nxt.f[EY][id] = ey;
  a.nxt.f[EZ][id] = ez;
}

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------
static inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

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
                               const int box[6], float s, float dx) {
  const int a0 = box[0] < 0 ? 0 : box[0], a1 = box[1] >= g.I ? g.I - 1 : box[1];
  const int b0 = box[2] < 0 ? 0 : box[2], b1 = box[3] >= g.J ? g.J - 1 : box[3];
  const int c0 = box[4] < 0 ? 0 : box[4], c1 = box[5] >= g.K ? g.K - 1 : box[5];
  if (a1 < a0 || b1 < b0 || c1 < c0) return hipSuccess;
  const int ni = a1 - a0 + 1, nj = b1 - b0 + 1, nk = c1 - c0 + 1;
  const unsigned nb = cdiv((long)ni * nj * nk, 256);
  if (co.mat) hipLaunchKernelGGL(k_add_currents<true>, dim3(nb), dim3(256), 0, st, g, F, co, comp, a0, b0, c0, ni, nj, nk, s, dx);
  else hipLaunchKernelGGL(k_add_currents<false>, dim3(nb), dim3(256), 0, st, g, F, co, comp, a0, b0, c0, ni, nj, nk, s, dx);
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

bool fused_supported(const GridDev &g) { return g.I >= 2 * g.P + 2 && g.J >= 2 * g.P + 2 && g.K >= 2 * g.P + 2; }

template <bool IDX>
static hipError_t launch_interior_t(hipStream_t st, const StepArgs *dargs, int nwg, int n) {
  hipLaunchKernelGGL((k_interior<IDX>), dim3(nwg), dim3(256), 0, st, dargs, n);
  return hipGetLastError();
}

size_t step_args_size() { return sizeof(StepArgs); }

// Argument block of the interior kernel (host copy, uploaded once per state parity / plane-wave
// mode); returns the number of interior workgroups (0 if the interior is empty), -1 if unsupported.
int make_step_args(const FusedDev &f, const StepTuning &tune, void *out) {
  const GridDev &g = f.g;
  if (!fused_supported(g)) return -1;
  if (f.mat && f.ntab > 256) return -1;
  const int ty = 4;  // rows (waves) per workgroup
  const int cI = g.I - 2 * g.P, cJ = g.J - 2 * g.P, cK = g.K - 2 * g.P;
  const int nTx = (cI + kOwnCols - 1) / kOwnCols, nTy = (cJ + ty - 1) / ty;
  int kc = tune.z_chunk;
  if (kc <= 0) {
    const int want = 4096;  // workgroups per launch (several per CU)
    int nkc = (want + nTx * nTy - 1) / (nTx * nTy);
    if (nkc < 1) nkc = 1;
    kc = (cK + nkc - 1) / nkc;
    if (kc < 16) kc = 16;
  }
  const int nKc = (cK + kc - 1) / kc;
  StepArgs a;
  std::memset(&a, 0, sizeof(a));
  a.I = g.I; a.J = g.J; a.K = g.K; a.P = g.P;
  a.i0 = g.i0; a.i1 = g.i1; a.j0 = g.j0; a.j1 = g.j1; a.k0 = g.k0; a.k1 = g.k1;
  a.nTx = nTx; a.nTy = nTy; a.KC = kc; a.ntab = f.ntab;
  a.act_h = f.act_h; a.act_e = f.act_e; a.N1 = f.N1; a.pad = ty;
  a.IJK = (size_t)g.I * g.J * g.K;
  a.cs = f.cs;
  a.sx = (size_t)2 * g.P * g.J * g.K; a.sy = (size_t)2 * g.I * g.P * g.K; a.sz = (size_t)2 * g.I * g.J * g.P;
  a.Fc = f.Fc; a.Fn = f.Fn; a.Pc = f.Pc; a.Pn = f.Pn;
  a.mat = f.mat; a.coef = f.coef; a.prm = f.prm; a.pwtab = f.pwtab;
  a.Eh = f.Eh; a.Hh = f.Hh;
  a.Chh = f.Chh; a.Che = f.Che; a.dx = f.dx;
  for (int c = 0; c < 3; ++c) { a.polE[c] = f.polE[c]; a.polH[c] = f.polH[c]; }
  std::memcpy(out, &a, sizeof(a));
  return (cI > 0 && cJ > 0 && cK > 0) ? nTx * nTy * nKc : 0;
}

static SlabArgs slab_args(const FusedDev &f, int n) {
  SlabArgs s;
  const GridDev &g = f.g;
  const size_t IJK = (size_t)g.I * g.J * g.K;
  const size_t sx = (size_t)2 * g.P * g.J * g.K, sy = (size_t)2 * g.I * g.P * g.K, sz = (size_t)2 * g.I * g.J * g.P;
  s.g = g;
  for (int c = 0; c < 6; ++c) {
    s.cur.f[c] = const_cast<float *>(f.Fc) + c * f.cs;
    s.nxt.f[c] = f.Fn + c * f.cs;
  }
  auto psi = [&](float *base, PsiSet &ps) {
    float *px = base, *py = px + 4 * sx, *pz = py + 4 * sy;
    ps.p[P_EYX] = px; ps.p[P_EZX] = px + sx; ps.p[P_HYX] = px + 2 * sx; ps.p[P_HZX] = px + 3 * sx;
    ps.p[P_EXY] = py; ps.p[P_EZY] = py + sy; ps.p[P_HXY] = py + 2 * sy; ps.p[P_HZY] = py + 3 * sy;
    ps.p[P_EXZ] = pz; ps.p[P_EYZ] = pz + sz; ps.p[P_HXZ] = pz + 2 * sz; ps.p[P_HYZ] = pz + 3 * sz;
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
  return s;
}

// One time step: A (slab H), B (interior H and E), C (slab E).
hipError_t launch_step_fused(hipStream_t st, const FusedDev &f, const void *dargs, const void *hargs, int nwg, int n) {
  const GridDev &g = f.g;
  const SlabArgs sa = slab_args(f, n);
  const long nslab = 2L * g.P * g.I * g.J + 2L * g.P * g.I * (g.K - 2 * g.P) + 2L * g.P * (g.J - 2 * g.P) * (g.K - 2 * g.P);
  const unsigned nb = cdiv(nslab, 256);
  hipLaunchKernelGGL(k_slab_h, dim3(nb), dim3(256), 0, st, sa);
  if (nwg > 0) {
    const StepArgs &a = *(const StepArgs *)hargs;
    const StepArgs *d = (const StepArgs *)dargs;
    const hipError_t e = a.mat ? launch_interior_t<true>(st, d, nwg, n) : launch_interior_t<false>(st, d, nwg, n);
    if (e != hipSuccess) return e;
  }
  if (f.mat) hipLaunchKernelGGL(k_slab_e<true>, dim3(nb), dim3(256), 0, st, sa);
  else hipLaunchKernelGGL(k_slab_e<false>, dim3(nb), dim3(256), 0, st, sa);
  return hipGetLastError();
}

}  // namespace fdtd
