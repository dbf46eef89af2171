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
// fused step kernel
// ------------------------------------------------------------------------------------------
struct StepArgs {
  GridDev g;
  FieldSet cur, nxt;
  PsiSet pcur, pnxt;
  CoefDev co;
  PWDev pw;
  int n;
  int KC;   // z chunk depth
  int nTy;  // row tiles
};

// Workgroups are dealt to the 8 XCDs round-robin; remap so that each XCD walks a contiguous range
// of tiles (neighbouring row tiles of one z chunk share halo rows in that XCD's L2).
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int per = nb / 8, rem = nb % 8;
  const int x = b % 8, q = b / 8;
  return x < rem ? x * (per + 1) + q : rem * (per + 1) + (x - rem) * per + q;
}

template <bool IDX, int NC, int TY>
__global__ void __launch_bounds__(64 * (TY + 2)) k_step_fused(const StepArgs a) {
  constexpr int W = TY + 2;   // waves: halo row, TY own rows, halo row
  constexpr int IP = NC * 64; // LDS row length
  __shared__ float sE[3][W][IP];
  __shared__ float sH[3][TY + 1][IP];
  __shared__ float sCoef[IDX ? 256 * 6 : 6];

  const GridDev &g = a.g;
  const int I = g.I, J = g.J, K = g.K;
  const size_t IJ = (size_t)I * J;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int ty = L % a.nTy, tz = L / a.nTy;
  const int jb = ty * TY, kb = tz * a.KC, ke = min(kb + a.KC, K);
  const int j = jb - 1 + w;
  const bool row_ok = (j >= 0) && (j < J);
  const bool h_row = row_ok && (w <= TY);
  const bool own_row = row_ok && (w >= 1) && (w <= TY);
  const float *rowE = a.pw.active_h ? a.pw.Eh + (size_t)a.n * a.pw.N1 : nullptr;
  const float *rowH = a.pw.active_e ? a.pw.Hh + (size_t)a.n * a.pw.N1 : nullptr;

  if (IDX) {
    for (int q = threadIdx.x; q < a.co.ntab * 6; q += 64 * W) sCoef[q] = a.co.tab[q];
  }

  float ecx[NC], ecy[NC], ecz[NC];
  float enx[NC], eny[NC], enz[NC];
  float hpx[NC], hpy[NC];
  float nhx[NC], nhy[NC], nhz[NC];

  int k = kb > 0 ? kb - 1 : 0;
  {
    const size_t base = (size_t)j * I + (size_t)k * IJ;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int i = lane + 64 * c;
      const bool ok = row_ok && i < I;
      ecx[c] = ok ? a.cur.f[EX][base + i] : 0.0f;
      ecy[c] = ok ? a.cur.f[EY][base + i] : 0.0f;
      ecz[c] = ok ? a.cur.f[EZ][base + i] : 0.0f;
      hpx[c] = 0.0f;
      hpy[c] = 0.0f;
    }
  }

  for (; k < ke; ++k) {
    const bool kp1 = (k + 1 < K);
    const size_t base = (size_t)j * I + (size_t)k * IJ;
    // E(k+1) of this row (feeds the z differences of H(k) and next plane's LDS)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int i = lane + 64 * c;
      const bool ok = row_ok && kp1 && i < I;
      enx[c] = ok ? a.cur.f[EX][base + IJ + i] : 0.0f;
      eny[c] = ok ? a.cur.f[EY][base + IJ + i] : 0.0f;
      enz[c] = ok ? a.cur.f[EZ][base + IJ + i] : 0.0f;
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int i = lane + 64 * c;
      sE[0][w][i] = ecx[c];
      sE[1][w][i] = ecy[c];
      sE[2][w][i] = ecz[c];
    }
    __syncthreads();

    // ---- H(k) for rows jb-1 .. jb+TY-1 ----
    const bool own_plane = (k >= kb);
    if (h_row) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int i = lane + 64 * c;
        if (i < I) {
          const size_t id = base + i;
          float hx = a.cur.f[HX][id], hy = a.cur.f[HY][id], hz = a.cur.f[HZ][id];
          HNb e;
          e.ex = ecx[c]; e.ey = ecy[c]; e.ez = ecz[c];
          const bool jn = (j + 1 < J), in = (i + 1 < I);
          e.ex_j1 = jn ? sE[0][w + 1][i] : 0.0f;
          e.ez_j1 = jn ? sE[2][w + 1][i] : 0.0f;
          e.ey_i1 = in ? sE[1][w][i + 1] : 0.0f;
          e.ez_i1 = in ? sE[2][w][i + 1] : 0.0f;
          e.ex_k1 = enx[c]; e.ey_k1 = eny[c];
          const bool own = own_row && own_plane;
          h_cell(g, a.co, a.pw, rowE, a.pcur, a.pnxt, own, i, j, k, IJ, e, hx, hy, hz);
          sH[0][w][i] = hx;
          sH[1][w][i] = hy;
          sH[2][w][i] = hz;
          if (own) {
            a.nxt.f[HX][id] = hx;
            a.nxt.f[HY][id] = hy;
            a.nxt.f[HZ][id] = hz;
          }
          nhx[c] = hx; nhy[c] = hy; nhz[c] = hz;
        } else {
          nhx[c] = 0.0f; nhy[c] = 0.0f; nhz[c] = 0.0f;
        }
      }
    }
    __syncthreads();

    // ---- E(k) for own rows ----
    if (own_row && own_plane) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int i = lane + 64 * c;
        if (i < I) {
          const size_t id = base + i;
          ENb h;
          h.hx = nhx[c]; h.hy = nhy[c]; h.hz = nhz[c];
          h.hx_jm = sH[0][w - 1][i];
          h.hz_jm = sH[2][w - 1][i];
          h.hy_im = i > 0 ? sH[1][w][i - 1] : 0.0f;
          h.hz_im = i > 0 ? sH[2][w][i - 1] : 0.0f;
          h.hx_km = hpx[c]; h.hy_km = hpy[c];
          float ce[3], ch[3];
          cell_coef<IDX>(a.co, sCoef, id, i, I, ce, ch);
          float ex = ecx[c], ey = ecy[c], ez = ecz[c];
          e_cell(g, a.co, a.pw, rowH, a.pcur, a.pnxt, i, j, k, IJ, ce, ch, h, ex, ey, ez);
          a.nxt.f[EX][id] = ex;
          a.nxt.f[EY][id] = ey;
          a.nxt.f[EZ][id] = ez;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      hpx[c] = nhx[c]; hpy[c] = nhy[c];
      ecx[c] = enx[c]; ecy[c] = eny[c]; ecz[c] = enz[c];
    }
  }
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

bool fused_supported(const GridDev &g) { return g.I <= 512 && g.I >= 2 * g.P + 2; }

template <bool IDX, int NC, int TY>
static hipError_t launch_fused_t(hipStream_t st, const StepArgs &a, int nwg) {
  hipLaunchKernelGGL((k_step_fused<IDX, NC, TY>), dim3(nwg), dim3(64 * (TY + 2)), 0, st, a);
  return hipGetLastError();
}

template <bool IDX, int TY>
static hipError_t launch_fused_nc(hipStream_t st, const StepArgs &a, int nwg, int nc) {
  switch (nc) {
    case 1: return launch_fused_t<IDX, 1, TY>(st, a, nwg);
    case 2: return launch_fused_t<IDX, 2, TY>(st, a, nwg);
    case 3: return launch_fused_t<IDX, 3, TY>(st, a, nwg);
    case 4: return launch_fused_t<IDX, 4, TY>(st, a, nwg);
    case 5: return launch_fused_t<IDX, 5, TY>(st, a, nwg);
    case 6: return launch_fused_t<IDX, 6, TY>(st, a, nwg);
    case 7: return launch_fused_t<IDX, 7, TY>(st, a, nwg);
    case 8: return launch_fused_t<IDX, 8, TY>(st, a, nwg);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_step_fused(hipStream_t st, const GridDev &g, const FieldSet &cur, const FieldSet &nxt,
                             const PsiSet &pcur, const PsiSet &pnxt, const CoefDev &co, const PWDev &pw, int n,
                             const StepTuning &tune) {
  if (!fused_supported(g)) return hipErrorInvalidValue;
  const int nc = (g.I + 63) / 64;
  // LDS budget: rows x components x row length; wide rows use 4-row tiles.
  int ty = tune.tile_rows == 4 ? 4 : 8;
  if (nc > 5) ty = 4;
  const int nTy = (g.J + ty - 1) / ty;
  int kc = tune.z_chunk;
  if (kc <= 0) {
    const int want = 1024;  // workgroups to launch (a few per CU)
    int nkc = (want + nTy - 1) / nTy;
    if (nkc < 1) nkc = 1;
    kc = (g.K + nkc - 1) / nkc;
    if (kc < 8) kc = 8;
  }
  const int nKc = (g.K + kc - 1) / kc;
  StepArgs a;
  a.g = g; a.cur = cur; a.nxt = nxt; a.pcur = pcur; a.pnxt = pnxt; a.co = co; a.pw = pw; a.n = n;
  a.KC = kc; a.nTy = nTy;
  const int nwg = nTy * nKc;
  const bool idx = co.mat != nullptr;
  if (idx && co.ntab > 256) return hipErrorInvalidValue;
  if (idx) return ty == 4 ? launch_fused_nc<true, 4>(st, a, nwg, nc) : launch_fused_nc<true, 8>(st, a, nwg, nc);
  return ty == 4 ? launch_fused_nc<false, 4>(st, a, nwg, nc) : launch_fused_nc<false, 8>(st, a, nwg, nc);
}

}  // namespace fdtd
