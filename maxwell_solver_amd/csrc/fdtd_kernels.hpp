// fdtd_kernels.hpp -- device-side argument blocks and host launchers of the FDTD kernels.
#pragma once

#include <hip/hip_runtime_api.h>

#include "fdtd_internal.hpp"

namespace fdtd {

struct GridDev {
  int I, J, K, P;
  int i0, i1, j0, j1, k0, k1;  // TF/SF box
  // Row pitch of the per-cell device arrays (fields, materials, per-cell coefficients): rows of
  // Ip >= I elements, planes of Ip*J, so that every row starts on a 128-byte line when I is not
  // a multiple of 32; 0 means I.  The psi slabs and face sheets keep their own unpadded layouts.
  int Ip;
};


struct CoefDev {
  const uint8_t *mat = nullptr;  // index mode: per-cell material id
  const float *tab = nullptr;    // index mode: [ntab][6] = Ce x,y,z, Ch x,y,z (Ch with 1/dx)
  int ntab = 0;
  const float *Ce[3] = {nullptr, nullptr, nullptr};  // full mode: per-cell arrays
  const float *Ch[3] = {nullptr, nullptr, nullptr};
  const float *ikE[3] = {nullptr, nullptr, nullptr}, *ikH[3] = {nullptr, nullptr, nullptr};
  const float *bE = nullptr, *cE = nullptr, *bH = nullptr, *cH = nullptr;
  const float *Chmu[3] = {nullptr, nullptr, nullptr};  // per-cell mu (CreateMu): K2 H coefficients
  float Chh = 1, Che = 0, dx = 0;
};

struct FieldSet {
  float *f[6];
};
struct PsiSet {
  float *p[12];
};

struct LineDev {
  int N, T, P, src;
  float Ch, ds;
  const float *Ce, *Cb, *ikE, *ikH;  // per node
  const float *bE, *cE, *bH, *cH;    // [2P]
  const float *pw;                   // [T] source samples (already scaled by rate)
  float *Eh, *Hh;                    // [T][N] histories
};

// Per-op kernels (state updated in place), one launcher per reference op.
hipError_t launch_update_h_all(hipStream_t st, const GridDev &g, const FieldSet &F, const CoefDev &co);
hipError_t launch_update_e_all(hipStream_t st, const GridDev &g, const FieldSet &F, const CoefDev &co);
hipError_t launch_update_h_pml(hipStream_t st, const GridDev &g, const FieldSet &F, const PsiSet &ps, const CoefDev &co);
hipError_t launch_update_e_pml(hipStream_t st, const GridDev &g, const FieldSet &F, const PsiSet &ps, const CoefDev &co);
hipError_t launch_pw_h(hipStream_t st, const GridDev &g, const FieldSet &F, const CoefDev &co, const PWDev &pw, int n);
hipError_t launch_pw_e(hipStream_t st, const GridDev &g, const FieldSet &F, const CoefDev &co, const PWDev &pw, int n);
// src (optional): take the sample from src[*dn + q] instead of s (launches recorded into a graph)
hipError_t launch_add_currents(hipStream_t st, const GridDev &g, const FieldSet &F, const CoefDev &co, int comp,
                               const int box[6], float s, float dx, const float *src = nullptr, const int *dn = nullptr,
                               int q = 0);
// Cells an AddCurrents box writes (upper ends clipped to the grid); false when it writes none
// (an empty box, or a negative start, as the reference's size_t box, src/gpu.h:55-82)
bool current_box(const GridDev &g, const int box[6], int &a0, int &a1, int &b0, int &b1, int &c0, int &c1);
// dn[0] += steps, dn[1] += steps (one thread): the time index and record row of graph-recorded
// launches, advanced once per replay
hipError_t launch_advance(hipStream_t st, int *dn, int steps);
hipError_t launch_line(hipStream_t st, const LineDev &L);
hipError_t launch_probes(hipStream_t st, const FieldSet &F, const int *idx, int nAnt, const int *comps, int ncomp,
                         float *port, int row, int T);

// Fused step: reads the cur state, writes the nxt state (packed buffers).
struct StepTuning {
  int tile_rows = 0;  // rows per workgroup of the row kernel (1, 2, 4, 8, 16; 0: automatic)
  int z_chunk = 0;    // planes per chunk (0: automatic)
  int cpl = 0;        // columns per lane (1, 2, 4; 0: automatic)
  bool no_split = false;  // 1 column per lane without the split edge segment
  int slab_chunk = 0;  // planes per chunk of the slab part (0: as the interior part; the z slabs whole when thin)
};
struct FusedDev {
  GridDev g;
  bool force_global = false;  // the row kernel's 64-bit global addressing even where buffers fit
  const float *Fc = nullptr;  // six fields, component c at c*cs
  size_t cs = 0;
  float *Fn = nullptr;
  const float *Pc = nullptr;  // psi groups: x slabs [Eyx Ezx Hyx Hzx], y [Exy Ezy Hxy Hzy], z [Exz Eyz Hxz Hyz]
  float *Pn = nullptr;
  const uint8_t *mat = nullptr;
  const float *coef = nullptr;  // index mode: table [ntab][6]; full mode: [6][Ip*J*K]
  int ntab = 0;
  const float *prm = nullptr;   // bE cE bH cH [2P each], 1/kappa E x,y,z, 1/kappa H x,y,z
  const float *pwtab = nullptr; // incident-field projection tables (E then H; x,y,z; integer, half)
  const float *Eh = nullptr, *Hh = nullptr;
  float *xs = nullptr;           // x-face sheets of the TF/SF box: 8 x (k1-k0+1) x (j1-j0+1) floats
  int N1 = 0, act_h = 0, act_e = 0;
  float polE[3] = {0, 0, 0}, polH[3] = {0, 0, 0};
  float Chh = 1, Che = 0, dx = 0;
  const float *Chmu = nullptr;  // per-cell H coefficients of CreateMu [3][Ip*J*K] (K2), or null
  bool gr = false;              // graded mesh: the per-axis tables carry dx / spacing in every cell
};
// Argument block of the fused kernel: built on the host (make_step_args returns the workgroup
// count, -1 if unsupported), uploaded once per (state parity, plane-wave mode), then launched.
size_t step_args_size();
int make_step_args(const FusedDev &f, const StepTuning &tune, void *out);
// Work folded into the launches of a fused step by a run of consecutive steps: the face sheets of
// step n filled by the previous step (sheets_ready), the sheets of step n + 1 filled into xs_next
// by the slab part's launch (next_filled reports whether it did), and the probe gather of the
// state the step reads into record row `row` (-1: none) by the same launch.
struct StepFold {
  bool sheets_ready = false;
  float *xs_next = nullptr;
  bool next_filled = false;
  FieldSet F{};
  const int *idx = nullptr, *comps = nullptr;
  float *port = nullptr;
  int nAnt = 0, ncomp = 0, row = -1, T = 0;
  // Step currents (cur_comp >= 0: E component): AddCurrents on the new state after the step, taken
  // into the slab part's launch when the box lies in the interior part's cells (cur_done reports
  // it; otherwise the caller launches them).  Sample cur_s, or cur_src[*cur_dn + cur_q].
  int cur_comp = -1, cur_box[6] = {0, 0, 0, 0, 0, 0}, cur_q = 0;
  float cur_s = 0.0f, cur_dx = 0.0f;
  const float *cur_src = nullptr;
  const int *cur_dn = nullptr;
  bool cur_done = false;
  // Launches recorded into a graph of several steps (gidx set): the step is the gq-th of the
  // recording; the sheets filled are those of time index gidx[0] + gq + 1 (none past T) and, with
  // probe_rel, the record row gathered is gidx[1] + gq - 1 (none before gidx[2]).
  const int *gidx = nullptr;
  int gq = 0;
  bool probe_rel = false;
};
// ev_rows (optional): two events recorded on st around the row kernel's interior part
hipError_t launch_step_fused(hipStream_t st, const FusedDev &f, const void *dargs, const void *hargs, int nwg, int n,
                             hipEvent_t *ev_rows = nullptr, StepFold *fold = nullptr);
bool fused_supported(const GridDev &g);
// The TF/SF face sheets of time index n into f.xs (nothing without a plane wave).
hipError_t launch_sheets(hipStream_t st, const FusedDev &f, int n);
// The stream ceiling: Fn[c][id] = Fc[c][id] + mat[id] for the six components (cs apart), n cells.
hipError_t launch_stream(hipStream_t st, const float *Fc, float *Fn, const uint8_t *mat, size_t cs, size_t n);

// Device scratch of the allocation-free run (fdtd_run_noalloc): kScratchBytes per GPU, a file-scope
// device array; scratch_region returns its device address on the current device (looked up once
// per device).  Layout: two TF/SF face-sheet buffers of kScratchSheetBytes each, then the row
// kernel's argument blocks of the two state sets (kScratchArgBytes each).
constexpr size_t kScratchSheetBytes = (size_t)24 << 20;
constexpr size_t kScratchArgBytes = 4096;
constexpr size_t kScratchBytes = 2 * kScratchSheetBytes + 2 * kScratchArgBytes;
hipError_t scratch_region(void **ptr);

// ---- batched fused step: one launch per kernel covers many grids ----
// Per grid of a batched launch: first workgroup, current set and time index at batch step 0, and
// the argument block (device copy) to use while the current set is 0 or 1.  A table has ng + 1
// entries; the last holds the total workgroup count in `off`.
struct BatchGrid {
  int off = 0, cur0 = 0, n0 = 0, pad = 0;
  const void *arg[2] = {nullptr, nullptr};
};
// Probe gather of one grid in a batch: the field sets 0 and 1, probe cells, components (0..5),
// record [ncomp][T][nAnt] and the record row at batch step 0.
struct ProbeJob {
  FieldSet F[2];
  const int *idx = nullptr;
  const int *comps = nullptr;
  float *port = nullptr;
  int nAnt = 0, ncomp = 0, row0 = 0, T = 0, cur0 = 0, pad = 0;
};
// Row-kernel variant of a grid (grids of one batched row launch share it).
struct RowsKey {
  bool idx = false, buf = false;
  int cpl = 1, threads = 256, threads0 = 256;  // workgroup sizes of the slab part and the interior part
  unsigned lds = 0, lds0 = 0;                    // their dynamic LDS bytes (row exchange)
  bool ntf = false;                              // the slab part needs no TF/SF code
  bool operator<(const RowsKey &o) const {
    if (ntf != o.ntf) return ntf < o.ntf;
    if (idx != o.idx) return idx < o.idx;
    if (buf != o.buf) return buf < o.buf;
    if (cpl != o.cpl) return cpl < o.cpl;
    if (threads != o.threads) return threads < o.threads;
    if (threads0 != o.threads0) return threads0 < o.threads0;
    if (lds != o.lds) return lds < o.lds;
    return lds0 < o.lds0;
  }
};
RowsKey rows_key(const void *hargs);
// Face-sheet argument block of f (host copy; n is supplied per step) and its workgroup count.
size_t slab_args_size();
void make_slab_args(const FusedDev &f, void *out, unsigned *nb_sheets);
hipError_t launch_rows_batch(hipStream_t st, const RowsKey &key, int part, const BatchGrid *tab, int ng, int nwg,
                             int step);
// The same tiles pulled from device work queues (one per XCD) by resident workgroups, 1/share of
// those that fit at once: queue[8 slot .. 8 slot + 7] count tickets (zero at launch; the launch
// clears those of slot ^ 1 for the next one).
hipError_t launch_rows_queue(hipStream_t st, const RowsKey &key, int part, const BatchGrid *tab, int ng, int nwg,
                             int step, int *queue, int slot, int share);
// Workgroups of part 0 (interior rows and chunks) or part 1 (the slab rows and chunks) of the row kernel.
int rows_part_nwg(const void *hargs, int part);
hipError_t launch_sheets_batch(hipStream_t st, const BatchGrid *tab, int ng, int nwg, int step);
hipError_t launch_probes_batch(hipStream_t st, const ProbeJob *jobs, int njobs, int max_ant, int step);

}  // namespace fdtd
