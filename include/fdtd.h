/*
 * fdtd.h -- C ABI of the MI355X (gfx950) 3D Yee FDTD library, libfdtd_hip.so.
 *
 * One entry point per op of the reference MATLAB MEX gateway (src/simFDTD.cu:19-187), with the
 * op's argument meaning and order kept, plus the extensions the hot path needs (fused step/run,
 * runtime CPML parameters, oblique and layered plane waves, uint8 material index, probes,
 * snapshots).  Plain pointers and sizes only; all arrays are host arrays in the reference's
 * x-fastest layout (index i + j*I + k*I*J, as MATLAB's column-major x(:) of an I x J x K array).
 *
 * Every function returns 0 on success or an FDTD_ERR_* code; fdtd_last_error() gives a
 * thread-local message and fdtd_last_error_id() the MEX error identifier the reference would
 * raise ("simFDTD:InvalidInput", "simFDTD:AlreadyInitialized", ...).  GPU work is queued on the
 * simulation's own HIP stream and is asynchronous except where a function returns host data.
 */
#ifndef MAXWELL_FDTD_H
#define MAXWELL_FDTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdtd_sim fdtd_sim; /* one simulation; reference: class simulation, src/simulation.h:24.
                                   * Every entry point rejects a handle that fdtd_create did not return
                                   * or fdtd_destroy has destroyed (simFDTD:InvalidInput); the reference
                                   * leaves that undefined (src/simFDTD.cu:32-52). */
typedef struct fdtd_batch fdtd_batch; /* a set of simulations advanced together (no reference counterpart) */

enum {
  FDTD_OK = 0,
  FDTD_ERR_INVALID = 1,     /* bad argument                        -> simFDTD:InvalidInput */
  FDTD_ERR_STATE = 2,       /* op called in the wrong order        -> simFDTD:InvalidInput */
  FDTD_ERR_ALREADY = 3,     /* repeated one-time op                -> simFDTD:AlreadyInitialized */
  FDTD_ERR_DEVICE = 4,      /* HIP runtime failure                 -> simFDTD:DeviceError */
  FDTD_ERR_NOMEM = 5,       /* allocation failed / over budget     -> simFDTD:OutOfMemory */
  FDTD_ERR_UNKNOWN_OP = 6   /* unknown op string (gateway level)   -> simFDTD:UnknownOperation */
};

/* Field / probe component ids: 1..3 = Ex, Ey, Ez; 4..6 = Hx, Hy, Hz. */

const char *fdtd_last_error(void);
const char *fdtd_last_error_id(void);
const char *fdtd_version(void);

/* ---------------- ops of the reference MEX gateway ---------------- */

/* 'CreateSimulation': src/simFDTD.cu:30-38, constructor src/simulation.h:70-77 */
int fdtd_create(fdtd_sim **out);
/* 'DeleteSimulation': src/simFDTD.cu:49-52, destructor src/simulation.h:1153-1269 */
int fdtd_destroy(fdtd_sim *s);
/* 'SetConst': freqRes = [fctr, dx, dt], timeDim = [T, I, J, K, nAnt]; src/simulation.h:79-131 */
int fdtd_set_const(fdtd_sim *s, const float freqRes[3], const int32_t timeDim[5]);
/* 'SetSource' (update = 0) / update mode with a 4th MEX argument (update = 1);
 * source has T values; src/simFDTD.cu:111-125, src/simulation.h:1068-1095 */
int fdtd_set_source(fdtd_sim *s, const float *source, int32_t update);
/* 'SetAntennaLocations': locations int32[nAnt x 3] column-major (x, y, z offsets from the CPML);
 * src/simulation.h:1047-1067 */
int fdtd_set_antenna_locations(fdtd_sim *s, const int32_t *locations);
/* 'UpdateMaterial': flag 1..3 = sigma x,y,z; 4..6 = eps_inf x,y,z; values has I*J*K entries;
 * src/simulation.h:288-328 */
int fdtd_update_material(fdtd_sim *s, int32_t flag, const float *values);
/* 'ReadySimulation': CPML profiles, coefficients, device state; src/simulation.h:640-1045 */
int fdtd_ready(fdtd_sim *s);
/* 'Init1D': incident-line histories for all T steps; src/simulation.h:134-254 */
int fdtd_init_1d(fdtd_sim *s, float rate, const float *pw_source);
/* 'UpdateCuda': recompute the E coefficients from the material arrays; src/simulation.h:1097-1127 */
int fdtd_update_coeffs(fdtd_sim *s);
/* 'ResetFieldAll': zero the six fields and twelve CPML psi arrays; src/simulation.h:1128-1151 */
int fdtd_reset_fields(fdtd_sim *s);
/* 'UpdateH_All' src/simulation.h:371-385 (gpu.h:575-602); 'UpdateE_All' :386-391 (gpu.h:727-763) */
int fdtd_update_h_all(fdtd_sim *s);
int fdtd_update_e_all(fdtd_sim *s);
/* 'UpdateH_PML' src/simulation.h:392-403 (gpu.h:604-725); 'UpdateE_PML' :404-418 (gpu.h:765-868).
 * UpdateE_PML also closes a step for the probe recorder (row = steps since ResetFieldAll). */
int fdtd_update_h_pml(fdtd_sim *s);
int fdtd_update_e_pml(fdtd_sim *s);
/* 'UpdateH_PlaneWave' / 'UpdateE_PlaneWave': time = [n, mode]; modes 1..4 as the reference
 * (1 ZY, 2 ZX, 3 XZ, 4 XY), 5 = oblique; src/simulation.h:330-369, :421-570.  Modes 6 (YX) and
 * 7 (YZ): the y-propagating modes the reference defines but never dispatches
 * (src/simulation.h:446-467, 519-544; kernels src/gpu.h:294-422), on the forward 1D line with
 * node offset N1 - J (the z and x modes use N1 - K and N1 - I). */
int fdtd_update_h_planewave(fdtd_sim *s, const int32_t time[2]);
int fdtd_update_e_planewave(fdtd_sim *s, const int32_t time[2]);
/* 'AddCurrents': time = [n, pol], box = [i0,i1,j0,j1,k0,k1]; src/simulation.h:256-286 (gpu.h:55-82) */
int fdtd_add_currents(fdtd_sim *s, const int32_t time[2], const int32_t box[6]);
/* 'OutputAnt': out has `size` floats, filled with the probe record [ncomp][T][nAnt];
 * src/simulation.h:621-638.  Without antennas (or before ReadySimulation) it prints "Antenna data
 * is not initialized." to stderr and leaves out untouched, as the reference (:634-637). */
int fdtd_output_ant(fdtd_sim *s, float *out, uint32_t size);
/* 'OutputVis': flag 1..3 = Ex,Ey,Ez, 4..6 = Hx,Hy,Hz, 0/-1/-2 = sigma x/y/z; out has `size` >= I*J*K
 * floats; src/simulation.h:572-619 */
int fdtd_output_vis(fdtd_sim *s, float *out, uint32_t size, int32_t flag);

/* ---------------- extensions (no counterpart in the reference) ---------------- */

/* CPML parameters [P, poly_m, kappa_max, alpha_max, alpha_m, sigma_factor]; between SetConst and
 * ReadySimulation.  sigma_max = sigma_factor*(m+1)/(eta0*dx) as src/simulation.h:87.  The
 * reference values are [10, 3, 1, 0.08, 1, 0.8]; kappa is applied per axis. */
int fdtd_set_cpml(fdtd_sim *s, const double params[6]);
/* TF/SF box inset from the grid faces (reference 15, src/simulation.h:102); before ReadySimulation */
int fdtd_set_tfsf(fdtd_sim *s, int32_t inset);
/* Graded (tensor-product) mesh, between SetConst and ReadySimulation: dx[I], dy[J], dz[K] are the
 * distances between consecutive grid planes along x, y, z (entry n: planes n and n+1); a null
 * pointer keeps the uniform dx of SetConst along that axis.  Replaces the reference's
 * non-uniform-mesh scaffolding (FDTD_UNIFORM_MESH and den_ey/den_ez/den_hy/den_hz,
 * src/simulation.h:2, 679-685, 718-724, allocated but never read): the H update differences E
 * over the cell n (spacing d[n]), the E update differences H over the dual cell n
 * ((d[n-1] + d[n]) / 2), both as factors dx/spacing on the per-axis den tables (den_h* and den_e*
 * of :861-912).  The CPML cells (P+1 at each end of an axis, the last entry included) must keep
 * dx; every spacing must be positive, and SetConst's dt within the Courant limit of the finest
 * cells, dt <= 1/(c sqrt(1/min dx^2 + 1/min dy^2 + 1/min dz^2)) (checked at ReadySimulation).  A graded mesh takes no plane wave (modes 1..7 are refused:
 * the 1D incident line is uniform); AddCurrents and step currents keep the reference's source
 * scaling (src/gpu.h:55-82).  Fused steps of a graded grid run the row kernel's full-slab march
 * (one column per lane; with per-cell mu, the per-op operations). */
int fdtd_set_mesh(fdtd_sim *s, const float *dx, const float *dy, const float *dz);
/* Plane-wave set-up used by Init1D.  mode 1..4, 6, 7: params[0] != 0 makes the 1D line carry the
 * grid's centre-column material profile (layered media).  mode 5: oblique incidence,
 * params = [theta, phi, psi] (radians), dispersion-matched 1D line. */
int fdtd_set_plane_wave(fdtd_sim *s, int32_t mode, const double params[4]);
/* 'CreateMu' src/simulation.h:28-68 (gateway src/simFDTD.cu:133-141): per-cell relative
 * permeabilities (three arrays of I*J*K floats, each > 0).  The H update then uses the per-cell
 * coefficients of K2 GPU_updateH_Full (src/gpu.h:4-30) instead of K1's scalar ones, in the per-op
 * kernels and in the fused step (12 more bytes per cell).  The reference's initialisation order is fixed: it sets init_mu
 * only before ReadySimulation but needs den_hx, which ReadySimulation fills (:30-32, :45), so it
 * never takes effect there; here CreateMu may come any time after SetConst and takes effect at
 * ReadySimulation, or at once when that has run. */
int fdtd_create_mu(fdtd_sim *s, const float *mu_x, const float *mu_y, const float *mu_z);
/* Step currents: after every step of fdtd_step / fdtd_run, AddCurrents(n, pol, box) with the
 * sample source[n] of SetSource -- the loop of GICsimulation.m:119-124 (UpdateH_All, UpdateH_PML,
 * UpdateE_All, UpdateE_PML, AddCurrents).  pol 0 turns them off; box as for 'AddCurrents'.
 * Probes sample a step before its currents, as in that loop. */
int fdtd_set_step_currents(fdtd_sim *s, int32_t pol, const int32_t box[6]);
/* Graph replay: fdtd_run records `steps` consecutive fused steps into one HIP graph and replays it
 * (the time index, probe record row and current sample of each recorded launch are read on the
 * device from a counter the replay advances), for runs with or without a plane wave and probes;
 * not with snapshots or timing, nor probes together with step currents; other steps run one by
 * one.  steps: an even count, 0 = off (the default). */
int fdtd_set_graph_steps(fdtd_sim *s, int32_t steps);
/* Per-cell uint8 material index into a table of n rows [sigma_x, sigma_y, sigma_z, eps_x, eps_y,
 * eps_z] (n <= 256); replaces the six UpdateMaterial arrays. */
int fdtd_set_material_index(fdtd_sim *s, const uint8_t *index, const float *table, int32_t n);
/* Probe components (1..6), recorded after every step; default {3} (Ez). */
int fdtd_set_probe_components(fdtd_sim *s, const int32_t *comps, int32_t n);
/* One full time step n (H update, plane wave, CPML, E update, plane wave, CPML, probes) as the
 * fused step (over a ping-ponged state: the TF/SF face sheets, then H and E of every cell in one
 * pass of the row kernel, interior and CPML-slab parts), or the per-op kernels for grids outside
 * its layouts; mode 0 = no plane wave. */
int fdtd_step(fdtd_sim *s, int32_t n, int32_t mode);
/* Steps n0 .. n0+count-1 (fused).  Within a run the slab part's launch of step n also fills the
 * face sheets of step n+1 and gathers the probes of step n-1, so a step is two launches.  If
 * snap_every > 0, component snap_comp (1..6) is copied to a device snapshot buffer after every
 * snap_every-th step (counted since ResetFieldAll). */
int fdtd_run(fdtd_sim *s, int32_t n0, int32_t count, int32_t mode, int32_t snap_every, int32_t snap_comp);
/* fdtd_run without any allocation (no counterpart in the reference): for callers that cannot
 * allocate device memory -- code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Steps n0 .. n0+count-1 exactly as fdtd_run (same kernels and results) with no
 * snapshots, timing or graph replay; the TF/SF face sheets and the row kernel's argument blocks
 * live in a file-scope device array of the library (48 MiB + 8 KiB per GPU, held until the process
 * ends) instead of buffers of the simulation.  Limit: a grid whose face sheets exceed 24 MiB per
 * buffer (8 (nj nk + ni nk + ni nj) floats for the TF/SF box extents ni, nj, nk; about a 512^2 box
 * face) is refused with FDTD_ERR_INVALID before any work.  All calls on one GPU share that array:
 * calls that overlap on one GPU (other threads or streams) race.  The simulation's state (both
 * state sets) is allocated by ReadySimulation. */
int fdtd_run_noalloc(fdtd_sim *s, int32_t n0, int32_t count, int32_t mode);
/* Number of snapshots held and their copy to host: out has count*I*J*K floats. */
int fdtd_snapshot_count(fdtd_sim *s, int32_t *count);
int fdtd_output_snapshots(fdtd_sim *s, float *out, uint64_t size);
/* Device bytes the simulation holds now (state, materials, tables, line histories, probe record,
 * argument blocks and snapshot buffers).  Repeated UpdateCuda, Init1D and SetAntennaLocations calls
 * reuse or replace their buffers, so the figure does not grow with them. */
int fdtd_device_bytes(fdtd_sim *s, uint64_t *bytes);
/* Measurement (no counterpart in the reference): the stream ceiling of this grid -- a kernel that
 * reads the six field components and the material byte of every cell from one state set and
 * writes the six components to the other (49 bytes per cell, the fused step's compulsory traffic
 * with no stencil), in the fused step's layout; *ms = its average time over reps launches.  It
 * overwrites the fields (call ResetFieldAll before stepping again).  Needs the uint8 material
 * index form. */
int fdtd_stream_ceiling(fdtd_sim *s, int32_t reps, double *ms);
/* Wait for all queued work of this simulation. */
int fdtd_synchronize(fdtd_sim *s);
/* Read any internal array to host for verification: what = 0..5 fields, 6..11 Ce/Ch per cell,
 * 20..31 psi (reference order Exy,Exz,Eyx,Eyz,Ezx,Ezy,Hxy,Hxz,Hyx,Hyz,Hzx,Hzy), 40/41 E/H line
 * history, 42/43 inverse line history.  *count receives the element count when out is NULL. */
int fdtd_read_array(fdtd_sim *s, int32_t what, float *out, uint64_t *count);
/* Host-only part of ReadySimulation (CPML profiles, 1/kappa, material table or per-cell
 * coefficients, oblique plane-wave tables); makes no device call, so it can run without a GPU.
 * ReadySimulation calls it itself. */
int fdtd_prepare(fdtd_sim *s);
/* Read a host set-up array (after fdtd_prepare): 0..3 bE,cE,bH,cH [2P]; 4,5 kappa E,H [P];
 * 6..8 1/kappa E along x,y,z; 9..11 1/kappa H; 12 scalars [Chh, Che, sigma_max, N1, i0, i1, j0,
 * j1, k0, k1, P, dx/ds, source node]; 13..15 Ce x,y,z per cell; 16..18 Ch x,y,z per cell;
 * 19 material id per cell; 20 + 4*axis + {0,1,2,3} oblique tables (E integer, E half, H integer,
 * H half); 32 polarisation [eE x,y,z, eH x,y,z].  *count receives the element count. */
int fdtd_read_setup(fdtd_sim *s, int32_t what, float *out, uint64_t *count);
/* Launch tuning of the fused step's row kernel: rows per workgroup (1, 2, 4, 8 or 16, at most 1024
 * threads; rows of a workgroup exchange their neighbour values through LDS; 0 = automatic),
 * z-chunk depth in planes (0 = automatic) and columns per lane (1, 2 or 4; 0 = automatic; the
 * grid's I must be a multiple of it and a row at most 256 of it).  With 1 column per lane and I a
 * multiple of 64 the row's two ends share one wave (split edge segment); -1 = 1 column per lane
 * without that split.  An unsupported choice makes
 * the next step fail with FDTD_ERR_INVALID. */
int fdtd_set_tuning(fdtd_sim *s, int32_t rows_per_group, int32_t z_chunk, int32_t cols_per_lane);
/* The launch tuning in force: out = [rows per group, z chunk, columns per lane (-1: one column
 * without the split edge segment), slab chunk]; 0 means automatic. */
int fdtd_get_tuning(fdtd_sim *s, int32_t out[4]);
/* Planes per z chunk of the row kernel's slab part (its y-band rows over the interior planes, and
 * the z slabs, at most the whole slab); 0 (default) = the interior part's chunk depth for the
 * y-band rows and the whole z slab where it is at most two chunks deep.  Results do not depend on
 * it.  (No counterpart in the reference.) */
int fdtd_set_slab_chunk(fdtd_sim *s, int32_t planes);
/* Addressing of the fused step's row kernel: 0 (default) = buffer resources with 32-bit offsets
 * where every array it touches is under 4 GiB, 64-bit global addresses otherwise; 1 = the 64-bit
 * global path on any grid (it serves grids of more than about 178 M cells; forcing it lets a
 * small grid test that path).  Results do not depend on it.  (No counterpart in the reference.) */
int fdtd_set_addressing(fdtd_sim *s, int32_t mode);
/* Batched engine (no counterpart in the reference; for many independent grids, each much smaller
 * than the GPU).  fdtd_batch_add queues `count` steps of a ready simulation from time index n0
 * with plane-wave mode `mode`; fdtd_batch_run advances every queued simulation to the end of its
 * steps.  Each batch step launches every kernel of the fused step once for all grids still
 * running, so the grids share the GPU and the launch count does not grow with their number.
 * Results, psi, probe records and time steps are those of fdtd_run on each simulation alone
 * (snapshots are not taken in a batch).  A simulation is queued at most once per batch; grids the
 * fused step does not cover run alone after the batch. */
int fdtd_batch_create(fdtd_batch **out);
int fdtd_batch_destroy(fdtd_batch *b);
int fdtd_batch_add(fdtd_batch *b, fdtd_sim *s, int32_t n0, int32_t count, int32_t mode);
int fdtd_batch_clear(fdtd_batch *b);
int fdtd_batch_run(fdtd_batch *b);
/* How fdtd_batch_run advances the grids: 1 = merged launches (each kernel of the step launched once
 * for all grids, in up to four interleaved sub-batches); 2 = one stream per grid (each grid's
 * steps queued on its own stream); 3 = device work queue (as 1, but each row-kernel launch is a
 * persistent launch of as many workgroups as fit at once, which pull (grid, z-chunk tile) work
 * items from an atomic counter in device memory until none is left); 0 (default) = automatic, at
 * present merged launches (the fastest of the forms on a mix of 32..268-cell grids).  Results are
 * the same either way. */
int fdtd_batch_set_strategy(fdtd_batch *b, int32_t strategy);
/* Request server (no counterpart in the reference; the served form of the solver).  Solves arrive
 * one at a time and are advanced together with continuous batching on one GPU.
 *   fdtd_server_create: hbm_budget = device bytes the admitted requests may hold together,
 *     max_active = requests advanced at once, steps_per_round = steps per round.
 *   fdtd_server_submit: queue a solve.  s has its host set-up done (SetConst, source, materials or
 *     material index, probe locations and components, CPML / TF/SF / plane-wave options) and is not
 *     ready; the server owns it from here on (do not destroy it).  At admission the server runs
 *     ReadySimulation and, for mode != 0, Init1D(rate, pw_source) (pw_source: T samples, copied
 *     here), then steps n = 0 .. count-1 in plane-wave mode `mode`.  t_arrive: arrival time on the
 *     server clock (fdtd_server_now), or < 0 for now.  A request is charged an upper bound of its
 *     device bytes from admission to release; one above the budget alone is refused
 *     (FDTD_ERR_NOMEM).
 *   fdtd_server_round: one round: admit queued requests in arrival order while slots and budget
 *     allow, advance every admitted request by up to steps_per_round steps (batched engine), retire
 *     the ones that reached their last step.  Returns after the round's GPU work is done.
 *   fdtd_server_collect: up to max completed requests: ids[k] and times[3k..3k+2] = arrival,
 *     admission and completion times (server clock, seconds).  Their results stay readable through
 *     the simulation handle (OutputAnt, OutputVis, fdtd_read_array) until fdtd_server_release,
 *     which destroys the simulation and returns its bytes to the budget.
 * Each request's fields, psi and probe record equal those of fdtd_run of the same simulation alone. */
typedef struct fdtd_server fdtd_server;
int fdtd_server_create(fdtd_server **out, uint64_t hbm_budget, int32_t max_active, int32_t steps_per_round);
int fdtd_server_destroy(fdtd_server *srv);
int fdtd_server_now(fdtd_server *srv, double *t);
int fdtd_server_submit(fdtd_server *srv, fdtd_sim *s, float rate, const float *pw_source, int32_t count, int32_t mode,
                       double t_arrive, int64_t *id);
int fdtd_server_round(fdtd_server *srv, int32_t *active, int32_t *queued);
int fdtd_server_collect(fdtd_server *srv, int64_t *ids, double *times, int32_t max, int32_t *n);
int fdtd_server_release(fdtd_server *srv, int64_t id);
/* Error of a request whose set-up failed at admission (ReadySimulation / Init1D): *code = its
 * FDTD_ERR_* (0 for a request admitted normally) and its message in msg (msg_len bytes, may be
 * NULL).  Such a request is completed at once: collect returns it with admission time -1, it is
 * charged nothing against the budget, and release frees what its set-up allocated. */
int fdtd_server_request_error(fdtd_server *srv, int64_t id, int32_t *code, char *msg, int32_t msg_len);
/* Device bytes a request is charged while admitted (its upper bound, fixed at submission). */
int fdtd_server_request_bytes(fdtd_server *srv, int64_t id, uint64_t *bytes);
/* Bytes charged to admitted requests now and at most so far; cell updates done. */
int fdtd_server_stats(fdtd_server *srv, uint64_t *bytes_in_use, uint64_t *bytes_peak, int64_t *cell_updates);
/* Time the fused step at time index n, plane-wave mode `mode`, for each available columns-per-lane
 * and z-chunk choice and two tile heights (reps steps each, in the launch form of a run) and keep
 * the fastest, unless the automatic choice is within 1% of it (*best_ms: its average step time).
 * The steps are run dry: the fields, psi, probe record and time step are unchanged.  The choice
 * is also entered in the size-class table (below). */
int fdtd_autotune(fdtd_sim *s, int32_t n, int32_t mode, int32_t reps, double *best_ms);
/* The size-class tuning table: the autotuner's choices per class of grid (I; floor(log2 J);
 * floor(log2 K); P; index-form materials; plane wave; per-cell mu), shared by the simulations of
 * the process and read from / written to the JSON file FDTD_TUNE_CACHE names; without it the
 * table shipped beside the library (tuning_gfx950.json, written by tune_table.py) is read.  fdtd_run
 * and the batched engine consult it for a simulation whose tuning is automatic: policy 1 (the
 * default) applies the class's entry when there is one; policy 2 also autotunes a class that has
 * none (at the run's first time index) and adds it; 0 keeps the automatic choice. */
int fdtd_set_autotune(fdtd_sim *s, int32_t policy);
/* Average device time of the fused step over the last fdtd_run with timing on: HIP events on the
 * simulation's stream around every `every`-th step of the run (steps 0, every, 2 every, ...;
 * every = 1: all steps, 0: off).  Each pair of events costs the stream a few microseconds, so a
 * long run is timed on a sample of its steps. *launches = the steps timed. */
int fdtd_set_timing(fdtd_sim *s, int32_t every);
int fdtd_kernel_time(fdtd_sim *s, double *avg_ms, int64_t *launches);
/* Average device time of the row kernel's interior part alone (the rows j in [P+1, J-P) of the
 * planes k in [P, K-P), the step's dominant launch) over the same run. */
int fdtd_row_kernel_time(fdtd_sim *s, double *avg_ms);

#ifdef __cplusplus
}
#endif

#endif /* MAXWELL_FDTD_H */
