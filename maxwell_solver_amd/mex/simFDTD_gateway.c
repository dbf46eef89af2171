/*
 * simFDTD_gateway.c -- MATLAB MEX gateway over libfdtd_hip.so.
 *
 * Drop-in for the reference's src/simFDTD.cu:19-187: the same op strings (case-insensitive), the
 * same argument order and MATLAB classes, the same outputs (uint64 handle, single row vectors) and
 * the same error identifiers.  Build against MATLAB's mex.h:
 *     mex -I../../include simFDTD_gateway.c -L.. -lfdtd_hip
 * Extension ops (no counterpart in the reference): SetCPML, SetTFSF, SetPlaneWave,
 * SetMaterialIndex, SetProbeComponents, SetMesh, Step, Run, OutputSnapshots.
 */
#include <ctype.h>
#include <stdint.h>
#include <string.h>

#include "fdtd.h"
#include "mex.h"

static void lower_ascii(char *s) {
  for (; *s; ++s) *s = (char)tolower((unsigned char)*s);
}

static void check(int rc) {
  if (rc) mexErrMsgIdAndTxt(fdtd_last_error_id(), "%s", fdtd_last_error());
}

static void need(int nrhs, int n, const char *msg) {
  if (nrhs < n) mexErrMsgIdAndTxt("simFDTD:InvalidInput", "%s", msg);
}

static int32_t i32_at(const mxArray *a, int k) { return ((const int32_t *)mxGetData(a))[k]; }

static mxArray *single_row(size_t n) {
  const mwSize dims[2] = {1, (mwSize)n};
  return mxCreateNumericArray(2, dims, mxSINGLE_CLASS, mxREAL);
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  (void)nlhs;
  if (nrhs < 1) mexErrMsgIdAndTxt("simFDTD:InvalidInput", "At least one input required: operation string.");
  char *op = mxArrayToString(prhs[0]);
  if (!op) mexErrMsgIdAndTxt("simFDTD:InvalidInput", "At least one input required: operation string.");
  lower_ascii(op);

  if (strcmp(op, "createsimulation") == 0) {
    fdtd_sim *s = NULL;
    check(fdtd_create(&s));
    const mwSize dims[2] = {1, 1};
    mxArray *out = mxCreateNumericArray(2, dims, mxUINT64_CLASS, mxREAL);
    *(uint64_t *)mxGetData(out) = (uint64_t)(uintptr_t)s;
    plhs[0] = out;
    mxFree(op);
    return;
  }
  need(nrhs, 2, "Simulation pointer required for this operation.");
  fdtd_sim *s = (fdtd_sim *)(uintptr_t)(*(const uint64_t *)mxGetData(prhs[1]));

  if (strcmp(op, "deletesimulation") == 0) check(fdtd_destroy(s));
  else if (strcmp(op, "updatee_all") == 0) check(fdtd_update_e_all(s));
  else if (strcmp(op, "updateh_all") == 0) check(fdtd_update_h_all(s));
  else if (strcmp(op, "updatee_pml") == 0) check(fdtd_update_e_pml(s));
  else if (strcmp(op, "updateh_pml") == 0) check(fdtd_update_h_pml(s));
  else if (strcmp(op, "updatecuda") == 0) check(fdtd_update_coeffs(s));
  else if (strcmp(op, "addcurrents") == 0) {
    need(nrhs, 4, "time and indices required.");
    check(fdtd_add_currents(s, (const int32_t *)mxGetData(prhs[2]), (const int32_t *)mxGetData(prhs[3])));
  } else if (strcmp(op, "readysimulation") == 0) check(fdtd_ready(s));
  else if (strcmp(op, "resetfieldall") == 0) check(fdtd_reset_fields(s));
  else if (strcmp(op, "updatee_planewave") == 0) {
    need(nrhs, 3, "time required.");
    check(fdtd_update_e_planewave(s, (const int32_t *)mxGetData(prhs[2])));
  } else if (strcmp(op, "updateh_planewave") == 0) {
    need(nrhs, 3, "time required.");
    check(fdtd_update_h_planewave(s, (const int32_t *)mxGetData(prhs[2])));
  } else if (strcmp(op, "setconst") == 0) {
    need(nrhs, 4, "freqRes and TimeDim required.");
    check(fdtd_set_const(s, (const float *)mxGetData(prhs[2]), (const int32_t *)mxGetData(prhs[3])));
  } else if (strcmp(op, "setsource") == 0) {
    need(nrhs, 3, "SourceIn required.");
    if (nrhs > 3) mexPrintf("Update Source Mode!\n");
    check(fdtd_set_source(s, (const float *)mxGetData(prhs[2]), nrhs > 3));
  } else if (strcmp(op, "setantennalocations") == 0) {
    need(nrhs, 3, "Locations required.");
    check(fdtd_set_antenna_locations(s, (const int32_t *)mxGetData(prhs[2])));
  } else if (strcmp(op, "createmu") == 0) {
    need(nrhs, 5, "mu_x, mu_y, mu_z required.");
    check(fdtd_create_mu(s, (const float *)mxGetData(prhs[2]), (const float *)mxGetData(prhs[3]),
                         (const float *)mxGetData(prhs[4])));
  } else if (strcmp(op, "init1d") == 0) {
    need(nrhs, 4, "rate and PW_source required.");
    check(fdtd_init_1d(s, *(const float *)mxGetData(prhs[2]), (const float *)mxGetData(prhs[3])));
  } else if (strcmp(op, "updatematerial") == 0) {
    need(nrhs, 4, "flag, values required.");
    check(fdtd_update_material(s, i32_at(prhs[2], 0), (const float *)mxGetData(prhs[3])));
  } else if (strcmp(op, "outputant") == 0) {
    need(nrhs, 3, "OutAntSize required.");
    const uint32_t n = *(const uint32_t *)mxGetData(prhs[2]);
    plhs[0] = single_row(n);
    check(fdtd_output_ant(s, (float *)mxGetData(plhs[0]), n));
  } else if (strcmp(op, "outputvis") == 0) {
    need(nrhs, 4, "OutAntSize and flag required.");
    const uint32_t n = *(const uint32_t *)mxGetData(prhs[2]);
    plhs[0] = single_row(n);
    check(fdtd_output_vis(s, (float *)mxGetData(plhs[0]), n, i32_at(prhs[3], 0)));
  }
  /* ---- extensions ---- */
  else if (strcmp(op, "setcpml") == 0) {
    need(nrhs, 3, "CPML parameters required.");
    check(fdtd_set_cpml(s, (const double *)mxGetData(prhs[2])));
  } else if (strcmp(op, "settfsf") == 0) {
    need(nrhs, 3, "inset required.");
    check(fdtd_set_tfsf(s, i32_at(prhs[2], 0)));
  } else if (strcmp(op, "setplanewave") == 0) {
    need(nrhs, 4, "mode and parameters required.");
    check(fdtd_set_plane_wave(s, i32_at(prhs[2], 0), (const double *)mxGetData(prhs[3])));
  } else if (strcmp(op, "setmaterialindex") == 0) {
    need(nrhs, 4, "index and table required.");
    check(fdtd_set_material_index(s, (const uint8_t *)mxGetData(prhs[2]), (const float *)mxGetData(prhs[3]),
                                  (int32_t)(mxGetNumberOfElements(prhs[3]) / 6)));
  } else if (strcmp(op, "setmesh") == 0) {
    /* single spacings along x, y, z (I, J, K entries), an empty array for a uniform axis */
    need(nrhs, 5, "spacings along x, y and z required (empty: uniform).");
    const float *d[3];
    for (int a = 0; a < 3; ++a) d[a] = mxGetNumberOfElements(prhs[2 + a]) ? (const float *)mxGetData(prhs[2 + a]) : NULL;
    check(fdtd_set_mesh(s, d[0], d[1], d[2]));
  } else if (strcmp(op, "setprobecomponents") == 0) {
    need(nrhs, 3, "components required.");
    check(fdtd_set_probe_components(s, (const int32_t *)mxGetData(prhs[2]), (int32_t)mxGetNumberOfElements(prhs[2])));
  } else if (strcmp(op, "step") == 0) {
    need(nrhs, 3, "time required.");
    check(fdtd_step(s, i32_at(prhs[2], 0), i32_at(prhs[2], 1)));
  } else if (strcmp(op, "run") == 0) {
    need(nrhs, 3, "[n0, count, mode, snap_every, snap_comp] required.");
    const size_t m = mxGetNumberOfElements(prhs[2]);
    check(fdtd_run(s, i32_at(prhs[2], 0), i32_at(prhs[2], 1), i32_at(prhs[2], 2), m > 3 ? i32_at(prhs[2], 3) : 0,
                   m > 4 ? i32_at(prhs[2], 4) : 3));
  } else if (strcmp(op, "outputsnapshots") == 0) {
    need(nrhs, 3, "OutSize required.");
    const uint32_t n = *(const uint32_t *)mxGetData(prhs[2]);
    plhs[0] = single_row(n);
    check(fdtd_output_snapshots(s, (float *)mxGetData(plhs[0]), n));
  } else {
    mexErrMsgIdAndTxt("simFDTD:UnknownOperation", "Unknown operation: %s", op);
  }
  mxFree(op);
}
