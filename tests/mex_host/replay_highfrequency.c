/* Test driver: replays the op sequence of highfrequency.m:63-128 through the C MEX gateway
 * (simFDTD_gateway.c) with the fake mx layer, on a small vacuum grid.
 *   replay_highfrequency errors                         -> error-identifier checks (no GPU needed)
 *   replay_highfrequency run MODE N T src.bin out.bin   -> writes the polarised E component after T steps  */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

int call_mex(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
extern char g_err_id[];
extern char g_err_msg[];

static mxArray *str(const char *s) {
  const mwSize d[2] = {1, strlen(s)};
  mxArray *a = mxCreateNumericArray(2, d, mxCHAR_CLASS, mxREAL);
  memcpy(mxGetData(a), s, strlen(s));
  return a;
}
static mxArray *vec(mxClassID c, size_t n) {
  const mwSize d[2] = {1, n};
  return mxCreateNumericArray(2, d, c, mxREAL);
}
static mxArray *i32s(int n, const int *v) {
  mxArray *a = vec(mxINT32_CLASS, n);
  memcpy(mxGetData(a), v, n * sizeof(int));
  return a;
}
static mxArray *f32s(int n, const float *v) {
  mxArray *a = vec(mxSINGLE_CLASS, n);
  memcpy(mxGetData(a), v, n * sizeof(float));
  return a;
}

static int op(const char *name, mxArray *h, int nx, mxArray **x, mxArray **out) {
  const mxArray *in[8];
  int n = 0;
  in[n++] = str(name);
  if (h) in[n++] = h;
  for (int q = 0; q < nx; ++q) in[n++] = x[q];
  mxArray *pl[1] = {NULL};
  int rc = call_mex(1, pl, n, in);
  if (out) *out = pl[0];
  return rc;
}

static int expect_err(const char *what, int rc, const char *id) {
  if (rc == 1 && strcmp(g_err_id, id) == 0) {
    printf("ok %s -> %s\n", what, id);
    return 0;
  }
  printf("FAIL %s: rc=%d id=%s msg=%s\n", what, rc, g_err_id, g_err_msg);
  return 1;
}

int main(int argc, char **argv) {
  if (argc >= 2 && strcmp(argv[1], "errors") == 0) {
    int bad = 0;
    mxArray *h = NULL;
    if (op("CreateSimulation", NULL, 0, NULL, &h)) { printf("FAIL create %s\n", g_err_msg); return 1; }
    const mxArray *none[1] = {NULL};
    bad |= expect_err("no args", call_mex(0, NULL, 0, none), "simFDTD:InvalidInput");
    bad |= expect_err("no handle", op("UpdateH_All", NULL, 0, NULL, NULL), "simFDTD:InvalidInput");
    bad |= expect_err("unknown op", op("FooBar", h, 0, NULL, NULL), "simFDTD:UnknownOperation");
    bad |= expect_err("setconst arity", op("SetConst", h, 0, NULL, NULL), "simFDTD:InvalidInput");
    float fr[3] = {3e9f, 0.01f, 1.9e-11f};
    int td[5] = {10, 40, 40, 40, 0};
    mxArray *a[2] = {f32s(3, fr), i32s(5, td)};
    if (op("SetConst", h, 2, a, NULL)) { printf("FAIL setconst %s\n", g_err_msg); bad = 1; }
    bad |= expect_err("setconst twice", op("setconst", h, 2, a, NULL), "simFDTD:AlreadyInitialized");
    bad |= expect_err("update before ready", op("UpdateE_All", h, 0, NULL, NULL), "simFDTD:InvalidInput");
    {
      /* SetMesh (graded mesh extension): arity, a non-positive spacing, then a valid mesh and a
       * uniform one (empty arrays) */
      bad |= expect_err("setmesh arity", op("SetMesh", h, 0, NULL, NULL), "simFDTD:InvalidInput");
      float d[40];
      for (int q = 0; q < 40; ++q) d[q] = 0.01f;
      mxArray *dx = f32s(40, d), *dy = f32s(40, d), *none = vec(mxSINGLE_CLASS, 0);
      d[20] = 0.0f;
      mxArray *dz_bad = f32s(40, d);
      mxArray *m_bad[3] = {dx, dy, dz_bad};
      bad |= expect_err("setmesh zero spacing", op("SetMesh", h, 3, m_bad, NULL), "simFDTD:InvalidInput");
      mxArray *m_ok[3] = {dx, none, dy};
      if (op("SetMesh", h, 3, m_ok, NULL)) { printf("FAIL setmesh %s\n", g_err_msg); bad = 1; }
      else printf("ok setmesh\n");
      mxArray *m_uni[3] = {none, none, none};
      if (op("SetMesh", h, 3, m_uni, NULL)) { printf("FAIL setmesh uniform %s\n", g_err_msg); bad = 1; }
      else printf("ok setmesh uniform\n");
    }
    {
      /* OutputAnt without antennas: the reference prints a message and returns the zero-filled
       * array (src/simulation.h:634-637) */
      mxArray *sz = vec(mxUINT32_CLASS, 1), *out = NULL;
      *(unsigned *)mxGetData(sz) = 10;
      mxArray *a[1] = {sz};
      if (op("OutputAnt", h, 1, a, &out) || !out || mxGetNumberOfElements(out) != 10) {
        printf("FAIL OutputAnt without antennas: %s\n", g_err_msg);
        bad = 1;
      } else {
        const float *v = (const float *)mxGetData(out);
        for (int q = 0; q < 10; ++q)
          if (v[q] != 0.0f) { printf("FAIL OutputAnt without antennas: nonzero\n"); bad = 1; break; }
        printf("ok OutputAnt without antennas -> zeros\n");
      }
    }
    if (op("DeleteSimulation", h, 0, NULL, NULL)) bad = 1;
    /* a deleted handle is rejected (the reference leaves its use undefined) */
    bad |= expect_err("use after delete", op("UpdateH_All", h, 0, NULL, NULL), "simFDTD:InvalidInput");
    bad |= expect_err("delete twice", op("DeleteSimulation", h, 0, NULL, NULL), "simFDTD:InvalidInput");
    {
      mxArray *bogus = vec(mxUINT64_CLASS, 1);
      *(unsigned long long *)mxGetData(bogus) = 0x12345678ull;
      bad |= expect_err("never created", op("ResetFieldAll", bogus, 0, NULL, NULL), "simFDTD:InvalidInput");
    }
    printf(bad ? "ERRORS FAILED\n" : "ERRORS OK\n");
    return bad;
  }
  if (argc < 7 || strcmp(argv[1], "run") != 0) {
    fprintf(stderr, "usage: %s errors | run MODE N T src.bin out.bin\n", argv[0]);
    return 2;
  }
  const int mode = atoi(argv[2]), N = atoi(argv[3]), T = atoi(argv[4]);
  float *src = (float *)malloc(sizeof(float) * T);
  FILE *f = fopen(argv[5], "rb");
  if (!f || fread(src, sizeof(float), T, f) != (size_t)T) { fprintf(stderr, "cannot read %s\n", argv[5]); return 2; }
  fclose(f);
  const size_t IJK = (size_t)N * N * N;
  const float c0 = 299792458.0f, dx = 0.01f;
  const float fr[3] = {3e9f, dx, (float)(0.99 * 0.01 / (1.7320508075688772 * 299792458.0))};
  (void)c0;
  const int td[5] = {T, N, N, N, 0};
  mxArray *h = NULL;
#define OK(x)                                                        \
  do {                                                               \
    if (x) { fprintf(stderr, "op failed: %s %s\n", g_err_id, g_err_msg); return 1; } \
  } while (0)
  OK(op("CreateSimulation", NULL, 0, NULL, &h));
  { mxArray *a[2] = {f32s(3, fr), i32s(5, td)}; OK(op("SetConst", h, 2, a, NULL)); }
  { mxArray *a[1] = {f32s(T, src)}; OK(op("SetSource", h, 1, a, NULL)); }
  float *ones = (float *)malloc(sizeof(float) * IJK), *zeros = (float *)calloc(IJK, sizeof(float));
  for (size_t q = 0; q < IJK; ++q) ones[q] = 1.0f;
  for (int flag = 1; flag <= 6; ++flag) {
    mxArray *a[2] = {i32s(1, &flag), f32s((int)IJK, flag <= 3 ? zeros : ones)};
    OK(op("UpdateMaterial", h, 2, a, NULL));
  }
  OK(op("ReadySimulation", h, 0, NULL, NULL));
  {
    /* no antennas were set: OutputAnt leaves the zero-filled array (src/simulation.h:634-637) */
    mxArray *sz = vec(mxUINT32_CLASS, 1), *ant = NULL;
    *(unsigned *)mxGetData(sz) = 16;
    mxArray *a[1] = {sz};
    OK(op("OutputAnt", h, 1, a, &ant));
    for (int q = 0; q < 16; ++q)
      if (((const float *)mxGetData(ant))[q] != 0.0f) { fprintf(stderr, "OutputAnt without antennas: nonzero\n"); return 1; }
  }
  { const float one = 1.0f; mxArray *a[2] = {f32s(1, &one), f32s(T, src)}; OK(op("Init1D", h, 2, a, NULL)); }
  OK(op("UpdateCuda", h, 0, NULL, NULL));
  OK(op("ResetFieldAll", h, 0, NULL, NULL));
  for (int n = 0; n < T; ++n) {
    int t[2] = {n, mode};
    mxArray *tm[1] = {i32s(2, t)};
    OK(op("UpdateH_All", h, 0, NULL, NULL));
    OK(op("UpdateH_PlaneWave", h, 1, tm, NULL));
    OK(op("UpdateH_PML", h, 0, NULL, NULL));
    OK(op("UpdateE_All", h, 0, NULL, NULL));
    OK(op("UpdateE_PlaneWave", h, 1, tm, NULL));
    OK(op("UpdateE_PML", h, 0, NULL, NULL));
  }
  mxArray *out = NULL;
  {
    mxArray *sz = vec(mxUINT32_CLASS, 1);
    *(unsigned *)mxGetData(sz) = (unsigned)IJK;
    /* the polarised component: mode 1/4 Ey, mode 2/6 Ex, mode 3/7 Ez */
    const int flag = (mode == 2 || mode == 6) ? 1 : (mode == 3 || mode == 7) ? 3 : 2;
    mxArray *a[2] = {sz, i32s(1, &flag)};
    OK(op("OutputVis", h, 2, a, &out));
  }
  f = fopen(argv[6], "wb");
  fwrite(mxGetData(out), sizeof(float), IJK, f);
  fclose(f);
  OK(op("DeleteSimulation", h, 0, NULL, NULL));
  printf("replayed %d steps of mode %d on %d^3\n", T, mode, N);
  return 0;
}
