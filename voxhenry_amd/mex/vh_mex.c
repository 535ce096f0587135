/*
 * vh_mex.c -- MATLAB gateway of libvoxhenry (interleaved complex API).
 *
 *   mex -R2018a -I<repo>/include vh_mex.c -L<repo>/voxhenry_amd -lvoxhenry
 *
 * The shadow .m files in matlab/ (same names and signatures as the
 * reference's, first on the MATLAB path) forward the executer's calls here:
 *
 *   lse_sparse_precon_prepare   -> vh_mex('prepare', dx, z_real, z_imag, idxS, st_sparse_precon, Ae [, type])
 *   lse_matvect_mult            -> JOut = vh_mex('matvec', J [, jwmu])
 *   lse_sparse_precon_multiply  -> y = vh_mex('precon_apply', v [, prectol])
 *   lse_sparse_schur_compl_multiply -> y = vh_mex('schur_multiply', v)
 *   pgmres / fpgmres            -> [x,flag,relres,iter,resvec] = vh_mex('pgmres' | 'fpgmres', b, restart, tol, maxit)
 *
 * the batched tier (every port x frequency system in one device solve):
 *
 *   [X,flag,relres,iters,inner] = vh_mex('solve_batch', B, jwmu, pidx, 'fp64' | 'fp64_true' | 'mixed', tol)
 *
 * and the lower-level commands the prepare step is made of:
 *
 *   st2 = vh_mex('geom', L, M, N, dx, idxS)   grid, 1-based voxel indices; the
 *         Green's tensor is generated on the GPU (lse_generate_circulant_tensor
 *         with ko = 1); st2 = st_sparse_precon2 (1x3 complex)
 *   vh_mex('incidence', Ae)                  Ae with the port rows removed (num_node x 5nv sparse)
 *   vh_mex('impedance', z_real [, z_imag])   full L x M x N grids (VoxHenry_executer.m:231-259)
 *   vh_mex('reset')                          release the device objects
 *
 * One process-wide context is created on the first command that needs the
 * device; the gateway then locks itself (mexLock) and releases everything
 * from its mexAtExit handler or on 'reset'.  Every argument is checked before
 * the device is touched: a bad call raises 'voxhenry:usage', a call before
 * the state it needs 'voxhenry:state', a failure inside the library
 * 'voxhenry:lib' with the library's message (vh_last_error).  Temporaries are
 * mxMalloc'ed and freed before returning; on an error MATLAB frees them.
 */
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"
#include "voxhenry.h"

static struct {
  vh_ctx* ctx;
  vh_geom* geom;
  vh_precon* pre;
  int64_t dims[3];
  double dx;
  int64_t nv;
  int64_t* idx; /* 0-based voxel indices, malloc'ed: lives across calls */
  int64_t num_curr, num_node, n;
  double st2[6];
  int has_inc, has_z, prepared, has_jw;
  int z_has_imag; /* the impedance was set with a nonzero z_imag (two-fluid: depends on f) */
  double jw[2]; /* j*omega*mu of the last prepare */
  int pc_type;
  double prectol;
} S = {0};

static char msgbuf[512];

static void fail(const char* id, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msgbuf, sizeof msgbuf, fmt, ap);
  va_end(ap);
  mexErrMsgIdAndTxt(id, "%s", msgbuf);
}

static void lib(int rc) {
  if (rc != 0) fail("voxhenry:lib", "libvoxhenry error %d: %s", rc, vh_last_error());
}

static void release(void) {
  if (S.pre) vh_precon_destroy(S.pre);
  if (S.geom) vh_geom_destroy(S.geom);
  if (S.ctx) vh_ctx_destroy(S.ctx);
  free(S.idx);
  memset(&S, 0, sizeof S);
}

static void at_exit(void) { release(); }

static void ensure_ctx(void) {
  if (S.ctx) return;
  vh_ctx* c = NULL;
  lib(vh_ctx_create(0, &c));
  S.ctx = c;
  S.prectol = 0.1;
  mexLock();
  mexAtExit(at_exit);
}

/* ---- argument checks ----------------------------------------------------- */

static int is_real_double(const mxArray* a) { return mxIsDouble(a) && !mxIsComplex(a) && !mxIsSparse(a); }

static double real_scalar(const mxArray* a, const char* what) {
  if (!is_real_double(a) || mxGetNumberOfElements(a) != 1) fail("voxhenry:usage", "%s must be a real double scalar", what);
  const double v = mxGetDoubles(a)[0];
  if (!isfinite(v)) fail("voxhenry:usage", "%s must be finite", what);
  return v;
}

static int64_t int_scalar(const mxArray* a, const char* what, int64_t lo) {
  const double v = real_scalar(a, what);
  if (v != floor(v) || v < (double)lo || v > 9.0e15) fail("voxhenry:usage", "%s must be an integer >= %lld", what, (long long)lo);
  return (int64_t)v;
}

/* first element of a double array as a complex number (st_sparse_precon(1), jwmu) */
static void complex_first(const mxArray* a, const char* what, double out[2]) {
  if (!mxIsDouble(a) || mxIsSparse(a) || mxGetNumberOfElements(a) < 1) fail("voxhenry:usage", "%s must be a double array", what);
  if (mxIsComplex(a)) {
    const mxComplexDouble* c = mxGetComplexDoubles(a);
    out[0] = c[0].real;
    out[1] = c[0].imag;
  } else {
    out[0] = mxGetDoubles(a)[0];
    out[1] = 0.0;
  }
  if (!isfinite(out[0]) || !isfinite(out[1])) fail("voxhenry:usage", "%s must be finite", what);
}

/* Column block of `rows` rows as interleaved complex data: the array's own
 * storage when it is complex, else an mxMalloc'ed copy (*tmp) with zero
 * imaginary parts -- the executer's first Arnoldi vector is real-typed. */
static const double* columns(const mxArray* a, int64_t rows, const char* what, int64_t* ncol, double** tmp) {
  *tmp = NULL;
  if (!mxIsDouble(a) || mxIsSparse(a)) fail("voxhenry:usage", "%s must be a full double array", what);
  if ((int64_t)mxGetM(a) != rows || mxGetN(a) < 1)
    fail("voxhenry:usage", "%s must have %lld rows (has %lld)", what, (long long)rows, (long long)mxGetM(a));
  *ncol = (int64_t)mxGetN(a);
  if (mxIsComplex(a)) return (const double*)mxGetComplexDoubles(a);
  const size_t cnt = (size_t)rows * (size_t)*ncol;
  const double* re = mxGetDoubles(a);
  double* c = (double*)mxMalloc(2 * cnt * sizeof(double));
  for (size_t i = 0; i < cnt; ++i) {
    c[2 * i] = re[i];
    c[2 * i + 1] = 0.0;
  }
  *tmp = c;
  return c;
}

/* idxS (1-based doubles, strictly increasing, <= L*M*N) -> 0-based int64 */
static int64_t* voxel_indices(const mxArray* a, int64_t nvox_grid, int64_t* nv) {
  if (!is_real_double(a)) fail("voxhenry:usage", "idxS must be a real double vector");
  const int64_t m = (int64_t)mxGetNumberOfElements(a);
  if (m < 1) fail("voxhenry:usage", "idxS is empty");
  const double* d = mxGetDoubles(a);
  int64_t* idx = (int64_t*)mxMalloc((size_t)m * sizeof(int64_t));
  for (int64_t i = 0; i < m; ++i) {
    const double v = d[i];
    if (!(v >= 1.0 && v <= (double)nvox_grid && v == floor(v)))
      fail("voxhenry:usage", "idxS(%lld) = %g is not a voxel index in 1..%lld", (long long)(i + 1), v, (long long)nvox_grid);
    idx[i] = (int64_t)v - 1;
    if (i > 0 && idx[i] <= idx[i - 1]) fail("voxhenry:usage", "idxS must be strictly increasing (entry %lld)", (long long)(i + 1));
  }
  *nv = m;
  return idx;
}

static void need_geom(void) {
  if (!S.geom) fail("voxhenry:state", "no geometry: call vh_mex('geom', ...) or vh_mex('prepare', ...) first");
}

static void need_operator(void) {
  need_geom();
  if (!S.has_inc) fail("voxhenry:state", "no incidence matrix: call vh_mex('incidence', Ae) first");
  if (!S.has_z) fail("voxhenry:state", "no impedance: call vh_mex('impedance', z_real, z_imag) first");
}

static void need_prepared(void) {
  need_operator();
  if (!S.prepared) fail("voxhenry:state", "preconditioner not prepared: call vh_mex('prepare', ...) first");
}

static void nargs(int nrhs, int lo, int hi, int nlhs, int maxout, const char* cmd) {
  if (nrhs < lo || nrhs > hi) fail("voxhenry:usage", "'%s' takes %d to %d arguments after the command", cmd, lo - 1, hi - 1);
  if (nlhs > maxout) fail("voxhenry:usage", "'%s' returns at most %d outputs", cmd, maxout);
}

/* ---- state changes (all arguments already checked) ------------------------ */

static void make_geom(int64_t L, int64_t M, int64_t N, double dx, int64_t nv, const int64_t* idx) {
  ensure_ctx();
  if (S.pre) { vh_precon_destroy(S.pre); S.pre = NULL; }
  if (S.geom) { vh_geom_destroy(S.geom); S.geom = NULL; }
  free(S.idx);
  S.idx = NULL;
  S.has_inc = S.has_z = S.prepared = S.has_jw = S.z_has_imag = 0;
  /* the geometry is published only once it is complete: a failure leaves no geometry */
  int64_t* keep = (int64_t*)malloc((size_t)nv * sizeof(int64_t));
  if (!keep) fail("voxhenry:lib", "out of host memory");
  memcpy(keep, idx, (size_t)nv * sizeof(int64_t));
  vh_geom* g = NULL;
  int rc = vh_geom_create(S.ctx, L, M, N, dx, nv, idx, 1, &g);
  if (rc == 0) {
    rc = vh_geom_generate_green(g, S.st2);
    if (rc != 0) {
      snprintf(msgbuf, sizeof msgbuf, "%s", vh_last_error());
      vh_geom_destroy(g);
    }
  } else {
    snprintf(msgbuf, sizeof msgbuf, "%s", vh_last_error());
  }
  if (rc != 0) {
    free(keep);
    char m[sizeof msgbuf];
    memcpy(m, msgbuf, sizeof m);
    fail("voxhenry:lib", "libvoxhenry error %d: %s", rc, m);
  }
  S.geom = g;
  S.idx = keep;
  S.dims[0] = L;
  S.dims[1] = M;
  S.dims[2] = N;
  S.dx = dx;
  S.nv = nv;
  S.num_curr = 5 * nv;
}

static void check_incidence(const mxArray* a) {
  if (!mxIsSparse(a) || !mxIsDouble(a) || mxIsComplex(a)) fail("voxhenry:usage", "Ae must be a real sparse double matrix");
  if ((int64_t)mxGetN(a) != S.num_curr)
    fail("voxhenry:usage", "Ae must have 5*numel(idxS) = %lld columns (has %lld)", (long long)S.num_curr, (long long)mxGetN(a));
  if (mxGetM(a) < 1) fail("voxhenry:usage", "Ae has no rows");
}

static void set_incidence(const mxArray* a) {
  const int64_t rows = (int64_t)mxGetM(a), cols = (int64_t)mxGetN(a);
  const mwIndex* jc = mxGetJc(a);
  const mwIndex* ir = mxGetIr(a);
  const double* pr = mxGetDoubles(a);
  const int64_t nnz = (int64_t)jc[cols];
  int64_t* cp = (int64_t*)mxMalloc((size_t)(cols + 1) * sizeof(int64_t));
  int64_t* ri = (int64_t*)mxMalloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int64_t));
  for (int64_t j = 0; j <= cols; ++j) cp[j] = (int64_t)jc[j];
  for (int64_t k = 0; k < nnz; ++k) ri[k] = (int64_t)ir[k];
  if (S.pre) { vh_precon_destroy(S.pre); S.pre = NULL; }
  S.has_inc = S.prepared = 0;
  lib(vh_geom_set_incidence(S.geom, rows, nnz, cp, ri, pr));
  mxFree(cp);
  mxFree(ri);
  S.num_node = rows;
  S.n = S.num_curr + rows;
  S.has_inc = 1;
}

/* z = z_real(idxS) + z_imag(idxS); z_imag may be empty, real or complex */
static void check_impedance(const mxArray* zr, const mxArray* zi, int64_t grid) {
  if (!is_real_double(zr) || (int64_t)mxGetNumberOfElements(zr) != grid)
    fail("voxhenry:usage", "z_real must be a real double array of L*M*N = %lld entries", (long long)grid);
  if (zi && !mxIsEmpty(zi) && (!mxIsDouble(zi) || mxIsSparse(zi) || (int64_t)mxGetNumberOfElements(zi) != grid))
    fail("voxhenry:usage", "z_imag must be empty or a double array of L*M*N = %lld entries", (long long)grid);
}

static void set_impedance(const mxArray* zr, const mxArray* zi) {
  const double* r = mxGetDoubles(zr);
  double* z = (double*)mxMalloc((size_t)(2 * S.nv) * sizeof(double));
  for (int64_t v = 0; v < S.nv; ++v) {
    z[2 * v] = r[S.idx[v]];
    z[2 * v + 1] = 0.0;
  }
  S.z_has_imag = 0;
  if (zi && !mxIsEmpty(zi)) {
    const double* zd = mxIsComplex(zi) ? NULL : mxGetDoubles(zi);
    const mxComplexDouble* zc = mxIsComplex(zi) ? mxGetComplexDoubles(zi) : NULL;
    for (int64_t v = 0; v < S.nv && !S.z_has_imag; ++v)
      S.z_has_imag = zc ? (zc[S.idx[v]].real != 0.0 || zc[S.idx[v]].imag != 0.0) : zd[S.idx[v]] != 0.0;
    if (mxIsComplex(zi)) {
      const mxComplexDouble* c = mxGetComplexDoubles(zi);
      for (int64_t v = 0; v < S.nv; ++v) {
        z[2 * v] += c[S.idx[v]].real;
        z[2 * v + 1] += c[S.idx[v]].imag;
      }
    } else {
      const double* d = mxGetDoubles(zi);
      for (int64_t v = 0; v < S.nv; ++v) z[2 * v] += d[S.idx[v]];
    }
  }
  S.has_z = S.prepared = 0;
  lib(vh_geom_set_impedance(S.geom, z));
  mxFree(z);
  S.has_z = 1;
}

static int precon_type_of(const mxArray* a) {
  char t[64];
  if (!mxIsChar(a) || mxGetString(a, t, sizeof t) != 0) fail("voxhenry:usage", "the preconditioner type must be a string");
  if (!strcmp(t, "schur_invert")) return VH_PC_SCHUR_INVERT;
  if (!strcmp(t, "schur_gmres")) return VH_PC_SCHUR_GMRES;
  fail("voxhenry:usage", "preconditioner type '%s' is not supported (schur_invert, schur_gmres)", t);
  return -1;
}

static void prepare_precon(const double jw[2], int type) {
  if (!S.pre) {
    vh_precon* p = NULL;
    lib(vh_precon_create(S.geom, &p));
    S.pre = p;
  }
  S.prepared = 0;
  lib(vh_precon_set_type(S.pre, type, S.prectol));
  S.pc_type = type;
  lib(vh_precon_prepare(S.pre, jw[0], jw[1], NULL));
  S.jw[0] = jw[0];
  S.jw[1] = jw[1];
  S.has_jw = 1;
  S.prepared = 1;
}

/* ---- commands ------------------------------------------------------------- */

static void cmd_geom(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  nargs(nrhs, 6, 6, nlhs, 1, "geom");
  const int64_t L = int_scalar(prhs[1], "L", 1), M = int_scalar(prhs[2], "M", 1), N = int_scalar(prhs[3], "N", 1);
  const double dx = real_scalar(prhs[4], "dx");
  if (!(dx > 0)) fail("voxhenry:usage", "dx must be positive");
  int64_t nv = 0;
  int64_t* idx = voxel_indices(prhs[5], L * M * N, &nv);
  make_geom(L, M, N, dx, nv, idx);
  mxFree(idx);
  plhs[0] = mxCreateDoubleMatrix(1, 3, mxCOMPLEX);
  memcpy(mxGetComplexDoubles(plhs[0]), S.st2, 6 * sizeof(double));
}

static void cmd_incidence(int nlhs, int nrhs, const mxArray* prhs[]) {
  nargs(nrhs, 2, 2, nlhs, 0, "incidence");
  need_geom();
  check_incidence(prhs[1]);
  set_incidence(prhs[1]);
}

static void cmd_impedance(int nlhs, int nrhs, const mxArray* prhs[]) {
  nargs(nrhs, 2, 3, nlhs, 0, "impedance");
  need_geom();
  const mxArray* zi = nrhs > 2 ? prhs[2] : NULL;
  check_impedance(prhs[1], zi, S.dims[0] * S.dims[1] * S.dims[2]);
  set_impedance(prhs[1], zi);
}

/* vh_mex('prepare', dx, z_real, z_imag, idxS, st_sparse_precon, Ae [, type]):
 * the whole per-frequency set-up of lse_sparse_precon_prepare.m.  The grid is
 * size(z_real); the geometry (and its Green's tensor and incidence) is rebuilt
 * only when the grid, dx or idxS change; jwmu = st_sparse_precon(1) /
 * st_sparse_precon2(1), since the executer scales both by j*omega*mu
 * (VoxHenry_executer.m:266-267). */
static void cmd_prepare(int nlhs, int nrhs, const mxArray* prhs[]) {
  nargs(nrhs, 7, 8, nlhs, 0, "prepare");
  const double dx = real_scalar(prhs[1], "dx");
  if (!(dx > 0)) fail("voxhenry:usage", "dx must be positive");
  const mxArray* zr = prhs[2];
  const mxArray* zi = prhs[3];
  if (!is_real_double(zr)) fail("voxhenry:usage", "z_real must be a real double array");
  const mwSize nd = mxGetNumberOfDimensions(zr);
  const mwSize* d = mxGetDimensions(zr);
  if (nd > 3) fail("voxhenry:usage", "z_real must be an L x M x N array");
  const int64_t L = (int64_t)d[0], M = (int64_t)d[1], N = nd > 2 ? (int64_t)d[2] : 1;
  if (L < 1 || M < 1 || N < 1) fail("voxhenry:usage", "z_real is empty");
  int64_t nv = 0;
  int64_t* idx = voxel_indices(prhs[4], L * M * N, &nv);
  double st[2];
  complex_first(prhs[5], "st_sparse_precon", st);
  const int type = nrhs > 7 ? precon_type_of(prhs[7]) : VH_PC_SCHUR_INVERT;
  const mxArray* ae = prhs[6];
  if (!mxIsSparse(ae) || !mxIsDouble(ae) || mxIsComplex(ae) || (int64_t)mxGetN(ae) != 5 * nv || mxGetM(ae) < 1)
    fail("voxhenry:usage", "Ae must be a real sparse double matrix with 5*numel(idxS) = %lld columns", (long long)(5 * nv));
  check_impedance(zr, zi, L * M * N);
  const int same = S.geom && S.dims[0] == L && S.dims[1] == M && S.dims[2] == N && S.dx == dx && S.nv == nv &&
                   memcmp(S.idx, idx, (size_t)nv * sizeof(int64_t)) == 0;
  if (!same) make_geom(L, M, N, dx, nv, idx);
  mxFree(idx);
  if (!same || !S.has_inc) set_incidence(ae);
  set_impedance(zr, zi);
  const double c = S.st2[0], e = S.st2[1], den = c * c + e * e;
  if (!(den > 0)) fail("voxhenry:lib", "st_sparse_precon2(1) is zero");
  const double jw[2] = {(st[0] * c + st[1] * e) / den, (st[1] * c - st[0] * e) / den};
  prepare_precon(jw, type);
}

static void cmd_matvec(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  nargs(nrhs, 2, 3, nlhs, 1, "matvec");
  need_operator();
  double jw[2];
  if (nrhs > 2) {
    complex_first(prhs[2], "jwmu", jw);
  } else {
    if (!S.has_jw) fail("voxhenry:state", "no frequency: pass jwmu or call vh_mex('prepare', ...) first");
    jw[0] = S.jw[0];
    jw[1] = S.jw[1];
  }
  int64_t k = 0;
  double* tmp = NULL;
  const double* x = columns(prhs[1], S.n, "J", &k, &tmp);
  plhs[0] = mxCreateDoubleMatrix((mwSize)S.n, (mwSize)k, mxCOMPLEX);
  lib(vh_matvec(S.geom, jw[0], jw[1], VH_C128, 0, k, x, S.n, mxGetComplexDoubles(plhs[0]), S.n, 0));
  if (tmp) mxFree(tmp);
}

static void cmd_precon_apply(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  nargs(nrhs, 2, 3, nlhs, 1, "precon_apply");
  need_prepared();
  if (nrhs > 2) {
    const double t = real_scalar(prhs[2], "prectol");
    if (!(t > 0)) fail("voxhenry:usage", "prectol must be positive");
    if (t != S.prectol) {
      lib(vh_precon_set_type(S.pre, S.pc_type, t));
      S.prectol = t;
    }
  }
  int64_t k = 0;
  double* tmp = NULL;
  const double* x = columns(prhs[1], S.n, "v", &k, &tmp);
  plhs[0] = mxCreateDoubleMatrix((mwSize)S.n, (mwSize)k, mxCOMPLEX);
  lib(vh_precon_apply(S.pre, VH_C128, k, x, S.n, mxGetComplexDoubles(plhs[0]), S.n, 0));
  if (tmp) mxFree(tmp);
}

static void cmd_schur_multiply(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  nargs(nrhs, 2, 2, nlhs, 1, "schur_multiply");
  need_prepared();
  int64_t k = 0;
  double* tmp = NULL;
  const double* x = columns(prhs[1], S.num_node, "v", &k, &tmp);
  plhs[0] = mxCreateDoubleMatrix((mwSize)S.num_node, (mwSize)k, mxCOMPLEX);
  lib(vh_precon_schur_multiply(S.pre, k, x, S.num_node, mxGetComplexDoubles(plhs[0]), S.num_node, 0));
  if (tmp) mxFree(tmp);
}

/* [x, flag, relres, iter, resvec] as pgmres.m / fpgmres.m return them:
 * iter = [inner steps of the last cycle, outer cycles] (pgmres.m:178),
 * resvec = relative residual history (column). */
static void cmd_gmres(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[], int flexible) {
  const char* name = flexible ? "fpgmres" : "pgmres";
  nargs(nrhs, 5, 5, nlhs, 5, name);
  const int64_t restart_in = int_scalar(prhs[2], "restart", 1);
  const double tol = real_scalar(prhs[3], "tol");
  const int64_t maxit = int_scalar(prhs[4], "maxit", 1);
  if (!(tol > 0)) fail("voxhenry:usage", "tol must be positive");
  if (maxit > 100000) fail("voxhenry:usage", "maxit too large");
  need_prepared();
  if (!S.has_jw) fail("voxhenry:state", "no frequency: call vh_mex('prepare', ...) first");
  int64_t k = 0;
  double* tmp = NULL;
  const double* b = columns(prhs[1], S.n, "b", &k, &tmp);
  if (k != 1) fail("voxhenry:usage", "b must be one column");
  const int restart = (int)(restart_in < S.n ? restart_in : S.n);
  const int64_t cap = (int64_t)restart * maxit + 1;
  double* rv = (double*)mxMalloc((size_t)cap * sizeof(double));
  int flag = 0, iter[2] = {0, 0}, niters = 0, nres = 0;
  double relres = 0;
  plhs[0] = mxCreateDoubleMatrix((mwSize)S.n, 1, mxCOMPLEX);
  lib(vh_gmres(S.geom, S.pre, S.jw[0], S.jw[1], VH_C128, flexible, 1, b, mxGetComplexDoubles(plhs[0]), 0, restart, tol,
               (int)maxit, &flag, &relres, iter, &niters, rv, cap, &nres));
  if (nlhs > 1) plhs[1] = mxCreateDoubleScalar((double)flag);
  if (nlhs > 2) plhs[2] = mxCreateDoubleScalar(relres);
  if (nlhs > 3) {
    plhs[3] = mxCreateDoubleMatrix(1, 2, mxREAL);
    mxGetDoubles(plhs[3])[0] = iter[0];
    mxGetDoubles(plhs[3])[1] = iter[1];
  }
  if (nlhs > 4) {
    const int64_t m = nres < cap ? nres : cap;
    plhs[4] = mxCreateDoubleMatrix((mwSize)m, 1, mxREAL);
    if (m > 0) memcpy(mxGetDoubles(plhs[4]), rv, (size_t)m * sizeof(double));
  }
  mxFree(rv);
  if (tmp) mxFree(tmp);
}

/* [X, flag, relres, iters, inner_iters] = vh_mex('solve_batch', B, jwmu, pidx, method, tol):
 * the batched tier (vh_solve_batch): column s of B (n x nsys, real or complex)
 * is solved at frequency jwmu(pidx(s)) (jwmu: the j*omega*mu of nf
 * frequencies, pidx 1-based) with that frequency's preconditioner, prepared
 * here for the geometry's current impedance; method 'fp64' (pgmres
 * semantics, restart 100, 10 cycles), 'fp64_true' (the same, then the
 * columns it leaves above tol continue in FP64 FGMRES to the true residual)
 * or 'mixed' (FP64 flexible outer, restart 8, 4 cycles; FP32 inner, restart
 * 30); tol a scalar or one per column.  Outputs are 1 x nsys rows. */
static void cmd_solve_batch(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  nargs(nrhs, 6, 6, nlhs, 5, "solve_batch");
  need_operator();
  /* every frequency's preconditioner and matvec use the geometry's one
   * impedance; a two-fluid impedance (VoxHenry_executer.m:243-254) changes
   * with f, so it would be wrong at every frequency but the one it was set for */
  if (S.z_has_imag)
    fail("voxhenry:usage", "solve_batch uses one impedance for all frequencies; the impedance has a nonzero z_imag "
                           "(two-fluid conductors, frequency dependent): solve per frequency instead");
  if (!mxIsDouble(prhs[2]) || mxIsSparse(prhs[2]) || mxGetNumberOfElements(prhs[2]) < 1)
    fail("voxhenry:usage", "jwmu must be a double array of one value per frequency");
  const int64_t nf = (int64_t)mxGetNumberOfElements(prhs[2]);
  if (!is_real_double(prhs[3])) fail("voxhenry:usage", "pidx must be a real double vector");
  char mth[16];
  if (!mxIsChar(prhs[4]) || mxGetString(prhs[4], mth, sizeof mth) != 0 ||
      (strcmp(mth, "fp64") && strcmp(mth, "mixed") && strcmp(mth, "fp64_true")))
    fail("voxhenry:usage", "method must be 'fp64', 'fp64_true' or 'mixed'");
  const int method = strcmp(mth, "mixed") == 0 ? 1 : strcmp(mth, "fp64_true") == 0 ? 2 : 0;
  int64_t ns = 0;
  double* tmp = NULL;
  const double* b = columns(prhs[1], S.n, "B", &ns, &tmp);
  if ((int64_t)mxGetNumberOfElements(prhs[3]) != ns) fail("voxhenry:usage", "pidx must have one entry per column of B");
  if (!is_real_double(prhs[5]) || (mxGetNumberOfElements(prhs[5]) != 1 && (int64_t)mxGetNumberOfElements(prhs[5]) != ns))
    fail("voxhenry:usage", "tol must be a scalar or one value per column of B");
  const double* tv = mxGetDoubles(prhs[5]);
  const int per_col = mxGetNumberOfElements(prhs[5]) != 1;
  for (int64_t s = 0; s < (per_col ? ns : 1); ++s)
    if (!(tv[s] > 0)) fail("voxhenry:usage", "tol must be positive");
  double* jw = (double*)mxMalloc((size_t)(2 * nf) * sizeof(double));
  for (int64_t f = 0; f < nf; ++f) {
    if (mxIsComplex(prhs[2])) {
      jw[2 * f] = mxGetComplexDoubles(prhs[2])[f].real;
      jw[2 * f + 1] = mxGetComplexDoubles(prhs[2])[f].imag;
    } else {
      jw[2 * f] = mxGetDoubles(prhs[2])[f];
      jw[2 * f + 1] = 0.0;
    }
  }
  int* pidx = (int*)mxMalloc((size_t)ns * sizeof(int));
  double* jws = (double*)mxMalloc((size_t)(2 * ns) * sizeof(double));
  const double* pd = mxGetDoubles(prhs[3]);
  for (int64_t s = 0; s < ns; ++s) {
    if (!(pd[s] >= 1 && pd[s] <= (double)nf && pd[s] == floor(pd[s])))
      fail("voxhenry:usage", "pidx(%lld) = %g is not a frequency index in 1..%lld", (long long)(s + 1), pd[s], (long long)nf);
    pidx[s] = (int)pd[s] - 1;
    jws[2 * s] = jw[2 * pidx[s]];
    jws[2 * s + 1] = jw[2 * pidx[s] + 1];
  }
  double* tols = NULL;
  if (per_col) tols = (double*)mxMalloc((size_t)ns * sizeof(double));
  if (per_col) memcpy(tols, tv, (size_t)ns * sizeof(double));
  double tmax = tv[0];
  for (int64_t s = 1; s < (per_col ? ns : 1); ++s) tmax = tv[s] > tmax ? tv[s] : tmax;
  /* outputs first, then the preconditioners: from here on every failure
   * releases them before raising the error */
  plhs[0] = mxCreateDoubleMatrix((mwSize)S.n, (mwSize)ns, mxCOMPLEX);
  int* flag = (int*)mxMalloc((size_t)ns * sizeof(int));
  int* its = (int*)mxMalloc((size_t)ns * sizeof(int));
  int* iits = (int*)mxMalloc((size_t)ns * sizeof(int));
  double* rr = (double*)mxMalloc((size_t)ns * sizeof(double));
  vh_precon** ps = (vh_precon**)mxCalloc((size_t)nf, sizeof(vh_precon*));
  int rc = 0;
  for (int64_t f = 0; f < nf && rc == 0; ++f) rc = vh_precon_create(S.geom, &ps[f]);
  if (rc == 0) rc = vh_precon_prepare_many(ps, nf, jw, NULL, 0);
  if (rc == 0)
    rc = vh_solve_batch(S.geom, nf, ps, ns, jws, pidx, b, (double*)mxGetComplexDoubles(plhs[0]), 0, 0, method, tmax,
                        tols, method == 1 ? 8 : 100, method == 1 ? 4 : 10, 30, 1e-5, 0.0, flag, rr, its, iits, NULL);
  if (rc != 0) snprintf(msgbuf, sizeof msgbuf, "%s", vh_last_error());
  for (int64_t f = 0; f < nf; ++f)
    if (ps[f]) vh_precon_destroy(ps[f]);
  if (rc != 0) {
    char m[sizeof msgbuf];
    memcpy(m, msgbuf, sizeof m);
    fail("voxhenry:lib", "libvoxhenry error %d: %s", rc, m);
  }
  mxArray* outs[4] = {NULL, NULL, NULL, NULL};
  for (int q = 0; q < 4 && q + 1 < (nlhs > 1 ? nlhs : 1); ++q) {
    outs[q] = mxCreateDoubleMatrix(1, (mwSize)ns, mxREAL);
    double* o = mxGetDoubles(outs[q]);
    for (int64_t s = 0; s < ns; ++s) o[s] = q == 0 ? flag[s] : q == 1 ? rr[s] : q == 2 ? its[s] : iits[s];
    plhs[q + 1] = outs[q];
  }
  mxFree(ps);
  mxFree(rr);
  mxFree(iits);
  mxFree(its);
  mxFree(flag);
  if (tols) mxFree(tols);
  mxFree(jws);
  mxFree(pidx);
  mxFree(jw);
  if (tmp) mxFree(tmp);
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  char cmd[32];
  if (nrhs < 1 || !mxIsChar(prhs[0]) || mxGetString(prhs[0], cmd, sizeof cmd) != 0)
    fail("voxhenry:usage", "first argument must be a command string ('prepare', 'matvec', 'precon_apply', ...)");
  if (!strcmp(cmd, "geom")) cmd_geom(nlhs, plhs, nrhs, prhs);
  else if (!strcmp(cmd, "incidence")) cmd_incidence(nlhs, nrhs, prhs);
  else if (!strcmp(cmd, "impedance")) cmd_impedance(nlhs, nrhs, prhs);
  else if (!strcmp(cmd, "prepare")) cmd_prepare(nlhs, nrhs, prhs);
  else if (!strcmp(cmd, "matvec")) cmd_matvec(nlhs, plhs, nrhs, prhs);
  else if (!strcmp(cmd, "precon_apply")) cmd_precon_apply(nlhs, plhs, nrhs, prhs);
  else if (!strcmp(cmd, "schur_multiply")) cmd_schur_multiply(nlhs, plhs, nrhs, prhs);
  else if (!strcmp(cmd, "pgmres")) cmd_gmres(nlhs, plhs, nrhs, prhs, 0);
  else if (!strcmp(cmd, "fpgmres")) cmd_gmres(nlhs, plhs, nrhs, prhs, 1);
  else if (!strcmp(cmd, "solve_batch")) cmd_solve_batch(nlhs, plhs, nrhs, prhs);
  else if (!strcmp(cmd, "reset")) {
    nargs(nrhs, 1, 1, nlhs, 0, "reset");
    const int locked = S.ctx != NULL;
    release();
    if (locked) mexUnlock();
  } else {
    fail("voxhenry:usage", "unknown command '%s'", cmd);
  }
}
