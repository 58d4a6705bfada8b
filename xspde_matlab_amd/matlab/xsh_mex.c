/*
 * xsh_mex.c - MEX gateway of the MI355X ensemble integration path
 * (include/xspde_hip.h) for the MATLAB host of xSPDE.
 *
 * Build (MATLAB R2018a+, interleaved complex):
 *   mex -R2018a -I<repo>/include xsh_mex.c -L<repo>/xspde_matlab_amd -lxspde_hip
 *
 * Commands (the handle is a uint64 scalar):
 *   h    = xsh_mex('create', d)             d: struct from xgpudesc(p)
 *   info = xsh_mex('query', h)
 *   [av, a] = xsh_mex('run_path', h, npar, j, nc, a0)
 *          the p.path contract (Simulator/xpath.m:1): av is the [avlen, 1]
 *          complex block of p.d.av{o} arrays (time fastest, xpreferences.m:340);
 *          a is the final field [1, nx, ny, ensembles(1)]; a0 may be [] for
 *          the built-in initial form
 *   xsh_mex('run_ensemble', h, npar)        every (j, nc) path of ensemble npar
 *   av   = xsh_mex('read_averages', h)      [avlen, ensembles(2)*nchk] complex,
 *                                           column (j-1)*nchk + nc
 *   flag = xsh_mex('nonfinite', h)          1 if a stored field value of the last run
 *                                           was NaN or Inf (xsh_nonfinite)
 *   info = xsh_mex('plan', d)               what create would build, with no device
 *                                           (xsh_plan; xSPDE:gpu:unsupported: CPU path)
 *   xsh_mex('destroy', h)
 *
 * Errors map the library's status to an identifier: xSPDE:gpu:unsupported
 * (XSH_EUNSUPPORTED: the caller stays on the CPU xensemble), xSPDE:gpu:invalid,
 * xSPDE:gpu:device, with the library's message (the analogue of the
 * reference's error() aborts, Applications/SPDE/xtransform.m:51).
 *
 * MATLAB and mex.h are not part of this repository's build image, so this file
 * is not compiled or run by its tests; it marshals MATLAB arrays to the C ABI
 * that the tests exercise through ctypes.
 */
#include <stdint.h>
#include <string.h>

#include "mex.h"
#include "xspde_hip.h"

static void fail_status(int rc, const char* what) {
  const char* id = rc == XSH_EUNSUPPORTED ? "xSPDE:gpu:unsupported"
                   : rc == XSH_EINVAL     ? "xSPDE:gpu:invalid"
                   : rc == XSH_ENOMEM     ? "xSPDE:gpu:memory"
                                          : "xSPDE:gpu:device";
  mexErrMsgIdAndTxt(id, "%s: %s", what, xsh_last_error());
}

static const mxArray* field(const mxArray* s, const char* name) {
  return mxIsStruct(s) ? mxGetField(s, 0, name) : NULL;
}

static double num(const mxArray* s, const char* name, double dflt) {
  const mxArray* v = field(s, name);
  return v && !mxIsEmpty(v) ? mxGetScalar(v) : dflt;
}

/* copy up to n doubles of field name (real part) into out */
static void vec(const mxArray* s, const char* name, double* out, int n) {
  const mxArray* v = field(s, name);
  if (!v || mxIsEmpty(v)) return;
  const mwSize m = mxGetNumberOfElements(v);
  if (mxIsComplex(v)) {
    const mxComplexDouble* z = mxGetComplexDoubles(v);
    for (int i = 0; i < n && (mwSize)i < m; ++i) out[i] = z[i].real;
  } else {
    const double* r = mxGetDoubles(v);
    for (int i = 0; i < n && (mwSize)i < m; ++i) out[i] = r[i];
  }
}

/* complex field name as interleaved (re, im) pairs, n complex values */
static void cvec(const mxArray* s, const char* name, double* out, int n) {
  const mxArray* v = field(s, name);
  if (!v || mxIsEmpty(v)) return;
  const mwSize m = mxGetNumberOfElements(v);
  for (int i = 0; i < n && (mwSize)i < m; ++i) {
    if (mxIsComplex(v)) {
      out[2 * i] = mxGetComplexDoubles(v)[i].real;
      out[2 * i + 1] = mxGetComplexDoubles(v)[i].imag;
    } else {
      out[2 * i] = mxGetDoubles(v)[i];
      out[2 * i + 1] = 0.0;
    }
  }
}

static const double* ctable(const mxArray* s, const char* name, mwSize* count) {
  const mxArray* v = field(s, name);
  *count = 0;
  if (!v || mxIsEmpty(v)) return NULL;
  if (!mxIsComplex(v)) mexErrMsgIdAndTxt("xSPDE:gpu:invalid", "%s must be a complex array", name);
  *count = mxGetNumberOfElements(v);
  return (const double*)mxGetComplexDoubles(v);
}

static void make_desc(const mxArray* s, xsh_desc* d) {
  double t[64];
  memset(d, 0, sizeof *d);
  d->struct_size = (int)sizeof *d;
  d->precision = (int)num(s, "precision", XSH_FP32);
  d->dimensions = (int)num(s, "dimensions", 1);
  d->points = (int)num(s, "points", 51);
  d->nx = (int)num(s, "nx", 1);
  d->ny = (int)num(s, "ny", 1);
  d->steps = (int)num(s, "steps", 1);
  d->checks = (int)num(s, "checks", 1);
  d->method = (int)num(s, "method", XSH_MP);
  d->iterations = (int)num(s, "iterations", 4);
  t[0] = t[1] = t[2] = 1.0;
  vec(s, "ensembles", t, 3);
  for (int i = 0; i < 3; ++i) d->ensembles[i] = (int)t[i];
  d->seed = (int)num(s, "seed", 0);
  d->noises = (int)num(s, "noises", 0);
  d->inrandoms = (int)num(s, "inrandoms", 0);
  d->urandoms = (int)num(s, "urandoms", 0);
  memset(t, 0, sizeof t);
  vec(s, "boundaries", t, 4); /* [2 x 2] column-major: (dim, side) */
  for (int k = 0; k < 2; ++k)
    for (int side = 0; side < 2; ++side) d->boundaries[k][side] = (int)t[k + 2 * side];
  vec(s, "ranges", d->ranges, 3);
  vec(s, "origins", d->origins, 3);
  d->deriv = (int)num(s, "deriv", XSH_DERIV_ZERO);
  vec(s, "deriv_par", d->deriv_par, 6);
  cvec(s, "lin_l0", d->lin_l0, 1);
  cvec(s, "lin_l1", &d->lin_l1[0][0], 2);
  cvec(s, "lin_l2", &d->lin_l2[0][0], 2);
  d->has_linear = (int)num(s, "has_linear", 0);
  d->initial = (int)num(s, "initial", XSH_INIT_CONST);
  vec(s, "init_par", d->init_par, 6);
  memset(t, 0, sizeof t);
  vec(s, "observe", t, XSH_MAX_OBS);
  const mxArray* ob = field(s, "observe");
  d->nobserve = ob ? (int)mxGetNumberOfElements(ob) : 0;
  for (int o = 0; o < d->nobserve && o < XSH_MAX_OBS; ++o) d->observe[o] = (int)t[o];
  d->seq = (int)num(s, "seq", 0);
  d->inst = (int)num(s, "inst", 0);
  d->traj_out = (int)num(s, "traj_out", 0);
  memset(t, 0, sizeof t);
  vec(s, "modes", t, 2 * XSH_MAX_MODES); /* [2 x nmodes]: (kx, ky) per column */
  const mxArray* md = field(s, "modes");
  d->nmodes = md ? (int)(mxGetNumberOfElements(md) / 2) : 0;
  for (int m = 0; m < d->nmodes && m < XSH_MAX_MODES; ++m) {
    d->modes[m][0] = (int)t[2 * m];
    d->modes[m][1] = (int)t[2 * m + 1];
  }
  double bv[8] = {0};
  cvec(s, "boundval", bv, 4); /* [2 x 2] complex column-major: (dim, side) */
  for (int k = 0; k < 2; ++k)
    for (int side = 0; side < 2; ++side) {
      d->boundval[k][side][0] = bv[2 * (k + 2 * side)];
      d->boundval[k][side][1] = bv[2 * (k + 2 * side) + 1];
    }
  d->raw = (int)num(s, "raw", 0);
  mwSize n = 0;
  /* boundtab: [nb, 2, 2, levels] complex, column-major = [level][dim][side][b] */
  d->boundtab = ctable(s, "boundtab", &n);
  if (d->boundtab) {
    const mwSize nb = d->dimensions == 3 ? (mwSize)(d->nx > d->ny ? d->nx : d->ny) : 1;
    d->boundlevels = (int)(n / (4 * nb));
  }
  /* nfilter, rfilter: [nx*ny, count] complex (k sites in FFT order) */
  d->nfilter = ctable(s, "nfilter", &n);
  d->knoises = d->nfilter ? (int)(n / ((mwSize)d->nx * d->ny)) : 0;
  d->rfilter = ctable(s, "rfilter", &n);
  d->krandoms = d->rfilter ? (int)(n / ((mwSize)d->nx * d->ny)) : 0;
  /* tspec: per observable, 1 for the time-domain spectrum (p.transforms{o}(1)) */
  memset(t, 0, sizeof t);
  vec(s, "tspec", t, XSH_MAX_OBS);
  for (int o = 0; o < XSH_MAX_OBS; ++o) d->tspec[o] = (int)t[o];
  /* coupled SDE components (xgpudesc.m, matrix form): fields; fmat [F x F],
     fnl, fl0, fc0 [F] complex; fnoise [F x noises] real; obs_field per observable */
  d->fields = (int)num(s, "fields", 0);
  if (d->fields > 1 && d->fields <= XSH_MAX_FIELDS) {
    const int F = d->fields;
    double m[2 * XSH_MAX_FIELDS * XSH_MAX_FIELDS] = {0};
    cvec(s, "fmat", m, F * F); /* column-major: (i, j) at i + F j */
    for (int i = 0; i < F; ++i)
      for (int j = 0; j < F; ++j) {
        d->fmat[i][j][0] = m[2 * (i + F * j)];
        d->fmat[i][j][1] = m[2 * (i + F * j) + 1];
      }
    cvec(s, "fnl", &d->fnl[0][0], F);
    cvec(s, "fl0", &d->fl0[0][0], F);
    cvec(s, "fc0", &d->fc0[0][0], F);
    double sn[4 * XSH_MAX_FIELDS] = {0};
    vec(s, "fnoise", sn, 4 * F);
    const mxArray* fn = field(s, "fnoise");
    const int nw = fn ? (int)(mxGetNumberOfElements(fn) / F) : 0;
    for (int i = 0; i < F; ++i)
      for (int k = 0; k < nw && k < 4; ++k) d->fnoise[i][k] = sn[i + F * k];
    memset(t, 0, sizeof t);
    vec(s, "obs_field", t, XSH_MAX_OBS);
    for (int o = 0; o < XSH_MAX_OBS; ++o) d->obs_field[o] = (int)t[o];
  }
}

static mxArray* info_struct(const xsh_info* info) {
  const char* names[] = {"nsites", "nchk", "avlen", "obs_offset", "obs_sites", "traj_steps", "kernel_kind",
                         "ensembles", "points"};
  mxArray* r = mxCreateStructMatrix(1, 1, 9, names);
  mxSetField(r, 0, "nsites", mxCreateDoubleScalar(info->nsites));
  mxSetField(r, 0, "nchk", mxCreateDoubleScalar(info->nchk));
  mxSetField(r, 0, "avlen", mxCreateDoubleScalar((double)info->avlen));
  mxArray* off = mxCreateDoubleMatrix(1, XSH_MAX_OBS, mxREAL);
  mxArray* ns = mxCreateDoubleMatrix(1, XSH_MAX_OBS, mxREAL);
  for (int o = 0; o < XSH_MAX_OBS; ++o) {
    mxGetDoubles(off)[o] = (double)info->obs_offset[o];
    mxGetDoubles(ns)[o] = info->obs_sites[o];
  }
  mxSetField(r, 0, "obs_offset", off);
  mxSetField(r, 0, "obs_sites", ns);
  mxSetField(r, 0, "traj_steps", mxCreateDoubleScalar((double)info->traj_steps));
  mxSetField(r, 0, "kernel_kind", mxCreateDoubleScalar(info->kernel_kind));
  mxArray* ens = mxCreateDoubleMatrix(1, 3, mxREAL);
  for (int i = 0; i < 3; ++i) mxGetDoubles(ens)[i] = info->ensembles[i];
  mxSetField(r, 0, "ensembles", ens);
  mxSetField(r, 0, "points", mxCreateDoubleScalar(info->points));
  return r;
}

static xsh_handle* handle_arg(const mxArray* a) {
  if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1)
    mexErrMsgIdAndTxt("xSPDE:gpu:invalid", "handle must be a uint64 scalar from xsh_mex('create', d)");
  return (xsh_handle*)(uintptr_t)(*(const uint64_t*)mxGetData(a));
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  char cmd[32];
  if (nrhs < 1 || mxGetString(prhs[0], cmd, sizeof cmd))
    mexErrMsgIdAndTxt("xSPDE:gpu:invalid", "first argument: command string");
  if (!strcmp(cmd, "create")) {
    if (nrhs < 2) mexErrMsgIdAndTxt("xSPDE:gpu:invalid", "create needs the descriptor struct");
    xsh_desc d;
    make_desc(prhs[1], &d);
    xsh_handle* h = NULL;
    const int rc = xsh_create(&d, &h);
    if (rc) fail_status(rc, "xsh_create");
    plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
    *(uint64_t*)mxGetData(plhs[0]) = (uint64_t)(uintptr_t)h;
    return;
  }
  if (!strcmp(cmd, "plan")) {
    if (nrhs < 2) mexErrMsgIdAndTxt("xSPDE:gpu:invalid", "plan needs the descriptor struct");
    xsh_desc d;
    make_desc(prhs[1], &d);
    xsh_info pinfo;
    const int rc = xsh_plan(&d, &pinfo);
    if (rc) fail_status(rc, "xsh_plan");
    plhs[0] = info_struct(&pinfo);
    return;
  }
  if (nrhs < 2) mexErrMsgIdAndTxt("xSPDE:gpu:invalid", "%s needs a handle", cmd);
  xsh_handle* h = handle_arg(prhs[1]);
  xsh_info info;
  if (!strcmp(cmd, "destroy")) {
    xsh_destroy(h);
    return;
  }
  int rc = xsh_query(h, &info);
  if (rc) fail_status(rc, "xsh_query");
  if (!strcmp(cmd, "query")) {
    plhs[0] = info_struct(&info);
    return;
  }
  if (!strcmp(cmd, "nonfinite")) {
    int flag = 0;
    rc = xsh_nonfinite(h, &flag);
    if (rc) fail_status(rc, "xsh_nonfinite");
    plhs[0] = mxCreateDoubleScalar(flag);
    return;
  }
  if (!strcmp(cmd, "run_path")) {
    if (nrhs < 5) mexErrMsgIdAndTxt("xSPDE:gpu:invalid", "run_path(h, npar, j, nc[, a0])");
    const int npar = (int)mxGetScalar(prhs[2]), j = (int)mxGetScalar(prhs[3]), nc = (int)mxGetScalar(prhs[4]);
    const mwSize ncell = (mwSize)info.nsites * (mwSize)info.ensembles[0];
    const double* a0 = NULL;
    if (nrhs > 5 && !mxIsEmpty(prhs[5])) {
      if (!mxIsComplex(prhs[5]) || mxGetNumberOfElements(prhs[5]) != ncell)
        mexErrMsgIdAndTxt("xSPDE:gpu:invalid", "a0 must be complex with nx*ny*ensembles(1) elements");
      a0 = (const double*)mxGetComplexDoubles(prhs[5]);
    }
    plhs[0] = mxCreateDoubleMatrix((mwSize)info.avlen, 1, mxCOMPLEX);
    double* aout = NULL;
    if (nlhs > 1) {
      /* [nx*ny, ensembles(1)]: the caller reshapes to p.d.ca = [1, nx, ny, E] */
      plhs[1] = mxCreateDoubleMatrix((mwSize)info.nsites, (mwSize)info.ensembles[0], mxCOMPLEX);
      aout = (double*)mxGetComplexDoubles(plhs[1]);
    }
    rc = xsh_run_path(h, npar, j, nc, a0, aout, (double*)mxGetComplexDoubles(plhs[0]));
    if (rc) fail_status(rc, "xsh_run_path");
    return;
  }
  if (!strcmp(cmd, "run_ensemble")) {
    if (nrhs < 3) mexErrMsgIdAndTxt("xSPDE:gpu:invalid", "run_ensemble(h, npar)");
    rc = xsh_run_ensemble(h, (int)mxGetScalar(prhs[2]));
    if (!rc) rc = xsh_synchronize(h, NULL);
    if (rc) fail_status(rc, "xsh_run_ensemble");
    return;
  }
  if (!strcmp(cmd, "read_averages")) {
    /* one column per path, column (j-1)*nchk + nc */
    plhs[0] = mxCreateDoubleMatrix((mwSize)info.avlen, (mwSize)info.ensembles[1] * info.nchk, mxCOMPLEX);
    rc = xsh_read_averages(h, (double*)mxGetComplexDoubles(plhs[0]));
    if (rc) fail_status(rc, "xsh_read_averages");
    return;
  }
  if (!strcmp(cmd, "read_raw")) {
    /* one column per path as read_averages; each column reshapes to
       p.d.raw{1} = [1, points, nx, ny, ensembles(1)] */
    const mwSize len = (mwSize)info.points * info.nsites * (mwSize)info.ensembles[0];
    plhs[0] = mxCreateDoubleMatrix(len, (mwSize)info.ensembles[1] * info.nchk, mxCOMPLEX);
    rc = xsh_read_raw(h, (double*)mxGetComplexDoubles(plhs[0]));
    if (rc) fail_status(rc, "xsh_read_raw");
    return;
  }
  mexErrMsgIdAndTxt("xSPDE:gpu:invalid", "unknown command '%s'", cmd);
}
