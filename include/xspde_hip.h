/*
 * xspde_hip.h - C ABI of the MI355X (gfx950) ensemble integration path of
 * xSPDE (peterddrummond/xspde_matlab, xSPDE 4.21).
 *
 * The library replaces, for a library of named built-in drift/linear/
 * initial/observe forms, the reference's path plug-in slot
 *
 *     [a,av,raw] = p.path(a,nc,p)          Simulator/xpath.m:1,
 *                                          called at Simulator/xensemble.m:79
 *
 * together with the pieces that slot calls on the hot loop:
 *     p.method  RK4 / MP                   Applications/SDE/RK4.m:1, MP.m:1
 *     p.prop    xprop -> xtransform        Applications/SPDE/xprop.m:1, xtransform.m:1
 *     p.noisegen xnoise, p.randomgen       Simulator/xnoise.m:1, xrandom.m:1
 *     xdata (ensemble means)               Simulator/xdata.m:1
 * and offers a batched entry that runs all ensembles(2) x checks paths of
 * one parallel ensemble (xensemble.m:25-80) in one launch.
 *
 * Conventions: plain pointers and sizes, no exceptions across the ABI; every
 * entry returns an xsh_status and leaves a message for xsh_last_error()
 * (thread-local), the analogue of the reference's error() aborts
 * (e.g. xtransform.m:51, spdpreferences.m:67-68).  Host buffers are
 * MATLAB column-major, interleaved complex double (mxGetComplexDoubles):
 *   fields    a0, a_out   [1, nx, ny, ensembles(1)]   (p.d.ca, xpreferences.m:145)
 *   averages  per observable o, at offset obs_offset[o] of a path's [avlen]
 *             block: [1, points, sites_o] with sites_o = nx*ny (x before y) for
 *             per-site forms, 1 for lattice means (p.d.av{o}, xpreferences.m:340),
 *             i.e. element (n, s) at obs_offset[o] + n + points*s: time fastest
 *   raw       per path [1, points, nx, ny, ensembles(1)]  (p.d.raw, xpreferences.m:148)
 */
#ifndef XSPDE_HIP_H
#define XSPDE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define XSH_ABI_VERSION 1

typedef enum {
  XSH_OK = 0,
  XSH_EINVAL = 1,       /* invalid argument or descriptor field        */
  XSH_EUNSUPPORTED = 2, /* valid xSPDE input, not a built-in GPU form  */
  XSH_ENOMEM = 3,       /* device allocation refused                   */
  XSH_EDEVICE = 4,      /* HIP runtime error / no device               */
  XSH_ESTATE = 5        /* call out of order (e.g. nothing run yet)    */
} xsh_status;

enum { XSH_FP32 = 0, XSH_FP64 = 1 };
enum { XSH_RK4 = 0, XSH_MP = 1 };
enum { XSH_DERIV_ZERO = 0, XSH_DERIV_KUBO = 1, XSH_DERIV_GP = 2, XSH_DERIV_MATRIX = 3 };
enum {
  XSH_INIT_CONST = 0,
  XSH_INIT_SECH = 1,
  XSH_INIT_GAUSS = 2,
  XSH_INIT_PHASE = 3,
  XSH_INIT_COHERENT = 4
};
enum {
  XSH_OBS_A = 0,        /* a(x)             per site   */
  XSH_OBS_ABS2 = 1,     /* |a(x)|^2         per site   */
  XSH_OBS_A_AVE = 2,    /* lattice mean a              */
  XSH_OBS_ABS2_AVE = 3, /* lattice mean |a|^2          */
  XSH_OBS_ABS2_INT = 4, /* sum |a|^2 dV                */
  XSH_OBS_ABS4_AVE = 5, /* lattice mean |a|^4          */
  XSH_OBS_CORR = 6,     /* (1/Ns) sum_x a*(x) a(x+r), per r (periodic) */
  XSH_OBS_LOWK = 7,     /* sum over the nonzero noise modes k of (|a_k|^2 - 1/2),
                           a_k = sqrt(dv/Ns) sum_x a(x) e^{-ik.x}: occupation of the
                           lowest lattice modes (periodic, nmodes > 0) */
  XSH_OBS_ABS2K = 8     /* |a_k|^2 per k-space site, a_k = fft(a) prod_d dx_d/sqrt(2 pi)
                           (the observable |a|^2 with p.transforms set, xnft.m), FFT order */
};
#define XSH_MAX_OBS 8
#define XSH_MAX_MODES 32
#define XSH_MAX_FIELDS 4

/* Problem descriptor: the subset of the reference's struct p that the path
 * reads (xpreferences.m:42-219, spdpreferences.m:10-72).  Lattice sizes are
 * the refined sizes points{1}(2:nd) = 1+(points-1)*steps(2:nd)
 * (xpreferences.m:137). */
typedef struct {
  int struct_size;      /* sizeof(xsh_desc)                                  */
  int precision;        /* XSH_FP32 | XSH_FP64 (field arithmetic)            */
  int dimensions;       /* 1 = SDE, 2 = t+x, 3 = t+x+y  (p.dimensions)       */
  int points;           /* p.points{1}(1): stored time points                */
  int nx, ny;           /* lattice points in x, y (1 if absent)              */
  int steps;            /* p.steps(1)                                        */
  int checks;           /* p.checks(1): 0 or 1                               */
  int method;           /* XSH_RK4 | XSH_MP  (p.method)                      */
  int iterations;       /* p.iterations (MP)                                 */
  int ensembles[3];     /* p.ensembles                                       */
  int seed;             /* p.seed                                            */
  int noises;           /* real noise components per site (p.noises)        */
  int inrandoms;        /* Gaussian initial randoms per site (p.inrandoms)   */
  int urandoms;         /* uniform initial randoms per site (p.urandoms)     */
  int boundaries[2][2]; /* [dim][lower,upper]: 0 periodic, -1 Neumann, 1 Dirichlet */
  double ranges[3];     /* p.ranges                                          */
  double origins[3];    /* p.origins                                         */
  int deriv;            /* XSH_DERIV_*                                       */
  double deriv_par[6];  /* KUBO: c.  GP: g.re g.im u.re u.im s               */
  double lin_l0[2];     /* L = l0 + sum_d l1[d] D_d + l2[d] D_d^2 (complex)  */
  double lin_l1[2][2];
  double lin_l2[2][2];
  int has_linear;       /* 0: propagator is 1 (xpropfactor.m:27,37)        */
  int initial;          /* XSH_INIT_*                                        */
  double init_par[6];   /* CONST: c.re c.im; SECH: amp width eps center
                           (amp sech((x - center)/width)); GAUSS: amp width
                           width_y (amp exp(-x^2/2w^2 - y^2/2w_y^2), width_y 0:
                           = width); PHASE: amp; COHERENT: alpha.re alpha.im  */
  int nobserve;
  int observe[XSH_MAX_OBS];
  int seq;              /* sequence index (random stream key)                */
  int inst;             /* instance id (random stream counter)               */
  int traj_out;         /* 1: keep lattice-mean observables per trajectory    */
  int nmodes;           /* Fourier modes of the initial noise kept per trajectory */
  int modes[XSH_MAX_MODES][2]; /* (kx, ky) FFT indices of those modes          */
  double boundval[2][2][2]; /* [dim][lower,upper][re,im]: constant boundary values
                               (p.boundval, spdpreferences.m:26-43): the field at
                               Dirichlet ends, its derivative at Neumann ends     */
  int raw;              /* 1: keep every trajectory's field at every stored point
                           (p.rawdata, xpath.m:73-78), read with xsh_read_raw  */
  /* Boundary values that change in time or along the boundary (p.boundfun,
     called by xpatch.m:37 and xsetbound.m:39 at the method's p.t), tabulated by
     the caller: boundtab[level][dim][lower,upper][b][re,im] with
     b < (dimensions == 3 ? max(nx, ny) : 1) running along the other space
     dimension (y for the x ends, x for the y ends).  Level m is the time
     origins[0] + m * dtr / 2 with dtr the fine step ranges[0] / ((points - 1) *
     steps * (1 + checks)): boundlevels is 1 (fixed values) or
     2 * (points * steps - 1) * (1 + checks) + 1 (one level per half fine step of
     the path; coarse steps use the even levels).  NULL: the boundval constants. */
  const double* boundtab;
  int boundlevels;
  /* k-space filtered noises and randoms (p.knoises, p.krandoms; xnoise.m:39-50,
     xrandom.m:24-29): knoises 0..2, krandoms 0 or 2, with the filters p.nfilter,
     p.rfilter as complex tables nfilter[knoises][nx*ny], rfilter[krandoms][nx*ny]
     of the multiplier at every k site in FFT order (site kx + nx*ky).  The
     drift's noise term (w1 + i w2, or w1) and the initial forms' randoms
     (v1 + i v2) then come from these cells (noises, inrandoms 0).  Periodic
     lattices on the register-resident kernels; not in batches. */
  int knoises, krandoms;
  const double* nfilter;
  const double* rfilter;
  /* Time-domain spectra (p.transforms{o}(1), xpath.m:45-68,90-106): tspec[o] = 1
     takes observable o of the field's time transform instead of the field:
     a(w_m) = fsc/M sum_n a_n exp(+2 pi i j_m n / M) over the M = points * steps
     step fields a_n of each trajectory (a_0 the initial field; a fine run with
     checks stores the mean of each step's two half-step results), at the
     points frequencies j_m = m - floor(points/2), w_m = j_m dw,
     dw = 2 pi / (ranges[0] + dt), fsc = sqrt(2 pi) / dw (xpreferences.m:199-203).
     Its averages fill the observable's [points] block with frequency in place of
     time.  SDEs (dimensions = 1); not in batches or xsh_run_path_noalloc. */
  int tspec[XSH_MAX_OBS];
  /* Coupled field components of an SDE (p.fields > 1 with dimensions = 1,
     xpreferences.m:44-50; e.g. Examples/SDE_1D/SHO.m): fields 2..XSH_MAX_FIELDS
     (0 or 1: one component and the forms above).  The drift is then
     deriv = XSH_DERIV_MATRIX,
       da_i/dt = sum_j M[i][j] a_j + g[i] |a_i|^2 a_i + sum_k S[i][k] w_k
     (fmat = M and fnl = g complex, fnoise = S real over the noises <= 4 real
     noise components w_k of xnoise.m:30-32); the interaction picture
     propagator exp(fl0[i] dt / ipsteps) per component (p.linear returning one
     row per field, xpropfactor.m:26-39); the initial state the constants fc0[i]
     (initial = XSH_INIT_CONST) or a0 as [ensembles(1)][fields] (components
     fastest, p.d.ca); observable o (XSH_OBS_A / _A_AVE: a_i, _ABS2 / _ABS2_AVE:
     |a_i|^2, _ABS4_AVE: |a_i|^4) reads component obs_field[o] (0-based).
     xsh_info.nsites is then the number of components (the size of a0, a_out
     per trajectory).  Not with tspec, in batches or xsh_run_path_noalloc. */
  int fields;
  double fmat[XSH_MAX_FIELDS][XSH_MAX_FIELDS][2];
  double fnl[XSH_MAX_FIELDS][2];
  double fnoise[XSH_MAX_FIELDS][4];
  double fl0[XSH_MAX_FIELDS][2];
  double fc0[XSH_MAX_FIELDS][2];
  int obs_field[XSH_MAX_OBS];
} xsh_desc;

typedef struct xsh_handle xsh_handle;

typedef struct {
  int nsites;               /* lattice sites per trajectory                     */
  int nchk;                 /* 1 + checks                                       */
  long long avlen;          /* complex values of av per (j, nc) path            */
  long long obs_offset[XSH_MAX_OBS]; /* complex offset of observable o in av     */
  int obs_sites[XSH_MAX_OBS];        /* sites per time point of observable o      */
  long long traj_steps;     /* trajectory-steps executed by xsh_run_ensemble    */
  int threads, traj_per_block; /* launch shape of the path kernel              */
  int kernel_kind;          /* 0 register-resident FFT, 1 streaming (fields in
                               device memory), 2 dense transforms, 3 SDE       */
  int ensembles[3];         /* the descriptor's ensembles (sizes of a0, a_out
                               and of the averages of xsh_read_averages)       */
  int points;               /* stored time points                              */
} xsh_info;

/* Library version (XSH_ABI_VERSION). */
int xsh_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* xsh_last_error(void);

/* Validate a descriptor, derive the lattice (xpreferences.m:133-175,
 * spdpreferences.m:16-72), build propagator and twiddle tables
 * (xpropfactor.m:23-58) and allocate device buffers.
 * Replaces the per-(j,nc) setup of xensemble.m:41-53. */
int xsh_create(const xsh_desc* desc, xsh_handle** out);

/* Everything xsh_create derives on the host, without a device: validates the
 * descriptor, derives the lattice, picks the kernel, builds every table, and
 * fills *info as xsh_query would for the handle.  XSH_OK means xsh_create
 * would get as far as device allocation; XSH_EUNSUPPORTED is the reference's
 * CPU xpath (no reference counterpart: lets a MATLAB host choose the path
 * before any device is touched). */
int xsh_plan(const xsh_desc* desc, xsh_info* info);

/* Release device memory and the handle. */
void xsh_destroy(xsh_handle* h);

/* Sizes of the outputs of a handle. */
int xsh_query(const xsh_handle* h, xsh_info* info);

/* One path, the contract of p.path (xpath.m:1, called xensemble.m:79):
 * integrates ensembles(1) trajectories of sub-ensemble j (1-based), check nc
 * (1 fine, 2 coarse) of parallel ensemble npar (1-based), from a0
 * ([1, nx, ny, ensembles(1)] complex, or NULL to use the built-in initial
 * form, xensemble.m:53-60), and writes the final fields a_out (same shape,
 * may be NULL) and the averages av_out ([avlen] complex, the p.d.av{o} blocks of
 * xpath.m:79-83 one after another, each time fastest). */
int xsh_run_path(xsh_handle* h, int npar, int j, int nc, const double* a0, double* a_out,
                 double* av_out);

/* No counterpart in the reference: xsh_run_path for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops, hosts
 * that forbid hipMalloc).  Same path and kernels, but inputs and outputs are
 * the caller's device buffers - a0_dev (NULL: the built-in initial form),
 * a_out_dev (may be NULL), av_dev laid out as xsh_run_path's av_out - and the
 * scratch (kernel arguments, the averages as the kernels accumulate them, the
 * streaming kernel's field planes, k-space noise planes) is a static device
 * array of XSH_SCRATCH_BYTES per device, held until the process ends.  Paths
 * needing more scratch fail with XSH_EINVAL; the streaming kernel's persistent
 * grid shrinks to the workgroups whose planes fit.  Asynchronous on stream (a
 * hipStream_t, NULL: the handle's stream); nothing is allocated or freed.
 * Every call on a device shares that device's array: calls that overlap on one
 * device race. */
#define XSH_SCRATCH_BYTES (64 * 1024 * 1024)
int xsh_run_path_noalloc(xsh_handle* h, int npar, int j, int nc, const double* a0_dev, double* a_out_dev,
                         double* av_dev, void* stream);

/* All ensembles(2) x (1+checks) paths of parallel ensemble npar in one
 * launch (xensemble.m:25-80), asynchronously on the handle's stream.  The
 * averages stay in device memory until xsh_read_averages. */
int xsh_run_ensemble(xsh_handle* h, int npar);

/* Wait for the handle's stream; returns the last kernel time in *ms if non-NULL. */
int xsh_synchronize(xsh_handle* h, double* ms);

/* Copy the averages of the last xsh_run_ensemble: one [avlen] block per path,
 * path (j, nc) at block (j-1)*nchk + nc-1, each laid out as xsh_run_path's av_out. */
int xsh_read_averages(xsh_handle* h, double* av_out);

/* Divergence check (no reference counterpart: MATLAB carries NaN into the
 * averages without notice): waits for the handle's stream and sets *flag to 1
 * if a trajectory's field held a NaN or infinite value at a stored time point
 * of the last xsh_run_path or xsh_run_ensemble, else 0.  The kernels test
 * every site at every stored point (xsh_run_path_noalloc does not check). */
int xsh_nonfinite(xsh_handle* h, int* flag);

/* Per-trajectory data of the last xsh_run_ensemble (no reference counterpart;
 * feeds the control-variate estimator):
 *   obs  [ensembles(2)][nchk][ensembles(1)][points][nobserve] real parts of the
 *        lattice-mean observables (traj_out = 1, NULL to skip)
 *   feat [ensembles(2)][nchk][ensembles(1)][2*nmodes] Re, Im of the chosen
 *        Fourier modes of the initial noise, scaled to unit variance (NULL to skip) */
int xsh_read_trajectories(xsh_handle* h, double* obs, float* feat);

/* Raw trajectories of the last xsh_run_ensemble (p.rawdata, xpath.m:73-78, stored
 * by the reference in raw{s,nc,j}, xensemble.m:79): one block per path in the order of
 * xsh_read_averages, each p.d.raw = [1, points, nx, ny, ensembles(1)] column-major
 * (time fastest, then x, y, trajectory); the descriptor must set raw = 1. */
int xsh_read_raw(xsh_handle* h, double* out);

/* Test hooks (no reference counterpart; used by the parity tests). */
/* Standard normals of one draw: out[t][L], L < count, for trajectories
 * traj0 .. traj0+ntraj-1, in the working precision of the handle. */
int xsh_probe_normals(xsh_handle* h, int nsp, int tag, int draw, int traj0, int ntraj, int count,
                      double* out);
/* Apply the handle's interaction-picture propagator for check nc to nplanes
 * planes ([nplanes][nx*ny] complex, in/out on the host):
 * a -> T^-1 (P . T a) with the handle's boundary transforms. */
int xsh_probe_propagate(xsh_handle* h, int nc, int nplanes, double* planes);

/* ---- batches of independent instances -----------------------------------
 * No single reference counterpart: the reference runs one project per xsim
 * call (xsim.m:1) and each parallel ensemble as a parfor worker
 * (xsim.m:51-56); a batch runs many projects - uneven lattices, ensemble sizes
 * and drift parameters - in one captured HIP graph: one persistent kernel per
 * lattice kernel family, the families concurrently.  Every instance keeps the
 * semantics and the averages layout of xsh_run_ensemble. */
typedef struct xsh_batch xsh_batch;
#define XSH_BATCH_AUTOTUNE 1  /* time every kernel variant (threads, trajectories
                                 per workgroup, first FFT radix) and persistent
                                 grid per family on a shortened run; keep the fastest */
#define XSH_BATCH_SERIAL 2    /* run the families one after another inside the graph
                                 (default: concurrently on forked streams) */
int xsh_batch_create(int n, const xsh_desc* descs, int flags, xsh_batch** out);
void xsh_batch_destroy(xsh_batch* b);
/* All instances, parallel ensemble npar, asynchronously (one graph launch). */
int xsh_batch_run(xsh_batch* b, int npar);
/* Wait; *ms = time of the last xsh_batch_run if non-NULL. */
int xsh_batch_synchronize(xsh_batch* b, double* ms);
/* Averages of instance i, laid out as xsh_read_averages. */
int xsh_batch_read_averages(xsh_batch* b, int i, double* av_out);
int xsh_batch_query(const xsh_batch* b, int i, xsh_info* info);
/* xsh_nonfinite of instance i for the last xsh_batch_run. */
int xsh_batch_nonfinite(xsh_batch* b, int i, int* flag);
/* Kernel plan: *ngroups families; plan[8*g ..] = precision, nx, ny, instances,
 * threads, trajectories per workgroup, first x radix, persistent grid. */
int xsh_batch_plan(const xsh_batch* b, int* ngroups, int* plan);

/* ---- control-variate network (truncated-Wigner estimator) -------------
 * No reference counterpart: the reference has no learned component.  A
 * 64 -> 256 -> 256 -> 1 GELU network on MFMA (bf16 operands, FP32
 * accumulation) predicts a trajectory's observable from the lowest Fourier
 * modes of its initial noise; its mean over cheap Gaussian samples plus the
 * ensemble mean of (observable - prediction) is an unbiased, lower-variance
 * estimate of the observable's mean.  Parameters are a flat float array
 * W1[nin][nh] b1[nh] W2[nh][nh] b2[nh] w3[nh] b3. */
typedef struct xsh_mlp xsh_mlp;
int xsh_mlp_create(int nin, int nh, int batch_max, const float* theta0, xsh_mlp** out);
void xsh_mlp_destroy(xsh_mlp* h);
const char* xsh_mlp_last_error(void);
int xsh_mlp_get_params(xsh_mlp* h, float* theta);
/* y[i] = MLP(x[i]) for n rows of nin floats (host arrays). */
int xsh_mlp_forward(xsh_mlp* h, long long n, const float* x, float* y);
/* Loss sum((y-t)^2)/(2n) and its gradient (host arrays, n <= batch_max). */
int xsh_mlp_loss_grad(xsh_mlp* h, int n, const float* x, const float* t, double* loss, float* grad);
/* Adam over `epochs` passes of minibatches; order[e*n + i] = sample at position i of epoch e. */
int xsh_mlp_train(xsh_mlp* h, int n, const float* x, const float* t, int epochs, int batch, const int* order,
                  float lr, double* epoch_loss);
/* Translation augmentation for xsh_mlp_train: input pair (2m, 2m+1) holds Re, Im of
 * the Fourier mode of wavenumber (kxy[2m], kxy[2m+1]) of an nx x ny periodic
 * lattice; every epoch each training sample is moved by a random lattice
 * translation (its modes multiplied by exp(-2 pi i (kx sx/nx + ky sy/ny))), an
 * exact symmetry of translation-invariant observables.  nmodes = 0 turns it off. */
int xsh_mlp_set_translations(xsh_mlp* h, int nmodes, const int* kxy, int nx, int ny, int seed);
/* Sum and sum of squares of MLP(z) over n standard-normal inputs drawn on the device. */
int xsh_mlp_gaussian_moments(xsh_mlp* h, long long n, int seed, double* sum, double* sumsq);

#ifdef __cplusplus
}
#endif
#endif /* XSPDE_HIP_H */
