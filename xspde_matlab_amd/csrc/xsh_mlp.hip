// Control-variate network of the truncated-Wigner estimator: y = MLP(z),
// 64 -> 256 -> 256 -> 1 with GELU, trained and evaluated on the GPU.
//
// Every matrix product runs on MFMA: v_mfma_f32_16x16x16_bf16 tiles, one
// 16x16 output tile per wavefront, operands rounded to bfloat16 (nearest
// even) when they are loaded - the weights the tiles consume are bf16 - and
// FP32 accumulation.  Epilogues fuse the bias, GELU and the GELU derivative.
// The FP32 master weights and the Adam moments stay in device memory.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/xspde_hip.h"
#include "xsh_common.hpp"

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_bits(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (short)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

enum { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_DGELU = 3 };

struct Mat {
  const float* p;
  long long s0, s1;  // element (i, j) at p[i*s0 + j*s1]
};

// C[M x N] = A[M x K] B[K x N] (+ epilogue); one 16x16 tile per wave, 4 waves per block.
// MFMA 16x16x16 operand layout: lane l supplies A[l%16][4*(l/16)+j] and
// B[4*(l/16)+j][l%16], j < 4, and receives D[4*(l/16)+j][l%16].
template <int EPI>
__global__ __launch_bounds__(256) void gemm16(int M, int N, int K, Mat A, Mat B, float* C, long long ldc,
                                              const float* bias, float* Zout, const float* Zin) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tilesN = (N + 15) >> 4;
  const long long tile = (long long)blockIdx.x * 4 + wave;
  const long long tm = tile / tilesN;
  const int tn = (int)(tile % tilesN);
  if (tm * 16 >= M) return;
  const int r = lane & 15, g = lane >> 4;
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  const long long am = tm * 16 + r;
  const int bn = tn * 16 + r;
  for (int k0 = 0; k0 < K; k0 += 16) {
    v4s a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + 4 * g + j;
      const float av = (am < M && k < K) ? A.p[am * A.s0 + k * A.s1] : 0.f;
      const float bv = (k < K && bn < N) ? B.p[(long long)k * B.s0 + (long long)bn * B.s1] : 0.f;
      a[j] = bf16_bits(av);
      b[j] = bf16_bits(bv);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const long long m = tm * 16 + 4 * g + j;
    const int n = tn * 16 + r;
    if (m >= M || n >= N) continue;
    float v = acc[j];
    const long long o = m * ldc + n;
    if (EPI == EPI_BIAS || EPI == EPI_BIAS_GELU) v += bias[n];
    if (EPI == EPI_BIAS_GELU) {
      Zout[o] = v;
      v = gelu(v);
    }
    if (EPI == EPI_DGELU) v *= gelu_grad(Zin[o]);
    C[o] = v;
  }
}

template <int EPI>
hipError_t gemm(hipStream_t s, int M, int N, int K, Mat A, Mat B, float* C, long long ldc,
                const float* bias = nullptr, float* Zout = nullptr, const float* Zin = nullptr) {
  const long long tiles = (long long)((M + 15) / 16) * ((N + 15) / 16);
  const long long blocks = (tiles + 3) / 4;
  hipLaunchKernelGGL(gemm16<EPI>, dim3((unsigned)blocks), dim3(256), 0, s, M, N, K, A, B, C, ldc, bias, Zout, Zin);
  return hipGetLastError();
}

__global__ void zero_dbl(double* p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ void fill_kernel(float* p, long long n, float v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// dy = (y - t) / n and per-block partial sums of (y - t)^2 (loss)
__global__ void loss_kernel(const float* y, const float* t, float* dy, int n, float inv_n, double* loss) {
  __shared__ double red[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e2 = 0.0;
  if (i < n) {
    const float e = y[i] - t[i];
    dy[i] = e * inv_n;
    e2 = (double)e * e;
  }
  for (int o = 32; o > 0; o >>= 1) e2 += __shfl_down(e2, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e2;
  __syncthreads();
  if (threadIdx.x == 0) unsafeAtomicAdd(loss, 0.5 * (red[0] + red[1] + red[2] + red[3]) * inv_n);
}

// Per-epoch training state in device memory, written by set_state before
// each epoch (so one captured epoch replays with new values):
//   [0] epoch index into the caller's order array, [1] translation epoch,
//   [2] Adam steps taken before this epoch
__global__ void set_state(int* st, int e, int tepoch, int t0) {
  st[0] = e;
  st[1] = tepoch;
  st[2] = t0;
}

// Adam step number st[2] + b + 1 (b: the minibatch within the epoch), bias
// corrections 1 - beta^t computed here
__global__ void adam_kernel(float* th, const float* g, float* m, float* v, int n, float lr, float b1, float b2,
                            const int* st, int b, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float t = (float)(st[2] + b + 1);
  const float c1 = 1.f - powf(b1, t), c2 = 1.f - powf(b2, t);
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  th[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// gather rows idx[0..n) of X (nin wide) and t into contiguous buffers
// rows of position r0 .. r0 + n of epoch st[0] in the order array (ntot rows per epoch)
__global__ void gather_kernel(const float* X, const float* t, const int* order, const int* st, long long ntot, int r0,
                              int n, int nin, float* Xb, float* tb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * nin) return;
  const int* idx = order + (long long)st[0] * ntot + r0;
  const int r = i / nin, c = i % nin;
  Xb[i] = X[(long long)idx[r] * nin + c];
  if (c == 0) tb[r] = t[idx[r]];
}

// gather rows idx[0..n) as above, each moved by a random lattice translation:
// input pair (2m, 2m+1) is the Fourier mode of wavenumber (kxy[2m], kxy[2m+1])
// of an nx x ny periodic lattice, so a translation by (sx, sy) multiplies it by
// exp(-2 pi i (kx sx / nx + ky sy / ny)).  The shift of sample idx in epoch e is
// drawn from the counter stream, key (seed, 4), counter (e, idx, 0, 0).
__global__ void gather_translate(const float* X, const float* t, const int* order, const int* st, long long ntot,
                                 int r0, int n, int nin, float* Xb, float* tb, const int* kxy, int nx, int ny,
                                 uint32_t seed) {
  const int np = nin / 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * np) return;
  const int* idx = order + (long long)st[0] * ntot + r0;
  const uint32_t epoch = (uint32_t)st[1];
  const int r = i / np, m = i % np;
  const int src = idx[r];
  const xsh::u4 u = xsh::philox10(xsh::u4{epoch, (uint32_t)src, 0u, 0u}, seed, 4u);
  const int sx = (int)(u.x % (uint32_t)nx), sy = (int)(u.y % (uint32_t)ny);
  // phase in turns, reduced exactly in integers before the float conversion
  const long long tx = ((long long)kxy[2 * m] * sx) % nx, ty = ((long long)kxy[2 * m + 1] * sy) % ny;
  const float turns = -((float)tx / (float)nx + (float)ty / (float)ny);
  float sn, cs;
  sincospif(2.0f * turns, &sn, &cs);
  const float a = X[(long long)src * nin + 2 * m], b = X[(long long)src * nin + 2 * m + 1];
  Xb[(long long)r * nin + 2 * m] = a * cs - b * sn;
  Xb[(long long)r * nin + 2 * m + 1] = a * sn + b * cs;
  if (m == 0) tb[r] = t[src];
}

// standard normal features z[s][j] from the counter stream (key (seed, 3))
__global__ void gauss_features(float* Z, int n, int nin, uint32_t seed, uint64_t s0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int groups = (nin + 3) / 4;
  if (i >= n * groups) return;
  const int r = i / groups, gq = i % groups;
  const uint64_t sidx = s0 + (uint64_t)r;
  float z[4];
  xsh::normals4<float>(xsh::u4{(uint32_t)gq, (uint32_t)sidx, (uint32_t)(sidx >> 32), 0u}, seed, 3u, z);
  for (int l = 0; l < 4; ++l)
    if (4 * gq + l < nin) Z[(long long)r * nin + 4 * gq + l] = z[l];
}

__global__ void sum_kernel(const float* y, int n, double* acc) {
  __shared__ double red[8];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double a = 0.0, b = 0.0;
  if (i < n) {
    a = y[i];
    b = (double)y[i] * y[i];
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    b += __shfl_down(b, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[2 * (threadIdx.x >> 6)] = a;
    red[2 * (threadIdx.x >> 6) + 1] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsafeAtomicAdd(acc, red[0] + red[2] + red[4] + red[6]);
    unsafeAtomicAdd(acc + 1, red[1] + red[3] + red[5] + red[7]);
  }
}

thread_local std::string g_mlp_err;
int mfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_mlp_err = buf;
  return code;
}
#define MCHK(expr)                                                                                \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) return mfail(XSH_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_));      \
  } while (0)

}  // namespace

struct xsh_mlp {
  int nin, nh, P, bmax;
  float *th, *g, *m, *v;                 // parameters, gradient, Adam moments
  float *X, *T, *Y, *Z1, *H1, *Z2, *H2;  // batch buffers
  float *dY, *dZ2, *dZ1, *ones;
  float *Xall, *Tall;                    // training set
  int* idx;
  long long nall;
  double* dacc;
  int adam_t;
  hipStream_t s;
  int* kxy;                 // translation augmentation: mode wavenumbers (device), or NULL
  int tnx, tny;
  uint32_t tseed, tepoch;   // stream key and epochs drawn so far
  int* st;                  // per-epoch state (set_state)
  hipGraphExec_t ep_exec;   // one captured training epoch, or NULL
  hipGraph_t ep_graph;
  long long idx_cap;        // ints allocated for the order array
  // what the captured epoch was built with (buffers and baked-in arguments)
  struct EpochKey {
    long long n;
    int batch, tnx, tny;
    uint32_t tseed;
    float lr;
    const void *X, *T, *idx, *kxy;
    bool operator==(const EpochKey& o) const {
      return n == o.n && batch == o.batch && tnx == o.tnx && tny == o.tny && tseed == o.tseed && lr == o.lr &&
             X == o.X && T == o.T && idx == o.idx && kxy == o.kxy;
    }
  } ep_key;
};

extern "C" const char* xsh_mlp_last_error(void) { return g_mlp_err.c_str(); }

static void mlp_free(xsh_mlp* h) {
  if (!h) return;
  float** bufs[] = {&h->th, &h->g, &h->m, &h->v, &h->X, &h->T, &h->Y, &h->Z1, &h->H1, &h->Z2,
                    &h->H2, &h->dY, &h->dZ2, &h->dZ1, &h->ones, &h->Xall, &h->Tall};
  for (auto b : bufs)
    if (*b) (void)hipFree(*b);
  if (h->idx) (void)hipFree(h->idx);
  if (h->kxy) (void)hipFree(h->kxy);
  if (h->dacc) (void)hipFree(h->dacc);
  if (h->st) (void)hipFree(h->st);
  if (h->ep_exec) (void)hipGraphExecDestroy(h->ep_exec);
  if (h->ep_graph) (void)hipGraphDestroy(h->ep_graph);
  if (h->s) (void)hipStreamDestroy(h->s);
  delete h;
}

extern "C" int xsh_mlp_create(int nin, int nh, int batch_max, const float* theta0, xsh_mlp** out) {
  if (!out || !theta0) return mfail(XSH_EINVAL, "NULL argument");
  *out = nullptr;
  if (nin < 1 || nh < 1 || batch_max < 1) return mfail(XSH_EINVAL, "sizes must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return mfail(XSH_EDEVICE, "no HIP device");
  auto* h = new xsh_mlp();
  std::memset(h, 0, sizeof(*h));
  h->nin = nin;
  h->nh = nh;
  h->bmax = batch_max;
  h->P = nin * nh + nh + nh * nh + nh + nh + 1;
  const size_t P = h->P, B = batch_max;
  struct {
    float** p;
    size_t n;
  } al[] = {{&h->th, P}, {&h->g, P}, {&h->m, P}, {&h->v, P}, {&h->X, B * nin}, {&h->T, B}, {&h->Y, B},
            {&h->Z1, B * nh}, {&h->H1, B * nh}, {&h->Z2, B * nh}, {&h->H2, B * nh}, {&h->dY, B},
            {&h->dZ2, B * nh}, {&h->dZ1, B * nh}, {&h->ones, B}};
  for (auto& a : al)
    if (hipMalloc(a.p, a.n * sizeof(float)) != hipSuccess) {
      mlp_free(h);
      return mfail(XSH_ENOMEM, "hipMalloc (%zu floats)", a.n);
    }
  if (hipMalloc(&h->dacc, 4 * sizeof(double)) != hipSuccess || hipMalloc(&h->st, 4 * sizeof(int)) != hipSuccess ||
      hipStreamCreate(&h->s) != hipSuccess) {
    mlp_free(h);
    return mfail(XSH_EDEVICE, "allocation");
  }
  if (hipMemcpy(h->th, theta0, P * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(h->m, 0, P * sizeof(float)) != hipSuccess || hipMemset(h->v, 0, P * sizeof(float)) != hipSuccess) {
    mlp_free(h);
    return mfail(XSH_EDEVICE, "initial parameters");
  }
  hipLaunchKernelGGL(fill_kernel, dim3((B + 255) / 256), dim3(256), 0, h->s, h->ones, (long long)B, 1.0f);
  if (hipStreamSynchronize(h->s) != hipSuccess) {
    mlp_free(h);
    return mfail(XSH_EDEVICE, "initialisation");
  }
  *out = h;
  return XSH_OK;
}

extern "C" void xsh_mlp_destroy(xsh_mlp* h) {
  if (h && h->s) (void)hipStreamSynchronize(h->s);
  mlp_free(h);
}

// parameter views
static float* W1(xsh_mlp* h) { return h->th; }
static float* b1(xsh_mlp* h) { return h->th + h->nin * h->nh; }
static float* W2(xsh_mlp* h) { return b1(h) + h->nh; }
static float* b2(xsh_mlp* h) { return W2(h) + h->nh * h->nh; }
static float* w3(xsh_mlp* h) { return b2(h) + h->nh; }
static float* b3(xsh_mlp* h) { return w3(h) + h->nh; }

// forward of n <= bmax rows held in h->X; output in h->Y
static int forward_dev(xsh_mlp* h, int n) {
  const int nin = h->nin, nh = h->nh;
  MCHK((gemm<EPI_BIAS_GELU>(h->s, n, nh, nin, Mat{h->X, nin, 1}, Mat{W1(h), nh, 1}, h->H1, nh, b1(h), h->Z1)));
  MCHK((gemm<EPI_BIAS_GELU>(h->s, n, nh, nh, Mat{h->H1, nh, 1}, Mat{W2(h), nh, 1}, h->H2, nh, b2(h), h->Z2)));
  MCHK((gemm<EPI_BIAS>(h->s, n, 1, nh, Mat{h->H2, nh, 1}, Mat{w3(h), 1, 1}, h->Y, 1, b3(h))));
  return XSH_OK;
}

// gradient of L = sum (y - t)^2 / (2n) into h->g (targets in h->T); loss added to dacc[0]
static int backward_dev(xsh_mlp* h, int n) {
  const int nin = h->nin, nh = h->nh;
  float* g = h->g;
  float* gW1 = g;
  float* gb1 = gW1 + nin * nh;
  float* gW2 = gb1 + nh;
  float* gb2 = gW2 + nh * nh;
  float* gw3 = gb2 + nh;
  float* gb3 = gw3 + nh;
  hipLaunchKernelGGL(loss_kernel, dim3((n + 255) / 256), dim3(256), 0, h->s, h->Y, h->T, h->dY, n, 1.0f / n, h->dacc);
  // layer 3: gw3 = H2^T dy, gb3 = 1^T dy
  MCHK((gemm<EPI_STORE>(h->s, nh, 1, n, Mat{h->H2, 1, nh}, Mat{h->dY, 1, 1}, gw3, 1)));
  MCHK((gemm<EPI_STORE>(h->s, 1, 1, n, Mat{h->ones, 0, 1}, Mat{h->dY, 1, 1}, gb3, 1)));
  // dZ2 = (dy w3^T) * gelu'(Z2)
  MCHK((gemm<EPI_DGELU>(h->s, n, nh, 1, Mat{h->dY, 1, 1}, Mat{w3(h), 1, 1}, h->dZ2, nh, nullptr, nullptr, h->Z2)));
  MCHK((gemm<EPI_STORE>(h->s, nh, nh, n, Mat{h->H1, 1, nh}, Mat{h->dZ2, nh, 1}, gW2, nh)));
  MCHK((gemm<EPI_STORE>(h->s, 1, nh, n, Mat{h->ones, 0, 1}, Mat{h->dZ2, nh, 1}, gb2, nh)));
  // dZ1 = (dZ2 W2^T) * gelu'(Z1)
  MCHK((gemm<EPI_DGELU>(h->s, n, nh, nh, Mat{h->dZ2, nh, 1}, Mat{W2(h), 1, nh}, h->dZ1, nh, nullptr, nullptr, h->Z1)));
  MCHK((gemm<EPI_STORE>(h->s, nin, nh, n, Mat{h->X, 1, nin}, Mat{h->dZ1, nh, 1}, gW1, nh)));
  MCHK((gemm<EPI_STORE>(h->s, 1, nh, n, Mat{h->ones, 0, 1}, Mat{h->dZ1, nh, 1}, gb1, nh)));
  return XSH_OK;
}

extern "C" int xsh_mlp_get_params(xsh_mlp* h, float* theta) {
  if (!h || !theta) return mfail(XSH_EINVAL, "NULL argument");
  MCHK(hipStreamSynchronize(h->s));
  MCHK(hipMemcpy(theta, h->th, (size_t)h->P * sizeof(float), hipMemcpyDeviceToHost));
  return XSH_OK;
}

extern "C" int xsh_mlp_forward(xsh_mlp* h, long long n, const float* x, float* y) {
  if (!h || !x || !y || n < 0) return mfail(XSH_EINVAL, "bad argument");
  for (long long r0 = 0; r0 < n; r0 += h->bmax) {
    const int nb = (int)std::min<long long>(h->bmax, n - r0);
    MCHK(hipMemcpyAsync(h->X, x + r0 * h->nin, (size_t)nb * h->nin * sizeof(float), hipMemcpyHostToDevice, h->s));
    int rc = forward_dev(h, nb);
    if (rc) return rc;
    MCHK(hipMemcpyAsync(y + r0, h->Y, (size_t)nb * sizeof(float), hipMemcpyDeviceToHost, h->s));
    MCHK(hipStreamSynchronize(h->s));
  }
  return XSH_OK;
}

extern "C" int xsh_mlp_loss_grad(xsh_mlp* h, int n, const float* x, const float* t, double* loss, float* grad) {
  if (!h || !x || !t || !loss || !grad) return mfail(XSH_EINVAL, "NULL argument");
  if (n < 1 || n > h->bmax) return mfail(XSH_EINVAL, "n %d outside 1..%d", n, h->bmax);
  MCHK(hipMemcpyAsync(h->X, x, (size_t)n * h->nin * sizeof(float), hipMemcpyHostToDevice, h->s));
  MCHK(hipMemcpyAsync(h->T, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice, h->s));
  MCHK(hipMemsetAsync(h->dacc, 0, sizeof(double), h->s));
  int rc = forward_dev(h, n);
  if (!rc) rc = backward_dev(h, n);
  if (rc) return rc;
  MCHK(hipMemcpyAsync(grad, h->g, (size_t)h->P * sizeof(float), hipMemcpyDeviceToHost, h->s));
  MCHK(hipMemcpyAsync(loss, h->dacc, sizeof(double), hipMemcpyDeviceToHost, h->s));
  MCHK(hipStreamSynchronize(h->s));
  return XSH_OK;
}

// Adam on minibatches of a training set held on the device; order[e*n + i]
// is the sample of position i in epoch e (the caller's shuffle).
extern "C" int xsh_mlp_train(xsh_mlp* h, int n, const float* x, const float* t, int epochs, int batch,
                             const int* order, float lr, double* epoch_loss) {
  if (!h || !x || !t || !order) return mfail(XSH_EINVAL, "NULL argument");
  if (n < 1 || epochs < 1 || batch < 1 || batch > h->bmax) return mfail(XSH_EINVAL, "bad training sizes");
  if (h->nall < n) {
    if (h->Xall) (void)hipFree(h->Xall);
    if (h->Tall) (void)hipFree(h->Tall);
    h->Xall = h->Tall = nullptr;
    h->nall = 0;
    if (hipMalloc(&h->Xall, (size_t)n * h->nin * sizeof(float)) != hipSuccess ||
        hipMalloc(&h->Tall, (size_t)n * sizeof(float)) != hipSuccess)
      return mfail(XSH_ENOMEM, "training set");
    h->nall = n;
  }
  if (h->idx_cap < (long long)n * epochs) {
    if (h->idx) (void)hipFree(h->idx);
    h->idx = nullptr;
    h->idx_cap = 0;
    if (hipMalloc(&h->idx, (size_t)n * epochs * sizeof(int)) != hipSuccess) return mfail(XSH_ENOMEM, "order");
    h->idx_cap = (long long)n * epochs;
  }
  MCHK(hipMemcpyAsync(h->Xall, x, (size_t)n * h->nin * sizeof(float), hipMemcpyHostToDevice, h->s));
  MCHK(hipMemcpyAsync(h->Tall, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice, h->s));
  MCHK(hipMemcpyAsync(h->idx, order, (size_t)n * epochs * sizeof(int), hipMemcpyHostToDevice, h->s));
  // One epoch: clear the loss, then per minibatch gather, forward, backward
  // and Adam - about 15 small kernels each, so the epoch is captured once as a
  // HIP graph and replayed (the per-epoch values live in h->st, set by a
  // kernel before each replay; the loss is cleared by a fill kernel, not a
  // memset node).  The captured sequence is exactly the eager one.
  const int nbat = n / batch;
  auto epoch_body = [&](hipStream_t s) -> int {
    hipLaunchKernelGGL(zero_dbl, dim3(1), dim3(64), 0, s, h->dacc, 1LL);
    for (int b = 0; b < nbat; ++b) {
      const int r0 = b * batch;
      if (h->kxy)
        hipLaunchKernelGGL(gather_translate, dim3((batch * (h->nin / 2) + 255) / 256), dim3(256), 0, s, h->Xall,
                           h->Tall, h->idx, h->st, (long long)n, r0, batch, h->nin, h->X, h->T, h->kxy, h->tnx,
                           h->tny, h->tseed);
      else
        hipLaunchKernelGGL(gather_kernel, dim3((batch * h->nin + 255) / 256), dim3(256), 0, s, h->Xall, h->Tall,
                           h->idx, h->st, (long long)n, r0, batch, h->nin, h->X, h->T);
      int rc = forward_dev(h, batch);
      if (!rc) rc = backward_dev(h, batch);
      if (rc) return rc;
      hipLaunchKernelGGL(adam_kernel, dim3((h->P + 255) / 256), dim3(256), 0, s, h->th, h->g, h->m, h->v, h->P, lr,
                         0.9f, 0.999f, h->st, b, 1e-8f);
    }
    MCHK(hipGetLastError());
    return XSH_OK;
  };
  const xsh_mlp::EpochKey key{n, batch, h->tnx, h->tny, h->tseed, lr, h->Xall, h->Tall, h->idx, h->kxy};
  if (h->ep_exec && !(h->ep_key == key)) {
    (void)hipGraphExecDestroy(h->ep_exec);
    (void)hipGraphDestroy(h->ep_graph);
    h->ep_exec = nullptr;
    h->ep_graph = nullptr;
  }
  const bool graph = epochs > 1 && nbat > 0;
  if (graph && !h->ep_exec) {
    MCHK(hipStreamBeginCapture(h->s, hipStreamCaptureModeThreadLocal));
    const int rc = epoch_body(h->s);
    hipGraph_t gr = nullptr;
    const hipError_t e2 = hipStreamEndCapture(h->s, &gr);
    if (rc || e2 != hipSuccess) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc ? rc : mfail(XSH_EDEVICE, "capture of the training epoch: %s", hipGetErrorString(e2));
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t e3 = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    if (e3 != hipSuccess) {
      (void)hipGraphDestroy(gr);
      return mfail(XSH_EDEVICE, "instantiate the training epoch: %s", hipGetErrorString(e3));
    }
    h->ep_graph = gr;
    h->ep_exec = ex;
    h->ep_key = key;
  }
  for (int e = 0; e < epochs; ++e) {
    hipLaunchKernelGGL(set_state, dim3(1), dim3(1), 0, h->s, h->st, e, (int)(h->tepoch + (uint32_t)e), h->adam_t);
    MCHK(hipGetLastError());
    if (graph) {
      MCHK(hipGraphLaunch(h->ep_exec, h->s));
    } else {
      const int rc = epoch_body(h->s);
      if (rc) return rc;
    }
    h->adam_t += nbat;
    if (epoch_loss) {
      double l = 0.0;
      MCHK(hipMemcpyAsync(&l, h->dacc, sizeof(double), hipMemcpyDeviceToHost, h->s));
      MCHK(hipStreamSynchronize(h->s));
      epoch_loss[e] = nbat ? l / nbat : 0.0;
    }
  }
  MCHK(hipStreamSynchronize(h->s));
  h->tepoch += (uint32_t)epochs;
  return XSH_OK;
}

extern "C" int xsh_mlp_set_translations(xsh_mlp* h, int nmodes, const int* kxy, int nx, int ny, int seed) {
  if (!h) return mfail(XSH_EINVAL, "NULL handle");
  if (h->kxy) {
    MCHK(hipStreamSynchronize(h->s));
    (void)hipFree(h->kxy);
    h->kxy = nullptr;
  }
  if (nmodes == 0) return XSH_OK;
  if (!kxy) return mfail(XSH_EINVAL, "NULL wavenumbers");
  if (2 * nmodes != h->nin) return mfail(XSH_EINVAL, "nmodes %d must be nin/2 = %d", nmodes, h->nin / 2);
  if (nx < 1 || ny < 1) return mfail(XSH_EINVAL, "lattice %d x %d", nx, ny);
  for (int m = 0; m < 2 * nmodes; ++m)
    if (kxy[m] <= -(1 << 20) || kxy[m] >= (1 << 20)) return mfail(XSH_EINVAL, "wavenumber %d out of range", kxy[m]);
  std::vector<int> k(kxy, kxy + 2 * nmodes);
  for (int m = 0; m < nmodes; ++m) {  // non-negative residues keep the integer phase reduction exact
    k[2 * m] = ((k[2 * m] % nx) + nx) % nx;
    k[2 * m + 1] = ((k[2 * m + 1] % ny) + ny) % ny;
  }
  if (hipMalloc(&h->kxy, 2 * nmodes * sizeof(int)) != hipSuccess) return mfail(XSH_ENOMEM, "wavenumbers");
  MCHK(hipMemcpy(h->kxy, k.data(), 2 * nmodes * sizeof(int), hipMemcpyHostToDevice));
  h->tnx = nx;
  h->tny = ny;
  h->tseed = (uint32_t)seed;
  h->tepoch = 0;
  return XSH_OK;
}

// Mean and second moment of MLP(z) over n standard-normal feature vectors
// drawn on the device (the cheap samples of the control variate).
extern "C" int xsh_mlp_gaussian_moments(xsh_mlp* h, long long n, int seed, double* sum, double* sumsq) {
  if (!h || !sum || !sumsq || n < 1) return mfail(XSH_EINVAL, "bad argument");
  MCHK(hipMemsetAsync(h->dacc + 2, 0, 2 * sizeof(double), h->s));
  const int groups = (h->nin + 3) / 4;
  for (long long r0 = 0; r0 < n; r0 += h->bmax) {
    const int nb = (int)std::min<long long>(h->bmax, n - r0);
    hipLaunchKernelGGL(gauss_features, dim3((nb * groups + 255) / 256), dim3(256), 0, h->s, h->X, nb, h->nin,
                       (uint32_t)seed, (uint64_t)r0);
    int rc = forward_dev(h, nb);
    if (rc) return rc;
    hipLaunchKernelGGL(sum_kernel, dim3((nb + 255) / 256), dim3(256), 0, h->s, h->Y, nb, h->dacc + 2);
  }
  double acc[2];
  MCHK(hipMemcpyAsync(acc, h->dacc + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, h->s));
  MCHK(hipStreamSynchronize(h->s));
  *sum = acc[0];
  *sumsq = acc[1];
  return XSH_OK;
}
