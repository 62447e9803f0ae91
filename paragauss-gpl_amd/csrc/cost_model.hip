// Learned cost model of the work items: a 3-layer MLP (F -> H -> H -> 1, tanh
// hidden units) that predicts log(item time / us) from the item's features
// (kind, precision, tile, screened and dense K lengths, ...).  Training and
// inference run in FP64 on the GPU through the kernels below; every reduction
// over samples runs in a fixed order, so training is bit-reproducible.
//
// Parameter vector (flat, FP64): W1[F][H], b1[H], W2[H][H], b2[H], w3[H], b3.
// Inputs and the target are standardised with statistics of the first
// training set (kept with the model).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pgxc.h"
#include "pgxc_internal.h"

namespace pgxc {

constexpr int kH = 64;       // hidden width = one wave: lane = hidden unit
constexpr int kMaxF = 32;    // largest feature count
constexpr int kChunk = 4096; // samples per partial gradient sum

struct MlpView {
  const double *W1, *b1, *W2, *b2, *w3, *b3;
};

__device__ __forceinline__ MlpView mlp_view(const double *P, int F) {
  MlpView v;
  v.W1 = P;
  v.b1 = v.W1 + F * kH;
  v.W2 = v.b1 + kH;
  v.b2 = v.W2 + kH * kH;
  v.w3 = v.b2 + kH;
  v.b3 = v.w3 + kH;
  return v;
}

// wave-wide sum in a fixed butterfly order
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off);
  return x;
}

// Forward pass; wave = sample, lane = hidden unit.  Stores the hidden
// activations (for the backward pass) when H1/H2 are given.
__global__ __launch_bounds__(kThreads) void k_mlp_forward(const double *__restrict__ P, int F,
                                                          const double *__restrict__ X, long n,
                                                          double *H1, double *H2, double *y) {
  __shared__ double W2s[kH * kH];
  __shared__ double h1s[kThreads / 64][kH];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const MlpView m = mlp_view(P, F);
  for (int i = tid; i < kH * kH; i += kThreads) W2s[i] = m.W2[i];
  __syncthreads();
  const long nwaves = (long)gridDim.x * (kThreads / 64);
  const long rounds = (n + nwaves - 1) / nwaves;
  for (long r = 0; r < rounds; ++r) {
    const long s = r * nwaves + (long)blockIdx.x * (kThreads / 64) + wave;
    const bool ok = s < n;
    double a = m.b1[lane];
    if (ok)
      for (int f = 0; f < F; ++f) a += X[s * F + f] * m.W1[f * kH + lane];
    const double h1 = tanh(a);
    h1s[wave][lane] = h1;
    __syncthreads();
    double b = m.b2[lane];
    for (int i = 0; i < kH; ++i) b += h1s[wave][i] * W2s[i * kH + lane];
    const double h2 = tanh(b);
    const double out = wave_sum(h2 * m.w3[lane]) + m.b3[0];
    if (ok) {
      if (H1) H1[s * kH + lane] = h1;
      if (H2) H2[s * kH + lane] = h2;
      if (lane == 0) y[s] = out;
    }
    __syncthreads();
  }
}

// Backward pass of L = (1/n) sum_s (y_s - t_s)^2 down to the hidden units:
// dy_s, dA2 = dL/d(pre-activation 2), dA1 = dL/d(pre-activation 1).
__global__ __launch_bounds__(kThreads) void k_mlp_backward(const double *__restrict__ P, int F,
                                                           long n, const double *__restrict__ H1,
                                                           const double *__restrict__ H2,
                                                           const double *__restrict__ y,
                                                           const double *__restrict__ t,
                                                           double *dy, double *dA1, double *dA2) {
  __shared__ double W2s[kH * kH];
  __shared__ double d2s[kThreads / 64][kH];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const MlpView m = mlp_view(P, F);
  for (int i = tid; i < kH * kH; i += kThreads) W2s[i] = m.W2[i];
  __syncthreads();
  const long nwaves = (long)gridDim.x * (kThreads / 64);
  const long rounds = (n + nwaves - 1) / nwaves;
  for (long r = 0; r < rounds; ++r) {
    const long s = r * nwaves + (long)blockIdx.x * (kThreads / 64) + wave;
    const bool ok = s < n;
    double g = 0.0, h1 = 0.0, h2 = 0.0;
    if (ok) {
      g = 2.0 * (y[s] - t[s]) / (double)n;
      h1 = H1[s * kH + lane];
      h2 = H2[s * kH + lane];
    }
    const double d2 = g * m.w3[lane] * (1.0 - h2 * h2);
    d2s[wave][lane] = d2;
    __syncthreads();
    double acc = 0.0;
    for (int j = 0; j < kH; ++j) acc += W2s[lane * kH + j] * d2s[wave][j];
    const double d1 = acc * (1.0 - h1 * h1);
    if (ok) {
      dA1[s * kH + lane] = d1;
      dA2[s * kH + lane] = d2;
      if (lane == 0) dy[s] = g;
    }
    __syncthreads();
  }
}

// Partial parameter gradients over one chunk of samples: thread = parameter,
// blockIdx.y = chunk; sums run over the chunk's samples in order.
__global__ __launch_bounds__(kThreads) void k_mlp_wgrad(int F, long n, const double *__restrict__ X,
                                                        const double *__restrict__ H1,
                                                        const double *__restrict__ H2,
                                                        const double *__restrict__ dy,
                                                        const double *__restrict__ dA1,
                                                        const double *__restrict__ dA2, int np,
                                                        double *partial) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  const long c = blockIdx.y;
  if (p >= np) return;
  const long s0 = c * kChunk, s1 = min(n, s0 + kChunk);
  const int oW1 = 0, ob1 = F * kH, oW2 = ob1 + kH, ob2 = oW2 + kH * kH, ow3 = ob2 + kH,
            ob3 = ow3 + kH;
  double acc = 0.0;
  if (p < ob1) {
    const int f = p / kH, j = p % kH;
    for (long s = s0; s < s1; ++s) acc += X[s * F + f] * dA1[s * kH + j];
  } else if (p < oW2) {
    const int j = p - ob1;
    for (long s = s0; s < s1; ++s) acc += dA1[s * kH + j];
  } else if (p < ob2) {
    const int i = (p - oW2) / kH, j = (p - oW2) % kH;
    for (long s = s0; s < s1; ++s) acc += H1[s * kH + i] * dA2[s * kH + j];
  } else if (p < ow3) {
    const int j = p - ob2;
    for (long s = s0; s < s1; ++s) acc += dA2[s * kH + j];
  } else if (p < ob3) {
    const int j = p - ow3;
    for (long s = s0; s < s1; ++s) acc += H2[s * kH + j] * dy[s];
  } else {
    for (long s = s0; s < s1; ++s) acc += dy[s];
  }
  (void)oW1;
  partial[c * np + p] = acc;
}

// grad = sum over chunks of the partials (in chunk order)
__global__ __launch_bounds__(kThreads) void k_mlp_reduce(const double *__restrict__ partial,
                                                         int nchunk, int np, double *grad) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= np) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += partial[(long)c * np + p];
  grad[p] = s;
}

// Adam step (Kingma & Ba, 2015) with bias correction; step = 1, 2, ...
__global__ __launch_bounds__(kThreads) void k_mlp_adam(double *P, const double *__restrict__ g,
                                                       double *m1, double *m2, int np, double lr,
                                                       int step) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= np) return;
  const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
  const double gp = g[p];
  const double a = b1 * m1[p] + (1.0 - b1) * gp;
  const double v = b2 * m2[p] + (1.0 - b2) * gp * gp;
  m1[p] = a;
  m2[p] = v;
  const double ah = a / (1.0 - pow(b1, (double)step)), vh = v / (1.0 - pow(b2, (double)step));
  P[p] -= lr * ah / (sqrt(vh) + eps);
}

}  // namespace pgxc

using namespace pgxc;

struct pgxc_costmodel {
  int F = 0, np = 0;
  std::string err;
  std::vector<double> host;            // parameters
  std::vector<double> xmu, xsd;        // input standardisation
  double ymu = 0.0, ysd = 1.0;         // target standardisation
  bool have_stats = false;
  int adam_step = 0;
  long version = 0;  // bumped whenever the parameters change
  double *P = nullptr, *m1 = nullptr, *m2 = nullptr, *g = nullptr;
  hipStream_t stream = nullptr;
};

namespace pgxc {
long costmodel_version(const pgxc_costmodel *m) { return m ? m->version : -1; }
int costmodel_features(const pgxc_costmodel *m) { return m ? m->F : 0; }
}  // namespace pgxc

namespace {

int cm_fail(pgxc_costmodel *m, int code, const std::string &msg) {
  if (m) m->err = msg;
  return code;
}

#define CMCHK(m, call)                                                                 \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return cm_fail(m, e_ == hipErrorOutOfMemory ? PGXC_ERR_NOMEM : PGXC_ERR_HIP,     \
                     std::string(#call) + ": " + hipGetErrorString(e_));               \
  } while (0)

// splitmix64 (Steele, Lea & Flood, 2014): a deterministic initialisation
uint64_t splitmix(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

unsigned grid_for(long n_waves) {
  const long wg = (n_waves + 3) / 4;
  return (unsigned)std::max<long>(1, std::min<long>(wg, 1024));
}

// device scratch of one pass over n samples
struct Work {
  double *X = nullptr, *t = nullptr, *H1 = nullptr, *H2 = nullptr, *y = nullptr, *dy = nullptr,
         *dA1 = nullptr, *dA2 = nullptr, *partial = nullptr;
  ~Work() {
    for (double *p : {X, t, H1, H2, y, dy, dA1, dA2, partial})
      if (p) (void)hipFree(p);
  }
};

int standardise(pgxc_costmodel *m, const double *X, long n, std::vector<double> &Xs) {
  const int F = m->F;
  Xs.resize((size_t)n * F);
  for (long s = 0; s < n; ++s)
    for (int f = 0; f < F; ++f) {
      const double v = X[s * F + f];
      if (!std::isfinite(v)) return cm_fail(m, PGXC_ERR_ARG, "non-finite feature");
      Xs[(size_t)s * F + f] = (v - m->xmu[f]) / m->xsd[f];
    }
  return PGXC_OK;
}

int alloc_work(pgxc_costmodel *m, Work &w, long n, bool train) {
  const int F = m->F;
  CMCHK(m, hipMalloc((void **)&w.X, sizeof(double) * std::max<long>(1, n * F)));
  CMCHK(m, hipMalloc((void **)&w.y, sizeof(double) * std::max<long>(1, n)));
  if (!train) return PGXC_OK;
  const long nchunk = (n + kChunk - 1) / kChunk;
  CMCHK(m, hipMalloc((void **)&w.t, sizeof(double) * n));
  CMCHK(m, hipMalloc((void **)&w.H1, sizeof(double) * n * kH));
  CMCHK(m, hipMalloc((void **)&w.H2, sizeof(double) * n * kH));
  CMCHK(m, hipMalloc((void **)&w.dy, sizeof(double) * n));
  CMCHK(m, hipMalloc((void **)&w.dA1, sizeof(double) * n * kH));
  CMCHK(m, hipMalloc((void **)&w.dA2, sizeof(double) * n * kH));
  CMCHK(m, hipMalloc((void **)&w.partial, sizeof(double) * nchunk * m->np));
  return PGXC_OK;
}

// forward + backward + gradient (into m->g) of the standardised problem
void launch_gradient(pgxc_costmodel *m, Work &w, long n) {
  hipStream_t s = m->stream;
  const unsigned gw = grid_for(n);
  hipLaunchKernelGGL(k_mlp_forward, dim3(gw), dim3(kThreads), 0, s, m->P, m->F, w.X, n, w.H1,
                     w.H2, w.y);
  hipLaunchKernelGGL(k_mlp_backward, dim3(gw), dim3(kThreads), 0, s, m->P, m->F, n, w.H1, w.H2,
                     w.y, w.t, w.dy, w.dA1, w.dA2);
  const int nchunk = (int)((n + kChunk - 1) / kChunk);
  const unsigned gp = (unsigned)((m->np + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(gp, nchunk), dim3(kThreads), 0, s, m->F, n, w.X, w.H1,
                     w.H2, w.dy, w.dA1, w.dA2, m->np, w.partial);
  hipLaunchKernelGGL(k_mlp_reduce, dim3(gp), dim3(kThreads), 0, s, w.partial, nchunk, m->np, m->g);
}

int upload_problem(pgxc_costmodel *m, Work &w, const double *X, const double *y, long n) {
  std::vector<double> Xs;
  int rc = standardise(m, X, n, Xs);
  if (rc) return rc;
  CMCHK(m, hipMemcpy(w.X, Xs.data(), sizeof(double) * Xs.size(), hipMemcpyHostToDevice));
  if (y) {
    std::vector<double> ts((size_t)n);
    for (long s = 0; s < n; ++s) {
      if (!std::isfinite(y[s])) return cm_fail(m, PGXC_ERR_ARG, "non-finite target");
      ts[(size_t)s] = (y[s] - m->ymu) / m->ysd;
    }
    CMCHK(m, hipMemcpy(w.t, ts.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  }
  return PGXC_OK;
}

double host_loss(pgxc_costmodel *m, Work &w, long n, const double *y) {
  std::vector<double> out((size_t)n);
  if (hipMemcpy(out.data(), w.y, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
    return NAN;
  double l = 0.0;
  for (long s = 0; s < n; ++s) {
    const double d = out[(size_t)s] - (y[s] - m->ymu) / m->ysd;
    l += d * d;
  }
  return l / (double)n;
}

}  // namespace

extern "C" {

int pgxc_costmodel_create(pgxc_costmodel **out, int32_t n_features, uint64_t seed) {
  if (!out) return PGXC_ERR_ARG;
  *out = nullptr;
  pgxc_costmodel *m = new pgxc_costmodel();
  *out = m;
  if (n_features <= 0 || n_features > kMaxF)
    return cm_fail(m, PGXC_ERR_ARG, "n_features must be 1.." + std::to_string(kMaxF));
  m->F = n_features;
  m->np = n_features * kH + kH + kH * kH + kH + kH + 1;
  m->host.assign((size_t)m->np, 0.0);
  m->xmu.assign((size_t)m->F, 0.0);
  m->xsd.assign((size_t)m->F, 1.0);
  // Glorot-uniform weights, zero biases
  uint64_t st = seed;
  auto uni = [&](double lim) { return lim * (2.0 * ((splitmix(st) >> 11) * 0x1.0p-53) - 1.0); };
  double *W1 = m->host.data(), *W2 = W1 + m->F * kH + kH, *w3 = W2 + kH * kH + kH;
  for (int i = 0; i < m->F * kH; ++i) W1[i] = uni(std::sqrt(6.0 / (m->F + kH)));
  for (int i = 0; i < kH * kH; ++i) W2[i] = uni(std::sqrt(6.0 / (2.0 * kH)));
  for (int i = 0; i < kH; ++i) w3[i] = uni(std::sqrt(6.0 / (kH + 1.0)));
  CMCHK(m, hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  for (double **p : {&m->P, &m->m1, &m->m2, &m->g})
    CMCHK(m, hipMalloc((void **)p, sizeof(double) * m->np));
  CMCHK(m, hipMemcpy(m->P, m->host.data(), sizeof(double) * m->np, hipMemcpyHostToDevice));
  CMCHK(m, hipMemset(m->m1, 0, sizeof(double) * m->np));
  CMCHK(m, hipMemset(m->m2, 0, sizeof(double) * m->np));
  return PGXC_OK;
}

int32_t pgxc_costmodel_nparams(const pgxc_costmodel *m) { return m ? m->np : -1; }

int pgxc_costmodel_params(pgxc_costmodel *m, double *params, double *norm) {
  if (!m || !params) return PGXC_ERR_ARG;
  CMCHK(m, hipMemcpy(params, m->P, sizeof(double) * m->np, hipMemcpyDeviceToHost));
  if (norm) {
    for (int f = 0; f < m->F; ++f) {
      norm[f] = m->xmu[f];
      norm[m->F + f] = m->xsd[f];
    }
    norm[2 * m->F] = m->ymu;
    norm[2 * m->F + 1] = m->ysd;
  }
  return PGXC_OK;
}

int pgxc_costmodel_set_params(pgxc_costmodel *m, const double *params, const double *norm) {
  if (!m || !params) return PGXC_ERR_ARG;
  for (int i = 0; i < m->np; ++i)
    if (!std::isfinite(params[i])) return cm_fail(m, PGXC_ERR_ARG, "non-finite parameter");
  CMCHK(m, hipMemcpy(m->P, params, sizeof(double) * m->np, hipMemcpyHostToDevice));
  if (norm) {
    for (int f = 0; f < m->F; ++f) {
      if (!(norm[m->F + f] > 0.0)) return cm_fail(m, PGXC_ERR_ARG, "feature scale must be > 0");
      m->xmu[f] = norm[f];
      m->xsd[f] = norm[m->F + f];
    }
    if (!(norm[2 * m->F + 1] > 0.0)) return cm_fail(m, PGXC_ERR_ARG, "target scale must be > 0");
    m->ymu = norm[2 * m->F];
    m->ysd = norm[2 * m->F + 1];
    m->have_stats = true;
  }
  m->adam_step = 0;
  ++m->version;
  CMCHK(m, hipMemset(m->m1, 0, sizeof(double) * m->np));
  CMCHK(m, hipMemset(m->m2, 0, sizeof(double) * m->np));
  return PGXC_OK;
}

int pgxc_costmodel_gradient(pgxc_costmodel *m, const double *X, const double *y, int64_t n,
                            double *grad, double *loss) {
  if (!m || !X || !y || !grad || n <= 0) return PGXC_ERR_ARG;
  Work w;
  int rc = alloc_work(m, w, n, true);
  if (rc) return rc;
  if ((rc = upload_problem(m, w, X, y, n))) return rc;
  launch_gradient(m, w, n);
  CMCHK(m, hipGetLastError());
  CMCHK(m, hipStreamSynchronize(m->stream));
  CMCHK(m, hipMemcpy(grad, m->g, sizeof(double) * m->np, hipMemcpyDeviceToHost));
  if (loss) *loss = host_loss(m, w, n, y);
  return PGXC_OK;
}

int pgxc_costmodel_train(pgxc_costmodel *m, const double *X, const double *y, int64_t n,
                         int32_t steps, double lr, double *loss) {
  if (!m || !X || !y || n <= 0 || steps < 0 || !(lr > 0.0)) return PGXC_ERR_ARG;
  if (!m->have_stats) {  // standardisation from the first training set
    const int F = m->F;
    for (int f = 0; f < F; ++f) {
      double mu = 0.0, sd = 0.0;
      for (long s = 0; s < n; ++s) mu += X[s * F + f];
      mu /= (double)n;
      for (long s = 0; s < n; ++s) sd += (X[s * F + f] - mu) * (X[s * F + f] - mu);
      sd = std::sqrt(sd / (double)n);
      m->xmu[f] = mu;
      m->xsd[f] = sd > 1e-12 ? sd : 1.0;
    }
    double mu = 0.0, sd = 0.0;
    for (long s = 0; s < n; ++s) mu += y[s];
    mu /= (double)n;
    for (long s = 0; s < n; ++s) sd += (y[s] - mu) * (y[s] - mu);
    sd = std::sqrt(sd / (double)n);
    m->ymu = mu;
    m->ysd = sd > 1e-12 ? sd : 1.0;
    m->have_stats = true;
  }
  Work w;
  int rc = alloc_work(m, w, n, true);
  if (rc) return rc;
  if ((rc = upload_problem(m, w, X, y, n))) return rc;
  const unsigned gp = (unsigned)((m->np + kThreads - 1) / kThreads);
  ++m->version;
  for (int k = 0; k < steps; ++k) {
    launch_gradient(m, w, n);
    ++m->adam_step;
    hipLaunchKernelGGL(k_mlp_adam, dim3(gp), dim3(kThreads), 0, m->stream, m->P, m->g, m->m1,
                       m->m2, m->np, lr, m->adam_step);
  }
  CMCHK(m, hipGetLastError());
  if (loss) {
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid_for(n)), dim3(kThreads), 0, m->stream, m->P,
                       m->F, w.X, (long)n, (double *)nullptr, (double *)nullptr, w.y);
    CMCHK(m, hipStreamSynchronize(m->stream));
    *loss = host_loss(m, w, n, y);
  }
  CMCHK(m, hipStreamSynchronize(m->stream));
  return PGXC_OK;
}

int pgxc_costmodel_predict(pgxc_costmodel *m, const double *X, int64_t n, double *out) {
  if (!m || !X || !out || n < 0) return PGXC_ERR_ARG;
  if (n == 0) return PGXC_OK;
  Work w;
  int rc = alloc_work(m, w, n, false);
  if (rc) return rc;
  if ((rc = upload_problem(m, w, X, nullptr, n))) return rc;
  hipLaunchKernelGGL(k_mlp_forward, dim3(grid_for(n)), dim3(kThreads), 0, m->stream, m->P, m->F,
                     w.X, (long)n, (double *)nullptr, (double *)nullptr, w.y);
  CMCHK(m, hipGetLastError());
  CMCHK(m, hipStreamSynchronize(m->stream));
  CMCHK(m, hipMemcpy(out, w.y, sizeof(double) * n, hipMemcpyDeviceToHost));
  for (long s = 0; s < n; ++s) out[s] = out[s] * m->ysd + m->ymu;
  return PGXC_OK;
}

const char *pgxc_costmodel_last_error(const pgxc_costmodel *m) {
  return m ? m->err.c_str() : "null cost model";
}

int pgxc_costmodel_destroy(pgxc_costmodel *m) {
  if (!m) return PGXC_OK;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (double *p : {m->P, m->m1, m->m2, m->g})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return PGXC_OK;
}

}  // extern "C"
