// atus_net.hip -- the warm-start network of the Sobolev minimiser: a
// four-layer perceptron (tanh hidden layers, linear output) trained and
// evaluated with the BF16 MFMA kernels of k_mlp.hip.  Parameters, gradients
// and Adam moments are fp32; matrix-product operands are rounded to bf16.
// No counterpart in the reference: the network only proposes the start
// state that CSOB_Base's iteration then polishes.
#include <cmath>
#include <cstdint>
#include <vector>

#include "ctx.hpp"
#include "mlp.hpp"

using namespace atus;
using atus_host::fail;

struct atus_net {
  int dims[5] = {0, 0, 0, 0, 0};  // inputs, three hidden widths, outputs
  int max_batch = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  long long np = 0;
  long long off_w[4] = {0, 0, 0, 0}, off_b[4] = {0, 0, 0, 0};
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  float* x[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // activations, x[0] the input
  float* d[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // gradients of the activations
  float* t = nullptr;                                           // targets
  float* work = nullptr;
  size_t work_floats = 0;
  float* part = nullptr;
  std::vector<float> part_host;
  int nblk = 256;
  float* data_x = nullptr;
  float* data_t = nullptr;
  int ndata = 0;
  int* idx = nullptr;
  long long step = 0;
  bool split = true;  // matrix products as hi/lo bf16 pairs (atus_net_set_precision)
};

namespace {

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

int check_net(atus_net* n) {
  if (!n) return fail(ATUS_ERR_ARG, "null network");
  return hipSetDevice(n->device) == hipSuccess ? ATUS_OK : fail(ATUS_ERR_HIP, "hipSetDevice failed");
}

int forward(atus_net* n, int B) {
  for (int l = 0; l < 4; ++l) {
    const int in = n->dims[l], out = n->dims[l + 1];
    HIP_TRY(mlp_gemm(n->x[l], in, 1, n->p + n->off_w[l], out, 1, n->x[l + 1], B, out, in, n->p + n->off_b[l],
                     l < 3 ? 1 : 0, n->work, n->work_floats, n->stream, n->split));
  }
  return ATUS_OK;
}

// gradients of the mean squared error of x[4] against t into n->g; loss
int backward(atus_net* n, int B, double* loss) {
  const long long ny = (long long)B * n->dims[4];
  HIP_TRY(mlp_mse(n->x[4], n->t, ny, n->d[4], n->part, n->nblk, n->stream));
  for (int l = 3; l >= 0; --l) {
    const int in = n->dims[l], out = n->dims[l + 1];
    float* dz = n->d[l + 1];
    if (l < 3) HIP_TRY(mlp_tanh_back(dz, n->x[l + 1], dz, (long long)B * out, n->stream));
    // dW = x^T dZ, db = column sums of dZ, dX = dZ W^T
    HIP_TRY(mlp_gemm(n->x[l], 1, in, dz, out, 1, n->g + n->off_w[l], in, out, B, nullptr, 0, n->work,
                     n->work_floats, n->stream, n->split));
    HIP_TRY(mlp_colsum(dz, B, out, n->g + n->off_b[l], n->stream));
    if (l > 0)
      HIP_TRY(mlp_gemm(dz, out, 1, n->p + n->off_w[l], 1, out, n->d[l], B, in, out, nullptr, 0, n->work,
                       n->work_floats, n->stream, n->split));
  }
  HIP_TRY(hipMemcpyAsync(n->part_host.data(), n->part, sizeof(float) * n->nblk, hipMemcpyDeviceToHost, n->stream));
  HIP_TRY(hipStreamSynchronize(n->stream));
  double s = 0.0;
  for (int b = 0; b < n->nblk; ++b) s += n->part_host[b];
  *loss = s / (double)ny;
  return ATUS_OK;
}

int adam_step(atus_net* n, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  ++n->step;
  const float c1 = 1.f - powf(b1, (float)n->step), c2 = 1.f - powf(b2, (float)n->step);
  HIP_TRY(mlp_adam(n->p, n->g, n->m, n->v, n->np, lr, b1, b2, eps, c1, c2, n->stream));
  return ATUS_OK;
}

int check_batch(atus_net* n, int B) {
  if (B < 1 || B > n->max_batch)
    return fail(ATUS_ERR_ARG, "batch must be 1.." + std::to_string(n->max_batch));
  return ATUS_OK;
}

}  // namespace

extern "C" {

int atus_net_create(atus_net** out, int n_in, int h1, int h2, int h3, int n_out, int max_batch,
                    unsigned long long seed) {
  if (!out) return fail(ATUS_ERR_ARG, "null output pointer");
  *out = nullptr;
  const int dims[5] = {n_in, h1, h2, h3, n_out};
  for (int l = 0; l < 5; ++l)
    if (dims[l] < 1 || dims[l] > (1 << 20)) return fail(ATUS_ERR_ARG, "layer widths must be 1..2^20");
  if (max_batch < 1 || max_batch > 65536) return fail(ATUS_ERR_ARG, "max_batch must be 1..65536");
  atus_net* n = new atus_net();
  for (int l = 0; l < 5; ++l) n->dims[l] = dims[l];
  n->max_batch = max_batch;
  long long off = 0;
  for (int l = 0; l < 4; ++l) {
    n->off_w[l] = off;
    off += (long long)dims[l] * dims[l + 1];
    n->off_b[l] = off;
    off += dims[l + 1];
  }
  n->np = off;
  auto bail = [&](int code, const std::string& m) {
    atus_net_destroy(n);
    return fail(code, m);
  };
  if (hipGetDevice(&n->device) != hipSuccess) return bail(ATUS_ERR_HIP, "no HIP device");
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) return bail(ATUS_ERR_HIP, "stream");
  size_t work = 0;
  for (int l = 0; l < 4; ++l) {
    const int in = dims[l], o = dims[l + 1];
    work = std::max(work, mlp_gemm_work(max_batch, o, in));  // forward
    work = std::max(work, mlp_gemm_work(in, o, max_batch));  // dW
    work = std::max(work, mlp_gemm_work(max_batch, in, o));  // dX
  }
  n->work_floats = work;
  n->part_host.resize(n->nblk);
  bool ok = hipMalloc(&n->p, sizeof(float) * n->np) == hipSuccess && hipMalloc(&n->g, sizeof(float) * n->np) == hipSuccess &&
            hipMalloc(&n->m, sizeof(float) * n->np) == hipSuccess && hipMalloc(&n->v, sizeof(float) * n->np) == hipSuccess &&
            hipMalloc(&n->work, sizeof(float) * work) == hipSuccess &&
            hipMalloc(&n->part, sizeof(float) * n->nblk) == hipSuccess &&
            hipMalloc(&n->t, sizeof(float) * (size_t)max_batch * n_out) == hipSuccess &&
            hipMalloc(&n->idx, sizeof(int) * max_batch) == hipSuccess;
  for (int l = 0; ok && l < 5; ++l) {
    ok = hipMalloc(&n->x[l], sizeof(float) * (size_t)max_batch * dims[l]) == hipSuccess;
    if (ok && l > 0) ok = hipMalloc(&n->d[l], sizeof(float) * (size_t)max_batch * dims[l]) == hipSuccess;
  }
  if (!ok) return bail(ATUS_ERR_NOMEM, "network buffers");
  // Glorot-uniform weights, zero biases (SplitMix64 stream from `seed`)
  std::vector<float> h(n->np, 0.f);
  uint64_t s = seed;
  for (int l = 0; l < 4; ++l) {
    const double a = sqrt(6.0 / (double)(dims[l] + dims[l + 1]));
    for (long long e = 0; e < (long long)dims[l] * dims[l + 1]; ++e)
      h[n->off_w[l] + e] = (float)((2.0 * uniform01(s) - 1.0) * a);
  }
  if (hipMemcpy(n->p, h.data(), sizeof(float) * n->np, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(n->m, 0, sizeof(float) * n->np) != hipSuccess || hipMemset(n->v, 0, sizeof(float) * n->np) != hipSuccess)
    return bail(ATUS_ERR_HIP, "network initialisation");
  *out = n;
  return ATUS_OK;
}

int atus_net_destroy(atus_net* n) {
  if (!n) return ATUS_OK;
  (void)hipSetDevice(n->device);
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  for (float* q : {n->p, n->g, n->m, n->v, n->work, n->part, n->t, n->data_x, n->data_t}) (void)hipFree(q);
  for (int l = 0; l < 5; ++l) {
    (void)hipFree(n->x[l]);
    (void)hipFree(n->d[l]);
  }
  (void)hipFree(n->idx);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
  return ATUS_OK;
}

int atus_net_set_precision(atus_net* n, int bf16_terms) {
  ATUS_TRY(check_net(n));
  if (bf16_terms != 1 && bf16_terms != 3) return fail(ATUS_ERR_ARG, "bf16_terms must be 1 or 3");
  n->split = bf16_terms == 3;
  return ATUS_OK;
}

int atus_net_param_count(atus_net* n, long long* count) {
  ATUS_TRY(check_net(n));
  if (!count) return fail(ATUS_ERR_ARG, "null output");
  *count = n->np;
  return ATUS_OK;
}

int atus_net_get_params(atus_net* n, float* host) {
  ATUS_TRY(check_net(n));
  if (!host) return fail(ATUS_ERR_ARG, "null host pointer");
  HIP_TRY(hipMemcpyAsync(host, n->p, sizeof(float) * n->np, hipMemcpyDeviceToHost, n->stream));
  HIP_TRY(hipStreamSynchronize(n->stream));
  return ATUS_OK;
}

int atus_net_set_params(atus_net* n, const float* host) {
  ATUS_TRY(check_net(n));
  if (!host) return fail(ATUS_ERR_ARG, "null host pointer");
  for (long long e = 0; e < n->np; ++e)
    if (!std::isfinite(host[e])) return fail(ATUS_ERR_NUMERIC, "non-finite parameter");
  HIP_TRY(hipMemcpyAsync(n->p, host, sizeof(float) * n->np, hipMemcpyHostToDevice, n->stream));
  HIP_TRY(hipMemsetAsync(n->m, 0, sizeof(float) * n->np, n->stream));
  HIP_TRY(hipMemsetAsync(n->v, 0, sizeof(float) * n->np, n->stream));
  HIP_TRY(hipStreamSynchronize(n->stream));
  n->step = 0;
  return ATUS_OK;
}

int atus_net_forward(atus_net* n, const float* x, int batch, float* y) {
  ATUS_TRY(check_net(n));
  ATUS_TRY(check_batch(n, batch));
  if (!x || !y) return fail(ATUS_ERR_ARG, "null host pointer");
  HIP_TRY(hipMemcpyAsync(n->x[0], x, sizeof(float) * batch * n->dims[0], hipMemcpyHostToDevice, n->stream));
  ATUS_TRY(forward(n, batch));
  HIP_TRY(hipMemcpyAsync(y, n->x[4], sizeof(float) * (size_t)batch * n->dims[4], hipMemcpyDeviceToHost, n->stream));
  HIP_TRY(hipStreamSynchronize(n->stream));
  return ATUS_OK;
}

int atus_net_gradients(atus_net* n, const float* x, const float* t, int batch, float* loss, float* grads) {
  ATUS_TRY(check_net(n));
  ATUS_TRY(check_batch(n, batch));
  if (!x || !t || !loss) return fail(ATUS_ERR_ARG, "null argument");
  HIP_TRY(hipMemcpyAsync(n->x[0], x, sizeof(float) * batch * n->dims[0], hipMemcpyHostToDevice, n->stream));
  HIP_TRY(hipMemcpyAsync(n->t, t, sizeof(float) * (size_t)batch * n->dims[4], hipMemcpyHostToDevice, n->stream));
  ATUS_TRY(forward(n, batch));
  double l = 0.0;
  ATUS_TRY(backward(n, batch, &l));
  *loss = (float)l;
  if (grads) {
    HIP_TRY(hipMemcpyAsync(grads, n->g, sizeof(float) * n->np, hipMemcpyDeviceToHost, n->stream));
    HIP_TRY(hipStreamSynchronize(n->stream));
  }
  return ATUS_OK;
}

int atus_net_set_data(atus_net* n, const float* x, const float* t, int count) {
  ATUS_TRY(check_net(n));
  if (!x || !t || count < 1) return fail(ATUS_ERR_ARG, "need inputs, targets and count >= 1");
  HIP_TRY(hipStreamSynchronize(n->stream));
  (void)hipFree(n->data_x);
  (void)hipFree(n->data_t);
  n->data_x = n->data_t = nullptr;
  n->ndata = 0;
  if (hipMalloc(&n->data_x, sizeof(float) * (size_t)count * n->dims[0]) != hipSuccess ||
      hipMalloc(&n->data_t, sizeof(float) * (size_t)count * n->dims[4]) != hipSuccess)
    return fail(ATUS_ERR_NOMEM, "training set");
  HIP_TRY(hipMemcpyAsync(n->data_x, x, sizeof(float) * (size_t)count * n->dims[0], hipMemcpyHostToDevice, n->stream));
  HIP_TRY(hipMemcpyAsync(n->data_t, t, sizeof(float) * (size_t)count * n->dims[4], hipMemcpyHostToDevice, n->stream));
  HIP_TRY(hipStreamSynchronize(n->stream));
  n->ndata = count;
  return ATUS_OK;
}

int atus_net_train(atus_net* n, int steps, int batch, float lr, float lr_final, unsigned long long seed,
                   float* losses) {
  ATUS_TRY(check_net(n));
  ATUS_TRY(check_batch(n, batch));
  if (n->ndata < batch) return fail(ATUS_ERR_STATE, "training set smaller than the batch (atus_net_set_data)");
  if (steps < 1 || !(lr > 0.f) || !(lr_final >= 0.f)) return fail(ATUS_ERR_ARG, "need steps >= 1 and lr > 0");
  std::vector<int> perm(n->ndata);
  for (int i = 0; i < n->ndata; ++i) perm[i] = i;
  uint64_t s = seed;
  int pos = n->ndata;  // reshuffle before the first batch
  for (int it = 0; it < steps; ++it) {
    if (pos + batch > n->ndata) {
      for (int i = n->ndata - 1; i > 0; --i) std::swap(perm[i], perm[(int)(splitmix64(s) % (uint64_t)(i + 1))]);
      pos = 0;
    }
    HIP_TRY(hipMemcpyAsync(n->idx, perm.data() + pos, sizeof(int) * batch, hipMemcpyHostToDevice, n->stream));
    pos += batch;
    HIP_TRY(mlp_gather(n->data_x, n->idx, batch, n->dims[0], n->x[0], n->stream));
    HIP_TRY(mlp_gather(n->data_t, n->idx, batch, n->dims[4], n->t, n->stream));
    ATUS_TRY(forward(n, batch));
    double l = 0.0;
    ATUS_TRY(backward(n, batch, &l));
    if (!std::isfinite(l)) return fail(ATUS_ERR_NUMERIC, "training loss is not finite");
    if (losses) losses[it] = (float)l;
    // cosine decay from lr to lr_final
    const double c = 0.5 * (1.0 + cos(M_PI * (double)it / (double)steps));
    ATUS_TRY(adam_step(n, (float)(lr_final + (lr - lr_final) * c)));
  }
  HIP_TRY(hipStreamSynchronize(n->stream));
  return ATUS_OK;
}

int atus_net_warm_start(atus_net* n, const float* x, const double half_width[3], atus_ctx* c, int comp) {
  ATUS_TRY(check_net(n));
  if (!c || !x || !half_width) return fail(ATUS_ERR_ARG, "null argument");
  if (n->dims[4] != 32 * 32 * 32) return fail(ATUS_ERR_STATE, "the network does not output a 32^3 grid");
  if (c->dim != 3 || c->nbatch != 1) return fail(ATUS_ERR_UNSUPPORTED, "warm starts fill one 3D run");
  if (comp < 0 || comp >= c->ncomp) return fail(ATUS_ERR_ARG, "component out of range");
  for (int a = 0; a < 3; ++a)
    if (!(half_width[a] > 0.0)) return fail(ATUS_ERR_ARG, "half widths must be positive");
  // per axis: the trigonometric interpolant of the 32 periodic samples at
  // x_j = (j - 16) h, h = s/16, evaluated at the context's grid points;
  // zero outside [-s, s).  Weight of sample j at t = (x - x_j)/h:
  // sin(pi t) / (32 tan(pi t / 32)), 1 at t = 0.
  std::vector<double> E[3];
  for (int a = 0; a < 3; ++a) {
    const int na = c->n[a];
    const double s = half_width[a], h = s / 16.0;
    E[a].assign((size_t)na * 32, 0.0);
    for (int i = 0; i < na; ++i) {
      const double xi = (double)(i - na / 2) * c->dx[a];
      if (xi < -s || xi >= s) continue;
      for (int j = 0; j < 32; ++j) {
        const double t = (xi - (double)(j - 16) * h) / h;
        const double rt = std::nearbyint(t);
        double w;
        if (std::fabs(t - rt) < 1e-12)
          w = ((long long)rt % 32 == 0) ? 1.0 : 0.0;
        else
          w = std::sin(M_PI * t) / (32.0 * std::tan(M_PI * t / 32.0));
        E[a][(size_t)i * 32 + j] = w;
      }
    }
  }
  HIP_TRY(hipMemcpyAsync(n->x[0], x, sizeof(float) * n->dims[0], hipMemcpyHostToDevice, n->stream));
  ATUS_TRY(forward(n, 1));
  HIP_TRY(hipStreamSynchronize(n->stream));
  if (hipSetDevice(c->device) != hipSuccess) return fail(ATUS_ERR_HIP, "hipSetDevice failed");
  const size_t ne = (size_t)32 * (c->n[0] + c->n[1] + c->n[2]);
  const size_t nt = (size_t)c->n[0] * 32 * 32 + (size_t)c->n[0] * c->n[1] * 32;
  double* buf = nullptr;
  if (hipMalloc(&buf, sizeof(double) * (ne + nt)) != hipSuccess) return fail(ATUS_ERR_NOMEM, "interpolation buffers");
  double* e0 = buf;
  double* e1 = e0 + (size_t)32 * c->n[0];
  double* e2 = e1 + (size_t)32 * c->n[1];
  double* t1 = e2 + (size_t)32 * c->n[2];
  double* t2 = t1 + (size_t)c->n[0] * 32 * 32;
  int rc = ATUS_OK;
  if (hipMemcpyAsync(e0, E[0].data(), sizeof(double) * E[0].size(), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(e1, E[1].data(), sizeof(double) * E[1].size(), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(e2, E[2].data(), sizeof(double) * E[2].size(), hipMemcpyHostToDevice, c->stream) != hipSuccess)
    rc = fail(ATUS_ERR_HIP, "interpolation weights");
  if (rc == ATUS_OK) {
    hipError_t e = mlp_interp32(n->x[4], e0, e1, e2, t1, t2, c->psi + (long long)comp * c->npts, c->n, c->stream);
    if (e != hipSuccess) rc = fail(ATUS_ERR_HIP, std::string("interpolation: ") + hipGetErrorString(e));
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess && rc == ATUS_OK) rc = fail(ATUS_ERR_HIP, "synchronize");
  (void)hipFree(buf);
  return rc;
}

}  // extern "C"
