/*
 * shx_mlp.hip -- the learned warm start of libshellhip: a small MLP
 * (inputs -> hidden layers with GELU -> outputs) trained and evaluated with
 * hand-written gfx950 kernels.  Every GEMM runs on the matrix cores
 * (v_mfma_f32_16x16x16_bf16: BF16 operands rounded to nearest even from
 * FP32 on load, FP32 accumulation); bias, GELU and their derivatives are
 * fused into GEMM epilogues or small elementwise kernels; Adam updates FP32
 * master weights.  Activations are kept in FP32 between layers.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "shellhip.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

/* FP32 -> BF16, round to nearest even (the operands here are finite) */
__device__ __forceinline__ short bf16_rne(float x)
{
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

__device__ __forceinline__ float gelu(float x)
{
  return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float x)
{
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

/* C[M x N] = op(A)[M x K] * op(B)[K x N] (+ bias[N]) with the epilogue
 * `epi`: 0 plain, 1 store the sum in Z (unless Z is null) and GELU(sum) in
 * C, 2 accumulate into C.  op(A)[m][k] = A[m * lda + k] (ta = 0) or A[k * lda + m] (ta = 1);
 * op(B)[k][n] = B[k * ldb + n] (tb = 0) or B[n * ldb + k] (tb = 1).
 * One wave per 16 x 16 tile of C, four waves per block (32 x 32); the K
 * loop takes 16 at a time.  MFMA operand layout (16x16x16): lane l holds
 * A[l % 16][4 (l / 16) + j] and B[4 (l / 16) + j][l % 16], j = 0..3, and
 * receives D[4 (l / 16) + i][l % 16], i = 0..3. */
__global__ void __launch_bounds__(256) k_gemm_bf16(int M, int N, int K, const float* A, int lda, int ta,
                                                   const float* B, int ldb, int tb, const float* bias, float* C,
                                                   int ldc, float* Z, int epi)
{
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int m0 = blockIdx.y * 32 + (wave / 2) * 16, n0 = blockIdx.x * 32 + (wave % 2) * 16;
  const int r = lane % 16, kb = 4 * (lane / 16);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + kb + j;
      const int m = m0 + r, n = n0 + r;
      float av = 0.f, bv = 0.f;
      if (m < M && k < K) av = ta ? A[(size_t)k * lda + m] : A[(size_t)m * lda + k];
      if (k < K && n < N) bv = tb ? B[(size_t)n * ldb + k] : B[(size_t)k * ldb + n];
      a[j] = bf16_rne(av);
      b[j] = bf16_rne(bv);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + kb + i, n = n0 + r;
    if (m >= M || n >= N) continue;
    float v = acc[i] + (bias ? bias[n] : 0.f);
    const size_t o = (size_t)m * ldc + n;
    if (epi == 1) {
      if (Z) Z[o] = v;
      C[o] = gelu(v);
    } else if (epi == 2) {
      C[o] += v;
    } else {
      C[o] = v;
    }
  }
}

/* dZ = dA * GELU'(Z), elementwise */
__global__ void k_gelu_back(int n, const float* dA, const float* Z, float* dZ)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dZ[i] = dA[i] * gelu_grad(Z[i]);
}

/* mean-squared error over a [B x n] output against targets; dY = 2 (Y - T) /
 * (B n); per-block partial sums of the squared error (fixed order) */
__global__ void k_mse(int total, float scale, const float* Y, const float* T, float* dY, float* part)
{
  __shared__ float s[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.f;
  if (i < total) {
    const float d = Y[i] - T[i];
    dY[i] = 2.f * d * scale;
    e2 = d * d;
  }
  s[threadIdx.x] = e2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

/* db[n] = sum over rows of dZ[B x N] (column sums, rows in order) */
__global__ void k_colsum(int B, int N, const float* dZ, float* db)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int b = 0; b < B; ++b) s += dZ[(size_t)b * N + n];
  db[n] = s;
}

/* dst[j] = src[idx[j]] for B rows of w floats */
__global__ void k_gather_rows(int B, int w, const int* idx, const float* src, float* dst)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * w) return;
  const int j = i / w, c = i % w;
  dst[i] = src[(size_t)idx[j] * w + c];
}

/* Adam on the FP32 master weights (Kingma & Ba 2015, bias-corrected) */
__global__ void k_adam(int n, float* w, const float* g, float* m1, float* m2, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  w[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

}  // namespace

/* Scratch of the allocation-free forward pass (shx_mlp_forward_noalloc):
 * the hidden activations, ping-ponged between its two halves.  One copy per
 * GPU, held until the process ends; host code never passes this name to a
 * kernel, only the per-device address hipGetSymbolAddress gives for it. */
constexpr size_t kNoAllocFloats = (size_t)1 << 20;   /* 4 MiB */
__device__ __align__(256) float scratch[kNoAllocFloats];

/* ------------------------------------------------------------------ host */

struct shx_mlp {
  int nl = 0;                  /* layers */
  int dev = 0;                 /* the GPU that holds the network */
  float* scratch = nullptr;    /* that GPU's copy of `scratch` (device address) */
  std::vector<int> width;      /* nl + 1 */
  std::vector<size_t> woff, boff;
  size_t nparam = 0;
  int bcap = 0;                /* batch capacity of the activation buffers */
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  std::vector<float*> act, pre, dact;   /* per layer boundary: [bcap x width] */
  float *tgt = nullptr, *part = nullptr, *dtmp = nullptr;
  int step = 0;
  hipStream_t st = nullptr;
  std::string err;
};

namespace {

#define MCHK(x)                                                                            \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      m->err = std::string(#x) + ": " + hipGetErrorString(e_);                             \
      return SHX_E_HIP;                                                                    \
    }                                                                                      \
  } while (0)

thread_local std::string g_mlp_err;

/* device address of `scratch` on each GPU, looked up once per device with
 * that device current (each GPU holds its own copy at its own address) */
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_dev[kMaxDevices] = {};

hipError_t scratch_address(int dev, float** out)
{
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_dev[dev];
  return hipSuccess;
}

void gemm(hipStream_t st, int M, int N, int K, const float* A, int lda, int ta, const float* B, int ldb, int tb,
          const float* bias, float* C, int ldc, float* Z, int epi)
{
  dim3 grid((N + 31) / 32, (M + 31) / 32);
  k_gemm_bf16<<<grid, 256, 0, st>>>(M, N, K, A, lda, ta, B, ldb, tb, bias, C, ldc, Z, epi);
}

int ensure_batch(shx_mlp* m, int B)
{
  if (B <= m->bcap) return 0;
  for (auto* v : {&m->act, &m->pre, &m->dact})
    for (float* q : *v) (void)hipFree(q);
  m->act.assign(m->nl + 1, nullptr);
  m->pre.assign(m->nl + 1, nullptr);
  m->dact.assign(m->nl + 1, nullptr);
  (void)hipFree(m->tgt);
  (void)hipFree(m->part);
  (void)hipFree(m->dtmp);
  int wmax = 0;
  for (int w : m->width) wmax = std::max(wmax, w);
  for (int l = 0; l <= m->nl; ++l) {
    const size_t n = (size_t)B * m->width[l];
    MCHK(hipMalloc(&m->act[l], n * sizeof(float)));
    MCHK(hipMalloc(&m->pre[l], n * sizeof(float)));
    MCHK(hipMalloc(&m->dact[l], n * sizeof(float)));
  }
  MCHK(hipMalloc(&m->tgt, (size_t)B * m->width[m->nl] * sizeof(float)));
  MCHK(hipMalloc(&m->part, ((size_t)B * m->width[m->nl] / 256 + 2) * sizeof(float)));
  MCHK(hipMalloc(&m->dtmp, (size_t)B * wmax * sizeof(float)));
  m->bcap = B;
  return 0;
}

/* forward pass of B rows already in act[0] */
void forward(shx_mlp* m, int B)
{
  for (int l = 0; l < m->nl; ++l) {
    const int K = m->width[l], N = m->width[l + 1];
    const bool last = l == m->nl - 1;
    gemm(m->st, B, N, K, m->act[l], K, 0, m->p + m->woff[l], N, 0, m->p + m->boff[l], m->act[l + 1], N,
         m->pre[l + 1], last ? 0 : 1);
  }
}

/* loss and gradients of B rows against tgt: g = d(mean squared error)/dp */
int backward(shx_mlp* m, int B, double* loss)
{
  const int L = m->nl, nout = m->width[L];
  const int total = B * nout, nb = (total + 255) / 256;
  k_mse<<<nb, 256, 0, m->st>>>(total, 1.f / (float)total, m->act[L], m->tgt, m->dact[L], m->part);
  for (int l = L - 1; l >= 0; --l) {
    const int K = m->width[l], N = m->width[l + 1];
    const float* dZ = m->dact[l + 1];
    if (l < L - 1) {
      k_gelu_back<<<(B * N + 255) / 256, 256, 0, m->st>>>(B * N, m->dact[l + 1], m->pre[l + 1], m->dtmp);
      dZ = m->dtmp;
    }
    /* dW[K x N] = act[l]^T dZ ; db = column sums ; dact[l] = dZ W^T */
    gemm(m->st, K, N, B, m->act[l], K, 1, dZ, N, 0, nullptr, m->g + m->woff[l], N, nullptr, 0);
    k_colsum<<<(N + 255) / 256, 256, 0, m->st>>>(B, N, dZ, m->g + m->boff[l]);
    if (l > 0) gemm(m->st, B, K, N, dZ, N, 0, m->p + m->woff[l], N, 1, nullptr, m->dact[l], K, nullptr, 0);
  }
  std::vector<float> part(nb);
  MCHK(hipMemcpyAsync(part.data(), m->part, nb * sizeof(float), hipMemcpyDeviceToHost, m->st));
  MCHK(hipStreamSynchronize(m->st));
  double s = 0;
  for (float x : part) s += x;
  if (loss) *loss = s / total;
  return 0;
}

}  // namespace

extern "C" {

const char* shx_mlp_last_error(const shx_mlp* m) { return m ? m->err.c_str() : g_mlp_err.c_str(); }

int shx_mlp_create(shx_mlp** out, int n_layers, const int* width, uint64_t seed, int device)
{
  if (!out || n_layers < 1 || !width) { g_mlp_err = "invalid argument"; return SHX_E_ARG; }
  for (int l = 0; l <= n_layers; ++l)
    if (width[l] <= 0) { g_mlp_err = "invalid width"; return SHX_E_ARG; }
  shx_mlp* m = new shx_mlp();
  m->nl = n_layers;
  m->width.assign(width, width + n_layers + 1);
  size_t off = 0;
  for (int l = 0; l < n_layers; ++l) {
    m->woff.push_back(off);
    off += (size_t)width[l] * width[l + 1];
    m->boff.push_back(off);
    off += width[l + 1];
  }
  m->nparam = off;
  /* He-normal weights (Box-Muller on a xorshift stream), zero biases */
  std::vector<float> h(off, 0.f);
  uint64_t s = seed ? seed : 1;
  auto uni = [&]() {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    return ((double)(s >> 11) + 0.5) / 9007199254740992.0;
  };
  for (int l = 0; l < n_layers; ++l) {
    const double sd = std::sqrt(2.0 / width[l]);
    for (size_t i = 0; i < (size_t)width[l] * width[l + 1]; ++i) {
      const double u1 = uni(), u2 = uni();
      h[m->woff[l] + i] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    }
  }
  auto fail = [&](const char* what) {
    g_mlp_err = what;
    shx_mlp_destroy(m);
    return SHX_E_HIP;
  };
  m->dev = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&m->st) != hipSuccess) return fail("hip setup failed");
  if (scratch_address(device, &m->scratch) != hipSuccess) return fail("hipGetSymbolAddress(scratch) failed");
  if (hipMalloc(&m->p, off * sizeof(float)) != hipSuccess || hipMalloc(&m->g, off * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->m1, off * sizeof(float)) != hipSuccess || hipMalloc(&m->m2, off * sizeof(float)) != hipSuccess)
    return fail("hipMalloc failed");
  if (hipMemcpy(m->p, h.data(), off * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->g, 0, off * sizeof(float)) != hipSuccess || hipMemset(m->m1, 0, off * sizeof(float)) != hipSuccess ||
      hipMemset(m->m2, 0, off * sizeof(float)) != hipSuccess)
    return fail("upload failed");
  *out = m;
  return SHX_OK;
}

int shx_mlp_destroy(shx_mlp* m)
{
  if (!m) return SHX_OK;
  for (auto* v : {&m->act, &m->pre, &m->dact})
    for (float* q : *v) (void)hipFree(q);
  for (float* q : {m->p, m->g, m->m1, m->m2, m->tgt, m->part, m->dtmp}) (void)hipFree(q);
  if (m->st) (void)hipStreamDestroy(m->st);
  delete m;
  return SHX_OK;
}

int shx_mlp_num_params(const shx_mlp* m) { return m ? (int)m->nparam : SHX_E_ARG; }

int shx_mlp_get_params(shx_mlp* m, float* out)
{
  if (!m || !out) return SHX_E_ARG;
  MCHK(hipMemcpy(out, m->p, m->nparam * sizeof(float), hipMemcpyDeviceToHost));
  return SHX_OK;
}

int shx_mlp_set_params(shx_mlp* m, const float* in)
{
  if (!m || !in) return SHX_E_ARG;
  MCHK(hipMemcpy(m->p, in, m->nparam * sizeof(float), hipMemcpyHostToDevice));
  m->step = 0;
  MCHK(hipMemset(m->m1, 0, m->nparam * sizeof(float)));
  MCHK(hipMemset(m->m2, 0, m->nparam * sizeof(float)));
  return SHX_OK;
}

int shx_mlp_forward(shx_mlp* m, int B, const float* x, float* y)
{
  if (!m || B <= 0 || !x || !y) return SHX_E_ARG;
  if (int r = ensure_batch(m, B)) return r;
  MCHK(hipMemcpyAsync(m->act[0], x, (size_t)B * m->width[0] * sizeof(float), hipMemcpyHostToDevice, m->st));
  forward(m, B);
  MCHK(hipGetLastError());
  MCHK(hipMemcpyAsync(y, m->act[m->nl], (size_t)B * m->width[m->nl] * sizeof(float), hipMemcpyDeviceToHost,
                      m->st));
  MCHK(hipStreamSynchronize(m->st));
  return SHX_OK;
}

int shx_mlp_forward_noalloc(shx_mlp* m, int B, const float* x_dev, float* y_dev)
{
  if (!m || B <= 0 || !x_dev || !y_dev) return SHX_E_ARG;
  if (!m->scratch) { m->err = "no scratch address for this device"; return SHX_E_STATE; }
  int wh = 0;                                   /* widest hidden layer */
  for (int l = 1; l < m->nl; ++l) wh = std::max(wh, m->width[l]);
  if ((size_t)2 * B * wh > kNoAllocFloats) {
    m->err = "shx_mlp_forward_noalloc: batch " + std::to_string(B) + " exceeds the scratch limit of " +
             std::to_string(wh ? kNoAllocFloats / (2 * (size_t)wh) : 0) + " rows for hidden width " +
             std::to_string(wh);
    return SHX_E_ARG;
  }
  MCHK(hipSetDevice(m->dev));
  float* half[2] = {m->scratch, m->scratch + kNoAllocFloats / 2};
  const float* in = x_dev;
  for (int l = 0; l < m->nl; ++l) {
    const int K = m->width[l], N = m->width[l + 1];
    const bool last = l == m->nl - 1;
    float* out = last ? y_dev : half[l & 1];
    gemm(m->st, B, N, K, in, K, 0, m->p + m->woff[l], N, 0, m->p + m->boff[l], out, N, nullptr, last ? 0 : 1);
    in = out;
  }
  MCHK(hipGetLastError());
  MCHK(hipStreamSynchronize(m->st));
  return SHX_OK;
}

int shx_mlp_gradient(shx_mlp* m, int B, const float* x, const float* t, float* grad, double* loss)
{
  if (!m || B <= 0 || !x || !t) return SHX_E_ARG;
  if (int r = ensure_batch(m, B)) return r;
  MCHK(hipMemcpyAsync(m->act[0], x, (size_t)B * m->width[0] * sizeof(float), hipMemcpyHostToDevice, m->st));
  MCHK(hipMemcpyAsync(m->tgt, t, (size_t)B * m->width[m->nl] * sizeof(float), hipMemcpyHostToDevice, m->st));
  forward(m, B);
  if (int r = backward(m, B, loss)) return r;
  MCHK(hipGetLastError());
  if (grad) MCHK(hipMemcpy(grad, m->g, m->nparam * sizeof(float), hipMemcpyDeviceToHost));
  return SHX_OK;
}

int shx_mlp_train(shx_mlp* m, int n, const float* x, const float* t, int batch, int epochs, float lr,
                  uint64_t seed, double* loss_out)
{
  if (!m || n <= 0 || !x || !t || batch <= 0 || epochs <= 0) return SHX_E_ARG;
  if (batch > n) {
    m->err = "minibatch larger than the training set (batch " + std::to_string(batch) + " > n " +
             std::to_string(n) + "): no optimizer step would run";
    return SHX_E_ARG;
  }
  const int nin = m->width[0], nout = m->width[m->nl];
  if (int r = ensure_batch(m, batch)) return r;
  float *dx = nullptr, *dt = nullptr;
  int* didx = nullptr;
  MCHK(hipMalloc(&dx, (size_t)n * nin * sizeof(float)));
  MCHK(hipMalloc(&dt, (size_t)n * nout * sizeof(float)));
  MCHK(hipMalloc(&didx, (size_t)n * sizeof(int)));
  MCHK(hipMemcpy(dx, x, (size_t)n * nin * sizeof(float), hipMemcpyHostToDevice));
  MCHK(hipMemcpy(dt, t, (size_t)n * nout * sizeof(float), hipMemcpyHostToDevice));
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  uint64_t s = seed ? seed : 1;
  double last = 0;
  int rc = SHX_OK;
  for (int ep = 0; ep < epochs && rc == SHX_OK; ++ep) {
    for (int i = n - 1; i > 0; --i) {   /* Fisher-Yates shuffle */
      s ^= s << 13; s ^= s >> 7; s ^= s << 17;
      std::swap(order[i], order[s % (uint64_t)(i + 1)]);
    }
    double esum = 0;
    int nbat = 0;
    if (hipMemcpyAsync(didx, order.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, m->st) != hipSuccess) {
      rc = SHX_E_HIP;
      break;
    }
    for (int b0 = 0; b0 + batch <= n; b0 += batch) {
      k_gather_rows<<<(batch * nin + 255) / 256, 256, 0, m->st>>>(batch, nin, didx + b0, dx, m->act[0]);
      k_gather_rows<<<(batch * nout + 255) / 256, 256, 0, m->st>>>(batch, nout, didx + b0, dt, m->tgt);
      forward(m, batch);
      double l = 0;
      if ((rc = backward(m, batch, &l))) break;
      ++m->step;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
      k_adam<<<(int)((m->nparam + 255) / 256), 256, 0, m->st>>>((int)m->nparam, m->p, m->g, m->m1, m->m2, lr, b1,
                                                                 b2, 1e-8f, c1, c2);
      esum += l;
      ++nbat;
    }
    last = nbat ? esum / nbat : 0;
  }
  (void)hipStreamSynchronize(m->st);
  (void)hipFree(dx);
  (void)hipFree(dt);
  (void)hipFree(didx);
  if (rc != SHX_OK) { m->err = "training step failed"; return rc; }
  MCHK(hipGetLastError());
  MCHK(hipStreamSynchronize(m->st));
  if (loss_out) *loss_out = last;
  return SHX_OK;
}

int shx_mlp_gemm(int M, int N, int K, const float* A, int ta, const float* B, int tb, float* C)
{
  /* matrix-core GEMM on host arrays (layout checks): C = op(A) op(B), row major */
  float *dA = nullptr, *dB = nullptr, *dC = nullptr;
  if (hipMalloc(&dA, (size_t)M * K * sizeof(float)) != hipSuccess ||
      hipMalloc(&dB, (size_t)K * N * sizeof(float)) != hipSuccess ||
      hipMalloc(&dC, (size_t)M * N * sizeof(float)) != hipSuccess)
    return SHX_E_HIP;
  int rc = SHX_OK;
  if (hipMemcpy(dA, A, (size_t)M * K * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dB, B, (size_t)K * N * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    rc = SHX_E_HIP;
  if (rc == SHX_OK) {
    gemm(nullptr, M, N, K, dA, ta ? M : K, ta, dB, tb ? K : N, tb, nullptr, dC, N, nullptr, 0);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(C, dC, (size_t)M * N * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      rc = SHX_E_HIP;
  }
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  return rc;
}

}  // extern "C"
