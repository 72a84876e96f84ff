// Learned Newton predictor network: an MLP with BF16 weights and activations
// and FP32 accumulation on the CDNA matrix cores, with hand-written forward,
// backward (mean masked squared error) and Adam kernels.  It has no
// counterpart in the reference: it proposes the Newton initial guess that
// OneStep::obtainPredictor (src/TimeIntegrationPKG/N_TIA_OneStep.C:135)
// computes by polynomial extrapolation, which stays the fallback.
//
// Layout (all row-major, rows = samples):
//   parameters  per layer l: W_l [out_l][in_l] then b_l [out_l]; one FP32
//               master vector, one BF16 copy (the GEMM operand), FP32
//               gradients and Adam moments of the same layout
//   activations X [rows][in] bf16; Z_l [rows][width] fp32 (pre-activation,
//               kept for the backward pass); H_l [rows][width] bf16
//   output      Y [rows][out] fp32 (last layer linear)
// Every dimension is a multiple of 64 (rows, widths) so the 64x64 tiles of
// the GEMM need no edge handling; the host checks this before each launch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "xg_internal.h"

#define HIPCHK(x)                          \
  do {                                     \
    hipError_t e_ = (x);                   \
    if (e_ != hipSuccess) return XG_E_HIP; \
  } while (0)

namespace xgm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  // round to nearest even; NaN stays NaN
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// GELU, erf form: z * Phi(z); derivative Phi(z) + z * phi(z)
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum Epi { EPI_F32 = 0, EPI_BIAS_F32 = 1, EPI_BIAS_GELU = 2, EPI_GELU_BWD = 3 };

struct GemmArgs {
  int M, N, K;
  const unsigned short* A; int lda;   // A(m,k) = TA ? A[k*lda+m] : A[m*lda+k]
  const unsigned short* B; int ldb;   // B(n,k) = TB ? B[k*ldb+n] : B[n*ldb+k]
  float* C; int ldc;                  // EPI_F32 / EPI_BIAS_F32 output; EPI_BIAS_GELU: Z
  const float* bias;
  unsigned short* H;                  // EPI_BIAS_GELU: gelu(Z) bf16; EPI_GELU_BWD: dZ bf16
  const float* Zin;                   // EPI_GELU_BWD: pre-activation of this layer
};

constexpr int BM = 64, BN = 64, BK = 32, LDS_K = BK + 8;

// C = A B^T on 64x64 tiles: 4 waves, each a 32x32 block of
// v_mfma_f32_32x32x16_bf16 (A: lane l holds row l%32, k = 8*(l/32)..+7;
// B likewise by column; D register r of lane l is row 8*(r/4) + 4*(l/32) +
// r%4, column l%32).
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) unsigned short As[BM * LDS_K];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[BN * LDS_K];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  f32x16 acc = {};
  for (int k0 = 0; k0 < g.K; k0 += BK) {
    // ---- global -> LDS (16-byte loads)
    if (!TA) {
      const int r = tid >> 2, kc = (tid & 3) * 8;
      const uint4 v = *reinterpret_cast<const uint4*>(g.A + (size_t)(m0 + r) * g.lda + k0 + kc);
      *reinterpret_cast<uint4*>(&As[r * LDS_K + kc]) = v;
    } else {
      const int kr = tid >> 3, mc = (tid & 7) * 8;
      const uint4 v = *reinterpret_cast<const uint4*>(g.A + (size_t)(k0 + kr) * g.lda + m0 + mc);
      const unsigned short* e = reinterpret_cast<const unsigned short*>(&v);
      for (int j = 0; j < 8; ++j) As[(mc + j) * LDS_K + kr] = e[j];
    }
    if (!TB) {
      const int r = tid >> 2, kc = (tid & 3) * 8;
      const uint4 v = *reinterpret_cast<const uint4*>(g.B + (size_t)(n0 + r) * g.ldb + k0 + kc);
      *reinterpret_cast<uint4*>(&Bs[r * LDS_K + kc]) = v;
    } else {
      const int kr = tid >> 3, nc = (tid & 7) * 8;
      const uint4 v = *reinterpret_cast<const uint4*>(g.B + (size_t)(k0 + kr) * g.ldb + n0 + nc);
      const unsigned short* e = reinterpret_cast<const unsigned short*>(&v);
      for (int j = 0; j < 8; ++j) Bs[(nc + j) * LDS_K + kr] = e[j];
    }
    __syncthreads();
    // ---- matrix cores
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[(wm * 32 + (lane & 31)) * LDS_K + kk + 8 * (lane >> 5)]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[(wn * 32 + (lane & 31)) * LDS_K + kk + 8 * (lane >> 5)]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // ---- epilogue
  const int col = n0 + wn * 32 + (lane & 31);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const size_t o = (size_t)row * g.ldc + col;
    const float v = acc[r];
    if (EPI == EPI_F32) {
      g.C[o] = v;
    } else if (EPI == EPI_BIAS_F32) {
      g.C[o] = v + g.bias[col];
    } else if (EPI == EPI_BIAS_GELU) {
      const float z = v + g.bias[col];
      g.C[o] = z;
      g.H[o] = f2bf(gelu(z));
    } else {
      g.H[o] = f2bf(v * gelu_d(g.Zin[o]));
    }
  }
}

// loss = sum_rows sum_j mask (y - t)^2 / count, dY = 2 mask (y - t) / count
// (bf16), per-block partial sums of the loss (deterministic two-pass sum)
__global__ void k_mse(int n, int count, const float* y, const float* t, const float* mask, unsigned short* dy,
                      double* partial) {
  __shared__ double red[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (i < n) {
    const float d = (y[i] - t[i]) * mask[i];
    l = (double)d * d;
    dy[i] = f2bf(2.0f * d / (float)count);
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
__global__ void k_sum(int n, const double* p, double* out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += p[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// bias gradient: column sums of a bf16 [rows][N] matrix, two passes
// (chunks of 256 rows, then over chunks)
__global__ void k_colsum_part(int rows, int N, const unsigned short* d, float* part) {
  const int col = blockIdx.x * 64 + (threadIdx.x & 63);
  const int r0 = blockIdx.y * 256 + (threadIdx.x >> 6) * 64;
  float s = 0.0f;
  for (int r = r0; r < r0 + 64 && r < rows; ++r) s += bf2f(d[(size_t)r * N + col]);
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < 64) {
    const float t = red[threadIdx.x] + red[threadIdx.x + 64] + red[threadIdx.x + 128] + red[threadIdx.x + 192];
    part[(size_t)blockIdx.y * N + col] = t;
  }
}
__global__ void k_colsum_fin(int chunks, int N, const float* part, float* out) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= N) return;
  float s = 0.0f;
  for (int c = 0; c < chunks; ++c) s += part[(size_t)c * N + col];
  out[col] = s;
}

// Adam on the FP32 master parameters, refreshing the BF16 copy
__global__ void k_adam(long long n, float* w, unsigned short* wb, const float* g, float* m, float* v, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  wb[i] = f2bf(wi);
}
__global__ void k_to_bf16(long long n, const float* w, unsigned short* wb) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = f2bf(w[i]);
}

inline unsigned nblk(long long n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

template <class T>
struct Buf {
  T* p = nullptr;
  ~Buf() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t n) {
    if (p) { (void)hipFree(p); p = nullptr; }
    return n ? hipMalloc((void**)&p, n * sizeof(T)) : hipSuccess;
  }
};

}  // namespace xgm

using namespace xgm;

struct xg_mlp {
  int in_dim = 0, width = 0, hidden = 0, out_dim = 0, max_rows = 0;
  int n_layers = 0;                       // hidden + 1
  std::vector<int> din, dout;             // per layer
  std::vector<long long> woff, boff;      // parameter offsets
  long long n_params = 0;
  long long adam_t = 0;
  hipStream_t stream = nullptr;
  Buf<float> w, g, m, v, Y, part, Zall;
  Buf<unsigned short> wb, Hall, dH, dY;
  Buf<double> lpart, lsum;
  // Z_l / H_l at offset l * max_rows * width
  float* Z(int l) { return Zall.p + (size_t)l * max_rows * width; }
  unsigned short* H(int l) { return Hall.p + (size_t)l * max_rows * width; }
  ~xg_mlp() { if (stream) (void)hipStreamDestroy(stream); }

  template <bool TA, bool TB, int EPI>
  int gemm(const GemmArgs& a) {
    if (a.M % BM || a.N % BN || a.K % BK || a.M <= 0 || a.N <= 0 || a.K <= 0) return XG_E_ARG;
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), dim3(a.N / BN, a.M / BM), dim3(256), 0, stream, a);
    HIPCHK(hipGetLastError());
    return XG_OK;
  }

  // forward of `rows` samples; keeps Z/H for the backward pass; Y fp32
  int forward(int rows, const unsigned short* X, float* Yout) { return forward_on(rows, X, Yout, stream); }
  int forward_on(int rows, const unsigned short* X, float* Yout, hipStream_t s) {
    hipStream_t keep = stream;
    stream = s;
    const int st = forward_layers(rows, X, Yout);
    stream = keep;
    return st;
  }
  int forward_layers(int rows, const unsigned short* X, float* Yout) {
    const unsigned short* in = X;
    for (int l = 0; l < n_layers; ++l) {
      GemmArgs a{};
      a.M = rows; a.N = dout[l]; a.K = din[l];
      a.A = in; a.lda = din[l];
      a.B = wb.p + woff[l]; a.ldb = din[l];
      a.bias = w.p + boff[l];
      int st;
      if (l + 1 < n_layers) {
        a.C = Z(l); a.ldc = width; a.H = H(l);
        st = gemm<false, false, EPI_BIAS_GELU>(a);
        in = H(l);
      } else {
        a.C = Yout; a.ldc = out_dim;
        st = gemm<false, false, EPI_BIAS_F32>(a);
      }
      if (st != XG_OK) return st;
    }
    return XG_OK;
  }

  int colsum(int rows, int N, const unsigned short* d, float* out) {
    const int chunks = (rows + 255) / 256;
    hipLaunchKernelGGL(k_colsum_part, dim3(N / 64, chunks), dim3(256), 0, stream, rows, N, d, part.p);
    hipLaunchKernelGGL(k_colsum_fin, dim3(nblk(N)), dim3(256), 0, stream, chunks, N, part.p, out);
    HIPCHK(hipGetLastError());
    return XG_OK;
  }

  // backward of the loss whose output gradient is in dY (bf16 [rows][out])
  int backward(int rows, const unsigned short* X) {
    const unsigned short* dcur = dY.p;   // gradient w.r.t. this layer's output (pre-activation)
    for (int l = n_layers - 1; l >= 0; --l) {
      const unsigned short* in = l == 0 ? X : H(l - 1);
      // dW_l = dcur^T in : C[dout][din] = sum_m dcur[m][o] in[m][i]
      GemmArgs a{};
      a.M = dout[l]; a.N = din[l]; a.K = rows;
      a.A = dcur; a.lda = dout[l];
      a.B = in; a.ldb = din[l];
      a.C = g.p + woff[l]; a.ldc = din[l];
      int st = gemm<true, true, EPI_F32>(a);
      if (st != XG_OK) return st;
      st = colsum(rows, dout[l], dcur, g.p + boff[l]);
      if (st != XG_OK) return st;
      if (l == 0) break;
      // d(pre-activation of layer l-1) = (dcur W_l) * gelu'(Z_{l-1})
      GemmArgs b{};
      b.M = rows; b.N = din[l]; b.K = dout[l];
      b.A = dcur; b.lda = dout[l];
      b.B = wb.p + woff[l]; b.ldb = din[l];
      unsigned short* dst = dH.p + (size_t)((l - 1) & 1) * max_rows * width;
      b.H = dst; b.ldc = width; b.Zin = Z(l - 1);
      st = gemm<false, true, EPI_GELU_BWD>(b);
      if (st != XG_OK) return st;
      dcur = dst;
    }
    return XG_OK;
  }
};

namespace xg {
int mlp_forward_on_stream(xg_mlp* net, int rows, const unsigned short* x, float* y, void* stream) {
  if (!net || rows <= 0 || rows > net->max_rows || rows % 64) return XG_E_ARG;
  return net->forward_on(rows, x, y, (hipStream_t)stream);
}
int mlp_dims(const xg_mlp* net, int* in_dim, int* out_dim, int* max_rows) {
  if (!net) return XG_E_ARG;
  *in_dim = net->in_dim; *out_dim = net->out_dim; *max_rows = net->max_rows;
  return XG_OK;
}
}  // namespace xg

extern "C" {

int xg_mlp_create(int in_dim, int width, int hidden_layers, int out_dim, int max_rows, unsigned long long seed,
                  xg_mlp** out) {
  if (!out || in_dim <= 0 || width <= 0 || out_dim <= 0 || hidden_layers < 1 || max_rows <= 0) return XG_E_ARG;
  if (in_dim % 64 || width % 64 || out_dim % 64 || max_rows % 64) return XG_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return XG_E_NOGPU;
  auto net = std::make_unique<xg_mlp>();
  net->in_dim = in_dim; net->width = width; net->hidden = hidden_layers; net->out_dim = out_dim;
  net->max_rows = max_rows;
  net->n_layers = hidden_layers + 1;
  long long off = 0;
  for (int l = 0; l < net->n_layers; ++l) {
    const int di = l == 0 ? in_dim : width, dn = l + 1 < net->n_layers ? width : out_dim;
    net->din.push_back(di); net->dout.push_back(dn);
    net->woff.push_back(off); off += (long long)di * dn;
    net->boff.push_back(off); off += dn;
  }
  net->n_params = off;
  // uniform(-1/sqrt(in), 1/sqrt(in)) weights, zero biases; the output layer
  // scaled by 0.1 so an untrained net proposes almost no correction
  std::vector<float> h(off, 0.0f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < net->n_layers; ++l) {
    const float a = 1.0f / std::sqrt((float)net->din[l]) * (l + 1 == net->n_layers ? 0.1f : 1.0f);
    std::uniform_real_distribution<float> u(-a, a);
    for (long long i = 0; i < (long long)net->din[l] * net->dout[l]; ++i) h[net->woff[l] + i] = u(rng);
  }
  if (hipStreamCreate(&net->stream) != hipSuccess) return XG_E_HIP;
  HIPCHK(net->w.alloc(off)); HIPCHK(net->g.alloc(off)); HIPCHK(net->m.alloc(off)); HIPCHK(net->v.alloc(off));
  HIPCHK(net->wb.alloc(off));
  HIPCHK(net->Zall.alloc((size_t)hidden_layers * max_rows * width));
  HIPCHK(net->Hall.alloc((size_t)hidden_layers * max_rows * width));
  HIPCHK(net->dH.alloc((size_t)2 * max_rows * width));
  HIPCHK(net->dY.alloc((size_t)max_rows * out_dim));
  HIPCHK(net->Y.alloc((size_t)max_rows * out_dim));
  const int maxw = std::max(width, std::max(in_dim, out_dim));
  HIPCHK(net->part.alloc((size_t)((max_rows + 255) / 256) * maxw));
  HIPCHK(net->lpart.alloc(nblk((long long)max_rows * out_dim)));
  HIPCHK(net->lsum.alloc(1));
  HIPCHK(hipMemcpy(net->w.p, h.data(), off * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(net->m.p, 0, off * sizeof(float)));
  HIPCHK(hipMemset(net->v.p, 0, off * sizeof(float)));
  HIPCHK(hipMemset(net->g.p, 0, off * sizeof(float)));
  hipLaunchKernelGGL(k_to_bf16, dim3(nblk(off)), dim3(256), 0, net->stream, off, net->w.p, net->wb.p);
  HIPCHK(hipStreamSynchronize(net->stream));
  *out = net.release();
  return XG_OK;
}

int xg_mlp_destroy(xg_mlp* net) {
  if (!net) return XG_E_ARG;
  delete net;
  return XG_OK;
}

int xg_mlp_params(xg_mlp* net, int set, float* params, long long* count) {
  if (!net) return XG_E_ARG;
  if (count) *count = net->n_params;
  if (!params) return XG_OK;
  if (set) {
    HIPCHK(hipMemcpyAsync(net->w.p, params, net->n_params * sizeof(float), hipMemcpyHostToDevice, net->stream));
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(net->n_params)), dim3(256), 0, net->stream, net->n_params, net->w.p,
                       net->wb.p);
    HIPCHK(hipMemsetAsync(net->m.p, 0, net->n_params * sizeof(float), net->stream));
    HIPCHK(hipMemsetAsync(net->v.p, 0, net->n_params * sizeof(float), net->stream));
    net->adam_t = 0;
  } else {
    HIPCHK(hipMemcpyAsync(params, net->w.p, net->n_params * sizeof(float), hipMemcpyDeviceToHost, net->stream));
  }
  HIPCHK(hipStreamSynchronize(net->stream));
  return XG_OK;
}

int xg_mlp_forward(xg_mlp* net, int rows, const unsigned short* x_dev, float* y_dev) {
  if (!net || !x_dev || !y_dev || rows <= 0 || rows > net->max_rows || rows % 64) return XG_E_ARG;
  int st = net->forward(rows, x_dev, y_dev);
  if (st != XG_OK) return st;
  HIPCHK(hipStreamSynchronize(net->stream));
  return XG_OK;
}

int xg_mlp_loss_backward(xg_mlp* net, int rows, const unsigned short* x_dev, const float* t_dev,
                         const float* mask_dev, int count, double* loss, float* grads) {
  if (!net || !x_dev || !t_dev || !mask_dev || rows <= 0 || rows > net->max_rows || rows % 64 || count <= 0)
    return XG_E_ARG;
  int st = net->forward(rows, x_dev, net->Y.p);
  if (st != XG_OK) return st;
  const long long n = (long long)rows * net->out_dim;
  hipLaunchKernelGGL(k_mse, dim3(nblk(n)), dim3(256), 0, net->stream, (int)n, count, net->Y.p, t_dev, mask_dev,
                     net->dY.p, net->lpart.p);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, net->stream, (int)nblk(n), net->lpart.p, net->lsum.p);
  HIPCHK(hipGetLastError());
  st = net->backward(rows, x_dev);
  if (st != XG_OK) return st;
  double l = 0.0;
  HIPCHK(hipMemcpyAsync(&l, net->lsum.p, sizeof(double), hipMemcpyDeviceToHost, net->stream));
  if (grads) HIPCHK(hipMemcpyAsync(grads, net->g.p, net->n_params * sizeof(float), hipMemcpyDeviceToHost, net->stream));
  HIPCHK(hipStreamSynchronize(net->stream));
  if (loss) *loss = l / count;
  return XG_OK;
}

int xg_mlp_adam(xg_mlp* net, float lr, float beta1, float beta2, float eps) {
  if (!net || !(lr > 0.0f)) return XG_E_ARG;
  net->adam_t += 1;
  const float c1 = 1.0f - std::pow(beta1, (float)net->adam_t), c2 = 1.0f - std::pow(beta2, (float)net->adam_t);
  hipLaunchKernelGGL(k_adam, dim3(nblk(net->n_params)), dim3(256), 0, net->stream, net->n_params, net->w.p,
                     net->wb.p, net->g.p, net->m.p, net->v.p, lr, beta1, beta2, eps, c1, c2);
  HIPCHK(hipGetLastError());
  return XG_OK;
}

int xg_mlp_sync(xg_mlp* net) {
  if (!net) return XG_E_ARG;
  HIPCHK(hipStreamSynchronize(net->stream));
  return XG_OK;
}

}  // extern "C"
