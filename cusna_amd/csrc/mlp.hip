// mlp.hip -- the surrogate perceptron on gfx950 BF16 matrix cores.
//
// Every product is one kernel: a 256-thread workgroup computes a 64 x 64
// block of C = A B (A row-major M x K, B given as B^T row-major N x K, both
// bfloat16), each wave a 16 x 64 strip as four 16 x 16 MFMA tiles; per
// 32-wide K step a lane loads 8 contiguous bfloat16 of its A row and of each
// of its four B columns (16-byte loads) and issues four
// v_mfma_f32_16x16x32_bf16. Operand layout of that instruction (checked on
// the device): lane l holds A[l % 16][8 (l / 16) + e] and B[8 (l / 16) + e][l % 16],
// e = 0..7, and receives D[4 (l / 16) + v][l % 16], v = 0..3. The epilogue
// adds the bias, keeps the float32 pre-activation, applies SiLU (forward) or
// multiplies by SiLU'(z) (backward), and writes bfloat16 row-major and/or
// transposed copies, so that every later product reads contiguous K.
//
// Latency, not throughput, bounds these products (a 1024 x 256 x 256 product
// is ~0.13 GFLOP): the K loop is unrolled so that a wave's loads for U K
// steps are all in flight before its MFMAs, and products with a long K and a
// plain float32 result (the weight gradients, K = batch) are split over K
// into partial blocks that a second kernel sums in split order
// (deterministic). The bfloat16 outputs pass through LDS so that the
// row-major and transposed copies are written as 16-byte stores.
#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

namespace cusna_impl {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kMlpThreads = 256;

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

struct GemmArgs {
  const __bf16* A;   // M x K
  const __bf16* Bt;  // N x K
  int M, N, K;
  const float* bias;  // N or null
  const float* dz;    // M x N pre-activations: multiply by SiLU'(dz), or null
  int act;            // 1: SiLU (forward)
  float* Cf;          // M x N float32 result before the activation, or null
  __bf16* Cb;         // M x N bfloat16 result, or null
  __bf16* CbT;        // N x M bfloat16 result transposed, or null
  int Ks;             // K per split (blockIdx.z); Ks = K without a split
  float* part;        // split products: partial blocks [gridDim.z][M][N]
};

// U: K steps of 32 whose loads are issued together (Ks a multiple of 32 U).
template <int U>
__global__ void __launch_bounds__(kMlpThreads) gemm_bf16_kernel(const GemmArgs g) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m0 = blockIdx.x * 64 + 16 * w, n0 = blockIdx.y * 64;
  const int r = l & 15, q = l >> 4;
  floatx4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  const __bf16* arow = g.A + (size_t)(m0 + r) * g.K + 8 * q;
  const __bf16* brow = g.Bt + (size_t)(n0 + r) * g.K + 8 * q;
  const int kbeg = blockIdx.z * g.Ks, kend = kbeg + g.Ks;
  for (int k0 = kbeg; k0 < kend; k0 += 32 * U) {
    bf16x8 a[U], b[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      a[u] = *reinterpret_cast<const bf16x8*>(arow + k0 + 32 * u);
#pragma unroll
      for (int t = 0; t < 4; ++t)
        b[u][t] = *reinterpret_cast<const bf16x8*>(brow + (size_t)16 * t * g.K + k0 + 32 * u);
    }
    // every load of the U steps is issued before the first MFMA (the default
    // schedule keeps only ~7 loads in flight to save registers)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int t = 0; t < 4; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[u], b[u][t], acc[t], 0, 0, 0);
    }
  }
  const int mt = blockIdx.x * 64;  // the workgroup's 64 x 64 tile: rows mt.., columns n0..
  if (gridDim.z > 1) {  // partial block of split blockIdx.z
    const size_t MN = (size_t)g.M * g.N;
    float* part = g.part + (size_t)blockIdx.z * MN;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int v = 0; v < 4; ++v)
        part[(size_t)(m0 + 4 * q + v) * g.N + n0 + 16 * t + r] = acc[t][v];
    }
    return;
  }
  __shared__ __align__(16) __bf16 tile[64][72];  // bfloat16 outputs, rows padded
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int col = n0 + 16 * t + r;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = m0 + 4 * q + v;
      const size_t ix = (size_t)row * g.N + col;
      float val = acc[t][v];
      if (g.bias) val += g.bias[col];
      if (g.Cf) g.Cf[ix] = val;
      if (g.dz) val *= dsilu(g.dz[ix]);
      else if (g.act) val = silu(val);
      tile[row - mt][col - n0] = (__bf16)val;
    }
  }
  if (!g.Cb && !g.CbT) return;
  __syncthreads();
  const int a = threadIdx.x >> 2, c16 = 16 * (threadIdx.x & 3);
  if (g.Cb) {  // row a of the tile, columns c16..c16+15
    const bf16x8* src = reinterpret_cast<const bf16x8*>(&tile[a][c16]);
    bf16x8* dst = reinterpret_cast<bf16x8*>(g.Cb + (size_t)(mt + a) * g.N + n0 + c16);
    dst[0] = src[0];
    dst[1] = src[1];
  }
  if (g.CbT) {  // column a of the tile, rows c16..c16+15, as a row of the transpose
    bf16x8 lo, hi;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      lo[e] = tile[c16 + e][a];
      hi[e] = tile[c16 + 8 + e][a];
    }
    bf16x8* dst = reinterpret_cast<bf16x8*>(g.CbT + (size_t)(n0 + a) * g.M + mt + c16);
    dst[0] = lo;
    dst[1] = hi;
  }
}

// out[i] = sum over s = 0..S-1 of part[s][i], in that order.
__global__ void sum_parts_kernel(const float* part, int S, int MN, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MN) return;
  float v = part[i];
  for (int s = 1; s < S; ++s) v += part[(size_t)s * MN + i];
  out[i] = v;
}

// x (B x p_in, float32) -> X (B x Kin) and X^T (Kin x B), bfloat16, zero padded.
__global__ void prep_input_kernel(const float* x, int B, int p_in, int Kin, __bf16* X,
                                  __bf16* XT) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * Kin) return;
  const int b = i / Kin, k = i - b * Kin;
  const __bf16 v = (__bf16)(k < p_in ? x[(size_t)b * p_in + k] : 0.0f);
  X[i] = v;
  XT[(size_t)k * B + b] = v;
}

// Gaussian negative log-likelihood per sample and its gradient with respect
// to the outputs (mu in columns 0..p-1, log-variance in p..2p-1, clamped to
// [-12, 12] with zero gradient outside), scaled by 1/B.
__global__ void nll_kernel(const float* O, const float* Y, int B, int p, int Nout, float* loss,
                           __bf16* dO, __bf16* dOT) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float invb = 1.0f / (float)B;
  float L = 0.0f;
  for (int a = 0; a < Nout; ++a) {
    float gv = 0.0f;
    if (a < p) {
      const float mu = O[(size_t)b * Nout + a];
      const float lvr = O[(size_t)b * Nout + p + a];
      const float lv = fminf(fmaxf(lvr, -12.0f), 12.0f);
      const float e = expf(-lv);
      const float res = Y[(size_t)b * p + a] - mu;
      L += 0.5f * (res * res * e + lv);
      gv = -res * e * invb;
    } else if (a < 2 * p) {
      const int c = a - p;
      const float mu = O[(size_t)b * Nout + c];
      const float lvr = O[(size_t)b * Nout + a];
      const float lv = fminf(fmaxf(lvr, -12.0f), 12.0f);
      const float res = Y[(size_t)b * p + c] - mu;
      gv = (lvr > -12.0f && lvr < 12.0f) ? 0.5f * (1.0f - res * res * expf(-lv)) * invb : 0.0f;
    }
    const __bf16 hv = (__bf16)gv;
    dO[(size_t)b * Nout + a] = hv;
    dOT[(size_t)a * B + b] = hv;
  }
  loss[b] = L;
}

// out[n] = sum_b XT[n][b] (bias gradients from the transposed bfloat16
// gradients), one workgroup per n, fixed tree order.
__global__ void __launch_bounds__(kMlpThreads) rowsum_kernel(const __bf16* XT, int B, float* out) {
  __shared__ float sh[kMlpThreads];
  const int n = blockIdx.x;
  float v = 0.0f;
  for (int b = threadIdx.x; b < B; b += kMlpThreads) v += (float)XT[(size_t)n * B + b];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = kMlpThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = sh[0];
}

// Where the weights of each layer sit in the flat parameter vector and their
// bfloat16 copies W [K][N] and W^T [N][K].
struct CastMap {
  size_t off[kMlpLayers];
  int K[kMlpLayers], N[kMlpLayers];
  __bf16* Wb[kMlpLayers];
  __bf16* Wtb[kMlpLayers];
};

// Adam (bias-corrected): w -= lr * m_hat / (sqrt(v_hat) + eps); a weight's
// new value is also written to its two bfloat16 copies (as cast_kernel).
__global__ void adam_kernel(float* w, float* m, float* v, const float* g, size_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2, const CastMap cm) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  w[i] = wi;
#pragma unroll
  for (int l = 0; l < kMlpLayers; ++l) {
    const size_t e = i - cm.off[l];
    if (i >= cm.off[l] && e < (size_t)cm.K[l] * cm.N[l]) {
      const int k = (int)(e / cm.N[l]), c = (int)(e - (size_t)k * cm.N[l]);
      const __bf16 hv = (__bf16)wi;
      cm.Wb[l][e] = hv;
      cm.Wtb[l][(size_t)c * cm.K[l] + k] = hv;
    }
  }
}

// Master weights W [K][N] -> bfloat16 W [K][N] and W^T [N][K].
__global__ void cast_kernel(const float* W, int K, int N, __bf16* Wb, __bf16* Wtb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K * N) return;
  const int k = i / N, n = i - k * N;
  const __bf16 v = (__bf16)W[i];
  Wb[i] = v;
  Wtb[(size_t)n * K + k] = v;
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xFF51AFD7ED558CCDull;
  x ^= x >> 33;
  x *= 0xC4CEB9FE1A85EC53ull;
  return x ^ (x >> 33);
}

// Minibatch by sampling with replacement: row b of step s is sample
// mix64(seed + s * B + b) mod n.
__global__ void gather_kernel(const float* x_all, const float* y_all, long long n, int p_in,
                              int p_out, int B, uint64_t seed, long long step, float* xb,
                              float* yb) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const uint64_t h = mix64(seed + (uint64_t)step * (uint64_t)B + (uint64_t)b);
  const long long idx = (long long)(h % (uint64_t)n);
  for (int a = 0; a < p_in; ++a) xb[(size_t)b * p_in + a] = x_all[(size_t)idx * p_in + a];
  for (int a = 0; a < p_out; ++a) yb[(size_t)b * p_out + a] = y_all[(size_t)idx * p_out + a];
}

int pad64(int x) { return (x + 63) / 64 * 64; }

}  // namespace

cusna_status Mlp::init(int p_in, int p_out, int width, int batch, uint64_t seed, hipStream_t s) {
  if (p_in < 1 || p_in > 64 || p_out < 1 || p_out > 64 || width < 64 || width % 64 ||
      width > 1024 || batch < 64 || batch % 64) {
    set_error("mlp: p_in, p_out in 1..64, width and batch multiples of 64");
    return CUSNA_ERR_SHAPE;
  }
  p_in_ = p_in;
  p_out_ = p_out;
  width_ = width;
  B_ = batch;
  s_ = s;
  K_[0] = pad64(p_in);
  N_[0] = width;
  K_[1] = N_[1] = K_[2] = N_[2] = width;
  K_[3] = width;
  N_[3] = pad64(2 * p_out);
  size_t o = 0;
  for (int l = 0; l < kMlpLayers; ++l) {
    off_w_[l] = o;
    o += (size_t)K_[l] * N_[l];
    off_b_[l] = o;
    o += (size_t)N_[l];
  }
  n_params_ = o;
  const size_t P4 = n_params_ * 4;
  const size_t o_p = D_.add(P4), o_m = D_.add(P4), o_v = D_.add(P4), o_g = D_.add(P4);
  size_t o_wb[kMlpLayers], o_wtb[kMlpLayers];
  for (int l = 0; l < kMlpLayers; ++l) {
    o_wb[l] = D_.add((size_t)K_[l] * N_[l] * 2);
    o_wtb[l] = D_.add((size_t)K_[l] * N_[l] * 2);
  }
  const int B = B_;
  const size_t o_x = D_.add((size_t)B * K_[0] * 2), o_xt = D_.add((size_t)B * K_[0] * 2);
  size_t o_z[3], o_h[3], o_ht[3], o_dz[3], o_dzt[3];
  for (int l = 0; l < 3; ++l) {
    o_z[l] = D_.add((size_t)B * N_[l] * 4);
    o_h[l] = D_.add((size_t)B * N_[l] * 2);
    o_ht[l] = D_.add((size_t)B * N_[l] * 2);
    o_dz[l] = D_.add((size_t)B * N_[l] * 2);
    o_dzt[l] = D_.add((size_t)B * N_[l] * 2);
  }
  const size_t o_o = D_.add((size_t)B * N_[3] * 4), o_do = D_.add((size_t)B * N_[3] * 2);
  const size_t o_dot = D_.add((size_t)B * N_[3] * 2), o_loss = D_.add((size_t)B * 4);
  const size_t o_yb = D_.add((size_t)B * p_out * 4), o_xb = D_.add((size_t)B * p_in * 4);
  // partial blocks of the split weight-gradient products (K = B)
  part_floats_ = 0;
  if (B >= 512 && B % 128 == 0)
    for (int l = 0; l < kMlpLayers; ++l)
      part_floats_ = std::max(part_floats_, (size_t)(B / 128) * K_[l] * N_[l]);
  const size_t o_part = D_.add(part_floats_ * 4);
  HIP_TRY(D_.alloc());
  part_ = D_.at<float>(o_part);
  params_ = D_.at<float>(o_p);
  m_ = D_.at<float>(o_m);
  v_ = D_.at<float>(o_v);
  grads_ = D_.at<float>(o_g);
  for (int l = 0; l < kMlpLayers; ++l) {
    Wb_[l] = D_.at<__bf16>(o_wb[l]);
    Wtb_[l] = D_.at<__bf16>(o_wtb[l]);
  }
  X_ = D_.at<__bf16>(o_x);
  XT_ = D_.at<__bf16>(o_xt);
  for (int l = 0; l < 3; ++l) {
    Z_[l] = D_.at<float>(o_z[l]);
    H_[l] = D_.at<__bf16>(o_h[l]);
    HT_[l] = D_.at<__bf16>(o_ht[l]);
    dZ_[l] = D_.at<__bf16>(o_dz[l]);
    dZT_[l] = D_.at<__bf16>(o_dzt[l]);
  }
  O_ = D_.at<float>(o_o);
  dO_ = D_.at<__bf16>(o_do);
  dOT_ = D_.at<__bf16>(o_dot);
  loss_ = D_.at<float>(o_loss);
  Yb_ = D_.at<float>(o_yb);
  Xb_ = D_.at<float>(o_xb);

  return reset(seed);
}

cusna_status Mlp::reset(uint64_t seed) {
  // Glorot-uniform weights over the real (unpadded) fan-in / fan-out, zero
  // biases and padding, from a splitmix64 stream
  const int p_in = p_in_, p_out = p_out_, width = width_;
  std::vector<float> h(n_params_, 0.0f);
  uint64_t st = seed;
  auto next = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int l = 0; l < kMlpLayers; ++l) {
    const int kin = l == 0 ? p_in : width, nout = l == kMlpLayers - 1 ? 2 * p_out : width;
    const double a = std::sqrt(6.0 / (double)(kin + nout));
    for (int k = 0; k < kin; ++k)
      for (int n = 0; n < nout; ++n)
        h[off_w_[l] + (size_t)k * N_[l] + n] = (float)((2.0 * next() - 1.0) * a);
  }
  return set_params(h.data());
}

cusna_status Mlp::set_params(const float* host_flat) {
  HIP_TRY(hipMemcpyAsync(params_, host_flat, n_params_ * 4, hipMemcpyHostToDevice, s_));
  std::vector<float> z(n_params_, 0.0f);
  HIP_TRY(hipMemcpyAsync(m_, z.data(), n_params_ * 4, hipMemcpyHostToDevice, s_));
  HIP_TRY(hipMemcpyAsync(v_, z.data(), n_params_ * 4, hipMemcpyHostToDevice, s_));
  HIP_TRY(hipMemcpyAsync(grads_, z.data(), n_params_ * 4, hipMemcpyHostToDevice, s_));
  t_ = 0;
  cusna_status st = cast_weights();
  if (st != CUSNA_OK) return st;
  HIP_TRY(hipStreamSynchronize(s_));
  return CUSNA_OK;
}

cusna_status Mlp::get_params(float* host_flat) {
  HIP_TRY(hipMemcpyAsync(host_flat, params_, n_params_ * 4, hipMemcpyDeviceToHost, s_));
  HIP_TRY(hipStreamSynchronize(s_));
  return CUSNA_OK;
}

cusna_status Mlp::cast_weights() {
  for (int l = 0; l < kMlpLayers; ++l) {
    const int n = K_[l] * N_[l];
    hipLaunchKernelGGL(cast_kernel, dim3((n + 255) / 256), dim3(256), 0, s_,
                       params_ + off_w_[l], K_[l], N_[l], Wb_[l], Wtb_[l]);
  }
  HIP_TRY(hipGetLastError());
  return CUSNA_OK;
}

cusna_status Mlp::gemm(const __bf16* A, const __bf16* Bt, int M, int N, int K, const float* bias,
                       const float* dz, int act, float* Cf, __bf16* Cb, __bf16* CbT) {
  if (M % 64 || N % 64 || K % 32) {
    set_error("mlp: product shape not a multiple of the 64 x 64 x 32 tile");
    return CUSNA_ERR_INTERNAL;
  }
  // split K (128 per split) for long plain float32 products
  const bool plain = Cf && !bias && !dz && !act && !Cb && !CbT;
  const int S = (plain && K >= 512 && K % 128 == 0) ? K / 128 : 1;
  if (S > 1 && (size_t)S * M * N > part_floats_) {
    set_error("mlp: split-product scratch too small");
    return CUSNA_ERR_INTERNAL;
  }
  GemmArgs g{A, Bt, M, N, K, bias, dz, act, Cf, Cb, CbT, K / S, part_};
  const dim3 grid(M / 64, N / 64, S);
  if (g.Ks % 256 == 0)
    hipLaunchKernelGGL(gemm_bf16_kernel<8>, grid, dim3(kMlpThreads), 0, s_, g);
  else if (g.Ks % 128 == 0)
    hipLaunchKernelGGL(gemm_bf16_kernel<4>, grid, dim3(kMlpThreads), 0, s_, g);
  else if (g.Ks % 64 == 0)
    hipLaunchKernelGGL(gemm_bf16_kernel<2>, grid, dim3(kMlpThreads), 0, s_, g);
  else
    hipLaunchKernelGGL(gemm_bf16_kernel<1>, grid, dim3(kMlpThreads), 0, s_, g);
  if (S > 1)
    hipLaunchKernelGGL(sum_parts_kernel, dim3((M * N + 255) / 256), dim3(256), 0, s_, part_, S,
                       M * N, Cf);
  HIP_TRY(hipGetLastError());
  return CUSNA_OK;
}

cusna_status Mlp::forward(const float* x_dev) {
  const int B = B_;
  hipLaunchKernelGGL(prep_input_kernel, dim3((B * K_[0] + 255) / 256), dim3(256), 0, s_, x_dev,
                     B, p_in_, K_[0], X_, XT_);
  cusna_status st;
  const __bf16* in = X_;
  for (int l = 0; l < kMlpLayers - 1; ++l) {
    st = gemm(in, Wtb_[l], B, N_[l], K_[l], params_ + off_b_[l], nullptr, 1, Z_[l], H_[l],
              HT_[l]);
    if (st != CUSNA_OK) return st;
    in = H_[l];
  }
  return gemm(in, Wtb_[3], B, N_[3], K_[3], params_ + off_b_[3], nullptr, 0, O_, nullptr,
              nullptr);
}

cusna_status Mlp::loss_and_grad(const float* y_dev) {
  hipLaunchKernelGGL(nll_kernel, dim3((B_ + 63) / 64), dim3(64), 0, s_, O_, y_dev, B_, p_out_,
                     N_[3], loss_, dO_, dOT_);
  HIP_TRY(hipGetLastError());
  return CUSNA_OK;
}

cusna_status Mlp::forward_backward(const float* x_dev, const float* y_dev) {
  cusna_status st = forward(x_dev);
  if (st != CUSNA_OK) return st;
  if ((st = loss_and_grad(y_dev)) != CUSNA_OK) return st;
  const int B = B_;
  // layer l gradient G_l (B x N_l): dO for the output layer, dZ below
  const __bf16* G = dO_;
  const __bf16* GT = dOT_;
  for (int l = kMlpLayers - 1; l >= 0; --l) {
    const __bf16* inT = l == 0 ? XT_ : HT_[l - 1];
    // dW_l = H_{l-1}^T G_l  (K_l x N_l), reduction over the batch
    st = gemm(inT, GT, K_[l], N_[l], B, nullptr, nullptr, 0, grads_ + off_w_[l], nullptr,
              nullptr);
    if (st != CUSNA_OK) return st;
    hipLaunchKernelGGL(rowsum_kernel, dim3(N_[l]), dim3(kMlpThreads), 0, s_, GT, B,
                       grads_ + off_b_[l]);
    if (l == 0) break;
    // G_{l-1} = (G_l W_l^T) * SiLU'(Z_{l-1})
    st = gemm(G, Wb_[l], B, K_[l], N_[l], nullptr, Z_[l - 1], 0, nullptr, dZ_[l - 1],
              dZT_[l - 1]);
    if (st != CUSNA_OK) return st;
    G = dZ_[l - 1];
    GT = dZT_[l - 1];
  }
  HIP_TRY(hipGetLastError());
  return CUSNA_OK;
}

cusna_status Mlp::adam_step(float lr) {
  ++t_;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)t_)));
  const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)t_)));
  CastMap cm;
  for (int l = 0; l < kMlpLayers; ++l) {
    cm.off[l] = off_w_[l];
    cm.K[l] = K_[l];
    cm.N[l] = N_[l];
    cm.Wb[l] = Wb_[l];
    cm.Wtb[l] = Wtb_[l];
  }
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n_params_ + 255) / 256)), dim3(256), 0, s_,
                     params_, m_, v_, grads_, n_params_, lr, b1, b2, eps, c1, c2, cm);
  HIP_TRY(hipGetLastError());
  return CUSNA_OK;
}

cusna_status Mlp::train(const float* x_all, const float* y_all, int64_t n, int steps, float lr,
                        uint64_t seed) {
  if (n < 1) return CUSNA_ERR_SHAPE;
  for (int sidx = 0; sidx < steps; ++sidx) {
    hipLaunchKernelGGL(gather_kernel, dim3((B_ + 63) / 64), dim3(64), 0, s_, x_all, y_all,
                       (long long)n, p_in_, p_out_, B_, seed, (long long)sidx, Xb_, Yb_);
    cusna_status st = forward_backward(Xb_, Yb_);
    if (st != CUSNA_OK) return st;
    const float lr_s = lr * (1.0f - 0.9f * (float)sidx / (float)(steps > 1 ? steps - 1 : 1));
    if ((st = adam_step(lr_s)) != CUSNA_OK) return st;
  }
  return CUSNA_OK;
}

cusna_status Mlp::last_loss(double* loss) {
  std::vector<float> h(B_);
  HIP_TRY(hipMemcpyAsync(h.data(), loss_, (size_t)B_ * 4, hipMemcpyDeviceToHost, s_));
  HIP_TRY(hipStreamSynchronize(s_));
  double s = 0.0;
  for (float x : h) s += x;
  *loss = s / B_;
  return CUSNA_OK;
}

}  // namespace cusna_impl

// ---------------------------------------------------------------- C ABI
using cusna_impl::Mlp;

struct cusna_mlp {
  Mlp net;
  hipStream_t s = nullptr;
  ~cusna_mlp() {
    if (s) (void)hipStreamDestroy(s);
  }
};

extern "C" {

CUSNA_API cusna_status cusna_mlp_create(int32_t p_in, int32_t p_out, int32_t width,
                                        int32_t batch, uint64_t seed, cusna_mlp** out) {
  if (!out) return CUSNA_ERR_NULL_ARG;
  if (cusna_impl::device_count() < 1) return CUSNA_ERR_NO_CUDA;
  std::unique_ptr<cusna_mlp> m(new cusna_mlp());
  HIP_TRY(hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking));
  cusna_status st = m->net.init(p_in, p_out, width, batch, seed, m->s);
  if (st != CUSNA_OK) return st;
  *out = m.release();
  return CUSNA_OK;
}

CUSNA_API int64_t cusna_mlp_param_count(const cusna_mlp* m) {
  return m ? (int64_t)m->net.param_count() : -1;
}

CUSNA_API cusna_status cusna_mlp_layer_shape(const cusna_mlp* m, int32_t layer, int32_t* K,
                                             int32_t* N) {
  if (!m || !K || !N) return CUSNA_ERR_NULL_ARG;
  if (layer < 0 || layer >= cusna_impl::kMlpLayers) return CUSNA_ERR_SHAPE;
  *K = m->net.K(layer);
  *N = m->net.N(layer);
  return CUSNA_OK;
}

CUSNA_API cusna_status cusna_mlp_get_params(cusna_mlp* m, float* flat) {
  if (!m || !flat) return CUSNA_ERR_NULL_ARG;
  return m->net.get_params(flat);
}

CUSNA_API cusna_status cusna_mlp_set_params(cusna_mlp* m, const float* flat) {
  if (!m || !flat) return CUSNA_ERR_NULL_ARG;
  return m->net.set_params(flat);
}

CUSNA_API cusna_status cusna_mlp_forward_backward(cusna_mlp* m, const float* x, const float* y,
                                                  float* out, float* grads, double* loss) {
  if (!m || !x || !y) return CUSNA_ERR_NULL_ARG;
  Mlp& net = m->net;
  const int B = net.batch();
  cusna_impl::DeviceBlock D;
  const size_t ox = D.add((size_t)B * net.p_in() * 4), oy = D.add((size_t)B * net.p_out() * 4);
  HIP_TRY(D.alloc());
  HIP_TRY(hipMemcpyAsync(D.at<char>(ox), x, (size_t)B * net.p_in() * 4, hipMemcpyHostToDevice,
                         m->s));
  HIP_TRY(hipMemcpyAsync(D.at<char>(oy), y, (size_t)B * net.p_out() * 4, hipMemcpyHostToDevice,
                         m->s));
  cusna_status st = net.forward_backward(D.at<float>(ox), D.at<float>(oy));
  if (st != CUSNA_OK) return st;
  if (out)
    HIP_TRY(hipMemcpyAsync(out, net.outputs(), (size_t)B * net.N(3) * 4, hipMemcpyDeviceToHost,
                           m->s));
  if (grads)
    HIP_TRY(hipMemcpyAsync(grads, net.grads(), net.param_count() * 4, hipMemcpyDeviceToHost,
                           m->s));
  HIP_TRY(hipStreamSynchronize(m->s));
  if (loss) return net.last_loss(loss);
  return CUSNA_OK;
}

CUSNA_API cusna_status cusna_mlp_train(cusna_mlp* m, const float* x_all, const float* y_all,
                                       int64_t n, int32_t steps, float lr, uint64_t seed,
                                       double* last_loss) {
  if (!m || !x_all || !y_all) return CUSNA_ERR_NULL_ARG;
  if (n < 1 || steps < 0) return CUSNA_ERR_SHAPE;
  Mlp& net = m->net;
  cusna_impl::DeviceBlock D;
  const size_t ox = D.add((size_t)n * net.p_in() * 4), oy = D.add((size_t)n * net.p_out() * 4);
  HIP_TRY(D.alloc());
  HIP_TRY(hipMemcpyAsync(D.at<char>(ox), x_all, (size_t)n * net.p_in() * 4,
                         hipMemcpyHostToDevice, m->s));
  HIP_TRY(hipMemcpyAsync(D.at<char>(oy), y_all, (size_t)n * net.p_out() * 4,
                         hipMemcpyHostToDevice, m->s));
  cusna_status st = net.train(D.at<float>(ox), D.at<float>(oy), n, steps, lr, seed);
  if (st != CUSNA_OK) return st;
  HIP_TRY(hipStreamSynchronize(m->s));
  if (last_loss) return net.last_loss(last_loss);
  return CUSNA_OK;
}

CUSNA_API cusna_status cusna_mlp_predict(cusna_mlp* m, const float* x, int64_t n, float* out) {
  if (!m || !x || !out) return CUSNA_ERR_NULL_ARG;
  Mlp& net = m->net;
  const int B = net.batch(), pin = net.p_in(), N3 = net.N(3), po = net.p_out();
  cusna_impl::DeviceBlock D;
  const size_t ox = D.add((size_t)B * pin * 4);
  HIP_TRY(D.alloc());
  std::vector<float> xb((size_t)B * pin), ob((size_t)B * N3);
  for (int64_t b0 = 0; b0 < n; b0 += B) {
    const int64_t nb = std::min<int64_t>(B, n - b0);
    std::fill(xb.begin(), xb.end(), 0.0f);
    std::memcpy(xb.data(), x + b0 * pin, (size_t)nb * pin * 4);
    HIP_TRY(hipMemcpyAsync(D.at<char>(ox), xb.data(), xb.size() * 4, hipMemcpyHostToDevice,
                           m->s));
    cusna_status st = net.forward(D.at<float>(ox));
    if (st != CUSNA_OK) return st;
    HIP_TRY(hipMemcpyAsync(ob.data(), net.outputs(), ob.size() * 4, hipMemcpyDeviceToHost, m->s));
    HIP_TRY(hipStreamSynchronize(m->s));
    for (int64_t b = 0; b < nb; ++b)
      std::memcpy(out + (b0 + b) * 2 * po, ob.data() + b * N3, (size_t)2 * po * 4);
  }
  return CUSNA_OK;
}

CUSNA_API void cusna_mlp_destroy(cusna_mlp* m) { delete m; }

}  // extern "C"
