// mlp.hip -- the surrogate perceptron on gfx950 BF16 matrix cores.
//
// A training step is two launches.
//
// mlp_slice_kernel: one workgroup per 16 batch rows runs the whole network
// for its rows -- the four forward products, the loss and its output
// gradient, and the three backward products of the activation gradients --
// with every activation, SiLU'(z) and gradient of the slice in LDS. A
// product is A (LDS, rows of the slice) times W^T or W (device memory, read
// once per workgroup from L2): the four waves take interleaved 16-column
// tiles of the output, each lane loading 16 bytes of A and of B per 32-wide
// K step for v_mfma_f32_16x16x32_bf16 (float32 accumulation). Operand layout
// of that instruction (checked on the device): lane l holds
// A[l % 16][8 (l / 16) + e] and B[8 (l / 16) + e][l % 16], e = 0..7, and
// receives D[4 (l / 16) + v][l % 16], v = 0..3. The slice's inputs and the
// transposed bfloat16 activations and gradients (X^T, H_l^T, G_l^T) go to
// device memory for the weight gradients.
//
// mlp_wgrad_adam_kernel: one workgroup per 64 x 64 tile of a weight matrix
// forms dW = H^T G over the whole batch (the four waves take interleaved
// pairs of 32-row K steps, eight steps' loads in flight at once, and their partial
// tiles are added in wave order: the result does not depend on timing), the
// bias gradients of its columns, and applies Adam to the tile (its state
// read before the product), writing the float32 master weights and their
// row-major and transposed bfloat16 copies through LDS as 16-byte stores.
//
// Rounding points: inputs, weights, each hidden activation and each
// back-propagated gradient are rounded to bfloat16 as they enter a product;
// pre-activations, outputs, SiLU'(z), the loss and the weight gradients are
// float32.
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
constexpr int kSlice = 16;              // batch rows per workgroup of the slice kernel
constexpr int kSliceTM = kSlice / 16;   // 16-row tiles per slice
constexpr int kMaxWidth = 256;          // hidden width bound of the fused step (LDS)

// sigmoid(z) from the hardware exp2 and reciprocal (about 1 ulp each; the
// epilogues run at one wave per SIMD, where IEEE division and expf's range
// reduction were most of a layer's time); SiLU(z) = z s, SiLU'(z) = s (1 + z (1 - s)).
__device__ __forceinline__ float sigmoid_fast(float z) {
  const float e = __builtin_amdgcn_exp2f(-1.44269504088896341f * z);
  return __builtin_amdgcn_rcpf(1.0f + e);
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xFF51AFD7ED558CCDull;
  x ^= x >> 33;
  x *= 0xC4CEB9FE1A85EC53ull;
  return x ^ (x >> 33);
}

struct SliceArgs {
  const __bf16* Wt[kMlpLayers];  // W_l^T [N_l][K_l]: forward operands
  const __bf16* W[kMlpLayers];   // W_l [K_l][N_l]: backward operands
  const float* bias[kMlpLayers];
  int K[kMlpLayers], N[kMlpLayers];
  int p_in, p_out, B;
  int mode;                      // 0 forward, 1 forward + backward, 2 training step
  const float* x;                // modes 0, 1: B x p_in; mode 2: n x p_in
  const float* y;                // mode 1: B x p_out; mode 2: n x p_out
  long long n;                   // mode 2: samples to draw from
  uint64_t seed;                 // mode 2: row b of step s is sample
  long long step;                //   mix64(seed + s B + b) mod n
  float* O;                      // B x N_3 outputs
  float* loss;                   // B per-sample losses (modes 1, 2)
  __bf16* XT;                    // K_0 x B
  __bf16* HT[kMlpLayers - 1];    // N_l x B
  __bf16* GT[kMlpLayers];        // N_l x B: gradient of layer l's pre-activation (output)
};

// LDS of the slice kernel: byte offsets and row strides (elements).
struct SliceLds {
  int xs, ha, hb, d, os, dos, idx, total;
  int ldx, ldh, ldd, ldo;
};

__host__ __device__ inline SliceLds slice_lds(int K0, int width, int N3) {
  SliceLds L;
  L.ldx = K0 + 8;      // bfloat16 rows padded by 16 bytes (banks)
  L.ldh = width + 8;
  L.ldd = width + 1;   // float32 rows padded by one word
  L.ldo = N3 + 8;
  int o = 0;
  L.xs = o;
  o += kSlice * L.ldx * 2;
  L.ha = o;  // activations / gradients, ping-pong
  o += kSlice * L.ldh * 2;
  L.hb = o;
  o += kSlice * L.ldh * 2;
  L.d = o;   // SiLU'(z) of the three hidden layers
  o += 3 * kSlice * L.ldd * 4;
  L.os = o;  // float32 outputs
  o += kSlice * N3 * 4;
  L.dos = o; // output gradient, bfloat16
  o += kSlice * L.ldo * 2;
  L.idx = o; // sample index of each row
  o += kSlice * 8;
  L.total = o;
  return L;
}

// B operand fragments of one product for the wave: columns 16 (w + 4 j)..
// (j < TNW) of B, given as Bt [N][K] in device memory with K = 32 KS, every
// K step. Loaded before they are needed: weights do not depend on the
// activations, so a product's fragments are in flight while the previous
// product's epilogue, barrier and stores run.
template <int TNW, int KS>
struct BFrag {
  bf16x8 v[KS][TNW];
};

template <int TNW, int KS>
__device__ __forceinline__ void load_b(const __bf16* Bt, BFrag<TNW, KS>& f) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6, r = l & 15, q = l >> 4;
  constexpr int K = 32 * KS;
  const __bf16* brow = Bt + (size_t)(16 * w + r) * K + 8 * q;
#pragma unroll
  for (int s = 0; s < KS; ++s)
#pragma unroll
    for (int j = 0; j < TNW; ++j)
      f.v[s][j] = *reinterpret_cast<const bf16x8*>(brow + (size_t)64 * j * K + 32 * s);
}

// acc[i][j] = rows 16 i.. of A (LDS, row stride lda) times the fragments' columns.
template <int TNW, int KS>
__device__ __forceinline__ void slice_mma(const __bf16* A, int lda, const BFrag<TNW, KS>& f,
                                          floatx4 (&acc)[kSliceTM][TNW]) {
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
#pragma unroll
  for (int i = 0; i < kSliceTM; ++i)
#pragma unroll
    for (int j = 0; j < TNW; ++j) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  const __bf16* arow = A + r * lda + 8 * q;
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    bf16x8 a[kSliceTM];
#pragma unroll
    for (int i = 0; i < kSliceTM; ++i)
      a[i] = *reinterpret_cast<const bf16x8*>(arow + 16 * i * lda + 32 * s);
#pragma unroll
    for (int i = 0; i < kSliceTM; ++i)
#pragma unroll
      for (int j = 0; j < TNW; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], f.v[s][j], acc[i][j], 0, 0, 0);
  }
}

// dst [ncols][B], columns b0..b0+kSlice-1 = src (LDS [kSlice][ld]) transposed,
// 16-byte stores.
__device__ __forceinline__ void store_transposed(const __bf16* src, int ld, int ncols,
                                                 __bf16* dst, int B, int b0) {
  constexpr int kChunks = kSlice / 8;
  for (int it = threadIdx.x; it < ncols * kChunks; it += kMlpThreads) {
    const int c = it / kChunks, r8 = 8 * (it - c * kChunks);
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = src[(r8 + e) * ld + c];
    *reinterpret_cast<bf16x8*>(dst + (size_t)c * B + b0 + r8) = v;
  }
}

// Hidden-layer epilogue: H = bf16(SiLU(z)), z = acc + bias; SiLU'(z) kept
// (by the lane that will need it in the backward pass) when bwd.
template <int TNW>
__device__ __forceinline__ void load_bias(const float* bias, float (&bc)[TNW]) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6, r = l & 15;
#pragma unroll
  for (int j = 0; j < TNW; ++j) bc[j] = bias[16 * w + 64 * j + r];
}

template <int NWW>
__device__ __forceinline__ void hidden_epilogue(const floatx4 (&acc)[kSliceTM][NWW],
                                                const float (&bias)[NWW], __bf16* out, int ldh,
                                                float* Dl, int ldd, bool bwd) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6, r = l & 15, q = l >> 4;
#pragma unroll
  for (int i = 0; i < kSliceTM; ++i)
#pragma unroll
    for (int j = 0; j < NWW; ++j) {
      const int col = 16 * w + 64 * j + r;
      const float bc = bias[j];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * i + 4 * q + v;
        const float z = acc[i][j][v] + bc;
        const float sg = sigmoid_fast(z);
        if (bwd) Dl[row * ldd + col] = sg * (1.0f + z * (1.0f - sg));
        out[row * ldh + col] = (__bf16)(z * sg);
      }
    }
}

// Backward epilogue: G = bf16(acc * SiLU'(z)) at the same positions.
template <int NWW>
__device__ __forceinline__ void grad_epilogue(const floatx4 (&acc)[kSliceTM][NWW], __bf16* out,
                                              int ldh, const float* Dl, int ldd) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6, r = l & 15, q = l >> 4;
#pragma unroll
  for (int i = 0; i < kSliceTM; ++i)
#pragma unroll
    for (int j = 0; j < NWW; ++j) {
      const int col = 16 * w + 64 * j + r;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * i + 4 * q + v;
        out[row * ldh + col] = (__bf16)(acc[i][j][v] * Dl[row * ldd + col]);
      }
    }
}

// NWW = width / 64, N3W = N_3 / 64 (K_0 = 64: p_in <= 64).
template <int NWW, int N3W>
__global__ void __launch_bounds__(kMlpThreads) mlp_slice_kernel(const SliceArgs g) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int K0 = 64, width = 64 * NWW, N3 = 64 * N3W;
  const int B = g.B, p = g.p_out;
  const SliceLds L = slice_lds(K0, width, N3);
  __bf16* Xs = reinterpret_cast<__bf16*>(smem + L.xs);
  __bf16* Ha = reinterpret_cast<__bf16*>(smem + L.ha);
  __bf16* Hb = reinterpret_cast<__bf16*>(smem + L.hb);
  float* D = reinterpret_cast<float*>(smem + L.d);
  float* Os = reinterpret_cast<float*>(smem + L.os);
  __bf16* dOs = reinterpret_cast<__bf16*>(smem + L.dos);
  long long* idx = reinterpret_cast<long long*>(smem + L.idx);
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6, r = l & 15, q = l >> 4;
  const int b0 = blockIdx.x * kSlice;
  const bool bwd = g.mode != 0;
  const int dstride = kSlice * L.ldd;
  BFrag<NWW, 2> f0;
  load_b(g.Wt[0], f0);
  float bc0[NWW], bc1[NWW], bc2[NWW], bc3[N3W];  // biases, read with the first weights
  load_bias(g.bias[0], bc0);
  load_bias(g.bias[1], bc1);
  load_bias(g.bias[2], bc2);
  load_bias(g.bias[kMlpLayers - 1], bc3);
  if (tid < kSlice) {
    long long src = b0 + tid;
    if (g.mode == 2)
      src = (long long)(mix64(g.seed + (uint64_t)g.step * (uint64_t)B + (uint64_t)(b0 + tid)) %
                        (uint64_t)g.n);
    idx[tid] = src;
  }
  __syncthreads();
  for (int e = tid; e < kSlice * K0; e += kMlpThreads) {
    const int row = e / K0, k = e - row * K0;
    Xs[row * L.ldx + k] = (__bf16)(k < g.p_in ? g.x[(size_t)idx[row] * g.p_in + k] : 0.0f);
  }
  // the loss phase's targets, read now: thread tid takes output column
  // tid % N_3 of rows tid / N_3 + (256 / N_3) c
  constexpr int kNll = kSlice * N3 / kMlpThreads, kRowStep = kMlpThreads / N3;
  const int ac = tid % N3, rbase = tid / N3;
  float yv[kNll];
#pragma unroll
  for (int c = 0; c < kNll; ++c)
    yv[c] = (bwd && ac < 2 * p) ? g.y[(size_t)idx[rbase + kRowStep * c] * p + (ac < p ? ac : ac - p)]
                                : 0.0f;
  __syncthreads();
  if (bwd) store_transposed(Xs, L.ldx, K0, g.XT, B, b0);
  // hidden layers: H_l = bf16(SiLU(A W_l + b_l)), SiLU'(z) kept for the backward pass
  floatx4 acc[kSliceTM][NWW];
  BFrag<NWW, 2 * NWW> fh;
  slice_mma(Xs, L.ldx, f0, acc);
  load_b(g.Wt[1], fh);
  hidden_epilogue(acc, bc0, Ha, L.ldh, D, L.ldd, bwd);
  __syncthreads();
  if (bwd) store_transposed(Ha, L.ldh, width, g.HT[0], B, b0);
  slice_mma(Ha, L.ldh, fh, acc);
  load_b(g.Wt[2], fh);
  hidden_epilogue(acc, bc1, Hb, L.ldh, D + dstride, L.ldd, bwd);
  __syncthreads();
  if (bwd) store_transposed(Hb, L.ldh, width, g.HT[1], B, b0);
  BFrag<N3W, 2 * NWW> fo;
  slice_mma(Hb, L.ldh, fh, acc);
  load_b(g.Wt[3], fo);
  hidden_epilogue(acc, bc2, Ha, L.ldh, D + 2 * dstride, L.ldd, bwd);
  __syncthreads();
  if (bwd) store_transposed(Ha, L.ldh, width, g.HT[2], B, b0);
  BFrag<NWW, 2 * N3W> f3;
  {  // outputs O = H_2 W_3 + b_3
    floatx4 acco[kSliceTM][N3W];
    slice_mma(Ha, L.ldh, fo, acco);
    if (bwd) load_b(g.W[3], f3);
#pragma unroll
    for (int i = 0; i < kSliceTM; ++i)
#pragma unroll
      for (int j = 0; j < N3W; ++j) {
        const int col = 16 * w + 64 * j + r;
        const float bc = bc3[j];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 16 * i + 4 * q + v;
          const float o = acco[i][j][v] + bc;
          Os[row * N3 + col] = o;
          g.O[(size_t)(b0 + row) * N3 + col] = o;
        }
      }
  }
  if (!bwd) return;
  __syncthreads();
  // Gaussian negative log-likelihood (mu in columns 0..p-1, log-variance in
  // p..2p-1, clamped to [-12, 12] with zero gradient outside) and its
  // gradient with respect to the outputs, scaled by 1/B; (row, column) items
  // as the targets were read, the row's loss terms added in column order below. The terms
  // take the two activation buffers (H_1, H_2 are stored and consumed;
  // kSlice p floats fit in them as p <= width + 8).
  float* terms = reinterpret_cast<float*>(Ha);
  const float invb = 1.0f / (float)B;
#pragma unroll
  for (int c = 0; c < kNll; ++c) {
    const int row = rbase + kRowStep * c, a = ac;
    float gv = 0.0f;
    if (a < p) {
      const float mu = Os[row * N3 + a];
      const float lvr = Os[row * N3 + p + a];
      const float lv = fminf(fmaxf(lvr, -12.0f), 12.0f);
      const float e = expf(-lv);
      const float res = yv[c] - mu;
      terms[row * p + a] = 0.5f * (res * res * e + lv);
      gv = -res * e * invb;
    } else if (a < 2 * p) {
      const int cc = a - p;
      const float mu = Os[row * N3 + cc];
      const float lvr = Os[row * N3 + a];
      const float lv = fminf(fmaxf(lvr, -12.0f), 12.0f);
      const float res = yv[c] - mu;
      gv = (lvr > -12.0f && lvr < 12.0f) ? 0.5f * (1.0f - res * res * expf(-lv)) * invb : 0.0f;
    }
    dOs[row * L.ldo + a] = (__bf16)gv;
  }
  __syncthreads();
  if (tid < kSlice) {
    float s = 0.0f;
    for (int a = 0; a < p; ++a) s += terms[tid * p + a];
    g.loss[b0 + tid] = s;
  }
  store_transposed(dOs, L.ldo, N3, g.GT[kMlpLayers - 1], B, b0);
  __syncthreads();
  // G_{l-1} = bf16((G_l W_l^T) * SiLU'(z_{l-1})), l = 3, 2, 1
  slice_mma(dOs, L.ldo, f3, acc);
  load_b(g.W[2], fh);
  grad_epilogue(acc, Hb, L.ldh, D + 2 * dstride, L.ldd);
  __syncthreads();
  store_transposed(Hb, L.ldh, width, g.GT[2], B, b0);
  slice_mma(Hb, L.ldh, fh, acc);
  load_b(g.W[1], fh);
  grad_epilogue(acc, Ha, L.ldh, D + dstride, L.ldd);
  __syncthreads();
  store_transposed(Ha, L.ldh, width, g.GT[1], B, b0);
  slice_mma(Ha, L.ldh, fh, acc);
  grad_epilogue(acc, Hb, L.ldh, D, L.ldd);
  __syncthreads();
  store_transposed(Hb, L.ldh, width, g.GT[0], B, b0);
}

struct WgradArgs {
  const __bf16* AT[kMlpLayers];  // layer inputs transposed: X^T, H_0^T, H_1^T, H_2^T (K_l x B)
  const __bf16* GT[kMlpLayers];  // N_l x B
  int K[kMlpLayers], N[kMlpLayers], B;
  int first[kMlpLayers + 1];     // first workgroup of each layer's 64 x 64 tiles
  float *params, *m, *v, *grads;
  size_t off_w[kMlpLayers], off_b[kMlpLayers];
  __bf16* Wb[kMlpLayers];        // W_l [K][N]
  __bf16* Wtb[kMlpLayers];       // W_l^T [N][K]
  int adam;                      // 0: gradients only
  float lr, b1, b2, eps, c1, c2; // Adam: bias corrections c1 = 1/(1-b1^t), c2 = 1/(1-b2^t)
};

constexpr int kWgradThreads = 256;            // four waves per weight tile (512 registers each)
constexpr int kWgradWaves = kWgradThreads / 64;
constexpr int kWgradPer = 64 * 64 / kWgradThreads;  // tile elements per thread
constexpr int kWgradU = 8;  // K steps per wave whose loads are in flight together
constexpr int kRedLd = 65;  // row stride of the partial tiles (floats)
constexpr size_t kWgradLds =
    (size_t)kWgradWaves * 64 * kRedLd * 4 + 64 * 72 * 2 + kWgradWaves * 64 * 4;

__global__ void __launch_bounds__(kWgradThreads) mlp_wgrad_adam_kernel(const WgradArgs g) {
  extern __shared__ __align__(16) unsigned char smem[];
  float* red = reinterpret_cast<float*>(smem);  // [waves][64][kRedLd]
  __bf16* tile = reinterpret_cast<__bf16*>(smem + (size_t)kWgradWaves * 64 * kRedLd * 4);
  float* bpart = reinterpret_cast<float*>(smem + (size_t)kWgradWaves * 64 * kRedLd * 4 +
                                          64 * 72 * 2);  // [waves][64]
  const int bid = blockIdx.x;
  int layer = 0;
  while (layer < kMlpLayers - 1 && bid >= g.first[layer + 1]) ++layer;
  const int N = g.N[layer], K = g.K[layer], B = g.B;
  const int t = bid - g.first[layer], ntn = N / 64;
  const int k0 = 64 * (t / ntn), n0 = 64 * (t - (t / ntn) * ntn);
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6, r = l & 15, q = l >> 4;
  const bool has_bias = k0 == 0;  // the first row of tiles also takes the bias columns
  // Adam state of the thread's tile elements (tid + 256 c), read ahead of the product
  float pw[kWgradPer], pm[kWgradPer], pv[kWgradPer];
  float bw = 0.0f, bm = 0.0f, bv = 0.0f;
  if (g.adam) {
#pragma unroll
    for (int c = 0; c < kWgradPer; ++c) {
      const int e = tid + kWgradThreads * c;
      const size_t ix = g.off_w[layer] + (size_t)(k0 + (e >> 6)) * N + n0 + (e & 63);
      pw[c] = g.params[ix];
      pm[c] = g.m[ix];
      pv[c] = g.v[ix];
    }
    if (has_bias && tid < 64) {
      const size_t ix = g.off_b[layer] + n0 + tid;
      bw = g.params[ix];
      bm = g.m[ix];
      bv = g.v[ix];
    }
  }
  const __bf16* A = g.AT[layer];
  const __bf16* G = g.GT[layer];
  floatx4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  float bsum[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // column n0 + 16 j + r, the lane's batch rows
  // K steps of 32 batch rows; wave w takes the pairs of steps 2 w, 2 w + 1,
  // 2 w + 8, 2 w + 9, ... (a pair reads whole 128-byte lines), eight at a time
  const int S = B / 32;
  for (int s0 = 2 * w; s0 < S; s0 += kWgradU * kWgradWaves) {
    bf16x8 a[kWgradU][4], b[kWgradU][4];
#pragma unroll
    for (int u = 0; u < kWgradU; ++u) {
      const int sx = s0 + 2 * kWgradWaves * (u >> 1) + (u & 1);
      const int su = sx < S ? sx : s0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[u][i] = *reinterpret_cast<const bf16x8*>(A + (size_t)(k0 + 16 * i + r) * B + 32 * su +
                                                   8 * q);
        b[u][i] = *reinterpret_cast<const bf16x8*>(G + (size_t)(n0 + 16 * i + r) * B + 32 * su +
                                                   8 * q);
      }
    }
    __builtin_amdgcn_sched_barrier(0);  // every step's loads before the MFMAs
#pragma unroll
    for (int u = 0; u < kWgradU; ++u) {
      if (s0 + 2 * kWgradWaves * (u >> 1) + (u & 1) >= S) break;
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] =
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[u][i], b[u][j], acc[i][j], 0, 0, 0);
      if (has_bias)  // the bias gradient is the sum of the same gradient fragments
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int e = 0; e < 8; ++e) bsum[j] += (float)b[u][j][e];
    }
  }
  float* mine = red + (size_t)w * 64 * kRedLd;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) mine[(16 * i + 4 * q + v) * kRedLd + 16 * j + r] = acc[i][j][v];
  // bias gradients of the tile's columns: the four lanes of a column (q)
  // added as a fixed tree, the waves in wave order below
  if (has_bias) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float v = bsum[j];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      if (q == 0) bpart[w * 64 + 16 * j + r] = v;
    }
  }
  __syncthreads();
  const float b1 = g.b1, b2 = g.b2;
#pragma unroll
  for (int c = 0; c < kWgradPer; ++c) {
    const int e = tid + kWgradThreads * c, row = e >> 6, col = e & 63;
    float gs = red[row * kRedLd + col];
#pragma unroll
    for (int u = 1; u < kWgradWaves; ++u) gs += red[(size_t)u * 64 * kRedLd + row * kRedLd + col];
    const size_t ix = g.off_w[layer] + (size_t)(k0 + row) * N + n0 + col;
    g.grads[ix] = gs;
    if (g.adam) {
      const float mi = b1 * pm[c] + (1.0f - b1) * gs;
      const float vi = b2 * pv[c] + (1.0f - b2) * gs * gs;
      const float wi = pw[c] - g.lr * (mi * g.c1) / (sqrtf(vi * g.c2) + g.eps);
      g.m[ix] = mi;
      g.v[ix] = vi;
      g.params[ix] = wi;
      tile[row * 72 + col] = (__bf16)wi;
    }
  }
  if (has_bias && tid < 64) {
    float gb = bpart[tid];
#pragma unroll
    for (int u = 1; u < kWgradWaves; ++u) gb += bpart[u * 64 + tid];
    const size_t ix = g.off_b[layer] + n0 + tid;
    g.grads[ix] = gb;
    if (g.adam) {
      const float mi = b1 * bm + (1.0f - b1) * gb;
      const float vi = b2 * bv + (1.0f - b2) * gb * gb;
      g.m[ix] = mi;
      g.v[ix] = vi;
      g.params[ix] = bw - g.lr * (mi * g.c1) / (sqrtf(vi * g.c2) + g.eps);
    }
  }
  if (!g.adam) return;
  __syncthreads();
  const int a = tid >> 2, c16 = 16 * (tid & 3);
  {  // row a of the tile into W [K][N]
    const bf16x8* src = reinterpret_cast<const bf16x8*>(&tile[a * 72 + c16]);
    bf16x8* dst = reinterpret_cast<bf16x8*>(g.Wb[layer] + (size_t)(k0 + a) * N + n0 + c16);
    dst[0] = src[0];
    dst[1] = src[1];
  }
  {  // column a of the tile as a row of W^T [N][K]
    bf16x8 lo, hi;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      lo[e] = tile[(c16 + e) * 72 + a];
      hi[e] = tile[(c16 + 8 + e) * 72 + a];
    }
    bf16x8* dst = reinterpret_cast<bf16x8*>(g.Wtb[layer] + (size_t)(n0 + a) * K + k0 + c16);
    dst[0] = lo;
    dst[1] = hi;
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

int pad64(int x) { return (x + 63) / 64 * 64; }

typedef void (*SliceKernel)(const SliceArgs);

template <int NWW, int N3W>
SliceKernel slice_instance() {
  return mlp_slice_kernel<NWW, N3W>;
}

// The slice kernel for width / 64 and N_3 / 64, its LDS limit raised once.
SliceKernel slice_kernel(int nww, int n3w, size_t lds) {
  static const SliceKernel table[4][2] = {
      {slice_instance<1, 1>(), slice_instance<1, 2>()},
      {slice_instance<2, 1>(), slice_instance<2, 2>()},
      {slice_instance<3, 1>(), slice_instance<3, 2>()},
      {slice_instance<4, 1>(), slice_instance<4, 2>()}};
  if (nww < 1 || nww > 4 || n3w < 1 || n3w > 2) return nullptr;
  SliceKernel k = table[nww - 1][n3w - 1];
  static bool raised[4][2] = {};
  if (!raised[nww - 1][n3w - 1]) {
    if (hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess)
      return nullptr;
    raised[nww - 1][n3w - 1] = true;
  }
  return k;
}

}  // namespace

cusna_status Mlp::init(int p_in, int p_out, int width, int batch, uint64_t seed, hipStream_t s) {
  if (p_in < 1 || p_in > 64 || p_out < 1 || p_out > 64 || width < 64 || width % 64 ||
      width > kMaxWidth || batch < 64 || batch % 64) {
    set_error("mlp: p_in, p_out in 1..64, width a multiple of 64 up to 256, batch a multiple "
              "of 64");
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
  size_t o_wb[kMlpLayers], o_wtb[kMlpLayers], o_gt[kMlpLayers], o_ht[kMlpLayers - 1];
  for (int l = 0; l < kMlpLayers; ++l) {
    o_wb[l] = D_.add((size_t)K_[l] * N_[l] * 2);
    o_wtb[l] = D_.add((size_t)K_[l] * N_[l] * 2);
    o_gt[l] = D_.add((size_t)N_[l] * B_ * 2);
  }
  for (int l = 0; l < kMlpLayers - 1; ++l) o_ht[l] = D_.add((size_t)N_[l] * B_ * 2);
  const size_t o_xt = D_.add((size_t)K_[0] * B_ * 2);
  const size_t o_o = D_.add((size_t)B_ * N_[3] * 4), o_loss = D_.add((size_t)B_ * 4);
  HIP_TRY(D_.alloc());
  params_ = D_.at<float>(o_p);
  m_ = D_.at<float>(o_m);
  v_ = D_.at<float>(o_v);
  grads_ = D_.at<float>(o_g);
  for (int l = 0; l < kMlpLayers; ++l) {
    Wb_[l] = D_.at<__bf16>(o_wb[l]);
    Wtb_[l] = D_.at<__bf16>(o_wtb[l]);
    GT_[l] = D_.at<__bf16>(o_gt[l]);
  }
  for (int l = 0; l < kMlpLayers - 1; ++l) HT_[l] = D_.at<__bf16>(o_ht[l]);
  XT_ = D_.at<__bf16>(o_xt);
  O_ = D_.at<float>(o_o);
  loss_ = D_.at<float>(o_loss);
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

cusna_status Mlp::slice(int mode, const float* x, const float* y, int64_t n, uint64_t seed,
                        long long step) {
  const SliceLds L = slice_lds(K_[0], width_, N_[3]);
  SliceKernel k = slice_kernel(width_ / 64, N_[3] / 64, (size_t)L.total);
  if (!k || (size_t)L.total > 160 * 1024) {
    set_error("mlp: no fused kernel for this shape");
    return CUSNA_ERR_INTERNAL;
  }
  SliceArgs a;
  for (int l = 0; l < kMlpLayers; ++l) {
    a.Wt[l] = Wtb_[l];
    a.W[l] = Wb_[l];
    a.bias[l] = params_ + off_b_[l];
    a.K[l] = K_[l];
    a.N[l] = N_[l];
    a.GT[l] = GT_[l];
  }
  for (int l = 0; l < kMlpLayers - 1; ++l) a.HT[l] = HT_[l];
  a.p_in = p_in_;
  a.p_out = p_out_;
  a.B = B_;
  a.mode = mode;
  a.x = x;
  a.y = y;
  a.n = (long long)n;
  a.seed = seed;
  a.step = step;
  a.O = O_;
  a.loss = loss_;
  a.XT = XT_;
  hipLaunchKernelGGL(k, dim3(B_ / kSlice), dim3(kMlpThreads), (size_t)L.total, s_, a);
  HIP_TRY(hipGetLastError());
  return CUSNA_OK;
}

cusna_status Mlp::wgrad(bool adam, float lr) {
  static bool raised = false;
  if (!raised) {
    HIP_TRY(hipFuncSetAttribute((const void*)mlp_wgrad_adam_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kWgradLds));
    raised = true;
  }
  WgradArgs a;
  int tiles = 0;
  for (int l = 0; l < kMlpLayers; ++l) {
    a.AT[l] = l == 0 ? XT_ : HT_[l - 1];
    a.GT[l] = GT_[l];
    a.K[l] = K_[l];
    a.N[l] = N_[l];
    a.first[l] = tiles;
    tiles += (K_[l] / 64) * (N_[l] / 64);
    a.off_w[l] = off_w_[l];
    a.off_b[l] = off_b_[l];
    a.Wb[l] = Wb_[l];
    a.Wtb[l] = Wtb_[l];
  }
  a.first[kMlpLayers] = tiles;
  a.B = B_;
  a.params = params_;
  a.m = m_;
  a.v = v_;
  a.grads = grads_;
  a.adam = adam ? 1 : 0;
  a.b1 = 0.9f;
  a.b2 = 0.999f;
  a.eps = 1e-8f;
  a.lr = lr;
  a.c1 = a.c2 = 1.0f;
  if (adam) {
    ++t_;
    a.c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)t_)));
    a.c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)t_)));
  }
  hipLaunchKernelGGL(mlp_wgrad_adam_kernel, dim3(tiles), dim3(kWgradThreads), kWgradLds, s_,
                     a);
  HIP_TRY(hipGetLastError());
  return CUSNA_OK;
}

cusna_status Mlp::forward(const float* x_dev) {
  return slice(0, x_dev, nullptr, B_, 0, 0);
}

cusna_status Mlp::forward_backward(const float* x_dev, const float* y_dev) {
  cusna_status st = slice(1, x_dev, y_dev, B_, 0, 0);
  if (st != CUSNA_OK) return st;
  return wgrad(false, 0.0f);
}

cusna_status Mlp::train(const float* x_all, const float* y_all, int64_t n, int steps, float lr,
                        uint64_t seed) {
  if (n < 1) return CUSNA_ERR_SHAPE;
  for (int sidx = 0; sidx < steps; ++sidx) {
    cusna_status st = slice(2, x_all, y_all, n, seed, sidx);
    if (st != CUSNA_OK) return st;
    const float lr_s = lr * (1.0f - 0.9f * (float)sidx / (float)(steps > 1 ? steps - 1 : 1));
    if ((st = wgrad(true, lr_s)) != CUSNA_OK) return st;
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
