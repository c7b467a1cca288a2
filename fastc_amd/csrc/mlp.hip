// BC7 candidate ranker: a 64 -> 256 -> 256 -> 512 MLP on the bf16 matrix
// cores of gfx950 (v_mfma_f32_32x32x16_bf16, fp32 accumulate), with the
// forward, backward and Adam kernels that train it.  No counterpart in the
// reference; it ranks the (mode, slot) candidates of a block so that only
// the best few are searched exactly.
//
// Every matrix product is "NT": C[M x N] = A[M x K] * Bt[N x K]^T with both
// operands row-major bf16 (K contiguous), so each lane loads its eight
// consecutive K values of a 32x32x16 MFMA operand with one 16-byte load.
// Layers keep their weights twice in bf16, [out][in] for the forward
// product and [in][out] for the input-gradient product; activations and
// output gradients are written row-major and transposed by the product that
// makes them, so the weight-gradient products are NT as well.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "internal.h"
#include "ranker_tile.h"

namespace fh {
namespace mlp {

// ---------------------------------------------------------------------------
// NT product with fused epilogue.  A workgroup of four waves computes a
// 128 x 64 tile; each wave a 32 x 64 strip as two 32x32 MFMA tiles.
// Requires M % 32 == 0, N % 32 == 0, K % 16 == 0 and 16-byte aligned rows.
// ---------------------------------------------------------------------------
struct Gemm {
  const uint16_t *A;
  const uint16_t *Bt;
  int M, N, K, lda, ldb;
  const float *bias;     // [N] or null
  int relu;
  const uint16_t *mask;  // bf16 [M x N] (ld ldm): zero where mask <= 0, or null
  int ldm;
  float *C32;            // fp32 [M x N] (ld ldc32) or null
  int ldc32;
  uint16_t *C16;         // bf16 [M x N] (ld ldc16) or null
  int ldc16;
  uint16_t *C16T;        // bf16 [N x M] (ld ldc16t) or null
  int ldc16t;
};

__global__ void __launch_bounds__(256) gemm_nt_kernel(Gemm g) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row0 = blockIdx.y * 128 + wave * 32;
  const int col0 = blockIdx.x * 64;
  if (row0 >= g.M) return;  // whole wave
  const int r = lane & 31, kh = (lane >> 5) * 8;
  const uint16_t *pa = g.A + (size_t)(row0 + r) * g.lda + kh;
  const bool has1 = col0 + 32 < g.N;
  const uint16_t *pb0 = g.Bt + (size_t)(col0 + r) * g.ldb + kh;
  const uint16_t *pb1 = g.Bt + (size_t)(col0 + 32 + (has1 ? r : 0)) * g.ldb + kh;
  floatx16 acc0 = {}, acc1 = {};
  for (int k = 0; k < g.K; k += 16) {
    const bf16x8 a = *reinterpret_cast<const bf16x8 *>(pa + k);
    const bf16x8 b0 = *reinterpret_cast<const bf16x8 *>(pb0 + k);
    acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b0, acc0, 0, 0, 0);
    if (has1) {
      const bf16x8 b1 = *reinterpret_cast<const bf16x8 *>(pb1 + k);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b1, acc1, 0, 0, 0);
    }
  }
  // D[i][j]: j = lane % 32, i = 8 * (v / 4) + 4 * (lane / 32) + v % 4
  for (int t = 0; t < 2; t++) {
    if (t == 1 && !has1) break;
    const int col = col0 + 32 * t + (lane & 31);
    const float bias = g.bias ? g.bias[col] : 0.0f;
#pragma unroll
    for (int v = 0; v < 16; v++) {
      const int row = row0 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
      float x = (t == 0 ? acc0[v] : acc1[v]) + bias;
      if (g.relu && x < 0.0f) x = 0.0f;
      if (g.mask && !(bf2f(g.mask[(size_t)row * g.ldm + col]) > 0.0f)) x = 0.0f;
      if (g.C32) g.C32[(size_t)row * g.ldc32 + col] = x;
      const uint16_t h = f2bf(x);
      if (g.C16) g.C16[(size_t)row * g.ldc16 + col] = h;
      if (g.C16T) g.C16T[(size_t)col * g.ldc16t + row] = h;
    }
  }
}

// Network inputs: the 16 pixels' RGBA / 255 in bf16, pixel-major
// (x[4 i + c]); rows past n (padding) are zero.
__global__ void __launch_bounds__(256) features_kernel(BlockRange r, const uint32_t *idx,
                                                       uint32_t n, uint32_t rows, uint16_t *X,
                                                       uint16_t *XT) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;  // row * 16 + pixel
  const uint32_t row = t >> 4, p = t & 15;
  if (row >= rows) return;
  uint32_t px = 0;
  if (row < n) {
    const uint32_t b = r.first_block + (idx ? idx[row] : row);
    const uint32_t bx = b % r.blocks_wide, by = b / r.blocks_wide;
    px = r.img[(size_t)(by * 4 + (p >> 2)) * r.width + bx * 4 + (p & 3)];
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const uint16_t h = row < n ? f2bf((float)((px >> (8 * c)) & 0xFFu) / 255.0f) : (uint16_t)0;
    X[(size_t)row * 64 + p * 4 + c] = h;
    if (XT) XT[(size_t)(p * 4 + c) * rows + row] = h;
  }
}

// Logit j = mode * 64 + slot is a real candidate (16 shapes for mode 0,
// 8 rotation/index-mode pairs for mode 4, 4 rotations for mode 5, one for
// mode 6, 64 shapes otherwise).
__device__ __forceinline__ bool logit_valid(int j) {
  const int mode = j >> 6, slot = j & 63;
  switch (mode) {
    case 0: return slot < 16;
    case 4: return slot < 8;
    case 5: return slot < 4;
    case 6: return slot == 0;
    default: return true;
  }
}

// Softmax cross-entropy over the valid logits, one wave per row:
// d = (softmax - onehot) * inv_b in bf16 (row-major and transposed) and
// the row's loss.  Rows >= n (padding) get zero gradient and loss.
__global__ void __launch_bounds__(64) xent_kernel(const float *logits, const int32_t *labels,
                                                  const uint32_t *idx, uint32_t n, uint32_t rows,
                                                  float inv_b, uint16_t *d, uint16_t *dT,
                                                  float *loss) {
  const uint32_t row = blockIdx.x;
  const int lane = threadIdx.x;
  if (row >= rows) return;
  float z[8];
  float mx = -INFINITY;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int j = lane + 64 * i;
    z[i] = logits[(size_t)row * 512 + j];
    if (logit_valid(j)) mx = fmaxf(mx, z[i]);
  }
  for (int off = 32; off >= 1; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; i++)
    if (logit_valid(lane + 64 * i)) s += expf(z[i] - mx);
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
  const bool real = row < n;
  const int label = real ? labels[idx ? idx[row] : row] : -1;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int j = lane + 64 * i;
    float p = logit_valid(j) ? expf(z[i] - mx) / s : 0.0f;
    if (j == label) p -= 1.0f;
    const uint16_t h = real ? f2bf(p * inv_b) : (uint16_t)0;
    d[(size_t)row * 512 + j] = h;
    dT[(size_t)j * rows + row] = h;
    if (j == label && loss) loss[row] = logf(s) + mx - z[i];
  }
  if (!real && lane == 0 && loss) loss[row] = 0.0f;
}

// db[j] = sum over rows of dT[j][row], one wave per j.
__global__ void __launch_bounds__(64) colsum_kernel(const uint16_t *dT, uint32_t rows,
                                                    float *db) {
  const uint32_t j = blockIdx.x;
  float s = 0.0f;
  for (uint32_t r = threadIdx.x; r < rows; r += 64) s += bf2f(dT[(size_t)j * rows + r]);
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
  if (threadIdx.x == 0) db[j] = s;
}

// Adam (Kingma & Ba, 2015) with bias correction c1 = 1 - beta1^t,
// c2 = 1 - beta2^t.
__global__ void __launch_bounds__(256) adam_kernel(float *w, float *m, float *v, const float *g,
                                                   uint32_t n, float lr, float beta1,
                                                   float beta2, float eps, float c1, float c2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
  const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 copies of one layer's fp32 weights W[in][out]: Wio = W, Woi = W^T.
__global__ void __launch_bounds__(256) cast_kernel(const float *W, int nin, int nout,
                                                   uint16_t *Wio, uint16_t *Woi) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nin * nout) return;
  const int i = t / nout, o = t % nout;
  const uint16_t h = f2bf(W[t]);
  Wio[t] = h;
  Woi[(size_t)o * nin + i] = h;
}



}  // namespace mlp

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

namespace {

hipError_t launch_gemm(const mlp::Gemm &g, hipStream_t s) {
  if (g.M % 32 || g.N % 32 || g.K % 16) return hipErrorInvalidValue;
  const dim3 grid((g.N + 63) / 64, (g.M + 127) / 128);
  hipLaunchKernelGGL(mlp::gemm_nt_kernel, grid, dim3(256), 0, s, g);
  return hipGetLastError();
}

mlp::Gemm gemm(const uint16_t *A, int lda, const uint16_t *Bt, int ldb, int M, int N, int K) {
  mlp::Gemm g = {};
  g.A = A;
  g.lda = lda;
  g.Bt = Bt;
  g.ldb = ldb;
  g.M = M;
  g.N = N;
  g.K = K;
  return g;
}

}  // namespace

constexpr int kIn = 64, kH = 256, kOut = 512;
constexpr size_t kOffW1 = 0, kOffB1 = kOffW1 + kIn * kH, kOffW2 = kOffB1 + kH,
                 kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kOut,
                 kNumParams = kOffB3 + kOut;
static_assert(mlp::kRankOffB1 == (int)kOffB1 && mlp::kRankOffB2 == (int)kOffB2 &&
                  mlp::kRankOffB3 == (int)kOffB3,
              "bias offsets of the inference tile");

size_t ranker_num_params() { return kNumParams; }

hipError_t Ranker::init(int dev) {
  device = dev;
  hipError_t e = hipMalloc(&master, kNumParams * 4 * 4);  // w, m, v, g
  if (e != hipSuccess) return e;
  m = master + kNumParams;
  v = m + kNumParams;
  grad = v + kNumParams;
  const size_t wsz = (size_t)(kIn * kH + kH * kH + kH * kOut);
  e = hipMalloc(&wbf, wsz * 2 * 2);
  if (e != hipSuccess) return e;
  w1_io = wbf;
  w2_io = w1_io + kIn * kH;
  w3_io = w2_io + kH * kH;
  w1_oi = w3_io + kH * kOut;
  w2_oi = w1_oi + kIn * kH;
  w3_oi = w2_oi + kH * kH;
  return hipMemset(m, 0, kNumParams * 4 * 3);
}

void Ranker::release() {
  if (master) (void)hipFree(master);
  if (wbf) (void)hipFree(wbf);
  if (work) (void)hipFree(work);
  master = nullptr;
  wbf = nullptr;
  work = nullptr;
}

hipError_t Ranker::refresh_bf16(hipStream_t s) {
  hipLaunchKernelGGL(mlp::cast_kernel, dim3((kIn * kH + 255) / 256), dim3(256), 0, s,
                     master + kOffW1, kIn, kH, w1_io, w1_oi);
  hipLaunchKernelGGL(mlp::cast_kernel, dim3((kH * kH + 255) / 256), dim3(256), 0, s,
                     master + kOffW2, kH, kH, w2_io, w2_oi);
  hipLaunchKernelGGL(mlp::cast_kernel, dim3((kH * kOut + 255) / 256), dim3(256), 0, s,
                     master + kOffW3, kH, kOut, w3_io, w3_oi);
  return hipGetLastError();
}

// Workspace for `rows` rows (a multiple of 32).
hipError_t Ranker::reserve(uint32_t rows, bool training) {
  const size_t per = (size_t)kIn * 2 * 2 + (size_t)kH * 2 * 4 + (size_t)kOut * (4 + 2 * 2) +
                     (size_t)kH * 2 * 4 + 4 + 16;
  const size_t need = per * rows + 64 * 256;
  if (need <= work_size && training <= work_training) return hipSuccess;
  if (work) (void)hipFree(work);
  work = nullptr;
  work_size = 0;
  hipError_t e = hipMalloc(&work, need);
  if (e != hipSuccess) return e;
  work_size = need;
  work_training = training;
  return hipSuccess;
}

hipError_t Ranker::forward(const BlockRange &r, const uint32_t *idx, uint32_t n, uint32_t rows,
                           bool training, hipStream_t s) {
  uint8_t *p = reinterpret_cast<uint8_t *>(work);
  auto carve = [&](size_t bytes) {
    uint8_t *q = p;
    p += (bytes + 255) & ~(size_t)255;
    return q;
  };
  X = reinterpret_cast<uint16_t *>(carve((size_t)rows * kIn * 2));
  XT = reinterpret_cast<uint16_t *>(carve((size_t)rows * kIn * 2));
  h1 = reinterpret_cast<uint16_t *>(carve((size_t)rows * kH * 2));
  h1T = reinterpret_cast<uint16_t *>(carve((size_t)rows * kH * 2));
  h2 = reinterpret_cast<uint16_t *>(carve((size_t)rows * kH * 2));
  h2T = reinterpret_cast<uint16_t *>(carve((size_t)rows * kH * 2));
  logits = reinterpret_cast<float *>(carve((size_t)rows * kOut * 4));
  dl = reinterpret_cast<uint16_t *>(carve((size_t)rows * kOut * 2));
  dlT = reinterpret_cast<uint16_t *>(carve((size_t)rows * kOut * 2));
  d2 = reinterpret_cast<uint16_t *>(carve((size_t)rows * kH * 2));
  d2T = reinterpret_cast<uint16_t *>(carve((size_t)rows * kH * 2));
  d1T = reinterpret_cast<uint16_t *>(carve((size_t)rows * kH * 2));
  loss = reinterpret_cast<float *>(carve((size_t)rows * 4));
  hipLaunchKernelGGL(mlp::features_kernel, dim3((rows * 16 + 255) / 256), dim3(256), 0, s, r,
                     idx, n, rows, X, training ? XT : (uint16_t *)nullptr);
  mlp::Gemm g = gemm(X, kIn, w1_oi, kIn, (int)rows, kH, kIn);
  g.bias = master + kOffB1;
  g.relu = 1;
  g.C16 = h1;
  g.ldc16 = kH;
  if (training) {
    g.C16T = h1T;
    g.ldc16t = (int)rows;
  }
  hipError_t e = launch_gemm(g, s);
  if (e != hipSuccess) return e;
  g = gemm(h1, kH, w2_oi, kH, (int)rows, kH, kH);
  g.bias = master + kOffB2;
  g.relu = 1;
  g.C16 = h2;
  g.ldc16 = kH;
  if (training) {
    g.C16T = h2T;
    g.ldc16t = (int)rows;
  }
  e = launch_gemm(g, s);
  if (e != hipSuccess) return e;
  g = gemm(h2, kH, w3_oi, kH, (int)rows, kOut, kH);
  g.bias = master + kOffB3;
  g.C32 = logits;
  g.ldc32 = kOut;
  return launch_gemm(g, s);
}


hipError_t Ranker::backward(const int32_t *labels, const uint32_t *idx, uint32_t n,
                            uint32_t rows, hipStream_t s) {
  hipLaunchKernelGGL(mlp::xent_kernel, dim3(rows), dim3(64), 0, s, logits, labels, idx, n, rows,
                     1.0f / (float)n, dl, dlT, loss);
  // layer 3
  mlp::Gemm g = gemm(h2T, (int)rows, dlT, (int)rows, kH, kOut, (int)rows);
  g.C32 = grad + kOffW3;
  g.ldc32 = kOut;
  hipError_t e = launch_gemm(g, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mlp::colsum_kernel, dim3(kOut), dim3(64), 0, s, dlT, rows,
                     grad + kOffB3);
  g = gemm(dl, kOut, w3_io, kOut, (int)rows, kH, kOut);
  g.mask = h2;
  g.ldm = kH;
  g.C16 = d2;
  g.ldc16 = kH;
  g.C16T = d2T;
  g.ldc16t = (int)rows;
  e = launch_gemm(g, s);
  if (e != hipSuccess) return e;
  // layer 2
  g = gemm(h1T, (int)rows, d2T, (int)rows, kH, kH, (int)rows);
  g.C32 = grad + kOffW2;
  g.ldc32 = kH;
  e = launch_gemm(g, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mlp::colsum_kernel, dim3(kH), dim3(64), 0, s, d2T, rows, grad + kOffB2);
  g = gemm(d2, kH, w2_io, kH, (int)rows, kH, kH);
  g.mask = h1;
  g.ldm = kH;
  g.C16T = d1T;
  g.ldc16t = (int)rows;
  e = launch_gemm(g, s);
  if (e != hipSuccess) return e;
  // layer 1
  g = gemm(XT, (int)rows, d1T, (int)rows, kIn, kH, (int)rows);
  g.C32 = grad + kOffW1;
  g.ldc32 = kH;
  e = launch_gemm(g, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mlp::colsum_kernel, dim3(kH), dim3(64), 0, s, d1T, rows, grad + kOffB1);
  return hipGetLastError();
}

hipError_t Ranker::adam(float lr, hipStream_t s) {
  step++;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = (float)(1.0 - pow((double)b1, step));
  const float c2 = (float)(1.0 - pow((double)b2, step));
  hipLaunchKernelGGL(mlp::adam_kernel, dim3((uint32_t)((kNumParams + 255) / 256)), dim3(256), 0,
                     s, master, m, v, grad, (uint32_t)kNumParams, lr, b1, b2, eps, c1, c2);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return refresh_bf16(s);
}

}  // namespace fh
