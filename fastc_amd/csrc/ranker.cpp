// C-ABI entry points of the BC7 candidate ranker (include/fasthip.h):
// create / initialise / load / store the weights, forward logits, gradients
// for checks, and the training driver (epochs of shuffled batches, one Adam
// step per batch).  No counterpart in the reference.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/fasthip.h"
#include "internal.h"

namespace fh {
namespace {

fh_status rfail(fh_status s, const char *msg) {
  set_error(msg);
  return s;
}

fh_status rfail_hip(hipError_t e, const char *what) {
  char buf[256];
  snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
  set_error(buf);
  return FH_ERR_HIP;
}

uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double uniform(uint64_t &x) { return (double)(splitmix64(x) >> 11) * (1.0 / 9007199254740992.0); }

// rows padded to the product tiles
uint32_t pad_rows(uint32_t n) { return (n + 127u) & ~127u; }

// blocks: n x 16 pixels, seen as a 4 x 4n image one block wide
BlockRange block_array(const uint32_t *blocks, uint32_t n) {
  BlockRange r;
  r.img = blocks;
  r.out = nullptr;
  r.width = 4;
  r.blocks_wide = 1;
  r.first_block = 0;
  r.num_blocks = n;
  return r;
}

}  // namespace
}  // namespace fh

using namespace fh;

extern "C" {

fh_status fh_ranker_create(fh_ranker **out) {
  if (!out) return rfail(FH_ERR_INVALID_ARGUMENT, "null output");
  *out = nullptr;
  int n = 0, dev = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return rfail(FH_ERR_NO_DEVICE, "no HIP device available");
  if (hipGetDevice(&dev) != hipSuccess) return rfail(FH_ERR_NO_DEVICE, "hipGetDevice failed");
  fh_ranker *r = new fh_ranker();
  hipError_t e = r->net.init(dev);
  if (e != hipSuccess) {
    r->net.release();
    delete r;
    return rfail_hip(e, "ranker allocation");
  }
  *out = r;
  return FH_OK;
}

fh_status fh_ranker_destroy(fh_ranker *r) {
  if (!r) return FH_OK;
  r->net.release();
  delete r;
  return FH_OK;
}

size_t fh_ranker_num_params(void) { return ranker_num_params(); }

fh_status fh_ranker_set_params(fh_ranker *r, const float *params, size_t count) {
  if (!r || !params) return rfail(FH_ERR_INVALID_ARGUMENT, "null argument");
  if (count != ranker_num_params()) return rfail(FH_ERR_INVALID_ARGUMENT, "wrong parameter count");
  std::lock_guard<std::mutex> g(r->mu);
  hipError_t e = hipMemcpy(r->net.master, params, count * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(r->net.m, 0, count * 4 * 2);
  if (e == hipSuccess) e = r->net.refresh_bf16(nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return rfail_hip(e, "ranker parameter upload");
  r->net.step = 0;
  return FH_OK;
}

fh_status fh_ranker_get_params(fh_ranker *r, float *params, size_t count) {
  if (!r || !params) return rfail(FH_ERR_INVALID_ARGUMENT, "null argument");
  if (count != ranker_num_params()) return rfail(FH_ERR_INVALID_ARGUMENT, "wrong parameter count");
  std::lock_guard<std::mutex> g(r->mu);
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(params, r->net.master, count * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return rfail_hip(e, "ranker parameter download");
  return FH_OK;
}

fh_status fh_ranker_init(fh_ranker *r, uint64_t seed) {
  if (!r) return rfail(FH_ERR_INVALID_ARGUMENT, "null ranker");
  // He-uniform weights U(-sqrt(6/fan_in), +sqrt(6/fan_in)), zero biases
  std::vector<float> p(ranker_num_params(), 0.0f);
  uint64_t s = seed;
  const int dims[4] = {64, 256, 256, 512};
  size_t off = 0;
  for (int l = 0; l < 3; l++) {
    const double a = sqrt(6.0 / dims[l]);
    const size_t nw = (size_t)dims[l] * dims[l + 1];
    for (size_t i = 0; i < nw; i++) p[off + i] = (float)((2.0 * uniform(s) - 1.0) * a);
    off += nw + dims[l + 1];
  }
  return fh_ranker_set_params(r, p.data(), p.size());
}

fh_status fh_ranker_forward(fh_ranker *r, const uint32_t *blocks, uint32_t n, float *logits,
                            void *hip_stream) {
  if (!r || !blocks || !logits) return rfail(FH_ERR_INVALID_ARGUMENT, "null argument");
  if (n == 0) return FH_OK;
  std::lock_guard<std::mutex> g(r->mu);
  hipStream_t s = reinterpret_cast<hipStream_t>(hip_stream);
  const uint32_t rows = pad_rows(n);
  hipError_t e = r->net.reserve(rows, false);
  if (e == hipSuccess) e = r->net.forward(block_array(blocks, n), nullptr, n, rows, false, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(logits, r->net.logits, (size_t)n * 512 * 4, hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return rfail_hip(e, "ranker forward");
  return FH_OK;
}

fh_status fh_ranker_gradients(fh_ranker *r, const uint32_t *blocks, const int32_t *labels,
                              uint32_t n, float *grad, double *loss) {
  if (!r || !blocks || !labels || !grad) return rfail(FH_ERR_INVALID_ARGUMENT, "null argument");
  if (n == 0) return rfail(FH_ERR_NO_DATA, "empty batch");
  std::lock_guard<std::mutex> g(r->mu);
  const uint32_t rows = pad_rows(n);
  hipError_t e = r->net.reserve(rows, true);
  if (e == hipSuccess) e = r->net.forward(block_array(blocks, n), nullptr, n, rows, true, nullptr);
  if (e == hipSuccess) e = r->net.backward(labels, nullptr, n, rows, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess)
    e = hipMemcpy(grad, r->net.grad, ranker_num_params() * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && loss) {
    std::vector<float> l(n);
    e = hipMemcpy(l.data(), r->net.loss, (size_t)n * 4, hipMemcpyDeviceToHost);
    double t = 0.0;
    for (float x : l) t += x;
    *loss = t / n;
  }
  if (e != hipSuccess) return rfail_hip(e, "ranker gradients");
  return FH_OK;
}

fh_status fh_ranker_train(fh_ranker *r, const uint32_t *blocks, const int32_t *labels, uint32_t n,
                          const fh_train_config *cfg, fh_train_report *rep) {
  if (!r || !blocks || !labels || !cfg) return rfail(FH_ERR_INVALID_ARGUMENT, "null argument");
  if (n == 0 || cfg->batch == 0 || cfg->epochs == 0)
    return rfail(FH_ERR_INVALID_ARGUMENT, "empty training set, batch or epoch count");
  if (cfg->batch > n) return rfail(FH_ERR_INVALID_ARGUMENT, "batch larger than the training set");
  std::lock_guard<std::mutex> g(r->mu);
  hipStream_t s = nullptr;
  const uint32_t batch = cfg->batch;
  const uint32_t rows = pad_rows(batch);
  hipError_t e = r->net.reserve(rows, true);
  uint32_t *d_perm = nullptr;
  if (e == hipSuccess) e = hipMalloc(&d_perm, (size_t)n * 4);
  if (e != hipSuccess) return rfail_hip(e, "training workspace");
  std::vector<uint32_t> perm(n);
  std::vector<float> lrow(batch);
  uint64_t rs = cfg->shuffle_seed;
  const uint32_t steps_per_epoch = n / batch;  // the tail of each permutation is dropped
  hipEvent_t t0, t1;
  (void)hipEventCreate(&t0);
  (void)hipEventCreate(&t1);
  (void)hipEventRecord(t0, s);
  uint32_t total_steps = 0;
  for (uint32_t ep = 0; ep < cfg->epochs && e == hipSuccess; ep++) {
    for (uint32_t i = 0; i < n; i++) perm[i] = i;
    for (uint32_t i = n - 1; i > 0; i--) {  // Fisher-Yates
      const uint32_t j = (uint32_t)(splitmix64(rs) % (uint64_t)(i + 1));
      const uint32_t t = perm[i];
      perm[i] = perm[j];
      perm[j] = t;
    }
    e = hipMemcpyAsync(d_perm, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, s);
    double ep_loss = 0.0;
    for (uint32_t st = 0; st < steps_per_epoch && e == hipSuccess; st++) {
      const uint32_t *idx = d_perm + (size_t)st * batch;
      e = r->net.forward(block_array(blocks, n), idx, batch, rows, true, s);
      if (e == hipSuccess) e = r->net.backward(labels, idx, batch, rows, s);
      if (e == hipSuccess) e = r->net.adam(cfg->lr, s);
      if (e == hipSuccess && rep && (st % 64 == 0 || st + 1 == steps_per_epoch)) {
        e = hipMemcpyAsync(lrow.data(), r->net.loss, (size_t)batch * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        double t = 0.0;
        for (float x : lrow) t += x;
        ep_loss = t / batch;
        if (st == 0 && ep == 0) rep->first_loss = ep_loss;
      }
      total_steps++;
    }
    if (rep && ep < 16) rep->epoch_loss[ep] = ep_loss;
    // the next permutation overwrites the host array the copy reads
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  (void)hipEventRecord(t1, s);
  if (e == hipSuccess) e = hipEventSynchronize(t1);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, t0, t1);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  (void)hipFree(d_perm);
  if (e != hipSuccess) return rfail_hip(e, "training step");
  if (rep) {
    rep->steps = total_steps;
    rep->ms = ms;
  }
  return FH_OK;
}

}  // extern "C"
