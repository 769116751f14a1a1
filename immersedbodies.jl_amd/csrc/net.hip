// Warm-start network of the CG surface-force solve: host side of include/ibpm.h's
// ibpm_net_* entry points and the hooks a solver step calls (net_features, net_predict,
// net_record). No counterpart in the reference, whose cg! starts from the previous F̃b
// (coupling.jl:194-202, :246-247).
//
// Parameters live in HBM as FP32 master copies, per layer [W (cout x K x cin) | b (cout)]
// with channel counts padded to multiples of 32 (padded entries stay exactly zero: their
// gradients are zero). BF16 copies W and Wt (flipped, transposed) feed the MFMA kernels.
// A training step gathers a minibatch from the training set, runs forward, loss and
// backward kernels and an Adam update, all queued on the network's stream without host
// synchronisation; the minibatch indices of a whole ibpm_net_train call are drawn on the
// host and uploaded once.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/ibpm.h"
#include "ibpm_internal.h"

using namespace ibpm;

struct ibpm_net {
  ibpm_net_config cfg{};
  int L = 0, K = 0, C = 0;
  std::vector<int> cin, cout, cin_real, cout_real;
  std::vector<size_t> poff;  // layer offsets in p/g/m/v (floats)
  std::vector<size_t> woff;  // layer offsets in wb/wt (bf16)
  size_t np = 0, nw = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *wb = nullptr, *wt = nullptr;
  float out_scale = 0.f;
  int adam_t = 0;
  int maxRpad = 0, rows_per_chunk = 128;
  std::vector<uint16_t*> act;  // input rows of each layer
  uint16_t *dz0 = nullptr, *dz1 = nullptr;
  float* out = nullptr;
  float* tgt = nullptr;
  int* seg = nullptr;
  float* part = nullptr;
  double* lpart = nullptr;
  double* loss = nullptr;
  int loss_cap = 0;
  float* hx = nullptr;  // host rows staged on the device
  int* hseg = nullptr;
  float* feat = nullptr;  // features of the current solver step
  float *ds_feat = nullptr, *ds_tgt = nullptr;
  int *ds_body = nullptr, *ds_count = nullptr;
  std::vector<int> count;
  int held = 0, next = 0;
  int* bidx = nullptr;
  int* boff = nullptr;
  size_t bcap = 0;
  double* red = nullptr;  // reduction partials (output scale)
  hipStream_t st = nullptr;
  uint64_t rng = 0;
};

namespace {

#define NCHK(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return set_error(IBPM_E_HIP, (std::string(#expr) + ": " + hipGetErrorString(e_)).c_str()); \
  } while (0)

int argfail(const char* msg) { return set_error(IBPM_E_ARG, msg); }

int ceil_to(int x, int q) { return (x + q - 1) / q * q; }

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

template <typename T>
hipError_t nalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e == hipSuccess) e = hipMemset(*p, 0, n * sizeof(T));
  return e;
}

void net_free(ibpm_net* n) {
  void* ptrs[] = {n->p, n->g, n->m, n->v, n->wb, n->wt, n->dz0, n->dz1, n->out, n->tgt, n->seg,
                  n->part, n->lpart, n->loss, n->hx, n->hseg, n->feat, n->ds_feat, n->ds_tgt,
                  n->ds_body, n->ds_count, n->bidx, n->boff, n->red};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  for (uint16_t* a : n->act)
    if (a) (void)hipFree(a);
  if (n->st) (void)hipStreamDestroy(n->st);
}

int cast_all(ibpm_net* n, hipStream_t st) {
  for (int l = 0; l < n->L; ++l)
    NCHK(launch_cast_weights(n->p + n->poff[l], n->cout[l], n->K, n->cin[l], n->wb + n->woff[l],
                             n->wt + n->woff[l], st));
  return IBPM_OK;
}

const float* bias_of(const ibpm_net* n, int l) {
  return n->p + n->poff[l] + (size_t)n->cout[l] * n->K * n->cin[l];
}

// Forward over R rows of act[0] / seg (Rpad rows allocated).
int forward(ibpm_net* n, int R, int Rpad, hipStream_t st) {
  for (int l = 0; l < n->L; ++l) {
    const bool last = l == n->L - 1;
    ConvArgs a;
    a.X = n->act[l];
    a.Cin = n->cin[l];
    a.W = n->wb + n->woff[l];
    a.Cout = n->cout[l];
    a.K = n->K;
    a.bias = bias_of(n, l);
    a.seg = n->seg;
    a.R = R;
    a.Rpad = Rpad;
    a.Y = last ? (void*)n->out : (void*)n->act[l + 1];
    a.mask = nullptr;
    NCHK(launch_conv(a, last ? CONV_LINEAR_F32 : CONV_RELU, st));
  }
  return IBPM_OK;
}

// Loss into *loss_slot and parameter gradients into g, after forward (targets in n->tgt).
int backward(ibpm_net* n, int R, int Rpad, double* loss_slot, hipStream_t st) {
  uint16_t* dz = n->dz0;
  uint16_t* dzn = n->dz1;
  NCHK(launch_mse(n->out, kNetOutPad, n->tgt, R, Rpad, dz, n->lpart, loss_slot, st));
  for (int l = n->L - 1; l >= 0; --l) {
    WgradArgs w;
    w.dY = dz;
    w.Cout = n->cout[l];
    w.X = n->act[l];
    w.Cin = n->cin[l];
    w.seg = n->seg;
    w.R = R;
    w.K = n->K;
    w.rows_per_chunk = n->rows_per_chunk;
    w.part = n->part;
    NCHK(launch_wgrad(w, n->g + n->poff[l], st));
    if (l > 0) {
      ConvArgs a;
      a.X = dz;
      a.Cin = n->cout[l];
      a.W = n->wt + n->woff[l];
      a.Cout = n->cin[l];
      a.K = n->K;
      a.bias = nullptr;
      a.seg = n->seg;
      a.R = R;
      a.Rpad = Rpad;
      a.Y = dzn;
      a.mask = n->act[l];
      NCHK(launch_conv(a, CONV_MASK, st));
      std::swap(dz, dzn);
    }
  }
  return IBPM_OK;
}

int adam_step(ibpm_net* n, double lr, hipStream_t st) {
  NCHK(launch_adam(n->p, n->g, n->m, n->v, n->np, (float)lr, 0.9f, 0.999f, 1e-8f, ++n->adam_t, st));
  return cast_all(n, st);
}

// Host rows (x: R x kNetFeatures, seg, target R x 2 or null) into the device row buffers.
int stage_rows(ibpm_net* n, const float* x, const int* seg, const float* target, int R, int* Rpad) {
  if (!x || !seg || R < 1) return argfail("ibpm_net: rows, features and sequence ids required");
  if ((size_t)R > (size_t)n->cfg.max_batch * n->cfg.max_points)
    return argfail("ibpm_net: more rows than max_batch * max_points");
  *Rpad = ceil_to(R, 64);
  NCHK(hipMemcpyAsync(n->hx, x, (size_t)R * kNetFeatures * sizeof(float), hipMemcpyHostToDevice, n->st));
  NCHK(hipMemcpyAsync(n->hseg, seg, (size_t)R * sizeof(int), hipMemcpyHostToDevice, n->st));
  if (target)
    NCHK(hipMemcpyAsync(n->tgt, target, (size_t)R * 2 * sizeof(float), hipMemcpyHostToDevice, n->st));
  NCHK(launch_pack_rows(n->hx, n->hseg, R, *Rpad, n->act[0], n->seg, n->st));
  return IBPM_OK;
}

void to_public(const ibpm_net* n, const std::vector<float>& dev, float* out) {
  size_t o = 0;
  for (int l = 0; l < n->L; ++l) {
    const int co_r = n->cout_real[l], ci_r = n->cin_real[l], ci = n->cin[l], K = n->K;
    const float* w = dev.data() + n->poff[l];
    for (int co = 0; co < co_r; ++co)
      for (int c = 0; c < ci_r; ++c)
        for (int k = 0; k < K; ++k) out[o++] = w[((size_t)co * K + k) * ci + c];
    const float* b = w + (size_t)n->cout[l] * K * ci;
    for (int co = 0; co < co_r; ++co) out[o++] = b[co];
  }
}

void from_public(const ibpm_net* n, const float* in, std::vector<float>& dev) {
  std::fill(dev.begin(), dev.end(), 0.f);
  size_t o = 0;
  for (int l = 0; l < n->L; ++l) {
    const int co_r = n->cout_real[l], ci_r = n->cin_real[l], ci = n->cin[l], K = n->K;
    float* w = dev.data() + n->poff[l];
    for (int co = 0; co < co_r; ++co)
      for (int c = 0; c < ci_r; ++c)
        for (int k = 0; k < K; ++k) w[((size_t)co * K + k) * ci + c] = in[o++];
    float* b = w + (size_t)n->cout[l] * K * ci;
    for (int co = 0; co < co_r; ++co) b[co] = in[o++];
  }
}

size_t public_count(const ibpm_net* n) {
  size_t c = 0;
  for (int l = 0; l < n->L; ++l)
    c += (size_t)n->cout_real[l] * n->cin_real[l] * n->K + n->cout_real[l];
  return c + 1;
}

}  // namespace

namespace ibpm {

int net_device(const ibpm_net* n) { return n->cfg.device; }
int net_max_points(const ibpm_net* n) { return n->cfg.max_points; }

int net_features(ibpm_net* n, WsArgs& a, hipStream_t st) {
  a.feat = n->feat;
  a.A0 = n->act[0];
  a.seg = n->seg;
  a.Rpad = ceil_to(a.n, 64);
  NCHK(launch_ws_features(a, st));
  return IBPM_OK;
}

int net_predict(ibpm_net* n, const WsArgs& a, double* x0, hipStream_t st) {
  if (!(n->out_scale > 0.f))
    return argfail("IBPM_WS_LEARNED needs a trained network (ibpm_net_train or ibpm_net_set_params)");
  int rc = forward(n, a.n, a.Rpad, st);
  if (rc) return rc;
  NCHK(launch_ws_apply(a.n, a.xe, n->out, kNetOutPad, a.scal, n->out_scale, x0, st));
  return IBPM_OK;
}

int net_record(ibpm_net* n, const WsArgs& a, hipStream_t st) {
  const int stride = n->cfg.max_points;
  const int slot = n->next;
  n->next = (n->next + 1) % n->cfg.capacity;
  n->held = std::min(n->held + 1, n->cfg.capacity);
  n->count[slot] = a.n;
  NCHK(launch_ws_record(a, stride, n->ds_feat + (size_t)slot * stride * kNetFeatures,
                        n->ds_tgt + (size_t)slot * stride * 2, n->ds_body + (size_t)slot * stride,
                        st));
  return IBPM_OK;
}

}  // namespace ibpm

extern "C" {

void ibpm_net_default_config(ibpm_net_config* c) {
  if (!c) return;
  c->channels = 128;
  c->layers = 4;
  c->kernel = 5;
  c->max_points = 1024;
  c->capacity = 20000;
  c->max_batch = 64;
  c->seed = 1234;
  c->device = 0;
}

int ibpm_net_create(const ibpm_net_config* c, ibpm_net** out) {
  if (!c || !out) return argfail("null argument");
  *out = nullptr;
  if (c->channels < 32 || c->channels % 32) return argfail("ibpm_net: channels must be a positive multiple of 32");
  if (c->layers < 2) return argfail("ibpm_net: at least 2 layers");
  if (c->kernel != 3 && c->kernel != 5 && c->kernel != 7) return argfail("ibpm_net: kernel must be 3, 5 or 7");
  if (c->max_points < 1 || c->capacity < 1 || c->max_batch < 1)
    return argfail("ibpm_net: max_points, capacity and max_batch must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_error(IBPM_E_NO_DEVICE, "no HIP device");
  if (c->device < 0 || c->device >= ndev) return argfail("ibpm_net: bad device ordinal");
  NCHK(hipSetDevice(c->device));
  ibpm_net* n = new ibpm_net;
  n->cfg = *c;
  n->L = c->layers;
  n->K = c->kernel;
  n->C = c->channels;
  n->rng = c->seed ^ 0x5DEECE66Dull;
  for (int l = 0; l < n->L; ++l) {
    n->cin_real.push_back(l == 0 ? kNetFeatures : n->C);
    n->cout_real.push_back(l == n->L - 1 ? IBPM_NET_OUTPUTS : n->C);
    n->cin.push_back(l == 0 ? kNetInPad : n->C);
    n->cout.push_back(l == n->L - 1 ? kNetOutPad : n->C);
    n->poff.push_back(n->np);
    n->woff.push_back(n->nw);
    const size_t w = (size_t)n->cout[l] * n->K * n->cin[l];
    n->np += w + n->cout[l];
    n->nw += w;
  }
  const size_t maxR = (size_t)c->max_batch * c->max_points;
  n->maxRpad = ceil_to((int)std::min<size_t>(maxR, 1u << 30), 64);
  n->rows_per_chunk = std::max(128, ceil_to((n->maxRpad + 255) / 256, 32));
  const int nchunk = wgrad_chunks(n->maxRpad, n->rows_per_chunk);
  size_t pstride = 0;
  for (int l = 0; l < n->L; ++l)
    pstride = std::max(pstride, (size_t)n->cout[l] * n->K * n->cin[l] + n->cout[l]);
  const int wide = std::max(n->C, kNetOutPad);
  const size_t stride = c->max_points;
  const size_t cap = c->capacity;
  hipError_t e = hipStreamCreateWithFlags(&n->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = nalloc(&n->p, n->np);
  if (e == hipSuccess) e = nalloc(&n->g, n->np);
  if (e == hipSuccess) e = nalloc(&n->m, n->np);
  if (e == hipSuccess) e = nalloc(&n->v, n->np);
  if (e == hipSuccess) e = nalloc(&n->wb, n->nw);
  if (e == hipSuccess) e = nalloc(&n->wt, n->nw);
  for (int l = 0; l < n->L && e == hipSuccess; ++l) {
    uint16_t* a = nullptr;
    e = nalloc(&a, (size_t)n->maxRpad * n->cin[l]);
    n->act.push_back(a);
  }
  if (e == hipSuccess) e = nalloc(&n->dz0, (size_t)n->maxRpad * wide);
  if (e == hipSuccess) e = nalloc(&n->dz1, (size_t)n->maxRpad * wide);
  if (e == hipSuccess) e = nalloc(&n->out, (size_t)n->maxRpad * kNetOutPad);
  if (e == hipSuccess) e = nalloc(&n->tgt, (size_t)n->maxRpad * 2);
  if (e == hipSuccess) e = nalloc(&n->seg, (size_t)n->maxRpad);
  if (e == hipSuccess) e = nalloc(&n->part, (size_t)nchunk * pstride);
  if (e == hipSuccess) e = nalloc(&n->lpart, (size_t)mse_blocks(n->maxRpad, kNetOutPad));
  if (e == hipSuccess) e = nalloc(&n->loss, 1);
  n->loss_cap = 1;
  if (e == hipSuccess) e = nalloc(&n->hx, (size_t)n->maxRpad * kNetFeatures);
  if (e == hipSuccess) e = nalloc(&n->hseg, (size_t)n->maxRpad);
  if (e == hipSuccess) e = nalloc(&n->feat, stride * kNetFeatures);
  if (e == hipSuccess) e = nalloc(&n->ds_feat, cap * stride * kNetFeatures);
  if (e == hipSuccess) e = nalloc(&n->ds_tgt, cap * stride * 2);
  if (e == hipSuccess) e = nalloc(&n->ds_body, cap * stride);
  if (e == hipSuccess) e = nalloc(&n->ds_count, cap);
  if (e == hipSuccess) e = nalloc(&n->red, 256);
  if (e != hipSuccess) {
    net_free(n);
    delete n;
    return set_error(IBPM_E_HIP, (std::string("ibpm_net_create: ") + hipGetErrorString(e)).c_str());
  }
  n->count.assign(cap, 0);
  // He-uniform initialisation (ReLU layers), LeCun-uniform for the output layer; zero bias
  std::vector<float> host(n->np, 0.f);
  for (int l = 0; l < n->L; ++l) {
    const double fan_in = (double)n->cin_real[l] * n->K;
    const double bound = sqrt((l == n->L - 1 ? 3.0 : 6.0) / fan_in);
    float* w = host.data() + n->poff[l];
    for (int co = 0; co < n->cout_real[l]; ++co)
      for (int k = 0; k < n->K; ++k)
        for (int ci = 0; ci < n->cin_real[l]; ++ci)
          w[((size_t)co * n->K + k) * n->cin[l] + ci] = (float)((2.0 * uniform01(n->rng) - 1.0) * bound);
  }
  e = hipMemcpy(n->p, host.data(), n->np * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess || cast_all(n, n->st) != IBPM_OK || hipStreamSynchronize(n->st) != hipSuccess) {
    net_free(n);
    delete n;
    return set_error(IBPM_E_HIP, "ibpm_net_create: initialisation failed");
  }
  *out = n;
  return IBPM_OK;
}

void ibpm_net_destroy(ibpm_net* n) {
  if (!n) return;
  (void)hipSetDevice(n->cfg.device);
  (void)hipStreamSynchronize(n->st);
  net_free(n);
  delete n;
}

int ibpm_net_dataset_size(const ibpm_net* n) { return n ? n->held : 0; }

int ibpm_net_clear_dataset(ibpm_net* n) {
  if (!n) return argfail("null network");
  n->held = 0;
  n->next = 0;
  std::fill(n->count.begin(), n->count.end(), 0);
  return IBPM_OK;
}

size_t ibpm_net_param_count(const ibpm_net* n) { return n ? public_count(n) : 0; }

int ibpm_net_get_params(const ibpm_net* n, float* p, size_t cnt) {
  if (!n || !p) return argfail("null argument");
  if (cnt != public_count(n)) return argfail("ibpm_net_get_params: wrong parameter count");
  NCHK(hipSetDevice(n->cfg.device));
  NCHK(hipStreamSynchronize(n->st));
  std::vector<float> dev(n->np);
  NCHK(hipMemcpy(dev.data(), n->p, n->np * sizeof(float), hipMemcpyDeviceToHost));
  to_public(n, dev, p);
  p[cnt - 1] = n->out_scale;
  return IBPM_OK;
}

int ibpm_net_set_params(ibpm_net* n, const float* p, size_t cnt) {
  if (!n || !p) return argfail("null argument");
  if (cnt != public_count(n)) return argfail("ibpm_net_set_params: wrong parameter count");
  NCHK(hipSetDevice(n->cfg.device));
  NCHK(hipStreamSynchronize(n->st));
  std::vector<float> dev(n->np);
  from_public(n, p, dev);
  NCHK(hipMemcpy(n->p, dev.data(), n->np * sizeof(float), hipMemcpyHostToDevice));
  NCHK(hipMemset(n->m, 0, n->np * sizeof(float)));
  NCHK(hipMemset(n->v, 0, n->np * sizeof(float)));
  n->adam_t = 0;
  n->out_scale = p[cnt - 1];
  int rc = cast_all(n, n->st);
  if (rc) return rc;
  NCHK(hipStreamSynchronize(n->st));
  return IBPM_OK;
}

int ibpm_net_forward(ibpm_net* n, const float* x, const int* seg, int rows, float* y) {
  if (!n || !y) return argfail("null argument");
  NCHK(hipSetDevice(n->cfg.device));
  int Rpad = 0;
  int rc = stage_rows(n, x, seg, nullptr, rows, &Rpad);
  if (rc) return rc;
  rc = forward(n, rows, Rpad, n->st);
  if (rc) return rc;
  std::vector<float> o((size_t)rows * kNetOutPad);
  NCHK(hipMemcpyAsync(o.data(), n->out, o.size() * sizeof(float), hipMemcpyDeviceToHost, n->st));
  NCHK(hipStreamSynchronize(n->st));
  for (int r = 0; r < rows; ++r) {
    y[2 * r] = o[(size_t)r * kNetOutPad];
    y[2 * r + 1] = o[(size_t)r * kNetOutPad + 1];
  }
  return IBPM_OK;
}

static int grads_on_rows(ibpm_net* n, const float* x, const int* seg, const float* target, int rows,
                         double* loss) {
  if (!target) return argfail("ibpm_net: targets required");
  int Rpad = 0;
  int rc = stage_rows(n, x, seg, target, rows, &Rpad);
  if (rc) return rc;
  rc = forward(n, rows, Rpad, n->st);
  if (rc) return rc;
  rc = backward(n, rows, Rpad, n->loss, n->st);
  if (rc) return rc;
  if (loss) {
    NCHK(hipMemcpyAsync(loss, n->loss, sizeof(double), hipMemcpyDeviceToHost, n->st));
  }
  return IBPM_OK;
}

int ibpm_net_gradients(ibpm_net* n, const float* x, const int* seg, const float* target, int rows,
                       float* grad, size_t cnt, double* loss) {
  if (!n || !grad) return argfail("null argument");
  if (cnt != public_count(n)) return argfail("ibpm_net_gradients: wrong parameter count");
  NCHK(hipSetDevice(n->cfg.device));
  int rc = grads_on_rows(n, x, seg, target, rows, loss);
  if (rc) return rc;
  std::vector<float> dev(n->np);
  NCHK(hipMemcpyAsync(dev.data(), n->g, n->np * sizeof(float), hipMemcpyDeviceToHost, n->st));
  NCHK(hipStreamSynchronize(n->st));
  to_public(n, dev, grad);
  grad[cnt - 1] = 0.f;
  return IBPM_OK;
}

int ibpm_net_step(ibpm_net* n, const float* x, const int* seg, const float* target, int rows,
                  double lr, double* loss) {
  if (!n) return argfail("null network");
  NCHK(hipSetDevice(n->cfg.device));
  int rc = grads_on_rows(n, x, seg, target, rows, loss);
  if (rc) return rc;
  rc = adam_step(n, lr, n->st);
  if (rc) return rc;
  NCHK(hipStreamSynchronize(n->st));
  return IBPM_OK;
}

int ibpm_net_train(ibpm_net* n, int iters, int batch, double lr, double* loss) {
  if (!n) return argfail("null network");
  if (iters < 0 || batch < 1 || batch > n->cfg.max_batch)
    return argfail("ibpm_net_train: need iters >= 0 and 1 <= batch <= max_batch");
  if (n->held < 1) return argfail("ibpm_net_train: the training set is empty");
  if (iters == 0) return IBPM_OK;
  NCHK(hipSetDevice(n->cfg.device));
  const int stride = n->cfg.max_points;
  NCHK(hipMemcpyAsync(n->ds_count, n->count.data(), n->count.size() * sizeof(int),
                      hipMemcpyHostToDevice, n->st));
  if (!(n->out_scale > 0.f)) {
    // output scale: rms of the held targets (rows past each sample's count are zero)
    const size_t cnt = (size_t)n->held * stride * 2;
    NCHK(launch_sumsq(n->ds_tgt, cnt, n->red, 256, n->st));
    double part[256];
    NCHK(hipMemcpyAsync(part, n->red, sizeof part, hipMemcpyDeviceToHost, n->st));
    NCHK(hipStreamSynchronize(n->st));
    double s = 0.0;
    for (double q : part) s += q;
    size_t rows = 0;
    for (int i = 0; i < n->held; ++i) rows += n->count[i];
    const double rms = rows ? sqrt(s / (2.0 * rows)) : 0.0;
    n->out_scale = rms > 0.0 ? (float)rms : 1.f;
  }
  // minibatches of the whole call
  std::vector<int> idx((size_t)iters * batch), off((size_t)iters * (batch + 1));
  std::vector<int> R(iters);
  for (int it = 0; it < iters; ++it) {
    int r = 0;
    for (int s = 0; s < batch; ++s) {
      const int j = (int)(uniform01(n->rng) * n->held) % n->held;
      idx[(size_t)it * batch + s] = j;
      off[(size_t)it * (batch + 1) + s] = r;
      r += n->count[j];
    }
    off[(size_t)it * (batch + 1) + batch] = r;
    R[it] = r;
    if (ceil_to(std::max(r, 1), 64) > n->maxRpad) return argfail("ibpm_net_train: batch too large");
  }
  if (n->bcap < off.size()) {
    if (n->bidx) (void)hipFree(n->bidx);
    if (n->boff) (void)hipFree(n->boff);
    n->bidx = nullptr;
    n->boff = nullptr;
    NCHK(nalloc(&n->bidx, off.size()));
    NCHK(nalloc(&n->boff, off.size()));
    n->bcap = off.size();
  }
  if (n->loss_cap < iters) {
    (void)hipFree(n->loss);
    n->loss = nullptr;
    NCHK(nalloc(&n->loss, (size_t)iters));
    n->loss_cap = iters;
  }
  NCHK(hipMemcpyAsync(n->bidx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, n->st));
  NCHK(hipMemcpyAsync(n->boff, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, n->st));
  for (int it = 0; it < iters; ++it) {
    const int Rr = R[it];
    if (Rr < 1) continue;
    const int Rpad = ceil_to(Rr, 64);
    GatherBatchArgs ga;
    ga.feat = n->ds_feat;
    ga.tgt = n->ds_tgt;
    ga.body = n->ds_body;
    ga.counts = n->ds_count;
    ga.stride = stride;
    ga.idx = n->bidx + (size_t)it * batch;
    ga.off = n->boff + (size_t)it * (batch + 1);
    ga.S = batch;
    ga.R = Rr;
    ga.Rpad = Rpad;
    ga.inv_scale = 1.f / n->out_scale;
    ga.A0 = n->act[0];
    ga.tgt_out = n->tgt;
    ga.seg = n->seg;
    NCHK(launch_gather_batch(ga, stride, n->st));
    int rc = forward(n, Rr, Rpad, n->st);
    if (!rc) rc = backward(n, Rr, Rpad, n->loss + it, n->st);
    const double lr_it = lr * (0.1 + 0.45 * (1.0 + cos(M_PI * it / (double)iters)));
    if (!rc) rc = adam_step(n, lr_it, n->st);
    if (rc) return rc;
  }
  if (loss) NCHK(hipMemcpyAsync(loss, n->loss, (size_t)iters * sizeof(double), hipMemcpyDeviceToHost, n->st));
  NCHK(hipStreamSynchronize(n->st));
  return IBPM_OK;
}

}  // extern "C"
