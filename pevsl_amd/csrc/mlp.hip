// mlp.hip -- the learned slicer's network: a fully connected MLP with GELU
// hidden layers, trained and evaluated by hand-written HIP kernels in FP32.
// It has no counterpart in the reference.
//
// Layout.  Layer l maps din[l] -> dout[l] with W_l (dout x din, row-major)
// and b_l; the parameters are one packed vector [W_0, b_0, W_1, b_1, ...].
// Activations are row-major (batch x width).
//
// Output head.  The network output z (nout values per sample) is added to a
// per-sample prior q before the loss: y = z + q.  The first nsoft values of y
// form a softmax group, the remaining nout - nsoft are plain regressions.
// Loss per sample = -sum_i t_i log softmax(y)_i  +  sum_r (y_r - t_r)^2, and
// the batch loss is its mean, so d loss / d y = (p - t)/batch on the softmax
// group and 2 (y - t)/batch on the regressions.
//
// Optimizer: Adam (Kingma & Ba 2015, beta1 0.9, beta2 0.999, eps 1e-8) with
// a cosine learning-rate decay to 2% over the run in pEVSL_MlpTrain.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "internal.hpp"

namespace pevsl {
namespace {

constexpr int kT = 64;   // output tile (kT x kT) per block, 4 x 4 per thread
constexpr int kKc = 16;  // K chunk staged through LDS

// C (M x N, row stride ldc) = A (M x K) * B (K x N) [+ bias[n]], where A and
// B are addressed through element strides (sam, sak) and (sbk, sbn) so the
// transposed operands of the backward pass need no copies.  With act = 1 the
// pre-activation goes to Z and C = gelu(Z) (exact erf form).
__global__ void __launch_bounds__(256)
k_gemm(int M, int N, int K, const float *__restrict__ A, long sam, long sak,
       const float *__restrict__ B, long sbk, long sbn, float *__restrict__ C, int ldc,
       const float *__restrict__ bias, float *__restrict__ Z, int act) {
  __shared__ float As[kKc][kT + 1];
  __shared__ float Bs[kKc][kT + 1];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.y * kT, n0 = blockIdx.x * kT;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += kKc) {
    for (int i = threadIdx.x; i < kKc * kT; i += 256) {
      // when the fast axis of an operand is k, let consecutive threads walk k
      int kk, mm;
      if (sak == 1) { kk = i % kKc; mm = i / kKc; } else { kk = i / kT; mm = i % kT; }
      const int gm = m0 + mm, gk = k0 + kk;
      As[kk][mm] = (gm < M && gk < K) ? A[(long)gm * sam + (long)gk * sak] : 0.f;
      if (sbk == 1) { kk = i % kKc; mm = i / kKc; } else { kk = i / kT; mm = i % kT; }
      const int gn = n0 + mm, gk2 = k0 + kk;
      Bs[kk][mm] = (gn < N && gk2 < K) ? B[(long)gk2 * sbk + (long)gn * sbn] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kKc; kk++) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] += a[i] * b[j];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int gm = m0 + ty * 4 + i;
    if (gm >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int gn = n0 + tx * 4 + j;
      if (gn >= N) continue;
      float z = acc[i][j] + (bias ? bias[gn] : 0.f);
      if (act) {
        Z[(size_t)gm * ldc + gn] = z;
        z = 0.5f * z * (1.f + erff(z * 0.70710678118654752f));
      }
      C[(size_t)gm * ldc + gn] = z;
    }
  }
}

// d = d * gelu'(z), in place
__global__ void k_gelu_bwd(size_t n, const float *__restrict__ Z, float *d) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float z = Z[i];
  const float cdf = 0.5f * (1.f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  d[i] = d[i] * (cdf + z * pdf);
}

// db[c] = sum over rows of dZ[:, c]
__global__ void k_colsum(int rows, int cols, const float *__restrict__ dZ,
                         float *__restrict__ db) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= cols) return;
  float s = 0.f;
  for (int r = 0; r < rows; r++) s += dZ[(size_t)r * cols + c];
  db[c] = s;
}

__device__ float block_reduce(float v, float *red, bool is_max) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      red[threadIdx.x] = is_max ? fmaxf(red[threadIdx.x], red[threadIdx.x + w])
                                : red[threadIdx.x] + red[threadIdx.x + w];
    __syncthreads();
  }
  const float r = red[0];
  __syncthreads();
  return r;
}

// One block per sample: y = z + q, softmax over the first nsoft values,
// loss and d loss / d z against T when T is given, Y = (p | regressions).
__global__ void __launch_bounds__(256)
k_head(int nout, int nsoft, int batch, const float *__restrict__ Zl,
       const float *__restrict__ Q, const float *__restrict__ T, float *__restrict__ Y,
       float *__restrict__ dZ, float *__restrict__ loss_rows) {
  __shared__ float red[256];
  const int r = blockIdx.x;
  const size_t o = (size_t)r * nout;
  auto yv = [&](int c) { return Zl[o + c] + (Q ? Q[o + c] : 0.f); };
  float l = 0.f;
  if (nsoft > 0) {
    float mx = -INFINITY;
    for (int c = threadIdx.x; c < nsoft; c += 256) mx = fmaxf(mx, yv(c));
    mx = block_reduce(mx, red, true);
    float se = 0.f;
    for (int c = threadIdx.x; c < nsoft; c += 256) se += expf(yv(c) - mx);
    se = block_reduce(se, red, false);
    const float lse = logf(se) + mx;
    for (int c = threadIdx.x; c < nsoft; c += 256) {
      const float lp = yv(c) - lse, p = expf(lp);
      if (Y) Y[o + c] = p;
      if (T) {
        const float t = T[o + c];
        if (t > 0.f) l -= t * lp;
        if (dZ) dZ[o + c] = (p - t) / (float)batch;
      }
    }
  }
  for (int c = nsoft + threadIdx.x; c < nout; c += 256) {
    const float y = yv(c);
    if (Y) Y[o + c] = y;
    if (T) {
      const float e = y - T[o + c];
      l += e * e;
      if (dZ) dZ[o + c] = 2.f * e / (float)batch;
    }
  }
  l = block_reduce(l, red, false);
  if (threadIdx.x == 0 && loss_rows) loss_rows[r] = l;
}

// acc[0] += scale * sum(v[0..n))   (one block)
__global__ void __launch_bounds__(256)
k_accum(int n, const float *__restrict__ v, float scale, float *__restrict__ acc) {
  __shared__ float red[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) s += v[i];
  s = block_reduce(s, red, false);
  if (threadIdx.x == 0) acc[0] += scale * s;
}

// dst row i = src row idx[i] (width w)
__global__ void k_gather(int rows, int w, const int *__restrict__ idx,
                         const float *__restrict__ src, float *__restrict__ dst) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)rows * w) return;
  const int r = (int)(i / w), c = (int)(i % w);
  dst[i] = src[(size_t)idx[r] * w + c];
}

__global__ void k_fill(size_t n, float *__restrict__ p, float v) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_adam(size_t n, float *__restrict__ w, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

int nblk(size_t n) { return (int)((n + 255) / 256); }

constexpr char kMagic[8] = {'P', 'E', 'V', 'S', 'L', 'M', 'L', 'P'};

}  // namespace

struct Mlp {
  int nin = 0, nh = 0, nl = 0, nout = 0, nsoft = 0;
  std::vector<int> din, dout;      // per layer (nl + 1 layers)
  std::vector<size_t> woff, boff;  // offsets into the packed parameters
  size_t nw = 0;
  DBuf<float> w, g, m, v;          // parameters, gradients, Adam moments
  long long step = 0;
  hipStream_t s = nullptr;
  int cap = 0;                     // batch capacity of the work buffers
  std::vector<DBuf<float>> act, pre;  // act[l] = input of layer l; pre = pre-activation
  DBuf<float> out, y, dout_, tmpa, tmpb, q, t, lossr, lacc;

  explicit Mlp(int layers) : act((size_t)layers), pre((size_t)layers) {}

  void reserve(int batch) {
    if (batch <= cap) return;
    cap = batch;
    for (int l = 0; l <= nl; l++) {
      act[(size_t)l].ensure((size_t)batch * din[(size_t)l]);
      pre[(size_t)l].ensure((size_t)batch * dout[(size_t)l]);
    }
    const int wmax = std::max(nin, std::max(nh, nout));
    for (DBuf<float> *b : {&out, &y, &dout_, &q, &t}) b->ensure((size_t)batch * nout);
    tmpa.ensure((size_t)batch * wmax);
    tmpb.ensure((size_t)batch * wmax);
    lossr.ensure((size_t)batch);
  }
  void gemm(int M, int N, int K, const float *A, long sam, long sak, const float *B, long sbk,
            long sbn, float *C, int ldc, const float *bias, float *Z, int act_) {
    dim3 grid((unsigned)((N + kT - 1) / kT), (unsigned)((M + kT - 1) / kT));
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk, sbn, C,
                       ldc, bias, Z, act_);
  }
  // act[0] holds the batch's inputs; the raw network output lands in `out`
  void forward(int batch) {
    for (int l = 0; l <= nl; l++) {
      const int in = din[(size_t)l], o = dout[(size_t)l];
      const float *W = w.p + woff[(size_t)l], *b = w.p + boff[(size_t)l];
      float *dst = l < nl ? act[(size_t)l + 1].p : out.p;
      // dst = act W^T + b
      gemm(batch, o, in, act[(size_t)l].p, in, 1, W, 1, in, dst, o, b, pre[(size_t)l].p,
           l < nl ? 1 : 0);
    }
  }
  // from d loss / d out (in dout_) to the parameter gradients g
  void backward(int batch) {
    float *dcur = dout_.p;
    for (int l = nl; l >= 0; l--) {
      const int in = din[(size_t)l], o = dout[(size_t)l];
      const float *W = w.p + woff[(size_t)l];
      // gW = dcur^T act (o x in), gb = column sums of dcur
      gemm(o, in, batch, dcur, 1, o, act[(size_t)l].p, in, 1, g.p + woff[(size_t)l], in,
           nullptr, nullptr, 0);
      hipLaunchKernelGGL(k_colsum, dim3(nblk((size_t)o)), dim3(256), 0, s, batch, o,
                         (const float *)dcur, g.p + boff[(size_t)l]);
      if (l == 0) break;
      // d act_l = dcur W (batch x in), then through the GELU of layer l-1
      float *dA = dcur == tmpa.p ? tmpb.p : tmpa.p;
      gemm(batch, in, o, dcur, o, 1, W, in, 1, dA, in, nullptr, nullptr, 0);
      hipLaunchKernelGGL(k_gelu_bwd, dim3(nblk((size_t)batch * in)), dim3(256), 0, s,
                         (size_t)batch * in, (const float *)pre[(size_t)l - 1].p, dA);
      dcur = dA;
    }
  }
  // head + backward on the batch already in act[0], q, t
  void loss_grad(int batch) {
    forward(batch);
    hipLaunchKernelGGL(k_head, dim3((unsigned)batch), dim3(256), 0, s, nout, nsoft, batch,
                       (const float *)out.p, (const float *)q.p, (const float *)t.p, y.p,
                       dout_.p, lossr.p);
    backward(batch);
  }
  void adam(float lr) {
    step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - (float)std::pow(0.9, (double)step);
    const float c2 = 1.f - (float)std::pow(0.999, (double)step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(nw)), dim3(256), 0, s, nw, w.p, (const float *)g.p,
                       m.p, v.p, lr, b1, b2, 1e-8f, c1, c2);
  }
  void upload(float *dst, const float *src, size_t cnt) {
    if (src)
      PV_HIP(hipMemcpyAsync(dst, src, sizeof(float) * cnt, hipMemcpyHostToDevice, s));
    else
      hipLaunchKernelGGL(k_fill, dim3(nblk(cnt)), dim3(256), 0, s, cnt, dst, 0.f);
  }
};

Mlp *mlp_of(pevsl_Mlp *p) { return reinterpret_cast<Mlp *>(p); }

}  // namespace pevsl

using namespace pevsl;

extern "C" {

int pEVSL_MlpCreate(int nin, int nhidden, int nlayers, int nout, int nsoft,
                    unsigned long long seed, pevsl_Mlp **mo) {
  if (!mo || nin < 1 || nhidden < 1 || nlayers < 0 || nout < 1 || nsoft < 0 || nsoft > nout)
    return -1;
  Mlp *m = new Mlp(nlayers + 1);
  m->nin = nin;
  m->nh = nhidden;
  m->nl = nlayers;
  m->nout = nout;
  m->nsoft = nsoft;
  PV_HIP(hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking));
  size_t off = 0;
  for (int l = 0; l <= nlayers; l++) {
    const int in = l == 0 ? nin : nhidden, o = l == nlayers ? nout : nhidden;
    m->din.push_back(in);
    m->dout.push_back(o);
    m->woff.push_back(off);
    off += (size_t)in * o;
    m->boff.push_back(off);
    off += (size_t)o;
  }
  m->nw = off;
  // He-normal hidden weights (Box-Muller on a splitmix64 stream), zero biases
  // and a zero output layer: an untrained network returns the prior.
  std::vector<float> h(off, 0.f);
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
  auto unif = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((double)(z >> 11) + 0.5) / 9007199254740992.0;
  };
  for (int l = 0; l < nlayers; l++) {
    const double sc = std::sqrt(2.0 / m->din[(size_t)l]);
    const size_t cnt = (size_t)m->din[(size_t)l] * m->dout[(size_t)l];
    for (size_t i = 0; i < cnt; i++) {
      const double u1 = unif(), u2 = unif();
      h[m->woff[(size_t)l] + i] =
          (float)(sc * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2));
    }
  }
  for (DBuf<float> *b : {&m->w, &m->g, &m->m, &m->v}) b->ensure(off);
  m->lacc.ensure(1);
  PV_HIP(hipMemcpyAsync(m->w.p, h.data(), sizeof(float) * off, hipMemcpyHostToDevice, m->s));
  m->upload(m->m.p, nullptr, off);
  m->upload(m->v.p, nullptr, off);
  m->upload(m->g.p, nullptr, off);
  PV_HIP(hipStreamSynchronize(m->s));
  *mo = reinterpret_cast<pevsl_Mlp *>(m);
  return 0;
}

int pEVSL_MlpFree(pevsl_Mlp *mp) {
  Mlp *m = mlp_of(mp);
  if (!m) return 0;
  (void)hipStreamSynchronize(m->s);
  hipStream_t s = m->s;
  delete m;
  (void)hipStreamDestroy(s);
  return 0;
}

int pEVSL_MlpShape(pevsl_Mlp *mp, int *nin, int *nhidden, int *nlayers, int *nout, int *nsoft,
                   size_t *nweights) {
  Mlp *m = mlp_of(mp);
  if (!m) return -1;
  if (nin) *nin = m->nin;
  if (nhidden) *nhidden = m->nh;
  if (nlayers) *nlayers = m->nl;
  if (nout) *nout = m->nout;
  if (nsoft) *nsoft = m->nsoft;
  if (nweights) *nweights = m->nw;
  return 0;
}

int pEVSL_MlpSetWeights(pevsl_Mlp *mp, const float *w, size_t nw) {
  Mlp *m = mlp_of(mp);
  if (!m || !w || nw != m->nw) return -1;
  PV_HIP(hipMemcpyAsync(m->w.p, w, sizeof(float) * nw, hipMemcpyHostToDevice, m->s));
  m->upload(m->m.p, nullptr, nw);
  m->upload(m->v.p, nullptr, nw);
  PV_HIP(hipStreamSynchronize(m->s));
  m->step = 0;
  return 0;
}

int pEVSL_MlpGetWeights(pevsl_Mlp *mp, float *w, size_t nw) {
  Mlp *m = mlp_of(mp);
  if (!m || !w || nw != m->nw) return -1;
  PV_HIP(hipMemcpyAsync(w, m->w.p, sizeof(float) * nw, hipMemcpyDeviceToHost, m->s));
  PV_HIP(hipStreamSynchronize(m->s));
  return 0;
}

int pEVSL_MlpForward(pevsl_Mlp *mp, const float *X, const float *Q, int batch, float *Y) {
  Mlp *m = mlp_of(mp);
  if (!m || !X || !Y || batch < 1) return -1;
  m->reserve(batch);
  m->upload(m->act[0].p, X, (size_t)batch * m->nin);
  m->upload(m->q.p, Q, (size_t)batch * m->nout);
  m->forward(batch);
  hipLaunchKernelGGL(k_head, dim3((unsigned)batch), dim3(256), 0, m->s, m->nout, m->nsoft,
                     batch, (const float *)m->out.p, (const float *)m->q.p,
                     (const float *)nullptr, m->y.p, (float *)nullptr, (float *)nullptr);
  PV_HIP(hipMemcpyAsync(Y, m->y.p, sizeof(float) * batch * m->nout, hipMemcpyDeviceToHost,
                        m->s));
  PV_HIP(hipStreamSynchronize(m->s));
  return 0;
}

int pEVSL_MlpLossGrad(pevsl_Mlp *mp, const float *X, const float *Q, const float *T,
                      int batch, float *loss, float *grad) {
  Mlp *m = mlp_of(mp);
  if (!m || !X || !T || batch < 1) return -1;
  m->reserve(batch);
  m->upload(m->act[0].p, X, (size_t)batch * m->nin);
  m->upload(m->q.p, Q, (size_t)batch * m->nout);
  m->upload(m->t.p, T, (size_t)batch * m->nout);
  m->upload(m->lacc.p, nullptr, 1);
  m->loss_grad(batch);
  hipLaunchKernelGGL(k_accum, dim3(1), dim3(256), 0, m->s, batch, (const float *)m->lossr.p,
                     1.f / batch, m->lacc.p);
  float l = 0.f;
  PV_HIP(hipMemcpyAsync(&l, m->lacc.p, sizeof(float), hipMemcpyDeviceToHost, m->s));
  if (grad)
    PV_HIP(hipMemcpyAsync(grad, m->g.p, sizeof(float) * m->nw, hipMemcpyDeviceToHost, m->s));
  PV_HIP(hipStreamSynchronize(m->s));
  if (loss) *loss = l;
  return 0;
}

int pEVSL_MlpTrainStep(pevsl_Mlp *mp, const float *X, const float *Q, const float *T,
                       int batch, float lr, float *loss) {
  Mlp *m = mlp_of(mp);
  if (!m || !X || !T || batch < 1) return -1;
  const int rc = pEVSL_MlpLossGrad(mp, X, Q, T, batch, loss, nullptr);
  if (rc) return rc;
  m->adam(lr);
  PV_HIP(hipStreamSynchronize(m->s));
  return 0;
}

int pEVSL_MlpTrain(pevsl_Mlp *mp, const float *X, const float *Q, const float *T, int nsamples,
                   int batch, int epochs, float lr, unsigned long long seed, float *losses,
                   int verbose) {
  Mlp *m = mlp_of(mp);
  if (!m || !X || !T || nsamples < 1 || batch < 1 || epochs < 1) return -1;
  batch = std::min(batch, nsamples);
  m->reserve(batch);
  // the whole set stays resident; each step gathers its rows on the device
  DBuf<float> dX, dQ, dT, dl;
  DBuf<int> perm;
  dX.ensure((size_t)nsamples * m->nin);
  dQ.ensure((size_t)nsamples * m->nout);
  dT.ensure((size_t)nsamples * m->nout);
  dl.ensure((size_t)epochs);
  perm.ensure((size_t)nsamples);
  m->upload(dX.p, X, (size_t)nsamples * m->nin);
  m->upload(dQ.p, Q, (size_t)nsamples * m->nout);
  m->upload(dT.p, T, (size_t)nsamples * m->nout);
  m->upload(dl.p, nullptr, (size_t)epochs);
  std::mt19937_64 rng(seed);
  std::vector<int> order((size_t)nsamples);
  std::iota(order.begin(), order.end(), 0);
  const int steps_per_epoch = (nsamples + batch - 1) / batch;
  const double total = (double)steps_per_epoch * epochs;
  long long it = 0;
  for (int e = 0; e < epochs; e++) {
    std::shuffle(order.begin(), order.end(), rng);
    // the stream must be done with the previous order before it is replaced
    PV_HIP(hipStreamSynchronize(m->s));
    PV_HIP(hipMemcpyAsync(perm.p, order.data(), sizeof(int) * nsamples, hipMemcpyHostToDevice,
                          m->s));
    for (int b0 = 0; b0 < nsamples; b0 += batch, it++) {
      const int bs = std::min(batch, nsamples - b0);
      const int *idx = perm.p + b0;
      hipLaunchKernelGGL(k_gather, dim3(nblk((size_t)bs * m->nin)), dim3(256), 0, m->s, bs,
                         m->nin, idx, (const float *)dX.p, m->act[0].p);
      hipLaunchKernelGGL(k_gather, dim3(nblk((size_t)bs * m->nout)), dim3(256), 0, m->s, bs,
                         m->nout, idx, (const float *)dQ.p, m->q.p);
      hipLaunchKernelGGL(k_gather, dim3(nblk((size_t)bs * m->nout)), dim3(256), 0, m->s, bs,
                         m->nout, idx, (const float *)dT.p, m->t.p);
      m->loss_grad(bs);
      hipLaunchKernelGGL(k_accum, dim3(1), dim3(256), 0, m->s, bs, (const float *)m->lossr.p,
                         1.f / ((float)bs * steps_per_epoch), dl.p + e);
      const double f = 0.02 + 0.98 * 0.5 * (1.0 + std::cos(M_PI * (double)it / total));
      m->adam((float)(lr * f));
    }
    if (verbose > 0 && ((e + 1) % verbose == 0 || e + 1 == epochs)) {
      float le = 0.f;
      PV_HIP(hipMemcpyAsync(&le, dl.p + e, sizeof(float), hipMemcpyDeviceToHost, m->s));
      PV_HIP(hipStreamSynchronize(m->s));
      fprintf(stdout, "MlpTrain: epoch %d/%d mean loss %.6f\n", e + 1, epochs, le);
      fflush(stdout);
    }
  }
  if (losses)
    PV_HIP(hipMemcpyAsync(losses, dl.p, sizeof(float) * epochs, hipMemcpyDeviceToHost, m->s));
  PV_HIP(hipStreamSynchronize(m->s));
  return 0;
}

int pEVSL_MlpSave(pevsl_Mlp *mp, const char *path) {
  Mlp *m = mlp_of(mp);
  if (!m || !path) return -1;
  std::vector<float> h(m->nw);
  if (pEVSL_MlpGetWeights(mp, h.data(), m->nw)) return -1;
  FILE *f = fopen(path, "wb");
  if (!f) return -2;
  const int32_t hdr[6] = {1, m->nin, m->nh, m->nl, m->nout, m->nsoft};
  bool ok = fwrite(kMagic, 1, 8, f) == 8 && fwrite(hdr, sizeof(int32_t), 6, f) == 6 &&
            fwrite(h.data(), sizeof(float), m->nw, f) == m->nw;
  ok = (fclose(f) == 0) && ok;
  return ok ? 0 : -3;
}

int pEVSL_MlpLoad(const char *path, pevsl_Mlp **mo) {
  if (!path || !mo) return -1;
  FILE *f = fopen(path, "rb");
  if (!f) return -2;
  char mg[8];
  int32_t hdr[6];
  if (fread(mg, 1, 8, f) != 8 || memcmp(mg, kMagic, 8) != 0 ||
      fread(hdr, sizeof(int32_t), 6, f) != 6 || hdr[0] != 1) {
    fclose(f);
    return -3;
  }
  pevsl_Mlp *mp = nullptr;
  if (pEVSL_MlpCreate(hdr[1], hdr[2], hdr[3], hdr[4], hdr[5], 0, &mp)) {
    fclose(f);
    return -3;
  }
  Mlp *m = mlp_of(mp);
  std::vector<float> h(m->nw);
  const bool ok = fread(h.data(), sizeof(float), m->nw, f) == m->nw;
  fclose(f);
  if (!ok || pEVSL_MlpSetWeights(mp, h.data(), m->nw)) {
    pEVSL_MlpFree(mp);
    return -3;
  }
  *mo = mp;
  return 0;
}

}  // extern "C"
