// select.hip -- the learned strategy selector: a per-instance sketch of the
// index distribution (12 features) and a small fp32 MLP (12 -> 64 -> 64 ->
// 3 strategy + 4 variant + 3 task-size logits) with hand-written inference and
// training kernels.  It replaces the exhaustive tuner's search; the tuner stays
// as fallback and as the source of the training labels.  The arithmetic of the
// histogram / gather itself is untouched by it.
#include <math.h>
#include <string.h>

#include "bale_internal.h"

namespace bale {

constexpr int kFeat = 12;
constexpr int kHid = 64;
constexpr int kOut = 10;  // 3 strategy + 4 variant + 3 task size
constexpr int kHeadOff[3] = {0, 3, 7};
constexpr int kHeadLen[3] = {3, 4, 3};
// parameter layout: W1[64][12], b1[64], W2[64][64], b2[64], W3[10][64], b3[10]
constexpr int kW1 = 0, kB1 = kW1 + kHid * kFeat, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid,
              kW3 = kB2 + kHid, kB3 = kW3 + kOut * kHid, kParams = kB3 + kOut;
constexpr int kWin = 64, kWinLen = 64, kSample = kWin * kWinLen;

struct SketchIn {
  const int64_t* idx;
  int64_t n, table_len;
  int32_t kind, dtype;
};

__device__ inline float block_sum(float v, float* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  float s = 0.f;
  for (int i = 0; i < nw; i++) s += red[i];
  return s;
}

__device__ inline float block_max(float v, float* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  float s = -3.4e38f;
  for (int i = 0; i < nw; i++) s = fmaxf(s, red[i]);
  return s;
}

// One workgroup of 256 per instance.
__global__ __launch_bounds__(256) void k_sketch(const SketchIn* __restrict__ ins, int64_t count,
                                                float* __restrict__ feat) {
  __shared__ uint32_t s[kSample];
  __shared__ float cnt[kSample];
  __shared__ float red[8];
  __shared__ int claim;
  const int64_t inst = blockIdx.x;
  if (inst >= count) return;
  const SketchIn I = ins[inst];
  const int T = blockDim.x, t = threadIdx.x;
  // 64 windows of 64 consecutive indices spread over the array (all of it when
  // shorter than 4096)
  const int m = (int)min<int64_t>(I.n, kSample);
  float eq = 0.f;
  for (int k = t; k < kSample; k += T) {
    uint32_t v = 0xFFFFFFFFu;
    if (k < m) {
      int64_t pos;
      if (I.n <= kSample) {
        pos = k;
      } else {
        int w = k / kWinLen, o = k % kWinLen;
        pos = (int64_t)(((unsigned __int128)w * (uint64_t)(I.n - kWinLen)) / (kWin - 1)) + o;
      }
      v = (uint32_t)(uint64_t)I.idx[pos];
    }
    s[k] = v;
  }
  __syncthreads();
  for (int k = t; k < m; k += T)
    if (k % kWinLen != 0 && s[k] == s[k - 1]) eq += 1.f;
  const float n_eq = block_sum(eq, red);
  __syncthreads();
  // bitonic sort of s
  for (int size = 2; size <= kSample; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int k = t; k < kSample / 2; k += T) {
        int lo = 2 * k - (k & (stride - 1)), hi = lo + stride;
        bool up = (lo & size) == 0;
        uint32_t a = s[lo], b = s[hi];
        if ((a > b) == up) {
          s[lo] = b;
          s[hi] = a;
        }
      }
      __syncthreads();
    }
  // value frequencies: the first element of each run of equal values holds its count
  float ent = 0.f, distinct = 0.f, dup = 0.f, top1 = 0.f;
  for (int k = t; k < kSample; k += T) {
    float c = 0.f;
    if (k < m && (k == 0 || s[k - 1] != s[k])) {
      int e = k + 1;
      while (e < m && s[e] == s[k]) e++;
      c = (float)(e - k);
      float p = c / (float)m;
      ent -= p * log2f(p);
      distinct += 1.f;
      if (c >= 2.f) dup += c;
      top1 = fmaxf(top1, c);
    }
    cnt[k] = c;
  }
  ent = block_sum(ent, red);
  distinct = block_sum(distinct, red);
  dup = block_sum(dup, red);
  top1 = block_max(top1, red);
  // mass of the 16 most frequent values: 16 rounds of "take the max"
  float top16 = 0.f;
  for (int r = 0; r < 16; r++) {
    float best = -1.f;
    int arg = -1;
    for (int k = t; k < kSample; k += T)
      if (cnt[k] > best) {
        best = cnt[k];
        arg = k;
      }
    float mx = block_max(best, red);
    if (mx <= 0.f) break;
    top16 += mx;
    __syncthreads();
    // one holder of the maximum clears it
    if (t == 0) claim = -1;
    __syncthreads();
    if (best == mx && arg >= 0) atomicCAS(&claim, -1, arg);
    __syncthreads();
    if (t == 0 && claim >= 0) cnt[claim] = 0.f;
    __syncthreads();
  }
  if (t == 0) {
    const float fm = (float)max(m, 1);
    const float nwin = I.n <= kSample ? 1.f : (float)kWin;
    const float runs = fmaxf(1.f, nwin + ((float)m - nwin - n_eq));
    float* f = feat + inst * kFeat;
    f[0] = log2f((float)max<int64_t>(I.table_len, 1)) / 24.f;
    f[1] = log2f((float)max<int64_t>(I.n, 1)) / 22.f;
    f[2] = (float)I.kind;
    f[3] = (float)I.dtype;
    f[4] = fminf(2.f, fmaxf(-2.f, (log2f((float)max<int64_t>(I.n, 1)) -
                                  log2f((float)max<int64_t>(I.table_len, 1))) / 12.f));
    f[5] = ent / 12.f;
    f[6] = top16 / fm;
    f[7] = top1 / fm;
    f[8] = distinct / fm;
    f[9] = n_eq / fm;
    f[10] = log2f(fm / runs) / 6.f;
    f[11] = dup / fm;
  }
}

// ---- MLP -------------------------------------------------------------------
// forward for one sample held by one thread; h1/h2 in registers
__device__ inline void mlp_forward(const float* __restrict__ P, const float* x, float* h1, float* h2,
                                   float* out) {
  for (int j = 0; j < kHid; j++) {
    float a = P[kB1 + j];
#pragma unroll
    for (int i = 0; i < kFeat; i++) a += P[kW1 + j * kFeat + i] * x[i];
    h1[j] = a > 0.f ? a : 0.f;
  }
  for (int j = 0; j < kHid; j++) {
    float a = P[kB2 + j];
    for (int i = 0; i < kHid; i++) a += P[kW2 + j * kHid + i] * h1[i];
    h2[j] = a > 0.f ? a : 0.f;
  }
  for (int o = 0; o < kOut; o++) {
    float a = P[kB3 + o];
    for (int i = 0; i < kHid; i++) a += P[kW3 + o * kHid + i] * h2[i];
    out[o] = a;
  }
}

// One thread per instance: logits and the argmax of each head.
__global__ __launch_bounds__(256) void k_mlp_infer(const float* __restrict__ P,
                                                   const float* __restrict__ feat, int64_t count,
                                                   float* __restrict__ logits,
                                                   int32_t* __restrict__ cls) {
  __shared__ float sp[kParams];
  for (int k = threadIdx.x; k < kParams; k += blockDim.x) sp[k] = P[k];
  __syncthreads();
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float x[kFeat], h1[kHid], h2[kHid], out[kOut];
  for (int k = 0; k < kFeat; k++) x[k] = feat[i * kFeat + k];
  mlp_forward(sp, x, h1, h2, out);
  for (int o = 0; o < kOut; o++) logits[i * kOut + o] = out[o];
  for (int h = 0; h < 3; h++) {
    int best = 0;
    for (int o = 1; o < kHeadLen[h]; o++)
      if (out[kHeadOff[h] + o] > out[kHeadOff[h] + best]) best = o;
    cls[i * 3 + h] = best;
  }
}

// Training: one workgroup of 256 threads runs every step.  Each step takes a
// minibatch (a seeded permutation slice), every thread does forward+backward
// for one sample and adds its gradient into LDS, then the workgroup applies an
// Adam update.  Loss: sum over the three heads of softmax cross-entropy
// (label < 0 means "no label" for that head).
// Candidates of one training instance, for the cost-aware loss: the three
// head classes of each candidate the tuner timed and its time over the best
// one (>= 1); s < 0 ends the list.
struct TrainCand {
  float s, v, t, r;
};
constexpr float kCostSharpness = 8.f;  // (t_best / t)^8: 10% slower weighs 0.47, 50% slower 0.04

// Training loss.  cands == nullptr: softmax cross-entropy of each head
// against the tuner's argmin (labels).  Otherwise a cross-entropy over the
// instance's timed candidates: a candidate's score is the sum of its heads'
// log-probabilities (the rule inference uses), q = softmax of the scores over
// the candidates, and the target weights each candidate by r^-kCostSharpness
// (normalised), so near-best candidates share the target and slow ones get
// almost none: loss = -sum_j w_j log q_j.  Its gradient with respect to head
// h's logit m is the sum over candidates j with class m in head h of
// q_j - w_j (the heads' own softmax terms cancel because both q and w sum
// to 1).  Unlike the expected relative time sum_j q_j r_j, whose gradient
// vanishes once q settles on any one candidate, this keeps pulling toward
// the fast ones.
__global__ __launch_bounds__(256) void k_mlp_train(float* __restrict__ P, float* __restrict__ adam,
                                                    const float* __restrict__ feat,
                                                    const int32_t* __restrict__ labels,
                                                    const TrainCand* __restrict__ cands, int maxc,
                                                    int64_t count, int steps, int batch, float lr,
                                                    uint64_t seed, float* __restrict__ loss_out) {
  __shared__ float g[kParams];
  __shared__ float sp[kParams];
  __shared__ float red[16];
  const int t = threadIdx.x, T = blockDim.x;
  for (int k = t; k < kParams; k += T) sp[k] = P[k];
  __syncthreads();
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float* m1 = adam;
  float* m2 = adam + kParams;
  for (int step = 0; step < steps; step++) {
    for (int k = t; k < kParams; k += T) g[k] = 0.f;
    __syncthreads();
    float lsum = 0.f;
    for (int r = t; r < batch; r += T) {
      // sample index: hash of (seed, step, r) reduced to [0,count)
      uint64_t h = mix64(seed + ((uint64_t)step * 1315423911ull + (uint64_t)r + 1) * kGolden);
      int64_t i = (int64_t)__umul64hi(h, (uint64_t)count);
      float x[kFeat], h1[kHid], h2[kHid], out[kOut], dout[kOut];
      for (int k = 0; k < kFeat; k++) x[k] = feat[i * kFeat + k];
      mlp_forward(sp, x, h1, h2, out);
      if (cands) {
        float lp[kOut];
        for (int hd = 0; hd < 3; hd++) {
          int o0 = kHeadOff[hd], L = kHeadLen[hd];
          float mx = out[o0];
          for (int o = 1; o < L; o++) mx = fmaxf(mx, out[o0 + o]);
          float z = 0.f;
          for (int o = 0; o < L; o++) z += expf(out[o0 + o] - mx);
          const float lz = logf(z) + mx;
          for (int o = 0; o < L; o++) lp[o0 + o] = out[o0 + o] - lz;
        }
        const TrainCand* ci = cands + i * maxc;
        float smax = -3.4e38f;
        int nc = 0;
        for (; nc < maxc && ci[nc].s >= 0.f; nc++) {
          const float sc = lp[(int)ci[nc].s] + lp[3 + (int)ci[nc].v] + lp[7 + (int)ci[nc].t];
          smax = fmaxf(smax, sc);
        }
        float z = 0.f, wz = 0.f;
        for (int j = 0; j < nc; j++) {
          z += expf(lp[(int)ci[j].s] + lp[3 + (int)ci[j].v] + lp[7 + (int)ci[j].t] - smax);
          wz += powf(fmaxf(ci[j].r, 1.f), -kCostSharpness);
        }
        const float lz = logf(z) + smax;
        for (int o = 0; o < kOut; o++) dout[o] = 0.f;
        for (int j = 0; j < nc; j++) {
          const float sc = lp[(int)ci[j].s] + lp[3 + (int)ci[j].v] + lp[7 + (int)ci[j].t];
          const float qj = expf(sc - lz);
          const float wj = powf(fmaxf(ci[j].r, 1.f), -kCostSharpness) / wz;
          const float gj = qj - wj;
          dout[(int)ci[j].s] += gj;
          dout[3 + (int)ci[j].v] += gj;
          dout[7 + (int)ci[j].t] += gj;
          lsum += wj * (lz - sc);
        }
      } else
      for (int hd = 0; hd < 3; hd++) {
        int lab = labels[i * 3 + hd];
        int o0 = kHeadOff[hd], L = kHeadLen[hd];
        if (lab < 0 || lab >= L) {
          for (int o = 0; o < L; o++) dout[o0 + o] = 0.f;
          continue;
        }
        float mx = out[o0];
        for (int o = 1; o < L; o++) mx = fmaxf(mx, out[o0 + o]);
        float z = 0.f;
        for (int o = 0; o < L; o++) z += expf(out[o0 + o] - mx);
        for (int o = 0; o < L; o++) {
          float p = expf(out[o0 + o] - mx) / z;
          dout[o0 + o] = p - (o == lab ? 1.f : 0.f);
        }
        lsum += logf(z) + mx - out[o0 + lab];
      }
      // backward
      float dh2[kHid];
      for (int i2 = 0; i2 < kHid; i2++) {
        float a = 0.f;
        for (int o = 0; o < kOut; o++) a += sp[kW3 + o * kHid + i2] * dout[o];
        dh2[i2] = h2[i2] > 0.f ? a : 0.f;
      }
      for (int o = 0; o < kOut; o++) {
        if (dout[o] == 0.f) continue;
        atomicAdd(&g[kB3 + o], dout[o]);
        for (int i2 = 0; i2 < kHid; i2++)
          if (h2[i2] != 0.f) atomicAdd(&g[kW3 + o * kHid + i2], dout[o] * h2[i2]);
      }
      float dh1[kHid];
      for (int i1 = 0; i1 < kHid; i1++) dh1[i1] = 0.f;
      for (int j = 0; j < kHid; j++) {
        if (dh2[j] == 0.f) continue;
        atomicAdd(&g[kB2 + j], dh2[j]);
        for (int i1 = 0; i1 < kHid; i1++) {
          if (h1[i1] != 0.f) atomicAdd(&g[kW2 + j * kHid + i1], dh2[j] * h1[i1]);
          dh1[i1] += sp[kW2 + j * kHid + i1] * dh2[j];
        }
      }
      for (int j = 0; j < kHid; j++) {
        if (h1[j] <= 0.f || dh1[j] == 0.f) continue;
        atomicAdd(&g[kB1 + j], dh1[j]);
        for (int i0 = 0; i0 < kFeat; i0++) atomicAdd(&g[kW1 + j * kFeat + i0], dh1[j] * x[i0]);
      }
    }
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
    if ((t & 63) == 0) red[t >> 6] = lsum;
    __syncthreads();
    if (t == 0 && loss_out) {
      float s = 0.f;
      for (int w = 0; w < T / 64; w++) s += red[w];
      loss_out[step] = s / (float)batch;
    }
    // Adam
    const float bc1 = 1.f - powf(b1, (float)(step + 1)), bc2 = 1.f - powf(b2, (float)(step + 1));
    for (int k = t; k < kParams; k += T) {
      float gr = g[k] / (float)batch;
      float a = b1 * m1[k] + (1.f - b1) * gr;
      float b = b2 * m2[k] + (1.f - b2) * gr * gr;
      m1[k] = a;
      m2[k] = b;
      sp[k] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
    }
    __syncthreads();
  }
  for (int k = t; k < kParams; k += T) P[k] = sp[k];
}

// Mean and standard deviation of each feature over count rows (one
// workgroup), written to stats[0..kFeat) and stats[kFeat..2 kFeat), and the
// standardised copy of the features to out.
__global__ __launch_bounds__(256) void k_feat_standardize(const float* __restrict__ feat,
                                                           int64_t count, float* __restrict__ stats,
                                                           float* __restrict__ out) {
  __shared__ float red[8];
  __shared__ float mu[kFeat], sd[kFeat];
  for (int k = 0; k < kFeat; k++) {
    float a = 0.f;
    for (int64_t i = threadIdx.x; i < count; i += blockDim.x) a += feat[i * kFeat + k];
    const float m = block_sum(a, red) / (float)count;
    float b = 0.f;
    for (int64_t i = threadIdx.x; i < count; i += blockDim.x) {
      const float d = feat[i * kFeat + k] - m;
      b += d * d;
    }
    const float v = block_sum(b, red) / (float)count;
    if (threadIdx.x == 0) {
      mu[k] = m;
      sd[k] = sqrtf(v) + 1e-6f;
      stats[k] = m;
      stats[kFeat + k] = sqrtf(v) + 1e-6f;
    }
  }
  __syncthreads();
  for (int64_t q = threadIdx.x; q < count * kFeat; q += blockDim.x) {
    const int k = (int)(q % kFeat);
    out[q] = (feat[q] - mu[k]) / sd[k];
  }
}

// Fold x' = (x - mu) / sd into the first layer: W1 <- W1 / sd,
// b1 <- b1 - W1 mu / sd, so inference takes the raw features.
__global__ __launch_bounds__(64) void k_mlp_fold(float* __restrict__ P,
                                                 const float* __restrict__ stats) {
  const int j = threadIdx.x;
  if (j >= kHid) return;
  float shift = 0.f;
  for (int k = 0; k < kFeat; k++) {
    const float w = P[kW1 + j * kFeat + k] / stats[kFeat + k];
    P[kW1 + j * kFeat + k] = w;
    shift += w * stats[k];
  }
  P[kB1 + j] -= shift;
}

}  // namespace bale

using namespace bale;

extern "C" {

int bale_mlp_param_count(void) { return kParams; }
int bale_sketch_feature_count(void) { return kFeat; }

int bale_sketch(const bale_instance* inst, int64_t count, float* features, void* stream) {
  if (count < 0 || (count > 0 && (!inst || !features)))
    return bale_set_error(BALE_ERR_ARG, "bad sketch arguments");
  if (count == 0) return BALE_OK;
  if (!device_state()) return bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
  hipStream_t s = (hipStream_t)stream;
  SketchIn* h = new SketchIn[count];
  for (int64_t i = 0; i < count; i++) {
    h[i].idx = inst[i].index;
    h[i].n = inst[i].n;
    h[i].table_len = inst[i].table_len;
    h[i].kind = inst[i].kind;
    h[i].dtype = inst[i].dtype;
    if (inst[i].n > 0 && !inst[i].index) {
      delete[] h;
      return bale_set_error(BALE_ERR_ARG, "instance %lld has no index", (long long)i);
    }
  }
  SketchIn* d = (SketchIn*)BALE_DEV_ALLOC(sizeof(SketchIn) * count);
  if (!d) {
    delete[] h;
    return bale_set_error(BALE_ERR_ALLOC, "sketch descriptors: device allocation failed");
  }
  int rc = (int)hipMemcpy(d, h, sizeof(SketchIn) * count, hipMemcpyHostToDevice);
  delete[] h;
  if (rc) {
    dev_free(d);
    return bale_set_error(BALE_ERR_HIP, "sketch descriptors: %s", hipGetErrorString((hipError_t)rc));
  }
  hipLaunchKernelGGL(k_sketch, dim3((unsigned)count), dim3(256), 0, s, d, count, features);
  rc = (int)hipStreamSynchronize(s);
  dev_free(d);
  return rc ? bale_set_error(BALE_ERR_HIP, "sketch: %s", hipGetErrorString((hipError_t)rc)) : BALE_OK;
}

int bale_mlp_infer(const float* params, const float* features, int64_t count, float* logits,
                   int32_t* classes, void* stream) {
  if (count < 0 || (count > 0 && (!params || !features || !logits || !classes)))
    return bale_set_error(BALE_ERR_ARG, "bad inference arguments");
  if (count == 0) return BALE_OK;
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)ceil_div(count, 256)), dim3(256), 0,
                     (hipStream_t)stream, params, features, count, logits, classes);
  int rc = (int)hipGetLastError();
  return rc ? bale_set_error(BALE_ERR_HIP, "mlp infer: %s", hipGetErrorString((hipError_t)rc)) : BALE_OK;
}

int bale_mlp_train(float* params, float* adam_state, const float* features, const int32_t* labels,
                   int64_t count, int steps, int batch, float lr, uint64_t seed, float* loss,
                   void* stream) {
  if (count <= 0 || steps < 0 || batch <= 0 || !params || !adam_state || !features || !labels)
    return bale_set_error(BALE_ERR_ARG, "bad training arguments");
  hipLaunchKernelGGL(k_mlp_train, dim3(1), dim3(256), 0, (hipStream_t)stream, params, adam_state,
                     features, labels, (const TrainCand*)nullptr, 0, count, steps, batch, lr, seed,
                     loss);
  int rc = (int)hipGetLastError();
  return rc ? bale_set_error(BALE_ERR_HIP, "mlp train: %s", hipGetErrorString((hipError_t)rc)) : BALE_OK;
}

int bale_mlp_train_cost(float* params, float* adam_state, const float* features,
                        const float* cands, int maxc, int64_t count, int steps, int batch,
                        float lr, uint64_t seed, float* loss, float* scratch, void* stream) {
  if (count <= 0 || steps < 0 || batch <= 0 || maxc <= 0 || !params || !adam_state || !features ||
      !cands || !scratch)
    return bale_set_error(BALE_ERR_ARG, "bad training arguments");
  hipStream_t s = (hipStream_t)stream;
  // scratch: 2 kFeat floats of statistics, then the standardised features
  float* stats = scratch;
  float* norm = scratch + 2 * kFeat;
  hipLaunchKernelGGL(k_feat_standardize, dim3(1), dim3(256), 0, s, features, count, stats, norm);
  hipLaunchKernelGGL(k_mlp_train, dim3(1), dim3(256), 0, s, params, adam_state, norm,
                     (const int32_t*)nullptr, (const TrainCand*)cands, maxc, count, steps, batch,
                     lr, seed, loss);
  hipLaunchKernelGGL(k_mlp_fold, dim3(1), dim3(64), 0, s, params, (const float*)stats);
  int rc = (int)hipGetLastError();
  return rc ? bale_set_error(BALE_ERR_HIP, "mlp train: %s", hipGetErrorString((hipError_t)rc)) : BALE_OK;
}

}  // extern "C"
