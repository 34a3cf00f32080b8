// HIP kernels of the learned selector (model/mlp.h): wave-per-sample
// forward, workgroup-per-sample argmin selection (the path inside the
// inspector), and training: per-wave partial gradients of the masked squared
// error, a fixed-order reduction (so training is bit-reproducible) and Adam.
//
// Inside a wavefront lane j owns hidden unit j (and output j < 30): every
// weight read W[k][j] is one coalesced row across the lanes, activations are
// broadcast from LDS. No MFMA: the network is 6942 parameters and one level
// needs 13.6 kFLOP.
#pragma once
#include "../abstraction/statistics.h"
#include "../model/mlp.h"
#include "../utils/wave.h"

namespace gsw {

using namespace gsw_mlp;

// per-wave LDS scratch of one sample
struct mlp_wave_sh_t {
  float x[16];
  float h1[NH];
  float h2[NH];
  float dy[32];
  float dh2[NH];
};

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Forward pass of the sample in s.x; returns output j (lanes >= 30: 0).
__device__ __forceinline__ float mlp_wave_forward(const float *w,
                                                  mlp_wave_sh_t &s) {
  const int j = lane_id();
  float a = w[OB1 + j];
#pragma unroll
  for (int k = 0; k < NIN; ++k) a += s.x[k] * w[OW1 + k * NH + j];
  s.h1[j] = a > 0.f ? a : 0.f;
  wave_lds_sync();
  a = w[OB2 + j];
#pragma unroll 8
  for (int k = 0; k < NH; ++k) a += s.h1[k] * w[OW2 + k * NH + j];
  s.h2[j] = a > 0.f ? a : 0.f;
  wave_lds_sync();
  float y = 0.f;
  if (j < NOUT) {
    y = w[OB3 + j];
#pragma unroll 8
    for (int k = 0; k < NH; ++k) y += s.h2[k] * w[OW3 + k * NOUT + j];
  }
  return y;
}

// argmin over the classes set in `mask`, ties to the lower id (all lanes)
__device__ __forceinline__ int wave_argmin(float y, uint32_t mask) {
  const int j = lane_id();
  float v = (j < NOUT && ((mask >> j) & 1u)) ? y : __builtin_inff();
  int id = (v < __builtin_inff()) ? j : 1 << 30;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const float ov = __shfl_xor(v, o, 64);
    const int oi = __shfl_xor(id, o, 64);
    if (ov < v || (ov == v && oi < id)) {
      v = ov;
      id = oi;
    }
  }
  return id >= NOUT ? -1 : id;
}

// One-sample inference by a whole 256-thread workgroup (the selection path
// inside the inspector and after fused expands). Every thread issues all of
// its weight loads before any arithmetic, so the three layers cost about one
// memory round trip; hidden layers are split over the four wavefronts
// (16 inputs each) and summed through LDS. Returns the argmin class over
// `mask` in every thread of wavefront 0 (others: undefined).
struct mlp_block_sh_t {
  float x[16];
  float h1[NH];
  float h2[NH];
  float part[BLOCK / WAVE][NH];
};

__device__ __forceinline__ int mlp_block_select(const float *w, uint32_t mask,
                                                mlp_block_sh_t &s) {
  const int t = threadIdx.x, j = t & (WAVE - 1), q = t >> LOG_WAVE;
  constexpr int KQ = NH / (BLOCK / WAVE);  // 16 inputs per wavefront
  float w1[NIN], w2[KQ], w3[KQ];
  float b1 = 0.f, b2 = 0.f, b3 = 0.f;
  if (q == 0) {
#pragma unroll
    for (int k = 0; k < NIN; ++k) w1[k] = w[OW1 + k * NH + j];
    b1 = w[OB1 + j];
    b2 = w[OB2 + j];
    if (j < NOUT) b3 = w[OB3 + j];
  }
#pragma unroll
  for (int k = 0; k < KQ; ++k) w2[k] = w[OW2 + (q * KQ + k) * NH + j];
#pragma unroll
  for (int k = 0; k < KQ; ++k)
    w3[k] = (j < NOUT) ? w[OW3 + (q * KQ + k) * NOUT + j] : 0.f;
  __syncthreads();  // s.x written by the caller
  if (q == 0) {
    float a = b1;
#pragma unroll
    for (int k = 0; k < NIN; ++k) a += s.x[k] * w1[k];
    s.h1[j] = a > 0.f ? a : 0.f;
  }
  __syncthreads();
  float a = 0.f;
#pragma unroll
  for (int k = 0; k < KQ; ++k) a += s.h1[q * KQ + k] * w2[k];
  s.part[q][j] = a;
  __syncthreads();
  if (q == 0) {
    float h = b2;
#pragma unroll
    for (int r = 0; r < BLOCK / WAVE; ++r) h += s.part[r][j];
    s.h2[j] = h > 0.f ? h : 0.f;
  }
  __syncthreads();
  a = 0.f;
#pragma unroll
  for (int k = 0; k < KQ; ++k) a += s.h2[q * KQ + k] * w3[k];
  s.part[q][j] = a;
  __syncthreads();
  float y = b3;
#pragma unroll
  for (int r = 0; r < BLOCK / WAVE; ++r) y += s.part[r][j];
  return wave_argmin(y, mask);
}

__device__ __forceinline__ void block_features(const sel_dev_t &d, float act,
                                               float inact, float pushw,
                                               float pullw, float maxa,
                                               mlp_block_sh_t &s) {
  if (threadIdx.x == 0) {
    counts_t c;
    c.n = d.n;
    c.m = d.m;
    c.avg_deg = d.avg_deg;
    c.std_deg = d.std_deg;
    c.gini = d.gini;
    c.level = d.level;
    c.active = act;
    c.inactive = inact;
    c.push_wl = pushw;
    c.pull_wl = pullw;
    c.max_deg_active = maxa;
    features(c, s.x);
  }
}

// One workgroup, launched on the engine's stream right after a pass that
// added the six inspector counters into the partial slots (the inspector,
// or a TM pull that folded the inspection in): wavefront 0 reduces the
// STAT_SLOTS slots (lane s reads slot s) into info[0..6), clears the slots
// for the next pass and, with the learned selector on (d.w set), turns the
// counters into the 12 features and writes the class the network picks to
// info[6], which the host reads with the counters. A kernel of its own
// rather than a tail of the pass, so the pass keeps its register budget.
__global__ void __launch_bounds__(BLOCK)
    k_stat_finish(sel_dev_t d, long long *slots, long long *info) {
  static_assert(STAT_SLOTS == WAVE, "one lane per partial slot");
  __shared__ mlp_block_sh_t s;
  __shared__ long long tot[6];
  if (threadIdx.x < WAVE) {
    long long *p = slots + (int64_t)threadIdx.x * STAT_STRIDE;
    long long v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = p[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) p[k] = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = wave_sum_any(v[k]);
    v[4] = wave_max_any(v[4]);
    v[5] = wave_max_any(v[5]);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        tot[k] = v[k];
        info[k] = v[k];
      }
    }
  }
  if (!d.w) return;  // uniform over the workgroup
  __syncthreads();
  block_features(d, (float)tot[0], (float)tot[1], (float)tot[2], (float)tot[3],
                 (float)tot[4], s);
  const int cls = mlp_block_select(d.w, d.mask, s);
  if (threadIdx.x == 0) info[6] = cls;
}

// ---- batch paths (training, tests) -----------------------------------------
__device__ __forceinline__ void load_sample_x(const float *x, int64_t i,
                                              mlp_wave_sh_t &s) {
  const int j = lane_id();
  if (j < NIN) s.x[j] = x[i * NIN + j];
  wave_lds_sync();
}

// y[i][0..30) for B samples, one wavefront per sample at a time
__global__ void __launch_bounds__(BLOCK)
    k_mlp_forward_batch(const float *w, const float *x, int64_t B, float *y) {
  __shared__ mlp_wave_sh_t sh[BLOCK / WAVE];
  mlp_wave_sh_t &s = sh[threadIdx.x >> LOG_WAVE];
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> LOG_WAVE;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> LOG_WAVE;
  for (int64_t i = wave; i < B; i += nwaves) {
    load_sample_x(x, i, s);
    const float v = mlp_wave_forward(w, s);
    if (lane_id() < NOUT) y[i * NOUT + lane_id()] = v;
    wave_lds_sync();
  }
}

// class[i] = argmin over `mask` of the forward pass (the inference path)
__global__ void __launch_bounds__(BLOCK)
    k_mlp_select_batch(const float *w, const float *x, int64_t B, uint32_t mask,
                       int32_t *cls) {
  // the workgroup-per-sample path of the inspector / fused selection
  __shared__ mlp_block_sh_t s;
  for (int64_t i = blockIdx.x; i < B; i += gridDim.x) {
    __syncthreads();  // the previous sample's reads of s are done
    if (threadIdx.x < NIN) s.x[threadIdx.x] = x[i * NIN + threadIdx.x];
    const int c = mlp_block_select(w, mask, s);
    if (threadIdx.x == 0) cls[i] = c;
  }
}

// Weights staged in LDS with transposed copies of W2 and W3 for the
// backward pass (lane j reads row j as a coalesced column).
struct mlp_lds_w_t {
  float w[NPARAM];
  float w2t[NH * NH];    // w2t[c][j] = W2[j][c]
  float w3t[NOUT * NH];  // w3t[c][j] = W3[j][c]
};

// Partial gradients of L = scale * sum_i sum_c mask (y - t)^2 over the
// samples idx[off + i] (idx == nullptr: off + i), i < B. Wavefront g
// handles samples g*per .. g*per+per-1 and writes its partial gradient to
// part[g][0..NPARAM) and its partial loss to part_loss[g].
__global__ void __launch_bounds__(BLOCK)
    k_mlp_grad_partial(const float *w, const float *x, const float *t,
                       const float *mask, const int32_t *idx, int64_t off,
                       int64_t B, int per, float scale, float *part,
                       float *part_loss) {
  extern __shared__ float dyn[];
  mlp_lds_w_t &L = *reinterpret_cast<mlp_lds_w_t *>(dyn);
  mlp_wave_sh_t *shw = reinterpret_cast<mlp_wave_sh_t *>(dyn + sizeof(mlp_lds_w_t) / 4);
  for (int p = threadIdx.x; p < NPARAM; p += blockDim.x) L.w[p] = w[p];
  for (int p = threadIdx.x; p < NH * NH; p += blockDim.x) {
    const int c = p / NH, j = p % NH;
    L.w2t[p] = w[OW2 + j * NH + c];
  }
  for (int p = threadIdx.x; p < NOUT * NH; p += blockDim.x) {
    const int c = p / NH, j = p % NH;
    L.w3t[p] = w[OW3 + j * NOUT + c];
  }
  __syncthreads();
  mlp_wave_sh_t &s = shw[threadIdx.x >> LOG_WAVE];
  const int j = lane_id();
  const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> LOG_WAVE;
  float g1[NIN], g2[NH], g3[NH];
  float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, lsum = 0.f;
#pragma unroll
  for (int k = 0; k < NIN; ++k) g1[k] = 0.f;
#pragma unroll
  for (int k = 0; k < NH; ++k) {
    g2[k] = 0.f;
    g3[k] = 0.f;
  }
  const int64_t i0 = g * per;
  const int64_t i1 = (i0 + per < B) ? i0 + per : B;
  for (int64_t i = i0; i < i1; ++i) {
    const int64_t r = idx ? (int64_t)idx[off + i] : off + i;
    load_sample_x(x, r, s);
    const float y = mlp_wave_forward(L.w, s);
    float dy = 0.f;
    if (j < NOUT) {
      const float mk = mask[r * NOUT + j];
      const float diff = y - t[r * NOUT + j];
      lsum += mk * diff * diff;
      dy = 2.f * scale * mk * diff;
      s.dy[j] = dy;
    }
    gb3 += dy;
#pragma unroll
    for (int k = 0; k < NH; ++k) g3[k] += s.h2[k] * dy;
    wave_lds_sync();
    float dh2 = 0.f;
    if (s.h2[j] > 0.f) {
#pragma unroll 6
      for (int c = 0; c < NOUT; ++c) dh2 += L.w3t[c * NH + j] * s.dy[c];
    }
    s.dh2[j] = dh2;
    gb2 += dh2;
#pragma unroll
    for (int k = 0; k < NH; ++k) g2[k] += s.h1[k] * dh2;
    wave_lds_sync();
    float dh1 = 0.f;
    if (s.h1[j] > 0.f) {
#pragma unroll 8
      for (int c = 0; c < NH; ++c) dh1 += L.w2t[c * NH + j] * s.dh2[c];
    }
    gb1 += dh1;
#pragma unroll
    for (int k = 0; k < NIN; ++k) g1[k] += s.x[k] * dh1;
    wave_lds_sync();
  }
  float *o = part + g * NPARAM;
#pragma unroll
  for (int k = 0; k < NIN; ++k) o[OW1 + k * NH + j] = g1[k];
  o[OB1 + j] = gb1;
#pragma unroll
  for (int k = 0; k < NH; ++k) o[OW2 + k * NH + j] = g2[k];
  o[OB2 + j] = gb2;
  if (j < NOUT) {
#pragma unroll
    for (int k = 0; k < NH; ++k) o[OW3 + k * NOUT + j] = g3[k];
    o[OB3 + j] = gb3;
  }
  lsum = wave_sum_any(lsum);
  if (j == 0) part_loss[g] = lsum;
}

constexpr size_t MLP_GRAD_LDS = sizeof(mlp_lds_w_t) + sizeof(mlp_wave_sh_t) * (BLOCK / WAVE);

// out[p] = sum over partials in ascending order (bit-reproducible)
__global__ void __launch_bounds__(BLOCK)
    k_mlp_reduce(const float *part, const float *part_loss, int nparts,
                 float *out, float *loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NPARAM) {
    float s = 0.f;
    for (int q = 0; q < nparts; ++q) s += part[(int64_t)q * NPARAM + p];
    out[p] = s;
  }
  if (p == 0 && loss) {
    float s = 0.f;
    for (int q = 0; q < nparts; ++q) s += part_loss[q];
    *loss = s;
  }
}

// Adam with bias correction folded into lr_t (computed on the host)
__global__ void __launch_bounds__(BLOCK)
    k_adam(float *w, const float *grad, float *m1, float *m2, int n, float lr_t,
           float b1, float b2, float eps) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float gr = grad[p];
  const float a = b1 * m1[p] + (1.f - b1) * gr;
  const float b = b2 * m2[p] + (1.f - b2) * gr * gr;
  m1[p] = a;
  m2[p] = b;
  w[p] -= lr_t * a / (sqrtf(b) + eps);
}

}  // namespace gsw
