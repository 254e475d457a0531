// Pass-limit network: from per-codeblock features (bitplane_kernel) to the
// predicted number of passes rate control keeps, in units of 8 passes.
//
// 16 -> 64 -> 64 -> 1 with ReLU. Layers 1 and 2 run on the matrix cores:
// v_mfma_f32_16x16x16bf16_1k, bf16 operands (weights and activations,
// rounded to nearest even), fp32 accumulation; the last layer is an fp32 dot
// product. One wavefront handles a tile of 16 codeblocks. Operand layout of
// the 16x16x16 instruction, lane l = 16 g + r (g = l / 16, r = l % 16):
//   A[m][k]: lane holds m = r, k = 4 g .. 4 g + 3
//   B[k][n]: lane holds n = r, k = 4 g .. 4 g + 3
//   D[m][n]: lane holds n = r, m = 4 g .. 4 g + 3
//
// Training (skipnet_train_kernel): one workgroup of four wavefronts; each
// step takes 256 samples (four tiles per wavefront), runs the forward pass,
// the backward pass (dH1 = dH2 W2, dW2 = dH2^T H1, dW1 = dH1^T X on the
// matrix cores, the rest in fp32), sums the gradients of the four
// wavefronts in LDS and applies Adam to the fp32 master weights.
// The loss is an asymmetric square, 4 (y - t)^2 below the target and
// (y - t)^2 above it, averaged over the batch: predicting too few passes
// costs further coding, too many only a little extra work.
#include <hip/hip_runtime.h>

#include <atomic>

#include <cstdio>
#include <utility>
#include <vector>

#include "gpu.hpp"

namespace bwc {
namespace gpu {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// flat parameter vector (fp32): W1[64][16] b1[64] W2[64][64] b2[64] w3[64] b3
// then feature means mu[16] and scales sc[16] (input = (f - mu) * sc)
constexpr int kW1 = 0, kB1 = 1024, kW2 = 1088, kB2 = 5184, kW3 = 5248, kB3 = 5312, kMu = 5313, kSc = 5329;
constexpr int kNTrain = 5313;   // trained parameters (W1 .. b3)
static_assert(kSc + 16 == kSkipModelSize, "parameter layout");

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even (finite values)
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Network state shared by the kernels (LDS): bf16 weight copies, fp32 biases,
// feature normalisation
struct NetLds {
  uint16_t W1[64 * 16];
  uint16_t W2[64 * 64];
  float b1[64], b2[64], w3[64], mu[16], sc[16];
  float b3;
};

__device__ void load_net(NetLds &n, const float *w) {
  for (int i = threadIdx.x; i < 1024; i += blockDim.x) n.W1[i] = to_bf16(w[kW1 + i]);
  for (int i = threadIdx.x; i < 4096; i += blockDim.x) n.W2[i] = to_bf16(w[kW2 + i]);
  for (int i = threadIdx.x; i < 64; i += blockDim.x) {
    n.b1[i] = w[kB1 + i];
    n.b2[i] = w[kB2 + i];
    n.w3[i] = w[kW3 + i];
  }
  for (int i = threadIdx.x; i < 16; i += blockDim.x) {
    n.mu[i] = w[kMu + i];
    n.sc[i] = w[kSc + i];
  }
  if (threadIdx.x == 0) n.b3 = w[kB3];
}

// Forward pass of one 16-row tile for this wavefront. X rows beyond `nrow`
// are zero. Leaves in the per-wave LDS buffers: Xb (16x16 bf16, normalised
// input), H1 (16x64 bf16); in registers: pre-activations z1, z2 (D layout,
// 4 column blocks) and the fp32 outputs y for rows 4g..4g+3 (on every lane of
// the group). `rel` is the tile-level mean subtracted from feature 15.
struct Fwd {
  f32x4 z1[4], z2[4];
  float y[4];
};

__device__ void forward_tile(const NetLds &n, const float *feat, uint64_t row0, uint64_t nrow, float rel,
                             uint16_t *Xb, uint16_t *H1, Fwd &F) {
  const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  // input tile, A layout: row r, features 4g..4g+3
  bf16x4 a;
  for (int j = 0; j < 4; ++j) {
    const int f = 4 * g + j;
    float x = 0.f;
    if (row0 + r < nrow) {
      x = feat[(row0 + r) * 16 + f];
      if (f == 15) x -= rel;
      x = (x - n.mu[f]) * n.sc[f];
    }
    const uint16_t h = to_bf16(x);
    a[j] = (short)h;
    Xb[r * 16 + f] = h;
  }
  // layer 1: B[k][n] = W1[unit = nb*16 + n][k]
  for (int nb = 0; nb < 4; ++nb) {
    bf16x4 b;
    for (int j = 0; j < 4; ++j) b[j] = (short)n.W1[(nb * 16 + r) * 16 + 4 * g + j];
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma16(a, b, acc);
    const float bias = n.b1[nb * 16 + r];
    for (int i = 0; i < 4; ++i) {
      acc[i] += bias;
      H1[(4 * g + i) * 64 + nb * 16 + r] = to_bf16(acc[i] > 0.f ? acc[i] : 0.f);
    }
    F.z1[nb] = acc;
  }
  __syncthreads();
  // layer 2: A[m = row][k = unit], B[k][n] = W2[unit2 = nb*16 + n][k]
  float part[4] = {0.f, 0.f, 0.f, 0.f};
  for (int nb = 0; nb < 4; ++nb) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kk = 0; kk < 4; ++kk) {
      bf16x4 a2, b2;
      for (int j = 0; j < 4; ++j) {
        a2[j] = (short)H1[r * 64 + kk * 16 + 4 * g + j];
        b2[j] = (short)n.W2[(nb * 16 + r) * 64 + kk * 16 + 4 * g + j];
      }
      acc = mfma16(a2, b2, acc);
    }
    const float bias = n.b2[nb * 16 + r], w3 = n.w3[nb * 16 + r];
    for (int i = 0; i < 4; ++i) {
      acc[i] += bias;
      part[i] += (acc[i] > 0.f ? acc[i] : 0.f) * w3;
    }
    F.z2[nb] = acc;
  }
  // layer 3: sum over the 16 lanes of the group (units r, r+16, r+32, r+48)
  for (int i = 0; i < 4; ++i) {
    float v = part[i];
    for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
    F.y[i] = v + n.b3;
  }
}

// ---------------------------------------------------------------------------
// inference: feat (n x 16) -> pass limits (passes = ceil(8 y) + margin,
// within 1 .. 3K - 2) or, when lim is null, the raw outputs y.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) skipnet_infer_kernel(const float *__restrict__ feat, uint64_t n,
                                                            const float *__restrict__ w,
                                                            const float *__restrict__ tstat,
                                                            const uint8_t *__restrict__ K, int margin,
                                                            uint8_t *__restrict__ lim, float *__restrict__ yout) {
  __shared__ NetLds net;
  __shared__ uint16_t Xb[4][16 * 16], H1[4][16 * 64];
  load_net(net, w);
  __syncthreads();
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  const uint64_t row0 = ((uint64_t)blockIdx.x * 4 + wv) * 16;
  const float rel = tstat ? tstat[0] / (tstat[1] > 0.f ? tstat[1] : 1.f) : 0.f;
  Fwd F;
  forward_tile(net, feat, row0, n, rel, Xb[wv], H1[wv], F);
  if (r == 0)
    for (int i = 0; i < 4; ++i) {
      const uint64_t row = row0 + 4 * g + i;
      if (row >= n) continue;
      if (yout) yout[row] = F.y[i];
      if (lim) {
        const int full = K[row] ? 3 * K[row] - 2 : 1;
        float p = ceilf(8.f * F.y[i]);
        p = p < 0.f ? 0.f : (p > 255.f ? 255.f : p);
        int z = (int)p + margin;
        z = z < 1 ? 1 : (z > full ? full : z);
        lim[row] = (uint8_t)z;
      }
    }
}

// ---------------------------------------------------------------------------
// training
// ---------------------------------------------------------------------------
struct TrainLds {
  NetLds net;
  uint16_t Xb[4][16 * 16], H1[4][16 * 64], dH2[4][16 * 64], dH1[4][16 * 64];
  float grad[kNTrain];
  float lossacc;
};

// feat (n x 16, feature 15 already relative), tgt (n) targets in units of 8
// passes, order (n) sample permutation; w: fp32 parameters (kSkipModelSize),
// m, v: Adam moments (kNTrain each), updated in place. mode 0: `steps`
// Adam steps over batches of 256 taken cyclically from `order` starting at
// `start`; mode 1: gradient of the batch order[start .. start + min(n, 256))
// into gout, no update. loss[step] = batch loss.
__global__ void __launch_bounds__(256) skipnet_train_kernel(const float *__restrict__ feat,
                                                            const float *__restrict__ tgt,
                                                            const uint32_t *__restrict__ order, uint64_t n,
                                                            uint64_t start, int steps, float lr, int t0, float *w,
                                                            float *m, float *v, float *gout,
                                                            float *__restrict__ loss, int mode) {
  extern __shared__ uint8_t lds_raw[];
  TrainLds &S = *reinterpret_cast<TrainLds *>(lds_raw);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  const uint64_t bsz = mode == 1 ? (n - start < 256 ? n - start : 256) : 256;
  for (int step = 0; step < steps; ++step) {
    load_net(S.net, w);
    for (int i = threadIdx.x; i < kNTrain; i += blockDim.x) S.grad[i] = 0.f;
    if (threadIdx.x == 0) S.lossacc = 0.f;
    __syncthreads();
    // register accumulators: dW2 (16 blocks, D layout), dW1 (4 blocks)
    f32x4 gW2[4][4], gW1[4];
    for (int a = 0; a < 4; ++a) {
      gW1[a] = f32x4{0.f, 0.f, 0.f, 0.f};
      for (int b = 0; b < 4; ++b) gW2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    for (int tl = 0; tl < 4; ++tl) {
      // gather this tile's 16 samples into a dense 16 x 16 block in LDS
      // (rows beyond the batch are zero and carry no loss)
      const uint64_t b0 = (uint64_t)(wv * 4 + tl) * 16;
      uint16_t *Xb = S.Xb[wv], *H1 = S.H1[wv], *dH2 = S.dH2[wv], *dH1 = S.dH1[wv];
      // forward_tile reads feat rows; map tile row -> sample through order[]
      float xrow[4];
      bool valid = false;
      uint64_t sidx = 0;
      {
        const uint64_t bi = b0 + r;
        valid = bi < bsz;
        sidx = valid ? order[(start + (uint64_t)step * 256 + bi) % n] : 0;
        for (int j = 0; j < 4; ++j) xrow[j] = valid ? feat[sidx * 16 + 4 * g + j] : 0.f;
      }
      // forward (same arithmetic as forward_tile, input from registers)
      Fwd F;
      {
        bf16x4 a;
        for (int j = 0; j < 4; ++j) {
          const int f = 4 * g + j;
          const float x = valid ? (xrow[j] - S.net.mu[f]) * S.net.sc[f] : 0.f;
          const uint16_t h = to_bf16(x);
          a[j] = (short)h;
          Xb[r * 16 + f] = h;
        }
        for (int nb = 0; nb < 4; ++nb) {
          bf16x4 b;
          for (int j = 0; j < 4; ++j) b[j] = (short)S.net.W1[(nb * 16 + r) * 16 + 4 * g + j];
          f32x4 acc = {0.f, 0.f, 0.f, 0.f};
          acc = mfma16(a, b, acc);
          const float bias = S.net.b1[nb * 16 + r];
          for (int i = 0; i < 4; ++i) {
            acc[i] += bias;
            H1[(4 * g + i) * 64 + nb * 16 + r] = to_bf16(acc[i] > 0.f ? acc[i] : 0.f);
          }
          F.z1[nb] = acc;
        }
      }
      __syncthreads();
      float part[4] = {0.f, 0.f, 0.f, 0.f};
      for (int nb = 0; nb < 4; ++nb) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kk = 0; kk < 4; ++kk) {
          bf16x4 a2, b2;
          for (int j = 0; j < 4; ++j) {
            a2[j] = (short)H1[r * 64 + kk * 16 + 4 * g + j];
            b2[j] = (short)S.net.W2[(nb * 16 + r) * 64 + kk * 16 + 4 * g + j];
          }
          acc = mfma16(a2, b2, acc);
        }
        const float bias = S.net.b2[nb * 16 + r], w3 = S.net.w3[nb * 16 + r];
        for (int i = 0; i < 4; ++i) {
          acc[i] += bias;
          part[i] += (acc[i] > 0.f ? acc[i] : 0.f) * w3;
        }
        F.z2[nb] = acc;
      }
      float dy[4];
      for (int i = 0; i < 4; ++i) {
        float vsum = part[i];
        for (int o = 1; o < 16; o <<= 1) vsum += __shfl_xor(vsum, o, 64);
        F.y[i] = vsum + S.net.b3;
        const uint64_t bi = b0 + (uint64_t)(4 * g + i);
        const uint64_t si = bi < bsz ? order[(start + (uint64_t)step * 256 + bi) % n] : 0;
        const float t = bi < bsz ? tgt[si] : 0.f;
        const float e = F.y[i] - t;
        const float wgt = e < 0.f ? 4.f : 1.f;
        dy[i] = bi < bsz ? 2.f * wgt * e / (float)bsz : 0.f;
        if (r == 0 && bi < bsz) atomicAdd(&S.lossacc, wgt * e * e / (float)bsz);
      }
      // dH2 = dy w3^T masked by z2 > 0 (D layout: rows 4g+i, unit nb*16 + r)
      for (int nb = 0; nb < 4; ++nb) {
        const int u = nb * 16 + r;
        const float w3 = S.net.w3[u];
        float gw3 = 0.f, gb2 = 0.f;
        for (int i = 0; i < 4; ++i) {
          const float z = F.z2[nb][i];
          const float h2 = z > 0.f ? z : 0.f;
          const float d = z > 0.f ? dy[i] * w3 : 0.f;
          gw3 += h2 * dy[i];
          gb2 += d;
          dH2[(4 * g + i) * 64 + u] = to_bf16(d);
        }
        // sum over the four lane groups (rows)
        gw3 += __shfl_xor(gw3, 16, 64);
        gw3 += __shfl_xor(gw3, 32, 64);
        gb2 += __shfl_xor(gb2, 16, 64);
        gb2 += __shfl_xor(gb2, 32, 64);
        if (g == 0) {
          atomicAdd(&S.grad[kW3 + u], gw3);
          atomicAdd(&S.grad[kB2 + u], gb2);
        }
      }
      {
        float gb3 = (dy[0] + dy[1] + dy[2] + dy[3]);
        gb3 += __shfl_xor(gb3, 16, 64);
        gb3 += __shfl_xor(gb3, 32, 64);
        if (lane == 0) atomicAdd(&S.grad[kB3], gb3);
      }
      __syncthreads();
      // dH1 = dH2 W2 (A = dH2 [row][u], B[k = u][n = v] = W2[u][v]) masked by z1 > 0
      for (int vb = 0; vb < 4; ++vb) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kk = 0; kk < 4; ++kk) {
          bf16x4 a2, b2;
          for (int j = 0; j < 4; ++j) {
            a2[j] = (short)dH2[r * 64 + kk * 16 + 4 * g + j];
            b2[j] = (short)S.net.W2[(kk * 16 + 4 * g + j) * 64 + vb * 16 + r];
          }
          acc = mfma16(a2, b2, acc);
        }
        const int vunit = vb * 16 + r;
        float gb1 = 0.f;
        for (int i = 0; i < 4; ++i) {
          const float d = F.z1[vb][i] > 0.f ? acc[i] : 0.f;
          gb1 += d;
          dH1[(4 * g + i) * 64 + vunit] = to_bf16(d);
        }
        gb1 += __shfl_xor(gb1, 16, 64);
        gb1 += __shfl_xor(gb1, 32, 64);
        if (g == 0) atomicAdd(&S.grad[kB1 + vunit], gb1);
      }
      __syncthreads();
      // dW2[u][v] += sum_row dH2[row][u] H1[row][v]  (K = 16 rows)
      for (int ub = 0; ub < 4; ++ub)
        for (int vb = 0; vb < 4; ++vb) {
          bf16x4 a2, b2;
          for (int j = 0; j < 4; ++j) {
            a2[j] = (short)dH2[(4 * g + j) * 64 + ub * 16 + r];
            b2[j] = (short)H1[(4 * g + j) * 64 + vb * 16 + r];
          }
          gW2[ub][vb] = mfma16(a2, b2, gW2[ub][vb]);
        }
      // dW1[v][f] += sum_row dH1[row][v] X[row][f]
      for (int vb = 0; vb < 4; ++vb) {
        bf16x4 a2, b2;
        for (int j = 0; j < 4; ++j) {
          a2[j] = (short)dH1[(4 * g + j) * 64 + vb * 16 + r];
          b2[j] = (short)Xb[(4 * g + j) * 16 + r];
        }
        gW1[vb] = mfma16(a2, b2, gW1[vb]);
      }
      __syncthreads();
    }
    // combine the wavefronts' matrix gradients (D layout: m = 4g + i, n = r)
    for (int ub = 0; ub < 4; ++ub)
      for (int vb = 0; vb < 4; ++vb)
        for (int i = 0; i < 4; ++i)
          atomicAdd(&S.grad[kW2 + (ub * 16 + 4 * g + i) * 64 + vb * 16 + r], gW2[ub][vb][i]);
    for (int vb = 0; vb < 4; ++vb)
      for (int i = 0; i < 4; ++i) atomicAdd(&S.grad[kW1 + (vb * 16 + 4 * g + i) * 16 + r], gW1[vb][i]);
    __syncthreads();
    if (threadIdx.x == 0 && loss) loss[step] = S.lossacc;
    if (mode == 1) {
      for (int i = threadIdx.x; i < kNTrain; i += blockDim.x) gout[i] = S.grad[i];
      return;
    }
    // Adam on the fp32 master weights
    const float b1c = 1.f - powf(0.9f, (float)(t0 + step + 1)), b2c = 1.f - powf(0.999f, (float)(t0 + step + 1));
    for (int i = threadIdx.x; i < kNTrain; i += blockDim.x) {
      const float gi = S.grad[i];
      const float mi = 0.9f * m[i] + 0.1f * gi;
      const float vi = 0.999f * v[i] + 0.001f * gi * gi;
      m[i] = mi;
      v[i] = vi;
      w[i] -= lr * (mi / b1c) / (sqrtf(vi / b2c) + 1e-8f);
    }
    __threadfence_block();
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
#define SN_OK(x)                                                                                      \
  do {                                                                                                \
    hipError_t e_ = (x);                                                                              \
    if (e_ != hipSuccess) {                                                                           \
      std::snprintf(err, errlen, "%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_));       \
      return 1;                                                                                       \
    }                                                                                                 \
  } while (0)

int skipnet_infer(hipStream_t s, const float *d_feat, uint64_t n, const float *d_w, const float *d_tstat,
                  const uint8_t *d_K, int margin, uint8_t *d_lim, float *d_y, char *err, size_t errlen) {
  if (!n) return 0;
  const uint32_t blocks = (uint32_t)((n + 63) / 64);
  hipLaunchKernelGGL(skipnet_infer_kernel, dim3(blocks), dim3(256), 0, s, d_feat, n, d_w, d_tstat, d_K, margin,
                     d_lim, d_y);
  SN_OK(hipGetLastError());
  return 0;
}

size_t skipnet_train_lds() { return sizeof(TrainLds); }

int skipnet_train(hipStream_t s, const float *d_feat, const float *d_tgt, const uint32_t *d_order, uint64_t n,
                  uint64_t start, int steps, float lr, int t0, float *d_w, float *d_m, float *d_v, float *d_g,
                  float *d_loss, int mode, char *err, size_t errlen) {
  static std::atomic<bool> attr{false};
  if (!attr.load()) {
    SN_OK(hipFuncSetAttribute((const void *)skipnet_train_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)sizeof(TrainLds)));
    attr.store(true);
  }
  hipLaunchKernelGGL(skipnet_train_kernel, dim3(1), dim3(256), sizeof(TrainLds), s, d_feat, d_tgt, d_order, n,
                     start, steps, lr, t0, d_w, d_m, d_v, d_g, d_loss, mode);
  SN_OK(hipGetLastError());
  return 0;
}

namespace {
// device buffers of one host-side call, freed on every exit
struct DevBufs {
  std::vector<void *> p;
  ~DevBufs() {
    for (void *q : p) (void)hipFree(q);
  }
  template <class T>
  hipError_t alloc(T **out, size_t count) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) + 16);
    if (e == hipSuccess) p.push_back(q);
    *out = static_cast<T *>(q);
    return e;
  }
};

uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
}  // namespace

int skipnet_fit(const float *feat, const float *tgt, uint64_t n, float *w, int epochs, float lr, uint64_t seed,
                float *loss_out, char *err, size_t errlen) {
  if (!n || epochs <= 0) return 0;
  DevBufs B;
  float *df, *dt, *dw, *dm, *dv, *dl;
  uint32_t *dord;
  const int steps = (int)((n + 255) / 256);
  SN_OK(B.alloc(&df, n * 16));
  SN_OK(B.alloc(&dt, n));
  SN_OK(B.alloc(&dw, kSkipModelSize));
  SN_OK(B.alloc(&dm, kNTrain));
  SN_OK(B.alloc(&dv, kNTrain));
  SN_OK(B.alloc(&dl, steps));
  SN_OK(B.alloc(&dord, n));
  SN_OK(hipMemcpy(df, feat, n * 16 * sizeof(float), hipMemcpyHostToDevice));
  SN_OK(hipMemcpy(dt, tgt, n * sizeof(float), hipMemcpyHostToDevice));
  SN_OK(hipMemcpy(dw, w, kSkipModelSize * sizeof(float), hipMemcpyHostToDevice));
  SN_OK(hipMemset(dm, 0, kNTrain * sizeof(float)));
  SN_OK(hipMemset(dv, 0, kNTrain * sizeof(float)));
  std::vector<uint32_t> order(n);
  std::vector<float> losses(steps);
  for (uint64_t i = 0; i < n; ++i) order[i] = (uint32_t)i;
  uint64_t s = seed;
  for (int ep = 0; ep < epochs; ++ep) {
    for (uint64_t i = n - 1; i > 0; --i) std::swap(order[i], order[splitmix(s) % (i + 1)]);
    SN_OK(hipMemcpy(dord, order.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (skipnet_train(nullptr, df, dt, dord, n, 0, steps, lr, ep * steps, dw, dm, dv, nullptr, dl, 0, err, errlen))
      return 1;
    SN_OK(hipMemcpy(losses.data(), dl, steps * sizeof(float), hipMemcpyDeviceToHost));
    double acc = 0;
    for (float l : losses) acc += l;
    if (loss_out) loss_out[ep] = (float)(acc / steps);
  }
  SN_OK(hipMemcpy(w, dw, kNTrain * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int skipnet_grad(const float *feat, const float *tgt, uint64_t n, const float *w, float *grad, float *loss,
                 char *err, size_t errlen) {
  if (!n || n > 256) {
    std::snprintf(err, errlen, "gradient batch must hold 1 to 256 samples");
    return 1;
  }
  DevBufs B;
  float *df, *dt, *dw, *dg, *dl, *dm, *dv;
  uint32_t *dord;
  SN_OK(B.alloc(&df, n * 16));
  SN_OK(B.alloc(&dt, n));
  SN_OK(B.alloc(&dw, kSkipModelSize));
  SN_OK(B.alloc(&dg, kNTrain));
  SN_OK(B.alloc(&dl, 1));
  SN_OK(B.alloc(&dm, kNTrain));
  SN_OK(B.alloc(&dv, kNTrain));
  SN_OK(B.alloc(&dord, n));
  std::vector<uint32_t> order(n);
  for (uint64_t i = 0; i < n; ++i) order[i] = (uint32_t)i;
  SN_OK(hipMemcpy(df, feat, n * 16 * sizeof(float), hipMemcpyHostToDevice));
  SN_OK(hipMemcpy(dt, tgt, n * sizeof(float), hipMemcpyHostToDevice));
  SN_OK(hipMemcpy(dw, w, kSkipModelSize * sizeof(float), hipMemcpyHostToDevice));
  SN_OK(hipMemcpy(dord, order.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (skipnet_train(nullptr, df, dt, dord, n, 0, 1, 0.f, 0, dw, dm, dv, dg, dl, 1, err, errlen)) return 1;
  SN_OK(hipMemcpy(grad, dg, kNTrain * sizeof(float), hipMemcpyDeviceToHost));
  if (loss) SN_OK(hipMemcpy(loss, dl, sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int skipnet_predict(const float *feat, uint64_t n, const float *w, float *y, char *err, size_t errlen) {
  if (!n) return 0;
  DevBufs B;
  float *df, *dw, *dy;
  SN_OK(B.alloc(&df, n * 16));
  SN_OK(B.alloc(&dw, kSkipModelSize));
  SN_OK(B.alloc(&dy, n));
  SN_OK(hipMemcpy(df, feat, n * 16 * sizeof(float), hipMemcpyHostToDevice));
  SN_OK(hipMemcpy(dw, w, kSkipModelSize * sizeof(float), hipMemcpyHostToDevice));
  if (skipnet_infer(nullptr, df, n, dw, nullptr, nullptr, 0, nullptr, dy, err, errlen)) return 1;
  SN_OK(hipMemcpy(y, dy, n * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

}  // namespace gpu
}  // namespace bwc
