// PCRD rate control on the device: the slope-threshold bisection of
// create_quality_layers (tier2.c:1115-1254), the per-threshold codeblock cuts
// and quality-layer size estimates of create_quality_layer (tier2.c:896-1079)
// with the packet-header estimate of create_packet (tier2.c:434-652: tag
// trees, delta_z and length codes, 0xFF bit stuffing), and the check that
// proves prefix codings give the cuts of complete codings.
//
// The records of every parameter of a tile stay in their device slots
// (Workspace::rec_slots). Codeblocks are numbered g in packet order: for each
// parameter, resolution and precinct (one packet P per layer), the codeblocks
// of every subband's precinct in turn, the order create_packet visits them.
//
// A probe evaluates one threshold T of layer q:
//   rc_cut_kernel    per (probe, codeblock): cut, body bytes, header bits
//   scan             exclusive scan of the header bits (bit offsets)
//   rc_write_kernel  per (probe, codeblock): its header bits into the packet's
//                    bit string (leading 1, tag-tree bits, delta_z, length)
//   rc_piece_kernel  per (probe, 64 Kibit piece): the byte count of JPEG 2000
//                    stuffing (after an 0xFF byte the next holds 7 bits) for
//                    each of the 16 states a piece can be entered in (next
//                    byte start 0-7 bits in, 8- or 7-bit byte), composed
//                    over the threads' chunks in LDS
//   rc_packet_kernel per (probe, packet): pieces composed from the packet's
//                    start; 1 byte for a packet no codeblock contributes to
// Every round evaluates the next five levels of the bisection's decision tree
// at once (31 thresholds); rc_advance_kernel walks it with the estimates as
// the reference's loop would and sets up the next round. The host launches a
// fixed sequence (rounds enough for a 16-bit range) and waits once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "gpu.hpp"

namespace bwc {
namespace gpu {

#define HIP_RC(x)                                                                                         \
  do {                                                                                                    \
    hipError_t e_ = (x);                                                                                  \
    if (e_ != hipSuccess) {                                                                               \
      std::snprintf(rc_err, sizeof rc_err, "%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
      return 1;                                                                                           \
    }                                                                                                     \
  } while (0)

static thread_local char rc_err[256];
const char *rc_last_error() { return rc_err; }

namespace {

constexpr int kTraceMax = 4096;
constexpr uint32_t kCbSlackBits = 96;              // dz flag, delta_z and length code bound per codeblock

// one codeblock in packet order
struct RcCb {
  uint32_t j;      // parameter
  uint32_t i;      // codeblock index within the parameter
  uint32_t P;      // packet (parameter, resolution, precinct)
  uint32_t leaf;   // node of its leaf in the tile's MSB tag trees
};

// device record arrays of one parameter slot
struct RcSlot {
  const uint8_t *K, *Kmsbs, *Z, *U;
  const uint32_t *L;
  const uint16_t *S;
  const double *mse, *bfut;
};

struct RcTrace {
  uint32_t layer, slope;
  int64_t est;
};

// bisection state of the tile being rate-controlled
struct RcState {
  int64_t est_tile, est_ql, target, hdr;
  int64_t f[32];            // estimates of this round's probes
  uint32_t T[32];           // this round's thresholds
  uint32_t smin, smax, tmp, slope, slope_min, slope_max;
  uint32_t m;               // probes in this round (0: nothing to evaluate)
  uint32_t done;            // the bisection of this layer has ended
  uint32_t layer, fallback, err, ntrace;
  RcTrace trace[kTraceMax];
};

__device__ __forceinline__ uint32_t ilog2u(uint32_t x) { return x ? 31u - (uint32_t)__builtin_clz(x) : 0u; }

// bits of the delta_z code (tier2.c:100-153); 293 or more new passes write none
__device__ __forceinline__ uint32_t dz_bits(int dz) {
  return dz == 1 ? 1 : dz == 2 ? 2 : dz < 6 ? 4 : dz < 37 ? 9 : dz < 293 ? 17 : 0;
}

// length code of a contribution of `len` bytes over dz passes with
// beta = beta_est (tier2.c:243-320): the run of ones, its 0, the length bits
struct LenCode {
  uint32_t ones, nb;
  uint16_t beta;   // beta after the code
};
__device__ __forceinline__ LenCode len_code(uint16_t beta, int dz, uint32_t len) {
  const uint32_t cpb = ilog2u((uint32_t)dz);
  const int32_t lb = (int32_t)ilog2u(len) + 1 - (int32_t)cpb;
  uint16_t maxb = beta;
  if ((int32_t)maxb < lb) maxb = (uint16_t)lb;
  LenCode c;
  c.ones = (uint32_t)(maxb - beta);
  c.beta = beta > maxb ? beta : maxb;
  const int nb = (int)c.beta + (int)cpb;
  c.nb = (nb > 0 && nb < 32) ? (uint32_t)nb : 0u;
  return c;
}

// Cut of codeblock g at threshold T (max{z : S[z] >= T}) and L[cut - 1] (0
// for a cut of 0): by bisection over its hull points when their slopes do
// not increase, else by a scan of S from the last coded pass down.
__device__ __forceinline__ int cut_at(const RcSlot &s, uint32_t i, int Z, const uint32_t *hz, const uint32_t *hl,
                                      uint64_t h0, uint32_t hc, bool scan, uint32_t T, uint32_t &len) {
  if (T >= 1 && !scan) {
    uint64_t lo = h0, hi = h0 + hc;   // the first point (pass 0) has a slope >= T
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) >> 1;
      if ((hz[mid] & 0xFFFFu) >= T) lo = mid;
      else hi = mid;
    }
    len = hl[lo];
    return (int)(hz[lo] >> 16);
  }
  const uint16_t *S = s.S + (uint64_t)i * (kMaxPasses + 1);
  int z = Z;
  while (z >= 0 && S[z] < T) --z;
  len = z > 0 ? s.L[(uint64_t)i * kMaxPasses + z - 1] : 0u;
  return z;
}

// ---------------------------------------------------------------------------
// set-up
// ---------------------------------------------------------------------------
__global__ void rc_state_init_kernel(RcState *st, int64_t hdr) {
  st->est_tile = 0;
  st->est_ql = 0;
  st->hdr = hdr;
  st->slope_min = 0xFFFF;
  st->slope_max = 0;
  st->m = 0;
  st->done = 1;
  st->fallback = 0;
  st->err = 0;
  st->ntrace = 0;
}

__global__ void rc_fill_u32_kernel(uint32_t *p, uint64_t n, uint32_t v) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
    p[k] = v;
}

// Per codeblock: its hull points (pass z << 16 | slope, L[z - 1]) at hoff[g],
// whether a scan is needed, Z and Kmsbs (0 for codeblocks without
// significant bits, which are never coded), the layer state, and the tile's
// slope bounds over nonzero slopes (tier1.c:3047-3061).
__global__ void rc_init_kernel(const RcCb *__restrict__ cb, uint64_t G, const RcSlot *__restrict__ slots,
                               const uint64_t *__restrict__ hoff, uint32_t *__restrict__ hz,
                               uint32_t *__restrict__ hl, uint32_t *__restrict__ hcnt, uint8_t *__restrict__ hscan,
                               uint8_t *__restrict__ kms, uint8_t *__restrict__ zg, int16_t *__restrict__ prev,
                               uint32_t *__restrict__ plen, uint16_t *__restrict__ beta, uint8_t *__restrict__ bad,
                               RcState *st) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= G) return;
  const RcCb c = cb[g];
  const RcSlot s = slots[c.j];
  const int K = s.K[c.i];
  const uint64_t o = hoff[g];
  uint32_t n = 0, smin = 0xFFFF, smax = 0;
  bool scan = false;
  int Z = 0;
  if (K == 0) {
    hz[o] = 0xFFFFu;
    hl[o] = 0;
    n = 1;
    kms[g] = 0;
  } else {
    Z = s.Z[c.i];
    kms[g] = s.Kmsbs[c.i];
    const uint16_t *S = s.S + (uint64_t)c.i * (kMaxPasses + 1);
    const uint32_t *L = s.L + (uint64_t)c.i * kMaxPasses;
    uint32_t pv = 0xFFFF;
    for (int z = 0; z <= Z; ++z) {
      const uint32_t v = S[z];
      if (!v) continue;
      if (v > pv) scan = true;
      pv = v;
      hz[o + n] = ((uint32_t)z << 16) | v;
      hl[o + n] = z > 0 ? L[z - 1] : 0u;
      ++n;
      smin = v < smin ? v : smin;
      smax = v > smax ? v : smax;
    }
    if (S[0] == 0) scan = true;
  }
  zg[g] = (uint8_t)Z;
  hcnt[g] = n;
  hscan[g] = scan;
  prev[g] = 0;
  plen[g] = 0;
  beta[g] = 3;
  bad[g] = 0;
  if (K && smin <= smax) {
    atomicMin(&st->slope_min, smin);
    atomicMax(&st->slope_max, smax);
  }
}

// MSB tag-tree values as tagtree_set_value leaves them once every leaf of a
// tree holds its codeblock's Kmsbs (tagtree.c:159-181): the minimum over
// the leaves below each node
__global__ void rc_msbs_set_kernel(const RcCb *__restrict__ cb, uint64_t G, const uint8_t *__restrict__ kms,
                                   const int32_t *__restrict__ parent, uint32_t *__restrict__ nodeval) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= G) return;
  const uint32_t v = kms[g];
  for (int32_t n = (int32_t)cb[g].leaf; n >= 0; n = parent[n]) atomicMin(&nodeval[n], v);
}

// MSB tag-tree bits of each codeblock's first contribution in an estimate
// (encode_tagtree for thresholds 1 .. Kmsbs + 1, tagtree.c:310-373; an
// estimate leaves node thresholds at 0): a run of zeros per node up to its
// value and a 1 when the value lies below the threshold. Into the
// codeblock's own `stride` words, MSB first; mcn[g] = bit count.
__global__ void rc_msbs_cache_kernel(const RcCb *__restrict__ cb, uint64_t G, const uint8_t *__restrict__ kms,
                                     const int32_t *__restrict__ parent, const uint32_t *__restrict__ nodeval,
                                     uint32_t stride, uint32_t *__restrict__ mcw, uint32_t *__restrict__ mcn) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= G) return;
  uint32_t *w = mcw + g * stride;
  for (uint32_t k = 0; k < stride; ++k) w[k] = 0;
  int32_t path[64];
  int np = 0;
  for (int32_t n = (int32_t)cb[g].leaf; n >= 0 && np < 64; n = parent[n]) path[np++] = n;
  const uint32_t T = kms[g] + 1u;
  uint32_t pos = 0;
  for (uint32_t t = 1; t <= T; ++t) {
    uint32_t tmin = 0;
    for (int a = np - 1; a >= 0; --a) {
      const uint32_t v = nodeval[path[a]];
      uint32_t th = tmin;
      if (v >= th && th < t) {
        const uint32_t hi = v < t ? v : t;
        pos += hi - th;
        if (v < t) {
          if (pos < stride * 32) w[pos >> 5] |= 1u << (31 - (pos & 31));
          ++pos;
        }
        th = (v + 1 < t) ? v + 1 : t;
      }
      tmin = v < th ? v : th;
    }
  }
  mcn[g] = pos;
}

// ---------------------------------------------------------------------------
// bisection control (one thread)
// ---------------------------------------------------------------------------
// thresholds of the next `depth` levels of the bisection's decision tree from
// (smin, smax), heap ordered: node n's children are 2n + 1 (estimate above
// the target: smin = slope) and 2n + 2 (below: smax = slope)
__device__ void rc_frontier(RcState *st, int depth) {
  uint32_t a[64], b[64];
  const uint32_t nn = (1u << depth) - 1;
  a[0] = st->smin;
  b[0] = st->smax;
  for (uint32_t n = 0; n < nn; ++n) {
    const uint32_t lo = a[n], hi = b[n];
    uint32_t s = lo;
    if (lo < hi) s = (lo + hi) >> 1;
    st->T[n] = s;
    st->f[n] = 0;
    if (2 * n + 2 < 64) {
      a[2 * n + 1] = s;
      b[2 * n + 1] = (lo < hi && lo == s) ? s : hi;
      a[2 * n + 2] = lo;
      b[2 * n + 2] = (lo < hi) ? s : hi;
    }
  }
  st->m = nn;
}

// start of layer l: the target and the bisection range (tier2.c:1160-1190)
__global__ void rc_layer_begin_kernel(RcState *st, uint32_t l, int64_t target, int depth) {
  st->layer = l;
  st->target = target;
  st->done = 1;
  st->m = 0;
  if (target > 0) {
    st->smin = st->slope_min;
    st->smax = st->slope_max;
    st->tmp = st->slope_min;
    st->slope = st->smin;
    if (st->smin < st->smax) {
      st->done = 0;
      rc_frontier(st, depth);
    }
  } else {
    st->slope = 0;
  }
}

// Walk this round's decision tree with its estimates, as the reference's
// loop probes (tier2.c:1193-1235); set up the next round where it leaves it.
__global__ void rc_advance_kernel(RcState *st, int depth) {
  if (st->done) {
    st->m = 0;
    return;
  }
  uint32_t n = 0;
  const int64_t target = st->target;
  while (st->smin < st->smax) {
    const uint32_t slope = (st->smax + st->smin) >> 1;
    if (n >= st->m) break;
    if (st->T[n] != slope) st->err |= 1;
    st->slope = slope;
    const int64_t est = st->f[n] + st->est_tile + st->hdr;
    st->est_ql = est;
    const uint32_t k = st->ntrace;
    if (k < kTraceMax) st->trace[k] = RcTrace{st->layer, slope, est};
    st->ntrace = k + 1;
    if (est > target) {
      if (st->smin == slope) st->smax = slope;
      st->smin = slope;
      n = 2 * n + 1;
    } else if (est < target) {
      if (st->smax == slope) st->smin = slope;
      st->smax = slope;
      st->tmp = slope;
      n = 2 * n + 2;
    } else {
      st->smax = st->smin = slope;
    }
  }
  if (st->smin >= st->smax) {
    st->done = 1;
    st->m = 0;
    return;
  }
  rc_frontier(st, depth);
}

// End of the bisection (tier2.c:1232-1235), then the probes that follow it:
// mode 1 the check's thresholds {slope, slope - 1}, mode 2 the final
// layer's estimate at the slope, mode 0 none.
__global__ void rc_bisect_end_kernel(RcState *st) {
  if (!st->done) st->err |= 2;
  if (st->target > 0 && st->est_ql > st->target) st->slope = st->tmp > st->slope_min ? st->tmp : st->slope_min;
  st->m = 0;
}

__global__ void rc_setup_probes_kernel(RcState *st, int mode) {
  st->m = 0;
  if (mode == 1 && st->target > 0) {
    st->T[0] = st->slope;
    st->f[0] = 0;
    st->m = 1;
    if (st->slope > 0) {
      st->T[1] = st->slope - 1;
      st->f[1] = 0;
      st->m = 2;
    }
  } else if (mode == 2) {
    st->T[0] = st->slope;
    st->f[0] = 0;
    st->m = 1;
  }
}

// ---------------------------------------------------------------------------
// probes
// ---------------------------------------------------------------------------
struct ProbeBufs {
  int16_t *cur;   // cut per (probe, codeblock)
  uint32_t *zl;   // L[cut - 1]
  uint32_t *nb;   // header bits (the packet's leading 1 with its first codeblock)
  uint32_t *off;  // exclusive scan of nb
  int64_t *body;  // per (probe, packet): body bytes its codeblocks add to the layer's estimate
  uint32_t *any;  // per (probe, packet): some codeblock contributes new passes
};

// zero the header bit strings and the per-packet sums of a round
__global__ void rc_round_zero_kernel(uint32_t *bits, uint64_t nw, int64_t *body, uint32_t *any, uint64_t nk) {
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < nw; k += step) bits[k] = 0;
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < nk; k += step) body[k] = 0, any[k] = 0;
}

// Sum v and OR a over the lanes of the wavefront whose key equals lane 0's
// (consecutive codeblocks share a packet); lanes with another key add
// their own. key ~0: nothing to add.
__device__ __forceinline__ void packet_sums(uint64_t key, int64_t v, uint32_t a, int64_t *body, uint32_t *any) {
  const uint64_t k0 = __shfl(key, 0);
  const bool same = key == k0;
  int64_t sv = same ? v : 0;
  uint32_t sa = same ? a : 0u;
  for (int d = 32; d >= 1; d >>= 1) {
    sv += __shfl_xor(sv, d);
    sa |= __shfl_xor(sa, d);
  }
  if (__lane_id() == 0 && k0 != ~0ull) {
    if (sv) atomicAdd((unsigned long long *)&body[k0], (unsigned long long)sv);
    if (sa) atomicOr(&any[k0], 1u);
  }
  if (!same && key != ~0ull) {
    if (v) atomicAdd((unsigned long long *)&body[key], (unsigned long long)v);
    if (a) atomicOr(&any[key], 1u);
  }
}

__global__ void __launch_bounds__(256) rc_cut_kernel(
    const RcState *__restrict__ st, uint32_t mcap, const RcCb *__restrict__ cb, uint64_t G, uint32_t NP,
    const RcSlot *__restrict__ slots, const uint64_t *__restrict__ hoff, const uint32_t *__restrict__ hz,
    const uint32_t *__restrict__ hl, const uint32_t *__restrict__ hcnt, const uint8_t *__restrict__ hscan,
    const uint8_t *__restrict__ zg, const int16_t *__restrict__ prev, const uint32_t *__restrict__ plen,
    const uint16_t *__restrict__ beta, const uint32_t *__restrict__ mcn, const uint32_t *__restrict__ pk_first,
    ProbeBufs pb) {
  const uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint32_t t = (uint32_t)(idx / G);
  const uint64_t g = idx - (uint64_t)t * G;
  uint64_t key = ~0ull;
  int64_t bd = 0;
  uint32_t contrib = 0;
  if (idx < (uint64_t)mcap * G && t >= st->m) pb.nb[idx] = 0;
  if (idx < (uint64_t)mcap * G && t < st->m) {
    const uint32_t T = st->T[t], q = st->layer;
    const RcCb c = cb[g];
    const RcSlot s = slots[c.j];
    uint32_t zl;
    const int z = cut_at(s, c.i, zg[g], hz, hl, hoff[g], hcnt[g], hscan[g], T, zl);
    const int pv = prev[g];
    int cur;
    if (z < pv) {
      cur = pv;
      if (cur == 1) bd = s.L[(uint64_t)c.i * kMaxPasses];
    } else {
      cur = z;
      if (z > 0) {
        bd = zl;
        if (z > 1 && pv > 0) bd -= plen[g];
      }
    }
    const int dz = cur - pv;
    uint32_t n = (pk_first[c.P] == g) ? 1u : 0u;
    if (pv == 0) n += (q + 1) + (cur ? mcn[g] : 0u);
    else n += 1;
    if (dz) {
      const LenCode lc = len_code(beta[g], dz, zl - (pv > 0 ? plen[g] : 0u));
      n += dz_bits(dz) + lc.ones + 1 + lc.nb;
    }
    pb.cur[idx] = (int16_t)cur;
    pb.zl[idx] = zl;
    pb.nb[idx] = n;
    key = (uint64_t)t * NP + c.P;
    contrib = dz > 0;
  }
  packet_sums(key, bd, contrib, pb.body, pb.any);
}

// exclusive scan of n u32 values, 1024 per workgroup: local scans and the
// workgroup totals, the totals scanned by one workgroup, then added back
__global__ void __launch_bounds__(256) rc_scan_local_kernel(const uint32_t *__restrict__ in, uint32_t *out,
                                                             uint64_t n, uint32_t *bsum) {
  __shared__ uint32_t part[256];
  const uint64_t base = blockIdx.x * 1024ull + threadIdx.x * 4ull;
  uint32_t v[4], s = 0;
  for (int k = 0; k < 4; ++k) {
    v[k] = base + k < n ? in[base + k] : 0u;
    s += v[k];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    const uint32_t x = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0u;
    __syncthreads();
    part[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - s;
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) bsum[blockIdx.x] = part[255];
}

__global__ void __launch_bounds__(1024) rc_scan_top_kernel(uint32_t *bsum, uint64_t nb) {
  __shared__ uint32_t part[1024];
  uint32_t carry = 0;
  for (uint64_t b0 = 0; b0 < nb; b0 += 1024) {
    const uint64_t k = b0 + threadIdx.x;
    const uint32_t v = k < nb ? bsum[k] : 0u;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
      const uint32_t x = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0u;
      __syncthreads();
      part[threadIdx.x] += x;
      __syncthreads();
    }
    if (k < nb) bsum[k] = carry + part[threadIdx.x] - v;
    const uint32_t tot = part[1023];
    __syncthreads();
    carry += tot;
  }
}

__global__ void rc_scan_add_kernel(uint32_t *out, uint64_t n, const uint32_t *bsum) {
  const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (k < n) out[k] += bsum[k >> 10];
}

// OR the n low bits of v (n <= 32), MSB first, at bit position pos of a
// string of big-endian 32-bit words
__device__ __forceinline__ void put_bits(uint32_t *w, uint64_t pos, uint32_t v, uint32_t n) {
  if (!n) return;
  v &= (n == 32) ? 0xFFFFFFFFu : ((1u << n) - 1u);
  if (!v) return;
  const uint64_t wi = pos >> 5;
  const uint32_t sh = (uint32_t)(pos & 31);
  const uint64_t x = (uint64_t)v << (64 - sh - n);
  const uint32_t hi = (uint32_t)(x >> 32), lo = (uint32_t)x;
  if (hi) atomicOr(&w[wi], hi);
  if (lo) atomicOr(&w[wi + 1], lo);
}

__global__ void rc_write_kernel(const RcState *__restrict__ st, uint32_t mcap, const RcCb *__restrict__ cb,
                                uint64_t G, const int16_t *__restrict__ prev, const uint32_t *__restrict__ plen,
                                const uint16_t *__restrict__ beta, const uint32_t *__restrict__ mcn,
                                const uint32_t *__restrict__ mcw, uint32_t stride,
                                const uint32_t *__restrict__ pk_first, const uint64_t *__restrict__ pbase,
                                uint64_t capT, ProbeBufs pb, uint32_t *__restrict__ bits) {
  const uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (idx >= (uint64_t)mcap * G) return;
  const uint32_t t = (uint32_t)(idx / G);
  const uint64_t g = idx - (uint64_t)t * G;
  if (t >= st->m) return;
  const uint32_t q = st->layer;
  const RcCb c = cb[g];
  const uint64_t f0 = pk_first[c.P];
  uint64_t pos = (uint64_t)t * capT + pbase[c.P] + (uint32_t)(pb.off[idx] - pb.off[(uint64_t)t * G + f0]);
  if (f0 == g) {
    put_bits(bits, pos, 1, 1);
    ++pos;
  }
  const int pv = prev[g], cur = pb.cur[idx], dz = cur - pv;
  if (pv == 0) {
    pos += q + 1;   // inclusion tag tree: its values are unset (0xFFFF) during estimates
    if (cur) {
      const uint32_t n = mcn[g];
      const uint32_t *src = mcw + g * stride;
      for (uint32_t k = 0; 32 * k < n; ++k) put_bits(bits, pos + 32 * k, src[k], 32);
      pos += n;
    }
  } else {
    put_bits(bits, pos, dz ? 1u : 0u, 1);
    ++pos;
  }
  if (dz) {
    if (dz == 1) {
      pos += 1;
    } else if (dz == 2) {
      put_bits(bits, pos, 2, 2);
      pos += 2;
    } else if (dz < 6) {
      put_bits(bits, pos, 0xCu | (uint32_t)(dz - 3), 4);
      pos += 4;
    } else if (dz < 37) {
      put_bits(bits, pos, (0xFu << 5) | (uint32_t)(dz - 6), 9);
      pos += 9;
    } else if (dz < 293) {
      put_bits(bits, pos, (0x1FFu << 8) | (uint32_t)(dz - 37), 17);
      pos += 17;
    }
    const uint32_t len = pb.zl[idx] - (pv > 0 ? plen[g] : 0u);
    const LenCode lc = len_code(beta[g], dz, len);
    for (uint32_t k = lc.ones; k > 0;) {
      const uint32_t n = k < 32 ? k : 32;
      put_bits(bits, pos, 0xFFFFFFFFu, n);
      pos += n;
      k -= n;
    }
    pos += 1;
    put_bits(bits, pos, len, lc.nb);
  }
}

// ---------------------------------------------------------------------------
// byte count with 0xFF stuffing
// ---------------------------------------------------------------------------
// The header writer (tier2.c, bitstream.c) fills bytes MSB first; after an
// 0xFF byte the next byte holds 7 bits. Over a range of a packet's header
// string, the bytes it starts depend on the state it is entered in: the next
// byte starts o = 0..7 bits in and holds 8 or 7 bits (e = o | (cap == 7) << 3).
// A range function maps each of the 16 entry states to the exit state and the
// bytes started in the range; starts at or past the packet's end pe are not
// bytes, except a 7-bit byte starting exactly at pe (the writer's flush after
// a final 0xFF). Functions of consecutive ranges compose.
//
// The chunk functions below work on a thread's kChunkWords words (plus one
// word of lookahead) held in registers: wv[k] is word k of the chunk.
constexpr int kChunkWords = 8;

// whether 8 ones start at some bit p < lim of the chunk
__device__ __forceinline__ bool chunk_has_ff(const uint32_t (&wv)[kChunkWords + 1], uint64_t lim) {
  bool any = false;
#pragma unroll
  for (int k = 0; k < kChunkWords; ++k) {
    const uint64_t x = ((uint64_t)wv[k] << 32) | (uint64_t)wv[k + 1];
    uint64_t y = x;
#pragma unroll
    for (int j = 1; j < 8; ++j) y &= x << j;
    // bit 63 - p of y: 8 ones from bit p of word k; keep 32k + p < lim
    uint32_t m = (uint32_t)(y >> 32);
    const int64_t room = (int64_t)lim - 32 * k;
    if (room <= 0) m = 0;
    else if (room < 32) m &= ~(0xFFFFFFFFu >> room);
    any |= m != 0;
  }
  return any;
}

// Range function of [a, b) for entry state e when no 0xFF byte can start in
// it: bytes at a + o, then every 8 bits after the first (c bits) one.
__device__ __forceinline__ void plain_range(uint32_t e, uint64_t a, uint64_t b, uint64_t pe, uint8_t &ex,
                                            uint32_t &cn) {
  if (a >= pe) {
    ex = (uint8_t)e, cn = 0;
    return;
  }
  const uint64_t lim = b < pe ? b : pe;
  uint64_t next = a + (e & 7);
  uint32_t cap = (e & 8) ? 7 : 8, cnt = 0;
  if (next < lim) {
    next += cap;
    cap = 8;
    cnt = 1;
    if (next < lim) {
      const uint64_t k = (lim - 1 - next) / 8 + 1;
      cnt += (uint32_t)k;
      next += 8 * k;
    }
  }
  if (pe <= b && next == pe && cap == 7) ++cnt;
  ex = (uint8_t)((next >= b ? (next - b) & 7 : 0) | (cap == 7 ? 8 : 0));
  cn = cnt;
}

// the same for a chunk starting at bit a of the packet, byte by byte (where
// 0xFF bytes can start)
__device__ __forceinline__ void step_chunk(const uint32_t (&wv)[kChunkWords + 1], uint32_t e, uint64_t a, uint64_t pe,
                                           uint8_t &ex, uint32_t &cn) {
  const uint32_t span = kChunkWords * 32;
  const uint32_t lim = (uint32_t)(pe - a < span ? pe - a : span);   // pe > a
  uint32_t pos = e & 7, cap = (e & 8) ? 7 : 8, cnt = 0;
#pragma unroll
  for (int k = 0; k < kChunkWords; ++k) {
    const uint64_t x = ((uint64_t)wv[k] << 32) | (uint64_t)wv[k + 1];
    while (pos < 32u * (k + 1) && pos < lim) {
      ++cnt;
      const uint32_t next = (cap == 8 && ((x >> (56 - (pos - 32u * k))) & 0xFFu) == 0xFFu) ? 7 : 8;
      pos += cap;
      cap = next;
    }
  }
  if (a + span >= pe && pos == lim && cap == 7) ++cnt;
  ex = (uint8_t)((pos >= span ? (pos - span) & 7 : 0) | (cap == 7 ? 8 : 0));
  cn = cnt;
}

constexpr uint64_t kPieceSpan = 256ull * kChunkWords * 32;   // bits per rc_piece_kernel workgroup

// Per (probe, piece of kPieceSpan bits of one packet's header string): the
// piece's range function, from the threads' 256-bit chunks composed in LDS;
// a piece where no 0xFF byte can start is computed in one go. Pieces at or
// past the packet's end are skipped (rc_packet_kernel stops before them).
__global__ void __launch_bounds__(256) rc_piece_kernel(const RcState *__restrict__ st,
                                                        const uint32_t *__restrict__ bits, uint64_t capT,
                                                        const uint64_t *__restrict__ pbase,
                                                        const uint32_t *__restrict__ piece_first, uint32_t NP,
                                                        uint32_t npiece, const uint32_t *__restrict__ pk_first,
                                                        uint64_t G, ProbeBufs pb, uint8_t *__restrict__ pf_ex,
                                                        uint32_t *__restrict__ pf_cn) {
  __shared__ uint8_t ex[256][16];
  __shared__ uint32_t cn[256][16];
  const uint32_t t = blockIdx.y, b = blockIdx.x;
  if (t >= st->m || b >= piece_first[NP]) return;
  uint32_t lo = 0, hi = NP;   // packet: piece_first[P] <= b < piece_first[P + 1]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (piece_first[mid] <= b) lo = mid;
    else hi = mid;
  }
  const uint32_t P = lo;
  const uint64_t a0 = pk_first[P], a1 = pk_first[P + 1];
  uint64_t pe = 0;   // the packet's header bits
  if (a1 > a0) {
    const uint64_t r0 = (uint64_t)t * G + a0, r1 = (uint64_t)t * G + a1 - 1;
    pe = (uint32_t)(pb.off[r1] + pb.nb[r1] - pb.off[r0]);
  }
  const uint64_t ps = (uint64_t)(b - piece_first[P]) * kPieceSpan;
  if (ps >= pe) return;
  const uint32_t *w = bits + (((uint64_t)t * capT + pbase[P]) >> 5);   // the packet's first word
  const uint64_t nw = (pbase[P + 1] - pbase[P]) >> 5;                   // words of its region
  const uint64_t ca = ps + (uint64_t)threadIdx.x * kChunkWords * 32, cb = ca + kChunkWords * 32;
  uint32_t wv[kChunkWords + 1];
#pragma unroll
  for (int k = 0; k <= kChunkWords; ++k) {
    const uint64_t wi = (ca >> 5) + k;
    wv[k] = (ca < pe && wi < nw) ? w[wi] : 0u;
  }
  const bool hot = ca < pe && chunk_has_ff(wv, pe - ca);
  if (!__syncthreads_or(hot)) {
    if (threadIdx.x < 16) {
      uint8_t e8;
      uint32_t c32;
      plain_range(threadIdx.x, ps, ps + kPieceSpan, pe, e8, c32);
      const uint64_t o = ((uint64_t)t * npiece + b) * 16 + threadIdx.x;
      pf_ex[o] = e8;
      pf_cn[o] = c32;
    }
    return;
  }
  for (uint32_t e = 0; e < 16; ++e) {
    uint8_t e8;
    uint32_t c32;
    if (hot) step_chunk(wv, e, ca, pe, e8, c32);
    else plain_range(e, ca, cb, pe, e8, c32);
    ex[threadIdx.x][e] = e8;
    cn[threadIdx.x][e] = c32;
  }
  __syncthreads();
  for (uint32_t d = 1; d < 256; d <<= 1) {
    uint8_t e8[16];
    uint32_t c32[16];
    const bool lead = (threadIdx.x & (2 * d - 1)) == 0;
    if (lead)
      for (int e = 0; e < 16; ++e) {
        const uint8_t m = ex[threadIdx.x][e];
        e8[e] = ex[threadIdx.x + d][m];
        c32[e] = cn[threadIdx.x][e] + cn[threadIdx.x + d][m];
      }
    __syncthreads();
    if (lead)
      for (int e = 0; e < 16; ++e) ex[threadIdx.x][e] = e8[e], cn[threadIdx.x][e] = c32[e];
    __syncthreads();
  }
  if (threadIdx.x < 16) {
    const uint64_t o = ((uint64_t)t * npiece + b) * 16 + threadIdx.x;
    pf_ex[o] = ex[0][threadIdx.x];
    pf_cn[o] = cn[0][threadIdx.x];
  }
}

// Per (probe, packet), one wavefront: the header bytes from its pieces
// composed from the packet's start, or 1 byte for a packet no codeblock
// contributes to (tier2.c:447-459), plus the body bytes its codeblocks add;
// into the probe's estimate (and per packet into pkt when given).
__global__ void __launch_bounds__(64) rc_packet_kernel(RcState *st, const uint32_t *__restrict__ piece_first,
                                                       uint32_t npiece, const uint32_t *__restrict__ pk_first,
                                                       uint64_t G, uint32_t NP, ProbeBufs pb,
                                                       const uint8_t *__restrict__ pf_ex,
                                                       const uint32_t *__restrict__ pf_cn, int64_t *pkt) {
  __shared__ uint8_t sex[64][16];
  __shared__ uint32_t scn[64][16];
  const uint32_t t = blockIdx.y, P = blockIdx.x;
  if (t >= st->m) return;
  const uint64_t key = (uint64_t)t * NP + P;
  int64_t bytes = 1;
  if (pb.any[key]) {
    const uint64_t a0 = pk_first[P], a1 = pk_first[P + 1];
    const uint64_t r0 = (uint64_t)t * G + a0, r1 = (uint64_t)t * G + a1 - 1;
    const uint64_t pe = (uint32_t)(pb.off[r1] + pb.nb[r1] - pb.off[r0]);
    const uint32_t k1 = piece_first[P] + (uint32_t)((pe + kPieceSpan - 1) / kPieceSpan);
    uint32_t e = 0, total = 0;
    for (uint32_t k0 = piece_first[P]; k0 < k1; k0 += 64) {
      const uint32_t k = k0 + threadIdx.x;
      if (k < k1) {
        const uint64_t o = ((uint64_t)t * npiece + k) * 16;
        for (int s = 0; s < 16; ++s) sex[threadIdx.x][s] = pf_ex[o + s], scn[threadIdx.x][s] = pf_cn[o + s];
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        const uint32_t kn = k1 - k0 < 64 ? k1 - k0 : 64;
        for (uint32_t x = 0; x < kn; ++x) {
          total += scn[x][e];
          e = sex[x][e];
        }
      }
      __syncthreads();
    }
    bytes = total;
  }
  if (threadIdx.x == 0) {
    const int64_t v = bytes + pb.body[key];
    if (pkt) pkt[key] = bytes;
    atomicAdd((unsigned long long *)&st->f[t], (unsigned long long)v);
  }
}

// ---------------------------------------------------------------------------
// the pass-skip check and the layer's final cuts
// ---------------------------------------------------------------------------
// a slope that slope2log maps to t or above is at least 2^e (see the host
// restatement in tier2.cpp): 0 below 3
__device__ __forceinline__ double slope_floor(uint32_t t) {
  if (t <= 2) return 0.0;
  const int num = (int)t - 65536 - 255;
  const int e = 32 + (num >= 0 ? (num + 255) / 256 : -((-num) / 256));
  return ldexp(1.0, e);
}

// Per codeblock coded to a prefix of its passes: its cuts at slope - 1 and
// slope must lie `margin` passes before its first unsettled pass U, and no
// point after them may reach those slopes in a complete coding (the bound
// on the distortion of the passes from U on and of the uncoded ones). The
// tile-level test: the estimate at slope is below the target and the one at
// slope - 1 above it (probes 0 and 1 of the check round).
__global__ void rc_check_kernel(RcState *st, const RcCb *__restrict__ cb, uint64_t G,
                                const RcSlot *__restrict__ slots, const uint64_t *__restrict__ hoff,
                                const uint32_t *__restrict__ hz, const uint32_t *__restrict__ hl,
                                const uint32_t *__restrict__ hcnt, const uint8_t *__restrict__ hscan,
                                const uint8_t *__restrict__ zg, int margin, uint8_t *__restrict__ bad) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint32_t slope = st->slope;
  if (g == 0 && st->target > 0) {
    const int64_t base = st->est_tile + st->hdr;
    const int64_t f0 = st->f[0] + base;
    const int64_t fm = slope > 0 ? st->f[1] + base : INT64_MAX;
    if (!(f0 < st->target && fm > st->target)) st->fallback = 1;
  }
  if (g >= G) return;
  const RcCb c = cb[g];
  const RcSlot s = slots[c.j];
  const int K = s.K[c.i];
  const int Z = zg[g];
  if (K == 0 || Z >= 3 * K - 2) return;
  const uint32_t ts[2] = {slope > 0 ? slope - 1 : 0u, slope};
  const int U = s.U[c.i];
  const int Zs = Z < U ? Z : U;
  const double *ms = s.mse + (uint64_t)c.i * kMaxPasses;
  const uint32_t *L = s.L + (uint64_t)c.i * kMaxPasses;
  for (int k = 0; k < 2; ++k) {
    uint32_t len;
    const int h = cut_at(s, c.i, Z, hz, hl, hoff[g], hcnt[g], hscan[g], ts[k], len);
    bool ok = h + margin <= U && h < Zs;
    if (ok) {
      double D = s.bfut[c.i];
      for (int z = h; z < Zs; ++z) D += ms[z];
      for (int z = Zs; z < Z; ++z) D += ms[z] > 0.0 ? ms[z] : 0.0;
      const double dL = (double)L[Zs - 1] - (h > 0 ? (double)L[h - 1] : 0.0);
      ok = D < slope_floor(ts[k]) * dL;
    }
    if (!ok) {
      bad[g] = 1;
      return;
    }
  }
}

// the layer's cuts at the final slope (create_quality_layer with est 0,
// tier2.c:960-1060): cp, L at the cut, the state the next layer starts from;
// beta_est as the layer's header estimate (est 1) leaves it
__global__ void rc_commit_kernel(const RcState *__restrict__ st, const RcCb *__restrict__ cb, uint64_t G,
                                 const RcSlot *__restrict__ slots, const uint64_t *__restrict__ hoff,
                                 const uint32_t *__restrict__ hz, const uint32_t *__restrict__ hl,
                                 const uint32_t *__restrict__ hcnt, const uint8_t *__restrict__ hscan,
                                 const uint8_t *__restrict__ zg, int16_t *__restrict__ prev,
                                 uint32_t *__restrict__ plen, uint16_t *__restrict__ beta, int nl,
                                 int16_t *__restrict__ cpo, uint32_t *__restrict__ lcut) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= G) return;
  const uint32_t q = st->layer;
  const RcCb c = cb[g];
  const RcSlot s = slots[c.j];
  uint32_t zl;
  const int z = cut_at(s, c.i, zg[g], hz, hl, hoff[g], hcnt[g], hscan[g], st->slope, zl);
  const int pv = prev[g];
  const int cur = z < pv ? pv : z;
  const uint32_t lc = cur > 0 ? s.L[(uint64_t)c.i * kMaxPasses + cur - 1] : 0u;
  if ((int)q + 1 < nl && cur > pv) beta[g] = len_code(beta[g], cur - pv, lc - (pv > 0 ? plen[g] : 0u)).beta;
  cpo[g * nl + q] = (int16_t)cur;
  lcut[g * nl + q] = lc;
  prev[g] = (int16_t)cur;
  plen[g] = lc;
}

// end of layer: the tile estimate (tier2.c:1242-1247); est1 = whether the
// layer's estimate (mode 2 probe) was made (every layer but the last)
__global__ void rc_layer_end_kernel(RcState *st, int est1) {
  st->est_ql = est1 ? st->f[0] : 0;
  st->est_tile += st->est_ql + st->hdr;
  st->m = 0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// packet-order tables of one tile (fixed for the codec's geometry)
struct RcGeo {
  uint64_t G = 0, NP = 0, N = 0;
  int depth = 1;
  std::vector<RcCb> cb;
  std::vector<uint32_t> pk_first;
  RcCb *d_cb = nullptr;
  uint32_t *d_pk_first = nullptr;
  int32_t *d_parent = nullptr;
};

struct RcCtx {
  const void *geo_of = nullptr;   // the Job the tables were built for
  std::vector<RcGeo> geo;         // per tile
  RcState *d_st = nullptr;
  RcState *h_st = nullptr;        // page-locked copy of the state
  // per-call buffers, grown on demand
  struct Buf {
    void *p = nullptr;
    size_t cap = 0;
  };
  Buf slots, hoff, hz, hl, hcnt, hscan, kms, zg, prev, plen, beta, bad, nodeval, mcw, mcn, cpo, lcut, pbase,
      piece_first, cur, zl, nb, off, pbody, pany, bsum, bits, pf_ex, pf_cn, pkt;
  uint8_t *h_io = nullptr;        // page-locked staging of uploads and results
  size_t cap_io = 0;
  ~RcCtx() {
    for (auto &g : geo) {
      if (g.d_cb) (void)hipFree(g.d_cb);
      if (g.d_pk_first) (void)hipFree(g.d_pk_first);
      if (g.d_parent) (void)hipFree(g.d_parent);
    }
    Buf *bs[] = {&slots, &hoff, &hz, &hl, &hcnt, &hscan, &kms, &zg, &prev, &plen, &beta, &bad, &nodeval, &mcw,
                 &mcn, &cpo, &lcut, &pbase, &piece_first, &cur, &zl, &nb, &off, &pbody, &pany, &bsum, &bits, &pf_ex, &pf_cn, &pkt};
    for (Buf *b : bs)
      if (b->p) (void)hipFree(b->p);
    if (d_st) (void)hipFree(d_st);
    if (h_st) (void)hipHostFree(h_st);
    if (h_io) (void)hipHostFree(h_io);
  }
};

int grow(RcCtx::Buf &b, size_t bytes) {
  bytes = std::max<size_t>(bytes, 256);
  if (b.p && b.cap >= bytes) return 0;
  if (b.p) HIP_RC(hipFree(b.p));
  b.p = nullptr;
  b.cap = 0;
  HIP_RC(hipMalloc(&b.p, bytes + bytes / 4));
  b.cap = bytes + bytes / 4;
  return 0;
}

template <class T>
T *ptr(RcCtx::Buf &b) {
  return static_cast<T *>(b.p);
}

// packet order of tile ti: parameters, resolutions, precincts; within a
// packet every subband's precinct codeblocks in turn. Each precinct-subband
// MSB tree gets its nodes at a running offset (leaves first).
int build_geo(bwc_codec *c, Job *job, uint64_t ti, RcGeo &g, hipStream_t s) {
  Tile &tl = job->plan.tiles[ti];
  g.cb.clear();
  g.pk_first.clear();
  std::vector<int32_t> parent;
  int depth = 1;
  for (int j = 0; j < c->info.nPar; ++j)
    for (auto &res : tl.par[j].res)
      for (uint64_t p = 0; p < res.nprec; ++p) {
        const uint32_t P = (uint32_t)g.pk_first.size();
        g.pk_first.push_back((uint32_t)g.cb.size());
        for (auto &sb : res.sb) {
          const Precinct &pr = sb.prec[p];
          const uint64_t base = parent.size();
          for (size_t n = 0; n < pr.msbs.parent.size(); ++n)
            parent.push_back(pr.msbs.parent[n] < 0 ? -1 : (int32_t)(base + (uint64_t)pr.msbs.parent[n]));
          for (uint64_t k = 0; k < pr.ncbs; ++k) {
            g.cb.push_back(RcCb{(uint32_t)j, (uint32_t)(pr.first_cb + k), P, (uint32_t)(base + k)});
            int d = 1;
            for (int64_t n = (int64_t)k; pr.msbs.parent[n] >= 0; n = pr.msbs.parent[n]) ++d;
            depth = std::max(depth, d);
          }
        }
      }
  g.pk_first.push_back((uint32_t)g.cb.size());
  g.G = g.cb.size();
  g.NP = g.pk_first.size() - 1;
  g.N = parent.size();
  g.depth = depth;
  if (g.G >= (1ull << 32) || g.N >= (1ull << 31)) {
    std::snprintf(rc_err, sizeof rc_err, "tile too large for device rate control");
    return 1;
  }
  HIP_RC(hipMalloc((void **)&g.d_cb, std::max<size_t>(1, g.G) * sizeof(RcCb)));
  HIP_RC(hipMalloc((void **)&g.d_pk_first, (g.NP + 1) * sizeof(uint32_t)));
  HIP_RC(hipMalloc((void **)&g.d_parent, std::max<size_t>(1, g.N) * sizeof(int32_t)));
  // copies on the call's own (non-blocking) stream: a copy on the legacy
  // stream would depend on every blocking stream, which another thread's
  // graph capture forbids
  if (g.G) HIP_RC(hipMemcpyAsync(g.d_cb, g.cb.data(), g.G * sizeof(RcCb), hipMemcpyHostToDevice, s));
  HIP_RC(hipMemcpyAsync(g.d_pk_first, g.pk_first.data(), (g.NP + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  if (g.N) HIP_RC(hipMemcpyAsync(g.d_parent, parent.data(), g.N * sizeof(int32_t), hipMemcpyHostToDevice, s));
  HIP_RC(hipStreamSynchronize(s));
  return 0;
}

inline uint32_t blocks(uint64_t n, uint32_t bs) { return (uint32_t)std::max<uint64_t>(1, (n + bs - 1) / bs); }

}  // namespace

void rc_release(Workspace *ws) {
  delete static_cast<RcCtx *>(ws->rc);
  ws->rc = nullptr;
}

int rc_trace(Workspace *ws, uint32_t *layer, uint32_t *slope, int64_t *est, int cap) {
  RcCtx *x = static_cast<RcCtx *>(ws->rc);
  if (!x || !x->h_st) return 0;
  const int n = (int)std::min<uint32_t>(x->h_st->ntrace, kTraceMax);
  for (int k = 0; k < n && k < cap; ++k) {
    layer[k] = x->h_st->trace[k].layer;
    slope[k] = x->h_st->trace[k].slope;
    est[k] = x->h_st->trace[k].est;
  }
  return n;
}

// Byte counts of n bit strings under the header writer's 0xFF stuffing, by
// the kernels rate control counts header estimates with (each string a
// one-codeblock packet of a single probe): string i is nbits[i] bits MSB
// first from word woff[i] of words (big-endian 32-bit words, host).
int rc_stuffed_bytes(const uint32_t *words, const uint64_t *woff, const uint64_t *nbits, int n, uint64_t *bytes) {
  if (n <= 0) return 0;
  std::vector<uint64_t> pbase(n + 1), hw;
  std::vector<uint32_t> pfirst(n + 1, 0), pk(n + 1), nb(n), off(n, 0), any(n, 1);
  uint64_t tot = 0;
  for (int i = 0; i < n; ++i) {
    if (nbits[i] >= (1ull << 32) || woff[i + 1] < woff[i] + (nbits[i] + 31) / 32) {
      std::snprintf(rc_err, sizeof rc_err, "bit string %d: bad length", i);
      return 1;
    }
    const uint64_t cw = woff[i + 1] - woff[i] + 2;
    pbase[i] = tot * 32;
    tot += cw;
    pfirst[i + 1] = pfirst[i] + (uint32_t)((cw * 32 + kPieceSpan - 1) / kPieceSpan);
    pk[i] = (uint32_t)i;
    nb[i] = (uint32_t)nbits[i];
  }
  pk[n] = (uint32_t)n;
  pbase[n] = tot * 32;
  hw.assign(tot + 2, 0);
  std::vector<uint32_t> wbuf((tot + 2) * 2, 0);
  for (int i = 0; i < n; ++i)
    std::memcpy(&wbuf[pbase[i] / 32], words + woff[i], (woff[i + 1] - woff[i]) * 4);
  const uint32_t npiece = pfirst[n];
  RcState *st = nullptr;
  uint32_t *d_w = nullptr, *d_pf = nullptr, *d_pk = nullptr, *d_nb = nullptr, *d_off = nullptr, *d_any = nullptr,
           *d_cn = nullptr;
  uint64_t *d_pb = nullptr;
  uint8_t *d_ex = nullptr;
  int64_t *d_body = nullptr, *d_pkt = nullptr;
  int rc = 1;
  do {
    RcState h{};
    h.m = 1;
    if (hipMalloc((void **)&st, sizeof(RcState)) != hipSuccess ||
        hipMalloc((void **)&d_w, wbuf.size() * 4) != hipSuccess ||
        hipMalloc((void **)&d_pf, (n + 1) * 4) != hipSuccess || hipMalloc((void **)&d_pk, (n + 1) * 4) != hipSuccess ||
        hipMalloc((void **)&d_nb, n * 4) != hipSuccess || hipMalloc((void **)&d_off, n * 4) != hipSuccess ||
        hipMalloc((void **)&d_any, n * 4) != hipSuccess || hipMalloc((void **)&d_pb, (n + 1) * 8) != hipSuccess ||
        hipMalloc((void **)&d_ex, (uint64_t)npiece * 16) != hipSuccess ||
        hipMalloc((void **)&d_cn, (uint64_t)npiece * 64) != hipSuccess ||
        hipMalloc((void **)&d_body, n * 8) != hipSuccess || hipMalloc((void **)&d_pkt, n * 8) != hipSuccess)
      break;
    std::vector<int64_t> zero(n, 0);
    if (hipMemcpy(st, &h, sizeof h, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_w, wbuf.data(), wbuf.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_pf, pfirst.data(), (n + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_pk, pk.data(), (n + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_nb, nb.data(), n * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_off, off.data(), n * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_any, any.data(), n * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_pb, pbase.data(), (n + 1) * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_body, zero.data(), n * 8, hipMemcpyHostToDevice) != hipSuccess)
      break;
    ProbeBufs pb{nullptr, nullptr, d_nb, d_off, d_body, d_any};
    const uint64_t capT = pbase[n] + 64;
    hipLaunchKernelGGL(rc_piece_kernel, dim3(npiece, 1), dim3(256), 0, 0, st, d_w, capT, d_pb, d_pf, (uint32_t)n,
                       npiece, d_pk, (uint64_t)n, pb, d_ex, d_cn);
    hipLaunchKernelGGL(rc_packet_kernel, dim3((uint32_t)n, 1), dim3(64), 0, 0, st, d_pf, npiece, d_pk, (uint64_t)n,
                       (uint32_t)n, pb, d_ex, d_cn, d_pkt);
    std::vector<int64_t> out(n);
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(out.data(), d_pkt, n * 8, hipMemcpyDeviceToHost) != hipSuccess)
      break;
    for (int i = 0; i < n; ++i) bytes[i] = (uint64_t)out[i];
    rc = 0;
  } while (0);
  if (rc) std::snprintf(rc_err, sizeof rc_err, "stuffed byte count: %s", hipGetErrorString(hipGetLastError()));
  void *ps[] = {st, d_w, d_pf, d_pk, d_nb, d_off, d_any, d_pb, d_ex, d_cn, d_body, d_pkt};
  for (void *q : ps)
    if (q) (void)hipFree(q);
  return rc;
}

int rc_tile(Workspace *ws, bwc_codec *c, Job *job, uint64_t ti, SkipCheck *chk) {
  if (!ws->rc) ws->rc = new RcCtx();
  RcCtx &x = *static_cast<RcCtx *>(ws->rc);
  if (x.geo_of != job || x.geo.size() != job->plan.tiles.size()) {
    for (auto &g : x.geo) {
      if (g.d_cb) HIP_RC(hipFree(g.d_cb));
      if (g.d_pk_first) HIP_RC(hipFree(g.d_pk_first));
      if (g.d_parent) HIP_RC(hipFree(g.d_parent));
    }
    x.geo.assign(job->plan.tiles.size(), RcGeo());
    x.geo_of = job;
  }
  RcGeo &g = x.geo[ti];
  if (!g.d_pk_first && build_geo(c, job, ti, g, ws->stream)) return 1;
  if (!x.d_st) {
    HIP_RC(hipMalloc((void **)&x.d_st, sizeof(RcState)));
    HIP_RC(hipHostMalloc((void **)&x.h_st, sizeof(RcState)));
  }
  Tile &tl = job->plan.tiles[ti];
  const int np = c->info.nPar, nl = c->control.nLayers;
  const uint64_t G = g.G, NP = g.NP;
  hipStream_t s = ws->stream;
  if ((int)ws->rec_slots.size() < np) {
    std::snprintf(rc_err, sizeof rc_err, "records of %d parameters expected on the device", np);
    return 1;
  }
  // per-call host tables: hull offsets (Z + 1 entries per codeblock), the
  // MSB-cache stride, per-packet header bit capacities and their pieces
  std::vector<uint64_t> hoff(G + 1, 0);
  uint32_t tmax = 1;
  std::vector<uint64_t> pcap(NP, 1);
  for (uint64_t k = 0; k < G; ++k) {
    const RcCb &cb = g.cb[k];
    const CblkRecords &R = job->rec[ti * np + cb.j];
    const int K = R.K[cb.i];
    hoff[k + 1] = hoff[k] + (K ? (uint64_t)R.Z[cb.i] + 1 : 1);
    const uint32_t T = (K ? R.Kmsbs[cb.i] : 0u) + 1u;
    tmax = std::max(tmax, T);
    pcap[cb.P] += (uint64_t)nl + kCbSlackBits + (uint64_t)T * (T + 1) / 2 + (uint64_t)T * g.depth;
  }
  const uint32_t stride = (uint32_t)((tmax * (tmax + 1) / 2 + (uint64_t)tmax * g.depth + 63) / 32);
  std::vector<uint64_t> pbase(NP + 1, 0);
  std::vector<uint32_t> pfirst(NP + 1, 0);
  for (uint64_t P = 0; P < NP; ++P) {
    const uint64_t cw = (pcap[P] + 31) / 32 + 2;   // whole words, and room for a window past the end
    pbase[P + 1] = pbase[P] + cw * 32;
    pfirst[P + 1] = pfirst[P] + (uint32_t)((cw * 32 + kPieceSpan - 1) / kPieceSpan);
  }
  const uint64_t capT = pbase[NP] + 64;
  const uint32_t npiece = pfirst[NP];
  const int depth = G * 31 <= (64ull << 20) / 8 ? 5 : 4;   // probes per round: 31, or 15 for large tiles
  const uint32_t mmax = (1u << depth) - 1;
  const int rounds = (17 + depth - 1) / depth;
  // buffers
  const uint64_t MG = (uint64_t)mmax * std::max<uint64_t>(G, 1);
  if (grow(x.slots, np * sizeof(RcSlot)) || grow(x.hoff, (G + 1) * 8) || grow(x.hz, hoff[G] * 4) ||
      grow(x.hl, hoff[G] * 4) || grow(x.hcnt, G * 4) || grow(x.hscan, G) || grow(x.kms, G) || grow(x.zg, G) ||
      grow(x.prev, G * 2) || grow(x.plen, G * 4) || grow(x.beta, G * 2) || grow(x.bad, G) ||
      grow(x.nodeval, g.N * 4) || grow(x.mcw, G * (uint64_t)stride * 4) || grow(x.mcn, G * 4) ||
      grow(x.cpo, G * nl * 2) || grow(x.lcut, G * nl * 4) || grow(x.pbase, (NP + 1) * 8) ||
      grow(x.piece_first, (NP + 1) * 4) || grow(x.cur, MG * 2) || grow(x.zl, MG * 4) || grow(x.nb, MG * 4) ||
      grow(x.off, MG * 4) || grow(x.pbody, (uint64_t)mmax * NP * 8) || grow(x.pany, (uint64_t)mmax * NP * 4) || grow(x.bsum, (MG / 1024 + 2) * 4) ||
      grow(x.bits, mmax * capT / 8 + 64) || grow(x.pf_ex, (uint64_t)mmax * npiece * 16) ||
      grow(x.pf_cn, (uint64_t)mmax * npiece * 64))
    return 1;
  // uploads through page-locked staging
  const size_t io_up = np * sizeof(RcSlot) + (G + 1) * 8 + (NP + 1) * 8 + (NP + 1) * 4;
  const size_t io_down = G * nl * 6 + G + 64;
  if (io_up + io_down > x.cap_io) {
    if (x.h_io) HIP_RC(hipHostFree(x.h_io));
    x.h_io = nullptr;
    x.cap_io = 0;
    HIP_RC(hipHostMalloc((void **)&x.h_io, io_up + io_down + (1 << 16)));
    x.cap_io = io_up + io_down + (1 << 16);
  }
  HIP_RC(hipStreamSynchronize(s));   // the staging may still feed the last call's copies
  uint8_t *up = x.h_io;
  RcSlot *hs = reinterpret_cast<RcSlot *>(up);
  for (int j = 0; j < np; ++j) {
    const Workspace::RecSlot &r = ws->rec_slots[j];
    hs[j] = RcSlot{r.K, r.Kmsbs, r.Z, r.U, r.L, r.S, r.mse, r.bfut};
  }
  size_t o = np * sizeof(RcSlot);
  std::memcpy(up + o, hoff.data(), (G + 1) * 8);
  const size_t o_hoff = o;
  o += (G + 1) * 8;
  std::memcpy(up + o, pbase.data(), (NP + 1) * 8);
  const size_t o_pbase = o;
  o += (NP + 1) * 8;
  std::memcpy(up + o, pfirst.data(), (NP + 1) * 4);
  const size_t o_pfirst = o;
  HIP_RC(hipMemcpyAsync(x.slots.p, up, np * sizeof(RcSlot), hipMemcpyHostToDevice, s));
  HIP_RC(hipMemcpyAsync(x.hoff.p, up + o_hoff, (G + 1) * 8, hipMemcpyHostToDevice, s));
  HIP_RC(hipMemcpyAsync(x.pbase.p, up + o_pbase, (NP + 1) * 8, hipMemcpyHostToDevice, s));
  HIP_RC(hipMemcpyAsync(x.piece_first.p, up + o_pfirst, (NP + 1) * 4, hipMemcpyHostToDevice, s));

  const uint64_t hdr = (uint64_t)std::ceil((float)((uint32_t)c->control.headerSize + tl.header_size) / nl);
  const uint64_t tsz = tl.box.volume() * (uint64_t)np;
  RcState *st = x.d_st;
  const RcSlot *dslots = ptr<RcSlot>(x.slots);
  const uint64_t *dhoff = ptr<uint64_t>(x.hoff);
  uint32_t *dhz = ptr<uint32_t>(x.hz), *dhl = ptr<uint32_t>(x.hl), *dhcnt = ptr<uint32_t>(x.hcnt);
  uint8_t *dhscan = ptr<uint8_t>(x.hscan), *dkms = ptr<uint8_t>(x.kms), *dzg = ptr<uint8_t>(x.zg);
  int16_t *dprev = ptr<int16_t>(x.prev);
  uint32_t *dplen = ptr<uint32_t>(x.plen), *dnodeval = ptr<uint32_t>(x.nodeval), *dmcw = ptr<uint32_t>(x.mcw),
           *dmcn = ptr<uint32_t>(x.mcn), *dlcut = ptr<uint32_t>(x.lcut), *dbits = ptr<uint32_t>(x.bits),
           *dpf_cn = ptr<uint32_t>(x.pf_cn), *dbsum = ptr<uint32_t>(x.bsum);
  uint16_t *dbeta = ptr<uint16_t>(x.beta);
  uint8_t *dbad = ptr<uint8_t>(x.bad), *dpf_ex = ptr<uint8_t>(x.pf_ex);
  int16_t *dcpo = ptr<int16_t>(x.cpo);
  const uint64_t *dpbase = ptr<uint64_t>(x.pbase);
  const uint32_t *dpfirst = ptr<uint32_t>(x.piece_first);
  ProbeBufs pb{ptr<int16_t>(x.cur), ptr<uint32_t>(x.zl), ptr<uint32_t>(x.nb), ptr<uint32_t>(x.off),
               ptr<int64_t>(x.pbody), ptr<uint32_t>(x.pany)};

  // the launch sequence, up to the copies of the results into page-locked
  // memory: recorded as a graph and replayed while everything it depends on
  // stays the same (pipeline graphs on)
  uint8_t *down = x.h_io + io_up;
  auto launch_all = [&]() -> int {
    hipLaunchKernelGGL(rc_state_init_kernel, dim3(1), dim3(1), 0, s, st, (int64_t)hdr);
    if (g.N) hipLaunchKernelGGL(rc_fill_u32_kernel, dim3(blocks(g.N, 256)), dim3(256), 0, s, dnodeval, g.N, 0xFFFFu);
    if (G) {
      hipLaunchKernelGGL(rc_init_kernel, dim3(blocks(G, 256)), dim3(256), 0, s, g.d_cb, G, dslots, dhoff, dhz, dhl,
                         dhcnt, dhscan, dkms, dzg, dprev, dplen, dbeta, dbad, st);
      hipLaunchKernelGGL(rc_msbs_cache_kernel, dim3(blocks(G, 256)), dim3(256), 0, s, g.d_cb, G, dkms, g.d_parent,
                         dnodeval, stride, dmcw, dmcn);
    }
    HIP_RC(hipGetLastError());

    // one evaluation of st->m (at most mcap) thresholds of the current layer
    auto probe_round = [&](uint32_t mcap) -> int {
      if (!G) return 0;
      const uint64_t n = (uint64_t)mcap * G;
      hipLaunchKernelGGL(rc_round_zero_kernel, dim3(blocks(mcap * capT / 32, 256 * 8)), dim3(256), 0, s, dbits,
                         mcap * capT / 32 + 2, pb.body, pb.any, (uint64_t)mcap * NP);
      hipLaunchKernelGGL(rc_cut_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, st, mcap, g.d_cb, G, (uint32_t)NP,
                         dslots, dhoff, dhz, dhl, dhcnt, dhscan, dzg, dprev, dplen, dbeta, dmcn, g.d_pk_first, pb);
      const uint64_t nblk = (n + 1023) / 1024;
      hipLaunchKernelGGL(rc_scan_local_kernel, dim3((uint32_t)nblk), dim3(256), 0, s, pb.nb, pb.off, n, dbsum);
      hipLaunchKernelGGL(rc_scan_top_kernel, dim3(1), dim3(1024), 0, s, dbsum, nblk);
      hipLaunchKernelGGL(rc_scan_add_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, pb.off, n, dbsum);
      hipLaunchKernelGGL(rc_write_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, st, mcap, g.d_cb, G, dprev, dplen,
                         dbeta, dmcn, dmcw, stride, g.d_pk_first, dpbase, capT, pb, dbits);
      hipLaunchKernelGGL(rc_piece_kernel, dim3(npiece, mcap), dim3(256), 0, s, st, dbits, capT, dpbase, dpfirst,
                         (uint32_t)NP, npiece, g.d_pk_first, G, pb, dpf_ex, dpf_cn);
      hipLaunchKernelGGL(rc_packet_kernel, dim3((uint32_t)NP, mcap), dim3(64), 0, s, st, dpfirst, npiece,
                         g.d_pk_first, G, (uint32_t)NP, pb, dpf_ex, dpf_cn, (int64_t *)nullptr);
      HIP_RC(hipGetLastError());
      return 0;
    };

    for (int l = 0; l < nl; ++l) {
      const int64_t target = (int64_t)std::ceil(c->control.bitrate[l] * tsz / 8.0f);
      hipLaunchKernelGGL(rc_layer_begin_kernel, dim3(1), dim3(1), 0, s, st, (uint32_t)l, target, depth);
      for (int r = 0; r < rounds; ++r) {
        if (probe_round(mmax)) return 1;
        hipLaunchKernelGGL(rc_advance_kernel, dim3(1), dim3(1), 0, s, st, depth);
      }
      hipLaunchKernelGGL(rc_bisect_end_kernel, dim3(1), dim3(1), 0, s, st);
      if (chk) {
        hipLaunchKernelGGL(rc_setup_probes_kernel, dim3(1), dim3(1), 0, s, st, 1);
        if (probe_round(2)) return 1;
        hipLaunchKernelGGL(rc_check_kernel, dim3(blocks(G, 256)), dim3(256), 0, s, st, g.d_cb, G, dslots, dhoff, dhz,
                           dhl, dhcnt, dhscan, dzg, chk->margin, dbad);
      }
      if (l == 0 && nl > 1 && G) {
        // the layer's final estimate sees the MSB trees set (create_quality_layer
        // sets them before create_packet, tier2.c:1040-1064)
        hipLaunchKernelGGL(rc_msbs_set_kernel, dim3(blocks(G, 256)), dim3(256), 0, s, g.d_cb, G, dkms, g.d_parent,
                           dnodeval);
        hipLaunchKernelGGL(rc_msbs_cache_kernel, dim3(blocks(G, 256)), dim3(256), 0, s, g.d_cb, G, dkms, g.d_parent,
                           dnodeval, stride, dmcw, dmcn);
      }
      const int est1 = l + 1 < nl;
      if (est1) {
        hipLaunchKernelGGL(rc_setup_probes_kernel, dim3(1), dim3(1), 0, s, st, 2);
        if (probe_round(1)) return 1;
      }
      if (G)
        hipLaunchKernelGGL(rc_commit_kernel, dim3(blocks(G, 256)), dim3(256), 0, s, st, g.d_cb, G, dslots, dhoff, dhz,
                           dhl, dhcnt, dhscan, dzg, dprev, dplen, dbeta, nl, dcpo, dlcut);
      hipLaunchKernelGGL(rc_layer_end_kernel, dim3(1), dim3(1), 0, s, st, est1);
      HIP_RC(hipGetLastError());
    }
    // results: cuts and L at the cuts per layer, the check's flags, the state
    HIP_RC(hipMemcpyAsync(down, dcpo, G * nl * 2, hipMemcpyDeviceToHost, s));
    HIP_RC(hipMemcpyAsync(down + G * nl * 2, dlcut, G * nl * 4, hipMemcpyDeviceToHost, s));
    if (chk) HIP_RC(hipMemcpyAsync(down + G * nl * 6, dbad, G, hipMemcpyDeviceToHost, s));
    HIP_RC(hipMemcpyAsync(x.h_st, st, sizeof(RcState), hipMemcpyDeviceToHost, s));
    return 0;
  };
  if (graphs_on()) {
    uint64_t key = 0xC6A4A7935BD1E995ull;
    auto mix = [&key](uint64_t v) { key = (key ^ v) * 0x100000001B3ull + 0x9E3779B97F4A7C15ull; };
    RcCtx::Buf *bs[] = {&x.slots, &x.hoff,  &x.hz,   &x.hl,          &x.hcnt, &x.hscan, &x.kms, &x.zg,
                        &x.prev,  &x.plen,  &x.beta, &x.bad,         &x.nodeval, &x.mcw, &x.mcn, &x.cpo,
                        &x.lcut,  &x.pbase, &x.piece_first, &x.cur,  &x.zl,   &x.nb,    &x.off, &x.pbody,
                        &x.pany,  &x.bsum,  &x.bits, &x.pf_ex,       &x.pf_cn};
    for (RcCtx::Buf *b : bs) mix((uint64_t)(uintptr_t)b->p);
    const void *more[] = {x.d_st, x.h_st, down, g.d_cb, g.d_pk_first, g.d_parent};
    for (const void *p : more) mix((uint64_t)(uintptr_t)p);
    mix(G), mix(NP), mix(g.N), mix(capT), mix(npiece), mix(stride), mix((uint64_t)depth << 8 | (uint64_t)rounds);
    mix((uint64_t)nl | (uint64_t)(chk ? chk->margin + 1 : 0) << 16), mix(hdr);
    for (int l = 0; l < nl; ++l) mix((uint64_t)(int64_t)std::ceil(c->control.bitrate[l] * tsz / 8.0f));
    if (graph_run(ws, key, launch_all)) return 1;
  } else if (launch_all()) {
    return 1;
  }
  HIP_RC(hipStreamSynchronize(s));
  if (x.h_st->err) {
    std::snprintf(rc_err, sizeof rc_err, "device rate control: bisection state error %u", x.h_st->err);
    return 1;
  }
  const int16_t *hcp = reinterpret_cast<const int16_t *>(down);
  const uint32_t *hlc = reinterpret_cast<const uint32_t *>(down + G * nl * 2);
  for (uint64_t k = 0; k < G; ++k) {
    const RcCb &cb = g.cb[k];
    CblkRecords &R = job->rec[ti * np + cb.j];
    int16_t *cp = &R.cp[(uint64_t)cb.i * (nl + 1) + 1];
    uint32_t *L = &R.L[(uint64_t)cb.i * kMaxPasses];
    for (int l = 0; l < nl; ++l) {
      cp[l] = hcp[k * nl + l];
      if (cp[l] > 0) L[cp[l] - 1] = hlc[k * nl + l];
    }
  }
  if (chk) {
    chk->fallback = x.h_st->fallback != 0;
    chk->flagged.assign(np, {});
    const uint8_t *hb = down + G * nl * 6;
    for (uint64_t k = 0; k < G; ++k)
      if (hb[k]) chk->flagged[g.cb[k].j].push_back(g.cb[k].i);
    for (auto &f : chk->flagged) std::sort(f.begin(), f.end());
  }
  return 0;
}

}  // namespace gpu
}  // namespace bwc
