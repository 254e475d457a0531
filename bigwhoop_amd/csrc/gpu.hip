// gfx950 kernels of the compression path and their host-side launchers.
//
//   fill_minmax_kernel    tile copy to f64 + min/max      (libbwc.c:622-812)
//   normalize_kernel      (x-a)*b, exponent += Qm          (libbwc.c:993-1078)
//   dwt97_axis_kernel     one lifting pass over many lines (dwt.c:133-404, :645-721, :1216-1724)
//   quantize_kernel       |c|/delta -> stripe columns, K   (tier1.c:526-729)
//   t1_encode_kernel      passes + MQ + lengths + slopes   (tier1.c:962-1511, :2100-2588; mq.c)
//   gather_kernel         pack included codeblock prefixes
//
// Arithmetic that decides quantisation indices is kept in IEEE binary64 with
// the reference's operand order; the library is compiled with
// -ffp-contract=off so no FMA is formed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <condition_variable>
#include <mutex>

#include "gpu.hpp"
#include "tuning.hpp"

namespace bwc {
namespace gpu {

static thread_local char g_err[256];
const char *last_error() { return g_err; }

// BWC_AMD_RC_TRACE=1: host-side phase times of tier-1 on stderr (measurement)
static bool trace_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("BWC_AMD_RC_TRACE");
    return e && e[0] == '1';
  }();
  return on;
}
static double trace_mark(const char *name, double last) {
  const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
  if (trace_enabled() && name) std::fprintf(stderr, "t1: %-22s %8.3f ms\n", name, t - last);
  return t;
}
static int process_setup();

#define HIP_OK(x)                                                                                      \
  do {                                                                                                 \
    hipError_t e_ = (x);                                                                               \
    if (e_ != hipSuccess) {                                                                            \
      std::snprintf(g_err, sizeof g_err, "%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
      std::fprintf(stderr, "bwc: HIP error: %s\n", g_err);                                             \
      return 1;                                                                                        \
    }                                                                                                  \
  } while (0)

#define CHK(x)       \
  do {               \
    if ((x)) return 1; \
  } while (0)

Workspace::~Workspace() {
  if (device < 0) return;
  (void)hipSetDevice(device);
  void *ptrs[] = {d_input, d_work,  d_work2, d_red,  d_norm, d_cbbox, d_cbsb, d_cbmeta, d_qoff,   d_doff, d_q,
                  d_state, d_data,  d_snap,  d_flags,
                  d_packed, d_gather, d_bp, d_bpoff, d_sym, d_symoff, d_pcnt, d_pmse, d_list, d_zlim,
                  d_feat,  d_tstat, d_skipw, d_bodies, dec.box, dec.sb, dec.meta, dec.K, dec.Z, dec.len,
                  dec.off, dec.data, dec.soff, dec.state, dec.out, d_poff, d_rpack, d_t1info};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (auto &r : rec_slots) {
    void *rp[] = {r.K, r.Kmsbs, r.Z, r.U, r.L, r.S, r.mse, r.bfut};
    for (void *p : rp)
      if (p) (void)hipFree(p);
  }
  rc_release(this);
  for (auto &b : bp_slots) {
    if (b.bp) (void)hipFree(b.bp);
    if (b.hist) (void)hipFree(b.hist);
  }
  for (void *p : d_datap)
    if (p) (void)hipFree(p);
  if (h_pinned) (void)hipHostFree(h_pinned);
  if (ev_up) {
    (void)hipEventSynchronize(ev_up);
    (void)hipEventDestroy(ev_up);
  }
  if (h_up) (void)hipHostFree(h_up);
  if (h_down) (void)hipHostFree(h_down);
  if (h_kz) (void)hipHostFree(h_kz);
  for (int b = 0; b < kStageBufs; ++b) {
    if (ev_stage[b]) {
      (void)hipEventSynchronize(ev_stage[b]);
      (void)hipEventDestroy(ev_stage[b]);
    }
    if (h_stage[b]) (void)hipHostFree(h_stage[b]);
  }
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (copy_stream) (void)hipStreamSynchronize(copy_stream);
  for (auto &e : ev_copy) (void)hipEventDestroy(e);
  if (ev_fill) (void)hipEventDestroy(ev_fill);
  if (copy_stream) (void)hipStreamDestroy(copy_stream);
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto &g : dwt_graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  for (auto &g : graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (stream) (void)hipStreamDestroy(stream);
}

static int ensure(void **p, size_t *cap, size_t bytes) {
  if (*p && *cap >= bytes) return 0;
  if (*p) HIP_OK(hipFree(*p));
  *p = nullptr;
  size_t nb = std::max<size_t>(bytes, 256);
  HIP_OK(hipMalloc(p, nb));
  *cap = nb;
  return 0;
}

// Point the workspace's record aliases (d_K, d_L, ...) at parameter slot
// `slot`, with room for ncb codeblocks (grown without keeping contents).
static int bind_slot(Workspace *ws, int slot, uint64_t ncb) {
  if (slot < 0 || slot > 255) {
    std::snprintf(g_err, sizeof g_err, "parameter slot out of range");
    return 1;
  }
  if (slot >= (int)ws->rec_slots.size()) ws->rec_slots.resize(slot + 1);
  Workspace::RecSlot &r = ws->rec_slots[slot];
  if (r.cap < ncb + 1) {
    void **rp[] = {(void **)&r.K, (void **)&r.Kmsbs, (void **)&r.Z,   (void **)&r.U,
                   (void **)&r.L, (void **)&r.S,     (void **)&r.mse, (void **)&r.bfut};
    for (void **p : rp)
      if (*p) {
        HIP_OK(hipFree(*p));
        *p = nullptr;
      }
    r.cap = 0;
    const uint64_t n = ncb + 1;
    HIP_OK(hipMalloc((void **)&r.K, n));
    HIP_OK(hipMalloc((void **)&r.Kmsbs, n));
    HIP_OK(hipMalloc((void **)&r.Z, n));
    HIP_OK(hipMalloc((void **)&r.U, n));
    HIP_OK(hipMalloc((void **)&r.L, n * kMaxPasses * sizeof(uint32_t)));
    HIP_OK(hipMalloc((void **)&r.S, n * (kMaxPasses + 1) * sizeof(uint16_t)));
    HIP_OK(hipMalloc((void **)&r.mse, n * kMaxPasses * sizeof(double)));
    HIP_OK(hipMalloc((void **)&r.bfut, n * sizeof(double)));
    r.cap = n;
  }
  ws->d_K = r.K, ws->d_Kmsbs = r.Kmsbs, ws->d_Z = r.Z, ws->d_U = r.U;
  ws->d_L = r.L, ws->d_S = r.S, ws->d_mse = r.mse, ws->d_bfut = r.bfut;
  return 0;
}

// Upload staging (page-locked): up_begin makes room for `bytes` of uploads
// once the previous call's uploads have left the buffer; up_copy stages one
// host array and queues its copy to dst on the workspace stream.
// upload bytes per codeblock of one tier-1 call: box 32, steps 24, meta 4,
// five offset tables 40, list 8, pass limit 1 (rounded up)
constexpr size_t kUpBytesPerCb = 128;
static int up_begin(Workspace *ws, size_t bytes) {
  if (!ws->ev_up) HIP_OK(hipEventCreateWithFlags(&ws->ev_up, hipEventDisableTiming));
  else HIP_OK(hipEventSynchronize(ws->ev_up));
  ws->used_up = 0;
  if (bytes <= ws->cap_up) return 0;
  if (ws->h_up) HIP_OK(hipHostFree(ws->h_up));
  ws->h_up = nullptr;
  ws->cap_up = 0;
  HIP_OK(hipHostMalloc((void **)&ws->h_up, bytes));
  ws->cap_up = bytes;
  return 0;
}
static int up_copy(Workspace *ws, void *dst, const void *src, size_t n) {
  const size_t at = (ws->used_up + 255) & ~(size_t)255;
  if (at + n > ws->cap_up) {
    std::snprintf(g_err, sizeof g_err, "upload staging too small (%zu of %zu bytes)", at + n, ws->cap_up);
    return 1;
  }
  std::memcpy(ws->h_up + at, src, n);
  ws->used_up = at + n;
  HIP_OK(hipMemcpyAsync(dst, ws->h_up + at, n, hipMemcpyHostToDevice, ws->stream));
  HIP_OK(hipEventRecord(ws->ev_up, ws->stream));
  return 0;
}

// ===========================================================================
// fill + min/max, normalize
// ===========================================================================
// Copy to f64 + per-block min/max over the tile-local indices [i0, i1) of one
// (tile, param). The source is the caller's field (strides sX, sY, sZ, tile
// offsets x0, y0, z0) or a tile-shaped staging copy of it (strides w, h, d,
// zero offsets). The time index is not offset by the tile's start, as in the
// reference's fill_buffer (libbwc.c:727-729: nZ * t). Partials go to
// red[2 * blockIdx.x], red[2 * blockIdx.x + 1].
__global__ void fill_minmax_kernel(const void *__restrict__ in, int prec, uint64_t sX, uint64_t sY, uint64_t sZ,
                                   uint64_t poff, uint64_t x0, uint64_t y0, uint64_t z0, uint64_t w, uint64_t h,
                                   uint64_t d, uint64_t i0, uint64_t i1, double *__restrict__ out,
                                   double *__restrict__ red) {
  double mx = -1.7976931348623157e+308, mn = 1.7976931348623157e+308;
  for (uint64_t i = i0 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < i1;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = i % w, r = i / w;
    const uint64_t y = r % h, r2 = r / h;
    const uint64_t z = r2 % d, t = r2 / d;
    const uint64_t src = poff + x0 + x + sX * (y0 + y + sY * (z0 + z + sZ * t));
    const double v = prec == 8 ? ((const double *)in)[src] : (double)((const float *)in)[src];
    out[i] = v;
    mx = (mx < v) ? v : mx;
    mn = (mn > v) ? v : mn;
  }
  __shared__ double smx[256], smn[256];
  smx[threadIdx.x] = mx;
  smn[threadIdx.x] = mn;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smx[threadIdx.x] = (smx[threadIdx.x] < smx[threadIdx.x + s]) ? smx[threadIdx.x + s] : smx[threadIdx.x];
      smn[threadIdx.x] = (smn[threadIdx.x] > smn[threadIdx.x + s]) ? smn[threadIdx.x + s] : smn[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    red[2 * blockIdx.x] = smn[0];
    red[2 * blockIdx.x + 1] = smx[0];
  }
}

// reduce the partials; alpha = (max+min)/2, beta = 2/(max-min) (libbwc.c:810-811)
__global__ void minmax_final_kernel(const double *red, int nblk, double *norm) {
  __shared__ double smx[256], smn[256];
  double mx = -1.7976931348623157e+308, mn = 1.7976931348623157e+308;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
    mn = (mn > red[2 * i]) ? red[2 * i] : mn;
    mx = (mx < red[2 * i + 1]) ? red[2 * i + 1] : mx;
  }
  smx[threadIdx.x] = mx;
  smn[threadIdx.x] = mn;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smx[threadIdx.x] = (smx[threadIdx.x] < smx[threadIdx.x + s]) ? smx[threadIdx.x + s] : smx[threadIdx.x];
      smn[threadIdx.x] = (smn[threadIdx.x] > smn[threadIdx.x + s]) ? smn[threadIdx.x + s] : smn[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    norm[0] = smn[0];
    norm[1] = smx[0];
    norm[2] = (smx[0] + smn[0]) / 2.0f;
    norm[3] = 2.0f / (smx[0] - smn[0]);
  }
}

// (x - alpha) * beta, then Qm added to the exponent field of every value
// other than +-0 (libbwc.c:993-1078; denormals are shifted the same way)
__device__ __forceinline__ double norm_value(double x, double alpha, double beta, uint64_t shift) {
  const double v = (x - alpha) * beta;
  uint64_t raw = __double_as_longlong(v);
  // A NaN here has the bits x86-64 SSE arithmetic gives it: a NaN sample
  // propagates quieted, an invalid operation (0 * inf for a constant field,
  // whose beta is 2/0) gives the default NaN with its sign bit set. The
  // exponent shift below turns those bits into a finite value that the
  // transform then carries.
  if (v != v) {
    const uint64_t xr = (uint64_t)__double_as_longlong(x);
    raw = (x != x) ? (xr | 0x0008000000000000ull) : 0xFFF8000000000000ull;
  }
  raw += ((raw & 0x7FF0000000000000ull) == 0 && (raw & 0x000FFFFFFFFFFFFFull) == 0) ? 0 : shift;
  return __longlong_as_double((long long)raw);
}

__global__ void normalize_kernel(double *__restrict__ data, uint64_t n, const double *__restrict__ norm,
                                 uint64_t shift) {
  const double alpha = norm[2], beta = norm[3];
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    data[i] = norm_value(data[i], alpha, beta, shift);
}

// partial min/max slots of one fill (all chunks together)
constexpr int kRedSlots = 8192;
// bytes per staged chunk of a host field: copies of later chunks overlap the
// min/max pass over earlier ones
constexpr uint64_t kFillChunkBytes = 64ull << 20;

int fill_minmax(Workspace *ws, const bwc_codec *c, const Tile &tl, int param, const void *input, bool on_dev,
                double *work) {
  const Param &par = tl.par[param];
  const uint64_t n = par.box.volume();
  const uint64_t nX = c->info.nX, nY = c->info.nY, nZ = c->info.nZ, nTS = c->info.nTS;
  const int prec = c->info.data_prec == bwc_precision_single ? 4 : 8;
  const uint64_t poff = nX * nY * nZ * nTS * (uint64_t)param;
  const uint64_t w = par.box.len(0), h = par.box.len(1), d = par.box.len(2), dt = par.box.len(3);
  const uint64_t x0 = tl.box.lo[0], y0 = tl.box.lo[1], z0 = tl.box.lo[2];
  if (!ws->d_red) HIP_OK(hipMalloc((void **)&ws->d_red, 2 * kRedSlots * sizeof(double)));
  if (!ws->d_norm) HIP_OK(hipMalloc((void **)&ws->d_norm, 4 * sizeof(double)));
  const int threads = 256;
  int slots = 0;
  // page-locked host memory the GPU can address: the kernel reads it over
  // PCIe itself (no staging copy), like a device field
  const void *src = input;
  bool direct = on_dev;
  if (!on_dev && n) {
    const char *zc = std::getenv("BWC_AMD_FILL_ZEROCOPY");
    if (!(zc && zc[0] == '0')) {
      hipPointerAttribute_t attr;
      if (hipPointerGetAttributes(&attr, input) == hipSuccess && attr.type == hipMemoryTypeHost && attr.devicePointer) {
        src = attr.devicePointer;
        direct = true;
      } else {
        (void)hipGetLastError();
      }
    }
  }
  if (direct || !n) {
    const int blocks = (int)std::max<uint64_t>(1, std::min<uint64_t>((n + threads - 1) / threads, 4096));
    hipLaunchKernelGGL(fill_minmax_kernel, dim3(blocks), dim3(threads), 0, ws->stream, src, prec, nX, nY, nZ, poff,
                       x0, y0, z0, w, h, d, (uint64_t)0, n, work, ws->d_red);
    HIP_OK(hipGetLastError());
    slots = blocks;
  } else {
    // Host field: only this tile's box crosses PCIe, into a tile-shaped
    // staging buffer, in chunks on a copy stream; the min/max pass over a
    // chunk starts as soon as it has landed. Contiguous source runs: whole
    // (z, t) volumes when the tile spans full X and Y rows and planes, whole
    // planes when it spans full rows, else rows (one 3-D copy per time step).
    CHK(ensure(&ws->d_input, &ws->input_cap, n * prec));
    if (!ws->copy_stream) HIP_OK(hipStreamCreateWithFlags(&ws->copy_stream, hipStreamNonBlocking));
    if (!ws->ev_fill) HIP_OK(hipEventCreateWithFlags(&ws->ev_fill, hipEventDisableTiming));
    // the staging buffer may still be read by the previous fill
    HIP_OK(hipEventRecord(ws->ev_fill, ws->stream));
    HIP_OK(hipStreamWaitEvent(ws->copy_stream, ws->ev_fill, 0));
    const char *src = static_cast<const char *>(input);
    char *dst = static_cast<char *>(ws->d_input);
    // (tile-local start, length) of each chunk, in values
    std::vector<std::pair<uint64_t, uint64_t>> chunks;
    const uint64_t cvals = std::max<uint64_t>(1, kFillChunkBytes / prec);
    if (w == nX && h == nY) {
      // per time step one contiguous range of w*h*d values (all of them
      // contiguous when d == nZ)
      const uint64_t run = w * h * d;
      const bool whole = d == nZ;
      // Pageable memory: the host's threads copy each chunk into one of the
      // page-locked bounce buffers (in parallel, one chunk after another),
      // and the DMA of a chunk runs while the next is copied. BWC_AMD_STAGE=0
      // hands the pageable source to hipMemcpyAsync instead.
      const char *se = std::getenv("BWC_AMD_STAGE");
      const bool bounce = !(se && se[0] == '0');
      if (bounce && ws->cap_stage < kFillChunkBytes) {
        for (int b = 0; b < Workspace::kStageBufs; ++b) {
          if (ws->h_stage[b]) HIP_OK(hipHostFree(ws->h_stage[b]));
          ws->h_stage[b] = nullptr;
        }
        ws->cap_stage = 0;
        for (int b = 0; b < Workspace::kStageBufs; ++b) {
          HIP_OK(hipHostMalloc((void **)&ws->h_stage[b], kFillChunkBytes));
          if (!ws->ev_stage[b]) HIP_OK(hipEventCreateWithFlags(&ws->ev_stage[b], hipEventDisableTiming));
        }
        ws->cap_stage = kFillChunkBytes;
      }
      const int hthreads = rc_thread_count();
      int nb = 0;
      for (uint64_t t = 0; t < (whole ? 1 : dt); ++t) {
        const uint64_t len = whole ? n : run;
        const uint64_t sbase = poff + nX * nY * (z0 + nZ * t);
        for (uint64_t o = 0; o < len; o += cvals) {
          const uint64_t m = std::min(cvals, len - o);
          if (bounce) {
            const int b = nb++ % Workspace::kStageBufs;
            // the buffer's last DMA has finished reading it
            HIP_OK(hipEventSynchronize(ws->ev_stage[b]));
            const char *from = src + (sbase + o) * prec;
            uint8_t *to = ws->h_stage[b];
            const uint64_t bytes = m * prec;
            const size_t parts = (size_t)std::max<uint64_t>(1, std::min<uint64_t>(16, bytes >> 22));
            const std::function<void(size_t)> cp = [&](size_t k) {
              const uint64_t a = (bytes * k / parts) & ~63ull, e = k + 1 == parts ? bytes : (bytes * (k + 1) / parts) & ~63ull;
              std::memcpy(to + a, from + a, e - a);
            };
            if (parts > 1 && hthreads > 1) host_pool_run(std::min(hthreads, 16) - 1, parts, cp);
            else cp(0);
            HIP_OK(hipMemcpyAsync(dst + (t * run + o) * prec, to, bytes, hipMemcpyHostToDevice, ws->copy_stream));
            HIP_OK(hipEventRecord(ws->ev_stage[b], ws->copy_stream));
          } else {
            HIP_OK(hipMemcpyAsync(dst + (t * run + o) * prec, src + (sbase + o) * prec, m * prec,
                                  hipMemcpyHostToDevice, ws->copy_stream));
          }
          chunks.push_back({t * run + o, m});
          if (ws->ev_copy.size() < chunks.size()) {
            hipEvent_t e;
            HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            ws->ev_copy.push_back(e);
          }
          HIP_OK(hipEventRecord(ws->ev_copy[chunks.size() - 1], ws->copy_stream));
        }
      }
    } else {
      // one strided 3-D copy per time step: rows of w values at pitch nX,
      // h rows per plane, d planes
      for (uint64_t t = 0; t < dt; ++t) {
        hipMemcpy3DParms cp = {};
        cp.srcPtr = make_hipPitchedPtr(const_cast<char *>(src) + (poff + nX * nY * nZ * t) * prec, nX * prec, nX, nY);
        cp.srcPos = make_hipPos(x0 * prec, y0, z0);
        cp.dstPtr = make_hipPitchedPtr(dst + w * h * d * t * prec, w * prec, w, h);
        cp.dstPos = make_hipPos(0, 0, 0);
        cp.extent = make_hipExtent(w * prec, h, d);
        cp.kind = hipMemcpyHostToDevice;
        HIP_OK(hipMemcpy3DAsync(&cp, ws->copy_stream));
        chunks.push_back({w * h * d * t, w * h * d});
        if (ws->ev_copy.size() < chunks.size()) {
          hipEvent_t e;
          HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
          ws->ev_copy.push_back(e);
        }
        HIP_OK(hipEventRecord(ws->ev_copy[chunks.size() - 1], ws->copy_stream));
      }
    }
    const int per = (int)std::max<size_t>(1, std::min<size_t>(512, kRedSlots / std::max<size_t>(1, chunks.size())));
    if ((size_t)per * chunks.size() > (size_t)kRedSlots) {
      std::snprintf(g_err, sizeof g_err, "field staged in too many chunks (%zu)", chunks.size());
      return 1;
    }
    for (size_t k = 0; k < chunks.size(); ++k) {
      HIP_OK(hipStreamWaitEvent(ws->stream, ws->ev_copy[k], 0));
      const int blocks = (int)std::max<uint64_t>(1, std::min<uint64_t>((chunks[k].second + threads - 1) / threads, per));
      hipLaunchKernelGGL(fill_minmax_kernel, dim3(blocks), dim3(threads), 0, ws->stream, ws->d_input, prec, w, h, d,
                         (uint64_t)0, (uint64_t)0, (uint64_t)0, (uint64_t)0, w, h, d, chunks[k].first,
                         chunks[k].first + chunks[k].second, work, ws->d_red + 2 * slots);
      HIP_OK(hipGetLastError());
      slots += blocks;
    }
  }
  hipLaunchKernelGGL(minmax_final_kernel, dim3(1), dim3(256), 0, ws->stream, ws->d_red, slots, ws->d_norm);
  HIP_OK(hipGetLastError());
  return 0;
}

static int normalize_on(hipStream_t stream, double *work, uint64_t n, const double *norm, uint8_t Qm) {
  const int threads = 256;
  const int blocks = (int)std::max<uint64_t>(1, std::min<uint64_t>((n + threads - 1) / threads, 8192));
  hipLaunchKernelGGL(normalize_kernel, dim3(blocks), dim3(threads), 0, stream, work, n, norm, (uint64_t)Qm << 52);
  HIP_OK(hipGetLastError());
  return 0;
}

int normalize(Workspace *ws, double *work, uint64_t n, uint8_t Qm) {
  return normalize_on(ws->stream, work, n, ws->d_norm, Qm);
}

// ===========================================================================
// DWT: one axis pass of one level over all lines of the current LL box.
// A workgroup stages G adjacent lines in LDS (line-major, E-sample whole-point
// symmetric extension on both ends), runs the lifting steps over exactly the
// reference's index ranges, and writes lowpass | highpass back in place.
// Lines are read either along the line (contiguous X pass) or across lines
// (strided Y/Z/T passes, adjacent lines adjacent in memory) for coalescing.
// ===========================================================================
#define DWT_E 8   // extension slots per side (reference uses 6 for 9/7, 4 for 5/3)

__device__ __forceinline__ int64_t mirror_index(int64_t j, int64_t len) {
  // whole-point symmetric reflection about 0 and len-1, repeated for short
  // lines (dwt.c:286-342)
  if (len == 1) return 0;
  while (j < 0 || j > len - 1) j = (j < 0) ? -j : 2 * (len - 1) - j;
  return j;
}

template <int KERN, bool FWD>
__global__ void dwt_axis_kernel(double *__restrict__ data, int64_t len, int par0, uint64_t sa, uint64_t n0,
                                uint64_t s0, uint64_t n1, uint64_t s1, uint64_t s2, int G, int contiguous) {
  extern __shared__ double lds[];
  const int64_t pitch = len + 2 * DWT_E + 1;
  const uint64_t g0 = (uint64_t)blockIdx.x * G;
  const uint64_t i1 = blockIdx.y % n1, i2 = blockIdx.y / n1;
  const int ng = (int)std::min<uint64_t>((uint64_t)G, n0 - g0);
  double *base = data + g0 * s0 + i1 * s1 + i2 * s2;
  const int64_t tot = (int64_t)ng * len;
  const int nt = blockDim.x;
  const int64_t half = par0 ? (len >> 1) : ((len + 1) >> 1);

  // ---- load ----
  for (int64_t e = threadIdx.x; e < tot; e += nt) {
    int64_t g, p;
    if (contiguous) {
      g = e / len;
      p = e % len;
    } else {
      g = e % ng;
      p = e / ng;
    }
    double v;
    if (FWD) {
      v = base[g * s0 + p * sa];
    } else {
      // inverse: interleave lowpass [0,half) and highpass [half,len) as
      // fill_inverse_buffer does (dwt.c:169-272): position p of the line
      // holds low[p/2] when p is even and high[(p-1)/2] when p is odd, for
      // either origin parity (an odd origin only shifts the lifting by one)
      const int64_t src = !(p & 1) ? (p >> 1) : (half + ((p - 1) >> 1));
      v = base[g * s0 + src * sa];
    }
    lds[g * pitch + DWT_E + p] = v;
  }
  __syncthreads();
  // ---- symmetric extension ----
  for (int64_t e = threadIdx.x; e < (int64_t)ng * 2 * DWT_E; e += nt) {
    const int64_t g = e / (2 * DWT_E), k = e % (2 * DWT_E);
    const int64_t p = (k < DWT_E) ? (k - DWT_E) : (len + (k - DWT_E));
    lds[g * pitch + DWT_E + p] = lds[g * pitch + DWT_E + mirror_index(p, len)];
  }
  __syncthreads();
  // working index i addresses line position i + par0 (dwt.c:1393-1396)
  const int64_t w0 = DWT_E + par0;
  if (KERN == 0) {
    const double A = -1.586134342059924f, B = -0.052980118572961f, Gm = 0.882911075530934f,
                 D = 0.360523644801462f, KH = 1.230174104914001f, KL = 0.812893066115961f;
    if (FWD) {
      // four lifting steps with their extended ranges (dwt.c:371-403)
      const int64_t lo[4] = {-3, -2, -1, 0}, hi[4] = {len + 3, len + 2, len + 1, len};
      for (int st = 0; st < 4; ++st) {
        const int64_t cnt = (hi[st] - lo[st] + 1) / 2;
        for (int64_t e = threadIdx.x; e < (int64_t)ng * cnt; e += nt) {
          const int64_t g = e / cnt, m = e % cnt;
          double *w = lds + g * pitch + w0;
          const int64_t i = lo[st] + 2 * m;
          if (st == 0) w[i] += A * (w[i - 1] + w[i + 1]);
          else if (st == 1) w[i] += B * (w[i - 1] + w[i + 1]);
          else if (st == 2) { w[i] += Gm * (w[i - 1] + w[i + 1]); w[i] = KH * w[i]; }
          else { w[i] += D * (w[i - 1] + w[i + 1]); w[i] = KL * w[i]; }
        }
        __syncthreads();
      }
    } else {
      // inverse steps (dwt.c:428-465)
      const int64_t lo[4] = {-4, -5, -4, -3}, hi[4] = {len + 4, len + 3, len + 2, len + 1};
      for (int st = 0; st < 4; ++st) {
        const int64_t cnt = (hi[st] - lo[st] + 1) / 2;
        for (int64_t e = threadIdx.x; e < (int64_t)ng * cnt; e += nt) {
          const int64_t g = e / cnt, m = e % cnt;
          double *w = lds + g * pitch + w0;
          const int64_t i = lo[st] + 2 * m;
          if (st == 0) { w[i] = KH * w[i]; w[i] -= D * (w[i - 1] + w[i + 1]); }
          else if (st == 1) { w[i] = KL * w[i]; w[i] -= Gm * (w[i - 1] + w[i + 1]); }
          else if (st == 2) w[i] -= B * (w[i - 1] + w[i + 1]);
          else w[i] -= A * (w[i - 1] + w[i + 1]);
        }
        __syncthreads();
      }
    }
  } else {
    if (FWD) {
      const int64_t lo[2] = {-1, 0}, hi[2] = {len + 1, len};
      for (int st = 0; st < 2; ++st) {
        const int64_t cnt = (hi[st] - lo[st] + 1) / 2;
        for (int64_t e = threadIdx.x; e < (int64_t)ng * cnt; e += nt) {
          const int64_t g = e / cnt, m = e % cnt;
          double *w = lds + g * pitch + w0;
          const int64_t i = lo[st] + 2 * m;
          if (st == 0) w[i] -= (w[i - 1] + w[i + 1]) / 2;
          else w[i] += (w[i - 1] + w[i + 1] + 2) / 4;
        }
        __syncthreads();
      }
    } else {
      const int64_t lo[2] = {-2, -1}, hi[2] = {len + 1, len};
      for (int st = 0; st < 2; ++st) {
        const int64_t cnt = (hi[st] - lo[st] + 1) / 2;
        for (int64_t e = threadIdx.x; e < (int64_t)ng * cnt; e += nt) {
          const int64_t g = e / cnt, m = e % cnt;
          double *w = lds + g * pitch + w0;
          const int64_t i = lo[st] + 2 * m;
          if (st == 0) w[i] -= (w[i - 1] + w[i + 1] + 2) / 4;
          else w[i] += (w[i - 1] + w[i + 1]) / 2;
        }
        __syncthreads();
      }
    }
  }
  // ---- store ----
  for (int64_t e = threadIdx.x; e < tot; e += nt) {
    int64_t g, o;
    if (contiguous) {
      g = e / len;
      o = e % len;
    } else {
      g = e % ng;
      o = e / ng;
    }
    double v;
    if (FWD) {
      // lowpass o < half from position 2o + par0, highpass from 2(o-half) + 1 - par0
      // (buffer_flush_forward, dwt.c:645-721)
      const int64_t p = (o < half) ? (2 * o + par0) : (2 * (o - half) + 1 - par0);
      v = lds[g * pitch + DWT_E + p];
    } else {
      v = lds[g * pitch + DWT_E + o];
    }
    base[g * s0 + o * sa] = v;
  }
}

// ---------------------------------------------------------------------------
// Forward CDF 9/7, fast kernels. Both compute, for every line, exactly the
// reference's lifting steps over its extended ranges (dwt.c:354-404) on the
// whole-point symmetric extension (dwt.c:286-342), so each coefficient is the
// same sequence of IEEE binary64 operations as in the reference; only the
// order in which independent positions are visited differs.
//
// Working index i addresses line position p = i + par0. Step s is applied at
// i = lo_s, lo_s + 2, ... (cnt_s positions): lo = {-3, -2, -1, 0},
// cnt = {(len+7)/2, (len+5)/2, (len+3)/2, (len+1)/2}.
//   st0 (odd i):  a_i = x_i + A (x_{i-1} + x_{i+1})
//   st1 (even i): b_i = a_i + B (a_{i-1} + a_{i+1})
//   st2 (odd i):  c_i = KH (b_i + G (b_{i-1} + b_{i+1}))
//   st3 (even i): d_i = KL (c_i + D (c_{i-1} + c_{i+1}))
// Lowpass o < half takes i = 2o, highpass o >= half takes the odd
// i = 2(o - half) + 1 - 2 par0 (buffer_flush_forward, dwt.c:645-721).
// ---------------------------------------------------------------------------
struct Lift97 {
  static constexpr double A = -1.586134342059924f, B = -0.052980118572961f, G = 0.882911075530934f,
                          D = 0.360523644801462f, KH = 1.230174104914001f, KL = 0.812893066115961f;
};

__device__ __forceinline__ int mirror32(int p, int len) {
  if (len == 1) return 0;
  while (p < 0 || p > len - 1) p = (p < 0) ? -p : 2 * (len - 1) - p;
  return p;
}

// X axis (contiguous lines): one wavefront per line. The extended line is
// split into even and odd working positions (Ev, Od) in LDS, so every step
// reads and writes consecutive LDS words and the output is two contiguous
// runs (lowpass = Ev, highpass = Od). src and dst may be the same buffer.
// Ev[j] holds i = 2j - 8, Od[j] holds i = 2j - 7.
// With `norm` set (the first pass over a freshly filled tile) every sample is
// normalised as it is read, as normalize_kernel would have done in place.
__global__ void __launch_bounds__(256) dwt97_x_kernel(const double *src, double *dst, int len, int par0, int half,
                                                      uint32_t n0, uint64_t s0, uint32_t n1, uint64_t s1,
                                                      uint64_t s2, const double *norm, uint64_t shift, int nt) {
  extern __shared__ double lds[];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int hp = (len + 16) / 2 + 1;
  double *Ev = lds + (size_t)wv * 2 * hp, *Od = Ev + hp;
  const uint32_t i0 = blockIdx.x * 4 + wv;
  const uint32_t i1 = blockIdx.y % n1, i2 = blockIdx.y / n1;
  const bool act = i0 < n0;
  const uint64_t base = (uint64_t)i0 * s0 + (uint64_t)i1 * s1 + (uint64_t)i2 * s2;
  if (act) {
    const double *s = src + base;
    const double alpha = norm ? norm[2] : 0, beta = norm ? norm[3] : 0;
    // eight loads in flight per lane before the first LDS write
    for (int e0 = 0; e0 < len + 16; e0 += 64 * 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + 64 * u + lane, p = e - 8 + par0;
        const double *q = &s[((unsigned)p < (unsigned)len) ? p : mirror32(p, len)];
        v[u] = (e < len + 16) ? ((nt & 2) ? __builtin_nontemporal_load(q) : *q) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + 64 * u + lane;
        if (e >= len + 16) break;
        const double x = norm ? norm_value(v[u], alpha, beta, shift) : v[u];
        if (e & 1) Od[e >> 1] = x;
        else Ev[e >> 1] = x;
      }
    }
  }
  __syncthreads();
  if (act)
    for (int m = lane; m < (len + 7) / 2; m += 64) Od[2 + m] += Lift97::A * (Ev[2 + m] + Ev[3 + m]);
  __syncthreads();
  if (act)
    for (int m = lane; m < (len + 5) / 2; m += 64) Ev[3 + m] += Lift97::B * (Od[2 + m] + Od[3 + m]);
  __syncthreads();
  if (act)
    for (int m = lane; m < (len + 3) / 2; m += 64) {
      const double t = Od[3 + m] + Lift97::G * (Ev[3 + m] + Ev[4 + m]);
      Od[3 + m] = Lift97::KH * t;
    }
  __syncthreads();
  if (act)
    for (int m = lane; m < (len + 1) / 2; m += 64) {
      const double t = Ev[4 + m] + Lift97::D * (Od[3 + m] + Od[4 + m]);
      Ev[4 + m] = Lift97::KL * t;
    }
  __syncthreads();
  if (act) {
    double *d = dst + base;
    if (nt & 1) {
      for (int o = lane; o < len; o += 64)
        __builtin_nontemporal_store((o < half) ? Ev[o + 4] : Od[o - half + 4 - par0], &d[o]);
    } else {
      for (int o = lane; o < len; o += 64) d[o] = (o < half) ? Ev[o + 4] : Od[o - half + 4 - par0];
    }
  }
}

// Strided axes: one thread per line segment, consecutive lanes on
// consecutive lines of the fastest other axis, so every load and store of a
// wavefront is one contiguous row segment. The four steps run as a pipeline
// along the line: after reading x_{2k} and x_{2k+1} the thread finishes
// a_{2k-1}, b_{2k-2}, c_{2k-3} and d_{2k-4} (outputs at i = 2k-4, 2k-3).
// Output iterations k = 1 .. kend-1 are split into nseg segments of ksz; a
// segment starts its pipeline 4 iterations early: the carried values start
// as zeros and, as each step reaches back at most two iterations, are exact
// again after three, so only exact values are stored. Highpass outputs land
// ahead of the read position, so the pass is out of place (src != dst).
// loads and stores of the streaming DWT passes, with the nontemporal hint when asked
template <bool NT>
__device__ __forceinline__ double st_ld(const double *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT>
__device__ __forceinline__ void st_st(double *p, double v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <int U, int NT>
__global__ void __launch_bounds__(256) dwt97_stream_kernel(const double *__restrict__ src, double *__restrict__ dst,
                                                           int len, int par0, int half, uint64_t sa, uint32_t n0,
                                                           uint64_t s0, uint32_t n1, uint64_t s1, uint64_t s2,
                                                           uint32_t nlines, int nseg, int ksz) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nlines * (uint32_t)nseg) return;
  const uint32_t l = tid % nlines, seg = tid / nlines;
  const uint32_t i0 = l % n0, r = l / n0;
  const uint32_t i1 = r % n1, i2 = r / n1;
  const uint64_t base = (uint64_t)i0 * s0 + (uint64_t)i1 * s1 + (uint64_t)i2 * s2;
  const double *s = src + base;
  double *d = dst + base;
  const int last0 = -3 + 2 * ((len + 7) / 2 - 1), last1 = -2 + 2 * ((len + 5) / 2 - 1);
  const int last2 = -1 + 2 * ((len + 3) / 2 - 1), last3 = 2 * ((len + 1) / 2 - 1);
  const int ilo = -par0, ihi = len - par0;   // output working indices [ilo, ihi)
  const int kend = (ihi + 5) / 2;            // first k with 2k - 4 >= ihi
  const int kA = 1 + (int)seg * ksz, kB = min(kend, kA + ksz);
  if (kA >= kB) return;
  auto ld = [&](int i) -> double {
    const int p = i + par0;
    return st_ld<(NT & 2) != 0>(&s[(int64_t)(((unsigned)p < (unsigned)len) ? p : mirror32(p, len)) * (int64_t)sa]);
  };
  // carried: x_{2k-2}, x_{2k-1}, a_{2k-3}, b_{2k-4}, c_{2k-5}
  const int k0 = kA - 4;
  double xm2 = ld(2 * k0 - 2), xm1 = ld(2 * k0 - 1), am3 = 0, bm4 = 0, cm5 = 0;
  for (int k = k0; k < kB; k += U) {
    double xv[2 * U];
#pragma unroll
    for (int u = 0; u < 2 * U; ++u) xv[u] = ld(2 * k + u);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int kk = k + u;
      if (kk >= kB) break;
      const double x0 = xv[2 * u], x1 = xv[2 * u + 1];
      int i = 2 * kk - 1;
      const double a1 = (i >= -3 && i <= last0) ? xm1 + Lift97::A * (xm2 + x0) : xm1;
      i = 2 * kk - 2;
      const double b2 = (i >= -2 && i <= last1) ? xm2 + Lift97::B * (am3 + a1) : xm2;
      i = 2 * kk - 3;
      const double c3 = (i >= -1 && i <= last2) ? Lift97::KH * (am3 + Lift97::G * (bm4 + b2)) : am3;
      i = 2 * kk - 4;
      const double d4 = (i >= 0 && i <= last3) ? Lift97::KL * (bm4 + Lift97::D * (cm5 + c3)) : bm4;
      if (kk >= kA && i >= ilo && i < ihi) st_st<(NT & 1) != 0>(&d[(int64_t)(i >> 1) * (int64_t)sa], d4);
      i = 2 * kk - 3;
      if (kk >= kA && i >= ilo && i < ihi)
        st_st<(NT & 1) != 0>(&d[(int64_t)(half + ((i + 2 * par0 - 1) >> 1)) * (int64_t)sa], c3);
      xm2 = x0;
      xm1 = x1;
      am3 = a1;
      bm4 = b2;
      cm5 = c3;
    }
  }
}

// Inverse 9/7 along X (contiguous lines), one wavefront per line as in
// dwt97_x_kernel: the extended interleaved line (idwt97_stream_kernel's
// sample at each working index) split into even and odd working positions in
// LDS (Ev[j] holds i = 2j - 8, Od[j] holds i = 2j - 7), the four inverse
// steps over their ranges, and the line written back in natural order.
// src and dst may be the same buffer.
__global__ void __launch_bounds__(256) idwt97_x_kernel(const double *src, double *dst, int len, int par0, int half,
                                                       uint32_t n0, uint64_t s0, uint32_t n1, uint64_t s1,
                                                       uint64_t s2) {
  extern __shared__ double lds[];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int hp = (len + 16) / 2 + 1;
  double *Ev = lds + (size_t)wv * 2 * hp, *Od = Ev + hp;
  const uint32_t i0 = blockIdx.x * 4 + wv;
  const uint32_t i1 = blockIdx.y % n1, i2 = blockIdx.y / n1;
  const bool act = i0 < n0;
  const uint64_t base = (uint64_t)i0 * s0 + (uint64_t)i1 * s1 + (uint64_t)i2 * s2;
  if (act) {
    const double *s = src + base;
    for (int e0 = 0; e0 < len + 16; e0 += 64 * 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + 64 * u + lane;
        int p = e - 8 + par0;
        p = ((unsigned)p < (unsigned)len) ? p : mirror32(p, len);
        const int from = !(p & 1) ? (p >> 1) : (half + ((p - 1) >> 1));
        v[u] = (e < len + 16) ? s[from] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + 64 * u + lane;
        if (e >= len + 16) break;
        if (e & 1) Od[e >> 1] = v[u];
        else Ev[e >> 1] = v[u];
      }
    }
  }
  __syncthreads();
  if (act)
    for (int m = lane; m < (len + 9) / 2; m += 64) {
      double t = Lift97::KH * Ev[2 + m];
      t -= Lift97::D * (Od[1 + m] + Od[2 + m]);
      Ev[2 + m] = t;
    }
  __syncthreads();
  if (act)
    for (int m = lane; m < (len + 9) / 2; m += 64) {
      double t = Lift97::KL * Od[1 + m];
      t -= Lift97::G * (Ev[1 + m] + Ev[2 + m]);
      Od[1 + m] = t;
    }
  __syncthreads();
  if (act)
    for (int m = lane; m < (len + 7) / 2; m += 64) Ev[2 + m] -= Lift97::B * (Od[1 + m] + Od[2 + m]);
  __syncthreads();
  if (act)
    for (int m = lane; m < (len + 5) / 2; m += 64) Od[2 + m] -= Lift97::A * (Ev[2 + m] + Ev[3 + m]);
  __syncthreads();
  if (act) {
    double *d = dst + base;
    for (int p = lane; p < len; p += 64) {
      const int e = p - par0 + 8;
      d[p] = (e & 1) ? Od[e >> 1] : Ev[e >> 1];
    }
  }
}

// Inverse 9/7 along a strided axis, one thread per line segment as in
// dwt97_stream_kernel (consecutive lanes on consecutive lines). The line is
// the interleaving of its lowpass half [0, half) and highpass half [half,
// len) (fill_inverse_buffer, dwt.c:169-272), extended by whole-point
// symmetry; the reference's inverse steps (dwt.c:428-465) on working index i
// (position i + par0) are
//   st0 (even i in [-4, m0]): e = KH x_i - D (x_{i-1} + x_{i+1})
//   st1 (odd i in [-5, m1]):  o = KL x_i - G (e_{i-1} + e_{i+1})
//   st2 (even i in [-4, m2]): f = e_i - B (o_{i-1} + o_{i+1})
//   st3 (odd i in [-3, m3]):  g = o_i - A (f_{i-1} + f_{i+1})
// with positions outside a step's range left as they are. After reading
// x_{2k} and x_{2k+1} the thread finishes e_{2k}, o_{2k-1}, f_{2k-2} and
// g_{2k-3}, and stores f_{2k-2} and g_{2k-3} at their positions. Segments
// start four iterations early with zero carried values, which are exact
// again after three. Outputs land where lowpass and highpass inputs may
// still be unread, so the pass is out of place.
template <int U, int NT>
__global__ void __launch_bounds__(256) idwt97_stream_kernel(const double *__restrict__ src, double *__restrict__ dst,
                                                            int len, int par0, int half, uint64_t sa, uint32_t n0,
                                                            uint64_t s0, uint32_t n1, uint64_t s1, uint64_t s2,
                                                            uint32_t nlines, int nseg, int ksz) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nlines * (uint32_t)nseg) return;
  const uint32_t l = tid % nlines, seg = tid / nlines;
  const uint32_t i0 = l % n0, r = l / n0;
  const uint32_t i1 = r % n1, i2 = r / n1;
  const uint64_t base = (uint64_t)i0 * s0 + (uint64_t)i1 * s1 + (uint64_t)i2 * s2;
  const double *s = src + base;
  double *d = dst + base;
  const int m0 = -4 + 2 * ((len + 9) / 2 - 1), m1 = -5 + 2 * ((len + 9) / 2 - 1);
  const int m2 = -4 + 2 * ((len + 7) / 2 - 1), m3 = -3 + 2 * ((len + 5) / 2 - 1);
  const int ilo = -par0, ihi = len - par0;   // output working indices [ilo, ihi)
  const int kend = (ihi + 2) / 2 + 1;        // past the last k with an output
  const int kA = 1 + (int)seg * ksz, kB = min(kend, kA + ksz);
  if (kA >= kB) return;
  // interleaved sample at working index i (position mirrored into the line)
  auto ld = [&](int i) -> double {
    int p = i + par0;
    p = ((unsigned)p < (unsigned)len) ? p : mirror32(p, len);
    const int from = !(p & 1) ? (p >> 1) : (half + ((p - 1) >> 1));
    return st_ld<(NT & 2) != 0>(&s[(int64_t)from * (int64_t)sa]);
  };
  // carried: x_{2k-1}, e_{2k-2}, o_{2k-3}, f_{2k-4}
  const int k0 = kA - 4;
  double xm1 = ld(2 * k0 - 1), em2 = 0, om3 = 0, fm4 = 0;
  for (int k = k0; k < kB; k += U) {
    double xv[2 * U];
#pragma unroll
    for (int u = 0; u < 2 * U; ++u) xv[u] = ld(2 * k + u);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int kk = k + u;
      if (kk >= kB) break;
      const double x0 = xv[2 * u], x1 = xv[2 * u + 1];
      int i = 2 * kk;
      double e0 = x0;
      if (i >= -4 && i <= m0) {
        e0 = Lift97::KH * x0;
        e0 -= Lift97::D * (xm1 + x1);
      }
      i = 2 * kk - 1;
      double o1 = xm1;
      if (i >= -5 && i <= m1) {
        o1 = Lift97::KL * xm1;
        o1 -= Lift97::G * (em2 + e0);
      }
      i = 2 * kk - 2;
      const double f2 = (i >= -4 && i <= m2) ? em2 - Lift97::B * (om3 + o1) : em2;
      i = 2 * kk - 3;
      const double g3 = (i >= -3 && i <= m3) ? om3 - Lift97::A * (fm4 + f2) : om3;
      if (kk >= kA) {
        i = 2 * kk - 2;
        if (i >= ilo && i < ihi) st_st<(NT & 1) != 0>(&d[(int64_t)(i + par0) * (int64_t)sa], f2);
        i = 2 * kk - 3;
        if (i >= ilo && i < ihi) st_st<(NT & 1) != 0>(&d[(int64_t)(i + par0) * (int64_t)sa], g3);
      }
      xm1 = x1;
      em2 = e0;
      om3 = o1;
      fm4 = f2;
    }
  }
}

// X and Y passes of one level in one kernel, for boxes whose X lines fit
// LDS: the box is read once and written once for both axes. A workgroup of
// 256 threads takes one (Z, T) plane and one segment of its Y iterations.
// Per batch, the four wavefronts X-transform 2U rows of the plane (the Y
// pipeline's next working indices, mirrored into the plane) in LDS with
// dwt97_x_kernel's steps; then every thread advances the Y pipeline of its
// X-transformed positions o = t, t + 256, ... over those rows with
// dwt97_stream_kernel's steps and stores the outputs. The next batch's rows
// are loaded into registers while the current batch's X and Y steps run. Every
// coefficient is the same sequence of binary64 operations as in the two
// separate passes. Row slot r of LDS: Ev = lds + 2 hp r, Od = Ev + hp, as in
// dwt97_x_kernel; slot r is loaded by wavefront r % 4.
template <int U, int NT, int CMAX>
__global__ void __launch_bounds__(256) dwt97_xy_kernel(const double *__restrict__ src, double *__restrict__ dst,
                                                       int lenx, int par0x, int halfx, int leny, int par0y, int halfy,
                                                       uint64_t sy, uint32_t n2, uint64_t s2, uint64_t s3, int nseg,
                                                       int ksz, const double *norm, uint64_t shift) {
  extern __shared__ double lds[];
  constexpr int NV = (256 * CMAX + 16 + 63) / 64;   // samples of an extended row per lane
  constexpr int SPW = (2 * U + 3) / 4;              // row slots per wavefront
  const int hp = (lenx + 16) / 2 + 1;
  const uint32_t plane = blockIdx.x / (uint32_t)nseg, seg = blockIdx.x % (uint32_t)nseg;
  const uint64_t base = (uint64_t)(plane % n2) * s2 + (uint64_t)(plane / n2) * s3;
  const double *s = src + base;
  double *d = dst + base;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const double alpha = norm ? norm[2] : 0, beta = norm ? norm[3] : 0;
  const int last0 = -3 + 2 * ((leny + 7) / 2 - 1), last1 = -2 + 2 * ((leny + 5) / 2 - 1);
  const int last2 = -1 + 2 * ((leny + 3) / 2 - 1), last3 = 2 * ((leny + 1) / 2 - 1);
  const int ilo = -par0y, ihi = leny - par0y;
  const int kend = (ihi + 5) / 2;
  const int kA = 1 + (int)seg * ksz, kB = min(kend, kA + ksz);
  if (kA >= kB) return;   // the whole workgroup
  // rows with Y working indices i0 .. i0 + nrows - 1 into registers: sample
  // e = 64 u + lane of the extended row (X position e - 8 + par0x, mirrored)
  double pf[SPW][NV];
  auto fetch = [&](int i0, int nrows) {
#pragma unroll
    for (int j = 0; j < SPW; ++j) {
      const int r = wv + 4 * j;
      if (r >= nrows) break;
      const int p = i0 + r + par0y;
      const double *row = s + (uint64_t)(((unsigned)p < (unsigned)leny) ? p : mirror32(p, leny)) * sy;
#pragma unroll
      for (int u = 0; u < NV; ++u) {
        const int e = 64 * u + lane, q = e - 8 + par0x;
        pf[j][u] = (e < lenx + 16) ? st_ld<(NT & 2) != 0>(&row[((unsigned)q < (unsigned)lenx) ? q : mirror32(q, lenx)])
                                   : 0.0;
      }
    }
  };
  // the fetched rows into their slots, normalised when asked
  auto stage = [&](int nrows) {
#pragma unroll
    for (int j = 0; j < SPW; ++j) {
      const int r = wv + 4 * j;
      if (r >= nrows) break;
      double *Ev = lds + (size_t)2 * hp * r, *Od = Ev + hp;
#pragma unroll
      for (int u = 0; u < NV; ++u) {
        const int e = 64 * u + lane;
        if (e >= lenx + 16) break;
        const double x = norm ? norm_value(pf[j][u], alpha, beta, shift) : pf[j][u];
        if (e & 1) Od[e >> 1] = x;
        else Ev[e >> 1] = x;
      }
    }
  };
  // the staged rows X-transformed
  auto lift = [&](int nrows) {
    __syncthreads();
    for (int r = wv; r < nrows; r += 4) {
      double *Ev = lds + (size_t)2 * hp * r, *Od = Ev + hp;
      for (int m = lane; m < (lenx + 7) / 2; m += 64) Od[2 + m] += Lift97::A * (Ev[2 + m] + Ev[3 + m]);
    }
    __syncthreads();
    for (int r = wv; r < nrows; r += 4) {
      double *Ev = lds + (size_t)2 * hp * r, *Od = Ev + hp;
      for (int m = lane; m < (lenx + 5) / 2; m += 64) Ev[3 + m] += Lift97::B * (Od[2 + m] + Od[3 + m]);
    }
    __syncthreads();
    for (int r = wv; r < nrows; r += 4) {
      double *Ev = lds + (size_t)2 * hp * r, *Od = Ev + hp;
      for (int m = lane; m < (lenx + 3) / 2; m += 64) {
        const double t = Od[3 + m] + Lift97::G * (Ev[3 + m] + Ev[4 + m]);
        Od[3 + m] = Lift97::KH * t;
      }
    }
    __syncthreads();
    for (int r = wv; r < nrows; r += 4) {
      double *Ev = lds + (size_t)2 * hp * r, *Od = Ev + hp;
      for (int m = lane; m < (lenx + 1) / 2; m += 64) {
        const double t = Ev[4 + m] + Lift97::D * (Od[3 + m] + Od[4 + m]);
        Ev[4 + m] = Lift97::KL * t;
      }
    }
    __syncthreads();
  };
  // X output o of slot r
  auto xval = [&](int r, int o) -> double {
    const double *Ev = lds + (size_t)2 * hp * r, *Od = Ev + hp;
    return (o < halfx) ? Ev[o + 4] : Od[o - halfx + 4 - par0x];
  };
  // carried per column: x_{2k-2}, x_{2k-1}, a_{2k-3}, b_{2k-4}, c_{2k-5}
  double xm2[CMAX], xm1[CMAX], am3[CMAX], bm4[CMAX], cm5[CMAX];
  const int k0 = kA - 4;
  fetch(2 * k0 - 2, 2);
  stage(2);
  lift(2);
#pragma unroll
  for (int c = 0; c < CMAX; ++c) {
    const int o = threadIdx.x + 256 * c;
    xm2[c] = o < lenx ? xval(0, o) : 0.0;
    xm1[c] = o < lenx ? xval(1, o) : 0.0;
    am3[c] = bm4[c] = cm5[c] = 0.0;
  }
  fetch(2 * k0, 2 * U);
  __syncthreads();
  for (int k = k0; k < kB; k += U) {
    stage(2 * U);
    if (k + U < kB) fetch(2 * (k + U), 2 * U);
    lift(2 * U);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int kk = k + u;
      if (kk >= kB) break;
#pragma unroll
      for (int c = 0; c < CMAX; ++c) {
        const int o = threadIdx.x + 256 * c;
        if (o >= lenx) continue;
        const double x0 = xval(2 * u, o), x1 = xval(2 * u + 1, o);
        int i = 2 * kk - 1;
        const double a1 = (i >= -3 && i <= last0) ? xm1[c] + Lift97::A * (xm2[c] + x0) : xm1[c];
        i = 2 * kk - 2;
        const double b2 = (i >= -2 && i <= last1) ? xm2[c] + Lift97::B * (am3[c] + a1) : xm2[c];
        i = 2 * kk - 3;
        const double c3 = (i >= -1 && i <= last2) ? Lift97::KH * (am3[c] + Lift97::G * (bm4[c] + b2)) : am3[c];
        i = 2 * kk - 4;
        const double d4 = (i >= 0 && i <= last3) ? Lift97::KL * (bm4[c] + Lift97::D * (cm5[c] + c3)) : bm4[c];
        if (kk >= kA && i >= ilo && i < ihi) st_st<(NT & 1) != 0>(&d[(uint64_t)(i >> 1) * sy + o], d4);
        i = 2 * kk - 3;
        if (kk >= kA && i >= ilo && i < ihi)
          st_st<(NT & 1) != 0>(&d[(uint64_t)(halfy + ((i + 2 * par0y - 1) >> 1)) * sy + o], c3);
        xm2[c] = x0;
        xm1[c] = x1;
        am3[c] = a1;
        bm4[c] = b2;
        cm5[c] = c3;
      }
    }
    __syncthreads();
  }
}

// copy the box [0,n0) x [0,n1) x [0,n2) x [0,n3) (strides 1, s1, s2, s3)
__global__ void box_copy_kernel(const double *__restrict__ src, double *__restrict__ dst, uint32_t n0, uint32_t n1,
                                uint32_t n2, uint64_t s1, uint64_t s2, uint64_t s3, uint64_t total) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t r = e;
    const uint64_t x = r % n0;
    r /= n0;
    const uint64_t y = r % n1;
    r /= n1;
    const uint64_t z = r % n2, t = r / n2;
    const uint64_t o = x + y * s1 + z * s2 + t * s3;
    dst[o] = src[o];
  }
}

// Where one DWT call runs: its stream, the second buffer of the out-of-place
// passes (grown from ws when the call may allocate, else of fixed capacity),
// the {min, max, alpha, beta} the fused normalisation reads, and the number
// of kernels it launched.
struct DwtCtx {
  hipStream_t stream = nullptr;
  double *work2 = nullptr;
  size_t work2_cap = 0;     // bytes
  const double *norm = nullptr;
  Workspace *grow = nullptr;
  int launches = 0;
};

static int dwt_passes(DwtCtx &x, const bwc_codec *c, const Param &par, double *work, bool fwd, int norm_qm) {
  const bwc_gl_ctrl &ctl = c->control;
  // normalisation still to apply (forward only): fused into the first X pass
  // when that pass runs on the fast kernel, else a separate in-place pass
  bool norm_pending = fwd && norm_qm >= 0;
  const uint64_t W = par.box.len(0), H = par.box.len(1), D = par.box.len(2);
  const uint64_t inc[4] = {1, W, W * H, W * H * D};
  const uint8_t dec[4] = {ctl.decompX, ctl.decompY, ctl.decompZ, ctl.decompTS};
  const int kern[4] = {ctl.KernelX, ctl.KernelY, ctl.KernelZ, ctl.KernelTS};
  const int nd = ctl.nDecomp;
  x.launches = 0;
  // BWC_AMD_DWT=0 keeps every pass on the generic in-place kernel
  const char *dwt_env = std::getenv("BWC_AMD_DWT");
  const bool generic_only = dwt_env && dwt_env[0] == '0';
  // launch shape of the streaming passes: pair iterations unrolled per load
  // batch, thread target 2^seg_log2 and minimum iterations per segment,
  // nontemporal hints
  struct {
    int unroll, seg_log2, seg_min, nt;
    bool no_xy;
    int xy_min;
    bool no_inv;
  } tune = {tuning().dwt_unroll, tuning().dwt_seg_log2, tuning().dwt_seg_min, tuning().dwt_nt,
            tuning().dwt_xy == 0, tuning().dwt_xy_min_log2, tuning().dwt_inv_stream == 0};
  for (int step = 0; step < nd; ++step) {
    const int level = fwd ? step : nd - 1 - step;
    const Box &rb = par.res[nd - level].box;
    uint64_t sz[4];
    for (int a = 0; a < 4; ++a) sz[a] = rb.len(a);
    auto active = [&](int a) { return level < dec[a] && sz[a] > 1; };
    auto fast97 = [&](int a) { return fwd && !generic_only && kern[a] == bwc_dwt_9_7 && active(a); };
    // X lines that fit the LDS kernel take it; longer ones stream like the
    // strided axes (one thread per line segment)
    auto x_lds = [&](int a) {
      return a == 0 && fast97(0) && (size_t)4 * 2 * ((sz[0] + 16) / 2 + 1) * sizeof(double) <= 160 * 1024;
    };
    auto streamed = [&](int a) { return fast97(a) && !x_lds(a); };
    // every streamed pass moves the box to the other buffer; an odd count
    // is evened out by running the X pass out of place, else by a copy back
    const int nstrided = (int)streamed(0) + (int)streamed(1) + (int)streamed(2) + (int)streamed(3);
    const bool x_oop = (nstrided & 1) && x_lds(0);
    // X and Y in one pass (dwt97_xy_kernel) when Y streams, X fits the LDS
    // kernel with at most 1024 samples, the streamed passes are even in
    // number (the fused pass is out of place and counts as Y's) and the box
    // holds at least 2^dwt_xy_min_log2 values (2^24 by default: smaller ones
    // are latency-bound, and the separate passes run more workgroups on them)
    const bool xy = x_lds(0) && streamed(1) && !(nstrided & 1) && sz[0] <= 1024 &&
                    sz[0] * sz[1] * sz[2] * sz[3] >= (1ull << tune.xy_min) && !tune.no_xy;
    double *cur = work, *oth = x.work2;
    // inverse 9/7 along the strided axes on the streaming kernel, when the
    // level has an even number of such passes (each moves the box to the
    // other buffer; X stays on the in-place kernel)
    auto istream = [&](int a) {
      return !fwd && !generic_only && a != 0 && kern[a] == bwc_dwt_9_7 && active(a) && !tune.no_inv;
    };
    const int ninv = (int)istream(1) + (int)istream(2) + (int)istream(3);
    const bool inv = ninv > 0 && !(ninv & 1);
    if ((nstrided || inv) && x.work2_cap < par.box.volume() * sizeof(double)) {
      if (!x.grow) {
        std::snprintf(g_err, sizeof g_err, "DWT second buffer of %zu bytes is smaller than the %llu the box needs",
                      x.work2_cap, (unsigned long long)(par.box.volume() * sizeof(double)));
        return 1;
      }
      size_t cap = x.grow->work2_cap;
      CHK(ensure((void **)&x.grow->d_work2, &cap, par.box.volume() * sizeof(double)));
      x.grow->work2_cap = cap;
      x.work2 = x.grow->d_work2;
      x.work2_cap = cap;
      oth = x.work2;
    }
    if (norm_pending && !x_lds(0)) {
      CHK(normalize_on(x.stream, work, par.box.volume(), x.norm, (uint8_t)norm_qm));
      norm_pending = false;
    }
    for (int oi = 0; oi < 4; ++oi) {
      const int a = fwd ? oi : 3 - oi;
      if (!active(a)) continue;
      if (kern[a] == bwc_dwt_haar) {
        std::snprintf(g_err, sizeof g_err, "Haar kernel has no boundary extension in the codec");
        return 1;
      }
      // the three other axes; axis o[0] is the fastest in memory
      int o[3], k = 0;
      for (int b = 0; b < 4; ++b)
        if (b != a) o[k++] = b;
      const int64_t len = (int64_t)sz[a];
      const int par0 = (int)((fwd ? rb.lo[a] : rb.lo[a]) & 1);
      const bool contiguous = (a == 0);
      // lines per workgroup: ~24 KiB of LDS for contiguous lines; at least
      // 16 adjacent lines for strided axes so each row read spans 128 bytes
      const int64_t pitch = len + 2 * DWT_E + 1;
      int G = contiguous ? (int)std::max<int64_t>(1, std::min<int64_t>(64, 3072 / pitch))
                         : (int)std::max<int64_t>(16, std::min<int64_t>(64, 8192 / pitch));
      while (G > 1 && (size_t)G * pitch * sizeof(double) > 96 * 1024) --G;
      if ((uint64_t)G > sz[o[0]]) G = (int)sz[o[0]];
      const size_t lds = (size_t)G * (len + 2 * DWT_E + 1) * sizeof(double);
      const uint64_t gx = (sz[o[0]] + G - 1) / G;
      const uint64_t gy = sz[o[1]] * sz[o[2]];
      const int threads = 256;
      // forward 9/7 takes the fast kernels: LDS wavefront-per-line along X
      // (in place, or out of place to keep the level's buffer parity), a
      // streaming thread-per-line pass from the current buffer into the
      // other one along the strided axes
      const int half = par0 ? (int)(len / 2) : (int)((len + 1) / 2);
      const size_t xlds = (size_t)4 * 2 * ((len + 16) / 2 + 1) * sizeof(double);
      if (xy && a == 1) continue;   // done with X
      if (xy && a == 0) {
        const int leny = (int)sz[1], par0y = (int)(rb.lo[1] & 1);
        const int halfy = par0y ? leny / 2 : (leny + 1) / 2;
        const uint64_t planes = sz[2] * sz[3];
        // Y iterations split into segments until about four workgroups per
        // CU of work are there, keeping >= 16 pair iterations each
        const int kout = (leny - par0y + 5) / 2 - 1;
        int nseg = 1;
        while (planes * nseg < 1024 && kout / (nseg * 2) >= 16) nseg *= 2;
        const int ksz = (kout + nseg - 1) / nseg;
        constexpr int XU = 4;
        const size_t xylds = (size_t)2 * XU * 2 * ((len + 16) / 2 + 1) * sizeof(double);
        const uint64_t shift = (uint64_t)(norm_pending ? norm_qm : 0) << 52;
        const double *nrm = norm_pending ? x.norm : nullptr;
#define XY_LAUNCH(NT, CM)                                                                                           \
  hipLaunchKernelGGL((dwt97_xy_kernel<XU, NT, CM>), dim3((uint32_t)(planes * nseg)), dim3(256), xylds, x.stream, cur, \
                     oth, (int)len, par0, half, leny, par0y, halfy, inc[1], (uint32_t)sz[2], inc[2], inc[3], nseg,   \
                     ksz, nrm, shift)
#define XY_LAUNCH_C(NT)                   \
  if (len <= 256) XY_LAUNCH(NT, 1);       \
  else if (len <= 512) XY_LAUNCH(NT, 2);  \
  else XY_LAUNCH(NT, 4);
        switch (tune.nt) {
          case 1: XY_LAUNCH_C(1); break;
          case 2: XY_LAUNCH_C(2); break;
          case 3: XY_LAUNCH_C(3); break;
          default: XY_LAUNCH_C(0); break;
        }
#undef XY_LAUNCH_C
#undef XY_LAUNCH
        HIP_OK(hipGetLastError());
        norm_pending = false;
        std::swap(cur, oth);
        ++x.launches;
        continue;
      }
      if (x_lds(a)) {
        double *dst = x_oop ? oth : cur;
        hipLaunchKernelGGL(dwt97_x_kernel, dim3((uint32_t)((sz[o[0]] + 3) / 4), (uint32_t)gy), dim3(256), xlds,
                           x.stream, cur, dst, (int)len, par0, half, (uint32_t)sz[o[0]], inc[o[0]],
                           (uint32_t)sz[o[1]], inc[o[1]], inc[o[2]], norm_pending ? x.norm : nullptr,
                           (uint64_t)(norm_pending ? norm_qm : 0) << 52, tune.nt);
        HIP_OK(hipGetLastError());
        norm_pending = false;
        if (x_oop) std::swap(cur, oth);
        ++x.launches;
        continue;
      }
      if (streamed(a)) {
        const uint64_t nlines = sz[o[0]] * sz[o[1]] * sz[o[2]];
        // split lines into segments until about 128K threads run (short
        // lines of the coarse levels), keeping >= 8 pair iterations each
        const int kout = (int)((len - par0 + 5) / 2) - 1;
        int nseg = 1;
        while ((uint64_t)nseg * 2 * nlines <= (1ull << tune.seg_log2) && kout / (nseg * 2) >= tune.seg_min) nseg *= 2;
        const int ksz = (kout + nseg - 1) / nseg;
        const uint64_t nthreads = nlines * (uint64_t)nseg;
        const uint32_t nblk = (uint32_t)((nthreads + 255) / 256);
#define STREAM_LAUNCH(UU, NT)                                                                                      \
  hipLaunchKernelGGL((dwt97_stream_kernel<UU, NT>), dim3(nblk), dim3(256), 0, x.stream, cur, oth, (int)len, par0, \
                     half, inc[a], (uint32_t)sz[o[0]], inc[o[0]], (uint32_t)sz[o[1]], inc[o[1]], inc[o[2]],        \
                     (uint32_t)nlines, nseg, ksz)
#define STREAM_LAUNCH_U(NT)                  \
  if (tune.unroll == 2) STREAM_LAUNCH(2, NT); \
  else if (tune.unroll == 8) STREAM_LAUNCH(8, NT); \
  else STREAM_LAUNCH(4, NT);
        switch (tune.nt) {
          case 1: STREAM_LAUNCH_U(1); break;
          case 2: STREAM_LAUNCH_U(2); break;
          case 3: STREAM_LAUNCH_U(3); break;
          default: STREAM_LAUNCH_U(0); break;
        }
#undef STREAM_LAUNCH_U
#undef STREAM_LAUNCH
        HIP_OK(hipGetLastError());
        std::swap(cur, oth);
        ++x.launches;
        continue;
      }
      if (!fwd && !generic_only && a == 0 && kern[0] == bwc_dwt_9_7 && !tune.no_inv && xlds <= 160 * 1024) {
        // inverse X on the wavefront-per-line LDS kernel, in place
        hipLaunchKernelGGL(idwt97_x_kernel, dim3((uint32_t)((sz[o[0]] + 3) / 4), (uint32_t)gy), dim3(256), xlds,
                           x.stream, cur, cur, (int)len, par0, half, (uint32_t)sz[o[0]], inc[o[0]],
                           (uint32_t)sz[o[1]], inc[o[1]], inc[o[2]]);
        HIP_OK(hipGetLastError());
        ++x.launches;
        continue;
      }
      if (inv && istream(a)) {
        const uint64_t nlines = sz[o[0]] * sz[o[1]] * sz[o[2]];
        const int kout = (int)((len - par0 + 2) / 2);
        int nseg = 1;
        while ((uint64_t)nseg * 2 * nlines <= (1ull << tune.seg_log2) && kout / (nseg * 2) >= tune.seg_min) nseg *= 2;
        const int ksz = (kout + nseg - 1) / nseg;
        const uint32_t nblk = (uint32_t)((nlines * (uint64_t)nseg + 255) / 256);
#define ISTREAM_LAUNCH(NT)                                                                                           \
  hipLaunchKernelGGL((idwt97_stream_kernel<4, NT>), dim3(nblk), dim3(256), 0, x.stream, cur, oth, (int)len, par0,    \
                     half, inc[a], (uint32_t)sz[o[0]], inc[o[0]], (uint32_t)sz[o[1]], inc[o[1]], inc[o[2]],           \
                     (uint32_t)nlines, nseg, ksz)
        switch (tune.nt) {
          case 1: ISTREAM_LAUNCH(1); break;
          case 2: ISTREAM_LAUNCH(2); break;
          case 3: ISTREAM_LAUNCH(3); break;
          default: ISTREAM_LAUNCH(0); break;
        }
#undef ISTREAM_LAUNCH
        HIP_OK(hipGetLastError());
        std::swap(cur, oth);
        ++x.launches;
        continue;
      }
      if (lds > 160 * 1024) {
        std::snprintf(g_err, sizeof g_err, "DWT line of %lld samples exceeds the LDS line buffer", (long long)len);
        return 1;
      }
      auto launch = [&](auto kfn) {
        hipLaunchKernelGGL(kfn, dim3((uint32_t)gx, (uint32_t)gy), dim3(threads), lds, x.stream, cur, len, par0,
                           inc[a], sz[o[0]], inc[o[0]], sz[o[1]], inc[o[1]], inc[o[2]], G, contiguous ? 1 : 0);
      };
      if (kern[a] == bwc_dwt_9_7) {
        if (fwd) launch(dwt_axis_kernel<0, true>);
        else launch(dwt_axis_kernel<0, false>);
      } else {
        if (fwd) launch(dwt_axis_kernel<1, true>);
        else launch(dwt_axis_kernel<1, false>);
      }
      HIP_OK(hipGetLastError());
      ++x.launches;
    }
    if (cur != work) {
      const uint64_t total = sz[0] * sz[1] * sz[2] * sz[3];
      hipLaunchKernelGGL(box_copy_kernel, dim3((uint32_t)std::min<uint64_t>((total + 255) / 256, 16384)), dim3(256),
                         0, x.stream, cur, work, (uint32_t)sz[0], (uint32_t)sz[1], (uint32_t)sz[2], inc[1], inc[2],
                         inc[3], total);
      HIP_OK(hipGetLastError());
      ++x.launches;
    }
  }
  if (norm_pending) CHK(normalize_on(x.stream, work, par.box.volume(), x.norm, (uint8_t)norm_qm));
  return 0;
}

static int dwt_run(Workspace *ws, const bwc_codec *c, const Param &par, double *work, bool fwd, int norm_qm) {
  DwtCtx x;
  x.stream = ws->stream;
  x.work2 = ws->d_work2;
  x.work2_cap = ws->work2_cap;
  x.norm = ws->d_norm;
  x.grow = ws;
  const int rc = dwt_passes(x, c, par, work, fwd, norm_qm);
  ws->last_dwt_launches = x.launches;
  return rc;
}



// Everything the forward DWT's launch sequence depends on: the parameter's
// geometry, the buffers, the normalisation and the launch knobs.
static uint64_t dwt_key(const Workspace *ws, const bwc_codec *c, const Param &par, const double *work, int norm_qm) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) {
    h ^= v;
    h *= 1099511628211ull;
  };
  mix((uint64_t)(uintptr_t)&par);
  for (int a = 0; a < 4; ++a) mix(par.box.lo[a]), mix(par.box.hi[a]);
  mix(c->control.nDecomp);
  mix(c->control.decompX | (uint64_t)c->control.decompY << 8 | (uint64_t)c->control.decompZ << 16 |
      (uint64_t)c->control.decompTS << 24);
  mix(c->control.KernelX | (uint64_t)c->control.KernelY << 8 | (uint64_t)c->control.KernelZ << 16 |
      (uint64_t)c->control.KernelTS << 24);
  mix((uint64_t)(uintptr_t)work);
  mix((uint64_t)(uintptr_t)ws->d_work2);
  mix((uint64_t)(uintptr_t)ws->d_norm);
  mix((uint64_t)(int64_t)norm_qm);
  mix((uint64_t)tuning().dwt_unroll | (uint64_t)tuning().dwt_seg_log2 << 16 | (uint64_t)tuning().dwt_seg_min << 32 |
      (uint64_t)tuning().dwt_nt << 56 | (uint64_t)tuning().dwt_xy << 60);
  mix((uint64_t)tuning().dwt_xy_min_log2 | (uint64_t)tuning().dwt_inv_stream << 8);
  const char *d = std::getenv("BWC_AMD_DWT");
  mix(d && d[0] == '0');
  return h;
}

// The forward DWT of one tile parameter is 15-30 short launches for a 3-D
// box. With BWC_AMD_GRAPH=1 a compression repeated on the same codec replays
// them as one HIP graph (kernels only: no memset or allocation is recorded):
// the first call with a key runs them directly (allocating the second
// buffer), the second records the graph, later ones replay it. Off by
// default: the launches are asynchronous already and the stage is bound by
// the kernels, so a replay measured no faster (64^3: 0.083 ms direct, 0.088 ms
// replayed; 512^3: 1.516 / 1.526 ms).
int dwt_forward(Workspace *ws, const bwc_codec *c, const Param &par, double *work, int norm_qm) {
  const char *ge = std::getenv("BWC_AMD_GRAPH");
  if (!(ge && ge[0] == '1')) return dwt_run(ws, c, par, work, true, norm_qm);
  const uint64_t key = dwt_key(ws, c, par, work, norm_qm);
  for (Workspace::DwtGraph &g : ws->dwt_graphs) {
    if (g.key != key) continue;
    if (!g.exec) {
      hipGraph_t graph = nullptr;
      HIP_OK(hipStreamBeginCapture(ws->stream, hipStreamCaptureModeThreadLocal));
      const int rc = dwt_run(ws, c, par, work, true, norm_qm);
      const hipError_t e = hipStreamEndCapture(ws->stream, &graph);
      if (rc || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (!rc) std::snprintf(g_err, sizeof g_err, "DWT graph capture: %s", hipGetErrorString(e));
        return 1;
      }
      const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        g.exec = nullptr;
        std::snprintf(g_err, sizeof g_err, "DWT graph instantiation: %s", hipGetErrorString(ei));
        return 1;
      }
      g.launches = ws->last_dwt_launches;
    }
    HIP_OK(hipGraphLaunch(g.exec, ws->stream));
    ws->last_dwt_launches = g.launches;
    return 0;
  }
  CHK(dwt_run(ws, c, par, work, true, norm_qm));
  // keyed after the call, which allocates the second buffer on first use
  if (ws->dwt_graphs.size() < 64) ws->dwt_graphs.push_back({dwt_key(ws, c, par, work, norm_qm), nullptr, 0});
  return 0;
}
int dwt_forward_direct(Workspace *ws, const bwc_codec *c, const Param &par, double *work, int norm_qm) {
  return dwt_run(ws, c, par, work, true, norm_qm);
}
uint64_t dwt_launch_key(const Workspace *ws, const bwc_codec *c, const Param &par, const double *work, int norm_qm) {
  return dwt_key(ws, c, par, work, norm_qm);
}

namespace {
thread_local bool t_graphs_off = false;
}
void set_thread_graphs(bool on) { t_graphs_off = !on; }

bool graphs_on() {
  if (t_graphs_off) return false;
  const char *e = std::getenv("BWC_AMD_GRAPH");
  return !(e && e[0] == '0');
}

int graph_run(Workspace *ws, uint64_t key, const std::function<int()> &launch) {
  for (Workspace::Graph &g : ws->graphs) {
    if (g.key != key) continue;
    if (!g.exec) {
      hipGraph_t graph = nullptr;
      HIP_OK(hipStreamBeginCapture(ws->stream, hipStreamCaptureModeThreadLocal));
      const int rc = launch();
      const hipError_t e = hipStreamEndCapture(ws->stream, &graph);
      if (rc || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (!rc) std::snprintf(g_err, sizeof g_err, "graph capture: %s", hipGetErrorString(e));
        return 1;
      }
      const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        g.exec = nullptr;
        std::snprintf(g_err, sizeof g_err, "graph instantiation: %s", hipGetErrorString(ei));
        return 1;
      }
      ++ws->graph_records;
    }
    HIP_OK(hipGraphLaunch(g.exec, ws->stream));
    ++ws->graph_replays;
    return 0;
  }
  CHK(launch());
  if (ws->graphs.size() >= 256) {
    for (auto &g : ws->graphs)
      if (g.exec) HIP_OK(hipGraphExecDestroy(g.exec));
    ws->graphs.clear();
  }
  ws->graphs.push_back({key, nullptr});
  return 0;
}

int dwt_inverse(Workspace *ws, const bwc_codec *c, const Param &par, double *work) {
  return dwt_run(ws, c, par, work, false, -1);
}

// ===========================================================================
// Fill + forward DWT with no allocation: the min/max partials and the second
// buffer of the out-of-place passes live in `scratch`, a static device array
// (each GPU the library runs on holds its own copy until the process ends).
// Its device address is looked up once per device with hipGetSymbolAddress;
// host code uses only that address, never the array's name. Calls that
// overlap on one GPU share the array and race.
// ===========================================================================
constexpr size_t kScratchRed = 2 * (size_t)kRedSlots * sizeof(double);   // min/max partials
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64];

static int scratch_of_device(void **out) {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    std::snprintf(g_err, sizeof g_err, "device %d out of range for the scratch table", dev);
    return 1;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

int dwt_noalloc(const bwc_codec *c, const Tile &tl, int param, const void *input, double *out, double *norm,
                hipStream_t stream) {
  const Param &par = tl.par[param];
  const uint64_t n = par.box.volume();
  if (n * sizeof(double) > kNoAllocScratchBytes - kScratchRed) {
    std::snprintf(g_err, sizeof g_err,
                  "tile parameter of %llu values exceeds the allocation-free limit of %llu values",
                  (unsigned long long)n, (unsigned long long)((kNoAllocScratchBytes - kScratchRed) / sizeof(double)));
    return 1;
  }
  if (!is_device_pointer(input) || !is_device_pointer(out) || !is_device_pointer(norm)) {
    std::snprintf(g_err, sizeof g_err, "the allocation-free DWT takes device pointers only");
    return 1;
  }
  CHK(process_setup());
  void *sp = nullptr;
  CHK(scratch_of_device(&sp));
  double *red = static_cast<double *>(sp);
  double *work2 = reinterpret_cast<double *>(static_cast<unsigned char *>(sp) + kScratchRed);
  const uint64_t nX = c->info.nX, nY = c->info.nY, nZ = c->info.nZ, nTS = c->info.nTS;
  const int prec = c->info.data_prec == bwc_precision_single ? 4 : 8;
  const int threads = 256;
  const int blocks = (int)std::max<uint64_t>(1, std::min<uint64_t>((n + threads - 1) / threads, 4096));
  hipLaunchKernelGGL(fill_minmax_kernel, dim3(blocks), dim3(threads), 0, stream, input, prec, nX, nY, nZ,
                     nX * nY * nZ * nTS * (uint64_t)param, (uint64_t)tl.box.lo[0], (uint64_t)tl.box.lo[1],
                     (uint64_t)tl.box.lo[2], par.box.len(0), par.box.len(1), par.box.len(2), (uint64_t)0, n, out,
                     red);
  hipLaunchKernelGGL(minmax_final_kernel, dim3(1), dim3(256), 0, stream, red, blocks, norm);
  HIP_OK(hipGetLastError());
  DwtCtx x;
  x.stream = stream;
  x.work2 = work2;
  x.work2_cap = kNoAllocScratchBytes - kScratchRed;
  x.norm = norm;
  return dwt_passes(x, c, par, out, true, c->control.Qm);
}

// ===========================================================================
// quantisation into stripe columns (tier1.c:526-729)
// ===========================================================================
// (uint64_t) of a non-negative double exactly as x86-64 code compiled by gcc
// computes it: truncation in range, 2^63 for NaN, 0 at or above 2^64.
__device__ __forceinline__ uint64_t x86_to_u64(double v) {
  const double two63 = 9223372036854775808.0;
  if (v < two63) return (uint64_t)(int64_t)v;
  if (v != v) return 0x8000000000000000ull;
  const double r = v - two63;
  if (r < two63) return ((uint64_t)(int64_t)r) ^ 0x8000000000000000ull;
  return 0;
}

__global__ void quantize_kernel(const double *__restrict__ work, uint64_t W, uint64_t H, uint64_t D,
                                const uint32_t *__restrict__ cbbox, const double *__restrict__ cbsb,
                                const uint64_t *__restrict__ qoff, uint64_t *__restrict__ q,
                                uint16_t *__restrict__ state, uint8_t *__restrict__ Kout, uint64_t ncb) {
  __shared__ int smax[256];
  for (uint64_t cb = blockIdx.x; cb < ncb; cb += gridDim.x) {
    const uint32_t *b = cbbox + 8 * cb;
    const uint64_t cw = b[4] - b[0], ch = b[5] - b[1], cd = b[6] - b[2], ct = b[7] - b[3];
    const uint64_t ns = (ch + 3) / 4;
    const uint64_t ncol = cw * ns * cd * ct;
    const double step = cbsb[3 * cb];
    const uint64_t qo = qoff[cb];
    int kb = 0;
    for (uint64_t k = threadIdx.x; k < ncol; k += blockDim.x) {
      const uint64_t x = k % cw, j = (k / cw) % ns, sl = k / (cw * ns);
      const uint64_t z = sl % cd, t = sl / cd;
      uint16_t xi = 0;
      for (int r = 0; r < 4; ++r) {
        const uint64_t row = 4 * j + r;
        uint64_t qq = 0;
        if (row < ch) {
          const uint64_t idx = (b[0] + x) + W * ((b[1] + row) + H * ((b[2] + z) + D * (b[3] + t)));
          uint64_t raw = (uint64_t)__double_as_longlong(work[idx]);
          xi |= (uint16_t)((raw >> 63) << (3 - r));
          raw &= 0x7FFFFFFFFFFFFFFFull;
          qq = x86_to_u64(__longlong_as_double((long long)raw) / step);
          const int nb = qq ? 64 - __clzll((long long)qq) : 0;
          kb = nb > kb ? nb : kb;
        }
        q[qo + 4 * k + (3 - r)] = qq;
      }
      state[qo / 4 + k] = (uint16_t)(xi << 12);
    }
    smax[threadIdx.x] = kb;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) smax[threadIdx.x] = max(smax[threadIdx.x], smax[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) Kout[cb] = (uint8_t)smax[0];
    __syncthreads();
  }
}

// ===========================================================================
// tier-1: one lane per codeblock (reference-order passes + MQ coder)
// ===========================================================================
// MQ probability states (ITU-T T.800 Table C.2): qe | nmps << 16 | nlps << 22 | switch << 28
__constant__ uint32_t c_mq[47];
__constant__ int32_t c_tm[64];
__constant__ uint8_t c_sig[3][45];
__constant__ uint8_t c_sig2xi[16];
__constant__ uint8_t c_xi2cont[11][2];
__constant__ uint16_t c_frac2log[256];
__constant__ uint64_t c_sigpk[3][3];   // c_sig packed as 4-bit entries, 16 per word

struct MqState {
  int64_t L;
  uint32_t C;
  uint16_t A;
  int8_t t;
  uint8_t T;
};

struct Coder {
  MqState s;
  uint8_t *buf;
  uint64_t cap;
  uint32_t ovf;
  uint8_t *ctx;   // LDS, stride 64: state index | mps << 7
  int lane;
  __device__ __forceinline__ uint8_t &cx(int k) { return ctx[k * 64 + lane]; }

  __device__ __forceinline__ void put(int64_t L, uint8_t v) {
    if (L >= 0) {
      if ((uint64_t)L < cap) buf[L] = v;
      else ovf = 1;
    }
  }
  // transfer_byte (mq.c:85-127)
  __device__ __forceinline__ void byte_out() {
    if (s.T == 0xFF) {
      put(s.L, s.T);
      s.L++;
      s.T = (uint8_t)(s.C >> 20);
      s.C &= 0xFFFFF;
      s.t = 7;
    } else {
      if (s.C & 0x8000000) {
        s.T++;
        s.C &= 0x7FFFFFF;
      }
      put(s.L, s.T);
      s.L++;
      if (s.T == 0xFF) {
        s.T = (uint8_t)(s.C >> 20);
        s.C &= 0xFFFFF;
        s.t = 7;
      } else {
        s.T = (uint8_t)(s.C >> 19);
        s.C &= 0x7FFFF;
        s.t = 8;
      }
    }
  }
  // mq_bit_encode (mq.c:477-556)
  __device__ __forceinline__ void encode(uint32_t sym, int k) {
    uint8_t &st = cx(k);
    const uint32_t e = c_mq[st & 0x3F];
    const uint16_t p = (uint16_t)(e & 0xFFFF);
    const uint32_t mps = st >> 7;
    s.A -= p;
    if (sym == mps) {
      if (s.A >= 0x8000) {
        s.C += p;
        return;
      }
      if (s.A < p) s.A = p;
      else s.C += p;
      st = (uint8_t)(((e >> 16) & 0x3F) | (mps << 7));
    } else {
      if (s.A < p) s.C += p;
      else s.A = p;
      const uint32_t nm = ((e >> 28) & 1) ? (mps ^ 1) : mps;
      st = (uint8_t)(((e >> 22) & 0x3F) | (nm << 7));
    }
    do {
      s.A <<= 1;
      s.C <<= 1;
      s.t--;
      if (s.t == 0) byte_out();
    } while (s.A < 0x8000);
  }
};

struct Blk {
  uint64_t *q;       // 4 per stripe column (l = 3 is the stripe's top row)
  uint16_t *st;      // state per stripe column
  uint64_t cw, ns, nslice, ch;
  __device__ __forceinline__ uint32_t sig(int64_t k) const { return k < 0 ? 0u : (st[k] & 0xF); }
  __device__ __forceinline__ uint32_t xi(int64_t k) const { return k < 0 ? 0u : ((st[k] >> 12) & 0xF); }
  __device__ __forceinline__ uint32_t nib(int64_t k, int b) const {
    const uint64_t *p = q + 4 * k;
    return (uint32_t)(((p[0] >> b) & 1) | (((p[1] >> b) & 1) << 1) | (((p[2] >> b) & 1) << 2) |
                      (((p[3] >> b) & 1) << 3));
  }
};

// neighbour links of stripe column k (negative = insignificant dummy)
struct Nb {
  int64_t l, r, u, d, ul, ur, dl, dr;
};
__device__ __forceinline__ Nb links(const Blk &B, int64_t k) {
  const uint64_t x = (uint64_t)k % B.cw, j = ((uint64_t)k / B.cw) % B.ns;
  Nb n;
  n.l = x > 0 ? k - 1 : -1;
  n.r = x + 1 < B.cw ? k + 1 : -1;
  n.u = j > 0 ? k - (int64_t)B.cw : -1;
  n.d = j + 1 < B.ns ? k + (int64_t)B.cw : -1;
  n.ul = n.u >= 0 && x > 0 ? n.u - 1 : -1;
  n.ur = n.u >= 0 && x + 1 < B.cw ? n.u + 1 : -1;
  n.dl = n.d >= 0 && x > 0 ? n.d - 1 : -1;
  n.dr = n.d >= 0 && x + 1 < B.cw ? n.d + 1 : -1;
  return n;
}

__device__ __forceinline__ uint32_t sig_index(const Blk &B, int64_t k, const Nb &n, uint32_t bm, int l, bool top,
                                              bool bot) {
  const uint32_t s = B.sig(k), L = B.sig(n.l), R = B.sig(n.r);
  uint32_t kh = (L & bm) + (R & bm);
  uint32_t kv = (bm & (s >> 1)) + (bm & (s << 1));
  uint32_t kd = (bm & (L >> 1)) + (bm & (L << 1)) + (bm & (R >> 1)) + (bm & (R << 1));
  if (top) {
    kv += (0x01 & B.sig(n.u)) << 3;
    kd += ((0x01 & B.sig(n.ul)) + (0x01 & B.sig(n.ur))) << 3;
  } else if (bot) {
    kv += (0x08 & B.sig(n.d)) >> 3;
    kd += ((0x08 & B.sig(n.dl)) + (0x08 & B.sig(n.dr))) >> 3;
  }
  return (15 * kh + 5 * kv + kd) >> l;
}

__device__ __forceinline__ void code_sign(Coder &C, const Blk &B, int64_t k, const Nb &n, uint32_t bm, int l, bool top,
                                          bool bot) {
  const uint32_t s = B.sig(k), x = B.xi(k);
  const uint32_t L = B.sig(n.l), R = B.sig(n.r), xL = B.xi(n.l), xR = B.xi(n.r);
  uint32_t xh = ((bm & L) | ((bm & xL) << 2) | ((bm & R) << 1) | ((bm & xR) << 3)) >> l;
  uint32_t xv = ((bm & (s >> 1)) | ((bm & (x >> 1)) << 2) | ((bm & (s << 1)) << 1) | ((bm & (x << 1)) << 3)) >> l;
  if (top) xv |= (0x01 & B.sig(n.u)) | ((0x01 & B.xi(n.u)) << 2);
  else if (bot) xv |= ((0x08 & B.sig(n.d)) >> 2) | (0x08 & B.xi(n.d));
  xh = c_sig2xi[xh];
  xv = c_sig2xi[xv];
  const int idx = (int)((xh << 2) | xv);
  C.encode(((bm & x) >> l) ^ c_xi2cont[idx][1], c_xi2cont[idx][0]);
}

__device__ int64_t pass_spp(Coder &C, Blk &B, int b, const uint8_t *s2c) {
  int64_t mse = 0;
  const uint32_t rest = 4 - (uint32_t)(B.ch & 3);
  const uint32_t smask = (uint32_t)(~((0x0F >> (B.ch & 3)) | 0xF0)) & 0xFF;
  const uint32_t dsh = b < 5 ? 0 : b - 5, dcor = b < 5 ? 5 - b : 0;
  const uint64_t nfull = B.ch >> 2;
  int64_t k = 0;
  for (uint64_t sl = 0; sl < B.nslice; ++sl) {
    for (uint64_t rowblk = 0; rowblk < B.ns; ++rowblk) {
      const bool full = rowblk < nfull;
      for (uint64_t x = 0; x < B.cw; ++x, ++k) {
        const uint32_t s0 = B.sig(k);
        if (full ? (s0 ^ 0xF) : ((s0 ^ 0xF) & smask)) {
          const Nb n = links(B, k);
          const uint32_t bits = B.nib(k, b);
          const int lmin = full ? 0 : (int)rest;
          for (int l = 3; l >= lmin; --l) {
            const uint32_t bm = 1u << l;
            if (!(B.sig(k) & bm)) {
              const uint32_t ks = sig_index(B, k, n, bm, l, l == 3, full && l == 0);
              if (ks) {
                C.encode((bits >> l) & 1, s2c[ks]);
                if (bits & bm) {
                  code_sign(C, B, k, n, bm, l, l == 3, full && l == 0);
                  mse += 49152 + 6144 * (int32_t)(((B.q[4 * k + l] >> dsh) << dcor) & 0x1F);
                  B.st[k] |= (uint16_t)bm;
                }
                B.st[k] |= (uint16_t)(bm << 4);
              }
            } else {
              B.st[k] &= (uint16_t)~(bm << 4);
            }
          }
        } else {
          B.st[k] &= (uint16_t)~0xF0;
        }
      }
    }
  }
  return mse;
}

__device__ int64_t pass_mrp(Coder &C, Blk &B, int b) {
  int64_t mse = 0;
  const uint32_t rest = 4 - (uint32_t)(B.ch & 3);
  const uint32_t smask = (uint32_t)(~((0x0F >> (B.ch & 3)) | 0xF0)) & 0xFF;
  const uint32_t dsh = b < 6 ? 0 : b - 6, dcor = b < 6 ? 6 - b : 0;
  const uint64_t nfull = B.ch >> 2;
  int64_t k = 0;
  for (uint64_t sl = 0; sl < B.nslice; ++sl) {
    for (uint64_t rowblk = 0; rowblk < B.ns; ++rowblk) {
      const bool full = rowblk < nfull;
      for (uint64_t x = 0; x < B.cw; ++x, ++k) {
        const uint16_t stv = B.st[k];
        uint32_t sg = (stv & 0xF) & ~((stv >> 4) & 0xF);
        if (!full) sg &= smask;
        if (!sg) continue;
        const Nb n = links(B, k);
        const uint32_t bits = B.nib(k, b);
        const uint32_t s = stv & 0xF, L = B.sig(n.l), R = B.sig(n.r);
        const int lmin = full ? 0 : (int)rest;
        for (int l = 3; l >= lmin; --l) {
          const uint32_t bm = 1u << l;
          if (!(sg & bm)) continue;
          int km = 15;
          if ((B.st[k] >> 8) & bm) {
            km += 2;
          } else if ((bm & (s >> 1)) || (bm & (s << 1)) || (L & bm) || (R & bm) || (bm & (L >> 1)) ||
                     (bm & (L << 1)) || (bm & (R >> 1)) || (bm & (R << 1))) {
            km += 1;
          } else if (l == 3 && ((0x01 & B.sig(n.u)) || (0x01 & B.sig(n.ul)) || (0x01 & B.sig(n.ur)))) {
            km += 1;
          } else if (full && l == 0 && ((0x08 & B.sig(n.d)) || (0x08 & B.sig(n.dl)) || (0x08 & B.sig(n.dr)))) {
            km += 1;
          }
          C.encode((bits >> l) & 1, km);
          mse += c_tm[((B.q[4 * k + l] >> dsh) << dcor) & 0x3F];
          B.st[k] |= (uint16_t)(bm << 8);
        }
      }
    }
  }
  return mse;
}

__device__ int64_t pass_cup(Coder &C, Blk &B, int b, const uint8_t *s2c) {
  int64_t mse = 0;
  const uint32_t rest = 4 - (uint32_t)(B.ch & 3);
  const uint32_t dsh = b < 5 ? 0 : b - 5, dcor = b < 5 ? 5 - b : 0;
  const uint64_t nfull = B.ch >> 2;
  int64_t k = 0;
  for (uint64_t sl = 0; sl < B.nslice; ++sl) {
    for (uint64_t rowblk = 0; rowblk < B.ns; ++rowblk) {
      const bool full = rowblk < nfull;
      for (uint64_t x = 0; x < B.cw; ++x, ++k) {
        const Nb n = links(B, k);
        const uint32_t bits = B.nib(k, b);
        int r = -1;
        if (full) {
          if (!(0x08 & B.sig(n.dl)) && !(0x08 & B.sig(n.d)) && !(0x08 & B.sig(n.dr)) && !(0x01 & B.sig(n.u)) &&
              !(0x01 & B.sig(n.ul)) && !B.sig(n.l) && !(0x01 & B.sig(n.ur)) && !B.sig(n.r) && !B.sig(k)) {
            if (bits) {
              uint32_t sm = bits | (bits >> 1);
              sm |= sm >> 2;
              r = sm == 1 ? 3 : (sm == 3 ? 2 : (sm == 7 ? 1 : 0));
              C.encode(1, 9);
              C.encode(((uint32_t)r >> 1) & 1, 18);
              C.encode((uint32_t)r & 1, 18);
            } else {
              r = 4;
              C.encode(0, 9);
            }
          }
        }
        const int lmin = full ? 0 : (int)rest;
        for (int l = 3; l >= lmin; --l) {
          const uint32_t bm = 1u << l;
          const uint16_t stv = B.st[k];
          if (!(stv & bm) && !((stv >> 4) & bm)) {
            if (r >= 0) --r;
            else C.encode((bits >> l) & 1, s2c[sig_index(B, k, n, bm, l, l == 3, full && l == 0)]);
            if (bits & bm) {
              code_sign(C, B, k, n, bm, l, l == 3, full && l == 0);
              mse += 49152 + 6144 * (int32_t)(((B.q[4 * k + l] >> dsh) << dcor) & 0x1F);
              B.st[k] |= (uint16_t)bm;
            }
          }
        }
      }
    }
  }
  return mse;
}

// minimal truncation length of one pass snapshot (mq.c:586-675, FP64 build)
__device__ __forceinline__ uint32_t byte_at(const uint8_t *buf, uint64_t cap, int64_t i) {
  return (i >= 0 && (uint64_t)i < cap) ? buf[i] : 0u;
}

// *reach: one past the highest byte index the length search read
__device__ int64_t trunc_len(MqState s, const uint8_t *buf, uint64_t cap, int64_t *reach) {
  int64_t Cr = ((int64_t)s.T << 27) + ((int64_t)s.C << s.t);
  int64_t CAr = ((int64_t)s.A << s.t) + Cr;
  int sh = 8;
  // bounded: each round consumes one byte; the terminated codeword ends
  // within the buffer, so the loop exits long before 2^20 rounds
  for (int guard = 0; guard < (1 << 20) && (((Cr >> 27) > 0xff) || ((CAr >> 27) <= 0xff)); ++guard) {
    const int64_t v = byte_at(buf, cap, s.L);
    s.L++;
    Cr -= v << (35 - sh);
    CAr -= v << (35 - sh);
    Cr <<= sh;
    CAr <<= sh;
    sh = (v == 0xFF) ? 7 : 8;
  }
  *reach = s.L;
  if (s.L > 0 && byte_at(buf, cap, s.L - 1) == 0xFF) s.L--;
  while (s.L > 1 && byte_at(buf, cap, s.L - 2) == 0xFF && byte_at(buf, cap, s.L - 1) == 0x7F) s.L -= 2;
  return s.L;
}

__device__ uint16_t slope2log(double lambda) {
  const uint64_t raw = (uint64_t)__double_as_longlong(lambda);
  const uint64_t ex = ((raw & 0x7FF0000000000000ull) >> 44) - 0x41F00;
  const uint64_t ma = (raw & 0x000FFFFFFFFFFFFFull) >> 44;
  const uint64_t v = (ex + c_frac2log[ma]) + 0x10000;
  if (v > 0xFFFF) return 0xFFFF;
  if (v < 2) return 2;
  return (uint16_t)v;
}

// nfinal: bytes [0, nfinal) of buf are final (a longer coding of the same
// codeblock starts with them); *U receives the first pass whose length
// search read past them (Z if none)
__device__ void finish_block(uint8_t *buf, uint64_t cap, const uint64_t *sn, const double *mse, int Z,
                             uint32_t *L, uint16_t *S, uint32_t *flags, int64_t nfinal, uint8_t *U);

__global__ void __launch_bounds__(64) t1_encode_kernel(
    uint64_t ncb, const uint32_t *__restrict__ cbbox, const double *__restrict__ cbsb,
    const uint32_t *__restrict__ cbmeta, const uint64_t *__restrict__ qoff, const uint64_t *__restrict__ doff,
    uint64_t *__restrict__ q, uint16_t *__restrict__ state, uint8_t *__restrict__ data, uint8_t *__restrict__ Kp,
    uint8_t *__restrict__ Kmsbsp, uint8_t *__restrict__ Zp, uint32_t *__restrict__ Lp, uint16_t *__restrict__ Sp,
    uint64_t *__restrict__ snap, double *__restrict__ msep, uint32_t *__restrict__ flags, int erres) {
  __shared__ uint8_t ctx[19 * 64];
  const uint64_t cb = (uint64_t)blockIdx.x * 64 + threadIdx.x;
  if (cb >= ncb) return;
  const int lane = threadIdx.x;
  const uint32_t *bx = cbbox + 8 * cb;
  const uint64_t cw = bx[4] - bx[0], ch = bx[5] - bx[1], cd = bx[6] - bx[2], ct = bx[7] - bx[3];
  const int K = Kp[cb];
  uint16_t *S = Sp + cb * 191;
  uint32_t *L = Lp + cb * 190;
  if (K == 0) {
    S[0] = 0xFFFF;
    Zp[cb] = 0;
    Kmsbsp[cb] = 0;
    return;
  }
  const uint32_t meta = cbmeta[cb];
  const uint32_t hb = meta & 0xFF, Kmax = (meta >> 8) & 0xFF;
  const int sidx = (hb == 1 || hb == 5 || hb == 9 || hb == 13) ? 1 : ((hb & 3) == 3 ? 2 : 0);
  const uint8_t *s2c = c_sig[sidx];
  Blk B;
  B.q = q + qoff[cb];
  B.st = state + qoff[cb] / 4;
  B.cw = cw;
  B.ns = (ch + 3) / 4;
  B.nslice = cd * ct;
  B.ch = ch;
  Coder C;
  C.buf = data + doff[cb];
  C.cap = doff[cb + 1] - doff[cb];
  C.ovf = 0;
  C.ctx = ctx;
  C.lane = lane;
  for (int i = 0; i < 19; ++i) C.cx(i) = 0;
  C.cx(0) = 4;
  C.cx(9) = 3;
  C.cx(18) = 46;
  C.s.A = 0x8000;
  C.s.t = 12;
  C.s.L = -1;
  C.s.C = 0;
  C.s.T = 0;
  uint64_t *sn = snap + cb * 2 * 190;
  double *mse = msep + cb * 190;
  double scale = cbsb[3 * cb + 2];
  scale *= (cbsb[3 * cb + 1] * cbsb[3 * cb + 1]);
  scale *= ldexp(1.0, 2 * K);
  scale /= 0x10000;
  int j = 0, i = 2;
  for (int k = K; k-- > 0; i = 0) {
    for (; i < 3; ++i, ++j) {
      // pass j starts from the live state; its end state is snapshotted below
      const int64_t dm = (i == 0) ? pass_spp(C, B, k, s2c) : (i == 1 ? pass_mrp(C, B, k) : pass_cup(C, B, k, s2c));
      mse[j] = (double)dm * scale;
      if (!(i == 2 && erres)) {
        sn[2 * j] = (uint64_t)C.s.L;
        sn[2 * j + 1] = (uint64_t)C.s.C | ((uint64_t)C.s.A << 32) | ((uint64_t)(uint8_t)C.s.t << 48) |
                        ((uint64_t)C.s.T << 56);
      }
    }
    if (erres) {
      C.encode(1, 18);
      C.encode(0, 18);
      C.encode(1, 18);
      C.encode(0, 18);
      sn[2 * (j - 1)] = (uint64_t)C.s.L;
      sn[2 * (j - 1) + 1] = (uint64_t)C.s.C | ((uint64_t)C.s.A << 32) | ((uint64_t)(uint8_t)C.s.t << 48) |
                            ((uint64_t)C.s.T << 56);
    }
    scale *= 0.25;
  }
  const int Z = j;
  // termination of the last state (mq.c:706-741); the live state is restored
  {
    MqState keep = C.s;
    int n = 12 - C.s.t;
    C.s.C <<= C.s.t;
    while (n > 0) {
      C.byte_out();
      n -= C.s.t;
      C.s.C <<= C.s.t;
    }
    C.byte_out();
    C.s = keep;
  }
  if (C.ovf) {
    atomicOr(&flags[0], 1u);
    Zp[cb] = 0;
    return;
  }
  Kmsbsp[cb] = (uint8_t)(Kmax - K);
  Zp[cb] = (uint8_t)Z;
  finish_block(C.buf, C.cap, sn, mse, Z, L, S, flags, INT64_MAX, nullptr);
}

// Truncation lengths of every pass snapshot (mq.c:706-759), the convex hull
// and the log slopes (tier1.c:2161-2267) of one finished codeblock; folds its
// non-zero slopes into the tile bounds (tier1.c:3047-3061).
__device__ void finish_block(uint8_t *buf, uint64_t cap, const uint64_t *sn, const double *mse, int Z,
                             uint32_t *L, uint16_t *S, uint32_t *flags, int64_t nfinal, uint8_t *U) {
  int u = Z;
  for (int z = 0; z < Z; ++z) {
    MqState s;
    s.L = (int64_t)sn[2 * z];
    const uint64_t w = sn[2 * z + 1];
    s.C = (uint32_t)w;
    s.A = (uint16_t)(w >> 32);
    s.t = (int8_t)(uint8_t)(w >> 48);
    s.T = (uint8_t)(w >> 56);
    int64_t reach = 0;
    L[z] = (uint32_t)trunc_len(s, buf, cap, &reach);
    if (reach > nfinal && z < u) u = z;
  }
  if (U) *U = (uint8_t)u;
  uint8_t hull[192];
  double lam[191];
  for (int z = 0; z <= Z; ++z) lam[z] = 0;
  int h = 0;
  hull[0] = 0;
  int hlast = 0;
  lam[0] = 18446744073709551616.0;
  for (int ii = 0; ii < Z; ++ii) {
    double dD = 0;
    for (int jj = hlast; jj <= ii; ++jj) dD += mse[jj];
    uint64_t dL = hlast > 0 ? (uint64_t)(L[ii] - L[hlast - 1]) : (uint64_t)L[ii];
    if (dD > 0) {
      while (h > 0 && dD >= lam[hlast] * (double)dL) {
        lam[hlast] = 0;
        --h;
        hlast = hull[h];
        dD = 0;
        for (int jj = hlast; jj <= ii; ++jj) dD += mse[jj];
        dL = hlast > 0 ? (uint64_t)(L[ii] - L[hlast - 1]) : (uint64_t)L[ii];
      }
      ++h;
      hlast = ii + 1;
      hull[h] = (uint8_t)hlast;
      if (dL == 0) lam[hlast] = dD < 0 ? -1.0 : (dD == 0 ? 0.0 : 18446744073709551616.0);
      else lam[hlast] = dD / (double)dL;
    } else {
      lam[ii + 1] = 0;
    }
  }
  uint32_t smin = 0xFFFF, smax = 0;
  for (int z = 0; z <= Z; ++z) {
    uint16_t v;
    if (lam[z] < 0) v = 1;
    else if (lam[z] == 0) v = (z == Z) ? 1 : 0;
    else v = slope2log(lam[z]);
    S[z] = v;
    if (v) {
      smin = v < smin ? v : smin;
      smax = v > smax ? v : smax;
    }
  }
  atomicMin(&flags[1], smin);
  atomicMax(&flags[2], smax);
}

// quantisation straight into bitplanes for the fast tier-1 path
// (tier1.c:630-728): per codeblock, plane b holds one nibble per stripe
// column (row 0 in bit 3), two columns per byte; plane 63 holds the signs.
// Per-codeblock features for the pass-limit network (written when `feat` is
// set; 16 floats per codeblock, entry 15 filled later from tile statistics):
//  0 K/64   1 (Kmax-K)/16   2 resolution/8   3 highpass axes/4
//  4 mean magnitude bit length/64   5 log2(mean q + 1)/64
//  6 log2(rms q + 1)/64   7..12 fraction of values whose bit length is at
//  least K - {0,1,2,4,6,9}   13 log2(gain * step^2)/64   14 log2(bits/value)
// and in tstat[0..1] the sum and count of log2(rms q + 1) + log2(gain *
// step^2)/2 (the codeblock's distortion-weighted level).
struct BpGeom {
  uint64_t W, H, D;
  uint32_t b0, b1, b2, b3, cw, ch, cd, ns, ncol;
  double step;
};
// The four raw values of stripe column k (row r at index r); values outside
// the codeblock (rows past its height, columns k >= ncol) are 0.
__device__ __forceinline__ void bp_load(const double *__restrict__ work, const BpGeom &g, uint32_t k,
                                        uint64_t raw[4]) {
  const uint32_t x = k % g.cw, rest = k / g.cw;
  const uint32_t j = rest % g.ns, sl = rest / g.ns;
  const uint32_t z = sl % g.cd, t = sl / g.cd;
  const bool colv = k < g.ncol;
  // slots outside the codeblock read its first value and drop it
  const uint64_t first = (uint64_t)g.b0 + g.W * ((uint64_t)g.b1 + g.H * ((uint64_t)g.b2 + g.D * g.b3));
  const uint64_t base = (uint64_t)(g.b0 + x) + g.W * ((uint64_t)(g.b1 + 4 * j) + g.H * ((uint64_t)(g.b2 + z) +
                                                                                       g.D * (g.b3 + t)));
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const bool v = colv && 4 * j + r < g.ch;
    const uint64_t w = (uint64_t)__double_as_longlong(work[v ? base + g.W * r : first]);
    raw[r] = v ? w : 0;
  }
}
// |value| / step, the quotient that quantisation converts
__device__ __forceinline__ double bp_quot(uint64_t raw, double step) {
  return __longlong_as_double((long long)(raw & 0x7FFFFFFFFFFFFFFFull)) / step;
}

// Bit transpose across the eight lanes of an octet, lane c holding stripe
// column 8g + c: on entry m[3 - r] is a 32-bit slice of row r's magnitude
// (value bits as columns), on exit m[i] is the word of plane 4c + i for the
// octet's four column pairs (byte u = pair 4g + u, bit 3 - r + 4h for row r
// of column 8g + 2u + h). Exchanges off-diagonal blocks of 16, 8 and 4 bits
// between lanes (rows 4c + i and 4c + i +- 16, 8, 4), then of 2 and 1 bits
// within a lane.
__device__ __forceinline__ void octet_transpose(uint32_t m[4]) {
#pragma unroll
  for (int st = 0; st < 3; ++st) {
    const int s = 16 >> st, x = 4 >> st;
    const uint32_t mask = st == 0 ? 0x0000FFFFu : st == 1 ? 0x00FF00FFu : 0x0F0F0F0Fu;
    const bool up = threadIdx.x & x;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t p = __shfl_xor(m[i], x, 64);
      const uint32_t t = (((up ? p : m[i]) >> s) ^ (up ? m[i] : p)) & mask;
      m[i] ^= up ? t : t << s;
    }
  }
#pragma unroll
  for (int s = 2, k = 0; s > 0; s >>= 1, ++k) {
    const uint32_t mask = k == 0 ? 0x33333333u : 0x55555555u;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (!(i & s)) {
        const uint32_t t = ((m[i] >> s) ^ m[i + s]) & mask;
        m[i] ^= t << s;
        m[i + s] ^= t;
      }
  }
}

__global__ void __launch_bounds__(256) bitplane_kernel(const double *__restrict__ work, uint64_t W, uint64_t H,
                                                       uint64_t D, const uint32_t *__restrict__ cbbox,
                                                       const double *__restrict__ cbsb,
                                                       const uint64_t *__restrict__ bpoff, uint8_t *__restrict__ bp,
                                                       uint8_t *__restrict__ Kout, uint64_t ncb,
                                                       const uint32_t *__restrict__ cbmeta, float *__restrict__ feat,
                                                       float *__restrict__ tstat, float rate,
                                                       uint32_t *__restrict__ hist_out) {
  __shared__ int smax[16];
  __shared__ uint32_t hist[65];
  __shared__ double s1[16], s2[16];
  const bool want_hist = feat || hist_out;
  for (uint64_t cb = blockIdx.x; cb < ncb; cb += gridDim.x) {
    if (want_hist)
      for (int i = threadIdx.x; i < 65; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    double sq = 0, sqq = 0;
    const uint32_t *b = cbbox + 8 * cb;
    BpGeom g;
    g.W = W, g.H = H, g.D = D;
    g.b0 = b[0], g.b1 = b[1], g.b2 = b[2], g.b3 = b[3];
    g.cw = b[4] - b[0], g.ch = b[5] - b[1], g.cd = b[6] - b[2];
    const uint32_t ct = b[7] - b[3];
    g.ns = (g.ch + 3) / 4;
    g.ncol = g.cw * g.ns * g.cd * ct;
    g.step = cbsb[3 * cb];
    const uint32_t ncolh = (g.ncol + 1) / 2;
    uint8_t *planes = bp + bpoff[cb];
    // thread per stripe column; with a plane stride of whole words the eight
    // columns of a lane octet (four column pairs) are transposed together and
    // each lane stores four plane words, else each even lane stores its
    // pair's plane bytes. Columns past the last form whole octets or pairs
    // of zero values.
    const bool octets = (ncolh & 3) == 0;
    const uint32_t kend = octets ? (g.ncol + 7) & ~7u : 2 * ncolh;
    // K first: x86_to_u64 is monotone on quotients in [0, 2^64), maps NaN to
    // 2^63 (bit length 64) and larger quotients to 0, so K is the bit length
    // of the largest quotient below 2^64, or 64 with a NaN
    // Division by the step is monotone, so the largest quotient is the
    // thread's largest magnitude divided once; only when that reaches 2^64
    // does the thread look for the largest quotient below it value by value.
    // Two columns per iteration keep eight loads in flight.
    double dmax = 0;
    bool nan = false;
    uint64_t amax = 0;   // largest magnitude bits of the thread's non-NaN values
#pragma unroll 1
    for (uint32_t k = threadIdx.x; k < kend; k += 2 * blockDim.x) {
      uint64_t raw[8];
      bp_load(work, g, k, raw);
      bp_load(work, g, k + blockDim.x, raw + 4);   // past the codeblock: zeros
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const uint64_t a = raw[r] & 0x7FFFFFFFFFFFFFFFull;
        const bool isn = a > 0x7FF0000000000000ull;
        nan |= isn;
        amax = (!isn && a > amax) ? a : amax;
      }
    }
    dmax = __longlong_as_double((long long)amax) / g.step;
    if (!(dmax < 18446744073709551616.0)) {
      dmax = 0;
#pragma unroll 1
      for (uint32_t k = threadIdx.x; k < kend; k += blockDim.x) {
        uint64_t raw[4];
        bp_load(work, g, k, raw);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double d = bp_quot(raw[r], g.step);
          dmax = d < 18446744073709551616.0 && d > dmax ? d : dmax;
        }
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double od = __shfl_xor(dmax, o, 64);
      dmax = od > dmax ? od : dmax;
    }
    nan = __any(nan);
    const uint32_t wv = threadIdx.x / 64, nw = blockDim.x / 64;
    if ((threadIdx.x & 63) == 0) {
      s1[wv] = dmax;
      smax[wv] = nan;
    }
    __syncthreads();
    for (uint32_t w = 0; w < nw; ++w) {
      dmax = s1[w] > dmax ? s1[w] : dmax;
      nan |= smax[w] != 0;
    }
    const uint64_t qmax = x86_to_u64(dmax);
    const int K = __builtin_amdgcn_readfirstlane(nan ? 64 : (qmax ? 64 - __clzll((long long)qmax) : 0));
    __syncthreads();
    // quantisation, statistics and the planes: planes 0..K-1 (at most 62)
    // and the sign plane; planes K..62 are never read
#pragma unroll 1
    for (uint32_t k = threadIdx.x; k < kend; k += blockDim.x) {
      uint64_t raw[4];
      bp_load(work, g, k, raw);
      uint32_t lo[4], hi[4], sg = 0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        sg |= (uint32_t)(raw[r] >> 63) << (3 - r);
        const uint64_t qq = x86_to_u64(bp_quot(raw[r], g.step));
        lo[r] = (uint32_t)qq;
        hi[r] = (uint32_t)(qq >> 32);
        // values outside the codeblock are 0: bit length 0, nothing to the
        // sums; the histogram takes their count back below
        if (want_hist) atomicAdd(&hist[qq ? 64 - __clzll((long long)qq) : 0], 1u);
        if (feat) {
          const double dq = (double)qq;
          sq += dq;
          sqq += dq * dq;
        }
      }
      // sign plane: the pair's byte from its even column's lane
      const uint32_t sgo = __shfl_xor(sg, 1, 64);
      const uint32_t pr = k >> 1;
      if (!(k & 1)) planes[63 * (uint64_t)ncolh + pr] = (uint8_t)(sg | sgo << 4);
      if (octets) {
        const uint32_t c = threadIdx.x & 7;
        uint8_t *grp = planes + ((k >> 1) & ~3u);
        uint32_t m[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) m[3 - r] = lo[r];
        octet_transpose(m);
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if ((int)(4 * c + i) < K) *reinterpret_cast<uint32_t *>(grp + (uint64_t)(4 * c + i) * ncolh) = m[i];
        if (K > 32) {
#pragma unroll
          for (int r = 0; r < 4; ++r) m[3 - r] = hi[r];
          octet_transpose(m);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int pl = 32 + 4 * c + i;
            if (pl < K && pl < 63) *reinterpret_cast<uint32_t *>(grp + (uint64_t)pl * ncolh) = m[i];
          }
        }
      } else {
        for (int pl = 0; pl < K && pl < 63; ++pl) {
          uint32_t v = 0;
          if (pl < 32) {
#pragma unroll
            for (int r = 0; r < 4; ++r) v |= ((lo[r] >> pl) & 1u) << (3 - r);
          } else {
#pragma unroll
            for (int r = 0; r < 4; ++r) v |= ((hi[r] >> (pl - 32)) & 1u) << (3 - r);
          }
          const uint32_t vo = __shfl_xor(v, 1, 64);
          if (!(k & 1)) planes[(uint64_t)pl * ncolh + pr] = (uint8_t)(v | vo << 4);
        }
      }
    }
    const uint32_t nvals = g.cw * g.ch * g.cd * ct;
    if (want_hist && threadIdx.x == 0) atomicSub(&hist[0], 8 * ncolh - nvals);
    for (int o = 32; o > 0; o >>= 1) {
      sq += __shfl_xor(sq, o, 64);
      sqq += __shfl_xor(sqq, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      s1[wv] = sq;
      s2[wv] = sqq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      Kout[cb] = (uint8_t)K;
      if (feat) {
        const uint32_t meta = cbmeta[cb];
        const uint32_t hb = meta & 0xFF, Kmax = (meta >> 8) & 0xFF, res = (meta >> 16) & 0xFF;
        const double nv = (double)nvals;
        for (uint32_t w = 1; w < nw; ++w) {
          s1[0] += s1[w];
          s2[0] += s2[w];
        }
        double sb = 0;
        uint32_t cum[6] = {0, 0, 0, 0, 0, 0};
        // (loops kept rolled: one thread runs them, unrolled they would
        // hold the histogram in registers and set the kernel's register use)
#pragma unroll 1
        for (int nb = 0; nb <= 64; ++nb) sb += (double)nb * hist[nb];
        const int offs[6] = {0, 1, 2, 4, 6, 9};
#pragma unroll 1
        for (int o = 0; o < 6; ++o)
#pragma unroll 1
          for (int nb = K - offs[o] > 1 ? K - offs[o] : 1; nb <= 64; ++nb) cum[o] += hist[nb];
        const double gain = cbsb[3 * cb + 2], stp = cbsb[3 * cb + 1];
        const float lvl_w = (float)(log2(gain) + 2.0 * log2(stp));
        const float lrms = (float)log2(sqrt(s2[0] / nv) + 1.0);
        float *f = feat + 16 * cb;
        f[0] = K / 64.0f;
        f[1] = ((float)Kmax - (float)K) / 16.0f;
        f[2] = res / 8.0f;
        f[3] = __popc(hb) / 4.0f;
        f[4] = (float)(sb / nv) / 64.0f;
        f[5] = (float)log2(s1[0] / nv + 1.0) / 64.0f;
        f[6] = lrms / 64.0f;
        for (int o = 0; o < 6; ++o) f[7 + o] = (float)(cum[o] / nv);
        f[13] = lvl_w / 64.0f;
        f[14] = log2f(rate);
        f[15] = lrms + 0.5f * lvl_w;   // made relative to the tile mean by the network kernel
        atomicAdd(&tstat[0], f[15]);
        atomicAdd(&tstat[1], 1.0f);
      }
    }
    // bit-length histogram (count of values per number of magnitude bits)
    if (hist_out)
      for (int i = threadIdx.x; i < 65; i += blockDim.x) hist_out[65 * cb + i] = hist[i];
    __syncthreads();
  }
}

#include "t1v2.inc"
#include "t1dec.inc"
#include "t1dec2.inc"
#include "t1dec3.inc"

__global__ void zero_f32_kernel(float *p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.f;
}

__global__ void init_flags_kernel(uint32_t *flags) {
  if (threadIdx.x == 0) {
    flags[0] = 0;
    flags[1] = 0xFFFF;
    flags[2] = 0;
    flags[3] = 0;
  }
}

// copy each codeblock's included prefix into one packed buffer
__global__ void gather_kernel(const uint8_t *__restrict__ data, const uint64_t *__restrict__ g, uint64_t ncb,
                              uint8_t *__restrict__ out) {
  for (uint64_t cb = blockIdx.x; cb < ncb; cb += gridDim.x) {
    const uint64_t dst = g[3 * cb], len = g[3 * cb + 1];
    const uint8_t *src = data + g[3 * cb + 2];
    for (uint64_t i = threadIdx.x; i < len; i += blockDim.x) out[dst + i] = src[i];
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static const uint16_t kQe[47] = {0x5601, 0x3401, 0x1801, 0x0AC1, 0x0521, 0x0221, 0x5601, 0x5401, 0x4801, 0x3801,
                                 0x3001, 0x2401, 0x1C01, 0x1601, 0x5601, 0x5401, 0x5101, 0x4801, 0x3801, 0x3401,
                                 0x3001, 0x2801, 0x2401, 0x2201, 0x1C01, 0x1801, 0x1601, 0x1401, 0x1201, 0x1101,
                                 0x0AC1, 0x09C1, 0x08A1, 0x0521, 0x0441, 0x02A1, 0x0221, 0x0141, 0x0111, 0x0085,
                                 0x0049, 0x0025, 0x0015, 0x0009, 0x0005, 0x0001, 0x5601};
static const uint8_t kNmps[47] = {1,  2,  3,  4,  5,  38, 7,  8,  9,  10, 11, 12, 13, 29, 15, 16,
                                  17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32,
                                  33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 45, 46};
static const uint8_t kNlps[47] = {1,  6,  9,  12, 29, 33, 6,  14, 14, 14, 17, 18, 20, 21, 14, 14,
                                  15, 16, 17, 18, 19, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29,
                                  30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 46};
// significance context maps by 15*kh + 5*kv + kd for L(L/LH), H(HL) and HH
// subbands, sign-context tables and the distortion / log tables
// (tier1.c:139-275)
static const uint8_t kSig[3][45] = {
    {0, 1, 2, 2, 2, 3, 3, 3, 3, 3, 4, 4, 4, 4, 4, 5, 6, 6, 6, 6, 7, 7, 7,
     7, 7, 7, 7, 7, 7, 7, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8},
    {0, 1, 2, 2, 2, 5, 6, 6, 6, 6, 8, 8, 8, 8, 8, 3, 3, 3, 3, 3, 7, 7, 7,
     7, 7, 8, 8, 8, 8, 8, 4, 4, 4, 4, 4, 7, 7, 7, 7, 7, 8, 8, 8, 8, 8},
    {0, 3, 6, 8, 8, 1, 4, 7, 8, 8, 2, 5, 7, 8, 8, 1, 4, 7, 8, 8, 2, 5, 7,
     8, 8, 2, 5, 7, 8, 8, 2, 5, 7, 8, 8, 2, 5, 7, 8, 8, 2, 5, 7, 8, 8}};
static const uint8_t kSig2xi[16] = {1, 2, 2, 2, 1, 0, 2, 1, 1, 2, 0, 1, 1, 0, 0, 0};
static const uint8_t kXi2cont[11][2] = {{14, 1}, {13, 1}, {12, 1}, {0, 0}, {11, 1}, {10, 0},
                                        {11, 0}, {0, 0},  {12, 0}, {13, 0}, {14, 0}};
static const int32_t kTm[64] = {49152,  47104,  45056,  43008,  40960,  38912,  36864,  34816,  32768,  30720,
                                28672,  26624,  24576,  22528,  20480,  18432,  16384,  14336,  12288,  10240,
                                8192,   6144,   4096,   2048,   0,      -2048,  -4096,  -6144,  -8192,  -10240,
                                -12288, -14336, -16384, -14336, -12288, -10240, -8192,  -6144,  -4096,  -2048,
                                0,      2048,   4096,   6144,   8192,   10240,  12288,  14336,  16384,  18432,
                                20480,  22528,  24576,  26624,  28672,  30720,  32768,  34816,  36864,  38912,
                                40960,  43008,  45056,  47104};
static const uint16_t kFrac2log[256] = {
    0,   1,   2,   4,   5,   7,   8,   9,   11,  12,  14,  15,  16,  18,  19,  21,  22,  23,  25,  26,  27,  29,
    30,  31,  33,  34,  35,  37,  38,  39,  40,  42,  43,  44,  46,  47,  48,  49,  51,  52,  53,  54,  56,  57,
    58,  59,  61,  62,  63,  64,  65,  67,  68,  69,  70,  71,  73,  74,  75,  76,  77,  78,  80,  81,  82,  83,
    84,  85,  87,  88,  89,  90,  91,  92,  93,  94,  96,  97,  98,  99,  100, 101, 102, 103, 104, 105, 106, 108,
    109, 110, 111, 112, 113, 114, 115, 116, 117, 118, 119, 120, 121, 122, 123, 124, 125, 126, 127, 128, 129, 131,
    132, 133, 134, 135, 136, 137, 138, 139, 140, 140, 141, 142, 143, 144, 145, 146, 147, 148, 149, 150, 151, 152,
    153, 154, 155, 156, 157, 158, 159, 160, 161, 162, 162, 163, 164, 165, 166, 167, 168, 169, 170, 171, 172, 173,
    173, 174, 175, 176, 177, 178, 179, 180, 181, 181, 182, 183, 184, 185, 186, 187, 188, 188, 189, 190, 191, 192,
    193, 194, 194, 195, 196, 197, 198, 199, 200, 200, 201, 202, 203, 204, 205, 205, 206, 207, 208, 209, 209, 210,
    211, 212, 213, 214, 214, 215, 216, 217, 218, 218, 219, 220, 221, 222, 222, 223, 224, 225, 225, 226, 227, 228,
    229, 229, 230, 231, 232, 232, 233, 234, 235, 235, 236, 237, 238, 239, 239, 240, 241, 242, 242, 243, 244, 245,
    245, 246, 247, 247, 248, 249, 250, 250, 251, 252, 253, 253, 254, 255};
static const uint8_t kSwitch0 = 0, kSwitch6 = 6, kSwitch14 = 14;

static int upload_tables() {
  uint32_t mq[47];
  for (int i = 0; i < 47; ++i) {
    const uint32_t sw = (i == kSwitch0 || i == kSwitch6 || i == kSwitch14) ? 1u : 0u;
    mq[i] = kQe[i] | ((uint32_t)kNmps[i] << 16) | ((uint32_t)kNlps[i] << 22) | (sw << 28);
  }
  HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_mq), mq, sizeof mq));
  HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_tm), kTm, sizeof kTm));
  HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_sig), kSig, sizeof kSig));
  HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_sig2xi), kSig2xi, sizeof kSig2xi));
  HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_xi2cont), kXi2cont, sizeof kXi2cont));
  HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_frac2log), kFrac2log, sizeof kFrac2log));
  uint64_t pk[3][3] = {};
  for (int m = 0; m < 3; ++m)
    for (int i = 0; i < 45; ++i) pk[m][i / 16] |= (uint64_t)kSig[m][i] << ((i % 16) * 4);
  HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_sigpk), pk, sizeof pk));
  return 0;
}

Workspace *workspace_of(Job *job);

// Per-codeblock tables of one (tile, param), rebuilt identically for every
// launch that touches it: quantisation steps and gain, hb | Kmax << 8, the
// offsets of its stripe columns, MQ byte buffer (64-byte aligned) and
// bitplanes.
struct CbTables {
  std::vector<double> sb;
  std::vector<uint32_t> meta;
  std::vector<uint64_t> qoff, doff, bpoff;
  uint64_t qo = 0, dof = 0, bo = 0;
};

static void build_tables(const Param &par, CbTables &t) {
  const uint64_t ncb = par.ncb;
  t.sb.assign(3 * ncb, 0);
  t.meta.assign(ncb, 0);
  t.qoff.assign(ncb + 1, 0);
  t.doff.assign(ncb + 1, 0);
  t.bpoff.assign(ncb + 1, 0);
  t.qo = t.dof = t.bo = 0;
  for (uint64_t i = 0; i < ncb; ++i) {
    const uint32_t *b = &par.cb_box[8 * i];
    const Subband &s = par.res[par.cb_sbref[i] >> 8].sb[par.cb_sbref[i] & 0xFF];
    t.sb[3 * i] = s.step_eff;
    t.sb[3 * i + 1] = s.step;
    t.sb[3 * i + 2] = s.gain;
    t.meta[i] = (uint32_t)s.hb | ((uint32_t)s.Kmax << 8) | ((par.cb_sbref[i] >> 8) << 16);
    const uint64_t cw = b[4] - b[0], ch = b[5] - b[1], cd = b[6] - b[2], ct = b[7] - b[3];
    const uint64_t ncol = cw * ((ch + 3) / 4) * cd * ct;
    t.qoff[i] = t.qo;
    t.qo += 4 * ncol;
    t.doff[i] = t.dof;
    t.dof += ((cw * ch * cd * ct) * 10 + 256 + 63) & ~63ull;   // 64-byte aligned buffers
    t.bpoff[i] = t.bo;
    t.bo += 64 * (ncol + 1) / 2 + 64;
  }
  t.qoff[ncb] = t.qo;
  t.doff[ncb] = t.dof;
  t.bpoff[ncb] = t.bo;
}

// symbol stream offsets from the actual K: per plane one significance or
// refinement symbol per value, one sign per value overall, three run
// symbols per stripe column, four segmentation symbols; only the first
// zlim[i] passes when limits are given
static uint64_t stream_offsets(const Param &par, const CbTables &t, const uint8_t *K, const uint8_t *zlim,
                               std::vector<uint64_t> &symoff) {
  const uint64_t ncb = par.ncb;
  symoff.assign(ncb + 1, 0);
  uint64_t so = 0;
  for (uint64_t i = 0; i < ncb; ++i) {
    symoff[i] = so;
    const uint32_t *b = &par.cb_box[8 * i];
    const uint64_t nv = (uint64_t)(b[4] - b[0]) * (b[5] - b[1]) * (b[6] - b[2]) * (b[7] - b[3]);
    const uint64_t ncol = (t.qoff[i + 1] - t.qoff[i]) / 4;
    uint64_t planes = K[i];
    if (zlim && K[i] && zlim[i] < 3u * K[i] - 2) planes = (zlim[i] + 2) / 3 + 1;
    so += (nv * (planes + 1) + 3 * ncol * planes + 4 * planes + 32 + 15) & ~15ull;
  }
  symoff[ncb] = so;
  return so;
}

// The same on the device, with the list of codeblocks to code: one workgroup
// of 1024 threads, each over a contiguous range of codeblocks, then a
// workgroup scan of the ranges' sums. Writes, in codeblock order, list[]
// (codeblocks with K > 0), poff[] (their first coded pass in the packed
// records, as fetch_rows lays them out), symoff[] (every codeblock's symbol
// stream offset, as stream_offsets) and info = {list length, symbol bytes,
// passes, overflow}: bit 0 of overflow when the symbols exceed cap_sym
// bytes, bit 1 when the passes exceed cap_pass; the coding kernels then do
// nothing and the host grows the buffers and launches again.
__device__ __forceinline__ uint32_t dev_passes_coded(uint32_t K, const uint8_t *zlim, uint64_t cb) {
  if (K == 0) return 0;
  const uint32_t full = 3 * K - 2;
  return (zlim && zlim[cb] < full) ? (zlim[cb] > 0 ? zlim[cb] : 1u) : full;
}
__device__ __forceinline__ uint64_t dev_sym_bytes(uint64_t i, uint32_t K, const uint8_t *zlim,
                                                  const uint32_t *cbbox, const uint64_t *qoff) {
  const uint32_t *b = cbbox + 8 * i;
  const uint64_t nv = (uint64_t)(b[4] - b[0]) * (b[5] - b[1]) * (b[6] - b[2]) * (b[7] - b[3]);
  const uint64_t ncol = (qoff[i + 1] - qoff[i]) / 4;
  uint64_t planes = K;
  if (zlim && K && zlim[i] < 3u * K - 2) planes = (zlim[i] + 2) / 3 + 1;
  return (nv * (planes + 1) + 3 * ncol * planes + 4 * planes + 32 + 15) & ~15ull;
}
__global__ void __launch_bounds__(1024) t1_list_kernel(uint64_t ncb, const uint8_t *__restrict__ Kp,
                                                       const uint8_t *__restrict__ zlim,
                                                       const uint32_t *__restrict__ cbbox,
                                                       const uint64_t *__restrict__ qoff, uint64_t cap_sym,
                                                       uint64_t cap_pass, uint64_t *__restrict__ list,
                                                       uint64_t *__restrict__ symoff, uint64_t *__restrict__ poff,
                                                       uint64_t *__restrict__ info) {
  __shared__ uint64_t ssym[1024], scnt[1024], spas[1024];
  const uint32_t t = threadIdx.x;
  const uint64_t a = ncb * t / 1024, e = ncb * (t + 1) / 1024;
  uint64_t sy = 0, cn = 0, pa = 0;
  for (uint64_t i = a; i < e; ++i) {
    const uint32_t K = Kp[i];
    sy += dev_sym_bytes(i, K, zlim, cbbox, qoff);
    if (K) {
      ++cn;
      pa += dev_passes_coded(K, zlim, i);
    }
  }
  ssym[t] = sy, scnt[t] = cn, spas[t] = pa;
  __syncthreads();
  // inclusive scan of the three sums
  for (uint32_t d = 1; d < 1024; d <<= 1) {
    const uint64_t y0 = t >= d ? ssym[t - d] : 0, y1 = t >= d ? scnt[t - d] : 0, y2 = t >= d ? spas[t - d] : 0;
    __syncthreads();
    ssym[t] += y0, scnt[t] += y1, spas[t] += y2;
    __syncthreads();
  }
  uint64_t so = ssym[t] - sy, c = scnt[t] - cn, po = spas[t] - pa;
  for (uint64_t i = a; i < e; ++i) {
    const uint32_t K = Kp[i];
    symoff[i] = so;
    so += dev_sym_bytes(i, K, zlim, cbbox, qoff);
    if (K) {
      list[c] = i;
      poff[c] = po;
      ++c;
      po += dev_passes_coded(K, zlim, i);
    }
  }
  if (t == 1023) {
    symoff[ncb] = ssym[t];
    poff[scnt[t]] = spas[t];
    info[0] = scnt[t];
    info[1] = ssym[t];
    info[2] = spas[t];
    info[3] = (ssym[t] + 64 > cap_sym ? 1u : 0u) | (spas[t] > cap_pass ? 2u : 0u);
  }
}

// LDS of the symbol generator: per codeblock it packs P = 64 / pow2ceil(width)
// slices side by side, each after the first with its own pass-output buffer;
// the buffers take the largest, over the codeblocks, of P - 1 times the
// codeblock's per-slice bound
struct GenParams {
  uint32_t maxcol = 0, obuf_bytes = 0;
  size_t lds = 0;
};
static GenParams gen_params(const Param &par) {
  GenParams g;
  for (uint64_t i = 0; i < par.ncb; ++i) {
    const uint32_t *b = &par.cb_box[8 * i];
    const uint32_t cw = b[4] - b[0], ch = b[5] - b[1], ns = (ch + 3) / 4;
    const uint32_t ncol = cw * ns * (b[6] - b[2]) * (b[7] - b[3]);
    g.maxcol = std::max(g.maxcol, ncol);
    uint32_t gw = 1;
    while (gw < cw) gw <<= 1;
    const uint32_t nsl = (b[6] - b[2]) * (b[7] - b[3]);
    const uint32_t P = gw <= 64 ? std::max(1u, std::min(64 / gw, nsl)) : 1u;
    g.obuf_bytes = std::max(g.obuf_bytes, (P - 1) * (cw * ch * 2 + cw * ns * 3 + 16));
  }
  // column states, the sign plane and the current plane (each to 4 bytes),
  // then the pass-output buffers
  const size_t hq = ((g.maxcol + 1) / 2 + 3) & ~3u;
  g.lds = (((size_t)g.maxcol * 2 + 3) & ~(size_t)3) + 2 * hq + (size_t)g.obuf_bytes;
  return g;
}

// Buffers of symbol generation and MQ coding: the symbol streams (sym_bytes
// bytes), per-codeblock stream offsets, pass ends and distortions, the list.
static int v2_buffers(Workspace *ws, uint64_t ncb, uint64_t sym_bytes, uint64_t nlist) {
  size_t cap = ws->cap_sym;
  CHK(ensure((void **)&ws->d_sym, &cap, sym_bytes));
  ws->cap_sym = cap;
  cap = ws->cap_symoff;
  CHK(ensure((void **)&ws->d_symoff, &cap, (ncb + 1) * sizeof(uint64_t)));
  ws->cap_symoff = cap;
  cap = ws->cap_pc;
  CHK(ensure((void **)&ws->d_pcnt, &cap, (ncb + 1) * kMaxPasses * sizeof(uint32_t)));
  ws->cap_pc = cap;
  cap = ws->cap_pm;
  CHK(ensure((void **)&ws->d_pmse, &cap, (ncb + 1) * kMaxPasses * sizeof(int64_t)));
  ws->cap_pm = cap;
  cap = ws->cap_list;
  CHK(ensure((void **)&ws->d_list, &cap, (nlist + 1) * sizeof(uint64_t)));
  ws->cap_list = cap;
  return 0;
}

// Symbol generation and MQ coding of the codeblocks in ws->d_list: grid for
// `nl` entries; info (device) holds the list's real length when the list
// was built on the device, else null (nl entries). dz: the pass limits on
// the device or null.
static int v2_launch(Workspace *ws, const bwc_codec *c, uint64_t nl, const uint8_t *dz, uint8_t *d_data,
                     const uint8_t *d_bp, const uint32_t *d_hist, const GenParams &gp, const uint64_t *info) {
  if (!nl) return 0;
  const int erres = (int)c->control.error_resilience;
  // codeblocks per wavefront of the MQ kernel (lanes beyond them idle), for
  // complete codings and for codings to a pass limit (tuning.hpp); halved
  // while fewer than 256 wavefronts (one per CU) would run, since lanes of
  // one wavefront share its issue slots and a short list leaves CUs idle
  int lpw = dz ? tuning().mq_lanes_limited : tuning().mq_lanes_full;
  while (lpw > 1 && (nl + lpw - 1) / lpw < 256) lpw >>= 1;
  const uint64_t *dl = ws->d_list;
  hipLaunchKernelGGL(t1_gen_kernel, dim3((uint32_t)nl), dim3(64), gp.lds, ws->stream, dl, ws->d_cbbox, ws->d_cbmeta,
                     ws->d_K, ws->d_bpoff, d_bp, ws->d_symoff, ws->d_sym, ws->d_pcnt, ws->d_pmse, ws->d_flags, erres,
                     gp.maxcol, gp.obuf_bytes, dz, info);
  HIP_OK(hipGetLastError());
#define MQ_LAUNCH(N)                                                                                                 \
  hipLaunchKernelGGL(t1_mq_kernel<N>, dim3((uint32_t)((nl + N - 1) / N)), dim3(64), 0, ws->stream, dl, nl,           \
                     ws->d_cbmeta, ws->d_cbsb, ws->d_K, ws->d_symoff, ws->d_sym, ws->d_pcnt, ws->d_pmse, ws->d_doff, \
                     d_data, ws->d_Kmsbs, ws->d_Z, ws->d_L, ws->d_S, ws->d_snap, ws->d_mse, ws->d_flags, dz, ws->d_U, \
                     d_hist, ws->d_bfut, info)
  switch (lpw) {
    case 1: MQ_LAUNCH(1); break;
    case 2: MQ_LAUNCH(2); break;
    case 3: MQ_LAUNCH(3); break;
    case 4: MQ_LAUNCH(4); break;
    case 5: MQ_LAUNCH(5); break;
    case 6: MQ_LAUNCH(6); break;
    case 8: MQ_LAUNCH(8); break;
    case 32: MQ_LAUNCH(32); break;
    default: MQ_LAUNCH(16); break;
  }
#undef MQ_LAUNCH
  HIP_OK(hipGetLastError());
  return 0;
}

// Symbol generation and MQ coding of the codeblocks in `list` (host); the
// (tile, param) tables, K and bitplane offsets are on the device already.
// zlim: per-codeblock pass limits (host array of ncb) or null for all passes.
static int v2_code(Workspace *ws, const bwc_codec *c, const Param &par, const CbTables &tb, const uint8_t *Kh,
                   const std::vector<uint64_t> &list, const uint8_t *zlim, uint8_t *d_data, const uint8_t *d_bp,
                   const uint32_t *d_hist, const GenParams &gp) {
  const uint64_t ncb = par.ncb;
  std::vector<uint64_t> symoff;
  const uint64_t so = stream_offsets(par, tb, Kh, zlim, symoff);
  CHK(v2_buffers(ws, ncb, so + 64, list.size()));
  CHK(up_copy(ws, ws->d_symoff, symoff.data(), (ncb + 1) * sizeof(uint64_t)));
  if (zlim) CHK(up_copy(ws, ws->d_zlim, zlim, ncb));
  if (list.empty()) return 0;
  CHK(up_copy(ws, ws->d_list, list.data(), list.size() * sizeof(uint64_t)));
  return v2_launch(ws, c, list.size(), zlim ? ws->d_zlim : nullptr, d_data, d_bp, d_hist, gp, nullptr);
}

// passes tier-1 codes for a codeblock with K bitplanes: all 3K - 2, or the
// first zlim[cb] (at least one) when limits are given
static uint32_t passes_coded(int K, const uint8_t *zlim, uint64_t cb) {
  if (K == 0) return 0;
  const uint32_t full = 3 * (uint32_t)K - 2;
  return (zlim && zlim[cb] < full) ? (zlim[cb] > 0 ? zlim[cb] : 1u) : full;
}

// The records of the codeblocks in d_list, packed for one copy to the host:
// per listed codeblock the first m entries of L and mse and m + 1 of S (m =
// passes coded, poff[r + 1] - poff[r]), the bound on its uncoded passes'
// distortion, then its Z, U and Kmsbs.
__global__ void pack_records_kernel(const uint64_t *__restrict__ rows, uint32_t nrows,
                                    const uint64_t *__restrict__ poff, const uint32_t *__restrict__ L,
                                    const uint16_t *__restrict__ S, const double *__restrict__ mse,
                                    const double *__restrict__ bfut, const uint8_t *__restrict__ Z,
                                    const uint8_t *__restrict__ U, const uint8_t *__restrict__ Km,
                                    uint32_t *__restrict__ pL, uint16_t *__restrict__ pS, double *__restrict__ pM,
                                    double *__restrict__ pB, uint8_t *__restrict__ pb,
                                    const uint64_t *__restrict__ info) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (info) {   // a list built on the device (t1_list_kernel)
    if (info[3]) return;
    nrows = (uint32_t)info[0];
  }
  if (r >= nrows) return;
  const uint64_t cb = rows[r], o = poff[r], m = poff[r + 1] - o;
  const uint32_t *l = L + cb * kMaxPasses;
  const uint16_t *sl = S + cb * (kMaxPasses + 1);
  const double *ml = mse + cb * kMaxPasses;
  for (uint64_t j = 0; j < m; ++j) pL[o + j] = l[j];
  for (uint64_t j = 0; j < m; ++j) pM[o + j] = ml[j];
  for (uint64_t j = 0; j <= m; ++j) pS[o + r + j] = sl[j];
  pB[r] = bfut[cb];
  pb[3 * r] = Z[cb];
  pb[3 * r + 1] = U[cb];
  pb[3 * r + 2] = Km[cb];
}

// Byte offsets of the packed records of n rows and tl passes in all (L,
// then S, mse, the bounds, Z/U/Kmsbs), then 16 bytes of flag words and 32
// of list totals.
struct PackLayout {
  size_t oS, oM, oD, oB, oF, total;
};
static PackLayout pack_layout(uint64_t n, uint64_t tl) {
  PackLayout p;
  p.oS = (4 * tl + 255) & ~(size_t)255;
  p.oM = (p.oS + 2 * (tl + n) + 255) & ~(size_t)255;
  p.oD = (p.oM + 8 * tl + 255) & ~(size_t)255;
  p.oB = (p.oD + 8 * n + 255) & ~(size_t)255;
  p.oF = (p.oB + 3 * n + 255) & ~(size_t)255;
  p.total = p.oF + 48;
  return p;
}

// device and page-locked host buffers of a packed-record layout
static int pack_alloc(Workspace *ws, const PackLayout &lo, uint64_t n) {
  size_t cap = ws->cap_poff;
  CHK(ensure((void **)&ws->d_poff, &cap, (n + 1) * sizeof(uint64_t)));
  ws->cap_poff = cap;
  cap = ws->cap_rpack;
  CHK(ensure((void **)&ws->d_rpack, &cap, lo.total));
  ws->cap_rpack = cap;
  if (lo.total > ws->cap_down) {
    if (ws->h_down) HIP_OK(hipHostFree(ws->h_down));
    ws->h_down = nullptr;
    ws->cap_down = 0;
    HIP_OK(hipHostMalloc((void **)&ws->h_down, lo.total));
    ws->cap_down = lo.total;
  }
  return 0;
}

static void pack_launch(Workspace *ws, const PackLayout &lo, uint64_t ngrid, const uint64_t *info) {
  hipLaunchKernelGGL(pack_records_kernel, dim3((uint32_t)((ngrid + 255) / 256)), dim3(256), 0, ws->stream,
                     ws->d_list, (uint32_t)ngrid, ws->d_poff, ws->d_L, ws->d_S, ws->d_mse, ws->d_bfut, ws->d_Z,
                     ws->d_U, ws->d_Kmsbs, (uint32_t *)ws->d_rpack, (uint16_t *)(ws->d_rpack + lo.oS),
                     (double *)(ws->d_rpack + lo.oM), (double *)(ws->d_rpack + lo.oD), ws->d_rpack + lo.oB, info);
}

// The packed records in h_down (layout lo) into rec: rows[r] coded to m[r]
// passes, from pass offset poff[r].
static int unpack_rows(Workspace *ws, CblkRecords &rec, const std::vector<uint64_t> &rows,
                       const std::vector<uint32_t> &m, const std::vector<uint64_t> &poff, const PackLayout &lo,
                       double tmf) {
  const uint64_t n = rows.size();
  const uint32_t *pL = (const uint32_t *)ws->h_down;
  const uint16_t *pS = (const uint16_t *)(ws->h_down + lo.oS);
  const double *pM = (const double *)(ws->h_down + lo.oM), *pD = (const double *)(ws->h_down + lo.oD);
  const uint8_t *pb = ws->h_down + lo.oB;
  if (rec.mse.size() != rec.n * kMaxPasses) rec.mse.assign(rec.n * kMaxPasses, 0.0);
  if (rec.bfut.size() != rec.n) rec.bfut.assign(rec.n, 0.0);
  // into the per-codeblock record arrays (strided by the pass maximum, so
  // mostly cache and TLB misses): in chunks on the rate-control pool
  std::atomic<uint64_t> bad{UINT64_MAX};
  const size_t nch = (size_t)std::min<uint64_t>(64, (n + 511) / 512);
  const std::function<void(size_t)> unpack = [&](size_t ch) {
    for (uint64_t r = n * ch / nch, e = n * (ch + 1) / nch; r < e; ++r) {
      const uint64_t cb = rows[r];
      if (pb[3 * r] != m[r]) {
        uint64_t cur = bad.load();
        while (r < cur && !bad.compare_exchange_weak(cur, r)) {
        }
        continue;
      }
      rec.Z[cb] = pb[3 * r];
      rec.U[cb] = pb[3 * r + 1];
      rec.Kmsbs[cb] = pb[3 * r + 2];
      std::memcpy(&rec.L[cb * kMaxPasses], pL + poff[r], m[r] * sizeof(uint32_t));
      std::memcpy(&rec.S[cb * (kMaxPasses + 1)], pS + poff[r] + r, (m[r] + 1) * sizeof(uint16_t));
      std::memcpy(&rec.mse[cb * kMaxPasses], pM + poff[r], m[r] * sizeof(double));
      rec.bfut[cb] = pD[r];
    }
  };
  const int threads = rc_thread_count();
  if (nch > 1 && threads > 1) {
    host_pool_run(threads - 1, nch, unpack);
  } else {
    for (size_t ch = 0; ch < nch; ++ch) unpack(ch);
  }
  if (bad.load() != UINT64_MAX) {
    const uint64_t r = bad.load();
    std::snprintf(g_err, sizeof g_err, "codeblock %llu: %u passes coded, %u expected", (unsigned long long)rows[r],
                  (unsigned)pb[3 * r], m[r]);
    return 1;
  }
  trace_mark("records: unpack", tmf);
  return 0;
}

static int batch_join(Workspace *ws, const T1Field *f, size_t lds, bool limited);

// page-locked room for K, the pass limits, Z and Kmsbs of ncb codeblocks and
// 16 bytes of flag words
static int kz_room(Workspace *ws, uint64_t ncb) {
  const size_t need = 4 * ncb + 16;
  if (need <= ws->cap_kz) return 0;
  if (ws->h_kz) HIP_OK(hipHostFree(ws->h_kz));
  ws->h_kz = nullptr;
  ws->cap_kz = 0;
  HIP_OK(hipHostMalloc((void **)&ws->h_kz, need));
  ws->cap_kz = need;
  return 0;
}

// Z and Kmsbs of the listed rows (coded to m[r] passes) from the device
// arrays' copies Zh and Kh; a row whose Z is not m[r] is an error.
static int take_z_kmsbs(CblkRecords &rec, const std::vector<uint64_t> &rows, const std::vector<uint32_t> &m,
                        const uint8_t *Zh, const uint8_t *Kh) {
  for (size_t r = 0; r < rows.size(); ++r) {
    const uint64_t cb = rows[r];
    if (Zh[cb] != m[r]) {
      std::snprintf(g_err, sizeof g_err, "codeblock %llu: %u passes coded, %u expected", (unsigned long long)cb,
                    (unsigned)Zh[cb], m[r]);
      return 1;
    }
    rec.Z[cb] = Zh[cb];
    rec.Kmsbs[cb] = Kh[cb];
  }
  return 0;
}

// Bring the records of `rows` (the host list v2_code launched, in d_list),
// coded to m[r] passes, into rec, together with the flag words (hfl); ends
// with a stream synchronisation. The offsets go up through the open upload
// staging.
static int fetch_rows(Workspace *ws, CblkRecords &rec, const std::vector<uint64_t> &rows,
                      const std::vector<uint32_t> &m, uint32_t hfl[4], bool records) {
  const uint64_t n = rows.size();
  if (!records) {
    // the records stay on the device (rate control runs there): Z and Kmsbs
    // of the rows, and the flag words
    const uint64_t ncb = rec.n;
    CHK(kz_room(ws, ncb));
    HIP_OK(hipMemcpyAsync(ws->h_kz + 2 * ncb, ws->d_Z, ncb, hipMemcpyDeviceToHost, ws->stream));
    HIP_OK(hipMemcpyAsync(ws->h_kz + 3 * ncb, ws->d_Kmsbs, ncb, hipMemcpyDeviceToHost, ws->stream));
    HIP_OK(hipMemcpyAsync(ws->h_kz + 4 * ncb, ws->d_flags, 16, hipMemcpyDeviceToHost, ws->stream));
    HIP_OK(hipStreamSynchronize(ws->stream));
    std::memcpy(hfl, ws->h_kz + 4 * ncb, 16);
    if (hfl[0]) return 0;
    return take_z_kmsbs(rec, rows, m, ws->h_kz + 2 * ncb, ws->h_kz + 3 * ncb);
  }
  std::vector<uint64_t> poff(n + 1, 0);
  for (uint64_t r = 0; r < n; ++r) poff[r + 1] = poff[r] + m[r];
  const PackLayout lo = pack_layout(n, poff[n]);
  CHK(pack_alloc(ws, lo, n));
  if (n) {
    CHK(up_copy(ws, ws->d_poff, poff.data(), (n + 1) * sizeof(uint64_t)));
    pack_launch(ws, lo, n, nullptr);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(ws->h_down, ws->d_rpack, lo.oB + 3 * n, hipMemcpyDeviceToHost, ws->stream));
  }
  HIP_OK(hipMemcpyAsync(ws->h_down + lo.oF, ws->d_flags, 16, hipMemcpyDeviceToHost, ws->stream));
  double tmf = trace_mark(nullptr, 0);
  HIP_OK(hipStreamSynchronize(ws->stream));
  tmf = trace_mark("records: wait for the GPU", tmf);
  std::memcpy(hfl, ws->h_down + lo.oF, 16);
  if (hfl[0]) return 0;
  return unpack_rows(ws, rec, rows, m, poff, lo, tmf);
}

// Tier-1 coding of every codeblock with K > 0 with the list, symbol-stream
// offsets and packed-record offsets built on the device from K and the pass
// limits dz (device, or null): the host launches the whole chain without
// waiting and waits once, for the totals; records then come back by section.
// When the symbol or record buffers (kept from earlier calls) are too small
// the coding kernels skip their work; the buffers grow to the totals and the
// chain runs again (once). On return rec.K, zh (the limits, when dz), list
// and m are on the host and the records unpacked.
static int t1_code_dev(Workspace *ws, const bwc_codec *c, const Param &par, const uint8_t *dz, uint8_t *d_data,
                       const uint8_t *d_bp, const uint32_t *d_hist, const GenParams &gp, CblkRecords &rec,
                       std::vector<uint8_t> &zh, std::vector<uint64_t> &list, std::vector<uint32_t> &m,
                       uint32_t hfl[4], bool records, const std::function<int()> *pre, uint64_t pre_key) {
  const uint64_t ncb = par.ncb;
  if (!ws->d_t1info) HIP_OK(hipMalloc((void **)&ws->d_t1info, 4 * sizeof(uint64_t)));
  CHK(kz_room(ws, ncb));
  uint64_t info[4] = {0, 0, 0, 0};
  PackLayout lo{};
  double tm = trace_mark(nullptr, 0);
  CHK(v2_buffers(ws, ncb, ws->cap_sym, ncb));
  lo = pack_layout(ncb, ws->cap_pass);
  CHK(pack_alloc(ws, lo, ncb));
  // the device chain: what comes before it (DWT, bitplanes, pass limits),
  // the list, generation, MQ coding, [record packing], and the copies of the
  // totals, flags, K, the pass limits, Z and Kmsbs to the host
  auto chain = [&](bool with_pre, bool queued) -> int {
    if (with_pre && pre) CHK((*pre)());
    hipLaunchKernelGGL(t1_list_kernel, dim3(1), dim3(1024), 0, ws->stream, ncb, ws->d_K, dz, ws->d_cbbox,
                       ws->d_qoff, (uint64_t)ws->cap_sym, ws->cap_pass, ws->d_list, ws->d_symoff, ws->d_poff,
                       ws->d_t1info);
    HIP_OK(hipGetLastError());
    if (queued) {
      // generation and MQ coding through the batch's global codeblock queue
      T1Field f{ws->d_cbbox, ws->d_cbmeta, ws->d_cbsb,  ws->d_K,     ws->d_bpoff, d_bp,      ws->d_symoff,
                ws->d_sym,   ws->d_pcnt,   ws->d_pmse,  ws->d_flags, dz,          ws->d_doff, d_data,
                ws->d_Kmsbs, ws->d_Z,      ws->d_U,     ws->d_L,     ws->d_S,     ws->d_snap, ws->d_mse,
                ws->d_bfut,  d_hist,       ws->d_t1info, ws->d_list, (int)c->control.error_resilience,
                gp.maxcol,   gp.obuf_bytes};
      CHK(batch_join(ws, &f, gp.lds, dz != nullptr));
    } else {
      CHK(v2_launch(ws, c, ncb, dz, d_data, d_bp, d_hist, gp, ws->d_t1info));
    }
    if (records) {
      pack_launch(ws, lo, ncb, ws->d_t1info);
      HIP_OK(hipGetLastError());
    }
    HIP_OK(hipMemcpyAsync(ws->h_down + lo.oF, ws->d_flags, 16, hipMemcpyDeviceToHost, ws->stream));
    HIP_OK(hipMemcpyAsync(ws->h_down + lo.oF + 16, ws->d_t1info, 32, hipMemcpyDeviceToHost, ws->stream));
    HIP_OK(hipMemcpyAsync(ws->h_kz, ws->d_K, ncb, hipMemcpyDeviceToHost, ws->stream));
    if (dz) HIP_OK(hipMemcpyAsync(ws->h_kz + ncb, dz, ncb, hipMemcpyDeviceToHost, ws->stream));
    if (!records) {
      HIP_OK(hipMemcpyAsync(ws->h_kz + 2 * ncb, ws->d_Z, ncb, hipMemcpyDeviceToHost, ws->stream));
      HIP_OK(hipMemcpyAsync(ws->h_kz + 3 * ncb, ws->d_Kmsbs, ncb, hipMemcpyDeviceToHost, ws->stream));
    }
    return 0;
  };
  if (pre_key) {
    // everything the chain's launches depend on
    uint64_t h = pre_key ^ 0x9E3779B97F4A7C15ull;
    auto mix = [&h](uint64_t v) { h = (h ^ v) * 0x100000001B3ull + 0x9E3779B97F4A7C15ull; };
    const void *ptrs[] = {d_data, d_bp, d_hist, dz, ws->d_K, ws->d_Kmsbs, ws->d_Z, ws->d_U, ws->d_L, ws->d_S,
                          ws->d_mse, ws->d_bfut, ws->d_cbbox, ws->d_cbsb, ws->d_cbmeta, ws->d_qoff, ws->d_doff,
                          ws->d_bpoff, ws->d_snap, ws->d_flags, ws->d_t1info, ws->d_list, ws->d_symoff, ws->d_sym,
                          ws->d_pcnt, ws->d_pmse, ws->d_poff, ws->d_rpack, ws->h_down, ws->h_kz, ws->d_feat,
                          ws->d_tstat, ws->d_skipw, ws->d_zlim};
    for (const void *p : ptrs) mix((uint64_t)(uintptr_t)p);
    mix(ncb);
    mix(ws->cap_sym);
    mix(ws->cap_pass);
    mix(lo.oF | (uint64_t)records << 62);
    mix((uint64_t)gp.lds | (uint64_t)gp.maxcol << 32);
    mix(gp.obuf_bytes);
    mix((uint64_t)tuning().mq_lanes_full | (uint64_t)tuning().mq_lanes_limited << 16 |
        (uint64_t)c->control.error_resilience << 32);
    CHK(graph_run(ws, h, [&]() { return chain(true, false); }));
  } else {
    CHK(chain(true, ws->batch && !ws->batch_joined));
  }
  tm = trace_mark("list + offsets + launch", tm);
  for (int attempt = 0;; ++attempt) {
    HIP_OK(hipStreamSynchronize(ws->stream));
    tm = trace_mark("records: wait for the GPU", tm);
    std::memcpy(info, ws->h_down + lo.oF + 16, 32);
    if (!info[3]) break;
    if (attempt > 0) {
      std::snprintf(g_err, sizeof g_err, "tier-1 buffers too small after growing them");
      return 1;
    }
    // the coding kernels did nothing: grow the buffers to the totals and run
    // the chain from the list again (directly)
    CHK(v2_buffers(ws, ncb, info[1] + 64, ncb));
    ws->cap_pass = std::max<uint64_t>(ws->cap_pass, info[2]);
    lo = pack_layout(ncb, ws->cap_pass);
    CHK(pack_alloc(ws, lo, ncb));
    CHK(chain(false, false));
  }
  std::memcpy(hfl, ws->h_down + lo.oF, 16);
  std::memcpy(rec.K.data(), ws->h_kz, ncb);
  if (dz) zh.assign(ws->h_kz + ncb, ws->h_kz + 2 * ncb);
  const uint8_t *zl = dz ? zh.data() : nullptr;
  // the same list and offsets on the host, for unpacking
  list.clear();
  m.clear();
  for (uint64_t i = 0; i < ncb; ++i)
    if (rec.K[i]) {
      list.push_back(i);
      m.push_back(passes_coded(rec.K[i], zl, i));
    }
  const uint64_t n = list.size();
  std::vector<uint64_t> poff(n + 1, 0);
  for (uint64_t r = 0; r < n; ++r) poff[r + 1] = poff[r] + m[r];
  if (n != info[0] || poff[n] != info[2]) {
    std::snprintf(g_err, sizeof g_err, "device codeblock list (%llu, %llu passes) differs from the host's",
                  (unsigned long long)info[0], (unsigned long long)info[2]);
    return 1;
  }
  if (hfl[0] || !n) return 0;
  if (!records) return take_z_kmsbs(rec, list, m, ws->h_kz + 2 * ncb, ws->h_kz + 3 * ncb);
  const uint64_t tl = poff[n];
  HIP_OK(hipMemcpyAsync(ws->h_down, ws->d_rpack, 4 * tl, hipMemcpyDeviceToHost, ws->stream));
  HIP_OK(hipMemcpyAsync(ws->h_down + lo.oS, ws->d_rpack + lo.oS, 2 * (tl + n), hipMemcpyDeviceToHost, ws->stream));
  HIP_OK(hipMemcpyAsync(ws->h_down + lo.oM, ws->d_rpack + lo.oM, 8 * tl, hipMemcpyDeviceToHost, ws->stream));
  HIP_OK(hipMemcpyAsync(ws->h_down + lo.oD, ws->d_rpack + lo.oD, 8 * n, hipMemcpyDeviceToHost, ws->stream));
  HIP_OK(hipMemcpyAsync(ws->h_down + lo.oB, ws->d_rpack + lo.oB, 3 * n, hipMemcpyDeviceToHost, ws->stream));
  HIP_OK(hipStreamSynchronize(ws->stream));
  tm = trace_mark("records: copy", tm);
  return unpack_rows(ws, rec, list, m, poff, lo, tm);
}

int t1_encode(Workspace *ws, const bwc_codec *c, const Param &par, const double *work, CblkRecords &rec,
              uint16_t *slope_min, uint16_t *slope_max, int slot, const PassLimits &lim) {
  const uint8_t *zlim = lim.fixed;
  std::vector<uint8_t> zpred;
  const uint64_t ncb = par.ncb;
  double tm0 = trace_mark(nullptr, 0);
  rec.resize(ncb, c->control.nLayers);
  if (!ncb) return 0;
  CbTables tb;
  build_tables(par, tb);
  tm0 = trace_mark("records + tables", tm0);
  const std::vector<double> &sb = tb.sb;
  const std::vector<uint32_t> &meta = tb.meta;
  const std::vector<uint64_t> &qoff = tb.qoff, &doff = tb.doff;
  const uint64_t qo = tb.qo, dof = tb.dof;
  size_t cap;
  if (ncb > ws->cap_cb) {
    void **tabs[] = {(void **)&ws->d_cbbox, (void **)&ws->d_cbsb, (void **)&ws->d_cbmeta, (void **)&ws->d_qoff,
                     (void **)&ws->d_doff,  (void **)&ws->d_snap, (void **)&ws->d_zlim};
    for (void **p : tabs)
      if (*p) {
        HIP_OK(hipFree(*p));
        *p = nullptr;
      }
    const uint64_t n = ncb + 1;
    HIP_OK(hipMalloc((void **)&ws->d_zlim, n));
    HIP_OK(hipMalloc((void **)&ws->d_cbbox, 8 * n * sizeof(uint32_t)));
    HIP_OK(hipMalloc((void **)&ws->d_cbsb, 3 * n * sizeof(double)));
    HIP_OK(hipMalloc((void **)&ws->d_cbmeta, n * sizeof(uint32_t)));
    HIP_OK(hipMalloc((void **)&ws->d_qoff, n * sizeof(uint64_t)));
    HIP_OK(hipMalloc((void **)&ws->d_doff, n * sizeof(uint64_t)));
    HIP_OK(hipMalloc((void **)&ws->d_snap, n * 2 * kMaxPasses * sizeof(uint64_t)));
    ws->cap_cb = ncb;
  }
  CHK(bind_slot(ws, slot, ncb));
  cap = ws->cap_datap[slot];
  CHK(ensure((void **)&ws->d_datap[slot], &cap, dof));
  ws->cap_datap[slot] = cap;
  uint8_t *const d_data = ws->d_datap[slot];
  rec.slot = slot;
  rec.src_off = doff;
  if (!ws->d_flags) HIP_OK(hipMalloc((void **)&ws->d_flags, 16));
  CHK(up_begin(ws, kUpBytesPerCb * (ncb + 1) + kSkipModelSize * sizeof(float) + 16 * 256));
  CHK(up_copy(ws, ws->d_cbbox, par.cb_box.data(), 8 * ncb * sizeof(uint32_t)));
  CHK(up_copy(ws, ws->d_cbsb, sb.data(), 3 * ncb * sizeof(double)));
  CHK(up_copy(ws, ws->d_cbmeta, meta.data(), ncb * sizeof(uint32_t)));
  CHK(up_copy(ws, ws->d_qoff, qoff.data(), (ncb + 1) * sizeof(uint64_t)));
  CHK(up_copy(ws, ws->d_doff, doff.data(), (ncb + 1) * sizeof(uint64_t)));
  const uint64_t W = par.box.len(0), H = par.box.len(1), D = par.box.len(2);

  // fast path: wavefront-per-codeblock symbol generation + few-lane MQ, when
  // every codeblock's state fits LDS (rows wider than a wavefront are coded
  // in 64-column chunks)
  const GenParams gp = gen_params(par);
  const char *force = std::getenv("BWC_AMD_T1");
  const bool v2 = gp.lds <= 64 * 1024 && !(force && force[0] == '1');
  const int erres = (int)c->control.error_resilience;
  rec.U.assign(ncb, 0);
  rec.Zfull.assign(ncb, 0);
  uint32_t hfl[4] = {0, 0, 0, 0};
  if (v2) {
    if (slot >= (int)ws->bp_slots.size()) ws->bp_slots.resize(slot + 1);
    Workspace::BpSlot &bs = ws->bp_slots[slot];
    cap = bs.cap_bp;
    CHK(ensure((void **)&bs.bp, &cap, tb.bo));
    bs.cap_bp = cap;
    cap = ws->cap_bpoff;
    CHK(ensure((void **)&ws->d_bpoff, &cap, (ncb + 1) * sizeof(uint64_t)));
    ws->cap_bpoff = cap;
    CHK(up_copy(ws, ws->d_bpoff, tb.bpoff.data(), (ncb + 1) * sizeof(uint64_t)));
    float *d_feat = nullptr;
    if (lim.model || lim.features) {
      cap = ws->cap_feat;
      CHK(ensure((void **)&ws->d_feat, &cap, (ncb + 1) * 16 * sizeof(float)));
      ws->cap_feat = cap;
      if (!ws->d_tstat) HIP_OK(hipMalloc((void **)&ws->d_tstat, 4 * sizeof(float)));
      if (!ws->d_skipw) HIP_OK(hipMalloc((void **)&ws->d_skipw, kSkipModelSize * sizeof(float)));
      if (lim.model) CHK(up_copy(ws, ws->d_skipw, lim.model, kSkipModelSize * sizeof(float)));
      d_feat = ws->d_feat;
    }
    const int nl = c->control.nLayers;
    const float rate = nl > 0 && c->control.bitrate ? c->control.bitrate[nl - 1] : 1.0f;
    // codings to a pass limit need the bit-length histogram for the bound on
    // the distortion of the passes left out (kept with the bitplanes)
    uint32_t *d_hist = nullptr;
    if (lim.model || lim.fixed) {
      cap = bs.cap_hist;
      CHK(ensure((void **)&bs.hist, &cap, (ncb + 1) * 65 * sizeof(uint32_t)));
      bs.cap_hist = cap;
      d_hist = bs.hist;
    }
    // pass limits on the device: the network's or the fixed ones
    const uint8_t *dz = nullptr;
    if (lim.model) {
      dz = ws->d_zlim;
    } else if (lim.fixed) {
      CHK(up_copy(ws, ws->d_zlim, lim.fixed, ncb));
      dz = ws->d_zlim;
    }
    tm0 = trace_mark("uploads + buffers", tm0);
    // the device chain up to the codeblock list: the caller's launches (the
    // DWT), the flag words, bitplanes and features, the network's limits
    const float rate_arg = rate > 0 ? rate : 1.0f;
    uint8_t *const bp = bs.bp;
    const std::function<int()> pre = [&]() -> int {
      if (lim.pre) CHK((*lim.pre)());
      hipLaunchKernelGGL(init_flags_kernel, dim3(1), dim3(64), 0, ws->stream, ws->d_flags);
      if (d_feat) hipLaunchKernelGGL(zero_f32_kernel, dim3(1), dim3(64), 0, ws->stream, ws->d_tstat, 4);
      hipLaunchKernelGGL(bitplane_kernel, dim3((uint32_t)std::min<uint64_t>(ncb, 65535)), dim3(256), 0, ws->stream,
                         work, W, H, D, ws->d_cbbox, ws->d_cbsb, ws->d_bpoff, bp, ws->d_K, ncb, ws->d_cbmeta,
                         d_feat, ws->d_tstat, rate_arg, d_hist);
      HIP_OK(hipGetLastError());
      if (lim.model) {
        char e[200];
        if (skipnet_infer(ws->stream, ws->d_feat, ncb, ws->d_skipw, ws->d_tstat, ws->d_K, lim.margin, ws->d_zlim,
                          nullptr, e, sizeof e)) {
          std::snprintf(g_err, sizeof g_err, "%s", e);
          return 1;
        }
      }
      return 0;
    };
    // with a key from the caller (pipeline graphs), the key of this part
    uint64_t key = 0;
    if (lim.pre_key && !lim.features) {
      key = lim.pre_key;
      auto mix = [&key](uint64_t v) { key = (key ^ v) * 0x100000001B3ull + 0x9E3779B97F4A7C15ull; };
      mix((uint64_t)(uintptr_t)work);
      mix(W), mix(H), mix(D);
      uint32_t rb;
      std::memcpy(&rb, &rate_arg, 4);
      mix(rb | (uint64_t)(lim.model != nullptr) << 32 | (uint64_t)(lim.fixed != nullptr) << 33 |
          (uint64_t)(uint32_t)lim.margin << 34);
    }
    // list, offsets, coding and record packing on the device, one wait
    std::vector<uint64_t> list;
    std::vector<uint32_t> m;
    CHK(t1_code_dev(ws, c, par, dz, d_data, bs.bp, d_hist, gp, rec, zpred, list, m, hfl, lim.records, &pre, key));
    if (lim.model) zlim = zpred.data();
    if (lim.out && zlim) std::memcpy(lim.out, zlim, ncb);
    if (lim.features) {
      // decisions per codeblock: the symbol position where its last coded
      // pass ends (measurement only)
      std::vector<uint32_t> pc((size_t)ncb * kMaxPasses);
      HIP_OK(hipMemcpyAsync(pc.data(), ws->d_pcnt, pc.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, ws->stream));
      HIP_OK(hipStreamSynchronize(ws->stream));
      rec.decisions.assign(ncb, 0);
      for (size_t r = 0; r < list.size(); ++r)
        if (m[r]) rec.decisions[list[r]] = pc[list[r] * kMaxPasses + m[r] - 1];
    }
  } else {
    if (ws->batch && !ws->batch_joined) batch_leave(ws);
    if (lim.pre) CHK((*lim.pre)());
    hipLaunchKernelGGL(init_flags_kernel, dim3(1), dim3(64), 0, ws->stream, ws->d_flags);
    cap = ws->cap_q;
    CHK(ensure((void **)&ws->d_q, &cap, qo * sizeof(uint64_t)));
    ws->cap_q = cap;
    cap = ws->cap_state;
    CHK(ensure((void **)&ws->d_state, &cap, (qo / 4 + 1) * sizeof(uint16_t)));
    ws->cap_state = cap;
    hipLaunchKernelGGL(quantize_kernel, dim3((uint32_t)std::min<uint64_t>(ncb, 65535)), dim3(256), 0, ws->stream,
                       work, W, H, D, ws->d_cbbox, ws->d_cbsb, ws->d_qoff, ws->d_q, ws->d_state, ws->d_K, ncb);
    HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(t1_encode_kernel, dim3((uint32_t)((ncb + 63) / 64)), dim3(64), 0, ws->stream, ncb,
                       ws->d_cbbox, ws->d_cbsb, ws->d_cbmeta, ws->d_qoff, ws->d_doff, ws->d_q, ws->d_state,
                       d_data, ws->d_K, ws->d_Kmsbs, ws->d_Z, ws->d_L, ws->d_S, ws->d_snap, ws->d_mse,
                       ws->d_flags, erres);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(rec.K.data(), ws->d_K, ncb, hipMemcpyDeviceToHost, ws->stream));
    HIP_OK(hipMemcpyAsync(rec.Kmsbs.data(), ws->d_Kmsbs, ncb, hipMemcpyDeviceToHost, ws->stream));
    HIP_OK(hipMemcpyAsync(rec.Z.data(), ws->d_Z, ncb, hipMemcpyDeviceToHost, ws->stream));
    HIP_OK(hipMemcpyAsync(rec.L.data(), ws->d_L, ncb * kMaxPasses * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          ws->stream));
    HIP_OK(hipMemcpyAsync(rec.S.data(), ws->d_S, ncb * (kMaxPasses + 1) * sizeof(uint16_t), hipMemcpyDeviceToHost,
                          ws->stream));
    HIP_OK(hipMemcpyAsync(hfl, ws->d_flags, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, ws->stream));
    HIP_OK(hipStreamSynchronize(ws->stream));
  }
  if (hfl[0]) {
    std::snprintf(g_err, sizeof g_err, "tier-1 output exceeded a per-codeblock buffer (flags %u)", hfl[0]);
    return 1;
  }
  for (uint64_t i = 0; i < ncb; ++i) {
    rec.Zfull[i] = rec.K[i] ? (uint8_t)(3 * rec.K[i] - 2) : 0;
    if (!v2) rec.U[i] = rec.Z[i];
    // codeblocks without significant bits are never launched on the fast path
    if (rec.K[i] == 0) {
      rec.Z[i] = 0;
      rec.U[i] = 0;
      rec.Kmsbs[i] = 0;
      rec.S[i * (kMaxPasses + 1)] = 0xFFFF;
    }
  }
  if (hfl[1] < *slope_min) *slope_min = (uint16_t)hfl[1];
  if (hfl[2] > *slope_max) *slope_max = (uint16_t)hfl[2];
  return 0;
}

// Re-code the codeblocks in `cbs` of slot `slot` (the parameter coded last
// into that slot by t1_encode, bitplanes still on the device) to the pass
// limits in `zlim` (ncb entries; those of other codeblocks are ignored), and
// refresh their records.
int t1_extend(Workspace *ws, const bwc_codec *c, const Param &par, CblkRecords &rec, int slot,
              const std::vector<uint64_t> &cbs, const uint8_t *zlim, bool records) {
  const uint64_t ncb = par.ncb;
  if (cbs.empty()) return 0;
  if (slot < 0 || slot >= (int)ws->bp_slots.size() || !ws->bp_slots[slot].bp || ncb > ws->cap_cb ||
      slot >= (int)ws->rec_slots.size() || ws->rec_slots[slot].cap < ncb + 1) {
    std::snprintf(g_err, sizeof g_err, "no bitplanes kept for parameter slot %d", slot);
    return 1;
  }
  CHK(bind_slot(ws, slot, ncb));
  CbTables tb;
  build_tables(par, tb);
  const GenParams gp = gen_params(par);
  CHK(up_begin(ws, kUpBytesPerCb * (ncb + 1) + 16 * 256));
  CHK(up_copy(ws, ws->d_cbbox, par.cb_box.data(), 8 * ncb * sizeof(uint32_t)));
  CHK(up_copy(ws, ws->d_cbsb, tb.sb.data(), 3 * ncb * sizeof(double)));
  CHK(up_copy(ws, ws->d_cbmeta, tb.meta.data(), ncb * sizeof(uint32_t)));
  CHK(up_copy(ws, ws->d_doff, tb.doff.data(), (ncb + 1) * sizeof(uint64_t)));
  CHK(up_copy(ws, ws->d_bpoff, tb.bpoff.data(), (ncb + 1) * sizeof(uint64_t)));
  CHK(up_copy(ws, ws->d_K, rec.K.data(), ncb));
  hipLaunchKernelGGL(init_flags_kernel, dim3(1), dim3(64), 0, ws->stream, ws->d_flags);
  CHK(v2_code(ws, c, par, tb, rec.K.data(), cbs, zlim, ws->d_datap[slot], ws->bp_slots[slot].bp,
              ws->bp_slots[slot].hist, gp));
  // the re-coded rows: Z, U, Kmsbs, L, S (the device arrays hold other
  // parameters' rows as well; only the listed rows reach the records)
  std::vector<uint32_t> m(cbs.size());
  for (size_t r = 0; r < cbs.size(); ++r) m[r] = passes_coded(rec.K[cbs[r]], zlim, cbs[r]);
  uint32_t hfl[4];
  CHK(fetch_rows(ws, rec, cbs, m, hfl, records));
  if (hfl[0]) {
    std::snprintf(g_err, sizeof g_err, "tier-1 output exceeded a per-codeblock buffer (flags %u)", hfl[0]);
    return 1;
  }
  return 0;
}

int gather(Workspace *ws, CblkRecords &rec, const std::vector<uint32_t> &len) {
  const uint64_t ncb = rec.n;
  std::vector<uint64_t> g(3 * ncb);
  uint64_t tot = 0;
  for (uint64_t i = 0; i < ncb; ++i) {
    rec.data_off[i] = tot;
    g[3 * i] = tot;
    g[3 * i + 1] = len[i];
    g[3 * i + 2] = rec.src_off[i];
    tot += len[i];
  }
  rec.data.assign(tot ? tot : 1, 0);
  if (!ncb || !tot) return 0;
  size_t cap = ws->cap_packed;
  CHK(ensure((void **)&ws->d_packed, &cap, tot));
  ws->cap_packed = cap;
  cap = ws->cap_gather;
  CHK(ensure((void **)&ws->d_gather, &cap, 3 * ncb * sizeof(uint64_t)));
  ws->cap_gather = cap;
  HIP_OK(hipMemcpyAsync(ws->d_gather, g.data(), 3 * ncb * sizeof(uint64_t), hipMemcpyHostToDevice, ws->stream));
  hipLaunchKernelGGL(gather_kernel, dim3((uint32_t)std::min<uint64_t>(ncb, 65535)), dim3(256), 0, ws->stream,
                     ws->d_datap[rec.slot], ws->d_gather, ncb, ws->d_packed);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(rec.data.data(), ws->d_packed, tot, hipMemcpyDeviceToHost, ws->stream));
  HIP_OK(hipStreamSynchronize(ws->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// the global codeblock queue of a batch
// ---------------------------------------------------------------------------
struct BatchQueue {
  std::mutex m;
  std::condition_variable cv;
  int expected = 0, joined = 0, err = 0, slots = 1;
  bool launched = false;
  std::vector<T1Field> fields;
  std::vector<hipEvent_t> ready;   // each field's list built
  size_t lds = 0;
  bool limited = false;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  uint64_t max_items = 0;
  T1Field *d_fields = nullptr;
  uint64_t *d_items = nullptr;
  uint32_t *d_misc = nullptr;   // [0] items, [1..2] queue counters, [4..67] bucket counts, [68..131] offsets
};

// work buckets: bucket of a codeblock = floor(log2) of its symbol-stream
// bound (largest first in the queue)
__global__ void batch_count_kernel(const T1Field *__restrict__ F, uint32_t *misc) {
  const T1Field &f = F[blockIdx.y];
  if (f.info[3]) return;
  const uint64_t n = f.info[0];
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t cb = f.list[r];
    const uint64_t sz = f.symoff[cb + 1] - f.symoff[cb];
    atomicAdd(&misc[4 + (63 - __clzll(sz | 1))], 1u);
  }
}
__global__ void batch_offsets_kernel(uint32_t *misc) {
  uint32_t o = 0;
  for (int b = 63; b >= 0; --b) {
    misc[68 + b] = o;
    o += misc[4 + b];
  }
  misc[0] = o;
  misc[1] = 0;
  misc[2] = 0;
}
__global__ void batch_fill_kernel(const T1Field *__restrict__ F, uint32_t *misc, uint64_t *items) {
  const T1Field &f = F[blockIdx.y];
  if (f.info[3]) return;
  const uint64_t n = f.info[0];
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t cb = f.list[r];
    const uint64_t sz = f.symoff[cb + 1] - f.symoff[cb];
    const uint32_t k = atomicAdd(&misc[68 + (63 - __clzll(sz | 1))], 1u);
    items[k] = (uint64_t)blockIdx.y << 32 | cb;
  }
}
__global__ void batch_zero_kernel(uint32_t *misc, int n) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) misc[k] = 0;
}

BatchQueue *batch_create(int fields, uint64_t max_items, int slots) {
  BatchQueue *q = new BatchQueue();
  q->expected = fields;
  q->slots = std::max(1, slots);
  q->max_items = max_items;
  (void)hipGetDevice(&q->device);
  if (hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&q->done, hipEventDisableTiming) != hipSuccess ||
      hipMalloc((void **)&q->d_fields, std::max(1, fields) * sizeof(T1Field)) != hipSuccess ||
      hipMalloc((void **)&q->d_items, std::max<uint64_t>(1, max_items) * sizeof(uint64_t)) != hipSuccess ||
      hipMalloc((void **)&q->d_misc, 256 * sizeof(uint32_t)) != hipSuccess) {
    std::snprintf(g_err, sizeof g_err, "batch queue: %s", hipGetErrorString(hipGetLastError()));
    (void)batch_destroy(q);
    return nullptr;
  }
  return q;
}

// the last field to join: the queue and the two launches on the batch stream
static int batch_launch(BatchQueue *q) {
  const int nf = (int)q->fields.size();
  HIP_OK(hipSetDevice(q->device));
  for (hipEvent_t e : q->ready) HIP_OK(hipStreamWaitEvent(q->stream, e, 0));
  if (!nf) return 0;
  HIP_OK(hipMemcpyAsync(q->d_fields, q->fields.data(), nf * sizeof(T1Field), hipMemcpyHostToDevice, q->stream));
  hipLaunchKernelGGL(batch_zero_kernel, dim3(1), dim3(256), 0, q->stream, q->d_misc, 256);
  hipLaunchKernelGGL(batch_count_kernel, dim3(64, nf), dim3(256), 0, q->stream, q->d_fields, q->d_misc);
  hipLaunchKernelGGL(batch_offsets_kernel, dim3(1), dim3(1), 0, q->stream, q->d_misc);
  hipLaunchKernelGGL(batch_fill_kernel, dim3(64, nf), dim3(256), 0, q->stream, q->d_fields, q->d_misc, q->d_items);
  HIP_OK(hipGetLastError());
  // persistent grids: symbol generation as many workgroups as fit (the
  // largest per-codeblock LDS of the fields), MQ coding four wavefronts per
  // SIMD with the lanes per wavefront of the codings' kind
  const uint32_t per_cu = (uint32_t)std::max<size_t>(1, std::min<size_t>(16, (160 * 1024) / std::max<size_t>(1, q->lds)));
  const uint32_t ggen = (uint32_t)std::min<uint64_t>(q->max_items, 256ull * per_cu);
  hipLaunchKernelGGL(t1_gen_queue_kernel, dim3(std::max(1u, ggen)), dim3(64), q->lds, q->stream, q->d_fields,
                     q->d_items, q->d_misc, q->d_misc + 1);
  HIP_OK(hipGetLastError());
  int lpw = q->limited ? tuning().mq_lanes_limited : tuning().mq_lanes_full;
#define MQQ(N)                                                                                                   \
  hipLaunchKernelGGL(t1_mq_queue_kernel<N>,                                                                       \
                     dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((q->max_items + N - 1) / N, 4096))), \
                     dim3(64), 0, q->stream, q->d_fields, q->d_items, q->d_misc, q->d_misc + 1)
  switch (lpw) {
    case 1: MQQ(1); break;
    case 2: MQQ(2); break;
    case 3: MQQ(3); break;
    case 4: MQQ(4); break;
    case 5: MQQ(5); break;
    case 6: MQQ(6); break;
    case 8: MQQ(8); break;
    case 32: MQQ(32); break;
    default: MQQ(16); break;
  }
#undef MQQ
  HIP_OK(hipGetLastError());
  return 0;
}

static int batch_arrive(Workspace *ws, const T1Field *f, size_t lds, bool limited) {
  BatchQueue *q = ws->batch;
  ws->batch_joined = true;
  hipEvent_t e = nullptr;
  if (f) {
    HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_OK(hipEventRecord(e, ws->stream));
  }
  int err = 0;
  {
    std::unique_lock<std::mutex> lk(q->m);
    if (f) {
      q->fields.push_back(*f);
      q->ready.push_back(e);
      q->lds = std::max(q->lds, lds);
      q->limited |= limited;
    }
    if (++q->joined == q->expected) {
      if (batch_launch(q)) q->err = 1;
      if (hipEventRecord(q->done, q->stream) != hipSuccess) q->err = 1;
      q->launched = true;
      q->cv.notify_all();
    } else {
      q->cv.wait(lk, [q] { return q->launched; });
    }
    err = q->err;
    if (f) {
      q->cv.wait(lk, [q] { return q->slots > 0; });
      --q->slots;
      ws->batch_slot = true;
    }
  }
  if (err) {
    std::snprintf(g_err, sizeof g_err, "batch queue launch failed");
    return 1;
  }
  if (f) {
    HIP_OK(hipStreamWaitEvent(ws->stream, q->done, 0));
    ++ws->batch_queued;
  }
  return 0;
}

static int batch_join(Workspace *ws, const T1Field *f, size_t lds, bool limited) {
  return batch_arrive(ws, f, lds, limited);
}

void batch_leave(Workspace *ws) {
  if (ws->batch && !ws->batch_joined) (void)batch_arrive(ws, nullptr, 0, false);
  if (ws->batch && ws->batch_slot) {
    std::lock_guard<std::mutex> lk(ws->batch->m);
    ++ws->batch->slots;
    ws->batch_slot = false;
    ws->batch->cv.notify_all();
  }
}

int batch_destroy(BatchQueue *q) {
  if (!q) return 0;
  int rc = q->err;
  if (q->stream && hipStreamSynchronize(q->stream) != hipSuccess) rc = 1;
  for (hipEvent_t e : q->ready)
    if (e) (void)hipEventDestroy(e);
  if (q->done) (void)hipEventDestroy(q->done);
  if (q->stream) (void)hipStreamDestroy(q->stream);
  if (q->d_fields) (void)hipFree(q->d_fields);
  if (q->d_items) (void)hipFree(q->d_items);
  if (q->d_misc) (void)hipFree(q->d_misc);
  delete q;
  return rc;
}

// ---------------------------------------------------------------------------
// decompression: tier-1 decoding of one (tile, param) into the zeroed tile
// buffer, then (after the inverse DWT) denormalisation into the output
// ---------------------------------------------------------------------------
int t1_decode(Workspace *ws, const bwc_codec *c, const Param &par, const DecodeRecords &rec, double *work) {
  const uint64_t ncb = par.ncb, n = par.box.volume();
  HIP_OK(hipMemsetAsync(work, 0, n * sizeof(double), ws->stream));
  if (!ncb) return 0;
  if (rec.n != ncb) {
    std::snprintf(g_err, sizeof g_err, "decode records do not match the codec's codeblocks");
    return 1;
  }
  CbTables tb;
  build_tables(par, tb);
  std::vector<uint8_t> K(ncb);
  std::vector<uint64_t> soff(ncb + 1);
  for (uint64_t i = 0; i < ncb; ++i) {
    const Subband &s = par.res[par.cb_sbref[i] >> 8].sb[par.cb_sbref[i] & 0xFF];
    K[i] = (uint8_t)(s.Kmax - rec.Kmsbs[i]);   // uint8 arithmetic, as tier1.c:3316
    soff[i] = tb.qoff[i] / 4;
  }
  soff[ncb] = tb.qo / 4;
  auto up = [&](void **p, size_t *cap, const void *src, size_t bytes) -> int {
    CHK(ensure(p, cap, bytes));
    HIP_OK(hipMemcpyAsync(*p, src, bytes, hipMemcpyHostToDevice, ws->stream));
    return 0;
  };
  CHK(up((void **)&ws->dec.box, &ws->dec.cap_box, par.cb_box.data(), 8 * ncb * sizeof(uint32_t)));
  CHK(up((void **)&ws->dec.sb, &ws->dec.cap_sb, tb.sb.data(), 3 * ncb * sizeof(double)));
  CHK(up((void **)&ws->dec.meta, &ws->dec.cap_meta, tb.meta.data(), ncb * sizeof(uint32_t)));
  CHK(up((void **)&ws->dec.K, &ws->dec.cap_K, K.data(), ncb));
  CHK(up((void **)&ws->dec.Z, &ws->dec.cap_Z, rec.Z.data(), ncb));
  CHK(up((void **)&ws->dec.len, &ws->dec.cap_len, rec.len.data(), ncb * sizeof(uint32_t)));
  CHK(up((void **)&ws->dec.off, &ws->dec.cap_off, rec.off.data(), (ncb + 1) * sizeof(uint64_t)));
  CHK(up((void **)&ws->dec.data, &ws->dec.cap_data, rec.data.get(), rec.data_size));
  CHK(up((void **)&ws->dec.soff, &ws->dec.cap_soff, soff.data(), (ncb + 1) * sizeof(uint64_t)));
  CHK(ensure((void **)&ws->dec.state, &ws->dec.cap_state, (soff[ncb] + 1) * sizeof(uint16_t)));
  const uint64_t W = par.box.len(0), H = par.box.len(1), D = par.box.len(2);
  // codeblocks per wavefront: each lane is a chain of dependent decisions,
  // so more wavefronts with fewer lanes hide more of the latency (512^3,
  // 2 bpv: 1 -> 137 ms, 2 -> 134, 4 -> 155, 8 -> 187, 16 -> 205 ms per
  // decompression)
  uint32_t ncolmax = 0;
  for (uint64_t i = 0; i < ncb; ++i) ncolmax = std::max<uint32_t>(ncolmax, (uint32_t)(soff[i + 1] - soff[i]));
  if (trace_enabled()) {
    const double tm = trace_mark(nullptr, 0);
    HIP_OK(hipStreamSynchronize(ws->stream));
    trace_mark("dec: uploads wait", tm);
  }
  // the scalar decoder (one codeblock per wavefront, state in registers)
  // when every codeblock is at most 64 columns wide and has at most 256
  // stripe rows (stripes x slices)
  uint32_t cwmax = 0, nrowmax = 0;
  for (uint64_t i = 0; i < ncb; ++i) {
    const uint32_t *bx = par.cb_box.data() + 8 * i;
    cwmax = std::max(cwmax, bx[4] - bx[0]);
    nrowmax = std::max(nrowmax, ((bx[5] - bx[1] + 3) / 4) * (bx[6] - bx[2]) * (bx[7] - bx[3]));
  }
  const char *force_sc = std::getenv("BWC_AMD_DEC_SC");
  if (cwmax <= 64 && nrowmax <= 256 && !(force_sc && force_sc[0] == '0')) {
    const uint32_t *d32 = reinterpret_cast<const uint32_t *>(ws->dec.data);
#define SC_LAUNCH(W64, NB)                                                                                          \
  hipLaunchKernelGGL((t1_decode_sc_kernel<W64, NB>), dim3((uint32_t)ncb), dim3(64), 0, ws->stream, ncb, ws->dec.box,  \
                     ws->dec.sb, ws->dec.meta, ws->dec.K, ws->dec.Z, ws->dec.len, ws->dec.off, d32, W, H, D, work,    \
                     (int)c->control.error_resilience)
    const int nb = nrowmax <= 64 ? 1 : (nrowmax <= 128 ? 2 : 4);
    if (cwmax <= 32) {
      if (nb == 1) SC_LAUNCH(false, 1);
      else if (nb == 2) SC_LAUNCH(false, 2);
      else SC_LAUNCH(false, 4);
    } else {
      if (nb == 1) SC_LAUNCH(true, 1);
      else if (nb == 2) SC_LAUNCH(true, 2);
      else SC_LAUNCH(true, 4);
    }
#undef SC_LAUNCH
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(ws->stream));
    return 0;
  }
  // the LDS-state decoder when every lane's column states fit 32 KB
  int lanes = 2;
  if (const char *e = std::getenv("BWC_AMD_DEC_LANES")) lanes = std::max(1, std::min(4, std::atoi(e)));
  if (lanes == 3) lanes = 2;
  const size_t lds = (size_t)lanes * ncolmax * sizeof(uint16_t);
  const char *force = std::getenv("BWC_AMD_DEC_LDS");
  if (lds <= 32 * 1024 && !(force && force[0] == '0')) {
#define DEC_LAUNCH(N)                                                                                              \
  hipLaunchKernelGGL(t1_decode_lds_kernel<N>, dim3((uint32_t)((ncb + N - 1) / N)), dim3(64), lds, ws->stream, ncb, \
                     ws->dec.box, ws->dec.sb, ws->dec.meta, ws->dec.K, ws->dec.Z, ws->dec.len, ws->dec.off,          \
                     ws->dec.data, ncolmax, W, H, D, work, (int)c->control.error_resilience)
    if (lanes == 1) DEC_LAUNCH(1);
    else if (lanes == 4) DEC_LAUNCH(4);
    else DEC_LAUNCH(2);
#undef DEC_LAUNCH
  } else {
    constexpr int LPW = 2;
    hipLaunchKernelGGL(t1_decode_kernel<LPW>, dim3((uint32_t)((ncb + LPW - 1) / LPW)), dim3(64), 0, ws->stream, ncb,
                       ws->dec.box, ws->dec.sb, ws->dec.meta, ws->dec.K, ws->dec.Z, ws->dec.len, ws->dec.off,
                       ws->dec.data, ws->dec.soff, ws->dec.state, W, H, D, work, (int)c->control.error_resilience);
  }
  HIP_OK(hipGetLastError());
  // the host records are read by the uploads above before this returns
  HIP_OK(hipStreamSynchronize(ws->stream));
  return 0;
}

int decode_output(Workspace *ws, const bwc_codec *c, const Tile &tl, int p, const Param &par, const double *work,
                  void *out) {
  const uint64_t n = par.box.volume();
  if (!n) return 0;
  const double alpha = (par.pmax + par.pmin) / 2.0f, beta = (par.pmax - par.pmin) / 2.0f;
  const bwc_gl_inf &inf = c->info;
  const uint64_t poff = inf.nX * inf.nY * inf.nZ * inf.nTS * (uint64_t)p;
  const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(denorm_flush_kernel, dim3((uint32_t)blocks), dim3(256), 0, ws->stream, work, par.box.len(0),
                     par.box.len(1), par.box.len(2), n, alpha, beta, (uint64_t)c->control.Qm << 52, inf.nX, inf.nY,
                     inf.nZ, tl.box.lo[0], tl.box.lo[1], tl.box.lo[2], poff, (int)inf.data_prec, out);
  HIP_OK(hipGetLastError());
  return 0;
}

// Copy the samples the decoder wrote for (tile, param) from the device
// staging buffer `stage` (the whole field's layout) to the caller's host
// buffer; samples no tile writes keep what the caller had there, as with the
// reference's flush_buffer (libbwc.c:827-978). The time range starts at 0, as
// flush_buffer indexes it.
int output_box_d2h(Workspace *ws, const bwc_codec *c, const Tile &tl, int p, const Param &par, const void *stage,
                   void *out) {
  const bwc_gl_inf &inf = c->info;
  const uint64_t nX = inf.nX, nY = inf.nY, nZ = inf.nZ;
  const uint64_t prec = (uint64_t)inf.data_prec;
  const uint64_t w = par.box.len(0), h = par.box.len(1), d = par.box.len(2), dt = par.box.len(3);
  if (!w || !h || !d || !dt) return 0;
  const uint64_t poff = nX * nY * nZ * inf.nTS * (uint64_t)p;
  const uint64_t x0 = tl.box.lo[0], y0 = tl.box.lo[1], z0 = tl.box.lo[2];
  const char *src = static_cast<const char *>(stage);
  char *dst = static_cast<char *>(out);
  if (w == nX && h == nY) {
    const bool whole = d == nZ;
    for (uint64_t t = 0; t < (whole ? 1 : dt); ++t) {
      const uint64_t o = (poff + nX * nY * (z0 + nZ * t)) * prec;
      HIP_OK(hipMemcpyAsync(dst + o, src + o, (whole ? w * h * d * dt : w * h * d) * prec, hipMemcpyDeviceToHost,
                            ws->stream));
    }
    return 0;
  }
  for (uint64_t t = 0; t < dt; ++t) {
    const uint64_t base = (poff + nX * nY * nZ * t) * prec;
    hipMemcpy3DParms cp = {};
    cp.srcPtr = make_hipPitchedPtr(const_cast<char *>(src) + base, nX * prec, nX, nY);
    cp.srcPos = make_hipPos(x0 * prec, y0, z0);
    cp.dstPtr = make_hipPitchedPtr(dst + base, nX * prec, nX, nY);
    cp.dstPos = make_hipPos(x0 * prec, y0, z0);
    cp.extent = make_hipExtent(w * prec, h, d);
    cp.kind = hipMemcpyDeviceToHost;
    HIP_OK(hipMemcpy3DAsync(&cp, ws->stream));
  }
  return 0;
}

int ensure_output(Workspace *ws, size_t bytes, void **out) {
  CHK(ensure((void **)&ws->dec.out, &ws->dec.cap_out, bytes));
  *out = ws->dec.out;
  return 0;
}

// chunk k: desc[3k] = offset in the body buffer, desc[3k + 1] = length,
// desc[3k + 2] = source address
__global__ void gather_abs_kernel(const uint64_t *__restrict__ desc, uint64_t n, uint8_t *__restrict__ out) {
  for (uint64_t k = blockIdx.x; k < n; k += gridDim.x) {
    const uint64_t dst = desc[3 * k], len = desc[3 * k + 1];
    const uint8_t *src = reinterpret_cast<const uint8_t *>(desc[3 * k + 2]);
    for (uint64_t i = threadIdx.x; i < len; i += blockDim.x) out[dst + i] = src[i];
  }
}

int gather_bodies(Workspace *ws, const std::vector<uint64_t> &desc, const std::vector<int> &slots, uint64_t end) {
  const uint64_t n = desc.size() / 3;
  if (slots.size() != n) {
    std::snprintf(g_err, sizeof g_err, "gather_bodies: one slot per chunk expected");
    return 1;
  }
  if (end > ws->cap_bodies) {
    // grow, keeping what earlier tiles placed
    uint8_t *nb = nullptr;
    const size_t cap = std::max<size_t>(end + end / 4, 1 << 20);
    HIP_OK(hipMalloc((void **)&nb, cap));
    if (ws->d_bodies) {
      HIP_OK(hipMemcpyAsync(nb, ws->d_bodies, ws->cap_bodies, hipMemcpyDeviceToDevice, ws->stream));
      HIP_OK(hipStreamSynchronize(ws->stream));
      HIP_OK(hipFree(ws->d_bodies));
    }
    ws->d_bodies = nb;
    ws->cap_bodies = cap;
  }
  if (!n) return 0;
  std::vector<uint64_t> g(desc);
  CHK(up_begin(ws, g.size() * sizeof(uint64_t) + 256));
  for (uint64_t k = 0; k < n; ++k) {
    const int s = slots[k];
    if (s < 0 || s > 255 || !ws->d_datap[s] || g[3 * k + 2] + g[3 * k + 1] > ws->cap_datap[s] ||
        g[3 * k] + g[3 * k + 1] > end) {
      std::snprintf(g_err, sizeof g_err, "gather_bodies: chunk %llu out of range", (unsigned long long)k);
      return 1;
    }
    g[3 * k + 2] += (uint64_t)(uintptr_t)ws->d_datap[s];
  }
  size_t cap = ws->cap_gather;
  CHK(ensure((void **)&ws->d_gather, &cap, g.size() * sizeof(uint64_t)));
  ws->cap_gather = cap;
  // the descriptors go through the page-locked staging, so nothing here
  // waits for the copy or the kernel (later work on the stream follows them)
  CHK(up_copy(ws, ws->d_gather, g.data(), g.size() * sizeof(uint64_t)));
  hipLaunchKernelGGL(gather_abs_kernel, dim3((uint32_t)std::min<uint64_t>(n, 65535)), dim3(256), 0, ws->stream,
                     ws->d_gather, n, ws->d_bodies);
  HIP_OK(hipGetLastError());
  return 0;
}

int body_d2h(Workspace *ws, uint8_t *dst, uint64_t off, uint64_t n) {
  if (off + n > ws->cap_bodies) {
    std::snprintf(g_err, sizeof g_err, "body_d2h: range outside the body buffer");
    return 1;
  }
  HIP_OK(hipMemcpyAsync(dst, ws->d_bodies + off, n, hipMemcpyDeviceToHost, ws->stream));
  return 0;
}

int sync(Workspace *ws) {
  HIP_OK(hipStreamSynchronize(ws->stream));
  return 0;
}

int init_device(Workspace *ws) {
  if (ws->device >= 0) return 0;
  int dev = 0, n = 0;
  HIP_OK(hipGetDeviceCount(&n));
  if (n <= 0) {
    std::snprintf(g_err, sizeof g_err, "no HIP device");
    return 1;
  }
  HIP_OK(hipGetDevice(&dev));
  ws->device = dev;
  HIP_OK(hipStreamCreateWithFlags(&ws->stream, hipStreamNonBlocking));
  for (auto &e : ws->ev) HIP_OK(hipEventCreate(&e));
  return process_setup();
}

// constant tables and kernel attributes, once per process (workspaces may be
// created from several host threads)
static int process_setup() {
  static std::mutex m;
  static bool tables = false;
  std::lock_guard<std::mutex> lk(m);
  if (!tables) {
    if (upload_tables()) return 1;
    const int lds_max = 160 * 1024;
    HIP_OK(hipFuncSetAttribute((const void *)dwt_axis_kernel<0, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
    HIP_OK(hipFuncSetAttribute((const void *)dwt_axis_kernel<0, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
    HIP_OK(hipFuncSetAttribute((const void *)dwt_axis_kernel<1, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
    HIP_OK(hipFuncSetAttribute((const void *)dwt_axis_kernel<1, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
    HIP_OK(hipFuncSetAttribute((const void *)dwt97_x_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
    HIP_OK(hipFuncSetAttribute((const void *)idwt97_x_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
#define XY_ATTR(NT)                                                                                                 \
  HIP_OK(hipFuncSetAttribute((const void *)dwt97_xy_kernel<4, NT, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                             lds_max));                                                                          \
  HIP_OK(hipFuncSetAttribute((const void *)dwt97_xy_kernel<4, NT, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                             lds_max));                                                                          \
  HIP_OK(hipFuncSetAttribute((const void *)dwt97_xy_kernel<4, NT, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                             lds_max));
    XY_ATTR(0) XY_ATTR(1) XY_ATTR(2) XY_ATTR(3)
#undef XY_ATTR
    tables = true;
  }
  return 0;
}

static void ws_delete(void *p) { delete static_cast<Workspace *>(p); }

Workspace *workspace_of(Job *job) {
  if (!job->ws) {
    auto *ws = new Workspace();
    if (init_device(ws)) {
      delete ws;
      return nullptr;
    }
    job->ws = ws;
    job->ws_free = ws_delete;
  }
  return static_cast<Workspace *>(job->ws);
}

int ensure_work(Workspace *ws, uint64_t n) {
  size_t cap = ws->work_cap;
  if (ensure((void **)&ws->d_work, &cap, n * sizeof(double))) return 1;
  ws->work_cap = cap;
  return 0;
}

int record(Workspace *ws, int i) {
  HIP_OK(hipEventRecord(ws->ev[i], ws->stream));
  return 0;
}

double elapsed(Workspace *ws, int a, int b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, ws->ev[a], ws->ev[b]) != hipSuccess) return -1;
  return ms;
}

int last_features(Workspace *ws, uint64_t ncb, float *out) {
  if (!ws->d_feat || !ws->d_tstat || ws->cap_feat < ncb * 16 * sizeof(float)) {
    std::snprintf(g_err, sizeof g_err, "no codeblock features were computed");
    return 1;
  }
  float ts[4];
  HIP_OK(hipMemcpyAsync(out, ws->d_feat, ncb * 16 * sizeof(float), hipMemcpyDeviceToHost, ws->stream));
  HIP_OK(hipMemcpyAsync(ts, ws->d_tstat, sizeof ts, hipMemcpyDeviceToHost, ws->stream));
  HIP_OK(hipStreamSynchronize(ws->stream));
  const float mean = ts[1] > 0.f ? ts[0] / ts[1] : 0.f;
  for (uint64_t i = 0; i < ncb; ++i) out[16 * i + 15] -= mean;
  return 0;
}

int read_norm(Workspace *ws, double mm[4]) {
  if (!ws->h_pinned) HIP_OK(hipHostMalloc((void **)&ws->h_pinned, 64, hipHostMallocDefault));
  HIP_OK(hipMemcpyAsync(ws->h_pinned, ws->d_norm, 4 * sizeof(double), hipMemcpyDeviceToHost, ws->stream));
  HIP_OK(hipStreamSynchronize(ws->stream));
  std::memcpy(mm, ws->h_pinned, 4 * sizeof(double));
  return 0;
}

void *host_alloc(size_t bytes) {
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    std::snprintf(g_err, sizeof g_err, "hipHostMalloc of %zu bytes failed", bytes);
    return nullptr;
  }
  return p;
}

void host_free(void *p) {
  if (p) (void)hipHostFree(p);
}

int is_device_pointer(const void *p) {
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return attr.type == hipMemoryTypeDevice ? 1 : 0;
}

int copy_h2d(Workspace *ws, void *dst, const void *src, size_t n) {
  HIP_OK(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, ws->stream));
  HIP_OK(hipStreamSynchronize(ws->stream));
  return 0;
}

int copy_d2h(Workspace *ws, void *dst, const void *src, size_t n) {
  HIP_OK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, ws->stream));
  HIP_OK(hipStreamSynchronize(ws->stream));
  return 0;
}

int device_count() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

}  // namespace gpu
}  // namespace bwc
