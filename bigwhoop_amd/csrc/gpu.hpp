// Host-side interface of the gfx950 kernels (gpu.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <vector>

#include "codec.hpp"

namespace bwc {
namespace gpu {

// Device buffers owned by one codec; grown on demand, freed with the codec.
struct Workspace {
  int device = -1;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {};
  // host fields: the tile's box is staged into d_input in chunks on
  // copy_stream, one event per chunk; ev_fill orders a fill after the last
  // reader of the staging buffer
  hipStream_t copy_stream = nullptr;
  std::vector<hipEvent_t> ev_copy;
  hipEvent_t ev_fill = nullptr;
  // pageable host fields: page-locked bounce buffers the host threads copy
  // chunks into (several at once) while earlier chunks cross PCIe; ev_stage
  // marks the end of the last copy out of each
  static constexpr int kStageBufs = 4;
  uint8_t *h_stage[kStageBufs] = {};
  size_t cap_stage = 0;
  hipEvent_t ev_stage[kStageBufs] = {};
  // field staging (host input copied here) and the tile working buffer (f64)
  void *d_input = nullptr;
  size_t input_cap = 0;
  double *d_work = nullptr;
  size_t work_cap = 0;
  double *d_work2 = nullptr;      // second buffer for out-of-place DWT passes
  size_t work2_cap = 0;
  // reductions
  double *d_red = nullptr;        // partial min/max
  double *d_norm = nullptr;       // {min, max, alpha, beta}
  // codeblock tables + tier-1 outputs for one (tile, param)
  uint32_t *d_cbbox = nullptr;    // 8 per codeblock
  double *d_cbsb = nullptr;       // step_eff, step, gain per codeblock
  uint32_t *d_cbmeta = nullptr;   // hb | Kmax << 8 per codeblock
  uint64_t *d_qoff = nullptr;     // offset of each codeblock's stripe columns
  uint64_t *d_doff = nullptr;     // offset of each codeblock's byte buffer
  uint64_t *d_q = nullptr;        // quantised magnitudes, 4 per stripe column
  uint16_t *d_state = nullptr;    // sigma | pi << 4 | delta << 8 | xi << 12
  uint8_t *d_data = nullptr;      // unused; MQ bytes live in one slot per parameter:
  uint8_t *d_datap[256] = {};     // slot p holds parameter p's codeblock byte buffers until
  size_t cap_datap[256] = {};     // rate control has chosen what to gather
  // tier-1 records of the parameter slot being coded (aliases of one
  // rec_slots entry, see bind_slot): K, Kmsbs, Z, U per codeblock, L[190],
  // S[191], mse[190] per codeblock, the bound on the uncoded passes'
  // distortion
  uint8_t *d_K = nullptr, *d_Kmsbs = nullptr, *d_Z = nullptr;
  uint32_t *d_L = nullptr;
  uint16_t *d_S = nullptr;
  uint64_t *d_snap = nullptr;     // 2 words per pass per codeblock
  double *d_mse = nullptr;
  uint32_t *d_flags = nullptr;    // [0] overflow, [1] slope_min, [2] slope_max
  uint8_t *d_packed = nullptr;
  uint64_t *d_gather = nullptr;
  uint8_t *d_bodies = nullptr;    // packet bodies of the codestream being written
  size_t cap_bodies = 0;
  // decompression: codeblock tables and records of one (tile, param), the
  // decoder's stripe-column state, the output staging buffer
  struct {
    uint32_t *box = nullptr, *meta = nullptr, *len = nullptr;
    double *sb = nullptr;
    uint8_t *K = nullptr, *Z = nullptr, *data = nullptr;
    uint64_t *off = nullptr, *soff = nullptr;
    uint16_t *state = nullptr;
    void *out = nullptr;
    size_t cap_box = 0, cap_meta = 0, cap_len = 0, cap_sb = 0, cap_K = 0, cap_Z = 0, cap_data = 0, cap_off = 0,
           cap_soff = 0, cap_state = 0, cap_out = 0;
  } dec;
  size_t cap_cb = 0, cap_q = 0, cap_data = 0, cap_packed = 0, cap_gather = 0, cap_state = 0;
  // fast tier-1 path: bitplanes (kept per parameter slot until the next
  // compression, so rate control can have codeblocks coded further),
  // symbol streams, per-pass counts/distortion, pass limits, first
  // unsettled pass per codeblock
  struct BpSlot {
    uint8_t *bp = nullptr;
    size_t cap_bp = 0;
    uint32_t *hist = nullptr;     // bit-length histogram, 65 per codeblock
    size_t cap_hist = 0;
  };
  std::vector<BpSlot> bp_slots;
  // tier-1 records per parameter slot, kept on the device until rate
  // control (which runs on the device over every parameter of a tile) is done
  struct RecSlot {
    uint8_t *K = nullptr, *Kmsbs = nullptr, *Z = nullptr, *U = nullptr;
    uint32_t *L = nullptr;
    uint16_t *S = nullptr;
    double *mse = nullptr, *bfut = nullptr;
    size_t cap = 0;   // codeblocks
  };
  std::vector<RecSlot> rec_slots;
  // device rate control (pcrd.hip): its buffers and per-tile tables
  void *rc = nullptr;
  // the batch whose global codeblock queue codes this workspace's next
  // tier-1 call (bwc_amd_compress_batch), and whether the call has joined it
  struct BatchQueue *batch = nullptr;
  bool batch_joined = false, batch_slot = false;
  uint64_t batch_queued = 0;   // tier-1 calls coded through a batch's queue
  uint8_t *d_U = nullptr, *d_zlim = nullptr;
  // pass-limit network: features (16 per codeblock), tile statistics,
  // weights on the device
  float *d_feat = nullptr, *d_tstat = nullptr, *d_skipw = nullptr;
  size_t cap_feat = 0;
  uint8_t *d_bp = nullptr;
  uint64_t *d_bpoff = nullptr;
  uint8_t *d_sym = nullptr;
  uint64_t *d_symoff = nullptr;
  uint32_t *d_pcnt = nullptr;
  int64_t *d_pmse = nullptr;
  uint64_t *d_list = nullptr;
  size_t cap_bp = 0, cap_bpoff = 0, cap_sym = 0, cap_symoff = 0, cap_pc = 0, cap_pm = 0, cap_list = 0;
  uint64_t *h_pinned = nullptr;   // small pinned staging for flags / min-max
  // page-locked staging of one tier-1 call: the uploads (codeblock tables,
  // offsets, lists) are then asynchronous DMA that the host does not wait
  // for, and the packed records come back the same way; ev_up marks the end
  // of the last upload from h_up
  uint8_t *h_up = nullptr, *h_down = nullptr;
  size_t cap_up = 0, used_up = 0, cap_down = 0;
  hipEvent_t ev_up = nullptr;
  double *d_bfut = nullptr;       // bound on the distortion of uncoded passes, per codeblock
  uint64_t *d_poff = nullptr;     // packed-record offsets per listed codeblock
  uint8_t *d_rpack = nullptr;     // packed records
  size_t cap_poff = 0, cap_rpack = 0;
  // list built on the device: {length, symbol bytes, passes, overflow}, and
  // the packed-record pass capacity the record layout was sized for
  uint64_t *d_t1info = nullptr;
  uint64_t cap_pass = 0;
  uint8_t *h_kz = nullptr;        // page-locked: K, then the pass limits, of every codeblock
  size_t cap_kz = 0;
  int last_dwt_launches = 0;      // kernel launches of the last DWT call
  // forward DWT launch sequences as HIP graphs, one per (parameter geometry,
  // buffers, normalisation, launch knobs): recorded the second time a key is
  // seen, replayed from then on (exec null: seen once, not recorded yet)
  struct DwtGraph {
    uint64_t key;
    hipGraphExec_t exec;
    int launches;
  };
  std::vector<DwtGraph> dwt_graphs;
  // pipeline graphs: the forward DWT with tier-1 of one (tile, parameter),
  // and the rate control of one tile, each keyed by everything its launch
  // sequence depends on (exec null: seen once, recorded the next time)
  struct Graph {
    uint64_t key;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  uint64_t graph_records = 0, graph_replays = 0;   // pipeline graphs recorded / launched
  ~Workspace();
};

Workspace *workspace_of(Job *job);

// Stage entry points (all asynchronous on ws->stream unless stated).
int fill_minmax(Workspace *ws, const bwc_codec *c, const Tile &tl, int param, const void *input, bool input_on_device,
                double *work);
int normalize(Workspace *ws, double *work, uint64_t n, uint8_t Qm);
// norm_qm >= 0: the tile still needs normalize(..., norm_qm); the DWT applies
// it (fused into its first pass where it can)
int dwt_forward(Workspace *ws, const bwc_codec *c, const Param &par, double *work, int norm_qm = -1);
int dwt_inverse(Workspace *ws, const bwc_codec *c, const Param &par, double *work);
// Fill + min/max + normalised forward DWT of (tile, param) from the caller's
// device field into `out` (device, the parameter's tile-local layout), with
// {min, max, alpha, beta} into `norm` (device, 4 doubles), asynchronously on
// `stream`, without allocating: the partials and the second DWT buffer use a
// static device scratch array of kNoAllocScratchBytes, so the parameter may
// hold at most (kNoAllocScratchBytes - 128 KiB) / 8 values. Concurrent calls
// on one GPU race on that array.
constexpr size_t kNoAllocScratchBytes = (64ull << 20) + (128ull << 10);
int dwt_noalloc(const bwc_codec *c, const Tile &tl, int param, const void *input, double *out, double *norm,
                hipStream_t stream);
// quantise + tier-1 for all codeblocks of (tile, param); results land in rec
// Pass limits for t1_encode: none, given per codeblock (host, par.ncb
// entries), or predicted on the device by the pass-limit network from the
// bitplane kernel's features (weights: kSkipModelSize floats, host).
struct PassLimits {
  const uint8_t *fixed = nullptr;
  const float *model = nullptr;
  int margin = 0;
  uint8_t *out = nullptr;   // optional: the limits used (host, par.ncb)
  bool features = false;    // compute the network's features even without a model
  // copy the records (L, S, distortions, U, bound) to the host; without it
  // they stay on the device for rate control there and only K, Z and Kmsbs
  // come back
  bool records = true;
  // launched first in the same device chain (the forward DWT), with the key
  // of its launch sequence; with pipeline graphs on, the chain from it to
  // tier-1's results is recorded once and replayed
  const std::function<int()> *pre = nullptr;
  uint64_t pre_key = 0;
};
// pipeline graphs (BWC_AMD_GRAPH, on unless set to 0)
bool graphs_on();
// pipeline graphs for the calling thread's compressions (on unless set off;
// the batch's threads run without them, so that no thread captures while
// the others launch)
void set_thread_graphs(bool on);
// the forward DWT's launches without a graph of its own, and their key
int dwt_forward_direct(Workspace *ws, const bwc_codec *c, const Param &par, double *work, int norm_qm);
uint64_t dwt_launch_key(const Workspace *ws, const bwc_codec *c, const Param &par, const double *work, int norm_qm);
// run `launch` (launches and copies on ws->stream that neither allocate nor
// wait) directly the first time `key` is seen, record it as a graph the
// second time, replay that graph from then on
int graph_run(Workspace *ws, uint64_t key, const std::function<int()> &launch);
int t1_encode(Workspace *ws, const bwc_codec *c, const Param &par, const double *work, CblkRecords &rec,
              uint16_t *slope_min, uint16_t *slope_max, int slot, const PassLimits &lim = PassLimits());
// features of the last t1_encode with a model (par.ncb x 16, feature 15
// relative to the tile mean)
int last_features(Workspace *ws, uint64_t ncb, float *out);

// pass-limit network kernels (skipnet.hip)
int skipnet_infer(hipStream_t s, const float *d_feat, uint64_t n, const float *d_w, const float *d_tstat,
                  const uint8_t *d_K, int margin, uint8_t *d_lim, float *d_y, char *err, size_t errlen);
size_t skipnet_train_lds();
int skipnet_train(hipStream_t s, const float *d_feat, const float *d_tgt, const uint32_t *d_order, uint64_t n,
                  uint64_t start, int steps, float lr, int t0, float *d_w, float *d_m, float *d_v, float *d_g,
                  float *d_loss, int mode, char *err, size_t errlen);
// host-buffer drivers of the network kernels on the current device
int skipnet_fit(const float *feat, const float *tgt, uint64_t n, float *w, int epochs, float lr, uint64_t seed,
                float *loss_out, char *err, size_t errlen);
int skipnet_grad(const float *feat, const float *tgt, uint64_t n, const float *w, float *grad, float *loss,
                 char *err, size_t errlen);
int skipnet_predict(const float *feat, uint64_t n, const float *w, float *y, char *err, size_t errlen);
// re-code codeblocks `cbs` of the parameter last coded into `slot` with new
// pass limits (zlim has par.ncb entries) and refresh their records
int t1_extend(Workspace *ws, const bwc_codec *c, const Param &par, CblkRecords &rec, int slot,
              const std::vector<uint64_t> &cbs, const uint8_t *zlim, bool records = true);
// PCRD rate control of tile `tile` on the device (pcrd.hip) over the records
// kept in the parameter slots 0 .. nPar-1: fills R.cp of every parameter and
// R.L at the cut points; with chk, the pass-skip check as well
int rc_tile(Workspace *ws, bwc_codec *c, Job *job, uint64_t tile, SkipCheck *chk);
// bisection probes of the last rc_tile call: (layer, threshold, estimate)
// triples, at most cap; returns the count
int rc_trace(Workspace *ws, uint32_t *layer, uint32_t *slope, int64_t *est, int cap);
void rc_release(Workspace *ws);
// A batch's global codeblock queue: the tier-1 calls of `fields` workspaces
// (each with batch set) join it after their codeblock lists are built; the
// last to join launches one symbol-generation and one MQ launch over every
// field's codeblocks (largest codings first) on the batch's stream, and all
// go on once those are queued. max_items bounds the codeblocks of all fields.
// After the launches at most `slots` of the joined calls go on at a time
// (the host side of rate control, packets and assembly).
BatchQueue *batch_create(int fields, uint64_t max_items, int slots);
// at the end of a batch call's compression: a call that never joined (an
// error before tier-1, or a field coded on the general tier-1 path) counts
// as joined with no codeblocks; a call that joined gives back its slot
void batch_leave(Workspace *ws);
// wait for the batch's launches and free it; nonzero when a launch failed
int batch_destroy(BatchQueue *q);
const char *rc_last_error();
// the header writer's stuffed byte count of n bit strings, by the device
// rate control's kernels (string i: nbits[i] bits from word woff[i])
int rc_stuffed_bytes(const uint32_t *words, const uint64_t *woff, const uint64_t *nbits, int n, uint64_t *bytes);
// gather each codeblock's prefix [0, len[i]) into rec.data / rec.data_off
int gather(Workspace *ws, CblkRecords &rec, const std::vector<uint32_t> &len);
// Packet bodies on the device: desc holds (offset in the body buffer,
// length, source offset in the coded data of parameter slot slots[k]) per
// chunk, all below `end`; the body buffer keeps its contents when it grows.
// body_d2h copies n body bytes at offset off to host memory.
int gather_bodies(Workspace *ws, const std::vector<uint64_t> &desc, const std::vector<int> &slots, uint64_t end);
int body_d2h(Workspace *ws, uint8_t *dst, uint64_t off, uint64_t n);
// decompression of one (tile, param): tier-1 decoding and reconstruction
// into work (zeroed first), and after the inverse DWT denormalisation into
// out (device memory holding the whole field, [param][t][z][y][x]);
// ensure_output gives a device staging buffer of `bytes`
int t1_decode(Workspace *ws, const bwc_codec *c, const Param &par, const DecodeRecords &rec, double *work);
int decode_output(Workspace *ws, const bwc_codec *c, const Tile &tl, int p, const Param &par, const double *work,
                  void *out);
int ensure_output(Workspace *ws, size_t bytes, void **out);
// the samples decode_output wrote for (tile, param) into the staging buffer,
// copied to the same places of a host output buffer
int output_box_d2h(Workspace *ws, const bwc_codec *c, const Tile &tl, int p, const Param &par, const void *stage,
                   void *out);
int sync(Workspace *ws);
const char *last_error();
int ensure_work(Workspace *ws, uint64_t n);
int record(Workspace *ws, int i);
double elapsed(Workspace *ws, int a, int b);
int read_norm(Workspace *ws, double mm[4]);
int is_device_pointer(const void *p);
void *host_alloc(size_t bytes);   // page-locked
void host_free(void *p);
int copy_h2d(Workspace *ws, void *dst, const void *src, size_t n);
int copy_d2h(Workspace *ws, void *dst, const void *src, size_t n);
int device_count();

}  // namespace gpu
}  // namespace bwc
