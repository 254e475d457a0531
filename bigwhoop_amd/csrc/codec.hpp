// Internal data model of libbwc (MI355X build).
//
// The public bwc_codec (include/bwc.h) keeps the reference layout; its opaque
// `tile` pointer holds a Job: the codec plan (tiles -> parameters ->
// resolutions -> subbands -> precincts -> codeblocks, flattened into arrays
// so device kernels and tier-2 share one codeblock index) plus the encoder
// state between bwc_create_compression and bwc_compress.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <functional>
#include <memory>
#include <vector>

#include "../../include/bwc.h"

namespace bwc {

constexpr int kMaxPasses = 190;     // (64 * 3) - 2 fractional bitplanes, FP64 build
constexpr int kPrecBits = 63;       // PREC_BIT of the FP64 build
constexpr int kPrecBytes = 8;
// pass-limit network 16 -> 64 -> 64 -> 1 (+ 16 feature means and scales)
constexpr int kSkipFeatures = 16, kSkipHidden = 64;
constexpr int kSkipModelSize = kSkipHidden * kSkipFeatures + kSkipHidden + kSkipHidden * kSkipHidden + kSkipHidden +
                               kSkipHidden + 1 + 2 * kSkipFeatures;

struct Box {
  uint64_t lo[4] = {0, 0, 0, 0};
  uint64_t hi[4] = {0, 0, 0, 0};
  uint64_t len(int a) const { return hi[a] - lo[a]; }
  uint64_t volume() const { return len(0) * len(1) * len(2) * len(3); }
};

// One tag tree (JPEG 2000 style, 4-D leaves) stored as flat node arrays.
struct TagTree {
  std::vector<uint16_t> value, threshold;
  std::vector<int64_t> parent;   // -1 at the root
  void build(uint64_t lx, uint64_t ly, uint64_t lz, uint64_t lt);
  void set(uint64_t leaf, uint16_t v);
  void reset_thresholds();
};

struct Precinct {
  uint32_t box[8];           // precinct bounds in subband coordinates (x0..t0, x1..t1)
  uint16_t ncb[4];
  uint64_t ncbs = 0;
  uint64_t first_cb = 0;     // index of its first codeblock in Param::cb_* arrays
  TagTree incl, msbs;
};

struct Subband {
  Box box;
  double gain = 1, step = 1, step_eff = 1;
  uint8_t hb = 0;
  uint16_t Kmax = 0;
  uint8_t qt_exp = 0;
  uint16_t qt_mant = 0;
  std::vector<Precinct> prec;   // one per precinct of the resolution
};

struct Packet {
  std::vector<uint8_t> header;
  uint64_t body_size = 0;
  uint8_t e = 0;
  uint32_t size = 0;
  bool built = false;
  uint64_t dev_off = 0;   // body offset in the device body buffer (bodies gathered on the GPU)
};

struct Resolution {
  Box box;
  uint16_t np[4] = {0, 0, 0, 0};
  uint64_t nprec = 0;
  std::vector<Subband> sb;
  std::vector<Packet> pk;       // [layer * nprec + p]
};

struct Param {
  Box box;
  double pmin = 0, pmax = 0, alpha = 0, beta = 0;
  std::vector<Resolution> res;
  // codeblocks in the reference's access order (subband, precinct, raster)
  uint64_t ncb = 0;
  std::vector<uint32_t> cb_box;      // 8 per codeblock, tile-local buffer coordinates
  std::vector<uint32_t> cb_sbref;    // (resolution << 8) | subband index
  std::vector<uint32_t> cb_prec;     // precinct index within the resolution
  std::vector<uint32_t> cb_idx;      // index within its precinct
};

// Decoder-side records of one (tile, param), per codeblock in access order:
// bit-planes skipped (Kmsbs), passes included (Z), coded bytes (len) and
// their offset in `data`.
struct DecodeRecords {
  uint64_t n = 0;
  std::vector<uint8_t> Kmsbs, Z;
  std::vector<uint32_t> len;
  std::vector<uint64_t> off;   // n + 1
  // the codeblocks' bytes back to back (off), 512 zeroed bytes of slack
  // after the last; data_size bytes
  std::unique_ptr<uint8_t[]> data;
  uint64_t data_size = 0;
};

struct Tile {
  Box box;
  uint32_t index = 0;
  uint32_t header_size = 18;
  uint64_t body_size = 0;
  uint64_t npackets = 0;
  uint16_t slope_min = 0xFFFF, slope_max = 0;
  std::vector<Param> par;
  std::vector<Packet *> seq;
};

// Per-codeblock tier-1 output consumed by tier-2 (host copy).
struct CblkRecords {
  uint64_t n = 0;
  std::vector<uint8_t> K, Kmsbs, Z;
  // Z = passes coded; Zfull = 3K - 2 passes of a complete coding; U = first
  // pass whose length may still change when the codeblock is coded further
  // (Z when coded completely)
  std::vector<uint8_t> Zfull, U;
  std::vector<uint32_t> L;      // kMaxPasses per codeblock, bytes
  std::vector<uint16_t> S;      // kMaxPasses + 1 per codeblock
  // fast tier-1 path: distortion reduction of each coded pass (kMaxPasses
  // per codeblock, the values the hull was built from) and, for a codeblock
  // coded to a prefix of its passes, an upper bound on what the passes left
  // out can add
  std::vector<double> mse, bfut;
  // tier-2 state
  std::vector<uint16_t> beta, beta_est;
  std::vector<int16_t> cp;      // (nLayers + 1) per codeblock, cp[0] == 0 is layer -1
  // included bytes, packed after rate control: data_off[i] into `data`
  std::vector<uint64_t> data_off;
  std::vector<uint8_t> data;
  // where tier-1 left the full MQ bytes on the device: workspace slot and
  // per-codeblock offsets inside it
  int slot = 0;
  std::vector<uint64_t> src_off;
  // pass limits the codeblocks were coded with (pass skipping)
  std::vector<uint8_t> zlim;
  // pass-limit network features (16 per codeblock) when collected
  std::vector<float> feat;
  // MQ decisions of each codeblock's last coding (fast tier-1 path), when
  // collected with the features
  std::vector<uint32_t> decisions;
  void resize(uint64_t ncb, int nlayers);
};

struct Plan {
  bwc_codec *codec = nullptr;
  std::vector<Tile> tiles;
};

// Whole library-side state hanging off bwc_codec::tile.
struct Job {
  Plan plan;
  double lut[3][10];                 // gain LUT snapshot used for this codec
  std::vector<float> bitrate;        // owned copy; codec->control.bitrate points here
  bool created = false;
  // per (tile, param) records, in tile-major order
  std::vector<CblkRecords> rec;
  // decoder-side: per (tile, param) records from the packets
  std::vector<DecodeRecords> drec;
  bool decodable = false;
  // device workspace (gpu::Workspace), created on first use
  void *ws = nullptr;
  void (*ws_free)(void *) = nullptr;
  // stage timings of the last compression (bwc_amd_last_timings)
  double ms[10] = {0};
  int launches[4] = {0};
  // pass skipping (bwc_amd_set_pass_skip): 0 off, 1 fixed pass limit
  // (skip_param passes), 2 learned limit (skip_param margin passes)
  int skip_mode = 0, skip_param = 0;
  std::vector<float> skip_w;        // network weights, fp32 (mode 2)
  bool collect_features = false;    // keep the network's features per codeblock
  // last compression: [0] codeblocks coded, [1] passes coded (all rounds),
  // [2] passes of complete codings, [3] extra rate-control rounds,
  // [4] codeblocks coded further, [5] fallbacks to complete coding
  uint64_t skip_stats[8] = {0};
  ~Job() {
    if (ws && ws_free) ws_free(ws);
  }
};

// ---- configuration / geometry (geometry.cpp) ----
void gain_lut_accumulate();                    // one initialize_gain_lut call (process global)
const double (*gain_lut())[10];
void reset_gain_lut();                         // back to the state of a fresh process
double energy_gain(const bwc_codec *c, const double lut[3][10], uint8_t hb, uint16_t level);
void compute_qt(bwc_gl_ctrl *ctl, double delta);
int build_plan(bwc_codec *codec, Job *job);
Job *job_of(bwc_codec *c);

// ---- tier-2 + codestream (tier2.cpp) ----
int t2_encode_tile(bwc_codec *codec, Job *job, uint64_t tile);
// Outcome of rate control over prefix-coded records (pass skipping):
// codeblocks per parameter that must be coded further, or `fallback` when
// the tile's thresholds are not provably those of complete codings.
struct SkipCheck {
  int margin = 3;
  bool fallback = false;
  std::vector<std::vector<uint64_t>> flagged;
};
int t2_rate_control(bwc_codec *codec, Job *job, uint64_t tile, SkipCheck *chk = nullptr);
int t2_packets(bwc_codec *codec, Job *job, uint64_t tile);
// the packet flags, body sizes and tag-tree values create_quality_layer
// leaves for the cuts already in the records (R.cp, R.L at the cuts): the
// host's share of rate control done on the device
int t2_apply_cuts(bwc_codec *codec, Job *job, uint64_t tile);
// the host thread pool of rate control: fn(0 .. n-1) on `workers` pool
// threads and the caller; rc_thread_count() is the rate-control thread count
// (BWC_AMD_RC_THREADS, else one less than the host threads, at most 15)
void host_pool_run(int workers, size_t n, const std::function<void(size_t)> &fn);
int rc_thread_count();
// Packet bodies already laid out elsewhere (on the device): copy(ctx, dst,
// dev_off, n) writes the n body bytes of a packet, found at Packet::dev_off,
// to dst; nonzero return = failure.
struct BodySource {
  void *ctx = nullptr;
  int (*copy)(void *ctx, uint8_t *dst, uint64_t dev_off, uint64_t n) = nullptr;
};
// Codestream assembly into out. Bodies come from each record's data (host)
// or, with `bodies`, from the packets' device layout.
size_t assemble_codestream(bwc_codec *codec, Job *job, const bwc_stream *st, uint8_t *out,
                           const BodySource *bodies = nullptr);
// The (source offset in the parameter's coded data, length) of every body
// chunk of packet (param j, resolution r, precinct p, layer l) of a tile, in
// codestream order.
void packet_body_chunks(bwc_codec *codec, Job *job, uint64_t tile, int j, int r, uint64_t p, int l,
                        std::vector<std::pair<uint64_t, uint64_t>> &out);
uint64_t header_size(const bwc_codec *codec, const bwc_stream *st);

// ---- decode (decode.cpp) ----
// bwc_create_decompression: main header into the codec (configuration,
// geometry, parameter ranges), then every packet into per-codeblock decode
// records (job->drec). bwc_open_header uses the header part alone.
int parse_codestream(bwc_codec *codec, bwc_stream *data, uint8_t layer);
int parse_header_only(bwc_codec *codec, bwc_stream *data);

void warn(const char *msg);

}  // namespace bwc
