// C API of libbwc (include/bwc.h, include/bwc_amd.h).
//
// Configuration and setters keep the reference's validation, warnings and
// codestream control bits (libbwc.c:1828-2828); bwc_compress runs the hot
// stages on one GPU (gpu.hip) and tier-2 / codestream assembly on the host
// (tier2.cpp).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <atomic>
#include <mutex>
#include <string>
#include <thread>

#include "codec.hpp"
#include "gpu.hpp"
#include "tuning.hpp"

using namespace bwc;

namespace {

thread_local char g_api_err[512];

void set_err(const char *m) {
  std::snprintf(g_api_err, sizeof g_api_err, "%s", m);
  std::fprintf(stderr, "bwc: error: %s\n", m);
}

bwc_codec *alloc_codec() {
  auto *c = static_cast<bwc_codec *>(std::calloc(1, sizeof(bwc_codec)));
  if (!c) return nullptr;
  c->info.codec_prec = bwc_precision_double;
  c->tile = new Job();
  return c;
}

void refresh_qt(bwc_codec *c) {
  double lut[3][10];
  std::memcpy(lut, gain_lut(), sizeof lut);
  const double delta = 1 / (std::pow(2, 2 + kPrecBits) * std::sqrt(energy_gain(c, lut, 0, c->control.nDecomp)));
  compute_qt(&c->control, delta);
}

uint8_t levels_of(uint64_t n) { return (uint8_t)(std::log((double)n) / std::log(2)); }

bool trace_on() {
  static const bool on = [] {
    const char *e = std::getenv("BWC_AMD_RC_TRACE");
    return e && e[0] == '1';
  }();
  return on;
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void reset_for_encode(bwc_codec *c, Job *job) {
  for (auto &tl : job->plan.tiles) {
    tl.slope_min = 0xFFFF;
    tl.slope_max = 0;
    for (auto &par : tl.par)
      for (auto &res : par.res) {
        for (auto &pk : res.pk) pk = Packet();
        for (auto &sb : res.sb)
          for (auto &pr : sb.prec) {
            std::fill(pr.incl.value.begin(), pr.incl.value.end(), 0xFFFF);
            std::fill(pr.incl.threshold.begin(), pr.incl.threshold.end(), 0);
            std::fill(pr.msbs.value.begin(), pr.msbs.value.end(), 0xFFFF);
            std::fill(pr.msbs.threshold.begin(), pr.msbs.threshold.end(), 0);
          }
      }
  }
  // records keep their allocations from one call to the next (tier-1 fills
  // them again); only what a call may leave unset is cleared
  const size_t nrec = job->plan.tiles.size() * c->info.nPar;
  if (job->rec.size() != nrec) {
    job->rec.assign(nrec, CblkRecords());
    return;
  }
  for (auto &R : job->rec) {
    R.n = 0;
    R.slot = 0;
    R.data.clear();
    R.src_off.clear();
    R.zlim.clear();
    R.feat.clear();
    R.decisions.clear();
  }
}

// tier-2 state of one tile back to "before rate control", so rate control
// can run again after codeblocks were coded further; slopes: recompute the
// tile's slope bounds from the records (tier-1 reports them for its own
// codings, so only codings extended since need it)
void reset_tile_t2(bwc_codec *c, Job *job, uint64_t ti, bool slopes = true) {
  Tile &tl = job->plan.tiles[ti];
  const int nl = c->control.nLayers;
  for (auto &par : tl.par)
    for (auto &res : par.res) {
      for (auto &pk : res.pk) pk = Packet();
      for (auto &sb : res.sb)
        for (auto &pr : sb.prec) {
          std::fill(pr.incl.value.begin(), pr.incl.value.end(), 0xFFFF);
          std::fill(pr.incl.threshold.begin(), pr.incl.threshold.end(), 0);
          std::fill(pr.msbs.value.begin(), pr.msbs.value.end(), 0xFFFF);
          std::fill(pr.msbs.threshold.begin(), pr.msbs.threshold.end(), 0);
        }
    }
  // slope bounds over the current records (nonzero slopes, as tier-1 reports)
  if (slopes) {
    tl.slope_min = 0xFFFF;
    tl.slope_max = 0;
  }
  for (int p = 0; p < c->info.nPar; ++p) {
    CblkRecords &R = job->rec[ti * c->info.nPar + p];
    R.beta.assign(R.n, 3);
    R.beta_est.assign(R.n, 3);
    R.cp.assign(R.n * (nl + 1), 0);
    if (!slopes) continue;
    for (uint64_t i = 0; i < R.n; ++i) {
      if (!R.K[i]) continue;
      const uint16_t *S = &R.S[i * (kMaxPasses + 1)];
      for (int z = 0; z <= R.Z[i]; ++z)
        if (S[z]) {
          tl.slope_min = std::min(tl.slope_min, S[z]);
          tl.slope_max = std::max(tl.slope_max, S[z]);
        }
    }
  }
}

// Gather the packet bodies of tile ti on the device, from [body_end] on:
// each packet's chunks back to back (Packet::dev_off, its body_size bytes).
int gather_tile_bodies(bwc_codec *c, Job *job, gpu::Workspace *ws, uint64_t ti, uint64_t &body_end) {
  Tile &tl = job->plan.tiles[ti];
  const int nl = c->control.nLayers;
  std::vector<uint64_t> desc;
  std::vector<int> slots;
  std::vector<std::pair<uint64_t, uint64_t>> ch;
  uint64_t off = body_end;
  for (int j = 0; j < c->info.nPar; ++j) {
    const CblkRecords &R = job->rec[ti * c->info.nPar + j];
    for (size_t r = 0; r < tl.par[j].res.size(); ++r) {
      Resolution &res = tl.par[j].res[r];
      for (uint64_t p = 0; p < res.nprec; ++p)
        for (int l = 0; l < nl; ++l) {
          Packet &pk = res.pk[l * res.nprec + p];
          pk.dev_off = off;
          if (!pk.e) continue;
          ch.clear();
          packet_body_chunks(c, job, ti, j, (int)r, p, l, ch);
          uint64_t sum = 0;
          for (auto &x : ch) {
            if (!x.second) continue;
            desc.push_back(off + sum);
            desc.push_back(x.second);
            desc.push_back(x.first);
            slots.push_back(R.slot);
            sum += x.second;
          }
          if (sum != pk.body_size) {
            set_err("packet body size differs from its chunks");
            return 1;
          }
          off += sum;
        }
    }
  }
  if (gpu::gather_bodies(ws, desc, slots, off)) return 1;
  body_end = off;
  return 0;
}

// Rate control of one tile whose codeblocks may be coded to a prefix of
// their passes: run it, and while its outcome is not provably that of
// complete codings, code the flagged codeblocks further (limits grow 9, 18,
// 36 passes per round; complete coding from the fourth round or for the
// whole tile on a fallback) and run it again.
int rate_control_prefix(bwc_codec *c, Job *job, gpu::Workspace *ws, uint64_t ti) {
  Tile &tl = job->plan.tiles[ti];
  const int np = c->info.nPar;
  SkipCheck chk;
  chk.margin = 3;
  // BWC_AMD_SKIP_SETTLE: passes a cut must lie before the first unsettled
  // pass (measurement of the proof's slack; at least 1)
  if (const char *e = std::getenv("BWC_AMD_SKIP_SETTLE")) chk.margin = std::max(1, std::atoi(e));
  for (int round = 0;; ++round) {
    double tr = now_ms();
    reset_tile_t2(c, job, ti, round > 0);
    // every codeblock coded completely: the outcome is exact as it stands
    uint64_t ntrunc = 0;
    for (int p = 0; p < np; ++p) {
      const CblkRecords &R = job->rec[ti * np + p];
      for (uint64_t i = 0; i < R.n; ++i) ntrunc += R.Z[i] < R.Zfull[i];
    }
    if (trace_on()) std::fprintf(stderr, "rc: reset + count          %8.3f ms\n", now_ms() - tr);
    if (!ntrunc) {
      t2_rate_control(c, job, ti);
      return 0;
    }
    tr = now_ms();
    t2_rate_control(c, job, ti, &chk);
    if (trace_on()) std::fprintf(stderr, "rc: t2_rate_control total  %8.3f ms\n", now_ms() - tr);
    uint64_t nflag = 0;
    for (auto &f : chk.flagged) nflag += f.size();
    if (!chk.fallback && !nflag) return 0;
    // the tile-level test speaks about settled cuts only: code flagged
    // codeblocks further first, fall back only once none is flagged
    if (nflag) chk.fallback = false;
    job->skip_stats[3] += 1;
    if (chk.fallback) job->skip_stats[5] += 1;
    for (int p = 0; p < np; ++p) {
      CblkRecords &R = job->rec[ti * np + p];
      std::vector<uint64_t> list;
      if (chk.fallback) {
        for (uint64_t i = 0; i < R.n; ++i)
          if (R.Z[i] < R.Zfull[i]) list.push_back(i);
      } else {
        list = chk.flagged[p];
      }
      if (list.empty()) continue;
      for (uint64_t i : list) {
        const int step = 9 << std::min(round, 3);
        int nz = (chk.fallback || round >= 3) ? R.Zfull[i] : std::min<int>(R.Zfull[i], R.Z[i] + step);
        R.zlim[i] = (uint8_t)nz;
        job->skip_stats[1] += (uint64_t)nz;   // passes coded again
      }
      job->skip_stats[4] += list.size();
      if (gpu::t1_extend(ws, c, tl.par[p], R, p, list, R.zlim.data())) return 1;
    }
  }
}

// Rate control on the device (the default; BWC_AMD_RC=host keeps it on the
// host pool over records copied back from tier-1)
bool rc_on_device() {
  const char *e = std::getenv("BWC_AMD_RC");
  return !(e && std::strcmp(e, "host") == 0);
}

// Rate control of tile ti on the device, with the records left there by
// tier-1: as rate_control_prefix, the pass-skip check runs while codeblocks
// are coded to a prefix of their passes, and flagged codeblocks are coded
// further until the outcome is that of complete codings. The host then only
// sets the packet flags, body sizes and tag trees from the cuts.
int rate_control_dev(bwc_codec *c, Job *job, gpu::Workspace *ws, uint64_t ti) {
  Tile &tl = job->plan.tiles[ti];
  const int np = c->info.nPar;
  SkipCheck chk;
  chk.margin = 3;
  if (const char *e = std::getenv("BWC_AMD_SKIP_SETTLE")) chk.margin = std::max(1, std::atoi(e));
  for (int round = 0;; ++round) {
    reset_tile_t2(c, job, ti, false);
    uint64_t ntrunc = 0;
    for (int p = 0; p < np; ++p) {
      const CblkRecords &R = job->rec[ti * np + p];
      for (uint64_t i = 0; i < R.n; ++i) ntrunc += R.Z[i] < R.Zfull[i];
    }
    const double tr = now_ms();
    if (gpu::rc_tile(ws, c, job, ti, ntrunc ? &chk : nullptr)) {
      set_err(gpu::rc_last_error());
      return 1;
    }
    if (trace_on()) std::fprintf(stderr, "rc: device rate control   %8.3f ms\n", now_ms() - tr);
    uint64_t nflag = 0;
    if (ntrunc)
      for (auto &f : chk.flagged) nflag += f.size();
    if (!ntrunc || (!chk.fallback && !nflag)) {
      t2_apply_cuts(c, job, ti);
      return 0;
    }
    if (nflag) chk.fallback = false;
    job->skip_stats[3] += 1;
    if (chk.fallback) job->skip_stats[5] += 1;
    for (int p = 0; p < np; ++p) {
      CblkRecords &R = job->rec[ti * np + p];
      std::vector<uint64_t> list;
      if (chk.fallback) {
        for (uint64_t i = 0; i < R.n; ++i)
          if (R.Z[i] < R.Zfull[i]) list.push_back(i);
      } else {
        list = chk.flagged[p];
      }
      if (list.empty()) continue;
      for (uint64_t i : list) {
        const int step = 9 << std::min(round, 3);
        int nz = (chk.fallback || round >= 3) ? R.Zfull[i] : std::min<int>(R.Zfull[i], R.Z[i] + step);
        R.zlim[i] = (uint8_t)nz;
        job->skip_stats[1] += (uint64_t)nz;
      }
      job->skip_stats[4] += list.size();
      if (gpu::t1_extend(ws, c, tl.par[p], R, p, list, R.zlim.data(), false)) {
        set_err(gpu::last_error());
        return 1;
      }
    }
  }
}

}  // namespace

extern "C" {

bwc_stream *bwc_init_stream(void *const inpbuf, void *const outbuf, bwc_mode const mode) {
  auto *s = static_cast<bwc_stream *>(std::calloc(1, sizeof(bwc_stream)));
  if (!s) return nullptr;
  s->inp = inpbuf;
  s->out = outbuf;
  s->mode = mode;
  return s;
}

uchar bwc_set_com(bwc_stream *const data, char const *const com, uint16 const size) {
  data->codestream.com = static_cast<bwc_span *>(std::calloc(1, sizeof(bwc_span)));
  if (!data->codestream.com) return 1;
  data->codestream.com->memory = static_cast<uchar *>(std::calloc(size ? size : 1, 1));
  if (!data->codestream.com->memory) return 1;
  std::memcpy(data->codestream.com->memory, com, size);
  data->codestream.com->access = data->codestream.com->memory;
  data->codestream.com->size = size;
  return 0;
}

uchar bwc_set_aux(bwc_stream *const data, char const *const aux, uint32 const size) {
  data->codestream.aux = static_cast<bwc_span *>(std::calloc(1, sizeof(bwc_span)));
  if (!data->codestream.aux) return 1;
  data->codestream.aux->memory = static_cast<uchar *>(std::calloc(size ? size : 1, 1));
  if (!data->codestream.aux->memory) return 1;
  std::memcpy(data->codestream.aux->memory, aux, size);
  data->codestream.aux->access = data->codestream.aux->memory;
  data->codestream.aux->size = size;
  return 0;
}

// configure_codec (libbwc.c:1828-1982)
bwc_codec *configure_codec(bwc_codec *const codec, uint64 const nX, uint64 const nY, uint64 const nZ, uint64 const nTS,
                           uint8 const nPar, bwc_precision const prec) {
  bwc_gl_ctrl &c = codec->control;
  bwc_gl_inf &i = codec->info;
  i.nX = nX, i.nY = nY, i.nZ = nZ, i.nTS = nTS, i.nPar = nPar, i.data_prec = prec;
  c.tileSizeX = nX, c.tileSizeY = nY, c.tileSizeZ = nZ, c.tileSizeTS = nTS;
  c.nTiles = 1;
  c.precSizeX = (nX >> 1) ? 15 : 0, c.precSizeY = (nY >> 1) ? 15 : 0;
  c.precSizeZ = (nZ >> 1) ? 15 : 0, c.precSizeTS = (nTS >> 1) ? 15 : 0;
  c.cbX = (nX >> 1) ? 5 : 0, c.cbY = (nY >> 1) ? 5 : 0, c.cbZ = (nZ >> 1) ? 5 : 0, c.cbTS = (nTS >> 1) ? 5 : 0;
  c.KernelX = c.KernelY = c.KernelZ = c.KernelTS = bwc_dwt_9_7;
  c.progression = bwc_prog_LRCP;
  c.quantization_style = bwc_qt_derived;
  c.nLayers = 0;
  c.Qm = kPrecBits - 1;
  c.bitrate = nullptr;
  c.guard_bits = 2;
  const uint8_t lx = levels_of(nX), ly = levels_of(nY), lz = levels_of(nZ), lt = levels_of(nTS);
  c.decompX = (nX >> 1) ? (4 < lx ? 4 : lx) : 0;
  c.decompY = (nY >> 1) ? (4 < ly ? 4 : ly) : 0;
  c.decompZ = (nZ >> 1) ? (4 < lz ? 4 : lz) : 0;
  c.decompTS = (nTS >> 1) ? (4 < lt ? 4 : lt) : 0;
  c.nDecomp = std::max(c.decompX, std::max(c.decompY, std::max(c.decompZ, c.decompTS)));
  gain_lut_accumulate();
  refresh_qt(codec);
  return codec;
}

bwc_codec *bwc_alloc_coder(uint64 const nX, uint64 const nY, uint64 const nZ, uint64 const nTS, uint8 const nPar,
                           bwc_precision const prec) {
  bwc_codec *c = alloc_codec();
  if (!c) return nullptr;
  configure_codec(c, nX, nY, nZ, nTS, nPar, prec);
  c->mode = comp;
  return c;
}

bwc_codec *bwc_alloc_decoder(void) {
  bwc_codec *c = alloc_codec();
  if (!c) return nullptr;
  c->mode = decomp;
  return c;
}

void bwc_free_codec(bwc_codec *const codec) {
  if (!codec) return;
  delete job_of(codec);
  std::free(codec->control.bitrate);
  std::free(codec);
}

void bwc_set_error_resilience(bwc_codec *const codec) {
  codec->control.error_resilience ^= 0x01;
  codec->control.CSsgc ^= (0x01 << 0);
}

void set_quant_style(bwc_codec *const codec, bwc_quant_st const style) {
  codec->control.quantization_style = style;
  codec->control.CSsgc |= (0x01 << 1);
}

void set_quant_step_size(bwc_codec *const codec, double const delta) {
  if ((delta <= 0) || (delta >= 2)) {
    warn("invalid quantization step size; 0 < step size < 2");
    return;
  }
  compute_qt(&codec->control, delta);
  codec->control.CSsgc |= (0x01 << 2);
}

void set_progression(bwc_codec *const codec, bwc_prog_ord const progression) {
  codec->control.progression = progression;
  codec->control.CSsgc |= (0x01 << 3);
}

void set_kernels(bwc_codec *const codec, bwc_dwt_filter const KernelX, bwc_dwt_filter const KernelY,
                 bwc_dwt_filter const KernelZ, bwc_dwt_filter const KernelTS) {
  bwc_gl_ctrl &c = codec->control;
  c.KernelX = KernelX ? KernelX : bwc_dwt_9_7;
  c.KernelY = KernelY ? KernelY : bwc_dwt_9_7;
  c.KernelZ = KernelZ ? KernelZ : bwc_dwt_9_7;
  c.KernelTS = KernelTS ? KernelTS : bwc_dwt_haar;
  c.CSsgc |= (0x01 << 4);
}

// bwc_set_decomp (libbwc.c:2293-2421): accumulates the gain table again
void bwc_set_decomp(bwc_codec *const codec, uint8 const dX, uint8 const dY, uint8 const dZ, uint8 const dTS) {
  bwc_gl_ctrl &c = codec->control;
  bwc_gl_inf &i = codec->info;
  const uint8_t lx = levels_of(i.nX), ly = levels_of(i.nY), lz = levels_of(i.nZ), lt = levels_of(i.nTS);
  c.decompX = (i.nX >> 1) ? (dX < lx ? dX : lx) : 0;
  c.decompY = (i.nY >> 1) ? (dY < ly ? dY : ly) : 0;
  c.decompZ = (i.nZ >> 1) ? (dZ < lz ? dZ : lz) : 0;
  c.decompTS = (i.nTS >> 1) ? (dTS < lt ? dTS : lt) : 0;
  c.nDecomp = std::max(c.decompX, std::max(c.decompY, std::max(c.decompZ, c.decompTS)));
  if (c.decompX > 63 || c.decompY > 63 || c.decompZ > 63 || c.decompTS > 31) {
    warn("invalid decomposition level value; maximum is 63 (spatial) and 31 (temporal)");
    c.decompX = (i.nX >> 1) ? 4 : 0;
    c.decompY = (i.nY >> 1) ? 4 : 0;
    c.decompZ = (i.nZ >> 1) ? 4 : 0;
    c.decompTS = (i.nTS >> 1) ? 4 : 0;
    return;
  }
  gain_lut_accumulate();
  refresh_qt(codec);
  c.CSsgc |= (0x01 << 5);
}

void bwc_set_precincts(bwc_codec *const codec, uint8 const pX, uint8 const pY, uint8 const pZ, uint8 const pTS) {
  bwc_gl_ctrl &c = codec->control;
  bwc_gl_inf &i = codec->info;
  c.precSizeX = (i.nX >> 1) ? (pX ? pX : 15) : 0;
  c.precSizeY = (i.nY >> 1) ? (pY ? pY : 15) : 0;
  c.precSizeZ = (i.nZ >> 1) ? (pZ ? pZ : 15) : 0;
  c.precSizeTS = (i.nTS >> 1) ? (pTS ? pTS : 15) : 0;
  if ((c.precSizeX < 1 && (i.nX >> 1)) || (c.precSizeY < 1 && (i.nY >> 1)) || (c.precSizeZ < 1 && (i.nZ >> 1)) ||
      (c.precSizeTS < 1 && (i.nTS >> 1)) || c.precSizeX > 15 || c.precSizeY > 15 || c.precSizeZ > 15 ||
      c.precSizeTS > 15) {
    warn("invalid precinct size; valid sizes are 2^1 to 2^15");
    c.precSizeX = (i.nX >> 1) ? 15 : 0;
    c.precSizeY = (i.nY >> 1) ? 15 : 0;
    c.precSizeZ = (i.nZ >> 1) ? 15 : 0;
    c.precSizeTS = (i.nTS >> 1) ? 15 : 0;
    return;
  }
  c.CSsgc |= (0x01 << 6);
}

void bwc_set_codeblocks(bwc_codec *const codec, uint8 const cbX, uint8 const cbY, uint8 const cbZ, uint8 const cbTS) {
  bwc_gl_ctrl &c = codec->control;
  bwc_gl_inf &i = codec->info;
  c.cbX = (i.nX >> 1) ? (cbX ? cbX : 5) : 0;
  c.cbY = (i.nY >> 1) ? (cbY ? cbY : 5) : 0;
  c.cbZ = (i.nZ >> 1) ? (cbZ ? cbZ : 5) : 0;
  c.cbTS = (i.nTS >> 1) ? (cbTS ? cbTS : 5) : 0;
  const int s = c.cbX + c.cbY + c.cbZ + c.cbTS;
  if (c.cbX > 10 || c.cbY > 10 || c.cbZ > 10 || c.cbTS > 10 || s < 4 || s > 20) {
    warn("invalid codeblock size; at most 2^20 values with at most 2^10 per axis, at least 2^4");
    c.cbX = (i.nX >> 1) ? 5 : 0;
    c.cbY = (i.nY >> 1) ? 5 : 0;
    c.cbZ = (i.nZ >> 1) ? 5 : 0;
    c.cbTS = (i.nTS >> 1) ? 5 : 0;
    return;
  }
  c.CSsgc |= (0x01 << 7);
}

void bwc_set_qm(bwc_codec *const codec, uint8 const Qm) {
  if ((int8_t)(kPrecBits - Qm) < 1) {
    warn("invalid Q number format range; larger than the permitted 62 bits");
    return;
  }
  codec->control.Qm = Qm;
  codec->control.CSsgc |= (0x01 << 8);
}

// bwc_set_tiles (libbwc.c:2648-2828)
void bwc_set_tiles(bwc_codec *const codec, uint64 const tX, uint64 const tY, uint64 const tZ, uint64 const tTS,
                   bwc_tile_instr const instr) {
  bwc_gl_ctrl &c = codec->control;
  bwc_gl_inf &i = codec->info;
  if (instr == bwc_tile_sizeof) {
    if ((c.tileSizeX < 16 && c.tileSizeX > i.nX) || (c.tileSizeY < 16 && c.tileSizeY > i.nY) ||
        (c.tileSizeZ < 16 && c.tileSizeZ > i.nZ) || (c.tileSizeTS < 16 && c.tileSizeTS > i.nTS)) {
      warn("invalid tile dimensions; 16 <= tile size <= grid points");
      return;
    }
    c.tileSizeX = (i.nX >> 1) ? (tX ? tX : i.nX) : i.nX;
    c.tileSizeY = (i.nY >> 1) ? (tY ? tY : i.nY) : i.nY;
    c.tileSizeZ = (i.nZ >> 1) ? (tZ ? tZ : i.nZ) : i.nZ;
    c.tileSizeTS = (i.nTS >> 1) ? (tTS ? tTS : i.nTS) : i.nTS;
    const uint64_t nx = (uint64_t)std::ceil(((float)i.nX / c.tileSizeX));
    const uint64_t ny = (uint64_t)std::ceil(((float)i.nY / c.tileSizeY));
    const uint64_t nz = (uint64_t)std::ceil(((float)i.nZ / c.tileSizeZ));
    const uint64_t nt = (uint64_t)std::ceil(((float)i.nTS / c.tileSizeTS));
    c.nTiles = nx * ny * nz * nt;
    if (((double)nx * ny * nz * nt) > (double)0xFFFFFFFFFFFFFFFFull) {
      warn("invalid tile dimensions; the number of tiles must stay below 2^64");
      c.tileSizeX = i.nX, c.tileSizeY = i.nY, c.tileSizeZ = i.nZ, c.tileSizeTS = i.nTS;
      return;
    }
  } else {
    if (((double)tX * tY * tZ * tTS) > (double)0xFFFFFFFFFFFFFFFFull) {
      warn("invalid number of tiles; must stay below 2^64");
      return;
    }
    c.tileSizeX = (i.nX >> 1) ? (tX ? (uint64_t)std::ceil(((float)i.nX / tX)) : i.nX) : i.nX;
    c.tileSizeY = (i.nY >> 1) ? (tY ? (uint64_t)std::ceil(((float)i.nY / tY)) : i.nY) : i.nY;
    c.tileSizeZ = (i.nZ >> 1) ? (tZ ? (uint64_t)std::ceil(((float)i.nZ / tZ)) : i.nZ) : i.nZ;
    c.tileSizeTS = (i.nTS >> 1) ? (tTS ? (uint64_t)std::ceil(((float)i.nTS / tTS)) : i.nTS) : i.nTS;
    if ((c.tileSizeX < 16 && c.tileSizeX > i.nX) || (c.tileSizeY < 16 && c.tileSizeY > i.nY) ||
        (c.tileSizeZ < 16 && c.tileSizeZ > i.nZ) || (c.tileSizeTS < 16 && c.tileSizeTS > i.nTS)) {
      warn("invalid number of tiles; tile sizes must satisfy 16 <= size <= grid points");
      c.tileSizeX = i.nX, c.tileSizeY = i.nY, c.tileSizeZ = i.nZ, c.tileSizeTS = i.nTS;
      return;
    }
    const uint64_t nx = (uint64_t)std::ceil(((float)i.nX / c.tileSizeX));
    const uint64_t ny = (uint64_t)std::ceil(((float)i.nY / c.tileSizeY));
    const uint64_t nz = (uint64_t)std::ceil(((float)i.nZ / c.tileSizeZ));
    const uint64_t nt = (uint64_t)std::ceil(((float)i.nTS / c.tileSizeTS));
    c.nTiles = nx * ny * nz * nt;
  }
  c.CSsgc |= (0x01 << 9);
}

uchar create_codec(bwc_codec *const codec) {
  if (!codec->tile) codec->tile = new Job();
  Job *job = job_of(codec);
  // the gain table is read when the geometry is built, whatever was
  // accumulated into it by then (libbwc.c:450 via create_codec)
  std::memcpy(job->lut, gain_lut(), sizeof job->lut);
  job->plan = Plan();
  const int rc = build_plan(codec, job);
  if (rc == 2)
    set_err("a tile bound above 2^24 rounds in single precision so that the resolution boxes "
            "(libbwc.c:1539-1546) would reach past the tile's samples; such extents are not supported");
  if (rc) return 1;
  job->created = true;
  return 0;
}

// bwc_create_compression (libbwc.c:2970-3131)
uchar bwc_create_compression(bwc_codec *const codec, bwc_stream *const stream, char *const rate_control) {
  bwc_gl_ctrl &c = codec->control;
  std::vector<float> br;
  if (rate_control == nullptr) {
    br.push_back(kPrecBits + 1);
  } else {
    char *save = nullptr;
    char *tok = strtok_r(rate_control, "-, ", &save);
    if (!tok) br.push_back(kPrecBits + 1);
    for (; tok; tok = strtok_r(nullptr, "-, ", &save)) {
      const float f = std::strtof(tok, nullptr);
      if (f <= 0 || f > 64) {
        std::fprintf(stderr, "bwc: error: invalid bitrate value %f; valid range 0 < bitrate <= 64\n", f);
        return 1;
      }
      // sorted insert, dropping values within 0.01 of an existing one
      int l = (int)br.size();
      br.push_back(0.0f);
      for (; l >= 0; --l) {
        if (l && f < br[l - 1]) {
          br[l] = br[l - 1];
        } else if (l && std::fabs(f - br[l - 1]) < 0.01f) {
          for (; l < (int)br.size() - 1; ++l) br[l] = br[l + 1];
          br.pop_back();
          break;
        } else {
          br[l] = f;
          break;
        }
      }
    }
  }
  std::free(c.bitrate);
  c.bitrate = static_cast<float *>(std::malloc(br.size() * sizeof(float)));
  std::memcpy(c.bitrate, br.data(), br.size() * sizeof(float));
  c.nLayers = (uint8_t)br.size();
  if (create_codec(codec)) return 1;
  c.headerSize = header_size(codec, stream);
  if (c.headerSize >= 0xFFFFFFFF) {
    set_err("main header exceeds the maximum size of 4294967295 bytes");
    return 1;
  }
  return 0;
}

// bwc_compress (libbwc.c:3145-3438)
size_t bwc_compress(bwc_codec *const codec, bwc_stream *const stream) {
  Job *job = job_of(codec);
  if (!job || !job->created) {
    set_err("bwc_compress called before bwc_create_compression");
    return 0;
  }
  gpu::Workspace *ws = gpu::workspace_of(job);
  if (!ws) {
    set_err(gpu::last_error());
    return 0;
  }
  bwc_gl_ctrl &ctl = codec->control;
  const bool rc_dev = rc_on_device();
  const double t_start = now_ms();
  // a call of a batch's global codeblock queue that ends before its tier-1
  // joined the queue still counts as joined (with no codeblocks)
  struct LeaveBatch {
    gpu::Workspace *ws;
    ~LeaveBatch() { gpu::batch_leave(ws); }
  } leave_batch{ws};
  reset_for_encode(codec, job);
  std::fill(std::begin(job->skip_stats), std::end(job->skip_stats), 0);
  const bool on_dev = gpu::is_device_pointer(stream->inp) != 0;
  uint64_t maxsz = 0;
  for (auto &tl : job->plan.tiles) maxsz = std::max<uint64_t>(maxsz, tl.box.volume());
  if (gpu::ensure_work(ws, maxsz)) {
    set_err(gpu::last_error());
    return 0;
  }
  double ms[10] = {0};
  int dwt_launch = 0;
  uint64_t body_end = 0;   // packet bodies gathered so far (device body buffer)
  BodySource bodies;
  bodies.ctx = ws;
  bodies.copy = [](void *w, uint8_t *dst, uint64_t off, uint64_t n) {
    return gpu::body_d2h(static_cast<gpu::Workspace *>(w), dst, off, n);
  };
  for (uint64_t ti = 0; ti < job->plan.tiles.size(); ++ti) {
    Tile &tl = job->plan.tiles[ti];
    for (int p = 0; p < codec->info.nPar; ++p) {
      Param &par = tl.par[p];
      gpu::record(ws, 0);
      if (gpu::fill_minmax(ws, codec, tl, p, stream->inp, on_dev, ws->d_work)) goto fail;
      gpu::record(ws, 1);
      // normalisation is fused into the DWT's first pass (stage time 0 here)
      // With pipeline graphs the DWT is launched as the head of tier-1's
      // device chain (recorded with it; its stage time is then counted in
      // tier-1's), otherwise on its own.
      const bool chain = gpu::graphs_on() && !job->collect_features && !ws->batch;
      const std::function<int()> dwt_launch_fn = [&]() {
        return gpu::dwt_forward_direct(ws, codec, par, ws->d_work, ctl.Qm);
      };
      gpu::record(ws, 2);
      if (!chain && gpu::dwt_forward(ws, codec, par, ws->d_work, ctl.Qm)) goto fail;
      gpu::record(ws, 3);
      {
        CblkRecords &R = job->rec[ti * codec->info.nPar + p];
        std::vector<uint8_t> zl(par.ncb, 255);
        gpu::PassLimits pl;
        if (job->skip_mode == 1) {
          std::fill(zl.begin(), zl.end(), (uint8_t)std::min(255, std::max(1, job->skip_param)));
          pl.fixed = zl.data();
        } else if (job->skip_mode == 2) {
          pl.model = job->skip_w.data();
          pl.margin = job->skip_param;
          pl.out = zl.data();
        }
        pl.features = job->collect_features;
        pl.records = !rc_dev;
        if (chain) {
          pl.pre = &dwt_launch_fn;
          pl.pre_key = gpu::dwt_launch_key(ws, codec, par, ws->d_work, ctl.Qm) | 1;
        }
        gpu::record(ws, 4);
        if (gpu::t1_encode(ws, codec, par, ws->d_work, R, &tl.slope_min, &tl.slope_max, p, pl)) goto fail;
        {
          double mm[4];
          if (gpu::read_norm(ws, mm)) goto fail;
          par.pmin = mm[0];
          par.pmax = mm[1];
          par.alpha = mm[2];
          par.beta = mm[3];
        }
        R.zlim = zl;
        if (job->collect_features) {
          R.feat.resize(par.ncb * 16);
          if (par.ncb && gpu::last_features(ws, par.ncb, R.feat.data())) goto fail;
        }
        for (uint64_t i = 0; i < R.n; ++i) {
          if (!R.K[i]) continue;
          job->skip_stats[0] += 1;
          job->skip_stats[1] += R.Z[i];
          job->skip_stats[2] += R.Zfull[i];
        }
      }
      gpu::record(ws, 5);
      gpu::sync(ws);
      ms[0] += gpu::elapsed(ws, 0, 1);
      ms[1] += gpu::elapsed(ws, 1, 2);
      ms[2] += gpu::elapsed(ws, 2, 3);
      dwt_launch += ws->last_dwt_launches;
      ms[4] += gpu::elapsed(ws, 4, 5);
    }
    {
      const double t0 = now_ms();
      if (rc_dev) {
        if (rate_control_dev(codec, job, ws, ti)) {
          ms[6] += now_ms() - t0;
          return 0;
        }
      } else if (job->skip_mode) {
        if (rate_control_prefix(codec, job, ws, ti)) goto fail;
      } else {
        t2_rate_control(codec, job, ti);
      }
      ms[6] += now_ms() - t0;
      // packet headers, then every packet body gathered on the device into
      // one contiguous range of the body buffer (the parameter slots are
      // coded over by the next tile)
      // (the gather runs on the GPU while the host writes the headers)
      const double t1 = now_ms();
      if (gather_tile_bodies(codec, job, ws, ti, body_end)) goto fail;
      const double t2 = now_ms();
      t2_packets(codec, job, ti);
      ms[7] += now_ms() - t1;
      if (trace_on())
        std::fprintf(stderr, "t2: body gather %8.3f ms, packet headers %8.3f ms\n", t2 - t1, now_ms() - t2);
    }
  }
  {
    const double t3 = now_ms();
    size_t n = assemble_codestream(codec, job, stream, static_cast<uint8_t *>(stream->out), &bodies);
    if (gpu::sync(ws)) n = 0;
    ms[8] += now_ms() - t3;
    ms[9] = now_ms() - t_start;
    std::memcpy(job->ms, ms, sizeof ms);
    job->launches[0] = dwt_launch;
    if (!n) set_err("codestream assembly failed");
    return n;
  }
fail:
  set_err(gpu::last_error());
  return 0;
}

// bwc_create_decompression (libbwc.c:3453-3470): the codestream's header
// configures the codec, its packets become per-codeblock decode records
uchar bwc_create_decompression(bwc_codec *const codec, bwc_stream *const stream, uint8 const layer) {
  if (!codec || !stream || !stream->inp) {
    set_err("bwc_create_decompression: codec and codestream required");
    return 1;
  }
  const double t0 = now_ms();
  if (parse_codestream(codec, stream, layer)) {
    set_err("invalid codestream");
    return 1;
  }
  if (trace_on()) std::fprintf(stderr, "dec: parse codestream      %8.3f ms\n", now_ms() - t0);
  return 0;
}

// bwc_decompress (libbwc.c:3485-3710): per tile and parameter, tier-1
// decoding, inverse DWT, denormalisation and output, all on the GPU; the
// field goes to stream->out (host or device memory)
uchar bwc_decompress(bwc_codec *const codec, bwc_stream *const stream) {
  Job *job = job_of(codec);
  if (!job || !job->decodable || !stream || !stream->out) {
    set_err("bwc_decompress called before bwc_create_decompression");
    return 1;
  }
  gpu::Workspace *ws = gpu::workspace_of(job);
  if (!ws) {
    set_err(gpu::last_error());
    return 1;
  }
  const bwc_gl_inf &inf = codec->info;
  const size_t bytes = (size_t)(inf.nX * inf.nY * inf.nZ * inf.nTS) * inf.nPar * (size_t)inf.data_prec;
  const bool on_dev = gpu::is_device_pointer(stream->out) != 0;
  void *dst = stream->out;
  uint64_t maxsz = 0;
  for (auto &tl : job->plan.tiles) maxsz = std::max<uint64_t>(maxsz, tl.box.volume());
  if (gpu::ensure_work(ws, maxsz)) goto fail;
  if (!on_dev && gpu::ensure_output(ws, bytes, &dst)) goto fail;
  for (uint64_t ti = 0; ti < job->plan.tiles.size(); ++ti) {
    Tile &tl = job->plan.tiles[ti];
    for (int p = 0; p < inf.nPar; ++p) {
      Param &par = tl.par[p];
      const double t0 = now_ms();
      if (gpu::t1_decode(ws, codec, par, job->drec[ti * inf.nPar + p], ws->d_work)) goto fail;
      const double t1 = now_ms();
      if (gpu::dwt_inverse(ws, codec, par, ws->d_work)) goto fail;
      if (gpu::decode_output(ws, codec, tl, p, par, ws->d_work, dst)) goto fail;
      if (trace_on()) {
        gpu::sync(ws);
        std::fprintf(stderr, "dec: tier-1 (host + kernel) %8.3f ms, inverse DWT + output %8.3f ms\n", t1 - t0,
                     now_ms() - t1);
      }
    }
  }
  // host output: only the samples a tile wrote cross back (the others keep
  // the caller's contents, as the reference leaves them)
  if (!on_dev)
    for (uint64_t ti = 0; ti < job->plan.tiles.size(); ++ti)
      for (int p = 0; p < inf.nPar; ++p)
        if (gpu::output_box_d2h(ws, codec, job->plan.tiles[ti], p, job->plan.tiles[ti].par[p], dst, stream->out))
          goto fail;
  if (gpu::sync(ws)) goto fail;
  return 0;
fail:
  set_err(gpu::last_error());
  return 1;
}

// bwc_open_header (libbwc.c:2840-2935): the main header alone
bwc_header *bwc_open_header(void *const inpbuf) {
  if (!inpbuf) return nullptr;
  bwc_codec *codec = bwc_alloc_decoder();
  bwc_stream *data = bwc_init_stream(inpbuf, nullptr, decomp);
  bwc_header *h = nullptr;
  if (codec && data && !parse_header_only(codec, data)) {
    h = static_cast<bwc_header *>(std::calloc(1, sizeof(bwc_header)));
    if (h) {
      h->info = codec->info;
      h->control = codec->control;
      h->control.bitrate = static_cast<float *>(std::calloc(codec->control.nLayers ? codec->control.nLayers : 1,
                                                            sizeof(float)));
      if (codec->control.nLayers)
        std::memcpy(h->control.bitrate, codec->control.bitrate, codec->control.nLayers * sizeof(float));
      if (data->codestream.aux) {
        h->aux.memory = data->codestream.aux->memory;
        h->aux.size = data->codestream.aux->size;
        std::free(data->codestream.aux);
      }
      if (data->codestream.com) {
        h->com.memory = data->codestream.com->memory;
        h->com.size = data->codestream.com->size;
        std::free(data->codestream.com);
      }
    }
  } else {
    set_err("invalid codestream header");
  }
  bwc_free_codec(codec);
  std::free(data);
  return h;
}

void bwc_close_header(bwc_header *const header) {
  if (!header) return;
  std::free(header->aux.memory);
  std::free(header->com.memory);
  std::free(header->control.bitrate);
  std::free(header);
}

// ---------------------------------------------------------------------------
// bwc_amd.h
// ---------------------------------------------------------------------------
int bwc_amd_version(void) { return gpu::device_count() > 0 ? 1 : 0; }

void bwc_amd_reset_gain_lut(void) { reset_gain_lut(); }
int bwc_amd_device_count(void) { return gpu::device_count(); }
void *bwc_amd_host_alloc(size_t bytes) { return gpu::host_alloc(bytes); }
void bwc_amd_host_free(void *p) { gpu::host_free(p); }
const char *bwc_amd_last_error(void) { return g_api_err[0] ? g_api_err : gpu::last_error(); }

int bwc_amd_compress_batch(bwc_codec *const *codecs, bwc_stream *const *streams, int n, size_t *sizes,
                           int workers) {
  if (n <= 0) return 0;
  if (!codecs || !streams || !sizes) {
    set_err("bwc_amd_compress_batch: null argument");
    return 1;
  }
  {  // one codec's workspace and records serve one compression at a time;
     // two workers must not write one stream or one output buffer
    std::vector<const void *> seen[3];
    for (int i = 0; i < n; ++i) {
      seen[0].push_back(codecs[i]);
      seen[1].push_back(streams[i]);
      if (streams[i]) seen[2].push_back(streams[i]->out);
    }
    static const char *what[3] = {"a codec", "a stream", "an output buffer"};
    for (int k = 0; k < 3; ++k) {
      std::sort(seen[k].begin(), seen[k].end());
      if (std::adjacent_find(seen[k].begin(), seen[k].end()) != seen[k].end()) {
        std::snprintf(g_api_err, sizeof g_api_err, "bwc_amd_compress_batch: %s appears twice", what[k]);
        std::fprintf(stderr, "bwc: error: %s\n", g_api_err);
        return 1;
      }
    }
  }
  // workspaces (device state, streams) made here, on the calling thread
  for (int i = 0; i < n; ++i) {
    sizes[i] = 0;
    Job *job = codecs[i] ? job_of(codecs[i]) : nullptr;
    if (!job || !job->created || !streams[i]) {
      set_err("bwc_amd_compress_batch: a codec is not ready for compression");
      return 1;
    }
    if (!gpu::workspace_of(job)) {
      set_err(gpu::last_error());
      return 1;
    }
  }
  if (workers <= 0) workers = 8;
  workers = std::min(workers, n);
  int device = 0;
  (void)hipGetDevice(&device);
  // Fields of one tile and one variable code their tier-1 through one global
  // codeblock queue (BWC_AMD_BATCH=threads: each field on its own): every
  // such field's compression runs on a thread of its own up to its codeblock
  // list, the last to get there launches symbol generation and MQ coding
  // once over all of them, and each goes on with its rate control.
  const char *bm = std::getenv("BWC_AMD_BATCH");
  const bool queue = !(bm && std::strcmp(bm, "threads") == 0);
  std::vector<int> in_queue;
  uint64_t items = 0;
  for (int i = 0; queue && i < n; ++i) {
    Job *job = job_of(codecs[i]);
    if (job->plan.tiles.size() == 1 && codecs[i]->info.nPar == 1) {
      in_queue.push_back(i);
      items += job->plan.tiles[0].par[0].ncb;
    }
  }
  gpu::BatchQueue *bq = nullptr;
  const int slots = workers;
  if (in_queue.size() > 1 && in_queue.size() <= 512) {
    bq = gpu::batch_create((int)in_queue.size(), items, slots);
    if (!bq) {
      set_err(gpu::last_error());
      return 1;
    }
    for (int i : in_queue) {
      gpu::Workspace *ws = gpu::workspace_of(job_of(codecs[i]));
      ws->batch = bq;
      ws->batch_joined = false;
    }
    // every queued field needs a thread of its own until the launch; after it
    // `workers` of them go on at a time
    workers = std::max(workers, (int)in_queue.size());
  }
  std::atomic<int> next{0};
  std::atomic<bool> failed{false};
  std::mutex em;
  std::string first_err;
  auto work = [&] {
    (void)hipSetDevice(device);
    gpu::set_thread_graphs(false);
    for (int i; !failed.load() && (i = next.fetch_add(1)) < n;) {
      sizes[i] = bwc_compress(codecs[i], streams[i]);
      if (!sizes[i]) {
        failed.store(true);   // no further codec is started
        std::lock_guard<std::mutex> lk(em);
        if (first_err.empty()) first_err = bwc_amd_last_error();
      }
    }
  };
  std::vector<std::thread> th;
  if (bq) {
    // queued fields first, one per thread, then the others on the same threads
    std::vector<int> order = in_queue;
    for (int i = 0; i < n; ++i)
      if (std::find(in_queue.begin(), in_queue.end(), i) == in_queue.end()) order.push_back(i);
    auto qwork = [&](int first) {
      (void)hipSetDevice(device);
      gpu::set_thread_graphs(false);
      for (size_t k = (size_t)first; k < order.size(); k += (size_t)workers) {
        const int i = order[k];
        gpu::Workspace *ws = gpu::workspace_of(job_of(codecs[i]));
        if (failed.load() && !(ws->batch && !ws->batch_joined)) continue;
        sizes[i] = bwc_compress(codecs[i], streams[i]);
        if (!sizes[i]) {
          failed.store(true);
          std::lock_guard<std::mutex> lk(em);
          if (first_err.empty()) first_err = bwc_amd_last_error();
        }
      }
    };
    for (int w = 1; w < workers; ++w) th.emplace_back(qwork, w);
    qwork(0);
    gpu::set_thread_graphs(true);
    for (auto &t : th) t.join();
    for (int i : in_queue) {
      gpu::Workspace *ws = gpu::workspace_of(job_of(codecs[i]));
      ws->batch = nullptr;
      ws->batch_joined = false;
    }
    if (gpu::batch_destroy(bq) && first_err.empty()) first_err = "batch queue failed";
  } else {
    for (int w = 1; w < workers; ++w) th.emplace_back(work);
    work();
    gpu::set_thread_graphs(true);
    for (auto &t : th) t.join();
  }
  if (!first_err.empty()) {
    std::snprintf(g_api_err, sizeof g_api_err, "%s", first_err.c_str());
    return 1;
  }
  return 0;
}

int bwc_amd_set_tuning(const char *name, int value) {
  if (set_tuning(name, value)) {
    set_err("unknown tuning knob or value out of range");
    return 1;
  }
  return 0;
}
int bwc_amd_get_tuning(const char *name, int *value) {
  if (get_tuning(name, value)) {
    set_err("unknown tuning knob");
    return 1;
  }
  return 0;
}
const char *bwc_amd_tuning_source(void) { return tuning_source(); }

uint64_t bwc_amd_ncblk(bwc_codec *codec, uint64_t tile, uint8_t param) {
  Job *job = job_of(codec);
  if (!job || !job->created || tile >= job->plan.tiles.size() || param >= codec->info.nPar) return 0;
  return job->plan.tiles[tile].par[param].ncb;
}

int bwc_amd_cblk_info(bwc_codec *codec, uint64_t tile, uint8_t param, uint32_t *box, double *sb, uint32_t *meta) {
  Job *job = job_of(codec);
  if (!job || !job->created || tile >= job->plan.tiles.size() || param >= codec->info.nPar) return 1;
  const Param &par = job->plan.tiles[tile].par[param];
  for (uint64_t i = 0; i < par.ncb; ++i) {
    for (int k = 0; k < 8; ++k) box[8 * i + k] = par.cb_box[8 * i + k];
    const Subband &s = par.res[par.cb_sbref[i] >> 8].sb[par.cb_sbref[i] & 0xFF];
    sb[3 * i] = s.gain;
    sb[3 * i + 1] = s.step;
    sb[3 * i + 2] = s.step_eff;
    meta[4 * i] = s.hb;
    meta[4 * i + 1] = s.Kmax;
    meta[4 * i + 2] = s.qt_exp;
    meta[4 * i + 3] = par.cb_idx[i];
  }
  return 0;
}

int bwc_amd_param_dims(bwc_codec *codec, uint64_t tile, uint64_t dims[4]) {
  Job *job = job_of(codec);
  if (!job || !job->created || tile >= job->plan.tiles.size()) return 1;
  for (int a = 0; a < 4; ++a) dims[a] = job->plan.tiles[tile].box.len(a);
  return 0;
}

int bwc_amd_stage_dwt(bwc_codec *codec, uint64_t tile, uint8_t param, const void *input, double *out,
                      double mm[4]) {
  Job *job = job_of(codec);
  if (!job || !job->created || tile >= job->plan.tiles.size() || param >= codec->info.nPar) return 1;
  gpu::Workspace *ws = gpu::workspace_of(job);
  if (!ws) return 1;
  Tile &tl = job->plan.tiles[tile];
  Param &par = tl.par[param];
  const uint64_t n = par.box.volume();
  if (gpu::ensure_work(ws, n)) return 1;
  if (gpu::fill_minmax(ws, codec, tl, param, input, gpu::is_device_pointer(input) != 0, ws->d_work)) return 1;
  if (gpu::dwt_forward(ws, codec, par, ws->d_work, codec->control.Qm)) return 1;
  if (gpu::read_norm(ws, mm)) return 1;
  return gpu::copy_d2h(ws, out, ws->d_work, n * sizeof(double));
}

int bwc_amd_stage_dwt_noalloc(bwc_codec *codec, uint64_t tile, uint8_t param, const void *input, double *out,
                              double *norm, void *stream) {
  Job *job = job_of(codec);
  if (!job || !job->created || tile >= job->plan.tiles.size() || param >= codec->info.nPar) {
    set_err("bwc_amd_stage_dwt_noalloc: codec not created or tile / parameter out of range");
    return 1;
  }
  if (!input || !out || !norm) {
    set_err("bwc_amd_stage_dwt_noalloc: null buffer");
    return 1;
  }
  if (gpu::dwt_noalloc(codec, job->plan.tiles[tile], param, input, out, norm, static_cast<hipStream_t>(stream))) {
    set_err(gpu::last_error());
    return 1;
  }
  return 0;
}

int bwc_amd_stage_idwt(bwc_codec *codec, uint64_t tile, uint8_t param, double *coeffs) {
  Job *job = job_of(codec);
  if (!job || !job->created || tile >= job->plan.tiles.size() || param >= codec->info.nPar) return 1;
  gpu::Workspace *ws = gpu::workspace_of(job);
  if (!ws) return 1;
  Param &par = job->plan.tiles[tile].par[param];
  const uint64_t n = par.box.volume();
  if (gpu::ensure_work(ws, n)) return 1;
  if (gpu::copy_h2d(ws, ws->d_work, coeffs, n * sizeof(double))) return 1;
  if (gpu::dwt_inverse(ws, codec, par, ws->d_work)) return 1;
  return gpu::copy_d2h(ws, coeffs, ws->d_work, n * sizeof(double));
}

int bwc_amd_stage_t1(bwc_codec *codec, uint64_t tile, uint8_t param, const double *coeffs, uint8_t *K,
                     uint8_t *Kmsbs, uint8_t *Z, uint32_t *L, uint16_t *S, uint8_t *data, uint64_t data_cap,
                     uint64_t *data_off, uint16_t slopes[2]) {
  Job *job = job_of(codec);
  if (!job || !job->created || tile >= job->plan.tiles.size() || param >= codec->info.nPar) return 1;
  gpu::Workspace *ws = gpu::workspace_of(job);
  if (!ws) return 1;
  Param &par = job->plan.tiles[tile].par[param];
  const uint64_t n = par.box.volume();
  if (gpu::ensure_work(ws, n)) return 1;
  if (gpu::copy_h2d(ws, ws->d_work, coeffs, n * sizeof(double))) return 1;
  CblkRecords R;
  uint16_t smin = 0xFFFF, smax = 0;
  if (gpu::t1_encode(ws, codec, par, ws->d_work, R, &smin, &smax, 0)) return 1;
  std::vector<uint32_t> len(R.n);
  uint64_t tot = 0;
  for (uint64_t i = 0; i < R.n; ++i) {
    len[i] = R.Z[i] ? R.L[i * kMaxPasses + R.Z[i] - 1] : 0;
    tot += len[i];
  }
  if (tot > data_cap) {
    set_err("data buffer too small");
    return 2;
  }
  if (gpu::gather(ws, R, len)) return 1;
  std::memcpy(K, R.K.data(), R.n);
  std::memcpy(Kmsbs, R.Kmsbs.data(), R.n);
  std::memcpy(Z, R.Z.data(), R.n);
  std::memcpy(L, R.L.data(), R.n * kMaxPasses * sizeof(uint32_t));
  std::memcpy(S, R.S.data(), R.n * (kMaxPasses + 1) * sizeof(uint16_t));
  if (tot) std::memcpy(data, R.data.data(), tot);
  for (uint64_t i = 0; i < R.n; ++i) data_off[i] = R.data_off[i];
  data_off[R.n] = tot;
  slopes[0] = smin;
  slopes[1] = smax;
  return 0;
}

int bwc_amd_set_records(bwc_codec *codec, uint64_t tile, uint8_t param, const uint8_t *K, const uint8_t *Kmsbs,
                        const uint8_t *Z, const uint32_t *L, const uint16_t *S, const uint8_t *data,
                        const uint64_t *data_off, const double mm[4]) {
  Job *job = job_of(codec);
  if (!job || !job->created || tile >= job->plan.tiles.size() || param >= codec->info.nPar) return 1;
  if (job->rec.size() != job->plan.tiles.size() * codec->info.nPar) reset_for_encode(codec, job);
  Tile &tl = job->plan.tiles[tile];
  Param &par = tl.par[param];
  CblkRecords &R = job->rec[tile * codec->info.nPar + param];
  R.resize(par.ncb, codec->control.nLayers);
  std::memcpy(R.K.data(), K, par.ncb);
  std::memcpy(R.Kmsbs.data(), Kmsbs, par.ncb);
  std::memcpy(R.Z.data(), Z, par.ncb);
  std::memcpy(R.Zfull.data(), Z, par.ncb);   // complete codings
  std::memcpy(R.U.data(), Z, par.ncb);
  std::memcpy(R.L.data(), L, par.ncb * kMaxPasses * sizeof(uint32_t));
  std::memcpy(R.S.data(), S, par.ncb * (kMaxPasses + 1) * sizeof(uint16_t));
  R.data.assign(data, data + data_off[par.ncb]);
  for (uint64_t i = 0; i < par.ncb; ++i) {
    R.data_off[i] = data_off[i];
    if (R.K[i] == 0) continue;
    for (int z = 0; z <= R.Z[i]; ++z) {
      const uint16_t s = R.S[i * (kMaxPasses + 1) + z];
      if (s) {
        tl.slope_min = std::min(tl.slope_min, s);
        tl.slope_max = std::max(tl.slope_max, s);
      }
    }
  }
  par.pmin = mm[0], par.pmax = mm[1], par.alpha = mm[2], par.beta = mm[3];
  return 0;
}

size_t bwc_amd_encode_records(bwc_codec *codec, bwc_stream *stream) {
  Job *job = job_of(codec);
  if (!job || !job->created || job->rec.size() != job->plan.tiles.size() * codec->info.nPar) return 0;
  for (uint64_t ti = 0; ti < job->plan.tiles.size(); ++ti) {
    t2_rate_control(codec, job, ti);
    t2_packets(codec, job, ti);
  }
  return assemble_codestream(codec, job, stream, static_cast<uint8_t *>(stream->out));
}

void bwc_amd_begin_records(bwc_codec *codec) {
  Job *job = job_of(codec);
  if (job && job->created) reset_for_encode(codec, job);
}

int bwc_amd_last_timings(bwc_codec *codec, double ms[10], int launches[4]) {
  Job *job = job_of(codec);
  if (!job) return 1;
  std::memcpy(ms, job->ms, sizeof job->ms);
  std::memcpy(launches, job->launches, sizeof job->launches);
  return 0;
}

int bwc_amd_set_pass_skip(bwc_codec *codec, int mode, int param) {
  Job *job = job_of(codec);
  if (!job || mode < 0 || mode > 2 || param < 0 || param > 255) {
    set_err("bwc_amd_set_pass_skip: invalid codec, mode or parameter");
    return 1;
  }
  if (mode == 2 && job->skip_w.empty()) {
    set_err("bwc_amd_set_pass_skip: mode 2 needs bwc_amd_set_skip_model first");
    return 1;
  }
  job->skip_mode = mode;
  job->skip_param = param;
  return 0;
}

int bwc_amd_set_skip_model(bwc_codec *codec, const float *w, size_t n) {
  Job *job = job_of(codec);
  if (!job || !w || n != (size_t)kSkipModelSize) {
    set_err("bwc_amd_set_skip_model: invalid codec or weight count");
    return 1;
  }
  job->skip_w.assign(w, w + n);
  return 0;
}

int bwc_amd_skip_stats(bwc_codec *codec, uint64_t stats[8]) {
  Job *job = job_of(codec);
  if (!job) return 1;
  std::memcpy(stats, job->skip_stats, sizeof job->skip_stats);
  return 0;
}

int bwc_amd_collect_features(bwc_codec *codec, int on) {
  Job *job = job_of(codec);
  if (!job) return 1;
  job->collect_features = on != 0;
  return 0;
}

int bwc_amd_cblk_features(bwc_codec *codec, uint64_t tile, uint8_t param, float *feat) {
  Job *job = job_of(codec);
  if (!job || tile >= job->plan.tiles.size() || param >= codec->info.nPar) return 1;
  const uint64_t r = tile * codec->info.nPar + param;
  if (r >= job->rec.size() || job->rec[r].feat.size() != job->rec[r].n * 16) {
    set_err("no features collected for this tile and parameter");
    return 1;
  }
  std::memcpy(feat, job->rec[r].feat.data(), job->rec[r].feat.size() * sizeof(float));
  return 0;
}

int bwc_amd_cblk_decisions(bwc_codec *codec, uint64_t tile, uint8_t param, uint32_t *out) {
  Job *job = job_of(codec);
  if (!job || tile >= job->plan.tiles.size() || param >= codec->info.nPar) return 1;
  const uint64_t r = tile * codec->info.nPar + param;
  if (r >= job->rec.size() || job->rec[r].decisions.size() != job->rec[r].n) {
    set_err("no decision counts collected for this tile and parameter");
    return 1;
  }
  std::memcpy(out, job->rec[r].decisions.data(), job->rec[r].decisions.size() * sizeof(uint32_t));
  return 0;
}

int bwc_amd_skipnet_train(const float *feat, const float *target, uint64_t n, float *weights, int epochs, float lr,
                          uint64_t seed, float *loss) {
  char e[256] = {0};
  if (gpu::skipnet_fit(feat, target, n, weights, epochs, lr, seed, loss, e, sizeof e)) {
    set_err(e);
    return 1;
  }
  return 0;
}

int bwc_amd_skipnet_grad(const float *feat, const float *target, uint64_t n, const float *weights, float *grad,
                         float *loss) {
  char e[256] = {0};
  if (gpu::skipnet_grad(feat, target, n, weights, grad, loss, e, sizeof e)) {
    set_err(e);
    return 1;
  }
  return 0;
}

int bwc_amd_skipnet_predict(const float *feat, uint64_t n, const float *weights, float *y) {
  char e[256] = {0};
  if (gpu::skipnet_predict(feat, n, weights, y, e, sizeof e)) {
    set_err(e);
    return 1;
  }
  return 0;
}

int bwc_amd_stuffed_bytes(const uint32_t *words, const uint64_t *woff, const uint64_t *nbits, int n,
                          uint64_t *bytes) {
  if (gpu::rc_stuffed_bytes(words, woff, nbits, n, bytes)) {
    set_err(gpu::rc_last_error());
    return 1;
  }
  return 0;
}

uint64_t bwc_amd_batch_queued(bwc_codec *codec) {
  Job *job = job_of(codec);
  if (!job || !job->ws) return 0;
  return static_cast<const gpu::Workspace *>(job->ws)->batch_queued;
}

int bwc_amd_graph_counts(bwc_codec *codec, uint64_t counts[2]) {
  Job *job = job_of(codec);
  if (!job || !job->ws) return 1;
  const gpu::Workspace *ws = static_cast<const gpu::Workspace *>(job->ws);
  counts[0] = ws->graph_records;
  counts[1] = ws->graph_replays;
  return 0;
}

int bwc_amd_rc_trace(bwc_codec *codec, uint32_t *layer, uint32_t *slope, int64_t *est, int cap) {
  Job *job = job_of(codec);
  if (!job || !job->ws) return -1;
  return gpu::rc_trace(static_cast<gpu::Workspace *>(job->ws), layer, slope, est, cap);
}

int bwc_amd_cblk_cuts(bwc_codec *codec, uint64_t tile, uint8_t param, int16_t *cuts) {
  Job *job = job_of(codec);
  if (!job || tile >= job->plan.tiles.size() || param >= codec->info.nPar) return 1;
  const uint64_t r = tile * codec->info.nPar + param;
  if (r >= job->rec.size()) return 1;
  const CblkRecords &R = job->rec[r];
  const int nl = codec->control.nLayers;
  for (uint64_t i = 0; i < R.n; ++i) cuts[i] = R.cp.empty() ? 0 : R.cp[i * (nl + 1) + nl];
  return 0;
}

}  // extern "C"
