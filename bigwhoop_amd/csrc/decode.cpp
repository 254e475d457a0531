// Decompression, host side: the main header into the codec, and the
// packets of every tile into per-codeblock decode records (Kmsbs, passes,
// coded bytes). Tier-1 decoding, the inverse DWT and the output run on the
// GPU (gpu.hip, driven from api.cpp).
//
// Reference: parse_codestream / parse_main_header / parse_tile / parse_body
// (codestream.c:601-1260, :1448-1510), parse_packet with decode_delta_z and
// decode_length (tier2.c:155-200, :344-375, :1364-1634), decode_tagtree
// (tagtree.c:387-430), the bit reader (bitstream.c:80-90, :452-545).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "codec.hpp"

namespace bwc {

namespace {

// byte / bit reader over a codestream. Lmax is the limit the reference's
// reader checks (for the whole stream the size declared in the header, for a
// packet header the tile's remaining declared body); cap is the number of
// bytes actually there, past which nothing is read.
struct Reader {
  const uint8_t *m;
  uint64_t L = 0, Lmax = 0, cap = 0;
  uint16_t T = 0;
  uint8_t t = 0;
  bool err = false;
  uint64_t used() const { return T == 0xFF ? L + 1 : L; }
  bool can(uint64_t n) const { return L + n <= Lmax; }
  uint64_t sym(int n) {
    if (n > 8 || used() + (uint64_t)n > Lmax || L + (uint64_t)n > cap) {
      err = true;
      return 0;
    }
    uint64_t v = 0;
    for (int i = n; i-- > 0;) v |= (uint64_t)m[L++] << (8 * i);
    return v;
  }
  uint8_t bit() {
    if (!t) {
      if (L == Lmax || L >= cap) return 1;
      t = (T == 0xFF) ? 7 : 8;
      T = m[L++];
    }
    t--;
    return (uint8_t)((T >> t) & 1);
  }
};

// decode_tagtree: 0 when the leaf's value is below threshold
int tagtree_decode(TagTree &tt, Reader &r, uint32_t threshold, uint64_t leaf) {
  int64_t path[64];
  int np = 0;
  for (int64_t n = (int64_t)leaf; n >= 0; n = tt.parent[n]) {
    if (tt.value[n] == 0xFFFF) tt.value[n] = 0;
    path[np++] = n;
  }
  uint32_t tmin = 0;
  int64_t n = (int64_t)leaf;
  for (int a = np - 1; a >= 0; --a) {
    n = path[a];
    if (tt.threshold[n] < tmin) {
      tt.value[n] = (uint16_t)tmin;
      tt.threshold[n] = (uint16_t)tmin;
    }
    while (tt.value[n] == tt.threshold[n] && tt.threshold[n] < threshold) {
      tt.threshold[n]++;
      if (!r.bit()) tt.value[n]++;
    }
    tmin = std::min<uint32_t>(tt.value[n], tt.threshold[n]);
  }
  return tt.value[n] < threshold ? 0 : 1;
}

// main header (codestream.c:601-1053): configure the codec as the stream
// says, in the order of the reference's setter calls
int parse_main_header(bwc_codec *codec, bwc_stream *data, Reader &r) {
  bwc_gl_ctrl &c = codec->control;
  bool sgi = false, sgc = false;
  uint8_t codec_prec = 8;
  for (;;) {
    if (!r.can(2)) return 1;
    const uint16_t marker = (uint16_t)r.sym(2);
    if (marker < 0xFF00) return 1;
    switch (marker) {
      case 0xFF50:   // SOC: the declared codestream size
        r.Lmax = r.sym(8);
        break;
      case 0xFF51: {   // SGI
        r.sym(2);
        uint64_t n[4];
        for (auto &v : n) v = r.sym(8);
        const uint8_t npar = (uint8_t)r.sym(1);
        const uint8_t dprec = (uint8_t)r.sym(1);
        codec_prec = (uint8_t)r.sym(1);
        if (r.err || !npar || (dprec != 4 && dprec != 8)) return 1;
        codec->info.codec_prec = (bwc_precision)codec_prec;
        configure_codec(codec, n[0], n[1], n[2], n[3], npar, (bwc_precision)dprec);
        c.codestreamSize = r.Lmax;
        sgi = true;
        break;
      }
      case 0xFF52: {   // SGC
        if (!sgi) return 1;
        r.sym(2);
        const uint16_t cs = (uint16_t)r.sym(2);
        r.sym(1);
        if (cs & (1 << 0)) bwc_set_error_resilience(codec);
        uint64_t v = r.sym(1);
        if (cs & (1 << 1)) set_quant_style(codec, (bwc_quant_st)v);
        r.sym(1);   // guard bits: not read back
        const uint32_t qe = (uint32_t)r.sym(1), qm = (uint32_t)r.sym(2);
        if (cs & (1 << 2)) c.qt_exponent = qe, c.qt_mantissa = qm;
        v = r.sym(1);
        if (cs & (1 << 3)) set_progression(codec, (bwc_prog_ord)v);
        v = r.sym(1);
        if (cs & (1 << 4))
          set_kernels(codec, (bwc_dwt_filter)((v >> 6) & 3), (bwc_dwt_filter)((v >> 4) & 3),
                      (bwc_dwt_filter)((v >> 2) & 3), (bwc_dwt_filter)(v & 3));
        v = r.sym(4);
        if (cs & (1 << 5))
          bwc_set_decomp(codec, (uint8)((v >> 24) & 0xFF), (uint8)((v >> 16) & 0xFF), (uint8)((v >> 8) & 0xFF),
                         (uint8)(v & 0xFF));
        v = r.sym(2);
        if (cs & (1 << 6))
          bwc_set_precincts(codec, (uint8)((v >> 8) & 0x0F), (uint8)((v >> 12) & 0x0F), (uint8)(v & 0x0F),
                            (uint8)((v >> 4) & 0x0F));
        v = r.sym(4);
        if (cs & (1 << 7))
          bwc_set_codeblocks(codec, (uint8)((v >> 24) & 0xFF), (uint8)((v >> 16) & 0xFF), (uint8)((v >> 8) & 0xFF),
                             (uint8)(v & 0xFF));
        v = r.sym(1);
        if (cs & (1 << 8)) bwc_set_qm(codec, (uint8)v);
        uint64_t ts[4];
        for (auto &x : ts) x = r.sym(8);
        if (cs & (1 << 9)) bwc_set_tiles(codec, ts[0], ts[1], ts[2], ts[3], bwc_tile_sizeof);
        c.nLayers = (uint8)r.sym(1);
        std::free(c.bitrate);
        c.bitrate = static_cast<float *>(std::calloc(c.nLayers ? c.nLayers : 1, sizeof(float)));
        for (int l = 0; l < c.nLayers; ++l) {
          const uint32_t bits = (uint32_t)r.sym(4);
          std::memcpy(&c.bitrate[l], &bits, 4);
        }
        if (r.err || create_codec(codec)) return 1;
        sgc = true;
        break;
      }
      case 0xFF53: {   // SAX
        const uint64_t n = r.sym(4);
        if (r.err || n < 4 || !r.can(n - 4) || r.L + (n - 4) > r.cap) return 1;
        if (data) {
          data->codestream.aux = static_cast<bwc_span *>(std::calloc(1, sizeof(bwc_span)));
          data->codestream.aux->memory = static_cast<uchar *>(std::malloc(n - 4 ? n - 4 : 1));
          std::memcpy(data->codestream.aux->memory, r.m + r.L, n - 4);
          data->codestream.aux->size = n - 4;
        }
        r.L += n - 4;
        break;
      }
      case 0xFF57: {   // COM
        const uint64_t n = r.sym(2);
        if (r.err || n < 2 || !r.can(n - 2) || r.L + (n - 2) > r.cap) return 1;
        if (data) {
          data->codestream.com = static_cast<bwc_span *>(std::calloc(1, sizeof(bwc_span)));
          data->codestream.com->memory = static_cast<uchar *>(std::malloc(n - 2 ? n - 2 : 1));
          std::memcpy(data->codestream.com->memory, r.m + r.L, n - 2);
          data->codestream.com->size = n - 2;
        }
        r.L += n - 2;
        break;
      }
      case 0xFF58: {   // EOH: parameter ranges of every tile
        if (!sgi || !sgc) return 1;
        r.sym(2);
        Job *job = job_of(codec);
        for (auto &tl : job->plan.tiles)
          for (auto &par : tl.par) {
            double mm[2];
            for (double &x : mm) {
              if (codec_prec == 8) {
                const uint64_t raw = r.sym(8);
                std::memcpy(&x, &raw, 8);
              } else {
                const uint32_t raw = (uint32_t)r.sym(4);
                float f;
                std::memcpy(&f, &raw, 4);
                x = f;
              }
            }
            par.pmin = mm[0];
            par.pmax = mm[1];
          }
        return r.err ? 1 : 0;
      }
      default:
        return 1;
    }
    if (r.err) return 1;
  }
}

// one packet body segment of a codeblock: bytes [start, start + n) of its
// coded data, at src in the codestream
struct BodyChunk {
  uint64_t cb;
  uint32_t start, n;
  const uint8_t *src;
};
std::vector<BodyChunk> &empty_chunks() {
  static std::vector<BodyChunk> e;
  return e;
}

// parse_packet (tier2.c:1364-1634) for packet (param j, resolution res,
// precinct p, layer l) at `at`; `avail` is the tile's remaining declared
// body, `real` the bytes there. Returns header + body bytes, or -1.
int64_t parse_packet(bwc_codec *codec, Job *job, uint64_t ti, int j, Resolution &res, uint64_t p, int l,
                     std::vector<std::vector<int16_t>> &cp, std::vector<std::vector<uint16_t>> &beta,
                     std::vector<std::vector<uint32_t>> &Lcut, std::vector<std::vector<BodyChunk>> &chunks,
                     const uint8_t *at, uint64_t avail, uint64_t real) {
  const int nl = codec->control.nLayers;
  DecodeRecords &D = job->drec[ti * codec->info.nPar + j];
  Reader h;
  h.m = at, h.Lmax = avail, h.cap = real;
  const uint8_t e = h.bit();
  for (auto &sb : res.sb) {
    Precinct &pr = sb.prec[p];
    for (uint64_t k = 0; k < pr.ncbs; ++k) {
      const uint64_t i = pr.first_cb + k;
      int16_t *c = &cp[j][i * (nl + 1) + 1];
      if (!e) {
        c[l] = c[l - 1];
        continue;
      }
      int dz = 0;
      if (c[l - 1] == 0) {
        tagtree_decode(pr.incl, h, (uint32_t)l + 1, k);
        if (pr.incl.value[k] <= l) {
          dz = 1;
          uint32_t t = 1;
          while (tagtree_decode(pr.msbs, h, t++, k)) {
          }
          D.Kmsbs[i] = (uint8_t)pr.msbs.value[k];
        }
      } else {
        dz = h.bit();
      }
      c[l] = c[l - 1];
      if (!dz) continue;
      // decode_delta_z
      if (!h.bit()) {
        dz = 1;
      } else if (!h.bit()) {
        dz = 2;
      } else {
        dz = (h.bit() << 1) | h.bit();
        if (dz < 3) {
          dz += 3;
        } else {
          dz = 0;
          for (int b = 0; b < 5; ++b) dz = (dz << 1) | h.bit();
          if (dz < 31) {
            dz += 6;
          } else {
            dz = 0;
            for (int b = 0; b < 8; ++b) dz = (dz << 1) | h.bit();
            dz += 37;
          }
        }
      }
      c[l] = (int16_t)(c[l] + dz);
      D.Z[i] = (uint8_t)c[l];
      // decode_length
      while (h.bit()) beta[j][i]++;
      const int16_t zc = (int16_t)(c[l] - 1), zp = (int16_t)(c[l - 1] - 1);
      uint16_t nb = (uint16_t)(beta[j][i] + (uint16_t)std::log2((double)(zc - zp)));
      uint32_t v = 0;
      for (; nb > 0; --nb) v = (v << 1) | h.bit();
      if (zc < 0 || zc >= kMaxPasses) return -1;
      Lcut[j][i * kMaxPasses + zc] = ((zp >= 0) ? Lcut[j][i * kMaxPasses + zp] : 0) + v;
    }
  }
  uint64_t hsize = h.used();
  const uint8_t *body = at + hsize;
  if (codec->control.error_resilience) {
    if (hsize + 2 > real || body[0] != 0xFF || body[1] != 0x92) return -1;
    body += 2;
    hsize += 2;
  }
  uint64_t bsize = 0;
  if (e)
    for (auto &sb : res.sb) {
      Precinct &pr = sb.prec[p];
      for (uint64_t k = 0; k < pr.ncbs; ++k) {
        const uint64_t i = pr.first_cb + k;
        const int16_t *c = &cp[j][i * (nl + 1) + 1];
        if (c[l] <= c[l - 1]) continue;
        const uint32_t *L = &Lcut[j][i * kMaxPasses];
        const uint32_t start = (c[l - 1] == 0) ? 0 : L[c[l - 1] - 1];
        const uint32_t end = (c[l] == 0) ? 0 : L[c[l] - 1];
        if (end < start || hsize + bsize + (end - start) > real) return -1;
        if (end > start) chunks[j].push_back({i, start, end - start, body + bsize});
        bsize += end - start;
      }
    }
  return (int64_t)(hsize + bsize);
}

// Per-tile packet state: cut passes per layer, beta, lengths at the cuts,
// and per parameter the body chunks read so far (in codestream order).
struct TileState {
  bool seen = false;
  std::vector<std::vector<int16_t>> cp;
  std::vector<std::vector<uint16_t>> beta;
  std::vector<std::vector<uint32_t>> Lcut;
  std::vector<std::vector<BodyChunk>> chunks;
};

// one tile (parse_tile, codestream.c:1054-1200). Returns 0, 1 when the
// reference would stop parsing here, or -1 when a read would leave the
// buffer.
int parse_tile(bwc_codec *codec, Reader &r, std::vector<TileState> &ts) {
  Job *job = job_of(codec);
  const bwc_gl_ctrl &ctl = codec->control;
  const int np = codec->info.nPar, nl = ctl.nLayers;
  if (!r.can(2)) return 1;
  r.sym(2);
  const uint64_t ti = r.sym(4);
  if (r.err || ti >= job->plan.tiles.size()) return 1;
  Tile &tl = job->plan.tiles[ti];
  uint64_t body = r.sym(8);
  if (r.err || (uint16_t)r.sym(2) != 0xFF93) return 1;
  TileState &st = ts[ti];
  if (!st.seen) {
    st.seen = true;
    st.cp.assign(np, {}), st.beta.assign(np, {}), st.Lcut.assign(np, {}), st.chunks.assign(np, {});
    for (int j = 0; j < np; ++j) {
      const uint64_t n = tl.par[j].ncb;
      st.cp[j].assign(n * (nl + 1), 0);
      st.beta[j].assign(n, 3);
      st.Lcut[j].assign(n * kMaxPasses, 0);
    }
  }
  const bool rlcp = ctl.progression == bwc_prog_RLCP;
  const int nA = rlcp ? ctl.nDecomp + 1 : nl, nB = rlcp ? nl : ctl.nDecomp + 1;
  uint64_t npk = 0, idx = 0;
  for (int j = 0; j < np; ++j)
    for (auto &res : tl.par[j].res) npk += res.nprec * (uint64_t)nl;
  for (int a = 0; a < nA; ++a)
    for (int b = 0; b < nB; ++b)
      for (int j = 0; j < np; ++j) {
        const int rr = rlcp ? a : b, l = rlcp ? b : a;
        Resolution &res = tl.par[j].res[rr];
        for (uint64_t p = 0; p < res.nprec; ++p, ++idx) {
          if (!(body != 0 && r.can(body) && idx < npk)) return 0;
          if (r.L + 2 > r.cap) return -1;
          if ((uint16_t)r.sym(2) == 0xFF91) {
            if (!r.can(6)) return 1;
            if ((uint16_t)r.sym(2) != 4) return 1;
            r.sym(4);
            body -= 8;
          } else {
            r.L -= 2;
          }
          if (r.err) return -1;
          const int64_t sz = parse_packet(codec, job, ti, j, res, p, l, st.cp, st.beta, st.Lcut, st.chunks,
                                          r.m + r.L, body, r.cap - r.L);
          if (sz < 0) return -1;
          r.L += (uint64_t)sz;
          body -= (uint64_t)sz;
        }
      }
  return 0;
}

// tiles and packets (parse_body, codestream.c:1204-1260). As there, a bad
// or unexpected marker ends the parse and what was read is decoded.
int parse_body(bwc_codec *codec, Reader &r) {
  Job *job = job_of(codec);
  const int np = codec->info.nPar;
  job->drec.clear();
  job->drec.resize(job->plan.tiles.size() * np);
  for (size_t ti = 0; ti < job->plan.tiles.size(); ++ti)
    for (int j = 0; j < np; ++j) {
      DecodeRecords &D = job->drec[ti * np + j];
      D.n = job->plan.tiles[ti].par[j].ncb;
      D.Kmsbs.assign(D.n, 0), D.Z.assign(D.n, 0), D.len.assign(D.n, 0), D.off.assign(D.n + 1, 0);
    }
  std::vector<TileState> ts(job->plan.tiles.size());
  for (auto &tl : job->plan.tiles)
    for (auto &par : tl.par)
      for (auto &res : par.res)
        for (auto &sb : res.sb)
          for (auto &pr : sb.prec) {
            std::fill(pr.incl.value.begin(), pr.incl.value.end(), 0xFFFF);
            std::fill(pr.incl.threshold.begin(), pr.incl.threshold.end(), 0);
            std::fill(pr.msbs.value.begin(), pr.msbs.value.end(), 0xFFFF);
            std::fill(pr.msbs.threshold.begin(), pr.msbs.threshold.end(), 0);
          }
  int rc = 0;
  while (r.can(2)) {
    if (r.L + 2 > r.cap) {
      rc = -1;
      break;
    }
    const uint16_t marker = (uint16_t)r.sym(2);
    if (marker < 0xFF00 || marker == 0xFFFF) break;
    if (marker == 0xFF90) {
      const int t = parse_tile(codec, r, ts);
      if (t < 0) rc = -1;
      if (t) break;
    } else if (marker >= 0xFF50 && marker <= 0xFF58) {
      break;   // a main-header marker in the body
    } else {
      const uint64_t n = r.sym(2);
      if (r.err || n < 2 || !r.can(n - 2)) break;
      r.L += n - 2;
    }
  }
  if (rc) return 1;
  // every tile's records: bytes back to back in access order
  for (size_t ti = 0; ti < job->plan.tiles.size(); ++ti)
    for (int j = 0; j < np; ++j) {
      DecodeRecords &D = job->drec[ti * np + j];
      const uint64_t n = D.n;
      uint64_t tot = 0;
      for (uint64_t i = 0; i < n; ++i) {
        D.off[i] = tot;
        D.len[i] = (ts[ti].seen && D.Z[i]) ? ts[ti].Lcut[j][i * kMaxPasses + D.Z[i] - 1] : 0;
        tot += D.len[i];
      }
      D.off[n] = tot;
      // each codeblock's bytes [0, len): zero, then its chunks in the order
      // they were read (a later one overwrites an earlier one where they
      // meet, bytes past len are dropped), as the reference assembles them;
      // codeblocks in parallel on the host pool. 512 bytes of zeroed slack:
      // the decoders read 256-byte chunks up to one chunk past the end.
      D.data_size = tot + 512;
      D.data.reset(new uint8_t[D.data_size]);
      std::memset(D.data.get() + tot, 0, 512);
      std::vector<BodyChunk> &ch = ts[ti].seen ? ts[ti].chunks[j] : empty_chunks();
      std::vector<uint64_t> first(n + 1, 0);
      for (const BodyChunk &c : ch) first[c.cb + 1]++;
      for (uint64_t i = 0; i < n; ++i) first[i + 1] += first[i];
      std::vector<uint32_t> order(ch.size());
      {
        std::vector<uint64_t> at(first.begin(), first.end() - 1);
        for (uint32_t k = 0; k < (uint32_t)ch.size(); ++k) order[at[ch[k].cb]++] = k;
      }
      const size_t nparts = (size_t)std::min<uint64_t>(64, (n + 255) / 256);
      const std::function<void(size_t)> fill = [&](size_t part) {
        for (uint64_t i = n * part / nparts, e = n * (part + 1) / nparts; i < e; ++i) {
          uint8_t *dst = D.data.get() + D.off[i];
          std::memset(dst, 0, D.len[i]);
          for (uint64_t q = first[i]; q < first[i + 1]; ++q) {
            const BodyChunk &c = ch[order[q]];
            if (c.start < D.len[i]) std::memcpy(dst + c.start, c.src, std::min<uint64_t>(c.n, D.len[i] - c.start));
          }
        }
      };
      const int threads = rc_thread_count();
      if (nparts > 1 && threads > 1) {
        host_pool_run(threads - 1, nparts, fill);
      } else {
        for (size_t part = 0; part < nparts; ++part) fill(part);
      }
    }
  return 0;
}

}  // namespace

int parse_header_only(bwc_codec *codec, bwc_stream *data) {
  Reader r;
  r.m = static_cast<const uint8_t *>(data->inp);
  r.Lmax = 10;
  r.cap = ~0ull;   // the caller's buffer size is not known; the header bounds every read
  return parse_main_header(codec, data, r);
}

int parse_codestream(bwc_codec *codec, bwc_stream *data, uint8_t layer) {
  Reader r;
  r.m = static_cast<const uint8_t *>(data->inp);
  r.Lmax = 10;
  r.cap = 10;
  // the declared size bounds the stream (the caller's buffer holds it)
  {
    Reader hd;
    hd.m = r.m, hd.Lmax = 10, hd.cap = 10;
    if ((uint16_t)hd.sym(2) != 0xFF50) return 1;
    r.cap = hd.sym(8);
  }
  if (parse_main_header(codec, data, r)) return 1;
  // useLayer is kept as the reference keeps it; every layer is decoded
  codec->control.useLayer = codec->control.nLayers - 1;
  if (layer < codec->control.nLayers && layer > 0) codec->control.useLayer = layer;
  Job *job = job_of(codec);
  if (parse_body(codec, r)) return 1;
  job->decodable = true;
  return 0;
}

}  // namespace bwc
