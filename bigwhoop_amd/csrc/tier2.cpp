// Tier-2 on the host: PCRD rate control over the slope threshold, packet
// headers (tag trees, Δz codes, β/Lblock lengths with 0xFF bit stuffing) and
// the codestream writer.
//
// Reference: tier2.c:100-153 (Δz), :240-312 (lengths), :433-711 (packets with
// estimate modes 2 = probe, 1 = final estimate, 0 = real header), :750-849,
// :895-1254 (quality layers and the bisection); tagtree.c:310-373;
// bitstream.c:277-363, :561-612; codestream.c:139-570, :1329-1413.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <unordered_map>

#include "codec.hpp"

namespace bwc {

namespace {

// MSB-first bit writer with JPEG 2000 style stuffing: after an emitted 0xFF
// byte the next byte carries only 7 bits.
//
// put(v, n) appends the n low bits of v, most significant first; it handles
// a byte at a time, which gives the same bytes as n calls of bit().
template <class Sink>
struct StuffedBits {
  Sink sink;
  uint8_t T = 0;
  int t = 8;
  void emit() {
    sink(T);
    t = (T == 0xFF) ? 7 : 8;
    T = 0;
  }
  void bit(uint64_t b) {
    --t;
    if (b) T |= (uint8_t)(1u << t);
    if (!t) emit();
  }
  void put(uint64_t v, int n) {
    while (n > 0) {
      const int k = n < t ? n : t;
      const uint32_t chunk = (uint32_t)(v >> (n - k)) & ((1u << k) - 1u);
      T |= (uint8_t)(chunk << (t - k));
      t -= k;
      n -= k;
      if (!t) emit();
    }
  }
  void ones(int n) {
    while (n > 0) {
      const int k = n < 32 ? n : 32;
      put(0xFFFFFFFFull, k);
      n -= k;
    }
  }
  void zeros(int n) {
    while (n > 0) {
      const int k = n < 32 ? n : 32;
      put(0, k);
      n -= k;
    }
  }
  void flush() {
    if (t != 8) sink(T);
  }
};

struct ByteSink {
  std::vector<uint8_t> out;
  void operator()(uint8_t b) { out.push_back(b); }
};
struct CountSink {
  uint64_t n = 0;
  void operator()(uint8_t) { ++n; }
};

using BitWriter = StuffedBits<ByteSink>;
// Counting writer for the estimation modes: same bit-stuffing arithmetic,
// no buffer.
using BitCounter = StuffedBits<CountSink>;

// Raw (unstuffed) bit string, used to cache the tag-tree bits of one
// codeblock while the tree states stay fixed during a layer's probes.
struct BitString {
  std::vector<uint64_t> *words;
  size_t base;           // first word of this string in `words`
  uint64_t nbits = 0;
  explicit BitString(std::vector<uint64_t> *w) : words(w), base(w->size()) {}
  void bit(uint64_t b) { put(b ? 1 : 0, 1); }
  void put(uint64_t v, int n) {
    for (int i = n; i-- > 0;) {
      const size_t w = base + (size_t)(nbits >> 6);
      if (w >= words->size()) words->push_back(0);
      if ((v >> i) & 1) (*words)[w] |= 1ull << (63 - (nbits & 63));
      ++nbits;
    }
  }
  void ones(int n) { for (int i = 0; i < n; ++i) put(1, 1); }
  void zeros(int n) { for (int i = 0; i < n; ++i) put(0, 1); }
};

// Tag-tree coding of one leaf up to `threshold` (tagtree.c:310-373). At each
// node from the root down the reference emits one bit per threshold step: a
// 0 while the node value is still >= the raised threshold, then a 1 once it
// passes the value (if that happens below `threshold`). That is a run of
// min(v, threshold) - th zeros followed by a 1 when v < threshold.
template <class W>
void tagtree_encode(TagTree &tt, W &w, uint32_t threshold, uint64_t leaf, int est) {
  int64_t branch[64];
  int nb = 0;
  int64_t n = (int64_t)leaf;
  while (tt.parent[n] >= 0) {
    branch[nb++] = n;
    n = tt.parent[n];
  }
  uint32_t tmin = 0;
  for (;;) {
    uint32_t th = tt.threshold[n];
    if (th < tmin) th = tmin;
    const uint16_t v = tt.value[n];
    if (v >= th && th < threshold) {
      const uint32_t hi = std::min<uint32_t>(v, threshold);
      w.zeros((int)(hi - th));
      if (v < threshold) w.put(1, 1);
      th = std::min<uint32_t>((uint32_t)v + 1, threshold);
    }
    tmin = std::min<uint32_t>(v, th);
    if (!est) tt.threshold[n] = (uint16_t)th;
    if (n == (int64_t)leaf) break;
    n = branch[--nb];
  }
}

// Δz code of the number of new passes (tier2.c:100-153); nothing is written
// for 293 or more, as in the reference.
template <class W>
void delta_z_code(W &w, int dz) {
  if (dz == 1) {
    w.put(0, 1);
  } else if (dz == 2) {
    w.put(2, 2);
  } else if (dz < 6) {
    w.put(0xCu | (uint32_t)(dz - 3), 4);
  } else if (dz < 37) {
    w.put((0xFu << 5) | (uint32_t)(dz - 6), 9);
  } else if (dz < 293) {
    w.put((0x1FFu << 8) | (uint32_t)(dz - 37), 17);
  }
}

// floor(log2(x)) of a positive integer as the reference computes it through
// the double log2 (exact for 32-bit values); x == 0 keeps the library call
inline uint16_t ilog2_u32(uint32_t x) { return x ? (uint16_t)(31 - __builtin_clz(x)) : (uint16_t)std::log2(x); }

struct T2 {
  bwc_codec *c;
  Job *job;
  Tile *tl;
  uint64_t tile;
  int nl;
  CblkRecords &rec(int j) { return job->rec[tile * c->info.nPar + j]; }

  template <class W>
  void length_code(W &w, CblkRecords &R, uint64_t i, int q, int est) {
    int16_t *cp = &R.cp[i * (nl + 1) + 1];
    const int16_t zc = (int16_t)(cp[q] - 1), zp = (int16_t)(cp[q - 1] - 1);
    uint16_t beta = est ? R.beta_est[i] : R.beta[i];
    uint16_t maxb = beta;
    const uint32_t *L = &R.L[i * kMaxPasses];
    const uint32_t len = L[zc] - ((zp >= 0) ? L[zp] : 0);
    const uint16_t cpb = ilog2_u32((uint32_t)(zc - zp));
    int32_t lb = ilog2_u32(len) + 1 - cpb;
    if (maxb < lb) maxb = (uint16_t)lb;
    // β increments as a run of ones, a zero, then the length in β + ⌊log2 Δz⌋ bits
    w.ones(maxb - beta);
    beta = std::max(beta, maxb);
    w.put(0, 1);
    // (a shift of 31 or more turns the reference's int mask negative and it
    // writes no length bits; kept as is)
    const int nb = beta + cpb;
    if (nb > 0 && nb < 32) w.put(len & ((1u << nb) - 1u), nb);
    if (est == 0) {
      R.beta[i] = beta;
      R.beta_est[i] = beta;
    } else if (est == 1) {
      R.beta_est[i] = beta;
    }
  }

  // header bytes of packet (param j, resolution r, precinct p, layer q)
  template <class W>
  void packet_header(W &w, int j, Resolution &res, uint64_t p, int q, int est, uint64_t *body) {
    CblkRecords &R = rec(j);
    w.bit(1);
    for (auto &sb : res.sb) {
      Precinct &pr = sb.prec[p];
      for (uint64_t k = 0; k < pr.ncbs; ++k) {
        const uint64_t i = pr.first_cb + k;
        int16_t *cp = &R.cp[i * (nl + 1) + 1];
        const int dz = cp[q] - cp[q - 1];
        if (cp[q - 1] == 0) {
          if (est == 2 && lc) {
            lc->incl[j].feed(w, i);
            if (cp[q]) lc->msbs[j].feed(w, i);
          } else {
            tagtree_encode(pr.incl, w, (uint32_t)q + 1, k, est);
            if (cp[q])
              for (uint32_t t = 1; t <= R.Kmsbs[i] + 1u; ++t) tagtree_encode(pr.msbs, w, t, k, est);
          }
        } else {
          w.bit((uint64_t)dz);
        }
        if (dz) {
          delta_z_code(w, dz);
          length_code(w, R, i, q, est);
          if (est <= 1 && body) {
            const uint32_t *L = &R.L[i * kMaxPasses];
            const uint32_t start = (cp[q - 1] == 0) ? 0 : L[cp[q - 1] - 1];
            const uint32_t end = (cp[q] == 0) ? 0 : L[cp[q] - 1];
            *body += end - start;
          }
        }
      }
    }
    w.flush();
  }

  // create_packet for est 1/2: returns the header estimate in bytes
  int64_t estimate_packet(int j, Resolution &res, uint64_t p, int q, int est) {
    Packet &pk = res.pk[q * res.nprec + p];
    if (!pk.e) {
      if (est == 2) pk.e = 0;
      return 1;
    }
    BitCounter w;
    uint64_t body = 0;
    packet_header(w, j, res, p, q, est, est == 1 ? &body : nullptr);
    if (est == 1) pk.body_size += body;
    if (est == 2) pk.e = 0;
    return (int64_t)w.sink.n;
  }

  void real_packet(int j, Resolution &res, uint64_t p, int q) {
    Packet &pk = res.pk[q * res.nprec + p];
    pk.built = true;
    if (!pk.e) {
      pk.header.assign(1, 0);
      pk.size = 1;
      return;
    }
    BitWriter w;
    packet_header(w, j, res, p, q, 0, nullptr);
    pk.header = std::move(w.sink.out);
    pk.size = (uint32_t)(pk.body_size + pk.header.size());
  }

  // Per parameter: the passes with a nonzero slope (the convex-hull points)
  // in pass order. Hull slopes never increase along the hull, so the cut for
  // a threshold T >= 1, max{z : S[z] >= T}, is found by bisection over these
  // points instead of a scan over all passes. A codeblock whose nonzero
  // slopes are not ordered that way keeps the scan (flag `scan`).
  struct HullIndex {
    std::vector<uint32_t> off;
    std::vector<uint8_t> z;
    std::vector<uint16_t> s;
    std::vector<uint8_t> scan;
  };
  std::vector<HullIndex> hull;

  void build_hull_index() {
    hull.assign(c->info.nPar, HullIndex{});
    for (int j = 0; j < c->info.nPar; ++j) {
      CblkRecords &R = rec(j);
      HullIndex &H = hull[j];
      H.off.assign(R.n + 1, 0);
      H.scan.assign(R.n, 0);
      for (uint64_t i = 0; i < R.n; ++i) {
        H.off[i] = (uint32_t)H.z.size();
        const uint16_t *S = &R.S[i * (kMaxPasses + 1)];
        uint16_t prev = 0xFFFF;
        for (int z = 0; z <= R.Z[i]; ++z) {
          if (!S[z]) continue;
          if (S[z] > prev) H.scan[i] = 1;
          prev = S[z];
          H.z.push_back((uint8_t)z);
          H.s.push_back(S[z]);
        }
        if (S[0] == 0) H.scan[i] = 1;
      }
      H.off[R.n] = (uint32_t)H.z.size();
    }
  }

  int16_t cut_of(int j, const CblkRecords &R, uint64_t i, uint16_t threshold) const {
    if (threshold >= 1 && !hull.empty() && !hull[j].scan[i]) {
      const HullIndex &H = hull[j];
      uint32_t lo = H.off[i], hi = H.off[i + 1];   // S[z[lo]] = S[0] >= threshold
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (H.s[mid] >= threshold) lo = mid;
        else hi = mid;
      }
      return H.z[lo];
    }
    const uint16_t *S = &R.S[i * (kMaxPasses + 1)];
    int16_t z;
    for (z = R.Z[i]; z >= 0 && S[z] < threshold; --z) {
    }
    return z;
  }

  // Tag-tree bits of every not-yet-included codeblock for the probes of one
  // layer: inclusion to threshold q + 1 and the MSB sequence. Probes do not
  // change tree states, so these bits are the same for every probe.
  struct BitCache {
    std::vector<uint64_t> words;
    std::vector<uint32_t> woff, nbits;
    template <class W>
    void feed(W &w, uint64_t i) const {
      uint32_t n = nbits[i];
      const uint64_t *p = &words[woff[i]];
      while (n >= 64) {
        w.put(*p >> 32, 32);
        w.put(*p & 0xFFFFFFFFull, 32);
        ++p;
        n -= 64;
      }
      if (n) w.put(*p >> (64 - n), (int)n);
    }
  };
  struct LayerCache {
    std::vector<BitCache> incl, msbs;
  };
  LayerCache layer_cache;
  const LayerCache *lc = nullptr;

  void build_layer_cache(int q) {
    // BWC_AMD_RC_NOCACHE=1 encodes the trees in every probe instead
    if (const char *e = std::getenv("BWC_AMD_RC_NOCACHE"))
      if (e[0] == '1') {
        lc = nullptr;
        return;
      }
    layer_cache.incl.assign(c->info.nPar, BitCache{});
    layer_cache.msbs.assign(c->info.nPar, BitCache{});
    for (int j = 0; j < c->info.nPar; ++j) {
      Param &par = tl->par[j];
      CblkRecords &R = rec(j);
      BitCache &BI = layer_cache.incl[j], &BM = layer_cache.msbs[j];
      BI.woff.assign(R.n, 0), BI.nbits.assign(R.n, 0);
      BM.woff.assign(R.n, 0), BM.nbits.assign(R.n, 0);
      for (auto &res : par.res)
        for (auto &sb : res.sb)
          for (uint64_t p = 0; p < res.nprec; ++p) {
            Precinct &pr = sb.prec[p];
            for (uint64_t k = 0; k < pr.ncbs; ++k) {
              const uint64_t i = pr.first_cb + k;
              if (R.cp[i * (nl + 1) + q] != 0) continue;   // cp[q - 1]
              BI.woff[i] = (uint32_t)BI.words.size();
              BitString bi{&BI.words};
              tagtree_encode(pr.incl, bi, (uint32_t)q + 1, k, 2);
              BI.nbits[i] = (uint32_t)bi.nbits;
              if (bi.nbits == 0) BI.words.push_back(0);
              BM.woff[i] = (uint32_t)BM.words.size();
              BitString bm{&BM.words};
              for (uint32_t t = 1; t <= R.Kmsbs[i] + 1u; ++t) tagtree_encode(pr.msbs, bm, t, k, 2);
              BM.nbits[i] = (uint32_t)bm.nbits;
              if (bm.nbits == 0) BM.words.push_back(0);
            }
          }
    }
    lc = &layer_cache;
  }

  int64_t quality_layer(uint16_t threshold, int q, int est) {
    int64_t ql = 0;
    for (int j = 0; j < c->info.nPar; ++j) {
      Param &par = tl->par[j];
      CblkRecords &R = rec(j);
      for (auto &res : par.res) {
        for (auto &sb : res.sb)
          for (uint64_t p = 0; p < res.nprec; ++p) {
            Precinct &pr = sb.prec[p];
            for (uint64_t k = 0; k < pr.ncbs; ++k) {
              const uint64_t i = pr.first_cb + k;
              int16_t *cp = &R.cp[i * (nl + 1) + 1];
              const uint32_t *L = &R.L[i * kMaxPasses];
              const int16_t z = cut_of(j, R, i, threshold);
              if (z < cp[q - 1]) {
                cp[q] = cp[q - 1];
                if (cp[q] == 1) ql += L[0];
              } else {
                cp[q] = z;
                if (z > 0) {
                  ql += L[z - 1];
                  if (z > 1 && cp[q - 1] > 0) ql -= L[cp[q - 1] - 1];
                }
              }
              if (cp[q] > cp[q - 1]) res.pk[q * res.nprec + p].e = 1;
              if (!est) {
                if (q == 0) pr.msbs.set(k, R.Kmsbs[i]);
                if (q + 1 == nl) {
                  uint8_t l = 0;
                  while (l < nl && cp[l] == 0) ++l;
                  pr.incl.set(k, l);
                }
              }
            }
          }
        for (uint64_t p = 0; p < res.nprec; ++p) ql += estimate_packet(j, res, p, q, est + 1);
      }
    }
    return ql;
  }

  // ---- side-effect-free probes -------------------------------------------
  // probe(T, q) returns what quality_layer(T, q, 1) returns, reading the
  // shared state only (hull index, layer cache, previous layers' cuts,
  // beta_est) and writing into its own buffer, so several thresholds can be
  // estimated at once on host threads. quality_layer(.., 1) leaves no state
  // behind that a later probe reads (packet flags are cleared by every
  // estimate, cp[q] is rewritten for every codeblock), so the values are the
  // same as the sequential probes'.
  struct ProbeBuf {
    std::vector<std::vector<int16_t>> cq;   // per param: cut of layer q per codeblock
    std::vector<std::vector<uint8_t>> e;    // per param: contribution flag per (resolution, precinct)
    std::vector<std::vector<uint64_t>> eoff;
  };
  void probe_init(ProbeBuf &pb) {
    const int np = c->info.nPar;
    pb.cq.assign(np, {});
    pb.e.assign(np, {});
    pb.eoff.assign(np, {});
    for (int j = 0; j < np; ++j) {
      pb.cq[j].assign(rec(j).n, 0);
      uint64_t o = 0;
      for (auto &res : tl->par[j].res) {
        pb.eoff[j].push_back(o);
        o += res.nprec;
      }
      pb.e[j].assign(o, 0);
    }
  }

  template <class W>
  void length_code_probe(W &w, const CblkRecords &R, uint64_t i, int16_t prev, int16_t cur) const {
    const int16_t zc = (int16_t)(cur - 1), zp = (int16_t)(prev - 1);
    uint16_t beta = R.beta_est[i];
    uint16_t maxb = beta;
    const uint32_t *L = &R.L[i * kMaxPasses];
    const uint32_t len = L[zc] - ((zp >= 0) ? L[zp] : 0);
    const uint16_t cpb = ilog2_u32((uint32_t)(zc - zp));
    int32_t lb = ilog2_u32(len) + 1 - cpb;
    if (maxb < lb) maxb = (uint16_t)lb;
    w.ones(maxb - beta);
    beta = std::max(beta, maxb);
    w.put(0, 1);
    const int nb = beta + cpb;
    if (nb > 0 && nb < 32) w.put(len & ((1u << nb) - 1u), nb);
  }

  int64_t probe(uint16_t T, int q, ProbeBuf &pb) {
    int64_t ql = 0;
    for (int j = 0; j < c->info.nPar; ++j) {
      Param &par = tl->par[j];
      CblkRecords &R = rec(j);
      int16_t *cq = pb.cq[j].data();
      uint8_t *ef = pb.e[j].data();
      std::fill(pb.e[j].begin(), pb.e[j].end(), 0);
      for (size_t r = 0; r < par.res.size(); ++r) {
        Resolution &res = par.res[r];
        uint8_t *er = ef + pb.eoff[j][r];
        for (auto &sb : res.sb)
          for (uint64_t p = 0; p < res.nprec; ++p) {
            const Precinct &pr = sb.prec[p];
            for (uint64_t k = 0; k < pr.ncbs; ++k) {
              const uint64_t i = pr.first_cb + k;
              const int16_t prev = R.cp[i * (nl + 1) + q];
              const uint32_t *L = &R.L[i * kMaxPasses];
              const int16_t z = cut_of(j, R, i, T);
              int16_t cur;
              if (z < prev) {
                cur = prev;
                if (cur == 1) ql += L[0];
              } else {
                cur = z;
                if (z > 0) {
                  ql += L[z - 1];
                  if (z > 1 && prev > 0) ql -= L[prev - 1];
                }
              }
              cq[i] = cur;
              if (cur > prev) er[p] = 1;
            }
          }
        for (uint64_t p = 0; p < res.nprec; ++p) {
          if (!er[p]) {
            ql += 1;
            continue;
          }
          BitCounter w;
          w.bit(1);
          for (auto &sb : res.sb) {
            const Precinct &pr = sb.prec[p];
            for (uint64_t k = 0; k < pr.ncbs; ++k) {
              const uint64_t i = pr.first_cb + k;
              const int16_t prev = R.cp[i * (nl + 1) + q], cur = cq[i];
              const int dz = cur - prev;
              if (prev == 0) {
                if (lc) {
                  lc->incl[j].feed(w, i);
                  if (cur) lc->msbs[j].feed(w, i);
                } else {   // probes leave tree thresholds alone (est 2)
                  tagtree_encode(sb.prec[p].incl, w, (uint32_t)q + 1, k, 2);
                  if (cur)
                    for (uint32_t t = 1; t <= R.Kmsbs[i] + 1u; ++t) tagtree_encode(sb.prec[p].msbs, w, t, k, 2);
                }
              } else {
                w.bit((uint64_t)dz);
              }
              if (dz) {
                delta_z_code(w, dz);
                length_code_probe(w, R, i, prev, cur);
              }
            }
          }
          w.flush();
          ql += (int64_t)w.sink.n;
        }
      }
    }
    return ql;
  }

  // evaluate several thresholds of layer q at once
  int threads = 1;
  std::vector<ProbeBuf> pbufs;
  void probe_many(const std::vector<uint16_t> &Ts, int q, std::unordered_map<uint32_t, int64_t> &memo) {
    if ((int)pbufs.size() < std::max<int>(1, (int)Ts.size())) {
      pbufs.resize(std::max<size_t>(1, Ts.size()));
      for (auto &pb : pbufs)
        if (pb.cq.empty()) probe_init(pb);
    }
    std::vector<int64_t> out(Ts.size());
    if (Ts.size() == 1 || threads <= 1) {
      for (size_t a = 0; a < Ts.size(); ++a) out[a] = probe(Ts[a], q, pbufs[0]);
    } else {
      std::vector<std::thread> th;
      for (size_t a = 0; a < Ts.size(); ++a) th.emplace_back([&, a] { out[a] = probe(Ts[a], q, pbufs[a]); });
      for (auto &t : th) t.join();
    }
    for (size_t a = 0; a < Ts.size(); ++a) memo[Ts[a]] = out[a];
  }

  // thresholds the reference bisection may probe in its next `depth` steps
  // from (smin, smax): every step goes on with smin = slope or smax = slope
  // (an equal estimate ends it)
  static void frontier(uint16_t smin, uint16_t smax, int depth, std::vector<uint16_t> &out) {
    if (depth == 0 || smin >= smax) return;
    const uint16_t s = (uint16_t)((smax + smin) >> 1);
    if (std::find(out.begin(), out.end(), s) == out.end()) out.push_back(s);
    {  // estimate above the target
      uint16_t a = smin, b = smax;
      if (a == s) b = s;
      a = s;
      frontier(a, b, depth - 1, out);
    }
    {  // below
      uint16_t a = smin, b = smax;
      if (b == s) a = s;
      b = s;
      frontier(a, b, depth - 1, out);
    }
  }

  // Records coded only to a prefix of their passes (Z < Zfull): the layer's
  // threshold T and cuts equal those of complete codings when
  //  - the size estimate is strictly above the target at T - 1 and strictly
  //    below it at T: estimates do not grow with the threshold, so any
  //    bisection over an interval holding T - 1 returns T (a complete coding
  //    has a slope <= T - 1, as its estimate changes between T - 1 and T);
  //  - every prefix-coded codeblock cuts at T - 1 at least `margin` passes
  //    before its first pass whose length may still change (U), so the
  //    passes that decide its cuts at T - 1 and T are final.
  // Otherwise the codeblocks are flagged (or the tile, `fallback`) for
  // further coding.
  SkipCheck *chk = nullptr;
  void check_layer(uint16_t slope, int l, int64_t target, int64_t base) {
    if (target > 0) {
      std::unordered_map<uint32_t, int64_t> m;
      std::vector<uint16_t> Ts{slope};
      if (slope > 0) Ts.push_back((uint16_t)(slope - 1));
      if (threads > 0) {
        probe_many(Ts, l, m);
      } else {
        for (uint16_t T : Ts) m[T] = quality_layer(T, l, 1);
      }
      const int64_t f0 = m[slope] + base;
      const int64_t fm = slope > 0 ? m[(uint16_t)(slope - 1)] + base : INT64_MAX;
      if (!(f0 < target && fm > target)) chk->fallback = true;
    }
    const uint16_t t = slope > 0 ? (uint16_t)(slope - 1) : 0;
    for (int j = 0; j < c->info.nPar; ++j) {
      CblkRecords &R = rec(j);
      for (uint64_t i = 0; i < R.n; ++i) {
        if (R.Z[i] >= R.Zfull[i]) continue;
        const int z = cut_of(j, R, i, t);
        if (z + chk->margin > (int)R.U[i]) chk->flagged[j].push_back(i);
      }
    }
  }

  void rate_control() {
    int64_t est_tile = 0, est_ql = 0;
    const uint64_t tsz = tl->box.volume() * (uint64_t)c->info.nPar;
    const uint64_t hdr = (uint64_t)std::ceil((float)((uint32_t)c->control.headerSize + tl->header_size) / nl);
    build_hull_index();
    for (int l = 0; l < nl; ++l) {
      const int64_t target = (int64_t)std::ceil(c->control.bitrate[l] * tsz / 8.0f);
      uint16_t slope;
      if (target > 0.0f) {
        build_layer_cache(l);
        uint16_t smin = tl->slope_min, smax = tl->slope_max, tmp = tl->slope_min;
        slope = smin;
        // the reference's bisection, its probes estimated ahead on `threads`
        // host threads: the next three levels of its decision tree at once
        std::unordered_map<uint32_t, int64_t> memo;
        while (smin < smax) {
          slope = (uint16_t)((smax + smin) >> 1);
          int64_t f;
          if (threads > 0) {
            if (!memo.count(slope)) {
              std::vector<uint16_t> Ts;
              frontier(smin, smax, threads >= 7 ? 3 : (threads >= 3 ? 2 : 1), Ts);
              probe_many(Ts, l, memo);
            }
            f = memo[slope];
          } else {
            f = quality_layer(slope, l, 1);
          }
          est_ql = f + est_tile + (int64_t)hdr;
          if (est_ql > target) {
            if (smin == slope) smax = slope;
            smin = slope;
          } else if (est_ql < target) {
            if (smax == slope) smin = slope;
            smax = slope;
            tmp = slope;
          } else {
            smax = smin = slope;
          }
        }
        if (est_ql > target) slope = std::max(tmp, tl->slope_min);
      } else {
        slope = 0;
      }
      if (chk) check_layer(slope, l, target, est_tile + (int64_t)hdr);
      lc = nullptr;
      est_ql = quality_layer(slope, l, 0);
      est_tile += est_ql + (int64_t)hdr;
    }
  }

  void packets() {
    for (int j = 0; j < c->info.nPar; ++j)
      for (auto &res : tl->par[j].res)
        for (uint64_t p = 0; p < res.nprec; ++p)
          for (int l = 0; l < nl; ++l) real_packet(j, res, p, l);
  }
};

}  // namespace

// Rate control for one tile: fills cp (cut points per layer) and the packet
// contribution flags / body sizes. Needs tile slope bounds and records.
int t2_rate_control(bwc_codec *codec, Job *job, uint64_t tile, SkipCheck *chk) {
  T2 t{codec, job, &job->plan.tiles[tile], tile, codec->control.nLayers};
  // host threads for the bisection probes: BWC_AMD_RC_THREADS (0 = the
  // sequential in-place probes, 1 = side-effect-free probes on one thread)
  int rc_threads;
  if (const char *e = std::getenv("BWC_AMD_RC_THREADS")) {
    rc_threads = std::max(0, std::min(16, std::atoi(e)));
  } else {
    const unsigned hc = std::thread::hardware_concurrency();
    rc_threads = (int)std::max(1u, std::min(7u, hc > 1 ? hc - 1 : 1u));
  }
  t.threads = rc_threads;
  if (chk) {
    chk->fallback = false;
    chk->flagged.assign(codec->info.nPar, {});
    t.chk = chk;
  }
  t.rate_control();
  if (chk)
    for (auto &f : chk->flagged) {
      std::sort(f.begin(), f.end());
      f.erase(std::unique(f.begin(), f.end()), f.end());
    }
  return 0;
}

// Packet headers; needs the codeblock data only at assembly time.
int t2_packets(bwc_codec *codec, Job *job, uint64_t tile) {
  T2 t{codec, job, &job->plan.tiles[tile], tile, codec->control.nLayers};
  t.packets();
  return 0;
}

int t2_encode_tile(bwc_codec *codec, Job *job, uint64_t tile) {
  t2_rate_control(codec, job, tile);
  return t2_packets(codec, job, tile);
}

uint64_t header_size(const bwc_codec *c, const bwc_stream *st) {
  uint64_t h = 111 + (uint64_t)c->info.nPar * c->control.nTiles * 2 * kPrecBytes + (uint64_t)c->control.nLayers * 4;
  if (st && st->codestream.aux) h += 6 + st->codestream.aux->size;
  if (st && st->codestream.com) h += 6 + st->codestream.com->size;
  return h;
}

namespace {
struct ByteWriter {
  uint8_t *m;
  uint64_t L = 0, Lmax;
  bool err = false;
  void sym(uint64_t v, int n) {
    if (err) return;
    if (L + (uint64_t)n > Lmax) {
      err = true;
      return;
    }
    for (int i = n; i-- > 0;) m[L++] = (uint8_t)((v >> (8 * i)) & 0xFF);
  }
  void chunk(const uint8_t *p, uint64_t n) {
    if (err) return;
    if (L + n > Lmax) {
      err = true;
      return;
    }
    if (n) std::memcpy(m + L, p, n);
    L += n;
  }
};
}  // namespace

// Packet sequencing and codestream assembly (codestream.c:139-570,
// :1329-1413). Body bytes are copied straight from the packed codeblock data.
size_t assemble_codestream(bwc_codec *codec, Job *job, const bwc_stream *st, uint8_t *out) {
  bwc_gl_ctrl &ctl = codec->control;
  bwc_gl_inf &inf = codec->info;
  const int nl = ctl.nLayers;
  const bool rlcp = ctl.progression == bwc_prog_RLCP;
  ctl.codestreamSize = ctl.headerSize + 2;
  for (auto &tl : job->plan.tiles) {
    tl.seq.clear();
    uint64_t size = 0;
    const int nA = rlcp ? ctl.nDecomp + 1 : nl, nB = rlcp ? nl : ctl.nDecomp + 1;
    for (int a = 0; a < nA; ++a)
      for (int b = 0; b < nB; ++b)
        for (int j = 0; j < inf.nPar; ++j) {
          const int r = rlcp ? a : b, l = rlcp ? b : a;
          Resolution &res = tl.par[j].res[r];
          for (uint64_t p = 0; p < res.nprec; ++p) {
            Packet &pk = res.pk[l * res.nprec + p];
            tl.seq.push_back(&pk);
            size += pk.size + (ctl.error_resilience ? 6 : 0);
          }
        }
    tl.body_size = size;
    ctl.codestreamSize += tl.header_size + tl.body_size;
  }
  ByteWriter w{out, 0, ctl.codestreamSize};
  w.sym(0xFF50, 2);
  w.sym(ctl.codestreamSize, 8);
  w.sym(0xFF51, 2);
  w.sym(31, 2);
  w.sym(inf.nX, 8), w.sym(inf.nY, 8), w.sym(inf.nZ, 8), w.sym(inf.nTS, 8);
  w.sym(inf.nPar, 1);
  w.sym((uint8_t)inf.data_prec, 1);
  w.sym((uint8_t)inf.codec_prec, 1);
  w.sym(0xFF52, 2);
  w.sym(50 + nl * 4, 2);
  w.sym(ctl.CSsgc, 2);
  w.sym(ctl.error_resilience, 1);
  w.sym(ctl.quantization_style, 1);
  w.sym(ctl.guard_bits, 1);
  w.sym(ctl.qt_exponent, 1);
  w.sym(ctl.qt_mantissa, 2);
  w.sym(ctl.progression, 1);
  w.sym((uint64_t)(ctl.KernelX << 6 | ctl.KernelY << 4 | ctl.KernelZ << 2 | ctl.KernelTS), 1);
  w.sym(ctl.decompX, 1), w.sym(ctl.decompY, 1), w.sym(ctl.decompZ, 1), w.sym(ctl.decompTS, 1);
  w.sym((uint64_t)(ctl.precSizeY << 4 | ctl.precSizeX), 1);
  w.sym((uint64_t)(ctl.precSizeTS << 4 | ctl.precSizeZ), 1);
  w.sym(ctl.cbX, 1), w.sym(ctl.cbY, 1), w.sym(ctl.cbZ, 1), w.sym(ctl.cbTS, 1);
  w.sym(ctl.Qm, 1);
  w.sym(ctl.tileSizeX, 8), w.sym(ctl.tileSizeY, 8), w.sym(ctl.tileSizeZ, 8), w.sym(ctl.tileSizeTS, 8);
  w.sym(nl, 1);
  for (int l = 0; l < nl; ++l) {
    uint32_t bits;
    std::memcpy(&bits, &ctl.bitrate[l], 4);
    w.sym(bits, 4);
  }
  if (st && st->codestream.aux) {
    w.sym(0xFF53, 2);
    w.sym(st->codestream.aux->size + 4, 4);
    w.chunk(st->codestream.aux->memory, st->codestream.aux->size);
  }
  if (st && st->codestream.com) {
    w.sym(0xFF57, 2);
    w.sym(st->codestream.com->size + 2, 2);
    w.chunk(st->codestream.com->memory, st->codestream.com->size);
  }
  w.sym(0xFF58, 2);
  w.sym((uint64_t)inf.nPar * ctl.nTiles * 2 * kPrecBytes, 2);
  for (auto &tl : job->plan.tiles)
    for (auto &par : tl.par) {
      uint64_t r;
      std::memcpy(&r, &par.pmin, 8);
      w.sym(r, kPrecBytes);
      std::memcpy(&r, &par.pmax, 8);
      w.sym(r, kPrecBytes);
      par.pmax = -1.7976931348623157e+308;
      par.pmin = 1.7976931348623157e+308;
    }
  for (uint64_t ti = 0; ti < job->plan.tiles.size(); ++ti) {
    Tile &tl = job->plan.tiles[ti];
    w.sym(0xFF90, 2);
    w.sym(14, 2);
    w.sym(tl.index, 4);
    w.sym(tl.body_size, 8);
    w.sym(0xFF93, 2);
    // body slices follow the packet order: for packet (j, r, p, l) every
    // codeblock of every subband of precinct p contributes [L[cp[l-1]-1], L[cp[l]-1])
    uint64_t k = 0;
    const int nA = rlcp ? ctl.nDecomp + 1 : nl, nB = rlcp ? nl : ctl.nDecomp + 1;
    for (int a = 0; a < nA; ++a)
      for (int b = 0; b < nB; ++b)
        for (int j = 0; j < inf.nPar; ++j) {
          const int r = rlcp ? a : b, l = rlcp ? b : a;
          Resolution &res = tl.par[j].res[r];
          CblkRecords &R = job->rec[ti * inf.nPar + j];
          for (uint64_t p = 0; p < res.nprec; ++p, ++k) {
            Packet &pk = res.pk[l * res.nprec + p];
            if (ctl.error_resilience) {
              w.sym(0xFF91, 2);
              w.sym(4, 2);
              w.sym(k % 0x100000000ull, 4);
            }
            w.chunk(pk.header.data(), pk.header.size());
            if (ctl.error_resilience) w.sym(0xFF92, 2);
            if (pk.e) {
              for (auto &sb : res.sb) {
                Precinct &pr = sb.prec[p];
                for (uint64_t c = 0; c < pr.ncbs; ++c) {
                  const uint64_t i = pr.first_cb + c;
                  const int16_t *cp = &R.cp[i * (nl + 1) + 1];
                  if (cp[l] - cp[l - 1] <= 0) continue;
                  const uint32_t *L = &R.L[i * kMaxPasses];
                  const uint32_t start = (cp[l - 1] == 0) ? 0 : L[cp[l - 1] - 1];
                  const uint32_t end = (cp[l] == 0) ? 0 : L[cp[l] - 1];
                  w.chunk(R.data.data() + R.data_off[i] + start, end - start);
                }
              }
            }
          }
        }
  }
  w.sym(0xFFFF, 2);
  return w.err ? 0 : (size_t)w.L;
}

}  // namespace bwc
