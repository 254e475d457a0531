// Tier-2 on the host: PCRD rate control over the slope threshold, packet
// headers (tag trees, Δz codes, β/Lblock lengths with 0xFF bit stuffing) and
// the codestream writer.
//
// Reference: tier2.c:100-153 (Δz), :240-312 (lengths), :433-711 (packets with
// estimate modes 2 = probe, 1 = final estimate, 0 = real header), :750-849,
// :895-1254 (quality layers and the bisection); tagtree.c:310-373;
// bitstream.c:277-363, :561-612; codestream.c:139-570, :1329-1413.
#include <algorithm>
#include <chrono>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "codec.hpp"

namespace bwc {

namespace {

// Worker threads for the rate-control probes, started once and kept for the
// process. run(n, fn) calls fn(0) .. fn(n - 1) on the workers and the calling
// thread and returns when all are done; every worker takes part in every run
// (possibly with no task), so none can carry a finished run's function into
// the next.
class ProbePool {
 public:
  explicit ProbePool(int n) : nworkers_(n) {
    for (int i = 0; i < n; ++i) th_.emplace_back([this] { worker(); });
  }
  ~ProbePool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  int size() const { return nworkers_; }
  void run(size_t n, const std::function<void(size_t)> &fn) {
    {
      std::lock_guard<std::mutex> lk(m_);
      fn_ = &fn;
      n_ = n;
      next_ = 0;
      finished_ = 0;
      acked_ = 0;
      ++gen_;
    }
    cv_.notify_all();
    size_t done = 0;
    for (size_t a; (a = next_.fetch_add(1)) < n;) {
      fn(a);
      ++done;
    }
    std::unique_lock<std::mutex> lk(m_);
    finished_ += done;
    done_cv_.wait(lk, [&] { return finished_ == n_ && acked_ == nworkers_; });
  }

 private:
  void worker() {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(size_t)> *f;
      size_t n;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        f = fn_;
        n = n_;
      }
      size_t done = 0;
      for (size_t a; (a = next_.fetch_add(1)) < n;) {
        (*f)(a);
        ++done;
      }
      {
        std::lock_guard<std::mutex> lk(m_);
        finished_ += done;
        ++acked_;
      }
      done_cv_.notify_all();
    }
  }
  int nworkers_;
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(size_t)> *fn_ = nullptr;
  size_t n_ = 0;
  std::atomic<size_t> next_{0};
  size_t finished_ = 0;
  int acked_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// Runs fn(0 .. n-1) on the process's pool of `workers` threads (rebuilt
// when another size is asked). Runs from different host threads (concurrent
// compressions) take the pool one after the other.
void pool_run(int workers, size_t n, const std::function<void(size_t)> &fn) {
  static std::unique_ptr<ProbePool> pool;
  static std::mutex m;
  std::lock_guard<std::mutex> lk(m);
  if (!pool || pool->size() != workers) pool.reset(new ProbePool(workers));
  pool->run(n, fn);
}

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
using BitWriter = StuffedBits<ByteSink>;

// Counting writer for the estimation modes: the byte count StuffedBits
// would emit, from a 64-bit accumulator drained a byte at a time. A byte is
// 0xFF only when it holds 8 bits (a 7-bit byte's top bit is the stuffed 0),
// and flush() counts a last byte when bits are pending or when the byte
// before was 0xFF (StuffedBits then holds a 7-bit byte with nothing in it).
struct BitCounter {
  struct {
    uint64_t n = 0;
  } sink;
  uint64_t acc = 0;   // pending bits in the low `na` positions
  int na = 0;
  int cap = 8;        // bits of the byte being filled: 8, or 7 after 0xFF
  void drain() {
    while (na >= cap) {
      const uint32_t b = (uint32_t)(acc >> (na - cap)) & ((1u << cap) - 1u);
      na -= cap;
      ++sink.n;
      cap = (cap == 8 && b == 0xFF) ? 7 : 8;
    }
  }
  void put(uint64_t v, int n) {   // n <= 32
    acc = (acc << n) | (v & ((1ull << n) - 1ull));
    na += n;
    drain();
  }
  void bit(uint64_t b) { put(b ? 1 : 0, 1); }
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
    if (na > 0 || cap == 7) ++sink.n;
  }
};

// Raw (unstuffed) bit string, used to cache the tag-tree bits of one
// codeblock while the tree states stay fixed during a layer's probes.
struct BitString {
  std::vector<uint64_t> *words;
  size_t base;           // first word of this string in `words`
  uint64_t nbits = 0;
  explicit BitString(std::vector<uint64_t> *w) : words(w), base(w->size()) {}
  // words up to and including the one holding bit `nbits - 1`
  void cover() {
    const size_t need = base + (size_t)((nbits + 63) >> 6);
    if (words->size() < need) words->resize(need, 0);
  }
  void bit(uint64_t b) { put(b ? 1 : 0, 1); }
  void put(uint64_t v, int n) {   // n <= 32
    if (n <= 0) return;
    v &= (1ull << n) - 1ull;
    const size_t w = base + (size_t)(nbits >> 6);
    const int room = 64 - (int)(nbits & 63);
    nbits += (uint64_t)n;
    cover();
    if (n <= room) {
      (*words)[w] |= v << (room - n);
    } else {
      (*words)[w] |= v >> (n - room);
      (*words)[w + 1] |= v << (64 - (n - room));
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
    if (n <= 0) return;
    nbits += (uint64_t)n;
    cover();
  }
};

// Bits into preallocated, zeroed words (no growth, so several threads can
// write their own ranges of one array), and a bit count alone.
struct BitSpan {
  uint64_t *p;
  uint64_t nbits = 0;
  void bit(uint64_t b) { put(b ? 1 : 0, 1); }
  void put(uint64_t v, int n) {   // n <= 32
    if (n <= 0) return;
    v &= (1ull << n) - 1ull;
    const size_t w = (size_t)(nbits >> 6);
    const int room = 64 - (int)(nbits & 63);
    nbits += (uint64_t)n;
    if (n <= room) {
      p[w] |= v << (room - n);
    } else {
      p[w] |= v >> (n - room);
      p[w + 1] |= v << (64 - (n - room));
    }
  }
  void zeros(int n) {
    if (n > 0) nbits += (uint64_t)n;
  }
};
struct BitTally {
  uint64_t nbits = 0;
  void bit(uint64_t) { ++nbits; }
  void put(uint64_t, int n) {
    if (n > 0) nbits += (uint64_t)n;
  }
  void zeros(int n) {
    if (n > 0) nbits += (uint64_t)n;
  }
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

// tagtree_encode(tt, w, t, leaf, est) for t = 1 .. T in turn (the MSB code
// of a codeblock), with the leaf's path to the root found once
template <class W>
void tagtree_encode_upto(TagTree &tt, W &w, uint32_t T, uint64_t leaf, int est) {
  int64_t path[64];   // leaf first, root last
  int np = 0;
  for (int64_t n = (int64_t)leaf;; n = tt.parent[n]) {
    path[np++] = n;
    if (tt.parent[n] < 0) break;
  }
  for (uint32_t t = 1; t <= T; ++t) {
    uint32_t tmin = 0;
    for (int a = np - 1; a >= 0; --a) {
      const int64_t n = path[a];
      uint32_t th = tt.threshold[n];
      if (th < tmin) th = tmin;
      const uint16_t v = tt.value[n];
      if (v >= th && th < t) {
        const uint32_t hi = std::min<uint32_t>(v, t);
        w.zeros((int)(hi - th));
        if (v < t) w.put(1, 1);
        th = std::min<uint32_t>((uint32_t)v + 1, t);
      }
      tmin = std::min<uint32_t>(v, th);
      if (!est) tt.threshold[n] = (uint16_t)th;
    }
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
    w.bit(1);
    for (auto &sb : res.sb) subband_bits(w, j, sb, p, q, est, body);
    w.flush();
  }

  // the bits one subband's codeblocks add to a packet header (after its
  // leading 1): inclusion, zero bitplanes, passes and lengths
  template <class W>
  void subband_bits(W &w, int j, Subband &sb, uint64_t p, int q, int est, uint64_t *body) {
    CblkRecords &R = rec(j);
    {
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
              tagtree_encode_upto(pr.msbs, w, R.Kmsbs[i] + 1u, k, est);
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

  // what estimate_packet(.., est 1) adds to the packet's body size, without
  // the header estimate; returns the body bytes
  int64_t body_size_only(int j, Resolution &res, uint64_t p, int q) {
    Packet &pk = res.pk[q * res.nprec + p];
    if (!pk.e) return 1;
    CblkRecords &R = rec(j);
    uint64_t body = 0;
    for (auto &sb : res.sb) {
      const Precinct &pr = sb.prec[p];
      for (uint64_t k = 0; k < pr.ncbs; ++k) {
        const uint64_t i = pr.first_cb + k;
        const int16_t *cp = &R.cp[i * (nl + 1) + 1];
        if (cp[q] - cp[q - 1] == 0) continue;
        const uint32_t *L = &R.L[i * kMaxPasses];
        body += (cp[q] == 0 ? 0 : L[cp[q] - 1]) - (cp[q - 1] == 0 ? 0 : L[cp[q - 1] - 1]);
      }
    }
    pk.body_size += body;
    return (int64_t)body;
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
  // Per parameter, the passes with a nonzero slope of every codeblock
  // (entries off[i] .. off[i+1]): pass z, slope s and body length len =
  // L[z - 1] (0 for z = 0), kept together so a probe reads only these
  // compact arrays; l0 = L[0] per codeblock; scan = slopes not decreasing
  // (cuts found by a scan of S instead).
  struct HullIndex {
    std::vector<uint32_t> off;
    std::vector<uint8_t> z;
    std::vector<uint16_t> s;
    std::vector<uint32_t> len;
    std::vector<uint32_t> l0;
    std::vector<uint8_t> scan;
  };
  std::vector<HullIndex> hull;

  // fn(a, b) over [0, n) in chunks, on the probe threads when there are any
  void par_for(uint64_t n, const std::function<void(uint64_t, uint64_t)> &fn) {
    const uint64_t nch = threads > 1 ? std::min<uint64_t>(n, 64) : 1;
    if (nch <= 1) {
      fn(0, n);
      return;
    }
    const std::function<void(size_t)> task = [&](size_t k) { fn(n * k / nch, n * (k + 1) / nch); };
    pool_run(threads - 1, (size_t)nch, task);
  }

  void build_hull_index() {
    hull.assign(c->info.nPar, HullIndex{});
    for (int j = 0; j < c->info.nPar; ++j) {
      CblkRecords &R = rec(j);
      HullIndex &H = hull[j];
      const uint64_t n = R.n;
      H.off.assign(n + 1, 0);
      H.scan.assign(n, 0);
      H.l0.assign(n, 0);
      // entries per codeblock, then their offsets, then the entries
      par_for(n, [&](uint64_t a, uint64_t b) {
        for (uint64_t i = a; i < b; ++i) {
          const uint16_t *S = &R.S[i * (kMaxPasses + 1)];
          uint32_t k = 0;
          for (int z = 0; z <= R.Z[i]; ++z) k += S[z] != 0;
          H.off[i + 1] = k;
        }
      });
      for (uint64_t i = 0; i < n; ++i) H.off[i + 1] += H.off[i];
      const uint64_t tot = H.off[n];
      H.z.assign(tot, 0), H.s.assign(tot, 0), H.len.assign(tot, 0);
      par_for(n, [&](uint64_t a, uint64_t b) {
        for (uint64_t i = a; i < b; ++i) {
          const uint16_t *S = &R.S[i * (kMaxPasses + 1)];
          const uint32_t *L = &R.L[i * kMaxPasses];
          H.l0[i] = L[0];
          uint32_t k = H.off[i];
          uint16_t prev = 0xFFFF;
          for (int z = 0; z <= R.Z[i]; ++z) {
            if (!S[z]) continue;
            if (S[z] > prev) H.scan[i] = 1;
            prev = S[z];
            H.z[k] = (uint8_t)z;
            H.s[k] = S[z];
            H.len[k] = z > 0 ? L[z - 1] : 0;
            ++k;
          }
          if (S[0] == 0) H.scan[i] = 1;
        }
      });
    }
  }

  // cut of codeblock i at threshold T (as cut_of) and its body length
  // L[z - 1] (0 for z = 0)
  int16_t cut_len(int j, const CblkRecords &R, uint64_t i, uint16_t T, uint32_t &len) const {
    const HullIndex &H = hull[j];
    if (T >= 1 && !H.scan[i]) {
      uint32_t lo = H.off[i], hi = H.off[i + 1];
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (H.s[mid] >= T) lo = mid;
        else hi = mid;
      }
      len = H.len[lo];
      return H.z[lo];
    }
    const int16_t z = cut_of(j, R, i, T);
    len = z > 0 ? R.L[i * kMaxPasses + z - 1] : 0;
    return z;
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
      // the codeblocks not included yet, with their precinct and leaf: the
      // trees are only read here (estimates), so leaves are coded in
      // parallel, counted first and then written into their own words
      struct Leaf {
        uint64_t i, k;
        Precinct *pr;
      };
      std::vector<Leaf> leaves;
      for (auto &res : par.res)
        for (auto &sb : res.sb)
          for (uint64_t p = 0; p < res.nprec; ++p) {
            Precinct &pr = sb.prec[p];
            for (uint64_t k = 0; k < pr.ncbs; ++k) {
              const uint64_t i = pr.first_cb + k;
              if (R.cp[i * (nl + 1) + q] == 0) leaves.push_back({i, k, &pr});   // cp[q - 1]
            }
          }
      par_for(leaves.size(), [&](uint64_t a, uint64_t b) {
        for (uint64_t x = a; x < b; ++x) {
          const Leaf &lf = leaves[x];
          BitTally ti, tm;
          tagtree_encode(lf.pr->incl, ti, (uint32_t)q + 1, lf.k, 2);
          tagtree_encode_upto(lf.pr->msbs, tm, R.Kmsbs[lf.i] + 1u, lf.k, 2);
          BI.nbits[lf.i] = (uint32_t)ti.nbits;
          BM.nbits[lf.i] = (uint32_t)tm.nbits;
        }
      });
      // words per leaf: at least one (as a leaf with no bits had before)
      uint64_t wi = 0, wm = 0;
      for (const Leaf &lf : leaves) {
        BI.woff[lf.i] = (uint32_t)wi;
        wi += std::max<uint64_t>(1, (BI.nbits[lf.i] + 63) / 64);
        BM.woff[lf.i] = (uint32_t)wm;
        wm += std::max<uint64_t>(1, (BM.nbits[lf.i] + 63) / 64);
      }
      BI.words.assign(wi + 1, 0);
      BM.words.assign(wm + 1, 0);
      par_for(leaves.size(), [&](uint64_t a, uint64_t b) {
        for (uint64_t x = a; x < b; ++x) {
          const Leaf &lf = leaves[x];
          BitSpan bi{&BI.words[BI.woff[lf.i]]}, bm{&BM.words[BM.woff[lf.i]]};
          tagtree_encode(lf.pr->incl, bi, (uint32_t)q + 1, lf.k, 2);
          tagtree_encode_upto(lf.pr->msbs, bm, R.Kmsbs[lf.i] + 1u, lf.k, 2);
        }
      });
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
        if (est == 0 && q + 1 == nl) {
          // the last layer's header estimate is not used by anything after
          // it (the packets are written for real next); only the body
          // sizes it would accumulate are
          for (uint64_t p = 0; p < res.nprec; ++p) ql += body_size_only(j, res, p, q);
        } else {
          for (uint64_t p = 0; p < res.nprec; ++p) ql += estimate_packet(j, res, p, q, est + 1);
        }
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
    std::vector<std::vector<uint32_t>> cl;  // per param: body length L[cut - 1] where the cut is new
    std::vector<std::vector<uint8_t>> e;    // per param: contribution flag per (resolution, precinct)
    std::vector<std::vector<uint64_t>> eoff;
  };
  // per param: L[cut - 1] of the previous layer's cut per codeblock (0 when
  // nothing is included), set before each layer's probes
  std::vector<std::vector<uint32_t>> plen;
  void set_prev_lengths(int q) {
    plen.assign(c->info.nPar, {});
    for (int j = 0; j < c->info.nPar; ++j) {
      const CblkRecords &R = rec(j);
      plen[j].assign(R.n, 0);
      for (uint64_t i = 0; i < R.n; ++i) {
        const int16_t prev = R.cp[i * (nl + 1) + q];
        if (prev > 0) plen[j][i] = R.L[i * kMaxPasses + prev - 1];
      }
    }
  }
  void probe_init(ProbeBuf &pb) {
    const int np = c->info.nPar;
    pb.cq.assign(np, {});
    pb.cl.assign(np, {});
    pb.e.assign(np, {});
    pb.eoff.assign(np, {});
    for (int j = 0; j < np; ++j) {
      pb.cq[j].assign(rec(j).n, 0);
      pb.cl[j].assign(rec(j).n, 0);
      uint64_t o = 0;
      for (auto &res : tl->par[j].res) {
        pb.eoff[j].push_back(o);
        o += res.nprec;
      }
      pb.e[j].assign(o, 0);
    }
  }

  // length_code for est 2: lcur = L[cur - 1], lprev = L[prev - 1] (0 for prev 0)
  template <class W>
  void length_code_probe(W &w, const CblkRecords &R, uint64_t i, int16_t prev, int16_t cur, uint32_t lcur,
                         uint32_t lprev) const {
    const int16_t zc = (int16_t)(cur - 1), zp = (int16_t)(prev - 1);
    uint16_t beta = R.beta_est[i];
    uint16_t maxb = beta;
    const uint32_t len = lcur - lprev;
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
      const HullIndex &H = hull[j];
      const uint32_t *pl = plen[j].data();
      int16_t *cq = pb.cq[j].data();
      uint32_t *cl = pb.cl[j].data();
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
              uint32_t zl;
              const int16_t z = cut_len(j, R, i, T, zl);
              int16_t cur;
              if (z < prev) {
                cur = prev;
                if (cur == 1) ql += H.l0[i];
              } else {
                cur = z;
                if (z > 0) {
                  ql += zl;
                  if (z > 1 && prev > 0) ql -= pl[i];
                }
              }
              cq[i] = cur;
              cl[i] = zl;
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
                    tagtree_encode_upto(sb.prec[p].msbs, w, R.Kmsbs[i] + 1u, k, 2);
                }
              } else {
                w.bit((uint64_t)dz);
              }
              if (dz) {   // dz > 0: cur is this probe's cut, cl[i] its length
                delta_z_code(w, dz);
                length_code_probe(w, R, i, prev, cur, cl[i], prev > 0 ? pl[i] : 0);
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
      const std::function<void(size_t)> fn = [&](size_t a) { out[a] = probe(Ts[a], q, pbufs[a]); };
      pool_run(threads - 1, Ts.size(), fn);
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
  void check_layer(uint16_t slope, int l, int64_t target, int64_t base, std::unordered_map<uint32_t, int64_t> &m) {
    if (target > 0) {
      // the bisection has usually probed both thresholds already
      std::vector<uint16_t> Ts;
      if (!m.count(slope)) Ts.push_back(slope);
      if (slope > 0 && !m.count(slope - 1u)) Ts.push_back((uint16_t)(slope - 1));
      if (threads > 0) {
        if (!Ts.empty()) probe_many(Ts, l, m);
      } else {
        for (uint16_t T : Ts) m[T] = quality_layer(T, l, 1);
      }
      const int64_t f0 = m[slope] + base;
      const int64_t fm = slope > 0 ? m[(uint16_t)(slope - 1)] + base : INT64_MAX;
      if (!(f0 < target && fm > target)) chk->fallback = true;
    }
    // Per codeblock coded to a prefix of its passes, at thresholds t = T - 1
    // and T, with h its cut at t in the prefix coding:
    // (b) h lies `margin` passes before U, so the body bytes and the lengths
    //     up to the cut are those of a complete coding;
    // (c) no point after h can reach slope t in a complete coding, so the
    //     hull points up to h stay (a later point pops h only with a slope
    //     from h at least h's own, which is >= t) and none after h has slope
    //     >= t (slopes fall along the hull). A point i after h has slope
    //     (D[h..i]) / (L[i] - L[h - 1]) from h; D of the passes from U on
    //     (lengths there may change) and of the uncoded ones is bounded by
    //     their positive part and the tier-1 bound, and L[i] >= L[Zs - 1]
    //     with Zs = min(Z, U). Slopes below 2^e map below t (slope2log,
    //     tier1.c:2100-2130), so D < 2^e (L[Zs - 1] - L[h - 1]) suffices.
    // Codeblocks failing either are flagged (each once) to be coded further.
    const uint16_t ts[2] = {slope > 0 ? (uint16_t)(slope - 1) : (uint16_t)0, slope};
    const double lams[2] = {slope_floor(ts[0]), slope_floor(ts[1])};
    for (int j = 0; j < c->info.nPar; ++j) {
      CblkRecords &R = rec(j);
      // per codeblock failing at either threshold: flagged once (in chunks
      // on the probe threads; the flags are merged in codeblock order)
      std::vector<uint8_t> bad(R.n, 0);
      par_for(R.n, [&](uint64_t a, uint64_t b) {
        for (uint64_t i = a; i < b; ++i) {
          if (R.Z[i] >= R.Zfull[i]) continue;
          for (int k = 0; k < 2 && !bad[i]; ++k) {
            const int h = cut_of(j, R, i, ts[k]);
            const int Zs = std::min<int>(R.Z[i], R.U[i]);
            bool ok = h + chk->margin <= (int)R.U[i] && h < Zs && i < R.bfut.size();
            if (ok) {
              const double *ms = &R.mse[i * kMaxPasses];
              double D = R.bfut[i];
              for (int z = h; z < Zs; ++z) D += ms[z];
              for (int z = Zs; z < R.Z[i]; ++z) D += std::max(ms[z], 0.0);
              const uint32_t *L = &R.L[i * kMaxPasses];
              const double dL = (double)L[Zs - 1] - (h > 0 ? (double)L[h - 1] : 0.0);
              ok = D < lams[k] * dL;
            }
            if (!ok) bad[i] = 1;
          }
        }
      });
      for (uint64_t i = 0; i < R.n; ++i)
        if (bad[i]) chk->flagged[j].push_back(i);
    }
  }

  // A slope that slope2log maps to t or above is at least 2^e with
  // e = 32 + ceil((t - 65536 - 255) / 256): slope2log gives
  // 256 (e - 32) + f + 65536 for slope 1.m * 2^e, with f in [0, 255]. Below
  // 3 the value clamps (every positive slope reaches it): 0.
  static double slope_floor(uint16_t t) {
    if (t <= 2) return 0.0;
    const int num = (int)t - 65536 - 255;
    const int e = 32 + (num >= 0 ? (num + 255) / 256 : -((-num) / 256));
    return std::ldexp(1.0, e);
  }

  // BWC_AMD_RC_TRACE=1: phase times of rate control on stderr (measurement)
  bool trace = false;
  double t_last = 0;
  void phase(const char *name) {
    if (!trace) return;
    const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
                         .count();
    if (name) std::fprintf(stderr, "rc: %-22s %8.3f ms\n", name, t - t_last);
    t_last = t;
  }

  void rate_control() {
    if (const char *e = std::getenv("BWC_AMD_RC_TRACE")) trace = e[0] == '1';
    phase(nullptr);
    int64_t est_tile = 0, est_ql = 0;
    const uint64_t tsz = tl->box.volume() * (uint64_t)c->info.nPar;
    const uint64_t hdr = (uint64_t)std::ceil((float)((uint32_t)c->control.headerSize + tl->header_size) / nl);
    build_hull_index();
    phase("hull index");
    // size estimates of the thresholds probed for the current layer
    std::unordered_map<uint32_t, int64_t> memo;
    for (int l = 0; l < nl; ++l) {
      memo.clear();
      const int64_t target = (int64_t)std::ceil(c->control.bitrate[l] * tsz / 8.0f);
      uint16_t slope;
      set_prev_lengths(l);
      phase("prev lengths");
      if (target > 0.0f) {
        build_layer_cache(l);
        phase("layer cache");
        uint16_t smin = tl->slope_min, smax = tl->slope_max, tmp = tl->slope_min;
        slope = smin;
        // the reference's bisection, its probes estimated ahead on `threads`
        // host threads: the next three levels of its decision tree at once
        memo.clear();
        while (smin < smax) {
          slope = (uint16_t)((smax + smin) >> 1);
          int64_t f;
          if (threads > 0) {
            if (!memo.count(slope)) {
              std::vector<uint16_t> Ts;
              frontier(smin, smax, threads >= 15 ? 4 : (threads >= 7 ? 3 : (threads >= 3 ? 2 : 1)), Ts);
              probe_many(Ts, l, memo);
              phase("probe round");
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
      if (chk) check_layer(slope, l, target, est_tile + (int64_t)hdr, memo);
      phase("check layer");
      lc = nullptr;
      est_ql = quality_layer(slope, l, 0);
      est_tile += est_ql + (int64_t)hdr;
      phase("final layer");
    }
  }

  // What quality_layer(.., est 0) leaves besides the cuts, for cuts made
  // elsewhere (on the device): contribution flags, the MSB tree values (at
  // the first layer), the inclusion tree values (at the last) and the body
  // sizes of the packets (tier2.c:1030-1064, :751-849)
  void apply_cuts() {
    for (int j = 0; j < c->info.nPar; ++j) {
      Param &par = tl->par[j];
      CblkRecords &R = rec(j);
      for (int q = 0; q < nl; ++q)
        for (auto &res : par.res) {
          for (auto &sb : res.sb)
            for (uint64_t p = 0; p < res.nprec; ++p) {
              Precinct &pr = sb.prec[p];
              for (uint64_t k = 0; k < pr.ncbs; ++k) {
                const uint64_t i = pr.first_cb + k;
                const int16_t *cp = &R.cp[i * (nl + 1) + 1];
                if (cp[q] > cp[q - 1]) res.pk[q * res.nprec + p].e = 1;
                if (q == 0) pr.msbs.set(k, R.Kmsbs[i]);
                if (q + 1 == nl) {
                  uint8_t l = 0;
                  while (l < nl && cp[l] == 0) ++l;
                  pr.incl.set(k, l);
                }
              }
            }
          for (uint64_t p = 0; p < res.nprec; ++p) body_size_only(j, res, p, q);
        }
    }
  }

  // the packets of each (parameter, resolution, precinct) in layer order;
  // different precincts share no state (their codeblocks' beta, their tag
  // trees, their packets), so they are written on the probe threads
  //
  // A subband's precinct has its own tag trees, so the header bits of one
  // packet are the concatenation of independent per-subband bit strings:
  // those are written on the threads (each subband's layers in order), then
  // each packet's strings go through the 0xFF-stuffing writer in order.
  void packets() {
    struct Group {
      int j;
      Resolution *res;
      uint64_t p;
    };
    std::vector<Group> gs;
    struct Part {
      size_t g;
      size_t s;
    };
    std::vector<Part> parts;
    std::vector<size_t> first;   // first part of each group
    for (int j = 0; j < c->info.nPar; ++j)
      for (auto &res : tl->par[j].res)
        for (uint64_t p = 0; p < res.nprec; ++p) {
          first.push_back(parts.size());
          for (size_t sb = 0; sb < res.sb.size(); ++sb) parts.push_back({gs.size(), sb});
          gs.push_back({j, &res, p});
        }
    first.push_back(parts.size());
    // raw[part * nl + layer]: the part's bits in that layer's packet
    std::vector<std::vector<uint64_t>> raw(parts.size() * nl);
    std::vector<uint64_t> rawbits(parts.size() * nl, 0);
    const std::function<void(size_t)> sub = [&](size_t k) {
      const Group &g = gs[parts[k].g];
      Subband &sb = g.res->sb[parts[k].s];
      for (int l = 0; l < nl; ++l) {
        if (!g.res->pk[l * g.res->nprec + g.p].e) continue;
        BitString w(&raw[k * nl + l]);
        subband_bits(w, g.j, sb, g.p, l, 0, nullptr);
        rawbits[k * nl + l] = w.nbits;
      }
    };
    const std::function<void(size_t)> join = [&](size_t m) {
      const size_t gi = m / nl;
      const int l = (int)(m % nl);
      const Group &g = gs[gi];
      Packet &pk = g.res->pk[l * g.res->nprec + g.p];
      pk.built = true;
      if (!pk.e) {
        pk.header.assign(1, 0);
        pk.size = 1;
        return;
      }
      BitWriter w;
      w.bit(1);
      for (size_t k = first[gi]; k < first[gi + 1]; ++k) {
        const std::vector<uint64_t> &words = raw[k * nl + l];
        uint64_t left = rawbits[k * nl + l];
        for (size_t i = 0; left; ++i) {
          const int n = left < 64 ? (int)left : 64;
          const uint64_t v = words[i] >> (64 - n);
          if (n > 32) {
            w.put(v >> 32, n - 32);
            w.put(v & 0xFFFFFFFFull, 32);
          } else {
            w.put(v, n);
          }
          left -= (uint64_t)n;
        }
      }
      w.flush();
      pk.header = std::move(w.sink.out);
      pk.size = (uint32_t)(pk.body_size + pk.header.size());
    };
    if (threads > 1 && parts.size() > 1) {
      pool_run(threads - 1, parts.size(), sub);
      pool_run(threads - 1, gs.size() * nl, join);
    } else {
      for (size_t k = 0; k < parts.size(); ++k) sub(k);
      for (size_t m = 0; m < gs.size() * nl; ++m) join(m);
    }
  }
};

}  // namespace

// Rate control for one tile: fills cp (cut points per layer) and the packet
// contribution flags / body sizes. Needs tile slope bounds and records.
// host threads for the bisection probes: BWC_AMD_RC_THREADS (0 = the
// sequential in-place probes, 1 = side-effect-free probes on one thread); by
// default one less than the host threads, at most 15 (the bisection's next
// four levels at once)
void host_pool_run(int workers, size_t n, const std::function<void(size_t)> &fn) { pool_run(workers, n, fn); }

int rc_thread_count() {
  if (const char *e = std::getenv("BWC_AMD_RC_THREADS")) return std::max(0, std::min(16, std::atoi(e)));
  const unsigned hc = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(15u, hc > 1 ? hc - 1 : 1u));
}

int t2_rate_control(bwc_codec *codec, Job *job, uint64_t tile, SkipCheck *chk) {
  T2 t{codec, job, &job->plan.tiles[tile], tile, codec->control.nLayers};
  t.threads = rc_thread_count();
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
  t.threads = rc_thread_count();
  t.packets();
  return 0;
}

int t2_apply_cuts(bwc_codec *codec, Job *job, uint64_t tile) {
  T2 t{codec, job, &job->plan.tiles[tile], tile, codec->control.nLayers};
  t.apply_cuts();
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

void packet_body_chunks(bwc_codec *codec, Job *job, uint64_t tile, int j, int r, uint64_t p, int l,
                        std::vector<std::pair<uint64_t, uint64_t>> &out) {
  const int nl = codec->control.nLayers;
  Resolution &res = job->plan.tiles[tile].par[j].res[r];
  const CblkRecords &R = job->rec[tile * codec->info.nPar + j];
  for (auto &sb : res.sb) {
    const Precinct &pr = sb.prec[p];
    for (uint64_t c = 0; c < pr.ncbs; ++c) {
      const uint64_t i = pr.first_cb + c;
      const int16_t *cp = &R.cp[i * (nl + 1) + 1];
      if (cp[l] - cp[l - 1] <= 0) continue;
      const uint32_t *L = &R.L[i * kMaxPasses];
      const uint32_t start = (cp[l - 1] == 0) ? 0 : L[cp[l - 1] - 1];
      const uint32_t end = (cp[l] == 0) ? 0 : L[cp[l] - 1];
      out.emplace_back(R.src_off[i] + start, (uint64_t)(end - start));
    }
  }
}

// Packet sequencing and codestream assembly (codestream.c:139-570,
// :1329-1413). Body bytes are copied straight from the packed codeblock data,
// or per packet through `bodies`.
size_t assemble_codestream(bwc_codec *codec, Job *job, const bwc_stream *st, uint8_t *out,
                           const BodySource *bodies) {
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
            if (pk.e && bodies) {
              if (pk.body_size) {
                if (w.err || w.L + pk.body_size > w.Lmax ||
                    bodies->copy(bodies->ctx, w.m + w.L, pk.dev_off, pk.body_size)) {
                  w.err = true;
                } else {
                  w.L += pk.body_size;
                }
              }
            } else if (pk.e) {
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
