// Codec configuration helpers and the codec plan (tiles, resolutions,
// subbands, precincts, codeblocks).
//
// Semantics follow the reference exactly, including its single-precision
// float casts in the geometry (libbwc.c:209-216, :292-299, :430-437,
// :556-587, :1396-1399, :1539-1561) which diverge from integer arithmetic
// above 2^24, and the process-global energy-gain table that accumulates on
// every initialisation (dwt.c:80, :808-1066).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "codec.hpp"

namespace bwc {

namespace {
double g_lut[3][10];
std::mutex g_lut_mu;

// synthesis filters of the three kernels as float literals (dwt.h:77-117)
const double kF97[2][5] = {{-0.091271763114250f, -0.057543526228500f, 0.591271763114250f, 1.115087052457000f, 0},
                           {0.053497514821620f, 0.033728236885750f, -0.156446533057980f, -0.533728236885750f,
                            1.205898036472721f}};
const double kF53[2][5] = {{0.5f, 1.0f, 0, 0, 0}, {-0.125f, -0.25f, 0.75f, 0, 0}};
const double kFHaar[2][5] = {{1, 1, 0, 0, 0}, {-0.5, 0.5, 0, 0, 0}};
}  // namespace

void warn(const char *msg) { std::fprintf(stderr, "bwc: warning: %s\n", msg); }

// Energy of the iterated synthesis filters for levels 1..4 of each kernel.
// Mirrors initialize_gain_lut (dwt.c:808-1066): entries 0 and 5 are reset to
// one, all others are accumulated into the global table in the same order.
void gain_lut_accumulate() {
  std::lock_guard<std::mutex> lk(g_lut_mu);
  std::vector<double> A(512, 0.0), B(512, 0.0);
  double *cur = A.data(), *nxt = B.data();
  for (int f = 0; f < 3; ++f) g_lut[f][0] = g_lut[f][5] = 1.0f;
  for (int kernel = 0; kernel < 2; ++kernel) {
    const int lens[2] = {kernel ? 1 : 3, kernel ? 2 : 4};
    const double *taps = kernel ? &kF53[0][0] : &kF97[0][0];
    double *entry = &g_lut[kernel][1];
    const int lfb = lens[0];
    int lgb = lens[0];
    for (int band = 0; band < 2; ++band) {
      for (int k = 0; k <= lgb; ++k) cur[k] = cur[2 * lgb - k] = taps[k + 5 * band];
      for (int k = 0; k < 2 * lgb + 1; ++k) entry[0] += cur[k] * cur[k];
      for (int lvl = 1; lvl < 4; ++lvl) {
        for (int l = 0; l < 2 * lgb + 1; ++l) {
          nxt[2 * l + lfb] += cur[l] * taps[lfb];
          for (int m = 0; m < lfb; ++m) {
            nxt[2 * l + m] += cur[l] * taps[m];
            nxt[2 * l + (2 * lfb - m)] += cur[l] * taps[m];
          }
        }
        std::swap(cur, nxt);
        std::memset(nxt, 0, 512 * sizeof(double));
        lgb = 2 * lgb + lfb;
        for (int l = 0; l < 2 * lgb + 1; ++l) entry[lvl] += cur[l] * cur[l];
      }
      entry += 5;
      lgb = lens[1];
    }
    std::memset(cur, 0, 512 * sizeof(double));
  }
  {
    const double *taps = &kFHaar[0][0];
    double *entry = &g_lut[2][1];
    const int lfb = 2;
    int lgb = 2;
    for (int band = 0; band < 2; ++band) {
      for (int k = 0; k < lgb; ++k) cur[k] = taps[k + 5 * band];
      for (int k = 0; k < lgb; ++k) entry[0] += cur[k] * cur[k];
      for (int lvl = 1; lvl < 4; ++lvl) {
        for (int l = 0; l < lgb; ++l)
          for (int m = 0; m < lfb; ++m) nxt[2 * l + m] += cur[l] * taps[m];
        std::swap(cur, nxt);
        std::memset(nxt, 0, 512 * sizeof(double));
        lgb = lgb * lfb;
        for (int l = 0; l < lgb; ++l) entry[lvl] += cur[l] * cur[l];
      }
      entry += 5;
      lgb = 2;
    }
    std::memset(cur, 0, 512 * sizeof(double));
  }
}

const double (*gain_lut())[10] { return g_lut; }

void reset_gain_lut() {
  std::lock_guard<std::mutex> lk(g_lut_mu);
  std::memset(g_lut, 0, sizeof g_lut);
}

// get_dwt_energy_gain (dwt.c:1082-1202)
double energy_gain(const bwc_codec *c, const double lut[3][10], uint8_t hb, uint16_t level) {
  const uint8_t dec[4] = {c->control.decompX, c->control.decompY, c->control.decompZ, c->control.decompTS};
  const int kern[4] = {c->control.KernelX, c->control.KernelY, c->control.KernelZ, c->control.KernelTS};
  double G = 1.0f;
  for (int a = 0; a < 4; ++a) {
    int lv = level > dec[a] ? dec[a] : level;
    if (lv == 0) continue;
    while (lv > 4) {
      G *= 2.0f;
      --lv;
    }
    G *= lut[kern[a]][lv + ((hb >> a) & 1) * 5];
  }
  return G;
}

// quantisation exponent/mantissa of a step size (libbwc.c:1957-1979)
void compute_qt(bwc_gl_ctrl *ctl, double delta) {
  for (ctl->qt_exponent = 0; delta < 1; ++ctl->qt_exponent, delta *= 2) {
  }
  ctl->qt_mantissa = (uint16_t)std::floor(0.5f + ((delta - 1.0f) * (1 << 16)));
  if (ctl->qt_exponent > kPrecBits) {
    ctl->qt_exponent = kPrecBits;
    ctl->qt_mantissa = 0;
  }
  if (ctl->qt_mantissa >= 1 << 16) {
    if (ctl->qt_exponent == 0) {
      ctl->qt_mantissa = (1 << 16) - 1;
    } else {
      ctl->qt_exponent--;
      ctl->qt_mantissa = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// tag trees (tagtree.c:106-298)
// ---------------------------------------------------------------------------
void TagTree::build(uint64_t lx, uint64_t ly, uint64_t lz, uint64_t lt) {
  std::vector<uint32_t> n[4];
  uint32_t nx = (uint32_t)lx, ny = (uint32_t)ly, nz = (uint32_t)lz, nt = (uint32_t)lt;
  uint64_t total = 1;
  n[0].push_back(nx), n[1].push_back(ny), n[2].push_back(nz), n[3].push_back(nt);
  while ((uint64_t)nx * ny * nz * nt > 1) {
    total += (uint64_t)nx * ny * nz * nt;
    nx = (nx + 1) >> 1, ny = (ny + 1) >> 1, nz = (nz + 1) >> 1, nt = (nt + 1) >> 1;
    n[0].push_back(nx), n[1].push_back(ny), n[2].push_back(nz), n[3].push_back(nt);
  }
  value.assign(total, 0xFFFF);
  threshold.assign(total, 0);
  parent.assign(total, -1);
  uint64_t node = 0, base = 0;
  for (size_t lev = 0; lev + 1 < n[0].size(); ++lev) {
    const uint64_t cnt = (uint64_t)n[0][lev] * n[1][lev] * n[2][lev] * n[3][lev];
    const uint64_t pb = base + cnt;
    const uint64_t px = n[0][lev + 1], py = n[1][lev + 1], pz = n[2][lev + 1];
    for (uint32_t l = 0; l < n[3][lev]; ++l)
      for (uint32_t k = 0; k < n[2][lev]; ++k)
        for (uint32_t j = 0; j < n[1][lev]; ++j)
          for (uint32_t i = 0; i < n[0][lev]; ++i)
            parent[node++] = (int64_t)(pb + (i >> 1) + px * ((j >> 1) + py * ((k >> 1) + pz * (l >> 1))));
    base = pb;
  }
}

void TagTree::set(uint64_t leaf, uint16_t v) {
  int64_t n = (int64_t)leaf;
  do {
    value[n] = v;
    n = parent[n];
  } while (n >= 0 && value[n] > v);
}

// ---------------------------------------------------------------------------
// plan (create_codec, libbwc.c:1345-1716; initialize_subband :360-607;
// initialize_precinct :161-306)
// ---------------------------------------------------------------------------
namespace {

uint8_t popcount4(uint8_t v) { return (uint8_t)__builtin_popcount(v & 0xF); }

struct Ctx {
  bwc_codec *c;
  Job *job;
  uint8_t dec[4], ps[4], cbs[4];
  int ndec;
};

void init_precinct(Ctx &x, Param &par, Precinct &pr, uint32_t r, uint32_t m, uint32_t p, const uint32_t d[4]) {
  const uint64_t cs[4] = {1ull << x.cbs[0], 1ull << x.cbs[1], 1ull << x.cbs[2], 1ull << x.cbs[3]};
  for (int a = 0; a < 4; ++a) {
    const uint32_t lo = pr.box[a], hi = pr.box[4 + a];
    pr.ncb[a] = (uint16_t)((lo == hi) ? 0 : (std::ceil((float)hi / cs[a]) - std::floor((float)lo / cs[a])));
  }
  pr.ncbs = (uint64_t)pr.ncb[0] * pr.ncb[1] * pr.ncb[2] * pr.ncb[3];
  pr.first_cb = par.ncb;
  if (!pr.ncbs) return;
  pr.msbs.build(pr.ncb[0], pr.ncb[1], pr.ncb[2], pr.ncb[3]);
  pr.incl.build(pr.ncb[0], pr.ncb[1], pr.ncb[2], pr.ncb[3]);
  uint32_t fl[4];
  for (int a = 0; a < 4; ++a) fl[a] = (uint32_t)std::floor((float)pr.box[a] / cs[a]);
  uint64_t k = 0;
  for (uint16_t ct = 0; ct < pr.ncb[3]; ++ct)
    for (uint16_t cz = 0; cz < pr.ncb[2]; ++cz)
      for (uint16_t cy = 0; cy < pr.ncb[1]; ++cy)
        for (uint16_t cx = 0; cx < pr.ncb[0]; ++cx, ++k) {
          const uint16_t ci[4] = {cx, cy, cz, ct};
          for (int a = 0; a < 4; ++a) {
            const uint64_t lo = std::max<uint64_t>(pr.box[a], cs[a] * (ci[a] + fl[a]));
            const uint64_t hi = std::min<uint64_t>(pr.box[4 + a], cs[a] * (ci[a] + fl[a] + 1));
            par.cb_box.push_back(d[a] + (uint32_t)lo);
            (void)hi;
          }
          for (int a = 0; a < 4; ++a) {
            const uint64_t hi = std::min<uint64_t>(pr.box[4 + a], cs[a] * (ci[a] + fl[a] + 1));
            par.cb_box.push_back(d[a] + (uint32_t)hi);
          }
          par.cb_sbref.push_back((r << 8) | m);
          par.cb_prec.push_back(p);
          par.cb_idx.push_back((uint32_t)k);
        }
  par.ncb += pr.ncbs;
}

void init_subband(Ctx &x, Param &par, uint32_t r, uint32_t m, uint8_t hb) {
  bwc_gl_ctrl &ctl = x.c->control;
  Resolution &res = par.res[r];
  Subband &sb = res.sb[m];
  const int dl = x.ndec - (int)r + 1;
  for (int a = 0; a < 4; ++a) {
    const int lv = dl > x.dec[a] ? x.dec[a] : dl;
    const int h = (hb >> a) & 1;
    sb.box.lo[a] = (uint64_t)std::ceil(((float)par.box.lo[a] / (1 << lv)) - 0.5 * h);
    sb.box.hi[a] = (uint64_t)std::ceil(((float)par.box.hi[a] / (1 << lv)) - 0.5 * h);
  }
  const uint32_t Rb = kPrecBits + popcount4(hb);
  sb.hb = hb;
  sb.gain = energy_gain(x.c, x.job->lut, hb, (uint16_t)dl);
  if (ctl.quantization_style == bwc_qt_none) {
    sb.qt_exp = (uint8_t)Rb;
    sb.qt_mant = 0;
    sb.step = 1;
    sb.step_eff = 1;
  } else {
    sb.qt_exp = (uint8_t)((uint8_t)ctl.qt_exponent + ((r == 0) ? 0 : 1 - (int)r));
    sb.qt_mant = (uint16_t)ctl.qt_mantissa;
    sb.step = (1.0f + ((double)sb.qt_mant / (1 << 11))) / std::pow(2, sb.qt_exp);
    sb.step_eff = (1.0f + ((double)sb.qt_mant / (1 << 11)));
    if (sb.qt_exp > Rb)
      sb.step_eff /= std::pow(2, sb.qt_exp - Rb);
    else
      sb.step_eff *= std::pow(2, Rb - sb.qt_exp);
  }
  sb.Kmax = (uint16_t)std::max(0, ctl.guard_bits + sb.qt_exp - 1);

  uint64_t lowsize[4] = {0, 0, 0, 0};
  for (int a = 0; a < 4; ++a)
    if ((hb >> a) & 1) lowsize[a] = par.res[r - 1].box.len(a);
  if (!res.nprec) return;
  sb.prec.resize(res.nprec);
  const uint64_t psz[4] = {1ull << x.ps[0], 1ull << x.ps[1], 1ull << x.ps[2], 1ull << x.ps[3]};
  uint32_t d[4];
  for (int a = 0; a < 4; ++a) d[a] = (uint32_t)((uint32_t)lowsize[a] - sb.box.lo[a]);
  uint64_t p = 0;
  for (uint16_t pt = 0; pt < res.np[3]; ++pt)
    for (uint16_t pz = 0; pz < res.np[2]; ++pz)
      for (uint16_t py = 0; py < res.np[1]; ++py)
        for (uint16_t px = 0; px < res.np[0]; ++px, ++p) {
          Precinct &pr = sb.prec[p];
          const uint16_t pi[4] = {px, py, pz, pt};
          for (int a = 0; a < 4; ++a) {
            const uint64_t rl = res.box.lo[a], rh = res.box.hi[a];
            uint32_t lo = (uint32_t)std::max<uint64_t>(rl, psz[a] * (pi[a] + (uint32_t)std::floor(rl / psz[a])));
            uint32_t hi = (uint32_t)std::min<uint64_t>(rh, psz[a] * (pi[a] + (uint32_t)std::floor(rl / psz[a]) + 1));
            if ((x.ndec - x.dec[a]) < (int)r) {
              const int h = (hb >> a) & 1;
              lo = (uint32_t)std::ceil(((float)lo / 2) - 0.5 * h);
              hi = (uint32_t)std::ceil(((float)hi / 2) - 0.5 * h);
            }
            pr.box[a] = lo;
            pr.box[4 + a] = hi;
          }
          init_precinct(x, par, pr, r, m, (uint32_t)p, d);
        }
}

}  // namespace

int build_plan(bwc_codec *codec, Job *job) {
  bwc_gl_ctrl &ctl = codec->control;
  bwc_gl_inf &inf = codec->info;
  Ctx x{codec, job, {ctl.decompX, ctl.decompY, ctl.decompZ, ctl.decompTS},
        {ctl.precSizeX, ctl.precSizeY, ctl.precSizeZ, ctl.precSizeTS},
        {ctl.cbX, ctl.cbY, ctl.cbZ, ctl.cbTS}, ctl.nDecomp};
  const uint64_t n[4] = {inf.nX, inf.nY, inf.nZ, inf.nTS};
  const uint64_t ts[4] = {ctl.tileSizeX, ctl.tileSizeY, ctl.tileSizeZ, ctl.tileSizeTS};
  uint32_t nt[4];
  for (int a = 0; a < 4; ++a) nt[a] = (uint32_t)(int)std::ceil(((float)n[a] / ts[a]));
  // Resolution boxes are computed in single precision (libbwc.c:1539-1546),
  // so tile bounds above 2^24 round to the nearest float. Where that makes a
  // tile's finest resolution box longer than the tile, the transform would
  // run past the tile's samples (the reference reads and writes beyond its
  // tile buffer there): such extents are refused. Shorter boxes are kept,
  // as in the reference.
  for (int a = 0; a < 4; ++a)
    for (uint32_t t = 0; t < nt[a]; ++t) {
      const uint64_t lo = (uint64_t)t * ts[a], hi = std::min<uint64_t>(n[a], (uint64_t)(t + 1) * ts[a]);
      if ((uint64_t)std::ceil((float)hi) - (uint64_t)std::ceil((float)lo) > hi - lo) return 2;
    }
  Plan &plan = job->plan;
  plan.codec = codec;
  plan.tiles.assign(ctl.nTiles, Tile());
  uint32_t i = 0;
  for (uint32_t t = 0; t < nt[3]; ++t)
    for (uint32_t z = 0; z < nt[2]; ++z)
      for (uint32_t y = 0; y < nt[1]; ++y)
        for (uint32_t xx = 0; xx < nt[0]; ++xx, ++i) {
          if (i >= plan.tiles.size()) return 1;
          Tile &tl = plan.tiles[i];
          const uint32_t ti[4] = {xx, y, z, t};
          for (int a = 0; a < 4; ++a) {
            tl.box.lo[a] = (uint64_t)ti[a] * ts[a];
            tl.box.hi[a] = std::min<uint64_t>(n[a], (uint64_t)(ti[a] + 1) * ts[a]);
          }
          tl.index = i;
          tl.header_size = 18;
          tl.slope_max = 0x0000;
          tl.slope_min = 0xFFFF;
          tl.par.assign(inf.nPar, Param());
          uint64_t npk = 0;
          for (uint8_t j = 0; j < inf.nPar; ++j) {
            Param &par = tl.par[j];
            par.pmax = -1.7976931348623157e+308;
            par.pmin = 1.7976931348623157e+308;
            par.box = tl.box;
            par.res.assign(x.ndec + 1, Resolution());
            for (int r = 0; r < x.ndec + 1; ++r) {
              Resolution &res = par.res[r];
              for (int a = 0; a < 4; ++a) {
                const int sh = std::min(x.ndec - r, (int)x.dec[a]);
                res.box.lo[a] = (uint64_t)std::ceil((float)par.box.lo[a] / (1 << sh));
                res.box.hi[a] = (uint64_t)std::ceil((float)par.box.hi[a] / (1 << sh));
                res.np[a] = (uint16_t)((res.box.lo[a] == res.box.hi[a])
                                           ? 0
                                           : (std::ceil((float)res.box.hi[a] / (1 << x.ps[a]) -
                                                        std::floor((float)res.box.lo[a] / (1 << x.ps[a])))));
              }
              res.nprec = (uint64_t)res.np[0] * res.np[1] * res.np[2] * res.np[3];
              npk += res.nprec * ctl.nLayers;
              uint8_t hb = 0;
              for (int a = 0; a < 4; ++a)
                if (x.ndec - x.dec[a] < r && r > 0) hb |= (uint8_t)(1 << a);
              int nsb = 1 << popcount4(hb);
              nsb -= (nsb & 1) ? 0 : 1;
              res.sb.assign(nsb, Subband());
              res.pk.assign(res.nprec * ctl.nLayers, Packet());
              for (int l = 0, m = 0; m < nsb; ++l) {
                if (r == 0 || ((l & hb) == l && l != 0)) {
                  init_subband(x, par, (uint32_t)r, (uint32_t)m, (uint8_t)l);
                  ++m;
                }
              }
            }
          }
          tl.npackets = npk;
          tl.header_size += ctl.error_resilience ? (uint32_t)(8 * npk) : 0;
        }
  return 0;
}

Job *job_of(bwc_codec *c) { return static_cast<Job *>(c->tile); }

void CblkRecords::resize(uint64_t ncb, int nlayers) {
  n = ncb;
  K.assign(ncb, 0);
  Kmsbs.assign(ncb, 0);
  Z.assign(ncb, 0);
  Zfull.assign(ncb, 0);
  U.assign(ncb, 0);
  // tier-1 writes the rows of L and S that are read (L[0, Z), S[0, Z]);
  // a record reused at the same size keeps its allocation unfilled
  if (L.size() != ncb * kMaxPasses) L.assign(ncb * kMaxPasses, 0);
  if (S.size() != ncb * (kMaxPasses + 1)) S.assign(ncb * (kMaxPasses + 1), 0);
  beta.assign(ncb, 3);
  beta_est.assign(ncb, 3);
  cp.assign(ncb * (nlayers + 1), 0);
  data_off.assign(ncb, 0);
}

}  // namespace bwc
