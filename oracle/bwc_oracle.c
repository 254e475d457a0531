/*
 * bwc_oracle.c — CPU restatement of the BigWhoop compression path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker the HIP product is compared
 * against. It is loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product library.
 *
 * The code is a plain scalar restatement written from the reference sources
 * (paths relative to the reference checkout). Floating-point expressions keep
 * the reference's operand types and order (float vs double casts in the
 * geometry, float-rounded lifting constants, no FMA: built with
 * -ffp-contract=off), because quantisation and slope rounding depend on them.
 *
 * Parity: pinned byte-for-byte against the reference built from source
 * (oracle/_ref) through tests/golden fixtures and tests/test_oracle_*.py.
 */
#include "bwc_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXP        190          /* (64*3)-2 coding passes, prim_types_double.h */
#define PREC_BITS   63
#define PREC_MANT   52
#define PREC_BYTES  8
#define EXP_MASK    0x7FF0000000000000ull
#define MANT_MASK   0x000FFFFFFFFFFFFFull
#define FLTMAX      1.7976931348623157e+308
#define MX(a, b)    (((a) < (b)) ? (b) : (a))
#define MN(a, b)    (((a) > (b)) ? (b) : (a))

/* ====================================================================== */
/* gain LUT (dwt.c:80, dwt.c:808-1066, dwt.c:1082-1202)                   */
/* ====================================================================== */

/* Lifting and synthesis-filter constants are float literals promoted to
   double in the reference (dwt.h:77-117); keep them as float literals. */
#define K_ALPHA   -1.586134342059924f
#define K_BETA    -0.052980118572961f
#define K_GAMMA    0.882911075530934f
#define K_DELTA    0.360523644801462f
#define K_KAPPA_H  1.230174104914001f
#define K_KAPPA_L  0.812893066115961f

static double g_gain_lut[3][10];   /* process-global like DWT_ENERGY_GAIN_LUT */

/* One call of initialize_gain_lut: entries [0] and [5] are set to 1, every
   other entry is *accumulated* into the process global (dwt.c:864-1058). */
static void gain_lut_accumulate(void)
{
  static const int len97[2] = {3, 4}, len53[2] = {1, 2}, lenhaar[2] = {2, 2};
  static const double f97[2][5] = {{-0.091271763114250f, -0.057543526228500f, 0.591271763114250f, 1.115087052457000f, 0},
                                   {0.053497514821620f, 0.033728236885750f, -0.156446533057980f, -0.533728236885750f,
                                    1.205898036472721f}};
  static const double f53[2][5] = {{0.5f, 1.0f, 0, 0, 0}, {-0.125f, -0.25f, 0.75f, 0, 0}};
  static const double fhaar[2][5] = {{1, 1, 0, 0, 0}, {-0.5, 0.5, 0, 0, 0}};
  double *m1 = calloc(512, sizeof(double)), *m2 = calloc(512, sizeof(double));
  double *b1 = m1, *b2 = m2, *tmp;
  for (int f = 0; f < 3; ++f) g_gain_lut[f][0] = g_gain_lut[f][5] = 1.0f;

  for (int i = 0; i < 2; ++i)
  {
    double *lut = &g_gain_lut[i][1];
    const int *len = i ? len53 : len97;
    const double *taps = i ? &f53[0][0] : &f97[0][0];
    int lfb = len[0], lgb = len[0];
    for (int j = 0; j < 2; ++j)
    {
      for (int k = 0; k < lgb + 1; ++k) b1[k] = b1[2 * lgb - k] = taps[k + 5 * j];
      for (int k = 0; k < 2 * lgb + 1; ++k) lut[0] += b1[k] * b1[k];
      for (int k = 1; k < 4; ++k)
      {
        for (int l = 0; l < 2 * lgb + 1; ++l)
        {
          b2[2 * l + lfb] += b1[l] * taps[lfb];
          for (int m = 0; m < lfb; ++m)
          {
            b2[2 * l + m] += b1[l] * taps[m];
            b2[2 * l + (2 * lfb - m)] += b1[l] * taps[m];
          }
        }
        tmp = b1; b1 = b2; b2 = tmp;
        memset(b2, 0, 512 * sizeof(double));
        lgb = 2 * lgb + lfb;
        for (int l = 0; l < 2 * lgb + 1; ++l) lut[k] += b1[l] * b1[l];
      }
      lut += 5;
      lgb = len[1];
    }
    memset(b1, 0, 512 * sizeof(double));
  }
  {
    double *lut = &g_gain_lut[2][1];
    const int *len = lenhaar;
    const double *taps = &fhaar[0][0];
    int lfb = len[0], lgb = len[0];
    for (int j = 0; j < 2; ++j)
    {
      for (int k = 0; k < lgb; ++k) b1[k] = taps[k + 5 * j];
      for (int k = 0; k < lgb; ++k) lut[0] += b1[k] * b1[k];
      for (int k = 1; k < 4; ++k)
      {
        for (int l = 0; l < lgb; ++l)
          for (int m = 0; m < lfb; m++) b2[2 * l + m] += b1[l] * taps[m];
        tmp = b1; b1 = b2; b2 = tmp;
        memset(b2, 0, 512 * sizeof(double));
        lgb = lgb * lfb;
        for (int l = 0; l < lgb; ++l) lut[k] += b1[l] * b1[l];
      }
      lut += 5;
      lgb = len[1];
    }
    memset(b1, 0, 512 * sizeof(double));
  }
  free(m1);
  free(m2);
}

void oc_gain_lut(int ncalls, double lut[30])
{
  memset(g_gain_lut, 0, sizeof g_gain_lut);
  for (int i = 0; i < ncalls; ++i) gain_lut_accumulate();
  memcpy(lut, g_gain_lut, sizeof g_gain_lut);
}

/* ====================================================================== */
/* codec data model                                                        */
/* ====================================================================== */
typedef struct
{
  uint64_t x0, y0, z0, t0, x1, y1, z1, t1;
} box_t;

typedef struct
{
  uint16_t value, threshold;
  int64_t parent;   /* -1 for the root */
} tt_node;

typedef struct
{
  uint64_t nleaves;
  uint64_t nnodes;
  tt_node *nodes;
} tagtree;

typedef struct
{
  box_t b;                 /* tile-local buffer coordinates */
  uint32_t idx;
  uint64_t aidx;           /* position in the parameter's access order */
  uint8_t K, Kmsbs, Z;
  uint64_t L[MAXP];
  uint16_t S[MAXP + 1];
  uint8_t *data;
  uint16_t beta, beta_est;
  int16_t cp_mem[256 + 1]; /* cp[-1] == 0 (libbwc.c:274-282) */
} cblk_t;

typedef struct
{
  uint32_t x0, y0, z0, t0, x1, y1, z1, t1;
  uint16_t ncb[4];
  uint64_t ncbs;
  cblk_t *cb;
  tagtree *incl, *msbs;
} prec_t;

typedef struct
{
  box_t b;
  double dwt_gain, step, step_eff;
  uint8_t hb;
  uint16_t Kmax;
  uint8_t qt_exp;
  uint16_t qt_mant;
  prec_t *prec;
} subb_t;

typedef struct
{
  uint8_t *hdr;
  uint64_t hdr_size;
  uint64_t body_size;
  uint8_t *body;
  uint64_t body_pos;
  uint8_t e;
  uint32_t size;
} packet_t;

typedef struct
{
  box_t b;
  uint16_t np[4];
  uint64_t nprec;
  uint8_t nsb;
  subb_t *sb;
  packet_t *pk;
} res_t;

typedef struct
{
  subb_t *sb;
  prec_t *pr;
  cblk_t *cb;
} access_t;

typedef struct
{
  box_t b;
  double pmin, pmax, alpha, beta;
  res_t *res;
  uint64_t ncb;
  access_t *acc;
} param_t;

typedef struct
{
  box_t b;
  uint32_t index;
  uint32_t header_size;
  uint64_t body_size;
  uint64_t npk;
  uint16_t slope_min, slope_max;
  param_t *par;
  packet_t **seq;
} tile_t;

struct oc_codec
{
  uint64_t n[4];
  uint8_t nPar;
  uint8_t data_prec;
  uint16_t CSsgc;
  uint8_t erres;
  uint64_t ts[4];
  uint64_t ntiles;
  uint8_t ps[4], cbs[4], kern[4], dec[4], ndec;
  uint32_t qt_exp, qt_mant;
  uint8_t Qm, guard;
  float bitrate[256];
  uint8_t nlayers;
  uint64_t header_size, cs_size;
  uint8_t qstyle, prog;
  uint8_t *aux, *com;       /* SAX / COM payloads (codestream.c:239-258) */
  uint32_t naux;
  uint16_t ncom;
  tile_t *tile;
  double lut[3][10];   /* snapshot of the global LUT at geometry creation */
};

static double energy_gain(const oc_codec *c, uint8_t hb, uint16_t level)
{
  double G = 1.0f;
  for (int a = 0; a < 4; ++a)
  {
    int lv = (level > c->dec[a]) ? c->dec[a] : level;
    if (lv != 0)
    {
      while (lv > 4) { G *= 2.0f; lv--; }
      G *= c->lut[c->kern[a]][lv + ((hb >> a) & 1) * 5];
    }
  }
  return G;
}

/* quantisation exponent/mantissa from delta (libbwc.c:1957-1979) */
static void set_qt(oc_codec *c, double delta)
{
  for (c->qt_exp = 0; delta < 1; ++c->qt_exp, delta *= 2);
  c->qt_mant = (uint16_t)floor(0.5f + ((delta - 1.0f) * (1 << 16)));
  if (c->qt_exp > PREC_BITS) { c->qt_exp = PREC_BITS; c->qt_mant = 0; }
  if (c->qt_mant >= 1 << 16)
  {
    if (c->qt_exp == 0) { c->qt_exp = 0; c->qt_mant = (1 << 16) - 1; }
    else { c->qt_exp--; c->qt_mant = 0; }
  }
}

/* --- tag tree (tagtree.c:106-298) ------------------------------------- */
static tagtree *tt_create(uint64_t lx, uint64_t ly, uint64_t lz, uint64_t lt)
{
  uint32_t nx[32] = {0}, ny[32] = {0}, nz[32] = {0}, nt[32] = {0};
  uint64_t levels, nn;
  for (nx[0] = (uint32_t)lx, ny[0] = (uint32_t)ly, nz[0] = (uint32_t)lz, nt[0] = (uint32_t)lt, levels = 0, nn = 1;
       (uint64_t)nx[levels] * ny[levels] * nz[levels] * nt[levels] > 1; ++levels)
  {
    nn += (uint64_t)nx[levels] * ny[levels] * nz[levels] * nt[levels];
    nx[levels + 1] = (nx[levels] + 1) >> 1;
    ny[levels + 1] = (ny[levels] + 1) >> 1;
    nz[levels + 1] = (nz[levels] + 1) >> 1;
    nt[levels + 1] = (nt[levels] + 1) >> 1;
  }
  tagtree *t = calloc(1, sizeof *t);
  t->nleaves = lx * ly * lz * lt;
  t->nnodes = nn;
  t->nodes = calloc(nn, sizeof(tt_node));
  for (uint64_t i = 0; i < nn; ++i) t->nodes[i].parent = -1;
  uint64_t node = 0, level_base = 0;
  for (uint64_t n = 0; n < levels; ++n)
  {
    uint64_t cnt = (uint64_t)nx[n] * ny[n] * nz[n] * nt[n];
    uint64_t pbase = level_base + cnt;
    for (uint32_t l = 0; l < nt[n]; ++l)
      for (uint32_t k = 0; k < nz[n]; ++k)
        for (uint32_t j = 0; j < ny[n]; ++j)
          for (uint32_t i = 0; i < nx[n]; ++i)
          {
            uint64_t par = pbase + (uint64_t)(i >> 1) + (uint64_t)nx[n + 1] * ((j >> 1) + (uint64_t)ny[n + 1] * ((k >> 1) + (uint64_t)nz[n + 1] * (l >> 1)));
            t->nodes[node++].parent = (int64_t)par;
          }
    level_base = pbase;
  }
  for (uint64_t i = 0; i < nn; ++i) { t->nodes[i].value = 0xFFFF; t->nodes[i].threshold = 0; }
  return t;
}

static void tt_free(tagtree *t)
{
  if (t) { free(t->nodes); free(t); }
}

static void tt_set(tagtree *t, uint64_t leaf, uint16_t v)
{
  int64_t n = (int64_t)leaf;
  do
  {
    t->nodes[n].value = v;
    n = t->nodes[n].parent;
  } while (n >= 0 && t->nodes[n].value > v);
}

/* --- packet-header bit writer (bitstream.c:79-139, :277-363, :561-612) - */
typedef struct
{
  uint8_t *mem;
  uint64_t L, Lmax;
  uint8_t T;
  int8_t t;
} bitw;

static void bw_init(bitw *w, uint64_t cap)
{
  w->mem = calloc(cap, 1);
  w->L = 0;
  w->Lmax = cap;
  w->T = 0;
  w->t = 8;
}

static void bw_grow(bitw *w)
{
  if (w->L + 1 + (w->T == 0xFF) > w->Lmax)
  {
    uint64_t nl = w->Lmax * 2 + 16;
    w->mem = realloc(w->mem, nl);
    memset(w->mem + w->Lmax, 0, nl - w->Lmax);
    w->Lmax = nl;
  }
}

static void bw_bit(bitw *w, uint64_t bit)
{
  w->t--;
  if (bit) w->T |= (uint8_t)(1 << w->t);
  if (!w->t)
  {
    bw_grow(w);
    w->mem[w->L++] = w->T;
    w->t = (w->T == 0xFF) ? 7 : 8;
    w->T = 0;
  }
}

static void bw_flush(bitw *w)
{
  if (w->t != 8)
  {
    bw_grow(w);
    w->mem[w->L++] = w->T;
  }
}

/* tag-tree encoding, no state update when estimating (tagtree.c:310-373) */
static void tt_encode(tagtree *t, bitw *w, uint32_t threshold, uint64_t leaf, int est)
{
  int64_t branch[64];
  int nb = 0;
  int64_t n = (int64_t)leaf;
  while (t->nodes[n].parent >= 0)
  {
    branch[nb++] = n;
    n = t->nodes[n].parent;
  }
  uint32_t tmin = 0;
  for (;;)
  {
    tt_node *nd = &t->nodes[n];
    uint32_t th = nd->threshold;
    if (th < tmin) th = tmin;
    while ((nd->value >= th) && (th < threshold))
    {
      th++;
      bw_bit(w, (nd->value >= th) ? 0 : 1);
    }
    tmin = MN(nd->value, th);
    if (!est) nd->threshold = (uint16_t)th;
    if (n == (int64_t)leaf) break;
    n = branch[--nb];
  }
}

/* ====================================================================== */
/* geometry (libbwc.c:161-306, :360-607, :1345-1716, :1828-1982)           */
/* ====================================================================== */
static uint8_t popcount4(uint8_t hb)
{
  uint8_t t = hb;
  t -= (t >> 1) & 0x55;
  t = (t & 0x33) + ((t >> 2) & 0x33);
  return (t + (t >> 4)) & 0x0F;
}

static void init_precinct(oc_codec *c, prec_t *pr, uint32_t dX, uint32_t dY, uint32_t dZ, uint32_t dT)
{
  uint64_t cs[4] = {1ull << c->cbs[0], 1ull << c->cbs[1], 1ull << c->cbs[2], 1ull << c->cbs[3]};
  pr->ncb[0] = (uint16_t)((pr->x0 == pr->x1) ? 0 : (ceil((float)pr->x1 / cs[0]) - floor((float)pr->x0 / cs[0])));
  pr->ncb[1] = (uint16_t)((pr->y0 == pr->y1) ? 0 : (ceil((float)pr->y1 / cs[1]) - floor((float)pr->y0 / cs[1])));
  pr->ncb[2] = (uint16_t)((pr->z0 == pr->z1) ? 0 : (ceil((float)pr->z1 / cs[2]) - floor((float)pr->z0 / cs[2])));
  pr->ncb[3] = (uint16_t)((pr->t0 == pr->t1) ? 0 : (ceil((float)pr->t1 / cs[3]) - floor((float)pr->t0 / cs[3])));
  pr->ncbs = (uint64_t)pr->ncb[0] * pr->ncb[1] * pr->ncb[2] * pr->ncb[3];
  if (!pr->ncbs) return;
  pr->msbs = tt_create(pr->ncb[0], pr->ncb[1], pr->ncb[2], pr->ncb[3]);
  pr->incl = tt_create(pr->ncb[0], pr->ncb[1], pr->ncb[2], pr->ncb[3]);
  pr->cb = calloc(pr->ncbs, sizeof(cblk_t));
  uint64_t k = 0;
  for (uint16_t ct = 0; ct < pr->ncb[3]; ++ct)
    for (uint16_t cz = 0; cz < pr->ncb[2]; ++cz)
      for (uint16_t cy = 0; cy < pr->ncb[1]; ++cy)
        for (uint16_t cx = 0; cx < pr->ncb[0]; ++cx, ++k)
        {
          cblk_t *b = &pr->cb[k];
          b->beta = b->beta_est = 3;
          b->b.x0 = dX + (uint32_t)MX(pr->x0, cs[0] * (cx + (uint32_t)floor((float)pr->x0 / cs[0])));
          b->b.y0 = dY + (uint32_t)MX(pr->y0, cs[1] * (cy + (uint32_t)floor((float)pr->y0 / cs[1])));
          b->b.z0 = dZ + (uint32_t)MX(pr->z0, cs[2] * (cz + (uint32_t)floor((float)pr->z0 / cs[2])));
          b->b.t0 = dT + (uint32_t)MX(pr->t0, cs[3] * (ct + (uint32_t)floor((float)pr->t0 / cs[3])));
          b->b.x1 = dX + (uint32_t)MN(pr->x1, cs[0] * (cx + (uint32_t)floor((float)pr->x0 / cs[0]) + 1));
          b->b.y1 = dY + (uint32_t)MN(pr->y1, cs[1] * (cy + (uint32_t)floor((float)pr->y0 / cs[1]) + 1));
          b->b.z1 = dZ + (uint32_t)MN(pr->z1, cs[2] * (cz + (uint32_t)floor((float)pr->z0 / cs[2]) + 1));
          b->b.t1 = dT + (uint32_t)MN(pr->t1, cs[3] * (ct + (uint32_t)floor((float)pr->t0 / cs[3]) + 1));
        }
}

static void init_subband(oc_codec *c, param_t *par, res_t *res, subb_t *sb, int r, uint8_t hb)
{
  int dl = c->ndec - r + 1;
  int lv[4];
  for (int a = 0; a < 4; ++a) lv[a] = (dl > c->dec[a]) ? c->dec[a] : dl;
  sb->b.x0 = (uint64_t)ceil(((float)par->b.x0 / (1 << lv[0])) - 0.5 * ((hb & 1) >> 0));
  sb->b.y0 = (uint64_t)ceil(((float)par->b.y0 / (1 << lv[1])) - 0.5 * ((hb & 2) >> 1));
  sb->b.z0 = (uint64_t)ceil(((float)par->b.z0 / (1 << lv[2])) - 0.5 * ((hb & 4) >> 2));
  sb->b.t0 = (uint64_t)ceil(((float)par->b.t0 / (1 << lv[3])) - 0.5 * ((hb & 8) >> 3));
  sb->b.x1 = (uint64_t)ceil(((float)par->b.x1 / (1 << lv[0])) - 0.5 * ((hb & 1) >> 0));
  sb->b.y1 = (uint64_t)ceil(((float)par->b.y1 / (1 << lv[1])) - 0.5 * ((hb & 2) >> 1));
  sb->b.z1 = (uint64_t)ceil(((float)par->b.z1 / (1 << lv[2])) - 0.5 * ((hb & 4) >> 2));
  sb->b.t1 = (uint64_t)ceil(((float)par->b.t1 / (1 << lv[3])) - 0.5 * ((hb & 8) >> 3));

  uint32_t Rb = PREC_BITS + popcount4(hb);
  sb->hb = hb;
  sb->dwt_gain = energy_gain(c, hb, (uint16_t)dl);
  if (c->qstyle == 0)
  {
    sb->qt_exp = (uint8_t)Rb;
    sb->qt_mant = 0;
    sb->step = 1;
    sb->step_eff = 1;
  }
  else
  {
    sb->qt_exp = (uint8_t)((uint8_t)c->qt_exp + ((r == 0) ? 0 : 1 - r));
    sb->qt_mant = (uint16_t)c->qt_mant;
    sb->step = (1.0f + ((double)sb->qt_mant / (1 << 11))) / pow(2, sb->qt_exp);
    sb->step_eff = (1.0f + ((double)sb->qt_mant / (1 << 11)));
    if (sb->qt_exp > Rb) sb->step_eff /= pow(2, sb->qt_exp - Rb);
    else sb->step_eff *= pow(2, Rb - sb->qt_exp);
  }
  sb->Kmax = (uint16_t)MX(0, c->guard + sb->qt_exp - 1);

  uint64_t psz[4] = {1ull << c->ps[0], 1ull << c->ps[1], 1ull << c->ps[2], 1ull << c->ps[3]};
  uint64_t sbs[4] = {0, 0, 0, 0};
  if (hb & 1) sbs[0] = par->res[r - 1].b.x1 - par->res[r - 1].b.x0;
  if (hb & 2) sbs[1] = par->res[r - 1].b.y1 - par->res[r - 1].b.y0;
  if (hb & 4) sbs[2] = par->res[r - 1].b.z1 - par->res[r - 1].b.z0;
  if (hb & 8) sbs[3] = par->res[r - 1].b.t1 - par->res[r - 1].b.t0;
  if (!res->nprec) return;
  sb->prec = calloc(res->nprec, sizeof(prec_t));
  uint64_t p = 0;
  box_t *rb = &res->b;
  for (uint16_t pt = 0; pt < res->np[3]; ++pt)
    for (uint16_t pz = 0; pz < res->np[2]; ++pz)
      for (uint16_t py = 0; py < res->np[1]; ++py)
        for (uint16_t px = 0; px < res->np[0]; ++px, ++p)
        {
          prec_t *pr = &sb->prec[p];
          pr->x0 = (uint32_t)MX(rb->x0, psz[0] * (px + (uint32_t)floor(rb->x0 / psz[0])));
          pr->y0 = (uint32_t)MX(rb->y0, psz[1] * (py + (uint32_t)floor(rb->y0 / psz[1])));
          pr->z0 = (uint32_t)MX(rb->z0, psz[2] * (pz + (uint32_t)floor(rb->z0 / psz[2])));
          pr->t0 = (uint32_t)MX(rb->t0, psz[3] * (pt + (uint32_t)floor(rb->t0 / psz[3])));
          pr->x1 = (uint32_t)MN(rb->x1, psz[0] * (px + (uint32_t)floor(rb->x0 / psz[0]) + 1));
          pr->y1 = (uint32_t)MN(rb->y1, psz[1] * (py + (uint32_t)floor(rb->y0 / psz[1]) + 1));
          pr->z1 = (uint32_t)MN(rb->z1, psz[2] * (pz + (uint32_t)floor(rb->z0 / psz[2]) + 1));
          pr->t1 = (uint32_t)MN(rb->t1, psz[3] * (pt + (uint32_t)floor(rb->t0 / psz[3]) + 1));
          if ((c->ndec - c->dec[0]) < r)
          {
            pr->x0 = (uint32_t)ceil(((float)pr->x0 / 2) - 0.5 * ((hb & 1) >> 0));
            pr->x1 = (uint32_t)ceil(((float)pr->x1 / 2) - 0.5 * ((hb & 1) >> 0));
          }
          if ((c->ndec - c->dec[1]) < r)
          {
            pr->y0 = (uint32_t)ceil(((float)pr->y0 / 2) - 0.5 * ((hb & 2) >> 1));
            pr->y1 = (uint32_t)ceil(((float)pr->y1 / 2) - 0.5 * ((hb & 2) >> 1));
          }
          if ((c->ndec - c->dec[2]) < r)
          {
            pr->z0 = (uint32_t)ceil(((float)pr->z0 / 2) - 0.5 * ((hb & 4) >> 2));
            pr->z1 = (uint32_t)ceil(((float)pr->z1 / 2) - 0.5 * ((hb & 4) >> 2));
          }
          if ((c->ndec - c->dec[3]) < r)
          {
            pr->t0 = (uint32_t)ceil(((float)pr->t0 / 2) - 0.5 * ((hb & 8) >> 3));
            pr->t1 = (uint32_t)ceil(((float)pr->t1 / 2) - 0.5 * ((hb & 8) >> 3));
          }
          init_precinct(c, pr, (uint32_t)((uint32_t)sbs[0] - sb->b.x0), (uint32_t)((uint32_t)sbs[1] - sb->b.y0),
                        (uint32_t)((uint32_t)sbs[2] - sb->b.z0), (uint32_t)((uint32_t)sbs[3] - sb->b.t0));
          par->ncb += pr->ncbs;
        }
}

static int create_geometry(oc_codec *c)
{
  uint32_t ntx = (int)ceil(((float)c->n[0] / c->ts[0]));
  uint32_t nty = (int)ceil(((float)c->n[1] / c->ts[1]));
  uint32_t ntz = (int)ceil(((float)c->n[2] / c->ts[2]));
  uint32_t ntt = (int)ceil(((float)c->n[3] / c->ts[3]));
  /* the float resolution boxes below round tile bounds above 2^24; a finest
     box longer than its tile would take the transform past the samples
     (undefined in the reference), so such extents are refused */
  const uint32_t nt[4] = {ntx, nty, ntz, ntt};
  for (int a = 0; a < 4; ++a)
    for (uint32_t t = 0; t < nt[a]; ++t)
    {
      const uint64_t lo = (uint64_t)t * c->ts[a], hi = MN(c->n[a], (uint64_t)(t + 1) * c->ts[a]);
      if ((uint64_t)ceil((float)hi) - (uint64_t)ceil((float)lo) > hi - lo) return 1;
    }
  c->tile = calloc(c->ntiles, sizeof(tile_t));
  uint32_t i = 0;
  for (uint32_t t = 0; t < ntt; ++t)
    for (uint32_t z = 0; z < ntz; ++z)
      for (uint32_t y = 0; y < nty; ++y)
        for (uint32_t x = 0; x < ntx; ++x, ++i)
        {
          uint64_t npk = 0;
          tile_t *tl = &c->tile[i];
          tl->index = i;
          tl->b.x0 = (uint64_t)x * c->ts[0];
          tl->b.y0 = (uint64_t)y * c->ts[1];
          tl->b.z0 = (uint64_t)z * c->ts[2];
          tl->b.t0 = (uint64_t)t * c->ts[3];
          tl->b.x1 = MN(c->n[0], (x + 1) * c->ts[0]);
          tl->b.y1 = MN(c->n[1], (y + 1) * c->ts[1]);
          tl->b.z1 = MN(c->n[2], (z + 1) * c->ts[2]);
          tl->b.t1 = MN(c->n[3], (t + 1) * c->ts[3]);
          tl->header_size = 18;
          tl->slope_max = 0x0000;
          tl->slope_min = 0xFFFF;
          tl->par = calloc(c->nPar, sizeof(param_t));
          for (uint8_t j = 0; j < c->nPar; ++j)
          {
            param_t *par = &tl->par[j];
            par->pmax = -FLTMAX;
            par->pmin = FLTMAX;
            par->b = tl->b;
            par->res = calloc((size_t)c->ndec + 1, sizeof(res_t));
            uint64_t cblk = 0;
            for (int r = 0; r < c->ndec + 1; ++r)
            {
              res_t *res = &par->res[r];
              res->b.x0 = (uint64_t)ceil((float)par->b.x0 / (1 << MN(c->ndec - r, c->dec[0])));
              res->b.y0 = (uint64_t)ceil((float)par->b.y0 / (1 << MN(c->ndec - r, c->dec[1])));
              res->b.z0 = (uint64_t)ceil((float)par->b.z0 / (1 << MN(c->ndec - r, c->dec[2])));
              res->b.t0 = (uint64_t)ceil((float)par->b.t0 / (1 << MN(c->ndec - r, c->dec[3])));
              res->b.x1 = (uint64_t)ceil((float)par->b.x1 / (1 << MN(c->ndec - r, c->dec[0])));
              res->b.y1 = (uint64_t)ceil((float)par->b.y1 / (1 << MN(c->ndec - r, c->dec[1])));
              res->b.z1 = (uint64_t)ceil((float)par->b.z1 / (1 << MN(c->ndec - r, c->dec[2])));
              res->b.t1 = (uint64_t)ceil((float)par->b.t1 / (1 << MN(c->ndec - r, c->dec[3])));
              res->np[0] = (uint16_t)((res->b.x0 == res->b.x1) ? 0 : (ceil((float)res->b.x1 / (1 << c->ps[0]) - floor((float)res->b.x0 / (1 << c->ps[0])))));
              res->np[1] = (uint16_t)((res->b.y0 == res->b.y1) ? 0 : (ceil((float)res->b.y1 / (1 << c->ps[1]) - floor((float)res->b.y0 / (1 << c->ps[1])))));
              res->np[2] = (uint16_t)((res->b.z0 == res->b.z1) ? 0 : (ceil((float)res->b.z1 / (1 << c->ps[2]) - floor((float)res->b.z0 / (1 << c->ps[2])))));
              res->np[3] = (uint16_t)((res->b.t0 == res->b.t1) ? 0 : (ceil((float)res->b.t1 / (1 << c->ps[3]) - floor((float)res->b.t0 / (1 << c->ps[3])))));
              res->nprec = (uint64_t)res->np[0] * res->np[1] * res->np[2] * res->np[3];
              npk += res->nprec * c->nlayers;
              uint8_t hb = 0;
              for (int a = 0; a < 4; ++a)
                if (c->ndec - c->dec[a] < r && r > 0) hb |= (uint8_t)(1 << a);
              res->nsb = (uint8_t)(1 << ((hb & 1) + ((hb & 2) >> 1) + ((hb & 4) >> 2) + ((hb & 8) >> 3)));
              res->nsb -= (res->nsb & 1) ? 0 : 1;
              res->pk = calloc(res->nprec * c->nlayers + 1, sizeof(packet_t));
              res->sb = calloc(res->nsb, sizeof(subb_t));
              for (int l = 0, m = 0; m < res->nsb; ++l)
              {
                if (r == 0 || ((l & hb) == l && l != 0))
                {
                  init_subband(c, par, res, &res->sb[m], r, (uint8_t)l);
                  par->acc = realloc(par->acc, (par->ncb ? par->ncb : 1) * sizeof(access_t));
                  for (uint64_t p = 0; p < res->nprec; ++p)
                    for (uint64_t k = 0; k < res->sb[m].prec[p].ncbs; ++k, ++cblk)
                    {
                      par->acc[cblk].sb = &res->sb[m];
                      par->acc[cblk].pr = &res->sb[m].prec[p];
                      par->acc[cblk].cb = &res->sb[m].prec[p].cb[k];
                      par->acc[cblk].cb->idx = (uint32_t)k;
                      par->acc[cblk].cb->aidx = cblk;
                    }
                  m++;
                }
              }
            }
          }
          tl->seq = calloc(npk + 1, sizeof(packet_t *));
          tl->header_size += c->erres ? 8 * npk : 0;
          tl->npk = npk;
        }
  return 0;
}

/* configure_codec + setters + bwc_create_compression (libbwc.c:1828-3131) */
oc_codec *oc_create(const oc_params *p)
{
  oc_codec *c = calloc(1, sizeof *c);
  memset(g_gain_lut, 0, sizeof g_gain_lut);
  for (int a = 0; a < 4; ++a) c->n[a] = p->n[a];
  c->nPar = p->nPar;
  c->data_prec = p->data_prec;
  for (int a = 0; a < 4; ++a)
  {
    int valid = (c->n[a] >> 1) != 0;
    c->ts[a] = c->n[a];
    c->ps[a] = valid ? 15 : 0;
    c->cbs[a] = valid ? 5 : 0;
    c->kern[a] = 0;
  }
  c->ntiles = 1;
  c->prog = 0;
  c->qstyle = 1;
  c->nlayers = 0;
  c->Qm = PREC_BITS - 1;
  c->guard = 2;
  uint8_t lv[4];
  for (int a = 0; a < 4; ++a)
  {
    lv[a] = (uint8_t)(log((double)c->n[a]) / log(2));
    c->dec[a] = (c->n[a] >> 1) ? ((4 < lv[a]) ? 4 : lv[a]) : 0;
  }
  c->ndec = MX(c->dec[0], MX(c->dec[1], MX(c->dec[2], c->dec[3])));
  gain_lut_accumulate();
  memcpy(c->lut, g_gain_lut, sizeof c->lut);
  {
    double delta = (double)(1 / (pow(2, 2 + PREC_BITS) * sqrt(energy_gain(c, 0, c->ndec))));
    set_qt(c, delta);
  }

  /* header-carried settings (codestream.c:710-790): quantisation style,
     step size, progression, kernels; the decomposition setter below
     recomputes the step size again */
  if (p->have_qstyle) c->qstyle = p->qstyle, c->CSsgc |= (0x01 << 1);
  if (p->have_qt) c->qt_exp = p->qt_exp, c->qt_mant = p->qt_mant;
  if (p->have_prog) c->prog = p->prog, c->CSsgc |= (0x01 << 3);
  if (p->have_kern)
  {
    c->kern[0] = p->kern[0] ? p->kern[0] : 0;
    c->kern[1] = p->kern[1] ? p->kern[1] : 0;
    c->kern[2] = p->kern[2] ? p->kern[2] : 0;
    c->kern[3] = p->kern[3] ? p->kern[3] : 2;   /* set_kernels: Haar when 0 for the temporal axis */
    c->CSsgc |= (0x01 << 4);
  }
  if (p->have_decomp)
  {
    uint8_t d[4];
    int bad;
    for (int a = 0; a < 4; ++a) d[a] = (c->n[a] >> 1) ? ((p->decomp[a] < lv[a]) ? p->decomp[a] : lv[a]) : 0;
    bad = (d[0] > 63) || (d[1] > 63) || (d[2] > 63) || (d[3] > 31);
    if (bad)
    {
      for (int a = 0; a < 4; ++a) c->dec[a] = (c->n[a] >> 1) ? 4 : 0;
      c->ndec = MX(d[0], MX(d[1], MX(d[2], d[3])));
    }
    else
    {
      for (int a = 0; a < 4; ++a) c->dec[a] = d[a];
      c->ndec = MX(c->dec[0], MX(c->dec[1], MX(c->dec[2], c->dec[3])));
      gain_lut_accumulate();
      memcpy(c->lut, g_gain_lut, sizeof c->lut);
      double delta = 1 / (pow(2, 2 + PREC_BITS) * sqrt(energy_gain(c, 0, c->ndec)));
      set_qt(c, delta);
      c->CSsgc |= (0x01 << 5);
    }
  }
  if (p->have_prec)
  {
    uint8_t ps[4];
    for (int a = 0; a < 4; ++a) ps[a] = (c->n[a] >> 1) ? (p->prec[a] ? p->prec[a] : 15) : 0;
    int bad = 0;
    for (int a = 0; a < 4; ++a) bad |= ((ps[a] < 1) && (c->n[a] >> 1)) || (ps[a] > 15);
    for (int a = 0; a < 4; ++a) c->ps[a] = bad ? ((c->n[a] >> 1) ? 15 : 0) : ps[a];
    if (!bad) c->CSsgc |= (0x01 << 6);
  }
  if (p->have_cb)
  {
    uint8_t cb[4];
    for (int a = 0; a < 4; ++a) cb[a] = (c->n[a] >> 1) ? (p->cb[a] ? p->cb[a] : 5) : 0;
    int s = cb[0] + cb[1] + cb[2] + cb[3];
    int bad = (cb[0] > 10) || (cb[1] > 10) || (cb[2] > 10) || (cb[3] > 10) || s < 4 || s > 20;
    for (int a = 0; a < 4; ++a) c->cbs[a] = bad ? ((c->n[a] >> 1) ? 5 : 0) : cb[a];
    if (!bad) c->CSsgc |= (0x01 << 7);
  }
  if (p->have_qm)
  {
    if ((int8_t)(PREC_BITS - p->qm) >= 1)
    {
      c->Qm = p->qm;
      c->CSsgc |= (0x01 << 8);
    }
  }
  if (p->have_tiles)
  {
    /* bwc_set_tiles(bwc_tile_sizeof) (libbwc.c:2678-2748); the validity test
       reads the sizes before they are replaced, exactly as written there */
    int bad = 0;
    for (int a = 0; a < 4; ++a) bad |= (c->ts[a] < 16 && c->ts[a] > c->n[a]);
    if (!bad)
    {
      for (int a = 0; a < 4; ++a) c->ts[a] = (c->n[a] >> 1) ? (p->tiles[a] ? p->tiles[a] : c->n[a]) : c->n[a];
      uint64_t nt = 1;
      for (int a = 0; a < 4; ++a) nt *= (uint64_t)ceil(((float)c->n[a] / c->ts[a]));
      c->ntiles = nt;
      c->CSsgc |= (0x01 << 9);
    }
  }
  if (p->erres)
  {
    c->erres ^= 0x01;
    c->CSsgc ^= (0x01 << 0);
  }
  /* set_quant_step_size (libbwc.c:2134-2184): 0 < delta < 2, else a warning
     and no change */
  if (p->have_qstep && p->qstep > 0 && p->qstep < 2)
  {
    set_qt(c, p->qstep);
    c->CSsgc |= (0x01 << 2);
  }
  if (p->naux)
  {
    c->aux = malloc(p->naux);
    memcpy(c->aux, p->aux, p->naux);
    c->naux = p->naux;
  }
  if (p->ncom)
  {
    c->com = malloc(p->ncom);
    memcpy(c->com, p->com, p->ncom);
    c->ncom = p->ncom;
  }

  /* rate-control string (libbwc.c:3024-3089), or the layer rates as read
     from a codestream header */
  if (p->nrates)
  {
    for (int l = 0; l < p->nrates; ++l) c->bitrate[l] = p->rates[l];
    c->nlayers = p->nrates;
  }
  else if (p->rate == NULL)
  {
    c->bitrate[0] = PREC_BITS + 1;
    c->nlayers = 1;
  }
  else
  {
    char *rc = strdup(p->rate), *save = NULL;
    char *tok = strtok_r(rc, "-, ", &save);
    if (tok)
    {
      for (c->nlayers = 0; tok; tok = strtok_r(NULL, "-, ", &save))
      {
        float f = strtof(tok, NULL);
        if (f <= 0 || f > 64) { free(rc); oc_destroy(c); return NULL; }
        for (int l = c->nlayers; l >= 0; --l)
        {
          if (l && f < c->bitrate[l - 1]) c->bitrate[l] = c->bitrate[l - 1];
          else if (l && fabs(f - c->bitrate[l - 1]) < 0.01f)
          {
            for (; l < c->nlayers; ++l) c->bitrate[l] = c->bitrate[l + 1];
            break;
          }
          else
          {
            c->bitrate[l] = f;
            c->nlayers++;
            break;
          }
        }
      }
    }
    else
    {
      c->bitrate[0] = PREC_BITS + 1;
      c->nlayers = 1;
    }
    free(rc);
  }
  if (create_geometry(c))
  {
    oc_destroy(c);
    return NULL;
  }
  /* headerSize (libbwc.c:3096-3107): 6 bytes plus the payload for each of
     AUX and COM, although a COM segment takes 4 bytes plus its payload */
  c->header_size = 111 + (uint64_t)c->nPar * c->ntiles * 2 * PREC_BYTES + (uint64_t)c->nlayers * 4;
  if (c->naux) c->header_size += 6 + (uint64_t)c->naux;
  if (c->ncom) c->header_size += 6 + (uint64_t)c->ncom;
  return c;
}

void oc_destroy(oc_codec *c)
{
  if (!c) return;
  free(c->aux);
  free(c->com);
  for (uint64_t i = 0; c->tile && i < c->ntiles; ++i)
  {
    tile_t *tl = &c->tile[i];
    for (uint8_t j = 0; tl->par && j < c->nPar; ++j)
    {
      param_t *par = &tl->par[j];
      for (int r = 0; par->res && r < c->ndec + 1; ++r)
      {
        res_t *res = &par->res[r];
        for (uint8_t m = 0; res->sb && m < res->nsb; ++m)
        {
          subb_t *sb = &res->sb[m];
          for (uint64_t p = 0; sb->prec && p < res->nprec; ++p)
          {
            prec_t *pr = &sb->prec[p];
            for (uint64_t k = 0; k < pr->ncbs; ++k) free(pr->cb[k].data);
            free(pr->cb);
            tt_free(pr->incl);
            tt_free(pr->msbs);
          }
          free(sb->prec);
        }
        for (uint64_t k = 0; res->pk && k < res->nprec * c->nlayers; ++k)
        {
          free(res->pk[k].hdr);
          free(res->pk[k].body);
        }
        free(res->pk);
        free(res->sb);
      }
      free(par->res);
      free(par->acc);
    }
    free(tl->par);
    free(tl->seq);
  }
  free(c->tile);
  free(c);
}

uint64_t oc_ntiles(const oc_codec *c) { return c->ntiles; }
uint8_t oc_nlayers(const oc_codec *c) { return c->nlayers; }

void oc_param_dims(const oc_codec *c, uint64_t tile, uint64_t d[4])
{
  const box_t *b = &c->tile[tile].par[0].b;
  d[0] = b->x1 - b->x0;
  d[1] = b->y1 - b->y0;
  d[2] = b->z1 - b->z0;
  d[3] = b->t1 - b->t0;
}

uint64_t oc_param_size(const oc_codec *c, uint64_t tile)
{
  uint64_t d[4];
  oc_param_dims(c, tile, d);
  return d[0] * d[1] * d[2] * d[3];
}

uint64_t oc_ncblk(const oc_codec *c, uint64_t tile, uint32_t param) { return c->tile[tile].par[param].ncb; }

void oc_cblk_geometry(const oc_codec *c, uint64_t tile, uint32_t param, uint64_t cb, uint64_t box[8], double sb[3],
                      uint32_t meta[4])
{
  const access_t *a = &c->tile[tile].par[param].acc[cb];
  const box_t *b = &a->cb->b;
  box[0] = b->x0; box[1] = b->y0; box[2] = b->z0; box[3] = b->t0;
  box[4] = b->x1; box[5] = b->y1; box[6] = b->z1; box[7] = b->t1;
  sb[0] = a->sb->dwt_gain;
  sb[1] = a->sb->step;
  sb[2] = a->sb->step_eff;
  meta[0] = a->sb->hb;
  meta[1] = a->sb->Kmax;
  meta[2] = a->sb->qt_exp;
  meta[3] = a->cb->idx;
}

/* ====================================================================== */
/* fill + normalize (libbwc.c:622-812, :993-1078)                          */
/* ====================================================================== */
void oc_fill_normalize(oc_codec *c, uint64_t tile, uint32_t param, const void *input, double *work, double mm[4])
{
  tile_t *tl = &c->tile[tile];
  param_t *par = &tl->par[param];
  uint64_t w = par->b.x1 - par->b.x0, h = par->b.y1 - par->b.y0, d = par->b.z1 - par->b.z0, dt = par->b.t1 - par->b.t0;
  uint64_t poff = c->n[0] * c->n[1] * c->n[2] * c->n[3] * param;
  double mx = par->pmax, mn = par->pmin;
  for (uint64_t t = 0; t < dt; ++t)
    for (uint64_t z = 0; z < d; ++z)
      for (uint64_t y = 0; y < h; ++y)
      {
        uint64_t src = poff + tl->b.x0 + c->n[0] * (tl->b.y0 + y + c->n[1] * (tl->b.z0 + z + c->n[2] * t));
        double *dst = &work[w * (y + h * (z + d * t))];
        for (uint64_t x = 0; x < w; ++x)
        {
          dst[x] = (c->data_prec == 8) ? ((const double *)input)[src + x] : (double)((const float *)input)[src + x];
          mx = MX(mx, dst[x]);
          mn = MN(mn, dst[x]);
        }
      }
  par->pmax = mx;
  par->pmin = mn;
  par->alpha = (mx + mn) / 2.0f;
  par->beta = 2.0f / (mx - mn);
  uint64_t shift = (uint64_t)c->Qm << PREC_MANT;
  uint64_t n = w * h * d * dt;
  for (uint64_t x = 0; x < n; ++x)
  {
    double v = (double)(work[x] - par->alpha) * par->beta;
    uint64_t raw;
    memcpy(&raw, &v, 8);
    raw += (((raw & EXP_MASK) == 0) && ((raw & MANT_MASK) == 0)) ? 0 : shift;
    memcpy(&work[x], &raw, 8);
  }
  mm[0] = mn;
  mm[1] = mx;
  mm[2] = par->alpha;
  mm[3] = par->beta;
}

/* ====================================================================== */
/* DWT (dwt.c:133-789, :1216-2246)                                          */
/* ====================================================================== */
/* whole-point symmetric extension of `len` samples at buf[0..len) by nt on
   both sides, including the repeated reflection for short lines (dwt.c:286) */
static void wps_extend(double *buf, int64_t len, int nt)
{
  if (nt < len)
  {
    for (int i = 0; i < nt; ++i)
    {
      buf[i - nt] = buf[nt - i];
      buf[len + i] = buf[len - (i + 2)];
    }
  }
  else
  {
    int64_t mn = nt, mx = nt + len - 1;
    for (int64_t i = 0; i < nt; ++i)
    {
      int64_t j = i;
      while (j < mn || j > mx) j = (j < mn) ? (mn << 1) - j : (mx << 1) - j;
      buf[i - nt] = buf[j - nt];
      j = i + len + nt;
      while (j < mn || j > mx) j = (j < mn) ? (mn << 1) - j : (mx << 1) - j;
      buf[i + len] = buf[j - nt];
    }
  }
}

static void fwd97(double *w, int64_t len)
{
  for (int64_t i = -3; i < len + 3; i += 2) w[i] += K_ALPHA * (w[i - 1] + w[i + 1]);
  for (int64_t i = -2; i < len + 2; i += 2) w[i] += K_BETA * (w[i - 1] + w[i + 1]);
  for (int64_t i = -1; i < len + 1; i += 2)
  {
    w[i] += K_GAMMA * (w[i - 1] + w[i + 1]);
    w[i] = K_KAPPA_H * w[i];
  }
  for (int64_t i = 0; i < len; i += 2)
  {
    w[i] += K_DELTA * (w[i - 1] + w[i + 1]);
    w[i] = K_KAPPA_L * w[i];
  }
}

static void inv97(double *w, int64_t len)
{
  for (int64_t i = -4; i < len + 4; i += 2)
  {
    w[i] = K_KAPPA_H * w[i];
    w[i] -= K_DELTA * (w[i - 1] + w[i + 1]);
  }
  for (int64_t i = -5; i < len + 3; i += 2)
  {
    w[i] = K_KAPPA_L * w[i];
    w[i] -= K_GAMMA * (w[i - 1] + w[i + 1]);
  }
  for (int64_t i = -4; i < len + 2; i += 2) w[i] -= K_BETA * (w[i - 1] + w[i + 1]);
  for (int64_t i = -3; i < len + 1; i += 2) w[i] -= K_ALPHA * (w[i - 1] + w[i + 1]);
}

static void fwd53(double *w, int64_t len)
{
  for (int64_t i = -1; i < len + 1; i += 2) w[i] -= (w[i - 1] + w[i + 1]) / 2;
  for (int64_t i = 0; i < len; i += 2) w[i] += (w[i - 1] + w[i + 1] + 2) / 4;
}

static void inv53(double *w, int64_t len)
{
  for (int64_t i = -2; i < len + 1; i += 2) w[i] -= (w[i - 1] + w[i + 1] + 2) / 4;
  for (int64_t i = -1; i < len; i += 2) w[i] += (w[i - 1] + w[i + 1]) / 2;
}

/* one 1-D forward transform of a strided line of length r1-r0 (dwt.c:1378-1457) */
static void line_fwd(double *data, uint64_t r0, uint64_t r1, uint64_t inc, int kern, double *mem, int nt)
{
  int64_t len = (int64_t)(r1 - r0);
  double *base = mem + nt;
  for (int64_t i = 0; i < len; ++i) base[i] = data[(uint64_t)i * inc];
  wps_extend(base, len, nt);
  double *w = base + ((r0 % 2) ? 1 : 0);
  if (kern == 0) fwd97(w, len);
  else fwd53(w, len);
  /* buffer_flush_forward (dwt.c:645-721) */
  int64_t half = (r0 % 2) ? len >> 1 : (len + 1) >> 1;
  double *lo = data, *hi = data + (uint64_t)half * inc;
  int64_t i;
  if (r0 % 2) { *hi = w[-1]; hi += inc; }
  for (i = 0; i < half - 1; ++i)
  {
    *lo = w[i << 1];
    *hi = w[(i << 1) + 1];
    lo += inc;
    hi += inc;
  }
  if (r1 % 2) *lo = w[i << 1];
  else { *lo = w[i << 1]; *hi = w[(i << 1) + 1]; }
}

/* one 1-D inverse transform (dwt.c:169-272, :735-789, :1900-2232) */
static void line_inv(double *data, uint64_t r0, uint64_t r1, uint64_t inc, int kern, double *mem, int nt)
{
  int64_t len = (int64_t)(r1 - r0);
  double *base = mem + nt;
  int64_t half = (r0 % 2) ? len >> 1 : (len + 1) >> 1;
  double *lo = data, *hi = data + (uint64_t)half * inc;
  if ((r0 % 2) == 0)
  {
    for (int64_t i = 0; i < half; ++i) { base[i << 1] = *lo; lo += inc; }
    if (len % 2) --half;
    for (int64_t i = 0; i < half; ++i) { base[(i << 1) + 1] = *hi; hi += inc; }
  }
  else
  {
    for (int64_t i = 0; i < half; ++i) { base[i << 1] = *lo; lo += inc; }
    if (len % 2) ++half;
    for (int64_t i = 0; i < half; ++i) { base[(i << 1) + 1] = *hi; hi += inc; }
  }
  wps_extend(base, len, nt);
  double *w = base + ((r0 % 2) ? 1 : 0);
  if (kern == 0) inv97(w, len);
  else inv53(w, len);
  double *m = data;
  int64_t n = len;
  if (r0 % 2) { *m = w[-1]; m += inc; n--; }
  for (int64_t i = 0, j = 0; i < n; ++i, j += (int64_t)inc) m[j] = w[i];
}

static int kern_taps(int k) { return k == 0 ? 4 : (k == 1 ? 2 : 1); }

static void dwt_run(oc_codec *c, uint64_t tile, uint32_t param, double *data, int inverse)
{
  param_t *par = &c->tile[tile].par[param];
  uint64_t w = par->b.x1 - par->b.x0, h = par->b.y1 - par->b.y0, d = par->b.z1 - par->b.z0, dt = par->b.t1 - par->b.t0;
  uint64_t inc[4] = {1, w, w * h, w * h * d};
  uint64_t mx = MX(MX(w, h), MX(d, dt));
  int nt[4];
  for (int a = 0; a < 4; ++a) nt[a] = kern_taps(c->kern[a]) + 2;
  int ntmax = MX(MX(nt[0], nt[1]), MX(nt[2], nt[3]));
  double *mem = calloc(mx + 2 * (uint64_t)ntmax + 2, sizeof(double));
  int order_f[4] = {0, 1, 2, 3}, order_i[4] = {3, 2, 1, 0};
  int *order = inverse ? order_i : order_f;
  for (int lv = 0; lv < c->ndec; ++lv)
  {
    int level = inverse ? (c->ndec - 1 - lv) : lv;
    box_t *rb = &par->res[c->ndec - level].b;
    uint64_t r0[4] = {rb->x0, rb->y0, rb->z0, rb->t0}, r1[4] = {rb->x1, rb->y1, rb->z1, rb->t1};
    uint64_t sz[4] = {r1[0] - r0[0], r1[1] - r0[1], r1[2] - r0[2], r1[3] - r0[3]};
    for (int oi = 0; oi < 4; ++oi)
    {
      int a = order[oi];
      if (!(level < c->dec[a] && sz[a] > 1)) continue;
      if (c->kern[a] == 2) continue; /* Haar: no boundary handler in the reference (dwt.c:1259-1262) */
      int o[3], k = 0;
      for (int b = 0; b < 4; ++b)
        if (b != a) o[k++] = b;
      for (uint64_t i2 = 0; i2 < sz[o[2]]; ++i2)
        for (uint64_t i1 = 0; i1 < sz[o[1]]; ++i1)
          for (uint64_t i0 = 0; i0 < sz[o[0]]; ++i0)
          {
            double *line = data + i0 * inc[o[0]] + i1 * inc[o[1]] + i2 * inc[o[2]];
            if (inverse) line_inv(line, r0[a], r1[a], inc[a], c->kern[a], mem, nt[a]);
            else line_fwd(line, r0[a], r1[a], inc[a], c->kern[a], mem, nt[a]);
          }
    }
  }
  free(mem);
}

void oc_dwt_forward(oc_codec *c, uint64_t tile, uint32_t param, double *work) { dwt_run(c, tile, param, work, 0); }
void oc_dwt_inverse(oc_codec *c, uint64_t tile, uint32_t param, double *work) { dwt_run(c, tile, param, work, 1); }

#include "bwc_oracle_t1.inc"
#include "bwc_oracle_t2.inc"
