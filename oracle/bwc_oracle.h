/*
 * bwc_oracle — CPU restatement of BigWhoop's compression path (test
 * infrastructure; the checker, never the product).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library. Every function restates the reference algorithm from its
 * source (file:line citations in bwc_oracle.c) in plain scalar C.
 *
 * Pinning status (what fixes each function independently of this file):
 *   geometry, gain LUT, fill/normalize, DWT, tier-1 + MQ + truncation
 *   lengths + hull/slopes, tier-2 PCRD, packets, tag-trees, codestream:
 *     byte-identical codestreams and per-codeblock records against the
 *     reference built from source (oracle/_ref, tests/golden fixtures),
 *     plus closed-form checks (DWT DC gain and vanishing moments, MQ
 *     round trip through the decoder, tag-tree textbook cases).
 *   decompression (oc_decompress: header and packet parsing, tier-1
 *   decoding, reconstruction, inverse DWT, denormalisation): the output
 *     bytes of the reference's bwc_decompress on every golden codestream
 *     (SHA-256 and samples in tests/golden), and decoded quality growing
 *     with the rate.
 */
#ifndef BWC_ORACLE_H
#define BWC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oc_params
{
  uint64_t n[4];       /* nX, nY, nZ, nTS */
  uint8_t  nPar;
  uint8_t  data_prec;  /* 4 or 8 bytes per input value */
  uint8_t  have_decomp, decomp[4];
  uint8_t  have_prec, prec[4];
  uint8_t  have_cb, cb[4];
  uint8_t  have_qm, qm;
  uint8_t  have_tiles;
  uint64_t tiles[4];   /* tile sizes (bwc_tile_sizeof) */
  uint8_t  erres;
  const char *rate;    /* e.g. "2" or "1,2,4,8"; NULL = lossless-ish default */
  /* settings a codestream header can carry beyond the above (decoder side):
     applied in the order of codestream.c:710-790, before the decomposition */
  uint8_t  have_qstyle, qstyle;
  uint8_t  have_qt, qt_exp;
  uint16_t qt_mant;
  uint8_t  have_prog, prog;
  uint8_t  have_kern, kern[4];
  uint8_t  nrates;     /* > 0: these layer rates as read (no parsing of rate) */
  const float *rates;
  /* encoder-side setters applied after error resilience, in this order:
     set_quant_step_size (libbwc.c:2120-2184), then the stream's AUX and COM
     segments (bwc_set_aux / bwc_set_com, libbwc.c:1264-1340) */
  uint8_t  have_qstep;
  double   qstep;
  const uint8_t *aux;
  uint32_t naux;
  const uint8_t *com;
  uint16_t ncom;
} oc_params;

typedef struct oc_codec oc_codec;

/* configure + create geometry (bwc_alloc_coder, setters, bwc_create_compression) */
oc_codec *oc_create(const oc_params *p);
void      oc_destroy(oc_codec *c);

/* whole compression: returns codestream bytes written to out (0 on error) */
size_t    oc_compress(oc_codec *c, const void *input, uint8_t *out, size_t cap);

/* --- stage access for parity tests ----------------------------------- */
uint64_t  oc_ntiles(const oc_codec *c);
uint64_t  oc_param_size(const oc_codec *c, uint64_t tile);            /* values per (tile,param) */
void      oc_param_dims(const oc_codec *c, uint64_t tile, uint64_t dims[4]);
uint64_t  oc_ncblk(const oc_codec *c, uint64_t tile, uint32_t param);
/* per codeblock (access order): box[8] (x0 y0 z0 t0 x1 y1 z1 t1, tile-local),
   f64 {dwt_gain, qt_step_size, qt_effective_step_size},
   u32 {highband_flag, Kmax, qt_exponent, index-in-precinct} */
void      oc_cblk_geometry(const oc_codec *c, uint64_t tile, uint32_t param, uint64_t cb,
                           uint64_t box[8], double sb[3], uint32_t meta[4]);
/* fill + normalize one (tile,param) into work (tile-local, x fastest);
   returns min/max/alpha/beta in mm[4] */
void      oc_fill_normalize(oc_codec *c, uint64_t tile, uint32_t param, const void *input,
                            double *work, double mm[4]);
void      oc_dwt_forward(oc_codec *c, uint64_t tile, uint32_t param, double *work);
void      oc_dwt_inverse(oc_codec *c, uint64_t tile, uint32_t param, double *work);
/* tier-1 over all codeblocks of (tile,param); work is modified (|c|) */
void      oc_t1_encode(oc_codec *c, uint64_t tile, uint32_t param, double *work);
/* record of one encoded codeblock; returns number of data bytes */
uint64_t  oc_cblk_record(const oc_codec *c, uint64_t tile, uint32_t param, uint64_t cb,
                         uint8_t kkz[3], uint16_t *S, uint64_t *L, const uint8_t **data);
void      oc_tile_slopes(const oc_codec *c, uint64_t tile, uint16_t mm[2]);
/* tier-2 rate control + packets for a tile; cuts[cb*nLayers+l] = cp_contr after PCRD */
int       oc_t2_encode(oc_codec *c, uint64_t tile, int16_t *cuts_or_null);
/* (layer, threshold, estimate) of every probe of the last tile's PCRD
 * bisection, at most cap; returns the count */
int       oc_rc_trace(uint32_t *layer, uint32_t *slope, int64_t *est, int cap);
size_t    oc_assemble(oc_codec *c, uint8_t *out, size_t cap);
uint8_t   oc_nlayers(const oc_codec *c);

/* --- standalone kernels ------------------------------------------------ */
/* encode one codeblock taken from a (width,height,depth) float64 buffer */
int       oc_t1_encode_block(const double *buf, uint64_t width, uint64_t height, uint64_t depth,
                             const uint64_t box[8], double step_eff, double step, double gain,
                             uint8_t highband, uint16_t Kmax, uint8_t erres,
                             uint8_t kkz[3], uint16_t *S, uint64_t *L, uint8_t *data, uint64_t cap);
/* one codeblock coded to a pass limit (zlim > 0; 0 = every pass), for the
   pass-skip tests: records as oc_t1_encode_block, plus the integer
   distortion sum of each pass (dm), the read extent of each pass's
   truncation-length search (reach: one past the last byte index read), the
   bytes emitted before termination (nfinal) and the histogram of the
   quantised magnitudes' bit lengths (hist[65]) */
int       oc_t1_probe_block(const double *buf, uint64_t width, uint64_t height, uint64_t depth,
                            const uint64_t box[8], double step_eff, double step, double gain,
                            uint8_t highband, uint16_t Kmax, uint8_t erres, int zlim,
                            uint8_t kkz[3], uint16_t *S, uint64_t *L, int64_t *dm, int64_t *reach,
                            int64_t *nfinal, uint32_t hist[65]);
/* MQ round trip helper: encode n (bit,ctx) symbols; returns bytes */
uint64_t  oc_mq_encode_symbols(const uint8_t *bits, const uint8_t *ctx, uint64_t n,
                               uint8_t *out, uint64_t cap);
void      oc_mq_decode_symbols(const uint8_t *in, uint64_t len, const uint8_t *ctx, uint64_t n,
                               uint8_t *bits_out);
/* gain LUT after n accumulating initialisations (3 kernels x 10 entries) */
void      oc_gain_lut(int ncalls, double lut[30]);
/* slope2log of one lambda */
uint16_t  oc_slope2log(double lambda);

/* --- decompression (restated bwc_decompress) -------------------------- */
/* decode a codestream into out (nX*nY*nZ*nTS*nPar values of data_prec);
   layer as in bwc_create_decompression; returns 0 on success */
int       oc_decompress(const uint8_t *cs, size_t len, uint8_t layer, void *out, uint64_t dims_out[6]);

#ifdef __cplusplus
}
#endif
#endif
