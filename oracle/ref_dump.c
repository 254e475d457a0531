/*
 * ref_dump — test-infrastructure driver around the reference BigWhoop library.
 *
 * This file is project code. It is compiled against the reference's private
 * headers and objects (include/library/private, src/library/*.c), which are
 * read in place under the reference checkout by oracle/Makefile; nothing of
 * the reference is copied here. The binary lands in oracle/_ref/ and is only
 * ever run by tests/golden/make_golden.py and the oracle pinning tests.
 *
 * One compression per process (the reference's energy-gain LUT accumulates in
 * a process-global, dwt.c:80 / dwt.c:864-1045), with the fixed setter order
 * alloc -> kernels -> progression -> quant style -> decomp -> precincts ->
 * codeblocks -> qm -> tiles -> erres -> quant step size, then COM / AUX on
 * the stream before bwc_create_compression (the header size counts them).
 *
 * Modes
 *   full   : bwc_create_compression + bwc_compress (libbwc.c:2970, :3145);
 *            writes <prefix>.bwc
 *   stages : the same stage loop as bwc_compress (libbwc.c:3231-3383) driven
 *            from here so intermediate results can be dumped:
 *              <prefix>.dwt   float64 coefficients per (tile, param) after the
 *                             forward DWT (dwt.c:1216)
 *              <prefix>.t1    per-codeblock tier-1 records (tier1.c:2824)
 *              <prefix>.geo   per-codeblock geometry and subband constants
 *              <prefix>.bwc   the assembled codestream (codestream.c:1329)
 *            fill_buffer / normalize_param are file-static in libbwc.c, so
 *            their two loops (libbwc.c:716-811, :1064-1077) are restated
 *            below; the stages codestream is compared byte-for-byte with the
 *            full-mode one by the golden script, which pins that restatement.
 *   decode : bwc_create_decompression + bwc_decompress on <prefix>.bwc,
 *            writes <prefix>.dec (float64 or float32, like the input)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

#include "libbwc.h"
#include "dwt.h"
#include "tier1.h"
#include "tier2.h"

static void die(const char *m) { fprintf(stderr, "ref_dump: %s\n", m); exit(2); }

static void *read_file(const char *path, size_t *n)
{
  FILE *f = fopen(path, "rb");
  if (!f) die("cannot open input");
  fseek(f, 0, SEEK_END);
  *n = (size_t)ftell(f);
  fseek(f, 0, SEEK_SET);
  void *p = malloc(*n ? *n : 1);
  if (*n && fread(p, 1, *n, f) != *n) die("short read");
  fclose(f);
  return p;
}

static void write_file(const char *prefix, const char *ext, const void *p, size_t n)
{
  char path[4096];
  snprintf(path, sizeof path, "%s%s", prefix, ext);
  FILE *f = fopen(path, "wb");
  if (!f) die("cannot open output");
  if (n && fwrite(p, 1, n, f) != n) die("short write");
  fclose(f);
}

static int parse4(const char *s, unsigned long long v[4])
{
  return sscanf(s, "%llu,%llu,%llu,%llu", &v[0], &v[1], &v[2], &v[3]) == 4;
}

int main(int argc, char **argv)
{
  if (argc < 10)
    die("usage: ref_dump full|stages|decode input nX nY nZ nTS nPar prec rate prefix [--decomp a,b,c,d]"
        " [--prec a,b,c,d] [--cb a,b,c,d] [--qm N] [--tiles a,b,c,d] [--erres] [--layer L]"
        " [--kernels a,b,c,d] [--prog N] [--qstyle N] [--qstep D] [--com TEXT] [--aux TEXT]");
  const char *mode = argv[1];
  const char *input = argv[2];
  uint64 nX = strtoull(argv[3], 0, 10), nY = strtoull(argv[4], 0, 10);
  uint64 nZ = strtoull(argv[5], 0, 10), nTS = strtoull(argv[6], 0, 10);
  uint8 nPar = (uint8)atoi(argv[7]);
  int prec = atoi(argv[8]);
  char rate[256];
  snprintf(rate, sizeof rate, "%s", argv[9]);
  const char *prefix = argv[10];

  unsigned long long dec[4], pr[4], cb[4], tl[4], kn[4];
  int have_dec = 0, have_pr = 0, have_cb = 0, have_qm = 0, have_tl = 0, erres = 0, layer = 0;
  int have_kn = 0, have_prog = 0, have_qstyle = 0, have_qstep = 0;
  int qm = 0, prog = 0, qstyle = 0;
  double qstep = 0;
  const char *com = NULL, *aux = NULL;
  for (int i = 11; i < argc; ++i)
  {
    if (!strcmp(argv[i], "--decomp") && i + 1 < argc) have_dec = parse4(argv[++i], dec);
    else if (!strcmp(argv[i], "--prec") && i + 1 < argc) have_pr = parse4(argv[++i], pr);
    else if (!strcmp(argv[i], "--cb") && i + 1 < argc) have_cb = parse4(argv[++i], cb);
    else if (!strcmp(argv[i], "--qm") && i + 1 < argc) { have_qm = 1; qm = atoi(argv[++i]); }
    else if (!strcmp(argv[i], "--tiles") && i + 1 < argc) have_tl = parse4(argv[++i], tl);
    else if (!strcmp(argv[i], "--erres")) erres = 1;
    else if (!strcmp(argv[i], "--layer") && i + 1 < argc) layer = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--kernels") && i + 1 < argc) have_kn = parse4(argv[++i], kn);
    else if (!strcmp(argv[i], "--prog") && i + 1 < argc) { have_prog = 1; prog = atoi(argv[++i]); }
    else if (!strcmp(argv[i], "--qstyle") && i + 1 < argc) { have_qstyle = 1; qstyle = atoi(argv[++i]); }
    else if (!strcmp(argv[i], "--qstep") && i + 1 < argc) { have_qstep = 1; qstep = strtod(argv[++i], 0); }
    else if (!strcmp(argv[i], "--com") && i + 1 < argc) com = argv[++i];
    else if (!strcmp(argv[i], "--aux") && i + 1 < argc) aux = argv[++i];
    else die("bad option");
  }

  if (!strcmp(mode, "decode"))
  {
    size_t n;
    char path[4096];
    snprintf(path, sizeof path, "%s.bwc", prefix);
    void *cs = read_file(path, &n);
    size_t nout = (size_t)(nX * nY * nZ * nTS * nPar) * (size_t)prec;
    void *out = calloc(nout ? nout : 1, 1);
    bwc_codec *codec = bwc_alloc_decoder();
    bwc_stream *stream = bwc_init_stream(cs, out, decomp);
    if (bwc_create_decompression(codec, stream, (uint8)layer)) die("create_decompression failed");
    if (bwc_decompress(codec, stream)) die("decompress failed");
    write_file(prefix, ".dec", out, nout);
    return 0;
  }

  size_t nin;
  void *inp = read_file(input, &nin);
  size_t nelem = (size_t)(nX * nY * nZ * nTS * nPar);
  if (nin != nelem * (size_t)prec) die("input size mismatch");
  size_t outcap = nin * 2 + (1u << 20);
  unsigned char *out = calloc(outcap, 1);

  bwc_codec *codec = bwc_alloc_coder(nX, nY, nZ, nTS, nPar, (bwc_precision)prec);
  if (!codec) die("alloc_coder failed");
  if (have_kn)
    set_kernels(codec, (bwc_dwt_filter)kn[0], (bwc_dwt_filter)kn[1], (bwc_dwt_filter)kn[2], (bwc_dwt_filter)kn[3]);
  if (have_prog) set_progression(codec, (bwc_prog_ord)prog);
  if (have_qstyle) set_quant_style(codec, (bwc_quant_st)qstyle);
  if (have_dec) bwc_set_decomp(codec, (uint8)dec[0], (uint8)dec[1], (uint8)dec[2], (uint8)dec[3]);
  if (have_pr) bwc_set_precincts(codec, (uint8)pr[0], (uint8)pr[1], (uint8)pr[2], (uint8)pr[3]);
  if (have_cb) bwc_set_codeblocks(codec, (uint8)cb[0], (uint8)cb[1], (uint8)cb[2], (uint8)cb[3]);
  if (have_qm) bwc_set_qm(codec, (uint8)qm);
  if (have_tl) bwc_set_tiles(codec, tl[0], tl[1], tl[2], tl[3], bwc_tile_sizeof);
  if (erres) bwc_set_error_resilience(codec);
  if (have_qstep) set_quant_step_size(codec, qstep);

  bwc_stream *stream = bwc_init_stream(inp, out, comp);
  if (com && bwc_set_com(stream, com, (uint16)strlen(com))) die("set_com failed");
  if (aux && bwc_set_aux(stream, aux, (uint32)strlen(aux))) die("set_aux failed");
  if (bwc_create_compression(codec, stream, rate)) die("create_compression failed");

  if (!strcmp(mode, "full"))
  {
    size_t sz = bwc_compress(codec, stream);
    if (!sz) die("compress failed");
    write_file(prefix, ".bwc", out, sz);
    return 0;
  }
  if (strcmp(mode, "stages")) die("unknown mode");

  bwc_gl_ctrl *control = &codec->control;
  bwc_gl_inf *info = &codec->info;
  uint64 bsz = control->tileSizeX * control->tileSizeY * control->tileSizeZ * control->tileSizeTS;
  bwc_sample *wb = calloc(bsz, sizeof(bwc_sample));

  char path[4096];
  snprintf(path, sizeof path, "%s.dwt", prefix);
  FILE *fdwt = fopen(path, "wb");
  snprintf(path, sizeof path, "%s.t1", prefix);
  FILE *ft1 = fopen(path, "wb");
  snprintf(path, sizeof path, "%s.geo", prefix);
  FILE *fgeo = fopen(path, "wb");
  if (!fdwt || !ft1 || !fgeo) die("cannot open stage outputs");

  for (uint64 i = 0; i < control->nTiles; ++i)
  {
    bwc_tile *tile = &codec->tile[i];
    for (uint16 p = 0; p < info->nPar; ++p)
    {
      bwc_parameter *par = &tile->parameter[p];
      bwc_param_inf *pi = &par->info;
      uint64 w = pi->X1 - pi->X0, h = pi->Y1 - pi->Y0, d = pi->Z1 - pi->Z0, dt = pi->TS1 - pi->TS0;
      uint64 poff = info->nX * info->nY * info->nZ * info->nTS * p;
      bwc_float mx = pi->parameter_max, mn = pi->parameter_min;
      /* fill (libbwc.c:716-811) */
      for (uint64 t = 0; t < dt; ++t)
        for (uint64 z = 0; z < d; ++z)
          for (uint64 y = 0; y < h; ++y)
          {
            uint64 src = poff + tile->info.X0 + info->nX * (tile->info.Y0 + y + info->nY * (tile->info.Z0 + z + info->nZ * t));
            bwc_sample *dst = &wb[w * (y + h * (z + d * t))];
            for (uint64 x = 0; x < w; ++x)
            {
              dst[x].f = (prec == 8) ? (bwc_float)((double *)inp)[src + x] : (bwc_float)((float *)inp)[src + x];
              mx = MAX(mx, dst[x].f);
              mn = MIN(mn, dst[x].f);
            }
          }
      pi->parameter_max = mx;
      pi->parameter_min = mn;
      par->control.alpha = (mx + mn) / 2.0f;
      par->control.beta = 2.0f / (mx - mn);
      par->data = wb;
      /* normalize (libbwc.c:1064-1077) */
      bwc_raw shift = (uint64)control->Qm << PREC_MANTISSA;
      uint64 psz = w * h * d * dt;
      for (uint64 x = 0; x < psz; ++x)
      {
        wb[x].f = (bwc_float)(wb[x].f - par->control.alpha) * par->control.beta;
        wb[x].raw += (((wb[x].raw & EXPONENT) == 0) && ((wb[x].raw & MANTISSA) == 0)) ? 0 : shift;
      }
      if (forward_wavelet_transform(codec, par)) die("dwt failed");
      fwrite(wb, sizeof(bwc_sample), psz, fdwt);

      /* geometry: one record per codeblock in access order */
      for (uint64 c = 0; c < par->control.number_of_codeblocks; ++c)
      {
        bwc_cblk_access *a = &par->access[c];
        uint64 g[8] = {a->codeblock->info.X0, a->codeblock->info.Y0, a->codeblock->info.Z0, a->codeblock->info.TS0,
                       a->codeblock->info.X1, a->codeblock->info.Y1, a->codeblock->info.Z1, a->codeblock->info.TS1};
        fwrite(g, sizeof g, 1, fgeo);
        double sb[3] = {a->subband->info.dwt_gain, a->subband->control.qt_step_size,
                        a->subband->control.qt_effective_step_size};
        fwrite(sb, sizeof sb, 1, fgeo);
        uint32 m[4] = {a->subband->control.highband_flag, a->subband->control.Kmax,
                       a->subband->control.qt_exponent, a->codeblock->info.idx};
        fwrite(m, sizeof m, 1, fgeo);
      }

      if (t1_encode(codec, tile, par)) die("t1 failed");
      for (uint64 c = 0; c < par->control.number_of_codeblocks; ++c)
      {
        bwc_encoded_cblk *e = par->access[c].codeblock->encoded_block;
        uint8 hdr[4] = {e->K, e->Kmsbs, e->Z, 0};
        fwrite(hdr, 1, 4, ft1);
        fwrite(e->S, sizeof(uint16), (size_t)e->Z + 1, ft1);
        fwrite(e->L, sizeof(uint64), e->Z, ft1);
        uint64 nb = (e->Z && e->data) ? e->L[e->Z - 1] : 0;
        fwrite(&nb, sizeof nb, 1, ft1);
        if (nb) fwrite(e->data, 1, nb, ft1);
      }
      uint16 sl[2] = {tile->control.slope_min, tile->control.slope_max};
      fwrite(sl, sizeof sl, 1, ft1);
      memset(wb, 0, bsz * sizeof(bwc_sample));
      par->data = NULL;
    }
    if (t2_encode(codec, tile)) die("t2 failed");
  }
  fclose(fdwt);
  fclose(ft1);
  fclose(fgeo);
  size_t sz = assemble_codestream(codec, stream);
  if (!sz) die("assemble failed");
  write_file(prefix, ".bwc", out, sz);
  return 0;
}
