/*
 * h5z_replay — the BigWhoop HDF5 filter's calls to libbwc, replayed without
 * HDF5 (test driver; HDF5 is not installed here).
 *
 * Compiled against include/bwc.h only and linked with -lbwc, so the same
 * source runs against this project's libbwc.so and against the reference's
 * own library built from source: a drop-in check of the C boundary as the
 * filter uses it.
 *
 *   h5z_chunk()       the filter callback's two branches
 *                     (src/interfaces/HDF5/H5Zbwc.c:435-650): compress reads
 *                     the chunk geometry and options from cd_values, applies
 *                     the setters in the filter's order (error resilience,
 *                     codeblocks, precincts, tiles, decomposition, Qm), turns
 *                     the bitrate into the "%05.3f" string and compresses
 *                     into a buffer of the input's size; decompress opens the
 *                     header for the output size and decodes the layer
 *                     cd_values names.
 *   set_*()           the cd_values setter macros of
 *                     include/interfaces/HDF5/H5Zbwc.h:75-275, restated.
 *   set_local()       H5Z_bwc_set_local (H5Zbwc.c:319-430): non-unity
 *                     dataset dimensions, fastest first, data type size,
 *                     one parameter.
 *
 * Usage: h5z_replay OUTDIR TRIGFILE
 *   TRIGFILE holds 5 x 128 doubles: sin(i dX), cos(i dX), sin(j dY),
 *   cos(j dY), cos(k dZ) for i, j, k < 128 with dX = (2.0f * PI) / 128 and PI
 *   the float-rounded constant of test_hdf5.c:121 (computed with the C
 *   library's sin and cos), so that the fields below do not depend on the
 *   libm of the machine that runs this.
 * For every chunk it writes OUTDIR/<name>.bwc (the codestream) and
 * OUTDIR/<name>.dec (the decoded chunk) and prints one line
 *   <name> <input bytes> <codestream bytes>
 * then "tgv_psnr <value>" for the five Taylor-Green variables (PSNR as
 * test_hdf5.c:419-443 computes it). Exit status 0 when every call succeeded.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bwc.h"

enum { FG = 0, NX = 1, NY = 3, NZ = 5, NT = 7, NP = 9, DT = 10, ER = 11, BR = 12, CB = 13, PC = 14, TX = 15,
       TY = 17, TZ = 19, TT = 21, DL = 23, QM = 24, UL = 25, NPARMS = 26 };

/* ---- cd_values setters (H5Zbwc.h:75-275) ---- */
static void set_error_resilience(unsigned *cd)
{
  if ((cd[FG] & (1u << 0)) == 0) { cd[FG] ^= 1u << 0; cd[ER] = 1; }
}
static void set_bitrate(float br, unsigned *cd)
{
  if ((cd[FG] & (1u << 1)) == 0) { cd[FG] ^= 1u << 1; memcpy(&cd[BR], &br, 4); }
}
static unsigned pack4(unsigned a, unsigned b, unsigned c, unsigned d)
{
  return ((0xFFu & a) << 24) | ((0xFFu & b) << 16) | ((0xFFu & c) << 8) | (0xFFu & d);
}
static void set_codeblocks_d(unsigned x, unsigned y, unsigned z, unsigned t, unsigned *cd)
{
  if ((cd[FG] & (1u << 2)) == 0 && x < 512 && y < 512 && z < 512 && t < 512)
  { cd[FG] ^= 1u << 2; cd[CB] = pack4(x, y, z, t); }
}
static void set_precinct_d(unsigned x, unsigned y, unsigned z, unsigned t, unsigned *cd)
{
  if ((cd[FG] & (1u << 3)) == 0 && x < 512 && y < 512 && z < 512 && t < 512)
  { cd[FG] ^= 1u << 3; cd[PC] = pack4(x, y, z, t); }
}
static void set_tile_d(uint64_t x, uint64_t y, uint64_t z, uint64_t t, unsigned *cd)
{
  if ((cd[FG] & (1u << 4)) == 0)
  {
    cd[FG] ^= 1u << 4;
    cd[TX] = (unsigned)(x & 0xFFFFFFFF); cd[TX + 1] = (unsigned)(x >> 32);
    cd[TY] = (unsigned)(y & 0xFFFFFFFF); cd[TY + 1] = (unsigned)(y >> 32);
    cd[TZ] = (unsigned)(z & 0xFFFFFFFF); cd[TZ + 1] = (unsigned)(z >> 32);
    cd[TT] = (unsigned)(t & 0xFFFFFFFF); cd[TT + 1] = (unsigned)(t >> 32);
  }
}
static void set_decomp_d(unsigned x, unsigned y, unsigned z, unsigned t, unsigned *cd)
{
  if ((cd[FG] & (1u << 5)) == 0 && x < 255 && y < 255 && z < 255 && t < 255)
  { cd[FG] ^= 1u << 5; cd[DL] = pack4(x, y, z, t); }
}
static void set_qm(unsigned qm, unsigned *cd)
{
  if ((cd[FG] & (1u << 6)) == 0) { cd[FG] ^= 1u << 6; cd[QM] = qm; }
}
static void set_ul(unsigned ul, unsigned *cd)
{
  if ((cd[FG] & (1u << 7)) == 0) { cd[FG] ^= 1u << 7; cd[UL] = ul; }
}

/* ---- H5Z_bwc_set_local (H5Zbwc.c:319-430): dims slowest first ---- */
static int set_local(unsigned *cd, const uint64_t *dims, int ndims, unsigned dtsize)
{
  uint64_t us[8], vd[4] = {1, 1, 1, 1};
  int n = 0;
  if (dtsize != 4 && dtsize != 8) return 1;
  cd[DT] = dtsize;
  for (int i = 0; i < ndims; ++i)
    if (dims[i] > 1) us[n++] = dims[i];
  if (n <= 1 || n > 4) return 1;
  for (int i = 0; i < n; ++i) vd[i] = us[n - 1 - i];
  cd[NX] = (unsigned)(vd[0] & 0xFFFFFFFF); cd[NX + 1] = (unsigned)(vd[0] >> 32);
  cd[NY] = (unsigned)(vd[1] & 0xFFFFFFFF); cd[NY + 1] = (unsigned)(vd[1] >> 32);
  cd[NZ] = (unsigned)(vd[2] & 0xFFFFFFFF); cd[NZ + 1] = (unsigned)(vd[2] >> 32);
  cd[NT] = (unsigned)(vd[3] & 0xFFFFFFFF); cd[NT + 1] = (unsigned)(vd[3] >> 32);
  cd[NP] = 1;
  return 0;
}

/* ---- the filter callback (H5Zbwc.c:435-650); 0 = error ---- */
#define REVERSE 0x0100u /* H5Z_FLAG_REVERSE */
static size_t h5z_chunk(unsigned flags, size_t cd_nelmts, const unsigned *cd, size_t nbytes, size_t *buf_size,
                        void **buf)
{
  size_t ret = 0;
  unsigned char *input = (unsigned char *)(*buf), *output = NULL;
  bwc_stream *stream = NULL;
  bwc_codec *coder = NULL;
  char rate[20] = {0};
  if (cd_nelmts != NPARMS) return 0;
  if (flags & REVERSE)
  {
    bwc_header *header = bwc_open_header(input);
    if (!header) return 0;
    uint64_t d = header->info.nX * header->info.nY * header->info.nZ * header->info.nTS * header->info.nPar;
    if (header->info.data_prec == bwc_precision_double) d *= sizeof(double);
    else if (header->info.data_prec == bwc_precision_single) d *= sizeof(float);
    bwc_close_header(header);
    output = calloc(d, 1);
    unsigned layer = (cd[FG] & (1u << 7)) ? cd[UL] : 0;
    if (!output || !(coder = bwc_alloc_decoder()) || !(stream = bwc_init_stream(input, output, decomp)))
      goto done;
    if (bwc_create_decompression(coder, stream, (uint8)layer)) goto done;
    if (bwc_decompress(coder, stream)) goto done;
    free(*buf);
    *buf = output;
    output = NULL;
    *buf_size = (size_t)d;
    ret = (size_t)d;
  }
  else
  {
    uint64_t nX = ((uint64_t)cd[NX + 1] << 32) | cd[NX], nY = ((uint64_t)cd[NY + 1] << 32) | cd[NY];
    uint64_t nZ = ((uint64_t)cd[NZ + 1] << 32) | cd[NZ], nT = ((uint64_t)cd[NT + 1] << 32) | cd[NT];
    unsigned nPar = cd[NP];
    size_t dt = cd[DT];
    uint64_t d = nX * nY * nZ * nT * nPar * dt;
    if (nbytes != d) return 0;
    if (!(output = calloc(d, 1)) || !(stream = bwc_init_stream(input, output, comp))) goto done;
    if (!(coder = bwc_alloc_coder(nX, nY, nZ, nT, (uint8)nPar, (bwc_precision)dt))) goto done;
    unsigned fg = cd[FG];
    if (fg & (1u << 0)) bwc_set_error_resilience(coder);
    if (fg & (1u << 1))
    {
      float br;
      memcpy(&br, &cd[BR], 4);
      sprintf(rate, "%05.3f", br);
    }
    else
      sprintf(rate, "%05.3f", (float)(8.0f * dt));
    if (fg & (1u << 2))
      bwc_set_codeblocks(coder, (cd[CB] >> 24) & 0xFF, (cd[CB] >> 16) & 0xFF, (cd[CB] >> 8) & 0xFF, cd[CB] & 0xFF);
    if (fg & (1u << 3))
      bwc_set_precincts(coder, (cd[PC] >> 24) & 0xFF, (cd[PC] >> 16) & 0xFF, (cd[PC] >> 8) & 0xFF, cd[PC] & 0xFF);
    if (fg & (1u << 4))
      bwc_set_tiles(coder, ((uint64_t)cd[TX + 1] << 32) | cd[TX], ((uint64_t)cd[TY + 1] << 32) | cd[TY],
                    ((uint64_t)cd[TZ + 1] << 32) | cd[TZ], ((uint64_t)cd[TT + 1] << 32) | cd[TT], bwc_tile_sizeof);
    if (fg & (1u << 5))
      bwc_set_decomp(coder, (cd[DL] >> 24) & 0xFF, (cd[DL] >> 16) & 0xFF, (cd[DL] >> 8) & 0xFF, cd[DL] & 0xFF);
    if (fg & (1u << 6)) bwc_set_qm(coder, (uint8)cd[QM]);
    if (bwc_create_compression(coder, stream, rate)) goto done;
    d = bwc_compress(coder, stream);
    if (d == 0 || d > nbytes) goto done;
    free(*buf);
    *buf = output;
    output = NULL;
    *buf_size = (size_t)d;
    ret = (size_t)d;
  }
done:
  if (coder) bwc_free_codec(coder);
  if (stream) free(stream);
  free(output);
  return ret;
}

/* ---- driver ---- */
static const char *g_out;

static int write_file(const char *name, const char *ext, const void *p, size_t n)
{
  char path[4096];
  snprintf(path, sizeof path, "%s/%s%s", g_out, name, ext);
  FILE *f = fopen(path, "wb");
  if (!f || fwrite(p, 1, n, f) != n) return 1;
  fclose(f);
  return 0;
}

/* one chunk through the filter, both ways; the decoded chunk in *dec */
static int run_chunk(const char *name, const unsigned *cd, const void *data, size_t nbytes, void **dec)
{
  size_t bs = nbytes;
  void *buf = malloc(nbytes);
  memcpy(buf, data, nbytes);
  size_t n = h5z_chunk(0, NPARMS, cd, nbytes, &bs, &buf);
  if (!n) { fprintf(stderr, "%s: compression failed\n", name); free(buf); return 1; }
  if (write_file(name, ".bwc", buf, n)) return 1;
  size_t m = h5z_chunk(REVERSE, NPARMS, cd, n, &bs, &buf);
  if (!m) { fprintf(stderr, "%s: decompression failed\n", name); free(buf); return 1; }
  if (write_file(name, ".dec", buf, m)) return 1;
  printf("%s %zu %zu\n", name, nbytes, n);
  *dec = buf;
  return 0;
}

/* deterministic values in [0, 1) from an integer hash (splitmix64) */
static double hash01(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

int main(int argc, char **argv)
{
  if (argc < 3) { fprintf(stderr, "usage: h5z_replay OUTDIR TRIGFILE\n"); return 2; }
  g_out = argv[1];
  double tab[5][128];
  FILE *f = fopen(argv[2], "rb");
  if (!f || fread(tab, sizeof(double), 5 * 128, f) != 5 * 128) { fprintf(stderr, "bad trig file\n"); return 2; }
  fclose(f);
  const double *sx = tab[0], *cx = tab[1], *sy = tab[2], *cy = tab[3], *cz = tab[4];
  int rc = 0;

  /* 1. test_hdf5.c: five 128^3 variables of a Taylor-Green-like state
        (init_3d_field, :94-175), one dataset and chunk each, error
        resilience, 0.01 bits per value, Qm 32 (:305-315) */
  {
    const uint64_t n = 128;
    const size_t np = (size_t)(n * n * n);
    double *var[5];
    for (int p = 0; p < 5; ++p) var[p] = calloc(np, sizeof(double));
    const double GAMMA = 1.4f;
    for (uint64_t i = 0; i < n; ++i)
      for (uint64_t j = 0; j < n; ++j)
        for (uint64_t k = 0; k < n; ++k)
        {
          const size_t m = (size_t)(i + n * (j + n * k));
          const double u = sx[i] * cy[j] * cz[k];
          const double v = -cx[i] * sy[j] * cz[k];
          var[0][m] = 1.0f;
          var[1][m] = u;
          var[2][m] = v;
          var[3][m] = 0.0f;
          var[4][m] = 1.0f / ((GAMMA - 1.0f)) + (0.5f * (u * u + v * v));
        }
    unsigned cd[NPARMS] = {0};
    float br = 0.01f;
    set_error_resilience(cd);
    set_bitrate(br, cd);
    set_qm(32, cd);
    const uint64_t dims[3] = {n, n, n};
    if (set_local(cd, dims, 3, 8)) return 1;
    static const char *names[5] = {"tgv_rho", "tgv_rhou", "tgv_rhov", "tgv_rhow", "tgv_rhoE"};
    double peak = -1.7976931348623157e+308, mse = 0;
    for (int p = 0; p < 5 && !rc; ++p)
    {
      void *dec = NULL;
      rc |= run_chunk(names[p], cd, var[p], np * sizeof(double), &dec);
      if (rc) break;
      const double *vc = dec;
      double mn = 1.7976931348623157e+308, mx = -1.7976931348623157e+308;
      for (size_t i = 0; i < np; ++i)
      {
        mn = vc[i] < mn ? vc[i] : mn;
        mx = vc[i] > mx ? vc[i] : mx;
        const double d = vc[i] - var[p][i];
        mse += d * d;
      }
      peak = (mx - mn) > peak ? (mx - mn) : peak;
      free(dec);
    }
    if (!rc)
    {
      mse /= (double)np * 5;
      printf("tgv_psnr %.6f\n", 20 * log10(peak / (2 * sqrt(mse))));
    }
    for (int p = 0; p < 5; ++p) free(var[p]);
  }

  /* 2. other chunk shapes and options through the same callback */
  struct
  {
    const char *name;
    uint64_t dims[4];   /* dataset dims, slowest first */
    int ndims, dt;
    float br;           /* 0: the filter's default (8 x type size) */
    int erres, qm, ul;
    unsigned cb[4], pc[4], dl[4];
    uint64_t tl[4];
  } C[] = {
    {"chunk3d_f32_opts", {40, 48, 64}, 3, 4, 2.5f, 0, 0, 0, {5, 4, 3, 0}, {4, 4, 3, 0}, {3, 3, 2, 0}, {32, 32, 24, 0}},
    {"chunk2d_f64_rate", {1, 100, 200}, 3, 8, 1.75f, 1, 40, 0, {0}, {0}, {0}, {0}},
    {"chunk4d_f64_ul", {12, 16, 20, 24}, 4, 8, 0.5f, 0, 0, 1, {4, 4, 3, 2}, {0}, {2, 2, 2, 1}, {0}},
    {"chunk3d_f64_default", {24, 32, 40}, 3, 8, 0.0f, 0, 0, 0, {0}, {0}, {0}, {0}},
  };
  for (size_t c = 0; c < sizeof C / sizeof C[0] && !rc; ++c)
  {
    unsigned cd[NPARMS] = {0};
    if (C[c].erres) set_error_resilience(cd);
    if (C[c].br > 0) set_bitrate(C[c].br, cd);
    if (C[c].cb[0] | C[c].cb[1] | C[c].cb[2] | C[c].cb[3])
      set_codeblocks_d(C[c].cb[0], C[c].cb[1], C[c].cb[2], C[c].cb[3], cd);
    if (C[c].pc[0] | C[c].pc[1] | C[c].pc[2] | C[c].pc[3])
      set_precinct_d(C[c].pc[0], C[c].pc[1], C[c].pc[2], C[c].pc[3], cd);
    if (C[c].tl[0] | C[c].tl[1] | C[c].tl[2] | C[c].tl[3])
      set_tile_d(C[c].tl[0], C[c].tl[1], C[c].tl[2], C[c].tl[3], cd);
    if (C[c].dl[0] | C[c].dl[1] | C[c].dl[2] | C[c].dl[3])
      set_decomp_d(C[c].dl[0], C[c].dl[1], C[c].dl[2], C[c].dl[3], cd);
    if (C[c].qm) set_qm((unsigned)C[c].qm, cd);
    if (C[c].ul) set_ul((unsigned)C[c].ul, cd);
    if (set_local(cd, C[c].dims, C[c].ndims, (unsigned)C[c].dt)) return 1;
    uint64_t nx = ((uint64_t)cd[NX + 1] << 32) | cd[NX], ny = ((uint64_t)cd[NY + 1] << 32) | cd[NY];
    uint64_t nz = ((uint64_t)cd[NZ + 1] << 32) | cd[NZ], nt = ((uint64_t)cd[NT + 1] << 32) | cd[NT];
    const size_t np = (size_t)(nx * ny * nz * nt);
    void *data = malloc(np * (size_t)C[c].dt);
    /* a smooth field from the tables plus 1% hash noise, x fastest */
    for (size_t m = 0; m < np; ++m)
    {
      const uint64_t x = m % nx, y = (m / nx) % ny, z = (m / (nx * ny)) % nz, t = m / (nx * ny * nz);
      const double v = 3.0 + sx[(x * 3) & 127] * cy[(y * 5) & 127] + 0.5 * cz[(z * 7 + t * 11) & 127] +
                       0.01 * hash01(m + 1000003ull * c);
      if (C[c].dt == 8) ((double *)data)[m] = v;
      else ((float *)data)[m] = (float)v;
    }
    void *dec = NULL;
    rc |= run_chunk(C[c].name, cd, data, np * (size_t)C[c].dt, &dec);
    free(dec);
    free(data);
  }
  return rc;
}
