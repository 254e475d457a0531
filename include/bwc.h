/*
 * bwc.h — public C interface of the MI355X BigWhoop library (libbwc.so).
 *
 * Drop-in for the reference library's public header, which the reference
 * generates from include/library/private/{libbwc,types,constants}.h
 * (public_header.py:96-384). Function names, argument meaning, return codes and
 * the layouts of every struct a caller may dereference are the same, so code
 * built against the reference header links against this library unchanged.
 *
 * Interfaces replaced (reference file:line):
 *   bwc_init_stream            libbwc.h:67   / libbwc.c:1229
 *   bwc_set_com, bwc_set_aux   libbwc.h:71,75 / libbwc.c:1264,1306
 *   create_codec               libbwc.h:79   / libbwc.c:1345
 *   configure_codec            libbwc.h:81   / libbwc.c:1828
 *   bwc_alloc_coder            libbwc.h:89   / libbwc.c:1999
 *   bwc_alloc_decoder          libbwc.h:96   / libbwc.c:2020
 *   bwc_free_codec             libbwc.h:98   / libbwc.c:1726
 *   bwc_set_error_resilience   libbwc.h:100  / libbwc.c:2039
 *   set_quant_style            libbwc.h:102  / libbwc.c:2079
 *   set_quant_step_size        libbwc.h:105  / libbwc.c:2120
 *   set_progression            libbwc.h:108  / libbwc.c:2198
 *   set_kernels                libbwc.h:111  / libbwc.c:2244
 *   bwc_set_decomp             libbwc.h:117  / libbwc.c:2293
 *   bwc_set_precincts          libbwc.h:123  / libbwc.c:2435
 *   bwc_set_codeblocks         libbwc.h:129  / libbwc.c:2513
 *   bwc_set_qm                 libbwc.h:135  / libbwc.c:2588
 *   bwc_set_tiles              libbwc.h:138  / libbwc.c:2648
 *   bwc_open_header            libbwc.h:145  / libbwc.c:2840
 *   bwc_close_header           libbwc.h:147  / libbwc.c:2937
 *   bwc_create_compression     libbwc.h:149  / libbwc.c:2970
 *   bwc_compress               libbwc.h:153  / libbwc.c:3145
 *   bwc_create_decompression   libbwc.h:156  / libbwc.c:3453
 *   bwc_decompress             libbwc.h:160  / libbwc.c:3485
 *
 * Difference in behaviour a caller can rely on: stream->inp (compression) may
 * also be a HIP device pointer; the library then reads the field in place
 * instead of copying it over PCIe. stream->out is always host memory.
 */
#ifndef BWC_H
#define BWC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned char uchar;
typedef unsigned short ushort;
typedef unsigned int uint;
typedef int8_t int8;
typedef uint8_t uint8;
typedef int16_t int16;
typedef uint16_t uint16;
typedef int32_t int32;
typedef uint32_t uint32;
typedef int64_t int64;
typedef uint64_t uint64;
typedef double bwc_float;
typedef uint64 bwc_raw;

/* enums (constants.h:53-118) */
typedef enum { bwc_dwt_9_7, bwc_dwt_5_3, bwc_dwt_haar } bwc_dwt_filter;
typedef enum { bwc_prog_LRCP, bwc_prog_RLCP, bwc_prog_RPCL, bwc_prog_PCRL, bwc_prog_CPRL } bwc_prog_ord;
typedef enum { bwc_qt_none, bwc_qt_derived } bwc_quant_st;
typedef enum { bwc_tile_sizeof, bwc_tile_numbof } bwc_tile_instr;
typedef enum { bwc_precision_half = 2, bwc_precision_single = 4, bwc_precision_double = 8 } bwc_precision;
typedef enum { comp, decomp } bwc_mode;

/* byte span (types.h:89-96) */
typedef struct
{
  uint64 size;
  uint64 position;
  uchar *access;
  uchar *memory;
} bwc_span;

/* global data-set information (types.h:137-146) */
typedef struct
{
  uint64 nX, nY, nZ;
  uint64 nTS;
  uint8 nPar;
  bwc_precision data_prec;
  bwc_precision codec_prec;
} bwc_gl_inf;

/* user stream (types.h:155-167) */
typedef struct
{
  struct codestream
  {
    bwc_span *aux;
    bwc_span *com;
  } codestream;
  void *inp;
  void *out;
  bwc_mode mode;
} bwc_stream;

/* global control parameters (types.h:532-570) */
typedef struct
{
  uint16 CSsgc;
  uchar error_resilience;
  uint64 tileSizeX, tileSizeY;
  uint64 tileSizeZ, tileSizeTS;
  uint64 nTiles;
  uint8 precSizeX, precSizeY;
  uint8 precSizeZ, precSizeTS;
  uint8 cbX, cbY;
  uint8 cbZ, cbTS;
  bwc_dwt_filter KernelX, KernelY;
  bwc_dwt_filter KernelZ, KernelTS;
  uint8 decompX, decompY;
  uint8 decompZ, decompTS;
  uint8 nDecomp;
  uint32 qt_exponent, qt_mantissa;
  uint8 Qm;
  float *bitrate;
  uint8 nLayers;
  uint8 useLayer;
  uint8 guard_bits;
  uint64 headerSize;
  uint64 codestreamSize;
  bwc_quant_st quantization_style;
  bwc_prog_ord progression;
} bwc_gl_ctrl;

/* header probe result (types.h:578-584) */
typedef struct
{
  bwc_gl_inf info;
  bwc_gl_ctrl control;
  bwc_span aux;
  bwc_span com;
} bwc_header;

/* codec (types.h:592-600); `tile` is owned by the library and opaque here */
typedef struct
{
  bwc_gl_inf info;
  bwc_gl_ctrl control;
  void *tile;
  bwc_mode mode;
} bwc_codec;

bwc_stream *bwc_init_stream(void *const inpbuf, void *const outbuf, bwc_mode const mode);
uchar bwc_set_com(bwc_stream *const stream, char const *const com, uint16 const size);
uchar bwc_set_aux(bwc_stream *const stream, char const *const aux, uint32 const size);
uchar create_codec(bwc_codec *const codec);
bwc_codec *configure_codec(bwc_codec *const codec, uint64 const nX, uint64 const nY, uint64 const nZ,
                           uint64 const nTS, uint8 const nPar, bwc_precision const prec);
bwc_codec *bwc_alloc_coder(uint64 const nX, uint64 const nY, uint64 const nZ, uint64 const nTS, uint8 const nPar,
                           bwc_precision const prec);
bwc_codec *bwc_alloc_decoder(void);
void bwc_free_codec(bwc_codec *const codec);
void bwc_set_error_resilience(bwc_codec *const codec);
void set_quant_style(bwc_codec *const codec, bwc_quant_st const quantization_style);
void set_quant_step_size(bwc_codec *const codec, double const delta);
void set_progression(bwc_codec *const codec, bwc_prog_ord const progression);
void set_kernels(bwc_codec *const codec, bwc_dwt_filter const KernelX, bwc_dwt_filter const KernelY,
                 bwc_dwt_filter const KernelZ, bwc_dwt_filter const KernelTS);
void bwc_set_decomp(bwc_codec *const codec, uint8 const decompX, uint8 const decompY, uint8 const decompZ,
                    uint8 const decompTS);
void bwc_set_precincts(bwc_codec *const codec, uint8 const pX, uint8 const pY, uint8 const pZ, uint8 const pTS);
void bwc_set_codeblocks(bwc_codec *const codec, uint8 const cbX, uint8 const cbY, uint8 const cbZ,
                        uint8 const cbTS);
void bwc_set_qm(bwc_codec *const codec, uint8 const Qm);
void bwc_set_tiles(bwc_codec *const codec, uint64 const tilesX, uint64 const tilesY, uint64 const tilesZ,
                   uint64 const tilesTS, bwc_tile_instr const instr);
bwc_header *bwc_open_header(void *const inpbuf);
void bwc_close_header(bwc_header *const header);
uchar bwc_create_compression(bwc_codec *const codec, bwc_stream *const data, char *const rate_control);
size_t bwc_compress(bwc_codec *const codec, bwc_stream *const data);
uchar bwc_create_decompression(bwc_codec *const codec, bwc_stream *const stream, uint8 const layer);
uchar bwc_decompress(bwc_codec *const codec, bwc_stream *const stream);

#ifdef __cplusplus
}
#endif
#endif
