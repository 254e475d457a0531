/*
 * abi_layout — offsets and sizes of the public structs a caller of libbwc
 * dereferences (reference include/library/private/types.h:89-600). Built
 * once against the reference's own headers (-DREF_HEADERS, golden values in
 * tests/golden/abi_layout.txt) and once against this project's
 * include/bwc.h; the two outputs must be identical.
 */
#include <stddef.h>
#include <stdio.h>

#ifdef REF_HEADERS
#include "libbwc.h"
#else
#include "bwc.h"
#endif

#define S(T) printf("sizeof %s %zu\n", #T, sizeof(T))
#define F(T, m) printf("%s.%s %zu %zu\n", #T, #m, offsetof(T, m), sizeof(((T *)0)->m))

int main(void)
{
  S(bwc_span);
  F(bwc_span, size); F(bwc_span, position); F(bwc_span, access); F(bwc_span, memory);
  S(bwc_gl_inf);
  F(bwc_gl_inf, nX); F(bwc_gl_inf, nY); F(bwc_gl_inf, nZ); F(bwc_gl_inf, nTS); F(bwc_gl_inf, nPar);
  F(bwc_gl_inf, data_prec); F(bwc_gl_inf, codec_prec);
  S(bwc_stream);
  F(bwc_stream, codestream.aux); F(bwc_stream, codestream.com); F(bwc_stream, inp); F(bwc_stream, out);
  F(bwc_stream, mode);
  S(bwc_gl_ctrl);
  F(bwc_gl_ctrl, CSsgc); F(bwc_gl_ctrl, error_resilience); F(bwc_gl_ctrl, tileSizeX); F(bwc_gl_ctrl, tileSizeY);
  F(bwc_gl_ctrl, tileSizeZ); F(bwc_gl_ctrl, tileSizeTS); F(bwc_gl_ctrl, nTiles); F(bwc_gl_ctrl, precSizeX);
  F(bwc_gl_ctrl, precSizeY); F(bwc_gl_ctrl, precSizeZ); F(bwc_gl_ctrl, precSizeTS); F(bwc_gl_ctrl, cbX);
  F(bwc_gl_ctrl, cbY); F(bwc_gl_ctrl, cbZ); F(bwc_gl_ctrl, cbTS); F(bwc_gl_ctrl, KernelX); F(bwc_gl_ctrl, KernelY);
  F(bwc_gl_ctrl, KernelZ); F(bwc_gl_ctrl, KernelTS); F(bwc_gl_ctrl, decompX); F(bwc_gl_ctrl, decompY);
  F(bwc_gl_ctrl, decompZ); F(bwc_gl_ctrl, decompTS); F(bwc_gl_ctrl, nDecomp); F(bwc_gl_ctrl, qt_exponent);
  F(bwc_gl_ctrl, qt_mantissa); F(bwc_gl_ctrl, Qm); F(bwc_gl_ctrl, bitrate); F(bwc_gl_ctrl, nLayers);
  F(bwc_gl_ctrl, useLayer); F(bwc_gl_ctrl, guard_bits); F(bwc_gl_ctrl, headerSize); F(bwc_gl_ctrl, codestreamSize);
  F(bwc_gl_ctrl, quantization_style); F(bwc_gl_ctrl, progression);
  S(bwc_header);
  F(bwc_header, info); F(bwc_header, control); F(bwc_header, aux); F(bwc_header, com);
  S(bwc_codec);
  F(bwc_codec, info); F(bwc_codec, control); F(bwc_codec, tile); F(bwc_codec, mode);
  printf("enum %d %d %d | %d %d %d %d %d | %d %d | %d %d | %d %d %d | %d %d\n", bwc_dwt_9_7, bwc_dwt_5_3,
         bwc_dwt_haar, bwc_prog_LRCP, bwc_prog_RLCP, bwc_prog_RPCL, bwc_prog_PCRL, bwc_prog_CPRL, bwc_qt_none,
         bwc_qt_derived, bwc_tile_sizeof, bwc_tile_numbof, bwc_precision_half, bwc_precision_single,
         bwc_precision_double, comp, decomp);
  return 0;
}
