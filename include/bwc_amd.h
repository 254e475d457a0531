/*
 * bwc_amd.h — extra C-ABI entry points of the MI355X libbwc.so.
 *
 * None of these has a counterpart in the reference's public API: they expose
 * single stages of bwc_compress (reference libbwc.c:3231-3383) on the GPU so
 * that parity tests can compare each stage with the CPU oracle, and report the
 * device-side stage timings of the last bwc_compress call for benchmarking.
 * All pointers are plain host pointers unless stated; sizes are element
 * counts. Return 0 on success, non-zero on error (bwc_amd_last_error()).
 */
#ifndef BWC_AMD_H
#define BWC_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "bwc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* library identification; 0 when no HIP device is usable */
int bwc_amd_version(void);
int bwc_amd_device_count(void);
const char *bwc_amd_last_error(void);

/* Page-locked host memory (no counterpart in the reference). A codestream
   output buffer allocated here receives the packet bodies by DMA straight
   from the GPU; any other host buffer works too, through a staged copy.
   bwc_amd_host_alloc returns NULL on failure. */
void *bwc_amd_host_alloc(size_t bytes);
void bwc_amd_host_free(void *p);

/* The energy-gain table is process-global and accumulates on every codec
   configuration and bwc_set_decomp call, as in the reference (dwt.c:80,
   :808-1066). This returns it to the state of a fresh process, so a test or
   benchmark can run many compressions in one process with the gains a single
   compression per process would see. */
void bwc_amd_reset_gain_lut(void);

/* Compress n independently configured codecs (each after
   bwc_create_compression, each with its own stream) concurrently: `workers`
   host threads (<= 0: 8) take the codecs in turn, each codec's kernels run
   on its own HIP stream, so the GPU works on several fields at once and rate
   control of one field overlaps the GPU work of others. No counterpart in
   the reference; each codestream is byte-identical to what bwc_compress
   writes for that codec alone. sizes[i] receives the bytes of codestream i
   (0 on failure, or when the codec was not started because another one
   failed first: no further codec starts after a failure). Returns 0 when
   every codec succeeded, else 1 with the first error in
   bwc_amd_last_error(). A codec that appears twice is an error. */
int bwc_amd_compress_batch(bwc_codec *const *codecs, bwc_stream *const *streams, int n, size_t *sizes,
                           int workers);

/* Launch-shape knobs of the kernels (no counterpart in the reference); none
   changes a result. Names: "dwt_unroll" (2, 4, 8: pair iterations per load
   batch of the strided DWT passes), "dwt_seg_log2" (8-24: thread target
   2^n of those passes), "dwt_seg_min" (>= 2: minimum pair iterations per
   line segment), "dwt_nt" (0-3: nontemporal hint on the streaming DWT
   passes' stores, bit 0, and loads, bit 1), "dwt_xy" (0, 1: X and Y passes
   of a level fused into one kernel where it applies), "dwt_xy_min_log2"
   (0-40: the fusion only on boxes of at least 2^n values), "dwt_inv_stream"
   (0, 1: inverse 9/7 along strided axes on the streaming kernel),
   "mq_lanes_full" (1, 2,
   4, 8, 16, 32) and "mq_lanes_limited" (1-6, 8, 16, 32): codeblocks per MQ
   wavefront for complete codings and for codings to a pass limit. Values
   come from the defaults, then the tuning file the autotuner
   writes (python -m bigwhoop_amd.autotune; BWC_AMD_TUNING names another
   file), then environment variables (BWC_AMD_DWT_UNROLL, ...), then these
   calls. bwc_amd_tuning_source gives the file read ("" when none). Setting a
   knob while a compression runs in another thread is not supported. */
int bwc_amd_set_tuning(const char *name, int value);
int bwc_amd_get_tuning(const char *name, int *value);
const char *bwc_amd_tuning_source(void);

/* geometry of a codec after bwc_create_compression */
uint64_t bwc_amd_ncblk(bwc_codec *codec, uint64_t tile, uint8_t param);
/* per codeblock in access order (libbwc.c:1651-1683): box[8] tile-local
   x0 y0 z0 t0 x1 y1 z1 t1, sb[3] {dwt_gain, qt_step_size,
   qt_effective_step_size}, meta[4] {highband_flag, Kmax, qt_exponent,
   index in precinct} */
int bwc_amd_cblk_info(bwc_codec *codec, uint64_t tile, uint8_t param, uint32_t *box, double *sb, uint32_t *meta);
/* tile-local size of (tile, param): dims[4] */
int bwc_amd_param_dims(bwc_codec *codec, uint64_t tile, uint64_t dims[4]);

/* fill + normalize + forward DWT of one (tile, param) on the GPU
   (libbwc.c:622-1078, dwt.c:1216-1724); out receives the tile-local float64
   coefficients; mm receives {min, max, alpha, beta} */
int bwc_amd_stage_dwt(bwc_codec *codec, uint64_t tile, uint8_t param, const void *input, double *out, double mm[4]);

/* Allocation-free variant of bwc_amd_stage_dwt. No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc). input: the field in device memory ([param][t][z][y][x], as
   bwc_compress reads it); out: device buffer of the parameter's tile-local
   size (bwc_amd_param_dims) receiving the float64 coefficients; norm: device
   buffer of 4 doubles receiving {min, max, alpha, beta}; stream: a
   hipStream_t (NULL: the default stream). Asynchronous on that stream, no
   host synchronisation. Nothing is allocated: the min/max partials and the
   second buffer of the out-of-place DWT passes live in a static device
   array of 64 MiB + 128 KiB, so a tile parameter may hold at most 8,388,608
   values (larger ones return 1 with an error). Every call on a GPU shares
   that GPU's array: calls that overlap on one GPU race. */
int bwc_amd_stage_dwt_noalloc(bwc_codec *codec, uint64_t tile, uint8_t param, const void *input, double *out,
                              double *norm, void *stream);

/* inverse DWT of tile-local float64 coefficients in place on the GPU
   (dwt.c:1738-2246) */
int bwc_amd_stage_idwt(bwc_codec *codec, uint64_t tile, uint8_t param, double *coeffs);

/* tier-1 of all codeblocks of (tile, param) from tile-local coefficients
   (tier1.c:2824-3095). Per codeblock in access order: K, Kmsbs, Z (uint8),
   L (190 x uint32 bytes), S (191 x uint16); data receives each codeblock's
   L[Z-1] MQ bytes back to back, data_off[ncb+1] their offsets; slopes[2] the
   tile slope {min, max}. data_cap is the capacity of data in bytes. */
int bwc_amd_stage_t1(bwc_codec *codec, uint64_t tile, uint8_t param, const double *coeffs, uint8_t *K,
                     uint8_t *Kmsbs, uint8_t *Z, uint32_t *L, uint16_t *S, uint8_t *data, uint64_t data_cap,
                     uint64_t *data_off, uint16_t slopes[2]);

/* Host tier-2 on externally supplied tier-1 records (tier2.c:1297-1324,
   codestream.c:1329-1413): bwc_amd_begin_records resets the per-call state,
   bwc_amd_set_records installs one (tile, param) in the layout of
   bwc_amd_stage_t1 (data holds each codeblock's bytes at data_off[i],
   data_off[ncb] = total) plus {min, max, alpha, beta}, and
   bwc_amd_encode_records runs rate control, packets and codestream assembly
   into stream->out, returning the codestream size (0 on error). Used to test
   tier-2 and the codestream writer without a GPU. */
void bwc_amd_begin_records(bwc_codec *codec);
int bwc_amd_set_records(bwc_codec *codec, uint64_t tile, uint8_t param, const uint8_t *K, const uint8_t *Kmsbs,
                        const uint8_t *Z, const uint32_t *L, const uint16_t *S, const uint8_t *data,
                        const uint64_t *data_off, const double mm[4]);
size_t bwc_amd_encode_records(bwc_codec *codec, bwc_stream *stream);

/* device-side stage times of the last bwc_compress on this codec, in ms:
   [0] fill+minmax [1] normalize [2] DWT [3] quantize [4] tier-1
   [5] records D2H [6] rate control (host) [7] packet headers (host) and the
   device gather of packet bodies [8] codestream assembly with the bodies'
   D2H [9] total; launches[0] = number of DWT kernel launches */
int bwc_amd_last_timings(bwc_codec *codec, double ms[10], int launches[4]);

/* Pass skipping (no counterpart in the reference; the codestream stays the
   one bwc_compress writes without it). Tier-1 codes each codeblock only to a
   pass limit; rate control then checks that its thresholds and cuts are
   provably those of complete codings and codes flagged codeblocks further
   until they are.
   mode 0: off (default). mode 1: limit = `param` passes for every codeblock.
   mode 2: limit = the network's predicted cut (bwc_amd_set_skip_model) plus
   `param` passes. Returns 0 on success. */
int bwc_amd_set_pass_skip(bwc_codec *codec, int mode, int param);
/* Weights of the pass-limit network, fp32, in the order W1[64][16], b1[64],
   W2[64][64], b2[64], w3[64], b3, then 16 feature means and 16 feature
   scales (n = 5313 + 32 = 5345 values). */
int bwc_amd_set_skip_model(bwc_codec *codec, const float *w, size_t n);
/* Statistics of the last bwc_compress: [0] codeblocks coded, [1] passes
   coded (all rounds), [2] passes of complete codings, [3] extra
   rate-control rounds, [4] codeblocks coded further, [5] fallbacks to
   complete coding. */
int bwc_amd_skip_stats(bwc_codec *codec, uint64_t stats[8]);
/* Last bwc_compress: number of passes each codeblock of (tile, param)
   contributes to the last quality layer (ncblk entries). */
int bwc_amd_cblk_cuts(bwc_codec *codec, uint64_t tile, uint8_t param, int16_t *cuts);
/* Last bwc_compress whose rate control ran on the device (the default;
   BWC_AMD_RC=host runs it on the host): the probes the PCRD bisection of its
   last tile made, in the reference's order (tier2.c:1193-1235), as (layer,
   slope threshold, size estimate including the tile's earlier layers and
   header bytes); at most cap entries written, the count returned (-1 for a
   codec that has not compressed). */
int bwc_amd_rc_trace(bwc_codec *codec, uint32_t *layer, uint32_t *slope, int64_t *est, int cap);
/* Tier-1 calls of this codec whose symbol generation and MQ coding ran in
   the global codeblock queue of a bwc_amd_compress_batch call (fields of one
   tile and one variable; BWC_AMD_BATCH=threads codes each field on its
   own). */
uint64_t bwc_amd_batch_queued(bwc_codec *codec);
/* Pipeline graphs of a codec (BWC_AMD_GRAPH unset or 1; 0 turns them off):
   the forward DWT with tier-1 of each (tile, parameter), and the rate
   control of each tile, are recorded as HIP graphs the second time their
   launch sequence repeats and replayed after that. counts[0] = graphs
   recorded, counts[1] = graph launches, since the codec was created. */
int bwc_amd_graph_counts(bwc_codec *codec, uint64_t counts[2]);
/* Test entry point of the device rate control's header-size count: the
   bytes the packet-header writer emits for each of n bit strings (string i:
   nbits[i] bits, MSB first, from 32-bit big-endian word woff[i] of words;
   woff has n + 1 entries), with its 0xFF stuffing (after an 0xFF byte the
   next byte holds 7 bits; tier2.c / bitstream.c). Returns 0 on success. */
int bwc_amd_stuffed_bytes(const uint32_t *words, const uint64_t *woff, const uint64_t *nbits, int n,
                          uint64_t *bytes);
/* Keep the pass-limit network's 16 features per codeblock during
   bwc_compress (on != 0), and read them for (tile, param) afterwards
   (ncblk x 16 floats; feature 15 relative to the tile mean). */
int bwc_amd_collect_features(bwc_codec *codec, int on);
int bwc_amd_cblk_features(bwc_codec *codec, uint64_t tile, uint8_t param, float *feat);
/* With features collected: the MQ decisions of each codeblock's last coding
   in the last bwc_compress (ncblk entries; fast tier-1 path, 0 for
   codeblocks without significant bits), for measurement. */
int bwc_amd_cblk_decisions(bwc_codec *codec, uint64_t tile, uint8_t param, uint32_t *out);
/* Pass-limit network on the GPU (BF16 matrix cores, fp32 accumulation).
   Targets are cuts in units of 8 passes; weights as for
   bwc_amd_set_skip_model, feature means and scales set by the caller.
   train: `epochs` passes of Adam (batches of 256, order shuffled from
   `seed`), weights updated in place, mean batch loss per epoch in loss.
   grad: loss gradient of the first n <= 256 samples (5313 values, the
   weight order without means and scales), no update.
   predict: network outputs (units of 8 passes). */
int bwc_amd_skipnet_train(const float *feat, const float *target, uint64_t n, float *weights, int epochs, float lr,
                          uint64_t seed, float *loss);
int bwc_amd_skipnet_grad(const float *feat, const float *target, uint64_t n, const float *weights, float *grad,
                         float *loss);
int bwc_amd_skipnet_predict(const float *feat, uint64_t n, const float *weights, float *y);

#ifdef __cplusplus
}
#endif
#endif
