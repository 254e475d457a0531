// Launch-shape knobs of the kernels: defaults, then the tuning file written
// by the autotuner (python -m bigwhoop_amd.autotune), then environment
// variables, then bwc_amd_set_tuning. None of them changes a result, only
// how the work is laid out on the GPU.
#pragma once

namespace bwc {

struct Tuning {
  // streaming DWT passes (strided axes): pair iterations unrolled per load
  // batch (2, 4 or 8), thread target 2^dwt_seg_log2 per pass, minimum pair
  // iterations per line segment
  int dwt_unroll = 4, dwt_seg_log2 = 17, dwt_seg_min = 8;
  // forward 9/7 passes: bit 0 stores, bit 1 loads of the streaming kernels
  // with the nontemporal hint (data read or written once per pass)
  int dwt_nt = 3;
  // 1: X and Y passes of a level fused into one kernel where possible, on
  // boxes of at least 2^dwt_xy_min_log2 values
  int dwt_xy = 1, dwt_xy_min_log2 = 24;
  // 1: inverse 9/7 along strided axes on the streaming kernel
  int dwt_inv_stream = 1;
  // codeblocks per MQ wavefront: complete codings (1, 2, 4, 8, 16 or 32) and
  // codings to a pass limit (1-6, 8, 16 or 32)
  int mq_lanes_full = 16, mq_lanes_limited = 5;
};

// the current values (loaded on first use)
const Tuning &tuning();
// set one knob by name; 0 on success, 1 for an unknown name or a value out of
// its range. Not synchronised with compressions running in other threads.
int set_tuning(const char *name, int value);
int get_tuning(const char *name, int *value);
// the file the values were read from ("" when none)
const char *tuning_source();

}  // namespace bwc
