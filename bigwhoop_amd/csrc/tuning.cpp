// Kernel launch-shape knobs (tuning.hpp).
#include "tuning.hpp"

#include <dlfcn.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

namespace bwc {

namespace {

Tuning g_tune;
std::string g_source;
std::once_flag g_once;

struct Knob {
  const char *name, *env;
  int Tuning::*field;
  int lo, hi;
  bool pow2;
  uint64_t only;   // when nonzero: the valid values, bit v for value v
};

// the MQ kernel's instantiations: 1-6, 8, 16 or 32 codeblocks per wavefront
constexpr uint64_t kMqLanes = 0x7Eull | 1ull << 8 | 1ull << 16 | 1ull << 32;

const Knob kKnobs[] = {
    {"dwt_unroll", "BWC_AMD_DWT_UNROLL", &Tuning::dwt_unroll, 2, 8, true, 0},
    {"dwt_seg_log2", "BWC_AMD_DWT_SEG_LOG2", &Tuning::dwt_seg_log2, 8, 24, false, 0},
    {"dwt_seg_min", "BWC_AMD_DWT_SEG_MIN", &Tuning::dwt_seg_min, 2, 1 << 20, false, 0},
    {"dwt_nt", "BWC_AMD_DWT_NT", &Tuning::dwt_nt, 0, 3, false, 0},
    {"dwt_xy", "BWC_AMD_DWT_XY", &Tuning::dwt_xy, 0, 1, false, 0},
    {"dwt_xy_min_log2", "BWC_AMD_DWT_XY_MIN_LOG2", &Tuning::dwt_xy_min_log2, 0, 40, false, 0},
    {"dwt_inv_stream", "BWC_AMD_DWT_INV_STREAM", &Tuning::dwt_inv_stream, 0, 1, false, 0},
    {"mq_lanes_full", "BWC_AMD_MQ_LANES_FULL", &Tuning::mq_lanes_full, 1, 32, true, 0},
    {"mq_lanes_limited", "BWC_AMD_MQ_LANES_LIMITED", &Tuning::mq_lanes_limited, 1, 32, false, kMqLanes},
};

const Knob *find(const char *name) {
  for (const Knob &k : kKnobs)
    if (name && std::strcmp(k.name, name) == 0) return &k;
  return nullptr;
}

bool valid(const Knob &k, int v) {
  if (v < k.lo || v > k.hi) return false;
  if (k.only) return v < 64 && ((k.only >> v) & 1);
  return !k.pow2 || (v & (v - 1)) == 0;
}

// the tuning file: BWC_AMD_TUNING, else data/tuning_gfx950.txt next to the
// library's lib/ directory. Lines "name value"; '#' starts a comment.
std::string tuning_path() {
  if (const char *e = std::getenv("BWC_AMD_TUNING")) return e;
  Dl_info info;
  if (dladdr(reinterpret_cast<void *>(&tuning_path), &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    const size_t s = p.rfind('/');
    p = (s == std::string::npos) ? std::string(".") : p.substr(0, s);
    return p + "/../data/tuning_gfx950.txt";
  }
  return "";
}

void load() {
  const std::string path = tuning_path();
  if (!path.empty())
    if (FILE *f = std::fopen(path.c_str(), "r")) {
      char line[256];
      while (std::fgets(line, sizeof line, f)) {
        if (char *h = std::strchr(line, '#')) *h = 0;
        char name[64];
        int v;
        if (std::sscanf(line, "%63s %d", name, &v) == 2)
          if (const Knob *k = find(name))
            if (valid(*k, v)) g_tune.*(k->field) = v;
      }
      std::fclose(f);
      g_source = path;
    }
  for (const Knob &k : kKnobs)
    if (const char *e = std::getenv(k.env)) {
      const int v = std::atoi(e);
      if (valid(k, v)) g_tune.*(k.field) = v;
    }
}

}  // namespace

const Tuning &tuning() {
  std::call_once(g_once, load);
  return g_tune;
}

int set_tuning(const char *name, int value) {
  std::call_once(g_once, load);
  const Knob *k = find(name);
  if (!k || !valid(*k, value)) return 1;
  g_tune.*(k->field) = value;
  return 0;
}

int get_tuning(const char *name, int *value) {
  std::call_once(g_once, load);
  const Knob *k = find(name);
  if (!k || !value) return 1;
  *value = g_tune.*(k->field);
  return 0;
}

const char *tuning_source() {
  std::call_once(g_once, load);
  return g_source.c_str();
}

}  // namespace bwc
