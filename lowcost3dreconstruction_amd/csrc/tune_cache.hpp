// tune_cache.hpp -- the autotuner's cache file: one line per tuned workload,
// "<key>\t<tile> <precision_mode> <cell_factor> <ms>", the last line for a key
// winning.  Host-only and free of HIP calls, so the CPU tests exercise the same
// code through liblcr_mathcheck.so.
#pragma once
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>

namespace lcr {

// The workload's shape: the device (arch and CU count), clouds, pairs, pose
// hypotheses per pair, points over all clouds, the radius, the estimator and the
// k of the normals.
inline std::string tune_key(const std::string& device, int clouds, int pairs, int hyp,
                            int64_t points, double radius, int point_to_plane, int normals_k) {
  char buf[512];
  snprintf(buf, sizeof buf, "dev=%s clouds=%d pairs=%d hyp=%d points=%lld r=%.6g p2pl=%d k=%d",
           device.c_str(), clouds, pairs, hyp, (long long)points, radius, point_to_plane ? 1 : 0,
           normals_k);
  return buf;
}

// A cached choice is usable when its tile is one the kernels are built for, its
// precision mode is one of LCR_PREC_FP32 .. LCR_PREC_FP32_TILES (0..3) and, for
// the wavefront-tile modes (1..3), its tile is 64.
inline bool tune_entry_valid(int tile, int prec, double cell_factor) {
  if (tile != 64 && tile != 128 && tile != 256) return false;
  if (prec < 0 || prec > 3) return false;
  if (prec != 0 && tile != 64) return false;
  return cell_factor > 0.0;
}

// Parses one cache line; true when it is a valid entry for `key`.
inline bool tune_parse_line(const char* line, const std::string& key, int* tile, int* prec,
                            double* cell_factor) {
  const char* tab = strchr(line, '\t');
  if (!tab || std::string(line, (size_t)(tab - line)) != key) return false;
  int t = 0, p = 0;
  double cf = 0.0;
  if (sscanf(tab + 1, "%d %d %lf", &t, &p, &cf) != 3 || !tune_entry_valid(t, p, cf)) return false;
  *tile = t;
  *prec = p;
  *cell_factor = cf;
  return true;
}

// Looks `key` up in the cache file at `path`; the last valid entry wins.
inline bool tune_lookup_file(const char* path, const std::string& key, int* tile, int* prec,
                             double* cell_factor) {
  FILE* f = fopen(path, "r");
  if (!f) return false;
  char line[1024];
  bool found = false;
  while (fgets(line, sizeof line, f)) {
    int t, p;
    double cf;
    if (tune_parse_line(line, key, &t, &p, &cf)) {
      *tile = t;
      *prec = p;
      *cell_factor = cf;
      found = true;
    }
  }
  fclose(f);
  return found;
}

}  // namespace lcr
