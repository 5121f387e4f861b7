// rhg_synth.cpp -- synthetic framework positions for benchmark workloads:
// random sequential addition of uniform fractional positions under a
// minimum-image minimum-separation rule, with a cell list over fractional
// coordinates.  No counterpart in RASPA2 (its frameworks come from CIF files).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/rhg.h"
#include "rhg_common.h"

using namespace rhg;

extern "C" int rhg_synthetic_positions(uint64_t seed, const double box[9], int32_t n, double d_min,
                                       int64_t max_attempts, double *xyz, int32_t *n_placed) {
  if (!box || !xyz || n < 0 || !(d_min >= 0.0)) return RHG_ERR_INVALID;
  double inv[9];
  {
    const double *in = box;
    inv[AX] = in[BY] * in[CZ] - in[BZ] * in[CY];
    inv[AY] = in[AZ] * in[CY] - in[AY] * in[CZ];
    inv[AZ] = in[AY] * in[BZ] - in[AZ] * in[BY];
    inv[BX] = in[BZ] * in[CX] - in[BX] * in[CZ];
    inv[BY] = in[AX] * in[CZ] - in[AZ] * in[CX];
    inv[BZ] = in[AZ] * in[BX] - in[AX] * in[BZ];
    inv[CX] = in[BX] * in[CY] - in[BY] * in[CX];
    inv[CY] = in[AY] * in[CX] - in[AX] * in[CY];
    inv[CZ] = in[AX] * in[BY] - in[AY] * in[BX];
    double d = in[AX] * inv[AX] + in[BX] * inv[AY] + in[CX] * inv[AZ];
    if (d == 0.0) return RHG_ERR_INVALID;
    for (int i = 0; i < 9; ++i) inv[i] /= d;
  }
  // perpendicular widths = 1/|row of inverse|
  double w[3];
  w[0] = 1.0 / std::sqrt(inv[AX] * inv[AX] + inv[BX] * inv[BX] + inv[CX] * inv[CX]);
  w[1] = 1.0 / std::sqrt(inv[AY] * inv[AY] + inv[BY] * inv[BY] + inv[CY] * inv[CY]);
  w[2] = 1.0 / std::sqrt(inv[AZ] * inv[AZ] + inv[BZ] * inv[BZ] + inv[CZ] * inv[CZ]);
  int nc[3];
  for (int d = 0; d < 3; ++d) {
    nc[d] = d_min > 0.0 ? (int)std::floor(w[d] / d_min) : 1;
    if (nc[d] < 1) nc[d] = 1;
    if (nc[d] > 256) nc[d] = 256;
  }
  const int ncell = nc[0] * nc[1] * nc[2];
  std::vector<std::vector<int32_t>> cells((size_t)ncell);
  std::vector<double> fr;
  fr.reserve((size_t)n * 3);
  const double dmin2 = d_min * d_min;
  const uint64_t key = stream_key(seed, 0x5EED);
  uint64_t ctr = 0;
  int32_t placed = 0;
  int64_t attempts = 0;
  while (placed < n && attempts < max_attempts) {
    ++attempts;
    double s[3];
    for (int d = 0; d < 3; ++d) {
      s[d] = stream_uniform(key, ctr++);
      if (s[d] >= 1.0) s[d] = 0.0;
    }
    int ci[3];
    for (int d = 0; d < 3; ++d) {
      ci[d] = (int)(s[d] * nc[d]);
      if (ci[d] >= nc[d]) ci[d] = nc[d] - 1;
    }
    bool ok = true;
    const int rx = nc[0] >= 3 ? 1 : nc[0] / 2 + 1, ry = nc[1] >= 3 ? 1 : nc[1] / 2 + 1,
              rz = nc[2] >= 3 ? 1 : nc[2] / 2 + 1;
    for (int a = -rx; a <= rx && ok; ++a)
      for (int b = -ry; b <= ry && ok; ++b)
        for (int c = -rz; c <= rz && ok; ++c) {
          int x = ((ci[0] + a) % nc[0] + nc[0]) % nc[0];
          int y = ((ci[1] + b) % nc[1] + nc[1]) % nc[1];
          int z = ((ci[2] + c) % nc[2] + nc[2]) % nc[2];
          for (int32_t j : cells[(size_t)(x * nc[1] + y) * nc[2] + z]) {
            double t[3];
            for (int d = 0; d < 3; ++d) {
              t[d] = s[d] - fr[3 * j + d];
              t[d] -= std::nearbyint(t[d]);
            }
            double dx = box[AX] * t[0] + box[BX] * t[1] + box[CX] * t[2];
            double dy = box[AY] * t[0] + box[BY] * t[1] + box[CY] * t[2];
            double dz = box[AZ] * t[0] + box[BZ] * t[1] + box[CZ] * t[2];
            if (dx * dx + dy * dy + dz * dz < dmin2) {
              ok = false;
              break;
            }
          }
        }
    if (!ok) continue;
    fr.push_back(s[0]);
    fr.push_back(s[1]);
    fr.push_back(s[2]);
    cells[(size_t)(ci[0] * nc[1] + ci[1]) * nc[2] + ci[2]].push_back(placed);
    xyz[3 * placed + 0] = box[AX] * s[0] + box[BX] * s[1] + box[CX] * s[2];
    xyz[3 * placed + 1] = box[AY] * s[0] + box[BY] * s[1] + box[CY] * s[2];
    xyz[3 * placed + 2] = box[AZ] * s[0] + box[BZ] * s[1] + box[CZ] * s[2];
    ++placed;
  }
  if (n_placed) *n_placed = placed;
  return placed == n ? RHG_OK : RHG_ERR_INVALID;
}
