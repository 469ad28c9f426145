// rb_analysis.h -- spot analysis after the trace (device): the TH2D fill of
// the rays' terminal (x, y) and AGeoUtil::ContainmentRadius
// (src/AGeoUtil.cxx:18-42, 198-308) over it.
//
// The search reads the histogram only through SumInRadius, an integer count
// (summed exactly in 64 bits), and is written without FMA contraction, so
// that given the same counts and statistics it retraces a plain FP64 CPU
// evaluation of the reference's loop step for step. SumInRadius runs over
// the non-empty bins only, compacted row by row beforehand.
#pragma once

#include "rb_device.h"

namespace rbd {

struct HistParams {
  int32_t nx, ny;
  double xlo, xhi, ylo, yhi;
  int32_t status;
};

// TAxis::FindBin / GetBinCenter for fixed bins (ROOT): below the range 0,
// at or above it n + 1.
__device__ inline int hist_bin(double x, int n, double lo, double hi) {
#pragma clang fp contract(off)
  if (x < lo) return 0;
  if (!(x < hi)) return n + 1;
  return 1 + int(n * (x - lo) / (hi - lo));
}
__device__ inline double hist_center(int b, int n, double lo, double hi) {
#pragma clang fp contract(off)
  const double w = (hi - lo) / double(n);
  return lo + (b - 1) * w + 0.5 * w;
}

// Sum over the block of one 64-bit value per thread; every thread gets it.
__device__ inline uint64_t block_sum_u64(uint64_t v, uint64_t* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const uint64_t t = red[0];
  __syncthreads();
  return t;
}

// The non-empty bins, in bin order (row by row), with the first entry of
// each row: SumInRadius visits only the rows that can reach the circle.
struct SparseHist {
  const int32_t* bin;        // bin index (by - 1) * nx + bx - 1 of entry e
  const uint32_t* count;     // its content
  const int32_t* row_start;  // first entry of row by - 1; row_start[ny] = entries
};

// SumInRadius (src/AGeoUtil.cxx:18-42): contents of the bins whose centres
// lie within r of (x, y). All threads of the block call it together. Rows
// whose centres lie more than r (plus one bin) from y hold no such bin and
// are skipped; every other non-empty bin gets the reference's test, so the
// sum is the full scan's.
__device__ inline double sum_in_radius(const SparseHist& Hs, const HistParams& H, double x, double y, double r,
                                       uint64_t* red) {
#pragma clang fp contract(off)
  const double r2 = r * r;
  const double w = (H.yhi - H.ylo) / double(H.ny);
  const double lo = fmin(fmax(floor((y - r - H.ylo) / w + 0.5) - 1.0, 1.0), double(H.ny));
  const double hi = fmin(fmax(ceil((y + r - H.ylo) / w + 0.5) + 1.0, 1.0), double(H.ny));
  const int e0 = Hs.row_start[int(lo) - 1], e1 = Hs.row_start[int(hi)];
  uint64_t acc = 0;
  for (int e = e0 + int(threadIdx.x); e < e1; e += blockDim.x) {
    const int b = Hs.bin[e];
    const int iy = b / H.nx, ix = b - iy * H.nx;
    const double cx = hist_center(ix + 1, H.nx, H.xlo, H.xhi), cy = hist_center(iy + 1, H.ny, H.ylo, H.yhi);
    const double d2 = (cx - x) * (cx - x) + (cy - y) * (cy - y);
    if (d2 <= r2) acc += Hs.count[e];
  }
  return double(block_sum_u64(acc, red));
}

// Each hill-climbing run of the centre is capped at kMaxClimb steps (the
// reference's while loops have no cap; each step raises the contained count,
// so the cap only bounds the run time on adversarial histograms).
constexpr int kMaxClimb = 1 << 16;

// ContainmentRadius (src/AGeoUtil.cxx:198-308), run by every thread of one
// block on the same (uniform) values; st = {sum w, x, x^2, y, y^2} of the fill
// (TH1::GetMean / GetStdDev), integral = TH1::Integral. out = {r, x, y}.
__device__ inline void containment_search(const SparseHist& Hs, const HistParams& H, const double* st,
                                          double integral, double fraction, uint64_t* red, double* out) {
#pragma clang fp contract(off)
  auto sir = [&](double px, double py, double pr) { return sum_in_radius(Hs, H, px, py, pr, red); };
  int no_shift = 0, no_stable = 0;
  const double sw = st[0];
  double x = sw == 0 ? 0. : st[1] / sw;
  double y = sw == 0 ? 0. : st[3] / sw;
  const double sx = sqrt(sw == 0 ? 0. : fmax(st[2] / sw - x * x, 0.0));
  const double sy = sqrt(sw == 0 ? 0. : fmax(st[4] / sw - y * y, 0.0));
  double r = sqrt(sx * sx + sy * sy) * 1.5;
  double dr = 0.1 * r;
  const double sum_goal = integral * fraction;
  for (int i = 0; i < 100 && no_shift < 30; i++) {
    bool stable_r = false, stable_x = true, stable_y = true;
    const double sum0 = sir(x, y, r);
    double next_r = r;
    if (sum0 < sum_goal) {
      const double sum1 = sir(x, y, r + dr);
      if (sum1 == sum0) {
        dr *= 2.;
        continue;
      }
      next_r = r + dr * (sum_goal - sum0) / (sum1 - sum0);
    } else if (sum0 != sum_goal) {
      const double sum1 = sir(x, y, r - dr);
      if (sum1 == sum0) {
        dr *= 2.;
        continue;
      }
      next_r = r - dr * (sum0 - sum_goal) / (sum0 - sum1);
    }
    if (next_r < 0.) next_r = 0.5 * r;
    if (next_r < 0.5 * r) next_r = 0.5 * r;
    if (next_r > 2. * r) next_r = 2. * r;
    stable_r = fabs(next_r - r) < 0.0001 * r;
    r = next_r;
    {
      double sum1 = sir(x, y, r);
      dr *= sum0 != sum_goal ? fabs((sum1 - sum_goal) / (sum0 - sum_goal)) : 0.5;
      if (dr > 0.5 * r) dr = 0.5 * r;
      if (dr < 0.0005 * r) dr = 0.0005 * r;
      no_shift++;
      for (double dx = 0.25 * r; dx > 0.1 * dr; dx *= 0.25) {
        double sum_x1 = sir(x + dx, y, r);
        double sum_x2 = sir(x - dx, y, r);
        for (int climb = 0; climb < kMaxClimb && sum_x1 > sum1; climb++) {
          no_shift = 0;
          x += dx;
          sum_x2 = sum1;
          sum1 = sum_x1;
          sum_x1 = sir(x + dx, y, r);
          stable_x = false;
        }
        for (int climb = 0; climb < kMaxClimb && sum_x2 > sum1; climb++) {
          no_shift = 0;
          x -= dx;
          sum_x1 = sum1;
          sum1 = sum_x2;
          sum_x2 = sir(x - dx, y, r);
          stable_x = false;
        }
      }
    }
    for (double dy = 0.1 * r; dy > 0.1 * dr; dy *= 0.25) {
      double sum1 = sir(x, y, r);
      double sum_y1 = sir(x, y + dy, r);
      double sum_y2 = sir(x, y - dy, r);
      for (int climb = 0; climb < kMaxClimb && sum_y1 > sum1; climb++) {
        no_shift = 0;
        y += dy;
        sum_y2 = sum1;
        sum1 = sum_y1;
        sum_y1 = sir(x, y + dy, r);
        stable_y = false;
      }
      for (int climb = 0; climb < kMaxClimb && sum_y2 > sum1; climb++) {
        no_shift = 0;
        y -= dy;
        sum_y1 = sum1;
        sum1 = sum_y2;
        sum_y2 = sir(x, y - dy, r);
        stable_y = false;
      }
    }
    if (stable_r && stable_x && stable_y) no_stable++;
    else no_stable = 0;
    if (no_stable >= 4) break;
  }
  out[0] = r;
  out[1] = x;
  out[2] = y;
}

}  // namespace rbd
