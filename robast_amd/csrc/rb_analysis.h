// rb_analysis.h -- spot analysis after the trace (device): the TH2D fill of
// the rays' terminal (x, y) and AGeoUtil::ContainmentRadius
// (src/AGeoUtil.cxx:18-42, 198-308) over it.
//
// The search reads the histogram only through SumInRadius, an integer count
// (summed exactly in 64 bits), and is written without FMA contraction, so
// that given the same counts and statistics it retraces a plain FP64 CPU
// evaluation of the reference's loop step for step. SumInRadius reads each
// row's bins in the circle as one run of columns, from per-row prefix sums.
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

// Sum over the wavefront of one 64-bit value per lane; every lane gets it.
__device__ inline uint64_t wave_sum_u64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned lo = __shfl_xor(unsigned(v), off, 64), hi = __shfl_xor(unsigned(v >> 32), off, 64);
    v += (uint64_t(hi) << 32) | lo;
  }
  return v;
}

// Sum over the block (whole wavefronts) of one 64-bit value per thread:
// wavefront sums by shuffles, then one LDS slot per wavefront. Every thread
// gets the total. red holds blockDim.x / 64 entries.
__device__ inline uint64_t block_sum_u64(uint64_t v, uint64_t* red) {
  v = wave_sum_u64(v);
  const int waves = int(blockDim.x) >> 6;
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  uint64_t t = 0;
  for (int k = 0; k < waves; k++) t += red[k];
  __syncthreads();
  return t;
}

// Per-row inclusive prefix sums of the bin contents: p[(by - 1) nx + bx - 1]
// = contents of bins 1..bx of row by.
struct RowPrefix {
  const uint64_t* p;
};

// One SumInRadius call's circle, with the column geometry it needs: bin
// centres cx(b) = xlo + (b - 1) w + 0.5 w as hist_center computes them, and
// a = the number of columns whose centre lies left of x.
struct Circle {
  double x, y, r2, xlo, wx, inv_wx;
  int nx, a;
  __device__ double cx(int b) const {
#pragma clang fp contract(off)
    return xlo + (b - 1) * wx + 0.5 * wx;
  }
};

// Contents of the bins of row `by` whose centres lie within r of (x, y), by
// the reference's test d2 = (cx - x)^2 + (cy - y)^2 <= r^2 in FP64. Along a
// row fl(cx - x) is monotone in the column, so the test holds on one run of
// columns: on the columns left of x (1..a) it turns true once, on those from
// a + 1 it turns false once. Each end is estimated from sqrt(r^2 - dy^2) and
// settled with the exact test (a few steps, or a binary search past 8); two
// prefix sums give the run's contents.
__device__ inline uint64_t row_in_radius(const RowPrefix& P, const HistParams& H, const Circle& C, int by) {
#pragma clang fp contract(off)
  const double cy = hist_center(by, H.ny, H.ylo, H.yhi);
  const double dy2 = (cy - C.y) * (cy - C.y);
  if (!(dy2 <= C.r2)) return 0;  // (cx - x)^2 + dy2 >= dy2 > r^2 for every column
  auto in = [&](int b) {
    const double d = C.cx(b) - C.x;
    return d * d + dy2 <= C.r2;
  };
  const double half = sqrt(C.r2 - dy2);
  const int a = C.a, nx = C.nx;
  // first inside column of 1..a (a + 1: none); the test is false..true there
  double est = ceil((C.x - half - C.xlo) * C.inv_wx + 0.5);
  int f = est < 1.0 ? 1 : est > double(a + 1) ? a + 1 : int(est);
  if (f <= a && in(f)) {
    int k = 0;
    while (k < 8 && f > 1 && in(f - 1)) { f--; k++; }
    if (k == 8 && f > 1 && in(f - 1)) {
      int lo = 1, hi = f - 1;  // the first true lies in [lo, hi]
      while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (in(mid)) hi = mid; else lo = mid + 1;
      }
      f = lo;
    }
  } else {
    int k = 0;
    while (k < 8 && f <= a && !in(f)) { f++; k++; }
    if (f <= a && !in(f)) {
      int lo = f + 1, hi = a + 1;  // the first true lies in [lo, hi] (a + 1: none)
      while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (in(mid)) hi = mid; else lo = mid + 1;
      }
      f = lo;
    }
  }
  // last inside column of a + 1..nx (a: none); the test is true..false there
  est = floor((C.x + half - C.xlo) * C.inv_wx + 0.5);
  int l = est < double(a) ? a : est > double(nx) ? nx : int(est);
  if (l >= a + 1 && in(l)) {
    int k = 0;
    while (k < 8 && l < nx && in(l + 1)) { l++; k++; }
    if (k == 8 && l < nx && in(l + 1)) {
      int lo = l + 1, hi = nx;  // the last true lies in [lo, hi]
      while (lo < hi) {
        const int mid = hi - (hi - lo) / 2;
        if (in(mid)) lo = mid; else hi = mid - 1;
      }
      l = lo;
    }
  } else {
    int k = 0;
    while (k < 8 && l >= a + 1 && !in(l)) { l--; k++; }
    if (l >= a + 1 && !in(l)) {
      int lo = a, hi = l - 1;  // the last true lies in [lo, hi] (a: none)
      while (lo < hi) {
        const int mid = hi - (hi - lo) / 2;
        if (in(mid)) lo = mid; else hi = mid - 1;
      }
      l = lo;
    }
  }
  const int b0 = f <= a ? f : a + 1, b1 = l >= a + 1 ? l : a;
  if (b1 < b0) return 0;
  const uint64_t* row = P.p + int64_t(by - 1) * nx;
  return row[b1 - 1] - (b0 > 1 ? row[b0 - 2] : 0);
}

// SumInRadius (src/AGeoUtil.cxx:18-42): contents of the bins whose centres
// lie within r of (x, y). All threads of the block call it together and
// share the rows; rows whose centres lie more than r (plus one bin) from y
// hold no such bin and are skipped. The sum is the full scan's.
__device__ inline double sum_in_radius(const RowPrefix& P, const HistParams& H, double x, double y, double r,
                                       uint64_t* red) {
#pragma clang fp contract(off)
  const double r2 = r * r;
  if (!(x == x) || !(y == y) || !(r2 == r2)) return 0.0;  // NaN: the reference's test fails for every bin
  Circle C;
  C.x = x;
  C.y = y;
  C.r2 = r2;
  C.xlo = H.xlo;
  C.wx = (H.xhi - H.xlo) / double(H.nx);
  C.inv_wx = 1.0 / C.wx;
  C.nx = H.nx;
  const double ea = floor((x - H.xlo) * C.inv_wx + 0.5);  // columns with cx < x, then settled exactly
  int a = ea < 0.0 ? 0 : ea > double(H.nx) ? H.nx : int(ea);
  while (a < H.nx && C.cx(a + 1) < x) a++;
  while (a > 0 && !(C.cx(a) < x)) a--;
  C.a = a;
  const double w = (H.yhi - H.ylo) / double(H.ny);
  const double lo = fmin(fmax(floor((y - r - H.ylo) / w + 0.5) - 1.0, 1.0), double(H.ny));
  const double hi = fmin(fmax(ceil((y + r - H.ylo) / w + 0.5) + 1.0, 1.0), double(H.ny));
  uint64_t acc = 0;
  for (int by = int(lo) + int(threadIdx.x); by <= int(hi); by += int(blockDim.x)) acc += row_in_radius(P, H, C, by);
  return double(block_sum_u64(acc, red));
}

// Each hill-climbing run of the centre is capped at kMaxClimb steps (the
// reference's while loops have no cap; each step raises the contained count,
// so the cap only bounds the run time on adversarial histograms).
constexpr int kMaxClimb = 1 << 16;

// ContainmentRadius (src/AGeoUtil.cxx:198-308), run by every thread of one
// block on the same (uniform) values; st = {sum w, x, x^2, y, y^2} of the
// fill (TH1::GetMean / GetStdDev), integral = TH1::Integral. out = {r, x, y}.
__device__ inline void containment_search(const RowPrefix& P, const HistParams& H, const double* st,
                                          double integral, double fraction, uint64_t* red, double* out) {
#pragma clang fp contract(off)
  auto sir = [&](double px, double py, double pr) { return sum_in_radius(P, H, px, py, pr, red); };
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
