// rb_host_emulation.h -- lets rb_device.h / rb_trace.h compile as plain C++
// on the host (g++), so the device algorithms can be stepped through and
// checked on the CPU. Not used by librbgpu.
#pragma once
#include <cmath>
#include <cstdint>
#include <algorithm>
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
using std::fabs;
using std::sqrt;
using std::fmax;
using std::isfinite;
using std::min;
static inline uint32_t __umulhi(uint32_t a, uint32_t b) { return uint32_t((uint64_t(a) * b) >> 32); }
