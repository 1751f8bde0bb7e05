// Left-to-right floating-point summation, bit-identical to
//     double s = 0.0; for (k = 0; k < n; k++) s += t[k];
// computed by one 1024-thread workgroup.
//
// Idea ("binade walker").  Let s be a normal double with |s| in
// [2^e, 2^(e+1)) and u = 2^(e-52) its ulp; s = S*u with the integer
// 2^52 <= |S| < 2^53.  For any addend t, write q = t/u (exact: a power-of-two
// scaling) and n = round-to-nearest(q).  If q is not a tie (frac(q) != 1/2)
// and |S + n| stays in [2^52 + 1, 2^53 - 1], the exact sum s + t lies inside
// the same binade more than u/2 from its ends, where the doubles are exactly
// the multiples of u, so fl(s + t) = (S + n)*u.  Along a run of additions that
// keeps the running sum in one binade, the sequential result is therefore
// S + (n_1 + ... + n_k) in exact integer arithmetic -- a prefix sum the
// whole workgroup computes at once.
//
// The walker takes 16 consecutive addends per thread (16384 per chunk), forms
// the integer prefix for the binade of the current s, and finds the first
// addend where the reasoning above does not apply (tie, out-of-range prefix,
// huge or non-finite addend, or s zero/subnormal/non-finite).  Everything
// before it is exact integer work; that one addend is added with an ordinary
// IEEE addition, and the walk continues from the next element with the new
// binade.  Each walker step costs two block-wide reductions; the number of
// steps is the number of chunks plus the number of binade changes of the
// running sum.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace cs {

constexpr int kSumThreads = 1024;
constexpr int kSumPerThread = 16;
constexpr int kSumChunk = kSumThreads * kSumPerThread;

struct OrderedSumShared {
  long long wave_tot[kSumThreads / 64];
  int wave_min[kSumThreads / 64];
  long long scan_base;
  int first_fail;
  double s;
};

__device__ __forceinline__ long long wave_incl_scan(long long x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const long long y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x;
}

__device__ __forceinline__ int wave_min_int(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o, 64));
  return x;
}

// exclusive block scan (1024 threads), all threads receive their offset
__device__ __forceinline__ long long block_excl_scan64(long long x, OrderedSumShared& sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const long long inc = wave_incl_scan(x);
  if (lane == 63) sh.wave_tot[wid] = inc;
  __syncthreads();
  if (wid == 0) {
    long long w = lane < kSumThreads / 64 ? sh.wave_tot[lane] : 0;
    const long long wi = wave_incl_scan(w);
    if (lane < kSumThreads / 64) sh.wave_tot[lane] = wi - w;
  }
  __syncthreads();
  return sh.wave_tot[wid] + inc - x;
}

__device__ __forceinline__ int block_min_int(int x, OrderedSumShared& sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int w = wave_min_int(x);
  if (lane == 0) sh.wave_min[wid] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = sh.wave_min[0];
    for (int k = 1; k < kSumThreads / 64; k++) m = min(m, sh.wave_min[k]);
    sh.first_fail = m;
  }
  __syncthreads();
  return sh.first_fail;
}

// Must be called by all kSumThreads threads of the block; every thread
// returns the sum.
__device__ double ordered_sum_exact(const double* __restrict__ t, int n) {
  __shared__ OrderedSumShared sh;
  const int tid = threadIdx.x;
  const double two62 = 4611686018427387904.0;  // 2^62
  const long long lo52 = (1LL << 52) + 1, hi53 = (1LL << 53) - 1;
  if (tid == 0) sh.s = 0.0;
  __syncthreads();
  for (int base = 0; base < n; base += kSumChunk) {
    double v[kSumPerThread];
#pragma unroll
    for (int j = 0; j < kSumPerThread; j++) {
      const int idx = base + tid * kSumPerThread + j;
      v[j] = idx < n ? t[idx] : 0.0;
    }
    const int end = min(base + kSumChunk, n);
    int pos = base;
    while (pos < end) {
      const double s = sh.s;
      if (isnan(s)) break;  // NaN absorbs every later addend
      // modes: 0 zero, 1 subnormal, 2 normal, 3 infinite
      int mode;
      int scale = 0;
      long long S = 0;
      if (s == 0.0) {
        mode = 0;
      } else if (isinf(s)) {
        mode = 3;
      } else if (fabs(s) < 2.2250738585072014e-308) {
        mode = 1;
      } else {
        mode = 2;
        scale = 52 - ilogb(s);
        S = (long long)ldexp(s, scale);
      }
      long long pref[kSumPerThread];
      long long run = 0;
      int bad = 0x7fffffff;
#pragma unroll
      for (int j = 0; j < kSumPerThread; j++) {
        const int idx = base + tid * kSumPerThread + j;
        pref[j] = run;
        if (idx < pos || idx >= end || bad != 0x7fffffff) continue;
        const double x = v[j];
        if (mode == 2) {
          const double q = ldexp(x, scale);
          const double r = rint(q);
          if (!(fabs(q) < two62) || fabs(q - r) == 0.5) {
            bad = idx;
            continue;
          }
          run += (long long)r;
          pref[j] = run;
        } else if (mode == 0) {
          if (x != 0.0) bad = idx;  // zeros keep s = +0
        } else if (mode == 3) {
          if (isnan(x) || (isinf(x) && ((x < 0) != (s < 0)))) bad = idx;
        } else {
          bad = idx;  // subnormal running sum: one addend at a time
        }
      }
      const long long off = block_excl_scan64(run, sh);
      int fail = bad;
      if (mode == 2) {
        const bool neg = s < 0;
#pragma unroll
        for (int j = 0; j < kSumPerThread; j++) {
          const int idx = base + tid * kSumPerThread + j;
          if (idx < pos || idx >= end || idx >= fail) continue;
          const long long P = S + off + pref[j];
          const long long a = neg ? -P : P;
          if (a < lo52 || a > hi53) fail = idx;
        }
      }
      const int m = block_min_int(fail, sh);
      if (m == 0x7fffffff) {
        // no failure up to the end of the chunk
        if (mode == 2 && tid == kSumThreads - 1) sh.s = ldexp((double)(S + off + run), -scale);
        __syncthreads();
        pos = end;
      } else {
        const int owner = (m - base) / kSumPerThread;
        if (tid == owner) {
          const int j = (m - base) % kSumPerThread;
          long long pj = 0;  // integer prefix of this thread before element j
          double vj = 0.0;
#pragma unroll
          for (int k = 0; k < kSumPerThread; k++) {
            if (k + 1 == j) pj = pref[k];
            if (k == j) vj = v[k];
          }
          // the elements before m leave s unchanged in modes 0, 1 and 3
          const double before = mode == 2 ? ldexp((double)(S + off + pj), -scale) : s;
          sh.s = before + vj;
        }
        __syncthreads();
        pos = m + 1;
      }
    }
    if (isnan(sh.s)) break;
  }
  __syncthreads();
  return sh.s;
}

}  // namespace cs
