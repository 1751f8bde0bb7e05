// Left-to-right floating-point summation, bit-identical to
//     double s = 0.0; for (k = 0; k < n; k++) s += t[k];
// (the reference's InnerProduct, matrix_square_sparse.cpp:313-325), computed
// by two kernel launches over tiles of kSumTile elements.
//
// Integer view of a binade.  Let s be a normal double with |s| in
// [2^e, 2^(e+1)) and u = 2^(e-52) its ulp; s = S*u with the integer
// 2^52 <= |S| < 2^53.  For an addend t let q = t/u (exact: a power-of-two
// scaling) and n = round-to-nearest(q).  If q is not a tie (frac(q) != 1/2)
// and |S + n| stays in [2^52 + 1, 2^53 - 1], the exact sum s + t lies inside
// the binade more than u/2 from its ends, where the doubles are exactly the
// multiples of u, so fl(s + t) = (S + n)*u.  Along a run of additions that
// keeps the running sum in one binade, the sequential result is therefore
// S + (n_1 + ... + n_k) in exact integer arithmetic.
//
// Tiles in parallel (k_sum_exact, one block per tile).  Each tile guesses the
// binade of the running sum at its start from an approximate prefix (the sum
// of the approximate tile sums before it, k_sum_tiles) and records, in that
// binade's ulps, its integer total and the minimum and maximum of its
// inclusive integer prefixes, or "no summary" if some addend is a tie or too
// large.  None of this depends on the exact running sum.
//
// Chain (the last block to finish).  One thread walks the tiles in order with
// the exact running sum s.  When s lies in the guessed binade and
// S + min >= 2^52 + 1 and S + max <= 2^53 - 1 (mirrored for negative s),
// every addition of the tile stays in the binade, so the tile is applied as
// S += total.  Otherwise the tile is summed exactly by the whole block with
// the binade walker below.  A wrong guess costs time, never accuracy.
//
// Binade walker (ordered_sum_walk).  Per step the block forms the integer
// prefix for the binade of the current s over the rest of the tile and finds
// the first addend where the reasoning above does not apply (tie,
// out-of-range prefix, huge or non-finite addend, s zero/subnormal/
// non-finite).  Everything before it is exact integer work.  Where binade
// changes cluster (the running sum swinging through zero under heavy
// cancellation) one thread continues with plain IEEE additions from an LDS
// copy of the tile (a "burst") until the running sum has kept its binade for
// kBurstQuiet elements.  Both paths perform exactly the sequential additions.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <limits.h>
#include <math.h>

namespace cs {

constexpr int kSumBlock = 256;                  // threads per block
constexpr int kSumPer = 4;                      // consecutive elements per thread
constexpr int kSumTile = kSumBlock * kSumPer;   // elements per tile
constexpr int kSumBuf = kSumTile + kSumTile / 16;
constexpr int kSumWaves = kSumBlock / 64;
constexpr int kChainWin = 1024;   // tile summaries staged in LDS at a time
constexpr int kBurstMin = 64;     // elements a burst always covers
constexpr int kBurstQuiet = 64;   // binade-stable elements that end a burst
constexpr int kBurstFinish = 256; // a burst this close to the tile end finishes it

// device scratch of the reduction, one entry per tile
struct SumScratch {
  double* tsum;          // approximate tile sums
  int* code;             // guessed sign+exponent of the running sum, -1: no summary
  long long* tot;        // integer tile total in the guessed binade's ulps
  long long* pmin;       // minimum inclusive integer prefix inside the tile
  long long* pmax;       // maximum inclusive integer prefix
  unsigned int* ticket;  // blocks finished; zero between launches
  // optional counters (null: off): sums, tiles, tiles applied in O(1),
  // walker steps, burst elements
  unsigned long long* stats;
};
enum SumStat : int { kStatSums = 0, kStatTiles, kStatFast, kStatSteps, kStatBurst, kNumSumStat };

__host__ __device__ inline int sum_tiles(int n) { return n > 0 ? (n + kSumTile - 1) / kSumTile : 1; }

struct SumShared {
  long long wv[kSumWaves];
  long long wv2[kSumWaves];
  double wd[kSumWaves];
  int wi[kSumWaves];
  int pos;
  int tile;
  int last;
  double s;
  double buf[kSumBuf];
  int code[kChainWin];
  long long tot[kChainWin];
  long long pmin[kChainWin];
  long long pmax[kChainWin];
};

__device__ __forceinline__ int sum_pad(int i) { return i + (i >> 4); }

__device__ __forceinline__ int sign_exp(double x) {
  return (int)((unsigned long long)__double_as_longlong(x) >> 52);
}

__device__ __forceinline__ long long wave_incl_scan64(long long x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const long long y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x;
}

// exclusive block scan; every thread also receives the block total
__device__ __forceinline__ long long block_excl_scan64(long long x, SumShared& sh,
                                                       long long& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const long long inc = wave_incl_scan64(x);
  if (lane == 63) sh.wv[wid] = inc;
  __syncthreads();
  long long base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kSumWaves; w++) {
    const long long v = sh.wv[w];
    if (w < wid) base += v;
    tot += v;
  }
  total = tot;
  __syncthreads();
  return base + inc - x;
}

__device__ __forceinline__ int block_min_int(int x, SumShared& sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o, 64));
  if (lane == 0) sh.wi[wid] = x;
  __syncthreads();
  int m = sh.wi[0];
#pragma unroll
  for (int w = 1; w < kSumWaves; w++) m = min(m, sh.wi[w]);
  __syncthreads();
  return m;
}

__device__ __forceinline__ void block_minmax64(long long& mn, long long& mx, SumShared& sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mn = min(mn, (long long)__shfl_xor(mn, o, 64));
    mx = max(mx, (long long)__shfl_xor(mx, o, 64));
  }
  if (lane == 0) {
    sh.wv[wid] = mn;
    sh.wv2[wid] = mx;
  }
  __syncthreads();
  mn = sh.wv[0];
  mx = sh.wv2[0];
#pragma unroll
  for (int w = 1; w < kSumWaves; w++) {
    mn = min(mn, sh.wv[w]);
    mx = max(mx, sh.wv2[w]);
  }
  __syncthreads();
}

// any-order sum (only used to guess binades)
__device__ __forceinline__ double block_sum_approx(double x, SumShared& sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  if (lane == 0) sh.wd[wid] = x;
  __syncthreads();
  double s = sh.wd[0];
#pragma unroll
  for (int w = 1; w < kSumWaves; w++) s += sh.wd[w];
  __syncthreads();
  return s;
}

// Sequential IEEE additions of buf[m - base ...] onto s by one thread;
// returns the first element not yet added.
__device__ inline int ordered_sum_burst(const double* buf, int base, int m, int end, double& s) {
  int k = m, quiet = m;
  while (k < end) {
    const int hb = sign_exp(s);
    if (end - k >= 16) {
      double w[16];
#pragma unroll
      for (int j = 0; j < 16; j++) w[j] = buf[sum_pad(k - base + j)];
#pragma unroll
      for (int j = 0; j < 16; j++) s = s + w[j];
      k += 16;
    } else {
      for (; k < end; k++) s = s + buf[sum_pad(k - base)];
    }
    if (sign_exp(s) != hb) quiet = k;
    if (k - m >= kBurstMin && k - quiet >= kBurstQuiet && end - k > kBurstFinish) break;
  }
  return k;
}

// Adds t[begin, end) (at most one tile) onto sh.s exactly; called by all
// threads of the block, the result is in sh.s.
__device__ inline void ordered_sum_walk(const double* __restrict__ t, int begin, int end,
                                        SumShared& sh, unsigned long long* stats) {
  unsigned long long steps = 0, burst = 0;
  const int tid = threadIdx.x;
  const double two62 = 4611686018427387904.0;  // 2^62
  const long long lo52 = (1LL << 52) + 1, hi53 = (1LL << 53) - 1;
  double v[kSumPer];
#pragma unroll
  for (int j = 0; j < kSumPer; j++) {
    const int idx = begin + tid * kSumPer + j;
    v[j] = idx < end ? t[idx] : 0.0;
    sh.buf[sum_pad(tid * kSumPer + j)] = v[j];
  }
  __syncthreads();
  int pos = begin;
  while (pos < end) {
    const double s = sh.s;
    if (isnan(s)) break;  // NaN absorbs every later addend
    steps++;
    // modes: 0 zero, 1 subnormal, 2 normal, 3 infinite
    int mode;
    int scale = 0;
    long long S = 0;
    if (s == 0.0) {
      mode = 0;
    } else if (isinf(s)) {
      mode = 3;
    } else if (fabs(s) < DBL_MIN) {
      mode = 1;
    } else {
      mode = 2;
      scale = 52 - ilogb(s);
      S = (long long)ldexp(s, scale);
    }
    long long pref[kSumPer];
    long long run = 0;
    int bad = INT_MAX;
#pragma unroll
    for (int j = 0; j < kSumPer; j++) {
      const int idx = begin + tid * kSumPer + j;
      pref[j] = run;
      if (idx < pos || idx >= end || bad != INT_MAX) continue;
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
        bad = idx;  // subnormal running sum: handled by the burst
      }
    }
    long long total;
    const long long off = block_excl_scan64(run, sh, total);
    int fail = bad;
    if (mode == 2) {
      const bool neg = s < 0;
#pragma unroll
      for (int j = 0; j < kSumPer; j++) {
        const int idx = begin + tid * kSumPer + j;
        if (idx < pos || idx >= end || idx >= fail) continue;
        const long long P = S + off + pref[j];
        const long long a = neg ? -P : P;
        if (a < lo52 || a > hi53) fail = idx;
      }
    }
    const int m = block_min_int(fail, sh);
    if (m == INT_MAX) {
      // no failure up to the end of the range
      if (mode == 2 && tid == 0) sh.s = ldexp((double)(S + total), -scale);
      __syncthreads();
      pos = end;
    } else {
      // the running sum just before element m is exact: S + off + (this
      // thread's prefix before m), or s itself in modes 0, 1 and 3
      const int owner = (m - begin) / kSumPer;
      if (tid == owner) {
        const int j = (m - begin) % kSumPer;
        long long pj = 0;
#pragma unroll
        for (int k = 0; k < kSumPer; k++)
          if (k + 1 == j) pj = pref[k];
        sh.s = mode == 2 ? ldexp((double)(S + off + pj), -scale) : s;
      }
      __syncthreads();
      if (tid == 0) {
        double sb = sh.s;
        sh.pos = ordered_sum_burst(sh.buf, begin, m, end, sb);
        sh.s = sb;
      }
      __syncthreads();
      burst += sh.pos - m;
      pos = sh.pos;
    }
  }
  if (stats && tid == 0) {
    atomicAdd(stats + kStatSteps, steps);
    atomicAdd(stats + kStatBurst, burst);
  }
  __syncthreads();  // buf and sh.s are reused by the caller
}

// One reduction: the addends, their count, the scratch and an optional flag
// that cancels it.  A source policy Src provides
//   __device__ SumJob job() const        -- the reduction of grid row blockIdx.y
//   __device__ void finish(double) const -- consumes the exact sum (one thread)
struct SumJob {
  const double* t;
  int n;
  SumScratch w;
  const int* skip;
};

// approximate sum of every tile (any order); grid = (tiles of the largest job, jobs)
template <class Src>
__global__ __launch_bounds__(kSumBlock) void k_sum_tiles(Src src) {
  const SumJob job = src.job();
  if (job.skip && *job.skip) return;
  if ((int)blockIdx.x >= sum_tiles(job.n)) return;
  const double* __restrict__ t = job.t;
  const int n = job.n;
  __shared__ double wd[kSumWaves];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int base = blockIdx.x * kSumTile + tid * kSumPer;
  double a = 0.0;
#pragma unroll
  for (int j = 0; j < kSumPer; j++)
    if (base + j < n) a += t[base + j];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
  if (lane == 0) wd[wid] = a;
  __syncthreads();
  if (tid == 0) {
    double s = wd[0];
    for (int k = 1; k < kSumWaves; k++) s += wd[k];
    job.w.tsum[blockIdx.x] = s;
  }
}

// Tile summaries, then the chain in the last block of the job to finish;
// src.finish(s) runs on one thread with the exact sum.  Same grid as k_sum_tiles.
template <class Src>
__global__ __launch_bounds__(kSumBlock) void k_sum_exact(Src src) {
  const SumJob job = src.job();
  if (job.skip && *job.skip) return;
  const double* __restrict__ t = job.t;
  const int n = job.n;
  const SumScratch& w = job.w;
  const int b = blockIdx.x, nt = sum_tiles(n), tid = threadIdx.x;
  if (b >= nt) return;
  __shared__ SumShared sh;
  const double two62 = 4611686018427387904.0;  // 2^62
  const long long lo52 = (1LL << 52) + 1, hi53 = (1LL << 53) - 1;
  {
    // approximate running sum at the start of this tile -> binade guess
    double p = 0.0;
    for (int j = tid; j < b; j += kSumBlock) p += w.tsum[j];
    p = block_sum_approx(p, sh);
    int code = -1, scale = 0;
    if (fabs(p) >= DBL_MIN && fabs(p) <= DBL_MAX) {
      code = sign_exp(p);
      scale = 52 - ilogb(p);
    }
    long long pref[kSumPer];
    long long run = 0;
    int ok = code >= 0;
#pragma unroll
    for (int j = 0; j < kSumPer; j++) {
      const int idx = b * kSumTile + tid * kSumPer + j;
      const double x = idx < n ? t[idx] : 0.0;
      if (ok) {
        const double q = ldexp(x, scale);
        const double r = rint(q);
        if (!(fabs(q) < two62) || fabs(q - r) == 0.5)
          ok = 0;
        else
          run += (long long)r;
      }
      pref[j] = run;
    }
    ok = __syncthreads_and(ok);
    long long total;
    const long long off = block_excl_scan64(run, sh, total);
    long long mn = LLONG_MAX, mx = LLONG_MIN;
#pragma unroll
    for (int j = 0; j < kSumPer; j++) {
      mn = min(mn, off + pref[j]);
      mx = max(mx, off + pref[j]);
    }
    block_minmax64(mn, mx, sh);
    if (tid == 0) {
      w.code[b] = ok ? code : -1;
      w.tot[b] = total;
      w.pmin[b] = mn;
      w.pmax[b] = mx;
    }
  }
  // the last block to finish chains the tiles
  __threadfence();
  if (tid == 0) sh.last = atomicAdd(w.ticket, 1u) == (unsigned)(nt - 1);
  __syncthreads();
  if (!sh.last) return;
  __threadfence();
  if (tid == 0) {
    *w.ticket = 0u;
    sh.s = 0.0;
  }
  __syncthreads();
  int fallback = 0;
  for (int b0 = 0; b0 < nt; b0 += kChainWin) {
    const int wend = min(nt, b0 + kChainWin);
    for (int j = b0 + tid; j < wend; j += kSumBlock) {
      sh.code[j - b0] = w.code[j];
      sh.tot[j - b0] = w.tot[j];
      sh.pmin[j - b0] = w.pmin[j];
      sh.pmax[j - b0] = w.pmax[j];
    }
    __syncthreads();
    int bb = b0;
    while (bb < wend) {
      if (tid == 0) {
        double s = sh.s;
        int k = bb;
        for (; k < wend; k++) {
          if (isnan(s)) {
            k = wend;  // NaN absorbs everything that follows
            break;
          }
          const int code = sh.code[k - b0];
          if (code < 0 || code != sign_exp(s)) break;
          const int e = ilogb(s);
          const long long S = (long long)ldexp(s, 52 - e);
          const long long lo = S + sh.pmin[k - b0], hi = S + sh.pmax[k - b0];
          const bool inside = s > 0 ? (lo >= lo52 && hi <= hi53) : (lo >= -hi53 && hi <= -lo52);
          if (!inside) break;
          s = ldexp((double)(S + sh.tot[k - b0]), e - 52);
        }
        sh.s = s;
        sh.tile = k;
      }
      __syncthreads();
      bb = sh.tile;
      if (bb >= wend) break;
      ordered_sum_walk(t, bb * kSumTile, min(n, (bb + 1) * kSumTile), sh, w.stats);
      fallback++;
      bb++;
    }
    __syncthreads();  // the window arrays are refilled
  }
  if (tid == 0) {
    src.finish(sh.s);
    if (w.stats) {
      atomicAdd(w.stats + kStatSums, 1ull);
      atomicAdd(w.stats + kStatTiles, (unsigned long long)nt);
      atomicAdd(w.stats + kStatFast, (unsigned long long)(nt - fallback));
    }
  }
}

// a plain sum stored to memory
struct SumStore {
  SumJob j;
  double* out;
  __device__ SumJob job() const { return j; }
  __device__ void finish(double s) const { out[0] = s; }
};

// jobs = grid rows; max_n = the largest job's length
template <class Src>
inline void launch_ordered_sum(const Src& src, int max_n, int jobs, hipStream_t stream) {
  const dim3 grid(sum_tiles(max_n), jobs);
  hipLaunchKernelGGL(k_sum_tiles<Src>, grid, dim3(kSumBlock), 0, stream, src);
  hipLaunchKernelGGL(k_sum_exact<Src>, grid, dim3(kSumBlock), 0, stream, src);
}

}  // namespace cs
