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
// Chain (the first wave of the last block to finish).  The wave walks the
// tiles in order with the exact running sum s, held by every lane.  When s
// lies in a tile's guessed binade and S + min >= 2^52 + 1 and
// S + max <= 2^53 - 1 (mirrored for negative s), every addition of the tile
// stays in the binade and the tile is applied as S += total.  Sixty-four
// tiles are tested per step: a prefix scan of their totals gives each tile
// its S, a ballot the first that fails.  A failing tile is summed exactly by
// the binade walker below.  A wrong guess costs time, never accuracy.
//
// Binade walker (wave_walk).  Per step the wave forms the integer prefix for
// the binade of the current s over the rest of the tile, kWalkChunk addends
// (kWalkPer per lane) at a time, and finds the first addend where the reasoning above
// does not apply (tie, out-of-range prefix, huge or non-finite addend, s
// zero/subnormal/non-finite).  Everything before it is exact integer work.
// Where binade changes cluster (the running sum swinging through zero under
// heavy cancellation) the wave continues with plain IEEE additions from an
// LDS copy of the tile (a "burst") until the running sum has kept its binade
// for kBurstQuiet elements.  Both paths perform exactly the sequential
// additions.  The chain uses no block barrier.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <limits.h>
#include <math.h>

namespace cs {

// Loads and stores with global-memory instructions.  The compiler proves a
// kernel argument's pointers global, but not the pointers of a Many table
// entry, and accesses them with flat instructions; a flat access also counts
// against the LDS wait counter, so in a loop that keeps loads in flight
// across LDS traffic (the sweeps, the reductions) every LDS wait would wait
// for those loads too.  Every buffer of a cloth is device (global) memory.
template <class T>
struct GlobalWord {
  typedef T type;
};
template <>
struct GlobalWord<double2> {
  typedef double __attribute__((ext_vector_type(2))) type;
};
template <>
struct GlobalWord<int4> {
  typedef int __attribute__((ext_vector_type(4))) type;
};
template <>
struct GlobalWord<float4> {
  typedef float __attribute__((ext_vector_type(4))) type;
};
template <class T>
__device__ __forceinline__ T gload(const T* p) {
  typedef typename GlobalWord<T>::type W;
  const W w = *(const __attribute__((address_space(1))) W*)p;
  T v;
  __builtin_memcpy(&v, &w, sizeof(T));
  return v;
}
template <class T>
__device__ __forceinline__ void gstore(T* p, T v) {
  *(__attribute__((address_space(1))) T*)p = v;
}
// the same pointer in the global address space (for scalar element types)
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gptr(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gptr(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

// relaxed device-scope accesses: each is coherent across the XCDs by itself
// (no cache writeback or invalidate around it)
template <class T>
__device__ __forceinline__ void store_device(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T load_device(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr int kSumBlock = 256;                  // threads per block
constexpr int kSumPer = 4;                      // consecutive elements per thread
constexpr int kSumTile = kSumBlock * kSumPer;   // elements per tile
constexpr int kWalkPer = 4;                     // addends per lane in a walker chunk (8 measured slower)
constexpr int kWalkChunk = 64 * kWalkPer;       // addends per walker chunk
constexpr int kSumBuf = kSumTile + kWalkChunk;  // a tile plus the overreach of the last chunk
constexpr int kSumWaves = kSumBlock / 64;
constexpr int kChainWin = 1024;   // tile summaries staged in LDS at a time
constexpr int kBurstMin = 32;     // elements a burst always covers
constexpr int kBurstQuiet = 32;   // binade-stable elements that end a burst
constexpr int kBurstFinish = 64;  // a burst this close to the tile end finishes it

// device scratch of the reduction, one entry per tile
struct SumScratch {
  double* tsum;          // approximate tile sums
  int* code;             // guessed sign+exponent of the running sum, -1: no summary
  long long* tot;        // integer tile total in the guessed binade's ulps
  long long* pmin;       // minimum inclusive integer prefix inside the tile
  long long* pmax;       // maximum inclusive integer prefix
  unsigned int* ticket;  // blocks finished; zero between launches
  // optional counters (null: off): sums, tiles, tiles applied in O(1),
  // walker steps, burst elements; device wall-clock ticks spent in the
  // chain, in the walker (bursts included) and in bursts
  unsigned long long* stats;
};
enum SumStat : int {
  kStatSums = 0, kStatTiles, kStatFast, kStatSteps, kStatBurst,
  kStatChainTicks, kStatWalkTicks, kStatBurstTicks, kNumSumStat
};

__host__ __device__ inline int sum_tiles(int n) { return n > 0 ? (n + kSumTile - 1) / kSumTile : 1; }

struct SumShared {
  long long wv[kSumWaves];
  long long wv2[kSumWaves];
  double wd[kSumWaves];
  int wi[kSumWaves];
  int last;
  alignas(16) double buf[kSumBuf];
  int code[kChainWin];
  long long tot[kChainWin];
  long long pmin[kChainWin];
  long long pmax[kChainWin];
};

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

// ---------------------------------------------------------------------------
// one-wave chain

// LDS written by some lanes of the wave and read by others: the fences keep
// the compiler from moving accesses across (one wave's LDS operations
// execute in order)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int kCtrl, int kRowMask>
__device__ __forceinline__ long long dpp_add64(long long x) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)x, kCtrl, kRowMask, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), kCtrl, kRowMask, 0xf, false);
  return x + (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// inclusive prefix sum over the 64 lanes: Hillis-Steele inside each row of
// 16 lanes (row_shr 1, 2, 4, 8), then the row totals (row_bcast 15 into rows
// 1 and 3, row_bcast 31 into rows 2 and 3); source lanes outside a row or
// rows outside the mask contribute 0
__device__ __forceinline__ long long wave_scan_dpp64(long long x) {
  x = dpp_add64<0x111, 0xf>(x);
  x = dpp_add64<0x112, 0xf>(x);
  x = dpp_add64<0x114, 0xf>(x);
  x = dpp_add64<0x118, 0xf>(x);
  x = dpp_add64<0x142, 0xa>(x);
  x = dpp_add64<0x143, 0xc>(x);
  return x;
}

template <int kCtrl, int kRowMask>
__device__ __forceinline__ double dpp_f64(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), kCtrl, kRowMask, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), kCtrl, kRowMask, 0xf, false);
  return __hiloint2double(hi, lo);
}

// exclusive prefix sum over the 64 lanes in doubles: shift by one lane
// (wave_shr 1, lane 0 reads +0), then the inclusive scan above
__device__ __forceinline__ double wave_excl_scan_f64(double x) {
  x = dpp_f64<0x138, 0xf>(x);
  x = x + dpp_f64<0x111, 0xf>(x);
  x = x + dpp_f64<0x112, 0xf>(x);
  x = x + dpp_f64<0x114, 0xf>(x);
  x = x + dpp_f64<0x118, 0xf>(x);
  x = x + dpp_f64<0x142, 0xa>(x);
  x = x + dpp_f64<0x143, 0xc>(x);
  return x;
}

__device__ __forceinline__ double readlane_f64(double x, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), lane),
                          __builtin_amdgcn_readlane(__double2loint(x), lane));
}

__device__ __forceinline__ long long readlane64(long long x, int lane) {
  const int lo = __builtin_amdgcn_readlane((int)x, lane);
  const int hi = __builtin_amdgcn_readlane((int)(x >> 32), lane);
  return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// s is a normal double (its integer view exists)
__device__ __forceinline__ bool sum_normal(double s) {
  return fabs(s) >= DBL_MIN && fabs(s) <= DBL_MAX;
}

// Sequential IEEE additions of buf[m ...] onto s; every lane performs the
// same additions (uniform LDS reads).  Ends once the running sum kept its
// binade for kBurstQuiet elements (after kBurstMin at least), or at len when
// that is within kBurstFinish; returns the first element not yet added.
// Reads up to buf[len + 31].
__device__ __forceinline__ void burst_load(const double* buf, int k, double (&w)[16]) {
#pragma unroll
  for (int j = 0; j < 16; j++) w[j] = buf[k + j];
}

// adds w[0 .. min(16, len - k)) onto s (k < len) and loads the next 16
// into wn behind the first addition: waiting for w then waits only for loads
// issued a whole block earlier.  True when the burst ends after these.
__device__ __forceinline__ bool burst_block(const double (&w)[16], double (&wn)[16],
                                            const double* buf, int m, int len, int& k,
                                            int& quiet, double& s) {
  const int hb = sign_exp(s);
  s = s + w[0];
  __builtin_amdgcn_sched_barrier(0);
  burst_load(buf, k + 16, wn);
  __builtin_amdgcn_sched_barrier(0);
  if (len - k >= 16) {
#pragma unroll
    for (int j = 1; j < 16; j++) s = s + w[j];
    k += 16;
  } else {
#pragma unroll
    for (int j = 1; j < 16; j++)
      if (k + j < len) s = s + w[j];
    k = len;
  }
  if (sign_exp(s) != hb) quiet = k;
  return k >= len || (k - m >= kBurstMin && k - quiet >= kBurstQuiet && len - k > kBurstFinish);
}

__device__ inline int wave_burst(const double* buf, int m, int len, double& s) {
  int k = m, quiet = m;
  double wa[16], wb[16];  // the block being added and the next, in turn
  burst_load(buf, k, wa);
  while (true) {
    if (burst_block(wa, wb, buf, m, len, k, quiet, s)) break;
    if (burst_block(wb, wa, buf, m, len, k, quiet, s)) break;
  }
  return k;
}

// Adds buf[0, len) (a tile, len <= kSumTile) onto s exactly and returns the
// result; one wave, s the same in every lane.  Reads up to buf[kSumBuf - 1].
__device__ inline double wave_walk(const double* buf, int len, double s,
                                   unsigned long long* stats) {
  const int lane = threadIdx.x & 63;
  const double two52 = 4503599627370496.0;  // 2^52
  const double lo52 = two52 + 1.0, hi53 = 2.0 * two52 - 1.0;
  unsigned long long steps = 0, burst = 0, burst_ticks = 0;
  const unsigned long long walk_t0 = stats ? wall_clock64() : 0;
  int pos = 0;
  while (pos < len) {
    if (isnan(s)) break;  // NaN absorbs every later addend
    steps++;
    int m = len;  // first addend the integer view (or the s-keeping test) rejects
    if (sum_normal(s)) {
      // the integer view in doubles: every value used is an integer of
      // magnitude below 2^53 (exact) up to the first rejected addend
      const int scale = 52 - ilogb(s);
      const double sg = s < 0.0 ? -1.0 : 1.0;
      double S = ldexp(s, scale);
      double Sm = 0.0;
      for (int c0 = pos & ~(kWalkPer - 1); c0 < len; c0 += kWalkChunk) {
        const int i0 = c0 + kWalkPer * lane;
        double x[kWalkPer];
#pragma unroll
        for (int j = 0; j < kWalkPer; j += 2) {
          const double2 xy = *reinterpret_cast<const double2*>(buf + i0 + j);
          x[j] = xy.x;
          x[j + 1] = xy.y;
        }
        double inc[kWalkPer];
        double run = 0.0;
        unsigned usem = 0, failm = 0;  // bit j: addend j in [pos, len) / rejected
        // branch-free per addend; |r| >= 2^52 always leaves the binade, so it
        // is rejected with the ties
#pragma unroll
        for (int j = 0; j < kWalkPer; j++) {
          const int i = i0 + j;
          const bool use = i >= pos && i < len;
          const double q = ldexp(x[j], scale);
          const double r = rint(q);
          const bool conv = fabs(q) < two52 && fabs(q - r) != 0.5;
          run = run + ((use && conv) ? r : 0.0);
          inc[j] = run;
          usem |= use ? (1u << j) : 0u;
          failm |= (use && !conv) ? (1u << j) : 0u;
        }
        // exclusive prefix of the lane sums: exact up to the first lane
        // holding a rejected addend, which is all that is used
        const double off = S + wave_excl_scan_f64(run);
#pragma unroll
        for (int j = 0; j < kWalkPer; j++) {
          const double a = sg * (off + inc[j]);
          failm |= (((usem >> j) & 1u) && (a < lo52 || a > hi53)) ? (1u << j) : 0u;
        }
        const int jf = failm ? __builtin_ctz(failm) : kWalkPer;
        const unsigned long long bal = __ballot(jf < kWalkPer);
        if (bal) {
          const int f = __ffsll((long long)bal) - 1;
          double before = off;
#pragma unroll
          for (int j = 1; j < kWalkPer; j++) before = jf == j ? off + inc[j - 1] : before;
          Sm = readlane_f64(before, f);
          m = __builtin_amdgcn_readlane(i0 + jf, f);
          break;
        }
        S = readlane_f64(off + run, 63);
      }
      s = ldexp(m < len ? Sm : S, -scale);
    } else if (s == 0.0 || isinf(s)) {
      // zero addends keep s = +0; an infinite s stays unless the addend is
      // NaN or the opposite infinity
      for (int c0 = pos & ~3; c0 < len; c0 += 4 * 64) {
        const int i0 = c0 + 4 * lane;
        int jf = 4;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int i = i0 + j;
          const double x = buf[i];
          const bool keep = s == 0.0 ? x == 0.0
                                     : !(isnan(x) || (isinf(x) && ((x < 0.0) != (s < 0.0))));
          if (jf == 4 && i >= pos && i < len && !keep) jf = j;
        }
        const unsigned long long bal = __ballot(jf < 4);
        if (bal) {
          m = __builtin_amdgcn_readlane(i0 + jf, __ffsll((long long)bal) - 1);
          break;
        }
      }
    } else {
      m = pos;  // subnormal running sum: one by one
    }
    if (m >= len) break;
    const unsigned long long b0 = stats ? wall_clock64() : 0;
    pos = wave_burst(buf, m, len, s);
    burst += pos - m;
    if (stats) burst_ticks += wall_clock64() - b0;
  }
  if (stats && lane == 0) {
    atomicAdd(stats + kStatSteps, steps);
    atomicAdd(stats + kStatBurst, burst);
    atomicAdd(stats + kStatBurstTicks, burst_ticks);
    atomicAdd(stats + kStatWalkTicks, wall_clock64() - walk_t0);
  }
  return s;
}

// The chain over tiles b0 .. wend - 1 on one wave (every lane holds s):
// tiles whose summary applies are added 64 per step, a failing tile is
// walked from an LDS copy in buf.  code/tot/pmin/pmax hold the tiles'
// summaries at [b - b0]; `fallback` counts the walked tiles.
__device__ __forceinline__ double chain_tiles(double s, int b0, int wend, const int* code,
                                              const long long* tot, const long long* pmin,
                                              const long long* pmax,
                                              const double* __restrict__ t, int n, double* buf,
                                              unsigned long long* stats, int& fallback) {
  const int lane = threadIdx.x & 63;
  const long long lo52 = (1LL << 52) + 1, hi53 = (1LL << 53) - 1;
  int bb = b0;
  while (bb < wend && !isnan(s)) {
    // tiles applied from their summaries, 64 per step
    while (bb < wend && sum_normal(s)) {
      const int e = ilogb(s);
      const long long S = (long long)ldexp(s, 52 - e);
      const int k = bb + lane;
      const bool in = k < wend;
      const int kk = min(k, wend - 1) - b0;
      const long long tt = in ? tot[kk] : 0;
      const long long incl = wave_scan_dpp64(tt);
      const long long base = S + incl - tt;  // S at the start of tile k
      const long long lo = base + pmin[kk], hi = base + pmax[kk];
      const bool inside = s > 0.0 ? (lo >= lo52 && hi <= hi53) : (lo >= -hi53 && hi <= -lo52);
      const bool ok = in && code[kk] == sign_exp(s) && inside;
      const unsigned long long bad = __ballot(!ok);
      const int f = bad ? __ffsll((long long)bad) - 1 : 64;
      if (f > 0) s = ldexp((double)(S + readlane64(incl, f - 1)), e - 52);
      bb += f;
      if (f < 64) break;
    }
    if (bb >= wend || isnan(s)) break;
    // tile bb by the walker, from an LDS copy; len is 0 only for the one
    // tile of an empty sum, which adds nothing (and must not load)
    const int begin = bb * kSumTile, len = min(n, begin + kSumTile) - begin;
    if (len > 0) {
      // every load issued before the first store (past len: copies of the
      // last addend, never added)
#pragma unroll
      for (int r = 0; r < kSumTile / 64; r++) {
        const int i = r * 64 + lane;
        buf[i] = gload(t + begin + min(i, len - 1));
      }
      wave_lds_sync();
      s = wave_walk(buf, len, s, stats);
      wave_lds_sync();  // the next walk rewrites buf
    }
    fallback++;
    bb++;
  }
  return s;
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
    if (base + j < n) a += gload(t + base + j);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
  if (lane == 0) wd[wid] = a;
  __syncthreads();
  if (tid == 0) {
    double s = wd[0];
    for (int k = 1; k < kSumWaves; k++) s += wd[k];
    gstore(job.w.tsum + blockIdx.x, s);
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
  {
    // approximate running sum at the start of this tile -> binade guess
    double p = 0.0;
    for (int j = tid; j < b; j += kSumBlock) p += gload(w.tsum + j);
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
      const double x = idx < n ? gload(t + idx) : 0.0;
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
      store_device(w.code + b, ok ? code : -1);
      store_device(w.tot + b, total);
      store_device(w.pmin + b, mn);
      store_device(w.pmax + b, mx);
      // The last block to finish chains the tiles, on its first wave.  The
      // summary is stored and read with device-scope accesses (coherent per
      // instruction), and its stores have completed before the ticket is
      // taken, so the last block reads every summary without a device-scope
      // fence: on this GPU such a fence writes back and invalidates the L2
      // of the XCD, which slowed every kernel running beside the reduction.
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      sh.last = atomicAdd(w.ticket, 1u) == (unsigned)(nt - 1);
    }
  }
  __syncthreads();
  if (!sh.last || tid >= 64) return;
  const unsigned long long chain_t0 = w.stats ? wall_clock64() : 0;
  const int lane = tid;
  if (lane == 0) store_device(w.ticket, 0u);
  double s = 0.0;
  int fallback = 0;
  for (int b0 = 0; b0 < nt && !isnan(s); b0 += kChainWin) {
    const int wend = min(nt, b0 + kChainWin);
    wave_lds_sync();  // the previous window is no longer read
    // 256 summaries per pass, every load issued before the first store
    for (int j0 = b0; j0 < wend; j0 += 256) {
      int cd[4];
      long long tt[4], mn[4], mx[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int j = min(j0 + r * 64 + lane, wend - 1);
        cd[r] = load_device(w.code + j);
        tt[r] = load_device(w.tot + j);
        mn[r] = load_device(w.pmin + j);
        mx[r] = load_device(w.pmax + j);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int j = j0 + r * 64 + lane;
        if (j < wend) {
          sh.code[j - b0] = cd[r];
          sh.tot[j - b0] = tt[r];
          sh.pmin[j - b0] = mn[r];
          sh.pmax[j - b0] = mx[r];
        }
      }
    }
    wave_lds_sync();
    s = chain_tiles(s, b0, wend, sh.code, sh.tot, sh.pmin, sh.pmax, t, n, sh.buf, w.stats,
                    fallback);
  }
  if (lane == 0) {
    src.finish(s);
    if (w.stats) {
      atomicAdd(w.stats + kStatSums, 1ull);
      atomicAdd(w.stats + kStatTiles, (unsigned long long)nt);
      atomicAdd(w.stats + kStatFast, (unsigned long long)(nt - fallback));
      atomicAdd(w.stats + kStatChainTicks, wall_clock64() - chain_t0);
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

// ---------------------------------------------------------------------------
// The same exact sum computed by one workgroup, for kernels that run a whole
// solve in one workgroup: the tile summaries are formed one tile per wave
// (sixteen consecutive addends per lane instead of four per thread of a
// block: the same inclusive integer prefixes, so the same summaries), then
// the first wave chains them.  Every thread of the workgroup must call it.

constexpr int kWgMaxTiles = 64;  // sums of at most 64 tiles (65,536 addends)

struct WgSum {
  double buf[kSumBuf];  // a walked tile
  double tsum[kWgMaxTiles];
  long long tot[kWgMaxTiles];
  long long pmin[kWgMaxTiles];
  long long pmax[kWgMaxTiles];
  int code[kWgMaxTiles];
  double result;
};

// two's-complement addition (the integer views wrap like the block version's)
__device__ __forceinline__ long long add_wrap(long long a, long long b) {
  return (long long)((unsigned long long)a + (unsigned long long)b);
}

__device__ __forceinline__ double wg_ordered_sum(const double* __restrict__ t, int n, WgSum& w) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int nt = sum_tiles(n);
  // approximate tile sums (any order; they only steer the binade guesses)
  for (int b = wid; b < nt; b += nw) {
    double a = 0.0;
#pragma unroll 4
    for (int j = 0; j < kSumTile / 64; j++) {
      const int idx = b * kSumTile + j * 64 + lane;
      if (idx < n) a += gload(t + idx);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if (lane == 0) w.tsum[b] = a;
  }
  __syncthreads();
  // binade guess and integer summary of each tile, one wave per tile
  constexpr int kPer = kSumTile / 64;
  const double two62 = 4611686018427387904.0;  // 2^62
  for (int b = wid; b < nt; b += nw) {
    double p = 0.0;
    for (int j = lane; j < b; j += 64) p += w.tsum[j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) p += __shfl_xor(p, o, 64);
    int code = -1, scale = 0;
    if (fabs(p) >= DBL_MIN && fabs(p) <= DBL_MAX) {
      code = sign_exp(p);
      scale = 52 - ilogb(p);
    }
    long long pref[kPer];
    long long run = 0;
    bool ok = code >= 0;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const int idx = b * kSumTile + lane * kPer + j;
      const double x = idx < n ? gload(t + idx) : 0.0;
      if (ok) {
        const double q = ldexp(x, scale);
        const double r = rint(q);
        if (!(fabs(q) < two62) || fabs(q - r) == 0.5)
          ok = false;
        else
          run = add_wrap(run, (long long)r);
      }
      pref[j] = run;
    }
    const bool all_ok = __ballot(!ok) == 0ull;
    long long incl = run;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const long long y = __shfl_up(incl, o, 64);
      if (lane >= o) incl = add_wrap(incl, y);
    }
    const long long off = add_wrap(incl, -run);
    long long mn = LLONG_MAX, mx = LLONG_MIN;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const long long v = add_wrap(off, pref[j]);
      mn = min(mn, v);
      mx = max(mx, v);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      mn = min(mn, (long long)__shfl_xor(mn, o, 64));
      mx = max(mx, (long long)__shfl_xor(mx, o, 64));
    }
    const long long total = __shfl(incl, 63, 64);
    if (lane == 0) {
      w.code[b] = all_ok ? code : -1;
      w.tot[b] = total;
      w.pmin[b] = mn;
      w.pmax[b] = mx;
    }
  }
  __syncthreads();
  if (wid == 0) {
    int fallback = 0;
    const double s = chain_tiles(0.0, 0, nt, w.code, w.tot, w.pmin, w.pmax, t, n, w.buf, nullptr,
                                 fallback);
    if (lane == 0) w.result = s;
  }
  __syncthreads();
  return w.result;
}

}  // namespace cs
