// Element routines over arrays (cloth_internal_physics.h) and a device
// arithmetic self-check, exposed through the C-ABI for tests and callers
// that evaluate single elements.
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/clothsim.h"
#include "cs_device.hpp"
#include "cs_math.hpp"
#include "cs_ordered_sum.hpp"

namespace cs {

__global__ void k_elem_cst(int n, const double* C, const double* x, double lam, double mu,
                           double* dW, double* ddW) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  CstState s;
  cst_state(C + 9 * (size_t)e, x + 9 * (size_t)e, lam, mu, s);
  for (int a = 0; a < 3; a++)
    for (int i = 0; i < 3; i++) dW[9 * (size_t)e + a * 3 + i] = cst_grad(s, a, i);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) cst_hess_block(s, a, b, ddW + 81 * (size_t)e + (a * 3 + b) * 9);
}

// Hessian laid out [4][4][3][3] like the reference (w_a w_b scale on the diagonal)
__global__ void k_elem_bend(int n, const double* C, const double* x, double stiff, double* dW,
                            double* ddW) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double w[4], s, kab[16];
  bend_weights(C + 12 * (size_t)e, stiff, w, &s);
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) kab[a * 4 + b] = w[a] * w[b] * s;
  double* H = ddW + 144 * (size_t)e;
  for (int q = 0; q < 144; q++) H[q] = 0;
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++)
      for (int i = 0; i < 3; i++) H[((a * 4 + b) * 3 + i) * 3 + i] = kab[a * 4 + b];
  for (int a = 0; a < 4; a++)
    for (int i = 0; i < 3; i++) dW[12 * (size_t)e + a * 3 + i] = bend_grad(kab, x + 12 * (size_t)e, a, i);
}

struct Obst8 {
  double v[8];
};

// energy of each triangle, W = 0.5 A (E . S) (cloth_internal_physics.h:125)
__global__ void k_elem_cst_energy(int n, const double* C, const double* x, double lam, double mu,
                                  double* W) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  CstState s;
  cst_state(C + 9 * (size_t)e, x + 9 * (size_t)e, lam, mu, s);
  W[e] = 0.5 * s.area * (s.strain[0] * s.stress[0] + s.strain[1] * s.stress[1] +
                         s.strain[2] * s.stress[2]);
}

// penalty of each point from its obstacle depth and direction as a
// penetrationDepth callback reported them (cloth_internal_physics.h:254-279)
__global__ void k_elem_contact_depth(int n, const double* pd, const double* nrm, double k,
                                     double clearance, double* W, double* dW, double* ddW) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const bool on = contact_penalty(pd[e], nrm + 3 * (size_t)e, k, clearance, dW + 3 * (size_t)e,
                                  ddW + 9 * (size_t)e);
  const double d = pd[e] + clearance;
  W[e] = on ? 0.5 * k * d * d : 0.0;
}

__global__ void k_elem_contact(int n, const double* p, double k, double clearance, int kind,
                               Obst8 o, double* dW, double* ddW) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double* q = o.v;
  double pd, nrm[3];
  obstacle_depth(kind, q, p + 3 * (size_t)e, &pd, nrm);
  contact_penalty(pd, nrm, k, clearance, dW + 3 * (size_t)e, ddW + 9 * (size_t)e);
}

__global__ void k_divsqrt(int n, const double* a, const double* b, double* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  out[2 * e] = a[e] / b[e];
  out[2 * e + 1] = sqrt(a[e]);
}

namespace {

// device round trip for the element entry points
struct Scratch {
  std::vector<void*> ptrs;
  ~Scratch() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  int get(T** p, size_t n) {
    CS_HIP(hipMalloc((void**)p, sizeof(T) * (n ? n : 1)));
    ptrs.push_back(*p);
    return kOk;
  }
};

unsigned blocks(int n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace cs

using namespace cs;

extern "C" int cs_element_cst(int n, const double* C, const double* c, double lambda, double myu,
                              double* dW, double* ddW) {
  if (n < 0 || (n > 0 && (!C || !c || !dW || !ddW))) return kBadArg;
  if (n == 0) return kOk;
  Scratch s;
  double *dC, *dc, *dg, *dh;
  if (s.get(&dC, 9L * n) || s.get(&dc, 9L * n) || s.get(&dg, 9L * n) || s.get(&dh, 81L * n))
    return kHipError;
  CS_HIP(hipMemcpy(dC, C, sizeof(double) * 9 * n, hipMemcpyHostToDevice));
  CS_HIP(hipMemcpy(dc, c, sizeof(double) * 9 * n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_elem_cst, dim3(blocks(n)), dim3(256), 0, 0, n, dC, dc, lambda, myu, dg, dh);
  CS_HIP(hipGetLastError());
  CS_HIP(hipMemcpy(dW, dg, sizeof(double) * 9 * n, hipMemcpyDeviceToHost));
  CS_HIP(hipMemcpy(ddW, dh, sizeof(double) * 81 * n, hipMemcpyDeviceToHost));
  return kOk;
}

extern "C" int cs_element_cst_energy(int n, const double* C, const double* c, double lambda,
                                     double myu, double* W) {
  if (n < 0 || (n > 0 && (!C || !c || !W))) return kBadArg;
  if (n == 0) return kOk;
  Scratch s;
  double *dC, *dc, *dw;
  if (s.get(&dC, 9L * n) || s.get(&dc, 9L * n) || s.get(&dw, (long)n)) return kHipError;
  CS_HIP(hipMemcpy(dC, C, sizeof(double) * 9 * n, hipMemcpyHostToDevice));
  CS_HIP(hipMemcpy(dc, c, sizeof(double) * 9 * n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_elem_cst_energy, dim3(blocks(n)), dim3(256), 0, 0, n, dC, dc, lambda, myu, dw);
  CS_HIP(hipGetLastError());
  CS_HIP(hipMemcpy(W, dw, sizeof(double) * n, hipMemcpyDeviceToHost));
  return kOk;
}

extern "C" int cs_element_contact_depth(int n, const double* pd, const double* nrm, double stiff,
                                        double clearance, double* W, double* dW, double* ddW) {
  if (n < 0 || (n > 0 && (!pd || !nrm || !W || !dW || !ddW))) return kBadArg;
  if (n == 0) return kOk;
  Scratch s;
  double *dpd, *dn, *dw, *dg, *dh;
  if (s.get(&dpd, (long)n) || s.get(&dn, 3L * n) || s.get(&dw, (long)n) || s.get(&dg, 3L * n) ||
      s.get(&dh, 9L * n))
    return kHipError;
  CS_HIP(hipMemcpy(dpd, pd, sizeof(double) * n, hipMemcpyHostToDevice));
  CS_HIP(hipMemcpy(dn, nrm, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_elem_contact_depth, dim3(blocks(n)), dim3(256), 0, 0, n, dpd, dn, stiff,
                     clearance, dw, dg, dh);
  CS_HIP(hipGetLastError());
  CS_HIP(hipMemcpy(W, dw, sizeof(double) * n, hipMemcpyDeviceToHost));
  CS_HIP(hipMemcpy(dW, dg, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  CS_HIP(hipMemcpy(ddW, dh, sizeof(double) * 9 * n, hipMemcpyDeviceToHost));
  return kOk;
}

extern "C" int cs_element_bend(int n, const double* C, const double* c, double stiff, double* dW,
                               double* ddW) {
  if (n < 0 || (n > 0 && (!C || !c || !dW || !ddW))) return kBadArg;
  if (n == 0) return kOk;
  Scratch s;
  double *dC, *dc, *dg, *dh;
  if (s.get(&dC, 12L * n) || s.get(&dc, 12L * n) || s.get(&dg, 12L * n) || s.get(&dh, 144L * n))
    return kHipError;
  CS_HIP(hipMemcpy(dC, C, sizeof(double) * 12 * n, hipMemcpyHostToDevice));
  CS_HIP(hipMemcpy(dc, c, sizeof(double) * 12 * n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_elem_bend, dim3(blocks(n)), dim3(256), 0, 0, n, dC, dc, stiff, dg, dh);
  CS_HIP(hipGetLastError());
  CS_HIP(hipMemcpy(dW, dg, sizeof(double) * 12 * n, hipMemcpyDeviceToHost));
  CS_HIP(hipMemcpy(ddW, dh, sizeof(double) * 144 * n, hipMemcpyDeviceToHost));
  return kOk;
}

extern "C" int cs_element_contact(int n, const double* p, double stiff, double clearance,
                                  int kind, const double* obstacle, double* dW, double* ddW) {
  if (n < 0 || (n > 0 && (!p || !obstacle || !dW || !ddW)) || kind < 0 || kind > CS_OBST_MAX)
    return kBadArg;
  if (n == 0) return kOk;
  Scratch s;
  double *dp, *dg, *dh;
  if (s.get(&dp, 3L * n) || s.get(&dg, 3L * n) || s.get(&dh, 9L * n)) return kHipError;
  CS_HIP(hipMemcpy(dp, p, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  Obst8 o8;
  for (int k = 0; k < 8; k++) o8.v[k] = obstacle[k];
  hipLaunchKernelGGL(k_elem_contact, dim3(blocks(n)), dim3(256), 0, 0, n, dp, stiff, clearance,
                     kind, o8, dg, dh);
  CS_HIP(hipGetLastError());
  CS_HIP(hipMemcpy(dW, dg, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  CS_HIP(hipMemcpy(ddW, dh, sizeof(double) * 9 * n, hipMemcpyDeviceToHost));
  return kOk;
}

extern "C" int cs_selftest_divsqrt(int n, const double* a, const double* b, double* out) {
  if (n < 0 || (n > 0 && (!a || !b || !out))) return kBadArg;
  if (n == 0) return kOk;
  Scratch s;
  double *da, *db, *dout;
  if (s.get(&da, n) || s.get(&db, n) || s.get(&dout, 2L * n)) return kHipError;
  CS_HIP(hipMemcpy(da, a, sizeof(double) * n, hipMemcpyHostToDevice));
  CS_HIP(hipMemcpy(db, b, sizeof(double) * n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_divsqrt, dim3(blocks(n)), dim3(256), 0, 0, n, da, db, dout);
  CS_HIP(hipGetLastError());
  CS_HIP(hipMemcpy(out, dout, sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
  return kOk;
}

extern "C" int cs_selftest_ordered_sum(int n, const double* t, double* out) {
  if (n < 0 || (n > 0 && !t) || !out) return kBadArg;
  Scratch s;
  double *dt, *dout;
  SumScratch w{};
  const int nt = sum_tiles(n);
  if (s.get(&dt, n) || s.get(&dout, 1) || s.get(&w.tsum, nt) || s.get(&w.code, nt) ||
      s.get(&w.tot, nt) || s.get(&w.pmin, nt) || s.get(&w.pmax, nt) || s.get(&w.ticket, 1))
    return kHipError;
  CS_HIP(hipMemset(w.ticket, 0, sizeof(unsigned int)));
  if (n) CS_HIP(hipMemcpy(dt, t, sizeof(double) * n, hipMemcpyHostToDevice));
  launch_ordered_sum(SumStore{SumJob{dt, n, w, nullptr}, dout}, n, 1, 0);
  CS_HIP(hipGetLastError());
  CS_HIP(hipMemcpy(out, dout, sizeof(double), hipMemcpyDeviceToHost));
  return kOk;
}

namespace cs {

// plain streaming copy, 16 bytes per thread and iteration
__global__ void k_stream_copy(long long n2, const double2* __restrict__ x, double2* __restrict__ y) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride)
    y[i] = x[i];
}

// Streaming reads by a few workgroups of 512 threads (one per CU): the
// memory ceiling of a kernel that runs in one workgroup, as the ILU sweeps
// do.  Sixteen 16-byte loads per thread are in flight per pass; n2 is a
// multiple of 512 * 16 * blocks, each block reads its own contiguous part.
constexpr int kProbeThreads = 512, kProbeUnroll = 16;
__global__ __launch_bounds__(kProbeThreads) void k_read_probe(long long per_block,
                                                              const double2* __restrict__ x,
                                                              double* out) {
  const double2* a = x + (long long)blockIdx.x * per_block;
  double acc = 0.0;
  for (long long i = threadIdx.x; i < per_block; i += kProbeThreads * kProbeUnroll) {
    double2 v[kProbeUnroll];
#pragma unroll
    for (int u = 0; u < kProbeUnroll; u++) v[u] = a[i + u * kProbeThreads];
#pragma unroll
    for (int u = 0; u < kProbeUnroll; u++) acc += v[u].x + v[u].y;
  }
  if (acc == -1.0) out[blockIdx.x] = acc;  // keeps the loads; never taken for zeros
}

}  // namespace cs

extern "C" int cs_read_probe(long long n, int reps, int blocks, double* gbs) {
  using namespace cs;
  const long long unit = 2LL * kProbeThreads * kProbeUnroll * blocks;  // doubles per pass
  if (n <= 0 || n % unit || reps <= 0 || blocks <= 0 || blocks > 4096 || !gbs) return kBadArg;
  Scratch s;
  double *x, *out;
  if (s.get(&x, (size_t)n) || s.get(&out, (size_t)blocks)) return kHipError;
  CS_HIP(hipMemset(x, 0, sizeof(double) * n));
  const long long per_block = n / 2 / blocks;  // double2 per block
  hipEvent_t e0, e1;
  CS_HIP(hipEventCreate(&e0));
  CS_HIP(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_read_probe, dim3(blocks), dim3(kProbeThreads), 0, 0, per_block,
                     (const double2*)x, out);
  CS_HIP(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(k_read_probe, dim3(blocks), dim3(kProbeThreads), 0, 0, per_block,
                       (const double2*)x, out);
  CS_HIP(hipEventRecord(e1, 0));
  CS_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  CS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  CS_HIP(hipGetLastError());
  *gbs = 8.0 * (double)n * reps / (ms * 1e-3) / 1e9;
  return kOk;
}

namespace cs {

// Vector-load issue cost of one workgroup shaped like the ILU sweep: `waves`
// waves each issue kIssueLoads 16-byte loads per step into one of three
// register sets in rotation (two steps of loads in flight), quads of lanes
// reading a row's three components (lane 3 repeats lane 2), an LDS-only
// barrier per step.  Step s reads block s mod kIssueBlocks of 8 KiB: 64 MiB
// in all, a factor-sized working set that the first launch brings into the
// device caches.
constexpr int kIssueLoads = 12, kIssueBlocks = 8192, kIssueBlock = 512;  // double2 per block

struct IssueSet {
  double2 v[kIssueLoads];
};

__device__ __forceinline__ void issue_load(IssueSet& s, const double2* x, unsigned step,
                                           unsigned off) {
#pragma unroll
  for (int k = 0; k < kIssueLoads; k++) {
    const unsigned blk = (step * kIssueLoads + k) & (kIssueBlocks - 1);
    s.v[k] = x[blk * kIssueBlock + off];
  }
}

__device__ __forceinline__ double issue_use(const IssueSet& s) {
  double a = 0.0;
#pragma unroll
  for (int k = 0; k < kIssueLoads; k++) a += s.v[k].x + s.v[k].y;
  return a;
}

__device__ __forceinline__ void issue_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// steps: a multiple of three
__global__ __launch_bounds__(512) void k_issue_probe(const double2* __restrict__ x, int steps,
                                                     double* out) {
  const int q = threadIdx.x >> 2, l = threadIdx.x & 3;
  const unsigned off = (unsigned)(3 * q + (l < 3 ? l : 2));  // < 3 * 128 <= kIssueBlock
  IssueSet a, b, c;
  // the scheduling barriers keep the sets' loads in issue order (a, b, c
  // before the loop as in it, so the loop needs only the oldest set) and each
  // set's loads between its use and the step barrier, as in the sweep
  issue_load(a, x, 0, off);
  __builtin_amdgcn_sched_barrier(0);
  issue_load(b, x, 1, off);
  __builtin_amdgcn_sched_barrier(0);
  issue_load(c, x, 2, off);
  __builtin_amdgcn_sched_barrier(0);
  double acc = 0.0;
  for (int st = 0; st < steps; st += 3) {
    acc += issue_use(a);
    __builtin_amdgcn_sched_barrier(0);
    issue_load(a, x, st + 3, off);
    __builtin_amdgcn_sched_barrier(0);
    issue_barrier();
    acc += issue_use(b);
    __builtin_amdgcn_sched_barrier(0);
    issue_load(b, x, st + 4, off);
    __builtin_amdgcn_sched_barrier(0);
    issue_barrier();
    acc += issue_use(c);
    __builtin_amdgcn_sched_barrier(0);
    issue_load(c, x, st + 5, off);
    __builtin_amdgcn_sched_barrier(0);
    issue_barrier();
  }
  acc += issue_use(a) + issue_use(b) + issue_use(c);
  if (acc == -1.0) out[threadIdx.x] = acc;  // keeps the loads; never taken for zeros
}

}  // namespace cs

extern "C" int cs_issue_probe(int waves, int steps, double* ns_per_load) {
  using namespace cs;
  if (waves < 1 || waves > 8 || steps < 3 || steps % 3 || !ns_per_load) return kBadArg;
  Scratch s;
  double *x, *out;
  const size_t n = 2 * (size_t)kIssueBlocks * kIssueBlock;  // doubles
  if (s.get(&x, n) || s.get(&out, 512)) return kHipError;
  CS_HIP(hipMemset(x, 0, sizeof(double) * n));
  hipEvent_t e0, e1;
  CS_HIP(hipEventCreate(&e0));
  CS_HIP(hipEventCreate(&e1));
  const dim3 tb(64 * waves);
  hipLaunchKernelGGL(k_issue_probe, dim3(1), tb, 0, 0, (const double2*)x, steps, out);
  CS_HIP(hipEventRecord(e0, 0));
  hipLaunchKernelGGL(k_issue_probe, dim3(1), tb, 0, 0, (const double2*)x, steps, out);
  CS_HIP(hipEventRecord(e1, 0));
  CS_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  CS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  CS_HIP(hipGetLastError());
  *ns_per_load = (double)ms * 1e6 / ((double)steps * waves * kIssueLoads);
  return kOk;
}

extern "C" int cs_bandwidth_probe(long long n, int reps, double* gbs) {
  using namespace cs;
  if (n <= 0 || (n & 1) || reps <= 0 || !gbs) return kBadArg;
  Scratch s;
  double *x, *y;
  if (s.get(&x, (size_t)n) || s.get(&y, (size_t)n)) return kHipError;
  CS_HIP(hipMemset(x, 0, sizeof(double) * n));
  const long long n2 = n / 2;
  const int grid = 256 * 8;  // 8 blocks per CU, grid-stride
  hipEvent_t e0, e1;
  CS_HIP(hipEventCreate(&e0));
  CS_HIP(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, 0, n2, (const double2*)x, (double2*)y);
  CS_HIP(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, 0, n2, (const double2*)x,
                       (double2*)y);
  CS_HIP(hipEventRecord(e1, 0));
  CS_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  CS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  CS_HIP(hipGetLastError());
  *gbs = 2.0 * 8.0 * (double)n * reps / (ms * 1e-3) / 1e9;
  return kOk;
}
