// Implicit internal solve of the cloth step on the device:
//   element kernels -> gather assembly of the backward-Euler system
//   -> block ILU(0) -> ILU-preconditioned CG
// (solve_internal_sparse.h:255-326, ilu_sparse.cpp, matrix_square_sparse.cpp).
//
// Bit-exactness: every value is produced by one thread in the reference's
// operation order.  Assembly gathers each block's contributions in element
// order (no atomics); the ILU factor and the triangular sweeps run one row
// per thread over a precomputed level schedule; the PCG reductions are summed
// left to right exactly as the reference's InnerProduct.
#include <hip/hip_runtime.h>

#include "cs_device.hpp"
#include "cs_math.hpp"
#include "cs_ordered_sum.hpp"
#include "cs_solve.hpp"

namespace cs {

namespace {
constexpr int kTB = 256;
inline unsigned grid_for(long n, int tb = kTB) { return (unsigned)((n + tb - 1) / tb); }
}  // namespace

// ---------------------------------------------------------------------------
// element stage

// constant bending Hessian scalars w_a w_b scale per stencil
__global__ void k_bend_setup(DevCloth c, double stiff) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= c.nq) return;
  double X[12];
  for (int a = 0; a < 4; a++) {
    const int v = c.quad[4 * q + a];
    for (int k = 0; k < 3; k++) X[a * 3 + k] = c.xyz0[3 * v + k];
  }
  double w[4], s;
  bend_weights(X, stiff, w, &s);
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) c.kq[q * 16 + a * 4 + b] = w[a] * w[b] * s;
}

// membrane element: 81 Hessian values (block (a,b) at (a*3+b)*9) + 9 gradient values
template <class Sel>
__global__ void k_tri_elem(Sel sel, int with_hess) {
  const DevCloth c = sel.cloth();
  const double lam = sel.params().lam, mu = sel.params().mu;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.nt) return;
  double X[9], x[9];
  for (int a = 0; a < 3; a++) {
    const int v = c.tri[3 * t + a];
    for (int k = 0; k < 3; k++) {
      X[a * 3 + k] = c.xyz0[3 * v + k];
      x[a * 3 + k] = c.x[3 * v + k];
    }
  }
  CstState s;
  cst_state(X, x, lam, mu, s);
  double* o = c.etri + (size_t)t * 90;
  if (with_hess) {
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) cst_hess_block(s, a, b, o + (a * 3 + b) * 9);
  }
  for (int a = 0; a < 3; a++)
    for (int i = 0; i < 3; i++) o[81 + a * 3 + i] = cst_grad(s, a, i);
}

template <class Sel>
__global__ void k_quad_grad(Sel sel) {
  const DevCloth c = sel.cloth();
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= c.nq) return;
  double x[12];
  for (int a = 0; a < 4; a++) {
    const int v = c.quad[4 * q + a];
    for (int k = 0; k < 3; k++) x[a * 3 + k] = c.x[3 * v + k];
  }
  const double* kab = c.kq + (size_t)q * 16;
  for (int a = 0; a < 4; a++)
    for (int i = 0; i < 3; i++) c.eq[(size_t)q * 12 + a * 3 + i] = bend_grad(kab, x, a, i);
}

// ---------------------------------------------------------------------------
// gather assembly

__device__ __forceinline__ void add_hess_src(const DevCloth& c, int src, double acc[9]) {
  if (src >= 0) {
    const double* e = c.etri + (size_t)(src / 9) * 90 + (src % 9) * 9;
#pragma unroll
    for (int q = 0; q < 9; q++) acc[q] += e[q];
  } else {
    // bending block: kab on the diagonal, exact zeros elsewhere
    const double kv = c.kq[-src - 1];
#pragma unroll
    for (int q = 0; q < 9; q++) acc[q] += (q == 0 || q == 4 || q == 8) ? kv : 0.0;
  }
}

// off-diagonal blocks; rows and columns of fixed vertices are zeroed
// (matrix_square_sparse.cpp:272-302)
template <class Sel>
__global__ void k_asm_crs(Sel sel) {
  const DevCloth c = sel.cloth();
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= c.nnz) return;
  const int i = c.rowof[k], j = c.colind[k];
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; q++) acc[q] = 0.0;
  if (c.bc[i] == 0 && c.bc[j] == 0) {
    for (int s = c.crs_src_ptr[k]; s < c.crs_src_ptr[k + 1]; s++) add_hess_src(c, c.crs_src[s], acc);
  }
  double* o = c.crs + (size_t)k * 9;
#pragma unroll
  for (int q = 0; q < 9; q++) o[q] = acc[q];
}

// diagonal blocks + right-hand side (solve_internal_sparse.h:285-313)
template <class Sel>
__global__ void k_asm_dia_rhs(Sel sel) {
  const DevCloth c = sel.cloth();
  const Params P = sel.params();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.np) return;
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; q++) acc[q] = 0.0;
  for (int s = c.dia_src_ptr[i]; s < c.dia_src_ptr[i + 1]; s++) add_hess_src(c, c.dia_src[s], acc);
  double pd, n[3];
  if (c.host_pd) {
    pd = c.host_pd[i];
    n[0] = c.host_n[3 * i];
    n[1] = c.host_n[3 * i + 1];
    n[2] = c.host_n[3 * i + 2];
  } else {
    obstacle_depth(P.obstacle_kind, P.obstacle, c.x + 3 * i, &pd, n);
  }
  double cg[3], ch[9];
  contact_penalty(pd, n, P.k_contact, P.clearance, cg, ch);
#pragma unroll
  for (int q = 0; q < 9; q++) acc[q] += ch[q];
  const double md = P.mass / (P.dt * P.dt);
  acc[0] += md;
  acc[4] += md;
  acc[8] += md;
  const bool fixed = c.bc[i] != 0;
  double* o = c.dia + (size_t)i * 9;
#pragma unroll
  for (int q = 0; q < 9; q++) o[q] = fixed ? ((q == 0 || q == 4 || q == 8) ? 1.0 : 0.0) : acc[q];
  double g[3] = {0.0, 0.0, 0.0};
  for (int s = c.grad_src_ptr[i]; s < c.grad_src_ptr[i + 1]; s++) {
    const int src = c.grad_src[s];
    const double* e = (src >= 0) ? c.etri + (size_t)(src / 3) * 90 + 81 + (src % 3) * 3
                                 : c.eq + (size_t)(-src - 1) * 3;
    g[0] += e[0];
    g[1] += e[1];
    g[2] += e[2];
  }
  for (int k = 0; k < 3; k++) g[k] += cg[k];
  for (int k = 0; k < 3; k++) g[k] -= P.mass * P.grav[k];
  for (int k = 0; k < 3; k++)
    c.b[3 * i + k] = fixed ? 0.0 : -g[k] + P.mass * c.v[3 * i + k] / P.dt;
}

// ---------------------------------------------------------------------------
// block ILU(0), IKJ, one row per thread over the lower-part level schedule
// (ilu_sparse.cpp:387-401, 507-569).  Rows after `stop` keep the matrix
// values, row `stop` keeps its unscaled upper part: the reference returns
// right after its 11th singular pivot.

__device__ __forceinline__ int find_in_row(const int* colind, int b, int e, int j) {
  while (b < e) {
    const int m = (b + e) >> 1;
    const int v = colind[m];
    if (v == j) return m;
    if (v < j)
      b = m + 1;
    else
      e = m;
  }
  return -1;
}

__device__ void ilu_row(const DevCloth& c, int i, int stop, int* ising) {
  const int rb = c.rowptr[i], re = c.rowptr[i + 1], di = c.diaind[i];
  double* F = c.fcrs;
  for (int ik = rb; ik < di; ik++) {
    const int kr = c.colind[ik];
    const double* l = F + (size_t)ik * 9;
    for (int kj = c.diaind[kr]; kj < c.rowptr[kr + 1]; kj++) {
      const int j = c.colind[kj];
      const double* u = F + (size_t)kj * 9;
      double* a;
      if (j != i) {
        const int pos = find_in_row(c.colind, rb, re, j);
        if (pos < 0) continue;
        a = F + (size_t)pos * 9;
      } else {
        a = c.fdia + (size_t)i * 9;
      }
      for (int r = 0; r < 3; r++) {
        const double l0 = l[r * 3 + 0], l1 = l[r * 3 + 1], l2 = l[r * 3 + 2];
        a[r * 3 + 0] -= l0 * u[0] + l1 * u[3] + l2 * u[6];
        a[r * 3 + 1] -= l0 * u[1] + l1 * u[4] + l2 * u[7];
        a[r * 3 + 2] -= l0 * u[2] + l1 * u[5] + l2 * u[8];
      }
    }
  }
  double* d = c.fdia + (size_t)i * 9;
  const double det = +d[0] * d[4] * d[8] + d[3] * d[7] * d[2] + d[6] * d[1] * d[5] -
                     d[0] * d[7] * d[5] - d[6] * d[4] * d[2] - d[3] * d[1] * d[8];
  if (fabs(det) > 1.0e-30) {
    const double id = 1.0 / det;
    double t[9];
    for (int q = 0; q < 9; q++) t[q] = d[q];
    d[0] = id * (t[4] * t[8] - t[5] * t[7]);
    d[1] = id * (t[2] * t[7] - t[1] * t[8]);
    d[2] = id * (t[1] * t[5] - t[2] * t[4]);
    d[3] = id * (t[5] * t[6] - t[3] * t[8]);
    d[4] = id * (t[0] * t[8] - t[2] * t[6]);
    d[5] = id * (t[2] * t[3] - t[0] * t[5]);
    d[6] = id * (t[3] * t[7] - t[4] * t[6]);
    d[7] = id * (t[1] * t[6] - t[0] * t[7]);
    d[8] = id * (t[0] * t[4] - t[1] * t[3]);
  } else if (ising) {
    ising[i] = 1;
    atomicAdd(&c.flags[kNsing], 1);
  }
  if (i == stop) return;
  for (int ij = di; ij < re; ij++) {
    double* u = F + (size_t)ij * 9;
    double t[9];
    for (int q = 0; q < 9; q++) t[q] = u[q];
    for (int r = 0; r < 3; r++) {
      u[r * 3 + 0] = d[r * 3 + 0] * t[0] + d[r * 3 + 1] * t[3] + d[r * 3 + 2] * t[6];
      u[r * 3 + 1] = d[r * 3 + 0] * t[1] + d[r * 3 + 1] * t[4] + d[r * 3 + 2] * t[7];
      u[r * 3 + 2] = d[r * 3 + 0] * t[2] + d[r * 3 + 1] * t[5] + d[r * 3 + 2] * t[8];
    }
  }
}

// The factorisation for meshes with a row too wide for the LDS-staged kernel
// (k_ilu_factor_staged, which computes the same values faster).
// copy_in: also copy A into the factor storage and clear the pivot flags
// first (batched launches; the single-cloth path uses device copies)
template <class Sel>
__global__ __launch_bounds__(1024) void k_ilu_factor(Sel sel, int copy_in) {
  const DevCloth c = sel.cloth();
  if (copy_in) {
    for (long q = threadIdx.x; q < 9L * c.np; q += blockDim.x) c.fdia[q] = c.dia[q];
    for (long q = threadIdx.x; q < 9L * c.nnz; q += blockDim.x) c.fcrs[q] = c.crs[q];
    for (int q = threadIdx.x; q < c.np; q += blockDim.x) c.ising[q] = 0;
    if (threadIdx.x == 0) c.flags[kNsing] = 0;
    __syncthreads();
  }
  for (int l = 0; l < c.nlf; l++) {
    const int b = c.lf_ptr[l], e = c.lf_ptr[l + 1];
    for (int t = b + (int)threadIdx.x; t < e; t += blockDim.x) ilu_row(c, c.lf_rows[t], c.np, c.ising);
    __syncthreads();
  }
}

// Block ILU(0), one row per group of kFgLanes lanes with the row's blocks in
// LDS.  The arithmetic is ilu_row's, operation for operation: the updates of
// row i come from its precomputed list (the reference's loop order, so no
// column search).  The updates that use the same lower block L_ik target
// distinct blocks of row i, so they run at once, one per lane; the lower
// blocks are taken in ascending k, which gives every target block its
// updates in the reference's order and makes each L_ik final before it is
// used (it is itself a target only of updates with smaller k).  Each lane
// loads its update entries' upper blocks (rows of earlier levels, final)
// together with the row.  LDS slot per group: the row's blocks from
// rowptr[i] at their index, the diagonal at nb - 1.
constexpr int kStageDoubles = 15360;  // 120 KiB of LDS for the row slots
constexpr int kFgLanes = 16;          // lanes per row
constexpr int kFgLayers = 2;          // update entries each lane keeps in registers

int ilu_stage_blocks(int widest_row) {
  const int nb = widest_row + 1;
  // at least one wave of rows, and every block index fits the 8-bit codes
  return (widest_row < kFuDiag && kStageDoubles / (9 * nb) >= 64 / kFgLanes) ? nb : 0;
}

static int ilu_stage_threads(int nb) {
  return std::min(1024, kStageDoubles / (9 * nb) * kFgLanes / 64 * 64);
}

// LDS written by one lane of a wave is read by another after this
__device__ __forceinline__ void fg_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A_ij -= L_ik U'_kj for one update entry (code: ik - rowptr[i] | target << 8)
__device__ __forceinline__ void fg_apply(double* sl, int nb, int code, const double* u) {
  const double* l = sl + (code & 0xff) * 9;
  const int tg = (code >> 8) & 0xff;
  double* a = sl + (tg == kFuDiag ? nb - 1 : tg) * 9;
  for (int r = 0; r < 3; r++) {
    const double l0 = l[r * 3 + 0], l1 = l[r * 3 + 1], l2 = l[r * 3 + 2];
    a[r * 3 + 0] -= l0 * u[0] + l1 * u[3] + l2 * u[6];
    a[r * 3 + 1] -= l0 * u[1] + l1 * u[4] + l2 * u[7];
    a[r * 3 + 2] -= l0 * u[2] + l1 * u[5] + l2 * u[8];
  }
}

// row i by lane t of its group (every lane of the group calls this)
__device__ void ilu_row_grouped(const DevCloth& c, int i, int* ising, double* sl, int nb, int t) {
  const int rb = gptr(c.rowptr)[i], re = gptr(c.rowptr)[i + 1], di = gptr(c.diaind)[i];
  const auto F = gptr(c.fcrs);
  const auto fdia = gptr(c.fdia);
  const auto fu_code = gptr(c.fu_code);
  const auto fu_kj = gptr(c.fu_kj);
  const int u0 = gptr(c.fu_ptr)[i], u1 = gptr(c.fu_ptr)[i + 1];
  int code[kFgLayers];
  double ub[kFgLayers][9];
#pragma unroll
  for (int m = 0; m < kFgLayers; m++) {
    const int e = u0 + t + m * kFgLanes;
    code[m] = -1;
    if (e < u1) {
      code[m] = fu_code[e];
      const size_t kj = (size_t)fu_kj[e];
#pragma unroll
      for (int q = 0; q < 9; q++) ub[m][q] = F[kj * 9 + q];
    }
  }
  const int deg = re - rb;
  for (int q = t; q < 9 * deg; q += kFgLanes) sl[q] = F[(size_t)rb * 9 + q];
  double* d = sl + (nb - 1) * 9;
  for (int q = t; q < 9; q += kFgLanes) d[q] = fdia[(size_t)i * 9 + q];
  fg_wave_sync();
  for (int k = 0; k < di - rb; k++) {
#pragma unroll
    for (int m = 0; m < kFgLayers; m++)
      if (code[m] >= 0 && (code[m] & 0xff) == k) fg_apply(sl, nb, code[m], ub[m]);
    // entries beyond the registers (rows with more than kFgLayers * kFgLanes updates)
    for (int e = u0 + t + kFgLayers * kFgLanes; e < u1; e += kFgLanes) {
      const int cd = fu_code[e];
      if ((cd & 0xff) != k) continue;
      const size_t kj = (size_t)fu_kj[e];
      double u[9];
      for (int q = 0; q < 9; q++) u[q] = F[kj * 9 + q];
      fg_apply(sl, nb, cd, u);
    }
    fg_wave_sync();
  }
  if (t == 0) {
    const double det = +d[0] * d[4] * d[8] + d[3] * d[7] * d[2] + d[6] * d[1] * d[5] -
                       d[0] * d[7] * d[5] - d[6] * d[4] * d[2] - d[3] * d[1] * d[8];
    if (fabs(det) > 1.0e-30) {
      const double id = 1.0 / det;
      double w[9];
      for (int q = 0; q < 9; q++) w[q] = d[q];
      d[0] = id * (w[4] * w[8] - w[5] * w[7]);
      d[1] = id * (w[2] * w[7] - w[1] * w[8]);
      d[2] = id * (w[1] * w[5] - w[2] * w[4]);
      d[3] = id * (w[5] * w[6] - w[3] * w[8]);
      d[4] = id * (w[0] * w[8] - w[2] * w[6]);
      d[5] = id * (w[2] * w[3] - w[0] * w[5]);
      d[6] = id * (w[3] * w[7] - w[4] * w[6]);
      d[7] = id * (w[1] * w[6] - w[0] * w[7]);
      d[8] = id * (w[0] * w[4] - w[1] * w[3]);
    } else if (ising) {
      ising[i] = 1;
      atomicAdd(&c.flags[kNsing], 1);
    }
  }
  fg_wave_sync();
  for (int q = t; q < 9; q += kFgLanes) fdia[(size_t)i * 9 + q] = d[q];
  for (int q = t; q < 9 * (di - rb); q += kFgLanes) F[(size_t)rb * 9 + q] = sl[q];  // lower blocks
  for (int q = t; q < 3 * (re - di); q += kFgLanes) {  // upper blocks: D^-1 U, one block row each
    const int b = q / 3, r = q - 3 * b;
    const double* w = sl + (di - rb + b) * 9;
    const auto u = F + (size_t)(di + b) * 9;
    u[r * 3 + 0] = d[r * 3 + 0] * w[0] + d[r * 3 + 1] * w[3] + d[r * 3 + 2] * w[6];
    u[r * 3 + 1] = d[r * 3 + 0] * w[1] + d[r * 3 + 1] * w[4] + d[r * 3 + 2] * w[7];
    u[r * 3 + 2] = d[r * 3 + 0] * w[2] + d[r * 3 + 1] * w[5] + d[r * 3 + 2] * w[8];
  }
  fg_wave_sync();  // the slot is reused by the group's next row
}

// copy_in as k_ilu_factor; nb: blocks per row slot (at least every row's
// blocks plus one), blockDim.x / kFgLanes * nb * 9 <= kStageDoubles
template <class Sel>
__global__ __launch_bounds__(1024) void k_ilu_factor_staged(Sel sel, int copy_in, int nb) {
  const DevCloth c = sel.cloth();
  __shared__ double stage[kStageDoubles];
  if (copy_in) {
    for (long q = threadIdx.x; q < 9L * c.np; q += blockDim.x) c.fdia[q] = c.dia[q];
    for (long q = threadIdx.x; q < 9L * c.nnz; q += blockDim.x) c.fcrs[q] = c.crs[q];
    for (int q = threadIdx.x; q < c.np; q += blockDim.x) c.ising[q] = 0;
    if (threadIdx.x == 0) c.flags[kNsing] = 0;
    __syncthreads();
  }
  const int grp = threadIdx.x / kFgLanes, t = threadIdx.x % kFgLanes;
  const int ngrp = blockDim.x / kFgLanes;
  double* sl = stage + (size_t)grp * nb * 9;
  for (int l = 0; l < c.nlf; l++) {
    const int b = c.lf_ptr[l], e = c.lf_ptr[l + 1];
    for (int r = b + grp; r < e; r += ngrp) ilu_row_grouped(c, c.lf_rows[r], c.ising, sl, nb, t);
    __syncthreads();
  }
}

// re-run of the factorisation when more than 10 pivots were singular: stop at
// the 11th singular row in row order, leave later rows untouched
template <class Sel>
__global__ __launch_bounds__(1024) void k_ilu_redo(Sel sel) {
  const DevCloth c = sel.cloth();
  int* ising = c.ising;
  if (c.flags[kNsing] <= 10) return;
  __shared__ int stop;
  if (threadIdx.x == 0) {
    int cnt = 0, s = c.np;
    for (int i = 0; i < c.np; i++)
      if (ising[i] && ++cnt == 11) {
        s = i;
        break;
      }
    stop = s;
    c.flags[kFirstSing11] = s;
  }
  __syncthreads();
  const int s = stop;
  for (long q = threadIdx.x; q < 9L * c.np; q += blockDim.x) c.fdia[q] = c.dia[q];
  for (long q = threadIdx.x; q < 9L * c.nnz; q += blockDim.x) c.fcrs[q] = c.crs[q];
  __syncthreads();
  for (int l = 0; l < c.nlf; l++) {
    const int b = c.lf_ptr[l], e = c.lf_ptr[l + 1];
    for (int t = b + (int)threadIdx.x; t < e; t += blockDim.x) {
      const int i = c.lf_rows[t];
      if (i <= s) ilu_row(c, i, s, nullptr);
    }
    __syncthreads();
  }
}

// forward then backward substitution, dst = (LU)^-1 src; optionally
// prod = src .* dst for the next reduction (ilu_sparse.cpp:185-215, 295-324)
// src == nullptr: the PCG operands (r in, z out, r.*z into prod)
template <class Sel>
__global__ __launch_bounds__(1024) void k_ilu_apply(Sel sel, const double* src, double* dst,
                                                    double* prod) {
  const DevCloth c = sel.cloth();
  if (c.flags[kSkip]) return;
  if (!src) {
    src = c.r;
    dst = c.z;
    prod = c.prod;
  }
  for (int l = 0; l < c.nlf; l++) {
    const int b = c.lf_ptr[l], e = c.lf_ptr[l + 1];
    for (int t = b + (int)threadIdx.x; t < e; t += blockDim.x) {
      const int i = c.lf_rows[t];
      double t0 = src[3 * i], t1 = src[3 * i + 1], t2 = src[3 * i + 2];
      for (int k = c.rowptr[i]; k < c.diaind[i]; k++) {
        const int j = c.colind[k];
        const double* a = c.fcrs + (size_t)k * 9;
        const double y0 = dst[3 * j], y1 = dst[3 * j + 1], y2 = dst[3 * j + 2];
        t0 -= a[0] * y0 + a[1] * y1 + a[2] * y2;
        t1 -= a[3] * y0 + a[4] * y1 + a[5] * y2;
        t2 -= a[6] * y0 + a[7] * y1 + a[8] * y2;
      }
      const double* d = c.fdia + (size_t)i * 9;
      dst[3 * i] = d[0] * t0 + d[1] * t1 + d[2] * t2;
      dst[3 * i + 1] = d[3] * t0 + d[4] * t1 + d[5] * t2;
      dst[3 * i + 2] = d[6] * t0 + d[7] * t1 + d[8] * t2;
    }
    __syncthreads();
  }
  for (int l = 0; l < c.nlb; l++) {
    const int b = c.lb_ptr[l], e = c.lb_ptr[l + 1];
    for (int t = b + (int)threadIdx.x; t < e; t += blockDim.x) {
      const int i = c.lb_rows[t];
      double t0 = dst[3 * i], t1 = dst[3 * i + 1], t2 = dst[3 * i + 2];
      for (int k = c.diaind[i]; k < c.rowptr[i + 1]; k++) {
        const int j = c.colind[k];
        const double* a = c.fcrs + (size_t)k * 9;
        const double y0 = dst[3 * j], y1 = dst[3 * j + 1], y2 = dst[3 * j + 2];
        t0 -= a[0] * y0 + a[1] * y1 + a[2] * y2;
        t1 -= a[3] * y0 + a[4] * y1 + a[5] * y2;
        t2 -= a[6] * y0 + a[7] * y1 + a[8] * y2;
      }
      dst[3 * i] = t0;
      dst[3 * i + 1] = t1;
      dst[3 * i + 2] = t2;
    }
    __syncthreads();
  }
  if (prod)
    for (int e = threadIdx.x; e < 3 * c.np; e += blockDim.x) prod[e] = src[e] * dst[e];
}

// ---------------------------------------------------------------------------
// packed sweeps.  The rows of a sub-level (at most kQuadRows rows of one
// level) are consecutive "slots".  Each row is handled by four lanes: lane r
// (0..2) computes output component r, lane 3 idles.  The packed factor gives
// every lane its own entries as 16-byte pairs: value v = k * 3 + j (entry
// (r, j) of block k) at [(((v / 2) * np + slot) * 3 + r) * 2 + v % 2] (lane 3
// reads lane 2's), so one load instruction of a wave reads 768 contiguous
// bytes and a row component costs ceil(3K / 2) block loads.  The load
// instructions issued per sub-level, not HBM, bound the sweep.  The forward
// sweep reads its right-hand side in forward-slot order (inside the packed
// D^-1) and writes its result in backward-slot order (sb), so that no load of
// either sweep depends on another load.

// 16-byte pairs of a slot's values per lane (see above)
__host__ __device__ constexpr int pack_pairs(int K) { return (3 * K + 1) / 2; }
__device__ __forceinline__ size_t pack_index(int v, int sl, int r, int np) {
  return ((((size_t)(v >> 1) * np + sl) * 3 + r) << 1) + (v & 1);
}
// mixed precision: the same values rounded to FP32, four per 16 bytes, in the
// same buffer (it holds the FP64 pairs otherwise): value v of (slot, r) at
// float index [(((v / 4) * np + slot) * 3 + r) * 4 + v % 4]
__host__ __device__ constexpr int pack_quads(int K) { return (3 * K + 3) / 4; }
__device__ __forceinline__ size_t pack_index32(int v, int sl, int r, int np) {
  return ((((size_t)(v >> 2) * np + sl) * 3 + r) << 2) + (v & 3);
}

template <class Sel>
__global__ void k_pack_factor(Sel sel) {
  const DevCloth c = sel.cloth();
  const int sl = blockIdx.x * blockDim.x + threadIdx.x;
  const int np = c.np;
  if (sl >= np || c.kl == 0) return;
  {
    const int i = c.f_rci[4 * sl], cnt = c.f_rci[4 * sl + 1];
    const double* src = c.fcrs + (size_t)c.rowptr[i] * 9;
    if (c.precision)
      for (int v = 0; v < 4 * pack_quads(c.kpack); v++)
        for (int r = 0; r < 3; r++) {
          const int k = v / 3, j = v % 3;
          reinterpret_cast<float*>(c.f_blk)[pack_index32(v, sl, r, np)] =
              (k < cnt && k < c.kpack) ? (float)src[k * 9 + r * 3 + j] : 0.0f;
        }
    else
      for (int v = 0; v < 2 * pack_pairs(c.kpack); v++)
        for (int r = 0; r < 3; r++) {
          const int k = v / 3, j = v % 3;
          c.f_blk[pack_index(v, sl, r, np)] = (k < cnt && k < c.kpack) ? src[k * 9 + r * 3 + j] : 0.0;
        }
    const double* d = c.fdia + (size_t)i * 9;
    for (int v = 0; v < 4; v++)
      for (int r = 0; r < 3; r++) c.f_dinv[pack_index(v, sl, r, np)] = v < 3 ? d[r * 3 + v] : 0.0;
  }
  {
    const int i = c.b_rc[2 * sl], cnt = c.b_rc[2 * sl + 1];
    const double* src = c.fcrs + (size_t)c.diaind[i] * 9;
    if (c.precision)
      for (int v = 0; v < 4 * pack_quads(c.kpack); v++)
        for (int r = 0; r < 3; r++) {
          const int k = v / 3, j = v % 3;
          reinterpret_cast<float*>(c.b_blk)[pack_index32(v, sl, r, np)] =
              (k < cnt && k < c.kpack) ? (float)src[k * 9 + r * 3 + j] : 0.0f;
        }
    else
      for (int v = 0; v < 2 * pack_pairs(c.kpack); v++)
        for (int r = 0; r < 3; r++) {
          const int k = v / 3, j = v % 3;
          c.b_blk[pack_index(v, sl, r, np)] = (k < cnt && k < c.kpack) ? src[k * 9 + r * 3 + j] : 0.0;
        }
  }
}

// The forward sweep's right-hand side rides in the free fourth value of each
// slot's D^-1 pairs (value 3 of the packed D^-1): one load gives a lane its
// last D^-1 entry and its right-hand side.  Written after k_pack_factor.
__device__ __forceinline__ size_t fwd_rhs_index(int np, int slot, int r) {
  return pack_index(3, slot, r, np);
}

// x (row order) -> the forward sweep's right-hand side (forward-slot order)
template <class Sel>
__global__ void k_gather_fslots(Sel sel, const double* x) {
  const DevCloth c = sel.cloth();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * c.np) return;
  const int sl = e / 3, r = e - 3 * sl;
  c.f_dinv[fwd_rhs_index(c.np, sl, r)] = x[3 * c.f_rci[4 * sl] + r];
}

constexpr int kQuadThreads = 4 * kQuadRows;  // at most 512: four lanes per row of a sub-level

// value of lane j of this lane's quad
template <int j>
__device__ __forceinline__ double quad_bcast(double v) {
  constexpr int ctrl = j | (j << 2) | (j << 4) | (j << 6);
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctrl, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Workgroup barrier that orders LDS only: the ring writes of a level are
// visible after it, while the loads issued for later levels stay in flight
// (a full __syncthreads() would wait for them).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

struct QuadRow {
  bool live; // some quad of this wave has a row (wave-uniform)
  bool has;  // this quad has a row at this sub-level
  int s;     // slot in this sweep (the row's ring entry is s mod kRingRows)
  int out;   // output row: forward, the row's backward slot; backward, the row
  double t;  // this lane's component of the right-hand side
  int col[kMaxBlocksPerRow];  // byte offsets of the column rows' ring entries
  double a[kMaxBlocksPerRow * 3];
  double d[3];
};

// Bases of one sweep's operands, formed once per launch: every array is
// addressed as base[unsigned 32-bit index], which the loads take as a
// scalar base plus a per-lane offset (no 64-bit address arithmetic per load).
// F32: the coefficients as FP32 quads (mixed precision), else FP64 pairs
template <bool kFwd, int K, bool F32>
struct QuadSrc {
  const int4* meta[meta_groups(K)];  // sweep metadata, 16-byte groups per slot
  const double* rhs;
  const double2* blk[pack_pairs(K)];
  const float4* blk4[pack_quads(K)];
  const double2* d0;  // D^-1 values 0-1 and 2-3 (forward)
  const double2* d1;
};

template <bool kFwd, int K, bool F32>
__device__ __forceinline__ QuadSrc<kFwd, K, F32> quad_src(const DevCloth& c, const double* rhs) {
  QuadSrc<kFwd, K, F32> q;
  const size_t np = (size_t)c.np;
  const int4* meta = reinterpret_cast<const int4*>(kFwd ? c.f_meta : c.b_meta);
#pragma unroll
  for (int g = 0; g < meta_groups(K); g++) q.meta[g] = meta + g * np;
  q.rhs = rhs;
  const double2* blk = reinterpret_cast<const double2*>(kFwd ? c.f_blk : c.b_blk);
#pragma unroll
  for (int p = 0; p < pack_pairs(K); p++) q.blk[p] = blk + p * np * 3;
  const float4* blk4 = reinterpret_cast<const float4*>(kFwd ? c.f_blk : c.b_blk);
#pragma unroll
  for (int p = 0; p < pack_quads(K); p++) q.blk4[p] = blk4 + p * np * 3;
  q.d0 = reinterpret_cast<const double2*>(c.f_dinv);
  q.d1 = q.d0 + np * 3;
  return q;
}

// base[i] for an index whose byte offset fits 32 bits: the load then takes the
// scalar base plus a 32-bit per-lane offset
template <class T>
__device__ __forceinline__ T ld32(const T* base, unsigned i) {
  return gload(reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                          i * (unsigned)sizeof(T)));
}

// Sub-level cuts of one sweep, read at wave-uniform indices.  CutsLds reads
// a copy of one component's cuts in LDS (indices from its first sub-level
// q0); CutsGlobal reads them from global memory as scalar loads (any number
// of sub-levels).  A scalar load that misses the caches costs about a
// microsecond, and each new cache line of cuts was paid for by the waves
// waiting on it before they could issue their loads.
struct CutsLds {
  const int* p;
  int q0;
  __device__ int at(int l) const { return __builtin_amdgcn_readfirstlane(p[l - q0]); }
  __device__ int any() const { return q0; }  // an index that may be read
};
struct CutsGlobal {
  const int* p;
  __device__ int at(int l) const { return p[l]; }
  __device__ int any() const { return 0; }
};

// Branch-free on purpose: a quad without a row at this sub-level loads slot 0
// and is only marked (i = -1), lane 3 loads lane 2's right-hand side.  Loads
// on one path only (even a wave-uniform skip) make the compiler wait for every
// outstanding load where the paths merge, defeating the prefetch.
template <bool kFwd, int K, bool F32, class Cuts>
__device__ __forceinline__ void quad_load(const QuadSrc<kFwd, K, F32>& src, const Cuts& cuts, int q1,
                                          int l, QuadRow& q) {
  const int lane = threadIdx.x & 3, qi = threadIdx.x >> 2;
  const int lc = l < q1 ? l : cuts.any();  // past the last sub-level: marked empty
  const int first = cuts.at(lc), end = cuts.at(lc + 1);
  const int s0 = first + qi;
  const bool valid = l < q1 && s0 < end;
  // the quads of a wave are consecutive: the wave has a row iff its first does
  q.live = __builtin_amdgcn_readfirstlane(
               (int)(l < q1 && first + (int)(threadIdx.x >> 6) * 16 < end)) != 0;
  const unsigned sl = valid ? (unsigned)s0 : 0u;
  q.s = (int)sl;
  q.has = valid;
  unsigned mw[4 * meta_groups(K)];  // metadata as 32-bit words (two 16-bit entries each)
#pragma unroll
  for (int g = 0; g < meta_groups(K); g++) {
    const int4 v = ld32(src.meta[g], sl);
    mw[4 * g] = (unsigned)v.x;
    mw[4 * g + 1] = (unsigned)v.y;
    mw[4 * g + 2] = (unsigned)v.z;
    mw[4 * g + 3] = (unsigned)v.w;
  }
  auto word16 = [&](int h) { return (mw[h / 2] >> (16 * (h % 2))) & 0xffffu; };
#pragma unroll
  for (int k = 0; k < K; k++) q.col[k] = (int)word16(k);
  q.out = (int)(word16(K) | (word16(K + 1) << 16));
  const unsigned lr = lane < 3 ? lane : 2;  // lane 3 shadows lane 2 (same cache line)
  const unsigned e = 3 * sl + lr;
  if (!kFwd) q.t = ld32(src.rhs, e);  // forward: with the last D^-1 entry below
  if constexpr (F32) {
#pragma unroll
    for (int p = 0; p < pack_quads(K); p++) {
      const float4 w = ld32(src.blk4[p], e);
      q.a[4 * p] = (double)w.x;
      if (4 * p + 1 < 3 * K) q.a[4 * p + 1] = (double)w.y;
      if (4 * p + 2 < 3 * K) q.a[4 * p + 2] = (double)w.z;
      if (4 * p + 3 < 3 * K) q.a[4 * p + 3] = (double)w.w;
    }
  } else {
#pragma unroll
    for (int p = 0; p < pack_pairs(K); p++) {
      const double2 w = ld32(src.blk[p], e);
      q.a[2 * p] = w.x;
      if (2 * p + 1 < 3 * K) q.a[2 * p + 1] = w.y;
    }
  }
  if (kFwd) {
    const double2 w0 = ld32(src.d0, e);
    const double2 w1 = ld32(src.d1, e);
    q.d[0] = w0.x;
    q.d[1] = w0.y;
    q.d[2] = w1.x;
    q.t = w1.y;
  }
}

// one row component; every lane of a wave with rows runs it (the DPP
// exchange needs whole quads).  Branch-free up to the global store: a block
// past the row's count has zero coefficients and reads the ring's zero
// entry, so its update t - ((0*0 + 0*0) + 0*0) = t - (+0) leaves t unchanged
// bit for bit (also -0, infinities and NaN); lanes without a row (and lane 3)
// write their value to the ring's dummy entry.  Everything loaded is thus
// consumed on every path, which keeps the compiler from waiting for the next
// sub-level's loads.
template <bool kFwd, int K>
__device__ __forceinline__ void quad_compute(double (*ring)[kRingLen], double* out,
                                             const QuadRow& q) {
  constexpr int kMask = kRingRows - 1;
  if (!q.live) return;  // no row in this wave: nothing to compute (its loads are not skipped)
  const int lane = threadIdx.x & 3;
  double y[K][3];
  const char* rb = reinterpret_cast<const char*>(&ring[0][0]);
  constexpr int kRow = kRingLen * (int)sizeof(double);  // bytes per ring row
#pragma unroll
  for (int k = 0; k < K; k++) {
    const char* at = rb + q.col[k];
    y[k][0] = *reinterpret_cast<const double*>(at);
    y[k][1] = *reinterpret_cast<const double*>(at + kRow);
    y[k][2] = *reinterpret_cast<const double*>(at + 2 * kRow);
  }
  double t = q.t;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const double* a = q.a + k * 3;
    t = t - (a[0] * y[k][0] + a[1] * y[k][1] + a[2] * y[k][2]);
  }
  double z = t;
  if (kFwd) {
    const double t0 = quad_bcast<0>(t), t1 = quad_bcast<1>(t), t2 = quad_bcast<2>(t);
    z = q.d[0] * t0 + q.d[1] * t1 + q.d[2] * t2;
  }
  const bool own = q.has && lane < 3;
  ring[lane < 3 ? lane : 2][own ? (q.s & kMask) : kRingDummy] = z;
  if (own) gstore(out + 3 * q.out + lane, z);  // forward: backward-slot order; backward: row order
}

// sub-level cuts of one component's sweep kept in LDS (a sweep with more
// sub-levels reads them from global memory)
constexpr int kCutsLds = 4096;

// this wave has no row in sub-levels s .. s + 2 (those past q1 have none);
// wave-uniform
template <class Cuts>
__device__ __forceinline__ bool wave_idle3(const Cuts& cuts, int q1, int s) {
  const int w16 = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * 16;
  bool idle = true;
#pragma unroll
  for (int j = 0; j < 3; j++)
    if (s + j < q1) idle = idle && cuts.at(s + j) + w16 >= cuts.at(s + j + 1);
  return idle;
}

// one sweep over the sub-levels; three register sets rotate so that the
// loads of sub-levels l + 1 and l + 2 are in flight while l is computed
// over the sub-levels q0 .. q1 - 1 (one component's)
template <bool kFwd, int K, bool F32, class Cuts>
__device__ __forceinline__ void quad_sweep(const DevCloth& c, const Cuts& cuts, int q0, int q1,
                                           const double* rhs, double* out,
                                           double (*ring)[kRingLen]) {
  const QuadSrc<kFwd, K, F32> src = quad_src<kFwd, K, F32>(c, rhs);
  QuadRow qa, qb, qc;
  quad_load<kFwd, K, F32>(src, cuts, q1, q0, qa);
  quad_load<kFwd, K, F32>(src, cuts, q1, q0 + 1, qb);
  quad_load<kFwd, K, F32>(src, cuts, q1, q0 + 2, qc);
  // all three thirds unconditional (a count that is not a multiple of three
  // ends with empty sub-levels): a skipped third would change which loads are
  // the newest on one path, and the compiler would then wait for all loads at
  // the top of every iteration
  // a set is reloaded right after its compute, before the barrier: waves
  // that finish computing early issue their loads while others still compute
  for (int l = q0; l < q1; l += 3) {
    if (!qa.live && !qb.live && !qc.live) {
      // This wave has no row in sub-levels l .. l + 2 (narrow levels leave
      // the high waves idle).  Its loads would cost as much issue time as a
      // working wave's, so while the next group is idle too it only takes
      // the group's three barriers and moves on; the group before the next
      // working one then runs below (nothing to compute) and loads that
      // one.  No set is touched here, so both paths reach the code below
      // with the same loads in flight.
      while (l + 3 < q1 && wave_idle3(cuts, q1, l + 3)) {
        lds_barrier();
        lds_barrier();
        lds_barrier();
        l += 3;
      }
    }
    quad_compute<kFwd, K>(ring, out, qa);
    quad_load<kFwd, K, F32>(src, cuts, q1, l + 3, qa);
    lds_barrier();
    quad_compute<kFwd, K>(ring, out, qb);
    quad_load<kFwd, K, F32>(src, cuts, q1, l + 4, qb);
    lds_barrier();
    quad_compute<kFwd, K>(ring, out, qc);
    quad_load<kFwd, K, F32>(src, cuts, q1, l + 5, qc);
    lds_barrier();
  }
}

// M^-1 (L+D)(I+D^-1 U) solve of the factor: forward sweep from the
// right-hand side in forward-slot order (fwd_rhs_index) into c.sb, backward
// sweep into dst (row order); prod = src_row .* dst when prod is given.
// One workgroup per connected component (blockIdx.x): the matrix is block
// diagonal over them, so their sweeps are independent.
// K: the packed blocks per row (DevCloth::kpack)
// Forward then backward sweep of component `comp` by the whole workgroup:
// the forward result into c.sb (backward-slot order), the backward result
// into dst (row order).  ring: the LDS ring, its zero entry already set;
// cuts: LDS for the component's sub-level cuts (kCutsLds + 1 entries).
template <int K, bool F32 = false>
__device__ __forceinline__ void quad_component(const DevCloth& c, int comp, double* dst,
                                               double (*ring)[kRingLen], int* cuts) {
  const int f0 = c.qf_comp[comp], f1 = c.qf_comp[comp + 1];
  const bool fl = f1 - f0 <= kCutsLds;
  if (fl)
    for (int i = threadIdx.x; i <= f1 - f0; i += blockDim.x) cuts[i] = gload(c.qf_ptr + f0 + i);
  __syncthreads();
  if (fl)
    quad_sweep<true, K, F32>(c, CutsLds{cuts, f0}, f0, f1, nullptr, c.sb, ring);
  else
    quad_sweep<true, K, F32>(c, CutsGlobal{c.qf_ptr}, f0, f1, nullptr, c.sb, ring);
  __syncthreads();  // the backward sweep reads the component's complete forward result
  const int b0 = c.qb_comp[comp], b1 = c.qb_comp[comp + 1];
  const bool bl = b1 - b0 <= kCutsLds;
  if (bl)
    for (int i = threadIdx.x; i <= b1 - b0; i += blockDim.x) cuts[i] = gload(c.qb_ptr + b0 + i);
  __syncthreads();
  if (bl)
    quad_sweep<false, K, F32>(c, CutsLds{cuts, b0}, b0, b1, c.sb, dst, ring);
  else
    quad_sweep<false, K, F32>(c, CutsGlobal{c.qb_ptr}, b0, b1, c.sb, dst, ring);
  __syncthreads();
}

template <int K, bool F32, class Sel>
__global__ __launch_bounds__(kQuadThreads) void k_ilu_apply_quad(Sel sel, const double* src_row,
                                                                 double* dst, double* prod) {
  const DevCloth c = sel.cloth();
  const int comp = blockIdx.x;
  if (c.flags[kSkip] || comp >= c.ncomp) return;  // batch grids fit the most components
  if (!src_row) {
    src_row = c.r;
    dst = c.z;
    prod = c.prod;
  }
  __shared__ double ring[3][kRingLen];  // + the dummy and the zero entry
  __shared__ int cuts[kCutsLds + 1];
  if (threadIdx.x < 3) ring[threadIdx.x][kRingZero] = 0.0;  // never written again
  quad_component<K, F32>(c, comp, dst, ring, cuts);
  if (!prod) return;
  if (c.ncomp == 1) {
    for (int e = threadIdx.x; e < 3 * c.np; e += blockDim.x)
      gstore(prod + e, gload(src_row + e) * gload(dst + e));
  } else {  // this component's rows: its forward slots
    const int s0 = c.qf_ptr[c.qf_comp[comp]], s1 = c.qf_ptr[c.qf_comp[comp + 1]];
    for (int k = threadIdx.x; k < 3 * (s1 - s0); k += blockDim.x) {
      const int sl = s0 + k / 3;
      const int e = 3 * gload(c.f_rci + 4 * sl) + (k - 3 * (k / 3));
      gstore(prod + e, gload(src_row + e) * gload(dst + e));
    }
  }
}

// threads of a quad sweep: four per row of the widest sub-level (at most
// kQuadRows rows), in whole waves; waves past it would only load and wait
inline int quad_threads(int qwidth) {
  const int rows = std::min(std::max(qwidth, 1), kQuadRows);
  return 4 * ((rows + 15) / 16 * 16);
}

template <bool F32, class Sel>
static void launch_quad_p(int kq, dim3 grid, dim3 tb, const Sel& sel, const double* src,
                          double* dst, double* prod, hipStream_t s) {
  switch (kq) {
    case 1: hipLaunchKernelGGL((k_ilu_apply_quad<1, F32, Sel>), grid, tb, 0, s, sel, src, dst, prod); break;
    case 2: hipLaunchKernelGGL((k_ilu_apply_quad<2, F32, Sel>), grid, tb, 0, s, sel, src, dst, prod); break;
    case 3: hipLaunchKernelGGL((k_ilu_apply_quad<3, F32, Sel>), grid, tb, 0, s, sel, src, dst, prod); break;
    case 4: hipLaunchKernelGGL((k_ilu_apply_quad<4, F32, Sel>), grid, tb, 0, s, sel, src, dst, prod); break;
    case 5: hipLaunchKernelGGL((k_ilu_apply_quad<5, F32, Sel>), grid, tb, 0, s, sel, src, dst, prod); break;
    case 6: hipLaunchKernelGGL((k_ilu_apply_quad<6, F32, Sel>), grid, tb, 0, s, sel, src, dst, prod); break;
    case 7: hipLaunchKernelGGL((k_ilu_apply_quad<7, F32, Sel>), grid, tb, 0, s, sel, src, dst, prod); break;
    default: hipLaunchKernelGGL((k_ilu_apply_quad<8, F32, Sel>), grid, tb, 0, s, sel, src, dst, prod); break;
  }
}

// f32: the cloths' packed coefficients are FP32 (mixed precision)
template <class Sel>
static void launch_quad(int kq, bool f32, int qwidth, dim3 grid, const Sel& sel, const double* src,
                        double* dst, double* prod, hipStream_t s) {
  const dim3 tb(quad_threads(qwidth));
  if (f32)
    launch_quad_p<true>(kq, grid, tb, sel, src, dst, prod, s);
  else
    launch_quad_p<false>(kq, grid, tb, sel, src, dst, prod, s);
}

bool sweep_shape(const DevCloth& c, int* waves, int* loads_fwd, int* loads_bwd) {
  if (!quad_ok(c) || c.sweep_variant == 1) return false;
  const int K = std::min(std::max(c.kpack, 1), 8);  // the kernel instance launch_quad picks
  *waves = quad_threads(c.qwidth) / 64;
  // per lane: the metadata groups and coefficient pairs, then two D^-1 pairs
  // (forward, the right-hand side in the second) or the right-hand side
  const int blk = c.precision ? pack_quads(K) : pack_pairs(K);
  *loads_fwd = meta_groups(K) + blk + 2;
  *loads_bwd = meta_groups(K) + blk + 1;
  return true;
}

// one forward+backward substitution launch (variants give identical results):
// 0 automatic (the quad sweep when the mesh allows it), 1 row-by-row CSR
// sweep.  `rhs_ready`: the forward sweep's right-hand side already holds src
// (the PCG producers write it).
static void launch_sweep(const DevCloth& c, const double* src, double* dst, double* prod,
                         hipStream_t s, bool rhs_ready = false) {
  const One one{c, {}};
  if (quad_ok(c) && c.sweep_variant != 1) {
    if (!rhs_ready)
      hipLaunchKernelGGL(k_gather_fslots<One>, dim3(grid_for(3L * c.np)), dim3(kTB), 0, s, one, src);
    launch_quad(c.kpack, c.precision != 0, c.qwidth, dim3(c.ncomp), one, src, dst, prod, s);
  } else {
    hipLaunchKernelGGL(k_ilu_apply<One>, dim3(1), dim3(1024), 0, s, one, src, dst, prod);
  }
}

// y = alpha A x + beta y over the block-CSR + diagonal (matrix_square_sparse.cpp:218-242)
__global__ void k_matvec(int np, const int* rowptr, const int* colind, const double* dia,
                         const double* crs, double alpha, const double* x, double beta,
                         double* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double y0 = y[3 * i] * beta, y1 = y[3 * i + 1] * beta, y2 = y[3 * i + 2] * beta;
  for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
    const int j = colind[k];
    const double* a = crs + (size_t)k * 9;
    const double x0 = x[3 * j], x1 = x[3 * j + 1], x2 = x[3 * j + 2];
    y0 += alpha * (a[0] * x0 + a[1] * x1 + a[2] * x2);
    y1 += alpha * (a[3] * x0 + a[4] * x1 + a[5] * x2);
    y2 += alpha * (a[6] * x0 + a[7] * x1 + a[8] * x2);
  }
  const double* a = dia + (size_t)i * 9;
  const double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
  y0 += alpha * (a[0] * x0 + a[1] * x1 + a[2] * x2);
  y1 += alpha * (a[3] * x0 + a[4] * x1 + a[5] * x2);
  y2 += alpha * (a[6] * x0 + a[7] * x1 + a[8] * x2);
  y[3 * i] = y0;
  y[3 * i + 1] = y1;
  y[3 * i + 2] = y2;
}

// PCG search direction and product in one pass:
//   p_new = z (first iteration) or z + beta p_old   (ilu_sparse.cpp:663, 702)
//   ap    = 0*z + 1.0*(A p_new)  -- the reference writes A p into the buffer
//           that holds z (ilu_sparse.cpp:670-672), scaling it by beta = 0
//   prod  = p_new .* ap
// iteration `it`: p_new = c.p[it & 1], p_old = c.p[(it + 1) & 1]
// one row on one lane: p_new[i] (= z[i] or z[i] + beta p_old[i]), Ap[i]
// and p_new[i] .* Ap[i]
__device__ __forceinline__ void spmv_dir_row(const DevCloth& c, int i, bool first, double beta,
                                             const double* __restrict__ pold_,
                                             double* __restrict__ pnew_) {
  const auto z = gptr(c.z);
  const auto pold = gptr(pold_);
  const auto pnew = gptr(pnew_);
  const auto rowptr = gptr(c.rowptr);
  const auto colind = gptr(c.colind);
  double y0 = z[3 * i] * 0.0, y1 = z[3 * i + 1] * 0.0, y2 = z[3 * i + 2] * 0.0;
  for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
    const int j = colind[k];
    const auto a = gptr(c.crs) + (size_t)k * 9;
    double x0, x1, x2;
    if (first) {
      x0 = z[3 * j];
      x1 = z[3 * j + 1];
      x2 = z[3 * j + 2];
    } else {
      x0 = z[3 * j] + beta * pold[3 * j];
      x1 = z[3 * j + 1] + beta * pold[3 * j + 1];
      x2 = z[3 * j + 2] + beta * pold[3 * j + 2];
    }
    y0 += 1.0 * (a[0] * x0 + a[1] * x1 + a[2] * x2);
    y1 += 1.0 * (a[3] * x0 + a[4] * x1 + a[5] * x2);
    y2 += 1.0 * (a[6] * x0 + a[7] * x1 + a[8] * x2);
  }
  double x0, x1, x2;
  if (first) {
    x0 = z[3 * i];
    x1 = z[3 * i + 1];
    x2 = z[3 * i + 2];
  } else {
    x0 = z[3 * i] + beta * pold[3 * i];
    x1 = z[3 * i + 1] + beta * pold[3 * i + 1];
    x2 = z[3 * i + 2] + beta * pold[3 * i + 2];
  }
  const auto a = gptr(c.dia) + (size_t)i * 9;
  y0 += 1.0 * (a[0] * x0 + a[1] * x1 + a[2] * x2);
  y1 += 1.0 * (a[3] * x0 + a[4] * x1 + a[5] * x2);
  y2 += 1.0 * (a[6] * x0 + a[7] * x1 + a[8] * x2);
  const auto ap = gptr(c.ap);
  ap[3 * i] = y0;
  ap[3 * i + 1] = y1;
  ap[3 * i + 2] = y2;
  pnew[3 * i] = x0;
  pnew[3 * i + 1] = x1;
  pnew[3 * i + 2] = x2;
  const auto prod = gptr(c.prod);
  prod[3 * i] = x0 * y0;
  prod[3 * i + 1] = x1 * y1;
  prod[3 * i + 2] = x2 * y2;
}

// p_new, A p_new and their products with a group of kSpG lanes per row: lane l forms block l's
// product (the diagonal block on lane deg), reading the row's contiguous
// blocks side by side; the products are then added in the reference's order
// (off-diagonal blocks in column order, the diagonal last) as spmv_dir_row
// adds them, so every result is the same bit for bit.  Rows with more than
// 2 kSpG - 1 blocks fall back to spmv_dir_row on the group's first lane.
constexpr int kSpG = 16;
static_assert(kTB % kSpG == 0 && kTB / kSpG == kSpmvRows, "SpMV rows per block");
static_assert(kTB == kVecBlock, "vector-update entries per block");
template <class Sel>
__global__ __launch_bounds__(kTB) void k_spmv_dir_g(Sel sel, int it) {
  const DevCloth c = sel.cloth();
  if (c.flags[kDone]) return;
  const int i = sel.block_row0(kTB / kSpG) + (int)(threadIdx.x / kSpG);
  const int l = (int)(threadIdx.x % kSpG);
  if (i >= c.np) return;  // a whole group at a time
  const bool first = it == 0;
  const double beta = c.scal[kBeta];
  const auto z = gptr(c.z);
  const auto pold = gptr(pcg_dir(c, it + 1));
  const int rb = gptr(c.rowptr)[i], deg = gptr(c.rowptr)[i + 1] - rb;
  if (deg >= 2 * kSpG) {
    if (l == 0) spmv_dir_row(c, i, first, beta, pcg_dir(c, it + 1), pcg_dir(c, it));
    return;
  }
  // the product of block k (k == deg: the diagonal) with p_new of its column
  auto product = [&](int k, double& t0, double& t1, double& t2) {
    t0 = t1 = t2 = 0.0;
    if (k > deg) return;
    const int j = k < deg ? gptr(c.colind)[rb + k] : i;
    const auto a = k < deg ? gptr(c.crs) + (size_t)(rb + k) * 9 : gptr(c.dia) + (size_t)i * 9;
    double x0, x1, x2;
    if (first) {
      x0 = z[3 * j];
      x1 = z[3 * j + 1];
      x2 = z[3 * j + 2];
    } else {
      x0 = z[3 * j] + beta * pold[3 * j];
      x1 = z[3 * j + 1] + beta * pold[3 * j + 1];
      x2 = z[3 * j + 2] + beta * pold[3 * j + 2];
    }
    t0 = a[0] * x0 + a[1] * x1 + a[2] * x2;
    t1 = a[3] * x0 + a[4] * x1 + a[5] * x2;
    t2 = a[6] * x0 + a[7] * x1 + a[8] * x2;
  };
  double a0, a1, a2, b0, b1, b2;
  product(l, a0, a1, a2);
  product(l + kSpG, b0, b1, b2);
  double y0 = z[3 * i] * 0.0, y1 = z[3 * i + 1] * 0.0, y2 = z[3 * i + 2] * 0.0;
  for (int k = 0; k <= deg; k++) {  // deg is the same on every lane of the group
    const bool hi = k >= kSpG;
    const int src = k - (hi ? kSpG : 0);
    const double t0 = __shfl(hi ? b0 : a0, src, kSpG);
    const double t1 = __shfl(hi ? b1 : a1, src, kSpG);
    const double t2 = __shfl(hi ? b2 : a2, src, kSpG);
    y0 += 1.0 * t0;
    y1 += 1.0 * t1;
    y2 += 1.0 * t2;
  }
  if (l != 0) return;
  double x0, x1, x2;
  if (first) {
    x0 = z[3 * i];
    x1 = z[3 * i + 1];
    x2 = z[3 * i + 2];
  } else {
    x0 = z[3 * i] + beta * pold[3 * i];
    x1 = z[3 * i + 1] + beta * pold[3 * i + 1];
    x2 = z[3 * i + 2] + beta * pold[3 * i + 2];
  }
  const auto ap = gptr(c.ap), pnew = gptr(pcg_dir(c, it)), prod = gptr(c.prod);
  ap[3 * i] = y0;
  ap[3 * i + 1] = y1;
  ap[3 * i + 2] = y2;
  pnew[3 * i] = x0;
  pnew[3 * i + 1] = x1;
  pnew[3 * i + 2] = x2;
  prod[3 * i] = x0 * y0;
  prod[3 * i + 1] = x1 * y1;
  prod[3 * i + 2] = x2 * y2;
}

// one entry of k_update_rx: r -= alpha Ap, dx += alpha p, r .* r, and r as
// the next forward sweep's right-hand side
__device__ __forceinline__ void update_rx_elem(const DevCloth& c, int e, double alpha,
                                               const double* __restrict__ p) {
  const auto rr = gptr(c.r);
  const double r = rr[e] + -alpha * gptr(c.ap)[e];
  rr[e] = r;
  const auto dx = gptr(c.dx);
  dx[e] = dx[e] + alpha * gptr(p)[e];
  gptr(c.prod2)[e] = r * r;
  if (c.fslot) {
    const int i = e / 3;
    gptr(c.f_dinv)[fwd_rhs_index(c.np, gptr(c.fslot)[i], e - 3 * i)] = r;  // the next sweep's input
  }
}

// r -= alpha A p; x += alpha p; prod = r .* r  (matrix_square_sparse.cpp:327-336)
// Also snapshots kDone into kSkip for the kernels that run beside the next
// convergence test (no kernel changes kDone while this one runs).
template <class Sel>
__global__ void k_update_rx(Sel sel, int it) {
  const DevCloth c = sel.cloth();
  const double* __restrict__ p = pcg_dir(c, it);
  const int done = c.flags[kDone];
  const int e0 = sel.block_row0(kTB);  // first entry of this block
  if (e0 == 0 && threadIdx.x == 0) c.flags[kSkip] = done;
  if (done) return;
  const int e = e0 + (int)threadIdx.x;
  if (e >= 3 * c.np) return;
  update_rx_elem(c, e, c.scal[kAlpha], p);
}

template <class Sel>
__global__ void k_pcg_init(Sel sel, double tol2) {
  const DevCloth c = sel.cloth();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e == 0) {
    // the solve's flags (the factorisation's singular-pivot count stays)
    c.flags[kDone] = 0;
    c.flags[kIters] = 0;
    c.flags[kSkip] = 0;
    c.flags[kTicketErr] = 0;
    c.flags[kFlagNonFinite] = 0;
    c.scal[kTol2] = tol2;
    c.scal[kBeta] = 0.0;
  }
  if (e >= 3 * c.np) return;
  const double b = c.b[e];
  c.r[e] = b;
  c.dx[e] = 0;
  c.prod[e] = b * b;
  if (c.fslot) {
    const int i = e / 3;
    c.f_dinv[fwd_rhs_index(c.np, c.fslot[i], e - 3 * i)] = b;
  }
}

// ---------------------------------------------------------------------------
// reductions: the reference sums left to right (matrix_square_sparse.cpp:313-323);
// ordered_sum_exact (cs_ordered_sum.hpp) reproduces that loop bit for bit.

// what the PCG does with a finished reduction (ilu_sparse.cpp:648-708)
enum DotMode : int { kDotRr0 = 0, kDotRz0 = 1, kDotPap = 2, kDotRr = 3, kDotRz = 4, kDotFinal = 5 };

// exact left-to-right sum of prod, then the PCG bookkeeping of `mode`
// (ilu_sparse.cpp:646-700); every mode but the first is skipped once the
// solve has converged.  The (r, z) reductions run beside the convergence test
// that may set kDone, so they read its snapshot kSkip (k_update_rx), which is
// constant while they run: every block of a launch takes the same branch.
//
// Which products, scratch and skip flag a mode uses is decided on the host
// (dot_src) and passed as two plain values, so job() is two two-way selects.
// (A three-way choice of the skip pointer written as a conditional on `mode`
// was compiled for gfx950 into a branch structure that left the addend and
// skip-flag pointers undefined for the (p, Ap) mode, and the kernel faulted
// on the load of the flag.)
template <class Sel>
struct DotSrc {
  Sel sel;
  int mode;
  int iter;
  int aux;   // 1: the convergence test's products and scratch (prod2, sum2)
  int skip;  // flag that cancels the reduction, -1: none
  __device__ SumJob job() const {
    const DevCloth c = sel.cloth();
    const bool a = aux != 0;
    // the convergence test reduces its own products with its own scratch: it
    // runs beside the sweep and the (r, z) reduction
    return SumJob{a ? c.prod2 : c.prod, 3 * c.np, a ? c.sum2 : c.sum,
                  skip < 0 ? nullptr : c.flags + skip};
  }
  __device__ void finish(double s) const {
    const DevCloth c = sel.cloth();
    c.scal[kLastDot] = s;
    switch (mode) {
      case kDotRr0:
        if (s < 1.0e-30) {
          c.scal[kConv] = 0.0;
          c.flags[kIters] = 0;
          c.flags[kDone] = 1;
          c.flags[kSkip] = 1;  // nothing runs beside this one
        } else {
          c.scal[kInvRr0] = 1.0 / s;
        }
        break;
      case kDotRz0:
        c.scal[kRz] = s;
        break;
      case kDotPap:
        c.scal[kAlpha] = c.scal[kRz] / s;
        break;
      case kDotRr: {
        const double q = s * c.scal[kInvRr0];
        if (q < c.scal[kTol2]) {
          c.scal[kConv] = sqrt(q);
          c.flags[kIters] = iter;
          c.flags[kDone] = 1;
        }
        break;
      }
      case kDotRz: {
        c.scal[kBeta] = s / c.scal[kRz];
        c.scal[kRz] = s;
        break;
      }
      case kDotFinal:
        c.scal[kConv] = sqrt(s * c.scal[kInvRr0]);
        c.flags[kIters] = iter;
        break;
    }
  }
};

// the reduction of PCG mode `mode` (see DotSrc)
template <class Sel>
DotSrc<Sel> dot_src(const Sel& sel, int mode, int iter) {
  const int aux = mode == kDotRr ? 1 : 0;
  const int skip = mode == kDotRr0 ? -1 : (mode == kDotRz0 || mode == kDotRz) ? kSkip : kDone;
  return DotSrc<Sel>{sel, mode, iter, aux, skip};
}

static void launch_dot(const DevCloth& c, int mode, int iter, hipStream_t s) {
  launch_ordered_sum(dot_src(One{c, {}}, mode, iter), 3 * c.np, 1, s);
}

template <class Sel>
__global__ void k_square(Sel sel) {
  const DevCloth c = sel.cloth();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * c.np) return;
  c.prod[e] = c.r[e] * c.r[e];
}

// after a solve every reduction has run to its end, which leaves its ticket
// at zero; a nonzero ticket would mean blocks of one launch disagreed on
// skipping it, and is reported (kTicketErr) instead of corrupting the next solve
template <class Sel>
__global__ void k_pcg_check(Sel sel) {
  const DevCloth c = sel.cloth();
  if (threadIdx.x == 0)
    c.flags[kTicketErr] = (*c.sum.ticket != 0u || *c.sum2.ticket != 0u) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Solve_PCG (ilu_sparse.cpp:628-709) of one cloth in one workgroup: the
// search direction and A p, the update of x and r, the preconditioner sweeps
// and the four exact reductions of an iteration run back to back with
// workgroup barriers between them, and the convergence test decides inside
// the kernel.  Each operation is the one of the multi-launch solve (the same
// row and entry functions, the same sweeps, sums equal to the left-to-right
// loop), so x, r, the ratio and the count are identical.  One launch per
// solve; in a batch every cloth iterates to its own convergence, none waits
// for another.  The LDS holds the sweep ring and cuts and the sums' staging.

constexpr int kWgThreads = 512;

// z = M^-1 r (every component, forward then backward) and prod = r .* z
template <int K>
__device__ __forceinline__ void wg_precondition(const DevCloth& c, double (*ring)[kRingLen],
                                                int* cuts) {
  for (int comp = 0; comp < c.ncomp; comp++) quad_component<K>(c, comp, c.z, ring, cuts);
  const auto r = gptr(c.r), z = gptr(c.z);
  const auto prod = gptr(c.prod);
  for (int e = threadIdx.x; e < 3 * c.np; e += blockDim.x) prod[e] = r[e] * z[e];
  __syncthreads();
}

template <int K, class Sel>
__global__ __launch_bounds__(kWgThreads) void k_pcg_wg(Sel sel, double tol2, int maxit) {
  const DevCloth c = sel.cloth();
  __shared__ double ring[3][kRingLen];
  __shared__ int cuts[kCutsLds + 1];
  __shared__ WgSum ws;
  const int tid = threadIdx.x, nth = blockDim.x, n3 = 3 * c.np;
  if (tid < 3) ring[tid][kRingZero] = 0.0;  // never written again
  // r = b, dx = 0, b .* b, r as the forward sweep's right-hand side (k_pcg_init)
  for (int e = tid; e < n3; e += nth) {
    const double b = gptr(c.b)[e];
    gptr(c.r)[e] = b;
    gptr(c.dx)[e] = 0;
    gptr(c.prod)[e] = b * b;
    const int i = e / 3;
    gptr(c.f_dinv)[fwd_rhs_index(c.np, gptr(c.fslot)[i], e - 3 * i)] = b;
  }
  __syncthreads();
  double conv = 0.0, inv_rr0 = 0.0, rz = 0.0, beta = 0.0;
  int iters = 0;
  bool done = false;
  const double rr0 = wg_ordered_sum(c.prod, n3, ws);
  if (rr0 < 1.0e-30) {
    done = true;
  } else {
    inv_rr0 = 1.0 / rr0;
    wg_precondition<K>(c, ring, cuts);
    rz = wg_ordered_sum(c.prod, n3, ws);
  }
  for (int it = 0; it < maxit && !done; it++) {
    const double* __restrict__ pold = pcg_dir(c, it + 1);
    double* __restrict__ pnew = pcg_dir(c, it);
    for (int i = tid; i < c.np; i += nth) spmv_dir_row(c, i, it == 0, beta, pold, pnew);
    __syncthreads();
    const double alpha = rz / wg_ordered_sum(c.prod, n3, ws);
    for (int e = tid; e < n3; e += nth) update_rx_elem(c, e, alpha, pnew);
    __syncthreads();
    const double q = wg_ordered_sum(c.prod2, n3, ws) * inv_rr0;
    if (q < tol2) {
      conv = sqrt(q);
      iters = it;
      done = true;
      break;
    }
    wg_precondition<K>(c, ring, cuts);
    const double rzn = wg_ordered_sum(c.prod, n3, ws);
    beta = rzn / rz;
    rz = rzn;
  }
  if (!done) {
    // cap reached: the achieved ratio (ilu_sparse.cpp:706-707)
    for (int e = tid; e < n3; e += nth) gptr(c.prod)[e] = gptr(c.r)[e] * gptr(c.r)[e];
    __syncthreads();
    conv = sqrt(wg_ordered_sum(c.prod, n3, ws) * inv_rr0);
    iters = maxit;
  }
  if (tid == 0) {
    c.scal[kConv] = conv;
    c.scal[kInvRr0] = inv_rr0;
    c.scal[kRz] = rz;
    c.scal[kBeta] = beta;
    c.scal[kTol2] = tol2;
    c.flags[kIters] = iters;
    c.flags[kDone] = done ? 1 : 0;
    c.flags[kSkip] = 0;
    c.flags[kTicketErr] = 0;
    c.flags[kFlagNonFinite] = 0;
  }
}

// the one-workgroup solve applies: the packed sweep with its slot-order
// right-hand side, and reductions of at most kWgMaxTiles tiles
bool pcg_wg_ok(const DevCloth& c) {
  return quad_ok(c) && c.sweep_variant != 1 && c.fslot != nullptr && c.precision == 0 &&
         sum_tiles(3 * c.np) <= kWgMaxTiles;
}

template <class Sel>
static void launch_pcg_wg(int kq, dim3 grid, const Sel& sel, double tol2, int maxit,
                          hipStream_t s) {
  const dim3 tb(kWgThreads);
  switch (kq) {
    case 1: hipLaunchKernelGGL((k_pcg_wg<1, Sel>), grid, tb, 0, s, sel, tol2, maxit); break;
    case 2: hipLaunchKernelGGL((k_pcg_wg<2, Sel>), grid, tb, 0, s, sel, tol2, maxit); break;
    case 3: hipLaunchKernelGGL((k_pcg_wg<3, Sel>), grid, tb, 0, s, sel, tol2, maxit); break;
    case 4: hipLaunchKernelGGL((k_pcg_wg<4, Sel>), grid, tb, 0, s, sel, tol2, maxit); break;
    case 5: hipLaunchKernelGGL((k_pcg_wg<5, Sel>), grid, tb, 0, s, sel, tol2, maxit); break;
    case 6: hipLaunchKernelGGL((k_pcg_wg<6, Sel>), grid, tb, 0, s, sel, tol2, maxit); break;
    case 7: hipLaunchKernelGGL((k_pcg_wg<7, Sel>), grid, tb, 0, s, sel, tol2, maxit); break;
    default: hipLaunchKernelGGL((k_pcg_wg<8, Sel>), grid, tb, 0, s, sel, tol2, maxit); break;
  }
}

// x += dx, v = dx / dt (solve_internal_sparse.h:323-325)
template <class Sel>
__global__ void k_finish_internal(Sel sel) {
  const DevCloth c = sel.cloth();
  const double dt = sel.params().dt;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * c.np) return;
  const double d = c.dx[e];
  c.x[e] = c.x[e] + d;
  c.v[e] = d / dt;
  // the reference's ILU(0) can overflow on large meshes (its factors grow
  // along the ordering); such a step is reported instead of handing inf/NaN
  // positions to the contact pass
  if (!isfinite(d)) c.flags[kFlagNonFinite] = 1;
}

// prescribed positions of swept vertices: x = rest + delta * s
__global__ void k_apply_sweep(DevCloth c, const int* verts, int n, double o0, double o1, double o2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int v = verts[i];
  c.x[3 * v] = c.xyz0[3 * v] + o0;
  c.x[3 * v + 1] = c.xyz0[3 * v + 1] + o1;
  c.x[3 * v + 2] = c.xyz0[3 * v + 2] + o2;
}

// ---------------------------------------------------------------------------
// host launchers

int launch_apply_sweep(const DevCloth& c, const int* verts, int n, const double off[3],
                       hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_apply_sweep, dim3(grid_for(n)), dim3(kTB), 0, s, c, verts, n, off[0],
                       off[1], off[2]);
  CS_HIP(hipGetLastError());
  return kOk;
}

int launch_bend_setup(const DevCloth& c, const Params& P, hipStream_t s) {
  if (c.nq > 0) hipLaunchKernelGGL(k_bend_setup, dim3(grid_for(c.nq)), dim3(kTB), 0, s, c, P.stiff_bend);
  CS_HIP(hipGetLastError());
  return kOk;
}

int launch_assemble(const DevCloth& c, const Params& P, hipStream_t s, Prof* prof) {
  ProfScope ps(prof, kProfAssemble);
  const One one{c, P};
  hipLaunchKernelGGL(k_tri_elem<One>, dim3(grid_for(c.nt)), dim3(kTB), 0, s, one, 1);
  if (c.nq > 0) hipLaunchKernelGGL(k_quad_grad<One>, dim3(grid_for(c.nq)), dim3(kTB), 0, s, one);
  if (c.nnz > 0) hipLaunchKernelGGL(k_asm_crs<One>, dim3(grid_for(c.nnz)), dim3(kTB), 0, s, one);
  hipLaunchKernelGGL(k_asm_dia_rhs<One>, dim3(grid_for(c.np)), dim3(kTB), 0, s, one);
  CS_HIP(hipGetLastError());
  return kOk;
}

int launch_ilu(const DevCloth& c, hipStream_t s, Prof* prof) {
  ProfScope ps(prof, kProfFactor);
  CS_HIP(hipMemcpyAsync(c.fdia, c.dia, sizeof(double) * 9 * c.np, hipMemcpyDeviceToDevice, s));
  if (c.nnz > 0)
    CS_HIP(hipMemcpyAsync(c.fcrs, c.crs, sizeof(double) * 9 * (size_t)c.nnz,
                          hipMemcpyDeviceToDevice, s));
  CS_HIP(hipMemsetAsync(c.ising, 0, sizeof(int) * c.np, s));
  CS_HIP(hipMemsetAsync(c.flags + kNsing, 0, sizeof(int), s));
  if (c.fu_blocks > 0)
    hipLaunchKernelGGL(k_ilu_factor_staged<One>, dim3(1), dim3(ilu_stage_threads(c.fu_blocks)), 0,
                       s, One{c, {}}, 0, c.fu_blocks);
  else
    hipLaunchKernelGGL(k_ilu_factor<One>, dim3(1), dim3(1024), 0, s, One{c, {}}, 0);
  hipLaunchKernelGGL(k_ilu_redo<One>, dim3(1), dim3(1024), 0, s, One{c, {}});
  CS_HIP(hipGetLastError());
  return launch_pack_factor(c, s);
}

int launch_pack_factor(const DevCloth& c, hipStream_t s) {
  if (c.kl > 0)
    hipLaunchKernelGGL(k_pack_factor<One>, dim3(grid_for(c.np)), dim3(kTB), 0, s, One{c, {}});
  CS_HIP(hipGetLastError());
  return kOk;
}

int launch_matvec(const DevCloth& c, double alpha, const double* x, double beta, double* y,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_matvec, dim3(grid_for(c.np)), dim3(kTB), 0, s, c.np, c.rowptr, c.colind,
                     c.dia, c.crs, alpha, x, beta, y);
  CS_HIP(hipGetLastError());
  return kOk;
}

int launch_ilu_apply(const DevCloth& c, const double* src, double* dst, hipStream_t s) {
  launch_sweep(c, src, dst, nullptr, s);
  CS_HIP(hipGetLastError());
  return kOk;
}

// Solve_PCG on the assembled system (c.b) with the factor in fdia/fcrs.
// Result in c.dx; conv/iters in the device scalars.  `host_flags` is pinned
// memory used to poll for convergence every `poll` iterations.
int run_pcg(const DevCloth& c, double tol, int maxit, int* host_flags, hipStream_t s, int poll,
            Prof* prof, const PcgAux* aux) {
  const long n3 = 3L * c.np;
  const bool wg = pcg_wg_ok(c) && c.pcg_variant == 2;
  if (wg) {
    ProfScope ps(prof, kProfPcgWg);
    launch_pcg_wg(c.kpack, dim3(1), One{c, {}}, tol * tol, maxit, s);
    CS_HIP(hipGetLastError());
    return kOk;
  }
  {
    ProfScope ps(prof, kProfOther);
    hipLaunchKernelGGL(k_pcg_init<One>, dim3(grid_for(n3)), dim3(kTB), 0, s, One{c, {}}, tol * tol);
  }
  {
    ProfScope ps(prof, kProfReduce);
    launch_dot(c, kDotRr0, 0, s);
  }
  {
    ProfScope ps(prof, kProfSweep);
    launch_sweep(c, c.r, c.z, c.prod, s, true);
  }
  {
    ProfScope ps(prof, kProfReduce);
    launch_dot(c, kDotRz0, 0, s);
  }
  CS_HIP(hipGetLastError());
  bool done = false;
  int it = 0;
  for (; it < maxit && !done; it++) {
    {
      ProfScope ps(prof, kProfSpmv);
      hipLaunchKernelGGL(k_spmv_dir_g<One>, dim3(grid_for((long)kSpG * c.np)), dim3(kTB), 0, s,
                         One{c, {}}, it);
    }
    {
      ProfScope ps(prof, kProfReduce);
      launch_dot(c, kDotPap, it, s);
    }
    {
      ProfScope ps(prof, kProfAxpy);
      hipLaunchKernelGGL(k_update_rx<One>, dim3(grid_for(n3)), dim3(kTB), 0, s, One{c, {}}, it);
    }
    if (aux) {
      // (r, r) and the convergence test beside the sweep; every later kernel
      // of the solve checks the flag it sets, and x, r and the reported
      // ratio and count do not depend on z
      CS_HIP(hipEventRecord(aux->updated, s));
      CS_HIP(hipStreamWaitEvent(aux->stream, aux->updated, 0));
      {
        ProfScope ps(prof, kProfReduce, aux->stream);
        launch_dot(c, kDotRr, it, aux->stream);
      }
      CS_HIP(hipEventRecord(aux->tested, aux->stream));
    } else {
      ProfScope ps(prof, kProfReduce);
      launch_dot(c, kDotRr, it, s);
    }
    {
      ProfScope ps(prof, kProfSweep);
      launch_sweep(c, c.r, c.z, c.prod, s, true);
    }
    {
      ProfScope ps(prof, kProfReduce);
      launch_dot(c, kDotRz, it, s);
    }
    if (aux) CS_HIP(hipStreamWaitEvent(s, aux->tested, 0));
    CS_HIP(hipGetLastError());
    if ((it + 1) % poll == 0 || it + 1 == maxit) {
      CS_HIP(hipMemcpyAsync(host_flags, c.flags, sizeof(int) * kNumFlag, hipMemcpyDeviceToHost, s));
      CS_HIP(hipStreamSynchronize(s));
      done = host_flags[kDone] != 0;
    }
  }
  if (!done) {
    // cap reached: report the achieved ratio (ilu_sparse.cpp:706-707)
    hipLaunchKernelGGL(k_square<One>, dim3(grid_for(n3)), dim3(kTB), 0, s, One{c, {}});
    ProfScope ps(prof, kProfReduce);
    launch_dot(c, kDotFinal, maxit, s);
    CS_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_pcg_check<One>, dim3(1), dim3(64), 0, s, One{c, {}});
  CS_HIP(hipGetLastError());
  return kOk;
}

int launch_gradients(const DevCloth& c, const Params& P, hipStream_t s) {
  const One one{c, P};
  hipLaunchKernelGGL(k_tri_elem<One>, dim3(grid_for(c.nt)), dim3(kTB), 0, s, one, 0);
  if (c.nq > 0) hipLaunchKernelGGL(k_quad_grad<One>, dim3(grid_for(c.nq)), dim3(kTB), 0, s, one);
  CS_HIP(hipGetLastError());
  return kOk;
}

int launch_finish_internal(const DevCloth& c, const Params& P, hipStream_t s) {
  hipLaunchKernelGGL(k_finish_internal<One>, dim3(grid_for(3L * c.np)), dim3(kTB), 0, s, One{c, P});
  CS_HIP(hipGetLastError());
  return kOk;
}

}  // namespace cs

// ---------------------------------------------------------------------------
// batched internal step: the kernels above launched once for every cloth of a
// batch (grid row = cloth, Many selector).  Per cloth the arithmetic and the
// convergence handling are those of the single-cloth path, so each cloth's
// result is identical to stepping it alone.

namespace cs {

template <class Sel>
__global__ void k_save_start(Sel sel) {
  const DevCloth c = sel.cloth();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * c.np) return;
  c.x1[e] = c.x[e];
}

__global__ void k_gather_done(const DevCloth* tab, int n, int* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = tab[i].flags[kDone];
}

static void launch_sweep_batch(const BatchView& b, hipStream_t s) {
  const Many m{b.tab, b.ptab};
  if (b.piped) {
    // every cloth packed: r is already in slot order (k_pcg_init / k_update_rx)
    launch_quad(b.kpack, b.f32, b.max_qwidth, dim3(b.max_ncomp, b.n), m, nullptr, nullptr, nullptr, s);
  } else {
    hipLaunchKernelGGL(k_ilu_apply<Many>, dim3(1, b.n), dim3(1024), 0, s, m, nullptr, nullptr, nullptr);
  }
}

static void launch_dot_batch(const BatchView& b, int mode, int iter, hipStream_t s) {
  launch_ordered_sum(dot_src(Many{b.tab, b.ptab}, mode, iter), 3 * b.max_np, b.n, s);
}

int batch_internal_step(const BatchView& b, double tol, int maxit, int poll, int* dgather,
                        int* hgather, hipStream_t s, Prof* prof, const PcgAux* aux) {
  const Many m{b.tab, b.ptab};
  const int B = b.n;
  const unsigned gv = grid_for(3L * b.max_np), gp = grid_for(b.max_np);
  {
    ProfScope ps(prof, kProfAssemble);
    hipLaunchKernelGGL(k_save_start<Many>, dim3(gv, B), dim3(kTB), 0, s, m);
    hipLaunchKernelGGL(k_tri_elem<Many>, dim3(grid_for(b.max_nt), B), dim3(kTB), 0, s, m, 1);
    if (b.max_nq > 0)
      hipLaunchKernelGGL(k_quad_grad<Many>, dim3(grid_for(b.max_nq), B), dim3(kTB), 0, s, m);
    if (b.max_nnz > 0)
      hipLaunchKernelGGL(k_asm_crs<Many>, dim3(grid_for(b.max_nnz), B), dim3(kTB), 0, s, m);
    hipLaunchKernelGGL(k_asm_dia_rhs<Many>, dim3(gp, B), dim3(kTB), 0, s, m);
  }
  {
    ProfScope ps(prof, kProfFactor);
    if (b.max_fu_blocks > 0)
      hipLaunchKernelGGL(k_ilu_factor_staged<Many>, dim3(1, B),
                         dim3(ilu_stage_threads(b.max_fu_blocks)), 0, s, m, 1, b.max_fu_blocks);
    else
      hipLaunchKernelGGL(k_ilu_factor<Many>, dim3(1, B), dim3(1024), 0, s, m, 1);
    hipLaunchKernelGGL(k_ilu_redo<Many>, dim3(1, B), dim3(1024), 0, s, m);
    hipLaunchKernelGGL(k_pack_factor<Many>, dim3(gp, B), dim3(kTB), 0, s, m);
  }
  CS_HIP(hipGetLastError());
  if (b.wg) {
    // one workgroup per cloth runs its whole solve: no cloth waits for another
    {
      ProfScope ps(prof, kProfPcgWg);
      launch_pcg_wg(b.kpack, dim3(1, B), m, tol * tol, maxit, s);
    }
    hipLaunchKernelGGL(k_finish_internal<Many>, dim3(gv, B), dim3(kTB), 0, s, m);
    CS_HIP(hipGetLastError());
    return kOk;
  }
  {
    ProfScope ps(prof, kProfOther);
    hipLaunchKernelGGL(k_pcg_init<Many>, dim3(gv, B), dim3(kTB), 0, s, m, tol * tol);
  }
  {
    ProfScope ps(prof, kProfReduce);
    launch_dot_batch(b, kDotRr0, 0, s);
  }
  {
    ProfScope ps(prof, kProfSweep);
    launch_sweep_batch(b, s);
  }
  {
    ProfScope ps(prof, kProfReduce);
    launch_dot_batch(b, kDotRz0, 0, s);
  }
  CS_HIP(hipGetLastError());
  bool done = false;
  for (int it = 0; it < maxit && !done; it++) {
    {
      ProfScope ps(prof, kProfSpmv);
      hipLaunchKernelGGL(k_spmv_dir_g<ManyRows>, dim3(b.nrowblk), dim3(kTB), 0, s,
                         ManyRows{b.tab, b.ptab, b.rowmap}, it);
    }
    {
      ProfScope ps(prof, kProfReduce);
      launch_dot_batch(b, kDotPap, it, s);
    }
    {
      ProfScope ps(prof, kProfAxpy);
      hipLaunchKernelGGL(k_update_rx<ManyRows>, dim3(b.nvecblk), dim3(kTB), 0, s,
                         ManyRows{b.tab, b.ptab, b.vecmap}, it);
    }
    if (aux) {
      // as in run_pcg: the convergence test of every cloth beside the sweep
      CS_HIP(hipEventRecord(aux->updated, s));
      CS_HIP(hipStreamWaitEvent(aux->stream, aux->updated, 0));
      {
        ProfScope ps(prof, kProfReduce, aux->stream);
        launch_dot_batch(b, kDotRr, it, aux->stream);
      }
      CS_HIP(hipEventRecord(aux->tested, aux->stream));
    } else {
      ProfScope ps(prof, kProfReduce);
      launch_dot_batch(b, kDotRr, it, s);
    }
    {
      ProfScope ps(prof, kProfSweep);
      launch_sweep_batch(b, s);
    }
    {
      ProfScope ps(prof, kProfReduce);
      launch_dot_batch(b, kDotRz, it, s);
    }
    if (aux) CS_HIP(hipStreamWaitEvent(s, aux->tested, 0));
    CS_HIP(hipGetLastError());
    if ((it + 1) % poll == 0 || it + 1 == maxit) {
      hipLaunchKernelGGL(k_gather_done, dim3(grid_for(B)), dim3(kTB), 0, s, b.tab, B, dgather);
      CS_HIP(hipMemcpyAsync(hgather, dgather, sizeof(int) * B, hipMemcpyDeviceToHost, s));
      CS_HIP(hipStreamSynchronize(s));
      done = true;
      for (int i = 0; i < B; i++) done = done && hgather[i] != 0;
    }
  }
  if (!done) {
    // cloths at the cap report their achieved ratio; converged ones skip
    hipLaunchKernelGGL(k_square<Many>, dim3(gv, B), dim3(kTB), 0, s, m);
    ProfScope ps(prof, kProfReduce);
    launch_dot_batch(b, kDotFinal, maxit, s);
  }
  hipLaunchKernelGGL(k_pcg_check<Many>, dim3(1, B), dim3(64), 0, s, m);
  hipLaunchKernelGGL(k_finish_internal<Many>, dim3(gv, B), dim3(kTB), 0, s, m);
  CS_HIP(hipGetLastError());
  return kOk;
}

}  // namespace cs
