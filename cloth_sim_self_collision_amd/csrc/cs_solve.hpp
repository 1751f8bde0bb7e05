// Host launchers of the implicit internal solve (cs_solve.hip).
#pragma once

#include "cs_device.hpp"

namespace cs {

// the packed (quad) sweep applies: packed factor, LDS ring proven safe,
// sub-level pointers fit
// the quad sweep addresses its operands with 32-bit byte offsets (at most
// 3 * np 16-byte pairs per array)
inline bool quad_ok(const DevCloth& c) {
  return c.kpack > 0 && c.ring_ok &&
         48LL * c.np < (1LL << 32);
}

// the quad sweep's workgroup waves and vector loads per wave per forward /
// backward sub-level; false when the row-by-row sweep runs instead
bool sweep_shape(const DevCloth& c, int* waves, int* loads_fwd, int* loads_bwd);
// blocks of a staged ILU row (the widest row plus the diagonal) that the
// staged factorisation can hold, 0 if a row is too wide for it
int ilu_stage_blocks(int widest_row);
int launch_bend_setup(const DevCloth& c, const Params& P, hipStream_t s);
// element kernels + gather assembly of A and b (after boundary conditions)
int launch_assemble(const DevCloth& c, const Params& P, hipStream_t s, Prof* prof = nullptr);
// copy A into the factor storage and factorise in place
int launch_ilu(const DevCloth& c, hipStream_t s, Prof* prof = nullptr);
// rearrange the factor for the packed sweeps (after any change of fdia/fcrs)
int launch_pack_factor(const DevCloth& c, hipStream_t s);
int launch_matvec(const DevCloth& c, double alpha, const double* x, double beta, double* y,
                  hipStream_t s);
int launch_ilu_apply(const DevCloth& c, const double* src, double* dst, hipStream_t s);
// the one-workgroup PCG solve (k_pcg_wg) applies to this cloth
bool pcg_wg_ok(const DevCloth& c);
// Solve_PCG from dx = 0 with r = b; poll convergence every `poll` iterations
// (the one-workgroup solve needs no poll)
int run_pcg(const DevCloth& c, double tol, int maxit, int* host_flags, hipStream_t s, int poll,
            Prof* prof = nullptr, const PcgAux* aux = nullptr);
// x[v] = rest[v] + off for the swept vertices
int launch_apply_sweep(const DevCloth& c, const int* verts, int n, const double off[3],
                       hipStream_t s);
// membrane and bending gradients only (UpdateIntermidiateVelocity)
int launch_gradients(const DevCloth& c, const Params& P, hipStream_t s);
int launch_finish_internal(const DevCloth& c, const Params& P, hipStream_t s);

// a batch of cloths as the batched kernels see it: device tables of the
// cloths' DevCloth and Params, and the largest sizes (grid extents)
struct BatchView {
  const DevCloth* tab;
  const Params* ptab;
  int n;
  int max_np, max_nt, max_nq, max_nnz, max_wmax;
  int max_qwidth, max_ncomp;  // widest sweep sub-level, most components of a cloth
  int max_fu_blocks;  // staged factorisation: blocks per row slot (0: some cloth cannot)
  bool piped;  // every cloth can use the packed (quad) sweep ...
  int kpack;   // ... with this many packed blocks per row
  bool wg;     // every cloth takes the one-workgroup PCG solve (one workgroup per cloth)
  bool f32;    // every cloth's packed sweep coefficients are FP32 (mixed precision)
  const int2* rowmap;  // the SpMV's blocks: (cloth, first row), kSpmvRows rows each
  int nrowblk;         // entries of rowmap
  const int2* vecmap;  // the vector update's blocks: (cloth, first of 3 np entries)
  int nvecblk;         // entries of vecmap
};
// rows per block of the batched SpMV and entries per block of the batched
// vector update (their block maps are built with these steps)
constexpr int kSpmvRows = 16;
constexpr int kVecBlock = 256;
// StepTime_InternalDynamicsILU for every cloth of the batch (start positions
// saved, assembly, ILU(0), PCG, x += dx, v = dx/dt); dgather/hgather hold n
// ints (device / pinned host) for the convergence poll; with aux, the (r, r)
// reductions and tests run on aux->stream beside the sweeps
int batch_internal_step(const BatchView& b, double tol, int maxit, int poll, int* dgather,
                        int* hgather, hipStream_t s, Prof* prof = nullptr,
                        const PcgAux* aux = nullptr);

}  // namespace cs
