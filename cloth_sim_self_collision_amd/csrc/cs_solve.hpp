// Host launchers of the implicit internal solve (cs_solve.hip).
#pragma once

#include "cs_device.hpp"

namespace cs {

int launch_bend_setup(const DevCloth& c, const Params& P, hipStream_t s);
// element kernels + gather assembly of A and b (after boundary conditions)
int launch_assemble(const DevCloth& c, const Params& P, hipStream_t s, Prof* prof = nullptr);
// copy A into the factor storage and factorise in place
int launch_ilu(const DevCloth& c, int* ising, hipStream_t s, Prof* prof = nullptr);
// rearrange the factor for the packed sweeps (after any change of fdia/fcrs)
int launch_pack_factor(const DevCloth& c, hipStream_t s);
int launch_matvec(const DevCloth& c, double alpha, const double* x, double beta, double* y,
                  hipStream_t s);
int launch_ilu_apply(const DevCloth& c, const double* src, double* dst, hipStream_t s);
// Solve_PCG from dx = 0 with r = b; poll convergence every `poll` iterations
int run_pcg(const DevCloth& c, double tol, int maxit, int* host_flags, hipStream_t s, int poll,
            Prof* prof = nullptr);
// x[v] = rest[v] + off for the swept vertices
int launch_apply_sweep(const DevCloth& c, const int* verts, int n, const double off[3],
                       hipStream_t s);
// membrane and bending gradients only (UpdateIntermidiateVelocity)
int launch_gradients(const DevCloth& c, const Params& P, hipStream_t s);
int launch_finish_internal(const DevCloth& c, const Params& P, hipStream_t s);

}  // namespace cs
