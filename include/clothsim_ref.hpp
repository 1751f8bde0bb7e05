// C++ host mirror of the reference's step-loop functions over the clothsim C
// ABI.  Same names, argument order and meaning as
//   StepTime_InternalDynamicsILU            solve_internal_sparse.h:255-275
//   GetIntermidiateVelocityContactResolved  self_collision_cloth.h:21-34
//   UpdateIntermidiateVelocity              solve_internal_sparse.h:329-343
// so the reference's StepTime (self_contact_sparse/main.cpp:213-263) keeps
// its shape.  The reference's host workspaces -- CMatrixSquareSparse mat_A,
// CPreconditionerILU ilu_A (first call), CJaggedArray aEdge, iroot_bvh,
// aNodeBVH, aBB (second call) -- become one clothsim::Cloth, which keeps the
// pattern, matrix, factor, BVH and 1-ring resident on the GPU; the third call
// takes it as an extra last argument.
//
// The penetrationDepth callback is evaluated on the host for every vertex at
// the step's positions and uploaded, exactly the values the reference's
// WdWddW_Contact would see.
//
// Like the reference the functions return nothing.  A call that fails prints
// the library's message and aborts (the reference asserts).  Singular ILU
// pivots and an exhausted PCG are not failures: the reference carries on
// with them too.  PCG runs with the reference's hard-coded tolerance 1e-4 and
// cap 100 (solve_internal_sparse.h:318-319) unless Cloth::set_pcg says
// otherwise.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "clothsim.h"

namespace clothsim {

inline void check(int status, const char* what) {
  if (status == CS_OK || status == CS_ERR_SINGULAR || status == CS_ERR_NOT_CONVERGED) return;
  std::fprintf(stderr, "clothsim: %s failed (status %d): %s\n", what, status, cs_last_error());
  std::abort();
}

class Cloth {
 public:
  // the mesh of the reference's setup (self_contact_sparse/main.cpp:561-583);
  // aXYZ seeds the BVH topology like MakeBVHTopology_TopDown
  Cloth(const std::vector<double>& aXYZ0, const std::vector<double>& aXYZ,
        const std::vector<int>& aTri, const std::vector<int>& aQuad,
        const std::vector<int>& aBCFlag) {
    np_ = (int)aXYZ0.size() / 3;
    cs_mesh m{};
    m.np = np_;
    m.nt = (int)aTri.size() / 3;
    m.nq = (int)aQuad.size() / 4;
    m.xyz0 = aXYZ0.data();
    m.xyz = aXYZ.data();
    m.uvw = nullptr;
    m.tri = aTri.data();
    m.quad = aQuad.data();
    m.bc = aBCFlag.empty() ? nullptr : aBCFlag.data();
    p_.lambda = 1.0;
    p_.myu = 1.0;
    p_.mass_point = 1.0;
    p_.dt = 0.01;
    p_.pcg_tol = 1.0e-4;
    p_.pcg_maxit = 100;
    p_.obstacle_kind = CS_OBST_NONE;
    check(cs_cloth_create(&m, &p_, &h_), "cs_cloth_create");
  }
  ~Cloth() {
    if (h_) cs_cloth_destroy(h_);
  }
  Cloth(const Cloth&) = delete;
  Cloth& operator=(const Cloth&) = delete;

  cs_cloth* handle() const { return h_; }
  int np() const { return np_; }
  void set_pcg(double tol, int maxit) {
    p_.pcg_tol = tol;
    p_.pcg_maxit = maxit;
  }
  // what the last internal step's Solve_PCG reported
  double conv_ratio() const { return conv_; }
  int iteration() const { return iter_; }

  // internal: parameters of one call
  cs_params& params() { return p_; }
  void apply() { check(cs_cloth_set_params(h_, &p_), "cs_cloth_set_params"); }
  void record(double conv, int iter) {
    conv_ = conv;
    iter_ = iter;
  }
  // which CPreconditionerILU's factor the device holds now (clothsim_ref_ops.hpp);
  // the internal step refactors on the device and clears it
  const void* factor_owner = nullptr;

 private:
  cs_cloth* h_ = nullptr;
  cs_params p_{};
  int np_ = 0;
  double conv_ = 0.0;
  int iter_ = 0;
};

// solve_internal_sparse.h:255-275, with (mat_A, ilu_A) -> cloth
inline void StepTime_InternalDynamicsILU(
    std::vector<double>& aXYZ, std::vector<double>& aUVW, Cloth& cloth,
    const std::vector<double>& aXYZ0, const std::vector<int>& aBCFlag,
    const std::vector<int>& aTri, const std::vector<int>& aQuad, const double dt,
    double lambda, double myu, double stiff_bend, const double gravity[3], double mass_point,
    double stiff_contact, double contact_clearance,
    void (*penetrationDepth)(double&, double*, const double*)) {
  (void)aXYZ0, (void)aBCFlag, (void)aTri, (void)aQuad;  // resident since construction
  cs_params& p = cloth.params();
  p.lambda = lambda;
  p.myu = myu;
  p.stiff_bend = stiff_bend;
  for (int i = 0; i < 3; i++) p.gravity[i] = gravity[i];
  p.mass_point = mass_point;
  p.stiff_contact = stiff_contact;
  p.contact_clearance = contact_clearance;
  p.dt = dt;
  p.obstacle_kind = CS_OBST_NONE;
  cloth.apply();
  const int np = cloth.np();
  std::vector<double> pd(np), n(3 * (size_t)np);
  for (int ip = 0; ip < np; ip++) penetrationDepth(pd[ip], &n[3 * ip], &aXYZ[3 * ip]);
  check(cs_set_obstacle_values(cloth.handle(), pd.data(), n.data()), "cs_set_obstacle_values");
  check(cs_set_state(cloth.handle(), aXYZ.data(), aUVW.data()), "cs_set_state");
  double conv = 0.0;
  int iter = 0;
  check(cs_step_internal(cloth.handle(), &conv, &iter), "cs_step_internal");
  cloth.factor_owner = nullptr;
  cloth.record(conv, iter);
  check(cs_get_state(cloth.handle(), aXYZ.data(), aUVW.data()), "cs_get_state");
}

// self_collision_cloth.h:21-34, with (aEdge, iroot_bvh, aNodeBVH, aBB) -> cloth
inline void GetIntermidiateVelocityContactResolved(
    std::vector<double>& aUVWm, bool& is_impulse_applied, double dt, double contact_clearance,
    double mass_point, double cloth_contact_stiffness, const std::vector<double>& aXYZ,
    const std::vector<int>& aTri, Cloth& cloth) {
  (void)aTri;
  cs_params& p = cloth.params();
  p.dt = dt;
  p.contact_clearance = contact_clearance;
  p.mass_point = mass_point;
  p.stiff_contact = cloth_contact_stiffness;
  cloth.apply();
  check(cs_set_step_start(cloth.handle(), aXYZ.data()), "cs_set_step_start");
  check(cs_set_state(cloth.handle(), nullptr, aUVWm.data()), "cs_set_state");
  int applied = 0;
  check(cs_contact_resolve(cloth.handle(), &applied, nullptr), "cs_contact_resolve");
  is_impulse_applied = applied != 0;
  check(cs_get_state(cloth.handle(), nullptr, aUVWm.data()), "cs_get_state");
}

// solve_internal_sparse.h:329-343, plus the cloth
inline void UpdateIntermidiateVelocity(
    std::vector<double>& aUVW, const std::vector<double>& aXYZ,
    const std::vector<double>& aXYZ0, const std::vector<int>& aBCFlag,
    const std::vector<int>& aTri, const std::vector<int>& aQuad, const double dt,
    double lambda, double myu, double stiff_bend, const double gravity[3], double mass_point,
    Cloth& cloth) {
  (void)aXYZ0, (void)aBCFlag, (void)aTri, (void)aQuad;
  cs_params& p = cloth.params();
  p.lambda = lambda;
  p.myu = myu;
  p.stiff_bend = stiff_bend;
  for (int i = 0; i < 3; i++) p.gravity[i] = gravity[i];
  p.mass_point = mass_point;
  p.dt = dt;
  cloth.apply();
  check(cs_set_state(cloth.handle(), aXYZ.data(), aUVW.data()), "cs_set_state");
  check(cs_update_velocity(cloth.handle()), "cs_update_velocity");
  check(cs_get_state(cloth.handle(), nullptr, aUVW.data()), "cs_get_state");
}

}  // namespace clothsim
