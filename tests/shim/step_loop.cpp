// Drives the C++ shim exactly like the reference's StepTime
// (self_contact_sparse/main.cpp:213-263): internal dynamics, contact-resolved
// velocities, and, when an impulse was applied, x = x1 + v dt followed by
// UpdateIntermidiateVelocity.
//
//   step_loop <input> <output>
// input (little endian): int32 np, nt, nq, steps, pcg_maxit;
//   float64 dt, lambda, myu, stiff_bend, gx, gy, gz, mass_point,
//   stiff_contact, clearance, pcg_tol, sphere cx, cy, cz, r;
//   float64 xyz0[3np], xyz[3np], uvw[3np]; int32 tri[3nt], quad[4nq], bc[np]
// output: float64 xyz[3np], uvw[3np] after `steps` steps
#include <cmath>
#include <cstdio>
#include <vector>

#include "clothsim_ref.hpp"

static double g_sphere[4];

// outward-normal sphere, pd > 0 inside (the form of internal_cloth_sparse/main.cpp:120-132)
static void penetrationDepth_Sphere(double& pd, double* n, const double* p) {
  n[0] = p[0] - g_sphere[0];
  n[1] = p[1] - g_sphere[1];
  n[2] = p[2] - g_sphere[2];
  const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  n[0] /= len;
  n[1] /= len;
  n[2] /= len;
  pd = g_sphere[3] - len;
}

template <typename T>
static bool rd(FILE* f, T* p, size_t n) {
  return std::fread(p, sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: step_loop <input> <output>\n");
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[5];
  double prm[15];
  if (!rd(f, hdr, 5) || !rd(f, prm, 15)) return 2;
  const int np = hdr[0], nt = hdr[1], nq = hdr[2], steps = hdr[3], maxit = hdr[4];
  std::vector<double> aXYZ0(3 * np), aXYZ(3 * np), aUVW(3 * np);
  std::vector<int> aTri(3 * nt), aQuad(4 * nq), aBCFlag(np);
  if (!rd(f, aXYZ0.data(), aXYZ0.size()) || !rd(f, aXYZ.data(), aXYZ.size()) ||
      !rd(f, aUVW.data(), aUVW.size()) || !rd(f, aTri.data(), aTri.size()) ||
      !rd(f, aQuad.data(), aQuad.size()) || !rd(f, aBCFlag.data(), aBCFlag.size()))
    return 2;
  std::fclose(f);
  const double dt = prm[0], lambda = prm[1], myu = prm[2], stiff_bend = prm[3];
  const double gravity[3] = {prm[4], prm[5], prm[6]};
  const double mass_point = prm[7], stiff_contact = prm[8], contact_clearance = prm[9];
  for (int i = 0; i < 4; i++) g_sphere[i] = prm[11 + i];

  clothsim::Cloth cloth(aXYZ0, aXYZ, aTri, aQuad, aBCFlag);
  cloth.set_pcg(prm[10], maxit);
  for (int k = 0; k < steps; k++) {
    std::vector<double> aXYZ1 = aXYZ;
    clothsim::StepTime_InternalDynamicsILU(aXYZ, aUVW, cloth, aXYZ0, aBCFlag, aTri, aQuad, dt,
                                           lambda, myu, stiff_bend, gravity, mass_point,
                                           stiff_contact, contact_clearance,
                                           penetrationDepth_Sphere);
    bool is_impulse_applied = false;
    clothsim::GetIntermidiateVelocityContactResolved(aUVW, is_impulse_applied, dt,
                                                     contact_clearance, mass_point,
                                                     stiff_contact, aXYZ1, aTri, cloth);
    if (is_impulse_applied) {
      for (int ip = 0; ip < np; ip++)
        for (int i = 0; i < 3; i++) aXYZ[ip * 3 + i] = aXYZ1[ip * 3 + i] + aUVW[ip * 3 + i] * dt;
      clothsim::UpdateIntermidiateVelocity(aUVW, aXYZ, aXYZ0, aBCFlag, aTri, aQuad, dt, lambda,
                                           myu, stiff_bend, gravity, mass_point, cloth);
    }
  }
  FILE* o = std::fopen(argv[2], "wb");
  if (!o) return 2;
  std::fwrite(aXYZ.data(), sizeof(double), aXYZ.size(), o);
  std::fwrite(aUVW.data(), sizeof(double), aUVW.size(), o);
  std::fclose(o);
  return 0;
}
