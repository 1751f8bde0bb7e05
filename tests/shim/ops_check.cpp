// Drives the reference-named lower-level API (include/clothsim_ref_ops.hpp)
// the way the reference's StepTime_InternalDynamicsILU body drives its own
// classes (solve_internal_sparse.h:277-325): the pattern handed to
// SetPattern, the system assembled element by element with WdWddW_CST /
// WdWddW_Bend / WdWddW_Contact and Mearge, the mass term poked into m_valDia,
// SetBoundaryCondition, SetValueILU + DoILUDecomp, Solve_PCG; then MatVec,
// CPreconditionerILU::Solve, the BVH topology, both refits and both contact
// queries.  Writes every result for the test to compare with the oracle.
//
//   ops_check <input> <output>
// input (little endian): int32 np, nt, nq, pcg_maxit, obstacle (0 none,
//   2 sphere); float64 lambda, myu, stiff_bend, gx, gy, gz, mass_point,
//   stiff_contact, clearance, dt, pcg_tol, sphere cx, cy, cz, r;
//   float64 xyz0[3np], xyz_setup[3np] (the positions the setup saw: they
//   seed the BVH topology), xyz[3np], uvw[3np]; int32 tri[3nt], quad[4nq],
//   bc[np]
// output: see the writes at the end (the test reads them in the same order)
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <set>
#include <vector>

#include "clothsim_ref_ops.hpp"

using namespace clothsim;

static double g_sphere[4];
static int g_obstacle = 0;

// the sphere of CS_OBST_SPHERE, pd > 0 inside; no obstacle: far outside
static void penetrationDepth(double& pd, double* n, const double* p) {
  if (g_obstacle != 2) {
    pd = -100;
    n[0] = n[1] = n[2] = 0;
    return;
  }
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
static bool rd(FILE* f, std::vector<T>& v) {
  return std::fread(v.data(), sizeof(T), v.size(), f) == v.size();
}
template <typename T>
static void wr(FILE* f, const T* p, size_t n) {
  std::fwrite(p, sizeof(T), n, f);
}
template <typename T>
static void wr(FILE* f, const std::vector<T>& v) {
  wr(f, v.data(), v.size());
}

// the vertex adjacency of the bending stencils, self excluded, columns
// ascending: the pattern the reference's setup hands to SetPattern
static void stencil_pattern(int np, const std::vector<int>& quad, std::vector<int>& offs,
                            std::vector<int>& cols) {
  std::vector<std::vector<int>> adj(np);
  for (size_t q = 0; q + 3 < quad.size(); q += 4)
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++)
        if (a != b) adj[quad[q + a]].push_back(quad[q + b]);
  offs.assign(np + 1, 0);
  cols.clear();
  for (int i = 0; i < np; i++) {
    std::sort(adj[i].begin(), adj[i].end());
    adj[i].erase(std::unique(adj[i].begin(), adj[i].end()), adj[i].end());
    cols.insert(cols.end(), adj[i].begin(), adj[i].end());
    offs[i + 1] = (int)cols.size();
  }
}

static void write_set(FILE* o, const std::set<CContactElement>& s) {
  const int n = (int)s.size();
  wr(o, &n, 1);
  for (const CContactElement& e : s) {
    const int r[5] = {e.is_fv ? 1 : 0, e.ino0, e.ino1, e.ino2, e.ino3};
    wr(o, r, 5);
  }
}

static void write_boxes(FILE* o, const std::vector<CAABB3D>& bb) {
  for (const CAABB3D& b : bb) {
    const double r[6] = {b.x_min, b.x_max, b.y_min, b.y_max, b.z_min, b.z_max};
    wr(o, r, 6);
  }
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: ops_check <input> <output>\n");
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  std::vector<int> hdr(5);
  std::vector<double> prm(15);
  if (!rd(f, hdr) || !rd(f, prm)) return 2;
  const int np = hdr[0], nt = hdr[1], nq = hdr[2], maxit = hdr[3];
  g_obstacle = hdr[4];
  std::vector<double> aXYZ0(3 * np), aXYZ_setup(3 * np), aXYZ(3 * np), aUVW(3 * np);
  std::vector<int> aTri(3 * nt), aQuad(4 * nq), aBCFlag(np);
  if (!rd(f, aXYZ0) || !rd(f, aXYZ_setup) || !rd(f, aXYZ) || !rd(f, aUVW) || !rd(f, aTri) ||
      !rd(f, aQuad) ||
      !rd(f, aBCFlag))
    return 2;
  std::fclose(f);
  const double lambda = prm[0], myu = prm[1], stiff_bend = prm[2];
  const double gravity[3] = {prm[3], prm[4], prm[5]};
  const double mass_point = prm[6], stiff_contact = prm[7], clearance = prm[8], dt = prm[9];
  const double tol = prm[10];
  for (int i = 0; i < 4; i++) g_sphere[i] = prm[11 + i];

  Cloth cloth(aXYZ0, aXYZ_setup, aTri, aQuad, aBCFlag);

  // setup: the matrix and the preconditioner
  CMatrixSquareSparse mat_A(cloth);
  mat_A.Initialize(np, 3);
  std::vector<int> offs, cols;
  stencil_pattern(np, aQuad, offs, cols);
  std::vector<int> bad_cols = cols;
  if (!bad_cols.empty()) bad_cols[0] = (bad_cols[0] + 1) % np;
  const int rejected = mat_A.SetPattern(offs, bad_cols) ? 0 : 1;
  const int accepted = mat_A.SetPattern(offs, cols) ? 1 : 0;
  CPreconditionerILU ilu_A;
  ilu_A.Initialize_ILU0(mat_A);

  // the system, element by element
  double W = 0;
  std::vector<double> vec_b(3 * (size_t)np, 0.0);
  mat_A.SetZero();
  std::vector<int> buf(np, -1);
  for (int t = 0; t < nt; t++) {
    const int ip[3] = {aTri[3 * t], aTri[3 * t + 1], aTri[3 * t + 2]};
    double C[3][3], c[3][3];
    for (int a = 0; a < 3; a++)
      for (int i = 0; i < 3; i++) C[a][i] = aXYZ0[3 * ip[a] + i], c[a][i] = aXYZ[3 * ip[a] + i];
    double e, de[3][3], dde[3][3][3][3];
    WdWddW_CST(e, de, dde, C, c, lambda, myu);
    W += e;
    for (int a = 0; a < 3; a++)
      for (int i = 0; i < 3; i++) vec_b[3 * ip[a] + i] += de[a][i];
    mat_A.Mearge(3, ip, 3, ip, 9, &dde[0][0][0][0], buf);
  }
  for (int q = 0; q < nq; q++) {
    const int ip[4] = {aQuad[4 * q], aQuad[4 * q + 1], aQuad[4 * q + 2], aQuad[4 * q + 3]};
    double C[4][3], c[4][3];
    for (int a = 0; a < 4; a++)
      for (int i = 0; i < 3; i++) C[a][i] = aXYZ0[3 * ip[a] + i], c[a][i] = aXYZ[3 * ip[a] + i];
    double e = 0, de[4][3], dde[4][4][3][3];
    WdWddW_Bend(e, de, dde, C, c, stiff_bend);
    W += e;
    for (int a = 0; a < 4; a++)
      for (int i = 0; i < 3; i++) vec_b[3 * ip[a] + i] += de[a][i];
    mat_A.Mearge(4, ip, 4, ip, 9, &dde[0][0][0][0], buf);
  }
  for (int p = 0; p < np; p++) {
    double e, de[3], dde[3][3];
    WdWddW_Contact(e, de, dde, &aXYZ[3 * p], stiff_contact, clearance, penetrationDepth);
    W += e;
    for (int i = 0; i < 3; i++) vec_b[3 * p + i] += de[i];
    mat_A.Mearge(1, &p, 1, &p, 9, &dde[0][0], buf);
  }
  for (int p = 0; p < np; p++)
    for (int i = 0; i < 3; i++) vec_b[3 * p + i] -= mass_point * gravity[i];
  for (int i = 0; i < 3 * np; i++) vec_b[i] = -vec_b[i] + mass_point * aUVW[i] / dt;
  for (int p = 0; p < np; p++)
    for (int i = 0; i < 3; i++) mat_A.m_valDia[9 * p + 4 * i] += mass_point / (dt * dt);
  mat_A.SetBoundaryCondition(aBCFlag);
  for (int p = 0; p < np; p++)
    if (aBCFlag[p] != 0) vec_b[3 * p] = vec_b[3 * p + 1] = vec_b[3 * p + 2] = 0;

  ilu_A.SetValueILU(mat_A);
  ilu_A.DoILUDecomp();
  std::vector<double> b0 = vec_b, vec_x;
  double conv_ratio = tol;
  int iteration = maxit;
  Solve_PCG(conv_ratio, iteration, mat_A, ilu_A, vec_b, vec_x);

  // the operators on their own
  std::vector<double> y = b0;
  mat_A.MatVec(1.5, vec_x, 0.5, y);
  std::vector<double> z = b0;
  ilu_A.Solve(z);

  // BVH and contacts: proximity at the positions, CCD over the solve's velocities
  std::vector<CNodeBVH> aNodeBVH;
  const int iroot = MakeBVHTopology_TopDown(aTri, aXYZ_setup, aNodeBVH, cloth);
  std::vector<CAABB3D> bb_prx, bb_ccd;
  BuildBoundingBoxChild_Prx(iroot, clearance, aXYZ, aTri, aNodeBVH, bb_prx, cloth);
  std::set<CContactElement> prx;
  GetContactElement_Proximity(prx, clearance, aXYZ, aTri, iroot, aNodeBVH, bb_prx, cloth);
  std::vector<double> uvw_m(3 * (size_t)np);
  for (int i = 0; i < 3 * np; i++) uvw_m[i] = vec_x[i] / dt;
  BuildBoundingBoxChild_CCD(iroot, dt, aXYZ, uvw_m, aTri, aNodeBVH, bb_ccd, cloth);
  std::set<CContactElement> ccd;
  GetContactElement_CCD(ccd, dt, clearance, aXYZ, uvw_m, aTri, iroot, aNodeBVH, bb_ccd, cloth);

  FILE* o = std::fopen(argv[2], "wb");
  if (!o) return 2;
  const int flags[4] = {rejected, accepted, iteration, ilu_A.n_singular};
  wr(o, flags, 4);
  wr(o, &conv_ratio, 1);
  wr(o, &W, 1);
  wr(o, mat_A.m_valDia, 9 * (size_t)np);
  wr(o, mat_A.m_valCrs, 9 * (size_t)mat_A.m_ncrs);
  wr(o, b0);
  wr(o, ilu_A.mat.m_valDia, 9 * (size_t)np);
  wr(o, ilu_A.mat.m_valCrs, 9 * (size_t)mat_A.m_ncrs);
  wr(o, ilu_A.m_diaInd, (size_t)np);
  wr(o, vec_x);
  wr(o, vec_b);
  wr(o, y);
  wr(o, z);
  const int nnode = (int)aNodeBVH.size();
  wr(o, &nnode, 1);
  for (const CNodeBVH& n : aNodeBVH) {
    const int r[3] = {n.iroot, n.ichild[0], n.ichild[1]};
    wr(o, r, 3);
  }
  write_boxes(o, bb_prx);
  write_boxes(o, bb_ccd);
  write_set(o, prx);
  write_set(o, ccd);
  std::fclose(o);
  return 0;
}
