// Test-only harness around the compiled reference (nobuyuki83/cloth_sim_self_collision).
//
// This file is test infrastructure: it is linked with the reference's own
// library sources (matrix_square_sparse.cpp, ilu_sparse.cpp, bvh_aabb.cpp,
// self_collision_cloth.cpp) and its header-only physics
// (cloth_internal_physics.h, solve_internal_sparse.h) by oracle/Makefile, and
// exposes them behind a flat extern "C" API so that golden vectors can be
// written and the C restatement in oracle/ can be checked call for call.
// Nothing in the product links it.
//
// Setup mirrors self_contact_sparse/main.cpp:561-583; the step mirrors
// StepTime at main.cpp:213-263.  `ref_step_internal_tol` replays the body of
// StepTime_InternalDynamicsILU (solve_internal_sparse.h:255-326) with the
// PCG tolerance/cap as arguments (the reference hard-codes 1e-4 / 100 there);
// `ref_step_internal_native` calls the reference function itself so the
// replay can be checked against it bit for bit.
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <new>
#include <set>
#include <sstream>
#include <streambuf>
#include <vector>

#include "solve_internal_sparse.h"
#include "self_collision_cloth.h"
#include "bvh_aabb.h"
#include "jagged_array.h"

namespace {

class NullBuf : public std::streambuf {
 protected:
  int overflow(int c) override { return c; }
};
NullBuf g_null;
std::streambuf* g_saved = nullptr;
struct Quiet {
  Quiet() { g_saved = std::cout.rdbuf(&g_null); }
  ~Quiet() { std::cout.rdbuf(g_saved); }
};

// obstacle for the penetrationDepth callback (set per call)
double g_obst[9];

void pd_none(double& pd, double* n, const double* p) {
  (void)p;
  pd = -100;
  n[0] = n[1] = n[2] = 0;
}
// pd = d - dot(nrm, p): the form of penetrationDepth_Plane (main.cpp:110-114)
void pd_plane(double& pd, double* n, const double* p) {
  n[0] = g_obst[1];
  n[1] = g_obst[2];
  n[2] = g_obst[3];
  pd = g_obst[4] - (n[0] * p[0] + n[1] * p[1] + n[2] * p[2]);
}
// the form of penetrationDepth_Sphere in internal_cloth_sparse/main.cpp:120-132
void pd_sphere(double& pd, double* n, const double* p) {
  n[0] = p[0] - g_obst[1];
  n[1] = p[1] - g_obst[2];
  n[2] = p[2] - g_obst[3];
  double len = Length3D(n);
  n[0] /= len;
  n[1] /= len;
  n[2] /= len;
  pd = g_obst[4] - len;
}
// the sign-flipped sphere of self_contact_sparse/main.cpp:117-133
void pd_sphere_inv(double& pd, double* n, const double* p) {
  pd_sphere(pd, n, p);
  n[0] *= -1;
  n[1] *= -1;
  n[2] *= -1;
  pd *= -1;
}

// axis-aligned box (center g_obst[1..3], half-extents g_obst[4..6]); the
// scene convention of the other callbacks (pd > 0 inside, n push-out)
void pd_box(double& pd, double* n, const double* p) {
  double s[3], d[3];
  for (int i = 0; i < 3; i++) {
    s[i] = p[i] - g_obst[1 + i];
    d[i] = std::fabs(s[i]) - g_obst[4 + i];
  }
  if (d[0] <= 0.0 && d[1] <= 0.0 && d[2] <= 0.0) {
    int m = 0;
    if (d[1] > d[m]) m = 1;
    if (d[2] > d[m]) m = 2;
    n[0] = n[1] = n[2] = 0.0;
    n[m] = s[m] < 0.0 ? -1.0 : 1.0;
    pd = -d[m];
  } else {
    double o[3];
    for (int i = 0; i < 3; i++) o[i] = d[i] > 0.0 ? d[i] : 0.0;
    const double len = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
    for (int i = 0; i < 3; i++) n[i] = (s[i] < 0.0 ? -o[i] : o[i]) / len;
    pd = -len;
  }
}

// bar: the capsule around the segment g_obst[1..3]-g_obst[4..6] with radius
// g_obst[7]; with the ground z = g_obst[8] (normal +z) the deeper of the two,
// the bar on ties (same convention)
void pd_bar(double& pd, double* n, const double* p) {
  double e[3], w[3];
  for (int i = 0; i < 3; i++) {
    e[i] = g_obst[4 + i] - g_obst[1 + i];
    w[i] = p[i] - g_obst[1 + i];
  }
  const double ee = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
  double t = (w[0] * e[0] + w[1] * e[1] + w[2] * e[2]) / ee;
  t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
  for (int i = 0; i < 3; i++) n[i] = w[i] - t * e[i];
  const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  n[0] /= len;
  n[1] /= len;
  n[2] /= len;
  pd = g_obst[7] - len;
}
void pd_bar_ground(double& pd, double* n, const double* p) {
  pd_bar(pd, n, p);
  const double pg = g_obst[8] - p[2];
  if (pg > pd) {
    n[0] = 0.0;
    n[1] = 0.0;
    n[2] = 1.0;
    pd = pg;
  }
}

typedef void (*PdFn)(double&, double*, const double*);
PdFn pick_pd(const double* obst) {
  for (int i = 0; i < 9; i++) g_obst[i] = obst[i];
  switch ((int)obst[0]) {
    case 1: return pd_plane;
    case 2: return pd_sphere;
    case 3: return pd_sphere_inv;
    case 4: return pd_box;
    case 5: return pd_bar;
    case 6: return pd_bar_ground;
    default: return pd_none;
  }
}

struct RefCtx {
  int np, nt, nq;
  std::vector<double> xyz0;
  std::vector<int> tri, quad, bc;
  double lam, mu, stiff_bend, grav[3], mass_point, stiff_contact, clearance, dt;
  double tol;
  int maxit;
  double obst[9];
  CMatrixSquareSparse mat_A;
  CPreconditionerILU ilu_A;
  int iroot_bvh;
  std::vector<CNodeBVH> aNodeBVH;
  std::vector<CAABB3D> aBB;
  CJaggedArray aEdge;
  // last step diagnostics
  double last_conv;
  int last_iter;
};

void load_matrix(CMatrixSquareSparse& m, const double* dia, const double* crs) {
  std::memcpy(m.m_valDia, dia, sizeof(double) * 9 * m.m_nblk);
  std::memcpy(m.m_valCrs, crs, sizeof(double) * 9 * m.m_ncrs);
}

void contact_list(const std::set<CContactElement>& s, int* out, int cap, int* n) {
  int k = 0;
  for (std::set<CContactElement>::const_iterator it = s.begin(); it != s.end(); ++it) {
    if (k < cap) {
      out[k * 5 + 0] = it->is_fv ? 1 : 0;
      out[k * 5 + 1] = it->ino0;
      out[k * 5 + 2] = it->ino1;
      out[k * 5 + 3] = it->ino2;
      out[k * 5 + 4] = it->ino3;
    }
    k++;
  }
  *n = k;
}

// the body of StepTime_InternalDynamicsILU (solve_internal_sparse.h:277-325)
// with tolerance and cap as arguments
void internal_step(RefCtx* c, std::vector<double>& aXYZ, std::vector<double>& aUVW, double tol,
                   int maxit, std::vector<double>* b_out) {
  PdFn pd = pick_pd(c->obst);
  const int np = c->np;
  const int nDof = np * 3;
  double W = 0;
  std::vector<double> vec_b(nDof, 0);
  c->mat_A.SetZero();
  std::vector<int> tmp_buffer(np, -1);
  AddWdWddW_Cloth(W, vec_b, c->mat_A, tmp_buffer, aXYZ, c->xyz0, c->tri, c->quad, c->lam, c->mu,
                  c->stiff_bend);
  AddWdWddW_Contact(W, vec_b, c->mat_A, tmp_buffer, aXYZ, c->stiff_contact, c->clearance, pd);
  AddWdW_Gravity(W, vec_b, aXYZ, c->grav, c->mass_point);
  for (int i = 0; i < nDof; i++) vec_b[i] = -vec_b[i] + c->mass_point * aUVW[i] / c->dt;
  for (int ip = 0; ip < np; ip++) {
    c->mat_A.m_valDia[ip * 9 + 0 * 3 + 0] += c->mass_point / (c->dt * c->dt);
    c->mat_A.m_valDia[ip * 9 + 1 * 3 + 1] += c->mass_point / (c->dt * c->dt);
    c->mat_A.m_valDia[ip * 9 + 2 * 3 + 2] += c->mass_point / (c->dt * c->dt);
  }
  c->mat_A.SetBoundaryCondition(c->bc);
  for (int ip = 0; ip < np; ip++) {
    if (c->bc[ip] == 0) continue;
    vec_b[ip * 3 + 0] = 0;
    vec_b[ip * 3 + 1] = 0;
    vec_b[ip * 3 + 2] = 0;
  }
  if (b_out) {
    *b_out = vec_b;
    return;
  }
  c->ilu_A.SetValueILU(c->mat_A);
  c->ilu_A.DoILUDecomp();
  std::vector<double> vec_x;
  double conv_ratio = tol;
  int iteration = maxit;
  Solve_PCG(conv_ratio, iteration, c->mat_A, c->ilu_A, vec_b, vec_x);
  c->last_conv = conv_ratio;
  c->last_iter = iteration;
  for (int i = 0; i < nDof; i++) aXYZ[i] += vec_x[i];
  for (int i = 0; i < nDof; i++) aUVW[i] = vec_x[i] / c->dt;
}

}  // namespace

extern "C" {

// params: lam, mu, stiff_bend, gx, gy, gz, mass_point, stiff_contact, clearance, dt, tol, maxit
void* ref_create(int np, const double* xyz0, const double* xyz_init, int nt, const int* tri, int nq,
                 const int* quad, const int* bc, const double* params, const double* obst) {
  Quiet q;
  // zero-filled storage: CPreconditionerILU's constructor leaves m_diaInd
  // unset and Initialize_ILU0 deletes it when non-zero (ilu_sparse.cpp:102-121)
  void* mem = std::calloc(1, sizeof(RefCtx));
  RefCtx* c = new (mem) RefCtx();
  c->np = np;
  c->nt = nt;
  c->nq = nq;
  c->xyz0.assign(xyz0, xyz0 + 3 * np);
  c->tri.assign(tri, tri + 3 * nt);
  c->quad.assign(quad, quad + 4 * nq);
  c->bc.assign(bc, bc + np);
  c->lam = params[0];
  c->mu = params[1];
  c->stiff_bend = params[2];
  c->grav[0] = params[3];
  c->grav[1] = params[4];
  c->grav[2] = params[5];
  c->mass_point = params[6];
  c->stiff_contact = params[7];
  c->clearance = params[8];
  c->dt = params[9];
  c->tol = params[10];
  c->maxit = (int)params[11];
  for (int i = 0; i < 9; i++) c->obst[i] = obst[i];
  // main.cpp:574-582
  std::vector<double> xinit(xyz_init, xyz_init + 3 * np);
  c->iroot_bvh = MakeBVHTopology_TopDown(c->tri, xinit, c->aNodeBVH);
  c->aEdge.SetEdgeOfElem(c->tri, nt, 3, np, false);
  c->mat_A.Initialize(np, 3);
  CJaggedArray crs;
  crs.SetEdgeOfElem(c->quad, nq, 4, np, false);
  crs.Sort();
  c->mat_A.SetPattern(crs.index, crs.array);
  c->ilu_A.Initialize_ILU0(c->mat_A);
  return c;
}

void ref_destroy(void* p) {
  RefCtx* c = (RefCtx*)p;
  c->~RefCtx();
  std::free(c);
}

int ref_nnz(void* p) { return ((RefCtx*)p)->mat_A.m_ncrs; }
int ref_nnode(void* p) { return (int)((RefCtx*)p)->aNodeBVH.size(); }
int ref_root(void* p) { return ((RefCtx*)p)->iroot_bvh; }
int ref_nedge(void* p) { return (int)((RefCtx*)p)->aEdge.array.size(); }

void ref_get_pattern(void* p, int* rowptr, int* colind, int* diaind) {
  RefCtx* c = (RefCtx*)p;
  for (int i = 0; i <= c->np; i++) rowptr[i] = c->mat_A.m_colInd[i];
  for (int i = 0; i < c->mat_A.m_ncrs; i++) colind[i] = c->mat_A.m_rowPtr[i];
  for (int i = 0; i < c->np; i++) diaind[i] = c->ilu_A.m_diaInd[i];
}

void ref_get_bvh(void* p, int* nodes) {
  RefCtx* c = (RefCtx*)p;
  for (size_t i = 0; i < c->aNodeBVH.size(); i++) {
    nodes[i * 3 + 0] = c->aNodeBVH[i].iroot;
    nodes[i * 3 + 1] = c->aNodeBVH[i].ichild[0];
    nodes[i * 3 + 2] = c->aNodeBVH[i].ichild[1];
  }
}

void ref_get_edge(void* p, int* index, int* array) {
  RefCtx* c = (RefCtx*)p;
  for (size_t i = 0; i < c->aEdge.index.size(); i++) index[i] = c->aEdge.index[i];
  for (size_t i = 0; i < c->aEdge.array.size(); i++) array[i] = c->aEdge.array[i];
}

// the assembled backward-Euler system A dx = b after boundary conditions
void ref_assemble(void* p, const double* xyz, const double* uvw, double* dia, double* crs,
                  double* b) {
  Quiet q;
  RefCtx* c = (RefCtx*)p;
  std::vector<double> X(xyz, xyz + 3 * c->np), V(uvw, uvw + 3 * c->np), B;
  internal_step(c, X, V, c->tol, c->maxit, &B);
  std::memcpy(dia, c->mat_A.m_valDia, sizeof(double) * 9 * c->np);
  std::memcpy(crs, c->mat_A.m_valCrs, sizeof(double) * 9 * c->mat_A.m_ncrs);
  std::memcpy(b, B.data(), sizeof(double) * 3 * c->np);
}

// SetValueILU + DoILUDecomp on a given matrix; returns the factor
void ref_ilu(void* p, const double* dia, const double* crs, double* fdia, double* fcrs) {
  Quiet q;
  RefCtx* c = (RefCtx*)p;
  load_matrix(c->mat_A, dia, crs);
  c->ilu_A.SetValueILU(c->mat_A);
  c->ilu_A.DoILUDecomp();
  std::memcpy(fdia, c->ilu_A.mat.m_valDia, sizeof(double) * 9 * c->np);
  std::memcpy(fcrs, c->ilu_A.mat.m_valCrs, sizeof(double) * 9 * c->mat_A.m_ncrs);
}

// Solve_PCG on a given matrix + given factor; r is consumed (in/out), x out
void ref_pcg(void* p, const double* dia, const double* crs, const double* fdia, const double* fcrs,
             double* r, double* x, double tol, int maxit, double* conv, int* iters) {
  Quiet q;
  RefCtx* c = (RefCtx*)p;
  load_matrix(c->mat_A, dia, crs);
  load_matrix(c->ilu_A.mat, fdia, fcrs);
  std::vector<double> R(r, r + 3 * c->np), X;
  double cr = tol;
  int it = maxit;
  Solve_PCG(cr, it, c->mat_A, c->ilu_A, R, X);
  std::memcpy(r, R.data(), sizeof(double) * 3 * c->np);
  std::memcpy(x, X.data(), sizeof(double) * 3 * c->np);
  *conv = cr;
  *iters = it;
}

// ILU forward+backward substitution only
void ref_ilu_solve(void* p, const double* fdia, const double* fcrs, double* vec) {
  RefCtx* c = (RefCtx*)p;
  load_matrix(c->ilu_A.mat, fdia, fcrs);
  std::vector<double> v(vec, vec + 3 * c->np);
  c->ilu_A.Solve(v);
  std::memcpy(vec, v.data(), sizeof(double) * 3 * c->np);
}

void ref_matvec(void* p, const double* dia, const double* crs, double alpha, const double* x,
                double beta, double* y) {
  RefCtx* c = (RefCtx*)p;
  load_matrix(c->mat_A, dia, crs);
  std::vector<double> X(x, x + 3 * c->np), Y(y, y + 3 * c->np);
  c->mat_A.MatVec(alpha, X, beta, Y);
  std::memcpy(y, Y.data(), sizeof(double) * 3 * c->np);
}

void ref_step_internal_tol(void* p, double* xyz, double* uvw, double tol, int maxit, double* conv,
                           int* iters) {
  Quiet q;
  RefCtx* c = (RefCtx*)p;
  std::vector<double> X(xyz, xyz + 3 * c->np), V(uvw, uvw + 3 * c->np);
  internal_step(c, X, V, tol, maxit, nullptr);
  std::memcpy(xyz, X.data(), sizeof(double) * 3 * c->np);
  std::memcpy(uvw, V.data(), sizeof(double) * 3 * c->np);
  *conv = c->last_conv;
  *iters = c->last_iter;
}

void ref_step_internal_native(void* p, double* xyz, double* uvw) {
  Quiet q;
  RefCtx* c = (RefCtx*)p;
  PdFn pd = pick_pd(c->obst);
  std::vector<double> X(xyz, xyz + 3 * c->np), V(uvw, uvw + 3 * c->np);
  StepTime_InternalDynamicsILU(X, V, c->mat_A, c->ilu_A, c->xyz0, c->bc, c->tri, c->quad, c->dt,
                               c->lam, c->mu, c->stiff_bend, c->grav, c->mass_point,
                               c->stiff_contact, c->clearance, pd);
  std::memcpy(xyz, X.data(), sizeof(double) * 3 * c->np);
  std::memcpy(uvw, V.data(), sizeof(double) * 3 * c->np);
}

void ref_refit_prx(void* p, const double* xyz, double delta, double* bb) {
  RefCtx* c = (RefCtx*)p;
  std::vector<double> X(xyz, xyz + 3 * c->np);
  BuildBoundingBoxChild_Prx(c->iroot_bvh, delta, X, c->tri, c->aNodeBVH, c->aBB);
  for (size_t i = 0; i < c->aBB.size(); i++) {
    bb[i * 6 + 0] = c->aBB[i].x_min;
    bb[i * 6 + 1] = c->aBB[i].x_max;
    bb[i * 6 + 2] = c->aBB[i].y_min;
    bb[i * 6 + 3] = c->aBB[i].y_max;
    bb[i * 6 + 4] = c->aBB[i].z_min;
    bb[i * 6 + 5] = c->aBB[i].z_max;
  }
}

void ref_refit_ccd(void* p, const double* xyz, const double* uvw, double dt, double* bb) {
  RefCtx* c = (RefCtx*)p;
  std::vector<double> X(xyz, xyz + 3 * c->np), V(uvw, uvw + 3 * c->np);
  BuildBoundingBoxChild_CCD(c->iroot_bvh, dt, X, V, c->tri, c->aNodeBVH, c->aBB);
  for (size_t i = 0; i < c->aBB.size(); i++) {
    bb[i * 6 + 0] = c->aBB[i].x_min;
    bb[i * 6 + 1] = c->aBB[i].x_max;
    bb[i * 6 + 2] = c->aBB[i].y_min;
    bb[i * 6 + 3] = c->aBB[i].y_max;
    bb[i * 6 + 4] = c->aBB[i].z_min;
    bb[i * 6 + 5] = c->aBB[i].z_max;
  }
}

// proximity contact set (refit + query), main path of self_collision_cloth.cpp:409-421
int ref_prox_set(void* p, const double* xyz, double delta, int* out, int cap) {
  RefCtx* c = (RefCtx*)p;
  std::vector<double> X(xyz, xyz + 3 * c->np);
  BuildBoundingBoxChild_Prx(c->iroot_bvh, delta, X, c->tri, c->aNodeBVH, c->aBB);
  std::set<CContactElement> s;
  GetContactElement_Proximity(s, delta, X, c->tri, c->iroot_bvh, c->aNodeBVH, c->aBB);
  int n;
  contact_list(s, out, cap, &n);
  return n;
}

int ref_ccd_set(void* p, const double* xyz, const double* uvw, double dt, double delta, int* out,
                int cap) {
  RefCtx* c = (RefCtx*)p;
  std::vector<double> X(xyz, xyz + 3 * c->np), V(uvw, uvw + 3 * c->np);
  BuildBoundingBoxChild_CCD(c->iroot_bvh, dt, X, V, c->tri, c->aNodeBVH, c->aBB);
  std::set<CContactElement> s;
  GetContactElement_CCD(s, dt, delta, X, V, c->tri, c->iroot_bvh, c->aNodeBVH, c->aBB);
  int n;
  contact_list(s, out, cap, &n);
  return n;
}

// GetIntermidiateVelocityContactResolved on (start-of-step positions, velocities)
int ref_contact_resolve(void* p, const double* xyz1, double* uvw) {
  Quiet q;
  RefCtx* c = (RefCtx*)p;
  std::vector<double> X(xyz1, xyz1 + 3 * c->np), V(uvw, uvw + 3 * c->np);
  bool applied = false;
  GetIntermidiateVelocityContactResolved(V, applied, c->dt, c->clearance, c->mass_point,
                                         c->stiff_contact, X, c->tri, c->aEdge, c->iroot_bvh,
                                         c->aNodeBVH, c->aBB);
  std::memcpy(uvw, V.data(), sizeof(double) * 3 * c->np);
  return applied ? 1 : 0;
}

void ref_update_velocity(void* p, double* uvw, const double* xyz) {
  RefCtx* c = (RefCtx*)p;
  std::vector<double> X(xyz, xyz + 3 * c->np), V(uvw, uvw + 3 * c->np);
  UpdateIntermidiateVelocity(V, X, c->xyz0, c->bc, c->tri, c->quad, c->dt, c->lam, c->mu,
                             c->stiff_bend, c->grav, c->mass_point);
  std::memcpy(uvw, V.data(), sizeof(double) * 3 * c->np);
}

// StepTime (self_contact_sparse/main.cpp:213-261) with the context's PCG settings;
// stats: [impulse_applied, pcg_iters]; conv: achieved ratio
void ref_step(void* p, double* xyz, double* uvw, int* stats, double* conv) {
  Quiet q;
  RefCtx* c = (RefCtx*)p;
  std::vector<double> X(xyz, xyz + 3 * c->np), V(uvw, uvw + 3 * c->np);
  std::vector<double> X1 = X;
  internal_step(c, X, V, c->tol, c->maxit, nullptr);
  bool applied = false;
  GetIntermidiateVelocityContactResolved(V, applied, c->dt, c->clearance, c->mass_point,
                                         c->stiff_contact, X1, c->tri, c->aEdge, c->iroot_bvh,
                                         c->aNodeBVH, c->aBB);
  if (applied) {
    for (int ip = 0; ip < c->np; ip++) {
      X[ip * 3 + 0] = X1[ip * 3 + 0] + V[ip * 3 + 0] * c->dt;
      X[ip * 3 + 1] = X1[ip * 3 + 1] + V[ip * 3 + 1] * c->dt;
      X[ip * 3 + 2] = X1[ip * 3 + 2] + V[ip * 3 + 2] * c->dt;
    }
    UpdateIntermidiateVelocity(V, X, c->xyz0, c->bc, c->tri, c->quad, c->dt, c->lam, c->mu,
                               c->stiff_bend, c->grav, c->mass_point);
  }
  std::memcpy(xyz, X.data(), sizeof(double) * 3 * c->np);
  std::memcpy(uvw, V.data(), sizeof(double) * 3 * c->np);
  stats[0] = applied ? 1 : 0;
  stats[1] = c->last_iter;
  *conv = c->last_conv;
}

// element routines (cloth_internal_physics.h)
void ref_cst(const double* C, const double* cc, double lam, double mu, double* W, double* dW,
             double* ddW) {
  double Ca[3][3], ca[3][3], dWa[3][3], ddWa[3][3][3][3];
  std::memcpy(Ca, C, sizeof(Ca));
  std::memcpy(ca, cc, sizeof(ca));
  WdWddW_CST(*W, dWa, ddWa, Ca, ca, lam, mu);
  std::memcpy(dW, dWa, sizeof(dWa));
  std::memcpy(ddW, ddWa, sizeof(ddWa));
}

void ref_bend(const double* C, const double* cc, double stiff, double* dW, double* ddW) {
  double Ca[4][3], ca[4][3], dWa[4][3], ddWa[4][4][3][3];
  std::memcpy(Ca, C, sizeof(Ca));
  std::memcpy(ca, cc, sizeof(ca));
  double W = 0;  // the reference accumulates W into the caller's variable
  WdWddW_Bend(W, dWa, ddWa, Ca, ca, stiff);
  std::memcpy(dW, dWa, sizeof(dWa));
  std::memcpy(ddW, ddWa, sizeof(ddWa));
}

void ref_contact_elem(const double* cc, double k, double clearance, const double* obst, double* W,
                      double* dW, double* ddW) {
  PdFn pd = pick_pd(obst);
  double dWa[3], ddWa[3][3], ca[3] = {cc[0], cc[1], cc[2]};
  WdWddW_Contact(*W, dWa, ddWa, ca, k, clearance, pd);
  std::memcpy(dW, dWa, sizeof(dWa));
  std::memcpy(ddW, ddWa, sizeof(ddWa));
}

void ref_psd22(const double* s2, double* s3) { MakePositiveDefinite_Sim22(s2, s3); }

double ref_coplanar(const double* s, const double* e) {
  CVector3D S[4], E[4];
  for (int i = 0; i < 4; i++) {
    S[i] = CVector3D(s[i * 3 + 0], s[i * 3 + 1], s[i * 3 + 2]);
    E[i] = CVector3D(e[i * 3 + 0], e[i * 3 + 1], e[i * 3 + 2]);
  }
  return FindCoplanerInterp(S[0], S[1], S[2], S[3], E[0], E[1], E[2], E[3]);
}

double ref_dist_fv(const double* p, double* w) {
  CVector3D P[4];
  for (int i = 0; i < 4; i++) P[i] = CVector3D(p[i * 3 + 0], p[i * 3 + 1], p[i * 3 + 2]);
  return DistanceFaceVertex(P[0], P[1], P[2], P[3], w[0], w[1]);
}

double ref_dist_ee(const double* p, double* w) {
  CVector3D P[4];
  for (int i = 0; i < 4; i++) P[i] = CVector3D(p[i * 3 + 0], p[i * 3 + 1], p[i * 3 + 2]);
  return DistanceEdgeEdge(P[0], P[1], P[2], P[3], w[0], w[1]);
}

}  // extern "C"
