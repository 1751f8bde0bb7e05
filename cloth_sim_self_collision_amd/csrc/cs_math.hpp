// Device-side arithmetic of the cloth step: 3-vectors, boxes, element
// physics and the contact narrow phase.
//
// Every expression keeps the evaluation order of the reference
// (nobuyuki83/cloth_sim_self_collision); with -ffp-contract=off and IEEE
// division/sqrt each result is bit-identical to the reference's x86-64 build.
// Functions cite the reference lines whose arithmetic they follow.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#define CS_HD __host__ __device__ __forceinline__

namespace cs {

struct Vec3 {
  double x, y, z;
};

CS_HD Vec3 mk(double x, double y, double z) {
  Vec3 r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}
CS_HD Vec3 operator+(Vec3 a, Vec3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
CS_HD Vec3 operator-(Vec3 a, Vec3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
// scalar * vector multiplies each component by the scalar (component first),
// as CVector3D::operator*= does (vector3d.h:59-62, 128-132)
CS_HD Vec3 operator*(double s, Vec3 a) { return mk(a.x * s, a.y * s, a.z * s); }
CS_HD double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CS_HD Vec3 cross(Vec3 a, Vec3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
CS_HD double norm(Vec3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
CS_HD Vec3 unit(Vec3 a) {  // vector3d.h:93-99
  const double s = 1.0 / norm(a);
  return mk(a.x * s, a.y * s, a.z * s);
}
CS_HD Vec3 at(const double* p, int i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
// position advanced by dt*velocity, per component (bvh_aabb.cpp:87-89)
CS_HD Vec3 advanced(const double* p, const double* u, double dt, int i) {
  return mk(p[3 * i] + dt * u[3 * i], p[3 * i + 1] + dt * u[3 * i + 1],
            p[3 * i + 2] + dt * u[3 * i + 2]);
}
CS_HD double triple(Vec3 a, Vec3 b, Vec3 c) {  // vector3d.h:216-218
  return a.x * (b.y * c.z - b.z * c.y) + a.y * (b.z * c.x - b.x * c.z) +
         a.z * (b.x * c.y - b.y * c.x);
}

// ---------------------------------------------------------------------------
// axis-aligned box: lo/hi per axis (aabb.h:16-148)
struct Box {
  double lx, hx, ly, hy, lz, hz;
};

CS_HD Box box_point(Vec3 p, double e) {
  Box b;
  b.lx = p.x - e;
  b.hx = p.x + e;
  b.ly = p.y - e;
  b.hy = p.y + e;
  b.lz = p.z - e;
  b.hz = p.z + e;
  return b;
}
CS_HD void box_grow(Box& b, Vec3 p, double e) {  // aabb.h:104-111
  b.lx = (b.lx < p.x - e) ? b.lx : p.x - e;
  b.ly = (b.ly < p.y - e) ? b.ly : p.y - e;
  b.lz = (b.lz < p.z - e) ? b.lz : p.z - e;
  b.hx = (b.hx > p.x + e) ? b.hx : p.x + e;
  b.hy = (b.hy > p.y + e) ? b.hy : p.y + e;
  b.hz = (b.hz > p.z + e) ? b.hz : p.z + e;
}
CS_HD Box box_merge(const Box& a, const Box& b) {  // aabb.h:65-70
  Box r;
  r.hx = (a.hx > b.hx) ? a.hx : b.hx;
  r.lx = (a.lx < b.lx) ? a.lx : b.lx;
  r.hy = (a.hy > b.hy) ? a.hy : b.hy;
  r.ly = (a.ly < b.ly) ? a.ly : b.ly;
  r.hz = (a.hz > b.hz) ? a.hz : b.hz;
  r.lz = (a.lz < b.lz) ? a.lz : b.lz;
  return r;
}
CS_HD bool overlap(const Box& a, const Box& b) {  // aabb.h:73-84
  return !(a.hx < b.lx || a.lx > b.hx || a.hy < b.ly || a.ly > b.hy || a.hz < b.lz ||
           a.lz > b.hz);
}
CS_HD bool inside(const Box& b, Vec3 p) {  // aabb.h:137-144
  return p.x >= b.lx && p.x <= b.hx && p.y >= b.ly && p.y <= b.hy && p.z >= b.lz && p.z <= b.hz;
}

// ---------------------------------------------------------------------------
// element physics (cloth_internal_physics.h)

// positive semi-definite part of a symmetric 2x2 tensor (xx, yy, xy); :15-67
CS_HD void psd_part(const double s[3], double o[3]) {
  const double b = (s[0] + s[1]) * 0.5;
  const double d = (s[0] - s[1]) * (s[0] - s[1]) * 0.25 + s[2] * s[2];
  const double e = sqrt(d);
  if (b - e > 1.0e-20) {
    o[0] = s[0];
    o[1] = s[1];
    o[2] = s[2];
    return;
  }
  o[0] = o[1] = o[2] = 0;
  if (b + e < 0) return;
  const double l = b + e;
  const double a0 = s[0] - l, a1 = s[2];  // two candidate eigenvectors
  const double b0 = s[2], b1 = s[1] - l;
  const double na = a0 * a0 + a1 * a1;
  const double nb = b0 * b0 + b1 * b1;
  const bool use_a = na > nb;
  const double n2 = use_a ? na : nb;
  if (n2 < 1.0e-20) return;
  const double inv = 1.0 / sqrt(n2);
  const double u0 = (use_a ? a0 : b0) * inv;
  const double u1 = (use_a ? a1 : b1) * inv;
  o[0] = l * u0 * u0;
  o[1] = l * u1 * u1;
  o[2] = l * u0 * u1;
}

// linear-triangle shape gradients in (r, s): N0 = 1-r-s, N1 = r, N2 = s
__device__ __constant__ static const double kShapeGrad[3][2] = {
    {-1.0, -1.0}, {+1.0, +0.0}, {+0.0, +1.0}};
// summation order of the 16 material-stiffness terms (cloth_internal_physics.h:146-161):
// term k = d[a][i] * N[ino][b] * Cmat[r][s] * d[c][j] * N[jno][e], (a,b,r,s,c,e) = row k
__device__ __constant__ static const signed char kMatTerm[16][6] = {
    {0, 0, 0, 0, 0, 0}, {0, 0, 0, 1, 1, 1}, {0, 0, 0, 2, 0, 1}, {0, 0, 0, 2, 1, 0},
    {1, 1, 1, 0, 0, 0}, {1, 1, 1, 1, 1, 1}, {1, 1, 1, 2, 0, 1}, {1, 1, 1, 2, 1, 0},
    {0, 1, 2, 0, 0, 0}, {0, 1, 2, 1, 1, 1}, {0, 1, 2, 2, 0, 1}, {0, 1, 2, 2, 1, 0},
    {1, 0, 2, 0, 0, 0}, {1, 0, 2, 1, 1, 1}, {1, 0, 2, 2, 0, 1}, {1, 0, 2, 2, 1, 0}};

// Quantities of one St.Venant-Kirchhoff constant-strain triangle shared by
// its gradient and Hessian blocks (cloth_internal_physics.h:70-139).
struct CstState {
  double area;
  double d[2][3];     // current edge vectors c1-c0, c2-c0
  double cmat[3][3];  // constitutive tensor in the contravariant metric
  double stress[3];   // 2nd Piola-Kirchhoff (xx, yy, xy)
  double psd[3];      // PSD part of the stress (geometric stiffness)
  double strain[3];
};

// X: rest corners (9), x: current corners (9)
__device__ __forceinline__ void cst_state(const double* X, const double* x, double lam, double mu,
                                          CstState& s) {
  double e0[3], e1[3], nrm[3];
  for (int k = 0; k < 3; k++) {
    e0[k] = X[3 + k] - X[k];
    e1[k] = X[6 + k] - X[k];
  }
  {  // utility.h:58-73 (unit normal and area of the rest triangle)
    const double* a = X;
    const double* b = X + 3;
    const double* c = X + 6;
    nrm[0] = (b[1] - a[1]) * (c[2] - a[2]) - (c[1] - a[1]) * (b[2] - a[2]);
    nrm[1] = (b[2] - a[2]) * (c[0] - a[0]) - (c[2] - a[2]) * (b[0] - a[0]);
    nrm[2] = (b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1]);
    s.area = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]) * 0.5;
    const double il = 0.5 / s.area;
    nrm[0] *= il;
    nrm[1] *= il;
    nrm[2] *= il;
  }
  // contravariant basis g0 = (e1 x n)/((e1 x n).e0), g1 = (n x e0)/((n x e0).e1)
  double g0[3], g1[3];
  g0[0] = e1[1] * nrm[2] - nrm[1] * e1[2];
  g0[1] = e1[2] * nrm[0] - nrm[2] * e1[0];
  g0[2] = e1[0] * nrm[1] - nrm[0] * e1[1];
  {
    const double f = 1.0 / (g0[0] * e0[0] + g0[1] * e0[1] + g0[2] * e0[2]);
    g0[0] *= f;
    g0[1] *= f;
    g0[2] *= f;
  }
  g1[0] = nrm[1] * e0[2] - e0[1] * nrm[2];
  g1[1] = nrm[2] * e0[0] - e0[2] * nrm[0];
  g1[2] = nrm[0] * e0[1] - e0[0] * nrm[1];
  {
    const double f = 1.0 / (g1[0] * e1[0] + g1[1] * e1[1] + g1[2] * e1[2]);
    g1[0] *= f;
    g1[1] *= f;
    g1[2] *= f;
  }
  for (int k = 0; k < 3; k++) {
    s.d[0][k] = x[3 + k] - x[k];
    s.d[1][k] = x[6 + k] - x[k];
  }
  const double* d0 = s.d[0];
  const double* d1 = s.d[1];
  s.strain[0] = 0.5 * ((d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]) -
                       (e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2]));
  s.strain[1] = 0.5 * ((d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2]) -
                       (e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]));
  s.strain[2] = 1.0 * ((d0[0] * d1[0] + d0[1] * d1[1] + d0[2] * d1[2]) -
                       (e0[0] * e1[0] + e0[1] * e1[1] + e0[2] * e1[2]));
  const double m00 = g0[0] * g0[0] + g0[1] * g0[1] + g0[2] * g0[2];
  const double m11 = g1[0] * g1[0] + g1[1] * g1[1] + g1[2] * g1[2];
  const double m01 = g1[0] * g0[0] + g1[1] * g0[1] + g1[2] * g0[2];
  double(*C)[3] = s.cmat;
  C[0][0] = lam * m00 * m00 + 2 * mu * (m00 * m00);
  C[0][1] = lam * m00 * m11 + 2 * mu * (m01 * m01);
  C[0][2] = lam * m00 * m01 + 2 * mu * (m00 * m01);
  C[1][0] = lam * m11 * m00 + 2 * mu * (m01 * m01);
  C[1][1] = lam * m11 * m11 + 2 * mu * (m11 * m11);
  C[1][2] = lam * m11 * m01 + 2 * mu * (m01 * m11);
  C[2][0] = lam * m01 * m00 + 2 * mu * (m00 * m01);
  C[2][1] = lam * m01 * m11 + 2 * mu * (m01 * m11);
  C[2][2] = lam * m01 * m01 + 1 * mu * (m00 * m11 + m01 * m01);
  for (int r = 0; r < 3; r++)
    s.stress[r] = C[r][0] * s.strain[0] + C[r][1] * s.strain[1] + C[r][2] * s.strain[2];
  psd_part(s.stress, s.psd);
}

// gradient of corner a, component i (cloth_internal_physics.h:128-136)
__device__ __forceinline__ double cst_grad(const CstState& s, int a, int i) {
  const double n0 = kShapeGrad[a][0], n1 = kShapeGrad[a][1];
  return s.area * (+s.stress[0] * s.d[0][i] * n0 + s.stress[2] * s.d[0][i] * n1 +
                   s.stress[2] * s.d[1][i] * n0 + s.stress[1] * s.d[1][i] * n1);
}

// Hessian block (a, b): 9 values row-major (cloth_internal_physics.h:142-174)
__device__ __forceinline__ void cst_hess_block(const CstState& s, int a, int b, double out[9]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double acc = 0;
#pragma unroll
      for (int k = 0; k < 16; k++) {
        const signed char* t = kMatTerm[k];
        acc += s.d[t[0]][i] * kShapeGrad[a][t[1]] * s.cmat[t[2]][t[3]] * s.d[t[4]][j] *
               kShapeGrad[b][t[5]];
      }
      out[i * 3 + j] = acc * s.area;
    }
  const double na0 = kShapeGrad[a][0], na1 = kShapeGrad[a][1];
  const double nb0 = kShapeGrad[b][0], nb1 = kShapeGrad[b][1];
  const double geo = s.area * (+s.psd[0] * na0 * nb0 + s.psd[2] * na0 * nb1 +
                               s.psd[2] * na1 * nb0 + s.psd[1] * na1 * nb1);
  out[0] += geo;
  out[4] += geo;
  out[8] += geo;
}

CS_HD double tri_area3(const double* a, const double* b, const double* c) {  // utility.h:43-53
  const double n0 = (b[1] - a[1]) * (c[2] - a[2]) - (c[1] - a[1]) * (b[2] - a[2]);
  const double n1 = (b[2] - a[2]) * (c[0] - a[0]) - (c[2] - a[2]) * (b[0] - a[0]);
  const double n2 = (b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1]);
  return sqrt(n0 * n0 + n1 * n1 + n2 * n2) * 0.5;
}

// Rest-shape weights of a bending stencil [opp0, opp1, e0, e1]:
// w[a] = K[a], scale = stiff/((A0+A1) L^2); the Hessian is w_a w_b scale * I3
// (cloth_internal_physics.h:188-225).  X: 12 rest coordinates.
CS_HD void bend_weights(const double* X, double stiff, double w[4], double* scale) {
  const double* P0 = X;
  const double* P1 = X + 3;
  const double* P2 = X + 6;
  const double* P3 = X + 9;
  const double A0 = tri_area3(P0, P2, P3);
  const double A1 = tri_area3(P1, P3, P2);
  const double L = sqrt((P3[0] - P2[0]) * (P3[0] - P2[0]) + (P3[1] - P2[1]) * (P3[1] - P2[1]) +
                        (P3[2] - P2[2]) * (P3[2] - P2[2]));
  const double h0 = A0 * 2.0 / L;
  const double h1 = A1 * 2.0 / L;
  double e23[3], e02[3], e03[3], e12[3], e13[3];
  for (int k = 0; k < 3; k++) {
    e23[k] = P3[k] - P2[k];
    e02[k] = P2[k] - P0[k];
    e03[k] = P3[k] - P0[k];
    e12[k] = P2[k] - P1[k];
    e13[k] = P3[k] - P1[k];
  }
  const double c023 = -(e02[0] * e23[0] + e02[1] * e23[1] + e02[2] * e23[2]) / h0;
  const double c032 = +(e03[0] * e23[0] + e03[1] * e23[1] + e03[2] * e23[2]) / h0;
  const double c123 = -(e12[0] * e23[0] + e12[1] * e23[1] + e12[2] * e23[2]) / h1;
  const double c132 = +(e13[0] * e23[0] + e13[1] * e23[1] + e13[2] * e23[2]) / h1;
  *scale = stiff / ((A0 + A1) * L * L);
  w[0] = -c023 - c032;
  w[1] = -c123 - c132;
  w[2] = c032 + c132;
  w[3] = c023 + c123;
}

// bending gradient of corner a, component i from the 16 Hessian scalars
// kab[a*4+b] (the Hessian block (a,b) is kab*I3): the reference sums all 12
// (b, j) terms including the zero off-diagonal ones (cloth_internal_physics.h:227-234)
CS_HD double bend_grad(const double* kab, const double* x, int a, int i) {
  double acc = 0;
  for (int b = 0; b < 4; b++)
    for (int j = 0; j < 3; j++) acc += (i == j ? kab[a * 4 + b] : 0.0) * x[b * 3 + j];
  return acc;
}

// obstacle: kind 0 none, 1 plane (nx,ny,nz,d), 2 sphere (cx,cy,cz,r),
// 3 sign-flipped sphere; pd > 0 inside, n = push-out direction
CS_HD void obstacle_depth(int kind, const double* q, const double* p, double* pd, double* n) {
  if (kind == 1) {
    n[0] = q[0];
    n[1] = q[1];
    n[2] = q[2];
    *pd = q[3] - (n[0] * p[0] + n[1] * p[1] + n[2] * p[2]);
  } else if (kind == 2 || kind == 3) {
    n[0] = p[0] - q[0];
    n[1] = p[1] - q[1];
    n[2] = p[2] - q[2];
    const double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] /= len;
    n[1] /= len;
    n[2] /= len;
    *pd = q[3] - len;
    if (kind == 3) {
      n[0] *= -1;
      n[1] *= -1;
      n[2] *= -1;
      *pd *= -1;
    }
  } else if (kind == 4) {
    // axis-aligned box: center q[0..2], half-extents q[3..5]
    double s[3], d[3];
    for (int i = 0; i < 3; i++) {
      s[i] = p[i] - q[i];
      d[i] = fabs(s[i]) - q[3 + i];
    }
    if (d[0] <= 0.0 && d[1] <= 0.0 && d[2] <= 0.0) {
      int m = 0;
      if (d[1] > d[m]) m = 1;
      if (d[2] > d[m]) m = 2;
      n[0] = n[1] = n[2] = 0.0;
      n[m] = s[m] < 0.0 ? -1.0 : 1.0;
      *pd = -d[m];
    } else {
      double o[3];
      for (int i = 0; i < 3; i++) o[i] = d[i] > 0.0 ? d[i] : 0.0;
      const double len = sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
      for (int i = 0; i < 3; i++) n[i] = (s[i] < 0.0 ? -o[i] : o[i]) / len;
      *pd = -len;
    }
  } else if (kind == 5 || kind == 6) {
    // bar: the capsule around the segment q[0..2]-q[3..5] with radius q[6],
    // pd = r - |p - c| with c the closest point of the segment, n = (p - c)/|p - c|
    double e[3], w[3];
    for (int i = 0; i < 3; i++) {
      e[i] = q[3 + i] - q[i];
      w[i] = p[i] - q[i];
    }
    const double ee = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
    double t = (w[0] * e[0] + w[1] * e[1] + w[2] * e[2]) / ee;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    for (int i = 0; i < 3; i++) n[i] = w[i] - t * e[i];
    const double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] /= len;
    n[1] /= len;
    n[2] /= len;
    *pd = q[6] - len;
    if (kind == 6) {
      // and the ground z = q[7] (normal +z): the deeper of the two, the bar on ties
      const double pg = q[7] - p[2];
      if (pg > *pd) {
        n[0] = 0.0;
        n[1] = 0.0;
        n[2] = 1.0;
        *pd = pg;
      }
    }
  } else {
    *pd = -100;
    n[0] = n[1] = n[2] = 0;
  }
}

// penalty contact of one vertex (cloth_internal_physics.h:243-280) given
// (pd, n) of the obstacle; returns whether active
CS_HD bool contact_penalty(double pd, const double* n, double k, double clearance, double* grad,
                           double* hess) {
  pd += clearance;
  if (pd < 0) {
    for (int i = 0; i < 3; i++) grad[i] = 0;
    for (int i = 0; i < 9; i++) hess[i] = 0;
    return false;
  }
  for (int i = 0; i < 3; i++) grad[i] = -k * pd * n[i];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) hess[i * 3 + j] = k * n[i] * n[j];
  return true;
}

// ---------------------------------------------------------------------------
// narrow phase (bvh_aabb.cpp)

CS_HD double tet_height(Vec3 a, Vec3 b, Vec3 c, Vec3 d) {  // vector3d.h:254-267
  double nx = (b.y - a.y) * (c.z - a.z) - (b.z - a.z) * (c.y - a.y);
  double ny = (b.z - a.z) * (c.x - a.x) - (b.x - a.x) * (c.z - a.z);
  double nz = (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x);
  const double s = 1.0 / sqrt(nx * nx + ny * ny + nz * nz);
  nx *= s;
  ny *= s;
  nz *= s;
  return (d.x - a.x) * nx + (d.y - a.y) * ny + (d.z - a.z) * nz;
}

CS_HD Vec3 closest_on_segment(Vec3 p, Vec3 s, Vec3 e) {  // vector3d.h:282-295
  const Vec3 d = e - s;
  const double a = dot(d, d);
  const double b = dot(d, s - p);
  double t = -b / a;
  if (t < 0) t = 0;
  if (t > 1) t = 1;
  return s + t * d;
}

// vertex p3 against the plane of (p0,p1,p2), unclamped barycentrics (:321-340)
CS_HD double face_vertex(Vec3 p0, Vec3 p1, Vec3 p2, Vec3 p3, double& w0, double& w1) {
  const Vec3 a = p0 - p2, b = p1 - p2;
  const double t0 = dot(a, a), t1 = dot(b, b), t2 = dot(a, b);
  const double t3 = dot(a, p3 - p2), t4 = dot(b, p3 - p2);
  const double det = t0 * t1 - t2 * t2;
  const double inv = 1.0 / det;
  w0 = (+t1 * t3 - t2 * t4) * inv;
  w1 = (-t2 * t3 + t0 * t4) * inv;
  const double w2 = 1 - w0 - w1;
  const Vec3 pw = w0 * p0 + w1 * p1 + w2 * p2;
  return norm(pw - p3);
}

// closest points of the lines (p0,p1), (q0,q1) with a parallel branch (:343-384)
CS_HD double edge_edge(Vec3 p0, Vec3 p1, Vec3 q0, Vec3 q1, double& rp, double& rq) {
  const Vec3 vp = p1 - p0, vq = q1 - q0;
  if (norm(cross(vp, vq)) < 1.0e-10) {
    const Vec3 pq0 = p0 - q0;
    const Vec3 u = unit(vp);
    const Vec3 perp = pq0 - dot(pq0, u) * u;
    const double dist = norm(perp);
    const double lp0 = dot(p0, u), lp1 = dot(p1, u);
    const double lq0 = dot(q0, u), lq1 = dot(q1, u);
    const double pmin = (lp0 < lp1) ? lp0 : lp1, pmax = (lp0 > lp1) ? lp0 : lp1;
    const double qmin = (lq0 < lq1) ? lq0 : lq1, qmax = (lq0 > lq1) ? lq0 : lq1;
    double mid;
    if (pmax < qmin)
      mid = (pmax + qmin) * 0.5;
    else if (qmax < pmin)
      mid = (qmax + pmin) * 0.5;
    else if (pmax < qmax)
      mid = (pmax + qmin) * 0.5;
    else
      mid = (qmax + pmin) * 0.5;
    rp = (mid - lp0) / (lp1 - lp0);
    rq = (mid - lq0) / (lq1 - lq0);
    return dist;
  }
  const double t0 = dot(vp, vp), t1 = dot(vq, vq), t2 = dot(vp, vq);
  const double t3 = dot(vp, q0 - p0), t4 = dot(vq, q0 - p0);
  const double det = t0 * t1 - t2 * t2;
  const double inv = 1.0 / det;
  rp = (+t1 * t3 - t2 * t4) * inv;
  rq = (+t2 * t3 - t0 * t4) * inv;
  const Vec3 pc = p0 + rp * vp;
  const Vec3 qc = q0 + rq * vq;
  return norm(pc - qc);
}

// vertex-face proximity (:392-412); the reference's margin is zero
CS_HD bool fv_near(int i0, int i1, int i2, int i3, Vec3 p0, Vec3 p1, Vec3 p2, Vec3 p3,
                   const Box& b, double delta) {
  if (i3 == i0 || i3 == i1 || i3 == i2) return false;
  if (!inside(b, p3)) return false;
  if (fabs(tet_height(p0, p1, p2, p3)) > delta) return false;
  double w0, w1;
  const double dist = face_vertex(p0, p1, p2, p3, w0, w1);
  const double w2 = 1 - w0 - w1;
  if (dist > delta) return false;
  if (w0 < -0.0 || w0 > 1 + 0.0) return false;
  if (w1 < -0.0 || w1 > 1 + 0.0) return false;
  if (w2 < -0.0 || w2 > 1 + 0.0) return false;
  return true;
}

// edge-edge proximity (:415-438)
CS_HD bool ee_near(int i0, int i1, int j0, int j1, Vec3 p0, Vec3 p1, Vec3 q0, Vec3 q1, double d) {
  if (i0 == j0 || i0 == j1 || i1 == j0 || i1 == j1) return false;
  if (q0.x + d < p0.x && q0.x + d < p1.x && q1.x + d < p0.x && q1.x + d < p1.x) return false;
  if (q0.x - d > p0.x && q0.x - d > p1.x && q1.x - d > p0.x && q1.x - d > p1.x) return false;
  if (q0.y + d < p0.y && q0.y + d < p1.y && q1.y + d < p0.y && q1.y + d < p1.y) return false;
  if (q0.y - d > p0.y && q0.y - d > p1.y && q1.y - d > p0.y && q1.y - d > p1.y) return false;
  if (q0.z + d < p0.z && q0.z + d < p1.z && q1.z + d < p0.z && q1.z + d < p1.z) return false;
  if (q0.z - d > p0.z && q0.z - d > p1.z && q1.z - d > p0.z && q1.z - d > p1.z) return false;
  double rp, rq;
  const double dist = edge_edge(p0, p1, q0, q1, rp, rq);
  if (dist > d) return false;
  if (rp < 0 || rp > 1 || rq < 0 || rq > 1) return false;
  const Vec3 pm = (1 - rp) * p0 + rp * p1;
  const Vec3 qm = (1 - rq) * q0 + rq * q1;
  if (norm(pm - qm) > d) return false;
  return true;
}

CS_HD double cubic_at(double r, double k0, double k1, double k2, double k3) {  // :563-568
  return k0 + k1 * r + k2 * r * r + k3 * r * r * r;
}

// 14 halvings; the sign reference stays the value at the first left end (:571-595)
CS_HD double bisect_root(double r0, double r1, double f0, double k0, double k1, double k2,
                         double k3) {
  for (int n = 0; n < 14; n++) {
    const double rm = 0.5 * (r0 + r1);
    const double fm = cubic_at(rm, k0, k1, k2, k3);
    if (f0 * fm < 0)
      r1 = rm;
    else
      r0 = rm;
  }
  return 0.5 * (r0 + r1);
}

// first coplanarity time of four linearly moving points in [0,1], -1 if none (:598-662)
CS_HD double coplanar_time(Vec3 s0, Vec3 s1, Vec3 s2, Vec3 s3, Vec3 e0, Vec3 e1, Vec3 e2,
                           Vec3 e3) {
  const Vec3 x1 = s1 - s0, x2 = s2 - s0, x3 = s3 - s0;
  const Vec3 v1 = e1 - e0 - x1, v2 = e2 - e0 - x2, v3 = e3 - e0 - x3;
  const double k0 = triple(x3, x1, x2);
  const double k1 = triple(v3, x1, x2) + triple(x3, v1, x2) + triple(x3, x1, v2);
  const double k2 = triple(v3, v1, x2) + triple(v3, x1, v2) + triple(x3, v1, v2);
  const double k3 = triple(v3, v1, v2);
  const double r0 = -0.0, r1 = +1.0;
  const double f0 = cubic_at(r0, k0, k1, k2, k3);
  const double f1 = cubic_at(r1, k0, k1, k2, k3);
  const double disc = k2 * k2 - 3 * k1 * k3;
  if (fabs(k3) < 1.0e-10 && fabs(k2) > 1.0e-10) {
    const double r2 = -k1 / (2 * k2);
    const double f2 = cubic_at(r2, k0, k1, k2, k3);
    if (r2 > 0 && r2 < 1) {
      if (f0 * f2 < 0) return bisect_root(r0, r2, f0, k0, k1, k2, k3);
      if (f2 * f1 < 0) return bisect_root(r2, r1, f2, k0, k1, k2, k3);
    }
  }
  if (disc > 0 && fabs(k3) > 1.0e-10) {
    const double r3 = (-k2 - sqrt(disc)) / (3 * k3);
    const double f3 = cubic_at(r3, k0, k1, k2, k3);
    if (r3 > 0 && r3 < 1) {
      if (f0 * f3 < 0) return bisect_root(r0, r3, f0, k0, k1, k2, k3);
      if (f3 * f1 < 0) return bisect_root(r3, r1, f3, k0, k1, k2, k3);
    }
    const double r4 = (-k2 + sqrt(disc)) / (3 * k3);
    const double f4 = cubic_at(r4, k0, k1, k2, k3);
    if (r3 > 0 && r3 < 1 && r4 > 0 && r4 < 1) {
      if (f3 * f4 < 0) return bisect_root(r3, r4, f3, k0, k1, k2, k3);
    }
    if (r4 > 0 && r4 < 1) {
      if (f0 * f4 < 0) return bisect_root(r0, r4, f0, k0, k1, k2, k3);
      if (f4 * f1 < 0) return bisect_root(r4, r1, f4, k0, k1, k2, k3);
    }
  }
  if (f0 * f1 > 0) return -1;
  return bisect_root(r0, r1, f0, k0, k1, k2, k3);
}

// continuous vertex-face test over one step (:666-719)
CS_HD bool fv_sweep(int i0, int i1, int i2, int i3, Vec3 p0, Vec3 p1, Vec3 p2, Vec3 p3, Vec3 q0,
                    Vec3 q1, Vec3 q2, Vec3 q3, const Box& b) {
  const double eps = 1.0e-10;
  if (i3 == i0 || i3 == i1 || i3 == i2) return false;
  Box bp = box_point(p3, eps);
  box_grow(bp, q3, eps);
  if (!overlap(b, bp)) return false;
  {
    const Vec3 n = cross(p1 - p0, p2 - p0);
    const double t0 = dot(p0 - p3, n);
    const double t1 = dot(q0 - q3, n);
    const double t2 = dot(q1 - q3, n);
    const double t3 = dot(q2 - q3, n);
    if (t0 * t1 > 0 && t0 * t2 > 0 && t0 * t3 > 0) return false;
  }
  double r0, r1;
  const double dist = face_vertex(p0, p1, p2, p3, r0, r1);
  {
    const double m0 = norm(p0 - q0), m1 = norm(p1 - q1), m2 = norm(p2 - q2), m3 = norm(p3 - q3);
    double mf = (m0 > m1) ? m0 : m1;
    mf = (m2 > mf) ? m2 : mf;
    const double reach = mf + m3;
    const double r2 = 1 - r0 - r1;
    if (dist > reach) return false;
    if (r0 < 0 || r0 > 1 || r1 < 0 || r1 > 1 || r2 < 0 || r2 > 1) {
      const double d01 = norm(closest_on_segment(p3, p0, p1) - p3);
      const double d12 = norm(closest_on_segment(p3, p1, p2) - p3);
      const double d20 = norm(closest_on_segment(p3, p2, p0) - p3);
      if (d01 > reach && d12 > reach && d20 > reach) return false;
    }
  }
  const double t = coplanar_time(p0, p1, p2, p3, q0, q1, q2, q3);
  if (t < 0 || t > 1) return false;
  const Vec3 m0 = (1 - t) * p0 + t * q0;
  const Vec3 m1 = (1 - t) * p1 + t * q1;
  const Vec3 m2 = (1 - t) * p2 + t * q2;
  const Vec3 m3 = (1 - t) * p3 + t * q3;
  double w0, w1;
  face_vertex(m0, m1, m2, m3, w0, w1);
  const double w2 = 1 - w0 - w1;
  if (w0 < 0 || w0 > 1) return false;
  if (w1 < 0 || w1 > 1) return false;
  if (w2 < 0 || w2 > 1) return false;
  return true;
}

// continuous edge-edge test over one step (:723-793); 1e-2 is the reference's
// hard-coded distance threshold
CS_HD bool ee_sweep(int i0, int i1, int j0, int j1, Vec3 p0s, Vec3 p1s, Vec3 q0s, Vec3 q1s,
                    Vec3 p0e, Vec3 p1e, Vec3 q0e, Vec3 q1e) {
  const double eps = 1.0e-10;
  if (i0 == j0 || i0 == j1 || i1 == j0 || i1 == j1) return false;
  Box bq = box_point(q0s, eps);
  box_grow(bq, q1s, eps);
  box_grow(bq, q0e, eps);
  box_grow(bq, q1e, eps);
  Box bp = box_point(p0s, eps);
  box_grow(bp, p1s, eps);
  box_grow(bp, p0e, eps);
  box_grow(bp, p1e, eps);
  if (!overlap(bp, bq)) return false;
  const double t = coplanar_time(p0s, p1s, q0s, q1s, p0e, p1e, q0e, q1e);
  if (t < 0 || t > 1) return false;
  const Vec3 a0 = (1 - t) * p0s + t * p0e;
  const Vec3 a1 = (1 - t) * p1s + t * p1e;
  const Vec3 b0 = (1 - t) * q0s + t * q0e;
  const Vec3 b1 = (1 - t) * q1s + t * q1e;
  double w0, w1;
  const double dist = edge_edge(a0, a1, b0, b1, w0, w1);
  if (w0 < 0 || w0 > 1) return false;
  if (w1 < 0 || w1 > 1) return false;
  if (dist > 1.0e-2) return false;
  return true;
}

}  // namespace cs
