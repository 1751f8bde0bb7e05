/* CPU oracle: single-threaded FP64 C restatement of the reference's implicit
 * cloth step with self-contact.  TEST INFRASTRUCTURE ONLY (see cloth_oracle.h):
 * it is the checker for the HIP library, never part of it.
 *
 * Every double operation is written in the order the reference evaluates it
 * (C/C++ left-to-right for equal precedence), and the file is compiled with
 * -ffp-contract=off on x86-64 (SSE2, no FMA), so results are bit-identical to
 * the reference's own g++ build.  Each function cites the reference lines it
 * restates ("ref:" = file:line in nobuyuki83/cloth_sim_self_collision).
 *
 * Parity pins (tests/): compiled reference (oracle/_ref) call for call,
 * golden fixtures in tests/golden, closed forms / finite differences for the
 * element routines, dense solves for ILU/PCG on tiny systems, brute force for
 * BVH refit and contact sets.
 */
#include "cloth_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* 3-vectors, evaluated like CVector3D (ref: vector3d.h:33-295) and the
 * array helpers (ref: utility.h:15-73) */

typedef struct {
  double x, y, z;
} v3;

static inline v3 V(double x, double y, double z) {
  v3 r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
/* d*v : CVector3D temp = v; temp *= d  (component * scalar) */
static inline v3 vscl(double d, v3 a) { return V(a.x * d, a.y * d, a.z * d); }
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
  return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double vlen(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline v3 vnormalized(v3 a) { /* SetNormalizedVector, vector3d.h:93-99 */
  double inv = 1.0 / vlen(a);
  return V(a.x * inv, a.y * inv, a.z * inv);
}
static inline double vcomp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline v3 vload(const double* p, int i) { return V(p[i * 3 + 0], p[i * 3 + 1], p[i * 3 + 2]); }
/* x + dt*v per component (ref: bvh_aabb.cpp:87-89, 848-853) */
static inline v3 vload_adv(const double* p, const double* u, double dt, int i) {
  return V(p[i * 3 + 0] + dt * u[i * 3 + 0], p[i * 3 + 1] + dt * u[i * 3 + 1],
           p[i * 3 + 2] + dt * u[i * 3 + 2]);
}
/* ScalarTripleProduct (ref: vector3d.h:216-218) */
static inline double stp(v3 a, v3 b, v3 c) {
  return a.x * (b.y * c.z - b.z * c.y) + a.y * (b.z * c.x - b.x * c.z) +
         a.z * (b.x * c.y - b.y * c.x);
}
/* tetra height (ref: vector3d.h:254-267) */
static double height4(v3 a, v3 b, v3 c, v3 d) {
  double nx = (b.y - a.y) * (c.z - a.z) - (b.z - a.z) * (c.y - a.y);
  double ny = (b.z - a.z) * (c.x - a.x) - (b.x - a.x) * (c.z - a.z);
  double nz = (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x);
  const double inv = 1.0 / sqrt(nx * nx + ny * ny + nz * nz);
  nx *= inv;
  ny *= inv;
  nz *= inv;
  return (d.x - a.x) * nx + (d.y - a.y) * ny + (d.z - a.z) * nz;
}
/* closest point on segment s-e to p (ref: vector3d.h:282-295) */
static v3 seg_closest(v3 p, v3 s, v3 e) {
  v3 d = vsub(e, s);
  double a = vdot(d, d);
  double b = vdot(d, vsub(s, p));
  double t = -b / a;
  if (t < 0) t = 0;
  if (t > 1) t = 1;
  return vadd(s, vscl(t, d));
}

static inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void cross3(double* r, const double* a, const double* b) {
  r[0] = a[1] * b[2] - b[1] * a[2];
  r[1] = a[2] * b[0] - b[2] * a[0];
  r[2] = a[0] * b[1] - b[0] * a[1];
}
static inline double tri_area(const double* a, const double* b, const double* c) {
  double n0 = (b[1] - a[1]) * (c[2] - a[2]) - (c[1] - a[1]) * (b[2] - a[2]);
  double n1 = (b[2] - a[2]) * (c[0] - a[0]) - (c[2] - a[2]) * (b[0] - a[0]);
  double n2 = (b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1]);
  return sqrt(n0 * n0 + n1 * n1 + n2 * n2) * 0.5;
}

/* ------------------------------------------------------------------------ */
/* axis-aligned boxes (ref: aabb.h:16-148) */

typedef struct {
  double x0, x1, y0, y1, z0, z1; /* min/max per axis */
  int valid;
} box;

static inline void box_add(box* b, double x, double y, double z, double eps) {
  /* CAABB3D::AddPoint, aabb.h:102-122 */
  if (b->valid) {
    b->x0 = (b->x0 < x - eps) ? b->x0 : x - eps;
    b->y0 = (b->y0 < y - eps) ? b->y0 : y - eps;
    b->z0 = (b->z0 < z - eps) ? b->z0 : z - eps;
    b->x1 = (b->x1 > x + eps) ? b->x1 : x + eps;
    b->y1 = (b->y1 > y + eps) ? b->y1 : y + eps;
    b->z1 = (b->z1 > z + eps) ? b->z1 : z + eps;
  } else {
    b->valid = 1;
    b->x0 = x - eps;
    b->y0 = y - eps;
    b->z0 = z - eps;
    b->x1 = x + eps;
    b->y1 = y + eps;
    b->z1 = z + eps;
  }
}
static inline void box_addv(box* b, v3 p, double eps) { box_add(b, p.x, p.y, p.z, eps); }
static inline void box_union(box* b, const box* o) { /* operator+=, aabb.h:55-72 */
  if (!o->valid) return;
  if (!b->valid) {
    *b = *o;
    return;
  }
  b->x1 = (b->x1 > o->x1) ? b->x1 : o->x1;
  b->x0 = (b->x0 < o->x0) ? b->x0 : o->x0;
  b->y1 = (b->y1 > o->y1) ? b->y1 : o->y1;
  b->y0 = (b->y0 < o->y0) ? b->y0 : o->y0;
  b->z1 = (b->z1 > o->z1) ? b->z1 : o->z1;
  b->z0 = (b->z0 < o->z0) ? b->z0 : o->z0;
}
static inline int box_overlap(const box* a, const box* b) { /* aabb.h:73-84 */
  if (!a->valid || !b->valid) return 0;
  if (a->x1 < b->x0) return 0;
  if (a->x0 > b->x1) return 0;
  if (a->y1 < b->y0) return 0;
  if (a->y0 > b->y1) return 0;
  if (a->z1 < b->z0) return 0;
  if (a->z0 > b->z1) return 0;
  return 1;
}
static inline int box_inside(const box* b, v3 p) { /* aabb.h:137-144 */
  if (!b->valid) return 0;
  return p.x >= b->x0 && p.x <= b->x1 && p.y >= b->y0 && p.y <= b->y1 && p.z >= b->z0 &&
         p.z <= b->z1;
}

/* ------------------------------------------------------------------------ */
/* element physics (ref: cloth_internal_physics.h) */

/* PSD part of a symmetric 2x2 stress (xx, yy, xy); ref: :15-67 */
void orc_psd22(const double* s2, double* s3) {
  const double b = (s2[0] + s2[1]) * 0.5;
  const double d = (s2[0] - s2[1]) * (s2[0] - s2[1]) * 0.25 + s2[2] * s2[2];
  const double e = sqrt(d);
  if (b - e > 1.0e-20) {
    s3[0] = s2[0];
    s3[1] = s2[1];
    s3[2] = s2[2];
    return;
  }
  if (b + e < 0) {
    s3[0] = s3[1] = s3[2] = 0;
    return;
  }
  const double l = b + e; /* the one non-negative eigenvalue */
  double u[2] = {s2[0] - l, s2[2]};
  double w[2] = {s2[2], s2[1] - l};
  const double su = u[0] * u[0] + u[1] * u[1];
  const double sw = w[0] * w[0] + w[1] * w[1];
  double* t = (su > sw) ? u : w;
  const double st = (su > sw) ? su : sw;
  if (st < 1.0e-20) {
    s3[0] = s3[1] = s3[2] = 0;
    return;
  }
  const double inv = 1.0 / sqrt(st);
  t[0] *= inv;
  t[1] *= inv;
  s3[0] = l * t[0] * t[0];
  s3[1] = l * t[1] * t[1];
  s3[2] = l * t[0] * t[1];
}

/* shape-function derivatives of the linear triangle: N0=1-r-s, N1=r, N2=s */
static const double kdN[3][2] = {{-1.0, -1.0}, {+1.0, +0.0}, {+0.0, +1.0}};

/* The 16 terms of the material Hessian in the reference's summation order
 * (ref: cloth_internal_physics.h:146-161).  Term k is
 *   d[a][i] * dN[ino][b] * Cons[r][s] * d[c][j] * dN[jno][e]
 * with (a,b,r,s,c,e) = kMat[k]. */
static const int kMat[16][6] = {
    {0, 0, 0, 0, 0, 0}, {0, 0, 0, 1, 1, 1}, {0, 0, 0, 2, 0, 1}, {0, 0, 0, 2, 1, 0},
    {1, 1, 1, 0, 0, 0}, {1, 1, 1, 1, 1, 1}, {1, 1, 1, 2, 0, 1}, {1, 1, 1, 2, 1, 0},
    {0, 1, 2, 0, 0, 0}, {0, 1, 2, 1, 1, 1}, {0, 1, 2, 2, 0, 1}, {0, 1, 2, 2, 1, 0},
    {1, 0, 2, 0, 0, 0}, {1, 0, 2, 1, 1, 1}, {1, 0, 2, 2, 0, 1}, {1, 0, 2, 2, 1, 0}};

/* StVK constant-strain triangle; ref: cloth_internal_physics.h:70-176.
 * C, c: 3x3 rest / current corners; dW[3][3]; ddW[3][3][3][3] (ino,jno,i,j). */
void orc_cst(const double* C, const double* c, double lam, double mu, double* W, double* dW,
             double* ddW) {
  double D[3][3];
  for (int k = 0; k < 3; k++) {
    D[0][k] = C[3 + k] - C[k];
    D[1][k] = C[6 + k] - C[k];
  }
  /* unit normal and area of the rest triangle (utility.h:58-73) */
  double area;
  {
    const double *a = C, *b = C + 3, *e = C + 6;
    D[2][0] = (b[1] - a[1]) * (e[2] - a[2]) - (e[1] - a[1]) * (b[2] - a[2]);
    D[2][1] = (b[2] - a[2]) * (e[0] - a[0]) - (e[2] - a[2]) * (b[0] - a[0]);
    D[2][2] = (b[0] - a[0]) * (e[1] - a[1]) - (e[0] - a[0]) * (b[1] - a[1]);
    area = sqrt(D[2][0] * D[2][0] + D[2][1] * D[2][1] + D[2][2] * D[2][2]) * 0.5;
    const double inv = 0.5 / area;
    D[2][0] *= inv;
    D[2][1] *= inv;
    D[2][2] *= inv;
  }
  /* dual (contravariant) basis of the rest edges */
  double G[2][3];
  cross3(G[0], D[1], D[2]);
  {
    const double s = 1.0 / dot3(G[0], D[0]);
    G[0][0] *= s;
    G[0][1] *= s;
    G[0][2] *= s;
  }
  cross3(G[1], D[2], D[0]);
  {
    const double s = 1.0 / dot3(G[1], D[1]);
    G[1][0] *= s;
    G[1][1] *= s;
    G[1][2] *= s;
  }
  double d[2][3];
  for (int k = 0; k < 3; k++) {
    d[0][k] = c[3 + k] - c[k];
    d[1][k] = c[6 + k] - c[k];
  }
  /* Green strain in engineering notation (xx, yy, 2xy) */
  const double E[3] = {0.5 * (dot3(d[0], d[0]) - dot3(D[0], D[0])),
                       0.5 * (dot3(d[1], d[1]) - dot3(D[1], D[1])),
                       1.0 * (dot3(d[0], d[1]) - dot3(D[0], D[1]))};
  /* contravariant metric g^ab */
  const double g0 = dot3(G[0], G[0]), g1 = dot3(G[1], G[1]), g2 = dot3(G[1], G[0]);
  double K[3][3];
  K[0][0] = lam * g0 * g0 + 2 * mu * (g0 * g0);
  K[0][1] = lam * g0 * g1 + 2 * mu * (g2 * g2);
  K[0][2] = lam * g0 * g2 + 2 * mu * (g0 * g2);
  K[1][0] = lam * g1 * g0 + 2 * mu * (g2 * g2);
  K[1][1] = lam * g1 * g1 + 2 * mu * (g1 * g1);
  K[1][2] = lam * g1 * g2 + 2 * mu * (g2 * g1);
  K[2][0] = lam * g2 * g0 + 2 * mu * (g0 * g2);
  K[2][1] = lam * g2 * g1 + 2 * mu * (g2 * g1);
  K[2][2] = lam * g2 * g2 + 1 * mu * (g0 * g1 + g2 * g2);
  double S[3];
  for (int r = 0; r < 3; r++) S[r] = K[r][0] * E[0] + K[r][1] * E[1] + K[r][2] * E[2];
  *W = 0.5 * area * (E[0] * S[0] + E[1] * S[1] + E[2] * S[2]);
  for (int a = 0; a < 3; a++) {
    for (int i = 0; i < 3; i++) {
      dW[a * 3 + i] = area * (+S[0] * d[0][i] * kdN[a][0] + S[2] * d[0][i] * kdN[a][1] +
                              S[2] * d[1][i] * kdN[a][0] + S[1] * d[1][i] * kdN[a][1]);
    }
  }
  double P[3];
  orc_psd22(S, P);
  for (int a = 0; a < 3; a++) {
    for (int b = 0; b < 3; b++) {
      for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
          double acc = 0;
          for (int k = 0; k < 16; k++) {
            const int* m = kMat[k];
            acc += d[m[0]][i] * kdN[a][m[1]] * K[m[2]][m[3]] * d[m[4]][j] * kdN[b][m[5]];
          }
          ddW[((a * 3 + b) * 3 + i) * 3 + j] = acc * area;
        }
      }
      const double geo = area * (+P[0] * kdN[a][0] * kdN[b][0] + P[2] * kdN[a][0] * kdN[b][1] +
                                 P[2] * kdN[a][1] * kdN[b][0] + P[1] * kdN[a][1] * kdN[b][1]);
      for (int i = 0; i < 3; i++) ddW[((a * 3 + b) * 3 + i) * 3 + i] += geo;
    }
  }
}

/* rest-shape cotangent weights of a bending stencil [opp0, opp1, e0, e1]
 * and the stiffness scale; ref: cloth_internal_physics.h:188-215 */
static void bend_weights(const double* C, double stiff, double K[4], double* scale) {
  const double* C0 = C;
  const double* C1 = C + 3;
  const double* C2 = C + 6;
  const double* C3 = C + 9;
  const double A0 = tri_area(C0, C2, C3);
  const double A1 = tri_area(C1, C3, C2);
  const double L0 = sqrt((C3[0] - C2[0]) * (C3[0] - C2[0]) + (C3[1] - C2[1]) * (C3[1] - C2[1]) +
                         (C3[2] - C2[2]) * (C3[2] - C2[2]));
  const double H0 = A0 * 2.0 / L0;
  const double H1 = A1 * 2.0 / L0;
  double e23[3], e02[3], e03[3], e12[3], e13[3];
  for (int k = 0; k < 3; k++) {
    e23[k] = C3[k] - C2[k];
    e02[k] = C2[k] - C0[k];
    e03[k] = C3[k] - C0[k];
    e12[k] = C2[k] - C1[k];
    e13[k] = C3[k] - C1[k];
  }
  const double c023 = -dot3(e02, e23) / H0;
  const double c032 = +dot3(e03, e23) / H0;
  const double c123 = -dot3(e12, e23) / H1;
  const double c132 = +dot3(e13, e23) / H1;
  *scale = stiff / ((A0 + A1) * L0 * L0);
  K[0] = -c023 - c032;
  K[1] = -c123 - c132;
  K[2] = c032 + c132;
  K[3] = c023 + c123;
}

/* quadratic bending; ref: cloth_internal_physics.h:180-238.
 * dW[4][3]; ddW[4][4][3][3].  (The reference also accumulates an energy into
 * an uninitialised caller variable; energy is not part of the step's output.) */
void orc_bend(const double* C, const double* c, double stiff, double* dW, double* ddW) {
  double K[4], s;
  bend_weights(C, stiff, K, &s);
  for (int i = 0; i < 144; i++) ddW[i] = 0;
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) {
      const double v = K[a] * K[b] * s;
      for (int i = 0; i < 3; i++) ddW[((a * 4 + b) * 3 + i) * 3 + i] = v;
    }
  for (int a = 0; a < 4; a++)
    for (int i = 0; i < 3; i++) {
      double acc = 0;
      for (int b = 0; b < 4; b++)
        for (int j = 0; j < 3; j++) acc += ddW[((a * 4 + b) * 3 + i) * 3 + j] * c[b * 3 + j];
      dW[a * 3 + i] = acc;
    }
}

/* obstacle penetration depth: pd > 0 inside, n = push-out direction.
 * Plane and sphere follow the demo callbacks (self_contact_sparse/main.cpp:110-133,
 * internal_cloth_sparse/main.cpp:113-132). */
static void obstacle_depth(const double* obst, const double* p, double* pd, double* n) {
  const int kind = (int)obst[0];
  if (kind == 1) {
    n[0] = obst[1];
    n[1] = obst[2];
    n[2] = obst[3];
    *pd = obst[4] - (n[0] * p[0] + n[1] * p[1] + n[2] * p[2]);
  } else if (kind == 2 || kind == 3) {
    n[0] = p[0] - obst[1];
    n[1] = p[1] - obst[2];
    n[2] = p[2] - obst[3];
    const double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] /= len;
    n[1] /= len;
    n[2] /= len;
    *pd = obst[4] - len;
    if (kind == 3) {
      n[0] *= -1;
      n[1] *= -1;
      n[2] *= -1;
      *pd *= -1;
    }
  } else if (kind == 4) {
    /* axis-aligned box (center obst[1..3], half-extents obst[4..6]); not in
     * the reference, which has only plane and sphere callbacks: the scene
     * convention pd > 0 inside, n = push-out direction */
    double s[3], d[3];
    for (int i = 0; i < 3; i++) {
      s[i] = p[i] - obst[1 + i];
      d[i] = fabs(s[i]) - obst[4 + i];
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
    /* bar: the capsule around the segment obst[1..3]-obst[4..6] with radius
     * obst[7]; kind 6 adds the ground z = obst[8] (normal +z), the deeper of
     * the two, the bar on ties.  Not in the reference, same convention. */
    double e[3], w[3];
    for (int i = 0; i < 3; i++) {
      e[i] = obst[4 + i] - obst[1 + i];
      w[i] = p[i] - obst[1 + i];
    }
    const double ee = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
    double t = (w[0] * e[0] + w[1] * e[1] + w[2] * e[2]) / ee;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    for (int i = 0; i < 3; i++) n[i] = w[i] - t * e[i];
    const double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] /= len;
    n[1] /= len;
    n[2] /= len;
    *pd = obst[7] - len;
    if (kind == 6) {
      const double pg = obst[8] - p[2];
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

/* penalty contact against the obstacle; ref: cloth_internal_physics.h:243-280 */
void orc_contact_elem(const double* c, double k, double clearance, const double* obst, double* W,
                      double* dW, double* ddW) {
  double pd, n[3];
  obstacle_depth(obst, c, &pd, n);
  pd += clearance;
  if (pd < 0) {
    *W = 0;
    for (int i = 0; i < 3; i++) dW[i] = 0;
    for (int i = 0; i < 9; i++) ddW[i] = 0;
    return;
  }
  *W = 0.5 * k * pd * pd;
  for (int i = 0; i < 3; i++) dW[i] = -k * pd * n[i];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) ddW[i * 3 + j] = k * n[i] * n[j];
}

/* ------------------------------------------------------------------------ */
/* context */

typedef struct {
  int* zone;  /* sorted vertex ids */
  int n, cap;
} zone_t;

typedef struct {
  int np, nt, nq;
  double *xyz0;
  int *tri, *quad, *bc;
  double lam, mu, bend, grav[3], mass, kc, clear, dt, tol;
  int maxit;
  double obst[9];
  /* 3x3-block CSR without the diagonal + separate diagonal (ref: matrix_square_sparse.h:14-48) */
  int nnz;
  int *rowptr, *colind, *diaind;
  double *dia, *crs;    /* A */
  double *fdia, *fcrs;  /* ILU factor */
  int* marker;          /* np, -1 */
  /* BVH */
  int nnode, root;
  int* node;  /* 3 per node: parent, child0, child1 (child1 == -1: leaf, child0 = triangle) */
  box* bb;
  /* triangle 1-ring of vertices (used by impact zones) */
  int *eidx, *earr;
  int dot_mode;
  /* diagnostics of the last internal solve */
  double last_conv;
  int last_iter;
} ctx_t;

static void* xmalloc(size_t n) {
  void* p = malloc(n ? n : 1);
  if (!p) abort();
  return p;
}
static void* xcalloc(size_t n, size_t s) {
  void* p = calloc(n ? n : 1, s ? s : 1);
  if (!p) abort();
  return p;
}

/* vertex -> elements in element order (ref: jagged_array.h:96-126) */
static void node_to_elem(const int* elem, int ne, int nn, int np, int** pidx, int** parr) {
  int* idx = (int*)xcalloc(np + 1, sizeof(int));
  for (int e = 0; e < ne; e++)
    for (int k = 0; k < nn; k++) idx[elem[e * nn + k] + 1]++;
  for (int i = 0; i < np; i++) idx[i + 1] += idx[i];
  int* arr = (int*)xmalloc(sizeof(int) * idx[np]);
  for (int e = 0; e < ne; e++)
    for (int k = 0; k < nn; k++) {
      int v = elem[e * nn + k];
      arr[idx[v]++] = e;
    }
  for (int i = np; i > 0; i--) idx[i] = idx[i - 1];
  idx[0] = 0;
  *pidx = idx;
  *parr = arr;
}

/* vertex adjacency through element stencils, self excluded, in discovery
 * order (ref: jagged_array.h:158-207 with is_dia=false) */
static void edge_of_elem(const int* elem, int ne, int nn, int np, int** pidx, int** parr) {
  int *ni, *na;
  node_to_elem(elem, ne, nn, np, &ni, &na);
  int* flag = (int*)xmalloc(sizeof(int) * (np ? np : 1));
  int* idx = (int*)xcalloc(np + 1, sizeof(int));
  for (int i = 0; i < np; i++) flag[i] = -1;
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      for (int i = 0; i < np; i++) idx[i + 1] += idx[i];
      for (int i = 0; i < np; i++) flag[i] = -1;
    }
    int* arr = (pass == 1) ? (int*)xmalloc(sizeof(int) * (idx[np] ? idx[np] : 1)) : NULL;
    for (int v = 0; v < np; v++) {
      flag[v] = v;
      for (int k = ni[v]; k < ni[v + 1]; k++) {
        const int e = na[k];
        for (int j = 0; j < nn; j++) {
          const int w = elem[e * nn + j];
          if (flag[w] != v) {
            flag[w] = v;
            if (pass == 0)
              idx[v + 1]++;
            else
              arr[idx[v]++] = w;
          }
        }
      }
    }
    if (pass == 1) {
      for (int i = np; i > 0; i--) idx[i] = idx[i - 1];
      idx[0] = 0;
      *parr = arr;
    }
  }
  *pidx = idx;
  free(flag);
  free(ni);
  free(na);
}

/* per-row ascending order (the reference uses a bubble sort, jagged_array.h:71-90;
 * any sort gives the same sorted rows) */
static int cmp_int(const void* a, const void* b) {
  int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

/* triangle adjacency across oppositely oriented edges (ref: bvh_aabb.cpp:109-138) */
static int* tri_adjacency(const int* tri, int nt, int np) {
  static const int e2n[3][2] = {{1, 2}, {2, 0}, {0, 1}};
  int *ni, *na;
  node_to_elem(tri, nt, 3, np, &ni, &na);
  int* sur = (int*)xmalloc(sizeof(int) * 3 * (nt ? nt : 1));
  for (int i = 0; i < 3 * nt; i++) sur[i] = -1;
  for (int t = 0; t < nt; t++)
    for (int e = 0; e < 3; e++) {
      const int a = tri[t * 3 + e2n[e][0]], b = tri[t * 3 + e2n[e][1]];
      for (int k = ni[a]; k < ni[a + 1]; k++) {
        const int u = na[k];
        if (u == t) continue;
        for (int f = 0; f < 3; f++) {
          if (a == tri[u * 3 + e2n[f][1]] && b == tri[u * 3 + e2n[f][0]]) {
            sur[t * 3 + e] = u;
            break;
          }
        }
        if (sur[t * 3 + e] == u) break;
      }
    }
  free(ni);
  free(na);
  return sur;
}

/* ---- BVH topology: top-down connectivity-preserving split
 * (ref: bvh_aabb.cpp:142-313) */

typedef struct {
  int* node;
  int n, cap;
} nodes_t;

static int nodes_push2(nodes_t* s, int parent) {
  if (s->n + 2 > s->cap) {
    s->cap = s->cap ? s->cap * 2 : 64;
    s->node = (int*)realloc(s->node, sizeof(int) * 3 * s->cap);
    if (!s->node) abort();
  }
  int a = s->n;
  for (int k = 0; k < 2; k++) {
    s->node[(a + k) * 3 + 0] = parent;
    s->node[(a + k) * 3 + 1] = 0;
    s->node[(a + k) * 3 + 2] = 0;
  }
  s->n += 2;
  return a;
}

static v3 tri_centroid(const int* tri, const double* xyz, int t) {
  const int a = tri[t * 3 + 0], b = tri[t * 3 + 1], c = tri[t * 3 + 2];
  return V((xyz[a * 3 + 0] + xyz[b * 3 + 0] + xyz[c * 3 + 0]) / 3.0,
           (xyz[a * 3 + 1] + xyz[b * 3 + 1] + xyz[c * 3 + 1]) / 3.0,
           (xyz[a * 3 + 2] + xyz[b * 3 + 2] + xyz[c * 3 + 2]) / 3.0);
}

static void bvh_split(int parent, int* owner, nodes_t* s, const int* list, int nlist,
                      const int* tri, const double* xyz, const int* sur, int* stack) {
  const double eps = 1.0e-10;
  box b;
  memset(&b, 0, sizeof(b));
  for (int i = 0; i < nlist; i++) box_addv(&b, tri_centroid(tri, xyz, list[i]), eps);
  const double lx = b.x1 - b.x0, ly = b.y1 - b.y0, lz = b.z1 - b.z0;
  v3 dir = V(0, 0, 0); /* the strictly longest axis; (0,0,0) on ties */
  if (lx > ly && lx > lz) dir = V(1, 0, 0);
  if (ly > lz && ly > lx) dir = V(0, 1, 0);
  if (lz > lx && lz > ly) dir = V(0, 0, 1);
  v3 org = V((b.x0 + b.x1) * 0.5, (b.y0 + b.y1) * 0.5, (b.z0 + b.z1) * 0.5);
  int seed = -1;
  for (int i = 0; i < nlist; i++) {
    const v3 g = tri_centroid(tri, xyz, list[i]);
    if (fabs(vdot(vsub(g, org), dir)) < 1.0e-10) continue;
    if (vdot(vsub(g, org), dir) < 0) dir = vscl(-1, dir);
    seed = list[i];
    break;
  }
  if (seed == -1) {
    /* principal axis of the centroid covariance by 10 power iterations from (1,1,1) */
    org = V(0, 0, 0);
    for (int i = 0; i < nlist; i++) org = vadd(org, tri_centroid(tri, xyz, list[i]));
    {
      const double dn = (double)nlist;
      if (!(fabs(dn) < 1.e-16)) org = V(org.x / dn, org.y / dn, org.z / dn);
    }
    double m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < nlist; i++) {
      const v3 g = vsub(tri_centroid(tri, xyz, list[i]), org);
      for (int r = 0; r < 3; r++)
        for (int q = 0; q < 3; q++) m[r][q] += vcomp(g, r) * vcomp(g, q);
    }
    double dv[3] = {1, 1, 1};
    for (int it = 0; it < 10; it++) {
      double t[3];
      for (int r = 0; r < 3; r++) t[r] = m[r][0] * dv[0] + m[r][1] * dv[1] + m[r][2] * dv[2];
      const double len = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
      for (int r = 0; r < 3; r++) dv[r] = t[r] / len;
    }
    dir = V(dv[0], dv[1], dv[2]);
    for (int i = 0; i < nlist; i++) {
      const v3 g = tri_centroid(tri, xyz, list[i]);
      if (fabs(vdot(vsub(g, org), dir)) < 1.0e-10) continue;
      if (vdot(vsub(g, org), dir) < 0) dir = vscl(-1, dir);
      seed = list[i];
      break;
    }
  }
  if (seed < 0) abort(); /* degenerate cluster (reference: undefined behaviour) */
  const int c0 = nodes_push2(s, parent);
  const int c1 = c0 + 1;
  s->node[parent * 3 + 1] = c0;
  s->node[parent * 3 + 2] = c1;
  /* child 0: flood fill from the seed over edge-adjacent triangles on the
   * positive side, LIFO order */
  int* l0 = (int*)xmalloc(sizeof(int) * nlist);
  int n0 = 0, sp = 0;
  owner[seed] = c0;
  l0[n0++] = seed;
  stack[sp++] = seed;
  while (sp > 0) {
    const int t = stack[--sp];
    for (int e = 0; e < 3; e++) {
      const int u = sur[t * 3 + e];
      if (u == -1) continue;
      if (owner[u] != parent) continue;
      const v3 g = tri_centroid(tri, xyz, u);
      if (vdot(vsub(g, org), dir) < 0) continue;
      stack[sp++] = u;
      owner[u] = c0;
      l0[n0++] = u;
    }
  }
  int* l1 = (int*)xmalloc(sizeof(int) * nlist);
  int n1 = 0;
  for (int i = 0; i < nlist; i++) {
    const int t = list[i];
    if (owner[t] == c0) continue;
    owner[t] = c1;
    l1[n1++] = t;
  }
  if (n1 == 0) abort(); /* reference asserts a non-empty second child */
  if (n0 == 1) {
    s->node[c0 * 3 + 1] = l0[0];
    s->node[c0 * 3 + 2] = -1;
  } else {
    bvh_split(c0, owner, s, l0, n0, tri, xyz, sur, stack);
  }
  free(l0);
  if (n1 == 1) {
    s->node[c1 * 3 + 1] = l1[0];
    s->node[c1 * 3 + 2] = -1;
  } else {
    bvh_split(c1, owner, s, l1, n1, tri, xyz, sur, stack);
  }
  free(l1);
}

static void build_bvh(ctx_t* c, const double* xyz) {
  int* sur = tri_adjacency(c->tri, c->nt, c->np);
  nodes_t s;
  memset(&s, 0, sizeof(s));
  nodes_push2(&s, -1);
  s.n = 1; /* only the root; push2 reserved room for two */
  int* list = (int*)xmalloc(sizeof(int) * c->nt);
  for (int i = 0; i < c->nt; i++) list[i] = i;
  int* owner = (int*)xcalloc(c->nt, sizeof(int));
  int* stack = (int*)xmalloc(sizeof(int) * (c->nt + 1));
  bvh_split(0, owner, &s, list, c->nt, c->tri, xyz, sur, stack);
  c->node = s.node;
  c->nnode = s.n;
  c->root = 0;
  free(list);
  free(owner);
  free(stack);
  free(sur);
}

/* ------------------------------------------------------------------------ */

void* orc_create(int np, const double* xyz0, const double* xyz_init, int nt, const int* tri,
                 int nq, const int* quad, const int* bc, const double* params, const double* obst) {
  ctx_t* c = (ctx_t*)xcalloc(1, sizeof(ctx_t));
  c->np = np;
  c->nt = nt;
  c->nq = nq;
  c->xyz0 = (double*)xmalloc(sizeof(double) * 3 * np);
  memcpy(c->xyz0, xyz0, sizeof(double) * 3 * np);
  c->tri = (int*)xmalloc(sizeof(int) * 3 * (nt ? nt : 1));
  memcpy(c->tri, tri, sizeof(int) * 3 * nt);
  c->quad = (int*)xmalloc(sizeof(int) * 4 * (nq ? nq : 1));
  memcpy(c->quad, quad, sizeof(int) * 4 * nq);
  c->bc = (int*)xmalloc(sizeof(int) * np);
  memcpy(c->bc, bc, sizeof(int) * np);
  c->lam = params[0];
  c->mu = params[1];
  c->bend = params[2];
  c->grav[0] = params[3];
  c->grav[1] = params[4];
  c->grav[2] = params[5];
  c->mass = params[6];
  c->kc = params[7];
  c->clear = params[8];
  c->dt = params[9];
  c->tol = params[10];
  c->maxit = (int)params[11];
  memcpy(c->obst, obst, sizeof(double) * 9);
  /* setup order of self_contact_sparse/main.cpp:574-582 */
  build_bvh(c, xyz_init);
  edge_of_elem(c->tri, nt, 3, np, &c->eidx, &c->earr);
  edge_of_elem(c->quad, nq, 4, np, &c->rowptr, &c->colind);
  for (int i = 0; i < np; i++)
    qsort(c->colind + c->rowptr[i], c->rowptr[i + 1] - c->rowptr[i], sizeof(int), cmp_int);
  c->nnz = c->rowptr[np];
  /* first entry right of the diagonal (ref: ilu_sparse.cpp:114-134) */
  c->diaind = (int*)xmalloc(sizeof(int) * np);
  for (int i = 0; i < np; i++) {
    c->diaind[i] = c->rowptr[i + 1];
    for (int k = c->rowptr[i]; k < c->rowptr[i + 1]; k++)
      if (c->colind[k] > i) {
        c->diaind[i] = k;
        break;
      }
  }
  c->dia = (double*)xcalloc(9 * np, sizeof(double));
  c->crs = (double*)xcalloc(9 * (size_t)c->nnz, sizeof(double));
  c->fdia = (double*)xcalloc(9 * np, sizeof(double));
  c->fcrs = (double*)xcalloc(9 * (size_t)c->nnz, sizeof(double));
  c->marker = (int*)xmalloc(sizeof(int) * np);
  for (int i = 0; i < np; i++) c->marker[i] = -1;
  c->bb = (box*)xcalloc(c->nnode, sizeof(box));
  return c;
}

void orc_destroy(void* h) {
  ctx_t* c = (ctx_t*)h;
  free(c->xyz0);
  free(c->tri);
  free(c->quad);
  free(c->bc);
  free(c->rowptr);
  free(c->colind);
  free(c->diaind);
  free(c->dia);
  free(c->crs);
  free(c->fdia);
  free(c->fcrs);
  free(c->marker);
  free(c->node);
  free(c->bb);
  free(c->eidx);
  free(c->earr);
  free(c);
}

int orc_nnz(void* h) { return ((ctx_t*)h)->nnz; }
int orc_nnode(void* h) { return ((ctx_t*)h)->nnode; }
int orc_root(void* h) { return ((ctx_t*)h)->root; }
int orc_nedge(void* h) {
  ctx_t* c = (ctx_t*)h;
  return c->eidx[c->np];
}
void orc_set_dot_mode(void* h, int mode) { ((ctx_t*)h)->dot_mode = mode; }

void orc_get_pattern(void* h, int* rowptr, int* colind, int* diaind) {
  ctx_t* c = (ctx_t*)h;
  memcpy(rowptr, c->rowptr, sizeof(int) * (c->np + 1));
  memcpy(colind, c->colind, sizeof(int) * c->nnz);
  memcpy(diaind, c->diaind, sizeof(int) * c->np);
}
void orc_get_bvh(void* h, int* nodes) {
  ctx_t* c = (ctx_t*)h;
  memcpy(nodes, c->node, sizeof(int) * 3 * c->nnode);
}
void orc_get_edge(void* h, int* index, int* array) {
  ctx_t* c = (ctx_t*)h;
  memcpy(index, c->eidx, sizeof(int) * (c->np + 1));
  memcpy(array, c->earr, sizeof(int) * c->eidx[c->np]);
}

/* ------------------------------------------------------------------------ */
/* block sparse matrix (ref: matrix_square_sparse.cpp) */

/* scatter an n-vertex element matrix laid out [n][n][3][3]; entries outside
 * the pattern are dropped (ref: :88-138) */
static void merge(ctx_t* c, double* dia, double* crs, int n, const int* ip, const double* em) {
  for (int a = 0; a < n; a++) {
    const int i = ip[a];
    for (int k = c->rowptr[i]; k < c->rowptr[i + 1]; k++) c->marker[c->colind[k]] = k;
    for (int b = 0; b < n; b++) {
      const double* src = em + (a * n + b) * 9;
      double* dst;
      if (a == b) {
        dst = dia + i * 9;
      } else {
        const int k = c->marker[ip[b]];
        if (k == -1) continue;
        dst = crs + (size_t)k * 9;
      }
      for (int q = 0; q < 9; q++) dst[q] += src[q];
    }
    for (int k = c->rowptr[i]; k < c->rowptr[i + 1]; k++) c->marker[c->colind[k]] = -1;
  }
}

/* y = alpha*A*x + beta*y (ref: :167-243, 3x3 branch :218-242) */
static void matvec(const ctx_t* c, const double* dia, const double* crs, double alpha,
                   const double* x, double beta, double* y) {
  for (int i = 0; i < c->np; i++) {
    y[i * 3 + 0] *= beta;
    y[i * 3 + 1] *= beta;
    y[i * 3 + 2] *= beta;
    for (int k = c->rowptr[i]; k < c->rowptr[i + 1]; k++) {
      const int j = c->colind[k];
      const double* v = crs + (size_t)k * 9;
      y[i * 3 + 0] += alpha * (v[0] * x[j * 3 + 0] + v[1] * x[j * 3 + 1] + v[2] * x[j * 3 + 2]);
      y[i * 3 + 1] += alpha * (v[3] * x[j * 3 + 0] + v[4] * x[j * 3 + 1] + v[5] * x[j * 3 + 2]);
      y[i * 3 + 2] += alpha * (v[6] * x[j * 3 + 0] + v[7] * x[j * 3 + 1] + v[8] * x[j * 3 + 2]);
    }
    const double* v = dia + i * 9;
    y[i * 3 + 0] += alpha * (v[0] * x[i * 3 + 0] + v[1] * x[i * 3 + 1] + v[2] * x[i * 3 + 2]);
    y[i * 3 + 1] += alpha * (v[3] * x[i * 3 + 0] + v[4] * x[i * 3 + 1] + v[5] * x[i * 3 + 2]);
    y[i * 3 + 2] += alpha * (v[6] * x[i * 3 + 0] + v[7] * x[i * 3 + 1] + v[8] * x[i * 3 + 2]);
  }
}

/* fixed rows and columns zeroed, identity diagonal (ref: :272-302) */
static void set_bc(const ctx_t* c, double* dia, double* crs) {
  for (int i = 0; i < c->np; i++) {
    if (c->bc[i] == 0) continue;
    for (int r = 0; r < 3; r++) {
      for (int q = 0; q < 3; q++) dia[i * 9 + r * 3 + q] = 0.0;
      dia[i * 9 + r * 3 + r] = 1.0;
    }
    for (int k = c->rowptr[i]; k < c->rowptr[i + 1]; k++)
      for (int q = 0; q < 9; q++) crs[(size_t)k * 9 + q] = 0.0;
  }
  for (int k = 0; k < c->nnz; k++) {
    if (c->bc[c->colind[k]] == 0) continue;
    for (int q = 0; q < 9; q++) crs[(size_t)k * 9 + q] = 0.0;
  }
}

/* (ref: matrix_square_sparse.cpp:313-336) */
static double dot_seq(const double* a, const double* b, int n) {
  double r = 0.0;
  for (int i = 0; i < n; i++) r += a[i] * b[i];
  return r;
}
/* blocked pairwise sum: a deliberately different rounding order, used only to
 * measure how sensitive trajectories are to the order of the PCG reductions */
static double dot_pair(const double* a, const double* b, int n) {
  if (n <= 64) return dot_seq(a, b, n);
  int h = n / 2;
  return dot_pair(a, b, h) + dot_pair(a + h, b + h, n - h);
}
static double dot_(const ctx_t* c, const double* a, const double* b, int n) {
  return c->dot_mode == 1 ? dot_pair(a, b, n) : dot_seq(a, b, n);
}

/* ------------------------------------------------------------------------ */
/* ILU(0) with 3x3 blocks (ref: ilu_sparse.cpp) */

/* cofactor inverse (ref: ilu_sparse.cpp:77-97) */
static void inv3_inplace(double* a) {
  const double det = +a[0] * a[4] * a[8] + a[3] * a[7] * a[2] + a[6] * a[1] * a[5] -
                     a[0] * a[7] * a[5] - a[6] * a[4] * a[2] - a[3] * a[1] * a[8];
  const double id = 1.0 / det;
  double t[9];
  memcpy(t, a, sizeof(t));
  a[0] = id * (t[4] * t[8] - t[5] * t[7]);
  a[1] = id * (t[2] * t[7] - t[1] * t[8]);
  a[2] = id * (t[1] * t[5] - t[2] * t[4]);
  a[3] = id * (t[5] * t[6] - t[3] * t[8]);
  a[4] = id * (t[0] * t[8] - t[2] * t[6]);
  a[5] = id * (t[2] * t[3] - t[0] * t[5]);
  a[6] = id * (t[3] * t[7] - t[4] * t[6]);
  a[7] = id * (t[1] * t[6] - t[0] * t[7]);
  a[8] = id * (t[0] * t[4] - t[1] * t[3]);
}

/* numeric factorisation, IKJ order; the diagonal blocks hold D^-1 and the
 * upper blocks D^-1 U.  More than 10 singular pivots end the factorisation
 * early (ref: ilu_sparse.cpp:387-401, 507-569). Returns the singular count. */
static int ilu_factor(ctx_t* c, double* vd, double* vc) {
  int nsing = 0;
  int* mark = c->marker;
  for (int i = 0; i < c->np; i++) {
    for (int k = c->rowptr[i]; k < c->rowptr[i + 1]; k++) mark[c->colind[k]] = k;
    for (int ik = c->rowptr[i]; ik < c->diaind[i]; ik++) {
      const int kr = c->colind[ik];
      const double* lik = vc + (size_t)ik * 9;
      for (int kj = c->diaind[kr]; kj < c->rowptr[kr + 1]; kj++) {
        const int j = c->colind[kj];
        const double* ukj = vc + (size_t)kj * 9;
        double* aij;
        if (j != i) {
          const int ij = mark[j];
          if (ij == -1) continue;
          aij = vc + (size_t)ij * 9;
        } else {
          aij = vd + i * 9;
        }
        for (int r = 0; r < 3; r++) {
          aij[r * 3 + 0] -= lik[r * 3 + 0] * ukj[0] + lik[r * 3 + 1] * ukj[3] + lik[r * 3 + 2] * ukj[6];
          aij[r * 3 + 1] -= lik[r * 3 + 0] * ukj[1] + lik[r * 3 + 1] * ukj[4] + lik[r * 3 + 2] * ukj[7];
          aij[r * 3 + 2] -= lik[r * 3 + 0] * ukj[2] + lik[r * 3 + 1] * ukj[5] + lik[r * 3 + 2] * ukj[8];
        }
      }
    }
    double* dii = vd + i * 9;
    const double det = +dii[0] * dii[4] * dii[8] + dii[3] * dii[7] * dii[2] +
                       dii[6] * dii[1] * dii[5] - dii[0] * dii[7] * dii[5] -
                       dii[6] * dii[4] * dii[2] - dii[3] * dii[1] * dii[8];
    if (fabs(det) > 1.0e-30) {
      inv3_inplace(dii);
    } else {
      nsing++;
      if (nsing > 10) {
        for (int k = c->rowptr[i]; k < c->rowptr[i + 1]; k++) mark[c->colind[k]] = -1;
        return nsing;
      }
    }
    for (int ij = c->diaind[i]; ij < c->rowptr[i + 1]; ij++) {
      double* u = vc + (size_t)ij * 9;
      double t[9];
      memcpy(t, u, sizeof(t));
      for (int r = 0; r < 3; r++) {
        u[r * 3 + 0] = dii[r * 3 + 0] * t[0] + dii[r * 3 + 1] * t[3] + dii[r * 3 + 2] * t[6];
        u[r * 3 + 1] = dii[r * 3 + 0] * t[1] + dii[r * 3 + 1] * t[4] + dii[r * 3 + 2] * t[7];
        u[r * 3 + 2] = dii[r * 3 + 0] * t[2] + dii[r * 3 + 1] * t[5] + dii[r * 3 + 2] * t[8];
      }
    }
    for (int k = c->rowptr[i]; k < c->rowptr[i + 1]; k++) mark[c->colind[k]] = -1;
  }
  return nsing;
}

/* forward then backward substitution, in place (ref: ilu_sparse.cpp:185-215, 295-324) */
static void ilu_apply(const ctx_t* c, const double* vd, const double* vc, double* v) {
  for (int i = 0; i < c->np; i++) {
    double t0 = v[i * 3 + 0], t1 = v[i * 3 + 1], t2 = v[i * 3 + 2];
    for (int k = c->rowptr[i]; k < c->diaind[i]; k++) {
      const int j = c->colind[k];
      const double* a = vc + (size_t)k * 9;
      const double y0 = v[j * 3 + 0], y1 = v[j * 3 + 1], y2 = v[j * 3 + 2];
      t0 -= a[0] * y0 + a[1] * y1 + a[2] * y2;
      t1 -= a[3] * y0 + a[4] * y1 + a[5] * y2;
      t2 -= a[6] * y0 + a[7] * y1 + a[8] * y2;
    }
    const double* d = vd + i * 9;
    v[i * 3 + 0] = d[0] * t0 + d[1] * t1 + d[2] * t2;
    v[i * 3 + 1] = d[3] * t0 + d[4] * t1 + d[5] * t2;
    v[i * 3 + 2] = d[6] * t0 + d[7] * t1 + d[8] * t2;
  }
  for (int i = c->np - 1; i >= 0; i--) {
    double t0 = v[i * 3 + 0], t1 = v[i * 3 + 1], t2 = v[i * 3 + 2];
    for (int k = c->diaind[i]; k < c->rowptr[i + 1]; k++) {
      const int j = c->colind[k];
      const double* a = vc + (size_t)k * 9;
      const double y0 = v[j * 3 + 0], y1 = v[j * 3 + 1], y2 = v[j * 3 + 2];
      t0 -= a[0] * y0 + a[1] * y1 + a[2] * y2;
      t1 -= a[3] * y0 + a[4] * y1 + a[5] * y2;
      t2 -= a[6] * y0 + a[7] * y1 + a[8] * y2;
    }
    v[i * 3 + 0] = t0;
    v[i * 3 + 1] = t1;
    v[i * 3 + 2] = t2;
  }
}

/* ILU-preconditioned CG from x0 = 0; r (in: b) is consumed
 * (ref: ilu_sparse.cpp:628-709) */
static void pcg(ctx_t* c, const double* A_d, const double* A_c, const double* F_d,
                const double* F_c, double* r, double* x, double tol, int maxit, double* conv,
                int* iters) {
  const int n = c->np * 3;
  for (int i = 0; i < n; i++) x[i] = 0;
  const double rr0 = dot_(c, r, r, n);
  if (rr0 < 1.0e-30) {
    *conv = 0.0;
    *iters = 0;
    return;
  }
  const double inv_rr0 = 1.0 / rr0;
  double* z = (double*)xmalloc(sizeof(double) * n); /* preconditioned residual; also holds A p */
  double* p = (double*)xmalloc(sizeof(double) * n);
  memcpy(z, r, sizeof(double) * n);
  ilu_apply(c, F_d, F_c, z);
  memcpy(p, z, sizeof(double) * n);
  double rz = dot_(c, r, z, n);
  *iters = maxit;
  for (int it = 0; it < maxit; it++) {
    matvec(c, A_d, A_c, 1.0, p, 0.0, z);
    const double pAp = dot_(c, p, z, n);
    const double alpha = rz / pAp;
    for (int i = 0; i < n; i++) r[i] += -alpha * z[i];
    for (int i = 0; i < n; i++) x[i] += +alpha * p[i];
    const double rr = dot_(c, r, r, n);
    if (rr * inv_rr0 < tol * tol) {
      *conv = sqrt(rr * inv_rr0);
      *iters = it;
      free(z);
      free(p);
      return;
    }
    memcpy(z, r, sizeof(double) * n);
    ilu_apply(c, F_d, F_c, z);
    const double rz1 = dot_(c, r, z, n);
    const double beta = rz1 / rz;
    rz = rz1;
    for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
  }
  *conv = sqrt(dot_(c, r, r, n) * inv_rr0);
  free(z);
  free(p);
}

/* ------------------------------------------------------------------------ */
/* assembly (ref: solve_internal_sparse.h:19-183, 277-313) */

static void gather_elem(const ctx_t* c, const double* xyz, const int* ip, int n, double* C_,
                        double* c_) {
  for (int a = 0; a < n; a++)
    for (int k = 0; k < 3; k++) {
      C_[a * 3 + k] = c->xyz0[ip[a] * 3 + k];
      c_[a * 3 + k] = xyz[ip[a] * 3 + k];
    }
}

/* dW (and, when dia != NULL, the Hessian) of membrane + bending */
static void add_cloth(ctx_t* c, const double* xyz, double* dW, double* dia, double* crs) {
  double C_[12], c_[12], de[12], dde[144], W;
  for (int t = 0; t < c->nt; t++) {
    const int* ip = c->tri + t * 3;
    gather_elem(c, xyz, ip, 3, C_, c_);
    orc_cst(C_, c_, c->lam, c->mu, &W, de, dde);
    for (int a = 0; a < 3; a++)
      for (int k = 0; k < 3; k++) dW[ip[a] * 3 + k] += de[a * 3 + k];
    if (dia) merge(c, dia, crs, 3, ip, dde);
  }
  for (int q = 0; q < c->nq; q++) {
    const int* ip = c->quad + q * 4;
    gather_elem(c, xyz, ip, 4, C_, c_);
    orc_bend(C_, c_, c->bend, de, dde);
    for (int a = 0; a < 4; a++)
      for (int k = 0; k < 3; k++) dW[ip[a] * 3 + k] += de[a * 3 + k];
    if (dia) merge(c, dia, crs, 4, ip, dde);
  }
}

static void add_contact(ctx_t* c, const double* xyz, double* dW, double* dia, double* crs) {
  double W, de[3], dde[9];
  for (int i = 0; i < c->np; i++) {
    orc_contact_elem(xyz + i * 3, c->kc, c->clear, c->obst, &W, de, dde);
    for (int k = 0; k < 3; k++) dW[i * 3 + k] += de[k];
    merge(c, dia, crs, 1, &i, dde);
  }
}

static void add_gravity(ctx_t* c, double* dW) {
  for (int i = 0; i < c->np; i++)
    for (int k = 0; k < 3; k++) dW[i * 3 + k] -= c->mass * c->grav[k];
}

/* backward-Euler system A dx = b after boundary conditions */
static void assemble_system(ctx_t* c, const double* xyz, const double* uvw, double* b) {
  const int n = c->np * 3;
  memset(c->dia, 0, sizeof(double) * 9 * c->np);
  memset(c->crs, 0, sizeof(double) * 9 * (size_t)c->nnz);
  for (int i = 0; i < n; i++) b[i] = 0;
  add_cloth(c, xyz, b, c->dia, c->crs);
  add_contact(c, xyz, b, c->dia, c->crs);
  add_gravity(c, b);
  for (int i = 0; i < n; i++) b[i] = -b[i] + c->mass * uvw[i] / c->dt;
  for (int i = 0; i < c->np; i++)
    for (int k = 0; k < 3; k++) c->dia[i * 9 + k * 3 + k] += c->mass / (c->dt * c->dt);
  set_bc(c, c->dia, c->crs);
  for (int i = 0; i < c->np; i++) {
    if (c->bc[i] == 0) continue;
    b[i * 3 + 0] = b[i * 3 + 1] = b[i * 3 + 2] = 0;
  }
}

void orc_assemble(void* h, const double* xyz, const double* uvw, double* dia, double* crs,
                  double* b) {
  ctx_t* c = (ctx_t*)h;
  assemble_system(c, xyz, uvw, b);
  memcpy(dia, c->dia, sizeof(double) * 9 * c->np);
  memcpy(crs, c->crs, sizeof(double) * 9 * (size_t)c->nnz);
}

void orc_ilu(void* h, const double* dia, const double* crs, double* fdia, double* fcrs) {
  ctx_t* c = (ctx_t*)h;
  memcpy(fdia, dia, sizeof(double) * 9 * c->np);
  memcpy(fcrs, crs, sizeof(double) * 9 * (size_t)c->nnz);
  ilu_factor(c, fdia, fcrs);
}

void orc_pcg(void* h, const double* dia, const double* crs, const double* fdia, const double* fcrs,
             double* r, double* x, double tol, int maxit, double* conv, int* iters) {
  pcg((ctx_t*)h, dia, crs, fdia, fcrs, r, x, tol, maxit, conv, iters);
}

void orc_ilu_solve(void* h, const double* fdia, const double* fcrs, double* vec) {
  ilu_apply((ctx_t*)h, fdia, fcrs, vec);
}

void orc_matvec(void* h, const double* dia, const double* crs, double alpha, const double* x,
                double beta, double* y) {
  matvec((ctx_t*)h, dia, crs, alpha, x, beta, y);
}

/* StepTime_InternalDynamicsILU with tolerance/cap as arguments
 * (ref: solve_internal_sparse.h:255-326) */
static void internal_step(ctx_t* c, double* xyz, double* uvw, double tol, int maxit) {
  const int n = c->np * 3;
  double* b = (double*)xmalloc(sizeof(double) * n);
  double* x = (double*)xmalloc(sizeof(double) * n);
  assemble_system(c, xyz, uvw, b);
  memcpy(c->fdia, c->dia, sizeof(double) * 9 * c->np);
  memcpy(c->fcrs, c->crs, sizeof(double) * 9 * (size_t)c->nnz);
  ilu_factor(c, c->fdia, c->fcrs);
  pcg(c, c->dia, c->crs, c->fdia, c->fcrs, b, x, tol, maxit, &c->last_conv, &c->last_iter);
  for (int i = 0; i < n; i++) xyz[i] += x[i];
  for (int i = 0; i < n; i++) uvw[i] = x[i] / c->dt;
  free(b);
  free(x);
}

void orc_step_internal_tol(void* h, double* xyz, double* uvw, double tol, int maxit, double* conv,
                           int* iters) {
  ctx_t* c = (ctx_t*)h;
  internal_step(c, xyz, uvw, tol, maxit);
  *conv = c->last_conv;
  *iters = c->last_iter;
}

/* ------------------------------------------------------------------------ */
/* BVH refit (ref: bvh_aabb.cpp:23-106) */

static void refit_prx(ctx_t* c, int n, const double* xyz, double delta) {
  const int c0 = c->node[n * 3 + 1], c1 = c->node[n * 3 + 2];
  box* b = &c->bb[n];
  if (c1 == -1) {
    const int* t = c->tri + c0 * 3;
    b->valid = 0;
    for (int a = 0; a < 3; a++) box_addv(b, vload(xyz, t[a]), delta * 0.5);
    return;
  }
  refit_prx(c, c0, xyz, delta);
  refit_prx(c, c1, xyz, delta);
  *b = c->bb[c0];
  box_union(b, &c->bb[c1]);
}

static void refit_ccd(ctx_t* c, int n, const double* xyz, const double* uvw, double dt) {
  const double eps = 1.0e-10;
  const int c0 = c->node[n * 3 + 1], c1 = c->node[n * 3 + 2];
  box* b = &c->bb[n];
  if (c1 == -1) {
    const int* t = c->tri + c0 * 3;
    b->valid = 0;
    for (int a = 0; a < 3; a++) box_addv(b, vload(xyz, t[a]), eps);
    for (int a = 0; a < 3; a++) box_addv(b, vload_adv(xyz, uvw, dt, t[a]), eps);
    return;
  }
  refit_ccd(c, c0, xyz, uvw, dt);
  refit_ccd(c, c1, xyz, uvw, dt);
  *b = c->bb[c0];
  box_union(b, &c->bb[c1]);
}

static void export_bb(const ctx_t* c, double* bb) {
  for (int i = 0; i < c->nnode; i++) {
    bb[i * 6 + 0] = c->bb[i].x0;
    bb[i * 6 + 1] = c->bb[i].x1;
    bb[i * 6 + 2] = c->bb[i].y0;
    bb[i * 6 + 3] = c->bb[i].y1;
    bb[i * 6 + 4] = c->bb[i].z0;
    bb[i * 6 + 5] = c->bb[i].z1;
  }
}

void orc_refit_prx(void* h, const double* xyz, double delta, double* bb) {
  ctx_t* c = (ctx_t*)h;
  refit_prx(c, c->root, xyz, delta);
  export_bb(c, bb);
}
void orc_refit_ccd(void* h, const double* xyz, const double* uvw, double dt, double* bb) {
  ctx_t* c = (ctx_t*)h;
  refit_ccd(c, c->root, xyz, uvw, dt);
  export_bb(c, bb);
}

/* ------------------------------------------------------------------------ */
/* narrow phase (ref: bvh_aabb.cpp:321-438, 563-793) */

/* vertex p3 against the plane of p0 p1 p2: unclamped barycentrics (ref: :321-340) */
static double dist_fv(v3 p0, v3 p1, v3 p2, v3 p3, double* w0, double* w1) {
  const v3 a = vsub(p0, p2), b = vsub(p1, p2);
  const double t0 = vdot(a, a), t1 = vdot(b, b), t2 = vdot(a, b);
  const double t3 = vdot(a, vsub(p3, p2)), t4 = vdot(b, vsub(p3, p2));
  const double det = t0 * t1 - t2 * t2;
  const double inv = 1.0 / det;
  *w0 = (+t1 * t3 - t2 * t4) * inv;
  *w1 = (-t2 * t3 + t0 * t4) * inv;
  const double w2 = 1 - *w0 - *w1;
  const v3 pw = vadd(vadd(vscl(*w0, p0), vscl(*w1, p1)), vscl(w2, p2));
  return vlen(vsub(pw, p3));
}

/* closest points of two lines, with a parallel-edge branch (ref: :343-384) */
static double dist_ee(v3 p0, v3 p1, v3 q0, v3 q1, double* rp, double* rq) {
  const v3 vp = vsub(p1, p0), vq = vsub(q1, q0);
  if (vlen(vcross(vp, vq)) < 1.0e-10) {
    const v3 pq0 = vsub(p0, q0);
    const v3 nvp = vnormalized(vp);
    const v3 vert = vsub(pq0, vscl(vdot(pq0, nvp), nvp));
    const double dist = vlen(vert);
    const double lp0 = vdot(p0, nvp), lp1 = vdot(p1, nvp);
    const double lq0 = vdot(q0, nvp), lq1 = vdot(q1, nvp);
    const double pmin = (lp0 < lp1) ? lp0 : lp1, pmax = (lp0 > lp1) ? lp0 : lp1;
    const double qmin = (lq0 < lq1) ? lq0 : lq1, qmax = (lq0 > lq1) ? lq0 : lq1;
    double lm;
    if (pmax < qmin)
      lm = (pmax + qmin) * 0.5;
    else if (qmax < pmin)
      lm = (qmax + pmin) * 0.5;
    else if (pmax < qmax)
      lm = (pmax + qmin) * 0.5;
    else
      lm = (qmax + pmin) * 0.5;
    *rp = (lm - lp0) / (lp1 - lp0);
    *rq = (lm - lq0) / (lq1 - lq0);
    return dist;
  }
  const double t0 = vdot(vp, vp), t1 = vdot(vq, vq), t2 = vdot(vp, vq);
  const double t3 = vdot(vp, vsub(q0, p0)), t4 = vdot(vq, vsub(q0, p0));
  const double det = t0 * t1 - t2 * t2;
  const double inv = 1.0 / det;
  *rp = (+t1 * t3 - t2 * t4) * inv;
  *rq = (+t2 * t3 - t0 * t4) * inv;
  const v3 pc = vadd(p0, vscl(*rp, vp));
  const v3 qc = vadd(q0, vscl(*rq, vq));
  return vlen(vsub(pc, qc));
}

double orc_dist_fv(const double* p, double* w) {
  return dist_fv(vload(p, 0), vload(p, 1), vload(p, 2), vload(p, 3), &w[0], &w[1]);
}
double orc_dist_ee(const double* p, double* w) {
  return dist_ee(vload(p, 0), vload(p, 1), vload(p, 2), vload(p, 3), &w[0], &w[1]);
}

static int fv_proximity(int i0, int i1, int i2, int i3, v3 p0, v3 p1, v3 p2, v3 p3, const box* b,
                        double delta) {
  if (i3 == i0 || i3 == i1 || i3 == i2) return 0;
  if (!box_inside(b, p3)) return 0;
  const double h = fabs(height4(p0, p1, p2, p3));
  if (h > delta) return 0;
  double w0, w1;
  const double dist = dist_fv(p0, p1, p2, p3, &w0, &w1);
  const double w2 = 1 - w0 - w1;
  if (dist > delta) return 0;
  /* the reference computes a margin and then sets it to 0 (ref: :406-407) */
  if (w0 < -0.0 || w0 > 1 + 0.0) return 0;
  if (w1 < -0.0 || w1 > 1 + 0.0) return 0;
  if (w2 < -0.0 || w2 > 1 + 0.0) return 0;
  return 1;
}

static int ee_proximity(int i0, int i1, int j0, int j1, v3 p0, v3 p1, v3 q0, v3 q1, double d) {
  if (i0 == j0 || i0 == j1 || i1 == j0 || i1 == j1) return 0;
  if (q0.x + d < p0.x && q0.x + d < p1.x && q1.x + d < p0.x && q1.x + d < p1.x) return 0;
  if (q0.x - d > p0.x && q0.x - d > p1.x && q1.x - d > p0.x && q1.x - d > p1.x) return 0;
  if (q0.y + d < p0.y && q0.y + d < p1.y && q1.y + d < p0.y && q1.y + d < p1.y) return 0;
  if (q0.y - d > p0.y && q0.y - d > p1.y && q1.y - d > p0.y && q1.y - d > p1.y) return 0;
  if (q0.z + d < p0.z && q0.z + d < p1.z && q1.z + d < p0.z && q1.z + d < p1.z) return 0;
  if (q0.z - d > p0.z && q0.z - d > p1.z && q1.z - d > p0.z && q1.z - d > p1.z) return 0;
  double rp, rq;
  const double dist = dist_ee(p0, p1, q0, q1, &rp, &rq);
  if (dist > d) return 0;
  if (rp < 0) return 0;
  if (rp > 1) return 0;
  if (rq < 0) return 0;
  if (rq > 1) return 0;
  const v3 pm = vadd(vscl(1 - rp, p0), vscl(rp, p1));
  const v3 qm = vadd(vscl(1 - rq, q0), vscl(rq, q1));
  if (vlen(vsub(pm, qm)) > d) return 0;
  return 1;
}

static inline double cubic(double r, double k0, double k1, double k2, double k3) {
  return k0 + k1 * r + k2 * r * r + k3 * r * r * r; /* ref: :563-568 */
}

/* bisection with a counter of 15 (14 halvings).  The sign reference stays
 * the value at the initial left end throughout (ref: :571-595). */
static double cubic_root(double r0, double r1, double v0, double k0, double k1, double k2,
                         double k3) {
  for (int cnt = 15;;) {
    cnt--;
    if (cnt <= 0) break;
    const double r2 = 0.5 * (r0 + r1);
    const double v2 = cubic(r2, k0, k1, k2, k3);
    if (v0 * v2 < 0)
      r1 = r2;
    else
      r0 = r2;
  }
  return 0.5 * (r0 + r1);
}

/* time in [0,1] at which the four moving points become coplanar, -1 if none
 * found (ref: :598-662) */
static double coplanar_t(v3 s0, v3 s1, v3 s2, v3 s3, v3 e0, v3 e1, v3 e2, v3 e3) {
  const v3 x1 = vsub(s1, s0), x2 = vsub(s2, s0), x3 = vsub(s3, s0);
  const v3 v1 = vsub(vsub(e1, e0), x1), v2 = vsub(vsub(e2, e0), x2), v3_ = vsub(vsub(e3, e0), x3);
  const double k0 = stp(x3, x1, x2);
  const double k1 = stp(v3_, x1, x2) + stp(x3, v1, x2) + stp(x3, x1, v2);
  const double k2 = stp(v3_, v1, x2) + stp(v3_, x1, v2) + stp(x3, v1, v2);
  const double k3 = stp(v3_, v1, v2);
  const double r0 = -0.0, r1 = +1.0;
  const double f0 = cubic(r0, k0, k1, k2, k3);
  const double f1 = cubic(r1, k0, k1, k2, k3);
  const double det = k2 * k2 - 3 * k1 * k3;
  if (fabs(k3) < 1.0e-10 && fabs(k2) > 1.0e-10) { /* quadratic: split at the extremum */
    const double r2 = -k1 / (2 * k2);
    const double f2 = cubic(r2, k0, k1, k2, k3);
    if (r2 > 0 && r2 < 1) {
      if (f0 * f2 < 0) return cubic_root(r0, r2, f0, k0, k1, k2, k3);
      if (f2 * f1 < 0) return cubic_root(r2, r1, f2, k0, k1, k2, k3);
    }
  }
  if (det > 0 && fabs(k3) > 1.0e-10) { /* cubic with two extrema */
    const double r3 = (-k2 - sqrt(det)) / (3 * k3);
    const double f3 = cubic(r3, k0, k1, k2, k3);
    if (r3 > 0 && r3 < 1) {
      if (f0 * f3 < 0) return cubic_root(r0, r3, f0, k0, k1, k2, k3);
      if (f3 * f1 < 0) return cubic_root(r3, r1, f3, k0, k1, k2, k3);
    }
    const double r4 = (-k2 + sqrt(det)) / (3 * k3);
    const double f4 = cubic(r4, k0, k1, k2, k3);
    if (r3 > 0 && r3 < 1 && r4 > 0 && r4 < 1) {
      if (f3 * f4 < 0) return cubic_root(r3, r4, f3, k0, k1, k2, k3);
    }
    if (r4 > 0 && r4 < 1) {
      if (f0 * f4 < 0) return cubic_root(r0, r4, f0, k0, k1, k2, k3);
      if (f4 * f1 < 0) return cubic_root(r4, r1, f4, k0, k1, k2, k3);
    }
  }
  if (f0 * f1 > 0) return -1;
  return cubic_root(r0, r1, f0, k0, k1, k2, k3);
}

double orc_coplanar(const double* s, const double* e) {
  return coplanar_t(vload(s, 0), vload(s, 1), vload(s, 2), vload(s, 3), vload(e, 0), vload(e, 1),
                    vload(e, 2), vload(e, 3));
}

static int fv_ccd(int i0, int i1, int i2, int i3, v3 p0, v3 p1, v3 p2, v3 p3, v3 q0, v3 q1, v3 q2,
                  v3 q3, const box* b) {
  const double eps = 1.0e-10;
  if (i3 == i0 || i3 == i1 || i3 == i2) return 0;
  box bp;
  bp.valid = 0;
  box_addv(&bp, p3, eps);
  box_addv(&bp, q3, eps);
  if (!box_overlap(b, &bp)) return 0;
  {
    const v3 n = vcross(vsub(p1, p0), vsub(p2, p0));
    const double t0 = vdot(vsub(p0, p3), n);
    const double t1 = vdot(vsub(q0, q3), n);
    const double t2 = vdot(vsub(q1, q3), n);
    const double t3 = vdot(vsub(q2, q3), n);
    if (t0 * t1 > 0 && t0 * t2 > 0 && t0 * t3 > 0) return 0;
  }
  double r0, r1;
  const double dist = dist_fv(p0, p1, p2, p3, &r0, &r1);
  {
    const double vn0 = vlen(vsub(p0, q0)), vn1 = vlen(vsub(p1, q1));
    const double vn2 = vlen(vsub(p2, q2)), vn3 = vlen(vsub(p3, q3));
    double vnt = (vn0 > vn1) ? vn0 : vn1;
    vnt = (vn2 > vnt) ? vn2 : vnt;
    const double reach = vnt + vn3;
    const double r2 = 1 - r0 - r1;
    if (dist > reach) return 0;
    if (r0 < 0 || r0 > 1 || r1 < 0 || r1 > 1 || r2 < 0 || r2 > 1) {
      const double d01 = vlen(vsub(seg_closest(p3, p0, p1), p3));
      const double d12 = vlen(vsub(seg_closest(p3, p1, p2), p3));
      const double d20 = vlen(vsub(seg_closest(p3, p2, p0), p3));
      if (d01 > reach && d12 > reach && d20 > reach) return 0;
    }
  }
  const double t = coplanar_t(p0, p1, p2, p3, q0, q1, q2, q3);
  if (t < 0 || t > 1) return 0;
  const v3 m0 = vadd(vscl(1 - t, p0), vscl(t, q0));
  const v3 m1 = vadd(vscl(1 - t, p1), vscl(t, q1));
  const v3 m2 = vadd(vscl(1 - t, p2), vscl(t, q2));
  const v3 m3 = vadd(vscl(1 - t, p3), vscl(t, q3));
  double w0, w1;
  dist_fv(m0, m1, m2, m3, &w0, &w1);
  const double w2 = 1 - w0 - w1;
  if (w0 < 0 || w0 > 1) return 0;
  if (w1 < 0 || w1 > 1) return 0;
  if (w2 < 0 || w2 > 1) return 0;
  return 1;
}

static int ee_ccd(int i0, int i1, int j0, int j1, v3 p0s, v3 p1s, v3 q0s, v3 q1s, v3 p0e, v3 p1e,
                  v3 q0e, v3 q1e) {
  const double eps = 1.0e-10;
  if (i0 == j0 || i0 == j1 || i1 == j0 || i1 == j1) return 0;
  box bq, bp;
  bq.valid = 0;
  bp.valid = 0;
  box_addv(&bq, q0s, eps);
  box_addv(&bq, q1s, eps);
  box_addv(&bq, q0e, eps);
  box_addv(&bq, q1e, eps);
  box_addv(&bp, p0s, eps);
  box_addv(&bp, p1s, eps);
  box_addv(&bp, p0e, eps);
  box_addv(&bp, p1e, eps);
  if (!box_overlap(&bp, &bq)) return 0;
  const double t = coplanar_t(p0s, p1s, q0s, q1s, p0e, p1e, q0e, q1e);
  if (t < 0 || t > 1) return 0;
  const v3 a0 = vadd(vscl(1 - t, p0s), vscl(t, p0e));
  const v3 a1 = vadd(vscl(1 - t, p1s), vscl(t, p1e));
  const v3 b0 = vadd(vscl(1 - t, q0s), vscl(t, q0e));
  const v3 b1 = vadd(vscl(1 - t, q1s), vscl(t, q1e));
  double w0, w1;
  const double dist = dist_ee(a0, a1, b0, b1, &w0, &w1);
  if (w0 < 0 || w0 > 1) return 0;
  if (w1 < 0 || w1 > 1) return 0;
  if (dist > 1.0e-2) return 0; /* hard-coded in the reference (ref: :791) */
  return 1;
}

/* ------------------------------------------------------------------------ */
/* contact sets: canonical 4-index keys, first insertion wins on equal keys,
 * listed in key order (std::set<CContactElement>, ref: bvh_aabb.h:65-106) */

typedef struct {
  int k[4];
  int is_fv;
  long seq;
} ce_t;

typedef struct {
  ce_t* v;
  long n, cap;
} ceset_t;

static void ce_push(ceset_t* s, int is_fv, int j0, int j1, int j2, int j3) {
  ce_t e;
  e.is_fv = is_fv;
  if (is_fv) { /* face sorted ascending, vertex last */
    int a = j0, b = j1, cc = j2, t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > cc) { t = b; b = cc; cc = t; }
    if (a > b) { t = a; a = b; b = t; }
    e.k[0] = a;
    e.k[1] = b;
    e.k[2] = cc;
    e.k[3] = j3;
  } else { /* edge holding the smallest index first, smaller end first;
              the other edge ordered ascending */
    const int m01 = j0 < j1 ? j0 : j1, m23 = j2 < j3 ? j2 : j3;
    if (m01 < m23) {
      e.k[0] = j0 < j1 ? j0 : j1;
      e.k[1] = j0 < j1 ? j1 : j0;
      e.k[2] = j2 < j3 ? j2 : j3;
      e.k[3] = j2 < j3 ? j3 : j2;
    } else {
      e.k[0] = j2 < j3 ? j2 : j3;
      e.k[1] = j2 < j3 ? j3 : j2;
      e.k[2] = j0 < j1 ? j0 : j1;
      e.k[3] = j0 < j1 ? j1 : j0;
    }
  }
  if (s->n == s->cap) {
    s->cap = s->cap ? s->cap * 2 : 1024;
    s->v = (ce_t*)realloc(s->v, sizeof(ce_t) * s->cap);
    if (!s->v) abort();
  }
  e.seq = s->n;
  s->v[s->n++] = e;
}

static int ce_cmp(const void* a, const void* b) {
  const ce_t* x = (const ce_t*)a;
  const ce_t* y = (const ce_t*)b;
  for (int i = 0; i < 4; i++)
    if (x->k[i] != y->k[i]) return x->k[i] < y->k[i] ? -1 : 1;
  return (x->seq > y->seq) - (x->seq < y->seq);
}

/* sort by key then insertion order and keep the first of each key */
static void ce_finish(ceset_t* s) {
  if (s->n == 0) return;
  qsort(s->v, s->n, sizeof(ce_t), ce_cmp);
  long m = 0;
  for (long i = 0; i < s->n; i++) {
    if (m > 0 && memcmp(s->v[m - 1].k, s->v[i].k, sizeof(int) * 4) == 0) continue;
    s->v[m++] = s->v[i];
  }
  s->n = m;
}

/* pair traversal (ref: bvh_aabb.cpp:441-537 proximity, 797-902 CCD) */
typedef struct {
  ctx_t* c;
  const double* xyz;
  const double* uvw;
  double dt, delta;
  int ccd;
  ceset_t* out;
} trav_t;

static void leaf_prox(trav_t* T, int itri, int jtri) {
  const ctx_t* c = T->c;
  const double* X = T->xyz;
  const double d = T->delta;
  const int in0 = c->tri[itri * 3 + 0], in1 = c->tri[itri * 3 + 1], in2 = c->tri[itri * 3 + 2];
  const int jn0 = c->tri[jtri * 3 + 0], jn1 = c->tri[jtri * 3 + 1], jn2 = c->tri[jtri * 3 + 2];
  const v3 p0 = vload(X, in0), p1 = vload(X, in1), p2 = vload(X, in2);
  const v3 q0 = vload(X, jn0), q1 = vload(X, jn1), q2 = vload(X, jn2);
  /* the face test uses the box of the node whose id equals the triangle id
   * (ref: :490-507) */
  const box* bi = &c->bb[itri];
  const box* bj = &c->bb[jtri];
  const int iv[3] = {in0, in1, in2}, jv[3] = {jn0, jn1, jn2};
  const v3 P[3] = {p0, p1, p2}, Q[3] = {q0, q1, q2};
  for (int a = 0; a < 3; a++)
    if (fv_proximity(in0, in1, in2, jv[a], p0, p1, p2, Q[a], bi, d))
      ce_push(T->out, 1, in0, in1, in2, jv[a]);
  for (int a = 0; a < 3; a++)
    if (fv_proximity(jn0, jn1, jn2, iv[a], q0, q1, q2, P[a], bj, d))
      ce_push(T->out, 1, jn0, jn1, jn2, iv[a]);
  /* edges in the order (0,1), (1,2), (2,0) on both sides */
  for (int a = 0; a < 3; a++) {
    const int a1 = (a + 1) % 3;
    for (int b = 0; b < 3; b++) {
      const int b1 = (b + 1) % 3;
      if (ee_proximity(iv[a], iv[a1], jv[b], jv[b1], P[a], P[a1], Q[b], Q[b1], d))
        ce_push(T->out, 0, iv[a], iv[a1], jv[b], jv[b1]);
    }
  }
}

static void leaf_ccd(trav_t* T, int ni, int nj, int itri, int jtri) {
  const ctx_t* c = T->c;
  const double* X = T->xyz;
  const double* U = T->uvw;
  const double dt = T->dt;
  const int iv[3] = {c->tri[itri * 3 + 0], c->tri[itri * 3 + 1], c->tri[itri * 3 + 2]};
  const int jv[3] = {c->tri[jtri * 3 + 0], c->tri[jtri * 3 + 1], c->tri[jtri * 3 + 2]};
  v3 Ps[3], Qs[3], Pe[3], Qe[3];
  for (int a = 0; a < 3; a++) {
    Ps[a] = vload(X, iv[a]);
    Qs[a] = vload(X, jv[a]);
    Pe[a] = vload_adv(X, U, dt, iv[a]);
    Qe[a] = vload_adv(X, U, dt, jv[a]);
  }
  const box* bi = &c->bb[ni];
  const box* bj = &c->bb[nj];
  for (int a = 0; a < 3; a++)
    if (fv_ccd(iv[0], iv[1], iv[2], jv[a], Ps[0], Ps[1], Ps[2], Qs[a], Pe[0], Pe[1], Pe[2], Qe[a],
               bi))
      ce_push(T->out, 1, iv[0], iv[1], iv[2], jv[a]);
  for (int a = 0; a < 3; a++)
    if (fv_ccd(jv[0], jv[1], jv[2], iv[a], Qs[0], Qs[1], Qs[2], Ps[a], Qe[0], Qe[1], Qe[2], Pe[a],
               bj))
      ce_push(T->out, 1, jv[0], jv[1], jv[2], iv[a]);
  for (int a = 0; a < 3; a++) {
    const int a1 = (a + 1) % 3;
    for (int b = 0; b < 3; b++) {
      const int b1 = (b + 1) % 3;
      if (ee_ccd(iv[a], iv[a1], jv[b], jv[b1], Ps[a], Ps[a1], Qs[b], Qs[b1], Pe[a], Pe[a1], Qe[b],
                 Qe[b1]))
        ce_push(T->out, 0, iv[a], iv[a1], jv[b], jv[b1]);
    }
  }
}

static void trav_pair(trav_t* T, int A, int B) {
  const ctx_t* c = T->c;
  if (!box_overlap(&c->bb[A], &c->bb[B])) return;
  const int a0 = c->node[A * 3 + 1], a1 = c->node[A * 3 + 2];
  const int b0 = c->node[B * 3 + 1], b1 = c->node[B * 3 + 2];
  const int la = (a1 == -1), lb = (b1 == -1);
  if (!la && !lb) {
    trav_pair(T, a0, b0);
    trav_pair(T, a1, b0);
    trav_pair(T, a0, b1);
    trav_pair(T, a1, b1);
  } else if (!la && lb) {
    trav_pair(T, a0, B);
    trav_pair(T, a1, B);
  } else if (la && !lb) {
    trav_pair(T, A, b0);
    trav_pair(T, A, b1);
  } else {
    if (T->ccd)
      leaf_ccd(T, A, B, a0, b0);
    else
      leaf_prox(T, a0, b0);
  }
}

static void trav_self(trav_t* T, int N) {
  const int c0 = T->c->node[N * 3 + 1], c1 = T->c->node[N * 3 + 2];
  if (c1 == -1) return;
  trav_pair(T, c0, c1);
  trav_self(T, c0);
  trav_self(T, c1);
}

static void query_prox(ctx_t* c, const double* xyz, double delta, ceset_t* out) {
  refit_prx(c, c->root, xyz, delta);
  trav_t T = {c, xyz, NULL, 0.0, delta, 0, out};
  trav_self(&T, c->root);
  ce_finish(out);
}

static void query_ccd(ctx_t* c, const double* xyz, const double* uvw, double dt, double delta,
                      ceset_t* out) {
  refit_ccd(c, c->root, xyz, uvw, dt);
  trav_t T = {c, xyz, uvw, dt, delta, 1, out};
  trav_self(&T, c->root);
  ce_finish(out);
}

static int export_set(const ceset_t* s, int* out, int cap) {
  for (long i = 0; i < s->n && i < cap; i++) {
    out[i * 5 + 0] = s->v[i].is_fv;
    for (int k = 0; k < 4; k++) out[i * 5 + 1 + k] = s->v[i].k[k];
  }
  return (int)s->n;
}

int orc_prox_set(void* h, const double* xyz, double delta, int* out, int cap) {
  ceset_t s = {NULL, 0, 0};
  query_prox((ctx_t*)h, xyz, delta, &s);
  int n = export_set(&s, out, cap);
  free(s.v);
  return n;
}

int orc_ccd_set(void* h, const double* xyz, const double* uvw, double dt, double delta, int* out,
                int cap) {
  ceset_t s = {NULL, 0, 0};
  query_ccd((ctx_t*)h, xyz, uvw, dt, delta, &s);
  int n = export_set(&s, out, cap);
  free(s.v);
  return n;
}

/* ------------------------------------------------------------------------ */
/* impulses (ref: self_collision_cloth.cpp:20-212), Gauss-Seidel in key order */

static inline void vel_add(double* u, int i, v3 d) {
  u[i * 3 + 0] += d.x;
  u[i * 3 + 1] += d.y;
  u[i * 3 + 2] += d.z;
}

static void impulse_prox(double* U, double delta, double k, double dt, double m, const double* X,
                         const ceset_t* s) {
  for (long e = 0; e < s->n; e++) {
    const ce_t* ce = &s->v[e];
    const int i0 = ce->k[0], i1 = ce->k[1], i2 = ce->k[2], i3 = ce->k[3];
    const v3 p0 = vload(X, i0), p1 = vload(X, i1), p2 = vload(X, i2), p3 = vload(X, i3);
    const v3 v0 = vload(U, i0), v1 = vload(U, i1), v2 = vload(U, i2), v3v = vload(U, i3);
    if (ce->is_fv) {
      double w0, w1;
      {
        const double dist = dist_fv(p0, p1, p2, p3, &w0, &w1);
        if (w0 < 0 || w0 > 1) continue;
        if (w1 < 0 || w1 > 1) continue;
        if (dist > delta) continue;
      }
      const double w2 = 1.0 - w0 - w1;
      const v3 pc = vadd(vadd(vscl(w0, p0), vscl(w1, p1)), vscl(w2, p2));
      const v3 nrm = vnormalized(vsub(p3, pc));
      const double depth = delta - vdot(vsub(p3, pc), nrm);
      const double rel =
          vdot(vsub(vsub(vsub(v3v, vscl(w0, v0)), vscl(w1, v1)), vscl(w2, v2)), nrm);
      if (rel > 0.1 * depth / dt) continue;
      const double el = dt * k * depth;
      const double ie = m * (0.1 * depth / dt - rel);
      const double imin = (el < ie) ? el : ie;
      double j = 2 * imin / (1 + w0 * w0 + w1 * w1 + w2 * w2);
      j /= m;
      j *= 0.25;
      vel_add(U, i0, V(-nrm.x * j * w0, -nrm.y * j * w0, -nrm.z * j * w0));
      vel_add(U, i1, V(-nrm.x * j * w1, -nrm.y * j * w1, -nrm.z * j * w1));
      vel_add(U, i2, V(-nrm.x * j * w2, -nrm.y * j * w2, -nrm.z * j * w2));
      vel_add(U, i3, V(+nrm.x * j, +nrm.y * j, +nrm.z * j));
    } else {
      double wa, wb;
      {
        const double dist = dist_ee(p0, p1, p2, p3, &wa, &wb);
        if (wa < 0 || wa > 1) continue;
        if (wb < 0 || wb > 1) continue;
        if (dist > delta) continue;
      }
      const v3 ca = vadd(vscl(1 - wa, p0), vscl(wa, p1));
      const v3 cb = vadd(vscl(1 - wb, p2), vscl(wb, p3));
      const v3 nrm = vnormalized(vsub(cb, ca));
      const double depth = delta - vlen(vsub(cb, ca));
      const double rel = vdot(vsub(vsub(vadd(vscl(1 - wb, v2), vscl(wb, v3v)), vscl(1 - wa, v0)),
                                   vscl(wa, v1)),
                              nrm);
      if (rel > 0.1 * depth / dt) continue;
      const double el = dt * k * depth;
      const double ie = m * (0.1 * depth / dt - rel);
      const double imin = (el < ie) ? el : ie;
      double j = 2 * imin / (wa * wa + (1 - wa) * (1 - wa) + wb * wb + (1 - wb) * (1 - wb));
      j /= m;
      j *= 0.25;
      vel_add(U, i0, V(-nrm.x * j * (1 - wa), -nrm.y * j * (1 - wa), -nrm.z * j * (1 - wa)));
      vel_add(U, i1, V(-nrm.x * j * wa, -nrm.y * j * wa, -nrm.z * j * wa));
      vel_add(U, i2, V(+nrm.x * j * (1 - wb), +nrm.y * j * (1 - wb), +nrm.z * j * (1 - wb)));
      vel_add(U, i3, V(+nrm.x * j * wb, +nrm.y * j * wb, +nrm.z * j * wb));
    }
  }
}

static void impulse_ccd(double* U, double delta, double dt, double m, const double* X,
                        const ceset_t* s) {
  for (long e = 0; e < s->n; e++) {
    const ce_t* ce = &s->v[e];
    const int i0 = ce->k[0], i1 = ce->k[1], i2 = ce->k[2], i3 = ce->k[3];
    const v3 p0 = vload(X, i0), p1 = vload(X, i1), p2 = vload(X, i2), p3 = vload(X, i3);
    const v3 v0 = vload(U, i0), v1 = vload(U, i1), v2 = vload(U, i2), v3v = vload(U, i3);
    /* the end positions are p + v, not p + dt v (ref: :141) */
    const double t = coplanar_t(p0, p1, p2, p3, vadd(p0, v0), vadd(p1, v1), vadd(p2, v2),
                                vadd(p3, v3v));
    if (t < 0 || t > 1) continue;
    if (ce->is_fv) {
      double w0, w1;
      {
        const v3 m0 = vadd(p0, vscl(t, v0)), m1 = vadd(p1, vscl(t, v1));
        const v3 m2 = vadd(p2, vscl(t, v2)), m3 = vadd(p3, vscl(t, v3v));
        const double dist = dist_fv(m0, m1, m2, m3, &w0, &w1);
        if (w0 < 0 || w0 > 1) continue;
        if (w1 < 0 || w1 > 1) continue;
        if (dist > delta) continue;
      }
      const double w2 = 1.0 - w0 - w1;
      const v3 pc = vadd(vadd(vscl(w0, p0), vscl(w1, p1)), vscl(w2, p2));
      const v3 nrm = vnormalized(vsub(p3, pc));
      const double rel =
          vdot(vsub(vsub(vsub(v3v, vscl(w0, v0)), vscl(w1, v1)), vscl(w2, v2)), nrm);
      if (rel > 0.1 * delta / dt) continue;
      const double imp = m * (0.1 * delta / dt - rel);
      double j = 2 * imp / (1.0 + w0 * w0 + w1 * w1 + w2 * w2);
      j /= m;
      j *= 0.1;
      vel_add(U, i0, V(-nrm.x * j * w0, -nrm.y * j * w0, -nrm.z * j * w0));
      vel_add(U, i1, V(-nrm.x * j * w1, -nrm.y * j * w1, -nrm.z * j * w1));
      vel_add(U, i2, V(-nrm.x * j * w2, -nrm.y * j * w2, -nrm.z * j * w2));
      vel_add(U, i3, V(+nrm.x * j, +nrm.y * j, +nrm.z * j));
    } else {
      double wa, wb;
      {
        const v3 m0 = vadd(p0, vscl(t, v0)), m1 = vadd(p1, vscl(t, v1));
        const v3 m2 = vadd(p2, vscl(t, v2)), m3 = vadd(p3, vscl(t, v3v));
        const double dist = dist_ee(m0, m1, m2, m3, &wa, &wb);
        if (wa < 0 || wa > 1) continue;
        if (wb < 0 || wb > 1) continue;
        if (dist > delta) continue;
      }
      const v3 ca = vadd(vscl(1 - wa, p0), vscl(wa, p1));
      const v3 cb = vadd(vscl(1 - wb, p2), vscl(wb, p3));
      const v3 nrm = vnormalized(vsub(cb, ca));
      const double rel = vdot(vsub(vsub(vadd(vscl(1 - wb, v2), vscl(wb, v3v)), vscl(1 - wa, v0)),
                                   vscl(wa, v1)),
                              nrm);
      if (rel > 0.1 * delta / dt) continue;
      const double imp = m * (0.1 * delta / dt - rel);
      double j = 2 * imp / (wa * wa + (1 - wa) * (1 - wa) + wb * wb + (1 - wb) * (1 - wb));
      j /= m;
      j *= 0.1;
      vel_add(U, i0, V(-nrm.x * j * (1 - wa), -nrm.y * j * (1 - wa), -nrm.z * j * (1 - wa)));
      vel_add(U, i1, V(-nrm.x * j * wa, -nrm.y * j * wa, -nrm.z * j * wa));
      vel_add(U, i2, V(+nrm.x * j * (1 - wb), +nrm.y * j * (1 - wb), +nrm.z * j * (1 - wb)));
      vel_add(U, i3, V(+nrm.x * j * wb, +nrm.y * j * wb, +nrm.z * j * wb));
    }
  }
}

/* ------------------------------------------------------------------------ */
/* rigid impact zones (ref: self_collision_cloth.cpp:220-348) */

typedef struct {
  zone_t* z;
  int n, cap;
} zones_t;

static int zone_has(const zone_t* z, int v) {
  int lo = 0, hi = z->n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (z->zone[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < z->n && z->zone[lo] == v;
}
static void zone_ins(zone_t* z, int v) {
  int lo = 0, hi = z->n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (z->zone[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo < z->n && z->zone[lo] == v) return;
  if (z->n == z->cap) {
    z->cap = z->cap ? z->cap * 2 : 8;
    z->zone = (int*)realloc(z->zone, sizeof(int) * z->cap);
    if (!z->zone) abort();
  }
  memmove(z->zone + lo + 1, z->zone + lo, sizeof(int) * (z->n - lo));
  z->zone[lo] = v;
  z->n++;
}

/* merge contacts into zones: a contact touches a zone through one of its
 * vertices or a 1-ring neighbour of one; touching several zones merges them
 * into a new zone appended after the untouched ones (ref: :220-274) */
static void riz_make(zones_t* Z, const ceset_t* s, const int* eidx, const int* earr) {
  int* hit = NULL;
  int hcap = 0;
  for (long e = 0; e < s->n; e++) {
    const int* n4 = s->v[e].k;
    if (hcap < Z->n + 1) {
      hcap = 2 * (Z->n + 1);
      hit = (int*)realloc(hit, sizeof(int) * hcap);
    }
    int nh = 0;
    for (int iz = 0; iz < Z->n; iz++) hit[iz] = 0;
    for (int a = 0; a < 4; a++) {
      const int v = n4[a];
      for (int iz = 0; iz < Z->n; iz++) {
        if (zone_has(&Z->z[iz], v)) {
          hit[iz] = 1;
        } else {
          for (int k = eidx[v]; k < eidx[v + 1]; k++)
            if (zone_has(&Z->z[iz], earr[k])) {
              hit[iz] = 1;
              break;
            }
        }
      }
    }
    for (int iz = 0; iz < Z->n; iz++) nh += hit[iz];
    if (nh == 0) {
      if (Z->n == Z->cap) {
        Z->cap = Z->cap ? 2 * Z->cap : 8;
        Z->z = (zone_t*)realloc(Z->z, sizeof(zone_t) * Z->cap);
      }
      zone_t nz = {NULL, 0, 0};
      for (int a = 0; a < 4; a++) zone_ins(&nz, n4[a]);
      Z->z[Z->n++] = nz;
    } else if (nh == 1) {
      int iz = 0;
      while (!hit[iz]) iz++;
      for (int a = 0; a < 4; a++) zone_ins(&Z->z[iz], n4[a]);
    } else {
      zone_t merged = {NULL, 0, 0};
      int m = 0;
      for (int iz = 0; iz < Z->n; iz++) {
        if (hit[iz]) {
          for (int q = 0; q < Z->z[iz].n; q++) zone_ins(&merged, Z->z[iz].zone[q]);
          free(Z->z[iz].zone);
        } else {
          Z->z[m++] = Z->z[iz];
        }
      }
      for (int a = 0; a < 4; a++) zone_ins(&merged, n4[a]);
      Z->z[m++] = merged;
      Z->n = m;
    }
  }
  free(hit);
}

/* cofactor inverse into a separate array (ref: self_collision_cloth.cpp:278-296) */
static void inv3(double* r, const double* a) {
  const double det = +a[0] * a[4] * a[8] + a[3] * a[7] * a[2] + a[6] * a[1] * a[5] -
                     a[0] * a[7] * a[5] - a[6] * a[4] * a[2] - a[3] * a[1] * a[8];
  const double id = 1.0 / det;
  r[0] = id * (a[4] * a[8] - a[5] * a[7]);
  r[1] = id * (a[2] * a[7] - a[1] * a[8]);
  r[2] = id * (a[1] * a[5] - a[2] * a[4]);
  r[3] = id * (a[5] * a[6] - a[3] * a[8]);
  r[4] = id * (a[0] * a[8] - a[2] * a[6]);
  r[5] = id * (a[2] * a[3] - a[0] * a[5]);
  r[6] = id * (a[3] * a[7] - a[4] * a[6]);
  r[7] = id * (a[1] * a[6] - a[0] * a[7]);
  r[8] = id * (a[0] * a[4] - a[1] * a[3]);
}

/* rigid velocity field per zone from the pre-zone velocities U0: mean
 * velocity plus omega x r, omega = I^-1 L.  The inertia diagonal uses |v|^2
 * where a rigid inertia would use |q|^2 (ref: :327-329). */
static void riz_apply(double* U, const zones_t* Z, const double* X, const double* U0) {
  for (int iz = 0; iz < Z->n; iz++) {
    const zone_t* z = &Z->z[iz];
    v3 gc = V(0, 0, 0), av = V(0, 0, 0);
    for (int q = 0; q < z->n; q++) {
      gc = vadd(gc, vload(X, z->zone[q]));
      av = vadd(av, vload(U0, z->zone[q]));
    }
    {
      const double dn = (double)z->n;
      if (!(fabs(dn) < 1.e-16)) {
        gc = V(gc.x / dn, gc.y / dn, gc.z / dn);
        av = V(av.x / dn, av.y / dn, av.z / dn);
      }
    }
    v3 L = V(0, 0, 0);
    double I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < z->n; q++) {
      const v3 p = vload(X, z->zone[q]);
      const v3 v = vload(U0, z->zone[q]);
      L = vadd(L, vcross(vsub(p, gc), vsub(v, av)));
      const v3 r = vsub(p, gc);
      const double vv = vdot(v, v);
      I[0] += vv - r.x * r.x;
      I[1] += -r.x * r.y;
      I[2] += -r.x * r.z;
      I[3] += -r.y * r.x;
      I[4] += vv - r.y * r.y;
      I[5] += -r.y * r.z;
      I[6] += -r.z * r.x;
      I[7] += -r.z * r.y;
      I[8] += vv - r.z * r.z;
    }
    double Ii[9];
    inv3(Ii, I);
    const v3 om = V(Ii[0] * L.x + Ii[1] * L.y + Ii[2] * L.z, Ii[3] * L.x + Ii[4] * L.y + Ii[5] * L.z,
                    Ii[6] * L.x + Ii[7] * L.y + Ii[8] * L.z);
    for (int q = 0; q < z->n; q++) {
      const int i = z->zone[q];
      const v3 rot = vscl(-1.0, vcross(vsub(vload(X, i), gc), om));
      U[i * 3 + 0] = av.x + rot.x;
      U[i * 3 + 1] = av.y + rot.y;
      U[i * 3 + 2] = av.z + rot.z;
    }
  }
}

/* GetIntermidiateVelocityContactResolved (ref: self_collision_cloth.cpp:355-456):
 * one proximity pass, up to 5 CCD passes, then up to 100 impact-zone passes.
 * X = start-of-step positions; U = velocities (in/out). Returns whether any
 * impulse was applied. */
static int contact_resolve(ctx_t* c, const double* X, double* U) {
  int applied;
  {
    ceset_t s = {NULL, 0, 0};
    query_prox(c, X, c->clear, &s);
    applied = s.n > 0;
    impulse_prox(U, c->clear, c->kc, c->dt, c->mass, X, &s);
    free(s.v);
  }
  for (int it = 0; it < 5; it++) {
    ceset_t s = {NULL, 0, 0};
    query_ccd(c, X, U, c->dt, c->clear, &s);
    if (s.n == 0) {
      free(s.v);
      return applied;
    }
    applied = 1;
    impulse_ccd(U, c->clear, c->dt, c->mass, X, &s);
    free(s.v);
  }
  const int n = c->np * 3;
  double* U0 = (double*)xmalloc(sizeof(double) * n);
  memcpy(U0, U, sizeof(double) * n);
  zones_t Z = {NULL, 0, 0};
  for (int it = 0; it < 100; it++) {
    ceset_t s = {NULL, 0, 0};
    query_ccd(c, X, U, c->dt, c->clear, &s);
    if (s.n == 0) {
      free(s.v);
      break;
    }
    riz_make(&Z, &s, c->eidx, c->earr);
    riz_apply(U, &Z, X, U0);
    free(s.v);
  }
  for (int iz = 0; iz < Z.n; iz++) free(Z.z[iz].zone);
  free(Z.z);
  free(U0);
  return applied;
}

int orc_contact_resolve(void* h, const double* xyz1, double* uvw) {
  return contact_resolve((ctx_t*)h, xyz1, uvw);
}

/* UpdateIntermidiateVelocity (ref: solve_internal_sparse.h:329-369) */
static void update_velocity(ctx_t* c, double* U, const double* X) {
  const int n = c->np * 3;
  double* dW = (double*)xcalloc(n, sizeof(double));
  add_cloth(c, X, dW, NULL, NULL);
  add_gravity(c, dW);
  for (int i = 0; i < c->np; i++) {
    if (c->bc[i] == 0) continue;
    U[i * 3 + 0] = U[i * 3 + 1] = U[i * 3 + 2] = 0;
  }
  for (int i = 0; i < c->np; i++)
    for (int k = 0; k < 3; k++) U[i * 3 + k] += -(0.5 * c->dt / c->mass) * dW[i * 3 + k];
  free(dW);
}

void orc_update_velocity(void* h, double* uvw, const double* xyz) {
  update_velocity((ctx_t*)h, uvw, xyz);
}

/* StepTime (ref: self_contact_sparse/main.cpp:213-261) */
void orc_step(void* h, double* xyz, double* uvw, int* stats, double* conv) {
  ctx_t* c = (ctx_t*)h;
  const int n = c->np * 3;
  double* X1 = (double*)xmalloc(sizeof(double) * n);
  memcpy(X1, xyz, sizeof(double) * n);
  internal_step(c, xyz, uvw, c->tol, c->maxit);
  const int applied = contact_resolve(c, X1, uvw);
  if (applied) {
    for (int i = 0; i < n; i++) xyz[i] = X1[i] + uvw[i] * c->dt;
    update_velocity(c, uvw, xyz);
  }
  free(X1);
  stats[0] = applied;
  stats[1] = c->last_iter;
  *conv = c->last_conv;
}
