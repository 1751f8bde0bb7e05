// Contact-resolved velocity pass on the device
// (self_collision_cloth.cpp:355-456, bvh_aabb.cpp).
//
//  refit      one launch, bottom-up: every leaf thread climbs while it is the
//             second child to arrive (arrival counters are never reset: each
//             refit adds exactly two arrivals per internal node, so the parity
//             of the counter tells first from second).
//  traverse   one thread per leaf walks the BVH in pre-order (stackless, the
//             subtree end of every node is its skip link) and emits every
//             later leaf whose box overlaps: exactly the leaf pairs the
//             reference's recursion visits, oriented the same way (the leaf
//             under the LCA's child 0 first).  The 6 face-vertex and 9
//             edge-edge tests run on the pair in the reference's order.
//  select     records are radix-sorted by canonical key; for a key found as
//             both face-vertex and edge-edge the visit that comes first in the
//             reference's depth-first order wins (std::set keeps the first
//             insertion, bvh_aabb.h:96-102).
//  impulses   Gauss-Seidel in key order, scheduled in rounds: a contact runs
//             once every earlier contact sharing a vertex has run, so each
//             velocity sees exactly the updates it sees sequentially.
//  zones      rigid impact zones by union-find in one workgroup, zone sums in
//             ascending vertex order.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>

#include "cs_contact.hpp"
#include "cs_math.hpp"
#include "cs_solve.hpp"

namespace cs {

namespace {
constexpr int kTB = 256;
inline unsigned grid_for(long n, int tb = kTB) { return (unsigned)((n + tb - 1) / tb); }
enum Counter : int {
  kRaw = 0,       // narrow-phase records emitted
  kGroups = 1,    // distinct contacts
  kDoneCnt = 2,   // contacts whose impulse has run
  kCand = 3,      // overlapping leaf pairs found by the traversal
  kReady = 4,     // contacts with no earlier contact on their vertices
  kQueryCounters = 8,  // a query clears counters [0, kQueryCounters)
  kZoneVerts = 8, // vertices in impact zones after the last zone pass
  kNumCounters = 16
};
}  // namespace

__device__ __forceinline__ Box load_box(const double* bb, int n) {
  Box b;
  b.lx = bb[6 * n + 0];
  b.hx = bb[6 * n + 1];
  b.ly = bb[6 * n + 2];
  b.hy = bb[6 * n + 3];
  b.lz = bb[6 * n + 4];
  b.hz = bb[6 * n + 5];
  return b;
}
__device__ __forceinline__ void store_box(double* bb, int n, const Box& b) {
  bb[6 * n + 0] = b.lx;
  bb[6 * n + 1] = b.hx;
  bb[6 * n + 2] = b.ly;
  bb[6 * n + 3] = b.hy;
  bb[6 * n + 4] = b.lz;
  bb[6 * n + 5] = b.hz;
}

// ---------------------------------------------------------------------------
// refit (bvh_aabb.cpp:23-106)

// Leaf boxes from their triangles; the internal nodes follow level by level
// (k_refit_merge), deepest first.  A box merge is a componentwise min / max,
// exact in any order, so the boxes are the reference's.
__global__ void k_refit_leaves(DevCloth c, const double* X, const double* U, double delta,
                               double dt, int ccd) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.nt) return;
  const int i0 = c.tri[3 * t], i1 = c.tri[3 * t + 1], i2 = c.tri[3 * t + 2];
  Box b;
  if (ccd) {
    const double eps = 1.0e-10;
    b = box_point(at(X, i0), eps);
    box_grow(b, at(X, i1), eps);
    box_grow(b, at(X, i2), eps);
    box_grow(b, advanced(X, U, dt, i0), eps);
    box_grow(b, advanced(X, U, dt, i1), eps);
    box_grow(b, advanced(X, U, dt, i2), eps);
  } else {
    const double eps = delta * 0.5;
    b = box_point(at(X, i0), eps);
    box_grow(b, at(X, i1), eps);
    box_grow(b, at(X, i2), eps);
  }
  const int L = c.leaf_of[t];
  store_box(c.bb, L, b);
  store_box(c.rbb, c.pre[L], b);
}

// the internal nodes among refit_order[b .. e) (one depth): box = merge of
// the children's boxes, written by earlier launches
__global__ void k_refit_merge(DevCloth c, int b, int e) {
  const int k = b + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= e) return;
  const int n = c.refit_order[k];
  if (c.node[3 * n + 2] == -1) return;  // a leaf: boxed from its triangle
  const Box m = box_merge(load_box(c.bb, c.node[3 * n + 1]), load_box(c.bb, c.node[3 * n + 2]));
  store_box(c.bb, n, m);
  store_box(c.rbb, c.pre[n], m);
}

// ---------------------------------------------------------------------------
// narrow phase

// canonical key of a contact element (bvh_aabb.h:68-95): a face-vertex key
// lists the face ascending then the vertex; an edge-edge key starts with the
// edge holding the smallest index, each edge ascending
__device__ __forceinline__ void canonical(bool fv, int j0, int j1, int j2, int j3, int k[4]) {
  if (fv) {
    int a = j0, b = j1, d = j2, t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > d) { t = b; b = d; d = t; }
    if (a > b) { t = a; a = b; b = t; }
    k[0] = a;
    k[1] = b;
    k[2] = d;
    k[3] = j3;
  } else {
    const int a0 = min(j0, j1), a1 = max(j0, j1), b0 = min(j2, j3), b1 = max(j2, j3);
    if (a0 < b0) {
      k[0] = a0; k[1] = a1; k[2] = b0; k[3] = b1;
    } else {
      k[0] = b0; k[1] = b1; k[2] = a0; k[3] = a1;
    }
  }
}

struct Emit {
  int* counters;
  int cap;
  int* rkey;
  int* rinfo;
  int* rpair;
  __device__ void operator()(bool fv, int test, int first, int second, int j0, int j1, int j2,
                             int j3) const {
    const int slot = atomicAdd(&counters[kRaw], 1);
    if (slot >= cap) return;
    int k[4];
    canonical(fv, j0, j1, j2, j3, k);
    rkey[4 * slot + 0] = k[0];
    rkey[4 * slot + 1] = k[1];
    rkey[4 * slot + 2] = k[2];
    rkey[4 * slot + 3] = k[3];
    rinfo[slot] = (fv ? 1 : 0) | (test << 1);
    rpair[2 * slot] = first;
    rpair[2 * slot + 1] = second;
  }
};

// leaf-leaf tests of one pair: face-vertex 0..5 then edge-edge 6..14
// (bvh_aabb.cpp:475-536 proximity, 833-900 continuous)
__device__ void narrow_pair(const DevCloth& c, const double* X, const double* U, double dt,
                            double delta, int ccd, int L, int M, const Emit& emit) {
  const int it = c.node[3 * L + 1], jt = c.node[3 * M + 1];
  const int iv[3] = {c.tri[3 * it], c.tri[3 * it + 1], c.tri[3 * it + 2]};
  const int jv[3] = {c.tri[3 * jt], c.tri[3 * jt + 1], c.tri[3 * jt + 2]};
  Vec3 P[3], Q[3];
  for (int a = 0; a < 3; a++) {
    P[a] = at(X, iv[a]);
    Q[a] = at(X, jv[a]);
  }
  if (!ccd) {
    // the face tests read the box of the node whose id equals the triangle id
    // (bvh_aabb.cpp:490-507)
    const Box bi = load_box(c.bb, it), bj = load_box(c.bb, jt);
    for (int a = 0; a < 3; a++)
      if (fv_near(iv[0], iv[1], iv[2], jv[a], P[0], P[1], P[2], Q[a], bi, delta))
        emit(true, a, L, M, iv[0], iv[1], iv[2], jv[a]);
    for (int a = 0; a < 3; a++)
      if (fv_near(jv[0], jv[1], jv[2], iv[a], Q[0], Q[1], Q[2], P[a], bj, delta))
        emit(true, 3 + a, L, M, jv[0], jv[1], jv[2], iv[a]);
    for (int a = 0; a < 3; a++) {
      const int a1 = a == 2 ? 0 : a + 1;
      for (int b = 0; b < 3; b++) {
        const int b1 = b == 2 ? 0 : b + 1;
        if (ee_near(iv[a], iv[a1], jv[b], jv[b1], P[a], P[a1], Q[b], Q[b1], delta))
          emit(false, 6 + a * 3 + b, L, M, iv[a], iv[a1], jv[b], jv[b1]);
      }
    }
  } else {
    Vec3 Pe[3], Qe[3];
    for (int a = 0; a < 3; a++) {
      Pe[a] = advanced(X, U, dt, iv[a]);
      Qe[a] = advanced(X, U, dt, jv[a]);
    }
    const Box bi = load_box(c.bb, L), bj = load_box(c.bb, M);
    for (int a = 0; a < 3; a++)
      if (fv_sweep(iv[0], iv[1], iv[2], jv[a], P[0], P[1], P[2], Q[a], Pe[0], Pe[1], Pe[2], Qe[a],
                   bi))
        emit(true, a, L, M, iv[0], iv[1], iv[2], jv[a]);
    for (int a = 0; a < 3; a++)
      if (fv_sweep(jv[0], jv[1], jv[2], iv[a], Q[0], Q[1], Q[2], P[a], Qe[0], Qe[1], Qe[2], Pe[a],
                   bj))
        emit(true, 3 + a, L, M, jv[0], jv[1], jv[2], iv[a]);
    for (int a = 0; a < 3; a++) {
      const int a1 = a == 2 ? 0 : a + 1;
      for (int b = 0; b < 3; b++) {
        const int b1 = b == 2 ? 0 : b + 1;
        if (ee_sweep(iv[a], iv[a1], jv[b], jv[b1], P[a], P[a1], Q[b], Q[b1], Pe[a], Pe[a1], Qe[b],
                     Qe[b1]))
          emit(false, 6 + a * 3 + b, L, M, iv[a], iv[a1], jv[b], jv[b1]);
      }
    }
  }
}

// Overlapping leaf pairs (L, M), M after L in pre-order, as the reference's
// pair recursion finds them (bvh_aabb.cpp:461-474, 544-556): the records are
// sorted by key afterwards and k_select orders visits by the tree alone, so
// the order in which pairs are found does not matter.
//
// One wave per 64 consecutive leaves in pre-order.  Each lane's own walk is a
// pre-order scan from the root that skips a subtree (jumps to its end) when
// it holds no leaf after the lane's leaf or its box does not overlap the
// leaf's box, and records every leaf it reaches after its own.  The wave walks
// the union of its lanes' walks: at rank r a lane takes part iff r is at or
// past the end of the last subtree it skipped, so every lane tests exactly
// the nodes its own walk tests; the wave goes on to r + 1 if any lane
// descends, else to the end of r's subtree, the first rank at which any lane
// takes part again.  The node data are read by wave-uniform rank (scalar
// loads of rend / rbb / by_pre), the leaf's box stays in registers, and the
// pairs go through an LDS buffer, one counter atomic per kTravBuf pairs.
constexpr int kTravBuf = 512;

// overlap() evaluated without short-circuit branches (each comparison as
// written there, so NaN coordinates behave the same), which keeps all six
// bounds of a box in one round of loads
__device__ __forceinline__ bool overlap_all(const Box& a, const Box& b) {
  return !(a.hx < b.lx) & !(a.lx > b.hx) & !(a.hy < b.ly) & !(a.ly > b.hy) & !(a.hz < b.lz) &
         !(a.lz > b.hz);
}

__device__ __forceinline__ void trav_flush(const int2* buf, int fill, int* __restrict__ counters,
                                           int cap, int2* __restrict__ cand) {
  __syncthreads();  // the buffer's entries, written by every lane
  int base = 0;
  if (threadIdx.x == 0) base = atomicAdd(&counters[kCand], fill);
  base = __shfl(base, 0);
  for (int i = (int)threadIdx.x; i < fill; i += 64)
    if (base + i < cap) cand[base + i] = buf[i];
  __syncthreads();  // read out before the buffer is reused
}

__global__ __launch_bounds__(64) void k_traverse(int nnode, int nt, int per,
                                                 const int* __restrict__ leaf_rank,
                                                 const int* __restrict__ rend,
                                                 const double* __restrict__ rbb,
                                                 const int* __restrict__ by_pre,
                                                 int* __restrict__ counters, int cap,
                                                 int2* __restrict__ cand) {
  __shared__ int2 buf[kTravBuf];
  // per (1..64) leaves per wave: fewer leaves make shorter union walks and
  // more waves, which pays where the tree is small and few waves would run
  const int k = (int)(blockIdx.x * per + threadIdx.x);
  int rl = 0, L = -1, skip = INT_MAX;  // lanes past the last leaf never take part
  Box bl{};
  if ((int)threadIdx.x < per && k < nt) {
    rl = leaf_rank[k];
    L = by_pre[rl];
    bl = load_box(rbb, rl);
    skip = 0;
  }
  int fill = 0;
  int r = 0;
  while (r < nnode) {
    // all three loads issue together (no branch between them and their use)
    const int end = rend[r];
    const Box b = load_box(rbb, r);
    const int m = by_pre[r];
    const bool alive = r >= skip;
    const bool down = alive & (end > rl + 1) & overlap_all(bl, b);
    const bool hit = down & (end == r + 1) & (r > rl);
    skip = (alive & !down) ? end : skip;
    const unsigned long long hm = __ballot(hit);
    if (hm) {
      const int cnt = __popcll(hm);
      if (fill + cnt > kTravBuf) {
        trav_flush(buf, fill, counters, cap, cand);
        fill = 0;
      }
      if (hit) {
        const int below = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(hm >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((unsigned)hm, 0u));
        buf[fill + below] = make_int2(L, m);
      }
      fill += cnt;
    }
    r = __ballot(down) ? r + 1 : end;
  }
  if (fill) trav_flush(buf, fill, counters, cap, cand);
}

// n overlapping leaf pairs (n <= the buffer's capacity)
__global__ void k_narrow(DevCloth c, const double* X, const double* U, double dt, double delta,
                         int ccd, int n, const int* cand, Emit emit) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  narrow_pair(c, X, U, dt, delta, ccd, cand[2 * p], cand[2 * p + 1], emit);
}

// ---------------------------------------------------------------------------
// selection: sort by key, then per key the first visit in depth-first order

__global__ void k_make_keys(const int* rkey, int n, int nbits, int pass, uint64_t* key, int* idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k0 = (uint64_t)rkey[4 * i], k1 = (uint64_t)rkey[4 * i + 1];
  const uint64_t k2 = (uint64_t)rkey[4 * i + 2], k3 = (uint64_t)rkey[4 * i + 3];
  uint64_t k;
  if (pass == 0)  // whole key fits in 64 bits
    k = (((k0 << nbits | k1) << nbits | k2) << nbits) | k3;
  else if (pass == 1)  // least significant part of a two-pass sort
    k = k3;
  else
    k = ((k0 << nbits | k1) << nbits) | k2;
  key[i] = k;
  if (pass != 2) idx[i] = i;
}

__global__ void k_rekey(const int* rkey, const int* order, int n, int nbits, uint64_t* key) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = order[i];
  const uint64_t k0 = (uint64_t)rkey[4 * r], k1 = (uint64_t)rkey[4 * r + 1];
  const uint64_t k2 = (uint64_t)rkey[4 * r + 2];
  key[i] = ((k0 << nbits | k1) << nbits) | k2;
}

__device__ __forceinline__ bool same_key(const int* rkey, int a, int b) {
  return rkey[4 * a] == rkey[4 * b] && rkey[4 * a + 1] == rkey[4 * b + 1] &&
         rkey[4 * a + 2] == rkey[4 * b + 2] && rkey[4 * a + 3] == rkey[4 * b + 3];
}

__global__ void k_heads(const int* rkey, const int* order, int n, int* head) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  head[i] = (i == 0 || !same_key(rkey, order[i], order[i - 1])) ? 1 : 0;
}

__device__ __forceinline__ bool in_subtree(const DevCloth& c, int n, int leaf) {
  const int r = c.pre[leaf];
  return c.pre[n] <= r && r < c.pre_end[n];
}

// Does visit (L0, M0, test t0) come before (L1, M1, t1) in the reference's
// traversal?  self(N) visits the pairs whose LCA is N before those of its
// subtrees (LCAs in pre-order); within one LCA the pair recursion descends
// both sides at once and orders children pairs (00, 10, 01, 11)
// (bvh_aabb.cpp:461-474, 550-556).
__device__ bool visited_before(const DevCloth& c, int L0, int M0, int t0, int L1, int M1, int t1) {
  if (L0 == L1 && M0 == M1) return t0 < t1;
  int A0 = L0;
  while (!in_subtree(c, A0, M0)) A0 = c.node[3 * A0];
  int A1 = L1;
  while (!in_subtree(c, A1, M1)) A1 = c.node[3 * A1];
  if (A0 != A1) return c.pre[A0] < c.pre[A1];
  int X = c.node[3 * A0 + 1], Y = c.node[3 * A0 + 2];
  for (int guard = 0; guard < 4 * c.nnode; guard++) {
    const bool lx = c.node[3 * X + 2] == -1, ly = c.node[3 * Y + 2] == -1;
    const int xb0 = lx ? 0 : (in_subtree(c, c.node[3 * X + 1], L0) ? 0 : 1);
    const int xb1 = lx ? 0 : (in_subtree(c, c.node[3 * X + 1], L1) ? 0 : 1);
    const int yb0 = ly ? 0 : (in_subtree(c, c.node[3 * Y + 1], M0) ? 0 : 1);
    const int yb1 = ly ? 0 : (in_subtree(c, c.node[3 * Y + 1], M1) ? 0 : 1);
    const int d0 = xb0 + 2 * yb0, d1 = xb1 + 2 * yb1;
    if (d0 != d1) return d0 < d1;
    if (lx && ly) return false;  // identical pairs (handled above)
    if (!lx) X = c.node[3 * X + 1 + xb0];
    if (!ly) Y = c.node[3 * Y + 1 + yb0];
  }
  return false;
}

__global__ void k_select(DevCloth c, const int* rkey, const int* rinfo, const int* rpair,
                         const int* order, const int* head, const int* gpos, int n, int* ce,
                         int* counters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i == n - 1) counters[kGroups] = gpos[i] + head[i];
  if (!head[i]) return;
  int e = i + 1;
  while (e < n && !head[e]) e++;
  int w = order[i];
  bool mixed = false;
  for (int k = i + 1; k < e; k++) mixed |= ((rinfo[order[k]] ^ rinfo[w]) & 1) != 0;
  if (mixed) {
    for (int k = i + 1; k < e; k++) {
      const int r = order[k];
      if (visited_before(c, rpair[2 * r], rpair[2 * r + 1], rinfo[r] >> 1, rpair[2 * w],
                         rpair[2 * w + 1], rinfo[w] >> 1))
        w = r;
    }
  }
  const int g = gpos[i];
  ce[5 * g] = rinfo[w] & 1;
  for (int q = 0; q < 4; q++) ce[5 * g + 1 + q] = rkey[4 * w + q];
}

// ---------------------------------------------------------------------------
// impulses (self_collision_cloth.cpp:20-212)

__device__ __forceinline__ void vel_add(double* U, int i, double a, double b, double d) {
  U[3 * i] += a;
  U[3 * i + 1] += b;
  U[3 * i + 2] += d;
}

__device__ void impulse_one(const int* e, const double* X, double* U, double delta, double k,
                            double dt, double m, int ccd) {
  const int i0 = e[1], i1 = e[2], i2 = e[3], i3 = e[4];
  const Vec3 p0 = at(X, i0), p1 = at(X, i1), p2 = at(X, i2), p3 = at(X, i3);
  const Vec3 v0 = at(U, i0), v1 = at(U, i1), v2 = at(U, i2), v3 = at(U, i3);
  double t = 0;
  if (ccd) {
    // end positions are p + v, not p + dt v (self_collision_cloth.cpp:141)
    t = coplanar_time(p0, p1, p2, p3, p0 + v0, p1 + v1, p2 + v2, p3 + v3);
    if (t < 0 || t > 1) return;
  }
  if (e[0]) {  // face (i0,i1,i2) - vertex i3
    double w0, w1;
    {
      double dist;
      if (ccd)
        dist = face_vertex(p0 + t * v0, p1 + t * v1, p2 + t * v2, p3 + t * v3, w0, w1);
      else
        dist = face_vertex(p0, p1, p2, p3, w0, w1);
      if (w0 < 0 || w0 > 1) return;
      if (w1 < 0 || w1 > 1) return;
      if (dist > delta) return;
    }
    const double w2 = 1.0 - w0 - w1;
    const Vec3 pc = w0 * p0 + w1 * p1 + w2 * p2;
    const Vec3 nrm = unit(p3 - pc);
    const double rel = dot(v3 - w0 * v0 - w1 * v1 - w2 * v2, nrm);
    double j;
    if (!ccd) {
      const double depth = delta - dot(p3 - pc, nrm);
      if (rel > 0.1 * depth / dt) return;
      const double el = dt * k * depth;
      const double ie = m * (0.1 * depth / dt - rel);
      const double imin = (el < ie) ? el : ie;
      j = 2 * imin / (1 + w0 * w0 + w1 * w1 + w2 * w2);
      j /= m;
      j *= 0.25;
    } else {
      if (rel > 0.1 * delta / dt) return;
      const double imp = m * (0.1 * delta / dt - rel);
      j = 2 * imp / (1.0 + w0 * w0 + w1 * w1 + w2 * w2);
      j /= m;
      j *= 0.1;
    }
    vel_add(U, i0, -nrm.x * j * w0, -nrm.y * j * w0, -nrm.z * j * w0);
    vel_add(U, i1, -nrm.x * j * w1, -nrm.y * j * w1, -nrm.z * j * w1);
    vel_add(U, i2, -nrm.x * j * w2, -nrm.y * j * w2, -nrm.z * j * w2);
    vel_add(U, i3, +nrm.x * j, +nrm.y * j, +nrm.z * j);
  } else {  // edge (i0,i1) - edge (i2,i3)
    double wa, wb;
    {
      double dist;
      if (ccd)
        dist = edge_edge(p0 + t * v0, p1 + t * v1, p2 + t * v2, p3 + t * v3, wa, wb);
      else
        dist = edge_edge(p0, p1, p2, p3, wa, wb);
      if (wa < 0 || wa > 1) return;
      if (wb < 0 || wb > 1) return;
      if (dist > delta) return;
    }
    const Vec3 ca = (1 - wa) * p0 + wa * p1;
    const Vec3 cb = (1 - wb) * p2 + wb * p3;
    const Vec3 nrm = unit(cb - ca);
    const double rel = dot((1 - wb) * v2 + wb * v3 - (1 - wa) * v0 - wa * v1, nrm);
    double j;
    if (!ccd) {
      const double depth = delta - norm(cb - ca);
      if (rel > 0.1 * depth / dt) return;
      const double el = dt * k * depth;
      const double ie = m * (0.1 * depth / dt - rel);
      const double imin = (el < ie) ? el : ie;
      j = 2 * imin / (wa * wa + (1 - wa) * (1 - wa) + wb * wb + (1 - wb) * (1 - wb));
      j /= m;
      j *= 0.25;
    } else {
      if (rel > 0.1 * delta / dt) return;
      const double imp = m * (0.1 * delta / dt - rel);
      j = 2 * imp / (wa * wa + (1 - wa) * (1 - wa) + wb * wb + (1 - wb) * (1 - wb));
      j /= m;
      j *= 0.1;
    }
    vel_add(U, i0, -nrm.x * j * (1 - wa), -nrm.y * j * (1 - wa), -nrm.z * j * (1 - wa));
    vel_add(U, i1, -nrm.x * j * wa, -nrm.y * j * wa, -nrm.z * j * wa);
    vel_add(U, i2, +nrm.x * j * (1 - wb), +nrm.y * j * (1 - wb), +nrm.z * j * (1 - wb));
    vel_add(U, i3, +nrm.x * j * wb, +nrm.y * j * wb, +nrm.z * j * wb);
  }
}

__global__ void k_inc_count(const int* ce, int g, int* vcount) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g) return;
  for (int a = 0; a < 4; a++) atomicAdd(&vcount[ce[5 * i + 1 + a]], 1);
}

__global__ void k_inc_fill(const int* ce, int g, const int* vstart, int* vfill, int* inc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g) return;
  for (int a = 0; a < 4; a++) {
    const int v = ce[5 * i + 1 + a];
    inc[vstart[v] + atomicAdd(&vfill[v], 1)] = 4 * i + a;
  }
}

// order each vertex's incidences by contact and link every incidence to the
// previous and the next contact on that vertex
__global__ void k_inc_link(int np, const int* vstart, int* inc, int* pred, int* succ) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= np) return;
  const int b = vstart[v], e = vstart[v + 1];
  for (int i = b + 1; i < e; i++) {
    const int x = inc[i];
    int j = i - 1;
    while (j >= b && inc[j] > x) {
      inc[j + 1] = inc[j];
      j--;
    }
    inc[j + 1] = x;
  }
  for (int i = b; i < e; i++) {
    pred[inc[i]] = (i == b) ? -1 : (inc[i - 1] >> 2);
    succ[inc[i]] = (i + 1 == e) ? -1 : (inc[i + 1] >> 2);
  }
}

// contacts with no earlier contact on their vertices start the schedule
__global__ void k_impulse_ready(int g, const int* pred, int* npred, int* ready, int* counters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g) return;
  int c = 0;
  for (int a = 0; a < 4; a++) c += pred[4 * i + a] >= 0 ? 1 : 0;
  npred[i] = c;
  if (c == 0) ready[atomicAdd(&counters[kReady], 1)] = i;
}

// Runs the ready contacts and every contact that becomes ready, round by
// round, in one workgroup: a contact is ready once every earlier contact on
// its vertices has run.  The contacts of a round share no vertex and each
// reads its vertices' velocities after every earlier contact on them has
// written them, so the velocities are those of the key-order Gauss-Seidel
// sweep.  Every contact enters a list once, so the loop ends after at most
// g contacts; the count of those run goes to counters[kDoneCnt].
//
// A round's chain of memory latencies is what the schedule waits for, so
// the lists after the first live in LDS (up to kDrainLds entries; the rest
// of a longer round in the global list), and a contact releases its
// successors before its impulse: they only run in the next round, after
// the barrier that ends this one, and their counter atomics overlap the
// impulse's loads.  lds_cap (at most kDrainLds) is the entries per LDS list;
// a smaller value (the CLOTHSIM_DRAIN_LDS environment variable, a test hook)
// sends more of every round through the global lists.
constexpr int kDrainLds = 4096;

__global__ __launch_bounds__(1024) void k_impulse_drain(const int* ce, const int* succ, int* npred,
                                                        int* ready0, int* ready1, const double* X,
                                                        double* U, double delta, double k,
                                                        double dt, double m, int ccd,
                                                        int* counters, int lds_cap) {
  __shared__ int n_next;
  __shared__ int lists[2][kDrainLds];
  int n = counters[kReady];
  int *gcur = ready0, *gnxt = ready1;  // the first round's list and the spills
  int cur = 0;                         // the LDS list of this round
  bool first = true;                   // the first round reads k_impulse_ready's list
  int done = 0;
  while (n > 0) {
    if (threadIdx.x == 0) n_next = 0;
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
      const int i = (first || t >= lds_cap) ? gcur[t] : lists[cur][t];
      for (int a = 0; a < 4; a++) {
        const int sc = succ[4 * i + a];
        if (sc >= 0 && atomicSub(&npred[sc], 1) == 1) {
          const int pos = atomicAdd(&n_next, 1);
          if (pos < lds_cap)
            lists[cur ^ 1][pos] = sc;
          else
            gnxt[pos] = sc;
        }
      }
      impulse_one(ce + 5 * i, X, U, delta, k, dt, m, ccd);
    }
    done += n;
    __syncthreads();  // the round's velocities and the next list are complete
    n = n_next;
    int* tmp = gcur;
    gcur = gnxt;
    gnxt = tmp;
    cur ^= 1;
    first = false;
    __syncthreads();  // every thread has read n_next before it is reset
  }
  if (threadIdx.x == 0) counters[kDoneCnt] = done;
}

// ---------------------------------------------------------------------------
// rigid impact zones (self_collision_cloth.cpp:220-348), one workgroup

__device__ int zfind(int* par, int v) {
  while (par[v] != v) {
    par[v] = par[par[v]];
    v = par[v];
  }
  return v;
}
__device__ int zfind_ro(const int* par, int v) {
  while (par[v] != v) v = par[v];
  return v;
}

// block-wide exclusive scan of one int per thread; the block total goes to *total
__device__ int block_excl_scan(int x, int* total) {
  __shared__ int ws[16];
  __shared__ int tot;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int v = x;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  if (lane == 63) ws[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
      const int t = ws[w];
      ws[w] = acc;
      acc += t;
    }
    tot = acc;
  }
  __syncthreads();
  const int ex = v - x + ws[wid];
  const int t = tot;
  __syncthreads();
  *total = t;
  return ex;
}

__global__ __launch_bounds__(1024) void k_riz_pass(DevCloth c, const int* ce, int g, int* zpar,
                                                   int* zroot, int* zmemb, int* zroots,
                                                   const double* X, const double* U0, double* U,
                                                   int* counters) {
  // 1. zones touched by each contact through a vertex or its 1-ring; merge
  //    (only the partition matters: zone order never changes a result).
  //    Each touched zone is joined to the first one found as it is found, so
  //    any number of zones merges; membership (zpar >= 0) is what the
  //    reference tests, and a merge does not change it.
  if (threadIdx.x == 0) {
    for (int e = 0; e < g; e++) {
      const int* n4 = ce + 5 * e + 1;
      int root = -1;
      for (int a = 0; a < 4; a++) {
        const int v = n4[a];
        if (zpar[v] >= 0) {
          const int r = zfind(zpar, v);
          if (root < 0) root = r; else if (r != root) zpar[r] = root;
        }
        for (int q = c.ring_ptr[v]; q < c.ring_ptr[v + 1]; q++) {
          const int w = c.ring[q];
          if (zpar[w] < 0) continue;
          const int r = zfind(zpar, w);
          if (root < 0) root = r; else if (r != root) zpar[r] = root;
        }
      }
      if (root < 0) {
        root = n4[0];
        zpar[root] = root;
      }
      for (int a = 0; a < 4; a++)
        if (zpar[n4[a]] < 0) zpar[n4[a]] = root;
    }
  }
  __syncthreads();
  // 2. zone members in ascending vertex order, and the zone roots
  __shared__ int s_off, s_roff;
  if (threadIdx.x == 0) {
    s_off = 0;
    s_roff = 0;
  }
  __syncthreads();
  for (int base = 0; base < c.np; base += blockDim.x) {
    const int v = base + threadIdx.x;
    int in = 0, isroot = 0;
    if (v < c.np && zpar[v] >= 0) {
      const int r = zfind_ro(zpar, v);
      in = 1;
      isroot = (r == v);
      zroot[v] = r;
    }
    int tot, rtot;
    const int pos = block_excl_scan(in, &tot);
    const int rpos = block_excl_scan(isroot, &rtot);
    if (in) zmemb[s_off + pos] = v;
    if (isroot) zroots[s_roff + rpos] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      s_off += tot;
      s_roff += rtot;
    }
    __syncthreads();
  }
  const int nm = s_off, nz = s_roff;
  if (threadIdx.x == 0) counters[kZoneVerts] = nm;
  // 3. rigid velocity per zone from the pre-zone velocities U0
  for (int z = threadIdx.x; z < nz; z += blockDim.x) {
    const int R = zroots[z];
    Vec3 gc = mk(0, 0, 0), av = mk(0, 0, 0);
    int cnt = 0;
    for (int q = 0; q < nm; q++) {
      const int v = zmemb[q];
      if (zroot[v] != R) continue;
      gc = gc + at(X, v);
      av = av + at(U0, v);
      cnt++;
    }
    const double dn = (double)cnt;
    if (!(fabs(dn) < 1.e-16)) {  // operator/= guard (vector3d.h:63-67)
      gc = mk(gc.x / dn, gc.y / dn, gc.z / dn);
      av = mk(av.x / dn, av.y / dn, av.z / dn);
    }
    Vec3 L = mk(0, 0, 0);
    double I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < nm; q++) {
      const int v = zmemb[q];
      if (zroot[v] != R) continue;
      const Vec3 p = at(X, v), u = at(U0, v);
      L = L + cross(p - gc, u - av);
      const Vec3 r = p - gc;
      const double uu = dot(u, u);  // |v|^2 where a rigid inertia uses |r|^2 (:327-329)
      I[0] += uu - r.x * r.x;
      I[1] += -r.x * r.y;
      I[2] += -r.x * r.z;
      I[3] += -r.y * r.x;
      I[4] += uu - r.y * r.y;
      I[5] += -r.y * r.z;
      I[6] += -r.z * r.x;
      I[7] += -r.z * r.y;
      I[8] += uu - r.z * r.z;
    }
    const double det = +I[0] * I[4] * I[8] + I[3] * I[7] * I[2] + I[6] * I[1] * I[5] -
                       I[0] * I[7] * I[5] - I[6] * I[4] * I[2] - I[3] * I[1] * I[8];
    const double id = 1.0 / det;
    double J[9];
    J[0] = id * (I[4] * I[8] - I[5] * I[7]);
    J[1] = id * (I[2] * I[7] - I[1] * I[8]);
    J[2] = id * (I[1] * I[5] - I[2] * I[4]);
    J[3] = id * (I[5] * I[6] - I[3] * I[8]);
    J[4] = id * (I[0] * I[8] - I[2] * I[6]);
    J[5] = id * (I[2] * I[3] - I[0] * I[5]);
    J[6] = id * (I[3] * I[7] - I[4] * I[6]);
    J[7] = id * (I[1] * I[6] - I[0] * I[7]);
    J[8] = id * (I[0] * I[4] - I[1] * I[3]);
    const Vec3 om = mk(J[0] * L.x + J[1] * L.y + J[2] * L.z, J[3] * L.x + J[4] * L.y + J[5] * L.z,
                       J[6] * L.x + J[7] * L.y + J[8] * L.z);
    for (int q = 0; q < nm; q++) {
      const int v = zmemb[q];
      if (zroot[v] != R) continue;
      const Vec3 rr = cross(at(X, v) - gc, om);
      U[3 * v] = av.x + -rr.x;
      U[3 * v + 1] = av.y + -rr.y;
      U[3 * v + 2] = av.z + -rr.z;
    }
  }
}

// ---------------------------------------------------------------------------
// after the contact pass (self_contact_sparse/main.cpp:246-261,
// solve_internal_sparse.h:329-369)

__global__ void k_advance(DevCloth c, double dt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * c.np) return;
  c.x[e] = c.x1[e] + c.v[e] * dt;
}

__global__ void k_velocity_update(DevCloth c, Params P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.np) return;
  double g[3] = {0.0, 0.0, 0.0};
  for (int s = c.grad_src_ptr[i]; s < c.grad_src_ptr[i + 1]; s++) {
    const int src = c.grad_src[s];
    const double* e = (src >= 0) ? c.etri + (size_t)(src / 3) * 90 + 81 + (src % 3) * 3
                                 : c.eq + (size_t)(-src - 1) * 3;
    g[0] += e[0];
    g[1] += e[1];
    g[2] += e[2];
  }
  for (int k = 0; k < 3; k++) g[k] -= P.mass * P.grav[k];
  double* v = c.v + 3 * i;
  if (c.bc[i] != 0) v[0] = v[1] = v[2] = 0;
  for (int k = 0; k < 3; k++) v[k] += -(0.5 * P.dt / P.mass) * g[k];
}

// ---------------------------------------------------------------------------
// host side

namespace {

// device buffers of the contact state: allocated and released in the order
// of the cloth's stream s, without synchronising the device
template <typename T>
int dalloc(T** p, size_t n, hipStream_t s) {
  return dev_alloc((void**)p, sizeof(T) * (n ? n : 1), s);
}
template <typename T>
void dfree(T*& p, hipStream_t s) {
  dev_free(p, s);
  p = nullptr;
}

int grow_records(ContactState& cs, int need, hipStream_t s) {
  if (need <= cs.cap) return kOk;
  CS_HIP(hipStreamSynchronize(s));
  int cap = std::max(need + need / 2, 1 << 16);
  dfree(cs.rkey, s);
  dfree(cs.rinfo, s);
  dfree(cs.rpair, s);
  dfree(cs.skey[0], s);
  dfree(cs.skey[1], s);
  dfree(cs.sidx[0], s);
  dfree(cs.sidx[1], s);
  dfree(cs.head, s);
  dfree(cs.gpos, s);
  dfree(cs.ce, s);
  dfree(cs.pred, s);
  dfree(cs.succ, s);
  dfree(cs.npred, s);
  dfree(cs.ready[0], s);
  dfree(cs.ready[1], s);
  dfree(cs.inc, s);
  dev_free(cs.tmp, s);
  cs.tmp = nullptr;
  cs.cap = 0;
  if (int st = dalloc(&cs.rkey, 4L * cap, s)) return st;
  if (int st = dalloc(&cs.rinfo, cap, s)) return st;
  if (int st = dalloc(&cs.rpair, 2L * cap, s)) return st;
  if (int st = dalloc(&cs.skey[0], cap, s)) return st;
  if (int st = dalloc(&cs.skey[1], cap, s)) return st;
  if (int st = dalloc(&cs.sidx[0], cap, s)) return st;
  if (int st = dalloc(&cs.sidx[1], cap, s)) return st;
  if (int st = dalloc(&cs.head, cap, s)) return st;
  if (int st = dalloc(&cs.gpos, cap, s)) return st;
  if (int st = dalloc(&cs.ce, 5L * cap, s)) return st;
  if (int st = dalloc(&cs.pred, 4L * cap, s)) return st;
  if (int st = dalloc(&cs.succ, 4L * cap, s)) return st;
  if (int st = dalloc(&cs.npred, cap, s)) return st;
  if (int st = dalloc(&cs.ready[0], cap, s)) return st;
  if (int st = dalloc(&cs.ready[1], cap, s)) return st;
  if (int st = dalloc(&cs.inc, 4L * cap, s)) return st;
  // temporary storage for the largest sort / scan
  size_t a = 0, b = 0, c2 = 0;
  CS_HIP(rocprim::radix_sort_pairs(nullptr, a, cs.skey[0], cs.skey[1], cs.sidx[0], cs.sidx[1],
                                   (unsigned)cap, 0, 64, s));
  CS_HIP(rocprim::exclusive_scan(nullptr, b, cs.head, cs.gpos, 0, (size_t)cap,
                                 rocprim::plus<int>(), s));
  CS_HIP(rocprim::exclusive_scan(nullptr, c2, cs.vcount, cs.vstart, 0, (size_t)cs.np + 1,
                                 rocprim::plus<int>(), s));
  cs.tmp_bytes = std::max(a, std::max(b, c2)) + 256;
  if (int st = dev_alloc(&cs.tmp, cs.tmp_bytes, s)) return st;
  cs.cap = cap;
  return kOk;
}

int grow_cand(ContactState& cs, int need, hipStream_t s) {
  if (need <= cs.cand_cap) return kOk;
  CS_HIP(hipStreamSynchronize(s));
  const int cap = std::max(need + need / 2, 1 << 16);
  dfree(cs.cand, s);
  cs.cand_cap = 0;
  if (int st = dalloc(&cs.cand, 2L * cap, s)) return st;
  cs.cand_cap = cap;
  return kOk;
}

int read_counters(ContactState& cs, hipStream_t s) {
  CS_HIP(hipMemcpyAsync(cs.hcounters, cs.counters, sizeof(int) * kNumCounters,
                        hipMemcpyDeviceToHost, s));
  CS_HIP(hipStreamSynchronize(s));
  return kOk;
}

int refit(ContactState& cs, const DevCloth& d, hipStream_t s, const double* X, const double* U,
          double delta, double dt, int ccd) {
  hipLaunchKernelGGL(k_refit_leaves, dim3(grid_for(d.nt)), dim3(kTB), 0, s, d, X, U, delta, dt,
                     ccd);
  for (size_t l = 0; l + 1 < cs.refit_lvl.size(); l++) {
    const int b = cs.refit_lvl[l], e = cs.refit_lvl[l + 1];
    if (e > b) hipLaunchKernelGGL(k_refit_merge, dim3(grid_for(e - b)), dim3(kTB), 0, s, d, b, e);
  }
  CS_HIP(hipGetLastError());
  return kOk;
}

// refit + traverse + select: the contact set into cs.ce, its size into *g
int query(ContactState& cs, const DevCloth& d, hipStream_t s, const double* X, const double* U,
          double delta, double dt, int ccd, int* g) {
  if (int st = refit(cs, d, s, X, U, delta, dt, ccd)) return st;
  int ncand = 0;
  for (int attempt = 0; attempt < 2; attempt++) {
    CS_HIP(hipMemsetAsync(cs.counters, 0, sizeof(int) * kQueryCounters, s));
    // leaves per wave: 64, halved (down to 16) while the launch would have
    // fewer than 2048 waves, so that a small tree's walks spread over more waves
    int per = 64;
    while (per > 16 && (d.nt + per - 1) / per < 2048) per /= 2;
    hipLaunchKernelGGL(k_traverse, dim3((d.nt + per - 1) / per), dim3(64), 0, s, d.nnode, d.nt, per,
                       d.leaf_rank, d.rend, d.rbb, d.by_pre, cs.counters, cs.cand_cap,
                       reinterpret_cast<int2*>(cs.cand));
    CS_HIP(hipGetLastError());
    if (int st = read_counters(cs, s)) return st;
    ncand = cs.hcounters[kCand];
    if (ncand <= cs.cand_cap) break;
    if (int st = grow_cand(cs, ncand, s)) return st;
  }
  int raw = 0;
  for (int attempt = 0; attempt < 2 && ncand > 0; attempt++) {
    CS_HIP(hipMemsetAsync(cs.counters + kRaw, 0, sizeof(int), s));
    Emit em{cs.counters, cs.cap, cs.rkey, cs.rinfo, cs.rpair};
    hipLaunchKernelGGL(k_narrow, dim3(grid_for(ncand)), dim3(kTB), 0, s, d, X, U, dt, delta, ccd,
                       ncand, cs.cand, em);
    CS_HIP(hipGetLastError());
    if (int st = read_counters(cs, s)) return st;
    raw = cs.hcounters[kRaw];
    if (raw <= cs.cap) break;
    if (int st = grow_records(cs, raw, s)) return st;
  }
  if (raw == 0) {
    *g = 0;
    return kOk;
  }
  // radix sort of the records by canonical key
  const int nb = cs.nbits;
  const int* order = nullptr;
  size_t tb = cs.tmp_bytes;
  if (4 * nb <= 64) {
    hipLaunchKernelGGL(k_make_keys, dim3(grid_for(raw)), dim3(kTB), 0, s, cs.rkey, raw, nb, 0,
                       cs.skey[0], cs.sidx[0]);
    CS_HIP(rocprim::radix_sort_pairs(cs.tmp, tb, cs.skey[0], cs.skey[1], cs.sidx[0], cs.sidx[1],
                                     (unsigned)raw, 0, 4 * nb, s));
    order = cs.sidx[1];
  } else {
    hipLaunchKernelGGL(k_make_keys, dim3(grid_for(raw)), dim3(kTB), 0, s, cs.rkey, raw, nb, 1,
                       cs.skey[0], cs.sidx[0]);
    CS_HIP(rocprim::radix_sort_pairs(cs.tmp, tb, cs.skey[0], cs.skey[1], cs.sidx[0], cs.sidx[1],
                                     (unsigned)raw, 0, nb, s));
    hipLaunchKernelGGL(k_rekey, dim3(grid_for(raw)), dim3(kTB), 0, s, cs.rkey, cs.sidx[1], raw, nb,
                       cs.skey[0]);
    tb = cs.tmp_bytes;
    CS_HIP(rocprim::radix_sort_pairs(cs.tmp, tb, cs.skey[0], cs.skey[1], cs.sidx[1], cs.sidx[0],
                                     (unsigned)raw, 0, 3 * nb, s));
    order = cs.sidx[0];
  }
  hipLaunchKernelGGL(k_heads, dim3(grid_for(raw)), dim3(kTB), 0, s, cs.rkey, order, raw, cs.head);
  tb = cs.tmp_bytes;
  CS_HIP(rocprim::exclusive_scan(cs.tmp, tb, cs.head, cs.gpos, 0, (size_t)raw,
                                 rocprim::plus<int>(), s));
  hipLaunchKernelGGL(k_select, dim3(grid_for(raw)), dim3(kTB), 0, s, d, cs.rkey, cs.rinfo,
                     cs.rpair, order, cs.head, cs.gpos, raw, cs.ce, cs.counters);
  CS_HIP(hipGetLastError());
  if (int st = read_counters(cs, s)) return st;
  *g = cs.hcounters[kGroups];
  return kOk;
}

// Gauss-Seidel impulses over cs.ce[0..g) in key order
int impulses(ContactState& cs, const DevCloth& d, const Params& P, hipStream_t s, const double* X,
             double* U, int g, int ccd) {
  if (g == 0) return kOk;
  CS_HIP(hipMemsetAsync(cs.vcount, 0, sizeof(int) * (d.np + 1), s));
  CS_HIP(hipMemsetAsync(cs.vfill, 0, sizeof(int) * (d.np + 1), s));
  hipLaunchKernelGGL(k_inc_count, dim3(grid_for(g)), dim3(kTB), 0, s, cs.ce, g, cs.vcount);
  size_t tb = cs.tmp_bytes;
  CS_HIP(rocprim::exclusive_scan(cs.tmp, tb, cs.vcount, cs.vstart, 0, (size_t)d.np + 1,
                                 rocprim::plus<int>(), s));
  hipLaunchKernelGGL(k_inc_fill, dim3(grid_for(g)), dim3(kTB), 0, s, cs.ce, g, cs.vstart, cs.vfill,
                     cs.inc);
  hipLaunchKernelGGL(k_inc_link, dim3(grid_for(d.np)), dim3(kTB), 0, s, d.np, cs.vstart, cs.inc,
                     cs.pred, cs.succ);
  CS_HIP(hipMemsetAsync(cs.counters + kDoneCnt, 0, sizeof(int), s));
  CS_HIP(hipMemsetAsync(cs.counters + kReady, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_impulse_ready, dim3(grid_for(g)), dim3(kTB), 0, s, g, cs.pred, cs.npred,
                     cs.ready[0], cs.counters);
  int lds_cap = kDrainLds;
  if (const char* e = getenv("CLOTHSIM_DRAIN_LDS")) lds_cap = std::min(std::max(atoi(e), 1), kDrainLds);
  hipLaunchKernelGGL(k_impulse_drain, dim3(1), dim3(1024), 0, s, cs.ce, cs.succ, cs.npred,
                     cs.ready[0], cs.ready[1], X, U, P.clearance, P.k_contact, P.dt, P.mass, ccd,
                     cs.counters, lds_cap);
  CS_HIP(hipGetLastError());
  if (int st = read_counters(cs, s)) return st;
  if (cs.hcounters[kDoneCnt] != g) {
    set_error("impulse schedule did not drain");
    return kHipError;
  }
  return kOk;
}

}  // namespace

int contact_init(ContactState& cs, const MeshSetup& S, const DevCloth& d, hipStream_t s) {
  cs.np = S.np;
  cs.nt = S.nt;
  cs.nnode = S.nnode;
  int nb = 1;
  while ((1 << nb) < S.np) nb++;
  cs.nbits = nb;
  if (3 * nb > 64) {
    set_error("mesh too large for the contact keys");
    return kBadArg;
  }
  if (int st = dalloc(&cs.vcount, S.np + 1, s)) return st;
  if (int st = dalloc(&cs.vstart, S.np + 1, s)) return st;
  if (int st = dalloc(&cs.vfill, S.np + 1, s)) return st;
  if (int st = dalloc(&cs.counters, kNumCounters, s)) return st;
  cs.refit_lvl = S.refit_lvl_ptr;
  if (int st = dalloc(&cs.zpar, S.np, s)) return st;
  if (int st = dalloc(&cs.zroot, S.np, s)) return st;
  if (int st = dalloc(&cs.zmemb, S.np, s)) return st;
  if (int st = dalloc(&cs.zroots, S.np, s)) return st;
  if (int st = dalloc(&cs.u0, 3L * S.np, s)) return st;
  if (int st = dalloc(&cs.qx, 3L * S.np, s)) return st;
  if (int st = dalloc(&cs.qv, 3L * S.np, s)) return st;
  if (int st = pinned_alloc((void**)&cs.hcounters, sizeof(int) * kNumCounters)) return st;
  CS_HIP(hipMemsetAsync(cs.counters, 0, sizeof(int) * kNumCounters, s));
  if (int st = grow_cand(cs, std::max(8 * S.nt, 1 << 16), s)) return st;
  return grow_records(cs, std::max(8 * S.nt, 1 << 16), s);
}

void contact_release(ContactState& cs, hipStream_t s) {
  dfree(cs.cand, s);
  cs.cand_cap = 0;
  dfree(cs.rkey, s);
  dfree(cs.rinfo, s);
  dfree(cs.rpair, s);
  dfree(cs.skey[0], s);
  dfree(cs.skey[1], s);
  dfree(cs.sidx[0], s);
  dfree(cs.sidx[1], s);
  dfree(cs.head, s);
  dfree(cs.gpos, s);
  dfree(cs.ce, s);
  dfree(cs.pred, s);
  dfree(cs.succ, s);
  dfree(cs.npred, s);
  dfree(cs.ready[0], s);
  dfree(cs.ready[1], s);
  dfree(cs.inc, s);
  dfree(cs.vcount, s);
  dfree(cs.vstart, s);
  dfree(cs.vfill, s);
  dfree(cs.counters, s);
  dfree(cs.zpar, s);
  dfree(cs.zroot, s);
  dfree(cs.zmemb, s);
  dfree(cs.zroots, s);
  dfree(cs.u0, s);
  dfree(cs.qx, s);
  dfree(cs.qv, s);
  dev_free(cs.tmp, s);
  cs.tmp = nullptr;
  pinned_free(cs.hcounters);
  cs.hcounters = nullptr;
}

int contact_resolve(ContactState& cs, const MeshSetup& S, const DevCloth& d, const Params& P,
                    hipStream_t s, int* applied, cs_step_stats* stats, Prof* prof) {
  (void)S;
  ProfScope ps(prof, kProfContact);
  const double* X = d.x1;
  double* U = d.v;
  int g = 0;
  if (stats) {
    stats->n_prox = 0;
    stats->n_ccd_pass = 0;
    for (int k = 0; k < 5; k++) stats->n_ccd[k] = 0;
    stats->n_riz_pass = 0;
    stats->n_riz_vertices = 0;
  }
  // one proximity pass
  if (int st = query(cs, d, s, X, U, P.clearance, P.dt, 0, &g)) return st;
  *applied = g > 0;
  if (stats) stats->n_prox = g;
  if (int st = impulses(cs, d, P, s, X, U, g, 0)) return st;
  // up to five continuous passes
  for (int it = 0; it < 5; it++) {
    if (int st = query(cs, d, s, X, U, P.clearance, P.dt, 1, &g)) return st;
    if (g == 0) return kOk;
    *applied = 1;
    if (stats) {
      stats->n_ccd[it] = g;
      stats->n_ccd_pass = it + 1;
    }
    if (int st = impulses(cs, d, P, s, X, U, g, 1)) return st;
  }
  // rigid impact zones, up to 100 passes
  CS_HIP(hipMemcpyAsync(cs.u0, U, sizeof(double) * 3 * d.np, hipMemcpyDeviceToDevice, s));
  CS_HIP(hipMemsetAsync(cs.zpar, 0xff, sizeof(int) * d.np, s));
  CS_HIP(hipMemsetAsync(cs.counters + kZoneVerts, 0, sizeof(int), s));
  for (int it = 0; it < 100; it++) {
    if (int st = query(cs, d, s, X, U, P.clearance, P.dt, 1, &g)) return st;
    if (g == 0) break;
    hipLaunchKernelGGL(k_riz_pass, dim3(1), dim3(1024), 0, s, d, cs.ce, g, cs.zpar, cs.zroot,
                       cs.zmemb, cs.zroots, X, (const double*)cs.u0, U, cs.counters);
    CS_HIP(hipGetLastError());
    if (stats) stats->n_riz_pass = it + 1;
  }
  if (stats && stats->n_riz_pass > 0) {
    if (int st = read_counters(cs, s)) return st;
    stats->n_riz_vertices = cs.hcounters[kZoneVerts];
  }
  return kOk;
}

int update_velocity(const DevCloth& d, const Params& P, hipStream_t s) {
  if (int st = launch_gradients(d, P, s)) return st;
  hipLaunchKernelGGL(k_velocity_update, dim3(grid_for(d.np)), dim3(kTB), 0, s, d, P);
  CS_HIP(hipGetLastError());
  return kOk;
}

int post_contact_update(const DevCloth& d, const Params& P, hipStream_t s) {
  hipLaunchKernelGGL(k_advance, dim3(grid_for(3L * d.np)), dim3(kTB), 0, s, d, P.dt);
  CS_HIP(hipGetLastError());
  return update_velocity(d, P, s);
}

int contact_refit_host(ContactState& cs, const MeshSetup& S, const DevCloth& d, hipStream_t s,
                       const double* xyz, const double* uvw, double delta, double dt, double* bb) {
  const size_t n = sizeof(double) * 3 * d.np;
  CS_HIP(hipMemcpyAsync(cs.qx, xyz, n, hipMemcpyHostToDevice, s));
  if (uvw) CS_HIP(hipMemcpyAsync(cs.qv, uvw, n, hipMemcpyHostToDevice, s));
  if (int st = refit(cs, d, s, cs.qx, cs.qv, delta, dt, uvw ? 1 : 0)) return st;
  if (bb) CS_HIP(hipMemcpyAsync(bb, d.bb, sizeof(double) * 6 * S.nnode, hipMemcpyDeviceToHost, s));
  CS_HIP(hipStreamSynchronize(s));
  return kOk;
}

int contact_query_host(ContactState& cs, const MeshSetup& S, const DevCloth& d, hipStream_t s,
                       const double* xyz, const double* uvw, double delta, double dt, int* out,
                       int cap, int* n) {
  (void)S;
  const size_t bytes = sizeof(double) * 3 * d.np;
  CS_HIP(hipMemcpyAsync(cs.qx, xyz, bytes, hipMemcpyHostToDevice, s));
  if (uvw) CS_HIP(hipMemcpyAsync(cs.qv, uvw, bytes, hipMemcpyHostToDevice, s));
  int g = 0;
  if (int st = query(cs, d, s, cs.qx, cs.qv, delta, dt, uvw ? 1 : 0, &g)) return st;
  if (n) *n = g;
  const int m = std::min(g, cap);
  if (out && m > 0)
    CS_HIP(hipMemcpyAsync(out, cs.ce, sizeof(int) * 5 * m, hipMemcpyDeviceToHost, s));
  CS_HIP(hipStreamSynchronize(s));
  return kOk;
}

}  // namespace cs
