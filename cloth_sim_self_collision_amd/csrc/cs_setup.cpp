// Host-side, per-mesh setup of the cloth step (see cs_setup.hpp).
#include "cs_setup.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <string>
#include <vector>

namespace cs {
namespace {

// vertex -> incident elements, in element order
void incidence(const int* elem, int ne, int nn, int np, std::vector<int>& ptr,
               std::vector<int>& idx) {
  ptr.assign(np + 1, 0);
  for (int e = 0; e < ne; e++)
    for (int k = 0; k < nn; k++) ptr[elem[e * nn + k] + 1]++;
  for (int i = 0; i < np; i++) ptr[i + 1] += ptr[i];
  idx.resize(ptr[np]);
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int e = 0; e < ne; e++)
    for (int k = 0; k < nn; k++) idx[fill[elem[e * nn + k]]++] = e;
}

// vertices sharing an element with each vertex, self excluded, in discovery
// order over the incident elements (jagged_array.h:158-207, is_dia=false)
void neighbours(const int* elem, int ne, int nn, int np, std::vector<int>& ptr,
                std::vector<int>& out) {
  std::vector<int> vptr, vidx;
  incidence(elem, ne, nn, np, vptr, vidx);
  std::vector<int> seen(np, -1);
  ptr.assign(np + 1, 0);
  out.clear();
  for (int v = 0; v < np; v++) {
    seen[v] = v;
    for (int k = vptr[v]; k < vptr[v + 1]; k++) {
      const int e = vidx[k];
      for (int a = 0; a < nn; a++) {
        const int w = elem[e * nn + a];
        if (seen[w] != v) {
          seen[w] = v;
          out.push_back(w);
        }
      }
    }
    ptr[v + 1] = (int)out.size();
  }
}

// ---- top-down BVH whose node numbering follows the reference's creation
// order: children are created in pairs, child 0 is the flood-filled side
// (bvh_aabb.cpp:109-313)

struct Centroid {
  double x, y, z;
};

struct BvhBuilder {
  const int* tri;
  const double* xyz;
  std::vector<int> adj;    // 3 per triangle, -1 on boundary
  std::vector<int> owner;  // node currently owning each triangle
  std::vector<int>& node;
  std::vector<int> stack;

  BvhBuilder(std::vector<int>& n) : node(n) {}

  Centroid centroid(int t) const {
    const int a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
    Centroid g;
    g.x = (xyz[3 * a] + xyz[3 * b] + xyz[3 * c]) / 3.0;
    g.y = (xyz[3 * a + 1] + xyz[3 * b + 1] + xyz[3 * c + 1]) / 3.0;
    g.z = (xyz[3 * a + 2] + xyz[3 * b + 2] + xyz[3 * c + 2]) / 3.0;
    return g;
  }
  static double proj(const Centroid& g, const double o[3], const double d[3]) {
    return (g.x - o[0]) * d[0] + (g.y - o[1]) * d[1] + (g.z - o[2]) * d[2];
  }

  // first triangle (in list order) off the splitting plane; orients d towards it
  int pick_seed(const std::vector<int>& list, const double o[3], double d[3]) const {
    for (int t : list) {
      const Centroid g = centroid(t);
      if (std::fabs(proj(g, o, d)) < 1.0e-10) continue;
      if (proj(g, o, d) < 0) {
        d[0] *= -1;
        d[1] *= -1;
        d[2] *= -1;
      }
      return t;
    }
    return -1;
  }

  bool split(int parent, const std::vector<int>& list) {
    const double eps = 1.0e-10;
    // box of the centroids, grown point by point with margin eps (aabb.h:102-122)
    double lo[3], hi[3];
    for (size_t i = 0; i < list.size(); i++) {
      const Centroid g = centroid(list[i]);
      const double p[3] = {g.x, g.y, g.z};
      for (int k = 0; k < 3; k++) {
        if (i == 0) {
          lo[k] = p[k] - eps;
          hi[k] = p[k] + eps;
        } else {
          lo[k] = (lo[k] < p[k] - eps) ? lo[k] : p[k] - eps;
          hi[k] = (hi[k] > p[k] + eps) ? hi[k] : p[k] + eps;
        }
      }
    }
    const double lx = hi[0] - lo[0], ly = hi[1] - lo[1], lz = hi[2] - lo[2];
    double d[3] = {0, 0, 0};  // strictly longest axis; stays zero on a tie
    if (lx > ly && lx > lz) d[0] = 1;
    if (ly > lz && ly > lx) d[1] = 1;
    if (lz > lx && lz > ly) d[2] = 1;
    double o[3] = {(lo[0] + hi[0]) * 0.5, (lo[1] + hi[1]) * 0.5, (lo[2] + hi[2]) * 0.5};
    int seed = pick_seed(list, o, d);
    if (seed < 0) {
      // principal axis of the centroid scatter: 10 power iterations from (1,1,1)
      double m[3] = {0, 0, 0};
      for (int t : list) {
        const Centroid g = centroid(t);
        m[0] += g.x;
        m[1] += g.y;
        m[2] += g.z;
      }
      const double n = (double)list.size();
      if (!(std::fabs(n) < 1.e-16)) {
        m[0] /= n;
        m[1] /= n;
        m[2] /= n;
      }
      o[0] = m[0];
      o[1] = m[1];
      o[2] = m[2];
      double cov[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int t : list) {
        const Centroid g = centroid(t);
        const double r[3] = {g.x - o[0], g.y - o[1], g.z - o[2]};
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) cov[i][j] += r[i] * r[j];
      }
      d[0] = d[1] = d[2] = 1;
      for (int it = 0; it < 10; it++) {
        double t[3];
        for (int i = 0; i < 3; i++) t[i] = cov[i][0] * d[0] + cov[i][1] * d[1] + cov[i][2] * d[2];
        const double len = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        for (int i = 0; i < 3; i++) d[i] = t[i] / len;
      }
      seed = pick_seed(list, o, d);
      if (seed < 0) return false;
    }
    const int c0 = (int)node.size() / 3;
    const int c1 = c0 + 1;
    node.insert(node.end(), {parent, 0, 0, parent, 0, 0});
    node[3 * parent + 1] = c0;
    node[3 * parent + 2] = c1;
    // child 0: LIFO flood fill from the seed over adjacent triangles that are
    // still owned by the parent and lie on the non-negative side
    std::vector<int> l0;
    owner[seed] = c0;
    l0.push_back(seed);
    stack.clear();
    stack.push_back(seed);
    while (!stack.empty()) {
      const int t = stack.back();
      stack.pop_back();
      for (int e = 0; e < 3; e++) {
        const int u = adj[3 * t + e];
        if (u == -1 || owner[u] != parent) continue;
        if (proj(centroid(u), o, d) < 0) continue;
        stack.push_back(u);
        owner[u] = c0;
        l0.push_back(u);
      }
    }
    std::vector<int> l1;
    for (int t : list) {
      if (owner[t] == c0) continue;
      owner[t] = c1;
      l1.push_back(t);
    }
    if (l1.empty()) return false;
    if (l0.size() == 1) {
      node[3 * c0 + 1] = l0[0];
      node[3 * c0 + 2] = -1;
    } else if (!split(c0, l0)) {
      return false;
    }
    l0.clear();
    l0.shrink_to_fit();
    if (l1.size() == 1) {
      node[3 * c1 + 1] = l1[0];
      node[3 * c1 + 2] = -1;
    } else if (!split(c1, l1)) {
      return false;
    }
    return true;
  }
};

// triangle adjacency across oppositely oriented shared edges, first match in
// incidence order (bvh_aabb.cpp:109-138)
void triangle_adjacency(const int* tri, int nt, int np, std::vector<int>& adj) {
  static const int ev[3][2] = {{1, 2}, {2, 0}, {0, 1}};
  std::vector<int> vptr, vidx;
  incidence(tri, nt, 3, np, vptr, vidx);
  adj.assign(3 * nt, -1);
  for (int t = 0; t < nt; t++)
    for (int e = 0; e < 3; e++) {
      const int a = tri[3 * t + ev[e][0]], b = tri[3 * t + ev[e][1]];
      for (int k = vptr[a]; k < vptr[a + 1] && adj[3 * t + e] == -1; k++) {
        const int u = vidx[k];
        if (u == t) continue;
        for (int f = 0; f < 3; f++)
          if (tri[3 * u + ev[f][1]] == a && tri[3 * u + ev[f][0]] == b) {
            adj[3 * t + e] = u;
            break;
          }
      }
    }
}

void levels_to_lists(const std::vector<int>& lvl, int nlev, std::vector<int>& ptr,
                     std::vector<int>& rows, int* maxw) {
  ptr.assign(nlev + 1, 0);
  for (int l : lvl) ptr[l + 1]++;
  for (int i = 0; i < nlev; i++) ptr[i + 1] += ptr[i];
  rows.resize(lvl.size());
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int i = 0; i < (int)lvl.size(); i++) rows[fill[lvl[i]]++] = i;
  *maxw = 0;
  for (int i = 0; i < nlev; i++) *maxw = std::max(*maxw, ptr[i + 1] - ptr[i]);
}

// Live interval of a row in a sweep: [level written, last level read]; rows
// of one component whose sweep slots share a ring position must have disjoint
// intervals (components are swept by separate workgroups, each with its own
// ring).  Slots run in level order within a component, so the ring is a FIFO
// whose window only has to cover the rows live across a few levels.
bool ring_fits(const MeshSetup& s, const std::vector<int>& lvl, bool lower,
               const std::vector<int>& slot_of) {
  const int np = s.np;
  std::vector<int> last(np);
  for (int i = 0; i < np; i++) last[i] = lvl[i];
  for (int i = 0; i < np; i++) {
    const int b = lower ? s.rowptr[i] : s.diaind[i];
    const int e = lower ? s.diaind[i] : s.rowptr[i + 1];
    for (int k = b; k < e; k++) last[s.colind[k]] = std::max(last[s.colind[k]], lvl[i]);
  }
  // (component, ring entry, first level, last level) per row
  std::vector<std::array<int, 4>> iv(np);
  for (int i = 0; i < np; i++) iv[i] = {s.comp[i], slot_of[i] & (kRingRows - 1), lvl[i], last[i]};
  std::sort(iv.begin(), iv.end());
  for (int q = 1; q < np; q++)
    if (iv[q][0] == iv[q - 1][0] && iv[q][1] == iv[q - 1][1] && iv[q][2] <= iv[q - 1][3])
      return false;
  return true;
}

}  // namespace

std::string build_setup(MeshSetup& s, int np, int nt, int nq, const int* tri, const int* quad,
                        const double* xyz_init) {
  if (np <= 0 || nt <= 1 || nq < 0) return "mesh needs np > 0, nt > 1, nq >= 0";
  for (int i = 0; i < 3 * nt; i++)
    if (tri[i] < 0 || tri[i] >= np) return "triangle index out of range";
  for (int i = 0; i < 4 * nq; i++)
    if (quad[i] < 0 || quad[i] >= np) return "bending stencil index out of range";
  for (int t = 0; t < nt; t++) {
    const int a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
    if (a == b || b == c || a == c) return "degenerate triangle (repeated vertex)";
  }
  s.np = np;
  s.nt = nt;
  s.nq = nq;
  // pattern: union of the bending stencils, columns ascending
  neighbours(quad, nq, 4, np, s.rowptr, s.colind);
  for (int i = 0; i < np; i++) std::sort(s.colind.begin() + s.rowptr[i], s.colind.begin() + s.rowptr[i + 1]);
  s.nnz = s.rowptr[np];
  s.rowof.resize(s.nnz);
  s.diaind.resize(np);
  for (int i = 0; i < np; i++) {
    s.diaind[i] = s.rowptr[i + 1];
    for (int k = s.rowptr[i]; k < s.rowptr[i + 1]; k++) {
      s.rowof[k] = i;
      if (s.colind[k] > i && s.diaind[i] == s.rowptr[i + 1]) s.diaind[i] = k;
    }
  }
  // ILU(0) update lists (the pattern is fixed, so the column searches of
  // the factorisation are done once here)
  s.fu_ptr.assign(np + 1, 0);
  s.fu_kj.clear();
  s.fu_code.clear();
  for (int i = 0; i < np; i++) {
    const int rb = s.rowptr[i], re = s.rowptr[i + 1];
    for (int ik = rb; ik < s.diaind[i]; ik++) {
      const int kr = s.colind[ik];
      for (int kj = s.diaind[kr]; kj < s.rowptr[kr + 1]; kj++) {
        const int j = s.colind[kj];
        int tgt = kFuDiag;
        if (j != i) {
          const int* p = std::lower_bound(s.colind.data() + rb, s.colind.data() + re, j);
          if (p == s.colind.data() + re || *p != j) continue;  // no fill-in
          tgt = (int)(p - (s.colind.data() + rb));
        }
        s.fu_kj.push_back(kj);
        s.fu_code.push_back((ik - rb) | (tgt << 8));
      }
    }
    s.fu_ptr[i + 1] = (int)s.fu_kj.size();
  }
  // level schedules
  {
    std::vector<int> lf(np, 0), lb(np, 0);
    int nf = 0, nb = 0;
    for (int i = 0; i < np; i++) {
      int l = 0;
      for (int k = s.rowptr[i]; k < s.diaind[i]; k++) l = std::max(l, lf[s.colind[k]] + 1);
      lf[i] = l;
      nf = std::max(nf, l + 1);
    }
    for (int i = np - 1; i >= 0; i--) {
      int l = 0;
      for (int k = s.diaind[i]; k < s.rowptr[i + 1]; k++) l = std::max(l, lb[s.colind[k]] + 1);
      lb[i] = l;
      nb = std::max(nb, l + 1);
    }
    s.nlf = nf;
    s.nlb = nb;
    levels_to_lists(lf, nf, s.lf_ptr, s.lf_rows, &s.max_lf);
    levels_to_lists(lb, nb, s.lb_ptr, s.lb_rows, &s.max_lb);
    // components (union-find over the pattern), numbered by smallest row
    std::vector<int> up(np);
    for (int i = 0; i < np; i++) up[i] = i;
    auto find = [&up](int i) {
      while (up[i] != i) {
        up[i] = up[up[i]];
        i = up[i];
      }
      return i;
    };
    for (int i = 0; i < np; i++)
      for (int k = s.rowptr[i]; k < s.rowptr[i + 1]; k++) {
        const int a = find(i), b = find(s.colind[k]);
        if (a != b) up[std::max(a, b)] = std::min(a, b);
      }
    s.comp.assign(np, -1);
    std::vector<int> id_of_root(np, -1);
    s.ncomp = 0;
    for (int i = 0; i < np; i++) {
      const int r = find(i);
      if (id_of_root[r] < 0) id_of_root[r] = s.ncomp++;
      s.comp[i] = id_of_root[r];
    }
    // quad sweep slot orders: the level lists regrouped by component (stable,
    // so each component keeps level order and row order within a level)
    auto by_comp = [&](const std::vector<int>& rows, const std::vector<int>& lvl,
                       std::vector<int>& out, std::vector<int>& q, std::vector<int>& qc) {
      out.assign(np, 0);
      std::vector<int> start(s.ncomp + 1, 0);
      for (int i = 0; i < np; i++) start[s.comp[i] + 1]++;
      for (int c = 0; c < s.ncomp; c++) start[c + 1] += start[c];
      std::vector<int> fill(start.begin(), start.end() - 1);
      for (int k = 0; k < np; k++) out[fill[s.comp[rows[k]]]++] = rows[k];
      q.assign(1, 0);
      qc.assign(1, 0);
      for (int c = 0; c < s.ncomp; c++) {
        for (int b = start[c]; b < start[c + 1];) {
          int e = b + 1;  // the rest of this level in this component, at most kQuadRows rows
          while (e < start[c + 1] && e - b < kQuadRows && lvl[out[e]] == lvl[out[b]]) e++;
          q.push_back(e);
          s.qwidth = std::max(s.qwidth, e - b);
          b = e;
        }
        qc.push_back((int)q.size() - 1);
      }
    };
    s.qwidth = 0;
    by_comp(s.lf_rows, lf, s.sf_rows, s.qf_ptr, s.qf_comp);
    by_comp(s.lb_rows, lb, s.sb_rows, s.qb_ptr, s.qb_comp);
    s.nqf = (int)s.qf_ptr.size() - 1;
    s.nqb = (int)s.qb_ptr.size() - 1;
  }
  // packed sweep layout and the ring-buffer safety check
  {
    s.kl = s.ku = 0;
    for (int i = 0; i < np; i++) {
      s.kl = std::max(s.kl, s.diaind[i] - s.rowptr[i]);
      s.ku = std::max(s.ku, s.rowptr[i + 1] - s.diaind[i]);
    }
    if (s.kl <= kMaxBlocksPerRow && s.ku <= kMaxBlocksPerRow) {
      s.f_rci.assign(4 * (size_t)np, 0);
      s.b_rc.assign(2 * (size_t)np, 0);
      s.fslot.assign(np, 0);
      std::vector<int> bslot(np);
      for (int sl = 0; sl < np; sl++) {
        s.fslot[s.sf_rows[sl]] = sl;
        bslot[s.sb_rows[sl]] = sl;
      }
      const int kq = std::max(std::max(s.kl, s.ku), 1);  // both sweeps padded alike
      const size_t mw = (size_t)meta_groups(kq) * 8;    // metadata words per slot
      s.f_meta.assign(mw * np, 0);
      s.b_meta.assign(mw * np, 0);
      auto midx = [np](int h, int sl) { return ((size_t)(h / 8) * np + sl) * 8 + h % 8; };
      // byte offset of a slot's entry in a row of the sweep's LDS ring
      auto ring_byte = [](int slot) {
        return (unsigned short)((slot & (kRingRows - 1)) * (int)sizeof(double));
      };
      const unsigned short zero_byte = (unsigned short)(kRingZero * (int)sizeof(double));
      auto put32 = [&](std::vector<unsigned short>& m, int sl, unsigned v) {
        m[midx(kq, sl)] = (unsigned short)(v & 0xffffu);
        m[midx(kq + 1, sl)] = (unsigned short)(v >> 16);
      };
      for (int sl = 0; sl < np; sl++) {
        const int i = s.sf_rows[sl];
        const int fc = s.diaind[i] - s.rowptr[i];
        s.f_rci[4 * sl] = i;
        s.f_rci[4 * sl + 1] = fc;
        s.f_rci[4 * sl + 2] = bslot[i];
        for (int k = 0; k < kq; k++)
          s.f_meta[midx(k, sl)] = k < fc ? ring_byte(s.fslot[s.colind[s.rowptr[i] + k]]) : zero_byte;
        put32(s.f_meta, sl, (unsigned)bslot[i]);
        const int j = s.sb_rows[sl];
        const int bc = s.rowptr[j + 1] - s.diaind[j];
        s.b_rc[2 * sl] = j;
        s.b_rc[2 * sl + 1] = bc;
        for (int k = 0; k < kq; k++)
          s.b_meta[midx(k, sl)] = k < bc ? ring_byte(bslot[s.colind[s.diaind[j] + k]]) : zero_byte;
        put32(s.b_meta, sl, (unsigned)j);
      }
      std::vector<int> lf(np), lb(np);
      for (int l = 0; l < s.nlf; l++)
        for (int q = s.lf_ptr[l]; q < s.lf_ptr[l + 1]; q++) lf[s.lf_rows[q]] = l;
      for (int l = 0; l < s.nlb; l++)
        for (int q = s.lb_ptr[l]; q < s.lb_ptr[l + 1]; q++) lb[s.lb_rows[q]] = l;
      s.ring_ok = ring_fits(s, lf, true, s.fslot) && ring_fits(s, lb, false, bslot);
    } else {
      s.kl = s.ku = 0;
    }
  }
  // gather lists: sources of every Hessian block and gradient entry in the
  // order the reference's element loops add them (solve_internal_sparse.h:19-76)
  {
    std::vector<std::vector<int>> crs(s.nnz), dia(np), grad(np);
    auto find_col = [&](int i, int j) -> int {
      const int* b = s.colind.data() + s.rowptr[i];
      const int* e = s.colind.data() + s.rowptr[i + 1];
      const int* f = std::lower_bound(b, e, j);
      return (f != e && *f == j) ? (int)(f - s.colind.data()) : -1;
    };
    for (int t = 0; t < nt; t++)
      for (int a = 0; a < 3; a++) {
        const int i = tri[3 * t + a];
        grad[i].push_back(t * 3 + a);
        for (int b = 0; b < 3; b++) {
          if (a == b) {
            dia[i].push_back(t * 9 + a * 3 + b);
          } else {
            const int k = find_col(i, tri[3 * t + b]);
            if (k >= 0) crs[k].push_back(t * 9 + a * 3 + b);
          }
        }
      }
    for (int q = 0; q < nq; q++)
      for (int a = 0; a < 4; a++) {
        const int i = quad[4 * q + a];
        grad[i].push_back(-(q * 4 + a) - 1);
        for (int b = 0; b < 4; b++) {
          if (a == b) {
            dia[i].push_back(-(q * 16 + a * 4 + b) - 1);
          } else {
            const int k = find_col(i, quad[4 * q + b]);
            if (k >= 0) crs[k].push_back(-(q * 16 + a * 4 + b) - 1);
          }
        }
      }
    auto flatten = [](std::vector<std::vector<int>>& v, std::vector<int>& ptr,
                      std::vector<int>& out) {
      ptr.assign(v.size() + 1, 0);
      for (size_t i = 0; i < v.size(); i++) ptr[i + 1] = ptr[i] + (int)v[i].size();
      out.resize(ptr.back());
      for (size_t i = 0; i < v.size(); i++)
        std::copy(v[i].begin(), v[i].end(), out.begin() + ptr[i]);
    };
    flatten(crs, s.crs_src_ptr, s.crs_src);
    flatten(dia, s.dia_src_ptr, s.dia_src);
    flatten(grad, s.grad_src_ptr, s.grad_src);
  }
  neighbours(tri, nt, 3, np, s.ring_ptr, s.ring);
  // BVH topology
  {
    s.node.clear();
    s.node.insert(s.node.end(), {-1, 0, 0});
    BvhBuilder B(s.node);
    B.tri = tri;
    B.xyz = xyz_init;
    triangle_adjacency(tri, nt, np, B.adj);
    B.owner.assign(nt, 0);
    std::vector<int> all(nt);
    for (int t = 0; t < nt; t++) all[t] = t;
    if (!B.split(0, all)) return "BVH split failed (degenerate triangle cluster)";
    s.nnode = (int)s.node.size() / 3;
    s.root = 0;
  }
  // pre-order ranks, subtree extents, depths, leaf of each triangle
  {
    const int n = s.nnode;
    s.pre.assign(n, 0);
    s.pre_end.assign(n, 0);
    s.by_pre.assign(n, 0);
    s.depth.assign(n, 0);
    s.leaf_of.assign(nt, -1);
    std::vector<int> st;
    st.push_back(s.root);
    int r = 0;
    while (!st.empty()) {
      const int v = st.back();
      st.pop_back();
      s.pre[v] = r;
      s.by_pre[r] = v;
      r++;
      const int c0 = s.node[3 * v + 1], c1 = s.node[3 * v + 2];
      if (c1 == -1) {
        s.leaf_of[c0] = v;
      } else {
        s.depth[c0] = s.depth[c1] = s.depth[v] + 1;
        st.push_back(c1);
        st.push_back(c0);
      }
    }
    // subtree end: pre + size, sizes accumulated children-first (reverse pre-order)
    std::vector<int> size(n, 1);
    for (int k = n - 1; k >= 0; k--) {
      const int v = s.by_pre[k];
      if (s.node[3 * v + 2] != -1) size[v] = 1 + size[s.node[3 * v + 1]] + size[s.node[3 * v + 2]];
    }
    for (int v = 0; v < n; v++) s.pre_end[v] = s.pre[v] + size[v];
    for (int t = 0; t < nt; t++)
      if (s.leaf_of[t] < 0) return "triangle without a BVH leaf";
    s.max_depth = 0;
    for (int v = 0; v < n; v++) s.max_depth = std::max(s.max_depth, s.depth[v]);
    // bottom-up refit order: deepest level first
    s.refit_lvl_ptr.assign(s.max_depth + 2, 0);
    for (int v = 0; v < n; v++) s.refit_lvl_ptr[s.max_depth - s.depth[v] + 1]++;
    for (int l = 0; l <= s.max_depth; l++) s.refit_lvl_ptr[l + 1] += s.refit_lvl_ptr[l];
    s.refit_order.resize(n);
    std::vector<int> fill(s.refit_lvl_ptr.begin(), s.refit_lvl_ptr.end() - 1);
    for (int v = 0; v < n; v++) s.refit_order[fill[s.max_depth - s.depth[v]]++] = v;
  }
  return "";
}

// ---------------------------------------------------------------------------
// structural validation of a MeshSetup against its mesh

namespace {

bool is_perm(const std::vector<int>& v, int n) {
  if ((int)v.size() != n) return false;
  std::vector<char> seen(n, 0);
  for (int x : v) {
    if (x < 0 || x >= n || seen[x]) return false;
    seen[x] = 1;
  }
  return true;
}

bool is_ptr(const std::vector<int>& p, int n, int total) {
  if ((int)p.size() != n + 1 || p[0] != 0 || p[n] != total) return false;
  for (int i = 0; i < n; i++)
    if (p[i + 1] < p[i]) return false;
  return true;
}

// level of each row from a level list; false unless it lists every row once
bool level_of(const std::vector<int>& ptr, const std::vector<int>& rows, int nl, int np,
              std::vector<int>& lvl) {
  if (!is_ptr(ptr, nl, np) || !is_perm(rows, np)) return false;
  lvl.assign(np, -1);
  for (int l = 0; l < nl; l++)
    for (int q = ptr[l]; q < ptr[l + 1]; q++) lvl[rows[q]] = l;
  return true;
}

// slot order `rows` cut by `q` into sub-levels of one level and one
// component, at most kQuadRows rows each; component c's sub-levels are
// qc[c] .. qc[c + 1]
bool subs_ok(const MeshSetup& s, const std::vector<int>& rows, const std::vector<int>& q,
             const std::vector<int>& qc, const std::vector<int>& lvl, int nsub) {
  if (!is_perm(rows, s.np) || !is_ptr(q, nsub, s.np) || !is_ptr(qc, s.ncomp, nsub)) return false;
  for (int c = 0; c < s.ncomp; c++)
    for (int l = qc[c]; l < qc[c + 1]; l++) {
      if (q[l + 1] - q[l] < 1 || q[l + 1] - q[l] > kQuadRows || q[l + 1] - q[l] > s.qwidth)
        return false;
      for (int k = q[l]; k < q[l + 1]; k++)
        if (s.comp[rows[k]] != c || lvl[rows[k]] != lvl[rows[q[l]]]) return false;
      if (l > qc[c] && lvl[rows[q[l]]] < lvl[rows[q[l - 1]]]) return false;
    }
  return true;
}

bool src_ok(const std::vector<int>& ptr, const std::vector<int>& src, int n, long long tri_lim,
            long long quad_lim) {
  if (!is_ptr(ptr, n, (int)src.size())) return false;
  for (int v : src)
    if (v >= 0 ? (long long)v >= tri_lim : -(long long)v - 1 >= quad_lim) return false;
  return true;
}

}  // namespace

std::string validate_setup(const MeshSetup& s, int np, int nt, int nq, const int* tri,
                           const int* quad) {
  if (s.np != np || s.nt != nt || s.nq != nq) return "sizes differ from the mesh";
  // pattern: the union of the bending stencils, sorted, no diagonal
  if (!is_ptr(s.rowptr, np, s.nnz) || (int)s.colind.size() != s.nnz) return "bad row offsets";
  if ((int)s.diaind.size() != np || (int)s.rowof.size() != s.nnz) return "bad pattern sizes";
  for (int i = 0; i < np; i++) {
    int first_upper = s.rowptr[i + 1];
    for (int k = s.rowptr[i]; k < s.rowptr[i + 1]; k++) {
      const int j = s.colind[k];
      if (j < 0 || j >= np || j == i || s.rowof[k] != i) return "bad column index";
      if (k > s.rowptr[i] && j <= s.colind[k - 1]) return "columns not ascending";
      if (j > i && first_upper == s.rowptr[i + 1]) first_upper = k;
    }
    if (s.diaind[i] != first_upper) return "bad diagonal index";
  }
  {
    // every stencil pair is an entry and every entry comes from a stencil
    std::vector<char> covered(s.nnz, 0);
    for (int e = 0; e < nq; e++)
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) {
          const int i = quad[4 * e + a], j = quad[4 * e + b];
          if (i == j) continue;
          if (i < 0 || i >= np || j < 0 || j >= np) return "bending stencil index out of range";
          const int* lo = s.colind.data() + s.rowptr[i];
          const int* hi = s.colind.data() + s.rowptr[i + 1];
          const int* f = std::lower_bound(lo, hi, j);
          if (f == hi || *f != j) return "pattern misses a stencil pair";
          covered[f - s.colind.data()] = 1;
        }
    for (int k = 0; k < s.nnz; k++)
      if (!covered[k]) return "pattern entry outside the bending stencils";
  }
  // ILU(0) update lists
  if (!is_ptr(s.fu_ptr, np, (int)s.fu_kj.size()) || s.fu_code.size() != s.fu_kj.size())
    return "bad ILU update lists";
  for (int i = 0; i < np; i++)
    for (int u = s.fu_ptr[i]; u < s.fu_ptr[i + 1]; u++) {
      const int ik = s.fu_code[u] & 0xff, tgt = (s.fu_code[u] >> 8) & 0xff;
      const int kj = s.fu_kj[u];
      if (ik >= s.diaind[i] - s.rowptr[i] || kj < 0 || kj >= s.nnz) return "bad ILU update";
      const int k = s.colind[s.rowptr[i] + ik];
      if (kj < s.diaind[k] || kj >= s.rowptr[k + 1]) return "ILU update outside the upper row";
      const int j = s.colind[kj];
      if (tgt == kFuDiag ? j != i : (tgt >= s.rowptr[i + 1] - s.rowptr[i] ||
                                      s.colind[s.rowptr[i] + tgt] != j))
        return "ILU update target is not the column";
    }
  // level schedules: every dependency lies in a strictly earlier level, so
  // the rows of one level are independent
  std::vector<int> lf, lb;
  if (!level_of(s.lf_ptr, s.lf_rows, s.nlf, np, lf) || !level_of(s.lb_ptr, s.lb_rows, s.nlb, np, lb))
    return "bad level lists";
  for (int i = 0; i < np; i++) {
    for (int k = s.rowptr[i]; k < s.diaind[i]; k++)
      if (lf[s.colind[k]] >= lf[i]) return "forward level schedule breaks a dependency";
    for (int k = s.diaind[i]; k < s.rowptr[i + 1]; k++)
      if (lb[s.colind[k]] >= lb[i]) return "backward level schedule breaks a dependency";
  }
  // components: rows joined by a block share one
  if ((int)s.comp.size() != np || s.ncomp < 1) return "bad components";
  for (int i = 0; i < np; i++) {
    if (s.comp[i] < 0 || s.comp[i] >= s.ncomp) return "bad component id";
    for (int k = s.rowptr[i]; k < s.rowptr[i + 1]; k++)
      if (s.comp[s.colind[k]] != s.comp[i]) return "a block joins two components";
  }
  if (s.qwidth < 1 || s.qwidth > kQuadRows) return "bad sub-level width";
  if ((int)s.qf_ptr.size() != s.nqf + 1 || (int)s.qb_ptr.size() != s.nqb + 1 ||
      !subs_ok(s, s.sf_rows, s.qf_ptr, s.qf_comp, lf, s.nqf) ||
      !subs_ok(s, s.sb_rows, s.qb_ptr, s.qb_comp, lb, s.nqb))
    return "bad sweep sub-levels";
  // packed sweeps: slots, counts and LDS ring offsets (the kernel indexes
  // its ring with these offsets, so they must stay inside it)
  if (s.kl < 0 || s.ku < 0 || s.kl > kMaxBlocksPerRow || s.ku > kMaxBlocksPerRow)
    return "bad packed block counts";
  if (s.kl > 0) {
    int kl = 0, ku = 0;
    for (int i = 0; i < np; i++) {
      kl = std::max(kl, s.diaind[i] - s.rowptr[i]);
      ku = std::max(ku, s.rowptr[i + 1] - s.diaind[i]);
    }
    if (kl != s.kl || ku != s.ku) return "packed block counts differ from the pattern";
    const int kq = std::max(std::max(s.kl, s.ku), 1);
    const size_t mw = (size_t)meta_groups(kq) * 8;
    if (s.f_rci.size() != 4 * (size_t)np || s.b_rc.size() != 2 * (size_t)np ||
        !is_perm(s.fslot, np) || s.f_meta.size() != mw * np || s.b_meta.size() != mw * np)
      return "bad packed sweep sizes";
    std::vector<int> bslot(np);
    for (int sl = 0; sl < np; sl++) bslot[s.sb_rows[sl]] = sl;
    auto midx = [np](int h, int sl) { return ((size_t)(h / 8) * np + sl) * 8 + h % 8; };
    const unsigned short zero_byte = (unsigned short)(kRingZero * (int)sizeof(double));
    for (int sl = 0; sl < np; sl++) {
      const int i = s.sf_rows[sl], j = s.sb_rows[sl];
      if (s.fslot[i] != sl || s.f_rci[4 * sl] != i || s.f_rci[4 * sl + 1] != s.diaind[i] - s.rowptr[i] ||
          s.f_rci[4 * sl + 2] != bslot[i] || s.b_rc[2 * sl] != j ||
          s.b_rc[2 * sl + 1] != s.rowptr[j + 1] - s.diaind[j])
        return "bad packed slot record";
      for (int k = 0; k < kq; k++) {
        const unsigned short fo = s.f_meta[midx(k, sl)], bo = s.b_meta[midx(k, sl)];
        const int fc = s.diaind[i] - s.rowptr[i], bc = s.rowptr[j + 1] - s.diaind[j];
        const unsigned short fw =
            k < fc ? (unsigned short)((s.fslot[s.colind[s.rowptr[i] + k]] & (kRingRows - 1)) * 8)
                   : zero_byte;
        const unsigned short bw =
            k < bc ? (unsigned short)((bslot[s.colind[s.diaind[j] + k]] & (kRingRows - 1)) * 8)
                   : zero_byte;
        if (fo != fw || bo != bw) return "bad sweep ring offset";
      }
      const unsigned fout = s.f_meta[midx(kq, sl)] | ((unsigned)s.f_meta[midx(kq + 1, sl)] << 16);
      const unsigned bout = s.b_meta[midx(kq, sl)] | ((unsigned)s.b_meta[midx(kq + 1, sl)] << 16);
      if (fout != (unsigned)bslot[i] || bout != (unsigned)j) return "bad sweep output index";
    }
    if (s.ring_ok && !(ring_fits(s, lf, true, s.fslot) && ring_fits(s, lb, false, bslot)))
      return "sweep dependencies do not fit the LDS ring";
  }
  // gather lists
  if (!src_ok(s.crs_src_ptr, s.crs_src, s.nnz, 9LL * nt, 16LL * nq) ||
      !src_ok(s.dia_src_ptr, s.dia_src, np, 9LL * nt, 16LL * nq) ||
      !src_ok(s.grad_src_ptr, s.grad_src, np, 3LL * nt, 4LL * nq))
    return "bad assembly gather lists";
  if (!is_ptr(s.ring_ptr, np, (int)s.ring.size())) return "bad vertex rings";
  for (int v : s.ring)
    if (v < 0 || v >= np) return "bad vertex ring entry";
  // BVH: a binary tree over the triangles rooted at node 0
  const int nn = s.nnode;
  if (nn != 2 * nt - 1 || (int)s.node.size() != 3 * nn || s.root != 0) return "bad BVH size";
  if ((int)s.pre.size() != nn || (int)s.pre_end.size() != nn || !is_perm(s.by_pre, nn) ||
      (int)s.depth.size() != nn || (int)s.leaf_of.size() != nt || !is_perm(s.refit_order, nn) ||
      !is_ptr(s.refit_lvl_ptr, s.max_depth + 1, nn))
    return "bad BVH arrays";
  std::vector<int> leaf_seen(nt, 0);
  for (int v = 0; v < nn; v++) {
    const int p = s.node[3 * v], c0 = s.node[3 * v + 1], c1 = s.node[3 * v + 2];
    if (v == 0 ? p != -1 : (p < 0 || p >= nn)) return "bad BVH parent";
    if (s.by_pre[s.pre[v]] != v || s.pre_end[v] <= s.pre[v] || s.pre_end[v] > nn)
      return "bad BVH pre-order";
    if (s.depth[v] < 0 || s.depth[v] > s.max_depth) return "bad BVH depth";
    if (c1 == -1) {
      if (c0 < 0 || c0 >= nt || leaf_seen[c0]++ || s.leaf_of[c0] != v) return "bad BVH leaf";
    } else {
      if (c0 <= 0 || c0 >= nn || c1 <= 0 || c1 >= nn || s.node[3 * c0] != v || s.node[3 * c1] != v ||
          s.depth[c0] != s.depth[v] + 1 || s.depth[c1] != s.depth[v] + 1)
        return "bad BVH child";
    }
  }
  for (int l = 0; l <= s.max_depth; l++)
    for (int k = s.refit_lvl_ptr[l]; k < s.refit_lvl_ptr[l + 1]; k++)
      if (s.depth[s.refit_order[k]] != s.max_depth - l) return "bad BVH refit order";
  (void)tri;
  return "";
}

}  // namespace cs
