// C++ host mirror of the reference's lower-level classes and functions on the
// step path, over the clothsim C ABI.  Same names, argument order and meaning
// as
//   CMatrixSquareSparse                     matrix_square_sparse.h:14-48
//   CPreconditionerILU, Solve_PCG           ilu_sparse.h:16-43
//   CAABB3D (bounds and isnt_empty)         aabb.h:16-36
//   CNodeBVH, MakeBVHTopology_TopDown,
//   BuildBoundingBoxChild_Prx / _CCD,
//   CContactElement,
//   GetContactElement_Proximity / _CCD      bvh_aabb.h:18-130
//   WdWddW_CST, WdWddW_Bend, WdWddW_Contact cloth_internal_physics.h:70-280
// with one difference: every object and function is bound to a
// clothsim::Cloth (clothsim_ref.hpp), the mesh whose pattern, factor and BVH
// live on the GPU, passed as the constructor argument or an extra last
// argument.  A matrix's pattern must be that cloth's (the union of its
// bending stencils, as the reference's setup builds it); SetPattern checks
// it.  BVH calls start at the root node, as every reference call site does.
//
// Arithmetic (element derivatives, MatVec, ILU factor and substitutions, PCG,
// box refits, contact tests) runs on the GPU and is bit-identical to the
// reference's.  The matrix values are host arrays, like the reference's
// public members, so the caller may poke them (solve_internal_sparse.h:303-305
// adds the mass term to m_valDia this way); SetZero, Mearge and
// SetBoundaryCondition act on these host arrays, and each device operation
// that reads the matrix uploads them first.  Element calls cost one launch
// each: the step path (StepTime_InternalDynamicsILU) assembles every element
// on the device in one pass instead.
//
// Failures print the library's message and abort, as the reference asserts.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <set>
#include <utility>
#include <vector>

#include "clothsim_ref.hpp"

namespace clothsim {

// matrix_square_sparse.h:14-48.  Naming as in the reference: m_colInd holds
// the row offsets (nblk + 1), m_rowPtr the column indices (m_ncrs).
class CMatrixSquareSparse {
 public:
  CMatrixSquareSparse() = default;
  explicit CMatrixSquareSparse(Cloth& cloth) : cloth_(&cloth) {}
  CMatrixSquareSparse(const CMatrixSquareSparse&) = delete;
  CMatrixSquareSparse& operator=(const CMatrixSquareSparse&) = delete;

  void Bind(Cloth& cloth) { cloth_ = &cloth; }
  Cloth& cloth() const {
    if (!cloth_) fail("the matrix is not bound to a cloth");
    return *cloth_;
  }

  // matrix_square_sparse.cpp:46-57: nblk blocks of len x len; the cloth's
  // vertex count and 3
  void Initialize(int nblk, int len) {
    if (nblk != cloth().np() || len != 3) fail("Initialize: the cloth has np blocks of 3x3");
    m_nblk = nblk;
    m_len = len;
    m_ncrs = 0;
    colInd_.assign(nblk + 1, 0);
    rowPtr_.clear();
    valDia_.assign(9 * (size_t)nblk, 0.0);
    valCrs_.clear();
    refresh();
  }

  // matrix_square_sparse.cpp:140-163.  Returns false (and changes nothing)
  // when the pattern is not the cloth's.
  bool SetPattern(const std::vector<int>& colind, const std::vector<int>& rowptr) {
    std::vector<int> rp, ci, di;
    cloth_pattern(rp, ci, di);
    if (colind.size() != (size_t)m_nblk + 1 || colind[m_nblk] != (int)ci.size() ||
        rowptr.size() < ci.size() || !std::equal(rp.begin(), rp.end(), colind.begin()) ||
        !std::equal(ci.begin(), ci.end(), rowptr.begin())) {
      std::fprintf(stderr, "clothsim: SetPattern: not the cloth's pattern\n");
      return false;
    }
    colInd_ = rp;
    rowPtr_ = ci;
    m_ncrs = (int)ci.size();
    valCrs_.assign(9 * (size_t)m_ncrs, 0.0);
    refresh();
    return true;
  }

  // matrix_square_sparse.cpp:80-86
  bool SetZero() {
    std::fill(valDia_.begin(), valDia_.end(), 0.0);
    std::fill(valCrs_.begin(), valCrs_.end(), 0.0);
    return true;
  }

  // matrix_square_sparse.cpp:88-138: adds an element matrix laid out
  // [i][j][len][len] into the host values; blocks outside the pattern are
  // dropped.  marge_buffer holds nblk entries of -1 between calls.
  bool Mearge(int nblkel_col, const int* blkel_col, int nblkel_row, const int* blkel_row,
              int blksize, const double* emat, std::vector<int>& marge_buffer) {
    if (nblkel_col != nblkel_row || blksize != m_len * m_len) return false;
    for (int a = 0; a < nblkel_col; a++) {
      const int row = blkel_col[a];
      for (int k = colInd_[row]; k < colInd_[row + 1]; k++) marge_buffer[rowPtr_[k]] = k;
      for (int b = 0; b < nblkel_row; b++) {
        double* dst;
        if (a == b) {
          dst = &valDia_[(size_t)row * blksize];
        } else {
          const int k = marge_buffer[blkel_row[b]];
          if (k < 0) continue;
          dst = &valCrs_[(size_t)k * blksize];
        }
        const double* src = emat + (size_t)(a * nblkel_row + b) * blksize;
        for (int q = 0; q < blksize; q++) dst[q] += src[q];
      }
      for (int k = colInd_[row]; k < colInd_[row + 1]; k++) marge_buffer[rowPtr_[k]] = -1;
    }
    return true;
  }

  // matrix_square_sparse.cpp:167-270: y = alpha A x + beta y on the GPU
  void MatVec(double alpha, const std::vector<double>& x, double beta,
              std::vector<double>& y) const {
    upload();
    check(cs_matvec(cloth().handle(), alpha, x.data(), beta, y.data()), "cs_matvec");
  }

  // matrix_square_sparse.cpp:272-302: rows and columns of fixed blocks
  // zeroed, their diagonal blocks the identity
  void SetBoundaryCondition(const std::vector<int>& bc_flag) {
    for (int i = 0; i < m_nblk; i++) {
      if (bc_flag[i] == 0) continue;
      double* d = &valDia_[9 * (size_t)i];
      for (int q = 0; q < 9; q++) d[q] = (q % 4 == 0) ? 1.0 : 0.0;
      for (int k = colInd_[i]; k < colInd_[i + 1]; k++)
        for (int q = 0; q < 9; q++) valCrs_[9 * (size_t)k + q] = 0.0;
    }
    for (int k = 0; k < m_ncrs; k++) {
      if (bc_flag[rowPtr_[k]] == 0) continue;
      for (int q = 0; q < 9; q++) valCrs_[9 * (size_t)k + q] = 0.0;
    }
  }

  // not in the reference: the cloth's device system (after the device
  // assembled it) into the host values
  void Download() {
    check(cs_get_system(cloth().handle(), valDia_.data(), m_ncrs ? valCrs_.data() : nullptr,
                        nullptr),
          "cs_get_system");
  }
  // the host values into the cloth's device system
  void upload() const {
    check(cs_set_system(cloth().handle(), valDia_.data(), m_ncrs ? valCrs_.data() : nullptr,
                        nullptr),
          "cs_set_system");
  }
  // the cloth's pattern (row offsets, columns, first column above the diagonal)
  void cloth_pattern(std::vector<int>& rp, std::vector<int>& ci, std::vector<int>& di) const {
    int np = 0, nt = 0, nq = 0, nnz = 0, nnode = 0;
    check(cs_cloth_sizes(cloth().handle(), &np, &nt, &nq, &nnz, &nnode), "cs_cloth_sizes");
    rp.assign(np + 1, 0);
    ci.assign(nnz, 0);
    di.assign(np, 0);
    check(cs_get_pattern(cloth().handle(), rp.data(), ci.data(), di.data()), "cs_get_pattern");
  }

 public:
  int m_nblk = 0;
  int m_len = 0;
  int m_ncrs = 0;
  int* m_colInd = nullptr;
  int* m_rowPtr = nullptr;
  double* m_valCrs = nullptr;
  double* m_valDia = nullptr;

 private:
  friend class CPreconditionerILU;
  [[noreturn]] static void fail(const char* msg) {
    std::fprintf(stderr, "clothsim: %s\n", msg);
    std::abort();
  }
  void refresh() {
    m_colInd = colInd_.data();
    m_rowPtr = rowPtr_.empty() ? nullptr : rowPtr_.data();
    m_valDia = valDia_.data();
    m_valCrs = valCrs_.empty() ? nullptr : valCrs_.data();
  }
  Cloth* cloth_ = nullptr;
  std::vector<int> colInd_, rowPtr_;
  std::vector<double> valDia_, valCrs_;
};

// ilu_sparse.h:16-34.  mat holds the factor (D^-1 on the diagonal, L and
// D^-1 U off it, as DoILUDecomp leaves them), m_diaInd the first column above
// the diagonal of each row.
class CPreconditionerILU {
 public:
  CPreconditionerILU() = default;
  CPreconditionerILU(const CPreconditionerILU&) = delete;
  CPreconditionerILU& operator=(const CPreconditionerILU&) = delete;
  ~CPreconditionerILU() {
    if (mat.cloth_ && mat.cloth_->factor_owner == this) mat.cloth_->factor_owner = nullptr;
  }

  // ilu_sparse.cpp:114-134
  void Initialize_ILU0(const CMatrixSquareSparse& m) {
    mat.Bind(m.cloth());
    mat.Initialize(m.m_nblk, m.m_len);
    std::vector<int> rp, ci;
    mat.cloth_pattern(rp, ci, diaInd_);
    if (!mat.SetPattern(rp, ci)) CMatrixSquareSparse::fail("Initialize_ILU0: pattern");
    m_diaInd = diaInd_.data();
  }
  // ilu_sparse.cpp:352-384: the values to factor
  void SetValueILU(const CMatrixSquareSparse& m) {
    mat.valDia_ = m.valDia_;
    mat.valCrs_ = m.valCrs_;
    mat.refresh();
  }
  // ilu_sparse.cpp:387-624 on the GPU; the factor stays resident there and
  // comes back into mat.  Pivots with |det| <= 1e-30 are counted, not fatal.
  void DoILUDecomp() {
    Cloth& c = mat.cloth();
    mat.upload();
    check(cs_ilu_factor(c.handle(), &n_singular), "cs_ilu_factor");
    check(cs_get_factor(c.handle(), mat.valDia_.data(),
                        mat.m_ncrs ? mat.valCrs_.data() : nullptr),
          "cs_get_factor");
    c.factor_owner = this;
  }
  // ilu_sparse.h:23-26: forward then backward substitution, in place
  void Solve(std::vector<double>& vec) const {
    resident();
    check(cs_ilu_solve(mat.cloth().handle(), vec.data()), "cs_ilu_solve");
  }
  // the factor on the device is this one's (another preconditioner of the
  // same cloth, or its internal step, may have replaced it)
  void resident() const {
    Cloth& c = mat.cloth();
    if (c.factor_owner == this) return;
    check(cs_set_factor(c.handle(), mat.valDia_.data(), mat.m_ncrs ? mat.valCrs_.data() : nullptr),
          "cs_set_factor");
    c.factor_owner = this;
  }

 public:
  CMatrixSquareSparse mat;
  int* m_diaInd = nullptr;
  int n_singular = 0;  // not in the reference: singular pivots of the last DoILUDecomp

 private:
  std::vector<int> diaInd_;
};

// ilu_sparse.cpp:628-709.  In: conv_ratio = tolerance, iteration = cap,
// r_vec = right-hand side.  Out: the achieved ratio and iteration as the
// reference reports them, x_vec the solution, r_vec the last residual.
inline void Solve_PCG(double& conv_ratio, int& iteration, const CMatrixSquareSparse& mat,
                      const CPreconditionerILU& ilu, std::vector<double>& r_vec,
                      std::vector<double>& x_vec) {
  mat.upload();
  ilu.resident();
  x_vec.assign(r_vec.size(), 0.0);
  double conv = 0.0;
  int it = 0;
  check(cs_pcg_solve(mat.cloth().handle(), r_vec.data(), conv_ratio, iteration, x_vec.data(),
                     r_vec.data(), &conv, &it),
        "cs_pcg_solve");
  conv_ratio = conv;
  iteration = it;
}

// aabb.h:16-36 (the six bounds and the flag)
struct CAABB3D {
  double x_min = 0, x_max = 0, y_min = 0, y_max = 0, z_min = 0, z_max = 0;
  bool isnt_empty = false;
};

// bvh_aabb.h:18-23: iroot = parent (-1 at the root); a leaf has
// ichild[1] == -1 and holds triangle ichild[0]
struct CNodeBVH {
  int iroot;
  int ichild[2];
};

// bvh_aabb.cpp:294-313: the cloth's topology, built at its creation from the
// same triangles and positions; returns the root (node 0)
inline int MakeBVHTopology_TopDown(const std::vector<int>& aTri, const std::vector<double>& aXYZ,
                                   std::vector<CNodeBVH>& aNodeBVH, const Cloth& cloth) {
  (void)aTri, (void)aXYZ;
  int np = 0, nt = 0, nq = 0, nnz = 0, nnode = 0;
  check(cs_cloth_sizes(cloth.handle(), &np, &nt, &nq, &nnz, &nnode), "cs_cloth_sizes");
  std::vector<int> nodes(3 * (size_t)nnode);
  check(cs_get_bvh(cloth.handle(), nodes.data()), "cs_get_bvh");
  aNodeBVH.resize(nnode);
  for (int i = 0; i < nnode; i++)
    aNodeBVH[i] = CNodeBVH{nodes[3 * i], {nodes[3 * i + 1], nodes[3 * i + 2]}};
  return 0;
}

namespace detail {
inline void root_only(int ibvh, const char* what) {
  if (ibvh == 0) return;
  std::fprintf(stderr, "clothsim: %s starts at the root node (0), not %d\n", what, ibvh);
  std::abort();
}
inline void boxes_out(const std::vector<double>& bb, std::vector<CAABB3D>& aBB) {
  aBB.resize(bb.size() / 6);
  for (size_t i = 0; i < aBB.size(); i++) {
    const double* b = &bb[6 * i];
    aBB[i] = CAABB3D{b[0], b[1], b[2], b[3], b[4], b[5], true};
  }
}
inline int node_count(const Cloth& cloth) {
  int np = 0, nt = 0, nq = 0, nnz = 0, nnode = 0;
  check(cs_cloth_sizes(cloth.handle(), &np, &nt, &nq, &nnz, &nnode), "cs_cloth_sizes");
  return nnode;
}
}  // namespace detail

// bvh_aabb.cpp:23-61: leaf boxes are the triangles grown by delta / 2
inline void BuildBoundingBoxChild_Prx(int ibvh, double delta, const std::vector<double>& aXYZ,
                                      const std::vector<int>& aTri,
                                      const std::vector<CNodeBVH>& aNodeBVH,
                                      std::vector<CAABB3D>& aBB, Cloth& cloth) {
  (void)aTri, (void)aNodeBVH;
  detail::root_only(ibvh, "BuildBoundingBoxChild_Prx");
  std::vector<double> bb(6 * (size_t)detail::node_count(cloth));
  check(cs_bvh_refit_prx(cloth.handle(), aXYZ.data(), delta, bb.data()), "cs_bvh_refit_prx");
  detail::boxes_out(bb, aBB);
}

// bvh_aabb.cpp:64-106: leaf boxes hold each vertex at p and p + dt v, grown
// by 1e-10
inline void BuildBoundingBoxChild_CCD(int ibvh, double dt, const std::vector<double>& aXYZ,
                                      const std::vector<double>& aUVW,
                                      const std::vector<int>& aTri,
                                      const std::vector<CNodeBVH>& aNodeBVH,
                                      std::vector<CAABB3D>& aBB, Cloth& cloth) {
  (void)aTri, (void)aNodeBVH;
  detail::root_only(ibvh, "BuildBoundingBoxChild_CCD");
  std::vector<double> bb(6 * (size_t)detail::node_count(cloth));
  check(cs_bvh_refit_ccd(cloth.handle(), aXYZ.data(), aUVW.data(), dt, bb.data()),
        "cs_bvh_refit_ccd");
  detail::boxes_out(bb, aBB);
}

// bvh_aabb.h:65-106: the canonical index order (face vertices ascending then
// the vertex; each edge ascending, the edge holding the smallest index
// first); ordering ignores is_fv, as the reference's does
class CContactElement {
 public:
  CContactElement(bool is_fv, int j0, int j1, int j2, int j3) : is_fv(is_fv) {
    if (is_fv) {
      int f[3] = {j0, j1, j2};
      if (f[1] < f[0]) std::swap(f[0], f[1]);
      if (f[2] < f[1]) std::swap(f[1], f[2]);
      if (f[1] < f[0]) std::swap(f[0], f[1]);
      ino0 = f[0], ino1 = f[1], ino2 = f[2], ino3 = j3;
    } else {
      int a0 = std::min(j0, j1), a1 = std::max(j0, j1);
      int b0 = std::min(j2, j3), b1 = std::max(j2, j3);
      if (b0 < a0) std::swap(a0, b0), std::swap(a1, b1);
      ino0 = a0, ino1 = a1, ino2 = b0, ino3 = b1;
    }
  }
  bool operator<(const CContactElement& o) const {
    if (ino0 != o.ino0) return ino0 < o.ino0;
    if (ino1 != o.ino1) return ino1 < o.ino1;
    if (ino2 != o.ino2) return ino2 < o.ino2;
    return ino3 < o.ino3;
  }

 public:
  bool is_fv;
  int ino0, ino1, ino2, ino3;
};

namespace detail {
template <typename Query>
inline void contacts_into(std::set<CContactElement>& out, Query q) {
  std::vector<int> buf(5 * 4096);
  int n = 0;
  check(q(buf.data(), (int)(buf.size() / 5), &n), "contact query");
  if ((size_t)n > buf.size() / 5) {
    buf.resize(5 * (size_t)n);
    check(q(buf.data(), n, &n), "contact query");
  }
  for (int k = 0; k < n; k++) {
    const int* e = &buf[5 * (size_t)k];
    out.insert(CContactElement(e[0] != 0, e[1], e[2], e[3], e[4]));
  }
}
}  // namespace detail

// bvh_aabb.cpp:540-557 (self query from the root): the proximity contacts of
// the positions, in the reference's set; the cloth refits its own boxes, so
// aBB is not read
inline void GetContactElement_Proximity(std::set<CContactElement>& aContactElem, double delta,
                                        const std::vector<double>& aXYZ,
                                        const std::vector<int>& aTri, int ibvh,
                                        const std::vector<CNodeBVH>& aBVH,
                                        const std::vector<CAABB3D>& aBB, Cloth& cloth) {
  (void)aTri, (void)aBVH, (void)aBB;
  detail::root_only(ibvh, "GetContactElement_Proximity");
  detail::contacts_into(aContactElem, [&](int* out, int cap, int* n) {
    return cs_query_proximity(cloth.handle(), aXYZ.data(), delta, out, cap, n);
  });
}

// bvh_aabb.cpp:904-923: the continuous contacts over [0, dt]
inline void GetContactElement_CCD(std::set<CContactElement>& aContactElem, double dt,
                                  double delta, const std::vector<double>& aXYZ,
                                  const std::vector<double>& aUVW, const std::vector<int>& aTri,
                                  int ibvh, const std::vector<CNodeBVH>& aBVH,
                                  const std::vector<CAABB3D>& aBB, Cloth& cloth) {
  (void)aTri, (void)aBVH, (void)aBB;
  detail::root_only(ibvh, "GetContactElement_CCD");
  detail::contacts_into(aContactElem, [&](int* out, int cap, int* n) {
    return cs_query_ccd(cloth.handle(), aXYZ.data(), aUVW.data(), dt, delta, out, cap, n);
  });
}

// cloth_internal_physics.h:70-176
inline void WdWddW_CST(double& W, double dW[3][3], double ddW[3][3][3][3], const double C[3][3],
                       const double c[3][3], const double lambda, const double myu) {
  check(cs_element_cst(1, &C[0][0], &c[0][0], lambda, myu, &dW[0][0], &ddW[0][0][0][0]),
        "cs_element_cst");
  check(cs_element_cst_energy(1, &C[0][0], &c[0][0], lambda, myu, &W), "cs_element_cst_energy");
}

// cloth_internal_physics.h:180-238.  Like the reference, W is increased by
// dW . c, one term at a time in corner-then-axis order, into the caller's W.
inline void WdWddW_Bend(double& W, double dW[4][3], double ddW[4][4][3][3], const double C[4][3],
                        const double c[4][3], double stiff) {
  check(cs_element_bend(1, &C[0][0], &c[0][0], stiff, &dW[0][0], &ddW[0][0][0][0]),
        "cs_element_bend");
  for (int a = 0; a < 4; a++)
    for (int i = 0; i < 3; i++) W += dW[a][i] * c[a][i];
}

// cloth_internal_physics.h:243-280; the callback runs on the host, the
// penalty on the GPU
inline void WdWddW_Contact(double& W, double dW[3], double ddW[3][3], const double c[3],
                           double stiff_contact, double contact_clearance,
                           void (*penetrationDepth)(double&, double*, const double*)) {
  double pd = 0.0, n[3] = {0, 0, 0};
  penetrationDepth(pd, n, c);
  check(cs_element_contact_depth(1, &pd, n, stiff_contact, contact_clearance, &W, dW, &ddW[0][0]),
        "cs_element_contact_depth");
}

}  // namespace clothsim
