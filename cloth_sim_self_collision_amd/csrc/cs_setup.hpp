// Host-side, per-mesh setup of the cloth step: everything that depends only
// on the mesh connectivity (and, for the BVH, on the initial positions) and is
// built once before the device step loop runs.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cs {

// source encodings of the gather lists
//   Hessian block sources: >= 0 -> triangle block t*9 + a*3 + b
//                          <  0 -> bending block -(q*16 + a*4 + b) - 1
//   gradient sources:      >= 0 -> triangle corner t*3 + a
//                          <  0 -> bending corner -(q*4 + a) - 1
struct MeshSetup {
  int np = 0, nt = 0, nq = 0, nnz = 0;
  // 3x3-block CSR pattern without the diagonal: union of the bending stencils
  // (self_contact_sparse/main.cpp:577-581)
  std::vector<int> rowptr, colind, diaind, rowof;
  // level schedules: forward substitution / ILU rows (dependencies through the
  // strictly lower part) and backward substitution (strictly upper part)
  std::vector<int> lf_ptr, lf_rows, lb_ptr, lb_rows;
  int nlf = 0, nlb = 0, max_lf = 0, max_lb = 0;
  // ILU(0) updates of each row in the factorisation's loop order
  // (ilu_sparse.cpp:507-569: for each lower block ik of row i in column
  // order, for each upper block kj of row k = colind[ik] in column order
  // whose column j is i or a column of row i): fu_ptr[i] .. fu_ptr[i + 1],
  // fu_kj = kj, fu_code = (ik - rowptr[i]) | target << 8 with target the
  // block's index in row i from rowptr[i], or kFuDiag for j == i
  std::vector<int> fu_ptr, fu_kj, fu_code;
  // connected components of the pattern (rows joined by a block), numbered
  // in order of their smallest row; the matrix is block diagonal over them,
  // so each component's sweeps are independent of the others'
  int ncomp = 0;
  std::vector<int> comp;  // component of each row
  // quad sweep slot orders: rows by (component, level, row) for the forward
  // and the backward sweep, cut into sub-levels of at most kQuadRows rows of
  // one level of one component (the rows of a level are independent, so
  // running a level in pieces keeps every dependency satisfied);
  // qf_comp[c] .. qf_comp[c + 1] are component c's sub-levels (same backward)
  std::vector<int> sf_rows, sb_rows;
  std::vector<int> qf_ptr, qb_ptr, qf_comp, qb_comp;
  int nqf = 0, nqb = 0;
  int qwidth = 0;  // rows of the widest sub-level
  // packed sweeps: rows in level order ("slots"), at most kl lower / ku upper
  // blocks per row.  f_rci[4*slot] = {row, block count, the row's backward
  // slot, 0}, b_rc[2*slot] = {row, block count}, fslot[row] = forward slot.
  // Sweep metadata per slot (f_meta, b_meta), 16-bit words in 16-byte groups,
  // word h at [((h / 8) * np + slot) * 8 + h % 8]: for each of the
  // kq = max(kl, ku) blocks the byte offset of its column row's entry in a
  // row of the LDS ring (entry = that row's slot in the same sweep mod
  // kRingRows; blocks past the row's count: the ring's zero entry), then a
  // 32-bit value as low and high word (forward: the row's backward slot;
  // backward: the row)
  int kl = 0, ku = 0;
  std::vector<int> f_rci, b_rc, fslot;
  std::vector<unsigned short> f_meta, b_meta;
  // the LDS ring of kRingRows entries (entry = slot mod kRingRows) can hold the
  // sweep dependencies: no two rows of one component sharing an entry are
  // live at the same level (each component is swept by its own workgroup)
  bool ring_ok = false;
  // gather lists in element order (assembly without atomics, in the
  // reference's summation order)
  std::vector<int> crs_src_ptr, crs_src;
  std::vector<int> dia_src_ptr, dia_src;
  std::vector<int> grad_src_ptr, grad_src;
  // vertex 1-ring through triangles (rigid impact zones)
  std::vector<int> ring_ptr, ring;
  // BVH: node i = (parent, child0, child1); child1 == -1 marks a leaf whose
  // child0 is its triangle.  Node numbering is the reference's creation order.
  int nnode = 0, root = 0;
  std::vector<int> node;
  std::vector<int> pre;       // pre-order rank of each node
  std::vector<int> pre_end;   // one past the last pre-order rank of the subtree
  std::vector<int> by_pre;    // node id at each pre-order rank
  std::vector<int> leaf_of;   // leaf node of each triangle
  std::vector<int> depth;     // depth of each node (root 0)
  std::vector<int> refit_order;  // nodes ordered by decreasing depth (bottom-up refit)
  std::vector<int> refit_lvl_ptr;  // level boundaries inside refit_order
  int max_depth = 0;
};

constexpr int kRingRows = 4096;
// a ring row holds kRingRows entries, then a write-only dummy entry (lanes
// without a row) and an entry that always holds +0.0 (read by padded blocks)
constexpr int kRingDummy = kRingRows, kRingZero = kRingRows + 1, kRingLen = kRingRows + 2;
constexpr int kQuadRows = 128;  // rows per sub-level of the quad sweep (4 lanes each)
constexpr int kMaxBlocksPerRow = 8;
constexpr int kFuDiag = 0xff;  // update target: the diagonal block
// 16-byte groups of sweep metadata per slot for k blocks per row
inline constexpr int meta_groups(int k) { return (k + 2 + 7) / 8; }

// Returns an empty string on success, else a description of the problem.
std::string build_setup(MeshSetup& s, int np, int nt, int nq, const int* tri, const int* quad,
                        const double* xyz_init);

// Structural check of a setup against its mesh: the pattern is the stencils'
// union, every index is in range, the level schedules respect every
// dependency (rows of a level are independent), sub-levels and sweep slots
// are consistent, the LDS ring offsets stay inside the ring (and ring_ok
// holds when claimed), gather lists and the BVH are well formed.  Empty
// string when it holds.  Every setup loaded from the cache passes it before
// use; fresh builds are checked when CLOTHSIM_VALIDATE_SETUP is set.
std::string validate_setup(const MeshSetup& s, int np, int nt, int nq, const int* tri,
                           const int* quad);

// build_setup through the on-disk cache of setup products (cs_setup_cache.cpp;
// no cache directory set: plain build_setup)
std::string build_setup_cached(MeshSetup& s, int np, int nt, int nq, const int* tri,
                               const int* quad, const double* xyz_init);
void set_setup_cache_dir(const char* dir);  // null or "": no cache
void setup_cache_stats(long long* hits, long long* misses, long long* writes);

}  // namespace cs
