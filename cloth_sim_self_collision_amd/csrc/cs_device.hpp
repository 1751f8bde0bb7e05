// Device-resident state of one cloth and the host orchestration around it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "cs_ordered_sum.hpp"
#include "cs_setup.hpp"

namespace cs {

// status codes of the C-ABI (include/clothsim.h)
enum Status : int {
  kOk = 0,
  kBadArg = 1,
  kHipError = 2,
  kOutOfMemory = 3,
  kSingularPivot = 4,
  kNotConverged = 5,
  kNonFinite = 6,
  kCapacity = 7,
  kNoDevice = 8,
};

void set_error(const std::string& msg);
const char* last_error();

// Memory that is released without synchronising the device.  hipFree and
// hipHostFree wait for every stream of the device, so releasing a cloth, a
// batch or a contact buffer while another stream runs a long kernel (another
// batch's solve) would wait for that kernel.  Device buffers come from the
// device's stream-ordered pool, which keeps freed memory for later
// allocations (plain hipMalloc/hipFree where the device has no pools); small
// pinned host buffers come from a free list that is never handed back.
// dev_alloc: the memory is ready for work on s, and for any stream after s
// has been synchronised.  dev_free: after the work queued on s; work on
// other streams that used the memory must have completed.
int dev_alloc(void** p, size_t bytes, hipStream_t s);
void dev_free(void* p, hipStream_t s);
// zero-filled pinned host memory; pinned_free after its last use completed
int pinned_alloc(void** p, size_t bytes);
void pinned_free(void* p);

#define CS_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) {                                                               \
      ::cs::set_error(std::string(#call) + ": " + hipGetErrorString(e_));                 \
      return ::cs::kHipError;                                                             \
    }                                                                                     \
  } while (0)

struct Params {
  double lam, mu, stiff_bend, grav[3], mass, k_contact, clearance, dt;
  double pcg_tol;
  int pcg_maxit;
  int obstacle_kind;
  double obstacle[8];
};

// device scalar slots of the linear solve
enum Scal : int {
  kRz = 0,      // (r, M^-1 r) of the current iteration
  kInvRr0 = 1,  // 1 / (r0, r0)
  kAlpha = 2,
  kBeta = 3,
  kConv = 4,    // achieved residual ratio
  kTol2 = 5,    // tolerance squared
  kLastDot = 6,
  kNumScal = 16
};
enum Flag : int {
  kDone = 0,    // the solve has converged (or had a zero right-hand side)
  kIters = 1,   // iteration count reported like Solve_PCG
  kNsing = 2,   // ILU pivots with |det| <= 1e-30
  kFirstSing11 = 3,  // row of the 11th singular pivot, np if none
  // kDone as the main stream saw it before the convergence test of the
  // iteration started beside the sweep: the sweep and the (r, z) reduction
  // read this copy, which nothing changes while they run.  (A reduction whose
  // blocks disagreed on a flag changing under them would leave its ticket
  // short and corrupt every later reduction on the same scratch.)
  kSkip = 4,
  kTicketErr = 5,  // a reduction ticket was found nonzero after the solve
  kFlagNonFinite = 6,  // the step's displacement has an infinite or NaN entry
  kNumFlag = 16
};

// Plain pointers to every device array of a cloth; passed to kernels by value.
struct DevCloth {
  int np, nt, nq, nnz, nnode;
  int nlf, nlb;
  // mesh and setup
  const int* tri;
  const int* quad;
  const int* bc;
  const double* xyz0;
  const int* rowptr;
  const int* colind;
  const int* diaind;
  const int* rowof;
  const int* lf_ptr;
  const int* lf_rows;
  // ILU(0) update lists in the factorisation's loop order (MeshSetup::fu_*)
  // and the blocks of the widest row plus the diagonal (0: the staged
  // factorisation is not used)
  const int* fu_ptr;
  const int* fu_kj;
  const int* fu_code;
  int fu_blocks;
  const int* lb_ptr;
  const int* lb_rows;
  const int* qf_ptr;  // sub-levels of at most kQuadRows slots
  const int* qb_ptr;
  const int* qf_comp;  // each component's first sub-level (ncomp + 1 entries)
  const int* qb_comp;
  int nqf, nqb;
  int ncomp;   // connected components (one sweep workgroup each)
  int qwidth;  // rows of the widest sub-level
  const int* crs_src_ptr;
  const int* crs_src;
  const int* dia_src_ptr;
  const int* dia_src;
  const int* grad_src_ptr;
  const int* grad_src;
  const int* ring_ptr;
  const int* ring;
  const int* node;
  const int* pre;
  const int* pre_end;
  const int* by_pre;
  const int* leaf_of;
  const int* refit_order;
  // the traversal's view of the tree by pre-order rank: the subtree end of
  // the node at each rank (a leaf: rank + 1) and the rank of each leaf in
  // pre-order (nt entries)
  const int* rend;
  const int* leaf_rank;
  // state
  double* x;   // positions (3 np)
  double* v;   // velocities (3 np)
  double* x1;  // positions at the start of the step
  // linear system
  double* dia;
  double* crs;
  double* b;
  double* fdia;
  double* fcrs;
  double* etri;  // per triangle: 81 Hessian values then 9 gradient values
  double* kq;    // per bending stencil: 16 Hessian scalars (constant in time)
  double* eq;    // per bending stencil: 12 gradient values
  const double* host_pd;  // optional per-vertex obstacle depth (np) from a host callback
  const double* host_n;   // and normal (3 np); null -> analytic obstacle
  // PCG
  double* r;
  double* dx;
  double* z;
  double* ap;
  double* p[2];
  double* prod;  // element-wise products feeding the next reduction
  SumScratch sum;  // tile summaries of the ordered reductions over prod
  double* prod2;   // r.*r of the convergence test, reduced on the auxiliary stream
  SumScratch sum2;
  double* scal;
  int* flags;
  // BVH boxes (6 per node), by node id and by pre-order rank
  double* bb;
  double* rbb;
  // packed sweeps (null when a row has more than kMaxBlocksPerRow blocks)
  int kl, ku, ring_ok;
  int kpack;          // packed blocks per row of both sweeps: max(kl, ku), 0 = not packed
  int wmax;           // widest level of either sweep
  int sweep_variant;  // 0 automatic (quad sweep), 1 row-by-row
  int pcg_variant;    // 0 automatic, 1 one launch per PCG operation, 2 one-workgroup solve
  int precision;      // 0 FP64 (bit-exact), 1 mixed: the packed sweep's coefficients in FP32
  const int* f_rci;  // {row, block count, backward slot, 0} per forward slot
  const int* b_rc;   // {row, block count} per backward slot
  // sweep metadata per slot: kpack 16-bit ring byte offsets of the column
  // rows (padded blocks: the ring's zero entry), then the forward slot's
  // backward slot / the backward slot's row as two 16-bit halves; word h at
  // [((h / 8) * np + slot) * 8 + h % 8]
  const unsigned short* f_meta;
  const unsigned short* b_meta;
  const int* fslot;  // forward slot of each row
  double* sb;        // forward result in backward-slot order
  // packed layout: value v = k * 3 + j (entry (r, j) of block k of a slot's
  // row, r the output component) at [(((v / 2) * np + slot) * 3 + r) * 2 + v % 2]
  double* f_blk;
  double* f_dinv;  // D^-1 entry (r, j) the same way with v = j; v = 3: the forward sweep's right-hand side
  double* b_blk;
  int* ising;      // per-row singular-pivot flags of the last factorisation
};

// Which cloth a block works on.  One: the cloth handed to the launch by value
// (single-cloth launches).  Many: entry blockIdx.y of device tables, one grid
// row per cloth (batched launches); grids are sized for the largest cloth and
// blocks past a cloth's own size return at once.
// search direction of PCG iteration `it` (p_new = pcg_dir(c, it), p_old =
// pcg_dir(c, it + 1)).  A select, not c.p[it & 1]: a dynamic index into the
// struct would keep a per-thread copy of a batch table entry in scratch memory.
__device__ __forceinline__ double* pcg_dir(const DevCloth& c, int it) {
  return (it & 1) ? c.p[1] : c.p[0];
}

struct One {
  DevCloth c;
  Params P;
  __device__ __forceinline__ const DevCloth& cloth() const { return c; }
  __device__ __forceinline__ const Params& params() const { return P; }
  // first row of this block in a kernel that takes `rows` rows per block
  __device__ __forceinline__ int block_row0(int rows) const { return (int)blockIdx.x * rows; }
};

struct Many {
  const DevCloth* tab;
  const Params* ptab;
  __device__ __forceinline__ DevCloth cloth() const { return tab[blockIdx.y]; }
  __device__ __forceinline__ Params params() const { return ptab[blockIdx.y]; }
  __device__ __forceinline__ int block_row0(int rows) const { return (int)blockIdx.x * rows; }
};

// A batch's row kernel launched over the rows that exist rather than over
// (rows of the largest cloth) x (cloths): block k takes the rows from
// rowmap[k].y of cloth rowmap[k].x (one-dimensional grid of rowmap's length)
struct ManyRows {
  const DevCloth* tab;
  const Params* ptab;
  const int2* rowmap;
  __device__ __forceinline__ DevCloth cloth() const { return tab[rowmap[blockIdx.x].x]; }
  __device__ __forceinline__ Params params() const { return ptab[rowmap[blockIdx.x].x]; }
  __device__ __forceinline__ int block_row0(int) const { return rowmap[blockIdx.x].y; }
};


// Optional per-kernel-class timing with HIP events recorded on the cloth's
// stream around each launch (cs_profile_enable / cs_profile_read).
enum ProfKind : int {
  kProfSweep = 0,     // ILU forward+backward substitution
  kProfReduce = 1,    // ordered PCG reductions
  kProfSpmv = 2,      // search direction + A p
  kProfAxpy = 3,      // r, x updates
  kProfFactor = 4,    // ILU(0) factorisation
  kProfAssemble = 5,  // element kernels + gather assembly
  kProfContact = 6,   // refit, traversal, selection, impulses, zones
  kProfOther = 7,
  kProfPcgWg = 8,     // the whole PCG solve in one workgroup (k_pcg_wg)
  kNumProf = 9
};

struct Prof {
  bool on = false;
  unsigned mask = ~0u;  // kernel classes timed while on
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<int> kinds;
  hipStream_t stream = nullptr;
  double total_ms[kNumProf] = {0};
  long count[kNumProf] = {0};
  // fold the recorded pairs into the totals (waits for every pair: scopes
  // may sit on more than one stream)
  void harvest() {
    if (used == 0) return;
    for (size_t k = 0; k < kinds.size(); k++) {
      float ms = 0.f;
      (void)hipEventSynchronize(pool[2 * k + 1]);
      if (hipEventElapsedTime(&ms, pool[2 * k], pool[2 * k + 1]) == hipSuccess) {
        total_ms[kinds[k]] += ms;
        count[kinds[k]] += 1;
      }
    }
    used = 0;
    kinds.clear();
  }
  void reset() {
    harvest();
    for (int k = 0; k < kNumProf; k++) {
      total_ms[k] = 0;
      count[k] = 0;
    }
  }
  void release() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    pool.clear();
    used = 0;
    kinds.clear();
  }
  int begin(int kind, hipStream_t on_stream = nullptr) {
    if (!on || !((mask >> kind) & 1u)) return -1;
    if (used + 2 > 4096) harvest();
    if (used + 2 > pool.size()) {
      for (int k = 0; k < 64; k++) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        pool.push_back(e);
      }
    }
    const int id = (int)used;
    used += 2;
    kinds.push_back(kind);
    (void)hipEventRecord(pool[id], on_stream ? on_stream : stream);
    return id;
  }
  void end(int id, hipStream_t on_stream = nullptr) {
    if (id >= 0) (void)hipEventRecord(pool[id + 1], on_stream ? on_stream : stream);
  }
};

struct ProfScope {
  Prof* p;
  int id;
  hipStream_t s;
  ProfScope(Prof* pr, int kind, hipStream_t on_stream = nullptr)
      : p(pr), id(pr ? pr->begin(kind, on_stream) : -1), s(on_stream) {}
  ~ProfScope() {
    if (p) p->end(id, s);
  }
};

// a second stream for the convergence-test reduction, which runs beside the
// ILU sweep of the same PCG iteration (null: everything on one stream)
struct PcgAux {
  hipStream_t stream = nullptr;
  hipEvent_t updated = nullptr;   // r and r.*r written
  hipEvent_t tested = nullptr;    // (r, r) reduced and tested
};

struct Cloth;  // host handle (cs_api.cpp)

}  // namespace cs
