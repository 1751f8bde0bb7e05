// C-ABI of the cloth step (include/clothsim.h): handle lifetime, uploads,
// and the host orchestration of StepTime (self_contact_sparse/main.cpp:213-263).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/clothsim.h"
#include "cs_contact.hpp"
#include "cs_device.hpp"
#include "cs_setup.hpp"
#include "cs_solve.hpp"

namespace cs {

namespace {
thread_local std::string g_err;
}
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

namespace {
std::mutex g_mem_mu;
std::vector<signed char> g_pool;  // per device: 0 not yet asked, 1 pool, -1 plain
std::map<size_t, std::vector<void*>> g_pinned_free;  // free pinned blocks by size
std::map<void*, size_t> g_pinned_size;               // every pinned block's size

bool use_pool() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return false;
  std::lock_guard<std::mutex> lk(g_mem_mu);
  if ((int)g_pool.size() <= dev) g_pool.resize(dev + 1, 0);
  if (g_pool[dev] == 0) {
    g_pool[dev] = -1;
    int ok = 0;
    hipMemPool_t pool = nullptr;
    uint64_t keep = UINT64_MAX;  // freed memory stays in the pool
    if (hipDeviceGetAttribute(&ok, hipDeviceAttributeMemoryPoolsSupported, dev) == hipSuccess &&
        ok && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess &&
        hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess)
      g_pool[dev] = 1;
  }
  return g_pool[dev] == 1;
}
}  // namespace

int dev_alloc(void** p, size_t bytes, hipStream_t s) {
  *p = nullptr;
  if (use_pool())
    CS_HIP(hipMallocAsync(p, bytes, s));
  else
    CS_HIP(hipMalloc(p, bytes));
  return kOk;
}

void dev_free(void* p, hipStream_t s) {
  if (!p) return;
  if (use_pool())
    (void)hipFreeAsync(p, s);
  else
    (void)hipFree(p);
}

int pinned_alloc(void** p, size_t bytes) {
  size_t sz = 256;
  while (sz < bytes) sz <<= 1;
  void* q = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mem_mu);
    auto& fl = g_pinned_free[sz];
    if (!fl.empty()) {
      q = fl.back();
      fl.pop_back();
    }
  }
  if (!q) {
    CS_HIP(hipHostMalloc(&q, sz));
    std::lock_guard<std::mutex> lk(g_mem_mu);
    g_pinned_size[q] = sz;
  }
  std::memset(q, 0, sz);
  *p = q;
  return kOk;
}

void pinned_free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_mem_mu);
  auto it = g_pinned_size.find(p);
  if (it != g_pinned_size.end()) g_pinned_free[it->second].push_back(p);
}

// One device allocation carved into aligned sub-arrays.
class Arena {
 public:
  template <typename T>
  void reserve(T** slot, size_t count) {
    Item it;
    it.slot = (void**)slot;
    it.offset = total_;
    items_.push_back(it);
    total_ += ((count * sizeof(T) + 255) / 256) * 256;
    if (count == 0) total_ += 256;
  }
  int commit(hipStream_t s) {
    if (total_ == 0) return kOk;
    if (int st = dev_alloc(&base_, total_, s)) return st;
    for (auto& it : items_) *it.slot = (char*)base_ + it.offset;
    return kOk;
  }
  void release(hipStream_t s) {
    dev_free(base_, s);
    base_ = nullptr;
  }
  size_t bytes() const { return total_; }

 private:
  struct Item {
    void** slot;
    size_t offset;
  };
  std::vector<Item> items_;
  size_t total_ = 0;
  void* base_ = nullptr;
};

struct Cloth {
  MeshSetup S;
  Params P;
  DevCloth d;
  Arena arena;
  hipStream_t stream = nullptr;
  int* hflags = nullptr;  // pinned
  double* hscal = nullptr;
  int* ising = nullptr;   // per-row singular-pivot flags
  PcgAux aux;             // convergence-test reduction beside the sweep
  unsigned long long* sum_stats = nullptr;  // reduction counters (while profiling)
  ContactState contact;
  bool have_x1 = false;
  int poll = 8;
  Prof prof;
  double time = 0.0;  // simulated time, advanced by dt per cs_step
  // kinematic sweep: vertices placed at rest + delta * clamp((t+dt)/duration, 0, 1)
  int* sweep_verts = nullptr;
  int n_sweep = 0;
  double sweep_delta[3] = {0, 0, 0};
  double sweep_duration = 1.0;
};

namespace {

template <typename T>
int upload(T* dst, const std::vector<T>& v, hipStream_t s) {
  if (!v.empty()) CS_HIP(hipMemcpyAsync(dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
  return kOk;
}

Params to_params(const cs_params* p) {
  Params P;
  P.lam = p->lambda;
  P.mu = p->myu;
  P.stiff_bend = p->stiff_bend;
  for (int k = 0; k < 3; k++) P.grav[k] = p->gravity[k];
  P.mass = p->mass_point;
  P.k_contact = p->stiff_contact;
  P.clearance = p->contact_clearance;
  P.dt = p->dt;
  P.pcg_tol = p->pcg_tol;
  P.pcg_maxit = p->pcg_maxit;
  P.obstacle_kind = p->obstacle_kind;
  for (int k = 0; k < 8; k++) P.obstacle[k] = p->obstacle[k];
  return P;
}

int check_params(const cs_params* p) {
  if (!p) {
    set_error("params is NULL");
    return kBadArg;
  }
  if (!(p->dt > 0) || !(p->mass_point > 0) || p->pcg_maxit < 0 || !(p->pcg_tol >= 0)) {
    set_error("params need dt > 0, mass_point > 0, pcg_maxit >= 0, pcg_tol >= 0");
    return kBadArg;
  }
  if (p->obstacle_kind < 0 || p->obstacle_kind > CS_OBST_MAX) {
    set_error("unknown obstacle kind");
    return kBadArg;
  }
  return kOk;
}

int build(Cloth* c, const cs_mesh* m) {
  MeshSetup& S = c->S;
  std::string err = build_setup_cached(S, m->np, m->nt, m->nq, m->tri, m->quad, m->xyz);
  if (!err.empty()) {
    set_error(err);
    return kBadArg;
  }
  const int np = S.np, nt = S.nt, nq = S.nq, nnz = S.nnz, nn = S.nnode;
  DevCloth& d = c->d;
  std::memset(&d, 0, sizeof(d));
  d.np = np;
  d.nt = nt;
  d.nq = nq;
  d.nnz = nnz;
  d.nnode = nn;
  d.nlf = S.nlf;
  d.nlb = S.nlb;
  d.wmax = 0;
  for (int l = 0; l < S.nlf; l++) d.wmax = std::max(d.wmax, S.lf_ptr[l + 1] - S.lf_ptr[l]);
  for (int l = 0; l < S.nlb; l++) d.wmax = std::max(d.wmax, S.lb_ptr[l + 1] - S.lb_ptr[l]);
  Arena& A = c->arena;
  int *tri, *quad, *bc, *rowptr, *colind, *diaind, *rowof, *lfp, *lfr, *lbp, *lbr;
  int *fup, *fuk, *fuc;
  int *csp, *cs_, *dsp, *ds, *gsp, *gs, *rp, *rg, *node, *pre, *pree, *bypre, *leafof, *rord;
  int *rend, *lrank;
  double* xyz0;
  A.reserve(&tri, 3 * (size_t)nt);
  A.reserve(&quad, 4 * (size_t)nq);
  A.reserve(&bc, np);
  A.reserve(&xyz0, 3 * (size_t)np);
  A.reserve(&rowptr, np + 1);
  A.reserve(&colind, nnz);
  A.reserve(&diaind, np);
  A.reserve(&rowof, nnz);
  A.reserve(&fup, S.fu_ptr.size());
  A.reserve(&fuk, S.fu_kj.size());
  A.reserve(&fuc, S.fu_code.size());
  A.reserve(&lfp, S.lf_ptr.size());
  A.reserve(&lfr, S.lf_rows.size());
  A.reserve(&lbp, S.lb_ptr.size());
  A.reserve(&lbr, S.lb_rows.size());
  A.reserve(&csp, S.crs_src_ptr.size());
  A.reserve(&cs_, S.crs_src.size());
  A.reserve(&dsp, S.dia_src_ptr.size());
  A.reserve(&ds, S.dia_src.size());
  A.reserve(&gsp, S.grad_src_ptr.size());
  A.reserve(&gs, S.grad_src.size());
  A.reserve(&rp, S.ring_ptr.size());
  A.reserve(&rg, S.ring.size());
  A.reserve(&node, 3 * (size_t)nn);
  A.reserve(&pre, nn);
  A.reserve(&pree, nn);
  A.reserve(&bypre, nn);
  A.reserve(&leafof, nt);
  A.reserve(&rend, nn);
  A.reserve(&lrank, nt);
  A.reserve(&rord, nn);
  A.reserve(&d.x, 3 * (size_t)np);
  A.reserve(&d.v, 3 * (size_t)np);
  A.reserve(&d.x1, 3 * (size_t)np);
  A.reserve(&d.dia, 9 * (size_t)np);
  A.reserve(&d.crs, 9 * (size_t)nnz);
  A.reserve(&d.b, 3 * (size_t)np);
  A.reserve(&d.fdia, 9 * (size_t)np);
  A.reserve(&d.fcrs, 9 * (size_t)nnz);
  A.reserve(&d.etri, 90 * (size_t)nt);
  A.reserve(&d.kq, 16 * (size_t)nq);
  A.reserve(&d.eq, 12 * (size_t)nq);
  A.reserve(&d.r, 3 * (size_t)np);
  A.reserve(&d.dx, 3 * (size_t)np);
  A.reserve(&d.z, 3 * (size_t)np);
  A.reserve(&d.ap, 3 * (size_t)np);
  A.reserve(&d.p[0], 3 * (size_t)np);
  A.reserve(&d.p[1], 3 * (size_t)np);
  A.reserve(&d.prod, 3 * (size_t)np);
  const size_t ntile = (size_t)sum_tiles(3 * np);
  A.reserve(&d.sum.tsum, ntile);
  A.reserve(&d.sum.code, ntile);
  A.reserve(&d.sum.tot, ntile);
  A.reserve(&d.sum.pmin, ntile);
  A.reserve(&d.sum.pmax, ntile);
  A.reserve(&d.sum.ticket, 1);
  A.reserve(&d.prod2, 3 * (size_t)np);
  A.reserve(&d.sum2.tsum, ntile);
  A.reserve(&d.sum2.code, ntile);
  A.reserve(&d.sum2.tot, ntile);
  A.reserve(&d.sum2.pmin, ntile);
  A.reserve(&d.sum2.pmax, ntile);
  A.reserve(&d.sum2.ticket, 1);
  A.reserve(&c->sum_stats, kNumSumStat);
  A.reserve(&d.scal, kNumScal);
  A.reserve(&d.flags, kNumFlag);
  A.reserve(&d.bb, 6 * (size_t)nn);
  A.reserve(&d.rbb, 6 * (size_t)nn);
  A.reserve(&c->ising, np);
  int *frc, *brc, *fsl;
  unsigned short *fmeta, *bmeta;
  const int kl = S.kl > 0 ? S.kl : 0, ku = S.ku > 0 ? S.ku : 0;
  A.reserve(&frc, kl ? 4 * (size_t)np : 0);
  A.reserve(&fsl, kl ? (size_t)np : 0);
  A.reserve(&d.sb, 3 * (size_t)np);
  A.reserve(&fmeta, kl ? (size_t)meta_groups(std::max(kl, ku)) * 8 * np : 0);
  A.reserve(&brc, kl ? 2 * (size_t)np : 0);
  A.reserve(&bmeta, kl ? (size_t)meta_groups(std::max(kl, ku)) * 8 * np : 0);
  const int kq = kl > 0 ? std::max(kl, ku) : 0;
  // value pairs per lane: (3 kq + 1) / 2 pairs of 3 lanes x 2 doubles per slot
  const size_t pk = kq ? (size_t)(3 * kq + 1) / 2 * 6 * np : 0;
  A.reserve(&d.f_blk, pk);
  A.reserve(&d.f_dinv, kq ? 12 * (size_t)np : 0);
  A.reserve(&d.b_blk, pk);
  int *qfp, *qbp;
  A.reserve(&qfp, S.qf_ptr.size());
  A.reserve(&qbp, S.qb_ptr.size());
  int *qfc, *qbc;
  A.reserve(&qfc, S.qf_comp.size());
  A.reserve(&qbc, S.qb_comp.size());
  double* host_pd;
  A.reserve(&host_pd, 4 * (size_t)np);
  if (int st = A.commit(c->stream)) return st;
  d.tri = tri;
  d.quad = quad;
  d.bc = bc;
  d.xyz0 = xyz0;
  d.rowptr = rowptr;
  d.colind = colind;
  d.diaind = diaind;
  d.rowof = rowof;
  d.lf_ptr = lfp;
  d.lf_rows = lfr;
  d.fu_ptr = fup;
  d.fu_kj = fuk;
  d.fu_code = fuc;
  {
    int widest = 0;
    for (int i = 0; i < np; i++) widest = std::max(widest, S.rowptr[i + 1] - S.rowptr[i]);
    d.fu_blocks = ilu_stage_blocks(widest);
  }
  d.lb_ptr = lbp;
  d.lb_rows = lbr;
  d.crs_src_ptr = csp;
  d.crs_src = cs_;
  d.dia_src_ptr = dsp;
  d.dia_src = ds;
  d.grad_src_ptr = gsp;
  d.grad_src = gs;
  d.ring_ptr = rp;
  d.ring = rg;
  d.node = node;
  d.pre = pre;
  d.pre_end = pree;
  d.by_pre = bypre;
  d.leaf_of = leafof;
  d.rend = rend;
  d.leaf_rank = lrank;
  d.refit_order = rord;
  d.kl = S.kl;
  d.ku = S.ku;
  d.kpack = kq;
  d.ring_ok = S.ring_ok ? 1 : 0;
  d.f_rci = frc;
  d.fslot = kl ? fsl : nullptr;  // null: no packed sweep, no slot-order copy of r
  d.f_meta = fmeta;
  d.b_rc = brc;
  d.b_meta = bmeta;
  d.ising = c->ising;
  d.qf_ptr = qfp;
  d.qb_ptr = qbp;
  d.qf_comp = qfc;
  d.qb_comp = qbc;
  d.nqf = S.nqf;
  d.nqb = S.nqb;
  d.ncomp = S.ncomp;
  d.qwidth = S.qwidth;
  hipStream_t s = c->stream;
  CS_HIP(hipMemsetAsync(d.sum.ticket, 0, sizeof(unsigned int), s));
  CS_HIP(hipMemsetAsync(d.sum2.ticket, 0, sizeof(unsigned int), s));
  CS_HIP(hipMemcpyAsync(tri, m->tri, sizeof(int) * 3 * nt, hipMemcpyHostToDevice, s));
  if (nq) CS_HIP(hipMemcpyAsync(quad, m->quad, sizeof(int) * 4 * nq, hipMemcpyHostToDevice, s));
  std::vector<int> bcv(np, 0);
  if (m->bc)
    for (int i = 0; i < np; i++) bcv[i] = m->bc[i] != 0;
  if (int st = upload(bc, bcv, s)) return st;
  CS_HIP(hipMemcpyAsync(xyz0, m->xyz0, sizeof(double) * 3 * np, hipMemcpyHostToDevice, s));
  if (int st = upload(rowptr, S.rowptr, s)) return st;
  if (int st = upload(colind, S.colind, s)) return st;
  if (int st = upload(diaind, S.diaind, s)) return st;
  if (int st = upload(rowof, S.rowof, s)) return st;
  if (int st = upload(fup, S.fu_ptr, s)) return st;
  if (int st = upload(fuk, S.fu_kj, s)) return st;
  if (int st = upload(fuc, S.fu_code, s)) return st;
  if (int st = upload(lfp, S.lf_ptr, s)) return st;
  if (int st = upload(lfr, S.lf_rows, s)) return st;
  if (int st = upload(lbp, S.lb_ptr, s)) return st;
  if (int st = upload(lbr, S.lb_rows, s)) return st;
  if (int st = upload(csp, S.crs_src_ptr, s)) return st;
  if (int st = upload(cs_, S.crs_src, s)) return st;
  if (int st = upload(dsp, S.dia_src_ptr, s)) return st;
  if (int st = upload(ds, S.dia_src, s)) return st;
  if (int st = upload(gsp, S.grad_src_ptr, s)) return st;
  if (int st = upload(gs, S.grad_src, s)) return st;
  if (int st = upload(rp, S.ring_ptr, s)) return st;
  if (int st = upload(rg, S.ring, s)) return st;
  if (int st = upload(node, S.node, s)) return st;
  if (int st = upload(pre, S.pre, s)) return st;
  if (int st = upload(pree, S.pre_end, s)) return st;
  if (int st = upload(bypre, S.by_pre, s)) return st;
  if (int st = upload(leafof, S.leaf_of, s)) return st;
  // by pre-order rank: subtree ends and the leaves (held until the stream
  // synchronises below, as the copies read them)
  std::vector<int> rank_end(nn), leaf_ranks;
  leaf_ranks.reserve(nt);
  for (int r = 0; r < nn; r++) {
    rank_end[r] = S.pre_end[S.by_pre[r]];
    if (rank_end[r] == r + 1) leaf_ranks.push_back(r);
  }
  if ((int)leaf_ranks.size() != nt) {
    set_error("cs_cloth_create: the BVH does not have one leaf per triangle");
    return kBadArg;
  }
  if (int st = upload(rend, rank_end, s)) return st;
  if (int st = upload(lrank, leaf_ranks, s)) return st;
  if (int st = upload(rord, S.refit_order, s)) return st;
  if (S.kl > 0) {
    if (int st = upload(qfp, S.qf_ptr, s)) return st;
    if (int st = upload(qfc, S.qf_comp, s)) return st;
    if (int st = upload(qbc, S.qb_comp, s)) return st;
    if (int st = upload(qbp, S.qb_ptr, s)) return st;
    if (int st = upload(frc, S.f_rci, s)) return st;
    if (int st = upload(fsl, S.fslot, s)) return st;
    if (int st = upload(fmeta, S.f_meta, s)) return st;
    if (int st = upload(brc, S.b_rc, s)) return st;
    if (int st = upload(bmeta, S.b_meta, s)) return st;
  }
  CS_HIP(hipMemcpyAsync(d.x, m->xyz, sizeof(double) * 3 * np, hipMemcpyHostToDevice, s));
  if (m->uvw)
    CS_HIP(hipMemcpyAsync(d.v, m->uvw, sizeof(double) * 3 * np, hipMemcpyHostToDevice, s));
  else
    CS_HIP(hipMemsetAsync(d.v, 0, sizeof(double) * 3 * np, s));
  CS_HIP(hipMemsetAsync(d.flags, 0, sizeof(int) * kNumFlag, s));
  CS_HIP(hipMemsetAsync(d.scal, 0, sizeof(double) * kNumScal, s));
  c->contact.host_pd_buf = host_pd;
  if (int st = launch_bend_setup(d, c->P, s)) return st;
  if (int st = contact_init(c->contact, S, d, s)) return st;
  CS_HIP(hipStreamSynchronize(s));
  return kOk;
}

Cloth* H(cs_cloth* p) { return reinterpret_cast<Cloth*>(p); }
const Cloth* H(const cs_cloth* p) { return reinterpret_cast<const Cloth*>(p); }

int sync(Cloth* c) {
  CS_HIP(hipStreamSynchronize(c->stream));
  return kOk;
}

int check_tickets(const int* hflags) {
  if (hflags[kTicketErr]) {
    set_error("a PCG reduction ended with its ticket nonzero (inconsistent skip flag)");
    return kHipError;
  }
  return kOk;
}

// StepTime_InternalDynamicsILU on the resident state
int internal_step(Cloth* c, double* conv, int* iters) {
  DevCloth& d = c->d;
  hipStream_t s = c->stream;
  CS_HIP(hipMemcpyAsync(d.x1, d.x, sizeof(double) * 3 * d.np, hipMemcpyDeviceToDevice, s));
  c->have_x1 = true;
  Prof* prof = c->prof.on ? &c->prof : nullptr;
  if (int st = launch_assemble(d, c->P, s, prof)) return st;
  if (int st = launch_ilu(d, s, prof)) return st;
  if (int st = run_pcg(d, c->P.pcg_tol, c->P.pcg_maxit, c->hflags, s, c->poll, prof, &c->aux))
    return st;
  if (int st = launch_finish_internal(d, c->P, s)) return st;
  CS_HIP(hipMemcpyAsync(c->hflags, d.flags, sizeof(int) * kNumFlag, hipMemcpyDeviceToHost, s));
  CS_HIP(hipMemcpyAsync(c->hscal, d.scal, sizeof(double) * kNumScal, hipMemcpyDeviceToHost, s));
  CS_HIP(hipStreamSynchronize(s));
  if (int st = check_tickets(c->hflags)) return st;
  if (conv) *conv = c->hscal[kConv];
  if (iters) *iters = c->hflags[kDone] ? c->hflags[kIters] : c->P.pcg_maxit;
  if (c->hflags[kFlagNonFinite]) {
    set_error("the internal solve produced an infinite or NaN displacement (the ILU(0)-PCG "
              "iterate overflowed)");
    return kNonFinite;
  }
  return kOk;
}

}  // namespace
}  // namespace cs

using namespace cs;

extern "C" {

const char* cs_last_error(void) { return last_error(); }

int cs_device_count(int* n) {
  int k = 0;
  hipError_t e = hipGetDeviceCount(&k);
  if (e != hipSuccess) k = 0;
  if (n) *n = k;
  return k > 0 ? kOk : kNoDevice;
}

int cs_cloth_create(const cs_mesh* mesh, const cs_params* params, cs_cloth** out) {
  if (!mesh || !out || !mesh->xyz0 || !mesh->xyz || !mesh->tri || (mesh->nq > 0 && !mesh->quad)) {
    set_error("cs_cloth_create: missing mesh arrays");
    return kBadArg;
  }
  if (int st = check_params(params)) return st;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error("no HIP device");
    return kNoDevice;
  }
  Cloth* c = new (std::nothrow) Cloth();
  if (!c) {
    set_error("host allocation failed");
    return kOutOfMemory;
  }
  c->P = to_params(params);
  int st = kOk;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->aux.stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->aux.updated, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->aux.tested, hipEventDisableTiming) != hipSuccess ||
      pinned_alloc((void**)&c->hflags, sizeof(int) * kNumFlag) != kOk ||
      pinned_alloc((void**)&c->hscal, sizeof(double) * kNumScal) != kOk) {
    set_error("stream or pinned buffer creation failed");
    st = kHipError;
  }
  if (st == kOk) st = build(c, mesh);
  if (st != kOk) {
    cs_cloth_destroy(reinterpret_cast<cs_cloth*>(c));
    return st;
  }
  *out = reinterpret_cast<cs_cloth*>(c);
  return kOk;
}

int cs_cloth_destroy(cs_cloth* p) {
  if (!p) return kOk;
  Cloth* c = H(p);
  // every stream that used the cloth's memory is idle before it is released
  // (the device memory after the work queued on the cloth's stream)
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->aux.stream) (void)hipStreamSynchronize(c->aux.stream);
  contact_release(c->contact, c->stream);
  c->prof.release();
  dev_free(c->sweep_verts, c->stream);
  c->arena.release(c->stream);
  pinned_free(c->hflags);
  pinned_free(c->hscal);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->aux.stream) (void)hipStreamDestroy(c->aux.stream);
  if (c->aux.updated) (void)hipEventDestroy(c->aux.updated);
  if (c->aux.tested) (void)hipEventDestroy(c->aux.tested);
  delete c;
  return kOk;
}

int cs_cloth_set_params(cs_cloth* p, const cs_params* params) {
  if (!p) return kBadArg;
  if (int st = check_params(params)) return st;
  Cloth* c = H(p);
  const double old_bend = c->P.stiff_bend;
  c->P = to_params(params);
  if (c->P.stiff_bend != old_bend) {
    if (int st = launch_bend_setup(c->d, c->P, c->stream)) return st;
    return sync(c);
  }
  return kOk;
}

int cs_cloth_memory(const cs_cloth* p, long long* bytes) {
  if (!p || !bytes) return kBadArg;
  const Cloth* c = H(p);
  const ContactState& k = c->contact;
  // the arena, the contact records at their current capacity (keys, tests,
  // leaf pairs, two sort buffers, groups, the set, impulse links), sort
  // scratch and the per-vertex contact arrays
  *bytes = (long long)c->arena.bytes() + (long long)k.cap * 88 + (long long)k.tmp_bytes +
           (long long)c->S.np * 64;
  return kOk;
}

int cs_cloth_sizes(const cs_cloth* p, int* np, int* nt, int* nq, int* nnz, int* nnode) {
  if (!p) return kBadArg;
  const Cloth* c = H(p);
  if (np) *np = c->S.np;
  if (nt) *nt = c->S.nt;
  if (nq) *nq = c->S.nq;
  if (nnz) *nnz = c->S.nnz;
  if (nnode) *nnode = c->S.nnode;
  return kOk;
}

int cs_set_state(cs_cloth* p, const double* xyz, const double* uvw) {
  if (!p) return kBadArg;
  Cloth* c = H(p);
  const size_t n = sizeof(double) * 3 * c->S.np;
  if (xyz) CS_HIP(hipMemcpyAsync(c->d.x, xyz, n, hipMemcpyHostToDevice, c->stream));
  if (uvw) CS_HIP(hipMemcpyAsync(c->d.v, uvw, n, hipMemcpyHostToDevice, c->stream));
  return sync(c);
}

int cs_get_state(cs_cloth* p, double* xyz, double* uvw) {
  if (!p) return kBadArg;
  Cloth* c = H(p);
  const size_t n = sizeof(double) * 3 * c->S.np;
  if (xyz) CS_HIP(hipMemcpyAsync(xyz, c->d.x, n, hipMemcpyDeviceToHost, c->stream));
  if (uvw) CS_HIP(hipMemcpyAsync(uvw, c->d.v, n, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int cs_state_device(cs_cloth* p, double** xyz, double** uvw) {
  if (!p) return kBadArg;
  Cloth* c = H(p);
  if (xyz) *xyz = c->d.x;
  if (uvw) *uvw = c->d.v;
  return kOk;
}

int cs_set_step_start(cs_cloth* p, const double* xyz1) {
  if (!p || !xyz1) return kBadArg;
  Cloth* c = H(p);
  CS_HIP(hipMemcpyAsync(c->d.x1, xyz1, sizeof(double) * 3 * c->S.np, hipMemcpyHostToDevice,
                        c->stream));
  c->have_x1 = true;
  return sync(c);
}

int cs_set_obstacle_values(cs_cloth* p, const double* pd, const double* n) {
  if (!p) return kBadArg;
  Cloth* c = H(p);
  if (!pd || !n) {
    c->d.host_pd = nullptr;
    c->d.host_n = nullptr;
    return kOk;
  }
  const int np = c->S.np;
  double* buf = c->contact.host_pd_buf;
  CS_HIP(hipMemcpyAsync(buf, pd, sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
  CS_HIP(hipMemcpyAsync(buf + np, n, sizeof(double) * 3 * np, hipMemcpyHostToDevice, c->stream));
  c->d.host_pd = buf;
  c->d.host_n = buf + np;
  return sync(c);
}

int cs_step_internal(cs_cloth* p, double* conv_ratio, int* iteration) {
  if (!p) return kBadArg;
  return internal_step(H(p), conv_ratio, iteration);
}

int cs_contact_resolve(cs_cloth* p, int* impulse_applied, cs_step_stats* stats) {
  if (!p) return kBadArg;
  Cloth* c = H(p);
  if (!c->have_x1) {
    set_error("cs_contact_resolve needs the start-of-step positions of cs_step_internal");
    return kBadArg;
  }
  int applied = 0;
  if (int st = contact_resolve(c->contact, c->S, c->d, c->P, c->stream, &applied, stats)) return st;
  if (impulse_applied) *impulse_applied = applied;
  return kOk;
}

int cs_post_contact_update(cs_cloth* p) {
  if (!p) return kBadArg;
  Cloth* c = H(p);
  if (int st = post_contact_update(c->d, c->P, c->stream)) return st;
  return sync(c);
}

int cs_update_velocity(cs_cloth* p) {
  if (!p) return kBadArg;
  Cloth* c = H(p);
  if (int st = update_velocity(c->d, c->P, c->stream)) return st;
  return sync(c);
}

int cs_step(cs_cloth* p, int n, cs_step_stats* stats) {
  if (!p || n < 0) return kBadArg;
  Cloth* c = H(p);
  if (c->d.host_pd) {
    set_error("cs_step: host obstacle values describe one state; use cs_step_internal");
    return kBadArg;
  }
  for (int k = 0; k < n; k++) {
    cs_step_stats st_local;
    std::memset(&st_local, 0, sizeof(st_local));
    double conv = 0;
    int iters = 0;
    if (c->n_sweep > 0) {
      double sv = (c->time + c->P.dt) / c->sweep_duration;
      sv = sv < 0.0 ? 0.0 : (sv > 1.0 ? 1.0 : sv);
      const double off[3] = {c->sweep_delta[0] * sv, c->sweep_delta[1] * sv, c->sweep_delta[2] * sv};
      if (int st = launch_apply_sweep(c->d, c->sweep_verts, c->n_sweep, off, c->stream)) return st;
    }
    if (int st = internal_step(c, &conv, &iters)) {
      if (st == kNonFinite) {
        c->time += c->P.dt;
        st_local.pcg_conv = conv;
        st_local.pcg_iters = iters;
        st_local.ilu_singular = c->hflags[kNsing];
        st_local.status = kNonFinite;
        if (stats) *stats = st_local;
      }
      return st;
    }
    int applied = 0;
    Prof* prof = c->prof.on ? &c->prof : nullptr;
    if (int st = contact_resolve(c->contact, c->S, c->d, c->P, c->stream, &applied, &st_local, prof))
      return st;
    if (applied) {
      if (int st = post_contact_update(c->d, c->P, c->stream)) return st;
    }
    st_local.pcg_conv = conv;
    st_local.pcg_iters = iters;
    st_local.ilu_singular = c->hflags[kNsing];
    st_local.impulse_applied = applied;
    if (stats) *stats = st_local;
    c->time += c->P.dt;
  }
  return sync(c);
}

int cs_set_sweep(cs_cloth* p, int n, const int* verts, const double* delta, double duration) {
  if (!p || n < 0 || (n > 0 && (!verts || !delta)) || !(duration > 0)) return kBadArg;
  Cloth* c = H(p);
  for (int i = 0; i < n; i++)
    if (verts[i] < 0 || verts[i] >= c->S.np) {
      set_error("sweep vertex out of range");
      return kBadArg;
    }
  CS_HIP(hipStreamSynchronize(c->stream));  // no step still reads the old list
  dev_free(c->sweep_verts, c->stream);
  c->sweep_verts = nullptr;
  c->n_sweep = 0;
  if (n > 0) {
    if (int st = dev_alloc((void**)&c->sweep_verts, sizeof(int) * n, c->stream)) return st;
    CS_HIP(hipMemcpyAsync(c->sweep_verts, verts, sizeof(int) * n, hipMemcpyHostToDevice,
                          c->stream));
    CS_HIP(hipStreamSynchronize(c->stream));
  }
  c->n_sweep = n;
  for (int k = 0; k < 3; k++) c->sweep_delta[k] = n > 0 ? delta[k] : 0.0;
  c->sweep_duration = duration;
  return kOk;
}

int cs_set_option(cs_cloth* p, int option, int value) {
  if (!p) return kBadArg;
  Cloth* c = H(p);
  if (option == CS_OPT_SWEEP && value >= 0 && value <= 1) {
    if (value == 1 && c->d.precision != 0) {
      set_error("the row-by-row sweep is FP64 only: set CS_OPT_PRECISION to 0 first");
      return kBadArg;
    }
    c->d.sweep_variant = value;
    return kOk;
  }
  if (option == CS_OPT_PCG && value >= 0 && value <= 2) {
    if (value == 2 && !pcg_wg_ok(c->d)) {
      set_error("the one-workgroup PCG solve needs the packed sweep, FP64 and 3 np <= 65536");
      return kBadArg;
    }
    c->d.pcg_variant = value;
    return kOk;
  }
  if (option == CS_OPT_PRECISION && value >= 0 && value <= 1) {
    if (value == 1 && (!quad_ok(c->d) || c->d.sweep_variant == 1 || c->d.pcg_variant == 2)) {
      set_error("mixed precision needs the packed sweep and the multi-launch PCG solve");
      return kBadArg;
    }
    c->d.precision = value;
    // the packed coefficients change format: repack the current factor
    if (int st = launch_pack_factor(c->d, c->stream)) return st;
    return sync(c);
  }
  set_error("unknown option or value");
  return kBadArg;
}

int cs_set_time(cs_cloth* p, double t) {
  if (!p) return kBadArg;
  H(p)->time = t;
  return kOk;
}

int cs_get_time(const cs_cloth* p, double* t) {
  if (!p || !t) return kBadArg;
  *t = H(p)->time;
  return kOk;
}

int cs_profile_enable(cs_cloth* p, int on) {
  if (!p || on < CS_PROFILE_OFF || on > CS_PROFILE_SWEEP) return kBadArg;
  Cloth* c = H(p);
  c->prof.stream = c->stream;
  c->prof.reset();
  c->prof.on = on != CS_PROFILE_OFF;
  c->prof.mask = on == CS_PROFILE_SWEEP ? 1u << kProfSweep : ~0u;
  CS_HIP(hipMemsetAsync(c->sum_stats, 0, sizeof(unsigned long long) * kNumSumStat, c->stream));
  c->d.sum.stats = on == CS_PROFILE_ALL ? c->sum_stats : nullptr;
  c->d.sum2.stats = c->d.sum.stats;
  return sync(c);
}

int cs_reduction_stats(cs_cloth* p, long long* out) {
  if (!p || !out) return kBadArg;
  Cloth* c = H(p);
  unsigned long long h[kNumSumStat];
  CS_HIP(hipMemcpyAsync(h, c->sum_stats, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  if (int st = sync(c)) return st;
  for (int k = 0; k < kStatChainTicks; k++) out[k] = (long long)h[k];
  // wall-clock ticks to nanoseconds (the rate is in kHz)
  int dev = 0, khz = 0;
  CS_HIP(hipGetDevice(&dev));
  CS_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  for (int k = kStatChainTicks; k < kNumSumStat; k++)
    out[k] = khz > 0 ? (long long)((double)h[k] * 1.0e6 / khz) : 0;
  return kOk;
}

int cs_profile_read(cs_cloth* p, int kind, double* total_ms, long long* launches) {
  if (!p || kind < 0 || kind >= kNumProf) return kBadArg;
  Cloth* c = H(p);
  c->prof.harvest();
  if (total_ms) *total_ms = c->prof.total_ms[kind];
  if (launches) *launches = c->prof.count[kind];
  return kOk;
}

int cs_get_pattern(const cs_cloth* p, int* rowptr, int* colind, int* diaind) {
  if (!p) return kBadArg;
  const Cloth* c = H(p);
  if (rowptr) std::memcpy(rowptr, c->S.rowptr.data(), sizeof(int) * (c->S.np + 1));
  if (colind) std::memcpy(colind, c->S.colind.data(), sizeof(int) * c->S.nnz);
  if (diaind) std::memcpy(diaind, c->S.diaind.data(), sizeof(int) * c->S.np);
  return kOk;
}

int cs_assemble(cs_cloth* p, const double* xyz, const double* uvw) {
  if (!p) return kBadArg;
  Cloth* c = H(p);
  const size_t n = sizeof(double) * 3 * c->S.np;
  if (xyz) CS_HIP(hipMemcpyAsync(c->d.x, xyz, n, hipMemcpyHostToDevice, c->stream));
  if (uvw) CS_HIP(hipMemcpyAsync(c->d.v, uvw, n, hipMemcpyHostToDevice, c->stream));
  if (int st = launch_assemble(c->d, c->P, c->stream)) return st;
  return sync(c);
}

int cs_get_system(cs_cloth* p, double* dia, double* crs, double* b) {
  if (!p) return kBadArg;
  Cloth* c = H(p);
  hipStream_t s = c->stream;
  if (dia) CS_HIP(hipMemcpyAsync(dia, c->d.dia, sizeof(double) * 9 * c->S.np, hipMemcpyDeviceToHost, s));
  if (crs && c->S.nnz)
    CS_HIP(hipMemcpyAsync(crs, c->d.crs, sizeof(double) * 9 * c->S.nnz, hipMemcpyDeviceToHost, s));
  if (b) CS_HIP(hipMemcpyAsync(b, c->d.b, sizeof(double) * 3 * c->S.np, hipMemcpyDeviceToHost, s));
  return sync(c);
}

int cs_set_system(cs_cloth* p, const double* dia, const double* crs, const double* b) {
  if (!p) return kBadArg;
  Cloth* c = H(p);
  hipStream_t s = c->stream;
  if (dia) CS_HIP(hipMemcpyAsync(c->d.dia, dia, sizeof(double) * 9 * c->S.np, hipMemcpyHostToDevice, s));
  if (crs && c->S.nnz)
    CS_HIP(hipMemcpyAsync(c->d.crs, crs, sizeof(double) * 9 * c->S.nnz, hipMemcpyHostToDevice, s));
  if (b) CS_HIP(hipMemcpyAsync(c->d.b, b, sizeof(double) * 3 * c->S.np, hipMemcpyHostToDevice, s));
  return sync(c);
}

int cs_ilu_factor(cs_cloth* p, int* n_singular) {
  if (!p) return kBadArg;
  Cloth* c = H(p);
  CS_HIP(hipMemsetAsync(c->d.flags, 0, sizeof(int) * kNumFlag, c->stream));
  if (int st = launch_ilu(c->d, c->stream)) return st;
  CS_HIP(hipMemcpyAsync(c->hflags, c->d.flags, sizeof(int) * kNumFlag, hipMemcpyDeviceToHost, c->stream));
  if (int st = sync(c)) return st;
  if (n_singular) *n_singular = c->hflags[kNsing];
  return kOk;
}

int cs_get_factor(cs_cloth* p, double* fdia, double* fcrs) {
  if (!p) return kBadArg;
  Cloth* c = H(p);
  hipStream_t s = c->stream;
  if (fdia) CS_HIP(hipMemcpyAsync(fdia, c->d.fdia, sizeof(double) * 9 * c->S.np, hipMemcpyDeviceToHost, s));
  if (fcrs && c->S.nnz)
    CS_HIP(hipMemcpyAsync(fcrs, c->d.fcrs, sizeof(double) * 9 * c->S.nnz, hipMemcpyDeviceToHost, s));
  return sync(c);
}

int cs_set_factor(cs_cloth* p, const double* fdia, const double* fcrs) {
  if (!p) return kBadArg;
  Cloth* c = H(p);
  hipStream_t s = c->stream;
  if (fdia) CS_HIP(hipMemcpyAsync(c->d.fdia, fdia, sizeof(double) * 9 * c->S.np, hipMemcpyHostToDevice, s));
  if (fcrs && c->S.nnz)
    CS_HIP(hipMemcpyAsync(c->d.fcrs, fcrs, sizeof(double) * 9 * c->S.nnz, hipMemcpyHostToDevice, s));
  if (int st = launch_pack_factor(c->d, s)) return st;
  return sync(c);
}

int cs_ilu_solve(cs_cloth* p, double* vec) {
  if (!p || !vec) return kBadArg;
  Cloth* c = H(p);
  hipStream_t s = c->stream;
  const size_t n = sizeof(double) * 3 * c->S.np;
  CS_HIP(hipMemsetAsync(c->d.flags, 0, sizeof(int) * kNumFlag, s));
  CS_HIP(hipMemcpyAsync(c->d.r, vec, n, hipMemcpyHostToDevice, s));
  if (int st = launch_ilu_apply(c->d, c->d.r, c->d.z, s)) return st;
  CS_HIP(hipMemcpyAsync(vec, c->d.z, n, hipMemcpyDeviceToHost, s));
  return sync(c);
}

int cs_matvec(cs_cloth* p, double alpha, const double* x, double beta, double* y) {
  if (!p || !x || !y) return kBadArg;
  Cloth* c = H(p);
  hipStream_t s = c->stream;
  const size_t n = sizeof(double) * 3 * c->S.np;
  CS_HIP(hipMemcpyAsync(c->d.r, x, n, hipMemcpyHostToDevice, s));
  CS_HIP(hipMemcpyAsync(c->d.z, y, n, hipMemcpyHostToDevice, s));
  if (int st = launch_matvec(c->d, alpha, c->d.r, beta, c->d.z, s)) return st;
  CS_HIP(hipMemcpyAsync(y, c->d.z, n, hipMemcpyDeviceToHost, s));
  return sync(c);
}

int cs_pcg_solve(cs_cloth* p, const double* b, double tol, int maxit, double* x, double* r,
                 double* conv, int* iters) {
  if (!p || !b || maxit < 0) return kBadArg;
  Cloth* c = H(p);
  hipStream_t s = c->stream;
  const size_t n = sizeof(double) * 3 * c->S.np;
  CS_HIP(hipMemcpyAsync(c->d.b, b, n, hipMemcpyHostToDevice, s));
  if (int st = run_pcg(c->d, tol, maxit, c->hflags, s, c->poll, nullptr, &c->aux)) return st;
  if (x) CS_HIP(hipMemcpyAsync(x, c->d.dx, n, hipMemcpyDeviceToHost, s));
  if (r) CS_HIP(hipMemcpyAsync(r, c->d.r, n, hipMemcpyDeviceToHost, s));
  CS_HIP(hipMemcpyAsync(c->hflags, c->d.flags, sizeof(int) * kNumFlag, hipMemcpyDeviceToHost, s));
  CS_HIP(hipMemcpyAsync(c->hscal, c->d.scal, sizeof(double) * kNumScal, hipMemcpyDeviceToHost, s));
  if (int st = sync(c)) return st;
  if (int st = check_tickets(c->hflags)) return st;
  if (conv) *conv = c->hscal[kConv];
  if (iters) *iters = c->hflags[kDone] ? c->hflags[kIters] : maxit;
  return kOk;
}

int cs_inspect_setup(const cs_mesh* m, int* sizes, int* rowptr, int* colind, int* diaind,
                     int* nodes) {
  if (!m || !m->xyz || !m->tri || (m->nq > 0 && !m->quad)) {
    set_error("cs_inspect_setup: missing mesh arrays");
    return kBadArg;
  }
  MeshSetup S;
  std::string err = build_setup_cached(S, m->np, m->nt, m->nq, m->tri, m->quad, m->xyz);
  if (!err.empty()) {
    set_error(err);
    return kBadArg;
  }
  if (sizes) {
    sizes[0] = S.np;
    sizes[1] = S.nt;
    sizes[2] = S.nq;
    sizes[3] = S.nnz;
    sizes[4] = S.nnode;
    sizes[5] = S.nlf;
    sizes[6] = S.nlb;
    sizes[7] = S.kl;
    sizes[8] = S.ku;
    sizes[9] = S.ring_ok ? 1 : 0;
  }
  if (rowptr) std::memcpy(rowptr, S.rowptr.data(), sizeof(int) * (S.np + 1));
  if (colind) std::memcpy(colind, S.colind.data(), sizeof(int) * S.nnz);
  if (diaind) std::memcpy(diaind, S.diaind.data(), sizeof(int) * S.np);
  if (nodes) std::memcpy(nodes, S.node.data(), sizeof(int) * 3 * S.nnode);
  return kOk;
}

// vector-load sets one wave of a quad sweep workgroup issues over the
// sub-levels q0 .. q1 - 1: three before the loop, then three per group of
// three sub-levels, except for the groups an idle wave passes without loading
// (it loads only while the next group has a row of its own)
static long long sweep_wave_sets(const std::vector<int>& qptr, int q0, int q1, int wave) {
  auto idle = [&](int l) { return l >= q1 || qptr[l] + 16 * wave >= qptr[l + 1]; };
  auto idle3 = [&](int l) { return idle(l) && idle(l + 1) && idle(l + 2); };
  long long sets = 3;
  for (int l = q0; l < q1; l += 3) {
    if (idle3(l))
      while (l + 3 < q1 && idle3(l + 3)) l += 3;
    sets += 3;
  }
  return sets;
}

int cs_sweep_shape(const cs_cloth* p, int* out) {
  if (!p || !out) return kBadArg;
  const Cloth* c = H(p);
  for (int i = 0; i < 8; i++) out[i] = 0;
  out[5] = c->d.ncomp;
  if (!sweep_shape(c->d, &out[2], &out[3], &out[4])) return kOk;
  long long most = 0;
  for (int k = 0; k < c->S.ncomp; k++) {
    const int f0 = c->S.qf_comp[k], f1 = c->S.qf_comp[k + 1];
    const int b0 = c->S.qb_comp[k], b1 = c->S.qb_comp[k + 1];
    out[0] = std::max(out[0], f1 - f0);
    out[1] = std::max(out[1], b1 - b0);
    long long loads = 0;
    for (int w = 0; w < out[2]; w++)
      loads += sweep_wave_sets(c->S.qf_ptr, f0, f1, w) * out[3] +
               sweep_wave_sets(c->S.qb_ptr, b0, b1, w) * out[4];
    most = std::max(most, loads);
  }
  out[6] = 1;
  out[7] = (int)std::min<long long>(most, INT_MAX);
  return kOk;
}

int cs_get_bvh(const cs_cloth* p, int* nodes) {
  if (!p || !nodes) return kBadArg;
  const Cloth* c = H(p);
  std::memcpy(nodes, c->S.node.data(), sizeof(int) * 3 * c->S.nnode);
  return kOk;
}

int cs_bvh_refit_prx(cs_cloth* p, const double* xyz, double delta, double* bb) {
  if (!p || !xyz) return kBadArg;
  Cloth* c = H(p);
  return contact_refit_host(c->contact, c->S, c->d, c->stream, xyz, nullptr, delta, 0.0, bb);
}

int cs_bvh_refit_ccd(cs_cloth* p, const double* xyz, const double* uvw, double dt, double* bb) {
  if (!p || !xyz || !uvw) return kBadArg;
  Cloth* c = H(p);
  return contact_refit_host(c->contact, c->S, c->d, c->stream, xyz, uvw, 0.0, dt, bb);
}

int cs_query_proximity(cs_cloth* p, const double* xyz, double delta, int* out, int cap, int* n) {
  if (!p || !xyz || cap < 0) return kBadArg;
  Cloth* c = H(p);
  return contact_query_host(c->contact, c->S, c->d, c->stream, xyz, nullptr, delta, 0.0, out, cap, n);
}

int cs_query_ccd(cs_cloth* p, const double* xyz, const double* uvw, double dt, double delta,
                 int* out, int cap, int* n) {
  if (!p || !xyz || !uvw || cap < 0) return kBadArg;
  Cloth* c = H(p);
  return contact_query_host(c->contact, c->S, c->d, c->stream, xyz, uvw, delta, dt, out, cap, n);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batches: many cloths stepped by shared launches

namespace cs {

struct Batch {
  std::vector<Cloth*> cl;
  DevCloth* dtab = nullptr;  // device copies of the cloths' DevCloth / Params
  Params* ptab = nullptr;
  int2* rowmap = nullptr;  // the batched SpMV's (cloth, first row) per block
  int2* vecmap = nullptr;  // the batched vector update's (cloth, first entry) per block
  int* dgather = nullptr;
  int* hgather = nullptr;  // pinned
  hipStream_t stream = nullptr;
  PcgAux aux;  // convergence-test reductions beside the sweeps
  BatchView view{};
  double tol = 0.0;
  int maxit = 0;
  int poll = 8;
  std::vector<DevCloth> htab;
  std::vector<Params> hptab;
  hipEvent_t internal_done = nullptr;    // the batched internal step of this step
  std::vector<hipEvent_t> cloth_done;    // each cloth's contact pass and update
  Prof prof;                             // per-kernel-class timing of the batched launches
};

// host threads that run the contact passes of a batch's cloths
constexpr int kContactThreads = 8;

// fn(i) for i in [0, n) on up to `threads` host threads; returns the status
// of the first failing index (its message becomes this thread's error)
template <class F>
int parallel_for(int n, int threads, F&& fn) {
  if (n <= 0) return kOk;
  const int t = std::max(1, std::min(threads, n));
  std::vector<int> status(n, kOk);
  std::vector<std::string> msg(n);
  std::atomic<int> next{0};
  auto work = [&]() {
    for (int i = next++; i < n; i = next++) {
      status[i] = fn(i);
      if (status[i] != kOk) msg[i] = last_error();
    }
  };
  std::vector<std::thread> pool;
  for (int k = 1; k < t; k++) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  for (int i = 0; i < n; i++)
    if (status[i] != kOk) {
      set_error(msg[i]);
      return status[i];
    }
  return kOk;
}

namespace {

Batch* HB(cs_batch* p) { return reinterpret_cast<Batch*>(p); }

void batch_free(Batch* b) {
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->aux.stream) {
    (void)hipStreamSynchronize(b->aux.stream);
    (void)hipStreamDestroy(b->aux.stream);
  }
  if (b->aux.updated) (void)hipEventDestroy(b->aux.updated);
  if (b->aux.tested) (void)hipEventDestroy(b->aux.tested);
  dev_free(b->dtab, b->stream);
  dev_free(b->ptab, b->stream);
  dev_free(b->rowmap, b->stream);
  dev_free(b->vecmap, b->stream);
  dev_free(b->dgather, b->stream);
  pinned_free(b->hgather);
  if (b->internal_done) (void)hipEventDestroy(b->internal_done);
  for (hipEvent_t e : b->cloth_done) (void)hipEventDestroy(e);
  b->prof.release();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// the member cloths' current device views and parameters into the tables
int batch_refresh(Batch* b) {
  for (size_t i = 0; i < b->cl.size(); i++) {
    b->htab[i] = b->cl[i]->d;
    b->hptab[i] = b->cl[i]->P;
  }
  const size_t n = b->cl.size();
  CS_HIP(hipMemcpyAsync(b->dtab, b->htab.data(), sizeof(DevCloth) * n, hipMemcpyHostToDevice,
                        b->stream));
  CS_HIP(hipMemcpyAsync(b->ptab, b->hptab.data(), sizeof(Params) * n, hipMemcpyHostToDevice,
                        b->stream));
  return kOk;
}

}  // namespace
}  // namespace cs

extern "C" int cs_batch_create(int n, cs_cloth* const* cloths, cs_batch** out) {
  using namespace cs;
  if (n < 1 || !cloths || !out) return kBadArg;
  *out = nullptr;
  Batch* b = new Batch();
  for (int i = 0; i < n; i++) {
    if (!cloths[i]) {
      delete b;
      set_error("cs_batch_create: NULL cloth");
      return kBadArg;
    }
    Cloth* c = H(cloths[i]);
    for (Cloth* o : b->cl)
      if (o == c) {
        delete b;
        set_error("cs_batch_create: a cloth appears twice");
        return kBadArg;
      }
    b->cl.push_back(c);
  }
  b->tol = b->cl[0]->P.pcg_tol;
  b->maxit = b->cl[0]->P.pcg_maxit;
  BatchView& v = b->view;
  v.n = n;
  v.piped = true;
  for (Cloth* c : b->cl) {
    if (c->P.pcg_tol != b->tol || c->P.pcg_maxit != b->maxit) {
      delete b;
      set_error("cs_batch_create: the cloths must share the PCG tolerance and iteration cap");
      return kBadArg;
    }
    const DevCloth& d = c->d;
    v.max_np = std::max(v.max_np, d.np);
    v.max_nt = std::max(v.max_nt, d.nt);
    v.max_nq = std::max(v.max_nq, d.nq);
    v.max_nnz = std::max(v.max_nnz, d.nnz);
    v.max_wmax = std::max(v.max_wmax, d.wmax);
    v.max_qwidth = std::max(v.max_qwidth, d.qwidth);
    v.max_ncomp = std::max(v.max_ncomp, d.ncomp);
    v.max_fu_blocks = (v.max_fu_blocks < 0 || d.fu_blocks == 0)
                          ? -1
                          : std::max(v.max_fu_blocks, d.fu_blocks);
    v.piped = v.piped && quad_ok(d) && (v.kpack == 0 || v.kpack == d.kpack);
    v.kpack = d.kpack;
  }
  v.wg = v.piped;
  for (Cloth* c : b->cl) v.wg = v.wg && pcg_wg_ok(c->d) && c->d.pcg_variant == 2;
  b->htab.resize(n);
  b->hptab.resize(n);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->aux.stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->aux.updated, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->aux.tested, hipEventDisableTiming) != hipSuccess ||
      dev_alloc((void**)&b->dtab, sizeof(DevCloth) * n, b->stream) != kOk ||
      dev_alloc((void**)&b->ptab, sizeof(Params) * n, b->stream) != kOk ||
      dev_alloc((void**)&b->dgather, sizeof(int) * n, b->stream) != kOk ||
      pinned_alloc((void**)&b->hgather, sizeof(int) * n) != kOk ||
      hipStreamSynchronize(b->stream) != hipSuccess ||
      hipEventCreateWithFlags(&b->internal_done, hipEventDisableTiming) != hipSuccess) {
    batch_free(b);
    set_error("cs_batch_create: device allocation failed");
    return kOutOfMemory;
  }
  b->cloth_done.assign(n, nullptr);
  for (int i = 0; i < n; i++)
    if (hipEventCreateWithFlags(&b->cloth_done[i], hipEventDisableTiming) != hipSuccess) {
      batch_free(b);
      set_error("cs_batch_create: event creation failed");
      return kHipError;
    }
  // the SpMV's blocks over the rows that exist (kSpmvRows rows each) and the
  // vector update's over the entries that exist (kVecBlock each)
  std::vector<int2> rowmap, vecmap;
  for (int i = 0; i < n; i++) {
    for (int r0 = 0; r0 < b->cl[i]->d.np; r0 += kSpmvRows) rowmap.push_back(make_int2(i, r0));
    for (int e0 = 0; e0 < 3 * b->cl[i]->d.np; e0 += kVecBlock) vecmap.push_back(make_int2(i, e0));
  }
  if (dev_alloc((void**)&b->rowmap, sizeof(int2) * rowmap.size(), b->stream) != kOk ||
      dev_alloc((void**)&b->vecmap, sizeof(int2) * vecmap.size(), b->stream) != kOk ||
      hipMemcpyAsync(b->rowmap, rowmap.data(), sizeof(int2) * rowmap.size(),
                     hipMemcpyHostToDevice, b->stream) != hipSuccess ||
      hipMemcpyAsync(b->vecmap, vecmap.data(), sizeof(int2) * vecmap.size(),
                     hipMemcpyHostToDevice, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess) {
    batch_free(b);
    set_error("cs_batch_create: device allocation failed");
    return kOutOfMemory;
  }
  v.tab = b->dtab;
  v.ptab = b->ptab;
  v.rowmap = b->rowmap;
  v.nrowblk = (int)rowmap.size();
  v.vecmap = b->vecmap;
  v.nvecblk = (int)vecmap.size();
  *out = reinterpret_cast<cs_batch*>(b);
  return kOk;
}

extern "C" int cs_batch_destroy(cs_batch* p) {
  if (!p) return cs::kOk;
  cs::batch_free(cs::HB(p));
  return cs::kOk;
}

extern "C" int cs_batch_size(cs_batch* p, int* n) {
  if (!p || !n) return cs::kBadArg;
  *n = (int)cs::HB(p)->cl.size();
  return cs::kOk;
}

extern "C" int cs_batch_step(cs_batch* p, int nsteps, cs_step_stats* stats) {
  using namespace cs;
  if (!p || nsteps < 0) return kBadArg;
  Batch* b = HB(p);
  for (Cloth* c : b->cl)
    if (c->d.host_pd) {
      set_error("cs_batch_step: host obstacle values describe one state");
      return kBadArg;
    }
  hipStream_t s = b->stream;
  // member cloths may have pending work on their own streams
  for (Cloth* c : b->cl) CS_HIP(hipStreamSynchronize(c->stream));
  if (int st = batch_refresh(b)) return st;
  // the PCG variant follows the members' current options (results are the
  // same either way): the one-workgroup solve when every member selects it
  b->view.wg = b->view.piped;
  for (Cloth* c : b->cl) b->view.wg = b->view.wg && pcg_wg_ok(c->d) && c->d.pcg_variant == 2;
  // one sweep instance serves every member: their coefficient formats agree
  b->view.f32 = b->cl[0]->d.precision != 0;
  for (Cloth* c : b->cl)
    if ((c->d.precision != 0) != b->view.f32) {
      set_error("cs_batch_step: the cloths of a batch must share the precision");
      return kBadArg;
    }
  const size_t n = b->cl.size();
  for (int k = 0; k < nsteps; k++) {
    for (Cloth* c : b->cl)
      if (c->n_sweep > 0) {
        double sv = (c->time + c->P.dt) / c->sweep_duration;
        sv = sv < 0.0 ? 0.0 : (sv > 1.0 ? 1.0 : sv);
        const double off[3] = {c->sweep_delta[0] * sv, c->sweep_delta[1] * sv,
                               c->sweep_delta[2] * sv};
        if (int st = launch_apply_sweep(c->d, c->sweep_verts, c->n_sweep, off, s)) return st;
      }
    if (int st = batch_internal_step(b->view, b->tol, b->maxit, b->poll, b->dgather, b->hgather, s,
                                     b->prof.on ? &b->prof : nullptr, &b->aux))
      return st;
    for (Cloth* c : b->cl) {
      CS_HIP(hipMemcpyAsync(c->hflags, c->d.flags, sizeof(int) * kNumFlag, hipMemcpyDeviceToHost, s));
      CS_HIP(hipMemcpyAsync(c->hscal, c->d.scal, sizeof(double) * kNumScal, hipMemcpyDeviceToHost, s));
    }
    CS_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < n; i++)
      if (int st = check_tickets(b->cl[i]->hflags)) return st;
    // contact passes: every cloth on its own stream, from up to
    // kContactThreads host threads (each pass waits on the host for its
    // contact counts, so passes of different cloths overlap those waits)
    CS_HIP(hipEventRecord(b->internal_done, s));
    auto one = [&](int i) -> int {
      Cloth* c = b->cl[i];
      c->have_x1 = true;
      cs_step_stats st_local;
      std::memset(&st_local, 0, sizeof(st_local));
      st_local.pcg_conv = c->hscal[kConv];
      st_local.pcg_iters = c->hflags[kDone] ? c->hflags[kIters] : c->P.pcg_maxit;
      st_local.ilu_singular = c->hflags[kNsing];
      c->time += c->P.dt;
      if (c->hflags[kFlagNonFinite]) {
        st_local.status = kNonFinite;
        if (stats) stats[i] = st_local;
        return kOk;
      }
      CS_HIP(hipStreamWaitEvent(c->stream, b->internal_done, 0));
      int applied = 0;
      if (int st = contact_resolve(c->contact, c->S, c->d, c->P, c->stream, &applied, &st_local,
                                   nullptr))
        return st;
      if (applied)
        if (int st = post_contact_update(c->d, c->P, c->stream)) return st;
      st_local.impulse_applied = applied;
      CS_HIP(hipEventRecord(b->cloth_done[i], c->stream));
      if (stats) stats[i] = st_local;
      return kOk;
    };
    if (int st = parallel_for((int)n, kContactThreads, one)) return st;
    for (size_t i = 0; i < n; i++)
      if (!b->cl[i]->hflags[kFlagNonFinite]) CS_HIP(hipStreamWaitEvent(s, b->cloth_done[i], 0));
  }
  CS_HIP(hipStreamSynchronize(s));
  return kOk;
}

extern "C" int cs_batch_profile_enable(cs_batch* p, int on) {
  using namespace cs;
  if (!p || on < CS_PROFILE_OFF || on > CS_PROFILE_SWEEP) return kBadArg;
  Batch* b = HB(p);
  b->prof.stream = b->stream;
  b->prof.reset();
  b->prof.on = on != CS_PROFILE_OFF;
  b->prof.mask = on == CS_PROFILE_SWEEP ? 1u << kProfSweep : ~0u;
  return kOk;
}

extern "C" int cs_batch_profile_read(cs_batch* p, int kind, double* total_ms, long long* launches) {
  using namespace cs;
  if (!p || kind < 0 || kind >= kNumProf) return kBadArg;
  Batch* b = HB(p);
  b->prof.harvest();
  if (total_ms) *total_ms = b->prof.total_ms[kind];
  if (launches) *launches = b->prof.count[kind];
  return kOk;
}

extern "C" int cs_set_setup_cache(const char* dir) {
  cs::set_setup_cache_dir(dir);
  return cs::kOk;
}

extern "C" int cs_setup_cache_stats(long long* hits, long long* misses, long long* writes) {
  cs::setup_cache_stats(hits, misses, writes);
  return cs::kOk;
}
