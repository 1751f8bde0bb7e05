// Contact-resolved velocity pass on the device
// (GetIntermidiateVelocityContactResolved, self_collision_cloth.cpp:355-456).
#pragma once

#include <vector>

#include <cstdint>

#include "../../include/clothsim.h"
#include "cs_device.hpp"
#include "cs_setup.hpp"

namespace cs {

struct ContactState {
  int np = 0, nt = 0, nnode = 0;
  int nbits = 1;          // bits per vertex index in the sort keys
  int cap = 0;            // capacity of raw contact records
  // leaf pairs whose boxes overlap (traversal output, narrow-phase input):
  // 2 ints per pair (first and second leaf node)
  int* cand = nullptr;
  int cand_cap = 0;
  // raw records from the narrow phase (unordered)
  int* rkey = nullptr;    // 4 per record: canonical key
  int* rinfo = nullptr;   // is_fv | test index << 1
  int* rpair = nullptr;   // 2 per record: first and second leaf node
  // sorting / grouping
  uint64_t* skey[2] = {nullptr, nullptr};
  int* sidx[2] = {nullptr, nullptr};
  int* head = nullptr;
  int* gpos = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  // the contact set in key order: 5 ints (is_fv, i0, i1, i2, i3)
  int* ce = nullptr;
  // impulse schedule
  int* pred = nullptr;       // 4 per contact: previous contact on the same vertex
  int* succ = nullptr;       // 4 per contact: next contact on the same vertex
  int* npred = nullptr;      // earlier contacts on its vertices not yet run
  int* ready[2] = {nullptr, nullptr};  // contacts ready to run, this round and the next
  int* inc = nullptr;        // 4 per contact: vertex incidences grouped by vertex
  int* vcount = nullptr;     // np + 1
  int* vstart = nullptr;     // np + 1
  int* vfill = nullptr;      // np + 1
  // scalars
  int* counters = nullptr;   // device
  int* hcounters = nullptr;  // pinned mirror
  // refit levels: nodes of depth max_depth - l at refit_order[refit_lvl[l]
  // .. refit_lvl[l + 1]) (MeshSetup::refit_lvl_ptr, host copy)
  std::vector<int> refit_lvl;
  // rigid impact zones
  int* zpar = nullptr;       // union-find parent per vertex, -1 outside zones
  int* zroot = nullptr;
  int* zmemb = nullptr;
  int* zroots = nullptr;
  double* u0 = nullptr;      // velocities before the zone passes
  // inputs of the host-array query entry points
  double* qx = nullptr;
  double* qv = nullptr;
  double* qbb = nullptr;     // unused spare
  double* host_pd_buf = nullptr;  // owned by the cloth arena
};

int contact_init(ContactState& cs, const MeshSetup& S, const DevCloth& d, hipStream_t s);
// after every stream that used the buffers is idle (they are freed on s)
void contact_release(ContactState& cs, hipStream_t s);
// full contact pass on (d.x1, d.v); stats may be NULL
int contact_resolve(ContactState& cs, const MeshSetup& S, const DevCloth& d, const Params& P,
                    hipStream_t s, int* applied, cs_step_stats* stats, Prof* prof = nullptr);
// x = x1 + v dt, then UpdateIntermidiateVelocity
// UpdateIntermidiateVelocity (solve_internal_sparse.h:329-369) at the current x
int update_velocity(const DevCloth& d, const Params& P, hipStream_t s);
int post_contact_update(const DevCloth& d, const Params& P, hipStream_t s);
// host-array helpers of the C-ABI (uvw == NULL -> proximity)
int contact_refit_host(ContactState& cs, const MeshSetup& S, const DevCloth& d, hipStream_t s,
                       const double* xyz, const double* uvw, double delta, double dt, double* bb);
int contact_query_host(ContactState& cs, const MeshSetup& S, const DevCloth& d, hipStream_t s,
                       const double* xyz, const double* uvw, double delta, double dt, int* out,
                       int cap, int* n);

}  // namespace cs
