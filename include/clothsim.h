/* clothsim: MI355X (gfx950) implementation of the implicit cloth step with
 * self-contact of nobuyuki83/cloth_sim_self_collision, behind a C ABI.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Arrays follow the reference layouts:
 *    positions/velocities are AoS double[3*np]; triangles int[3*nt];
 *    bending stencils int[4*nq] laid out [opp0, opp1, e0, e1]; boundary flags
 *    int[np] (0 free, 1 fixed); the matrix is 3x3-block CSR without the
 *    diagonal plus a separate diagonal (CMatrixSquareSparse,
 *    matrix_square_sparse.h:14-48) -- row offsets int[np+1], column indices
 *    int[nnz] ascending per row, values double[9*nnz] and double[9*np],
 *    blocks row-major.
 *  - Every call returns a status (CS_OK = 0); cs_last_error() describes the
 *    most recent failure of the calling thread.  No C++ exception crosses
 *    the ABI.
 *  - One HIP stream per cloth; calls on distinct cloths may run from
 *    different threads, calls on one cloth must not overlap.
 *  - Host-pointer entry points copy in/out and synchronise; *_device entry
 *    points take device pointers and stay on the cloth's stream.
 *  - FP64 throughout; results are bit-identical to the reference's x86-64
 *    build on the same inputs (contact sets, matrices, factors, positions).
 */
#ifndef CLOTHSIM_H
#define CLOTHSIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CS_OK 0
#define CS_ERR_BAD_ARG 1
#define CS_ERR_HIP 2
#define CS_ERR_OOM 3
#define CS_ERR_SINGULAR 4      /* informational: ILU met singular pivots */
#define CS_ERR_NOT_CONVERGED 5 /* informational: PCG hit its iteration cap */
#define CS_ERR_NONFINITE 6
#define CS_ERR_CAPACITY 7
#define CS_ERR_NO_DEVICE 8

/* obstacle kinds: penetrationDepth(pd, n, p) with pd > 0 inside and n the
 * push-out direction (cloth_internal_physics.h:251-255) */
#define CS_OBST_NONE 0
#define CS_OBST_PLANE 1      /* params (nx, ny, nz, d): pd = d - n.p        */
#define CS_OBST_SPHERE 2     /* params (cx, cy, cz, r): pd = r - |p - c|    */
#define CS_OBST_SPHERE_IN 3  /* sign-flipped sphere of self_contact_sparse/main.cpp:117-133 */
#define CS_OBST_BOX 4        /* params (cx, cy, cz, hx, hy, hz), axis-aligned, half-extents h:
                              * inside pd = distance to the nearest face (n its outward
                              * normal, the first axis on ties), outside pd = -distance */
#define CS_OBST_BAR 5        /* params (ax, ay, az, bx, by, bz, r): the capsule of radius r
                              * around segment a-b; pd = r - |p - c|, c the segment's
                              * closest point, n = (p - c)/|p - c| */
#define CS_OBST_BAR_GROUND 6 /* params (ax, ay, az, bx, by, bz, r, h): the bar and the
                              * ground z = h (normal +z); the deeper of the two, the bar
                              * on ties.  Kinds 4-6 are not in the reference (obstacle
                              * callbacks with its penetrationDepth convention) */
#define CS_OBST_MAX 6

typedef struct cs_cloth cs_cloth;

typedef struct {
  int np, nt, nq;
  const double* xyz0;    /* rest positions, 3*np */
  const double* xyz;     /* initial positions (also seed the BVH topology), 3*np */
  const double* uvw;     /* initial velocities, 3*np (NULL = zero) */
  const int* tri;        /* 3*nt */
  const int* quad;       /* 4*nq */
  const int* bc;         /* np, NULL = all free */
} cs_mesh;

typedef struct {
  double lambda, myu;         /* Lame parameters */
  double stiff_bend;
  double gravity[3];
  double mass_point;          /* per-vertex mass (area*density/np in the demos) */
  double stiff_contact;       /* obstacle penalty and self-contact stiffness */
  double contact_clearance;
  double dt;
  double pcg_tol;             /* the reference hard-codes 1e-4 ... */
  int pcg_maxit;              /* ... and 100 (solve_internal_sparse.h:318-319) */
  int obstacle_kind;
  double obstacle[8];
} cs_params;

typedef struct {
  double pcg_conv;        /* achieved residual ratio */
  int pcg_iters;          /* as Solve_PCG reports it */
  int ilu_singular;       /* pivots with |det| <= 1e-30 */
  int impulse_applied;    /* the contact pass changed velocities */
  int n_prox;             /* proximity contacts */
  int n_ccd_pass;         /* CCD passes that found contacts (<= 5) */
  int n_ccd[5];           /* contacts per CCD pass */
  int n_riz_pass;         /* impact-zone passes run (<= 100) */
  int n_riz_vertices;     /* vertices in impact zones at the end */
  int status;             /* CS_OK, or why this cloth's step stopped: CS_ERR_NONFINITE when
                           * the solve produced an infinite or NaN displacement (the
                           * contact pass is then skipped) */
} cs_step_stats;

const char* cs_last_error(void);
int cs_device_count(int* n);

/* Create a cloth: uploads the mesh and builds the pattern (union of the
 * bending stencils, self_contact_sparse/main.cpp:577-581), the ILU(0)
 * diagonal index (ilu_sparse.cpp:114-134), level schedules, assembly gather
 * lists, the BVH topology (MakeBVHTopology_TopDown, bvh_aabb.cpp:294-313) and
 * the vertex 1-ring (main.cpp:575). */
int cs_cloth_create(const cs_mesh* mesh, const cs_params* params, cs_cloth** out);
int cs_cloth_destroy(cs_cloth* c);
int cs_cloth_set_params(cs_cloth* c, const cs_params* params);
int cs_cloth_sizes(const cs_cloth* c, int* np, int* nt, int* nq, int* nnz, int* nnode);
/* device memory held by the cloth now (its arena plus the contact buffers at
 * their current capacity); what a server admits against its HBM budget */
int cs_cloth_memory(const cs_cloth* c, long long* bytes);

/* state (host arrays of 3*np doubles) */
int cs_set_state(cs_cloth* c, const double* xyz, const double* uvw);
int cs_get_state(cs_cloth* c, double* xyz, double* uvw);
/* device pointers of the resident state (valid for the cloth's lifetime) */
int cs_state_device(cs_cloth* c, double** xyz, double** uvw);
/* positions at the start of the step, read by cs_contact_resolve
 * (cs_step_internal stores them itself) */
int cs_set_step_start(cs_cloth* c, const double* xyz1);
/* per-vertex obstacle values from a host penetrationDepth callback: pd[np],
 * n[3*np], used by every assembly until cleared (cs_step refuses them, since
 * they describe one state); NULL restores the analytic obstacle */
int cs_set_obstacle_values(cs_cloth* c, const double* pd, const double* n);

/* ---- the three calls of StepTime (self_contact_sparse/main.cpp:213-263) ---- */

/* StepTime_InternalDynamicsILU (solve_internal_sparse.h:255-326): assemble,
 * factor, PCG with the cloth's tolerance/cap, x += dx, v = dx/dt.  Also
 * stores the start-of-step positions used by the contact pass. */
int cs_step_internal(cs_cloth* c, double* conv_ratio, int* iteration);
/* GetIntermidiateVelocityContactResolved (self_collision_cloth.h:21-34) on the
 * stored start-of-step positions and the current velocities. */
int cs_contact_resolve(cs_cloth* c, int* impulse_applied, cs_step_stats* stats);
/* x = x_start + v dt followed by UpdateIntermidiateVelocity
 * (solve_internal_sparse.h:329-369); StepTime does this only when an impulse
 * was applied. */
int cs_post_contact_update(cs_cloth* c);
/* UpdateIntermidiateVelocity alone (solve_internal_sparse.h:329-369): v -=
 * dt/(2m) grad W(x) at the current positions (set them with cs_set_state) */
int cs_update_velocity(cs_cloth* c);
/* n full StepTime iterations; stats of the last step (may be NULL).  Before
 * each step the swept vertices (cs_set_sweep) are placed at
 * rest + delta * clamp((t + dt) / duration, 0, 1); t then advances by dt.
 * A step whose solve yields an infinite or NaN displacement (the reference's
 * ILU(0) overflows on some large meshes) stops the loop with
 * CS_ERR_NONFINITE before its contact pass; the state then holds that step's
 * internal result. */
int cs_step(cs_cloth* c, int n, cs_step_stats* stats);
/* kinematic boundary: n vertices (which should carry bc = 1) moved linearly by
 * delta[3] over `duration` seconds, then held.  The reference has only fixed
 * boundaries; this drives the fold workload. n = 0 clears it. */
int cs_set_sweep(cs_cloth* c, int n, const int* verts, const double* delta, double duration);
int cs_set_time(cs_cloth* c, double t);
int cs_get_time(const cs_cloth* c, double* t);

/* ---- batches: many cloths stepped by shared kernel launches ----
 * Every kernel of the internal step (assembly, ILU(0), each PCG operation) is
 * launched once for all cloths of the batch, one grid row per cloth, so a
 * batch fills the GPU where one cloth occupies a few CUs.  Each cloth keeps
 * its own convergence test and its result is bit-identical to stepping it
 * alone; the contact pass runs per cloth.  The caller keeps ownership of the
 * cloths; they must share pcg_tol and pcg_maxit, and must not be stepped on
 * their own while a cs_batch_step on them runs. */
typedef struct cs_batch cs_batch;
int cs_batch_create(int n, cs_cloth* const* cloths, cs_batch** out);
int cs_batch_destroy(cs_batch* b);
int cs_batch_size(cs_batch* b, int* n);
/* per-kernel-class timing of a batch's shared launches with HIP events on the
 * batch's stream (kinds and modes as cs_profile_enable; the contact passes,
 * on the cloths' own streams, are not included) */
int cs_batch_profile_enable(cs_batch* b, int on);
int cs_batch_profile_read(cs_batch* b, int kind, double* total_ms, long long* launches);
/* nsteps StepTime iterations of every cloth; stats[i] (n entries, may be
 * NULL) receives cloth i's last step.  The contact passes of the cloths run
 * concurrently, each on its cloth's stream from a host worker thread.  A
 * cloth whose solve is not finite gets stats[i].status = CS_ERR_NONFINITE
 * and no contact pass; the others are unaffected and the call returns CS_OK
 * (remove such a cloth from the batch: it stays non-finite). */
int cs_batch_step(cs_batch* b, int nsteps, cs_step_stats* stats);

/* implementation choices (results are identical for every value):
 * CS_OPT_SWEEP: 0 automatic (the packed four-lanes-per-row sweep when
 *               the mesh allows it), 1 row-by-row CSR sweep
 * CS_OPT_PCG:   0 automatic (= 1), 1 one kernel launch per PCG operation
 *               with a host poll every 8 iterations, 2 the whole solve in one
 *               workgroup (one launch; needs the packed sweep and
 *               3 np <= 65536; a batch uses it for a step when all its cloths
 *               select it at that step,
 *               one workgroup per cloth, each iterating to its own
 *               convergence).  Measured: 2 is not faster than 1 for a 32^2
 *               cloth and slower for a 64-cloth batch of 24^2..128^2 cloths
 *               at the iteration cap (each cloth's SpMV and reductions then
 *               run on one CU) */
/* CS_OPT_PRECISION: 0 FP64, bit-identical to the reference (default); 1
 *               mixed: the packed ILU sweep reads the factor's coefficients
 *               rounded to FP32 (D^-1, the vectors, SpMV, the reductions and
 *               everything else stay FP64).  Not bit-identical: the position
 *               error against the FP64 reference is small only where
 *               ILU(0)-PCG converges (a non-converging solve stops at its
 *               iteration cap on a different iterate).  Needs the packed sweep
 *               and CS_OPT_PCG != 2; all cloths of a batch share it.  Measured:
 *               no faster than FP64 (the sweep waits on its instruction
 *               stream, and the conversions add to it). */
#define CS_OPT_SWEEP 1
#define CS_OPT_PCG 2
#define CS_OPT_PRECISION 3
int cs_set_option(cs_cloth* c, int option, int value);

/* per-kernel-class timing with HIP events on the cloth's stream:
 * 0 ILU sweeps, 1 PCG reductions, 2 SpMV, 3 vector updates, 4 ILU factor,
 * 5 assembly, 6 contact pass, 7 other, 8 the one-workgroup PCG solve (all
 * of 0-3 in one kernel).  Enabling resets the totals.
 * on = CS_PROFILE_OFF, CS_PROFILE_ALL (every class, plus the reduction
 * counters of cs_reduction_stats) or CS_PROFILE_SWEEP (the ILU sweeps only:
 * two events per sweep launch, no counters, for timing that barely
 * disturbs the step) */
#define CS_PROFILE_OFF 0
#define CS_PROFILE_ALL 1
#define CS_PROFILE_SWEEP 2
int cs_profile_enable(cs_cloth* c, int on);
int cs_profile_read(cs_cloth* c, int kind, double* total_ms, long long* launches);
/* counters of the exact PCG reductions while profiling is on: out[8] =
 * reductions, tiles, tiles applied from their summary, binade-walker steps,
 * elements added one by one, then nanoseconds (device wall clock, summed
 * over reductions) spent in the serial chain, in the binade walker (bursts
 * included) and in one-by-one bursts */
int cs_reduction_stats(cs_cloth* c, long long* out);

/* ---- lower-level entry points (host arrays unless noted) ---- */
int cs_get_pattern(const cs_cloth* c, int* rowptr, int* colind, int* diaind);
/* assemble A and b at the given state without solving */
int cs_assemble(cs_cloth* c, const double* xyz, const double* uvw);
int cs_get_system(cs_cloth* c, double* dia, double* crs, double* b);
int cs_set_system(cs_cloth* c, const double* dia, const double* crs, const double* b);
/* SetValueILU + DoILUDecomp on the current system (ilu_sparse.cpp:352-624) */
int cs_ilu_factor(cs_cloth* c, int* n_singular);
int cs_get_factor(cs_cloth* c, double* fdia, double* fcrs);
int cs_set_factor(cs_cloth* c, const double* fdia, const double* fcrs);
/* CPreconditionerILU::Solve (forward + backward substitution), in place */
int cs_ilu_solve(cs_cloth* c, double* vec);
/* CMatrixSquareSparse::MatVec: y = alpha A x + beta y */
int cs_matvec(cs_cloth* c, double alpha, const double* x, double beta, double* y);
/* Solve_PCG (ilu_sparse.cpp:628-709) on the current system and factor with
 * right-hand side b; x out, r (residual) out; conv/iters like the reference */
int cs_pcg_solve(cs_cloth* c, const double* b, double tol, int maxit, double* x, double* r,
                 double* conv, int* iters);

/* BVH: node array 3 ints per node (parent, child0, child1; child1 == -1 marks
 * a leaf holding triangle child0), numbered like the reference */
int cs_get_bvh(const cs_cloth* c, int* nodes);
/* BuildBoundingBoxChild_Prx / _CCD (bvh_aabb.cpp:23-106); boxes out as 6
 * doubles per node (xmin, xmax, ymin, ymax, zmin, zmax) when bb != NULL */
int cs_bvh_refit_prx(cs_cloth* c, const double* xyz, double delta, double* bb);
int cs_bvh_refit_ccd(cs_cloth* c, const double* xyz, const double* uvw, double dt, double* bb);
/* GetContactElement_Proximity / _CCD (bvh_aabb.cpp:441-557, 797-923) as the
 * std::set<CContactElement> contents in order: 5 ints per contact
 * (is_fv, i0, i1, i2, i3).  *n receives the count; at most cap are written. */
int cs_query_proximity(cs_cloth* c, const double* xyz, double delta, int* out, int cap, int* n);
int cs_query_ccd(cs_cloth* c, const double* xyz, const double* uvw, double dt, double delta,
                 int* out, int cap, int* n);

/* ---- element routines over arrays (cloth_internal_physics.h) ---- */
/* WdWddW_CST for n triangles: C, c 9 doubles each; W[n], dW[9n], ddW[81n] */
int cs_element_cst(int n, const double* C, const double* c, double lambda, double myu,
                   double* dW, double* ddW);
/* the energy W[n] of WdWddW_CST (cloth_internal_physics.h:125) */
int cs_element_cst_energy(int n, const double* C, const double* c, double lambda, double myu,
                          double* W);
/* WdWddW_Bend for n stencils: C, c 12 doubles each; dW[12n], ddW[144n] */
int cs_element_bend(int n, const double* C, const double* c, double stiff, double* dW,
                    double* ddW);
/* WdWddW_Contact for n points against an obstacle (parameters as in
 * cs_params, 8 doubles): dW[3n], ddW[9n] */
int cs_element_contact(int n, const double* p, double stiff, double clearance, int kind,
                       const double* obstacle, double* dW, double* ddW);
/* WdWddW_Contact (cloth_internal_physics.h:243-280) for n points whose
 * penetrationDepth callback gave depth pd[n] and direction nrm[3n]:
 * W[n], dW[3n], ddW[9n] */
int cs_element_contact_depth(int n, const double* pd, const double* nrm, double stiff,
                             double clearance, double* W, double* dW, double* ddW);

/* On-disk cache of the per-mesh setup (pattern, level schedules, sweep
 * slots, gather lists, BVH topology), the work cs_cloth_create and
 * cs_inspect_setup do on the host before anything reaches the device
 * (self_contact_sparse/main.cpp:561-583 plus this library's schedules; 2 s at
 * 512^2).  Files are keyed by the mesh (connectivity and initial positions)
 * and the layout version; a file that is missing, truncated or for another
 * mesh is rebuilt.  dir NULL or "": no cache (the default).  Process-wide;
 * safe to call from any thread. */
int cs_set_setup_cache(const char* dir);
/* cache lookups that hit, missed, and files written since the process began */
int cs_setup_cache_stats(long long* hits, long long* misses, long long* writes);
/* host-only: the per-mesh setup of cs_cloth_create without a device.
 * sizes[10] = np, nt, nq, nnz, nnode, forward levels, backward levels, max
 * lower / upper blocks per row of the packed sweeps (0 = not packed), 1 if
 * the sweep dependencies fit the LDS ring; any
 * output array may be NULL (call once with NULLs to learn the sizes):
 * rowptr[np+1], colind[nnz], diaind[np], nodes[3*nnode] */
int cs_inspect_setup(const cs_mesh* mesh, int* sizes, int* rowptr, int* colind, int* diaind,
                     int* nodes);

/* device arithmetic check: out[2i] = a[i]/b[i], out[2i+1] = sqrt(a[i]) */
int cs_selftest_divsqrt(int n, const double* a, const double* b, double* out);
/* the PCG reduction on the device: *out = ((0.0 + t[0]) + t[1]) + ... */
int cs_selftest_ordered_sum(int n, const double* t, double* out);
/* cs_ordered_sum_noalloc has no counterpart in the reference.  It is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the reduction of
 * cs_selftest_ordered_sum, *out = ((0.0 + t[0]) + t[1]) + ... bit for bit,
 * on device pointers t (n doubles) and out, enqueued on `stream` (a
 * hipStream_t; NULL = the null stream), allocating nothing.  Its scratch is
 * a static device array of CS_NOALLOC_MAX_TILES tile summaries (147 KB per
 * GPU, held until the process ends), so n is at most
 * CS_NOALLOC_MAX_TILES * 1024 = 4,194,304 (CS_ERR_CAPACITY beyond).  All
 * calls on one GPU share that array: calls that overlap on one GPU race. */
#define CS_NOALLOC_MAX_TILES 4096
int cs_ordered_sum_noalloc(int n, const double* t, double* out, void* stream);
/* device copy bandwidth: reps + 1 launches of a streaming copy of n doubles
 * (n even) between two fresh buffers; the last reps are timed with events.
 * *gbs = 16 n reps / seconds / 1e9 (read + write) */
int cs_bandwidth_probe(long long n, int reps, double* gbs);
/* read-only streaming of n doubles by `blocks` workgroups of 512 threads
 * (blocks = 1: one CU, the ceiling of the one-workgroup ILU sweep), reps
 * times; n a multiple of 16384 * blocks; GB/s of reads */
int cs_read_probe(long long n, int reps, int blocks, double* gbs);
/* vector-load issue cost of one workgroup of `waves` waves (1..8) shaped like
 * the ILU sweep: each wave issues twelve 16-byte loads per step with two
 * steps of loads in flight and an LDS-only barrier per step, quads of lanes
 * reading three consecutive 16-byte values (lane 3 repeats lane 2) from a
 * 64 MiB buffer; steps a multiple of 3.  *ns_per_load = time of one wave's
 * load instruction (the launch time over steps * waves * 12) */
int cs_issue_probe(int waves, int steps, double* ns_per_load);
/* shape of a cloth's ILU sweeps, for pricing them against cs_issue_probe:
 * out[8] = sub-levels of the longest component's forward and backward
 * sweeps, waves per sweep workgroup, vector loads per wave per forward and
 * per backward sub-level, components (one workgroup each), 1 if the packed
 * quad sweep is in use (0: the row-by-row sweep; the fields but the
 * components are then 0), and the wave-load instructions one launch issues in
 * its busiest workgroup (forward and backward; waves without a row in three
 * consecutive sub-levels skip their loads) */
int cs_sweep_shape(const cs_cloth* cloth, int* out);

#ifdef __cplusplus
}
#endif
#endif /* CLOTHSIM_H */
