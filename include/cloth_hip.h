/*
 * cloth_hip.h — C ABI of the MI355X implicit cloth step (HipSolver backend).
 *
 * This is the drop-in boundary for closim's `ClothSolver` backend interface
 * (src/backend/solver.ts:5-13).  Every `cs_*` entry point below replaces one
 * member of that interface as implemented by `CpuSolver`
 * (src/backend/cpu-solver.ts:7-49); the citation is on each declaration.
 *
 * Conventions
 *   - Plain pointers + sizes, no C++/torch types.  All arrays are host memory
 *     owned by the caller; the library copies what it needs.
 *   - Every function returns an int status (CS_OK == 0).  On failure the
 *     message for the calling thread is available from cs_last_error().
 *     No C++ exception ever crosses the ABI.
 *   - A solver is not re-entrant: one thread at a time per cs_solver, exactly
 *     like the synchronous single-threaded CpuSolver.
 *   - Layouts follow the reference ClothScene / ClothMeshData
 *     (src/physics/scene.ts:5-14, src/mesh/mesh.ts:3-14): interleaved xyz
 *     (3n doubles), triangle indices (3m u32), invDm as [a,b,c,d] per triangle
 *     and areas/masses as float32 exactly as the reference stores them.
 */
#ifndef CLOTH_HIP_H
#define CLOTH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CS_ABI_VERSION 2

/* Opaque handle: one simulated ClothScene resident on one GPU. */
typedef struct cs_solver cs_solver;

/* ---- status codes (thrown JS Errors in the reference) ------------------ */
enum {
  CS_OK = 0,
  CS_E_NOT_INIT = 1,    /* "CpuSolver not initialized" (cpu-solver.ts:16,22,35,41,46) */
  CS_E_INVALID = 2,     /* bad argument / malformed scene */
  CS_E_NONFINITE = 3,   /* "non-finite gradient — blow up; reduce dt/stiffness" (newton.ts:129) */
  CS_E_HIP = 4,         /* HIP runtime error (no device, launch failure, ...) */
  CS_E_OVERFLOW = 5,    /* a bounded device buffer could not be grown */
  CS_E_NOMEM = 6,
  CS_E_UNSUPPORTED = 7
};

/* ---- solver configuration ------------------------------------------------ */
enum { CS_PREC_FP64 = 0, CS_PREC_MIXED = 1 };
/* CS_HVP_FD: the reference CpuSolver operator (central FD of the membrane
 * gradient, src/physics/fem.ts:108-130).  CS_HVP_ANALYTIC: the analytic
 * membrane HVP (src/physics/membrane-hvp.ts:18-108). */
enum { CS_HVP_FD = 0, CS_HVP_ANALYTIC = 1 };
/* pcg() preconditioner (src/math/pcg.ts:11-95 `applyPreconditioner`). */
enum { CS_PRECOND_JACOBI = 0, CS_PRECOND_BLOCK_JACOBI = 1, CS_PRECOND_SCHWARZ = 2 };
/* Coarse level added to the Schwarz preconditioner (two-level additive
 * Schwarz, src/solver/mas.ts, src/solver/coarse-csr.ts; the reference's GPU
 * path selects them with useMas / useCoarseC0 / useCoarsePcg,
 * src/backend/webgpu/gpu-newton.ts:213-242): z = Schwarz1(r) + P C(R r) with
 * one coarse dof per domain and axis and C the damped Jacobi omega Dc^-1
 * (applyMasTwoLevel), the diagonal-block inverse (C0) or coarse_iters
 * iterations of block-Jacobi PCG on the assembled Ac (C1).  Needs
 * CS_PRECOND_SCHWARZ. */
enum { CS_COARSE_NONE = 0, CS_COARSE_JACOBI = 1, CS_COARSE_BLOCK = 2, CS_COARSE_PCG = 3 };

#define CS_FLAG_TIMING 1u   /* record per-kernel HIP event timings (cs_kernel_timing) */
/* Root polish (not in the reference): when every Armijo trial of a Newton
 * iteration fails, take a full Newton step of the complete Hessian instead
 * (analytic membrane, central differences of bending and contact) if it passes
 * the contact check and at least halves the Newton gradient norm, and keep
 * iterating.  The
 * energy test stops resolving decreases near the root; rollouts that are
 * differentiated with cs_step_adjoint need the root to grad_tol. */
#define CS_FLAG_POLISH_ROOT 2u
/* Launch the mixed-precision PCG iterations one kernel at a time.  By default
 * they run as a captured HIP graph of 8 iterations replayed until the solve
 * converges (same kernels, same order, same results); CS_FLAG_TIMING also
 * selects the one-by-one launches. */
#define CS_FLAG_EAGER 4u
/* Newton direction with the analytic membrane Hessian whose stress term is
 * projected to positive semidefinite per triangle (tension only; not in the
 * reference): PCG then solves a positive definite system and Newton descends
 * through compressed, buckling states where the unprojected operator stalls.
 * Takes effect with CS_HVP_ANALYTIC; the gradient, energy, line search and
 * the adjoint's Hessian are unchanged. */
#define CS_FLAG_PSD_NEWTON 8u
/* Device Newton control (no reference counterpart; the reference's GPU path
 * does the same, src/backend/webgpu/gpu-newton.ts:1889-2200): the descent
 * safeguard and the trust region are decided on the device, and each Armijo
 * trial's contact check and energies come back in one host read instead of
 * two to three.  The trials are judged in the reference's order and rules,
 * so positions, energies, trial counts and contact diagnostics are those of
 * the sequential line search. */
#define CS_FLAG_DEVICE_NEWTON 32u
/* With CS_FLAG_DEVICE_NEWTON: evaluate the Armijo trials speculatively in
 * batches of 1, 2, 4, 8 (the reference's batched candidates, shaders/
 * armijo.wgsl:48-238) with one host read per batch; the same results.  Off
 * by default: a trial costs a whole contact check, so the trials computed
 * past the accepted one cost more than the host reads they save (the
 * 512x512 fold runs 13.0 steps/s without, 11.6 with). */
#define CS_FLAG_ARMIJO_SPECULATE 128u
/* The descent gate of the reference's WebGPU Newton (src/backend/webgpu/
 * gpu-solver.ts:1168-1188, shaders/newton-control.wgsl:58-73) instead of
 * CpuSolver's (src/solver/newton.ts:154-160): the Jacobi-scaled steepest
 * descent -g / diag replaces the PCG direction also when the PCG stopped on
 * non-positive curvature (p.Ap <= 1e-30), not only when the direction is
 * not a descent direction.  Off by default (CpuSolver's trajectories). */
#define CS_FLAG_BREAKDOWN_DESCENT 64u

/* ClothMaterial (src/physics/types.ts:1-14). */
typedef struct cs_material {
  double areal_density;    /* arealDensityKgM2 */
  double thickness;
  double stretch_warp;     /* C00 */
  double stretch_weft;     /* C11 */
  double stretch_coupling; /* C01 */
  double shear;            /* G */
  double bend_warp;
  double bend_weft;
  double damping;
} cs_material;

/* ContactParams (src/collision/types.ts:27-39). */
typedef struct cs_contact_params {
  double d_hat;         /* dHatM */
  double kappa;         /* kappaJ */
  double d_min;         /* dMinM */
  double friction_mu;   /* frictionMu */
  double friction_eps;  /* frictionEpsM */
} cs_contact_params;

/* ClothScene (src/physics/scene.ts:5-14) + ClothMeshData (src/mesh/mesh.ts:3-14)
 * + the ContactSystem configuration it carries (contact-assembly.ts:46-73). */
typedef struct cs_scene_desc {
  uint32_t n_verts;
  uint32_t n_tris;
  uint32_t n_hinges;
  const double* positions;        /* 3n, scene.positions */
  const double* velocities;       /* 3n, scene.velocities */
  const uint32_t* indices;        /* 3m */
  const float* inv_dm;            /* 4m  [a,b,c,d] */
  const float* areas;             /* m */
  const float* masses;            /* n  lumped */
  const float* rest_positions;    /* 3n mesh.restPositions (NULL: use positions) */
  const uint32_t* hinge_verts;    /* 4h (v0,v1 shared edge; v2,v3 opposite) */
  const double* hinge_rest_angle; /* h */
  const double* hinge_edge_len;   /* h */
  const double* hinge_area_sum;   /* h */
  cs_material material;
  double gravity[3];
  uint32_t n_pins;
  const uint32_t* pin_ids;        /* scene.pinned keys (insertion order) */
  const double* pin_pos;          /* 3 * n_pins */
  int32_t contact_enabled;        /* scene.contact != null */
  cs_contact_params contact;
  int32_t has_floor;              /* ContactSystem.floorY != null */
  double floor_y;
  uint32_t n_static_verts;        /* ContactSystem.setStaticMesh */
  uint32_t n_static_tris;
  const float* static_pos;        /* 3 * n_static_verts */
  const uint32_t* static_idx;     /* 3 * n_static_tris */
} cs_scene_desc;

/* StepperOptions (src/solver/newton.ts:20-25) + backend selection. */
typedef struct cs_options {
  int32_t newton_iters;       /* default 10 */
  int32_t pcg_max_iters;      /* default 60 */
  int32_t line_search_iters;  /* default 10 */
  double grad_tol;            /* default 1e-5 (absolute) */
  int32_t precision;          /* CS_PREC_* (default FP64) */
  int32_t hvp_mode;           /* CS_HVP_* (default FD = CpuSolver operator) */
  int32_t precond;            /* CS_PRECOND_* (default Jacobi = CpuSolver) */
  int32_t schwarz_target;     /* buildSchwarzDomains(target, min, max) */
  int32_t schwarz_min;
  int32_t schwarz_max;
  int32_t device;             /* HIP device ordinal */
  uint32_t flags;             /* CS_FLAG_* */
  int32_t coarse;             /* CS_COARSE_* (default none = one-level Schwarz) */
  int32_t coarse_iters;       /* CS_COARSE_PCG iterations (0: 8, the reference's c1-8) */
  double coarse_omega;        /* CS_COARSE_JACOBI damping (0: 0.5, masOmega) */
} cs_options;

/* ContactDiagnostics (src/collision/types.ts:49-61). */
typedef struct cs_contact_diag {
  uint64_t candidate_pairs;
  uint64_t active_pairs;
  uint64_t vt_pairs;
  uint64_t ee_pairs;
  uint64_t floor_pairs;
  double min_distance;
  double min_toi;
  double barrier_energy;
  double friction_work;
  uint64_t rejected_line_search_steps;
  uint64_t ccd_failures;
  uint64_t filter_rounds;  /* rounds of the parallel restitution-0 velocity filter (not in the reference) */
} cs_contact_diag;

/* StepStats (src/solver/newton.ts:10-18) plus a few counters the reference
 * does not report (newton_used, line_search_trials). */
typedef struct cs_step_stats {
  int32_t newton_iters;       /* configured maximum, as the reference reports (newton.ts:231) */
  int32_t pcg_iters;          /* total PCG iterations in the step */
  double grad_norm;
  double energy;
  double max_strain;
  int32_t converged;
  int32_t has_contact;
  cs_contact_diag contact;
  int32_t newton_used;        /* Newton iterations actually entered */
  int32_t line_search_trials; /* Armijo trials evaluated */
} cs_step_stats;

/* ---- lifecycle ----------------------------------------------------------- */

/* Fill `opt` with the CpuSolver defaults (newton.ts:30-33; FP64, FD HVP, Jacobi). */
void cs_options_default(cs_options* opt);

/* ClothSolver.initialize(scene) — cpu-solver.ts:11-13.  Copies the scene
 * (mesh, material, pins, contact configuration) to the device. */
int cs_create(const cs_scene_desc* scene, const cs_options* opt, cs_solver** out);
/* (no reference counterpart: releases device memory) */
void cs_destroy(cs_solver* s);

/* ClothSolver.step(dt) — cpu-solver.ts:15-18 → implicitStep (newton.ts:27-232).
 * Synchronous.  CS_E_NONFINITE mirrors the reference's non-finite gradient throw. */
int cs_step(cs_solver* s, double dt);

/* ClothSolver.setMaterial — cpu-solver.ts:20-23 (copies the material). */
int cs_set_material(cs_solver* s, const cs_material* mat);

/* ClothSolver.pinVertex — cpu-solver.ts:25-33.  pos == NULL pins at the current
 * position; pins are enforced on the positions immediately. */
int cs_pin_vertex(cs_solver* s, uint32_t vertex_id, const double* pos_or_null);

/* ContactSystem.setStaticMesh(pos, idx) — contact-assembly.ts:71-73, called
 * between steps; a moving rigid collider is the collider's mesh sampled at
 * each tick and attached this way (toCollisionObject + attachCollisionObject,
 * avatar-collider.ts:205-230).  The static triangles (count and corner ids)
 * are fixed at cs_create; new positions (f32, 3 n_verts) take effect at the
 * next step.  CS_E_UNSUPPORTED for another topology, CS_E_INVALID without a
 * contact system. */
int cs_set_static_mesh(cs_solver* s, const float* pos, uint32_t n_verts, const uint32_t* idx, uint32_t n_tris);

/* ClothSolver.unpinVertex — cpu-solver.ts:35-37. */
int cs_unpin_vertex(cs_solver* s, uint32_t vertex_id);

/* ClothSolver.getPositions / getVelocities — cpu-solver.ts:39-48.  Copy-out of
 * 3n doubles (the host wrapper writes them into the live scene arrays). */
int cs_get_positions(cs_solver* s, double* out, size_t n3);
int cs_get_velocities(cs_solver* s, double* out, size_t n3);

/* Overwrite scene.positions / scene.velocities (checkpoint restore; the
 * reference mutates these arrays directly, scene.ts:10-11). */
int cs_set_state(cs_solver* s, const double* x, const double* v, size_t n3);

/* CpuSolver.lastStats — cpu-solver.ts:9. */
int cs_last_stats(const cs_solver* s, cs_step_stats* out);

/* Thread-local message of the last failing call on this thread: the text of
 * the Error the reference throws (e.g. newton.ts:129, cpu-solver.ts:16). */
const char* cs_last_error(void);

/* Device name / ABI version (diagnostics only; no reference counterpart). */
int cs_device_info(int device, char* name, size_t cap);
int cs_abi_version(void);

/* ---- measurement hooks (no reference counterpart) ------------------------ */
enum {
  CS_KERNEL_HVP = 0,       /* one membrane(+mass+contact) Hessian-vector product */
  CS_KERNEL_GRADIENT = 1,  /* Newton gradient assembly */
  CS_KERNEL_BROADPHASE = 2,/* candidate generation (LBVH build + traverse) */
  CS_KERNEL_CCD = 3,       /* trial validation (distances + CCD) */
  CS_KERNEL_PRECOND = 4,   /* preconditioner apply */
  CS_KERNEL_COUNT = 5
};
/* Accumulated HIP-event time (ms) and launch count of a kernel family since the
 * last reset; only recorded when CS_FLAG_TIMING is set.  `bytes` receives the
 * algorithmic bytes per launch for that family (0 when not defined). */
int cs_kernel_timing(const cs_solver* s, int kernel, double* total_ms, uint64_t* count, double* bytes);
int cs_reset_timing(cs_solver* s);

/* Measurement hook, no counterpart in the reference: mean duration (ms) of
 * `reps` back-to-back launches of the PCG Hessian-vector kernel (k_mx_hvp in
 * mixed precision, k_pcg_hvp in FP64) on the operands the last cs_step left
 * resident, bracketed by HIP events on the solver stream; `bytes` receives
 * the algorithmic bytes of one launch.  Requires a previous cs_step. */
int cs_time_hvp(cs_solver* s, int reps, double* mean_ms, double* bytes);
/* (no reference counterpart: measurement hook) Bytes of one HVP launch on the
 * current operands: `algorithmic` the compulsory traffic (FP32 40n + 32m +
 * 64C, FP64 80n + 52m + 96C for n vertices, m triangles, C active contacts),
 * `read` what this build's kernel reads and writes, index maps and prepared
 * per-triangle records included.  cs_time_hvp's `bytes` is `algorithmic`. */
int cs_hvp_bytes(const cs_solver* s, double* algorithmic, double* read);
/* (no reference counterpart: measurement hook) The grid caps the launch
 * autotuner chose for the contact kernels whose results do not depend on
 * their grid (k_trial_pairs, k_trial_ccd, k_active_pairs), as text such as
 * "k_trial_pairs=2048 k_trial_ccd=256 k_active_pairs=1024" ("tuning" while it still measures
 * the candidates on the first calls; empty without contact). */
int cs_tuned_launch(const cs_solver* s, char* buf, size_t cap);

/* One step of each of `count` independent solvers (no reference counterpart:
 * the reference steps one scene).  Host threads (`threads`, at least 1) take
 * the instances from a queue, largest first, and each steps its instance on
 * that instance's own HIP stream, so the instances' kernels overlap on the
 * GPU.  `status` (optional, `count` entries) receives each cs_step result; the
 * return value is the first failure (cs_last_error names the instance) or
 * CS_OK.  A solver may appear only once in `solvers`. */
int cs_batch_step(cs_solver* const* solvers, uint32_t count, double dt, int threads, int32_t* status);
/* The same batch step with the instances' inner solves packed (no reference
 * counterpart): each instance steps in its own host thread, and at every
 * Newton iteration the PCG systems of all instances still stepping are
 * solved together, each PCG stage one launch over every instance's blocks
 * (a block table gives each block its instance, local block and local grid,
 * so each instance keeps its own material, partial sums and reduction trees:
 * the trajectories are those of cs_step).  Mixed-precision solvers with the
 * Jacobi or block-Jacobi preconditioner on one device (CS_E_INVALID
 * otherwise). */
int cs_batch_step_packed(cs_solver* const* solvers, uint32_t count, double dt, int32_t* status);

/* Adjoint of one step (no reference counterpart: the reference has no
 * differentiation).  Given the state before the step (x_prev, v_prev), the
 * state the step produced (x_next; the velocity follows from it), the same
 * pins and material as that step, and a loss's gradients dL/dx_next,
 * dL/dv_next (3n each), solves H lambda = dL/dx_next + (damp/dt) dL/dv_next on
 * the free DOFs, where H = d(Newton gradient)/dx at x_next is applied as the
 * analytic inertia + membrane Hessian plus central differences of the
 * bending and contact gradients, by block-Jacobi PCG to relative residual
 * tol (at most max_iters iterations), and writes dL/dx_prev, dL/dv_prev (3n)
 * and dL_dtheta[5] = dL/d(stretchWarp, stretchWeft, stretchCoupling, shear,
 * kb) with kb = (bendWarp + bendWeft) / 2, the hinge stiffness of
 * src/physics/fem.ts:44.  *iters / *rel_residual report the solve.  Contact,
 * when the scene has it, enters with the active set frozen at x_next and the
 * lagged friction normal and force held; the friction slip is measured from
 * x_prev (contact-assembly.ts:344-367), which dL/dx_prev includes; the
 * velocity filter is not differentiated. */
int cs_step_adjoint(cs_solver* s, double dt, const double* x_prev, const double* v_prev, const double* x_next,
                    const double* dL_dx_next, const double* dL_dv_next, double tol, int max_iters,
                    double* dL_dx_prev, double* dL_dv_prev, double* dL_dtheta, int* iters, double* rel_residual);

/* ---- component probes for parity tests (no reference counterpart) ------- */
/* buildSchwarzDomains (src/solver/schwarz.ts:44-140) on the host, without a
 * device: vertex -> domain id and local index (n each); *domains = count. */
int cs_schwarz_partition(const cs_scene_desc* scene, int target, int min_size, int max_size, uint32_t* domain_of,
                         uint32_t* local_of, uint32_t* domains);

/* Layout probe (host only, no device): the vertex-owner tiles of the mixed
 * solve's HVP.  Vertices in Morton order of the rest positions are taken
 * greedily into tiles of at most 256 owned vertices, 640 triangles (every
 * triangle with an owned corner, ascending) and 384 local vertices (the owned
 * ones first, then the others in order of appearance).  sizes[5] receives
 * (tiles, local vertices, tile triangles, owned-vertex offsets, entries), all
 * 0 when some vertex alone exceeds a tile; when all five buffers are given
 * they receive, per tile, 8 u32 (local begin, owned count, local count,
 * triangle begin, triangle count, entry begin, offset begin, 0), the local
 * vertex ids, per tile triangle 2 u32 (a | b << 16, c) of local corners, and
 * per owned vertex its first entry (one extra per tile) and the entries
 * slot * 4 + corner. */
int cs_hvp_tiles(const cs_scene_desc* scene, uint64_t* sizes, uint32_t* hdr, uint32_t* loc, uint32_t* lidx,
                 uint16_t* off, uint16_t* ent);
/* The HVP layout a mixed solver uses: its tile count (0: not tiled, the
 * per-vertex gather kernel runs). */
int cs_hvp_layout(cs_solver* s, uint32_t* tiles);

/* Component probe (membraneHvp with the analytic Hessian,
 * membrane-hvp.ts:18-108, plus the inertia term): one launch of the mixed
 * solve's FP32 HVP kernel at (x, p), no contact, pinned rows zeroed:
 * out = M p inv_h2 + H_membrane(x) p in FP32, returned as FP64.  Mixed
 * precision solvers only. */
int cs_probe_hvp(cs_solver* s, const double* x, const double* p, double inv_h2, double* out);
/* The same launch with the frozen barrier term (applyHvp,
 * contact-assembly.ts:371-392) of the active set built at x for a step begun
 * at x0 (updateActiveSet, :164-220), its coefficients and Jacobians rounded
 * to FP32 as the mixed solve applies them; *active (optional) receives the
 * active contact count.  Mixed precision solvers with a contact system. */
int cs_probe_hvp_contact(cs_solver* s, const double* x0, const double* x, const double* p, double inv_h2,
                         double* out, uint32_t* active);

/* The coarse level's block pattern (buildCoarsePattern, coarse-csr.ts:30-50):
 * block rows `off` (domains + 1) and up to `cap` block columns; *nnz and
 * *domains receive the sizes.  Solvers created with a coarse level only. */
int cs_coarse_pattern(cs_solver* s, uint32_t* off, uint32_t* col, uint64_t cap, uint64_t* nnz, uint32_t* domains);
/* Component probe of the two-level preconditioner at (x, diag): the FP64
 * factors and Ac, then z64 = Schwarz1(r) + P C(R r) in FP64 and z32 the same
 * with the mixed solve's FP32 factors (r rounded to FP32); raw / sym receive
 * the raw and symmetrised coarse blocks (assembleCoarseRaw,
 * assembleCoarseValues, 9 per block), dc the coarse diagonal (buildCoarseDiag,
 * mas.ts:68-101).  Outputs may be null. */
int cs_probe_two_level(cs_solver* s, const double* x, const double* diag, const double* r, double* z64, double* z32,
                       double* raw, double* sym, double* dc);
/* buildSchwarzFactors + applySchwarz (src/solver/schwarz.ts:154-309) with the
 * solver's domains (created with CS_PRECOND_SCHWARZ) at positions x and
 * Jacobi diagonal diag: z64 = the FP64 apply in the reference's order, z32 =
 * the mixed solve's FP32 factors and apply (either may be NULL); flags64 /
 * flags32 receive one byte per domain (1 Cholesky, 0 Jacobi fallback) and
 * *domains the domain count. */
int cs_probe_schwarz(cs_solver* s, const double* x, const double* diag, const double* r, double* z64, double* z32,
                     uint8_t* flags64, uint8_t* flags32, uint32_t* domains);

/* Begin a step at x0 and build the frozen active set at x (updateActiveSet,
 * contact-assembly.ts:164-220).  Writes min(count, cap) records in active-list
 * order: ids6 = (kind, p, a, b, c, d), vals7 = (w0, w1, w2, w3, nx, ny, nz). */
int cs_probe_active_set(cs_solver* s, const double* x0, const double* x, int32_t* ids6, double* vals7,
                        uint64_t cap, uint64_t* count, uint64_t* n_candidates);
/* Begin a step at x0 and list the triangle pairs the contact system expands
 * into candidates at x (buildCandidates, contact-assembly.ts:87-121: swept
 * boxes x0 -> x padded by dMinM when pad_is_dmin, else by dHatM): cloth pairs
 * (tA < tB) and cloth/static pairs (cloth triangle, static triangle), 2 u32
 * each, unordered; *n_self / *n_static receive the counts (up to the caps are
 * written). */
int cs_probe_candidates(cs_solver* s, const double* x0, const double* x, int32_t pad_is_dmin, uint32_t* self_pairs,
                        uint64_t self_cap, uint64_t* n_self, uint32_t* static_pairs, uint64_t static_cap,
                        uint64_t* n_static);
/* Component probe of the step adjoint's contact terms (no reference
 * counterpart): begins a step at x_step (beginStep, contact-assembly.ts:75-81,
 * which clears the lagged table), builds the active set at x and writes grad =
 * energyGrad at x (barrier + lagged friction, contact-assembly.ts:262-285,
 * 3n), and, when lam is not NULL, vjp = -(d grad / d x_step)^T lam: the lagged
 * friction's share of dL/dx_prev that cs_step_adjoint adds (the friction slip
 * is measured from x_step, :344-367). */
int cs_probe_contact_adjoint(cs_solver* s, const double* x_step, const double* x, const double* lam, double* grad,
                             double* vjp);
/* Begin a step at x0 and validate trial xt (checkTrial, contact-assembly.ts:489-553).
 * min_dist / min_toi are the minima recorded when the trial is valid. */
int cs_probe_check_trial(cs_solver* s, const double* x0, const double* xt, int32_t* valid, double* min_dist,
                         double* min_toi);
/* Batched primitives on device `device`: kinds[i] 0 = VT (p,a,b,c), 1 = EE
 * (a,b,c,d); ids4 index interleaved xyz arrays x0/x1 of n_verts vertices.
 * cs_probe_ccd -> vtCCD/eeCCD time of impact (ccd-vt.ts:20-115,
 * ccd-ee.ts:24-101); cs_probe_closest -> (s, t, dist) of the closest point at
 * x0 (closest-point.ts:14-112). */
int cs_probe_ccd(int device, const double* x0, const double* x1, uint32_t n_verts, const int32_t* kinds,
                 const int32_t* ids4, uint64_t count, double thickness, double* toi);
int cs_probe_closest(int device, const double* x0, uint32_t n_verts, const int32_t* kinds, const int32_t* ids4,
                     uint64_t count, double* out3);

/* ---- device buffers (no reference counterpart) ----------------------------
 * For callers of the device-pointer entry points below that hold no HIP
 * runtime of their own (a Node-API addon, a ctypes caller): allocation, a
 * synchronous copy and release through the library's own runtime, so the
 * pointers are valid for the same runtime that runs the kernels. */
enum { CS_COPY_HOST_TO_DEVICE = 1, CS_COPY_DEVICE_TO_HOST = 2, CS_COPY_DEVICE_TO_DEVICE = 3 };
int cs_device_alloc(int device, size_t bytes, void** out);
int cs_device_free(int device, void* ptr);
int cs_device_copy(int device, void* dst, const void* src, size_t bytes, int direction);

/* Component probe of the analytic hinge bending terms (no reference
 * counterpart: the reference differences hingeEnergy, bending.ts:19-42): for
 * `count` independent hinges (X12: the positions of v0, v1 (shared edge), v2,
 * v3 (opposite vertices), 12 doubles each; rest angle, edge length, area sum
 * and stiffness k as in hingeEnergy, bending.ts:7-16), grad12 = the analytic
 * gradient of the hinge energy and hvp12 = its Hessian times P12.  Host
 * arrays. */
int cs_probe_hinges(int device, const double* X12, const double* P12, const double* rest, const double* len,
                    const double* asum, double k, uint64_t count, double* grad12, double* hvp12);

/* ---- batched CCD queries ---------------------------------------------------
 * vtCCD / eeCCD (src/collision/ccd-vt.ts:20-115, ccd-ee.ts:24-101, the
 * time-of-impact queries checkTrial issues per candidate primitive,
 * contact-assembly.ts:489-553) over DEVICE arrays on `device`: kinds[i] 0 = VT
 * (p, a, b, c), 1 = EE (a, b, c, d); ids4 (16-byte aligned) index the
 * interleaved xyz arrays x0 (start) and x1 (end) of n_verts vertices; toi[i]
 * receives the time of impact in {0} u (0, 1] u {Infinity}.  An id outside
 * [0, n_verts) is CS_E_INVALID.  Synchronous on the null stream; kernel_ms
 * (optional) receives the kernel's HIP-event time. */
int cs_ccd_queries(int device, const double* x0, const double* x1, uint32_t n_verts, const int32_t* kinds,
                   const int32_t* ids4, uint64_t count, double thickness, double* toi, double* kernel_ms);

/* ---- standalone broadphase -------------------------------------------------
 * BroadPhase.build(x0, x1) (src/collision/broadphase.ts:61-64) for the cloth's
 * self pairs: the CPU golden TriBvh.build + selfPairs (src/collision/bvh.ts:
 * 19-124, exclusion of buildAdjacencyExclusions :133-151) and the GPU mirror
 * GpuBroadPhase.build (src/backend/webgpu/gpu-broadphase.ts:193-388).  Every
 * pair tA < tB of triangles whose swept boxes sweptTriAabb(x0, x1, pad)
 * (src/collision/aabb.ts:9-33) overlap and that share no vertex, sorted by
 * (tA, tB), as `pairs` = 2 u32 each.  x0, x1 (3 n_verts doubles), indices
 * (3 n_tris u32) and pairs (2 cap u32) are DEVICE pointers on `device`; the
 * call is synchronous on the null stream.  *count receives the number of
 * pairs; when it exceeds `cap` nothing is written and CS_E_OVERFLOW is
 * returned (the reference's explicit candidateOverflow, broadphase.ts:40-41),
 * so the caller can retry with a buffer of *count pairs.  A corner index >=
 * n_verts is CS_E_INVALID.  The scratch (about 132 bytes per triangle) is
 * allocated for the call; build_ms / query_ms (optional) receive the
 * hierarchy build and the pair query times (HIP events). */
int cs_broadphase_self_pairs(int device, const double* x0, const double* x1, uint32_t n_verts,
                             const uint32_t* indices, uint32_t n_tris, double pad, uint32_t* pairs, uint64_t cap,
                             uint64_t* count, double* build_ms, double* query_ms);

/* No counterpart in the reference: cs_broadphase_self_pairs for callers that
 * cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results, without the timing outputs; it
 * calls no hipMalloc / hipFree.  Its scratch is a fixed device array inside
 * the library (one copy per GPU, about 69 MB, held until the process ends),
 * which limits n_tris to 524,288 (CS_E_INVALID beyond).  Every call on a GPU
 * uses that GPU's one array: calls that overlap on one GPU (several host
 * threads, or work still running on another stream) race. */
int cs_broadphase_self_pairs_noalloc(int device, const double* x0, const double* x1, uint32_t n_verts,
                                     const uint32_t* indices, uint32_t n_tris, double pad, uint32_t* pairs,
                                     uint64_t cap, uint64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* CLOTH_HIP_H */
