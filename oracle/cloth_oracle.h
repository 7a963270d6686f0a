/*
 * cloth_oracle.h — CPU FP64 restatement of closim's CpuSolver.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product (closim_amd/) never links, loads or calls it.
 *
 * Every function follows a reference file line by line (citations on each
 * definition in cloth_oracle.cpp).  Scenes are passed with the public scene
 * struct of the product ABI (include/cloth_hip.h), so both backends consume
 * byte-identical inputs.
 */
#ifndef CLOTH_ORACLE_H
#define CLOTH_ORACLE_H

#include "../include/cloth_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_solver orc_solver;

/* ---- mesh (src/mesh/mesh.ts) ---- */
void orc_build_grid(int nx, int ny, double w, double h, float* pos, float* uv, uint32_t* idx);
/* preprocess(): returns hinge count (<= max_hinges) or -1 (degenerate triangle). */
int orc_preprocess(uint32_t n, uint32_t m, const float* pos, const float* uv, const uint32_t* idx,
                   double areal_density, float* inv_dm, float* areas, float* masses,
                   uint32_t* hinge_v, double* rest, double* len, double* asum, int max_hinges);
double orc_dihedral(const double* x, uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3);
double orc_js_hypot3(double a, double b, double c);

/* ---- element kernels (src/physics) ---- */
void orc_tri_energy_grad(const double* x0, const double* x1, const double* x2, const double* inv,
                         double area, const cs_material* mat, double* energy, double* grad9, double* E3);
void orc_tri_hvp(const double* x0, const double* x1, const double* x2, const double* v0,
                 const double* v1, const double* v2, const double* inv, double area,
                 const cs_material* mat, double* out9);
void orc_tri_hessian9(const double* x0, const double* x1, const double* x2, const double* inv,
                      double area, const cs_material* mat, double* h81);
double orc_hinge_energy(const double* x, uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3,
                        double rest, double len, double asum, double k);

/* Mesh-level (operate on a solver's scene). */
void orc_eval_internal(orc_solver* s, const double* x, double* grad, double* energy, double* max_strain);
double orc_internal_energy(orc_solver* s, const double* x);
void orc_membrane_hvp_fd(orc_solver* s, const double* x, const double* p, double* out);
void orc_membrane_hvp_analytic(orc_solver* s, const double* x, const double* p, double* out);
void orc_membrane_blocks(orc_solver* s, const double* x, double* blocks9n);

/* ---- collision (src/collision) ---- */
void orc_closest_vt(const double* p, const double* a, const double* b, const double* c, double* out6);
void orc_closest_ee(const double* a, const double* b, const double* c, const double* d, double* out6);
int orc_solve_cubic01(double a3, double a2, double a1, double a0, double* roots3);
double orc_vt_ccd(const double* x0, const double* x1, int p, int a, int b, int c, double thickness);
double orc_ee_ccd(const double* x0, const double* x1, int a, int b, int c, int d, double thickness);
double orc_barrier_value(double d, double dhat);
double orc_barrier_grad(double d, double dhat, double kappa, double w);
void orc_coulomb(const double* u, const double* n, double lambda_n, double mu, double eps, double* f3);
/* TriBvh build + selfPairs (bvh.ts): sorted (tA<tB) pairs, returns count (may exceed cap). */
int64_t orc_self_pairs(const double* x0, const double* x1, const uint32_t* idx, uint32_t m,
                       double pad, uint32_t* pairs, int64_t cap);

/* ---- solver (src/backend/cpu-solver.ts + src/solver/newton.ts) ---- */
orc_solver* orc_create(const cs_scene_desc* scene, const cs_options* opt);
void orc_destroy(orc_solver* s);
int orc_step(orc_solver* s, double dt);           /* 0 ok, CS_E_NONFINITE on blow-up */
int orc_set_static_mesh(orc_solver* o, const float* pos, uint32_t ns, const uint32_t* idx, uint32_t ms);
int orc_pin_vertex(orc_solver* s, uint32_t id, const double* pos_or_null);
int orc_unpin_vertex(orc_solver* s, uint32_t id);
void orc_set_material(orc_solver* s, const cs_material* mat);
void orc_get_positions(orc_solver* s, double* out);
void orc_get_velocities(orc_solver* s, double* out);
void orc_set_state(orc_solver* s, const double* x, const double* v);
void orc_last_stats(orc_solver* s, cs_step_stats* out);
const char* orc_last_error(void);

/* Contact-system probes for set-equality tests: begin a step at x0 and build the
 * frozen active set at x.  Returns the active count; records are written as
 * 6 ints (kind,p,a,b,c,d) + 7 doubles (w0..w3,nx,ny,nz) per contact. */
int64_t orc_coarse_pattern(orc_solver* o, int target, int minsz, int maxsz, uint32_t* off, uint32_t* col,
                           int64_t cap);
int orc_coarse_values(orc_solver* o, int target, int minsz, int maxsz, const double* x, const double* diag,
                      double* raw, double* sym, double* dc);
int orc_coarse_transfer(orc_solver* o, int target, int minsz, int maxsz, int dir, const double* in, double* out);
int orc_coarse_exact(orc_solver* o, int target, int minsz, int maxsz, const double* x, const double* diag,
                     double* ac, double* chol);
int orc_coarse_solve(orc_solver* o, int target, int minsz, int maxsz, const double* x, const double* diag, int mode,
                     int iters, double tol, const double* rhs, double* out, double* residual);
int orc_two_level_apply(orc_solver* o, int target, int minsz, int maxsz, const double* x, const double* diag,
                        int mode, int iters, double omega, const double* r, double* z);
int64_t orc_contact_hvp(orc_solver* o, const double* x0, const double* x, const double* p, double* out);
int64_t orc_contact_grad_frozen(orc_solver* o, const double* x0, const double* x, const double* y, double* grad);
int64_t orc_candidates(orc_solver* o, const double* x0, const double* x, int pad_is_dmin, int32_t* ids6, int64_t cap);
int64_t orc_contact_probe(orc_solver* s, const double* x0, const double* x, double pad_is_dmin,
                          int32_t* ids6, double* vals7, int64_t cap, int64_t* n_candidates);
/* checkTrial(xT) from step start x0: returns valid (0/1); fills minDist/minTOI. */
int orc_check_trial(orc_solver* s, const double* x0, const double* xT, double* min_dist, double* min_toi);

/* Schwarz domains (src/solver/schwarz.ts:44-140) for tests. */
int orc_schwarz_domains(orc_solver* s, int target, int minsz, int maxsz, uint32_t* domain_of,
                        uint32_t* local_of);
/* buildSchwarzFactors at x with Jacobi diagonal `diag`, then applySchwarz to r
 * (src/solver/schwarz.ts:154-309); flags[d] = 1 Cholesky, 0 Jacobi fallback.
 * Returns the domain count. */
int orc_schwarz_apply(orc_solver* s, int target, int minsz, int maxsz, const double* x, const double* diag,
                      const double* r, double* z, uint8_t* flags);

/* Summation mode of array-wide reductions: 0 (default) correctly rounded
 * (double-double), 1 the reference's plain left-to-right sums. */
void orc_set_sum_mode(int faithful);
/* Accepted Armijo trials of the last orc_step (newton.ts:204-208): 4 doubles
 * each (energy before, energy after, alpha, g.dx); returns the count. */
int64_t orc_armijo_trace(orc_solver* s, double* out4, int64_t cap);

/* ---- preconditioners and pcg (src/solver/block-jacobi.ts, src/math/pcg.ts) ---- */
void orc_block_factors(orc_solver* s, const double* x, const double* diag, double* inv9n, uint8_t* flags);
void orc_apply_block_factors(const double* inv9n, const double* r, double* z, uint32_t n);
int orc_pcg_dense(const double* A, uint32_t ndof, const double* b, const double* diag, const uint32_t* pins,
                  uint32_t npins, int max_iters, double tol, double* x, int* iters, double* residual);

/* ---- bending, static pairs, velocity filter ---- */
double orc_hinge_grad_fd(const double* x12, double rest, double len, double asum, double k, double* grad12);
int64_t orc_static_pairs(const double* x0, const double* x1, const uint32_t* idx, uint32_t m, const float* spos,
                         uint32_t ns, const uint32_t* sidx, uint32_t ms, double pad, uint32_t* pairs, int64_t cap);
int orc_kill_velocity(orc_solver* s, const double* x0, const double* x, double* v);
/* commit() (contact-assembly.ts:475-485) after freezing the active set at x:
 * the lagged table as key5 (kind, ids...) + (nx, ny, nz, lambdaN) entries. */
int64_t orc_commit_probe(orc_solver* s, const double* x0, const double* x, int64_t* keys5, double* vals4,
                         int64_t cap);

/* Fault hooks: plant a known mistake in the oracle so tests can show that the
 * checks that pin it are discriminating.  0 = none. */
void orc_set_fault(int fault);
enum { ORC_FAULT_NONE = 0, ORC_FAULT_SHEAR_4G = 1, ORC_FAULT_CCD_NO_FALLBACK = 2,
       ORC_FAULT_BARRIER_SIGN = 3, ORC_FAULT_EE_TRANSPOSE = 4, ORC_FAULT_SCHWARZ_BACK_ROW = 5,
       ORC_FAULT_PCG_BETA = 6, ORC_FAULT_BJ_COFACTOR = 7, ORC_FAULT_HINGE_HALF = 8,
       ORC_FAULT_KILL_VT_QUARTER = 9, ORC_FAULT_STATIC_NO_PAD = 10 };

#ifdef __cplusplus
}
#endif
#endif
