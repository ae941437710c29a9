/*
 * oracle.h -- FP64 CPU restatement of FeeLLGood's per-timestep LLG path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the parity tests, the smoke check and
 * the CPU-baseline leg of the benchmark compare against. The product (libfgmi) never links,
 * loads or calls it.
 *
 * Every function restates the reference algorithm at the cited file:line; the reference itself
 * (Eigen, TBB, ScalFMM, gmsh, yaml-cpp, Duktape) cannot be built in this image. The demag
 * potential is computed by exact direct summation instead of ScalFMM's FMM.
 */
#ifndef FG_ORACLE_H
#define FG_ORACLE_H

#include <stdint.h>
#include "../include/fgmi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_problem orc_problem;

typedef struct orc_solve_info {
    int32_t status;
    int32_t iterations;
    double residual;
    double rhs_norm;
    double v_max;
} orc_solve_info;

/* ---- whole-problem API (mirrors fgmi.h) ---- */
orc_problem *orc_create(const fg_mesh *mesh, double tol, int max_iter);
void orc_destroy(orc_problem *p);
const char *orc_last_error(void);
int orc_n_edges(const orc_problem *p);
int orc_n_mag_tris(const orc_problem *p);
void orc_set_state(orc_problem *p, int which, const double *u, const double *v, const double *phi,
                   const double *phiv);
void orc_get_state(const orc_problem *p, int which, double *u, double *v, double *phi,
                   double *phiv);
void orc_base_projection(orc_problem *p, double r);
void orc_get_basis(const orc_problem *p, double *ep, double *eq);
void orc_prepare_elements(orc_problem *p, const double Hext[3], double dt, double prefactor);
/* Returns 1 on solver failure (reference LinAlgebra::solve returns true), else 0. */
int orc_solve(orc_problem *p, double dt, orc_solve_info *info);
/* algebra::bicg_dir on an arbitrary CSR matrix (sorted columns); returns the status. */
int orc_bicg_dir(int64_t n, const int64_t *rowptr, const int32_t *col, const double *val,
                 const double *rhs, double *x, int n_mask, const int32_t *mask_idx, double tol,
                 int max_iter, orc_solve_info *info);
/* algebra::cg_dir on an arbitrary CSR matrix with Dirichlet values xd on mask_idx; returns the status. */
int orc_cg_dir(int64_t n, const int64_t *rowptr, const int32_t *col, const double *val, const double *rhs,
               const double *xd, double *x, int n_mask, const int32_t *mask_idx, double tol, int max_iter,
               orc_solve_info *info);
/* The electrostatic potential V on the nodes (electrostatSolver::solve): sigma per volume region,
   jn and V per surface region (NaN: none). Returns -1 on bad boundary conditions, else the status. */
int orc_electrostatic_potential(const fg_mesh *m, const double *sigma, const double *jn, const double *Vbc,
                                double tol, int max_iter, double *V, orc_solve_info *info);
/* The spin-accumulation field Hst of the element pass from s[3 n_nodes]; NULL removes it. */
void orc_set_spin_accumulation(orc_problem *p, const double *s, const fg_spin_volume *vp);
/* The spin accumulation s[3 n_nodes] (spinAcc::solve): returns -1 on bad boundary conditions,
   else the bicg_dir status. */
int orc_spin_accumulation(const fg_mesh *m, const fg_spin_volume *vp, const fg_spin_surface *sp, const double *V,
                          const double *u, const double *u_src, double tol, int max_iter, double *s_out,
                          orc_solve_info *info);
/* The k = nhist-th solve of a run (the reference's solver never clears K and L_rhs between the
   calls of spinAcc::solve): the system summed over the magnetizations uh[nhist][3 n_nodes] and
   sources ush (NULL: uh). */
int orc_spin_accumulation_hist(const fg_mesh *m, const fg_spin_volume *vp, const fg_spin_surface *sp,
                               const double *V, int nhist, const double *uh, const double *ush, double tol,
                               int max_iter, double *s_out, orc_solve_info *info);
void orc_system_size(const orc_problem *p, int64_t *rows, int64_t *nnz);
void orc_get_system(const orc_problem *p, int64_t *rowptr, int32_t *col, double *val,
                    double *rhs);
void orc_get_solution(const orc_problem *p, double *x);
/* Element matrices of tet k after orc_prepare_elements: Kp[64] row-major, Lp[8]. */
void orc_get_tet_element(const orc_problem *p, int k, double *Kp, double *Lp);
void orc_calc_demag(orc_problem *p);
/* Demag method of orc_calc_demag: order 0 = exact direct sums (default); order > 0 = the CPU FMM
 * (fmm_cpu.cpp) of that order, leaves of `leaf` sources on average. Returns 0, or -1 on bad values. */
int orc_set_demag(orc_problem *p, int order, int leaf);
/* The CPU FMM on raw sources and targets: sum_s q_s / |x_t - y_s| for both charge sets (no
 * corrections, no 1/(4 pi)); returns the leaf level. */
int orc_fmm_potentials(const double *sxyz, int64_t S, const double *qu, const double *qv, const double *txyz,
                       int64_t T, int order, int leaf, double *pu, double *pv);
/* Charges of u and v NEXT and the exact potentials at n targets only (state untouched). */
void orc_demag_sample(orc_problem *p, const int32_t *targets, int n, double *phi, double *phiv);
/* Charges of the last demag call: sources [n_src] (tets then tris), corrections [n_nodes]. */
int64_t orc_n_sources(const orc_problem *p);
void orc_get_sources(const orc_problem *p, double *xyz, double *q_u, double *q_v);
void orc_get_corrections(const orc_problem *p, double *corr_u, double *corr_v);
void orc_energy(orc_problem *p, const double Hext[3], double E[4]);
/* Space-dependent applied field (R4toR3): h[15 t + 3 g + k] at Gauss point g of tet t, A/m; NULL
 * removes it. The _amp forms use amp * h in place of the uniform field
 * (src/linear_algebra.cpp:54-75, src/energy.cpp:11-16, src/tetra.cpp:381-391). */
void orc_set_space_field(orc_problem *p, const double *h);
void orc_prepare_elements_amp(orc_problem *p, double amp, double dt, double prefactor);
void orc_energy_amp(orc_problem *p, double amp, double E[4]);
void orc_evolution(orc_problem *p);
double orc_max_angle(const orc_problem *p);
void orc_average(const orc_problem *p, int which, double out[3]);

/* ---- elementary restatements, used by the known-answer tests ---- */
/* Tet constructor: orientation (swaps ind[2], ind[3] when needed), da = dadu*inv(J), weights.
 * pos[12] holds the 4 vertex positions in the order of ind. Returns 0, or -1 when singular. */
int orc_tet_geometry(const double *pos, int32_t *ind, double *da, double *w);
/* Tetra::calc_alpha_eff for NPI values of uHeff. */
void orc_alpha_eff(double dt, double alpha, const double *uHeff, int n, double *out);
/* Tet::lumping: AE[144] row-major (+=). */
void orc_tet_lumping(const double *da, const double *w, const double *u_nod,
                     const double *alpha_eff, double prefactor, double *AE);
/* Tet::integrales for one isolated tet: state arrays are per local node (u[12], v[12], phi[4],
 * phiv[4], ep[12], eq[12]); Hext[3] uniform; outputs Kp[64], Lp[8]. */
void orc_tet_integrales(const double *da, const double *w, const double *u, const double *v,
                        const double *phi, const double *phiv, const double *ep,
                        const double *eq, const fg_volume_region *prm, const double *Hext,
                        double dt, double prefactor, double *Kp, double *Lp);
/* Tet::calc_aniso_uniax / calc_aniso_cub on Gauss-point values U[15], V[15] ([5][3]): H[15] and
 * the uHeff contributions contrib[5]. */
void orc_aniso_uniax(const double *U, const double *V, const double *uk, double Kbis, double s_dt,
                     double *H, double *contrib);
void orc_aniso_cub(const double *U, const double *V, const double *ex, const double *ey, const double *ez,
                   double K3bis, double s_dt, double *H, double *contrib);
/* Energies of one tetrahedron, e[5] = exchange, uniaxial, cubic, demag, Zeeman (field Hg[15] at the
 * Gauss points), and of one triangle, e[2] = surface anisotropy, demag. */
void orc_tet_energy(const double *da, const double *w, const double *u, const double *phi,
                    const fg_volume_region *prm, const double *Hg, double *e);
void orc_tri_energy(const double *pos, const double *u, const double *phi, double dMs, double Ks,
                    const double *uk, double *e);
/* Charges of one tetrahedron (q[5]) and one triangle (q[4]). */
void orc_tet_charges(const double *da, const double *w, const double *u, double Ms, double *q);
void orc_tri_charges(const double *pos, const double *u, double dMs, double *q);
/* Tri::integrales of one triangle: Lp[6]. */
void orc_tri_integrales(const double *pos, const double *u, const double *ep, const double *eq, double dMs,
                        double Ks, const double *uk, double *Lp);
/* Gauss tables a[4][5], pds[5] (tetrahedron) and a[3][4], pds[4] (triangle). */
void orc_gauss_tables(double *ta, double *tpds, double *fa, double *fpds);
/* Tri::potential at local vertex i: pos[9], s[3] = getter(node).n values, surf, dMs. */
double orc_tri_potential(const double *pos, const double *s, double dMs, int i);
/* Node::setBasis on one node. */
void orc_set_basis(const double *u, double r, double *ep, double *eq);

/* ---- reproducible random numbers ---- */
/* n draws of std::uniform_real_distribution<>(0,1) from std::mt19937(seed). */
void orc_mt_uniform(uint32_t seed, int n, double *out);
/* srand(seed), then n angles M_2_PI * U(0,1) drawn as LinAlgebra::base_projection does. */
void orc_basis_angles(uint32_t seed, int n, double *out);

/* ---- the reference time loop (src/time_integration.cpp:136-254) ---- */
typedef struct orc_run_settings {
    double final_time;
    double evol_time_step;
    double dt_min;
    double dt_max;
    double du_max;
    double tol;
    int32_t max_iter;
    uint32_t seed;
    double Bext[3];           /* tesla, constant */
    int32_t max_rows;         /* capacity of evol (rows) */
    int32_t verbose;
} orc_run_settings;
/* evol rows of 12 doubles: t, <Mx>, <My>, <Mz>, E_ex, E_aniso, E_demag, E_zeeman, E_tot, dt,
 * max_dm, nt. u0 [3*n_nodes] initial magnetisation. Returns the exit status (0 ok, 1 aborted)
 * and writes the number of rows to *n_rows. */
int orc_run(const fg_mesh *mesh, const double *u0, const orc_run_settings *s, double *evol,
            int32_t *n_rows);

/* The engine's learned-guess network restated (see oracle.cpp): predictions v_pred [3n], loss,
 * the reference guess's loss and the gradient [FG_GUESS_NPARAM - 3]; any output may be NULL, and
 * v_base [3n] (the base the network corrects) may be NULL for the previous v. */
void orc_guess_eval(const float *params, int round_bf16, int n, const float *feat, const float *v_base,
                    const int64_t *adj_ptr, const int32_t *adj_idx, const int32_t *mag, const double *v_target,
                    double *v_pred, double *loss, double *loss_ref, double *grad);

#ifdef __cplusplus
}
#endif

#endif
