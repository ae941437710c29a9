/*
 * fgmi.h -- C-ABI of the MI355X (gfx950) LLG time-step engine.
 *
 * This is the drop-in boundary for FeeLLGood's per-timestep hot path. Every entry point is
 * extern "C", takes plain pointers and sizes, returns an int status and never exits or throws.
 * The context owns all device memory; host arrays passed in are copied.
 *
 * Reference interfaces replaced (FeeLLGood, C++ classes called from the time loop):
 *   fg_create             Mesh::mesh + LinAlgebra ctor + scal_fmm::fmm ctor
 *                         (src/mesh.h:38-135, src/linear_algebra.h:40-64, src/fmm_demag.h:64-90)
 *   fg_set_state          Mesh::mesh::init_distrib / readSol (src/mesh.h:221-277, src/read.cpp:199-249)
 *   fg_get_state          node getters used by savesol (src/save.cpp:164-186)
 *   fg_base_projection    LinAlgebra::base_projection (src/linear_algebra.h:105,
 *                         src/linear_algebra.cpp:3-11); the random angle is passed in explicitly
 *   fg_prepare_elements   LinAlgebra::prepareElements(const Vector3d&, const timing&)
 *                         (src/linear_algebra.h:75-76, src/linear_algebra.cpp:26-52)
 *   fg_solve              bool LinAlgebra::solve(const timing&) + get_v_max
 *                         (src/linear_algebra.h:94,101-102, src/solver.cpp:6-90)
 *   fg_calc_demag         scal_fmm::fmm::calc_demag (src/fmm_demag.h:96-101)
 *   fg_energy             Fem::energy (src/fem.h:152, src/energy.cpp:5-68)
 *   fg_evolution          Fem::evolution (src/fem.h:159-163)
 *   fg_max_angle          Mesh::mesh::max_angle (src/mesh.h:295-306)
 *   fg_step               the loop body of Fem::time_integration for a uniform field without spin
 *                         accumulation (src/time_integration.cpp:193-242): the calls above in order
 *   fg_average            Mesh::mesh::avg (src/mesh.cpp:89-106)
 *   fg_set_space_field    Mesh::mesh::setExtSpaceField (src/mesh.cpp:377-389), with
 *                         fg_prepare_elements_amp / fg_energy_amp for the R4toR3 field
 *   fg_electrostatic_potential  electrostatSolver (src/electrostatSolver.cpp:10-136)
 *   fg_spin_accumulation  spinAcc::solve (src/spinAccumulationSolver.cpp:151-287)
 *   fg_spin_create / fg_spin_solve / fg_spin_solve_ctx / fg_spin_destroy
 *                         spinAcc::spinAcc, compute, solve: the persistent solver of a run,
 *                         with the reference's accumulation of K and L_rhs (src/solver.h:110-142)
 *   fg_set_spin_accumulation    spinAcc::prepareExtraField + Tetra::calc_Hst
 *                         (src/spinAccumulationSolver.cpp:138-149, src/tetra.cpp:93-106)
 *
 * Threading: one context per host thread; calls on one context are serialised on the
 * context's HIP stream.
 */
#ifndef FGMI_H
#define FGMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (int return value of every entry point) ---- */
#define FG_OK 0
#define FG_SOLVE_FAILED 1     /* fg_solve: the Krylov solve did not converge (reference returns true) */
#define FG_STEP_REJECTED 2    /* fg_step: dt * v_max above dumax, the step is not taken */
#define FG_ERR_ARG (-1)       /* invalid argument */
#define FG_ERR_DEVICE (-2)    /* HIP runtime error */
#define FG_ERR_MESH (-3)      /* degenerate element or inconsistent mesh */
#define FG_ERR_STATE (-4)     /* call order violated */
#define FG_ERR_RANGE (-5)     /* size beyond a documented limit */

/* ---- Krylov status, same values as algebra::algoStatus (src/algebra/iter.h:24-30) ---- */
#define FG_ITER_UNDEFINED (-1)
#define FG_ITER_CONVERGED 0
#define FG_ITER_OVERFLOW 1
#define FG_ITER_CANNOT_CONVERGE 2

/* ---- which node state a call addresses (Nodes::step, src/node.h:25-29) ---- */
#define FG_CURRENT 0
#define FG_NEXT 1
#define FG_BOTH 2

/* ---- energy slots (ENERGY_TYPE, src/fem.h:30-36) ---- */
#define FG_E_EXCHANGE 0
#define FG_E_ANISOTROPY 1
#define FG_E_DEMAG 2
#define FG_E_ZEEMAN 3

/* ---- demag methods ---- */
#define FG_DEMAG_DIRECT 0     /* exact O(N*S) summation */
#define FG_DEMAG_FMM 1        /* octree fast multipole method */

/* ---- precision modes ---- */
#define FG_PREC_FP64 0        /* plain FP64 everywhere */
#define FG_PREC_MIXED 1       /* FMM demag in FP32: near-field coefficients, expansions, the
                                 translations on the FP32 matrix cores, the multipole-to-target
                                 evaluation, the far field of the v charges to degree 4 (FP64
                                 potentials and sums); everything else FP64 */

/* ---- Krylov solve (fg_solve) ---- */
#define FG_KRYLOV_FP64 0      /* bicg_dir in FP64 (src/algebra/bicg.h:162-234) */
#define FG_KRYLOV_MIXED 1     /* FP64 iterative refinement around an FP32 bicg_dir: r = b - A x in
                                 FP64, the FP32 solve of A d = r from d = 0 (FP32 stencil values,
                                 basis, blocks, preconditioner and vectors, FP64 dot products) to
                                 the relative target krylov_inner_tol, or what tol still asks if
                                 that is looser, x += d, until the FP64 residual meets tol ||b|| */

/* ---- Krylov solve of a batch (fg_create_batch) ---- */
#define FG_BATCH_UNION 0      /* one bicg_dir over the union, ||r|| <= tol min_k ||b_k|| */
#define FG_BATCH_PERSISTENT 1 /* one launch: a block of batch_block threads per instance runs that
                                 instance's bicg_dir to ||r_k|| <= tol ||b_k|| with block-level
                                 reductions and no host read between iterations (FP64; not with
                                 FG_KRYLOV_MIXED) */

/* Material of a volume region (Tetra::prm, src/tetra.h:226-252). Ms <= 0: non-magnetic. */
typedef struct fg_volume_region {
    double alpha;      /* alpha_LLG, dimensionless */
    double Ae;         /* exchange stiffness [J/m] */
    double Ms;         /* saturation magnetisation [A/m] */
    double K;          /* uniaxial anisotropy [J/m^3] */
    double uk[3];      /* uniaxial axis (unit) */
    double K3;         /* cubic anisotropy [J/m^3] */
    double ex[3];      /* cubic axes (unit) */
    double ey[3];
    double ez[3];
} fg_volume_region;

/* Material of a surface region (Triangle::prm, src/triangle.h:69-101). */
typedef struct fg_surface_region {
    double Ks;                 /* surface (Neel) anisotropy [J/m^2] */
    double uk[3];              /* surface anisotropy axis (unit) */
    int32_t suppress_charges;  /* nonzero: the triangles of this region carry no charges */
    int32_t reserved;
} fg_surface_region;

/*
 * A preprocessed mesh, as Mesh::mesh holds it after its constructor: node positions in metres,
 * nodes sorted along the longest axis, tetrahedra with positive orientation, every boundary and
 * interface triangle present with its signed dMs (src/mesh.cpp:121-245,334-367).
 */
typedef struct fg_mesh {
    int32_t n_nodes;
    int32_t n_tets;
    int32_t n_tris;
    int32_t n_vol_regions;
    int32_t n_surf_regions;
    int32_t reserved;
    const double *pos;          /* [3*n_nodes] x,y,z per node */
    const int32_t *tet_nodes;   /* [4*n_tets] 0-based node indices */
    const int32_t *tet_region;  /* [n_tets] index into vol */
    const int32_t *tri_nodes;   /* [3*n_tris] */
    const int32_t *tri_region;  /* [n_tris] index into surf */
    const double *tri_dMs;      /* [n_tris] signed Ms jump, dMs*n = Ms*n_outward */
    const fg_volume_region *vol;
    const fg_surface_region *surf;
} fg_mesh;

/* Solver and demag options. Zero-initialise, then call fg_default_options. */
typedef struct fg_options {
    double tol;            /* finite_element_solver.tolerance (default 1e-6) */
    int32_t max_iter;      /* finite_element_solver.max(iter) (default 700) */
    int32_t demag_method;  /* FG_DEMAG_DIRECT or FG_DEMAG_FMM */
    int32_t fmm_order;     /* expansion degree p of the FMM: 4, 6, 8 or 10 (default 8) */
    int32_t fmm_leaf;      /* target mean number of sources per leaf (default 64; 0 = default) */
    int32_t precision;     /* FG_PREC_FP64 or FG_PREC_MIXED */
    int32_t device;        /* HIP device ordinal */
    int32_t profile;       /* nonzero: time every kernel with HIP events */
    int32_t fmm_far_order; /* degree of the multipole evaluations of source cells three cell widths
                              from all their targets: fmm_order (one degree), or the lower far
                              degree 2, 4, 5, 6 for fmm_order 4, 6, 8, 10 (FMM error up by under
                              5 %); 0 = default: fmm_order in FP64, the lower degree in
                              FG_PREC_MIXED */
    int32_t krylov;        /* FG_KRYLOV_FP64 (default) or FG_KRYLOV_MIXED */
    double krylov_inner_tol; /* FG_KRYLOV_MIXED: relative target of each FP32 solve (0 = 3e-7) */
    int32_t batch_solver;  /* batches: FG_BATCH_UNION (default) or FG_BATCH_PERSISTENT */
    int32_t batch_block;   /* FG_BATCH_PERSISTENT: threads per instance, 256, 512 or 1024 (0 = 256) */
} fg_options;

/* Result of fg_solve (iteration<T> state, src/algebra/iter.h:36-179, and LinAlgebra::v_max). */
typedef struct fg_solve_info {
    int32_t status;        /* FG_ITER_* */
    int32_t iterations;    /* BiCGSTAB iterations (nit) */
    double residual;       /* last monitored residual norm */
    double rhs_norm;       /* ||b|| */
    double v_max;          /* gamma0 * max_i |(x_2i, x_2i+1)| [1/s] */
    int32_t refinements;   /* FG_KRYLOV_MIXED: FP32 solves, each followed by an FP64 correction;
                              else 0 */
    int32_t inner_iterations; /* FG_KRYLOV_MIXED: FP32 iterations in all (= iterations); else 0 */
} fg_solve_info;

typedef struct fg_ctx fg_ctx;

void fg_default_options(fg_options *opt);

/* Build the device plan (matrix pattern, gather lists, geometry, octree) and upload the mesh. */
int fg_create(const fg_mesh *mesh, const fg_options *opt, fg_ctx **out);
/* Batched independent instances: k meshes held side by side in one context, so that every verb
 * below runs each kernel once over all of them. Nothing couples the instances: each has its own
 * octree (the demag of one never sees another), and BiCGSTAB stops when ||r|| <= tol min_k ||b_k||
 * (k over instances with b_k != 0), which meets every instance's own criterion ||r_k|| <= tol
 * ||b_k||. States are exchanged for the union: node ranges as fg_batch_layout reports them, in
 * input order. One basis angle and one applied field serve all instances. No counterpart in the
 * reference (one geometry per process there). */
int fg_create_batch(const fg_mesh *meshes, int k, const fg_options *opt, fg_ctx **out);
/* Instance layout of a context: k, node offsets [k+1] and magnetic-tetrahedron offsets [k+1]
 * (either array may be NULL; a plain context reports k = 1). */
int fg_batch_layout(fg_ctx *ctx, int64_t *node_off, int64_t *mtet_off, int *k);
int fg_destroy(fg_ctx *ctx);

/* Copy a node state in (which = FG_CURRENT, FG_NEXT or FG_BOTH). NULL pointers write zeros.
 * u, v: [3*n_nodes]; phi, phiv: [n_nodes]. */
int fg_set_state(fg_ctx *ctx, int which, const double *u, const double *v, const double *phi,
                 const double *phiv);
/* Copy a node state out (which = FG_CURRENT or FG_NEXT). NULL pointers are skipped. */
int fg_get_state(fg_ctx *ctx, int which, double *u, double *v, double *phi, double *phiv);

/* Tangent basis of every node from u CURRENT, rotated by the angle r (the reference draws
 * r = M_2_PI * U(0,1) from std::mt19937(rand())). */
int fg_base_projection(fg_ctx *ctx, double r);
/* Element pass for a uniform applied field Hext [A/m] (B/mu0); dt and the timing prefactor. */
int fg_prepare_elements(fg_ctx *ctx, const double Hext[3], double dt, double prefactor);
/* Assemble, build the initial guess, run BiCGSTAB, update the NEXT state. Returns FG_OK,
 * FG_SOLVE_FAILED (status ITER_OVERFLOW / CANNOT_CONVERGE / residual above tol) or an error. */
int fg_solve(fg_ctx *ctx, double dt, fg_solve_info *info);
/* Demag potentials phi (from u NEXT) and phi_v (from v NEXT), written to the NEXT state. */
int fg_calc_demag(fg_ctx *ctx);
/* Energies of the NEXT state for a uniform applied field Hext [A/m]; E[4] in joules. */
int fg_energy(fg_ctx *ctx, const double Hext[3], double E[4]);
/* Space-dependent applied field (the reference's R4toR3 Bext: H(x, t) = A(t) h(x), evaluated at the
 * tetrahedra's Gauss points, src/mesh.cpp:377-389): h[15 * t + 3 g + k] is component k of h at
 * Gauss point g of mesh tetrahedron t, in A/m, for every tetrahedron of the fg_mesh (non-magnetic
 * ones are ignored); NULL removes it. The Gauss point positions come from fg_gauss_points. */
int fg_set_space_field(fg_ctx *ctx, const double *h);
/* Positions of the 5 Gauss points of every tetrahedron of a mesh (the engine's rule, src/tetra.h:50-60,
 * node order as in the mesh): pg[15 * t + 3 * g + k]. Needs no context. */
int fg_gauss_points(const fg_mesh *mesh, double *pg);
/* LinAlgebra::prepareElements(double A_Hext, const timing&) (src/linear_algebra.cpp:54-75): the
 * element pass with the applied field amp * h at each Gauss point. FG_ERR_STATE without a space
 * field. */
int fg_prepare_elements_amp(fg_ctx *ctx, double amp, double dt, double prefactor);
/* Fem::energy with an R4toR3 field (src/energy.cpp:11-16, src/tetra.cpp:381-391): the Zeeman
 * term from amp * h at the Gauss points. FG_ERR_STATE without a space field. */
int fg_energy_amp(fg_ctx *ctx, double amp, double E[4]);

/* Spin-diffusion parameters of a volume region (Tetra::prm, src/tetra.h: sigma [S/m], N0 density
 * of states [J^-1 m^-3], P polarization rate in [0, 1), lsd and lsf [m]) and of a surface region
 * (Triangle::prm, src/triangle.h: jn [A/m^2] with the polarization direction uP, or the spin
 * accumulation s [A/m] on the surface; NaN where the region sets none). */
typedef struct fg_spin_volume {
    double sigma, N0, P, lsd, lsf;
} fg_spin_volume;
typedef struct fg_spin_surface {
    double jn;
    double uP[3];
    double s[3];
} fg_spin_surface;

/* The spin-accumulation field of the element pass (spinAcc::prepareExtraField and Tetra::calc_Hst,
 * src/spinAccumulationSolver.cpp:138-149, src/tetra.cpp:93-106): Hst = D0 / (lsd^2 gamma0 Ms)
 * sum_i s_i a_ig at the Gauss points of every magnetic tetrahedron, from the node spin
 * accumulation s[3 n_nodes] (fg_spin_accumulation) and the regions' spin parameters. Every later
 * fg_prepare_elements / fg_prepare_elements_amp adds Hst to the effective field of alpha_eff and
 * to the right-hand side with its second-order term (Tet::integrales, src/tetra.cpp:248-255,
 * 296-302) until it is replaced or removed (s = NULL). */
int fg_set_spin_accumulation(fg_ctx *ctx, const double *s, const fg_spin_volume *vol);

/* The electrostatic potential of the current-injection problem on every node of a mesh
 * (electrostatSolver::checkBoundaryConditions, solve, src/electrostatSolver.cpp:10-28,96-136,
 * built by spinAcc::spinAcc, src/spinAccumulationSolver.cpp:21-30): conductivity sigma per volume
 * region (the tetrahedra's Tetra::prm::sigma), current density jn and potential V per surface
 * region (Triangle::prm::jn, V; NaN where the region sets none), algebra::cg_dir (src/algebra/cg.h:
 * 67-121) from V = 0 with tolerance tol and max_iter iterations on HIP device `device` (made
 * current). Returns FG_ERR_ARG unless exactly one surface region has a finite jn and one a finite V,
 * FG_ERR_MESH on an out-of-range index, FG_ERR_DEVICE on a HIP error, else FG_OK with the potential
 * in V_out[n_nodes] and the solver's status, iterations, residual and ||b|| in *info (the reference
 * exits unless the status is FG_ITER_CONVERGED). Needs no context. */
int fg_electrostatic_potential(const fg_mesh *mesh, const double *sigma, const double *jn, const double *V,
                               double tol, int32_t max_iter, int32_t device, double *V_out, fg_solve_info *info);
/* fg_electrostatic_potential without any device allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). Same arguments, kernels, results and return codes; every device
 * buffer of the set-up and the solve comes from one static scratch array of FG_ES_SCRATCH_BYTES per
 * GPU, held until the process ends. A mesh whose set-up needs more (about 1.8 kB per node of a
 * tetrahedral mesh, i.e. some 35,000 nodes) returns FG_ERR_RANGE before any kernel runs. Calls on
 * one GPU share that array: calls that overlap on one GPU race. */
#define FG_ES_SCRATCH_BYTES ((size_t)64 << 20)
int fg_electrostatic_potential_noalloc(const fg_mesh *mesh, const double *sigma, const double *jn, const double *V,
                                      double tol, int32_t max_iter, int32_t device, double *V_out,
                                      fg_solve_info *info);
/* The spin accumulation of the spin-diffusion problem (spinAcc::checkBoundaryConditions,
 * boundaryConditions, solve, integrales, src/spinAccumulationSolver.cpp:41-287): the 3 n_nodes
 * system of the diffusion (D0 = 2 sigma / (e^2 N0), spin-flip length lsf) and, in the tetrahedra
 * whose region has Ms > 0, the s-d exchange with the node magnetization u[3 n_nodes] (length lsd)
 * and the source of the potential V[n_nodes] (fg_electrostatic_potential) along u_src[3 n_nodes]
 * (NULL: u). The reference reads the exchange from the CURRENT state and the source from the NEXT
 * one (Tet::calcOffDiagBlock, src/tetra.cpp:142-148; mesh::getNode_u, src/mesh.h:156), which differ
 * when compute_all runs between a solve and the evolution (src/fem.h:203-224). Dirichlet values on the
 * surfaces with s, or with jn (s = -jn mu_B / e uP). algebra::bicg_dir (6-argument overload,
 * src/algebra/bicg.h:82-158) from s = 0, tolerance tol, max_iter iterations, on HIP device `device`
 * (made current). Returns FG_ERR_ARG unless exactly one surface region has a finite jn and one a
 * finite s and every volume region a finite nonzero N0 and 0 <= P < 1, FG_ERR_MESH on an
 * out-of-range index, FG_ERR_DEVICE on a HIP error, else FG_OK with s in s_out[3 n_nodes]
 * (interleaved x, y, z per node) and the solver's status in *info (the reference zeroes s unless
 * the status is FG_ITER_CONVERGED). Needs no context. */
int fg_spin_accumulation(const fg_mesh *mesh, const fg_spin_volume *vol, const fg_spin_surface *surf, const double *V,
                         const double *u, const double *u_src, double tol, int32_t max_iter, int32_t device,
                         double *s_out, fg_solve_info *info);
/* The persistent spin-diffusion solver of a run (spinAcc, src/spinAccumulationSolver.h; built by
 * spinAcc::spinAcc, src/spinAccumulationSolver.cpp:8-36, and called by every compute_all,
 * src/fem.h:203-224). fg_spin_create checks the boundary conditions and builds the node pattern,
 * the Dirichlet values, the surface right-hand side and every device buffer once, on HIP device
 * `device`, from the mesh, the regions' spin parameters and the node potential V[n_nodes]; return
 * codes as fg_spin_accumulation's, *out NULL on error. */
typedef struct fg_spin fg_spin;
int fg_spin_create(const fg_mesh *mesh, const fg_spin_volume *vol, const fg_spin_surface *surf, const double *V,
                   double tol, int32_t max_iter, int32_t device, fg_spin **out);
/* One solve (spinAcc::solve, src/spinAccumulationSolver.cpp:165-209) at the node magnetizations
 * u[3 n_nodes] (s-d exchange) and u_src[3 n_nodes] (source; NULL: u), from s = 0, with the
 * BiCGSTAB scalars kept on the device (one host read per 16 iterations). With accumulate nonzero
 * the new element and surface terms are ADDED to the system of the earlier solves, as in the
 * reference, whose solver::buildMat / buildVect add into K and L_rhs (src/solver.h:110-142) and
 * whose spinAcc never clears them, so that its k-th solve is that of the sum of k assemblies; with
 * accumulate zero they replace it. s_out[3 n_nodes] and *info as fg_spin_accumulation's (s is
 * returned whatever the status). */
int fg_spin_solve(fg_spin *sp, const double *u, const double *u_src, int32_t accumulate, double *s_out,
                  fg_solve_info *info);
/* fg_spin_solve at an engine's own state, as compute_all does (src/fem.h:203-224): the s-d
 * exchange from the CURRENT state and the source from the NEXT one, read on the device and
 * solved on the context's stream; s then becomes the context's spin-accumulation field (as
 * fg_set_spin_accumulation with the solver's regions), or zero when the solver did not converge
 * (spinAcc::compute, src/spinAccumulationSolver.cpp:151-163). The solver must be built on the
 * context's mesh, on its device (FG_ERR_ARG otherwise). s_out[3 n_nodes] (may be NULL) receives s. */
int fg_spin_solve_ctx(fg_spin *sp, fg_ctx *ctx, int32_t accumulate, double *s_out, fg_solve_info *info);
int fg_spin_destroy(fg_spin *sp);
/* NEXT -> CURRENT. */
int fg_evolution(fg_ctx *ctx);
/* acos(min over magnetic edges of u_i . u_j) on the NEXT state. */
int fg_max_angle(fg_ctx *ctx, double *angle);
/* One step of the reference's time loop body for a uniform applied field Hext [A/m] and no spin
 * accumulation (src/time_integration.cpp:193-242): base_projection(r), prepareElements(Hext, dt),
 * solve and get_v_max; then, when the solve succeeded and dt * v_max <= dumax (dumax <= 0: no
 * limit), compute_all (calc_demag, energy, evolution; src/fem.h:203-224) and max_angle. The same
 * kernels as the separate calls, with one host read of the results at the end instead of one per
 * call. Returns FG_OK (step taken: info, E[4] of the new state, *max_angle), FG_SOLVE_FAILED (the
 * solve failed, nothing evolved: the reference retries with a smaller dt), FG_STEP_REJECTED
 * (dt * v_max > dumax: NEXT holds the solution, nothing evolved) or an error. */
int fg_step(fg_ctx *ctx, const double Hext[3], double dt, double prefactor, double r, double dumax,
            fg_solve_info *info, double E[4], double *max_angle);
/* Volume average of u NEXT (which = 0) or v NEXT (which = 1) over the magnetic tetrahedra. */
int fg_average(fg_ctx *ctx, int which, double out[3]);

/* ---- introspection used by the parity tests and the benchmark ---- */
/* Basis vectors [3*n_nodes] each. */
int fg_get_basis(fg_ctx *ctx, double *ep, double *eq);
/* Number of rows (2*n_nodes) and stored entries of the assembled matrix (4*(N + 2E)). */
int fg_system_size(fg_ctx *ctx, int64_t *rows, int64_t *nnz);
/* Materialise the assembled (unmasked) matrix in CSR with sorted columns and the masked RHS, as
 * solver<2>::K and L_rhs hold them after assembly (src/solver.cpp:11-48). */
int fg_get_system(fg_ctx *ctx, int64_t *rowptr, int32_t *col, double *val, double *rhs);
/* The last solution vector x (= v/gamma0 in the tangent basis), [2*n_nodes]. */
int fg_get_solution(fg_ctx *ctx, double *x);
/* Device-side per-kernel timing (opt.profile != 0): name, launches and summed milliseconds of
 * kernel k. Returns FG_ERR_RANGE past the last kernel. */
int fg_kernel_stats(fg_ctx *ctx, int k, char *name, int name_len, int64_t *launches, double *ms);
int fg_reset_kernel_stats(fg_ctx *ctx);
/* Host (CPU) evaluation of the same octree FMM as the device pass: potentials sum_s q_s/|x_t - y_s|
 * at T targets from S sources (xyz arrays of 3*n doubles), for two charge sets. No GPU needed;
 * used to check the device pass and the expansion orders (far_order as fg_options.fmm_far_order).
 * leaf_level receives the tree depth. */
int fg_fmm_reference(const double *src_xyz, int64_t S, const double *tgt_xyz, int64_t T,
                     const double *q_u, const double *q_v, int order, int far_order, int leaf,
                     double *pot_u, double *pot_v, int *leaf_level);
/* ---- learned initial guess ----
 * No counterpart in the reference: when enabled, a per-node network (features u, the tangent
 * right-hand side over the lumped mass, the previous v, and their mean over mesh-edge neighbours;
 * 18 -> 64 -> 64 -> 3, GELU, BF16 weights, FP32 arithmetic) replaces the initial guess that
 * LinAlgebra::buildInitGuess forms (src/linear_algebra.cpp:13-24, the previous v) by a correction
 * of its linear extrapolation v + (dt / dt_prev)(v - v_prev) over the last two solutions (the
 * previous v alone until two solves have followed fg_set_state). BiCGSTAB then solves to the same
 * tolerance, so the guess changes iteration counts, not results beyond the tolerance.
 * Its inputs are the node's u, tangent right-hand side over the lumped mass, previous v and the
 * second difference of the last three solutions, and their means over the mesh neighbours.
 * Parameter block (FP32): W1[64][24], b1[64], W2[64][64], b2[64], W3[3][64], b3[3], then the
 * scales s_h, s_v, s_out, s_d (set from the first training batch when zero). */
#define FG_GUESS_NPARAM 5959
/* Allocate the network with random weights (W3 = 0: it starts as the extrapolation). Every
 * later fg_solve records the features of its assembly, so plain solves provide training data. */
int fg_guess_init(fg_ctx *ctx, uint64_t seed);
/* Load or save the parameter block (n must be FG_GUESS_NPARAM); set allocates when needed. */
int fg_guess_set(fg_ctx *ctx, const float *params, int64_t n);
int fg_guess_get(fg_ctx *ctx, float *params, int64_t n);
/* Use the network's guess in fg_solve (on != 0) or the reference guess (on == 0). */
int fg_guess_enable(fg_ctx *ctx, int on);
/* One Adam step (learning rate lr) on the last successful solve: features of its assembly,
 * target its solution. loss: the network's loss before the step; loss_ref: the reference
 * guess's loss on the same batch (both normalised by the extrapolation's error on the first
 * training batch, so that an untrained network scores about 1 there). */
int fg_guess_train(fg_ctx *ctx, float lr, double *loss, double *loss_ref);
/* Gradient of the loss on the last solve at the current parameters, [FG_GUESS_NPARAM - 4]. */
int fg_guess_gradient(fg_ctx *ctx, float *grad, double *loss, double *loss_ref);
/* Raw features of the last solve, [16 * n_nodes]: u, the tangent RHS over the mass, previous v,
 * the second difference of the last three solutions, the extrapolated base the network corrects,
 * and one unused value. */
int fg_guess_features(fg_ctx *ctx, float *feat);

/* Octree FMM plan of an FG_DEMAG_FMM context (FG_ERR_STATE otherwise): info[0] leaf level,
 * [1] cells, [2] leaves, [3] evaluation groups, [4] source-list entries, [5] multipole-to-target
 * pairs, [6] M2L pairs, [7] near-field coefficients, [8] near-field shared columns, [9] targets,
 * [10] sources, [11] expansion order, [12] leaves with targets, [13] M2M parent cells, [14] lane
 * slots of the multipole-to-target launch (pairs / slots = its SIMD efficiency), [15] pairs of [5]
 * evaluated at the far degree. Used by the benchmark's work models. */
int fg_fmm_info(fg_ctx *ctx, int64_t info[16]);
/* Wait for all work queued on the context's stream. */
int fg_synchronize(fg_ctx *ctx);
/* Last error message of this context (or of fg_create when ctx is NULL). */
const char *fg_last_error(fg_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* FGMI_H */
