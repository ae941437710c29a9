// ctx.h -- the engine context: device buffers, the host-built plan and the launchers.
#pragma once

#include <string>
#include <vector>

#include "../../include/fgmi.h"
#include "common.h"

struct FmmPlan;  // fmm.hip

namespace fg {

struct GuessNet;  // guess.hip

struct KernelStat {
    std::string name;
    int64_t launches = 0;
    double ms = 0;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // a second stream (default priority) on which the FMM's near field, M2L and L2L run beside M2P:
    // ev_mid (multipoles complete) forks it, ev_join joins it before the final per-target sum
    hipStream_t aux = nullptr;
    hipEvent_t ev_mid = nullptr, ev_join = nullptr;
    fg_options opt{};
    std::string err;

    // sizes
    int N = 0;         // nodes
    int Tm = 0;        // magnetic tetrahedra
    std::vector<int32_t> h_tet_mesh;  // mesh index of each magnetic tetrahedron
    double *hsp = nullptr;            // space-dependent applied field [15][Tm] (fg_set_space_field)
    double *hst = nullptr;            // spin-accumulation field Hst [15][Tm] (fg_set_spin_accumulation)
    double *spin_s = nullptr;         // node spin accumulation [3 N] and Hst prefactor per region
    double *spin_pref = nullptr;
    std::vector<double> h_reg_Ms;     // Ms per volume region
    int Fm = 0;        // charged magnetic triangles (magTri)
    int Fk = 0;        // triangles with surface anisotropy in the RHS
    int64_t S = 0;     // demag sources = 5 Tm + 4 Fm
    int nslices = 0;   // SELL slices of the stiffness stencil
    int64_t nnzS = 0;  // stored stencil entries (with padding)
    int64_t nnz_real = 0;  // stencil entries without padding = N + 2 E_mag
    int n_regions = 0;
    double totalMagVol = 0;

    // host copies kept for introspection
    std::vector<int32_t> h_mag;          // [N]
    // independent instances held side by side (one unless the context is a batch): node,
    // magnetic-tetrahedron and charged-triangle ranges, size K + 1
    std::vector<int64_t> inst_node_off, inst_mtet_off, inst_mtri_off;
    std::vector<int64_t> in_node_off, in_tet_off, in_tri_off;  // the same ranges in the input mesh
    int64_t *d_inst_node_off = nullptr;  // batch only: node ranges on the device
    double *inst_bn2 = nullptr;          // batch only: ||b_k||^2 per instance
    int32_t *d_node_inst = nullptr;      // batch only: instance of each node
    longlong4 *d_inst_src = nullptr;     // batch only: tetrahedron and triangle source ranges
    int32_t *d_inst_order = nullptr;     // batch only: instances by decreasing node count
    int2 *d_inst_stat = nullptr;         // batch only (FG_BATCH_PERSISTENT): status, iterations
    double2 *d_inst_resn = nullptr;      //   and (residual, ||b_k||) per instance, pinned mirrors
    int2 *h_inst_stat = nullptr;
    double2 *h_inst_resn = nullptr;
    double charge_unit = 1.0;            // typical |charge| (Ms_max w^(2/3)), scales the FP32 FMM pass
    std::vector<int64_t> h_rowptr;       // stencil CSR (unpadded), for fg_get_system
    std::vector<int32_t> h_col;
    std::vector<double> h_val;

    // node data
    NodeRec *rec[2] = {nullptr, nullptr};  // CURRENT = rec[cur], NEXT = rec[1-cur]
    int cur = 0;
    double4 *pos4 = nullptr;  // node positions
    double4 *ep4 = nullptr, *eq4 = nullptr;
    int32_t *mag = nullptr;
    double *mass = nullptr;   // lumped mass m_i = sum_t sum_g a_ig w_g
    double *sdiag = nullptr;  // S'_ii
    double4 *Dg = nullptr;    // per-node 2x2 block without the stencil: D_i B(i,i) + m_i X(i)
    double2 *dprec = nullptr; // Jacobi preconditioner, masked
    // Krylov vectors (2 unknowns per node)
    double2 *b = nullptr, *x = nullptr, *r = nullptr, *rt = nullptr, *p = nullptr, *ph = nullptr,
            *v = nullptr, *s = nullptr, *sh = nullptr, *t = nullptr;
    double4 *w = nullptr;     // lifted 3-D vector fed to the stencil
    KrylovCtl *ctl = nullptr;
    KrylovCtl *h_ctl = nullptr;  // pinned host mirror
    // mixed-precision Krylov (FG_KRYLOV_MIXED): FP32 stencil values, per-step basis, diagonal
    // blocks and preconditioner, the FP32 loop's vectors and its control block
    float *svalf = nullptr;
    float4 *ep4f = nullptr, *eq4f = nullptr, *Dgf = nullptr, *wf = nullptr;
    float2 *dprecf = nullptr, *rf = nullptr, *rtf = nullptr, *pf = nullptr, *phf = nullptr, *vf = nullptr,
           *sf = nullptr, *shf = nullptr, *tf = nullptr, *df = nullptr;
    KrylovCtl *ic = nullptr;
    KrylovCtl *h_ic = nullptr;
    double *part = nullptr;      // [8][MAX_PART] partial sums
    double *h_part = nullptr;    // pinned host mirror of small results

    // stiffness stencil S' (SELL-64)
    int64_t *slice_ptr = nullptr;
    int32_t *slice_w = nullptr;
    int32_t *scol = nullptr;
    double *sval = nullptr;

    // tetrahedra (magnetic only)
    int4 *tet_nodes = nullptr;
    double *geo = nullptr;    // SoA [17][Tm]: da (12, row-major 4x3), w (5)
    int32_t *tet_reg = nullptr;
    DevRegion *regions = nullptr;
    double4 *elem_out = nullptr;  // [4 Tm + 3 Fk] (D contribution, BE xyz), in node order
    int32_t *inc_ptr = nullptr;   // node i's contributions: elem_out[inc_ptr[i] .. inc_ptr[i+1])
    int32_t *epos = nullptr;      // element output entry (a Tm + t, 4 Tm + a Fk + f) -> its slot

    // surface anisotropy triangles
    int4 *trik_nodes = nullptr;
    double *trik_geo = nullptr;  // SoA [9][Fk]: w0..3, Kbis, uk xyz, pad

    // charged triangles
    int4 *tri_nodes = nullptr;
    double *tri_geo = nullptr;   // SoA [16][Fm]: n xyz, w0..3, dMs, surf, pad x3, Ks, uk xyz
    double2 *tri_corr = nullptr; // [3][Fm] vertex corrections (u, v)
    int32_t *cinc_ptr = nullptr, *cinc_idx = nullptr;  // node -> tri_corr entries
    double2 *corr = nullptr;     // [N]

    // demag sources
    double4 *src_pos = nullptr;  // [S] in tree order when the FMM is used
    int64_t *src_rank = nullptr; // [S] source index (5 per magnetic tet, 4 per charged tri) -> slot
    double2 *src_q = nullptr;    // [S] (q from u, q from v)
    float4 *src_pos_f = nullptr; // FP32 copies for the mixed-precision P2P
    double2 *qe = nullptr;       // [Tm + 3 Fm] element charge values for the FMM near-field operator:
                                 // -Ms div u per tetrahedron, u_a . n per triangle vertex (u, v)
    FmmPlan *fmm = nullptr;
    GuessNet *guess = nullptr;   // learned initial guess (allocated on first use)

    // energies / reductions
    double *epart = nullptr;  // [8][MAX_PART]
    double *res = nullptr;    // [32] small results: energies 0..5, min dot 8, average 16..18

    // profiling
    std::vector<KernelStat> stats;
    struct Pending {
        int id;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;

    int blocks_for(int64_t n) const
    {
        int64_t b = (n + BLOCK - 1) / BLOCK;
        return (int)(b < MAX_PART ? (b > 0 ? b : 1) : MAX_PART);
    }
    int kid(const char *name);
    void tic(int id, hipEvent_t *a);
    void toc(int id, hipEvent_t a);
    void flush_stats();
};

// launch-time profiling scope: records HIP events on the context stream around a kernel
struct Prof {
    Ctx &c;
    int id;
    hipEvent_t a = nullptr;
    Prof(Ctx &ctx, const char *name) : c(ctx), id(ctx.opt.profile ? ctx.kid(name) : -1)
    {
        if (id >= 0) c.tic(id, &a);
    }
    ~Prof()
    {
        if (id >= 0) c.toc(id, a);
    }
};

// ---- launchers (kernels.hip) ----
// hsp: optional applied field per magnetic tet and Gauss point ([15][Tm], A/m), scaled by amp
// and used in place of Hext
void launch_element(Ctx &c, const double Hext[3], double dt, const double *hsp = nullptr, double amp = 0);
void launch_basis(Ctx &c, double r);
void launch_node_assemble(Ctx &c, double r, double c_stencil, double dt);
void launch_r0(Ctx &c, double c_stencil);
// one BiCGSTAB iteration (part 0), its first phase alone (1: the convergence test of the previous
// iteration) or the other four (2)
void launch_iteration(Ctx &c, double c_stencil, int part = 0);
// mixed-precision Krylov: FP32 copies and count reset; FP64 residual and outer test; one FP32
// iteration; x += d
void launch_mixed_start(Ctx &c);
void launch_mixed_residual(Ctx &c, double c_stencil, double tol_in, int max_inner);
void launch_inner_iteration(Ctx &c, double c_stencil);
void launch_refine(Ctx &c);
// FG_BATCH_PERSISTENT: every instance's bicg_dir in one launch (one block per instance)
void launch_batch_solve(Ctx &c, double c_stencil, int block);
// gated: only if the control block says the solve was accepted (launched with the last
// iterations, before the host has read it)
void launch_update(Ctx &c, double dt, bool gated = false);
void launch_spmv_debug(Ctx &c, double c_stencil, const double2 *xin, double2 *yout);
void launch_charges(Ctx &c);
void launch_direct_p2p(Ctx &c);
void launch_energy(Ctx &c, const double Hext[3], const double *hsp = nullptr, double amp = 0);
// c.hst from the node spin accumulation c.spin_s and the per-region prefactors c.spin_pref
void launch_spin_field(Ctx &c);
void launch_max_angle(Ctx &c);
void launch_average(Ctx &c, int which);
void launch_evolution(Ctx &c);

// ---- learned initial guess (guess.hip) ----
void guess_init(Ctx &c, uint64_t seed);
void guess_set(Ctx &c, const float *params);
void guess_get(Ctx &c, float *params);
void guess_destroy(Ctx &c);
void guess_on_assembly(Ctx &c, double dt);
void guess_restart(Ctx &c);
void guess_on_solved(Ctx &c);
void guess_train(Ctx &c, float lr, double *loss, double *loss_ref);
void guess_gradient(Ctx &c, float *grad, double *loss, double *loss_ref);
void guess_batch(Ctx &c, float *feat);
bool guess_enable(Ctx &c, int on);

// ---- FMM (fmm.hip) ----
int fmm_create(Ctx &c);
void fmm_destroy(Ctx &c);
void launch_fmm(Ctx &c);
void fmm_info(const Ctx &c, int64_t info[16]);

// ---- persistent spin-diffusion solver (transport.hip) ----
// its volume regions, node count and device
const fg_spin_volume *spin_volumes(const fg_spin *sp, int *n_regions, int *n_nodes, int *device);
// one solve from the node states cur (s-d exchange) and nxt (source) on stream st; s, zero unless
// converged, copied to s_dev[3 N] on the stream
int spin_solve_nodes(fg_spin *sp, const NodeRec *cur, const NodeRec *nxt, hipStream_t st, int accumulate,
                     double *s_dev, fg_solve_info *info);

}  // namespace fg
