// kernels.hip -- element pass, node assembly, Krylov, update, demag sources, energies.
//
// Data layout (all FP64):
//   node state     NodeRec[N] x 2 (CURRENT / NEXT; evolution copies NEXT)      64 B per node
//   basis          ep4[N], eq4[N] (double4)
//   tet geometry   SoA geo[17][Tm] (da 4x3 row-major, w[5]); nodes int4[Tm]
//   element output elem_out[4 Tm (+3 Fk)] double4 = (alpha_eff-weighted mass, BE x, y, z), stored in
//                  node order (each node's contributions contiguous, in mesh order of their
//                  elements): the element pass scatters its writes so the node pass reads runs
//   stiffness S'   SELL-64 (col int32, val f64), S'_ij = sum_t Abis_t sum_g w_g (da_a . da_b)
//   Krylov vectors double2[N] (the two tangent unknowns of a node side by side)
//
// The LLG matrix of FeeLLGood (src/tetra.cpp:210-307 assembled by src/solver.h:319-339) factors as
//   K[(i,di),(j,dj)] = c S'_ij (e_r(di),i . e_s(dj),j) + delta_ij (D_i B_ii + m_i X_i)[di][dj]
// with c = prefactor * theta * dt * gamma0, rows r = (eq, ep), columns s = (ep, eq), B_ij the
// basis dot products, X_i the projection of [u_i]x and D_i the alpha_eff-weighted lumped mass.
// The SpMV therefore lifts x to a 3-D field w = x_p ep + x_q eq, applies the scalar stencil S' to
// it and projects back, reading 12 bytes per stored entry instead of a 2x2 FP64 block.
#include "ctx.h"

#include <cmath>

namespace fg {

// ------------------------------------------------------------------ reductions
__device__ inline double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wave_min(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

// This lane's share (p[lane], p[lane + 64], ...) of nb partials, combined by OP (a deterministic
// order): eight independent accumulators keep eight loads in flight instead of one after another.
template <class OP>
__device__ inline double lane_fold(const double *__restrict__ p, int nb, double init, OP op)
{
    double a0 = init, a1 = init, a2 = init, a3 = init, a4 = init, a5 = init, a6 = init, a7 = init;
    int i = threadIdx.x & 63;
    for (; i + 7 * 64 < nb; i += 8 * 64) {
        a0 = op(a0, p[i]);
        a1 = op(a1, p[i + 64]);
        a2 = op(a2, p[i + 128]);
        a3 = op(a3, p[i + 192]);
        a4 = op(a4, p[i + 256]);
        a5 = op(a5, p[i + 320]);
        a6 = op(a6, p[i + 384]);
        a7 = op(a7, p[i + 448]);
    }
    for (; i < nb; i += 64) a0 = op(a0, p[i]);
    return op(op(op(a0, a1), op(a2, a3)), op(op(a4, a5), op(a6, a7)));
}
struct AddOp {
    __device__ double operator()(double a, double b) const { return a + b; }
};
struct MaxOp {
    __device__ double operator()(double a, double b) const { return fmax(a, b); }
};
struct MinOp {
    __device__ double operator()(double a, double b) const { return fmin(a, b); }
};

// Sum of NV per-thread values over the block, stored to part[slot[k] * MAX_PART + blk], blk being
// the block index the kernel assigned its items by (blockIdx.x or lblock()).
template <int NV>
__device__ inline void block_store(const double (&v)[NV], const int (&slot)[NV], double *part, int blk)
{
    __shared__ double sh[NV][BLOCK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double r[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) r[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) sh[k][wid] = r[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = sh[k][0];
            for (int q = 1; q < BLOCK / 64; q++) s += sh[k][q];
            part[slot[k] * MAX_PART + blk] = s;
        }
}

// Deterministic sum of nb partials of a slot; every thread of the block gets the result.
template <int NV>
__device__ inline void block_load(const double *part, const int (&slot)[NV], int nb, double (&out)[NV])
{
    __shared__ double res[NV];
    if (threadIdx.x < 64) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const double s = wave_sum(lane_fold(part + slot[k] * MAX_PART, nb, 0.0, AddOp()));
            if (threadIdx.x == 0) res[k] = s;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) out[k] = res[k];
}

// ------------------------------------------------------------------ element pass
// Tetra::calc_alpha_eff (src/tetra.cpp:47-75)
__device__ inline double alpha_eff(double dt, double alpha, double h)
{
    const double reduced_dt = GAMMA0 * dt;
    const double M = 2. * alpha * 0.1 / reduced_dt;
    if (h > 0.) return alpha + reduced_dt / 2. * (h > M ? M : h);
    if (h < -M) return alpha / (1. + reduced_dt / (2. * alpha) * M);
    return alpha / (1. - reduced_dt / (2. * alpha) * h);
}

// One thread per magnetic tetrahedron: the basis-free part of Tet::integrales
// (src/tetra.cpp:210-307): for every local node a, the alpha_eff-weighted lumped mass
// sum_g a_ag w_g alpha_eff,g and the 3-D right-hand side BE_a.
__global__ __launch_bounds__(BLOCK) void k_element(int Tm, const int4 *__restrict__ tn,
                                                   const double *__restrict__ geo,
                                                   const int32_t *__restrict__ treg,
                                                   const DevRegion *__restrict__ regs,
                                                   const NodeRec *__restrict__ cur, double Hx,
                                                   double Hy, double Hz, const double *__restrict__ hsp,
                                                   double amp, const double *__restrict__ hst, double dt,
                                                   const int32_t *__restrict__ epos, double4 *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Tm) return;
    const int4 q = tn[t];
    const int nd[4] = {q.x, q.y, q.z, q.w};
    double da[4][3], w[5];
#pragma unroll
    for (int c = 0; c < 12; c++) da[c / 3][c % 3] = geo[(size_t)c * Tm + t];
#pragma unroll
    for (int g = 0; g < 5; g++) w[g] = geo[(size_t)(12 + g) * Tm + t];
    const DevRegion R = regs[treg[t]];
    d3 u[4], v[4];
    double ph[4], phv[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const NodeRec n = cur[nd[a]];
        u[a] = mk(n.ux, n.uy, n.uz);
        v[a] = mk(n.vx, n.vy, n.vz);
        ph[a] = n.phi;
        phv[a] = n.phiv;
    }
    d3 dU[3];
#pragma unroll
    for (int k = 0; k < 3; k++) dU[k] = da[0][k] * u[0] + da[1][k] * u[1] + da[2][k] * u[2] + da[3][k] * u[3];
    d3 Hd = mk(0, 0, 0), Hv = mk(0, 0, 0);
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const d3 g = mk(da[a][0], da[a][1], da[a][2]);
        Hd = Hd - ph[a] * g;
        Hv = Hv - phv[a] * g;
    }
    const double exch = -R.Abis * (dot(dU[0], dU[0]) + dot(dU[1], dU[1]) + dot(dU[2], dU[2]));
    // applied field: uniform, or amp times the space field at each Gauss point (hsp, [15][Tm])
    const d3 Hu = Hd + mk(Hx, Hy, Hz);
    const double sdt = THETA * dt;  // s_dt / gamma0
    const d3 uk = mk(R.uk[0], R.uk[1], R.uk[2]);
    const d3 ex = mk(R.ex[0], R.ex[1], R.ex[2]), ey = mk(R.ey[0], R.ey[1], R.ey[2]),
             ez = mk(R.ez[0], R.ez[1], R.ez[2]);
    double contrib[4] = {0, 0, 0, 0};
    d3 BE[4];
    d3 gex[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
        BE[a] = mk(0, 0, 0);
        gex[a] = da[a][0] * dU[0] + da[a][1] * dU[1] + da[a][2] * dU[2];
    }
#pragma unroll
    for (int g = 0; g < 5; g++) {
        d3 U = mk(0, 0, 0), V = mk(0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; a++) {
            U = U + tet_a(a, g) * u[a];
            V = V + tet_a(a, g) * v[a];
        }
        const d3 H0 = hsp ? Hd + amp * mk(hsp[(size_t)(3 * g) * Tm + t], hsp[(size_t)(3 * g + 1) * Tm + t],
                                          hsp[(size_t)(3 * g + 2) * Tm + t])
                          : Hu;
        double uH = exch;
        d3 Han = mk(0, 0, 0);
        if (R.hasK) {  // Tet::calc_aniso_uniax (src/tetra.cpp:171-181)
            Han = Han + (R.Kbis * dot(uk, U + sdt * V)) * uk;
            const double c = dot(U, uk);
            uH += R.Kbis * (c * c);
        }
        if (R.hasK3) {  // Tet::calc_aniso_cub (src/tetra.cpp:183-208)
            const d3 bu = mk(dot(ex, U), dot(ey, U), dot(ez, U));
            const d3 bv = mk(dot(ex, V), dot(ey, V), dot(ez, V));
            const d3 b3 = mk(bu.x * (1.0 - bu.x * bu.x), bu.y * (1.0 - bu.y * bu.y), bu.z * (1.0 - bu.z * bu.z));
            const d3 tm = mk(bv.x * ex.x * (1 - 3 * bu.x * bu.x), bv.y * ex.y * (1 - 3 * bu.y * bu.y),
                             bv.z * ex.z * (1 - 3 * bu.z * bu.z));
            Han = Han + (-R.K3bis) * (b3.x * ex + b3.y * ey + b3.z * ez + sdt * tm);
            uH += -R.K3bis * dot(bu, b3);
        }
        // the spin-accumulation field (Tet::integrales, src/tetra.cpp:248-255,296-302): in the
        // effective field of alpha_eff, in the right-hand side, and its second-order term
        // -(Hst . U) s_dt V; zero without one
        const d3 Hs = hst ? mk(hst[(size_t)(3 * g) * Tm + t], hst[(size_t)(3 * g + 1) * Tm + t],
                               hst[(size_t)(3 * g + 2) * Tm + t])
                          : mk(0.0, 0.0, 0.0);
        uH += dot(U, H0 + Hs);
        const double ae = alpha_eff(dt, R.alpha, uH);
        const d3 H = H0 + (Han + sdt * Hv);
        const double wg = w[g];
        const double hsu = dot(Hs, U);
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const double aw = wg * tet_a(a, g);
            contrib[a] += tet_a(a, g) * (wg * ae);
            BE[a] = BE[a] - (wg * R.Abis) * gex[a];
            BE[a] = BE[a] + aw * (H + Hs);
            if (hst) BE[a] = BE[a] - ((aw * hsu) * (sdt * GAMMA0)) * V;
        }
    }
#pragma unroll
    for (int a = 0; a < 4; a++) out[epos[(size_t)a * Tm + t]] = make_double4(contrib[a], BE[a].x, BE[a].y, BE[a].z);
}

// Tri::integrales (src/triangle.cpp:6-36): surface anisotropy right-hand side, 3-D form.
__global__ __launch_bounds__(BLOCK) void k_tri_aniso(int Fk, const int4 *__restrict__ tn,
                                                     const double *__restrict__ tg,
                                                     const NodeRec *__restrict__ cur,
                                                     const int32_t *__restrict__ epos, double4 *__restrict__ out)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= Fk) return;
    const int4 q = tn[f];
    const int nd[3] = {q.x, q.y, q.z};
    d3 u[3];
    for (int a = 0; a < 3; a++) {
        const NodeRec n = cur[nd[a]];
        u[a] = mk(n.ux, n.uy, n.uz);
    }
    const double Kbis = tg[(size_t)4 * Fk + f];
    const d3 uk = mk(tg[(size_t)5 * Fk + f], tg[(size_t)6 * Fk + f], tg[(size_t)7 * Fk + f]);
    d3 BE[3] = {mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)};
    for (int g = 0; g < 4; g++) {
        const d3 U = tri_a(0, g) * u[0] + tri_a(1, g) * u[1] + tri_a(2, g) * u[2];
        const double pre = tg[(size_t)g * Fk + f] * Kbis * dot(uk, U);
        for (int a = 0; a < 3; a++) BE[a] = BE[a] + (pre * tri_a(a, g)) * uk;
    }
    for (int a = 0; a < 3; a++) out[epos[(size_t)a * Fk + f]] = make_double4(0.0, BE[a].x, BE[a].y, BE[a].z);
}

// Node::setBasis (src/node.h:73-102)
__device__ inline void set_basis(d3 u, double cr, double sr, d3 &ep, d3 &eq)
{
    const double ax = fabs(u.x), ay = fabs(u.y), az = fabs(u.z);
    int k = 0;
    if (ay < ax) k = 1;
    if (az < (k == 0 ? ax : ay)) k = 2;
    d3 e = mk(k == 0, k == 1, k == 2);
    e = e - dot(e, u) * u;
    const double n2 = dot(e, e);
    if (n2 > 0) {
        const double n = sqrt(n2);
        e = mk(e.x / n, e.y / n, e.z / n);
    }
    const d3 qv = cross(u, e);
    ep = cr * e - sr * qv;
    eq = sr * e + cr * qv;
}

__global__ __launch_bounds__(BLOCK) void k_basis(int N, double cr, double sr, const NodeRec *__restrict__ cur,
                                                 double4 *__restrict__ ep4, double4 *__restrict__ eq4)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const NodeRec n = cur[i];
        d3 ep, eq;
        set_basis(mk(n.ux, n.uy, n.uz), cr, sr, ep, eq);
        ep4[i] = make_double4(ep.x, ep.y, ep.z, 0.0);
        eq4[i] = make_double4(eq.x, eq.y, eq.z, 0.0);
    }
}

// One thread per node: basis, gathered element contributions, diagonal block, Jacobi
// preconditioner, right-hand side and initial guess (LinAlgebra::solve assembly part,
// src/solver.cpp:11-59; buildInitGuess src/linear_algebra.cpp:13-24).
__global__ __launch_bounds__(BLOCK) void k_node_assemble(
    int N, double cr, double sr, double cst, const NodeRec *__restrict__ cur,
    const NodeRec *__restrict__ vsrc, const int32_t *__restrict__ mag,
    const int32_t *__restrict__ inc_ptr,
    const double4 *__restrict__ eout, const double *__restrict__ mass,
    const double *__restrict__ sdiag, double4 *__restrict__ ep4, double4 *__restrict__ eq4,
    double4 *__restrict__ Dg, double2 *__restrict__ dprec, double2 *__restrict__ b,
    double2 *__restrict__ x, double4 *__restrict__ w, double *__restrict__ part)
{
    double bb = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const NodeRec n = cur[i];
        const d3 u = mk(n.ux, n.uy, n.uz);
        d3 ep, eq;
        set_basis(u, cr, sr, ep, eq);
        ep4[i] = make_double4(ep.x, ep.y, ep.z, 0.0);
        eq4[i] = make_double4(eq.x, eq.y, eq.z, 0.0);
        if (!mag[i]) {
            Dg[i] = make_double4(0, 0, 0, 0);
            dprec[i] = make_double2(0, 0);
            b[i] = make_double2(0, 0);
            x[i] = make_double2(0, 0);
            w[i] = make_double4(0, 0, 0, 0);
            continue;
        }
        double D = 0;
        d3 F = mk(0, 0, 0);
        for (int k = inc_ptr[i]; k < inc_ptr[i + 1]; k++) {
            const double4 e = eout[k];
            D += e.x;
            F = F + mk(e.y, e.z, e.w);
        }
        const double m = mass[i];
        // B = [[eq.ep, eq.eq], [ep.ep, ep.eq]];  X = projection of [u]x
        const double B00 = dot(eq, ep), B01 = dot(eq, eq), B10 = dot(ep, ep), B11 = dot(ep, eq);
        const d3 uep = cross(u, ep), ueq = cross(u, eq);
        const double X00 = dot(eq, uep), X01 = dot(eq, ueq), X10 = dot(ep, uep), X11 = dot(ep, ueq);
        const double4 dg = make_double4(D * B00 + m * X00, D * B01 + m * X01, D * B10 + m * X10, D * B11 + m * X11);
        Dg[i] = dg;
        const double sd = cst * sdiag[i];
        dprec[i] = make_double2(1.0 / (sd * B00 + dg.x), 1.0 / (sd * B11 + dg.w));
        const double2 bi = make_double2(dot(eq, F), dot(ep, F));
        b[i] = bi;
        bb += bi.x * bi.x + bi.y * bi.y;
        const NodeRec nv = vsrc[i];
        const d3 vv = mk(nv.vx, nv.vy, nv.vz);
        const double2 xi = make_double2(dot(vv, ep) / GAMMA0, dot(vv, eq) / GAMMA0);
        x[i] = xi;
        const d3 wi = xi.x * ep + xi.y * eq;
        w[i] = make_double4(wi.x, wi.y, wi.z, 0.0);
    }
    const double vals[1] = {bb};
    const int slots[1] = {0};
    block_store<1>(vals, slots, part, blockIdx.x);
}

// ------------------------------------------------------------------ stencil SpMV
// y_i = c [eq_i . z_i, ep_i . z_i] + Dg_i x_i with z_i = sum_j S'_ij w_j, zero for
// non-magnetic rows (applyMask after mult, src/algebra/bicg.h:201,216).
__device__ inline double2 stencil_row(int i, double cst, const int64_t *__restrict__ sptr,
                                      const int32_t *__restrict__ swid, const int32_t *__restrict__ scol,
                                      const double *__restrict__ sval, const double4 *__restrict__ w,
                                      const double4 *__restrict__ ep4, const double4 *__restrict__ eq4,
                                      const double4 *__restrict__ Dg, double2 xi)
{
    const int sl = i >> 6, lane = i & 63;
    const int64_t off = sptr[sl];
    const int width = swid[sl];
    double zx = 0, zy = 0, zz = 0;
    for (int k = 0; k < width; k++) {
        const int64_t e = off + (int64_t)k * SELL_C + lane;
        const int j = scol[e];
        const double sv = sval[e];
        const double4 wj = w[j];
        zx += sv * wj.x;
        zy += sv * wj.y;
        zz += sv * wj.z;
    }
    const double4 p = ep4[i], q = eq4[i], g = Dg[i];
    double2 y;
    y.x = cst * (q.x * zx + q.y * zy + q.z * zz) + (g.x * xi.x + g.y * xi.y);
    y.y = cst * (p.x * zx + p.y * zy + p.z * zz) + (g.z * xi.x + g.w * xi.y);
    return y;
}

#define STENCIL_ARGS                                                                          \
    const int64_t *__restrict__ sptr, const int32_t *__restrict__ swid,                       \
        const int32_t *__restrict__ scol, const double *__restrict__ sval,                    \
        const double4 *__restrict__ ep4, const double4 *__restrict__ eq4,                     \
        const double4 *__restrict__ Dg, const int32_t *__restrict__ mag
#define STENCIL_PASS c.slice_ptr, c.slice_w, c.scol, c.sval, c.ep4, c.eq4, c.Dg, c.mag

// r = b - A x (masked); rt = r; p = r; partials ||r||^2, rt.r  (src/algebra/bicg.h:172-184)
__global__ __launch_bounds__(BLOCK) void k_r0(int N, double cst, STENCIL_ARGS, const double4 *__restrict__ w,
                                              const double2 *__restrict__ x, const double2 *__restrict__ b,
                                              double2 *__restrict__ r, double2 *__restrict__ rt,
                                              double2 *__restrict__ p, KrylovCtl *ctl, double *__restrict__ part)
{
    double rr = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        double2 ri = make_double2(0, 0);
        if (mag[i]) {
            const double2 y = stencil_row(i, cst, sptr, swid, scol, sval, w, ep4, eq4, Dg, x[i]);
            const double2 bi = b[i];
            ri = make_double2(bi.x - y.x, bi.y - y.y);
        }
        r[i] = ri;
        rt[i] = ri;
        p[i] = ri;
        rr += ri.x * ri.x + ri.y * ri.y;
    }
    // rt = r, so rt . r = ||r||^2 for the first iteration
    const double vals[2] = {rr, rr};
    const int slots[2] = {1, 2};
    block_store<2>(vals, slots, part, blockIdx.x);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->nit = 0;
        ctl->status = -1;
        ctl->done = ctl->stop = 0;
        ctl->rho1 = ctl->rho2 = ctl->alpha = ctl->omega = ctl->beta = 0.0;
    }
}

// ||b|| into the control block (one block)
__global__ void k_set_rhsn(int nb, KrylovCtl *ctl, const double *__restrict__ part)
{
    const int slots[1] = {0};
    double v[1];
    block_load<1>(part, slots, nb, v);
    if (threadIdx.x == 0) ctl->rhsn = sqrt(fabs(v[0]));
}

// Batched instances: ||b_k||^2 of each instance (one block per instance), then the stopping scale
// min over instances with b_k != 0, so that ||r|| <= tol min_k ||b_k|| gives every instance its own
// criterion ||r_k|| <= tol ||b_k||.
__global__ __launch_bounds__(BLOCK) void k_inst_norm2(const int64_t *__restrict__ off, const double2 *__restrict__ b,
                                                     double *__restrict__ out)
{
    __shared__ double red[BLOCK];
    double s = 0;
    for (int64_t i = off[blockIdx.x] + threadIdx.x; i < off[blockIdx.x + 1]; i += blockDim.x) {
        const double2 v = b[i];
        s += v.x * v.x + v.y * v.y;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = BLOCK / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ void k_inst_rhsn(int K, const double *__restrict__ n2, KrylovCtl *ctl)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double mn = -1;
    for (int k = 0; k < K; k++)
        if (n2[k] > 0 && (mn < 0 || n2[k] < mn)) mn = n2[k];
    if (mn > 0) ctl->rhsn = sqrt(mn);
}

// (C) loop head: convergence on ||r||, rho_1 = rt.r, beta, p = r + beta (p - omega v),
// p^ = D p, w = lift(p^)  (src/algebra/bicg.h:180-199)
__global__ __launch_bounds__(BLOCK) void k_iter_p(int N, int nb, KrylovCtl *ctl, const double4 *__restrict__ ep4,
                                                  const double4 *__restrict__ eq4,
                                                  const double2 *__restrict__ dprec,
                                                  const double2 *__restrict__ r, const double2 *__restrict__ v,
                                                  double2 *__restrict__ p, double2 *__restrict__ ph,
                                                  double4 *__restrict__ w, const double *__restrict__ part)
{
    if (ctl->done) return;
    const int slots[2] = {1, 2};
    double red[2];
    block_load<2>(part, slots, nb, red);
    const double res = sqrt(fabs(red[0]));
    const double rho1 = red[1];
    const int nit = ctl->nit;
    int status = ctl->status, done = 0;
    if (isnan(res)) {
        status = 2;
        done = 1;
    } else if (res <= ctl->rhsn * ctl->tol) {
        status = 0;
        done = 1;
    }
    double beta = 0;
    const double omega = ctl->omega;
    if (!done && nit > 0) {
        if (ctl->rho2 == 0 || omega == 0) {
            status = 2;
            done = 1;
        } else {
            beta = (rho1 / ctl->rho2) * (ctl->alpha / omega);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->res = res;
        ctl->status = status;
        ctl->rho1 = rho1;
        ctl->beta = beta;
        if (done) ctl->done = 1;
    }
    if (done) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        double2 pi;
        const double2 ri = r[i];
        if (nit > 0) {
            const double2 vi = v[i];
            pi = p[i];
            pi.x = (pi.x - omega * vi.x) * beta + ri.x;
            pi.y = (pi.y - omega * vi.y) * beta + ri.y;
        } else {
            pi = ri;
        }
        p[i] = pi;
        const double2 d = dprec[i];
        const double2 h = make_double2(d.x * pi.x, d.y * pi.y);
        ph[i] = h;
        const double4 e = ep4[i], q = eq4[i];
        w[i] = make_double4(h.x * e.x + h.y * q.x, h.x * e.y + h.y * q.y, h.x * e.z + h.y * q.z, 0.0);
    }
}

// (D) v = A p^ (masked); partial v . rt
__global__ __launch_bounds__(BLOCK) void k_spmv_v(int N, double cst, STENCIL_ARGS, const KrylovCtl *ctl,
                                                  const double4 *__restrict__ w, const double2 *__restrict__ ph,
                                                  const double2 *__restrict__ rt, double2 *__restrict__ v,
                                                  double *__restrict__ part)
{
    if (ctl->done) return;
    double vr = 0;
    for (int i = lblock() * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        double2 y = make_double2(0, 0);
        if (mag[i]) y = stencil_row(i, cst, sptr, swid, scol, sval, w, ep4, eq4, Dg, ph[i]);
        v[i] = y;
        const double2 q = rt[i];
        vr += y.x * q.x + y.y * q.y;
    }
    const double vals[1] = {vr};
    const int slots[1] = {3};
    block_store<1>(vals, slots, part, lblock());
}

// (E) alpha = rho_1 / (v . rt); s = r - alpha v; s^ = D s; w = lift(s^); partial ||s||^2
__global__ __launch_bounds__(BLOCK) void k_iter_s(int N, int nb, KrylovCtl *ctl, const double4 *__restrict__ ep4,
                                                  const double4 *__restrict__ eq4,
                                                  const double2 *__restrict__ dprec,
                                                  const double2 *__restrict__ r, const double2 *__restrict__ v,
                                                  double2 *__restrict__ s, double2 *__restrict__ sh,
                                                  double4 *__restrict__ w, double *__restrict__ part)
{
    if (ctl->done) return;
    const int slots[1] = {3};
    double red[1];
    block_load<1>(part, slots, nb, red);
    const double alpha = ctl->rho1 / red[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->alpha = alpha;
    double ss = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const double2 ri = r[i], vi = v[i];
        const double2 si = make_double2(ri.x + (-alpha) * vi.x, ri.y + (-alpha) * vi.y);
        s[i] = si;
        ss += si.x * si.x + si.y * si.y;
        const double2 d = dprec[i];
        const double2 h = make_double2(d.x * si.x, d.y * si.y);
        sh[i] = h;
        const double4 e = ep4[i], q = eq4[i];
        w[i] = make_double4(h.x * e.x + h.y * q.x, h.x * e.y + h.y * q.y, h.x * e.z + h.y * q.z, 0.0);
    }
    const double vals[1] = {ss};
    const int slots2[1] = {4};
    block_store<1>(vals, slots2, part, blockIdx.x);
}

// (F) convergence on ||s|| (x += alpha p^ when converged), else t = A s^ (masked);
// partials t.s, t.t  (src/algebra/bicg.h:206-217)
__global__ __launch_bounds__(BLOCK) void k_spmv_t(int N, int nb, double cst, STENCIL_ARGS, KrylovCtl *ctl,
                                                  const double4 *__restrict__ w, const double2 *__restrict__ sh,
                                                  const double2 *__restrict__ s, const double2 *__restrict__ ph,
                                                  double2 *__restrict__ x, double2 *__restrict__ t,
                                                  double *__restrict__ part)
{
    if (ctl->done) return;
    const int slots[1] = {4};
    double red[1];
    block_load<1>(part, slots, nb, red);
    const double res = sqrt(fabs(red[0]));
    int status = ctl->status, done = 0;
    bool conv = false;
    if (isnan(res)) {
        status = 2;
        done = 1;
    } else if (res <= ctl->rhsn * ctl->tol) {
        status = 0;
        done = 1;
        conv = true;
    } else if (status == 1 || status == 2) {
        done = 1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && done) {
        ctl->res = res;
        ctl->status = status;
        ctl->stop = 1;
    }
    if (done) {
        if (conv) {
            const double alpha = ctl->alpha;
            for (int i = lblock() * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
                double2 xi = x[i];
                const double2 h = ph[i];
                xi.x += alpha * h.x;
                xi.y += alpha * h.y;
                x[i] = xi;
            }
        }
        return;
    }
    double ts = 0, tt = 0;
    for (int i = lblock() * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        double2 y = make_double2(0, 0);
        if (mag[i]) y = stencil_row(i, cst, sptr, swid, scol, sval, w, ep4, eq4, Dg, sh[i]);
        t[i] = y;
        const double2 si = s[i];
        ts += y.x * si.x + y.y * si.y;
        tt += y.x * y.x + y.y * y.y;
    }
    const double vals[2] = {ts, tt};
    const int slots2[2] = {5, 6};
    block_store<2>(vals, slots2, part, lblock());
}

// (G) omega = (t.s)/(t.t); x += alpha p^ + omega s^; r = s - omega t; partials ||r||^2, rt.r
__global__ __launch_bounds__(BLOCK) void k_iter_x(int N, int nb, KrylovCtl *ctl, const double2 *__restrict__ ph,
                                                  const double2 *__restrict__ sh, const double2 *__restrict__ s,
                                                  const double2 *__restrict__ t, const double2 *__restrict__ rt,
                                                  double2 *__restrict__ x, double2 *__restrict__ r,
                                                  double *__restrict__ part)
{
    if (ctl->done) return;
    if (ctl->stop) {  // stopped on ||s|| by the previous phase
        if (blockIdx.x == 0 && threadIdx.x == 0) ctl->done = 1;
        return;
    }
    const int slots[2] = {5, 6};
    double red[2];
    block_load<2>(part, slots, nb, red);
    const double omega = red[0] / red[1];
    const double alpha = ctl->alpha;
    double rr = 0, rtr = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        double2 xi = x[i];
        const double2 h = ph[i], hs = sh[i], si = s[i], ti = t[i], q = rt[i];
        xi.x += alpha * h.x;
        xi.y += alpha * h.y;
        xi.x += omega * hs.x;
        xi.y += omega * hs.y;
        x[i] = xi;
        const double2 ri = make_double2(si.x - omega * ti.x, si.y - omega * ti.y);
        r[i] = ri;
        rr += ri.x * ri.x + ri.y * ri.y;
        rtr += q.x * ri.x + q.y * ri.y;
    }
    const double vals[2] = {rr, rtr};
    const int slots2[2] = {1, 2};
    block_store<2>(vals, slots2, part, blockIdx.x);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->omega = omega;
        ctl->rho2 = ctl->rho1;
        const int nit = ctl->nit + 1;
        ctl->nit = nit;
        if (nit >= ctl->max_iter) ctl->status = 1;
    }
}

// ------------------------------------------------------------------ batched instances, one launch
// FG_BATCH_PERSISTENT: one block per instance runs that instance's whole bicg_dir
// (src/algebra/bicg.h:162-234) in a single launch for the batch, with block-level reductions (a
// fixed tree), its own criterion ||r_k|| <= tol ||b_k|| and no host read or grid-wide step between
// iterations: the stencil never crosses instances, so a block needs only its own rows and their
// lifted vectors. The phases and scalars are those of the grid kernels above. Blocks take the
// instances largest first (order[]), so the long solves start early.

// sums of NV values over the block (fixed tree), the same in every thread
template <int NV, int BS>
__device__ inline void bsum(double (&v)[NV], double (*sh)[BS / 64])
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const double r = wave_sum(v[k]);
        if (lane == 0) sh[k][wid] = r;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = sh[k][0];
        for (int q = 1; q < BS / 64; q++) s += sh[k][q];
        v[k] = s;
    }
    __syncthreads();  // sh is free again
}

template <int BS>
__global__ __launch_bounds__(BS) void k_bicg_batch(const int32_t *__restrict__ order,
                                                   const int64_t *__restrict__ off, double cst, double tol,
                                                   int max_iter, STENCIL_ARGS, const double2 *__restrict__ dprec,
                                                   const double2 *__restrict__ b, double2 *__restrict__ x,
                                                   double2 *__restrict__ r, double2 *__restrict__ rt,
                                                   double2 *__restrict__ p, double2 *__restrict__ ph,
                                                   double2 *__restrict__ v, double2 *__restrict__ s,
                                                   double2 *__restrict__ sh, double2 *__restrict__ t,
                                                   double4 *__restrict__ w, int2 *__restrict__ stat,
                                                   double2 *__restrict__ resn)
{
    __shared__ double red[2][BS / 64];
    const int k = order[blockIdx.x];
    const int n0 = (int)off[k], n1 = (int)off[k + 1];
    // r = b - A x0 (w = lift(x0) from the assembly or the guess), rt = p = r
    double a2[2] = {0, 0};
    for (int i = n0 + threadIdx.x; i < n1; i += BS) {
        double2 ri = make_double2(0, 0);
        const double2 bi = b[i];
        if (mag[i]) {
            const double2 y = stencil_row(i, cst, sptr, swid, scol, sval, w, ep4, eq4, Dg, x[i]);
            ri = make_double2(bi.x - y.x, bi.y - y.y);
        }
        r[i] = ri;
        rt[i] = ri;
        p[i] = ri;
        a2[0] += ri.x * ri.x + ri.y * ri.y;
        a2[1] += bi.x * bi.x + bi.y * bi.y;
    }
    bsum<2, BS>(a2, red);
    const double rhsn = sqrt(fabs(a2[1]));
    double res = sqrt(fabs(a2[0])), rho1 = a2[0], rho2 = 0, alpha = 0, omega = 0, beta = 0;
    int nit = 0, status = -1;
    while (true) {
        // loop head: the test on ||r||, then rho and beta (src/algebra/bicg.h:180-199)
        if (isnan(res)) {
            status = FG_ITER_CANNOT_CONVERGE;
            break;
        }
        if (res <= rhsn * tol) {
            status = FG_ITER_CONVERGED;
            break;
        }
        if (nit > 0) {
            if (rho2 == 0 || omega == 0) {
                status = FG_ITER_CANNOT_CONVERGE;
                break;
            }
            beta = (rho1 / rho2) * (alpha / omega);
        }
        for (int i = n0 + threadIdx.x; i < n1; i += BS) {
            double2 pi;
            const double2 ri = r[i];
            if (nit > 0) {
                const double2 vi = v[i];
                pi = p[i];
                pi.x = (pi.x - omega * vi.x) * beta + ri.x;
                pi.y = (pi.y - omega * vi.y) * beta + ri.y;
            } else {
                pi = ri;
            }
            p[i] = pi;
            const double2 d = dprec[i];
            const double2 h = make_double2(d.x * pi.x, d.y * pi.y);
            ph[i] = h;
            const double4 e = ep4[i], q = eq4[i];
            w[i] = make_double4(h.x * e.x + h.y * q.x, h.x * e.y + h.y * q.y, h.x * e.z + h.y * q.z, 0.0);
        }
        __syncthreads();  // the instance's w
        double a1[1] = {0};
        for (int i = n0 + threadIdx.x; i < n1; i += BS) {
            double2 y = make_double2(0, 0);
            if (mag[i]) y = stencil_row(i, cst, sptr, swid, scol, sval, w, ep4, eq4, Dg, ph[i]);
            v[i] = y;
            const double2 q = rt[i];
            a1[0] += y.x * q.x + y.y * q.y;
        }
        bsum<1, BS>(a1, red);  // every thread has read w before it is rewritten below
        alpha = rho1 / a1[0];
        double a3[1] = {0};
        for (int i = n0 + threadIdx.x; i < n1; i += BS) {
            const double2 ri = r[i], vi = v[i];
            const double2 si = make_double2(ri.x + (-alpha) * vi.x, ri.y + (-alpha) * vi.y);
            s[i] = si;
            a3[0] += si.x * si.x + si.y * si.y;
            const double2 d = dprec[i];
            const double2 h = make_double2(d.x * si.x, d.y * si.y);
            sh[i] = h;
            const double4 e = ep4[i], q = eq4[i];
            w[i] = make_double4(h.x * e.x + h.y * q.x, h.x * e.y + h.y * q.y, h.x * e.z + h.y * q.z, 0.0);
        }
        bsum<1, BS>(a3, red);  // also publishes the instance's w
        // the test on ||s|| (src/algebra/bicg.h:206-212)
        const double rs = sqrt(fabs(a3[0]));
        if (isnan(rs)) {
            status = FG_ITER_CANNOT_CONVERGE;
            res = rs;
            break;
        }
        if (rs <= rhsn * tol) {
            for (int i = n0 + threadIdx.x; i < n1; i += BS) {
                double2 xi = x[i];
                const double2 h = ph[i];
                xi.x += alpha * h.x;
                xi.y += alpha * h.y;
                x[i] = xi;
            }
            status = FG_ITER_CONVERGED;
            res = rs;
            break;
        }
        if (status == FG_ITER_OVERFLOW || status == FG_ITER_CANNOT_CONVERGE) {
            res = rs;
            break;
        }
        double a4[2] = {0, 0};
        for (int i = n0 + threadIdx.x; i < n1; i += BS) {
            double2 y = make_double2(0, 0);
            if (mag[i]) y = stencil_row(i, cst, sptr, swid, scol, sval, w, ep4, eq4, Dg, sh[i]);
            t[i] = y;
            const double2 si = s[i];
            a4[0] += y.x * si.x + y.y * si.y;
            a4[1] += y.x * y.x + y.y * y.y;
        }
        bsum<2, BS>(a4, red);
        omega = a4[0] / a4[1];
        double a5[2] = {0, 0};
        for (int i = n0 + threadIdx.x; i < n1; i += BS) {
            double2 xi = x[i];
            const double2 h = ph[i], hs = sh[i], si = s[i], ti = t[i], q = rt[i];
            xi.x += alpha * h.x;
            xi.y += alpha * h.y;
            xi.x += omega * hs.x;
            xi.y += omega * hs.y;
            x[i] = xi;
            const double2 ri = make_double2(si.x - omega * ti.x, si.y - omega * ti.y);
            r[i] = ri;
            a5[0] += ri.x * ri.x + ri.y * ri.y;
            a5[1] += q.x * ri.x + q.y * ri.y;
        }
        bsum<2, BS>(a5, red);
        res = sqrt(fabs(a5[0]));
        rho2 = rho1;
        rho1 = a5[1];
        if (++nit >= max_iter) status = FG_ITER_OVERFLOW;
    }
    if (threadIdx.x == 0) {
        stat[k] = make_int2(status, nit);
        resn[k] = make_double2(res, rhsn);
    }
}

// ------------------------------------------------------------------ mixed-precision Krylov
// FG_KRYLOV_MIXED: BiCGSTAB on A d = r in FP32 (stencil values, basis, diagonal blocks,
// preconditioner and vectors in FP32, dot products summed in FP64) inside FP64 iterative
// refinement: r = b - A x in FP64, the FP32 solve from d = 0 to a loose tolerance, x += d, until
// the FP64 residual meets tol ||b||. The FP32 loop has its own control block (ic) with the FP64
// loop's scalars and stopping rules (src/algebra/bicg.h:162-234); the outer block (ctl) keeps
// ||b||, the tolerance, the status and the counts.

__device__ inline float2 stencil_row_f(int i, float cst, const int64_t *__restrict__ sptr,
                                       const int32_t *__restrict__ swid, const int32_t *__restrict__ scol,
                                       const float *__restrict__ sval, const float4 *__restrict__ w,
                                       const float4 *__restrict__ ep4, const float4 *__restrict__ eq4,
                                       const float4 *__restrict__ Dg, float2 xi)
{
    const int sl = i >> 6, lane = i & 63;
    const int64_t off = sptr[sl];
    const int width = swid[sl];
    float zx = 0, zy = 0, zz = 0;
    for (int k = 0; k < width; k++) {
        const int64_t e = off + (int64_t)k * SELL_C + lane;
        const int j = scol[e];
        const float sv = sval[e];
        const float4 wj = w[j];
        zx += sv * wj.x;
        zy += sv * wj.y;
        zz += sv * wj.z;
    }
    const float4 p = ep4[i], q = eq4[i], g = Dg[i];
    float2 y;
    y.x = cst * (q.x * zx + q.y * zy + q.z * zz) + (g.x * xi.x + g.y * xi.y);
    y.y = cst * (p.x * zx + p.y * zy + p.z * zz) + (g.z * xi.x + g.w * xi.y);
    return y;
}

#define STENCIL_ARGS_F                                                                        \
    const int64_t *__restrict__ sptr, const int32_t *__restrict__ swid,                       \
        const int32_t *__restrict__ scol, const float *__restrict__ sval,                     \
        const float4 *__restrict__ ep4, const float4 *__restrict__ eq4,                       \
        const float4 *__restrict__ Dg, const int32_t *__restrict__ mag
#define STENCIL_PASS_F c.slice_ptr, c.slice_w, c.scol, c.svalf, c.ep4f, c.eq4f, c.Dgf, c.mag

__device__ inline float4 lift_f(float2 h, float4 e, float4 q)
{
    return make_float4(h.x * e.x + h.y * q.x, h.x * e.y + h.y * q.y, h.x * e.z + h.y * q.z, 0.f);
}

// FP32 copies of the per-step basis, diagonal blocks and preconditioner
__global__ __launch_bounds__(BLOCK) void k_to_f32(int N, const double4 *__restrict__ ep4,
                                                  const double4 *__restrict__ eq4, const double4 *__restrict__ Dg,
                                                  const double2 *__restrict__ dprec, float4 *__restrict__ ep4f,
                                                  float4 *__restrict__ eq4f, float4 *__restrict__ Dgf,
                                                  float2 *__restrict__ dprecf)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const double4 e = ep4[i], q = eq4[i], g = Dg[i];
        const double2 d = dprec[i];
        ep4f[i] = make_float4((float)e.x, (float)e.y, (float)e.z, 0.f);
        eq4f[i] = make_float4((float)q.x, (float)q.y, (float)q.z, 0.f);
        Dgf[i] = make_float4((float)g.x, (float)g.y, (float)g.z, (float)g.w);
        dprecf[i] = make_float2((float)d.x, (float)d.y);
    }
}

// the outer loop's start (||b||, or the batch's criterion, is set by the assembly): counts and
// status reset
__global__ void k_mixed_init(KrylovCtl *ctl)
{
    if (threadIdx.x == 0) {
        ctl->nit = 0;
        ctl->status = -1;
        ctl->done = ctl->stop = 0;
        ctl->refinements = 0;
        ctl->res = 0;
    }
}

// FP64 residual r = b - A x (masked, with w = lift(x)); r in FP32 as the inner right-hand side,
// rt = p = r, d = 0; partials: slot 0 the FP64 ||r||^2, slots 1 and 2 the FP32 vector's (the inner
// loop's first ||r|| and rt . r)
__global__ __launch_bounds__(BLOCK) void k_resid(int N, double cst, STENCIL_ARGS, const KrylovCtl *ctl,
                                                 const double4 *__restrict__ w, const double2 *__restrict__ x,
                                                 const double2 *__restrict__ b, float2 *__restrict__ rf,
                                                 float2 *__restrict__ rtf, float2 *__restrict__ pf,
                                                 float2 *__restrict__ df, double *__restrict__ part)
{
    if (ctl->done) return;
    double rr = 0, ff = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        double2 ri = make_double2(0, 0);
        if (mag[i]) {
            const double2 y = stencil_row(i, cst, sptr, swid, scol, sval, w, ep4, eq4, Dg, x[i]);
            const double2 bi = b[i];
            ri = make_double2(bi.x - y.x, bi.y - y.y);
        }
        rr += ri.x * ri.x + ri.y * ri.y;
        const float2 f = make_float2((float)ri.x, (float)ri.y);
        rf[i] = f;
        rtf[i] = f;
        pf[i] = f;
        df[i] = make_float2(0.f, 0.f);
        ff += (double)f.x * f.x + (double)f.y * f.y;
    }
    const double vals[3] = {rr, ff, ff};
    const int slots[3] = {0, 1, 2};
    block_store<3>(vals, slots, part, blockIdx.x);
}

// The outer test on the FP64 residual (one block): converged, NaN or out of iterations ends both
// loops; otherwise the inner loop starts from d = 0 with ||r|| as its scale and tolerance
// tol_in, or what the outer tolerance still asks if that is looser.
__global__ void k_outer(int nb, KrylovCtl *ctl, KrylovCtl *ic, double tol_in, int max_inner,
                        const double *__restrict__ part)
{
    if (ctl->done) return;
    const int slots[1] = {0};
    double v[1];
    block_load<1>(part, slots, nb, v);
    if (threadIdx.x) return;
    const double res = sqrt(fabs(v[0]));
    ctl->res = res;
    int st = -1;
    if (isnan(res)) st = FG_ITER_CANNOT_CONVERGE;
    else if (res <= ctl->rhsn * ctl->tol) st = FG_ITER_CONVERGED;
    else if (ctl->nit >= ctl->max_iter) st = FG_ITER_OVERFLOW;
    if (st >= 0) {
        ctl->status = st;
        ctl->done = 1;
        ic->done = 1;
        return;
    }
    ic->rhsn = res;
    ic->tol = fmax(tol_in, 0.5 * ctl->rhsn * ctl->tol / res);
    ic->max_iter = min(max_inner, ctl->max_iter - ctl->nit);
    ic->nit = 0;
    ic->status = -1;
    ic->done = ic->stop = 0;
    ic->rho1 = ic->rho2 = ic->alpha = ic->omega = ic->beta = 0.0;
    ctl->refinements++;
}

// x += d, w = lift(x) (FP64); the inner iterations added to the outer count
__global__ __launch_bounds__(BLOCK) void k_refine(int N, KrylovCtl *ctl, const KrylovCtl *ic,
                                                  const double4 *__restrict__ ep4, const double4 *__restrict__ eq4,
                                                  const float2 *__restrict__ df, double2 *__restrict__ x,
                                                  double4 *__restrict__ w)
{
    if (ctl->done) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        double2 xi = x[i];
        const float2 d = df[i];
        xi.x += (double)d.x;
        xi.y += (double)d.y;
        x[i] = xi;
        const double4 e = ep4[i], q = eq4[i];
        w[i] = make_double4(xi.x * e.x + xi.y * q.x, xi.x * e.y + xi.y * q.y, xi.x * e.z + xi.y * q.z, 0.0);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->nit += ic->nit;
}

// The FP32 loop: the same five phases as the FP64 one, on ic
__global__ __launch_bounds__(BLOCK) void kf_iter_p(int N, int nb, KrylovCtl *ic, const float4 *__restrict__ ep4,
                                                   const float4 *__restrict__ eq4, const float2 *__restrict__ dprec,
                                                   const float2 *__restrict__ r, const float2 *__restrict__ v,
                                                   float2 *__restrict__ p, float2 *__restrict__ ph,
                                                   float4 *__restrict__ w, const double *__restrict__ part)
{
    if (ic->done) return;
    const int slots[2] = {1, 2};
    double red[2];
    block_load<2>(part, slots, nb, red);
    const double res = sqrt(fabs(red[0]));
    const double rho1 = red[1];
    const int nit = ic->nit;
    int status = ic->status, done = 0;
    if (isnan(res)) {
        status = 2;
        done = 1;
    } else if (res <= ic->rhsn * ic->tol) {
        status = 0;
        done = 1;
    }
    double beta = 0;
    const double omega = ic->omega;
    if (!done && nit > 0) {
        if (ic->rho2 == 0 || omega == 0) {
            status = 2;
            done = 1;
        } else {
            beta = (rho1 / ic->rho2) * (ic->alpha / omega);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ic->res = res;
        ic->status = status;
        ic->rho1 = rho1;
        ic->beta = beta;
        if (done) ic->done = 1;
    }
    if (done) return;
    const float om = (float)omega, be = (float)beta;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        float2 pi;
        const float2 ri = r[i];
        if (nit > 0) {
            const float2 vi = v[i];
            pi = p[i];
            pi.x = (pi.x - om * vi.x) * be + ri.x;
            pi.y = (pi.y - om * vi.y) * be + ri.y;
        } else {
            pi = ri;
        }
        p[i] = pi;
        const float2 d = dprec[i];
        const float2 h = make_float2(d.x * pi.x, d.y * pi.y);
        ph[i] = h;
        w[i] = lift_f(h, ep4[i], eq4[i]);
    }
}

__global__ __launch_bounds__(BLOCK) void kf_spmv_v(int N, float cst, STENCIL_ARGS_F, const KrylovCtl *ic,
                                                   const float4 *__restrict__ w, const float2 *__restrict__ ph,
                                                   const float2 *__restrict__ rt, float2 *__restrict__ v,
                                                   double *__restrict__ part)
{
    if (ic->done) return;
    double vr = 0;
    for (int i = lblock() * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        float2 y = make_float2(0, 0);
        if (mag[i]) y = stencil_row_f(i, cst, sptr, swid, scol, sval, w, ep4, eq4, Dg, ph[i]);
        v[i] = y;
        const float2 q = rt[i];
        vr += (double)y.x * q.x + (double)y.y * q.y;
    }
    const double vals[1] = {vr};
    const int slots[1] = {3};
    block_store<1>(vals, slots, part, lblock());
}

__global__ __launch_bounds__(BLOCK) void kf_iter_s(int N, int nb, KrylovCtl *ic, const float4 *__restrict__ ep4,
                                                   const float4 *__restrict__ eq4, const float2 *__restrict__ dprec,
                                                   const float2 *__restrict__ r, const float2 *__restrict__ v,
                                                   float2 *__restrict__ s, float2 *__restrict__ sh,
                                                   float4 *__restrict__ w, double *__restrict__ part)
{
    if (ic->done) return;
    const int slots[1] = {3};
    double red[1];
    block_load<1>(part, slots, nb, red);
    const double alpha = ic->rho1 / red[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) ic->alpha = alpha;
    const float al = (float)alpha;
    double ss = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const float2 ri = r[i], vi = v[i];
        const float2 si = make_float2(ri.x + (-al) * vi.x, ri.y + (-al) * vi.y);
        s[i] = si;
        ss += (double)si.x * si.x + (double)si.y * si.y;
        const float2 d = dprec[i];
        const float2 h = make_float2(d.x * si.x, d.y * si.y);
        sh[i] = h;
        w[i] = lift_f(h, ep4[i], eq4[i]);
    }
    const double vals[1] = {ss};
    const int slots2[1] = {4};
    block_store<1>(vals, slots2, part, blockIdx.x);
}

__global__ __launch_bounds__(BLOCK) void kf_spmv_t(int N, int nb, float cst, STENCIL_ARGS_F, KrylovCtl *ic,
                                                   const float4 *__restrict__ w, const float2 *__restrict__ sh,
                                                   const float2 *__restrict__ s, const float2 *__restrict__ ph,
                                                   float2 *__restrict__ x, float2 *__restrict__ t,
                                                   double *__restrict__ part)
{
    if (ic->done) return;
    const int slots[1] = {4};
    double red[1];
    block_load<1>(part, slots, nb, red);
    const double res = sqrt(fabs(red[0]));
    int status = ic->status, done = 0;
    bool conv = false;
    if (isnan(res)) {
        status = 2;
        done = 1;
    } else if (res <= ic->rhsn * ic->tol) {
        status = 0;
        done = 1;
        conv = true;
    } else if (status == 1 || status == 2) {
        done = 1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && done) {
        ic->res = res;
        ic->status = status;
        ic->stop = 1;
    }
    if (done) {
        if (conv) {
            const float al = (float)ic->alpha;
            for (int i = lblock() * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
                float2 xi = x[i];
                const float2 h = ph[i];
                xi.x += al * h.x;
                xi.y += al * h.y;
                x[i] = xi;
            }
        }
        return;
    }
    double ts = 0, tt = 0;
    for (int i = lblock() * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        float2 y = make_float2(0, 0);
        if (mag[i]) y = stencil_row_f(i, cst, sptr, swid, scol, sval, w, ep4, eq4, Dg, sh[i]);
        t[i] = y;
        const float2 si = s[i];
        ts += (double)y.x * si.x + (double)y.y * si.y;
        tt += (double)y.x * y.x + (double)y.y * y.y;
    }
    const double vals[2] = {ts, tt};
    const int slots2[2] = {5, 6};
    block_store<2>(vals, slots2, part, lblock());
}

__global__ __launch_bounds__(BLOCK) void kf_iter_x(int N, int nb, KrylovCtl *ic, const float2 *__restrict__ ph,
                                                   const float2 *__restrict__ sh, const float2 *__restrict__ s,
                                                   const float2 *__restrict__ t, const float2 *__restrict__ rt,
                                                   float2 *__restrict__ x, float2 *__restrict__ r,
                                                   double *__restrict__ part)
{
    if (ic->done) return;
    if (ic->stop) {  // stopped on ||s|| by the previous phase
        if (blockIdx.x == 0 && threadIdx.x == 0) ic->done = 1;
        return;
    }
    const int slots[2] = {5, 6};
    double red[2];
    block_load<2>(part, slots, nb, red);
    const double omega = red[0] / red[1];
    const float al = (float)ic->alpha, om = (float)omega;
    double rr = 0, rtr = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        float2 xi = x[i];
        const float2 h = ph[i], hs = sh[i], si = s[i], ti = t[i], q = rt[i];
        xi.x += al * h.x;
        xi.y += al * h.y;
        xi.x += om * hs.x;
        xi.y += om * hs.y;
        x[i] = xi;
        const float2 ri = make_float2(si.x - om * ti.x, si.y - om * ti.y);
        r[i] = ri;
        rr += (double)ri.x * ri.x + (double)ri.y * ri.y;
        rtr += (double)q.x * ri.x + (double)q.y * ri.y;
    }
    const double vals[2] = {rr, rtr};
    const int slots2[2] = {1, 2};
    block_store<2>(vals, slots2, part, blockIdx.x);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ic->omega = omega;
        ic->rho2 = ic->rho1;
        const int nit = ic->nit + 1;
        ic->nit = nit;
        if (nit >= ic->max_iter) ic->status = 1;
    }
}

// debug: y = A x without masking side effects
__global__ __launch_bounds__(BLOCK) void k_spmv_plain(int N, double cst, STENCIL_ARGS, const double4 *__restrict__ w,
                                                      const double2 *__restrict__ xin, double2 *__restrict__ y)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        double2 yi = make_double2(0, 0);
        if (mag[i]) yi = stencil_row(i, cst, sptr, swid, scol, sval, w, ep4, eq4, Dg, xin[i]);
        y[i] = yi;
    }
}

__global__ __launch_bounds__(BLOCK) void k_lift(int N, const double4 *__restrict__ ep4,
                                                const double4 *__restrict__ eq4,
                                                const double2 *__restrict__ xin, double4 *__restrict__ w)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const double2 h = xin[i];
        const double4 e = ep4[i], q = eq4[i];
        w[i] = make_double4(h.x * e.x + h.y * q.x, h.x * e.y + h.y * q.y, h.x * e.z + h.y * q.z, 0.0);
    }
}

// ------------------------------------------------------------------ update
// v = gamma0 (x_p ep + x_q eq); u_next = normalize(u + dt v); partial max |x_i|^2
// (src/solver.cpp:74-88, src/node.h:116-122)
// The solve ended and is accepted: converged, no overflow or breakdown, final residual within tol
// (the host's test of the same control block)
__device__ inline bool solved(const KrylovCtl *ctl)
{
    return ctl->done && ctl->status != 1 && ctl->status != 2 && ctl->res <= ctl->tol;
}

__global__ __launch_bounds__(BLOCK) void k_update(int N, double dt, const int32_t *__restrict__ mag,
                                                  const double2 *__restrict__ x, const double4 *__restrict__ ep4,
                                                  const double4 *__restrict__ eq4,
                                                  const NodeRec *__restrict__ cur, NodeRec *__restrict__ next,
                                                  double *__restrict__ part, const KrylovCtl *gate)
{
    if (gate && !solved(gate)) return;
    double m2 = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        if (!mag[i]) continue;
        const double2 xi = x[i];
        m2 = fmax(m2, xi.x * xi.x + xi.y * xi.y);
        const double4 e = ep4[i], q = eq4[i];
        const double vp = xi.x * GAMMA0, vq = xi.y * GAMMA0;
        const d3 v = vp * mk(e.x, e.y, e.z) + vq * mk(q.x, q.y, q.z);
        const NodeRec n = cur[i];
        d3 u = mk(n.ux, n.uy, n.uz) + dt * v;
        const double n2 = dot(u, u);
        if (n2 > 0) {
            const double nn = sqrt(n2);
            u = mk(u.x / nn, u.y / nn, u.z / nn);
        }
        NodeRec o = next[i];
        o.ux = u.x;
        o.uy = u.y;
        o.uz = u.z;
        o.vx = v.x;
        o.vy = v.y;
        o.vz = v.z;
        next[i] = o;
    }
    __shared__ double sh[BLOCK / 64];
    m2 = wave_max(m2);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = sh[0];
        for (int q = 1; q < BLOCK / 64; q++) s = fmax(s, sh[q]);
        part[7 * MAX_PART + blockIdx.x] = s;
    }
}

__global__ void k_finish_update(int nb, KrylovCtl *ctl, const double *__restrict__ part, bool gated)
{
    if (gated && !solved(ctl)) return;
    if (threadIdx.x < 64) {
        const double m = wave_max(lane_fold(part + 7 * MAX_PART, nb, 0.0, MaxOp()));
        if (threadIdx.x == 0) ctl->vmax2 = m;
    }
}

// ------------------------------------------------------------------ demag sources
// Tet::charges (src/tetra.cpp:347-359): q_g = -Ms w_g div(u), for u and v NEXT at once.
__global__ __launch_bounds__(BLOCK) void k_charges_tet(int Tm, const int4 *__restrict__ tn,
                                                       const double *__restrict__ geo,
                                                       const int32_t *__restrict__ treg,
                                                       const DevRegion *__restrict__ regs,
                                                       const NodeRec *__restrict__ nxt,
                                                       const int64_t *__restrict__ rank,
                                                       double2 *__restrict__ q, double2 *__restrict__ qe)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Tm) return;
    const int4 qn = tn[t];
    const int nd[4] = {qn.x, qn.y, qn.z, qn.w};
    double du = 0, dv = 0, duy = 0, dvy = 0, duz = 0, dvz = 0;
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const NodeRec n = nxt[nd[a]];
        const double gx = geo[(size_t)(3 * a) * Tm + t], gy = geo[(size_t)(3 * a + 1) * Tm + t],
                     gz = geo[(size_t)(3 * a + 2) * Tm + t];
        du += n.ux * gx;
        dv += n.vx * gx;
        duy += n.uy * gy;
        dvy += n.vy * gy;
        duz += n.uz * gz;
        dvz += n.vz * gz;
    }
    du = du + duy + duz;
    dv = dv + dvy + dvz;
    const double Ms = regs[treg[t]].Ms;
    if (q)  // per-source charges for the direct sum (the FMM takes the element values)
#pragma unroll
        for (int g = 0; g < 5; g++) {
            const double wg = geo[(size_t)(12 + g) * Tm + t];
            q[rank[(size_t)5 * t + g]] = make_double2(-Ms * wg * du, -Ms * wg * dv);
        }
    if (qe) qe[t] = make_double2(-Ms * du, -Ms * dv);
}

// Tri::potential (src/triangle.cpp:87-125): analytic potential at vertex i of the linear surface
// charge density dMs * sum_j s_j lambda_j
__device__ inline double tri_potential(const d3 P[3], const double s[3], double surf, double dMs, int i)
{
    const int ii = (i + 1) % 3, iii = (i + 2) % 3;
    const d3 p12 = P[ii] - P[i], p13 = P[iii] - P[i];
    const double b = sqrt(dot(p12, p12));
    const double t = dot(p12, p13) / b;
    const double s2 = 2. * surf;
    const double h = s2 / b;
    const double c = (t - b) / h;
    const double fth = sqrt(1.0 + (t / h) * (t / h)), fc = sqrt(1.0 + c * c);
    const double r = h * fth;
    const double lg = log((c * t + h + fc * r) / (b * (c + fc)));
    const double xi = b * lg / fc;
    const double pot = xi * s[i] + ((xi * (h + c * t) - b * (r - b)) * s[ii] + b * (r - b - c * xi) * s[iii]) * b /
                                       (s2 * (1 + c * c));
    return 0.5 * dMs * pot;
}

// Tri::charges + Tri::correctionCharges (src/triangle.cpp:45-78) for u and v NEXT at once.
__global__ __launch_bounds__(BLOCK) void k_charges_tri(int Fm, int64_t src_off, const int4 *__restrict__ tn,
                                                       const double *__restrict__ tg,
                                                       const double4 *__restrict__ pos4,
                                                       const NodeRec *__restrict__ nxt,
                                                       const int64_t *__restrict__ rank,
                                                       double2 *__restrict__ q, double2 *__restrict__ corr3,
                                                       double2 *__restrict__ qe, int Tm)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= Fm) return;
    const int4 qn = tn[f];
    const int nd[3] = {qn.x, qn.y, qn.z};
    const d3 n = mk(tg[f], tg[(size_t)Fm + f], tg[(size_t)2 * Fm + f]);
    const double dMs = tg[(size_t)7 * Fm + f], surf = tg[(size_t)8 * Fm + f];
    d3 P[3], u[3], v[3];
    for (int a = 0; a < 3; a++) {
        const double4 p = pos4[nd[a]];
        P[a] = mk(p.x, p.y, p.z);
        const NodeRec r = nxt[nd[a]];
        u[a] = mk(r.ux, r.uy, r.uz);
        v[a] = mk(r.vx, r.vy, r.vz);
    }
    double qu[4], qv[4];
    d3 G[4];
    for (int g = 0; g < 4; g++) {
        const double wg = tg[(size_t)(3 + g) * Fm + f];
        const d3 U = tri_a(0, g) * u[0] + tri_a(1, g) * u[1] + tri_a(2, g) * u[2];
        const d3 V = tri_a(0, g) * v[0] + tri_a(1, g) * v[1] + tri_a(2, g) * v[2];
        G[g] = tri_a(0, g) * P[0] + tri_a(1, g) * P[1] + tri_a(2, g) * P[2];
        qu[g] = dMs != 0.0 ? dMs * (wg * dot(U, n)) : 0.0;
        qv[g] = dMs != 0.0 ? dMs * (wg * dot(V, n)) : 0.0;
        if (q) q[rank[src_off + (int64_t)4 * f + g]] = make_double2(qu[g], qv[g]);
    }
    const double su[3] = {dot(u[0], n), dot(u[1], n), dot(u[2], n)};
    const double sv[3] = {dot(v[0], n), dot(v[1], n), dot(v[2], n)};
    if (qe)
        for (int a = 0; a < 3; a++) qe[Tm + 3 * f + a] = make_double2(su[a], sv[a]);
    for (int a = 0; a < 3; a++) {
        double cu = 0, cv = 0;
        for (int g = 0; g < 4; g++) {
            const d3 d = P[a] - G[g];
            const double dist = sqrt(dot(d, d));
            cu -= qu[g] / dist;
            cv -= qv[g] / dist;
        }
        cu += tri_potential(P, su, surf, dMs, a);
        cv += tri_potential(P, sv, surf, dMs, a);
        corr3[(size_t)a * Fm + f] = make_double2(cu, cv);
    }
}

__global__ __launch_bounds__(BLOCK) void k_corr_gather(int N, const int32_t *__restrict__ cptr,
                                                       const int32_t *__restrict__ cidx,
                                                       const double2 *__restrict__ corr3, double2 *__restrict__ corr)
{
    for (int i = lblock() * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        double cu = 0, cv = 0;
        for (int k = cptr[i]; k < cptr[i + 1]; k++) {
            const double2 e = corr3[cidx[k]];
            cu += e.x;
            cv += e.y;
        }
        corr[i] = make_double2(cu, cv);
    }
}

// Exact potentials phi_i = (sum_s q_s / |x_i - y_s| + corr_i) / (4 pi) (src/fmm_demag.h:211-223),
// FP64, sources staged through LDS in tiles of BLOCK. In a batch (node_inst != nullptr) a node
// sees only its own instance's sources: tetrahedron sources [r.x, r.y), triangle sources [r.z, r.w)
// with r = inst_src[instance].
__global__ __launch_bounds__(BLOCK) void k_p2p_direct(int N, int64_t S, const int32_t *__restrict__ mag,
                                                      const double4 *__restrict__ pos4,
                                                      const double4 *__restrict__ spos,
                                                      const double2 *__restrict__ q,
                                                      const double2 *__restrict__ corr, NodeRec *__restrict__ nxt,
                                                      const int32_t *__restrict__ node_inst,
                                                      const longlong4 *__restrict__ inst_src)
{
    __shared__ double4 ts[BLOCK];
    __shared__ double2 tq[BLOCK];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < N && mag[i];
    double4 xi = make_double4(0, 0, 0, 0);
    if (i < N) xi = pos4[i];
    longlong4 r = make_longlong4(0, S, 0, 0);
    if (node_inst && i < N) r = inst_src[node_inst[i]];
    double pu = 0, pv = 0;
    for (int64_t base = 0; base < S; base += BLOCK) {
        const int64_t k = base + threadIdx.x;
        if (k < S) {
            ts[threadIdx.x] = spos[k];
            tq[threadIdx.x] = q[k];
        } else {
            ts[threadIdx.x] = make_double4(1e30, 1e30, 1e30, 0);
            tq[threadIdx.x] = make_double2(0, 0);
        }
        __syncthreads();
        if (act) {
            const bool all = (base >= r.x && base + BLOCK <= r.y) || (base >= r.z && base + BLOCK <= r.w);
            if (all) {
#pragma unroll 8
                for (int j = 0; j < BLOCK; j++) {
                    const double4 y = ts[j];
                    const double dx = xi.x - y.x, dy = xi.y - y.y, dz = xi.z - y.z;
                    const double ir = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
                    const double2 qq = tq[j];
                    pu += qq.x * ir;
                    pv += qq.y * ir;
                }
            } else {
                for (int j = 0; j < BLOCK; j++) {
                    const int64_t s = base + j;
                    if (!((s >= r.x && s < r.y) || (s >= r.z && s < r.w))) continue;
                    const double4 y = ts[j];
                    const double dx = xi.x - y.x, dy = xi.y - y.y, dz = xi.z - y.z;
                    const double ir = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
                    const double2 qq = tq[j];
                    pu += qq.x * ir;
                    pv += qq.y * ir;
                }
            }
        }
        __syncthreads();
    }
    if (act) {
        const double2 c = corr[i];
        NodeRec r = nxt[i];
        r.phi = (pu + c.x) / (4 * M_PI);
        r.phiv = (pv + c.y) / (4 * M_PI);
        nxt[i] = r;
    }
}

// ------------------------------------------------------------------ energies
// Fem::energy (src/energy.cpp:5-68), element energies src/tetra.cpp:309-391 and
// src/triangle.cpp:38-43,80-85, on the NEXT state; partials per block, slots 0..3 = E[4].
__global__ __launch_bounds__(BLOCK) void k_energy_tet(int Tm, const int4 *__restrict__ tn,
                                                      const double *__restrict__ geo,
                                                      const int32_t *__restrict__ treg,
                                                      const DevRegion *__restrict__ regs,
                                                      const NodeRec *__restrict__ nxt, double Hx, double Hy,
                                                      double Hz, const double *__restrict__ hsp, double amp,
                                                      double *__restrict__ part)
{
    double e[4] = {0, 0, 0, 0};
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < Tm; t += gridDim.x * blockDim.x) {
        const int4 q = tn[t];
        const int nd[4] = {q.x, q.y, q.z, q.w};
        const DevRegion R = regs[treg[t]];
        d3 u[4];
        double ph[4];
        for (int a = 0; a < 4; a++) {
            const NodeRec n = nxt[nd[a]];
            u[a] = mk(n.ux, n.uy, n.uz);
            ph[a] = n.phi;
        }
        double da[4][3], w[5];
        for (int c = 0; c < 12; c++) da[c / 3][c % 3] = geo[(size_t)c * Tm + t];
        for (int g = 0; g < 5; g++) w[g] = geo[(size_t)(12 + g) * Tm + t];
        d3 dU[3];
        for (int k = 0; k < 3; k++) dU[k] = da[0][k] * u[0] + da[1][k] * u[1] + da[2][k] * u[2] + da[3][k] * u[3];
        const double dens = dot(dU[0], dU[0]) + dot(dU[1], dU[1]) + dot(dU[2], dU[2]);
        const double divu = dU[0].x + dU[1].y + dU[2].z;
        double ex = 0, dm = 0, an = 0, ze = 0;
        const d3 H = mk(Hx, Hy, Hz);
        for (int g = 0; g < 5; g++) {
            d3 U = mk(0, 0, 0);
            double pg = 0;
            for (int a = 0; a < 4; a++) {
                U = U + tet_a(a, g) * u[a];
                pg += ph[a] * tet_a(a, g);
            }
            ex += w[g] * dens;
            dm += w[g] * (divu * pg);
            if (R.hasK) {
                const double c = dot(mk(R.uk[0], R.uk[1], R.uk[2]), U);
                an += w[g] * (-R.K) * (c * c);
            }
            if (R.hasK3) {
                const double a0 = dot(U, mk(R.ex[0], R.ex[1], R.ex[2])), a1 = dot(U, mk(R.ey[0], R.ey[1], R.ey[2])),
                             a2 = dot(U, mk(R.ez[0], R.ez[1], R.ez[2]));
                an += w[g] * R.K3 * ((a0 * a1) * (a0 * a1) + (a1 * a2) * (a1 * a2) + (a2 * a0) * (a2 * a0));
            }
            const d3 Hg = hsp ? amp * mk(hsp[(size_t)(3 * g) * Tm + t], hsp[(size_t)(3 * g + 1) * Tm + t],
                                         hsp[(size_t)(3 * g + 2) * Tm + t])
                              : H;
            ze += w[g] * dot(U, Hg);
        }
        e[0] += R.Ae * ex;
        e[1] += an;
        e[2] += -0.5 * MU0 * R.Ms * dm;
        e[3] += -MU0 * R.Ms * ze;
    }
    const int slots[4] = {0, 1, 2, 3};
    block_store<4>(e, slots, part, blockIdx.x);
}

__global__ __launch_bounds__(BLOCK) void k_energy_tri(int Fm, const int4 *__restrict__ tn,
                                                      const double *__restrict__ tg,
                                                      const double *__restrict__ tks, // Ks, uk per tri (4 SoA)
                                                      const NodeRec *__restrict__ nxt, double *__restrict__ part)
{
    double e[2] = {0, 0};
    for (int f = lblock() * blockDim.x + threadIdx.x; f < Fm; f += gridDim.x * blockDim.x) {
        const int4 q = tn[f];
        const int nd[3] = {q.x, q.y, q.z};
        const d3 n = mk(tg[f], tg[(size_t)Fm + f], tg[(size_t)2 * Fm + f]);
        const double dMs = tg[(size_t)7 * Fm + f];
        const double Ks = tks[f];
        const d3 uk = mk(tks[(size_t)Fm + f], tks[(size_t)2 * Fm + f], tks[(size_t)3 * Fm + f]);
        d3 u[3];
        double ph[3];
        for (int a = 0; a < 3; a++) {
            const NodeRec r = nxt[nd[a]];
            u[a] = mk(r.ux, r.uy, r.uz);
            ph[a] = r.phi;
        }
        double d = 0, an = 0;
        for (int g = 0; g < 4; g++) {
            const d3 U = tri_a(0, g) * u[0] + tri_a(1, g) * u[1] + tri_a(2, g) * u[2];
            const double pg = ph[0] * tri_a(0, g) + ph[1] * tri_a(1, g) + ph[2] * tri_a(2, g);
            const double wg = tg[(size_t)(3 + g) * Fm + f];
            d += (dot(U, n) * pg) * wg;
            if (Ks != 0.0) {
                const double c = dot(U, uk);
                an += wg * (c * c);
            }
        }
        e[0] += -Ks * an;
        e[1] += 0.5 * MU0 * dMs * d;
    }
    const int slots[2] = {4, 5};
    block_store<2>(e, slots, part, lblock());
}

// final sums of the partials: out[k] = sum of slot k, k < nslots; slots >= nfirst use nb_b blocks
__global__ void k_sum_finish(int nslots, int nfirst, int nb_a, int nb_b, const double *__restrict__ part,
                             double *__restrict__ out)
{
    if (threadIdx.x < 64) {
        for (int k = 0; k < nslots; k++) {
            const int nb = k < nfirst ? nb_a : nb_b;
            const double s = wave_sum(lane_fold(part + k * MAX_PART, nb, 0.0, AddOp()));
            if (threadIdx.x == 0) out[k] = s;
        }
    }
}

// mesh::max_angle (src/mesh.h:295-306) over the magnetic edges of the stencil pattern
__global__ __launch_bounds__(BLOCK) void k_max_angle(int N, const int64_t *__restrict__ sptr,
                                                     const int32_t *__restrict__ swid,
                                                     const int32_t *__restrict__ scol,
                                                     const NodeRec *__restrict__ nxt, double *__restrict__ part)
{
    double mn = 1.0;
    for (int i = lblock() * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const int sl = i >> 6, lane = i & 63;
        const int64_t off = sptr[sl];
        const NodeRec a = nxt[i];
        for (int k = 0; k < swid[sl]; k++) {
            const int j = scol[off + (int64_t)k * SELL_C + lane];
            if (j <= i) continue;
            const NodeRec bq = nxt[j];
            mn = fmin(mn, a.ux * bq.ux + a.uy * bq.uy + a.uz * bq.uz);
        }
    }
    __shared__ double sh[BLOCK / 64];
    mn = wave_min(mn);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mn;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = sh[0];
        for (int q = 1; q < BLOCK / 64; q++) s = fmin(s, sh[q]);
        part[lblock()] = s;
    }
}

__global__ void k_min_finish(int nb, const double *__restrict__ part, double *__restrict__ out)
{
    if (threadIdx.x < 64) {
        const double m = wave_min(lane_fold(part, nb, 1.0, MinOp()));
        if (threadIdx.x == 0) out[0] = m;
    }
}

// mesh::avg (src/mesh.cpp:89-106): sum_t sum_g w_g U_g (which 0: u, 1: v), slots 0..2
__global__ __launch_bounds__(BLOCK) void k_average(int Tm, int which, const int4 *__restrict__ tn,
                                                   const double *__restrict__ geo,
                                                   const NodeRec *__restrict__ nxt, double *__restrict__ part)
{
    double e[3] = {0, 0, 0};
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < Tm; t += gridDim.x * blockDim.x) {
        const int4 q = tn[t];
        const int nd[4] = {q.x, q.y, q.z, q.w};
        d3 u[4];
        for (int a = 0; a < 4; a++) {
            const NodeRec n = nxt[nd[a]];
            u[a] = which ? mk(n.vx, n.vy, n.vz) : mk(n.ux, n.uy, n.uz);
        }
        for (int g = 0; g < 5; g++) {
            const double wg = geo[(size_t)(12 + g) * Tm + t];
            const d3 U = tet_a(0, g) * u[0] + tet_a(1, g) * u[1] + tet_a(2, g) * u[2] + tet_a(3, g) * u[3];
            e[0] += wg * U.x;
            e[1] += wg * U.y;
            e[2] += wg * U.z;
        }
    }
    const int slots[3] = {0, 1, 2};
    block_store<3>(e, slots, part, blockIdx.x);
}

// Node::evolution (src/node.h:107): CURRENT = NEXT
__global__ __launch_bounds__(BLOCK) void k_evolution(int N, const NodeRec *__restrict__ nxt, NodeRec *__restrict__ cur)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) cur[i] = nxt[i];
}

// ------------------------------------------------------------------ launchers
static inline int grid1(int64_t n) { return (int)((n + BLOCK - 1) / BLOCK); }

// Tetra::calc_Hst (src/tetra.cpp:93-106): Hst_g = pref_region sum_i s_i a_ig at the Gauss points of
// each magnetic tetrahedron, [15][Tm]
__global__ __launch_bounds__(BLOCK) void k_spin_field(int Tm, const int4 *__restrict__ tn, const int32_t *__restrict__ treg,
                                                      const double *__restrict__ pref, const double *__restrict__ s,
                                                      double *__restrict__ hst)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Tm) return;
    const int4 q = tn[t];
    const int nd[4] = {q.x, q.y, q.z, q.w};
    const double p = pref[treg[t]];
    double sn[4][3];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) sn[i][k] = s[3 * (size_t)nd[i] + k];
#pragma unroll
    for (int g = 0; g < 5; g++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double v = 0.0;
#pragma unroll
            for (int i = 0; i < 4; i++) v += sn[i][k] * tet_a(i, g);
            hst[(size_t)(3 * g + k) * Tm + t] = p * v;
        }
}

void launch_spin_field(Ctx &c)
{
    if (c.Tm > 0)
        k_spin_field<<<grid1(c.Tm), BLOCK, 0, c.stream>>>(c.Tm, c.tet_nodes, c.tet_reg, c.spin_pref, c.spin_s, c.hst);
}

void launch_element(Ctx &c, const double Hext[3], double dt, const double *hsp, double amp)
{
    const NodeRec *cur = c.rec[c.cur];
    if (c.Tm > 0) {
        Prof pr(c, "element");
        k_element<<<grid1(c.Tm), BLOCK, 0, c.stream>>>(c.Tm, c.tet_nodes, c.geo, c.tet_reg, c.regions, cur, Hext[0],
                                                       Hext[1], Hext[2], hsp, amp, c.hst, dt, c.epos, c.elem_out);
    }
    if (c.Fk > 0) {
        Prof pr(c, "tri_aniso");
        k_tri_aniso<<<grid1(c.Fk), BLOCK, 0, c.stream>>>(c.Fk, c.trik_nodes, c.trik_geo, cur,
                                                         c.epos + (size_t)4 * c.Tm, c.elem_out);
    }
}

void launch_node_assemble(Ctx &c, double r, double cst, double /*dt*/)
{
    const int nb = c.blocks_for(c.N);
    const NodeRec *vsrc = c.rec[1 - c.cur];  // buildInitGuess reads v NEXT
    {
        Prof pr(c, "node_assemble");
        k_node_assemble<<<nb, BLOCK, 0, c.stream>>>(c.N, cos(r), sin(r), cst, c.rec[c.cur], vsrc, c.mag, c.inc_ptr,
                                                    c.elem_out, c.mass, c.sdiag, c.ep4, c.eq4, c.Dg,
                                                    c.dprec, c.b, c.x, c.w, c.part);
    }
    k_set_rhsn<<<1, 64, 0, c.stream>>>(nb, c.ctl, c.part);
    const int K = (int)c.inst_node_off.size() - 1;
    if (K > 1) {
        k_inst_norm2<<<K, BLOCK, 0, c.stream>>>(c.d_inst_node_off, c.b, c.inst_bn2);
        k_inst_rhsn<<<1, 64, 0, c.stream>>>(K, c.inst_bn2, c.ctl);
    }
}

void launch_basis(Ctx &c, double r)
{
    k_basis<<<c.blocks_for(c.N), BLOCK, 0, c.stream>>>(c.N, cos(r), sin(r), c.rec[c.cur], c.ep4, c.eq4);
}

void launch_r0(Ctx &c, double cst)
{
    const int nb = c.blocks_for(c.N);
    Prof pr(c, "spmv_r0");
    k_r0<<<nb, BLOCK, 0, c.stream>>>(c.N, cst, STENCIL_PASS, c.w, c.x, c.b, c.r, c.rt, c.p, c.ctl, c.part);
}

void launch_iteration(Ctx &c, double cst, int part)
{
    const int nb = c.blocks_for(c.N);
    if (part != 2) {
        Prof pr(c, "iter_p");
        k_iter_p<<<nb, BLOCK, 0, c.stream>>>(c.N, nb, c.ctl, c.ep4, c.eq4, c.dprec, c.r, c.v, c.p, c.ph, c.w, c.part);
    }
    if (part == 1) return;
    {
        Prof pr(c, "spmv");
        k_spmv_v<<<nb, BLOCK, 0, c.stream>>>(c.N, cst, STENCIL_PASS, c.ctl, c.w, c.ph, c.rt, c.v, c.part);
    }
    {
        Prof pr(c, "iter_s");
        k_iter_s<<<nb, BLOCK, 0, c.stream>>>(c.N, nb, c.ctl, c.ep4, c.eq4, c.dprec, c.r, c.v, c.s, c.sh, c.w, c.part);
    }
    {
        Prof pr(c, "spmv");
        k_spmv_t<<<nb, BLOCK, 0, c.stream>>>(c.N, nb, cst, STENCIL_PASS, c.ctl, c.w, c.sh, c.s, c.ph, c.x, c.t, c.part);
    }
    {
        Prof pr(c, "iter_x");
        k_iter_x<<<nb, BLOCK, 0, c.stream>>>(c.N, nb, c.ctl, c.ph, c.sh, c.s, c.t, c.rt, c.x, c.r, c.part);
    }
}

void launch_mixed_start(Ctx &c)
{
    const int nb = c.blocks_for(c.N);
    {
        Prof pr(c, "to_f32");
        k_to_f32<<<nb, BLOCK, 0, c.stream>>>(c.N, c.ep4, c.eq4, c.Dg, c.dprec, c.ep4f, c.eq4f, c.Dgf, c.dprecf);
    }
    k_mixed_init<<<1, 64, 0, c.stream>>>(c.ctl);
}

void launch_mixed_residual(Ctx &c, double cst, double tol_in, int max_inner)
{
    const int nb = c.blocks_for(c.N);
    {
        Prof pr(c, "spmv_r0");
        k_resid<<<nb, BLOCK, 0, c.stream>>>(c.N, cst, STENCIL_PASS, c.ctl, c.w, c.x, c.b, c.rf, c.rtf, c.pf, c.df,
                                            c.part);
    }
    k_outer<<<1, 64, 0, c.stream>>>(nb, c.ctl, c.ic, tol_in, max_inner, c.part);
}

void launch_inner_iteration(Ctx &c, double cst)
{
    const int nb = c.blocks_for(c.N);
    const float cf = (float)cst;
    {
        Prof pr(c, "iter_p");
        kf_iter_p<<<nb, BLOCK, 0, c.stream>>>(c.N, nb, c.ic, c.ep4f, c.eq4f, c.dprecf, c.rf, c.vf, c.pf, c.phf, c.wf,
                                              c.part);
    }
    {
        Prof pr(c, "spmv");
        kf_spmv_v<<<nb, BLOCK, 0, c.stream>>>(c.N, cf, STENCIL_PASS_F, c.ic, c.wf, c.phf, c.rtf, c.vf, c.part);
    }
    {
        Prof pr(c, "iter_s");
        kf_iter_s<<<nb, BLOCK, 0, c.stream>>>(c.N, nb, c.ic, c.ep4f, c.eq4f, c.dprecf, c.rf, c.vf, c.sf, c.shf, c.wf,
                                              c.part);
    }
    {
        Prof pr(c, "spmv");
        kf_spmv_t<<<nb, BLOCK, 0, c.stream>>>(c.N, nb, cf, STENCIL_PASS_F, c.ic, c.wf, c.shf, c.sf, c.phf, c.df, c.tf,
                                              c.part);
    }
    {
        Prof pr(c, "iter_x");
        kf_iter_x<<<nb, BLOCK, 0, c.stream>>>(c.N, nb, c.ic, c.phf, c.shf, c.sf, c.tf, c.rtf, c.df, c.rf, c.part);
    }
}

void launch_batch_solve(Ctx &c, double cst, int block)
{
    const int K = (int)c.inst_node_off.size() - 1;
    Prof pr(c, "bicg_batch");
#define FG_BATCH_ARGS                                                                                         \
    c.d_inst_order, c.d_inst_node_off, cst, c.opt.tol, c.opt.max_iter, STENCIL_PASS, c.dprec, c.b, c.x, c.r, c.rt, \
        c.p, c.ph, c.v, c.s, c.sh, c.t, c.w, c.d_inst_stat, c.d_inst_resn
    if (block == 1024) k_bicg_batch<1024><<<K, 1024, 0, c.stream>>>(FG_BATCH_ARGS);
    else if (block == 512) k_bicg_batch<512><<<K, 512, 0, c.stream>>>(FG_BATCH_ARGS);
    else k_bicg_batch<256><<<K, 256, 0, c.stream>>>(FG_BATCH_ARGS);
#undef FG_BATCH_ARGS
}

void launch_refine(Ctx &c)
{
    Prof pr(c, "refine");
    k_refine<<<c.blocks_for(c.N), BLOCK, 0, c.stream>>>(c.N, c.ctl, c.ic, c.ep4, c.eq4, c.df, c.x, c.w);
}

void launch_update(Ctx &c, double dt, bool gated)
{
    const int nb = c.blocks_for(c.N);
    {
        Prof pr(c, "update");
        k_update<<<nb, BLOCK, 0, c.stream>>>(c.N, dt, c.mag, c.x, c.ep4, c.eq4, c.rec[c.cur], c.rec[1 - c.cur], c.part,
                                             gated ? c.ctl : nullptr);
    }
    k_finish_update<<<1, 64, 0, c.stream>>>(nb, c.ctl, c.part, gated);
}

void launch_spmv_debug(Ctx &c, double cst, const double2 *xin, double2 *yout)
{
    const int nb = c.blocks_for(c.N);
    k_lift<<<nb, BLOCK, 0, c.stream>>>(c.N, c.ep4, c.eq4, xin, c.w);
    k_spmv_plain<<<nb, BLOCK, 0, c.stream>>>(c.N, cst, STENCIL_PASS, c.w, xin, yout);
}

void launch_charges(Ctx &c)
{
    const NodeRec *nxt = c.rec[1 - c.cur];
    double2 *q = c.fmm ? nullptr : c.src_q;  // the FMM's P2M and near field take the element values
    if (c.Tm > 0) {
        Prof pr(c, "charges_tet");
        k_charges_tet<<<grid1(c.Tm), BLOCK, 0, c.stream>>>(c.Tm, c.tet_nodes, c.geo, c.tet_reg, c.regions, nxt,
                                                           c.src_rank, q, c.qe);
    }
    if (c.Fm > 0) {
        Prof pr(c, "charges_tri");
        k_charges_tri<<<grid1(c.Fm), BLOCK, 0, c.stream>>>(c.Fm, (int64_t)5 * c.Tm, c.tri_nodes, c.tri_geo, c.pos4,
                                                           nxt, c.src_rank, q, c.tri_corr, c.qe, c.Tm);
    }
    Prof pr(c, "corr_gather");
    k_corr_gather<<<c.blocks_for(c.N), BLOCK, 0, c.stream>>>(c.N, c.cinc_ptr, c.cinc_idx, c.tri_corr, c.corr);
}

void launch_direct_p2p(Ctx &c)
{
    Prof pr(c, "p2p_direct");
    k_p2p_direct<<<grid1(c.N), BLOCK, 0, c.stream>>>(c.N, c.S, c.mag, c.pos4, c.src_pos, c.src_q, c.corr,
                                                     c.rec[1 - c.cur], c.d_node_inst, c.d_inst_src);
}

void launch_energy(Ctx &c, const double Hext[3], const double *hsp, double amp)
{
    const NodeRec *nxt = c.rec[1 - c.cur];
    const int nbt = c.blocks_for(c.Tm), nbf = c.blocks_for(c.Fm);
    {
        Prof pr(c, "energy");
        k_energy_tet<<<nbt, BLOCK, 0, c.stream>>>(c.Tm, c.tet_nodes, c.geo, c.tet_reg, c.regions, nxt, Hext[0], Hext[1],
                                                  Hext[2], hsp, amp, c.epart);
        if (c.Fm > 0)
            k_energy_tri<<<nbf, BLOCK, 0, c.stream>>>(c.Fm, c.tri_nodes, c.tri_geo, c.tri_geo + (size_t)12 * c.Fm, nxt,
                                                      c.epart);
    }
    k_sum_finish<<<1, 64, 0, c.stream>>>(6, 4, nbt, c.Fm > 0 ? nbf : 0, c.epart, c.res);
}

void launch_max_angle(Ctx &c)
{
    const int nb = c.blocks_for(c.N);
    Prof pr(c, "max_angle");
    k_max_angle<<<nb, BLOCK, 0, c.stream>>>(c.N, c.slice_ptr, c.slice_w, c.scol, c.rec[1 - c.cur], c.epart);
    k_min_finish<<<1, 64, 0, c.stream>>>(nb, c.epart, c.res + 8);
}

void launch_evolution(Ctx &c)
{
    Prof pr(c, "evolution");
    k_evolution<<<c.blocks_for(c.N), BLOCK, 0, c.stream>>>(c.N, c.rec[1 - c.cur], c.rec[c.cur]);
}

void launch_average(Ctx &c, int which)
{
    const int nb = c.blocks_for(c.Tm);
    k_average<<<nb, BLOCK, 0, c.stream>>>(c.Tm, which, c.tet_nodes, c.geo, c.rec[1 - c.cur], c.epart);
    k_sum_finish<<<1, 64, 0, c.stream>>>(3, 3, nb, 0, c.epart, c.res + 16);
}

}  // namespace fg
