// transport.hip -- the current-driven side problems of spin-transfer runs: the electrostatic
// potential of the current injection (FeeLLGood's electrostatSolver, src/electrostatSolver.cpp:
// 10-136) and the spin accumulation of the spin-diffusion problem (spinAcc,
// src/spinAccumulationSolver.cpp:41-287). Each assembles its finite-element matrix on the device
// (every entry summed from its tetrahedra in mesh order, from a list the host builds once), takes
// the surface terms on the host (surface-sized work), and runs its Krylov solver on the device with
// the scalars of each iteration read back by the host: conjugate gradient with the diagonal
// preconditioner and Dirichlet masking (algebra::cg_dir, src/algebra/cg.h:67-121) for the
// potential, BiCGSTAB with Dirichlet values (algebra::bicg_dir, src/algebra/bicg.h:82-158) for the
// spin accumulation.
//
// Both are solved at set-up (src/spinAccumulationSolver.cpp:21-36), so the kernels favour
// determinism over speed: every dot product is a fixed tree in each block with the block partials
// summed in order on the host.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/fgmi.h"
#include "common.h"

using namespace fg;

namespace {

#define ESCHK(expr)                                                                        \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) throw std::string(#expr) + ": " + hipGetErrorString(e_);    \
    } while (0)

constexpr int ESB = 256;  // threads per block

// The shape-function gradients da and det J of tetrahedron t, with the nodes reordered for a positive
// Jacobian as Tetra::Tet does (src/tetra.h:280-303); perm[k] is the caller's local position of
// the k-th reordered node, id[] the reordered node indices.
__device__ inline double tet_grad(const int4 q, const double *__restrict__ pos, double da[4][3], int perm[4], int id[4])
{
    id[0] = q.x;
    id[1] = q.y;
    id[2] = q.z;
    id[3] = q.w;
    d3 P[4];
#pragma unroll
    for (int a = 0; a < 4; a++) P[a] = mk(pos[3 * (size_t)id[a]], pos[3 * (size_t)id[a] + 1], pos[3 * (size_t)id[a] + 2]);
    d3 e1 = P[1] - P[0], e2 = P[2] - P[0], e3 = P[3] - P[0];
#pragma unroll
    for (int a = 0; a < 4; a++) perm[a] = a;
    if (dot(e1, cross(e2, e3)) < 0) {
        const d3 s = e2;
        e2 = e3;
        e3 = s;
        perm[2] = 3;
        perm[3] = 2;
        const int k = id[2];
        id[2] = id[3];
        id[3] = k;
    }
    const double J[3][3] = {{e1.x, e2.x, e3.x}, {e1.y, e2.y, e3.y}, {e1.z, e2.z, e3.z}};
    const double detJ = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                        J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                        J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double cf[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            cf[i][j] = J[i1][j1] * J[i2][j2] - J[i1][j2] * J[i2][j1];
        }
    const double inv = 1.0 / (cf[0][0] * J[0][0] + cf[1][0] * J[1][0] + cf[2][0] * J[2][0]);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double j0 = cf[k][0] * inv, j1 = cf[k][1] * inv, j2 = cf[k][2] * inv;  // column k of J^-1
        da[0][k] = -j0 - j1 - j2;
        da[1][k] = j0;
        da[2][k] = j1;
        da[3][k] = j2;
    }
    return detJ;
}

// K_e[i][j] = sigma sum_g w_g da_i . da_j (electrostatSolver::integrales, src/electrostatSolver.cpp:42-65),
// stored at the caller's local positions, [16][T].
__global__ __launch_bounds__(ESB) void k_es_element(int T, const int4 *__restrict__ tn, const double *__restrict__ pos,
                                                    const double *__restrict__ tsig, double *__restrict__ ke)
{
    const int t = blockIdx.x * ESB + threadIdx.x;
    if (t >= T) return;
    double da[4][3];
    int perm[4], id[4];
    const double detJ = tet_grad(tn[t], pos, da, perm, id);
    const double sg = tsig[t];
    double AE[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) AE[i][j] = 0.0;
#pragma unroll
    for (int g = 0; g < NPI_TET; g++) {
        const double sw = sg * (detJ * TET_PDS[g]);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                AE[i][j] += sw * (da[i][0] * da[j][0] + da[i][1] * da[j][1] + da[i][2] * da[j][2]);
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) ke[(size_t)(perm[i] * 4 + perm[j]) * T + t] = AE[i][j];
}

// each CSR entry is the sum of its tetrahedra's K_e entries in mesh order (solver::buildMat,
// src/solver.h:110-131)
__global__ __launch_bounds__(ESB) void k_es_gather(int64_t nnz, const int64_t *__restrict__ cptr,
                                                   const int64_t *__restrict__ cslot, const double *__restrict__ ke,
                                                   double *__restrict__ val)
{
    const int64_t k = (int64_t)blockIdx.x * ESB + threadIdx.x;
    if (k >= nnz) return;
    double s = 0.0;
    for (int64_t c = cptr[k]; c < cptr[k + 1]; c++) s += ke[cslot[c]];
    val[k] = s;
}

// sum of the block's values, a fixed tree; thread 0 writes the block's partial
__device__ inline void block_sum(double v, double *sh, double *out)
{
    sh[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int s = ESB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// D = 1 / diag(K), zero on the Dirichlet nodes (build_diag_precond, src/algebra/sparseMat.h:173-183,
// then applyMask)
__global__ __launch_bounds__(ESB) void k_es_diag(int n, const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                 const double *__restrict__ val, const uint8_t *__restrict__ dir,
                                                 double *__restrict__ D)
{
    const int i = blockIdx.x * ESB + threadIdx.x;
    if (i >= n) return;
    double c = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; k++)
        if (col[k] == i) c = val[k];
    D[i] = dir[i] ? 0.0 : 1.0 / c;
}

// y = K x, zero on the Dirichlet nodes when mask is set; partial sums of y . w (w may be null)
__global__ __launch_bounds__(ESB) void k_es_spmv(int n, const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                 const double *__restrict__ val, const double *__restrict__ x,
                                                 const uint8_t *__restrict__ dir, int mask, double *__restrict__ y,
                                                 const double *__restrict__ w, double *__restrict__ part)
{
    __shared__ double sh[ESB];
    const int i = blockIdx.x * ESB + threadIdx.x;
    double yw = 0.0;
    if (i < n) {
        double s = 0.0;
        for (int64_t k = rp[i]; k < rp[i + 1]; k++) s += val[k] * x[col[k]];
        if (mask && dir[i]) s = 0.0;
        y[i] = s;
        if (w) yw = s * w[i];
    }
    if (part) block_sum(yw, sh, part + blockIdx.x);
}

// b = rhs - K xd, masked; r = b; z = D r; p = z; partials of r . r and z . r
__global__ __launch_bounds__(ESB) void k_es_init(int n, const double *__restrict__ rhs, const double *__restrict__ kxd,
                                                 const uint8_t *__restrict__ dir, const double *__restrict__ D,
                                                 double *__restrict__ r, double *__restrict__ z, double *__restrict__ p,
                                                 double *__restrict__ prr, double *__restrict__ pzr)
{
    __shared__ double sh[ESB];
    const int i = blockIdx.x * ESB + threadIdx.x;
    double rr = 0.0, zr = 0.0;
    if (i < n) {
        double b = rhs[i];
        b -= kxd[i];
        if (dir[i]) b = 0.0;
        const double zi = D[i] * b;
        r[i] = b;
        z[i] = zi;
        p[i] = zi;
        rr = b * b;
        zr = zi * b;
    }
    block_sum(rr, sh, prr + blockIdx.x);
    block_sum(zr, sh, pzr + blockIdx.x);
}

// p = p (rho / rho_1) + z
__global__ __launch_bounds__(ESB) void k_es_dir(int n, double ratio, const double *__restrict__ z, double *__restrict__ p)
{
    const int i = blockIdx.x * ESB + threadIdx.x;
    if (i >= n) return;
    double v = p[i];
    v *= ratio;
    v += z[i];
    p[i] = v;
}

// x += a p; r += (-a) q; z = D r; partials of r . r and z . r
__global__ __launch_bounds__(ESB) void k_es_update(int n, double a, const double *__restrict__ p,
                                                   const double *__restrict__ q, const double *__restrict__ D,
                                                   double *__restrict__ x, double *__restrict__ r, double *__restrict__ z,
                                                   double *__restrict__ prr, double *__restrict__ pzr)
{
    __shared__ double sh[ESB];
    const int i = blockIdx.x * ESB + threadIdx.x;
    double rr = 0.0, zr = 0.0;
    if (i < n) {
        x[i] += a * p[i];
        const double ri = r[i] + (-a) * q[i];
        r[i] = ri;
        const double zi = D[i] * ri;
        z[i] = zi;
        rr = ri * ri;
        zr = zi * ri;
    }
    block_sum(rr, sh, prr + blockIdx.x);
    block_sum(zr, sh, pzr + blockIdx.x);
}

__global__ __launch_bounds__(ESB) void k_es_add(int n, const double *__restrict__ xd, double *__restrict__ x)
{
    const int i = blockIdx.x * ESB + threadIdx.x;
    if (i < n) x[i] += xd[i];
}

__global__ __launch_bounds__(ESB) void k_es_zero(int n, double *__restrict__ x)
{
    const int i = blockIdx.x * ESB + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

struct DevBufs {
    std::vector<void *> ptrs;
    template <class T>
    T *alloc(size_t n)
    {
        void *p = nullptr;
        ESCHK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        ptrs.push_back(p);
        return (T *)p;
    }
    template <class T>
    T *upload(const std::vector<T> &h)
    {
        T *d = alloc<T>(h.size());
        if (!h.empty()) ESCHK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        return d;
    }
    ~DevBufs()
    {
        for (void *p : ptrs) (void)hipFree(p);
    }
};

// Device slices carved from one buffer the caller does not allocate (the static scratch array
// below): 256-byte aligned, never freed. A slice that would pass the end throws ScratchFull before
// it is handed out, so nothing is written beyond the buffer.
struct ScratchFull {};
struct ArenaBufs {
    char *base;
    size_t cap, off = 0;
    template <class T>
    T *alloc(size_t n)
    {
        const size_t bytes = (std::max<size_t>(n, 1) * sizeof(T) + 255) & ~(size_t)255;
        if (bytes > cap - off) throw ScratchFull{};
        T *p = (T *)(base + off);
        off += bytes;
        return p;
    }
    template <class T>
    T *upload(const std::vector<T> &h)
    {
        T *d = alloc<T>(h.size());
        if (!h.empty()) ESCHK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        return d;
    }
};

double host_sum(const std::vector<double> &h)
{
    double s = 0.0;
    for (double v : h) s += v;
    return s;
}


// ---- spin accumulation ----

// Per tetrahedron (spinAcc::integrales, src/spinAccumulationSolver.cpp:216-287), at the caller's
// local positions: rows 0-15 the diffusion block A_ij = (D0 sum w) da_i . da_j + [i == j] D0 / lsf^2
// a_w,i (Tet::calcDiagBlock, src/tetra.cpp:133-140); rows 16-27 the s-d exchange of node i,
// D0 / lsd^2 a_w,i u_i (X, Y, Z); rows 28-39 the source cst0 sum_g w_g (grad V . da_i) usrc_i; the
// last two zero outside magnetic regions. prm = (D0, D0 / lsf^2, D0 / lsd^2, cst0).
__global__ __launch_bounds__(ESB) void k_sa_element(int T, const int4 *__restrict__ tn, const double *__restrict__ pos,
                                                    const double4 *__restrict__ prm, const uint8_t *__restrict__ mag,
                                                    const double *__restrict__ V, const double *__restrict__ u,
                                                    const double *__restrict__ usrc, double *__restrict__ ke)
{
    const int t = blockIdx.x * ESB + threadIdx.x;
    if (t >= T) return;
    double da[4][3];
    int perm[4], id[4];
    const double detJ = tet_grad(tn[t], pos, da, perm, id);
    const double4 pr = prm[t];
    double w[NPI_TET], wsum = 0.0, aw[4];
#pragma unroll
    for (int g = 0; g < NPI_TET; g++) {
        w[g] = detJ * TET_PDS[g];
        wsum += w[g];
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        aw[i] = 0.0;
#pragma unroll
        for (int g = 0; g < NPI_TET; g++) aw[i] += tet_a(i, g) * w[g];
    }
    const double c = pr.x * wsum;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            double v = (da[i][0] * da[j][0] + da[i][1] * da[j][1] + da[i][2] * da[j][2]) * c;
            if (i == j) v += pr.y * aw[i];
            ke[(size_t)(perm[i] * 4 + perm[j]) * T + t] = v;
        }
    const bool mg = mag[t];
    double gV[3] = {0.0, 0.0, 0.0};
    if (mg)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const double vi = V[id[i]];
#pragma unroll
            for (int k = 0; k < 3; k++) gV[k] += vi * da[i][k];
        }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const double ui[3] = {u[3 * (size_t)id[i]], u[3 * (size_t)id[i] + 1], u[3 * (size_t)id[i] + 2]};
        const double us[3] = {usrc[3 * (size_t)id[i]], usrc[3 * (size_t)id[i] + 1], usrc[3 * (size_t)id[i] + 2]};
        double be[3] = {0.0, 0.0, 0.0};
        if (mg)
#pragma unroll
            for (int g = 0; g < NPI_TET; g++) {
                const double cw = pr.w * w[g];
                const double tmp = (cw * gV[0]) * da[i][0] + (cw * gV[1]) * da[i][1] + (cw * gV[2]) * da[i][2];
#pragma unroll
                for (int k = 0; k < 3; k++) be[k] += tmp * us[k];
            }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            ke[(size_t)(16 + 3 * perm[i] + k) * T + t] = mg ? pr.z * (aw[i] * ui[k]) : 0.0;
            ke[(size_t)(28 + 3 * perm[i] + k) * T + t] = be[k];
        }
    }
}

// the assembly of one call added to the kept system (acc) or replacing it: matrix values in the
// pattern's order, each entry's tetrahedra in mesh order after the kept value, as solver::buildMat
// adds into K (src/solver.h:110-129)
__global__ __launch_bounds__(ESB) void k_sp_gather(int64_t nnz, const int64_t *__restrict__ cptr,
                                                   const int64_t *__restrict__ cslot, const double *__restrict__ ke,
                                                   int acc, double *__restrict__ val)
{
    const int64_t k = (int64_t)blockIdx.x * ESB + threadIdx.x;
    if (k >= nnz) return;
    double s = acc ? val[k] : 0.0;
    for (int64_t c = cptr[k]; c < cptr[k + 1]; c++) s += ke[cslot[c]];
    val[k] = s;
}

// per node, in mesh order of its tetrahedra: the s-d exchange sums S (X, Y, Z) and the volume
// right-hand side after the kept values (acc), then the surface right-hand side Ls of the
// triangles (solver::buildVect, src/solver.h:131-142, volume terms first, :173-197);
// slot = local position * T + tetrahedron
__global__ __launch_bounds__(ESB) void k_sa_node(int N, int T, const int64_t *__restrict__ nptr,
                                                 const int64_t *__restrict__ nslot, const double *__restrict__ ke,
                                                 const double *__restrict__ Ls, int acc, double *__restrict__ S,
                                                 double *__restrict__ L)
{
    const int i = blockIdx.x * ESB + threadIdx.x;
    if (i >= N) return;
    double sk[3], b[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        sk[k] = acc ? S[3 * (size_t)i + k] : 0.0;
        b[k] = acc ? L[3 * (size_t)i + k] : 0.0;
    }
    for (int64_t c = nptr[i]; c < nptr[i + 1]; c++) {
        const int64_t l = nslot[c] / T, t = nslot[c] % T;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            sk[k] += ke[(size_t)(16 + 3 * l + k) * T + t];
            b[k] += ke[(size_t)(28 + 3 * l + k) * T + t];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        S[3 * (size_t)i + k] = sk[k];
        L[3 * (size_t)i + k] = b[k] + Ls[3 * (size_t)i + k];
    }
}

// the node magnetizations of an engine state: u from CURRENT (the s-d exchange) and u_src from
// NEXT (the source)
__global__ __launch_bounds__(ESB) void k_sp_rec_u(int N, const NodeRec *__restrict__ cur,
                                                  const NodeRec *__restrict__ nxt, double *__restrict__ u,
                                                  double *__restrict__ us)
{
    const int i = blockIdx.x * ESB + threadIdx.x;
    if (i >= N) return;
    const NodeRec a = cur[i], b = nxt[i];
    u[3 * (size_t)i] = a.ux;
    u[3 * (size_t)i + 1] = a.uy;
    u[3 * (size_t)i + 2] = a.uz;
    us[3 * (size_t)i] = b.ux;
    us[3 * (size_t)i + 1] = b.uy;
    us[3 * (size_t)i + 2] = b.uz;
}

// Scalars of algebra::bicg_dir and algebra::iteration (src/algebra/bicg.h:82-158,
// src/algebra/iter.h:119-178) in device memory, updated by one-block kernels between the vector
// phases so that chunks of iterations run with no host read. stop: the loop has ended; fin = 1:
// its last half step (x += alpha phat) is due, 2: taken.
struct SpinCtl {
    double rho1, rho2, alpha, omega, beta, rhsn, res, tol;
    int nit, max_iter, status, stop, fin, pad[3];
};

__global__ void k_sp_reset(SpinCtl *c, double tol, int max_iter)
{
    if (threadIdx.x) return;
    *c = SpinCtl{};
    c->tol = tol;
    c->max_iter = max_iter;
    c->status = FG_ITER_UNDEFINED;
}

enum SpinPhase { SP_INIT, SP_ALPHA, SP_S, SP_OMEGA, SP_LOOP };

// entry (d, e) of node i's own 3x3 block: the diffusion diagonal, and the antisymmetric s-d
// exchange (rows y, z of X; x, z of Y; x, y of Z, with the signs of spinAcc::integrales)
__device__ inline double self_block(int d, int e, double a, const double *S)
{
    if (d == e) return a;
    if (d == 0) return e == 1 ? S[2] : -S[1];
    if (d == 1) return e == 0 ? -S[2] : S[0];
    return e == 0 ? S[1] : -S[0];
}

// y = K x over the 3 N rows (row 3 i + d), zero on the Dirichlet rows when mask is set; partials of
// y . w (w may be null) and, when pyy is set, of y . y. Nothing once the solver has stopped.
__global__ __launch_bounds__(ESB) void k_sa_spmv(int n3, const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                 const double *__restrict__ a, const double *__restrict__ S,
                                                 const double *__restrict__ x, const uint8_t *__restrict__ dir, int mask,
                                                 double *__restrict__ y, const double *__restrict__ w,
                                                 double *__restrict__ pyw, double *__restrict__ pyy,
                                                 const SpinCtl *__restrict__ ctl)
{
    __shared__ double sh[ESB];
    if (ctl && ctl->stop) return;
    const int r = blockIdx.x * ESB + threadIdx.x;
    double yw = 0.0, yy = 0.0;
    if (r < n3) {
        const int i = r / 3, d = r - 3 * i;
        double s = 0.0;
        for (int64_t k = rp[i]; k < rp[i + 1]; k++) {
            const int j = col[k];
            if (j != i) {
                s += a[k] * x[3 * (size_t)j + d];
            } else {
#pragma unroll
                for (int e = 0; e < 3; e++) s += self_block(d, e, a[k], S + 3 * (size_t)i) * x[3 * (size_t)i + e];
            }
        }
        if (mask && dir[i]) s = 0.0;
        y[r] = s;
        if (w) yw = s * w[r];
        yy = s * s;
    }
    if (pyw) block_sum(yw, sh, pyw + blockIdx.x);
    if (pyy) block_sum(yy, sh, pyy + blockIdx.x);
}

// D = 1 / diag(K), zero on the Dirichlet rows
__global__ __launch_bounds__(ESB) void k_sa_diag(int n3, const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                 const double *__restrict__ a, const uint8_t *__restrict__ dir,
                                                 double *__restrict__ D)
{
    const int r = blockIdx.x * ESB + threadIdx.x;
    if (r >= n3) return;
    const int i = r / 3;
    double c = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; k++)
        if (col[k] == i) c = a[k];
    D[r] = dir[i] ? 0.0 : 1.0 / c;
}

// b = L - K xd, masked; r = rt = p = b; x = 0; partials of b . b
__global__ __launch_bounds__(ESB) void k_sa_init(int n3, const double *__restrict__ L, const double *__restrict__ kxd,
                                                 const uint8_t *__restrict__ dir, double *__restrict__ r,
                                                 double *__restrict__ rt, double *__restrict__ p, double *__restrict__ x,
                                                 double *__restrict__ pbb)
{
    __shared__ double sh[ESB];
    const int k = blockIdx.x * ESB + threadIdx.x;
    double bb = 0.0;
    if (k < n3) {
        double b = L[k];
        b -= kxd[k];
        if (dir[k / 3]) b = 0.0;
        r[k] = b;
        rt[k] = b;
        p[k] = b;
        x[k] = 0.0;
        bb = b * b;
    }
    block_sum(bb, sh, pbb + blockIdx.x);
}

// The scalar step that follows a reduction: the partials of one (two for SP_OMEGA and SP_LOOP)
// dot products, nb each, summed in a fixed order per thread and a fixed tree; thread 0 then takes
// the step of bicg_dir that needs them.
__global__ __launch_bounds__(ESB) void k_sp_ctl(int nb, const double *__restrict__ part, SpinCtl *__restrict__ c,
                                                int phase)
{
    __shared__ double sh[2][ESB];
    if (c->stop && !(phase == SP_LOOP && c->fin == 1)) return;
    const int nsum = (phase == SP_OMEGA || phase == SP_LOOP) ? 2 : 1;
    for (int q = 0; q < nsum; q++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += ESB) s += part[(size_t)q * nb + b];
        sh[q][threadIdx.x] = s;
    }
    __syncthreads();
    for (int w = ESB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < nsum; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x) return;
    const double a0 = sh[0][0], a1 = nsum > 1 ? sh[1][0] : 0.0;
    // iteration::converged (src/algebra/iter.h:145-166): NaN sets CANNOT_CONVERGE without ending
    // the loop, as in the reference
    auto converged = [c](double nr) {
        c->res = fabs(nr);
        if (isnan(c->res)) {
            c->status = FG_ITER_CANNOT_CONVERGE;
            return false;
        }
        const bool cv = c->res <= c->rhsn * c->tol;
        if (cv) c->status = FG_ITER_CONVERGED;
        return cv;
    };
    switch (phase) {
    case SP_INIT:  // ||b||, the loop's first test, rho1 = rt . r with rt = r = b
        c->rhsn = sqrt(a0);
        if (converged(sqrt(a0))) {
            c->stop = 1;
            return;
        }
        c->rho1 = a0;
        break;
    case SP_ALPHA:  // alpha = rho1 / (v . rt)
        c->alpha = c->rho1 / a0;
        break;
    case SP_S:  // the test on ||s||: the last half step, or a break on overflow / breakdown
        if (converged(sqrt(a0))) {
            c->stop = 1;
            c->fin = 1;
        } else if (c->status == FG_ITER_OVERFLOW || c->status == FG_ITER_CANNOT_CONVERGE) {
            c->stop = 1;
        }
        break;
    case SP_OMEGA:  // omega = (t . s) / (t . t)
        c->omega = a0 / a1;
        break;
    case SP_LOOP:  // rho2 = rho1, ++iter, the loop's test on ||r||, rho1 = rt . r, beta
        if (c->stop) {
            c->fin = 2;
            return;
        }
        c->rho2 = c->rho1;
        if (++c->nit >= c->max_iter) c->status = FG_ITER_OVERFLOW;
        if (converged(sqrt(a0))) {
            c->stop = 1;
            return;
        }
        c->rho1 = a1;
        if (c->rho2 == 0 || c->omega == 0) {
            c->status = FG_ITER_CANNOT_CONVERGE;
            c->stop = 1;
            return;
        }
        c->beta = (c->rho1 / c->rho2) * (c->alpha / c->omega);
        break;
    }
}

// p = (p - omega v) beta + r after the first iteration; phat = D p
__global__ __launch_bounds__(ESB) void k_sp_p(int n3, const SpinCtl *__restrict__ c, const double *__restrict__ v,
                                              const double *__restrict__ r, const double *__restrict__ D,
                                              double *__restrict__ p, double *__restrict__ phat)
{
    if (c->stop) return;
    const int k = blockIdx.x * ESB + threadIdx.x;
    if (k >= n3) return;
    double pv = p[k];
    if (c->nit > 0) {
        pv -= v[k] * c->omega;
        pv *= c->beta;
        pv += r[k];
        p[k] = pv;
    }
    phat[k] = D[k] * pv;
}

// s = r - alpha v; shat = D s; partials of s . s
__global__ __launch_bounds__(ESB) void k_sp_s(int n3, const SpinCtl *__restrict__ c, const double *__restrict__ r,
                                              const double *__restrict__ v, const double *__restrict__ D,
                                              double *__restrict__ sv, double *__restrict__ shat, double *__restrict__ pss)
{
    __shared__ double sh[ESB];
    if (c->stop) return;
    const int k = blockIdx.x * ESB + threadIdx.x;
    double ss = 0.0;
    if (k < n3) {
        const double s = r[k] + (-c->alpha) * v[k];
        sv[k] = s;
        shat[k] = D[k] * s;
        ss = s * s;
    }
    block_sum(ss, sh, pss + blockIdx.x);
}

// x += alpha phat; x += omega shat; r = s - omega t; partials of r . r and rt . r. Once stopped:
// the last half step x += alpha phat if it is due.
__global__ __launch_bounds__(ESB) void k_sp_x(int n3, const SpinCtl *__restrict__ c, const double *__restrict__ phat,
                                              const double *__restrict__ shat, const double *__restrict__ sv,
                                              const double *__restrict__ tv, const double *__restrict__ rt,
                                              double *__restrict__ x, double *__restrict__ r, double *__restrict__ prr,
                                              double *__restrict__ prt)
{
    __shared__ double sh[ESB];
    const int k = blockIdx.x * ESB + threadIdx.x;
    if (c->stop) {
        if (c->fin == 1 && k < n3) x[k] += c->alpha * phat[k];
        return;
    }
    const double alpha = c->alpha, omega = c->omega;
    double rr = 0.0, rtr = 0.0;
    if (k < n3) {
        double xv = x[k];
        xv += alpha * phat[k];
        xv += omega * shat[k];
        x[k] = xv;
        const double rv = sv[k] - tv[k] * omega;
        r[k] = rv;
        rr = rv * rv;
        rtr = rt[k] * rv;
    }
    block_sum(rr, sh, prr + blockIdx.x);
    block_sum(rtr, sh, prt + blockIdx.x);
}

// x += xd (the Dirichlet values); zero everywhere when zero_failed is set and the solver did not
// converge (spinAcc::compute, src/spinAccumulationSolver.cpp:151-163)
__global__ __launch_bounds__(ESB) void k_sp_finish(int n3, const SpinCtl *__restrict__ c, int zero_failed,
                                                   const double *__restrict__ xd, double *__restrict__ x)
{
    const int k = blockIdx.x * ESB + threadIdx.x;
    if (k >= n3) return;
    x[k] = (zero_failed && c->status != FG_ITER_CONVERGED) ? 0.0 : x[k] + xd[k];
}

// The CSR pattern of the node graph of the tetrahedra, and for each entry the list of its
// tetrahedra's element-matrix slots ((i * 4 + j) * T + t) in mesh order.
struct NodePattern {
    std::vector<int64_t> rp, cptr, cslot;
    std::vector<int32_t> col;
};

NodePattern node_pattern(const fg_mesh *m)
{
    const int N = m->n_nodes, T = m->n_tets;
    std::vector<std::vector<std::pair<int32_t, int64_t>>> rows(N);
    for (int t = 0; t < T; t++) {
        const int32_t *q = m->tet_nodes + 4 * (size_t)t;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) rows[q[i]].push_back({q[j], (int64_t)(i * 4 + j) * T + t});
    }
    NodePattern P;
    P.rp.assign(N + 1, 0);
    P.cptr.assign(1, 0);
    P.cslot.reserve((size_t)16 * T);
    for (int i = 0; i < N; i++) {
        auto &r = rows[i];
        std::stable_sort(r.begin(), r.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        for (size_t k = 0; k < r.size(); k++) {
            if (k == 0 || r[k].first != r[k - 1].first) {
                if (k > 0) P.cptr.push_back((int64_t)P.cslot.size());
                P.col.push_back(r[k].first);
            }
            P.cslot.push_back(r[k].second);
        }
        if (!r.empty()) P.cptr.push_back((int64_t)P.cslot.size());
        P.rp[i + 1] = (int64_t)P.col.size();
        std::vector<std::pair<int32_t, int64_t>>().swap(r);
    }
    return P;
}

// tetrahedra nodes checked against the mesh; FG_OK or FG_ERR_MESH
int check_tets(const fg_mesh *m, std::vector<int4> &tn)
{
    tn.resize(m->n_tets);
    for (int t = 0; t < m->n_tets; t++) {
        const int32_t *q = m->tet_nodes + 4 * (size_t)t;
        for (int a = 0; a < 4; a++)
            if (q[a] < 0 || q[a] >= m->n_nodes) return FG_ERR_MESH;
        const int rg = m->tet_region[t];
        if (rg < 0 || rg >= m->n_vol_regions) return FG_ERR_MESH;
        tn[t] = make_int4(q[0], q[1], q[2], q[3]);
    }
    for (int64_t f = 0; f < 3 * (int64_t)m->n_tris; f++)
        if (m->tri_nodes[f] < 0 || m->tri_nodes[f] >= m->n_nodes) return FG_ERR_MESH;
    return FG_OK;
}

double tri_area(const fg_mesh *m, int f)
{
    const int32_t *q = m->tri_nodes + 3 * (size_t)f;
    const double *p0 = m->pos + 3 * (size_t)q[0], *p1 = m->pos + 3 * (size_t)q[1], *p2 = m->pos + 3 * (size_t)q[2];
    const d3 e1 = mk(p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]), e2 = mk(p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]);
    const d3 c = cross(e1, e2);
    return 0.5 * std::sqrt(dot(c, c));
}

// Host side of the potential's set-up, shared by both entry points: the arguments,
// electrostatSolver::checkBoundaryConditions (src/electrostatSolver.cpp:10-28: one surface region
// with a current density and one with a potential), the node pattern and the surface terms.
struct EsSetup {
    std::vector<int4> tn;
    std::vector<double> tsig, L, Vd;
    std::vector<uint8_t> dir;
    NodePattern P;
};

int es_setup(const fg_mesh *m, const double *sigma, const double *jn, const double *Vbc, double tol, int32_t max_iter,
             const double *V_out, EsSetup &S)
{
    if (!m || !sigma || !jn || !Vbc || !V_out || !m->pos || !m->tet_nodes || !m->tet_region || m->n_nodes <= 0 ||
        m->n_tets <= 0 || (m->n_tris > 0 && (!m->tri_nodes || !m->tri_region)) || !(tol > 0) || max_iter < 1)
        return FG_ERR_ARG;
    int nJ = 0, nV = 0;
    for (int k = 0; k < m->n_surf_regions; k++) {
        nJ += std::isfinite(jn[k]);
        nV += std::isfinite(Vbc[k]);
    }
    if (nJ != 1 || nV != 1) return FG_ERR_ARG;
    if (check_tets(m, S.tn) != FG_OK) return FG_ERR_MESH;
    const int N = m->n_nodes, T = m->n_tets;
    S.tsig.resize(T);
    for (int t = 0; t < T; t++) S.tsig[t] = sigma[m->tet_region[t]];
    // CSR pattern of the node graph; the entries of each (row, column) listed in mesh order
    S.P = node_pattern(m);
    if (S.P.rp[N] == 0) return FG_ERR_MESH;
    // L = - sum over the triangles of jn a_ig w_g (src/electrostatSolver.cpp:67-76, buildVect), and
    // the Dirichlet nodes with their potentials (:112-127)
    S.L.assign(N, 0.0);
    S.Vd.assign(N, 0.0);
    S.dir.assign(N, 0);
    for (int f = 0; f < m->n_tris; f++) {
        const int32_t *q = m->tri_nodes + 3 * (size_t)f;
        const int rg = m->tri_region[f];
        const bool has = rg >= 0 && rg < m->n_surf_regions;
        double J = has ? jn[rg] : NAN;
        if (!std::isfinite(J)) J = 0.0;
        const double surf = tri_area(m, f);
        double BE[3] = {0.0, 0.0, 0.0};
        for (int g = 0; g < NPI_TRI; g++) {
            const double Jw = J * (2.0 * surf * TRI_PDS[g]);
            for (int i = 0; i < 3; i++) BE[i] -= tri_a(i, g) * Jw;
        }
        for (int i = 0; i < 3; i++) S.L[q[i]] += BE[i];
        if (has && !std::isnan(Vbc[rg]))
            for (int i = 0; i < 3; i++) {
                S.Vd[q[i]] = Vbc[rg];
                S.dir[q[i]] = 1;
            }
    }
    return FG_OK;
}

// Device side: assembly and algebra::cg_dir on the null stream of the current device, with every
// buffer taken from B (all of them before the first kernel). Throws std::string on a HIP error.
template <class Bufs>
void es_solve(Bufs &B, const fg_mesh *m, const EsSetup &S, double tol, int32_t max_iter, double *V_out,
              fg_solve_info *info)
{
    const int N = m->n_nodes, T = m->n_tets;
    const NodePattern &P = S.P;
    const int64_t nnz = P.rp[N];
    const int nb = (N + ESB - 1) / ESB;
    int4 *d_tn = B.upload(S.tn);
    double *d_pos = B.upload(std::vector<double>(m->pos, m->pos + 3 * (size_t)N));
    double *d_sig = B.upload(S.tsig);
    double *d_ke = B.template alloc<double>((size_t)16 * T);
    int64_t *d_cptr = B.upload(P.cptr), *d_cslot = B.upload(P.cslot), *d_rp = B.upload(P.rp);
    int32_t *d_col = B.upload(P.col);
    double *d_val = B.template alloc<double>(nnz);
    uint8_t *d_dir = B.upload(S.dir);
    double *d_L = B.upload(S.L), *d_Vd = B.upload(S.Vd);
    double *d_D = B.template alloc<double>(N), *d_x = B.template alloc<double>(N), *d_r = B.template alloc<double>(N),
           *d_z = B.template alloc<double>(N), *d_p = B.template alloc<double>(N), *d_q = B.template alloc<double>(N);
    double *d_part = B.template alloc<double>((size_t)2 * nb);
    k_es_zero<<<nb, ESB>>>(N, d_x);
    k_es_element<<<(T + ESB - 1) / ESB, ESB>>>(T, d_tn, d_pos, d_sig, d_ke);
    ESCHK(hipGetLastError());
    k_es_gather<<<(unsigned)((nnz + ESB - 1) / ESB), ESB>>>(nnz, d_cptr, d_cslot, d_ke, d_val);
    k_es_diag<<<nb, ESB>>>(N, d_rp, d_col, d_val, d_dir, d_D);
    // b = L - K Vd (masked), r = b (the start is V = 0), z = D r, p = z
    k_es_spmv<<<nb, ESB>>>(N, d_rp, d_col, d_val, d_Vd, d_dir, 0, d_q, nullptr, nullptr);
    k_es_init<<<nb, ESB>>>(N, d_L, d_q, d_dir, d_D, d_r, d_z, d_p, d_part, d_part + nb);
    ESCHK(hipGetLastError());
    std::vector<double> hp(2 * (size_t)nb);
    auto read2 = [&](double &a, double &b) {
        ESCHK(hipMemcpy(hp.data(), d_part, hp.size() * sizeof(double), hipMemcpyDeviceToHost));
        a = host_sum(std::vector<double>(hp.begin(), hp.begin() + nb));
        b = host_sum(std::vector<double>(hp.begin() + nb, hp.end()));
    };
    double rr, zr;
    read2(rr, zr);
    // iteration<double> (src/algebra/iter.h:119-166)
    const double rhsn = std::sqrt(rr);
    double res = 0, rho = zr, rho_1 = 0;
    int nit = 0, status = FG_ITER_UNDEFINED;
    auto converged = [&](double nr) {
        res = std::fabs(nr);
        if (std::isnan(res)) {
            status = FG_ITER_CANNOT_CONVERGE;
            return false;
        }
        const bool cv = res <= rhsn * tol;
        if (cv) status = FG_ITER_CONVERGED;
        return cv;
    };
    while (!converged(std::sqrt(rr)) && status != FG_ITER_OVERFLOW && status != FG_ITER_CANNOT_CONVERGE) {
        if (nit > 0) {
            rho = zr;  // z = D r and z . r came with the last update
            k_es_dir<<<nb, ESB>>>(N, rho / rho_1, d_z, d_p);
        }
        k_es_spmv<<<nb, ESB>>>(N, d_rp, d_col, d_val, d_p, d_dir, 1, d_q, d_p, d_part);
        ESCHK(hipMemcpy(hp.data(), d_part, nb * sizeof(double), hipMemcpyDeviceToHost));
        const double qp = host_sum(std::vector<double>(hp.begin(), hp.begin() + nb));
        if (qp == 0.0) {
            status = FG_ITER_CANNOT_CONVERGE;
            break;
        }
        const double a = rho / qp;
        k_es_update<<<nb, ESB>>>(N, a, d_p, d_q, d_D, d_x, d_r, d_z, d_part, d_part + nb);
        read2(rr, zr);
        rho_1 = rho;
        if (++nit >= max_iter) status = FG_ITER_OVERFLOW;
    }
    k_es_add<<<nb, ESB>>>(N, d_Vd, d_x);
    ESCHK(hipGetLastError());
    ESCHK(hipMemcpy(V_out, d_x, (size_t)N * sizeof(double), hipMemcpyDeviceToHost));
    if (info) {
        *info = fg_solve_info{};
        info->status = status;
        info->iterations = nit;
        info->residual = res;
        info->rhs_norm = rhsn;
    }
}

// The scratch of fg_electrostatic_potential_noalloc: one copy per GPU, held until the process ends.
// Host code uses the device address hipGetSymbolAddress returns for the current device, never the
// array's name.
constexpr size_t ES_SCRATCH_BYTES = FG_ES_SCRATCH_BYTES;
static __device__ __align__(256) char scratch[ES_SCRATCH_BYTES];
constexpr int ES_MAX_DEVICES = 64;
std::mutex es_scratch_mu;
void *es_scratch_ptr[ES_MAX_DEVICES] = {};

// The scratch's address on `device`, taken once per device with that device current.
void *es_scratch(int device)
{
    std::lock_guard<std::mutex> lk(es_scratch_mu);
    if (device < 0 || device >= ES_MAX_DEVICES) throw std::string("device index out of range");
    if (!es_scratch_ptr[device]) {
        void *p = nullptr;
        ESCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        es_scratch_ptr[device] = p;
    }
    return es_scratch_ptr[device];
}
}  // namespace

extern "C" int fg_electrostatic_potential(const fg_mesh *m, const double *sigma, const double *jn, const double *Vbc,
                                          double tol, int32_t max_iter, int32_t device, double *V_out,
                                          fg_solve_info *info)
{
    EsSetup S;
    const int rc = es_setup(m, sigma, jn, Vbc, tol, max_iter, V_out, S);
    if (rc != FG_OK) return rc;
    try {
        ESCHK(hipSetDevice(device));
        DevBufs B;
        es_solve(B, m, S, tol, max_iter, V_out, info);
        return FG_OK;
    } catch (const std::string &) {
        return FG_ERR_DEVICE;
    }
}

extern "C" int fg_electrostatic_potential_noalloc(const fg_mesh *m, const double *sigma, const double *jn,
                                                  const double *Vbc, double tol, int32_t max_iter, int32_t device,
                                                  double *V_out, fg_solve_info *info)
{
    EsSetup S;
    const int rc = es_setup(m, sigma, jn, Vbc, tol, max_iter, V_out, S);
    if (rc != FG_OK) return rc;
    try {
        ESCHK(hipSetDevice(device));
        ArenaBufs B{(char *)es_scratch(device), ES_SCRATCH_BYTES};
        es_solve(B, m, S, tol, max_iter, V_out, info);
        return FG_OK;
    } catch (const ScratchFull &) {
        return FG_ERR_RANGE;
    } catch (const std::string &) {
        return FG_ERR_DEVICE;
    }
}

// ---- the persistent spin-diffusion solver ----

// spinAcc (src/spinAccumulationSolver.h): what does not change between solves, built once (the
// parameters, the node pattern, the Dirichlet values, the surface right-hand side, the device
// buffers), and the system K, L that the solves with accumulate set add into.
struct fg_spin {
    int device = 0, N = 0, T = 0, n3 = 0, nb = 0, nb3 = 0;
    int64_t nnz = 0;
    double tol = 0;
    int32_t max_iter = 0;
    int assemblies = 0;  // assemblies summed in K and L
    std::vector<fg_spin_volume> vol;
    hipStream_t stream = nullptr;
    DevBufs B;
    int4 *tn = nullptr;
    double4 *prm = nullptr;
    uint8_t *mag = nullptr, *dir = nullptr;
    int64_t *cptr = nullptr, *cslot = nullptr, *rp = nullptr, *nptr = nullptr, *nslot = nullptr;
    int32_t *col = nullptr;
    double *pos = nullptr, *V = nullptr, *u = nullptr, *us = nullptr, *ke = nullptr, *a = nullptr, *S = nullptr,
           *L = nullptr, *Ls = nullptr, *Vd = nullptr, *D = nullptr, *x = nullptr, *r = nullptr, *rt = nullptr,
           *p = nullptr, *ph = nullptr, *v = nullptr, *s = nullptr, *sh = nullptr, *t = nullptr, *part = nullptr;
    SpinCtl *ctl = nullptr, *h_ctl = nullptr;
    ~fg_spin()
    {
        if (h_ctl) (void)hipHostFree(h_ctl);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// spinAcc::spinAcc up to the first solve (src/spinAccumulationSolver.cpp:8-36,41-116): the
// boundary-condition check, the per-tetrahedron parameters, the node pattern, the Dirichlet values
// and the surface terms of solve (:173-197, the same in every call), then the device buffers.
int spin_setup(const fg_mesh *m, const fg_spin_volume *vol, const fg_spin_surface *surf, const double *Vn,
               double tol, int32_t max_iter, int32_t device, fg_spin &sp)
{
    if (!m || !vol || !surf || !Vn || !m->pos || !m->tet_nodes || !m->tet_region || !m->vol || m->n_nodes <= 0 ||
        m->n_tets <= 0 || (m->n_tris > 0 && (!m->tri_nodes || !m->tri_region)) || !(tol > 0) || max_iter < 1)
        return FG_ERR_ARG;
    // spinAcc::checkBoundaryConditions (src/spinAccumulationSolver.cpp:41-79)
    auto finite3 = [](const double *v) { return std::isfinite(std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])); };
    int nJ = 0, nS = 0;
    for (int k = 0; k < m->n_surf_regions; k++) {
        nJ += std::isfinite(surf[k].jn);
        nS += finite3(surf[k].s);
    }
    bool ok = nJ == 1 && nS == 1;
    for (int k = 0; k < m->n_vol_regions; k++)
        ok = ok && std::isfinite(vol[k].N0) && vol[k].N0 != 0 && std::isfinite(vol[k].P) && vol[k].P >= 0 && vol[k].P < 1;
    if (!ok) return FG_ERR_ARG;
    std::vector<int4> tn;
    if (check_tets(m, tn) != FG_OK) return FG_ERR_MESH;
    // physical constants (src/config.h.in:33-37)
    const double E = 1.602176634e-19, HBAR = 6.62607015e-34 / (2 * M_PI), ME = 9.1093837139e-31;
    const double MUB = E * HBAR / (2. * ME);
    const int N = m->n_nodes, T = m->n_tets, n3 = 3 * N;
    std::vector<double4> prm(T);
    std::vector<uint8_t> mag(T);
    for (int t = 0; t < T; t++) {
        const int rg = m->tet_region[t];
        const fg_spin_volume &p = vol[rg];
        const double D0 = 2.0 * p.sigma / ((E * E) * p.N0);  // spinAcc::getDiffusionCst
        prm[t] = make_double4(D0, D0 / (p.lsf * p.lsf), D0 / (p.lsd * p.lsd), MUB * p.P * p.sigma / E);
        mag[t] = m->vol[rg].Ms > 0;  // mesh::isMagnetic
    }
    const NodePattern P = node_pattern(m);
    if (P.rp[N] == 0) return FG_ERR_MESH;
    // each node's (local position, tetrahedron) slots in mesh order
    std::vector<int64_t> nptr(N + 1, 0), nslot((size_t)4 * T);
    for (int64_t k = 0; k < 4 * (int64_t)T; k++) nptr[m->tet_nodes[k] + 1]++;
    for (int i = 0; i < N; i++) nptr[i + 1] += nptr[i];
    {
        std::vector<int64_t> fill(nptr.begin(), nptr.end() - 1);
        for (int t = 0; t < T; t++)
            for (int l = 0; l < 4; l++) nslot[fill[m->tet_nodes[4 * (size_t)t + l]]++] = (int64_t)l * T + t;
    }
    // Dirichlet values (spinAcc::boundaryConditions, :91-116) and the surface right-hand side
    // (solve, :173-197), both in triangle order
    std::vector<double> Vd(n3, 0.0), Ls(n3, 0.0);
    std::vector<uint8_t> dir(N, 0);
    for (int f = 0; f < m->n_tris; f++) {
        const int rg = m->tri_region[f];
        if (rg < 0 || rg >= m->n_surf_regions) continue;
        const fg_spin_surface &p = surf[rg];
        const int32_t *q = m->tri_nodes + 3 * (size_t)f;
        double dv[3];
        bool d = false;
        if (std::isnan(std::sqrt(p.s[0] * p.s[0] + p.s[1] * p.s[1] + p.s[2] * p.s[2])) && std::isfinite(p.jn)) {
            for (int k = 0; k < 3; k++) dv[k] = -p.jn * (MUB / E) * p.uP[k];
            d = true;
        } else if (finite3(p.s) && std::isnan(p.jn)) {
            for (int k = 0; k < 3; k++) dv[k] = p.s[k];
            d = true;
        }
        if (d)
            for (int i = 0; i < 3; i++) {
                for (int k = 0; k < 3; k++) Vd[3 * (size_t)q[i] + k] = dv[k];
                dir[q[i]] = 1;
            }
        double sv[3] = {p.s[0], p.s[1], p.s[2]};
        if (std::isfinite(p.jn) && finite3(p.uP))
            for (int k = 0; k < 3; k++) sv[k] = -p.jn * (MUB / E) * p.uP[k];
        if (!finite3(sv)) continue;
        const double surfa = tri_area(m, f);
        double Le[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int g = 0; g < NPI_TRI; g++) {
            const double wg = 2.0 * surfa * TRI_PDS[g];
            for (int i = 0; i < 3; i++) {
                const double aiw = wg * tri_a(i, g);
                for (int k = 0; k < 3; k++) Le[3 * k + i] -= sv[k] * aiw;
            }
        }
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) Ls[3 * (size_t)q[i] + k] += Le[3 * k + i];
    }
    sp.device = device;
    sp.N = N;
    sp.T = T;
    sp.n3 = n3;
    sp.nb = (N + ESB - 1) / ESB;
    sp.nb3 = (n3 + ESB - 1) / ESB;
    sp.nnz = P.rp[N];
    sp.tol = tol;
    sp.max_iter = max_iter;
    sp.vol.assign(vol, vol + m->n_vol_regions);
    ESCHK(hipSetDevice(device));
    ESCHK(hipStreamCreateWithFlags(&sp.stream, hipStreamNonBlocking));
    DevBufs &B = sp.B;
    sp.tn = B.upload(tn);
    sp.pos = B.upload(std::vector<double>(m->pos, m->pos + 3 * (size_t)N));
    sp.prm = B.upload(prm);
    sp.mag = B.upload(mag);
    sp.dir = B.upload(dir);
    sp.V = B.upload(std::vector<double>(Vn, Vn + N));
    sp.u = B.alloc<double>(n3);
    sp.us = B.alloc<double>(n3);
    sp.ke = B.alloc<double>((size_t)40 * T);
    sp.cptr = B.upload(P.cptr);
    sp.cslot = B.upload(P.cslot);
    sp.rp = B.upload(P.rp);
    sp.col = B.upload(P.col);
    sp.nptr = B.upload(nptr);
    sp.nslot = B.upload(nslot);
    sp.a = B.alloc<double>(sp.nnz);
    sp.Vd = B.upload(Vd);
    sp.Ls = B.upload(Ls);
    for (double **q : {&sp.S, &sp.L, &sp.D, &sp.x, &sp.r, &sp.rt, &sp.p, &sp.ph, &sp.v, &sp.s, &sp.sh, &sp.t})
        *q = B.alloc<double>(n3);
    sp.part = B.alloc<double>((size_t)2 * sp.nb3);
    sp.ctl = B.alloc<SpinCtl>(1);
    ESCHK(hipHostMalloc((void **)&sp.h_ctl, sizeof(SpinCtl), hipHostMallocDefault));
    return FG_OK;
}

// spinAcc::solve (src/spinAccumulationSolver.cpp:165-209) at the magnetizations in sp.u (s-d
// exchange) and sp.us (source), on stream st: the assembly, added to the kept system when acc is
// set, then bicg_dir from s = 0 with its scalars on the device; s in sp.x.
void spin_run(fg_spin &sp, hipStream_t st, int acc, int zero_failed, fg_solve_info *info)
{
    const int T = sp.T, N = sp.N, n3 = sp.n3, nb3 = sp.nb3;
    acc = acc && sp.assemblies > 0;
    k_sp_reset<<<1, 64, 0, st>>>(sp.ctl, sp.tol, sp.max_iter);
    k_sa_element<<<(T + ESB - 1) / ESB, ESB, 0, st>>>(T, sp.tn, sp.pos, sp.prm, sp.mag, sp.V, sp.u, sp.us, sp.ke);
    k_sp_gather<<<(unsigned)((sp.nnz + ESB - 1) / ESB), ESB, 0, st>>>(sp.nnz, sp.cptr, sp.cslot, sp.ke, acc, sp.a);
    k_sa_node<<<sp.nb, ESB, 0, st>>>(N, T, sp.nptr, sp.nslot, sp.ke, sp.Ls, acc, sp.S, sp.L);
    sp.assemblies = acc ? sp.assemblies + 1 : 1;
    k_sa_diag<<<nb3, ESB, 0, st>>>(n3, sp.rp, sp.col, sp.a, sp.dir, sp.D);
    k_sa_spmv<<<nb3, ESB, 0, st>>>(n3, sp.rp, sp.col, sp.a, sp.S, sp.Vd, sp.dir, 0, sp.v, nullptr, nullptr, nullptr,
                                   nullptr);
    k_sa_init<<<nb3, ESB, 0, st>>>(n3, sp.L, sp.v, sp.dir, sp.r, sp.rt, sp.p, sp.x, sp.part);
    k_sp_ctl<<<1, ESB, 0, st>>>(nb3, sp.part, sp.ctl, SP_INIT);
    ESCHK(hipGetLastError());
    // chunks of iterations between host reads of the control block; bicg_dir ends by itself after
    // at most max_iter + 1 iterations (the overflow status breaks the one after the max_iter-th)
    constexpr int CHUNK = 16;
    const int max_chunks = sp.max_iter / CHUNK + 2;
    double *pv = sp.part, *pt = sp.part + nb3;
    for (int ch = 0;; ch++) {
        ESCHK(hipMemcpyAsync(sp.h_ctl, sp.ctl, sizeof(SpinCtl), hipMemcpyDeviceToHost, st));
        ESCHK(hipStreamSynchronize(st));
        if (sp.h_ctl->stop) break;
        if (ch >= max_chunks) throw std::string("spin solver: bicg_dir did not end");
        for (int it = 0; it < CHUNK; it++) {
            k_sp_p<<<nb3, ESB, 0, st>>>(n3, sp.ctl, sp.v, sp.r, sp.D, sp.p, sp.ph);
            k_sa_spmv<<<nb3, ESB, 0, st>>>(n3, sp.rp, sp.col, sp.a, sp.S, sp.ph, sp.dir, 1, sp.v, sp.rt, pv, nullptr,
                                           sp.ctl);
            k_sp_ctl<<<1, ESB, 0, st>>>(nb3, sp.part, sp.ctl, SP_ALPHA);
            k_sp_s<<<nb3, ESB, 0, st>>>(n3, sp.ctl, sp.r, sp.v, sp.D, sp.s, sp.sh, pv);
            k_sp_ctl<<<1, ESB, 0, st>>>(nb3, sp.part, sp.ctl, SP_S);
            k_sa_spmv<<<nb3, ESB, 0, st>>>(n3, sp.rp, sp.col, sp.a, sp.S, sp.sh, sp.dir, 1, sp.t, sp.s, pv, pt, sp.ctl);
            k_sp_ctl<<<1, ESB, 0, st>>>(nb3, sp.part, sp.ctl, SP_OMEGA);
            k_sp_x<<<nb3, ESB, 0, st>>>(n3, sp.ctl, sp.ph, sp.sh, sp.s, sp.t, sp.rt, sp.x, sp.r, pv, pt);
            k_sp_ctl<<<1, ESB, 0, st>>>(nb3, sp.part, sp.ctl, SP_LOOP);
        }
        ESCHK(hipGetLastError());
    }
    k_sp_finish<<<nb3, ESB, 0, st>>>(n3, sp.ctl, zero_failed, sp.Vd, sp.x);
    ESCHK(hipGetLastError());
    if (info) {
        const SpinCtl &h = *sp.h_ctl;
        *info = fg_solve_info{};
        info->status = h.status;
        info->iterations = h.nit;
        info->residual = h.res;
        info->rhs_norm = h.rhsn;
    }
}

}  // namespace

namespace fg {

const fg_spin_volume *spin_volumes(const fg_spin *sp, int *n_regions, int *n_nodes, int *device)
{
    *n_regions = (int)sp->vol.size();
    *n_nodes = sp->N;
    *device = sp->device;
    return sp->vol.data();
}

int spin_solve_nodes(fg_spin *sp, const NodeRec *cur, const NodeRec *nxt, hipStream_t st, int accumulate,
                     double *s_dev, fg_solve_info *info)
{
    try {
        ESCHK(hipSetDevice(sp->device));
        k_sp_rec_u<<<sp->nb, ESB, 0, st>>>(sp->N, cur, nxt, sp->u, sp->us);
        spin_run(*sp, st, accumulate, 1, info);
        ESCHK(hipMemcpyAsync(s_dev, sp->x, (size_t)sp->n3 * sizeof(double), hipMemcpyDeviceToDevice, st));
        return FG_OK;
    } catch (const std::string &) {
        return FG_ERR_DEVICE;
    }
}

}  // namespace fg

extern "C" int fg_spin_create(const fg_mesh *m, const fg_spin_volume *vol, const fg_spin_surface *surf,
                              const double *Vn, double tol, int32_t max_iter, int32_t device, fg_spin **out)
{
    if (!out) return FG_ERR_ARG;
    *out = nullptr;
    fg_spin *sp = new fg_spin;
    try {
        const int rc = spin_setup(m, vol, surf, Vn, tol, max_iter, device, *sp);
        if (rc != FG_OK) {
            delete sp;
            return rc;
        }
    } catch (const std::string &) {
        delete sp;
        return FG_ERR_DEVICE;
    }
    *out = sp;
    return FG_OK;
}

extern "C" int fg_spin_solve(fg_spin *sp, const double *u, const double *u_src, int32_t accumulate, double *s_out,
                             fg_solve_info *info)
{
    if (!sp || !u || !s_out) return FG_ERR_ARG;
    try {
        ESCHK(hipSetDevice(sp->device));
        const size_t bytes = (size_t)sp->n3 * sizeof(double);
        ESCHK(hipMemcpyAsync(sp->u, u, bytes, hipMemcpyHostToDevice, sp->stream));
        ESCHK(hipMemcpyAsync(sp->us, u_src ? u_src : u, bytes, hipMemcpyHostToDevice, sp->stream));
        spin_run(*sp, sp->stream, accumulate, 0, info);
        ESCHK(hipMemcpyAsync(s_out, sp->x, bytes, hipMemcpyDeviceToHost, sp->stream));
        ESCHK(hipStreamSynchronize(sp->stream));
        return FG_OK;
    } catch (const std::string &) {
        return FG_ERR_DEVICE;
    }
}

extern "C" int fg_spin_destroy(fg_spin *sp)
{
    if (!sp) return FG_ERR_ARG;
    (void)hipSetDevice(sp->device);
    (void)hipStreamSynchronize(sp->stream);
    delete sp;
    return FG_OK;
}

extern "C" int fg_spin_accumulation(const fg_mesh *m, const fg_spin_volume *vol, const fg_spin_surface *surf,
                                    const double *Vn, const double *u, const double *u_src, double tol,
                                    int32_t max_iter, int32_t device, double *s_out, fg_solve_info *info)
{
    if (!u || !s_out) return FG_ERR_ARG;
    fg_spin *sp = nullptr;
    int rc = fg_spin_create(m, vol, surf, Vn, tol, max_iter, device, &sp);
    if (rc != FG_OK) return rc;
    rc = fg_spin_solve(sp, u, u_src, 0, s_out, info);
    fg_spin_destroy(sp);
    return rc;
}
