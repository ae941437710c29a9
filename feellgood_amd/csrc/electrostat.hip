// electrostat.hip -- the electrostatic potential of the current-injection problem (FeeLLGood's
// electrostatSolver, src/electrostatSolver.cpp:10-136): the P1 conductivity matrix assembled on the
// device, the current-density boundary term on the host (surface triangles only), and conjugate
// gradient with the diagonal preconditioner and Dirichlet masking (algebra::cg_dir,
// src/algebra/cg.h:67-121) on the device, the scalars of each iteration read back by the host.
//
// The potential is solved once per run (the reference builds it in the spin-accumulation set-up,
// src/spinAccumulationSolver.cpp:21-30), so the kernels favour determinism over speed: every
// matrix entry is summed from its tetrahedra in mesh order, and every dot product is a fixed tree
// in each block with the block partials summed in order on the host.
#include <algorithm>
#include <cmath>
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

// K_e[i][j] = sigma sum_g w_g da_i . da_j (electrostatSolver::integrales, src/electrostatSolver.cpp:42-65),
// computed with the nodes reordered for a positive Jacobian (Tetra::Tet, src/tetra.h:280-303) and
// stored at the caller's local positions, [16][T].
__global__ __launch_bounds__(ESB) void k_es_element(int T, const int4 *__restrict__ tn, const double *__restrict__ pos,
                                                    const double *__restrict__ tsig, double *__restrict__ ke)
{
    const int t = blockIdx.x * ESB + threadIdx.x;
    if (t >= T) return;
    const int4 q = tn[t];
    const int id[4] = {q.x, q.y, q.z, q.w};
    d3 P[4];
#pragma unroll
    for (int a = 0; a < 4; a++) P[a] = mk(pos[3 * (size_t)id[a]], pos[3 * (size_t)id[a] + 1], pos[3 * (size_t)id[a] + 2]);
    d3 e1 = P[1] - P[0], e2 = P[2] - P[0], e3 = P[3] - P[0];
    int perm[4] = {0, 1, 2, 3};
    if (dot(e1, cross(e2, e3)) < 0) {
        const d3 s = e2;
        e2 = e3;
        e3 = s;
        perm[2] = 3;
        perm[3] = 2;
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
    double da[4][3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double j0 = cf[k][0] * inv, j1 = cf[k][1] * inv, j2 = cf[k][2] * inv;  // column k of J^-1
        da[0][k] = -j0 - j1 - j2;
        da[1][k] = j0;
        da[2][k] = j1;
        da[3][k] = j2;
    }
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

double host_sum(const std::vector<double> &h)
{
    double s = 0.0;
    for (double v : h) s += v;
    return s;
}

}  // namespace

extern "C" int fg_electrostatic_potential(const fg_mesh *m, const double *sigma, const double *jn, const double *Vbc,
                                          double tol, int32_t max_iter, int32_t device, double *V_out,
                                          fg_solve_info *info)
{
    if (!m || !sigma || !jn || !Vbc || !V_out || !m->pos || !m->tet_nodes || !m->tet_region || m->n_nodes <= 0 ||
        m->n_tets <= 0 || (m->n_tris > 0 && (!m->tri_nodes || !m->tri_region)) || !(tol > 0) || max_iter < 1)
        return FG_ERR_ARG;
    // electrostatSolver::checkBoundaryConditions (src/electrostatSolver.cpp:10-28): one surface
    // region with a current density and one with a potential
    int nJ = 0, nV = 0;
    for (int k = 0; k < m->n_surf_regions; k++) {
        nJ += std::isfinite(jn[k]);
        nV += std::isfinite(Vbc[k]);
    }
    if (nJ != 1 || nV != 1) return FG_ERR_ARG;
    const int N = m->n_nodes, T = m->n_tets;
    std::vector<int4> tn(T);
    std::vector<double> tsig(T);
    for (int t = 0; t < T; t++) {
        const int32_t *q = m->tet_nodes + 4 * (size_t)t;
        for (int a = 0; a < 4; a++)
            if (q[a] < 0 || q[a] >= N) return FG_ERR_MESH;
        const int rg = m->tet_region[t];
        if (rg < 0 || rg >= m->n_vol_regions) return FG_ERR_MESH;
        tn[t] = make_int4(q[0], q[1], q[2], q[3]);
        tsig[t] = sigma[rg];
    }
    // CSR pattern of the node graph; the entries of each (row, column) listed in mesh order
    std::vector<std::vector<std::pair<int32_t, int64_t>>> rows(N);
    for (int t = 0; t < T; t++) {
        const int32_t *q = m->tet_nodes + 4 * (size_t)t;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) rows[q[i]].push_back({q[j], (int64_t)(i * 4 + j) * T + t});
    }
    std::vector<int64_t> rp(N + 1, 0), cptr(1, 0), cslot;
    std::vector<int32_t> col;
    cslot.reserve((size_t)16 * T);
    for (int i = 0; i < N; i++) {
        auto &r = rows[i];
        std::stable_sort(r.begin(), r.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        for (size_t k = 0; k < r.size(); k++) {
            if (k == 0 || r[k].first != r[k - 1].first) {
                if (k > 0) cptr.push_back((int64_t)cslot.size());
                col.push_back(r[k].first);
            }
            cslot.push_back(r[k].second);
        }
        if (!r.empty()) cptr.push_back((int64_t)cslot.size());
        rp[i + 1] = (int64_t)col.size();
        std::vector<std::pair<int32_t, int64_t>>().swap(r);
    }
    if (rp[N] == 0) return FG_ERR_MESH;
    // L = - sum over the triangles of jn a_ig w_g (src/electrostatSolver.cpp:67-76, buildVect), and
    // the Dirichlet nodes with their potentials (:112-127)
    std::vector<double> L(N, 0.0), Vd(N, 0.0);
    std::vector<uint8_t> dir(N, 0);
    for (int f = 0; f < m->n_tris; f++) {
        const int32_t *q = m->tri_nodes + 3 * (size_t)f;
        for (int a = 0; a < 3; a++)
            if (q[a] < 0 || q[a] >= N) return FG_ERR_MESH;
        const int rg = m->tri_region[f];
        const bool has = rg >= 0 && rg < m->n_surf_regions;
        double J = has ? jn[rg] : NAN;
        if (!std::isfinite(J)) J = 0.0;
        const double *p0 = m->pos + 3 * (size_t)q[0], *p1 = m->pos + 3 * (size_t)q[1], *p2 = m->pos + 3 * (size_t)q[2];
        const d3 e1 = mk(p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]), e2 = mk(p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]);
        const d3 c = cross(e1, e2);
        const double surf = 0.5 * std::sqrt(dot(c, c));
        double BE[3] = {0.0, 0.0, 0.0};
        for (int g = 0; g < NPI_TRI; g++) {
            const double Jw = J * (2.0 * surf * TRI_PDS[g]);
            for (int i = 0; i < 3; i++) BE[i] -= tri_a(i, g) * Jw;
        }
        for (int i = 0; i < 3; i++) L[q[i]] += BE[i];
        if (has && !std::isnan(Vbc[rg]))
            for (int i = 0; i < 3; i++) {
                Vd[q[i]] = Vbc[rg];
                dir[q[i]] = 1;
            }
    }
    try {
        ESCHK(hipSetDevice(device));
        DevBufs B;
        const int64_t nnz = rp[N];
        const int nb = (N + ESB - 1) / ESB;
        int4 *d_tn = B.upload(tn);
        double *d_pos = B.upload(std::vector<double>(m->pos, m->pos + 3 * (size_t)N));
        double *d_sig = B.upload(tsig);
        double *d_ke = B.alloc<double>((size_t)16 * T);
        int64_t *d_cptr = B.upload(cptr), *d_cslot = B.upload(cslot), *d_rp = B.upload(rp);
        int32_t *d_col = B.upload(col);
        double *d_val = B.alloc<double>(nnz);
        uint8_t *d_dir = B.upload(dir);
        double *d_L = B.upload(L), *d_Vd = B.upload(Vd);
        double *d_D = B.alloc<double>(N), *d_x = B.alloc<double>(N), *d_r = B.alloc<double>(N),
               *d_z = B.alloc<double>(N), *d_p = B.alloc<double>(N), *d_q = B.alloc<double>(N);
        double *d_part = B.alloc<double>((size_t)2 * nb);
        ESCHK(hipMemset(d_x, 0, (size_t)N * sizeof(double)));
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
        return FG_OK;
    } catch (const std::string &) {
        return FG_ERR_DEVICE;
    }
}
