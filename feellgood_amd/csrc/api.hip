// api.hip -- the C-ABI of include/fgmi.h: plan construction on the host, device buffers, and the
// per-step call sequence of FeeLLGood's time loop (src/time_integration.cpp:166-245).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <utility>
#include <vector>

#include "ctx.h"

using namespace fg;

struct fg_ctx {
    Ctx c;
    // last call parameters
    double r = 0;
    bool basis_valid = false;
    bool elements_ready = false;
    double dt = 0, prefactor = 1;
    double Hext[3] = {0, 0, 0};
    int last_nit = 8;
};

namespace {

thread_local std::string g_create_err;

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            throw std::string(#expr) + ": " + hipGetErrorString(e_);                       \
        }                                                                                  \
    } while (0)

template <class T>
T *dalloc(size_t n)
{
    void *p = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc(&p, n * sizeof(T)));
    return (T *)p;
}

template <class T>
void upload(T *d, const std::vector<T> &h, hipStream_t s)
{
    if (!h.empty()) HIPCHK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

struct P3 {
    double x, y, z;
};
inline P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 crs(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dt3(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Plan construction: geometry of every magnetic element, the stiffness stencil, the gather lists.
void build_plan(Ctx &c, const fg_mesh *m)
{
    const int N = m->n_nodes;
    c.N = N;
    std::vector<P3> pos(N);
    for (int i = 0; i < N; i++) pos[i] = {m->pos[3 * i], m->pos[3 * i + 1], m->pos[3 * i + 2]};
    // regions
    c.n_regions = m->n_vol_regions;
    std::vector<DevRegion> regs(std::max(1, m->n_vol_regions));
    for (int k = 0; k < m->n_vol_regions; k++) {
        const fg_volume_region &v = m->vol[k];
        DevRegion &R = regs[k];
        std::memset(&R, 0, sizeof(R));
        R.alpha = v.alpha;
        R.Ms = v.Ms;
        R.Ae = v.Ae;
        R.K = v.K;
        R.K3 = v.K3;
        if (v.Ms > 0) {
            R.Abis = 2.0 * v.Ae / (MU0 * v.Ms);
            R.Kbis = 2.0 * v.K / (MU0 * v.Ms);
            R.K3bis = 2.0 * v.K3 / (MU0 * v.Ms);
        }
        for (int d = 0; d < 3; d++) {
            R.uk[d] = v.uk[d];
            R.ex[d] = v.ex[d];
            R.ey[d] = v.ey[d];
            R.ez[d] = v.ez[d];
        }
        R.hasK = v.K != 0;
        R.hasK3 = v.K3 != 0;
    }
    // magnetic tetrahedra, oriented, with geometry (Tet ctor, src/tetra.h:280-303)
    std::vector<int32_t> mtets;  // 4 per magnetic tet
    std::vector<int32_t> mreg;
    std::vector<double> da_w;    // 17 per magnetic tet, AoS during construction
    std::vector<char> magn(N, 0), tet_is_mag(m->n_tets, 0), tri_is_mag(std::max(0, m->n_tris), 0);
    std::vector<std::pair<int, int>> edges;
    edges.reserve((size_t)6 * m->n_tets);
    for (int t = 0; t < m->n_tets; t++) {
        int q[4];
        for (int a = 0; a < 4; a++) {
            q[a] = m->tet_nodes[4 * t + a];
            if (q[a] < 0 || q[a] >= N) throw std::string("tetrahedron node index out of range");
        }
        for (int a = 0; a < 3; a++)
            for (int b = a + 1; b < 4; b++) edges.push_back(std::minmax(q[a], q[b]));
        const int rg = m->tet_region[t];
        if (rg < 0 || rg >= m->n_vol_regions) throw std::string("tetrahedron region out of range");
        if (!(m->vol[rg].Ms > 0)) continue;
        P3 e1 = sub(pos[q[1]], pos[q[0]]), e2 = sub(pos[q[2]], pos[q[0]]), e3 = sub(pos[q[3]], pos[q[0]]);
        double mixed = dt3(e1, crs(e2, e3));
        if (std::fabs(mixed) < 1e-40) throw std::string("singular tetrahedron ") + std::to_string(t);
        if (mixed < 0) {
            std::swap(q[2], q[3]);
            e2 = sub(pos[q[2]], pos[q[0]]);
            e3 = sub(pos[q[3]], pos[q[0]]);
        }
        // J = [e1 e2 e3] (columns); grad lambda_{1..3} = rows of J^-1, grad lambda_0 = -sum
        const double J[3][3] = {{e1.x, e2.x, e3.x}, {e1.y, e2.y, e3.y}, {e1.z, e2.z, e3.z}};
        const double detJ = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                            J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                            J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        auto cof = [&J](int i, int j) {
            int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            return J[i1][j1] * J[i2][j2] - J[i1][j2] * J[i2][j1];
        };
        const double inv = 1.0 / (cof(0, 0) * J[0][0] + cof(1, 0) * J[1][0] + cof(2, 0) * J[2][0]);
        double Ji[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) Ji[j][i] = cof(i, j) * inv;
        double g[17];
        for (int k = 0; k < 3; k++) {
            g[k] = -Ji[0][k] - Ji[1][k] - Ji[2][k];
            g[3 + k] = Ji[0][k];
            g[6 + k] = Ji[1][k];
            g[9 + k] = Ji[2][k];
        }
        for (int p = 0; p < 5; p++) g[12 + p] = detJ * TET_PDS[p];
        mtets.insert(mtets.end(), q, q + 4);
        tet_is_mag[t] = 1;
        c.h_tet_mesh.push_back(t);
        mreg.push_back(rg);
        da_w.insert(da_w.end(), g, g + 17);
        for (int a = 0; a < 4; a++) magn[q[a]] = 1;
    }
    const int Tm = (int)mreg.size();
    c.Tm = Tm;
    c.h_reg_Ms.resize(std::max(1, m->n_vol_regions));
    for (int k = 0; k < m->n_vol_regions; k++) c.h_reg_Ms[k] = m->vol[k].Ms;
    if (Tm == 0) throw std::string("no magnetic tetrahedron");
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    c.h_mag.assign(N, 0);
    for (int i = 0; i < N; i++) c.h_mag[i] = magn[i];

    // stiffness stencil pattern: diagonal of every node + edges between magnetic nodes
    // (solver<2>::build_shape, src/solver.h:285-314), values S'_ij = sum_t Abis_t sum_g w_g da_a.da_b
    std::vector<int32_t> deg(N, 1);
    for (auto &e : edges)
        if (magn[e.first] && magn[e.second]) {
            deg[e.first]++;
            deg[e.second]++;
        }
    c.h_rowptr.assign(N + 1, 0);
    for (int i = 0; i < N; i++) c.h_rowptr[i + 1] = c.h_rowptr[i] + deg[i];
    c.h_col.assign(c.h_rowptr[N], 0);
    {
        std::vector<int64_t> fill(c.h_rowptr.begin(), c.h_rowptr.end() - 1);
        for (int i = 0; i < N; i++) c.h_col[fill[i]++] = i;
        for (auto &e : edges)
            if (magn[e.first] && magn[e.second]) {
                c.h_col[fill[e.first]++] = e.second;
                c.h_col[fill[e.second]++] = e.first;
            }
        for (int i = 0; i < N; i++) std::sort(c.h_col.begin() + c.h_rowptr[i], c.h_col.begin() + c.h_rowptr[i + 1]);
    }
    c.nnz_real = c.h_rowptr[N];
    c.h_val.assign(c.nnz_real, 0.0);
    std::vector<double> mass(N, 0.0);
    c.totalMagVol = 0;
    for (int t = 0; t < Tm; t++) {
        const int32_t *q = &mtets[4 * t];
        const double *g = &da_w[17 * (size_t)t];
        double wsum = 0;
        for (int p = 0; p < 5; p++) wsum += g[12 + p];
        c.totalMagVol += wsum;
        const double Abis = regs[mreg[t]].Abis;
        for (int a = 0; a < 4; a++) {
            double ma = 0;
            for (int p = 0; p < 5; p++) ma += tet_a(a, p) * g[12 + p];
            mass[q[a]] += ma;
            for (int b = 0; b < 4; b++) {
                const double s = g[3 * a] * g[3 * b] + g[3 * a + 1] * g[3 * b + 1] + g[3 * a + 2] * g[3 * b + 2];
                auto bg = c.h_col.begin() + c.h_rowptr[q[a]], en = c.h_col.begin() + c.h_rowptr[q[a] + 1];
                auto it = std::lower_bound(bg, en, q[b]);
                c.h_val[it - c.h_col.begin()] += Abis * wsum * s;
            }
        }
    }
    std::vector<double> sdiag(N, 0.0);
    for (int i = 0; i < N; i++)
        for (int64_t k = c.h_rowptr[i]; k < c.h_rowptr[i + 1]; k++)
            if (c.h_col[k] == i) sdiag[i] = c.h_val[k];
    // SELL-64
    c.nslices = (N + SELL_C - 1) / SELL_C;
    std::vector<int64_t> sptr(c.nslices + 1, 0);
    std::vector<int32_t> swid(c.nslices, 0);
    for (int sl = 0; sl < c.nslices; sl++) {
        int wmax = 0;
        for (int i = sl * SELL_C; i < std::min(N, (sl + 1) * SELL_C); i++) wmax = std::max(wmax, deg[i]);
        swid[sl] = wmax;
        sptr[sl + 1] = sptr[sl] + (int64_t)wmax * SELL_C;
    }
    c.nnzS = sptr[c.nslices];
    std::vector<int32_t> scol(c.nnzS, 0);
    std::vector<double> sval(c.nnzS, 0.0);
    for (int sl = 0; sl < c.nslices; sl++)
        for (int lane = 0; lane < SELL_C; lane++) {
            const int i = sl * SELL_C + lane;
            const int row = i < N ? i : N - 1;
            for (int k = 0; k < swid[sl]; k++) {
                const int64_t e = sptr[sl] + (int64_t)k * SELL_C + lane;
                if (i < N && k < deg[i]) {
                    scol[e] = c.h_col[c.h_rowptr[i] + k];
                    sval[e] = c.h_val[c.h_rowptr[i] + k];
                } else {
                    scol[e] = row;
                    sval[e] = 0.0;
                }
            }
        }
    // node -> element output gather list (deterministic: ascending tet index)
    std::vector<int32_t> incp(N + 1, 0);
    for (int t = 0; t < Tm; t++)
        for (int a = 0; a < 4; a++) incp[mtets[4 * t + a] + 1]++;

    // surface triangles: charged (magTri) and anisotropic (Ks != 0, in magTri)
    std::vector<int> magtri, kstri;
    for (int f = 0; f < m->n_tris; f++) {
        const int32_t *q = &m->tri_nodes[3 * f];
        for (int a = 0; a < 3; a++)
            if (q[a] < 0 || q[a] >= N) throw std::string("triangle node index out of range");
        const bool mag = magn[q[0]] && magn[q[1]] && magn[q[2]];
        const int rg = m->tri_region[f];
        const bool supp = (m->n_surf_regions > 0 && rg >= 0 && rg < m->n_surf_regions) && m->surf[rg].suppress_charges;
        if (mag && !supp) {
            magtri.push_back(f);
            tri_is_mag[f] = 1;
            if (m->n_surf_regions > 0 && rg >= 0 && rg < m->n_surf_regions && m->surf[rg].Ks != 0) kstri.push_back(f);
        }
    }
    c.Fm = (int)magtri.size();
    c.Fk = (int)kstri.size();
    // instance ranges: the input ranges (one instance unless a batch) mapped to magnetic elements
    if (c.in_node_off.empty()) {
        c.in_node_off = {0, N};
        c.in_tet_off = {0, m->n_tets};
        c.in_tri_off = {0, std::max(0, m->n_tris)};
    }
    {
        const size_t K = c.in_node_off.size() - 1;
        c.inst_node_off = c.in_node_off;
        c.inst_mtet_off.assign(K + 1, 0);
        c.inst_mtri_off.assign(K + 1, 0);
        for (size_t k = 0; k < K; k++) {
            int64_t nt = 0, nf = 0;
            for (int64_t t = c.in_tet_off[k]; t < c.in_tet_off[k + 1]; t++) nt += tet_is_mag[t];
            for (int64_t f = c.in_tri_off[k]; f < c.in_tri_off[k + 1]; f++) nf += tri_is_mag[f];
            c.inst_mtet_off[k + 1] = c.inst_mtet_off[k] + nt;
            c.inst_mtri_off[k + 1] = c.inst_mtri_off[k] + nf;
        }
    }
    for (int f : kstri)
        for (int a = 0; a < 3; a++) incp[m->tri_nodes[3 * f + a] + 1]++;
    for (int i = 0; i < N; i++) incp[i + 1] += incp[i];
    std::vector<int32_t> inci(incp[N]);
    {
        std::vector<int32_t> fill(incp.begin(), incp.end() - 1);
        for (int t = 0; t < Tm; t++)
            for (int a = 0; a < 4; a++) {
                const int i = mtets[4 * t + a];
                inci[fill[i]++] = a * Tm + t;
            }
        for (int k = 0; k < c.Fk; k++)
            for (int a = 0; a < 3; a++) {
                const int i = m->tri_nodes[3 * kstri[k] + a];
                inci[fill[i]++] = 4 * Tm + a * c.Fk + k;
            }
    }
    auto tri_geom = [&](int f, P3 &n, double &surf) {
        const int32_t *q = &m->tri_nodes[3 * f];
        P3 nv = crs(sub(pos[q[1]], pos[q[0]]), sub(pos[q[2]], pos[q[0]]));
        double len = std::sqrt(dt3(nv, nv));
        surf = 0.5 * len;
        n = len > 0 ? P3{nv.x / len, nv.y / len, nv.z / len} : nv;
    };
    const int Fm = c.Fm, Fk = c.Fk;
    std::vector<int4> trin(std::max(1, Fm)), trikn(std::max(1, Fk));
    std::vector<double> tgeo((size_t)16 * std::max(1, Fm), 0.0), tkgeo((size_t)9 * std::max(1, Fk), 0.0);
    for (int k = 0; k < Fm; k++) {
        const int f = magtri[k];
        const int32_t *q = &m->tri_nodes[3 * f];
        trin[k] = make_int4(q[0], q[1], q[2], 0);
        P3 n;
        double surf;
        tri_geom(f, n, surf);
        tgeo[0 * (size_t)Fm + k] = n.x;
        tgeo[1 * (size_t)Fm + k] = n.y;
        tgeo[2 * (size_t)Fm + k] = n.z;
        for (int g = 0; g < 4; g++) tgeo[(3 + g) * (size_t)Fm + k] = 2.0 * surf * TRI_PDS[g];
        tgeo[7 * (size_t)Fm + k] = m->tri_dMs[f];
        tgeo[8 * (size_t)Fm + k] = surf;
        const int rg = m->tri_region[f];
        const bool has = m->n_surf_regions > 0 && rg >= 0 && rg < m->n_surf_regions;
        tgeo[12 * (size_t)Fm + k] = has ? m->surf[rg].Ks : 0.0;
        for (int d = 0; d < 3; d++) tgeo[(13 + d) * (size_t)Fm + k] = has ? m->surf[rg].uk[d] : 0.0;
    }
    for (int k = 0; k < Fk; k++) {
        const int f = kstri[k];
        const int32_t *q = &m->tri_nodes[3 * f];
        trikn[k] = make_int4(q[0], q[1], q[2], 0);
        P3 n;
        double surf;
        tri_geom(f, n, surf);
        for (int g = 0; g < 4; g++) tkgeo[g * (size_t)Fk + k] = 2.0 * surf * TRI_PDS[g];
        const int rg = m->tri_region[f];
        tkgeo[4 * (size_t)Fk + k] = 2.0 * m->surf[rg].Ks / m->tri_dMs[f];
        for (int d = 0; d < 3; d++) tkgeo[(5 + d) * (size_t)Fk + k] = m->surf[rg].uk[d];
    }
    // node -> triangle-vertex corrections
    std::vector<int32_t> cincp(N + 1, 0);
    for (int k = 0; k < Fm; k++)
        for (int a = 0; a < 3; a++) cincp[m->tri_nodes[3 * magtri[k] + a] + 1]++;
    for (int i = 0; i < N; i++) cincp[i + 1] += cincp[i];
    std::vector<int32_t> cinci(std::max(1, cincp[N]));
    {
        std::vector<int32_t> fill(cincp.begin(), cincp.end() - 1);
        for (int k = 0; k < Fm; k++)
            for (int a = 0; a < 3; a++) cinci[fill[m->tri_nodes[3 * magtri[k] + a]]++] = a * Fm + k;
    }
    // demag sources: Gauss points of magnetic tets then charged triangles (fmm ctor,
    // src/fmm_demag.h:85-86)
    c.S = (int64_t)5 * Tm + (int64_t)4 * Fm;
    std::vector<double4> spos(c.S);
    for (int t = 0; t < Tm; t++) {
        const int32_t *q = &mtets[4 * t];
        for (int g = 0; g < 5; g++) {
            double x = 0, y = 0, z = 0;
            for (int a = 0; a < 4; a++) {
                x += tet_a(a, g) * pos[q[a]].x;
                y += tet_a(a, g) * pos[q[a]].y;
                z += tet_a(a, g) * pos[q[a]].z;
            }
            spos[(size_t)5 * t + g] = make_double4(x, y, z, 0.0);
        }
    }
    for (int k = 0; k < Fm; k++) {
        const int32_t *q = &m->tri_nodes[3 * magtri[k]];
        for (int g = 0; g < 4; g++) {
            double x = 0, y = 0, z = 0;
            for (int a = 0; a < 3; a++) {
                x += tri_a(a, g) * pos[q[a]].x;
                y += tri_a(a, g) * pos[q[a]].y;
                z += tri_a(a, g) * pos[q[a]].z;
            }
            spos[(size_t)5 * Tm + (size_t)4 * k + g] = make_double4(x, y, z, 0.0);
        }
    }
    // ---- device allocation and upload ----
    hipStream_t st = c.stream;
    c.rec[0] = dalloc<NodeRec>(N);
    c.rec[1] = dalloc<NodeRec>(N);
    HIPCHK(hipMemsetAsync(c.rec[0], 0, sizeof(NodeRec) * N, st));
    HIPCHK(hipMemsetAsync(c.rec[1], 0, sizeof(NodeRec) * N, st));
    std::vector<double4> p4(N);
    for (int i = 0; i < N; i++) p4[i] = make_double4(pos[i].x, pos[i].y, pos[i].z, 0.0);
    c.pos4 = dalloc<double4>(N);
    upload(c.pos4, p4, st);
    c.ep4 = dalloc<double4>(N);
    c.eq4 = dalloc<double4>(N);
    c.mag = dalloc<int32_t>(N);
    upload(c.mag, c.h_mag, st);
    c.mass = dalloc<double>(N);
    upload(c.mass, mass, st);
    c.sdiag = dalloc<double>(N);
    upload(c.sdiag, sdiag, st);
    c.Dg = dalloc<double4>(N);
    c.dprec = dalloc<double2>(N);
    double2 **vecs[] = {&c.b, &c.x, &c.r, &c.rt, &c.p, &c.ph, &c.v, &c.s, &c.sh, &c.t};
    for (auto v : vecs) {
        *v = dalloc<double2>(N);
        HIPCHK(hipMemsetAsync(*v, 0, sizeof(double2) * N, st));
    }
    c.w = dalloc<double4>(N);
    c.ctl = dalloc<KrylovCtl>(1);
    HIPCHK(hipHostMalloc((void **)&c.h_ctl, sizeof(KrylovCtl), hipHostMallocDefault));
    std::memset(c.h_ctl, 0, sizeof(KrylovCtl));
    c.h_ctl->tol = c.opt.tol;
    c.h_ctl->max_iter = c.opt.max_iter;
    HIPCHK(hipMemcpyAsync(c.ctl, c.h_ctl, sizeof(KrylovCtl), hipMemcpyHostToDevice, st));
    c.part = dalloc<double>(8 * MAX_PART);
    c.epart = dalloc<double>(8 * MAX_PART);
    c.res = dalloc<double>(32);
    HIPCHK(hipHostMalloc((void **)&c.h_part, 64 * sizeof(double), hipHostMallocDefault));
    c.slice_ptr = dalloc<int64_t>(c.nslices + 1);
    upload(c.slice_ptr, sptr, st);
    c.slice_w = dalloc<int32_t>(c.nslices);
    upload(c.slice_w, swid, st);
    c.scol = dalloc<int32_t>(c.nnzS);
    upload(c.scol, scol, st);
    c.sval = dalloc<double>(c.nnzS);
    upload(c.sval, sval, st);
    if (c.opt.krylov == FG_KRYLOV_MIXED) {
        std::vector<float> svf(sval.begin(), sval.end());
        c.svalf = dalloc<float>(c.nnzS);
        upload(c.svalf, svf, st);
        for (float4 **q : {&c.ep4f, &c.eq4f, &c.Dgf, &c.wf}) *q = dalloc<float4>(N);
        for (float2 **q : {&c.dprecf, &c.rf, &c.rtf, &c.pf, &c.phf, &c.vf, &c.sf, &c.shf, &c.tf, &c.df}) {
            *q = dalloc<float2>(N);
            HIPCHK(hipMemsetAsync(*q, 0, sizeof(float2) * N, st));
        }
        c.ic = dalloc<KrylovCtl>(1);
        HIPCHK(hipHostMalloc((void **)&c.h_ic, sizeof(KrylovCtl), hipHostMallocDefault));
        std::memset(c.h_ic, 0, sizeof(KrylovCtl));
        HIPCHK(hipMemcpyAsync(c.ic, c.h_ic, sizeof(KrylovCtl), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    std::vector<int4> tn(Tm);
    for (int t = 0; t < Tm; t++) tn[t] = make_int4(mtets[4 * t], mtets[4 * t + 1], mtets[4 * t + 2], mtets[4 * t + 3]);
    c.tet_nodes = dalloc<int4>(Tm);
    upload(c.tet_nodes, tn, st);
    std::vector<double> geo((size_t)17 * Tm);
    for (int t = 0; t < Tm; t++)
        for (int k = 0; k < 17; k++) geo[(size_t)k * Tm + t] = da_w[(size_t)17 * t + k];
    c.geo = dalloc<double>(geo.size());
    upload(c.geo, geo, st);
    c.tet_reg = dalloc<int32_t>(Tm);
    upload(c.tet_reg, mreg, st);
    c.regions = dalloc<DevRegion>(regs.size());
    upload(c.regions, regs, st);
    {
        // typical charge magnitude -Ms w div u ~ Ms w / h = Ms w^(2/3)
        double ms = 0, wsum = 0;
        for (auto &r : regs) ms = std::max(ms, std::fabs(r.Ms));
        for (int t = 0; t < Tm; t++)
            for (int g = 0; g < 5; g++) wsum += std::fabs(da_w[(size_t)17 * t + 12 + g]);
        const double w = wsum / (5.0 * Tm);
        c.charge_unit = (ms > 0 && w > 0) ? ms * std::cbrt(w * w) : 1.0;
    }
    c.elem_out = dalloc<double4>((size_t)4 * Tm + (size_t)3 * Fk);
    c.inc_ptr = dalloc<int32_t>(N + 1);
    upload(c.inc_ptr, incp, st);
    {
        // the slot of each element output entry in node order (inverse of the gather list)
        std::vector<int32_t> epos(inci.size());
        for (size_t k = 0; k < inci.size(); k++) epos[inci[k]] = (int32_t)k;
        c.epos = dalloc<int32_t>(epos.size());
        upload(c.epos, epos, st);
        HIPCHK(hipStreamSynchronize(st));  // the host vector goes at the end of this block
    }
    c.trik_nodes = dalloc<int4>(trikn.size());
    upload(c.trik_nodes, trikn, st);
    c.trik_geo = dalloc<double>(tkgeo.size());
    upload(c.trik_geo, tkgeo, st);
    c.tri_nodes = dalloc<int4>(trin.size());
    upload(c.tri_nodes, trin, st);
    c.tri_geo = dalloc<double>(tgeo.size());
    upload(c.tri_geo, tgeo, st);
    c.tri_corr = dalloc<double2>((size_t)3 * std::max(1, Fm));
    c.cinc_ptr = dalloc<int32_t>(N + 1);
    upload(c.cinc_ptr, cincp, st);
    c.cinc_idx = dalloc<int32_t>(cinci.size());
    upload(c.cinc_idx, cinci, st);
    c.corr = dalloc<double2>(N);
    c.src_pos = dalloc<double4>(c.S);
    upload(c.src_pos, spos, st);
    c.src_q = dalloc<double2>(c.S);
    std::vector<int64_t> rank(c.S);
    for (int64_t k = 0; k < c.S; k++) rank[k] = k;
    c.src_rank = dalloc<int64_t>(c.S);
    upload(c.src_rank, rank, st);
    if (c.inst_node_off.size() > 2) {
        const size_t K = c.inst_node_off.size() - 1;
        c.d_inst_node_off = dalloc<int64_t>(K + 1);
        upload(c.d_inst_node_off, c.inst_node_off, st);
        c.inst_bn2 = dalloc<double>(K + 1);
        std::vector<int32_t> ni(N);
        std::vector<longlong4> rs(K);
        for (size_t k = 0; k < K; k++) {
            for (int64_t i = c.inst_node_off[k]; i < c.inst_node_off[k + 1]; i++) ni[i] = (int32_t)k;
            rs[k] = make_longlong4(5 * c.inst_mtet_off[k], 5 * c.inst_mtet_off[k + 1],
                                   5 * (int64_t)Tm + 4 * c.inst_mtri_off[k], 5 * (int64_t)Tm + 4 * c.inst_mtri_off[k + 1]);
        }
        c.d_node_inst = dalloc<int32_t>(N);
        upload(c.d_node_inst, ni, st);
        c.d_inst_src = dalloc<longlong4>(K);
        upload(c.d_inst_src, rs, st);
        // instances by decreasing size (the persistent solve starts the long ones first)
        std::vector<int32_t> order(K);
        for (size_t k = 0; k < K; k++) order[k] = (int32_t)k;
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            return c.inst_node_off[a + 1] - c.inst_node_off[a] > c.inst_node_off[b + 1] - c.inst_node_off[b];
        });
        c.d_inst_order = dalloc<int32_t>(K);
        upload(c.d_inst_order, order, st);
        c.d_inst_stat = dalloc<int2>(K);
        c.d_inst_resn = dalloc<double2>(K);
        HIPCHK(hipHostMalloc((void **)&c.h_inst_stat, K * sizeof(int2), hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void **)&c.h_inst_resn, K * sizeof(double2), hipHostMallocDefault));
        HIPCHK(hipStreamSynchronize(st));  // order goes at the end of this block
    }
    HIPCHK(hipStreamSynchronize(st));
}

void free_all(Ctx &c)
{
    void *ptrs[] = {c.rec[0], c.rec[1], c.pos4, c.ep4, c.eq4, c.mag, c.mass, c.sdiag, c.Dg, c.dprec, c.b, c.x, c.r,
                    c.rt, c.p, c.ph, c.v, c.s, c.sh, c.t, c.w, c.ctl, c.part, c.epart, c.res, c.slice_ptr, c.slice_w,
                    c.scol, c.sval, c.tet_nodes, c.geo, c.tet_reg, c.regions, c.elem_out, c.inc_ptr, c.epos,
                    c.trik_nodes, c.trik_geo, c.tri_nodes, c.tri_geo, c.tri_corr, c.cinc_ptr, c.cinc_idx, c.corr,
                    c.src_pos, c.src_q, c.src_pos_f, c.src_rank, c.d_inst_node_off, c.inst_bn2, c.d_node_inst, c.d_inst_src,
                    c.hsp, c.hst, c.spin_s, c.spin_pref, c.svalf, c.ep4f, c.eq4f, c.Dgf, c.wf, c.dprecf, c.rf,
                    c.rtf, c.pf, c.phf, c.vf, c.sf, c.shf, c.tf, c.df, c.ic, c.d_inst_order, c.d_inst_stat,
                    c.d_inst_resn};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c.h_ctl) (void)hipHostFree(c.h_ctl);
    if (c.h_ic) (void)hipHostFree(c.h_ic);
    if (c.h_inst_stat) (void)hipHostFree(c.h_inst_stat);
    if (c.h_inst_resn) (void)hipHostFree(c.h_inst_resn);
    if (c.h_part) (void)hipHostFree(c.h_part);
    for (auto &pnd : c.pending) {
        (void)hipEventDestroy(pnd.a);
        (void)hipEventDestroy(pnd.b);
    }
    for (auto e : c.event_pool) (void)hipEventDestroy(e);
    if (c.ev_mid) (void)hipEventDestroy(c.ev_mid);
    if (c.ev_join) (void)hipEventDestroy(c.ev_join);
    if (c.aux) (void)hipStreamDestroy(c.aux);
    c.ev_mid = c.ev_join = nullptr;
    c.aux = nullptr;
}

double stencil_coef(const fg_ctx *x) { return x->prefactor * THETA * x->dt * GAMMA0; }

template <class F>
int guarded(fg_ctx *x, F &&f)
{
    if (!x) return FG_ERR_ARG;
    try {
        HIPCHK(hipSetDevice(x->c.device));
        return f();
    } catch (const std::string &e) {
        x->c.err = e;
        return FG_ERR_DEVICE;
    }
}

}  // namespace

// ---- profiling helpers ----
namespace fg {
int Ctx::kid(const char *name)
{
    for (size_t k = 0; k < stats.size(); k++)
        if (stats[k].name == name) return (int)k;
    stats.push_back({name, 0, 0.0});
    return (int)stats.size() - 1;
}
void Ctx::tic(int id, hipEvent_t *a)
{
    (void)id;
    if (event_pool.empty()) {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        event_pool.push_back(e);
    }
    *a = event_pool.back();
    event_pool.pop_back();
    (void)hipEventRecord(*a, stream);
}
void Ctx::toc(int id, hipEvent_t a)
{
    hipEvent_t b;
    if (event_pool.empty()) (void)hipEventCreate(&b);
    else {
        b = event_pool.back();
        event_pool.pop_back();
    }
    (void)hipEventRecord(b, stream);
    pending.push_back({id, a, b});
    if (pending.size() > 4096) flush_stats();
}
void Ctx::flush_stats()
{
    if (pending.empty()) return;
    (void)hipEventSynchronize(pending.back().b);
    for (auto &p : pending) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, p.a, p.b);
        stats[p.id].launches++;
        stats[p.id].ms += ms;
        event_pool.push_back(p.a);
        event_pool.push_back(p.b);
    }
    pending.clear();
}
}  // namespace fg

extern "C" {

void fg_default_options(fg_options *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->tol = 1e-6;
    o->max_iter = 700;
    o->demag_method = FG_DEMAG_DIRECT;
    o->fmm_order = 8;
    o->fmm_leaf = 64;
    o->precision = FG_PREC_FP64;
    o->device = 0;
    o->profile = 0;
}

namespace {

int create_impl(const fg_mesh *mesh, const fg_options *opt, fg_ctx **out, const std::vector<int64_t> *node_off,
                const std::vector<int64_t> *tet_off, const std::vector<int64_t> *tri_off)
{
    if (!mesh || !out || mesh->n_nodes <= 0 || mesh->n_tets <= 0 || !mesh->pos || !mesh->tet_nodes ||
        !mesh->tet_region || !mesh->vol || mesh->n_vol_regions <= 0 || (mesh->n_tris > 0 && (!mesh->tri_nodes || !mesh->tri_region || !mesh->tri_dMs))) {
        g_create_err = "fg_create: invalid mesh";
        return FG_ERR_ARG;
    }
    auto *x = new fg_ctx();
    Ctx &c = x->c;
    if (opt) c.opt = *opt;
    else fg_default_options(&c.opt);
    if (c.opt.max_iter <= 0 || !(c.opt.tol > 0) ||
        (c.opt.krylov != FG_KRYLOV_FP64 && c.opt.krylov != FG_KRYLOV_MIXED) || !(c.opt.krylov_inner_tol >= 0) ||
        c.opt.krylov_inner_tol >= 1 ||
        (c.opt.batch_solver != FG_BATCH_UNION && c.opt.batch_solver != FG_BATCH_PERSISTENT) ||
        (c.opt.batch_solver == FG_BATCH_PERSISTENT && c.opt.krylov == FG_KRYLOV_MIXED) ||
        (c.opt.batch_block != 0 && c.opt.batch_block != 256 && c.opt.batch_block != 512 &&
         c.opt.batch_block != 1024)) {
        g_create_err = "fg_create: invalid options";
        delete x;
        return FG_ERR_ARG;
    }
    if (node_off) {
        c.in_node_off = *node_off;
        c.in_tet_off = *tet_off;
        c.in_tri_off = *tri_off;
    }
    c.device = c.opt.device;
    try {
        HIPCHK(hipSetDevice(c.device));
        HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&c.aux, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&c.ev_mid, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
        build_plan(c, mesh);
        if (c.opt.demag_method == FG_DEMAG_FMM) {
            int rc = fmm_create(c);
            if (rc != FG_OK) throw std::string("fmm_create failed: ") + c.err;
        }
    } catch (const std::string &e) {
        g_create_err = e;
        if (c.stream) {
            (void)hipStreamSynchronize(c.stream);
            fmm_destroy(c);
            free_all(c);
            (void)hipStreamDestroy(c.stream);
        }
        delete x;
        return e.find("singular") != std::string::npos || e.find("range") != std::string::npos ? FG_ERR_MESH
                                                                                               : FG_ERR_DEVICE;
    }
    *out = x;
    return FG_OK;
}

}  // namespace

int fg_create(const fg_mesh *mesh, const fg_options *opt, fg_ctx **out)
{
    return create_impl(mesh, opt, out, nullptr, nullptr, nullptr);
}

int fg_create_batch(const fg_mesh *meshes, int k, const fg_options *opt, fg_ctx **out)
{
    if (!meshes || k <= 0 || !out) {
        g_create_err = "fg_create_batch: invalid arguments";
        return FG_ERR_ARG;
    }
    // the union of the instances: nodes, elements and regions side by side
    std::vector<double> pos, dMs;
    std::vector<int32_t> tets, treg, tris, freg;
    std::vector<fg_volume_region> vol;
    std::vector<fg_surface_region> surf;
    std::vector<int64_t> node_off(1, 0), tet_off(1, 0), tri_off(1, 0);
    int64_t surf_total = 0;
    for (int q = 0; q < k; q++) surf_total += std::max(0, meshes[q].n_surf_regions);
    for (int q = 0; q < k; q++) {
        const fg_mesh &m = meshes[q];
        if (m.n_nodes <= 0 || m.n_tets <= 0 || !m.pos || !m.tet_nodes || !m.tet_region || !m.vol ||
            m.n_vol_regions <= 0 || (m.n_tris > 0 && (!m.tri_nodes || !m.tri_region || !m.tri_dMs))) {
            g_create_err = "fg_create_batch: invalid mesh " + std::to_string(q);
            return FG_ERR_ARG;
        }
        const int32_t no = (int32_t)node_off.back(), vo = (int32_t)vol.size(), so = (int32_t)surf.size();
        pos.insert(pos.end(), m.pos, m.pos + 3 * (size_t)m.n_nodes);
        for (int64_t t = 0; t < 4 * (int64_t)m.n_tets; t++) tets.push_back(m.tet_nodes[t] + no);
        for (int t = 0; t < m.n_tets; t++)
            treg.push_back(m.tet_region[t] >= 0 && m.tet_region[t] < m.n_vol_regions ? m.tet_region[t] + vo : -1);
        for (int64_t f = 0; f < 3 * (int64_t)std::max(0, m.n_tris); f++) tris.push_back(m.tri_nodes[f] + no);
        for (int f = 0; f < m.n_tris; f++) {
            const int r = m.tri_region[f];
            // regions beyond the instance's list (created triangles) keep the default parameters
            freg.push_back(r >= 0 && r < m.n_surf_regions ? r + so : (int32_t)surf_total);
            dMs.push_back(m.tri_dMs[f]);
        }
        vol.insert(vol.end(), m.vol, m.vol + m.n_vol_regions);
        if (m.n_surf_regions > 0) surf.insert(surf.end(), m.surf, m.surf + m.n_surf_regions);
        node_off.push_back(node_off.back() + m.n_nodes);
        tet_off.push_back(tet_off.back() + m.n_tets);
        tri_off.push_back(tri_off.back() + std::max(0, m.n_tris));
    }
    if (node_off.back() > INT32_MAX || tet_off.back() > INT32_MAX || tri_off.back() > INT32_MAX) {
        g_create_err = "fg_create_batch: the union exceeds int32 indices";
        return FG_ERR_RANGE;
    }
    fg_mesh u{};
    u.n_nodes = (int32_t)node_off.back();
    u.n_tets = (int32_t)tet_off.back();
    u.n_tris = (int32_t)tri_off.back();
    u.n_vol_regions = (int32_t)vol.size();
    u.n_surf_regions = (int32_t)surf.size();
    u.pos = pos.data();
    u.tet_nodes = tets.data();
    u.tet_region = treg.data();
    u.tri_nodes = tris.empty() ? nullptr : tris.data();
    u.tri_region = freg.empty() ? nullptr : freg.data();
    u.tri_dMs = dMs.empty() ? nullptr : dMs.data();
    u.vol = vol.data();
    u.surf = surf.empty() ? nullptr : surf.data();
    return create_impl(&u, opt, out, &node_off, &tet_off, &tri_off);
}

int fg_batch_layout(fg_ctx *x, int64_t *node_off, int64_t *mtet_off, int *k)
{
    if (!x || !k) return FG_ERR_ARG;
    const Ctx &c = x->c;
    *k = (int)c.inst_node_off.size() - 1;
    for (int q = 0; q <= *k; q++) {
        if (node_off) node_off[q] = c.inst_node_off[q];
        if (mtet_off) mtet_off[q] = c.inst_mtet_off[q];
    }
    return FG_OK;
}

int fg_destroy(fg_ctx *x)
{
    if (!x) return FG_ERR_ARG;
    (void)hipSetDevice(x->c.device);
    (void)hipStreamSynchronize(x->c.stream);
    fmm_destroy(x->c);
    guess_destroy(x->c);
    free_all(x->c);
    (void)hipStreamDestroy(x->c.stream);
    delete x;
    return FG_OK;
}

int fg_set_state(fg_ctx *x, int which, const double *u, const double *v, const double *phi, const double *phiv)
{
    return guarded(x, [&]() {
        Ctx &c = x->c;
        if (which != FG_CURRENT && which != FG_NEXT && which != FG_BOTH) return FG_ERR_ARG;
        guess_restart(c);  // a new trajectory: no solve history to extrapolate
        std::vector<NodeRec> h(c.N);
        for (int i = 0; i < c.N; i++) {
            NodeRec &r = h[i];
            r.ux = u ? u[3 * i] : 0.0;
            r.uy = u ? u[3 * i + 1] : 0.0;
            r.uz = u ? u[3 * i + 2] : 0.0;
            r.vx = v ? v[3 * i] : 0.0;
            r.vy = v ? v[3 * i + 1] : 0.0;
            r.vz = v ? v[3 * i + 2] : 0.0;
            r.phi = phi ? phi[i] : 0.0;
            r.phiv = phiv ? phiv[i] : 0.0;
        }
        if (which == FG_CURRENT || which == FG_BOTH)
            HIPCHK(hipMemcpyAsync(c.rec[c.cur], h.data(), sizeof(NodeRec) * c.N, hipMemcpyHostToDevice, c.stream));
        if (which == FG_NEXT || which == FG_BOTH)
            HIPCHK(hipMemcpyAsync(c.rec[1 - c.cur], h.data(), sizeof(NodeRec) * c.N, hipMemcpyHostToDevice, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        x->basis_valid = false;
        return FG_OK;
    });
}

int fg_get_state(fg_ctx *x, int which, double *u, double *v, double *phi, double *phiv)
{
    return guarded(x, [&]() {
        Ctx &c = x->c;
        if (which != FG_CURRENT && which != FG_NEXT) return FG_ERR_ARG;
        std::vector<NodeRec> h(c.N);
        const NodeRec *src = which == FG_CURRENT ? c.rec[c.cur] : c.rec[1 - c.cur];
        HIPCHK(hipMemcpyAsync(h.data(), src, sizeof(NodeRec) * c.N, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        for (int i = 0; i < c.N; i++) {
            const NodeRec &r = h[i];
            if (u) {
                u[3 * i] = r.ux;
                u[3 * i + 1] = r.uy;
                u[3 * i + 2] = r.uz;
            }
            if (v) {
                v[3 * i] = r.vx;
                v[3 * i + 1] = r.vy;
                v[3 * i + 2] = r.vz;
            }
            if (phi) phi[i] = r.phi;
            if (phiv) phiv[i] = r.phiv;
        }
        return FG_OK;
    });
}

int fg_base_projection(fg_ctx *x, double r)
{
    return guarded(x, [&]() {
        x->r = r;
        x->basis_valid = false;
        return FG_OK;
    });
}

int fg_prepare_elements(fg_ctx *x, const double Hext[3], double dt, double prefactor)
{
    return guarded(x, [&]() {
        if (!Hext || !(dt > 0)) return FG_ERR_ARG;
        x->dt = dt;
        x->prefactor = prefactor;
        for (int k = 0; k < 3; k++) x->Hext[k] = Hext[k];
        launch_element(x->c, Hext, dt);
        x->elements_ready = true;
        return FG_OK;
    });
}

int fg_set_space_field(fg_ctx *x, const double *h)
{
    return guarded(x, [&]() {
        Ctx &c = x->c;
        if (!h) {
            if (c.hsp) HIPCHK(hipFree(c.hsp));
            c.hsp = nullptr;
            return FG_OK;
        }
        // the magnetic tetrahedra's entries, component-major for coalesced reads: [3 g + k][Tm]
        std::vector<double> soa((size_t)15 * c.Tm);
        for (int t = 0; t < c.Tm; t++) {
            const double *src = h + (size_t)15 * c.h_tet_mesh[t];
            for (int q = 0; q < 15; q++) soa[(size_t)q * c.Tm + t] = src[q];
        }
        if (!c.hsp) c.hsp = dalloc<double>(soa.size());
        HIPCHK(hipMemcpyAsync(c.hsp, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        return FG_OK;
    });
}

// The spin-accumulation buffers of a context and the prefactor of spinAcc::prepareExtraField
// (src/spinAccumulationSolver.cpp:138-149), D0 / (lsd^2 gamma0 Ms) per magnetic region, into pref
// (uploaded on the context's stream; the caller synchronises before pref goes)
static void spin_field_setup(Ctx &c, const fg_spin_volume *vol, std::vector<double> &pref)
{
    const double E = 1.602176634e-19;
    pref.assign(c.h_reg_Ms.size(), 0.0);
    for (size_t rg = 0; rg < pref.size(); rg++)
        if (c.h_reg_Ms[rg] > 0) {
            const fg_spin_volume &p = vol[rg];
            const double D0 = 2.0 * p.sigma / ((E * E) * p.N0);
            pref[rg] = D0 / ((p.lsd * p.lsd) * GAMMA0 * c.h_reg_Ms[rg]);
        }
    if (!c.hst) {
        c.hst = dalloc<double>((size_t)15 * c.Tm);
        c.spin_s = dalloc<double>((size_t)3 * c.N);
        c.spin_pref = dalloc<double>(pref.size());
    }
    HIPCHK(hipMemcpyAsync(c.spin_pref, pref.data(), pref.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
}

int fg_set_spin_accumulation(fg_ctx *x, const double *s, const fg_spin_volume *vol)
{
    return guarded(x, [&]() {
        Ctx &c = x->c;
        if (!s) {
            for (double **p : {&c.hst, &c.spin_s, &c.spin_pref}) {
                if (*p) HIPCHK(hipFree(*p));
                *p = nullptr;
            }
            return FG_OK;
        }
        if (!vol) return FG_ERR_ARG;
        std::vector<double> pref;
        spin_field_setup(c, vol, pref);
        HIPCHK(hipMemcpyAsync(c.spin_s, s, (size_t)3 * c.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
        launch_spin_field(c);  // Hst itself on the device
        HIPCHK(hipStreamSynchronize(c.stream));
        return FG_OK;
    });
}

int fg_spin_solve_ctx(fg_spin *sp, fg_ctx *x, int32_t accumulate, double *s_out, fg_solve_info *info)
{
    return guarded(x, [&]() {
        Ctx &c = x->c;
        if (!sp) return FG_ERR_ARG;
        int nreg = 0, nn = 0, dev = 0;
        const fg_spin_volume *vol = spin_volumes(sp, &nreg, &nn, &dev);
        if (nn != c.N || dev != c.device || nreg < (int)c.h_reg_Ms.size()) return FG_ERR_ARG;
        std::vector<double> pref;
        spin_field_setup(c, vol, pref);
        const int rc = spin_solve_nodes(sp, c.rec[c.cur], c.rec[1 - c.cur], c.stream, accumulate, c.spin_s, info);
        if (rc != FG_OK) return rc;
        launch_spin_field(c);
        if (s_out)
            HIPCHK(hipMemcpyAsync(s_out, c.spin_s, (size_t)3 * c.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        return FG_OK;
    });
}

int fg_gauss_points(const fg_mesh *m, double *pg)
{
    if (!m || !pg || !m->pos || !m->tet_nodes || m->n_tets < 0) return FG_ERR_ARG;
    for (int64_t t = 0; t < m->n_tets; t++)
        for (int g = 0; g < 5; g++)
            for (int k = 0; k < 3; k++) {
                double x = 0;
                for (int a = 0; a < 4; a++) x += tet_a(a, g) * m->pos[3 * (size_t)m->tet_nodes[4 * t + a] + k];
                pg[15 * t + 3 * g + k] = x;
            }
    return FG_OK;
}

int fg_prepare_elements_amp(fg_ctx *x, double amp, double dt, double prefactor)
{
    return guarded(x, [&]() {
        if (!x->c.hsp || !(dt > 0)) return x->c.hsp ? FG_ERR_ARG : FG_ERR_STATE;
        x->dt = dt;
        x->prefactor = prefactor;
        for (int k = 0; k < 3; k++) x->Hext[k] = 0;
        const double zero[3] = {0, 0, 0};
        launch_element(x->c, zero, dt, x->c.hsp, amp);
        x->elements_ready = true;
        return FG_OK;
    });
}

// After the solve (either Krylov mode): the update of NEXT, the guess's bookkeeping and v_max.
static int finish_solve(fg_ctx *x, double dt, fg_solve_info *info)
{
    Ctx &c = x->c;
    launch_update(c, dt);
    guess_on_solved(c);
    HIPCHK(hipMemcpyAsync(c.h_ctl, c.ctl, sizeof(KrylovCtl), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    if (info) info->v_max = GAMMA0 * std::sqrt(c.h_ctl->vmax2);
    if (c.opt.profile) c.flush_stats();
    return FG_OK;
}

// FG_KRYLOV_MIXED: FP64 residual, FP32 bicg_dir on the correction, x += d, until the FP64 residual
// meets tol ||b||. The FP32 loop runs in batches between host reads of its control block, the
// first batch one more than the last solve's FP32 count; the outer test follows each correction.
static int solve_mixed(fg_ctx *x, double dt, double cst, fg_solve_info *info)
{
    Ctx &c = x->c;
    // the FP32 loop's relative target (above its rounding floor) and its iterations per correction
    const double TOL_IN = c.opt.krylov_inner_tol > 0 ? c.opt.krylov_inner_tol : 3e-7;
    constexpr int MAX_INNER = 20;
    constexpr int MAX_REF = 8;  // corrections
    launch_mixed_start(c);
    launch_mixed_residual(c, cst, TOL_IN, MAX_INNER);
    bool done = false;
    for (int ref = 0; ref < MAX_REF && !done; ref++) {
        int launched = 0, batch = std::max(2, std::min(x->last_nit + 1, MAX_INNER));
        while (true) {
            for (int k = 0; k < batch && launched <= MAX_INNER; k++, launched++) launch_inner_iteration(c, cst);
            HIPCHK(hipMemcpyAsync(c.h_ic, c.ic, sizeof(KrylovCtl), hipMemcpyDeviceToHost, c.stream));
            HIPCHK(hipStreamSynchronize(c.stream));
            if (c.h_ic->done || launched > MAX_INNER) break;
            batch = 4;
        }
        launch_refine(c);
        launch_mixed_residual(c, cst, TOL_IN, MAX_INNER);
        HIPCHK(hipMemcpyAsync(c.h_ctl, c.ctl, sizeof(KrylovCtl), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        done = c.h_ctl->done != 0;
    }
    const KrylovCtl k = *c.h_ctl;
    x->last_nit = std::max(2, k.refinements > 0 ? k.nit / k.refinements : k.nit);
    const bool failed = !done || k.status != FG_ITER_CONVERGED;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->status = done ? k.status : FG_ITER_OVERFLOW;
        info->iterations = k.nit;
        info->residual = k.res;
        info->rhs_norm = k.rhsn;
        info->refinements = k.refinements;
        info->inner_iterations = k.nit;
    }
    if (failed) return FG_SOLVE_FAILED;
    return finish_solve(x, dt, info);
}

// FG_BATCH_PERSISTENT: one launch solves every instance to its own criterion; the batch has failed
// if any instance has. Reported: the worst status, the most iterations, the largest residual and
// the smallest nonzero ||b_k||.
static int solve_batch(fg_ctx *x, double dt, double cst, fg_solve_info *info)
{
    Ctx &c = x->c;
    const int K = (int)c.inst_node_off.size() - 1;
    launch_batch_solve(c, cst, c.opt.batch_block);
    HIPCHK(hipMemcpyAsync(c.h_inst_stat, c.d_inst_stat, K * sizeof(int2), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipMemcpyAsync(c.h_inst_resn, c.d_inst_resn, K * sizeof(double2), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    int status = FG_ITER_CONVERGED, nit = 0;
    double res = 0, rhsn = -1;
    for (int k = 0; k < K; k++) {
        const int2 st = c.h_inst_stat[k];
        const double2 rn = c.h_inst_resn[k];
        if (st.x != FG_ITER_CONVERGED && status == FG_ITER_CONVERGED) status = st.x;
        nit = std::max(nit, st.y);
        res = std::max(res, rn.x);
        if (rn.y > 0 && (rhsn < 0 || rn.y < rhsn)) rhsn = rn.y;
    }
    x->last_nit = std::max(2, nit);
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->status = status;
        info->iterations = nit;
        info->residual = res;
        info->rhs_norm = std::max(rhsn, 0.0);
    }
    if (status != FG_ITER_CONVERGED) return FG_SOLVE_FAILED;
    return finish_solve(x, dt, info);
}

int fg_solve(fg_ctx *x, double dt, fg_solve_info *info)
{
    return guarded(x, [&]() {
        Ctx &c = x->c;
        if (!x->elements_ready) return FG_ERR_STATE;
        if (dt != x->dt) return FG_ERR_ARG;
        const double cst = stencil_coef(x);
        launch_node_assemble(c, x->r, cst, dt);
        guess_on_assembly(c, dt);
        x->basis_valid = true;
        if (c.opt.krylov == FG_KRYLOV_MIXED) return solve_mixed(x, dt, cst, info);
        if (c.opt.batch_solver == FG_BATCH_PERSISTENT && c.inst_node_off.size() > 2)
            return solve_batch(x, dt, cst, info);
        launch_r0(c, cst);
        // BiCGSTAB: batches of iterations, the device control block says when to stop. Iteration
        // n's convergence is tested at the start of iteration n + 1, so the first batch is the
        // previous solve's count plus that test: a solve that converges like the last one needs a
        // single round trip to the host.
        // Each batch is followed by the node update gated on the control block, so that a solve
        // accepted in the batch is finished by the time the host reads it (no second round trip);
        // profiled runs update after the read, so the update's statistics hold real launches only.
        // A batch ends with the first phase of its next iteration, the test that may end the solve;
        // the next batch starts with that iteration's other four phases.
        const bool gated = !c.opt.profile;
        int launched = 0;   // iterations queued
        bool open = false;  // the last batch ended with the first phase of iteration launched + 1
        int batch = std::max(2, std::min(x->last_nit + 1, c.opt.max_iter));
        while (true) {
            if (open) {
                launch_iteration(c, cst, 2);
                launched++;
                open = false;
            }
            for (int k = 0; k + 1 < batch && launched <= c.opt.max_iter; k++, launched++) launch_iteration(c, cst);
            if (launched <= c.opt.max_iter) {
                launch_iteration(c, cst, 1);
                open = true;
            }
            if (gated) launch_update(c, dt, true);
            HIPCHK(hipMemcpyAsync(c.h_ctl, c.ctl, sizeof(KrylovCtl), hipMemcpyDeviceToHost, c.stream));
            HIPCHK(hipStreamSynchronize(c.stream));
            if (c.h_ctl->done || launched > c.opt.max_iter) break;
            batch = 4;
        }
        const KrylovCtl k = *c.h_ctl;
        x->last_nit = std::max(2, k.nit);
        bool failed = (k.status == FG_ITER_OVERFLOW || k.status == FG_ITER_CANNOT_CONVERGE || k.res > c.opt.tol ||
                       !k.done);
        if (info) {
            std::memset(info, 0, sizeof(*info));
            info->status = k.done ? k.status : FG_ITER_OVERFLOW;
            info->iterations = k.nit;
            info->residual = k.res;
            info->rhs_norm = k.rhsn;
            info->v_max = 0;
        }
        if (failed) return FG_SOLVE_FAILED;
        if (!gated) return finish_solve(x, dt, info);
        guess_on_solved(c);
        if (info) info->v_max = GAMMA0 * std::sqrt(k.vmax2);
        return FG_OK;
    });
}

int fg_calc_demag(fg_ctx *x)
{
    return guarded(x, [&]() {
        Ctx &c = x->c;
        launch_charges(c);
        if (c.opt.demag_method == FG_DEMAG_FMM) launch_fmm(c);
        else launch_direct_p2p(c);
        // no wait: every entry point works in order on the context's stream and waits where it
        // reads results back
        if (c.opt.profile) {
            HIPCHK(hipStreamSynchronize(c.stream));
            c.flush_stats();
        }
        return FG_OK;
    });
}

int fg_energy(fg_ctx *x, const double Hext[3], double E[4])
{
    return guarded(x, [&]() {
        Ctx &c = x->c;
        if (!Hext || !E) return FG_ERR_ARG;
        launch_energy(c, Hext);
        HIPCHK(hipMemcpyAsync(c.h_part, c.res, 8 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        E[FG_E_EXCHANGE] = c.h_part[0];
        E[FG_E_ANISOTROPY] = c.h_part[1] + c.h_part[4];
        E[FG_E_DEMAG] = c.h_part[2] + c.h_part[5];
        E[FG_E_ZEEMAN] = c.h_part[3];
        return FG_OK;
    });
}

int fg_energy_amp(fg_ctx *x, double amp, double E[4])
{
    return guarded(x, [&]() {
        Ctx &c = x->c;
        if (!E) return FG_ERR_ARG;
        if (!c.hsp) return FG_ERR_STATE;
        const double zero[3] = {0, 0, 0};
        launch_energy(c, zero, c.hsp, amp);
        HIPCHK(hipMemcpyAsync(c.h_part, c.res, 8 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        E[FG_E_EXCHANGE] = c.h_part[0];
        E[FG_E_ANISOTROPY] = c.h_part[1] + c.h_part[4];
        E[FG_E_DEMAG] = c.h_part[2] + c.h_part[5];
        E[FG_E_ZEEMAN] = c.h_part[3];
        return FG_OK;
    });
}

int fg_evolution(fg_ctx *x)
{
    return guarded(x, [&]() {
        launch_evolution(x->c);
        x->basis_valid = false;
        return FG_OK;
    });
}

int fg_max_angle(fg_ctx *x, double *angle)
{
    return guarded(x, [&]() {
        Ctx &c = x->c;
        if (!angle) return FG_ERR_ARG;
        launch_max_angle(c);
        HIPCHK(hipMemcpyAsync(c.h_part, c.res + 8, sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        *angle = std::acos(c.h_part[0]);
        return FG_OK;
    });
}

int fg_step(fg_ctx *x, const double Hext[3], double dt, double prefactor, double r, double dumax,
            fg_solve_info *info, double E[4], double *max_angle)
{
    return guarded(x, [&]() {
        Ctx &c = x->c;
        if (!Hext || !E || !max_angle) return FG_ERR_ARG;
        int rc = fg_base_projection(x, r);
        if (rc == FG_OK) rc = fg_prepare_elements(x, Hext, dt, prefactor);
        fg_solve_info si{};
        if (rc == FG_OK) rc = fg_solve(x, dt, &si);
        if (info && (rc == FG_OK || rc == FG_SOLVE_FAILED)) *info = si;
        if (rc != FG_OK) return rc;
        if (dumax > 0 && dt * si.v_max > dumax) return FG_STEP_REJECTED;
        // compute_all and max_angle, queued back to back and read back once
        launch_charges(c);
        if (c.opt.demag_method == FG_DEMAG_FMM) launch_fmm(c);
        else launch_direct_p2p(c);
        launch_energy(c, Hext);
        launch_evolution(c);
        x->basis_valid = false;
        launch_max_angle(c);
        HIPCHK(hipMemcpyAsync(c.h_part, c.res, 9 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        E[FG_E_EXCHANGE] = c.h_part[0];
        E[FG_E_ANISOTROPY] = c.h_part[1] + c.h_part[4];
        E[FG_E_DEMAG] = c.h_part[2] + c.h_part[5];
        E[FG_E_ZEEMAN] = c.h_part[3];
        *max_angle = std::acos(c.h_part[8]);
        if (c.opt.profile) c.flush_stats();
        return FG_OK;
    });
}

int fg_average(fg_ctx *x, int which, double out[3])
{
    return guarded(x, [&]() {
        Ctx &c = x->c;
        if (!out || (which != 0 && which != 1)) return FG_ERR_ARG;
        launch_average(c, which);
        HIPCHK(hipMemcpyAsync(c.h_part, c.res + 16, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        for (int k = 0; k < 3; k++) out[k] = c.h_part[k] / c.totalMagVol;
        return FG_OK;
    });
}

int fg_get_basis(fg_ctx *x, double *ep, double *eq)
{
    return guarded(x, [&]() {
        Ctx &c = x->c;
        if (!x->basis_valid) {
            launch_basis(c, x->r);
            x->basis_valid = true;
        }
        std::vector<double4> a(c.N), b(c.N);
        HIPCHK(hipMemcpyAsync(a.data(), c.ep4, sizeof(double4) * c.N, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipMemcpyAsync(b.data(), c.eq4, sizeof(double4) * c.N, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        for (int i = 0; i < c.N; i++) {
            if (ep) {
                ep[3 * i] = a[i].x;
                ep[3 * i + 1] = a[i].y;
                ep[3 * i + 2] = a[i].z;
            }
            if (eq) {
                eq[3 * i] = b[i].x;
                eq[3 * i + 1] = b[i].y;
                eq[3 * i + 2] = b[i].z;
            }
        }
        return FG_OK;
    });
}

int fg_system_size(fg_ctx *x, int64_t *rows, int64_t *nnz)
{
    if (!x || !rows || !nnz) return FG_ERR_ARG;
    *rows = 2 * (int64_t)x->c.N;
    *nnz = 4 * x->c.nnz_real;
    return FG_OK;
}

int fg_guess_init(fg_ctx *x, uint64_t seed)
{
    return guarded(x, [&]() {
        guess_init(x->c, seed);
        return FG_OK;
    });
}

int fg_guess_set(fg_ctx *x, const float *params, int64_t n)
{
    if (!params || n != FG_GUESS_NPARAM) return FG_ERR_ARG;
    return guarded(x, [&]() {
        if (!x->c.guess) guess_init(x->c, 0);
        guess_set(x->c, params);
        return FG_OK;
    });
}

int fg_guess_get(fg_ctx *x, float *params, int64_t n)
{
    if (!params || n != FG_GUESS_NPARAM) return FG_ERR_ARG;
    return guarded(x, [&]() {
        if (!x->c.guess) return FG_ERR_STATE;
        guess_get(x->c, params);
        return FG_OK;
    });
}

int fg_guess_enable(fg_ctx *x, int on)
{
    if (!x) return FG_ERR_ARG;
    return guess_enable(x->c, on) ? FG_OK : FG_ERR_STATE;
}

int fg_guess_train(fg_ctx *x, float lr, double *loss, double *loss_ref)
{
    return guarded(x, [&]() {
        if (!x->c.guess) return FG_ERR_STATE;
        guess_train(x->c, lr, loss, loss_ref);
        return FG_OK;
    });
}

int fg_guess_gradient(fg_ctx *x, float *grad, double *loss, double *loss_ref)
{
    return guarded(x, [&]() {
        if (!x->c.guess) return FG_ERR_STATE;
        guess_gradient(x->c, grad, loss, loss_ref);
        return FG_OK;
    });
}

int fg_guess_features(fg_ctx *x, float *feat)
{
    if (!feat) return FG_ERR_ARG;
    return guarded(x, [&]() {
        if (!x->c.guess) return FG_ERR_STATE;
        guess_batch(x->c, feat);
        return FG_OK;
    });
}

int fg_fmm_info(fg_ctx *x, int64_t info[16])
{
    if (!x || !info) return FG_ERR_ARG;
    if (!x->c.fmm) return FG_ERR_STATE;
    fmm_info(x->c, info);
    return FG_OK;
}

int fg_get_system(fg_ctx *x, int64_t *rowptr, int32_t *col, double *val, double *rhs)
{
    return guarded(x, [&]() {
        Ctx &c = x->c;
        if (!x->basis_valid) return FG_ERR_STATE;
        const int N = c.N;
        std::vector<double4> ep(N), eq(N), Dg(N);
        std::vector<double2> b(N);
        HIPCHK(hipMemcpyAsync(ep.data(), c.ep4, sizeof(double4) * N, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipMemcpyAsync(eq.data(), c.eq4, sizeof(double4) * N, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipMemcpyAsync(Dg.data(), c.Dg, sizeof(double4) * N, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipMemcpyAsync(b.data(), c.b, sizeof(double2) * N, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        const double cst = stencil_coef(x);
        int64_t pos = 0;
        if (rowptr) rowptr[0] = 0;
        for (int i = 0; i < N; i++) {
            const double4 ri[2] = {eq[i], ep[i]};  // row 2i: e_q equation, row 2i+1: e_p equation
            for (int di = 0; di < 2; di++) {
                for (int64_t k = c.h_rowptr[i]; k < c.h_rowptr[i + 1]; k++) {
                    const int j = c.h_col[k];
                    const double4 cj[2] = {ep[j], eq[j]};
                    for (int dj = 0; dj < 2; dj++) {
                        double v = cst * c.h_val[k] * (ri[di].x * cj[dj].x + ri[di].y * cj[dj].y + ri[di].z * cj[dj].z);
                        if (j == i) {
                            const double g[2][2] = {{Dg[i].x, Dg[i].y}, {Dg[i].z, Dg[i].w}};
                            v += g[di][dj];
                            if (!c.h_mag[i]) v = (di == dj) ? 1.0 : 0.0;
                        }
                        if (col) col[pos] = 2 * j + dj;
                        if (val) val[pos] = v;
                        pos++;
                    }
                }
                if (rowptr) rowptr[2 * i + di + 1] = pos;
            }
            if (rhs) {
                rhs[2 * i] = b[i].x;
                rhs[2 * i + 1] = b[i].y;
            }
        }
        return FG_OK;
    });
}

int fg_get_solution(fg_ctx *x, double *out)
{
    return guarded(x, [&]() {
        Ctx &c = x->c;
        if (!out) return FG_ERR_ARG;
        HIPCHK(hipMemcpyAsync(out, c.x, sizeof(double2) * c.N, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        return FG_OK;
    });
}

int fg_kernel_stats(fg_ctx *x, int k, char *name, int name_len, int64_t *launches, double *ms)
{
    if (!x) return FG_ERR_ARG;
    (void)hipSetDevice(x->c.device);
    x->c.flush_stats();
    if (k < 0 || k >= (int)x->c.stats.size()) return FG_ERR_RANGE;
    const KernelStat &s = x->c.stats[k];
    if (name && name_len > 0) {
        std::strncpy(name, s.name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (launches) *launches = s.launches;
    if (ms) *ms = s.ms;
    return FG_OK;
}

int fg_reset_kernel_stats(fg_ctx *x)
{
    if (!x) return FG_ERR_ARG;
    (void)hipSetDevice(x->c.device);
    x->c.flush_stats();
    for (auto &s : x->c.stats) {
        s.launches = 0;
        s.ms = 0;
    }
    return FG_OK;
}

int fg_synchronize(fg_ctx *x)
{
    return guarded(x, [&]() {
        HIPCHK(hipStreamSynchronize(x->c.stream));
        return FG_OK;
    });
}

const char *fg_last_error(fg_ctx *x)
{
    if (!x) return g_create_err.c_str();
    return x->c.err.c_str();
}

}  // extern "C"
