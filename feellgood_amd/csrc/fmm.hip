// fmm.hip -- octree fast multipole method for the demag potentials (both charge sets, u and v, in
// one pass). Replaces ScalFMM's FFmmAlgorithmThreadTsm with a rotation kernel (FeeLLGood
// src/fmm_demag.h:26-53,184-224): same potential sum_s q_s / |x_i - y_s|, a uniform-depth octree
// whose leaf level follows the source density instead of the fixed height 6.
//
// Plan (host, once): cube around all sources and targets, Morton-sorted sources and targets, the
// occupied cells of every level, parent/child links, interaction lists (children of the parent's
// neighbours that are not neighbours) and near lists (adjacent leaves), the irregular-harmonic
// table of the 316 M2L offsets and the regular table of the 8 child offsets.
//
// Device pass per call: near field (a precomputed sparse operator on the element charge values)
// -> P2M (leaves) -> M2M (upward) -> M2L (levels 2..L-2) -> L2L (downward) -> per group of targets
// of a level-(L-1) cell: M2P of the cells of levels L-1 and L in the far field, and L2P.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <type_traits>
#include <string>
#include <unordered_map>
#include <vector>

#include "ctx.h"
#include "fmm_math.h"

using namespace fg;

struct FmmPlan {
    int p = 8, pf = 6, L = 2, nc = 0, ni = 0;  // pf: degree of the far entries of the group lists
    double H = 0, C[3] = {0, 0, 0};
    std::vector<double> W;            // cell width per level
    std::vector<int> lev_off;         // first cell of each level (size L+2)
    // cells (all levels)
    std::vector<double4> cen;         // centre (x, y, z, level)
    std::vector<int32_t> parent, oct; // parent cell, octant of the cell in its parent
    std::vector<char> has_src, has_tgt;
    std::vector<int32_t> ch_ptr, ch_idx;      // children with sources (M2M)
    std::vector<int32_t> il_ptr, il_src, il_off;  // interaction lists (target cells)
    std::vector<int32_t> leaf_sb, leaf_se, leaf_tb, leaf_te;  // ranges per leaf (index - lev_off[L])
    std::vector<int32_t> near_ptr, near_leaf;  // per leaf: source leaves adjacent (leaf indices)
    // Evaluation groups: the targets of one level-(L-1) cell. M2L/L2L run on levels 2..ls-1, the
    // group's targets evaluate the local expansion of its level-(L-2) ancestor (lcell) and the
    // multipoles of its source list (the cell's interaction list and the children of its
    // neighbours) directly; u_mask bit o says whether a list entry applies to the targets in child
    // octant o (entries adjacent to that child belong to the near field). Each group's list holds
    // first the entries at Chebyshev distance 2 (in widths of the entry's level) from some of the
    // targets they serve, evaluated to degree p, then from u_far[g] on those at distance 3 from all
    // of them, evaluated to degree pf.
    int ls = 2;
    std::vector<int32_t> grp_tb, grp_te, grp_lcell, u_ptr, u_cell, u_far;
    std::vector<uint8_t> u_mask, tgt_oct;
    std::vector<int32_t> tgt_grp;
    // sorted points
    std::vector<int64_t> src_rank;    // original source index -> sorted position
    std::vector<double4> src_sorted;
    std::vector<int32_t> tgt_node;    // sorted target -> node index
    std::vector<int32_t> tgt_leaf;    // sorted target -> leaf index
    std::vector<double4> tgt_pos;
    // tables
    std::vector<double2> It;          // [343][ni]
    std::vector<double2> Rc;          // [8][nc]
    // device copies
    double4 *d_cen = nullptr;
    int32_t *d_leaf_sb = nullptr, *d_leaf_se = nullptr;
    int32_t *d_grp_tb = nullptr, *d_grp_te = nullptr, *d_grp_lcell = nullptr, *d_u_ptr = nullptr, *d_u_cell = nullptr,
            *d_u_far = nullptr;
    uint8_t *d_u_mask = nullptr, *d_tgt_oct = nullptr;
    // M2M and L2L as real linear operators per child octant (xlate_layout packing), and their
    // launch schedules: M2M parents per level with their children by octant (-1: none); L2L
    // parents per level with the children that receive a local expansion
    void *d_Tm2m = nullptr, *d_Tl2l = nullptr;  // FP64, or FP32 in the mixed-precision mode
    void *d_Tm2m_pair = nullptr;  // FP32 mode: M2M in the k order of k_fmm_m2m_all's pair loads
    // M2L on the matrix cores: the operator of every offset code (fragment order), the target
    // cells in tiles of 64 (one octant class per tile), each tile's offset codes in chunks of up
    // to M2L_CHUNK with the source cell of every (offset, target) (-1: none), and the partial sums
    // of every chunk, reduced per target in chunk order
    void *d_Tm2l = nullptr, *d_m2l_part = nullptr;  // FP64, or FP32 in the mixed-precision mode
    std::vector<int32_t> m2l_tile_cell, m2l_tile_chunk, m2l_chunk_code, m2l_chunk_src;
    int32_t *d_m2l_tile_cell = nullptr, *d_m2l_tile_chunk = nullptr, *d_m2l_chunk_code = nullptr,
            *d_m2l_chunk_src = nullptr;
    int m2l_ntiles = 0, m2l_nchunks = 0;
    std::vector<int32_t> m2m_child, l2l_par, l2l_par_off, l2l_child;
    int32_t *d_m2m_child = nullptr, *d_l2l_par = nullptr, *d_l2l_child = nullptr;
    // near-field operator: per sorted target, (element charge index, coefficient) entries
    int64_t *d_nf_ptr = nullptr;  // [T+1] start of each target's coefficient row
    double *d_nf_val = nullptr;
    int ntl = 0;                  // leaves with targets
    int32_t *d_tl_tb = nullptr, *d_tl_te = nullptr;  // their target ranges
    int64_t *d_lc_ptr = nullptr;  // [ntl+1] shared column list of each such leaf
    int32_t *d_lc_col = nullptr;
    double2 *d_pot = nullptr;  // [T] near-field + correction potential
    double2 *d_far = nullptr;  // [T] multipole (far) part of the potential, M2P's output
    int32_t *d_tgt_lc = nullptr;  // [T] local-expansion cell of each target's group (-1: none)
    float *d_nf_valf = nullptr;  // FP32 coefficients (mixed precision; d_nf_val is then released)
    bool mixed = false;
    double mscale = 1.0;         // expansions times mscale (a power of two) in the FP32 mode
    int64_t m2p_pairs = 0, m2l_pairs = 0, lc_cols = 0;
    int64_t m2p_slots = 0, m2p_far_pairs = 0;  // lane slots the M2P launch mapping executes (useful pairs / slots = SIMD efficiency)
    // M2P launches: the groups of more than M2P_HALF_MAX targets (one cell per row), then the
    // smaller ones (one cell per half row)
    std::vector<int32_t> m2p_gid;
    int m2p_n1 = 0;
    int32_t *d_m2p_gid = nullptr;
    int64_t nf_nnz = 0;
    int32_t *d_tgt_node = nullptr;
    double4 *d_tgt_pos = nullptr;
    // multipole and local expansions [ncell][2][nc]: double2, or float2 times mscale in the
    // mixed-precision mode (every translation then runs in FP32 on the matrix cores)
    void *d_M = nullptr, *d_L = nullptr;
    std::vector<int32_t> src_cells_lev_off, tgt_cells_lev_off;  // offsets into the lists
    std::vector<int32_t> src_cells, tgt_cells;
    // Launch schedule in global cell ids (a device plan may hold several independent trees, one
    // per instance, side by side): leaves, M2M parents per level, M2L cells, L2L cells per level.
    int Lmax = 2;
    std::vector<int32_t> leaf_cell, leaf_ix, leaf_iy, leaf_iz;
    std::vector<double> cell_iw;  // 1 / cell width
    std::vector<int32_t> m2m_lev_off, m2m_cells, m2l_cells, l2l_lev_off, l2l_cells;
    int32_t *d_leaf_cell = nullptr, *d_m2m_cells = nullptr;
    // P2M source records in tree order (k_fmm_p2m): double4 / float4 positions with the weight
    // w0, codes, and the two further weights of each triangle Gauss point (double2 / float2)
    void *d_p2m_pos = nullptr, *d_p2m_tw = nullptr;
    int32_t *d_p2m_code = nullptr;
    int ncell() const { return (int)cen.size(); }
};

namespace {

constexpr int LMAX = 10;

inline uint32_t spread3(uint32_t v)
{
    uint32_t r = 0;
    for (int b = 0; b < LMAX; b++) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}
inline uint32_t morton(uint32_t ix, uint32_t iy, uint32_t iz) { return (spread3(ix) << 2) | (spread3(iy) << 1) | spread3(iz); }
inline void unmorton(uint32_t c, int l, int &ix, int &iy, int &iz)
{
    ix = iy = iz = 0;
    for (int b = 0; b < l; b++) {
        ix |= ((c >> (3 * b + 2)) & 1u) << b;
        iy |= ((c >> (3 * b + 1)) & 1u) << b;
        iz |= ((c >> (3 * b)) & 1u) << b;
    }
}

}  // namespace

namespace fg {

// Lower degree of the far list entries (Chebyshev distance 3 from every target they serve): the
// truncation error of a cell three widths away at this degree stays below that of the cells two
// widths away at degree p (the film's FMM error grows by at most 5 % over one degree everywhere).
constexpr int fmm_far_degree(int p) { return p <= 6 ? p - 2 : p == 8 ? 5 : 6; }
// Degree of the far field of the second charge set (charges of v) in the mixed-precision mode: its
// potential enters the step only through theta dt H_d(v) (src/tetra.cpp:289), a term some 1e-4 of
// H_d(u) at the benchmark steps, so its multipole evaluation is cut to this degree.
constexpr int fmm_v_degree(int p) { return p <= 4 ? p : 4; }

// Host construction of the plan from source positions (original order) and target nodes.
std::string fmm_build_host(FmmPlan &P, const std::vector<double4> &src, const std::vector<int32_t> &tnode,
                           const std::vector<double4> &tpos, int p, int pf, int leaf_target)
{
    if (p != 4 && p != 6 && p != 8 && p != 10) return "fmm_order must be 4, 6, 8 or 10";
    if (pf == 0) pf = fmm_far_degree(p);  // callers resolve the precision-dependent default
    if (pf != p && pf != fmm_far_degree(p)) return "fmm_far_order must be 0, fmm_order or the default far degree";
    P.p = p;
    P.pf = pf;
    P.nc = fcount(p);
    P.ni = fcount(2 * p);
    const size_t S = src.size(), T = tpos.size();
    if (S == 0 || T == 0) return "empty FMM problem";
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    auto upd = [&](const double4 &q) {
        const double v[3] = {q.x, q.y, q.z};
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], v[k]);
            hi[k] = std::max(hi[k], v[k]);
        }
    };
    for (auto &q : src) upd(q);
    for (auto &q : tpos) upd(q);
    double ext = 0;
    for (int k = 0; k < 3; k++) {
        P.C[k] = 0.5 * (lo[k] + hi[k]);
        ext = std::max(ext, hi[k] - lo[k]);
    }
    P.H = 0.5 * ext * (1.0 + 1e-9) + 1e-300;
    const double corner[3] = {P.C[0] - P.H, P.C[1] - P.H, P.C[2] - P.H};
    const double scale = (double)(1u << LMAX) / (2.0 * P.H);
    auto code_of = [&](const double4 &q) {
        const double v[3] = {q.x, q.y, q.z};
        uint32_t ic[3];
        for (int k = 0; k < 3; k++) {
            double f = std::floor((v[k] - corner[k]) * scale);
            if (f < 0) f = 0;
            if (f > (1u << LMAX) - 1) f = (1u << LMAX) - 1;
            ic[k] = (uint32_t)f;
        }
        return morton(ic[0], ic[1], ic[2]);
    };
    std::vector<uint32_t> scode(S), tcode(T);
    for (size_t i = 0; i < S; i++) scode[i] = code_of(src[i]);
    for (size_t i = 0; i < T; i++) tcode[i] = code_of(tpos[i]);
    std::vector<int64_t> sord(S);
    for (size_t i = 0; i < S; i++) sord[i] = (int64_t)i;
    std::stable_sort(sord.begin(), sord.end(), [&](int64_t a, int64_t b) { return scode[a] < scode[b]; });
    // leaf level: first level whose mean occupancy is at most leaf_target sources
    int L = LMAX;
    for (int l = 2; l <= LMAX; l++) {
        const int sh = 3 * (LMAX - l);
        size_t cnt = 0;
        uint32_t prev = 0xffffffffu;
        for (size_t k = 0; k < S; k++) {
            const uint32_t c = scode[sord[k]] >> sh;
            if (c != prev) {
                cnt++;
                prev = c;
            }
        }
        if ((double)S / (double)cnt <= (double)leaf_target) {
            L = l;
            break;
        }
    }
    P.L = L;
    const int shL = 3 * (LMAX - L);
    P.src_rank.assign(S, 0);
    P.src_sorted.resize(S);
    for (size_t k = 0; k < S; k++) {
        P.src_rank[sord[k]] = (int64_t)k;
        P.src_sorted[k] = src[sord[k]];
    }
    std::vector<int64_t> tord(T);
    for (size_t i = 0; i < T; i++) tord[i] = (int64_t)i;
    std::stable_sort(tord.begin(), tord.end(), [&](int64_t a, int64_t b) { return tcode[a] < tcode[b]; });
    P.tgt_node.resize(T);
    P.tgt_pos.resize(T);
    for (size_t k = 0; k < T; k++) {
        P.tgt_node[k] = tnode[tord[k]];
        P.tgt_pos[k] = tpos[tord[k]];
    }
    // cells of every level: union of source and target codes
    P.lev_off.assign(L + 2, 0);
    P.W.assign(L + 1, 0);
    std::vector<std::vector<uint32_t>> codes(L + 1);
    std::vector<std::unordered_map<uint32_t, int32_t>> index(L + 1);
    for (int l = 0; l <= L; l++) {
        P.W[l] = 2.0 * P.H / (double)(1u << l);
        const int sh = 3 * (LMAX - l);
        std::vector<uint32_t> &cs = codes[l];
        cs.reserve(S / 8 + T);
        uint32_t prev = 0xffffffffu;
        for (size_t k = 0; k < S; k++) {
            const uint32_t c = scode[sord[k]] >> sh;
            if (c != prev) cs.push_back(c), prev = c;
        }
        for (size_t k = 0; k < T; k++) cs.push_back(tcode[tord[k]] >> sh);
        std::sort(cs.begin(), cs.end());
        cs.erase(std::unique(cs.begin(), cs.end()), cs.end());
        P.lev_off[l + 1] = P.lev_off[l] + (int)cs.size();
        for (size_t k = 0; k < cs.size(); k++) index[l][cs[k]] = P.lev_off[l] + (int32_t)k;
    }
    const int NC = P.lev_off[L + 1];
    P.cen.resize(NC);
    P.parent.assign(NC, -1);
    P.oct.assign(NC, 0);
    P.has_src.assign(NC, 0);
    P.has_tgt.assign(NC, 0);
    for (int l = 0; l <= L; l++)
        for (size_t k = 0; k < codes[l].size(); k++) {
            const int ci = P.lev_off[l] + (int)k;
            int ix, iy, iz;
            unmorton(codes[l][k], l, ix, iy, iz);
            P.cen[ci] = make_double4(corner[0] + (ix + 0.5) * P.W[l], corner[1] + (iy + 0.5) * P.W[l],
                                     corner[2] + (iz + 0.5) * P.W[l], (double)l);
            if (l > 0) {
                P.parent[ci] = index[l - 1].at(codes[l][k] >> 3);
                P.oct[ci] = (int32_t)(codes[l][k] & 7u);
            }
        }
    for (int l = 0; l <= L; l++) {
        const int sh = 3 * (LMAX - l);
        for (size_t k = 0; k < S; k++) P.has_src[index[l].at(scode[sord[k]] >> sh)] = 1;
        for (size_t k = 0; k < T; k++) P.has_tgt[index[l].at(tcode[tord[k]] >> sh)] = 1;
    }
    // children with sources
    P.ch_ptr.assign(NC + 1, 0);
    for (int c = 0; c < NC; c++)
        if (P.parent[c] >= 0 && P.has_src[c]) P.ch_ptr[P.parent[c] + 1]++;
    for (int c = 0; c < NC; c++) P.ch_ptr[c + 1] += P.ch_ptr[c];
    P.ch_idx.assign(P.ch_ptr[NC], 0);
    {
        std::vector<int32_t> fill(P.ch_ptr.begin(), P.ch_ptr.end() - 1);
        for (int c = 0; c < NC; c++)
            if (P.parent[c] >= 0 && P.has_src[c]) P.ch_idx[fill[P.parent[c]]++] = c;
    }
    // interaction lists
    P.il_ptr.assign(NC + 1, 0);
    std::vector<std::vector<std::pair<int32_t, int32_t>>> il(NC);
    for (int l = 2; l <= L; l++) {
        const int n = 1 << l;
        for (size_t k = 0; k < codes[l].size(); k++) {
            const int ci = P.lev_off[l] + (int)k;
            if (!P.has_tgt[ci]) continue;
            int ix, iy, iz;
            unmorton(codes[l][k], l, ix, iy, iz);
            const int px = ix >> 1, py = iy >> 1, pz = iz >> 1;
            for (int dx = -1; dx <= 1; dx++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dz = -1; dz <= 1; dz++) {
                        const int qx = px + dx, qy = py + dy, qz = pz + dz;
                        if (qx < 0 || qy < 0 || qz < 0 || qx >= n / 2 || qy >= n / 2 || qz >= n / 2) continue;
                        for (int o = 0; o < 8; o++) {
                            const int sx = 2 * qx + ((o >> 2) & 1), sy = 2 * qy + ((o >> 1) & 1), sz = 2 * qz + (o & 1);
                            // offset of the target from the source, c_t - c_s, in cell widths
                            const int ex = ix - sx, ey = iy - sy, ez = iz - sz;
                            if (std::abs(ex) <= 1 && std::abs(ey) <= 1 && std::abs(ez) <= 1) continue;
                            auto it = index[l].find(morton(sx, sy, sz));
                            if (it == index[l].end() || !P.has_src[it->second]) continue;
                            il[ci].push_back({it->second, (ex + 3) * 49 + (ey + 3) * 7 + (ez + 3)});
                        }
                    }
        }
    }
    for (int c = 0; c < NC; c++) P.il_ptr[c + 1] = P.il_ptr[c] + (int32_t)il[c].size();
    P.il_src.resize(P.il_ptr[NC]);
    P.il_off.resize(P.il_ptr[NC]);
    for (int c = 0; c < NC; c++)
        for (size_t k = 0; k < il[c].size(); k++) {
            P.il_src[P.il_ptr[c] + k] = il[c][k].first;
            P.il_off[P.il_ptr[c] + k] = il[c][k].second;
        }
    // leaves: ranges and near lists
    const int NL = P.lev_off[L + 1] - P.lev_off[L];
    P.leaf_sb.assign(NL, 0);
    P.leaf_se.assign(NL, 0);
    P.leaf_tb.assign(NL, 0);
    P.leaf_te.assign(NL, 0);
    for (size_t k = 0; k < S;) {
        const uint32_t c = scode[sord[k]] >> shL;
        size_t e = k;
        while (e < S && (scode[sord[e]] >> shL) == c) e++;
        const int li = index[L].at(c) - P.lev_off[L];
        P.leaf_sb[li] = (int32_t)k;
        P.leaf_se[li] = (int32_t)e;
        k = e;
    }
    P.tgt_leaf.resize(T);
    for (size_t k = 0; k < T;) {
        const uint32_t c = tcode[tord[k]] >> shL;
        size_t e = k;
        while (e < T && (tcode[tord[e]] >> shL) == c) e++;
        const int li = index[L].at(c) - P.lev_off[L];
        P.leaf_tb[li] = (int32_t)k;
        P.leaf_te[li] = (int32_t)e;
        for (size_t q = k; q < e; q++) P.tgt_leaf[q] = li;
        k = e;
    }
    P.near_ptr.assign(NL + 1, 0);
    std::vector<std::vector<int32_t>> nl(NL);
    {
        const int n = 1 << L;
        for (int li = 0; li < NL; li++) {
            const int ci = P.lev_off[L] + li;
            if (!P.has_tgt[ci]) continue;
            int ix, iy, iz;
            unmorton(codes[L][li], L, ix, iy, iz);
            for (int dx = -1; dx <= 1; dx++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dz = -1; dz <= 1; dz++) {
                        const int sx = ix + dx, sy = iy + dy, sz = iz + dz;
                        if (sx < 0 || sy < 0 || sz < 0 || sx >= n || sy >= n || sz >= n) continue;
                        auto it = index[L].find(morton(sx, sy, sz));
                        if (it == index[L].end() || !P.has_src[it->second]) continue;
                        nl[li].push_back(it->second - P.lev_off[L]);
                    }
        }
    }
    for (int li = 0; li < NL; li++) P.near_ptr[li + 1] = P.near_ptr[li] + (int32_t)nl[li].size();
    P.near_leaf.resize(P.near_ptr[NL]);
    for (int li = 0; li < NL; li++) std::copy(nl[li].begin(), nl[li].end(), P.near_leaf.begin() + P.near_ptr[li]);
    // per-level lists of source cells (upward) and target cells (downward)
    P.src_cells_lev_off.assign(L + 2, 0);
    P.tgt_cells_lev_off.assign(L + 2, 0);
    for (int l = 0; l <= L; l++) {
        for (int c = P.lev_off[l]; c < P.lev_off[l + 1]; c++) {
            if (P.has_src[c]) P.src_cells.push_back(c);
            if (P.has_tgt[c]) P.tgt_cells.push_back(c);
        }
        P.src_cells_lev_off[l + 1] = (int32_t)P.src_cells.size();
        P.tgt_cells_lev_off[l + 1] = (int32_t)P.tgt_cells.size();
    }
    // evaluation groups (level gl = L-1): the targets of one cell, contiguous in Morton order
    P.ls = std::max(2, L - 1);
    {
        const int gl = L - 1, shg = 3 * (LMAX - gl), n = 1 << gl, nL = 1 << L;
        P.tgt_oct.resize(T);
        P.tgt_grp.resize(T);
        P.u_ptr.assign(1, 0);
        for (size_t k = 0; k < T;) {
            const uint32_t gc = tcode[tord[k]] >> shg;
            size_t e = k;
            while (e < T && (tcode[tord[e]] >> shg) == gc) e++;
            const int g = (int)P.grp_tb.size();
            for (size_t q = k; q < e; q++) {
                P.tgt_oct[q] = (uint8_t)((tcode[tord[q]] >> shL) & 7u);
                P.tgt_grp[q] = g;
            }
            const int gci = index[gl].at(gc);
            P.grp_tb.push_back((int32_t)k);
            P.grp_te.push_back((int32_t)e);
            P.grp_lcell.push_back(L - 2 >= 2 ? P.parent[gci] : -1);
            std::vector<int32_t> far_cell;
            std::vector<uint8_t> far_mask;
            for (int q = P.il_ptr[gci]; q < P.il_ptr[gci + 1]; q++) {
                const int o = P.il_off[q], d = std::max({std::abs(o / 49 - 3), std::abs(o / 7 % 7 - 3), std::abs(o % 7 - 3)});
                (d >= 3 ? far_cell : P.u_cell).push_back(P.il_src[q]);
                (d >= 3 ? far_mask : P.u_mask).push_back(0xff);
            }
            int px, py, pz;
            unmorton(gc, gl, px, py, pz);
            for (int dx = -1; dx <= 1; dx++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dz = -1; dz <= 1; dz++) {
                        const int qx = px + dx, qy = py + dy, qz = pz + dz;
                        if (qx < 0 || qy < 0 || qz < 0 || qx >= n || qy >= n || qz >= n) continue;
                        for (int o = 0; o < 8; o++) {
                            const int sx = 2 * qx + ((o >> 2) & 1), sy = 2 * qy + ((o >> 1) & 1), sz = 2 * qz + (o & 1);
                            if (sx >= nL || sy >= nL || sz >= nL) continue;
                            auto it = index[L].find(morton(sx, sy, sz));
                            if (it == index[L].end() || !P.has_src[it->second]) continue;
                            uint8_t mn = 0, mf = 0;  // octants at distance 2, at distance 3
                            for (int t = 0; t < 8; t++) {
                                const int tx = 2 * px + ((t >> 2) & 1), ty = 2 * py + ((t >> 1) & 1), tz = 2 * pz + (t & 1);
                                const int d = std::max({std::abs(tx - sx), std::abs(ty - sy), std::abs(tz - sz)});
                                if (d >= 3)
                                    mf |= (uint8_t)(1u << t);
                                else if (d == 2)
                                    mn |= (uint8_t)(1u << t);
                            }
                            // an entry near to some of its octants is evaluated to degree p for all
                            // of them (splitting it would stage the cell twice)
                            if (mn) {
                                P.u_cell.push_back(it->second);
                                P.u_mask.push_back(mn | mf);
                            } else if (mf) {
                                far_cell.push_back(it->second);
                                far_mask.push_back(mf);
                            }
                        }
                    }
            P.u_far.push_back((int32_t)P.u_cell.size());
            P.u_cell.insert(P.u_cell.end(), far_cell.begin(), far_cell.end());
            P.u_mask.insert(P.u_mask.end(), far_mask.begin(), far_mask.end());
            P.u_ptr.push_back((int32_t)P.u_cell.size());
            k = e;
        }
    }
    // launch schedule of this tree
    P.Lmax = L;
    P.leaf_cell.resize(NL);
    P.leaf_ix.resize(NL);
    P.leaf_iy.resize(NL);
    P.leaf_iz.resize(NL);
    for (int li = 0; li < NL; li++) {
        P.leaf_cell[li] = P.lev_off[L] + li;
        unmorton(codes[L][li], L, P.leaf_ix[li], P.leaf_iy[li], P.leaf_iz[li]);
    }
    P.cell_iw.resize(NC);
    for (int ci = 0; ci < NC; ci++) P.cell_iw[ci] = 1.0 / P.W[(int)P.cen[ci].w];
    P.m2m_lev_off.assign(L + 2, 0);
    P.l2l_lev_off.assign(L + 2, 0);
    for (int l = 0; l <= L; l++) {
        for (int ci = P.lev_off[l]; ci < P.lev_off[l + 1]; ci++) {
            if (l >= 2 && l < L && P.has_src[ci]) P.m2m_cells.push_back(ci);
            if (l >= 3 && l < P.ls && P.has_tgt[ci]) P.l2l_cells.push_back(ci);
            if (l >= 2 && l < P.ls && P.has_tgt[ci]) P.m2l_cells.push_back(ci);
        }
        P.m2m_lev_off[l + 1] = (int32_t)P.m2m_cells.size();
        P.l2l_lev_off[l + 1] = (int32_t)P.l2l_cells.size();
    }
    // tables
    P.It.assign((size_t)343 * P.ni, make_double2(0, 0));
    for (int dx = -3; dx <= 3; dx++)
        for (int dy = -3; dy <= 3; dy++)
            for (int dz = -3; dz <= 3; dz++) {
                if (std::abs(dx) <= 1 && std::abs(dy) <= 1 && std::abs(dz) <= 1) continue;
                const int off = (dx + 3) * 49 + (dy + 3) * 7 + (dz + 3);
                irregular((double)dx, (double)dy, (double)dz, 2 * p, &P.It[(size_t)off * P.ni]);
            }
    P.Rc.assign((size_t)8 * P.nc, make_double2(0, 0));
    for (int o = 0; o < 8; o++) {
        const double d[3] = {((o >> 2) & 1) ? 0.25 : -0.25, ((o >> 1) & 1) ? 0.25 : -0.25, (o & 1) ? 0.25 : -0.25};
        regular(d[0], d[1], d[2], p, &P.Rc[(size_t)o * P.nc]);
    }
    return "";
}

// Host P2M of one leaf in the device's form: both column halves of p2m_accumulate_cols, then the
// factorial scale.
template <int P, unsigned MASK>
static void p2m_half_host(const FmmPlan &Pl, int li, int ci, const std::vector<double> (&qs)[2], std::vector<double2> &M)
{
    constexpr int NCo = fcount(P), NH = p2m_half_count<P>(MASK);
    const double4 c = Pl.cen[ci];
    const double iw = 1.0 / Pl.W[Pl.L];
    for (int s = 0; s < 2; s++) {
        double acc[2 * NH] = {};
        for (int k = Pl.leaf_sb[li]; k < Pl.leaf_se[li]; k++) {
            const double4 x = Pl.src_sorted[k];
            p2m_accumulate_cols<P, MASK>((x.x - c.x) * iw, (x.y - c.y) * iw, (x.z - c.z) * iw, qs[s][k], acc);
        }
        for (int m = 0; m <= P; m++) {
            if (!((MASK >> m) & 1)) continue;
            const int o = p2m_col_off<P>(MASK, m);
            for (int n = m; n <= P; n++)
                M[((size_t)ci * 2 + s) * NCo + fidx(n, m)] =
                    make_double2(acc[2 * (o + n - m)] * p2m_scale(n, m), acc[2 * (o + n - m) + 1] * p2m_scale(n, m));
        }
    }
}
template <int P>
static void p2m_leaf_host(const FmmPlan &Pl, int li, int ci, const std::vector<double> (&qs)[2], std::vector<double2> &M)
{
    p2m_half_host<P, p2m_half_mask<P>(0)>(Pl, li, ci, qs, M);
    p2m_half_host<P, p2m_half_mask<P>(1)>(Pl, li, ci, qs, M);
}

// Host evaluation of the whole FMM (reference for the device pass): qs[2][S] in the original
// source order; out[2][T] potentials sum q/r at the targets in the original target order.
void fmm_eval_host(const FmmPlan &P, const double *qu, const double *qv, std::vector<double> &pot_u,
                   std::vector<double> &pot_v, const std::vector<int32_t> &tnode_orig)
{
    const int NC = P.ncell(), nc = P.nc, p = P.p, L = P.L;
    const size_t S = P.src_sorted.size();
    std::vector<double> qs[2] = {std::vector<double>(S), std::vector<double>(S)};
    for (size_t i = 0; i < S; i++) {
        qs[0][P.src_rank[i]] = qu[i];
        qs[1][P.src_rank[i]] = qv[i];
    }
    std::vector<double2> M((size_t)NC * 2 * nc, make_double2(0, 0)), Lx((size_t)NC * 2 * nc, make_double2(0, 0));
    const int NL = P.lev_off[L + 1] - P.lev_off[L];
    for (int li = 0; li < NL; li++) {
        const int ci = P.lev_off[L] + li;
        switch (p) {
        case 4: p2m_leaf_host<4>(P, li, ci, qs, M); break;
        case 6: p2m_leaf_host<6>(P, li, ci, qs, M); break;
        case 8: p2m_leaf_host<8>(P, li, ci, qs, M); break;
        case 10: p2m_leaf_host<10>(P, li, ci, qs, M); break;
        default: break;
        }
    }
    for (int l = L - 1; l >= 2; l--)
        for (int ci = P.lev_off[l]; ci < P.lev_off[l + 1]; ci++)
            for (int k = P.ch_ptr[ci]; k < P.ch_ptr[ci + 1]; k++) {
                const int ch = P.ch_idx[k];
                for (int s = 0; s < 2; s++)
                    for (int n = 0; n <= p; n++)
                        for (int m = 0; m <= n; m++) {
                            const cplx v = m2m_coef(&M[((size_t)ch * 2 + s) * nc], &P.Rc[(size_t)P.oct[ch] * nc], n, m);
                            double2 &d = M[((size_t)ci * 2 + s) * nc + fidx(n, m)];
                            d.x += v.re;
                            d.y += v.im;
                        }
            }
    for (int ci = 0; ci < NC; ci++) {
        const int l = (int)P.cen[ci].w;
        if (l >= P.ls || !P.has_tgt[ci]) continue;
        const double iw = 1.0 / P.W[l];
        for (int k = P.il_ptr[ci]; k < P.il_ptr[ci + 1]; k++)
            for (int s = 0; s < 2; s++)
                for (int j = 0; j <= p; j++)
                    for (int kk = 0; kk <= j; kk++) {
                        const cplx v = m2l_coef(&M[((size_t)P.il_src[k] * 2 + s) * nc], &P.It[(size_t)P.il_off[k] * P.ni], p, j, kk);
                        double2 &d = Lx[((size_t)ci * 2 + s) * nc + fidx(j, kk)];
                        d.x += v.re * iw;
                        d.y += v.im * iw;
                    }
    }
    for (int l = 3; l < P.ls; l++)
        for (int ci = P.lev_off[l]; ci < P.lev_off[l + 1]; ci++) {
            if (!P.has_tgt[ci]) continue;
            const int pa = P.parent[ci];
            for (int s = 0; s < 2; s++)
                for (int j = 0; j <= p; j++)
                    for (int kk = 0; kk <= j; kk++) {
                        const cplx v = l2l_coef(&Lx[((size_t)pa * 2 + s) * nc], &P.Rc[(size_t)P.oct[ci] * nc], p, j, kk);
                        double2 &d = Lx[((size_t)ci * 2 + s) * nc + fidx(j, kk)];
                        d.x += v.re;
                        d.y += v.im;
                    }
        }
    const size_t T = P.tgt_pos.size();
    std::unordered_map<int32_t, size_t> pos_of;
    for (size_t k = 0; k < tnode_orig.size(); k++) pos_of[tnode_orig[k]] = k;
    pot_u.assign(T, 0.0);
    pot_v.assign(T, 0.0);
    for (size_t t = 0; t < T; t++) {
        const int li = P.tgt_leaf[t], g = P.tgt_grp[t];
        const double4 y = P.tgt_pos[t];
        double acc[2] = {0, 0};
        if (P.grp_lcell[g] >= 0) {
            const int ci = P.grp_lcell[g];
            const double4 c = P.cen[ci];
            const double iw = 1.0 / P.W[(int)c.w];
            for (int s = 0; s < 2; s++)
                acc[s] = l2p_eval(&Lx[((size_t)ci * 2 + s) * nc], p, (y.x - c.x) * iw, (y.y - c.y) * iw, (y.z - c.z) * iw);
        }
        for (int k = P.u_ptr[g]; k < P.u_ptr[g + 1]; k++) {
            if (!((P.u_mask[k] >> P.tgt_oct[t]) & 1)) continue;
            const int sc = P.u_cell[k];
            const double4 c = P.cen[sc];
            const double iw = 1.0 / P.W[(int)c.w];
            double su = 0, sv = 0;
            m2p_eval2(&M[(size_t)sc * 2 * nc], &M[((size_t)sc * 2 + 1) * nc], k < P.u_far[g] ? p : P.pf,
                      (y.x - c.x) * iw, (y.y - c.y) * iw, (y.z - c.z) * iw, su, sv);
            acc[0] += su * iw;
            acc[1] += sv * iw;
        }
        for (int k = P.near_ptr[li]; k < P.near_ptr[li + 1]; k++) {
            const int sl = P.near_leaf[k];
            for (int q = P.leaf_sb[sl]; q < P.leaf_se[sl]; q++) {
                const double4 x = P.src_sorted[q];
                const double dx = y.x - x.x, dy = y.y - x.y, dz = y.z - x.z;
                const double ir = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
                acc[0] += qs[0][q] * ir;
                acc[1] += qs[1][q] * ir;
            }
        }
        const size_t o = pos_of.at(P.tgt_node[t]);
        pot_u[o] = acc[0];
        pot_v[o] = acc[1];
    }
}


// Several independent trees side by side in one device plan: cells, leaves, sources, targets,
// groups and lists are concatenated with offsets; src_glob[k][s] is the global original index of
// part k's source s (the charge kernels address sources by that index).
void fmm_merge(std::vector<FmmPlan> &parts, const std::vector<std::vector<int64_t>> &src_glob, int64_t S_total,
               FmmPlan &M)
{
    const FmmPlan &P0 = parts[0];
    M.p = P0.p;
    M.pf = P0.pf;
    M.nc = P0.nc;
    M.ni = P0.ni;
    M.It = P0.It;
    M.Rc = P0.Rc;
    M.Lmax = 0;
    for (auto &P : parts) M.Lmax = std::max(M.Lmax, P.Lmax);
    M.src_rank.assign(S_total, -1);
    M.m2m_lev_off.assign(M.Lmax + 2, 0);
    M.l2l_lev_off.assign(M.Lmax + 2, 0);
    M.ch_ptr.assign(1, 0);
    M.il_ptr.assign(1, 0);
    M.near_ptr.assign(1, 0);
    M.u_ptr.assign(1, 0);
    std::vector<std::vector<int32_t>> m2m_lev(M.Lmax + 1), l2l_lev(M.Lmax + 1);
    for (size_t k = 0; k < parts.size(); k++) {
        const FmmPlan &P = parts[k];
        const int32_t co = (int32_t)M.cen.size(), lo = (int32_t)M.leaf_cell.size(), so = (int32_t)M.src_sorted.size(),
                      to = (int32_t)M.tgt_node.size(), go = (int32_t)M.grp_tb.size();
        const int NCk = P.ncell();
        for (int c = 0; c < NCk; c++) {
            M.cen.push_back(P.cen[c]);
            M.cell_iw.push_back(P.cell_iw[c]);
            M.parent.push_back(P.parent[c] >= 0 ? P.parent[c] + co : -1);
            M.oct.push_back(P.oct[c]);
            M.has_src.push_back(P.has_src[c]);
            M.has_tgt.push_back(P.has_tgt[c]);
            for (int e = P.ch_ptr[c]; e < P.ch_ptr[c + 1]; e++) M.ch_idx.push_back(P.ch_idx[e] + co);
            M.ch_ptr.push_back((int32_t)M.ch_idx.size());
            for (int e = P.il_ptr[c]; e < P.il_ptr[c + 1]; e++) {
                M.il_src.push_back(P.il_src[e] + co);
                M.il_off.push_back(P.il_off[e]);
            }
            M.il_ptr.push_back((int32_t)M.il_src.size());
        }
        const int NLk = (int)P.leaf_cell.size();
        for (int li = 0; li < NLk; li++) {
            M.leaf_cell.push_back(P.leaf_cell[li] + co);
            M.leaf_ix.push_back(P.leaf_ix[li]);
            M.leaf_iy.push_back(P.leaf_iy[li]);
            M.leaf_iz.push_back(P.leaf_iz[li]);
            M.leaf_sb.push_back(P.leaf_sb[li] + so);
            M.leaf_se.push_back(P.leaf_se[li] + so);
            M.leaf_tb.push_back(P.leaf_tb[li] + to);
            M.leaf_te.push_back(P.leaf_te[li] + to);
            for (int e = P.near_ptr[li]; e < P.near_ptr[li + 1]; e++) M.near_leaf.push_back(P.near_leaf[e] + lo);
            M.near_ptr.push_back((int32_t)M.near_leaf.size());
        }
        for (size_t s = 0; s < P.src_sorted.size(); s++) M.src_sorted.push_back(P.src_sorted[s]);
        for (size_t s = 0; s < P.src_rank.size(); s++) M.src_rank[src_glob[k][s]] = so + P.src_rank[s];
        for (size_t t = 0; t < P.tgt_node.size(); t++) {
            M.tgt_node.push_back(P.tgt_node[t]);
            M.tgt_pos.push_back(P.tgt_pos[t]);
            M.tgt_leaf.push_back(P.tgt_leaf[t] + lo);
            M.tgt_oct.push_back(P.tgt_oct[t]);
            M.tgt_grp.push_back(P.tgt_grp[t] + go);
        }
        for (size_t g = 0; g < P.grp_tb.size(); g++) {
            M.grp_tb.push_back(P.grp_tb[g] + to);
            M.grp_te.push_back(P.grp_te[g] + to);
            M.grp_lcell.push_back(P.grp_lcell[g] >= 0 ? P.grp_lcell[g] + co : -1);
            M.u_far.push_back(P.u_far[g] - P.u_ptr[g] + (int32_t)M.u_cell.size());
            for (int e = P.u_ptr[g]; e < P.u_ptr[g + 1]; e++) {
                M.u_cell.push_back(P.u_cell[e] + co);
                M.u_mask.push_back(P.u_mask[e]);
            }
            M.u_ptr.push_back((int32_t)M.u_cell.size());
        }
        for (int l = 0; l <= P.Lmax; l++) {
            for (int q = P.m2m_lev_off[l]; q < P.m2m_lev_off[l + 1]; q++) m2m_lev[l].push_back(P.m2m_cells[q] + co);
            for (int q = P.l2l_lev_off[l]; q < P.l2l_lev_off[l + 1]; q++) l2l_lev[l].push_back(P.l2l_cells[q] + co);
        }
        for (int32_t c : P.m2l_cells) M.m2l_cells.push_back(c + co);
    }
    for (int l = 0; l <= M.Lmax; l++) {
        M.m2m_cells.insert(M.m2m_cells.end(), m2m_lev[l].begin(), m2m_lev[l].end());
        M.m2m_lev_off[l + 1] = (int32_t)M.m2m_cells.size();
        M.l2l_cells.insert(M.l2l_cells.end(), l2l_lev[l].begin(), l2l_lev[l].end());
        M.l2l_lev_off[l + 1] = (int32_t)M.l2l_cells.size();
    }
}

}  // namespace fg

// =====================================================================================  device
namespace fg {

template <class T>
static T *dput(const std::vector<T> &h)
{
    void *d = nullptr;
    const size_t n = std::max<size_t>(1, h.size());
    if (hipMalloc(&d, n * sizeof(T)) != hipSuccess) throw std::string("hipMalloc failed (FMM plan)");
    if (!h.empty() && hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        throw std::string("hipMemcpy failed (FMM plan)");
    return (T *)d;
}


// Sum over each aligned group of 8 lanes, left in every lane of the group: data-parallel-primitive
// moves (swap within pairs, swap pairs within quads, mirror within the half row) of the two
// 32-bit halves, no LDS round trip.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int G>
__device__ __forceinline__ double sum_dpp(double v)
{
    static_assert(G == 4 || G == 8, "sums over aligned groups of 4 or 8 lanes");
    v += dpp_mov<0xB1>(v);                     // quad_perm [1,0,3,2]
    v += dpp_mov<0x4E>(v);                     // quad_perm [2,3,0,1]
    if constexpr (G == 8) v += dpp_mov<0x141>(v);  // row_half_mirror
    return v;
}
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, true));
}
template <int G>
__device__ __forceinline__ float sum_dpp(float v)
{
    static_assert(G == 4 || G == 8, "sums over aligned groups of 4 or 8 lanes");
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    if constexpr (G == 8) v += dpp_mov<0x141>(v);
    return v;
}

// P2M from the element values (the charges are linear in them, as in the near field): source s of
// a leaf carries its position and weight w0 (pos[s]) and code[s], a tetrahedron t >= 0, whose charges
// are w0 qe[t], or ~k for Gauss point k = 4 f + g of triangle f, whose charges are
// w0 qe[e] + tw[k].x qe[e + 1] + tw[k].y qe[e + 2] with e = Tm + 3 f (the vertex values u_a . n and
// weights dMs w_g lambda_a(g)). FP64 (S4 = double4, positions absolute, F = double) or FP32 (S4 =
// float4, positions relative to the leaf centre in cell widths and weights times mscale,
// F = float). G lanes per leaf, each a strided subset of its sources, the next source loaded
// while the current one is accumulated; both charge sets at once, columns of MASK only (two
// launches, one per half of the columns), combined by cross-lane sums.
template <class S4> struct P2mSrc;
template <> struct P2mSrc<double4> { typedef double2 W2; };
template <> struct P2mSrc<float4> { typedef float2 W2; };

template <int P, int G, unsigned MASK, class F, class S4>
__global__ __launch_bounds__(64) void k_fmm_p2m(int nleaf, int Tm, const int32_t *__restrict__ leaf_cell,
                                                const double4 *__restrict__ cen, const int32_t *__restrict__ sb,
                                                const int32_t *__restrict__ se, const S4 *__restrict__ pos,
                                                const int32_t *__restrict__ code,
                                                const typename P2mSrc<S4>::W2 *__restrict__ tw,
                                                const double2 *__restrict__ qe, F *__restrict__ M)
{
    constexpr int NCo = fcount(P), NH = p2m_half_count<P>(MASK);
    constexpr bool REL = std::is_same<S4, float4>::value;  // positions already relative, in cell widths
    const int lane = threadIdx.x & (G - 1);
    const int li = (blockIdx.x * 64 + threadIdx.x) / G;
    const bool act = li < nleaf;
    const int cell = act ? leaf_cell[li] : 0;
    F au[2 * NH], av[2 * NH];
#pragma unroll
    for (int k = 0; k < 2 * NH; k++) au[k] = av[k] = F(0);
    const double4 c = cen[cell];
    const double iw = c.w;
    const int kb = act ? sb[li] + lane : 0, ke = act ? se[li] : 0;
    S4 x = kb < ke ? pos[kb] : S4{0, 0, 0, 0};
    int cd = kb < ke ? code[kb] : 0;
    for (int k = kb; k < ke; k += G) {
        const S4 xc = x;
        const int cc = cd;
        if (k + G < ke) {  // the next source, loaded under this one's work
            x = pos[k + G];
            cd = code[k + G];
        }
        F qu, qv;
        if (cc >= 0) {
            const double2 e = qe[cc];
            qu = xc.w * (F)e.x;
            qv = xc.w * (F)e.y;
        } else {
            const int gk = ~cc, e0 = Tm + 3 * (gk >> 2);
            const auto w12 = tw[gk];
            const double2 a = qe[e0], b = qe[e0 + 1], d = qe[e0 + 2];
            qu = xc.w * (F)a.x + w12.x * (F)b.x + w12.y * (F)d.x;
            qv = xc.w * (F)a.y + w12.x * (F)b.y + w12.y * (F)d.y;
        }
        F rx, ry, rz;
        if constexpr (REL) {
            rx = xc.x;
            ry = xc.y;
            rz = xc.z;
        } else {
            rx = (xc.x - c.x) * iw;
            ry = (xc.y - c.y) * iw;
            rz = (xc.z - c.z) * iw;
        }
        p2m_accumulate2_cols<P, MASK, F>(rx, ry, rz, qu, qv, au, av);
    }
    static_assert(G == 4 || G == 8, "the cross-lane sum spans 4 or 8 lanes");
#pragma unroll
    for (int k = 0; k < 2 * NH; k++) {
        au[k] = sum_dpp<G>(au[k]);
        av[k] = sum_dpp<G>(av[k]);
    }
    if (act && lane == 0) {
        F *du = M + (size_t)cell * 4 * NCo, *dv = du + 2 * NCo;
#pragma unroll
        for (int m = 0; m <= P; m++) {
            if (!((MASK >> m) & 1)) continue;
            const int o = p2m_col_off<P>(MASK, m);
#pragma unroll
            for (int n = m; n <= P; n++) {
                const F sc = (F)p2m_scale(n, m);
                const int i = 2 * (o + n - m);
                du[2 * fidx(n, m)] = au[i] * sc;
                du[2 * fidx(n, m) + 1] = au[i + 1] * sc;
                dv[2 * fidx(n, m)] = av[i] * sc;
                dv[2 * fidx(n, m) + 1] = av[i + 1] * sc;
            }
        }
    }
}

// The P2M source records from the sources in tree order: position (and, in FP32, relative to the
// leaf centre in cell widths), weight w0, code and the triangle weights (src_w: the near field's
// per-source weights, (w_g, 0, 0) for a tetrahedron point and dMs w_g lambda_a(g) for a triangle
// point). One thread per leaf.
template <class S4>
__global__ void k_p2m_records(int nleaf, int Tm, int64_t S5, const int32_t *__restrict__ leaf_cell,
                              const double4 *__restrict__ cen, const int32_t *__restrict__ sb,
                              const int32_t *__restrict__ se, const double4 *__restrict__ spos,
                              const double4 *__restrict__ src_w, double wscale, S4 *__restrict__ pos)
{
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nleaf) return;
    const double4 c = cen[leaf_cell[li]];
    for (int k = sb[li]; k < se[li]; k++) {
        const double4 y = spos[k];
        const double w = src_w[k].x * wscale;
        if constexpr (std::is_same<S4, float4>::value)
            pos[k] = make_float4((float)((y.x - c.x) * c.w), (float)((y.y - c.y) * c.w), (float)((y.z - c.z) * c.w),
                                 (float)w);
        else
            pos[k] = make_double4(y.x, y.y, y.z, w);
    }
}

// Codes (tetrahedron index, or ~(4 f + g) for Gauss point g of triangle f) and triangle weights,
// one thread per source in the original order (tetrahedron points first, then triangle points).
template <class W2>
__global__ void k_p2m_codes(int Tm, int Fm, const int64_t *__restrict__ rank, const double4 *__restrict__ src_w,
                            double wscale, int32_t *__restrict__ code, W2 *__restrict__ tw)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s < (int64_t)5 * Tm) {
        code[rank[s]] = (int32_t)(s / 5);
    } else if (s < (int64_t)5 * Tm + (int64_t)4 * Fm) {
        const int k = (int)(s - (int64_t)5 * Tm);
        const int64_t r = rank[s];
        code[r] = ~k;
        const double4 w = src_w[r];
        tw[k].x = w.y * wscale;
        tw[k].y = w.z * wscale;
    }
}

// Translations between expansions of degree P in real form: NR = 2 fcount(P) reals, real j being
// component j & 1 of coefficient j >> 1. M2M (K = 0) maps child to parent, L2L (K = 1) parent to
// child; each is a fixed NR x NR matrix per child octant (expansions are normalised by the cell
// width, so it is the same at every level). They are applied on the matrix cores as products
// Y = T X with v_mfma_f64_16x16x4f64 (A: lane l holds T[l%16][l/16] of a 16x4 tile; B: lane l
// holds X[l/16][l%16] of a 4x16 tile; D: lane l holds Y[l/16 + 4t][l%16], t < 4). T is stored in
// fragment order [octant][row tile][k-step][lane] (rows padded to 16, columns to 4). M2M couples
// output degree n to input degrees <= n and L2L to degrees >= n, so each row tile runs only the
// k-steps [kb, ke) its degrees need.
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
// The 16x16x4 matrix-core products in FP64 and in FP32 take the same A and B fragments; the four
// accumulators of lane l hold column l % 16 of D, rows l / 16 + 4 t (FP64) or 4 (l / 16) + t (FP32).
template <class F> struct Mfma;
template <> struct Mfma<double> {
    typedef f64x4 V;
    __device__ static V mma(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static constexpr int row(int kq, int t) { return kq + 4 * t; }
};
template <> struct Mfma<float> {
    typedef f32x4 V;
    __device__ static V mma(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static constexpr int row(int kq, int t) { return 4 * kq + t; }
};
template <int P> struct XL {
    static constexpr int NCo = fcount(P), NR = 2 * NCo, NT = (NR + 15) / 16, NKS = (NR + 3) / 4;
    static constexpr int frag_len() { return 8 * NT * NKS * 64; }
    static constexpr int deg(int j)
    {
        int n = 0;
        while (fidx(n + 1, 0) <= (j >> 1)) n++;
        return n;
    }
    // k-steps [kb, ke) of row tile R: the input reals whose degrees its output degrees couple to
    static constexpr int kb(int K, int R) { return K == 0 ? 0 : (2 * fidx(deg(16 * R), 0)) / 4; }
    static constexpr int ke(int K, int R)
    {
        return K == 0 ? (2 * fcount(deg(16 * R + 15 < NR ? 16 * R + 15 : NR - 1)) + 3) / 4 : NKS;
    }
    // Pair order (FP32 M2M of k_fmm_m2m_all): lane group g of k-step q takes input real
    // kpair(q, g) = 8 (q / 2) + 2 g + q % 2, so one 8-byte load per lane and 8-real block feeds two
    // k-steps; NKS2 steps cover the NR reals, ke2(R) those M2M row tile R needs
    static constexpr int NKS2 = 2 * ((NR + 7) / 8);
    static constexpr int kpair(int q, int g) { return 8 * (q / 2) + 2 * g + q % 2; }
    static constexpr int ke2(int R) { return 2 * ((ke(0, R) + 1) / 2); }
};

// One block per (16-column tile, charge set, row tile); a column is a parent cell, and wavefront
// w takes child octants 2w and 2w + 1 (short chains: the top levels have few cells, so a level
// takes as long as one block). M2M (K = 0): parent
// = sum over octants of T_o child_o, the four partial tiles summed through LDS and written to the
// parent. L2L (K = 1): child_o += T_o parent for each octant with a receiving child.
template <int P, int K, int R, class F>
__device__ __forceinline__ void xlate_tile(int n, const int32_t *__restrict__ cells, const int32_t *__restrict__ child,
                                           const F *__restrict__ Top, F *__restrict__ X,
                                           typename Mfma<F>::V (*red)[64])
{
    using L = XL<P>;
    using MF = Mfma<F>;
    using V = typename MF::V;
    constexpr int NR = L::NR, NT = L::NT, NKS = L::NKS, kb = L::kb(K, R), ke = L::ke(K, R);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    // the row tiles and charge sets of a column tile on one XCD: they read the same children
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int s = (b / NT) & 1, ct = b / (2 * NT);
    const int col = l & 15, kq = l >> 4;
    const int pi = ct * 16 + col;
    const bool act = pi < n;
    const int pc = act ? cells[pi] : 0;
    const F *A = Top + (size_t)R * NKS * 64 + l;
    F pv[NKS];  // L2L: the parent's fragments, the same for every octant
    if (K == 1) {
        const F *xs = X + ((size_t)pc * 2 + s) * NR;
#pragma unroll
        for (int ks = 0; ks < NKS; ks++) {
            if (ks < kb || ks >= ke) continue;
            const int k = 4 * ks + kq;
            pv[ks] = act && k < NR ? xs[k] : F(0);
        }
    }
    V acc = {F(0), F(0), F(0), F(0)};
    for (int oo = 0; oo < 2; oo++) {
        const int o = 2 * w + oo;
        const int ch = act ? child[(size_t)pi * 8 + o] : -1;
        const F *Ao = A + (size_t)o * NT * NKS * 64;
        F av[NKS], bv[NKS];
        const F *xs = X + ((size_t)(ch < 0 ? 0 : ch) * 2 + s) * NR;
#pragma unroll
        for (int ks = 0; ks < NKS; ks++) {
            if (ks < kb || ks >= ke) continue;
            const int k = 4 * ks + kq;
            av[ks] = Ao[ks * 64];
            bv[ks] = K == 1 ? pv[ks] : (ch >= 0 && k < NR ? xs[k] : F(0));
        }
        if (K == 1) acc = V{F(0), F(0), F(0), F(0)};
#pragma unroll
        for (int ks = 0; ks < NKS; ks++) {
            if (ks < kb || ks >= ke) continue;
            acc = MF::mma(av[ks], bv[ks], acc);
        }
        if (K == 1 && ch >= 0) {
            F *d = X + ((size_t)ch * 2 + s) * NR;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int r = 16 * R + MF::row(kq, t);
                if (r < NR) d[r] += acc[t];
            }
        }
    }
    if (K == 0) {
        if (w > 0) red[w - 1][l] = acc;
        __syncthreads();
        if (w == 0 && act) {
#pragma unroll
            for (int q = 0; q < 3; q++) acc += red[q][l];
            F *d = X + ((size_t)pc * 2 + s) * NR;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int r = 16 * R + MF::row(kq, t);
                if (r < NR) d[r] = acc[t];
            }
        }
    }
}

template <int P, int K, int R, class F>
__device__ __forceinline__ void xlate_tile_dispatch(int r, int n, const int32_t *__restrict__ cells,
                                                    const int32_t *__restrict__ child, const F *__restrict__ Top,
                                                    F *__restrict__ X, typename Mfma<F>::V (*red)[64])
{
    if constexpr (R < XL<P>::NT) {
        if (r == R)
            xlate_tile<P, K, R, F>(n, cells, child, Top, X, red);
        else
            xlate_tile_dispatch<P, K, R + 1, F>(r, n, cells, child, Top, X, red);
    }
}

template <int P, int K, class F>
__global__ __launch_bounds__(256) void k_fmm_xlate(int n, const int32_t *__restrict__ cells,
                                                   const int32_t *__restrict__ child, const F *__restrict__ Top,
                                                   F *__restrict__ X)
{
    constexpr int NT = XL<P>::NT;
    __shared__ typename Mfma<F>::V red[3][64];  // partial tiles of wavefronts 1..3 (M2M)
    const int R = xcd_block(blockIdx.x, gridDim.x) % NT;
    xlate_tile_dispatch<P, K, 0, F>(R, n, cells, child, Top, X, red);
}

// Row tiles R.. of the M2M below: acc[R] += T_o[R rows] child_o for the wavefront's two octants,
// over the k-steps of the input degrees that row tile's output degrees take.
template <int P, int R, class F, bool PAIR, int NK>
__device__ __forceinline__ void m2m_rows(const F *__restrict__ A, const F (&bv)[2][NK],
                                         typename Mfma<F>::V (&acc)[XL<P>::NT])
{
    if constexpr (R < XL<P>::NT) {
        using L = XL<P>;
        constexpr int NKS = NK, ke = PAIR ? L::ke2(R) : L::ke(0, R);
        F av[2][ke];
#pragma unroll
        for (int oo = 0; oo < 2; oo++)
#pragma unroll
            for (int ks = 0; ks < ke; ks++) av[oo][ks] = A[((size_t)oo * L::NT + R) * NKS * 64 + ks * 64];
#pragma unroll
        for (int oo = 0; oo < 2; oo++)
#pragma unroll
            for (int ks = 0; ks < ke; ks++) acc[R] = Mfma<F>::mma(av[oo][ks], bv[oo][ks], acc[R]);
        m2m_rows<P, R + 1, F, PAIR, NK>(A, bv, acc);
    }
}

constexpr int M2M_ALL_MIN = 2048;  // parents of a level from which k_fmm_m2m_all runs it

// M2M for the levels with many parents. One block per (16 parents, charge set) takes every row
// tile, so each child expansion is read once and every wavefront has a row tile's worth of
// matrix-core work per load: wavefront w loads the B fragments of octants 2w and 2w + 1 (column =
// parent, k-step = four input reals), then applies those octants' operators tile by tile. The four
// partial results of each row tile go through LDS; wavefront w sums and writes tiles w, w + 4.
template <int P, class F>
__global__ __launch_bounds__(256) void k_fmm_m2m_all(int n, const int32_t *__restrict__ cells,
                                                     const int32_t *__restrict__ child, const F *__restrict__ Top,
                                                     F *__restrict__ X)
{
    using L = XL<P>;
    using MF = Mfma<F>;
    using V = typename MF::V;
    constexpr int NR = L::NR, NT = L::NT, NKS = L::NKS;
    __shared__ V red[NT][4][64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int b = xcd_block(blockIdx.x, gridDim.x);  // the two charge sets of a column tile together
    const int s = b & 1, ct = b >> 1;
    const int col = l & 15, kq = l >> 4;
    const int pi = ct * 16 + col;
    const bool act = pi < n;
    // FP32: pair order, one 8-byte load per 8-real block (NR is even and each expansion starts on
    // an 8-byte boundary); FP64: one real per k-step
    constexpr bool PAIR = std::is_same<F, float>::value;
    constexpr int NK = PAIR ? L::NKS2 : NKS;
    F bv[2][NK];
#pragma unroll
    for (int oo = 0; oo < 2; oo++) {
        const int ch = act ? child[(size_t)pi * 8 + 2 * w + oo] : -1;
        const F *xs = X + ((size_t)(ch < 0 ? 0 : ch) * 2 + s) * NR;
        if constexpr (PAIR) {
            static_assert(NR % 2 == 0, "pairs never straddle the end of an expansion");
#pragma unroll
            for (int j = 0; j < NK / 2; j++) {
                const int k = 8 * j + 2 * kq;
                const float2 v = ch >= 0 && k < NR ? reinterpret_cast<const float2 *>(xs)[k / 2] : make_float2(0.f, 0.f);
                bv[oo][2 * j] = v.x;
                bv[oo][2 * j + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int ks = 0; ks < NKS; ks++) {
                const int k = 4 * ks + kq;
                bv[oo][ks] = ch >= 0 && k < NR ? xs[k] : F(0);
            }
        }
    }
    V acc[NT];
#pragma unroll
    for (int R = 0; R < NT; R++) acc[R] = V{F(0), F(0), F(0), F(0)};
    m2m_rows<P, 0, F, PAIR, NK>(Top + (size_t)2 * w * NT * NK * 64 + l, bv, acc);
#pragma unroll
    for (int R = 0; R < NT; R++) red[R][w][l] = acc[R];
    __syncthreads();
    if (!act) return;
    const int pc = cells[pi];
    F *d = X + ((size_t)pc * 2 + s) * NR;
    for (int R = w; R < NT; R += 4) {
        V a = red[R][0][l];
#pragma unroll
        for (int q = 1; q < 4; q++) a += red[R][q][l];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int r = 16 * R + MF::row(kq, t);
            if (r < NR) d[r] = a[t];
        }
    }
}

// M2L on the matrix cores. The translation from a source cell to a target cell at offset d (in
// cell widths, d = c_t - c_s) is a fixed NR x NR real matrix T_d up to the factor 1/W of the level
// (the expansions are width-normalised), built on the host from m2l_coef and stored per offset
// code in v_mfma_f64_16x16x4f64 fragment order. One block per chunk of up to M2L_CHUNK offsets of
// a tile of 64 target cells: for each offset the block stages T_d in LDS and wavefront w applies
// it to the source expansions of targets 16w..16w+15 (both charge sets, two column tiles, every
// row tile), gathering each target's source from the chunk's table (zero columns where there is
// none). The chunk's partial sums go to a buffer; k_fmm_m2l_reduce adds a tile's chunks in order
// and writes L = (1/W) sum (deterministic).
constexpr int M2L_CHUNK = 4;
template <int P, class F>
__global__ __launch_bounds__(256) void k_fmm_m2l_mfma(const int32_t *__restrict__ tile_chunk, int ntiles,
                                                      const int32_t *__restrict__ chunk_code,
                                                      const int32_t *__restrict__ chunk_src,
                                                      const F *__restrict__ Tm2l, const F *__restrict__ X,
                                                      F *__restrict__ part)
{
    using L = XL<P>;
    using MF = Mfma<F>;
    using V = typename MF::V;
    constexpr int NR = L::NR, NT = L::NT, NKS = L::NKS, OL = NT * NKS * 64;
    __shared__ F sA[OL];
    const int ch = blockIdx.x, tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int col = l & 15, kq = l >> 4;
    V acc[NT][2];
#pragma unroll
    for (int R = 0; R < NT; R++) acc[R][0] = acc[R][1] = V{F(0), F(0), F(0), F(0)};
    for (int i = 0; i < M2L_CHUNK; i++) {
        const int code = chunk_code[ch * M2L_CHUNK + i];
        if (code < 0) break;  // uniform: the chunk's offsets are packed first
        __syncthreads();
        const F *A = Tm2l + (size_t)code * OL;
        for (int e = tid; e < OL; e += 256) sA[e] = A[e];
        __syncthreads();
        const int src = chunk_src[((size_t)ch * M2L_CHUNK + i) * 64 + 16 * w + col];
        F bv[2][NKS];
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const F *xs = X + ((size_t)(src < 0 ? 0 : src) * 2 + s) * NR;
#pragma unroll
            for (int ks = 0; ks < NKS; ks++) {
                const int k = 4 * ks + kq;
                bv[s][ks] = src >= 0 && k < NR ? xs[k] : F(0);
            }
        }
#pragma unroll
        for (int R = 0; R < NT; R++)
#pragma unroll
            for (int ks = 0; ks < NKS; ks++) {
                const F a = sA[(R * NKS + ks) * 64 + l];
                acc[R][0] = MF::mma(a, bv[0][ks], acc[R][0]);
                acc[R][1] = MF::mma(a, bv[1][ks], acc[R][1]);
            }
    }
    // partial sums of this chunk: [chunk][target slot][set][NR]
    F *d = part + ((size_t)ch * 64 + 16 * w + col) * 2 * NR;
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int R = 0; R < NT; R++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int r = 16 * R + MF::row(kq, t);
                if (r < NR) d[s * NR + r] = acc[R][s][t];
            }
}

// L of every M2L target: (1/W) times the sum of its tile's chunk partials, in chunk order.
template <int P, class F>
__global__ __launch_bounds__(256) void k_fmm_m2l_reduce(int ntiles, const int32_t *__restrict__ tile_cell,
                                                        const int32_t *__restrict__ tile_chunk,
                                                        const double4 *__restrict__ cen,
                                                        const F *__restrict__ part, F *__restrict__ Lx)
{
    constexpr int NR = 2 * fcount(P);
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t per_tile = 64 * 2 * NR;
    if (id >= (int64_t)ntiles * per_tile) return;
    const int tile = (int)(id / per_tile), rem = (int)(id - (int64_t)tile * per_tile);
    const int slot = rem / (2 * NR), sr = rem - slot * 2 * NR;
    const int cell = tile_cell[tile * 64 + slot];
    if (cell < 0) return;
    double sum = 0.0;
    for (int c = tile_chunk[tile]; c < tile_chunk[tile + 1]; c++) sum += part[((size_t)c * 64 + slot) * 2 * NR + sr];
    Lx[(size_t)cell * 2 * NR + sr] = (F)(sum * cen[cell].w);
}

__device__ inline double wsum64(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ void k_to_float(int64_t n, const double *__restrict__ a, float *__restrict__ b)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

// Near-field operator, per-source weights in tree order: a tetrahedron's Gauss point carries w_g
// (its charge is w_g times the element value -Ms div u); a triangle's Gauss point carries
// dMs w_g lambda_a(g) for its vertices a = 0..2 (the charge is their sum against u_a . n).
__global__ __launch_bounds__(256) void k_nf_weights(int Tm, int Fm, const double *__restrict__ geo,
                                                    const double *__restrict__ tg, const int64_t *__restrict__ rank,
                                                    double4 *__restrict__ wsrc)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s < (int64_t)5 * Tm) {
        const int t = (int)(s / 5), g = (int)(s % 5);
        wsrc[rank[s]] = make_double4(geo[(size_t)(12 + g) * Tm + t], 0, 0, 0);
    } else if (s < (int64_t)5 * Tm + (int64_t)4 * Fm) {
        const int64_t r = s - (int64_t)5 * Tm;
        const int f = (int)(r / 4), g = (int)(r % 4);
        const double w = tg[(size_t)(3 + g) * Fm + f] * tg[(size_t)7 * Fm + f];
        wsrc[rank[s]] = make_double4(w * tri_a(0, g), w * tri_a(1, g), w * tri_a(2, g), 0);
    }
}

// Near-field operator coefficients: one wavefront per target; lanes take the entries of the
// target's leaf. An entry is one element (a tetrahedron, or one vertex of a triangle) with the mask
// of its Gauss points that lie in leaves adjacent to the target's; its coefficient is
// sum over those points of weight / distance. The columns are the same for all targets of a leaf
// and are stored once per leaf.
__global__ __launch_bounds__(256) void k_nf_build(int T, const int32_t *__restrict__ tgt_tl,
                                                  const double4 *__restrict__ tpos,
                                                  const int64_t *__restrict__ lc_ptr,
                                                  const int32_t *__restrict__ lc_base,
                                                  const int32_t *__restrict__ lc_nk, const int64_t *__restrict__ rank,
                                                  const double4 *__restrict__ spos, const double4 *__restrict__ wsrc,
                                                  const int64_t *__restrict__ nf_ptr, double *__restrict__ nf_val)
{
    const int t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (t >= T) return;
    const int tl = tgt_tl[t];
    const double4 y = tpos[t];
    const int64_t b = lc_ptr[tl], n = lc_ptr[tl + 1] - b, pos = nf_ptr[t];
    for (int64_t j = lane; j < n; j += 64) {
        const int base = lc_base[b + j], nk = lc_nk[b + j];
        const int kind = nk >> 8;
        double v = 0;
        for (int g = 0; g < 5; g++) {
            if (!((nk >> g) & 1)) continue;
            const int64_t s = rank[(int64_t)base + g];
            const double4 x = spos[s], w = wsrc[s];
            const double dx = y.x - x.x, dy = y.y - x.y, dz = y.z - x.z;
            const double wk = kind <= 1 ? w.x : (kind == 2 ? w.y : w.z);
            v += wk / sqrt(dx * dx + dy * dy + dz * dz);
        }
        nf_val[pos + j] = v;
    }
}

// Near field + vertex corrections: one wavefront per leaf with targets. The leaf's targets share
// one column list (the runs of the adjacent leaves); each element value is gathered once and
// applied to up to four targets' coefficient rows at a time. Coefficients are FP64, or FP32 in the
// mixed-precision mode (products and sums stay FP64).
template <class TC>
__global__ __launch_bounds__(256) void k_nf_apply(int ntl, const int32_t *__restrict__ tl_tb,
                                                  const int32_t *__restrict__ tl_te, const int64_t *__restrict__ lc_ptr,
                                                  const int32_t *__restrict__ lc_col, const int64_t *__restrict__ nf_ptr,
                                                  const TC *__restrict__ nf_val, const double2 *__restrict__ qe,
                                                  const int32_t *__restrict__ tgt_node, const double2 *__restrict__ corr,
                                                  double2 *__restrict__ pot)
{
    const int w = ((int)xcd_block(blockIdx.x, gridDim.x) * (int)blockDim.x + (int)threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (w >= ntl) return;
    const int tb = tl_tb[w], te = tl_te[w];
    const int64_t cb = lc_ptr[w];
    const int nc = (int)(lc_ptr[w + 1] - cb);
    for (int t0 = tb; t0 < te; t0 += 4) {
        const int nt = min(4, te - t0);
        int64_t vb[4];
#pragma unroll
        for (int i = 0; i < 4; i++) vb[i] = nf_ptr[t0 + min(i, nt - 1)];
        double au[4] = {0, 0, 0, 0}, av[4] = {0, 0, 0, 0};
        constexpr int U = 2;  // column chunks per iteration: both chunks' gathers are issued first
        for (int j0 = lane; j0 < nc; j0 += 64 * U) {
            int col[U];
#pragma unroll
            for (int u = 0; u < U; u++) col[u] = j0 + 64 * u < nc ? lc_col[cb + j0 + 64 * u] : -1;
            double2 q[U];
#pragma unroll
            for (int u = 0; u < U; u++) q[u] = col[u] >= 0 ? qe[col[u]] : make_double2(0.0, 0.0);
#pragma unroll
            for (int i = 0; i < 4; i++)
                if (i < nt) {
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const double v = col[u] >= 0 ? (double)nf_val[vb[i] + j0 + 64 * u] : 0.0;
                        au[i] += v * q[u].x;
                        av[i] += v * q[u].y;
                    }
                }
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (i >= nt) break;
            const double su = wsum64(au[i]), sv = wsum64(av[i]);
            if (lane == 0) {
                const double2 c = corr[tgt_node[t0 + i]];
                pot[t0 + i] = make_double2(su + c.x, sv + c.y);
            }
        }
    }
}

// Far field: multipole-to-target evaluation (M2P) per evaluation group, one wavefront per group.
// The wavefront's four rows of 16 lanes are (target row, cell slot) pairs: with nt targets in the
// current chunk of at most 64, nr = ceil(nt / 16) rows cover the targets and cps = 4 / nr cells
// of the group's list are evaluated at once, one per slot. A row holds its cell's two expansions
// in VGPRs (row-broadcast layout below), loaded at each step; every product with a coefficient is
// a DPP row-broadcast fused multiply-add, so no lane re-reads the coefficients from memory or LDS.
// Entries from u_far[g] on are evaluated to degree PF. The cell slots of each target are summed in
// slot order, then the L2P of the group's level-(L-2) ancestor is added and phi, phi_v are
// written.

// Row-broadcast fused multiply-add: acc += c[lane L of this lane's row of 16] * x, through DPP
// row_newbcast on src0 of v_fmac_f64 / v_fmac_f32 (gfx950), written only in the lanes of the banks
// (groups of 4 lanes of a row) set in BM. Rules for the callers: EXEC is full wherever these run
// (the broadcasting lane must be active), and a VALU write of a coefficient register is separated
// from its first DPP read by row_fence (two wait states).
template <int L, int BM = 15>
__device__ __forceinline__ void fmac_row(double &acc, double c, double x)
{
    asm("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:%4"
        : "+v"(acc)
        : "v"(c), "v"(x), "i"(L), "i"(BM));
}
template <int L, int BM = 15>
__device__ __forceinline__ void fmac_row(float &acc, float c, float x)
{
    asm("v_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:%4"
        : "+v"(acc)
        : "v"(c), "v"(x), "i"(L), "i"(BM));
}
// acc += (coefficient double K of this lane's cell) * x. H = 1: one cell per row, K held by lane
// K % 16 of the row in register K / 16. H = 2: one cell per half row (lanes 0-7, 8-15), K held by
// lane 8 h + K % 8 of half h in register K / 8; each half takes its own broadcast (two masked
// instructions).
template <int H, int K, class R, int NQ>
__device__ __forceinline__ void fmac_coef(R &acc, const R (&c)[NQ], R x)
{
    if constexpr (H == 1) {
        fmac_row<K % 16>(acc, c[K / 16], x);
    } else {
        fmac_row<K % 8, 0x3>(acc, c[K / 8], x);
        fmac_row<8 + K % 8, 0xc>(acc, c[K / 8], x);
    }
}
template <class R>
__device__ __forceinline__ void row_fence(R &c)
{
    asm volatile("s_nop 1" : "+v"(c));
}

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>());
        static_for<B + 1, E>(f);
    }
}

// M2P of the two expansions of a row's cell at one target per lane (x in cell widths), the first
// to degree Q, the second to degree QV <= Q: su = sum M_u,n^m conj(I_n^m(x)), sv likewise.
// Coefficient k of a set (k = 2 fidx(n, m), + 1 for the imaginary part) sits where
// fmac_coef<H, k> reads it. The irregular harmonics are streamed column by column (diagonal
// I_m^m, then I_{m+1}^m and the three-term recurrence in n), the recurrence factors of each degree
// shared by all columns; the m = 0 column is real (M_n^0 and I_n^0 are real for real charges).
template <int Q, int QV, int H, class R, int NQ, int NQV>
__device__ __forceinline__ void m2p_rows(const R (&cu)[NQ], const R (&cv)[NQV], R x, R y, R z, R &su, R &sv)
{
    const R r2 = x * x + y * y + z * z;
    const R ir2 = fast_inv(r2), zr = z * ir2, zr2 = zr + zr, ir22 = ir2 + ir2;
    // tz[n] = (2n - 1) z / r^2 and un[n] = (n - 1)^2 / r^2; column m takes
    // un[n] - m^2 / r^2 = (n + m - 1)(n - m - 1) / r^2
    R tz[Q + 1], un[Q + 1];
    tz[1] = zr;
    un[1] = R(0);
    {
        R st = ir2;
#pragma unroll
        for (int n = 2; n <= Q; n++) {
            tz[n] = tz[n - 1] + zr2;
            un[n] = un[n - 1] + st;
            st += ir22;
        }
    }
    R dre = fast_rsqrt(r2), dim = R(0);
    R au = R(0), av = R(0);
    fmac_coef<H, 0>(au, cu, dre);
    fmac_coef<H, 0>(av, cv, dre);
    if constexpr (Q >= 1) {
        constexpr int k1 = 2 * fidx(1, 0);
        R p2 = dre, p1 = zr * dre;
        fmac_coef<H, k1>(au, cu, p1);
        if constexpr (QV >= 1) fmac_coef<H, k1>(av, cv, p1);
        static_for<2, Q + 1>([&](auto nc) {
            constexpr int n = decltype(nc)::value, k = 2 * fidx(n, 0);
            const R c = tz[n] * p1 - un[n] * p2;
            fmac_coef<H, k>(au, cu, c);
            if constexpr (n <= QV) fmac_coef<H, k>(av, cv, c);
            p2 = p1;
            p1 = c;
        });
    }
    R hu = R(0), hv = R(0);
    R fd = -ir2, vm = R(0), dv = ir2;  // fd = (1 - 2m) / r^2, vm = m^2 / r^2
    static_for<1, Q + 1>([&](auto mc) {
        constexpr int m = decltype(mc)::value, kd = 2 * fidx(m, m);
        const R nre = fd * (x * dre - y * dim), nim = fd * (x * dim + y * dre);
        fd -= ir22;
        vm += dv;
        dv += ir22;
        dre = nre;
        dim = nim;
        fmac_coef<H, kd>(hu, cu, dre);
        fmac_coef<H, kd + 1>(hu, cu, dim);
        if constexpr (m <= QV) {
            fmac_coef<H, kd>(hv, cv, dre);
            fmac_coef<H, kd + 1>(hv, cv, dim);
        }
        if constexpr (m + 1 <= Q) {
            constexpr int k1 = 2 * fidx(m + 1, m);
            R p2re = dre, p2im = dim, p1re = tz[m + 1] * dre, p1im = tz[m + 1] * dim;
            fmac_coef<H, k1>(hu, cu, p1re);
            fmac_coef<H, k1 + 1>(hu, cu, p1im);
            if constexpr (m + 1 <= QV) {
                fmac_coef<H, k1>(hv, cv, p1re);
                fmac_coef<H, k1 + 1>(hv, cv, p1im);
            }
            static_for<m + 2, Q + 1>([&](auto nc) {
                constexpr int n = decltype(nc)::value, k = 2 * fidx(n, m);
                const R bb = un[n] - vm;
                const R cre = tz[n] * p1re - bb * p2re, cim = tz[n] * p1im - bb * p2im;
                fmac_coef<H, k>(hu, cu, cre);
                fmac_coef<H, k + 1>(hu, cu, cim);
                if constexpr (n <= QV) {
                    fmac_coef<H, k>(hv, cv, cre);
                    fmac_coef<H, k + 1>(hv, cv, cim);
                }
                p2re = p1re;
                p2im = p1im;
                p1re = cre;
                p1im = cim;
            });
        }
    });
    su = au + (hu + hu);
    sv = av + (hv + hv);
}

// The source cell of a row (H = 1) or half row (H = 2) in a step: both expansions in
// row-broadcast layout (degree Q of order-P storage), its centre and 1/W, and whether it applies
// to this lane's target.
template <int P, int Q, int QV, int H, class R>
struct M2pCell {
    static constexpr int LPC = 16 / H, NK = 2 * fcount(Q), NQ = (NK + LPC - 1) / LPC;
    static constexpr int NKV = 2 * fcount(QV), NQV = (NKV + LPC - 1) / LPC;
    R raw[NQ], rawv[NQV];
    double4 c;
    bool on;
};

// List entries [b, e) of a group, evaluated to degree Q <= P (expansions stored to P): step s puts
// entry b + s cps + slot in each cell slot (other wavefronts on the SIMD hide the loads' latency).
// Every lane runs every step (invalid lanes evaluate at a dummy offset and add 0).
template <int P, int Q, int QV, int H, class R>
__device__ __forceinline__ void m2p_range(int b, int e, int cps, int slot, bool tv, int oc, double4 y,
                                          const int32_t *__restrict__ u_cell, const uint8_t *__restrict__ u_mask,
                                          const double4 *__restrict__ cen, const R *__restrict__ Md, double &pu,
                                          double &pv)
{
    using C = M2pCell<P, Q, QV, H, R>;
    constexpr int NCo = fcount(P), LPC = C::LPC, NK = C::NK, NQ = C::NQ, NKV = C::NKV, NQV = C::NQV;
    const int j = threadIdx.x & (LPC - 1);  // lane within the cell's lanes
    const int nst = (e - b + cps - 1) / cps;  // uniform over the wavefront
    auto fetch = [&](int s, C &cl) {
        const int ei = b + s * cps + slot;
        const bool ev = slot < cps && ei < e;
        const int ec = ev ? ei : b;
        const int cell = u_cell[ec];
        cl.c = cen[cell];
        cl.on = tv && ev && ((u_mask[ec] >> oc) & 1);
        const R *src = Md + (size_t)cell * 4 * NCo;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            // lanes past the last coefficient repeat it; no DPP read takes them
            cl.raw[q] = src[min(LPC * q + j, NK - 1)];
        }
#pragma unroll
        for (int q = 0; q < NQV; q++) cl.rawv[q] = src[2 * NCo + min(LPC * q + j, NKV - 1)];
    };
    auto eval = [&](C &cl) {
        R cf[NQ], cfv[NQV];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            cf[q] = cl.raw[q];
            row_fence(cf[q]);
        }
#pragma unroll
        for (int q = 0; q < NQV; q++) {
            cfv[q] = cl.rawv[q];
            row_fence(cfv[q]);
        }
        const double4 c = cl.c;
        const R x = cl.on ? (R)((y.x - c.x) * c.w) : R(2), yy = cl.on ? (R)((y.y - c.y) * c.w) : R(0),
                z = cl.on ? (R)((y.z - c.z) * c.w) : R(0);
        R su, sv;
        m2p_rows<Q, QV, H, R, NQ, NQV>(cf, cfv, x, yy, z, su, sv);
        pu += cl.on ? (double)su * c.w : 0.0;
        pv += cl.on ? (double)sv * c.w : 0.0;
    };
    for (int s = 0; s < nst; s++) {
        C A;
        fetch(s, A);
        eval(A);
    }
}

// H = 1: groups of any size, chunks of up to 64 targets as above. H = 2: groups of at most 8
// targets; each half row holds its own cell, lanes 8 h + i of row r evaluate cell slot 2 r + h for
// target i, so a step covers 8 cells (each coefficient product then takes two masked broadcasts).
// gid lists the groups of the launch.
template <int P, int PF, int PV, int H, class R>
__global__ __launch_bounds__(256) void k_fmm_m2p(int ng, const int32_t *__restrict__ gid,
                                                 const int32_t *__restrict__ grp_tb,
                                                 const int32_t *__restrict__ grp_te,
                                                 const int32_t *__restrict__ u_ptr, const int32_t *__restrict__ u_far,
                                                 const int32_t *__restrict__ u_cell,
                                                 const uint8_t *__restrict__ u_mask,
                                                 const uint8_t *__restrict__ tgt_oct, const double4 *__restrict__ tpos,
                                                 const double4 *__restrict__ cen, const R *__restrict__ M,
                                                 double2 *__restrict__ far)
{
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int gi = xcd_block(blockIdx.x, gridDim.x) * 4 + wv;
    if (gi >= ng) return;  // whole wavefront
    const int g = gid[gi];
    const int lane = threadIdx.x & 63, row = lane >> 4, j = lane & 15;
    const int tb = grp_tb[g], te = grp_te[g], ub = u_ptr[g], ue = u_ptr[g + 1], uf = u_far[g];
    for (int t0 = tb; t0 < te; t0 += 64) {
        const int nt = min(64, te - t0);
        // H = 1: nr target rows, cps = 4 / nr cells per step; H = 2: one target row of 8 lanes
        // per half, 8 cells per step
        const int nr = H == 1 ? (nt + 15) >> 4 : 1, cps = H == 1 ? 4 / nr : 8;
        const int tri = H == 1 ? row % nr : 0, slot = H == 1 ? row / nr : 2 * row + (j >> 3);
        const int ti = H == 1 ? 16 * tri + j : j & 7;
        const int t = t0 + ti;
        const bool tv = slot < cps && ti < nt;
        const double4 y = tpos[tv ? t : t0];
        const int oc = tv ? tgt_oct[t] : 0;
        double pu = 0, pv = 0;
        if constexpr (PF == P) {
            m2p_range<P, P, PV, H, R>(ub, ue, cps, slot, tv, oc, y, u_cell, u_mask, cen, M, pu, pv);
        } else {
            m2p_range<P, P, PV, H, R>(ub, uf, cps, slot, tv, oc, y, u_cell, u_mask, cen, M, pu, pv);
            m2p_range<P, PF, (PV < PF ? PV : PF), H, R>(uf, ue, cps, slot, tv, oc, y, u_cell, u_mask, cen, M, pu, pv);
        }
        // each target's cell slots, in slot order (all lanes take part in the shuffles)
        double su = 0, sv = 0;
        for (int q = 0; q < cps; q++) {
            const int src = H == 1 ? ((q * nr + tri) << 4) + j : ((q >> 1) << 4) + ((q & 1) << 3) + (j & 7);
            su += __shfl(pu, src, 64);
            sv += __shfl(pv, src, 64);
        }
        if (slot == 0 && tv) far[t] = make_double2(su, sv);
    }
}

// Per target: the multipole part (M2P), the local expansion of the group's ancestor (scaled like
// the multipoles), the scale undone, plus the near field and corrections, into the node's
// potentials (1/4 pi, src/fmm_demag.h:184-224).
template <int P, class R>
__global__ __launch_bounds__(256) void k_fmm_finish(int T, const int32_t *__restrict__ tgt_lc,
                                                    const double4 *__restrict__ tpos, const double4 *__restrict__ cen,
                                                    const R *__restrict__ Lx, const double2 *__restrict__ far,
                                                    const double2 *__restrict__ pot, const int32_t *__restrict__ tgt_node,
                                                    NodeRec *__restrict__ nxt, double mscale)
{
    constexpr int NCo = fcount(P), CW = 2 * NCo;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= T) return;
    const double2 f = far[t];
    double su = f.x, sv = f.y;
    const int lc = tgt_lc[t];
    if (lc >= 0) {
        const double4 c = cen[lc], y = tpos[t];
        const double iw = c.w;
        l2p_eval2_t<P>(Lx + (size_t)lc * 2 * CW, Lx + (size_t)lc * 2 * CW + 2 * NCo, (y.x - c.x) * iw,
                       (y.y - c.y) * iw, (y.z - c.z) * iw, su, sv);
    }
    su /= mscale;
    sv /= mscale;
    const double2 pn = pot[t];
    NodeRec *r = nxt + tgt_node[t];
    r->phi = (pn.x + su) / (4 * M_PI);
    r->phiv = (pn.y + sv) / (4 * M_PI);
}

void fmm_build_near(Ctx &c, FmmPlan &P);

// Groups of at most this many targets are evaluated with one source cell per half row.
constexpr int M2P_HALF_MAX = 8;

// Lane slots executed by k_fmm_m2p's mapping: per 64-target chunk of a group, one wavefront per
// step of cps cells (degree-P and far entries stepped separately).
static int64_t m2p_lane_slots(const FmmPlan &P)
{
    int64_t slots = 0;
    for (size_t g = 0; g < P.grp_tb.size(); g++)
        for (int t0 = P.grp_tb[g]; t0 < P.grp_te[g]; t0 += 64) {
            const int nt = std::min(64, P.grp_te[g] - t0);
            const int cps = P.grp_te[g] - P.grp_tb[g] <= M2P_HALF_MAX ? 8 : 4 / ((nt + 15) / 16);
            const int near = P.u_far[g] - P.u_ptr[g], far = P.u_ptr[g + 1] - P.u_far[g];
            slots += 64 * (int64_t)((near + cps - 1) / cps + (far + cps - 1) / cps);
        }
    return slots;
}

// The M2M (K = 0) or L2L (K = 1) operators of the 8 child octants in real form, in the fragment
// order k_fmm_xlate reads; built by applying the reference translations (m2m_coef, l2l_coef) to
// unit expansions, so device and host translations are the same linear maps.
template <int P>
static std::vector<double> xlate_pack(const FmmPlan &Pl, int K, bool pair)
{
    using L = XL<P>;
    constexpr int NR = L::NR, NCo = L::NCo, NT = L::NT;
    const int NKS = pair ? L::NKS2 : L::NKS;
    std::vector<double> out((size_t)8 * NT * NKS * 64, 0.0), T((size_t)NR * NR);
    std::vector<double2> e(NCo);
    for (int o = 0; o < 8; o++) {
        const double2 *Rd = &Pl.Rc[(size_t)o * Pl.nc];
        for (int j = 0; j < NR; j++) {
            std::fill(e.begin(), e.end(), make_double2(0, 0));
            e[j >> 1] = (j & 1) ? make_double2(0, 1) : make_double2(1, 0);
            for (int n = 0; n <= P; n++)
                for (int m = 0; m <= n; m++) {
                    const cplx v = K == 0 ? m2m_coef(e.data(), Rd, n, m) : l2l_coef(e.data(), Rd, P, n, m);
                    T[(size_t)(2 * fidx(n, m)) * NR + j] = v.re;
                    T[(size_t)(2 * fidx(n, m) + 1) * NR + j] = v.im;
                }
        }
        for (int R = 0; R < NT; R++)
            for (int ks = 0; ks < NKS; ks++)
                for (int l = 0; l < 64; l++) {
                    const int r = 16 * R + (l & 15), k = pair ? L::kpair(ks, l >> 4) : 4 * ks + (l >> 4);
                    out[(((size_t)o * NT + R) * NKS + ks) * 64 + l] = r < NR && k < NR ? T[(size_t)r * NR + k] : 0.0;
                }
    }
    return out;
}

static std::vector<double> xlate_pack(const FmmPlan &Pl, int K, bool pair = false)
{
    switch (Pl.p) {
    case 4: return xlate_pack<4>(Pl, K, pair);
    case 6: return xlate_pack<6>(Pl, K, pair);
    case 8: return xlate_pack<8>(Pl, K, pair);
    default: return xlate_pack<10>(Pl, K, pair);
    }
}

// Launch schedules of k_fmm_xlate: M2M children of each parent by octant; L2L parents per level
// (distinct parents of the level's receiving cells) with their receiving children.
static void xlate_schedules(FmmPlan &P)
{
    P.m2m_child.assign(P.m2m_cells.size() * 8, -1);
    for (size_t k = 0; k < P.m2m_cells.size(); k++) {
        const int ci = P.m2m_cells[k];
        for (int e = P.ch_ptr[ci]; e < P.ch_ptr[ci + 1]; e++) {
            const int ch = P.ch_idx[e];
            P.m2m_child[k * 8 + P.oct[ch]] = ch;
        }
    }
    P.l2l_par.clear();
    P.l2l_child.clear();
    P.l2l_par_off.assign(P.Lmax + 2, 0);
    for (int l = 0; l <= P.Lmax; l++) {
        std::unordered_map<int32_t, int32_t> slot;
        for (int q = P.l2l_lev_off[l]; q < P.l2l_lev_off[l + 1]; q++) {
            const int ci = P.l2l_cells[q], pa = P.parent[ci];
            auto it = slot.find(pa);
            int s;
            if (it == slot.end()) {
                s = (int)P.l2l_par.size();
                slot[pa] = s;
                P.l2l_par.push_back(pa);
                P.l2l_child.insert(P.l2l_child.end(), 8, -1);
            } else {
                s = it->second;
            }
            P.l2l_child[(size_t)s * 8 + P.oct[ci]] = ci;
        }
        P.l2l_par_off[l + 1] = (int32_t)P.l2l_par.size();
    }
}

// The M2L operator of every offset code (dx, dy, dz in [-3, 3]: (dx+3) 49 + (dy+3) 7 + dz+3) in
// fragment order, from m2l_coef on unit expansions (the level's 1/W is applied in the reduction);
// codes of adjacent offsets stay zero.
template <int P>
static std::vector<double> m2l_pack(const FmmPlan &Pl)
{
    using L = XL<P>;
    constexpr int NR = L::NR, NCo = L::NCo, NT = L::NT, NKS = L::NKS, OL = NT * NKS * 64;
    std::vector<double> out((size_t)343 * OL, 0.0), T((size_t)NR * NR);
    std::vector<double2> e(NCo);
    for (int code = 0; code < 343; code++) {
        const int dx = code / 49 - 3, dy = code / 7 % 7 - 3, dz = code % 7 - 3;
        if (std::max({std::abs(dx), std::abs(dy), std::abs(dz)}) <= 1) continue;
        const double2 *It = &Pl.It[(size_t)code * Pl.ni];
        for (int j = 0; j < NR; j++) {
            std::fill(e.begin(), e.end(), make_double2(0, 0));
            e[j >> 1] = (j & 1) ? make_double2(0, 1) : make_double2(1, 0);
            for (int n = 0; n <= P; n++)
                for (int m = 0; m <= n; m++) {
                    const cplx v = m2l_coef(e.data(), It, P, n, m);
                    T[(size_t)(2 * fidx(n, m)) * NR + j] = v.re;
                    T[(size_t)(2 * fidx(n, m) + 1) * NR + j] = v.im;
                }
        }
        for (int R = 0; R < NT; R++)
            for (int ks = 0; ks < NKS; ks++)
                for (int l = 0; l < 64; l++) {
                    const int r = 16 * R + (l & 15), k = 4 * ks + (l >> 4);
                    out[(size_t)code * OL + ((size_t)R * NKS + ks) * 64 + l] =
                        r < NR && k < NR ? T[(size_t)r * NR + k] : 0.0;
                }
    }
    return out;
}
static std::vector<double> m2l_pack(const FmmPlan &Pl)
{
    switch (Pl.p) {
    case 4: return m2l_pack<4>(Pl);
    case 6: return m2l_pack<6>(Pl);
    case 8: return m2l_pack<8>(Pl);
    default: return m2l_pack<10>(Pl);
    }
}

// Tiles of 64 M2L targets of one octant class (so the tile's targets share most offsets), each
// tile's distinct offset codes in chunks of M2L_CHUNK with the source of every target.
static void m2l_schedule(FmmPlan &P)
{
    std::vector<int32_t> cls[8];
    for (int32_t ci : P.m2l_cells) cls[P.oct[ci] & 7].push_back(ci);
    P.m2l_tile_cell.clear();
    P.m2l_tile_chunk.assign(1, 0);
    P.m2l_chunk_code.clear();
    P.m2l_chunk_src.clear();
    for (int o = 0; o < 8; o++)
        for (size_t t0 = 0; t0 < cls[o].size(); t0 += 64) {
            const size_t nt = std::min<size_t>(64, cls[o].size() - t0);
            std::map<int32_t, std::vector<int32_t>> by_code;  // offset code -> source per target slot
            for (size_t s = 0; s < nt; s++) {
                const int ci = cls[o][t0 + s];
                P.m2l_tile_cell.push_back(ci);
                for (int e = P.il_ptr[ci]; e < P.il_ptr[ci + 1]; e++) {
                    auto &v = by_code[P.il_off[e]];
                    if (v.empty()) v.assign(64, -1);
                    v[s] = P.il_src[e];
                }
            }
            for (size_t s = nt; s < 64; s++) P.m2l_tile_cell.push_back(-1);
            int k = 0;
            for (auto &kv : by_code) {
                if (k % M2L_CHUNK == 0) {
                    P.m2l_chunk_code.insert(P.m2l_chunk_code.end(), M2L_CHUNK, -1);
                    P.m2l_chunk_src.insert(P.m2l_chunk_src.end(), (size_t)M2L_CHUNK * 64, -1);
                }
                const size_t c = P.m2l_chunk_code.size() / M2L_CHUNK - 1, i = k % M2L_CHUNK;
                P.m2l_chunk_code[c * M2L_CHUNK + i] = kv.first;
                std::copy(kv.second.begin(), kv.second.end(), P.m2l_chunk_src.begin() + (c * M2L_CHUNK + i) * 64);
                k++;
            }
            P.m2l_tile_chunk.push_back((int32_t)(P.m2l_chunk_code.size() / M2L_CHUNK));
        }
    P.m2l_ntiles = (int)(P.m2l_tile_chunk.size() - 1);
    P.m2l_nchunks = (int)(P.m2l_chunk_code.size() / M2L_CHUNK);
}

// ---- host side of the device pass ----
int fmm_create(Ctx &c)
{
    try {
        std::vector<double4> src(c.S);
        if (hipMemcpy(src.data(), c.src_pos, sizeof(double4) * c.S, hipMemcpyDeviceToHost) != hipSuccess)
            throw std::string("hipMemcpy src");
        std::vector<double4> p4(c.N);
        if (hipMemcpy(p4.data(), c.pos4, sizeof(double4) * c.N, hipMemcpyDeviceToHost) != hipSuccess)
            throw std::string("hipMemcpy pos");
        const int leaf = c.opt.fmm_leaf > 0 ? c.opt.fmm_leaf : 64;
        // one tree per instance (one instance unless the context is a batch), merged side by side
        const int K = (int)c.inst_node_off.size() - 1;
        std::vector<FmmPlan> parts(K);
        std::vector<std::vector<int64_t>> glob(K);
        for (int k = 0; k < K; k++) {
            std::vector<double4> s;
            std::vector<int64_t> &gl = glob[k];
            for (int64_t t = c.inst_mtet_off[k]; t < c.inst_mtet_off[k + 1]; t++)
                for (int g = 0; g < 5; g++) gl.push_back(5 * t + g);
            for (int64_t q = c.inst_mtri_off[k]; q < c.inst_mtri_off[k + 1]; q++)
                for (int g = 0; g < 4; g++) gl.push_back(5 * (int64_t)c.Tm + 4 * q + g);
            s.reserve(gl.size());
            for (int64_t o : gl) s.push_back(src[o]);
            std::vector<int32_t> tnode;
            std::vector<double4> tpos;
            for (int64_t i = c.inst_node_off[k]; i < c.inst_node_off[k + 1]; i++)
                if (c.h_mag[i]) {
                    tnode.push_back((int32_t)i);
                    tpos.push_back(p4[i]);
                }
            // default far degree: one degree in FP64, the lower far degree in mixed precision
            const int far = c.opt.fmm_far_order ? c.opt.fmm_far_order
                                                : c.opt.precision == FG_PREC_MIXED ? 0 : c.opt.fmm_order;
            std::string e = fmm_build_host(parts[k], s, tnode, tpos, c.opt.fmm_order, far, leaf);
            if (!e.empty()) throw e + " (instance " + std::to_string(k) + ")";
        }
        auto *P = new FmmPlan();
        c.fmm = P;
        fmm_merge(parts, glob, c.S, *P);
        parts.clear();
        for (int64_t r : P->src_rank)
            if (r < 0) throw std::string("FMM plan: a source belongs to no instance");
        // sources in tree order; the charge kernels scatter through src_rank
        if (hipMemcpy(c.src_pos, P->src_sorted.data(), sizeof(double4) * c.S, hipMemcpyHostToDevice) != hipSuccess)
            throw std::string("hipMemcpy sorted sources");
        if (hipMemcpy(c.src_rank, P->src_rank.data(), sizeof(int64_t) * c.S, hipMemcpyHostToDevice) != hipSuccess)
            throw std::string("hipMemcpy rank");
        {
            // device centres carry the inverse cell width in w
            std::vector<double4> cw(P->cen);
            for (size_t q = 0; q < cw.size(); q++) cw[q].w = P->cell_iw[q];
            P->d_cen = dput(cw);
        }
        P->d_leaf_sb = dput(P->leaf_sb);
        P->d_leaf_se = dput(P->leaf_se);
        P->d_grp_tb = dput(P->grp_tb);
        P->m2p_gid.clear();
        for (int pass = 0; pass < 2; pass++) {
            for (size_t g = 0; g < P->grp_tb.size(); g++)
                if ((P->grp_te[g] - P->grp_tb[g] <= M2P_HALF_MAX) == (pass == 1)) P->m2p_gid.push_back((int32_t)g);
            if (pass == 0) P->m2p_n1 = (int)P->m2p_gid.size();
        }
        P->d_m2p_gid = dput(P->m2p_gid);
        P->d_grp_te = dput(P->grp_te);
        P->d_grp_lcell = dput(P->grp_lcell);
        P->d_u_ptr = dput(P->u_ptr);
        P->d_u_far = dput(P->u_far);
        P->d_u_cell = dput(P->u_cell);
        P->d_u_mask = dput(P->u_mask);
        P->d_tgt_oct = dput(P->tgt_oct);
        P->d_tgt_node = dput(P->tgt_node);
        P->d_tgt_pos = dput(P->tgt_pos);
        {
            std::vector<int32_t> tlc(P->tgt_node.size(), -1);
            for (size_t g = 0; g < P->grp_tb.size(); g++)
                for (int t = P->grp_tb[g]; t < P->grp_te[g]; t++) tlc[t] = P->grp_lcell[g];
            P->d_tgt_lc = dput(tlc);
            P->d_far = dput(std::vector<double2>(std::max<size_t>(1, P->tgt_node.size()), make_double2(0, 0)));
        }
        P->d_leaf_cell = dput(P->leaf_cell);
        P->d_m2m_cells = dput(P->m2m_cells);
        P->mixed = c.opt.precision == FG_PREC_MIXED;
        // FP32 mode: charges of order charge_unit times the unit-width regular harmonics; the
        // expansions are scaled to order one by a power of two (exact in both directions)
        P->mscale = P->mixed ? std::ldexp(1.0, -(int)std::lround(std::log2(c.charge_unit))) : 1.0;
        // translation operators in the precision of the expansions
        auto put_op = [&](const std::vector<double> &v) -> void * {
            if (!P->mixed) return dput(v);
            return dput(std::vector<float>(v.begin(), v.end()));
        };
        const size_t fsz = P->mixed ? sizeof(float) : sizeof(double);
        P->d_Tm2m = put_op(xlate_pack(*P, 0));
        if (P->mixed) P->d_Tm2m_pair = put_op(xlate_pack(*P, 0, true));
        P->d_Tl2l = put_op(xlate_pack(*P, 1));
        xlate_schedules(*P);
        P->d_Tm2l = put_op(m2l_pack(*P));
        m2l_schedule(*P);
        P->d_m2l_tile_cell = dput(P->m2l_tile_cell);
        P->d_m2l_tile_chunk = dput(P->m2l_tile_chunk);
        P->d_m2l_chunk_code = dput(P->m2l_chunk_code);
        P->d_m2l_chunk_src = dput(P->m2l_chunk_src);
        {
            void *pp = nullptr;
            const size_t np = std::max<size_t>(1, (size_t)P->m2l_nchunks * 64 * 2 * 2 * P->nc);
            if (hipMalloc(&pp, np * fsz) != hipSuccess) throw std::string("hipMalloc M2L partials");
            P->d_m2l_part = pp;
        }
        P->d_m2m_child = dput(P->m2m_child);
        P->d_l2l_par = dput(P->l2l_par);
        P->d_l2l_child = dput(P->l2l_child);
        const size_t nexp = (size_t)P->ncell() * 2 * P->nc;
        void *d = nullptr;
        if (hipMalloc(&d, nexp * 2 * fsz) != hipSuccess) throw std::string("hipMalloc M");
        P->d_M = d;
        if (hipMalloc(&d, nexp * 2 * fsz) != hipSuccess) throw std::string("hipMalloc L");
        P->d_L = d;
        fmm_build_near(c, *P);
        // interaction counts (benchmark work models)
        for (size_t g = 0; g < P->grp_tb.size(); g++)
            for (int t = P->grp_tb[g]; t < P->grp_te[g]; t++)
                for (int k = P->u_ptr[g]; k < P->u_ptr[g + 1]; k++) {
                    const int on = (P->u_mask[k] >> P->tgt_oct[t]) & 1;
                    P->m2p_pairs += on;
                    if (k >= P->u_far[g]) P->m2p_far_pairs += on;
                }
        P->m2p_slots = m2p_lane_slots(*P);
        for (int32_t ci : P->m2l_cells) P->m2l_pairs += P->il_ptr[ci + 1] - P->il_ptr[ci];
        return FG_OK;
    } catch (const std::string &e) {
        c.err = e;
        return FG_ERR_ARG;
    }
}

// Near-field operator. For each element (a tetrahedron; a triangle expanded to its three vertices)
// and each target leaf adjacent to a leaf holding one of its Gauss points, one entry with the mask
// of the Gauss points in leaves adjacent to that target leaf: an element split over several leaves
// gives one entry, not one per leaf. The entries are grouped by target leaf (shared columns); the
// coefficients sum_g weight_g / |y - x_g| are computed once on the device.
void fmm_build_near(Ctx &c, FmmPlan &P)
{
    const int NL = (int)P.leaf_cell.size();
    const int64_t S = c.S;
    const int Tm = c.Tm, Fm = c.Fm;
    std::vector<int32_t> leaf_of(S);
    for (int li = 0; li < NL; li++)
        for (int k = P.leaf_sb[li]; k < P.leaf_se[li]; k++) leaf_of[k] = li;
    const std::vector<int32_t> &lx = P.leaf_ix, &ly = P.leaf_iy, &lz = P.leaf_iz;
    // target leaves and, per source leaf, the target leaves adjacent to it
    std::vector<int32_t> tl_of_leaf(NL, -1), tl_tb, tl_te;
    for (int li = 0; li < NL; li++)
        if (P.leaf_te[li] > P.leaf_tb[li]) {
            tl_of_leaf[li] = (int32_t)tl_tb.size();
            tl_tb.push_back(P.leaf_tb[li]);
            tl_te.push_back(P.leaf_te[li]);
        }
    const int ntl = (int)tl_tb.size();
    std::vector<int32_t> rn_ptr(NL + 1, 0), rn_idx;
    for (int li = 0; li < NL; li++)
        for (int e = P.near_ptr[li]; e < P.near_ptr[li + 1]; e++) rn_ptr[P.near_leaf[e] + 1]++;
    for (int li = 0; li < NL; li++) rn_ptr[li + 1] += rn_ptr[li];
    rn_idx.resize(rn_ptr[NL]);
    {
        std::vector<int32_t> fill(rn_ptr.begin(), rn_ptr.end() - 1);
        for (int li = 0; li < NL; li++)
            for (int e = P.near_ptr[li]; e < P.near_ptr[li + 1]; e++) rn_idx[fill[P.near_leaf[e]]++] = li;
    }
    // entries: (target leaf, column, base source index, mask | kind << 8), two passes (count, fill)
    const int64_t nelem = (int64_t)Tm + Fm;
    auto visit = [&](auto &&emit) {
        std::vector<int32_t> cand;
        for (int64_t e = 0; e < nelem; e++) {
            const bool tet = e < Tm;
            const int ng = tet ? 5 : 4;
            const int64_t base = tet ? 5 * e : 5 * (int64_t)Tm + 4 * (e - Tm);
            int lg[5];
            cand.clear();
            for (int g = 0; g < ng; g++) {
                lg[g] = leaf_of[P.src_rank[base + g]];
                for (int q = rn_ptr[lg[g]]; q < rn_ptr[lg[g] + 1]; q++) cand.push_back(rn_idx[q]);
            }
            std::sort(cand.begin(), cand.end());
            cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
            for (int tli : cand) {
                int mask = 0;
                for (int g = 0; g < ng; g++)
                    if (std::abs(lx[tli] - lx[lg[g]]) <= 1 && std::abs(ly[tli] - ly[lg[g]]) <= 1 &&
                        std::abs(lz[tli] - lz[lg[g]]) <= 1)
                        mask |= 1 << g;
                const int t = tl_of_leaf[tli];
                if (tet) emit(t, (int32_t)e, (int32_t)base, mask);
                else
                    for (int a = 0; a < 3; a++)
                        emit(t, (int32_t)(Tm + 3 * (e - Tm) + a), (int32_t)base, mask | ((1 + a) << 8));
            }
        }
    };
    std::vector<int64_t> lc_ptr(ntl + 1, 0);
    visit([&](int t, int32_t, int32_t, int) { lc_ptr[t + 1]++; });
    for (int t = 0; t < ntl; t++) lc_ptr[t + 1] += lc_ptr[t];
    const int64_t ncol = lc_ptr[ntl];
    std::vector<int32_t> lc_col(ncol), lc_base(ncol), lc_nk(ncol);
    {
        std::vector<int64_t> fill(lc_ptr.begin(), lc_ptr.end() - 1);
        visit([&](int t, int32_t col, int32_t base, int nk) {
            const int64_t q = fill[t]++;
            lc_col[q] = col;
            lc_base[q] = base;
            lc_nk[q] = nk;
        });
    }
    // coefficient rows: one per target, as long as its leaf's column list
    const int T = (int)P.tgt_node.size();
    std::vector<int32_t> tgt_tl(T);
    std::vector<int64_t> nf_ptr(T + 1, 0);
    for (int t = 0; t < T; t++) {
        const int tl = tl_of_leaf[P.tgt_leaf[t]];
        tgt_tl[t] = tl;
        nf_ptr[t + 1] = nf_ptr[t] + (lc_ptr[tl + 1] - lc_ptr[tl]);
    }
    P.nf_nnz = nf_ptr[T];
    P.ntl = ntl;
    P.lc_cols = ncol;
    P.d_tl_tb = dput(tl_tb);
    P.d_tl_te = dput(tl_te);
    P.d_lc_ptr = dput(lc_ptr);
    P.d_lc_col = dput(lc_col);
    int32_t *d_lc_base = dput(lc_base), *d_lc_nk = dput(lc_nk), *d_tgt_tl = dput(tgt_tl);
    P.d_nf_ptr = dput(nf_ptr);
    void *d = nullptr;
    if (hipMalloc(&d, std::max<int64_t>(1, P.nf_nnz) * sizeof(double)) != hipSuccess) throw std::string("hipMalloc near-field values");
    P.d_nf_val = (double *)d;
    if (hipMalloc(&d, sizeof(double4) * S) != hipSuccess) throw std::string("hipMalloc weights");
    double4 *wsrc = (double4 *)d;
    if (hipMalloc(&d, sizeof(double2) * std::max(1, T)) != hipSuccess) throw std::string("hipMalloc pot");
    P.d_pot = (double2 *)d;
    if (hipMalloc(&d, sizeof(double2) * ((size_t)Tm + 3 * (size_t)Fm)) != hipSuccess) throw std::string("hipMalloc qe");
    c.qe = (double2 *)d;
    k_nf_weights<<<(unsigned)((S + 255) / 256), 256, 0, c.stream>>>(Tm, Fm, c.geo, c.tri_geo, c.src_rank, wsrc);
    if (T > 0)
        k_nf_build<<<(unsigned)(((int64_t)T * 64 + 255) / 256), 256, 0, c.stream>>>(
            T, d_tgt_tl, P.d_tgt_pos, P.d_lc_ptr, d_lc_base, d_lc_nk, c.src_rank, c.src_pos, wsrc, P.d_nf_ptr,
            P.d_nf_val);
    {
        // P2M source records, from the same per-source weights
        const int NL = (int)P.leaf_cell.size();
        const size_t s4 = P.mixed ? sizeof(float4) : sizeof(double4), w2 = P.mixed ? sizeof(float2) : sizeof(double2);
        if (hipMalloc(&P.d_p2m_pos, std::max<size_t>(1, (size_t)S) * s4) != hipSuccess ||
            hipMalloc(&P.d_p2m_tw, std::max<size_t>(1, 4 * (size_t)Fm) * w2) != hipSuccess ||
            hipMalloc((void **)&P.d_p2m_code, std::max<size_t>(1, (size_t)S) * sizeof(int32_t)) != hipSuccess)
            throw std::string("hipMalloc P2M records");
        const unsigned nbs = (unsigned)((S + 255) / 256), nbl = (unsigned)((NL + 255) / 256);
        if (P.mixed) {
            k_p2m_codes<float2><<<nbs, 256, 0, c.stream>>>(Tm, Fm, c.src_rank, wsrc, P.mscale, P.d_p2m_code,
                                                          (float2 *)P.d_p2m_tw);
            if (NL > 0)
                k_p2m_records<float4><<<nbl, 256, 0, c.stream>>>(NL, Tm, S, P.d_leaf_cell, P.d_cen, P.d_leaf_sb,
                                                                P.d_leaf_se, c.src_pos, wsrc, P.mscale,
                                                                (float4 *)P.d_p2m_pos);
        } else {
            k_p2m_codes<double2><<<nbs, 256, 0, c.stream>>>(Tm, Fm, c.src_rank, wsrc, 1.0, P.d_p2m_code,
                                                           (double2 *)P.d_p2m_tw);
            if (NL > 0)
                k_p2m_records<double4><<<nbl, 256, 0, c.stream>>>(NL, Tm, S, P.d_leaf_cell, P.d_cen, P.d_leaf_sb,
                                                                 P.d_leaf_se, c.src_pos, wsrc, 1.0,
                                                                 (double4 *)P.d_p2m_pos);
        }
    }
    if (P.mixed && P.nf_nnz > 0) {
        if (hipMalloc(&d, P.nf_nnz * sizeof(float)) != hipSuccess) throw std::string("hipMalloc near-field values");
        P.d_nf_valf = (float *)d;
        k_to_float<<<4096, 256, 0, c.stream>>>(P.nf_nnz, P.d_nf_val, P.d_nf_valf);
    }
    const hipError_t err = hipStreamSynchronize(c.stream);
    void *tmp[] = {d_lc_base, d_lc_nk, d_tgt_tl, wsrc};
    for (void *p : tmp) (void)hipFree(p);
    if (err != hipSuccess) throw std::string("near-field build: ") + hipGetErrorString(err);
    if (P.d_nf_valf) {
        (void)hipFree(P.d_nf_val);
        P.d_nf_val = nullptr;
    }
}

void fmm_destroy(Ctx &c)
{
    FmmPlan *P = c.fmm;
    if (!P) return;
    void *ptrs[] = {P->d_cen, P->d_m2p_gid, P->d_leaf_sb, P->d_leaf_se, P->d_grp_tb, P->d_grp_te, P->d_grp_lcell, P->d_u_ptr,
                    P->d_u_far, P->d_u_cell, P->d_u_mask, P->d_tgt_oct, P->d_tgt_node, P->d_tgt_pos, P->d_tgt_lc, P->d_far,
                    P->d_Tm2m, P->d_Tm2m_pair, P->d_Tl2l, P->d_Tm2l, P->d_m2l_part, P->d_m2l_tile_cell, P->d_m2l_tile_chunk,
                    P->d_m2l_chunk_code, P->d_m2l_chunk_src, P->d_m2m_child, P->d_l2l_par, P->d_l2l_child, P->d_M, P->d_L, P->d_leaf_cell, P->d_m2m_cells,
                    P->d_nf_ptr, P->d_nf_val, P->d_p2m_pos, P->d_p2m_tw, P->d_p2m_code,
                    P->d_nf_valf, P->d_tl_tb, P->d_tl_te, P->d_lc_ptr, P->d_lc_col, P->d_pot, c.qe};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    c.qe = nullptr;
    delete P;
    c.fmm = nullptr;
}

// One FMM pass in the expansions' precision F: FP64, or FP32 (expansions times mscale, FP32
// translations on the matrix cores and FP32 evaluation) in the mixed-precision mode.
template <int PP, class F>
static void run_fmm(Ctx &c)
{
    FmmPlan &P = *c.fmm;
    const int NL = (int)P.leaf_cell.size();
    F *M = (F *)P.d_M, *Lx = (F *)P.d_L;
    // The near field reads the element values and corrections and writes potentials that only the
    // final sum reads, and M2L and L2L need the multipoles only: once M2M is done all three run on
    // the second stream beside M2P (memory latency and matrix-core work against VALU issue). Beside
    // P2M and M2M the near field slowed them more than it gained. Profiled runs keep one stream, so
    // each kernel's time is its own.
    const bool fork = !c.opt.profile && c.aux;
    hipStream_t ns = fork ? c.aux : c.stream;
    {
        Prof pr(c, "fmm_p2m");
        constexpr int G = 8;  // lanes per leaf (4 measured slower: longer per-lane source loops)
        const unsigned nb = (unsigned)((NL * G + 63) / 64);
        if (NL > 0) {
            if constexpr (std::is_same<F, float>::value) {
                k_fmm_p2m<PP, G, p2m_acc_half_mask<PP>(0), float, float4><<<nb, 64, 0, c.stream>>>(
                    NL, c.Tm, P.d_leaf_cell, P.d_cen, P.d_leaf_sb, P.d_leaf_se, (const float4 *)P.d_p2m_pos,
                    P.d_p2m_code, (const float2 *)P.d_p2m_tw, c.qe, M);
                k_fmm_p2m<PP, G, p2m_acc_half_mask<PP>(1), float, float4><<<nb, 64, 0, c.stream>>>(
                    NL, c.Tm, P.d_leaf_cell, P.d_cen, P.d_leaf_sb, P.d_leaf_se, (const float4 *)P.d_p2m_pos,
                    P.d_p2m_code, (const float2 *)P.d_p2m_tw, c.qe, M);
            } else {
                k_fmm_p2m<PP, G, p2m_half_mask<PP>(0), double, double4><<<nb, 64, 0, c.stream>>>(
                    NL, c.Tm, P.d_leaf_cell, P.d_cen, P.d_leaf_sb, P.d_leaf_se, (const double4 *)P.d_p2m_pos,
                    P.d_p2m_code, (const double2 *)P.d_p2m_tw, c.qe, M);
                k_fmm_p2m<PP, G, p2m_half_mask<PP>(1), double, double4><<<nb, 64, 0, c.stream>>>(
                    NL, c.Tm, P.d_leaf_cell, P.d_cen, P.d_leaf_sb, P.d_leaf_se, (const double4 *)P.d_p2m_pos,
                    P.d_p2m_code, (const double2 *)P.d_p2m_tw, c.qe, M);
            }
        }
    }
    {
        Prof pr(c, "fmm_m2m");
        for (int l = P.Lmax - 1; l >= 2; l--) {
            const int n = P.m2m_lev_off[l + 1] - P.m2m_lev_off[l];
            if (n == 0) continue;
            // a block per row tile where parents are few (a level then takes as long as one
            // block), all row tiles in one block where they are many
            if (n >= M2M_ALL_MIN)
                k_fmm_m2m_all<PP, F><<<(n + 15) / 16 * 2, 256, 0, c.stream>>>(
                    n, P.d_m2m_cells + P.m2m_lev_off[l], P.d_m2m_child + (size_t)8 * P.m2m_lev_off[l],
                    (const F *)(std::is_same<F, float>::value ? P.d_Tm2m_pair : P.d_Tm2m), M);
            else
                k_fmm_xlate<PP, 0, F><<<(n + 15) / 16 * 2 * XL<PP>::NT, 256, 0, c.stream>>>(
                    n, P.d_m2m_cells + P.m2m_lev_off[l], P.d_m2m_child + (size_t)8 * P.m2m_lev_off[l],
                    (const F *)P.d_Tm2m, M);
        }
    }
    // the near field and the downward pass (M2L, L2L) on the second stream once the multipoles are
    // complete, beside M2P
    if (fork) {
        if (hipEventRecord(c.ev_mid, c.stream) != hipSuccess || hipStreamWaitEvent(c.aux, c.ev_mid, 0) != hipSuccess)
            throw std::string("hipEventRecord / hipStreamWaitEvent (downward-pass fork)");
    }
    {
        Prof pr(c, "fmm_near");
        const unsigned nb = (unsigned)(((int64_t)P.ntl * 64 + 255) / 256);
        if (P.ntl > 0 && P.d_nf_valf)
            k_nf_apply<float><<<nb, 256, 0, ns>>>(P.ntl, P.d_tl_tb, P.d_tl_te, P.d_lc_ptr, P.d_lc_col, P.d_nf_ptr,
                                                  P.d_nf_valf, c.qe, P.d_tgt_node, c.corr, P.d_pot);
        else if (P.ntl > 0)
            k_nf_apply<double><<<nb, 256, 0, ns>>>(P.ntl, P.d_tl_tb, P.d_tl_te, P.d_lc_ptr, P.d_lc_col, P.d_nf_ptr,
                                                   P.d_nf_val, c.qe, P.d_tgt_node, c.corr, P.d_pot);
    }

    {
        Prof pr(c, "fmm_m2l");
        if (P.m2l_nchunks > 0)
            k_fmm_m2l_mfma<PP, F><<<P.m2l_nchunks, 256, 0, ns>>>(P.d_m2l_tile_chunk, P.m2l_ntiles,
                                                                P.d_m2l_chunk_code, P.d_m2l_chunk_src,
                                                                (const F *)P.d_Tm2l, M, (F *)P.d_m2l_part);
        if (P.m2l_ntiles > 0) {
            const int64_t n = (int64_t)P.m2l_ntiles * 64 * 2 * 2 * P.nc;
            k_fmm_m2l_reduce<PP, F><<<(int)((n + 255) / 256), 256, 0, ns>>>(
                P.m2l_ntiles, P.d_m2l_tile_cell, P.d_m2l_tile_chunk, P.d_cen, (const F *)P.d_m2l_part, Lx);
        }
    }
    {
        Prof pr(c, "fmm_l2l");
        for (int l = 3; l <= P.Lmax; l++) {
            const int n = P.l2l_par_off[l + 1] - P.l2l_par_off[l];
            if (n == 0) continue;
            k_fmm_xlate<PP, 1, F><<<(n + 15) / 16 * 2 * XL<PP>::NT, 256, 0, ns>>>(
                n, P.d_l2l_par + P.l2l_par_off[l], P.d_l2l_child + (size_t)8 * P.l2l_par_off[l],
                (const F *)P.d_Tl2l, Lx);
        }
    }
    if (fork && hipEventRecord(c.ev_join, c.aux) != hipSuccess) throw std::string("hipEventRecord (join)");
    {
        Prof pr(c, "fmm_m2p");
        const int ng = (int)P.grp_tb.size();
        if (ng > 0) {
            // groups of more than M2P_HALF_MAX targets with one cell per row, then the others with
            // one cell per half row
            auto go = [&](auto k1, auto k2) {
                const int n1 = P.m2p_n1, n2 = ng - P.m2p_n1;
                if (n1 > 0)
                    k1<<<(n1 + 3) / 4, 256, 0, c.stream>>>(n1, P.d_m2p_gid, P.d_grp_tb, P.d_grp_te, P.d_u_ptr,
                                                          P.d_u_far, P.d_u_cell, P.d_u_mask, P.d_tgt_oct, P.d_tgt_pos,
                                                          P.d_cen, M, P.d_far);
                if (n2 > 0)
                    k2<<<(n2 + 3) / 4, 256, 0, c.stream>>>(n2, P.d_m2p_gid + n1, P.d_grp_tb, P.d_grp_te, P.d_u_ptr,
                                                          P.d_u_far, P.d_u_cell, P.d_u_mask, P.d_tgt_oct, P.d_tgt_pos,
                                                          P.d_cen, M, P.d_far);
            };
            constexpr int PL = fmm_far_degree(PP);
            // the second expansion (charges of v) to degree fmm_v_degree in the FP32 mode
            constexpr int PV = std::is_same<F, float>::value ? fmm_v_degree(PP) : PP;
            if (P.pf != PP)
                go(k_fmm_m2p<PP, PL, PV, 1, F>, k_fmm_m2p<PP, PL, PV, 2, F>);
            else
                go(k_fmm_m2p<PP, PP, PV, 1, F>, k_fmm_m2p<PP, PP, PV, 2, F>);
        }
    }
    // join: the near field, M2L and L2L are in
    if (fork && hipStreamWaitEvent(c.stream, c.ev_join, 0) != hipSuccess)
        throw std::string("hipStreamWaitEvent (join)");
    {
        Prof pr(c, "fmm_finish");
        const int T = (int)P.tgt_node.size();
        if (T > 0)
            k_fmm_finish<PP, F><<<(T + 255) / 256, 256, 0, c.stream>>>(T, P.d_tgt_lc, P.d_tgt_pos, P.d_cen, Lx, P.d_far,
                                                                        P.d_pot, P.d_tgt_node, c.rec[1 - c.cur],
                                                                        P.mscale);
    }
}

void fmm_info(const Ctx &c, int64_t info[16])
{
    const FmmPlan &P = *c.fmm;
    for (int k = 0; k < 16; k++) info[k] = 0;
    info[0] = P.Lmax;
    info[1] = P.ncell();
    info[2] = (int64_t)P.leaf_cell.size();
    info[3] = (int64_t)P.grp_tb.size();
    info[4] = (int64_t)P.u_cell.size();
    info[5] = P.m2p_pairs;
    info[6] = P.m2l_pairs;
    info[7] = P.nf_nnz;
    info[8] = P.lc_cols;
    info[9] = (int64_t)P.tgt_node.size();
    info[10] = (int64_t)P.src_sorted.size();
    info[11] = P.p;
    info[12] = P.ntl;
    info[13] = (int64_t)P.m2m_cells.size();  // M2M parent cells
    info[14] = P.m2p_slots;
    info[15] = P.m2p_far_pairs;
}

void launch_fmm(Ctx &c)
{
    const bool f32 = c.fmm->mixed;
    switch (c.fmm->p) {
    case 4: f32 ? run_fmm<4, float>(c) : run_fmm<4, double>(c); break;
    case 6: f32 ? run_fmm<6, float>(c) : run_fmm<6, double>(c); break;
    case 8: f32 ? run_fmm<8, float>(c) : run_fmm<8, double>(c); break;
    case 10: f32 ? run_fmm<10, float>(c) : run_fmm<10, double>(c); break;
    default: break;
    }
}

}  // namespace fg

// ---- host reference of the FMM (tests) ----
extern "C" int fg_fmm_reference(const double *src_xyz, int64_t S, const double *tgt_xyz, int64_t T, const double *q_u,
                                const double *q_v, int order, int far_order, int leaf, double *pot_u, double *pot_v,
                                int *leaf_level)
{
    try {
        std::vector<double4> src(S), tpos(T);
        std::vector<int32_t> tnode(T);
        for (int64_t i = 0; i < S; i++) src[i] = make_double4(src_xyz[3 * i], src_xyz[3 * i + 1], src_xyz[3 * i + 2], 0);
        for (int64_t i = 0; i < T; i++) {
            tpos[i] = make_double4(tgt_xyz[3 * i], tgt_xyz[3 * i + 1], tgt_xyz[3 * i + 2], 0);
            tnode[i] = (int32_t)i;
        }
        FmmPlan P;
        std::string e = fmm_build_host(P, src, tnode, tpos, order, far_order, leaf);
        if (!e.empty()) return FG_ERR_ARG;
        std::vector<double> pu, pv;
        fmm_eval_host(P, q_u, q_v, pu, pv, tnode);
        std::copy(pu.begin(), pu.end(), pot_u);
        std::copy(pv.begin(), pv.end(), pot_v);
        if (leaf_level) *leaf_level = P.L;
        return FG_OK;
    } catch (...) {
        return FG_ERR_ARG;
    }
}
