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
// Device pass per call: P2M (leaves) -> M2M (upward) -> M2L (every level) -> L2L (downward) ->
// L2P + P2P fused per target.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "ctx.h"
#include "fmm_math.h"

using namespace fg;

struct FmmPlan {
    int p = 8, L = 2, nc = 0, ni = 0;
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
    int ls = 3;                        // first level evaluated multipole-to-target (M2P)
    std::vector<int32_t> m2p_ptr, m2p_cell;  // per leaf: source cells of the ancestors' lists, levels >= ls
    std::vector<int32_t> lcell;        // per leaf: ancestor at level ls-1 holding the local expansion, or -1
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
    int32_t *d_parent = nullptr, *d_oct = nullptr, *d_ch_ptr = nullptr, *d_ch_idx = nullptr;
    int32_t *d_il_ptr = nullptr, *d_il_src = nullptr, *d_il_off = nullptr;
    int32_t *d_leaf_sb = nullptr, *d_leaf_se = nullptr, *d_near_ptr = nullptr, *d_near_leaf = nullptr;
    int32_t *d_m2p_ptr = nullptr, *d_m2p_cell = nullptr, *d_lcell = nullptr;
    int32_t *d_tgt_node = nullptr, *d_tgt_leaf = nullptr;
    double4 *d_tgt_pos = nullptr;
    double2 *d_It = nullptr, *d_Rc = nullptr;
    double2 *d_M = nullptr, *d_L = nullptr;  // [ncell][2][nc]
    std::vector<int32_t> src_cells_lev_off, tgt_cells_lev_off;  // offsets into the lists
    std::vector<int32_t> src_cells, tgt_cells;
    int32_t *d_src_cells = nullptr, *d_tgt_cells = nullptr;
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

// Mean targets per cell at or below which a level is evaluated multipole-to-target
// (FG_FMM_M2P_TARGETS overrides; 0 forces M2L everywhere).
static int fmm_m2p_threshold()
{
    const char *e = std::getenv("FG_FMM_M2P_TARGETS");
    return e ? std::atoi(e) : 24;
}

// Host construction of the plan from source positions (original order) and target nodes.
std::string fmm_build_host(FmmPlan &P, const std::vector<double4> &src, const std::vector<int32_t> &tnode,
                           const std::vector<double4> &tpos, int p, int leaf_target, int m2p_threshold)
{
    if (p < 2 || p > 10 || (p & 1)) return "fmm_order must be even, 2..10";
    P.p = p;
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
    // evaluation split: levels with few targets per cell are evaluated multipole-to-target
    // (one interaction serves few targets there, a translation would cost more than the
    // evaluations it saves); coarser levels use M2L, L2L and L2P at level ls-1
    P.ls = L + 1;
    for (int l = 2; l <= L; l++) {
        const int ncl = P.tgt_cells_lev_off[l + 1] - P.tgt_cells_lev_off[l];
        if (ncl > 0 && (double)T / ncl <= (double)m2p_threshold) {
            P.ls = l;
            break;
        }
    }
    P.m2p_ptr.assign(NL + 1, 0);
    P.lcell.assign(NL, -1);
    {
        std::vector<int32_t> anc(L + 1);
        for (int li = 0; li < NL; li++) {
            const int ci = P.lev_off[L] + li;
            int a = ci;
            for (int l = L; l >= 0; l--) {
                anc[l] = a;
                a = l > 0 ? P.parent[a] : -1;
            }
            if (P.ls - 1 >= 2) P.lcell[li] = anc[P.ls - 1];
            int cnt = 0;
            if (P.has_tgt[ci])
                for (int l = P.ls; l <= L; l++) cnt += P.il_ptr[anc[l] + 1] - P.il_ptr[anc[l]];
            P.m2p_ptr[li + 1] = P.m2p_ptr[li] + cnt;
        }
        P.m2p_cell.resize(P.m2p_ptr[NL]);
        for (int li = 0; li < NL; li++) {
            const int ci = P.lev_off[L] + li;
            if (!P.has_tgt[ci]) continue;
            int a = ci;
            for (int l = L; l >= 0; l--) {
                anc[l] = a;
                a = l > 0 ? P.parent[a] : -1;
            }
            int pos = P.m2p_ptr[li];
            for (int l = P.ls; l <= L; l++)
                for (int k = P.il_ptr[anc[l]]; k < P.il_ptr[anc[l] + 1]; k++) P.m2p_cell[pos++] = P.il_src[k];
        }
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
    std::vector<double2> R(nc);
    const int NL = P.lev_off[L + 1] - P.lev_off[L];
    for (int li = 0; li < NL; li++) {
        const int ci = P.lev_off[L] + li;
        const double4 c = P.cen[ci];
        const double iw = 1.0 / P.W[L];
        for (int k = P.leaf_sb[li]; k < P.leaf_se[li]; k++) {
            const double4 x = P.src_sorted[k];
            regular((x.x - c.x) * iw, (x.y - c.y) * iw, (x.z - c.z) * iw, p, R.data());
            for (int s = 0; s < 2; s++)
                for (int q = 0; q < nc; q++) {
                    double2 &m = M[((size_t)ci * 2 + s) * nc + q];
                    m.x += qs[s][k] * R[q].x;
                    m.y += qs[s][k] * R[q].y;
                }
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
        if (l >= P.ls) continue;
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
        const int li = P.tgt_leaf[t];
        const double4 y = P.tgt_pos[t];
        double acc[2] = {0, 0};
        if (P.lcell[li] >= 0) {
            const int ci = P.lcell[li];
            const double4 c = P.cen[ci];
            const double iw = 1.0 / P.W[(int)c.w];
            for (int s = 0; s < 2; s++)
                acc[s] = l2p_eval(&Lx[((size_t)ci * 2 + s) * nc], p, (y.x - c.x) * iw, (y.y - c.y) * iw, (y.z - c.z) * iw);
        }
        for (int k = P.m2p_ptr[li]; k < P.m2p_ptr[li + 1]; k++) {
            const int sc = P.m2p_cell[k];
            const double4 c = P.cen[sc];
            const double iw = 1.0 / P.W[(int)c.w];
            double su = 0, sv = 0;
            m2p_eval2(&M[(size_t)sc * 2 * nc], &M[((size_t)sc * 2 + 1) * nc], p, (y.x - c.x) * iw, (y.y - c.y) * iw,
                      (y.z - c.z) * iw, su, sv);
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

__device__ inline double wsum64(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// P2M: a group of G lanes per source leaf, each lane a strided subset of the leaf's sources; the
// lanes' accumulators are combined with cross-lane sums. blockIdx.y = charge set.
template <int P, int G>
__global__ __launch_bounds__(64) void k_fmm_p2m(int nleaf, int cell0, double iw, const double4 *__restrict__ cen,
                                                const int32_t *__restrict__ sb, const int32_t *__restrict__ se,
                                                const double4 *__restrict__ spos, const double2 *__restrict__ q,
                                                double2 *__restrict__ M)
{
    constexpr int NCo = fcount(P);
    const int lane = threadIdx.x & (G - 1);
    const int li = (blockIdx.x * 64 + threadIdx.x) / G;
    const int set = blockIdx.y;
    double acc[2 * NCo];
#pragma unroll
    for (int k = 0; k < 2 * NCo; k++) acc[k] = 0.0;
    const bool act = li < nleaf;
    if (act) {
        const double4 c = cen[cell0 + li];
        for (int k = sb[li] + lane; k < se[li]; k += G) {
            const double4 x = spos[k];
            const double2 qq = q[k];
            p2m_accumulate<P>((x.x - c.x) * iw, (x.y - c.y) * iw, (x.z - c.z) * iw, set ? qq.y : qq.x, acc);
        }
    }
#pragma unroll
    for (int k = 0; k < 2 * NCo; k++) {
        double v = acc[k];
        for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        acc[k] = v;
    }
    if (act && lane == 0) {
        double2 *dst = M + ((size_t)(cell0 + li) * 2 + set) * NCo;
#pragma unroll
        for (int k = 0; k < NCo; k++) dst[k] = make_double2(acc[2 * k], acc[2 * k + 1]);
    }
}

// M2M: one thread per (cell with sources at this level, set, coefficient)
__global__ __launch_bounds__(256) void k_fmm_m2m(int ncells, const int32_t *__restrict__ cells, int p, int nc,
                                                 const int32_t *__restrict__ ch_ptr, const int32_t *__restrict__ ch_idx,
                                                 const int32_t *__restrict__ oct, const double2 *__restrict__ Rc,
                                                 double2 *__restrict__ M)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = tid % nc, set = (tid / nc) & 1, k = tid / (2 * nc);
    if (k >= ncells) return;
    int n = 0;
    while (fidx(n + 1, 0) <= q) n++;
    const int m = q - fidx(n, 0);
    const int ci = cells[k];
    double re = 0, im = 0;
    for (int e = ch_ptr[ci]; e < ch_ptr[ci + 1]; e++) {
        const int ch = ch_idx[e];
        const cplx v = m2m_coef(M + ((size_t)ch * 2 + set) * nc, Rc + (size_t)oct[ch] * nc, n, m);
        re += v.re;
        im += v.im;
    }
    M[((size_t)ci * 2 + set) * nc + q] = make_double2(re, im);
}

// M2L: one block per target cell (any level), lanes = coefficients x sets; the source expansion
// and the offset's irregular table are staged in LDS.
template <int P>
__global__ __launch_bounds__(256) void k_fmm_m2l(int ncells, const int32_t *__restrict__ cells,
                                                 const double4 *__restrict__ cen, const double *__restrict__ Wlev,
                                                 const int32_t *__restrict__ il_ptr, const int32_t *__restrict__ il_src,
                                                 const int32_t *__restrict__ il_off, const double2 *__restrict__ It,
                                                 const double2 *__restrict__ M, double2 *__restrict__ Lx)
{
    constexpr int NCo = fcount(P), NIo = fcount(2 * P);
    __shared__ double2 sM[2 * NCo];
    __shared__ double2 sI[NIo];
    const int ci = cells[blockIdx.x];
    const int t = threadIdx.x;
    const int q = t % NCo, set = t / NCo;
    const bool act = t < 2 * NCo;
    int j = 0;
    while (fidx(j + 1, 0) <= q) j++;
    const int k = q - fidx(j, 0);
    double re = 0, im = 0;
    for (int e = il_ptr[ci]; e < il_ptr[ci + 1]; e++) {
        const int sc = il_src[e];
        const double2 *src = M + (size_t)sc * 2 * NCo;
        for (int u = t; u < 2 * NCo; u += blockDim.x) sM[u] = src[u];
        const double2 *tab = It + (size_t)il_off[e] * NIo;
        for (int u = t; u < NIo; u += blockDim.x) sI[u] = tab[u];
        __syncthreads();
        if (act) {
            const cplx v = m2l_coef(sM + set * NCo, sI, P, j, k);
            re += v.re;
            im += v.im;
        }
        __syncthreads();
    }
    if (act) {
        const double iw = 1.0 / Wlev[(int)cen[ci].w];
        Lx[((size_t)ci * 2 + set) * NCo + q] = make_double2(re * iw, im * iw);
    }
}

// L2L: one thread per (cell with targets at this level, set, coefficient): L += shift(parent)
__global__ __launch_bounds__(256) void k_fmm_l2l(int ncells, const int32_t *__restrict__ cells, int p, int nc,
                                                 const int32_t *__restrict__ parent, const int32_t *__restrict__ oct,
                                                 const double2 *__restrict__ Rc, double2 *__restrict__ Lx)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = tid % nc, set = (tid / nc) & 1, kk = tid / (2 * nc);
    if (kk >= ncells) return;
    int j = 0;
    while (fidx(j + 1, 0) <= q) j++;
    const int k = q - fidx(j, 0);
    const int ci = cells[kk];
    const cplx v = l2l_coef(Lx + ((size_t)parent[ci] * 2 + set) * nc, Rc + (size_t)oct[ci] * nc, p, j, k);
    double2 &d = Lx[((size_t)ci * 2 + set) * nc + q];
    d.x += v.re;
    d.y += v.im;
}

// L2P + M2P + P2P: one wavefront per target; lanes stride over the sources of the adjacent leaves
// and over the cells evaluated multipole-to-target.
__global__ __launch_bounds__(256) void k_fmm_eval(int T, int p, int nc, const double *__restrict__ Wlev,
                                                  const int32_t *__restrict__ tgt_node,
                                                  const int32_t *__restrict__ tgt_leaf,
                                                  const double4 *__restrict__ tpos, const double4 *__restrict__ cen,
                                                  const double2 *__restrict__ Lx, const double2 *__restrict__ M,
                                                  const int32_t *__restrict__ lcell,
                                                  const int32_t *__restrict__ m2p_ptr,
                                                  const int32_t *__restrict__ m2p_cell,
                                                  const int32_t *__restrict__ near_ptr,
                                                  const int32_t *__restrict__ near_leaf,
                                                  const int32_t *__restrict__ sb, const int32_t *__restrict__ se,
                                                  const double4 *__restrict__ spos, const double2 *__restrict__ q,
                                                  const double2 *__restrict__ corr, NodeRec *__restrict__ nxt)
{
    const int t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (t >= T) return;
    const int li = tgt_leaf[t];
    const double4 y = tpos[t];
    double pu = 0, pv = 0;
    for (int e = near_ptr[li]; e < near_ptr[li + 1]; e++) {
        const int sl = near_leaf[e];
        for (int k = sb[sl] + lane; k < se[sl]; k += 64) {
            const double4 x = spos[k];
            const double dx = y.x - x.x, dy = y.y - x.y, dz = y.z - x.z;
            const double ir = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
            const double2 qq = q[k];
            pu += qq.x * ir;
            pv += qq.y * ir;
        }
    }
    for (int e = m2p_ptr[li] + lane; e < m2p_ptr[li + 1]; e += 64) {
        const int sc = m2p_cell[e];
        const double4 c = cen[sc];
        const double iw = 1.0 / Wlev[(int)c.w];
        double su = 0, sv = 0;
        m2p_eval2(M + (size_t)sc * 2 * nc, M + ((size_t)sc * 2 + 1) * nc, p, (y.x - c.x) * iw, (y.y - c.y) * iw,
                  (y.z - c.z) * iw, su, sv);
        pu += su * iw;
        pv += sv * iw;
    }
    const int lc = lcell[li];
    if (lc >= 0 && lane < 2) {
        const double4 c = cen[lc];
        const double iw = 1.0 / Wlev[(int)c.w];
        const double v = l2p_eval(Lx + ((size_t)lc * 2 + lane) * nc, p, (y.x - c.x) * iw, (y.y - c.y) * iw, (y.z - c.z) * iw);
        if (lane == 0) pu += v;
        else pv += v;
    }
    pu = wsum64(pu);
    pv = wsum64(pv);
    if (lane == 0) {
        const int node = tgt_node[t];
        const double2 c = corr[node];
        NodeRec r = nxt[node];
        r.phi = (pu + c.x) / (4 * M_PI);
        r.phiv = (pv + c.y) / (4 * M_PI);
        nxt[node] = r;
    }
}

__global__ void k_zero2(size_t n, double2 *__restrict__ a)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = make_double2(0, 0);
}

// ---- host side of the device pass ----
int fmm_create(Ctx &c)
{
    try {
        std::vector<double4> src(c.S);
        if (hipMemcpy(src.data(), c.src_pos, sizeof(double4) * c.S, hipMemcpyDeviceToHost) != hipSuccess)
            throw std::string("hipMemcpy src");
        std::vector<int32_t> tnode;
        std::vector<double4> tpos;
        std::vector<double4> p4(c.N);
        if (hipMemcpy(p4.data(), c.pos4, sizeof(double4) * c.N, hipMemcpyDeviceToHost) != hipSuccess)
            throw std::string("hipMemcpy pos");
        for (int i = 0; i < c.N; i++)
            if (c.h_mag[i]) {
                tnode.push_back(i);
                tpos.push_back(p4[i]);
            }
        auto *P = new FmmPlan();
        c.fmm = P;
        const int leaf = c.opt.fmm_leaf > 0 ? c.opt.fmm_leaf : 128;
        std::string e = fmm_build_host(*P, src, tnode, tpos, c.opt.fmm_order, leaf, fmm_m2p_threshold());
        if (!e.empty()) throw e;
        // sources in tree order; the charge kernels scatter through src_rank
        if (hipMemcpy(c.src_pos, P->src_sorted.data(), sizeof(double4) * c.S, hipMemcpyHostToDevice) != hipSuccess)
            throw std::string("hipMemcpy sorted sources");
        if (hipMemcpy(c.src_rank, P->src_rank.data(), sizeof(int64_t) * c.S, hipMemcpyHostToDevice) != hipSuccess)
            throw std::string("hipMemcpy rank");
        P->d_cen = dput(P->cen);
        P->d_parent = dput(P->parent);
        P->d_oct = dput(P->oct);
        P->d_ch_ptr = dput(P->ch_ptr);
        P->d_ch_idx = dput(P->ch_idx);
        P->d_il_ptr = dput(P->il_ptr);
        P->d_il_src = dput(P->il_src);
        P->d_il_off = dput(P->il_off);
        P->d_leaf_sb = dput(P->leaf_sb);
        P->d_leaf_se = dput(P->leaf_se);
        P->d_near_ptr = dput(P->near_ptr);
        P->d_near_leaf = dput(P->near_leaf);
        P->d_m2p_ptr = dput(P->m2p_ptr);
        P->d_m2p_cell = dput(P->m2p_cell);
        P->d_lcell = dput(P->lcell);
        P->d_tgt_node = dput(P->tgt_node);
        P->d_tgt_leaf = dput(P->tgt_leaf);
        P->d_tgt_pos = dput(P->tgt_pos);
        P->d_It = dput(P->It);
        P->d_Rc = dput(P->Rc);
        P->d_src_cells = dput(P->src_cells);
        P->d_tgt_cells = dput(P->tgt_cells);
        const size_t nexp = (size_t)P->ncell() * 2 * P->nc;
        void *d = nullptr;
        if (hipMalloc(&d, nexp * sizeof(double2)) != hipSuccess) throw std::string("hipMalloc M");
        P->d_M = (double2 *)d;
        if (hipMalloc(&d, nexp * sizeof(double2)) != hipSuccess) throw std::string("hipMalloc L");
        P->d_L = (double2 *)d;
        std::vector<double> W(P->W.begin(), P->W.end());
        c.fmm_W = dput(W);
        return FG_OK;
    } catch (const std::string &e) {
        c.err = e;
        return FG_ERR_ARG;
    }
}

void fmm_destroy(Ctx &c)
{
    FmmPlan *P = c.fmm;
    if (!P) return;
    void *ptrs[] = {P->d_cen, P->d_parent, P->d_oct, P->d_ch_ptr, P->d_ch_idx, P->d_il_ptr, P->d_il_src,
                    P->d_il_off, P->d_leaf_sb, P->d_leaf_se, P->d_near_ptr, P->d_near_leaf, P->d_tgt_node,
                    P->d_tgt_leaf, P->d_tgt_pos, P->d_It, P->d_Rc, P->d_M, P->d_L, P->d_src_cells, P->d_tgt_cells,
                    P->d_m2p_ptr, P->d_m2p_cell, P->d_lcell,
                    c.fmm_W};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    c.fmm_W = nullptr;
    delete P;
    c.fmm = nullptr;
}

template <int PP>
static void run_fmm(Ctx &c)
{
    FmmPlan &P = *c.fmm;
    const int L = P.L, nc = P.nc;
    const int NL = P.lev_off[L + 1] - P.lev_off[L];
    const size_t nexp = (size_t)P.ncell() * 2 * nc;
    {
        Prof pr(c, "fmm_clear");
        k_zero2<<<1024, 256, 0, c.stream>>>(nexp, P.d_M);
        k_zero2<<<1024, 256, 0, c.stream>>>(nexp, P.d_L);
    }
    {
        Prof pr(c, "fmm_p2m");
        constexpr int G = 8;
        dim3 grid((NL * G + 63) / 64, 2);
        k_fmm_p2m<PP, G><<<grid, 64, 0, c.stream>>>(NL, P.lev_off[L], 1.0 / P.W[L], P.d_cen, P.d_leaf_sb, P.d_leaf_se,
                                                    c.src_pos, c.src_q, P.d_M);
    }
    {
        Prof pr(c, "fmm_m2m");
        for (int l = L - 1; l >= 2; l--) {
            const int n = P.src_cells_lev_off[l + 1] - P.src_cells_lev_off[l];
            if (n == 0) continue;
            const int threads = n * 2 * nc;
            k_fmm_m2m<<<(threads + 255) / 256, 256, 0, c.stream>>>(n, P.d_src_cells + P.src_cells_lev_off[l], PP, nc,
                                                                   P.d_ch_ptr, P.d_ch_idx, P.d_oct, P.d_Rc, P.d_M);
        }
    }
    {
        Prof pr(c, "fmm_m2l");
        const int n = P.tgt_cells_lev_off[std::min(P.ls, L + 1)] - P.tgt_cells_lev_off[2];
        const int bs = ((2 * nc + 63) / 64) * 64;
        if (n > 0)
            k_fmm_m2l<PP><<<n, bs, 0, c.stream>>>(n, P.d_tgt_cells + P.tgt_cells_lev_off[2], P.d_cen, c.fmm_W,
                                                   P.d_il_ptr, P.d_il_src, P.d_il_off, P.d_It, P.d_M, P.d_L);
    }
    {
        Prof pr(c, "fmm_l2l");
        for (int l = 3; l < P.ls; l++) {
            const int n = P.tgt_cells_lev_off[l + 1] - P.tgt_cells_lev_off[l];
            if (n == 0) continue;
            const int threads = n * 2 * nc;
            k_fmm_l2l<<<(threads + 255) / 256, 256, 0, c.stream>>>(n, P.d_tgt_cells + P.tgt_cells_lev_off[l], PP, nc,
                                                                   P.d_parent, P.d_oct, P.d_Rc, P.d_L);
        }
    }
    {
        Prof pr(c, "fmm_eval");
        const int T = (int)P.tgt_node.size();
        k_fmm_eval<<<(T * 64 + 255) / 256, 256, 0, c.stream>>>(T, PP, nc, c.fmm_W, P.d_tgt_node, P.d_tgt_leaf,
                                                              P.d_tgt_pos, P.d_cen, P.d_L, P.d_M, P.d_lcell,
                                                              P.d_m2p_ptr, P.d_m2p_cell, P.d_near_ptr, P.d_near_leaf,
                                                              P.d_leaf_sb, P.d_leaf_se, c.src_pos, c.src_q, c.corr,
                                                              c.rec[1 - c.cur]);
    }
}

void launch_fmm(Ctx &c)
{
    switch (c.fmm->p) {
    case 4: run_fmm<4>(c); break;
    case 6: run_fmm<6>(c); break;
    case 8: run_fmm<8>(c); break;
    case 10: run_fmm<10>(c); break;
    default: break;
    }
}

}  // namespace fg

// ---- host reference of the FMM (tests) ----
extern "C" int fg_fmm_reference(const double *src_xyz, int64_t S, const double *tgt_xyz, int64_t T, const double *q_u,
                                const double *q_v, int order, int leaf, double *pot_u, double *pot_v, int *leaf_level)
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
        std::string e = fmm_build_host(P, src, tnode, tpos, order, leaf, fmm_m2p_threshold());
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
