/*
 * oracle.cpp -- FP64 CPU restatement of FeeLLGood's per-timestep LLG path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Written from the reference's mathematics; each
 * function cites the reference file:line it restates. Arithmetic follows the reference's order
 * where that is observable (element matrices through P*AE*P^T, sequential Krylov vector ops,
 * row-ordered SpMV), so parity with the product is a matter of rounding, not of algorithm.
 *
 * Parity pinning: every function here is checked by tests/test_oracle.py against known answers
 * from the reference's unit tests (inputs regenerated with std::mt19937(5489)), closed forms and
 * invariants, and end to end against ci-tests/full_test.py's golden values.
 */
#include "oracle.h"
#include "fmm_cpu.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <utility>
#include <map>
#include <vector>

namespace {

/* ---------------- constants (src/config.h.in:31-32,57-68) ---------------- */
const double MU0 = 1.25663706127e-6;
const double GAMMA0 = 1.76085962784e11 * MU0;
const double THETA = 0.5;
const double EPSILON_TET = 1e-40;

/* Gauss tables, default build NPI = 5 / 4 (src/tetra.h:190-220, src/triangle.h:42-64) */
const int TN = 4, TNPI = 5;
const double TA = 1. / 4., TB = 1. / 6., TC = 1. / 2., TD = -2. / 15., TE = 3. / 40.;
const double tu[TNPI] = {TA, TB, TB, TB, TC};
const double tv[TNPI] = {TA, TB, TB, TC, TB};
const double tw[TNPI] = {TA, TB, TC, TB, TB};
const double tpds[TNPI] = {TD, TE, TE, TE, TE};
double ta[TN][TNPI];

const int FN = 3, FNPI = 4;
const double fu[FNPI] = {1 / 3., 1 / 5., 3 / 5., 1 / 5.};
const double fv[FNPI] = {1 / 3., 1 / 5., 1 / 5., 3 / 5.};
const double fpds[FNPI] = {-27 / 96., 25 / 96., 25 / 96., 25 / 96.};
double fa[FN][FNPI];

struct TablesInit {
    TablesInit()
    {
        for (int j = 0; j < TNPI; j++) {
            ta[0][j] = 1. - tu[j] - tv[j] - tw[j];
            ta[1][j] = tu[j];
            ta[2][j] = tv[j];
            ta[3][j] = tw[j];
        }
        for (int j = 0; j < FNPI; j++) {
            fa[0][j] = 1. - fu[j] - fv[j];
            fa[1][j] = fu[j];
            fa[2][j] = fv[j];
        }
    }
} tables_init;

thread_local std::string g_err;

struct V3 {
    double x, y, z;
};
inline V3 v3(const double *p) { return {p[0], p[1], p[2]}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline double comp(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
inline V3 normalized(V3 a)
{
    double n2 = dot(a, a);
    if (n2 > 0) {
        double n = std::sqrt(n2);
        return {a.x / n, a.y / n, a.z / n};
    }
    return a;
}
inline double sq(double x) { return x * x; }

/* ---------------- node state (src/node.h:47-122) ---------------- */
struct DataNode {
    V3 u{0, 0, 0}, v{0, 0, 0};
    double phi = 0, phiv = 0;
};
struct Node {
    V3 p{0, 0, 0}, ep{0, 0, 0}, eq{0, 0, 0};
    DataNode d[2];
};

/* Node::setBasis (src/node.h:73-102) */
void set_basis(const V3 &u, double r, V3 &ep, V3 &eq)
{
    double au[3] = {std::fabs(u.x), std::fabs(u.y), std::fabs(u.z)};
    int minIdx = 0;
    for (int k = 1; k < 3; k++)
        if (au[k] < au[minIdx]) minIdx = k;
    V3 e{0, 0, 0};
    if (minIdx == 0) e.x = 1; else if (minIdx == 1) e.y = 1; else e.z = 1;
    e = e - dot(e, u) * u;
    e = normalized(e);
    V3 q = cross(u, e);
    V3 new_ep = std::cos(r) * e - std::sin(r) * q;
    eq = std::sin(r) * e + std::cos(r) * q;
    ep = new_ep;
}

/* ---------------- tetrahedron (src/tetra.h:280-303, src/tetra.cpp:393-424) ---------------- */
struct TetGeo {
    int ind[4];
    int reg;
    double da[4][3];
    double w[TNPI];
};

/* returns -1 when singular */
int tet_geometry(const V3 P[4], int ind[4], double da[4][3], double w[TNPI], V3 *Pout)
{
    V3 p[4] = {P[0], P[1], P[2], P[3]};
    V3 p0p1 = p[1] - p[0], p0p2 = p[2] - p[0], p0p3 = p[3] - p[0];
    double mixed = dot(p0p1, cross(p0p2, p0p3));
    if (std::fabs(mixed) < EPSILON_TET) return -1;
    if (mixed < 0.0) {
        std::swap(ind[2], ind[3]);
        std::swap(p[2], p[3]);
    }
    p0p1 = p[1] - p[0];
    p0p2 = p[2] - p[0];
    p0p3 = p[3] - p[0];
    double J[3][3] = {{p0p1.x, p0p2.x, p0p3.x}, {p0p1.y, p0p2.y, p0p3.y}, {p0p1.z, p0p2.z, p0p3.z}};
    /* Matrix3d::determinant: expansion along the first row */
    double detJ = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                  J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                  J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    /* Matrix3d::inverse: adjugate times 1/det, det from the first-column cofactors */
    auto cof = [&J](int i, int j) {
        int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return J[i1][j1] * J[i2][j2] - J[i1][j2] * J[i2][j1];
    };
    double invdet = 1.0 / (cof(0, 0) * J[0][0] + cof(1, 0) * J[1][0] + cof(2, 0) * J[2][0]);
    double Ji[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Ji[j][i] = cof(i, j) * invdet;
    const double dadu[4][3] = {{-1., -1., -1.}, {1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.}};
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 3; k++) {
            double s = 0;
            for (int l = 0; l < 3; l++) s += dadu[i][l] * Ji[l][k];
            da[i][k] = s;
        }
    for (int j = 0; j < TNPI; j++) w[j] = detJ * tpds[j];
    if (Pout)
        for (int i = 0; i < 4; i++) Pout[i] = p[i];
    return 0;
}

/* Tetra::calc_alpha_eff (src/tetra.cpp:47-75) */
double alpha_eff(double dt, double alpha, double h)
{
    double reduced_dt = GAMMA0 * dt;
    const double r = 0.1;
    const double M = 2. * alpha * r / reduced_dt;
    if (h > 0.) {
        if (h > M) return alpha + reduced_dt / 2. * M;
        return alpha + reduced_dt / 2. * h;
    }
    if (h < -M) return alpha / (1. + reduced_dt / (2. * alpha) * M);
    return alpha / (1. - reduced_dt / (2. * alpha) * h);
}

/* Tet::lumping (src/tetra.cpp:108-131) with calcDiagBlock / calcOffDiagBlock (:133-148) */
void tet_lumping(const double da[4][3], const double w[TNPI], const V3 u_nod[4],
                 const double aeff[TNPI], double prefactor, double AE[12][12])
{
    double wsum = 0;
    for (int j = 0; j < TNPI; j++) wsum += w[j];
    double contrib[4], a_w[4];
    for (int i = 0; i < 4; i++) {
        double c = 0, m = 0;
        for (int j = 0; j < TNPI; j++) {
            c += ta[i][j] * (w[j] * aeff[j]);
            m += ta[i][j] * w[j];
        }
        contrib[i] = c;
        a_w[i] = m;
    }
    double blk[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double s = da[i][0] * da[j][0] + da[i][1] * da[j][1] + da[i][2] * da[j][2];
            blk[i][j] = s * (prefactor * wsum);
        }
    for (int i = 0; i < 4; i++) blk[i][i] += contrib[i];
    for (int b = 0; b < 3; b++)
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) AE[4 * b + i][4 * b + j] += blk[i][j];
    for (int i = 0; i < 4; i++) {
        double X = a_w[i] * u_nod[i].x, Y = a_w[i] * u_nod[i].y, Z = a_w[i] * u_nod[i].z;
        AE[4 + i][8 + i] -= X;
        AE[8 + i][4 + i] += X;
        AE[0 + i][8 + i] += Y;
        AE[8 + i][0 + i] -= Y;
        AE[0 + i][4 + i] -= Z;
        AE[4 + i][0 + i] += Z;
    }
}

/* Tet::integrales (src/tetra.cpp:210-307) with the applied field at each Gauss point (calc_Hext:
 * a uniform field in every column, src/linear_algebra.cpp:30-35, or A(t) times the space field,
 * :63-68), no spin transfer, no drift. */
/* Tet::calc_aniso_uniax (src/tetra.cpp:171-181): H_an,g += Kbis (uk . (U_g + s_dt V_g)) uk and
   the uHeff contribution Kbis (uk . U_g)^2, added to uHeff[g]. */
void aniso_uniax(const V3 U[TNPI], const V3 V[TNPI], V3 uk, double Kbis, double s_dt, V3 Han[TNPI],
                 double uHeff[TNPI])
{
    for (int g = 0; g < TNPI; g++) {
        Han[g] = Han[g] + (Kbis * dot(uk, U[g] + s_dt * V[g])) * uk;
        uHeff[g] += Kbis * sq(dot(U[g], uk));
    }
}

/* Tet::calc_aniso_cub (src/tetra.cpp:183-208): with beta = (ex.U, ey.U, ez.U) and
   beta' = (ex.V, ey.V, ez.V), H_an,g += -K3bis [sum_a beta_a (1 - beta_a^2) e_a
   + s_dt (ex o beta') o (1 - 3 beta^2)] (o: component-wise; the second term takes the components
   of ex only, as the reference writes it) and uHeff[g] += -K3bis sum_a beta_a^2 (1 - beta_a^2). */
void aniso_cub(const V3 U[TNPI], const V3 V[TNPI], V3 ex, V3 ey, V3 ez, double K3bis, double s_dt,
               V3 Han[TNPI], double uHeff[TNPI])
{
    for (int g = 0; g < TNPI; g++) {
        V3 uk_u{dot(ex, U[g]), dot(ey, U[g]), dot(ez, U[g])};
        V3 uk_v{dot(ex, V[g]), dot(ey, V[g]), dot(ez, V[g])};
        V3 uuu{uk_u.x * (1.0 - uk_u.x * uk_u.x), uk_u.y * (1.0 - uk_u.y * uk_u.y),
               uk_u.z * (1.0 - uk_u.z * uk_u.z)};
        V3 tmp{uk_v.x * ex.x, uk_v.y * ex.y, uk_v.z * ex.z};
        V3 one3{1 - 3 * uk_u.x * uk_u.x, 1 - 3 * uk_u.y * uk_u.y, 1 - 3 * uk_u.z * uk_u.z};
        V3 t2{tmp.x * one3.x, tmp.y * one3.y, tmp.z * one3.z};
        V3 h = uuu.x * ex + uuu.y * ey + uuu.z * ez + s_dt * t2;
        Han[g] = Han[g] + (-K3bis) * h;
        uHeff[g] += -K3bis * dot(uk_u, uuu);
    }
}

void tet_integrales(const double da[4][3], const double w[TNPI], const V3 u[4], const V3 v[4],
                    const double phi[4], const double phiv[4], const V3 ep[4], const V3 eq[4],
                    const fg_volume_region &prm, const V3 Hext[TNPI], double dt, double prefactor,
                    double Kp[8][8], double Lp[8], const V3 *Hst = nullptr)
{
    const double alpha = prm.alpha;
    const double Ms = prm.Ms;
    const double Abis = 2.0 * prm.Ae / (MU0 * Ms);
    const double s_dt = THETA * dt * GAMMA0;

    /* interpolation (src/tetra.h:323-358) */
    V3 U[TNPI], V[TNPI], dU[3];
    for (int g = 0; g < TNPI; g++) {
        V3 su{0, 0, 0}, sv{0, 0, 0};
        for (int i = 0; i < 4; i++) {
            su = su + ta[i][g] * u[i];
            sv = sv + ta[i][g] * v[i];
        }
        U[g] = su;
        V[g] = sv;
    }
    for (int k = 0; k < 3; k++) {
        V3 s{0, 0, 0};
        for (int i = 0; i < 4; i++) s = s + da[i][k] * u[i];
        dU[k] = s;
    }
    V3 Hd{0, 0, 0}, Hv{0, 0, 0};
    for (int i = 0; i < 4; i++) {
        V3 dai{da[i][0], da[i][1], da[i][2]};
        Hd = Hd - phi[i] * dai;
        Hv = Hv - phiv[i] * dai;
    }
    double uHeff[TNPI];
    double exch = -Abis * (dot(dU[0], dU[0]) + dot(dU[1], dU[1]) + dot(dU[2], dU[2]));
    V3 Han[TNPI];
    for (int g = 0; g < TNPI; g++) {
        uHeff[g] = exch;
        Han[g] = {0, 0, 0};
    }
    if (prm.K != 0) /* calc_aniso_uniax called with s_dt/gamma0 (src/tetra.cpp:236-241) */
        aniso_uniax(U, V, v3(prm.uk), 2.0 * prm.K / (MU0 * Ms), s_dt / GAMMA0, Han, uHeff);
    if (prm.K3 != 0) /* calc_aniso_cub (src/tetra.cpp:242-246) */
        aniso_cub(U, V, v3(prm.ex), v3(prm.ey), v3(prm.ez), 2.0 * prm.K3 / (MU0 * Ms), s_dt / GAMMA0, Han, uHeff);
    /* Heff = Hd + Hext + Hst (src/tetra.cpp:248-255) */
    for (int g = 0; g < TNPI; g++) uHeff[g] += dot(U[g], Hst ? (Hd + Hext[g]) + Hst[g] : Hd + Hext[g]);
    double aeff[TNPI];
    for (int g = 0; g < TNPI; g++) aeff[g] = alpha_eff(dt, alpha, uHeff[g]);

    double AE[12][12];
    std::memset(AE, 0, sizeof(AE));
    tet_lumping(da, w, u, aeff, prefactor * s_dt * Abis, AE);

    /* P (src/element.h:81-96) and Perm (src/tetra.cpp:270-271) */
    double P[8][12];
    std::memset(P, 0, sizeof(P));
    for (int i = 0; i < 4; i++) {
        P[i][i] = ep[i].x;
        P[i][4 + i] = ep[i].y;
        P[i][8 + i] = ep[i].z;
        P[4 + i][i] = eq[i].x;
        P[4 + i][4 + i] = eq[i].y;
        P[4 + i][8 + i] = eq[i].z;
    }
    double PAE[8][12];
    for (int r = 0; r < 8; r++)
        for (int c = 0; c < 12; c++) {
            double s = 0;
            for (int k = 0; k < 12; k++) s += P[r][k] * AE[k][c];
            PAE[r][c] = s;
        }
    for (int r = 0; r < 8; r++) {
        int rp = (r + 4) % 8; /* Perm swaps the ep and eq halves */
        for (int c = 0; c < 8; c++) {
            double s = 0;
            for (int k = 0; k < 12; k++) s += PAE[r][k] * P[c][k];
            Kp[rp][c] = s;
        }
    }

    /* BE (src/tetra.cpp:277-303) */
    V3 BE[4] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const double sdt_g = s_dt / GAMMA0;
    for (int g = 0; g < TNPI; g++) {
        V3 H = (Hd + Hext[g]) + (Han[g] + sdt_g * Hv);
        const double wg = w[g];
        for (int i = 0; i < 4; i++) {
            const double ai_w = wg * ta[i][g];
            V3 ex = da[i][0] * dU[0] + da[i][1] * dU[1] + da[i][2] * dU[2];
            BE[i] = BE[i] - (wg * Abis) * ex;
            if (Hst) { /* the Hst terms of src/tetra.cpp:296-302 */
                BE[i] = BE[i] + ai_w * (H + Hst[g]);
                BE[i] = BE[i] - ((ai_w * dot(Hst[g], U[g])) * s_dt) * V[g];
            } else {
                BE[i] = BE[i] + ai_w * H;
            }
        }
    }
    /* Lp = Perm * P * vec(BE) (src/tetra.cpp:306) */
    for (int i = 0; i < 4; i++) {
        Lp[4 + i] = ep[i].x * BE[i].x + ep[i].y * BE[i].y + ep[i].z * BE[i].z;
        Lp[i] = eq[i].x * BE[i].x + eq[i].y * BE[i].y + eq[i].z * BE[i].z;
    }
}

/* ---------------- triangle (src/triangle.h:113-127, src/triangle.cpp) ---------------- */
struct TriGeo {
    int ind[3];
    int reg;
    double surf;
    V3 n;
    double w[FNPI];
    double dMs;
};

/* Tri::potential (src/triangle.cpp:87-125) */
double tri_potential(const V3 P[3], const double s[3], double surf, double dMs, int i)
{
    int ii = (i + 1) % 3, iii = (i + 2) % 3;
    V3 p1p2 = P[ii] - P[i];
    V3 p1p3 = P[iii] - P[i];
    auto f = [](double x) { return std::sqrt(1.0 + x * x); };
    double b = norm(p1p2);
    double t = dot(p1p2, p1p3) / b;
    double _2s = 2. * surf;
    double h = _2s / b;
    double c = (t - b) / h;
    double r = h * f(t / h);
    double log_1 = std::log((c * t + h + f(c) * r) / (b * (c + f(c))));
    double xi = b * log_1 / f(c);
    double s1 = s[i], s2 = s[ii], s3 = s[iii];
    double pot = xi * s1 + ((xi * (h + c * t) - b * (r - b)) * s2 + b * (r - b - c * xi) * s3) * b /
                              (_2s * (1 + c * c));
    return 0.5 * dMs * pot;
}

/* Tri geometry (src/triangle.h:113-127): area, unit normal of (P1 - P0) x (P2 - P0), weights
   w_g = 2 area pds_g. */
void tri_geometry(const V3 P[3], TriGeo &f)
{
    V3 nv = cross(P[1] - P[0], P[2] - P[0]);
    f.surf = 0.5 * norm(nv);
    f.n = normalized(nv);
    for (int j = 0; j < FNPI; j++) f.w[j] = 2.0 * f.surf * fpds[j];
}

/* Tri::integrales (src/triangle.cpp:6-36): surface (Neel) anisotropy, Kbis = 2 Ks / dMs,
   BE_i = sum_g w_g Kbis (uk . U_g) a_ig uk, Lp = [eq_i . BE_i ; ep_i . BE_i] (Perm after P). */
void tri_integrales(const TriGeo &f, const V3 u[3], const V3 ep[3], const V3 eq[3], double Ks, V3 uk,
                    double Lp[6])
{
    double Kbis = 2.0 * Ks / f.dMs;
    V3 BE[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int g = 0; g < FNPI; g++) {
        V3 U{0, 0, 0};
        for (int i = 0; i < 3; i++) U = U + fa[i][g] * u[i];
        double pre = f.w[g] * Kbis * dot(uk, U);
        for (int i = 0; i < 3; i++) BE[i] = BE[i] + (pre * fa[i][g]) * uk;
    }
    for (int i = 0; i < 3; i++) {
        Lp[3 + i] = dot(ep[i], BE[i]);
        Lp[i] = dot(eq[i], BE[i]);
    }
}

/* Tri::charges (src/triangle.cpp:45-59): q_g = dMs w_g (U_g . n), zero when dMs = 0. */
void tri_charges(const TriGeo &f, const V3 u[3], double q[FNPI])
{
    for (int g = 0; g < FNPI; g++) {
        V3 U{0, 0, 0};
        for (int i = 0; i < 3; i++) U = U + fa[i][g] * u[i];
        q[g] = (f.dMs != 0.0) ? f.dMs * (f.w[g] * dot(U, f.n)) : 0.0;
    }
}

/* Tri::anisotropyEnergy and Tri::demagEnergy (src/triangle.cpp:38-43,80-85): e[0] = -Ks sum_g
   w_g (U_g . uk)^2, e[1] = 1/2 mu0 dMs sum_g (U_g . n) phi_g w_g. */
void tri_energy_terms(const TriGeo &f, const V3 u[3], const double phi[3], double Ks, V3 uk, double e[2])
{
    V3 U[FNPI];
    double ph[FNPI];
    for (int g = 0; g < FNPI; g++) {
        V3 s{0, 0, 0};
        double sp = 0;
        for (int i = 0; i < 3; i++) {
            s = s + fa[i][g] * u[i];
            sp += phi[i] * fa[i][g];
        }
        U[g] = s;
        ph[g] = sp;
    }
    double a = 0;
    for (int g = 0; g < FNPI; g++) a += f.w[g] * sq(dot(U[g], uk));
    e[0] = -Ks * a;
    double d = 0;
    for (int g = 0; g < FNPI; g++) d += (dot(U[g], f.n) * ph[g]) * f.w[g];
    e[1] = 0.5 * MU0 * f.dMs * d;
}

/* Tet::charges (src/tetra.cpp:347-359): q_g = -Ms w_g div(u), div u = sum_i da_i . u_i. */
void tet_charges(const double da[4][3], const double w[TNPI], const V3 vn[4], double Ms, double q[TNPI])
{
    double dud_sum = 0;
    for (int i = 0; i < 4; i++) dud_sum += vn[i].x * da[i][0];
    double sy = 0, sz = 0;
    for (int i = 0; i < 4; i++) sy += vn[i].y * da[i][1];
    for (int i = 0; i < 4; i++) sz += vn[i].z * da[i][2];
    dud_sum = dud_sum + sy + sz;
    for (int g = 0; g < TNPI; g++) q[g] = -Ms * w[g] * dud_sum;
}

/* The energies of one tetrahedron (src/tetra.cpp:309-391): e[0] exchange A sum_g w_g |grad U|^2,
   e[1] uniaxial -K sum_g w_g (uk . U_g)^2, e[2] cubic K3 sum_g w_g (b0^2 b1^2 + b1^2 b2^2 +
   b2^2 b0^2), e[3] demag -1/2 mu0 Ms sum_g w_g div(U) phi_g, e[4] Zeeman -mu0 Ms sum_g w_g U_g . H_g. */
void tet_energy_terms(const double da[4][3], const double w[TNPI], const V3 u[4], const double phi[4],
                      const fg_volume_region &prm, const V3 Hg[TNPI], double e[5])
{
    V3 U[TNPI], dU[3];
    double ph[TNPI];
    for (int g = 0; g < TNPI; g++) {
        V3 s{0, 0, 0};
        double sp = 0;
        for (int i = 0; i < 4; i++) {
            s = s + ta[i][g] * u[i];
            sp += phi[i] * ta[i][g];
        }
        U[g] = s;
        ph[g] = sp;
    }
    for (int d = 0; d < 3; d++) {
        V3 s{0, 0, 0};
        for (int i = 0; i < 4; i++) s = s + da[i][d] * u[i];
        dU[d] = s;
    }
    double dens = dot(dU[0], dU[0]) + dot(dU[1], dU[1]) + dot(dU[2], dU[2]);
    double ex = 0, dm = 0, an = 0, ze = 0, c = 0;
    for (int g = 0; g < TNPI; g++) ex += w[g] * dens;
    e[0] = prm.Ae * ex;
    double divu = dU[0].x + dU[1].y + dU[2].z;
    for (int g = 0; g < TNPI; g++) dm += w[g] * (divu * ph[g]);
    e[3] = -0.5 * MU0 * prm.Ms * dm;
    V3 uk = v3(prm.uk);
    for (int g = 0; g < TNPI; g++) an += w[g] * sq(dot(uk, U[g]));
    e[1] = -prm.K * an;
    for (int g = 0; g < TNPI; g++) {
        double a0 = dot(U[g], v3(prm.ex)), a1 = dot(U[g], v3(prm.ey)), a2 = dot(U[g], v3(prm.ez));
        c += w[g] * (sq(a0 * a1) + sq(a1 * a2) + sq(a2 * a0));
    }
    e[2] = prm.K3 * c;
    for (int g = 0; g < TNPI; g++) ze += w[g] * dot(U[g], Hg[g]);
    e[4] = -MU0 * prm.Ms * ze;
}

} // namespace

/* ====================================================================================== */
struct orc_problem {
    int N = 0;
    std::vector<Node> node;
    std::vector<TetGeo> tet;
    std::vector<TriGeo> tri;
    std::vector<fg_volume_region> vol;
    std::vector<fg_surface_region> surf;
    std::vector<std::pair<int, int>> edges;
    std::vector<char> magNode;
    std::vector<int> magTet, magTri;
    std::vector<double> hsp; /* space-dependent applied field [n_tets][5][3] (orc_set_space_field) */
    std::vector<double> hst; /* spin-accumulation field [n_tets][5][3] (orc_set_spin_accumulation) */
    std::vector<int> lvd;
    /* element matrices (src/element.h:61-65) */
    std::vector<double> Kp;  /* 64 per tet */
    std::vector<double> Lp;  /* 8 per tet */
    std::vector<double> LpT; /* 6 per tri */
    /* sparse matrix rows (src/algebra/sparseMat.h:59-75) */
    std::vector<int64_t> rowptr;
    std::vector<int32_t> col;
    std::vector<double> val;
    std::vector<double> L;
    std::vector<double> X;
    double v_max = 0;
    double tol = 1e-6;
    int max_iter = 700;
    /* demag bookkeeping */
    std::vector<double> src_xyz;
    std::vector<double> src_qu, src_qv;
    std::vector<double> corr_u, corr_v;
    double totalMagVol = 0;
    int fmm_order = 0, fmm_leaf = 64; /* orc_set_demag: 0 = exact direct sums */

    bool isMagTet(const TetGeo &t) const { return vol[t.reg].Ms > 0; }
    double &K(int i, int j)
    {
        auto b = col.begin() + rowptr[i], e = col.begin() + rowptr[i + 1];
        auto it = std::lower_bound(b, e, j);
        return val[it - col.begin()];
    }
};

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

orc_problem *orc_create(const fg_mesh *m, double tol, int max_iter)
{
    if (!m || m->n_nodes <= 0 || m->n_tets <= 0) {
        g_err = "empty mesh";
        return nullptr;
    }
    auto *p = new orc_problem();
    p->N = m->n_nodes;
    p->tol = tol;
    p->max_iter = max_iter;
    p->vol.assign(m->vol, m->vol + m->n_vol_regions);
    if (m->n_surf_regions > 0) p->surf.assign(m->surf, m->surf + m->n_surf_regions);
    p->node.resize(p->N);
    for (int i = 0; i < p->N; i++) p->node[i].p = v3(m->pos + 3 * i);
    p->tet.resize(m->n_tets);
    for (int k = 0; k < m->n_tets; k++) {
        TetGeo &t = p->tet[k];
        for (int i = 0; i < 4; i++) t.ind[i] = m->tet_nodes[4 * k + i];
        t.reg = m->tet_region[k];
        V3 P[4];
        for (int i = 0; i < 4; i++) P[i] = p->node[t.ind[i]].p;
        if (tet_geometry(P, t.ind, t.da, t.w, nullptr) != 0) {
            g_err = "singular tetrahedron " + std::to_string(k);
            delete p;
            return nullptr;
        }
    }
    p->tri.resize(m->n_tris);
    for (int k = 0; k < m->n_tris; k++) {
        TriGeo &f = p->tri[k];
        for (int i = 0; i < 3; i++) f.ind[i] = m->tri_nodes[3 * k + i];
        f.reg = m->tri_region[k];
        f.dMs = m->tri_dMs[k];
        const V3 P[3] = {p->node[f.ind[0]].p, p->node[f.ind[1]].p, p->node[f.ind[2]].p};
        tri_geometry(P, f);
    }
    /* edges (src/mesh.h:100-114) */
    p->edges.reserve(6 * p->tet.size());
    for (const TetGeo &t : p->tet)
        for (int i = 0; i < 3; i++)
            for (int j = i + 1; j < 4; j++) p->edges.push_back(std::minmax(t.ind[i], t.ind[j]));
    std::sort(p->edges.begin(), p->edges.end());
    p->edges.erase(std::unique(p->edges.begin(), p->edges.end()), p->edges.end());
    /* magnetic sets (src/mesh.h:115-131) */
    p->magNode.assign(p->N, 0);
    for (size_t k = 0; k < p->tet.size(); k++)
        if (p->isMagTet(p->tet[k])) {
            p->magTet.push_back((int)k);
            for (int i = 0; i < 4; i++) p->magNode[p->tet[k].ind[i]] = 1;
            double vol = 0;
            for (int j = 0; j < TNPI; j++) vol += p->tet[k].w[j];
            p->totalMagVol += vol;
        }
    for (size_t k = 0; k < p->tri.size(); k++) {
        const TriGeo &f = p->tri[k];
        bool mag = p->magNode[f.ind[0]] && p->magNode[f.ind[1]] && p->magNode[f.ind[2]];
        bool supp = (!p->surf.empty()) && p->surf[f.reg].suppress_charges;
        if (mag && !supp) p->magTri.push_back((int)k);
    }
    /* lvd (src/linear_algebra.h:55-63) */
    for (int i = 0; i < p->N; i++)
        if (!p->magNode[i]) {
            p->lvd.push_back(2 * i);
            p->lvd.push_back(2 * i + 1);
        }
    /* matrix shape (src/solver.h:285-314) */
    std::vector<std::vector<int>> shape(2 * p->N);
    auto add_block = [&](int i, int j) {
        for (int k = 0; k < 2; k++)
            for (int l = 0; l < 2; l++) shape[2 * i + k].push_back(2 * j + l);
    };
    for (int i = 0; i < p->N; i++) add_block(i, i);
    for (auto &e : p->edges)
        if (p->magNode[e.first] && p->magNode[e.second]) {
            add_block(e.first, e.second);
            add_block(e.second, e.first);
        }
    p->rowptr.assign(2 * p->N + 1, 0);
    for (int r = 0; r < 2 * p->N; r++) {
        auto &s = shape[r];
        std::sort(s.begin(), s.end());
        s.erase(std::unique(s.begin(), s.end()), s.end());
        p->rowptr[r + 1] = p->rowptr[r] + (int64_t)s.size();
    }
    p->col.resize(p->rowptr.back());
    for (int r = 0; r < 2 * p->N; r++) std::copy(shape[r].begin(), shape[r].end(), p->col.begin() + p->rowptr[r]);
    p->val.assign(p->col.size(), 0.0);
    p->L.assign(2 * p->N, 0.0);
    p->X.assign(2 * p->N, 0.0);
    p->Kp.assign(64 * p->tet.size(), 0.0);
    p->Lp.assign(8 * p->tet.size(), 0.0);
    p->LpT.assign(6 * p->tri.size(), 0.0);
    p->corr_u.assign(p->N, 0.0);
    p->corr_v.assign(p->N, 0.0);
    return p;
}

void orc_destroy(orc_problem *p) { delete p; }
int orc_n_edges(const orc_problem *p) { return (int)p->edges.size(); }
int orc_n_mag_tris(const orc_problem *p) { return (int)p->magTri.size(); }

void orc_set_state(orc_problem *p, int which, const double *u, const double *v, const double *phi,
                   const double *phiv)
{
    for (int s = 0; s < 2; s++) {
        if (which != FG_BOTH && which != s) continue;
        for (int i = 0; i < p->N; i++) {
            DataNode &d = p->node[i].d[s];
            d.u = u ? v3(u + 3 * i) : V3{0, 0, 0};
            d.v = v ? v3(v + 3 * i) : V3{0, 0, 0};
            d.phi = phi ? phi[i] : 0.0;
            d.phiv = phiv ? phiv[i] : 0.0;
        }
    }
}

void orc_get_state(const orc_problem *p, int which, double *u, double *v, double *phi, double *phiv)
{
    int s = which == FG_NEXT ? 1 : 0;
    for (int i = 0; i < p->N; i++) {
        const DataNode &d = p->node[i].d[s];
        if (u) { u[3 * i] = d.u.x; u[3 * i + 1] = d.u.y; u[3 * i + 2] = d.u.z; }
        if (v) { v[3 * i] = d.v.x; v[3 * i + 1] = d.v.y; v[3 * i + 2] = d.v.z; }
        if (phi) phi[i] = d.phi;
        if (phiv) phiv[i] = d.phiv;
    }
}

/* LinAlgebra::base_projection (src/linear_algebra.cpp:3-11) with r given; mesh::setBasis
 * (src/mesh.h:178-182) */
void orc_base_projection(orc_problem *p, double r)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < p->N; i++) set_basis(p->node[i].d[0].u, r, p->node[i].ep, p->node[i].eq);
}

void orc_get_basis(const orc_problem *p, double *ep, double *eq)
{
    for (int i = 0; i < p->N; i++) {
        const Node &n = p->node[i];
        ep[3 * i] = n.ep.x; ep[3 * i + 1] = n.ep.y; ep[3 * i + 2] = n.ep.z;
        eq[3 * i] = n.eq.x; eq[3 * i + 1] = n.eq.y; eq[3 * i + 2] = n.eq.z;
    }
}

/* LinAlgebra::prepareElements (src/linear_algebra.cpp:26-52; with a space field, the overload of
 * :54-75: amp times the field at each Gauss point) */
static void prepare_elements(orc_problem *p, const double Hext_[3], const double *hsp, double amp, double dt,
                             double prefactor)
{
    const int T = (int)p->tet.size();
#pragma omp parallel for schedule(static)
    for (int k = 0; k < T; k++) {
        const TetGeo &t = p->tet[k];
        if (!p->isMagTet(t)) continue;
        V3 u[4], v[4], ep[4], eq[4];
        double phi[4], phiv[4];
        for (int i = 0; i < 4; i++) {
            const Node &n = p->node[t.ind[i]];
            u[i] = n.d[0].u;
            v[i] = n.d[0].v;
            phi[i] = n.d[0].phi;
            phiv[i] = n.d[0].phiv;
            ep[i] = n.ep;
            eq[i] = n.eq;
        }
        V3 H[TNPI];
        for (int g = 0; g < TNPI; g++)
            H[g] = hsp ? amp * V3{hsp[15 * (size_t)k + 3 * g], hsp[15 * (size_t)k + 3 * g + 1],
                                  hsp[15 * (size_t)k + 3 * g + 2]}
                       : v3(Hext_);
        double Kp[8][8], Lp[8];
        V3 Hs[TNPI];
        const bool st = !p->hst.empty();
        if (st)
            for (int g = 0; g < TNPI; g++)
                Hs[g] = {p->hst[15 * (size_t)k + 3 * g], p->hst[15 * (size_t)k + 3 * g + 1], p->hst[15 * (size_t)k + 3 * g + 2]};
        tet_integrales(t.da, t.w, u, v, phi, phiv, ep, eq, p->vol[t.reg], H, dt, prefactor, Kp, Lp, st ? Hs : nullptr);
        std::memcpy(&p->Kp[64 * k], Kp, sizeof(Kp));
        std::memcpy(&p->Lp[8 * k], Lp, sizeof(Lp));
    }
    /* Tri::integrales (src/triangle.cpp:6-36) for magnetic triangles with Ks != 0 */
    for (size_t k = 0; k < p->tri.size(); k++) {
        const TriGeo &f = p->tri[k];
        bool mag = p->magNode[f.ind[0]] && p->magNode[f.ind[1]] && p->magNode[f.ind[2]];
        if (!mag || p->surf.empty() || p->surf[f.reg].Ks == 0) continue;
        const fg_surface_region &sp = p->surf[f.reg];
        V3 u[3], ep[3], eq[3];
        for (int i = 0; i < 3; i++) {
            const Node &n = p->node[f.ind[i]];
            u[i] = n.d[0].u;
            ep[i] = n.ep;
            eq[i] = n.eq;
        }
        tri_integrales(f, u, ep, eq, sp.Ks, v3(sp.uk), &p->LpT[6 * k]);
    }
}

void orc_prepare_elements(orc_problem *p, const double Hext[3], double dt, double prefactor)
{
    prepare_elements(p, Hext, nullptr, 0.0, dt, prefactor);
}

void orc_prepare_elements_amp(orc_problem *p, double amp, double dt, double prefactor)
{
    const double zero[3] = {0, 0, 0};
    prepare_elements(p, zero, p->hsp.empty() ? nullptr : p->hsp.data(), amp, dt, prefactor);
}

/* Tetra::calc_Hst with spinAcc::prepareExtraField's prefactor D0 / (lsd^2 gamma0 Ms)
   (src/tetra.cpp:93-106, src/spinAccumulationSolver.cpp:138-149) from the node spin accumulation
   s[3 n_nodes]; NULL removes it */
void orc_set_spin_accumulation(orc_problem *p, const double *sacc, const fg_spin_volume *vp)
{
    if (!sacc) {
        p->hst.clear();
        return;
    }
    const double E = 1.602176634e-19;
    p->hst.assign(15 * p->tet.size(), 0.0);
    for (size_t k = 0; k < p->tet.size(); k++) {
        const TetGeo &t = p->tet[k];
        if (!p->isMagTet(t)) continue;
        const fg_spin_volume &pr = vp[t.reg];
        const double D0 = 2.0 * pr.sigma / ((E * E) * pr.N0);
        const double pref = D0 / ((pr.lsd * pr.lsd) * GAMMA0 * p->vol[t.reg].Ms);
        for (int g = 0; g < TNPI; g++)
            for (int c = 0; c < 3; c++) {
                double v = 0.0;
                for (int i = 0; i < 4; i++) v += sacc[3 * (size_t)t.ind[i] + c] * ta[i][g];
                p->hst[15 * k + 3 * g + c] = pref * v;
            }
    }
}

void orc_set_space_field(orc_problem *p, const double *h)
{
    if (h)
        p->hsp.assign(h, h + 15 * p->tet.size());
    else
        p->hsp.clear();
}

} // extern "C"

namespace {
/* algebra:: vector helpers (src/algebra/algebra.h:34-65, src/algebra/algebraCore.h:8-17),
 * sequential like the reference */
double vdot(const std::vector<double> &X, const std::vector<double> &Y)
{
    double s = 0;
    for (size_t i = 0; i < X.size(); i++) s += X[i] * Y[i];
    return s;
}
double vnorm(const std::vector<double> &X) { return std::sqrt(std::fabs(vdot(X, X))); }
void mask(const std::vector<int> &ld, std::vector<double> &X)
{
    for (int i : ld) X[i] = 0;
}
struct CSR {
    int64_t n;
    const int64_t *rowptr;
    const int32_t *col;
    const double *val;
};
void spmv(const CSR &A, const std::vector<double> &X, std::vector<double> &Y)
{
    const int64_t R = A.n;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < R; r++) {
        double s = 0;
        for (int64_t k = A.rowptr[r]; k < A.rowptr[r + 1]; k++) s += A.val[k] * X[A.col[k]];
        Y[r] = s;
    }
}

/* iteration<double> (src/algebra/iter.h:36-179) */
struct Iter {
    double rhsn = 1.0, res = std::numeric_limits<double>::max(), resmax;
    int maxiter, nit = 0, status = FG_ITER_UNDEFINED;
    Iter(double r, int m) : resmax(r), maxiter(m) {}
    bool converged(double nr)
    {
        res = std::fabs(nr);
        if (std::isnan(res)) {
            status = FG_ITER_CANNOT_CONVERGE;
            return false;
        }
        bool cv = res <= rhsn * resmax;
        if (cv) status = FG_ITER_CONVERGED;
        return cv;
    }
    void inc()
    {
        nit++;
        if (nit >= maxiter) status = FG_ITER_OVERFLOW;
    }
};

/* algebra::bicg_dir, 5-argument overload (src/algebra/bicg.h:162-234) */
void bicg_dir(Iter &iter, const CSR &A, std::vector<double> &x, const std::vector<double> &rhs,
              const std::vector<int> &ld, const std::vector<double> *xd = nullptr)
{
    double rho_1 = 0, rho_2 = 0, alpha = 0, beta = 0, omega = 0;
    const size_t DIM = x.size();
    std::vector<double> pv(DIM), phat(DIM), shat(DIM), r(DIM), rt(DIM), s(DIM), t(DIM), v(DIM),
        dp(DIM), b(rhs);
    /* build_diag_precond (src/algebra/sparseMat.h:173-183) */
    for (size_t i = 0; i < DIM; i++) {
        double c = 0;
        for (int64_t k = A.rowptr[i]; k < A.rowptr[i + 1]; k++)
            if (A.col[k] == (int)i) c = A.val[k];
        dp[i] = 1.0 / c;
    }
    if (xd) { /* the 6-argument overload (src/algebra/bicg.h:82-158): b = rhs - A xd */
        spmv(A, *xd, v);
        for (size_t i = 0; i < DIM; i++) b[i] -= v[i];
    }
    mask(ld, b);
    mask(ld, dp);
    iter.rhsn = vnorm(b);
    r = b;
    spmv(A, x, v);
    for (size_t i = 0; i < DIM; i++) r[i] -= v[i];
    mask(ld, r);
    rt = r;
    pv = r;
    while (!iter.converged(vnorm(r))) {
        rho_1 = vdot(rt, r);
        if (iter.nit > 0) {
            if (rho_2 == 0 || omega == 0) {
                iter.status = FG_ITER_CANNOT_CONVERGE;
                break;
            }
            beta = (rho_1 / rho_2) * (alpha / omega);
            for (size_t i = 0; i < DIM; i++) v[i] *= omega;
            for (size_t i = 0; i < DIM; i++) pv[i] -= v[i];
            for (size_t i = 0; i < DIM; i++) pv[i] *= beta;
            for (size_t i = 0; i < DIM; i++) pv[i] += r[i];
        }
        for (size_t i = 0; i < DIM; i++) phat[i] = dp[i] * pv[i];
        spmv(A, phat, v);
        mask(ld, v);
        alpha = rho_1 / vdot(v, rt);
        s = r;
        for (size_t i = 0; i < DIM; i++) s[i] += (-alpha) * v[i];
        if (iter.converged(vnorm(s))) {
            for (size_t i = 0; i < DIM; i++) x[i] += alpha * phat[i];
            break;
        } else if (iter.status == FG_ITER_OVERFLOW || iter.status == FG_ITER_CANNOT_CONVERGE) {
            break;
        }
        for (size_t i = 0; i < DIM; i++) shat[i] = dp[i] * s[i];
        spmv(A, shat, t);
        mask(ld, t);
        omega = vdot(t, s) / vdot(t, t);
        for (size_t i = 0; i < DIM; i++) x[i] += alpha * phat[i];
        for (size_t i = 0; i < DIM; i++) x[i] += omega * shat[i];
        for (size_t i = 0; i < DIM; i++) t[i] *= omega;
        r = s;
        for (size_t i = 0; i < DIM; i++) r[i] -= t[i];
        rho_2 = rho_1;
        iter.inc();
    }
    if (xd)
        for (size_t i = 0; i < DIM; i++) x[i] += (*xd)[i];
}

/* algebra::cg_dir (src/algebra/cg.h:67-121): conjugate gradient with the diagonal preconditioner,
   Dirichlet values xd on the masked indices ld */
void cg_dir(Iter &iter, const CSR &A, std::vector<double> &x, const std::vector<double> &rhs,
            const std::vector<double> &xd, const std::vector<int> &ld)
{
    const size_t DIM = x.size();
    std::vector<double> p(DIM), q(DIM), r(DIM), z(DIM), dp(DIM), b(rhs), vt(DIM);
    for (size_t i = 0; i < DIM; i++) {
        double c = 0;
        for (int64_t k = A.rowptr[i]; k < A.rowptr[i + 1]; k++)
            if (A.col[k] == (int)i) c = A.val[k];
        dp[i] = 1.0 / c;
    }
    spmv(A, xd, z);
    for (size_t i = 0; i < DIM; i++) b[i] -= z[i];
    mask(ld, b);
    mask(ld, dp);
    iter.rhsn = vnorm(b);
    r = b;
    spmv(A, x, vt);
    for (size_t i = 0; i < DIM; i++) r[i] -= vt[i];
    mask(ld, r);
    for (size_t i = 0; i < DIM; i++) z[i] = dp[i] * r[i];
    double rho = vdot(z, r), rho_1 = 0;
    p = z;
    while (!iter.converged(vnorm(r)) && iter.status != FG_ITER_OVERFLOW && iter.status != FG_ITER_CANNOT_CONVERGE) {
        if (iter.nit > 0) {
            for (size_t i = 0; i < DIM; i++) z[i] = dp[i] * r[i];
            rho = vdot(z, r);
            for (size_t i = 0; i < DIM; i++) p[i] *= rho / rho_1;
            for (size_t i = 0; i < DIM; i++) p[i] += z[i];
        }
        spmv(A, p, q);
        mask(ld, q);
        const double qp = vdot(q, p);
        if (qp == 0.0) {
            iter.status = FG_ITER_CANNOT_CONVERGE;
            break;
        }
        const double a = rho / qp;
        for (size_t i = 0; i < DIM; i++) x[i] += a * p[i];
        for (size_t i = 0; i < DIM; i++) r[i] += (-a) * q[i];
        rho_1 = rho;
        iter.inc();
    }
    for (size_t i = 0; i < DIM; i++) x[i] += xd[i];
}
} // namespace

extern "C" {

/* LinAlgebra::solve (src/solver.cpp:6-90) */
int orc_solve(orc_problem *p, double dt, orc_solve_info *info)
{
    std::fill(p->val.begin(), p->val.end(), 0.0);
    /* buildMat<4> (src/solver.h:319-339): K(2*i+di, 2*j+dj) += Kp(di*4+ie, dj*4+je) */
    for (int k : p->magTet) {
        const TetGeo &t = p->tet[k];
        const double *Kp = &p->Kp[64 * k];
        for (int ie = 0; ie < 4; ie++)
            for (int je = 0; je < 4; je++)
                for (int di = 0; di < 2; di++)
                    for (int dj = 0; dj < 2; dj++)
                        p->K(2 * t.ind[ie] + di, 2 * t.ind[je] + dj) += Kp[(di * 4 + ie) * 8 + dj * 4 + je];
    }
    std::fill(p->L.begin(), p->L.end(), 0.0);
    for (int k : p->magTet) {
        const TetGeo &t = p->tet[k];
        for (int ie = 0; ie < 4; ie++)
            for (int di = 0; di < 2; di++) p->L[2 * t.ind[ie] + di] += p->Lp[8 * k + di * 4 + ie];
    }
    for (int k : p->magTri) {
        const TriGeo &f = p->tri[k];
        for (int ie = 0; ie < 3; ie++)
            for (int di = 0; di < 2; di++) p->L[2 * f.ind[ie] + di] += p->LpT[6 * k + di * 3 + ie];
    }
    mask(p->lvd, p->L);
    for (int i : p->lvd) p->K(i, i) = 1.0;
    /* buildInitGuess (src/linear_algebra.cpp:13-24) */
    std::fill(p->X.begin(), p->X.end(), 0.0);
    for (int i = 0; i < p->N; i++)
        if (p->magNode[i]) {
            const Node &n = p->node[i];
            p->X[2 * i] = dot(n.d[1].v, n.ep) / GAMMA0;
            p->X[2 * i + 1] = dot(n.d[1].v, n.eq) / GAMMA0;
        }
    Iter iter(p->tol, p->max_iter);
    CSR A{2 * (int64_t)p->N, p->rowptr.data(), p->col.data(), p->val.data()};
    bicg_dir(iter, A, p->X, p->L, p->lvd);
    if (info) {
        info->status = iter.status;
        info->iterations = iter.nit;
        info->residual = iter.res;
        info->rhs_norm = iter.rhsn;
    }
    if (iter.status == FG_ITER_OVERFLOW || iter.status == FG_ITER_CANNOT_CONVERGE || iter.res > iter.resmax)
        return 1;
    double v2max = 0.0;
    for (int i = 0; i < p->N; i++)
        if (p->magNode[i]) {
            double vp = p->X[2 * i], vq = p->X[2 * i + 1];
            double v2 = vp * vp + vq * vq;
            if (v2 > v2max) v2max = v2;
            /* mesh::updateNode -> Node::make_evol (src/mesh.h:185-186, src/node.h:116-122) */
            Node &n = p->node[i];
            n.d[1].v = (vp * GAMMA0) * n.ep + (vq * GAMMA0) * n.eq;
            n.d[1].u = normalized(n.d[0].u + dt * n.d[1].v);
        }
    p->v_max = GAMMA0 * std::sqrt(v2max);
    if (info) info->v_max = p->v_max;
    return 0;
}

/* algebra::bicg_dir (5-argument overload) on a caller-supplied CSR matrix */
int orc_bicg_dir(int64_t n, const int64_t *rowptr, const int32_t *col, const double *val,
                 const double *rhs, double *x, int n_mask, const int32_t *mask_idx, double tol,
                 int max_iter, orc_solve_info *info)
{
    CSR A{n, rowptr, col, val};
    std::vector<double> X(x, x + n), b(rhs, rhs + n);
    std::vector<int> ld(mask_idx, mask_idx + n_mask);
    Iter iter(tol, max_iter);
    bicg_dir(iter, A, X, b, ld);
    std::copy(X.begin(), X.end(), x);
    if (info) {
        info->status = iter.status;
        info->iterations = iter.nit;
        info->residual = iter.res;
        info->rhs_norm = iter.rhsn;
        info->v_max = 0;
    }
    return iter.status;
}

/* algebra::cg_dir on a caller-supplied CSR matrix, Dirichlet values xd[mask_idx] */
int orc_cg_dir(int64_t n, const int64_t *rowptr, const int32_t *col, const double *val, const double *rhs,
               const double *xd, double *x, int n_mask, const int32_t *mask_idx, double tol, int max_iter,
               orc_solve_info *info)
{
    CSR A{n, rowptr, col, val};
    std::vector<double> X(x, x + n), b(rhs, rhs + n), XD(xd, xd + n);
    std::vector<int> ld(mask_idx, mask_idx + n_mask);
    Iter iter(tol, max_iter);
    cg_dir(iter, A, X, b, XD, ld);
    std::copy(X.begin(), X.end(), x);
    if (info) {
        info->status = iter.status;
        info->iterations = iter.nit;
        info->residual = iter.res;
        info->rhs_norm = iter.rhsn;
        info->v_max = 0;
    }
    return iter.status;
}

/* electrostatSolver::solve (src/electrostatSolver.cpp:96-136) after checkBoundaryConditions
   (:10-28): K = sum over the tetrahedra of sigma * sum_g w_g da_i . da_j (integrales, :42-65),
   L = - sum over the triangles of jn * a_ig w_g (:67-76), V fixed on the nodes of the triangles
   with a potential, cg_dir from V = 0. Returns -1 on bad boundary conditions, else the status. */
int orc_electrostatic_potential(const fg_mesh *m, const double *sigma, const double *jn, const double *Vbc,
                                double tol, int max_iter, double *V, orc_solve_info *info)
{
    int nJ = 0, nV = 0;
    for (int k = 0; k < m->n_surf_regions; k++) {
        nJ += std::isfinite(jn[k]);
        nV += std::isfinite(Vbc[k]);
    }
    if (nJ != 1 || nV != 1) return -1;
    const int N = m->n_nodes;
    std::vector<std::map<int, double>> rows(N);
    for (int t = 0; t < m->n_tets; t++) {
        V3 P[4];
        int ind[4];
        for (int a = 0; a < 4; a++) {
            ind[a] = m->tet_nodes[4 * t + a];
            P[a] = {m->pos[3 * ind[a]], m->pos[3 * ind[a] + 1], m->pos[3 * ind[a] + 2]};
        }
        double da[4][3], w[TNPI];
        if (tet_geometry(P, ind, da, w, nullptr)) return -1;
        const double sg = sigma[m->tet_region[t]];
        double AE[4][4] = {};
        for (int g = 0; g < TNPI; g++) {
            const double sw = sg * w[g];
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++)
                    AE[i][j] += sw * (da[i][0] * da[j][0] + da[i][1] * da[j][1] + da[i][2] * da[j][2]);
        }
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) rows[ind[i]][ind[j]] += AE[i][j];
    }
    std::vector<double> L(N, 0.0), Vd(N, 0.0);
    std::vector<int> ld;
    for (int f = 0; f < m->n_tris; f++) {
        const int rg = m->tri_region[f];
        const int *q = m->tri_nodes + 3 * f;
        V3 P[3];
        for (int a = 0; a < 3; a++) P[a] = {m->pos[3 * q[a]], m->pos[3 * q[a] + 1], m->pos[3 * q[a] + 2]};
        const double surf = 0.5 * norm(cross(P[1] - P[0], P[2] - P[0]));
        double J = rg >= 0 && rg < m->n_surf_regions ? jn[rg] : NAN;
        if (!std::isfinite(J)) J = 0.0;
        double BE[3] = {0, 0, 0};
        for (int g = 0; g < FNPI; g++) {
            const double Jw = J * (2.0 * surf * fpds[g]);
            for (int i = 0; i < 3; i++) BE[i] -= fa[i][g] * Jw;
        }
        for (int i = 0; i < 3; i++) L[q[i]] += BE[i];
        if (rg >= 0 && rg < m->n_surf_regions && !std::isnan(Vbc[rg]))
            for (int i = 0; i < 3; i++) {
                Vd[q[i]] = Vbc[rg];
                ld.push_back(q[i]);
            }
    }
    std::sort(ld.begin(), ld.end());
    ld.erase(std::unique(ld.begin(), ld.end()), ld.end());
    std::vector<int64_t> rp(N + 1, 0);
    std::vector<int32_t> cl;
    std::vector<double> vl;
    for (int i = 0; i < N; i++) {
        for (const auto &e : rows[i]) {
            cl.push_back(e.first);
            vl.push_back(e.second);
        }
        rp[i + 1] = (int64_t)cl.size();
    }
    CSR A{N, rp.data(), cl.data(), vl.data()};
    std::vector<double> X(N, 0.0);
    Iter iter(tol, max_iter);
    cg_dir(iter, A, X, L, Vd, ld);
    std::copy(X.begin(), X.end(), V);
    if (info) {
        info->status = iter.status;
        info->iterations = iter.nit;
        info->residual = iter.res;
        info->rhs_norm = iter.rhsn;
        info->v_max = 0;
    }
    return iter.status;
}

/* spinAcc::checkBoundaryConditions, boundaryConditions, solve and integrales
   (src/spinAccumulationSolver.cpp:41-287): the spin accumulation s (3 per node, interleaved) of
   the diffusion problem with the node magnetization u and potential V. Returns -1 on bad
   boundary conditions, else the bicg_dir status. */
int orc_spin_accumulation(const fg_mesh *m, const fg_spin_volume *vp, const fg_spin_surface *sp, const double *Vn,
                          const double *u, const double *u_src, double tol, int max_iter, double *s_out,
                          orc_solve_info *info)
{
    return orc_spin_accumulation_hist(m, vp, sp, Vn, 1, u, u_src, tol, max_iter, s_out, info);
}

/* The k-th solve of a run: solver::buildMat / buildVect add into K and L_rhs (src/solver.h:110-142)
 * and spinAcc::solve never clears them (src/spinAccumulationSolver.cpp:165-197), so the system is
 * the sum of the assemblies of the k calls, each its element terms then its surface terms, in the
 * reference's order; uh / ush hold the k magnetizations (ush NULL: uh). */
int orc_spin_accumulation_hist(const fg_mesh *m, const fg_spin_volume *vp, const fg_spin_surface *sp,
                               const double *Vn, int nhist, const double *uh, const double *ush, double tol,
                               int max_iter, double *s_out, orc_solve_info *info)
{
    if (nhist < 1) return -1;
    const double E = 1.602176634e-19, HBAR = 6.62607015e-34 / (2 * M_PI), ME = 9.1093837139e-31;
    const double MUB = E * HBAR / (2. * ME); /* src/config.h.in:33-37 */
    auto finite3 = [](const double *v) { return std::isfinite(std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])); };
    int nJ = 0, nS = 0;
    for (int k = 0; k < m->n_surf_regions; k++) {
        nJ += std::isfinite(sp[k].jn);
        nS += finite3(sp[k].s);
    }
    bool ok = nJ == 1 && nS == 1;
    for (int k = 0; k < m->n_vol_regions; k++)
        ok = ok && std::isfinite(vp[k].N0) && vp[k].N0 != 0 && std::isfinite(vp[k].P) && vp[k].P >= 0 && vp[k].P < 1;
    if (!ok) return -1;
    const int N = m->n_nodes, D = 3 * N;
    std::vector<std::map<int, double>> rows(D);
    std::vector<double> L(D, 0.0);
    std::vector<double> Vd(D, 0.0);
    std::vector<int> ld;
    for (int h = 0; h < nhist; h++) {
        /* the source's magnetization (mesh::getNode_u, NEXT) */
        const double *u = uh + (size_t)h * D, *u_src = (ush ? ush : uh) + (size_t)h * D;
        for (int t = 0; t < m->n_tets; t++) {
            V3 P[4];
            int ind[4];
            for (int a = 0; a < 4; a++) {
                ind[a] = m->tet_nodes[4 * t + a];
                P[a] = {m->pos[3 * ind[a]], m->pos[3 * ind[a] + 1], m->pos[3 * ind[a] + 2]};
            }
            double da[4][3], w[TNPI];
            if (tet_geometry(P, ind, da, w, nullptr)) return -1;
            const int rg = m->tet_region[t];
            const fg_spin_volume &pr = vp[rg];
            const double D0 = 2.0 * pr.sigma / ((E * E) * pr.N0);
            double aw[4], wsum = 0;
            for (int g = 0; g < TNPI; g++) wsum += w[g];
            for (int i = 0; i < 4; i++) {
                aw[i] = 0;
                for (int g = 0; g < TNPI; g++) aw[i] += ta[i][g] * w[g];
            }
            /* calcDiagBlock (src/tetra.cpp:133-140): da da^T (D0 sum w) + diag(D0 / lsf^2 a_w) */
            double AE[12][12] = {};
            const double c = D0 * wsum, itsf = D0 / (pr.lsf * pr.lsf);
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    double v = (da[i][0] * da[j][0] + da[i][1] * da[j][1] + da[i][2] * da[j][2]) * c;
                    if (i == j) v += itsf * aw[i];
                    for (int d = 0; d < 3; d++) AE[4 * d + i][4 * d + j] += v;
                }
            double BE[12] = {};
            if (m->vol[rg].Ms > 0) {
                const double itsd = D0 / (pr.lsd * pr.lsd);
                for (int i = 0; i < 4; i++) {
                    const double *ui = u + 3 * ind[i];
                    const double X = itsd * (aw[i] * ui[0]), Y = itsd * (aw[i] * ui[1]), Z = itsd * (aw[i] * ui[2]);
                    AE[4 + i][8 + i] += X;
                    AE[8 + i][4 + i] -= X;
                    AE[0 + i][8 + i] -= Y;
                    AE[8 + i][0 + i] += Y;
                    AE[0 + i][4 + i] += Z;
                    AE[4 + i][0 + i] -= Z;
                }
                /* the source cst0 grad V . da_i m_i (:262-286) */
                const double cst0 = MUB * pr.P * pr.sigma / E;
                double gV[3] = {0, 0, 0};
                for (int i = 0; i < 4; i++)
                    for (int k = 0; k < 3; k++) gV[k] += Vn[ind[i]] * da[i][k];
                for (int g = 0; g < TNPI; g++) {
                    const double cw = cst0 * w[g];
                    const double cg[3] = {cw * gV[0], cw * gV[1], cw * gV[2]};
                    for (int i = 0; i < 4; i++) {
                        const double *ui = u_src + 3 * ind[i];
                        const double tmp = cg[0] * da[i][0] + cg[1] * da[i][1] + cg[2] * da[i][2];
                        BE[i] += tmp * ui[0];
                        BE[4 + i] += tmp * ui[1];
                        BE[8 + i] += tmp * ui[2];
                    }
                }
            }
            /* solver::buildMat / buildVect (src/solver.h:110-143) */
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++)
                    for (int di = 0; di < 3; di++)
                        for (int dj = 0; dj < 3; dj++) rows[3 * ind[i] + di][3 * ind[j] + dj] += AE[4 * di + i][4 * dj + j];
            for (int i = 0; i < 4; i++)
                for (int d = 0; d < 3; d++) L[3 * ind[i] + d] += BE[4 * d + i];
        }
        /* boundary terms and Dirichlet values (:91-116, :173-197) */
        for (int f = 0; f < m->n_tris; f++) {
            const int rg = m->tri_region[f];
            const int *q = m->tri_nodes + 3 * f;
            if (rg < 0 || rg >= m->n_surf_regions) continue;
            const fg_spin_surface &pr = sp[rg];
            V3 P[3];
            for (int a = 0; a < 3; a++) P[a] = {m->pos[3 * q[a]], m->pos[3 * q[a] + 1], m->pos[3 * q[a] + 2]};
            const double surf = 0.5 * norm(cross(P[1] - P[0], P[2] - P[0]));
            double sv[3] = {pr.s[0], pr.s[1], pr.s[2]};
            if (std::isfinite(pr.jn) && finite3(pr.uP))
                for (int k = 0; k < 3; k++) sv[k] = -pr.jn * (MUB / E) * pr.uP[k];
            double Le[9] = {};
            if (finite3(sv))
                for (int g = 0; g < FNPI; g++) {
                    const double wg = 2.0 * surf * fpds[g];
                    for (int i = 0; i < 3; i++) {
                        const double aiw = wg * fa[i][g];
                        for (int k = 0; k < 3; k++) Le[3 * k + i] -= sv[k] * aiw;
                    }
                }
            for (int i = 0; i < 3; i++)
                for (int k = 0; k < 3; k++) L[3 * q[i] + k] += Le[3 * k + i];
            double dv[3];
            bool dirichlet = false;
            if (std::isnan(std::sqrt(pr.s[0] * pr.s[0] + pr.s[1] * pr.s[1] + pr.s[2] * pr.s[2])) && std::isfinite(pr.jn)) {
                for (int k = 0; k < 3; k++) dv[k] = -pr.jn * (MUB / E) * pr.uP[k];
                dirichlet = true;
            } else if (finite3(pr.s) && std::isnan(pr.jn)) {
                for (int k = 0; k < 3; k++) dv[k] = pr.s[k];
                dirichlet = true;
            }
            if (dirichlet && h == 0)
                for (int i = 0; i < 3; i++)
                    for (int k = 0; k < 3; k++) {
                        Vd[3 * q[i] + k] = dv[k];
                        ld.push_back(3 * q[i] + k);
                    }
        }
    }
    std::sort(ld.begin(), ld.end());
    ld.erase(std::unique(ld.begin(), ld.end()), ld.end());
    std::vector<int64_t> rp(D + 1, 0);
    std::vector<int32_t> cl;
    std::vector<double> vl;
    for (int i = 0; i < D; i++) {
        for (const auto &e : rows[i]) {
            cl.push_back(e.first);
            vl.push_back(e.second);
        }
        rp[i + 1] = (int64_t)cl.size();
    }
    CSR A{D, rp.data(), cl.data(), vl.data()};
    std::vector<double> X(D, 0.0);
    Iter iter(tol, max_iter);
    bicg_dir(iter, A, X, L, ld, &Vd);
    std::copy(X.begin(), X.end(), s_out);
    if (info) {
        info->status = iter.status;
        info->iterations = iter.nit;
        info->residual = iter.res;
        info->rhs_norm = iter.rhsn;
        info->v_max = 0;
    }
    return iter.status;
}

void orc_system_size(const orc_problem *p, int64_t *rows, int64_t *nnz)
{
    *rows = 2 * p->N;
    *nnz = (int64_t)p->col.size();
}

void orc_get_system(const orc_problem *p, int64_t *rowptr, int32_t *col, double *val, double *rhs)
{
    if (rowptr) std::copy(p->rowptr.begin(), p->rowptr.end(), rowptr);
    if (col) std::copy(p->col.begin(), p->col.end(), col);
    if (val) std::copy(p->val.begin(), p->val.end(), val);
    if (rhs) std::copy(p->L.begin(), p->L.end(), rhs);
}

void orc_get_solution(const orc_problem *p, double *x) { std::copy(p->X.begin(), p->X.end(), x); }

void orc_get_tet_element(const orc_problem *p, int k, double *Kp, double *Lp)
{
    if (Kp) std::memcpy(Kp, &p->Kp[64 * k], 64 * sizeof(double));
    if (Lp) std::memcpy(Lp, &p->Lp[8 * k], 8 * sizeof(double));
}

} // extern "C"

namespace {
/* fmm::calc_charges + the exact potential (src/fmm_demag.h:155-224, src/tetra.cpp:347-359,
 * src/triangle.cpp:45-78): phi_i = (sum_s q_s / |x_i - y_s| + corr_i) / (4 pi) */
void demag_charges(orc_problem *p, bool of_v)
{
    const int nt = (int)p->magTet.size(), nf = (int)p->magTri.size();
    const size_t S = (size_t)TNPI * nt + (size_t)FNPI * nf;
    p->src_xyz.resize(3 * S);
    std::vector<double> &q = of_v ? p->src_qv : p->src_qu;
    std::vector<double> &corr = of_v ? p->corr_v : p->corr_u;
    q.assign(S, 0.0);
    size_t ns = 0;
    for (int k : p->magTet) { /* Tet::charges: -Ms * w_g * div(u) */
        const TetGeo &t = p->tet[k];
        V3 vn[4];
        V3 P[4];
        for (int i = 0; i < 4; i++) {
            const Node &n = p->node[t.ind[i]];
            vn[i] = of_v ? n.d[1].v : n.d[1].u;
            P[i] = n.p;
        }
        double qt[TNPI];
        tet_charges(t.da, t.w, vn, p->vol[t.reg].Ms, qt);
        for (int g = 0; g < TNPI; g++) {
            V3 G{0, 0, 0};
            for (int i = 0; i < 4; i++) G = G + ta[i][g] * P[i];
            p->src_xyz[3 * ns] = G.x;
            p->src_xyz[3 * ns + 1] = G.y;
            p->src_xyz[3 * ns + 2] = G.z;
            q[ns] = qt[g];
            ns++;
        }
    }
    std::fill(corr.begin(), corr.end(), 0.0);
    for (int k : p->magTri) { /* Tri::charges + correctionCharges */
        const TriGeo &f = p->tri[k];
        V3 vn[3], P[3];
        for (int i = 0; i < 3; i++) {
            const Node &n = p->node[f.ind[i]];
            vn[i] = of_v ? n.d[1].v : n.d[1].u;
            P[i] = n.p;
        }
        double qg[FNPI];
        V3 G[FNPI];
        tri_charges(f, vn, qg);
        for (int g = 0; g < FNPI; g++) {
            V3 Gp{0, 0, 0};
            for (int i = 0; i < 3; i++) Gp = Gp + fa[i][g] * P[i];
            G[g] = Gp;
            p->src_xyz[3 * ns] = Gp.x;
            p->src_xyz[3 * ns + 1] = Gp.y;
            p->src_xyz[3 * ns + 2] = Gp.z;
            q[ns] = qg[g];
            ns++;
        }
        double s[3] = {dot(vn[0], f.n), dot(vn[1], f.n), dot(vn[2], f.n)};
        for (int i = 0; i < 3; i++) {
            for (int g = 0; g < FNPI; g++) corr[f.ind[i]] -= qg[g] / norm(P[i] - G[g]);
            corr[f.ind[i]] += tri_potential(P, s, f.surf, f.dMs, i);
        }
    }
}

double direct_potential(const orc_problem *p, bool of_v, int i)
{
    const std::vector<double> &q = of_v ? p->src_qv : p->src_qu;
    const std::vector<double> &corr = of_v ? p->corr_v : p->corr_u;
    const V3 x = p->node[i].p;
    const double *sx = p->src_xyz.data();
    const size_t S = q.size();
    double pot = 0;
    for (size_t s2 = 0; s2 < S; s2++) {
        double dx = x.x - sx[3 * s2], dy = x.y - sx[3 * s2 + 1], dz = x.z - sx[3 * s2 + 2];
        pot += q[s2] / std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    return (pot + corr[i]) / (4 * M_PI);
}

void demag_one(orc_problem *p, bool of_v)
{
    demag_charges(p, of_v);
    const int N = p->N;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; i++) {
        if (!p->magNode[i]) continue;
        double val = direct_potential(p, of_v, i);
        if (of_v) p->node[i].d[1].phiv = val;
        else p->node[i].d[1].phi = val;
    }
}

/* Both potentials by the CPU FMM (fmm_cpu.cpp) at the magnetic nodes, plus the corrections */
void demag_fmm(orc_problem *p)
{
    demag_charges(p, false);
    demag_charges(p, true);
    std::vector<double> txyz;
    std::vector<int> tn;
    for (int i = 0; i < p->N; i++)
        if (p->magNode[i]) {
            tn.push_back(i);
            txyz.push_back(p->node[i].p.x);
            txyz.push_back(p->node[i].p.y);
            txyz.push_back(p->node[i].p.z);
        }
    std::vector<double> pu, pv;
    orc_fmm::potentials(p->src_xyz, p->src_qu, p->src_qv, txyz, p->fmm_order, p->fmm_leaf, pu, pv);
    for (size_t k = 0; k < tn.size(); k++) {
        p->node[tn[k]].d[1].phi = (pu[k] + p->corr_u[tn[k]]) / (4 * M_PI);
        p->node[tn[k]].d[1].phiv = (pv[k] + p->corr_v[tn[k]]) / (4 * M_PI);
    }
}
} // namespace

extern "C" {

/* fmm::calc_demag (src/fmm_demag.h:96-101): exact direct sums, or the CPU FMM after orc_set_demag */
void orc_calc_demag(orc_problem *p)
{
    if (p->fmm_order > 0) {
        demag_fmm(p);
        return;
    }
    demag_one(p, false);
    demag_one(p, true);
}

/* Demag method of orc_calc_demag: order 0 = exact direct sums (the default); order > 0 = the CPU
 * FMM of that order with leaves of `leaf` sources on average (the CPU-baseline timing). */
int orc_set_demag(orc_problem *p, int order, int leaf)
{
    if (order < 0 || order > 12 || leaf < 1) return -1;
    p->fmm_order = order;
    p->fmm_leaf = leaf;
    return 0;
}

/* The CPU FMM on raw sources and targets (tests): potentials without corrections or 1/(4 pi);
 * returns the leaf level. */
int orc_fmm_potentials(const double *sxyz, int64_t S, const double *qu, const double *qv, const double *txyz,
                       int64_t T, int order, int leaf, double *pu, double *pv)
{
    std::vector<double> a(sxyz, sxyz + 3 * S), b(qu, qu + S), c(qv, qv + S), t(txyz, txyz + 3 * T), ou, ov;
    const int L = orc_fmm::potentials(a, b, c, t, order, leaf, ou, ov);
    std::copy(ou.begin(), ou.end(), pu);
    std::copy(ov.begin(), ov.end(), pv);
    return L;
}

/* The charges of both potentials and the exact potentials at n given targets only (CPU-baseline
 * sample of the demag cost); the node state is not modified. */
void orc_demag_sample(orc_problem *p, const int32_t *targets, int n, double *phi, double *phiv)
{
    demag_charges(p, false);
    demag_charges(p, true);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < n; k++) {
        phi[k] = direct_potential(p, false, targets[k]);
        phiv[k] = direct_potential(p, true, targets[k]);
    }
}

int64_t orc_n_sources(const orc_problem *p) { return (int64_t)p->src_qu.size(); }

void orc_get_sources(const orc_problem *p, double *xyz, double *q_u, double *q_v)
{
    if (xyz) std::copy(p->src_xyz.begin(), p->src_xyz.end(), xyz);
    if (q_u) std::copy(p->src_qu.begin(), p->src_qu.end(), q_u);
    if (q_v) std::copy(p->src_qv.begin(), p->src_qv.end(), q_v);
}

void orc_get_corrections(const orc_problem *p, double *cu, double *cv)
{
    if (cu) std::copy(p->corr_u.begin(), p->corr_u.end(), cu);
    if (cv) std::copy(p->corr_v.begin(), p->corr_v.end(), cv);
}

/* Fem::energy (src/energy.cpp:5-68) for a uniform field; element energies src/tetra.cpp:309-391,
 * src/triangle.cpp:38-43,80-85 */
static void energy(orc_problem *p, const double Hext_[3], const double *hsp, double amp, double E[4])
{
    V3 Hext = v3(Hext_);
    for (int k = 0; k < 4; k++) E[k] = 0;
    for (int k : p->magTet) {
        const TetGeo &t = p->tet[k];
        const fg_volume_region &prm = p->vol[t.reg];
        V3 u[4];
        double phi[4];
        for (int i = 0; i < 4; i++) {
            u[i] = p->node[t.ind[i]].d[1].u;
            phi[i] = p->node[t.ind[i]].d[1].phi;
        }
        V3 Hg[TNPI];
        for (int g = 0; g < TNPI; g++) /* Tet::zeemanEnergy, src/tetra.cpp:374-391 */
            Hg[g] = hsp ? amp * V3{hsp[15 * (size_t)k + 3 * g], hsp[15 * (size_t)k + 3 * g + 1],
                                   hsp[15 * (size_t)k + 3 * g + 2]}
                        : Hext;
        double e[5];
        tet_energy_terms(t.da, t.w, u, phi, prm, Hg, e);
        E[FG_E_EXCHANGE] += e[0];
        E[FG_E_DEMAG] += e[3];
        if (prm.K != 0.0) E[FG_E_ANISOTROPY] += e[1];
        if (prm.K3 != 0.0) E[FG_E_ANISOTROPY] += e[2];
        E[FG_E_ZEEMAN] += e[4];
    }
    for (int k : p->magTri) {
        const TriGeo &f = p->tri[k];
        V3 u[3];
        double phi[3];
        for (int i = 0; i < 3; i++) {
            u[i] = p->node[f.ind[i]].d[1].u;
            phi[i] = p->node[f.ind[i]].d[1].phi;
        }
        const bool ks = !p->surf.empty() && p->surf[f.reg].Ks != 0.0;
        double e[2];
        tri_energy_terms(f, u, phi, ks ? p->surf[f.reg].Ks : 0.0, ks ? v3(p->surf[f.reg].uk) : V3{0, 0, 0}, e);
        if (ks) E[FG_E_ANISOTROPY] += e[0];
        E[FG_E_DEMAG] += e[1];
    }
}

void orc_energy(orc_problem *p, const double Hext[3], double E[4]) { energy(p, Hext, nullptr, 0.0, E); }

void orc_energy_amp(orc_problem *p, double amp, double E[4])
{
    const double zero[3] = {0, 0, 0};
    energy(p, zero, p->hsp.empty() ? nullptr : p->hsp.data(), amp, E);
}

/* Fem::evolution -> Node::evolution (src/node.h:107) */
void orc_evolution(orc_problem *p)
{
    for (auto &n : p->node) n.d[0] = n.d[1];
}

/* mesh::max_angle (src/mesh.h:295-306) */
double orc_max_angle(const orc_problem *p)
{
    double m = 1.0;
    for (auto &e : p->edges) {
        double d = dot(p->node[e.first].d[1].u, p->node[e.second].d[1].u);
        m = std::min(m, d);
    }
    return std::acos(m);
}

/* mesh::avg (src/mesh.cpp:89-106) on NEXT u (which 0) or v (which 1) */
void orc_average(const orc_problem *p, int which, double out[3])
{
    for (int c = 0; c < 3; c++) {
        double sum = 0;
        for (int k : p->magTet) {
            const TetGeo &t = p->tet[k];
            double s = 0;
            for (int g = 0; g < TNPI; g++) {
                double val = 0;
                for (int i = 0; i < 4; i++) {
                    const DataNode &d = p->node[t.ind[i]].d[1];
                    val += comp(which ? d.v : d.u, c) * ta[i][g];
                }
                s += t.w[g] * val;
            }
            sum += s;
        }
        out[c] = sum / p->totalMagVol;
    }
}

/* ---------------- elementary entry points for the known-answer tests ---------------- */
int orc_tet_geometry(const double *pos, int32_t *ind, double *da, double *w)
{
    V3 P[4];
    for (int i = 0; i < 4; i++) P[i] = v3(pos + 3 * i);
    int id[4] = {ind[0], ind[1], ind[2], ind[3]};
    double d[4][3];
    int rc = tet_geometry(P, id, d, w, nullptr);
    if (rc) return rc;
    for (int i = 0; i < 4; i++) {
        ind[i] = id[i];
        for (int k = 0; k < 3; k++) da[3 * i + k] = d[i][k];
    }
    return 0;
}

void orc_alpha_eff(double dt, double alpha, const double *uHeff, int n, double *out)
{
    for (int i = 0; i < n; i++) out[i] = alpha_eff(dt, alpha, uHeff[i]);
}

void orc_tet_lumping(const double *da_, const double *w, const double *u_nod, const double *aeff,
                     double prefactor, double *AE_)
{
    double da[4][3];
    V3 u[4];
    for (int i = 0; i < 4; i++) {
        for (int k = 0; k < 3; k++) da[i][k] = da_[3 * i + k];
        u[i] = v3(u_nod + 3 * i);
    }
    double AE[12][12];
    std::memcpy(AE, AE_, sizeof(AE));
    tet_lumping(da, w, u, aeff, prefactor, AE);
    std::memcpy(AE_, AE, sizeof(AE));
}

void orc_tet_integrales(const double *da_, const double *w, const double *u_, const double *v_,
                        const double *phi, const double *phiv, const double *ep_, const double *eq_,
                        const fg_volume_region *prm, const double *Hext, double dt, double prefactor,
                        double *Kp_, double *Lp)
{
    double da[4][3];
    V3 u[4], v[4], ep[4], eq[4];
    for (int i = 0; i < 4; i++) {
        for (int k = 0; k < 3; k++) da[i][k] = da_[3 * i + k];
        u[i] = v3(u_ + 3 * i);
        v[i] = v3(v_ + 3 * i);
        ep[i] = v3(ep_ + 3 * i);
        eq[i] = v3(eq_ + 3 * i);
    }
    double Kp[8][8];
    V3 H[TNPI];
    for (int g = 0; g < TNPI; g++) H[g] = v3(Hext);
    tet_integrales(da, w, u, v, phi, phiv, ep, eq, *prm, H, dt, prefactor, Kp, Lp);
    std::memcpy(Kp_, Kp, sizeof(Kp));
}

/* Tet::calc_aniso_uniax / calc_aniso_cub on Gauss-point values U, V [5][3]: H_aniso [5][3] and the
   uHeff contributions [5] (zeroed first). */
void orc_aniso_uniax(const double *U_, const double *V_, const double *uk, double Kbis, double s_dt, double *H,
                     double *contrib)
{
    V3 U[TNPI], V[TNPI], Han[TNPI];
    for (int g = 0; g < TNPI; g++) {
        U[g] = v3(U_ + 3 * g);
        V[g] = v3(V_ + 3 * g);
        Han[g] = {0, 0, 0};
        contrib[g] = 0.0;
    }
    aniso_uniax(U, V, v3(uk), Kbis, s_dt, Han, contrib);
    for (int g = 0; g < TNPI; g++) {
        H[3 * g] = Han[g].x;
        H[3 * g + 1] = Han[g].y;
        H[3 * g + 2] = Han[g].z;
    }
}

void orc_aniso_cub(const double *U_, const double *V_, const double *ex, const double *ey, const double *ez,
                   double K3bis, double s_dt, double *H, double *contrib)
{
    V3 U[TNPI], V[TNPI], Han[TNPI];
    for (int g = 0; g < TNPI; g++) {
        U[g] = v3(U_ + 3 * g);
        V[g] = v3(V_ + 3 * g);
        Han[g] = {0, 0, 0};
        contrib[g] = 0.0;
    }
    aniso_cub(U, V, v3(ex), v3(ey), v3(ez), K3bis, s_dt, Han, contrib);
    for (int g = 0; g < TNPI; g++) {
        H[3 * g] = Han[g].x;
        H[3 * g + 1] = Han[g].y;
        H[3 * g + 2] = Han[g].z;
    }
}

/* Per-element energies: orc_tet_energy e[5] (exchange, uniaxial, cubic, demag, Zeeman with the field
   Hg [5][3] at the Gauss points) from node values u [4][3], phi [4]; orc_tri_energy e[2] (surface
   anisotropy, demag) of the triangle of positions pos [3][3] with node values u [3][3], phi [3]. */
void orc_tet_energy(const double *da_, const double *w, const double *u_, const double *phi,
                    const fg_volume_region *prm, const double *Hg_, double *e)
{
    double da[4][3];
    V3 u[4], Hg[TNPI];
    for (int i = 0; i < 4; i++) {
        for (int k = 0; k < 3; k++) da[i][k] = da_[3 * i + k];
        u[i] = v3(u_ + 3 * i);
    }
    for (int g = 0; g < TNPI; g++) Hg[g] = v3(Hg_ + 3 * g);
    tet_energy_terms(da, w, u, phi, *prm, Hg, e);
}

void orc_tri_energy(const double *pos, const double *u_, const double *phi, double dMs, double Ks, const double *uk,
                    double *e)
{
    TriGeo f;
    V3 P[3] = {v3(pos), v3(pos + 3), v3(pos + 6)};
    tri_geometry(P, f);
    f.dMs = dMs;
    V3 u[3] = {v3(u_), v3(u_ + 3), v3(u_ + 6)};
    tri_energy_terms(f, u, phi, Ks, v3(uk), e);
}

/* Per-element charges: orc_tet_charges q[5] (Tet::charges) from da [4][3], w [5], u [4][3]; orc_tri_charges
   q[4] (Tri::charges) of the triangle pos [3][3] with node values u [3][3]. */
void orc_tet_charges(const double *da_, const double *w, const double *u_, double Ms, double *q)
{
    double da[4][3];
    V3 u[4];
    for (int i = 0; i < 4; i++) {
        for (int k = 0; k < 3; k++) da[i][k] = da_[3 * i + k];
        u[i] = v3(u_ + 3 * i);
    }
    tet_charges(da, w, u, Ms, q);
}

void orc_tri_charges(const double *pos, const double *u_, double dMs, double *q)
{
    TriGeo f;
    V3 P[3] = {v3(pos), v3(pos + 3), v3(pos + 6)};
    tri_geometry(P, f);
    f.dMs = dMs;
    V3 u[3] = {v3(u_), v3(u_ + 3), v3(u_ + 6)};
    tri_charges(f, u, q);
}

/* Tri::integrales (surface anisotropy) of the triangle pos [3][3]: Lp[6] from u, ep, eq [3][3]. */
void orc_tri_integrales(const double *pos, const double *u_, const double *ep_, const double *eq_, double dMs,
                        double Ks, const double *uk, double *Lp)
{
    TriGeo f;
    V3 P[3] = {v3(pos), v3(pos + 3), v3(pos + 6)};
    tri_geometry(P, f);
    f.dMs = dMs;
    V3 u[3], ep[3], eq[3];
    for (int i = 0; i < 3; i++) {
        u[i] = v3(u_ + 3 * i);
        ep[i] = v3(ep_ + 3 * i);
        eq[i] = v3(eq_ + 3 * i);
    }
    tri_integrales(f, u, ep, eq, Ks, v3(uk), Lp);
}

/* Gauss tables: a[4][5] of the tetrahedron and a[3][4] of the triangle, with their weights. */
void orc_gauss_tables(double *ta_, double *tpds_, double *fa_, double *fpds_)
{
    for (int i = 0; i < TN; i++)
        for (int g = 0; g < TNPI; g++) ta_[TNPI * i + g] = ta[i][g];
    for (int g = 0; g < TNPI; g++) tpds_[g] = tpds[g];
    for (int i = 0; i < FN; i++)
        for (int g = 0; g < FNPI; g++) fa_[FNPI * i + g] = fa[i][g];
    for (int g = 0; g < FNPI; g++) fpds_[g] = fpds[g];
}

double orc_tri_potential(const double *pos, const double *s, double dMs, int i)
{
    V3 P[3] = {v3(pos), v3(pos + 3), v3(pos + 6)};
    double surf = 0.5 * norm(cross(P[1] - P[0], P[2] - P[0]));
    return tri_potential(P, s, surf, dMs, i);
}

void orc_set_basis(const double *u, double r, double *ep, double *eq)
{
    V3 e, q;
    set_basis(v3(u), r, e, q);
    ep[0] = e.x; ep[1] = e.y; ep[2] = e.z;
    eq[0] = q.x; eq[1] = q.y; eq[2] = q.z;
}

void orc_mt_uniform(uint32_t seed, int n, double *out)
{
    std::mt19937 gen(seed);
    std::uniform_real_distribution<> distrib(0.0, 1.0);
    for (int i = 0; i < n; i++) out[i] = distrib(gen);
}

void orc_basis_angles(uint32_t seed, int n, double *out)
{
    srand(seed);
    for (int i = 0; i < n; i++) {
        std::mt19937 gen(rand());
        std::uniform_real_distribution<> distrib(0.0, 1.0);
        out[i] = M_2_PI * distrib(gen);
    }
}

/* ---------------- the reference time loop ---------------- */
int orc_run(const fg_mesh *mesh, const double *u0, const orc_run_settings *s, double *evol,
            int32_t *n_rows)
{
    *n_rows = 0;
    orc_problem *p = orc_create(mesh, s->tol, s->max_iter);
    if (!p) return -1;
    srand(s->seed); /* src/main.cpp:211 */
    auto next_angle = []() {
        std::mt19937 gen(rand());
        std::uniform_real_distribution<> distrib(0.0, 1.0);
        return M_2_PI * distrib(gen);
    };
    /* Mesh::init_distrib (src/mesh.h:221-277): NEXT u, phi = phiv = 0; CURRENT u too */
    std::vector<double> u(3 * p->N);
    for (int i = 0; i < p->N; i++) {
        V3 m = p->magNode[i] ? normalized(v3(u0 + 3 * i)) : V3{NAN, NAN, NAN};
        u[3 * i] = m.x; u[3 * i + 1] = m.y; u[3 * i + 2] = m.z;
    }
    orc_set_state(p, FG_BOTH, u.data(), nullptr, nullptr, nullptr);
    orc_base_projection(p, next_angle()); /* LinAlgebra ctor (src/linear_algebra.h:49) */
    V3 Hext = (1.0 / MU0) * v3(s->Bext);  /* Settings::getField (src/settings.h:290-297) */
    double Hx[3] = {Hext.x, Hext.y, Hext.z};
    /* timing (src/time_integration.h:6-59) */
    const double DTMIN = s->dt_min, DTMAX = s->dt_max, TAUR = 100. * DTMAX;
    double t = 0, dt = 0, prefactor = 1;
    auto set_dt = [&](double d) {
        dt = d;
        double tt = d / TAUR;
        prefactor = 1. + tt * std::fabs(std::log(tt));
    };
    set_dt(std::sqrt(DTMIN * DTMAX));
    double E[4];
    double vmax = 0;
    auto compute_all = [&](double) { /* src/fem.h:203-224 */
        orc_calc_demag(p);
        orc_energy(p, Hx, E);
        orc_evolution(p);
    };
    compute_all(t);
    const double t_initial = t, t_step = s->evol_time_step;
    const int step_count = (int)std::lround((s->final_time - t_initial) / t_step);
    /* TimeStepper (src/time_integration.cpp:11-38) */
    const double hard_min = DTMIN, hard_max = DTMAX * (1 + FLT_EPSILON);
    double soft_max = dt;
    auto stepper = [&](double stride) {
        double step = std::min(stride, soft_max);
        if (step > stride - 2 * hard_min && step < stride) step = stride - 2 * hard_min;
        soft_max = std::max(soft_max, std::min(step * 1.1, hard_max));
        return step;
    };
    auto set_soft_limit = [&](double mx) { soft_max = std::min(soft_max, mx); };
    int status = 0, nt = 0;
    for (int step_nb = 0; step_nb <= step_count; step_nb++) {
        double t_target = t_initial + step_nb * t_step;
        while (t < t_target) {
            set_dt(stepper(t_target - t));
            bool last_step = (dt == t_target - t);
            if (dt < DTMIN) {
                status = 1;
                goto bailout;
            }
            orc_base_projection(p, next_angle());
            orc_prepare_elements(p, Hx, dt, prefactor);
            orc_solve_info info;
            bool err = orc_solve(p, dt, &info) != 0;
            vmax = p->v_max;
            if (s->verbose)
                std::printf("t=%.6e dt=%.4e status=%d nit=%d res=%.3e\n", t, dt, info.status, info.iterations,
                            info.residual);
            if (err) {
                set_soft_limit(dt / 2);
                continue;
            }
            double dumax = dt * vmax;
            set_soft_limit((s->du_max / vmax) * 0.95);
            if (dumax > s->du_max) continue;
            compute_all(t);
            nt++;
            if (last_step) t = t_target;
            else t += dt;
        }
        if (*n_rows < s->max_rows) { /* Fem::saver (src/save.cpp:13-162) */
            double *row = evol + 12 * (*n_rows);
            double m[3];
            orc_average(p, 0, m);
            row[0] = t;
            row[1] = m[0]; row[2] = m[1]; row[3] = m[2];
            row[4] = E[FG_E_EXCHANGE]; row[5] = E[FG_E_ANISOTROPY]; row[6] = E[FG_E_DEMAG];
            row[7] = E[FG_E_ZEEMAN]; row[8] = E[0] + E[1] + E[2] + E[3];
            row[9] = dt; row[10] = vmax * dt; row[11] = nt;
            (*n_rows)++;
        }
    }
bailout:
    orc_destroy(p);
    return status;
}

} // extern "C"

/* ---------------- learned initial guess: the network of the engine's guess, restated ----------------
 * Same parameter layout and formulas as the engine (documented in include/fgmi.h): features
 * f = [u, s_h h, s_v v, s_d d], neighbour mean over the given adjacency (magnetic neighbours only),
 * input [f, mean], 24 -> 64 -> 64 -> 3 with the tanh GELU, v_pred = P_u (v_base + s_out o) with the
 * given base per node (the previous v when v_base is NULL), loss sum |v_pred - v*|^2 /
 * (n_mag s_out^2); loss_ref scores the reference guess P_u v the same way. Arithmetic in double; with round_bf16 the weight
 * matrices are first rounded to BF16 (round to nearest even) as the engine's forward pass uses
 * them. Gradients are with respect to the (rounded) weights and the biases. */
namespace {
const int GNF = 12, GNX = 24, GNH = 64, GNO = 3;
const int G_W1 = 0, G_B1 = G_W1 + GNH * GNX, G_W2 = G_B1 + GNH, G_B2 = G_W2 + GNH * GNH, G_W3 = G_B2 + GNH,
          G_B3 = G_W3 + GNO * GNH, G_NTRAIN = G_B3 + GNO, G_S = G_NTRAIN;
double g_bf16(float x)
{
    uint32_t b;
    std::memcpy(&b, &x, 4);
    b = (b + 0x7FFFu + ((b >> 16) & 1u)) & 0xFFFF0000u;
    float y;
    std::memcpy(&y, &b, 4);
    return y;
}
double g_gelu(double z) { return 0.5 * z * (1.0 + std::tanh(0.7978845608028654 * (z + 0.044715 * z * z * z))); }
double g_gelu_grad(double z)
{
    const double k = 0.7978845608028654, t = std::tanh(k * (z + 0.044715 * z * z * z));
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * k * (1.0 + 3.0 * 0.044715 * z * z);
}
} // namespace

extern "C" {

void orc_guess_eval(const float *params, int round_bf16, int n, const float *feat, const float *v_base,
                    const int64_t *adj_ptr, const int32_t *adj_idx, const int32_t *mag, const double *v_target,
                    double *v_pred, double *loss, double *loss_ref, double *grad)
{
    std::vector<double> W(G_NTRAIN);
    for (int k = 0; k < G_NTRAIN; k++) {
        const bool is_w = (k >= G_W1 && k < G_B1) || (k >= G_W2 && k < G_B2) || (k >= G_W3 && k < G_B3);
        W[k] = (round_bf16 && is_w) ? g_bf16(params[k]) : (double)params[k];
    }
    const double sh = params[G_S], sv = params[G_S + 1], so = params[G_S + 2], sd = params[G_S + 3];
    int nmag = 0;
    for (int i = 0; i < n; i++) nmag += mag[i] != 0;
    const double inv_norm = 1.0 / (std::max(1, nmag) * so * so);
    if (grad) std::fill(grad, grad + G_NTRAIN, 0.0);
    double L = 0, Lr = 0;
    const double sc[GNF] = {1, 1, 1, sh, sh, sh, sv, sv, sv, sd, sd, sd};
    for (int i = 0; i < n; i++) {
        if (!mag[i]) {
            if (v_pred) v_pred[3 * i] = v_pred[3 * i + 1] = v_pred[3 * i + 2] = 0;
            continue;
        }
        double x[GNX], a[GNF] = {0};
        int cnt = 0;
        for (int k = 0; k < GNF; k++) x[k] = feat[(size_t)i * GNF + k] * sc[k];
        for (int64_t e = adj_ptr[i]; e < adj_ptr[i + 1]; e++) {
            const int j = adj_idx[e];
            if (j == i || !mag[j]) continue;
            cnt++;
            for (int k = 0; k < GNF; k++) a[k] += feat[(size_t)j * GNF + k];
        }
        for (int k = 0; k < GNF; k++) x[GNF + k] = cnt ? a[k] / cnt * sc[k] : 0.0;
        double z1[GNH], h1[GNH], z2[GNH], h2[GNH], out[GNO];
        for (int o = 0; o < GNH; o++) {
            double z = W[G_B1 + o];
            for (int k = 0; k < GNX; k++) z += W[G_W1 + o * GNX + k] * x[k];
            z1[o] = z;
            h1[o] = g_gelu(z);
        }
        for (int o = 0; o < GNH; o++) {
            double z = W[G_B2 + o];
            for (int k = 0; k < GNH; k++) z += W[G_W2 + o * GNH + k] * h1[k];
            z2[o] = z;
            h2[o] = g_gelu(z);
        }
        for (int c = 0; c < GNO; c++) {
            double z = W[G_B3 + c];
            for (int k = 0; k < GNH; k++) z += W[G_W3 + c * GNH + k] * h2[k];
            out[c] = z;
        }
        const double u[3] = {feat[(size_t)i * GNF], feat[(size_t)i * GNF + 1], feat[(size_t)i * GNF + 2]};
        const double vp[3] = {feat[(size_t)i * GNF + 6], feat[(size_t)i * GNF + 7], feat[(size_t)i * GNF + 8]};
        double v[3], vr[3];
        for (int k = 0; k < 3; k++) {
            v[k] = (v_base ? (double)v_base[3 * i + k] : vp[k]) + so * out[k];
            vr[k] = vp[k];
        }
        const double un = u[0] * v[0] + u[1] * v[1] + u[2] * v[2], ur = u[0] * vr[0] + u[1] * vr[1] + u[2] * vr[2];
        for (int k = 0; k < 3; k++) {
            v[k] -= un * u[k];
            vr[k] -= ur * u[k];
        }
        if (v_pred)
            for (int k = 0; k < 3; k++) v_pred[3 * i + k] = v[k];
        if (!v_target) continue;
        double e[3];
        for (int k = 0; k < 3; k++) {
            e[k] = v[k] - v_target[3 * i + k];
            L += e[k] * e[k] * inv_norm;
            Lr += (vr[k] - v_target[3 * i + k]) * (vr[k] - v_target[3 * i + k]) * inv_norm;
        }
        if (!grad) continue;
        const double eu = u[0] * e[0] + u[1] * e[1] + u[2] * e[2];
        double d3o[GNO];
        for (int k = 0; k < 3; k++) d3o[k] = 2.0 * inv_norm * so * (e[k] - eu * u[k]);
        double dh2[GNH] = {0}, dh1[GNH] = {0};
        for (int c = 0; c < GNO; c++) {
            grad[G_B3 + c] += d3o[c];
            for (int k = 0; k < GNH; k++) {
                grad[G_W3 + c * GNH + k] += d3o[c] * h2[k];
                dh2[k] += W[G_W3 + c * GNH + k] * d3o[c];
            }
        }
        for (int o = 0; o < GNH; o++) {
            const double dz = dh2[o] * g_gelu_grad(z2[o]);
            grad[G_B2 + o] += dz;
            for (int k = 0; k < GNH; k++) {
                grad[G_W2 + o * GNH + k] += dz * h1[k];
                dh1[k] += W[G_W2 + o * GNH + k] * dz;
            }
        }
        for (int o = 0; o < GNH; o++) {
            const double dz = dh1[o] * g_gelu_grad(z1[o]);
            grad[G_B1 + o] += dz;
            for (int k = 0; k < GNX; k++) grad[G_W1 + o * GNX + k] += dz * x[k];
        }
    }
    if (loss) *loss = L;
    if (loss_ref) *loss_ref = Lr;
}

} // extern "C"
