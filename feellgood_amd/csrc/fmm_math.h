// fmm_math.h -- solid-harmonic expansions of the 1/r kernel, shared by the host reference and
// the gfx950 kernels.
//
// Scaled solid harmonics (Racah-normalised R, I divided / multiplied by sqrt((n+m)!(n-m)!)):
//   R_n^m(x) = r^n P_n^m(cos t) e^{i m f} / (n+m)!        (regular)
//   I_n^m(x) = (n-m)! P_n^m(cos t) e^{i m f} / r^{n+1}    (irregular)
// with X_n^{-m} = (-1)^m conj(X_n^m). They satisfy
//   1/|r - a|      = sum_{n,m} conj(I_n^m(r)) R_n^m(a)                         (|a| < |r|)
//   R_n^m(r + a)   = sum_{k,l} R_k^l(r) R_{n-k}^{m-l}(a)
//   I_n^m(r + a)   = sum_{k,l} (-1)^k conj(R_k^l(a)) I_{n+k}^{m+l}(r)          (|a| < |r|)
// from which the operators below follow:
//   P2M  M_n^m  = sum_s q_s R_n^m(x_s - c)
//   M2M  M_n^m(c') = sum_{k,l} R_{n-k}^{m-l}(c - c') M_k^l(c)
//   M2L  L_j^k  = (-1)^{j+k} sum_{n,m} (-1)^m M_n^m I_{n+j}^{-(m+k)}(c_t - c_s)
//   L2L  L_j^k(c') = sum_{n>=j,m} L_n^m(c) R_{n-j}^{m-k}(c' - c)
//   L2P  phi(y) = sum_{n,m} L_n^m R_n^m(y - c),     M2P  phi(y) = sum_{n,m} M_n^m conj(I_n^m(y - c))
// Only m >= 0 is stored (index n(n+1)/2 + m). Expansions are normalised per tree level by the cell
// width W: M^_n = M_n / W^n, L^_n = L_n W^n, so every operator works on unit-width geometry.
#pragma once

#include <hip/hip_runtime.h>

namespace fg {

constexpr int FMM_PMAX = 12;
constexpr int FMM_NC_MAX = (FMM_PMAX + 1) * (FMM_PMAX + 2) / 2;            // 91
constexpr int FMM_NI_MAX = (2 * FMM_PMAX + 1) * (2 * FMM_PMAX + 2) / 2;    // 325

__host__ __device__ constexpr inline int fidx(int n, int m) { return n * (n + 1) / 2 + m; }
__host__ __device__ constexpr inline int fcount(int p) { return (p + 1) * (p + 2) / 2; }

struct cplx {
    double re, im;
};
__host__ __device__ inline cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__host__ __device__ inline cplx cconj(cplx a) { return {a.re, -a.im}; }

// X_n^m for any sign of m from the m >= 0 storage
__host__ __device__ inline cplx fget(const double2 *X, int n, int m)
{
    if (m >= 0) {
        const double2 v = X[fidx(n, m)];
        return {v.x, v.y};
    }
    const double2 v = X[fidx(n, -m)];
    return (m & 1) ? cplx{-v.x, v.y} : cplx{v.x, -v.y};
}

// Regular harmonics up to order p at x, for m >= 0 (column recurrence over n for each m).
__host__ __device__ inline void regular(double x, double y, double z, int p, double2 *R)
{
    const double r2 = x * x + y * y + z * z;
    double dre = 1.0, dim = 0.0;  // R_m^m
    for (int m = 0; m <= p; m++) {
        if (m > 0) {
            const double s = -1.0 / (2.0 * m);
            const double nre = s * (x * dre - y * dim), nim = s * (x * dim + y * dre);
            dre = nre;
            dim = nim;
        }
        R[fidx(m, m)] = make_double2(dre, dim);
        if (m + 1 > p) break;
        double p2re = dre, p2im = dim;           // R_{n-2}^m
        double p1re = z * dre, p1im = z * dim;   // R_{n-1}^m
        R[fidx(m + 1, m)] = make_double2(p1re, p1im);
        for (int n = m + 2; n <= p; n++) {
            const double inv = 1.0 / ((double)(n + m) * (double)(n - m));
            const double a = (2.0 * n - 1.0) * z;
            const double cre = (a * p1re - r2 * p2re) * inv, cim = (a * p1im - r2 * p2im) * inv;
            R[fidx(n, m)] = make_double2(cre, cim);
            p2re = p1re;
            p2im = p1im;
            p1re = cre;
            p1im = cim;
        }
    }
}

// Irregular harmonics up to order q at x (x != 0), for m >= 0.
__host__ __device__ inline void irregular(double x, double y, double z, int q, double2 *I)
{
    const double r2 = x * x + y * y + z * z;
    const double ir2 = 1.0 / r2;
    double dre = 1.0 / sqrt(r2), dim = 0.0;  // I_m^m
    for (int m = 0; m <= q; m++) {
        if (m > 0) {
            const double s = -(2.0 * m - 1.0) * ir2;
            const double nre = s * (x * dre - y * dim), nim = s * (x * dim + y * dre);
            dre = nre;
            dim = nim;
        }
        I[fidx(m, m)] = make_double2(dre, dim);
        if (m + 1 > q) break;
        double p2re = dre, p2im = dim;
        const double a1 = (2.0 * m + 1.0) * z * ir2;
        double p1re = a1 * dre, p1im = a1 * dim;
        I[fidx(m + 1, m)] = make_double2(p1re, p1im);
        for (int n = m + 2; n <= q; n++) {
            const double a = (2.0 * n - 1.0) * z, b = (double)(n + m - 1) * (double)(n - m - 1);
            const double cre = (a * p1re - b * p2re) * ir2, cim = (a * p1im - b * p2im) * ir2;
            I[fidx(n, m)] = make_double2(cre, cim);
            p2re = p1re;
            p2im = p1im;
            p1re = cre;
            p1im = cim;
        }
    }
}

// P2M accumulation of one charge at x (relative to the cell centre, in cell widths): streaming
// evaluation of the regular harmonics so that only the accumulators stay live.
template <int P>
__host__ __device__ inline void p2m_accumulate(double x, double y, double z, double q, double (&acc)[2 * fcount(P)])
{
    const double r2 = x * x + y * y + z * z;
    double dre = q, dim = 0.0;
#pragma unroll
    for (int m = 0; m <= P; m++) {
        if (m > 0) {
            const double s = -1.0 / (2.0 * m);
            const double nre = s * (x * dre - y * dim), nim = s * (x * dim + y * dre);
            dre = nre;
            dim = nim;
        }
        acc[2 * fidx(m, m)] += dre;
        acc[2 * fidx(m, m) + 1] += dim;
        if (m + 1 <= P) {
            double p2re = dre, p2im = dim;
            double p1re = z * dre, p1im = z * dim;
            acc[2 * fidx(m + 1, m)] += p1re;
            acc[2 * fidx(m + 1, m) + 1] += p1im;
#pragma unroll
            for (int n = m + 2; n <= P; n++) {
                const double inv = 1.0 / ((double)(n + m) * (double)(n - m));
                const double a = (2.0 * n - 1.0) * z;
                const double cre = (a * p1re - r2 * p2re) * inv, cim = (a * p1im - r2 * p2im) * inv;
                acc[2 * fidx(n, m)] += cre;
                acc[2 * fidx(n, m) + 1] += cim;
                p2re = p1re;
                p2im = p1im;
                p1re = cre;
                p1im = cim;
            }
        }
    }
}

// Column split of the P2M accumulators between two lanes: columns m (P + 1 - m coefficients each),
// largest first, go to the half with the smaller total. p2m_half_mask<P>(h) is the bit set of
// half h's columns, p2m_col_off the position of column m's first coefficient among them.
template <int P>
__host__ __device__ constexpr unsigned p2m_half_mask(int h)
{
    unsigned a = 0;
    int na = 0, nb = 0;
    for (int m = 0; m <= P; m++) {
        if (na <= nb) {
            a |= 1u << m;
            na += P + 1 - m;
        } else {
            nb += P + 1 - m;
        }
    }
    return h == 0 ? a : ((1u << (P + 1)) - 1) & ~a;
}
// The same split weighted by accumulators for p2m_accumulate2_cols, whose m = 0 column is real
// (half the accumulators of a complex column of its length).
template <int P>
__host__ __device__ constexpr unsigned p2m_acc_half_mask(int h)
{
    unsigned a = 0;
    int na = 0, nb = 0;  // accumulated reals
    int order[P + 1] = {}, w[P + 1] = {};
    for (int m = 0; m <= P; m++) {
        order[m] = m;
        w[m] = (m == 0 ? 1 : 2) * (P + 1 - m);
    }
    for (int i = 0; i <= P; i++)  // columns by decreasing weight (stable)
        for (int j = i + 1; j <= P; j++)
            if (w[order[j]] > w[order[i]]) {
                const int t = order[i];
                order[i] = order[j];
                order[j] = t;
            }
    for (int i = 0; i <= P; i++) {
        const int m = order[i];
        if (na <= nb) {
            a |= 1u << m;
            na += w[m];
        } else {
            nb += w[m];
        }
    }
    return h == 0 ? a : ((1u << (P + 1)) - 1) & ~a;
}
template <int P>
__host__ __device__ constexpr int p2m_col_off(unsigned mask, int m)
{
    int o = 0;
    for (int k = 0; k < m; k++)
        if ((mask >> k) & 1) o += P + 1 - k;
    return o;
}
template <int P>
__host__ __device__ constexpr int p2m_half_count(unsigned mask)
{
    return p2m_col_off<P>(mask, P + 1);
}

// P2M accumulation of one charge into the columns of MASK only (acc holds their coefficients,
// column m's at p2m_col_off, real and imaginary parts interleaved). It accumulates the unscaled
// harmonics Q_n^m = (n+m)! (n-m)! R_n^m, whose recurrences have integer factors,
//   Q_m^m = -(2m-1)(x+iy) Q_{m-1}^{m-1},  Q_{m+1}^m = (2m+1) z Q_m^m,
//   Q_n^m = (2n-1) z Q_{n-1}^m - ((n-1)^2 - m^2) r^2 Q_{n-2}^m,
// carried as running sums (no per-term constants to hold in registers); p2m_finish applies the
// factorial scale once per expansion. The diagonal chain runs up to the highest column of MASK.
template <int P, unsigned MASK>
__host__ __device__ inline void p2m_accumulate_cols(double x, double y, double z, double q,
                                                    double (&acc)[2 * p2m_half_count<P>(MASK)])
{
    const double r2 = x * x + y * y + z * z, z2 = z + z, r22 = r2 + r2;
    double dre = q, dim = 0.0;
    double tm = 1.0;  // -(2m - 1) once stepped
#pragma unroll
    for (int m = 0; m <= P; m++) {
        if ((MASK >> m) == 0) break;
        if (m > 0) {
            tm -= 2.0;
            const double nre = tm * (x * dre - y * dim), nim = tm * (x * dim + y * dre);
            dre = nre;
            dim = nim;
        }
        if (!((MASK >> m) & 1)) continue;
        const int o = 2 * p2m_col_off<P>(MASK, m);
        acc[o] += dre;
        acc[o + 1] += dim;
        if (m + 1 <= P) {
            // (2m+1) z = -(tm - 2) z for m > 0, z for m = 0
            const double a1 = (m == 0 ? 1.0 : 2.0 - tm) * z;
            double p2re = dre, p2im = dim;
            double p1re = a1 * dre, p1im = a1 * dim;
            acc[o + 2] += p1re;
            acc[o + 3] += p1im;
            double az = a1 + z2;       // (2n-1) z at n = m+2
            double kr2 = (double)(2 * m + 1) * r2, dk = (double)(2 * m + 3) * r2;  // ((n-1)^2 - m^2) r^2 and its step
#pragma unroll
            for (int n = m + 2; n <= P; n++) {
                const double cre = az * p1re - kr2 * p2re, cim = az * p1im - kr2 * p2im;
                acc[o + 2 * (n - m)] += cre;
                acc[o + 2 * (n - m) + 1] += cim;
                p2re = p1re;
                p2im = p1im;
                p1re = cre;
                p1im = cim;
                az += z2;
                kr2 += dk;
                dk += r22;
            }
        }
    }
}

// P2M accumulation of one source into both expansions at once, columns of MASK only: the unit
// harmonics Q_n^m of the source (the recurrences of p2m_accumulate_cols started from Q_0^0 = 1)
// times its charges qu (into au) and qv (into av). T is the arithmetic type (FP64, or FP32 in the
// mixed-precision mode). The factors (2n - 1) z are formed once per source and shared by the
// columns, ((n - 1)^2 - m^2) r^2 is one product by an integer constant, and the m = 0 column is
// real (its imaginary accumulators are left untouched).
template <int P, unsigned MASK, class T>
__host__ __device__ inline void p2m_accumulate2_cols(T x, T y, T z, T qu, T qv, T (&au)[2 * p2m_half_count<P>(MASK)],
                                                     T (&av)[2 * p2m_half_count<P>(MASK)])
{
    const T r2 = x * x + y * y + z * z, z2 = z + z;
    T tz[P + 1];  // tz[n] = (2n - 1) z
    tz[0] = T(0);
    if (P >= 1) tz[1] = z;
#pragma unroll
    for (int n = 2; n <= P; n++) tz[n] = tz[n - 1] + z2;
    auto put = [&](int o, T re, T im) {
        au[o] += re * qu;
        au[o + 1] += im * qu;
        av[o] += re * qv;
        av[o + 1] += im * qv;
    };
    if (MASK & 1u) {
        const int o = 2 * p2m_col_off<P>(MASK, 0);
        au[o] += qu;
        av[o] += qv;
        if (P >= 1) {
            T p2 = T(1), p1 = z;
            au[o + 2] += p1 * qu;
            av[o + 2] += p1 * qv;
#pragma unroll
            for (int n = 2; n <= P; n++) {
                const T c = tz[n] * p1 - (T((n - 1) * (n - 1)) * r2) * p2;
                au[o + 2 * n] += c * qu;
                av[o + 2 * n] += c * qv;
                p2 = p1;
                p1 = c;
            }
        }
    }
    T dre = -x, dim = -y;  // Q_1^1 = -(x + i y)
#pragma unroll
    for (int m = 1; m <= P; m++) {
        if ((MASK >> m) == 0) break;
        if (m > 1) {
            const T f = T(-(2 * m - 1));
            const T nre = f * (x * dre - y * dim), nim = f * (x * dim + y * dre);
            dre = nre;
            dim = nim;
        }
        if (!((MASK >> m) & 1)) continue;
        const int o = 2 * p2m_col_off<P>(MASK, m);
        put(o, dre, dim);
        if (m + 1 <= P) {
            T p2re = dre, p2im = dim;
            T p1re = tz[m + 1] * dre, p1im = tz[m + 1] * dim;
            put(o + 2, p1re, p1im);
#pragma unroll
            for (int n = m + 2; n <= P; n++) {
                const T kr2 = T((n - 1) * (n - 1) - m * m) * r2;
                const T cre = tz[n] * p1re - kr2 * p2re, cim = tz[n] * p1im - kr2 * p2im;
                put(o + 2 * (n - m), cre, cim);
                p2re = p1re;
                p2im = p1im;
                p1re = cre;
                p1im = cim;
            }
        }
    }
}

// Factorial scale of p2m_accumulate_cols' sums: coefficient (n, m) times 1 / ((n+m)! (n-m)!).
__host__ __device__ constexpr double p2m_scale(int n, int m)
{
    double f = 1.0;
    for (int k = 2; k <= n + m; k++) f *= k;
    for (int k = 2; k <= n - m; k++) f *= k;
    return 1.0 / f;
}

// L2P: sum_{n,m} L_n^m R_n^m(x), x in cell widths.
__host__ __device__ inline double l2p_eval(const double2 *L, int p, double x, double y, double z)
{
    const double r2 = x * x + y * y + z * z;
    double dre = 1.0, dim = 0.0, s = 0.0;
    for (int m = 0; m <= p; m++) {
        if (m > 0) {
            const double f = -1.0 / (2.0 * m);
            const double nre = f * (x * dre - y * dim), nim = f * (x * dim + y * dre);
            dre = nre;
            dim = nim;
        }
        const double w = m == 0 ? 1.0 : 2.0;
        double2 Lv = L[fidx(m, m)];
        s += w * (Lv.x * dre - Lv.y * dim);
        if (m + 1 > p) continue;
        double p2re = dre, p2im = dim, p1re = z * dre, p1im = z * dim;
        Lv = L[fidx(m + 1, m)];
        s += w * (Lv.x * p1re - Lv.y * p1im);
        for (int n = m + 2; n <= p; n++) {
            const double inv = 1.0 / ((double)(n + m) * (double)(n - m));
            const double a = (2.0 * n - 1.0) * z;
            const double cre = (a * p1re - r2 * p2re) * inv, cim = (a * p1im - r2 * p2im) * inv;
            Lv = L[fidx(n, m)];
            s += w * (Lv.x * cre - Lv.y * cim);
            p2re = p1re;
            p2im = p1im;
            p1re = cre;
            p1im = cim;
        }
    }
    return s;
}

// M2P for two expansions at once: sum_{n,m} M_n^m conj(I_n^m(x)), x in cell widths, streaming the
// irregular harmonics with the column recurrence. Accumulates into su, sv (caller scales by 1/W).
__host__ __device__ inline void m2p_eval2(const double2 *Mu, const double2 *Mv, int p, double x, double y, double z,
                                          double &su, double &sv)
{
    const double r2 = x * x + y * y + z * z;
    const double ir2 = 1.0 / r2;
    double dre = 1.0 / sqrt(r2), dim = 0.0;
    for (int m = 0; m <= p; m++) {
        if (m > 0) {
            const double f = -(2.0 * m - 1.0) * ir2;
            const double nre = f * (x * dre - y * dim), nim = f * (x * dim + y * dre);
            dre = nre;
            dim = nim;
        }
        const double w = m == 0 ? 1.0 : 2.0;
        double2 a = Mu[fidx(m, m)], b = Mv[fidx(m, m)];
        su += w * (a.x * dre + a.y * dim);
        sv += w * (b.x * dre + b.y * dim);
        if (m + 1 > p) continue;
        const double a1 = (2.0 * m + 1.0) * z * ir2;
        double p2re = dre, p2im = dim, p1re = a1 * dre, p1im = a1 * dim;
        a = Mu[fidx(m + 1, m)];
        b = Mv[fidx(m + 1, m)];
        su += w * (a.x * p1re + a.y * p1im);
        sv += w * (b.x * p1re + b.y * p1im);
        for (int n = m + 2; n <= p; n++) {
            const double aa = (2.0 * n - 1.0) * z, bb = (double)(n + m - 1) * (double)(n - m - 1);
            const double cre = (aa * p1re - bb * p2re) * ir2, cim = (aa * p1im - bb * p2im) * ir2;
            a = Mu[fidx(n, m)];
            b = Mv[fidx(n, m)];
            su += w * (a.x * cre + a.y * cim);
            sv += w * (b.x * cre + b.y * cim);
            p2re = p1re;
            p2im = p1im;
            p1re = cre;
            p1im = cim;
        }
    }
}

// Reciprocal and reciprocal square root for the device evaluators: the hardware approximations
// refined by Newton steps (two in FP64, each doubling the correct bits, to full precision for the
// O(1) arguments here, in place of the IEEE-exact division sequences), used as is in FP32.
__device__ inline double fast_inv(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ inline double fast_rsqrt(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x * y, y, 1.0);  // 1 - x y^2
    y = fma(0.5 * y, e, y);
    e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
}
__device__ inline float fast_inv(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float fast_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }

// L2P of two expansions at once: pu += sum L_u R, pv += sum L_v R. The expansions are T arrays of
// interleaved real and imaginary parts (FP64, or FP32 in the mixed-precision mode); the sums are
// FP64. Not unrolled (it runs once per target); the recurrence divisors are formed from loop
// counters.
template <int P, class T>
__host__ __device__ inline void l2p_eval2_t(const T *__restrict__ Lu, const T *__restrict__ Lv, double x, double y,
                                            double z, double &pu, double &pv)
{
    const double r2 = x * x + y * y + z * z;
    double dre = 1.0, dim = 0.0;
    auto at = [](const T *L, int k) { return make_double2((double)L[2 * k], (double)L[2 * k + 1]); };
#pragma unroll 1
    for (int m = 0; m <= P; m++) {
        if (m > 0) {
            const double f = -0.5 / (double)m;
            const double nre = f * (x * dre - y * dim), nim = f * (x * dim + y * dre);
            dre = nre;
            dim = nim;
        }
        double2 a = at(Lu, fidx(m, m)), b = at(Lv, fidx(m, m));
        double au = a.x * dre - a.y * dim, av = b.x * dre - b.y * dim;
        if (m + 1 <= P) {
            double p2re = dre, p2im = dim, p1re = z * dre, p1im = z * dim;
            a = at(Lu, fidx(m + 1, m));
            b = at(Lv, fidx(m + 1, m));
            au += a.x * p1re - a.y * p1im;
            av += b.x * p1re - b.y * p1im;
#pragma unroll 1
            for (int n = m + 2; n <= P; n++) {
                const double inv = 1.0 / ((double)(n + m) * (double)(n - m));
                const double aa = (2.0 * n - 1.0) * z;
                const double cre = (aa * p1re - r2 * p2re) * inv, cim = (aa * p1im - r2 * p2im) * inv;
                a = at(Lu, fidx(n, m));
                b = at(Lv, fidx(n, m));
                au += a.x * cre - a.y * cim;
                av += b.x * cre - b.y * cim;
                p2re = p1re;
                p2im = p1im;
                p1re = cre;
                p1im = cim;
            }
        }
        const double w = m == 0 ? 1.0 : 2.0;
        pu += w * au;
        pv += w * av;
    }
}

// M2M target coefficient (n, m >= 0): sum over the child's (k, l) of R_{n-k}^{m-l}(d) M_k^l, with d
// the child-to-parent centre offset in parent widths and the child expansion normalised by the
// child width (factor 2^-k).
__host__ __device__ inline cplx m2m_coef(const double2 *Mc, const double2 *Rd, int n, int m)
{
    cplx s = {0, 0};
    double sc = 1.0;
    for (int k = 0; k <= n; k++, sc *= 0.5) {
        for (int l = -k; l <= k; l++) {
            const int ml = m - l;
            if (ml > n - k || -ml > n - k) continue;
            const cplx t = cmul(fget(Rd, n - k, ml), fget(Mc, k, l));
            s.re += sc * t.re;
            s.im += sc * t.im;
        }
    }
    return s;
}

// M2L target coefficient (j, k >= 0) from a source expansion of order p with the irregular table
// of the offset (unit widths); the caller applies the 1/W factor.
__host__ __device__ inline cplx m2l_coef(const double2 *M, const double2 *It, int p, int j, int k)
{
    cplx s = {0, 0};
    for (int n = 0; n <= p; n++) {
        for (int m = -n; m <= n; m++) {
            const cplx t = cmul(fget(M, n, m), fget(It, n + j, -(m + k)));
            if (m & 1) {
                s.re -= t.re;
                s.im -= t.im;
            } else {
                s.re += t.re;
                s.im += t.im;
            }
        }
    }
    if ((j + k) & 1) {
        s.re = -s.re;
        s.im = -s.im;
    }
    return s;
}

// L2L target coefficient (j, k >= 0): parent expansion (unit parent width) shifted by d (parent
// widths), result normalised by the child width (factor 2^-j).
__host__ __device__ inline cplx l2l_coef(const double2 *Lp, const double2 *Rd, int p, int j, int k)
{
    cplx s = {0, 0};
    for (int n = j; n <= p; n++) {
        for (int m = -n; m <= n; m++) {
            const int mk = m - k;
            if (mk > n - j || -mk > n - j) continue;
            const cplx t = cmul(fget(Lp, n, m), fget(Rd, n - j, mk));
            s.re += t.re;
            s.im += t.im;
        }
    }
    double sc = 1.0;
    for (int q = 0; q < j; q++) sc *= 0.5;
    s.re *= sc;
    s.im *= sc;
    return s;
}

}  // namespace fg
