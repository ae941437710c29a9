// guess.hip -- the learned initial guess of the BiCGSTAB solve.
//
// FeeLLGood starts each solve from the previous v projected on the new tangent plane
// (LinAlgebra::buildInitGuess, src/linear_algebra.cpp:13-24). This network proposes a correction
// to that guess; the Krylov solve to the same tolerance follows unchanged, so a poor proposal
// costs iterations, never accuracy.
//
// Network, per magnetic node i (weights stored in BF16, arithmetic in FP32):
//   features f_i = [u_i, s_h h_i, s_v v_i, s_d d_i]  with h_i = F_t,i / m_i the tangent part of the
//                  assembled right-hand side over the lumped mass, v_i the previous v and d_i the
//                  second difference v_i - 2 v_i,prev + v_i,prev2 of the last three solutions
//                  (zero unless they came at one fixed step) (12 values)
//   aggregation a_i = mean of f_j over the magnetic mesh-edge neighbours j of i
//   x = [f_i, a_i] (24) -> z1 = W1 x + b1 -> h1 = GELU(z1) (64) -> z2 = W2 h1 + b2 -> h2 = GELU(z2)
//   (64) -> o = W3 h2 + b3 (3) -> v_pred = P_u (v_base,i + s_out o), P_u the tangent projection
//   guess x0_i = (v_pred . ep_i, v_pred . eq_i) / gamma0
// The base the network corrects is the linear extrapolation of the last two solutions,
// v_base = v_i + (dt / dt_prev) (v_i - v_i,prev) (v_i itself until two solves follow a state
// reset): at a fixed step it cuts the initial error by about the ratio of the step to the time
// scale of the dynamics.
// GELU is the tanh form 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))).
// Training: loss = sum_i |v_pred,i - v*_i|^2 / (n_mag s_out^2) against the converged solution v*
// of the same solve; Adam on FP32 master weights; the forward pass uses their BF16 rounding.
// The scales are the root mean squares of h, v*, v* - P_u v_base and d over the first training
// batch, so the base has loss 1 on it; W3 and b3 start at zero (the network starts as the base).
// The second difference is what the linear extrapolation leaves out at a fixed step (the
// quadratic extrapolation is v_base + d), the term a trained network can add to the base.
// The reference guess (P_u v_i) is scored on the same batch as loss_ref.
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "ctx.h"

namespace fg {

namespace {

constexpr int NF = 12, NX = 24, NH = 64, NO = 3;
constexpr int FS = 16;  // feature row: the NF features, then the base v the network corrects, 1 pad
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NX, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NO * NH, NTRAIN = OFF_B3 + NO;
constexpr int OFF_S = NTRAIN, NSC = 4, NPARAM = NTRAIN + NSC;  // scales s_h, s_v, s_out, s_d
static_assert(NPARAM == FG_GUESS_NPARAM, "parameter layout and header disagree");
constexpr int GB = 128;  // threads per block of the per-node kernels
constexpr int NXP = 24;  // inputs padded to whole MFMA k-steps of 4
constexpr int KS1 = NXP / 4;  // k-steps of the first layer

// GELU in its tanh form, written as 0.5 (1 + tanh(a)) = 1 / (1 + exp(-2a)) with
// a = sqrt(2/pi) (z + 0.044715 z^3), evaluated by the hardware exp2 and reciprocal (about one ulp
// each): exp(-2a) = 2^(c (z + 0.044715 z^3)), c = -2 sqrt(2/pi) log2(e).
__device__ inline float gelu_sigma(float z)
{
    const float c = -2.3022082f;
    const float u = fmaf(0.044715f * z * z, z, z);
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(c * u));
}
__device__ inline float gelu(float z) { return z * gelu_sigma(z); }
// d/dz z s(z) = s + z s (1 - s) 2 sqrt(2/pi) (1 + 3 0.044715 z^2)
__device__ inline float gelu_grad(float z)
{
    const float s = gelu_sigma(z);
    return s + z * s * (1.0f - s) * 1.5957691f * (1.0f + 0.134145f * z * z);
}

// round-to-nearest-even BF16 value of a float, returned as a float
__host__ __device__ inline float bf16_round(float x)
{
    uint32_t b;
    __builtin_memcpy(&b, &x, 4);
    b = (b + 0x7FFFu + ((b >> 16) & 1u)) & 0xFFFF0000u;
    float y;
    __builtin_memcpy(&y, &b, 4);
    return y;
}

}  // namespace

struct GuessNet {
    int N = 0, nmag = 0, nblk = 0;
    // scaled: the scales of h, v* and the output set; scaled_d: that of the second difference,
    // which the first training batches (fewer than three earlier solutions) cannot give
    bool enabled = false, have_batch = false, scaled = false, scaled_d = false;
    int adam_t = 0;
    float *params = nullptr;                 // FP32 master + scales [NPARAM]
    uint16_t *wbf = nullptr;                 // BF16 copy of the trainable block [NTRAIN]
    float *adam_m = nullptr, *adam_v = nullptr, *grad = nullptr;
    float *feat = nullptr;                   // [N][FS] features of the last solve (u, h, v) and base
    // training activations and deltas, [N][k]
    float *xin = nullptr, *z1 = nullptr, *z2 = nullptr, *dz1 = nullptr, *dz2 = nullptr, *dout = nullptr;
    float *gpart = nullptr;                  // per-block partial gradients
    double *lpart = nullptr;                 // per-block loss partials (2 per block)
    double *h_l = nullptr;                   // pinned
    float *xq = nullptr;                     // [N][NXP] network inputs of the inference pass
    float *vhist = nullptr;                  // [N][6] previous v of the last two recorded solves
    float *ninv = nullptr;                   // [N] 1 / number of magnetic mesh neighbours
    int nhist = 0;                           // assemblies recorded since the last state reset
    double dt_hist = 0, dt_hist2 = 0;        // steps of the last two recorded solves
};

namespace {

// BF16 rounding of the trainable block (biases are used in FP32 from the master copy)
__global__ void k_guess_round(const float *__restrict__ p, uint16_t *__restrict__ wbf)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < NTRAIN; i += gridDim.x * blockDim.x) {
        const float r = bf16_round(p[i]);
        uint32_t b;
        __builtin_memcpy(&b, &r, 4);
        wbf[i] = (uint16_t)(b >> 16);
    }
}

__device__ inline float bf16_to_float(uint16_t h)
{
    const uint32_t b = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &b, 4);
    return f;
}

// raw features: u, h = (b_y ep + b_x eq) / m (b = (eq.F, ep.F)), previous v
__global__ __launch_bounds__(256) void k_guess_features(int N, const NodeRec *__restrict__ cur,
                                                        const NodeRec *__restrict__ vsrc,
                                                        const double2 *__restrict__ b,
                                                        const double4 *__restrict__ ep4,
                                                        const double4 *__restrict__ eq4,
                                                        const double *__restrict__ mass,
                                                        const int32_t *__restrict__ mag, float *__restrict__ feat,
                                                        float *__restrict__ vhist, float rho, int have_hist,
                                                        int have_d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float f[NF] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (mag[i]) {
        const NodeRec n = cur[i], v = vsrc[i];
        const double2 bi = b[i];
        const double4 p = ep4[i], q = eq4[i];
        const double im = 1.0 / mass[i];
        f[0] = (float)n.ux;
        f[1] = (float)n.uy;
        f[2] = (float)n.uz;
        f[3] = (float)((bi.y * p.x + bi.x * q.x) * im);
        f[4] = (float)((bi.y * p.y + bi.x * q.y) * im);
        f[5] = (float)((bi.y * p.z + bi.x * q.z) * im);
        f[6] = (float)v.vx;
        f[7] = (float)v.vy;
        f[8] = (float)v.vz;
    }
    // extrapolated base from this solve's previous v and the one before it, the second difference
    // with the one before that; shift the history
    float vb[3];
    for (int k = 0; k < 3; k++) {
        const float v1 = f[6 + k], v2 = have_hist ? vhist[(size_t)i * 6 + k] : v1;
        const float v3 = have_d ? vhist[(size_t)i * 6 + 3 + k] : v2;
        vb[k] = v1 + rho * (v1 - v2);
        f[9 + k] = have_d ? (v1 - v2) - (v2 - v3) : 0.f;
        vhist[(size_t)i * 6 + 3 + k] = v2;
        vhist[(size_t)i * 6 + k] = v1;
    }
    float4 *dst = reinterpret_cast<float4 *>(feat + (size_t)i * FS);
    dst[0] = make_float4(f[0], f[1], f[2], f[3]);
    dst[1] = make_float4(f[4], f[5], f[6], f[7]);
    dst[2] = make_float4(f[8], f[9], f[10], f[11]);
    dst[3] = make_float4(vb[0], vb[1], vb[2], 0.f);
}

struct SWeights {
    float W1[NH * NX], b1[NH], W2[NH * NH], b2[NH], W3[NO * NH], b3[NO], s[NSC];
};

__device__ inline void load_weights(SWeights &S, const float *__restrict__ p, const uint16_t *__restrict__ wbf)
{
    for (int k = threadIdx.x; k < NH * NX; k += blockDim.x) S.W1[k] = bf16_to_float(wbf[OFF_W1 + k]);
    for (int k = threadIdx.x; k < NH * NH; k += blockDim.x) S.W2[k] = bf16_to_float(wbf[OFF_W2 + k]);
    for (int k = threadIdx.x; k < NO * NH; k += blockDim.x) S.W3[k] = bf16_to_float(wbf[OFF_W3 + k]);
    for (int k = threadIdx.x; k < NH; k += blockDim.x) {
        S.b1[k] = p[OFF_B1 + k];
        S.b2[k] = p[OFF_B2 + k];
    }
    if (threadIdx.x < NO) S.b3[threadIdx.x] = p[OFF_B3 + threadIdx.x];
    if (threadIdx.x < NSC) S.s[threadIdx.x] = p[OFF_S + threadIdx.x];
    __syncthreads();
}

// network input x = [f_i, mean f_j] with the feature scales applied
// Non-magnetic nodes have zero features, so the sum runs over all stencil neighbours; ninv holds
// 1 / (number of magnetic neighbours), fixed by the mesh (k_guess_ncount).
__device__ inline void gather_input(int i, const float *__restrict__ feat, const float *__restrict__ ninv,
                                    const int64_t *__restrict__ sptr, const int32_t *__restrict__ swid,
                                    const int32_t *__restrict__ scol, const float *s, float x[NX])
{
    const float sc[NF] = {1, 1, 1, s[0], s[0], s[0], s[1], s[1], s[1], s[3], s[3], s[3]};
    for (int k = 0; k < NF; k++) x[k] = feat[(size_t)i * FS + k] * sc[k];
    float a[NF] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int sl = i >> 6, lane = i & 63;
    const int64_t off = sptr[sl];
    const int width = swid[sl];
#pragma unroll 4
    for (int k = 0; k < width; k++) {
        const int j = scol[off + (int64_t)k * SELL_C + lane];
        if (j == i) continue;
        const float4 *fj = reinterpret_cast<const float4 *>(feat + (size_t)j * FS);
        const float4 f0 = fj[0], f1 = fj[1], f2 = fj[2];
        a[0] += f0.x;
        a[1] += f0.y;
        a[2] += f0.z;
        a[3] += f0.w;
        a[4] += f1.x;
        a[5] += f1.y;
        a[6] += f1.z;
        a[7] += f1.w;
        a[8] += f2.x;
        a[9] += f2.y;
        a[10] += f2.z;
        a[11] += f2.w;
    }
    const float ic = ninv[i];
    for (int k = 0; k < NF; k++) x[NF + k] = a[k] * ic * sc[k];
}

// 1 / (number of magnetic stencil neighbours) of every magnetic node (0 without any)
__global__ __launch_bounds__(256) void k_guess_ncount(int N, const int32_t *__restrict__ mag,
                                                      const int64_t *__restrict__ sptr, const int32_t *__restrict__ swid,
                                                      const int32_t *__restrict__ scol, float *__restrict__ ninv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int sl = i >> 6, lane = i & 63;
    const int64_t off = sptr[sl];
    int cnt = 0;
    for (int k = 0; k < swid[sl]; k++) {
        const int j = scol[off + (int64_t)k * SELL_C + lane];
        cnt += j != i && mag[j];
    }
    ninv[i] = cnt > 0 ? 1.0f / (float)cnt : 0.0f;
}

// Network inputs of the inference pass, padded to NXP (zero rows for non-magnetic nodes).
__global__ __launch_bounds__(256) void k_guess_input(int N, const float *__restrict__ feat,
                                                     const int32_t *__restrict__ mag, const float *__restrict__ ninv,
                                                     const int64_t *__restrict__ sptr, const int32_t *__restrict__ swid,
                                                     const int32_t *__restrict__ scol, const float *__restrict__ p,
                                                     float *__restrict__ xq)
{
    const int i = lblock() * blockDim.x + threadIdx.x;  // XCD order: neighbouring rows share an L2
    if (i >= N) return;
    float x[NX] = {};
    if (mag[i]) {
        const float s[NSC] = {p[OFF_S], p[OFF_S + 1], p[OFF_S + 2], p[OFF_S + 3]};
        gather_input(i, feat, ninv, sptr, swid, scol, s, x);
    }
    static_assert(NX == NXP, "the inputs fill whole k-steps");
    float4 *dst = reinterpret_cast<float4 *>(xq + (size_t)i * NXP);
#pragma unroll
    for (int k = 0; k < NXP / 4; k++) dst[k] = make_float4(x[4 * k], x[4 * k + 1], x[4 * k + 2], x[4 * k + 3]);
}

// Inference on the matrix cores. Each wavefront takes 64 nodes as four 16-node column tiles and
// computes the layers transposed, Z^T = W X^T, with v_mfma_f32_16x16x16_bf16 (A: 16x16 weights,
// lane l holding row l%16, k 4(l/16)..4(l/16)+3; B: 16x16 activations, lane l holding those k of
// column l%16; D: lane l holding rows 4(l/16)+j, j < 4, column l%16, FP32). A hidden layer's 16-row
// tile R of D is therefore, lane for lane, the B operand of the next layer's k-block R: each layer
// feeds the next without moving data. The weights are BF16 already; each FP32 activation x is
// split into BF16 parts hi + lo (x - hi rounded again), and both parts are multiplied, so the
// products carry x to about 16 bits (a relative 2^-17) while running at the BF16 matrix rate.
// The weight fragments are staged in LDS in lane order (one conflict-free read per fragment).
// The three outputs (rows 0..2, lanes 0..15) go through LDS to one node per lane for v_pred.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short bf16x4 __attribute__((ext_vector_type(4)));
constexpr int KB1 = (NX + 15) / 16;  // 16-input k-blocks of the first layer
struct MfmaNet {
    bf16x4 W2b[16 * 64];      // [(R2 * 4 + R) * 64 + l]: W2[16 R2 + l%16][16 R + 4(l/16) + j]
    bf16x4 W1b[4 * KB1 * 64];  // [(R * KB1 + q) * 64 + l]: W1[16 R + l%16][16 q + 4(l/16) + j]
    bf16x4 W3b[4 * 64];       // [R * 64 + l]: W3[l%16][16 R + 4(l/16) + j], rows past NO zero
    float b1[NH], b2[NH], b3[4], s[NSC];
    float o[256 * NO];
};
__device__ inline short bf16_bits(float x)
{
    uint32_t b;
    __builtin_memcpy(&b, &x, 4);
    return (short)(b >> 16);
}
// hi and lo BF16 parts of four activations
__device__ inline void split_bf16(const f32x4 &x, bf16x4 &hi, bf16x4 &lo)
{
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float h = bf16_round(x[j]);
        hi[j] = bf16_bits(h);
        lo[j] = bf16_bits(bf16_round(x[j] - h));
    }
}
__device__ inline f32x4 mma2(const bf16x4 &a, const bf16x4 &bh, const bf16x4 &bl, f32x4 acc)
{
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bh, acc, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bl, acc, 0, 0, 0);
}

__global__ __launch_bounds__(256) void k_guess_infer(int N, const float *__restrict__ xq,
                                                     const float *__restrict__ feat, const int32_t *__restrict__ mag,
                                                     const float *__restrict__ p, const uint16_t *__restrict__ wbf,
                                                     const double4 *__restrict__ ep4, const double4 *__restrict__ eq4,
                                                     double2 *__restrict__ xg, double4 *__restrict__ wg)
{
    __shared__ MfmaNet S;
    const int tid = threadIdx.x, l = tid & 63, wv = tid >> 6, g4 = l >> 4, ln = l & 15;
    for (int e = tid; e < 16 * 64; e += 256) {
        const int f = e >> 6, q = e & 63, R2 = f >> 2, R = f & 3;
        const int row = 16 * R2 + (q & 15), col = 16 * R + 4 * (q >> 4);
        const uint16_t *w = wbf + OFF_W2 + row * NH + col;
        S.W2b[e] = bf16x4{(short)w[0], (short)w[1], (short)w[2], (short)w[3]};
    }
    for (int e = tid; e < 4 * KB1 * 64; e += 256) {
        const int f = e >> 6, q = e & 63, R = f / KB1, kb = f - KB1 * R, k0 = 16 * kb + 4 * (q >> 4);
        bf16x4 v = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (k0 + j < NX) v[j] = (short)wbf[OFF_W1 + (16 * R + (q & 15)) * NX + k0 + j];
        S.W1b[e] = v;
    }
    for (int e = tid; e < 4 * 64; e += 256) {
        const int R = e >> 6, q = e & 63, o = q & 15, col = 16 * R + 4 * (q >> 4);
        bf16x4 v = {0, 0, 0, 0};
        if (o < NO)
            for (int j = 0; j < 4; j++) v[j] = (short)wbf[OFF_W3 + o * NH + col + j];
        S.W3b[e] = v;
    }
    if (tid < NH) {
        S.b1[tid] = p[OFF_B1 + tid];
        S.b2[tid] = p[OFF_B2 + tid];
    }
    if (tid < 4) S.b3[tid] = tid < NO ? p[OFF_B3 + tid] : 0.f;
    if (tid < NSC) S.s[tid] = p[OFF_S + tid];
    __syncthreads();
    for (int base = blockIdx.x * 256; base < N; base += gridDim.x * 256) {
        const int wbase = base + 64 * wv;
        // B operands of layer 1: inputs 16 q + 4 g4 + j of node wbase + 16 T + ln, split
        bf16x4 xh[4][KB1], xl[4][KB1];
#pragma unroll
        for (int T = 0; T < 4; T++) {
            const int node = wbase + 16 * T + ln;
#pragma unroll
            for (int q = 0; q < KB1; q++) {
                const int k0 = 16 * q + 4 * g4;
                f32x4 v = {0.f, 0.f, 0.f, 0.f};
                static_assert(NXP % 4 == 0 && NX % 4 == 0, "whole float4 groups of inputs");
                if (node < N && k0 < NX) {
                    const float4 t = *reinterpret_cast<const float4 *>(xq + (size_t)node * NXP + k0);
                    v = f32x4{t.x, t.y, t.z, t.w};
                }
                split_bf16(v, xh[T][q], xl[T][q]);
            }
        }
        f32x4 h1[4][4];
#pragma unroll
        for (int R = 0; R < 4; R++) {
            f32x4 a0;
#pragma unroll
            for (int j = 0; j < 4; j++) a0[j] = S.b1[16 * R + 4 * g4 + j];
#pragma unroll
            for (int T = 0; T < 4; T++) h1[R][T] = a0;
#pragma unroll
            for (int q = 0; q < KB1; q++) {
                const bf16x4 a = S.W1b[(R * KB1 + q) * 64 + l];
#pragma unroll
                for (int T = 0; T < 4; T++) h1[R][T] = mma2(a, xh[T][q], xl[T][q], h1[R][T]);
            }
#pragma unroll
            for (int T = 0; T < 4; T++)
#pragma unroll
                for (int j = 0; j < 4; j++) h1[R][T][j] = gelu(h1[R][T][j]);
        }
        bf16x4 hh[4][4], hl[4][4];
#pragma unroll
        for (int R = 0; R < 4; R++)
#pragma unroll
            for (int T = 0; T < 4; T++) split_bf16(h1[R][T], hh[R][T], hl[R][T]);
        f32x4 h2[4][4];
#pragma unroll
        for (int R2 = 0; R2 < 4; R2++) {
            f32x4 a0;
#pragma unroll
            for (int j = 0; j < 4; j++) a0[j] = S.b2[16 * R2 + 4 * g4 + j];
#pragma unroll
            for (int T = 0; T < 4; T++) h2[R2][T] = a0;
#pragma unroll
            for (int R = 0; R < 4; R++) {
                const bf16x4 a = S.W2b[(R2 * 4 + R) * 64 + l];
#pragma unroll
                for (int T = 0; T < 4; T++) h2[R2][T] = mma2(a, hh[R][T], hl[R][T], h2[R2][T]);
            }
#pragma unroll
            for (int T = 0; T < 4; T++)
#pragma unroll
                for (int j = 0; j < 4; j++) h2[R2][T][j] = gelu(h2[R2][T][j]);
        }
#pragma unroll
        for (int R = 0; R < 4; R++)
#pragma unroll
            for (int T = 0; T < 4; T++) split_bf16(h2[R][T], hh[R][T], hl[R][T]);
        f32x4 o3[4];
#pragma unroll
        for (int T = 0; T < 4; T++)
#pragma unroll
            for (int j = 0; j < 4; j++) o3[T][j] = 4 * g4 + j < NO ? S.b3[4 * g4 + j] : 0.f;
#pragma unroll
        for (int R = 0; R < 4; R++) {
            const bf16x4 a = S.W3b[R * 64 + l];
#pragma unroll
            for (int T = 0; T < 4; T++) o3[T] = mma2(a, hh[R][T], hl[R][T], o3[T]);
        }
        if (g4 == 0)
#pragma unroll
            for (int T = 0; T < 4; T++)
#pragma unroll
                for (int j = 0; j < NO; j++) S.o[(64 * wv + 16 * T + ln) * NO + j] = o3[T][j];
        __syncthreads();
        const int i = base + tid;
        if (i < N && mag[i]) {
            // v_pred = P_u (v_base + s_out o), then the guess in the tangent basis
            const double u[3] = {feat[(size_t)i * FS], feat[(size_t)i * FS + 1], feat[(size_t)i * FS + 2]};
            const double so = S.s[2];
            double v[3];
            for (int k = 0; k < 3; k++) v[k] = (double)feat[(size_t)i * FS + NF + k] + so * (double)S.o[tid * NO + k];
            const double un = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
            for (int k = 0; k < 3; k++) v[k] -= un * u[k];
            const double4 ep = ep4[i], eq = eq4[i];
            const double2 xx = make_double2((v[0] * ep.x + v[1] * ep.y + v[2] * ep.z) / GAMMA0,
                                            (v[0] * eq.x + v[1] * eq.y + v[2] * eq.z) / GAMMA0);
            xg[i] = xx;
            wg[i] = make_double4(xx.x * ep.x + xx.y * eq.x, xx.x * ep.y + xx.y * eq.y, xx.x * ep.z + xx.y * eq.z, 0.0);
        }
        __syncthreads();
    }
}

// Forward pass per magnetic node. mode 0: write the guess x, w. mode 1 (training): store the
// input and pre-activations, and the output delta and loss partials against the solution v*
// (NEXT v) and the reference guess.
template <int MODE>
__global__ __launch_bounds__(GB) void k_guess_forward(
    int N, double inv_norm, const float *__restrict__ feat, const int32_t *__restrict__ mag,
    const float *__restrict__ ninv, const int64_t *__restrict__ sptr, const int32_t *__restrict__ swid, const int32_t *__restrict__ scol,
    const float *__restrict__ p, const uint16_t *__restrict__ wbf, const double4 *__restrict__ ep4,
    const double4 *__restrict__ eq4, const NodeRec *__restrict__ nxt, double2 *__restrict__ xg,
    double4 *__restrict__ wg, float *__restrict__ xin, float *__restrict__ z1o, float *__restrict__ z2o,
    float *__restrict__ dout, double *__restrict__ lpart)
{
    __shared__ SWeights S;
    __shared__ double red[2][GB];
    load_weights(S, p, wbf);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double l_net = 0, l_ref = 0;
    if (i < N && mag[i]) {
        float x[NX];
        gather_input(i, feat, ninv, sptr, swid, scol, S.s, x);
        float h1[NH];
#pragma unroll
        for (int o = 0; o < NH; o++) {
            float z = S.b1[o];
#pragma unroll
            for (int k = 0; k < NX; k++) z += S.W1[o * NX + k] * x[k];
            if (MODE == 1) z1o[(size_t)i * NH + o] = z;
            h1[o] = gelu(z);
        }
        float out[NO] = {S.b3[0], S.b3[1], S.b3[2]};
        for (int o = 0; o < NH; o++) {
            float z = S.b2[o];
#pragma unroll
            for (int k = 0; k < NH; k++) z += S.W2[o * NH + k] * h1[k];
            if (MODE == 1) z2o[(size_t)i * NH + o] = z;
            const float h = gelu(z);
            out[0] += S.W3[o] * h;
            out[1] += S.W3[NH + o] * h;
            out[2] += S.W3[2 * NH + o] * h;
        }
        // v_pred = P_u (v_base + s_out o); the reference guess P_u v
        const double u[3] = {feat[(size_t)i * FS], feat[(size_t)i * FS + 1], feat[(size_t)i * FS + 2]};
        const double vp[3] = {feat[(size_t)i * FS + 6], feat[(size_t)i * FS + 7], feat[(size_t)i * FS + 8]};
        const double so = S.s[2];
        double v[3], vr[3];
        for (int k = 0; k < 3; k++) {
            v[k] = (double)feat[(size_t)i * FS + NF + k] + so * (double)out[k];
            vr[k] = vp[k];
        }
        const double un = u[0] * v[0] + u[1] * v[1] + u[2] * v[2], ur = u[0] * vr[0] + u[1] * vr[1] + u[2] * vr[2];
        for (int k = 0; k < 3; k++) {
            v[k] -= un * u[k];
            vr[k] -= ur * u[k];
        }
        if (MODE == 0) {
            const double4 ep = ep4[i], eq = eq4[i];
            const double2 xx = make_double2((v[0] * ep.x + v[1] * ep.y + v[2] * ep.z) / GAMMA0,
                                            (v[0] * eq.x + v[1] * eq.y + v[2] * eq.z) / GAMMA0);
            xg[i] = xx;
            wg[i] = make_double4(xx.x * ep.x + xx.y * eq.x, xx.x * ep.y + xx.y * eq.y, xx.x * ep.z + xx.y * eq.z, 0.0);
        } else {
            for (int k = 0; k < NX; k++) xin[(size_t)i * NX + k] = x[k];
            const NodeRec n = nxt[i];
            const double vs[3] = {n.vx, n.vy, n.vz};
            double e[3];
            for (int k = 0; k < 3; k++) {
                e[k] = v[k] - vs[k];
                l_net += e[k] * e[k];
                l_ref += (vr[k] - vs[k]) * (vr[k] - vs[k]);
            }
            // dL/do = s_out P_u dL/dv_pred, dL/dv_pred = 2 e inv_norm
            const double eu = u[0] * e[0] + u[1] * e[1] + u[2] * e[2];
            for (int k = 0; k < 3; k++) dout[(size_t)i * NO + k] = (float)(2.0 * inv_norm * so * (e[k] - eu * u[k]));
        }
    } else if (MODE == 1 && i < N) {
        for (int k = 0; k < NX; k++) xin[(size_t)i * NX + k] = 0.f;
        for (int k = 0; k < NH; k++) {
            z1o[(size_t)i * NH + k] = 0.f;
            z2o[(size_t)i * NH + k] = 0.f;
        }
        for (int k = 0; k < NO; k++) dout[(size_t)i * NO + k] = 0.f;
    }
    if (MODE == 1) {
        red[0][threadIdx.x] = l_net;
        red[1][threadIdx.x] = l_ref;
        __syncthreads();
        for (int s = GB / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) {
                red[0][threadIdx.x] += red[0][threadIdx.x + s];
                red[1][threadIdx.x] += red[1][threadIdx.x + s];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            lpart[2 * blockIdx.x] = red[0][0] * inv_norm;
            lpart[2 * blockIdx.x + 1] = red[1][0] * inv_norm;
        }
    }
}

// Backward through the layers per node: dz2 = (W3^T dout) * GELU'(z2), dz1 = (W2^T dz2) * GELU'(z1)
__global__ __launch_bounds__(GB) void k_guess_backward(int N, const float *__restrict__ p,
                                                      const uint16_t *__restrict__ wbf,
                                                      const float *__restrict__ z1, const float *__restrict__ z2,
                                                      const float *__restrict__ dout, float *__restrict__ dz1,
                                                      float *__restrict__ dz2)
{
    __shared__ SWeights S;
    load_weights(S, p, wbf);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float d0 = dout[(size_t)i * NO], d1 = dout[(size_t)i * NO + 1], d2 = dout[(size_t)i * NO + 2];
    float g2[NH];
#pragma unroll
    for (int o = 0; o < NH; o++) {
        const float dh = S.W3[o] * d0 + S.W3[NH + o] * d1 + S.W3[2 * NH + o] * d2;
        g2[o] = dh * gelu_grad(z2[(size_t)i * NH + o]);
        dz2[(size_t)i * NH + o] = g2[o];
    }
    for (int k = 0; k < NH; k++) {
        float dh = 0;
#pragma unroll
        for (int o = 0; o < NH; o++) dh += S.W2[o * NH + k] * g2[o];
        dz1[(size_t)i * NH + k] = dh * gelu_grad(z1[(size_t)i * NH + k]);
    }
}

// Partial weight gradients over a chunk of nodes: G[r][c] = sum_n A[n][r] B[n][c] with an implicit
// last column B[n][rb] = 1 (the bias gradient). B is an activation array, or the GELU of a
// pre-activation array when act is set. One partial of ra*(rb+1) values per block.
constexpr int WG_CHUNK = 2048, WG_TILE = 32;
__global__ __launch_bounds__(256) void k_guess_wgrad(int N, const float *__restrict__ A, int ra,
                                                     const float *__restrict__ B, int rb, int act,
                                                     float *__restrict__ part)
{
    __shared__ float sA[WG_TILE][NH];
    __shared__ float sB[WG_TILE][NH + 1];
    const int n0 = blockIdx.x * WG_CHUNK, n1 = min(N, n0 + WG_CHUNK);
    const int ne = ra * (rb + 1);
    float acc[17];
    for (int q = 0; q < 17; q++) acc[q] = 0.f;
    for (int t0 = n0; t0 < n1; t0 += WG_TILE) {
        const int nt = min(WG_TILE, n1 - t0);
        for (int k = threadIdx.x; k < WG_TILE * NH; k += blockDim.x) {
            const int r = k / NH, c = k % NH;
            float a = 0.f, bv = 0.f;
            if (r < nt && c < ra) a = A[(size_t)(t0 + r) * ra + c];
            if (r < nt && c < rb) {
                bv = B[(size_t)(t0 + r) * rb + c];
                if (act) bv = gelu(bv);
            }
            sA[r][c] = a;
            sB[r][c] = bv;
        }
        __syncthreads();
        for (int q = 0; q < 17; q++) {
            const int e = threadIdx.x + q * 256;
            if (e >= ne) break;
            const int r = e / (rb + 1), c = e % (rb + 1);
            float s = 0.f;
            if (c < rb)
                for (int n = 0; n < nt; n++) s += sA[n][r] * sB[n][c];
            else
                for (int n = 0; n < nt; n++) s += sA[n][r];
            acc[q] += s;
        }
        __syncthreads();
    }
    for (int q = 0; q < 17; q++) {
        const int e = threadIdx.x + q * 256;
        if (e < ne) part[(size_t)blockIdx.x * ne + e] = acc[q];
    }
}

// grad[dst(e)] = sum over blocks of part[blk][e], scattered into the parameter layout:
// weights at wofs + r * rb + c, bias at bofs + r
__global__ void k_guess_reduce(int nblk, int ra, int rb, const float *__restrict__ part, int wofs, int bofs,
                               float *__restrict__ grad)
{
    const int ne = ra * (rb + 1);
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
        double s = 0;
        for (int b = 0; b < nblk; b++) s += part[(size_t)b * ne + e];
        const int r = e / (rb + 1), c = e % (rb + 1);
        grad[c < rb ? wofs + r * rb + c : bofs + r] = (float)s;
    }
}

__global__ void k_guess_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                             float *__restrict__ v, float lr, float c1, float c2)
{
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < NTRAIN; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        m[i] = b1 * m[i] + (1 - b1) * gi;
        v[i] = b2 * v[i] + (1 - b2) * gi * gi;
        p[i] -= lr * (m[i] * c1) / (sqrtf(v[i] * c2) + eps);
    }
}

// sums for the scales: |h|^2, |v*|^2, |v* - P_u v_base|^2, |d|^2 over magnetic nodes
constexpr int NMOM = 4;
__global__ __launch_bounds__(256) void k_guess_moments(int N, const float *__restrict__ feat,
                                                       const int32_t *__restrict__ mag,
                                                       const NodeRec *__restrict__ nxt, double *__restrict__ part)
{
    __shared__ double red[NMOM][256];
    double s[NMOM] = {0, 0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        if (!mag[i]) continue;
        const float *f = feat + (size_t)i * FS;
        const NodeRec n = nxt[i];
        const double vs[3] = {n.vx, n.vy, n.vz};
        const float *vb = f + NF;
        const double ur = (double)f[0] * vb[0] + (double)f[1] * vb[1] + (double)f[2] * vb[2];
        for (int k = 0; k < 3; k++) {
            s[0] += (double)f[3 + k] * f[3 + k];
            s[1] += vs[k] * vs[k];
            const double d = vs[k] - (vb[k] - ur * f[k]);
            s[2] += d * d;
            s[3] += (double)f[9 + k] * f[9 + k];
        }
    }
    for (int k = 0; k < NMOM; k++) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int k = 0; k < NMOM; k++) red[k][threadIdx.x] += red[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < NMOM; k++) part[NMOM * blockIdx.x + k] = red[k][0];
}

template <class T>
void galloc(T **p, size_t n)
{
    if (hipMalloc((void **)p, std::max<size_t>(1, n) * sizeof(T)) != hipSuccess) throw std::string("hipMalloc (guess)");
}

void sync_or_throw(Ctx &c, const char *what)
{
    const hipError_t e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) throw std::string(what) + ": " + hipGetErrorString(e);
}

}  // namespace

void guess_destroy(Ctx &c)
{
    GuessNet *g = c.guess;
    if (!g) return;
    void *ptrs[] = {g->params, g->wbf, g->adam_m, g->adam_v, g->grad, g->feat, g->xin, g->z1, g->z2,
                    g->dz1, g->dz2, g->dout, g->gpart, g->lpart, g->xq, g->vhist, g->ninv};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (g->h_l) (void)hipHostFree(g->h_l);
    delete g;
    c.guess = nullptr;
}

// Allocate the network and draw initial weights (W1, W2 ~ N(0, 1/fan_in), W3 = 0, biases 0).
void guess_init(Ctx &c, uint64_t seed)
{
    if (!c.guess) {
        auto *g = new GuessNet();
        c.guess = g;
        g->N = c.N;
        for (int32_t m : c.h_mag) g->nmag += m != 0;
        g->nblk = (c.N + WG_CHUNK - 1) / WG_CHUNK;
        galloc(&g->params, NPARAM);
        galloc(&g->wbf, NTRAIN);
        galloc(&g->adam_m, NTRAIN);
        galloc(&g->adam_v, NTRAIN);
        galloc(&g->grad, NTRAIN);
        galloc(&g->feat, (size_t)c.N * FS);
        galloc(&g->vhist, (size_t)c.N * 6);
        galloc(&g->ninv, (size_t)c.N);
        k_guess_ncount<<<(c.N + 255) / 256, 256, 0, c.stream>>>(c.N, c.mag, c.slice_ptr, c.slice_w, c.scol, g->ninv);
        galloc(&g->lpart, 2 * ((c.N + GB - 1) / GB) + NMOM * 1024);
        const size_t nl = std::max<size_t>(NMOM * 1024, 2 * (size_t)((c.N + GB - 1) / GB));
        if (hipHostMalloc((void **)&g->h_l, nl * sizeof(double), hipHostMallocDefault) != hipSuccess)
            throw std::string("hipHostMalloc (guess)");
    }
    GuessNet *g = c.guess;
    std::vector<float> p(NPARAM, 0.f);
    std::mt19937_64 gen(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    for (int k = 0; k < NH * NX; k++) p[OFF_W1 + k] = nd(gen) / std::sqrt((float)NX);
    for (int k = 0; k < NH * NH; k++) p[OFF_W2 + k] = nd(gen) / std::sqrt((float)NH);
    guess_set(c, p.data());
    g->scaled = g->scaled_d = false;
}

void guess_set(Ctx &c, const float *params)
{
    GuessNet *g = c.guess;
    if (hipMemcpy(g->params, params, NPARAM * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        throw std::string("hipMemcpy (guess)");
    (void)hipMemset(g->adam_m, 0, NTRAIN * sizeof(float));
    (void)hipMemset(g->adam_v, 0, NTRAIN * sizeof(float));
    g->adam_t = 0;
    g->scaled = params[OFF_S] > 0 && params[OFF_S + 1] > 0 && params[OFF_S + 2] > 0;
    g->scaled_d = params[OFF_S + 3] > 0;
    k_guess_round<<<32, 256, 0, c.stream>>>(g->params, g->wbf);
    sync_or_throw(c, "guess set");
}

void guess_get(Ctx &c, float *params)
{
    sync_or_throw(c, "guess get");
    if (hipMemcpy(params, c.guess->params, NPARAM * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::string("hipMemcpy (guess)");
}

// after the assembly of a solve: record the features (always, so that plain solves feed training)
// and, when enabled, overwrite the initial guess
void guess_on_assembly(Ctx &c, double dt)
{
    GuessNet *g = c.guess;
    if (!g) return;
    const NodeRec *vsrc = c.rec[1 - c.cur];
    {
        Prof pr(c, "guess_features");
        // the previous v is a solution from the second assembly after a reset on, the one before it
        // from the third: extrapolate only between two solutions
        const bool ext = g->nhist >= 2 && g->dt_hist > 0;
        const float rho = ext ? (float)(dt / g->dt_hist) : 0.f;
        // the second difference from the fourth assembly on, when the last three solutions came at
        // this same step
        const bool dif = g->nhist >= 3 && dt == g->dt_hist && dt == g->dt_hist2;
        k_guess_features<<<(c.N + 255) / 256, 256, 0, c.stream>>>(c.N, c.rec[c.cur], vsrc, c.b, c.ep4, c.eq4, c.mass,
                                                                  c.mag, g->feat, g->vhist, rho, ext ? 1 : 0,
                                                                  dif ? 1 : 0);
        g->nhist++;
        g->dt_hist2 = g->dt_hist;
        g->dt_hist = dt;
    }
    g->have_batch = false;  // the solution of this solve is not known yet
    if (g->enabled && g->scaled) {
        if (!g->xq) galloc(&g->xq, (size_t)c.N * NXP);
        {
            Prof pr(c, "guess_input");
            k_guess_input<<<(c.N + 255) / 256, 256, 0, c.stream>>>(c.N, g->feat, c.mag, g->ninv, c.slice_ptr, c.slice_w,
                                                                   c.scol,
                                                                   g->params, g->xq);
        }
        Prof pr(c, "guess_forward");
        // two blocks per CU, each staging the weights once and looping over node tiles
        const int nb = std::min((c.N + 255) / 256, 512);
        k_guess_infer<<<nb, 256, 0, c.stream>>>(c.N, g->xq, g->feat, c.mag, g->params, g->wbf, c.ep4, c.eq4, c.x, c.w);
    }
}

void guess_restart(Ctx &c)
{
    if (c.guess) c.guess->nhist = 0;
}

void guess_on_solved(Ctx &c)
{
    if (c.guess) c.guess->have_batch = true;
}

static void ensure_training_buffers(GuessNet *g)
{
    if (g->xin) return;
    galloc(&g->xin, (size_t)g->N * NX);
    galloc(&g->z1, (size_t)g->N * NH);
    galloc(&g->z2, (size_t)g->N * NH);
    galloc(&g->dz1, (size_t)g->N * NH);
    galloc(&g->dz2, (size_t)g->N * NH);
    galloc(&g->dout, (size_t)g->N * NO);
    galloc(&g->gpart, (size_t)g->nblk * NH * (NH + 1));
}

static void compute_scales(Ctx &c)
{
    GuessNet *g = c.guess;
    const int nb = 1024;
    k_guess_moments<<<nb, 256, 0, c.stream>>>(c.N, g->feat, c.mag, c.rec[1 - c.cur], g->lpart);
    if (hipMemcpyAsync(g->h_l, g->lpart, NMOM * nb * sizeof(double), hipMemcpyDeviceToHost, c.stream) != hipSuccess)
        throw std::string("hipMemcpy (guess moments)");
    sync_or_throw(c, "guess moments");
    double s[NMOM] = {0, 0, 0, 0};
    for (int b = 0; b < nb; b++)
        for (int k = 0; k < NMOM; k++) s[k] += g->h_l[NMOM * b + k];
    const double n = std::max(1, g->nmag);
    if (!g->scaled) {
        float sc[3];
        sc[0] = s[0] > 0 ? (float)(1.0 / std::sqrt(s[0] / n)) : 1.f;  // 1 / rms|h|
        sc[1] = s[1] > 0 ? (float)(1.0 / std::sqrt(s[1] / n)) : 1.f;  // 1 / rms|v*|
        sc[2] = s[2] > 0 ? (float)std::sqrt(s[2] / n) : 1.f;          // rms|v* - P v_base|
        if (hipMemcpy(g->params + OFF_S, sc, 3 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            throw std::string("hipMemcpy (guess scales)");
        g->scaled = true;
    }
    // 1 / rms|d| from the first batch that has the second difference; until then the scale stays
    // zero and so does the network's d input
    if (!g->scaled_d && s[3] > 0) {
        const float sd = (float)(1.0 / std::sqrt(s[3] / n));
        if (hipMemcpy(g->params + OFF_S + 3, &sd, sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            throw std::string("hipMemcpy (guess scales)");
        g->scaled_d = true;
    }
}

// loss of the network and of the reference guess on the last solve, gradient into g->grad
static void loss_and_gradient(Ctx &c, double *loss, double *loss_ref)
{
    GuessNet *g = c.guess;
    ensure_training_buffers(g);
    float so;
    if (hipMemcpy(&so, g->params + OFF_S + 2, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::string("hipMemcpy (guess)");
    const double inv_norm = 1.0 / (std::max(1, g->nmag) * (double)so * (double)so);
    const int nbf = (c.N + GB - 1) / GB;
    k_guess_forward<1><<<nbf, GB, 0, c.stream>>>(c.N, inv_norm, g->feat, c.mag, g->ninv, c.slice_ptr, c.slice_w, c.scol,
                                                 g->params, g->wbf, c.ep4, c.eq4, c.rec[1 - c.cur], nullptr, nullptr,
                                                 g->xin, g->z1, g->z2, g->dout, g->lpart);
    k_guess_backward<<<nbf, GB, 0, c.stream>>>(c.N, g->params, g->wbf, g->z1, g->z2, g->dout, g->dz1, g->dz2);
    // weight gradients: (dout, h2) -> W3, b3; (dz2, h1) -> W2, b2; (dz1, x) -> W1, b1
    struct L {
        const float *A;
        int ra;
        const float *B;
        int rb, act, wofs, bofs;
    } layers[3] = {{g->dout, NO, g->z2, NH, 1, OFF_W3, OFF_B3},
                   {g->dz2, NH, g->z1, NH, 1, OFF_W2, OFF_B2},
                   {g->dz1, NH, g->xin, NX, 0, OFF_W1, OFF_B1}};
    for (auto &l : layers) {
        k_guess_wgrad<<<g->nblk, 256, 0, c.stream>>>(c.N, l.A, l.ra, l.B, l.rb, l.act, g->gpart);
        k_guess_reduce<<<32, 256, 0, c.stream>>>(g->nblk, l.ra, l.rb, g->gpart, l.wofs, l.bofs, g->grad);
    }
    if (hipMemcpyAsync(g->h_l, g->lpart, 2 * nbf * sizeof(double), hipMemcpyDeviceToHost, c.stream) != hipSuccess)
        throw std::string("hipMemcpy (guess loss)");
    sync_or_throw(c, "guess training");
    double a = 0, b = 0;
    for (int k = 0; k < nbf; k++) {
        a += g->h_l[2 * k];
        b += g->h_l[2 * k + 1];
    }
    if (loss) *loss = a;
    if (loss_ref) *loss_ref = b;
}

void guess_gradient(Ctx &c, float *grad, double *loss, double *loss_ref)
{
    GuessNet *g = c.guess;
    if (!g->have_batch) throw std::string("no solved step to train on");
    if (!g->scaled || !g->scaled_d) compute_scales(c);
    loss_and_gradient(c, loss, loss_ref);
    if (grad && hipMemcpy(grad, g->grad, NTRAIN * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::string("hipMemcpy (guess gradient)");
}

void guess_train(Ctx &c, float lr, double *loss, double *loss_ref)
{
    GuessNet *g = c.guess;
    if (!g->have_batch) throw std::string("no solved step to train on");
    if (!g->scaled || !g->scaled_d) compute_scales(c);
    loss_and_gradient(c, loss, loss_ref);
    g->adam_t++;
    const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)g->adam_t));
    const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)g->adam_t));
    k_guess_adam<<<32, 256, 0, c.stream>>>(g->params, g->grad, g->adam_m, g->adam_v, lr, c1, c2);
    k_guess_round<<<32, 256, 0, c.stream>>>(g->params, g->wbf);
    sync_or_throw(c, "guess update");
}

bool guess_enable(Ctx &c, int on)
{
    if (!c.guess) return false;
    c.guess->enabled = on != 0;
    return true;
}

void guess_batch(Ctx &c, float *feat)
{
    GuessNet *g = c.guess;
    sync_or_throw(c, "guess batch");
    if (hipMemcpy(feat, g->feat, (size_t)c.N * FS * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::string("hipMemcpy (guess batch)");
}

}  // namespace fg
