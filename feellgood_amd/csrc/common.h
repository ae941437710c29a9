// common.h -- shared device/host definitions of the gfx950 LLG engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fg {

// physical constants (FeeLLGood src/config.h.in:31-32,59)
constexpr double MU0 = 1.25663706127e-6;
constexpr double GAMMA0 = 1.76085962784e11 * MU0;
constexpr double THETA = 0.5;

constexpr int NPI_TET = 5;
constexpr int NPI_TRI = 4;
constexpr int BLOCK = 256;        // threads per block of the streaming kernels
constexpr int MAX_PART = 1024;    // blocks of a grid-stride streaming kernel (partial sums)
constexpr int SELL_C = 64;        // rows per SELL slice = one wavefront

// Logical block index with the dispatcher's round-robin over the 8 XCDs undone: consecutive
// logical blocks (neighbouring nodes and elements, which share gathered data) run on the same XCD
// and its L2. A bijection of [0, nb).
__device__ inline int xcd_block(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + min(x, r) + (b >> 3);
}
// this block's logical index
__device__ inline int lblock() { return xcd_block((int)blockIdx.x, (int)gridDim.x); }

// shape functions at the Gauss points (src/tetra.h:190-220, src/triangle.h:42-64)
__host__ __device__ inline double tet_a(int i, int g)
{
    const double u[5] = {0.25, 1. / 6., 1. / 6., 1. / 6., 0.5};
    const double v[5] = {0.25, 1. / 6., 1. / 6., 0.5, 1. / 6.};
    const double w[5] = {0.25, 1. / 6., 0.5, 1. / 6., 1. / 6.};
    return i == 0 ? 1. - u[g] - v[g] - w[g] : (i == 1 ? u[g] : (i == 2 ? v[g] : w[g]));
}
__host__ __device__ inline double tri_a(int i, int g)
{
    const double u[4] = {1. / 3., 1. / 5., 3. / 5., 1. / 5.};
    const double v[4] = {1. / 3., 1. / 5., 1. / 5., 3. / 5.};
    return i == 0 ? 1. - u[g] - v[g] : (i == 1 ? u[g] : v[g]);
}
constexpr double TET_PDS[5] = {-2. / 15., 3. / 40., 3. / 40., 3. / 40., 3. / 40.};
constexpr double TRI_PDS[4] = {-27. / 96., 25. / 96., 25. / 96., 25. / 96.};

// One node state record (dataNode, src/node.h:47-53): 64 bytes, one cache line half.
struct __align__(64) NodeRec {
    double ux, uy, uz, vx, vy, vz, phi, phiv;
};

// Material of a volume region, in the form the element kernel consumes.
struct DevRegion {
    double alpha;   // alpha_LLG
    double Abis;    // 2 A / (mu0 Ms)
    double Kbis;    // 2 K / (mu0 Ms)
    double K3bis;   // 2 K3 / (mu0 Ms)
    double Ms;
    double Ae, K, K3;
    double uk[3];
    double ex[3], ey[3], ez[3];
    int hasK, hasK3;
};

// Krylov control block, lives in device memory; written by block 0 of the kernels.
struct KrylovCtl {
    double rho1, rho2, alpha, omega, beta;
    double rhsn;     // ||b||
    double res;      // last monitored residual norm
    double tol;
    double vmax2;    // max (x_2i^2 + x_2i+1^2) after the update
    int nit;
    int status;      // FG_ITER_*
    int done;
    int max_iter;
    int refinements;  // FP64 refinement sweeps of the mixed-precision solve
    // set by the t = A s phase when ||s|| meets the criterion, turned into done by the next phase:
    // a block of that launch can start after block 0 has written, so it may not test done itself
    int stop;
};

// small vector helpers
struct d3 {
    double x, y, z;
};
__host__ __device__ inline d3 mk(double x, double y, double z) { return {x, y, z}; }
__host__ __device__ inline d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline d3 operator*(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline d3 cross(d3 a, d3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

}  // namespace fg
