// Synthetic helical fan-beam projections of ellipsoid phantoms (gfx950).
//
// The reference ships no data (its readers, src/readBigPatientData.m and
// src/GetHeadofGeDataSet.m, need proprietary GE files), so the workloads are
// analytic: exact line integrals of sums of ellipsoids along the rays of a
// third-generation curved detector on a helical source path.
//
// Geometry (the conventions the reference's rebinning and backprojection use):
//   fan view i: source angle beta_i = BetaS + i * dBeta, source
//     S = (SO cos beta, SO sin beta, h beta / 2pi);
//   channel c (after the reference's channel flip, src/analyticalReconBigPatient.m:30):
//     fan angle gamma_c = (c - YLC0) dYA (src/ParallelRebinningCBCurve.cpp:40-42);
//   in-plane ray direction -(cos(beta - gamma), sin(beta - gamma)), i.e. the
//     parallel view theta = beta - gamma, offset t = SO sin gamma
//     (src/ParallelRebinningCBCurve.cpp:37-39);
//   row k: height (k - ZLC0) dZL on the cylinder at distance SD from the source.
// Output layout [ViewN][YL][ZL] (the rebinning input layout), FP32.
#include "arecon_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace arecon {
namespace {

struct Ellipsoid {
    double cx, cy, cz;      // centre, mm
    double ia, ib, ic;      // reciprocal semi-axes, 1/mm
    double cphi, sphi;      // rotation about z
    double mu;              // attenuation added inside, 1/mm
};

__global__ __launch_bounds__(256) void project_ellipsoids(
    float* __restrict__ out, const Ellipsoid* __restrict__ E, int nE, int YL, int ZL,
    int ViewN, double BetaS, double dBeta, double SO, double SD, double h, double YLC0,
    double dYA, double ZLC0, double dZL) {
    const long long n = (long long)ViewN * YL * ZL;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int k = (int)(id % ZL);
    const long long vc = id / ZL;
    const int c = (int)(vc % YL);
    const int i = (int)(vc / YL);
    // FP64 ray set-up (source positions over many rotations), FP32 per ellipsoid.
    const double beta = BetaS + i * dBeta;
    const double gamma = (c - YLC0) * dYA;
    const double theta = beta - gamma;
    const double sx = SO * cos(beta), sy = SO * sin(beta), sz = h * beta / (2.0 * M_PI);
    const double dxr = -cos(theta), dyr = -sin(theta);
    const double slope = (k - ZLC0) * dZL / SD;       // dz per unit in-plane length
    const double dlen = sqrt(1.0 + slope * slope);
    // FP64 throughout: rays start ~SO from the ellipsoid centres, and the
    // quadratic's discriminant cancels badly in FP32.
    double acc = 0.0;
    for (int e = 0; e < nE; ++e) {
        const Ellipsoid q = E[e];
        // Ray relative to the ellipsoid centre, rotated into its frame, scaled.
        const double px = sx - q.cx, py = sy - q.cy, pz = sz - q.cz;
        const double cp = q.cphi, sp = q.sphi;
        const double ax = (cp * px + sp * py) * q.ia;
        const double ay = (-sp * px + cp * py) * q.ib;
        const double az = pz * q.ic;
        const double bx = (cp * dxr + sp * dyr) * q.ia;
        const double by = (-sp * dxr + cp * dyr) * q.ib;
        const double bz = slope * q.ic;
        // closest approach in the scaled frame: m2 = |a + l0 b|^2, l0 = -(a.b)/(b.b)
        const double A = bx * bx + by * by + bz * bz;
        const double l0 = -(ax * bx + ay * by + az * bz) / A;
        const double mx = ax + l0 * bx, my = ay + l0 * by, mz = az + l0 * bz;
        const double m2 = mx * mx + my * my + mz * mz;
        if (m2 < 1.0) acc += q.mu * 2.0 * sqrt((1.0 - m2) / A) * dlen;
    }
    out[id] = (float)acc;
}

}  // namespace
}  // namespace arecon

using namespace arecon;

extern "C" int arecon_project_ellipsoids_dev(float* d_out, const float* ellipsoids, int nE,
                                             int YL, int ZL, int ViewN, double BetaS,
                                             double dBeta, double SO, double SD, double h,
                                             double YLC0, double dYA, double ZLC0, double dZL,
                                             void* stream) {
    return guarded([&] {
        ARECON_REQUIRE(d_out && (ellipsoids || nE == 0), "project: null pointer");
        ARECON_REQUIRE(nE >= 0 && YL > 0 && ZL > 0 && ViewN > 0, "project: bad sizes");
        hipStream_t s = static_cast<hipStream_t>(stream);
        // ellipsoids: nE rows of (cx, cy, cz, a, b, c, phi_radians, mu)
        std::vector<Ellipsoid> E(nE);
        for (int e = 0; e < nE; ++e) {
            const float* r = ellipsoids + 8 * e;
            ARECON_REQUIRE(r[3] > 0.f && r[4] > 0.f && r[5] > 0.f, "project: semi-axes must be positive");
            E[e] = Ellipsoid{r[0], r[1], r[2], 1.0 / r[3], 1.0 / r[4], 1.0 / r[5],
                             std::cos((double)r[6]), std::sin((double)r[6]), r[7]};
        }
        const long long n = (long long)ViewN * YL * ZL;
        const long long grid = (n + 255) / 256;
        ARECON_REQUIRE(grid < (1LL << 31), "project: too many samples");
        Ellipsoid* dE = nullptr;
        ARECON_HIP(hipMalloc(&dE, sizeof(Ellipsoid) * std::max(nE, 1)));
        struct Free { Ellipsoid* p; ~Free() { (void)hipFree(p); } } fr{dE};
        if (nE) ARECON_HIP(hipMemcpy(dE, E.data(), sizeof(Ellipsoid) * nE, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(project_ellipsoids, dim3((unsigned)grid), dim3(256), 0, s, d_out, dE, nE,
                           YL, ZL, ViewN, BetaS, dBeta, SO, SD, h, YLC0, dYA, ZLC0, dZL);
        ARECON_HIP(hipGetLastError());
        // The table is freed on return, so wait for the kernel.
        ARECON_HIP(hipStreamSynchronize(s));
    });
}
