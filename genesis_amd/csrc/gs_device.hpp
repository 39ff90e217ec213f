// gs_device.hpp — device-side helpers shared by the gfx950 kernels.
//
// Wave64 (CDNA4) throughout.  The library is compiled with
// -ffp-contract=off: every fused multiply-add in the hot loops is written as
// an explicit fma(), and the expressions that must be bit-exact against the
// reference (cell keys, wrap, integrator, g(r) distances and bins) are
// written with the reference's operation order and no contraction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/genesis_gpu.h"

namespace gs {

constexpr int WAVE = 64;
constexpr double F64_EPS = 2.220446049250313e-16;  // f64::EPSILON

// f64::rem_euclid (Rust std), src/systems/boundary.rs:28-34 and
// src/physics/grid.rs:230.
__device__ __forceinline__ double rem_euclid(double a, double b) {
    // fmod(a, b) is a itself when |a| < |b| (the case of every wrap after
    // one tick of motion), so the general remainder runs only otherwise
    const double r = fabs(a) < fabs(b) ? a : fmod(a, b);
    return r < 0.0 ? r + fabs(b) : r;
}

// `f as u32` (saturating, NaN -> 0).
__device__ __forceinline__ uint32_t f64_as_u32(double x) {
    if (!(x > 0.0)) return 0u;
    if (x >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

// Cell coordinate on one axis exactly as SpatialGrid::cell_of
// (src/physics/grid.rs:228-238).
__device__ __forceinline__ uint32_t cell_axis(double v, double period, double cw, uint32_t n) {
    double u = rem_euclid(v + 0.5 * period, period);
    uint32_t c = f64_as_u32(floor(u / cw));
    return c < n - 1 ? c : n - 1;
}

// min_image, src/physics/grid.rs:300-306 (round half away from zero).
__device__ __forceinline__ double min_image1(double d, double size) {
    return d - size * round(d / size);
}

// 1/sqrt(x) for x > 0, to ~1 ulp: hardware v_rsq_f64 seed plus one
// third-order Newton correction, y1 = y0 (1 + e/2 + 3e^2/8) with
// e = 1 - x y0^2 (the binomial series of (1 - e)^(-1/2)).
__device__ __forceinline__ double rsqrt_refined(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x * y, y, 1.0);
    double t = fma(0.375, e, 0.5);
    return fma(y * e, t, y);
}

// Wave-wide sum of a double; every lane gets the same total (the xor
// butterfly pairs the same operands on both partners, so the result is
// bitwise identical across lanes and across runs).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ uint32_t mask_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Per species-pair constants of the switched LJ law (LjTable,
// src/physics/forces.rs:213-237) pre-arranged for the pair kernel.
struct PairCoef {
    double sigma;      // sigma_ij = (sigma_i + sigma_j)/2
    double eps4;       // 4 eps_ij
    double c24;        // 24 eps_ij / sigma_ij
    double rcore;      // LJ_NUCLEUS * sigma_ij (hardened core radius)
    double inv_rcore;  // 1 / rcore
    double mm;         // m_i m_j (gravity)
    double bond_k;     // well_curvature(eps_ij, sigma_ij)
    double bond_req;   // equilibrium_distance(sigma_ij)
};

// Per-species factors of the Lorentz-Berthelot mixing rules, stored after
// the pair table in the same buffer: sigma_ij = (sigma_i + sigma_j) / 2
// (the same expression as PairCoef::sigma) and 4 eps_ij = a4_i a4_j with
// a4 = sqrt(4 kT eps_k).  A pair kernel reading these from LDS touches ten
// 16-byte entries instead of up to a hundred 64-byte rows, whose reads by
// the lanes of a wave fall into the same banks.
struct SpCoef {
    double sigma;
    double a4;
};
constexpr int COEF_SP_OFFSET = GS_NSPECIES * GS_NSPECIES;  // in PairCoef units
constexpr int COEF_LEN = COEF_SP_OFFSET + (int)((GS_NSPECIES * sizeof(SpCoef) + 63) / 64);

// Per-species-pair parameters of the differentiable model (adjoint).
struct AdjCoef {
    double sigma;            // sigma_ij
    double eps;              // eps_ij = k sqrt(epsK_i epsK_j)
    double eps_over_epsK_i;  // eps_ij / epsK_i
    double inv_sigma;        // 1 / sigma_ij
};

// Scalar parameters every kernel reads (passed by value).
struct KParams {
    double box[3];
    double half[3];
    double inv_half[3];    // 1 / half
    double cell[3];        // force-grid cell edge per axis
    uint32_t dims[3];      // force-grid cells per axis
    uint32_t ncell;
    int32_t n;             // atoms
    int32_t flags;
    double dt, half_dt;
    double rc, rc2, eps2;  // grid cutoff (largest enabled)
    double lj_rc, lj_ron, lj_inv_w;
    double c_rc, c_ron, c_inv_w, ke;
    double g_rc, g_ron, g_inv_w, gconst;
    double mass[GS_NSPECIES];
    double inv_mass[GS_NSPECIES];
};

enum : int32_t {
    F_COULOMB = 1,
    F_GRAVITY = 2,
    F_BONDS = 4,
    F_WRAP = 8,
    F_SHARED_SWITCH = 16,  // Coulomb switch identical to the LJ one
    F_REF_MULT = 32,       // reference multiplicity on 2-cell axes
    F_SMALL = 64,          // < 3 cells on some axis (brute-force pairs)
};

}  // namespace gs
