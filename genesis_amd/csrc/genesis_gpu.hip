// genesis_gpu.hip — MI355X (gfx950) force-and-integration path of GENESIS
// behind the C ABI of include/genesis_gpu.h.
//
// Resident state lives in HBM in cell-sorted order (structure of arrays):
//   pos4[k] = (x, y, z, q)   double4, the pair kernel's tile format
//   vx/vy/vz, ax/ay/az       double SoA, coalesced integrator traffic
//   type[k] (u8), id[k] (u32 entity slot of sorted position k)
// Every force pass re-sorts the atoms by cell: keys + counting (atomic u32
// histogram), one exclusive scan, a scatter, then a per-cell ordering by
// entity id that makes the layout independent of atomic arrival order.  The
// pair kernel works cell by cell: one 256-thread workgroup per home cell
// stages the 27-cell stencil in LDS (positions pre-shifted by the periodic
// image), each wave takes home atoms in turn, sweeps the tile 64 candidates
// at a time, compacts the in-range ones with a ballot into an LDS queue and
// evaluates full 64-lane batches of the switched LJ + Coulomb law.  Each atom
// owns its own sum (full shell), so no floating-point atomics exist and
// results are bitwise reproducible run to run.
#include "gs_device.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <type_traits>
#include <vector>

using namespace gs;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return fail(GS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int BLOCK = 256;
constexpr int PAIR_WAVES = 4;
constexpr int PAIR_THREADS = PAIR_WAVES * WAVE;
constexpr int TILE_CAP = 1024;  // candidates staged per LDS chunk
constexpr int BUILD_WAVES = 8;  // waves of a list-build workgroup
constexpr int BUILD_TILE = 1280;  // candidates the list build stages in LDS per chunk
#ifndef GS_BUILD_WG_PER_CU
#define GS_BUILD_WG_PER_CU 1
#endif
constexpr uint32_t BUILD_WG_PER_CU = GS_BUILD_WG_PER_CU;  // list-build workgroups per CU
constexpr int BUILD_THREADS = BUILD_WAVES * WAVE;
constexpr int QCAP = 128;       // per-wave compaction ring (u16 tile indices)

inline unsigned nblocks(int64_t n, int b = BLOCK) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// =============================================================================
// Kernels
// =============================================================================

// Cell key of a position on the force grid: SpatialGrid::cell_of + encode
// (src/physics/grid.rs:228-246).
__device__ __forceinline__ uint32_t cell_key(const KParams& P, double x, double y, double z) {
    uint32_t cx = cell_axis(x, P.box[0], P.cell[0], P.dims[0]);
    uint32_t cy = cell_axis(y, P.box[1], P.cell[1], P.dims[1]);
    uint32_t cz = cell_axis(z, P.box[2], P.cell[2], P.dims[2]);
    return cx + P.dims[0] * (cy + P.dims[1] * cz);
}

// Fused: thermostat scale of the previous tick (v *= lambda,
// src/systems/thermostat.rs:68-70), half kick (src/systems/integrate.rs:38-44),
// drift (:61-67), periodic wrap (src/systems/boundary.rs:28-34), then the cell
// key and the atomic histogram of the rebuild.  Every expression keeps the
// reference's operation order, so given the same accelerations the result
// is bitwise the reference's.
template <bool KICK_DRIFT>
__global__ __launch_bounds__(BLOCK) void k_integrate_key(KParams P, double4* __restrict__ pos4,
                                                         double* __restrict__ vx, double* __restrict__ vy,
                                                         double* __restrict__ vz, const double* __restrict__ ax,
                                                         const double* __restrict__ ay, const double* __restrict__ az,
                                                         const double* __restrict__ lam_ptr,
                                                         uint32_t* __restrict__ key, uint32_t* __restrict__ ticket,
                                                         uint32_t* __restrict__ count) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    const bool valid = k < P.n;
    double4 p = valid ? pos4[k] : make_double4(0.0, 0.0, 0.0, 0.0);
    if (KICK_DRIFT && valid) {
        double lam = *lam_ptr;
        double a = vx[k] * lam, b = vy[k] * lam, c = vz[k] * lam;
        a = a + ax[k] * P.half_dt;
        b = b + ay[k] * P.half_dt;
        c = c + az[k] * P.half_dt;
        vx[k] = a;
        vy[k] = b;
        vz[k] = c;
        p.x = p.x + a * P.dt;
        p.y = p.y + b * P.dt;
        p.z = p.z + c * P.dt;
        if (P.flags & F_WRAP) {
            p.x = rem_euclid(p.x, 2.0 * P.half[0]) - P.half[0];
            p.y = rem_euclid(p.y, 2.0 * P.half[1]) - P.half[1];
            p.z = rem_euclid(p.z, 2.0 * P.half[2]) - P.half[2];
        }
        pos4[k] = p;
    }
    const uint32_t kk = valid ? cell_key(P, p.x, p.y, p.z) : 0xFFFFFFFFu;
    // The state is cell-sorted, so a wave's atoms fall into a few runs of
    // equal cells: each run takes its tickets with one atomic (per-lane
    // atomics on one address would serialise).  The order of the tickets
    // within a cell is undone by k_sort_gather's per-cell ordering.
    const int lane = threadIdx.x & 63;
    const uint32_t prev = __shfl_up(kk, 1, WAVE);
    const bool head = valid && (lane == 0 || prev != kk);
    const uint64_t heads = __ballot(head);
    const int nvalid = __popcll(__ballot(valid));
    const uint64_t upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
    const uint64_t hb = heads & upto;
    const int rs = hb ? 63 - __builtin_clzll(hb) : 0;  // head of this lane's run
    const uint64_t ha = heads & ~upto;
    const int re = ha ? __builtin_ctzll(ha) : nvalid;  // one past this lane's run
    uint32_t base = 0;
    if (head) base = atomicAdd(&count[kk], (uint32_t)(re - rs));
    base = __shfl(base, rs, WAVE);
    if (valid) {
        key[k] = kk;
        ticket[k] = base + (uint32_t)(lane - rs);
    }
}

// Exclusive scan of the per-cell histogram into cell_start[0..ncell], and
// reset of the histogram for the next rebuild.  One 1024-thread workgroup,
// 16 cells per thread per pass, wave scans with shuffles.
__global__ __launch_bounds__(1024) void k_scan(uint32_t* __restrict__ count, uint32_t* __restrict__ start,
                                               uint32_t ncell) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < ncell; base += 1024u * 16u) {
        uint32_t v[16];
        uint32_t s = 0;
        const uint32_t i0 = base + (uint32_t)tid * 16u;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            uint32_t idx = i0 + e;
            v[e] = idx < ncell ? count[idx] : 0u;
            s += v[e];
        }
        uint32_t incl = s;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t t = __shfl_up(incl, off, 64);
            if (lane >= off) incl += t;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        if (w == 0) {
            uint32_t x = lane < 16 ? wsum[lane] : 0u;
            uint32_t xi = x;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                uint32_t t = __shfl_up(xi, off, 64);
                if (lane >= off) xi += t;
            }
            if (lane < 16) wsum[lane] = xi - x;
        }
        __syncthreads();
        uint32_t excl = carry + wsum[w] + incl - s;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            uint32_t idx = i0 + e;
            if (idx < ncell) {
                start[idx] = excl;
                count[idx] = 0u;
            }
            excl += v[e];
        }
        __syncthreads();
        if (tid == 1023) carry = excl;
        __syncthreads();
    }
    if (tid == 0) start[ncell] = carry;
}

// Multi-block exclusive scan of a large cell histogram (more than one
// k_scan pass of 16384 cells): k_scan_part scans each chunk of SCAN_CHUNK
// cells in its own 256-thread workgroup (and resets the histogram), k_scan
// of the chunk totals gives each chunk's offset, and k_scan_fix adds it.
// The cell starts are the same integers as k_scan's.
constexpr uint32_t SCAN_CHUNK = 256u * 8u;
__global__ __launch_bounds__(256) void k_scan_part(uint32_t* __restrict__ count, uint32_t* __restrict__ start,
                                                   uint32_t ncell, uint32_t* __restrict__ bsum) {
    __shared__ uint32_t wsum[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t i0 = blockIdx.x * SCAN_CHUNK + (uint32_t)tid * 8u;
    uint32_t v[8];
    uint32_t s = 0;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const uint32_t idx = i0 + e;
        v[e] = idx < ncell ? count[idx] : 0u;
        s += v[e];
    }
    uint32_t incl = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        before += k < w ? wsum[k] : 0u;
        total += wsum[k];
    }
    if (tid == 0) bsum[blockIdx.x] = total;
    uint32_t excl = before + incl - s;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const uint32_t idx = i0 + e;
        if (idx < ncell) {
            start[idx] = excl;
            count[idx] = 0u;
        }
        excl += v[e];
    }
}

__global__ __launch_bounds__(256) void k_scan_fix(uint32_t* __restrict__ start, uint32_t ncell,
                                                  const uint32_t* __restrict__ boff) {
    const uint32_t off = boff[blockIdx.x];
    const uint32_t i0 = blockIdx.x * SCAN_CHUNK + threadIdx.x * 8u;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const uint32_t idx = i0 + e;
        if (idx < ncell) start[idx] += off;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) start[ncell] = boff[gridDim.x];
}

// Scratch words the multi-block scan of ncell cells needs (chunk totals and
// their offsets).
inline size_t scan_tmp_words(uint32_t ncell) {
    const size_t nb = (ncell + SCAN_CHUNK - 1) / SCAN_CHUNK;
    return 2 * (nb + 1);
}

// Exclusive scan of count[0..ncell) into start[0..ncell] (start[ncell] =
// total) with count reset to zero; tmp: scan_tmp_words(ncell) words.
inline void launch_scan(hipStream_t st, uint32_t* count, uint32_t* start, uint32_t ncell, uint32_t* tmp) {
    if (ncell <= 16384u || !tmp) {
        k_scan<<<1, 1024, 0, st>>>(count, start, ncell);
        return;
    }
    const uint32_t nb = (ncell + SCAN_CHUNK - 1) / SCAN_CHUNK;
    k_scan_part<<<nb, 256, 0, st>>>(count, start, ncell, tmp);
    k_scan<<<1, 1024, 0, st>>>(tmp, tmp + nb + 1, nb);
    k_scan_fix<<<nb, 256, 0, st>>>(start, ncell, tmp + nb + 1);
}

__global__ __launch_bounds__(BLOCK) void k_scatter(int n, const uint32_t* __restrict__ key,
                                                   const uint32_t* __restrict__ ticket,
                                                   const uint32_t* __restrict__ start, uint32_t* __restrict__ perm) {
    int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n) return;
    perm[start[key[k]] + ticket[k]] = (uint32_t)k;
}

// One wave per cell: order the cell's atoms by entity id (rank by counting;
// ids are unique) and gather the state into the new sorted arrays.
__global__ __launch_bounds__(BLOCK) void k_sort_gather(uint32_t ncell, const uint32_t* __restrict__ start,
                                                       const uint32_t* __restrict__ perm,
                                                       const double4* __restrict__ p_in,
                                                       const double* __restrict__ vx_in,
                                                       const double* __restrict__ vy_in,
                                                       const double* __restrict__ vz_in,
                                                       const uint8_t* __restrict__ t_in,
                                                       const uint32_t* __restrict__ id_in, double4* __restrict__ p_out,
                                                       double* __restrict__ vx_out, double* __restrict__ vy_out,
                                                       double* __restrict__ vz_out, uint8_t* __restrict__ t_out,
                                                       uint32_t* __restrict__ id_out) {
    __shared__ uint32_t sid[BLOCK / WAVE][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t c = blockIdx.x * (BLOCK / WAVE) + w;
    const bool valid = c < ncell;
    uint32_t s = 0, m = 0;
    if (valid) {
        s = start[c];
        m = start[c + 1] - s;
    }
    const bool lds = m <= 256;
    if (valid && lds)
        for (uint32_t e = lane; e < m; e += 64) sid[w][e] = id_in[perm[s + e]];
    __syncthreads();
    if (!valid) return;
    for (uint32_t e = lane; e < m; e += 64) {
        const uint32_t k = perm[s + e];
        const uint32_t myid = lds ? sid[w][e] : id_in[k];
        uint32_t rank = 0;
        for (uint32_t f = 0; f < m; ++f) {
            uint32_t o = lds ? sid[w][f] : id_in[perm[s + f]];
            rank += (o < myid) ? 1u : 0u;
        }
        const uint32_t p = s + rank;
        p_out[p] = p_in[k];
        vx_out[p] = vx_in[k];
        vy_out[p] = vy_in[k];
        vz_out[p] = vz_in[k];
        t_out[p] = t_in[k];
        id_out[p] = myid;
    }
}

// ---- pair law ----------------------------------------------------------------

// Quintic switch of smooth_cutoff (src/physics/forces.rs:302-317) with u
// clamped to [0, 1]: sw = 1 - 10u^3 + 15u^4 - 6u^5, dsw = -30u^2(1-u)^2 / w.
// u <= 0 gives (1, 0) and u >= 1 gives (0, 0), the reference's two early
// returns, so no branch is needed.
__device__ __forceinline__ void switch_q(double r, double r_on, double inv_w, double& sw, double& dsw) {
    double u = (r - r_on) * inv_w;
    u = fmin(fmax(u, 0.0), 1.0);
    double u2 = u * u;
    double u3 = u2 * u;
    sw = fma(u3, fma(u, fma(-6.0, u, 15.0), -10.0), 1.0);
    double om = 1.0 - u;
    dsw = (-30.0 * inv_w) * (u2 * (om * om));
}

// Force magnitude along the normal and potential of one pair at distance d
// (d = r2 * inv_d): the fold body of ForceSystem::run
// (src/systems/forces.rs:270-299) composed of force_switched_scalar
// (src/physics/forces.rs:270-275, hardened core at LJ_NUCLEUS sigma), the
// Coulomb term with r = max(d, 0.5) switched at d, and gravity.
// GRAVM: 0 no gravity term, 1 gravity term, 2 as P.flags says.
template <bool COUL, bool SHARED, int GRAVM = 2>
__device__ __forceinline__ void pair_law(const KParams& P, const PairCoef& pc, double d, double inv_d, double qq,
                                         double& fmag, double& vv) {
    const bool core = d < pc.rcore;
    const double r = core ? pc.rcore : d;
    const double inv_r = core ? pc.inv_rcore : inv_d;
    const double s = pc.sigma * inv_r;
    const double s2 = s * s;
    const double s6 = s2 * s2 * s2;
    const double s12 = s6 * s6;
    const double vlj = pc.eps4 * (s12 - s6);
    const double flj = (pc.c24 * s) * fma(2.0, s12, -s6);
    double sw, dsw;
    switch_q(r, P.lj_ron, P.lj_inv_w, sw, dsw);
    if (COUL && SHARED) {
        const bool cc = d < 0.5;
        const double inv_c = cc ? 2.0 : inv_d;
        const double vc = qq * inv_c;
        const double fc = vc * inv_c;
        const double fs = flj + fc, vs = vlj + vc;
        fmag = fma(fs, sw, -vs * dsw);
        vv = vs * sw;
    } else {
        fmag = fma(flj, sw, -vlj * dsw);
        vv = vlj * sw;
        if (COUL) {
            const bool cc = d < 0.5;
            const double inv_c = cc ? 2.0 : inv_d;
            const double vc = qq * inv_c;
            const double fc = vc * inv_c;
            double swc, dswc;
            switch_q(d, P.c_ron, P.c_inv_w, swc, dswc);
            fmag += fma(fc, swc, -vc * dswc);
            vv += vc * swc;
        }
    }
    if (GRAVM == 1 || (GRAVM == 2 && (P.flags & F_GRAVITY))) {
        const bool cc = d < 0.5;
        const double inv_g = cc ? 2.0 : inv_d;
        const double vg = -(P.gconst * pc.mm) * inv_g;
        const double fg = vg * inv_g;
        double swg, dswg;
        switch_q(d, P.g_ron, P.g_inv_w, swg, dswg);
        fmag += fma(fg, swg, -vg * dswg);
        vv += vg * swg;
    }
}

// pair_law with the LJ coefficients of the pair formed from per-species
// factors (SpCoef): sigma_ij and the core radius sigma_ij / 2 as in the
// pair table, so the hardened-core test takes the same branch; 4 eps_ij =
// a4_i a4_j and 24 eps_ij s / sigma_ij = 6 (4 eps_ij) / r, equal to the
// table's values up to rounding.  inv_rc is the table's 1 / rcore of the
// pair (read by the caller whether or not the pair is inside its core),
// rowc[tj].mm the mass product of the gravity term.  The function has no
// branch and no masked region, and GRAV is the gravity switch as a
// template argument: a runtime test between the pairs of a group would
// split them into separate basic blocks and keep the compiler from
// interleaving their FP64 dependency chains.
template <bool COUL, bool SHARED, bool GRAV>
__device__ __forceinline__ void pair_law_sp(const KParams& P, double sig_i, double a4_i, const SpCoef& cj,
                                            double inv_rc, const PairCoef* rowc, int tj, double d, double inv_d,
                                            double qq, double& fmag, double& vv) {
    const double sig = 0.5 * (sig_i + cj.sigma);
    const double eps4 = a4_i * cj.a4;
    const double rcore = 0.5 * sig;
    const bool core = d < rcore;
    const double r = core ? rcore : d;
    const double inv_r = core ? inv_rc : inv_d;
    const double s = sig * inv_r;
    const double s2 = s * s;
    const double s6 = s2 * s2 * s2;
    const double s12 = s6 * s6;
    const double vlj = eps4 * (s12 - s6);
    const double flj = (6.0 * eps4 * inv_r) * fma(2.0, s12, -s6);
    double sw, dsw;
    switch_q(r, P.lj_ron, P.lj_inv_w, sw, dsw);
    if (COUL && SHARED) {
        const bool cc = d < 0.5;
        const double inv_c = cc ? 2.0 : inv_d;
        const double vc = qq * inv_c;
        const double fc = vc * inv_c;
        const double fs = flj + fc, vs = vlj + vc;
        fmag = fma(fs, sw, -vs * dsw);
        vv = vs * sw;
    } else {
        fmag = fma(flj, sw, -vlj * dsw);
        vv = vlj * sw;
        if (COUL) {
            const bool cc = d < 0.5;
            const double inv_c = cc ? 2.0 : inv_d;
            const double vc = qq * inv_c;
            const double fc = vc * inv_c;
            double swc, dswc;
            switch_q(d, P.c_ron, P.c_inv_w, swc, dswc);
            fmag += fma(fc, swc, -vc * dswc);
            vv += vc * swc;
        }
    }
    if (GRAV) {
        const bool cc = d < 0.5;
        const double inv_g = cc ? 2.0 : inv_d;
        const double vg = -(P.gconst * rowc[tj].mm) * inv_g;
        const double fg = vg * inv_g;
        double swg, dswg;
        switch_q(d, P.g_ron, P.g_inv_w, swg, dswg);
        fmag += fma(fg, swg, -vg * dswg);
        vv += vg * swg;
    }
}

// Bound-pair screen of the list pass: k_bind^2 narrowed and widened by
// 1e-10 relative.
struct BoundScreen {
    double kb2_lo = 0.0, kb2_hi = 0.0;  // k_bind^2 (1 -+ 1e-10)
};

// Per-species-pair factors of the list pass, formed once per workgroup in
// LDS with the expressions pair_law_sp evaluates per pair: sigma_ij =
// (sigma_i + sigma_j) / 2, 4 eps_ij = a4_i a4_j, its sixfold (the factor of
// the LJ force), the table's 1 / rcore, the core radius sigma_ij / 2, and
// the bound-pair screen's thresholds sigma_ij^2 k_bind^2 (1 -+ 1e-10).
// The table has LIST_COLS columns per row: a list entry's four species bits
// index it directly (columns past the last species repeat it), so a pair's
// row address is its entry shifted, with no clamp.
struct ListCoef {
    double sig, eps4;
    double e6, irc;
    double rcore, thlo;
    double thhi, pad;
};
constexpr int LIST_COLS = 16;

// Four pairs of one atom at once, statement by statement (each statement
// over the four pairs before the next), with the expressions of the list
// traversal and pair_law_sp: a dependent FP64 operation waits for its
// operand's result for several issue slots, so the four independent chains
// are interleaved to fill them.  p: the four neighbour positions; ent:
// their list entries (species in bits 28-31, the column of lrow).  A slot
// past the end of the list holds the atom's own entry (r^2 = 0), which the
// tests below reject: such a pair (and a coincident one, r^2 <= eps2) is
// evaluated at r^2 = eps2 and scaled by 0, so no liveness mask is carried.
// Q0, NQ: the pairs q0 .. q0 + nq - 1 of the group (a group is evaluated as
// two halves, which holds fewer values live at once; the sums take the
// pairs in the same order either way).
// INNER: every pair of the group is known to lie inside every switching
// radius (d < r_on) on this tick, where the quintic switch is exactly 1 with
// derivative 0: its evaluation is skipped (fma(f, 1, -v * -0) = f, v * 1 = v,
// so the sums are bit-identical to the switched law's).
// SCREEN: bits 0-3 of hitm mark the pairs that are bound pairs for certain
// (r^2 at least 1e-10 relative inside (k_bind sigma_ij)^2, so no rounding of
// the reference's own expressions can move them out), bits 4-7 the pairs
// within that margin of the threshold, which the exact test decides.
template <bool COUL, bool SHARED, bool GRAV, bool INNER, bool SCREEN, int Q0, int NQ>
__device__ __forceinline__ void pair_eval4(const KParams& P, double Lx, double Ly, double Lz, double iLx, double iLy,
                                           double iLz, const double4& pi, double kqi, const ListCoef* lrow,
                                           const PairCoef* row, const double4 (&p)[4], const uint4& ent, double& fx,
                                           double& fy, double& fz, double& vl, uint32_t& hitm) {
    bool core[4];
    uint32_t tj[4];
    double dx[4], dy[4], dz[4], r2[4], y[4], e[4], inv_d[4], d[4], sig[4], eps4[4], e6[4], irc[4], rcore[4], r[4];
    double inv_r[4], okf[4], traw[4], thlo[4], thhi[4];
    double s6[4], s12[4], vlj[4], flj[4], sw[4], dsw[4], fmag[4], vv[4];
    {
        const uint32_t f4[4] = {ent.x, ent.y, ent.z, ent.w};
#pragma unroll
        for (int q = Q0; q < Q0 + NQ; ++q) tj[q] = f4[q] >> 28;
    }
#pragma unroll
    for (int q = Q0; q < Q0 + NQ; ++q) {
        dx[q] = pi.x - p[q].x;
        dy[q] = pi.y - p[q].y;
        dz[q] = pi.z - p[q].z;
    }
#pragma unroll
    for (int q = Q0; q < Q0 + NQ; ++q) {
        dx[q] = fma(-Lx, rint(dx[q] * iLx), dx[q]);
        dy[q] = fma(-Ly, rint(dy[q] * iLy), dy[q]);
        dz[q] = fma(-Lz, rint(dz[q] * iLz), dz[q]);
    }
#pragma unroll
    for (int q = Q0; q < Q0 + NQ; ++q) {
        // a masked pair (r^2 <= eps2: the atom itself past the end of its
        // list, or a coincident pair) is evaluated at r^2 = eps2, where every
        // term is finite, and its force and energy are scaled by okf = 0
        const double t = fma(dz[q], dz[q], fma(dy[q], dy[q], dx[q] * dx[q]));
        okf[q] = t > P.eps2 ? 1.0 : 0.0;
        r2[q] = fmax(t, P.eps2);
        if (SCREEN) traw[q] = t;
    }
    // rsqrt_refined, four at a time
#pragma unroll
    for (int q = Q0; q < Q0 + NQ; ++q) y[q] = __builtin_amdgcn_rsq(r2[q]);
#pragma unroll
    for (int q = Q0; q < Q0 + NQ; ++q) e[q] = fma(-r2[q] * y[q], y[q], 1.0);
#pragma unroll
    for (int q = Q0; q < Q0 + NQ; ++q) inv_d[q] = fma(y[q] * e[q], fma(0.375, e[q], 0.5), y[q]);
#pragma unroll
    for (int q = Q0; q < Q0 + NQ; ++q) {
        d[q] = r2[q] * inv_d[q];
        const ListCoef& c = lrow[tj[q]];
        sig[q] = c.sig;
        eps4[q] = c.eps4;
        e6[q] = c.e6;
        irc[q] = c.irc;
        rcore[q] = c.rcore;
        if (SCREEN) {
            thlo[q] = c.thlo;
            thhi[q] = c.thhi;
        }
    }
    if (SCREEN) {
        uint32_t m = 0u;
#pragma unroll
        for (int q = Q0; q < Q0 + NQ; ++q) {
            // (k_bind sigma_ab <= k_bind max sigma: inside the pair's own
            // threshold is inside the cutoff; r^2 = 0 is no pair); the tests
            // are combined without short-circuit so no branch is formed
            const double t = traw[q];
            const uint32_t sure = (uint32_t)(t > 1e-24) & (uint32_t)(t < thlo[q]);
            const uint32_t maybe = (uint32_t)(t > 0.0) & (uint32_t)(t < thhi[q]);
            m |= sure << q;
            m |= (maybe & (sure ^ 1u)) << (4 + q);
        }
        hitm |= m;
    }
#pragma unroll
    for (int q = Q0; q < Q0 + NQ; ++q) {
        core[q] = d[q] < rcore[q];
        r[q] = core[q] ? rcore[q] : d[q];
        inv_r[q] = core[q] ? irc[q] : inv_d[q];
    }
#pragma unroll
    for (int q = Q0; q < Q0 + NQ; ++q) {
        const double s = sig[q] * inv_r[q];
        const double s2 = s * s;
        s6[q] = s2 * s2 * s2;
        s12[q] = s6[q] * s6[q];
    }
#pragma unroll
    for (int q = Q0; q < Q0 + NQ; ++q) {
        vlj[q] = eps4[q] * (s12[q] - s6[q]);
        flj[q] = (e6[q] * inv_r[q]) * fma(2.0, s12[q], -s6[q]);
    }
    if (!INNER) {
#pragma unroll
        for (int q = Q0; q < Q0 + NQ; ++q) switch_q(r[q], P.lj_ron, P.lj_inv_w, sw[q], dsw[q]);
    }
    if (COUL && SHARED) {
#pragma unroll
        for (int q = Q0; q < Q0 + NQ; ++q) {
            const bool cc = d[q] < 0.5;
            const double inv_c = cc ? 2.0 : inv_d[q];
            const double vc = (kqi * p[q].w) * inv_c;
            const double fc = vc * inv_c;
            const double fs = flj[q] + fc, vs = vlj[q] + vc;
            fmag[q] = INNER ? fs : fma(fs, sw[q], -vs * dsw[q]);
            vv[q] = INNER ? vs : vs * sw[q];
        }
    } else {
#pragma unroll
        for (int q = Q0; q < Q0 + NQ; ++q) {
            fmag[q] = INNER ? flj[q] : fma(flj[q], sw[q], -vlj[q] * dsw[q]);
            vv[q] = INNER ? vlj[q] : vlj[q] * sw[q];
        }
        if (COUL) {
#pragma unroll
            for (int q = Q0; q < Q0 + NQ; ++q) {
                const bool cc = d[q] < 0.5;
                const double inv_c = cc ? 2.0 : inv_d[q];
                const double vc = (kqi * p[q].w) * inv_c;
                const double fc = vc * inv_c;
                if (INNER) {
                    fmag[q] += fc;
                    vv[q] += vc;
                } else {
                    double swc, dswc;
                    switch_q(d[q], P.c_ron, P.c_inv_w, swc, dswc);
                    fmag[q] += fma(fc, swc, -vc * dswc);
                    vv[q] += vc * swc;
                }
            }
        }
    }
    if (GRAV) {
#pragma unroll
        for (int q = Q0; q < Q0 + NQ; ++q) {
            const bool cc = d[q] < 0.5;
            const double inv_g = cc ? 2.0 : inv_d[q];
            const double vg = -(P.gconst * row[min(tj[q], (uint32_t)GS_NSPECIES - 1u)].mm) * inv_g;
            const double fg = vg * inv_g;
            if (INNER) {
                fmag[q] += fg;
                vv[q] += vg;
            } else {
                double swg, dswg;
                switch_q(d[q], P.g_ron, P.g_inv_w, swg, dswg);
                fmag[q] += fma(fg, swg, -vg * dswg);
                vv[q] += vg * swg;
            }
        }
    }
#pragma unroll
    for (int q = Q0; q < Q0 + NQ; ++q) {
        const double fs = (fmag[q] * inv_d[q]) * okf[q];
        fx = fma(fs, dx[q], fx);
        fy = fma(fs, dy[q], fy);
        fz = fma(fs, dz[q], fz);
        vl = fma(vv[q], okf[q], vl);
    }
}

// Fixed-order pairwise sum of N (a power of two) values: ((v0 + v1) + (v2 +
// v3)) + ...
template <int N>
__device__ __forceinline__ double tree_sum(const double* v) {
    if constexpr (N == 1)
        return v[0];
    else
        return tree_sum<N / 2>(v) + tree_sum<N / 2>(v + N / 2);
}

// Full-shell cell-tile pair kernel (the rayon force pass of
// src/systems/forces.rs:228-345 reorganised for a GPU: every atom sums its
// own pairs, so the ordered merge of :330-341 disappears).  Requires >= 3
// cells on every axis so the 27 stencil cells are distinct and each pair is
// seen once per end; smaller boxes use k_pair_brute.
//
// BATCH: many independent universes share the launch.  Their cell grids are
// concatenated into one global cell index space (replica r owns cells
// [cell_off[r], cell_off[r+1])); a workgroup reads its replica's parameters
// from Ptab and skips cells of replicas marked small (those use the
// brute-force batch kernel).
template <bool COUL, bool SHARED, bool EXTRAS, bool SINGLE, bool BATCH = false, int NW = PAIR_WAVES>
__global__ __launch_bounds__(NW * WAVE) void k_pair(KParams Pk, const uint32_t* __restrict__ cstart,
                                                       const double4* __restrict__ pos4,
                                                       const uint8_t* __restrict__ type,
                                                       const PairCoef* __restrict__ coef, double* __restrict__ ax,
                                                       double* __restrict__ ay, double* __restrict__ az,
                                                       double* __restrict__ pe_atom, double* __restrict__ s_abs,
                                                       double* __restrict__ vpart,
                                                       const KParams* __restrict__ Ptab = nullptr,
                                                       const uint16_t* __restrict__ cell_rep = nullptr,
                                                       const uint32_t* __restrict__ cell_off = nullptr) {
    __shared__ double4 tile[TILE_CAP];
    __shared__ uint8_t ttype[TILE_CAP];
    __shared__ PairCoef sc[SINGLE ? 1 : GS_NSPECIES * GS_NSPECIES];
    __shared__ uint32_t seg_base[27];
    __shared__ uint32_t seg_pref[28];
    __shared__ double seg_sh[27][3];
    __shared__ uint16_t queue[NW][QCAP];
    __shared__ double wv[NW];

    const uint32_t cg = blockIdx.x;  // global cell index
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t rep = BATCH ? cell_rep[cg] : 0u;
    const KParams& P = BATCH ? Ptab[rep] : Pk;
    const uint32_t cbase = BATCH ? cell_off[rep] : 0u;
    const uint32_t c = cg - cbase;  // cell index within the universe
    const uint32_t h0 = cstart[cg], nh = cstart[cg + 1] - h0;
    if (nh == 0 || (BATCH && (P.flags & F_SMALL))) {
        if (tid == 0) vpart[cg] = 0.0;
        return;
    }
    if (SINGLE) {
        if (tid == 0) sc[0] = coef[0];
    } else {
        for (int t = tid; t < GS_NSPECIES * GS_NSPECIES; t += NW * WAVE) sc[t] = coef[t];
    }
    if (tid < 27) {
        const uint32_t nx = P.dims[0], ny = P.dims[1], nz = P.dims[2];
        const int cx = (int)(c % nx), cy = (int)((c / nx) % ny), cz = (int)(c / (nx * ny));
        int ix = cx + tid % 3 - 1, iy = cy + (tid / 3) % 3 - 1, iz = cz + tid / 9 - 1;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        // An atom of the wrapped neighbour is moved by one period so that the
        // plain difference is the minimum image (valid for >= 3 cells/axis).
        if (ix < 0) { ix += nx; sx = -P.box[0]; } else if (ix >= (int)nx) { ix -= nx; sx = P.box[0]; }
        if (iy < 0) { iy += ny; sy = -P.box[1]; } else if (iy >= (int)ny) { iy -= ny; sy = P.box[1]; }
        if (iz < 0) { iz += nz; sz = -P.box[2]; } else if (iz >= (int)nz) { iz -= nz; sz = P.box[2]; }
        const uint32_t nc = cbase + (uint32_t)ix + nx * ((uint32_t)iy + ny * (uint32_t)iz);
        const uint32_t b = cstart[nc];
        seg_base[tid] = b;
        seg_pref[tid + 1] = cstart[nc + 1] - b;
        seg_sh[tid][0] = sx;
        seg_sh[tid][1] = sy;
        seg_sh[tid][2] = sz;
    }
    __syncthreads();
    if (tid == 0) {
        seg_pref[0] = 0;
        for (int s = 0; s < 27; ++s) seg_pref[s + 1] += seg_pref[s];
    }
    __syncthreads();
    const uint32_t ncand = seg_pref[27];
    const int nchunks = (int)((ncand + TILE_CAP - 1) / TILE_CAP);
    double vlane = 0.0;

    for (int ch = 0; ch < nchunks; ++ch) {
        const uint32_t c0 = (uint32_t)ch * TILE_CAP;
        const uint32_t len = min((uint32_t)TILE_CAP, ncand - c0);
        __syncthreads();
        for (uint32_t t = tid; t < len; t += NW * WAVE) {
            const uint32_t idx = c0 + t;
            int s = 0;
            while (seg_pref[s + 1] <= idx) ++s;
            const uint32_t j = seg_base[s] + (idx - seg_pref[s]);
            double4 pj = pos4[j];
            pj.x += seg_sh[s][0];
            pj.y += seg_sh[s][1];
            pj.z += seg_sh[s][2];
            tile[t] = pj;
            ttype[t] = type[j];
        }
        __syncthreads();

        for (uint32_t i = w; i < nh; i += NW) {
            const double4 pi = pos4[h0 + i];
            const int ti = type[h0 + i];
            const double kqi = P.ke * pi.w;
            double fx = 0.0, fy = 0.0, fz = 0.0, pe = 0.0, sa = 0.0;
            uint32_t qn = 0, qh = 0;

            auto batch = [&](uint32_t cnt) {
                if ((uint32_t)lane < cnt) {
                    const uint32_t j = queue[w][(qh + lane) & (QCAP - 1)];
                    const double4 pj = tile[j];
                    const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                    const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                    const double inv_d = rsqrt_refined(r2);
                    const double d = r2 * inv_d;
                    const int tj = ttype[j];
                    const PairCoef& pc = SINGLE ? sc[0] : sc[ti * GS_NSPECIES + tj];
                    double fmag, vv;
                    pair_law<COUL, SHARED>(P, pc, d, inv_d, kqi * pj.w, fmag, vv);
                    const double fs = fmag * inv_d;
                    fx = fma(fs, dx, fx);
                    fy = fma(fs, dy, fy);
                    fz = fma(fs, dz, fz);
                    vlane += vv;
                    if (EXTRAS) {
                        pe += vv;
                        sa += fabs(fmag);
                    }
                }
                qh += cnt;
            };

            for (uint32_t base = 0; base < len; base += 64) {
                const uint32_t j = base + lane;
                bool in = false;
                if (j < len) {
                    const double4 pj = tile[j];
                    const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                    const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                    in = (r2 < P.rc2) && (r2 > P.eps2);
                }
                const uint64_t m = __ballot(in);
                if (in) queue[w][(qn + mask_rank(m)) & (QCAP - 1)] = (uint16_t)j;
                qn += (uint32_t)__popcll(m);
                __builtin_amdgcn_wave_barrier();
                if (qn - qh >= 64) batch(64);
            }
            if (qn > qh) batch(qn - qh);

            fx = wave_sum(fx);
            fy = wave_sum(fy);
            fz = wave_sum(fz);
            if (EXTRAS) {
                pe = wave_sum(pe);
                sa = wave_sum(sa);
            }
            if (lane == 0) {
                const uint32_t k = h0 + i;
                const double im = P.inv_mass[ti];
                if (nchunks == 1) {
                    ax[k] = fx * im;
                    ay[k] = fy * im;
                    az[k] = fz * im;
                    if (EXTRAS) {
                        pe_atom[k] = 0.5 * pe;
                        s_abs[k] = sa;
                    }
                } else {
                    // Several LDS chunks: partial sums travel through the
                    // output arrays (this wave owns atom k for every chunk).
                    if (ch > 0) {
                        fx += ax[k];
                        fy += ay[k];
                        fz += az[k];
                        if (EXTRAS) {
                            pe += pe_atom[k];
                            sa += s_abs[k];
                        }
                    }
                    const bool last = ch == nchunks - 1;
                    ax[k] = last ? fx * im : fx;
                    ay[k] = last ? fy * im : fy;
                    az[k] = last ? fz * im : fz;
                    if (EXTRAS) {
                        pe_atom[k] = last ? 0.5 * pe : pe;
                        s_abs[k] = sa;
                    }
                }
            }
        }
    }
    vlane = wave_sum(vlane);
    if (lane == 0) wv[w] = vlane;
    __syncthreads();
    if (tid == 0) vpart[cg] = 0.5 * tree_sum<NW>(wv);
}

// Replica tables of a batched launch (gs_batch): the list-grid parameters
// and first list-grid cell of each replica (a replica on brute-force pairs
// keeps its force grid there, for the sort only), the replica of each
// list-grid cell and of each atom, and each replica's box (w = 1 for a
// replica on Verlet lists, 0 for one on brute-force pairs) with its
// reciprocal.
// box.w is the replica's half skin (0: brute-force pairs), ibox.w its list
// cutoff squared; rflags holds four words per replica (list mode): [0] an
// atom is within reach of skin/2, [1] bin this tick, [2] binned for a build
// next tick, [3] build this tick.
struct BatchTabs {
    const KParams* PL;
    const uint32_t* lcell_off;
    const uint16_t* lcell_rep;
    const uint16_t* arep;
    const double4* box;
    const double4* ibox;
    uint32_t ncell;
    uint32_t* rflags;
};

// Workgroup of a batched launch over atoms: its replica and first atom
// (blocks never straddle two replicas).
struct BlockRef {
    uint32_t rep;
    uint32_t first;
};

constexpr int BRUTE_WAVES = 16;    // home atoms of a brute-force workgroup, one per wave
constexpr int BRUTE_TILE = 1024;   // candidate positions staged in LDS per chunk

// Small-box pair kernel (any axis with < 3 cells): every ordered pair, the
// reference's min_image, and in GS_GRID_REFERENCE mode the multiplicity the
// reference grid gives a pair on an axis with exactly 2 cells (each wrapping
// offset reaches the other cell, src/physics/grid.rs:134-151,275-295): x2 per
// such axis on which the two atoms' cells differ.  One wave per home atom,
// sixteen per workgroup: the universe's positions are staged in LDS chunks
// shared by the sixteen, the lanes sweep a chunk 64 candidates at a time
// with the distance test, a ballot compacts the in-range candidates into a
// per-wave LDS ring, and full 64-lane batches evaluate the law, so the
// expensive part runs on in-range pairs only.  The sweep's test takes the
// image with a product by 1/L (no divide) against a cutoff widened by 1e-12
// relative, and every compacted pair is tested again with the reference's
// min_image, so the pair set is exactly the reference's.  The atom's sums
// are wave reductions (fixed order).  vpart[workgroup] = half the pair
// energy of its home atoms.
// BATCH: workgroup b takes home atoms hb[b].first .. + BRUTE_WAVES - 1 of
// replica hb[b].rep (clipped at its end), with that replica's parameters and
// atom range.
template <bool COUL, bool EXTRAS, bool BATCH>
__global__ __launch_bounds__(BRUTE_WAVES * WAVE) void k_pair_brute(
    KParams Pk, const double4* __restrict__ pos4, const uint8_t* __restrict__ type,
    const PairCoef* __restrict__ coef, double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
    double* __restrict__ pe_atom, double* __restrict__ s_abs, double* __restrict__ vpart,
    const KParams* __restrict__ Ptab, const BlockRef* __restrict__ hb, const uint32_t* __restrict__ atom_off) {
    constexpr int NT = BRUTE_WAVES * WAVE;
    __shared__ PairCoef sc[GS_NSPECIES * GS_NSPECIES];
    __shared__ double4 tile[BRUTE_TILE];
    __shared__ uint8_t ttype[BRUTE_TILE];
    __shared__ uint16_t queue[BRUTE_WAVES][QCAP];
    __shared__ double wv[BRUTE_WAVES];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t rep = 0, lo = 0, hi = (uint32_t)Pk.n, first = blockIdx.x * BRUTE_WAVES;
    if (BATCH) {
        const BlockRef br = hb[blockIdx.x];
        rep = br.rep;
        lo = atom_off[rep];
        hi = atom_off[rep + 1];
        first = br.first;
    }
    const KParams& P = BATCH ? Ptab[rep] : Pk;
    for (int t = tid; t < GS_NSPECIES * GS_NSPECIES; t += NT) sc[t] = coef[t];
    const uint32_t i = first + (uint32_t)w;
    const bool act = i < hi;  // uniform over the wave
    const double4 pi = act ? pos4[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    const int ti = act ? type[i] : 0;
    const double kqi = P.ke * pi.w;
    const bool refm = (P.flags & F_REF_MULT) != 0;
    const uint32_t ki = refm && act ? cell_key(P, pi.x, pi.y, pi.z) : 0u;
    const uint32_t cxi = ki % P.dims[0], cyi = (ki / P.dims[0]) % P.dims[1], czi = ki / (P.dims[0] * P.dims[1]);
    const double iLx = 1.0 / P.box[0], iLy = 1.0 / P.box[1], iLz = 1.0 / P.box[2];
    const double rc2w = P.rc2 * (1.0 + 1e-12);
    double fx = 0.0, fy = 0.0, fz = 0.0, pe = 0.0, sa = 0.0;
    uint32_t qn = 0, qh = 0;
    auto batch = [&](uint32_t cnt) {
        if ((uint32_t)lane < cnt) {
            const uint32_t t = queue[w][(qh + lane) & (QCAP - 1)];
            const double4 pj = tile[t];
            const double dx = min_image1(pi.x - pj.x, P.box[0]);
            const double dy = min_image1(pi.y - pj.y, P.box[1]);
            const double dz = min_image1(pi.z - pj.z, P.box[2]);
            const double r2x = fma(dz, dz, fma(dy, dy, dx * dx));
            const bool ok = r2x < P.rc2 && r2x > P.eps2;
            const double r2 = ok ? r2x : 1.0;
            double mult = ok ? 1.0 : 0.0;
            if (refm) {
                const uint32_t kj = cell_key(P, pj.x, pj.y, pj.z);
                const uint32_t cxj = kj % P.dims[0], cyj = (kj / P.dims[0]) % P.dims[1],
                               czj = kj / (P.dims[0] * P.dims[1]);
                if (P.dims[0] == 2 && cxi != cxj) mult *= 2.0;
                if (P.dims[1] == 2 && cyi != cyj) mult *= 2.0;
                if (P.dims[2] == 2 && czi != czj) mult *= 2.0;
            }
            const double inv_d = rsqrt_refined(r2);
            const double d = r2 * inv_d;
            double fmag, vv;
            pair_law<COUL, false>(P, sc[ti * GS_NSPECIES + ttype[t]], d, inv_d, kqi * pj.w, fmag, vv);
            fmag *= mult;
            vv *= mult;
            const double fs = fmag * inv_d;
            fx = fma(fs, dx, fx);
            fy = fma(fs, dy, fy);
            fz = fma(fs, dz, fz);
            pe += vv;
            if (EXTRAS) sa += fabs(fmag);
        }
        qh += cnt;
    };
    for (uint32_t c0 = lo; c0 < hi; c0 += BRUTE_TILE) {
        const uint32_t len = min((uint32_t)BRUTE_TILE, hi - c0);
        __syncthreads();
        for (uint32_t t = tid; t < len; t += NT) {
            tile[t] = pos4[c0 + t];
            ttype[t] = type[c0 + t];
        }
        __syncthreads();
        if (act) {
            for (uint32_t base = 0; base < len; base += 64) {
                const uint32_t t = base + (uint32_t)lane;
                bool in = false;
                if (t < len && c0 + t != i) {
                    const double4 pj = tile[t];
                    double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                    dx = fma(-P.box[0], rint(dx * iLx), dx);
                    dy = fma(-P.box[1], rint(dy * iLy), dy);
                    dz = fma(-P.box[2], rint(dz * iLz), dz);
                    const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                    in = r2 < rc2w;
                }
                const uint64_t m = __ballot(in);
                if (in) queue[w][(qn + mask_rank(m)) & (QCAP - 1)] = (uint16_t)t;
                qn += (uint32_t)__popcll(m);
                __builtin_amdgcn_wave_barrier();
                if (qn - qh >= 64) batch(64);
            }
            // the queue refers to this chunk's tile: drain it
            if (qn > qh) batch(qn - qh);
        }
    }
    double vl = 0.0;
    if (act) {
        fx = wave_sum(fx);
        fy = wave_sum(fy);
        fz = wave_sum(fz);
        pe = wave_sum(pe);
        if (EXTRAS) sa = wave_sum(sa);
        vl = pe;
        if (lane == 0) {
            const double im = P.inv_mass[ti];
            ax[i] = fx * im;
            ay[i] = fy * im;
            az[i] = fz * im;
            if (EXTRAS) {
                pe_atom[i] = 0.5 * pe;
                s_abs[i] = sa;
            }
        }
    }
    if (lane == 0) wv[w] = vl;
    __syncthreads();
    if (tid == 0) vpart[blockIdx.x] = 0.5 * tree_sum<BRUTE_WAVES>(wv);
}

// Hand-off of per-workgroup partials to the last workgroup of a launch.
// A device-scope fence (__threadfence) compiles to a write-back and an
// invalidation of the whole L2 of the workgroup's XCD; issued by every
// workgroup of a launch it costs microseconds and evicts the data the
// other workgroups of that XCD are still reading.  Instead a partial is
// stored as a device-scope relaxed atomic (written through to the device's
// coherence point), the writing wave waits for its stores to complete
// before the ticket is taken, and the last workgroup reads the partials
// back with device-scope atomic loads.  Ticket increments are device-scope
// atomics, performed at the coherence point.
template <typename T>
__device__ __forceinline__ void publish(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T collect(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Waits until this wave's memory operations, its published stores among
// them, have completed (the signal fences keep the compiler from moving
// memory operations across the wait).
__device__ __forceinline__ void published() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
// Takes the launch's ticket once this wave's published stores are complete;
// the workgroup that draws gridDim.x - 1 is the last.
__device__ __forceinline__ uint32_t take_ticket(uint32_t* counter) {
    published();
    const uint32_t t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    return t;
}

// Deterministic block sum: fixed per-thread strides then a fixed tree.
// Sum of a[0..n) by a workgroup of NT threads (red: NT / 64 doubles of
// LDS); the order of the additions depends on n and NT only.  COH: a[] was
// published by other workgroups of the running launch.
template <int NT = BLOCK, bool COH = false>
__device__ double block_sum_fixed(const double* __restrict__ a, uint32_t n, double* red) {
    static_assert(NT % WAVE == 0, "whole waves");
    const int tid = threadIdx.x;
    auto at = [&](uint32_t k) { return COH ? collect(a + k) : a[k]; };
    // four independent strided accumulators keep four loads in flight
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    uint32_t k = tid;
    for (; k + 3 * NT < n; k += 4 * NT) {
        s0 += at(k);
        s1 += at(k + NT);
        s2 += at(k + 2 * NT);
        s3 += at(k + 3 * NT);
    }
    for (; k < n; k += NT) s0 += at(k);
    double s = (s0 + s1) + (s2 + s3);
    s = wave_sum(s);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    double t;
    if (NT == 4 * WAVE) {
        t = (red[0] + red[1]) + (red[2] + red[3]);
    } else {
        t = red[0];
        for (int w = 1; w < NT / WAVE; ++w) t += red[w];
    }
    __syncthreads();
    return t;
}

// block_sum_fixed of two arrays of the same length at once (each result
// bit-identical to block_sum_fixed of its array): the loads of both share
// one round of latency.  red: 2 NT / 64 doubles of LDS.
template <int NT = BLOCK, bool COH = false>
__device__ void block_sum2_fixed(const double* __restrict__ a, const double* __restrict__ b, uint32_t n,
                                 double* red, double& sa, double& sb) {
    static_assert(NT % WAVE == 0, "whole waves");
    const int tid = threadIdx.x;
    auto at = [&](const double* p, uint32_t k) { return COH ? collect(p + k) : p[k]; };
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;
    uint32_t k = tid;
    for (; k + 3 * NT < n; k += 4 * NT) {
        a0 += at(a, k);
        b0 += at(b, k);
        a1 += at(a, k + NT);
        b1 += at(b, k + NT);
        a2 += at(a, k + 2 * NT);
        b2 += at(b, k + 2 * NT);
        a3 += at(a, k + 3 * NT);
        b3 += at(b, k + 3 * NT);
    }
    for (; k < n; k += NT) {
        a0 += at(a, k);
        b0 += at(b, k);
    }
    double x = (a0 + a1) + (a2 + a3), y = (b0 + b1) + (b2 + b3);
    x = wave_sum(x);
    y = wave_sum(y);
    constexpr int W = NT / WAVE;
    if ((tid & 63) == 0) {
        red[tid >> 6] = x;
        red[W + (tid >> 6)] = y;
    }
    __syncthreads();
    if (W == 4) {
        sa = (red[0] + red[1]) + (red[2] + red[3]);
        sb = (red[4] + red[5]) + (red[6] + red[7]);
    } else {
        sa = red[0];
        sb = red[W];
        for (int w = 1; w < W; ++w) {
            sa += red[w];
            sb += red[W + w];
        }
    }
    __syncthreads();
}

// Berendsen lambda from the kinetic energy after the second kick
// (src/systems/thermostat.rs:47-71, lambda = sqrt(max(0, 1 + (T0/T - 1)/tau)),
// tau in ticks) and the tick's observables (src/systems/aggregate.rs:86-121,
// kinetic energy after the rescale).  The caller applies lambda at the start
// of the next tick or with k_apply_lambda at the end of a call.
struct ThermoArgs {
    int on;
    double target, tau, kconst;
};

// The universe's scalar block: [0] lambda, [1] V, [2] K, [3] V_bond, [4] the
// fault word (u64 bits: 1 a non-finite kinetic energy, 2 potential, 4 bond
// energy, set by close_tick and cleared at the start of each gs_step call).
constexpr int SCAL_WORDS = 5;
constexpr int SCAL_FAULT = 4;

__device__ __forceinline__ unsigned long long* fault_word(double* scal) {
    return reinterpret_cast<unsigned long long*>(scal + SCAL_FAULT);
}

// fault: the fault word to mark when a total is non-finite (NULL: none).  A
// NaN or infinite velocity, acceleration or position reaches K within the
// tick it appears, so the per-tick totals are where it is caught; one thread
// of the tick's last workgroup writes it and ticks are ordered on the stream.
__device__ void close_tick(double K, double V, double Vb, int n, const ThermoArgs& th, double* lam_out,
                           gs_observables* obs, unsigned long long* fault) {
    if (fault && !(isfinite(K) && isfinite(V) && isfinite(Vb)))
        *fault |= (isfinite(K) ? 0ull : 1ull) | (isfinite(V) ? 0ull : 2ull) | (isfinite(Vb) ? 0ull : 4ull);
    double lam = 1.0;
    const double T = (2.0 / 3.0) * (K / (double)n) / th.kconst;
    if (th.on && th.target > 0.0 && n > 0) {
        if (isfinite(T) && T > 0.0) {
            const double l2 = 1.0 + (1.0 / th.tau) * (th.target / T - 1.0);
            lam = sqrt(fmax(l2, 0.0));
        }
    }
    *lam_out = lam;
    if (obs) {
        const double Ka = lam * lam * K;
        obs->kinetic = Ka;
        obs->potential = V;
        obs->bond_energy = Vb;
        obs->temperature = th.kconst > 0.0 ? (2.0 / 3.0) * (Ka / (double)n) / th.kconst : 0.0;
        obs->lambda = lam;
        obs->energy_total = Ka + V + Vb;
    }
}

// Second half kick (src/systems/integrate.rs:38-44) fused with the kinetic
// energy reduction, the Berendsen lambda (src/systems/thermostat.rs:47-71,
// applied at the start of the next tick or by k_apply_lambda) and the tick's
// observables (src/systems/aggregate.rs:86-121).  The last workgroup to
// finish reduces the per-block partials in a fixed order.
template <bool KICK>
__global__ __launch_bounds__(BLOCK) void k_kick_ke(KParams P, double* __restrict__ vx, double* __restrict__ vy,
                                                   double* __restrict__ vz, const double* __restrict__ ax,
                                                   const double* __restrict__ ay, const double* __restrict__ az,
                                                   const uint8_t* __restrict__ type, double* __restrict__ kpart,
                                                   uint32_t* __restrict__ counter, const double* __restrict__ vpart,
                                                   uint32_t nvpart, const double* __restrict__ vbpart, uint32_t nvb,
                                                   double* __restrict__ lam_out,
                                                   gs_observables* __restrict__ obs, int thermostat, double target,
                                                   double tau, double kconst) {
    __shared__ double red[BLOCK / WAVE];
    __shared__ bool last;
    const int tid = threadIdx.x;
    const int k = blockIdx.x * BLOCK + tid;
    double ke = 0.0;
    if (k < P.n) {
        double a = vx[k], b = vy[k], c = vz[k];
        if (KICK) {
            a = a + ax[k] * P.half_dt;
            b = b + ay[k] * P.half_dt;
            c = c + az[k] * P.half_dt;
            vx[k] = a;
            vy[k] = b;
            vz[k] = c;
        }
        ke = 0.5 * P.mass[type[k]] * (a * a + b * b + c * c);
    }
    ke = wave_sum(ke);
    if ((tid & 63) == 0) red[tid >> 6] = ke;
    __syncthreads();
    if (tid == 0) {
        publish(kpart + blockIdx.x, (red[0] + red[1]) + (red[2] + red[3]));
        last = take_ticket(counter) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    const double K = block_sum_fixed<BLOCK, true>(kpart, gridDim.x, red);
    const double V = vpart ? block_sum_fixed(vpart, nvpart, red) : 0.0;
    const double Vb = vbpart ? block_sum_fixed(vbpart, nvb, red) : 0.0;
    if (tid == 0) {
        close_tick(K, V, Vb, P.n, ThermoArgs{thermostat, target, tau, kconst}, lam_out, obs, fault_word(lam_out));
        *counter = 0u;
    }
}

__global__ __launch_bounds__(BLOCK) void k_apply_lambda(int n, double* __restrict__ vx, double* __restrict__ vy,
                                                        double* __restrict__ vz, double* __restrict__ lam) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    const double l = *lam;
    if (k < n) {
        vx[k] = vx[k] * l;
        vy[k] = vy[k] * l;
        vz[k] = vz[k] * l;
    }
}

__global__ void k_set_scalar(double* p, double v) { *p = v; }

__global__ __launch_bounds__(BLOCK) void k_reduce_to(const double* __restrict__ a, uint32_t n, double* __restrict__ out) {
    __shared__ double red[BLOCK / WAVE];
    const double s = block_sum_fixed(a, n, red);
    if (threadIdx.x == 0) *out = s;
}

// Individual systems.
__global__ __launch_bounds__(BLOCK) void k_half_kick(KParams P, double* vx, double* vy, double* vz, const double* ax,
                                                     const double* ay, const double* az) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= P.n) return;
    vx[k] = vx[k] + ax[k] * P.half_dt;
    vy[k] = vy[k] + ay[k] * P.half_dt;
    vz[k] = vz[k] + az[k] * P.half_dt;
}
__global__ __launch_bounds__(BLOCK) void k_drift(KParams P, double4* pos4, const double* vx, const double* vy,
                                                 const double* vz) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= P.n) return;
    double4 p = pos4[k];
    p.x = p.x + vx[k] * P.dt;
    p.y = p.y + vy[k] * P.dt;
    p.z = p.z + vz[k] * P.dt;
    pos4[k] = p;
}
__global__ __launch_bounds__(BLOCK) void k_wrap(KParams P, double4* pos4) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= P.n) return;
    double4 p = pos4[k];
    p.x = rem_euclid(p.x, 2.0 * P.half[0]) - P.half[0];
    p.y = rem_euclid(p.y, 2.0 * P.half[1]) - P.half[1];
    p.z = rem_euclid(p.z, 2.0 * P.half[2]) - P.half[2];
    pos4[k] = p;
}

// Sorted -> entity order for downloads.
__global__ __launch_bounds__(BLOCK) void k_unpermute_state(int n, const uint32_t* __restrict__ id,
                                                           const double4* __restrict__ pos4,
                                                           const double* __restrict__ vx, const double* __restrict__ vy,
                                                           const double* __restrict__ vz, const double* __restrict__ ax,
                                                           const double* __restrict__ ay, const double* __restrict__ az,
                                                           double* __restrict__ out) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n) return;
    const uint32_t e = id[k];
    const double4 p = pos4[k];
    out[0 * (size_t)n + e] = p.x;
    out[1 * (size_t)n + e] = p.y;
    out[2 * (size_t)n + e] = p.z;
    out[3 * (size_t)n + e] = vx[k];
    out[4 * (size_t)n + e] = vy[k];
    out[5 * (size_t)n + e] = vz[k];
    out[6 * (size_t)n + e] = ax[k];
    out[7 * (size_t)n + e] = ay[k];
    out[8 * (size_t)n + e] = az[k];
}
__global__ __launch_bounds__(BLOCK) void k_unpermute2(int n, const uint32_t* __restrict__ id,
                                                      const double* __restrict__ a, const double* __restrict__ b,
                                                      double* __restrict__ out) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n) return;
    const uint32_t e = id[k];
    out[e] = a[k];
    out[(size_t)n + e] = b[k];
}
__global__ __launch_bounds__(BLOCK) void k_cells_entity(KParams P, const uint32_t* __restrict__ id,
                                                        const double4* __restrict__ pos4, uint32_t* __restrict__ out) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= P.n) return;
    const double4 p = pos4[k];
    out[id[k]] = cell_key(P, p.x, p.y, p.z);
}


// ---- Verlet-list path ---------------------------------------------------------
//
// Every atom carries the list of atoms within rc + skin, built from the
// 27-cell stencil of a grid of edge >= rc + skin.  A list stays valid while
// no atom has moved more than skin/2 from its position at the build, so
// every pair within rc is in some list: the pair set, and hence the forces,
// are the reference's; only the evaluation order differs.
//
// The device decides, tick by tick, whether the lists must be rebuilt: the
// integrate kernel measures every atom's displacement from its build
// position and raises flags[0]; the bin and build kernels that follow do
// their work only when flags[0] is set (otherwise they exit at once), and
// the pair kernel of the tick clears the flag.  No host round trip decides
// a rebuild.  A rebuild bins the atoms into fixed-capacity cell buckets
// (one atomic per distinct cell in a wave, no scan), and the build orders
// each bucket by atom index in LDS, so the lists do not depend on the order
// in which the atomics handed out the slots.
//
// flags, reset at every call: [0] rebuild this tick, [1] a list overflowed
//        its capacity, [2] a stencil held more candidates than the build
//        orders in LDS (LDS_IDX_CAP, _LOCAL, _BATCH),
//        [3] rebuilds done by the device, [4] a tick needed a rebuild that
//        its binning had not been requested for (predictive binning),
//        [5] last-workgroup counter of the pair kernel, [6] a cell bucket
//        overflowed its capacity, [7] a lane of the pair kernel screened
//        more bound-pair candidates than its LDS column holds (fused bond
//        search);
//        kept across calls: [8] some atom is within reach of skin/2 next
//        tick, [9] bin (and rebuild) this tick, [10] this tick binned for
//        a build one tick behind (k_list_step), [11] build this tick from
//        the previous tick's binning.
//
// Predictive binning: the integrate kernel of a tick bins the atoms and
// forces a rebuild when the previous tick found an atom whose displacement
// from its build position plus 1.5 times its own displacement of that tick
// reaches skin/2 (the pair kernel's last workgroup moves [8] to [9]),
// so no separate binning launch runs on the other ticks.  An atom beyond skin/2 on a tick without
// binning raises [4] and the host replays the call with a binning launch
// every tick.

constexpr uint32_t NBR_IDX_MASK = 0x0FFFFFFFu;  // entry = (species << 28) | sorted index
// Atoms the Verlet-list pass holds: below 2^27 an entry shifted left by 5 is
// its atom's byte offset in the 32-byte position array (the species bits
// and bit 27 leave the word), which the gathers use directly.
constexpr uint32_t LIST_ATOMS_MAX = 1u << 27;
// A candidate as the list build stages it in LDS: its position relative to
// the centre of the home cell in FP32 (a Verlet list may hold a superset of
// the pairs within rc + skin, so the build's distance test need not be FP64)
// and its list entry; one 16-byte LDS read per candidate test.
struct alignas(16) BuildCand {
    float x, y, z;
    uint32_t e;
};
constexpr int LIST_FLAGS = 16;                  // words of the list flags
constexpr int LIST_FLAGS_RESET = 8;             // the first ones are reset at every call
constexpr int LDS_IDX_CAP = 2048;               // stencil candidates held (and ordered) in LDS
constexpr int LDS_IDX_CAP_LOCAL = 2048;         // the same for the cell-staged lists (11-bit positions)
constexpr int LDS_IDX_CAP_BATCH = 4096;         // the same for a batch's lists (a build launch of its own)
constexpr int MAX_LOGT = 4;                     // up to 16 lanes per atom
constexpr uint32_t HOME_CAP = 256;              // home atoms whose list lengths the build keeps in LDS
constexpr uint32_t LIST_CAP_MAX = 0xFFFFu;       // entries per atom a list length word holds
constexpr int SPLIT_CAP = 512;                  // longest list the build writes split
// List length word nlen[i].  An unsplit list: the entries (bits 0-15).  A
// split list (bit 31 set): the entries (bits 0-9), the inner entries (bits
// 10-19) and, among those, the bind entries (bits 20-29).  A split list holds
// first the entries whose build distance lies within the bind radius (the
// bond cutoff + skin: the only pairs that can be bound while the list is in
// use), then the other entries within the inner radius (the smallest
// switching radius - skin: inside every switching radius while the list is
// in use), then the rest, each class in candidate order.
constexpr uint32_t NLEN_SPLIT = 0x80000000u;
static_assert(SPLIT_CAP < 1024, "a split list's counts fit ten bits");
__host__ __device__ __forceinline__ uint32_t nlen_entries(uint32_t w) {
    return (w & NLEN_SPLIT) ? (w & 0x3FFu) : (w & 0xFFFFu);
}
__host__ __device__ __forceinline__ uint32_t nlen_inner(uint32_t w) {
    return (w & NLEN_SPLIT) ? ((w >> 10) & 0x3FFu) : 0u;
}
__host__ __device__ __forceinline__ uint32_t nlen_bind(uint32_t w) {
    return (w & NLEN_SPLIT) ? ((w >> 20) & 0x3FFu) : 0u;
}

// Cell-staged list pass (pair_mode 2): entries are 16 bits,
// (species << CELL_IDX_BITS) | position in the home cell's ordered
// candidates; home atom h of a cell keeps its list in row h of the cell's
// region (cap entries per row).
constexpr int CELL_THREADS = 384;    // workgroup of the cell-staged pair kernel
constexpr int STAGE_CAP = 1280;      // candidate positions staged in LDS per chunk
constexpr uint32_t CELL_IDX_BITS = 11;
constexpr uint32_t CELL_IDX_MASK = (1u << CELL_IDX_BITS) - 1u;
static_assert(LDS_IDX_CAP_LOCAL <= (1 << CELL_IDX_BITS), "a candidate position fits the entry");
static_assert(GS_NSPECIES <= (1 << (16 - CELL_IDX_BITS)), "a species fits the entry");

// Streaming 16-byte load with a non-temporal hint (global_load ... nt).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 load_nt(const uint4* p) {
    const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}

// Entry k of atom i with T lanes per atom: lane s = k mod T owns entries
// s, s+T, s+2T, ...; its m-th entry (m = k / T) sits in group m/4, slot
// m mod 4 of an [groups][N][T][4] array, so a lane reads four entries with
// one 16-byte load and the lanes of consecutive atoms read consecutive
// 16-byte words.
__device__ __forceinline__ size_t nbr_slot(uint32_t k, uint32_t i, uint32_t N, uint32_t logT) {
    const uint32_t m = k >> logT, s = k & ((1u << logT) - 1u);
    return ((((size_t)(m >> 2) * N + i) << logT) + s) * 4u + (m & 3u);
}

// Slots of the atoms of a wave in the list-grid cell buckets: each run of
// equal cells among the wave's lanes takes its slots with one atomic, and
// all the runs of the wave issue theirs at once (one round trip per wave).
// kk of a lane past the end is ignored (valid lanes are a prefix of the
// wave).  A bucket that overflows its capacity raises flags[6].
__device__ __forceinline__ void bin_wave(uint32_t kk, bool valid, uint32_t k, uint32_t* __restrict__ bcount,
                                         uint32_t* __restrict__ bucket, uint32_t bcap, uint32_t* __restrict__ flags) {
    const int lane = threadIdx.x & 63;
    if (!valid) kk = 0xFFFFFFFFu;
    const uint32_t prev = __shfl_up(kk, 1, WAVE);
    const bool head = valid && (lane == 0 || prev != kk);
    const uint64_t heads = __ballot(head);
    const int nvalid = __popcll(__ballot(valid));
    const uint64_t upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
    const uint64_t hb = heads & upto;
    const int rs = hb ? 63 - __builtin_clzll(hb) : 0;  // head of this lane's run
    const uint64_t ha = heads & ~upto;
    const int re = ha ? __builtin_ctzll(ha) : nvalid;  // one past this lane's run
    uint32_t base = 0;
    if (head) base = atomicAdd(&bcount[kk], (uint32_t)(re - rs));
    base = __shfl(base, rs, WAVE);
    if (valid) {
        const uint32_t slot = base + (uint32_t)(lane - rs);
        if (slot < bcap)
            bucket[(size_t)kk * bcap + slot] = k;
        else
            atomicOr(&flags[6], 1u);
    }
}

// List-mode tick start: thermostat scale, half kick, drift, wrap (the same
// expressions as k_integrate_key) and the displacement from the build
// position, tested against skin/2.  PREBIN (predictive binning, see the
// flags above): bin the new positions and force a rebuild when flags[9]
// asks for it, raise flags[4] for an atom beyond skin/2 otherwise, raise
// flags[8] for an atom within reach of skin/2 next tick, and record the
// largest one-tick displacement.  Without PREBIN: raise flags[0] for an atom
// beyond skin/2 (k_list_bin then bins) and clear the bucket counts.
template <bool PREBIN>
__global__ __launch_bounds__(BLOCK) void k_list_integrate(
    KParams P, KParams PL, double4* __restrict__ pos4, double* __restrict__ vx, double* __restrict__ vy,
    double* __restrict__ vz, const double* __restrict__ ax, const double* __restrict__ ay,
    const double* __restrict__ az, const double* __restrict__ lam_ptr, const double4* __restrict__ xref,
    double half_skin, uint32_t* __restrict__ bcount, uint32_t* __restrict__ bucket, uint32_t bcap,
    uint32_t* __restrict__ flags) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int k = blockIdx.x * BLOCK + tid;
    if (!PREBIN)
        for (uint32_t c = (uint32_t)k; c < PL.ncell; c += gridDim.x * BLOCK) bcount[c] = 0u;
    const bool binreq = PREBIN && flags[9] != 0u;
    const bool valid = k < P.n;
    bool far = false, warn = false;
    uint32_t kk = 0;
    if (valid) {
        double4 p = pos4[k];
        const double lam = *lam_ptr;
        double a = vx[k] * lam, b = vy[k] * lam, c = vz[k] * lam;
        a = a + ax[k] * P.half_dt;
        b = b + ay[k] * P.half_dt;
        c = c + az[k] * P.half_dt;
        vx[k] = a;
        vy[k] = b;
        vz[k] = c;
        const double sx = a * P.dt, sy = b * P.dt, sz = c * P.dt;
        p.x = p.x + sx;
        p.y = p.y + sy;
        p.z = p.z + sz;
        if (P.flags & F_WRAP) {
            p.x = rem_euclid(p.x, 2.0 * P.half[0]) - P.half[0];
            p.y = rem_euclid(p.y, 2.0 * P.half[1]) - P.half[1];
            p.z = rem_euclid(p.z, 2.0 * P.half[2]) - P.half[2];
        }
        pos4[k] = p;
        // The wrap (rem_euclid(v, 2h) - h) translates the whole universe
        // rigidly by h on every axis each tick, so the displacement since
        // the build is taken modulo h: that removes the accumulated
        // translation exactly and keeps any motion below h/2.
        const double4 r = xref[k];
        double dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
        dx = fma(-P.half[0], rint(dx * P.inv_half[0]), dx);
        dy = fma(-P.half[1], rint(dy * P.inv_half[1]), dy);
        dz = fma(-P.half[2], rint(dz * P.inv_half[2]), dz);
        const double d2 = fma(dz, dz, fma(dy, dy, dx * dx));
        far = !(d2 <= half_skin * half_skin);  // NaN counts as far
        if (PREBIN) {
            // a tick that rebuilds measures the next tick's displacement
            // from the new build position, this one
            const double step = sqrt(fma(sz, sz, fma(sy, sy, sx * sx)));
            warn = !(fma(1.5, step, binreq ? 0.0 : sqrt(d2)) < half_skin);
            if (binreq) kk = cell_key(PL, p.x, p.y, p.z);
        }
    }
    if (!PREBIN) {
        if (__ballot(far) != 0 && lane == 0) atomicOr(&flags[0], 1u);
        return;
    }
    if (binreq) {
        bin_wave(kk, valid, (uint32_t)k, bcount, bucket, bcap, flags);
        if (k == 0) flags[0] = 1u;
    } else if (__ballot(far) != 0 && lane == 0) {
        atomicOr(&flags[4], 1u);
    }
    if (__ballot(warn) != 0 && lane == 0) atomicOr(&flags[8], 1u);
}

// Rebuild binning without predictive binning (only when flags[0] is set):
// the list-grid cell of every atom and its slot in the cell's bucket.
__global__ __launch_bounds__(BLOCK) void k_list_bin(KParams PL, const double4* __restrict__ pos4,
                                                    uint32_t* __restrict__ bcount, uint32_t* __restrict__ bucket,
                                                    uint32_t bcap, uint32_t* __restrict__ flags) {
    if (flags[0] == 0u) return;
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    const bool valid = k < PL.n;
    uint32_t kk = 0;
    if (valid) {
        const double4 p = pos4[k];
        kk = cell_key(PL, p.x, p.y, p.z);
    }
    bin_wave(kk, valid, (uint32_t)k, bcount, bucket, bcap, flags);
}

// Buckets of a freshly cell-sorted state: bucket c holds cstart[c] ..
// cstart[c+1]-1 in order.
__global__ __launch_bounds__(BLOCK) void k_list_fill_buckets(uint32_t ncell, const uint32_t* __restrict__ cstart,
                                                             uint32_t* __restrict__ bcount,
                                                             uint32_t* __restrict__ bucket, uint32_t bcap,
                                                             uint32_t* __restrict__ flags) {
    const uint32_t c = blockIdx.x * (BLOCK / WAVE) + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (c >= ncell) return;
    const uint32_t s = cstart[c], m = cstart[c + 1] - s;
    if (lane == 0) {
        bcount[c] = m;
        if (m > bcap) atomicOr(&flags[6], 1u);
    }
    for (uint32_t e = lane; e < m && e < bcap; e += 64) bucket[(size_t)c * bcap + e] = s + e;
}

__global__ void k_flag_set(uint32_t* flags, int idx, uint32_t v) { flags[idx] = v; }

// One workgroup per home cell (only when flags[0] is set): the 27 stencil
// buckets are staged in LDS and each is ordered by atom index (rank by
// counting); the candidates' positions, shifted by the image offset of
// their bucket, are then staged in LDS chunks and every wave compacts, for
// one home atom at a time, the candidates within rc + skin with a ballot.
// The home atoms' positions are recorded as the build positions of the
// displacement test.
//
// LOCAL (cell-staged pair pass): the ordered candidates are stored per cell
// (cand_store, cand_n) for the pair workgroup to stage, entries hold the
// candidate's position in that order, and each atom's entries are contiguous
// (nbr[i * cap + k]); otherwise entries hold the atom index in the
// lane-grouped layout of nbr_slot.
// The cells b0, b0 + stride, ... of a list build, by a workgroup of NT
// threads; pos4 holds the positions the lists are built from.
// BATCH (gs_batch): the cells are the list-grid cells of all replicas; a
// cell's stencil lies in its own replica's grid (bt), and cells of replicas
// on brute-force pairs have empty buckets.
// SPLIT (the Verlet lists of one universe): with rin2 > 0, a cell whose
// stencil fits one LDS chunk writes each list split (nlen_inner, nlen_bind):
// its accepted candidates are staged in the wave's LDS row and, after the
// sweep, classed and written out class by class: within the inner radius
// and the species pair's bind radius (rbt: its square widened by the FP32
// margin, per species pair), within the inner radius, the rest.  A list
// longer than SPLIT_CAP or a cell of several chunks is written unsplit.
template <bool LOCAL, int NT, bool BATCH = false, bool SPLIT = false>
__device__ __forceinline__ void build_cells(
    uint32_t b0, uint32_t stride, KParams PL, const uint32_t* __restrict__ bcount,
    const uint32_t* __restrict__ bucket, uint32_t bcap, const double4* __restrict__ pos4,
    const uint8_t* __restrict__ type, uint32_t* __restrict__ nbr, uint32_t* __restrict__ nlen,
    double4* __restrict__ xref, uint32_t N, uint32_t logT, uint32_t cap, double cut2, uint32_t* __restrict__ flags,
    uint32_t* __restrict__ cand_store, uint4* __restrict__ cinfo, uint2* __restrict__ hrec,
    uint16_t* __restrict__ nbr16, size_t region, const BatchTabs& bt = BatchTabs{}, double rin2 = 0.0,
    const float* __restrict__ rbt = nullptr) {
    constexpr int NWV = NT / WAVE;
    // SPLIT: the outer candidates of the wave's current home atom
    __shared__ uint32_t outb[SPLIT ? NWV : 1][SPLIT ? SPLIT_CAP : 1];
    __shared__ uint8_t outc[SPLIT ? NWV : 1][SPLIT ? SPLIT_CAP : 1];  // their classes
    __shared__ float s_rbt[SPLIT ? GS_NSPECIES * GS_NSPECIES : 1];    // bind radius^2 per species pair
    // the candidates' indices are staged, unordered, in the memory of the
    // position tile, which is filled only after they have been ordered
    __shared__ BuildCand tile[BUILD_TILE];
    __shared__ double org[3];  // centre of the home cell
    constexpr int IDX_CAP = LOCAL ? LDS_IDX_CAP_LOCAL : (BATCH ? LDS_IDX_CAP_BATCH : LDS_IDX_CAP);
    __shared__ uint32_t ord[IDX_CAP];
    // running list lengths between chunks: a cell of several chunks is built
    // unsplit, so SPLIT keeps them in the staging rows
    static_assert(!SPLIT || NWV * SPLIT_CAP >= (int)HOME_CAP, "the staging rows hold the list lengths");
    __shared__ uint32_t hcnt_own[SPLIT ? 1 : HOME_CAP];
    uint32_t* const hcnt = SPLIT ? &outb[0][0] : hcnt_own;
    __shared__ uint32_t seg_base[27];
    __shared__ uint32_t seg_pref[28];
    __shared__ double seg_sh[27][3];
    static_assert(IDX_CAP * sizeof(uint32_t) <= BUILD_TILE * sizeof(BuildCand), "index staging fits the tile");
    uint32_t* cand = reinterpret_cast<uint32_t*>(tile);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (b0 == 0 && tid == 0) flags[3] += 1u;  // rebuilds done on the device
    if (SPLIT)
        for (int t = tid; t < GS_NSPECIES * GS_NSPECIES; t += NT) s_rbt[t] = rbt ? rbt[t] : 0.0f;
    const uint32_t ncell = BATCH ? bt.ncell : PL.ncell;
    for (uint32_t c = b0; c < ncell; c += stride) {
    // BATCH: a cell of a replica that does not build now is skipped (a
    // full build, flags[0], takes every cell); cut2 is its replica's
    double cutc = cut2;
    if (BATCH) {
        const uint32_t rep = bt.lcell_rep[c];
        if (flags[0] == 0u && bt.rflags[4 * rep + 3] == 0u) continue;
        cutc = bt.ibox[rep].w;
    }
    // FP32 test with a relative margin of 1e-5 on r^2: every pair within
    // rc + skin passes (the FP32 error of r^2 on cell-relative coordinates is
    // below 2e-6 relative), a pair up to 5e-6 (rc + skin) farther may pass too;
    // it lies beyond rc, where the switched pair law is exactly zero
    const float cut2f = (float)(cutc * (1.0 + 1e-5));
    // the inner class with the same error bound the other way: an entry
    // classed inner lies within the inner radius; the bind class takes every
    // candidate that may lie within the bind radius
    const float rin2f = (float)(rin2 * (1.0 - 1e-5));
    __syncthreads();  // the previous cell's LDS readers are done
    const uint32_t nh = min(bcount[c], bcap);
    if (nh == 0) {
        if (LOCAL && tid == 0) cinfo[c] = make_uint4(0u, 0u, 0u, 0u);
        continue;
    }
    const uint32_t* home = bucket + (size_t)c * bcap;
    if (tid < 27) {
        const uint32_t rep = BATCH ? bt.lcell_rep[c] : 0u;
        const KParams& Q = BATCH ? bt.PL[rep] : PL;
        const uint32_t cbase = BATCH ? bt.lcell_off[rep] : 0u;
        const uint32_t lc = c - cbase;
        const uint32_t nx = Q.dims[0], ny = Q.dims[1], nz = Q.dims[2];
        const int cx = (int)(lc % nx), cy = (int)((lc / nx) % ny), cz = (int)(lc / (nx * ny));
        int ix = cx + tid % 3 - 1, iy = cy + (tid / 3) % 3 - 1, iz = cz + tid / 9 - 1;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        if (ix < 0) { ix += nx; sx = -Q.box[0]; } else if (ix >= (int)nx) { ix -= nx; sx = Q.box[0]; }
        if (iy < 0) { iy += ny; sy = -Q.box[1]; } else if (iy >= (int)ny) { iy -= ny; sy = Q.box[1]; }
        if (iz < 0) { iz += nz; sz = -Q.box[2]; } else if (iz >= (int)nz) { iz -= nz; sz = Q.box[2]; }
        const uint32_t nc = cbase + (uint32_t)ix + nx * ((uint32_t)iy + ny * (uint32_t)iz);
        seg_base[tid] = nc * bcap;
        seg_pref[tid + 1] = min(bcount[nc], bcap);
        seg_sh[tid][0] = sx;
        seg_sh[tid][1] = sy;
        seg_sh[tid][2] = sz;
        if (tid == 13) {
            org[0] = -Q.half[0] + ((double)cx + 0.5) * Q.cell[0];
            org[1] = -Q.half[1] + ((double)cy + 0.5) * Q.cell[1];
            org[2] = -Q.half[2] + ((double)cz + 0.5) * Q.cell[2];
        }
    }
    __syncthreads();
    if (tid == 0) {
        seg_pref[0] = 0;
        for (int s2 = 0; s2 < 27; ++s2) seg_pref[s2 + 1] += seg_pref[s2];
    }
    __syncthreads();
    const uint32_t ncand = seg_pref[27];
    if (ncand > (uint32_t)IDX_CAP) {
        if (tid == 0) atomicOr(&flags[2], 1u);
        continue;
    }
    // stencil segment of candidate t < ncand: the last s with seg_pref[s] <= t
    auto seg_of = [&](uint32_t t) {
        int lo = 0, hi = 27;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (seg_pref[mid] <= t)
                lo = mid;
            else
                hi = mid;
        }
        return lo;
    };
    for (uint32_t t = tid; t < ncand; t += NT) {
        const int sg = seg_of(t);
        cand[t] = bucket[seg_base[sg] + (t - seg_pref[sg])];
    }
    __syncthreads();
    for (uint32_t t = tid; t < ncand; t += NT) {
        const int sg = seg_of(t);
        const uint32_t v = cand[t], lo = seg_pref[sg], hi = seg_pref[sg + 1];
        // rank of v in its segment, eight independent reads per step
        uint32_t r = 0, f = lo;
        for (; f + 8 <= hi; f += 8) {
            uint32_t x[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) x[q] = cand[f + q];
#pragma unroll
            for (int q = 0; q < 8; ++q) r += x[q] < v ? 1u : 0u;
        }
        for (; f < hi; ++f) r += cand[f] < v ? 1u : 0u;
        ord[lo + r] = v;
    }
    for (uint32_t i = tid; i < nh; i += NT) {
        const uint32_t gi = home[i];
        xref[gi] = pos4[gi];
    }
    // LOCAL: the home atoms, in index order, are the centre segment (13) of
    // the ordered candidates
    const uint32_t hs = seg_pref[13];
    uint16_t* reg = LOCAL ? nbr16 + (size_t)c * region : nullptr;
    if (LOCAL) {
        __syncthreads();
        for (uint32_t t = tid; t < ncand; t += NT) cand_store[(size_t)c * LDS_IDX_CAP_LOCAL + t] = ord[t];
        if (tid == 0) cinfo[c] = make_uint4(ncand, hs, nh, 0u);
    }
    const int nchunks = (int)((ncand + BUILD_TILE - 1) / BUILD_TILE);
    const bool split = SPLIT && rin2 > 0.0 && nchunks == 1 && cap <= (uint32_t)SPLIT_CAP;
    for (int ch = 0; ch < nchunks; ++ch) {
        const uint32_t c0 = (uint32_t)ch * BUILD_TILE;
        const uint32_t len = min((uint32_t)BUILD_TILE, ncand - c0);
        __syncthreads();
        for (uint32_t t = tid; t < len; t += NT) {
            const uint32_t idx = c0 + t;
            const int sg = seg_of(idx);
            const uint32_t j = ord[idx];
            const double4 pj = pos4[j];
            BuildCand q;
            q.x = (float)((pj.x + seg_sh[sg][0]) - org[0]);
            q.y = (float)((pj.y + seg_sh[sg][1]) - org[1]);
            q.z = (float)((pj.z + seg_sh[sg][2]) - org[2]);
            q.e = ((uint32_t)type[j] << 28) | j;
            tile[t] = q;
        }
        __syncthreads();
        // the wave's next home atom is fetched while this one is swept, and
        // a list's running length stays in LDS between chunks (in nlen when
        // the cell has more than HOME_CAP home atoms)
        // (home atom i is the i-th of the centre segment, in index order;
        // its position comes from the tile when the segment lies in this
        // chunk, which it does whenever the stencil fits one chunk)
        const bool home_in_tile = hs >= c0 && hs + nh <= c0 + len;
        auto home_pos = [&](uint32_t ii) {
            if (home_in_tile) {
                const BuildCand q = tile[hs + ii - c0];
                return make_float3(q.x, q.y, q.z);
            }
            const double4 p = pos4[ord[hs + ii]];  // centre cell: no image shift
            return make_float3((float)(p.x - org[0]), (float)(p.y - org[1]), (float)(p.z - org[2]));
        };
        float3 pi_n = (uint32_t)w < nh ? home_pos((uint32_t)w) : make_float3(0.f, 0.f, 0.f);
        for (uint32_t i = w; i < nh; i += NWV) {
            const uint32_t gi = ord[hs + i];
            const float3 pi = pi_n;
            if (i + NWV < nh) pi_n = home_pos(i + NWV);
            uint32_t cnt = ch == 0 ? 0u : (nh <= HOME_CAP ? hcnt[i] : nlen[gi]);
            // the bind radii^2 of the home atom's species (split lists)
            const float* rbrow =
                s_rbt + (SPLIT ? min((uint32_t)type[gi], (uint32_t)GS_NSPECIES - 1u) * GS_NSPECIES : 0u);
            // 256 candidates per step, four per lane (independent distance
            // tests); the ballots are taken in candidate order, so the list
            // (each class of a split list) is ordered by candidate position
            // as with one per lane.  Split: every accepted candidate is
            // staged in the wave's LDS row (its tile position); after the
            // sweep each gets its class (0 within the inner radius and the
            // species pair's bind radius, 1 other inner, 2 the rest) and the
            // row is written out class by class.
            for (uint32_t base = 0; base < len; base += 4 * WAVE) {
                bool in[4];
                uint32_t e[4], jt[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t j = base + (uint32_t)(q * WAVE + lane);
                    const uint32_t jc = j < len ? j : len - 1u;
                    const BuildCand cj = tile[jc];
                    const uint32_t jg = cj.e & NBR_IDX_MASK;
                    e[q] = LOCAL ? ((cj.e >> 28) << CELL_IDX_BITS) | (c0 + jc) : cj.e;
                    jt[q] = jc;
                    const float dx = pi.x - cj.x, dy = pi.y - cj.y, dz = pi.z - cj.z;
                    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    in[q] = j < len && r2 < cut2f && jg != gi;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint64_t m = __ballot(in[q]);
                    if (in[q]) {
                        const uint32_t k = cnt + mask_rank(m);
                        if (SPLIT && split) {
                            // (the tile position; the flush classes it)
                            if (k < (uint32_t)SPLIT_CAP) outb[w][k] = jt[q];
                        } else if (k < cap) {
                            if (LOCAL)
                                reg[(size_t)i * cap + k] = (uint16_t)e[q];
                            else
                                nbr[nbr_slot(k, gi, N, logT)] = e[q];
                        }
                    }
                    cnt += (uint32_t)__popcll(m);
                }
            }
            uint32_t nA = 0u, nB = 0u;
            // more accepted candidates than the LDS row holds: the list
            // exceeds its capacity (cap <= SPLIT_CAP) and is written empty,
            // the call being replayed
            const bool lost = SPLIT && split && cnt > (uint32_t)SPLIT_CAP;
            if (SPLIT && split && !lost) {
                // the staged row, class by class (the wave's own LDS row: its
                // writes above precede these reads in the wave's LDS order)
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                // each staged candidate's class, from its tile entry with the
                // sweep's FP32 distance, and its list entry
                for (uint32_t k0 = 0; k0 < cnt; k0 += WAVE) {
                    const uint32_t k = k0 + (uint32_t)lane;
                    if (k < cnt) {
                        const BuildCand cj = tile[outb[w][k]];
                        const float dx = pi.x - cj.x, dy = pi.y - cj.y, dz = pi.z - cj.z;
                        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        const uint32_t sj = min(cj.e >> 28, (uint32_t)GS_NSPECIES - 1u);
                        outc[w][k] = (uint8_t)(r2 < rin2f ? (r2 < rbrow[sj] ? 0u : 1u) : 2u);
                        outb[w][k] = cj.e;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                uint32_t at = 0u;
                for (uint32_t c = 0; c < 3u; ++c) {
                    for (uint32_t k0 = 0; k0 < cnt; k0 += WAVE) {
                        const uint32_t k = k0 + (uint32_t)lane;
                        const bool mine = k < cnt && outc[w][k] == c;
                        const uint64_t mc = __ballot(mine);
                        if (mine) {
                            const uint32_t pos = at + mask_rank(mc);
                            if (pos < cap) nbr[nbr_slot(pos, gi, N, logT)] = outb[w][k];
                        }
                        at += (uint32_t)__popcll(mc);
                    }
                    if (c == 0u) nB = min(at, cap);
                    if (c == 1u) nA = min(at, cap);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            if (!LOCAL && ch == nchunks - 1) {
                // each lane's groups are padded with the atom's own entry
                // (r^2 = 0: no pair) to the end of its last group and one
                // group beyond it, which the pair pass reads past the end of
                // the lane's entries (entries [len, 4T (ceil(len / 4T) + 1)))
                const uint32_t lenf = (cnt > cap || lost) ? (lost ? 0u : cap) : cnt;
                const uint32_t T4 = 4u << logT;
                const uint32_t kend = ((lenf + T4 - 1u) / T4 + 1u) * T4;
                const uint32_t own = ((uint32_t)type[gi] << 28) | gi;
                for (uint32_t k = lenf + (uint32_t)lane; k < kend; k += WAVE) nbr[nbr_slot(k, gi, N, logT)] = own;
            }
            if (lane == 0) {
                if (ch == nchunks - 1 && (cnt > cap || lost)) {
                    atomicOr(&flags[1], 1u);
                    cnt = lost ? 0u : cap;
                }
                if (nchunks > 1 && nh <= HOME_CAP) hcnt[i] = cnt;
                if (ch == nchunks - 1 || nh > HOME_CAP)
                    nlen[gi] = (SPLIT && split) ? (cnt | (nA << 10) | (nB << 20) | NLEN_SPLIT) : cnt;
                // LOCAL: home record ((species << 28) | atom, list length)
                if (LOCAL && ch == nchunks - 1) hrec[(size_t)c * bcap + i] = make_uint2(((uint32_t)type[gi] << 28) | gi, cnt);
            }
        }
    }
    }  // cells
}

template <bool LOCAL>
__global__ __launch_bounds__(BUILD_THREADS) void k_list_build(
    KParams PL, const uint32_t* __restrict__ bcount, const uint32_t* __restrict__ bucket, uint32_t bcap,
    const double4* __restrict__ pos4, const uint8_t* __restrict__ type, uint32_t* __restrict__ nbr,
    uint32_t* __restrict__ nlen, double4* __restrict__ xref, uint32_t N, uint32_t logT, uint32_t cap, double cut2,
    uint32_t* __restrict__ flags, uint32_t* __restrict__ cand_store, uint4* __restrict__ cinfo,
    uint2* __restrict__ hrec, uint16_t* __restrict__ nbr16, size_t region, double rin2,
    const float* __restrict__ rbt) {
    if (flags[0] == 0u) return;
    build_cells<LOCAL, BUILD_THREADS, false, !LOCAL>(blockIdx.x, gridDim.x, PL, bcount, bucket, bcap, pos4, type,
                                                      nbr, nlen, xref, N, logT, cap, cut2, flags, cand_store, cinfo,
                                                      hrec, nbr16, region, BatchTabs{}, rin2, rbt);
}

// Start of a list tick with the build one tick behind the binning
// (predictive binning, Verlet list): workgroups [0, natom) run the tick's
// kick, drift, wrap and displacement test for their atoms as
// k_list_integrate<true>, the others build the lists when the previous tick
// binned (flags[11]).  That build takes the positions the previous tick
// binned (xbin), so the lists it makes are valid from this tick on, their
// build positions being one tick old: this tick measures displacements
// from xbin and nothing in this launch reads what the build writes.  A
// binning is requested two ticks ahead (own displacement plus three times
// this tick's), so the lists in use on the binning tick are still valid.
// A tick that builds does not bin again.
__global__ __launch_bounds__(BLOCK) void k_list_step(
    KParams P, KParams PL, double4* __restrict__ pos4, double* __restrict__ vx, double* __restrict__ vy,
    double* __restrict__ vz, const double* __restrict__ ax, const double* __restrict__ ay,
    const double* __restrict__ az, const double* __restrict__ lam_ptr, double4* __restrict__ xref,
    double4* __restrict__ xbin, double half_skin, uint32_t* __restrict__ bcount, uint32_t* __restrict__ bucket,
    uint32_t bcap, uint32_t* __restrict__ flags, uint32_t natom, const uint8_t* __restrict__ type,
    uint32_t* __restrict__ nbr, uint32_t* __restrict__ nlen, uint32_t N, uint32_t logT, uint32_t cap, double cut2,
    double rin2, const float* __restrict__ rbt) {
    const bool build_now = flags[11] != 0u;
    if (blockIdx.x >= natom) {
        if (!build_now) return;
        build_cells<false, BLOCK, false, true>(blockIdx.x - natom, gridDim.x - natom, PL, bcount, bucket, bcap, xbin,
                                               type, nbr, nlen, xref, N, logT, cap, cut2, flags, nullptr, nullptr,
                                               nullptr, nullptr, 0, BatchTabs{}, rin2, rbt);
        return;
    }
    const int tid = threadIdx.x, lane = tid & 63;
    const int k = blockIdx.x * BLOCK + tid;
    const bool binreq = !build_now && flags[9] != 0u;
    const bool valid = k < P.n;
    bool far = false, warn = false;
    uint32_t kk = 0;
    if (valid) {
        double4 p = pos4[k];
        const double lam = *lam_ptr;
        double a = vx[k] * lam, b = vy[k] * lam, c = vz[k] * lam;
        a = a + ax[k] * P.half_dt;
        b = b + ay[k] * P.half_dt;
        c = c + az[k] * P.half_dt;
        vx[k] = a;
        vy[k] = b;
        vz[k] = c;
        const double sx = a * P.dt, sy = b * P.dt, sz = c * P.dt;
        p.x = p.x + sx;
        p.y = p.y + sy;
        p.z = p.z + sz;
        if (P.flags & F_WRAP) {
            p.x = rem_euclid(p.x, 2.0 * P.half[0]) - P.half[0];
            p.y = rem_euclid(p.y, 2.0 * P.half[1]) - P.half[1];
            p.z = rem_euclid(p.z, 2.0 * P.half[2]) - P.half[2];
        }
        pos4[k] = p;
        // displacement from the build position of the lists this tick uses,
        // modulo h (the wrap translates the universe by h each tick)
        const double4 r = build_now ? xbin[k] : xref[k];
        double dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
        dx = fma(-P.half[0], rint(dx * P.inv_half[0]), dx);
        dy = fma(-P.half[1], rint(dy * P.inv_half[1]), dy);
        dz = fma(-P.half[2], rint(dz * P.inv_half[2]), dz);
        const double d2 = fma(dz, dz, fma(dy, dy, dx * dx));
        far = !(d2 <= half_skin * half_skin);  // NaN counts as far
        if (binreq) {
            xbin[k] = p;
            kk = cell_key(PL, p.x, p.y, p.z);
        } else {
            const double step = sqrt(fma(sz, sz, fma(sy, sy, sx * sx)));
            warn = !(fma(3.0, step, sqrt(d2)) < half_skin);
        }
    }
    if (binreq) {
        bin_wave(kk, valid, (uint32_t)k, bcount, bucket, bcap, flags);
        if (k == 0) flags[10] = 1u;  // the next tick builds
    }
    if (build_now && k == 0) flags[0] = 1u;  // the pair kernel empties the buckets
    if (__ballot(far) != 0 && lane == 0) atomicOr(&flags[4], 1u);
    if (__ballot(warn) != 0 && lane == 0) atomicOr(&flags[8], 1u);
}

// End of a list tick (the pair kernel's last workgroup): the rebuild flag
// and the workgroup counter are cleared, and an atom found within reach of
// skin/2 asks the next tick to bin.
__device__ __forceinline__ void close_list_tick(uint32_t* flags) {
    flags[5] = 0u;
    flags[0] = 0u;
    flags[9] = flags[8];
    flags[8] = 0u;
    flags[11] = flags[10];
    flags[10] = 0u;
}

// Bound-pair search fused into the list pass (bond observation on the list
// path): the list of an atom holds every pair within rc, so with the bond
// cutoff k_bind max(sigma) <= rc the pair kernel finds the candidates of
// bound_pairs_with_grid (src/analysis/pairs.rs:62-85) among the entries it
// evaluates anyway.
constexpr int BOUND_MAXT = 32;  // bound partners kept per owner (the tracker's episodes)
// Bond counters: 5 words (formed bonds, lifetime sum, bonded pairs, bonded
// entities, the overflow flag) and the species matrix, kept in 64 stripes:
// workgroup b adds to stripe b mod 64, so thousands of workgroups do not
// serialise on a few addresses; readers sum the stripes (the overflow word,
// [4], lives in stripe 0 only).
constexpr int BOND_CNT_WORDS = 5 + GS_NSPECIES * GS_NSPECIES;
constexpr int BOND_STRIPES = 64;
constexpr int BOND_CNT_ALL = BOND_STRIPES * BOND_CNT_WORDS;
constexpr int HIT_LANE = 16;    // screened candidates a lane's LDS column holds
// Bound-now lists by resident slot i, slot-major: word q of slot i at q n + i.
__host__ __device__ __forceinline__ size_t now_idx(uint32_t q, uint32_t i, size_t n) {
    return (size_t)q * n + i;
}
struct BoundArgs {
    uint32_t* hits = nullptr;   // [h * (N << LOGT) + lane]: the lane's h-th screened list entry
    uint32_t* hinfo = nullptr;  // [lane]: screened entries (clamped at 255) | band mask << 16
    BoundScreen scr;            // the list pass's screen around k_bind sigma_ab
    uint32_t screen_all = 1u;   // 0: the outer entries of split lists can hold no bound pair
};


// Per-atom list traversal fused with the second half of the tick: T lanes
// per atom split its list (entries s, s+T, ...) and sum with shuffles; the
// minimum image is taken per pair (dx - L rint(dx / L)); pairs of the skin
// shell beyond rc get exactly zero from the clamped switch, coincident pairs
// are masked.  The lane that owns the atom then stores a = F/m, applies the
// second half kick (src/systems/integrate.rs:38-44) and contributes the
// kinetic energy; the last workgroup reduces the per-block kinetic and
// potential partials in a fixed order, computes the Berendsen lambda and the
// observables (close_tick) and clears the rebuild flag for the next tick.
// Split lists: the groups that every lane of the wave has inside its inner
// class skip the switch; with the bound-pair screen, the groups past every
// lane's inner class skip the screen when no outer entry can be a bound pair.
// BATCH (gs_batch): the atoms of many replicas; each atom takes its
// replica's box from the batch tables and atoms of replicas on brute-force
// pairs are skipped.  The kernel stores a = F/m and the atom's potential
// energy (half its pair sum) in pe_out; the kick and the per-replica
// reductions run in the batch's own kernels.
#ifndef GS_LIST_WAVES
#define GS_LIST_WAVES 2  // minimum waves per SIMD the list pair kernel is compiled for
#endif
template <int LOGT, bool COUL, bool SHARED, bool SINGLE, bool GRAV, bool BATCH = false, bool BOUND = false>
__global__ __launch_bounds__(BLOCK, GS_LIST_WAVES) void k_pair_list(KParams P, uint32_t N, const double4* __restrict__ pos4,
                                                     const uint8_t* __restrict__ type,
                                                     const uint32_t* __restrict__ nbr,
                                                     const uint32_t* __restrict__ nlen,
                                                     const PairCoef* __restrict__ coef, double* __restrict__ ax,
                                                     double* __restrict__ ay, double* __restrict__ az,
                                                     double* __restrict__ vx, double* __restrict__ vy,
                                                     double* __restrict__ vz, double* __restrict__ vpart,
                                                     double* __restrict__ kpart, uint32_t* __restrict__ flags,
                                                     double* __restrict__ lam_out, gs_observables* __restrict__ obs,
                                                     ThermoArgs th, const double* __restrict__ abond,
                                                     const double* __restrict__ vbpart, uint32_t nvb, uint32_t rpc,
                                                     uint32_t* __restrict__ bcount, uint32_t nbucket,
                                                     BatchTabs bt = BatchTabs{}, double* __restrict__ pe_out = nullptr,
                                                     BoundArgs ba = BoundArgs{}) {
    constexpr int T = 1 << LOGT;
    __shared__ PairCoef sc[(SINGLE || GRAV) ? (SINGLE ? 1 : GS_NSPECIES * GS_NSPECIES) : 1];
    __shared__ ListCoef lc[SINGLE ? 1 : GS_NSPECIES * LIST_COLS];
    __shared__ double wv[BLOCK / WAVE];
    __shared__ double wk[BLOCK / WAVE];
    __shared__ double red[2 * BLOCK / WAVE];
    __shared__ bool last;
    // BOUND: each lane's screened bound-pair candidates (list entries), a
    // column per lane, then its accepted partners' entity ids
    // (slot HIT_LANE is the spare that the branch-free recording writes
    // once the column is full; it is never read)
    __shared__ uint32_t hl[BOUND ? HIT_LANE + 1 : 1][BOUND ? BLOCK : 1];
    const int tid = threadIdx.x;
    uint32_t nhit = 0u;   // BOUND: screened candidates of this lane
    uint32_t bandm = 0u;  // BOUND: column entries within 1e-10 of a threshold
    // Atom chunk of this workgroup.  Workgroup b runs on L2 slice b mod 8
    // and, when the grid holds rpc workgroups per CU of a 256-CU device,
    // shares its CU with b + 256, b + 512, ...: those get adjacent chunks
    // (their neighbour positions share the CU's L1) inside a contiguous
    // slab of chunks per L2 slice.  rpc = 0 keeps chunk = b.
    uint32_t chunk = blockIdx.x;
    if (rpc) {
        const uint32_t k = blockIdx.x >> 3;
        chunk = (blockIdx.x & 7u) * (gridDim.x >> 3) + rpc * (k & 31u) + (k >> 5);
    }
    const uint32_t g = chunk * BLOCK + tid;
    const uint32_t i = g >> LOGT;
    // the atom's own words are loaded before the coefficient tables are
    // staged, so the two latencies overlap
    double4 pi0 = make_double4(0.0, 0.0, 0.0, 0.0);
    int ti0 = 0;
    uint32_t nl0 = 0u;
    if (i < N) {
        pi0 = pos4[i];
        ti0 = type[i];
        nl0 = nlen[i];
    }
    if (SINGLE) {
        if (tid == 0) sc[0] = coef[0];
    } else {
        const SpCoef* spc = reinterpret_cast<const SpCoef*>(coef + COEF_SP_OFFSET);
        if (GRAV)
            for (int t = tid; t < GS_NSPECIES * GS_NSPECIES; t += BLOCK) sc[t] = coef[t];
        for (int t = tid; t < GS_NSPECIES * LIST_COLS; t += BLOCK) {
            const int ta = t / LIST_COLS, tb = min(t % LIST_COLS, GS_NSPECIES - 1);
            const SpCoef a = spc[ta], b = spc[tb];
            ListCoef c;
            c.sig = 0.5 * (a.sigma + b.sigma);
            c.eps4 = a.a4 * b.a4;
            c.e6 = 6.0 * c.eps4;
            c.irc = coef[ta * GS_NSPECIES + tb].inv_rcore;
            c.rcore = 0.5 * c.sig;
            const double th2 = c.sig * c.sig;
            c.thlo = th2 * ba.scr.kb2_lo;
            c.thhi = th2 * ba.scr.kb2_hi;
            c.pad = 0.0;
            lc[t] = c;
        }
    }
    __syncthreads();
    // a tick that rebuilt leaves the buckets empty for the next binning
    // (BATCH: those of the replicas that built, flags[11] if any did)
    if (flags[0] != 0u)
        for (uint32_t b = blockIdx.x * BLOCK + tid; b < nbucket; b += gridDim.x * BLOCK) bcount[b] = 0u;
    else if (BATCH && flags[11] != 0u)
        for (uint32_t b = blockIdx.x * BLOCK + tid; b < nbucket; b += gridDim.x * BLOCK)
            if (bt.rflags[4 * bt.lcell_rep[b] + 3] != 0u) bcount[b] = 0u;
    const uint32_t s = g & (T - 1);
    double fx = 0.0, fy = 0.0, fz = 0.0, vl = 0.0;
    int ti = 0;
    // a build that met a stencil too dense for LDS or a full bucket left some
    // lists unwritten: the host replays this call, so no list is read now
    const bool lists_ok = (flags[2] | flags[6]) == 0u;
    bool own = i < N;  // BATCH: and the atom's replica is on lists
    double Lx = P.box[0], Ly = P.box[1], Lz = P.box[2];
    double iLx = 0.0, iLy = 0.0, iLz = 0.0;
    if (i < N) {
        if (BATCH) {
            const uint32_t rep = bt.arep[i];
            const double4 bx = bt.box[rep], ib = bt.ibox[rep];
            Lx = bx.x;
            Ly = bx.y;
            Lz = bx.z;
            iLx = ib.x;
            iLy = ib.y;
            iLz = ib.z;
            own = bx.w != 0.0;
        } else {
            iLx = 1.0 / Lx;
            iLy = 1.0 / Ly;
            iLz = 1.0 / Lz;
        }
    }
    // this lane's entries: s, s+T, ... (cnt of them), four per group, the
    // first cntA of them in the list's inner class, the first cntB of those
    // in its bind class
    const uint32_t len = lists_ok && own ? nlen_entries(nl0) : 0u;
    const bool split = (nl0 & NLEN_SPLIT) != 0u;
    const uint32_t nA = min(nlen_inner(nl0), len), nB = min(nlen_bind(nl0), nA);
    const uint32_t cnt = len > s ? (len - s + T - 1) >> LOGT : 0u;
    const uint32_t cntA = nA > s ? (nA - s + T - 1) >> LOGT : 0u;
    const uint32_t cntB = nB > s ? (nB - s + T - 1) >> LOGT : 0u;
    // Wave-uniform group ranges: groups [0, gin) hold inner entries only on
    // every lane of the wave (a lane whose entries are all inner does not
    // limit it); BOUND: groups from gout on hold no entry the screen needs
    // (of a split list, only its bind class can hold bound pairs).
    uint32_t gin = 0u, gout = 0xFFFFFFFFu;
    if (!SINGLE) {
        uint32_t mn = cntA >= cnt ? 0xFFFFFFFFu : cntA;
        uint32_t mx = (BOUND && (!split || ba.screen_all)) ? cnt : cntB;
#pragma unroll
        for (int off = WAVE / 2; off > 0; off >>= 1) {
            mn = min(mn, (uint32_t)__shfl_xor((int)mn, off, WAVE));
            mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, WAVE));
        }
        gin = __builtin_amdgcn_readfirstlane(mn) >> 2;
        if (BOUND) gout = (__builtin_amdgcn_readfirstlane(mx) + 3u) >> 2;
    }
    if (i < N) ti = ti0;
    // (lists of a build that did not complete are not walked: they may hold
    // unwritten groups)
    if (i < N && lists_ok && own) {
        const double4 pi = pi0;
        const double kqi = P.ke * pi.w;
        const PairCoef* row = SINGLE ? sc : sc + ((SINGLE || GRAV) ? ti * GS_NSPECIES : 0);
        const ListCoef* lrow = lc + (SINGLE ? 0 : min(ti, GS_NSPECIES - 1) * LIST_COLS);
        // a group's four position gathers are all in flight before its four
        // pairs are evaluated
        const uint4* grp = reinterpret_cast<const uint4*>(nbr) + (((size_t)i << LOGT) + s);
        const uint32_t gstride = N << LOGT;  // < 2^31: N <= LIST_ATOMS_MAX
        // Software pipeline over the lane's entry groups, two groups per
        // loop trip with the registers of the two named A and B (so no
        // register receiving a load in flight is ever copied, which would
        // make the wave wait for it): a step for group k loads the entries of
        // group k+2, issues the four position gathers of group k+1 and then
        // evaluates group k, whose gathers have had the previous step's
        // arithmetic to arrive in.  Loads are unconditional (past the end a
        // lane reads its padding group, the atom's own entry, and gathers
        // its own position), so the waits count them statically.  The
        // entries stream from HBM once per tick and carry a non-temporal
        // hint, so they do not displace the positions the gathers reuse
        // from L2.
        const uint32_t ng = (cnt + 3) >> 2;  // groups of this lane
        // past its last group a lane reads group ng, which holds the atom's
        // own entry only (the build pads every lane's list with it)
        auto group_ptr = [&](uint32_t k) { return grp + (uint64_t)min(k, ng) * (uint64_t)gstride; };
        // BOUND: the pairs the pair law's screen marks (hitm: bound for
        // certain, or within 1e-10 of a threshold) join the lane's column;
        // bandm marks the column entries the exact test must decide
        // (branch-free: every entry is stored at the next free slot, which
        // only a hit keeps; once the column's HIT_LANE slots are taken the
        // entries go to the spare slot, and a column with more hits than
        // slots is never read: the host replays the call)
        auto record = [&](uint32_t hitm, const uint4& ee) {
            const uint32_t f4[4] = {ee.x, ee.y, ee.z, ee.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t slot = min(nhit, (uint32_t)HIT_LANE);
                const uint32_t h = (hitm >> q) & 0x11u;
                hl[slot][tid] = f4[q];
                bandm |= ((h >> 4) & 1u) << slot;
                nhit += h != 0u ? 1u : 0u;
            }
        };
        // positions of a group's four entries (past the end of the lane's
        // entries: the atom's own); entry << 5 is the byte offset of its
        // atom (N <= LIST_ATOMS_MAX)
        auto gather = [&](const uint4& e, double4 (&p)[4]) {
            const uint32_t f4[4] = {e.x, e.y, e.z, e.w};
            const char* base = reinterpret_cast<const char*>(pos4);
#pragma unroll
            for (int q = 0; q < 4; ++q) p[q] = *reinterpret_cast<const double4*>(base + (f4[q] << 5));
        };
        // the pair law over one group (SINGLE: the scalar law of the
        // one-coefficient table)
        auto evaluate = [&](auto inner_c, auto screen_c, const double4 (&p)[4], const uint4& ent) -> uint32_t {
            constexpr bool IN = decltype(inner_c)::value, SC = decltype(screen_c)::value;
            uint32_t hitm = 0u;
            if (SINGLE) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double4 pj = p[q];
                    double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                    dx = fma(-Lx, rint(dx * iLx), dx);
                    dy = fma(-Ly, rint(dy * iLy), dy);
                    dz = fma(-Lz, rint(dz * iLz), dz);
                    const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                    if (SC) {
                        const double th2 = row[0].sigma * row[0].sigma;
                        const bool sure = r2 > 1e-24 && r2 < th2 * ba.scr.kb2_lo;
                        const bool maybe = r2 > 0.0 && r2 < th2 * ba.scr.kb2_hi;
                        hitm |= (sure ? 1u : 0u) << q;
                        hitm |= (maybe && !sure ? 1u : 0u) << (4 + q);
                    }
                    const bool ok = r2 > P.eps2;
                    const double r2s = ok ? r2 : 1.0;
                    const double inv_d = rsqrt_refined(r2s);
                    const double d = r2s * inv_d;
                    double fmag, vv;
                    pair_law<COUL, SHARED, GRAV ? 1 : 0>(P, row[0], d, inv_d, kqi * pj.w, fmag, vv);
                    const double fs = ok ? fmag * inv_d : 0.0;
                    fx = fma(fs, dx, fx);
                    fy = fma(fs, dy, fy);
                    fz = fma(fs, dz, fz);
                    vl += ok ? vv : 0.0;
                }
            } else {
                pair_eval4<COUL, SHARED, GRAV, IN, SC, 0, 2>(P, Lx, Ly, Lz, iLx, iLy, iLz, pi, kqi, lrow, row, p, ent,
                                                             fx, fy, fz, vl, hitm);
                pair_eval4<COUL, SHARED, GRAV, IN, SC, 2, 2>(P, Lx, Ly, Lz, iLx, iLy, iLz, pi, kqi, lrow, row, p, ent,
                                                             fx, fy, fz, vl, hitm);
            }
            return hitm;
        };
        uint4 eA = load_nt(group_ptr(0));
        uint4 eB = load_nt(group_ptr(1));
        double4 pA[4], pB[4];
        uint4 jA = make_uint4(0u, 0u, 0u, 0u), jB = jA;  // a group's entries, kept for its law and screen
        // group 1's entries are issued before group 0's gathers, the order
        // every later step keeps
        __builtin_amdgcn_sched_barrier(0);
        gather(eA, pA);
        jA = eA;
        // Trips of two groups (k, k + 1) while k < kend, with the law of one
        // group class.  No exit between a step's loads and the next step (the
        // compiler would sink the loads past it): a lane with an odd number
        // of groups evaluates one group of four masked pairs at its end.
        // The classes run as separate loops, so each is scheduled and its
        // registers allocated on its own.
        uint32_t k = 0u;
        // (LAST: the lane's final phase; its trips stop while a group is
        // left, and the final trip, with one or two groups, issues no
        // gather past the lane's last group: a gathered lane costs the load
        // path the same whether its data is used or not)
        auto trips = [&](auto inner_c, auto screen_c, uint32_t kend, auto last_c) {
            constexpr bool SC = decltype(screen_c)::value;
            constexpr bool LAST = decltype(last_c)::value;
            for (; LAST ? k + 2u < kend : k < kend; k += 2) {
                // step A: group k
                __builtin_amdgcn_sched_barrier(0);
                eA = load_nt(group_ptr(k + 2));
                gather(eB, pB);
                jB = eB;
                __builtin_amdgcn_sched_barrier(0);
                {
                    const uint32_t hm = evaluate(inner_c, screen_c, pA, jA);
                    if (SC) record(hm, jA);
                }
                // step B: group k + 1
                __builtin_amdgcn_sched_barrier(0);
                eB = load_nt(group_ptr(k + 3));
                gather(eA, pA);
                jA = eA;
                __builtin_amdgcn_sched_barrier(0);
                {
                    const uint32_t hm = evaluate(inner_c, screen_c, pB, jB);
                    if (SC) record(hm, jB);
                }
            }
            if (LAST && k < kend) {
                // groups k and (if it exists) k + 1; k's gathers are in
                // flight, k + 1's entries loaded
                const bool two = k + 1u < kend;
                if (two) gather(eB, pB);
                jB = eB;
                {
                    const uint32_t hm = evaluate(inner_c, screen_c, pA, jA);
                    if (SC) record(hm, jA);
                }
                if (two) {
                    const uint32_t hm = evaluate(inner_c, screen_c, pB, jB);
                    if (SC) record(hm, jB);
                }
                k += 2;
            }
        };
        using yes = std::integral_constant<bool, true>;
        using no = std::integral_constant<bool, false>;
        using scr = std::integral_constant<bool, BOUND>;
        if (SINGLE) {
            trips(no{}, scr{}, ng, yes{});
        } else {
            // trips of two groups by class (gin rounded down, gout up, to
            // whole trips): inner and screened, inner, screened, the rest
            const uint32_t kin = gin & ~1u, kout = BOUND ? (gout + 1u) & ~1u : 0u;
            if (BOUND) trips(yes{}, scr{}, min(min(kin, kout), ng), no{});
            trips(yes{}, no{}, min(kin, ng), no{});
            if (BOUND) trips(no{}, scr{}, min(kout, ng), no{});
            trips(no{}, no{}, ng, yes{});
        }
    }
#pragma unroll
    for (int off = T / 2; off > 0; off >>= 1) {
        fx += __shfl_xor(fx, off, WAVE);
        fy += __shfl_xor(fy, off, WAVE);
        fz += __shfl_xor(fz, off, WAVE);
        if (BATCH) vl += __shfl_xor(vl, off, WAVE);
    }
    if (BOUND && i < N) {
        // the lane's column goes to global memory (k_bound_from_hits applies
        // the owner test and the exact test and orders the partners); a
        // column that overflowed is not stored, the call is replayed with the
        // aux-grid search
        const size_t NT = (size_t)N << LOGT;
        const uint32_t nh = nhit > (uint32_t)HIT_LANE ? 0u : nhit;
        for (uint32_t h = 0; h < nh; ++h) ba.hits[(size_t)h * NT + g] = hl[h][tid];
        ba.hinfo[g] = (bandm << 16) | min(nhit, 255u);
        if (nhit > (uint32_t)HIT_LANE) atomicOr(&flags[7], 1u);
    }
    if (BATCH) {
        if (own && s == 0) {
            const double im = P.inv_mass[ti];
            ax[i] = fx * im;
            ay[i] = fy * im;
            az[i] = fz * im;
            pe_out[i] = 0.5 * vl;
        }
        return;
    }
    double ke = 0.0;
    if (i < N && s == 0) {
        const double im = P.inv_mass[ti];
        double a0 = fx * im, a1 = fy * im, a2 = fz * im;
        if (abond) {
            a0 += abond[i];
            a1 += abond[N + i];
            a2 += abond[2 * (size_t)N + i];
        }
        ax[i] = a0;
        ay[i] = a1;
        az[i] = a2;
        double u0 = vx[i], u1 = vy[i], u2 = vz[i];
        u0 = u0 + a0 * P.half_dt;
        u1 = u1 + a1 * P.half_dt;
        u2 = u2 + a2 * P.half_dt;
        vx[i] = u0;
        vy[i] = u1;
        vz[i] = u2;
        ke = 0.5 * P.mass[ti] * (u0 * u0 + u1 * u1 + u2 * u2);
    }
    vl = wave_sum(vl);
    ke = wave_sum(ke);
    if ((tid & 63) == 0) {
        wv[tid >> 6] = vl;
        wk[tid >> 6] = ke;
    }
    __syncthreads();
    if (tid == 0) {
        publish(vpart + blockIdx.x, 0.5 * ((wv[0] + wv[1]) + (wv[2] + wv[3])));
        publish(kpart + blockIdx.x, (wk[0] + wk[1]) + (wk[2] + wk[3]));
        last = take_ticket(&flags[5]) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    double K, V;
    block_sum2_fixed<BLOCK, true>(kpart, vpart, gridDim.x, red, K, V);
    const double Vb = vbpart ? block_sum_fixed(vbpart, nvb, red) : 0.0;
    if (tid == 0) {
        close_tick(K, V, Vb, (int)N, th, lam_out, obs, fault_word(lam_out));
        close_list_tick(flags);
    }
}

// Bound pairs of the fused search, one thread per resident atom i: the
// screened entries of its T lanes (k_pair_list<.., BOUND>) pass the owner
// test (a larger entity id) and, when the screen put them within 1e-10 of the
// threshold, bound_pairs_with_grid's exact test (the reference's min_image
// with its divisions, IEEE sqrt, d^2 < cutoff^2, d > EPS, d < k_bind
// sigma_ab, src/analysis/pairs.rs:62-85 and :132-174).  The accepted
// partners are ordered in the thread's LDS row and written slot-major by
// entity, as the aux-grid search writes them.
template <int LOGT>
__global__ __launch_bounds__(BLOCK) void k_bound_from_hits(KParams P, uint32_t N, const double4* __restrict__ pos4,
                                                           const uint8_t* __restrict__ type,
                                                           const uint32_t* __restrict__ id,
                                                           const uint32_t* __restrict__ hits,
                                                           const uint32_t* __restrict__ hinfo,
                                                           const double* __restrict__ thr, double cd2,
                                                           uint32_t* __restrict__ nowp, uint32_t* __restrict__ nown,
                                                           unsigned long long* __restrict__ cnt,
                                                           uint32_t* __restrict__ clr_deg,
                                                           uint32_t* __restrict__ clr_bdeg) {
    constexpr int T = 1 << LOGT;
    __shared__ uint32_t rows[BOUND_MAXT][BLOCK];
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    // the tracker's inputs cleared for the tick (k_bond_clear's work: the
    // next set's degrees, the degrees by slot, the per-tick counter words;
    // the grid covers max(N, the counter words))
    if (i < N) {
        clr_deg[i] = 0u;
        clr_bdeg[i] = 0u;
    }
    if (i < (uint32_t)BOND_CNT_ALL) {
        const uint32_t w = i % (uint32_t)BOND_CNT_WORDS;
        if (w == 2u || w == 3u || w >= 5u) cnt[i] = 0ull;
    }
    if (i >= N) return;
    const size_t NT = (size_t)N << LOGT;
    const uint32_t e = id[i];
    const int ti = type[i];
    uint32_t m = 0u;
    bool pos_loaded = false;
    double4 pi = make_double4(0.0, 0.0, 0.0, 0.0);
    uint32_t info[T];
#pragma unroll
    for (int s = 0; s < T; ++s) info[s] = hinfo[(i << LOGT) + (uint32_t)s];
    // four hits at a time: their entry loads, then their id gathers, are in
    // flight together; the first four of every lane's column are all loaded
    // before any is considered (past a column's end: the atom's own slot)
    uint32_t f0[T][4], em0[T][4];
#pragma unroll
    for (int s = 0; s < T; ++s) {
        const uint32_t c = info[s] & 0xFFu, nh = c > (uint32_t)HIT_LANE ? 0u : c;
#pragma unroll
        for (int k = 0; k < 4; ++k) f0[s][k] = (uint32_t)k < nh ? hits[(size_t)k * NT + (i << LOGT) + (uint32_t)s] : i;
    }
#pragma unroll
    for (int s = 0; s < T; ++s)
#pragma unroll
        for (int k = 0; k < 4; ++k) em0[s][k] = id[f0[s][k] & NBR_IDX_MASK];
#pragma unroll
    for (int s = 0; s < T; ++s) {
        const uint32_t g = (i << LOGT) + (uint32_t)s;
        const uint32_t c = info[s] & 0xFFu, band = info[s] >> 16;
        const uint32_t nh = c > (uint32_t)HIT_LANE ? 0u : c;  // an overflowed column is replayed
        for (uint32_t h0 = 0; h0 < nh; h0 += 4) {
            uint32_t f[4], em[4];
            if (h0 == 0) {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    f[k] = f0[s][k];
                    em[k] = em0[s][k];
                }
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) f[k] = hits[(size_t)min(h0 + (uint32_t)k, nh - 1u) * NT + g];
#pragma unroll
                for (int k = 0; k < 4; ++k) em[k] = id[f[k] & NBR_IDX_MASK];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t h = h0 + (uint32_t)k;
                bool acc = h < nh && em[k] > e;
                if (acc && ((band >> h) & 1u)) {
                    if (!pos_loaded) {
                        pi = pos4[i];
                        pos_loaded = true;
                    }
                    const double4 q = pos4[f[k] & NBR_IDX_MASK];
                    const double dx = min_image1(pi.x - q.x, P.box[0]);
                    const double dy = min_image1(pi.y - q.y, P.box[1]);
                    const double dz = min_image1(pi.z - q.z, P.box[2]);
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    const double d = sqrt(d2);
                    acc = d2 < cd2 && d > F64_EPS && d < thr[ti * GS_NSPECIES + (f[k] >> 28)];
                }
                if (!acc) continue;
                if (m < (uint32_t)BOUND_MAXT) {
                    // insertion in ascending id order of partner words
                    // (species << 28 | entity id; the entry carries the species)
                    uint32_t w = m;
                    while (w > 0 && (rows[w - 1][threadIdx.x] & 0x0FFFFFFFu) > em[k]) {
                        rows[w][threadIdx.x] = rows[w - 1][threadIdx.x];
                        --w;
                    }
                    rows[w][threadIdx.x] = (f[k] & 0xF0000000u) | em[k];
                }
                ++m;
            }
        }
    }
    // by resident slot, so the writes of a wave are coalesced; the tracker
    // reads them through slot_of
    const uint32_t mm = min(m, (uint32_t)BOUND_MAXT);
    for (uint32_t r = 0; r < mm; ++r) nowp[now_idx(r, i, N)] = rows[r][threadIdx.x];
    nown[i] = mm;
    if (m > (uint32_t)BOUND_MAXT) atomicOr(cnt + 4, 1ull);
}

// Exclusive prefix sum of one value per thread over a workgroup of NT
// threads (wsum: NT / 64 words of LDS); total receives the sum.
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* wsum, uint32_t& total) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, WAVE);
        if (lane >= off) x += y;
    }
    if (lane == WAVE - 1) wsum[w] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int k = 0; k < NT / WAVE; ++k) {
        const uint32_t t = wsum[k];
        before += k < w ? t : 0u;
        all += t;
    }
    __syncthreads();
    total = all;
    return before + x - v;
}

// Cell-staged list pass (pair_mode 2), fused like k_pair_list with the
// second kick, the kinetic reduction and close_tick.  One workgroup per
// list-grid cell (workgroup b takes cell (b mod 8) * per + b / 8, so each
// of the eight L2 slices of the device sees a contiguous slab of cells and
// their shared stencil candidates): the positions of the cell's ordered
// stencil candidates are staged in LDS (in chunks of STAGE_CAP), and the
// lanes of each home atom evaluate its list (lane s of T: entries s, s+T,
// ...) with the candidate read from LDS; per-pair minimum image as in
// k_pair_list.  The workgroup's lanes are shared out among the home atoms
// in proportion to their list lengths (atom h gets lanes
// [L_h, L_h+1) with L_h = floor((NT - nh) E_h / E) + h, E_h the entries of
// the atoms before it), so every lane has about the same number of pairs;
// a cell with more than NT/2 home atoms takes them NT at a time, one lane
// each.  A list is ordered by candidate position, so a lane's entries
// inside a chunk are a prefix of those it has left.  The partial forces of
// a home atom's lanes are summed through LDS in lane order.
template <bool COUL, bool SHARED, bool SINGLE, bool GRAV>
__global__ __launch_bounds__(CELL_THREADS) __attribute__((amdgpu_waves_per_eu(5))) void k_pair_cell(
    KParams P, uint32_t ncell, uint32_t per, const double4* __restrict__ pos4, const uint16_t* __restrict__ nbr16,
    size_t region, uint32_t cap, const uint2* __restrict__ hrec, uint32_t bcap,
    const uint32_t* __restrict__ cand_store, const uint4* __restrict__ cinfo, const PairCoef* __restrict__ coef,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az, double* __restrict__ vx,
    double* __restrict__ vy, double* __restrict__ vz, double* __restrict__ vpart, double* __restrict__ kpart,
    uint32_t* __restrict__ flags, double* __restrict__ lam_out, gs_observables* __restrict__ obs, ThermoArgs th,
    const double* __restrict__ abond, const double* __restrict__ vbpart, uint32_t nvb, uint32_t N,
    uint32_t* __restrict__ bcount) {
    constexpr int NW = CELL_THREADS / WAVE;
    constexpr int SK = (STAGE_CAP + CELL_THREADS - 1) / CELL_THREADS;  // staged candidates per thread
    __shared__ PairCoef sc[SINGLE ? 1 : GS_NSPECIES * GS_NSPECIES];
    __shared__ SpCoef ssp[GS_NSPECIES];
    __shared__ double sirc[SINGLE ? 1 : GS_NSPECIES * GS_NSPECIES];  // 1 / rcore, 8-byte stride
    __shared__ double4 tile[STAGE_CAP];
    __shared__ uint2 srec[CELL_THREADS];       // home records of the pass
    __shared__ uint32_t sL[CELL_THREADS + 1];  // first lane of each home atom
    __shared__ uint32_t wsum[NW];
    __shared__ double wv[NW];
    __shared__ double wk[NW];
    __shared__ double red[2 * NW];
    __shared__ bool last;
    static_assert(3 * CELL_THREADS * sizeof(double) <= STAGE_CAP * sizeof(double4), "partials fit the tile");
    const int tid = threadIdx.x;
    if (SINGLE) {
        if (tid == 0) sc[0] = coef[0];
    } else {
        for (int t = tid; t < GS_NSPECIES * GS_NSPECIES; t += CELL_THREADS) {
            sc[t] = coef[t];
            sirc[t] = coef[t].inv_rcore;
        }
        if (tid < GS_NSPECIES) ssp[tid] = reinterpret_cast<const SpCoef*>(coef + COEF_SP_OFFSET)[tid];
    }
    const uint32_t c = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    const uint32_t cc = c < ncell ? c : 0u;
    // a tick that rebuilt leaves the buckets empty for the next binning
    if (tid == 0 && c < ncell && flags[0] != 0u) bcount[c] = 0u;
    // a build that met a stencil too dense for LDS or a full bucket left some
    // lists unwritten: the host replays this call, so no list is read now
    const uint32_t bad = flags[2] | flags[6];
    uint4 ci = cinfo[cc];
    if (bad != 0u || c >= ncell) ci = make_uint4(0u, 0u, 0u, 0u);
    const uint32_t ncand = ci.x, nh = ci.z;
    const bool prop = nh <= (uint32_t)CELL_THREADS / 2;  // proportional lanes, one pass
    const uint32_t* cand = cand_store + (size_t)cc * LDS_IDX_CAP_LOCAL;
    const uint16_t* reg = nbr16 + (size_t)cc * region;
    const uint2* hr0 = hrec + (size_t)cc * bcap;
    const int nchunks = (int)((ncand + STAGE_CAP - 1) / STAGE_CAP);
    const double Lx = P.box[0], Ly = P.box[1], Lz = P.box[2];
    const double iLx = 1.0 / Lx, iLy = 1.0 / Ly, iLz = 1.0 / Lz;
    double* part = reinterpret_cast<double*>(tile);
    double vl = 0.0, ke = 0.0;
    for (uint32_t h0 = 0; h0 < nh; h0 += CELL_THREADS) {  // one pass unless nh > CELL_THREADS
        const uint32_t nhp = min(nh - h0, (uint32_t)CELL_THREADS);
        // this pass's home records, and the lane split
        const uint2 myrec = (uint32_t)tid < nhp ? hr0[h0 + tid] : make_uint2(0u, 0u);
        uint32_t E = 0;
        const uint32_t Eh = block_exclusive_scan<CELL_THREADS>(myrec.y, wsum, E);
        if ((uint32_t)tid < nhp) {
            srec[tid] = myrec;
            sL[tid] = prop ? (E ? (uint32_t)(((uint64_t)(CELL_THREADS - nhp) * Eh) / E) : 0u) + (uint32_t)tid
                           : (uint32_t)tid;
        }
        if (tid == 0) sL[nhp] = prop ? (uint32_t)CELL_THREADS : nhp;
        __syncthreads();
        // this lane's home atom: sL[h] <= tid < sL[h + 1]
        uint32_t h = 0, T = 1;
        bool act;
        if (prop) {
            uint32_t lo = 0, hi = nhp;  // sL[lo] <= tid < sL[hi]
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (sL[mid] <= (uint32_t)tid)
                    lo = mid;
                else
                    hi = mid;
            }
            h = lo;
            T = sL[h + 1] - sL[h];
            act = nhp > 0;
        } else {
            h = (uint32_t)tid;
            act = h < nhp;
        }
        const uint32_t s = act ? (uint32_t)tid - sL[h] : 0u;
        const uint2 hr = act ? srec[h] : make_uint2(0u, 0u);
        const uint32_t gi = hr.x & NBR_IDX_MASK;
        const int ti = (int)(hr.x >> 28);
        const uint32_t cnt = hr.y > s ? (hr.y - s + T - 1) / T : 0u;
        const double4 pi = act ? pos4[gi] : make_double4(0.0, 0.0, 0.0, 0.0);
        const double kqi = P.ke * pi.w;
        const PairCoef* rowc = SINGLE ? sc : sc + ti * GS_NSPECIES;
        const double sig_i = SINGLE ? 0.0 : ssp[ti].sigma, a4_i = SINGLE ? 0.0 : ssp[ti].a4;
        const uint16_t* my = reg + (size_t)(h0 + h) * cap + s;  // entry m at my[m * T]
        double fx = 0.0, fy = 0.0, fz = 0.0;
        uint32_t m = 0;
        for (int ch = 0; ch < nchunks; ++ch) {
            const uint32_t c0 = (uint32_t)ch * STAGE_CAP;
            const uint32_t c1 = min(ncand, c0 + (uint32_t)STAGE_CAP);
            // the chunk's candidate indices, then their positions, and this
            // lane's next four entries are all in flight together
            uint32_t cix[SK];
#pragma unroll
            for (int k = 0; k < SK; ++k) {
                const uint32_t t = c0 + (uint32_t)tid + (uint32_t)k * CELL_THREADS;
                cix[k] = t < c1 ? cand[t] : 0u;
            }
            uint32_t en[4];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                en[q] = m + q < cnt ? __builtin_nontemporal_load(my + (size_t)(m + q) * T) : 0u;
            double4 pst[SK];
#pragma unroll
            for (int k = 0; k < SK; ++k) {
                const uint32_t t = c0 + (uint32_t)tid + (uint32_t)k * CELL_THREADS;
                pst[k] = t < c1 ? pos4[cix[k]] : make_double4(0.0, 0.0, 0.0, 0.0);
            }
            if (ch > 0 || h0 > 0) __syncthreads();  // the previous chunk's readers are done
#pragma unroll
            for (int k = 0; k < SK; ++k) {
                const uint32_t t = (uint32_t)tid + (uint32_t)k * CELL_THREADS;
                if (c0 + t < c1) tile[t] = pst[k];
            }
            __syncthreads();
            for (;;) {
                uint32_t e4[4];
                bool live[4];
                uint32_t nl = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    e4[q] = en[q];
                    const uint32_t j = e4[q] & CELL_IDX_MASK;
                    live[q] = m + q < cnt && j < c1 && j >= c0;
                    nl += live[q] ? 1u : 0u;
                }
                // the following four, assuming this group is whole (a short
                // group ends the chunk and the loads are made again)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    en[q] = m + 4 + q < cnt ? __builtin_nontemporal_load(my + (size_t)(m + 4 + q) * T) : 0u;
#pragma unroll 1
                for (int q = 0; q < 4; ++q) {
                    const double4 pj = tile[live[q] ? (e4[q] & CELL_IDX_MASK) - c0 : 0u];
                    const int tj = live[q] ? (int)(e4[q] >> CELL_IDX_BITS) : 0;
                    double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                    dx = fma(-Lx, rint(dx * iLx), dx);
                    dy = fma(-Ly, rint(dy * iLy), dy);
                    dz = fma(-Lz, rint(dz * iLz), dz);
                    const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                    const bool ok = live[q] && r2 > P.eps2;
                    const double r2s = ok ? r2 : 1.0;
                    const double inv_d = rsqrt_refined(r2s);
                    const double d = r2s * inv_d;
                    double fmag, vv;
                    if (SINGLE)
                        pair_law<COUL, SHARED, GRAV ? 1 : 0>(P, rowc[0], d, inv_d, kqi * pj.w, fmag, vv);
                    else
                        pair_law_sp<COUL, SHARED, GRAV>(P, sig_i, a4_i, ssp[tj], sirc[ti * GS_NSPECIES + tj], rowc,
                                                        tj, d, inv_d, kqi * pj.w, fmag, vv);
                    const double fs = ok ? fmag * inv_d : 0.0;
                    fx = fma(fs, dx, fx);
                    fy = fma(fs, dy, fy);
                    fz = fma(fs, dz, fz);
                    vl += ok ? vv : 0.0;
                }
                m += nl;
                if (nl < 4) break;
            }
        }
        __syncthreads();  // the tile's memory holds the partial forces now
        part[tid] = fx;
        part[CELL_THREADS + tid] = fy;
        part[2 * CELL_THREADS + tid] = fz;
        __syncthreads();
        if (act && s == 0) {
            double sx = 0.0, sy = 0.0, sz = 0.0;
            for (uint32_t q = 0; q < T; ++q) {
                sx += part[tid + q];
                sy += part[CELL_THREADS + tid + q];
                sz += part[2 * CELL_THREADS + tid + q];
            }
            const double im = P.inv_mass[ti];
            double a0 = sx * im, a1 = sy * im, a2 = sz * im;
            if (abond) {
                a0 += abond[gi];
                a1 += abond[N + gi];
                a2 += abond[2 * (size_t)N + gi];
            }
            ax[gi] = a0;
            ay[gi] = a1;
            az[gi] = a2;
            double u0 = vx[gi], u1 = vy[gi], u2 = vz[gi];
            u0 = u0 + a0 * P.half_dt;
            u1 = u1 + a1 * P.half_dt;
            u2 = u2 + a2 * P.half_dt;
            vx[gi] = u0;
            vy[gi] = u1;
            vz[gi] = u2;
            ke += 0.5 * P.mass[ti] * (u0 * u0 + u1 * u1 + u2 * u2);
        }
        __syncthreads();  // srec, sL and the tile are rewritten by the next pass
    }
    vl = wave_sum(vl);
    ke = wave_sum(ke);
    if ((tid & 63) == 0) {
        wv[tid >> 6] = vl;
        wk[tid >> 6] = ke;
    }
    __syncthreads();
    if (tid == 0) {
        double sv = wv[0], sk = wk[0];
        for (int w = 1; w < NW; ++w) {
            sv += wv[w];
            sk += wk[w];
        }
        publish(vpart + blockIdx.x, 0.5 * sv);
        publish(kpart + blockIdx.x, sk);
        last = take_ticket(&flags[5]) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    double K, V;
    block_sum2_fixed<CELL_THREADS, true>(kpart, vpart, gridDim.x, red, K, V);
    const double Vb = vbpart ? block_sum_fixed<CELL_THREADS>(vbpart, nvb, red) : 0.0;
    if (tid == 0) {
        close_tick(K, V, Vb, (int)N, th, lam_out, obs, fault_word(lam_out));
        close_list_tick(flags);
    }
}

// ---- observation lenses --------------------------------------------------------

// Auxiliary grid of a lens: SpatialGrid::new(size, min_cell) over a point set
// (src/physics/grid.rs:67-86).
// velocity_histogram (src/stats/mod.rs:269-297): |v| = sqrt(x^2 + y^2 + z^2)
// (Vec3::length, left to right); speeds >= max_speed are overflow, the
// others fall in bin min((s / bin_width) as usize, nbins - 1).  Integer
// counts, so the order of the atomic adds does not matter.
// StatsSystem's per-entity sums (src/systems/aggregate.rs:87-93): kinetic
// 0.5 m |v|^2 and speed |v| = sqrt(|v|^2), |v|^2 = x x + y y + z z left to
// right (Vec3::length_squared), one partial of each per workgroup in a fixed
// order; k_reduce_to finishes both.
__global__ __launch_bounds__(BLOCK) void k_speed_parts(int n, const double* __restrict__ vx,
                                                      const double* __restrict__ vy, const double* __restrict__ vz,
                                                      const uint8_t* __restrict__ type, KParams P,
                                                      double* __restrict__ kpart, double* __restrict__ spart) {
    __shared__ double rk[BLOCK / WAVE], rs[BLOCK / WAVE];
    const int tid = threadIdx.x, k = blockIdx.x * BLOCK + tid;
    double ke = 0.0, sp = 0.0;
    if (k < n) {
        const double a = vx[k], b = vy[k], c = vz[k];
        const double l2 = a * a + b * b + c * c;
        ke = 0.5 * P.mass[type[k]] * l2;
        sp = sqrt(l2);
    }
    ke = wave_sum(ke);
    sp = wave_sum(sp);
    if ((tid & 63) == 0) {
        rk[tid >> 6] = ke;
        rs[tid >> 6] = sp;
    }
    __syncthreads();
    if (tid == 0) {
        kpart[blockIdx.x] = (rk[0] + rk[1]) + (rk[2] + rk[3]);
        spart[blockIdx.x] = (rs[0] + rs[1]) + (rs[2] + rs[3]);
    }
}

constexpr int VHIST_MAX_BINS = 512;
__global__ __launch_bounds__(BLOCK) void k_vhist(int n, const double* __restrict__ vx, const double* __restrict__ vy,
                                                 const double* __restrict__ vz, double max_speed, double bin_width,
                                                 int nbins, unsigned long long* __restrict__ counts) {
    __shared__ uint32_t h[VHIST_MAX_BINS + 1];  // [nbins] = overflow
    for (int b = threadIdx.x; b <= nbins; b += BLOCK) h[b] = 0u;
    __syncthreads();
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k < n) {
        const double x = vx[k], y = vy[k], z = vz[k];
        const double sp = sqrt(x * x + y * y + z * z);
        int bin;
        if (sp >= max_speed) {
            bin = nbins;
        } else {
            const double q = sp / bin_width;
            const uint32_t i = q > 0.0 ? (q < 4294967295.0 ? (uint32_t)q : 0xFFFFFFFFu) : 0u;
            bin = (int)min(i, (uint32_t)(nbins - 1));
        }
        atomicAdd(&h[bin], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b <= nbins; b += BLOCK)
        if (h[b]) atomicAdd(&counts[b], (unsigned long long)h[b]);
}

struct AuxGrid {
    uint32_t dims[3];
    double cell[3];
    uint32_t ncell;
};

__device__ __forceinline__ uint32_t aux_key(const AuxGrid& G, const KParams& P, double x, double y, double z,
                                            uint32_t c[3]) {
    c[0] = cell_axis(x, P.box[0], G.cell[0], G.dims[0]);
    c[1] = cell_axis(y, P.box[1], G.cell[1], G.dims[1]);
    c[2] = cell_axis(z, P.box[2], G.cell[2], G.dims[2]);
    return c[0] + G.dims[0] * (c[1] + G.dims[1] * c[2]);
}

__global__ __launch_bounds__(BLOCK) void k_aux_key(int n, AuxGrid G, KParams P, const double4* __restrict__ pts,
                                                   uint32_t* __restrict__ key, uint32_t* __restrict__ ticket,
                                                   uint32_t* __restrict__ count) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n) return;
    const double4 p = pts[k];
    uint32_t c[3];
    const uint32_t kk = aux_key(G, P, p.x, p.y, p.z, c);
    key[k] = kk;
    ticket[k] = atomicAdd(&count[kk], 1u);
}

// Stencil offsets of one axis as the reference grid enumerates them:
// 1 cell -> {0} (canonical_offset zeroes the axis); 2 cells -> {-1, 0, +1}
// in reference mode, both non-zero offsets reaching the same cell (the
// pair multiplicity of src/physics/grid.rs:134-151), or {0, +1} in min-image
// mode; >= 3 cells -> {-1, 0, +1}.
__device__ __forceinline__ int axis_offsets(uint32_t dims, bool refm, int* o) {
    if (dims == 1) { o[0] = 0; return 1; }
    if (dims == 2 && !refm) { o[0] = 0; o[1] = 1; return 2; }
    o[0] = -1; o[1] = 0; o[2] = 1;
    return 3;
}

// g(r) pair counts (radial_distribution, src/analysis/mod.rs:103-117): every
// unordered pair of the subsample found by the grid of edge r_max, counted
// from its lower-index end, with the reference's min_image, length and
// (d / dr) as usize.  One wave per point, its lanes strided over each
// stencil cell's members.  Integer counts are order independent, so the
// u32 LDS histogram with atomics is exact.
template <bool BETWEEN>
__global__ __launch_bounds__(BLOCK) void k_gr_pairs(int n, AuxGrid G, KParams P, const double4* __restrict__ pts,
                                                    const uint8_t* __restrict__ ptype,
                                                    const uint32_t* __restrict__ start,
                                                    const uint32_t* __restrict__ perm, double r_max, double dr,
                                                    int nbins, int ta, int tb, int refm,
                                                    unsigned long long* __restrict__ bins) {
    extern __shared__ uint32_t hist[];
    const bool lds = nbins <= 4096;
    if (lds)
        for (int b = threadIdx.x; b < nbins; b += BLOCK) hist[b] = 0u;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * (BLOCK / WAVE) + (threadIdx.x >> 6);
    if (i < n) {
        const double4 p = pts[i];
        const int ti = BETWEEN ? ptype[i] : 0;
        uint32_t c[3];
        aux_key(G, P, p.x, p.y, p.z, c);
        int ox[3], oy[3], oz[3];
        const int nx = axis_offsets(G.dims[0], refm, ox), ny = axis_offsets(G.dims[1], refm, oy),
                  nz = axis_offsets(G.dims[2], refm, oz);
        const double cut2 = r_max * r_max;
        for (int a = 0; a < nz; ++a)
            for (int b = 0; b < ny; ++b)
                for (int e = 0; e < nx; ++e) {
                    const uint32_t ix = (c[0] + G.dims[0] + ox[e]) % G.dims[0];
                    const uint32_t iy = (c[1] + G.dims[1] + oy[b]) % G.dims[1];
                    const uint32_t iz = (c[2] + G.dims[2] + oz[a]) % G.dims[2];
                    const uint32_t nc = ix + G.dims[0] * (iy + G.dims[1] * iz);
                    for (uint32_t t = start[nc] + lane; t < start[nc + 1]; t += WAVE) {
                        const uint32_t j = perm[t];
                        if ((int)j <= i) continue;
                        if (BETWEEN) {
                            const int tj = ptype[j];
                            if (!((ti == ta && tj == tb) || (ti == tb && tj == ta))) continue;
                        }
                        const double4 q = pts[j];
                        const double dx = min_image1(p.x - q.x, P.box[0]);
                        const double dy = min_image1(p.y - q.y, P.box[1]);
                        const double dz = min_image1(p.z - q.z, P.box[2]);
                        const double d2 = dx * dx + dy * dy + dz * dz;
                        if (!(d2 < cut2)) continue;
                        const double d = sqrt(d2);
                        if (d <= F64_EPS || d >= r_max) continue;
                        double fb = d / dr;
                        uint32_t bin = fb >= (double)nbins ? (uint32_t)(nbins - 1) : (uint32_t)fb;
                        if (lds)
                            atomicAdd(&hist[bin], 1u);
                        else
                            atomicAdd(&bins[bin], 1ull);
                    }
                }
    }
    __syncthreads();
    if (lds)
        for (int b = threadIdx.x; b < nbins; b += BLOCK)
            if (hist[b]) atomicAdd(&bins[b], (unsigned long long)hist[b]);
}

// Subsample of g_subsample (src/analysis/mod.rs:32-46) gathered from the
// sorted state: entity e is kept when e % step == 0 and e / step < count.
__global__ __launch_bounds__(BLOCK) void k_gr_gather(int n, const uint32_t* __restrict__ id,
                                                     const double4* __restrict__ pos4,
                                                     const uint8_t* __restrict__ type, uint32_t step, uint32_t count,
                                                     double4* __restrict__ sub, uint8_t* __restrict__ stype) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n) return;
    const uint32_t e = id[k];
    if (e % step) return;
    const uint32_t s = e / step;
    if (s >= count) return;
    sub[s] = pos4[k];
    stype[s] = type[k];
}

// Friends-of-friends (clusters, src/analysis/mod.rs:221-284) as a concurrent
// union-find over entity ids: a root is only ever hooked under a smaller
// root, so each component ends rooted at its smallest entity id whatever the
// order of the hooks (the canonical label), and the partition is exact.
__device__ __forceinline__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
    uint32_t p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (p != x) {
        x = p;
        p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return x;
}

__device__ __forceinline__ void uf_union(uint32_t* parent, uint32_t a, uint32_t b) {
    while (true) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a > b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        const uint32_t old = atomicCAS(&parent[b], b, a);
        if (old == b) return;
        b = old;
    }
}

__global__ __launch_bounds__(BLOCK) void k_fof_edges(int n, AuxGrid G, KParams P, const double4* __restrict__ pos4,
                                                     const uint8_t* __restrict__ type, const uint32_t* __restrict__ id,
                                                     const uint32_t* __restrict__ start,
                                                     const uint32_t* __restrict__ perm,
                                                     const double* __restrict__ sigma, double bond_factor,
                                                     double cutoff, int refm, uint32_t* __restrict__ parent,
                                                     unsigned long long* __restrict__ bound) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    unsigned long long local = 0;
    if (k < n) {
        const double4 p = pos4[k];
        const uint32_t ek = id[k];
        const double sk = sigma[type[k]];
        uint32_t c[3];
        aux_key(G, P, p.x, p.y, p.z, c);
        int ox[3], oy[3], oz[3];
        const int nx = axis_offsets(G.dims[0], refm, ox), ny = axis_offsets(G.dims[1], refm, oy),
                  nz = axis_offsets(G.dims[2], refm, oz);
        const double cut2 = cutoff * cutoff;
        for (int a = 0; a < nz; ++a)
            for (int b = 0; b < ny; ++b)
                for (int e = 0; e < nx; ++e) {
                    const uint32_t ix = (c[0] + G.dims[0] + ox[e]) % G.dims[0];
                    const uint32_t iy = (c[1] + G.dims[1] + oy[b]) % G.dims[1];
                    const uint32_t iz = (c[2] + G.dims[2] + oz[a]) % G.dims[2];
                    const uint32_t nc = ix + G.dims[0] * (iy + G.dims[1] * iz);
                    for (uint32_t t = start[nc]; t < start[nc + 1]; ++t) {
                        const uint32_t m = perm[t];
                        const uint32_t em = id[m];
                        if (em <= ek) continue;
                        const double4 q = pos4[m];
                        const double dx = min_image1(p.x - q.x, P.box[0]);
                        const double dy = min_image1(p.y - q.y, P.box[1]);
                        const double dz = min_image1(p.z - q.z, P.box[2]);
                        const double d2 = dx * dx + dy * dy + dz * dz;
                        if (!(d2 < cut2)) continue;
                        const double d = sqrt(d2);
                        if (d <= F64_EPS) continue;
                        const double bond = bond_factor * 0.5 * (sk + sigma[type[m]]);
                        if (d < bond) {
                            local += 1;
                            uf_union(parent, ek, em);
                        }
                    }
                }
    }
    if (local) atomicAdd(bound, local);
}

__global__ __launch_bounds__(BLOCK) void k_iota(int n, uint32_t* __restrict__ a) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k < n) a[k] = (uint32_t)k;
}

__global__ __launch_bounds__(BLOCK) void k_fof_finalize(int n, uint32_t* __restrict__ parent,
                                                        uint32_t* __restrict__ size) {
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= n) return;
    const uint32_t r = uf_find(parent, (uint32_t)e);
    atomicAdd(&size[r], 1u);
}

__global__ __launch_bounds__(BLOCK) void k_fof_labels(int n, uint32_t* __restrict__ parent,
                                                      uint32_t* __restrict__ label) {
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= n) return;
    label[e] = uf_find(parent, (uint32_t)e);
}

// monomers, aggregates, largest, roots (u64 each).
__global__ __launch_bounds__(BLOCK) void k_fof_stats(int n, const uint32_t* __restrict__ size,
                                                     unsigned long long* __restrict__ st) {
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= n) return;
    const uint32_t s = size[e];
    if (s == 0) return;
    atomicAdd(&st[3], 1ull);
    if (s == 1)
        atomicAdd(&st[0], 1ull);
    else
        atomicAdd(&st[1], 1ull);
    atomicMax(&st[2], (unsigned long long)s);
}

// =============================================================================
// Host side
// =============================================================================

struct gs_universe;
struct AdjState;
struct BondState;
static void adj_free(gs_universe* u);
static void bond_free(gs_universe* u);

struct gs_universe {
    gs_params prm;
    KParams kp;
    int n = 0;
    int dev = 0;
    hipStream_t stream = nullptr;
    bool have_state = false;
    bool small_box = false;
    bool extras_valid = false;
    bool coul = false, shared_switch = false, single = false;
    int cur = 0;
    // resident state (double-buffered where the rebuild permutes)
    double4* pos4[2] = {nullptr, nullptr};
    double* v[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    uint8_t* type[2] = {nullptr, nullptr};
    uint32_t* id[2] = {nullptr, nullptr};
    double* a[3] = {nullptr, nullptr, nullptr};
    // cell list
    uint32_t *key = nullptr, *ticket = nullptr, *count = nullptr, *cstart = nullptr, *perm = nullptr;
    uint32_t* scan_tmp = nullptr;  // multi-block scan scratch (scan_tmp_words of the largest grid)
    size_t scan_tmp_cap = 0;
    // coefficients, reductions, scalars
    PairCoef* coef = nullptr;
    double* vpart = nullptr;
    uint32_t nvpart = 0;
    uint32_t npart = 0;  // slots of vpart and kpart: every pair pass's workgroups
    double* kpart = nullptr;
    uint32_t* counter = nullptr;
    double* scal = nullptr;  // SCAL_WORDS: [0] lambda, [1] V, [2] K, [3] V_bond, [4] fault word
    double *pe = nullptr, *sabs = nullptr;
    double* stage = nullptr;  // 9n doubles
    uint32_t* stage_u = nullptr;
    gs_observables* obs = nullptr;
    int64_t obs_cap = 0;
    // lens scratch (allocated on first use)
    double4* l_pts = nullptr;
    uint8_t* l_type = nullptr;
    uint32_t *l_key = nullptr, *l_ticket = nullptr, *l_perm = nullptr, *l_count = nullptr, *l_start = nullptr;
    uint32_t* l_id = nullptr;  // entity ids in lens-grid order
    uint32_t l_ncell_cap = 0;
    uint32_t* l_parent = nullptr;
    uint32_t* l_size = nullptr;
    unsigned long long* l_bins = nullptr;
    int l_bins_cap = 0;
    unsigned long long* l_u64 = nullptr;  // [0..3] FoF stats, [4] bound pairs
    double* l_sigma = nullptr;
    bool present[GS_NSPECIES] = {false};
    std::vector<uint8_t> species_h;  // host copies of the last upload
    std::vector<double> charge_h;
    // Verlet-list path
    KParams kpl;                // grid of edge >= rc + skin
    bool list_ok = false;       // the list geometry admits >= 3 cells per axis
    bool list_valid = false;    // lists match the current sorted order
    int pair_mode = 1;          // 0 = per-tick cell tiles, 1 = Verlet list, 2 = cell-staged list
    int n_cu = 0;               // compute units of the device
    bool prebin = true;         // predictive binning in the list integrate kernel
    int logT = 2;               // lanes per atom = 1 << logT
    double skin = 0.0;
    float* list_rbt = nullptr;  // squared bind radii per species pair (split lists)
    float list_rbt_host[GS_NSPECIES * GS_NSPECIES] = {};
    int64_t sort_every = 100;   // ticks between full re-sorts of the state in list mode
    uint64_t order_gen = 0;     // changes whenever the resident order (id[cur]) may have changed
    int64_t list_age = 0;       // ticks since the last full re-sort
    uint32_t nbr_cap = 0;       // list capacity per atom (entries)
    uint32_t nbr_have = 0;      // capacity the per-atom list array holds
    uint32_t* nbr = nullptr;
    // cell-staged lists: per-cell regions of 16-bit entries, ordered stencil
    // candidates and {candidates, home start, home count} per cell
    uint16_t* nbr16 = nullptr;
    size_t nbr16_have = 0;
    uint32_t* cand_store = nullptr;
    uint4* cinfo = nullptr;
    uint32_t cell_have = 0;
    uint2* hrec = nullptr;      // per cell and home slot: (species << 28) | atom, list length
    size_t hrec_have = 0;
    uint32_t* nlen = nullptr;
    double4* xref = nullptr;
    double4* xbin = nullptr;    // positions binned for a build one tick behind
    uint32_t* flags = nullptr;  // device list flags and counters (k_list_integrate)
    uint32_t* bcount = nullptr;  // list-grid cell buckets of a rebuild
    uint32_t* bucket = nullptr;
    uint32_t bcap = 0;
    int64_t list_builds = 0;
    int64_t list_replays = 0;   // calls replayed from the snapshot
    // snapshot for replays (allocated on first list-mode step)
    double4* snap_pos = nullptr;
    double* snap_v[3] = {nullptr, nullptr, nullptr};
    double* snap_a[3] = {nullptr, nullptr, nullptr};
    uint8_t* snap_type = nullptr;
    uint32_t* snap_id = nullptr;
    // observed bonds (gs_bonds.inc)
    BondState* bonds = nullptr;
    double* l_thr = nullptr;      // bound thresholds k_bind sigma_ab
    int64_t tick = 0;             // ctx.time.tick of the last tick
    double* ab_live = nullptr;    // this tick's bond accelerations (list path)
    double* vb_live = nullptr;    // this tick's V_bond partials
    uint32_t nvb_live = 0;
    // differentiable trajectory (gs_adjoint.inc)
    AdjState* adj = nullptr;
    uint8_t* adj_keep = nullptr;
    AdjCoef* adj_coef = nullptr;
    double* adj_qs = nullptr;
    double* adj_acc = nullptr;
    double* adj_gatom = nullptr;
    unsigned long long* adj_hist = nullptr;  // smoothed g(r): fixed-point bin sums
    size_t adj_nbins_cap = 0;
    double* adj_ck = nullptr;
    double* adj_l[3] = {nullptr, nullptr, nullptr};
    double* adj_o[3] = {nullptr, nullptr, nullptr};
    double* adj_lv[3] = {nullptr, nullptr, nullptr};
    double* adj_x[3] = {nullptr, nullptr, nullptr};
    // event timing of gs_step
    int timing = 0;  // 0 off, 1 pair kernel only, 2 every phase
    int timing_every = 1;  // time every k-th tick of a call
    std::vector<hipEvent_t> ev;
    gs_timing tm = {0, 0.0, 0.0, 0.0, 0.0, 0};
    // bound-pair search fused into the list pass (bond observation)
    bool bound_fused = false;    // this attempt of a gs_step call uses it
    bool bound_unfused = false;  // a hit column overflowed: the call replays on the aux grid
    bool bound_aux_only = false;  // gs_set_bound_search(u, 1)
    double bound_cd2 = 0.0;      // cutoff^2 of the bound-pair search
    BoundScreen bound_scr;       // the fused search's screen thresholds
};

namespace {

// observed bonds (gs_bonds.inc)
int bond_force_pass(gs_universe* u, bool accumulate);
int bond_observe_pass(gs_universe* u, uint64_t now, bool fused = false);
int bond_fused_prepare(gs_universe* u, bool list);
BoundArgs bond_fused_args(gs_universe* u);
int bond_call_begin(gs_universe* u);
int bond_call_restore(gs_universe* u, bool existed);
int bond_refresh_species(gs_universe* u, const uint8_t* species);
int bond_overflow_fetch(gs_universe* u, unsigned long long* host_word);

int validate_params(const gs_params* p, int64_t n) {
    if (!p) return fail(GS_E_INVALID, "params is NULL");
    if (n < 0 || n > (int64_t)0x7FFFFFFF) return fail(GS_E_INVALID, "n_atoms out of range");
    for (int k = 0; k < 3; ++k)
        if (!(p->box[k] > 0.0) || !std::isfinite(p->box[k])) return fail(GS_E_INVALID, "box must be finite and > 0");
    if (!std::isfinite(p->dt)) return fail(GS_E_INVALID, "dt must be finite");
    for (int s = 0; s < GS_NSPECIES; ++s) {
        if (!(p->sigma[s] > 0.0) || !std::isfinite(p->sigma[s])) return fail(GS_E_INVALID, "sigma must be > 0");
        if (!(p->epsilon_k[s] >= 0.0) || !std::isfinite(p->epsilon_k[s]))
            return fail(GS_E_INVALID, "epsilon_k must be >= 0");
        if (!(p->mass[s] > 0.0) || !std::isfinite(p->mass[s])) return fail(GS_E_INVALID, "mass must be > 0");
    }
    if (!(p->thermal_constant > 0.0)) return fail(GS_E_INVALID, "thermal_constant must be > 0");
    if (p->grid_semantics != GS_GRID_REFERENCE && p->grid_semantics != GS_GRID_MINIMAGE)
        return fail(GS_E_INVALID, "unknown grid_semantics");
    return GS_OK;
}

// Host part of LjTable::new and ForceSystem::new
// (src/physics/forces.rs:223-237, src/systems/forces.rs:100-142).
std::vector<PairCoef> derive(gs_universe* u) {
    const gs_params& p = u->prm;
    KParams& k = u->kp;
    std::memset(&k, 0, sizeof(k));
    double max_sigma = 0.0;
    for (int s = 0; s < GS_NSPECIES; ++s) max_sigma = std::fmax(max_sigma, p.sigma[s]);
    double max_pair_sigma = 0.0;
    std::vector<PairCoef> coef(GS_NSPECIES * GS_NSPECIES);
    const double c2m13 = std::pow(2.0, -1.0 / 3.0), c2p16 = std::pow(2.0, 1.0 / 6.0);
    double max_rcore = 0.0;
    for (int i = 0; i < GS_NSPECIES; ++i)
        for (int j = 0; j < GS_NSPECIES; ++j) {
            const double sig = 0.5 * (p.sigma[i] + p.sigma[j]);
            const double eps = p.thermal_constant * std::sqrt(p.epsilon_k[i] * p.epsilon_k[j]);
            max_pair_sigma = std::fmax(max_pair_sigma, sig);
            PairCoef& c = coef[i * GS_NSPECIES + j];
            c.sigma = sig;
            c.eps4 = 4.0 * eps;
            c.c24 = 24.0 * eps / sig;
            c.rcore = 0.5 * sig;
            c.inv_rcore = 1.0 / c.rcore;
            c.mm = p.mass[i] * p.mass[j];
            c.bond_k = 72.0 * c2m13 * eps / (sig * sig);
            c.bond_req = sig * c2p16;
            max_rcore = std::fmax(max_rcore, c.rcore);
        }
    const double lj_rc = std::fmax(p.lj_cutoff_factor, 1.0) * max_pair_sigma;
    const double c_rc = p.enable_electrostatics ? std::fmax(p.coulomb_cutoff, 1.0) * max_sigma : 0.0;
    const double g_rc = p.enable_gravity ? std::fmax(p.gravity_cutoff, 1.0) * max_sigma : 0.0;
    const double rc = std::fmax(std::fmax(lj_rc, c_rc), g_rc);
    for (int a = 0; a < 3; ++a) {
        k.box[a] = p.box[a];
        k.half[a] = p.box[a] * 0.5;
        k.inv_half[a] = 1.0 / k.half[a];
        double nc = std::floor(p.box[a] / rc);
        if (!(nc >= 1.0)) nc = 1.0;
        if (nc > 1024.0) nc = 1024.0;
        k.dims[a] = (uint32_t)nc;
        k.cell[a] = p.box[a] / (double)k.dims[a];
    }
    k.ncell = k.dims[0] * k.dims[1] * k.dims[2];
    k.n = u->n;
    k.dt = p.dt;
    k.half_dt = 0.5 * p.dt;
    k.rc = rc;
    k.rc2 = rc * rc;
    k.eps2 = F64_EPS * F64_EPS;
    k.lj_rc = lj_rc;
    k.lj_ron = 0.9 * lj_rc;
    k.lj_inv_w = 1.0 / (lj_rc - k.lj_ron);
    k.c_rc = c_rc;
    k.c_ron = 0.9 * c_rc;
    k.c_inv_w = c_rc > 0 ? 1.0 / (c_rc - k.c_ron) : 0.0;
    k.ke = p.enable_electrostatics ? p.coulomb_constant : 0.0;
    k.g_rc = g_rc;
    k.g_ron = 0.9 * g_rc;
    k.g_inv_w = g_rc > 0 ? 1.0 / (g_rc - k.g_ron) : 0.0;
    k.gconst = p.gravity_constant;
    for (int s = 0; s < GS_NSPECIES; ++s) {
        k.mass[s] = p.mass[s];
        k.inv_mass[s] = 1.0 / p.mass[s];
    }
    int flags = 0;
    if (p.enable_electrostatics) flags |= F_COULOMB;
    if (p.enable_gravity) flags |= F_GRAVITY;
    if (p.enable_boundaries) flags |= F_WRAP;
    if (p.grid_semantics == GS_GRID_REFERENCE) flags |= F_REF_MULT;
    u->coul = p.enable_electrostatics != 0;
    // The Coulomb term may reuse the LJ switch when both laws share the
    // cutoff and neither core clamp can reach the switching shell.
    u->shared_switch = u->coul && c_rc == lj_rc && max_rcore < k.lj_ron && 0.5 < k.lj_ron;
    if (u->shared_switch) flags |= F_SHARED_SWITCH;
    k.flags = flags;
    u->small_box = k.dims[0] < 3 || k.dims[1] < 3 || k.dims[2] < 3;
    // Single-species fast path: every pair coefficient equal.
    u->single = !u->coul && !p.enable_gravity;
    if (u->single)
        for (int s = 1; s < GS_NSPECIES * GS_NSPECIES; ++s)
            if (std::memcmp(&coef[s], &coef[0], sizeof(PairCoef)) != 0) {
                u->single = false;
                break;
            }
    static_assert(sizeof(PairCoef) == 64, "PairCoef layout");
    coef.resize(COEF_LEN);
    SpCoef* sp = reinterpret_cast<SpCoef*>(&coef[COEF_SP_OFFSET]);
    for (int j = 0; j < GS_NSPECIES; ++j) {
        sp[j].sigma = p.sigma[j];
        sp[j].a4 = std::sqrt(4.0 * p.thermal_constant * p.epsilon_k[j]);
    }
    u->nvpart = u->small_box ? nblocks(u->n, BRUTE_WAVES) : k.ncell;
    if (u->nvpart == 0) u->nvpart = 1;
    // the tiles (a cell each), the lists (N << MAX_LOGT lanes at most) and
    // the cell-staged pass (a list cell each, eight slabs)
    u->npart = std::max<uint32_t>(std::max<uint32_t>(u->nvpart, nblocks((int64_t)u->n << MAX_LOGT)), k.ncell + 8u);
    return coef;
}

template <typename T>
int dalloc(T** p, size_t count) {
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) return fail(GS_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    return GS_OK;
}

// Scan scratch for grids of up to ncell cells.
int scan_reserve(uint32_t** tmp, size_t* cap, uint32_t ncell) {
    const size_t need = scan_tmp_words(ncell);
    if (*cap >= need) return GS_OK;
    (void)hipFree(*tmp);
    *tmp = nullptr;
    *cap = 0;
    int rc;
    if ((rc = dalloc(tmp, need))) return rc;
    *cap = need;
    return GS_OK;
}

int check_state(gs_universe* u) {
    if (!u) return fail(GS_E_INVALID, "universe handle is NULL");
    if (!u->have_state) return fail(GS_E_STATE, "no state uploaded");
    return GS_OK;
}

// Cell-list rebuild: key + histogram (optionally fused with kick/drift/wrap),
// scan, scatter, per-cell id ordering + gather into the other buffer.
int rebuild(gs_universe* u, bool kick_drift, const KParams& P) {
    const int n = u->n;
    if (n == 0) return GS_OK;
    const int c = u->cur, o = 1 - c;
    if (kick_drift)
        k_integrate_key<true><<<nblocks(n), BLOCK, 0, u->stream>>>(P, u->pos4[c], u->v[c][0], u->v[c][1], u->v[c][2],
                                                                  u->a[0], u->a[1], u->a[2], u->scal, u->key,
                                                                  u->ticket, u->count);
    else
        k_integrate_key<false><<<nblocks(n), BLOCK, 0, u->stream>>>(P, u->pos4[c], u->v[c][0], u->v[c][1],
                                                                   u->v[c][2], u->a[0], u->a[1], u->a[2], u->scal,
                                                                   u->key, u->ticket, u->count);
    launch_scan(u->stream, u->count, u->cstart, P.ncell, u->scan_tmp_cap >= scan_tmp_words(P.ncell) ? u->scan_tmp : nullptr);
    k_scatter<<<nblocks(n), BLOCK, 0, u->stream>>>(n, u->key, u->ticket, u->cstart, u->perm);
    k_sort_gather<<<nblocks(P.ncell, BLOCK / WAVE), BLOCK, 0, u->stream>>>(
        P.ncell, u->cstart, u->perm, u->pos4[c], u->v[c][0], u->v[c][1], u->v[c][2], u->type[c], u->id[c],
        u->pos4[o], u->v[o][0], u->v[o][1], u->v[o][2], u->type[o], u->id[o]);
    u->cur = o;
    ++u->order_gen;
    HIPCHK(hipGetLastError());
    return GS_OK;
}

// Launches a kernel on the universe's stream; with e0 / e1 the events are
// bound to the dispatch itself (hipExtLaunchKernelGGL), so their interval is
// the kernel's execution without the dispatch gap an event record adds.
template <typename K, typename... A>
void launch_timed(K kernel, dim3 grid, dim3 block, hipStream_t stream, hipEvent_t e0, hipEvent_t e1, A... args) {
    if (e0)
        hipExtLaunchKernelGGL(kernel, grid, block, 0, stream, e0, e1, 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, 0, stream, args...);
}

template <bool EXTRAS>
int launch_pair(gs_universe* u, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    const KParams& P = u->kp;
    const int c = u->cur;
    if (u->small_box) {
        const dim3 grid(nblocks(u->n, BRUTE_WAVES)), blk(BRUTE_WAVES * WAVE);
        const KParams* nt = nullptr;
        const BlockRef* nb = nullptr;
        const uint32_t* no = nullptr;
        if (u->coul)
            launch_timed(k_pair_brute<true, EXTRAS, false>, grid, blk, u->stream, e0, e1, P,
                         (const double4*)u->pos4[c], (const uint8_t*)u->type[c], (const PairCoef*)u->coef, u->a[0],
                         u->a[1], u->a[2], u->pe, u->sabs, u->vpart, nt, nb, no);
        else
            launch_timed(k_pair_brute<false, EXTRAS, false>, grid, blk, u->stream, e0, e1, P,
                         (const double4*)u->pos4[c], (const uint8_t*)u->type[c], (const PairCoef*)u->coef, u->a[0],
                         u->a[1], u->a[2], u->pe, u->sabs, u->vpart, nt, nb, no);
    } else {
    // a grid of few cells gets 16 waves per cell (their home atoms shared
    // among them), so the device holds several waves per SIMD
    const bool wide = P.ncell < 2u * (uint32_t)std::max(1, u->n_cu);
#define GS_PAIR_NW(C, S, E, SG, NW)                                                                               \
    launch_timed(k_pair<C, S, E, SG, false, NW>, dim3(P.ncell), dim3(NW * WAVE), u->stream, e0, e1, P,            \
                 (const uint32_t*)u->cstart, (const double4*)u->pos4[c], (const uint8_t*)u->type[c],             \
                 (const PairCoef*)u->coef, u->a[0], u->a[1], u->a[2], u->pe, u->sabs, u->vpart,                   \
                 (const KParams*)nullptr, (const uint16_t*)nullptr, (const uint32_t*)nullptr)
#define GS_PAIR(C, S, E, SG)                    \
    do {                                        \
        if (wide)                               \
            GS_PAIR_NW(C, S, E, SG, 16);        \
        else                                    \
            GS_PAIR_NW(C, S, E, SG, PAIR_WAVES); \
    } while (0)
        if (u->single)
            GS_PAIR(false, false, EXTRAS, true);
        else if (!u->coul)
            GS_PAIR(false, false, EXTRAS, false);
        else if (u->shared_switch)
            GS_PAIR(true, true, EXTRAS, false);
        else
            GS_PAIR(true, false, EXTRAS, false);
#undef GS_PAIR
#undef GS_PAIR_NW
    }
    HIPCHK(hipGetLastError());
    return GS_OK;
}


// Geometry of the list grid for a skin (cells >= rc + skin, >= 3 per axis).
bool list_geometry(gs_universe* u, double skin) {
    KParams q = u->kp;
    const double rl = u->kp.rc + skin;
    for (int a = 0; a < 3; ++a) {
        double nc = std::floor(q.box[a] / rl);
        if (!(nc >= 3.0)) return false;
        if (nc > 1024.0) nc = 1024.0;
        q.dims[a] = (uint32_t)nc;
        q.cell[a] = q.box[a] / (double)q.dims[a];
    }
    q.ncell = q.dims[0] * q.dims[1] * q.dims[2];
    if (q.ncell > u->kp.ncell) return false;  // count/cstart are sized for the force grid
    u->kpl = q;
    u->skin = skin;
    return true;
}

// Entries per cell of the cell-staged lists: a row of cap per bucket slot.
size_t cell_region(const gs_universe* u) { return (size_t)u->bcap * (size_t)u->nbr_cap; }

// Inner radius of the split Verlet lists: the smallest switching radius in
// use minus the skin.  An entry built within it stays inside every switch
// while its list is valid (each end moves less than skin/2), 0 when none
// can.
double list_inner(const gs_universe* u) {
    const KParams& P = u->kp;
    double r = P.lj_ron;
    if (u->coul && !u->shared_switch) r = std::min(r, P.c_ron);
    if (u->prm.enable_gravity) r = std::min(r, P.g_ron);
    r -= u->skin;
    return r > 0.0 ? r : 0.0;
}
double list_inner2(const gs_universe* u) {
    const double r = list_inner(u);
    return r * r;
}
// Bind radii of the split lists per species pair: the pair's bond cutoff
// k_bind sigma_ab (src/analysis/pairs.rs:132-174) plus the skin (a pair built
// farther cannot be bound while its list is valid), squared and widened by
// the build's FP32 margin; 0 without bond observation.  The device copy
// list_rbt is uploaded whenever the table differs from the last upload.
int list_bind_table(gs_universe* u) {
    float t[GS_NSPECIES * GS_NSPECIES] = {};
    if (u->prm.enable_bond_observation)
        for (int a = 0; a < GS_NSPECIES; ++a)
            for (int b = 0; b < GS_NSPECIES; ++b) {
                const double rb = u->prm.bond_k_bind * (0.5 * (u->prm.sigma[a] + u->prm.sigma[b])) * (1.0 + 1e-9) +
                                  u->skin;
                t[a * GS_NSPECIES + b] = (float)(rb * rb * (1.0 + 1e-5));
            }
    if (u->list_rbt && std::memcmp(t, u->list_rbt_host, sizeof(t)) == 0) return GS_OK;
    int rc;
    if (!u->list_rbt && (rc = dalloc(&u->list_rbt, GS_NSPECIES * GS_NSPECIES))) return rc;
    std::memcpy(u->list_rbt_host, t, sizeof(t));
    HIPCHK(hipMemcpyAsync(u->list_rbt, u->list_rbt_host, sizeof(t), hipMemcpyHostToDevice, u->stream));
    HIPCHK(hipStreamSynchronize(u->stream));
    return GS_OK;
}

int list_alloc(gs_universe* u, uint32_t cap, uint32_t bcap_min = 0) {
    int rc;
    const size_t n = (size_t)std::max(1, u->n);
    if (!u->nlen) {
        if ((rc = dalloc(&u->nlen, n)) || (rc = dalloc(&u->xref, n)) || (rc = dalloc(&u->xbin, n)) ||
            (rc = dalloc(&u->flags, LIST_FLAGS)) ||
            (rc = dalloc(&u->snap_pos, n)) || (rc = dalloc(&u->snap_type, n)) ||
            (rc = dalloc(&u->snap_id, n)))
            return rc;
        for (int a = 0; a < 3; ++a)
            if ((rc = dalloc(&u->snap_v[a], n)) || (rc = dalloc(&u->snap_a[a], n))) return rc;
        HIPCHK(hipMemsetAsync(u->flags, 0, LIST_FLAGS * sizeof(uint32_t), u->stream));
        // an atom whose list was never built has an empty one
        HIPCHK(hipMemsetAsync(u->nlen, 0, sizeof(uint32_t) * n, u->stream));
    }
    // bucket capacity: twice the mean list-cell occupancy plus 32, doubled
    // on overflow; the force grid bounds the list grid's cell count
    const uint32_t want = std::max<uint32_t>(
        bcap_min, (uint32_t)(2.0 * (double)u->n / (double)std::max<uint32_t>(1, u->kpl.ncell)) + 32u);
    if (u->bcap < want) {
        (void)hipFree(u->bucket);
        u->bucket = nullptr;
        u->bcap = 0;
        if (!u->bcount) {
            if ((rc = dalloc(&u->bcount, (size_t)u->kp.ncell))) return rc;
            // binning fills empty buckets
            HIPCHK(hipMemsetAsync(u->bcount, 0, sizeof(uint32_t) * u->kp.ncell, u->stream));
        }
        if ((rc = dalloc(&u->bucket, (size_t)u->kp.ncell * want))) return rc;
        u->bcap = want;
    }
    if (cap > u->nbr_cap) u->nbr_cap = cap;
    cap = u->nbr_cap;
    if (u->pair_mode == 2) {
        const uint32_t cells = u->kpl.ncell;
        if (u->cell_have < cells) {
            (void)hipFree(u->cand_store);
            (void)hipFree(u->cinfo);
            u->cand_store = nullptr;
            u->cinfo = nullptr;
            u->cell_have = 0;
            if ((rc = dalloc(&u->cand_store, (size_t)cells * LDS_IDX_CAP_LOCAL)) || (rc = dalloc(&u->cinfo, cells)))
                return rc;
            // a cell never built has no candidates
            HIPCHK(hipMemsetAsync(u->cinfo, 0, sizeof(uint4) * cells, u->stream));
            u->cell_have = cells;
        }
        const size_t nrec = (size_t)cells * u->bcap;
        if (u->hrec_have < nrec) {
            (void)hipFree(u->hrec);
            u->hrec = nullptr;
            u->hrec_have = 0;
            if ((rc = dalloc(&u->hrec, nrec))) return rc;
            u->hrec_have = nrec;
        }
        const size_t need = (size_t)cells * cell_region(u);
        if (u->nbr16_have < need) {
            (void)hipFree(u->nbr16);
            u->nbr16 = nullptr;
            u->nbr16_have = 0;
            if ((rc = dalloc(&u->nbr16, need))) return rc;
            u->nbr16_have = need;
        }
    } else if (u->nbr_have < cap) {
        (void)hipFree(u->nbr);
        u->nbr = nullptr;
        u->nbr_have = 0;
        // groups of 4 entries per lane: (cap + 8 T) entries per atom hold
        // the padded lists (each lane's last group and one more) of any
        // lanes-per-atom setting up to 16
        if ((rc = dalloc(&u->nbr, ((size_t)cap + (8u << MAX_LOGT)) * n))) return rc;
        u->nbr_have = cap;
    }
    return GS_OK;
}

int snapshot(gs_universe* u, bool restore) {
    const int c = u->cur;
    const size_t n = (size_t)u->n, b = n * sizeof(double);
    auto cp = [&](void* live, void* snap, size_t bytes) {
        return restore ? hipMemcpyAsync(live, snap, bytes, hipMemcpyDeviceToDevice, u->stream)
                       : hipMemcpyAsync(snap, live, bytes, hipMemcpyDeviceToDevice, u->stream);
    };
    HIPCHK(cp(u->pos4[c], u->snap_pos, n * sizeof(double4)));
    for (int a = 0; a < 3; ++a) {
        HIPCHK(cp(u->v[c][a], u->snap_v[a], b));
        HIPCHK(cp(u->a[a], u->snap_a[a], b));
    }
    HIPCHK(cp(u->type[c], u->snap_type, n));
    HIPCHK(cp(u->id[c], u->snap_id, n * sizeof(uint32_t)));
    if (restore) ++u->order_gen;
    return GS_OK;
}

// List pair pass of one tick, fused with the second kick, the kinetic
// reduction and the Berendsen lambda; obs (device) receives the observables.
// bound (may be NULL): the fused bound-pair search, for 2 or 4 lanes per atom.
int launch_pair_list(gs_universe* u, gs_observables* obs, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr,
                     const BoundArgs* bound = nullptr) {
    const KParams& P = u->kp;
    const gs_params& p = u->prm;
    const int c = u->cur;
    const uint32_t N = (uint32_t)u->n;
    const unsigned nb = nblocks((int64_t)N << u->logT);
    // CU-aware chunk order for a grid of whole workgroups per CU
    const uint32_t rpc = (u->n_cu == 256 && nb % 256u == 0u) ? nb / 256u : 0u;
    const ThermoArgs th{p.enable_thermostat != 0 ? 1 : 0, p.thermostat_temperature, std::fmax(p.thermostat_tau, 1.0),
                        p.thermal_constant};
#define GS_PL(L, C, S, SG, G, BD)                                                                             \
    launch_timed(k_pair_list<L, C, S, SG, G, false, BD>, dim3(nb), dim3(BLOCK), u->stream, e0, e1, P, N,            \
                 (const double4*)u->pos4[c], (const uint8_t*)u->type[c], (const uint32_t*)u->nbr,               \
                 (const uint32_t*)u->nlen, (const PairCoef*)u->coef, u->a[0], u->a[1], u->a[2], u->v[c][0],      \
                 u->v[c][1], u->v[c][2], u->vpart, u->kpart, u->flags, u->scal, obs, th,                          \
                 (const double*)u->ab_live, (const double*)u->vb_live, u->nvb_live, rpc, u->bcount, u->kpl.ncell, \
                 BatchTabs{}, (double*)nullptr, bound ? *bound : BoundArgs{})
#define GS_PL_T(C, S, SG, G)                                  \
    if (bound) {                                              \
        if (u->logT == 1)                                     \
            GS_PL(1, C, S, SG, G, true);                      \
        else                                                  \
            GS_PL(2, C, S, SG, G, true);                      \
    } else                                                    \
        switch (u->logT) {                                    \
            case 0: GS_PL(0, C, S, SG, G, false); break;      \
            case 1: GS_PL(1, C, S, SG, G, false); break;      \
            case 2: GS_PL(2, C, S, SG, G, false); break;      \
            case 3: GS_PL(3, C, S, SG, G, false); break;      \
            default: GS_PL(4, C, S, SG, G, false); break;     \
        }
    const bool grav = u->prm.enable_gravity != 0;
    if (u->single) {
        GS_PL_T(false, false, true, false);
    } else if (!u->coul) {
        if (grav) {
            GS_PL_T(false, false, false, true);
        } else {
            GS_PL_T(false, false, false, false);
        }
    } else if (u->shared_switch) {
        if (grav) {
            GS_PL_T(true, true, false, true);
        } else {
            GS_PL_T(true, true, false, false);
        }
    } else {
        if (grav) {
            GS_PL_T(true, false, false, true);
        } else {
            GS_PL_T(true, false, false, false);
        }
    }
#undef GS_PL_T
#undef GS_PL
    HIPCHK(hipGetLastError());
    return GS_OK;
}

// Cell-staged pair pass of one tick (pair_mode 2), fused like launch_pair_list.
int launch_pair_cell(gs_universe* u, gs_observables* obs, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    const KParams& P = u->kp;
    const gs_params& p = u->prm;
    const int c = u->cur;
    const uint32_t ncell = u->kpl.ncell, per = (ncell + 7u) / 8u;
    const unsigned nb = 8u * per;
    if (nb > u->npart) return fail(GS_E_STATE, "cell-staged pass: more workgroups than partial slots");
    const size_t region = cell_region(u);
    const ThermoArgs th{p.enable_thermostat != 0 ? 1 : 0, p.thermostat_temperature, std::fmax(p.thermostat_tau, 1.0),
                        p.thermal_constant};
#define GS_PC(C, S, SG, G)                                                                                     \
    launch_timed(k_pair_cell<C, S, SG, G>, dim3(nb), dim3(CELL_THREADS), u->stream, e0, e1, P, ncell, per,          \
                 (const double4*)u->pos4[c], (const uint16_t*)u->nbr16, region, u->nbr_cap, (const uint2*)u->hrec, \
                 u->bcap,                                                                                           \
                 (const uint32_t*)u->cand_store, (const uint4*)u->cinfo, (const PairCoef*)u->coef, u->a[0], u->a[1], u->a[2], u->v[c][0], u->v[c][1], u->v[c][2],         \
                 u->vpart, u->kpart, u->flags, u->scal, obs, th, (const double*)u->ab_live,                       \
                 (const double*)u->vb_live, u->nvb_live, (uint32_t)u->n, u->bcount)
    const bool grav = u->prm.enable_gravity != 0;
    if (u->single)
        GS_PC(false, false, true, false);
    else if (!u->coul)
        grav ? GS_PC(false, false, false, true) : GS_PC(false, false, false, false);
    else if (u->shared_switch)
        grav ? GS_PC(true, true, false, true) : GS_PC(true, true, false, false);
    else
        grav ? GS_PC(true, false, false, true) : GS_PC(true, false, false, false);
#undef GS_PC
    HIPCHK(hipGetLastError());
    return GS_OK;
}

// Start of a list-mode tick.  A full re-sort (every sort_every ticks, or
// when the lists do not match the resident order) runs the per-tick rebuild
// on the list grid, which permutes the state into cell order, and forces a
// list build over the identity order.  Otherwise the device integrates,
// tests the displacements and rebuilds the lists only when an atom moved
// beyond skin/2.
int list_tick_start(gs_universe* u) {
    const uint32_t N = (uint32_t)u->n;
    const double rl = u->kp.rc + u->skin;
    const KParams& PL = u->kpl;
    int rc;
    if (u->pair_mode == 1 && (rc = list_bind_table(u))) return rc;
    if (!u->list_valid || u->list_age >= u->sort_every) {
        if ((rc = rebuild(u, true, PL))) return rc;
        k_list_fill_buckets<<<nblocks(PL.ncell, BLOCK / WAVE), BLOCK, 0, u->stream>>>(PL.ncell, u->cstart, u->bcount,
                                                                                      u->bucket, u->bcap, u->flags);
        k_flag_set<<<1, 1, 0, u->stream>>>(u->flags, 0, 1u);
        // fresh lists: no binning or delayed build pending
        HIPCHK(hipMemsetAsync(u->flags + 8, 0, 4 * sizeof(uint32_t), u->stream));
        u->list_valid = true;
        u->list_age = 0;
    } else if (u->prebin && u->pair_mode == 1) {
        // integrate, and the build one tick behind, in one launch
        const int c = u->cur;
        const uint32_t natom = nblocks(u->n);
        const uint32_t nbuild = u->n_cu > 0 ? std::min<uint32_t>(PL.ncell, 4u * (uint32_t)u->n_cu) : PL.ncell;
        k_list_step<<<natom + nbuild, BLOCK, 0, u->stream>>>(
            u->kp, PL, u->pos4[c], u->v[c][0], u->v[c][1], u->v[c][2], u->a[0], u->a[1], u->a[2], u->scal, u->xref,
            u->xbin, 0.5 * u->skin, u->bcount, u->bucket, u->bcap, u->flags, natom, u->type[c], u->nbr, u->nlen, N,
            (uint32_t)u->logT, u->nbr_cap, rl * rl, list_inner2(u), u->list_rbt);
        u->list_age += 1;
        HIPCHK(hipGetLastError());
        return GS_OK;
    } else {
        const int c = u->cur;
        const double h = 0.5 * u->skin;
        if (u->prebin) {
            k_list_integrate<true><<<nblocks(u->n), BLOCK, 0, u->stream>>>(
                u->kp, PL, u->pos4[c], u->v[c][0], u->v[c][1], u->v[c][2], u->a[0], u->a[1], u->a[2], u->scal,
                u->xref, h, u->bcount, u->bucket, u->bcap, u->flags);
        } else {
            k_list_integrate<false><<<nblocks(u->n), BLOCK, 0, u->stream>>>(
                u->kp, PL, u->pos4[c], u->v[c][0], u->v[c][1], u->v[c][2], u->a[0], u->a[1], u->a[2], u->scal,
                u->xref, h, u->bcount, u->bucket, u->bcap, u->flags);
            k_list_bin<<<nblocks(u->n), BLOCK, 0, u->stream>>>(PL, u->pos4[c], u->bcount, u->bucket, u->bcap,
                                                              u->flags);
        }
    }
    const int c = u->cur;
    const uint32_t bgrid = u->n_cu > 0 ? std::min<uint32_t>(PL.ncell, BUILD_WG_PER_CU * (uint32_t)u->n_cu) : PL.ncell;
    if (u->pair_mode == 2)
        k_list_build<true><<<bgrid, BUILD_THREADS, 0, u->stream>>>(
            PL, u->bcount, u->bucket, u->bcap, u->pos4[c], u->type[c], nullptr, u->nlen, u->xref, N, 0u, u->nbr_cap,
            rl * rl, u->flags, u->cand_store, u->cinfo, u->hrec, u->nbr16, cell_region(u), 0.0, nullptr);
    else
        k_list_build<false><<<bgrid, BUILD_THREADS, 0, u->stream>>>(
            PL, u->bcount, u->bucket, u->bcap, u->pos4[c], u->type[c], u->nbr, u->nlen, u->xref, N,
            (uint32_t)u->logT, u->nbr_cap, rl * rl, u->flags, nullptr, nullptr, nullptr, nullptr, 0, list_inner2(u),
            u->list_rbt);
    u->list_age += 1;
    HIPCHK(hipGetLastError());
    return GS_OK;
}

// The tick loop of gs_step: kick, drift, wrap, [cell list], forces, kick,
// Berendsen, in build_schedule order (src/universe/mod.rs:355-369).
int run_ticks(gs_universe* u, int64_t n_steps, bool observe, bool list) {
    int rc;
    const gs_params& p = u->prm;
    const bool thermo = p.enable_thermostat != 0;
    for (int64_t s = 0; s < n_steps; ++s) {
        const int tmode = (s % u->timing_every) == 0 ? u->timing : 0;
        const int64_t te = 4 * (s / u->timing_every);  // this timed tick's four events
        if (tmode > 1) HIPCHK(hipEventRecord(u->ev[te + 0], u->stream));
        const bool fused = list && u->bound_fused;
        if (list) {
            if ((rc = list_tick_start(u))) return rc;
            if ((rc = bond_force_pass(u, false))) return rc;
            hipEvent_t e0 = tmode ? u->ev[te + 1] : nullptr, e1 = tmode ? u->ev[te + 2] : nullptr;
            gs_observables* o = observe ? u->obs + s : nullptr;
            const BoundArgs ba = fused ? bond_fused_args(u) : BoundArgs{};
            if ((rc = u->pair_mode == 2 ? launch_pair_cell(u, o, e0, e1)
                                        : launch_pair_list(u, o, e0, e1, fused ? &ba : nullptr)))
                return rc;
            if (tmode > 1) HIPCHK(hipEventRecord(u->ev[te + 3], u->stream));
        } else {
            if ((rc = rebuild(u, true, u->kp))) return rc;
            u->list_valid = false;
            if ((rc = launch_pair<false>(u, tmode ? u->ev[te + 1] : nullptr, tmode ? u->ev[te + 2] : nullptr)))
                return rc;
            if ((rc = bond_force_pass(u, true))) return rc;
            const int c = u->cur;
            k_kick_ke<true><<<nblocks(u->n), BLOCK, 0, u->stream>>>(
                u->kp, u->v[c][0], u->v[c][1], u->v[c][2], u->a[0], u->a[1], u->a[2], u->type[c], u->kpart,
                u->counter, u->vpart, u->nvpart, u->vb_live, u->nvb_live, u->scal, observe ? u->obs + s : nullptr,
                thermo ? 1 : 0, p.thermostat_temperature, std::fmax(p.thermostat_tau, 1.0), p.thermal_constant);
            if (tmode > 1) HIPCHK(hipEventRecord(u->ev[te + 3], u->stream));
        }
        // BondObservationSystem runs after the thermostat (build_schedule,
        // src/universe/mod.rs:383-388); its Bonds feed the next tick.
        u->tick += 1;
        if (p.enable_bond_observation && (rc = bond_observe_pass(u, (uint64_t)u->tick, fused))) return rc;
    }
    HIPCHK(hipGetLastError());
    return GS_OK;
}

}  // namespace

// gs_noalloc.inc: address of the allocation-free entry point's scratch area on
// the current device, set up once per device.
static int noalloc_scratch_setup(int dev);

// =============================================================================
// C ABI
// =============================================================================
extern "C" {

const char* gs_last_error(void) { return g_err.c_str(); }

int gs_device_info(int* device, char* name, int name_len) {
    int d = 0;
    HIPCHK(hipGetDevice(&d));
    if (device) *device = d;
    if (name && name_len > 0) {
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, d));
        std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    return GS_OK;
}

int gs_create(const gs_params* params, int64_t n_atoms, gs_universe** out) {
    if (!out) return fail(GS_E_INVALID, "out is NULL");
    *out = nullptr;
    int rc = validate_params(params, n_atoms);
    if (rc) return rc;
    gs_universe* u = new gs_universe();
    u->n = (int)n_atoms;
    std::memcpy(&u->prm, params, sizeof(gs_params));
    std::vector<PairCoef> coef = derive(u);
    // Default list skin: 6% of the cutoff (adapted lifetime keeps it valid).
    u->list_ok = !u->small_box && n_atoms <= (int64_t)LIST_ATOMS_MAX && list_geometry(u, 0.06 * u->kp.rc);
    if (!u->list_ok) u->pair_mode = 0;
    (void)hipGetDevice(&u->dev);
    if ((rc = noalloc_scratch_setup(u->dev))) {
        delete u;
        return rc;
    }
    {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, u->dev) == hipSuccess) u->n_cu = ncu;
    }
    const size_t n = (size_t)std::max(1, u->n);
    const KParams& P = u->kp;
    auto bail = [&](int code) {
        gs_destroy(u);
        return code;
    };
    if (hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(GS_E_HIP, "hipStreamCreate failed"));
    for (int b = 0; b < 2; ++b) {
        if ((rc = dalloc(&u->pos4[b], n))) return bail(rc);
        for (int a = 0; a < 3; ++a)
            if ((rc = dalloc(&u->v[b][a], n))) return bail(rc);
        if ((rc = dalloc(&u->type[b], n))) return bail(rc);
        if ((rc = dalloc(&u->id[b], n))) return bail(rc);
    }
    for (int a = 0; a < 3; ++a)
        if ((rc = dalloc(&u->a[a], n))) return bail(rc);
    if ((rc = dalloc(&u->key, n)) || (rc = dalloc(&u->ticket, n)) || (rc = dalloc(&u->perm, n)) ||
        (rc = dalloc(&u->count, (size_t)P.ncell)) || (rc = dalloc(&u->cstart, (size_t)P.ncell + 1)) ||
        (rc = dalloc(&u->coef, coef.size())) || (rc = dalloc(&u->vpart, (size_t)u->npart)) ||
        (rc = dalloc(&u->kpart, (size_t)u->npart)) || (rc = dalloc(&u->counter, 4)) ||
        (rc = dalloc(&u->scal, (size_t)SCAL_WORDS)) || (rc = dalloc(&u->pe, n)) || (rc = dalloc(&u->sabs, n)) ||
        (rc = dalloc(&u->stage, 9 * n)) || (rc = dalloc(&u->stage_u, n)) ||
        (rc = scan_reserve(&u->scan_tmp, &u->scan_tmp_cap, P.ncell)))
        return bail(rc);
    if (hipMemsetAsync(u->count, 0, sizeof(uint32_t) * P.ncell, u->stream) != hipSuccess ||
        hipMemsetAsync(u->counter, 0, sizeof(uint32_t) * 4, u->stream) != hipSuccess ||
        hipMemsetAsync(u->vpart, 0, sizeof(double) * u->nvpart, u->stream) != hipSuccess ||
        hipMemcpyAsync(u->coef, coef.data(), coef.size() * sizeof(PairCoef), hipMemcpyHostToDevice, u->stream) !=
            hipSuccess)
        return bail(fail(GS_E_HIP, "device initialisation failed"));
    k_set_scalar<<<1, 1, 0, u->stream>>>(u->scal, 1.0);
    k_set_scalar<<<1, 1, 0, u->stream>>>(u->scal + SCAL_FAULT, 0.0);
    if (hipStreamSynchronize(u->stream) != hipSuccess) return bail(fail(GS_E_HIP, "device initialisation failed"));
    *out = u;
    return GS_OK;
}

int gs_destroy(gs_universe* u) {
    if (!u) return GS_OK;
    if (u->stream) (void)hipStreamSynchronize(u->stream);
    bond_free(u);
    for (int b = 0; b < 2; ++b) {
        (void)hipFree(u->pos4[b]);
        for (int a = 0; a < 3; ++a) (void)hipFree(u->v[b][a]);
        (void)hipFree(u->type[b]);
        (void)hipFree(u->id[b]);
    }
    for (int a = 0; a < 3; ++a) (void)hipFree(u->a[a]);
    (void)hipFree(u->key);
    (void)hipFree(u->ticket);
    (void)hipFree(u->perm);
    (void)hipFree(u->count);
    (void)hipFree(u->cstart);
    (void)hipFree(u->scan_tmp);
    (void)hipFree(u->coef);
    (void)hipFree(u->vpart);
    (void)hipFree(u->kpart);
    (void)hipFree(u->counter);
    (void)hipFree(u->scal);
    (void)hipFree(u->pe);
    (void)hipFree(u->sabs);
    (void)hipFree(u->stage);
    (void)hipFree(u->stage_u);
    (void)hipFree(u->obs);
    (void)hipFree(u->l_pts);
    (void)hipFree(u->l_id);
    (void)hipFree(u->l_type);
    (void)hipFree(u->l_key);
    (void)hipFree(u->l_ticket);
    (void)hipFree(u->l_perm);
    (void)hipFree(u->l_count);
    (void)hipFree(u->l_start);
    (void)hipFree(u->l_parent);
    (void)hipFree(u->l_size);
    (void)hipFree(u->l_bins);
    (void)hipFree(u->l_u64);
    (void)hipFree(u->l_sigma);
    for (hipEvent_t e : u->ev) (void)hipEventDestroy(e);
    (void)hipFree(u->nbr);
    (void)hipFree(u->nbr16);
    (void)hipFree(u->list_rbt);
    (void)hipFree(u->cand_store);
    (void)hipFree(u->cinfo);
    (void)hipFree(u->hrec);
    (void)hipFree(u->nlen);
    (void)hipFree(u->xref);
    (void)hipFree(u->xbin);
    (void)hipFree(u->flags);
    (void)hipFree(u->bcount);
    (void)hipFree(u->bucket);
    (void)hipFree(u->snap_pos);
    for (int a = 0; a < 3; ++a) (void)hipFree(u->snap_v[a]);
    for (int a = 0; a < 3; ++a) (void)hipFree(u->snap_a[a]);
    (void)hipFree(u->snap_type);
    (void)hipFree(u->snap_id);
    adj_free(u);
    if (u->stream) (void)hipStreamDestroy(u->stream);
    delete u;
    return GS_OK;
}

int gs_upload_state(gs_universe* u, const double* x, const double* y, const double* z, const double* vx,
                    const double* vy, const double* vz, const double* ax, const double* ay, const double* az,
                    const uint8_t* species, const double* charge) {
    if (!u) return fail(GS_E_INVALID, "universe handle is NULL");
    const int n = u->n;
    if (n > 0 && (!x || !y || !z || !vx || !vy || !vz || !species || !charge))
        return fail(GS_E_INVALID, "state arrays must not be NULL");
    std::vector<double4> p4((size_t)n);
    std::vector<uint32_t> id((size_t)n);
    bool present[GS_NSPECIES] = {false};
    for (int i = 0; i < n; ++i) {
        if (species[i] >= GS_NSPECIES) return fail(GS_E_INVALID, "species index out of range");
        present[species[i]] = true;
        if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(z[i]) || !std::isfinite(vx[i]) ||
            !std::isfinite(vy[i]) || !std::isfinite(vz[i]) || !std::isfinite(charge[i]))
            return fail(GS_E_NONFINITE, "non-finite value in the uploaded state");
        p4[i] = make_double4(x[i], y[i], z[i], charge[i]);
        id[i] = (uint32_t)i;
    }
    const int c = u->cur;
    const size_t b = sizeof(double) * (size_t)n;
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(u->pos4[c], p4.data(), sizeof(double4) * (size_t)n, hipMemcpyHostToDevice, u->stream));
        HIPCHK(hipMemcpyAsync(u->v[c][0], vx, b, hipMemcpyHostToDevice, u->stream));
        HIPCHK(hipMemcpyAsync(u->v[c][1], vy, b, hipMemcpyHostToDevice, u->stream));
        HIPCHK(hipMemcpyAsync(u->v[c][2], vz, b, hipMemcpyHostToDevice, u->stream));
        const double* acc[3] = {ax, ay, az};
        for (int a = 0; a < 3; ++a) {
            if (acc[a])
                HIPCHK(hipMemcpyAsync(u->a[a], acc[a], b, hipMemcpyHostToDevice, u->stream));
            else
                HIPCHK(hipMemsetAsync(u->a[a], 0, b, u->stream));
        }
        HIPCHK(hipMemcpyAsync(u->type[c], species, (size_t)n, hipMemcpyHostToDevice, u->stream));
        HIPCHK(hipMemcpyAsync(u->id[c], id.data(), sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice, u->stream));
    }
    k_set_scalar<<<1, 1, 0, u->stream>>>(u->scal, 1.0);
    HIPCHK(hipStreamSynchronize(u->stream));
    u->have_state = true;
    u->extras_valid = false;
    u->list_valid = false;
    u->species_h.assign(species, species + n);
    u->charge_h.assign(charge, charge + n);
    std::memcpy(u->present, present, sizeof(present));
    ++u->order_gen;
    return bond_refresh_species(u, species);
}

int gs_download_state(gs_universe* u, double* x, double* y, double* z, double* vx, double* vy, double* vz, double* ax,
                      double* ay, double* az) {
    int rc = check_state(u);
    if (rc) return rc;
    const int n = u->n;
    if (n == 0) return GS_OK;
    const int c = u->cur;
    k_unpermute_state<<<nblocks(n), BLOCK, 0, u->stream>>>(n, u->id[c], u->pos4[c], u->v[c][0], u->v[c][1],
                                                           u->v[c][2], u->a[0], u->a[1], u->a[2], u->stage);
    HIPCHK(hipGetLastError());
    double* outs[9] = {x, y, z, vx, vy, vz, ax, ay, az};
    for (int k = 0; k < 9; ++k)
        if (outs[k])
            HIPCHK(hipMemcpyAsync(outs[k], u->stage + (size_t)k * n, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost,
                                  u->stream));
    HIPCHK(hipStreamSynchronize(u->stream));
    return GS_OK;
}

int gs_forces(gs_universe* u, double* potential, double* bond_energy) {
    int rc = check_state(u);
    if (rc) return rc;
    if (u->n == 0) {
        if (potential) *potential = 0.0;
        if (bond_energy) *bond_energy = 0.0;
        return GS_OK;
    }
    if ((rc = rebuild(u, false, u->kp))) return rc;
    u->list_valid = false;
    if ((rc = launch_pair<true>(u))) return rc;
    if ((rc = bond_force_pass(u, true))) return rc;
    k_reduce_to<<<1, BLOCK, 0, u->stream>>>(u->vpart, u->nvpart, u->scal + 1);
    if (u->vb_live) k_reduce_to<<<1, BLOCK, 0, u->stream>>>(u->vb_live, u->nvb_live, u->scal + 3);
    HIPCHK(hipGetLastError());
    double v = 0.0, vb = 0.0;
    HIPCHK(hipMemcpyAsync(&v, u->scal + 1, sizeof(double), hipMemcpyDeviceToHost, u->stream));
    if (u->vb_live) HIPCHK(hipMemcpyAsync(&vb, u->scal + 3, sizeof(double), hipMemcpyDeviceToHost, u->stream));
    HIPCHK(hipStreamSynchronize(u->stream));
    if (potential) *potential = v;
    if (bond_energy) *bond_energy = vb;
    u->extras_valid = true;
    return GS_OK;
}

int gs_download_force_extras(gs_universe* u, double* pe_atom, double* s_abs) {
    int rc = check_state(u);
    if (rc) return rc;
    if (!u->extras_valid) return fail(GS_E_STATE, "no gs_forces result to download");
    const int n = u->n;
    if (n == 0) return GS_OK;
    k_unpermute2<<<nblocks(n), BLOCK, 0, u->stream>>>(n, u->id[u->cur], u->pe, u->sabs, u->stage);
    HIPCHK(hipGetLastError());
    if (pe_atom)
        HIPCHK(hipMemcpyAsync(pe_atom, u->stage, sizeof(double) * n, hipMemcpyDeviceToHost, u->stream));
    if (s_abs)
        HIPCHK(hipMemcpyAsync(s_abs, u->stage + n, sizeof(double) * n, hipMemcpyDeviceToHost, u->stream));
    HIPCHK(hipStreamSynchronize(u->stream));
    return GS_OK;
}

int gs_download_cells(gs_universe* u, uint32_t* cell, uint32_t* dims) {
    int rc = check_state(u);
    if (rc) return rc;
    if (dims) {
        dims[0] = u->kp.dims[0];
        dims[1] = u->kp.dims[1];
        dims[2] = u->kp.dims[2];
    }
    const int n = u->n;
    if (n == 0 || !cell) return GS_OK;
    k_cells_entity<<<nblocks(n), BLOCK, 0, u->stream>>>(u->kp, u->id[u->cur], u->pos4[u->cur], u->stage_u);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(cell, u->stage_u, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, u->stream));
    HIPCHK(hipStreamSynchronize(u->stream));
    return GS_OK;
}

int gs_half_kick(gs_universe* u) {
    int rc = check_state(u);
    if (rc) return rc;
    if (u->n == 0) return GS_OK;
    const int c = u->cur;
    k_half_kick<<<nblocks(u->n), BLOCK, 0, u->stream>>>(u->kp, u->v[c][0], u->v[c][1], u->v[c][2], u->a[0], u->a[1],
                                                         u->a[2]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(u->stream));
    return GS_OK;
}

int gs_drift(gs_universe* u) {
    int rc = check_state(u);
    if (rc) return rc;
    if (u->n == 0) return GS_OK;
    const int c = u->cur;
    k_drift<<<nblocks(u->n), BLOCK, 0, u->stream>>>(u->kp, u->pos4[c], u->v[c][0], u->v[c][1], u->v[c][2]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(u->stream));
    return GS_OK;
}

int gs_wrap(gs_universe* u) {
    int rc = check_state(u);
    if (rc) return rc;
    if (u->n == 0) return GS_OK;
    k_wrap<<<nblocks(u->n), BLOCK, 0, u->stream>>>(u->kp, u->pos4[u->cur]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(u->stream));
    return GS_OK;
}

static int reduce_ke(gs_universe* u, bool thermostat, gs_observables* obs_dev) {
    const int c = u->cur;
    const gs_params& p = u->prm;
    k_kick_ke<false><<<nblocks(u->n), BLOCK, 0, u->stream>>>(
        u->kp, u->v[c][0], u->v[c][1], u->v[c][2], u->a[0], u->a[1], u->a[2], u->type[c], u->kpart, u->counter,
        nullptr, 0, nullptr, 0, u->scal, obs_dev, thermostat ? 1 : 0, p.thermostat_temperature,
        std::fmax(p.thermostat_tau, 1.0), p.thermal_constant);
    HIPCHK(hipGetLastError());
    return GS_OK;
}

int gs_thermostat(gs_universe* u, double* lambda) {
    int rc = check_state(u);
    if (rc) return rc;
    if (u->n == 0) {
        if (lambda) *lambda = 1.0;
        return GS_OK;
    }
    if ((rc = reduce_ke(u, true, nullptr))) return rc;
    const int c = u->cur;
    k_apply_lambda<<<nblocks(u->n), BLOCK, 0, u->stream>>>(u->n, u->v[c][0], u->v[c][1], u->v[c][2], u->scal);
    double l = 1.0;
    HIPCHK(hipMemcpyAsync(&l, u->scal, sizeof(double), hipMemcpyDeviceToHost, u->stream));
    k_set_scalar<<<1, 1, 0, u->stream>>>(u->scal, 1.0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(u->stream));
    if (lambda) *lambda = l;
    return GS_OK;
}

int gs_kinetic(gs_universe* u, double* kinetic) {
    int rc = check_state(u);
    if (rc) return rc;
    if (u->n == 0) {
        if (kinetic) *kinetic = 0.0;
        return GS_OK;
    }
    if (!u->obs || u->obs_cap < 1) {
        (void)hipFree(u->obs);
        u->obs = nullptr;
        u->obs_cap = 0;
        if ((rc = dalloc(&u->obs, 1))) return rc;
        u->obs_cap = 1;
    }
    if ((rc = reduce_ke(u, false, u->obs))) return rc;
    k_set_scalar<<<1, 1, 0, u->stream>>>(u->scal, 1.0);
    gs_observables o;
    HIPCHK(hipMemcpyAsync(&o, u->obs, sizeof(o), hipMemcpyDeviceToHost, u->stream));
    HIPCHK(hipStreamSynchronize(u->stream));
    if (kinetic) *kinetic = o.kinetic;
    return GS_OK;
}

int gs_speed_stats(gs_universe* u, double* kinetic, double* speed_sum) {
    int rc = check_state(u);
    if (rc) return rc;
    double out[2] = {0.0, 0.0};
    if (u->n > 0) {
        const int c = u->cur;
        const uint32_t nb = nblocks(u->n);
        double* kp = u->stage;  // 2 nb partials and the two sums fit the 9 n staging words
        double* sp = u->stage + nb;
        double* res = u->stage + 2 * (size_t)nb;
        k_speed_parts<<<nb, BLOCK, 0, u->stream>>>(u->n, u->v[c][0], u->v[c][1], u->v[c][2], u->type[c], u->kp, kp,
                                                   sp);
        k_reduce_to<<<1, BLOCK, 0, u->stream>>>(kp, nb, res);
        k_reduce_to<<<1, BLOCK, 0, u->stream>>>(sp, nb, res + 1);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, res, sizeof(out), hipMemcpyDeviceToHost, u->stream));
        HIPCHK(hipStreamSynchronize(u->stream));
    }
    if (kinetic) *kinetic = out[0];
    if (speed_sum) *speed_sum = out[1];
    return GS_OK;
}

}  // extern "C"

namespace {

// Ticks per segment of a gs_step call.  The call takes its replay snapshot
// and reads its flags once per segment, so a capacity overflow replays at
// most one segment, and a long call holds a bounded observables buffer.
constexpr int64_t STEP_SEGMENT = 1024;

// One segment of gs_step (n_steps <= STEP_SEGMENT): the ticks, then the
// flags; a replay returns the state, the bonds and the tick to the
// segment's start.
int step_segment(gs_universe* u, int64_t n_steps, gs_observables* obs) {
    int rc;
    if (obs && u->obs_cap < n_steps) {
        (void)hipFree(u->obs);
        u->obs = nullptr;
        u->obs_cap = 0;
        if ((rc = dalloc(&u->obs, (size_t)n_steps))) return rc;
        u->obs_cap = n_steps;
    }
    bool list = (u->pair_mode == 1 || u->pair_mode == 2) && u->list_ok && !u->small_box;
    // four events per timed tick: ticks 0, every, 2 every, ... of the call
    const size_t nev = u->timing ? (size_t)(4 * ((n_steps + u->timing_every - 1) / u->timing_every)) : 0;
    while (u->ev.size() < nev) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        u->ev.push_back(e);
    }
    const int cur0 = u->cur;
    const int64_t tick0 = u->tick;
    const bool bonds0 = u->bonds != nullptr;
    if (list) {
        if ((rc = list_alloc(u, u->nbr_cap ? u->nbr_cap : 384))) return rc;
        if ((rc = snapshot(u, false))) return rc;
        if ((rc = bond_call_begin(u))) return rc;
    }
    u->bound_unfused = false;
    for (int attempt = 0;; ++attempt) {
        if (list) HIPCHK(hipMemsetAsync(u->flags, 0, LIST_FLAGS_RESET * sizeof(uint32_t), u->stream));
        k_set_scalar<<<1, 1, 0, u->stream>>>(u->scal + SCAL_FAULT, 0.0);  // all-zero bits: no fault
        if ((rc = bond_fused_prepare(u, list))) return rc;
        if ((rc = run_ticks(u, n_steps, obs != nullptr, list))) return rc;
        const int c = u->cur;
        k_apply_lambda<<<nblocks(u->n), BLOCK, 0, u->stream>>>(u->n, u->v[c][0], u->v[c][1], u->v[c][2], u->scal);
        k_set_scalar<<<1, 1, 0, u->stream>>>(u->scal, 1.0);
        HIPCHK(hipGetLastError());
        uint32_t fl[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (list) HIPCHK(hipMemcpyAsync(fl, u->flags, sizeof(fl), hipMemcpyDeviceToHost, u->stream));
        // the bond state's sticky overflow word: an entity with more than 32
        // bound partners, tracked episodes or bonds
        unsigned long long bond_ovf = 0, fault = 0;
        if ((rc = bond_overflow_fetch(u, &bond_ovf))) return rc;
        HIPCHK(hipMemcpyAsync(&fault, u->scal + SCAL_FAULT, sizeof(fault), hipMemcpyDeviceToHost, u->stream));
        HIPCHK(hipStreamSynchronize(u->stream));
        if (list && (fl[1] || fl[2] || fl[4] || fl[6] || fl[7])) {
            // A list or a cell bucket overflowed its capacity, a stencil
            // held more candidates than the build stages in LDS, a tick
            // needed a rebuild its predictive binning had not foreseen, or
            // a lane of the pair kernel screened more bound-pair candidates
            // than its LDS column holds: replay the call from the snapshot with larger
            // lists or buckets, on the per-tick cell tiles, with a binning
            // launch every tick, or with the aux-grid bound-pair search.
            if (attempt > 8) return fail(GS_E_OVERFLOW, "neighbour list could not be made valid");
            u->list_replays += 1;
            u->cur = cur0;
            u->tick = tick0;
            if ((rc = snapshot(u, true))) return rc;
            if ((rc = bond_call_restore(u, bonds0))) return rc;
            u->list_valid = false;
            if (fl[4]) u->prebin = false;
            if (fl[7]) u->bound_unfused = true;  // the aux-grid bound-pair search for this call
            if (fl[2]) {
                u->pair_mode = 0;
                list = false;
            } else {
                const uint32_t cap = fl[1] ? u->nbr_cap * 2 : u->nbr_cap;
                if (cap > LIST_CAP_MAX) return fail(GS_E_OVERFLOW, "neighbour list exceeds 65535 entries per atom");
                if ((rc = list_alloc(u, cap, fl[6] ? 2 * u->bcap : 0))) return rc;
            }
            continue;
        }
        if (list) u->list_builds += fl[3];
        u->prebin = true;  // a replay without predictive binning holds for one call
        if (bond_ovf) {
            // Partners past the capacity were dropped: the observed bonds and
            // the forces they feed are incomplete.  On the list path the call
            // returns to its start (state, bonds and tick); on the cell tiles
            // the state is left at the call's end.
            if (list) {
                u->cur = cur0;
                u->tick = tick0;
                if ((rc = snapshot(u, true))) return rc;
                if ((rc = bond_call_restore(u, bonds0))) return rc;
                u->list_valid = false;
                HIPCHK(hipStreamSynchronize(u->stream));
            }
            return fail(GS_E_OVERFLOW, "more than 32 bound partners, tracked episodes or bonds for one entity");
        }
        if (fault) {
            // the ticks ran to the call's end; the state keeps what they made
            std::string what;
            if (fault & 1ull) what += " kinetic energy";
            if (fault & 2ull) what += std::string(what.empty() ? "" : ",") + " potential energy";
            if (fault & 4ull) what += std::string(what.empty() ? "" : ",") + " bond energy";
            return fail(GS_E_NONFINITE, "non-finite" + what + " in ticks " + std::to_string(tick0 + 1) + ".." +
                                            std::to_string(u->tick) + " (a NaN or infinite velocity, force or position)");
        }
        break;
    }
    if (obs)
        HIPCHK(hipMemcpy(obs, u->obs, sizeof(gs_observables) * (size_t)n_steps, hipMemcpyDeviceToHost));
    u->extras_valid = false;
    if (u->timing) {
        int64_t timed = 0;
        for (int64_t s = 0; s < n_steps; s += u->timing_every) {
            const int64_t te = 4 * (s / u->timing_every);
            float a = 0.f, b = 0.f, c2 = 0.f;
            ++timed;
            HIPCHK(hipEventElapsedTime(&b, u->ev[te + 1], u->ev[te + 2]));
            if (u->timing > 1) {
                HIPCHK(hipEventElapsedTime(&a, u->ev[te + 0], u->ev[te + 1]));
                HIPCHK(hipEventElapsedTime(&c2, u->ev[te + 2], u->ev[te + 3]));
            }
            u->tm.rebuild_ms += a;
            u->tm.pair_ms += b;
            u->tm.kick_ms += c2;
            u->tm.step_ms += (double)a + b + c2;
        }
        u->tm.steps += timed;
    }
    return GS_OK;
}

}  // namespace

extern "C" {

int gs_step(gs_universe* u, int64_t n_steps, gs_observables* obs) {
    int rc = check_state(u);
    if (rc) return rc;
    if (n_steps < 0) return fail(GS_E_INVALID, "n_steps < 0");
    if (n_steps == 0 || u->n == 0) return GS_OK;
    for (int64_t done = 0; done < n_steps; done += STEP_SEGMENT) {
        const int64_t k = std::min(STEP_SEGMENT, n_steps - done);
        if ((rc = step_segment(u, k, obs ? obs + done : nullptr))) return rc;
    }
    return GS_OK;
}

int gs_get_tuning(gs_universe* u, gs_tuning* out) {
    if (!u || !out) return fail(GS_E_INVALID, "NULL argument");
    const bool list = (u->pair_mode == 1 || u->pair_mode == 2) && u->list_ok && !u->small_box;
    out->pair_mode = list ? u->pair_mode : 0;
    // the cell-staged pass picks the lanes per atom cell by cell
    out->lanes_per_atom = out->pair_mode == 2 ? 0 : 1 << u->logT;
    out->skin = u->skin;
    out->sort_every = u->sort_every;
    out->list_builds = u->list_builds;
    out->list_replays = u->list_replays;
    out->ms_per_tick = 0.0;
    return GS_OK;
}

int gs_set_tuning(gs_universe* u, const gs_tuning* t) {
    if (!u || !t) return fail(GS_E_INVALID, "NULL argument");
    if (t->pair_mode < 0 || t->pair_mode > 2) return fail(GS_E_INVALID, "pair_mode must be 0, 1 or 2");
    if (t->pair_mode != 0) {
        int logT = u->logT;
        if (t->pair_mode == 1) switch (t->lanes_per_atom) {
            case 1: logT = 0; break;
            case 2: logT = 1; break;
            case 4: logT = 2; break;
            case 8: logT = 3; break;
            case 16: logT = 4; break;
            default: return fail(GS_E_INVALID, "lanes_per_atom must be 1, 2, 4, 8 or 16");
        }
        if (!(t->skin >= 0.0) || !std::isfinite(t->skin)) return fail(GS_E_INVALID, "skin must be >= 0");
        if (t->pair_mode == 1 && (uint64_t)u->n > LIST_ATOMS_MAX)
            return fail(GS_E_INVALID, "Verlet-list mode holds at most 2^27 atoms");
        if (u->small_box || !list_geometry(u, t->skin))
            return fail(GS_E_INVALID, "list mode needs >= 3 cells of edge rc + skin on every axis");
        u->list_ok = true;
        u->logT = logT;
        if (t->sort_every > 0) u->sort_every = t->sort_every;
    }
    u->pair_mode = t->pair_mode;
    u->list_valid = false;
    return GS_OK;
}

int gs_set_bound_search(gs_universe* u, int32_t aux_grid_only) {
    if (!u) return fail(GS_E_INVALID, "universe handle is NULL");
    if (aux_grid_only != 0 && aux_grid_only != 1) return fail(GS_E_INVALID, "aux_grid_only must be 0 or 1");
    u->bound_aux_only = aux_grid_only != 0;
    return GS_OK;
}

int gs_autotune(gs_universe* u, int32_t trial_steps, gs_tuning* best) {
    int rc = check_state(u);
    if (rc) return rc;
    if (trial_steps < 4) return fail(GS_E_INVALID, "trial_steps must be >= 4");
    const int n = u->n;
    // Host copy of the state; every trial starts from it.
    std::vector<double> st(9 * (size_t)std::max(n, 1));
    double* p[9];
    for (int k = 0; k < 9; ++k) p[k] = st.data() + (size_t)k * n;
    if ((rc = gs_download_state(u, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8]))) return rc;
    const std::vector<uint8_t> sp = u->species_h;
    const std::vector<double> q = u->charge_h;
    std::vector<gs_tuning> cand;
    cand.push_back(gs_tuning{0, 1, 0.0, 0, 0, 0, 0.0});
    if (!u->small_box)
        for (double f : {0.04, 0.08, 0.12, 0.18})
            for (int T : {0, 2, 4, 8, 16})
                cand.push_back(gs_tuning{T ? 1 : 2, T, f * u->kp.rc, 0, 0, 0, 0.0});
    gs_tuning orig;
    if ((rc = gs_get_tuning(u, &orig))) return rc;
    const int was_timing = u->timing;
    u->timing = 0;
    gs_tuning win{0, 1, 0.0, 0, 0, 0, 1e300};
    bool any = false;
    int last_err = GS_OK;
    std::string last_msg;
    for (gs_tuning c : cand) {
        if (gs_set_tuning(u, &c) != GS_OK) continue;
        // a trial that fails (a capacity the candidate cannot meet) is skipped;
        // the state is restored either way
        int trc = gs_step(u, trial_steps / 2, nullptr);  // warm up: first lists, list capacity
        float ms = 0.f;
        if (trc == GS_OK) {
            hipEvent_t e0, e1;
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            HIPCHK(hipEventRecord(e0, u->stream));
            trc = gs_step(u, trial_steps, nullptr);
            HIPCHK(hipEventRecord(e1, u->stream));
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        if (trc == GS_E_HIP) {
            u->timing = was_timing;
            return trc;
        }
        if (trc == GS_OK) {
            any = true;
            c.ms_per_tick = ms / trial_steps;
            c.sort_every = u->sort_every;
            // a list setting that fell back to the cell tiles is not a candidate
            if (u->pair_mode == c.pair_mode && c.ms_per_tick < win.ms_per_tick) win = c;
        } else {
            last_err = trc;
            last_msg = g_err;
            if (u->bonds && (rc = gs_bond_reset(u))) return rc;
        }
        if ((rc = gs_upload_state(u, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], sp.data(), q.data())))
            return rc;
    }
    u->timing = was_timing;
    if (!any) {
        // every trial failed: the universe keeps its state and tuning
        orig.list_builds = orig.list_replays = 0;
        if ((rc = gs_set_tuning(u, &orig))) return rc;
        if (u->bonds && (rc = gs_bond_reset(u))) return rc;
        u->tick = 0;
        return fail(last_err, ("autotune: every candidate failed; last: " + last_msg).c_str());
    }
    if ((rc = gs_set_tuning(u, &win))) return rc;
    // the trials observed bonds on copies of the starting state: start the
    // observation afresh
    if (u->bonds && (rc = gs_bond_reset(u))) return rc;
    u->tick = 0;
    u->list_builds = 0;
    if (best) {
        *best = win;
        best->list_builds = 0;
    }
    return GS_OK;
}

int gs_timing_enable(gs_universe* u, int on, int every) {
    if (!u) return fail(GS_E_INVALID, "universe handle is NULL");
    if (on < 0 || on > 2) return fail(GS_E_INVALID, "timing mode must be 0, 1 or 2");
    if (every < 1) return fail(GS_E_INVALID, "timing stride must be >= 1");
    u->timing = on;
    u->timing_every = every;
    u->tm = gs_timing{0, 0.0, 0.0, 0.0, 0.0, 0};
    return GS_OK;
}

int gs_timing_read(gs_universe* u, gs_timing* out) {
    if (!u || !out) return fail(GS_E_INVALID, "NULL argument");
    *out = u->tm;
    return GS_OK;
}

}  // extern "C"

namespace {

// SpatialGrid::new on the host (src/physics/grid.rs:67-86).  Axes with many
// cells are coarsened so the cell count stays O(n): cells stay >= min_cell,
// so the enumerated candidate set still contains every pair within the
// cutoff, and an axis with 1 or 2 cells (where the reference's enumeration
// differs) is never touched.
AuxGrid aux_grid(const KParams& P, double min_cell, int64_t n) {
    AuxGrid G;
    for (int a = 0; a < 3; ++a) {
        double nc = std::floor(P.box[a] / min_cell);
        if (!(nc >= 1.0)) nc = 1.0;
        if (nc > 1024.0) nc = 1024.0;
        G.dims[a] = (uint32_t)nc;
    }
    const uint64_t cap = (uint64_t)std::max<int64_t>(n, 1) * 4 + 4096;
    while ((uint64_t)G.dims[0] * G.dims[1] * G.dims[2] > cap) {
        int a = 0;
        for (int b = 1; b < 3; ++b)
            if (G.dims[b] > G.dims[a]) a = b;
        if (G.dims[a] <= 3) break;
        G.dims[a] = std::max<uint32_t>(3, G.dims[a] / 2);
    }
    for (int a = 0; a < 3; ++a) G.cell[a] = P.box[a] / (double)G.dims[a];
    G.ncell = G.dims[0] * G.dims[1] * G.dims[2];
    return G;
}

int lens_alloc(gs_universe* u, uint32_t ncell) {
    int rc;
    const size_t n = (size_t)std::max(1, u->n);
    if (!u->l_pts) {
        if ((rc = dalloc(&u->l_pts, n)) || (rc = dalloc(&u->l_type, n)) || (rc = dalloc(&u->l_key, n)) ||
            (rc = dalloc(&u->l_id, n)) ||
            (rc = dalloc(&u->l_ticket, n)) || (rc = dalloc(&u->l_perm, n)) || (rc = dalloc(&u->l_parent, n)) ||
            (rc = dalloc(&u->l_size, n)) || (rc = dalloc(&u->l_u64, 8)) || (rc = dalloc(&u->l_sigma, GS_NSPECIES)))
            return rc;
    }
    if (u->l_ncell_cap < ncell) {
        (void)hipFree(u->l_count);
        (void)hipFree(u->l_start);
        u->l_count = nullptr;
        u->l_start = nullptr;
        u->l_ncell_cap = 0;
        if ((rc = dalloc(&u->l_count, ncell)) || (rc = dalloc(&u->l_start, (size_t)ncell + 1))) return rc;
        u->l_ncell_cap = ncell;
    }
    if ((rc = scan_reserve(&u->scan_tmp, &u->scan_tmp_cap, ncell))) return rc;
    HIPCHK(hipMemsetAsync(u->l_count, 0, sizeof(uint32_t) * ncell, u->stream));
    return GS_OK;
}

// Counting sort of m points into the lens grid: l_start, l_perm.
int lens_sort(gs_universe* u, const AuxGrid& G, const double4* pts, int m) {
    k_aux_key<<<nblocks(m), BLOCK, 0, u->stream>>>(m, G, u->kp, pts, u->l_key, u->l_ticket, u->l_count);
    launch_scan(u->stream, u->l_count, u->l_start, G.ncell, u->scan_tmp_cap >= scan_tmp_words(G.ncell) ? u->scan_tmp : nullptr);
    k_scatter<<<nblocks(m), BLOCK, 0, u->stream>>>(m, u->l_key, u->l_ticket, u->l_start, u->l_perm);
    HIPCHK(hipGetLastError());
    return GS_OK;
}

// full: every atom (no subsample).
int radial_impl(gs_universe* u, bool between, int ta, int tb, double r_max_in, int nbins_in, uint64_t* counts,
                double* g, double* r_max_out, double* dr_out, bool full = false) {
    int rc = check_state(u);
    if (rc) return rc;
    if (between && (ta < 0 || ta >= GS_NSPECIES || tb < 0 || tb >= GS_NSPECIES))
        return fail(GS_E_INVALID, "species index out of range");
    // radial_distribution(_between), src/analysis/mod.rs:93-97 and :152-156.
    const int nbins = std::min(std::max(nbins_in, 1), 1 << 16);
    const KParams& P = u->kp;
    const double half_min = std::min(std::min(P.box[0], P.box[1]), P.box[2]) * 0.5;
    const double r_max = std::min(std::fmax(r_max_in, 1e-9), half_min);
    const double dr = r_max / (double)nbins;
    const double vol = P.box[0] * P.box[1] * P.box[2];
    // g_subsample (src/analysis/mod.rs:32-46).
    const uint32_t N = (uint32_t)u->n;
    uint32_t step = 1, m = N;
    if (N > 8192u && !full) {
        step = N / 8192u + 1u;
        m = N / step;
    }
    std::vector<unsigned long long> hc((size_t)nbins, 0ull);
    if (m > 0) {
        const AuxGrid G = aux_grid(P, r_max, m);
        if ((rc = lens_alloc(u, G.ncell))) return rc;
        if (u->l_bins_cap < nbins) {
            (void)hipFree(u->l_bins);
            u->l_bins = nullptr;
            u->l_bins_cap = 0;
            if ((rc = dalloc(&u->l_bins, (size_t)nbins))) return rc;
            u->l_bins_cap = nbins;
        }
        HIPCHK(hipMemsetAsync(u->l_bins, 0, sizeof(unsigned long long) * nbins, u->stream));
        const int c = u->cur;
        k_gr_gather<<<nblocks(u->n), BLOCK, 0, u->stream>>>(u->n, u->id[c], u->pos4[c], u->type[c], step, m,
                                                             u->l_pts, u->l_type);
        if ((rc = lens_sort(u, G, u->l_pts, (int)m))) return rc;
        const int refm = u->prm.grid_semantics == GS_GRID_REFERENCE;
        const size_t shm = nbins <= 4096 ? sizeof(uint32_t) * (size_t)nbins : 0;
        if (between)
            k_gr_pairs<true><<<nblocks(m, BLOCK / WAVE), BLOCK, shm, u->stream>>>((int)m, G, P, u->l_pts, u->l_type, u->l_start,
                                                                    u->l_perm, r_max, dr, nbins, ta, tb, refm,
                                                                    u->l_bins);
        else
            k_gr_pairs<false><<<nblocks(m, BLOCK / WAVE), BLOCK, shm, u->stream>>>((int)m, G, P, u->l_pts, u->l_type, u->l_start,
                                                                     u->l_perm, r_max, dr, nbins, ta, tb, refm,
                                                                     u->l_bins);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(hc.data(), u->l_bins, sizeof(unsigned long long) * nbins, hipMemcpyDeviceToHost,
                              u->stream));
        HIPCHK(hipStreamSynchronize(u->stream));
    }
    // Normalisation, src/analysis/mod.rs:119-132 / :184-200, on the host with
    // the reference's operation order.
    double pair_norm;
    if (!between) {
        pair_norm = m < 2 ? 0.0 : (double)((uint64_t)m * (uint64_t)(m - 1)) * 0.5;
    } else {
        // species counts of the subsample
        std::vector<uint8_t> st((size_t)m);
        if (m > 0) {
            HIPCHK(hipMemcpyAsync(st.data(), u->l_type, m, hipMemcpyDeviceToHost, u->stream));
            HIPCHK(hipStreamSynchronize(u->stream));
        }
        uint64_t na = 0, nb = 0;
        for (uint8_t t : st) {
            na += (t == ta);
            nb += (t == tb);
        }
        pair_norm = ta == tb ? (double)na * (double)(na > 0 ? na - 1 : 0) * 0.5 : (double)na * (double)nb;
    }
    for (int i = 0; i < nbins; ++i) {
        if (counts) counts[i] = (uint64_t)hc[i];
        if (g) {
            if (pair_norm <= 0.0) {
                g[i] = 0.0;
                continue;
            }
            const double r_lo = (double)i * dr;
            const double r_hi = r_lo + dr;
            const double shell = 4.0 / 3.0 * M_PI * (r_hi * r_hi * r_hi - r_lo * r_lo * r_lo);
            const double expected = pair_norm * shell / vol;
            g[i] = (double)hc[i] / std::fmax(expected, 2.2250738585072014e-308);
        }
    }
    if (r_max_out) *r_max_out = r_max;
    if (dr_out) *dr_out = dr;
    return GS_OK;
}

}  // namespace

extern "C" {

int gs_velocity_histogram(gs_universe* u, double max_speed, int32_t nbins_in, uint64_t* counts,
                          uint64_t* samples, uint64_t* overflow, double* bin_width) {
    int rc = check_state(u);
    if (rc) return rc;
    const int nbins = std::min(std::max(nbins_in, 1), VHIST_MAX_BINS);
    const double ms = std::fmax(max_speed, 1e-9);
    const double bw = ms / (double)nbins;
    if ((rc = lens_alloc(u, 1))) return rc;
    unsigned long long* d = u->l_bins;
    if (!d || u->l_bins_cap < nbins + 1) {
        (void)hipFree(u->l_bins);
        u->l_bins = nullptr;
        u->l_bins_cap = 0;
        if ((rc = dalloc(&u->l_bins, (size_t)std::max(nbins + 1, VHIST_MAX_BINS + 1)))) return rc;
        u->l_bins_cap = std::max(nbins + 1, VHIST_MAX_BINS + 1);
        d = u->l_bins;
    }
    HIPCHK(hipMemsetAsync(d, 0, sizeof(unsigned long long) * (size_t)(nbins + 1), u->stream));
    const int c = u->cur;
    if (u->n > 0)
        k_vhist<<<nblocks(u->n), BLOCK, 0, u->stream>>>(u->n, u->v[c][0], u->v[c][1], u->v[c][2], ms, bw, nbins, d);
    HIPCHK(hipGetLastError());
    std::vector<unsigned long long> h((size_t)nbins + 1);
    HIPCHK(hipMemcpyAsync(h.data(), d, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, u->stream));
    HIPCHK(hipStreamSynchronize(u->stream));
    if (counts)
        for (int k = 0; k < nbins; ++k) counts[k] = h[k];
    if (samples) *samples = (uint64_t)u->n;
    if (overflow) *overflow = h[nbins];
    if (bin_width) *bin_width = bw;
    return GS_OK;
}

int gs_radial_distribution(gs_universe* u, double r_max, int32_t nbins, uint64_t* counts, double* g,
                           double* r_max_out, double* dr_out) {
    return radial_impl(u, false, -1, -1, r_max, nbins, counts, g, r_max_out, dr_out);
}

int gs_radial_distribution_between(gs_universe* u, int32_t species_a, int32_t species_b, double r_max,
                                   int32_t nbins, uint64_t* counts, double* g, double* r_max_out, double* dr_out) {
    return radial_impl(u, true, species_a, species_b, r_max, nbins, counts, g, r_max_out, dr_out);
}

int gs_radial_distribution_full(gs_universe* u, int32_t species_a, int32_t species_b, double r_max, int32_t nbins,
                                uint64_t* counts, double* g, double* r_max_out, double* dr_out) {
    if (species_a < 0) return radial_impl(u, false, -1, -1, r_max, nbins, counts, g, r_max_out, dr_out, true);
    return radial_impl(u, true, species_a, species_b, r_max, nbins, counts, g, r_max_out, dr_out, true);
}

int gs_clusters(gs_universe* u, double bond_factor, uint32_t* labels, gs_cluster_stats* stats) {
    int rc = check_state(u);
    if (rc) return rc;
    gs_cluster_stats st = {0, 0, 0, 0.0, 0};
    const int n = u->n;
    if (n == 0) {
        if (stats) *stats = st;
        return GS_OK;
    }
    // clusters, src/analysis/mod.rs:235-243: sigma of the present species.
    double max_sigma = 0.0;
    for (int s = 0; s < GS_NSPECIES; ++s)
        if (u->present[s]) max_sigma = std::fmax(max_sigma, u->prm.sigma[s]);
    const double cutoff = std::fmax(std::fmax(bond_factor, 0.1) * max_sigma, 1e-9);
    const KParams& P = u->kp;
    const AuxGrid G = aux_grid(P, cutoff, n);
    if ((rc = lens_alloc(u, G.ncell))) return rc;
    const int c = u->cur;
    HIPCHK(hipMemcpyAsync(u->l_sigma, u->prm.sigma, sizeof(double) * GS_NSPECIES, hipMemcpyHostToDevice, u->stream));
    HIPCHK(hipMemsetAsync(u->l_u64, 0, sizeof(unsigned long long) * 8, u->stream));
    HIPCHK(hipMemsetAsync(u->l_size, 0, sizeof(uint32_t) * (size_t)n, u->stream));
    k_iota<<<nblocks(n), BLOCK, 0, u->stream>>>(n, u->l_parent);
    if ((rc = lens_sort(u, G, u->pos4[c], n))) return rc;
    const int refm = u->prm.grid_semantics == GS_GRID_REFERENCE;
    k_fof_edges<<<nblocks(n), BLOCK, 0, u->stream>>>(n, G, P, u->pos4[c], u->type[c], u->id[c], u->l_start,
                                                     u->l_perm, u->l_sigma, bond_factor, cutoff, refm, u->l_parent,
                                                     u->l_u64 + 4);
    k_fof_finalize<<<nblocks(n), BLOCK, 0, u->stream>>>(n, u->l_parent, u->l_size);
    k_fof_stats<<<nblocks(n), BLOCK, 0, u->stream>>>(n, u->l_size, u->l_u64);
    if (labels) k_fof_labels<<<nblocks(n), BLOCK, 0, u->stream>>>(n, u->l_parent, u->l_key);
    HIPCHK(hipGetLastError());
    unsigned long long h[8];
    HIPCHK(hipMemcpyAsync(h, u->l_u64, sizeof(h), hipMemcpyDeviceToHost, u->stream));
    if (labels) HIPCHK(hipMemcpyAsync(labels, u->l_key, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToHost, u->stream));
    HIPCHK(hipStreamSynchronize(u->stream));
    st.monomers = (int64_t)h[0];
    st.aggregates = (int64_t)h[1];
    st.largest = (int64_t)h[2];
    st.mean_size = (double)n / (double)std::max<unsigned long long>(h[3], 1ull);
    st.bound_pairs = (int64_t)h[4];
    if (stats) *stats = st;
    return GS_OK;
}

}  // extern "C"

#include "gs_adjoint.inc"
#include "gs_batch.inc"
#include "gs_bonds.inc"
#include "gs_noalloc.inc"
#include "gs_snapshot.inc"
