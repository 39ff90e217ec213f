/*
 * oracle.cpp — CPU FP64 restatement of the reference force-and-integration
 * path of ValentinVaninetti/genesis (a Rust crate that cannot be built in this
 * image: no cargo/rustc).  TEST INFRASTRUCTURE ONLY: it is the checker for
 * the HIP path and the timed CPU baseline in bench.py.  Nothing in the
 * product (genesis_amd/) links, loads or calls it.
 *
 * Every function cites the reference file:line whose arithmetic it restates.
 * Evaluation order follows Rust's left-to-right semantics; the file is built
 * with -ffp-contract=off so no FMA is formed where Rust would round twice.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *  - pair laws, mixing, switch, curvature: the reference's own closed-form
 *    unit tests (src/physics/forces.rs:394-594) re-run against this file;
 *  - grid enumeration: exact pair sets against brute force
 *    (src/physics/grid.rs:313-388), plus the 2-cell multiplicity (Q1) whose
 *    counts {2,4,8} follow from canonical_offset (src/physics/grid.rs:275-295);
 *  - force pass: F = -dV/dx by central differences of the total potential,
 *    Newton's third law (sum of forces = 0), a brute-force O(N^2) pair sum;
 *  - integrator: NVE energy drift and momentum (src/systems/forces.rs:368-401);
 *  - thermostat: convergence toward the target (src/systems/thermostat.rs:94-116)
 *    and the closed-form lambda;
 *  - g(r) / clusters / bond components: the reference's hand fixtures
 *    (src/analysis/mod.rs:454-655) and brute-force histograms.
 */
#include "../include/genesis_gpu.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace {

constexpr double F64_EPS = 2.220446049250313e-16;  // f64::EPSILON
constexpr double F64_MIN_POSITIVE = 2.2250738585072014e-308;
constexpr double LJ_NUCLEUS = 0.5;                   // src/physics/forces.rs:173
constexpr uint32_t EMPTY = 0xFFFFFFFFu;              // src/physics/grid.rs:34
constexpr size_t G_SAMPLE_CAP = 8192;                // src/analysis/mod.rs:28

struct V3 {
    double x, y, z;
};
inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator*(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
// Vec3::length_squared, src/math/mod.rs:29-31 (left to right).
inline double len2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }

// f64::rem_euclid (Rust std): r = a % b; if r < 0 { r + |b| } else { r }.
inline double rem_euclid(double a, double b) {
    double r = std::fmod(a, b);
    return r < 0.0 ? r + std::fabs(b) : r;
}
// `x as u32` for f64 in Rust saturates (NaN -> 0).
inline uint32_t f64_as_u32(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
inline uint64_t f64_as_u64(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 18446744073709551615.0) return ~0ull;
    return (uint64_t)x;
}
// Rust f64::max: the non-NaN operand wins.
inline double fmax_r(double a, double b) { return std::fmax(a, b); }

// min_image, src/physics/grid.rs:300-306 (round = half away from zero).
inline V3 min_image(V3 d, V3 size) {
    return v3(d.x - size.x * std::round(d.x / size.x),
              d.y - size.y * std::round(d.y / size.y),
              d.z - size.z * std::round(d.z / size.z));
}

// ---- pair laws: src/physics/forces.rs ---------------------------------------

struct LjPair {
    double sigma, epsilon;
};

// lj_raw, src/physics/forces.rs:346-353.
inline void lj_raw(LjPair p, double r, double* m, double* v) {
    double s = p.sigma / r;
    double s6 = s * s * s * s * s * s;
    double s12 = s6 * s6;
    *v = 4.0 * p.epsilon * (s12 - s6);
    *m = (24.0 * p.epsilon / p.sigma) * s * (2.0 * s12 - s6);
}

// smooth_cutoff, src/physics/forces.rs:302-317.
inline void smooth_cutoff(double r, double r_on, double rc, double* sw, double* dsw) {
    if (r <= r_on) { *sw = 1.0; *dsw = 0.0; return; }
    if (r >= rc) { *sw = 0.0; *dsw = 0.0; return; }
    double u = (r - r_on) / (rc - r_on);
    double u2 = u * u;
    double u3 = u2 * u;
    double u4 = u3 * u;
    double u5 = u4 * u;
    *sw = 1.0 - 10.0 * u3 + 15.0 * u4 - 6.0 * u5;
    double dsdu = -30.0 * u2 * (1.0 - u) * (1.0 - u);
    *dsw = dsdu / (rc - r_on);
}

// switched, src/physics/forces.rs:322-324.
inline void switched(double f_raw, double v_raw, double sw, double dsw, double* f, double* v) {
    *f = f_raw * sw - v_raw * dsw;
    *v = v_raw * sw;
}

// coulomb_raw, src/physics/forces.rs:330-333.
inline void coulomb_raw(double k_e, double qa, double qb, double r, double* f, double* v) {
    double q = k_e * qa * qb;
    *f = q / (r * r);
    *v = q / r;
}

// gravity_raw, src/physics/forces.rs:338-341.
inline void gravity_raw(double g, double ma, double mb, double r, double* f, double* v) {
    double m = g * ma * mb;
    *f = -m / (r * r);
    *v = -m / r;
}

// bond_harmonic_raw, src/physics/forces.rs:290-293.
inline void bond_harmonic_raw(double k, double r_eq, double r, double* f, double* v) {
    double dr = r - r_eq;
    *f = -k * dr;
    *v = 0.5 * k * dr * dr;
}

// well_curvature, src/physics/forces.rs:194-196.
inline double well_curvature(double epsilon, double sigma) {
    return 72.0 * std::pow(2.0, -1.0 / 3.0) * epsilon / (sigma * sigma);
}
// equilibrium_distance, src/physics/forces.rs:281-283.
inline double equilibrium_distance(double sigma) { return sigma * std::pow(2.0, 1.0 / 6.0); }
// vib_period, src/physics/forces.rs:201-203.
inline double vib_period(double epsilon, double sigma, double reduced_mass) {
    return 2.0 * M_PI * std::sqrt(reduced_mass / well_curvature(epsilon, sigma));
}

// LjTable::new, src/physics/forces.rs:223-237.
struct LjTable {
    LjPair pair[GS_NSPECIES][GS_NSPECIES];
    double rc, r_on;
    LjTable(const gs_params& p) {
        double max_sigma = 0.0;
        for (int i = 0; i < GS_NSPECIES; ++i)
            for (int j = 0; j < GS_NSPECIES; ++j) {
                double sigma = 0.5 * (p.sigma[i] + p.sigma[j]);
                double epsilon = p.thermal_constant * std::sqrt(p.epsilon_k[i] * p.epsilon_k[j]);
                max_sigma = fmax_r(max_sigma, sigma);
                pair[i][j] = LjPair{sigma, epsilon};
            }
        rc = fmax_r(p.lj_cutoff_factor, 1.0) * max_sigma;
        r_on = 0.9 * rc;
    }
    // force_switched_scalar, src/physics/forces.rs:270-275.
    void force_switched_scalar(LjPair p, double r, double* m, double* v) const {
        r = fmax_r(r, LJ_NUCLEUS * p.sigma);
        double mr, vr, sw, dsw;
        lj_raw(p, r, &mr, &vr);
        smooth_cutoff(r, r_on, rc, &sw, &dsw);
        switched(mr, vr, sw, dsw, m, v);
    }
};

// ---- SpatialGrid: src/physics/grid.rs ---------------------------------------

const int ALL_OFFSETS[27][3] = {  // src/physics/grid.rs:37-47
    {-1, -1, -1}, {-1, -1, 0}, {-1, -1, 1}, {-1, 0, -1}, {-1, 0, 0}, {-1, 0, 1},
    {-1, 1, -1},  {-1, 1, 0},  {-1, 1, 1},  {0, -1, -1}, {0, -1, 0}, {0, -1, 1},
    {0, 0, -1},   {0, 0, 0},   {0, 0, 1},   {0, 1, -1},  {0, 1, 0},  {0, 1, 1},
    {1, -1, -1},  {1, -1, 0},  {1, -1, 1},  {1, 0, -1},  {1, 0, 0},  {1, 0, 1},
    {1, 1, -1},   {1, 1, 0},   {1, 1, 1},
};

struct Pair {
    uint32_t a, b;
    V3 normal;
};

struct SpatialGrid {
    uint32_t dims[3];
    V3 cell, size;
    std::vector<uint32_t> heads, touched, next, chain, nchain;

    // SpatialGrid::new, src/physics/grid.rs:67-86.
    SpatialGrid(V3 world, double min_cell) : size(world) {
        auto ncell = [&](double s) { return f64_as_u32(fmax_r(std::floor(s / min_cell), 1.0)); };
        dims[0] = ncell(world.x);
        dims[1] = ncell(world.y);
        dims[2] = ncell(world.z);
        cell = v3(world.x / (double)dims[0], world.y / (double)dims[1], world.z / (double)dims[2]);
        heads.assign((size_t)dims[0] * dims[1] * dims[2], EMPTY);
    }
    // cell_of, src/physics/grid.rs:228-238.
    void cell_of(V3 pos, uint32_t c[3]) const {
        auto axis = [](double v, double period, double cw, uint32_t n) {
            double u = rem_euclid(v + 0.5 * period, period);
            return std::min(f64_as_u32(std::floor(u / cw)), n - 1);
        };
        c[0] = axis(pos.x, size.x, cell.x, dims[0]);
        c[1] = axis(pos.y, size.y, cell.y, dims[1]);
        c[2] = axis(pos.z, size.z, cell.z, dims[2]);
    }
    // encode, src/physics/grid.rs:244-246.
    size_t encode(uint32_t ix, uint32_t iy, uint32_t iz) const {
        return (size_t)(ix + dims[0] * (iy + dims[1] * iz));
    }
    void decode(uint32_t idx, uint32_t c[3]) const {
        uint32_t nxny = dims[0] * dims[1];
        c[2] = idx / nxny;
        uint32_t rem = idx % nxny;
        c[1] = rem / dims[0];
        c[0] = rem % dims[0];
    }
    size_t neighbor(const uint32_t c[3], const int d[3]) const {
        uint32_t r[3];
        for (int k = 0; k < 3; ++k) {
            int64_t v = ((int64_t)c[k] + d[k]) % (int64_t)dims[k];
            if (v < 0) v += dims[k];
            r[k] = (uint32_t)v;
        }
        return encode(r[0], r[1], r[2]);
    }
    // canonical_offset, src/physics/grid.rs:275-295.
    void canonical(const int o[3], int out[3]) const {
        int d[3] = {o[0], o[1], o[2]};
        for (int k = 0; k < 3; ++k)
            if (dims[k] <= 1) d[k] = 0;
        int first = 0;
        for (int k = 0; k < 3; ++k)
            if (d[k] != 0) { first = d[k]; break; }
        if (first == 1) { out[0] = d[0]; out[1] = d[1]; out[2] = d[2]; }
        else if (first == -1) { out[0] = -d[0]; out[1] = -d[1]; out[2] = -d[2]; }
        else { out[0] = out[1] = out[2] = 0; }
    }
    // build, src/physics/grid.rs:99-113 (head insertion).
    void build(const std::vector<V3>& pos) {
        for (uint32_t ci : touched) heads[ci] = EMPTY;
        touched.clear();
        next.assign(pos.size(), EMPTY);
        for (size_t slot = 0; slot < pos.size(); ++slot) {
            uint32_t c[3];
            cell_of(pos[slot], c);
            size_t ci = encode(c[0], c[1], c[2]);
            if (heads[ci] == EMPTY) touched.push_back((uint32_t)ci);
            next[slot] = heads[ci];
            heads[ci] = (uint32_t)slot;
        }
    }
    static void gather(const std::vector<uint32_t>& nxt, uint32_t head, std::vector<uint32_t>& buf) {
        buf.clear();
        for (uint32_t cur = head; cur != EMPTY; cur = nxt[cur]) buf.push_back(cur);
    }
    // check_pair, src/physics/grid.rs:201-225.
    void check_pair(const std::vector<V3>& pos, uint32_t sa, uint32_t sb, double cd2,
                    std::vector<Pair>& out) const {
        V3 delta = min_image(pos[sa] - pos[sb], size);
        double d2 = len2(delta);
        if (d2 >= cd2) return;
        double d = std::sqrt(d2);
        if (d <= F64_EPS) return;
        out.push_back(Pair{sa, sb, delta * (1.0 / d)});
    }
    // neighbors, src/physics/grid.rs:119-153 (half shell, canonical offsets).
    void neighbors(const std::vector<V3>& pos, double cutoff, std::vector<Pair>& out) {
        double cd2 = cutoff * cutoff;
        out.clear();
        for (uint32_t ci : touched) {
            gather(next, heads[ci], chain);
            uint32_t c[3];
            decode(ci, c);
            for (size_t i = 0; i < chain.size(); ++i)
                for (size_t j = i + 1; j < chain.size(); ++j) check_pair(pos, chain[i], chain[j], cd2, out);
            for (int k = 0; k < 27; ++k) {
                const int* o = ALL_OFFSETS[k];
                if (o[0] == 0 && o[1] == 0 && o[2] == 0) continue;
                size_t nc = neighbor(c, o);
                if (nc == ci || heads[nc] == EMPTY) continue;
                int can[3];
                canonical(o, can);
                if (can[0] != o[0] || can[1] != o[1] || can[2] != o[2]) continue;
                gather(next, heads[nc], nchain);
                for (uint32_t sa : chain)
                    for (uint32_t sb : nchain) check_pair(pos, sa, sb, cd2, out);
            }
        }
    }
};

// brute_pairs, src/physics/grid.rs:313-326 (with check_pair's overlap guard
// so it is the same relation as the grid's).
void brute_pairs(const std::vector<V3>& pos, V3 size, double cutoff, std::vector<Pair>& out) {
    double cd2 = cutoff * cutoff;
    out.clear();
    for (size_t i = 0; i < pos.size(); ++i)
        for (size_t j = i + 1; j < pos.size(); ++j) {
            V3 delta = min_image(pos[i] - pos[j], size);
            double d2 = len2(delta);
            if (d2 >= cd2) continue;
            double d = std::sqrt(d2);
            if (d <= F64_EPS) continue;
            out.push_back(Pair{(uint32_t)i, (uint32_t)j, delta * (1.0 / d)});
        }
}

void enumerate_pairs(const gs_params& p, const std::vector<V3>& pos, double cutoff, std::vector<Pair>& out) {
    V3 size = v3(p.box[0], p.box[1], p.box[2]);
    if (p.grid_semantics == GS_GRID_MINIMAGE) {
        brute_pairs(pos, size, cutoff, out);
    } else {
        SpatialGrid g(size, cutoff);
        g.build(pos);
        g.neighbors(pos, cutoff, out);
    }
}

// ---- ForceSystem: src/systems/forces.rs --------------------------------------

struct ForceCutoffs {
    double rc, lj_rc, coulomb_rc, coulomb_r_on, gravity_rc, gravity_r_on;
};

// ForceSystem::new, src/systems/forces.rs:100-142.
ForceCutoffs force_cutoffs(const gs_params& p, const LjTable& lj) {
    double max_sigma = 0.0;
    for (int i = 0; i < GS_NSPECIES; ++i) max_sigma = fmax_r(max_sigma, p.sigma[i]);
    ForceCutoffs c;
    c.lj_rc = lj.rc;
    c.coulomb_rc = p.enable_electrostatics ? fmax_r(p.coulomb_cutoff, 1.0) * max_sigma : 0.0;
    c.gravity_rc = p.enable_gravity ? fmax_r(p.gravity_cutoff, 1.0) * max_sigma : 0.0;
    c.rc = fmax_r(fmax_r(c.lj_rc, c.coulomb_rc), c.gravity_rc);
    c.coulomb_r_on = 0.9 * c.coulomb_rc;
    c.gravity_r_on = 0.9 * c.gravity_rc;
    return c;
}

struct ForceOut {
    std::vector<V3> acc;
    double potential = 0.0, bond_potential = 0.0;
    std::vector<double> pe, sabs;
    int64_t npairs = 0;
};

struct Part {
    std::vector<std::pair<uint32_t, V3>> forces;
    double v = 0.0, vb = 0.0;
    std::vector<std::pair<uint32_t, double>> pe, sabs;
};

// One pair of the fold body, src/systems/forces.rs:256-324.
void fold_pair(const gs_params& p, const LjTable& lj, const ForceCutoffs& c, const std::vector<V3>& pos,
               const uint8_t* types, const double* q, const std::vector<std::vector<uint32_t>>* bond_of,
               const Pair& pair, Part& acc, bool extras) {
    V3 size = v3(p.box[0], p.box[1], p.box[2]);
    double rc2 = c.rc * c.rc;
    V3 delta = min_image(pos[pair.a] - pos[pair.b], size);
    double d2 = len2(delta);
    if (d2 >= rc2) return;
    double d = std::sqrt(d2);
    if (d <= F64_EPS) return;
    V3 normal = delta * (1.0 / d);
    double fmag = 0.0, vpair = 0.0, vbpair = 0.0;
    if (d < lj.rc) {
        LjPair lp = lj.pair[types[pair.a]][types[pair.b]];
        double m, vv;
        lj.force_switched_scalar(lp, d, &m, &vv);
        fmag += m;
        acc.v += vv;
        vpair += vv;
    }
    if (p.enable_electrostatics && d < c.coulomb_rc) {
        double r = fmax_r(d, 0.5);
        double m, vv, sw, dsw, m2, vv2;
        coulomb_raw(p.coulomb_constant, q[pair.a], q[pair.b], r, &m, &vv);
        smooth_cutoff(d, c.coulomb_r_on, c.coulomb_rc, &sw, &dsw);
        switched(m, vv, sw, dsw, &m2, &vv2);
        fmag += m2;
        acc.v += vv2;
        vpair += vv2;
    }
    if (p.enable_gravity && d < c.gravity_rc) {
        double r = fmax_r(d, 0.5);
        double m, vv, sw, dsw, m2, vv2;
        gravity_raw(p.gravity_constant, p.mass[types[pair.a]], p.mass[types[pair.b]], r, &m, &vv);
        smooth_cutoff(d, c.gravity_r_on, c.gravity_rc, &sw, &dsw);
        switched(m, vv, sw, dsw, &m2, &vv2);
        fmag += m2;
        acc.v += vv2;
        vpair += vv2;
    }
    if (p.enable_bond_interaction && bond_of) {
        const auto& lst = (*bond_of)[pair.a];
        if (std::find(lst.begin(), lst.end(), pair.b) != lst.end()) {
            LjPair lp = lj.pair[types[pair.a]][types[pair.b]];
            double r_eq = equilibrium_distance(lp.sigma);
            double rc_bond = p.bond_k_bind * lp.sigma;
            if (d < rc_bond) {
                double k = well_curvature(lp.epsilon, lp.sigma);
                double m, vv, sw, dsw, m2, vv2;
                bond_harmonic_raw(k, r_eq, d, &m, &vv);
                smooth_cutoff(d, r_eq, rc_bond, &sw, &dsw);
                switched(m, vv, sw, dsw, &m2, &vv2);
                fmag += m2;
                acc.vb += vv2;
                vbpair += vv2;
            }
        }
    }
    if (fmag != 0.0) {
        acc.forces.emplace_back(pair.a, normal * fmag);
        acc.forces.emplace_back(pair.b, normal * (-fmag));
    }
    if (extras) {
        double h = 0.5 * (vpair + vbpair);
        acc.pe.emplace_back(pair.a, h);
        acc.pe.emplace_back(pair.b, h);
        acc.sabs.emplace_back(pair.a, std::fabs(fmag));
        acc.sabs.emplace_back(pair.b, std::fabs(fmag));
    }
}

// ForceSystem::run, src/systems/forces.rs:163-353.  `nthreads` contiguous
// splits of the pair list play the role of rayon's fold splits; parts are
// merged in split order (src/systems/forces.rs:330-341).
void force_pass(const gs_params& p, const std::vector<V3>& pos, const uint8_t* types, const double* q,
                const std::vector<std::vector<uint32_t>>* bond_of, int nthreads, bool extras, ForceOut& out) {
    LjTable lj(p);
    ForceCutoffs c = force_cutoffs(p, lj);
    size_t n = pos.size();
    out.acc.assign(n, v3(0, 0, 0));
    out.potential = 0.0;
    out.bond_potential = 0.0;
    if (extras) {
        out.pe.assign(n, 0.0);
        out.sabs.assign(n, 0.0);
    }
    if (n == 0) return;
    std::vector<Pair> pairs;
    enumerate_pairs(p, pos, c.rc, pairs);
    out.npairs = (int64_t)pairs.size();
    int T = std::max(1, nthreads);
    if ((size_t)T > pairs.size()) T = std::max<size_t>(1, pairs.size());
    std::vector<Part> parts(T);
    auto work = [&](int t) {
        size_t lo = pairs.size() * (size_t)t / T, hi = pairs.size() * (size_t)(t + 1) / T;
        for (size_t k = lo; k < hi; ++k) fold_pair(p, lj, c, pos, types, q, bond_of, pairs[k], parts[t], extras);
    };
    if (T == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t) th.emplace_back(work, t);
        for (auto& x : th) x.join();
    }
    for (auto& part : parts) {
        out.potential += part.v;
        out.bond_potential += part.vb;
        for (auto& f : part.forces) {
            double inv = 1.0 / p.mass[types[f.first]];
            out.acc[f.first] = out.acc[f.first] + f.second * inv;
        }
        if (extras) {
            for (auto& e : part.pe) out.pe[e.first] += e.second;
            for (auto& e : part.sabs) out.sabs[e.first] += e.second;
        }
    }
}

// ---- integrator, wrap, thermostat, stats -------------------------------------

// VelocityHalfKick::run, src/systems/integrate.rs:38-44.
void half_kick(std::vector<V3>& v, const std::vector<V3>& a, double dt) {
    double half_dt = 0.5 * dt;
    for (size_t i = 0; i < v.size(); ++i) v[i] = v[i] + a[i] * half_dt;
}
// PositionDrift::run, src/systems/integrate.rs:61-67.
void drift(std::vector<V3>& x, const std::vector<V3>& v, double dt) {
    for (size_t i = 0; i < x.size(); ++i) x[i] = x[i] + v[i] * dt;
}
// BoundarySystem::run, src/systems/boundary.rs:22-35.
void wrap(std::vector<V3>& x, const double box[3]) {
    double h[3] = {box[0] * 0.5, box[1] * 0.5, box[2] * 0.5};
    auto w = [](double v, double hh) { return rem_euclid(v, 2.0 * hh) - hh; };
    for (auto& p : x) {
        p.x = w(p.x, h[0]);
        p.y = w(p.y, h[1]);
        p.z = w(p.z, h[2]);
    }
}
// Kinetic sum in slot order, src/systems/thermostat.rs:54-57 and
// src/systems/aggregate.rs:86-93.
double kinetic(const std::vector<V3>& v, const uint8_t* types, const double* mass) {
    double k = 0.0;
    for (size_t i = 0; i < v.size(); ++i) k += 0.5 * mass[types[i]] * len2(v[i]);
    return k;
}
// ThermostatSystem::run, src/systems/thermostat.rs:47-71.  Returns lambda
// (1 when the system returns early).
double thermostat(const gs_params& p, std::vector<V3>& v, const uint8_t* types) {
    double tau = fmax_r(p.thermostat_tau, 1.0);
    double target = p.thermostat_temperature;
    if (target <= 0.0 || v.empty()) return 1.0;
    double kin = kinetic(v, types, p.mass);
    double t = (2.0 / 3.0) * (kin / (double)v.size()) / p.thermal_constant;
    if (!std::isfinite(t) || t <= 0.0) return 1.0;
    double lambda2 = 1.0 + (1.0 / tau) * (target / t - 1.0);
    double lambda = std::sqrt(fmax_r(lambda2, 0.0));
    for (auto& x : v) x = x * lambda;
    return lambda;
}

// ---- analysis lenses: src/analysis/mod.rs ------------------------------------

// g_subsample, src/analysis/mod.rs:32-46.
std::vector<size_t> g_subsample_index(size_t n) {
    std::vector<size_t> idx;
    if (n <= G_SAMPLE_CAP) {
        idx.resize(n);
        for (size_t i = 0; i < n; ++i) idx[i] = i;
        return idx;
    }
    size_t step = n / G_SAMPLE_CAP + 1;
    size_t count = n / step;
    idx.resize(count);
    for (size_t i = 0; i < count; ++i) idx[i] = i * step;
    return idx;
}

// Pair counting shared by radial_distribution(_between),
// src/analysis/mod.rs:93-117 and :152-182.
// full: every point, no subsample (the GPU's full mode; the reference always
// subsamples above G_SAMPLE_CAP points).
void gr_counts(const std::vector<V3>& pos, const std::vector<int>& type, V3 size, double r_max_in,
               int nbins_in, int ta, int tb, std::vector<uint64_t>& counts, double* r_max_out, double* dr_out,
               double* pair_norm, bool full = false) {
    int nbins = std::min(std::max(nbins_in, 1), 1 << 16);
    double half_min = std::min(std::min(size.x, size.y), size.z) * 0.5;
    double r_max = std::min(fmax_r(r_max_in, 1e-9), half_min);
    double dr = r_max / (double)nbins;
    std::vector<size_t> idx = g_subsample_index(full ? 0 : pos.size());
    if (full) {
        idx.resize(pos.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
    }
    std::vector<V3> sp(idx.size());
    std::vector<int> st(idx.size());
    for (size_t i = 0; i < idx.size(); ++i) {
        sp[i] = pos[idx[i]];
        st[i] = type.empty() ? 0 : type[idx[i]];
    }
    SpatialGrid g(size, r_max);
    g.build(sp);
    std::vector<Pair> pairs;
    g.neighbors(sp, r_max, pairs);
    counts.assign(nbins, 0);
    for (auto& pr : pairs) {
        if (ta >= 0) {
            int a = st[pr.a], b = st[pr.b];
            if (!((a == ta && b == tb) || (a == tb && b == ta))) continue;
        }
        double d = std::sqrt(len2(min_image(sp[pr.a] - sp[pr.b], size)));
        if (d >= r_max) continue;
        uint64_t bin = f64_as_u64(d / dr);
        if (bin >= (uint64_t)nbins) bin = nbins - 1;  // the reference would panic here
        counts[bin] += 1;
    }
    *r_max_out = r_max;
    *dr_out = dr;
    size_t n = sp.size();
    if (ta < 0) {
        *pair_norm = n < 2 ? 0.0 : (double)(n * (n - 1)) * 0.5;
    } else {
        size_t na = 0, nb = 0;
        for (int t : st) {
            na += (t == ta);
            nb += (t == tb);
        }
        *pair_norm = ta == tb ? (double)na * (double)(na > 0 ? na - 1 : 0) * 0.5 : (double)na * (double)nb;
    }
}

// Normalisation, src/analysis/mod.rs:119-132 and :184-200.
void gr_normalise(const std::vector<uint64_t>& counts, double dr, V3 size, double pair_norm, double* g) {
    double vol = size.x * size.y * size.z;
    for (size_t i = 0; i < counts.size(); ++i) {
        if (pair_norm <= 0.0) { g[i] = 0.0; continue; }
        double r_lo = (double)i * dr;
        double r_hi = r_lo + dr;
        double shell = 4.0 / 3.0 * M_PI * (r_hi * r_hi * r_hi - r_lo * r_lo * r_lo);
        double expected = pair_norm * shell / vol;
        g[i] = (double)counts[i] / fmax_r(expected, F64_MIN_POSITIVE);
    }
}

// Union by rank + path halving, src/analysis/mod.rs:287-307.
size_t uf_find(std::vector<size_t>& parent, size_t x) {
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}
void uf_union(std::vector<size_t>& parent, std::vector<uint32_t>& rank, size_t a, size_t b) {
    size_t ra = uf_find(parent, a), rb = uf_find(parent, b);
    if (ra == rb) return;
    size_t hi, lo;
    if (rank[ra] < rank[rb]) { hi = rb; lo = ra; } else { hi = ra; lo = rb; }
    parent[lo] = hi;
    if (rank[hi] == rank[lo]) rank[hi] += 1;
}

std::vector<V3> load_pos(int64_t n, const double* x, const double* y, const double* z) {
    std::vector<V3> pos((size_t)n);
    for (int64_t i = 0; i < n; ++i) pos[i] = v3(x[i], y[i], z[i]);
    return pos;
}

std::vector<std::vector<uint32_t>> load_bonds(int64_t n, const int32_t* ptr, const int32_t* idx) {
    std::vector<std::vector<uint32_t>> bond_of((size_t)n);
    if (!ptr || !idx) return bond_of;
    // Same construction as src/systems/forces.rs:208-226: each undirected
    // pair once, pushed to both ends.
    for (int64_t e = 0; e < n; ++e)
        for (int32_t k = ptr[e]; k < ptr[e + 1]; ++k) {
            int64_t m = idx[k];
            if (e >= m) continue;
            bond_of[e].push_back((uint32_t)m);
            bond_of[m].push_back((uint32_t)e);
        }
    return bond_of;
}


// ---- BondObservationSystem: src/systems/bonds.rs, src/analysis/pairs.rs -----

constexpr uint64_t DEBOUNCE = 2;  // src/systems/bonds.rs:27

using IdPair = std::pair<uint32_t, uint32_t>;  // BoundPair, ids normalised a < b

// PairTracker::track_tick, src/analysis/pairs.rs:64-90.  Episodes closed by
// the debounce are not kept (only the observation reads the tracker).
struct PairTracker {
    uint64_t debounce = DEBOUNCE;
    std::map<IdPair, uint64_t> open, missing;
    void track_tick(const std::vector<IdPair>& bound) {
        std::set<IdPair> current(bound.begin(), bound.end());
        for (const IdPair& pr : current) {
            missing.erase(pr);
            open[pr] += 1;  // and_modify(+1).or_insert(1)
        }
        std::vector<IdPair> absent;
        for (const auto& kv : open)
            if (!current.count(kv.first)) absent.push_back(kv.first);
        for (const IdPair& pr : absent) {
            uint64_t& m = missing[pr];
            m += 1;
            if (m >= debounce) {
                open.erase(pr);
                missing.erase(pr);
            }
        }
    }
};

// BondObservationSystem state and its BondObservation summary
// (src/systems/bonds.rs:29-40, src/stats bond fields).
struct BondObserver {
    PairTracker tracker;
    std::map<IdPair, uint64_t> active_since;
    std::vector<std::vector<uint32_t>> bonds;  // Bonds component, per slot
    uint64_t bonds_formed = 0;
    double lifetime_sum_ticks = 0.0;
    uint64_t tick = 0, bonded_pairs = 0, bonded_entities = 0;
    double mean_coordination = 0.0;
    std::vector<uint64_t> matrix = std::vector<uint64_t>(GS_NSPECIES * GS_NSPECIES, 0);
};

// bind_cutoff, src/analysis/pairs.rs:115-117 (max over all ten elements).
double bind_cutoff(const gs_params& p) {
    double m = 0.0;
    for (int t = 0; t < GS_NSPECIES; ++t) m = fmax_r(m, p.sigma[t]);
    return p.bond_k_bind * m;
}

// bound_pairs_with_grid, src/analysis/pairs.rs:132-174: grid candidates
// within the bind cutoff (check_pair: d2 < cutoff^2, d > EPS), kept when
// |min_image(a - b)| < k_bind * sigma_ab.  Slot index = entity id.
void bound_pairs(const gs_params& p, const std::vector<V3>& pos, const uint8_t* types, std::vector<IdPair>& out) {
    out.clear();
    std::vector<Pair> cand;
    enumerate_pairs(p, pos, bind_cutoff(p), cand);
    V3 size = v3(p.box[0], p.box[1], p.box[2]);
    for (const Pair& pr : cand) {
        double d = std::sqrt(len2(min_image(pos[pr.a] - pos[pr.b], size)));
        double thr = p.bond_k_bind * (0.5 * (p.sigma[types[pr.a]] + p.sigma[types[pr.b]]));
        if (d < thr) out.emplace_back(std::min(pr.a, pr.b), std::max(pr.a, pr.b));
    }
}

// Persistence threshold of a species pair, src/systems/bonds.rs:92-97:
// ceil(min_periods * vib_period(eps_ab, sigma_ab, mu_ab) / dt) as u64.
uint64_t persistence_ticks(const gs_params& p, int a, int b) {
    double mu = p.mass[a] * p.mass[b] / (p.mass[a] + p.mass[b]);
    double eps = p.thermal_constant * std::sqrt(p.epsilon_k[a] * p.epsilon_k[b]);
    double sig = 0.5 * (p.sigma[a] + p.sigma[b]);
    return f64_as_u64(std::ceil(fmax_r(p.bond_min_periods, 0.1) * vib_period(eps, sig, mu) / p.dt));
}

// BondObservationSystem::run, src/systems/bonds.rs:73-185, at tick `now`.
void bond_observe(const gs_params& p, const std::vector<V3>& pos, const uint8_t* types, uint64_t now,
                  BondObserver& ob) {
    std::vector<IdPair> bound;
    bound_pairs(p, pos, types, bound);
    ob.tracker.track_tick(bound);
    std::map<uint32_t, std::vector<uint32_t>> persistent;
    for (const auto& kv : ob.tracker.open) {
        const IdPair& pr = kv.first;
        if (kv.second < persistence_ticks(p, types[pr.first], types[pr.second])) continue;
        persistent[pr.first].push_back(pr.second);
        persistent[pr.second].push_back(pr.first);
    }
    ob.bonds.assign(pos.size(), std::vector<uint32_t>());
    for (auto& kv : persistent) ob.bonds[kv.first] = kv.second;
    uint64_t bonded_pairs = 0;
    std::vector<uint64_t> matrix(GS_NSPECIES * GS_NSPECIES, 0);
    std::set<IdPair> current;
    for (auto& kv : persistent)
        for (uint32_t m : kv.second) {
            IdPair pr(std::min(kv.first, m), std::max(kv.first, m));
            if (!current.insert(pr).second) continue;
            bonded_pairs += 1;
            int ia = types[kv.first], ib = types[m];
            matrix[ia * GS_NSPECIES + ib] += 1;
            matrix[ib * GS_NSPECIES + ia] += 1;
        }
    std::vector<IdPair> broken;
    for (const auto& kv : ob.active_since)
        if (!current.count(kv.first)) broken.push_back(kv.first);
    for (const IdPair& pr : broken) {
        ob.bonds_formed += 1;
        ob.lifetime_sum_ticks += (double)(now - ob.active_since[pr]);
        ob.active_since.erase(pr);
    }
    for (const IdPair& pr : current) ob.active_since.emplace(pr, now);
    ob.tick = now;
    ob.bonded_pairs = bonded_pairs;
    ob.bonded_entities = persistent.size();
    ob.mean_coordination = ob.bonded_entities > 0 ? 2.0 * (double)bonded_pairs / (double)ob.bonded_entities : 0.0;
    ob.matrix = matrix;
}

}  // namespace

// =============================================================================
// extern "C" surface used by tests/ and bench.py (ctypes).
// =============================================================================
extern "C" {

// ---- known-answer helpers (scalar laws) ----
void or_lj_raw(double sigma, double epsilon, double r, double* m, double* v) {
    lj_raw(LjPair{sigma, epsilon}, r, m, v);
}
void or_smooth_cutoff(double r, double r_on, double rc, double* sw, double* dsw) {
    smooth_cutoff(r, r_on, rc, sw, dsw);
}
void or_switched(double f, double v, double sw, double dsw, double* fo, double* vo) { switched(f, v, sw, dsw, fo, vo); }
void or_coulomb_raw(double k, double qa, double qb, double r, double* f, double* v) { coulomb_raw(k, qa, qb, r, f, v); }
void or_gravity_raw(double g, double ma, double mb, double r, double* f, double* v) { gravity_raw(g, ma, mb, r, f, v); }
void or_bond_harmonic_raw(double k, double r_eq, double r, double* f, double* v) { bond_harmonic_raw(k, r_eq, r, f, v); }
double or_well_curvature(double eps, double sigma) { return well_curvature(eps, sigma); }
double or_vib_period(double eps, double sigma, double mu) { return vib_period(eps, sigma, mu); }
double or_equilibrium_distance(double sigma) { return equilibrium_distance(sigma); }

// LjTable + ForceSystem cutoffs: sigma_ij/eps_ij (10x10 row-major), rc_lj,
// r_on, rc (grid), coulomb rc/r_on.
void or_lj_table(const gs_params* p, double* sig, double* eps, double* cut6) {
    LjTable lj(*p);
    ForceCutoffs c = force_cutoffs(*p, lj);
    for (int i = 0; i < GS_NSPECIES; ++i)
        for (int j = 0; j < GS_NSPECIES; ++j) {
            sig[i * GS_NSPECIES + j] = lj.pair[i][j].sigma;
            eps[i * GS_NSPECIES + j] = lj.pair[i][j].epsilon;
        }
    cut6[0] = lj.rc;
    cut6[1] = lj.r_on;
    cut6[2] = c.rc;
    cut6[3] = c.coulomb_rc;
    cut6[4] = c.coulomb_r_on;
    cut6[5] = c.gravity_rc;
}

// force_switched_scalar for one pair of species.
void or_force_switched_scalar(const gs_params* p, int ta, int tb, double r, double* m, double* v) {
    LjTable lj(*p);
    lj.force_switched_scalar(lj.pair[ta][tb], r, m, v);
}

// min_image of one displacement.
void or_min_image(const double* d, const double* size, double* out) {
    V3 r = min_image(v3(d[0], d[1], d[2]), v3(size[0], size[1], size[2]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

// Grid dims and per-atom cell index (cell_of + encode).
void or_grid_cells(int64_t n, const double* x, const double* y, const double* z, const double* size,
                   double min_cell, uint32_t* dims, uint32_t* cell) {
    SpatialGrid g(v3(size[0], size[1], size[2]), min_cell);
    dims[0] = g.dims[0];
    dims[1] = g.dims[1];
    dims[2] = g.dims[2];
    for (int64_t i = 0; i < n; ++i) {
        uint32_t c[3];
        g.cell_of(v3(x[i], y[i], z[i]), c);
        cell[i] = (uint32_t)g.encode(c[0], c[1], c[2]);
    }
}

// SpatialGrid::neighbors (mode 0) or brute force (mode 1).  Writes up to cap
// pairs (a, b) in enumeration order; returns the total count.
int64_t or_pairs(int64_t n, const double* x, const double* y, const double* z, const double* size,
                 double cutoff, int mode, int64_t cap, uint32_t* a, uint32_t* b) {
    std::vector<V3> pos = load_pos(n, x, y, z);
    V3 s = v3(size[0], size[1], size[2]);
    std::vector<Pair> out;
    if (mode == 1) {
        brute_pairs(pos, s, cutoff, out);
    } else {
        SpatialGrid g(s, cutoff);
        g.build(pos);
        g.neighbors(pos, cutoff, out);
    }
    for (size_t k = 0; k < out.size() && (int64_t)k < cap; ++k) {
        a[k] = out[k].a;
        b[k] = out[k].b;
    }
    return (int64_t)out.size();
}

// ForceSystem::run on the given state.  Outputs in slot order.  bond_ptr /
// bond_idx: optional CSR adjacency of the Bonds component (slot indices).
int or_forces(const gs_params* p, int64_t n, const double* x, const double* y, const double* z,
              const uint8_t* species, const double* charge, const int32_t* bond_ptr, const int32_t* bond_idx,
              int nthreads, double* ax, double* ay, double* az, double* potential, double* bond_energy,
              double* pe_atom, double* s_abs, int64_t* npairs) {
    std::vector<V3> pos = load_pos(n, x, y, z);
    std::vector<std::vector<uint32_t>> bonds = load_bonds(n, bond_ptr, bond_idx);
    ForceOut out;
    bool extras = pe_atom || s_abs;
    force_pass(*p, pos, species, charge, bond_ptr ? &bonds : nullptr, nthreads, extras, out);
    for (int64_t i = 0; i < n; ++i) {
        if (ax) ax[i] = out.acc[i].x;
        if (ay) ay[i] = out.acc[i].y;
        if (az) az[i] = out.acc[i].z;
        if (pe_atom) pe_atom[i] = out.pe[i];
        if (s_abs) s_abs[i] = out.sabs[i];
    }
    if (potential) *potential = out.potential;
    if (bond_energy) *bond_energy = out.bond_potential;
    if (npairs) *npairs = out.npairs;
    return 0;
}

// Individual systems on SoA arrays (in place).
void or_half_kick(int64_t n, double* vx, double* vy, double* vz, const double* ax, const double* ay,
                  const double* az, double dt) {
    double half_dt = 0.5 * dt;
    for (int64_t i = 0; i < n; ++i) {
        vx[i] = vx[i] + ax[i] * half_dt;
        vy[i] = vy[i] + ay[i] * half_dt;
        vz[i] = vz[i] + az[i] * half_dt;
    }
}
void or_drift(int64_t n, double* x, double* y, double* z, const double* vx, const double* vy, const double* vz,
              double dt) {
    for (int64_t i = 0; i < n; ++i) {
        x[i] = x[i] + vx[i] * dt;
        y[i] = y[i] + vy[i] * dt;
        z[i] = z[i] + vz[i] * dt;
    }
}
void or_wrap(int64_t n, double* x, double* y, double* z, const double* box) {
    double h[3] = {box[0] * 0.5, box[1] * 0.5, box[2] * 0.5};
    for (int64_t i = 0; i < n; ++i) {
        x[i] = rem_euclid(x[i], 2.0 * h[0]) - h[0];
        y[i] = rem_euclid(y[i], 2.0 * h[1]) - h[1];
        z[i] = rem_euclid(z[i], 2.0 * h[2]) - h[2];
    }
}
double or_kinetic(int64_t n, const double* vx, const double* vy, const double* vz, const uint8_t* species,
                  const double* mass) {
    double k = 0.0;
    for (int64_t i = 0; i < n; ++i) k += 0.5 * mass[species[i]] * len2(v3(vx[i], vy[i], vz[i]));
    return k;
}
double or_thermostat(const gs_params* p, int64_t n, double* vx, double* vy, double* vz, const uint8_t* species) {
    std::vector<V3> v((size_t)n);
    for (int64_t i = 0; i < n; ++i) v[i] = v3(vx[i], vy[i], vz[i]);
    double lambda = thermostat(*p, v, species);
    for (int64_t i = 0; i < n; ++i) {
        vx[i] = v[i].x;
        vy[i] = v[i].y;
        vz[i] = v[i].z;
    }
    return lambda;
}

// Universe::tick x nsteps with the schedule of build_schedule
// (src/universe/mod.rs:141-159,355-392): kick, drift, [wrap], forces, kick,
// [thermostat], [bond observation], stats.  obs: 6 doubles per step (K, V,
// Vbond, T, lambda, E), may be NULL.  State arrays are updated in place.
// With bond observation on, the Bonds written at tick n feed the force pass
// of tick n + 1 (the tracker starts empty, at tick tick0 + 1); bstats gets
// per step: tick, bonded_pairs, bonded_entities, bonds_formed,
// lifetime_sum_ticks, mean_coordination, then the 100 species-matrix counts
// (106 doubles), and bond_ptr / bond_idx (n + 1 and up to bond_cap entries)
// the final Bonds adjacency, neighbours ascending.
int or_run_observed(const gs_params* p, int64_t n, double* x, double* y, double* z, double* vx, double* vy,
                    double* vz, double* ax, double* ay, double* az, const uint8_t* species, const double* charge,
                    int64_t nsteps, int nthreads, double* obs, int64_t tick0, double* bstats, int32_t* bond_ptr,
                    int32_t* bond_idx, int64_t bond_cap) {
    std::vector<V3> pos = load_pos(n, x, y, z), vel = load_pos(n, vx, vy, vz), acc = load_pos(n, ax, ay, az);
    ForceOut out;
    BondObserver ob;
    ob.bonds.assign((size_t)n, std::vector<uint32_t>());
    const bool observe_bonds = p->enable_bond_observation != 0;
    for (int64_t s = 0; s < nsteps; ++s) {
        half_kick(vel, acc, p->dt);
        drift(pos, vel, p->dt);
        if (p->enable_boundaries) wrap(pos, p->box);
        force_pass(*p, pos, species, charge, p->enable_bond_interaction ? &ob.bonds : nullptr, nthreads, false, out);
        acc = out.acc;
        half_kick(vel, acc, p->dt);
        double lambda = 1.0;
        if (p->enable_thermostat) lambda = thermostat(*p, vel, species);
        if (observe_bonds) {
            bond_observe(*p, pos, species, (uint64_t)(tick0 + s + 1), ob);
            if (bstats) {
                double* b = bstats + 106 * s;
                b[0] = (double)ob.tick;
                b[1] = (double)ob.bonded_pairs;
                b[2] = (double)ob.bonded_entities;
                b[3] = (double)ob.bonds_formed;
                b[4] = ob.lifetime_sum_ticks;
                b[5] = ob.mean_coordination;
                for (int k = 0; k < GS_NSPECIES * GS_NSPECIES; ++k) b[6 + k] = (double)ob.matrix[k];
            }
        }
        if (obs) {
            double k = kinetic(vel, species, p->mass);
            double t = p->thermal_constant > 0.0 ? (2.0 / 3.0) * (k / (double)n) / p->thermal_constant : 0.0;
            double* o = obs + 6 * s;
            o[0] = k;
            o[1] = out.potential;
            o[2] = out.bond_potential;
            o[3] = t;
            o[4] = lambda;
            o[5] = k + out.potential + out.bond_potential;
        }
    }
    for (int64_t i = 0; i < n; ++i) {
        x[i] = pos[i].x; y[i] = pos[i].y; z[i] = pos[i].z;
        vx[i] = vel[i].x; vy[i] = vel[i].y; vz[i] = vel[i].z;
        ax[i] = acc[i].x; ay[i] = acc[i].y; az[i] = acc[i].z;
    }
    if (bond_ptr) {
        int64_t off = 0;
        for (int64_t i = 0; i < n; ++i) {
            bond_ptr[i] = (int32_t)off;
            std::vector<uint32_t> nb = ob.bonds[(size_t)i];
            std::sort(nb.begin(), nb.end());
            for (uint32_t m : nb) {
                if (off >= bond_cap) return -1;
                bond_idx[off++] = (int32_t)m;
            }
        }
        bond_ptr[n] = (int32_t)off;
    }
    return 0;
}

int or_run(const gs_params* p, int64_t n, double* x, double* y, double* z, double* vx, double* vy, double* vz,
           double* ax, double* ay, double* az, const uint8_t* species, const double* charge, int64_t nsteps,
           int nthreads, double* obs) {
    return or_run_observed(p, n, x, y, z, vx, vy, vz, ax, ay, az, species, charge, nsteps, nthreads, obs, 0, nullptr,
                           nullptr, nullptr, 0);
}

// BondObservationSystem::run alone, driven by explicit position frames (the
// way its own test drives it, src/systems/bonds.rs:263-343): frame f is
// observed at tick tick0 + f + 1.  bstats as in or_run_observed.
int or_observe_frames(const gs_params* p, int64_t n, int64_t nframes, const double* x, const double* y,
                      const double* z, const uint8_t* species, int64_t tick0, double* bstats, int32_t* bond_ptr,
                      int32_t* bond_idx, int64_t bond_cap) {
    BondObserver ob;
    ob.bonds.assign((size_t)n, std::vector<uint32_t>());
    for (int64_t f = 0; f < nframes; ++f) {
        std::vector<V3> pos = load_pos(n, x + f * n, y + f * n, z + f * n);
        bond_observe(*p, pos, species, (uint64_t)(tick0 + f + 1), ob);
        double* b = bstats + 106 * f;
        b[0] = (double)ob.tick;
        b[1] = (double)ob.bonded_pairs;
        b[2] = (double)ob.bonded_entities;
        b[3] = (double)ob.bonds_formed;
        b[4] = ob.lifetime_sum_ticks;
        b[5] = ob.mean_coordination;
        for (int k = 0; k < GS_NSPECIES * GS_NSPECIES; ++k) b[6 + k] = (double)ob.matrix[k];
    }
    if (bond_ptr) {
        int64_t off = 0;
        for (int64_t i = 0; i < n; ++i) {
            bond_ptr[i] = (int32_t)off;
            std::vector<uint32_t> nb = ob.bonds[(size_t)i];
            std::sort(nb.begin(), nb.end());
            for (uint32_t m : nb) {
                if (off >= bond_cap) return -1;
                bond_idx[off++] = (int32_t)m;
            }
        }
        bond_ptr[n] = (int32_t)off;
    }
    return 0;
}

// bound_pairs_with_grid on one state: the normalised (a < b) pairs, sorted.
int64_t or_bound_pairs(const gs_params* p, int64_t n, const double* x, const double* y, const double* z,
                       const uint8_t* species, uint32_t* pa, uint32_t* pb, int64_t cap) {
    std::vector<V3> pos = load_pos(n, x, y, z);
    std::vector<IdPair> b;
    bound_pairs(*p, pos, species, b);
    std::set<IdPair> s(b.begin(), b.end());
    int64_t k = 0;
    for (const IdPair& pr : s) {
        if (k < cap) {
            pa[k] = pr.first;
            pb[k] = pr.second;
        }
        ++k;
    }
    return k;
}

// persistence threshold (ticks) of a species pair.
uint64_t or_persistence_ticks(const gs_params* p, int a, int b) { return persistence_ticks(*p, a, b); }

// PairTracker fed with explicit per-tick pair lists (the tracker's own
// unit-test interface): ticks concatenated in pa/pb with per-tick counts;
// returns the open count after the last tick and writes open (pair, ticks).
int64_t or_track(int64_t nticks, const int64_t* counts, const uint32_t* pa, const uint32_t* pb, uint64_t debounce,
                 uint32_t* open_a, uint32_t* open_b, uint64_t* open_ticks, int64_t cap) {
    PairTracker t;
    t.debounce = debounce < 1 ? 1 : debounce;
    int64_t off = 0;
    for (int64_t k = 0; k < nticks; ++k) {
        std::vector<IdPair> b;
        for (int64_t e = 0; e < counts[k]; ++e, ++off) b.emplace_back(pa[off], pb[off]);
        t.track_tick(b);
    }
    int64_t m = 0;
    for (const auto& kv : t.open) {
        if (m < cap) {
            open_a[m] = kv.first.first;
            open_b[m] = kv.first.second;
            open_ticks[m] = kv.second;
        }
        ++m;
    }
    return m;
}

// radial_distribution / radial_distribution_between.  species may be NULL
// for the total g(r); ta < 0 selects the total.
int or_radial_distribution(int64_t n, const double* x, const double* y, const double* z, const uint8_t* species,
                           const double* size, double r_max, int nbins, int ta, int tb, uint64_t* counts, double* g,
                           double* r_max_out, double* dr_out, int full) {
    std::vector<V3> pos = load_pos(n, x, y, z);
    std::vector<int> type;
    if (species) {
        type.resize((size_t)n);
        for (int64_t i = 0; i < n; ++i) type[i] = species[i];
    }
    std::vector<uint64_t> c;
    double rm, dr, norm;
    V3 s = v3(size[0], size[1], size[2]);
    gr_counts(pos, type, s, r_max, nbins, ta, tb, c, &rm, &dr, &norm, full != 0);
    if (counts) std::memcpy(counts, c.data(), c.size() * sizeof(uint64_t));
    if (g) gr_normalise(c, dr, s, norm, g);
    if (r_max_out) *r_max_out = rm;
    if (dr_out) *dr_out = dr;
    return (int)c.size();
}

// clusters, src/analysis/mod.rs:221-284, plus canonical labels (smallest
// slot index of each component).  stats: monomers, aggregates, largest,
// mean_size, bound_pairs (as doubles).
int or_clusters(int64_t n, const double* x, const double* y, const double* z, const uint8_t* species,
                const double* sigma_table, const double* size, double bond_factor, uint32_t* labels, double* stats) {
    for (int k = 0; k < 5; ++k) stats[k] = 0.0;
    if (n == 0) return 0;
    std::vector<V3> pos = load_pos(n, x, y, z);
    V3 s = v3(size[0], size[1], size[2]);
    std::vector<double> sig((size_t)n);
    double max_sigma = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        sig[i] = sigma_table[species[i]];
        max_sigma = fmax_r(max_sigma, sig[i]);
    }
    double cutoff = fmax_r(fmax_r(bond_factor, 0.1) * max_sigma, 1e-9);
    SpatialGrid g(s, cutoff);
    g.build(pos);
    std::vector<Pair> pairs;
    g.neighbors(pos, cutoff, pairs);
    std::vector<size_t> parent((size_t)n);
    for (int64_t i = 0; i < n; ++i) parent[i] = (size_t)i;
    std::vector<uint32_t> rank((size_t)n, 0);
    int64_t bound_pairs = 0;
    for (auto& pr : pairs) {
        double d = std::sqrt(len2(min_image(pos[pr.a] - pos[pr.b], s)));
        double bond = bond_factor * 0.5 * (sig[pr.a] + sig[pr.b]);
        if (d < bond) {
            uf_union(parent, rank, pr.a, pr.b);
            bound_pairs += 1;
        }
    }
    std::vector<size_t> sizes((size_t)n, 0);
    std::vector<uint32_t> minidx((size_t)n, 0xFFFFFFFFu);
    for (int64_t i = 0; i < n; ++i) {
        size_t r = uf_find(parent, (size_t)i);
        sizes[r] += 1;
        if (minidx[r] == 0xFFFFFFFFu) minidx[r] = (uint32_t)i;
    }
    if (labels)
        for (int64_t i = 0; i < n; ++i) labels[i] = minidx[uf_find(parent, (size_t)i)];
    int64_t monomers = 0, aggregates = 0, largest = 0, total = 0;
    for (size_t sz : sizes) {
        if (sz == 0) continue;
        total += 1;
        if (sz == 1) monomers += 1; else aggregates += 1;
        largest = std::max<int64_t>(largest, (int64_t)sz);
    }
    stats[0] = (double)monomers;
    stats[1] = (double)aggregates;
    stats[2] = (double)largest;
    stats[3] = (double)n / (double)std::max<int64_t>(total, 1);
    stats[4] = (double)bound_pairs;
    return 0;
}

// bond_components, src/analysis/mod.rs:337-412.  entities: ids + element
// index; edges: pairs of ids.  Writes, per component in output order (size
// desc, then formula asc), its size into comp_size and its members (sorted
// ids) concatenated into members.  Returns the number of components.
// pair_potential_raw, src/physics/forces.rs:360-384: the raw (unswitched)
// LJ term of the mixed pair, plus k_e q_a q_b / r with the table's species
// charges when with_coulomb, plus 0.5 k (r - r_eq)^2 with the pair's well
// curvature and equilibrium distance when with_bond.
double or_pair_potential_raw(const gs_params* p, int a, int b, double r, int with_coulomb, int with_bond) {
    const double sig = 0.5 * (p->sigma[a] + p->sigma[b]);
    const double eps = p->thermal_constant * std::sqrt(p->epsilon_k[a] * p->epsilon_k[b]);
    double m, v;
    lj_raw(LjPair{sig, eps}, r, &m, &v);
    double total = v;
    if (with_coulomb) {
        const double q = p->coulomb_constant * p->charge[a] * p->charge[b];
        total += q / r;
    }
    if (with_bond) {
        const double k = well_curvature(eps, sig);
        const double r_eq = equilibrium_distance(sig);
        total += 0.5 * k * (r - r_eq) * (r - r_eq);
    }
    return total;
}

// The binding energy of each bonded edge (a, b) as BondStructureSystem::run
// evaluates it (src/systems/bond_structure.rs:101-115): r =
// max(|min_image(x_a - x_b)|, 1e-9), pair_potential_raw with the system's
// Coulomb and bond switches.
void or_bond_binding(const gs_params* p, int64_t n, const double* x, const double* y, const double* z,
                     const uint8_t* species, int64_t nedge, const uint32_t* ea, const uint32_t* eb, double* out) {
    const V3 size = v3(p->box[0], p->box[1], p->box[2]);
    for (int64_t k = 0; k < nedge; ++k) {
        const uint32_t a = ea[k], b = eb[k];
        if ((int64_t)a >= n || (int64_t)b >= n) {
            out[k] = 0.0;
            continue;
        }
        const V3 d = min_image(v3(x[a], y[a], z[a]) - v3(x[b], y[b], z[b]), size);
        const double r = fmax_r(std::sqrt(len2(d)), 1e-9);
        out[k] = or_pair_potential_raw(p, species[a], species[b], r, p->enable_electrostatics,
                                       p->enable_bond_interaction);
    }
}

int or_bond_components(int64_t ne, const uint32_t* ent_id, const uint8_t* ent_species, int64_t nedge,
                       const uint32_t* ea, const uint32_t* eb, int64_t* comp_size, uint32_t* members,
                       char* formulas, int64_t formula_stride) {
    if (ne == 0) return 0;
    std::map<uint32_t, size_t> index;
    for (int64_t i = 0; i < ne; ++i) index[ent_id[i]] = (size_t)i;
    std::vector<size_t> parent((size_t)ne);
    for (int64_t i = 0; i < ne; ++i) parent[i] = (size_t)i;
    std::vector<uint32_t> rank((size_t)ne, 0);
    std::set<std::pair<uint32_t, uint32_t>> seen;
    for (int64_t k = 0; k < nedge; ++k) {
        auto ia = index.find(ea[k]), ib = index.find(eb[k]);
        if (ia == index.end() || ib == index.end()) continue;
        if (ia->second == ib->second) continue;
        auto edge = std::make_pair(std::min(ea[k], eb[k]), std::max(ea[k], eb[k]));
        if (!seen.insert(edge).second) continue;
        uf_union(parent, rank, ia->second, ib->second);
    }
    static const char* SYM[GS_NSPECIES] = {"H", "He", "C", "N", "O", "Na", "Si", "P", "S", "Fe"};
    std::map<size_t, std::pair<std::vector<uint32_t>, std::vector<uint32_t>>> comps;  // root -> (composition, members)
    for (int64_t i = 0; i < ne; ++i) {
        size_t r = uf_find(parent, (size_t)i);
        auto& c = comps[r];
        if (c.first.empty()) c.first.assign(GS_NSPECIES, 0);
        c.second.push_back(ent_id[i]);
        c.first[ent_species[i]] += 1;
    }
    struct Out {
        std::vector<uint32_t> members;
        std::string formula;
    };
    std::vector<Out> out;
    for (auto& kv : comps) {
        // formula, src/analysis/mod.rs:417-435.
        std::vector<std::pair<uint32_t, int>> parts;
        for (int t = 0; t < GS_NSPECIES; ++t)
            if (kv.second.first[t] > 0) parts.emplace_back(kv.second.first[t], t);
        std::sort(parts.begin(), parts.end(), [](const std::pair<uint32_t, int>& a, const std::pair<uint32_t, int>& b) {
            if (a.first != b.first) return a.first > b.first;
            return std::strcmp(SYM[a.second], SYM[b.second]) < 0;
        });
        std::string f;
        for (size_t k = 0; k < parts.size(); ++k) {
            if (k) f += "-";
            f += SYM[parts[k].second];
            if (parts[k].first > 1) f += std::to_string(parts[k].first);
        }
        Out o;
        o.members = kv.second.second;
        std::sort(o.members.begin(), o.members.end());
        o.formula = f;
        out.push_back(o);
    }
    // size desc, then formula (src/analysis/mod.rs:411); components equal on
    // both keep the reference's hash order there, here the smallest member
    std::sort(out.begin(), out.end(), [](const Out& a, const Out& b) {
        if (a.members.size() != b.members.size()) return a.members.size() > b.members.size();
        if (a.formula != b.formula) return a.formula < b.formula;
        return a.members.front() < b.members.front();
    });
    int64_t off = 0;
    for (size_t k = 0; k < out.size(); ++k) {
        comp_size[k] = (int64_t)out[k].members.size();
        for (uint32_t m : out[k].members) members[off++] = m;
        if (formulas) {
            std::strncpy(formulas + k * formula_stride, out[k].formula.c_str(), (size_t)formula_stride - 1);
            formulas[k * formula_stride + formula_stride - 1] = 0;
        }
    }
    return (int)out.size();
}

}  // extern "C"
