/*
 * adjoint.cpp — CPU FP64 reference gradient for the differentiable form of
 * the path (TEST INFRASTRUCTURE ONLY, like oracle.cpp).  The reference crate
 * has no adjoint; this file derives it from the reference's force law
 * (src/physics/forces.rs:270-353, src/systems/forces.rs:270-289):
 *
 *   phi(d) = V_LJ(r) sw(r) + V_C(d) sw(d),  r = max(d, sigma/2),
 *   V_LJ = 4 eps (s^12 - s^6), s = sigma/r,  V_C = qq / max(d, 1/2),
 *   sw = quintic switch on [0.9 rc, rc],  rc = cutoff_factor * max_s sigma_s
 *
 * with Lorentz-Berthelot sigma_ij = (sigma_i + sigma_j)/2,
 * eps_ij = k sqrt(epsK_i epsK_j) and qq = k_e q_i q_j.  Parameters theta are
 * the 10 per-species sigma, 10 epsK and 10 charges; an atom's charge is
 * keep_i * q_{s(i)} (keep_i in {0,1} records seeded neutralisation).
 *
 * Trajectory: velocity Verlet with a_0 = a(x_0), wrap after each drift.
 * Loss: mean V over ticks [win_from, T] + sum_k (g_k(x_T) - g0_k)^2 with a
 * Gaussian-smoothed g(r): g_k = sum_{i<j} c_k K(d_ij - r_k), K the normal
 * density of width sigma_b cut where it falls below 1.3e-14 of its peak
 * (bin centres within 8 sigma_b of d_ij).  Pairs are enumerated on a cell
 * grid (all pairs when an axis has fewer than 3 cells).  The gradient is the
 * discrete adjoint of the kick-drift-kick map; the two kicks at the same
 * positions are merged, so each tick costs one Hessian-vector product and one
 * parameter VJP.
 *
 * Pinned by: finite differences of every per-pair derivative, of the HVP
 * (forces at x +- h w), of the parameter VJP, and central differences of
 * the whole loss (tests/test_oracle_adjoint.py).
 */
#include "../include/genesis_gpu.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct P3 {
    double x, y, z;
};

// Derivatives of one pair term at distance d.
struct PairD {
    double phi, d1, d2;              // phi, dphi/dd, d2phi/dd2
    double phi_s, phi_e, phi_q, phi_rc;  // d phi / d(sigma_ij, eps_ij, qq, rc) at fixed d
    double f_s, f_e, f_q, f_rc;      // same for f = -dphi/dd
};

inline void quintic(double u, double& S, double& S1, double& S2) {
    if (u <= 0.0) { S = 1.0; S1 = 0.0; S2 = 0.0; return; }
    if (u >= 1.0) { S = 0.0; S1 = 0.0; S2 = 0.0; return; }
    double u2 = u * u, u3 = u2 * u;
    S = 1.0 - 10.0 * u3 + 15.0 * u3 * u - 6.0 * u3 * u2;
    S1 = -30.0 * u2 * (1.0 - u) * (1.0 - u);
    S2 = -60.0 * u * (1.0 - u) * (1.0 - 2.0 * u);
}

void pair_derivs(double d, double sigma, double eps, double qq, double rc, bool coul, PairD& o) {
    std::memset(&o, 0, sizeof(o));
    const double w = 0.1 * rc;  // rc - r_on
    double S, S1, S2;
    quintic((d - 0.9 * rc) / w, S, S1, S2);
    const double sw = S, sw1 = S1 / w, sw2 = S2 / (w * w);
    const double swr = -(d / rc) * sw1;            // d sw / d rc
    const double sw1r = -(d / rc) * sw2 - sw1 / rc;  // d sw1 / d rc
    if (d >= 0.5 * sigma) {
        const double s = sigma / d, s6 = s * s * s * s * s * s, s12 = s6 * s6;
        const double V = 4.0 * eps * (s12 - s6);
        const double V1 = -(24.0 * eps / d) * (2.0 * s12 - s6);
        const double V2 = (4.0 * eps / (d * d)) * (156.0 * s12 - 42.0 * s6);
        const double Vs = (24.0 * eps / sigma) * (2.0 * s12 - s6);
        const double V1s = -(24.0 * eps / (d * sigma)) * (24.0 * s12 - 6.0 * s6);
        const double Ve = 4.0 * (s12 - s6);
        const double V1e = -(24.0 / d) * (2.0 * s12 - s6);
        o.phi += V * sw;
        o.d1 += V1 * sw + V * sw1;
        o.d2 += V2 * sw + 2.0 * V1 * sw1 + V * sw2;
        o.phi_s += Vs * sw;
        o.phi_e += Ve * sw;
        o.phi_rc += V * swr;
        o.f_s += -(V1s * sw + Vs * sw1);
        o.f_e += -(V1e * sw + Ve * sw1);
        o.f_rc += -(V1 * swr + V * sw1r);
    } else {
        // hardened core: r = sigma/2, s = 2, below the switching shell
        o.phi += 4.0 * eps * (4096.0 - 64.0);
        o.phi_e += 4.0 * (4096.0 - 64.0);
    }
    if (coul) {
        const double rcl = std::fmax(d, 0.5);
        const double Vq = 1.0 / rcl;                      // V_C / qq
        const double V1q = d >= 0.5 ? -1.0 / (d * d) : 0.0;
        const double V2q = d >= 0.5 ? 2.0 / (d * d * d) : 0.0;
        o.phi += qq * Vq * sw;
        o.d1 += qq * (V1q * sw + Vq * sw1);
        o.d2 += qq * (V2q * sw + 2.0 * V1q * sw1 + Vq * sw2);
        o.phi_q += Vq * sw;
        o.phi_rc += qq * Vq * swr;
        o.f_q += -(V1q * sw + Vq * sw1);
        o.f_rc += -qq * (V1q * swr + Vq * sw1r);
    }
}

struct Model {
    int n;
    double box[3];
    double dt, k_e, kconst, cut_factor;
    bool coul;
    double sig[GS_NSPECIES], epsK[GS_NSPECIES], qs[GS_NSPECIES], mass[GS_NSPECIES];
    double sig_ij[GS_NSPECIES][GS_NSPECIES], eps_ij[GS_NSPECIES][GS_NSPECIES];
    double rc;
    int argmax;
    std::vector<uint8_t> type, keep;
    std::vector<double> q;  // per-atom charge

    void setup(const gs_params& p, const double* theta) {
        for (int s = 0; s < GS_NSPECIES; ++s) {
            sig[s] = theta[s];
            epsK[s] = theta[10 + s];
            qs[s] = theta[20 + s];
            mass[s] = p.mass[s];
        }
        box[0] = p.box[0];
        box[1] = p.box[1];
        box[2] = p.box[2];
        dt = p.dt;
        k_e = p.coulomb_constant;
        kconst = p.thermal_constant;
        cut_factor = p.lj_cutoff_factor;
        coul = p.enable_electrostatics != 0;
        double ms = 0.0;
        argmax = 0;
        for (int s = 0; s < GS_NSPECIES; ++s)
            if (sig[s] > ms) { ms = sig[s]; argmax = s; }
        rc = cut_factor * ms;
        for (int i = 0; i < GS_NSPECIES; ++i)
            for (int j = 0; j < GS_NSPECIES; ++j) {
                sig_ij[i][j] = 0.5 * (sig[i] + sig[j]);
                eps_ij[i][j] = kconst * std::sqrt(epsK[i] * epsK[j]);
            }
        q.resize(n);
        for (int i = 0; i < n; ++i) q[i] = keep[i] ? qs[type[i]] : 0.0;
    }
    P3 mi(const P3& a, const P3& b) const {
        P3 d{a.x - b.x, a.y - b.y, a.z - b.z};
        d.x -= box[0] * std::round(d.x / box[0]);
        d.y -= box[1] * std::round(d.y / box[1]);
        d.z -= box[2] * std::round(d.z / box[2]);
        return d;
    }
    void pd(int i, int j, double d, PairD& o) const {
        int ti = type[i], tj = type[j];
        pair_derivs(d, sig_ij[ti][tj], eps_ij[ti][tj], k_e * q[i] * q[j], rc, coul, o);
    }
};

// Every unordered pair (i < j) closer than `cutoff` (minimum image): a cell
// grid of edge >= cutoff when every axis has at least 3 cells, else all
// pairs.  The set is exactly the pairs the all-pairs loop would keep; only
// the order differs (sums then differ in rounding only).
struct PairList {
    std::vector<int> a, b;
    std::vector<double> d;
    std::vector<P3> dv;
};

void pair_list(const Model& M, const std::vector<P3>& x, double cutoff, bool keep_zero, PairList& out) {
    out.a.clear();
    out.b.clear();
    out.d.clear();
    out.dv.clear();
    int dims[3];
    for (int k = 0; k < 3; ++k) dims[k] = std::max(1, (int)std::floor(M.box[k] / cutoff));
    auto consider = [&](int i, int j) {
        P3 dv = M.mi(x[i], x[j]);
        double d = std::sqrt(dv.x * dv.x + dv.y * dv.y + dv.z * dv.z);
        if (d >= cutoff) return;
        if (!keep_zero && d <= 2.220446049250313e-16) return;
        out.a.push_back(std::min(i, j));
        out.b.push_back(std::max(i, j));
        if (i < j) {
            out.d.push_back(d);
            out.dv.push_back(dv);
        } else {
            out.d.push_back(d);
            out.dv.push_back(P3{-dv.x, -dv.y, -dv.z});
        }
    };
    if (dims[0] < 3 || dims[1] < 3 || dims[2] < 3) {
        for (int i = 0; i < M.n; ++i)
            for (int j = i + 1; j < M.n; ++j) consider(i, j);
        return;
    }
    const int ncell = dims[0] * dims[1] * dims[2];
    std::vector<std::vector<int>> cell(ncell);
    auto cidx = [&](const P3& p, int c[3]) {
        const double v[3] = {p.x, p.y, p.z};
        for (int k = 0; k < 3; ++k) {
            double u = std::fmod(v[k] + 0.5 * M.box[k], M.box[k]);
            if (u < 0.0) u += M.box[k];
            c[k] = std::min(dims[k] - 1, (int)std::floor(u / (M.box[k] / dims[k])));
        }
    };
    for (int i = 0; i < M.n; ++i) {
        int c[3];
        cidx(x[i], c);
        cell[c[0] + dims[0] * (c[1] + dims[1] * c[2])].push_back(i);
    }
    for (int i = 0; i < M.n; ++i) {
        int c[3];
        cidx(x[i], c);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int ix = (c[0] + dx + dims[0]) % dims[0], iy = (c[1] + dy + dims[1]) % dims[1],
                              iz = (c[2] + dz + dims[2]) % dims[2];
                    for (int j : cell[ix + dims[0] * (iy + dims[1] * iz)])
                        if (j > i) consider(i, j);
                }
    }
}

// Accelerations and V over the pairs within rc.
double accel(const Model& M, const std::vector<P3>& x, std::vector<P3>& a) {
    a.assign(M.n, P3{0, 0, 0});
    double V = 0.0;
    PairList pl;
    pair_list(M, x, M.rc, false, pl);
    for (size_t k = 0; k < pl.a.size(); ++k) {
        const int i = pl.a[k], j = pl.b[k];
        const P3 dv = pl.dv[k];
        const double d = pl.d[k];
        PairD o;
        M.pd(i, j, d, o);
        V += o.phi;
        double f = -o.d1 / d;  // along the unnormalised displacement
        a[i].x += f * dv.x; a[i].y += f * dv.y; a[i].z += f * dv.z;
        a[j].x -= f * dv.x; a[j].y -= f * dv.y; a[j].z -= f * dv.z;
    }
    for (int i = 0; i < M.n; ++i) {
        double im = 1.0 / M.mass[M.type[i]];
        a[i].x *= im; a[i].y *= im; a[i].z *= im;
    }
    return V;
}

// One adjoint pass at positions x with weight vector lam (an adjoint of the
// velocities): out_x = (da/dx)^T lam (a Hessian-vector product with
// u = lam/m), gth += (da/dtheta)^T lam; with vcoef != 0 also
// gth += vcoef * dV/dtheta and out_x -= vcoef * F (the loss's dV/dx).
void adjoint_pass(const Model& M, const std::vector<P3>& x, const std::vector<P3>& lam, double vcoef,
                  std::vector<P3>& out_x, double* gth) {
    out_x.assign(M.n, P3{0, 0, 0});
    double grc = 0.0;
    PairList pl;
    pair_list(M, x, M.rc, false, pl);
    for (size_t k = 0; k < 2 * pl.a.size(); ++k) {
        // every ordered pair: (a, b) then (b, a)
        const size_t m = k >> 1;
        const bool rev = (k & 1) != 0;
        const int i = rev ? pl.b[m] : pl.a[m], j = rev ? pl.a[m] : pl.b[m];
        const P3 dv = rev ? P3{-pl.dv[m].x, -pl.dv[m].y, -pl.dv[m].z} : pl.dv[m];
        const double d = pl.d[m];
        PairD o;
        M.pd(i, j, d, o);
        const int ti = M.type[i], tj = M.type[j];
        P3 r{dv.x / d, dv.y / d, dv.z / d};
        const double imi = 1.0 / M.mass[ti], imj = 1.0 / M.mass[tj];
        P3 du{lam[i].x * imi - lam[j].x * imj, lam[i].y * imi - lam[j].y * imj, lam[i].z * imi - lam[j].z * imj};
        const double rdu = r.x * du.x + r.y * du.y + r.z * du.z;
        // K du with K = f' rr^T + (f/d)(I - rr^T), f = -d1, f' = -d2
        const double fprime = -o.d2, fd = -o.d1 / d;
        out_x[i].x += fprime * rdu * r.x + fd * (du.x - rdu * r.x);
        out_x[i].y += fprime * rdu * r.y + fd * (du.y - rdu * r.y);
        out_x[i].z += fprime * rdu * r.z + fd * (du.z - rdu * r.z);
        // per ordered pair, credited to species ti
        const double Xs = rdu * o.f_s + vcoef * o.phi_s;
        const double Xe = rdu * o.f_e + vcoef * o.phi_e;
        const double Xq = rdu * o.f_q + vcoef * o.phi_q;
        const double Xr = rdu * o.f_rc + vcoef * o.phi_rc;
        gth[ti] += 0.5 * Xs;  // d sigma_ij / d sigma_ti = 1/2 (each ordered pair credits its first atom)
        gth[10 + ti] += 0.5 * Xe * M.eps_ij[ti][tj] / M.epsK[ti];
        gth[20 + ti] += Xq * M.k_e * (M.keep[i] ? 1.0 : 0.0) * M.q[j];
        grc += 0.5 * Xr;
        if (vcoef != 0.0) {
            // loss term vcoef * V: dV/dx_i = -F_i
            const double f = -o.d1;
            out_x[i].x -= vcoef * f * r.x;
            out_x[i].y -= vcoef * f * r.y;
            out_x[i].z -= vcoef * f * r.z;
        }
    }
    gth[M.argmax] += grc * M.cut_factor;
}

// Gaussian-smoothed g(r) on the final positions and its position gradient.
struct SmoothGr {
    int nb;
    double rmax, sb, dr;
    std::vector<double> c;  // dr / (npairs shell_k / vol)
    void setup(const Model& M, int nbins, double r_max, double sigma_b) {
        nb = nbins;
        rmax = r_max;
        sb = sigma_b;
        dr = rmax / nb;
        const double vol = M.box[0] * M.box[1] * M.box[2];
        const double np = 0.5 * (double)M.n * (double)(M.n - 1);
        c.resize(nb);
        for (int k = 0; k < nb; ++k) {
            double lo = k * dr, hi = lo + dr;
            double shell = 4.0 / 3.0 * M_PI * (hi * hi * hi - lo * lo * lo);
            c[k] = dr / (np * shell / vol);
        }
    }
    double K(double u) const { return std::exp(-u * u / (2.0 * sb * sb)) / (std::sqrt(2.0 * M_PI) * sb); }
    // The kernel's support: bins whose centre is within SGR_WIN sigma_b of d
    // (beyond it the Gaussian is below 1.3e-14 of its peak).
    static constexpr double SGR_WIN = 8.0;
    void window(double d, int& k0, int& k1) const {
        const double w = SGR_WIN * sb;
        k0 = std::max(0, (int)std::ceil((d - w) / dr - 0.5));
        k1 = std::min(nb - 1, (int)std::floor((d + w) / dr - 0.5));
    }
    double lim() const { return rmax + SGR_WIN * sb; }
    void eval(const Model& M, const std::vector<P3>& x, std::vector<double>& g) const {
        g.assign(nb, 0.0);
        PairList pl;
        pair_list(M, x, lim(), true, pl);
        for (size_t m = 0; m < pl.a.size(); ++m) {
            const double d = pl.d[m];
            int k0, k1;
            window(d, k0, k1);
            for (int k = k0; k <= k1; ++k) g[k] += c[k] * K(d - (k + 0.5) * dr);
        }
    }
    // out_x += d/dx sum_k (g_k - g0_k)^2
    void grad(const Model& M, const std::vector<P3>& x, const std::vector<double>& g, const double* g0,
              std::vector<P3>& out_x) const {
        PairList pl;
        pair_list(M, x, lim(), true, pl);
        for (size_t m = 0; m < pl.a.size(); ++m) {
            const int i = pl.a[m], j = pl.b[m];
            const P3 dv = pl.dv[m];
            const double d = pl.d[m];
            if (d <= 0.0) continue;
            int k0, k1;
            window(d, k0, k1);
            double s = 0.0;
            for (int k = k0; k <= k1; ++k) {
                double u = d - (k + 0.5) * dr;
                s += 2.0 * (g[k] - g0[k]) * c[k] * (-u / (sb * sb)) * K(u);
            }
            P3 r{dv.x / d, dv.y / d, dv.z / d};
            out_x[i].x += s * r.x; out_x[i].y += s * r.y; out_x[i].z += s * r.z;
            out_x[j].x -= s * r.x; out_x[j].y -= s * r.y; out_x[j].z -= s * r.z;
        }
    }
};

inline double wrap1(double v, double L) {
    double h = 0.5 * L;
    double r = std::fmod(v, 2.0 * h);
    r = r < 0.0 ? r + 2.0 * h : r;
    return r - h;
}

}  // namespace

extern "C" {

// Per-pair derivative probe (for finite-difference tests): out[11] =
// phi, d1, d2, phi_s, phi_e, phi_q, phi_rc, f_s, f_e, f_q, f_rc.
void or_pair_derivs(double d, double sigma, double eps, double qq, double rc, int coul, double* out) {
    PairD o;
    pair_derivs(d, sigma, eps, qq, rc, coul != 0, o);
    double v[11] = {o.phi, o.d1, o.d2, o.phi_s, o.phi_e, o.phi_q, o.phi_rc, o.f_s, o.f_e, o.f_q, o.f_rc};
    std::memcpy(out, v, sizeof(v));
}

// Accelerations / V of the parameterised model (theta: 30 values).
double or_adj_accel(const gs_params* p, int64_t n, const double* x, const double* y, const double* z,
                    const uint8_t* species, const uint8_t* keep, const double* theta, double* ax, double* ay,
                    double* az) {
    Model M;
    M.n = (int)n;
    M.type.assign(species, species + n);
    M.keep.assign(keep, keep + n);
    M.setup(*p, theta);
    std::vector<P3> X(n), A;
    for (int64_t i = 0; i < n; ++i) X[i] = P3{x[i], y[i], z[i]};
    double V = accel(M, X, A);
    for (int64_t i = 0; i < n; ++i) {
        ax[i] = A[i].x; ay[i] = A[i].y; az[i] = A[i].z;
    }
    return V;
}

// One adjoint pass (HVP + parameter VJP) at x with velocity adjoint lam.
void or_adj_pass(const gs_params* p, int64_t n, const double* x, const double* y, const double* z,
                 const uint8_t* species, const uint8_t* keep, const double* theta, const double* lx,
                 const double* ly, const double* lz, double vcoef, double* ox, double* oy, double* oz,
                 double* gtheta) {
    Model M;
    M.n = (int)n;
    M.type.assign(species, species + n);
    M.keep.assign(keep, keep + n);
    M.setup(*p, theta);
    std::vector<P3> X(n), L(n), O;
    for (int64_t i = 0; i < n; ++i) {
        X[i] = P3{x[i], y[i], z[i]};
        L[i] = P3{lx[i], ly[i], lz[i]};
    }
    for (int k = 0; k < 30; ++k) gtheta[k] = 0.0;
    adjoint_pass(M, X, L, vcoef, O, gtheta);
    for (int64_t i = 0; i < n; ++i) {
        ox[i] = O[i].x; oy[i] = O[i].y; oz[i] = O[i].z;
    }
}

// Loss and (if grad != NULL) its gradient w.r.t. theta.  win_from: first
// tick (1-based, <= nsteps) of the mean-V window.  g0 may be NULL (no g(r)
// term); g_out (nbins) receives the final smoothed g(r).
int or_adjoint(const gs_params* p, int64_t n, const double* x0, const double* y0, const double* z0,
               const double* vx0, const double* vy0, const double* vz0, const uint8_t* species,
               const uint8_t* keep, const double* theta, int64_t nsteps, int64_t win_from, int gr_nbins,
               double gr_rmax, double gr_sigma, const double* g0, double* loss, double* grad, double* g_out) {
    Model M;
    M.n = (int)n;
    M.type.assign(species, species + n);
    M.keep.assign(keep, keep + n);
    M.setup(*p, theta);
    const double dt = M.dt, h = 0.5 * dt;
    const int T = (int)nsteps;
    std::vector<std::vector<P3>> X(T + 1);
    X[0].resize(n);
    std::vector<P3> v(n), a;
    for (int64_t i = 0; i < n; ++i) {
        X[0][i] = P3{x0[i], y0[i], z0[i]};
        v[i] = P3{vx0[i], vy0[i], vz0[i]};
    }
    const int W = T - (int)win_from + 1;
    double Vsum = 0.0;
    accel(M, X[0], a);
    for (int t = 0; t < T; ++t) {
        for (int i = 0; i < M.n; ++i) {
            v[i].x += h * a[i].x; v[i].y += h * a[i].y; v[i].z += h * a[i].z;
        }
        X[t + 1].resize(n);
        for (int i = 0; i < M.n; ++i) {
            X[t + 1][i].x = wrap1(X[t][i].x + dt * v[i].x, M.box[0]);
            X[t + 1][i].y = wrap1(X[t][i].y + dt * v[i].y, M.box[1]);
            X[t + 1][i].z = wrap1(X[t][i].z + dt * v[i].z, M.box[2]);
        }
        double V = accel(M, X[t + 1], a);
        if (t + 1 >= win_from) Vsum += V;
        for (int i = 0; i < M.n; ++i) {
            v[i].x += h * a[i].x; v[i].y += h * a[i].y; v[i].z += h * a[i].z;
        }
    }
    double L = W > 0 ? Vsum / W : 0.0;
    SmoothGr G;
    std::vector<double> g;
    const bool use_g = gr_nbins > 0;
    if (use_g) {
        G.setup(M, gr_nbins, gr_rmax, gr_sigma);
        G.eval(M, X[T], g);
        if (g_out) std::memcpy(g_out, g.data(), sizeof(double) * gr_nbins);
        if (g0)
            for (int k = 0; k < gr_nbins; ++k) L += (g[k] - g0[k]) * (g[k] - g0[k]);
    }
    *loss = L;
    if (!grad) return 0;
    // reverse pass
    std::vector<P3> lx(n, P3{0, 0, 0}), lv(n, P3{0, 0, 0}), tmp;
    double gth[30] = {0};
    if (use_g && g0) G.grad(M, X[T], g, g0, lx);
    for (int t = T; t >= 0; --t) {
        // kicks at x_t: the second kick of tick t-1 (t >= 1) and the first
        // kick of tick t (t < T), merged
        const double kf = (t == T || t == 0) ? h : 2.0 * h;
        const double vcoef = (t >= win_from && t >= 1 && W > 0) ? 1.0 / W : 0.0;
        std::vector<P3> lam(n);
        for (int i = 0; i < M.n; ++i) lam[i] = P3{kf * lv[i].x, kf * lv[i].y, kf * lv[i].z};
        double gt[30];
        for (int k = 0; k < 30; ++k) gt[k] = 0.0;
        // the pass scales its V-derivative by vcoef; lam carries kf
        adjoint_pass(M, X[t], lam, vcoef, tmp, gt);
        for (int k = 0; k < 30; ++k) gth[k] += gt[k];
        for (int i = 0; i < M.n; ++i) {
            lx[i].x += tmp[i].x; lx[i].y += tmp[i].y; lx[i].z += tmp[i].z;
        }
        if (t == 0) break;
        // drift of tick t-1: x_t = x_{t-1} + dt v_half
        for (int i = 0; i < M.n; ++i) {
            lv[i].x += dt * lx[i].x; lv[i].y += dt * lx[i].y; lv[i].z += dt * lx[i].z;
        }
    }
    std::memcpy(grad, gth, sizeof(gth));
    return 0;
}

}  // extern "C"
