// Oracle: FP64 CPU restatement of Alea's batched VMC/DMC walker sweep.
//
// TEST INFRASTRUCTURE ONLY.  It is the checker for the HIP path, never the
// thing measured or shipped: only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it.  The product library (alea_amd) never
// links, loads or calls it.
//
// It restates the reference's algorithms in the reference's own layouts
// (walker-fastest SoA, [.. * nw + w]) and in the reference's order of
// operations where that matters for bit-level agreement (the minimum-image
// fold, the RNG, the tape order).  Each function cites the reference file:line
// it follows.  How each part is pinned (tests/test_oracle.py):
//   cell, distances, Padé (+ image sum), Drummond, Nosanow, kinetic, Ewald,
//   Coulomb, VMC sweep, RNG  -> against oracle/_ref (the reference's own
//                               sources compiled here) and the reference's
//                               golden vectors (xoshiro state {1,2,3,4},
//                               Madelung 1.6155426267, exciton E = -1)
//   Slater real/complex     -> the reference's characterization record
//                               (tests/golden/build_characterization.txt:
//                               log|psi| and E_L of four walkers of each of
//                               its configs the path holds), closed forms
//                               (N <= 2 determinants, plane-wave eigenstates:
//                               E_L = sum |k|^2/2 exactly), brute-force
//                               permutation-sum determinants, finite
//                               differences; the reference's own Slater files
//                               need Eigen (absent), so they are not built.
//   initial walkers         -> the same record; libstdc++'s own engine.
//   Eigen stand-in          -> the initial inverse uses the reference's own
//                               pivoted Gauss-Jordan (slater_determinant.cpp:
//                               905-993) in place of Eigen's D.inverse().
//   inverse power           -> oracle/_ref InversePowerTerm (dipole system);
//                               the Wigner-Seitz tail against a disk-cut
//                               lattice sum plus its continuum remainder
//   DMC weights, branch     -> oracle/_ref WalkerSet::branch and
//                               PopulationControl, bit-exact populations
//   McMillan Jastrow        -> no counterpart in the reference: closed form
//                               u = -(b/r)^n, u', u'' and finite differences of
//                               log psi ("parity pinned by closed form")

#include "oracle.h"

#include <algorithm>
#include <array>
#include <bit>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <exception>
#include <limits>
#include <memory>
#include <numbers>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;
constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kDetFloor = 1e-300;
using cplx = std::complex<double>;

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return 1;
    } catch (const std::exception& e) {
        g_error = e.what();
        return 2;
    }
}

// ── RNG: src/numerics/random.h:39-56, src/numerics/xoshiro_random.h:23-193 ──
constexpr uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
constexpr uint64_t stream_seed(uint64_t base, uint64_t stream, uint64_t index) {
    return splitmix64(splitmix64(splitmix64(base) ^ stream) ^ index);
}

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        for (uint64_t i = 0; i < 4; ++i) s[i] = splitmix64(seed + i * 0x9e3779b97f4a7c15ULL);
    }
    Xoshiro(const uint64_t* st) { std::memcpy(s, st, sizeof s); }
    uint64_t next() {
        const uint64_t result = std::rotl(s[0] + s[3], 23) + s[0];
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = std::rotl(s[3], 45);
        return result;
    }
};

struct Ziggurat {
    std::array<uint64_t, 256> k{};
    std::array<double, 256> w{}, f{};
    double r = 0.0;
};

// Marsaglia-Tsang tables with 256 equal-area regions, built as in
// xoshiro_random.h:119-163 (bisection on the top-layer mismatch).
Ziggurat make_ziggurat() {
    constexpr int n = 256;
    auto f = [](double x) { return std::exp(-0.5 * x * x); };
    auto v_r = [&](double r) {
        return r * f(r) + std::sqrt(0.5 * std::numbers::pi) * std::erfc(r / std::numbers::sqrt2);
    };
    std::array<double, n> x{};
    auto top_mismatch = [&](double r) {
        const double v = v_r(r);
        x[n - 1] = r;
        for (int i = n - 1; i >= 2; --i) {
            const double y = f(x[i]) + v / x[i];
            if (y >= 1.0) return 1.0;
            x[i - 1] = std::sqrt(-2.0 * std::log(y));
        }
        return f(x[1]) + v / x[1] - 1.0;
    };
    double lo = 3.0, hi = 4.5;
    for (;;) {
        const double mid = 0.5 * (lo + hi);
        if (mid == lo || mid == hi) break;
        (top_mismatch(mid) > 0.0 ? lo : hi) = mid;
    }
    Ziggurat z;
    z.r = hi;
    (void)top_mismatch(z.r);
    const double two52 = 0x1.0p52;
    const double base = v_r(z.r) / f(z.r);
    z.k[0] = static_cast<uint64_t>(two52 * (z.r / base));
    z.w[0] = base / two52;
    z.f[0] = 1.0;
    for (int i = 1; i < n; ++i) {
        z.k[i] = i == 1 ? 0 : static_cast<uint64_t>(two52 * (x[i - 1] / x[i]));
        z.w[i] = x[i] / two52;
        z.f[i] = f(x[i]);
    }
    return z;
}
const Ziggurat& ziggurat() {
    static const Ziggurat z = make_ziggurat();
    return z;
}

struct Rng {
    Xoshiro g;
    explicit Rng(uint64_t seed) : g(seed) {}
    double uniform() { return static_cast<double>(static_cast<int64_t>(g.next() >> 11)) * 0x1.0p-53; }
    static double with_sign(double x, uint64_t sign) {
        return std::bit_cast<double>(std::bit_cast<uint64_t>(x) ^ sign);
    }
    double tail(double r) {
        for (;;) {
            const double xi = -std::log1p(-uniform()) / r;
            const double y = -std::log1p(-uniform());
            if (y + y > xi * xi) return r + xi;
        }
    }
    double gaussian() {
        const Ziggurat& z = ziggurat();
        constexpr uint64_t mask = (uint64_t{1} << 52) - 1;
        for (;;) {
            const uint64_t bits = g.next();
            const std::size_t i = bits & 0xff;
            const uint64_t sign = (bits & 0x100) << 55;
            const uint64_t position = (bits >> 9) & mask;
            const double x = static_cast<double>(static_cast<int64_t>(position)) * z.w[i];
            if (position < z.k[i]) return with_sign(x, sign);
            if (i == 0) return with_sign(tail(z.r), sign);
            if (z.f[i] + uniform() * (z.f[i - 1] - z.f[i]) < std::exp(-0.5 * x * x))
                return with_sign(x, sign);
        }
    }
};

// draw_sweep_randoms (src/qmc/sweep_blocking.h:39-48).
void draw_sweep(Rng& rng, std::size_t np, std::size_t dim, std::size_t stride,
                std::size_t w0, std::size_t w1, double* out) {
    for (std::size_t p = 0; p < np; ++p) {
        double* row = out + p * (dim + 1) * stride;
        for (std::size_t d = 0; d < dim; ++d)
            for (std::size_t w = w0; w < w1; ++w) row[d * stride + w] = rng.gaussian();
        for (std::size_t w = w0; w < w1; ++w) row[dim * stride + w] = rng.uniform();
    }
}

// ── Cell: src/population/simulation_cell.{h:109-261, cpp:16-317} ────────────
struct Cell {
    bool periodic = false, ortho = false;
    std::size_t dim = 2, ncand = 1;
    double a[4]{}, ainv[4]{}, b[4]{}, cand[9 * 2]{}, box[2]{}, inv_box[2]{};
    double volume = 0.0, ws = 0.0;

    void init(std::size_t d, const double* v) {
        dim = d;
        if (!v) return;
        if (d != 2) throw std::invalid_argument("oracle: periodic cells are 2-D");
        periodic = true;
        for (int k = 0; k < 4; ++k) a[k] = v[k];
        // axis_align (simulation_cell.cpp:126-144)
        {
            std::size_t axis[2];
            bool used[2] = {false, false}, ok = true;
            for (std::size_t i = 0; i < 2 && ok; ++i) {
                std::size_t nonzero = 0, which = 0;
                for (std::size_t e = 0; e < 2; ++e)
                    if (a[i * 2 + e] != 0.0) { ++nonzero; which = e; }
                if (nonzero != 1 || used[which]) ok = false;
                else { axis[i] = which; used[which] = true; }
            }
            if (ok) {
                double re[4] = {};
                for (std::size_t i = 0; i < 2; ++i)
                    for (std::size_t e = 0; e < 2; ++e) re[axis[i] * 2 + e] = a[i * 2 + e];
                std::memcpy(a, re, sizeof re);
            }
            ortho = ok;
        }
        if (!ortho) {  // gauss_reduce_2d (simulation_cell.cpp:22-36)
            for (int guard = 0; guard < 64; ++guard) {
                const double n0 = a[0] * a[0] + a[1] * a[1];
                const double n1 = a[2] * a[2] + a[3] * a[3];
                if (n1 < n0) { std::swap(a[0], a[2]); std::swap(a[1], a[3]); continue; }
                if (n0 == 0.0) break;
                const double mu = std::nearbyint((a[0] * a[2] + a[1] * a[3]) / n0);
                if (mu == 0.0) break;
                a[2] -= mu * a[0];
                a[3] -= mu * a[1];
            }
        }
        const double det = a[0] * a[3] - a[1] * a[2];
        volume = std::abs(det);
        if (volume == 0.0) throw std::invalid_argument("oracle: singular cell");
        const double inv_det = 1.0 / det;
        ainv[0] = a[3] * inv_det;
        ainv[1] = -a[2] * inv_det;
        ainv[2] = -a[1] * inv_det;
        ainv[3] = a[0] * inv_det;
        for (int k = 0; k < 4; ++k) b[k] = kTwoPi * ainv[k];
        if (ortho)
            for (std::size_t e = 0; e < 2; ++e) {
                box[e] = std::abs(a[e * 2 + e]);
                inv_box[e] = 1.0 / box[e];
            }
        ncand = 1;
        cand[0] = cand[1] = 0.0;
        if (!ortho)
            for (int n0 = -1; n0 <= 1; ++n0)
                for (int n1 = -1; n1 <= 1; ++n1) {
                    if (n0 == 0 && n1 == 0) continue;
                    for (std::size_t e = 0; e < 2; ++e) {
                        double s = 0.0;
                        s += n0 * a[0 * 2 + e];
                        s += n1 * a[1 * 2 + e];
                        cand[ncand * 2 + e] = s;
                    }
                    ++ncand;
                }
        double shortest2 = 0.0;
        for (int n0 = -2; n0 <= 2; ++n0)
            for (int n1 = -2; n1 <= 2; ++n1) {
                if (n0 == 0 && n1 == 0) continue;
                double r2 = 0.0;
                for (std::size_t e = 0; e < 2; ++e) {
                    double s = 0.0;
                    s += n0 * a[0 * 2 + e];
                    s += n1 * a[1 * 2 + e];
                    r2 += s * s;
                }
                if (shortest2 == 0.0 || r2 < shortest2) shortest2 = r2;
            }
        ws = 0.5 * std::sqrt(shortest2);
    }

    // fold_bravais_2d (simulation_cell.h:201-234)
    void fold2(double& x0, double& x1) const {
        double s0 = 0.0;
        s0 += ainv[0] * x0;
        s0 += ainv[1] * x1;
        double s1 = 0.0;
        s1 += ainv[2] * x0;
        s1 += ainv[3] * x1;
        const double f0 = std::nearbyint(s0), f1 = std::nearbyint(s1);
        double t0 = 0.0;
        t0 += f0 * a[0];
        t0 += f1 * a[2];
        double t1 = 0.0;
        t1 += f0 * a[1];
        t1 += f1 * a[3];
        x0 -= t0;
        x1 -= t1;
        double best0 = x0, best1 = x1, best2 = 0.0;
        best2 += x0 * x0;
        best2 += x1 * x1;
        for (std::size_t c = 1; c < 9; ++c) {
            const double y0 = x0 + cand[c * 2 + 0];
            double r2 = 0.0;
            r2 += y0 * y0;
            const double y1 = x1 + cand[c * 2 + 1];
            r2 += y1 * y1;
            const double take = (r2 < best2) ? 1.0 : 0.0;
            best2 += take * (r2 - best2);
            best0 += take * (y0 - best0);
            best1 += take * (y1 - best1);
        }
        x0 = best0;
        x1 = best1;
    }
    // min_image_batch per lane (simulation_cell.cpp:252-299)
    void min_image(double& d0, double& d1) const {
        if (!periodic) return;
        if (ortho) {
            d0 = d0 - box[0] * std::nearbyint(d0 * inv_box[0]);
            d1 = d1 - box[1] * std::nearbyint(d1 * inv_box[1]);
            return;
        }
        fold2(d0, d1);
    }
    // wrap (simulation_cell.h:141-159)
    void wrap(double& x0, double& x1) const {
        if (!periodic) return;
        if (ortho) {
            x0 -= box[0] * std::nearbyint(x0 * inv_box[0]);
            x1 -= box[1] * std::nearbyint(x1 * inv_box[1]);
            return;
        }
        double f[2];
        for (std::size_t i = 0; i < 2; ++i) {
            double s = 0.0;
            s += ainv[i * 2 + 0] * x0;
            s += ainv[i * 2 + 1] * x1;
            f[i] = std::nearbyint(s);
        }
        double s0 = 0.0, s1 = 0.0;
        s0 += f[0] * a[0];
        s0 += f[1] * a[2];
        s1 += f[0] * a[1];
        s1 += f[1] * a[3];
        x0 -= s0;
        x1 -= s1;
    }
};

struct Group {
    std::size_t first, last;
    double charge, inv_mass;
    bool contains(std::size_t i) const { return i >= first && i < last; }
    std::size_t size() const { return last - first; }
};

// LongRangeSpline::eval (src/numerics/long_range_spline.h:41-60,66-110)
struct Spline {
    std::vector<double> scaled;
    double r_cut = 0.0, inv_delta = 0.0;
    int n_knots = 0;
    void eval(double r, double& f, double& df, double& d2f) const {
        if (r >= r_cut) { f = df = d2f = 0.0; return; }
        const double t = r * inv_delta;
        int i = static_cast<int>(t);
        if (i > n_knots - 2) i = n_knots - 2;
        const double x = t - i, y = 1.0 - x;
        const double* a = scaled.data() + 3 * i;
        const double* b = a + 3;
        const double x2 = x * x, x3 = x2 * x, x4 = x3 * x, x5 = x4 * x;
        const double y2 = y * y, y3 = y2 * y, y4 = y3 * y, y5 = y4 * y;
        const double p0 = 1.0 - 10.0 * x3 + 15.0 * x4 - 6.0 * x5;
        const double p1 = x - 6.0 * x3 + 8.0 * x4 - 3.0 * x5;
        const double p2 = 0.5 * x2 - 1.5 * x3 + 1.5 * x4 - 0.5 * x5;
        const double q0 = 1.0 - 10.0 * y3 + 15.0 * y4 - 6.0 * y5;
        const double q1 = y - 6.0 * y3 + 8.0 * y4 - 3.0 * y5;
        const double q2 = 0.5 * y2 - 1.5 * y3 + 1.5 * y4 - 0.5 * y5;
        const double dp0 = -30.0 * x2 + 60.0 * x3 - 30.0 * x4;
        const double dp1 = 1.0 - 18.0 * x2 + 32.0 * x3 - 15.0 * x4;
        const double dp2 = x - 4.5 * x2 + 6.0 * x3 - 2.5 * x4;
        const double dq0 = -30.0 * y2 + 60.0 * y3 - 30.0 * y4;
        const double dq1 = 1.0 - 18.0 * y2 + 32.0 * y3 - 15.0 * y4;
        const double dq2 = y - 4.5 * y2 + 6.0 * y3 - 2.5 * y4;
        const double ddp0 = -60.0 * x + 180.0 * x2 - 120.0 * x3;
        const double ddp1 = -36.0 * x + 96.0 * x2 - 60.0 * x3;
        const double ddp2 = 1.0 - 9.0 * x + 18.0 * x2 - 10.0 * x3;
        const double ddq0 = -60.0 * y + 180.0 * y2 - 120.0 * y3;
        const double ddq1 = -36.0 * y + 96.0 * y2 - 60.0 * y3;
        const double ddq2 = 1.0 - 9.0 * y + 18.0 * y2 - 10.0 * y3;
        f = a[0] * p0 + a[1] * p1 + a[2] * p2 + b[0] * q0 - b[1] * q1 + b[2] * q2;
        df = inv_delta * (a[0] * dp0 + a[1] * dp1 + a[2] * dp2 - (b[0] * dq0 - b[1] * dq1 + b[2] * dq2));
        d2f = inv_delta * inv_delta *
              (a[0] * ddp0 + a[1] * ddp1 + a[2] * ddp2 + b[0] * ddq0 - b[1] * ddq1 + b[2] * ddq2);
    }
    double value(double r) const {
        double f, df, d2f;
        eval(r, f, df, d2f);
        return f;
    }
};

// Pair radial functions: Padé (pade_jastrow.cpp:26-56) with the image-sum
// short-range transform (periodic_pair_sum.cpp:66-83), or Drummond
// (drummond_jastrow.cpp:27-80).
struct Radial {
    bool drummond = false;
    bool mcmillan = false;   // u = -(b/r)^n (no counterpart in the reference), b = c[0], n = num
    int num = 0, den = 0;
    std::vector<double> c;
    bool image = false;
    Spline sp;
    double u_inf = 0.0;
    int C = 3;
    double L = 0.0, alpha1 = 0.0, gamma = 0.0;

    static void poly(const double* c, int n, double r, double& v, double& dv, double& d2v) {
        double r_km1 = 1.0, r_km2 = 0.0;
        for (int i = 0; i < n; ++i) {
            const int k = i + 1;
            v += c[i] * r_km1 * r;
            dv += c[i] * k * r_km1;
            d2v += c[i] * k * (k - 1) * r_km2;
            r_km2 = r_km1;
            r_km1 *= r;
        }
    }
    void eval(double r, double& u, double& du, double& d2u) const {
        if (mcmillan) {
            const double x = c[0] / r;
            double xn = 1.0;
            for (int i = 0; i < num; ++i) xn *= x;
            const double n = static_cast<double>(num);
            u = -xn;
            du = n * xn / r;
            d2u = -n * (n + 1.0) * xn / (r * r);
            return;
        }
        if (drummond) {
            auto ipow = [](double b, int e) { double x = 1.0; for (int i = 0; i < e; ++i) x *= b; return x; };
            const double t = r - L;
            const double tCm2 = ipow(t, C - 2), tCm1 = tCm2 * t, tC = tCm1 * t;
            const double g = tC, dg = static_cast<double>(C) * tCm1,
                         d2g = static_cast<double>(C) * (C - 1) * tCm2;
            const int n = static_cast<int>(c.size());
            double P = c[0] + alpha1 * r, dP = alpha1, d2P = 0.0;
            double r_lm2 = 1.0, r_lm1 = r, r_l = r * r;
            for (int k = 1; k < n; ++k) {
                const int l = k + 1;
                P += c[k] * r_l;
                dP += c[k] * l * r_lm1;
                d2P += c[k] * l * (l - 1) * r_lm2;
                r_lm2 = r_lm1;
                r_lm1 = r_l;
                r_l *= r;
            }
            const double m = (r < L) ? 1.0 : 0.0;
            u = m * (g * P);
            du = m * (dg * P + g * dP);
            d2u = m * (d2g * P + 2.0 * dg * dP + g * d2P);
            return;
        }
        double p = 0.0, dp = 0.0, d2p = 0.0;
        poly(c.data(), num, r, p, dp, d2p);
        double q = 1.0, dq = 0.0, d2q = 0.0;
        poly(c.data() + num, den, r, q, dq, d2q);
        const double inv_q = 1.0 / q;
        u = p * inv_q;
        du = (dp * q - p * dq) * inv_q * inv_q;
        d2u = (d2p - (2.0 * dp * dq - 2.0 * p * dq * dq * inv_q + p * d2q) * inv_q) * inv_q;
        if (image) {
            if (r >= sp.r_cut) { u = du = d2u = 0.0; return; }
            double f, df, d2f;
            sp.eval(r, f, df, d2f);
            u -= u_inf + f;
            du -= df;
            d2u -= d2f;
        }
    }
};

struct Sys;

// Base factor caches (wavefunction_component.h:266-287).
struct Component {
    virtual ~Component() = default;
    virtual void evaluate(Sys& s) = 0;
    virtual void ratio(Sys& s, std::size_t p) = 0;
    virtual void accept(Sys& s, std::size_t p, const std::vector<double>& mask) = 0;
    virtual void refresh_particle(Sys&, std::size_t) {}
    // rebuild_inverse of every block for the walkers [w0, w1)
    // (slater_determinant.cpp:905-993); factors without an inverse do nothing
    virtual void refresh_inverse(Sys&, std::size_t, std::size_t) {}
    virtual void permute(const std::vector<std::size_t>& src, std::size_t old_nw) = 0;
    virtual bool complex() const { return false; }
    std::vector<double> log_psi_c, grad_c, lap_c, pgrad_c, plap_c;
    void resize_base(std::size_t nw, std::size_t np, std::size_t dim) {
        log_psi_c.assign(nw, 0.0);
        grad_c.assign(np * dim * nw, 0.0);
        lap_c.assign(np * dim * nw, 0.0);
        if (complex()) {
            pgrad_c.assign(np * dim * nw, 0.0);
            plap_c.assign(np * dim * nw, 0.0);
        }
    }
    static void gather(std::vector<double>& v, const std::vector<std::size_t>& src,
                       std::size_t old_nw) {
        if (v.empty()) return;
        const std::size_t width = v.size() / old_nw;
        std::vector<double> out(width * src.size());
        for (std::size_t k = 0; k < width; ++k)
            for (std::size_t j = 0; j < src.size(); ++j)
                out[k * src.size() + j] = v[k * old_nw + src[j]];
        v.swap(out);
    }
    void permute_base(const std::vector<std::size_t>& src, std::size_t old_nw) {
        gather(log_psi_c, src, old_nw);
        gather(grad_c, src, old_nw);
        gather(lap_c, src, old_nw);
        gather(pgrad_c, src, old_nw);
        gather(plap_c, src, old_nw);
    }
};

struct Ewald {
    bool active = false;
    std::vector<int> n1, n2;
    std::vector<double> kx, ky, w;
    Spline sp;
    double r_cut = 0.0, v_long_zero = 0.0, sr_k0 = 0.0, inv_eps = 1.0;
    int n1_max = 0, n2_max = 0;
};

// InversePowerTerm (inverse_power.h:17-95): C / r^p over the pairs at the
// minimum image plus the continuum tail outside the Wigner-Seitz cell.
struct InversePower {
    int ga = 0, gb = 0, p_int = 0;
    double c = 0.0, p = 0.0, tail = 0.0;
};

struct CoulombPair {
    int ga, gb;
    double inv_eps;
};

struct Sys {
    std::size_t nw = 0, np = 0, dim = 2;
    Cell cell;
    std::vector<Group> groups;
    std::vector<double> pos, dist, pending, new_pos;
    std::vector<double> log_ratios, phase_diffs, new_grad, new_lap, new_pgrad, new_plap;
    std::vector<std::unique_ptr<Component>> comps;
    // the wavefunction's factors in the order they were added, as the
    // product counts them: a pair or one-body factor, or one determinant
    // block (component, block index)
    std::vector<std::pair<Component*, int>> factors;
    Ewald ewald;
    std::vector<CoulombPair> coulomb;
    std::vector<InversePower> ipow;
    double lattice[4]{};           // primitive vectors as given
    std::vector<double> inv_mass;  // [np]
    std::vector<double> e_loc, e_loc_im;
    // DMC population state (walker_set.h:135-173): weights, E_L before the
    // sweep, rr sums, tapered walker-sweeps
    std::vector<double> weights, e_prev, rr_p, rr_a;
    std::vector<long long> n_cut;
    bool has_phase = false;
    // Estimators (PairCorrelation, StructureFactor): the block's running sums
    struct Gofr {
        bool on = false;
        double rmax = 0.0, inv_delta = 0.0, wsum = 0.0;
        std::size_t n_bins = 0, n_func = 0;
        std::vector<double> hist;   // [pid * n_bins + bin]
    } gofr;
    struct Sofk {
        bool on = false;
        std::vector<int> n1, n2, shell;
        std::vector<double> kx, ky;
        int m1 = 0, m2 = 0;
        double wsum = 0.0;
        std::vector<double> acc_s, acc_re, acc_im;   // [pid][k], [g][k]
    } sofk;
    struct Dens {
        bool on = false;
        double rmax = 0.0, inv_delta = 0.0, wsum = 0.0;
        std::size_t n_bins = 0, n_cells = 0;
        bool per_particle = false;
        std::vector<double> hist;   // [g * n_cells + cell], g a group or a particle
    } dens;

    double& P(std::size_t p, std::size_t d, std::size_t w) { return pos[(p * dim + d) * nw + w]; }
    double& D(std::size_t i, std::size_t j, std::size_t w) { return dist[(i * np + j) * nw + w]; }

    void alloc() {
        pos.assign(np * dim * nw, 0.0);
        dist.assign(np * np * nw, 0.0);
        pending.assign(np * nw, 0.0);
        new_pos.assign(dim * nw, 0.0);
        log_ratios.assign(nw, 0.0);
        phase_diffs.assign(nw, 0.0);
        new_grad.assign(dim * nw, 0.0);
        new_lap.assign(dim * nw, 0.0);
        new_pgrad.assign(dim * nw, 0.0);
        new_plap.assign(dim * nw, 0.0);
        e_loc.assign(nw, 0.0);
        e_loc_im.assign(nw, 0.0);
        weights.assign(nw, 1.0);
        e_prev.assign(nw, 0.0);
        rr_p.assign(nw, 0.0);
        rr_a.assign(nw, 0.0);
        n_cut.assign(nw, 0);
    }
    // BatchedDistanceTable::build_all (batched_distance_table.cpp:20-58)
    void build_dist() {
        for (std::size_t i = 0; i < np; ++i)
            for (std::size_t j = 0; j < np; ++j)
                for (std::size_t w = 0; w < nw; ++w) {
                    double d0 = P(i, 0, w) - P(j, 0, w), d1 = P(i, 1, w) - P(j, 1, w);
                    cell.min_image(d0, d1);
                    double r2 = 0.0;
                    r2 += d0 * d0;
                    r2 += d1 * d1;
                    D(i, j, w) = std::sqrt(r2);
                }
    }
    // update_batch (batched_distance_table.cpp:63-101)
    void propose(std::size_t p) {
        for (std::size_t k = 0; k < np; ++k)
            for (std::size_t w = 0; w < nw; ++w) {
                if (k == p) { pending[k * nw + w] = 0.0; continue; }
                double d0 = new_pos[w] - P(k, 0, w), d1 = new_pos[nw + w] - P(k, 1, w);
                cell.min_image(d0, d1);
                double r2 = 0.0;
                r2 += d0 * d0;
                r2 += d1 * d1;
                pending[k * nw + w] = std::sqrt(r2);
            }
    }
    // ParticleBatch::commit_move + WalkerSet::apply_move position part
    // (particle_batch.cpp:65-85, batched_distance_table.cpp:115-143)
    void commit(std::size_t p, const std::vector<double>& m) {
        for (std::size_t k = 0; k < np; ++k) {
            if (k == p) continue;
            for (std::size_t w = 0; w < nw; ++w) {
                double& pk = D(p, k, w);
                double& kp = D(k, p, w);
                const double nr = pending[k * nw + w];
                pk += m[w] * (nr - pk);
                kp += m[w] * (nr - kp);
            }
        }
        for (std::size_t w = 0; w < nw; ++w) {
            if (m[w] == 0.0) continue;
            double x0 = new_pos[w], x1 = new_pos[nw + w];
            cell.wrap(x0, x1);
            P(p, 0, w) = x0;
            P(p, 1, w) = x1;
        }
    }
};

// Displacement r_a - r_b folded by minimum image, for walker w.
inline void fold(const Sys& s, double xa0, double xa1, double xb0, double xb1, double& d0,
                 double& d1) {
    d0 = xa0 - xb0;
    d1 = xa1 - xb1;
    s.cell.min_image(d0, d1);
}

// ── KSpacePairSum (src/wavefunction/kspace_pair_sum.cpp:91-301) with
//    ReciprocalPhases (src/numerics/reciprocal_phases.h:26-75) ───────────────
struct KSpace {
    Group a, b;
    bool same = true;
    std::vector<int> n1, n2;
    std::vector<double> kx, ky, kx2, ky2, w;
    int m1 = 0, m2 = 0;
    double b1x = 0, b1y = 0, b2x = 0, b2y = 0;
    std::size_t nw = 0, n_cached = 0;
    std::vector<double> ra_re, ra_im, rb_re, rb_im, pc, ps, pend_c, pend_s;

    std::size_t nk() const { return kx.size(); }
    std::size_t slot(std::size_t i) const { return a.contains(i) ? i - a.first : a.size() + (i - b.first); }
    double* crow(std::size_t w, std::size_t i) { return pc.data() + (w * n_cached + slot(i)) * nk(); }
    double* srow(std::size_t w, std::size_t i) { return ps.data() + (w * n_cached + slot(i)) * nk(); }

    static void ladder(double theta, int m, std::vector<double>& re, std::vector<double>& im) {
        re.assign(2 * m + 1, 0.0);
        im.assign(2 * m + 1, 0.0);
        double* r = re.data() + m;
        double* i = im.data() + m;
        const double c = std::cos(theta), s = std::sin(theta);
        r[0] = 1.0;
        i[0] = 0.0;
        for (int n = 1; n <= m; ++n) {
            r[n] = r[n - 1] * c - i[n - 1] * s;
            i[n] = r[n - 1] * s + i[n - 1] * c;
            r[-n] = r[n];
            i[-n] = -i[n];
        }
    }
    void phases(double x, double y, double* c, double* s) const {
        std::vector<double> e1r, e1i, e2r, e2i;
        ladder(b1x * x + b1y * y, m1, e1r, e1i);
        ladder(b2x * x + b2y * y, m2, e2r, e2i);
        const double* f1r = e1r.data() + m1;
        const double* f1i = e1i.data() + m1;
        const double* f2r = e2r.data() + m2;
        const double* f2i = e2i.data() + m2;
        for (std::size_t k = 0; k < nk(); ++k) {
            const double ar = f1r[n1[k]], ai = f1i[n1[k]];
            const double br = f2r[n2[k]], bi = f2i[n2[k]];
            c[k] = ar * br - ai * bi;
            s[k] = ar * bi + ai * br;
        }
    }
    void prepare(std::size_t nw_) {
        nw = nw_;
        const std::size_t n = nw * nk();
        ra_re.assign(n, 0.0);
        ra_im.assign(n, 0.0);
        if (!same) { rb_re.assign(n, 0.0); rb_im.assign(n, 0.0); }
        pc.assign(n * n_cached, 0.0);
        ps.assign(n * n_cached, 0.0);
        pend_c.assign(n, 0.0);
        pend_s.assign(n, 0.0);
    }
    struct KT { double gx, gy, lx, ly; };
    KT kterms(const double* c, const double* s, const double* pr, const double* pi, double self) const {
        double gx = 0, gy = 0, lx = 0, ly = 0;
        for (std::size_t k = 0; k < nk(); ++k) {
            const double g = 2.0 * w[k] * (c[k] * pi[k] - s[k] * pr[k]);
            const double l = 2.0 * w[k] * (self - (c[k] * pr[k] + s[k] * pi[k]));
            gx += kx[k] * g;
            gy += ky[k] * g;
            lx += kx2[k] * l;
            ly += ky2[k] * l;
        }
        return {gx, gy, lx, ly};
    }
    double klog(const double* ar, const double* ai, const double* br, const double* bi) const {
        double t = 0.0;
        if (same)
            for (std::size_t k = 0; k < nk(); ++k) t += w[k] * (ar[k] * ar[k] + ai[k] * ai[k]);
        else
            for (std::size_t k = 0; k < nk(); ++k) t += 2.0 * w[k] * (ar[k] * br[k] + ai[k] * bi[k]);
        return t;
    }
    void evaluate(Sys& s, std::vector<double>& lp, std::vector<double>& grad, std::vector<double>& lap) {
        prepare(s.nw);
        const std::size_t K = nk();
        for (std::size_t wk = 0; wk < s.nw; ++wk) {
            auto build = [&](const Group& g, double* re, double* im) {
                std::fill_n(re, K, 0.0);
                std::fill_n(im, K, 0.0);
                for (std::size_t i = g.first; i < g.last; ++i) {
                    double* c = crow(wk, i);
                    double* sn = srow(wk, i);
                    phases(s.P(i, 0, wk), s.P(i, 1, wk), c, sn);
                    for (std::size_t k = 0; k < K; ++k) { re[k] += c[k]; im[k] += sn[k]; }
                }
            };
            build(a, ra_re.data() + wk * K, ra_im.data() + wk * K);
            if (!same) build(b, rb_re.data() + wk * K, rb_im.data() + wk * K);
            const double* ar = ra_re.data() + wk * K;
            const double* ai = ra_im.data() + wk * K;
            const double* br = same ? ar : rb_re.data() + wk * K;
            const double* bi = same ? ai : rb_im.data() + wk * K;
            lp[wk] += klog(ar, ai, br, bi);
            auto add = [&](std::size_t i, const double* pr, const double* pi, double self) {
                const KT t = kterms(crow(wk, i), srow(wk, i), pr, pi, self);
                grad[(i * 2 + 0) * s.nw + wk] += t.gx;
                grad[(i * 2 + 1) * s.nw + wk] += t.gy;
                lap[(i * 2 + 0) * s.nw + wk] += t.lx;
                lap[(i * 2 + 1) * s.nw + wk] += t.ly;
            };
            for (std::size_t i = a.first; i < a.last; ++i) add(i, br, bi, same ? 1.0 : 0.0);
            if (!same)
                for (std::size_t j = b.first; j < b.last; ++j) add(j, ar, ai, 0.0);
        }
    }
    void ratio(Sys& s, std::size_t p, std::vector<double>& lr, std::vector<double>& ng,
               std::vector<double>& nl) {
        const bool in_a = a.contains(p), in_b = !same && b.contains(p);
        if (!in_a && !in_b) return;
        const std::size_t K = nk(), nw_ = s.nw;
        std::vector<double> qr(K), qi(K);
        for (std::size_t wk = 0; wk < nw_; ++wk) {
            double* cn = pend_c.data() + wk * K;
            double* sn = pend_s.data() + wk * K;
            phases(s.new_pos[wk], s.new_pos[nw_ + wk], cn, sn);
            const double* co = crow(wk, p);
            const double* so = srow(wk, p);
            const double* mr = (in_a ? ra_re : rb_re).data() + wk * K;
            const double* mi = (in_a ? ra_im : rb_im).data() + wk * K;
            double dlog = 0.0;
            KT t{};
            if (same) {
                for (std::size_t k = 0; k < K; ++k) {
                    qr[k] = mr[k] + cn[k] - co[k];
                    qi[k] = mi[k] + sn[k] - so[k];
                    dlog += w[k] * (qr[k] * qr[k] + qi[k] * qi[k] - mr[k] * mr[k] - mi[k] * mi[k]);
                }
                t = kterms(cn, sn, qr.data(), qi.data(), 1.0);
            } else {
                const double* orr = (in_a ? rb_re : ra_re).data() + wk * K;
                const double* oi = (in_a ? rb_im : ra_im).data() + wk * K;
                for (std::size_t k = 0; k < K; ++k)
                    dlog += 2.0 * w[k] * ((cn[k] - co[k]) * orr[k] + (sn[k] - so[k]) * oi[k]);
                t = kterms(cn, sn, orr, oi, 0.0);
            }
            lr[wk] += dlog;
            ng[0 * nw_ + wk] += t.gx;
            ng[1 * nw_ + wk] += t.gy;
            nl[0 * nw_ + wk] += t.lx;
            nl[1 * nw_ + wk] += t.ly;
        }
    }
    void accept(Sys& s, std::size_t p, const std::vector<double>& mask, std::vector<double>& lp,
                std::vector<double>& grad, std::vector<double>& lap) {
        const bool in_a = a.contains(p), in_b = !same && b.contains(p);
        if (!in_a && !in_b) return;
        const std::size_t K = nk(), nw_ = s.nw;
        std::vector<double> dr(K), di(K);
        const Group& seers = same ? a : (in_a ? b : a);
        for (std::size_t wk = 0; wk < nw_; ++wk) {
            if (mask[wk] == 0.0) continue;
            const double* cn = pend_c.data() + wk * K;
            const double* sn = pend_s.data() + wk * K;
            double* co = crow(wk, p);
            double* so = srow(wk, p);
            double* mr = (in_a ? ra_re : rb_re).data() + wk * K;
            double* mi = (in_a ? ra_im : rb_im).data() + wk * K;
            double dlog = 0.0;
            if (same) {
                for (std::size_t k = 0; k < K; ++k) {
                    dr[k] = cn[k] - co[k];
                    di[k] = sn[k] - so[k];
                    dlog += w[k] * (2.0 * (mr[k] * dr[k] + mi[k] * di[k]) + dr[k] * dr[k] + di[k] * di[k]);
                }
            } else {
                const double* orr = (in_a ? rb_re : ra_re).data() + wk * K;
                const double* oi = (in_a ? rb_im : ra_im).data() + wk * K;
                for (std::size_t k = 0; k < K; ++k) {
                    dr[k] = cn[k] - co[k];
                    di[k] = sn[k] - so[k];
                    dlog += 2.0 * w[k] * (dr[k] * orr[k] + di[k] * oi[k]);
                }
            }
            lp[wk] += dlog;
            for (std::size_t i = seers.first; i < seers.last; ++i) {
                if (i == p) continue;
                const KT t = kterms(crow(wk, i), srow(wk, i), dr.data(), di.data(), 0.0);
                grad[(i * 2 + 0) * nw_ + wk] += t.gx;
                grad[(i * 2 + 1) * nw_ + wk] += t.gy;
                lap[(i * 2 + 0) * nw_ + wk] += t.lx;
                lap[(i * 2 + 1) * nw_ + wk] += t.ly;
            }
            for (std::size_t k = 0; k < K; ++k) {
                mr[k] += dr[k];
                mi[k] += di[k];
                co[k] = cn[k];
                so[k] = sn[k];
            }
        }
    }
    void permute(const std::vector<std::size_t>& src) {
        const std::size_t K = nk();
        auto carry = [&](std::vector<double>& v, std::size_t row) {
            if (v.empty()) return;
            std::vector<double> out(src.size() * row);
            for (std::size_t j = 0; j < src.size(); ++j)
                std::copy_n(v.data() + src[j] * row, row, out.data() + j * row);
            v.swap(out);
        };
        carry(ra_re, K); carry(ra_im, K); carry(rb_re, K); carry(rb_im, K);
        carry(pc, n_cached * K); carry(ps, n_cached * K);
        pend_c.assign(src.size() * K, 0.0);
        pend_s.assign(src.size() * K, 0.0);
        nw = src.size();
    }
};

// ── Pair Jastrow: PadeJastrow (pade_jastrow.cpp:250-634) and
//    RadialPairJastrow<Drummond> (radial_pair_jastrow_impl.h:47-350) ────────
struct PairJastrow : Component {
    Group ga, gb;
    bool same;
    Radial rad;
    bool has_k = false;
    KSpace ks;
    std::vector<double> pg, pl, prop;  // pending p grad/lap [dim*nw]; [k*6*nw + f*nw + w]

    template <class F>
    void partners(std::size_t p, F&& f) const {
        if (same) {
            if (ga.contains(p))
                for (std::size_t k = ga.first; k < ga.last; ++k)
                    if (k != p) f(k);
        } else if (ga.contains(p)) {
            for (std::size_t k = gb.first; k < gb.last; ++k) f(k);
        } else if (gb.contains(p)) {
            for (std::size_t k = ga.first; k < ga.last; ++k) f(k);
        }
    }
    void evaluate(Sys& s) override {
        const std::size_t nw = s.nw;
        resize_base(nw, s.np, s.dim);
        pg.assign(2 * nw, 0.0);
        pl.assign(2 * nw, 0.0);
        prop.assign(6 * s.np * nw, 0.0);
        auto pair = [&](std::size_t i, std::size_t j) {
            for (std::size_t w = 0; w < nw; ++w) {
                const double r = s.D(i, j, w);
                double u, du, d2u;
                rad.eval(r, u, du, d2u);
                const double invr = 1.0 / r;
                log_psi_c[w] += u;
                double del[2];
                fold(s, s.P(i, 0, w), s.P(i, 1, w), s.P(j, 0, w), s.P(j, 1, w), del[0], del[1]);
                for (std::size_t d = 0; d < 2; ++d) {
                    const double dor = del[d] * invr;
                    const double g = du * dor;
                    grad_c[(i * 2 + d) * nw + w] += g;
                    grad_c[(j * 2 + d) * nw + w] -= g;
                    const double ld = d2u * dor * dor + du * invr * (1.0 - dor * dor);
                    lap_c[(i * 2 + d) * nw + w] += ld;
                    lap_c[(j * 2 + d) * nw + w] += ld;
                }
            }
        };
        if (same) {
            for (std::size_t i = ga.first; i < ga.last; ++i)
                for (std::size_t j = i + 1; j < ga.last; ++j) pair(i, j);
        } else {
            for (std::size_t i = ga.first; i < ga.last; ++i)
                for (std::size_t j = gb.first; j < gb.last; ++j) pair(i, j);
        }
        if (has_k) ks.evaluate(s, log_psi_c, grad_c, lap_c);
    }
    void ratio(Sys& s, std::size_t p) override {
        const std::size_t nw = s.nw;
        std::fill(pg.begin(), pg.end(), 0.0);
        std::fill(pl.begin(), pl.end(), 0.0);
        partners(p, [&](std::size_t k) {
            for (std::size_t w = 0; w < nw; ++w) {
                double uo, duo, d2uo, un, dun, d2un;
                rad.eval(s.D(p, k, w), uo, duo, d2uo);
                const double rnew = s.pending[k * nw + w];
                rad.eval(rnew, un, dun, d2un);
                const double invr = 1.0 / rnew;
                s.log_ratios[w] += un - uo;
                double* pr = prop.data() + 6 * k * nw;
                pr[0 * nw + w] = uo; pr[1 * nw + w] = duo; pr[2 * nw + w] = d2uo;
                pr[3 * nw + w] = un; pr[4 * nw + w] = dun; pr[5 * nw + w] = d2un;
                double del[2];
                fold(s, s.new_pos[w], s.new_pos[nw + w], s.P(k, 0, w), s.P(k, 1, w), del[0], del[1]);
                for (std::size_t d = 0; d < 2; ++d) {
                    const double dor = del[d] * invr;
                    pg[d * nw + w] += dun * dor;
                    pl[d * nw + w] += d2un * dor * dor + dun * invr * (1.0 - dor * dor);
                }
            }
        });
        if (has_k) ks.ratio(s, p, s.log_ratios, pg, pl);
        for (std::size_t i = 0; i < 2 * nw; ++i) {
            s.new_grad[i] += pg[i];
            s.new_lap[i] += pl[i];
        }
    }
    void accept(Sys& s, std::size_t p, const std::vector<double>& m) override {
        const std::size_t nw = s.nw;
        bool acts = false;
        partners(p, [&](std::size_t k) {
            acts = true;
            for (std::size_t w = 0; w < nw; ++w) {
                const double* pr = prop.data() + 6 * k * nw;
                const double uo = pr[0 * nw + w], duo = pr[1 * nw + w], d2uo = pr[2 * nw + w];
                const double un = pr[3 * nw + w], dun = pr[4 * nw + w], d2un = pr[5 * nw + w];
                const double iro = 1.0 / s.D(p, k, w), irn = 1.0 / s.pending[k * nw + w];
                log_psi_c[w] += m[w] * (un - uo);
                double dlo[2], dln[2];
                fold(s, s.P(k, 0, w), s.P(k, 1, w), s.P(p, 0, w), s.P(p, 1, w), dlo[0], dlo[1]);
                fold(s, s.P(k, 0, w), s.P(k, 1, w), s.new_pos[w], s.new_pos[nw + w], dln[0], dln[1]);
                for (std::size_t d = 0; d < 2; ++d) {
                    const double dor_old = dlo[d] * iro, dor_new = dln[d] * irn;
                    grad_c[(k * 2 + d) * nw + w] += m[w] * (dun * dor_new - duo * dor_old);
                    const double lo = d2uo * dor_old * dor_old + duo * iro * (1.0 - dor_old * dor_old);
                    const double ln = d2un * dor_new * dor_new + dun * irn * (1.0 - dor_new * dor_new);
                    lap_c[(k * 2 + d) * nw + w] += m[w] * (ln - lo);
                }
            }
        });
        if (!acts) return;
        if (has_k) ks.accept(s, p, m, log_psi_c, grad_c, lap_c);
        for (std::size_t d = 0; d < 2; ++d)
            for (std::size_t w = 0; w < nw; ++w) {
                double& g = grad_c[(p * 2 + d) * nw + w];
                double& l = lap_c[(p * 2 + d) * nw + w];
                g += m[w] * (pg[d * nw + w] - g);
                l += m[w] * (pl[d * nw + w] - l);
            }
    }
    void permute(const std::vector<std::size_t>& src, std::size_t old_nw) override {
        permute_base(src, old_nw);
        pg.assign(2 * src.size(), 0.0);
        pl.assign(2 * src.size(), 0.0);
        prop.assign(6 * (grad_c.size() / (2 * src.size())) * src.size(), 0.0);
        if (has_k) ks.permute(src);
    }
};

// ── Nosanow one-body (src/wavefunction/nosanow_jastrow.cpp:58-174) ─────────
struct Nosanow : Component {
    Group g;
    double alpha = 0.0;
    std::vector<double> sites, value_c, pv, pg, pl;
    void terms(const Sys& s, double x0, double x1, std::size_t site, double& u, double* gr,
               double* lp) const {
        double d0 = x0 - sites[site * 2], d1 = x1 - sites[site * 2 + 1];
        if (s.cell.periodic) s.cell.min_image(d0, d1);
        u = 0.0;
        u -= alpha * d0 * d0;
        u -= alpha * d1 * d1;
        gr[0] = -2.0 * alpha * d0;
        gr[1] = -2.0 * alpha * d1;
        lp[0] = lp[1] = -2.0 * alpha;
    }
    void evaluate(Sys& s) override {
        const std::size_t nw = s.nw;
        resize_base(nw, s.np, 2);
        value_c.assign(s.np * nw, 0.0);
        pv.assign(nw, 0.0);
        pg.assign(2 * nw, 0.0);
        pl.assign(2 * nw, 0.0);
        for (std::size_t p = g.first; p < g.last; ++p)
            for (std::size_t w = 0; w < nw; ++w) {
                double u, gr[2], lp[2];
                terms(s, s.P(p, 0, w), s.P(p, 1, w), p - g.first, u, gr, lp);
                value_c[p * nw + w] = u;
                log_psi_c[w] += u;
                for (int d = 0; d < 2; ++d) {
                    grad_c[(p * 2 + d) * nw + w] = gr[d];
                    lap_c[(p * 2 + d) * nw + w] = lp[d];
                }
            }
    }
    void ratio(Sys& s, std::size_t p) override {
        if (!g.contains(p)) return;
        const std::size_t nw = s.nw;
        for (std::size_t w = 0; w < nw; ++w) {
            double u, gr[2], lp[2];
            terms(s, s.new_pos[w], s.new_pos[nw + w], p - g.first, u, gr, lp);
            pv[w] = u;
            s.log_ratios[w] += u - value_c[p * nw + w];
            for (int d = 0; d < 2; ++d) {
                pg[d * nw + w] = gr[d];
                pl[d * nw + w] = lp[d];
                s.new_grad[d * nw + w] += gr[d];
                s.new_lap[d * nw + w] += lp[d];
            }
        }
    }
    void accept(Sys& s, std::size_t p, const std::vector<double>& m) override {
        if (!g.contains(p)) return;
        const std::size_t nw = s.nw;
        for (std::size_t w = 0; w < nw; ++w) {
            const double dlog = m[w] * (pv[w] - value_c[p * nw + w]);
            log_psi_c[w] += dlog;
            if (m[w] != 0.0) {
                value_c[p * nw + w] = pv[w];
                for (int d = 0; d < 2; ++d) {
                    grad_c[(p * 2 + d) * nw + w] = pg[d * nw + w];
                    lap_c[(p * 2 + d) * nw + w] = pl[d * nw + w];
                }
            }
        }
    }
    void permute(const std::vector<std::size_t>& src, std::size_t old_nw) override {
        permute_base(src, old_nw);
        gather(value_c, src, old_nw);
        pv.assign(src.size(), 0.0);
        pg.assign(2 * src.size(), 0.0);
        pl.assign(2 * src.size(), 0.0);
    }
};

// ── Orbital sets ────────────────────────────────────────────────────────────
// Real: PlaneWaveOrbitalSet (orbital_set.h:427-474), GaussianOrbitalSet
// (orbital_set.h:164-224).  Values at one position, val[m], grad[m*2+d],
// lap[m*2+d].
struct RealOrbitals {
    bool gaussian = false;
    std::vector<double> k;          // [n*2]
    std::vector<unsigned char> is_sin;
    std::vector<double> centers, alphas;
    std::size_t n = 0;
    void eval(const Cell& cell, double x0, double x1, double* val, double* grad, double* lap) const {
        for (std::size_t i = 0; i < n; ++i) {
            if (gaussian) {
                double d0 = x0 - centers[i * 2], d1 = x1 - centers[i * 2 + 1];
                if (cell.periodic) cell.min_image(d0, d1);
                double r2 = 0.0;
                r2 += d0 * d0;
                r2 += d1 * d1;
                const double a = alphas[i];
                const double phi = std::exp(-a * r2);
                val[i] = phi;
                const double del[2] = {d0, d1};
                for (int d = 0; d < 2; ++d) {
                    grad[i * 2 + d] = phi * (-2.0 * a * del[d]);
                    lap[i * 2 + d] = phi * (4.0 * a * a * del[d] * del[d] - 2.0 * a);
                }
            } else {
                double phase = 0.0;
                phase += k[i * 2] * x0;
                phase += k[i * 2 + 1] * x1;
                const double s = std::sin(phase), c = std::cos(phase);
                val[i] = is_sin[i] ? s : c;
                for (int d = 0; d < 2; ++d) {
                    const double kd = k[i * 2 + d];
                    grad[i * 2 + d] = is_sin[i] ? kd * c : -kd * s;
                    lap[i * 2 + d] = is_sin[i] ? -kd * kd * s : -kd * kd * c;
                }
            }
        }
    }
};

// Pivoted Gauss-Jordan inverse of a row-major n x n matrix (the reference's
// rebuild_inverse, slater_determinant.cpp:905-993; also the stand-in for
// Eigen's D.inverse() at evaluate time).  Returns log|det|; throws if singular.
template <class T>
double gauss_jordan(std::size_t N, const T* d, T* a) {
    std::vector<T> m(d, d + N * N);
    for (std::size_t r = 0; r < N; ++r)
        for (std::size_t c = 0; c < N; ++c) a[r * N + c] = r == c ? T(1.0) : T(0.0);
    double log_det = 0.0;
    for (std::size_t k = 0; k < N; ++k) {
        std::size_t piv = k;
        for (std::size_t r = k + 1; r < N; ++r)
            if (std::abs(m[r * N + k]) > std::abs(m[piv * N + k])) piv = r;
        const T diag = m[piv * N + k];
        if (std::abs(diag) == 0.0 || !std::isfinite(std::abs(diag)))
            throw std::runtime_error("Slater determinant: accepted matrix is singular or nonfinite");
        log_det += std::log(std::abs(diag));
        if (piv != k) {
            std::swap_ranges(m.begin() + k * N + k, m.begin() + k * N + N, m.begin() + piv * N + k);
            std::swap_ranges(a + k * N, a + k * N + N, a + piv * N);
        }
        for (std::size_t c = k + 1; c < N; ++c) m[k * N + c] /= diag;
        for (std::size_t c = 0; c < N; ++c) a[k * N + c] /= diag;
        for (std::size_t r = 0; r < N; ++r) {
            if (r == k) continue;
            const T f = m[r * N + k];
            for (std::size_t c = k + 1; c < N; ++c) m[r * N + c] -= f * m[k * N + c];
            for (std::size_t c = 0; c < N; ++c) a[r * N + c] -= f * a[k * N + c];
        }
    }
    return log_det;
}

// ── SlaterDeterminant, real (slater_determinant.cpp:305-437, 744-900) ──────
// Matrices are contiguous per walker (the reference's layout for N >= 32 or
// nw >= 64, slater_determinant.cpp:26-44).
struct RealSlater : Component {
    struct Block {
        std::size_t first, last, N;
        RealOrbitals orb;
        std::vector<double> A, D, dD, d2D;  // [w*N*N + ...], dD [w*N*N*2 + (kr*N+m)*2+d]
        std::vector<double> logd, cond, row_abs;
        std::vector<int> updates;
        std::vector<double> nval, ngrad, nlap;  // proposal [w*N + m], [w*N*2 + m*2 + d]
    };
    std::vector<Block> blocks;
    std::size_t block_of(std::size_t p) const {
        for (std::size_t b = 0; b < blocks.size(); ++b)
            if (p >= blocks[b].first && p < blocks[b].last) return b;
        return SIZE_MAX;
    }
    void particle_derivs(const Sys& s, Block& B, std::size_t kr, std::size_t w) {
        const std::size_t N = B.N, nw = s.nw, particle = B.first + kr;
        const double* A = B.A.data() + w * N * N;
        for (std::size_t d = 0; d < 2; ++d) {
            double G = 0.0, Lr = 0.0;
            for (std::size_t m = 0; m < N; ++m) {
                const double a = A[m * N + kr];
                G += a * B.dD[w * N * N * 2 + (kr * N + m) * 2 + d];
                Lr += a * B.d2D[w * N * N * 2 + (kr * N + m) * 2 + d];
            }
            grad_c[(particle * 2 + d) * nw + w] = G;
            lap_c[(particle * 2 + d) * nw + w] = Lr - G * G;
        }
    }
    void evaluate(Sys& s) override {
        const std::size_t nw = s.nw;
        resize_base(nw, s.np, 2);
        for (auto& B : blocks) {
            const std::size_t N = B.N;
            B.A.assign(nw * N * N, 0.0);
            B.D.assign(nw * N * N, 0.0);
            B.dD.assign(nw * N * N * 2, 0.0);
            B.d2D.assign(nw * N * N * 2, 0.0);
            B.logd.assign(nw, 0.0);
            B.cond.assign(nw, 1.0);
            B.row_abs.assign(N * nw, 0.0);
            B.updates.assign(nw, 0);
            B.nval.assign(nw * N, 0.0);
            B.ngrad.assign(nw * N * 2, 0.0);
            B.nlap.assign(nw * N * 2, 0.0);
            for (std::size_t w = 0; w < nw; ++w) {
                for (std::size_t kr = 0; kr < N; ++kr) {
                    const std::size_t i = B.first + kr;
                    B.orb.eval(s.cell, s.P(i, 0, w), s.P(i, 1, w), B.D.data() + w * N * N + kr * N,
                               B.dD.data() + w * N * N * 2 + kr * N * 2,
                               B.d2D.data() + w * N * N * 2 + kr * N * 2);
                }
                const double* Dw = B.D.data() + w * N * N;
                double* Aw = B.A.data() + w * N * N;
                for (std::size_t kr = 0; kr < N; ++kr) {
                    double rs = 0.0;
                    for (std::size_t i = 0; i < N; ++i) rs += std::abs(Dw[kr * N + i]);
                    B.row_abs[kr * nw + w] = rs;
                }
                B.logd[w] = std::max(gauss_jordan<double>(N, Dw, Aw), std::log(kDetFloor));
                log_psi_c[w] += B.logd[w];
                double nd = 0.0, na = 0.0;
                for (std::size_t r = 0; r < N; ++r) {
                    double sd = 0.0, sa = 0.0;
                    for (std::size_t c = 0; c < N; ++c) { sd += std::abs(Dw[r * N + c]); sa += std::abs(Aw[r * N + c]); }
                    nd = std::max(nd, sd);
                    na = std::max(na, sa);
                }
                B.cond[w] = nd * na;
                for (std::size_t kr = 0; kr < N; ++kr) particle_derivs(s, B, kr, w);
            }
        }
    }
    void ratio(Sys& s, std::size_t p) override {
        const std::size_t b = block_of(p);
        if (b == SIZE_MAX) return;
        Block& B = blocks[b];
        const std::size_t N = B.N, j = p - B.first, nw = s.nw;
        for (std::size_t w = 0; w < nw; ++w) {
            double* v = B.nval.data() + w * N;
            double* g = B.ngrad.data() + w * N * 2;
            double* l = B.nlap.data() + w * N * 2;
            B.orb.eval(s.cell, s.new_pos[w], s.new_pos[nw + w], v, g, l);
            const double* A = B.A.data() + w * N * N;
            double R = 0.0;
            for (std::size_t m = 0; m < N; ++m) R += A[m * N + j] * v[m];
            s.log_ratios[w] += std::log(std::max(std::abs(R), kDetFloor));
            if (R < 0.0) s.phase_diffs[w] += kPi;
            const double invR = (R != 0.0) ? 1.0 / R : 0.0;
            for (std::size_t d = 0; d < 2; ++d) {
                double G = 0.0, Lr = 0.0;
                for (std::size_t m = 0; m < N; ++m) {
                    G += A[m * N + j] * g[m * 2 + d];
                    Lr += A[m * N + j] * l[m * 2 + d];
                }
                const double Gp = invR * G;
                s.new_grad[d * nw + w] += Gp;
                s.new_lap[d * nw + w] += invR * Lr - Gp * Gp;
            }
        }
    }
    double rebuild(Block& B, std::size_t w) {
        const std::size_t N = B.N;
        const double ld = gauss_jordan<double>(N, B.D.data() + w * N * N, B.A.data() + w * N * N);
        B.updates[w] = 0;
        return ld;
    }
    // commit_walker (slater_determinant.cpp:744-858), deferred partners.
    void accept(Sys& s, std::size_t p, const std::vector<double>& m) override {
        const std::size_t b = block_of(p);
        if (b == SIZE_MAX) return;
        Block& B = blocks[b];
        const std::size_t N = B.N, j = p - B.first, nw = s.nw;
        std::vector<double> column(N), rowp(N);
        for (std::size_t w = 0; w < nw; ++w) {
            if (m[w] == 0.0) continue;
            double* A = B.A.data() + w * N * N;
            double* Dv = B.D.data() + w * N * N;
            double* new_row = Dv + j * N;
            for (std::size_t mm = 0; mm < N; ++mm) new_row[mm] = B.nval[w * N + mm];
            for (std::size_t mm = 0; mm < N; ++mm)
                for (std::size_t d = 0; d < 2; ++d) {
                    B.dD[w * N * N * 2 + (j * N + mm) * 2 + d] = B.ngrad[w * N * 2 + mm * 2 + d];
                    B.d2D[w * N * N * 2 + (j * N + mm) * 2 + d] = B.nlap[w * N * 2 + mm * 2 + d];
                }
            double ratio = 0.0, scale = 0.0;
            std::fill(rowp.begin(), rowp.end(), 0.0);
            for (std::size_t mm = 0; mm < N; ++mm) {
                const double* ir = A + mm * N;
                const double v = new_row[mm];
                const double term = ir[j] * v;
                ratio += term;
                scale += std::abs(term);
                column[mm] = ir[j];
                for (std::size_t c = 0; c < N; ++c) rowp[c] += v * ir[c];
            }
            rowp[j] -= 1.0;
            const double old_log = B.logd[w];
            const bool rb = ++B.updates[w] >= 64 || !std::isfinite(ratio) || std::abs(ratio) < 1e-5 ||
                            B.cond[w] > 1e8 ||
                            std::abs(ratio) <= 1e4 * std::numeric_limits<double>::epsilon() * scale;
            double norm_a = 0.0;
            if (rb) {
                B.logd[w] = rebuild(B, w);
                for (std::size_t a = 0; a < N; ++a) {
                    double rs = 0.0;
                    for (std::size_t c = 0; c < N; ++c) rs += std::abs(A[a * N + c]);
                    norm_a = std::max(norm_a, rs);
                }
            } else {
                B.logd[w] += std::log(std::abs(ratio));
                for (std::size_t a = 0; a < N; ++a) {
                    const double f = column[a] / ratio;
                    double rs = 0.0;
                    for (std::size_t c = 0; c < N; ++c) {
                        A[a * N + c] = A[a * N + c] - f * rowp[c];
                        rs += std::abs(A[a * N + c]);
                    }
                    norm_a = std::max(norm_a, rs);
                }
            }
            double nrs = 0.0;
            for (std::size_t c = 0; c < N; ++c) nrs += std::abs(new_row[c]);
            B.row_abs[j * nw + w] = nrs;
            double norm_d = 0.0;
            for (std::size_t a = 0; a < N; ++a) norm_d = std::max(norm_d, B.row_abs[a * nw + w]);
            B.cond[w] = norm_d * norm_a;
            if (!rb && (!std::isfinite(B.cond[w]) || B.cond[w] > 1e8)) B.logd[w] = rebuild(B, w);
            log_psi_c[w] += B.logd[w] - old_log;
            particle_derivs(s, B, j, w);
        }
    }
    // rebuild_inverse (slater_determinant.cpp:905-993) of every block of the
    // walkers [w0, w1): log|det| floored as evaluate_batch floors it, the
    // condition estimate and the particles' derivatives from the new inverse
    void refresh_inverse(Sys& s, std::size_t w0, std::size_t w1) override {
        for (std::size_t w = w0; w < w1; ++w) {
            double total = 0.0;
            for (auto& B : blocks) {
                const std::size_t N = B.N;
                B.logd[w] = std::max(rebuild(B, w), std::log(kDetFloor));
                total += B.logd[w];
                const double* A = B.A.data() + w * N * N;
                double norm_a = 0.0, norm_d = 0.0;
                for (std::size_t a = 0; a < N; ++a) {
                    double rs = 0.0;
                    for (std::size_t c = 0; c < N; ++c) rs += std::abs(A[a * N + c]);
                    norm_a = std::max(norm_a, rs);
                    norm_d = std::max(norm_d, B.row_abs[a * s.nw + w]);
                }
                B.cond[w] = norm_d * norm_a;
                for (std::size_t kr = 0; kr < N; ++kr) particle_derivs(s, B, kr, w);
            }
            log_psi_c[w] = total;
        }
    }
    void refresh_particle(Sys& s, std::size_t p) override {
        const std::size_t b = block_of(p);
        if (b == SIZE_MAX) return;
        for (std::size_t w = 0; w < s.nw; ++w) particle_derivs(s, blocks[b], p - blocks[b].first, w);
    }
    void permute(const std::vector<std::size_t>& src, std::size_t old_nw) override {
        permute_base(src, old_nw);
        for (auto& B : blocks) {
            auto carry = [&](std::vector<double>& v) {
                const std::size_t row = v.size() / old_nw;
                std::vector<double> out(src.size() * row);
                for (std::size_t j = 0; j < src.size(); ++j)
                    std::copy_n(v.data() + src[j] * row, row, out.data() + j * row);
                v.swap(out);
            };
            carry(B.A); carry(B.D); carry(B.dD); carry(B.d2D);
            gather(B.logd, src, old_nw);
            gather(B.cond, src, old_nw);
            gather(B.row_abs, src, old_nw);
            std::vector<int> up(src.size());
            for (std::size_t j = 0; j < src.size(); ++j) up[j] = B.updates[src[j]];
            B.updates.swap(up);
            B.nval.assign(src.size() * B.N, 0.0);
            B.ngrad.assign(src.size() * B.N * 2, 0.0);
            B.nlap.assign(src.size() * B.N * 2, 0.0);
        }
    }
};

// ── ComplexSlaterDeterminant with twisted plane waves
//    (complex_slater_determinant.cpp:85-283, complex_orbital_set.h:91-124) ──
struct ComplexSlater : Component {
    struct Block {
        std::size_t first, last, N;
        std::vector<double> q;  // [N*2]
        std::vector<cplx> A;    // [w*N*N + a*N + c]
        std::vector<cplx> nval, ngrad, nlap;  // [w*N + m], [w*N*2 + m*2 + d]
    };
    std::vector<Block> blocks;
    bool complex() const override { return true; }
    std::size_t block_of(std::size_t p) const {
        for (std::size_t b = 0; b < blocks.size(); ++b)
            if (p >= blocks[b].first && p < blocks[b].last) return b;
        return SIZE_MAX;
    }
    static void orb(const Block& B, double x0, double x1, cplx* val, cplx* grad, cplx* lap) {
        for (std::size_t i = 0; i < B.N; ++i) {
            const double* qi = B.q.data() + i * 2;
            double arg = 0.0;
            arg += qi[0] * x0;
            arg += qi[1] * x1;
            const double c = std::cos(arg), s = std::sin(arg);
            val[i] = cplx(c, s);
            for (int d = 0; d < 2; ++d) {
                grad[i * 2 + d] = cplx(-qi[d] * s, qi[d] * c);
                lap[i * 2 + d] = cplx(-qi[d] * qi[d] * c, -qi[d] * qi[d] * s);
            }
        }
    }
    // Rows at a configuration, p at (x0,x1) if moved < N; inverse; derivatives.
    void full(Sys& s, Block& B, std::size_t w, std::size_t moved, bool set_log) {
        const std::size_t N = B.N, nw = s.nw;
        std::vector<cplx> D(N * N), G(N * N * 2), L(N * N * 2);
        for (std::size_t kr = 0; kr < N; ++kr) {
            const std::size_t i = B.first + kr;
            const double x0 = (i == moved) ? s.new_pos[w] : s.P(i, 0, w);
            const double x1 = (i == moved) ? s.new_pos[nw + w] : s.P(i, 1, w);
            orb(B, x0, x1, D.data() + kr * N, G.data() + kr * N * 2, L.data() + kr * N * 2);
        }
        cplx* A = B.A.data() + w * N * N;
        const double ld = gauss_jordan<cplx>(N, D.data(), A);
        if (set_log) log_psi_c[w] += std::max(ld, std::log(kDetFloor));
        for (std::size_t kr = 0; kr < N; ++kr) {
            const std::size_t i = B.first + kr;
            for (std::size_t d = 0; d < 2; ++d) {
                cplx g(0.0, 0.0), lr(0.0, 0.0);
                for (std::size_t m = 0; m < N; ++m) {
                    g += A[m * N + kr] * G[(kr * N + m) * 2 + d];
                    lr += A[m * N + kr] * L[(kr * N + m) * 2 + d];
                }
                const cplx l = lr - g * g;
                const std::size_t idx = (i * 2 + d) * nw + w;
                grad_c[idx] = g.real();
                lap_c[idx] = l.real();
                pgrad_c[idx] = g.imag();
                plap_c[idx] = l.imag();
            }
        }
    }
    // a fresh inverse of every block of the walkers [w0, w1) (the rebuild of
    // complex_slater_determinant.cpp:245-253), log|det| floored as
    // evaluate_batch floors it
    void refresh_inverse(Sys& s, std::size_t w0, std::size_t w1) override {
        for (std::size_t w = w0; w < w1; ++w) {
            log_psi_c[w] = 0.0;
            for (auto& B : blocks) full(s, B, w, SIZE_MAX, true);
        }
    }
    void evaluate(Sys& s) override {
        const std::size_t nw = s.nw;
        resize_base(nw, s.np, 2);
        for (auto& B : blocks) {
            B.A.assign(nw * B.N * B.N, cplx(0.0, 0.0));
            B.nval.assign(nw * B.N, cplx(0.0, 0.0));
            B.ngrad.assign(nw * B.N * 2, cplx(0.0, 0.0));
            B.nlap.assign(nw * B.N * 2, cplx(0.0, 0.0));
            for (std::size_t w = 0; w < nw; ++w) full(s, B, w, SIZE_MAX, true);
        }
    }
    void ratio(Sys& s, std::size_t p) override {
        const std::size_t b = block_of(p);
        if (b == SIZE_MAX) return;
        Block& B = blocks[b];
        const std::size_t N = B.N, j = p - B.first, nw = s.nw;
        for (std::size_t w = 0; w < nw; ++w) {
            cplx* v = B.nval.data() + w * N;
            cplx* g = B.ngrad.data() + w * N * 2;
            cplx* l = B.nlap.data() + w * N * 2;
            orb(B, s.new_pos[w], s.new_pos[nw + w], v, g, l);
            const cplx* A = B.A.data() + w * N * N;
            cplx R(0.0, 0.0);
            for (std::size_t m = 0; m < N; ++m) R += A[m * N + j] * v[m];
            const double absR = std::abs(R);
            s.log_ratios[w] += std::log(std::max(absR, kDetFloor));
            s.phase_diffs[w] += std::arg(R);
            const cplx invR = absR > 0.0 ? 1.0 / R : cplx(0.0, 0.0);
            for (std::size_t d = 0; d < 2; ++d) {
                cplx G(0.0, 0.0), Lr(0.0, 0.0);
                for (std::size_t m = 0; m < N; ++m) {
                    G += A[m * N + j] * g[m * 2 + d];
                    Lr += A[m * N + j] * l[m * 2 + d];
                }
                const cplx Gp = invR * G;
                const cplx Lp = invR * Lr - Gp * Gp;
                s.new_grad[d * nw + w] += Gp.real();
                s.new_lap[d * nw + w] += Lp.real();
                s.new_pgrad[d * nw + w] += Gp.imag();
                s.new_plap[d * nw + w] += Lp.imag();
            }
        }
    }
    void accept(Sys& s, std::size_t p, const std::vector<double>& m) override {
        const std::size_t b = block_of(p);
        if (b == SIZE_MAX) return;
        Block& B = blocks[b];
        const std::size_t N = B.N, j = p - B.first, nw = s.nw;
        for (std::size_t w = 0; w < nw; ++w) {
            if (m[w] == 0.0) continue;
            const cplx* A = B.A.data() + w * N * N;
            cplx R(0.0, 0.0);
            for (std::size_t mm = 0; mm < N; ++mm) R += A[mm * N + j] * B.nval[w * N + mm];
            log_psi_c[w] += std::log(std::max(std::abs(R), kDetFloor));
            full(s, B, w, p, false);
        }
    }
    void permute(const std::vector<std::size_t>& src, std::size_t old_nw) override {
        permute_base(src, old_nw);
        for (auto& B : blocks) {
            const std::size_t row = B.N * B.N;
            std::vector<cplx> out(src.size() * row);
            for (std::size_t j = 0; j < src.size(); ++j)
                std::copy_n(B.A.data() + src[j] * row, row, out.data() + j * row);
            B.A.swap(out);
            B.nval.assign(src.size() * B.N, cplx(0.0, 0.0));
            B.ngrad.assign(src.size() * B.N * 2, cplx(0.0, 0.0));
            B.nlap.assign(src.size() * B.N * 2, cplx(0.0, 0.0));
        }
    }
};

// ── Hamiltonian ────────────────────────────────────────────────────────────
// EwaldTerm::evaluate_potential (ewald_term.cpp:368-612), single Coulomb
// kernel: real-space V - V_long strictly inside r_cut, reciprocal sum from
// exponential ladders, self + background constants (ewald_term.cpp:186-262).
void ewald_energy(Sys& s, std::vector<double>& le) {
    const Ewald& e = s.ewald;
    const std::size_t nw = s.nw, np = s.np;
    std::vector<double> q(np);
    double q2 = 0.0, Q = 0.0;
    for (const auto& g : s.groups)
        for (std::size_t i = g.first; i < g.last; ++i) {
            q[i] = g.charge;
            q2 += g.charge * g.charge;
            Q += g.charge;
        }
    for (std::size_t i = 0; i < np; ++i)
        for (std::size_t j = i + 1; j < np; ++j) {
            const double coeff = q[i] * q[j];
            if (coeff == 0.0) continue;
            for (std::size_t w = 0; w < nw; ++w) {
                const double r = s.D(i, j, w);
                const double v = e.inv_eps / r;
                le[w] += r < e.r_cut ? coeff * (v - e.sp.value(r)) : 0.0;
            }
        }
    const double b1x = s.cell.b[0], b1y = s.cell.b[1], b2x = s.cell.b[2], b2y = s.cell.b[3];
    for (std::size_t w = 0; w < nw; ++w) {
        const int M1 = e.n1_max, M2 = e.n2_max;
        std::vector<double> e1r((M1 + 1) * np), e1i((M1 + 1) * np), e2r((M2 + 1) * np), e2i((M2 + 1) * np);
        auto build = [&](double bx, double by, int nmax, std::vector<double>& er, std::vector<double>& ei) {
            for (std::size_t i = 0; i < np; ++i) {
                er[i] = 1.0;
                ei[i] = 0.0;
                if (nmax < 1) continue;
                const double ph = bx * s.P(i, 0, w) + by * s.P(i, 1, w);
                er[np + i] = std::cos(ph);
                ei[np + i] = std::sin(ph);
                for (int n = 2; n <= nmax; ++n) {
                    const double pr = er[(n - 1) * np + i], pi = ei[(n - 1) * np + i];
                    er[n * np + i] = pr * er[np + i] - pi * ei[np + i];
                    ei[n * np + i] = pr * ei[np + i] + pi * er[np + i];
                }
            }
        };
        build(b1x, b1y, M1, e1r, e1i);
        build(b2x, b2y, M2, e2r, e2i);
        for (std::size_t kk = 0; kk < e.w.size(); ++kk) {
            const int n1 = e.n1[kk], n2 = e.n2[kk];
            const std::size_t a2 = static_cast<std::size_t>(std::abs(n2));
            const double sgn = n2 < 0 ? -1.0 : 1.0;
            double rr = 0.0, ri = 0.0;
            for (std::size_t i = 0; i < np; ++i) {
                if (q[i] == 0.0) continue;
                const double ar = e1r[n1 * np + i], ai = e1i[n1 * np + i];
                const double br = e2r[a2 * np + i], bim = sgn * e2i[a2 * np + i];
                rr += q[i] * (ar * br - ai * bim);
                ri += q[i] * (ar * bim + ai * br);
            }
            le[w] += e.w[kk] * (rr * rr + ri * ri);
        }
    }
    const double constants = -0.5 * e.v_long_zero * q2 - 0.5 * Q * Q * e.sr_k0;
    for (std::size_t w = 0; w < nw; ++w) le[w] += constants;
}

// CoulombTerm::evaluate_potential (coulomb.cpp:31-61), in-plane (d = 0).
void coulomb_energy(Sys& s, const CoulombPair& c, std::vector<double>& le) {
    const Group& a = s.groups[c.ga];
    const Group& b = s.groups[c.gb];
    const bool intra = c.ga == c.gb;
    const double coeff = a.charge * b.charge * c.inv_eps;
    for (std::size_t i = a.first; i < a.last; ++i)
        for (std::size_t j = intra ? i + 1 : b.first; j < (intra ? a.last : b.last); ++j)
            for (std::size_t w = 0; w < s.nw; ++w) {
                const double r = s.D(i, j, w);
                le[w] += coeff / std::sqrt(r * r + 0.0);
            }
}

// InversePowerTerm::evaluate_potential (inverse_power.cpp:130-168).
void inverse_power_energy(Sys& s, const InversePower& t, std::vector<double>& le) {
    const Group& a = s.groups[t.ga];
    const Group& b = s.groups[t.gb];
    const bool intra = t.ga == t.gb;
    for (std::size_t i = a.first; i < a.last; ++i)
        for (std::size_t j = intra ? i + 1 : b.first; j < (intra ? a.last : b.last); ++j)
            for (std::size_t w = 0; w < s.nw; ++w) {
                const double r = s.D(i, j, w);
                if (t.p_int > 0) {
                    const double inv_r = 1.0 / r;
                    double inv_rp = inv_r;
                    for (int k = 1; k < t.p_int; ++k) inv_rp *= inv_r;
                    le[w] += t.c * inv_rp;
                } else {
                    le[w] += t.c * std::pow(r, -t.p);
                }
            }
    if (t.tail != 0.0)
        for (std::size_t w = 0; w < s.nw; ++w) le[w] += t.tail;
}

// outside_wigner_seitz_integral in 2-D (inverse_power.cpp:28-41,
// wigner_seitz_exterior.h:58-83): (1/(p - 2)) times the midpoint rule on
// 2^18 angles of rho^(2 - p), rho from the bisectors of the 24 translations
// s a1 + t a2 with |s|, |t| <= 2.
double ws_exterior_integral(const double* a, double p) {
    std::vector<double> lx, ly;
    for (int si = -2; si <= 2; ++si)
        for (int ti = -2; ti <= 2; ++ti) {
            if (si == 0 && ti == 0) continue;
            lx.push_back(si * a[0] + ti * a[2]);
            ly.push_back(si * a[1] + ti * a[3]);
        }
    constexpr std::size_t n_theta = std::size_t{1} << 18;
    const double dtheta = 2.0 * kPi / static_cast<double>(n_theta);
    double sum = 0.0;
    for (std::size_t k = 0; k < n_theta; ++k) {
        const double theta = (static_cast<double>(k) + 0.5) * dtheta;
        const double ux = std::cos(theta), uy = std::sin(theta);
        double rho = std::numeric_limits<double>::infinity();
        for (std::size_t c = 0; c < lx.size(); ++c) {
            const double proj = ux * lx[c] + uy * ly[c];
            if (proj > 0.0) rho = std::min(rho, 0.5 * (lx[c] * lx[c] + ly[c] * ly[c]) / proj);
        }
        sum += std::pow(rho, 2.0 - p);
    }
    return sum * dtheta / (p - 2.0);
}

void totals(Sys& s, std::vector<double>& g, std::vector<double>& l, std::vector<double>* pg,
            std::vector<double>* pl) {
    const std::size_t n = s.np * 2 * s.nw;
    g.assign(n, 0.0);
    l.assign(n, 0.0);
    if (pg) pg->assign(n, 0.0);
    if (pl) pl->assign(n, 0.0);
    for (auto& c : s.comps) {
        for (std::size_t i = 0; i < n; ++i) { g[i] += c->grad_c[i]; l[i] += c->lap_c[i]; }
        if (c->complex() && pg && pl)
            for (std::size_t i = 0; i < n; ++i) { (*pg)[i] += c->pgrad_c[i]; (*pl)[i] += c->plap_c[i]; }
    }
}

void refresh_all(Sys& s) {
    for (auto& c : s.comps)
        for (std::size_t p = 0; p < s.np; ++p) c->refresh_particle(s, p);
}

// Kinetic::evaluate_batch + add_phase_terms (kinetic.cpp:45-110), then the
// potential terms: Hamiltonian::evaluate_batch (hamiltonian.cpp:18-29).
void local_energy(Sys& s) {
    refresh_all(s);
    std::vector<double> g, l, pg, pl;
    totals(s, g, l, &pg, &pl);
    const std::size_t nw = s.nw;
    std::fill(s.e_loc.begin(), s.e_loc.end(), 0.0);
    std::fill(s.e_loc_im.begin(), s.e_loc_im.end(), 0.0);
    for (std::size_t i = 0; i < s.np; ++i)
        for (std::size_t d = 0; d < 2; ++d) {
            const double im = s.inv_mass[i];
            const std::size_t off = (i * 2 + d) * nw;
            for (std::size_t w = 0; w < nw; ++w) s.e_loc[w] += -0.5 * im * (l[off + w] + g[off + w] * g[off + w]);
        }
    if (s.has_phase)
        for (std::size_t i = 0; i < s.np; ++i)
            for (std::size_t d = 0; d < 2; ++d) {
                const double im = s.inv_mass[i];
                const std::size_t off = (i * 2 + d) * nw;
                for (std::size_t w = 0; w < nw; ++w) {
                    s.e_loc[w] += 0.5 * im * pg[off + w] * pg[off + w];
                    s.e_loc_im[w] += -0.5 * im * (pl[off + w] + 2.0 * g[off + w] * pg[off + w]);
                }
            }
    if (s.ewald.active) ewald_energy(s, s.e_loc);
    for (const auto& c : s.coulomb) coulomb_energy(s, c, s.e_loc);
    for (const auto& t : s.ipow) inverse_power_energy(s, t, s.e_loc);
}

// TrialWavefunction::ratio_batch (trial_wavefunction.cpp:46-78).
void ratio_all(Sys& s, std::size_t p) {
    s.propose(p);
    std::fill(s.log_ratios.begin(), s.log_ratios.end(), 0.0);
    std::fill(s.phase_diffs.begin(), s.phase_diffs.end(), 0.0);
    std::fill(s.new_grad.begin(), s.new_grad.end(), 0.0);
    std::fill(s.new_lap.begin(), s.new_lap.end(), 0.0);
    std::fill(s.new_pgrad.begin(), s.new_pgrad.end(), 0.0);
    std::fill(s.new_plap.begin(), s.new_plap.end(), 0.0);
    for (auto& c : s.comps) c->ratio(s, p);
}

// Deferred commit, then WalkerSet::apply_move (vmc.cpp:147-152).
void accept_all(Sys& s, std::size_t p, const std::vector<double>& m) {
    for (auto& c : s.comps) c->accept(s, p, m);
    s.commit(p, m);
}

double log_psi(const Sys& s, std::size_t w) {
    double t = 0.0;
    for (const auto& c : s.comps) t += c->log_psi_c[w];
    return t;
}

Sys* S(oracle_system* h) { return reinterpret_cast<Sys*>(h); }

void check_index(const Sys& s, int g) {
    if (g < 0 || static_cast<std::size_t>(g) >= s.groups.size())
        throw std::invalid_argument("oracle: group index out of range");
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

uint64_t oracle_stream_seed(uint64_t base, uint64_t stream, uint64_t index) {
    return stream_seed(base, stream, index);
}

int oracle_xoshiro_state_u64(const uint64_t* state, int n, uint64_t* out) {
    Xoshiro g(state);
    for (int i = 0; i < n; ++i) out[i] = g.next();
    return 0;
}

int oracle_xoshiro_draws(uint64_t seed, int kind, int n, double* out) {
    Rng r(seed);
    for (int i = 0; i < n; ++i) out[i] = kind == 0 ? r.uniform() : r.gaussian();
    return 0;
}

// The per-walker device streams: walker w draws from its own stream
// stream_seed(seed, stream, w) (random.h:52-56), per sweep and per particle
// two Gaussians then a uniform (draw_sweep_randoms, sweep_blocking.h:39-48,
// over a one-walker range); out [sweep][np][3][nw].
int oracle_device_tape(uint64_t seed, uint64_t stream, int np, int nw, int n_sweeps, double* out) {
    for (int w = 0; w < nw; ++w) {
        Rng r(stream_seed(seed, stream, static_cast<uint64_t>(w)));
        for (int sw = 0; sw < n_sweeps; ++sw)
            for (int p = 0; p < np; ++p) {
                double* row = out + (static_cast<std::size_t>(sw) * np + p) * 3 * nw;
                row[w] = r.gaussian();
                row[nw + w] = r.gaussian();
                row[2 * nw + w] = r.uniform();
            }
    }
    return 0;
}

int oracle_sweep_tape(uint64_t seed, int np, int dim, int nw, int n_sweeps, double* out) {
    Rng r(seed);
    const std::size_t len = static_cast<std::size_t>(np) * (dim + 1) * nw;
    for (int s = 0; s < n_sweeps; ++s) draw_sweep(r, np, dim, nw, 0, nw, out + s * len);
    return 0;
}

// The reference's initial walkers (src/qmc/app.cpp:56-116, numerics::Random
// of src/numerics/random.h:11-27): one std::mt19937_64 seeded with `seed`,
// walker by walker and particle by particle; a periodic cell draws fractional
// coordinates f_i = uniform(-1/2, 1/2) and maps them through the cell's
// vectors() (pos_d = sum_i f_i a[i*2+d]): the basis after axis alignment and
// Gauss reduction, which for a hexagonal cell (on the reduction's tie) depends
// on rounding.  An open cell draws each coordinate in
// [-L/2, L/2); particles flagged in at_origin (may be null) sit at the origin
// and draw nothing.  The same libstdc++ engine and distribution as the
// reference's build.  Output AoS [w][p][d].  Crystal anchors and pairing
// partners (which also draw normals) are not restated.
int oracle_initial_positions(uint64_t seed, int np, int nw, const double* lattice, double L,
                             const unsigned char* at_origin, double* out) {
    std::mt19937_64 engine(seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    auto uniform = [&](double lo, double hi) { return lo + (hi - lo) * uni(engine); };
    Cell cell;
    if (lattice) {
        cell.init(2, lattice);
        lattice = cell.a;
    }
    for (int w = 0; w < nw; ++w)
        for (int p = 0; p < np; ++p) {
            double* x = out + (static_cast<std::size_t>(w) * np + p) * 2;
            if (at_origin && at_origin[p]) {
                x[0] = x[1] = 0.0;
            } else if (lattice) {
                double f[2];
                for (int i = 0; i < 2; ++i) f[i] = uniform(-0.5, 0.5);
                for (int d = 0; d < 2; ++d) {
                    double s = 0.0;
                    for (int i = 0; i < 2; ++i) s += f[i] * lattice[i * 2 + d];
                    x[d] = s;
                }
            } else {
                for (int d = 0; d < 2; ++d) x[d] = uniform(-0.5 * L, 0.5 * L);
            }
        }
    return 0;
}

// The 64-bit Mersenne Twister itself, n outputs from `seed`.
int oracle_mt19937_64(uint64_t seed, int n, uint64_t* out) {
    std::mt19937_64 engine(seed);
    for (int i = 0; i < n; ++i) out[i] = engine();
    return 0;
}

int oracle_ziggurat_tables(uint64_t* k, double* w, double* f, double* r) {
    const Ziggurat& z = ziggurat();
    for (int i = 0; i < 256; ++i) { k[i] = z.k[i]; w[i] = z.w[i]; f[i] = z.f[i]; }
    *r = z.r;
    return 0;
}

int oracle_create(oracle_system** out, int nw, int np, int dim, const double* lattice,
                  int precision) {
    return guarded([&] {
        (void)precision;
        if (dim != 2) throw std::invalid_argument("oracle: dim must be 2");
        if (nw <= 0 || np <= 0) throw std::invalid_argument("oracle: empty batch");
        auto* s = new Sys;
        s->nw = static_cast<std::size_t>(nw);
        s->np = static_cast<std::size_t>(np);
        s->dim = 2;
        s->cell.init(2, lattice);
        if (lattice) std::copy(lattice, lattice + 4, s->lattice);
        s->alloc();
        s->inv_mass.assign(s->np, 1.0);
        *out = reinterpret_cast<oracle_system*>(s);
    });
}

int oracle_destroy(oracle_system* h) {
    delete S(h);
    return 0;
}

int oracle_add_group(oracle_system* h, int first, int last, double charge, double inv_mass) {
    return guarded([&] {
        Sys& s = *S(h);
        if (first < 0 || last < first || static_cast<std::size_t>(last) > s.np)
            throw std::invalid_argument("oracle: group range out of bounds");
        s.groups.push_back({static_cast<std::size_t>(first), static_cast<std::size_t>(last), charge, inv_mass});
        for (int i = first; i < last; ++i) s.inv_mass[i] = inv_mass;
    });
}

int oracle_add_pade_jastrow(oracle_system* h, int ga, int gb, int num_order, int den_order,
                            const double* coeffs, int n_k, const int* n1, const int* n2,
                            const double* kx, const double* ky, const double* weights,
                            int n_knots, const double* spline, double r_cut, double inv_delta,
                            double u_inf) {
    return guarded([&] {
        Sys& s = *S(h);
        check_index(s, ga);
        check_index(s, gb);
        auto c = std::make_unique<PairJastrow>();
        c->ga = s.groups[ga];
        c->gb = s.groups[gb];
        c->same = ga == gb;
        c->rad.num = num_order;
        c->rad.den = den_order;
        c->rad.c.assign(coeffs, coeffs + num_order + den_order);
        if (n_k > 0) {
            if (!s.cell.periodic) throw std::invalid_argument("oracle: image sum needs a periodic cell");
            c->rad.image = true;
            c->rad.sp.scaled.assign(spline, spline + 3 * n_knots);
            c->rad.sp.n_knots = n_knots;
            c->rad.sp.r_cut = r_cut;
            c->rad.sp.inv_delta = inv_delta;
            c->rad.u_inf = u_inf;
            c->has_k = true;
            KSpace& k = c->ks;
            k.a = c->ga;
            k.b = c->gb;
            k.same = c->same;
            k.n1.assign(n1, n1 + n_k);
            k.n2.assign(n2, n2 + n_k);
            k.kx.assign(kx, kx + n_k);
            k.ky.assign(ky, ky + n_k);
            k.w.assign(weights, weights + n_k);
            for (int i = 0; i < n_k; ++i) {
                k.kx2.push_back(kx[i] * kx[i]);
                k.ky2.push_back(ky[i] * ky[i]);
                k.m1 = std::max(k.m1, std::abs(n1[i]));
                k.m2 = std::max(k.m2, std::abs(n2[i]));
            }
            k.b1x = s.cell.b[0]; k.b1y = s.cell.b[1]; k.b2x = s.cell.b[2]; k.b2y = s.cell.b[3];
            k.n_cached = k.a.size() + (k.same ? 0 : k.b.size());
        }
        s.factors.push_back({c.get(), -1});
        s.comps.push_back(std::move(c));
    });
}

int oracle_add_drummond_jastrow(oracle_system* h, int ga, int gb, int C, double gamma, double L,
                                int n_coeffs, const double* coeffs) {
    return guarded([&] {
        Sys& s = *S(h);
        check_index(s, ga);
        check_index(s, gb);
        if (n_coeffs < 1 || C < 2) throw std::invalid_argument("oracle: drummond needs n >= 1, C >= 2");
        auto c = std::make_unique<PairJastrow>();
        c->ga = s.groups[ga];
        c->gb = s.groups[gb];
        c->same = ga == gb;
        c->rad.drummond = true;
        c->rad.C = C;
        c->rad.L = L;
        c->rad.c.assign(coeffs, coeffs + n_coeffs);
        double ipw = 1.0;
        for (int i = 0; i < C; ++i) ipw *= -L;
        c->rad.alpha1 = gamma / ipw + coeffs[0] * static_cast<double>(C) / L;
        c->rad.gamma = gamma;
        s.factors.push_back({c.get(), -1});
        s.comps.push_back(std::move(c));
    });
}

int oracle_add_mcmillan_jastrow(oracle_system* h, int ga, int gb, double b, int n) {
    return guarded([&] {
        Sys& s = *S(h);
        check_index(s, ga);
        check_index(s, gb);
        if (!(b > 0.0) || !std::isfinite(b)) throw std::invalid_argument("oracle: mcmillan b must be positive");
        if (n < 1 || n > 16) throw std::invalid_argument("oracle: mcmillan exponent must be 1..16");
        auto c = std::make_unique<PairJastrow>();
        c->ga = s.groups[ga];
        c->gb = s.groups[gb];
        c->same = ga == gb;
        c->rad.mcmillan = true;
        c->rad.num = n;
        c->rad.c = {b};
        s.factors.push_back({c.get(), -1});
        s.comps.push_back(std::move(c));
    });
}

// InversePowerTerm constructor (inverse_power.cpp:43-83).
int oracle_add_inverse_power(oracle_system* h, int ga, int gb, double strength, double exponent) {
    return guarded([&] {
        Sys& s = *S(h);
        check_index(s, ga);
        check_index(s, gb);
        if (!std::isfinite(exponent) || exponent <= 0.0)
            throw std::invalid_argument("oracle: inverse_power exponent must be finite and positive");
        if (!std::isfinite(strength)) throw std::invalid_argument("oracle: inverse_power strength must be finite");
        if (strength < 0.0 && exponent >= 2.0)
            throw std::invalid_argument("oracle: inverse_power attractive core with exponent >= 2");
        InversePower t;
        t.ga = ga;
        t.gb = gb;
        t.c = strength;
        t.p = exponent;
        const double rounded = std::round(exponent);
        if (rounded == exponent && rounded <= 16.0) t.p_int = static_cast<int>(rounded);
        if (s.cell.periodic) {
            if (exponent <= 2.0) throw std::invalid_argument("oracle: inverse_power in a periodic cell needs p > 2");
            const double na = static_cast<double>(s.groups[ga].last - s.groups[ga].first);
            const double nb = static_cast<double>(s.groups[gb].last - s.groups[gb].first);
            const double pairs = ga == gb ? 0.5 * na * na : na * nb;
            const double area = std::fabs(s.lattice[0] * s.lattice[3] - s.lattice[1] * s.lattice[2]);
            t.tail = pairs / area * strength * ws_exterior_integral(s.lattice, exponent);
        }
        s.ipow.push_back(t);
    });
}

int oracle_add_nosanow_jastrow(oracle_system* h, int group, double alpha, const double* sites) {
    return guarded([&] {
        Sys& s = *S(h);
        check_index(s, group);
        // NosanowJastrow refuses a Gaussian that is still above 1e-6 of its
        // peak at the Wigner-Seitz boundary (nosanow_jastrow.h:51).
        if (s.cell.periodic && std::exp(-alpha * s.cell.ws * s.cell.ws) > 1e-6)
            throw std::invalid_argument("nosanow: alpha too small for the cell (folded tail above 1e-6)");
        auto c = std::make_unique<Nosanow>();
        c->g = s.groups[group];
        c->alpha = alpha;
        c->sites.assign(sites, sites + c->g.size() * 2);
        s.factors.push_back({c.get(), -1});
        s.comps.push_back(std::move(c));
    });
}

int oracle_add_slater_plane_wave(oracle_system* h, int first, int last, const double* kvecs,
                                 const unsigned char* is_sin) {
    return guarded([&] {
        Sys& s = *S(h);
        RealSlater::Block B;
        B.first = first;
        B.last = last;
        B.N = static_cast<std::size_t>(last - first);
        B.orb.n = B.N;
        B.orb.k.assign(kvecs, kvecs + 2 * B.N);
        B.orb.is_sin.assign(is_sin, is_sin + B.N);
        RealSlater* c = nullptr;
        for (auto& x : s.comps)
            if (auto* r = dynamic_cast<RealSlater*>(x.get())) c = r;
        if (!c) {
            s.comps.push_back(std::make_unique<RealSlater>());
            c = static_cast<RealSlater*>(s.comps.back().get());
        }
        c->blocks.push_back(std::move(B));
        s.factors.push_back({c, static_cast<int>(c->blocks.size()) - 1});
    });
}

int oracle_add_slater_gaussian(oracle_system* h, int first, int last, const double* centers,
                               const double* alphas) {
    return guarded([&] {
        Sys& s = *S(h);
        RealSlater::Block B;
        B.first = first;
        B.last = last;
        B.N = static_cast<std::size_t>(last - first);
        B.orb.n = B.N;
        B.orb.gaussian = true;
        B.orb.centers.assign(centers, centers + 2 * B.N);
        B.orb.alphas.assign(alphas, alphas + B.N);
        RealSlater* c = nullptr;
        for (auto& x : s.comps)
            if (auto* r = dynamic_cast<RealSlater*>(x.get())) c = r;
        if (!c) {
            s.comps.push_back(std::make_unique<RealSlater>());
            c = static_cast<RealSlater*>(s.comps.back().get());
        }
        c->blocks.push_back(std::move(B));
        s.factors.push_back({c, static_cast<int>(c->blocks.size()) - 1});
    });
}

int oracle_add_slater_twisted(oracle_system* h, int first, int last, const double* q) {
    return guarded([&] {
        Sys& s = *S(h);
        ComplexSlater::Block B;
        B.first = first;
        B.last = last;
        B.N = static_cast<std::size_t>(last - first);
        B.q.assign(q, q + 2 * B.N);
        ComplexSlater* c = nullptr;
        for (auto& x : s.comps)
            if (auto* r = dynamic_cast<ComplexSlater*>(x.get())) c = r;
        if (!c) {
            s.comps.push_back(std::make_unique<ComplexSlater>());
            c = static_cast<ComplexSlater*>(s.comps.back().get());
        }
        c->blocks.push_back(std::move(B));
        s.factors.push_back({c, static_cast<int>(c->blocks.size()) - 1});
        s.has_phase = true;
    });
}

int oracle_add_ewald(oracle_system* h, int n_k, const int* n1, const int* n2, const double* kx,
                     const double* ky, const double* weights, int n_knots, const double* spline,
                     double r_cut, double inv_delta, double v_long_zero, double sr_k0) {
    return guarded([&] {
        Sys& s = *S(h);
        if (!s.cell.periodic) throw std::invalid_argument("oracle: ewald needs a periodic cell");
        Ewald& e = s.ewald;
        e.active = true;
        e.n1.assign(n1, n1 + n_k);
        e.n2.assign(n2, n2 + n_k);
        e.kx.assign(kx, kx + n_k);
        e.ky.assign(ky, ky + n_k);
        e.w.assign(weights, weights + n_k);
        e.sp.scaled.assign(spline, spline + 3 * n_knots);
        e.sp.n_knots = n_knots;
        e.sp.r_cut = r_cut;
        e.sp.inv_delta = inv_delta;
        e.r_cut = r_cut;
        e.v_long_zero = v_long_zero;
        e.sr_k0 = sr_k0;
        e.inv_eps = 1.0;
        for (int i = 0; i < n_k; ++i) {
            if (n1[i] < 0) throw std::invalid_argument("oracle: ewald k-list must have n1 >= 0");
            e.n1_max = std::max(e.n1_max, n1[i]);
            e.n2_max = std::max(e.n2_max, std::abs(n2[i]));
        }
    });
}

int oracle_add_coulomb(oracle_system* h, int ga, int gb, double epsilon) {
    return guarded([&] {
        Sys& s = *S(h);
        check_index(s, ga);
        check_index(s, gb);
        s.coulomb.push_back({ga, gb, 1.0 / epsilon});
    });
}

int oracle_set_positions(oracle_system* h, const double* pos) {
    return guarded([&] {
        Sys& s = *S(h);
        std::copy(pos, pos + s.pos.size(), s.pos.begin());
        for (std::size_t w = 0; w < s.nw; ++w)
            for (std::size_t i = 0; i < s.np; ++i) s.cell.wrap(s.P(i, 0, w), s.P(i, 1, w));
        s.build_dist();
    });
}

int oracle_get_positions(oracle_system* h, double* pos) {
    Sys& s = *S(h);
    std::copy(s.pos.begin(), s.pos.end(), pos);
    return 0;
}

int oracle_evaluate(oracle_system* h) {
    return guarded([&] {
        Sys& s = *S(h);
        for (auto& c : s.comps) c->evaluate(s);
    });
}

int oracle_get_log_psi(oracle_system* h, double* out) {
    Sys& s = *S(h);
    for (std::size_t w = 0; w < s.nw; ++w) out[w] = log_psi(s, w);
    return 0;
}

int oracle_get_gradients(oracle_system* h, double* grad, double* lap) {
    return guarded([&] {
        Sys& s = *S(h);
        refresh_all(s);
        std::vector<double> g, l;
        totals(s, g, l, nullptr, nullptr);
        std::copy(g.begin(), g.end(), grad);
        std::copy(l.begin(), l.end(), lap);
    });
}

int oracle_get_phase_gradients(oracle_system* h, double* pgrad, double* plap) {
    return guarded([&] {
        Sys& s = *S(h);
        refresh_all(s);
        std::vector<double> g, l, pg, pl;
        totals(s, g, l, &pg, &pl);
        std::copy(pg.begin(), pg.end(), pgrad);
        std::copy(pl.begin(), pl.end(), plap);
    });
}

int oracle_local_energy(oracle_system* h, double* e_re, double* e_im) {
    return guarded([&] {
        Sys& s = *S(h);
        local_energy(s);
        std::copy(s.e_loc.begin(), s.e_loc.end(), e_re);
        if (e_im) std::copy(s.e_loc_im.begin(), s.e_loc_im.end(), e_im);
    });
}

int oracle_ratio(oracle_system* h, int p, const double* new_pos, double* log_ratio, double* phase,
                 double* new_grad, double* new_lap) {
    return guarded([&] {
        Sys& s = *S(h);
        if (p < 0 || static_cast<std::size_t>(p) >= s.np) throw std::invalid_argument("oracle: particle out of range");
        std::copy(new_pos, new_pos + 2 * s.nw, s.new_pos.begin());
        ratio_all(s, static_cast<std::size_t>(p));
        std::copy(s.log_ratios.begin(), s.log_ratios.end(), log_ratio);
        if (phase) std::copy(s.phase_diffs.begin(), s.phase_diffs.end(), phase);
        if (new_grad) std::copy(s.new_grad.begin(), s.new_grad.end(), new_grad);
        if (new_lap) std::copy(s.new_lap.begin(), s.new_lap.end(), new_lap);
    });
}

int oracle_accept(oracle_system* h, int p, const unsigned char* mask) {
    return guarded([&] {
        Sys& s = *S(h);
        std::vector<double> m(s.nw);
        for (std::size_t w = 0; w < s.nw; ++w) m[w] = mask[w] ? 1.0 : 0.0;
        accept_all(s, static_cast<std::size_t>(p), m);
    });
}

// ratio_batch / apply_move_batch over the walkers [w_begin, w_end)
// (trial_wavefunction.cpp:46-138): arrays keep the batch's walker stride and
// only the range's entries are read or written.  The restatement recomputes
// the whole batch with the other walkers' pending proposals left as they
// were, so calls for interleaved ranges (crowds) see their own proposals.
static std::pair<std::size_t, std::size_t> oracle_range(const Sys& s, int w_begin, int w_end) {
    const long long nw = static_cast<long long>(s.nw);
    const long long w1 = w_end < 0 ? nw : w_end;
    if (w_begin < 0 || w_begin > w1 || w1 > nw) throw std::invalid_argument("walker range outside the batch");
    return {static_cast<std::size_t>(w_begin), static_cast<std::size_t>(w1)};
}

int oracle_ratio_range(oracle_system* h, int p, int w_begin, int w_end, const double* new_pos, double* log_ratio,
                       double* phase, double* new_grad, double* new_lap) {
    return guarded([&] {
        Sys& s = *S(h);
        if (p < 0 || static_cast<std::size_t>(p) >= s.np) throw std::invalid_argument("oracle: particle out of range");
        const auto [w0, w1] = oracle_range(s, w_begin, w_end);
        const std::size_t nw = s.nw;
        for (std::size_t d = 0; d < 2; ++d)
            for (std::size_t w = w0; w < w1; ++w) s.new_pos[d * nw + w] = new_pos[d * nw + w];
        ratio_all(s, static_cast<std::size_t>(p));
        for (std::size_t w = w0; w < w1; ++w) {
            log_ratio[w] = s.log_ratios[w];
            if (phase) phase[w] = s.phase_diffs[w];
            for (std::size_t d = 0; d < 2; ++d) {
                if (new_grad) new_grad[d * nw + w] = s.new_grad[d * nw + w];
                if (new_lap) new_lap[d * nw + w] = s.new_lap[d * nw + w];
            }
        }
    });
}

int oracle_ratio_range_phase(oracle_system* h, int p, int w_begin, int w_end, const double* new_pos,
                             double* log_ratio, double* phase, double* new_grad, double* new_lap, double* new_pgrad,
                             double* new_plap) {
    return guarded([&] {
        Sys& s = *S(h);
        if (p < 0 || static_cast<std::size_t>(p) >= s.np) throw std::invalid_argument("oracle: particle out of range");
        const auto [w0, w1] = oracle_range(s, w_begin, w_end);
        const std::size_t nw = s.nw;
        for (std::size_t d = 0; d < 2; ++d)
            for (std::size_t w = w0; w < w1; ++w) s.new_pos[d * nw + w] = new_pos[d * nw + w];
        ratio_all(s, static_cast<std::size_t>(p));
        for (std::size_t w = w0; w < w1; ++w) {
            log_ratio[w] = s.log_ratios[w];
            if (phase) phase[w] = s.phase_diffs[w];
            for (std::size_t d = 0; d < 2; ++d) {
                const std::size_t i = d * nw + w;
                if (new_grad) new_grad[i] = s.new_grad[i];
                if (new_lap) new_lap[i] = s.new_lap[i];
                if (new_pgrad) new_pgrad[i] = s.new_pgrad.empty() ? 0.0 : s.new_pgrad[i];
                if (new_plap) new_plap[i] = s.new_plap.empty() ? 0.0 : s.new_plap[i];
            }
        }
    });
}

// log|psi| and the derivative planes of the walkers [w_begin, w_end) only.
int oracle_get_derivatives_range(oracle_system* h, int w_begin, int w_end, double* lp, double* grad, double* lap,
                                 double* pgrad, double* plap) {
    return guarded([&] {
        Sys& s = *S(h);
        const auto [w0, w1] = oracle_range(s, w_begin, w_end);
        for (std::size_t w = w0; w < w1; ++w)
            if (lp) lp[w] = log_psi(s, w);
        if (!grad && !lap && !pgrad && !plap) return;
        refresh_all(s);
        std::vector<double> g, l, pg, pl;
        totals(s, g, l, &pg, &pl);
        const std::size_t nw = s.nw;
        for (std::size_t r = 0; r < s.np * 2; ++r)
            for (std::size_t w = w0; w < w1; ++w) {
                const std::size_t i = r * nw + w;
                if (grad) grad[i] = g[i];
                if (lap) lap[i] = l[i];
                if (pgrad) pgrad[i] = pg[i];
                if (plap) plap[i] = pl[i];
            }
    });
}

// Particle p's rows of the derivative planes, [dim][nw], for the walkers
// [w_begin, w_end) only.
int oracle_get_particle_derivatives_range(oracle_system* h, int p, int w_begin, int w_end, double* grad, double* lap,
                                          double* pgrad, double* plap) {
    return guarded([&] {
        Sys& s = *S(h);
        if (p < 0 || static_cast<std::size_t>(p) >= s.np) throw std::invalid_argument("oracle: particle out of range");
        const auto [w0, w1] = oracle_range(s, w_begin, w_end);
        if (!grad && !lap && !pgrad && !plap) return;
        refresh_all(s);
        std::vector<double> g, l, pg, pl;
        totals(s, g, l, &pg, &pl);
        const std::size_t nw = s.nw;
        for (std::size_t d = 0; d < 2; ++d)
            for (std::size_t w = w0; w < w1; ++w) {
                const std::size_t i = (static_cast<std::size_t>(p) * 2 + d) * nw + w, o = d * nw + w;
                if (grad) grad[o] = g[i];
                if (lap) lap[o] = l[i];
                if (pgrad) pgrad[o] = pg[i];
                if (plap) plap[o] = pl[i];
            }
    });
}

// Hamiltonian::delta_potential_particle (hamiltonian.cpp:105-117) for each
// walker of [w_begin, w_end): the potential of the batch with particle p at
// new_pos[.][w] (wrapped into the cell) minus the potential now.  Ewald as
// EwaldTerm::delta_potential_particle (ewald_term.cpp:750-815): the moved
// particle's real-space pairs against the distance table, and
// Sum_k w_k [2 Re(delta rho*) + |delta|^2] with delta = q_p (e^{ik.r'} -
// e^{ik.r}) against rho of the current positions (computed here afresh; the
// reference keeps it in a cache that is exact by construction).  Coulomb
// (coulomb.cpp:148-155) through radial_particle_delta
// (hamiltonian_component.h:25-48); inverse powers, which the reference
// evaluates before and after (hamiltonian_component.h:216-260), through the
// same pair difference.
int oracle_delta_potential(oracle_system* h, int p, int w_begin, int w_end, const double* new_pos, double* delta) {
    return guarded([&] {
        Sys& s = *S(h);
        if (p < 0 || static_cast<std::size_t>(p) >= s.np) throw std::invalid_argument("oracle: particle out of range");
        const auto [w0, w1] = oracle_range(s, w_begin, w_end);
        const std::size_t nw = s.nw, np = s.np, pp = static_cast<std::size_t>(p);
        std::vector<double> q(np, 0.0);
        std::vector<int> group_of(np, -1);
        for (std::size_t g = 0; g < s.groups.size(); ++g)
            for (std::size_t i = s.groups[g].first; i < s.groups[g].last; ++i) {
                q[i] = s.groups[g].charge;
                group_of[i] = static_cast<int>(g);
            }
        for (std::size_t w = w0; w < w1; ++w) {
            double c0 = new_pos[w], c1 = new_pos[nw + w];
            s.cell.wrap(c0, c1);
            const double o0 = s.P(pp, 0, w), o1 = s.P(pp, 1, w);
            auto r_new = [&](std::size_t j) {
                double d0 = c0 - s.P(j, 0, w), d1 = c1 - s.P(j, 1, w);
                s.cell.min_image(d0, d1);
                return std::sqrt(d0 * d0 + d1 * d1);
            };
            double d = 0.0;
            if (s.ewald.active && q[pp] != 0.0) {
                const Ewald& e = s.ewald;
                auto short_range = [&](double r) { return r < e.r_cut ? e.inv_eps / r - e.sp.value(r) : 0.0; };
                for (std::size_t j = 0; j < np; ++j) {
                    if (j == pp || q[j] == 0.0) continue;
                    d += q[pp] * q[j] * (short_range(r_new(j)) - short_range(s.D(pp, j, w)));
                }
                for (std::size_t kk = 0; kk < e.w.size(); ++kk) {
                    double rr = 0.0, ri = 0.0;
                    for (std::size_t i = 0; i < np; ++i) {
                        if (q[i] == 0.0) continue;
                        const double ph = e.kx[kk] * s.P(i, 0, w) + e.ky[kk] * s.P(i, 1, w);
                        rr += q[i] * std::cos(ph);
                        ri += q[i] * std::sin(ph);
                    }
                    const double pn = e.kx[kk] * c0 + e.ky[kk] * c1, po = e.kx[kk] * o0 + e.ky[kk] * o1;
                    const double dr = q[pp] * (std::cos(pn) - std::cos(po)), di = q[pp] * (std::sin(pn) - std::sin(po));
                    d += e.w[kk] * (2.0 * (dr * rr + di * ri) + dr * dr + di * di);
                }
            }
            auto pairs = [&](int ga, int gb, auto radial) {
                auto one = [&](std::size_t j) {
                    if (j != pp) d += radial(r_new(j)) - radial(s.D(pp, j, w));
                };
                const Group& a = s.groups[ga];
                const Group& b = s.groups[gb];
                if (ga == gb) {
                    if (group_of[pp] == ga)
                        for (std::size_t j = a.first; j < a.last; ++j) one(j);
                } else {
                    if (group_of[pp] == ga)
                        for (std::size_t j = b.first; j < b.last; ++j) one(j);
                    if (group_of[pp] == gb)
                        for (std::size_t j = a.first; j < a.last; ++j) one(j);
                }
            };
            for (const auto& c : s.coulomb) {
                const double coeff = s.groups[c.ga].charge * s.groups[c.gb].charge * c.inv_eps;
                pairs(c.ga, c.gb, [&](double r) { return coeff / std::sqrt(r * r + 0.0); });
            }
            for (const auto& t : s.ipow)
                pairs(t.ga, t.gb, [&](double r) {
                    if (t.p_int > 0) {
                        const double inv_r = 1.0 / r;
                        double inv_rp = inv_r;
                        for (int k = 1; k < t.p_int; ++k) inv_rp *= inv_r;
                        return t.c * inv_rp;
                    }
                    return t.c * std::pow(r, -t.p);
                });
            delta[w] = d;
        }
    });
}

}  // extern "C"

namespace {

// ── optimizer derivatives ──────────────────────────────────────────────────
// f = du/dp of a pair factor's parameter q at distance r, with f' and f''.
// Padé (PadeFunction's parameter derivatives behind pade_jastrow.cpp:641-700,
// 806-880): du/da_k = r^e/Q (e = k+1), du/db_k = -P r^e/Q^2; summed over
// images the short-range part takes f - limit - spline(r) inside r_cut
// (periodic_pair_sum.cpp:85-98).  Drummond (drummond_jastrow.cpp:201-420),
// inside the cutoff, with t = r - L and g = t^C: du/dalpha_0 = g (1 + C r/L)
// (alpha_1 = gamma/(-L)^C + alpha_0 C/L follows alpha_0),
// du/dalpha_l = g r^l, du/dL = -C t^(C-1) P + g r dalpha_1/dL.
struct ParamRadial {
    const Radial& R;
    int index;
    const Spline* sp;     // the parameter's spline (image-summed Padé)
    double limit;
    void eval(double r, double& f, double& f1, double& f2) const {
        f = f1 = f2 = 0.0;
        if (R.drummond) {
            if (r >= R.L) return;
            const int C = R.C;
            const double Cd = static_cast<double>(C), t = r - R.L;
            auto ip = [](double x, int e) { double y = 1.0; for (int i = 0; i < e; ++i) y *= x; return y; };
            const double g = ip(t, C), g1 = Cd * ip(t, C - 1), g2 = Cd * (C - 1) * ip(t, C - 2);
            const int nc = static_cast<int>(R.c.size());
            if (index == 0) {
                const double k = Cd / R.L;
                f = g * (1.0 + k * r);
                f1 = g1 * (1.0 + k * r) + g * k;
                f2 = g2 * (1.0 + k * r) + 2.0 * g1 * k;
            } else if (index < nc) {
                const int l = index + 1;
                const double m = ip(r, l), m1 = l * ip(r, l - 1), m2 = l * (l - 1) * ip(r, l - 2);
                f = g * m;
                f1 = g1 * m + g * m1;
                f2 = g2 * m + 2.0 * g1 * m1 + g * m2;
            } else {
                double P = R.c[0] + R.alpha1 * r, P1 = R.alpha1, P2 = 0.0;
                for (int k = 1; k < nc; ++k) {
                    const int l = k + 1;
                    P += R.c[static_cast<std::size_t>(k)] * ip(r, l);
                    P1 += R.c[static_cast<std::size_t>(k)] * l * ip(r, l - 1);
                    P2 += R.c[static_cast<std::size_t>(k)] * l * (l - 1) * ip(r, l - 2);
                }
                const double h = -Cd * ip(t, C - 1), h1 = -Cd * (C - 1) * ip(t, C - 2),
                             h2 = C >= 3 ? -Cd * (C - 1) * (C - 2) * ip(t, C - 3) : 0.0;
                const double a = R.gamma * Cd / ip(-R.L, C + 1) - R.c[0] * Cd / (R.L * R.L);
                f = h * P + g * a * r;
                f1 = h1 * P + h * P1 + g1 * a * r + g * a;
                f2 = h2 * P + 2.0 * h1 * P1 + h * P2 + g2 * a * r + 2.0 * g1 * a;
            }
            return;
        }
        double P = 0.0, P1 = 0.0, P2 = 0.0, Q = 1.0, Q1 = 0.0, Q2 = 0.0;
        Radial::poly(R.c.data(), R.num, r, P, P1, P2);
        Radial::poly(R.c.data() + R.num, R.den, r, Q, Q1, Q2);
        const bool num = index < R.num;
        const int e = num ? index + 1 : index - R.num + 1;
        double m = 1.0;
        for (int i = 0; i < e; ++i) m *= r;
        const double m1 = e * m / r, m2 = e * (e - 1) * m / (r * r);
        const double iq = 1.0 / Q;
        if (num) {
            f = m * iq;
            f1 = m1 * iq - m * Q1 * iq * iq;
            f2 = m2 * iq - 2.0 * m1 * Q1 * iq * iq - m * Q2 * iq * iq + 2.0 * m * Q1 * Q1 * iq * iq * iq;
        } else {
            const double hh = P * m, h1 = P1 * m + P * m1, h2 = P2 * m + 2.0 * P1 * m1 + P * m2;
            const double iq2 = iq * iq, iq3 = iq2 * iq;
            f = -hh * iq2;
            f1 = -h1 * iq2 + 2.0 * hh * Q1 * iq3;
            f2 = -h2 * iq2 + 4.0 * h1 * Q1 * iq3 + 2.0 * hh * Q2 * iq3 - 6.0 * hh * Q1 * Q1 * iq3 * iq;
        }
        if (R.image) {
            if (r >= sp->r_cut) { f = f1 = f2 = 0.0; return; }
            double v, d1, d2;
            sp->eval(r, v, d1, d2);
            f -= limit + v;
            f1 -= d1;
            f2 -= d2;
        }
    }
};

// One listed parameter, resolved against the system's factors.
struct ResolvedParam {
    PairJastrow* pj = nullptr;
    Nosanow* ns = nullptr;
    RealSlater* det = nullptr;
    int block = -1;
    Spline sp;
};

ResolvedParam resolve_param(Sys& s, const oracle_param& q) {
    if (q.factor < 0 || static_cast<std::size_t>(q.factor) >= s.factors.size())
        throw std::invalid_argument("oracle: factor index out of range");
    ResolvedParam out;
    const auto [c, blk] = s.factors[static_cast<std::size_t>(q.factor)];
    if (auto* pj = dynamic_cast<PairJastrow*>(c)) {
        const Radial& R = pj->rad;
        if (R.mcmillan) throw std::invalid_argument("oracle: the McMillan factor has no free parameters");
        const int n_par = R.drummond ? static_cast<int>(R.c.size()) + 1 : R.num + R.den;
        if (q.index < 0 || q.index >= n_par) throw std::invalid_argument("oracle: parameter index out of range");
        if (R.image) {
            if (!q.weights || !q.spline) throw std::invalid_argument("oracle: image-summed parameter needs its tables");
            out.sp.scaled.assign(q.spline, q.spline + 3 * q.n_knots);
            out.sp.n_knots = q.n_knots;
            out.sp.r_cut = q.r_cut;
            out.sp.inv_delta = q.inv_delta;
        }
        out.pj = pj;
    } else if (auto* ns = dynamic_cast<Nosanow*>(c)) {
        if (q.index != 0) throw std::invalid_argument("oracle: the Nosanow factor has one parameter");
        out.ns = ns;
    } else if (auto* rs = dynamic_cast<RealSlater*>(c)) {
        const auto& B = rs->blocks[static_cast<std::size_t>(blk)];
        bool uniform = B.orb.gaussian;
        for (std::size_t m = 1; uniform && m < B.N; ++m) uniform = B.orb.alphas[m] == B.orb.alphas[0];
        if (!uniform) throw std::invalid_argument("oracle: this determinant has no free parameters");
        if (q.index != 0) throw std::invalid_argument("oracle: parameter index out of range");
        out.det = rs;
        out.block = blk;
    } else {
        throw std::invalid_argument("oracle: this determinant has no free parameters");
    }
    return out;
}

template <class F>
void for_pairs(const PairJastrow& pj, F&& f) {
    if (pj.same) {
        for (std::size_t a = pj.ga.first; a < pj.ga.last; ++a)
            for (std::size_t b = a + 1; b < pj.ga.last; ++b) f(a, b);
    } else {
        for (std::size_t a = pj.ga.first; a < pj.ga.last; ++a)
            for (std::size_t b = pj.gb.first; b < pj.gb.last; ++b) f(a, b);
    }
}

// dphi/dalpha of a Gaussian orbital at separation (d0, d1) from its centre
// (GaussianOrbitalSet::evaluate_parameter_derivatives, orbital_set.h:249-289):
// -r^2 phi, d/dalpha of its gradient -2 delta_d phi (1 - alpha r^2) and of its
// second derivatives phi [8 alpha delta_d^2 - 2 - r^2 (4 alpha^2 delta_d^2 -
// 2 alpha)].
void gauss_alpha_terms(double a, double d0, double d1, double& v, double* g, double* l) {
    const double r2 = d0 * d0 + d1 * d1, phi = std::exp(-a * r2);
    v = -r2 * phi;
    const double del[2] = {d0, d1};
    for (int d = 0; d < 2; ++d) {
        g[d] = -2.0 * del[d] * phi * (1.0 - a * r2);
        l[d] = phi * (8.0 * a * del[d] * del[d] - 2.0 - r2 * (4.0 * a * a * del[d] * del[d] - 2.0 * a));
    }
}

}  // namespace

extern "C" {

// TrialWavefunction::compute_parameter_derivatives (trial_wavefunction.cpp:
// 150-160) for the listed free parameters: derivs[i * nw + w] = d log|psi| /
// d p_i of walker w.  A pair factor sums du/dp over its pairs at the
// minimum-image distance (ParamRadial), plus, summed over images, the
// reciprocal part sum_k w_k |rho_a|^2 (one group) or 2 w_k Re(rho_a rho_b*)
// with the parameter's weights (kspace_pair_sum.cpp:303-313).  Nosanow
// (nosanow_jastrow.cpp:176-190): -sum_i |r_i - R_i|^2.  A Gaussian block's
// shared width (slater_determinant.cpp:98-157): sum_{c,a} A(a,c) dphi_a(r_c)/dalpha.
int oracle_parameter_derivatives(oracle_system* h, const oracle_param* params, int n, double* derivs) {
    return guarded([&] {
        Sys& s = *S(h);
        const std::size_t nw = s.nw;
        for (int i = 0; i < n; ++i) {
            const oracle_param& q = params[i];
            ResolvedParam rp = resolve_param(s, q);
            double* out = derivs + static_cast<std::size_t>(i) * nw;
            std::fill(out, out + nw, 0.0);
            if (rp.pj) {
                const PairJastrow& pj = *rp.pj;
                const ParamRadial pr{pj.rad, q.index, &rp.sp, q.limit};
                for_pairs(pj, [&](std::size_t a, std::size_t b) {
                    for (std::size_t w = 0; w < nw; ++w) {
                        double f, f1, f2;
                        pr.eval(s.D(a, b, w), f, f1, f2);
                        out[w] += f;
                    }
                });
                if (pj.rad.image) {
                    const KSpace& k = pj.ks;
                    const std::size_t K = k.nk();
                    for (std::size_t w = 0; w < nw; ++w) {
                        const double* ar = k.ra_re.data() + w * K;
                        const double* ai = k.ra_im.data() + w * K;
                        const double* br = k.same ? ar : k.rb_re.data() + w * K;
                        const double* bi = k.same ? ai : k.rb_im.data() + w * K;
                        double t = 0.0;
                        for (std::size_t kk = 0; kk < K; ++kk)
                            t += k.same ? q.weights[kk] * (ar[kk] * ar[kk] + ai[kk] * ai[kk])
                                        : 2.0 * q.weights[kk] * (ar[kk] * br[kk] + ai[kk] * bi[kk]);
                        out[w] += t;
                    }
                }
            } else if (rp.ns) {
                const Nosanow& ns = *rp.ns;
                for (std::size_t p = ns.g.first; p < ns.g.last; ++p)
                    for (std::size_t w = 0; w < nw; ++w) {
                        double d0 = s.P(p, 0, w) - ns.sites[(p - ns.g.first) * 2];
                        double d1 = s.P(p, 1, w) - ns.sites[(p - ns.g.first) * 2 + 1];
                        if (s.cell.periodic) s.cell.min_image(d0, d1);
                        out[w] -= d0 * d0 + d1 * d1;
                    }
            } else {
                const auto& B = rp.det->blocks[static_cast<std::size_t>(rp.block)];
                const std::size_t N = B.N;
                const double alpha = B.orb.alphas[0];
                for (std::size_t w = 0; w < nw; ++w) {
                    const double* A = B.A.data() + w * N * N;
                    double t = 0.0;
                    for (std::size_t c = 0; c < N; ++c)
                        for (std::size_t m = 0; m < N; ++m) {
                            double d0 = s.P(B.first + c, 0, w) - B.orb.centers[m * 2];
                            double d1 = s.P(B.first + c, 1, w) - B.orb.centers[m * 2 + 1];
                            if (s.cell.periodic) s.cell.min_image(d0, d1);
                            double v, g[2], l[2];
                            gauss_alpha_terms(alpha, d0, d1, v, g, l);
                            t += A[m * N + c] * v;
                        }
                    out[w] = t;
                }
            }
        }
    });
}

// Hamiltonian::local_energy_parameter_derivatives (hamiltonian.cpp:189-196)
// fed by TrialWavefunction::compute_parameter_gradient_derivatives
// (trial_wavefunction.cpp:162-185): for each listed parameter the change of
// every particle's gradient and Laplacian of log|psi|, reduced by the kinetic
// term (kinetic.cpp:176-200, 212-250) as
//   dE/dp = sum_{i,d} -1/2 (1/m_i) [dL + 2 G dG]   (real part)
//   dE_im/dp = sum_{i,d} -1/2 (1/m_i) 2 dG v        (v the phase gradient)
// with G the whole trial's gradient.  A pair factor's f = du/dp enters as u
// does: dG_i += f' delta/r (- for the partner), dL += f'' (delta_d/r)^2 +
// (f'/r)(1 - (delta_d/r)^2) for both (pade_jastrow.cpp:806-880); summed over
// images the reciprocal part takes the parameter's weights in the
// structure-factor gradient and Laplacian (kspace_pair_sum.cpp:315-320).
// Nosanow: dG = -2 delta, dL = -2 per dimension.  A Gaussian block's width
// (slater_determinant.cpp:159-251): dA = -A dD A, dG_j = sum_m dA(m,j)
// dphi_m(r_j) + A(m,j) d(dphi_m)/dalpha, the same for the second
// derivatives, and dL = dLraw - 2 G_det dG with the block's own gradient.
int oracle_local_energy_parameter_derivatives(oracle_system* h, const oracle_param* params, int n,
                                              double* de_re, double* de_im) {
    return guarded([&] {
        Sys& s = *S(h);
        const std::size_t nw = s.nw, np = s.np, stride = np * 2 * nw;
        refresh_all(s);
        std::vector<double> G, L, PG, PL;
        totals(s, G, L, &PG, &PL);
        std::vector<double> dg(stride), dl(stride);
        for (int i = 0; i < n; ++i) {
            const oracle_param& q = params[i];
            ResolvedParam rp = resolve_param(s, q);
            std::fill(dg.begin(), dg.end(), 0.0);
            std::fill(dl.begin(), dl.end(), 0.0);
            auto at = [&](std::size_t p, int d, std::size_t w) { return (p * 2 + d) * nw + w; };
            if (rp.pj) {
                const PairJastrow& pj = *rp.pj;
                const ParamRadial pr{pj.rad, q.index, &rp.sp, q.limit};
                for_pairs(pj, [&](std::size_t a, std::size_t b) {
                    for (std::size_t w = 0; w < nw; ++w) {
                        double d0 = s.P(a, 0, w) - s.P(b, 0, w), d1 = s.P(a, 1, w) - s.P(b, 1, w);
                        if (s.cell.periodic) s.cell.min_image(d0, d1);
                        const double r = s.D(a, b, w);
                        double f, f1, f2;
                        pr.eval(r, f, f1, f2);
                        const double del[2] = {d0, d1};
                        for (int d = 0; d < 2; ++d) {
                            const double o = del[d] / r, gd = f1 * o, ld = f2 * o * o + f1 / r * (1.0 - o * o);
                            dg[at(a, d, w)] += gd;
                            dg[at(b, d, w)] -= gd;
                            dl[at(a, d, w)] += ld;
                            dl[at(b, d, w)] += ld;
                        }
                    }
                });
                if (pj.rad.image) {
                    const KSpace& k = pj.ks;
                    const std::size_t K = k.nk();
                    std::vector<double> c(K), sn(K);
                    for (std::size_t w = 0; w < nw; ++w) {
                        const double* ar = k.ra_re.data() + w * K;
                        const double* ai = k.ra_im.data() + w * K;
                        const double* br = k.same ? ar : k.rb_re.data() + w * K;
                        const double* bi = k.same ? ai : k.rb_im.data() + w * K;
                        auto add = [&](std::size_t p, const double* pr_, const double* pi_, double self) {
                            k.phases(s.P(p, 0, w), s.P(p, 1, w), c.data(), sn.data());
                            for (std::size_t kk = 0; kk < K; ++kk) {
                                const double g = 2.0 * q.weights[kk] * (c[kk] * pi_[kk] - sn[kk] * pr_[kk]);
                                const double l = 2.0 * q.weights[kk] * (self - (c[kk] * pr_[kk] + sn[kk] * pi_[kk]));
                                dg[at(p, 0, w)] += k.kx[kk] * g;
                                dg[at(p, 1, w)] += k.ky[kk] * g;
                                dl[at(p, 0, w)] += k.kx2[kk] * l;
                                dl[at(p, 1, w)] += k.ky2[kk] * l;
                            }
                        };
                        for (std::size_t p = k.a.first; p < k.a.last; ++p) add(p, br, bi, k.same ? 1.0 : 0.0);
                        if (!k.same)
                            for (std::size_t p = k.b.first; p < k.b.last; ++p) add(p, ar, ai, 0.0);
                    }
                }
            } else if (rp.ns) {
                const Nosanow& ns = *rp.ns;
                for (std::size_t p = ns.g.first; p < ns.g.last; ++p)
                    for (std::size_t w = 0; w < nw; ++w) {
                        double d0 = s.P(p, 0, w) - ns.sites[(p - ns.g.first) * 2];
                        double d1 = s.P(p, 1, w) - ns.sites[(p - ns.g.first) * 2 + 1];
                        if (s.cell.periodic) s.cell.min_image(d0, d1);
                        dg[at(p, 0, w)] = -2.0 * d0;
                        dg[at(p, 1, w)] = -2.0 * d1;
                        dl[at(p, 0, w)] = -2.0;
                        dl[at(p, 1, w)] = -2.0;
                    }
            } else {
                const auto& B = rp.det->blocks[static_cast<std::size_t>(rp.block)];
                const std::size_t N = B.N;
                const double alpha = B.orb.alphas[0];
                std::vector<double> dD(N * N), Y(N * N), dA(N * N), val(N), og(N * 2), ol(N * 2);
                for (std::size_t w = 0; w < nw; ++w) {
                    const double* A = B.A.data() + w * N * N;
                    auto sep = [&](std::size_t c, std::size_t m, double& d0, double& d1) {
                        d0 = s.P(B.first + c, 0, w) - B.orb.centers[m * 2];
                        d1 = s.P(B.first + c, 1, w) - B.orb.centers[m * 2 + 1];
                        if (s.cell.periodic) s.cell.min_image(d0, d1);
                    };
                    for (std::size_t c = 0; c < N; ++c)
                        for (std::size_t m = 0; m < N; ++m) {
                            double d0, d1, v, g[2], l[2];
                            sep(c, m, d0, d1);
                            gauss_alpha_terms(alpha, d0, d1, v, g, l);
                            dD[c * N + m] = v;   // dD(particle c, orbital m)
                        }
                    // Y = dD A (particle x particle), dA = -A Y (orbital x particle)
                    for (std::size_t c = 0; c < N; ++c)
                        for (std::size_t j = 0; j < N; ++j) {
                            double t = 0.0;
                            for (std::size_t m = 0; m < N; ++m) t += dD[c * N + m] * A[m * N + j];
                            Y[c * N + j] = t;
                        }
                    for (std::size_t m = 0; m < N; ++m)
                        for (std::size_t j = 0; j < N; ++j) {
                            double t = 0.0;
                            for (std::size_t c = 0; c < N; ++c) t += A[m * N + c] * Y[c * N + j];
                            dA[m * N + j] = -t;
                        }
                    for (std::size_t j = 0; j < N; ++j) {
                        B.orb.eval(s.cell, s.P(B.first + j, 0, w), s.P(B.first + j, 1, w), val.data(), og.data(),
                                   ol.data());
                        for (int d = 0; d < 2; ++d) {
                            double Gd = 0.0, dG = 0.0, dLr = 0.0;
                            for (std::size_t m = 0; m < N; ++m) {
                                double d0, d1, v, g[2], l[2];
                                sep(j, m, d0, d1);
                                gauss_alpha_terms(alpha, d0, d1, v, g, l);
                                const double a = A[m * N + j], da = dA[m * N + j];
                                Gd += a * og[m * 2 + d];
                                dG += da * og[m * 2 + d] + a * g[d];
                                dLr += da * ol[m * 2 + d] + a * l[d];
                            }
                            dg[at(B.first + j, d, w)] = dG;
                            dl[at(B.first + j, d, w)] = dLr - 2.0 * Gd * dG;
                        }
                    }
                }
            }
            double* re = de_re + static_cast<std::size_t>(i) * nw;
            std::fill(re, re + nw, 0.0);
            double* im = de_im ? de_im + static_cast<std::size_t>(i) * nw : nullptr;
            if (im) std::fill(im, im + nw, 0.0);
            for (std::size_t p = 0; p < np; ++p)
                for (int d = 0; d < 2; ++d) {
                    const double mi = s.inv_mass[p];
                    for (std::size_t w = 0; w < nw; ++w) {
                        const std::size_t o = at(p, d, w);
                        re[w] += -0.5 * mi * (dl[o] + 2.0 * G[o] * dg[o]);
                        if (im && s.has_phase) im[w] += -0.5 * mi * (2.0 * dg[o] * PG[o]);
                    }
                }
        }
    });
}

int oracle_accept_range(oracle_system* h, int p, int w_begin, int w_end, const unsigned char* mask) {
    return guarded([&] {
        Sys& s = *S(h);
        const auto [w0, w1] = oracle_range(s, w_begin, w_end);
        std::vector<double> m(s.nw, 0.0);
        for (std::size_t w = w0; w < w1; ++w) m[w] = mask[w] ? 1.0 : 0.0;
        accept_all(s, static_cast<std::size_t>(p), m);
    });
}

// VMC::sweep_walkers with symmetric proposals (vmc.cpp:82-155).
int oracle_vmc_sweeps(oracle_system* h, int n_sweeps, double tau, const double* tape,
                      long long* n_accepted) {
    return guarded([&] {
        Sys& s = *S(h);
        const std::size_t nw = s.nw, np = s.np;
        long long acc = 0;
        std::vector<double> m(nw);
        for (int sw = 0; sw < n_sweeps; ++sw) {
            const double* randoms = tape + static_cast<std::size_t>(sw) * np * 3 * nw;
            for (std::size_t p = 0; p < np; ++p) {
                const double* gauss = randoms + p * 3 * nw;
                const double* uni = gauss + 2 * nw;
                const double sigma = std::sqrt(tau * s.inv_mass[p]);
                for (std::size_t d = 0; d < 2; ++d)
                    for (std::size_t w = 0; w < nw; ++w)
                        s.new_pos[d * nw + w] = s.P(p, d, w) + sigma * gauss[d * nw + w];
                ratio_all(s, p);
                for (std::size_t w = 0; w < nw; ++w) {
                    m[w] = uni[w] < std::exp(2.0 * s.log_ratios[w]) ? 1.0 : 0.0;
                    if (m[w] != 0.0) ++acc;
                }
                accept_all(s, p, m);
            }
        }
        *n_accepted = acc;
    });
}

// VMC::sweep_walkers with use_drift (vmc.cpp:98-139): the particle's
// derivatives refreshed, the bare Langevin drift (tau/m) grad log|psi| in the
// proposal, and Metropolis-Hastings with the per-dimension Green's function
// log-ratio -delta (g + g') + (tau/m)/2 (g^2 - g'^2)
// (numerics/green_function.h:20-26).  No node gate: the reference's VMC has
// none.
int oracle_vmc_drift_sweeps(oracle_system* h, int n_sweeps, double tau, const double* tape,
                            long long* n_accepted) {
    return guarded([&] {
        Sys& s = *S(h);
        const std::size_t nw = s.nw, np = s.np;
        long long acc = 0;
        std::vector<double> m(nw), g, l;
        for (int sw = 0; sw < n_sweeps; ++sw) {
            const double* randoms = tape + static_cast<std::size_t>(sw) * np * 3 * nw;
            for (std::size_t p = 0; p < np; ++p) {
                for (auto& c : s.comps) c->refresh_particle(s, p);
                totals(s, g, l, nullptr, nullptr);
                const double tm = tau * s.inv_mass[p];
                const double sigma = std::sqrt(tm);
                const double* gauss = randoms + p * 3 * nw;
                const double* uni = gauss + 2 * nw;
                for (std::size_t d = 0; d < 2; ++d)
                    for (std::size_t w = 0; w < nw; ++w)
                        s.new_pos[d * nw + w] = s.P(p, d, w) + tm * g[(p * 2 + d) * nw + w] + sigma * gauss[d * nw + w];
                ratio_all(s, p);
                for (std::size_t w = 0; w < nw; ++w) {
                    double lc = 0.0;
                    for (std::size_t d = 0; d < 2; ++d) {
                        const double delta = s.new_pos[d * nw + w] - s.P(p, d, w);
                        const double go = g[(p * 2 + d) * nw + w], gn = s.new_grad[d * nw + w];
                        lc += -delta * (go + gn) + 0.5 * tm * (go * go - gn * gn);
                    }
                    const double la = 2.0 * s.log_ratios[w] + lc;
                    m[w] = (la >= 0.0 || uni[w] < std::exp(la)) ? 1.0 : 0.0;
                    if (m[w] != 0.0) ++acc;
                }
                accept_all(s, p, m);
            }
        }
        *n_accepted = acc;
    });
}

// DMC::sweep_walkers (dmc.cpp:203-337): Umrigar-scaled drift, fixed-node
// gate (real trials) or fixed phase (complex), Metropolis-Hastings with the
// drift Green's function, rr accounting.  tape rows have stride tape_stride.
int oracle_dmc_sweeps(oracle_system* h, int n_sweeps, double tau, const double* tape,
                      int tape_stride, double* rr_proposed, double* rr_accepted,
                      long long* n_accepted) {
    return guarded([&] {
        Sys& s = *S(h);
        const std::size_t nw = s.nw, np = s.np, stride = static_cast<std::size_t>(tape_stride);
        long long acc = 0;
        std::vector<double> m(nw), scale(nw), g, l;
        for (int sw = 0; sw < n_sweeps; ++sw) {
            const double* randoms = tape + static_cast<std::size_t>(sw) * np * 3 * stride;
            for (std::size_t p = 0; p < np; ++p) {
                for (auto& c : s.comps) c->refresh_particle(s, p);
                totals(s, g, l, nullptr, nullptr);
                const double tm = tau * s.inv_mass[p];
                const double sigma = std::sqrt(tm);
                const double* gauss = randoms + p * 3 * stride;
                const double* uni = gauss + 2 * stride;
                auto umrigar = [](double t, double v2) {
                    if (v2 < 1e-10) return t;
                    return (-1.0 + std::sqrt(1.0 + 2.0 * t * v2)) / v2;
                };
                for (std::size_t w = 0; w < nw; ++w) {
                    double v2 = 0.0;
                    for (std::size_t d = 0; d < 2; ++d) {
                        const double gg = g[(p * 2 + d) * nw + w];
                        v2 += gg * gg;
                    }
                    scale[w] = umrigar(tm, v2);
                }
                for (std::size_t d = 0; d < 2; ++d)
                    for (std::size_t w = 0; w < nw; ++w)
                        s.new_pos[d * nw + w] = s.P(p, d, w) + scale[w] * g[(p * 2 + d) * nw + w] +
                                                sigma * gauss[d * stride + w];
                ratio_all(s, p);
                for (std::size_t w = 0; w < nw; ++w) {
                    if (!s.has_phase && std::cos(s.phase_diffs[w]) < std::numeric_limits<double>::epsilon()) {
                        m[w] = 0.0;
                        continue;
                    }
                    double vn = 0.0;
                    for (std::size_t d = 0; d < 2; ++d) vn += s.new_grad[d * nw + w] * s.new_grad[d * nw + w];
                    const double sb = umrigar(tm, vn);
                    double lc = 0.0;
                    for (std::size_t d = 0; d < 2; ++d) {
                        const double delta = s.new_pos[d * nw + w] - s.P(p, d, w);
                        const double df = scale[w] * g[(p * 2 + d) * nw + w];
                        const double db = sb * s.new_grad[d * nw + w];
                        lc += tm <= 0.0 ? 0.0 : (-delta * (df + db) + 0.5 * (df * df - db * db)) / tm;
                    }
                    const double la = 2.0 * s.log_ratios[w] + lc;
                    const double pa = la >= 0.0 ? 1.0 : std::exp(la);
                    m[w] = (la >= 0.0 || uni[w] < pa) ? 1.0 : 0.0;
                    if (m[w] != 0.0) ++acc;
                }
                for (std::size_t w = 0; w < nw; ++w) {
                    double rr = 0.0;
                    for (std::size_t d = 0; d < 2; ++d) {
                        const double dr = s.new_pos[d * nw + w] - s.P(p, d, w);
                        rr += tm * dr * dr;
                    }
                    (rr_proposed ? rr_proposed : s.rr_p.data())[w] += rr;
                    if (m[w] != 0.0) (rr_accepted ? rr_accepted : s.rr_a.data())[w] += rr;
                }
                accept_all(s, p, m);
            }
        }
        *n_accepted = acc;
    });
}

// WalkerSet::apply_permutation + permute_caches (walker_set.cpp:109-178,
// wavefunction_component.cpp:18-48).
int oracle_permute(oracle_system* h, int new_nw, const int* src) {
    return guarded([&] {
        Sys& s = *S(h);
        const std::size_t old_nw = s.nw;
        std::vector<std::size_t> sv(static_cast<std::size_t>(new_nw));
        for (int j = 0; j < new_nw; ++j) {
            if (src[j] < 0 || static_cast<std::size_t>(src[j]) >= old_nw)
                throw std::invalid_argument("oracle: permutation source out of range");
            sv[j] = static_cast<std::size_t>(src[j]);
        }
        Component::gather(s.pos, sv, old_nw);
        Component::gather(s.dist, sv, old_nw);
        for (auto* v : {&s.e_loc, &s.e_loc_im, &s.weights, &s.e_prev, &s.rr_p, &s.rr_a})
            Component::gather(*v, sv, old_nw);
        {
            std::vector<long long> c(sv.size());
            for (std::size_t j = 0; j < sv.size(); ++j) c[j] = s.n_cut[sv[j]];
            s.n_cut = c;
        }
        for (auto& c : s.comps) c->permute(sv, old_nw);
        s.nw = sv.size();
        const std::vector<double> pos = s.pos, dist = s.dist, el = s.e_loc, eli = s.e_loc_im,
                                  wt = s.weights, ep = s.e_prev, rp = s.rr_p, ra = s.rr_a;
        const std::vector<long long> nc = s.n_cut;
        s.alloc();
        s.pos = pos;
        s.dist = dist;
        s.e_loc = el;
        s.e_loc_im = eli;
        s.weights = wt;
        s.e_prev = ep;
        s.rr_p = rp;
        s.rr_a = ra;
        s.n_cut = nc;
    });
}

int oracle_refresh_inverse(oracle_system* h, int w_begin, int w_end) {
    return guarded([&] {
        Sys& s = *S(h);
        const long long nw = static_cast<long long>(s.nw);
        const long long w1 = w_end < 0 ? nw : w_end;
        if (w_begin < 0 || w_begin > w1 || w1 > nw) throw std::invalid_argument("walker range outside the batch");
        for (auto& c : s.comps) c->refresh_inverse(s, static_cast<std::size_t>(w_begin), static_cast<std::size_t>(w1));
    });
}

// The oracle throws at the singular rebuild itself (slater_determinant.cpp:
// 927-928), so nothing is pending here.
int oracle_check_status(oracle_system* h) {
    return guarded([&] { S(h); });
}

int oracle_get_num_walkers(oracle_system* h, int* nw) {
    return guarded([&] { *nw = static_cast<int>(S(h)->nw); });
}

// ForwardWalking's snapshot ring (forward_walking.cpp:99-108): the
// population's positions at a step, [slot][p][d][nw of that step], and
// ConfigurationBatch::load of the kept ancestors (forward_walking.cpp:148-150)
// into a measurement batch, evaluated with the descendant weights.
struct oracle_snapshots {
    Sys* pop = nullptr;
    std::vector<std::vector<double>> pos;
    std::vector<long long> nw;
};

int oracle_snapshots_create(oracle_system* population, int n_slots, oracle_snapshots** out) {
    return guarded([&] {
        if (n_slots < 1) throw std::invalid_argument("oracle: a snapshot ring needs at least one slot");
        auto* r = new oracle_snapshots;
        r->pop = S(population);
        r->pos.resize(n_slots);
        r->nw.assign(n_slots, -1);
        *out = r;
    });
}

int oracle_snapshots_destroy(oracle_snapshots* r) {
    delete r;
    return 0;
}

int oracle_snapshots_record(oracle_snapshots* r, int slot) {
    return guarded([&] {
        if (slot < 0 || static_cast<std::size_t>(slot) >= r->pos.size())
            throw std::invalid_argument("oracle: snapshot slot out of range");
        r->pos[slot] = r->pop->pos;
        r->nw[slot] = static_cast<long long>(r->pop->nw);
    });
}

int oracle_snapshots_load(oracle_snapshots* r, int slot, oracle_system* batch, const int* kept,
                          const double* weights, int n_kept) {
    return guarded([&] {
        if (slot < 0 || static_cast<std::size_t>(slot) >= r->pos.size() || r->nw[slot] < 0)
            throw std::invalid_argument("oracle: snapshot slot never recorded");
        Sys& b = *S(batch);
        if (b.np != r->pop->np) throw std::invalid_argument("oracle: the batch has another particle count");
        if (n_kept < 1) throw std::invalid_argument("oracle: nothing to load");
        const std::vector<double>& src = r->pos[slot];
        const std::size_t snw = static_cast<std::size_t>(r->nw[slot]), n = static_cast<std::size_t>(n_kept);
        for (std::size_t c = 0; c < n; ++c)
            if (kept[c] < 0 || static_cast<std::size_t>(kept[c]) >= snw)
                throw std::invalid_argument("oracle: kept index outside the snapshot");
        b.nw = n;
        b.alloc();
        for (std::size_t k = 0; k < b.np * b.dim; ++k)
            for (std::size_t c = 0; c < n; ++c) b.pos[k * n + c] = src[k * snw + static_cast<std::size_t>(kept[c])];
        b.build_dist();
        std::copy(weights, weights + n, b.weights.begin());
        for (auto& c : b.comps) c->evaluate(b);
    });
}

// DMC::advance_measure set-up (dmc.cpp:349-354).
int oracle_dmc_begin(oracle_system* h) {
    return guarded([&] {
        Sys& s = *S(h);
        local_energy(s);
        for (std::size_t w = 0; w < s.nw; ++w) {
            s.rr_p[w] = 0.0;
            s.rr_a[w] = 0.0;
            s.weights[w] = 1.0;
            s.e_prev[w] = s.e_loc[w];
            s.n_cut[w] = 0;
        }
    });
}

// H after a sweep and the bare branch factor (dmc.cpp:383-384, 404-408;
// PopulationControl::branch_weight and is_cut, population_control.cpp:58-67,
// population_control.h:189-191).
int oracle_dmc_weights(oracle_system* h, double e_trial, double e_ref, double e_cut, double e_max,
                       double tau) {
    return guarded([&] {
        Sys& s = *S(h);
        local_energy(s);
        for (std::size_t w = 0; w < s.nw; ++w) {
            const double e_new = s.e_loc[w], e_old = s.e_prev[w];
            double half_tau = 0.5 * tau;
            const double x = std::max(e_ref - e_new, e_ref - e_old);
            if (x > e_max)
                half_tau = 0.0;
            else if (x > e_cut)
                half_tau *= 1.0 - (x - e_cut) / (e_max - e_cut);
            s.weights[w] *= std::exp(half_tau * (2.0 * e_trial - e_new - e_old));
            if (e_ref - e_new > e_cut || e_ref - e_old > e_cut) ++s.n_cut[w];
            s.e_prev[w] = e_new;
        }
    });
}

// (sum w E_L, sum w) in walker order (dmc.cpp:412-418).
int oracle_dmc_sums(oracle_system* h, double* energy_sum, double* weight_sum, long long* n_cut) {
    return guarded([&] {
        Sys& s = *S(h);
        double es = 0.0, ws = 0.0;
        long long nc = 0;
        for (std::size_t w = 0; w < s.nw; ++w) {
            es += s.weights[w] * s.e_loc[w];
            ws += s.weights[w];
            nc += s.n_cut[w];
        }
        if (energy_sum) *energy_sum = es;
        if (weight_sum) *weight_sum = ws;
        if (n_cut) *n_cut = nc;
    });
}

int oracle_get_weights(oracle_system* h, double* w) {
    return guarded([&] { std::copy(S(h)->weights.begin(), S(h)->weights.end(), w); });
}

int oracle_set_weights(oracle_system* h, const double* w) {
    return guarded([&] {
        Sys& s = *S(h);
        for (std::size_t i = 0; i < s.nw; ++i)
            if (!(w[i] >= 0.0) || !std::isfinite(w[i]))
                throw std::invalid_argument("oracle: weights must be finite and non-negative");
        std::copy(w, w + s.nw, s.weights.begin());
    });
}

// ── Estimators ─────────────────────────────────────────────────────────────
// PairCorrelation (src/estimators/pair_correlation.cpp:16-64, 66-105): one
// function per unordered group pair a <= b, pid = a G - a(a-1)/2 + (b-a); every
// pair i < j of the distance table with r < rmax adds its walker's weight to
// bin int(r / delta); the block's weight sum adds each call's walker sum.
int oracle_add_gofr(oracle_system* h, double rmax, int n_bins) {
    return guarded([&] {
        Sys& s = *S(h);
        if (!(rmax > 0.0) || n_bins <= 0) throw std::invalid_argument("oracle: gofr needs rmax > 0, n_bins > 0");
        if (s.cell.periodic && rmax > s.cell.ws)
            throw std::invalid_argument("oracle: gofr rmax exceeds the Wigner-Seitz inradius");
        const std::size_t G = s.groups.size();
        auto& g = s.gofr;
        g.on = true;
        g.rmax = rmax;
        g.n_bins = static_cast<std::size_t>(n_bins);
        g.inv_delta = static_cast<double>(n_bins) / rmax;
        g.n_func = G * (G + 1) / 2;
        g.hist.assign(g.n_func * g.n_bins, 0.0);
    });
}

// StructureFactor (src/estimators/structure_factor.cpp:16-54): the half space
// n1 > 0, or n1 = 0 and n2 > 0, of the reciprocal lattice of the reduced
// basis, |k| <= kmax, sorted by |k| (std::sort, as the reference), shells of
// equal |k| within 1e-9 min|b|.
int oracle_add_sofk(oracle_system* h, double kmax, int* n_k) {
    return guarded([&] {
        Sys& s = *S(h);
        if (!s.cell.periodic) throw std::invalid_argument("oracle: sofk needs a periodic cell");
        const double* b = s.cell.b;
        const double b1 = std::hypot(b[0], b[1]), b2 = std::hypot(b[2], b[3]);
        const int nmax = static_cast<int>(std::ceil(kmax / std::min(b1, b2))) + 1;
        struct K { int n1, n2; double kx, ky, k; };
        std::vector<K> ks;
        for (int i = 0; i <= nmax; ++i)
            for (int j = -nmax; j <= nmax; ++j) {
                if (i == 0 && j <= 0) continue;
                const double kx = i * b[0] + j * b[2], ky = i * b[1] + j * b[3];
                const double k = std::hypot(kx, ky);
                if (k <= kmax) ks.push_back({i, j, kx, ky, k});
            }
        if (ks.empty()) throw std::invalid_argument("oracle: sofk kmax below the shortest reciprocal vector");
        std::sort(ks.begin(), ks.end(), [](const K& x, const K& y) { return x.k < y.k; });
        auto& f = s.sofk;
        f = Sys::Sofk{};
        f.on = true;
        const double tol = 1e-9 * std::min(b1, b2);
        double last = 0.0;
        int shell = -1;
        for (const K& q : ks) {
            if (shell < 0 || q.k - last > tol) {
                ++shell;
                last = q.k;
            }
            f.shell.push_back(shell);
            f.kx.push_back(q.kx);
            f.ky.push_back(q.ky);
            f.n1.push_back(q.n1);
            f.n2.push_back(q.n2);
            f.m1 = std::max(f.m1, std::abs(q.n1));
            f.m2 = std::max(f.m2, std::abs(q.n2));
        }
        const std::size_t nk = f.kx.size(), G = s.groups.size();
        f.acc_s.assign(G * (G + 1) / 2 * nk, 0.0);
        f.acc_re.assign(G * nk, 0.0);
        f.acc_im.assign(G * nk, 0.0);
        *n_k = static_cast<int>(nk);
    });
}

int oracle_get_sofk_vectors(oracle_system* h, double* kx, double* ky, int* shell) {
    return guarded([&] {
        const auto& f = S(h)->sofk;
        if (!f.on) throw std::invalid_argument("oracle: no sofk estimator");
        std::copy(f.kx.begin(), f.kx.end(), kx);
        std::copy(f.ky.begin(), f.ky.end(), ky);
        std::copy(f.shell.begin(), f.shell.end(), shell);
    });
}

// Density (src/estimators/density.cpp:16-64, 66-117): one n_bins x n_bins
// histogram per group, C row-major; a periodic cell bins the fractional
// coordinates f + 1/2 of the wrapped position, an open cell the window
// [-rmax, rmax)^2 (particles outside it drop); each call's walker weights add
// to the block's weight sum.
int oracle_add_density(oracle_system* h, double rmax, int n_bins, int per_particle) {
    return guarded([&] {
        Sys& s = *S(h);
        if (n_bins <= 0) throw std::invalid_argument("oracle: density needs n_bins > 0");
        if (!s.cell.periodic && !(rmax > 0.0)) throw std::invalid_argument("oracle: density needs rmax > 0");
        if (per_particle)
            for (std::size_t g = 0; g < s.groups.size(); ++g)
                if (s.groups[g].first != (g == 0 ? 0 : s.groups[g - 1].last))
                    throw std::invalid_argument("oracle: per-particle density needs the groups in ascending order");
        auto& d = s.dens;
        d.on = true;
        d.rmax = rmax;
        d.n_bins = static_cast<std::size_t>(n_bins);
        d.inv_delta = static_cast<double>(n_bins) / (2.0 * rmax);
        d.n_cells = d.n_bins * d.n_bins;
        d.per_particle = per_particle != 0;
        d.hist.assign((d.per_particle ? s.np : s.groups.size()) * d.n_cells, 0.0);
    });
}

int oracle_estimators_read_density(oracle_system* h, double* density, double* wsum) {
    return guarded([&] {
        const Sys& s = *S(h);
        if (!s.dens.on) throw std::invalid_argument("oracle: no density estimator");
        if (density) std::copy(s.dens.hist.begin(), s.dens.hist.end(), density);
        if (wsum) *wsum = s.dens.wsum;
    });
}

int oracle_estimators_start_block(oracle_system* h) {
    return guarded([&] {
        Sys& s = *S(h);
        std::fill(s.dens.hist.begin(), s.dens.hist.end(), 0.0);
        s.dens.wsum = 0.0;
        std::fill(s.gofr.hist.begin(), s.gofr.hist.end(), 0.0);
        s.gofr.wsum = 0.0;
        std::fill(s.sofk.acc_s.begin(), s.sofk.acc_s.end(), 0.0);
        std::fill(s.sofk.acc_re.begin(), s.sofk.acc_re.end(), 0.0);
        std::fill(s.sofk.acc_im.begin(), s.sofk.acc_im.end(), 0.0);
        s.sofk.wsum = 0.0;
    });
}

namespace {
// ReciprocalPhases::ladder (src/numerics/reciprocal_phases.h): e^{i n theta}
// for n in [-m, m] by repeated multiplication.
void phase_ladder(double theta, int m, std::vector<double>& re, std::vector<double>& im) {
    re.assign(2 * m + 1, 0.0);
    im.assign(2 * m + 1, 0.0);
    double* r = re.data() + m;
    double* i = im.data() + m;
    const double c = std::cos(theta), sn = std::sin(theta);
    r[0] = 1.0;
    i[0] = 0.0;
    for (int n = 1; n <= m; ++n) {
        r[n] = r[n - 1] * c - i[n - 1] * sn;
        i[n] = r[n - 1] * sn + i[n - 1] * c;
        r[-n] = r[n];
        i[-n] = -i[n];
    }
}
}  // namespace

// One measurement of the current walkers (Estimator::accumulate(wset)):
// weights 1 (use_weights = 0) or the walkers' DMC weights.
int oracle_estimators_accumulate(oracle_system* h, int use_weights) {
    return guarded([&] {
        Sys& s = *S(h);
        const std::size_t nw = s.nw, np = s.np, G = s.groups.size();
        std::vector<double> wt(nw, 1.0);
        if (use_weights) std::copy(s.weights.begin(), s.weights.begin() + nw, wt.begin());
        std::vector<std::size_t> group_of(np, 0);
        for (std::size_t g = 0; g < G; ++g)
            for (std::size_t p = s.groups[g].first; p < s.groups[g].last; ++p) group_of[p] = g;
        auto pair_id = [&](std::size_t a, std::size_t b) {
            if (b < a) std::swap(a, b);
            return a * G - a * (a - 1) / 2 + (b - a);
        };
        if (s.gofr.on) {   // pair_correlation.cpp:74-105
            auto& g = s.gofr;
            for (std::size_t i = 0; i < np; ++i)
                for (std::size_t j = i + 1; j < np; ++j) {
                    double* hst = g.hist.data() + pair_id(group_of[i], group_of[j]) * g.n_bins;
                    for (std::size_t w = 0; w < nw; ++w) {
                        const double r = s.D(i, j, w);
                        if (r < g.rmax) {
                            const std::size_t bin = static_cast<std::size_t>(r * g.inv_delta);
                            if (bin < g.n_bins) hst[bin] += wt[w];
                        }
                    }
                }
            double ws = 0.0;
            for (std::size_t w = 0; w < nw; ++w) ws += wt[w];
            g.wsum += ws;
        }
        if (s.dens.on) {   // density.cpp:66-117
            auto& d = s.dens;
            const double two_rmax = 2.0 * d.rmax;
            // per_particle: each particle is its own range (density.cpp:42-51)
            for (std::size_t gg = 0; gg < G; ++gg) {
                for (std::size_t p = s.groups[gg].first; p < s.groups[gg].last; ++p) {
                    double* hst = d.hist.data() + (d.per_particle ? p : gg) * d.n_cells;
                    for (std::size_t w = 0; w < nw; ++w) {
                        double r[2] = {s.P(p, 0, w), s.P(p, 1, w)};
                        std::size_t cell = 0;
                        if (s.cell.periodic) {
                            s.cell.wrap(r[0], r[1]);
                            for (std::size_t i = 0; i < 2; ++i) {
                                double f = 0.0;
                                f += s.cell.ainv[i * 2 + 0] * r[0];
                                f += s.cell.ainv[i * 2 + 1] * r[1];
                                const double x = f + 0.5;
                                std::size_t bin = static_cast<std::size_t>(x * static_cast<double>(d.n_bins));
                                if (bin >= d.n_bins) bin = d.n_bins - 1;
                                cell = cell * d.n_bins + bin;
                            }
                        } else {
                            bool in_box = true;
                            for (std::size_t i = 0; i < 2; ++i) {
                                const double x = r[i] + d.rmax;
                                if (x < 0.0 || x >= two_rmax) {
                                    in_box = false;
                                    break;
                                }
                                std::size_t bin = static_cast<std::size_t>(x * d.inv_delta);
                                if (bin >= d.n_bins) bin = d.n_bins - 1;
                                cell = cell * d.n_bins + bin;
                            }
                            if (!in_box) continue;
                        }
                        hst[cell] += wt[w];
                    }
                }
            }
            double ws = 0.0;
            for (std::size_t w = 0; w < nw; ++w) ws += wt[w];
            d.wsum += ws;
        }
        if (s.sofk.on) {   // structure_factor.cpp:62-105
            auto& f = s.sofk;
            const std::size_t nk = f.kx.size();
            const double* b = s.cell.b;
            std::vector<double> rre(G * nk), rim(G * nk), c(nk), sn(nk), e1r, e1i, e2r, e2i;
            for (std::size_t w = 0; w < nw; ++w) {
                for (std::size_t gg = 0; gg < G; ++gg) {
                    double* re = rre.data() + gg * nk;
                    double* im = rim.data() + gg * nk;
                    std::fill_n(re, nk, 0.0);
                    std::fill_n(im, nk, 0.0);
                    for (std::size_t i = s.groups[gg].first; i < s.groups[gg].last; ++i) {
                        const double x = s.P(i, 0, w), y = s.P(i, 1, w);
                        phase_ladder(b[0] * x + b[1] * y, f.m1, e1r, e1i);
                        phase_ladder(b[2] * x + b[3] * y, f.m2, e2r, e2i);
                        for (std::size_t k = 0; k < nk; ++k) {
                            const double ar = e1r[f.m1 + f.n1[k]], ai = e1i[f.m1 + f.n1[k]];
                            const double br = e2r[f.m2 + f.n2[k]], bi = e2i[f.m2 + f.n2[k]];
                            c[k] = ar * br - ai * bi;
                            sn[k] = ar * bi + ai * br;
                        }
                        for (std::size_t k = 0; k < nk; ++k) {
                            re[k] += c[k];
                            im[k] += sn[k];
                        }
                    }
                    for (std::size_t k = 0; k < nk; ++k) {
                        f.acc_re[gg * nk + k] += wt[w] * re[k];
                        f.acc_im[gg * nk + k] += wt[w] * im[k];
                    }
                }
                std::size_t pid = 0;
                for (std::size_t a = 0; a < G; ++a)
                    for (std::size_t bb = a; bb < G; ++bb, ++pid)
                        for (std::size_t k = 0; k < nk; ++k)
                            f.acc_s[pid * nk + k] += wt[w] * (rre[a * nk + k] * rre[bb * nk + k] +
                                                               rim[a * nk + k] * rim[bb * nk + k]);
                f.wsum += wt[w];
            }
        }
    });
}

// The block's running sums; wsum[0] the pair correlation's weight sum,
// wsum[1] the structure factor's (each summed in its reference's order).
int oracle_estimators_read(oracle_system* h, double* gofr, double* sofk, double* rho_re, double* rho_im,
                           double* wsum) {
    return guarded([&] {
        const Sys& s = *S(h);
        if (gofr) std::copy(s.gofr.hist.begin(), s.gofr.hist.end(), gofr);
        if (sofk) std::copy(s.sofk.acc_s.begin(), s.sofk.acc_s.end(), sofk);
        if (rho_re) std::copy(s.sofk.acc_re.begin(), s.sofk.acc_re.end(), rho_re);
        if (rho_im) std::copy(s.sofk.acc_im.begin(), s.sofk.acc_im.end(), rho_im);
        if (wsum) {
            wsum[0] = s.gofr.wsum;
            wsum[1] = s.sofk.wsum;
        }
    });
}

// WalkerSet::branch (walker_set.cpp:187-239) with the branch stream's state
// held by the caller (XoshiroRandom::uniform, xoshiro_random.h:90-92).
int oracle_branch(oracle_system* h, uint64_t* rng_state, int target, int* new_nw, int* src_out) {
    return guarded([&] {
        Sys& s = *S(h);
        if (target < 1) throw std::invalid_argument("oracle: target population must be positive");
        Xoshiro eng(static_cast<const uint64_t*>(rng_state));
        auto uniform = [&] { return static_cast<double>(static_cast<int64_t>(eng.next() >> 11)) * 0x1.0p-53; };
        const std::size_t n = s.nw, tgt = static_cast<std::size_t>(target);
        std::vector<std::size_t> src;
        for (std::size_t w = 0; w < n; ++w) {
            const int m = static_cast<int>(s.weights[w] + uniform());
            for (int c = 0; c < m; ++c) src.push_back(w);
        }
        const std::size_t n_max = 2 * tgt + 1, n_min = tgt / 5 + 1;
        if (src.empty()) {
            std::size_t best = 0;
            for (std::size_t w = 1; w < n; ++w)
                if (s.weights[w] > s.weights[best]) best = w;
            src.push_back(best);
        }
        while (src.size() > n_max) {
            const std::size_t i = static_cast<std::size_t>(uniform() * static_cast<double>(src.size()));
            src[i] = src.back();
            src.pop_back();
        }
        while (src.size() < n_min) {
            const std::size_t i = static_cast<std::size_t>(uniform() * static_cast<double>(src.size()));
            src.push_back(src[i]);
        }
        std::copy(eng.s, eng.s + 4, rng_state);
        std::vector<int> si(src.begin(), src.end());
        if (oracle_permute(h, static_cast<int>(si.size()), si.data()) != 0) throw std::runtime_error(g_error);
        std::fill(s.weights.begin(), s.weights.end(), 1.0);
        *new_nw = static_cast<int>(si.size());
        if (src_out) std::copy(si.begin(), si.end(), src_out);
    });
}

}  // extern "C"
