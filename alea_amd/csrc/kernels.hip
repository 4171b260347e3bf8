// HIP kernels for the batched VMC/DMC walker sweep on CDNA4 (gfx950).
//
// One 64-lane wavefront owns one walker (workgroup = one wave, so a
// __syncthreads() is a wave-local fence and costs nothing).  Inside a wave the
// lanes split whatever the current step iterates over: partner particles for
// the pair Jastrow, reciprocal vectors for the k-space Jastrow, orbitals for a
// determinant ratio, determinant columns for the Sherman-Morrison update.
// Per-walker state lives in HBM in walker-major records (model.h).
//
// Reference algorithms followed (file:line):
//   minimum image            src/population/simulation_cell.h:109-261
//   Padé / image-sum split   src/wavefunction/pade_jastrow.cpp:26-56,
//                            src/wavefunction/periodic_pair_sum.cpp:66-83,
//                            src/numerics/long_range_spline.h:41-110
//   Drummond                 src/wavefunction/drummond_jastrow.cpp:27-80
//   k-space pair sum         src/wavefunction/kspace_pair_sum.cpp:91-301
//   Nosanow                  src/wavefunction/nosanow_jastrow.cpp:58-174
//   orbitals                 src/wavefunction/orbital_set.h:164-224,427-474,
//                            src/wavefunction/complex_orbital_set.h:91-124
//   Slater ratio / SM / GJ   src/wavefunction/slater_determinant.cpp:395-437,
//                            744-858, 905-993
//   kinetic                  src/hamiltonian/kinetic.cpp:45-110
//   Ewald                    src/hamiltonian/ewald_term.cpp:368-612
//   VMC / DMC sweeps         src/qmc/vmc.cpp:82-155, src/qmc/dmc.cpp:203-337
//   RNG                      src/numerics/xoshiro_random.h:23-193
//
// Differences from the reference, by design:
//   * per-particle gradients are not carried incrementally: a move computes
//     the moved particle's old gradient (DMC drift) from the current state, and
//     measurement recomputes every particle's gradient and Laplacian.  The
//     O(N n_k) k-space "seer" update per accepted move disappears.
//   * the complex determinant is updated by Sherman-Morrison with the real
//     determinant's rebuild triggers, instead of a full re-inversion per
//     accepted move (complex_slater_determinant.cpp:245-253).
//   * the condition estimate uses the 1-norm of A (free in the column-per-lane
//     update pass) with the infinity-norm row sums of D.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "model.h"
#include "trig.h"

namespace alea_amd {

__constant__ uint64_t c_zig_k[256];
__constant__ double c_zig_w[256];
__constant__ double c_zig_f[256];
__constant__ double c_zig_r;

namespace {

constexpr int W = 64;   // wavefront

// Wavefront reductions (called with all 64 lanes active).  Within each row
// of 16 lanes, DPP moves (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror,
// row_mirror) combine partners in registers, leaving every lane of the row
// with the row's total; the four row totals are then read from lanes 0, 16,
// 32, 48 into scalar registers and combined in a fixed order, so every lane
// gets the same bits.
template <int CTRL>
__device__ inline double dpp_mov(double v) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ inline double lane_value(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}
// A row broadcast (row_bcast:15 into rows 1 and 3, row_bcast:31 into rows 2
// and 3); the rows it does not write hold no defined value, and only lane
// 63's sum is read afterwards.
template <int CTRL, int ROWS>
__device__ inline double dpp_bcast(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, ROWS, 0xF, false);
    return __hiloint2double(hi, lo);
}
template <int N>
__device__ inline void wsum_n(double (&v)[N]) {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += dpp_mov<0xB1>(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += dpp_mov<0x4E>(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += dpp_mov<0x141>(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += dpp_mov<0x140>(v[i]);
    // rows r0..r3 -> lane 63 holds (r3 + r2) + (r1 + r0), which IEEE
    // addition makes equal to (r0 + r1) + (r2 + r3) bit for bit
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += dpp_bcast<0x142, 0xA>(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += dpp_bcast<0x143, 0xC>(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = lane_value(v[i], 63);
}
__device__ inline double wsum(double v) {
    double a[1] = {v};
    wsum_n(a);
    return a[0];
}
__device__ inline double wmax(double v) {
    v = fmax(v, dpp_mov<0xB1>(v));
    v = fmax(v, dpp_mov<0x4E>(v));
    v = fmax(v, dpp_mov<0x141>(v));
    v = fmax(v, dpp_mov<0x140>(v));
    return fmax(fmax(lane_value(v, 0), lane_value(v, 16)), fmax(lane_value(v, 32), lane_value(v, 48)));
}

// ── cell ────────────────────────────────────────────────────────────────────
__device__ inline void min_image(const DevCell& c, double& d0, double& d1) {
#pragma clang fp contract(off)
    if (!c.periodic) return;
    if (c.ortho) {
        d0 = d0 - c.box[0] * rint(d0 * c.inv_box[0]);
        d1 = d1 - c.box[1] * rint(d1 * c.inv_box[1]);
        return;
    }
    double s0 = 0.0;
    s0 += c.ainv[0] * d0;
    s0 += c.ainv[1] * d1;
    double s1 = 0.0;
    s1 += c.ainv[2] * d0;
    s1 += c.ainv[3] * d1;
    const double f0 = rint(s0), f1 = rint(s1);
    double t0 = 0.0;
    t0 += f0 * c.a[0];
    t0 += f1 * c.a[2];
    double t1 = 0.0;
    t1 += f0 * c.a[1];
    t1 += f1 * c.a[3];
    d0 -= t0;
    d1 -= t1;
    double b0 = d0, b1 = d1, b2 = 0.0;
    b2 += d0 * d0;
    b2 += d1 * d1;
    for (int k = 1; k < 9; ++k) {
        const double y0 = d0 + c.cand[2 * k];
        double r2 = 0.0;
        r2 += y0 * y0;
        const double y1 = d1 + c.cand[2 * k + 1];
        r2 += y1 * y1;
        const double take = (r2 < b2) ? 1.0 : 0.0;
        b2 += take * (r2 - b2);
        b0 += take * (y0 - b0);
        b1 += take * (y1 - b1);
    }
    d0 = b0;
    d1 = b1;
}

__device__ inline void wrap(const DevCell& c, double& x0, double& x1) {
#pragma clang fp contract(off)
    if (!c.periodic) return;
    if (c.ortho) {
        x0 -= c.box[0] * rint(x0 * c.inv_box[0]);
        x1 -= c.box[1] * rint(x1 * c.inv_box[1]);
        return;
    }
    double f[2];
    for (int i = 0; i < 2; ++i) {
        double s = 0.0;
        s += c.ainv[i * 2 + 0] * x0;
        s += c.ainv[i * 2 + 1] * x1;
        f[i] = rint(s);
    }
    double s0 = 0.0, s1 = 0.0;
    s0 += f[0] * c.a[0];
    s0 += f[1] * c.a[2];
    s1 += f[0] * c.a[1];
    s1 += f[1] * c.a[3];
    x0 -= s0;
    x1 -= s1;
}

// sqrt of x >= 0 for a normal or zero x: the hardware reciprocal square
// root refined by Goldschmidt's iteration (three FMA corrections), the
// sequence of the library's correctly rounded sqrt without its rescaling of
// inputs below 2^-767 (an interparticle r^2 never is one), so equal to it
// bit for bit on every normal input.
__device__ inline double sqrt_normal(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return x > 0.0 ? g : 0.0;
}

__device__ inline double dist(const DevCell& c, double& d0, double& d1) {
    min_image(c, d0, d1);
    return sqrt_normal(d0 * d0 + d1 * d1);
}

// ── radial functions ───────────────────────────────────────────────────────
__device__ inline void spline_eval(const double* sc, int n_knots, double r_cut, double inv_delta,
                                   double r, double& f, double& df, double& d2f) {
    if (r >= r_cut) { f = df = d2f = 0.0; return; }
    const double t = r * inv_delta;
    int i = static_cast<int>(t);
    if (i > n_knots - 2) i = n_knots - 2;
    const double x = t - i, y = 1.0 - x;
    const double* a = sc + 3 * i;
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

// 1/q from the hardware reciprocal and two Newton steps (error below 1 ulp),
// in place of the IEEE division sequence.
__device__ inline double rcp_refined(double q) {
    double r = __builtin_amdgcn_rcp(q);
    double e = __builtin_fma(-q, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-q, r, 1.0);
    return __builtin_fma(r, e, r);
}

// The spline of spline_eval as monomials per panel (DevRadial::mono, stored
// coefficient-major): Horner in the panel coordinate for the value and both
// derivatives.
__device__ inline void spline_eval_mono(const double* mono, int n_knots, double r_cut, double inv_delta,
                                        double r, double& f, double& df, double& d2f) {
    if (r >= r_cut) { f = df = d2f = 0.0; return; }
    const double t = r * inv_delta;
    int i = static_cast<int>(t);
    if (i > n_knots - 2) i = n_knots - 2;
    const double x = t - i;
    // coefficient-major: c_k of panel i at k * (n_knots - 1) + i, so lanes
    // reading different panels touch different LDS banks
    const int np_ = n_knots - 1;
    const double c0 = mono[i], c1 = mono[np_ + i], c2 = mono[2 * np_ + i];
    const double c3 = mono[3 * np_ + i], c4 = mono[4 * np_ + i], c5 = mono[5 * np_ + i];
    f = ((((c5 * x + c4) * x + c3) * x + c2) * x + c1) * x + c0;
    df = inv_delta * ((((5.0 * c5 * x + 4.0 * c4) * x + 3.0 * c3) * x + 2.0 * c2) * x + c1);
    d2f = inv_delta * inv_delta * (((20.0 * c5 * x + 12.0 * c4) * x + 6.0 * c3) * x + 2.0 * c2);
}

// mono: an LDS copy of the factor's monomial spline (R.mono), or null for
// the knot form in R.spline
__device__ inline void radial_eval(const DevRadial& R, double r, double& u, double& du, double& d2u,
                                   const double* mono) {
    if (R.kind == 2) {   // McMillan: u = -(b/r)^n
        const double x = R.c[0] / r;
        double xn = 1.0;
        for (int i = 0; i < R.num; ++i) xn *= x;
        const double n = static_cast<double>(R.num);
        u = -xn;
        du = n * xn / r;
        d2u = -n * (n + 1.0) * xn / (r * r);
        return;
    }
    if (R.kind == 1) {
        const double t = r - R.L;
        double tCm2 = 1.0;
        for (int i = 0; i < R.C - 2; ++i) tCm2 *= t;
        const double tCm1 = tCm2 * t, tC = tCm1 * t;
        const double g = tC, dg = static_cast<double>(R.C) * tCm1,
                     d2g = static_cast<double>(R.C) * (R.C - 1) * tCm2;
        double P = R.c[0] + R.alpha1 * r, dP = R.alpha1, d2P = 0.0;
        double r_lm2 = 1.0, r_lm1 = r, r_l = r * r;
        for (int k = 1; k < R.num; ++k) {
            const int l = k + 1;
            P += R.c[k] * r_l;
            dP += R.c[k] * l * r_lm1;
            d2P += R.c[k] * l * (l - 1) * r_lm2;
            r_lm2 = r_lm1;
            r_lm1 = r_l;
            r_l *= r;
        }
        const double m = (r < R.L) ? 1.0 : 0.0;
        u = m * (g * P);
        du = m * (dg * P + g * dP);
        d2u = m * (d2g * P + 2.0 * dg * dP + g * d2P);
        return;
    }
    double p = 0.0, dp = 0.0, d2p = 0.0;
    double q = 1.0, dq = 0.0, d2q = 0.0;
    if (R.num == 1 && R.den == 1) {
        // u = a r / (1 + b r), the cusp Pade of every electron-gas input: the
        // general loops below give the same values
        p = R.c[0] * r;
        dp = R.c[0];
        q += R.c[1] * r;
        dq = R.c[1];
    } else {
        double r1 = 1.0, r2 = 0.0;
        for (int i = 0; i < R.num; ++i) {
            const int k = i + 1;
            p += R.c[i] * r1 * r;
            dp += R.c[i] * k * r1;
            d2p += R.c[i] * k * (k - 1) * r2;
            r2 = r1;
            r1 *= r;
        }
        r1 = 1.0;
        r2 = 0.0;
        for (int i = 0; i < R.den; ++i) {
            const int k = i + 1;
            q += R.c[R.num + i] * r1 * r;
            dq += R.c[R.num + i] * k * r1;
            d2q += R.c[R.num + i] * k * (k - 1) * r2;
            r2 = r1;
            r1 *= r;
        }
    }
    const double iq = rcp_refined(q);
    u = p * iq;
    du = (dp * q - p * dq) * iq * iq;
    d2u = (d2p - (2.0 * dp * dq - 2.0 * p * dq * dq * iq + p * d2q) * iq) * iq;
    if (R.image) {
        if (r >= R.r_cut) { u = du = d2u = 0.0; return; }
        double f, df, d2f;
        if (mono) spline_eval_mono(mono, R.n_knots, R.r_cut, R.inv_delta, r, f, df, d2f);
        else spline_eval(R.spline, R.n_knots, R.r_cut, R.inv_delta, r, f, df, d2f);
        u -= R.u_inf + f;
        du -= df;
        d2u -= d2f;
    }
}
__device__ inline void radial_eval(const DevRadial& R, double r, double& u, double& du, double& d2u) {
    radial_eval(R, r, u, du, d2u, nullptr);
}

// Pair contribution of a radial u to the gradient and per-dimension
// Laplacian of the particle at +delta (pade_jastrow.cpp:268-295).
__device__ inline void pair_terms(double du, double d2u, double r, double d0, double d1,
                                  double& g0, double& g1, double& l0, double& l1) {
    const double invr = 1.0 / r;
    const double o0 = d0 * invr, o1 = d1 * invr;
    g0 = du * o0;
    g1 = du * o1;
    l0 = d2u * o0 * o0 + du * invr * (1.0 - o0 * o0);
    l1 = d2u * o1 * o1 + du * invr * (1.0 - o1 * o1);
}

// ── orbitals (computed in FP64, stored in the determinant precision) ───────
struct Orb {
    double vr, vi;          // value
    double gr[2], gi[2];    // gradient
    double lr[2], li[2];    // per-dimension second derivative
};

__device__ inline void orbital(const DevModel& M, int b, int m, double x, double y, Orb& o) {
    const int kind = M.blk_kind[b];
    const double kx = M.orb_k[b][m][0], ky = M.orb_k[b][m][1];
    if (kind == kGaussian) {
        double d0 = x - kx, d1 = y - ky;
        if (M.cell.periodic) min_image(M.cell, d0, d1);
        const double a = M.orb_alpha[b][m];
        const double phi = exp(-a * (d0 * d0 + d1 * d1));
        o.vr = phi; o.vi = 0.0;
        o.gr[0] = phi * (-2.0 * a * d0); o.gr[1] = phi * (-2.0 * a * d1);
        o.lr[0] = phi * (4.0 * a * a * d0 * d0 - 2.0 * a);
        o.lr[1] = phi * (4.0 * a * a * d1 * d1 - 2.0 * a);
        o.gi[0] = o.gi[1] = o.li[0] = o.li[1] = 0.0;
        return;
    }
    double s, c;
    sincos_bounded(kx * x + ky * y, &s, &c);
    if (kind == kTwisted) {
        o.vr = c; o.vi = s;
        o.gr[0] = -kx * s; o.gi[0] = kx * c;
        o.gr[1] = -ky * s; o.gi[1] = ky * c;
        o.lr[0] = -kx * kx * c; o.li[0] = -kx * kx * s;
        o.lr[1] = -ky * ky * c; o.li[1] = -ky * ky * s;
        return;
    }
    const bool sn = M.orb_sin[b][m] != 0;
    o.vr = sn ? s : c; o.vi = 0.0;
    o.gr[0] = sn ? kx * c : -kx * s;
    o.gr[1] = sn ? ky * c : -ky * s;
    o.lr[0] = sn ? -kx * kx * s : -kx * kx * c;
    o.lr[1] = sn ? -ky * ky * s : -ky * ky * c;
    o.gi[0] = o.gi[1] = o.li[0] = o.li[1] = 0.0;
}

// Value only, in single precision, for the mixed form's ratio: plane-wave
// and twisted orbitals through sincos_bounded_f32 (the phase k.r reduced in
// double); Gaussian orbitals through the double-precision path.
__device__ inline void orbital_value_f32(const DevModel& M, int b, int m, double x, double y, float& vr,
                                         float& vi) {
    const int kind = M.blk_kind[b];
    if (kind == kGaussian) {
        Orb o;
        orbital(M, b, m, x, y, o);
        vr = static_cast<float>(o.vr);
        vi = 0.0f;
        return;
    }
    float s, c;
    sincos_bounded_f32(M.orb_k[b][m][0] * x + M.orb_k[b][m][1] * y, &s, &c);
    if (kind == kTwisted) {
        vr = c;
        vi = s;
    } else {
        vr = M.orb_sin[b][m] != 0 ? s : c;
        vi = 0.0f;
    }
}

// ── xoshiro256++ and the ziggurat, every lane stepping the same state ──────
struct Rng {
    uint64_t s[4];
    __device__ uint64_t next() {
        const uint64_t result = ((s[0] + s[3]) << 23 | (s[0] + s[3]) >> 41) + s[0];
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = (s[3] << 45) | (s[3] >> 19);
        return result;
    }
    __device__ double uniform() {
        return static_cast<double>(static_cast<int64_t>(next() >> 11)) * 0x1.0p-53;
    }
    __device__ double gaussian() {
        // the wedge test rounds as the host's does: no fused multiply-adds
#pragma clang fp contract(off)
        constexpr uint64_t mask = (uint64_t{1} << 52) - 1;
        for (;;) {
            const uint64_t bits = next();
            const int i = static_cast<int>(bits & 0xff);
            const uint64_t sign = (bits & 0x100) << 55;
            const uint64_t position = (bits >> 9) & mask;
            const double x = static_cast<double>(static_cast<int64_t>(position)) * c_zig_w[i];
            if (position < c_zig_k[i]) return __longlong_as_double(__double_as_longlong(x) ^ sign);
            if (i == 0) {
                for (;;) {
                    const double xi = -log1p(-uniform()) / c_zig_r;
                    const double y = -log1p(-uniform());
                    if (y + y > xi * xi)
                        return __longlong_as_double(__double_as_longlong(c_zig_r + xi) ^ sign);
                }
            }
            if (c_zig_f[i] + uniform() * (c_zig_f[i - 1] - c_zig_f[i]) < exp(-0.5 * x * x))
                return __longlong_as_double(__double_as_longlong(x) ^ sign);
        }
    }
};

// ── complex helpers in the determinant precision ───────────────────────────
template <typename T>
struct C2 {
    T re, im;
};
template <typename T>
__device__ inline C2<T> cmul(C2<T> a, C2<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ inline float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
// Entries m and m+1 (m even) of a complex LDS vector, in 16-byte reads: one
// read for two single-precision entries, one per double-precision entry.
// The vector starts 16-byte aligned.
template <typename T>
__device__ inline void lds_pair(const C2<T>* v, int m, C2<T> (&out)[2]) {
    if constexpr (sizeof(T) == 4) {
        const float4 q = reinterpret_cast<const float4*>(v)[m / 2];
        out[0] = C2<T>{q.x, q.y};
        out[1] = C2<T>{q.z, q.w};
    } else {
        const double2 q0 = reinterpret_cast<const double2*>(v)[m];
        const double2 q1 = reinterpret_cast<const double2*>(v)[m + 1];
        out[0] = C2<T>{q0.x, q0.y};
        out[1] = C2<T>{q1.x, q1.y};
    }
}

// Determinant inverse A = D^-1 (A[m][c]: orbital m, particle c), row-major
// per walker as in the reference (slater_determinant.cpp:363-365).  Lane c
// owns column c in the Sherman-Morrison update, so for a fixed row the wave
// reads and writes one contiguous run (coalesced); the row pointer is uniform
// across the wave, which lets the unrolled loads take a scalar base address.
// Complex entries are (re, im) pairs.
template <typename T, bool CPLX>
struct InvRef {
    T* base;
    int n;
    __device__ T* row(int m) const { return base + (CPLX ? 2 : 1) * m * n; }
    __device__ C2<T> get(int m, int c) const {
        if constexpr (CPLX) {
            const T* p = base + 2 * (m * n + c);
            return {p[0], p[1]};
        } else {
            return {base[m * n + c], T(0)};
        }
    }
    __device__ void set(int m, int c, C2<T> v) const {
        if constexpr (CPLX) {
            T* p = base + 2 * (m * n + c);
            p[0] = v.re;
            p[1] = v.im;
        } else {
            base[m * n + c] = v.re;
        }
    }
};

// One element of a row of A: a single 2-word load for a complex entry.
template <typename T, bool CPLX>
__device__ inline C2<T> load_elem_impl(const T* row, int c) {
    if constexpr (CPLX) return reinterpret_cast<const C2<T>*>(row)[c];
    else return C2<T>{row[c], T(0)};
}
template <typename T, bool CPLX>
__device__ inline void store_elem_impl(T* row, int c, C2<T> v) {
    if constexpr (CPLX) reinterpret_cast<C2<T>*>(row)[c] = v;
    else row[c] = v.re;
}

// LDS scratch of one walker, carved from the dynamic shared memory the host
// sizes per model (sweep_lds_bytes).
struct Shared {
    double* pos;   // [np*2]
    double* rho;   // [ng*nkj*2]
    double* v;     // [2*rows] proposed orbital values (stored in the determinant precision T)
    double* u;     // [2*rows] column j of A (old)
    double* f;     // [4*rows]: u / R, then v again, both as C2<T> (determinant precision)
    double* spl;   // [n_pair][spl_stride] pair-factor splines (copies of M.pair[f].rad.mono)
    double* kj;    // [nkj][2] k-space Jastrow vectors, then kjw [ng][ng][nkj]
};

__device__ inline Shared carve(const DevModel& M, double* base) {
    int nr = kRowPad;
    for (int b = 0; b < M.n_blocks; ++b) nr = M.blk_rows[b] > nr ? M.blk_rows[b] : nr;
    Shared sh;
    sh.pos = base;
    sh.rho = sh.pos + 2 * M.np;
    sh.v = sh.rho + 2 * M.ng * M.nkj;
    sh.u = sh.v + 2 * nr;
    sh.f = sh.u + 2 * nr;
    sh.spl = sh.f + 4 * nr;
    sh.kj = sh.spl + M.n_pair * M.spl_stride;
    return sh;
}


// Condition estimate of a freshly computed inverse, as stored: above 1e8 it
// is the block's own conditioning, which another rebuild cannot lower, and it
// is stored negated ("latched") so that the kappa > 1e8 rebuild trigger
// (slater_determinant.cpp:789, 825-826) stays quiet until the next batched
// refresh instead of rebuilding on every accepted move.
__device__ inline double fresh_cond(double kappa) { return kappa > 1e8 ? -kappa : kappa; }
// The same for walker w's block, counting a latched fresh inverse among the
// walker's rebuild triggers (kLatchedFresh).
__device__ inline void store_fresh_cond(const DevState& S, int w, int b, double kappa) {
    S.cond[w * kMaxBlocks + b] = fresh_cond(kappa);
    if (kappa > 1e8) atomicAdd(&S.rebuilds[w * kRebuildCauses + kLatchedFresh], 1);
}

// ── the per-walker engine ──────────────────────────────────────────────────
// NR = 0: the determinant inverses stay in HBM and every access goes there.
// NR > 0 (a multiple of kRowPad, >= every block's padded row count): the
// block being swept is held in registers, lane c holding column c (NR
// entries), from its first move in a sweep until a move of another block or
// the end of the kernel; HBM then sees one load and one store of each block
// per sweep instead of a read-modify-write of the block per accepted move.
template <typename T, bool CPLX, int NR = 0>
struct Walker {
    static_assert(NR % kRowPad == 0, "register tile is a whole number of row chunks");
    const DevModel& M;
    const DevState& S;
    Shared sh;
    int w, lane;
    double log_j;            // Jastrow + one-body part of log|psi|
    double log_b0 = 0.0, log_b1 = 0.0;   // determinant blocks (kMaxBlocks = 2)
    __device__ double& log_b(int b) { return b == 0 ? log_b0 : log_b1; }
    double kdr = 0.0, kdi = 0.0;  // this lane's k-space phase change of the last move
    C2<T> col[NR > 0 ? NR : 1];   // column `lane` of the resident block (NR > 0)
    int res_b = -1;               // resident block, -1 for none
    // the resident block's update count; HBM holds it again after spill()
    int res_upd = 0;

    // Load column `lane` of block b into registers (rows past the block's
    // padded count, and lanes past its orbital count, hold zeros).
    __device__ void fill(int b) {
        const InvRef<T, CPLX> A = inv(b);
        const int n = M.blk_n[b], nch = M.blk_rows[b] / kRowPad;
#pragma unroll
        for (int ch = 0; ch < NR / kRowPad; ++ch) {
#pragma unroll
            for (int i = 0; i < kRowPad; ++i) col[ch * kRowPad + i] = C2<T>{T(0), T(0)};
            if (ch < nch && lane < n) {
#pragma unroll
                for (int i = 0; i < kRowPad; ++i) col[ch * kRowPad + i] = A.get(ch * kRowPad + i, lane);
            }
        }
        res_upd = S.updates[w * kMaxBlocks + b];
        res_b = b;
        __syncthreads();
    }
    // Store the resident block back to HBM.
    __device__ void spill() {
        if (NR == 0 || res_b < 0) return;
        const InvRef<T, CPLX> A = inv(res_b);
        const int n = M.blk_n[res_b], nch = M.blk_rows[res_b] / kRowPad;
        if (lane < n) {
#pragma unroll
            for (int ch = 0; ch < NR / kRowPad; ++ch)
                if (ch < nch) {
#pragma unroll
                    for (int i = 0; i < kRowPad; ++i) A.set(ch * kRowPad + i, lane, col[ch * kRowPad + i]);
                }
        }
        if (lane == 0) S.updates[w * kMaxBlocks + res_b] = res_upd;
        res_b = -1;
    }
    __device__ void make_resident(int b) {
        if (res_b == b) return;
        spill();
        fill(b);
    }

    __device__ Walker(const DevModel& m, const DevState& s, const Shared& shm, int walker)
        : M(m), S(s), sh(shm), w(walker), lane(threadIdx.x) {}

    static __device__ C2<T> load_elem(const T* row, int c) { return load_elem_impl<T, CPLX>(row, c); }
    static __device__ void store_elem(T* row, int c, C2<T> v) { store_elem_impl<T, CPLX>(row, c, v); }

    __device__ InvRef<T, CPLX> inv(int b) const {
        T* base = static_cast<T*>(S.inv) + static_cast<size_t>(w) * S.inv_stride;
        int off = 0;
        for (int i = 0; i < b; ++i) off += (CPLX ? 2 : 1) * M.blk_rows[i] * M.blk_n[i];
        return {base + off, M.blk_n[b]};
    }

    __device__ void load() {
        const double* p = S.pos + static_cast<size_t>(w) * M.np * 2;
        for (int i = lane; i < 2 * M.np; i += W) sh.pos[i] = p[i];
        const int nr = M.ng * M.nkj * 2;
        const double* r = S.rho + static_cast<size_t>(w) * nr;
        for (int i = lane; i < nr; i += W) sh.rho[i] = r[i];
        log_j = S.logs[w * 4 + 0];
        if (M.n_blocks > 0) log_b0 = S.logs[w * 4 + 1];
        if (M.n_blocks > 1) log_b1 = S.logs[w * 4 + 2];
        // the tables every move gathers from: pair splines and k vectors
        for (int f = 0; f < M.n_pair; ++f) {
            const DevRadial& R = M.pair[f].rad;
            if (!R.image) continue;
            for (int i = lane; i < 6 * (R.n_knots - 1); i += W) sh.spl[f * M.spl_stride + i] = R.mono[i];
        }
        for (int k = lane; k < M.nkj; k += W) {
            sh.kj[2 * k] = M.kjx[k];
            sh.kj[2 * k + 1] = M.kjy[k];
            for (int g = 0; g < M.ng; ++g)
                for (int h = 0; h < M.ng; ++h) sh.kj[2 * M.nkj + (g * M.ng + h) * M.nkj + k] = M.kjw[g][h][k];
        }
        __syncthreads();
    }
    __device__ void store() {
        __syncthreads();
        double* p = S.pos + static_cast<size_t>(w) * M.np * 2;
        for (int i = lane; i < 2 * M.np; i += W) p[i] = sh.pos[i];
        const int nr = M.ng * M.nkj * 2;
        double* r = S.rho + static_cast<size_t>(w) * nr;
        for (int i = lane; i < nr; i += W) r[i] = sh.rho[i];
        if (lane == 0) {
            S.logs[w * 4 + 0] = log_j;
            if (M.n_blocks > 0) S.logs[w * 4 + 1] = log_b0;
            if (M.n_blocks > 1) S.logs[w * 4 + 2] = log_b1;
        }
    }

    // ── Jastrow + one-body: log-ratio, and the moved particle's gradient and
    //    Laplacian at the old (o*) and new (n*) positions ────────────────────
    struct JOut {
        double dlog;
        double og[2], ng[2], nl[2];
    };
    // Lane partial sums of the pair and k-space parts (reduced by the caller).
    // old_only: just the gradient at the current position (acc[1], acc[2]),
    // for a drift move's old gradient; nx is then ignored.
    template <bool derivs, bool old_only = false>
    __device__ void jastrow_lanes(int p, double x0, double x1, double nx0, double nx1, double (&acc)[7]) {
        const int gp = M.group_of[p];
        for (int i = 0; i < 7; ++i) acc[i] = 0.0;   // dlog, og0, og1, ng0, ng1, nl0, nl1
        // partners group by group (the groups partition the particles), so
        // the pair factors, and the loads of their parameters, are uniform
        // across the wavefront
        for (int h = 0; h < M.ng; ++h) {
          const int nf = M.pair_list_n[gp][h];
          if (nf == 0) continue;
          for (int k = M.g_first[h] + lane; k < M.g_last[h]; k += W) {
            if (k == p) continue;
            const double xk0 = sh.pos[2 * k], xk1 = sh.pos[2 * k + 1];
            double o0 = x0 - xk0, o1 = x1 - xk1;
            const double ro = dist(M.cell, o0, o1);
            if constexpr (old_only) {
                for (int fi = 0; fi < nf; ++fi) {
                    const int f = M.pair_list[gp][h][fi];
                    double uo, duo, d2uo;
                    radial_eval(M.pair[f].rad, ro, uo, duo, d2uo, sh.spl + f * M.spl_stride);
                    double g0, g1, l0, l1;
                    pair_terms(duo, d2uo, ro, o0, o1, g0, g1, l0, l1);
                    acc[1] += g0;
                    acc[2] += g1;
                }
                continue;
            }
            double n0 = nx0 - xk0, n1 = nx1 - xk1;
            const double rn = dist(M.cell, n0, n1);
            for (int fi = 0; fi < nf; ++fi) {
                const int f = M.pair_list[gp][h][fi];
                const DevRadial& R = M.pair[f].rad;
                const double* spl = sh.spl + f * M.spl_stride;   // LDS monomial copy (image factors)
                double uo, duo, d2uo, un, dun, d2un;
                radial_eval(R, ro, uo, duo, d2uo, spl);
                radial_eval(R, rn, un, dun, d2un, spl);
                acc[0] += un - uo;
                if constexpr (derivs) {
                    double g0, g1, l0, l1;
                    pair_terms(duo, d2uo, ro, o0, o1, g0, g1, l0, l1);
                    acc[1] += g0;
                    acc[2] += g1;
                    pair_terms(dun, d2un, rn, n0, n1, g0, g1, l0, l1);
                    acc[3] += g0;
                    acc[4] += g1;
                    acc[5] += l0;
                    acc[6] += l1;
                }
            }
          }
        }
        // k-space: lanes over k
        if (M.nkj > 0 && M.kj_active[gp]) {
            for (int k = lane; k < M.nkj; k += W) {   // one k per lane (nkj <= 64)
                const double kx = sh.kj[2 * k], ky = sh.kj[2 * k + 1];
                double so, co, sn, cn;
                sincos_bounded(kx * x0 + ky * x1, &so, &co);
                if constexpr (old_only) {
                    double gold = 0.0;
                    for (int h = 0; h < M.ng; ++h) {
                        if (!M.kj_active[h]) continue;
                        const double wgh = sh.kj[2 * M.nkj + (gp * M.ng + h) * M.nkj + k];
                        if (wgh == 0.0) continue;
                        const double rr = sh.rho[(h * M.nkj + k) * 2], ri = sh.rho[(h * M.nkj + k) * 2 + 1];
                        gold += 2.0 * wgh * (co * ri - so * rr);
                    }
                    acc[1] += kx * gold;
                    acc[2] += ky * gold;
                    continue;
                }
                sincos_bounded(kx * nx0 + ky * nx1, &sn, &cn);
                const double dr = cn - co, di = sn - so;
                kdr = dr;   // kept for jastrow_commit of this move
                kdi = di;
                double dl = 0.0;
                // moved particle's gradient at the old / new position
                double gold = 0.0, gnew = 0.0, lnew = 0.0;
                for (int h = 0; h < M.ng; ++h) {
                    if (!M.kj_active[h]) continue;
                    const double wgh = sh.kj[2 * M.nkj + (gp * M.ng + h) * M.nkj + k];
                    if (wgh == 0.0) continue;
                    const double rr = sh.rho[(h * M.nkj + k) * 2], ri = sh.rho[(h * M.nkj + k) * 2 + 1];
                    if (h == gp) {
                        dl += wgh * (2.0 * (rr * dr + ri * di) + dr * dr + di * di);
                        if constexpr (derivs) {
                            gold += 2.0 * wgh * (co * ri - so * rr);
                            const double qr = rr + dr, qi = ri + di;
                            gnew += 2.0 * wgh * (cn * qi - sn * qr);
                            lnew += 2.0 * wgh * (1.0 - (cn * qr + sn * qi));
                        }
                    } else {
                        dl += 2.0 * wgh * (dr * rr + di * ri);
                        if constexpr (derivs) {
                            gold += 2.0 * wgh * (co * ri - so * rr);
                            gnew += 2.0 * wgh * (cn * ri - sn * rr);
                            lnew += -2.0 * wgh * (cn * rr + sn * ri);
                        }
                    }
                }
                acc[0] += dl;
                if constexpr (derivs) {
                    acc[1] += kx * gold;
                    acc[2] += ky * gold;
                    acc[3] += kx * gnew;
                    acc[4] += ky * gnew;
                    acc[5] += kx * kx * lnew;
                    acc[6] += ky * ky * lnew;
                }
            }
        }
    }
    // The one-body part (uniform across lanes), added to the reduced sums.
    __device__ void jastrow_onebody(int p, double x0, double x1, double nx0, double nx1, double (&acc)[7]) {
        if (M.has_nosanow && M.nos_alpha[p] != 0.0) {
            const double a = M.nos_alpha[p];
            double o0 = x0 - M.nos_site[p][0], o1 = x1 - M.nos_site[p][1];
            double n0 = nx0 - M.nos_site[p][0], n1 = nx1 - M.nos_site[p][1];
            if (M.cell.periodic) { min_image(M.cell, o0, o1); min_image(M.cell, n0, n1); }
            acc[0] += (-a * n0 * n0 - a * n1 * n1) - (-a * o0 * o0 - a * o1 * o1);
            acc[1] += -2.0 * a * o0;
            acc[2] += -2.0 * a * o1;
            acc[3] += -2.0 * a * n0;
            acc[4] += -2.0 * a * n1;
            acc[5] += -2.0 * a;
            acc[6] += -2.0 * a;
        }
    }
    static __device__ void jastrow_out(const double (&acc)[7], JOut& out) {
        out.dlog = acc[0];
        out.og[0] = acc[1]; out.og[1] = acc[2];
        out.ng[0] = acc[3]; out.ng[1] = acc[4];
        out.nl[0] = acc[5]; out.nl[1] = acc[6];
    }

    // Commit the k-space densities of the accepted move of p, with the phase
    // change jastrow_lanes computed for it (one k per lane).
    __device__ void jastrow_commit(int p) {
        const int gp = M.group_of[p];
        if (M.nkj > 0 && M.kj_active[gp] && lane < M.nkj) {
            sh.rho[(gp * M.nkj + lane) * 2] += kdr;
            sh.rho[(gp * M.nkj + lane) * 2 + 1] += kdi;
        }
    }

    // ── determinant: ratio of moving p (column j of block b) ───────────────
    struct DOut {
        double R[2];       // ratio (complex)
        double og[2][2];   // old gradient of p, complex per dim: [d][re/im]
        double ol[2][2];
        double ng[2][2];   // Sum_m A_mj grad phi_m(r') (before the 1/R)
        double nl[2][2];
    };
    // Lane partial sums of the ratio of moving p (column j of block b), the
    // proposed orbital values v' (sh.v) and column j (sh.u) in LDS; reduced
    // by the caller, and sh.v / sh.u are read only after a later sync.
    template <bool derivs, bool old_derivs>
    __device__ void det_lanes(int b, int j, double x0, double x1, double nx0, double nx1, double (&acc)[20]) {
        const InvRef<T, CPLX> A = inv(b);
        const int n = M.blk_n[b];
        if constexpr (NR > 0) {
            // column j lives in lane j (the sweep loop made block b resident
            // before the move): stage it through LDS for the lanes over m
            const int nch = M.blk_rows[b] / kRowPad;
            // (in the determinant precision, 16 bytes per store)
            if (lane == j) {
#pragma unroll
                for (int ch = 0; ch < NR / kRowPad; ++ch)
                    if (ch < nch) {
                        const int m0 = ch * kRowPad;
                        if constexpr (sizeof(T) == 4) {
                            float4* s4 = reinterpret_cast<float4*>(sh.u);
#pragma unroll
                            for (int i = 0; i < kRowPad / 2; ++i)
                                s4[m0 / 2 + i] = make_float4(col[m0 + 2 * i].re, col[m0 + 2 * i].im,
                                                             col[m0 + 2 * i + 1].re, col[m0 + 2 * i + 1].im);
                        } else {
#pragma unroll
                            for (int i = 0; i < kRowPad; ++i) reinterpret_cast<C2<T>*>(sh.u)[m0 + i] = col[m0 + i];
                        }
                    }
            }
            __syncthreads();
        }
        for (int i = 0; i < 20; ++i) acc[i] = 0.0;
        for (int m = lane; m < n; m += W) {
            C2<T> a;
            if constexpr (NR > 0) a = reinterpret_cast<const C2<T>*>(sh.u)[m];
            else a = A.get(m, j);
            const double ar = static_cast<double>(a.re), ai = static_cast<double>(a.im);
            if constexpr (!derivs && old_derivs) {
                // the old gradient alone (a drift move's, r' = r): one
                // orbital evaluation, Sum_m A_mj grad phi_m(r) in acc[11..14]
                Orb oo;
                orbital(M, b, m, x0, x1, oo);
                for (int d = 0; d < 2; ++d) {
                    acc[11 + 2 * d] += ar * oo.gr[d] - ai * oo.gi[d];
                    acc[12 + 2 * d] += ar * oo.gi[d] + ai * oo.gr[d];
                }
                continue;
            }
            Orb on;
            T vr, vi;   // values in the determinant precision
            if constexpr (sizeof(T) == 4 && !derivs && !old_derivs) {
                orbital_value_f32(M, b, m, nx0, nx1, vr, vi);
                on.vr = static_cast<double>(vr);
                on.vi = static_cast<double>(vi);
            } else {
                orbital(M, b, m, nx0, nx1, on);
                vr = static_cast<T>(on.vr);
                vi = static_cast<T>(on.vi);
            }
            sh.v[2 * m] = static_cast<double>(vr);
            sh.v[2 * m + 1] = static_cast<double>(vi);
            if constexpr (NR == 0) reinterpret_cast<C2<T>*>(sh.u)[m] = a;
            const double tr = ar * static_cast<double>(vr) - ai * static_cast<double>(vi);
            const double ti = ar * static_cast<double>(vi) + ai * static_cast<double>(vr);
            acc[0] += tr;
            acc[1] += ti;
            // (the magnitude sums behind the rebuild thresholds are taken in
            // det_commit, for accepted moves only)
            if constexpr (derivs) {
                for (int d = 0; d < 2; ++d) {
                    acc[3 + 2 * d] += ar * on.gr[d] - ai * on.gi[d];
                    acc[4 + 2 * d] += ar * on.gi[d] + ai * on.gr[d];
                    acc[7 + 2 * d] += ar * on.lr[d] - ai * on.li[d];
                    acc[8 + 2 * d] += ar * on.li[d] + ai * on.lr[d];
                }
            }
            if constexpr (old_derivs) {
                Orb oo;
                orbital(M, b, m, x0, x1, oo);
                for (int d = 0; d < 2; ++d) {
                    acc[11 + 2 * d] += ar * oo.gr[d] - ai * oo.gi[d];
                    acc[12 + 2 * d] += ar * oo.gi[d] + ai * oo.gr[d];
                    acc[15 + 2 * d] += ar * oo.lr[d] - ai * oo.li[d];
                    acc[16 + 2 * d] += ar * oo.li[d] + ai * oo.lr[d];
                }
            }
        }
    }
    static __device__ void det_out(const double (&acc)[20], DOut& o) {
        o.R[0] = acc[0];
        o.R[1] = acc[1];
        for (int d = 0; d < 2; ++d) {
            o.ng[d][0] = acc[3 + 2 * d]; o.ng[d][1] = acc[4 + 2 * d];
            o.nl[d][0] = acc[7 + 2 * d]; o.nl[d][1] = acc[8 + 2 * d];
            o.og[d][0] = acc[11 + 2 * d]; o.og[d][1] = acc[12 + 2 * d];
            o.ol[d][0] = acc[15 + 2 * d]; o.ol[d][1] = acc[16 + 2 * d];
        }
    }

    // Pivoted Gauss-Jordan rebuild of block b from the current positions
    // (slater_determinant.cpp:905-993), in FP64, in the walker's workspace.
    // Returns log|det D|; sets status on a singular matrix.
    __device__ double rebuild(int b) {
        const int n = M.blk_n[b];
        const int e = CPLX ? 2 : 1;
        double* Mm = S.work + static_cast<size_t>(w) * S.work_stride;
        double* Aa = Mm + e * n * n;
        const int first = M.blk_first[b];
        for (int m = lane; m < n; m += W)
            for (int kr = 0; kr < n; ++kr) {
                Orb o;
                orbital(M, b, m, sh.pos[2 * (first + kr)], sh.pos[2 * (first + kr) + 1], o);
                Mm[e * (kr * n + m)] = o.vr;
                if (CPLX) Mm[e * (kr * n + m) + 1] = o.vi;
            }
        for (int c = lane; c < n; c += W)
            for (int r = 0; r < n; ++r) {
                Aa[e * (r * n + c)] = r == c ? 1.0 : 0.0;
                if (CPLX) Aa[e * (r * n + c) + 1] = 0.0;
            }
        __syncthreads();
        double log_det = 0.0;
        for (int k = 0; k < n; ++k) {
            // pivot: first row of largest |M[r][k]|, r >= k
            double best = -1.0;
            int piv = n;
            for (int r = k + lane; r < n; r += W) {
                const double re = Mm[e * (r * n + k)];
                const double im = CPLX ? Mm[e * (r * n + k) + 1] : 0.0;
                const double a = CPLX ? sqrt(re * re + im * im) : fabs(re);
                if (a > best) { best = a; piv = r; }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double ob = __shfl_xor(best, o, W);
                const int op = __shfl_xor(piv, o, W);
                if (ob > best || (ob == best && op < piv)) { best = ob; piv = op; }
            }
            const double dr = Mm[e * (piv * n + k)];
            const double di = CPLX ? Mm[e * (piv * n + k) + 1] : 0.0;
            const double dabs = CPLX ? sqrt(dr * dr + di * di) : fabs(dr);
            if (!(dabs > 0.0) || !isfinite(dabs)) {
                if (lane == 0) S.status[w] = 1;
                return log_det;
            }
            log_det += log(dabs);
            __syncthreads();
            if (piv != k) {
                for (int c = lane; c < n; c += W)
                    for (int q = 0; q < e; ++q) {
                        double t = Mm[e * (k * n + c) + q];
                        Mm[e * (k * n + c) + q] = Mm[e * (piv * n + c) + q];
                        Mm[e * (piv * n + c) + q] = t;
                        t = Aa[e * (k * n + c) + q];
                        Aa[e * (k * n + c) + q] = Aa[e * (piv * n + c) + q];
                        Aa[e * (piv * n + c) + q] = t;
                    }
                __syncthreads();
            }
            // row k /= diag
            double ir, ii;
            if (CPLX) {
                const double d2 = dr * dr + di * di;
                ir = dr / d2;
                ii = -di / d2;
            } else {
                ir = 1.0 / dr;
                ii = 0.0;
            }
            for (int c = lane; c < n; c += W) {
                if (CPLX) {
                    const double mr = Mm[2 * (k * n + c)], mi = Mm[2 * (k * n + c) + 1];
                    Mm[2 * (k * n + c)] = mr * ir - mi * ii;
                    Mm[2 * (k * n + c) + 1] = mr * ii + mi * ir;
                    const double ar = Aa[2 * (k * n + c)], ai = Aa[2 * (k * n + c) + 1];
                    Aa[2 * (k * n + c)] = ar * ir - ai * ii;
                    Aa[2 * (k * n + c) + 1] = ar * ii + ai * ir;
                } else {
                    Mm[k * n + c] /= dr;
                    Aa[k * n + c] /= dr;
                }
            }
            __syncthreads();
            // eliminate column k from every other row
            for (int c = lane; c < n; c += W) {
                const double mkr = Mm[e * (k * n + c)];
                const double mki = CPLX ? Mm[e * (k * n + c) + 1] : 0.0;
                const double akr = Aa[e * (k * n + c)];
                const double aki = CPLX ? Aa[e * (k * n + c) + 1] : 0.0;
                for (int r = 0; r < n; ++r) {
                    if (r == k) continue;
                    const double fr = Mm[e * (r * n + k)];
                    const double fi = CPLX ? Mm[e * (r * n + k) + 1] : 0.0;
                    if (c != k) {
                        Mm[e * (r * n + c)] -= fr * mkr - fi * mki;
                        if (CPLX) Mm[e * (r * n + c) + 1] -= fr * mki + fi * mkr;
                    }
                    Aa[e * (r * n + c)] -= fr * akr - fi * aki;
                    if (CPLX) Aa[e * (r * n + c) + 1] -= fr * aki + fi * akr;
                }
            }
            __syncthreads();
            // column k of the reduced matrix is only read through M[r][k]
            // above; set it to the identity column now that the step is done
            for (int r = lane; r < n; r += W) {
                Mm[e * (r * n + k)] = r == k ? 1.0 : 0.0;
                if (CPLX) Mm[e * (r * n + k) + 1] = 0.0;
            }
            __syncthreads();
        }
        // copy to the inverse in the determinant precision, row sums of D
        const InvRef<T, CPLX> A = inv(b);
        double colmax = 0.0;
        for (int c = lane; c < n; c += W) {
            double cs = 0.0;
            for (int m = 0; m < n; ++m) {
                const double ar = Aa[e * (m * n + c)];
                const double ai = CPLX ? Aa[e * (m * n + c) + 1] : 0.0;
                A.set(m, c, C2<T>{static_cast<T>(ar), static_cast<T>(ai)});
                cs += fabs(ar) + fabs(ai);
            }
            colmax = fmax(colmax, cs);
            for (int m = n; m < M.blk_rows[b]; ++m) A.set(m, c, C2<T>{T(0), T(0)});
        }
        colmax = wmax(colmax);
        double rowmax = 0.0;
        for (int kr = lane; kr < n; kr += W) {
            double rs = 0.0;
            for (int m = 0; m < n; ++m) {
                Orb o;
                orbital(M, b, m, sh.pos[2 * (first + kr)], sh.pos[2 * (first + kr) + 1], o);
                rs += CPLX ? sqrt(o.vr * o.vr + o.vi * o.vi) : fabs(o.vr);
            }
            rowmax = fmax(rowmax, rs);
        }
        rowmax = wmax(rowmax);
        if (lane == 0) {
            store_fresh_cond(S, w, b, rowmax * colmax);
            S.updates[w * kMaxBlocks + b] = 0;
        }
        __syncthreads();
        return log_det;
    }

    // Sherman-Morrison update of block b for the accepted move of column j
    // (slater_determinant.cpp:744-858), with the reference's rebuild
    // triggers.  sh.v holds v' (orbitals at r'), sh.u column j of A, and
    // sh.pos already holds the committed position of the moved particle.
    __device__ bool det_commit(int b, int j, const DOut& d) {
        const int n = M.blk_n[b];
        const double Rr = d.R[0], Ri = d.R[1];
        // The cancellation scale Sum_m |A_mj v'_m|, from column j (sh.u) and
        // v' (sh.v) as the ratio used them.  It feeds a threshold only, so a
        // single-precision square root is enough.
        double scale = 0.0;
        for (int m = lane; m < n; m += W) {
            const C2<T> a = reinterpret_cast<const C2<T>*>(sh.u)[m];
            const double ar = static_cast<double>(a.re), ai = static_cast<double>(a.im);
            const double vr = sh.v[2 * m], vi = sh.v[2 * m + 1];
            const double tr = ar * vr - ai * vi, ti = ar * vi + ai * vr;
            scale += CPLX ? static_cast<double>(__builtin_sqrtf(static_cast<float>(tr * tr + ti * ti))) : fabs(tr);
        }
        scale = wsum(scale);
        const double absR = CPLX ? sqrt(Rr * Rr + Ri * Ri) : fabs(Rr);
        const double eps = sizeof(T) == 8 ? 2.220446049250313e-16 : 1.1920928955078125e-07;
        int* upd = S.updates + w * kMaxBlocks + b;
        const int updates = (NR > 0 ? res_upd : *upd) + 1;
        __syncthreads();
        // The reference also rebuilds after 64 updates (slater_determinant.cpp:
        // 788-790); here that periodic rebuild is the batched refresh kernel,
        // which the host runs after every sweep.  A tiny or non-finite ratio,
        // and the cancellation guard (the rounding bound eps * sum|A_mj v_m|
        // on R above |R| / cancel_factor), rebuild here.  The condition
        // estimate is taken when an inverse is rebuilt (see fresh_cond); an
        // update rebuilds instead when its size bound max|r| max|f| is not
        // finite or could carry A past the precision's range, which is where
        // the reference's estimate turns non-finite (slater_determinant.cpp:
        // 825-826).
        int cause = -1;   // RebuildCause, uniform across the wave
        if (!isfinite(absR) || absR < 1e-5) cause = kRebuildTinyRatio;
        else if (absR <= M.cancel_factor * eps * scale) cause = kRebuildCancellation;
        bool rebuild_now = cause >= 0;
        if (!rebuild_now) {
            // r_c = sum_m v'_m A[m][c] - delta_cj, then A[m][c] -= f_m r_c
            // with f = u / R
            const T invRr = static_cast<T>(CPLX ? Rr / (absR * absR) : 1.0 / Rr);
            const T invRi = static_cast<T>(CPLX ? -Ri / (absR * absR) : 0.0);
            const int rows = M.blk_rows[b];
            C2<T>* fT = reinterpret_cast<C2<T>*>(sh.f);
            const C2<T>* vT = reinterpret_cast<const C2<T>*>(sh.f) + rows;   // packed after f
            T fmag = T(0);
            for (int m = lane; m < rows; m += W) {
                if (m < n) {
                    const C2<T> um = reinterpret_cast<const C2<T>*>(sh.u)[m];
                    const C2<T> fm = cmul(um, C2<T>{invRr, invRi});
                    fT[m] = fm;
                    fmag = fmax(fmag, fabs(fm.re) + fabs(fm.im));
                    reinterpret_cast<C2<T>*>(sh.f)[rows + m] =
                        C2<T>{static_cast<T>(sh.v[2 * m]), static_cast<T>(sh.v[2 * m + 1])};
                } else {
                    fT[m] = C2<T>{T(0), T(0)};
                    reinterpret_cast<C2<T>*>(sh.f)[rows + m] = C2<T>{T(0), T(0)};
                }
            }
            // |f| and |r| bound |f_m r_c| by max|f| max|r|: within this range
            // neither 64 updates nor a non-finite product can leave A
            // non-finite
            const double f_max = wmax(static_cast<double>(fmag));
            const double safe = sizeof(T) == 8 ? 1e300 : 1e30;
            bool bad = false;
            __syncthreads();
            // Lane c owns column c and holds it in registers, CH rows per
            // chunk: pass 1 issues a whole chunk of loads at once, and pass 2
            // updates the last chunk straight from registers (a column of
            // <= CH rows is read once and written once per accepted move).
            constexpr int CH = kRowPad;   // rows per chunk; rows padded with zeros to a multiple
            const int nch = M.blk_rows[b] / CH;
            if constexpr (NR > 0) {
                // the same two passes on the column held in registers; lanes
                // past n hold a zero column and stay zero.  v' and u/R are
                // LDS broadcasts, read 16 bytes at a time; each complex
                // multiply-accumulate is four FMAs.
                C2<T> r{T(0), T(0)};
#pragma unroll
                for (int ch = 0; ch < NR / CH; ++ch)
                    if (ch < nch) {
#pragma unroll
                        for (int i = 0; i < CH; i += 2) {
                            C2<T> v[2];
                            lds_pair(vT, ch * CH + i, v);
#pragma unroll
                            for (int q = 0; q < 2; ++q) {
                                const C2<T>& a = col[ch * CH + i + q];
                                r.re = fmadd(v[q].re, a.re, r.re);
                                if constexpr (CPLX) {
                                    r.re = fmadd(-v[q].im, a.im, r.re);
                                    r.im = fmadd(v[q].re, a.im, r.im);
                                    r.im = fmadd(v[q].im, a.re, r.im);
                                }
                            }
                        }
                    }
                if (lane == j) r.re -= T(1);
                bad = !(static_cast<double>(fabs(r.re) + fabs(r.im)) * f_max <= safe);
#pragma unroll
                for (int ch = NR / CH - 1; ch >= 0; --ch)
                    if (ch < nch) {
#pragma unroll
                        for (int i = 0; i < CH; i += 2) {
                            C2<T> f[2];
                            lds_pair(fT, ch * CH + i, f);
#pragma unroll
                            for (int q = 0; q < 2; ++q) {
                                C2<T>& a = col[ch * CH + i + q];
                                a.re = fmadd(-f[q].re, r.re, a.re);
                                if constexpr (CPLX) {
                                    a.re = fmadd(f[q].im, r.im, a.re);
                                    a.im = fmadd(-f[q].re, r.im, a.im);
                                    a.im = fmadd(-f[q].im, r.re, a.im);
                                }
                            }
                        }
                    }
            }
            const InvRef<T, CPLX> A = inv(b);
            for (int c = lane; NR == 0 && c < n; c += W) {
                C2<T> a[CH];
                C2<T> r{T(0), T(0)};
                for (int ch = 0; ch < nch; ++ch) {
                    const int m0 = ch * CH;
#pragma unroll
                    for (int i = 0; i < CH; ++i) a[i] = load_elem(A.row(m0 + i), c);
#pragma unroll
                    for (int i = 0; i < CH; ++i) {
                        const C2<T> pr = cmul(vT[m0 + i], a[i]);
                        r.re += pr.re;
                        r.im += pr.im;
                    }
                }
                if (c == j) r.re -= T(1);
                bad = bad || !(static_cast<double>(fabs(r.re) + fabs(r.im)) * f_max <= safe);
                // pass 2: A[m][c] -= f_m r_c, last chunk first (in registers)
                for (int ch = nch - 1; ch >= 0; --ch) {
                    const int m0 = ch * CH;
                    if (ch != nch - 1) {
#pragma unroll
                        for (int i = 0; i < CH; ++i) a[i] = load_elem(A.row(m0 + i), c);
                    }
#pragma unroll
                    for (int i = 0; i < CH; ++i) {
                        const C2<T> pr = cmul(fT[m0 + i], r);
                        a[i].re -= pr.re;
                        a[i].im -= pr.im;
                        store_elem(A.row(m0 + i), c, a[i]);
                    }
                }
            }
            if (__builtin_amdgcn_ballot_w64(bad) != 0) {
                // the rebuild below replaces the updated inverse
                rebuild_now = true;
                cause = kRebuildUpdateBound;
            } else {
                log_b(b) += log(absR);
                if constexpr (NR > 0) res_upd = updates;
                else if (lane == 0) *upd = updates;
            }
        }
        if (rebuild_now) {
            if (lane == 0) S.rebuilds[w * kRebuildCauses + cause] += 1;
            log_b(b) = rebuild(b);
            if constexpr (NR > 0) fill(b);   // the rebuild wrote the block to HBM
        }
        __syncthreads();
        return rebuild_now;
    }

    // A twisted plane wave is not periodic: phi(r - L) = phi(r) e^{-i q.L}.
    // Wrapping the accepted position re-gauges row j of D by that phase, so
    // column j of A = D^-1 takes the inverse phase (the reference re-inverts
    // from the wrapped rows on the next accept, complex_slater_determinant.cpp:
    // 222-253; every gauge-invariant quantity is unchanged).
    __device__ void regauge_column(int b, int j, double shift0, double shift1) {
        double s, c;
        sincos_bounded(M.orb_k[b][0][0] * shift0 + M.orb_k[b][0][1] * shift1, &s, &c);
        const InvRef<T, CPLX> A = inv(b);
        const C2<T> ph{static_cast<T>(c), static_cast<T>(s)};
        if constexpr (NR > 0) {
            // resident since this move's ratio; column j is lane j's registers
            if (lane == j) {
#pragma unroll
                for (int m = 0; m < NR; ++m) col[m] = cmul(col[m], ph);
            }
        } else {
            for (int m = lane; m < M.blk_n[b]; m += W) A.set(m, j, cmul(A.get(m, j), ph));
        }
        __syncthreads();
    }

    // ── one Metropolis / drift-diffusion move of particle p ────────────────
    // mode 0: sweep (decide from the random numbers); 1: ratio only (per-call
    // path, writes DevCall outputs); 2: commit where call.mask is set.
    // kind (mode 0): kSweepVmc, kSweepDmc or kSweepLangevin.
    template <bool derivs>
    __device__ bool move(int p, double nx0, double nx1, double uni, int mode, const DevCall* call,
                         double tau, int kind, double drift_scale, const double* gold,
                         double& rr_acc) {
        const double x0 = sh.pos[2 * p], x1 = sh.pos[2 * p + 1];
        const int b = M.block_of[p];
        // Jastrow and determinant lane sums, then one reduction for both
        double ja[7], da[20];
        jastrow_lanes<derivs>(p, x0, x1, nx0, nx1, ja);
        if (b >= 0) {
            det_lanes<derivs, false>(b, p - M.blk_first[b], x0, x1, nx0, nx1, da);
        } else {
            for (int i = 0; i < 20; ++i) da[i] = 0.0;
        }
        if constexpr (derivs) {
            // Jastrow dlog, gradients and Laplacians; R; Sum A grad/lap phi(r')
            double r[17];
            for (int i = 0; i < 7; ++i) r[i] = ja[i];
            r[7] = da[0];
            r[8] = da[1];
            for (int i = 0; i < 8; ++i) r[9 + i] = da[3 + i];
            wsum_n(r);
            for (int i = 0; i < 7; ++i) ja[i] = r[i];
            da[0] = r[7];
            da[1] = r[8];
            for (int i = 0; i < 8; ++i) da[3 + i] = r[9 + i];
        } else {
            double r[3] = {ja[0], da[0], da[1]};
            wsum_n(r);
            ja[0] = r[0]; da[0] = r[1]; da[1] = r[2];
        }
        jastrow_onebody(p, x0, x1, nx0, nx1, ja);
        JOut J;
        jastrow_out(ja, J);
        double lr = J.dlog, phase = 0.0;
        double ng[2] = {J.ng[0], J.ng[1]}, nl[2] = {J.nl[0], J.nl[1]};
        double npg[2] = {0, 0}, npl[2] = {0, 0};
        DOut D{};
        double R2 = 1.0;   // |R|^2 of the determinant ratio
        if (b >= 0) {
            det_out(da, D);
            __syncthreads();
            R2 = D.R[0] * D.R[0] + D.R[1] * D.R[1];
            const double absR = sqrt(R2);
            lr += log(fmax(absR, 1e-300));
            if (CPLX) {
                phase += atan2(D.R[1], D.R[0]);
            } else if (D.R[0] < 0.0) {
                phase += 3.14159265358979323846;
            }
            if constexpr (derivs) {
                // G' = (sum A grad phi)/R, L' = (sum A lap phi)/R - G'^2
                double ir = 0.0, ii = 0.0;
                if (absR > 0.0) { ir = D.R[0] / (absR * absR); ii = -D.R[1] / (absR * absR); }
                for (int d = 0; d < 2; ++d) {
                    const double gr = D.ng[d][0] * ir - D.ng[d][1] * ii;
                    const double gi = D.ng[d][0] * ii + D.ng[d][1] * ir;
                    const double lr_ = D.nl[d][0] * ir - D.nl[d][1] * ii;
                    const double li = D.nl[d][0] * ii + D.nl[d][1] * ir;
                    ng[d] += gr;
                    nl[d] += lr_ - (gr * gr - gi * gi);
                    npg[d] += gi;
                    npl[d] += li - 2.0 * gr * gi;
                }
            }
        }
        if (mode == 1) {
            if (lane == 0) {
                const int nw = M.nw;
                call->log_ratio[w] = lr;
                call->phase[w] = phase;
                for (int d = 0; d < 2; ++d) {
                    call->new_grad[d * nw + w] = ng[d];
                    call->new_lap[d * nw + w] = nl[d];
                    if (call->new_pgrad) {
                        call->new_pgrad[d * nw + w] = npg[d];
                        call->new_plap[d * nw + w] = npl[d];
                    }
                }
            }
            return false;
        }
        bool acc;
        if (mode == 2) {
            acc = call->mask[w] != 0;
        } else if (kind == kSweepVmc) {
            // uni < |psi'/psi|^2 = exp(2 lr), with the determinant's share as
            // |R|^2 rather than through its log; the mixed form, whose ratio
            // is single precision, takes the exponential in single precision
            // (its argument bounded to the float range)
            if constexpr (sizeof(T) == 4)
                acc = uni < static_cast<double>(expf(static_cast<float>(fmin(fmax(2.0 * J.dlog, -87.0), 88.0)))) * R2;
            else
                acc = uni < exp(2.0 * J.dlog) * R2;
        } else if (kind == kSweepLangevin) {
            // VMC with the bare drift (tau/m) g: no node gate, the Green's
            // function log-ratio per dimension -delta (g + g') + (tau/m)/2
            // (g^2 - g'^2) (numerics/green_function.h:20-26)
            const double tm = tau * M.inv_mass[p];
            double lc = 0.0;
            for (int d = 0; d < 2; ++d) {
                const double delta = (d == 0 ? nx0 - x0 : nx1 - x1);
                lc += -delta * (gold[d] + ng[d]) + 0.5 * tm * (gold[d] * gold[d] - ng[d] * ng[d]);
            }
            const double la = 2.0 * lr + lc;
            acc = la >= 0.0 || uni < exp(la);
        } else {
            const double tm = tau * M.inv_mass[p];
            // a real trial's node crossing: the phase change is 0 or pi, so
            // cos(phase) < eps is the determinant ratio's sign
            if (!CPLX && b >= 0 && D.R[0] < 0.0) {
                acc = false;
            } else {
                const double vn = ng[0] * ng[0] + ng[1] * ng[1];
                const double sb = vn < 1e-10 ? tm : (-1.0 + sqrt(1.0 + 2.0 * tm * vn)) / vn;
                double lc = 0.0;
                for (int d = 0; d < 2; ++d) {
                    const double delta = (d == 0 ? nx0 - x0 : nx1 - x1);
                    const double df = drift_scale * gold[d];
                    const double db = sb * ng[d];
                    lc += tm <= 0.0 ? 0.0 : (-delta * (df + db) + 0.5 * (df * df - db * db)) / tm;
                }
                const double la = 2.0 * lr + lc;
                acc = la >= 0.0 || uni < exp(la);
            }
            const double d0 = nx0 - x0, d1 = nx1 - x1;
            const double rr = tm * d0 * d0 + tm * d1 * d1;
            rr_acc += acc ? rr : 0.0;
        }
        if (!acc) return false;
        log_j += J.dlog;
        jastrow_commit(p);
        __syncthreads();
        if (lane == 0) {
            double w0 = nx0, w1 = nx1;
            wrap(M.cell, w0, w1);
            sh.pos[2 * p] = w0;
            sh.pos[2 * p + 1] = w1;
        }
        __syncthreads();
        if (b >= 0) {
            const bool rebuilt = det_commit(b, p - M.blk_first[b], D);
            const double s0 = nx0 - sh.pos[2 * p], s1 = nx1 - sh.pos[2 * p + 1];
            if (CPLX && !rebuilt && (s0 != 0.0 || s1 != 0.0)) regauge_column(b, p - M.blk_first[b], s0, s1);
        }
        return true;
    }

    // Old gradient of p (the drift of DMC and Langevin moves), all factors:
    // only the sums it needs are evaluated and reduced.
    __device__ void grad_of(int p, double g[2]) {
        const double x0 = sh.pos[2 * p], x1 = sh.pos[2 * p + 1];
        double ja[7];
        jastrow_lanes<true, true>(p, x0, x1, x0, x1, ja);
        double r[4] = {ja[1], ja[2], 0.0, 0.0};
        const int b = M.block_of[p];
        if (b >= 0) {
            double da[20];
            det_lanes<false, true>(b, p - M.blk_first[b], x0, x1, x0, x1, da);
            r[2] = da[11];   // Re Sum_m A_mj d_x phi_m(r): at r' = r the ratio is 1
            r[3] = da[13];
        }
        wsum_n(r);
        double acc[7] = {0.0, r[0], r[1], 0.0, 0.0, 0.0, 0.0};
        jastrow_onebody(p, x0, x1, x0, x1, acc);
        g[0] = acc[1] + r[2];
        g[1] = acc[2] + r[3];
        __syncthreads();
    }
};

// ── kernels ────────────────────────────────────────────────────────────────

// Full evaluation: Jastrow log from all pairs, k-space densities, determinant
// inverses by Gauss-Jordan (TrialWavefunction::evaluate_batch,
// trial_wavefunction.cpp:19-35).
template <typename T, bool CPLX>
__global__ __launch_bounds__(64) void k_evaluate(const DevModel* __restrict__ Mp, DevState S) {
    const DevModel& M = *Mp;
    extern __shared__ double lds_[];
    Shared sh = carve(M, lds_);
    const int w = blockIdx.x, lane = threadIdx.x;
    Walker<T, CPLX> wk(M, S, sh, w);
    const double* p = S.pos + static_cast<size_t>(w) * M.np * 2;
    for (int i = lane; i < 2 * M.np; i += W) sh.pos[i] = p[i];
    if (lane == 0) S.status[w] = 0;
    __syncthreads();
    double lj = 0.0;
    for (int i = lane; i < M.np; i += W) {
        const int gi = M.group_of[i];
        for (int k = i + 1; k < M.np; ++k) {
            const int gk = M.group_of[k];
            const int nf = M.pair_list_n[gi][gk];
            if (nf == 0) continue;
            double d0 = sh.pos[2 * i] - sh.pos[2 * k], d1 = sh.pos[2 * i + 1] - sh.pos[2 * k + 1];
            const double r = dist(M.cell, d0, d1);
            for (int fi = 0; fi < nf; ++fi) {
                double u, du, d2u;
                radial_eval(M.pair[M.pair_list[gi][gk][fi]].rad, r, u, du, d2u);
                lj += u;
            }
        }
        if (M.has_nosanow && M.nos_alpha[i] != 0.0) {
            double d0 = sh.pos[2 * i] - M.nos_site[i][0], d1 = sh.pos[2 * i + 1] - M.nos_site[i][1];
            if (M.cell.periodic) min_image(M.cell, d0, d1);
            lj += -M.nos_alpha[i] * d0 * d0 - M.nos_alpha[i] * d1 * d1;
        }
    }
    // k-space densities and their contribution
    double lk = 0.0;
    for (int k = lane; k < M.nkj; k += W) {
        for (int g = 0; g < M.ng; ++g) {
            double rr = 0.0, ri = 0.0;
            if (M.kj_active[g])
                for (int i = M.g_first[g]; i < M.g_last[g]; ++i) {
                    double s, c;
                    sincos_bounded(M.kjx[k] * sh.pos[2 * i] + M.kjy[k] * sh.pos[2 * i + 1], &s, &c);
                    rr += c;
                    ri += s;
                }
            sh.rho[(g * M.nkj + k) * 2] = rr;
            sh.rho[(g * M.nkj + k) * 2 + 1] = ri;
        }
        for (int g = 0; g < M.ng; ++g)
            for (int h = g; h < M.ng; ++h) {
                const double wgh = M.kjw[g][h][k];
                if (wgh == 0.0) continue;
                const double ar = sh.rho[(g * M.nkj + k) * 2], ai = sh.rho[(g * M.nkj + k) * 2 + 1];
                const double br = sh.rho[(h * M.nkj + k) * 2], bi = sh.rho[(h * M.nkj + k) * 2 + 1];
                lk += (g == h ? 1.0 : 2.0) * wgh * (ar * br + ai * bi);
            }
    }
    wk.log_j = wsum(lj) + wsum(lk);
    wk.log_b0 = wk.log_b1 = 0.0;   // set by k_refresh
    wk.store();
}

// The end of a batched rebuild: the inverse (held as the eliminated tile a,
// rows and columns in pivot order) written to the walker's record in the
// sweep's precision, the row sums of |M| (from part), the condition
// estimate, the cleared update count and log|det|.
template <typename T, bool CPLX, int NP, int GX>
__device__ __attribute__((always_inline)) inline void refresh_store(const DevModel& M, const DevState& S, int w, int b, int n, int tid,
                                     const double (&a)[NP / 16][NP / GX][CPLX ? 2 : 1], double (*part)[NP],
                                     const int* perm, const int* pinv, double (*red)[GX * 16 / W], double log_det) {
    constexpr int GY = 16;
    constexpr int TR = NP / GY, TC = NP / GX;
    constexpr int E = CPLX ? 2 : 1;
    constexpr int NT = GX * GY, NWAVES = NT / W;
    const int tx = tid / GY, ty = tid % GY;
    // column sums of |A| (a permutation of M's columns' sums), max over the
    // real columns; |re| + |im| as in the sweep's update
    double colmax = 0.0;
#pragma unroll
    for (int j = 0; j < TC; ++j) {
        double cs = 0.0;
#pragma unroll
        for (int i = 0; i < TR; ++i) cs += fabs(a[i][j][0]) + (CPLX ? fabs(a[i][j][E - 1]) : 0.0);
        for (int o = GY / 2; o > 0; o >>= 1) cs += __shfl_xor(cs, o, GY);
        if (tx * TC + j < n) colmax = fmax(colmax, cs);
    }
    T* out = static_cast<T*>(S.inv) + static_cast<size_t>(w) * S.inv_stride;
    for (int i = 0; i < b; ++i) out += E * M.blk_rows[i] * M.blk_n[i];
#pragma unroll
    for (int i = 0; i < TR; ++i) {
        const int r = ty * TR + i;
        if (r >= n) continue;
        const int m = pinv[r];
        if (m >= n) continue;   // cannot happen for a non-singular block
#pragma unroll
        for (int j = 0; j < TC; ++j) {
            const int k = tx * TC + j;
            if (k >= n || perm[k] >= n) continue;
            const size_t at = (static_cast<size_t>(m) * n + perm[k]) * E;
            out[at] = static_cast<T>(a[i][j][0]);
            if (CPLX) out[at + 1] = static_cast<T>(a[i][j][E - 1]);
        }
    }
    for (int idx = n * n + tid; idx < M.blk_rows[b] * n; idx += NT) {   // zero padding rows
        out[E * idx] = T(0);
        if (CPLX) out[E * idx + 1] = T(0);
    }
    double rowmax = 0.0;
    if (tid < n) {
        double rs = 0.0;
        for (int q = 0; q < GX; ++q) rs += part[q][tid];
        rowmax = rs;
    }
    colmax = wmax(colmax);
    rowmax = wmax(rowmax);
    if ((tid & (W - 1)) == 0) {
        red[0][tid / W] = colmax;
        red[1][tid / W] = rowmax;
    }
    __syncthreads();
    if (tid == 0) {
        for (int q = 1; q < NWAVES; ++q) {
            colmax = fmax(colmax, red[0][q]);
            rowmax = fmax(rowmax, red[1][q]);
        }
        store_fresh_cond(S, w, b, rowmax * colmax);
        S.updates[w * kMaxBlocks + b] = 0;
        S.logs[w * 4 + 1 + b] = fmax(log_det, log(1e-300));
    }
}

template <int CTRL>
__device__ inline void pivot_combine(double& best, int& bi) {
    const double ob = dpp_mov<CTRL>(best);
    const int oi = __builtin_amdgcn_update_dpp(bi, bi, CTRL, 0xF, 0xF, false);
    if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
    }
}

// Up to 64 orbitals the tile fits 128 registers: four waves per SIMD hide the
// two barriers of every elimination step better than three.
template <typename T, bool CPLX, int NP, int GX>
__global__ __launch_bounds__(GX * 16) __attribute__((amdgpu_waves_per_eu(NP <= 64 ? 4 : 1, 8))) void k_refresh_reg(const DevModel* __restrict__ Mp, DevState S, int min_updates) {
    constexpr int GY = 16;                 // owners of a column: 16 adjacent lanes
    constexpr int TR = NP / GY, TC = NP / GX;   // tile: TR rows x TC columns per thread
    constexpr int E = CPLX ? 2 : 1;
    constexpr int NWORDS = (NP + 63) / 64;
    constexpr int NT = GX * GY, NWAVES = NT / W;
    const DevModel& M = *Mp;
    __shared__ double prow[2][NP * E], fcol[2][NP * E], pval[2][2];
    __shared__ double part[GX][NP];
    __shared__ double best_sh[2], red[2][NWAVES];
    __shared__ int piv_sh[2], perm[NP], pinv[NP];
    const int nbk = M.n_blocks;
    const int w = blockIdx.x / nbk, b = blockIdx.x % nbk;
    // selective refresh: only blocks whose update count reached min_updates
    // (the whole workgroup reads the same count and leaves together)
    if (min_updates > 0 && S.updates[w * kMaxBlocks + b] < min_updates) return;
    const int tid = threadIdx.x, tx = tid / GY, ty = tid % GY;
    const int n = M.blk_n[b], first = M.blk_first[b];
    const double* pos = S.pos + static_cast<size_t>(w) * M.np * 2;

    double a[TR][TC][E];
#pragma unroll
    for (int i = 0; i < TR; ++i) {
        const int r = ty * TR + i;
        double rs = 0.0;
#pragma unroll
        for (int j = 0; j < TC; ++j) {
            const int c = tx * TC + j;
            if (r < n && c < n) {
                Orb o;
                orbital(M, b, c, pos[2 * (first + r)], pos[2 * (first + r) + 1], o);
                a[i][j][0] = o.vr;
                if (CPLX) a[i][j][E - 1] = o.vi;
                rs += CPLX ? sqrt(o.vr * o.vr + o.vi * o.vi) : fabs(o.vr);
            } else {
                a[i][j][0] = r == c ? 1.0 : 0.0;
                if (CPLX) a[i][j][E - 1] = 0.0;
            }
        }
        part[tx][r] = rs;
    }

    unsigned long long used[NWORDS];
#pragma unroll
    for (int q = 0; q < NWORDS; ++q) used[q] = 0ull;
    // log|det| as a running mantissa and binary exponent (one log at the end)
    double det_m = 1.0;
    int det_e = 0;
    bool singular = false;
    // Step k eliminates column k = kb*TC + kj; the TC steps of a column
    // block are unrolled so that column kj of the tile is a static register.
    // Only the n real columns are eliminated: the padding rows and columns
    // are identity and stay so.
    const int n_kb = (n + TC - 1) / TC;
    for (int kb = 0; kb < n_kb && !singular; ++kb) {
#pragma unroll
        for (int kj = 0; kj < TC; ++kj) {
            const int k = kb * TC + kj;
            if (k >= n) break;
            const int buf = k & 1;
            if (tx == kb) {
                double best = -1.0;
                int bi = NP;
#pragma unroll
                for (int i = 0; i < TR; ++i) {
                    const int r = ty * TR + i;
                    if ((used[r >> 6] >> (r & 63)) & 1ull) continue;
                    // |z|^2 orders the candidates as |z| does, without a sqrt
                    const double re = a[i][kj][0], im = CPLX ? a[i][kj][E - 1] : 0.0;
                    const double v = CPLX ? re * re + im * im : fabs(re);
                    if (v > best) {
                        best = v;
                        bi = r;
                    }
                }
                // the GY = 16 owners of column k are one DPP row: combine
                // (value, first row) pairs in registers
                pivot_combine<0xB1>(best, bi);
                pivot_combine<0x4E>(best, bi);
                pivot_combine<0x141>(best, bi);
                pivot_combine<0x140>(best, bi);
                if (ty == 0) {
                    piv_sh[buf] = bi;
                    best_sh[buf] = best;
                }
            }
            __syncthreads();
            const int p = piv_sh[buf];
            const double best = best_sh[buf];
            if (!(best > 0.0) || !isfinite(best) || p >= NP) {   // uniform: every thread reads the same
                singular = true;
                break;
            }
            {
                int e;
                det_m = frexp(det_m * best, &e);
                det_e += e;
            }
            // The pivot row goes to LDS with 1 in column k, its value to
            // pval, and its owners zero it in registers; the owners of column
            // k publish it with -1 in the pivot row and zero it.  One update
            // a -= f_r * (prow / pivot) then gives every row its new values:
            // the pivot row 0 + prow/pivot, column k -f_r/pivot (1/pivot in
            // the pivot row), the rest the rank-1 elimination.  LDS slots are
            // ordered by owner thread (column j*GX + tx, row i*GY + ty), so a
            // wavefront's reads fall in distinct banks.
            if (ty == p / TR) {
#pragma unroll
                for (int i = 0; i < TR; ++i)
                    if (i == p % TR) {
#pragma unroll
                        for (int j = 0; j < TC; ++j) {
                            const bool kk = j == kj && tx == kb;
                            if (kk) {
                                pval[buf][0] = a[i][j][0];
                                pval[buf][1] = CPLX ? a[i][j][E - 1] : 0.0;
                            }
                            prow[buf][(j * GX + tx) * E] = kk ? 1.0 : a[i][j][0];
                            if (CPLX) prow[buf][(j * GX + tx) * E + E - 1] = kk ? 0.0 : a[i][j][E - 1];
#pragma unroll
                            for (int q = 0; q < E; ++q) a[i][j][q] = 0.0;
                        }
                    }
            }
            if (tx == kb) {
#pragma unroll
                for (int i = 0; i < TR; ++i) {
                    const bool pr = ty * TR + i == p;
                    fcol[buf][(i * GY + ty) * E] = pr ? -1.0 : a[i][kj][0];
                    if (CPLX) fcol[buf][(i * GY + ty) * E + E - 1] = pr ? 0.0 : a[i][kj][E - 1];
#pragma unroll
                    for (int q = 0; q < E; ++q) a[i][kj][q] = 0.0;
                }
            }
            __syncthreads();
            const double dr = pval[buf][0], di = pval[buf][1];
            double ir, ii;
            if (CPLX) {
                const double inv_d2 = 1.0 / (dr * dr + di * di);
                ir = dr * inv_d2;
                ii = -di * inv_d2;
            } else {
                ir = 1.0 / dr;
                ii = 0.0;
            }
            double nr[TC][E];
#pragma unroll
            for (int j = 0; j < TC; ++j) {
                const double xr = prow[buf][(j * GX + tx) * E];
                const double xi = CPLX ? prow[buf][(j * GX + tx) * E + E - 1] : 0.0;
                nr[j][0] = CPLX ? xr * ir - xi * ii : xr * ir;
                if (CPLX) nr[j][E - 1] = xr * ii + xi * ir;
            }
#pragma unroll
            for (int i = 0; i < TR; ++i) {
                const double fr = fcol[buf][(i * GY + ty) * E], fi = CPLX ? fcol[buf][(i * GY + ty) * E + E - 1] : 0.0;
#pragma unroll
                for (int j = 0; j < TC; ++j) {   // a -= f * nr, four FMAs per complex entry
                    a[i][j][0] = fmadd(-fr, nr[j][0], a[i][j][0]);
                    if constexpr (CPLX) {
                        a[i][j][0] = fmadd(fi, nr[j][E - 1], a[i][j][0]);
                        a[i][j][E - 1] = fmadd(-fr, nr[j][E - 1], a[i][j][E - 1]);
                        a[i][j][E - 1] = fmadd(-fi, nr[j][0], a[i][j][E - 1]);
                    }
                }
            }
            used[p >> 6] |= 1ull << (p & 63);
            if (tid == 0) {
                perm[k] = p;
                pinv[p] = k;
            }
        }
    }
    const double log_det = (CPLX ? 0.5 : 1.0) * (log(det_m) + det_e * 0.69314718055994530942);
    if (singular) {
        if (tid == 0) S.status[w] = 1;
        return;
    }
    __syncthreads();   // perm, pinv
    refresh_store<T, CPLX, NP, GX>(M, S, w, b, n, tid, a, part, perm, pinv, red, log_det);
}

// ── the batched inverse recompute on the matrix cores ──────────────────────
//
// Blocks of up to 64 orbitals: the implicit-pivot Gauss-Jordan exchange of
// k_refresh_reg, blocked by 16 columns.  Four waves hold the 64 x 64 tile as
// four column panels in the FP64 MFMA accumulator layout.  For each panel in
// turn its owner wave sweeps the panel's pivots with one row per lane (pivot
// search by a wave max and a ballot, the pivot row through LDS, no workgroup
// barrier), and the other three waves apply the panel's exchange to their
// columns as one rank-16 product on v_mfma_f64_16x16x4f64:
//     X <- X + P' Y,   P' = P - sum_s e_{p_s} e_s^T,   Y = rows p_s of X,
// where P is the swept panel and p_s the pivot row of its step s; the -1 at
// (p_s, s) removes the old pivot rows, which the exchange replaces.
//
// v_mfma_f64_16x16x4f64 lane layout (lane l, checked on gfx950): A[i][k] at
// l = i + 16 k, B[k][j] at l = j + 16 k, D[i][j] at l = j + 16 (i % 4) in
// register i / 4.
typedef double dbl4 __attribute__((ext_vector_type(4)));

// Sum over the 16 lanes of one DPP row, left in every lane of the row.
__device__ inline double row16_sum(double v) {
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    v += dpp_mov<0x141>(v);
    v += dpp_mov<0x140>(v);
    return v;
}

// Largest unsigned value over the wave, as a uniform scalar: DPP row
// maxima, then the four rows' maxima.
__device__ inline unsigned wave_max_u32(unsigned v) {
    v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0xB1, 0xF, 0xF, false)));
    v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x4E, 0xF, 0xF, false)));
    v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x141, 0xF, 0xF, false)));
    v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x140, 0xF, 0xF, false)));
    const unsigned a = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 0));
    const unsigned b = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 16));
    const unsigned c = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 32));
    const unsigned d = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 48));
    return max(max(a, b), max(c, d));
}

// Partial pivoting over one wave (lane = row): the first row, among those
// not yet pivots, of the largest |z|^2 (|z|^2 orders as |z| does, without a
// sqrt), as a uniform lane index.  Two integer maxima over the IEEE bits of
// |z|^2 (non-negative, so they order as the values): the high words, then,
// among the rows that hold the largest high word, the low words with the
// row in their last six bits (ties within 2^-46 go to the first row).
__device__ inline int pivot_row(double re, double im, unsigned long long used, int lane) {
    const double val = re * re + im * im;
    const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(val));
    const bool free_row = !((used >> lane) & 1ull);
    const unsigned hi = free_row ? (static_cast<unsigned>(bits >> 32) | 0x80000000u) : 0u;
    const unsigned top = wave_max_u32(hi);
    const unsigned lo = hi == top ? ((static_cast<unsigned>(bits) & ~63u) | static_cast<unsigned>(63 - lane)) : 0u;
    return 63 - static_cast<int>(wave_max_u32(lo) & 63u);
}

// Orders the LDS accesses of one wave (no other wave takes part).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, bool CPLX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_refresh_mfma(const DevModel* __restrict__ Mp, DevState S, int min_updates) {
    constexpr int NP = 64, PB = 8, E = CPLX ? 2 : 1, NXI = CPLX ? 4 : 1;
    const DevModel& M = *Mp;
    // rows padded by one entry: a lane per row then reads across the banks
    __shared__ double pan[NP][17][E];       // the panel owner's 16-column tile, row-major
    __shared__ double ybuf[PB][NP + 1][E];  // the panel's pivot rows (before its exchange), every column
    __shared__ double part[4][NP];       // |D| row sums, per wave
    __shared__ double red[2][4];
    __shared__ double det_sh[2];         // log|det| as running mantissa and binary exponent
    __shared__ int perm[NP], pinv[NP], piv[PB], flag;
    const int nbk = M.n_blocks;
    const int w = blockIdx.x / nbk, b = blockIdx.x % nbk;
    // selective refresh: the whole workgroup reads the same count and leaves together
    if (min_updates > 0 && S.updates[w * kMaxBlocks + b] < min_updates) return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & (W - 1), j = lane & 15, g = lane >> 4;
    const int n = M.blk_n[b], first = M.blk_first[b];
    const int col = 16 * wave + j;       // this lane's column of the tile
    const double* pos = S.pos + static_cast<size_t>(w) * M.np * 2;

    // D (identity past n): register v of row tile rt holds row 16 rt + 4 v + g
    dbl4 xr[4], xi[NXI];
    double av[4][4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = 16 * rt + 4 * v + g;
            double vr = r == col ? 1.0 : 0.0, vi = 0.0, a = 0.0;
            if (r < n && col < n) {
                Orb o;
                orbital(M, b, col, pos[2 * (first + r)], pos[2 * (first + r) + 1], o);
                vr = o.vr;
                vi = o.vi;
                a = CPLX ? sqrt(vr * vr + vi * vi) : fabs(vr);
            }
            xr[rt][v] = vr;
            if constexpr (CPLX) xi[rt][v] = vi;
            av[rt][v] = a;
        }
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int v = 0; v < 4; ++v) av[rt][v] = row16_sum(av[rt][v]);
    if (j == 0) {
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int v = 0; v < 4; ++v) part[wave][16 * rt + 4 * v + g] = av[rt][v];
    }
    if (tid == 0) {
        det_sh[0] = 1.0;
        det_sh[1] = 0.0;
        flag = 0;
    }
    __syncthreads();

    unsigned long long used = 0ull;   // rows already pivots (uniform)
    const int n_panels = (n + PB - 1) / PB;
    for (int t = 0; t < n_panels; ++t) {
        const int owner = t / (16 / PB), c0 = PB * (t % (16 / PB));   // the panel's wave and its first tile column
        if (wave == owner) {
            // the whole tile to LDS (its registers are dead during the
            // steps), then one row of the panel per lane
#pragma unroll
            for (int rt = 0; rt < 4; ++rt)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = 16 * rt + 4 * v + g;
                    pan[r][j][0] = xr[rt][v];
                    if constexpr (CPLX) pan[r][j][E - 1] = xi[rt][v];
                }
            wave_sync();
            double pr[PB][E];   // row `lane` of the panel
#pragma unroll
            for (int c = 0; c < PB; ++c)
#pragma unroll
                for (int q = 0; q < E; ++q) pr[c][q] = pan[lane][c0 + c][q];
            double det_m = det_sh[0];
            int det_e = static_cast<int>(det_sh[1]);
            bool singular = false;
            // PB steps in every panel, past n too: the padding is the
            // identity, whose pivots are exact ones that change nothing, and
            // no real row is free by then.  The steps are pipelined: a step
            // updates the next pivot column first and starts the next pivot
            // search, which overlaps its updates of the other columns.  The
            // pivot row reaches the lanes by readlane (uniform p), so a step
            // has no LDS round trip and no branch.
            int p = pivot_row(pr[0][0], CPLX ? pr[0][E - 1] : 0.0, used, lane);
#pragma unroll
            for (int s = 0; s < PB; ++s) {
                const double re = pr[s][0], im = CPLX ? pr[s][E - 1] : 0.0;
                const double dr = lane_value(re, p), di = CPLX ? lane_value(im, p) : 0.0;
                const double d2 = CPLX ? dr * dr + di * di : fabs(dr);
                singular = singular || !(d2 > 0.0) || !isfinite(d2);
                {
                    int e;
                    det_m = frexp(det_m * d2, &e);
                    det_e += e;
                }
                double ir, ii;
                if (CPLX) {
                    const double inv_d2 = rcp_refined(d2);
                    ir = dr * inv_d2;
                    ii = -di * inv_d2;
                } else {
                    ir = rcp_refined(dr);
                    ii = 0.0;
                }
                // Every column c != s takes x - h y_c, y = the pivot row:
                // h = x_s / d off the pivot row gives the elimination, and
                // h = 1 - 1/d on it (where x = y) gives y / d; column s
                // becomes -h off the pivot row and 1/d on it.
                const bool is_p = lane == p;
                const double hr = is_p ? 1.0 - ir : (CPLX ? re * ir - im * ii : re * ir);
                const double hi = CPLX ? (is_p ? -ii : re * ii + im * ir) : 0.0;
                auto update = [&](int c) {
                    if (c == s) {
                        pr[c][0] = is_p ? ir : -hr;
                        if constexpr (CPLX) pr[c][E - 1] = is_p ? ii : -hi;
                        return;
                    }
                    const double yr = lane_value(pr[c][0], p);
                    double xr_ = fmadd(-hr, yr, pr[c][0]);
                    if constexpr (CPLX) {
                        const double yi = lane_value(pr[c][E - 1], p);
                        xr_ = fmadd(hi, yi, xr_);
                        double xi_ = fmadd(-hr, yi, pr[c][E - 1]);
                        xi_ = fmadd(-hi, yr, xi_);
                        pr[c][E - 1] = xi_;
                    }
                    pr[c][0] = xr_;
                };
                used |= 1ull << p;
                if (lane == 0) {
                    piv[s] = p;
                    perm[PB * t + s] = p;
                    pinv[p] = PB * t + s;
                }
                int p_next = 0;
                if (s + 1 < PB) {
                    update(s + 1);
                    p_next = pivot_row(pr[s + 1][0], CPLX ? pr[s + 1][E - 1] : 0.0, used, lane);
                }
#pragma unroll
                for (int c = 0; c < PB; ++c)
                    if (c != s + 1) update(c);
                p = p_next;
                __builtin_amdgcn_sched_barrier(0);   // one step at a time in registers
            }
#pragma unroll
            for (int c = 0; c < PB; ++c)
#pragma unroll
                for (int q = 0; q < E; ++q) pan[lane][c0 + c][q] = pr[c][q];
            wave_sync();
            // the tile back, with the swept panel in it
#pragma unroll
            for (int rt = 0; rt < 4; ++rt)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = 16 * rt + 4 * v + g;
                    xr[rt][v] = pan[r][j][0];
                    if constexpr (CPLX) xi[rt][v] = pan[r][j][E - 1];
                }
            if (lane == 0) {
                det_sh[0] = det_m;
                det_sh[1] = static_cast<double>(det_e);
                if (singular) flag = 1;
            }
        }
        __syncthreads();
        if (flag) break;   // uniform: read after the barrier
        unsigned long long pm = 0ull;
        for (int s = 0; s < PB; ++s) pm |= 1ull << piv[s];
        used |= pm;
        // Y: this wave's columns of the pivot rows, before the exchange (the
        // owner's panel columns, already final, take no update: B = 0 there)
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 16 * rt + 4 * v + g;
                if ((pm >> r) & 1ull) {
                    const int s = pinv[r] - PB * t;
                    ybuf[s][col][0] = xr[rt][v];
                    if constexpr (CPLX) ybuf[s][col][E - 1] = xi[rt][v];
                }
            }
        __syncthreads();
        {
            // X += P' Y over the panel's PB pivots (K = PB in steps of 4)
            const bool own = wave == owner && (j - c0) >= 0 && (j - c0) < PB;
#pragma unroll
            for (int ks = 0; ks < PB / 4; ++ks) {
                const int kk = 4 * ks + g;   // this lane's k in both operands
                const double br = own ? 0.0 : ybuf[kk][col][0];
                const double bi = (CPLX && !own) ? ybuf[kk][col][E - 1] : 0.0;
                const int pk = piv[kk];
#pragma unroll
                for (int rt = 0; rt < 4; ++rt) {
                    const int ra = 16 * rt + j;   // this lane's row of A
                    double ar = pan[ra][c0 + kk][0];
                    const double ai = CPLX ? pan[ra][c0 + kk][E - 1] : 0.0;
                    if (ra == pk) ar -= 1.0;
                    xr[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, xr[rt], 0, 0, 0);
                    if constexpr (CPLX) {
                        xr[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi, xr[rt], 0, 0, 0);
                        xi[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, xi[rt], 0, 0, 0);
                        xi[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, xi[rt], 0, 0, 0);
                    }
                }
            }
        }
        __syncthreads();   // pan, ybuf and piv are rewritten by the next panel
    }
    if (flag) {
        if (tid == 0) S.status[w] = 1;
        return;
    }
    const double log_det = (CPLX ? 0.5 : 1.0) * (log(det_sh[0]) + det_sh[1] * 0.69314718055994530942);

    // the inverse: A[pinv[r]][perm[c]] = X[r][c], in the sweep's precision
    T* out = static_cast<T*>(S.inv) + static_cast<size_t>(w) * S.inv_stride;
    for (int i = 0; i < b; ++i) out += E * M.blk_rows[i] * M.blk_n[i];
    double cs = 0.0;   // column sums of |X| (a permutation of A's), |re| + |im| as in the sweep
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = 16 * rt + 4 * v + g;
            cs += fabs(xr[rt][v]) + (CPLX ? fabs(xi[rt][v]) : 0.0);
            if (r < n && col < n) {
                const size_t at = (static_cast<size_t>(pinv[r]) * n + perm[col]) * E;
                out[at] = static_cast<T>(xr[rt][v]);
                if constexpr (CPLX) out[at + 1] = static_cast<T>(xi[rt][v]);
            }
        }
    cs += __shfl_xor(cs, 16, W);
    cs += __shfl_xor(cs, 32, W);
    double colmax = col < n ? cs : 0.0;
    for (int idx = n * n + tid; idx < M.blk_rows[b] * n; idx += 256) {   // zero padding rows
        out[E * idx] = T(0);
        if (CPLX) out[E * idx + 1] = T(0);
    }
    double rowmax = 0.0;
    if (tid < n) {
        const double rs = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
        rowmax = rs;
    }
    colmax = wmax(colmax);
    rowmax = wmax(rowmax);
    if (lane == 0) {
        red[0][wave] = colmax;
        red[1][wave] = rowmax;
    }
    __syncthreads();
    if (tid == 0) {
        for (int q = 1; q < 4; ++q) {
            colmax = fmax(colmax, red[0][q]);
            rowmax = fmax(rowmax, red[1][q]);
        }
        store_fresh_cond(S, w, b, rowmax * colmax);
        S.updates[w * kMaxBlocks + b] = 0;
        S.logs[w * 4 + 1 + b] = fmax(log_det, log(1e-300));
    }
}

// k_sweep's body for one walker w of a given system, for the fused launch
// below.  k_sweep keeps its own copy: compiled through this function its
// register allocation changes (more spills in the 64-row tile).
template <typename T, bool CPLX, bool DMC, int NR>
__device__ __attribute__((always_inline)) inline void sweep_walker(const DevModel* __restrict__ Mp, const DevState& S,
                                                                   const SweepParams& P, int w) {
    const DevModel& M = *Mp;
    extern __shared__ double lds_[];
    Shared sh = carve(M, lds_);
    const int lane = threadIdx.x;
    Walker<T, CPLX, NR> wk(M, S, sh, w);
    wk.load();
    long long nacc = 0;
    double rr_p = 0.0, rr_a = 0.0;
    const int np = M.np;
    // The random numbers come from a tape, [sweep][p][3][stride]: the
    // caller's, or the one k_draw wrote from the walker's own stream.  The
    // next move's three are loaded one move ahead, so the wait for them
    // overlaps the current move.
    const int ts = P.tape_stride;
    auto tape_row = [&](int s, int p) { return P.tape + (static_cast<size_t>(s) * np + p) * 3 * ts; };
    double nxt0 = tape_row(0, 0)[w], nxt1 = tape_row(0, 0)[ts + w], nxtu = tape_row(0, 0)[2 * ts + w];
    for (int s = 0; s < P.n_sweeps; ++s) {
        for (int p = 0; p < np; ++p) {
            const double xi0 = nxt0, xi1 = nxt1, uni = nxtu;
            {
                const bool more = p + 1 < np || s + 1 < P.n_sweeps;
                const double* row = more ? (p + 1 < np ? tape_row(s, p + 1) : tape_row(s + 1, 0)) : tape_row(s, p);
                nxt0 = row[w];
                nxt1 = row[ts + w];
                nxtu = row[2 * ts + w];
            }
            const double tm = P.tau * M.inv_mass[p];
            const double sigma = sqrt(tm);
            const double x0 = wk.sh.pos[2 * p], x1 = wk.sh.pos[2 * p + 1];
            double gold[2] = {0.0, 0.0}, scale = 0.0;
            // the moved particle's block into registers once per move, before
            // the old gradient and the ratio read it (one copy of the
            // spill-and-fill code, which the two reads would otherwise each
            // inline)
            if constexpr (NR > 0) {
                const int bp = M.block_of[p];
                if (bp >= 0) wk.make_resident(bp);
            }
            if constexpr (DMC) {
                wk.grad_of(p, gold);
                // DMC: the Umrigar-scaled drift; Langevin VMC: the bare tau/m
                const double v2 = gold[0] * gold[0] + gold[1] * gold[1];
                scale = (P.dmc == kSweepLangevin || v2 < 1e-10) ? tm : (-1.0 + sqrt(1.0 + 2.0 * tm * v2)) / v2;
            }
            const double nx0 = x0 + scale * gold[0] + sigma * xi0;
            const double nx1 = x1 + scale * gold[1] + sigma * xi1;
            if constexpr (DMC) {
                const double d0 = nx0 - x0, d1 = nx1 - x1;
                rr_p += tm * d0 * d0 + tm * d1 * d1;
            }
            const bool accepted =
                wk.template move<DMC>(p, nx0, nx1, uni, 0, nullptr, P.tau, DMC ? P.dmc : kSweepVmc, scale, gold, rr_a);
            if (accepted) ++nacc;
        }
    }
    wk.spill();
    wk.store();
    if (lane == 0) {
        S.n_accepted[w] += nacc;
        if (DMC && P.dmc == kSweepDmc) {
            S.rr[w * 2] += rr_p;
            S.rr[w * 2 + 1] += rr_a;
        }
    }
}

// Fused sweeps: every particle of the walker, n_sweeps times.  NR > 0 holds
// the block being swept in registers (see Walker).
template <typename T, bool CPLX, bool DMC, int NR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_sweep(const DevModel* __restrict__ Mp, DevState S,
                                              SweepParams P) {
    const DevModel& M = *Mp;
    extern __shared__ double lds_[];
    Shared sh = carve(M, lds_);
    const int w = blockIdx.x, lane = threadIdx.x;
    Walker<T, CPLX, NR> wk(M, S, sh, w);
    wk.load();
    long long nacc = 0;
    double rr_p = 0.0, rr_a = 0.0;
    const int np = M.np;
    // The random numbers come from a tape, [sweep][p][3][stride]: the
    // caller's, or the one k_draw wrote from the walker's own stream.  The
    // next move's three are loaded one move ahead, so the wait for them
    // overlaps the current move.
    const int ts = P.tape_stride;
    auto tape_row = [&](int s, int p) { return P.tape + (static_cast<size_t>(s) * np + p) * 3 * ts; };
    double nxt0 = tape_row(0, 0)[w], nxt1 = tape_row(0, 0)[ts + w], nxtu = tape_row(0, 0)[2 * ts + w];
    for (int s = 0; s < P.n_sweeps; ++s) {
        for (int p = 0; p < np; ++p) {
            const double xi0 = nxt0, xi1 = nxt1, uni = nxtu;
            {
                const bool more = p + 1 < np || s + 1 < P.n_sweeps;
                const double* row = more ? (p + 1 < np ? tape_row(s, p + 1) : tape_row(s + 1, 0)) : tape_row(s, p);
                nxt0 = row[w];
                nxt1 = row[ts + w];
                nxtu = row[2 * ts + w];
            }
            const double tm = P.tau * M.inv_mass[p];
            const double sigma = sqrt(tm);
            const double x0 = wk.sh.pos[2 * p], x1 = wk.sh.pos[2 * p + 1];
            double gold[2] = {0.0, 0.0}, scale = 0.0;
            // the moved particle's block into registers once per move, before
            // the old gradient and the ratio read it (one copy of the
            // spill-and-fill code, which the two reads would otherwise each
            // inline)
            if constexpr (NR > 0) {
                const int bp = M.block_of[p];
                if (bp >= 0) wk.make_resident(bp);
            }
            if constexpr (DMC) {
                wk.grad_of(p, gold);
                // DMC: the Umrigar-scaled drift; Langevin VMC: the bare tau/m
                const double v2 = gold[0] * gold[0] + gold[1] * gold[1];
                scale = (P.dmc == kSweepLangevin || v2 < 1e-10) ? tm : (-1.0 + sqrt(1.0 + 2.0 * tm * v2)) / v2;
            }
            const double nx0 = x0 + scale * gold[0] + sigma * xi0;
            const double nx1 = x1 + scale * gold[1] + sigma * xi1;
            if constexpr (DMC) {
                const double d0 = nx0 - x0, d1 = nx1 - x1;
                rr_p += tm * d0 * d0 + tm * d1 * d1;
            }
            const bool accepted =
                wk.template move<DMC>(p, nx0, nx1, uni, 0, nullptr, P.tau, DMC ? P.dmc : kSweepVmc, scale, gold, rr_a);
            if (accepted) ++nacc;
        }
    }
    wk.spill();
    wk.store();
    if (lane == 0) {
        S.n_accepted[w] += nacc;
        if (DMC && P.dmc == kSweepDmc) {
            S.rr[w * 2] += rr_p;
            S.rr[w * 2 + 1] += rr_a;
        }
    }
}

// Fused VMC sweeps of several systems that share the determinant precision,
// the complex form and the register tile: workgroup g runs walker
// g - jobs[j].first of system j, with system j's model, state and
// parameters, exactly as k_sweep would run it alone.
template <typename T, bool CPLX, int NR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_sweep_fused(
    const FusedArgs A) {
    const int g = blockIdx.x;
    int j = 0;
    while (j + 1 < A.n && g >= A.job[j + 1].first) ++j;
    const FusedJob& J = A.job[j];
    if (g - J.first >= J.nw) return;   // cannot happen: the grid is the sum of the walker counts
    sweep_walker<T, CPLX, false, NR>(J.m, J.s, J.p, g - J.first);
}

// One sweep of random numbers per walker from its xoshiro256++ stream, in
// the order the sweep consumes them (draw_sweep_randoms, sweep_blocking.h:
// 39-48, per walker): for each particle two Gaussians, then a uniform.
// out: [p][3][stride]; one thread per walker.
__global__ __launch_bounds__(64) void k_draw(DevState S, int np, int nw, double* __restrict__ out,
                                             int stride) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    Rng rng;
    for (int i = 0; i < 4; ++i) rng.s[i] = S.rng[w * 4 + i];
    for (int p = 0; p < np; ++p) {
        double* row = out + static_cast<size_t>(p) * 3 * stride;
        const double g0 = rng.gaussian();
        const double g1 = rng.gaussian();
        const double u = rng.uniform();
        row[w] = g0;
        row[stride + w] = g1;
        row[2 * stride + w] = u;
    }
    for (int i = 0; i < 4; ++i) S.rng[w * 4 + i] = rng.s[i];
}

// Per-call path: ratio of particle p at call.new_pos (mode 1) or commit of
// the masked walkers (mode 2).
template <typename T, bool CPLX>
__global__ __launch_bounds__(64) void k_call(const DevModel* __restrict__ Mp, DevState S, DevCall call,
                                             int p, int mode) {
    const DevModel& M = *Mp;
    extern __shared__ double lds_[];
    Shared sh = carve(M, lds_);
    const int w = blockIdx.x;
    Walker<T, CPLX> wk(M, S, sh, w);
    wk.load();
    double rr = 0.0;
    wk.template move<true>(p, call.new_pos[w], call.new_pos[M.nw + w], 0.0, mode, &call, 0.0, kSweepVmc, 0.0,
                           nullptr, rr);
    if (mode == 2) wk.store();
}

// ── the Ewald structure factor, kept up to date per walker ───────────────
//
// rho(k) = Sum_i q_i e^{ik.r_i} for the nke Ewald vectors, held per walker in
// HBM (ew_rho) with the positions it describes (ew_snap), as the reference's
// cache of delta_potential_particle does (EwaldTerm::sync_delta_cache,
// ewald_term.cpp:675-747, ewald_term.h:217-241): a call compares the current
// positions with the snapshot; up to `limit` changed particles are folded in
// as rho += q_i (e^{ik.r_new} - e^{ik.r_old}), more (or no valid cache, or
// 4096 increments since the last full sum, the bound on rounding drift)
// recompute the full sum.  No move has to notify it, so it cannot go stale:
// a permutation of the walkers or new positions from the host show up as
// changed particles.  Lane l gets rho for k = l + 64 q in rr[q], ri[q].
constexpr int kEwSlots = (kMaxKE + 63) / 64;
constexpr int kEwRebuildEvery = 4096;
__device__ void ewald_rho(const DevModel& M, const DevState& S, int w, const double* pos, int limit,
                          double (&rr)[kEwSlots], double (&ri)[kEwSlots]) {
    const int np = M.np, nke = M.nke, lane = threadIdx.x & (W - 1);
    double* rho = S.ew_rho + static_cast<size_t>(w) * nke * 2;
    double* snap = S.ew_snap + static_cast<size_t>(w) * np * 2;
    const int inc = S.ew_inc[w];
    unsigned long long chg[kMaxParticles / W];
    int n_changed = 0;
#pragma unroll
    for (int c = 0; c < kMaxParticles / W; ++c) {
        const int i = c * W + lane;
        const bool moved = i < np && (snap[2 * i] != pos[2 * i] || snap[2 * i + 1] != pos[2 * i + 1]);
        chg[c] = __ballot(moved);
        n_changed += __popcll(chg[c]);
    }
    const bool built = inc >= 0;
    const bool incremental = built && n_changed <= limit && inc + n_changed < kEwRebuildEvery;
#pragma unroll
    for (int q = 0; q < kEwSlots; ++q) {
        rr[q] = ri[q] = 0.0;
        const int k = lane + W * q;
        if (incremental && k < nke) {
            rr[q] = rho[2 * k];
            ri[q] = rho[2 * k + 1];
        }
    }
    if (built && n_changed == 0) return;
    if (incremental) {
#pragma unroll
        for (int c = 0; c < kMaxParticles / W; ++c) {
            for (unsigned long long m = chg[c]; m; m &= m - 1) {
                const int i = c * W + __builtin_ctzll(m);
                const double qi = M.charge[i];
                if (qi == 0.0) continue;
#pragma unroll
                for (int q = 0; q < kEwSlots; ++q) {
                    const int k = lane + W * q;
                    if (k >= nke) continue;
                    double so, co, sn, cn;
                    sincos_bounded(M.kex[k] * snap[2 * i] + M.key[k] * snap[2 * i + 1], &so, &co);
                    sincos_bounded(M.kex[k] * pos[2 * i] + M.key[k] * pos[2 * i + 1], &sn, &cn);
                    rr[q] += qi * (cn - co);
                    ri[q] += qi * (sn - so);
                }
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < kEwSlots; ++q) {
            const int k = lane + W * q;
            if (k >= nke) continue;
            for (int i = 0; i < np; ++i) {
                const double qi = M.charge[i];
                if (qi == 0.0) continue;
                double sn, cn;
                sincos_bounded(M.kex[k] * pos[2 * i] + M.key[k] * pos[2 * i + 1], &sn, &cn);
                rr[q] += qi * cn;
                ri[q] += qi * sn;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kEwSlots; ++q) {
        const int k = lane + W * q;
        if (k < nke) {
            rho[2 * k] = rr[q];
            rho[2 * k + 1] = ri[q];
        }
    }
    for (int i = lane; i < np; i += W) {
        snap[2 * i] = pos[2 * i];
        snap[2 * i + 1] = pos[2 * i + 1];
    }
    if (lane == 0) S.ew_inc[w] = incremental ? inc + n_changed : 0;
}

// Local energy: every particle's gradient and Laplacian from the current
// state, kinetic (+ fixed-phase terms), Ewald, open Coulomb.
template <typename T, bool CPLX>
__global__ __launch_bounds__(64) void k_measure(const DevModel* __restrict__ Mp, DevState S) {
    const DevModel& M = *Mp;
    extern __shared__ double lds_[];
    Shared sh = carve(M, lds_);
    const int w = blockIdx.x, lane = threadIdx.x;
    Walker<T, CPLX> wk(M, S, sh, w);
    wk.load();
    const int np = M.np, nw = M.nw;
    double kin = 0.0, kin_im = 0.0;
    for (int i = lane; i < np; i += W) {
        const int gi = M.group_of[i];
        const double x0 = sh.pos[2 * i], x1 = sh.pos[2 * i + 1];
        double g[2] = {0, 0}, l[2] = {0, 0}, pg[2] = {0, 0}, pl[2] = {0, 0};
        for (int k = 0; k < np; ++k) {
            if (k == i) continue;
            const int nf = M.pair_list_n[gi][M.group_of[k]];
            if (nf == 0) continue;
            double d0 = x0 - sh.pos[2 * k], d1 = x1 - sh.pos[2 * k + 1];
            const double r = dist(M.cell, d0, d1);
            for (int fi = 0; fi < nf; ++fi) {
                double u, du, d2u, g0, g1, l0, l1;
                radial_eval(M.pair[M.pair_list[gi][M.group_of[k]][fi]].rad, r, u, du, d2u);
                pair_terms(du, d2u, r, d0, d1, g0, g1, l0, l1);
                g[0] += g0; g[1] += g1; l[0] += l0; l[1] += l1;
            }
        }
        if (M.nkj > 0 && M.kj_active[gi])
            for (int k = 0; k < M.nkj; ++k) {
                double s, c;
                sincos_bounded(M.kjx[k] * x0 + M.kjy[k] * x1, &s, &c);
                double gk = 0.0, lk = 0.0;
                for (int h = 0; h < M.ng; ++h) {
                    const double wgh = M.kjw[gi][h][k];
                    if (wgh == 0.0 || !M.kj_active[h]) continue;
                    const double rr = sh.rho[(h * M.nkj + k) * 2], ri = sh.rho[(h * M.nkj + k) * 2 + 1];
                    gk += 2.0 * wgh * (c * ri - s * rr);
                    lk += 2.0 * wgh * ((h == gi ? 1.0 : 0.0) - (c * rr + s * ri));
                }
                g[0] += M.kjx[k] * gk; g[1] += M.kjy[k] * gk;
                l[0] += M.kjx[k] * M.kjx[k] * lk; l[1] += M.kjy[k] * M.kjy[k] * lk;
            }
        if (M.has_nosanow && M.nos_alpha[i] != 0.0) {
            const double a = M.nos_alpha[i];
            double d0 = x0 - M.nos_site[i][0], d1 = x1 - M.nos_site[i][1];
            if (M.cell.periodic) min_image(M.cell, d0, d1);
            g[0] += -2.0 * a * d0; g[1] += -2.0 * a * d1;
            l[0] += -2.0 * a; l[1] += -2.0 * a;
        }
        const int b = M.block_of[i];
        if (b >= 0) {
            const InvRef<T, CPLX> A = wk.inv(b);
            const int n = M.blk_n[b], kr = i - M.blk_first[b];
            double G[2][2] = {{0, 0}, {0, 0}}, L[2][2] = {{0, 0}, {0, 0}};
            for (int m = 0; m < n; ++m) {
                const C2<T> a = A.get(m, kr);
                const double ar = a.re, ai = a.im;
                Orb o;
                orbital(M, b, m, x0, x1, o);
                for (int d = 0; d < 2; ++d) {
                    G[d][0] += ar * o.gr[d] - ai * o.gi[d];
                    G[d][1] += ar * o.gi[d] + ai * o.gr[d];
                    L[d][0] += ar * o.lr[d] - ai * o.li[d];
                    L[d][1] += ar * o.li[d] + ai * o.lr[d];
                }
            }
            for (int d = 0; d < 2; ++d) {
                g[d] += G[d][0];
                l[d] += L[d][0] - (G[d][0] * G[d][0] - G[d][1] * G[d][1]);
                pg[d] += G[d][1];
                pl[d] += L[d][1] - 2.0 * G[d][0] * G[d][1];
            }
        }
        const double im = M.inv_mass[i];
        for (int d = 0; d < 2; ++d) {
            kin += -0.5 * im * (l[d] + g[d] * g[d]);
            if (CPLX) {
                kin += 0.5 * im * pg[d] * pg[d];
                kin_im += -0.5 * im * (pl[d] + 2.0 * g[d] * pg[d]);
            }
            if (S.grad_out) {
                S.grad_out[(i * 2 + d) * nw + w] = g[d];
                S.lap_out[(i * 2 + d) * nw + w] = l[d];
                S.pgrad_out[(i * 2 + d) * nw + w] = pg[d];
                S.plap_out[(i * 2 + d) * nw + w] = pl[d];
            }
        }
    }
    double pot = 0.0;
    if (M.ewald) {
        for (int i = lane; i < np; i += W)
            for (int k = i + 1; k < np; ++k) {
                const double coeff = M.charge[i] * M.charge[k];
                if (coeff == 0.0) continue;
                double d0 = sh.pos[2 * i] - sh.pos[2 * k], d1 = sh.pos[2 * i + 1] - sh.pos[2 * k + 1];
                const double r = dist(M.cell, d0, d1);
                if (r < M.ew_rcut) {
                    double f, df, d2f;
                    spline_eval(M.ew_spline, M.ew_knots, M.ew_rcut, M.ew_inv_delta, r, f, df, d2f);
                    pot += coeff * (M.ew_inv_eps / r - f);
                }
            }
        // the reciprocal sum from the walker's kept structure factor, folding
        // in the particles that moved since it was last brought up to date
        // while that is cheaper than the full sum (two phases per moved
        // particle against one per particle)
        double rr[kEwSlots], ri[kEwSlots];
        ewald_rho(M, S, w, sh.pos, np / 2, rr, ri);
#pragma unroll
        for (int q = 0; q < kEwSlots; ++q) {
            const int k = lane + W * q;
            if (k < M.nke) pot += M.kew[k] * (rr[q] * rr[q] + ri[q] * ri[q]);
        }
    }
    for (int cp = 0; cp < M.n_coul; ++cp) {
        const int ga = M.coul_ga[cp], gb = M.coul_gb[cp];
        const bool intra = ga == gb;
        for (int i = M.g_first[ga] + lane; i < M.g_last[ga]; i += W)
            for (int k = intra ? i + 1 : M.g_first[gb]; k < M.g_last[gb]; ++k) {
                double d0 = sh.pos[2 * i] - sh.pos[2 * k], d1 = sh.pos[2 * i + 1] - sh.pos[2 * k + 1];
                const double r = sqrt(d0 * d0 + d1 * d1);
                pot += M.coul_coeff[cp] / r;
            }
    }
    // InversePowerTerm::evaluate_potential (inverse_power.cpp:130-168)
    for (int t = 0; t < M.n_ipow; ++t) {
        const int ga = M.ipow_ga[t], gb = M.ipow_gb[t];
        const bool intra = ga == gb;
        const double c = M.ipow_c[t];
        const int pi = M.ipow_int[t];
        for (int i = M.g_first[ga] + lane; i < M.g_last[ga]; i += W)
            for (int k = intra ? i + 1 : M.g_first[gb]; k < M.g_last[gb]; ++k) {
                double d0 = sh.pos[2 * i] - sh.pos[2 * k], d1 = sh.pos[2 * i + 1] - sh.pos[2 * k + 1];
                const double r = dist(M.cell, d0, d1);
                if (pi > 0) {
                    const double inv_r = 1.0 / r;
                    double inv_rp = inv_r;
                    for (int q = 1; q < pi; ++q) inv_rp *= inv_r;
                    pot += c * inv_rp;
                } else {
                    pot += c * pow(r, -M.ipow_p[t]);
                }
            }
    }
    const double e = wsum(kin) + wsum(pot) + (M.ewald ? M.ew_const : 0.0) + M.ipow_tail;
    const double ei = wsum(kin_im);
    if (lane == 0) {
        S.e_loc[w] = e;
        S.e_loc_im[w] = ei;
    }
}

// Hamiltonian::delta_potential_particle (hamiltonian.cpp:105-117) for the
// walkers of the batch: the potential change of moving particle p of walker
// w to prop (wrapped), nothing committed.  Ewald
// (EwaldTerm::delta_potential_particle, ewald_term.cpp:750-815): the moved
// particle's real-space pairs, and the reciprocal part
// Sum_k w_k [2 Re(delta rho*) + |delta|^2], delta = q_p (e^{ik.r'} - e^{ik.r}),
// against the walker's kept structure factor (up to four moved particles
// folded in, as the reference's cache does).  Open Coulomb and inverse
// powers: the moved particle's pairs (radial_particle_delta,
// hamiltonian_component.h:25-48).  The constants cancel.
__global__ __launch_bounds__(64) void k_delta_potential(const DevModel* __restrict__ Mp, DevState S, int p,
                                                        const double* __restrict__ prop, int stride,
                                                        double* __restrict__ out) {
    const DevModel& M = *Mp;
    __shared__ double pos[2 * kMaxParticles];
    const int w = blockIdx.x, lane = threadIdx.x;
    const int np = M.np;
    const double* src = S.pos + static_cast<size_t>(w) * np * 2;
    for (int i = lane; i < 2 * np; i += W) pos[i] = src[i];
    __syncthreads();
    double c0 = prop[w], c1 = prop[stride + w];
    wrap(M.cell, c0, c1);
    const double o0 = pos[2 * p], o1 = pos[2 * p + 1];
    const int gp = M.group_of[p];
    double delta = 0.0;
    if (M.ewald) {
        const double qp = M.charge[p];
        if (qp != 0.0) {
            auto short_range = [&](double r) {
                if (r >= M.ew_rcut) return 0.0;
                double f, df, d2f;
                spline_eval(M.ew_spline, M.ew_knots, M.ew_rcut, M.ew_inv_delta, r, f, df, d2f);
                return M.ew_inv_eps / r - f;
            };
            for (int j = lane; j < np; j += W) {
                const double qj = M.charge[j];
                if (j == p || qj == 0.0) continue;
                double n0 = c0 - pos[2 * j], n1 = c1 - pos[2 * j + 1];
                double d0 = o0 - pos[2 * j], d1 = o1 - pos[2 * j + 1];
                const double rn = dist(M.cell, n0, n1), ro = dist(M.cell, d0, d1);
                delta += qp * qj * (short_range(rn) - short_range(ro));
            }
            double rr[kEwSlots], ri[kEwSlots];
            ewald_rho(M, S, w, pos, 4, rr, ri);
#pragma unroll
            for (int q = 0; q < kEwSlots; ++q) {
                const int k = lane + W * q;
                if (k >= M.nke) continue;
                double so, co, sn, cn;
                sincos_bounded(M.kex[k] * o0 + M.key[k] * o1, &so, &co);
                sincos_bounded(M.kex[k] * c0 + M.key[k] * c1, &sn, &cn);
                const double dr = qp * (cn - co), di = qp * (sn - so);
                delta += M.kew[k] * (2.0 * (dr * rr[q] + di * ri[q]) + dr * dr + di * di);
            }
        }
    }
    // the pairs of p in a term over groups (ga, gb)
    auto pairs = [&](int ga, int gb, auto radial) {
        auto one = [&](int j) {
            if (j == p) return;
            double n0 = c0 - pos[2 * j], n1 = c1 - pos[2 * j + 1];
            double d0 = o0 - pos[2 * j], d1 = o1 - pos[2 * j + 1];
            delta += radial(dist(M.cell, n0, n1)) - radial(dist(M.cell, d0, d1));
        };
        if (ga == gb) {
            if (gp == ga)
                for (int j = M.g_first[gb] + lane; j < M.g_last[gb]; j += W) one(j);
        } else {
            if (gp == ga)
                for (int j = M.g_first[gb] + lane; j < M.g_last[gb]; j += W) one(j);
            if (gp == gb)
                for (int j = M.g_first[ga] + lane; j < M.g_last[ga]; j += W) one(j);
        }
    };
    for (int cp = 0; cp < M.n_coul; ++cp) {
        const double coeff = M.coul_coeff[cp];
        pairs(M.coul_ga[cp], M.coul_gb[cp], [&](double r) { return coeff / r; });
    }
    for (int t = 0; t < M.n_ipow; ++t) {
        const double c = M.ipow_c[t], pw = M.ipow_p[t];
        const int pi = M.ipow_int[t];
        pairs(M.ipow_ga[t], M.ipow_gb[t], [&](double r) {
            if (pi > 0) {
                const double inv_r = 1.0 / r;
                double inv_rp = inv_r;
                for (int q = 1; q < pi; ++q) inv_rp *= inv_r;
                return c * inv_rp;
            }
            return c * pow(r, -pw);
        });
    }
    delta = wsum(delta);
    if (lane == 0) out[w] = delta;
}

// ── optimizer derivatives ──────────────────────────────────────────────────
// f = du/dp of a pair factor's parameter at distance r, with f' and f''.
// Padé (PadeJastrow, pade_jastrow.cpp:641-700, 806-880): du/da_k = r^e/Q
// (e = k+1), du/db_k = -P r^e/Q^2; summed over images the short-range part
// takes f - limit - spline(r) inside r_cut (periodic_pair_sum.cpp:85-98).
// Drummond (drummond_jastrow.cpp:201-420), inside the cutoff, with
// t = r - L and g = t^C: du/dalpha_0 = g (1 + C r/L) (alpha_1 follows
// alpha_0 through the cusp), du/dalpha_l = g r^l, du/dL = -C t^(C-1) P +
// g r dalpha_1/dL.
__device__ inline void param_radial(const DevRadial& R, const DevParam& q, const double* __restrict__ spl, double r,
                                    double& f, double& f1, double& f2) {
    f = f1 = f2 = 0.0;
    auto ip = [](double x, int e) {
        double y = 1.0;
        for (int i = 0; i < e; ++i) y *= x;
        return y;
    };
    if (R.kind == 1) {
        if (r >= R.L) return;
        const int C = R.C;
        const double Cd = static_cast<double>(C), t = r - R.L;
        const double g = ip(t, C), g1 = Cd * ip(t, C - 1), g2 = Cd * (C - 1) * ip(t, C - 2);
        if (q.index == 0) {
            const double k = Cd / R.L, m = 1.0 + k * r;
            f = g * m;
            f1 = g1 * m + g * k;
            f2 = g2 * m + 2.0 * g1 * k;
        } else if (q.index < R.num) {
            const int l = q.index + 1;
            const double m = ip(r, l), m1 = l * ip(r, l - 1), m2 = l * (l - 1) * ip(r, l - 2);
            f = g * m;
            f1 = g1 * m + g * m1;
            f2 = g2 * m + 2.0 * g1 * m1 + g * m2;
        } else {
            double P = R.c[0] + R.alpha1 * r, P1 = R.alpha1, P2 = 0.0;
            for (int k = 1; k < R.num; ++k) {
                const int l = k + 1;
                P += R.c[k] * ip(r, l);
                P1 += R.c[k] * l * ip(r, l - 1);
                P2 += R.c[k] * l * (l - 1) * ip(r, l - 2);
            }
            const double h = -Cd * ip(t, C - 1), h1 = -Cd * (C - 1) * ip(t, C - 2),
                         h2 = C >= 3 ? -Cd * (C - 1) * (C - 2) * ip(t, C - 3) : 0.0;
            const double a = q.da1_dl;
            f = h * P + g * a * r;
            f1 = h1 * P + h * P1 + g1 * a * r + g * a;
            f2 = h2 * P + 2.0 * h1 * P1 + h * P2 + g2 * a * r + 2.0 * g1 * a;
        }
        return;
    }
    double P = 0.0, P1 = 0.0, P2 = 0.0, Q = 1.0, Q1 = 0.0, Q2 = 0.0;
    {
        double rk2 = 0.0, rk1 = 1.0;   // r^(k-2), r^(k-1) for power k
        for (int i = 0; i < max(R.num, R.den); ++i) {
            const int k = i + 1;
            const double rk = rk1 * r;
            if (i < R.num) { P += R.c[i] * rk; P1 += R.c[i] * k * rk1; P2 += R.c[i] * k * (k - 1) * rk2; }
            if (i < R.den) {
                const double c = R.c[R.num + i];
                Q += c * rk; Q1 += c * k * rk1; Q2 += c * k * (k - 1) * rk2;
            }
            rk2 = rk1;
            rk1 = rk;
        }
    }
    const bool numer = q.index < R.num;
    const int e = numer ? q.index + 1 : q.index - R.num + 1;
    const double m = ip(r, e), m1 = e * ip(r, e - 1), m2 = e * (e - 1) * ip(r, e > 1 ? e - 2 : 0);
    const double iq = 1.0 / Q, iq2 = iq * iq, iq3 = iq2 * iq;
    if (numer) {
        f = m * iq;
        f1 = m1 * iq - m * Q1 * iq2;
        f2 = m2 * iq - 2.0 * m1 * Q1 * iq2 - m * Q2 * iq2 + 2.0 * m * Q1 * Q1 * iq3;
    } else {
        const double h = P * m, h1 = P1 * m + P * m1, h2 = P2 * m + 2.0 * P1 * m1 + P * m2;
        f = -h * iq2;
        f1 = -h1 * iq2 + 2.0 * h * Q1 * iq3;
        f2 = -h2 * iq2 + 4.0 * h1 * Q1 * iq3 + 2.0 * h * Q2 * iq3 - 6.0 * h * Q1 * Q1 * iq3 * iq;
    }
    if (R.image) {
        if (r >= q.r_cut) { f = f1 = f2 = 0.0; return; }
        double v, d1, d2;
        spline_eval(spl + q.sofs, q.n_knots, q.r_cut, q.inv_delta, r, v, d1, d2);
        f -= q.limit + v;
        f1 -= d1;
        f2 -= d2;
    }
}

// d phi / d alpha of a Gaussian orbital at separation (d0, d1) from its
// centre (GaussianOrbitalSet::evaluate_parameter_derivatives,
// orbital_set.h:249-289): -r^2 phi; of its gradient -2 delta_d phi
// (1 - alpha r^2); of its second derivatives phi [8 alpha delta_d^2 - 2 -
// r^2 (4 alpha^2 delta_d^2 - 2 alpha)].  Also the orbital's own gradient and
// second derivatives.
struct GaussAlpha {
    double v, ag[2], al[2], g[2], l[2];
};
__device__ inline GaussAlpha gauss_alpha(double a, double d0, double d1) {
    GaussAlpha o;
    const double r2 = d0 * d0 + d1 * d1, phi = exp(-a * r2);
    o.v = -r2 * phi;
    const double del[2] = {d0, d1};
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        o.ag[d] = -2.0 * del[d] * phi * (1.0 - a * r2);
        o.al[d] = phi * (8.0 * a * del[d] * del[d] - 2.0 - r2 * (4.0 * a * a * del[d] * del[d] - 2.0 * a));
        o.g[d] = phi * (-2.0 * a * del[d]);
        o.l[d] = phi * (4.0 * a * a * del[d] * del[d] - 2.0 * a);
    }
    return o;
}

// The walker's block b inverse, A(m, c) at [m * n + c].
template <typename T, bool CPLX>
__device__ inline const T* block_inverse(const DevModel& M, const DevState& S, int w, int b) {
    const T* base = static_cast<const T*>(S.inv) + static_cast<size_t>(w) * S.inv_stride;
    for (int i = 0; i < b; ++i) base += (CPLX ? 2 : 1) * M.blk_rows[i] * M.blk_n[i];
    return base;
}

// TrialWavefunction::compute_parameter_derivatives (trial_wavefunction.cpp:
// 150-160) for the free parameters listed in prm: out[i * stride + w] =
// d log|psi| / d p_i of walker w.  One wave per walker.  A pair factor sums
// du/dp (param_radial) over its pairs, lanes over the first particle, plus,
// summed over images, the reciprocal part sum_k w_k |rho_a|^2 (one group) or
// 2 w_k Re(rho_a rho_b*) with the parameter's weights
// (kspace_pair_sum.cpp:303-313), lanes over k.  Nosanow
// (nosanow_jastrow.cpp:176-190): -sum_i |r_i - R_i|^2.  A Gaussian block's
// shared width (slater_determinant.cpp:98-157): sum_{c,m} A(m,c)
// dphi_m(r_c)/dalpha, lanes over the block's particles (real blocks only).
template <typename T, bool CPLX>
__global__ __launch_bounds__(64) void k_param_derivs(const DevModel* __restrict__ Mp, DevState S,
                                                     const DevParam* __restrict__ prm, int n_prm,
                                                     const double* __restrict__ wts, const double* __restrict__ spl,
                                                     double* __restrict__ out, int stride) {
    const DevModel& M = *Mp;
    __shared__ double pos[2 * kMaxParticles];
    const int w = blockIdx.x, lane = threadIdx.x;
    const int np = M.np;
    const double* src = S.pos + static_cast<size_t>(w) * np * 2;
    for (int i = lane; i < 2 * np; i += W) pos[i] = src[i];
    __syncthreads();
    const double* rho = S.rho + static_cast<size_t>(w) * M.ng * M.nkj * 2;
    for (int i = 0; i < n_prm; ++i) {
        const DevParam& q = prm[i];
        double t = 0.0;
        if (q.pair >= 0) {
            const DevPair& F = M.pair[q.pair];
            const bool same = F.ga == F.gb;
            for (int a = M.g_first[F.ga] + lane; a < M.g_last[F.ga]; a += W)
                for (int b = same ? a + 1 : M.g_first[F.gb]; b < M.g_last[F.gb]; ++b) {
                    double d0 = pos[2 * a] - pos[2 * b], d1 = pos[2 * a + 1] - pos[2 * b + 1];
                    double f, f1, f2;
                    param_radial(F.rad, q, spl, dist(M.cell, d0, d1), f, f1, f2);
                    t += f;
                }
            if (F.rad.image)
                for (int k = lane; k < M.nkj; k += W) {
                    const double ar = rho[(F.ga * M.nkj + k) * 2], ai = rho[(F.ga * M.nkj + k) * 2 + 1];
                    const double br = rho[(F.gb * M.nkj + k) * 2], bi = rho[(F.gb * M.nkj + k) * 2 + 1];
                    t += (same ? 1.0 : 2.0) * wts[q.wofs + k] * (ar * br + ai * bi);
                }
        } else if (q.group >= 0) {
            for (int p = M.g_first[q.group] + lane; p < M.g_last[q.group]; p += W) {
                double d0 = pos[2 * p] - M.nos_site[p][0], d1 = pos[2 * p + 1] - M.nos_site[p][1];
                if (M.cell.periodic) min_image(M.cell, d0, d1);
                t -= d0 * d0 + d1 * d1;
            }
        } else if constexpr (!CPLX) {
            const int b = q.block, n = M.blk_n[b], first = M.blk_first[b];
            const T* A = block_inverse<T, CPLX>(M, S, w, b);
            for (int c = lane; c < n; c += W)
                for (int m = 0; m < n; ++m) {
                    double d0 = pos[2 * (first + c)] - M.orb_k[b][m][0];
                    double d1 = pos[2 * (first + c) + 1] - M.orb_k[b][m][1];
                    if (M.cell.periodic) min_image(M.cell, d0, d1);
                    const double r2 = d0 * d0 + d1 * d1;
                    t += static_cast<double>(A[m * n + c]) * (-r2 * exp(-q.alpha * r2));
                }
        }
        t = wsum(t);
        if (lane == 0) out[static_cast<size_t>(i) * stride + w] = t;
    }
}

// Hamiltonian::local_energy_parameter_derivatives (hamiltonian.cpp:189-196)
// over TrialWavefunction::compute_parameter_gradient_derivatives
// (trial_wavefunction.cpp:162-185) for the pair and one-body parameters of
// prm (a determinant's rows are left to k_gauss_eloc): every particle's
// dG and dL of log|psi| with respect to p, reduced on the spot by the kinetic
// term (kinetic.cpp:176-200, 212-250):
//   dE/dp = sum_{i,d} -1/2 (1/m_i) [dL + 2 G dG],  dE_im/dp = sum -1/2 (1/m_i) 2 dG v,
// G and v the whole trial's gradient and phase gradient ([i][d][nw], from
// k_measure).  A pair parameter's f = du/dp enters as u does (pade_jastrow.
// cpp:806-880): dG_i += f' delta/r, dL_i += f'' (delta_d/r)^2 + (f'/r)(1 -
// (delta_d/r)^2); summed over images its weights enter the reciprocal
// gradient and Laplacian (kspace_pair_sum.cpp:315-320).  Nosanow
// (nosanow_jastrow.cpp:192-210): dG = -2 delta, dL = -2.  One wave per
// walker, lanes over the factor's particles, each summing over its partners.
__global__ __launch_bounds__(64) void k_param_eloc(const DevModel* __restrict__ Mp, DevState S,
                                                   const DevParam* __restrict__ prm, int n_prm,
                                                   const double* __restrict__ wts, const double* __restrict__ spl,
                                                   const double* __restrict__ grad, const double* __restrict__ pgrad,
                                                   double* __restrict__ out_re, double* __restrict__ out_im,
                                                   int stride) {
    const DevModel& M = *Mp;
    __shared__ double pos[2 * kMaxParticles];
    const int w = blockIdx.x, lane = threadIdx.x;
    const int np = M.np, nw = M.nw;
    const double* src = S.pos + static_cast<size_t>(w) * np * 2;
    for (int i = lane; i < 2 * np; i += W) pos[i] = src[i];
    __syncthreads();
    const double* rho = S.rho + static_cast<size_t>(w) * M.ng * M.nkj * 2;
    for (int i = 0; i < n_prm; ++i) {
        const DevParam& q = prm[i];
        if (q.block >= 0) continue;
        double er = 0.0, ei = 0.0;
        auto reduce = [&](int p, const double dg[2], const double dl[2]) {
            const double im = M.inv_mass[p];
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const size_t o = static_cast<size_t>(p * 2 + d) * nw + w;
                er += -0.5 * im * (dl[d] + 2.0 * grad[o] * dg[d]);
                if (pgrad) ei += -0.5 * im * (2.0 * dg[d] * pgrad[o]);
            }
        };
        if (q.pair >= 0) {
            const DevPair& F = M.pair[q.pair];
            const bool same = F.ga == F.gb;
            // particle p of group gp against the factor's other group
            auto particle = [&](int p, int h) {
                double dg[2] = {0.0, 0.0}, dl[2] = {0.0, 0.0};
                for (int j = M.g_first[h]; j < M.g_last[h]; ++j) {
                    if (j == p) continue;
                    double d0 = pos[2 * p] - pos[2 * j], d1 = pos[2 * p + 1] - pos[2 * j + 1];
                    const double r = dist(M.cell, d0, d1);
                    double f, f1, f2, g0, g1, l0, l1;
                    param_radial(F.rad, q, spl, r, f, f1, f2);
                    pair_terms(f1, f2, r, d0, d1, g0, g1, l0, l1);
                    dg[0] += g0; dg[1] += g1; dl[0] += l0; dl[1] += l1;
                }
                if (F.rad.image) {
                    const double self = same ? 1.0 : 0.0;
                    for (int k = 0; k < M.nkj; ++k) {
                        double sn, cs;
                        sincos_bounded(M.kjx[k] * pos[2 * p] + M.kjy[k] * pos[2 * p + 1], &sn, &cs);
                        const double rr = rho[(h * M.nkj + k) * 2], ri = rho[(h * M.nkj + k) * 2 + 1];
                        const double wk = wts[q.wofs + k];
                        const double gk = 2.0 * wk * (cs * ri - sn * rr);
                        const double lk = 2.0 * wk * (self - (cs * rr + sn * ri));
                        dg[0] += M.kjx[k] * gk; dg[1] += M.kjy[k] * gk;
                        dl[0] += M.kjx[k] * M.kjx[k] * lk; dl[1] += M.kjy[k] * M.kjy[k] * lk;
                    }
                }
                reduce(p, dg, dl);
            };
            for (int p = M.g_first[F.ga] + lane; p < M.g_last[F.ga]; p += W) particle(p, F.gb);
            if (!same)
                for (int p = M.g_first[F.gb] + lane; p < M.g_last[F.gb]; p += W) particle(p, F.ga);
        } else {
            for (int p = M.g_first[q.group] + lane; p < M.g_last[q.group]; p += W) {
                double d0 = pos[2 * p] - M.nos_site[p][0], d1 = pos[2 * p + 1] - M.nos_site[p][1];
                if (M.cell.periodic) min_image(M.cell, d0, d1);
                const double dg[2] = {-2.0 * d0, -2.0 * d1}, dl[2] = {-2.0, -2.0};
                reduce(p, dg, dl);
            }
        }
        er = wsum(er);
        ei = wsum(ei);
        if (lane == 0) {
            out_re[static_cast<size_t>(i) * stride + w] = er;
            out_im[static_cast<size_t>(i) * stride + w] = ei;
        }
    }
}

// dE_L / d alpha for the shared width alpha of Gaussian block b
// (SlaterDeterminant::compute_parameter_gradient_derivatives,
// slater_determinant.cpp:159-251, reduced as in k_param_eloc).  The inverse
// moves with alpha: dA = -A dD A with dD(c, m) = dphi_m(r_c)/dalpha, so
//   dG_j = sum_m dA(m,j) dphi_m(r_j) + A(m,j) d(dphi_m(r_j))/dalpha,
// the same for the second derivatives, and dL = dLraw - 2 G_det dG with the
// block's own gradient G_det.  One workgroup of four waves per walker of the
// chunk [w0, w0 + gridDim.x), both products on v_mfma_f64_16x16x4f64 over the
// block padded to n16 = 16 ceil(n/16): the waves split the n16 x n16 output
// into quarters of 16 x 16 tiles, and the k dimension runs in panels of 16
// staged in LDS.  G1: Y = dD A, dD's panel evaluated in place (every exp
// once), Y written to ybuf (n16^2 per walker of the chunk); G2: dA = -A Y,
// whose tiles meet the orbitals in the epilogue: each lane evaluates
// orbital m at particle j for its own elements of dA and adds them into
// particle j's six sums, which lanes and then the two row quarters add up.
constexpr int kGaussThreads = 256;
constexpr int kGaussPanel = 16;
template <typename T>
__global__ __launch_bounds__(kGaussThreads) __attribute__((amdgpu_waves_per_eu(2))) void k_gauss_eloc(const DevModel* __restrict__ Mp, DevState S, int b,
                                                              double alpha, const double* __restrict__ grad,
                                                              double* __restrict__ ybuf, int w0,
                                                              double* __restrict__ out) {
    const DevModel& M = *Mp;
    constexpr int KP = kGaussPanel;
    __shared__ double px[kMaxOrb], py[kMaxOrb];
    __shared__ double xp[kMaxOrb][KP + 1];      // left operand, columns k0 .. k0 + 15
    __shared__ double zp[KP][kMaxOrb + 2];      // right operand, rows k0 .. k0 + 15
    __shared__ double red[2][kMaxOrb][6];
    __shared__ double esum[kGaussThreads / W];
    const int w = w0 + blockIdx.x, t = threadIdx.x, lane = t & (W - 1), wv = t / W;
    const int n = M.blk_n[b], first = M.blk_first[b], nw = M.nw;
    const int n16 = (n + 15) & ~15, tn = n16 / 16, th = (tn + 1) / 2;
    const int r0 = (wv >> 1) * th, c0 = (wv & 1) * th;   // this wave's first row and column tile
    const int li = lane & 15, lk = lane >> 4;
    const double* src = S.pos + static_cast<size_t>(w) * M.np * 2;
    for (int c = t; c < n; c += kGaussThreads) {
        px[c] = src[2 * (first + c)];
        py[c] = src[2 * (first + c) + 1];
    }
    const T* A = block_inverse<T, false>(M, S, w, b);
    double* Y = ybuf + static_cast<size_t>(blockIdx.x) * n16 * n16;
    auto a_at = [&](int m, int c) { return m < n && c < n ? static_cast<double>(A[m * n + c]) : 0.0; };
    auto sep = [&](int c, int m, double& d0, double& d1) {
        d0 = px[c] - M.orb_k[b][m][0];
        d1 = py[c] - M.orb_k[b][m][1];
        if (M.cell.periodic) min_image(M.cell, d0, d1);
    };
    __syncthreads();
    dbl4 acc[4][4];
    // acc = X Z over the k panels, fill(k0) staging X[:, k0:k0+16] in xp and
    // Z[k0:k0+16, :] in zp
    auto gemm = [&](auto fill) {
#pragma unroll
        for (int I = 0; I < 4; ++I)
#pragma unroll
            for (int J = 0; J < 4; ++J) acc[I][J] = dbl4{0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < n16; k0 += KP) {
            fill(k0);
            __syncthreads();
#pragma unroll
            for (int ks = 0; ks < KP / 4; ++ks) {
                double xa[4], zb[4];
#pragma unroll
                for (int I = 0; I < 4; ++I) xa[I] = r0 + I < tn ? xp[(r0 + I) * 16 + li][ks * 4 + lk] : 0.0;
#pragma unroll
                for (int J = 0; J < 4; ++J) zb[J] = c0 + J < tn ? zp[ks * 4 + lk][(c0 + J) * 16 + li] : 0.0;
#pragma unroll
                for (int I = 0; I < 4; ++I)
#pragma unroll
                    for (int J = 0; J < 4; ++J)
                        if (I < th && J < th) acc[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[I], zb[J], acc[I][J], 0, 0, 0);
            }
            __syncthreads();
        }
    };
    // G1: Y = dD A
    gemm([&](int k0) {
        for (int e = t; e < n16 * KP; e += kGaussThreads) {
            const int c = e / KP, m = k0 + e % KP;
            double v = 0.0;
            if (c < n && m < n) {
                double d0, d1;
                sep(c, m, d0, d1);
                const double r2 = d0 * d0 + d1 * d1;
                v = -r2 * exp(-alpha * r2);
            }
            xp[c][e % KP] = v;
        }
        for (int e = t; e < KP * n16; e += kGaussThreads) zp[e / n16][e % n16] = a_at(k0 + e / n16, e % n16);
    });
#pragma unroll
    for (int I = 0; I < 4; ++I)
#pragma unroll
        for (int J = 0; J < 4; ++J)
            if (I < th && J < th && r0 + I < tn && c0 + J < tn)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    Y[((r0 + I) * 16 + 4 * q + lk) * n16 + (c0 + J) * 16 + li] = acc[I][J][q];
    __syncthreads();   // Y is read by the other waves
    // G2: dA = -(A Y)
    gemm([&](int k0) {
        for (int e = t; e < n16 * KP; e += kGaussThreads) xp[e / KP][e % KP] = a_at(e / KP, k0 + e % KP);
        for (int e = t; e < KP * n16; e += kGaussThreads) zp[e / n16][e % n16] = Y[(k0 + e / n16) * n16 + e % n16];
    });
    double s6[4][6];
#pragma unroll
    for (int J = 0; J < 4; ++J)
#pragma unroll
        for (int q = 0; q < 6; ++q) s6[J][q] = 0.0;
#pragma unroll
    for (int J = 0; J < 4; ++J) {
        const int j = (c0 + J) * 16 + li;
        if (J >= th || c0 + J >= tn || j >= n) continue;
#pragma unroll
        for (int I = 0; I < 4; ++I) {
            if (I >= th || r0 + I >= tn) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = (r0 + I) * 16 + 4 * q + lk;
                if (m >= n) continue;
                const double da = -acc[I][J][q], a = static_cast<double>(A[m * n + j]);
                double d0, d1;
                sep(j, m, d0, d1);
                const GaussAlpha o = gauss_alpha(alpha, d0, d1);
#pragma unroll
                for (int d = 0; d < 2; ++d) {
                    s6[J][d] += a * o.g[d];
                    s6[J][2 + d] += da * o.g[d] + a * o.ag[d];
                    s6[J][4 + d] += da * o.l[d] + a * o.al[d];
                }
            }
        }
    }
    // the four lanes of one column (li, li + 16, li + 32, li + 48), then the
    // two row quarters through LDS
#pragma unroll
    for (int J = 0; J < 4; ++J)
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            double v = s6[J][q];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            s6[J][q] = v;
        }
    if (lk == 0)
#pragma unroll
        for (int J = 0; J < 4; ++J)
            if (J < th && c0 + J < tn)
#pragma unroll
                for (int q = 0; q < 6; ++q) red[wv >> 1][(c0 + J) * 16 + li][q] = s6[J][q];
    __syncthreads();
    double e = 0.0;
    if (t < n) {
        const int p = first + t;
        const double im = M.inv_mass[p];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const double G = red[0][t][d] + (th < tn ? red[1][t][d] : 0.0);
            const double dG = red[0][t][2 + d] + (th < tn ? red[1][t][2 + d] : 0.0);
            const double dLr = red[0][t][4 + d] + (th < tn ? red[1][t][4 + d] : 0.0);
            e += -0.5 * im * (dLr - 2.0 * G * dG + 2.0 * grad[static_cast<size_t>(p * 2 + d) * nw + w] * dG);
        }
    }
    e = wsum(e);
    if (lane == 0) esum[wv] = e;
    __syncthreads();
    if (t == 0) {
        double tot = 0.0;
        for (int k = 0; k < kGaussThreads / W; ++k) tot += esum[k];
        out[w] = tot;
    }
}

// Branching gather: new walker j takes old walker src[j]'s record.
__global__ void k_gather(const double* __restrict__ src_d, double* __restrict__ dst_d, int stride,
                         const int* __restrict__ src, int n_new) {
    const int j = blockIdx.x;
    if (j >= n_new) return;
    const double* a = src_d + static_cast<size_t>(src[j]) * stride;
    double* b = dst_d + static_cast<size_t>(j) * stride;
    for (int i = threadIdx.x; i < stride; i += blockDim.x) b[i] = a[i];
}

// One drift-diffusion sweep's branch factor (dmc.cpp:404-408): weight *=
// PopulationControl::branch_weight(E_new, E_old) (population_control.cpp:
// 58-67), the cut count of is_cut (population_control.h:189-191), then
// E_old = E_new for the next sweep.
__global__ void k_dmc_weights(const double* __restrict__ e_new, double* __restrict__ e_old,
                              double* __restrict__ weight, long long* __restrict__ n_cut, int nw,
                              double etrial, double eref, double e_cut, double e_max, double half_tau) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const double en = e_new[w], eo = e_old[w];
    double ht = half_tau;
    const double x = fmax(eref - en, eref - eo);
    if (x > e_max)
        ht = 0.0;
    else if (x > e_cut)
        ht *= 1.0 - (x - e_cut) / (e_max - e_cut);
    weight[w] *= exp(ht * (2.0 * etrial - en - eo));
    if (eref - en > e_cut || eref - eo > e_cut) n_cut[w] += 1;
    e_old[w] = en;
}

// Integer multiplicities m_w = int(weight_w + xi_w) (walker_set.cpp:193-197),
// then one workgroup turns them into the copy list: an exclusive scan over
// walkers in order and the expansion src[off_w + c] = w, c < m_w, so copies
// of a walker are adjacent and walkers keep their order.  The total is
// written even when it exceeds the list's capacity (the host then takes the
// slow path).
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void k_branch(const double* __restrict__ weight,
                                                        const double* __restrict__ xi,
                                                        int* __restrict__ src, int src_cap, int nw,
                                                        long long* __restrict__ total) {
    __shared__ long long part[kScanThreads];
    const int t = threadIdx.x;
    const int per = (nw + kScanThreads - 1) / kScanThreads;
    const int w0 = min(nw, t * per), w1 = min(nw, w0 + per);
    auto mult = [&](int w) -> long long {
        const double v = weight[w] + xi[w];
        return v < 2147483647.0 ? static_cast<long long>(static_cast<int>(v)) : 2147483647LL;
    };
    long long sum = 0;
    for (int w = w0; w < w1; ++w) sum += mult(w);
    part[t] = sum;
    __syncthreads();
    // Hillis-Steele inclusive scan over the per-thread sums
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const long long v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long pos = part[t] - sum;
    if (t == kScanThreads - 1) total[0] = part[t];
    for (int w = w0; w < w1; ++w) {
        const long long m = mult(w);
        for (long long c = 0; c < m && pos + c < src_cap; ++c) src[pos + c] = w;
        pos += m;
    }
}

// Fill a per-walker double array with a value (weights reset after branching).
__global__ void k_fill(double* __restrict__ a, int n, double v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_fill_int(int* __restrict__ a, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

}  // namespace

// ── host launch wrappers ───────────────────────────────────────────────────
int set_ziggurat(const uint64_t* k, const double* w, const double* f, double r) {
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_zig_k), k, 256 * sizeof(uint64_t)) != hipSuccess) return 1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_zig_w), w, 256 * sizeof(double)) != hipSuccess) return 1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_zig_f), f, 256 * sizeof(double)) != hipSuccess) return 1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_zig_r), &r, sizeof(double)) != hipSuccess) return 1;
    return 0;
}

#define DISPATCH(FN, ...)                                                              \
    do {                                                                               \
        if (mixed) {                                                                   \
            if (cplx) hipLaunchKernelGGL((FN<float, true>), __VA_ARGS__);              \
            else hipLaunchKernelGGL((FN<float, false>), __VA_ARGS__);                  \
        } else {                                                                       \
            if (cplx) hipLaunchKernelGGL((FN<double, true>), __VA_ARGS__);             \
            else hipLaunchKernelGGL((FN<double, false>), __VA_ARGS__);                 \
        }                                                                              \
    } while (0)

void launch_evaluate(const DevModel* m, const DevState& s, int nw, bool mixed, bool cplx, size_t lds,
                     hipStream_t st) {
    DISPATCH(k_evaluate, dim3(nw), dim3(64), lds, st, m, s);
}
// The register-tiled refresh: 16 x 16 threads up to 64 orbitals, 32 x 16 up
// to 128 (the most a block can have).
template <int NP, int GX>
static void launch_refresh_reg(const DevModel* m, const DevState& s, int nw, int n_blocks, bool mixed,
                               bool cplx, int min_updates, hipStream_t st) {
    const dim3 g(nw * n_blocks), t(GX * 16);
    if (mixed) {
        if (cplx) hipLaunchKernelGGL((k_refresh_reg<float, true, NP, GX>), g, t, 0, st, m, s, min_updates);
        else hipLaunchKernelGGL((k_refresh_reg<float, false, NP, GX>), g, t, 0, st, m, s, min_updates);
    } else {
        if (cplx) hipLaunchKernelGGL((k_refresh_reg<double, true, NP, GX>), g, t, 0, st, m, s, min_updates);
        else hipLaunchKernelGGL((k_refresh_reg<double, false, NP, GX>), g, t, 0, st, m, s, min_updates);
    }
}

// Blocks of up to 128 orbitals (the most the build accepts) refresh in
// registers: 16 x 16 threads up to 64 orbitals, 32 x 16 up to 128.  With
// mfma, blocks of 33 to 64 orbitals take the blocked exchange on the matrix
// cores (k_refresh_mfma) instead.  min_updates > 0 rebuilds only the blocks
// whose update count reached it.
void launch_refresh(const DevModel* m, const DevState& s, int nw, int n_blocks, int nb_max, bool mixed,
                    bool cplx, int min_updates, hipStream_t st, bool mfma) {
    if (nw == 0 || n_blocks == 0) return;
    if (mfma && nb_max > 32 && nb_max <= 64) {
        const dim3 g(nw * n_blocks), t(256);
        if (mixed) {
            if (cplx) hipLaunchKernelGGL((k_refresh_mfma<float, true>), g, t, 0, st, m, s, min_updates);
            else hipLaunchKernelGGL((k_refresh_mfma<float, false>), g, t, 0, st, m, s, min_updates);
        } else {
            if (cplx) hipLaunchKernelGGL((k_refresh_mfma<double, true>), g, t, 0, st, m, s, min_updates);
            else hipLaunchKernelGGL((k_refresh_mfma<double, false>), g, t, 0, st, m, s, min_updates);
        }
        return;
    }
    if (nb_max <= 16)
        launch_refresh_reg<16, 16>(m, s, nw, n_blocks, mixed, cplx, min_updates, st);
    else if (nb_max <= 32)
        launch_refresh_reg<32, 16>(m, s, nw, n_blocks, mixed, cplx, min_updates, st);
    else if (nb_max <= 64)
        launch_refresh_reg<64, 16>(m, s, nw, n_blocks, mixed, cplx, min_updates, st);
    else
        launch_refresh_reg<128, 32>(m, s, nw, n_blocks, mixed, cplx, min_updates, st);   // 512 threads, 8 x 4 tiles
}
// The register tile of the sweep: the smallest NR in {16, 32, 64} that
// covers every block's padded rows, within 128 32-bit registers per lane for
// the column; larger blocks (complex FP64 past 32 rows, anything past 64)
// take the HBM path (NR = 0).
template <typename T, bool CPLX>
constexpr int kColWords = (CPLX ? 2 : 1) * static_cast<int>(sizeof(T) / 4);

int sweep_tile_rows(bool mixed, bool cplx, int rows_max) {
    if (rows_max <= 0) return 0;
    const int words = (cplx ? 2 : 1) * (mixed ? 1 : 2);
    for (int nr = 16; nr <= 64; nr *= 2)
        if (rows_max <= nr) return nr * words <= 128 ? nr : 0;
    return 0;
}

// Register tile rows of a drift (DMC or Langevin) sweep, 0 for the HBM path.
// The drift move keeps the old and new gradients and Laplacians live across
// the move besides the tile, so it takes tiles only where they stay within
// the register file (spills of at most 8 registers): mixed blocks up to 48
// rows, FP64 real blocks up to 32, FP64 complex blocks up to 16.
int drift_tile_rows(bool mixed, bool cplx, int rows_max) {
    if (rows_max <= 0) return 0;
    const int cap = mixed ? 48 : (cplx ? 16 : 32);
    for (int nr = 16; nr <= cap; nr += 16)
        if (rows_max <= nr) return nr;
    return 0;
}

template <typename T, bool CPLX, bool DMC>
static void launch_sweep_t(const DevModel* m, const DevState& s, const SweepParams& p, int nw, int nr,
                           size_t lds, hipStream_t st) {
    const dim3 g(nw), t(64);
    if constexpr (DMC) {
        switch (nr) {
            case 16: hipLaunchKernelGGL((k_sweep<T, CPLX, DMC, 16>), g, t, lds, st, m, s, p); break;
            case 32: hipLaunchKernelGGL((k_sweep<T, CPLX, DMC, 32>), g, t, lds, st, m, s, p); break;
            case 48:
                if constexpr (sizeof(T) == 4) {
                    hipLaunchKernelGGL((k_sweep<T, CPLX, DMC, 48>), g, t, lds, st, m, s, p);
                    break;
                }
                [[fallthrough]];
            default: hipLaunchKernelGGL((k_sweep<T, CPLX, DMC, 0>), g, t, lds, st, m, s, p); break;
        }
        return;
    } else {
        switch (nr) {
            case 16: hipLaunchKernelGGL((k_sweep<T, CPLX, DMC, 16>), g, t, lds, st, m, s, p); break;
            case 32: hipLaunchKernelGGL((k_sweep<T, CPLX, DMC, 32>), g, t, lds, st, m, s, p); break;
            case 64:
                if constexpr (64 * kColWords<T, CPLX> <= 128) {
                    hipLaunchKernelGGL((k_sweep<T, CPLX, DMC, 64>), g, t, lds, st, m, s, p);
                    break;
                }
                [[fallthrough]];
            default: hipLaunchKernelGGL((k_sweep<T, CPLX, DMC, 0>), g, t, lds, st, m, s, p); break;
        }
    }
}

void launch_sweep(const DevModel* m, const DevState& s, const SweepParams& p, int nw, bool mixed,
                  bool cplx, int rows_max, size_t lds, hipStream_t st) {
    int nr = sweep_tile_rows(mixed, cplx, rows_max);
    if (p.dmc) {
        nr = drift_tile_rows(mixed, cplx, rows_max);
        if (mixed) {
            if (cplx) launch_sweep_t<float, true, true>(m, s, p, nw, nr, lds, st);
            else launch_sweep_t<float, false, true>(m, s, p, nw, nr, lds, st);
        } else {
            if (cplx) launch_sweep_t<double, true, true>(m, s, p, nw, nr, lds, st);
            else launch_sweep_t<double, false, true>(m, s, p, nw, nr, lds, st);
        }
        return;
    }
    if (mixed) {
        if (cplx) launch_sweep_t<float, true, false>(m, s, p, nw, nr, lds, st);
        else launch_sweep_t<float, false, false>(m, s, p, nw, nr, lds, st);
    } else {
        if (cplx) launch_sweep_t<double, true, false>(m, s, p, nw, nr, lds, st);
        else launch_sweep_t<double, false, false>(m, s, p, nw, nr, lds, st);
    }
}
static_assert(sizeof(FusedArgs) <= 4096, "the fused job table must fit the kernel arguments");

template <typename T, bool CPLX>
static void launch_sweep_fused_t(const FusedArgs& a, int total_nw, int nr, size_t lds, hipStream_t st) {
    const dim3 g(total_nw), t(64);
    switch (nr) {
        case 16: hipLaunchKernelGGL((k_sweep_fused<T, CPLX, 16>), g, t, lds, st, a); break;
        case 32: hipLaunchKernelGGL((k_sweep_fused<T, CPLX, 32>), g, t, lds, st, a); break;
        case 64:
            if constexpr (64 * kColWords<T, CPLX> <= 128) {
                hipLaunchKernelGGL((k_sweep_fused<T, CPLX, 64>), g, t, lds, st, a);
                break;
            }
            [[fallthrough]];
        default: hipLaunchKernelGGL((k_sweep_fused<T, CPLX, 0>), g, t, lds, st, a); break;
    }
}

void launch_sweep_fused(const FusedArgs& a, int total_nw, bool mixed, bool cplx, int rows_max, size_t lds,
                        hipStream_t st) {
    const int nr = sweep_tile_rows(mixed, cplx, rows_max);
    if (mixed) {
        if (cplx) launch_sweep_fused_t<float, true>(a, total_nw, nr, lds, st);
        else launch_sweep_fused_t<float, false>(a, total_nw, nr, lds, st);
    } else {
        if (cplx) launch_sweep_fused_t<double, true>(a, total_nw, nr, lds, st);
        else launch_sweep_fused_t<double, false>(a, total_nw, nr, lds, st);
    }
}

void launch_draw(const DevState& s, int np, int nw, double* out, int stride, hipStream_t st) {
    hipLaunchKernelGGL(k_draw, dim3((nw + 63) / 64), dim3(64), 0, st, s, np, nw, out, stride);
}
void launch_call(const DevModel* m, const DevState& s, const DevCall& c, int p, int mode, int nw,
                 bool mixed, bool cplx, size_t lds, hipStream_t st) {
    DISPATCH(k_call, dim3(nw), dim3(64), lds, st, m, s, c, p, mode);
}
void launch_measure(const DevModel* m, const DevState& s, int nw, bool mixed, bool cplx, size_t lds,
                    hipStream_t st) {
    DISPATCH(k_measure, dim3(nw), dim3(64), lds, st, m, s);
}
void launch_delta_potential(const DevModel* m, const DevState& s, int n, int p, const double* prop, int stride,
                            double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_delta_potential, dim3(n), dim3(64), 0, st, m, s, p, prop, stride, out);
}
void launch_param_derivs(const DevModel* m, const DevState& s, int nw, bool mixed, bool cplx, const DevParam* prm,
                         int n_prm, const double* wts, const double* spl, double* out, int stride, hipStream_t st) {
    DISPATCH(k_param_derivs, dim3(nw), dim3(64), 0, st, m, s, prm, n_prm, wts, spl, out, stride);
}
void launch_param_eloc(const DevModel* m, const DevState& s, int nw, const DevParam* prm, int n_prm,
                       const double* wts, const double* spl, const double* grad, const double* pgrad,
                       double* out_re, double* out_im, int stride, hipStream_t st) {
    hipLaunchKernelGGL(k_param_eloc, dim3(nw), dim3(64), 0, st, m, s, prm, n_prm, wts, spl, grad, pgrad, out_re,
                       out_im, stride);
}
void launch_gauss_eloc(const DevModel* m, const DevState& s, bool mixed, int b, double alpha, const double* grad,
                       double* ybuf, int w0, int n_walkers, double* out, hipStream_t st) {
    if (mixed)
        hipLaunchKernelGGL(k_gauss_eloc<float>, dim3(n_walkers), dim3(kGaussThreads), 0, st, m, s, b, alpha, grad,
                           ybuf, w0, out);
    else
        hipLaunchKernelGGL(k_gauss_eloc<double>, dim3(n_walkers), dim3(kGaussThreads), 0, st, m, s, b, alpha, grad,
                           ybuf, w0, out);
}
void launch_gather(const double* src_d, double* dst_d, int stride, const int* src, int n_new,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_gather, dim3(n_new), dim3(256), 0, st, src_d, dst_d, stride, src, n_new);
}

void launch_dmc_weights(const double* e_new, double* e_old, double* weight, long long* n_cut, int nw,
                        double etrial, double eref, double e_cut, double e_max, double half_tau,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_dmc_weights, dim3((nw + 255) / 256), dim3(256), 0, st, e_new, e_old, weight,
                       n_cut, nw, etrial, eref, e_cut, e_max, half_tau);
}

void launch_branch(const double* weight, const double* xi, int* src, int src_cap, int nw,
                   long long* total, hipStream_t st) {
    hipLaunchKernelGGL(k_branch, dim3(1), dim3(kScanThreads), 0, st, weight, xi, src, src_cap, nw, total);
}

void launch_fill(double* a, int n, double v, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, st, a, n, v);
}
void launch_fill_int(int* a, int n, int v, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, st, a, n, v);
}

// ── Estimators: pair correlation and structure factor ─────────────────────
// One measurement of every walker: each kernel processes a contiguous chunk
// of walkers per workgroup, in walker order, and writes the chunk's weighted
// sums to its own row of `part`; k_est_reduce then adds the rows in chunk
// order to the block's running sums.  The result does not depend on
// scheduling: the same walkers give the same bits.
namespace {

constexpr int kEstThreads = 256;
constexpr int kGofrPerThread = kMaxGofrEntries / kEstThreads;

// PairCorrelation::accumulate (pair_correlation.cpp:74-105): every pair i < j
// of a walker, the minimum-image distance of r_i - r_j as the distance table
// holds it (uncontracted r^2, correctly rounded sqrt), binned by
// int(r / delta) below rmax into function pid(g_i, g_j).  A walker's counts
// are exact integers in LDS; they join the thread's partial sums times the
// walker's weight.
__global__ __launch_bounds__(kEstThreads) void k_gofr(const DevModel* __restrict__ Mp, const double* __restrict__ pos,
                                                      const double* __restrict__ weights, int nw, int chunk,
                                                      double rmax, double inv_delta, int n_bins,
                                                      double* __restrict__ part, int part_stride) {
    const DevModel& M = *Mp;
    __shared__ int cnt[kMaxGofrEntries];
    __shared__ double sp[2 * kMaxParticles];
    __shared__ int gof[kMaxParticles];
    const int np = M.np, G = M.ng, nfb = G * (G + 1) / 2 * n_bins, tid = threadIdx.x;
    for (int k = tid; k < nfb; k += kEstThreads) cnt[k] = 0;
    for (int p = tid; p < np; p += kEstThreads) gof[p] = M.group_of[p];
    double acc[kGofrPerThread];
#pragma unroll
    for (int q = 0; q < kGofrPerThread; ++q) acc[q] = 0.0;
    const int w0 = blockIdx.x * chunk, w1 = min(nw, w0 + chunk);
    const int npairs = np * (np - 1) / 2, two_np = 2 * np;
    for (int w = w0; w < w1; ++w) {
        __syncthreads();   // the previous walker's positions and counts are done with
        for (int k = tid; k < 2 * np; k += kEstThreads) sp[k] = pos[static_cast<size_t>(w) * np * 2 + k];
        __syncthreads();
        // thread tid takes the contiguous pairs [t0, t1) of the upper
        // triangle in row order: row i holds start(i) .. start(i+1)-1 with
        // start(i) = i (2 np - i - 1) / 2; the first pair is located once,
        // the rest by stepping along the row
        const int per = (npairs + kEstThreads - 1) / kEstThreads;
        const int t0 = min(npairs, tid * per), t1 = min(npairs, t0 + per);
        int i = 0, j = 1;
        if (t0 < t1) {
            const double disc = static_cast<double>(two_np - 1) * (two_np - 1) - 8.0 * t0;
            i = static_cast<int>(0.5 * ((two_np - 1) - sqrt(fmax(disc, 0.0))));
            i = min(max(i, 0), np - 2);
            while (i > 0 && i * (two_np - i - 1) / 2 > t0) --i;
            while ((i + 1) * (two_np - i - 2) / 2 <= t0) ++i;
            j = t0 - i * (two_np - i - 1) / 2 + i + 1;
        }
        for (int t = t0; t < t1; ++t) {
            double d0 = sp[2 * i] - sp[2 * j], d1 = sp[2 * i + 1] - sp[2 * j + 1];
            min_image(M.cell, d0, d1);
            const double r = sqrt(__dadd_rn(__dmul_rn(d0, d0), __dmul_rn(d1, d1)));
            if (r < rmax) {
                const int bin = static_cast<int>(r * inv_delta);
                if (bin < n_bins) {
                    int a = gof[i], b = gof[j];
                    if (b < a) { const int x = a; a = b; b = x; }
                    atomicAdd(&cnt[(a * G - a * (a - 1) / 2 + (b - a)) * n_bins + bin], 1);
                }
            }
            if (++j == np) {
                ++i;
                j = i + 1;
            }
        }
        __syncthreads();
        const double wt = weights ? weights[w] : 1.0;
#pragma unroll
        for (int q = 0; q < kGofrPerThread; ++q) {
            const int k = tid + q * kEstThreads;
            if (k < nfb) {
                acc[q] += wt * static_cast<double>(cnt[k]);
                cnt[k] = 0;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kGofrPerThread; ++q) {
        const int k = tid + q * kEstThreads;
        if (k < nfb) part[static_cast<size_t>(blockIdx.x) * part_stride + k] = acc[q];
    }
}

// StructureFactor::accumulate (structure_factor.cpp:62-105): per walker and
// group, rho_g(k) = sum_i e^{i k.r_i}; the chunk sums w rho_g and
// w Re(rho_a rho_b^*) per group pair a <= b.  One wave per chunk and tile of
// 64 vectors, lanes over k.
__global__ __launch_bounds__(64) void k_sofk(const DevModel* __restrict__ Mp, const double* __restrict__ pos,
                                             const double* __restrict__ weights, int nw, int chunk,
                                             const double* __restrict__ kvec, int nk, double* __restrict__ part,
                                             int part_stride, int off) {
    const DevModel& M = *Mp;
    const int np = M.np, G = M.ng, lane = threadIdx.x;
    const int w0 = blockIdx.x * chunk, w1 = min(nw, w0 + chunk);
    constexpr int kPairs = kMaxGroups * (kMaxGroups + 1) / 2;
    double* row = part + static_cast<size_t>(blockIdx.x) * part_stride + off;
    const int npg = G * (G + 1) / 2;
    {   // one wave per (chunk, tile of 64 vectors)
        const int k = blockIdx.y * 64 + lane;
        const double kx = k < nk ? kvec[k] : 0.0, ky = k < nk ? kvec[nk + k] : 0.0;
        double as[kPairs], ar[kMaxGroups], ai[kMaxGroups];
#pragma unroll
        for (int q = 0; q < kPairs; ++q) as[q] = 0.0;
#pragma unroll
        for (int g = 0; g < kMaxGroups; ++g) ar[g] = ai[g] = 0.0;
        for (int w = w0; w < w1; ++w) {
            const double wt = weights ? weights[w] : 1.0;
            const double* pw = pos + static_cast<size_t>(w) * np * 2;
            double re[kMaxGroups], im[kMaxGroups];
#pragma unroll
            for (int g = 0; g < kMaxGroups; ++g) {
                re[g] = im[g] = 0.0;
                if (g < G)
                    for (int i = M.g_first[g]; i < M.g_last[g]; ++i) {
                        double sn, cs;
                        sincos_bounded(kx * pw[2 * i] + ky * pw[2 * i + 1], &sn, &cs);
                        re[g] += cs;
                        im[g] += sn;
                    }
                ar[g] += wt * re[g];
                ai[g] += wt * im[g];
            }
            int q = 0;
#pragma unroll
            for (int a = 0; a < kMaxGroups; ++a)
#pragma unroll
                for (int b = a; b < kMaxGroups; ++b, ++q)
                    as[q] += wt * (re[a] * re[b] + im[a] * im[b]);
        }
        if (k < nk) {
            // pair (a, b) of G groups sits at a G - a(a-1)/2 + (b - a)
            int q = 0;
#pragma unroll
            for (int a = 0; a < kMaxGroups; ++a)
#pragma unroll
                for (int b = a; b < kMaxGroups; ++b, ++q)
                    if (b < G) row[(a * G - a * (a - 1) / 2 + (b - a)) * nk + k] = as[q];
#pragma unroll
            for (int g = 0; g < kMaxGroups; ++g)
                if (g < G) {
                    row[npg * nk + g * nk + k] = ar[g];
                    row[npg * nk + G * nk + g * nk + k] = ai[g];
                }
        }
    }
}

// Density::accumulate (density.cpp:66-117): each particle's cell of an
// n_bins x n_bins grid per group, C row-major.  A periodic cell bins the
// fractional coordinates of the wrapped position (wrap as SimulationCell::
// wrap, f_i = sum_d ainv[i][d] r_d uncontracted, bin int((f + 1/2) n_bins)
// clamped to the last bin); an open cell bins [-rmax, rmax)^2 and drops the
// particles outside; per_particle gives every particle its own grid.  A
// walker's counts are exact integers in LDS; the thread
// that opens a cell for the walker folds count x weight into the chunk's
// FP64 histogram and clears the count, walker after walker, so the sums do
// not depend on scheduling.
__global__ __launch_bounds__(kEstThreads) void k_density(const DevModel* __restrict__ Mp, const double* __restrict__ pos,
                                                         const double* __restrict__ weights, int nw, int chunk,
                                                         double rmax, double inv_delta, int n_bins, int per_particle,
                                                         double* __restrict__ part, int part_stride, int off) {
    const DevModel& M = *Mp;
    __shared__ int cnt[kMaxDensityEntries];
    __shared__ double hist[kMaxDensityEntries];
    const int np = M.np, n_cells = n_bins * n_bins, tid = threadIdx.x;
    const int nden = (per_particle ? np : M.ng) * n_cells;
    for (int k = tid; k < nden; k += kEstThreads) {
        cnt[k] = 0;
        hist[k] = 0.0;
    }
    const double two_rmax = 2.0 * rmax, nb = static_cast<double>(n_bins);
    const int w0 = blockIdx.x * chunk, w1 = min(nw, w0 + chunk);
    for (int w = w0; w < w1; ++w) {
        __syncthreads();   // the previous walker's counts are folded and cleared
        // each thread's particles: tid, tid + 256, ... (np <= kMaxParticles)
        int idx[kMaxParticles / kEstThreads + 1];
#pragma unroll
        for (int q = 0; q * kEstThreads < kMaxParticles; ++q) {
            const int p = tid + q * kEstThreads;
            idx[q] = -1;
            if (p >= np) continue;
            double r0 = pos[(static_cast<size_t>(w) * np + p) * 2], r1 = pos[(static_cast<size_t>(w) * np + p) * 2 + 1];
            int cell = 0;
            if (M.cell.periodic) {
                wrap(M.cell, r0, r1);
                const double* ai = M.cell.ainv;
                const double f0 = __dadd_rn(__dadd_rn(0.0, __dmul_rn(ai[0], r0)), __dmul_rn(ai[1], r1));
                const double f1 = __dadd_rn(__dadd_rn(0.0, __dmul_rn(ai[2], r0)), __dmul_rn(ai[3], r1));
                const int b0 = min(static_cast<int>(__dmul_rn(__dadd_rn(f0, 0.5), nb)), n_bins - 1);
                const int b1 = min(static_cast<int>(__dmul_rn(__dadd_rn(f1, 0.5), nb)), n_bins - 1);
                cell = b0 * n_bins + b1;
            } else {
                const double x0 = __dadd_rn(r0, rmax), x1 = __dadd_rn(r1, rmax);
                if (x0 < 0.0 || x0 >= two_rmax || x1 < 0.0 || x1 >= two_rmax) continue;
                const int b0 = min(static_cast<int>(__dmul_rn(x0, inv_delta)), n_bins - 1);
                const int b1 = min(static_cast<int>(__dmul_rn(x1, inv_delta)), n_bins - 1);
                cell = b0 * n_bins + b1;
            }
            const int k = (per_particle ? p : M.group_of[p]) * n_cells + cell;
            // the thread whose increment opens the cell folds it
            idx[q] = atomicAdd(&cnt[k], 1) == 0 ? k : -1;
        }
        __syncthreads();
        const double wt = weights ? weights[w] : 1.0;
#pragma unroll
        for (int q = 0; q * kEstThreads < kMaxParticles; ++q) {
            const int k = idx[q];
            if (k < 0) continue;
            hist[k] += static_cast<double>(cnt[k]) * wt;
            cnt[k] = 0;
        }
    }
    __syncthreads();
    double* row = part + static_cast<size_t>(blockIdx.x) * part_stride + off;
    for (int k = tid; k < nden; k += kEstThreads) row[k] = hist[k];
}

// acc[e] += sum over chunks c (in order) of part[c][e]; the last workgroup
// adds the walkers' weight sum (a fixed per-thread then tree order) to
// acc[n_entries].
__global__ __launch_bounds__(kEstThreads) void k_est_reduce(const double* __restrict__ part, int n_chunks,
                                                            int part_stride, int n_entries,
                                                            double* __restrict__ acc,
                                                            const double* __restrict__ weights, int nw) {
    const int tid = threadIdx.x;
    const int e = blockIdx.x * kEstThreads + tid;
    if (e < n_entries) {
        double sum = 0.0;
        for (int c = 0; c < n_chunks; ++c) sum += part[static_cast<size_t>(c) * part_stride + e];
        acc[e] += sum;
    }
    if (blockIdx.x + 1 == gridDim.x) {
        __shared__ double red[kEstThreads];
        double sum = 0.0;
        for (int w = tid; w < nw; w += kEstThreads) sum += weights ? weights[w] : 1.0;
        red[tid] = sum;
        __syncthreads();
        for (int h = kEstThreads / 2; h > 0; h >>= 1) {
            if (tid < h) red[tid] += red[tid + h];
            __syncthreads();
        }
        if (tid == 0) acc[n_entries] += red[0];
    }
}

}  // namespace

void launch_estimators(const DevModel* m, const double* pos, const double* weights, int nw, int chunk,
                       bool gofr, double rmax, double inv_delta, int n_bins, int nfb, bool sofk,
                       const double* kvec, int nk, bool density, double d_rmax, double d_inv_delta, int d_bins,
                       bool d_per_particle, int d_off, double* part, int part_stride, int n_entries, double* acc,
                       hipStream_t st) {
    const int n_chunks = (nw + chunk - 1) / chunk;
    if (density)
        hipLaunchKernelGGL(k_density, dim3(n_chunks), dim3(kEstThreads), 0, st, m, pos, weights, nw, chunk, d_rmax,
                           d_inv_delta, d_bins, d_per_particle ? 1 : 0, part, part_stride, d_off);
    if (gofr)
        hipLaunchKernelGGL(k_gofr, dim3(n_chunks), dim3(kEstThreads), 0, st, m, pos, weights, nw, chunk, rmax,
                           inv_delta, n_bins, part, part_stride);
    if (sofk)
        hipLaunchKernelGGL(k_sofk, dim3(n_chunks, (nk + 63) / 64), dim3(64), 0, st, m, pos, weights, nw, chunk,
                           kvec, nk, part, part_stride, nfb);
    const int blocks = (n_entries + kEstThreads - 1) / kEstThreads + 1;
    hipLaunchKernelGGL(k_est_reduce, dim3(blocks), dim3(kEstThreads), 0, st, part, n_chunks, part_stride,
                       n_entries, acc, weights, nw);
}

}  // namespace alea_amd
