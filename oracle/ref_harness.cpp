// Harness over the reference implementation's own hot-path sources.
//
// oracle/Makefile compiles this file together with the reference sources where
// they lie (under /root/reference/src) into oracle/_ref/libalea_ref.so.  The
// harness only SEQUENCES the reference's public calls; every number it returns
// is computed by the reference's code.  It exists to pin the oracle
// (oracle/oracle.cpp) and the HIP path against the real thing: test
// infrastructure only, never linked into the product.
//
// What it cannot build: the Slater determinants (src/wavefunction/
// slater_determinant.cpp, complex_slater_determinant.cpp) include Eigen, which
// is absent from this image, so they are left out; determinant parity is pinned
// by closed forms and brute force instead.
//
// The registration entry points mirror include/alea_amd.h with a `ref_` prefix
// so one Python driver can run the same call sequence on all three libraries.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <span>
#include <string>
#include <vector>

#include "estimators/density.h"
#include "estimators/pair_correlation.h"
#include "estimators/structure_factor.h"
#include "hamiltonian/coulomb.h"
#include "hamiltonian/ewald_term.h"
#include "hamiltonian/hamiltonian.h"
#include "hamiltonian/inverse_power.h"
#include "hamiltonian/kinetic.h"
#include "numerics/ewald_breakup.h"
#include "numerics/green_function.h"
#include "numerics/long_range_breakup.h"
#include "numerics/pair_radial.h"
#include "numerics/random.h"
#include "numerics/xoshiro_random.h"
#include "numerics/statistics.h"
#include "population/simulation_cell.h"
#include "population/walker.h"
#include "population/walker_set.h"
#include "qmc/forward_walking.h"
#include "qmc/population_control.h"
#include "qmc/run_result.h"
#include "qmc/sweep_blocking.h"
#include "wavefunction/drummond_jastrow.h"
#include "wavefunction/nosanow_jastrow.h"
#include "wavefunction/pade_jastrow.h"
#include "wavefunction/trial_wavefunction.h"

namespace {

thread_local std::string g_error;

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return 2;
    }
}

struct RefSystem {
    std::size_t nw = 0, np = 0, dim = 2;
    bool periodic = false;
    std::vector<double> lattice;
    alea::SimulationCell cell;
    std::vector<alea::ParticleGroup> groups;
    alea::TrialWavefunction psi;
    alea::Hamiltonian ham;
    std::unique_ptr<alea::WalkerSet> ws;
    bool ham_built = false;
    double ewald_tol = 0.0, ewald_eps = 1.0;
    bool use_ewald = false;
    struct CoulombPair { int a, b; double eps; };
    std::vector<CoulombPair> coulomb;
    struct InversePowerSpec { int a, b; double c, p; };
    std::vector<InversePowerSpec> ipow;
    // free (optimizable) flags for the next factors' parameters, consumed in
    // order by the add calls; parameters past the list stay fixed
    std::vector<unsigned char> free_next;
    std::size_t free_used = 0;
    bool take_free() { return free_used < free_next.size() && free_next[free_used++] != 0; }
};

alea::ParticleGroup make_group(std::size_t first, std::size_t last, double charge,
                               double inv_mass, std::size_t dim) {
    alea::ParticleGroup g;
    g.name = "g" + std::to_string(first);
    g.mass.assign(dim, 1.0 / inv_mass);
    g.inv_mass.assign(dim, inv_mass);
    g.charge = charge;
    g.first = first;
    g.last = last;
    g.count = last - first;
    return g;
}

void build_hamiltonian(RefSystem& s) {
    if (s.ham_built) return;
    s.ham.add_component(std::make_unique<alea::Kinetic>(s.groups));
    if (s.use_ewald)
        s.ham.add_component(std::make_unique<alea::EwaldTerm>(s.groups, s.ewald_eps, s.cell,
                                                              s.ewald_tol));
    for (const auto& c : s.coulomb)
        s.ham.add_component(std::make_unique<alea::CoulombTerm>(
            s.groups[static_cast<std::size_t>(c.a)], s.groups[static_cast<std::size_t>(c.b)],
            c.eps));
    for (const auto& t : s.ipow)
        s.ham.add_component(std::make_unique<alea::InversePowerTerm>(
            s.groups[static_cast<std::size_t>(t.a)], s.groups[static_cast<std::size_t>(t.b)], t.c,
            t.p, s.cell));
    s.ham_built = true;
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_error.c_str(); }

// ── RNG (src/numerics/xoshiro_random.h, src/numerics/random.h) ──────────────
uint64_t ref_stream_seed(uint64_t base, uint64_t stream, uint64_t index) {
    return alea::numerics::stream_seed(base, stream, index);
}

int ref_xoshiro_state_u64(const uint64_t* state, int n, uint64_t* out) {
    std::array<std::uint64_t, 4> st{state[0], state[1], state[2], state[3]};
    alea::numerics::Xoshiro256pp g(st);
    for (int i = 0; i < n; ++i) out[i] = g();
    return 0;
}

int ref_xoshiro_seed_u64(uint64_t seed, int n, uint64_t* out) {
    alea::numerics::Xoshiro256pp g(seed);
    for (int i = 0; i < n; ++i) out[i] = g();
    return 0;
}

// kind 0: uniform(), 1: gaussian()
int ref_xoshiro_draws(uint64_t seed, int kind, int n, double* out) {
    alea::numerics::XoshiroRandom r(seed);
    for (int i = 0; i < n; ++i) out[i] = kind == 0 ? r.uniform() : r.gaussian();
    return 0;
}

// n_sweeps consecutive draw_sweep_randoms tapes of one stream (sweep_blocking.h:39).
int ref_sweep_tape(uint64_t seed, int np, int dim, int nw, int n_sweeps, double* out) {
    alea::numerics::XoshiroRandom r(seed);
    const std::size_t len = static_cast<std::size_t>(np) * (dim + 1) * nw;
    for (int s = 0; s < n_sweeps; ++s)
        alea::draw_sweep_randoms(r, np, dim, nw, 0, nw,
                                 std::span<double>(out + s * len, len));
    return 0;
}

// ── Breakup tables (setup; src/numerics/ewald_breakup.cpp:205) ─────────────
int ref_ewald_setup(const double* lattice, double tol, double eps, int cap_k, int* nk,
                    int* n1, int* n2, double* kx, double* ky, double* weight, int cap_knots,
                    int* n_knots, double* scaled, double* scalars /* r_cut, inv_delta,
                    v_long_zero, sr_k0, k_cut, area */) {
    return guarded([&] {
        alea::SimulationCell cell(2, std::span<const double>(lattice, 4));
        const auto s = alea::numerics::build_ewald(
            cell, tol, alea::numerics::PotentialSpec::coulomb(eps));
        if (static_cast<int>(s.n_k()) > cap_k) throw std::runtime_error("cap_k too small");
        *nk = static_cast<int>(s.n_k());
        for (std::size_t k = 0; k < s.n_k(); ++k) {
            n1[k] = s.n1[k];
            n2[k] = s.n2[k];
            kx[k] = s.kx[k];
            ky[k] = s.ky[k];
            weight[k] = s.weight[k];
        }
        if (static_cast<int>(s.long_range.scaled.size()) > 3 * cap_knots)
            throw std::runtime_error("cap_knots too small");
        *n_knots = s.long_range.n_knots;
        std::memcpy(scaled, s.long_range.scaled.data(),
                    s.long_range.scaled.size() * sizeof(double));
        scalars[0] = s.long_range.r_cut;
        scalars[1] = s.long_range.inv_delta;
        scalars[2] = s.v_long_zero;
        scalars[3] = s.sr_k0;
        scalars[4] = s.k_cut;
        scalars[5] = s.area;
    });
}

// The image-sum split of a Padé u = P/Q exactly as PeriodicPairSum builds it
// (src/wavefunction/periodic_pair_sum.cpp:16-64, src/wavefunction/pade_jastrow.cpp:204-238).
int ref_pade_breakup(const double* lattice, int na, const double* a, int nb, const double* b,
                     int cap_k, int* nk, int* n1, int* n2, double* kx, double* ky,
                     double* weight, int cap_knots, int* n_knots, double* scaled,
                     double* scalars /* r_cut, inv_delta, u_inf */) {
    return guarded([&] {
        alea::SimulationCell cell(2, std::span<const double>(lattice, 4));
        const double rc = cell.ws_radius();
        std::vector<double> av(a, a + na), bv(b, b + nb);
        alea::numerics::BreakupOptions opt{};
        opt.k_continuum = 20.0 * 3.14159265358979323846 / rc;
        opt.k_max = 300.0 / rc;
        opt.n_continuum = 150;
        opt.k_cut = 5.0 * std::min(std::hypot(cell.reciprocal()[0], cell.reciprocal()[1]),
                                   std::hypot(cell.reciprocal()[2], cell.reciprocal()[3]));
        const auto br = alea::numerics::build_long_range_breakup(
            cell, alea::numerics::pade_radial(av, bv, rc), opt);
        if (static_cast<int>(br.weight().size()) > cap_k)
            throw std::runtime_error("cap_k too small");
        *nk = static_cast<int>(br.weight().size());
        for (std::size_t k = 0; k < br.weight().size(); ++k) {
            n1[k] = br.n1()[k];
            n2[k] = br.n2()[k];
            kx[k] = br.kx()[k];
            ky[k] = br.ky()[k];
            weight[k] = br.weight()[k];
        }
        const auto& sp = br.spline();
        if (static_cast<int>(sp.scaled.size()) > 3 * cap_knots)
            throw std::runtime_error("cap_knots too small");
        *n_knots = sp.n_knots;
        std::memcpy(scaled, sp.scaled.data(), sp.scaled.size() * sizeof(double));
        scalars[0] = sp.r_cut;
        scalars[1] = sp.inv_delta;
        scalars[2] = alea::numerics::pade_limit(av, bv);
    });
}

// ── System registration (mirrors include/alea_amd.h) ───────────────────────
int ref_create(void** out, int nw, int np, int dim, const double* lattice) {
    return guarded([&] {
        auto* s = new RefSystem;
        s->nw = static_cast<std::size_t>(nw);
        s->np = static_cast<std::size_t>(np);
        s->dim = static_cast<std::size_t>(dim);
        s->periodic = lattice != nullptr;
        if (s->periodic) {
            s->lattice.assign(lattice, lattice + dim * dim);
            s->cell = alea::SimulationCell(s->dim, s->lattice);
        } else {
            s->cell = alea::SimulationCell(s->dim);
        }
        s->ws = std::make_unique<alea::WalkerSet>(s->nw, s->np, s->dim, s->cell);
        *out = s;
    });
}

int ref_destroy(void* h) {
    delete static_cast<RefSystem*>(h);
    return 0;
}

int ref_add_group(void* h, int first, int last, double charge, double inv_mass) {
    auto& s = *static_cast<RefSystem*>(h);
    s.groups.push_back(make_group(static_cast<std::size_t>(first), static_cast<std::size_t>(last),
                                  charge, inv_mass, s.dim));
    return 0;
}

int ref_add_pade_jastrow(void* h, int ga, int gb, int num_order, int den_order,
                         const double* coeffs, int image_sum) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        std::vector<alea::Parameter> params;
        for (int k = 0; k < num_order + den_order; ++k)
            params.push_back(alea::Parameter{"c" + std::to_string(k), coeffs[k], !s.take_free()});
        auto c = std::make_unique<alea::PadeJastrow>(num_order, den_order,
                                                     s.groups[static_cast<std::size_t>(ga)],
                                                     s.groups[static_cast<std::size_t>(gb)],
                                                     std::move(params));
        if (image_sum) c->sum_over_images(s.cell);
        s.psi.add_component(std::move(c));
    });
}

int ref_add_drummond_jastrow(void* h, int ga, int gb, int C, double gamma, double L,
                             int n_coeffs, const double* coeffs) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        std::vector<alea::Parameter> params;
        for (int k = 0; k < n_coeffs; ++k)
            params.push_back(alea::Parameter{"a" + std::to_string(k), coeffs[k], !s.take_free()});
        params.push_back(alea::Parameter{"L", L, !s.take_free()});
        s.psi.add_component(std::make_unique<alea::DrummondJastrow>(
            C, gamma, s.groups[static_cast<std::size_t>(ga)],
            s.groups[static_cast<std::size_t>(gb)], std::move(params)));
    });
}

int ref_add_nosanow_jastrow(void* h, int g, double alpha, const double* sites) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        const auto& grp = s.groups[static_cast<std::size_t>(g)];
        std::vector<double> st(sites, sites + grp.size() * s.dim);
        std::vector<alea::Parameter> params{alea::Parameter{"alpha", alpha, !s.take_free()}};
        s.psi.add_component(std::make_unique<alea::NosanowJastrow>(
            s.dim, grp.first, std::move(st), s.cell, std::move(params)));
    });
}

int ref_add_ewald(void* h, double tol, double eps) {
    auto& s = *static_cast<RefSystem*>(h);
    s.use_ewald = true;
    s.ewald_tol = tol;
    s.ewald_eps = eps;
    return 0;
}

int ref_add_coulomb(void* h, int ga, int gb, double eps) {
    auto& s = *static_cast<RefSystem*>(h);
    s.coulomb.push_back({ga, gb, eps});
    return 0;
}

// Positions in the reference SoA layout [p*dim*nw + d*nw + w]; they are wrapped
// into the cell and the distance tables rebuilt (WalkerSet::load_all,
// src/population/walker_set.cpp:51-70), then ψ is evaluated.
int ref_add_inverse_power(void* h, int ga, int gb, double c, double p) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        s.ipow.push_back({ga, gb, c, p});
    });
}

int ref_set_positions(void* h, const double* pos) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        auto p = s.ws->positions();
        std::copy(pos, pos + p.size(), p.begin());
        if (s.periodic) {
            double r[3];
            for (std::size_t w = 0; w < s.nw; ++w)
                for (std::size_t i = 0; i < s.np; ++i) {
                    for (std::size_t d = 0; d < s.dim; ++d) r[d] = p[(i * s.dim + d) * s.nw + w];
                    s.cell.wrap(r);
                    for (std::size_t d = 0; d < s.dim; ++d) p[(i * s.dim + d) * s.nw + w] = r[d];
                }
        }
        s.ws->dist_tables().build_all(p);
    });
}

int ref_evaluate(void* h) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        build_hamiltonian(s);
        s.psi.evaluate_batch(*s.ws);
    });
}

int ref_get_positions(void* h, double* out) {
    auto& s = *static_cast<RefSystem*>(h);
    auto p = s.ws->positions();
    std::copy(p.begin(), p.end(), out);
    return 0;
}

int ref_get_log_psi(void* h, double* out) {
    auto& s = *static_cast<RefSystem*>(h);
    auto p = s.ws->log_psi();
    std::copy(p.begin(), p.end(), out);
    return 0;
}

int ref_get_gradients(void* h, double* grad, double* lap) {
    auto& s = *static_cast<RefSystem*>(h);
    s.psi.refresh_derivatives(*s.ws);
    auto g = s.ws->gradients();
    auto l = s.ws->laplacians();
    std::copy(g.begin(), g.end(), grad);
    std::copy(l.begin(), l.end(), lap);
    return 0;
}

// refresh_derivatives + Hamiltonian::evaluate_batch (src/qmc/vmc.cpp:171-173).
int ref_local_energy(void* h, double* e_re) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        build_hamiltonian(s);
        s.psi.refresh_derivatives(*s.ws);
        s.ham.evaluate_batch(*s.ws);
        auto le = s.ws->local_energies();
        std::copy(le.begin(), le.end(), e_re);
    });
}

// Free flags for the parameters of the factors added next, in the order they
// are added (Padé: its coefficients; Drummond: its coefficients, then L;
// Nosanow: alpha).
int ref_set_free(void* h, int n, const unsigned char* flags) {
    auto& s = *static_cast<RefSystem*>(h);
    s.free_next.assign(flags, flags + n);
    s.free_used = 0;
    return 0;
}

// TrialWavefunction::compute_parameter_derivatives (trial_wavefunction.cpp:
// 150-160) after evaluate: derivs [n_opt][nw]; *n_opt the count.
int ref_parameter_derivatives(void* h, double* derivs, int cap, int* n_opt) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        const std::size_t n = s.psi.total_optimizable_params();
        *n_opt = static_cast<int>(n);
        if (static_cast<int>(n * s.nw) > cap) throw std::runtime_error("derivs too small");
        s.psi.compute_parameter_derivatives(*s.ws, std::span<double>(derivs, n * s.nw));
    });
}

// TrialWavefunction::compute_parameter_gradient_derivatives
// (trial_wavefunction.cpp:162-185) after refresh_derivatives, reduced by
// Hamiltonian::local_energy_parameter_derivatives (hamiltonian.cpp:189-196):
// de [n_opt][nw]; *n_opt the count.
int ref_local_energy_parameter_derivatives(void* h, double* de, int cap, int* n_opt) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        build_hamiltonian(s);
        s.psi.refresh_derivatives(*s.ws);
        const std::size_t n = s.psi.total_optimizable_params();
        *n_opt = static_cast<int>(n);
        if (static_cast<int>(n * s.nw) > cap) throw std::runtime_error("de too small");
        const std::size_t block = s.np * s.dim * s.nw;
        std::vector<double> dgrad(n * block), dlap(n * block);
        std::vector<char> supported(n);
        s.psi.compute_parameter_gradient_derivatives(*s.ws, dgrad, dlap, supported);
        for (std::size_t k = 0; k < n; ++k)
            if (!supported[k]) throw std::runtime_error("gradient derivatives not supported");
        std::fill(de, de + n * s.nw, 0.0);
        s.ham.local_energy_parameter_derivatives(*s.ws, dgrad, dlap, n, std::span<double>(de, n * s.nw));
    });
}

// Hamiltonian::delta_potential_particle (hamiltonian.cpp:105-117) for every
// walker: new_pos [dim][nw], out [nw].
int ref_delta_potential(void* h, int p, const double* new_pos, double* out) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        build_hamiltonian(s);
        std::vector<double> scratch(s.nw);
        double proposed[3];
        for (std::size_t k = 0; k < s.nw; ++k) {
            for (std::size_t d = 0; d < s.dim; ++d) proposed[d] = new_pos[d * s.nw + k];
            out[k] = s.ham.delta_potential_particle(*s.ws, static_cast<std::size_t>(p), k,
                                                    std::span<const double>(proposed, s.dim), scratch);
        }
    });
}

// Per-call path: propose particle p at new_pos ([dim*nw]) and return
// TrialWavefunction::ratio_batch's outputs without committing.
int ref_ratio(void* h, int p, const double* new_pos, double* log_ratio, double* new_grad,
              double* new_lap) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        auto np = s.ws->new_pos();
        std::copy(new_pos, new_pos + np.size(), np.begin());
        s.ws->propose_move_batch(static_cast<std::size_t>(p), np);
        s.psi.ratio_batch(*s.ws, static_cast<std::size_t>(p));
        auto lr = s.ws->log_ratios();
        std::copy(lr.begin(), lr.end(), log_ratio);
        auto g = s.ws->new_grad();
        auto l = s.ws->new_lap();
        std::copy(g.begin(), g.end(), new_grad);
        std::copy(l.begin(), l.end(), new_lap);
    });
}

// Accept the move proposed by the last ref_ratio where mask[w] != 0
// (apply_move_batch, then WalkerSet::apply_move; src/qmc/vmc.cpp:147-152).
int ref_accept(void* h, int p, const unsigned char* mask) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        auto m = s.ws->accept_mask();
        for (std::size_t w = 0; w < s.nw; ++w) m[w] = mask[w] != 0;
        s.psi.apply_move_batch(*s.ws, static_cast<std::size_t>(p), m);
        s.ws->apply_move(static_cast<std::size_t>(p), s.ws->new_pos(), s.ws->new_grad(),
                         s.ws->new_lap(), m);
    });
}

// VMC Metropolis sweeps from a given tape (one draw_sweep_randoms block per
// sweep, layout [np][dim+1][nw]), in the order of VMC::sweep_walkers
// (src/qmc/vmc.cpp:82-155) with symmetric proposals, sigma = sqrt(tau * inv_mass).
int ref_vmc_sweeps(void* h, int n_sweeps, double tau, const double* tape,
                   long long* n_accepted) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        const std::size_t nw = s.nw, np = s.np, dim = s.dim;
        std::vector<double> sigma(np * dim);
        for (const auto& g : s.groups)
            for (std::size_t i = g.first; i < g.last; ++i)
                for (std::size_t d = 0; d < dim; ++d)
                    sigma[i * dim + d] = std::sqrt(tau * g.inv_mass[d]);
        long long acc = 0;
        for (int sw = 0; sw < n_sweeps; ++sw) {
            const double* randoms = tape + static_cast<std::size_t>(sw) * np * (dim + 1) * nw;
            auto new_pos = s.ws->new_pos();
            auto mask = s.ws->accept_mask();
            for (std::size_t p = 0; p < np; ++p) {
                const double* gaussian = randoms + p * (dim + 1) * nw;
                const double* uniform = gaussian + dim * nw;
                const auto pos = s.ws->positions();
                for (std::size_t d = 0; d < dim; ++d)
                    for (std::size_t w = 0; w < nw; ++w)
                        new_pos[d * nw + w] = pos[p * dim * nw + d * nw + w] +
                                              sigma[p * dim + d] * gaussian[d * nw + w];
                s.ws->propose_move_batch(p, new_pos);
                s.psi.ratio_batch(*s.ws, p);
                const auto lr = s.ws->log_ratios();
                for (std::size_t w = 0; w < nw; ++w) {
                    mask[w] = uniform[w] < std::exp(2.0 * lr[w]);
                    if (mask[w]) ++acc;
                }
                s.psi.apply_move_batch_deferred(*s.ws, p, mask);
                s.ws->apply_move(p, new_pos, s.ws->new_grad(), s.ws->new_lap(), mask);
            }
        }
        *n_accepted = acc;
    });
}

// VMC::sweep_walkers with use_drift (vmc.cpp:98-139) restated around the
// reference's own wavefunction and walker set: refresh_particle_derivatives,
// the drift proposal, and numerics::langevin_log_ratio_d for the acceptance.
int ref_vmc_drift_sweeps(void* h, int n_sweeps, double tau, const double* tape,
                         long long* n_accepted) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        const std::size_t nw = s.nw, np = s.np, dim = s.dim;
        std::vector<double> tm(np * dim);
        for (const auto& g : s.groups)
            for (std::size_t i = g.first; i < g.last; ++i)
                for (std::size_t d = 0; d < dim; ++d) tm[i * dim + d] = tau * g.inv_mass[d];
        long long acc = 0;
        for (int sw = 0; sw < n_sweeps; ++sw) {
            const double* randoms = tape + static_cast<std::size_t>(sw) * np * (dim + 1) * nw;
            auto new_pos = s.ws->new_pos();
            auto mask = s.ws->accept_mask();
            for (std::size_t p = 0; p < np; ++p) {
                s.psi.refresh_particle_derivatives(*s.ws, p);
                const double* gaussian = randoms + p * (dim + 1) * nw;
                const double* uniform = gaussian + dim * nw;
                const auto pos = s.ws->positions();
                const auto grad = s.ws->gradients();
                for (std::size_t d = 0; d < dim; ++d)
                    for (std::size_t w = 0; w < nw; ++w) {
                        const std::size_t off = p * dim * nw + d * nw + w;
                        new_pos[d * nw + w] = pos[off] + tm[p * dim + d] * grad[off] +
                                              std::sqrt(tm[p * dim + d]) * gaussian[d * nw + w];
                    }
                s.ws->propose_move_batch(p, new_pos);
                s.psi.ratio_batch(*s.ws, p);
                const auto lr = s.ws->log_ratios();
                const auto ng = s.ws->new_grad();
                for (std::size_t w = 0; w < nw; ++w) {
                    double lc = 0.0;
                    for (std::size_t d = 0; d < dim; ++d) {
                        const std::size_t off = p * dim * nw + d * nw + w;
                        lc += alea::numerics::langevin_log_ratio_d(new_pos[d * nw + w] - pos[off], grad[off],
                                                                   ng[d * nw + w], tm[p * dim + d]);
                    }
                    const double la = 2.0 * lr[w] + lc;
                    mask[w] = la >= 0.0 || uniform[w] < std::exp(la);
                    if (mask[w]) ++acc;
                }
                s.psi.apply_move_batch_deferred(*s.ws, p, mask);
                s.ws->apply_move(p, new_pos, s.ws->new_grad(), s.ws->new_lap(), mask);
            }
        }
        *n_accepted = acc;
    });
}

// WalkerSet::branch (walker_set.cpp:187-239) on a one-particle walker set
// whose walkers carry the given weights; rng = XoshiroRandom(seed).  Writes
// the source map (the walker tagged w has x = w) and the next uniform the
// stream would give after the branch.
int ref_branch(const double* weights, int nw, uint64_t seed, int target, int* src_out,
               int* new_nw, double* next_uniform) {
    return guarded([&] {
        alea::WalkerSet ws(static_cast<std::size_t>(nw), 1, 2);
        std::vector<alea::Walker> walkers;
        walkers.reserve(static_cast<std::size_t>(nw));
        for (int w = 0; w < nw; ++w) walkers.emplace_back(1, 2);
        for (auto& w : walkers) ws.add_walker(w);
        ws.load_all();
        auto pos = ws.positions();
        for (int w = 0; w < nw; ++w) pos[w] = static_cast<double>(w);   // x of particle 0
        for (int w = 0; w < nw; ++w) ws.weights()[static_cast<std::size_t>(w)] = weights[w];
        alea::numerics::XoshiroRandom rng(seed);
        (void)ws.branch(rng, static_cast<std::size_t>(target));
        const std::size_t n = ws.num_walkers();
        auto p = ws.positions();
        for (std::size_t j = 0; j < n; ++j) src_out[j] = static_cast<int>(p[j]);
        *new_nw = static_cast<int>(n);
        *next_uniform = rng.uniform();
    });
}

// PopulationControl (population_control.cpp) driven through a scripted run:
// start(e0), set_energy_spread(sigma); then per step i the branch weights of
// n_pairs (e_new, e_old) pairs (and is_cut) and update(ens_e[i], wsum[i],
// accumulate[i]).  Writes the weights [n_steps][n_pairs], the cut flags, E_T
// after every update (which fixes E_ref) and e_cut.
int ref_popctl_run(double tau, int n_substeps, double generations, int target, int cutoff,
                   double alpha, int n_particles, double e0, double sigma, int n_steps,
                   const double* ens_e, const double* wsum, const int* accumulate, int n_pairs,
                   const double* e_new, const double* e_old, double* weights_out, int* cut_out,
                   double* etrial_out, double* ecut_out) {
    return guarded([&] {
        const alea::BranchCutoff kinds[3] = {alea::BranchCutoff::none, alea::BranchCutoff::drv,
                                             alea::BranchCutoff::universal};
        alea::PopulationControl pc(tau, n_substeps, generations, static_cast<std::size_t>(target),
                                   kinds[cutoff], alpha, static_cast<std::size_t>(n_particles));
        pc.start(e0);
        pc.set_energy_spread(sigma);
        *ecut_out = pc.energy_cutoff();
        for (int i = 0; i < n_steps; ++i) {
            for (int k = 0; k < n_pairs; ++k) {
                weights_out[i * n_pairs + k] = pc.branch_weight(e_new[k], e_old[k]);
                cut_out[i * n_pairs + k] = pc.is_cut(e_new[k], e_old[k]) ? 1 : 0;
            }
            pc.update(ens_e[i], wsum[i], accumulate[i] != 0);
            etrial_out[i] = pc.etrial();
        }
    });
}

// numerics::robust_spread (statistics.cpp:309-315).
double ref_robust_spread(const double* x, int n) {
    std::vector<double> scratch;
    return alea::numerics::robust_spread(std::span<const double>(x, static_cast<std::size_t>(n)), scratch);
}

// numerics::reblock / reblock_chains (n_chains = 0: reblock).  out[7] =
// mean, naive_error, error, tau_int, error_on_error, block_length, level;
// iout[3] = plateau_found, selected_level, selected_n_blocks.
int ref_reblock(const double* x, int n, int n_chains, double* out, int* iout) {
    return guarded([&] {
        const std::span<const double> s(x, static_cast<std::size_t>(n));
        const alea::numerics::ReblockResult r =
            n_chains > 0 ? alea::numerics::reblock_chains(s, static_cast<std::size_t>(n_chains))
                         : alea::numerics::reblock(s);
        out[0] = r.mean;
        out[1] = r.naive_error;
        out[2] = r.error;
        out[3] = r.tau_int;
        out[4] = r.error_on_error;
        out[5] = r.block_length;
        out[6] = r.level;
        iout[0] = r.plateau_found ? 1 : 0;
        iout[1] = r.selected_level;
        iout[2] = static_cast<int>(r.selected_n_blocks);
    });
}

// PairCorrelation (rmax > 0) and StructureFactor (kmax > 0) over the
// system's walkers: n_blocks blocks of `steps` measurements, measurement m
// at positions pos + m * np*dim*nw (SoA, wrapped and tabulated as
// ref_set_positions does) with weights w + m * nw; then finalize().  The
// RunResult series come back flattened: names joined by '\n' into `names`,
// values concatenated into `values`, lengths[n_series].
namespace {
int run_estimators(void* h, std::vector<std::unique_ptr<alea::Estimator>>& ests, int n_blocks, int steps,
                   const double* pos, const double* w, char* names, int names_cap, double* values,
                   int values_cap, int* lengths, int lengths_cap, int* n_series);
}

int ref_estimators(void* h, double rmax, int n_bins, double kmax, int n_blocks, int steps,
                   const double* pos, const double* w, char* names, int names_cap, double* values,
                   int values_cap, int* lengths, int lengths_cap, int* n_series) {
    std::vector<std::unique_ptr<alea::Estimator>> ests;
    const int rc = guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        if (rmax > 0.0)
            ests.push_back(std::make_unique<alea::PairCorrelation>(s.groups, s.dim, rmax, n_bins, s.cell));
        if (kmax > 0.0)
            ests.push_back(std::make_unique<alea::StructureFactor>(s.groups, s.cell, kmax));
    });
    if (rc) return rc;
    return run_estimators(h, ests, n_blocks, steps, pos, w, names, names_cap, values, values_cap, lengths,
                          lengths_cap, n_series);
}

// Density (density.h:55, per group, lab-frame grid or the periodic cell) over
// the same measurement sequence as ref_estimators.
int ref_density(void* h, double rmax, int n_bins, int per_particle, int n_blocks, int steps, const double* pos,
                const double* w,
                char* names, int names_cap, double* values, int values_cap, int* lengths, int lengths_cap,
                int* n_series) {
    std::vector<std::unique_ptr<alea::Estimator>> ests;
    const int rc = guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        ests.push_back(std::make_unique<alea::Density>(s.groups, s.dim, rmax, n_bins, per_particle != 0, s.cell));
    });
    if (rc) return rc;
    return run_estimators(h, ests, n_blocks, steps, pos, w, names, names_cap, values, values_cap, lengths,
                          lengths_cap, n_series);
}

// ForwardWalking (forward_walking.cpp) over a scripted population history on
// the harness's walker set.  Step t (numbered from 0) sets the population's
// nw_t walkers to positions pos_t [np*dim*nw_t] and weights w_t [nw_t], runs
// ForwardWalking::observe, then passes the step's source map src_t
// [nw_{t+1}] to ForwardWalking::branch.  Blocks of steps_per_block steps start at
// first_measured.  Every lag measures the potential energy, and a pair
// correlation when gofr_bins > 0.  Output: write_results' series, and its
// scalars as one-value series, in the layout of run_estimators.
int ref_forward_walking(void* h, int n_lags, const int* lags, double step_time, int interval, int first_measured,
                        int n_steps, int steps_per_block, double gofr_rmax, int gofr_bins, const int* nw_of_step,
                        const double* pos, const double* w, const int* src, char* names, int names_cap,
                        double* values, int values_cap, int* lengths, int lengths_cap, int* n_series) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        build_hamiltonian(s);
        std::vector<std::vector<std::unique_ptr<alea::Estimator>>> per_lag(static_cast<std::size_t>(n_lags));
        if (gofr_bins > 0)
            for (auto& set : per_lag) {
                set.push_back(std::make_unique<alea::PairCorrelation>(s.groups, s.dim, gofr_rmax, gofr_bins, s.cell));
                set.back()->set_output_label("fw" + std::to_string(&set - per_lag.data()));
            }
        std::vector<std::size_t> lag_steps(lags, lags + n_lags);
        alea::ForwardWalking fw(lag_steps, step_time, static_cast<std::size_t>(interval), std::move(per_lag));
        std::size_t capacity = 0;
        for (int t = 0; t < n_steps; ++t) capacity = std::max(capacity, static_cast<std::size_t>(nw_of_step[t]));
        fw.prepare(capacity, s.np, s.dim, s.cell, static_cast<std::size_t>(first_measured));
        std::size_t pos_off = 0, w_off = 0, src_off = 0;
        for (int t = 0; t < n_steps; ++t) {
            // every step sets every position and weight, so the population is
            // a fresh walker set of the step's size (the harness's sets carry
            // no Walker handles for apply_permutation to move)
            const std::size_t nw = static_cast<std::size_t>(nw_of_step[t]);
            s.ws = std::make_unique<alea::WalkerSet>(nw, s.np, s.dim, s.cell);
            s.nw = nw;
            if (ref_set_positions(h, pos + pos_off)) throw std::runtime_error(g_error);
            pos_off += s.np * s.dim * nw;
            auto wt = s.ws->weights();
            std::copy(w + w_off, w + w_off + nw, wt.begin());
            w_off += nw;
            const bool measured = t >= first_measured;
            const bool opens = measured && (t - first_measured) % steps_per_block == 0;
            const bool closes = measured && (t - first_measured + 1) % steps_per_block == 0;
            if (opens) fw.start_block();
            fw.observe(*s.ws, static_cast<std::size_t>(t), s.ham);
            if (closes) fw.stop_block();
            if (t + 1 < n_steps) {
                const std::size_t nn = static_cast<std::size_t>(nw_of_step[t + 1]);
                std::vector<std::size_t> map(src + src_off, src + src_off + nn);
                src_off += nn;
                fw.branch(map);
            }
        }
        alea::RunResult rr;
        fw.write_results(rr);
        for (const auto& [name, value] : rr.scalars) rr.series.emplace_back(name, std::vector<double>{value});
        std::string joined;
        std::size_t nv = 0;
        int ns = 0;
        for (const auto& [name, vals] : rr.series) {
            if (ns >= lengths_cap || nv + vals.size() > static_cast<std::size_t>(values_cap))
                throw std::runtime_error("ref_forward_walking: output capacity");
            joined += name + "\n";
            std::copy(vals.begin(), vals.end(), values + nv);
            nv += vals.size();
            lengths[ns++] = static_cast<int>(vals.size());
        }
        if (joined.size() + 1 > static_cast<std::size_t>(names_cap))
            throw std::runtime_error("ref_forward_walking: name capacity");
        std::copy(joined.begin(), joined.end(), names);
        names[joined.size()] = '\0';
        *n_series = ns;
    });
}

}  // extern "C"

namespace {
int run_estimators(void* h, std::vector<std::unique_ptr<alea::Estimator>>& ests, int n_blocks, int steps,
                   const double* pos, const double* w, char* names, int names_cap, double* values,
                   int values_cap, int* lengths, int lengths_cap, int* n_series) {
    return guarded([&] {
        auto& s = *static_cast<RefSystem*>(h);
        const std::size_t len = s.np * s.dim * s.nw;
        for (int b = 0; b < n_blocks; ++b) {
            for (auto& e : ests) e->start_block();
            for (int m = 0; m < steps; ++m) {
                const std::size_t idx = static_cast<std::size_t>(b) * steps + m;
                if (ref_set_positions(h, pos + idx * len)) throw std::runtime_error(g_error);
                const std::span<const double> wt(w + idx * s.nw, s.nw);
                for (auto& e : ests) e->accumulate(*s.ws, wt);
            }
            for (auto& e : ests) e->stop_block();
        }
        alea::RunResult rr;
        for (auto& e : ests) e->finalize(rr);
        std::string joined;
        std::size_t nv = 0;
        int ns = 0;
        for (const auto& [name, vals] : rr.series) {
            if (ns >= lengths_cap || nv + vals.size() > static_cast<std::size_t>(values_cap))
                throw std::runtime_error("ref_estimators: output capacity");
            joined += name + "\n";
            std::copy(vals.begin(), vals.end(), values + nv);
            nv += vals.size();
            lengths[ns++] = static_cast<int>(vals.size());
        }
        if (joined.size() + 1 > static_cast<std::size_t>(names_cap))
            throw std::runtime_error("ref_estimators: name capacity");
        std::memcpy(names, joined.c_str(), joined.size() + 1);
        *n_series = ns;
    });
}
}  // namespace
