// The device factors as components of the reference's own TrialWavefunction
// and Hamiltonian (include/alea_amd_component.hpp), with the oracle library
// behind the C-ABI (oracle_names.h), checked on the CPU against:
//   * the reference's native PadeJastrow, DrummondJastrow, NosanowJastrow and
//     EwaldTerm (compiled from its sources into oracle/_ref/libalea_ref.so):
//     the same trial and Hamiltonian built once from native components and
//     once with some of them replaced by one DeviceComponent / DevicePotential;
//   * the C-ABI system itself, for determinant blocks (the reference's
//     SlaterDeterminant needs Eigen and is not built): native Padé plus a
//     DeviceComponent of Slater blocks against one system holding all three.
// Each scenario drives the component contract the way VMC::sweep_walkers
// and the DMC crowds do (vmc.cpp:82-155, dmc.cpp:300-340): evaluate_batch,
// per-move ratio_batch / apply_move_batch over two crowd ranges, the deferred
// commit with refresh_derivatives (at the sweep's end, and per particle
// before its move), branching (apply_permutation +
// permute_caches, with copies and a shrink), clone, parameter derivatives and
// rebuild after a parameter change, local energies and
// delta_potential_particle.  Exit status 0 when every comparison holds to
// 1e-10 relative.  Test-only; never part of the product build.
#include <cmath>
#include <cstdio>
#include <deque>
#include <memory>
#include <random>
#include <span>
#include <string>
#include <vector>

#include "alea_amd_component.hpp"
#include "hamiltonian/ewald_term.h"
#include "hamiltonian/hamiltonian.h"
#include "hamiltonian/kinetic.h"
#include "population/simulation_cell.h"
#include "population/walker.h"
#include "population/walker_set.h"
#include "wavefunction/drummond_jastrow.h"
#include "wavefunction/nosanow_jastrow.h"
#include "wavefunction/pade_jastrow.h"
#include "wavefunction/trial_wavefunction.h"

extern "C" int ref_ewald_setup(const double* lattice, double tol, double eps, int cap_k, int* nk, int* n1, int* n2,
                               double* kx, double* ky, double* weight, int cap_knots, int* n_knots, double* scaled,
                               double* scalars);

namespace {

constexpr double kTol = 1e-10;
int g_failures = 0;
int g_checks = 0;

// max |a - b| over max(max |b|, floor): relative, with an absolute floor for
// a quantity that is zero up to roundoff
double rel(std::span<const double> a, std::span<const double> b, double floor = 1e-300) {
    double num = 0.0, den = floor;
    const std::size_t n = std::min(a.size(), b.size());
    for (std::size_t i = 0; i < n; ++i) {
        num = std::max(num, std::abs(a[i] - b[i]));
        den = std::max(den, std::abs(b[i]));
    }
    if (a.size() != b.size()) return INFINITY;
    return num / den;
}

void expect(const std::string& what, std::span<const double> got, std::span<const double> want,
            double tol = kTol, double floor = 1e-300) {
    ++g_checks;
    const double e = rel(got, want, floor);
    if (!(e <= tol)) {
        ++g_failures;
        std::printf("FAIL %s: rel err %.3e (n=%zu/%zu)\n", what.c_str(), e, got.size(), want.size());
    }
}

void expect_true(const std::string& what, bool ok) {
    ++g_checks;
    if (!ok) {
        ++g_failures;
        std::printf("FAIL %s\n", what.c_str());
    }
}

std::vector<double> copy_of(std::span<const double> s) { return {s.begin(), s.end()}; }

alea::ParticleGroup group(std::size_t first, std::size_t last) {
    alea::ParticleGroup g;
    g.name = "g" + std::to_string(first);
    g.mass.assign(2, 1.0);
    g.inv_mass.assign(2, 1.0);
    g.charge = -1.0;
    g.first = first;
    g.last = last;
    g.count = last - first;
    return g;
}

struct Cell {
    std::vector<double> lattice;
    alea::SimulationCell cell;
    std::vector<alea::ParticleGroup> groups;
    std::size_t np = 0;
};

// N electrons (two equal spin groups) at density rs in a square cell.
Cell make_cell(std::size_t np, double rs) {
    Cell c;
    const double L = std::sqrt(static_cast<double>(np) * M_PI * rs * rs);
    c.lattice = {L, 0.0, 0.0, L};
    c.cell = alea::SimulationCell(2, c.lattice);
    c.np = np;
    c.groups = {group(0, np / 2), group(np / 2, np)};
    return c;
}

void load(alea::WalkerSet& ws, const Cell& c, const std::vector<double>& pos) {
    auto p = ws.positions();
    std::copy(pos.begin(), pos.end(), p.begin());
    const std::size_t nw = ws.num_walkers();
    double r[3];
    for (std::size_t w = 0; w < nw; ++w)
        for (std::size_t i = 0; i < c.np; ++i) {
            for (std::size_t d = 0; d < 2; ++d) r[d] = p[(i * 2 + d) * nw + w];
            c.cell.wrap(r);
            for (std::size_t d = 0; d < 2; ++d) p[(i * 2 + d) * nw + w] = r[d];
        }
    ws.dist_tables().build_all(p);
}

std::vector<double> uniform_positions(const Cell& c, std::size_t nw, std::mt19937_64& rng) {
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<double> pos(c.np * 2 * nw);
    for (std::size_t i = 0; i < c.np; ++i)
        for (std::size_t w = 0; w < nw; ++w) {
            const double a = u(rng), b = u(rng);
            pos[(i * 2 + 0) * nw + w] = a * c.lattice[0] + b * c.lattice[2];
            pos[(i * 2 + 1) * nw + w] = a * c.lattice[1] + b * c.lattice[3];
        }
    return pos;
}

// ── Two trial wavefunctions driven by one random stream ──────────────────
// The walker set holds Walker handles, as every driver's does (branching
// carries them along).
struct Side {
    std::deque<alea::Walker> walkers;
    std::unique_ptr<alea::WalkerSet> ws;
    alea::TrialWavefunction psi;
    void make(std::size_t nw, const Cell& c) {
        ws = std::make_unique<alea::WalkerSet>(nw, c.np, 2, c.cell);
        walkers.clear();
        for (std::size_t w = 0; w < nw; ++w) ws->add_walker(walkers.emplace_back(c.np, 2));
    }
};

struct Pair {
    const Cell* c = nullptr;
    Side a, b;   // a: the reference's components (or the plain C-ABI system), b: with the adapter
};

void compare_totals(const std::string& tag, Side& a, Side& b) {
    expect(tag + " log psi", b.ws->log_psi(), a.ws->log_psi());
    expect(tag + " gradients", b.ws->gradients(), a.ws->gradients());
    expect(tag + " laplacians", b.ws->laplacians(), a.ws->laplacians());
    if (a.ws->has_phase() || b.ws->has_phase()) {
        expect_true(tag + " both carry phase planes", a.ws->has_phase() && b.ws->has_phase());
        expect(tag + " phase gradients", b.ws->phase_gradients(), a.ws->phase_gradients());
        // the twisted plane-wave blocks here have grad^2 phi = 0 identically
        // (their +-k column pairs make the determinant real up to
        // exp(i theta . sum r)): roundoff against roundoff, held absolutely
        expect(tag + " phase laplacians", b.ws->phase_laplacians(), a.ws->phase_laplacians(), kTol, 1.0);
    }
}

// One sweep of single-particle moves, both sides from the same proposals and
// side a's decisions; crowds: the walker ranges the calls are split into.
// deferred: commit with apply_move_batch_deferred and refresh at the end;
// refresh_each: also refresh each particle before its move (as the drift
// proposal of VMC::sweep_walkers does, vmc.cpp:98-139) and compare its rows.
void sweep(const std::string& tag, Pair& P, std::mt19937_64& rng, double sigma, bool deferred,
           const std::vector<std::pair<std::size_t, std::size_t>>& crowds, bool refresh_each = false) {
    const std::size_t nw = P.a.ws->num_walkers(), np = P.c->np;
    std::normal_distribution<double> gauss(0.0, 1.0);
    std::uniform_real_distribution<double> unif(0.0, 1.0);
    std::vector<double> newp(2 * nw), u(nw);
    for (std::size_t p = 0; p < np; ++p) {
        const auto pos = P.a.ws->positions();
        for (std::size_t d = 0; d < 2; ++d)
            for (std::size_t w = 0; w < nw; ++w)
                newp[d * nw + w] = pos[(p * 2 + d) * nw + w] + sigma * gauss(rng);
        for (auto& x : u) x = unif(rng);
        if (deferred && refresh_each) {
            for (Side* s : {&P.a, &P.b})
                for (const auto& [w0, w1] : crowds) s->psi.refresh_particle_derivatives(*s->ws, p, w0, w1);
            const std::size_t row = p * 2 * nw;
            const std::string rp = tag + " p=" + std::to_string(p) + " refreshed row";
            expect(rp + " gradient", P.b.ws->gradients().subspan(row, 2 * nw), P.a.ws->gradients().subspan(row, 2 * nw));
            expect(rp + " laplacian", P.b.ws->laplacians().subspan(row, 2 * nw),
                   P.a.ws->laplacians().subspan(row, 2 * nw));
        }
        for (Side* s : {&P.a, &P.b}) {
            auto np_ = s->ws->new_pos();
            std::copy(newp.begin(), newp.end(), np_.begin());
            s->ws->propose_move_batch(p, np_);
            for (const auto& [w0, w1] : crowds) s->psi.ratio_batch(*s->ws, p, w0, w1);
        }
        const std::string mv = tag + " p=" + std::to_string(p);
        expect(mv + " log ratios", P.b.ws->log_ratios(), P.a.ws->log_ratios());
        expect(mv + " new grad", P.b.ws->new_grad(), P.a.ws->new_grad());
        expect(mv + " new lap", P.b.ws->new_lap(), P.a.ws->new_lap());
        expect(mv + " phase diffs", P.b.ws->phase_diffs(), P.a.ws->phase_diffs());
        if (P.a.ws->has_phase()) {
            expect(mv + " new phase grad", P.b.ws->new_phase_grad(), P.a.ws->new_phase_grad());
            expect(mv + " new phase lap", P.b.ws->new_phase_lap(), P.a.ws->new_phase_lap(), kTol, 1.0);
        }
        const auto lr = P.a.ws->log_ratios();
        std::vector<char> acc(nw);
        for (std::size_t w = 0; w < nw; ++w) acc[w] = u[w] < std::exp(2.0 * lr[w]);
        for (Side* s : {&P.a, &P.b}) {
            auto m = s->ws->accept_mask();
            for (std::size_t w = 0; w < nw; ++w) m[w] = acc[w] != 0;
            for (const auto& [w0, w1] : crowds) {
                if (deferred) {
                    s->psi.apply_move_batch_deferred(*s->ws, p, m, w0, w1);
                } else {
                    s->psi.apply_move_batch(*s->ws, p, m, w0, w1);
                }
                s->ws->apply_move(p, s->ws->new_pos(), s->ws->new_grad(), s->ws->new_lap(), m, w0, w1);
            }
        }
        if (!deferred) compare_totals(mv, P.a, P.b);
    }
    if (deferred)
        for (Side* s : {&P.a, &P.b}) s->psi.refresh_derivatives(*s->ws);
    compare_totals(tag + " after the sweep", P.a, P.b);
    expect(tag + " positions", P.b.ws->positions(), P.a.ws->positions());
}

void branch(const std::string& tag, Pair& P, std::mt19937_64& rng, std::size_t new_nw) {
    const std::size_t nw = P.a.ws->num_walkers();
    std::uniform_int_distribution<std::size_t> pick(0, nw - 1);
    std::vector<std::size_t> src(new_nw);
    for (auto& x : src) x = pick(rng);
    src[0] = src[1];   // one walker twice at least
    for (Side* s : {&P.a, &P.b}) {
        s->ws->apply_permutation(src);
        s->psi.permute_caches(src);
    }
    compare_totals(tag, P.a, P.b);
}

// The carried totals against a fresh evaluation of the same configurations.
void reevaluate(const std::string& tag, Pair& P) {
    for (Side* s : {&P.a, &P.b}) {
        const auto lp = copy_of(s->ws->log_psi()), g = copy_of(s->ws->gradients()), l = copy_of(s->ws->laplacians());
        const bool ph = s->ws->has_phase();
        const auto pg = ph ? copy_of(s->ws->phase_gradients()) : std::vector<double>{};
        const auto pl = ph ? copy_of(s->ws->phase_laplacians()) : std::vector<double>{};
        s->psi.evaluate_batch(*s->ws);
        const std::string side = tag + (s == &P.a ? " side a" : " side b");
        expect(side + " carried log psi", s->ws->log_psi(), lp, 1e-9);
        expect(side + " carried gradients", s->ws->gradients(), g, 1e-9);
        expect(side + " carried laplacians", s->ws->laplacians(), l, 1e-9);
        if (ph) {
            expect(side + " carried phase gradients", s->ws->phase_gradients(), pg, 1e-9);
            expect(side + " carried phase laplacians", s->ws->phase_laplacians(), pl, 1e-9, 1.0);
        }
    }
    compare_totals(tag + " re-evaluated", P.a, P.b);
}

std::vector<double> param_derivs(Side& s) {
    std::vector<double> d(s.psi.total_optimizable_params() * s.ws->num_walkers());
    s.psi.compute_parameter_derivatives(*s.ws, d);
    return d;
}

// ── Scenario 1: Jastrow factors, native against adapted ──────────────────
struct JastrowSpec {
    double pade[2] = {0.2, 1.0};          // a_1, b_1 of the unlike-spin Padé
    double drummond[3] = {2e-3, -4e-4, 5e-5};
    double drummond_L = 0.0;
    double nosanow_alpha = 0.6;
    std::vector<double> sites;            // one per down-spin particle
};

std::unique_ptr<alea::PadeJastrow> native_pade(const Cell& c, int ga, int gb, double a1, double b1, bool free) {
    std::vector<alea::Parameter> prm{{"a1", a1, !free}, {"b1", b1, !free}};
    return std::make_unique<alea::PadeJastrow>(1, 1, c.groups[ga], c.groups[gb], std::move(prm));
}

std::vector<alea::Parameter> drummond_params(const JastrowSpec& j) {
    return {{"a0", j.drummond[0], false}, {"a2", j.drummond[1], false}, {"a3", j.drummond[2], true},
            {"L", j.drummond_L, false}};
}

std::unique_ptr<alea_amd::DeviceComponent> device_jastrows(const Cell& c, const JastrowSpec& j,
                                          std::vector<alea::Parameter> drm, std::size_t capacity) {
    alea_amd::DeviceSystemSpec spec;
    spec.np = c.np;
    spec.lattice = c.lattice;
    spec.capacity = capacity;
    const std::vector<double> sites = j.sites;
    const std::size_t n_up = c.np / 2;
    // the device factors read their current values from the parameters:
    // Drummond's alpha_0, alpha_2, alpha_3 and L, then Nosanow's alpha
    spec.registrar = [n_up, np = c.np, sites](alea_system* s, std::span<const alea::Parameter> prm) {
        alea_amd::component_check(alea_add_group(s, 0, static_cast<int>(n_up), -1.0, 1.0));
        alea_amd::component_check(alea_add_group(s, static_cast<int>(n_up), static_cast<int>(np), -1.0, 1.0));
        const double coeffs[3] = {prm[0].value, prm[1].value, prm[2].value};
        alea_amd::component_check(alea_add_drummond_jastrow(s, 0, 1, 3, 1.0, prm[3].value, 3, coeffs));
        alea_amd::component_check(alea_add_nosanow_jastrow(s, 1, prm[4].value, sites.data()));
    };
    std::vector<alea::Parameter> prm = std::move(drm);
    prm.push_back({"alpha", j.nosanow_alpha, false});
    std::vector<alea_param> ids;
    for (int k = 0; k < 4; ++k) ids.push_back(alea_param{0, k, nullptr, nullptr, 0, 0.0, 0.0, 0.0});
    ids.push_back(alea_param{1, 0, nullptr, nullptr, 0, 0.0, 0.0, 0.0});
    return std::make_unique<alea_amd::DeviceComponent>("device jastrows", std::move(spec), std::move(prm),
                                                        std::move(ids), false, false);
}

struct EwaldTables {
    int nk = 0, n_knots = 0;
    std::vector<int> n1, n2;
    std::vector<double> kx, ky, weight, scaled, scalars;
};

EwaldTables ewald_tables(const Cell& c, double tol) {
    EwaldTables t;
    const int cap_k = 20000, cap_knots = 20000;
    t.n1.resize(cap_k);
    t.n2.resize(cap_k);
    t.kx.resize(cap_k);
    t.ky.resize(cap_k);
    t.weight.resize(cap_k);
    t.scaled.resize(3 * cap_knots);
    t.scalars.resize(6);
    if (ref_ewald_setup(c.lattice.data(), tol, 1.0, cap_k, &t.nk, t.n1.data(), t.n2.data(), t.kx.data(), t.ky.data(),
                        t.weight.data(), cap_knots, &t.n_knots, t.scaled.data(), t.scalars.data()) != 0)
        throw std::runtime_error("ewald setup failed");
    return t;
}

void jastrow_scenario() {
    std::mt19937_64 rng(20260412);
    const Cell c = make_cell(10, 2.0);
    const std::size_t nw = 24, n_up = 5;
    JastrowSpec j;
    j.drummond_L = 0.45 * c.lattice[0];
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (std::size_t i = 0; i < c.np - n_up; ++i) {
        j.sites.push_back(u(rng) * c.lattice[0]);
        j.sites.push_back(u(rng) * c.lattice[3]);
    }
    Pair P;
    P.c = &c;
    for (Side* s : {&P.a, &P.b}) s->make(nw, c);
    // side a: Padé(0,1), Drummond(0,1), Nosanow(1), Padé(0,0), all native
    P.a.psi.add_component(native_pade(c, 0, 1, j.pade[0], j.pade[1], true));
    P.a.psi.add_component(std::make_unique<alea::DrummondJastrow>(3, 1.0, c.groups[0], c.groups[1],
                                                                   drummond_params(j)));
    P.a.psi.add_component(std::make_unique<alea::NosanowJastrow>(
        2, c.groups[1].first, j.sites, c.cell, std::vector<alea::Parameter>{{"alpha", j.nosanow_alpha, false}}));
    P.a.psi.add_component(native_pade(c, 0, 0, 0.25, 2.0, false));
    // side b: the same Padés native, Drummond and Nosanow as one device component
    P.b.psi.add_component(native_pade(c, 0, 1, j.pade[0], j.pade[1], true));
    P.b.psi.add_component(device_jastrows(c, j, drummond_params(j), nw + 8));
    P.b.psi.add_component(native_pade(c, 0, 0, 0.25, 2.0, false));
    expect_true("jastrow: same free parameter count",
                P.a.psi.total_optimizable_params() == P.b.psi.total_optimizable_params() &&
                    P.a.psi.total_optimizable_params() == 6);

    const auto pos = uniform_positions(c, nw, rng);
    for (Side* s : {&P.a, &P.b}) {
        load(*s->ws, c, pos);
        s->psi.prepare(nw, c.np, 2);
        s->psi.evaluate_batch(*s->ws);
    }
    compare_totals("jastrow: evaluate", P.a, P.b);
    const double sigma = 0.6;
    sweep("jastrow: eager sweep", P, rng, sigma, false, {{0, nw}});
    sweep("jastrow: deferred sweep over two crowds", P, rng, sigma, true, {{0, 11}, {11, nw}});
    sweep("jastrow: deferred sweep refreshing each particle", P, rng, sigma, true, {{0, 13}, {13, nw}}, true);
    reevaluate("jastrow: after two sweeps", P);
    expect("jastrow: d log psi / dp", param_derivs(P.b), param_derivs(P.a));

    // branching: copies, then a shrink, then a grow within the capacity
    branch("jastrow: permutation", P, rng, nw);
    sweep("jastrow: sweep after the permutation", P, rng, sigma, false, {{0, 7}, {7, nw}});
    branch("jastrow: shrink", P, rng, nw - 5);
    sweep("jastrow: sweep after the shrink", P, rng, sigma, true, {{0, nw - 5}});
    branch("jastrow: grow", P, rng, nw + 6);
    sweep("jastrow: sweep after the growth", P, rng, sigma, false, {{0, 15}, {15, nw + 6}});
    reevaluate("jastrow: after branching", P);
    expect("jastrow: d log psi / dp after branching", param_derivs(P.b), param_derivs(P.a));

    // clone: a fresh trial on its own walker set evaluates to the same totals
    {
        const std::size_t n = P.a.ws->num_walkers();
        Pair Q;
        Q.c = &c;
        const auto cur = copy_of(P.a.ws->positions());
        for (auto [src, dst] : {std::pair{&P.a, &Q.a}, std::pair{&P.b, &Q.b}}) {
            dst->psi = src->psi.clone();
            dst->make(n, c);
            load(*dst->ws, c, cur);
            dst->psi.prepare(n, c.np, 2);
            dst->psi.evaluate_batch(*dst->ws);
        }
        compare_totals("jastrow: clone", Q.a, Q.b);
        compare_totals("jastrow: clone against its source", Q.b, P.b);
        sweep("jastrow: sweep of the clone", Q, rng, sigma, false, {{0, n}});
    }

    // an optimizer step: new parameter values, rebuild, evaluate
    for (Side* s : {&P.a, &P.b}) {
        for (const auto& comp : s->psi.components())
            for (auto& prm : comp->parameters()) {
                if (prm.name == "a3") prm.value *= 1.5;
                if (prm.name == "L") prm.value = 0.4 * c.lattice[0];
                if (prm.name == "alpha") prm.value = 0.7;
            }
        s->psi.rebuild();
        s->psi.evaluate_batch(*s->ws);
    }
    compare_totals("jastrow: after rebuild", P.a, P.b);
    expect("jastrow: d log psi / dp after rebuild", param_derivs(P.b), param_derivs(P.a));

    // the Hamiltonian: Kinetic + EwaldTerm against Kinetic + the device Ewald
    const double tol = 1e-6;
    const EwaldTables t = ewald_tables(c, tol);
    alea::Hamiltonian ha, hb;
    ha.add_component(std::make_unique<alea::Kinetic>(c.groups));
    ha.add_component(std::make_unique<alea::EwaldTerm>(c.groups, 1.0, c.cell, tol));
    hb.add_component(std::make_unique<alea::Kinetic>(c.groups));
    alea_amd::DeviceSystemSpec vs;
    vs.np = c.np;
    vs.lattice = c.lattice;
    vs.registrar = [t, n_up, np = c.np](alea_system* s, std::span<const alea::Parameter>) {
        alea_amd::component_check(alea_add_group(s, 0, static_cast<int>(n_up), -1.0, 1.0));
        alea_amd::component_check(alea_add_group(s, static_cast<int>(n_up), static_cast<int>(np), -1.0, 1.0));
        alea_amd::component_check(alea_add_ewald(s, t.nk, t.n1.data(), t.n2.data(), t.kx.data(), t.ky.data(),
                                                 t.weight.data(), t.n_knots, t.scaled.data(), t.scalars[0],
                                                 t.scalars[1], t.scalars[2], t.scalars[3]));
    };
    hb.add_component(std::make_unique<alea_amd::DevicePotential>("device ewald", vs));
    const std::size_t n = P.a.ws->num_walkers();
    for (auto [s, h] : {std::pair{&P.a, &ha}, std::pair{&P.b, &hb}}) {
        s->psi.refresh_derivatives(*s->ws);
        auto le = s->ws->local_energies();
        std::fill(le.begin(), le.end(), 0.0);
        h->evaluate_batch(*s->ws, 0, 9);
        h->evaluate_batch(*s->ws, 9, n);
    }
    expect("jastrow: local energies", P.b.ws->local_energies(), P.a.ws->local_energies());
    {
        std::vector<double> va(n, 0.0), vb(n, 0.0);
        ha.evaluate_potential(*P.a.ws, va);
        hb.evaluate_potential(*P.a.ws, vb);
        expect("ewald: potential", vb, va);
        std::vector<double> da, db, scratch(n);
        for (std::size_t p : {0UL, 4UL, 7UL})
            for (std::size_t k : {0UL, 5UL, n - 1}) {
                const double prop[2] = {0.3 * c.lattice[0] + 0.1 * p, 0.8 * c.lattice[3] - 0.05 * k};
                da.push_back(ha.delta_potential_particle(*P.a.ws, p, k, prop, scratch));
                db.push_back(hb.delta_potential_particle(*P.a.ws, p, k, prop, scratch));
            }
        expect("ewald: delta_potential_particle", db, da);
    }
    std::printf("jastrow scenario: %d checks so far\n", g_checks);
}

// ── Scenario 2: determinant blocks, the adapter against the C-ABI system ─
// Side a is a TrialWavefunction over one "whole system" component: the
// C-ABI system with Padé(0,1) and both blocks.  Side b: the native Padé(0,1)
// plus a DeviceComponent holding the two blocks.
void slater_scenario(bool twisted) {
    const std::string tag = twisted ? "twisted slater" : "plane-wave slater";
    std::mt19937_64 rng(twisted ? 77 : 78);
    const Cell c = make_cell(10, 1.5);
    const std::size_t nw = 20, n_up = 5;
    const double g = 2.0 * M_PI / c.lattice[0];
    const double th[2] = {0.21 * g, -0.13 * g};
    const double kv[5][2] = {{0, 0}, {g, 0}, {-g, 0}, {0, g}, {0, -g}};
    std::vector<double> q, pw;
    std::vector<unsigned char> is_sin;
    for (auto& k : kv) {
        q.push_back(k[0] + th[0]);
        q.push_back(k[1] + th[1]);
    }
    // real plane waves: 1, cos gx, sin gx, cos gy, sin gy
    pw = {0, 0, g, 0, g, 0, 0, g, 0, g};
    is_sin = {0, 0, 1, 0, 1};
    const double pade[2] = {0.2, 1.0};
    auto blocks = [=](alea_system* s) {
        for (int b = 0; b < 2; ++b) {
            const int first = b * static_cast<int>(n_up), last = first + static_cast<int>(n_up);
            if (twisted)
                alea_amd::component_check(alea_add_slater_twisted(s, first, last, q.data()));
            else
                alea_amd::component_check(alea_add_slater_plane_wave(s, first, last, pw.data(), is_sin.data()));
        }
    };
    auto groups = [n_up, np = c.np](alea_system* s) {
        alea_amd::component_check(alea_add_group(s, 0, static_cast<int>(n_up), -1.0, 1.0));
        alea_amd::component_check(alea_add_group(s, static_cast<int>(n_up), static_cast<int>(np), -1.0, 1.0));
    };
    Pair P;
    P.c = &c;
    for (Side* s : {&P.a, &P.b}) s->make(nw, c);
    {
        alea_amd::DeviceSystemSpec all;
        all.np = c.np;
        all.lattice = c.lattice;
        all.registrar = [=](alea_system* s, std::span<const alea::Parameter>) {
            groups(s);
            alea_amd::component_check(alea_add_pade_jastrow(s, 0, 1, 1, 1, pade, 0, nullptr, nullptr, nullptr,
                                                            nullptr, nullptr, 0, nullptr, 0.0, 0.0, 0.0));
            blocks(s);
        };
        P.a.psi.add_component(std::make_unique<alea_amd::DeviceComponent>(
            "whole system", all, std::vector<alea::Parameter>{}, std::vector<alea_param>{}, true, twisted));
    }
    P.b.psi.add_component(native_pade(c, 0, 1, pade[0], pade[1], false));
    {
        alea_amd::DeviceSystemSpec dets;
        dets.np = c.np;
        dets.lattice = c.lattice;
        dets.registrar = [=](alea_system* s, std::span<const alea::Parameter>) {
            groups(s);
            blocks(s);
        };
        P.b.psi.add_component(std::make_unique<alea_amd::DeviceComponent>(
            "device determinants", dets, std::vector<alea::Parameter>{}, std::vector<alea_param>{}, true, twisted));
    }
    expect_true(tag + ": nodal", P.b.psi.components()[1]->is_nodal());
    expect_true(tag + ": complex flag", P.b.psi.is_complex() == twisted);

    // the same positions evaluated on a C-ABI system directly: the whole
    // system component reports exactly what the library reports
    const auto pos = uniform_positions(c, nw, rng);
    for (Side* s : {&P.a, &P.b}) {
        load(*s->ws, c, pos);
        s->psi.prepare(nw, c.np, 2);
        s->psi.evaluate_batch(*s->ws);
    }
    {
        alea_system* s = nullptr;
        alea_amd::component_check(alea_create(&s, static_cast<int>(nw), static_cast<int>(c.np), 2, c.lattice.data(),
                                              ALEA_FP64));
        groups(s);
        alea_amd::component_check(alea_add_pade_jastrow(s, 0, 1, 1, 1, pade, 0, nullptr, nullptr, nullptr, nullptr,
                                                        nullptr, 0, nullptr, 0.0, 0.0, 0.0));
        blocks(s);
        alea_amd::component_check(alea_set_positions(s, P.a.ws->positions().data()));
        alea_amd::component_check(alea_evaluate(s));
        std::vector<double> lp(nw), gr(c.np * 2 * nw), lap(c.np * 2 * nw);
        alea_amd::component_check(alea_get_log_psi(s, lp.data()));
        alea_amd::component_check(alea_get_gradients(s, gr.data(), lap.data()));
        expect(tag + ": whole-system component reports the library's log psi", P.a.ws->log_psi(), lp);
        expect(tag + ": whole-system component reports the library's gradients", P.a.ws->gradients(), gr);
        expect(tag + ": whole-system component reports the library's laplacians", P.a.ws->laplacians(), lap);
        alea_destroy(s);
    }
    compare_totals(tag + ": evaluate", P.a, P.b);
    const double sigma = 0.5;
    sweep(tag + ": eager sweep", P, rng, sigma, false, {{0, 8}, {8, nw}});
    sweep(tag + ": deferred sweep", P, rng, sigma, true, {{0, nw}});
    sweep(tag + ": deferred sweep refreshing each particle", P, rng, sigma, true, {{0, 6}, {6, nw}}, true);
    branch(tag + ": permutation", P, rng, nw - 3);
    sweep(tag + ": sweep after the permutation", P, rng, sigma, false, {{0, 9}, {9, nw - 3}});
    reevaluate(tag + ": after the sweeps", P);
    std::printf("%s scenario: %d checks so far\n", tag.c_str(), g_checks);
}

}  // namespace

int main() {
    try {
        jastrow_scenario();
        slater_scenario(false);
        slater_scenario(true);
    } catch (const std::exception& e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 2;
    }
    std::printf("%d checks, %d failures\n", g_checks, g_failures);
    return g_failures == 0 ? 0 : 1;
}
