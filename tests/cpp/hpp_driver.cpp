// Test driver of the C++20 host interface (include/alea_amd.hpp).
//
//   hpp_driver selftest
//       host-only checks (no device work): status -> exception mapping of
//       refused inputs, and the host restatements' values printed one per line
//       for the Python tests to compare (xoshiro seeding, robust_spread,
//       PopulationControl).
//   hpp_driver run <spec> <input> <output>
//       builds a walker batch from a text spec through WalkerSet /
//       TrialWavefunction / Hamiltonian, then: evaluate_batch; ratio_batch +
//       apply_move_batch over every particle (odd ones as two crowds); VMC sweeps from a tape; DMC
//       run_one_step with tapes.  Writes every result as "name count values".
//   hpp_driver estimators <spec> <input> <output>
//       the Estimators class (g(r), S(k), density) over given positions and weights;
//       writes every finalized series.
//
// Text files are whitespace-separated tokens; arrays are "count v1 .. vn".
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "alea_amd.hpp"

using namespace alea_amd;

namespace {

struct Reader {
    std::ifstream in;
    explicit Reader(const char* path) : in(path) {
        if (!in) throw std::runtime_error(std::string("cannot open ") + path);
    }
    std::string word() {
        std::string s;
        if (!(in >> s)) throw std::runtime_error("unexpected end of file");
        return s;
    }
    double num() {
        const std::string s = word();
        return std::stod(s);
    }
    long long integer() { return std::stoll(word()); }
    std::vector<double> vec() {
        const long long n = integer();
        std::vector<double> v(static_cast<std::size_t>(n));
        for (auto& x : v) x = num();
        return v;
    }
    std::vector<int> ivec() {
        const long long n = integer();
        std::vector<int> v(static_cast<std::size_t>(n));
        for (auto& x : v) x = static_cast<int>(integer());
        return v;
    }
    KSpaceTable table() {
        KSpaceTable t;
        t.n1 = ivec();
        t.n2 = ivec();
        t.kx = vec();
        t.ky = vec();
        t.weights = vec();
        t.spline = vec();
        t.r_cut = num();
        t.inv_delta = num();
        return t;
    }
};

struct Writer {
    std::ofstream out;
    explicit Writer(const char* path) : out(path) {
        if (!out) throw std::runtime_error(std::string("cannot open ") + path);
        out.precision(17);
    }
    template <class V>
    void put(const std::string& name, const V& v) {
        out << name << ' ' << v.size();
        for (const auto& x : v) out << ' ' << x;
        out << '\n';
    }
    void put(const std::string& name, double x) { out << name << " 1 " << x << '\n'; }
};

template <class E, class F>
bool throws(F&& f) {
    try {
        f();
    } catch (const E&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

int selftest() {
    int bad = 0;
    auto expect = [&](bool ok, const char* what) {
        if (!ok) {
            std::fprintf(stderr, "FAIL: %s\n", what);
            ++bad;
        }
    };
    // refused inputs map to std::invalid_argument before any device work
    expect(throws<std::invalid_argument>([] { WalkerSet ws(0, 4, 2); }), "empty batch refused");
    expect(throws<std::invalid_argument>([] { WalkerSet ws(8, 4, 3); }), "3-D refused");
    expect(throws<std::invalid_argument>([] { WalkerSet ws(8, 400, 2); }), "too many particles refused");
    expect(throws<std::invalid_argument>([] {
               const double lat[3] = {1, 0, 1};
               WalkerSet ws(8, 4, 2, lat);
           }),
           "short lattice refused");
    expect(throws<std::invalid_argument>([] { PopulationControl pc(0.01, 1, 10.0, 100, BranchCutoff::universal, 0.0); }),
           "alpha = 0 refused");
    expect(throws<std::invalid_argument>([] {
               PopulationControl pc(0.01, 1, 10.0, 100, BranchCutoff::none, 2.0, 0);
           }),
           "zero particles refused");
    try {
        check(ALEA_ERR_RUNTIME);
        expect(false, "runtime status throws");
    } catch (const std::runtime_error&) {
    }

    std::cout.precision(17);
    // xoshiro256++ seeding of stream_seed(7, "dmc-brch", 3)
    const XoshiroState x = XoshiroState::stream(7, 0x646d632d62726368ULL, 3);
    std::cout << "xoshiro";
    for (auto w : x.s) std::cout << ' ' << w;
    std::cout << '\n';
    std::cout << "robust_odd " << robust_spread({3.0, -1.0, 4.0, 1.5, 9.0, -2.5, 6.0}) << '\n';
    std::cout << "robust_even " << robust_spread({0.25, -0.5, 2.0, 8.0, 1.0, -3.0}) << '\n';
    // PopulationControl: universal cutoff from a spread, a tapered and a cut weight,
    // then a plain and two accumulating updates
    PopulationControl pc(0.02, 2, 8.0, 100, BranchCutoff::universal, 2.0, 26);
    pc.start(-1.25);
    pc.set_energy_spread(0.4);
    std::cout << "popctl_cut " << pc.energy_cutoff() << ' ' << pc.e_max() << '\n';
    std::cout << "popctl_w " << pc.branch_weight(-1.1, -1.32) << ' ' << pc.branch_weight(-1.25 - 1.1 * pc.energy_cutoff(), -1.2)
              << ' ' << pc.branch_weight(-1.25 - 2.0 * pc.energy_cutoff(), -1.2) << '\n';
    pc.update(-1.3, 97.5, false);
    std::cout << "popctl_u0 " << pc.eref() << ' ' << pc.etrial() << '\n';
    pc.update(-1.28, 103.0, true);
    pc.update(-1.31, 99.0, true);
    std::cout << "popctl_u2 " << pc.eref() << ' ' << pc.etrial() << '\n';
    PopulationControl drv(0.04, 1, 10.0, 50, BranchCutoff::drv);
    std::cout << "popctl_drv " << drv.energy_cutoff() << '\n';
    return bad == 0 ? 0 : 1;
}

int run(const char* spec_path, const char* in_path, const char* out_path) {
    Reader sp(spec_path);
    if (sp.word() != "system") throw std::runtime_error("spec must start with 'system'");
    const auto nw = static_cast<std::size_t>(sp.integer());
    const auto np = static_cast<std::size_t>(sp.integer());
    const auto precision = static_cast<Precision>(sp.integer());
    const std::vector<double> lattice = sp.vec();
    WalkerSet ws(nw, np, 2, lattice, precision);
    TrialWavefunction psi(ws);
    Hamiltonian ham(ws);
    // every parameter that needs no breakup tables: minimum-image Padé
    // coefficients, Drummond coefficients and L, Nosanow alpha, a uniform
    // Gaussian block's width
    std::vector<alea_param> free;
    int factor = 0;
    auto take = [&](int n) {
        for (int i = 0; i < n; ++i) free.push_back(alea_param{factor, i, nullptr, nullptr, 0, 0.0, 0.0, 0.0});
        ++factor;
    };
    for (std::string key = sp.word(); key != "end"; key = sp.word()) {
        if (key == "group") {
            ParticleGroup g;
            g.first = static_cast<std::size_t>(sp.integer());
            g.last = static_cast<std::size_t>(sp.integer());
            g.charge = sp.num();
            g.mass = sp.num();
            ws.add_group(g);
        } else if (key == "pade") {
            PadeJastrow j;
            j.group_a = static_cast<int>(sp.integer());
            j.group_b = static_cast<int>(sp.integer());
            j.num_order = static_cast<int>(sp.integer());
            j.den_order = static_cast<int>(sp.integer());
            j.params = sp.vec();
            j.image_sum = sp.integer() != 0;
            if (j.image_sum) {
                j.image = sp.table();
                j.u_inf = sp.num();
            }
            psi.add_component(j);
            take(j.image_sum ? 0 : j.num_order + j.den_order);
        } else if (key == "drummond") {
            DrummondJastrow j;
            j.group_a = static_cast<int>(sp.integer());
            j.group_b = static_cast<int>(sp.integer());
            j.C = static_cast<int>(sp.integer());
            j.gamma = sp.num();
            j.L = sp.num();
            j.params = sp.vec();
            psi.add_component(j);
            take(static_cast<int>(j.params.size()) + 1);
        } else if (key == "nosanow") {
            NosanowJastrow j;
            j.group = static_cast<int>(sp.integer());
            j.alpha = sp.num();
            j.sites = sp.vec();
            psi.add_component(j);
            take(1);
        } else if (key == "mcmillan") {
            McMillanJastrow j;
            j.group_a = static_cast<int>(sp.integer());
            j.group_b = static_cast<int>(sp.integer());
            j.b = sp.num();
            j.n = static_cast<int>(sp.integer());
            psi.add_component(j);
            take(0);
        } else if (key == "slater_pw") {
            PlaneWaveSlater d;
            d.first = static_cast<std::size_t>(sp.integer());
            d.last = static_cast<std::size_t>(sp.integer());
            d.kvecs = sp.vec();
            for (int v : sp.ivec()) d.is_sin.push_back(static_cast<unsigned char>(v));
            psi.add_component(d);
            take(0);
        } else if (key == "slater_twisted") {
            TwistedSlater d;
            d.first = static_cast<std::size_t>(sp.integer());
            d.last = static_cast<std::size_t>(sp.integer());
            d.q = sp.vec();
            psi.add_component(d);
            take(0);
        } else if (key == "slater_gauss") {
            GaussianSlater d;
            d.first = static_cast<std::size_t>(sp.integer());
            d.last = static_cast<std::size_t>(sp.integer());
            d.centers = sp.vec();
            d.alphas = sp.vec();
            psi.add_component(d);
            take(std::all_of(d.alphas.begin(), d.alphas.end(), [&](double a) { return a == d.alphas[0]; }) ? 1 : 0);
        } else if (key == "ewald") {
            EwaldTerm e;
            e.table = sp.table();
            e.v_long_zero = sp.num();
            e.sr_k0 = sp.num();
            ham.add_component(e);
        } else if (key == "coulomb") {
            CoulombTerm c;
            c.group_a = static_cast<int>(sp.integer());
            c.group_b = static_cast<int>(sp.integer());
            c.epsilon = sp.num();
            ham.add_component(c);
        } else if (key == "inverse_power") {
            InversePowerTerm t;
            t.group_a = static_cast<int>(sp.integer());
            t.group_b = static_cast<int>(sp.integer());
            t.strength = sp.num();
            t.exponent = sp.num();
            ham.add_component(t);
        } else {
            throw std::runtime_error("unknown spec entry " + key);
        }
    }

    Reader in(in_path);
    Writer out(out_path);
    // evaluate_batch + Hamiltonian::evaluate_batch
    ws.load_positions(in.vec());
    psi.evaluate_batch(ws);
    std::vector<double> grad, lap, e_im;
    psi.refresh_derivatives(ws, grad, lap);
    out.put("log_psi0", ws.log_psi());
    out.put("grad0", grad);
    out.put("lap0", lap);
    out.put("e0", ham.evaluate_batch(ws, &e_im));
    out.put("e0_im", e_im);
    // per-particle ratio_batch + apply_move_batch; odd particles go through
    // two crowds, [0, nw/2) and [nw/2, nw), which must fill the batch exactly
    // as one whole-batch call does
    for (std::size_t p = 0; p < np; ++p) {
        ws.propose_move_batch(p, in.vec());
        const std::size_t half = ws.num_walkers() / 2;
        const bool crowds = p % 2 == 1;
        if (crowds) {
            psi.ratio_batch(ws, p, 0, half);
            psi.ratio_batch(ws, p, half);
        } else {
            psi.ratio_batch(ws, p);
        }
        const std::string tag = std::to_string(p);
        out.put("lr" + tag, ws.log_ratios());
        out.put("ph" + tag, ws.phase_diffs());
        out.put("ng" + tag, ws.new_grad());
        out.put("nl" + tag, ws.new_lap());
        const std::vector<double> m = in.vec();
        std::unique_ptr<bool[]> mask(new bool[m.size()]);
        for (std::size_t w = 0; w < m.size(); ++w) mask[w] = m[w] != 0.0;
        const std::span<const bool> ms(mask.get(), m.size());
        if (crowds) {
            psi.apply_move_batch(ws, p, ms, half, TrialWavefunction::npos);
            psi.apply_move_batch(ws, p, ms, 0, half);
        } else {
            psi.apply_move_batch(ws, p, ms);
        }
    }
    out.put("log_psi1", ws.log_psi());
    out.put("e1", ham.evaluate_batch(ws));
    if (!free.empty()) {
        std::vector<double> dlp, de, de_im;
        psi.compute_parameter_derivatives(ws, free, dlp);
        ham.local_energy_parameter_derivatives(ws, free, de, &de_im);
        out.put("dlogpsi", dlp);
        out.put("de", de);
        out.put("de_im", de_im);
    }
    // VMC sweeps from a tape
    const double tau = in.num();
    const int n_sweeps = static_cast<int>(in.integer());
    const std::vector<double> tape = in.vec();
    VMC vmc(ws, tau);
    out.put("vmc_accepted", static_cast<double>(vmc.sweep(n_sweeps, tape.data())));
    out.put("pos2", ws.positions());
    out.put("log_psi2", ws.log_psi());
    out.put("e2", ham.evaluate_batch(ws));
    // DMC steps: target, tau, substeps, steps, the step from which E_ref
    // accumulates, seed, then steps x substeps tapes of [np][3][capacity]
    const auto target = static_cast<std::size_t>(in.integer());
    const double tau_dmc = in.num();
    const int n_sub = static_cast<int>(in.integer());
    const int n_steps = static_cast<int>(in.integer());
    const int measure_from = static_cast<int>(in.integer());
    const auto dmc_seed = static_cast<std::uint64_t>(in.integer());
    std::vector<std::size_t> first(target);
    for (std::size_t i = 0; i < target; ++i) first[i] = i;
    ws.apply_permutation(first);
    DMC dmc(ws, ham, tau_dmc, target, n_sub, 10.0, BranchCutoff::universal, 2.0, dmc_seed);
    for (int step = 0; step < n_steps; ++step) {
        std::vector<std::vector<double>> tapes(static_cast<std::size_t>(n_sub));
        std::vector<const double*> tp;
        for (auto& t : tapes) {
            t = in.vec();
            tp.push_back(t.data());
        }
        const auto [e_sum, w_sum] = dmc.run_one_step(step >= measure_from, tp.data(), ws.capacity());
        const std::string tag = std::to_string(step);
        out.put("dmc_e" + tag, e_sum);
        out.put("dmc_w" + tag, w_sum);
        out.put("dmc_nw" + tag, static_cast<double>(ws.num_walkers()));
        out.put("dmc_etrial" + tag, dmc.population().etrial());
        out.put("dmc_src" + tag, dmc.last_branch());
    }
    out.put("pos3", ws.positions());
    out.put("weights3", ws.weights());
    // Langevin VMC sweeps (use_drift) of the final population from the VMC
    // tape, read with the walker count as its row stride (the tape holds
    // capacity >= walker-count rows)
    VMC langevin(ws, tau, true);
    out.put("langevin_accepted", static_cast<double>(langevin.sweep(n_sweeps, tape.data())));
    out.put("pos_l", ws.positions());
    out.put("log_psi_l", ws.log_psi());
    return 0;
}

// hpp_driver estimators <spec> <input> <output>: the spec's groups only;
// input "rmax n_bins kmax n_blocks steps use_weights" then, per measurement,
// positions (SoA) and weights; output every finalized series.
int estimators(const char* spec_path, const char* in_path, const char* out_path) {
    Reader sp(spec_path);
    if (sp.word() != "system") throw std::runtime_error("spec must start with 'system'");
    const auto nw = static_cast<std::size_t>(sp.integer());
    const auto np = static_cast<std::size_t>(sp.integer());
    const auto precision = static_cast<Precision>(sp.integer());
    const std::vector<double> lattice = sp.vec();
    WalkerSet ws(nw, np, 2, lattice, precision);
    std::vector<EstimatorGroup> groups;
    for (std::string key = sp.word(); key != "end"; key = sp.word()) {
        if (key != "group") throw std::runtime_error("estimators: the spec holds groups only");
        ParticleGroup g;
        g.first = static_cast<std::size_t>(sp.integer());
        g.last = static_cast<std::size_t>(sp.integer());
        g.charge = sp.num();
        g.mass = sp.num();
        ws.add_group(g);
        groups.push_back({"g" + std::to_string(g.first), g.first, g.last});
    }
    Reader in(in_path);
    Writer out(out_path);
    const double rmax = in.num();
    const int n_bins = static_cast<int>(in.integer());
    const double kmax = in.num();
    const int n_blocks = static_cast<int>(in.integer());
    const int steps = static_cast<int>(in.integer());
    const bool use_weights = in.integer() != 0;
    const double d_rmax = in.num();
    const int d_bins = static_cast<int>(in.integer());
    Estimators est(ws, groups, lattice, rmax, n_bins, kmax);
    if (d_bins > 0) est.add_density(d_rmax, d_bins);
    for (int b = 0; b < n_blocks; ++b) {
        est.start_block();
        for (int m = 0; m < steps; ++m) {
            ws.load_positions(in.vec());
            const std::vector<double> w = in.vec();
            if (use_weights) ws.set_weights(w);
            est.accumulate(use_weights);
        }
        est.stop_block();
    }
    for (const Series& s : est.finalize()) out.put(s.name, s.values);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc == 2 && std::string(argv[1]) == "selftest") return selftest();
        if (argc == 5 && std::string(argv[1]) == "run") return run(argv[2], argv[3], argv[4]);
        if (argc == 5 && std::string(argv[1]) == "estimators") return estimators(argv[2], argv[3], argv[4]);
        std::fprintf(stderr, "usage: hpp_driver selftest | run|estimators <spec> <input> <output>\n");
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "hpp_driver: %s\n", e.what());
        return 1;
    }
}
