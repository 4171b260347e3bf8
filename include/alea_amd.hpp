// alea_amd C++20 host interface over the C-ABI (alea_amd.h).
//
// The reference's host-side interfaces for the walker sweep, kept so that its
// drivers can run on the device batch unchanged in shape:
//   WalkerSet           src/population/walker_set.h:30-223 (the batch, its SoA
//                       planes, apply_permutation, branch)
//   TrialWavefunction   src/wavefunction/trial_wavefunction.h:17-130
//                       (add_component, evaluate_batch, ratio_batch,
//                       apply_move_batch, refresh_derivatives)
//   Hamiltonian         src/hamiltonian/hamiltonian.h:20-40 (add_component,
//                       evaluate_batch)
//   PopulationControl   src/qmc/population_control.h:150-231
//   DMC::run_one_step   src/qmc/dmc.cpp:693-739 over advance_measure (bare
//                       reweighting, dmc.cpp:339-419), one crowd
// One WalkerSet is one device system: its TrialWavefunction and Hamiltonian
// register their factors on it, and every batch call runs HIP kernels.
// Status codes become the reference's exceptions: ALEA_ERR_INVALID ->
// std::invalid_argument, others -> std::runtime_error.
//
// Host arrays use the reference's walker-fastest SoA layouts:
//   positions, gradients, laplacians   [p*dim*nw + d*nw + w]
//   new_pos, new_grad, new_lap         [d*nw + w]
#ifndef ALEA_AMD_HPP
#define ALEA_AMD_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <limits>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "alea_amd.h"

namespace alea_amd {

inline void check(int rc) {
    if (rc == ALEA_OK) return;
    const std::string msg = alea_last_error();
    if (rc == ALEA_ERR_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

enum class Precision { fp64 = ALEA_FP64, mixed = ALEA_MIXED };

// ParticleGroup (src/population/particle_group.h:17), isotropic mass.
struct ParticleGroup {
    std::size_t first = 0, last = 0;
    double charge = -1.0;
    double mass = 1.0;
};

// ── factor descriptors: the reference's constructor arguments ─────────────
// Breakup tables of an image-summed factor or of the Ewald term: the k set,
// weights, quintic spline (3 per knot), r_cut, 1/delta.
struct KSpaceTable {
    std::vector<int> n1, n2;
    std::vector<double> kx, ky, weights, spline;
    double r_cut = 0.0, inv_delta = 0.0;
};

// PadeJastrow(num_order, den_order, group_a, group_b, params) — pade_jastrow.h:60;
// params = [a_1..a_N, b_1..b_M]; image set: PeriodicPairSum (periodic_pair_sum.h:45)
struct PadeJastrow {
    int num_order = 1, den_order = 1;
    int group_a = 0, group_b = 0;
    std::vector<double> params;
    bool image_sum = false;
    KSpaceTable image;
    double u_inf = 0.0;
};
// DrummondJastrow(C, gamma, group_a, group_b, params) — drummond_jastrow.h:77
struct DrummondJastrow {
    int C = 3;
    double gamma = 0.0, L = 0.0;
    int group_a = 0, group_b = 0;
    std::vector<double> params;
};
// NosanowJastrow(first, sites, {alpha}) — nosanow_jastrow.h:55
struct NosanowJastrow {
    int group = 0;
    double alpha = 0.0;
    std::vector<double> sites;
};
// McMillan u = -(b/r)^n (no counterpart in the reference)
struct McMillanJastrow {
    int group_a = 0, group_b = 0;
    double b = 1.0;
    int n = 5;
};
// SlaterDeterminant + PlaneWaveOrbitalSet / GaussianOrbitalSet, and
// ComplexSlaterDeterminant + TwistedPlaneWaveOrbitalSet (orbital_set.h,
// complex_orbital_set.h) on particles [first, last)
struct PlaneWaveSlater {
    std::size_t first = 0, last = 0;
    std::vector<double> kvecs;
    std::vector<unsigned char> is_sin;
};
struct GaussianSlater {
    std::size_t first = 0, last = 0;
    std::vector<double> centers, alphas;
};
struct TwistedSlater {
    std::size_t first = 0, last = 0;
    std::vector<double> q;
};
// EwaldTerm with its EwaldSetup (ewald_term.h:70, ewald_breakup.h:186)
struct EwaldTerm {
    KSpaceTable table;
    double v_long_zero = 0.0, sr_k0 = 0.0;
};
// CoulombTerm(group_a, group_b, epsilon) — coulomb.h:28
struct CoulombTerm {
    int group_a = 0, group_b = 0;
    double epsilon = 1.0;
};
// InversePowerTerm(group_a, group_b, strength, exponent, cell) — inverse_power.h:56
struct InversePowerTerm {
    int group_a = 0, group_b = 0;
    double strength = 1.0, exponent = 3.0;
};

// xoshiro256++ state of a stream seeded as XoshiroRandom(seed)
// (xoshiro_random.h:26-29) from stream_seed(seed, stream, index) (random.h:52-56).
struct XoshiroState {
    std::array<std::uint64_t, 4> s{};
    static constexpr std::uint64_t splitmix64(std::uint64_t x) {
        x += 0x9e3779b97f4a7c15ULL;
        x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
        x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
        return x ^ (x >> 31);
    }
    static XoshiroState from_seed(std::uint64_t seed) {
        XoshiroState r;
        for (std::uint64_t i = 0; i < 4; ++i) r.s[i] = splitmix64(seed + i * 0x9e3779b97f4a7c15ULL);
        return r;
    }
    static XoshiroState stream(std::uint64_t base, std::uint64_t stream, std::uint64_t index) {
        return from_seed(splitmix64(splitmix64(splitmix64(base) ^ stream) ^ index));
    }
};

// ── WalkerSet ──────────────────────────────────────────────────────────────
class WalkerSet {
public:
    // WalkerSet(nw, np, dim, cell): lattice = dim*dim primitive vectors, or
    // empty for an open cell.  nw is the capacity (DMC populations move below it).
    WalkerSet(std::size_t nw, std::size_t np, std::size_t dim, std::span<const double> lattice = {},
              Precision precision = Precision::fp64)
        : np_(np), dim_(dim) {
        if (!lattice.empty() && lattice.size() != dim * dim)
            throw std::invalid_argument("WalkerSet: lattice needs dim*dim components");
        check(alea_create(&h_, static_cast<int>(nw), static_cast<int>(np), static_cast<int>(dim),
                          lattice.empty() ? nullptr : lattice.data(), static_cast<int>(precision)));
        new_pos_.assign(dim * nw, 0.0);
        capacity_ = nw;
    }
    ~WalkerSet() { alea_destroy(h_); }
    WalkerSet(const WalkerSet&) = delete;
    WalkerSet& operator=(const WalkerSet&) = delete;
    WalkerSet(WalkerSet&& o) noexcept { *this = std::move(o); }
    WalkerSet& operator=(WalkerSet&& o) noexcept {
        std::swap(h_, o.h_);
        np_ = o.np_;
        dim_ = o.dim_;
        capacity_ = o.capacity_;
        new_pos_ = std::move(o.new_pos_);
        log_ratios_ = std::move(o.log_ratios_);
        phase_diffs_ = std::move(o.phase_diffs_);
        new_grad_ = std::move(o.new_grad_);
        new_lap_ = std::move(o.new_lap_);
        return *this;
    }

    std::size_t num_walkers() const {
        int n = 0;
        check(alea_get_num_walkers(h_, &n));
        return static_cast<std::size_t>(n);
    }
    std::size_t n_particles() const { return np_; }
    std::size_t dim() const { return dim_; }
    std::size_t capacity() const { return capacity_; }

    void add_group(const ParticleGroup& g) {
        check(alea_add_group(h_, static_cast<int>(g.first), static_cast<int>(g.last), g.charge, 1.0 / g.mass));
    }
    // WalkerSet::load_all (walker_set.cpp:51-70): positions wrapped into the cell
    void load_positions(std::span<const double> soa) {
        if (soa.size() != np_ * dim_ * num_walkers()) throw std::invalid_argument("load_positions: wrong size");
        check(alea_set_positions(h_, soa.data()));
    }
    std::vector<double> positions() const {
        std::vector<double> out(np_ * dim_ * num_walkers());
        check(alea_get_positions(h_, out.data()));
        return out;
    }
    std::vector<double> log_psi() const {
        std::vector<double> out(num_walkers());
        check(alea_get_log_psi(h_, out.data()));
        return out;
    }
    std::vector<double> weights() const {
        std::vector<double> out(num_walkers());
        check(alea_get_weights(h_, out.data()));
        return out;
    }
    void set_weights(std::span<const double> w) { check(alea_set_weights(h_, w.data())); }

    // ParticleBatch::propose_move_batch (particle_batch.cpp:43-57): the
    // proposed position of p, [dim*nw], read by ratio_batch
    std::span<double> new_pos() {
        new_pos_.resize(dim_ * num_walkers());
        return new_pos_;
    }
    void propose_move_batch(std::size_t p, std::span<const double> np) {
        proposed_ = p;
        std::copy(np.begin(), np.end(), new_pos().begin());
    }
    std::span<const double> log_ratios() const { return log_ratios_; }
    std::span<const double> phase_diffs() const { return phase_diffs_; }
    std::span<const double> new_grad() const { return new_grad_; }
    std::span<const double> new_lap() const { return new_lap_; }

    // WalkerSet::apply_permutation (walker_set.cpp:109-178)
    void apply_permutation(std::span<const std::size_t> src) {
        std::vector<int> s(src.begin(), src.end());
        check(alea_permute(h_, static_cast<int>(s.size()), s.data()));
    }
    // WalkerSet::branch (walker_set.cpp:187-239)
    void branch(XoshiroState& rng, std::size_t target, std::vector<std::size_t>& src) {
        std::vector<int> s(2 * target + 1);
        int n = 0;
        check(alea_branch(h_, rng.s.data(), static_cast<int>(target), &n, s.data()));
        src.assign(s.begin(), s.begin() + n);
    }

    alea_system* handle() const { return h_; }

private:
    friend class TrialWavefunction;
    alea_system* h_ = nullptr;
    std::size_t np_ = 0, dim_ = 0, capacity_ = 0, proposed_ = 0;
    std::vector<double> new_pos_, log_ratios_, phase_diffs_, new_grad_, new_lap_;
};

// ── TrialWavefunction ──────────────────────────────────────────────────────
class TrialWavefunction {
public:
    explicit TrialWavefunction(WalkerSet& ws) : h_(ws.handle()) {}

    void add_component(const PadeJastrow& c) {
        const KSpaceTable& t = c.image;
        const int nk = c.image_sum ? static_cast<int>(t.weights.size()) : 0;
        check(alea_add_pade_jastrow(h_, c.group_a, c.group_b, c.num_order, c.den_order, c.params.data(), nk,
                                    nk ? t.n1.data() : nullptr, nk ? t.n2.data() : nullptr,
                                    nk ? t.kx.data() : nullptr, nk ? t.ky.data() : nullptr,
                                    nk ? t.weights.data() : nullptr, nk ? static_cast<int>(t.spline.size() / 3) : 0,
                                    nk ? t.spline.data() : nullptr, t.r_cut, t.inv_delta, c.u_inf));
    }
    void add_component(const DrummondJastrow& c) {
        check(alea_add_drummond_jastrow(h_, c.group_a, c.group_b, c.C, c.gamma, c.L,
                                        static_cast<int>(c.params.size()), c.params.data()));
    }
    void add_component(const NosanowJastrow& c) {
        check(alea_add_nosanow_jastrow(h_, c.group, c.alpha, c.sites.data()));
    }
    void add_component(const McMillanJastrow& c) {
        check(alea_add_mcmillan_jastrow(h_, c.group_a, c.group_b, c.b, c.n));
    }
    void add_component(const PlaneWaveSlater& c) {
        check(alea_add_slater_plane_wave(h_, static_cast<int>(c.first), static_cast<int>(c.last), c.kvecs.data(),
                                         c.is_sin.data()));
    }
    void add_component(const GaussianSlater& c) {
        check(alea_add_slater_gaussian(h_, static_cast<int>(c.first), static_cast<int>(c.last), c.centers.data(),
                                       c.alphas.data()));
    }
    void add_component(const TwistedSlater& c) {
        check(alea_add_slater_twisted(h_, static_cast<int>(c.first), static_cast<int>(c.last), c.q.data()));
    }

    // evaluate_batch (trial_wavefunction.cpp:19-35)
    void evaluate_batch(WalkerSet& ws) {
        (void)ws;
        check(alea_evaluate(h_));
    }
    // ratio_batch (trial_wavefunction.cpp:46-78) for the move proposed on ws:
    // fills ws.log_ratios(), phase_diffs(), new_grad(), new_lap()
    void ratio_batch(WalkerSet& ws, std::size_t p) {
        const std::size_t nw = ws.num_walkers(), d = ws.dim();
        if (p != ws.proposed_) throw std::invalid_argument("ratio_batch: no move of this particle was proposed");
        ws.log_ratios_.resize(nw);
        ws.phase_diffs_.resize(nw);
        ws.new_grad_.resize(d * nw);
        ws.new_lap_.resize(d * nw);
        check(alea_ratio(h_, static_cast<int>(p), ws.new_pos_.data(), ws.log_ratios_.data(), ws.phase_diffs_.data(),
                         ws.new_grad_.data(), ws.new_lap_.data()));
    }
    // apply_move_batch + WalkerSet::apply_move (vmc.cpp:147-152)
    void apply_move_batch(WalkerSet& ws, std::size_t p, std::span<const bool> mask) {
        std::vector<unsigned char> m(mask.begin(), mask.end());
        if (m.size() != ws.num_walkers()) throw std::invalid_argument("apply_move_batch: mask size");
        check(alea_accept(h_, static_cast<int>(p), m.data()));
    }

    // The same over the walkers [w_begin, w_end) with the batch's walker
    // stride (wavefunction_component.h:96-110; npos: to the end): only the
    // range's entries of the log-ratio, phase and new-derivative planes are
    // written, so disjoint ranges (DMC crowds) fill one batch between them.
    static constexpr std::size_t npos = static_cast<std::size_t>(-1);
    void ratio_batch(WalkerSet& ws, std::size_t p, std::size_t w_begin, std::size_t w_end = npos) {
        const std::size_t nw = ws.num_walkers(), d = ws.dim();
        if (p != ws.proposed_) throw std::invalid_argument("ratio_batch: no move of this particle was proposed");
        if (ws.log_ratios_.size() != nw) ws.log_ratios_.assign(nw, 0.0);
        if (ws.phase_diffs_.size() != nw) ws.phase_diffs_.assign(nw, 0.0);
        if (ws.new_grad_.size() != d * nw) ws.new_grad_.assign(d * nw, 0.0);
        if (ws.new_lap_.size() != d * nw) ws.new_lap_.assign(d * nw, 0.0);
        check(alea_ratio_range(h_, static_cast<int>(p), range_int(w_begin), range_int(w_end), ws.new_pos_.data(),
                               ws.log_ratios_.data(), ws.phase_diffs_.data(), ws.new_grad_.data(),
                               ws.new_lap_.data()));
    }
    void apply_move_batch(WalkerSet& ws, std::size_t p, std::span<const bool> mask, std::size_t w_begin,
                          std::size_t w_end = npos) {
        std::vector<unsigned char> m(mask.begin(), mask.end());
        if (m.size() != ws.num_walkers()) throw std::invalid_argument("apply_move_batch: mask size");
        check(alea_accept_range(h_, static_cast<int>(p), range_int(w_begin), range_int(w_end), m.data()));
    }
    // refresh_derivatives: the gradient and Laplacian planes of log|psi|
    void refresh_derivatives(WalkerSet& ws, std::vector<double>& grad, std::vector<double>& lap) {
        const std::size_t n = ws.n_particles() * ws.dim() * ws.num_walkers();
        grad.resize(n);
        lap.resize(n);
        check(alea_get_gradients(h_, grad.data(), lap.data()));
    }
    // compute_parameter_derivatives (trial_wavefunction.cpp:150-160) for the
    // listed free parameters (factors counted in the order they were added,
    // each determinant block one factor): derivs [n][nw] of d log|psi| / d p
    void compute_parameter_derivatives(WalkerSet& ws, std::span<const alea_param> params,
                                       std::vector<double>& derivs) {
        derivs.resize(params.size() * ws.num_walkers());
        check(alea_parameter_derivatives(h_, params.data(), static_cast<int>(params.size()), derivs.data()));
    }

private:
    static int range_int(std::size_t w) { return w == npos ? -1 : static_cast<int>(w); }
    alea_system* h_;
};

// ── Hamiltonian ────────────────────────────────────────────────────────────
class Hamiltonian {
public:
    explicit Hamiltonian(WalkerSet& ws) : h_(ws.handle()) {}
    void add_component(const EwaldTerm& c) {
        const KSpaceTable& t = c.table;
        check(alea_add_ewald(h_, static_cast<int>(t.weights.size()), t.n1.data(), t.n2.data(), t.kx.data(),
                             t.ky.data(), t.weights.data(), static_cast<int>(t.spline.size() / 3), t.spline.data(),
                             t.r_cut, t.inv_delta, c.v_long_zero, c.sr_k0));
    }
    void add_component(const CoulombTerm& c) { check(alea_add_coulomb(h_, c.group_a, c.group_b, c.epsilon)); }
    void add_component(const InversePowerTerm& c) {
        check(alea_add_inverse_power(h_, c.group_a, c.group_b, c.strength, c.exponent));
    }
    // evaluate_batch (hamiltonian.cpp:18-29): E_L of every walker (and Im E_L
    // of a complex trial when im is given)
    std::vector<double> evaluate_batch(WalkerSet& ws, std::vector<double>* im = nullptr) {
        std::vector<double> e(ws.num_walkers());
        if (im) im->resize(e.size());
        check(alea_local_energy(h_, e.data(), im ? im->data() : nullptr));
        return e;
    }
    // local_energy_parameter_derivatives (hamiltonian.cpp:189-196) with the
    // wavefunction's gradient derivatives (trial_wavefunction.cpp:162-185)
    // formed on the device in the same call: out [n][nw] of dE_L / d p, and
    // its imaginary part for a complex trial in im when given
    void local_energy_parameter_derivatives(WalkerSet& ws, std::span<const alea_param> params,
                                            std::vector<double>& out, std::vector<double>* im = nullptr) {
        out.resize(params.size() * ws.num_walkers());
        if (im) im->resize(out.size());
        check(alea_local_energy_parameter_derivatives(h_, params.data(), static_cast<int>(params.size()),
                                                      out.data(), im ? im->data() : nullptr));
    }
    // delta_potential_particle (hamiltonian.cpp:105-117), for every walker of
    // [w_begin, w_end) at once: V with particle p at proposed ([dim][nw]) minus
    // V now, nothing committed
    std::vector<double> delta_potential_particle(WalkerSet& ws, int p, std::span<const double> proposed,
                                                 int w_begin = 0, int w_end = -1) {
        std::vector<double> d(ws.num_walkers());
        if (proposed.size() < 2 * d.size()) throw std::invalid_argument("proposal must be [dim][nw]");
        check(alea_delta_potential(h_, p, w_begin, w_end, proposed.data(), d.data()));
        return d;
    }

private:
    alea_system* h_;
};

// ── VMC sweeps ─────────────────────────────────────────────────────────────
// VMC::sweep_walkers (vmc.cpp:69-155): n fused sweeps, symmetric proposals of
// width sqrt(tau/m), or with use_drift (VMCConfig::use_drift, random_walk:
// langevin) the Langevin proposal and its Metropolis-Hastings correction;
// tape in draw_sweep_randoms order or the device streams seeded here with
// stream_seed(seed, "vmc-thrd", walker).
class VMC {
public:
    VMC(WalkerSet& ws, double tau, bool use_drift = false) : ws_(ws), tau_(tau), use_drift_(use_drift) {}
    void seed(std::uint64_t seed, std::uint64_t stream = 0x766d632d74687264ULL) {
        check(alea_seed(ws_.handle(), seed, stream));
    }
    long long sweep(int n_sweeps, const double* tape = nullptr) {
        long long acc = 0;
        if (use_drift_)
            check(alea_vmc_drift_sweeps(ws_.handle(), n_sweeps, tau_, tape, &acc));
        else
            check(alea_vmc_sweeps(ws_.handle(), n_sweeps, tau_, tape, &acc));
        return acc;
    }

private:
    WalkerSet& ws_;
    double tau_;
    bool use_drift_;
};

// ── PopulationControl (population_control.cpp:14-108) ────────────────────
enum class BranchCutoff { universal, drv, none };

class PopulationControl {
public:
    PopulationControl(double tau, int n_substeps, double generations, std::size_t target_walkers,
                      BranchCutoff cutoff = BranchCutoff::universal, double alpha = 2.0,
                      std::size_t n_particles = 1)
        : tau_(tau), cutoff_(cutoff), alpha_(alpha), n_particles_(static_cast<double>(n_particles)),
          target_walkers_(target_walkers) {
        if (cutoff == BranchCutoff::universal && !(std::isfinite(alpha) && alpha > 0.0))
            throw std::invalid_argument("dmc: the universal branch cutoff needs a finite alpha > 0");
        if (n_particles == 0) throw std::invalid_argument("dmc: the branch cutoff needs at least one particle");
        const double t_step = tau * static_cast<double>(std::max(1, n_substeps));
        feedback_rate_ = (t_step > 0.0 && generations > 0.0) ? 1.0 / (generations * t_step) : 0.0;
        if (cutoff == BranchCutoff::drv && tau > 0.0) set_cutoff(2.0 / std::sqrt(tau));
    }
    void start(double energy) {
        etrial_ = eref_ = energy;
        eref_sum_ = eref_wsum_ = 0.0;
    }
    void set_energy_spread(double sigma) {
        sigma_ = sigma;
        if (cutoff_ != BranchCutoff::universal) return;
        if (!(std::isfinite(sigma) && sigma > 0.0 && tau_ > 0.0)) {
            set_cutoff(kInf);
            return;
        }
        set_cutoff(alpha_ * std::sqrt(sigma * std::sqrt(n_particles_) / tau_));
    }
    double branch_weight(double e_new, double e_old) const {
        double half_tau = 0.5 * tau_;
        const double x = std::max(eref_ - e_new, eref_ - e_old);
        if (x > e_max_)
            half_tau = 0.0;
        else if (x > e_cut_)
            half_tau *= 1.0 - (x - e_cut_) / (e_max_ - e_cut_);
        return std::exp(half_tau * (2.0 * etrial_ - e_new - e_old));
    }
    void update(double ensemble_energy, double total_weight, bool accumulate) {
        if (accumulate) {
            eref_sum_ += ensemble_energy * total_weight;
            eref_wsum_ += total_weight;
            if (eref_wsum_ > 0.0) eref_ = eref_sum_ / eref_wsum_;
        } else {
            eref_ = ensemble_energy;
        }
        if (total_weight > 0.0)
            etrial_ = eref_ - feedback_rate_ * std::log(total_weight / static_cast<double>(target_walkers_));
    }
    void begin_accumulation() { eref_sum_ = eref_wsum_ = 0.0; }
    double etrial() const { return etrial_; }
    double eref() const { return eref_; }
    double energy_cutoff() const { return e_cut_; }
    double e_max() const { return e_max_; }
    double tau() const { return tau_; }
    std::size_t target_walkers() const { return target_walkers_; }

private:
    static constexpr double kInf = std::numeric_limits<double>::infinity();
    void set_cutoff(double e) {
        e_cut_ = e;
        e_max_ = 1.5 * e;
    }
    double tau_;
    BranchCutoff cutoff_;
    double alpha_, n_particles_;
    std::size_t target_walkers_;
    double feedback_rate_ = 0.0, sigma_ = 0.0;
    double e_cut_ = kInf, e_max_ = kInf;
    double etrial_ = 0.0, eref_ = 0.0, eref_sum_ = 0.0, eref_wsum_ = 0.0;
};

// robust_spread (statistics.cpp:297-315): 1.4826 x the median absolute deviation
inline double robust_spread(std::vector<double> x) {
    if (x.empty()) return 0.0;
    auto median = [](std::vector<double>& v) {
        const std::size_t n = v.size(), mid = n / 2;
        std::nth_element(v.begin(), v.begin() + static_cast<std::ptrdiff_t>(mid), v.end());
        const double upper = v[mid];
        if (n % 2 == 1) return upper;
        const double lower = *std::max_element(v.begin(), v.begin() + static_cast<std::ptrdiff_t>(mid));
        return 0.5 * (lower + upper);
    };
    std::vector<double> s = x;
    const double centre = median(s);
    for (std::size_t i = 0; i < x.size(); ++i) s[i] = std::abs(x[i] - centre);
    return 1.4826 * median(s);
}

// numerics::block_mean_error (statistics.cpp:18-34): mean and standard error
// of decorrelated block means, sums left to right.
inline std::pair<double, double> block_mean_error(std::span<const double> x) {
    const std::size_t m = x.size();
    if (m == 0) return {0.0, 0.0};
    double sum = 0.0;
    for (double v : x) sum += v;
    const double mean = sum / static_cast<double>(m);
    if (m < 2) return {mean, 0.0};
    double sq = 0.0;
    for (double v : x) sq += (v - mean) * (v - mean);
    return {mean, std::sqrt(sq / static_cast<double>(m - 1) / static_cast<double>(m))};
}

// ── Estimators ─────────────────────────────────────────────────────────────
// PairCorrelation (pair_correlation.cpp), StructureFactor
// (structure_factor.cpp) and Density (density.cpp) behind the Estimator
// interface (estimator.h:48-64).
// The device keeps each block's running sums (alea_estimators_*); stop_block
// normalizes them as the reference does and finalize() returns the
// reference's RunResult series: gofr_r, gofr_<A>_<B>[_error], sofk_k,
// sofk_shell_count, sofk_<A>_<B>[_error], sofk_kx, sofk_ky, sofk_shell,
// rhok_<A>_re/_im[_error], density_edges, density_<A>[_error].
struct EstimatorGroup {
    std::string name;
    std::size_t first = 0, last = 0;
};
struct Series {
    std::string name;
    std::vector<double> values;
};

class Estimators {
public:
    // rmax > 0 adds g(r) with n_bins bins; kmax > 0 adds S(k).  lattice: the
    // cell vectors as given to the WalkerSet (empty when open).
    Estimators(WalkerSet& ws, std::vector<EstimatorGroup> groups, std::span<const double> lattice,
               double rmax = 0.0, int n_bins = 0, double kmax = 0.0)
        : h_(ws.handle()), groups_(std::move(groups)) {
        G_ = groups_.size();
        npg_ = G_ * (G_ + 1) / 2;
        if (lattice.size() == 4) area_ = reduced_area(lattice);
        if (rmax > 0.0) {
            check(alea_add_gofr(h_, rmax, n_bins));
            gofr_ = true;
            n_bins_ = static_cast<std::size_t>(n_bins);
            delta_ = rmax / static_cast<double>(n_bins);
        }
        if (kmax > 0.0) {
            int nk = 0;
            check(alea_add_sofk(h_, kmax, &nk));
            sofk_ = true;
            nk_ = static_cast<std::size_t>(nk);
            kx_.resize(nk_);
            ky_.resize(nk_);
            shell_.resize(nk_);
            check(alea_get_sofk_vectors(h_, kx_.data(), ky_.data(), shell_.data()));
            for (std::size_t k = 0; k < nk_; ++k) {
                if (static_cast<std::size_t>(shell_[k]) == shell_count_.size()) {
                    shell_count_.push_back(0);
                    shell_k_.push_back(std::hypot(kx_[k], ky_[k]));
                }
                ++shell_count_[static_cast<std::size_t>(shell_[k])];
            }
        }
    }

    // Density(groups, dim, rmax, n_bins, per_particle, cell) (density.h:55):
    // n_bins^2 cells per group (per particle, named <group>_p<i>, when
    // per_particle) over the periodic cell, or over [-rmax, rmax)^2 when the
    // cell is open.
    void add_density(double rmax, int n_bins, bool per_particle = false) {
        check(alea_add_density(h_, rmax, n_bins, per_particle ? 1 : 0));
        density_ = true;
        d_names_.clear();
        for (const auto& g : groups_) {
            if (!per_particle) {
                d_names_.push_back(g.name);
                continue;
            }
            for (std::size_t p = g.first; p < g.last; ++p) d_names_.push_back(g.name + "_p" + std::to_string(p - g.first));
        }
        d_rmax_ = rmax;
        d_bins_ = static_cast<std::size_t>(n_bins);
        d_cells_ = d_bins_ * d_bins_;
        if (area_ > 0.0) {
            d_volume_ = area_ / static_cast<double>(d_cells_);
        } else {
            const double delta = 2.0 * rmax / static_cast<double>(n_bins);
            d_volume_ = 1.0 * delta * delta;
        }
        d_mean_.assign(d_names_.size() * d_cells_, 0.0);
        d_m2_.assign(d_names_.size() * d_cells_, 0.0);
        d_blocks_ = 0;
    }

    void start_block() { check(alea_estimators_start_block(h_)); }
    // one measurement: weight 1 per walker (VMC) or the walkers' DMC weights
    void accumulate(bool use_weights = false) { check(alea_estimators_accumulate(h_, use_weights ? 1 : 0)); }

    void stop_block() {
        if (density_) stop_density();
        if (!gofr_ && !sofk_) return;
        std::vector<double> g(npg_ * n_bins_), s(npg_ * nk_), re(G_ * nk_), im(G_ * nk_);
        double wsum[2] = {0.0, 0.0};
        check(alea_estimators_read(h_, gofr_ ? g.data() : nullptr, sofk_ ? s.data() : nullptr,
                                   sofk_ ? re.data() : nullptr, sofk_ ? im.data() : nullptr, wsum));
        if (gofr_) {   // pair_correlation.cpp:112-128
            std::vector<double> blk(g.size(), 0.0);
            if (wsum[0] > 0.0) {
                const double inv_w = 1.0 / wsum[0];
                for (std::size_t k = 0; k < g.size(); ++k) blk[k] = g[k] * inv_w;
            }
            hist_g_.push_back(std::move(blk));
        }
        if (sofk_) {   // structure_factor.cpp:113-141
            const std::size_t ns = shell_count_.size();
            std::vector<double> sb(npg_ * ns, 0.0), rb(G_ * nk_, 0.0), ib(G_ * nk_, 0.0);
            if (wsum[1] > 0.0) {
                const double inv_w = 1.0 / wsum[1];
                std::size_t pid = 0;
                for (std::size_t a = 0; a < G_; ++a)
                    for (std::size_t b = a; b < G_; ++b, ++pid) {
                        const double na = static_cast<double>(groups_[a].last - groups_[a].first);
                        const double nb = static_cast<double>(groups_[b].last - groups_[b].first);
                        const double norm = (na > 0 && nb > 0) ? 1.0 / std::sqrt(na * nb) : 0.0;
                        for (std::size_t k = 0; k < nk_; ++k) {
                            const auto q = static_cast<std::size_t>(shell_[k]);
                            sb[pid * ns + q] += s[pid * nk_ + k] * inv_w * norm / static_cast<double>(shell_count_[q]);
                        }
                    }
                for (std::size_t q = 0; q < G_ * nk_; ++q) {
                    rb[q] = re[q] * inv_w;
                    ib[q] = im[q] * inv_w;
                }
            }
            hist_s_.push_back(std::move(sb));
            hist_re_.push_back(std::move(rb));
            hist_im_.push_back(std::move(ib));
        }
    }

    std::vector<Series> finalize() const {
        std::vector<Series> out;
        if (density_ && d_blocks_ > 0) {   // density.cpp:220-266
            std::vector<double> edges(d_bins_ + 1);
            for (std::size_t b = 0; b <= d_bins_; ++b)
                edges[b] = area_ > 0.0 ? -0.5 + static_cast<double>(b) / static_cast<double>(d_bins_)
                                       : -d_rmax_ + static_cast<double>(b) * (2.0 * d_rmax_ / static_cast<double>(d_bins_));
            out.push_back({"density_edges", std::move(edges)});
            const double n = static_cast<double>(d_blocks_);
            const double err_norm = d_blocks_ >= 2 ? 1.0 / ((n - 1.0) * n) : 0.0;
            for (std::size_t g = 0; g < d_names_.size(); ++g) {
                std::vector<double> m(d_cells_), e(d_cells_);
                for (std::size_t c = 0; c < d_cells_; ++c) {
                    m[c] = d_mean_[g * d_cells_ + c];
                    e[c] = std::sqrt(d_m2_[g * d_cells_ + c] * err_norm);
                }
                out.push_back({"density_" + d_names_[g], std::move(m)});
                out.push_back({"density_" + d_names_[g] + "_error", std::move(e)});
            }
        }
        std::vector<double> per;
        auto stat = [&](const std::vector<std::vector<double>>& hist, std::size_t q) {
            per.resize(hist.size());
            for (std::size_t i = 0; i < hist.size(); ++i) per[i] = hist[i][q];
            return block_mean_error(per);
        };
        if (gofr_ && !hist_g_.empty()) {   // pair_correlation.cpp:140-190
            std::vector<double> r(n_bins_);
            for (std::size_t b = 0; b < n_bins_; ++b) r[b] = (static_cast<double>(b) + 0.5) * delta_;
            out.push_back({"gofr_r", std::move(r)});
            std::size_t pid = 0;
            for (std::size_t a = 0; a < G_; ++a)
                for (std::size_t b = a; b < G_; ++b, ++pid) {
                    const std::size_t na = groups_[a].last - groups_[a].first;
                    const std::size_t nb = groups_[b].last - groups_[b].first;
                    const std::size_t pc = a == b ? na * (na - 1) / 2 : na * nb;
                    if (pc == 0) continue;
                    std::vector<double> g(n_bins_), e(n_bins_);
                    for (std::size_t bin = 0; bin < n_bins_; ++bin) {
                        const double lo = static_cast<double>(bin) * delta_, hi = static_cast<double>(bin + 1) * delta_;
                        const double shell = kPi * (hi * hi - lo * lo);
                        const double dpc = static_cast<double>(pc);
                        const double norm = area_ > 0.0 ? dpc * shell / area_ : dpc * shell;
                        const auto [m, err] = stat(hist_g_, pid * n_bins_ + bin);
                        g[bin] = m / norm;
                        e[bin] = err / norm;
                    }
                    const std::string name = "gofr_" + groups_[a].name + "_" + groups_[b].name;
                    out.push_back({name, std::move(g)});
                    out.push_back({name + "_error", std::move(e)});
                }
        }
        if (sofk_ && !hist_s_.empty()) {   // structure_factor.cpp:169-221
            const std::size_t ns = shell_count_.size();
            auto emit = [&](const std::string& name, const std::vector<std::vector<double>>& hist,
                            std::size_t offset, std::size_t count) {
                std::vector<double> m(count), e(count);
                for (std::size_t q = 0; q < count; ++q) std::tie(m[q], e[q]) = stat(hist, offset + q);
                out.push_back({name, std::move(m)});
                out.push_back({name + "_error", std::move(e)});
            };
            out.push_back({"sofk_k", shell_k_});
            out.push_back({"sofk_shell_count", std::vector<double>(shell_count_.begin(), shell_count_.end())});
            std::size_t pid = 0;
            for (std::size_t a = 0; a < G_; ++a)
                for (std::size_t b = a; b < G_; ++b, ++pid)
                    emit("sofk_" + groups_[a].name + "_" + groups_[b].name, hist_s_, pid * ns, ns);
            out.push_back({"sofk_kx", kx_});
            out.push_back({"sofk_ky", ky_});
            out.push_back({"sofk_shell", std::vector<double>(shell_.begin(), shell_.end())});
            for (std::size_t g = 0; g < G_; ++g) {
                emit("rhok_" + groups_[g].name + "_re", hist_re_, g * nk_, nk_);
                emit("rhok_" + groups_[g].name + "_im", hist_im_, g * nk_, nk_);
            }
        }
        return out;
    }

    std::size_t n_vectors() const { return nk_; }

private:
    // density.cpp:120-170: x = counts / (W_block V_cell), then the singleton
    // Welford update of each cell's mean and second moment
    void stop_density() {
        std::vector<double> h(d_names_.size() * d_cells_);
        double w = 0.0;
        check(alea_estimators_read_density(h_, h.data(), &w));
        const double norm = w > 0.0 ? 1.0 / (w * d_volume_) : 0.0;
        ++d_blocks_;
        if (d_blocks_ == 1) {
            for (std::size_t k = 0; k < h.size(); ++k) {
                d_mean_[k] = h[k] * norm;
                d_m2_[k] = 0.0;
            }
            return;
        }
        const double inv_n = 1.0 / static_cast<double>(d_blocks_);
        const double vw = static_cast<double>(d_blocks_ - 1) * inv_n;
        for (std::size_t k = 0; k < h.size(); ++k) {
            const double x = h[k] * norm;
            const double d = x - d_mean_[k];
            d_mean_[k] += d * inv_n;
            d_m2_[k] += d * (d * vw);
        }
    }

    static constexpr double kPi = 3.14159265358979323846;
    // |det| of the basis SimulationCell keeps (simulation_cell.cpp:16-36,
    // 100-132): axis-aligned rows reordered, any other basis Gauss-reduced
    static double reduced_area(std::span<const double> v) {
        double a[4] = {v[0], v[1], v[2], v[3]};
        const bool row0_axis = (a[0] != 0.0) != (a[1] != 0.0), row1_axis = (a[2] != 0.0) != (a[3] != 0.0);
        const bool ortho = row0_axis && row1_axis && ((a[0] != 0.0) != (a[2] != 0.0));
        if (ortho) {
            if (a[0] == 0.0) {   // rows along y then x: put them in axis order
                std::swap(a[0], a[2]);
                std::swap(a[1], a[3]);
            }
        } else {
            for (int guard = 0; guard < 64; ++guard) {
                const double n0 = a[0] * a[0] + a[1] * a[1];
                const double n1 = a[2] * a[2] + a[3] * a[3];
                if (n1 < n0) {
                    std::swap(a[0], a[2]);
                    std::swap(a[1], a[3]);
                    continue;
                }
                if (n0 == 0.0) break;
                const double mu = std::nearbyint((a[0] * a[2] + a[1] * a[3]) / n0);
                if (mu == 0.0) break;
                a[2] -= mu * a[0];
                a[3] -= mu * a[1];
            }
        }
        return std::abs(a[0] * a[3] - a[1] * a[2]);
    }

    alea_system* h_;
    std::vector<EstimatorGroup> groups_;
    std::size_t G_ = 0, npg_ = 0, n_bins_ = 0, nk_ = 0;
    bool gofr_ = false, sofk_ = false;
    double delta_ = 0.0, area_ = 0.0;
    std::vector<double> kx_, ky_, shell_k_;
    std::vector<int> shell_;
    std::vector<std::size_t> shell_count_;
    std::vector<std::vector<double>> hist_g_, hist_s_, hist_re_, hist_im_;
    bool density_ = false;
    double d_rmax_ = 0.0, d_volume_ = 0.0;
    std::size_t d_bins_ = 0, d_cells_ = 0, d_blocks_ = 0;
    std::vector<double> d_mean_, d_m2_;
    std::vector<std::string> d_names_;
};

// ── DMC step ───────────────────────────────────────────────────────────────
// DMC::run_one_step (dmc.cpp:693-739) with one crowd over advance_measure
// (dmc.cpp:339-419, bare reweighting): n_substeps drift-diffusion sweeps each
// followed by E_L and the branch factor, the step's (sum w E_L, sum w), the
// integer branch from stream_seed(seed, "dmc-brch", 0) and the E_T update.
class DMC {
public:
    DMC(WalkerSet& ws, Hamiltonian& ham, double tau, std::size_t target, int n_substeps = 1,
        double generations = 10.0, BranchCutoff cutoff = BranchCutoff::universal, double alpha = 2.0,
        std::uint64_t seed = 42)
        : ws_(ws), tau_(tau), target_(target), n_substeps_(n_substeps),
          pc_(tau, n_substeps, generations, target, cutoff, alpha, ws.n_particles()),
          branch_rng_(XoshiroState::stream(seed, 0x646d632d62726368ULL, 0)) {
        const std::vector<double> e = ham.evaluate_batch(ws);
        double sum = 0.0;
        for (double x : e) sum += x;
        pc_.start(sum / static_cast<double>(e.size()));   // dmc.cpp:628-635
        pc_.set_energy_spread(robust_spread(e));          // dmc.cpp:640-644
    }
    // tapes: one [np][dim+1][stride] block per substep, or nullptr for the
    // device streams.  Returns (sum w E_L, sum w) of the step.
    std::pair<double, double> run_one_step(bool measure, const double* const* tapes = nullptr,
                                           std::size_t stride = 0) {
        alea_system* h = ws_.handle();
        check(alea_dmc_begin(h));
        long long acc = 0;
        for (int s = 0; s < n_substeps_; ++s) {
            const double* tape = tapes ? tapes[s] : nullptr;
            check(alea_dmc_sweeps(h, 1, tau_, tape, tape ? static_cast<int>(stride) : 0, nullptr, nullptr, &acc));
            check(alea_dmc_weights(h, pc_.etrial(), pc_.eref(), pc_.energy_cutoff(), pc_.e_max(), pc_.tau()));
        }
        double e = 0.0, w = 0.0;
        long long n_cut = 0;
        check(alea_dmc_sums(h, &e, &w, &n_cut));
        // the mixed estimators see the population before the branch, with
        // its DMC weights (dmc.cpp:702-707)
        if (measure && measure_hook_) measure_hook_();
        const double ens_e = w > 0.0 ? e / w : pc_.etrial();
        ws_.branch(branch_rng_, target_, src_);
        pc_.update(ens_e, w, measure);
        return {e, w};
    }
    const PopulationControl& population() const { return pc_; }
    const std::vector<std::size_t>& last_branch() const { return src_; }
    // DMC::set_measure_hook: called on measured steps before the branch,
    // e.g. [&] { estimators.accumulate(true); }
    void set_measure_hook(std::function<void()> hook) { measure_hook_ = std::move(hook); }

private:
    WalkerSet& ws_;
    double tau_;
    std::size_t target_;
    int n_substeps_;
    PopulationControl pc_;
    XoshiroState branch_rng_;
    std::vector<std::size_t> src_;
    std::function<void()> measure_hook_;
};

}  // namespace alea_amd

#endif
