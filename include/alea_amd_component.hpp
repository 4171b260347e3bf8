// alea_amd factors as components of the reference's own TrialWavefunction and
// Hamiltonian.
//
// Glue compiled inside the reference's build (it includes the reference's
// src/population/walker_set.h, src/wavefunction/wavefunction_component.h and
// src/hamiltonian/hamiltonian_component.h, so the include path is the
// reference's src/ plus this directory).  A DeviceComponent holds any subset
// of the device factors (pair Jastrows, Nosanow, determinant blocks) as ONE
// WavefunctionComponent (wavefunction_component.h:59-272), so it can sit
// beside the reference's CPU components in one TrialWavefunction; a
// DevicePotential holds device potential terms (Ewald, Coulomb, inverse
// powers) as one HamiltonianComponent (hamiltonian_component.h:352-403).
//
// The component contract, as the reference's components keep it:
//   evaluate_batch   adds log psi, gradient and Laplacian planes (and the phase
//                    planes of a complex trial) into the batch totals and keeps
//                    its own copies (log_psi_c_, grad_c_, lap_c_);
//   ratio_batch      adds the log-ratio, the phase change and the moved
//                    particle's new gradient and Laplacian over [w_begin, w_end);
//   apply_move_batch commits the accepted walkers of the range and pushes every
//                    particle's changed gradient and Laplacian into the totals
//                    (the "eager" form: partners included);
//   apply_move_batch_deferred commits and pushes log psi and the moved
//                    particle's rows only; refresh_particle_derivatives and
//                    refresh_derivatives bring the partners up to date;
//   permute_caches   carries the device walkers and the host copies through a
//                    source map (branching).
// Every range call touches only its range's walkers, on the device and here,
// so DMC crowds may share one instance (alea_ratio_range_phase,
// alea_accept_range and alea_get_derivatives_range are range-local).  The
// eager commit reads the range's gradient planes back after each accept, the
// deferred one only at a refresh; the fused sweeps (alea_vmc_sweeps,
// alea_dmc_sweeps) are the fast path.
#ifndef ALEA_AMD_COMPONENT_HPP
#define ALEA_AMD_COMPONENT_HPP

#include <algorithm>
#include <cstddef>
#include <functional>
#include <memory>
#include <mutex>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "alea_amd.h"
#include "hamiltonian/hamiltonian_component.h"
#include "population/walker_set.h"
#include "wavefunction/wavefunction_component.h"

namespace alea_amd {

inline void component_check(int rc) {
    if (rc == ALEA_OK) return;
    const std::string msg = alea_last_error();
    if (rc == ALEA_ERR_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// Registers the groups and factors on a fresh device system, with the
// parameters' current values (a rebuild() after an optimizer step passes the
// new ones).
using Registrar = std::function<void(alea_system*, std::span<const alea::Parameter>)>;

struct DeviceSystemSpec {
    std::size_t np = 0, dim = 2;
    std::vector<double> lattice;      // dim*dim primitive vectors, empty for an open cell
    int precision = ALEA_FP64;
    std::size_t capacity = 0;         // walkers the device system holds (0: the batch's nw)
    Registrar registrar;
};

// Owns one alea_system.
class DeviceSystem {
public:
    DeviceSystem() = default;
    DeviceSystem(const DeviceSystem&) = delete;
    DeviceSystem& operator=(const DeviceSystem&) = delete;
    ~DeviceSystem() { reset(); }
    void reset() {
        if (s_) (void)alea_destroy(s_);
        s_ = nullptr;
        nw_ = capacity_ = 0;
    }
    void create(const DeviceSystemSpec& spec, std::size_t capacity, std::span<const alea::Parameter> params) {
        reset();
        component_check(alea_create(&s_, static_cast<int>(capacity), static_cast<int>(spec.np),
                                    static_cast<int>(spec.dim), spec.lattice.empty() ? nullptr : spec.lattice.data(),
                                    spec.precision));
        capacity_ = nw_ = capacity;
        spec.registrar(s_, params);
    }
    // Positions [p][dim][nw] of nw <= capacity walkers, loaded and evaluated.
    // A population below the device count is loaded padded with copies of its
    // last walker, evaluated, and trimmed by a permutation (which wants an
    // evaluated system); a larger one is grown first (copies, overwritten by
    // the load).
    void load_evaluate(std::span<const double> pos, std::size_t nw) {
        if (nw == 0 || nw > capacity_)
            throw std::invalid_argument("alea_amd: walker count outside the device system's capacity");
        if (nw > nw_) resize(nw);
        if (nw == nw_) {
            component_check(alea_set_positions(s_, pos.data()));
            component_check(alea_evaluate(s_));
            return;
        }
        const std::size_t rows = pos.size() / nw;
        std::vector<double> padded(rows * nw_);
        for (std::size_t r = 0; r < rows; ++r)
            for (std::size_t j = 0; j < nw_; ++j) padded[r * nw_ + j] = pos[r * nw + std::min(j, nw - 1)];
        component_check(alea_set_positions(s_, padded.data()));
        component_check(alea_evaluate(s_));
        std::vector<int> src(nw);
        for (std::size_t j = 0; j < nw; ++j) src[j] = static_cast<int>(j);
        component_check(alea_permute(s_, static_cast<int>(nw), src.data()));
        nw_ = nw;
    }
    // The population to nw walkers (kept in order, the walkers past the old
    // count copies of the last one until they are loaded).
    void resize(std::size_t nw) {
        std::vector<int> src(nw);
        for (std::size_t j = 0; j < nw; ++j) src[j] = static_cast<int>(std::min(j, nw_ - 1));
        component_check(alea_permute(s_, static_cast<int>(nw), src.data()));
        nw_ = nw;
    }
    alea_system* get() const { return s_; }
    std::size_t nw() const { return nw_; }
    std::size_t capacity() const { return capacity_; }
    void set_nw(std::size_t nw) { nw_ = nw; }

private:
    alea_system* s_ = nullptr;
    std::size_t nw_ = 0, capacity_ = 0;
};

class DeviceComponent final : public alea::WavefunctionComponent {
public:
    // params: the device factors' parameters in the reference's order, with
    // `ids` the matching alea_param (factor, index and, for an image-summed
    // Padé, its breakup tables).  nodal / complex: whether the factors include
    // determinant blocks / twisted blocks.
    DeviceComponent(std::string name, DeviceSystemSpec spec, std::vector<alea::Parameter> params,
                    std::vector<alea_param> ids, bool nodal, bool complex)
        : name_(std::move(name)), spec_(std::move(spec)), params_(std::move(params)), ids_(std::move(ids)),
          nodal_(nodal), complex_(complex) {
        if (params_.size() != ids_.size()) throw std::invalid_argument("alea_amd: one alea_param per parameter");
        if (spec_.dim != 2) throw std::invalid_argument("alea_amd: dim must be 2");
    }

    std::string name() const override { return name_; }
    bool is_nodal() const override { return nodal_; }
    bool is_complex() const override { return complex_; }

    void prepare(std::size_t nw, std::size_t np, std::size_t dim) override {
        if (np != spec_.np || dim != spec_.dim) throw std::invalid_argument("alea_amd: batch shape differs from the device system's");
        resize_caches(nw, np, dim);
        const std::size_t plane = np * dim * nw;
        pgrad_c_.assign(complex_ ? plane : 0, 0.0);
        plap_c_.assign(complex_ ? plane : 0, 0.0);
        lr_.assign(nw, 0.0);
        ph_.assign(nw, 0.0);
        ng_.assign(dim * nw, 0.0);
        nl_.assign(dim * nw, 0.0);
        npg_.assign(dim * nw, 0.0);
        npl_.assign(dim * nw, 0.0);
        lp_.assign(nw, 0.0);
        g_.assign(plane, 0.0);
        l_.assign(plane, 0.0);
        pg_.assign(complex_ ? plane : 0, 0.0);
        pl_.assign(complex_ ? plane : 0, 0.0);
        mask_.assign(nw, 0);
        const std::size_t cap = std::max(nw, spec_.capacity);
        if (!dev_.get() || cap > dev_.capacity()) dev_.create(spec_, cap, params_);
    }

    void evaluate_batch(alea::WalkerSet& batch) override {
        const std::size_t nw = batch.num_walkers();
        prepare(nw, batch.n_particles(), batch.dim());
        dev_.load_evaluate(batch.positions(), nw);
        component_check(alea_get_derivatives_range(dev_.get(), 0, static_cast<int>(nw), log_psi_c_.data(),
                                                   grad_c_.data(), lap_c_.data(),
                                                   complex_ ? pgrad_c_.data() : nullptr,
                                                   complex_ ? plap_c_.data() : nullptr));
        add(batch.log_psi(), log_psi_c_);
        add(batch.gradients(), grad_c_);
        add(batch.laplacians(), lap_c_);
        if (complex_) {
            add(batch.phase_gradients(), pgrad_c_);
            add(batch.phase_laplacians(), plap_c_);
        }
    }

    void ratio_batch(alea::WalkerSet& batch, std::size_t p, std::size_t w_begin, std::size_t w_end) override {
        const std::size_t nw = batch.num_walkers(), dim = batch.dim();
        if (w_end > nw) w_end = nw;
        if (w_begin >= w_end) return;
        component_check(alea_ratio_range_phase(dev_.get(), static_cast<int>(p), static_cast<int>(w_begin),
                                               static_cast<int>(w_end), batch.new_pos().data(), lr_.data(),
                                               ph_.data(), ng_.data(), nl_.data(), npg_.data(), npl_.data()));
        auto lr = batch.log_ratios();
        auto ph = batch.phase_diffs();
        auto ng = batch.new_grad();
        auto nl = batch.new_lap();
        for (std::size_t w = w_begin; w < w_end; ++w) {
            lr[w] += lr_[w];
            ph[w] += ph_[w];
        }
        for (std::size_t d = 0; d < dim; ++d)
            for (std::size_t w = w_begin; w < w_end; ++w) {
                ng[d * nw + w] += ng_[d * nw + w];
                nl[d * nw + w] += nl_[d * nw + w];
            }
        if (complex_ && batch.has_phase()) {
            auto npg = batch.new_phase_grad();
            auto npl = batch.new_phase_lap();
            for (std::size_t d = 0; d < dim; ++d)
                for (std::size_t w = w_begin; w < w_end; ++w) {
                    npg[d * nw + w] += npg_[d * nw + w];
                    npl[d * nw + w] += npl_[d * nw + w];
                }
        }
    }

    void apply_move_batch(alea::WalkerSet& batch, std::size_t p, std::span<const bool> accept, std::size_t w_begin,
                          std::size_t w_end) override {
        const std::size_t nw = batch.num_walkers();
        if (w_end > nw) w_end = nw;
        if (w_begin >= w_end) return;
        for (std::size_t w = w_begin; w < w_end; ++w) mask_[w] = accept[w] ? 1 : 0;
        component_check(alea_accept_range(dev_.get(), static_cast<int>(p), static_cast<int>(w_begin),
                                          static_cast<int>(w_end), mask_.data()));
        component_check(alea_get_derivatives_range(dev_.get(), static_cast<int>(w_begin), static_cast<int>(w_end),
                                                   lp_.data(), g_.data(), l_.data(),
                                                   complex_ ? pg_.data() : nullptr, complex_ ? pl_.data() : nullptr));
        push(batch.log_psi(), log_psi_c_, lp_, 1, nw, w_begin, w_end);
        const std::size_t rows = batch.n_particles() * batch.dim();
        push(batch.gradients(), grad_c_, g_, rows, nw, w_begin, w_end);
        push(batch.laplacians(), lap_c_, l_, rows, nw, w_begin, w_end);
        if (complex_) {
            push(batch.phase_gradients(), pgrad_c_, pg_, rows, nw, w_begin, w_end);
            push(batch.phase_laplacians(), plap_c_, pl_, rows, nw, w_begin, w_end);
        }
    }

    // The deferred commit (wavefunction_component.h:236-252): the range's
    // accepted walkers committed on the device and log psi pushed; this
    // component's share of the moved particle's rows becomes its share of
    // the proposal's gradient and Laplacian (the walker set overwrites the
    // totals' row with the proposal's, WalkerSet::apply_move); partners stay
    // stale until refresh_particle_derivatives / refresh_derivatives read
    // the planes back.  A Metropolis sweep then reads no planes per move.
    void apply_move_batch_deferred(alea::WalkerSet& batch, std::size_t p, std::span<const bool> accept,
                                   std::size_t w_begin, std::size_t w_end) override {
        const std::size_t nw = batch.num_walkers(), dim = batch.dim();
        if (w_end > nw) w_end = nw;
        if (w_begin >= w_end) return;
        for (std::size_t w = w_begin; w < w_end; ++w) mask_[w] = accept[w] ? 1 : 0;
        component_check(alea_accept_range(dev_.get(), static_cast<int>(p), static_cast<int>(w_begin),
                                          static_cast<int>(w_end), mask_.data()));
        component_check(alea_get_derivatives_range(dev_.get(), static_cast<int>(w_begin), static_cast<int>(w_end),
                                                   lp_.data(), nullptr, nullptr, nullptr, nullptr));
        push(batch.log_psi(), log_psi_c_, lp_, 1, nw, w_begin, w_end);
        for (std::size_t d = 0; d < dim; ++d)
            for (std::size_t w = w_begin; w < w_end; ++w) {
                if (!accept[w]) continue;
                const std::size_t i = (p * dim + d) * nw + w, j = d * nw + w;
                grad_c_[i] = ng_[j];
                lap_c_[i] = nl_[j];
                if (complex_) {
                    pgrad_c_[i] = npg_[j];
                    plap_c_[i] = npl_[j];
                }
            }
    }
    // particle p's rows only (alea_get_particle_derivatives_range)
    void refresh_particle_derivatives(alea::WalkerSet& batch, std::size_t p, std::size_t w_begin,
                                      std::size_t w_end) override {
        const std::size_t nw = batch.num_walkers(), dim = batch.dim();
        if (w_end > nw) w_end = nw;
        if (w_begin >= w_end) return;
        component_check(alea_get_particle_derivatives_range(dev_.get(), static_cast<int>(p), static_cast<int>(w_begin),
                                                            static_cast<int>(w_end), g_.data(), l_.data(),
                                                            complex_ ? pg_.data() : nullptr,
                                                            complex_ ? pl_.data() : nullptr));
        auto row = [&](std::span<double> total, std::vector<double>& cache, const std::vector<double>& fresh) {
            for (std::size_t d = 0; d < dim; ++d)
                for (std::size_t w = w_begin; w < w_end; ++w) {
                    const std::size_t i = (p * dim + d) * nw + w, j = d * nw + w;
                    total[i] += fresh[j] - cache[i];
                    cache[i] = fresh[j];
                }
        };
        // the scratch planes, not the ratio's outputs, which a commit still reads
        row(batch.gradients(), grad_c_, g_);
        row(batch.laplacians(), lap_c_, l_);
        if (complex_) {
            row(batch.phase_gradients(), pgrad_c_, pg_);
            row(batch.phase_laplacians(), plap_c_, pl_);
        }
    }
    void refresh_derivatives(alea::WalkerSet& batch, std::size_t w_begin, std::size_t w_end) override {
        refresh_rows(batch, 0, batch.n_particles(), w_begin, w_end);
    }

    void permute_caches(std::span<const std::size_t> src) override {
        const std::size_t new_nw = src.size();
        if (new_nw > dev_.capacity())
            throw std::invalid_argument("alea_amd: the new population exceeds the device system's capacity");
        std::vector<int> s(src.begin(), src.end());
        component_check(alea_permute(dev_.get(), static_cast<int>(new_nw), s.data()));
        const std::size_t old_nw = log_psi_c_.size();
        dev_.set_nw(new_nw);
        alea::WavefunctionComponent::permute_caches(src);
        auto gather = [&](std::vector<double>& a) {
            if (a.empty() || old_nw == 0) return;
            const std::size_t width = a.size() / old_nw;
            std::vector<double> out(width * new_nw);
            for (std::size_t k = 0; k < width; ++k)
                for (std::size_t j = 0; j < new_nw; ++j) out[k * new_nw + j] = a[k * old_nw + src[j]];
            a = std::move(out);
        };
        gather(pgrad_c_);
        gather(plap_c_);
        const std::size_t plane = spec_.np * spec_.dim * new_nw;
        lr_.assign(new_nw, 0.0);
        ph_.assign(new_nw, 0.0);
        ng_.assign(spec_.dim * new_nw, 0.0);
        nl_.assign(spec_.dim * new_nw, 0.0);
        npg_.assign(spec_.dim * new_nw, 0.0);
        npl_.assign(spec_.dim * new_nw, 0.0);
        lp_.assign(new_nw, 0.0);
        g_.assign(plane, 0.0);
        l_.assign(plane, 0.0);
        pg_.assign(complex_ ? plane : 0, 0.0);
        pl_.assign(complex_ ? plane : 0, 0.0);
        mask_.assign(new_nw, 0);
    }

    // A deep copy: the same factors on a device system of its own, which its
    // next evaluate_batch loads.
    std::unique_ptr<alea::WavefunctionComponent> clone() const override {
        return std::make_unique<DeviceComponent>(name_, spec_, params_, ids_, nodal_, complex_);
    }

    std::span<alea::Parameter> parameters() override { return params_; }
    std::span<const alea::Parameter> parameters() const override { return params_; }

    // The factors again from the parameters' current values (after an
    // optimizer step); the next evaluate_batch loads the walkers.
    void rebuild() override {
        const std::size_t cap = dev_.capacity();
        if (cap) dev_.create(spec_, cap, params_);
    }

    // d log|psi| / dp of the free parameters, [n_free][nw] in parameter order.
    void compute_parameter_derivatives(const alea::WalkerSet& batch, std::span<double> derivs) const override {
        std::vector<alea_param> free;
        for (std::size_t k = 0; k < params_.size(); ++k)
            if (!params_[k].is_fixed) free.push_back(ids_[k]);
        if (free.empty()) return;
        const std::size_t nw = batch.num_walkers();
        if (derivs.size() < free.size() * nw) throw std::invalid_argument("alea_amd: derivs must be [n_free][nw]");
        component_check(alea_parameter_derivatives(dev_.get(), free.data(), static_cast<int>(free.size()),
                                                   derivs.data()));
    }

    std::size_t walker_state_bytes() const noexcept override {
        return alea::WavefunctionComponent::walker_state_bytes() +
               (pgrad_c_.size() + plap_c_.size()) / std::max<std::size_t>(log_psi_c_.size(), 1) * sizeof(double);
    }

private:
    // The rows of particles [p0, p1) over the walkers [w0, w1) read back and
    // pushed into the totals.
    void refresh_rows(alea::WalkerSet& batch, std::size_t p0, std::size_t p1, std::size_t w0, std::size_t w1) {
        const std::size_t nw = batch.num_walkers(), dim = batch.dim();
        if (w1 > nw) w1 = nw;
        if (w0 >= w1) return;
        component_check(alea_get_derivatives_range(dev_.get(), static_cast<int>(w0), static_cast<int>(w1), nullptr,
                                                   g_.data(), l_.data(), complex_ ? pg_.data() : nullptr,
                                                   complex_ ? pl_.data() : nullptr));
        const std::size_t r0 = p0 * dim, r1 = p1 * dim;
        push(batch.gradients(), grad_c_, g_, r1, nw, w0, w1, r0);
        push(batch.laplacians(), lap_c_, l_, r1, nw, w0, w1, r0);
        if (complex_) {
            push(batch.phase_gradients(), pgrad_c_, pg_, r1, nw, w0, w1, r0);
            push(batch.phase_laplacians(), plap_c_, pl_, r1, nw, w0, w1, r0);
        }
    }
    static void add(std::span<double> total, const std::vector<double>& part) {
        for (std::size_t k = 0; k < total.size() && k < part.size(); ++k) total[k] += part[k];
    }
    // cache <- fresh and total += fresh - cache over the range's columns of
    // `rows` planes
    static void push(std::span<double> total, std::vector<double>& cache, const std::vector<double>& fresh,
                     std::size_t rows, std::size_t nw, std::size_t w0, std::size_t w1, std::size_t r0 = 0) {
        for (std::size_t r = r0; r < rows; ++r)
            for (std::size_t w = w0; w < w1; ++w) {
                const std::size_t i = r * nw + w;
                total[i] += fresh[i] - cache[i];
                cache[i] = fresh[i];
            }
    }

    std::string name_;
    DeviceSystemSpec spec_;
    std::vector<alea::Parameter> params_;
    std::vector<alea_param> ids_;
    bool nodal_ = false, complex_ = false;
    mutable DeviceSystem dev_;
    std::vector<double> pgrad_c_, plap_c_;                        // the phase planes of a complex trial
    std::vector<double> lr_, ph_, ng_, nl_, npg_, npl_;            // a ratio's outputs
    std::vector<double> lp_, g_, l_, pg_, pl_;                     // a range's fresh planes
    std::vector<unsigned char> mask_;
};

// Device potential terms as one HamiltonianComponent: the registrar adds the
// groups and the terms (Ewald, Coulomb, inverse powers) and no wavefunction
// factor, so the device's local energy of a configuration is V(R).
// evaluate_potential is const and the reference may call it from several
// crowds at once: the device calls are serialized.
class DevicePotential final : public alea::HamiltonianComponent {
public:
    DevicePotential(std::string name, DeviceSystemSpec spec) : name_(std::move(name)), spec_(std::move(spec)) {}
    std::string name() const override { return name_; }

    void evaluate_potential(const alea::ParticleBatch& configs, std::span<double> v_out, std::size_t w_begin,
                            std::size_t w_end) const override {
        const std::size_t nw = configs.num_walkers();
        if (w_end > nw) w_end = nw;
        if (w_begin >= w_end) return;
        std::lock_guard<std::mutex> lock(*mutex_);
        const std::size_t cap = std::max(nw, spec_.capacity);
        if (!dev_.get() || cap > dev_.capacity()) dev_.create(spec_, cap, {});
        dev_.load_evaluate(configs.positions(), nw);
        v_.assign(nw, 0.0);
        component_check(alea_local_energy(dev_.get(), v_.data(), nullptr));
        for (std::size_t w = w_begin; w < w_end; ++w) v_out[w] += v_[w];
    }

    // V(p at proposed) - V(R) of configuration k (hamiltonian_component.h:
    // 216-246) through alea_delta_potential on the walker range [k, k+1):
    // nothing is committed, configs is not touched.
    double delta_potential_particle(alea::ParticleBatch& configs, std::size_t p, std::size_t k,
                                    std::span<const double> proposed, std::span<double> scratch) const override {
        (void)scratch;
        const std::size_t nw = configs.num_walkers();
        if (p >= configs.n_particles() || k >= nw || proposed.size() != configs.dim())
            throw std::invalid_argument("potential delta: invalid particle, slice or workspace");
        std::lock_guard<std::mutex> lock(*mutex_);
        const std::size_t cap = std::max(nw, spec_.capacity);
        if (!dev_.get() || cap > dev_.capacity()) dev_.create(spec_, cap, {});
        dev_.load_evaluate(configs.positions(), nw);
        newp_.assign(2 * nw, 0.0);
        newp_[k] = proposed[0];
        newp_[nw + k] = proposed[1];
        v_.assign(nw, 0.0);
        component_check(alea_delta_potential(dev_.get(), static_cast<int>(p), static_cast<int>(k),
                                             static_cast<int>(k + 1), newp_.data(), v_.data()));
        return v_[k];
    }

    std::unique_ptr<alea::HamiltonianComponent> clone() const override {
        return std::make_unique<DevicePotential>(name_, spec_);
    }

private:
    std::string name_;
    DeviceSystemSpec spec_;
    std::shared_ptr<std::mutex> mutex_ = std::make_shared<std::mutex>();
    mutable DeviceSystem dev_;
    mutable std::vector<double> v_, newp_;
};

}  // namespace alea_amd

#endif
