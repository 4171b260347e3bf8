// Host side of the C-ABI (include/alea_amd.h): builds the device model from
// the registration calls, owns the walker-major device state, converts
// between the reference's walker-fastest SoA and the device layout, and
// launches the kernels on one stream per system.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <limits>
#include <memory>
#include <mutex>
#include <numbers>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/alea_amd.h"
#include "model.h"

namespace alea_amd {
int set_ziggurat(const uint64_t* k, const double* w, const double* f, double r);
void launch_evaluate(const DevModel* m, const DevState& s, int nw, bool mixed, bool cplx, size_t lds,
                     hipStream_t st);
void launch_refresh(const DevModel* m, const DevState& s, int nw, int n_blocks, int nb_max, bool mixed,
                    bool cplx, int min_updates, hipStream_t st, bool mfma);
void launch_delta_potential(const DevModel* m, const DevState& s, int n, int p, const double* prop, int stride,
                            double* out, hipStream_t st);
void launch_param_derivs(const DevModel* m, const DevState& s, int nw, bool mixed, bool cplx, const DevParam* prm,
                         int n_prm, const double* wts, const double* spl, double* out, int stride, hipStream_t st);
void launch_param_eloc(const DevModel* m, const DevState& s, int nw, const DevParam* prm, int n_prm,
                       const double* wts, const double* spl, const double* grad, const double* pgrad,
                       double* out_re, double* out_im, int stride, hipStream_t st);
void launch_gauss_eloc(const DevModel* m, const DevState& s, bool mixed, int b, double alpha, const double* grad,
                       double* ybuf, int w0, int n_walkers, double* out, hipStream_t st);
void launch_sweep(const DevModel* m, const DevState& s, const SweepParams& p, int nw, bool mixed,
                  bool cplx, int rows_max, size_t lds, hipStream_t st);
void launch_draw(const DevState& s, int np, int nw, double* out, int stride, hipStream_t st);
int sweep_tile_rows(bool mixed, bool cplx, int rows_max);
void launch_sweep_fused(const FusedArgs& a, int total_nw, bool mixed, bool cplx, int rows_max, size_t lds,
                        hipStream_t st);
void launch_call(const DevModel* m, const DevState& s, const DevCall& c, int p, int mode, int nw,
                 bool mixed, bool cplx, size_t lds, hipStream_t st);
void launch_measure(const DevModel* m, const DevState& s, int nw, bool mixed, bool cplx, size_t lds,
                    hipStream_t st);
void launch_gather(const double* src_d, double* dst_d, int stride, const int* src, int n_new,
                   hipStream_t st);
void launch_dmc_weights(const double* e_new, double* e_old, double* weight, long long* n_cut, int nw,
                        double etrial, double eref, double e_cut, double e_max, double half_tau,
                        hipStream_t st);
void launch_branch(const double* weight, const double* xi, int* src, int src_cap, int nw,
                   long long* total, hipStream_t st);
void launch_fill(double* a, int n, double v, hipStream_t st);
void launch_fill_int(int* a, int n, int v, hipStream_t st);
void launch_estimators(const DevModel* m, const double* pos, const double* weights, int nw, int chunk,
                       bool gofr, double rmax, double inv_delta, int n_bins, int nfb, bool sofk,
                       const double* kvec, int nk, bool density, double d_rmax, double d_inv_delta, int d_bins,
                       bool d_per_particle, int d_off, double* part, int part_stride, int n_entries, double* acc,
                       hipStream_t st);
}  // namespace alea_amd

using namespace alea_amd;

namespace {

thread_local std::string g_error;

struct InvalidArg : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};
struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return ALEA_OK;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return ALEA_ERR_INVALID;
    } catch (const DeviceError& e) {
        g_error = e.what();
        return ALEA_ERR_DEVICE;
    } catch (const std::exception& e) {
        g_error = e.what();
        return ALEA_ERR_RUNTIME;
    }
}

// Ziggurat tables of XoshiroRandom (xoshiro_random.h:119-163): 256 equal-area
// regions, r from bisection on the top-layer mismatch.
struct Zig {
    std::array<uint64_t, 256> k{};
    std::array<double, 256> w{}, f{};
    double r = 0;
};
Zig make_zig() {
    constexpr int n = 256;
    auto f = [](double x) { return std::exp(-0.5 * x * x); };
    auto v_r = [&](double r) {
        return r * f(r) + std::sqrt(0.5 * std::numbers::pi) * std::erfc(r / std::numbers::sqrt2);
    };
    std::array<double, n> x{};
    auto top = [&](double r) {
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
        (top(mid) > 0.0 ? lo : hi) = mid;
    }
    Zig z;
    z.r = hi;
    (void)top(z.r);
    const double two52 = 0x1.0p52, base = v_r(z.r) / f(z.r);
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

constexpr uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

// A host <-> device copy ordered on a system's stream, complete on return.
// (A plain hipMemcpy runs on the legacy default stream, which does not wait
// for the systems' non-blocking streams.)
hipError_t copy_on(hipStream_t st, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

// Zero-filled device buffer from the stream-ordered allocator on the
// system's stream: allocating and freeing never synchronizes the device, so
// one system's set-up or teardown does not stall another's sweeps.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipStream_t stream = nullptr;
    void alloc(size_t count, hipStream_t st) {
        free();
        n = count;
        stream = st;
        if (count) {
            hip_check(hipMallocAsync(reinterpret_cast<void**>(&p), count * sizeof(T), st), "hipMallocAsync");
            hip_check(hipMemsetAsync(p, 0, count * sizeof(T), st), "hipMemsetAsync");
        }
    }
    void free() {
        if (p) (void)hipFreeAsync(p, stream);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { free(); }
};

}  // namespace

// The mixed form's cancellation factor.  Its inverse is stored in single
// precision, so a rebuild, which rounds the new inverse to single precision
// again, cannot lower the rounding floor eps32 * sum|A_mj v_m| of a ratio;
// the guard fires only where that floor reaches |R| (no correct digit).
constexpr double kMixedCancelFactor = 1.0;

struct alea_system {
    int capacity = 0, nw = 0, np = 0, precision = 0;
    double lattice[4] = {0, 0, 0, 0};   // primitive vectors as given (rows), periodic cells
    DevModel host;        // host copy of the model
    DevModel* dmodel = nullptr;
    bool built = false, evaluated = false;
    bool cplx = false;
    hipStream_t stream = nullptr;
    DevState st;
    DevBuf<double> pos, rho, logs, cond, work, e_loc, e_loc_im, rr;
    DevBuf<double> weights, e_prev, xi;   // DMC weights, E_L before the sweep, branch uniforms
    DevBuf<char> inv_bytes, inv_spare;    // inverse records; the gather target of a permutation
    DevBuf<double> perm_tmp;              // gather target of the other per-walker arrays
    DevBuf<int> updates, status, src, rebuilds;
    DevBuf<double> ew_rho, ew_snap;       // the Ewald structure factor kept per walker (ewald_rho)
    DevBuf<int> ew_inc;
    DevBuf<uint64_t> rng;
    DevBuf<long long> nacc, ncut, branch_total;
    bool energies_valid = false;          // e_loc holds E_L of the current configurations
    bool seeded = false;                  // device streams seeded (alea_seed)
    uint64_t seed_base = 0, seed_stream = 0, next_stream_index = 0;
    DevBuf<double> call_buf;       // new_pos, outputs
    DevBuf<unsigned char> call_mask;
    DevBuf<double> grad_buf;       // 4 * np * 2 * nw
    DevBuf<double> tape;
    DevBuf<double> dtape;          // one sweep of device-stream numbers, [np][3][nw]
    hipEvent_t ev0 = nullptr, ev1 = nullptr;   // brackets the last sweep launch (or graph block)
    // one pair per sweep of the last eager call: the sweep kernel alone
    std::vector<hipEvent_t> sweep_ev;
    int sweeps_timed = 0;
    bool timed = false, graph_timed = false;
    size_t walker_lds = 0;
    int rows_max = 0;             // largest padded block row count (0: no determinant)
    // After every sweep the batched refresh rebuilds each block whose update
    // count reached refresh_updates (the reference rebuilds at 64,
    // slater_determinant.cpp:788-790); refresh_every > 0 also rebuilds every
    // block every refresh_every sweeps.
    int refresh_updates = 64;
    int refresh_every = 0;
    // blocks of 33 to 64 orbitals refresh on the matrix cores (k_refresh_mfma)
    // unless this is cleared (alea_set_refresh_kernel)
    bool refresh_mfma = true;
    // the wavefunction factors in the order they were added (the reference's
    // component order): {0, pair factor}, {1, Nosanow group}, {2, block}
    std::vector<std::pair<int, int>> factors;
    long long since_refresh = 0;  // sweeps (or np accept calls) since the last full refresh

    // One captured block of device-stream sweeps (alea_set_sweep_graphs):
    // the key holds everything its launches bake in, compared bytewise; a
    // block whose key differs re-captures.
    struct SweepGraphKey {
        int n_sweeps, nw, np, dmc, refresh_every, refresh_updates, refresh_mfma, rows_max, precision, cplx;
        long long since_refresh;
        double tau;
        const double* dtape;
        const DevModel* dmodel;
        size_t walker_lds;
        DevState st;
    };
    bool sweep_graphs = false;
    hipGraphExec_t graph_exec = nullptr;
    SweepGraphKey graph_key{};
    long long graph_since_after = 0;   // since_refresh after the captured block
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        graph_exec = nullptr;
    }

    // Fused launches (alea_vmc_sweeps_fused): recorded on a member's stream
    // for the lead stream to wait on, or on the lead stream after the block
    // for the members to wait on.
    hipEvent_t fused_join = nullptr;

    // Estimators (pair correlation, structure factor): the block's running
    // sums on the device, [gofr: n_func*n_bins][sofk: n_pairs*nk][rho re:
    // G*nk][rho im: G*nk][weight sum]
    struct Estimators {
        bool gofr = false, sofk = false, density = false;
        double rmax = 0.0, inv_delta = 0.0;
        int n_bins = 0, nfb = 0, nk = 0;
        double d_rmax = 0.0, d_inv_delta = 0.0;   // density: window half-width (open cells), bins per length
        int d_bins = 0, d_rows = 0;   // density: grids (groups, or particles when per particle)
        bool d_per_particle = false;
        std::vector<double> kx, ky;
        std::vector<int> shell;
        DevBuf<double> acc, part, kvec;
        // [gofr][sofk, rho re, rho im][density]
        int sofk_entries(int ng) const { return sofk ? (ng * (ng + 1) / 2 + 2 * ng) * nk : 0; }
        int density_offset(int ng) const { return nfb + sofk_entries(ng); }
        int density_entries(int) const { return density ? d_rows * d_bins * d_bins : 0; }
        int entries(int ng) const { return density_offset(ng) + density_entries(ng); }
    } est;

    bool mixed() const { return precision == ALEA_MIXED; }
    size_t tsize() const { return mixed() ? 4 : 8; }

    void build() {
        if (built) return;
        DevModel& m = host;
        m.nw = capacity;
        m.np = np;
        // block membership
        for (int i = 0; i < np; ++i) m.block_of[i] = -1;
        int nb_max = 0, inv_elems = 0;
        rows_max = 0;
        for (int b = 0; b < m.n_blocks; ++b) {
            for (int i = m.blk_first[b]; i < m.blk_first[b] + m.blk_n[b]; ++i) m.block_of[i] = b;
            nb_max = std::max(nb_max, m.blk_n[b]);
            m.blk_rows[b] = (m.blk_n[b] + kRowPad - 1) / kRowPad * kRowPad;
            rows_max = std::max(rows_max, m.blk_rows[b]);
            inv_elems += (cplx ? 2 : 1) * m.blk_rows[b] * m.blk_n[b];
        }
        // pair lists
        for (int g = 0; g < kMaxGroups; ++g)
            for (int h = 0; h < kMaxGroups; ++h) m.pair_list_n[g][h] = 0;
        for (int f = 0; f < m.n_pair; ++f) {
            const int a = m.pair[f].ga, b = m.pair[f].gb;
            m.pair_list[a][b][m.pair_list_n[a][b]++] = f;
            if (a != b) m.pair_list[b][a][m.pair_list_n[b][a]++] = f;
        }
        // The sweep evaluates every image factor's spline in monomial form
        // from an LDS copy.  Panel i joins knots i and i+1, whose (f, f', f'')
        // triples a and b (LongRangeSpline::scaled order) weight the quintic
        // Hermite basis of spline_eval; expanded in powers of the panel
        // coordinate x that basis gives the coefficients below.
        m.spl_stride = 0;
        for (int f = 0; f < m.n_pair; ++f) {
            DevRadial& R = m.pair[f].rad;
            if (!R.image) continue;
            const int n_panels = R.n_knots - 1;   // coefficient k of panel i at k * n_panels + i
            for (int i = 0; i < n_panels; ++i) {
                const double* a = R.spline + 3 * i;
                const double* b = a + 3;
                double c[6];
                c[0] = a[0];
                c[1] = a[1];
                c[2] = 0.5 * a[2];
                c[3] = -10.0 * a[0] - 6.0 * a[1] - 1.5 * a[2] + 10.0 * b[0] - 4.0 * b[1] + 0.5 * b[2];
                c[4] = 15.0 * a[0] + 8.0 * a[1] + 1.5 * a[2] - 15.0 * b[0] + 7.0 * b[1] - b[2];
                c[5] = -6.0 * a[0] - 3.0 * a[1] - 0.5 * a[2] + 6.0 * b[0] - 3.0 * b[1] + 0.5 * b[2];
                for (int k = 0; k < 6; ++k) R.mono[k * n_panels + i] = c[k];
            }
            m.spl_stride = std::max(m.spl_stride, 6 * (R.n_knots - 1));
        }
        // the groups partition the particles (ParticleGroup ranges,
        // particle_group.h:21-27): the kernels walk pair partners group by group
        {
            std::vector<int> owner(np, -1);
            for (int g = 0; g < m.ng; ++g)
                for (int i = m.g_first[g]; i < m.g_last[g]; ++i) {
                    if (owner[i] >= 0) throw InvalidArg("particle groups overlap");
                    owner[i] = g;
                }
            for (int i = 0; i < np; ++i)
                if (owner[i] < 0) throw InvalidArg("particle " + std::to_string(i) + " is in no group");
        }
        for (int g = 0; g < m.ng; ++g)
            for (int i = m.g_first[g]; i < m.g_last[g]; ++i) m.group_of[i] = g;
        hip_check(hipMallocAsync(reinterpret_cast<void**>(&dmodel), sizeof(DevModel), stream), "hipMallocAsync model");
        hip_check(hipMemcpyAsync(dmodel, &m, sizeof(DevModel), hipMemcpyHostToDevice, stream), "copy model");
        model_nw_on_device = m.nw;
        const size_t nw_ = static_cast<size_t>(capacity);
        pos.alloc(nw_ * np * 2, stream);
        rho.alloc(std::max<size_t>(1, nw_ * m.ng * m.nkj * 2), stream);
        logs.alloc(nw_ * 4, stream);
        inv_elems += inv_elems & 1;   // whole doubles per record (branching gathers doubles)
        st.inv_stride = inv_elems;
        st.nb_max = std::max(nb_max, 1);
        inv_bytes.alloc(std::max<size_t>(1, nw_ * inv_elems * tsize()), stream);
        cond.alloc(nw_ * kMaxBlocks, stream);
        updates.alloc(nw_ * kMaxBlocks, stream);
        st.work_stride = 2 * (cplx ? 2 : 1) * nb_max * nb_max;
        work.alloc(std::max<size_t>(1, nw_ * st.work_stride), stream);
        rng.alloc(nw_ * 4, stream);
        e_loc.alloc(nw_, stream);
        e_loc_im.alloc(nw_, stream);
        rr.alloc(nw_ * 2, stream);
        weights.alloc(nw_, stream);
        launch_fill(weights.p, static_cast<int>(nw_), 1.0, stream);
        e_prev.alloc(nw_, stream);
        xi.alloc(nw_, stream);
        ncut.alloc(nw_, stream);
        branch_total.alloc(1, stream);
        status.alloc(nw_, stream);
        rebuilds.alloc(nw_ * kRebuildCauses, stream);
        // the per-walker Ewald structure factors, none computed yet
        ew_rho.alloc(m.ewald ? nw_ * m.nke * 2 : 1, stream);
        ew_snap.alloc(m.ewald ? nw_ * np * 2 : 1, stream);
        ew_inc.alloc(nw_, stream);
        launch_fill_int(ew_inc.p, static_cast<int>(nw_), -1, stream);
        nacc.alloc(nw_, stream);
        src.alloc(nw_, stream);
        // the per-call path's buffers (new_pos, outputs; mask), at the capacity
        call_buf.alloc(12 * nw_, stream);
        call_mask.alloc(nw_, stream);
        st.pos = pos.p;
        st.rho = rho.p;
        st.logs = logs.p;
        st.inv = inv_bytes.p;
        st.cond = cond.p;
        st.updates = updates.p;
        st.work = work.p;
        st.rng = rng.p;
        st.e_loc = e_loc.p;
        st.e_loc_im = e_loc_im.p;
        st.rr = rr.p;
        st.status = status.p;
        st.rebuilds = rebuilds.p;
        st.n_accepted = nacc.p;
        st.ew_rho = ew_rho.p;
        st.ew_snap = ew_snap.p;
        st.ew_inc = ew_inc.p;
        // LDS of one walker: positions, k-space densities, v', column j and
        // the Sherman-Morrison operands (these three over the padded rows),
        // then copies of the pair splines and the k-space Jastrow vectors and
        // weights (carve() in kernels.hip)
        walker_lds = sizeof(double) * (2 * static_cast<size_t>(np) + 2 * m.ng * m.nkj +
                                       8 * static_cast<size_t>(std::max(rows_max, kRowPad)) +
                                       static_cast<size_t>(m.n_pair) * m.spl_stride +
                                       (2 + static_cast<size_t>(m.ng) * m.ng) * m.nkj);
        // the batched refresh holds a block in registers (k_refresh_reg)
        if (m.n_blocks > 0 && nb_max > kMaxOrb) throw InvalidArg("determinant block larger than 128 orbitals");
        static bool zig_done = false;
        if (!zig_done) {
            const Zig z = make_zig();
            if (set_ziggurat(z.k.data(), z.w.data(), z.f.data(), z.r)) throw DeviceError("ziggurat upload");
            zig_done = true;
        }
        hip_check(hipStreamSynchronize(stream), "build");
        built = true;
    }

    void sync() { hip_check(hipStreamSynchronize(stream), "kernel"); }

    // The state of walkers [w0, capacity) as a batch of its own: every
    // per-walker array is a walker-major record, so a walker range is a
    // pointer offset (the SoA outputs of the measure kernel are not shifted).
    DevState shifted(int w0) const {
        DevState v = st;
        const size_t w = static_cast<size_t>(w0);
        const DevModel& m = host;
        v.pos += w * m.np * 2;
        v.rho += w * m.ng * m.nkj * 2;
        v.logs += w * 4;
        v.inv = static_cast<char*>(v.inv) + w * v.inv_stride * tsize();
        v.cond += w * kMaxBlocks;
        v.updates += w * kMaxBlocks;
        v.work += w * v.work_stride;
        v.rng += w * 4;
        v.e_loc += w;
        v.e_loc_im += w;
        v.rr += w * 2;
        v.status += w;
        v.n_accepted += w;
        v.rebuilds += w * kRebuildCauses;
        if (m.ewald) {
            v.ew_rho += w * m.nke * 2;
            v.ew_snap += w * m.np * 2;
        }
        v.ew_inc += w;
        v.grad_out = v.lap_out = v.pgrad_out = v.plap_out = nullptr;
        return v;
    }

    // The walker range [w_begin, w_end) of a range call; w_end < 0 means nw
    // (the reference's npos).
    std::pair<int, int> range(int w_begin, int w_end) const {
        if (w_end < 0) w_end = nw;
        if (w_begin < 0 || w_begin > w_end || w_end > nw)
            throw InvalidArg("walker range [" + std::to_string(w_begin) + ", " + std::to_string(w_end) +
                             ") outside the batch of " + std::to_string(nw));
        return {w_begin, w_end};
    }

    // Every block of every walker.
    void refresh(hipStream_t on = nullptr) {
        since_refresh = 0;
        if (host.n_blocks == 0) return;
        launch_refresh(dmodel, st, nw, host.n_blocks, st.nb_max, mixed(), cplx, 0, on ? on : stream, refresh_mfma);
        hip_check(hipGetLastError(), "launch refresh");
    }
    // After a sweep: the full refresh when refresh_every is due, else the
    // blocks whose update count reached refresh_updates.
    void refresh_after_sweep(hipStream_t on = nullptr) {
        if (refresh_every > 0 && ++since_refresh >= refresh_every) {
            refresh(on);
            return;
        }
        if (host.n_blocks == 0) return;
        launch_refresh(dmodel, st, nw, host.n_blocks, st.nb_max, mixed(), cplx, refresh_updates, on ? on : stream,
                       refresh_mfma);
        hip_check(hipGetLastError(), "launch refresh");
    }

    void check_status(int w0 = 0, int w1 = -1) {
        if (w1 < 0) w1 = nw;
        if (w1 <= w0) return;
        std::vector<int> s(static_cast<size_t>(w1 - w0));
        hip_check(copy_on(stream, s.data(), status.p + w0, s.size() * sizeof(int), hipMemcpyDeviceToHost), "status");
        for (int v : s)
            if (v) throw std::runtime_error("Slater determinant: accepted matrix is singular or nonfinite");
    }

    // The model's nw is the walker stride of the per-call SoA buffers.
    // ordered on the system's stream, and only when the count changed
    int model_nw_on_device = -1;
    void set_model_nw(int n) {
        host.nw = n;
        if (n == model_nw_on_device) return;
        model_nw_on_device = n;
        hip_check(hipMemcpyAsync(&dmodel->nw, &host.nw, sizeof(int), hipMemcpyHostToDevice, stream), "model nw");
    }
};

namespace {

alea_system* Sys(alea_system* s) {
    if (!s) throw InvalidArg("null system handle");
    return s;
}

void need_unbuilt(alea_system* s) {
    if (s->built) throw InvalidArg("components must be registered before positions are set");
}

void load_spline(DevRadial& R, int n_knots, const double* spline, double r_cut, double inv_delta) {
    if (n_knots < 2 || n_knots > kMaxKnots) throw InvalidArg("spline knot count out of range");
    std::copy(spline, spline + 3 * n_knots, R.spline);
    R.n_knots = n_knots;
    R.r_cut = r_cut;
    R.inv_delta = inv_delta;
}

void set_cell(DevCell& c, const double* v) {
    // SimulationCell::finalize_periodic (simulation_cell.cpp:146-244), 2-D.
    double a[4] = {v[0], v[1], v[2], v[3]};
    std::size_t axis[2];
    bool used[2] = {false, false}, ok = true;
    for (int i = 0; i < 2 && ok; ++i) {
        int nonzero = 0, which = 0;
        for (int e = 0; e < 2; ++e)
            if (a[i * 2 + e] != 0.0) { ++nonzero; which = e; }
        if (nonzero != 1 || used[which]) ok = false;
        else { axis[i] = which; used[which] = true; }
    }
    if (ok) {
        double re[4] = {};
        for (int i = 0; i < 2; ++i)
            for (int e = 0; e < 2; ++e) re[axis[i] * 2 + e] = a[i * 2 + e];
        std::memcpy(a, re, sizeof re);
    } else {
        for (int guard = 0; guard < 64; ++guard) {
            const double n0 = a[0] * a[0] + a[1] * a[1], n1 = a[2] * a[2] + a[3] * a[3];
            if (n1 < n0) { std::swap(a[0], a[2]); std::swap(a[1], a[3]); continue; }
            if (n0 == 0.0) break;
            const double mu = std::nearbyint((a[0] * a[2] + a[1] * a[3]) / n0);
            if (mu == 0.0) break;
            a[2] -= mu * a[0];
            a[3] -= mu * a[1];
        }
    }
    c.periodic = 1;
    c.ortho = ok ? 1 : 0;
    const double det = a[0] * a[3] - a[1] * a[2];
    if (det == 0.0) throw InvalidArg("lattice vectors are singular");
    const double id = 1.0 / det;
    c.ainv[0] = a[3] * id;
    c.ainv[1] = -a[2] * id;
    c.ainv[2] = -a[1] * id;
    c.ainv[3] = a[0] * id;
    for (int k = 0; k < 4; ++k) {
        c.a[k] = a[k];
        c.b[k] = 6.283185307179586476925286766559 * c.ainv[k];
    }
    if (c.ortho)
        for (int e = 0; e < 2; ++e) {
            c.box[e] = std::abs(a[e * 2 + e]);
            c.inv_box[e] = 1.0 / c.box[e];
        }
    int nc = 1;
    c.cand[0] = c.cand[1] = 0.0;
    if (!c.ortho)
        for (int n0 = -1; n0 <= 1; ++n0)
            for (int n1 = -1; n1 <= 1; ++n1) {
                if (n0 == 0 && n1 == 0) continue;
                for (int e = 0; e < 2; ++e) {
                    double s = 0.0;
                    s += n0 * a[0 * 2 + e];
                    s += n1 * a[1 * 2 + e];
                    c.cand[nc * 2 + e] = s;
                }
                ++nc;
            }
    double shortest2 = 0.0;
    for (int n0 = -2; n0 <= 2; ++n0)
        for (int n1 = -2; n1 <= 2; ++n1) {
            if (n0 == 0 && n1 == 0) continue;
            double r2 = 0.0;
            for (int e = 0; e < 2; ++e) {
                const double s = n0 * a[e] + n1 * a[2 + e];
                r2 += s * s;
            }
            if (shortest2 == 0.0 || r2 < shortest2) shortest2 = r2;
        }
    c.ws = 0.5 * std::sqrt(shortest2);
}

void add_kspace(alea_system* s, int ga, int gb, int n_k, const double* kx, const double* ky,
                const double* weights) {
    DevModel& m = s->host;
    if (n_k > kMaxKJ) throw InvalidArg("too many k-space Jastrow vectors");
    if (m.nkj == 0) {
        m.nkj = n_k;
        std::copy(kx, kx + n_k, m.kjx);
        std::copy(ky, ky + n_k, m.kjy);
    } else {
        if (m.nkj != n_k) throw InvalidArg("image-summed Jastrow factors must share one k set");
        for (int k = 0; k < n_k; ++k)
            if (m.kjx[k] != kx[k] || m.kjy[k] != ky[k])
                throw InvalidArg("image-summed Jastrow factors must share one k set");
    }
    for (int k = 0; k < n_k; ++k) {
        m.kjw[ga][gb][k] += weights[k];
        if (ga != gb) m.kjw[gb][ga][k] += weights[k];
    }
    m.kj_active[ga] = m.kj_active[gb] = 1;
}

void check_group(alea_system* s, int g) {
    if (g < 0 || g >= s->host.ng) throw InvalidArg("group index out of range");
}

void soa_to_walker(const double* soa, std::vector<double>& dev, int nw, int np, int stride_nw) {
    dev.assign(static_cast<size_t>(nw) * np * 2, 0.0);
    for (int p = 0; p < np; ++p)
        for (int d = 0; d < 2; ++d)
            for (int w = 0; w < nw; ++w)
                dev[(static_cast<size_t>(w) * np + p) * 2 + d] = soa[(p * 2 + d) * stride_nw + w];
}

}  // namespace

extern "C" {

const char* alea_last_error(void) { return g_error.c_str(); }

int alea_create(alea_system** out, int n_walkers, int n_particles, int dim, const double* lattice,
                int precision) {
    return guarded([&] {
        if (!out) throw InvalidArg("null output handle");
        if (dim != 2) throw InvalidArg("only 2-D systems are supported");
        if (n_walkers <= 0 || n_particles <= 0) throw InvalidArg("empty walker batch");
        if (n_particles > kMaxParticles) throw InvalidArg("too many particles (max 256)");
        if (precision != ALEA_FP64 && precision != ALEA_MIXED) throw InvalidArg("unknown precision");
        auto s = std::make_unique<alea_system>();
        s->capacity = s->nw = n_walkers;
        s->np = n_particles;
        s->precision = precision;
        s->host.cancel_factor = precision == ALEA_MIXED ? kMixedCancelFactor : 1e4;
        if (lattice) {
            set_cell(s->host.cell, lattice);
            std::copy(lattice, lattice + 4, s->lattice);
        }
        // non-blocking: copies on the legacy default stream never wait for
        // this system's kernels, nor its kernels for other systems' copies
        hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
        *out = s.release();
    });
}

int alea_destroy(alea_system* s) {
    if (!s) return ALEA_OK;
    hipStream_t st = s->stream;
    if (st) (void)hipStreamSynchronize(st);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    for (hipEvent_t e : s->sweep_ev) (void)hipEventDestroy(e);
    s->drop_graph();
    if (s->fused_join) (void)hipEventDestroy(s->fused_join);
    if (s->dmodel) (void)hipFreeAsync(s->dmodel, st);
    delete s;   // the buffers free on the stream, which must outlive them
    if (st) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    return ALEA_OK;
}

int alea_query(alea_system* s, int* done) {
    return guarded([&] {
        Sys(s);
        if (!done) throw InvalidArg("null output");
        const hipError_t e = hipStreamQuery(s->stream);
        if (e == hipErrorNotReady) {
            *done = 0;
            return;
        }
        hip_check(e, "stream query");
        *done = 1;
    });
}

int alea_add_group(alea_system* s, int first, int last, double charge, double inv_mass) {
    return guarded([&] {
        Sys(s);
        need_unbuilt(s);
        DevModel& m = s->host;
        if (m.ng >= kMaxGroups) throw InvalidArg("too many groups (max 4)");
        if (first < 0 || last < first || last > s->np) throw InvalidArg("group range out of bounds");
        m.g_first[m.ng] = first;
        m.g_last[m.ng] = last;
        for (int i = first; i < last; ++i) {
            m.inv_mass[i] = inv_mass;
            m.charge[i] = charge;
            m.group_of[i] = m.ng;
        }
        ++m.ng;
    });
}

int alea_add_pade_jastrow(alea_system* s, int ga, int gb, int num_order, int den_order,
                          const double* coeffs, int n_k, const int* n1, const int* n2,
                          const double* kx, const double* ky, const double* weights, int n_knots,
                          const double* spline, double r_cut, double inv_delta, double u_inf) {
    return guarded([&] {
        Sys(s);
        need_unbuilt(s);
        check_group(s, ga);
        check_group(s, gb);
        DevModel& m = s->host;
        if (m.n_pair >= kMaxPair) throw InvalidArg("too many pair Jastrow factors");
        if (num_order < 0 || den_order < 0 || num_order + den_order > kMaxCoef)
            throw InvalidArg("Padé order out of range");
        DevPair& P = m.pair[m.n_pair];
        P = DevPair{};
        P.ga = ga;
        P.gb = gb;
        P.rad.kind = 0;
        P.rad.num = num_order;
        P.rad.den = den_order;
        std::copy(coeffs, coeffs + num_order + den_order, P.rad.c);
        if (n_k > 0) {
            if (!m.cell.periodic) throw InvalidArg("an image sum needs a periodic cell");
            (void)n1;
            (void)n2;
            P.rad.image = 1;
            P.rad.u_inf = u_inf;
            load_spline(P.rad, n_knots, spline, r_cut, inv_delta);
            add_kspace(s, ga, gb, n_k, kx, ky, weights);
        }
        s->factors.push_back({0, m.n_pair});
        ++m.n_pair;
    });
}

int alea_add_drummond_jastrow(alea_system* s, int ga, int gb, int C, double gamma, double L,
                              int n_coeffs, const double* coeffs) {
    return guarded([&] {
        Sys(s);
        need_unbuilt(s);
        check_group(s, ga);
        check_group(s, gb);
        DevModel& m = s->host;
        if (m.n_pair >= kMaxPair) throw InvalidArg("too many pair Jastrow factors");
        if (n_coeffs < 1 || n_coeffs > kMaxCoef || C < 2) throw InvalidArg("Drummond needs 1..12 coefficients and C >= 2");
        if (m.cell.periodic && L > m.cell.ws) throw InvalidArg("Drummond cutoff exceeds the Wigner-Seitz inradius");
        DevPair& P = m.pair[m.n_pair];
        P = DevPair{};
        P.ga = ga;
        P.gb = gb;
        P.rad.kind = 1;
        P.rad.num = n_coeffs;
        P.rad.C = C;
        P.rad.L = L;
        std::copy(coeffs, coeffs + n_coeffs, P.rad.c);
        double ipw = 1.0;
        for (int i = 0; i < C; ++i) ipw *= -L;
        P.rad.alpha1 = gamma / ipw + coeffs[0] * static_cast<double>(C) / L;
        P.rad.gamma = gamma;
        s->factors.push_back({0, m.n_pair});
        ++m.n_pair;
    });
}

int alea_add_nosanow_jastrow(alea_system* s, int group, double alpha, const double* sites) {
    return guarded([&] {
        Sys(s);
        need_unbuilt(s);
        check_group(s, group);
        DevModel& m = s->host;
        if (m.cell.periodic && std::exp(-alpha * m.cell.ws * m.cell.ws) > 1e-6)
            throw InvalidArg("nosanow: alpha too small for the cell (folded tail above 1e-6)");
        for (int i = m.g_first[group]; i < m.g_last[group]; ++i) {
            m.nos_alpha[i] = alpha;
            m.nos_site[i][0] = sites[2 * (i - m.g_first[group])];
            m.nos_site[i][1] = sites[2 * (i - m.g_first[group]) + 1];
        }
        m.has_nosanow = 1;
        s->factors.push_back({1, group});
    });
}

static void add_block(alea_system* s, int first, int last, int kind) {
    DevModel& m = s->host;
    if (m.n_blocks >= kMaxBlocks) throw InvalidArg("too many determinant blocks (max 2)");
    const int n = last - first;
    if (first < 0 || n <= 0 || last > s->np) throw InvalidArg("block range out of bounds");
    if (n > kMaxOrb) throw InvalidArg("block larger than 128 orbitals");
    const bool c = kind == kTwisted;
    if (m.n_blocks > 0 && (m.det_complex != 0) != c) throw InvalidArg("mixing real and complex determinants");
    m.det_complex = c ? 1 : 0;
    s->cplx = c;
    m.blk_first[m.n_blocks] = first;
    m.blk_n[m.n_blocks] = n;
    m.blk_kind[m.n_blocks] = kind;
}

int alea_add_slater_plane_wave(alea_system* s, int first, int last, const double* kvecs,
                               const unsigned char* is_sin) {
    return guarded([&] {
        Sys(s);
        need_unbuilt(s);
        if (!s->host.cell.periodic) throw InvalidArg("plane-wave orbitals need a periodic cell");
        add_block(s, first, last, kPlaneWave);
        DevModel& m = s->host;
        const int b = m.n_blocks;
        for (int i = 0; i < last - first; ++i) {
            m.orb_k[b][i][0] = kvecs[2 * i];
            m.orb_k[b][i][1] = kvecs[2 * i + 1];
            m.orb_sin[b][i] = is_sin[i] ? 1 : 0;
        }
        s->factors.push_back({2, m.n_blocks});
        ++m.n_blocks;
    });
}

int alea_add_slater_gaussian(alea_system* s, int first, int last, const double* centers,
                             const double* alphas) {
    return guarded([&] {
        Sys(s);
        need_unbuilt(s);
        add_block(s, first, last, kGaussian);
        DevModel& m = s->host;
        const int b = m.n_blocks;
        for (int i = 0; i < last - first; ++i) {
            m.orb_k[b][i][0] = centers[2 * i];
            m.orb_k[b][i][1] = centers[2 * i + 1];
            m.orb_alpha[b][i] = alphas[i];
        }
        s->factors.push_back({2, m.n_blocks});
        ++m.n_blocks;
    });
}

int alea_add_slater_twisted(alea_system* s, int first, int last, const double* q) {
    return guarded([&] {
        Sys(s);
        need_unbuilt(s);
        if (!s->host.cell.periodic) throw InvalidArg("twisted plane waves need a periodic cell");
        add_block(s, first, last, kTwisted);
        DevModel& m = s->host;
        const int b = m.n_blocks;
        for (int i = 0; i < last - first; ++i) {
            m.orb_k[b][i][0] = q[2 * i];
            m.orb_k[b][i][1] = q[2 * i + 1];
        }
        s->factors.push_back({2, m.n_blocks});
        ++m.n_blocks;
    });
}

int alea_add_ewald(alea_system* s, int n_k, const int* n1, const int* n2, const double* kx,
                   const double* ky, const double* weights, int n_knots, const double* spline,
                   double r_cut, double inv_delta, double v_long_zero, double sr_k0) {
    return guarded([&] {
        Sys(s);
        need_unbuilt(s);
        DevModel& m = s->host;
        if (!m.cell.periodic) throw InvalidArg("Ewald needs a periodic cell");
        if (n_k > kMaxKE) throw InvalidArg("too many Ewald vectors");
        if (n_knots < 2 || n_knots > kMaxKnots) throw InvalidArg("spline knot count out of range");
        (void)n1;
        (void)n2;
        m.ewald = 1;
        m.nke = n_k;
        std::copy(kx, kx + n_k, m.kex);
        std::copy(ky, ky + n_k, m.key);
        std::copy(weights, weights + n_k, m.kew);
        std::copy(spline, spline + 3 * n_knots, m.ew_spline);
        m.ew_knots = n_knots;
        m.ew_rcut = r_cut;
        m.ew_inv_delta = inv_delta;
        // self and background constants (ewald_term.cpp:186-262, one kernel)
        double q2 = 0.0, Q = 0.0;
        for (int i = 0; i < s->np; ++i) {
            q2 += m.charge[i] * m.charge[i];
            Q += m.charge[i];
        }
        m.ew_const = -0.5 * v_long_zero * q2 - 0.5 * Q * Q * sr_k0;
        m.ew_inv_eps = 1.0;
    });
}

int alea_add_coulomb(alea_system* s, int ga, int gb, double epsilon) {
    return guarded([&] {
        Sys(s);
        need_unbuilt(s);
        check_group(s, ga);
        check_group(s, gb);
        DevModel& m = s->host;
        if (m.n_coul >= kMaxCoulomb) throw InvalidArg("too many Coulomb pairs");
        if (!(epsilon > 0.0)) throw InvalidArg("epsilon must be positive");
        const double qa = m.charge[m.g_first[ga]], qb = m.charge[m.g_first[gb]];
        m.coul_ga[m.n_coul] = ga;
        m.coul_gb[m.n_coul] = gb;
        m.coul_coeff[m.n_coul] = qa * qb / epsilon;
        ++m.n_coul;
    });
}

// McMillan pair Jastrow u(r) = -(b/r)^n at the minimum image (no
// counterpart in the reference).
int alea_add_mcmillan_jastrow(alea_system* s, int ga, int gb, double b, int n) {
    return guarded([&] {
        Sys(s);
        need_unbuilt(s);
        check_group(s, ga);
        check_group(s, gb);
        DevModel& m = s->host;
        if (m.n_pair >= kMaxPair) throw InvalidArg("too many pair Jastrow factors");
        if (!(b > 0.0) || !std::isfinite(b)) throw InvalidArg("mcmillan: b must be finite and positive");
        if (n < 1 || n > 16) throw InvalidArg("mcmillan: the exponent must be 1..16");
        s->factors.push_back({0, m.n_pair});
        DevPair& P = m.pair[m.n_pair++];
        P.ga = ga;
        P.gb = gb;
        P.rad = DevRadial{};
        P.rad.kind = 2;
        P.rad.num = n;
        P.rad.c[0] = b;
    });
}

namespace {
// The integral of r^-p over the plane outside the Wigner-Seitz cell of the
// lattice a (rows a1, a2), 1/(p - 2) times the integral over angle of
// rho(theta)^(2 - p), rho the distance from the centre to the cell boundary:
// the midpoint rule on 2^18 angles, the boundary from the bisectors of the
// 24 translations s a1 + t a2, |s|, |t| <= 2 (the method of
// outside_wigner_seitz_integral, inverse_power.cpp:28-41, and
// wigner_seitz_exterior.h:58-83).
double ws_exterior_integral(const double* a, double p) {
    double lx[24], ly[24];
    int nl = 0;
    for (int si = -2; si <= 2; ++si)
        for (int ti = -2; ti <= 2; ++ti) {
            if (si == 0 && ti == 0) continue;
            lx[nl] = si * a[0] + ti * a[2];
            ly[nl] = si * a[1] + ti * a[3];
            ++nl;
        }
    constexpr size_t n_theta = size_t{1} << 18;
    const double dtheta = 2.0 * 3.14159265358979323846 / static_cast<double>(n_theta);
    double sum = 0.0;
    for (size_t k = 0; k < n_theta; ++k) {
        const double theta = (static_cast<double>(k) + 0.5) * dtheta;
        const double ux = std::cos(theta), uy = std::sin(theta);
        double rho = std::numeric_limits<double>::infinity();
        for (int c = 0; c < nl; ++c) {
            const double proj = ux * lx[c] + uy * ly[c];
            if (proj > 0.0) rho = std::min(rho, 0.5 * (lx[c] * lx[c] + ly[c] * ly[c]) / proj);
        }
        sum += std::pow(rho, 2.0 - p);
    }
    return sum * dtheta / (p - 2.0);
}
}  // namespace

// InversePowerTerm(group_a, group_b, C, p, cell) (inverse_power.h:56-60,
// inverse_power.cpp:43-83): the same refusals, the same integer-power
// product, and in a periodic cell the continuum tail outside the
// Wigner-Seitz cell.
int alea_add_inverse_power(alea_system* s, int ga, int gb, double strength, double exponent) {
    return guarded([&] {
        Sys(s);
        need_unbuilt(s);
        check_group(s, ga);
        check_group(s, gb);
        DevModel& m = s->host;
        if (m.n_ipow >= kMaxInversePower) throw InvalidArg("too many inverse-power terms");
        if (!std::isfinite(exponent) || exponent <= 0.0)
            throw InvalidArg("inverse_power: exponent must be finite and positive");
        if (!std::isfinite(strength)) throw InvalidArg("inverse_power: strength must be finite");
        if (strength < 0.0 && exponent >= 2.0)
            throw InvalidArg("inverse_power: an attractive core (strength < 0) with exponent >= 2 falls to the centre");
        const int t = m.n_ipow;
        m.ipow_ga[t] = ga;
        m.ipow_gb[t] = gb;
        m.ipow_c[t] = strength;
        m.ipow_p[t] = exponent;
        const double rounded = std::round(exponent);
        m.ipow_int[t] = (rounded == exponent && rounded <= 16.0) ? static_cast<int>(rounded) : 0;
        if (m.cell.periodic) {
            if (exponent <= 2.0)
                throw InvalidArg("inverse_power: in a periodic 2-D cell the exponent must exceed 2");
            const double na = m.g_last[ga] - m.g_first[ga], nb = m.g_last[gb] - m.g_first[gb];
            const double pairs = ga == gb ? 0.5 * na * na : na * nb;
            const double area = std::fabs(s->lattice[0] * s->lattice[3] - s->lattice[1] * s->lattice[2]);
            const double outside = ws_exterior_integral(s->lattice, exponent);
            m.ipow_tail += pairs / area * strength * outside;
        }
        ++m.n_ipow;
    });
}

int alea_set_positions(alea_system* s, const double* pos) {
    return guarded([&] {
        Sys(s);
        s->build();
        std::vector<double> dev;
        soa_to_walker(pos, dev, s->nw, s->np, s->nw);
        // wrap into the cell (WalkerSet::load_all, walker_set.cpp:51-70)
        if (s->host.cell.periodic) {
            const DevCell& c = s->host.cell;
            for (size_t i = 0; i < dev.size(); i += 2) {
                double& x0 = dev[i];
                double& x1 = dev[i + 1];
                if (c.ortho) {
                    x0 -= c.box[0] * std::nearbyint(x0 * c.inv_box[0]);
                    x1 -= c.box[1] * std::nearbyint(x1 * c.inv_box[1]);
                } else {
                    double f0 = 0.0, f1 = 0.0;
                    f0 += c.ainv[0] * x0;
                    f0 += c.ainv[1] * x1;
                    f1 += c.ainv[2] * x0;
                    f1 += c.ainv[3] * x1;
                    f0 = std::nearbyint(f0);
                    f1 = std::nearbyint(f1);
                    double s0 = 0.0, s1 = 0.0;
                    s0 += f0 * c.a[0];
                    s0 += f1 * c.a[2];
                    s1 += f0 * c.a[1];
                    s1 += f1 * c.a[3];
                    x0 -= s0;
                    x1 -= s1;
                }
            }
        }
        hip_check(copy_on(s->stream, s->pos.p, dev.data(), dev.size() * sizeof(double), hipMemcpyHostToDevice),
                  "upload positions");
        s->evaluated = false;
        s->energies_valid = false;
    });
}

int alea_get_positions(alea_system* s, double* pos) {
    return guarded([&] {
        Sys(s);
        if (!s->built) throw InvalidArg("positions were never set");
        s->sync();
        std::vector<double> dev(static_cast<size_t>(s->nw) * s->np * 2);
        hip_check(copy_on(s->stream, dev.data(), s->pos.p, dev.size() * sizeof(double), hipMemcpyDeviceToHost),
                  "download positions");
        for (int p = 0; p < s->np; ++p)
            for (int d = 0; d < 2; ++d)
                for (int w = 0; w < s->nw; ++w)
                    pos[(p * 2 + d) * s->nw + w] = dev[(static_cast<size_t>(w) * s->np + p) * 2 + d];
    });
}

int alea_evaluate(alea_system* s) {
    return guarded([&] {
        Sys(s);
        if (!s->built) throw InvalidArg("positions were never set");
        s->set_model_nw(s->nw);
        launch_evaluate(s->dmodel, s->st, s->nw, s->mixed(), s->cplx, s->walker_lds, s->stream);
        hip_check(hipGetLastError(), "launch evaluate");
        s->refresh();
        s->sync();
        s->check_status();
        s->evaluated = true;
    });
}

int alea_get_log_psi(alea_system* s, double* out) {
    return guarded([&] {
        Sys(s);
        if (!s->evaluated) throw InvalidArg("evaluate first");
        s->sync();
        std::vector<double> l(static_cast<size_t>(s->nw) * 4);
        hip_check(copy_on(s->stream, l.data(), s->logs.p, l.size() * sizeof(double), hipMemcpyDeviceToHost), "logs");
        for (int w = 0; w < s->nw; ++w) {
            double t = l[w * 4];
            for (int b = 0; b < s->host.n_blocks; ++b) t += l[w * 4 + 1 + b];
            out[w] = t;
        }
    });
}

static void measure(alea_system* s, bool want_grads) {
    if (!s->evaluated) throw InvalidArg("evaluate first");
    DevState st = s->st;
    if (want_grads) {
        const size_t n = static_cast<size_t>(s->np) * 2 * s->nw;
        if (s->grad_buf.n < 4 * n) s->grad_buf.alloc(4 * n, s->stream);
        st.grad_out = s->grad_buf.p;
        st.lap_out = s->grad_buf.p + n;
        st.pgrad_out = s->grad_buf.p + 2 * n;
        st.plap_out = s->grad_buf.p + 3 * n;
    }
    s->set_model_nw(s->nw);
    launch_measure(s->dmodel, st, s->nw, s->mixed(), s->cplx, s->walker_lds, s->stream);
    hip_check(hipGetLastError(), "launch measure");
    s->sync();
    s->energies_valid = true;
}

int alea_get_gradients(alea_system* s, double* grad, double* lap) {
    return guarded([&] {
        Sys(s);
        measure(s, true);
        const size_t n = static_cast<size_t>(s->np) * 2 * s->nw;
        hip_check(copy_on(s->stream, grad, s->grad_buf.p, n * sizeof(double), hipMemcpyDeviceToHost), "grad");
        hip_check(copy_on(s->stream, lap, s->grad_buf.p + n, n * sizeof(double), hipMemcpyDeviceToHost), "lap");
    });
}

int alea_get_phase_gradients(alea_system* s, double* pgrad, double* plap) {
    return guarded([&] {
        Sys(s);
        measure(s, true);
        const size_t n = static_cast<size_t>(s->np) * 2 * s->nw;
        hip_check(copy_on(s->stream, pgrad, s->grad_buf.p + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost), "pgrad");
        hip_check(copy_on(s->stream, plap, s->grad_buf.p + 3 * n, n * sizeof(double), hipMemcpyDeviceToHost), "plap");
    });
}

int alea_local_energy(alea_system* s, double* e_re, double* e_im) {
    return guarded([&] {
        Sys(s);
        measure(s, false);
        hip_check(copy_on(s->stream, e_re, s->e_loc.p, s->nw * sizeof(double), hipMemcpyDeviceToHost), "e_loc");
        if (e_im)
            hip_check(copy_on(s->stream, e_im, s->e_loc_im.p, s->nw * sizeof(double), hipMemcpyDeviceToHost), "e_loc_im");
        // energies of a walker whose determinant went singular are not reported
        s->check_status();
    });
}

int alea_check_status(alea_system* s) {
    return guarded([&] {
        Sys(s);
        if (!s->built) return;
        s->sync();
        s->check_status();
    });
}

// The per-call path over the walkers [w0, w1): every array has the batch's
// walker stride nw, and only the range's entries are read or written (the
// reference's component contract, wavefunction_component.h:96-110).  The
// kernels see the range as a batch of its own (shifted state, offset call
// buffers); the call buffers are sized by the capacity at build, so range
// calls allocate nothing, and calls on disjoint ranges touch disjoint
// memory.
static void ratio_range(alea_system* s, int p, int w0, int w1, const double* new_pos, double* log_ratio,
                        double* phase, double* new_grad, double* new_lap, double* new_pgrad = nullptr,
                        double* new_plap = nullptr) {
    if (!s->evaluated) throw InvalidArg("evaluate first");
    if (p < 0 || p >= s->np) throw InvalidArg("particle index out of range");
    if (!new_pos || !log_ratio) throw InvalidArg("null proposal or output");
    const size_t nw = s->nw, n = static_cast<size_t>(w1 - w0);
    if (n == 0) return;
    double* b = s->call_buf.p;
    auto rows = [&](double* dev, const double* host, int n_rows, hipMemcpyKind kind, const char* what) {
        for (int d = 0; d < n_rows; ++d) {
            const size_t off = d * nw + w0;
            if (kind == hipMemcpyHostToDevice)
                hip_check(copy_on(s->stream, dev + off, host + off, n * sizeof(double), kind), what);
            else
                hip_check(copy_on(s->stream, const_cast<double*>(host) + off, dev + off, n * sizeof(double), kind), what);
        }
    };
    rows(b, new_pos, 2, hipMemcpyHostToDevice, "new_pos");
    DevCall c;
    c.new_pos = b + w0;
    c.log_ratio = b + 2 * nw + w0;
    c.phase = b + 3 * nw + w0;
    c.new_grad = b + 4 * nw + w0;
    c.new_lap = b + 6 * nw + w0;
    c.new_pgrad = b + 8 * nw + w0;
    c.new_plap = b + 10 * nw + w0;
    s->set_model_nw(s->nw);
    launch_call(s->dmodel, s->shifted(w0), c, p, 1, w1 - w0, s->mixed(), s->cplx, s->walker_lds, s->stream);
    hip_check(hipGetLastError(), "launch ratio");
    rows(b + 2 * nw, log_ratio, 1, hipMemcpyDeviceToHost, "log_ratio");
    if (phase) rows(b + 3 * nw, phase, 1, hipMemcpyDeviceToHost, "phase");
    if (new_grad) rows(b + 4 * nw, new_grad, 2, hipMemcpyDeviceToHost, "new_grad");
    if (new_lap) rows(b + 6 * nw, new_lap, 2, hipMemcpyDeviceToHost, "new_lap");
    if (new_pgrad) rows(b + 8 * nw, new_pgrad, 2, hipMemcpyDeviceToHost, "new_pgrad");
    if (new_plap) rows(b + 10 * nw, new_plap, 2, hipMemcpyDeviceToHost, "new_plap");
}

static void accept_range(alea_system* s, int p, int w0, int w1, const unsigned char* mask) {
    if (!s->evaluated) throw InvalidArg("evaluate first");
    if (p < 0 || p >= s->np) throw InvalidArg("particle index out of range");
    if (!mask) throw InvalidArg("null mask");
    const int n = w1 - w0;
    if (n == 0) return;
    hip_check(copy_on(s->stream, s->call_mask.p + w0, mask + w0, static_cast<size_t>(n), hipMemcpyHostToDevice), "mask");
    DevCall c;
    c.new_pos = s->call_buf.p + w0;
    c.mask = s->call_mask.p + w0;
    s->set_model_nw(s->nw);
    const DevState st = s->shifted(w0);
    launch_call(s->dmodel, st, c, p, 2, n, s->mixed(), s->cplx, s->walker_lds, s->stream);
    hip_check(hipGetLastError(), "launch accept");
    s->energies_valid = false;
    // blocks of the range whose update count reached refresh_updates are
    // rebuilt after each accept call (the reference rebuilds inside the call,
    // slater_determinant.cpp:788-790); a forced full refresh (whole batch)
    // counts one sweep's worth of calls as a sweep
    if (s->refresh_every > 0 && w0 == 0 && w1 == s->nw &&
        ++s->since_refresh >= static_cast<long long>(s->refresh_every) * s->np) {
        s->refresh();
    } else if (s->host.n_blocks > 0) {
        launch_refresh(s->dmodel, st, n, s->host.n_blocks, s->st.nb_max, s->mixed(), s->cplx, s->refresh_updates,
                       s->stream, s->refresh_mfma);
        hip_check(hipGetLastError(), "launch refresh");
    }
    s->sync();
    s->check_status(w0, w1);
}

int alea_ratio(alea_system* s, int p, const double* new_pos, double* log_ratio, double* phase,
               double* new_grad, double* new_lap) {
    return guarded([&] {
        Sys(s);
        ratio_range(s, p, 0, s->nw, new_pos, log_ratio, phase, new_grad, new_lap);
    });
}

int alea_accept(alea_system* s, int p, const unsigned char* mask) {
    return guarded([&] {
        Sys(s);
        accept_range(s, p, 0, s->nw, mask);
    });
}

int alea_ratio_range(alea_system* s, int p, int w_begin, int w_end, const double* new_pos, double* log_ratio,
                     double* phase, double* new_grad, double* new_lap) {
    return guarded([&] {
        Sys(s);
        const auto [w0, w1] = s->range(w_begin, w_end);
        ratio_range(s, p, w0, w1, new_pos, log_ratio, phase, new_grad, new_lap);
    });
}

int alea_ratio_range_phase(alea_system* s, int p, int w_begin, int w_end, const double* new_pos, double* log_ratio,
                           double* phase, double* new_grad, double* new_lap, double* new_pgrad, double* new_plap) {
    return guarded([&] {
        Sys(s);
        const auto [w0, w1] = s->range(w_begin, w_end);
        ratio_range(s, p, w0, w1, new_pos, log_ratio, phase, new_grad, new_lap, new_pgrad, new_plap);
    });
}

// k_measure on the walkers [w0, w1) (its grad_out planes offset to the
// range's first column, the batch's walker stride kept), then rows
// [row0, row0 + n_rows) of each plane copied out of the range's columns into
// outs[q] (rows from its start, the batch's walker stride).  Other walkers'
// entries are neither read nor written.  The staging planes are shared by the
// whole population: calls on one system are not made concurrently.
static void derivative_rows(alea_system* s, int w0, int w1, size_t row0, size_t n_rows, double* const outs[4]) {
    const size_t nw = s->nw, n = static_cast<size_t>(w1 - w0), rows = static_cast<size_t>(s->np) * 2;
    if (n == 0 || (!outs[0] && !outs[1] && !outs[2] && !outs[3])) return;
    const size_t plane = rows * nw;
    // sized by the capacity: after the first call a population change
    // (alea_permute) never reallocates it under a later range call
    const size_t staging = 4 * rows * static_cast<size_t>(s->capacity);
    if (s->grad_buf.n < staging) s->grad_buf.alloc(staging, s->stream);
    DevState st = s->shifted(w0);
    st.grad_out = s->grad_buf.p + w0;
    st.lap_out = s->grad_buf.p + plane + w0;
    st.pgrad_out = s->grad_buf.p + 2 * plane + w0;
    st.plap_out = s->grad_buf.p + 3 * plane + w0;
    s->set_model_nw(s->nw);
    launch_measure(s->dmodel, st, static_cast<int>(n), s->mixed(), s->cplx, s->walker_lds, s->stream);
    hip_check(hipGetLastError(), "launch measure");
    for (int q = 0; q < 4; ++q) {
        if (!outs[q]) continue;
        hip_check(hipMemcpy2DAsync(outs[q] + w0, nw * sizeof(double), s->grad_buf.p + q * plane + row0 * nw + w0,
                                   nw * sizeof(double), n * sizeof(double), n_rows, hipMemcpyDeviceToHost,
                                   s->stream), "derivative planes");
    }
    s->sync();
}

// log|psi| and the gradient planes of the walkers [w0, w1).
int alea_get_derivatives_range(alea_system* s, int w_begin, int w_end, double* log_psi, double* grad, double* lap,
                               double* pgrad, double* plap) {
    return guarded([&] {
        Sys(s);
        if (!s->evaluated) throw InvalidArg("evaluate first");
        const auto [w0, w1] = s->range(w_begin, w_end);
        const size_t n = static_cast<size_t>(w1 - w0);
        if (n == 0) return;
        if (log_psi) {
            std::vector<double> l(n * 4);
            hip_check(copy_on(s->stream, l.data(), s->logs.p + w0 * 4, l.size() * sizeof(double),
                              hipMemcpyDeviceToHost), "logs");
            for (size_t k = 0; k < n; ++k) {
                double t = l[k * 4];
                for (int b = 0; b < s->host.n_blocks; ++b) t += l[k * 4 + 1 + b];
                log_psi[w0 + k] = t;
            }
        }
        double* const outs[4] = {grad, lap, pgrad, plap};
        derivative_rows(s, w0, w1, 0, static_cast<size_t>(s->np) * 2, outs);
    });
}

// Particle p's rows of the same planes, [dim][nw].
int alea_get_particle_derivatives_range(alea_system* s, int p, int w_begin, int w_end, double* grad, double* lap,
                                        double* pgrad, double* plap) {
    return guarded([&] {
        Sys(s);
        if (!s->evaluated) throw InvalidArg("evaluate first");
        if (p < 0 || p >= s->np) throw InvalidArg("particle index out of range");
        const auto [w0, w1] = s->range(w_begin, w_end);
        double* const outs[4] = {grad, lap, pgrad, plap};
        derivative_rows(s, w0, w1, static_cast<size_t>(p) * 2, 2, outs);
    });
}

int alea_delta_potential(alea_system* s, int p, int w_begin, int w_end, const double* new_pos, double* delta) {
    return guarded([&] {
        Sys(s);
        const auto [w0, w1] = s->range(w_begin, w_end);
        if (!s->evaluated) throw InvalidArg("evaluate first");
        if (p < 0 || p >= s->np) throw InvalidArg("particle index out of range");
        if (!new_pos || !delta) throw InvalidArg("null proposal or output");
        const size_t nw = s->nw, n = static_cast<size_t>(w1 - w0);
        if (n == 0) return;
        double* b = s->call_buf.p;
        for (int d = 0; d < 2; ++d)
            hip_check(copy_on(s->stream, b + d * nw + w0, new_pos + d * nw + w0, n * sizeof(double),
                              hipMemcpyHostToDevice), "new_pos");
        launch_delta_potential(s->dmodel, s->shifted(w0), w1 - w0, p, b + w0, static_cast<int>(nw), b + 2 * nw + w0,
                               s->stream);
        hip_check(hipGetLastError(), "launch delta_potential");
        hip_check(copy_on(s->stream, delta + w0, b + 2 * nw + w0, n * sizeof(double), hipMemcpyDeviceToHost),
                  "delta");
    });
}

// The listed parameters resolved against the system's factors, packed for
// the derivative kernels, with the breakup tables of image-summed ones
// gathered into two device arrays.
struct PackedParams {
    std::vector<DevParam> host;
    DevBuf<DevParam> prm;
    DevBuf<double> w, spl;
};

static void pack_params(alea_system* s, const alea_param* params, int n, PackedParams& out) {
    const DevModel& m = s->host;
    std::vector<DevParam>& dp = out.host;
    dp.assign(static_cast<size_t>(n), DevParam{});
    std::vector<double> wts(1, 0.0), spl(1, 0.0);
    for (int i = 0; i < n; ++i) {
        const alea_param& q = params[i];
        if (q.factor < 0 || q.factor >= static_cast<int>(s->factors.size()))
            throw InvalidArg("factor index out of range");
        const auto [kind, idx] = s->factors[static_cast<size_t>(q.factor)];
        DevParam& d = dp[static_cast<size_t>(i)];
        d.index = q.index;
        if (kind == 0) {
            const DevRadial& R = m.pair[idx].rad;
            if (R.kind == 2) throw InvalidArg("the McMillan factor has no free parameters");
            const int n_par = R.kind == 1 ? R.num + 1 : R.num + R.den;
            if (q.index < 0 || q.index >= n_par) throw InvalidArg("parameter index out of range");
            d.pair = idx;
            if (R.kind == 1 && q.index == R.num) {
                double ipw = 1.0;   // (-L)^(C+1)
                for (int k = 0; k <= R.C; ++k) ipw *= -R.L;
                d.da1_dl = R.gamma * R.C / ipw - R.c[0] * R.C / (R.L * R.L);
            }
            if (R.image) {
                if (!q.weights || !q.spline) throw InvalidArg("a parameter of an image-summed factor needs its tables");
                if (q.n_knots < 2 || q.n_knots > kMaxKnots) throw InvalidArg("parameter spline: knot count out of range");
                d.wofs = static_cast<int>(wts.size());
                wts.insert(wts.end(), q.weights, q.weights + m.nkj);
                d.sofs = static_cast<int>(spl.size());
                spl.insert(spl.end(), q.spline, q.spline + 3 * q.n_knots);
                d.n_knots = q.n_knots;
                d.r_cut = q.r_cut;
                d.inv_delta = q.inv_delta;
                d.limit = q.limit;
            }
        } else if (kind == 1) {
            if (q.index != 0) throw InvalidArg("the Nosanow factor has one parameter");
            d.group = idx;
        } else {
            // GaussianOrbitalSet's shared width (orbital_set.h:228-247): only a
            // block whose Gaussians all have one width has it
            bool uniform = m.blk_kind[idx] == kGaussian && !s->cplx;
            for (int k = 1; uniform && k < m.blk_n[idx]; ++k) uniform = m.orb_alpha[idx][k] == m.orb_alpha[idx][0];
            if (!uniform) throw InvalidArg("this determinant has no free parameters");
            if (q.index != 0) throw InvalidArg("parameter index out of range");
            d.block = idx;
            d.alpha = m.orb_alpha[idx][0];
        }
    }
    out.prm.alloc(dp.size(), s->stream);
    out.w.alloc(wts.size(), s->stream);
    out.spl.alloc(spl.size(), s->stream);
    hip_check(hipMemcpyAsync(out.prm.p, dp.data(), dp.size() * sizeof(DevParam), hipMemcpyHostToDevice, s->stream),
              "parameters");
    hip_check(hipMemcpyAsync(out.w.p, wts.data(), wts.size() * sizeof(double), hipMemcpyHostToDevice, s->stream),
              "weights");
    hip_check(hipMemcpyAsync(out.spl.p, spl.data(), spl.size() * sizeof(double), hipMemcpyHostToDevice, s->stream),
              "splines");
}

int alea_parameter_derivatives(alea_system* s, const alea_param* params, int n, double* derivs) {
    return guarded([&] {
        Sys(s);
        if (!s->evaluated) throw InvalidArg("evaluate first");
        if (n < 0 || (n > 0 && (!params || !derivs))) throw InvalidArg("null parameters or output");
        if (n == 0) return;
        PackedParams pk;
        pack_params(s, params, n, pk);
        const size_t nw = static_cast<size_t>(s->nw);
        DevBuf<double> d_out;
        d_out.alloc(static_cast<size_t>(n) * nw, s->stream);
        s->set_model_nw(s->nw);
        launch_param_derivs(s->dmodel, s->st, s->nw, s->mixed(), s->cplx, pk.prm.p, n, pk.w.p, pk.spl.p, d_out.p,
                            static_cast<int>(nw), s->stream);
        hip_check(hipGetLastError(), "launch parameter derivatives");
        hip_check(copy_on(s->stream, derivs, d_out.p, static_cast<size_t>(n) * nw * sizeof(double),
                          hipMemcpyDeviceToHost), "derivatives");
    });
}

int alea_local_energy_parameter_derivatives(alea_system* s, const alea_param* params, int n, double* de_re,
                                            double* de_im) {
    return guarded([&] {
        Sys(s);
        if (!s->evaluated) throw InvalidArg("evaluate first");
        if (n < 0 || (n > 0 && (!params || !de_re))) throw InvalidArg("null parameters or output");
        if (n == 0) return;
        PackedParams pk;
        pack_params(s, params, n, pk);
        const size_t nw = static_cast<size_t>(s->nw), npnw = static_cast<size_t>(s->np) * 2 * nw;
        // every particle's gradient (and phase gradient) of the whole trial
        measure(s, true);
        const double* grad = s->grad_buf.p;
        const double* pgrad = s->cplx ? s->grad_buf.p + 2 * npnw : nullptr;
        DevBuf<double> d_re, d_im;
        d_re.alloc(static_cast<size_t>(n) * nw, s->stream);
        d_im.alloc(static_cast<size_t>(n) * nw, s->stream);
        launch_param_eloc(s->dmodel, s->st, s->nw, pk.prm.p, n, pk.w.p, pk.spl.p, grad, pgrad, d_re.p, d_im.p,
                          static_cast<int>(nw), s->stream);
        hip_check(hipGetLastError(), "launch local-energy parameter derivatives");
        DevBuf<double> ybuf;
        for (int i = 0; i < n; ++i) {
            const DevParam& q = pk.host[static_cast<size_t>(i)];
            if (q.block < 0) continue;
            // dD A for a chunk of walkers at a time, n16^2 doubles each (the
            // block padded to a multiple of 16), within 256 MiB
            const size_t nb = static_cast<size_t>((s->host.blk_n[q.block] + 15) & ~15);
            const int chunk = static_cast<int>(std::max<size_t>(1, std::min(nw, (size_t{256} << 20) / (nb * nb * 8))));
            if (ybuf.n < static_cast<size_t>(chunk) * nb * nb) ybuf.alloc(static_cast<size_t>(chunk) * nb * nb, s->stream);
            for (int w0 = 0; w0 < s->nw; w0 += chunk) {
                const int cnt = std::min(chunk, s->nw - w0);
                launch_gauss_eloc(s->dmodel, s->st, s->mixed(), q.block, q.alpha, grad, ybuf.p, w0, cnt,
                                  d_re.p + static_cast<size_t>(i) * nw, s->stream);
                hip_check(hipGetLastError(), "launch Gaussian width derivative");
            }
        }
        hip_check(copy_on(s->stream, de_re, d_re.p, static_cast<size_t>(n) * nw * sizeof(double),
                          hipMemcpyDeviceToHost), "dE/dp");
        if (de_im)
            hip_check(copy_on(s->stream, de_im, d_im.p, static_cast<size_t>(n) * nw * sizeof(double),
                              hipMemcpyDeviceToHost), "dE_im/dp");
    });
}

int alea_accept_range(alea_system* s, int p, int w_begin, int w_end, const unsigned char* mask) {
    return guarded([&] {
        Sys(s);
        const auto [w0, w1] = s->range(w_begin, w_end);
        accept_range(s, p, w0, w1, mask);
    });
}

// One sweep of the loop below: its k_draw on the device streams, the sweep,
// then the refresh.  With `events` the sweep kernel is bracketed by the i-th
// pair of the system's timing events (not inside a graph capture).
static void launch_one_sweep(alea_system* s, const SweepParams& P, int i, int stride, bool events) {
    SweepParams Pi = P;
    Pi.n_sweeps = 1;
    if (P.tape) {
        Pi.tape = P.tape + i * static_cast<size_t>(s->np) * 3 * stride;
    } else {
        launch_draw(s->st, s->np, s->nw, s->dtape.p, stride, s->stream);
        hip_check(hipGetLastError(), "launch draw");
        Pi.tape = s->dtape.p;
        Pi.tape_stride = stride;
    }
    if (events) hip_check(hipEventRecord(s->sweep_ev[2 * i], s->stream), "event record");
    launch_sweep(s->dmodel, s->st, Pi, s->nw, s->mixed(), s->cplx, s->rows_max, s->walker_lds, s->stream);
    hip_check(hipGetLastError(), "launch sweep");
    if (events) hip_check(hipEventRecord(s->sweep_ev[2 * i + 1], s->stream), "event record");
    s->refresh_after_sweep();
}

static void run_sweeps_eager(alea_system* s, const SweepParams& P, int stride, int n_sweeps) {
    while (s->sweep_ev.size() < 2 * static_cast<size_t>(n_sweeps)) {
        hipEvent_t e = nullptr;
        hip_check(hipEventCreate(&e), "event");
        s->sweep_ev.push_back(e);
    }
    for (int i = 0; i < n_sweeps; ++i) launch_one_sweep(s, P, i, stride, true);
    s->sweeps_timed = n_sweeps;
}

// The same launches captured once into a hipGraph and replayed while nothing
// they bake in changes (sweep count, step, walker count, refresh phase,
// device buffers).  The capture itself launches nothing: since_refresh is
// advanced by the capture as by the eager loop and restored on a replay.
// Nothing inside the graph is a memset (the acceptance counters are cleared
// before it, outside).
static void run_sweep_graph(alea_system* s, const SweepParams& P, int stride, int n_sweeps) {
    alea_system::SweepGraphKey k;
    std::memset(&k, 0, sizeof k);
    k.n_sweeps = n_sweeps;
    k.nw = s->nw;
    k.np = s->np;
    k.dmc = P.dmc;
    k.rows_max = s->rows_max;
    k.precision = s->precision;
    k.cplx = s->cplx ? 1 : 0;
    k.refresh_every = s->refresh_every;
    k.refresh_updates = s->refresh_updates;
    k.refresh_mfma = s->refresh_mfma ? 1 : 0;
    k.since_refresh = s->since_refresh;
    k.tau = P.tau;
    k.dtape = s->dtape.p;
    k.dmodel = s->dmodel;
    k.walker_lds = s->walker_lds;
    std::memcpy(&k.st, &s->st, sizeof(DevState));
    if (!s->graph_exec || std::memcmp(&k, &s->graph_key, sizeof k) != 0) {
        s->drop_graph();
        hipGraph_t graph = nullptr;
        hip_check(hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed), "begin capture");
        try {
            for (int i = 0; i < n_sweeps; ++i) launch_one_sweep(s, P, i, stride, false);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            s->since_refresh = k.since_refresh;
            throw;
        }
        hip_check(hipStreamEndCapture(s->stream, &graph), "end capture");
        const hipError_t e = hipGraphInstantiate(&s->graph_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) s->graph_exec = nullptr;
        hip_check(e, "graph instantiate");
        s->graph_key = k;
        s->graph_since_after = s->since_refresh;
    }
    hip_check(hipEventRecord(s->ev0, s->stream), "event record");
    hip_check(hipGraphLaunch(s->graph_exec, s->stream), "graph launch");
    hip_check(hipEventRecord(s->ev1, s->stream), "event record");
    s->since_refresh = s->graph_since_after;
    s->sweeps_timed = 0;
}

static void run_sweeps(alea_system* s, int n_sweeps, double tau, const double* tape, int stride,
                       SweepKind kind, long long* n_accepted) {
    if (!s->evaluated) throw InvalidArg("evaluate first");
    if (n_sweeps < 0) throw InvalidArg("negative sweep count");
    SweepParams P;
    P.n_sweeps = n_sweeps;
    P.tau = tau;
    P.dmc = kind;
    if (tape) {
        const size_t len = static_cast<size_t>(n_sweeps) * s->np * 3 * stride;
        if (s->tape.n < len) s->tape.alloc(len, s->stream);
        hip_check(copy_on(s->stream, s->tape.p, tape, len * sizeof(double), hipMemcpyHostToDevice), "tape");
        P.tape = s->tape.p;
        P.tape_stride = stride;
    } else if (n_sweeps > 0) {
        // device streams: k_draw writes each sweep's numbers before it
        const size_t len = static_cast<size_t>(s->np) * 3 * s->nw;
        if (s->dtape.n < len) s->dtape.alloc(static_cast<size_t>(s->np) * 3 * s->capacity, s->stream);
        stride = s->nw;
    }
    hip_check(hipMemsetAsync(s->nacc.p, 0, s->nw * sizeof(long long), s->stream), "memset");
    s->set_model_nw(s->nw);
    if (n_sweeps > 0) s->energies_valid = false;
    if (!s->ev0) {
        hip_check(hipEventCreate(&s->ev0), "event");
        hip_check(hipEventCreate(&s->ev1), "event");
    }
    // One launch per sweep (after its k_draw on the device streams); the
    // batched inverse refresh runs every refresh_every sweeps (at most
    // refresh_every * Nb Sherman-Morrison updates between rebuilds).  The
    // events bracket the draw and sweep launches only, or with sweep graphs
    // the whole block.
    if (!P.tape && s->sweep_graphs && n_sweeps > 0) {
        run_sweep_graph(s, P, stride, n_sweeps);
        s->timed = true;
        s->graph_timed = true;
    } else {
        run_sweeps_eager(s, P, stride, n_sweeps);
        s->timed = n_sweeps > 0;
        s->graph_timed = false;
    }
    if (n_accepted) {

        s->sync();
        std::vector<long long> a(s->nw);
        hip_check(copy_on(s->stream, a.data(), s->nacc.p, s->nw * sizeof(long long), hipMemcpyDeviceToHost), "nacc");
        long long t = 0;
        for (long long x : a) t += x;
        *n_accepted = t;
        s->check_status();
    }
}

// ── VMC sweeps that allocate nothing ─────────────────────────────────────
// The device-stream sweep stages each sweep's random numbers (np * 3 per
// walker) between k_draw and k_sweep.  alea_vmc_sweeps keeps that tape in a
// buffer it allocates on first use; alea_vmc_sweeps_static stages it in this
// file-scope array instead, a chunk of walkers at a time.  Each GPU holds its
// own copy of the array until the process ends.
namespace {
constexpr size_t kStaticTapeDoubles = size_t(4) << 20;   // 32 MiB per GPU
__device__ __align__(256) double scratch[kStaticTapeDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of this GPU's copy of `scratch` (the host-side name is
// only a stand-in): looked up once per device, with that device current.
double* device_scratch() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw DeviceError("device index out of range for the static tape");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}
}  // namespace

int alea_vmc_sweeps_static(alea_system* s, int n_sweeps, double tau) {
    return guarded([&] {
        Sys(s);
        if (!s->evaluated) throw InvalidArg("evaluate first");
        if (n_sweeps < 0) throw InvalidArg("negative sweep count");
        if (!s->seeded) throw InvalidArg("seed the device streams first (alea_seed)");
        const size_t per_walker = static_cast<size_t>(s->np) * 3;
        if (per_walker > kStaticTapeDoubles) throw InvalidArg("too many particles for the static tape");
        if (n_sweeps == 0 || s->nw == 0) return;
        double* tape = device_scratch();
        // walkers per chunk: as many as the array holds (positive: np <= 256)
        const int chunk = static_cast<int>(std::min<size_t>(kStaticTapeDoubles / per_walker, s->nw));
        s->energies_valid = false;
        s->set_model_nw(s->nw);
        SweepParams P;
        P.n_sweeps = 1;
        P.tau = tau;
        P.dmc = kSweepVmc;
        P.tape = tape;
        for (int k = 0; k < n_sweeps; ++k) {
            for (int w0 = 0; w0 < s->nw; w0 += chunk) {
                const int n = std::min(chunk, s->nw - w0);
                const DevState st = s->shifted(w0);
                // k_draw writes every number the sweep then reads: nothing
                // in the array needs clearing between calls
                launch_draw(st, s->np, n, tape, n, s->stream);
                hip_check(hipGetLastError(), "launch draw");
                P.tape_stride = n;
                launch_sweep(s->dmodel, st, P, n, s->mixed(), s->cplx, s->rows_max, s->walker_lds, s->stream);
                hip_check(hipGetLastError(), "launch sweep");
            }
            s->refresh_after_sweep();
        }
    });
}

int alea_vmc_sweeps(alea_system* s, int n_sweeps, double tau, const double* tape,
                    long long* n_accepted) {
    return guarded([&] {
        Sys(s);
        run_sweeps(s, n_sweeps, tau, tape, s->nw, kSweepVmc, n_accepted);
    });
}

int alea_vmc_drift_sweeps(alea_system* s, int n_sweeps, double tau, const double* tape,
                          long long* n_accepted) {
    return guarded([&] {
        Sys(s);
        run_sweeps(s, n_sweeps, tau, tape, s->nw, kSweepLangevin, n_accepted);
    });
}

int alea_dmc_sweeps(alea_system* s, int n_sweeps, double tau, const double* tape, int tape_stride,
                    double* rr_proposed, double* rr_accepted, long long* n_accepted) {
    return guarded([&] {
        Sys(s);
        if (tape && tape_stride < s->nw) throw InvalidArg("tape stride smaller than the walker count");
        if (!rr_proposed != !rr_accepted) throw InvalidArg("give both rr arrays or neither");
        const size_t nw = s->nw;
        std::vector<double> rr(2 * nw, 0.0);
        if (rr_proposed) {
            for (size_t w = 0; w < nw; ++w) {
                rr[2 * w] = rr_proposed[w];
                rr[2 * w + 1] = rr_accepted[w];
            }
            hip_check(copy_on(s->stream, s->rr.p, rr.data(), rr.size() * sizeof(double), hipMemcpyHostToDevice), "rr");
        }
        run_sweeps(s, n_sweeps, tau, tape, tape ? tape_stride : s->nw, kSweepDmc, n_accepted);
        if (rr_proposed) {
            s->sync();
            hip_check(copy_on(s->stream, rr.data(), s->rr.p, rr.size() * sizeof(double), hipMemcpyDeviceToHost), "rr");
            for (size_t w = 0; w < nw; ++w) {
                rr_proposed[w] = rr[2 * w];
                rr_accepted[w] = rr[2 * w + 1];
            }
        }
    });
}

int alea_refresh_inverse(alea_system* s, int w_begin, int w_end) {
    return guarded([&] {
        Sys(s);
        if (!s->evaluated) throw InvalidArg("evaluate first");
        const auto [w0, w1] = s->range(w_begin, w_end);
        if (s->host.n_blocks == 0 || w1 == w0) return;
        launch_refresh(s->dmodel, s->shifted(w0), w1 - w0, s->host.n_blocks, s->st.nb_max, s->mixed(), s->cplx, 0,
                       s->stream, s->refresh_mfma);
        hip_check(hipGetLastError(), "launch refresh");
        if (w0 == 0 && w1 == s->nw) s->since_refresh = 0;
        s->energies_valid = false;
    });
}

int alea_set_cancellation_factor(alea_system* s, double factor) {
    return guarded([&] {
        Sys(s);
        if (!(factor >= 0.0) || !std::isfinite(factor)) throw InvalidArg("cancellation factor must be finite and >= 0");
        s->host.cancel_factor = factor;
        if (s->built)
            hip_check(hipMemcpyAsync(&s->dmodel->cancel_factor, &s->host.cancel_factor, sizeof(double),
                                     hipMemcpyHostToDevice, s->stream),
                      "model cancellation factor");
        s->sync();
    });
}

int alea_set_refresh_interval(alea_system* s, int sweeps) {
    return guarded([&] {
        Sys(s);
        if (sweeps < 0) throw InvalidArg("refresh interval must not be negative");
        s->refresh_every = sweeps;
        s->since_refresh = 0;
    });
}

int alea_set_refresh_kernel(alea_system* s, int kind) {
    return guarded([&] {
        Sys(s);
        if (kind != 0 && kind != 1) throw InvalidArg("refresh kernel must be 0 (vector units) or 1 (matrix cores)");
        s->refresh_mfma = kind == 1;
    });
}

int alea_set_refresh_updates(alea_system* s, int updates) {
    return guarded([&] {
        Sys(s);
        if (updates < 1) throw InvalidArg("refresh update count must be at least one");
        s->refresh_updates = updates;
    });
}

int alea_vmc_sweeps_fused(alea_system* const* systems, int n_systems, int n_sweeps, const double* taus) {
    return guarded([&] {
        if (!systems || !taus) throw InvalidArg("null systems or step array");
        if (n_systems < 1 || n_systems > kMaxFused)
            throw InvalidArg("a fused launch takes 1 to " + std::to_string(kMaxFused) + " systems");
        if (n_sweeps < 0) throw InvalidArg("negative sweep count");
        alea_system* lead = Sys(systems[0]);
        const int nr = sweep_tile_rows(lead->mixed(), lead->cplx, lead->rows_max);
        size_t lds = 0;
        long long total = 0;
        for (int i = 0; i < n_systems; ++i) {
            alea_system* s = Sys(systems[i]);
            for (int k = 0; k < i; ++k)
                if (systems[k] == s) throw InvalidArg("a system appears twice in one fused launch");
            if (!s->evaluated) throw InvalidArg("evaluate first");
            if (s->precision != lead->precision || s->cplx != lead->cplx ||
                sweep_tile_rows(s->mixed(), s->cplx, s->rows_max) != nr)
                throw InvalidArg("fused systems must share the precision, the complex form and the register tile");
            lds = std::max(lds, s->walker_lds);
            total += s->nw;
        }
        if (total > std::numeric_limits<int>::max()) throw InvalidArg("too many walkers for one launch");
        if (n_sweeps == 0 || total == 0) return;
        hipStream_t st = lead->stream;
        // every member's pending work (and its buffers' allocation) comes
        // first: the lead stream waits for each member stream
        FusedArgs args;
        args.n = n_systems;
        for (int i = 0; i < n_systems; ++i) {
            alea_system* s = systems[i];
            s->set_model_nw(s->nw);
            const size_t len = static_cast<size_t>(s->np) * 3 * s->nw;
            if (s->dtape.n < len) s->dtape.alloc(static_cast<size_t>(s->np) * 3 * s->capacity, s->stream);
            if (!s->fused_join) hip_check(hipEventCreateWithFlags(&s->fused_join, hipEventDisableTiming), "event");
            if (i > 0) {
                hip_check(hipEventRecord(s->fused_join, s->stream), "event record");
                hip_check(hipStreamWaitEvent(st, s->fused_join, 0), "stream wait");
            }
            FusedJob& J = args.job[i];
            J.m = s->dmodel;
            J.s = s->st;
            J.p.n_sweeps = 1;
            J.p.tau = taus[i];
            J.p.tape = s->dtape.p;
            J.p.tape_stride = s->nw;
            J.p.dmc = kSweepVmc;
            J.first = i == 0 ? 0 : args.job[i - 1].first + systems[i - 1]->nw;
            J.nw = s->nw;
            s->energies_valid = false;
        }
        // per sweep: each member's draws, one fused sweep, each member's
        // refresh when it is due
        // (a member with no walkers has no work: its draws and refresh would
        // be empty launches, and the sweep kernel never reaches it)
        for (int k = 0; k < n_sweeps; ++k) {
            for (int i = 0; i < n_systems; ++i) {
                alea_system* s = systems[i];
                if (s->nw == 0) continue;
                launch_draw(s->st, s->np, s->nw, s->dtape.p, s->nw, st);
                hip_check(hipGetLastError(), "launch draw");
            }
            launch_sweep_fused(args, static_cast<int>(total), lead->mixed(), lead->cplx, lead->rows_max, lds, st);
            hip_check(hipGetLastError(), "launch fused sweep");
            for (int i = 0; i < n_systems; ++i) {
                alea_system* s = systems[i];
                if (s->nw == 0) continue;
                s->refresh_after_sweep(st);
            }
        }
        // each member's later work waits for the fused block
        hip_check(hipEventRecord(lead->fused_join, st), "event record");
        for (int i = 1; i < n_systems; ++i)
            hip_check(hipStreamWaitEvent(systems[i]->stream, lead->fused_join, 0), "stream wait");
    });
}

int alea_set_sweep_graphs(alea_system* s, int on) {
    return guarded([&] {
        Sys(s);
        s->sweep_graphs = on != 0;
        if (!s->sweep_graphs) s->drop_graph();
    });
}

int alea_synchronize(alea_system* s) {
    return guarded([&] {
        Sys(s);
        s->sync();
    });
}

int alea_rebuild_counts(alea_system* s, long long* counts, int* per_walker, int reset) {
    return guarded([&] {
        Sys(s);
        if (!s->built) throw InvalidArg("no walkers yet: set positions first");
        const size_t n = static_cast<size_t>(s->capacity) * kRebuildCauses;
        std::vector<int> c(n);
        hip_check(copy_on(s->stream, c.data(), s->rebuilds.p, n * sizeof(int), hipMemcpyDeviceToHost), "rebuilds");
        if (counts)
            for (int k = 0; k < kRebuildCauses; ++k) {
                long long t = 0;
                for (int w = 0; w < s->nw; ++w) t += c[static_cast<size_t>(w) * kRebuildCauses + k];
                counts[k] = t;
            }
        if (per_walker)
            for (int w = 0; w < s->nw; ++w) {
                int t = 0;
                for (int k = 0; k < kRebuildCauses; ++k) t += c[static_cast<size_t>(w) * kRebuildCauses + k];
                per_walker[w] = t;
            }
        if (reset) {
            // a fill kernel rather than a memset node, so a later graph capture never records one
            launch_fill_int(s->rebuilds.p, static_cast<int>(n), 0, s->stream);
            hip_check(hipGetLastError(), "launch fill");
            s->sync();
        }
    });
}

int alea_last_sweep_ms(alea_system* s, double* ms) {
    return guarded([&] {
        Sys(s);
        if (!s->timed) throw InvalidArg("no sweep has been launched");
        hipEvent_t a = s->ev0, b = s->ev1;
        if (!s->graph_timed) {
            a = s->sweep_ev[2 * (s->sweeps_timed - 1)];
            b = s->sweep_ev[2 * (s->sweeps_timed - 1) + 1];
        }
        hip_check(hipEventSynchronize(b), "event sync");
        float t = 0.0f;
        hip_check(hipEventElapsedTime(&t, a, b), "event elapsed");
        *ms = t;
    });
}

int alea_sweep_kernel_ms(alea_system* s, double* ms, int capacity, int* n) {
    return guarded([&] {
        Sys(s);
        if (!n || (!ms && capacity > 0)) throw InvalidArg("null output");
        if (!s->timed || s->graph_timed) throw InvalidArg("no eager sweeps have been launched");
        const int k = std::min(capacity, s->sweeps_timed);
        for (int i = 0; i < k; ++i) {
            hip_check(hipEventSynchronize(s->sweep_ev[2 * i + 1]), "event sync");
            float t = 0.0f;
            hip_check(hipEventElapsedTime(&t, s->sweep_ev[2 * i], s->sweep_ev[2 * i + 1]), "event elapsed");
            ms[i] = t;
        }
        *n = s->sweeps_timed;
    });
}

int alea_draw_tape(alea_system* s, int n_sweeps, double* out) {
    return guarded([&] {
        Sys(s);
        if (!s->seeded) throw InvalidArg("seed the device streams first (alea_seed)");
        if (n_sweeps < 0 || (n_sweeps > 0 && !out)) throw InvalidArg("bad sweep count or null output");
        const size_t len = static_cast<size_t>(s->np) * 3 * s->nw;
        if (s->dtape.n < len) s->dtape.alloc(static_cast<size_t>(s->np) * 3 * s->capacity, s->stream);
        for (int k = 0; k < n_sweeps; ++k) {
            launch_draw(s->st, s->np, s->nw, s->dtape.p, s->nw, s->stream);
            hip_check(hipGetLastError(), "launch draw");
            hip_check(copy_on(s->stream, out + k * len, s->dtape.p, len * sizeof(double), hipMemcpyDeviceToHost),
                      "draw tape");
        }
    });
}

int alea_seed(alea_system* s, uint64_t seed, uint64_t stream) {
    return guarded([&] {
        Sys(s);
        s->build();
        std::vector<uint64_t> st(static_cast<size_t>(s->capacity) * 4);
        for (int w = 0; w < s->capacity; ++w) {
            // stream_seed(seed, stream, w) (random.h:52-56), then splitmix64
            // seeding of xoshiro256++ (xoshiro_random.h:29-32)
            const uint64_t sd = splitmix64(splitmix64(splitmix64(seed) ^ stream) ^ static_cast<uint64_t>(w));
            for (uint64_t i = 0; i < 4; ++i) st[w * 4 + i] = splitmix64(sd + i * 0x9e3779b97f4a7c15ULL);
        }
        hip_check(copy_on(s->stream, s->rng.p, st.data(), st.size() * sizeof(uint64_t), hipMemcpyHostToDevice), "rng");
        s->seeded = true;
        s->seed_base = seed;
        s->seed_stream = stream;
        s->next_stream_index = static_cast<uint64_t>(s->capacity);
    });
}

// WalkerSet::apply_permutation + TrialWavefunction::permute_caches
// (walker_set.cpp:109-178): every per-walker record is gathered from src[j].
// The inverse records (the largest) are gathered into a spare buffer that then
// swaps with the live one; the others go through one scratch buffer.  Both
// are allocated on first use and kept.
static void permute_device(alea_system* s, int new_nw, const int* src_d) {
    const size_t cap = s->capacity;
    const DevModel& m = s->host;
    // (record, bytes per walker); every size is a whole number of doubles
    const std::pair<void*, size_t> recs[] = {
        {s->pos.p, sizeof(double) * m.np * 2},
        {s->rho.p, sizeof(double) * m.ng * m.nkj * 2},
        {s->logs.p, sizeof(double) * 4},
        {s->cond.p, sizeof(double) * kMaxBlocks},
        {s->updates.p, sizeof(int) * kMaxBlocks},
        {s->e_loc.p, sizeof(double)},
        {s->e_loc_im.p, sizeof(double)},
        {s->rr.p, sizeof(double) * 2},
        {s->weights.p, sizeof(double)},
        {s->e_prev.p, sizeof(double)},
        {s->ncut.p, sizeof(long long)},
        {s->rng.p, sizeof(uint64_t) * 4},
    };
    static_assert(sizeof(int) * kMaxBlocks % sizeof(double) == 0, "update counters pack into doubles");
    size_t widest = 0;
    for (const auto& r : recs) widest = std::max(widest, r.second / sizeof(double));
    if (s->perm_tmp.n < cap * widest) s->perm_tmp.alloc(cap * widest, s->stream);   // before any gather is queued
    for (const auto& r : recs) {
        if (r.second == 0) continue;
        launch_gather(static_cast<double*>(r.first), s->perm_tmp.p, static_cast<int>(r.second / sizeof(double)),
                      src_d, new_nw, s->stream);
        hip_check(hipMemcpyAsync(r.first, s->perm_tmp.p, static_cast<size_t>(new_nw) * r.second,
                                 hipMemcpyDeviceToDevice, s->stream), "permute copy");
    }
    if (s->st.inv_stride) {
        const size_t bytes = static_cast<size_t>(s->st.inv_stride) * s->tsize();
        if (s->inv_spare.n < s->inv_bytes.n) s->inv_spare.alloc(s->inv_bytes.n, s->stream);
        launch_gather(reinterpret_cast<const double*>(s->inv_bytes.p), reinterpret_cast<double*>(s->inv_spare.p),
                      static_cast<int>(bytes / sizeof(double)), src_d, new_nw, s->stream);
        std::swap(s->inv_bytes.p, s->inv_spare.p);
        std::swap(s->inv_bytes.n, s->inv_spare.n);
        s->st.inv = s->inv_bytes.p;
    }
    hip_check(hipGetLastError(), "launch gather");
    s->sync();
    s->nw = new_nw;
}

int alea_permute(alea_system* s, int new_nw, const int* src) {
    return guarded([&] {
        Sys(s);
        if (!s->evaluated) throw InvalidArg("evaluate first");
        if (new_nw <= 0 || new_nw > s->capacity) throw InvalidArg("new walker count exceeds the capacity");
        for (int j = 0; j < new_nw; ++j)
            if (src[j] < 0 || src[j] >= s->nw) throw InvalidArg("permutation source out of range");
        hip_check(copy_on(s->stream, s->src.p, src, new_nw * sizeof(int), hipMemcpyHostToDevice), "src");
        permute_device(s, new_nw, s->src.p);
    });
}

int alea_get_num_walkers(alea_system* s, int* nw) {
    return guarded([&] {
        Sys(s);
        *nw = s->nw;
    });
}

// ── Forward walking's snapshots (forward_walking.cpp:99-108, 148-150) ─────
// A ring of position snapshots of one population, kept on its device: a
// slot is a copy of the walker-major position records of the population's
// walkers at one step.  The ancestry (integers) stays with the caller; a
// load gathers the kept ancestors' configurations into a measurement batch.
struct alea_snapshots {
    alea_system* pop = nullptr;
    int n_slots = 0;
    size_t slot_doubles = 0;        // capacity * np * 2
    std::vector<int> nw;            // walkers recorded in each slot (-1: none)
    DevBuf<double> pos;             // [slot][capacity][np][2]
    DevBuf<int> idx;                // [capacity] a load's gather indices
};

int alea_snapshots_create(alea_system* population, int n_slots, alea_snapshots** out) {
    return guarded([&] {
        Sys(population);
        if (!out) throw InvalidArg("null output handle");
        if (!population->built) throw InvalidArg("positions were never set");
        if (n_slots < 1) throw InvalidArg("a snapshot ring needs at least one slot");
        auto r = std::make_unique<alea_snapshots>();
        r->pop = population;
        r->n_slots = n_slots;
        r->slot_doubles = static_cast<size_t>(population->capacity) * population->np * 2;
        r->nw.assign(n_slots, -1);
        r->pos.alloc(r->slot_doubles * n_slots, population->stream);
        r->idx.alloc(population->capacity, population->stream);
        population->sync();
        *out = r.release();
    });
}

int alea_snapshots_destroy(alea_snapshots* r) {
    return guarded([&] {
        if (!r) return;
        r->pop->sync();
        delete r;
    });
}

int alea_snapshots_record(alea_snapshots* r, int slot) {
    return guarded([&] {
        if (!r) throw InvalidArg("null snapshot handle");
        if (slot < 0 || slot >= r->n_slots) throw InvalidArg("snapshot slot out of range");
        alea_system* s = r->pop;
        const size_t n = static_cast<size_t>(s->nw) * s->np * 2;
        hip_check(hipMemcpyAsync(r->pos.p + slot * r->slot_doubles, s->pos.p, n * sizeof(double),
                                 hipMemcpyDeviceToDevice, s->stream),
                  "record snapshot");
        r->nw[slot] = s->nw;
    });
}

int alea_snapshots_load(alea_snapshots* r, int slot, alea_system* batch, const int* kept, const double* weights,
                        int n_kept) {
    return guarded([&] {
        if (!r) throw InvalidArg("null snapshot handle");
        Sys(batch);
        if (slot < 0 || slot >= r->n_slots || r->nw[slot] < 0) throw InvalidArg("snapshot slot never recorded");
        alea_system* s = r->pop;
        if (batch->np != s->np) throw InvalidArg("the batch has another particle count");
        if (n_kept < 1 || n_kept > batch->capacity) throw InvalidArg("kept count out of the batch's capacity");
        if (!kept || !weights) throw InvalidArg("null kept indices or weights");
        for (int c = 0; c < n_kept; ++c)
            if (kept[c] < 0 || kept[c] >= r->nw[slot]) throw InvalidArg("kept index outside the snapshot");
        for (int c = 0; c < n_kept; ++c)
            if (!(weights[c] >= 0.0) || !std::isfinite(weights[c]))
                throw InvalidArg("weights must be finite and non-negative");
        batch->build();
        hip_check(copy_on(s->stream, r->idx.p, kept, n_kept * sizeof(int), hipMemcpyHostToDevice), "kept");
        launch_gather(r->pos.p + slot * r->slot_doubles, batch->pos.p, s->np * 2, r->idx.p, n_kept, s->stream);
        hip_check(hipGetLastError(), "launch gather");
        s->sync();
        batch->nw = n_kept;
        hip_check(copy_on(batch->stream, batch->weights.p, weights, n_kept * sizeof(double), hipMemcpyHostToDevice),
                  "weights");
        batch->set_model_nw(n_kept);
        launch_evaluate(batch->dmodel, batch->st, n_kept, batch->mixed(), batch->cplx, batch->walker_lds,
                        batch->stream);
        hip_check(hipGetLastError(), "launch evaluate");
        batch->refresh();
        batch->sync();
        batch->check_status();
        batch->evaluated = true;
        batch->energies_valid = false;
    });
}

// DMC::advance_measure's set-up of a step (dmc.cpp:349-354): rr sums to zero,
// weights to one, E_old = the current local energies.
int alea_dmc_begin(alea_system* s) {
    return guarded([&] {
        Sys(s);
        if (!s->energies_valid) measure(s, false);
        hip_check(hipMemsetAsync(s->rr.p, 0, 2 * s->nw * sizeof(double), s->stream), "rr");
        hip_check(hipMemsetAsync(s->ncut.p, 0, s->nw * sizeof(long long), s->stream), "ncut");
        launch_fill(s->weights.p, s->nw, 1.0, s->stream);
        hip_check(hipMemcpyAsync(s->e_prev.p, s->e_loc.p, s->nw * sizeof(double), hipMemcpyDeviceToDevice,
                                 s->stream), "e_prev");
        hip_check(hipGetLastError(), "launch fill");
        s->sync();
    });
}

// After one drift-diffusion sweep: H (refresh_derivatives + evaluate_batch,
// dmc.cpp:383-384) and the bare branch factor of every walker (dmc.cpp:
// 404-408).  e_trial, e_ref, e_cut, e_max and tau are PopulationControl's
// state (population_control.h: etrial_, eref_, e_cut_, e_max_, tau_).
int alea_dmc_weights(alea_system* s, double e_trial, double e_ref, double e_cut, double e_max, double tau) {
    return guarded([&] {
        Sys(s);
        measure(s, false);
        launch_dmc_weights(s->e_loc.p, s->e_prev.p, s->weights.p, s->ncut.p, s->nw, e_trial, e_ref, e_cut,
                           e_max, 0.5 * tau, s->stream);
        hip_check(hipGetLastError(), "launch weights");
        s->sync();
    });
}

// The step's (sum w E_L, sum w) in walker order (dmc.cpp:412-418) and the
// number of tapered walker-sweeps.
int alea_dmc_sums(alea_system* s, double* energy_sum, double* weight_sum, long long* n_cut) {
    return guarded([&] {
        Sys(s);
        if (!s->energies_valid) throw InvalidArg("no local energies for the current configurations");
        s->sync();
        const size_t nw = s->nw;
        std::vector<double> w(nw), e(nw);
        std::vector<long long> c(nw);
        hip_check(copy_on(s->stream, w.data(), s->weights.p, nw * sizeof(double), hipMemcpyDeviceToHost), "weights");
        hip_check(copy_on(s->stream, e.data(), s->e_loc.p, nw * sizeof(double), hipMemcpyDeviceToHost), "e_loc");
        hip_check(copy_on(s->stream, c.data(), s->ncut.p, nw * sizeof(long long), hipMemcpyDeviceToHost), "ncut");
        double es = 0.0, ws = 0.0;
        long long nc = 0;
        for (size_t i = 0; i < nw; ++i) {
            es += w[i] * e[i];
            ws += w[i];
            nc += c[i];
        }
        if (energy_sum) *energy_sum = es;
        if (weight_sum) *weight_sum = ws;
        if (n_cut) *n_cut = nc;
    });
}

int alea_get_weights(alea_system* s, double* w) {
    return guarded([&] {
        Sys(s);
        if (!s->built) throw InvalidArg("positions were never set");
        s->sync();
        hip_check(copy_on(s->stream, w, s->weights.p, s->nw * sizeof(double), hipMemcpyDeviceToHost), "weights");
    });
}

int alea_set_weights(alea_system* s, const double* w) {
    return guarded([&] {
        Sys(s);
        if (!s->built) throw InvalidArg("positions were never set");
        for (int i = 0; i < s->nw; ++i)
            if (!(w[i] >= 0.0) || !std::isfinite(w[i])) throw InvalidArg("weights must be finite and non-negative");
        hip_check(copy_on(s->stream, s->weights.p, w, s->nw * sizeof(double), hipMemcpyHostToDevice), "weights");
    });
}

namespace {
// XoshiroRandom::uniform (xoshiro_random.h:90-92) on a caller-held state.
struct HostXoshiro {
    uint64_t* st;
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t result = rotl(st[0] + st[3], 23) + st[0];
        const uint64_t t = st[1] << 17;
        st[2] ^= st[0];
        st[3] ^= st[1];
        st[1] ^= st[2];
        st[0] ^= st[3];
        st[2] ^= t;
        st[3] = rotl(st[3], 45);
        return result;
    }
    double uniform() { return static_cast<double>(static_cast<int64_t>(next() >> 11)) * 0x1.0p-53; }
};
}  // namespace

// WalkerSet::branch (walker_set.cpp:187-239): m_w = int(w_w + xi) copies in
// walker order (xi drawn from rng_state in walker order), the bounds
// [target/5 + 1, 2 target + 1] by random removal / duplication, the single
// best walker on extinction; then the permutation and weights reset to one.
// Multiplicities and the integer scan run on the device; the rare bound
// fix-ups run here with the same stream.  Copies (second and later
// occurrences of a source) get fresh device streams when the device RNG is
// seeded.
int alea_branch(alea_system* s, uint64_t* rng_state, int target, int* new_nw, int* src_out) {
    return guarded([&] {
        Sys(s);
        if (!s->evaluated) throw InvalidArg("evaluate first");
        if (!rng_state) throw InvalidArg("null branch stream state");
        if (target < 1) throw InvalidArg("target population must be positive");
        const size_t n_max = 2 * static_cast<size_t>(target) + 1, n_min = static_cast<size_t>(target) / 5 + 1;
        if (n_max > static_cast<size_t>(s->capacity))
            throw InvalidArg("2 * target + 1 walkers exceed the capacity");
        const int n = s->nw;
        HostXoshiro rng{rng_state};
        std::vector<double> xi(n);
        for (int w = 0; w < n; ++w) xi[w] = rng.uniform();
        hip_check(hipMemcpyAsync(s->xi.p, xi.data(), n * sizeof(double), hipMemcpyHostToDevice, s->stream), "xi");
        launch_branch(s->weights.p, s->xi.p, s->src.p, s->capacity, n, s->branch_total.p, s->stream);
        hip_check(hipGetLastError(), "launch branch");
        long long total = 0;
        hip_check(hipMemcpyAsync(&total, s->branch_total.p, sizeof(long long), hipMemcpyDeviceToHost, s->stream),
                  "total");
        s->sync();
        std::vector<size_t> src;
        if (total > 0 && total <= s->capacity) {
            std::vector<int> d(total);
            hip_check(copy_on(s->stream, d.data(), s->src.p, total * sizeof(int), hipMemcpyDeviceToHost), "src");
            src.assign(d.begin(), d.end());
        } else {
            // slow path: recompute the list here (the same multiplicities)
            std::vector<double> w(n);
            hip_check(copy_on(s->stream, w.data(), s->weights.p, n * sizeof(double), hipMemcpyDeviceToHost), "weights");
            if (total == 0) {
                size_t best = 0;
                for (int i = 1; i < n; ++i)
                    if (w[i] > w[best]) best = i;
                src.push_back(best);
            } else {
                for (int i = 0; i < n; ++i) {
                    const int m = static_cast<int>(w[i] + xi[i]);
                    for (int c = 0; c < m; ++c) src.push_back(i);
                }
            }
        }
        while (src.size() > n_max) {
            const size_t i = static_cast<size_t>(rng.uniform() * static_cast<double>(src.size()));
            src[i] = src.back();
            src.pop_back();
        }
        while (src.size() < n_min) {
            const size_t i = static_cast<size_t>(rng.uniform() * static_cast<double>(src.size()));
            src.push_back(src[i]);
        }
        const int nn = static_cast<int>(src.size());
        std::vector<int> srci(src.begin(), src.end());
        hip_check(copy_on(s->stream, s->src.p, srci.data(), nn * sizeof(int), hipMemcpyHostToDevice), "src");
        permute_device(s, nn, s->src.p);
        launch_fill(s->weights.p, nn, 1.0, s->stream);
        hip_check(hipGetLastError(), "launch fill");
        if (s->seeded) {
            std::vector<char> seen(n, 0);
            std::vector<uint64_t> st(static_cast<size_t>(nn) * 4);
            hip_check(copy_on(s->stream, st.data(), s->rng.p, st.size() * sizeof(uint64_t), hipMemcpyDeviceToHost), "rng");
            for (int j = 0; j < nn; ++j) {
                if (!seen[srci[j]]) {
                    seen[srci[j]] = 1;
                    continue;
                }
                const uint64_t sd = splitmix64(splitmix64(splitmix64(s->seed_base) ^ s->seed_stream) ^
                                               s->next_stream_index++);
                for (uint64_t i = 0; i < 4; ++i) st[j * 4 + i] = splitmix64(sd + i * 0x9e3779b97f4a7c15ULL);
            }
            hip_check(copy_on(s->stream, s->rng.p, st.data(), st.size() * sizeof(uint64_t), hipMemcpyHostToDevice), "rng");
        }
        s->sync();
        *new_nw = nn;
        if (src_out) std::copy(srci.begin(), srci.end(), src_out);
    });
}

}  // extern "C"

// ── Estimators ──────────────────────────────────────────────────────────────
namespace {
constexpr int kEstMaxChunks = 2048;

void est_alloc(alea_system* s) {
    auto& e = s->est;
    const int n = e.entries(s->host.ng);
    e.acc.alloc(static_cast<size_t>(n) + 1, s->stream);
    e.part.alloc(static_cast<size_t>(std::min(s->capacity, kEstMaxChunks)) * std::max(n, 1), s->stream);
}
}  // namespace

extern "C" {

int alea_add_gofr(alea_system* s, double rmax, int n_bins) {
    return guarded([&] {
        Sys(s);
        const int G = s->host.ng;
        if (G <= 0) throw std::invalid_argument("alea_add_gofr: add the particle groups first");
        if (!(rmax > 0.0) || n_bins <= 0) throw std::invalid_argument("alea_add_gofr: needs rmax > 0 and n_bins > 0");
        if (G * (G + 1) / 2 * n_bins > kMaxGofrEntries)
            throw std::invalid_argument("alea_add_gofr: group pairs x bins exceeds 4096");
        if (s->host.cell.periodic && rmax > s->host.cell.ws)
            throw std::invalid_argument("alea_add_gofr: rmax exceeds the cell's Wigner-Seitz inradius");
        auto& e = s->est;
        e.gofr = true;
        e.rmax = rmax;
        e.n_bins = n_bins;
        e.inv_delta = static_cast<double>(n_bins) / rmax;
        e.nfb = G * (G + 1) / 2 * n_bins;
        est_alloc(s);
    });
}

// Density(groups, dim, rmax, n_bins, per_particle = false, cell)
// (density.h:55, density.cpp:16-64): one n_bins x n_bins grid per group over
// the periodic cell (rmax unused) or over [-rmax, rmax)^2 in an open cell.
int alea_add_density(alea_system* s, double rmax, int n_bins, int per_particle) {
    return guarded([&] {
        Sys(s);
        const int G = s->host.ng;
        if (G <= 0) throw std::invalid_argument("alea_add_density: add the particle groups first");
        if (n_bins <= 0) throw std::invalid_argument("alea_add_density: needs n_bins > 0");
        if (!s->host.cell.periodic && !(rmax > 0.0))
            throw std::invalid_argument("alea_add_density: an open cell needs rmax > 0");
        // per-particle grids are numbered in group declaration order
        // (density.cpp:46-56) and indexed by particle on the device: the two
        // agree only for groups declared in ascending particle order
        if (per_particle)
            for (int g = 0; g < G; ++g)
                if (s->host.g_first[g] != (g == 0 ? 0 : s->host.g_last[g - 1]))
                    throw std::invalid_argument(
                        "alea_add_density: per-particle grids need the groups declared in ascending particle order");
        const int rows = per_particle ? s->np : G;
        if (static_cast<long long>(rows) * n_bins * n_bins > kMaxDensityEntries)
            throw std::invalid_argument("alea_add_density: grids x n_bins^2 exceeds " +
                                        std::to_string(kMaxDensityEntries));
        auto& e = s->est;
        e.density = true;
        e.d_rows = rows;
        e.d_per_particle = per_particle != 0;
        e.d_rmax = rmax;
        e.d_bins = n_bins;
        e.d_inv_delta = static_cast<double>(n_bins) / (2.0 * rmax);
        est_alloc(s);
    });
}

// The half-space reciprocal vectors of the reduced basis with |k| <= kmax,
// sorted by |k|, in shells of equal |k| (structure_factor.cpp:16-54).
int alea_add_sofk(alea_system* s, double kmax, int* n_k) {
    return guarded([&] {
        Sys(s);
        if (!s->host.cell.periodic) throw std::invalid_argument("alea_add_sofk: needs a periodic cell");
        if (s->host.ng <= 0) throw std::invalid_argument("alea_add_sofk: add the particle groups first");
        const double* b = s->host.cell.b;
        const double b1 = std::hypot(b[0], b[1]), b2 = std::hypot(b[2], b[3]);
        const int nmax = static_cast<int>(std::ceil(kmax / std::min(b1, b2))) + 1;
        struct K { double kx, ky, k; };
        std::vector<K> ks;
        for (int i = 0; i <= nmax; ++i)
            for (int j = -nmax; j <= nmax; ++j) {
                if (i == 0 && j <= 0) continue;
                const double kx = i * b[0] + j * b[2], ky = i * b[1] + j * b[3];
                const double k = std::hypot(kx, ky);
                if (k <= kmax) ks.push_back({kx, ky, k});
            }
        if (ks.empty()) throw std::invalid_argument("alea_add_sofk: kmax is below the shortest reciprocal vector");
        if (static_cast<int>(ks.size()) > kMaxSofkVectors)
            throw std::invalid_argument("alea_add_sofk: more than 1024 vectors below kmax");
        std::sort(ks.begin(), ks.end(), [](const K& x, const K& y) { return x.k < y.k; });
        auto& e = s->est;
        e.sofk = true;
        e.kx.clear();
        e.ky.clear();
        e.shell.clear();
        const double tol = 1e-9 * std::min(b1, b2);
        double last = 0.0;
        int shell = -1;
        for (const K& q : ks) {
            if (shell < 0 || q.k - last > tol) {
                ++shell;
                last = q.k;
            }
            e.kx.push_back(q.kx);
            e.ky.push_back(q.ky);
            e.shell.push_back(shell);
        }
        e.nk = static_cast<int>(ks.size());
        std::vector<double> kv(e.kx);
        kv.insert(kv.end(), e.ky.begin(), e.ky.end());
        e.kvec.alloc(kv.size(), s->stream);
        hip_check(copy_on(s->stream, e.kvec.p, kv.data(), kv.size() * sizeof(double), hipMemcpyHostToDevice),
                  "upload sofk vectors");
        est_alloc(s);
        *n_k = e.nk;
    });
}

int alea_get_sofk_vectors(alea_system* s, double* kx, double* ky, int* shell) {
    return guarded([&] {
        Sys(s);
        const auto& e = s->est;
        if (!e.sofk) throw std::invalid_argument("alea_get_sofk_vectors: no structure factor");
        std::copy(e.kx.begin(), e.kx.end(), kx);
        std::copy(e.ky.begin(), e.ky.end(), ky);
        std::copy(e.shell.begin(), e.shell.end(), shell);
    });
}

int alea_estimators_start_block(alea_system* s) {
    return guarded([&] {
        Sys(s);
        auto& e = s->est;
        if (e.acc.p) launch_fill(e.acc.p, static_cast<int>(e.acc.n), 0.0, s->stream);
    });
}

int alea_estimators_accumulate(alea_system* s, int use_weights) {
    return guarded([&] {
        Sys(s);
        s->build();
        auto& e = s->est;
        if (!e.gofr && !e.sofk && !e.density) throw std::invalid_argument("alea_estimators_accumulate: no estimators");
        if (s->nw <= 0) return;
        const int chunk = (s->nw + kEstMaxChunks - 1) / kEstMaxChunks;
        const int n = e.entries(s->host.ng);
        launch_estimators(s->dmodel, s->pos.p, use_weights ? s->weights.p : nullptr, s->nw, chunk, e.gofr,
                          e.rmax, e.inv_delta, e.n_bins, e.nfb, e.sofk, e.kvec.p, e.nk, e.density, e.d_rmax,
                          e.d_inv_delta, e.d_bins, e.d_per_particle,
                          e.density_offset(s->host.ng), e.part.p, n, n, e.acc.p, s->stream);
        hip_check(hipGetLastError(), "estimator launch");
    });
}

int alea_estimators_read_density(alea_system* s, double* density, double* wsum) {
    return guarded([&] {
        Sys(s);
        auto& e = s->est;
        if (!e.density) throw std::invalid_argument("alea_estimators_read_density: no density estimator");
        const int n = e.entries(s->host.ng), off = e.density_offset(s->host.ng);
        std::vector<double> h(static_cast<size_t>(n) + 1, 0.0);
        if (e.acc.p)
            hip_check(copy_on(s->stream, h.data(), e.acc.p, h.size() * sizeof(double), hipMemcpyDeviceToHost),
                      "read density");
        if (density) std::copy(h.begin() + off, h.begin() + off + e.density_entries(s->host.ng), density);
        if (wsum) *wsum = h[n];
    });
}

int alea_estimators_read(alea_system* s, double* gofr, double* sofk, double* rho_re, double* rho_im,
                         double* wsum) {
    return guarded([&] {
        Sys(s);
        auto& e = s->est;
        const int n = e.entries(s->host.ng);
        std::vector<double> h(static_cast<size_t>(n) + 1, 0.0);
        if (e.acc.p) {
            hip_check(copy_on(s->stream, h.data(), e.acc.p, h.size() * sizeof(double), hipMemcpyDeviceToHost),
                      "read estimators");
            hip_check(hipStreamSynchronize(s->stream), "read estimators");
        }
        const int G = s->host.ng, npg = G * (G + 1) / 2;
        if (gofr) std::copy(h.begin(), h.begin() + e.nfb, gofr);
        const double* base = h.data() + e.nfb;
        if (sofk && e.sofk) std::copy(base, base + npg * e.nk, sofk);
        if (rho_re && e.sofk) std::copy(base + npg * e.nk, base + (npg + G) * e.nk, rho_re);
        if (rho_im && e.sofk) std::copy(base + (npg + G) * e.nk, base + (npg + 2 * G) * e.nk, rho_im);
        if (wsum) wsum[0] = wsum[1] = h[n];
    });
}

}  // extern "C"
