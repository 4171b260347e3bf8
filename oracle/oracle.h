/* Oracle C-ABI: the FP64 CPU restatement of the batched VMC/DMC walker sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only as the checker.
 * The entry points mirror include/alea_amd.h one for one (prefix oracle_
 * instead of alea_), with the same argument meaning and layouts, so a test can
 * run one call sequence on the HIP library and on the oracle and compare.
 *
 * Layouts follow the reference's walker-fastest SoA:
 *   positions / gradients / laplacians  [p*dim*nw + d*nw + w]
 *   proposed position, new_grad          [d*nw + w]
 *   per-walker scalars                   [w]
 */
#ifndef ALEA_ORACLE_H
#define ALEA_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_system oracle_system;

const char* oracle_last_error(void);

/* RNG: xoshiro256++ and the ziggurat Gaussian (src/numerics/xoshiro_random.h) */
uint64_t oracle_stream_seed(uint64_t base, uint64_t stream, uint64_t index);
int oracle_xoshiro_state_u64(const uint64_t* state, int n, uint64_t* out);
int oracle_xoshiro_draws(uint64_t seed, int kind, int n, double* out);
int oracle_sweep_tape(uint64_t seed, int np, int dim, int nw, int n_sweeps, double* out);
int oracle_ziggurat_tables(uint64_t* k, double* w, double* f, double* r);
int oracle_initial_positions(uint64_t seed, int np, int nw, const double* lattice, double L,
                             const unsigned char* at_origin, double* out);
int oracle_mt19937_64(uint64_t seed, int n, uint64_t* out);
int oracle_add_gofr(oracle_system* s, double rmax, int n_bins);
int oracle_add_sofk(oracle_system* s, double kmax, int* n_k);
int oracle_get_sofk_vectors(oracle_system* s, double* kx, double* ky, int* shell);
int oracle_add_density(oracle_system* s, double rmax, int n_bins, int per_particle);
int oracle_estimators_read_density(oracle_system* s, double* density, double* wsum);
int oracle_estimators_start_block(oracle_system* s);
int oracle_estimators_accumulate(oracle_system* s, int use_weights);
int oracle_estimators_read(oracle_system* s, double* gofr, double* sofk, double* rho_re, double* rho_im,
                           double* wsum);

int oracle_create(oracle_system** out, int nw, int np, int dim, const double* lattice,
                  int precision);
int oracle_destroy(oracle_system* s);
int oracle_add_group(oracle_system* s, int first, int last, double charge, double inv_mass);
int oracle_add_pade_jastrow(oracle_system* s, int ga, int gb, int num_order, int den_order,
                            const double* coeffs, int n_k, const int* n1, const int* n2,
                            const double* kx, const double* ky, const double* weights,
                            int n_knots, const double* spline, double r_cut,
                            double inv_delta, double u_inf);
int oracle_add_drummond_jastrow(oracle_system* s, int ga, int gb, int C, double gamma,
                                double L, int n_coeffs, const double* coeffs);
int oracle_add_nosanow_jastrow(oracle_system* s, int group, double alpha,
                               const double* sites);
int oracle_add_slater_plane_wave(oracle_system* s, int first, int last, const double* kvecs,
                                 const unsigned char* is_sin);
int oracle_add_slater_gaussian(oracle_system* s, int first, int last, const double* centers,
                               const double* alphas);
int oracle_add_slater_twisted(oracle_system* s, int first, int last, const double* q);
int oracle_add_ewald(oracle_system* s, int n_k, const int* n1, const int* n2,
                     const double* kx, const double* ky, const double* weights, int n_knots,
                     const double* spline, double r_cut, double inv_delta,
                     double v_long_zero, double sr_k0);
int oracle_add_coulomb(oracle_system* s, int ga, int gb, double epsilon);
int oracle_add_inverse_power(oracle_system* s, int ga, int gb, double strength, double exponent);
int oracle_add_mcmillan_jastrow(oracle_system* s, int ga, int gb, double b, int n);

int oracle_set_positions(oracle_system* s, const double* pos);
int oracle_get_positions(oracle_system* s, double* pos);
int oracle_evaluate(oracle_system* s);
int oracle_get_log_psi(oracle_system* s, double* out);
int oracle_get_gradients(oracle_system* s, double* grad, double* lap);
int oracle_get_phase_gradients(oracle_system* s, double* pgrad, double* plap);
int oracle_local_energy(oracle_system* s, double* e_re, double* e_im);
int oracle_ratio(oracle_system* s, int p, const double* new_pos, double* log_ratio,
                 double* phase, double* new_grad, double* new_lap);
int oracle_accept(oracle_system* s, int p, const unsigned char* mask);
int oracle_vmc_sweeps(oracle_system* s, int n_sweeps, double tau, const double* tape,
                      long long* n_accepted);
int oracle_vmc_drift_sweeps(oracle_system* s, int n_sweeps, double tau, const double* tape,
                            long long* n_accepted);
int oracle_dmc_sweeps(oracle_system* s, int n_sweeps, double tau, const double* tape,
                      int tape_stride, double* rr_proposed, double* rr_accepted,
                      long long* n_accepted);
int oracle_permute(oracle_system* s, int new_nw, const int* src);
typedef struct oracle_snapshots oracle_snapshots;
int oracle_snapshots_create(oracle_system* population, int n_slots, oracle_snapshots** out);
int oracle_snapshots_destroy(oracle_snapshots* r);
int oracle_snapshots_record(oracle_snapshots* r, int slot);
int oracle_snapshots_load(oracle_snapshots* r, int slot, oracle_system* batch, const int* kept,
                          const double* weights, int n_kept);
int oracle_get_num_walkers(oracle_system* s, int* nw);
int oracle_device_tape(uint64_t seed, uint64_t stream, int np, int nw, int n_sweeps, double* out);
int oracle_ratio_range(oracle_system* s, int p, int w_begin, int w_end, const double* new_pos, double* log_ratio,
                       double* phase, double* new_grad, double* new_lap);
int oracle_accept_range(oracle_system* s, int p, int w_begin, int w_end, const unsigned char* mask);
int oracle_ratio_range_phase(oracle_system* s, int p, int w_begin, int w_end, const double* new_pos,
                             double* log_ratio, double* phase, double* new_grad, double* new_lap, double* new_pgrad,
                             double* new_plap);
int oracle_get_derivatives_range(oracle_system* s, int w_begin, int w_end, double* log_psi, double* grad, double* lap,
                                 double* pgrad, double* plap);
int oracle_get_particle_derivatives_range(oracle_system* s, int p, int w_begin, int w_end, double* grad, double* lap,
                                          double* pgrad, double* plap);
int oracle_delta_potential(oracle_system* s, int p, int w_begin, int w_end, const double* new_pos, double* delta);
typedef struct oracle_param {
    int factor, index;
    const double* weights;
    const double* spline;
    int n_knots;
    double r_cut, inv_delta, limit;
} oracle_param;
int oracle_parameter_derivatives(oracle_system* s, const oracle_param* params, int n, double* derivs);
int oracle_local_energy_parameter_derivatives(oracle_system* s, const oracle_param* params, int n,
                                              double* de_re, double* de_im);
int oracle_refresh_inverse(oracle_system* s, int w_begin, int w_end);
int oracle_check_status(oracle_system* s);
int oracle_dmc_begin(oracle_system* s);
int oracle_dmc_weights(oracle_system* s, double e_trial, double e_ref, double e_cut, double e_max,
                       double tau);
int oracle_dmc_sums(oracle_system* s, double* energy_sum, double* weight_sum, long long* n_cut);
int oracle_get_weights(oracle_system* s, double* w);
int oracle_set_weights(oracle_system* s, const double* w);
int oracle_branch(oracle_system* s, uint64_t* rng_state, int target, int* new_nw, int* src_out);

#ifdef __cplusplus
}
#endif
#endif
