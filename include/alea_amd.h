/* alea_amd: MI355X-native batched VMC/DMC walker sweep — the C-ABI.
 *
 * One handle is one walker batch (the reference's WalkerSet with its
 * SimulationCell) plus its trial wavefunction and Hamiltonian, resident in
 * HBM.  Plain pointers and sizes only; every call returns ALEA_OK or an
 * error code, with the message in alea_last_error() (thread-local).  The
 * error codes map to the reference's exception types:
 *   ALEA_ERR_INVALID  std::invalid_argument (refused input)
 *   ALEA_ERR_RUNTIME  std::runtime_error   (e.g. a singular determinant,
 *                     slater_determinant.cpp:927-928)
 *   ALEA_ERR_DEVICE   a HIP runtime failure
 *
 * Host arrays use the reference's walker-fastest SoA layouts
 * (docs/architecture.md:141-156):
 *   positions, gradients, laplacians   [p*dim*nw + d*nw + w]
 *   proposed position, new_grad        [d*nw + w]
 *   per-walker scalars                 [w]
 * The device keeps walker-major records; conversion happens at these calls.
 *
 * Registration order: create, add_group (all groups), components, then
 * set_positions, evaluate.  Components cannot be added after positions are set.
 */
#ifndef ALEA_AMD_H
#define ALEA_AMD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define ALEA_OK 0
#define ALEA_ERR_INVALID 1
#define ALEA_ERR_RUNTIME 2
#define ALEA_ERR_DEVICE 3

/* Precision of the sweep. */
#define ALEA_FP64 0   /* plain: everything in FP64 */
#define ALEA_MIXED 1  /* determinant inverse, orbitals, ratios and rank-1 updates in FP32;
                         log-determinant, Jastrow, rebuilds, weights, energies in FP64 */

typedef struct alea_system alea_system;

const char* alea_last_error(void);

/* WalkerSet(nw, np, dim, cell) — src/population/walker_set.h:37;
 * SimulationCell(dim, vectors) — src/population/simulation_cell.h:66.
 * lattice: dim*dim primitive vectors, row-major, or NULL for an open cell.
 * n_walkers is the capacity (DMC populations may shrink below it). */
int alea_create(alea_system** out, int n_walkers, int n_particles, int dim,
                const double* lattice, int precision);
int alea_destroy(alea_system* s);

/* ParticleGroup — src/population/particle_group.h:17 (isotropic mass). */
int alea_add_group(alea_system* s, int first, int last, double charge, double inv_mass);

/* PadeJastrow(num_order, den_order, group_a, group_b, params) —
 * src/wavefunction/pade_jastrow.h:60; coeffs [a_1..a_N, b_1..b_M].
 * n_k > 0 turns on the image sum (PadeJastrow::sum_over_images,
 * pade_jastrow.h:69; PeriodicPairSum, periodic_pair_sum.h:45): the breakup's
 * k set (n1, n2, kx, ky), weights, quintic spline (LongRangeSpline::scaled,
 * 3 per knot), r_cut, 1/delta and u_inf. */
int alea_add_pade_jastrow(alea_system* s, int group_a, int group_b, int num_order, int den_order,
                          const double* coeffs, int n_k, const int* n1, const int* n2,
                          const double* kx, const double* ky, const double* weights,
                          int n_knots, const double* spline, double r_cut, double inv_delta,
                          double u_inf);

/* DrummondJastrow(C, gamma, group_a, group_b, params) —
 * src/wavefunction/drummond_jastrow.h:77; coeffs [alpha_0, alpha_2..alpha_N]. */
int alea_add_drummond_jastrow(alea_system* s, int group_a, int group_b, int C, double gamma,
                              double L, int n_coeffs, const double* coeffs);

/* NosanowJastrow(dim, first, sites, cell, {alpha}) —
 * src/wavefunction/nosanow_jastrow.h:55; sites [n*2], one per group particle. */
int alea_add_nosanow_jastrow(alea_system* s, int group, double alpha, const double* sites);

/* One determinant block [first, last):
 * SlaterDeterminant + PlaneWaveOrbitalSet — src/wavefunction/orbital_set.h:427
 *   (kvecs [n*2], is_sin [n]);
 * SlaterDeterminant + GaussianOrbitalSet — src/wavefunction/orbital_set.h:164
 *   (centers [n*2], alphas [n]);
 * ComplexSlaterDeterminant + TwistedPlaneWaveOrbitalSet —
 *   src/wavefunction/complex_orbital_set.h:91 (q = k + theta, [n*2]). */
int alea_add_slater_plane_wave(alea_system* s, int first, int last, const double* kvecs,
                               const unsigned char* is_sin);
int alea_add_slater_gaussian(alea_system* s, int first, int last, const double* centers,
                             const double* alphas);
int alea_add_slater_twisted(alea_system* s, int first, int last, const double* q);

/* EwaldTerm(groups, epsilon, cell, tol) — src/hamiltonian/ewald_term.h:70,
 * with the EwaldSetup of numerics::build_ewald (ewald_breakup.h:186):
 * k set, weights, long-range spline, r_cut, v_long_zero, sr_k0.  Charges come
 * from the groups. */
int alea_add_ewald(alea_system* s, int n_k, const int* n1, const int* n2, const double* kx,
                   const double* ky, const double* weights, int n_knots, const double* spline,
                   double r_cut, double inv_delta, double v_long_zero, double sr_k0);

/* CoulombTerm(group_a, group_b, epsilon) — src/hamiltonian/coulomb.h:28 (open cell). */
int alea_add_coulomb(alea_system* s, int group_a, int group_b, double epsilon);

/* InversePowerTerm(group_a, group_b, strength C, exponent p, cell) —
 * src/hamiltonian/inverse_power.h:56-60: sum of C / r^p over the pairs at the
 * minimum image, plus in a periodic cell the continuum tail outside the
 * Wigner-Seitz cell (inverse_power.cpp:43-83).  p = 3, C = 1: dipoles
 * perpendicular to the plane. */
int alea_add_inverse_power(alea_system* s, int group_a, int group_b, double strength,
                           double exponent);

/* McMillan pair Jastrow u(r) = -(b/r)^n at the minimum image (n = 5 is
 * McMillan's form for helium and the dipole gas).  No counterpart in the
 * reference: the dipole jobs of the serving stream name it. */
int alea_add_mcmillan_jastrow(alea_system* s, int group_a, int group_b, double b, int n);

/* WalkerSet::load_all (walker_set.cpp:51-70): positions are wrapped into the cell. */
int alea_set_positions(alea_system* s, const double* pos);
int alea_get_positions(alea_system* s, double* pos);

/* TrialWavefunction::evaluate_batch — src/wavefunction/trial_wavefunction.cpp:19. */
int alea_evaluate(alea_system* s);
/* WalkerSet::log_psi (walker_set.h:88): log|psi| per walker. */
int alea_get_log_psi(alea_system* s, double* out);
/* refresh_derivatives + WalkerSet::gradients/laplacians (and the phase planes
 * of a complex trial, walker_set.h:98-110). */
int alea_get_gradients(alea_system* s, double* grad, double* lap);
int alea_get_phase_gradients(alea_system* s, double* pgrad, double* plap);
/* refresh_derivatives + Hamiltonian::evaluate_batch — src/hamiltonian/hamiltonian.cpp:18;
 * e_im (may be NULL) is Im E_L of a complex trial. */
int alea_local_energy(alea_system* s, double* e_re, double* e_im);

/* Per-call path: ParticleBatch::propose_move_batch + TrialWavefunction::ratio_batch
 * (particle_batch.cpp:43-57, trial_wavefunction.cpp:46-78) for particle p at
 * new_pos [dim*nw]; outputs log|ratio|, its phase, and the new gradient and
 * Laplacian of p.  Then alea_accept commits the same proposal where mask[w] != 0
 * (TrialWavefunction::apply_move_batch + WalkerSet::apply_move, vmc.cpp:147-152). */
int alea_ratio(alea_system* s, int p, const double* new_pos, double* log_ratio, double* phase,
               double* new_grad, double* new_lap);
int alea_accept(alea_system* s, int p, const unsigned char* mask);
/* The same two calls over the walkers [w_begin, w_end) (w_end < 0: to nw,
 * the reference's npos): TrialWavefunction::ratio_batch(ws, p, w_begin,
 * w_end) and apply_move_batch(ws, p, accept, w_begin, w_end)
 * (trial_wavefunction.cpp:46-138, wavefunction_component.h:96-110).  Every
 * array keeps the batch's walker stride nw; only the range's entries are
 * read or written.  Nothing is allocated, and calls on disjoint ranges touch
 * disjoint memory, so DMC crowds can share one system. */
int alea_ratio_range(alea_system* s, int p, int w_begin, int w_end, const double* new_pos,
                     double* log_ratio, double* phase, double* new_grad, double* new_lap);
int alea_accept_range(alea_system* s, int p, int w_begin, int w_end, const unsigned char* mask);
/* alea_ratio_range plus the moved particle's new phase gradient and phase
 * Laplacian [dim][nw] (WalkerSet::new_phase_grad / new_phase_lap, which a
 * complex trial's ratio_batch fills, complex_slater_determinant.cpp:154-210;
 * zeros for a real trial).  Any output but log_ratio may be NULL. */
int alea_ratio_range_phase(alea_system* s, int p, int w_begin, int w_end, const double* new_pos, double* log_ratio,
                           double* phase, double* new_grad, double* new_lap, double* new_pgrad, double* new_plap);
/* log|psi| [nw] and the gradient, Laplacian, phase-gradient and
 * phase-Laplacian planes [np][dim][nw] of the walkers [w_begin, w_end) only
 * (WavefunctionComponent::refresh_derivatives over a crowd's range,
 * wavefunction_component.h:249-251), with the batch's walker stride: other
 * walkers' entries are neither read nor written.  Any output may be NULL.
 * The first call allocates device staging for the capacity; the calls share
 * it, so calls on one system are serialized by the caller. */
int alea_get_derivatives_range(alea_system* s, int w_begin, int w_end, double* log_psi, double* grad, double* lap,
                               double* pgrad, double* plap);
/* Particle p's rows of those planes, each [dim][nw] at the batch stride, for
 * the walkers [w_begin, w_end) (WavefunctionComponent::
 * refresh_particle_derivatives over a crowd's range,
 * wavefunction_component.h:245-248).  Any output may be NULL; calls share the
 * staging of alea_get_derivatives_range. */
int alea_get_particle_derivatives_range(alea_system* s, int p, int w_begin, int w_end, double* grad, double* lap,
                                        double* pgrad, double* plap);

/* Hamiltonian::delta_potential_particle (src/hamiltonian/hamiltonian.cpp:105-117),
 * for the walkers [w_begin, w_end): delta[w] = V(p at new_pos[.][w], wrapped
 * into the cell) - V(current positions), nothing committed.  new_pos is [2][nw]
 * and delta [nw] at the batch stride; only the range's entries are read or
 * written.  The Ewald part follows EwaldTerm::delta_potential_particle
 * (ewald_term.cpp:750-815): the moved particle's real-space pairs and the
 * reciprocal change against each walker's kept structure factor, which folds
 * in up to four particles moved since it last saw the walker and is
 * recomputed otherwise (sync_delta_cache, ewald_term.cpp:675-747); Coulomb and
 * inverse powers take the moved particle's pairs. */
int alea_delta_potential(alea_system* s, int p, int w_begin, int w_end, const double* new_pos, double* delta);

/* One free (optimizable) parameter of the trial wavefunction.  factor counts
 * the wavefunction factors in the order they were added (the reference's
 * component order, each determinant block counted as one factor); index is
 * the parameter within it: a Padé's coefficients a_1..a_N then b_1..b_M, a
 * Drummond's coefficients alpha_0, alpha_2, ... (as
 * alea_add_drummond_jastrow takes them) then L, a Nosanow's 0 (alpha), a
 * Gaussian block's 0 (the shared width, GaussianOrbitalSet's one parameter,
 * orbital_set.h:228-247; a block of several widths, and the plane-wave and
 * twisted blocks, have none).  For a Padé summed over images, the breakup of du/dp minus its
 * limit on the factor's k vectors (PeriodicPairSum's param_minus_limit,
 * periodic_pair_sum.cpp:16-64): weights [n_k], spline [3 n_knots] in the
 * knot form of alea_add_pade_jastrow, r_cut, 1/delta and the limit;
 * otherwise the pointers may be NULL. */
typedef struct alea_param {
    int factor, index;
    const double* weights;
    const double* spline;
    int n_knots;
    double r_cut, inv_delta, limit;
} alea_param;

/* TrialWavefunction::compute_parameter_derivatives
 * (src/wavefunction/trial_wavefunction.cpp:150-160), for the n parameters
 * listed: derivs[i * nw + w] = d log|psi| / d p_i of walker w, after
 * evaluate or sweeps.  Per factor as PadeJastrow (pade_jastrow.cpp:641-700),
 * DrummondJastrow (drummond_jastrow.cpp:201-320), NosanowJastrow
 * (nosanow_jastrow.cpp:176-190) and SlaterDeterminant's Gaussian width
 * (slater_determinant.cpp:98-157) compute it. */
int alea_parameter_derivatives(alea_system* s, const alea_param* params, int n, double* derivs);

/* Hamiltonian::local_energy_parameter_derivatives
 * (src/hamiltonian/hamiltonian.cpp:189-196) fed by
 * TrialWavefunction::compute_parameter_gradient_derivatives
 * (trial_wavefunction.cpp:162-185), in one call: de_re[i * nw + w] =
 * dE_L / d p_i of walker w, and for a complex trial de_im its imaginary part
 * (complex_local_energy_parameter_derivatives, kinetic.cpp:212-250; zeros
 * for a real trial; de_im may be NULL).  Only the kinetic term depends on
 * the parameters.  The per-particle gradient derivatives stay on the device. */
int alea_local_energy_parameter_derivatives(alea_system* s, const alea_param* params, int n, double* de_re,
                                            double* de_im);

/* Fused sweeps: VMC::sweep / sweep_walkers (vmc.cpp:69-155) with symmetric
 * Metropolis proposals of width sqrt(tau/m).  tape: [sweep][np][dim+1][nw]
 * random numbers in draw_sweep_randoms order (sweep_blocking.h:39-48), or NULL
 * for the per-walker device streams seeded by alea_seed.  With n_accepted
 * NULL the call returns once the sweeps are queued (see alea_query). */
int alea_vmc_sweeps(alea_system* s, int n_sweeps, double tau, const double* tape,
                    long long* n_accepted);
/* VMC::sweep_walkers with VMCConfig::use_drift (vmc.cpp:98-139): each
 * particle's gradient of log|psi| refreshed, the proposal
 * r + (tau/m) g + sqrt(tau/m) xi, acceptance min(1, |psi'/psi|^2 T(r|r')/T(r'|r))
 * with the Langevin Green's function (numerics/green_function.h:20-26).  No
 * node gate and no rr accounting.  Arguments as alea_vmc_sweeps. */
int alea_vmc_drift_sweeps(alea_system* s, int n_sweeps, double tau, const double* tape,
                          long long* n_accepted);
/* alea_vmc_sweeps on the device streams (tape NULL, no acceptance read-back)
 * that allocates nothing: the random numbers of each sweep are staged in a
 * fixed 32 MiB device array of the library (one per GPU, held until the
 * process ends) instead of a buffer allocated on first use, a chunk of
 * walkers at a time (4 Mi doubles / (3 np) walkers per chunk).  Results are
 * bit-identical to alea_vmc_sweeps.  Returns once the launches are queued on
 * the system's stream.  Calls on one GPU share the array: calls that overlap
 * in time on one GPU (systems on different streams) race and must be
 * ordered by the caller.  The system must be evaluated and seeded.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops). */
int alea_vmc_sweeps_static(alea_system* s, int n_sweeps, double tau);

/* DMC::sweep_walkers (dmc.cpp:203-337): Umrigar-scaled drift, node gate for a
 * real trial, Metropolis-Hastings with the drift Green's function.
 * rr_proposed and rr_accepted [nw] are accumulated into the given host arrays;
 * with both NULL they accumulate on the device (reset by alea_dmc_begin).
 * tape rows have stride tape_stride >= nw (the reference strides by the
 * population capacity). */
int alea_dmc_sweeps(alea_system* s, int n_sweeps, double tau, const double* tape, int tape_stride,
                    double* rr_proposed, double* rr_accepted, long long* n_accepted);

/* The periodic rebuild of SlaterDeterminant::commit_walker
 * (slater_determinant.cpp:788-790): the reference rebuilds one walker's
 * inverse inside the move once its block has taken `updates` Sherman-Morrison
 * updates (64).  Here, after every sweep (and after every alea_accept), one
 * batched kernel rebuilds every block whose update count reached `updates`
 * (default 64, both precisions), so a block carries at most updates + Nb - 1
 * updates.  The numerical triggers (tiny or non-finite ratio, the
 * cancellation guard) still rebuild inside the sweep. */
int alea_set_refresh_updates(alea_system* s, int updates);

/* Which kernel recomputes the inverses (the batched rebuild of
 * SlaterDeterminant::rebuild_inverse, slater_determinant.cpp:864-1060, the
 * same pivoted Gauss-Jordan either way): 1 (the default) runs blocks of 33 to
 * 64 orbitals on the FP64 matrix cores, as a Gauss-Jordan exchange blocked by
 * 16 columns whose rank-16 updates are MFMA products; 0 keeps every block on
 * the vector units.  Other block sizes always take the vector units. */
int alea_set_refresh_kernel(alea_system* s, int kind);

/* sweeps > 0 also rebuilds every block of every walker after every `sweeps`
 * sweeps (0, the default: only the update-count rule above). */
int alea_set_refresh_interval(alea_system* s, int sweeps);

/* n_sweeps VMC sweeps (as alea_vmc_sweeps on the device streams, tape NULL,
 * no acceptance read-back) of n_systems systems as fused launches: per sweep
 * each system's random numbers, then ONE sweep kernel over the walkers of all
 * of them (each workgroup runs one walker of one system with that system's
 * model, state and step taus[i]), then each system's refresh when it is due.
 * The systems must share the precision, the complex form and the register
 * tile (same block sizes), and may differ in everything else (density, twist,
 * factors, walker counts).  Every walker's result is bit-identical to
 * alea_vmc_sweeps on its own system.  The launches go on systems[0]'s stream
 * after the work already queued on every member's stream; each member's
 * later work waits for them.  Returns once they are queued.  At most 8
 * systems, none twice.  No counterpart in the reference: the serving loop's
 * size-bucketed launch. */
int alea_vmc_sweeps_fused(alea_system* const* systems, int n_systems, int n_sweeps, const double* taus);

/* on != 0: alea_vmc_sweeps / alea_dmc_sweeps on the device streams (tape
 * NULL) replay one captured hipGraph per block of n_sweeps (draw, sweep and
 * the periodic refresh of every sweep), re-captured when the sweep count,
 * step, walker count, refresh phase or any device buffer changes.  Results
 * are bit-identical to the eager launches; alea_last_sweep_ms then times the
 * whole block.  No counterpart in the reference: the serving loop's launch
 * path for repeated blocks. */
int alea_set_sweep_graphs(alea_system* s, int on);

/* Waits for the system's stream, then fails with ALEA_ERR_RUNTIME and the
 * reference's message (slater_determinant.cpp:927-928) if any walker's
 * determinant became singular or non-finite in a rebuild since evaluate.  The
 * asynchronous paths (sweeps with n_accepted NULL, fused sweeps, graph
 * replays) leave this check to the caller; alea_local_energy and the
 * synchronous sweeps make it themselves. */
int alea_check_status(alea_system* s);

/* Waits for the system's stream (hipStreamSynchronize). */
int alea_synchronize(alea_system* s);
/* Non-blocking: *done = 1 when all work queued on the system's stream has
 * finished, else 0 (hipStreamQuery).  Every system has its own non-blocking
 * stream, so work of different systems overlaps on the device; with
 * n_accepted NULL, alea_vmc_sweeps only queues its launches.  No counterpart
 * in the reference: the serving loop's completion poll. */
int alea_query(alea_system* s, int* done);
/* Device time of the sweep kernel of the last sweep of the last
 * alea_vmc_sweeps / alea_dmc_sweeps call (or, for a replayed sweep graph, of
 * the whole block), measured with HIP events on the system's stream (no
 * counterpart in the reference: the measurement hook of the benchmark). */
int alea_last_sweep_ms(alea_system* s, double* ms);
/* The sweep-kernel time of every sweep of the last eager call, ms[i] for
 * sweep i < min(capacity, *n); *n receives the call's sweep count (ms may be
 * NULL with capacity 0 to ask for the count alone). */
int alea_sweep_kernel_ms(alea_system* s, double* ms, int capacity, int* n);

/* Rebuild triggers of accepted moves since the last reset, by kind
 * (slater_determinant.cpp:615-617,788-790,825-826): counts[0] ratio below
 * 1e-5 or not finite, [1] rebuilt inverses whose condition estimate is above
 * 1e8 (latched: no further rebuild until the next periodic one), [2] the
 * cancellation guard, [3] an update whose size bound max|r| max|f| is not
 * finite or could carry the inverse out of range.  Kinds 0, 2 and 3 rebuild
 * the inverse inside the sweep.  per_walker [nw]
 * (may be NULL) receives each walker slot's total (a permutation does not
 * move the counts).  reset != 0 clears them.  Synchronous.  No counterpart in
 * the reference, which does not count its rebuilds: the measurement hook for
 * the rebuild rate. */
int alea_rebuild_counts(alea_system* s, long long* counts, int* per_walker, int reset);

/* SlaterDeterminant::rebuild_inverse (slater_determinant.cpp:905-993, declared
 * at slater_determinant.h:142) for every determinant block of the walkers
 * [w_begin, w_end) (w_end < 0: to nw, the reference's npos): D from the
 * current positions, A = D^-1 by pivoted Gauss-Jordan in FP64, stored in the
 * determinant precision with log|det|, the row sums and the condition
 * estimate; update counts to zero.  Asynchronous on the system's stream; a
 * singular block sets the walker's status (reported by the next synchronous
 * sweep or alea_check_status). */
int alea_refresh_inverse(alea_system* s, int w_begin, int w_end);

/* The cancellation guard of SlaterDeterminant::commit_walker
 * (slater_determinant.cpp:788-790): an accepted move rebuilds the inverse
 * when |R| <= factor * eps * sum_m |A_mj v'_m|, eps the unit roundoff of the
 * precision that stores A.  The reference's factor is 1e4 (FP64), the
 * default of ALEA_FP64.  ALEA_MIXED defaults to 1: its inverse is rounded to
 * FP32 again by any rebuild, so below that floor a rebuild buys nothing, and
 * the guard fires only where the ratio has no correct digit.  0 turns the
 * guard off. */
int alea_set_cancellation_factor(alea_system* s, double factor);

/* Per-walker xoshiro256++ streams: walker w gets stream_seed(seed, stream, w)
 * (random.h:52-56) expanded by splitmix64 (xoshiro_random.h:29-32). */
int alea_seed(alea_system* s, uint64_t seed, uint64_t stream);

/* The numbers the next n_sweeps device-stream sweeps would draw, into out
 * [sweep][np][dim+1][nw] (draw_sweep_randoms order, sweep_blocking.h:39-48,
 * over each walker's own stream: per particle two Gaussians then a uniform),
 * advancing the streams exactly as those sweeps would.  A run recorded this
 * way replays on any tape-driven path (the oracle's, say).  Synchronous. */
int alea_draw_tape(alea_system* s, int n_sweeps, double* out);

/* WalkerSet::apply_permutation + TrialWavefunction::permute_caches
 * (walker_set.cpp:109-178): new walker j takes old walker src[j]'s state;
 * new_nw <= capacity. */
int alea_permute(alea_system* s, int new_nw, const int* src);
/* WalkerSet::num_walkers (walker_set.h): the current population. */
int alea_get_num_walkers(alea_system* s, int* nw);

/* ── Forward walking's snapshots (ForwardWalking, forward_walking.h:60-147).
 * The ring of ForwardWalking::Snapshot::positions (forward_walking.h:
 * 116-121), held on the population's device: alea_snapshots_record is
 * ForwardWalking::record (forward_walking.cpp:99-108) without the ancestor
 * indices, which the caller keeps and carries through each branch's source
 * map (ForwardWalking::branch, forward_walking.cpp:175-182, integers on the
 * host).  alea_snapshots_load is ConfigurationBatch::load of the kept
 * ancestors (forward_walking.cpp:148-150): the batch's walkers become
 * configurations kept[0..n_kept) of the slot with the given weights (the
 * descendant weights D[a]), and the batch is evaluated, so its local energy
 * and estimators measure them (with no wavefunction factors in the batch,
 * E_L is Hamiltonian::evaluate_potential).  The batch must have the
 * population's particle count and capacity >= n_kept.  Destroy the ring
 * before its population. */
typedef struct alea_snapshots alea_snapshots;
int alea_snapshots_create(alea_system* population, int n_slots, alea_snapshots** out);
int alea_snapshots_destroy(alea_snapshots* r);
int alea_snapshots_record(alea_snapshots* r, int slot);
int alea_snapshots_load(alea_snapshots* r, int slot, alea_system* batch, const int* kept, const double* weights,
                        int n_kept);

/* ── DMC population step (DMC::advance_measure, dmc.cpp:339-419, bare
 * reweighting; DMC::run_one_step, dmc.cpp:693-739).  A step is:
 *   alea_dmc_begin; n_substeps x { alea_dmc_sweeps(rr NULL); alea_dmc_weights };
 *   alea_dmc_sums; alea_branch; PopulationControl::update on the host. */

/* Step set-up (dmc.cpp:349-354): rr sums to zero, weights to one, E_old to
 * the current local energies (evaluated here if not current). */
int alea_dmc_begin(alea_system* s);
/* After a drift-diffusion sweep: evaluate E_L (hamiltonian.cpp:18) and apply
 * weight *= PopulationControl::branch_weight(E_new, E_old)
 * (population_control.cpp:58-67) with the controller's state e_trial, e_ref,
 * e_cut, e_max = 1.5 e_cut and tau; then E_old = E_new (dmc.cpp:404-408). */
int alea_dmc_weights(alea_system* s, double e_trial, double e_ref, double e_cut, double e_max,
                     double tau);
/* (sum_w weight*E_L, sum_w weight) in walker order (dmc.cpp:412-418) and the
 * walker-sweeps tapered by the cutoff since alea_dmc_begin (any may be NULL). */
int alea_dmc_sums(alea_system* s, double* energy_sum, double* weight_sum, long long* n_cut);
/* WalkerSet::weights (walker_set.h:135) download / upload, [nw]. */
int alea_get_weights(alea_system* s, double* w);
int alea_set_weights(alea_system* s, const double* w);
/* WalkerSet::branch(rng, target, src) (walker_set.cpp:187-239): integer
 * multiplicities int(weight + xi) in walker order, population bounds
 * [target/5 + 1, 2 target + 1], the best walker on extinction; then
 * apply_permutation and weights reset to one.  rng_state is the branch
 * stream's xoshiro256++ state (4 words), advanced in place exactly as
 * XoshiroRandom::uniform draws it.  src_out (may be NULL) receives the
 * new_nw source indices; 2 target + 1 must not exceed the capacity. */
int alea_branch(alea_system* s, uint64_t* rng_state, int target, int* new_nw, int* src_out);

/* ── Estimators (src/estimators/estimator.h:48-64: start_block, accumulate,
 *    then the block's sums; stop_block's normalization and finalize run on
 *    the host) ─────────────────────────────────────────────────────────── */
/* PairCorrelation(groups, dim, rmax, n_bins, cell) —
 * src/estimators/pair_correlation.h:29-33: one function per unordered group
 * pair a <= b, pid = a G - a(a-1)/2 + (b - a), histogram [pid * n_bins + bin]
 * of r < rmax at bin int(r n_bins / rmax).  A periodic cell needs rmax at most
 * its Wigner-Seitz inradius; G(G+1)/2 * n_bins <= 4096. */
int alea_add_gofr(alea_system* s, double rmax, int n_bins);
/* StructureFactor(groups, cell, kmax) — src/estimators/structure_factor.h:52:
 * the half space (n1 > 0, or n1 = 0 and n2 > 0) of the reduced basis'
 * reciprocal lattice with |k| <= kmax, sorted by |k|; *n_k receives their
 * number (at most 1024).  Periodic cells only. */
int alea_add_sofk(alea_system* s, double kmax, int* n_k);
/* Density(groups, dim, rmax, n_bins, per_particle, cell) —
 * src/estimators/density.h:55: one n_bins x n_bins grid per group (or per
 * particle when per_particle != 0), C row-major [g * n_bins^2 + b0 * n_bins +
 * b1]; a periodic cell bins the fractional coordinates of the wrapped
 * position over the whole cell (rmax unused), an open cell the window
 * [-rmax, rmax)^2.  (G or N) * n_bins^2 <= 4096. */
int alea_add_density(alea_system* s, double rmax, int n_bins, int per_particle);
/* The structure factor's vectors kx, ky [n_k] and shell index [n_k]. */
int alea_get_sofk_vectors(alea_system* s, double* kx, double* ky, int* shell);
/* Estimator::start_block of every estimator: the running sums to zero. */
int alea_estimators_start_block(alea_system* s);
/* Estimator::accumulate(wset) (pair_correlation.cpp:74-105,
 * structure_factor.cpp:62-105, density.cpp:66-117): one measurement of the current walkers with
 * weight 1 each (use_weights = 0, VMC) or the walkers' weights (DMC).
 * Asynchronous on the system's stream; deterministic for given walkers. */
int alea_estimators_accumulate(alea_system* s, int use_weights);
/* The block's running sums (NULL skips): gofr [G(G+1)/2 * n_bins] sum of
 * weights per bin; sofk [G(G+1)/2 * n_k] sum of w Re(rho_a rho_b^*); rho_re,
 * rho_im [G * n_k] sum of w rho_g(k); wsum[2] the weight sum (for the pair
 * correlation's and the structure factor's normalization). */
int alea_estimators_read(alea_system* s, double* gofr, double* sofk, double* rho_re, double* rho_im,
                         double* wsum);
/* The density's running sums: density [(G or N) * n_bins^2] sum of weights per
 * cell (density.cpp:66-117), *wsum the weight sum. */
int alea_estimators_read_density(alea_system* s, double* density, double* wsum);

#ifdef __cplusplus
}
#endif
#endif
