/* Test-only: binds the C-ABI names to the oracle library (oracle/liboracle.so),
 * which exports the same entry points under the oracle_ prefix, so the C++
 * interface can be driven on the CPU.  Never part of the product build. */
#pragma once
#define alea_last_error oracle_last_error
#define alea_create oracle_create
#define alea_destroy oracle_destroy
#define alea_add_group oracle_add_group
#define alea_add_pade_jastrow oracle_add_pade_jastrow
#define alea_add_drummond_jastrow oracle_add_drummond_jastrow
#define alea_add_nosanow_jastrow oracle_add_nosanow_jastrow
#define alea_add_slater_plane_wave oracle_add_slater_plane_wave
#define alea_add_slater_gaussian oracle_add_slater_gaussian
#define alea_add_slater_twisted oracle_add_slater_twisted
#define alea_add_ewald oracle_add_ewald
#define alea_add_coulomb oracle_add_coulomb
#define alea_add_inverse_power oracle_add_inverse_power
#define alea_add_mcmillan_jastrow oracle_add_mcmillan_jastrow
#define alea_set_positions oracle_set_positions
#define alea_get_positions oracle_get_positions
#define alea_evaluate oracle_evaluate
#define alea_get_log_psi oracle_get_log_psi
#define alea_get_gradients oracle_get_gradients
#define alea_get_phase_gradients oracle_get_phase_gradients
#define alea_local_energy oracle_local_energy
#define alea_ratio oracle_ratio
#define alea_accept oracle_accept
#define alea_vmc_sweeps oracle_vmc_sweeps
#define alea_dmc_sweeps oracle_dmc_sweeps
#define alea_permute oracle_permute
#define alea_get_num_walkers oracle_get_num_walkers
#define alea_dmc_begin oracle_dmc_begin
#define alea_dmc_weights oracle_dmc_weights
#define alea_dmc_sums oracle_dmc_sums
#define alea_get_weights oracle_get_weights
#define alea_set_weights oracle_set_weights
#define alea_branch oracle_branch
#define alea_add_gofr oracle_add_gofr
#define alea_add_density oracle_add_density
#define alea_estimators_read_density oracle_estimators_read_density
#define alea_add_sofk oracle_add_sofk
#define alea_get_sofk_vectors oracle_get_sofk_vectors
#define alea_estimators_start_block oracle_estimators_start_block
#define alea_estimators_accumulate oracle_estimators_accumulate
#define alea_estimators_read oracle_estimators_read
#define alea_refresh_inverse oracle_refresh_inverse
#define alea_check_status oracle_check_status
