// Device-side description of one system (cell, groups, trial wavefunction,
// Hamiltonian) and the per-walker state layout.  Shared by the host C-ABI
// (capi.hip) and the kernels (kernels.hip).
//
// Walker-major layout in HBM: every walker owns a contiguous record in each
// state array, so one wavefront streams its own walker's state with coalesced
// loads.  (The reference's CPU layout is walker-fastest SoA; the C-ABI
// converts at upload/download.)
#pragma once
#include <cstdint>

namespace alea_amd {

constexpr int kMaxParticles = 256;
constexpr int kMaxGroups = 4;
constexpr int kMaxGofrEntries = 4096;   // pair-correlation functions x bins
constexpr int kMaxDensityEntries = 4096;   // density: groups x n_bins^2
constexpr int kMaxSofkVectors = 1024;
constexpr int kMaxPair = 8;        // pair-Jastrow factors
constexpr int kMaxCoef = 12;       // Padé / Drummond coefficients
constexpr int kMaxKnots = 64;      // breakup spline knots
constexpr int kMaxKJ = 64;         // k-space Jastrow vectors
constexpr int kMaxKE = 192;        // Ewald vectors
constexpr int kMaxBlocks = 2;      // determinant blocks
constexpr int kMaxOrb = 128;       // orbitals per block
constexpr int kMaxCoulomb = 8;
constexpr int kMaxInversePower = 4;
constexpr int kRowPad = 16;        // determinant rows padded to a multiple of this

struct DevCell {
    int periodic = 0, ortho = 0;
    double box[2] = {0, 0}, inv_box[2] = {0, 0};
    double a[4] = {0, 0, 0, 0}, ainv[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
    double cand[18] = {};
    double ws = 0.0;
};

// Radial pair function u(r): Padé P/Q (pade_jastrow.cpp:26-56) optionally
// reduced to its image-sum short-range part (periodic_pair_sum.cpp:66-83),
// Drummond (drummond_jastrow.cpp:27-80), or McMillan u = -(b/r)^n (c[0] = b,
// num = n).
struct DevRadial {
    int kind = 0;   // 0 Padé, 1 Drummond, 2 McMillan
    int num = 0, den = 0, image = 0;
    double c[kMaxCoef] = {};
    double r_cut = 0, inv_delta = 0, u_inf = 0;
    int n_knots = 0;
    int C = 3;
    double L = 0, alpha1 = 0, gamma = 0;
    double spline[3 * kMaxKnots] = {};
    // the same spline per panel as the monomial c0 + c1 x + ... + c5 x^5 in
    // the panel coordinate x (filled at build), evaluated by Horner in the sweep
    double mono[6 * (kMaxKnots - 1)] = {};
};

struct DevPair {
    int ga = 0, gb = 0;
    DevRadial rad;
};

enum OrbitalKind { kPlaneWave = 0, kGaussian = 1, kTwisted = 2 };

struct DevModel {
    int nw = 0, np = 0, ng = 0;
    DevCell cell;
    int group_of[kMaxParticles] = {};
    double inv_mass[kMaxParticles] = {};
    double charge[kMaxParticles] = {};
    int g_first[kMaxGroups] = {}, g_last[kMaxGroups] = {};

    // pair Jastrow: factors and, per (group of p, group of partner), the list
    int n_pair = 0;
    DevPair pair[kMaxPair];
    int pair_list_n[kMaxGroups][kMaxGroups] = {};
    int pair_list[kMaxGroups][kMaxGroups][kMaxPair] = {};
    int spl_stride = 0;   // doubles per factor of the splines the sweep copies to LDS

    // merged k-space Jastrow: U = sum_k sum_{g<=h} w_gh(k) T_gh
    int nkj = 0;
    double kjx[kMaxKJ] = {}, kjy[kMaxKJ] = {};
    double kjw[kMaxGroups][kMaxGroups][kMaxKJ] = {};   // symmetric in (g,h)
    int kj_active[kMaxGroups] = {};                   // group has a density

    // Nosanow one-body: per particle alpha (0: none) and site
    double nos_alpha[kMaxParticles] = {};
    double nos_site[kMaxParticles][2] = {};
    int has_nosanow = 0;

    // Slater determinant blocks
    int n_blocks = 0, det_complex = 0;
    int blk_kind[kMaxBlocks] = {}, blk_first[kMaxBlocks] = {}, blk_n[kMaxBlocks] = {};
    // rows of A stored per block: blk_n rounded up to kRowPad; the extra rows
    // are zero and stay zero, so the update loops run without tail guards
    int blk_rows[kMaxBlocks] = {};
    double orb_k[kMaxBlocks][kMaxOrb][2] = {};   // plane wave k, twisted q, Gaussian centre
    int orb_sin[kMaxBlocks][kMaxOrb] = {};
    double orb_alpha[kMaxBlocks][kMaxOrb] = {};
    int block_of[kMaxParticles] = {};            // -1: in no block

    // Ewald (single Coulomb kernel)
    int ewald = 0, nke = 0;
    double kex[kMaxKE] = {}, key[kMaxKE] = {}, kew[kMaxKE] = {};
    double ew_spline[3 * kMaxKnots] = {};
    int ew_knots = 0;
    double ew_rcut = 0, ew_inv_delta = 0, ew_const = 0, ew_inv_eps = 1.0;

    // open-cell Coulomb pairs
    int n_coul = 0;
    int coul_ga[kMaxCoulomb] = {}, coul_gb[kMaxCoulomb] = {};
    double coul_coeff[kMaxCoulomb] = {};

    // inverse-power pair terms C / r^p at the minimum image; ipow_int = p
    // when p is an integer <= 16 (repeated product), else 0 (pow); the
    // continuum tails of all terms summed in ipow_tail
    int n_ipow = 0;
    int ipow_ga[kMaxInversePower] = {}, ipow_gb[kMaxInversePower] = {}, ipow_int[kMaxInversePower] = {};
    double ipow_c[kMaxInversePower] = {}, ipow_p[kMaxInversePower] = {};
    double ipow_tail = 0.0;

    // The cancellation guard of an accepted move: rebuild when |R| <=
    // cancel_factor * eps * sum_m |A_mj v'_m|, eps the unit roundoff of the
    // determinant precision (the reference's 1e4 in FP64,
    // slater_determinant.cpp:788-790).
    double cancel_factor = 1e4;
};

// Per-walker state arrays (device pointers).  Strides are per walker.
struct DevState {
    double* pos = nullptr;        // [w][np][2]
    double* rho = nullptr;        // [w][ng][nkj][2]  k-space Jastrow densities
    double* logs = nullptr;       // [w][4]: 0 Jastrow+one-body, 1.. determinant blocks
    void* inv = nullptr;          // [w][blocks][Nb*Nb] (T, complex: 2 T per entry)
    double* cond = nullptr;       // [w][blocks]
    int* updates = nullptr;       // [w][blocks]
    double* work = nullptr;       // [w][2*Nb*Nb*cplx] Gauss-Jordan workspace (FP64)
    uint64_t* rng = nullptr;      // [w][4]
    double* e_loc = nullptr;      // [w]
    double* e_loc_im = nullptr;   // [w]
    double* rr = nullptr;         // [w][2] DMC rr_proposed, rr_accepted
    double* grad_out = nullptr;   // optional [np][2][w] SoA outputs of the measure kernel
    double* lap_out = nullptr;
    double* pgrad_out = nullptr;
    double* plap_out = nullptr;
    int* status = nullptr;        // [w] nonzero: singular determinant
    long long* n_accepted = nullptr;   // [w]
    int* rebuilds = nullptr;      // [w][4] in-sweep determinant rebuilds by trigger (RebuildCause)
    // Ewald structure factor rho(k) = Sum_i q_i e^{ik.r_i} of each walker,
    // with the positions it was computed from and the increments folded in
    // since its last full computation (-1: none yet); see ewald_rho
    double* ew_rho = nullptr;     // [w][nke][2]
    double* ew_snap = nullptr;    // [w][np][2]
    int* ew_inc = nullptr;        // [w]
    int inv_stride = 0;           // T elements per walker in inv
    int work_stride = 0;          // doubles per walker in work
    int nb_max = 0;
};

// One free parameter of alea_parameter_derivatives on the device: a pair
// factor's coefficient (pair >= 0) or a Nosanow factor's alpha (group >= 0),
// with the image-sum tables of its derivative for a Padé summed over images
// (weights at wofs in the weights buffer, the knot-form spline at sofs).
struct DevParam {
    int pair = -1, group = -1, block = -1, index = 0;
    int n_knots = 0, wofs = 0, sofs = 0;
    double r_cut = 0, inv_delta = 0, limit = 0;
    double da1_dl = 0;   // Drummond L: d alpha_1 / d L
    double alpha = 0;    // Gaussian block: the shared width
};

// Why an accepted move rebuilt a determinant inside the sweep
// (slater_determinant.cpp:615-617,788-790,825-826), and the latched fresh
// inverses: the index into DevState::rebuilds.
enum RebuildCause {
    kRebuildTinyRatio = 0,     // |R| < 1e-5 or R not finite
    kLatchedFresh = 1,         // a rebuilt inverse whose condition estimate is above 1e8 (no rebuild)
    kRebuildCancellation = 2,  // |R| within the rounding bound of its terms
    kRebuildUpdateBound = 3,   // the update's size bound max|r| max|f| not finite or out of range
    kRebuildCauses = 4
};

// Per-call (parity path) buffers, walker-fastest like the reference.
struct DevCall {
    const double* new_pos = nullptr;   // [2][nw]
    const unsigned char* mask = nullptr;
    double* log_ratio = nullptr;       // [nw]
    double* phase = nullptr;           // [nw]
    double* new_grad = nullptr;        // [2][nw]
    double* new_lap = nullptr;
    double* new_pgrad = nullptr;
    double* new_plap = nullptr;
};

// The move a sweep makes: Metropolis VMC (vmc.cpp:109-135 without drift),
// drift-diffusion DMC with the Umrigar-scaled drift and node gate
// (dmc.cpp:203-337), or VMC with the bare Langevin drift and its
// Metropolis-Hastings correction (vmc.cpp:98-129 with use_drift).
enum SweepKind : int { kSweepVmc = 0, kSweepDmc = 1, kSweepLangevin = 2 };

struct SweepParams {
    int n_sweeps = 1;
    double tau = 0.0;
    const double* tape = nullptr;   // [sweep][np][3][stride] or null (device RNG)
    int tape_stride = 0;
    int dmc = kSweepVmc;            // a SweepKind
};

// One system of a fused VMC sweep launch: its model, state and sweep
// parameters, and the range of the launch's workgroups that are its walkers.
struct FusedJob {
    const DevModel* m = nullptr;
    DevState s;
    SweepParams p;
    int first = 0, nw = 0;
};
constexpr int kMaxFused = 8;    // systems in one fused launch
// The job table travels as the kernel argument (invariant kernel-argument
// memory), about 1.9 KB.
struct FusedArgs {
    FusedJob job[kMaxFused];
    int n = 0;
};

}  // namespace alea_amd
