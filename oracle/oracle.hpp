// CPU FP64 oracle for the DISORT flux path -- TEST INFRASTRUCTURE ONLY.
//
// This directory is a from-scratch restatement of the DisORT++ algorithm
// (NewStrangeWorlds/DisORT, GPL-3.0; no reference text is copied).  It exists
// to check the GPU library: only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and never as the thing measured
// or shipped.  Every function cites the reference file:line whose behaviour
// it restates.
//
// Parity status of each function (what pins it):
//   planck_band / planck_band_deriv : bit-compared against the reference
//        src/Planck.cpp compiled into oracle/_ref (tests/test_oracle_numerics.py)
//        plus golden vectors in tests/golden/planck.json.
//   legendre_flat   : closed forms P0..P2 (tests/test_utils.cpp:150-210) and
//        the reference src/Legendre.cpp compiled into oracle/_ref.
//   gauss_quadrature: numpy.polynomial.legendre.leggauss.
//   band_factor / band_solve : brute force dense solve on small systems and
//        the reference src/BandSolver.hpp compiled into oracle/_ref
//        (same pivot sequence).
//   sym_eigen (Jacobi), cholesky, lu_solve : numpy.linalg on random SPD /
//        general matrices; invariants (orthogonality, reconstruction).
//   flux_solve      : DisORT++ self-regression values (13 digits,
//        tests/test_solver.cpp:5238-5434), C-DISORT disotest tables
//        (9a/9b, 6b/6c, thermal T1-T4), closed forms (Beer's law,
//        transparent medium), energy conservation at omega=1.
//   general_solve   : C-DISORT intensity tables quoted by the reference's
//        tests (IntensityValidation 1a-3b, disotest 7b/7c/9c/10a, spherical),
//        Nakajima-Tanaka corrected tables (IntensityCorrectionHG3a/3b), RPV
//        tables (RpvBRDF1, TransparentBrdfRpv, MultiLayerBrdfRpv) -- 23 cases
//        in tests/golden/intensity_cases.json; m=0 fluxes equal flux_solve's,
//        azimuthal symmetry I(phi) = I(360-phi), the azimuth mean equals the
//        m=0 term, the accuracy_fourier_series truncation, index_from_bottom.
//   rpv_reflectivity: reciprocity rho(mu1,mu2,phi) = rho(mu2,mu1,phi), and
//        the transparent-medium closed form (mu0 F0/pi rho at the top).
//   Strength of these pins: tests/test_oracle_pins.py plants a 1e-6 relative
//   error in each stage of flux_solve and general_solve (set_perturbation)
//   and checks that a fixture, asserted at the precision its literals carry,
//   fails.
#pragma once
#include <cstdint>

// The oracle is compiled twice: in FP64 (namespace oracle) and in x87
// extended precision (namespace oracle_ld, 64-bit significand) to serve as a
// higher-precision truth for judging FP64 rounding differences.
#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif
#ifndef ORACLE_NS
#define ORACLE_NS oracle
#endif
#include <string>
#include <vector>

namespace ORACLE_NS {
using Real = ORACLE_REAL;

// ---- mutation hooks (tests of the pins only) ------------------------------
// With a stage selected by set_perturbation, the oracle multiplies that
// stage's intermediate values by 1 + eps * w_i, w_i in {-1, -1/3, 1/3, 1}
// cycling with the element index (non-uniform, so that no gauge freedom such
// as the scale of an eigenvector absorbs it).  Stage 0 (the default) leaves
// every value untouched.  The flux-only solver's own stages are numbered
// 32 + stage (the shared numerics -- Planck, Chapman, RPV -- keep their
// number).  The pin tests show that each stage, perturbed by 1e-6, fails at
// least one fixture.
enum PerturbStage {
  PS_NONE = 0, PS_SCALED_COEF = 1, PS_SCALED_SSA = 2, PS_EIGVAL = 3, PS_EIGVEC = 4, PS_BEAM = 5,
  PS_THERMAL = 6, PS_PLANCK = 7, PS_CONSTANTS = 8, PS_QUAD = 9, PS_LEGENDRE = 10, PS_USER_EIGVEC = 11,
  PS_NT = 12, PS_BRDF = 13, PS_CHAPMAN = 14, PS_FLUX = 15, PS_BURAS = 16, PS_JACOBIAN = 17,
  PS_COUNT = 18
};
void set_perturbation(int stage, double eps);
void perturb(int stage, std::vector<Real>& v);
Real perturb1(int stage, Real x);

// ---- numerics (restating src/Planck.cpp, src/Legendre.cpp, ...) ----------
Real planck_band(Real wnumlo, Real wnumhi, Real temp);        // planckFunction2
Real planck_band_exact(Real wnumlo, Real wnumhi, Real temp);  // the same, never perturbed
Real planck_band_deriv(Real wnumlo, Real wnumhi, Real temp);  // planckFunctionDeriv2
void legendre_flat(int nmu, int m, int maxmu, int twonm1, const Real* mu, Real* ylm);
void gauss_quadrature(int nn, Real* cmu /*[2nn]*/, Real* cwt /*[2nn]*/);
Real chapman(int lc, Real tau_frac, const Real* level_altitudes,
               const Real* dtau_lyr, int num_layers, Real r, Real mu0);
Real chapman_exact(int lc, Real tau_frac, const Real* level_altitudes,
               const Real* dtau_lyr, int num_layers, Real r, Real mu0);

// LINPACK-semantics band LU (column-major, lda rows, 1-based ipvt).
void band_factor(Real* abd, int lda, int n, int ml, int mu, int* ipvt);
void band_solve(const Real* abd, int lda, int n, int ml, int mu, const int* ipvt, Real* b);

// Dense helpers, column-major n x n.
bool cholesky_lower(int n, const Real* a, Real* l);  // a = l l^T; false if not SPD
void sym_eigen_jacobi(int n, const Real* a, Real* evals, Real* evecs); // ascending
// Partial-pivot LU of a (in place) + solve of one right-hand side.
struct DenseLU {
  int n = 0;
  std::vector<Real> lu;  // column-major
  std::vector<int> piv;
  void factor(int n, const Real* a);
  void solve(Real* b) const;
  Real frob_norm_of_input = 0.0;
};

// ---- temperature Jacobians (src/FluxSolver.cpp:786-831, :1294-1497;
// src/DisortSolver.cpp:2130-2377), shared by both solvers -----------------
// Derivatives of one layer's thermal particular solution (reordered like
// z0/z1: dz0/dz1 [2][N], e = 0 top level, 1 bottom level) and of its Planck
// slope/intercept ([2]) from the LU of [I - C].
void thermal_source_derivs(const DenseLU& lu, int N, Real omm, Real dtau_s, Real tau_top, const Real* cmu, Real* dz0,
                           Real* dz1, Real* dslope, Real* dicpt);
struct ThermalJacobianInputs {
  int N = 0, L = 0, ncut = 0;
  bool lyrcut = false, diffusion_bc = false;
  const Real *cmu = nullptr, *cwt = nullptr;      // quadrature [N], [NN]
  const Real *taus = nullptr, *dtaus = nullptr;   // scaled [L+1], [L]
  const Real *kk = nullptr, *gc = nullptr;        // eigenvalues [L][N], eigenvectors [L][N x N] column-major
  const Real* band = nullptr;                     // the forward solve's band LU and pivots
  const int* ipvt = nullptr;
  const Real *dz0 = nullptr, *dz1 = nullptr;      // [L][2][N]
  const Real *dslope = nullptr, *dicpt = nullptr; // [L][2]
  const Real* ssa = nullptr;                      // unscaled [L]
  const Real* refl = nullptr;                     // [NN][NN] m = 0 surface reflectivity at quadrature angles
  const Real* emis = nullptr;                     // [NN] surface emissivity
  int nrows = 0;                                  // output levels
  const int* row_layer = nullptr;
  const Real* row_tau = nullptr;                  // scaled optical depth of each output level
  Real wn_lo = 0, wn_hi = 0, temp_bot = 0;
  const Real* temper = nullptr;                   // [L+1]
  int perturb_stage = 0;
};
// jac[f] = [nrows][L+2], f: flux_up, flux_down, mean_intensity,
// mean_intensity_down, mean_intensity_up, flux_tau_divergence.
void thermal_jacobian(const ThermalJacobianInputs& in, std::vector<Real> jac[6]);
void reverse_jacobian(std::vector<Real> jac[6], int nrows, int L);

// ---- flux solver (restating src/FluxSolver.cpp DisortFluxSolver<N>) -------
struct FluxConfig {
  int nlyr = 0, nstr = 0, nmom = 0;           // pmom rows hold max(nmom,nstr)+1 entries
  bool thermal = false, spherical = false, deltam_plus = false;
  bool diffusion_bc = false, index_from_bottom = false, jacobian = false;
  Real beam_flux = 0, beam_mu = 0.5, iso_top = 0, iso_bot = 0;
  Real temp_top = 0, temp_bot = 0, emis_top = 1, albedo = 0;
  Real wn_lo = 0, wn_hi = 0, radius = 0;
  std::vector<Real> dtau, ssa, temper, alt;
  std::vector<Real> pmom;                   // [nlyr][pstride]
  int pstride() const { return (nmom > nstr ? nmom : nstr) + 1; }
  Real pm(int k, int lc) const { return pmom[(size_t)lc * pstride() + k]; }
};

struct FluxOutput {
  // flux_direct_beam, flux_down, flux_up, flux_tau_divergence, mean_intensity,
  // mean_intensity_down, mean_intensity_up, mean_intensity_direct_beam
  std::vector<Real> f[8];
  // Temperature Jacobians [6][(nlyr+1)*(nlyr+2)] when requested.
  std::vector<Real> jac[6];
};

// Throws std::invalid_argument / std::runtime_error with the reference's messages.
void validate_flux_config(const FluxConfig& c);
FluxOutput flux_solve(const FluxConfig& c);

// ---- general solver (restating src/DisortSolver.cpp DisortSolver) ---------
// RPV parameters in the order rho0, k, theta, sigma, t1, t2, scale
// (src/DisortConfig.hpp RpvBrdfSpec).
// BRDF kinds; the parameter block rpv[7] holds the kind's parameters:
// RPV {rho0, k, theta, sigma, t1, t2, scale}, Hapke {b0, hh, w},
// Cox-Munk {u10, refractive index, shadowing 0/1}, Ross-Li {iso, vol, geo}.
enum { BRDF_NONE = 0, BRDF_RPV = 1, BRDF_HAPKE = 2, BRDF_COXMUNK = 3, BRDF_ROSSLI = 4 };
struct GeneralConfig {
  int nlyr = 0, nstr = 0, nmom = 0;
  bool thermal = false, spherical = false, deltam_plus = false, diffusion_bc = false;
  bool use_user_tau = false, use_user_mu = false, only_fluxes = false;
  bool lambertian = true, corr_buras = false, corr_nakajima = false, index_from_bottom = false;
  // special_bc: ibcnd = SPECIAL_BC (albedo and transmissivity of the medium);
  // no_cutoff: no absorption cut-off (the base solves of SPECIAL_BC).
  bool special_bc = false, no_cutoff = false, jacobian = false;
  int brdf = BRDF_NONE;
  Real rpv[7] = {0, 0, 0, 0, 0, 0, 1};
  Real beam_flux = 0, beam_mu = 0.5, beam_phi = 0, iso_top = 0, iso_bot = 0;
  Real temp_top = 0, temp_bot = 0, emis_top = 1, albedo = 0;
  Real wn_lo = 0, wn_hi = 0, radius = 0, accur = 0;
  std::vector<Real> dtau, ssa, temper, alt, pmom;  // pmom [nlyr][pstride]
  std::vector<Real> tau_user, mu_user, phi_user;
  // Tabulated phase function (num_phase_func_angles, mu_phase_function,
  // phase_function[layer][angle]) of the Buras-Emde correction.
  int nphase = 0;
  std::vector<Real> mu_phase, phase;  // [nphase], [nlyr][nphase]
  int pstride() const { return (nmom > nstr ? nmom : nstr) + 1; }
  Real pm(int k, int lc) const { return pmom[(size_t)lc * pstride() + k]; }
};

// Index j with xx[j] <= x < xx[j+1] in a sorted table (locate,
// src/Utils.hpp:27-64): 0 at or below an increasing table's first entry,
// n-2 at its last entry, n-1 beyond it.
int locate_sorted(const std::vector<Real>& xx, Real x);

struct GeneralOutput {
  int ntau = 0, numu = 0, nphi = 0, modes = 0;  // modes = Fourier modes computed
  std::vector<Real> f[8];                      // [ntau] each, FluxResult field order
  std::vector<Real> inten;                     // [ntau][numu][nphi]
  std::vector<Real> inten_avg;                 // [ntau][numu]
  std::vector<Real> mu_angles;                 // [numu]
  // SPECIAL_BC: albedo and transmissivity of the medium at the user polar
  // angles ([numu], the first numu/2 filled), spherical albedo and
  // transmissivity.
  std::vector<Real> albedo_medium, transmissivity_medium;
  Real spherical_albedo = 0, spherical_transmissivity = 0;
  // Temperature Jacobians [6][ntau][nlyr+2] when requested (thermal only).
  std::vector<Real> jac[6];
};

void validate_general_config(const GeneralConfig& c);
GeneralOutput general_solve(const GeneralConfig& c);
Real rpv_reflectivity(const Real* rpv, Real mu1, Real mu2, Real phi);
Real rpv_reflectivity_exact(const Real* rpv, Real mu1, Real mu2, Real phi);
Real hapke_reflectivity(const Real* p, Real mu, Real mup, Real dphi);
Real coxmunk_reflectivity(const Real* p, Real mu, Real mup, Real dphi);
Real rossli_reflectivity(const Real* p, Real mu, Real mup, Real dphi);
Real brdf_reflectivity(int kind, const Real* p, Real mu, Real mup, Real dphi);

}  // namespace ORACLE_NS
