// disortpp.hpp -- C++17 host facade with DisORT++'s class API, solving on the
// GPU through the C ABI of libdisort_amd.so (include/disort_amd.h).
//
// Names, members, defaults, validation messages and exception types follow
// the reference so that code written against DisORT++ compiles and behaves
// the same:
//   DisortFluxConfig   src/DisortFluxConfig.hpp:15-93, .cpp:9-143
//   FluxResult         src/FluxResult.hpp:14-51
//   DisortFluxSolver<N> src/FluxSolver.hpp:49 (solve), src/FluxSolver.cpp:23-171
//   DisortConfig & co  src/DisortConfig.hpp:13-243, .cpp:72-249
//   DisortResult       src/DisortResult.hpp:16-125
//   DisortSolver       src/DisortSolver.hpp:87, src/DisortSolver.cpp:21-399
//   spectral drivers   src/python_bindings.cpp:51-211
// There is no CPU solver behind this facade: every solve runs the gfx950
// kernels, and configurations the GPU path does not cover raise
// std::runtime_error instead of being solved some other way.
#pragma once

#include <algorithm>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace disortpp {

enum class PhaseFunction { Isotropic = 1, Rayleigh = 2, HenyeyGreenstein = 3, HazeGarciaSiewert = 4, CloudGarciaSiewert = 5 };
enum class BoundaryConditionType { General = 0, Special = 1 };
enum class BrdfType { None = 0, RPV = 1, CoxMunk = 2, Ambrals = 3, Hapke = 4 };

// Phase-moment fillers shared by both config classes
// (src/PhaseFunction.cpp:6-116).  lc < 0 fills every layer.
namespace phase_function {
void fillHenyeyGreenstein(std::vector<std::vector<double>>& m, int nmom_nstr, int nlyr, double g, int lc);
void fillIsotropic(std::vector<std::vector<double>>& m, int nmom_nstr, int nlyr, int lc);
void fillRayleigh(std::vector<std::vector<double>>& m, int nmom_nstr, int nlyr, int lc);
void fillHazeGarciaSiewert(std::vector<std::vector<double>>& m, int nmom_nstr, int nlyr, int lc);
void fillCloudGarciaSiewert(std::vector<std::vector<double>>& m, int nmom_nstr, int nlyr, int lc);
void fillPhaseFunction(std::vector<std::vector<double>>& m, int nmom_nstr, int nlyr, PhaseFunction type, double g,
                       int lc);
}  // namespace phase_function

// ---------------------------------------------------------------------------
// Flux-only configuration and result
// ---------------------------------------------------------------------------
class DisortFluxConfig {
 public:
  int num_layers = 0;
  int num_streams = 0;
  int num_phase_func_moments = 0;
  int nmomNstr() const { return std::max(num_phase_func_moments, num_streams); }

  bool use_thermal_emission = false;
  bool use_spherical_beam = false;
  bool use_delta_m_plus = false;
  bool use_diffusion_lower_bc = false;
  bool index_from_bottom = false;
  bool compute_temperature_jacobian = false;

  double direct_beam_flux = 0.0;
  double direct_beam_mu = 0.5;
  double isotropic_flux_top = 0.0;
  double isotropic_flux_bottom = 0.0;
  double temperature_top = 0.0;
  double temperature_bottom = 0.0;
  double emissivity_top = 1.0;
  double surface_albedo = 0.0;

  double wavenumber_low = 0.0;
  double wavenumber_high = 0.0;
  double bottom_radius = 0.0;

  std::vector<double> delta_tau;
  std::vector<double> single_scat_albedo;
  std::vector<std::vector<double>> phase_function_moments;  // [num_layers][nmomNstr()+1]
  std::vector<double> temperature;                          // [num_layers+1]
  std::vector<double> level_altitudes;                      // [num_layers+1]

  DisortFluxConfig() = default;
  DisortFluxConfig(int nlyr, int nstr, int nmom = 0);
  void allocate();
  void validate() const;

  double& phaseFunctionMoments(int k, int lc) { return phase_function_moments[lc][k]; }
  const double& phaseFunctionMoments(int k, int lc) const { return phase_function_moments[lc][k]; }
  void setHenyeyGreenstein(double g, int lc = -1);
  void setIsotropic(int lc = -1);
  void setRayleigh(int lc = -1);
  void setHazeGarciaSiewert(int lc = -1);
  void setCloudGarciaSiewert(int lc = -1);
  void setPhaseFunction(PhaseFunction type, double g = 0.0, int lc = -1);
};

struct FluxResult {
  std::vector<double> flux_direct_beam;
  std::vector<double> flux_down;
  std::vector<double> flux_up;
  std::vector<double> flux_tau_divergence;
  std::vector<double> mean_intensity;
  std::vector<double> mean_intensity_down;
  std::vector<double> mean_intensity_up;
  std::vector<double> mean_intensity_direct_beam;
  // Temperature Jacobians [level][dof], dof 0..num_layers the level
  // temperatures and num_layers+1 temperature_bottom; filled only when
  // compute_temperature_jacobian and use_thermal_emission are set.
  std::vector<std::vector<double>> flux_up_temperature_jac;
  std::vector<std::vector<double>> flux_down_temperature_jac;
  std::vector<std::vector<double>> mean_intensity_temperature_jac;
  std::vector<std::vector<double>> mean_intensity_down_temperature_jac;
  std::vector<std::vector<double>> mean_intensity_up_temperature_jac;
  std::vector<std::vector<double>> flux_divergence_temperature_jac;

  int num_levels() const { return static_cast<int>(flux_up.size()); }
  void allocate(int num_layers);
  double totalFluxDown(int lev) const { return flux_direct_beam[lev] + flux_down[lev]; }
  double netFlux(int lev) const { return flux_direct_beam[lev] + flux_down[lev] - flux_up[lev]; }
};

// ---------------------------------------------------------------------------
// General configuration and result (DisortSolver)
// ---------------------------------------------------------------------------
struct DisortFlags {
  bool use_user_tau = false;
  bool use_user_mu = false;
  BoundaryConditionType ibcnd = BoundaryConditionType::General;
  bool use_lambertian_surface = true;
  bool use_thermal_emission = false;
  bool use_spherical_beam = false;
  bool comp_only_fluxes = false;
  BrdfType brdf_type = BrdfType::None;
  bool intensity_corr_buras = false;
  bool intensity_corr_nakajima = false;
  bool use_delta_m_plus = false;
  bool use_diffusion_lower_bc = false;
  bool output_fourier_expansion = false;
  bool index_from_bottom = false;
  bool compute_temperature_jacobian = false;
};

struct BoundaryConditions {
  double direct_beam_flux = 0.0;
  double direct_beam_mu = 0.5;
  double direct_beam_phi = 0.0;
  double isotropic_flux_top = 0.0;
  double isotropic_flux_bottom = 0.0;
  double temperature_top = 0.0;
  double temperature_bottom = 0.0;
  double emissivity_top = 1.0;
  double surface_albedo = 0.0;
};

struct RpvBrdfSpec {
  double rho0 = 0.0, k = 0.0, theta = 0.0, sigma = 0.0, t1 = 0.0, t2 = 0.0, scale = 1.0;
};
struct AmbralsBrdfSpec {
  double iso = 0.0, vol = 0.0, geo = 0.0;
};
struct CoxMunkBrdfSpec {
  double u10 = 0.0, pcl = 0.0, xsal = 0.0, refractive_index = 1.34;
  bool do_shadow = true;
};
struct HapkeBrdfSpec {
  double b0 = 1.0, hh = 0.06, w = 0.6;
};
struct BrdfSpecification {
  std::optional<RpvBrdfSpec> rpv;
  std::optional<AmbralsBrdfSpec> ambrals;
  std::optional<CoxMunkBrdfSpec> cox_munk;
  std::optional<HapkeBrdfSpec> hapke;
  void setRpv(const RpvBrdfSpec& s) { rpv = s; }
  void setAmbrals(const AmbralsBrdfSpec& s) { ambrals = s; }
  void setCoxMunk(const CoxMunkBrdfSpec& s) { cox_munk = s; }
  void setHapke(const HapkeBrdfSpec& s) { hapke = s; }
};

class DisortConfig {
 public:
  DisortFlags flags;
  BoundaryConditions bc;
  BrdfSpecification brdf;

  int num_layers = 0;
  int num_phase_func_moments = 0;
  int num_streams = 0;
  int num_user_tau = 0;
  int num_user_mu = 0;
  int num_phi = 0;
  int num_phase_func_angles = 0;
  int nmomNstr() const { return std::max(num_phase_func_moments, num_streams); }

  double wavenumber_low = 0.0;
  double wavenumber_high = 0.0;
  double accuracy_fourier_series = 0.0;
  double bottom_radius = 0.0;

  std::vector<double> delta_tau;
  std::vector<double> single_scat_albedo;
  std::vector<std::vector<double>> phase_function_moments;
  std::vector<double> temperature;
  std::vector<double> level_altitudes;
  std::vector<double> tau_user;
  std::vector<double> mu_user;
  std::vector<double> phi_user;
  std::vector<double> mu_phase_function;
  std::vector<std::vector<double>> phase_function;

  DisortConfig() = default;
  DisortConfig(int nlyr, int nstr, int nmom = 0);
  void allocate();
  void validate() const;

  double& phaseFunctionMoments(int k, int lc) { return phase_function_moments[lc][k]; }
  const double& phaseFunctionMoments(int k, int lc) const { return phase_function_moments[lc][k]; }
  void setHenyeyGreenstein(double g, int lc = -1);
  void setIsotropic(int lc = -1);
  void setRayleigh(int lc = -1);
  void setHazeGarciaSiewert(int lc = -1);
  void setCloudGarciaSiewert(int lc = -1);
  void setPhaseFunction(PhaseFunction type, double g = 0.0, int lc = -1);
};

struct DisortResult {
  std::vector<double> flux_direct_beam;
  std::vector<double> flux_down;
  std::vector<double> flux_up;
  std::vector<double> flux_tau_divergence;
  std::vector<double> mean_intensity;
  std::vector<double> mean_intensity_down;
  std::vector<double> mean_intensity_up;
  std::vector<double> mean_intensity_direct_beam;
  std::vector<double> albedo_medium;
  std::vector<double> transmissivity_medium;
  double spherical_albedo = 0.0;
  double spherical_transmissivity = 0.0;
  std::vector<std::vector<double>> flux_up_temperature_jac;
  std::vector<std::vector<double>> flux_down_temperature_jac;
  std::vector<std::vector<double>> mean_intensity_temperature_jac;
  std::vector<std::vector<double>> mean_intensity_down_temperature_jac;
  std::vector<std::vector<double>> mean_intensity_up_temperature_jac;
  std::vector<std::vector<double>> flux_divergence_temperature_jac;
  std::vector<double> mu_angles;
  std::vector<std::vector<std::vector<double>>> intensity;  // [tau][mu][phi]
  std::vector<std::vector<double>> intensity_azimuth_avg;   // [tau][mu]
  std::vector<std::vector<std::vector<double>>> intensity_fourier_expansion;

  int num_user_tau() const { return static_cast<int>(flux_up.size()); }
  int num_user_mu() const { return intensity.empty() ? 0 : static_cast<int>(intensity[0].size()); }
  int num_phi() const {
    return (intensity.empty() || intensity[0].empty()) ? 0 : static_cast<int>(intensity[0][0].size());
  }
  double totalFluxDown(int lu) const { return flux_direct_beam[lu] + flux_down[lu]; }
  double netFlux(int lu) const { return flux_direct_beam[lu] + flux_down[lu] - flux_up[lu]; }
  double intensities(int iu, int lu, int j) const { return intensity[lu][iu][j]; }
  double intensitiesAzimuthAvg(int iu, int lu) const { return intensity_azimuth_avg[lu][iu]; }
};

// ---------------------------------------------------------------------------
// Solvers
// ---------------------------------------------------------------------------
// Per-band overrides of the spectral drivers (src/python_bindings.cpp:51-82).
struct SpectralOverrides {
  std::vector<std::vector<double>> delta_tau;
  std::vector<std::vector<double>> single_scat_albedo;
  std::vector<std::vector<std::vector<double>>> phase_function_moments;
  std::vector<double> surface_albedo;
  std::vector<double> emissivity_top;
  std::vector<double> direct_beam_flux;
  std::vector<double> isotropic_flux_top;
  std::vector<double> isotropic_flux_bottom;
};

// The same overrides as borrowed C-contiguous arrays (the NumPy form; no
// counterpart in the reference, whose bindings take nested lists).  A null
// pointer means "not given".  delta_tau and single_scat_albedo are
// [n][cols], phase_function_moments [n][rows][cols], the scalars [n]; row i
// overrides band i exactly as the nested-list entry i does, and n smaller
// than the band count raises std::out_of_range like the lists' .at(i).
struct ArrayView {
  const double* data = nullptr;
  int64_t n = 0, rows = 1, cols = 1;
  bool given() const { return data != nullptr; }
};
struct SpectralArrays {
  ArrayView delta_tau, single_scat_albedo, phase_function_moments;
  ArrayView surface_albedo, emissivity_top, direct_beam_flux, isotropic_flux_top, isotropic_flux_bottom;
  SpectralOverrides to_overrides(size_t bands) const;  // the nested-list form (copies)
};

// Solve many flux configurations as ONE GPU batch (ragged in layers and
// streams).  Configurations must share the batch-uniform flags
// (use_thermal_emission, use_spherical_beam, use_delta_m_plus,
// use_diffusion_lower_bc); each may use index_from_bottom independently.
std::vector<FluxResult> solveFluxBatch(const std::vector<DisortFluxConfig>& configs);

// Stream counts the reference instantiates (src/FluxSolver.cpp:1572-1580).
bool isSupportedStreamCount(int nstr);

template <int NStr>
class DisortFluxSolver {
 public:
  static_assert(NStr == 4 || NStr == 6 || NStr == 8 || NStr == 10 || NStr == 12 || NStr == 14 || NStr == 16 ||
                    NStr == 32 || NStr == 64,
                "unsupported stream count");
  DisortFluxSolver() = default;
  FluxResult solve(const DisortFluxConfig& config);
};

extern template class DisortFluxSolver<4>;
extern template class DisortFluxSolver<6>;
extern template class DisortFluxSolver<8>;
extern template class DisortFluxSolver<10>;
extern template class DisortFluxSolver<12>;
extern template class DisortFluxSolver<14>;
extern template class DisortFluxSolver<16>;
extern template class DisortFluxSolver<32>;
extern template class DisortFluxSolver<64>;

class DisortSolver {
 public:
  DisortSolver() = default;
  // Mutates `config` as the reference does: auto-allocation when delta_tau
  // is empty, num_user_tau/tau_user and num_user_mu/mu_user set when not
  // user-specified, pmom[0] = 1.
  DisortResult solve(DisortConfig& config);
};

std::vector<FluxResult> solveFluxSpectralBands(DisortFluxConfig& config,
                                               const std::vector<std::pair<double, double>>& wavenumber_bands,
                                               const SpectralOverrides& ov);
std::vector<FluxResult> solveFluxSpectral(DisortFluxConfig& config, const std::vector<double>& wavenumbers,
                                          const SpectralOverrides& ov);
std::vector<DisortResult> solveSpectralBands(DisortConfig config,
                                             const std::vector<std::pair<double, double>>& wavenumber_bands,
                                             const SpectralOverrides& ov);
std::vector<DisortResult> solveSpectral(DisortConfig config, const std::vector<double>& wavenumbers,
                                        const SpectralOverrides& ov);
// Array forms: the flux drivers pack the per-band arrays straight into the
// GPU batch (no per-band configuration is built; with top-down layers and
// rows of exactly num_layers (and nmomNstr()+1) entries the arrays are handed
// to the GPU copy as they are).
std::vector<FluxResult> solveFluxSpectralBands(DisortFluxConfig& config,
                                               const std::vector<std::pair<double, double>>& wavenumber_bands,
                                               const SpectralArrays& ov);
std::vector<FluxResult> solveFluxSpectral(DisortFluxConfig& config, const std::vector<double>& wavenumbers,
                                          const SpectralArrays& ov);
std::vector<DisortResult> solveSpectralBands(DisortConfig config,
                                             const std::vector<std::pair<double, double>>& wavenumber_bands,
                                             const SpectralArrays& ov);
std::vector<DisortResult> solveSpectral(DisortConfig config, const std::vector<double>& wavenumbers,
                                        const SpectralArrays& ov);

// GPU device used by the facade (default 0, or DISORT_AMD_DEVICE).
void setDevice(int device);

}  // namespace disortpp
