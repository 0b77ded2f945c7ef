// disortpp.cpp -- DisORT++ class API over the gfx950 C ABI (libdisort_amd.so).
//
// Every solve is packed into one structure-of-arrays batch (disort_flux_batch)
// and handed to disort_flux_solve; a spectral call over B bands is a single
// batch of B instances instead of the reference's OpenMP band loop
// (src/python_bindings.cpp:130-211).
#include "disortpp.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <mutex>

#include "../../include/disort_amd.h"
#include "garcia_siewert.hpp"

namespace disortpp {

// ---------------------------------------------------------------------------
// Phase functions (src/PhaseFunction.cpp:6-116)
// ---------------------------------------------------------------------------
namespace phase_function {
namespace {
template <class F>
void for_layers(int nlyr, int lc, F&& f) {
  const int lo = lc < 0 ? 0 : lc, hi = lc < 0 ? nlyr - 1 : lc;
  for (int l = lo; l <= hi; ++l) f(l);
}

// Moments chi_0 = 1, chi_k = beta_k / (2k+1) from a table of expansion
// coefficients beta_1..beta_count, zero beyond it.
void fill_tabulated(std::vector<std::vector<double>>& m, int nmom_nstr, int nlyr, int lc, const double* beta,
                    int count) {
  for_layers(nlyr, lc, [&](int l) {
    m[l][0] = 1.0;
    for (int k = 1; k <= nmom_nstr; ++k) m[l][k] = k <= count ? beta[k - 1] / (2 * k + 1) : 0.0;
  });
}
}  // namespace

void fillHenyeyGreenstein(std::vector<std::vector<double>>& m, int nmom_nstr, int nlyr, double g, int lc) {
  for_layers(nlyr, lc, [&](int l) {
    double gk = 1.0;  // successive products, as the reference forms g^k
    for (int k = 0; k <= nmom_nstr; ++k, gk *= g) m[l][k] = gk;
  });
}

void fillIsotropic(std::vector<std::vector<double>>& m, int nmom_nstr, int nlyr, int lc) {
  for_layers(nlyr, lc, [&](int l) {
    for (int k = 0; k <= nmom_nstr; ++k) m[l][k] = k == 0 ? 1.0 : 0.0;
  });
}

void fillRayleigh(std::vector<std::vector<double>>& m, int nmom_nstr, int nlyr, int lc) {
  for_layers(nlyr, lc, [&](int l) {
    for (int k = 0; k <= nmom_nstr; ++k) m[l][k] = k == 0 ? 1.0 : (k == 2 ? 0.1 : 0.0);
  });
}

void fillHazeGarciaSiewert(std::vector<std::vector<double>>& m, int nmom_nstr, int nlyr, int lc) {
  fill_tabulated(m, nmom_nstr, nlyr, lc, garcia_siewert::kHaze, garcia_siewert::kHazeCount);
}

void fillCloudGarciaSiewert(std::vector<std::vector<double>>& m, int nmom_nstr, int nlyr, int lc) {
  fill_tabulated(m, nmom_nstr, nlyr, lc, garcia_siewert::kCloud, garcia_siewert::kCloudUsed);
}

void fillPhaseFunction(std::vector<std::vector<double>>& m, int nmom_nstr, int nlyr, PhaseFunction type, double g,
                       int lc) {
  switch (type) {
    case PhaseFunction::Isotropic: return fillIsotropic(m, nmom_nstr, nlyr, lc);
    case PhaseFunction::Rayleigh: return fillRayleigh(m, nmom_nstr, nlyr, lc);
    case PhaseFunction::HenyeyGreenstein: return fillHenyeyGreenstein(m, nmom_nstr, nlyr, g, lc);
    case PhaseFunction::HazeGarciaSiewert:
      return fill_tabulated(m, nmom_nstr, nlyr, lc, garcia_siewert::kHaze, garcia_siewert::kHazeCount);
    case PhaseFunction::CloudGarciaSiewert:
      return fill_tabulated(m, nmom_nstr, nlyr, lc, garcia_siewert::kCloud, garcia_siewert::kCloudUsed);
  }
  throw std::invalid_argument("unknown phase function type");
}
}  // namespace phase_function

namespace {

void check_phase_alloc(const std::vector<std::vector<double>>& m, int nlyr, int nmom_nstr) {
  if (static_cast<int>(m.size()) < nlyr) throw std::invalid_argument("phase_function_moments not allocated");
  for (int l = 0; l < nlyr; ++l)
    if (static_cast<int>(m[l].size()) < nmom_nstr + 1) throw std::invalid_argument("phase_function_moments not allocated");
}

// ---------------------------------------------------------------------------
// The GPU behind the facade: one context, calls serialised.
// ---------------------------------------------------------------------------
struct Gpu {
  std::mutex mu;
  disort_ctx* ctx = nullptr;
  int device = -1;
  ~Gpu() {
    if (ctx) disort_ctx_destroy(ctx);
  }
};

Gpu& gpu() {
  static Gpu g;
  return g;
}

int default_device() {
  const char* e = std::getenv("DISORT_AMD_DEVICE");
  return e ? std::atoi(e) : 0;
}

// The per-instance record of one flux problem, in top-down order.
struct Instance {
  int nlyr, nstr, pstride;
  bool reverse;
  const std::vector<double>* dtau;
  const std::vector<double>* ssa;
  const std::vector<std::vector<double>>* pmom;
  const std::vector<double>* temper;
  const std::vector<double>* alt;
  double scal[DISORT_NSCAL];
  bool pmom0_one;  // DisortSolver sets pmom[0] = 1 (src/DisortSolver.cpp:669)
};

// Packs instances that share the batch-uniform flags into SoA arrays,
// applying index_from_bottom by reversal (src/IndexConversion.hpp:40-51):
// whole vectors are reversed, then the first nlyr (+1) entries are used.
struct Packed {
  int32_t flags = 0;
  std::vector<int32_t> nstr, nlyr, pstride, status;
  std::vector<int64_t> lay_off, mom_off;
  std::vector<double> dtau, ssa, pmom, temper, alt, scal, out;
  int64_t layers = 0, levels = 0, moments = 0;
  // Caller-owned arrays used in place of dtau / ssa / pmom when set (the
  // array-form spectral drivers with inputs already in the batch layout).
  const double *ext_dtau = nullptr, *ext_ssa = nullptr, *ext_pmom = nullptr;

  void add(const Instance& in) {
    const int L = in.nlyr, P = in.pstride;
    nstr.push_back(in.nstr);
    nlyr.push_back(L);
    pstride.push_back(P);
    lay_off.push_back(layers);
    mom_off.push_back(moments);
    auto pick = [&](const std::vector<double>& v, int count, std::vector<double>& dst) {
      const int64_t sz = static_cast<int64_t>(v.size());
      for (int j = 0; j < count; ++j) dst.push_back(in.reverse ? v[sz - 1 - j] : v[j]);
    };
    pick(*in.dtau, L, dtau);
    pick(*in.ssa, L, ssa);
    const int64_t rows = static_cast<int64_t>(in.pmom->size());
    for (int l = 0; l < L; ++l) {
      const auto& row = (*in.pmom)[in.reverse ? rows - 1 - l : l];
      for (int k = 0; k < P; ++k) pmom.push_back(k == 0 && in.pmom0_one ? 1.0 : row[k]);
    }
    if (flags & DISORT_F_THERMAL) pick(*in.temper, L + 1, temper);
    if (flags & DISORT_F_SPHERICAL) pick(*in.alt, L + 1, alt);
    scal.insert(scal.end(), in.scal, in.scal + DISORT_NSCAL);
    layers += L;
    levels += L + 1;
    moments += static_cast<int64_t>(L) * P;
  }

  // Solves the batch on the GPU; throws the reference's exception types.
  void solve() {
    const int64_t n = static_cast<int64_t>(nstr.size());
    if (n == 0) return;
    out.assign(static_cast<size_t>(DISORT_NOUT * levels), 0.0);
    status.assign(static_cast<size_t>(n), 0);
    disort_flux_batch b = view();
    b.out = out.data();
    b.status = status.data();

    run([&](disort_ctx* ctx) { return disort_flux_solve(ctx, &b, DISORT_MEM_HOST); });
  }

  // Temperature Jacobians of every instance (a thermal batch) into jac,
  // [instance][6][L+1][L+2] (disort_flux_jacobian's layout).
  void jacobian() {
    const int64_t n = static_cast<int64_t>(nstr.size());
    jac_off.assign(static_cast<size_t>(n) + 1, 0);
    for (int64_t i = 0; i < n; ++i) jac_off[i + 1] = jac_off[i] + 6LL * (nlyr[i] + 1) * (nlyr[i] + 2);
    jac.assign(static_cast<size_t>(jac_off[n]), 0.0);
    if (n == 0) return;
    disort_flux_batch b = view();
    run([&](disort_ctx* ctx) { return disort_flux_jacobian(ctx, &b, jac.data()); });
  }

  // Instance i's Jacobians into the six *_temperature_jac members, [level
  // lu][dof]; indexed from the bottom, rows and level columns are reversed
  // and the surface column stays last (src/IndexConversion.hpp:57-68).
  template <class R>
  void unpack_jacobian(int64_t i, bool reverse, R& r) const {
    const int L = nlyr[i], nd = L + 2;
    std::vector<std::vector<double>>* f[6] = {&r.flux_up_temperature_jac, &r.flux_down_temperature_jac,
                                              &r.mean_intensity_temperature_jac,
                                              &r.mean_intensity_down_temperature_jac,
                                              &r.mean_intensity_up_temperature_jac, &r.flux_divergence_temperature_jac};
    for (int q = 0; q < 6; ++q) {
      f[q]->assign(L + 1, std::vector<double>(nd, 0.0));
      const double* src = jac.data() + jac_off[i] + static_cast<int64_t>(q) * (L + 1) * nd;
      for (int lu = 0; lu <= L; ++lu)
        for (int k = 0; k < nd; ++k)
          (*f[q])[reverse ? L - lu : lu][reverse && k <= L ? L - k : k] = src[static_cast<int64_t>(lu) * nd + k];
    }
  }

  std::vector<double> jac;
  std::vector<int64_t> jac_off;

 private:
  disort_flux_batch view() {
    disort_flux_batch b{};
    b.n = static_cast<int64_t>(nstr.size());
    b.flags = flags;
    b.nstr = nstr.data();
    b.nlyr = nlyr.data();
    b.pstride = pstride.data();
    b.lay_off = lay_off.data();
    b.mom_off = mom_off.data();
    b.dtau = ext_dtau ? ext_dtau : dtau.data();
    b.ssa = ext_ssa ? ext_ssa : ssa.data();
    b.pmom = ext_pmom ? ext_pmom : pmom.data();
    b.temper = (flags & DISORT_F_THERMAL) ? temper.data() : nullptr;
    b.alt = (flags & DISORT_F_SPHERICAL) ? alt.data() : nullptr;
    b.scal = scal.data();
    b.total_levels = levels;
    return b;
  }

 public:
  // Runs one C-ABI call on the process's context (created on first use);
  // throws the reference's exception types.
  template <class F>
  void run(F call) {
    Gpu& g = gpu();
    std::lock_guard<std::mutex> lock(g.mu);
    if (!g.ctx) {
      const int dev = g.device >= 0 ? g.device : default_device();
      if (disort_ctx_create(dev, &g.ctx) != DISORT_OK) {
        g.ctx = nullptr;
        throw std::runtime_error("disortpp: cannot create a GPU context on device " + std::to_string(dev));
      }
      g.device = dev;
    }
    const int rc = call(g.ctx);
    if (rc == DISORT_OK) return;
    std::string msg = disort_last_error(g.ctx);
    if (nstr.size() == 1 && msg.rfind("instance 0: ", 0) == 0) msg = msg.substr(12);
    if (rc == DISORT_E_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error("disortpp: " + msg);
  }

 public:

  // Copies instance i's outputs into the eight result vectors (reversed
  // when the instance was indexed from the bottom).
  template <class R>
  void unpack(int64_t i, bool reverse, R& r) const {
    const int L = nlyr[i];
    const int64_t v0 = lay_off[i] + i;
    std::vector<double>* f[DISORT_NOUT] = {&r.flux_direct_beam, &r.flux_down, &r.flux_up,
                                           &r.flux_tau_divergence, &r.mean_intensity, &r.mean_intensity_down,
                                           &r.mean_intensity_up, &r.mean_intensity_direct_beam};
    for (int q = 0; q < DISORT_NOUT; ++q) {
      f[q]->resize(L + 1);
      const double* src = out.data() + q * levels + v0;
      for (int l = 0; l <= L; ++l) (*f[q])[reverse ? L - l : l] = src[l];
    }
  }
};

int32_t batch_flags(bool thermal, bool spherical, bool dmp, bool diffusion) {
  return (thermal ? DISORT_F_THERMAL : 0) | (spherical ? DISORT_F_SPHERICAL : 0) |
         (dmp ? DISORT_F_DELTA_M_PLUS : 0) | (diffusion ? DISORT_F_DIFFUSION_BC : 0);
}

void fill_scal(double* s, double beam, double mu0, double iso_top, double iso_bot, double t_top, double t_bot,
               double emis, double albedo, double wlo, double whi, double radius) {
  for (int k = 0; k < DISORT_NSCAL; ++k) s[k] = 0.0;
  s[DISORT_S_BEAM_FLUX] = beam;
  s[DISORT_S_BEAM_MU] = mu0;
  s[DISORT_S_ISO_TOP] = iso_top;
  s[DISORT_S_ISO_BOT] = iso_bot;
  s[DISORT_S_TEMP_TOP] = t_top;
  s[DISORT_S_TEMP_BOT] = t_bot;
  s[DISORT_S_EMIS_TOP] = emis;
  s[DISORT_S_ALBEDO] = albedo;
  s[DISORT_S_WN_LO] = wlo;
  s[DISORT_S_WN_HI] = whi;
  s[DISORT_S_RADIUS] = radius;
}

Instance flux_instance(const DisortFluxConfig& c) {
  Instance in{};
  in.nlyr = c.num_layers;
  in.nstr = c.num_streams;
  in.pstride = c.nmomNstr() + 1;
  in.reverse = c.index_from_bottom;
  in.dtau = &c.delta_tau;
  in.ssa = &c.single_scat_albedo;
  in.pmom = &c.phase_function_moments;
  in.temper = &c.temperature;
  in.alt = &c.level_altitudes;
  in.pmom0_one = false;
  fill_scal(in.scal, c.direct_beam_flux, c.direct_beam_mu, c.isotropic_flux_top, c.isotropic_flux_bottom,
            c.temperature_top, c.temperature_bottom, c.emissivity_top, c.surface_albedo, c.wavenumber_low,
            c.wavenumber_high, c.bottom_radius);
  return in;
}

// src/FluxSolver.cpp:1299-1307: Jacobians only with thermal emission on.
bool wants_jacobian(const DisortFluxConfig& c) { return c.compute_temperature_jacobian && c.use_thermal_emission; }

int32_t flux_flags(const DisortFluxConfig& c) {
  return batch_flags(c.use_thermal_emission, c.use_spherical_beam, c.use_delta_m_plus, c.use_diffusion_lower_bc);
}

// Validation in the order the reference's solve() runs it: on the
// top-down view (after reversal), then the GPU-path limits.
void validate_flux(const DisortFluxConfig& c) {
  if (c.index_from_bottom) {
    DisortFluxConfig r = c;
    std::reverse(r.delta_tau.begin(), r.delta_tau.end());
    std::reverse(r.single_scat_albedo.begin(), r.single_scat_albedo.end());
    std::reverse(r.phase_function_moments.begin(), r.phase_function_moments.end());
    std::reverse(r.temperature.begin(), r.temperature.end());
    std::reverse(r.level_altitudes.begin(), r.level_altitudes.end());
    r.validate();
  } else {
    c.validate();
  }
}

}  // namespace

void setDevice(int device) {
  Gpu& g = gpu();
  std::lock_guard<std::mutex> lock(g.mu);
  if (g.ctx && g.device != device) {
    disort_ctx_destroy(g.ctx);
    g.ctx = nullptr;
  }
  g.device = device;
}

bool isSupportedStreamCount(int n) {
  return n == 4 || n == 6 || n == 8 || n == 10 || n == 12 || n == 14 || n == 16 || n == 32 || n == 64;
}

// ---------------------------------------------------------------------------
// DisortFluxConfig (src/DisortFluxConfig.cpp:9-143)
// ---------------------------------------------------------------------------
DisortFluxConfig::DisortFluxConfig(int nlyr, int nstr, int nmom)
    : num_layers(nlyr), num_streams(nstr), num_phase_func_moments(nmom) {
  if (num_layers <= 0) throw std::invalid_argument("num_layers must be > 0");
  if (num_streams < 4 || num_streams % 2 != 0) throw std::invalid_argument("num_streams must be even and >= 4");
  if (num_phase_func_moments < 0) throw std::invalid_argument("num_phase_func_moments must be >= 0");
}

void DisortFluxConfig::allocate() {
  delta_tau.resize(num_layers, 0.0);
  single_scat_albedo.resize(num_layers, 0.0);
  phase_function_moments.resize(num_layers);
  for (auto& row : phase_function_moments) row.resize(nmomNstr() + 1, 0.0);
  if (use_thermal_emission) temperature.resize(num_layers + 1, 0.0);
  if (use_spherical_beam) level_altitudes.resize(num_layers + 1, 0.0);
}

void DisortFluxConfig::validate() const {
  auto bad = [](const std::string& m) { throw std::invalid_argument(m); };
  if (num_layers <= 0) bad("num_layers must be > 0");
  if (num_streams < 4 || num_streams % 2 != 0) bad("num_streams must be even and >= 4");
  if (num_phase_func_moments < 0) bad("num_phase_func_moments must be >= 0");
  if (static_cast<int>(delta_tau.size()) < num_layers)
    bad("delta_tau size (" + std::to_string(delta_tau.size()) + ") < num_layers (" + std::to_string(num_layers) + ")");
  if (static_cast<int>(single_scat_albedo.size()) < num_layers) bad("single_scat_albedo size < num_layers");
  if (static_cast<int>(phase_function_moments.size()) < num_layers) bad("phase_function_moments size < num_layers");
  for (int lc = 0; lc < num_layers; ++lc)
    if (static_cast<int>(phase_function_moments[lc].size()) < nmomNstr() + 1)
      bad("phase_function_moments[" + std::to_string(lc) + "] size < nmomNstr()+1");
  for (int lc = 0; lc < num_layers; ++lc) {
    if (delta_tau[lc] < 0.0) bad("delta_tau[" + std::to_string(lc) + "] must be >= 0");
    if (single_scat_albedo[lc] < 0.0 || single_scat_albedo[lc] > 1.0)
      bad("single_scat_albedo[" + std::to_string(lc) + "] must be in [0, 1]");
  }
  if (use_diffusion_lower_bc && !use_thermal_emission) bad("use_diffusion_lower_bc requires use_thermal_emission = true");
  if (direct_beam_flux < 0.0) bad("direct_beam_flux must be >= 0");
  if (direct_beam_flux > 0.0 && (direct_beam_mu <= 0.0 || direct_beam_mu > 1.0))
    bad("direct_beam_mu must be in (0, 1] when direct_beam_flux > 0");
  if (surface_albedo < 0.0 || surface_albedo > 1.0) bad("surface_albedo must be in [0, 1]");
  if (use_thermal_emission) {
    if (static_cast<int>(temperature.size()) < num_layers + 1) bad("temperature size < num_layers + 1");
    for (int l = 0; l <= num_layers; ++l)
      if (temperature[l] < 0.0) bad("temperature[" + std::to_string(l) + "] must be >= 0");
    if (wavenumber_low < 0.0 || wavenumber_low > wavenumber_high)
      bad("wavenumber: need 0 <= wavenumber_low <= wavenumber_high");
    if (temperature_top < 0.0 || temperature_bottom < 0.0) bad("temperature_top and temperature_bottom must be >= 0");
    if (emissivity_top < 0.0 || emissivity_top > 1.0) bad("emissivity_top must be in [0, 1]");
  }
  if (use_spherical_beam) {
    if (bottom_radius <= 0.0) bad("bottom_radius must be > 0 when use_spherical_beam is true");
    if (static_cast<int>(level_altitudes.size()) < num_layers + 1) bad("level_altitudes size < num_layers + 1");
  }
  if (use_delta_m_plus && num_phase_func_moments < num_streams + 1)
    bad("use_delta_m_plus requires num_phase_func_moments >= num_streams + 1");
}

void DisortFluxConfig::setHenyeyGreenstein(double g, int lc) {
  check_phase_alloc(phase_function_moments, num_layers, nmomNstr());
  phase_function::fillHenyeyGreenstein(phase_function_moments, nmomNstr(), num_layers, g, lc);
}
void DisortFluxConfig::setIsotropic(int lc) {
  check_phase_alloc(phase_function_moments, num_layers, nmomNstr());
  phase_function::fillIsotropic(phase_function_moments, nmomNstr(), num_layers, lc);
}
void DisortFluxConfig::setRayleigh(int lc) {
  check_phase_alloc(phase_function_moments, num_layers, nmomNstr());
  phase_function::fillRayleigh(phase_function_moments, nmomNstr(), num_layers, lc);
}
void DisortFluxConfig::setHazeGarciaSiewert(int lc) { setPhaseFunction(PhaseFunction::HazeGarciaSiewert, 0.0, lc); }
void DisortFluxConfig::setCloudGarciaSiewert(int lc) { setPhaseFunction(PhaseFunction::CloudGarciaSiewert, 0.0, lc); }
void DisortFluxConfig::setPhaseFunction(PhaseFunction type, double g, int lc) {
  check_phase_alloc(phase_function_moments, num_layers, nmomNstr());
  phase_function::fillPhaseFunction(phase_function_moments, nmomNstr(), num_layers, type, g, lc);
}

void FluxResult::allocate(int num_layers) {
  for (auto* v : {&flux_direct_beam, &flux_down, &flux_up, &flux_tau_divergence, &mean_intensity,
                  &mean_intensity_down, &mean_intensity_up, &mean_intensity_direct_beam})
    v->assign(num_layers + 1, 0.0);
}

// ---------------------------------------------------------------------------
// Flux solves
// ---------------------------------------------------------------------------
std::vector<FluxResult> solveFluxBatch(const std::vector<DisortFluxConfig>& configs) {
  std::vector<FluxResult> results(configs.size());
  for (const auto& c : configs) validate_flux(c);
  // One GPU batch per distinct set of batch-uniform flags.
  std::vector<int32_t> seen;
  for (const auto& c : configs)
    if (std::find(seen.begin(), seen.end(), flux_flags(c)) == seen.end()) seen.push_back(flux_flags(c));
  for (int32_t fl : seen) {
    Packed p;
    p.flags = fl;
    std::vector<size_t> idx;
    for (size_t i = 0; i < configs.size(); ++i)
      if (flux_flags(configs[i]) == fl) {
        p.add(flux_instance(configs[i]));
        idx.push_back(i);
      }
    p.solve();
    for (size_t j = 0; j < idx.size(); ++j) p.unpack(static_cast<int64_t>(j), configs[idx[j]].index_from_bottom, results[idx[j]]);
    // Temperature Jacobians for the instances that ask for them.
    Packed pj;
    pj.flags = fl;
    std::vector<size_t> jdx;
    for (size_t i : idx)
      if (wants_jacobian(configs[i])) {
        pj.add(flux_instance(configs[i]));
        jdx.push_back(i);
      }
    if (jdx.empty()) continue;
    pj.jacobian();
    for (size_t j = 0; j < jdx.size(); ++j)
      pj.unpack_jacobian(static_cast<int64_t>(j), configs[jdx[j]].index_from_bottom, results[jdx[j]]);
  }
  return results;
}

template <int NStr>
FluxResult DisortFluxSolver<NStr>::solve(const DisortFluxConfig& config) {
  validate_flux(config);
  if (config.num_streams != NStr)
    throw std::invalid_argument("DisortFluxSolver<" + std::to_string(NStr) + ">::solve: config.num_streams=" +
                                std::to_string(config.num_streams) + " does not match template NStr");
  Packed p;
  p.flags = flux_flags(config);
  p.add(flux_instance(config));
  p.solve();
  FluxResult r;
  p.unpack(0, config.index_from_bottom, r);
  if (wants_jacobian(config)) {
    p.jacobian();
    p.unpack_jacobian(0, config.index_from_bottom, r);
  }
  return r;
}

template class DisortFluxSolver<4>;
template class DisortFluxSolver<6>;
template class DisortFluxSolver<8>;
template class DisortFluxSolver<10>;
template class DisortFluxSolver<12>;
template class DisortFluxSolver<14>;
template class DisortFluxSolver<16>;
template class DisortFluxSolver<32>;
template class DisortFluxSolver<64>;

namespace {
void unsupported_streams(int nstr) {
  throw std::invalid_argument("Unsupported num_streams=" + std::to_string(nstr) +
                              ". Supported values: 4, 6, 8, 10, 12, 14, 16, 32, 64");
}

// The per-band overrides of applyOverrides (src/python_bindings.cpp:73-82).
template <class Cfg, class Set>
void apply_overrides(Cfg& c, const SpectralOverrides& ov, size_t i, Set&& set_scalars) {
  if (!ov.delta_tau.empty()) c.delta_tau = ov.delta_tau.at(i);
  if (!ov.single_scat_albedo.empty()) c.single_scat_albedo = ov.single_scat_albedo.at(i);
  if (!ov.phase_function_moments.empty()) c.phase_function_moments = ov.phase_function_moments.at(i);
  set_scalars(c);
}
}  // namespace

std::vector<FluxResult> solveFluxSpectralBands(DisortFluxConfig& config,
                                               const std::vector<std::pair<double, double>>& bands,
                                               const SpectralOverrides& ov) {
  if (!isSupportedStreamCount(config.num_streams)) unsupported_streams(config.num_streams);
  std::vector<DisortFluxConfig> cfgs(bands.size(), config);
  for (size_t i = 0; i < bands.size(); ++i) {
    DisortFluxConfig& c = cfgs[i];
    c.wavenumber_low = bands[i].first;
    c.wavenumber_high = bands[i].second;
    apply_overrides(c, ov, i, [&](DisortFluxConfig& x) {
      if (!ov.surface_albedo.empty()) x.surface_albedo = ov.surface_albedo.at(i);
      if (!ov.emissivity_top.empty()) x.emissivity_top = ov.emissivity_top.at(i);
      if (!ov.direct_beam_flux.empty()) x.direct_beam_flux = ov.direct_beam_flux.at(i);
      if (!ov.isotropic_flux_top.empty()) x.isotropic_flux_top = ov.isotropic_flux_top.at(i);
      if (!ov.isotropic_flux_bottom.empty()) x.isotropic_flux_bottom = ov.isotropic_flux_bottom.at(i);
    });
  }
  return solveFluxBatch(cfgs);
}

std::vector<FluxResult> solveFluxSpectral(DisortFluxConfig& config, const std::vector<double>& wavenumbers,
                                          const SpectralOverrides& ov) {
  std::vector<std::pair<double, double>> bands;
  bands.reserve(wavenumbers.size());
  for (double w : wavenumbers) bands.emplace_back(w, w);
  return solveFluxSpectralBands(config, bands, ov);
}

// ---------------------------------------------------------------------------
// Array-form spectral drivers
// ---------------------------------------------------------------------------
namespace {
void need_rows(const ArrayView& v, size_t bands, const char* name) {
  if (v.given() && v.n < static_cast<int64_t>(bands))
    throw std::out_of_range(std::string(name) + " has " + std::to_string(v.n) + " rows for " + std::to_string(bands) +
                            " bands");
}

void need_rows(const SpectralArrays& ov, size_t bands) {
  need_rows(ov.delta_tau, bands, "delta_tau");
  need_rows(ov.single_scat_albedo, bands, "single_scat_albedo");
  need_rows(ov.phase_function_moments, bands, "phase_function_moments");
  need_rows(ov.surface_albedo, bands, "surface_albedo");
  need_rows(ov.emissivity_top, bands, "emissivity_top");
  need_rows(ov.direct_beam_flux, bands, "direct_beam_flux");
  need_rows(ov.isotropic_flux_top, bands, "isotropic_flux_top");
  need_rows(ov.isotropic_flux_bottom, bands, "isotropic_flux_bottom");
}

double scalar_or(const ArrayView& v, size_t i, double dflt) { return v.given() ? v.data[i] : dflt; }

// Band i's configuration as the nested-list driver builds it.
DisortFluxConfig flux_band_config(const DisortFluxConfig& config, const std::pair<double, double>& band,
                                  const SpectralArrays& ov, size_t i) {
  DisortFluxConfig c = config;
  c.wavenumber_low = band.first;
  c.wavenumber_high = band.second;
  if (ov.delta_tau.given())
    c.delta_tau.assign(ov.delta_tau.data + i * ov.delta_tau.cols, ov.delta_tau.data + (i + 1) * ov.delta_tau.cols);
  if (ov.single_scat_albedo.given())
    c.single_scat_albedo.assign(ov.single_scat_albedo.data + i * ov.single_scat_albedo.cols,
                                ov.single_scat_albedo.data + (i + 1) * ov.single_scat_albedo.cols);
  if (ov.phase_function_moments.given()) {
    const ArrayView& v = ov.phase_function_moments;
    c.phase_function_moments.assign(v.rows, {});
    for (int64_t r = 0; r < v.rows; ++r)
      c.phase_function_moments[r].assign(v.data + (i * v.rows + r) * v.cols, v.data + (i * v.rows + r + 1) * v.cols);
  }
  c.surface_albedo = scalar_or(ov.surface_albedo, i, c.surface_albedo);
  c.emissivity_top = scalar_or(ov.emissivity_top, i, c.emissivity_top);
  c.direct_beam_flux = scalar_or(ov.direct_beam_flux, i, c.direct_beam_flux);
  c.isotropic_flux_top = scalar_or(ov.isotropic_flux_top, i, c.isotropic_flux_top);
  c.isotropic_flux_bottom = scalar_or(ov.isotropic_flux_bottom, i, c.isotropic_flux_bottom);
  return c;
}
}  // namespace

SpectralOverrides SpectralArrays::to_overrides(size_t bands) const {
  need_rows(*this, bands);
  SpectralOverrides o;
  auto mat = [](const ArrayView& v, std::vector<std::vector<double>>& dst) {
    if (!v.given()) return;
    dst.resize(v.n);
    for (int64_t i = 0; i < v.n; ++i) dst[i].assign(v.data + i * v.cols, v.data + (i + 1) * v.cols);
  };
  auto vec = [](const ArrayView& v, std::vector<double>& dst) {
    if (v.given()) dst.assign(v.data, v.data + v.n);
  };
  mat(delta_tau, o.delta_tau);
  mat(single_scat_albedo, o.single_scat_albedo);
  if (phase_function_moments.given()) {
    const ArrayView& v = phase_function_moments;
    o.phase_function_moments.resize(v.n);
    for (int64_t i = 0; i < v.n; ++i) {
      o.phase_function_moments[i].resize(v.rows);
      for (int64_t r = 0; r < v.rows; ++r)
        o.phase_function_moments[i][r].assign(v.data + (i * v.rows + r) * v.cols, v.data + (i * v.rows + r + 1) * v.cols);
    }
  }
  vec(surface_albedo, o.surface_albedo);
  vec(emissivity_top, o.emissivity_top);
  vec(direct_beam_flux, o.direct_beam_flux);
  vec(isotropic_flux_top, o.isotropic_flux_top);
  vec(isotropic_flux_bottom, o.isotropic_flux_bottom);
  return o;
}

// The bands share every shape (layers, streams, moment rows, flags), so the
// batch is packed from the arrays directly: one pass over the inputs, no
// per-band configuration, and rows already in the batch layout (top-down,
// num_layers entries, nmomNstr()+1 moments) are passed to the GPU copy in
// place.  Band 0 is validated as a whole configuration (the structure every
// band shares); the per-band values are validated by disort_flux_solve with
// the reference's rules, and a failing band is re-validated as a
// configuration so that the message is the reference's.
std::vector<FluxResult> solveFluxSpectralBands(DisortFluxConfig& config,
                                               const std::vector<std::pair<double, double>>& bands,
                                               const SpectralArrays& ov) {
  if (!isSupportedStreamCount(config.num_streams)) unsupported_streams(config.num_streams);
  need_rows(ov, bands.size());
  const int64_t nb = static_cast<int64_t>(bands.size());
  if (nb == 0) return {};
  validate_flux(flux_band_config(config, bands[0], ov, 0));
  const int L = config.num_layers, N = config.num_streams, P = config.nmomNstr() + 1;
  const bool rev = config.index_from_bottom;
  Packed p;
  p.flags = flux_flags(config);
  p.nstr.assign(nb, N);
  p.nlyr.assign(nb, L);
  p.pstride.assign(nb, P);
  p.lay_off.resize(nb);
  p.mom_off.resize(nb);
  for (int64_t i = 0; i < nb; ++i) {
    p.lay_off[i] = i * L;
    p.mom_off[i] = i * L * P;
  }
  p.layers = nb * L;
  p.levels = nb * (L + 1);
  p.moments = nb * L * P;
  // Layer rows: from the override array (row i) or the shared config.
  auto layer_rows = [&](const ArrayView& v, const std::vector<double>& dflt, std::vector<double>& dst,
                        const double** ext) {
    if (v.given() && !rev && v.cols == L) {
      *ext = v.data;
      return;
    }
    dst.resize(static_cast<size_t>(nb) * L);
    for (int64_t i = 0; i < nb; ++i) {
      const double* src = v.given() ? v.data + i * v.cols : dflt.data();
      const int64_t sz = v.given() ? v.cols : static_cast<int64_t>(dflt.size());
      for (int l = 0; l < L; ++l) dst[i * L + l] = src[rev ? sz - 1 - l : l];
    }
  };
  layer_rows(ov.delta_tau, config.delta_tau, p.dtau, &p.ext_dtau);
  layer_rows(ov.single_scat_albedo, config.single_scat_albedo, p.ssa, &p.ext_ssa);
  const ArrayView& pm = ov.phase_function_moments;
  if (pm.given() && !rev && pm.rows == L && pm.cols == P) {
    p.ext_pmom = pm.data;
  } else {
    p.pmom.resize(static_cast<size_t>(p.moments));
    const int64_t rows = pm.given() ? pm.rows : static_cast<int64_t>(config.phase_function_moments.size());
    for (int64_t i = 0; i < nb; ++i)
      for (int l = 0; l < L; ++l) {
        const int64_t r = rev ? rows - 1 - l : l;
        const double* src = pm.given() ? pm.data + (i * pm.rows + r) * pm.cols : config.phase_function_moments[r].data();
        std::copy(src, src + P, p.pmom.begin() + (i * L + l) * P);
      }
  }
  auto level_rows = [&](const std::vector<double>& v, std::vector<double>& dst) {
    dst.resize(static_cast<size_t>(p.levels));
    const int64_t sz = static_cast<int64_t>(v.size());
    for (int64_t i = 0; i < nb; ++i)
      for (int l = 0; l <= L; ++l) dst[i * (L + 1) + l] = v[rev ? sz - 1 - l : l];
  };
  if (p.flags & DISORT_F_THERMAL) level_rows(config.temperature, p.temper);
  if (p.flags & DISORT_F_SPHERICAL) level_rows(config.level_altitudes, p.alt);
  p.scal.resize(static_cast<size_t>(nb) * DISORT_NSCAL);
  for (int64_t i = 0; i < nb; ++i)
    fill_scal(p.scal.data() + i * DISORT_NSCAL, scalar_or(ov.direct_beam_flux, i, config.direct_beam_flux),
              config.direct_beam_mu, scalar_or(ov.isotropic_flux_top, i, config.isotropic_flux_top),
              scalar_or(ov.isotropic_flux_bottom, i, config.isotropic_flux_bottom), config.temperature_top,
              config.temperature_bottom, scalar_or(ov.emissivity_top, i, config.emissivity_top),
              scalar_or(ov.surface_albedo, i, config.surface_albedo), bands[i].first, bands[i].second,
              config.bottom_radius);
  try {
    p.solve();
    if (wants_jacobian(config)) p.jacobian();
  } catch (const std::invalid_argument& e) {
    // "instance i: <message>": re-validate band i as a configuration.
    const std::string m = e.what();
    if (m.rfind("instance ", 0) == 0) {
      const size_t i = std::stoul(m.substr(9));
      if (i < bands.size()) validate_flux(flux_band_config(config, bands[i], ov, i));
    }
    throw;
  }
  std::vector<FluxResult> results(nb);
  for (int64_t i = 0; i < nb; ++i) {
    p.unpack(i, rev, results[i]);
    if (wants_jacobian(config)) p.unpack_jacobian(i, rev, results[i]);
  }
  return results;
}

std::vector<FluxResult> solveFluxSpectral(DisortFluxConfig& config, const std::vector<double>& wavenumbers,
                                          const SpectralArrays& ov) {
  std::vector<std::pair<double, double>> bands;
  bands.reserve(wavenumbers.size());
  for (double w : wavenumbers) bands.emplace_back(w, w);
  return solveFluxSpectralBands(config, bands, ov);
}

// ---------------------------------------------------------------------------
// DisortConfig (src/DisortConfig.cpp:9-249)
// ---------------------------------------------------------------------------
DisortConfig::DisortConfig(int nlyr, int nstr, int nmom) : num_layers(nlyr), num_phase_func_moments(nmom), num_streams(nstr) {
  if (num_layers <= 0) throw std::invalid_argument("num_layers must be positive");
  if (num_streams < 4 || num_streams % 2 != 0) throw std::invalid_argument("num_streams must be even and >= 4");
  if (num_phase_func_moments < 0) throw std::invalid_argument("num_phase_func_moments must be non-negative");
}

void DisortConfig::allocate() {
  delta_tau.resize(num_layers, 0.0);
  single_scat_albedo.resize(num_layers, 0.0);
  phase_function_moments.resize(num_layers);
  for (auto& row : phase_function_moments) row.resize(nmomNstr() + 1, 0.0);
  if (flags.use_thermal_emission) temperature.resize(num_layers + 1, 0.0);
  if (num_user_tau > 0) tau_user.resize(num_user_tau, 0.0);
  if (num_user_mu > 0) mu_user.resize(num_user_mu, 0.0);
  if (!flags.comp_only_fluxes && num_phi > 0) phi_user.resize(num_phi, 0.0);
  if (flags.use_spherical_beam) {
    if (bottom_radius <= 0.0) throw std::invalid_argument("bottom_radius must be positive when use_spherical_beam=true");
    level_altitudes.resize(num_layers + 1, 0.0);
  }
  if (num_phase_func_angles > 0) {
    mu_phase_function.resize(num_phase_func_angles, 0.0);
    phase_function.assign(num_layers, std::vector<double>(num_phase_func_angles, 0.0));
  }
}

void DisortConfig::validate() const {
  auto bad = [](const std::string& m) { throw std::invalid_argument(m); };
  if (num_layers <= 0) bad("num_layers must be positive");
  if (num_streams < 4) bad("num_streams must be >= 4");
  if (num_streams % 2 != 0) bad("num_streams must be even");
  if (num_phase_func_moments < 0) bad("num_phase_func_moments must be non-negative");
  if (delta_tau.size() < static_cast<size_t>(num_layers)) bad("delta_tau array is too small");
  if (single_scat_albedo.size() < static_cast<size_t>(num_layers)) bad("single_scat_albedo array is too small");
  if (static_cast<int>(phase_function_moments.size()) < num_layers)
    bad("phase_function_moments must have at least num_layers layers");
  for (int lc = 0; lc < num_layers; ++lc)
    if (static_cast<int>(phase_function_moments[lc].size()) < nmomNstr() + 1)
      bad("phase_function_moments[" + std::to_string(lc) + "] must have at least nmomNstr()+1 elements");
  for (int lc = 0; lc < num_layers; ++lc) {
    if (delta_tau[lc] < 0.0) bad("delta_tau[" + std::to_string(lc) + "] is negative");
    if (single_scat_albedo[lc] < 0.0 || single_scat_albedo[lc] > 1.0)
      bad("single_scat_albedo[" + std::to_string(lc) + "] must be in [0, 1]");
  }
  if (flags.use_thermal_emission) {
    if (temperature.size() < static_cast<size_t>(num_layers + 1))
      bad("temperature array is too small (need num_layers+1 when use_thermal_emission=true)");
    for (int l = 0; l <= num_layers; ++l)
      if (temperature[l] < 0.0) bad("temperature[" + std::to_string(l) + "] is negative");
    if (wavenumber_low < 0.0 || wavenumber_high < wavenumber_low) bad("Invalid wavenumber range for Planck function");
  }
  if (flags.use_diffusion_lower_bc) {
    if (!flags.use_thermal_emission) bad("use_diffusion_lower_bc requires use_thermal_emission = true");
    if (!flags.use_lambertian_surface)
      bad("use_diffusion_lower_bc requires use_lambertian_surface = true (BRDF is incompatible)");
  }
  if (bc.direct_beam_flux < 0.0) bad("direct_beam_flux must be non-negative");
  if ((bc.direct_beam_mu <= 0.0 || bc.direct_beam_mu > 1.0) && bc.direct_beam_flux > 0.0)
    bad("direct_beam_mu must be in (0, 1] when direct_beam_flux > 0");
  if (flags.use_lambertian_surface && (bc.surface_albedo < 0.0 || bc.surface_albedo > 1.0))
    bad("surface_albedo must be in [0, 1] when use_lambertian_surface=true");
  if (flags.use_user_mu) {
    if (num_user_mu <= 0) bad("num_user_mu must be positive when use_user_mu=true");
    if (mu_user.size() < static_cast<size_t>(num_user_mu)) bad("mu_user array is too small");
    for (int iu = 0; iu < num_user_mu; ++iu) {
      if (std::abs(mu_user[iu]) > 1.0) bad("mu_user[" + std::to_string(iu) + "] out of range [-1, 1]");
      if (iu > 0 && mu_user[iu] <= mu_user[iu - 1]) bad("mu_user must be in increasing order");
    }
  }
  if (flags.use_spherical_beam) {
    if (bottom_radius <= 0.0) bad("bottom_radius must be positive when use_spherical_beam=true");
    if (level_altitudes.size() < static_cast<size_t>(num_layers + 1))
      bad("level_altitudes array is too small (need num_layers+1 when use_spherical_beam=true)");
  }
  if (!flags.comp_only_fluxes && flags.ibcnd != BoundaryConditionType::Special) {
    if (num_phi <= 0) bad("num_phi must be positive when comp_only_fluxes=false");
    if (phi_user.size() < static_cast<size_t>(num_phi)) bad("phi_user array is too small");
    for (int j = 0; j < num_phi; ++j)
      if (phi_user[j] < 0.0 || phi_user[j] > 360.0) bad("phi_user[" + std::to_string(j) + "] must be in [0, 360]");
  }
  if (flags.use_delta_m_plus && num_phase_func_moments < num_streams + 1)
    bad("use_delta_m_plus=true requires num_phase_func_moments >= num_streams + 1 (need moments M and M+1)");
  if (accuracy_fourier_series < 0.0) bad("accuracy_fourier_series must be non-negative (0 = disabled)");
  if (flags.use_thermal_emission) {
    if (wavenumber_low < 0.0 || wavenumber_high <= wavenumber_low)
      bad("use_thermal_emission=true requires 0 <= wavenumber_low < wavenumber_high");
    if (bc.temperature_top < 0.0 || bc.temperature_bottom < 0.0)
      bad("use_thermal_emission=true requires non-negative temperature_top and temperature_bottom");
    if (bc.emissivity_top < 0.0 || bc.emissivity_top > 1.0)
      bad("use_thermal_emission=true requires 0 <= emissivity_top <= 1");
    if (static_cast<int>(temperature.size()) < num_layers + 1)
      bad("use_thermal_emission=true requires temperature array of size num_layers+1");
  }
}

void DisortConfig::setHenyeyGreenstein(double g, int lc) { setPhaseFunction(PhaseFunction::HenyeyGreenstein, g, lc); }
void DisortConfig::setIsotropic(int lc) { setPhaseFunction(PhaseFunction::Isotropic, 0.0, lc); }
void DisortConfig::setRayleigh(int lc) { setPhaseFunction(PhaseFunction::Rayleigh, 0.0, lc); }
void DisortConfig::setHazeGarciaSiewert(int lc) { setPhaseFunction(PhaseFunction::HazeGarciaSiewert, 0.0, lc); }
void DisortConfig::setCloudGarciaSiewert(int lc) { setPhaseFunction(PhaseFunction::CloudGarciaSiewert, 0.0, lc); }
void DisortConfig::setPhaseFunction(PhaseFunction type, double g, int lc) {
  check_phase_alloc(phase_function_moments, num_layers, nmomNstr());
  phase_function::fillPhaseFunction(phase_function_moments, nmomNstr(), num_layers, type, g, lc);
}

// ---------------------------------------------------------------------------
// DisortSolver (src/DisortSolver.cpp:21-399).  Flux-only, Lambertian,
// general-BC problems at the layer boundaries go through the flux path
// (disort_flux_solve); everything else the GPU solves -- intensities, user
// optical depths, the RPV surface, the intensity corrections -- through
// disort_general_solve.
// ---------------------------------------------------------------------------
namespace {

// src/IndexConversion.hpp:16-38 (an involution: applying it twice restores).
void reverse_general_inputs(DisortConfig& c) {
  std::reverse(c.delta_tau.begin(), c.delta_tau.end());
  std::reverse(c.single_scat_albedo.begin(), c.single_scat_albedo.end());
  std::reverse(c.phase_function_moments.begin(), c.phase_function_moments.end());
  std::reverse(c.temperature.begin(), c.temperature.end());
  std::reverse(c.level_altitudes.begin(), c.level_altitudes.end());
  if (!c.phase_function.empty()) std::reverse(c.phase_function.begin(), c.phase_function.end());
  if (c.flags.use_user_tau && !c.tau_user.empty()) {
    double total = 0.0;
    for (double d : c.delta_tau) total += d;
    for (double& t : c.tau_user) t = total - t;
    std::reverse(c.tau_user.begin(), c.tau_user.end());
  }
}

bool flux_route(const DisortConfig& c) {
  return c.flags.comp_only_fluxes && c.flags.use_lambertian_surface && c.flags.ibcnd == BoundaryConditionType::General &&
         !c.flags.use_user_tau;
}

// src/DisortSolver.cpp:2143-2151: Jacobians only with thermal emission on.
bool wants_general_jacobian(const DisortConfig& c) {
  return c.flags.compute_temperature_jacobian && c.flags.use_thermal_emission &&
         c.flags.ibcnd == BoundaryConditionType::General;
}

void require_gpu_subset(const DisortConfig& c) {
  auto no = [](const std::string& what) {
    throw std::runtime_error("disortpp: " + what + " is not available on the GPU path");
  };
  if (!c.flags.use_lambertian_surface) {
    // Missing specifications fail as bidirReflectivity does
    // (src/DisortSolver.cpp:2902-2936); Hapke falls back to its defaults.
    switch (c.flags.brdf_type) {
      case BrdfType::RPV:
        if (!c.brdf.rpv) throw std::runtime_error("bidirReflectivity: RPV BRDF selected but rpv spec is null");
        break;
      case BrdfType::CoxMunk:
        if (!c.brdf.cox_munk)
          throw std::runtime_error("bidirReflectivity: Cox-Munk BRDF selected but cox_munk spec is null");
        break;
      case BrdfType::Ambrals:
        if (!c.brdf.ambrals)
          throw std::runtime_error("bidirReflectivity: Ross-Li/Ambrals BRDF selected but ambrals spec is null");
        break;
      case BrdfType::Hapke: break;
      default: throw std::runtime_error("bidirReflectivity: unknown BRDF type");
    }
  }
  if (c.flags.output_fourier_expansion) no("output_fourier_expansion");
}

// Everything DisortSolver::solve does to the config before the solve
// proper (src/DisortSolver.cpp:26-66, :641-669, :842-850); on return the
// config holds what the reference leaves in it.
void prepare_general(DisortConfig& c, std::vector<double>& cmu) {
  if (c.delta_tau.empty()) c.allocate();
  const bool rev = c.flags.index_from_bottom;
  if (rev) reverse_general_inputs(c);
  c.validate();
  require_gpu_subset(c);
  if (!isSupportedStreamCount(c.num_streams)) unsupported_streams(c.num_streams);
  const int L = c.num_layers, N = c.num_streams, NN = N / 2;
  if (!c.flags.use_user_tau) {
    c.num_user_tau = L + 1;
    c.tau_user.resize(L + 1);
    double t = 0.0;
    c.tau_user[0] = 0.0;
    for (int l = 0; l < L; ++l) c.tau_user[l + 1] = (t += c.delta_tau[l]);
  }
  if (!c.flags.use_user_mu || c.flags.comp_only_fluxes) {
    c.num_user_mu = N;
    c.mu_user.resize(N);
    std::vector<double> wt(N);
    cmu.resize(N);
    if (disort_quadrature(N, cmu.data(), wt.data()) != DISORT_OK)
      throw std::runtime_error("gaussianQuadrature: Failed to converge");
    for (int iu = 0; iu < NN; ++iu) c.mu_user[iu] = -cmu[NN - 1 - iu];
    for (int iu = NN; iu < N; ++iu) c.mu_user[iu] = cmu[iu - NN];
  }
  for (int l = 0; l < L; ++l) c.phase_function_moments[l][0] = 1.0;
  if (rev) reverse_general_inputs(c);
}

Instance general_instance(const DisortConfig& c) {
  Instance in{};
  in.nlyr = c.num_layers;
  in.nstr = c.num_streams;
  in.pstride = c.nmomNstr() + 1;
  in.reverse = c.flags.index_from_bottom;
  in.dtau = &c.delta_tau;
  in.ssa = &c.single_scat_albedo;
  in.pmom = &c.phase_function_moments;
  in.temper = &c.temperature;
  in.alt = &c.level_altitudes;
  in.pmom0_one = true;
  fill_scal(in.scal, c.bc.direct_beam_flux, c.bc.direct_beam_mu, c.bc.isotropic_flux_top, c.bc.isotropic_flux_bottom,
            c.bc.temperature_top, c.bc.temperature_bottom, c.bc.emissivity_top, c.bc.surface_albedo, c.wavenumber_low,
            c.wavenumber_high, c.bottom_radius);
  return in;
}

int32_t general_flags(const DisortConfig& c) {
  return batch_flags(c.flags.use_thermal_emission, c.flags.use_spherical_beam, c.flags.use_delta_m_plus,
                     c.flags.use_diffusion_lower_bc);
}

DisortResult to_result(const Packed& p, int64_t i, const DisortConfig& c) {
  DisortResult r;
  p.unpack(i, c.flags.index_from_bottom, r);
  r.mu_angles.assign(c.mu_user.begin(), c.mu_user.begin() + c.num_user_mu);
  return r;
}

// Full-solver batch: columns of one shape (layers, streams, moments, user
// grids, flags) packed into a disort_general_batch.
int32_t full_flags(const DisortConfig& c) {
  int32_t f = general_flags(c);
  if (c.flags.use_user_tau) f |= DISORT_G_USER_TAU;
  if (c.flags.use_user_mu) f |= DISORT_G_USER_MU;
  if (c.flags.comp_only_fluxes) f |= DISORT_G_ONLY_FLUXES;
  if (!c.flags.use_lambertian_surface) {
    switch (c.flags.brdf_type) {
      case BrdfType::Hapke: f |= DISORT_G_BRDF_HAPKE; break;
      case BrdfType::CoxMunk: f |= DISORT_G_BRDF_COXMUNK; break;
      case BrdfType::Ambrals: f |= DISORT_G_BRDF_ROSSLI; break;
      default: f |= DISORT_G_BRDF_RPV; break;
    }
  }
  if (c.flags.intensity_corr_buras) f |= DISORT_G_CORR_BURAS;
  if (c.flags.intensity_corr_nakajima) f |= DISORT_G_CORR_NAKAJIMA;
  return f;
}

struct GeneralPack {
  int32_t flags = 0;
  int32_t force_flags = 0;  // ABI flags added to every column's (DISORT_G_NO_CUTOFF)
  int N = 0, L = 0, P = 0, ntau = 0, numu = 0, nphi = 0;
  double accuracy = 0.0;
  std::vector<double> dtau, ssa, pmom, temper, alt, scal, rpv, tau_user, mu_user, phi_user, flux, inten, avg;
  std::vector<double> mu_phase, phase;  // tabulated phase function (Buras-Emde correction only)
  std::vector<int32_t> status;
  int64_t n = 0;

  // The tabulated phase function enters only the Buras-Emde correction.
  static bool uses_phase_table(const DisortConfig& c) {
    return c.flags.intensity_corr_buras && !c.flags.intensity_corr_nakajima && c.num_phase_func_angles > 0 &&
           !c.flags.comp_only_fluxes && c.flags.use_user_mu;
  }

  // Whether config c can join this batch.
  bool fits(const DisortConfig& c) const {
    if (n == 0) return true;
    return (full_flags(c) | force_flags) == flags && c.num_streams == N && c.num_layers == L && c.nmomNstr() + 1 == P &&
           c.num_user_tau == ntau && c.num_user_mu == numu && (c.flags.comp_only_fluxes ? 0 : c.num_phi) == nphi &&
           c.accuracy_fourier_series == accuracy &&
           std::equal(mu_user.begin(), mu_user.end(), c.mu_user.begin()) &&
           std::equal(phi_user.begin(), phi_user.end(), c.phi_user.begin()) &&
           (uses_phase_table(c) ? c.num_phase_func_angles == (int)mu_phase.size() &&
                                      c.mu_phase_function.size() >= mu_phase.size() &&
                                      std::equal(mu_phase.begin(), mu_phase.end(), c.mu_phase_function.begin())
                                : mu_phase.empty());
  }

  void add(const DisortConfig& cin) {
    DisortConfig c = cin;
    if (c.flags.index_from_bottom) reverse_general_inputs(c);
    if (n == 0) {
      flags = full_flags(c) | force_flags;
      N = c.num_streams;
      L = c.num_layers;
      P = c.nmomNstr() + 1;
      ntau = c.num_user_tau;
      numu = c.num_user_mu;
      nphi = c.flags.comp_only_fluxes ? 0 : c.num_phi;
      accuracy = c.accuracy_fourier_series;
      mu_user.assign(c.mu_user.begin(), c.mu_user.begin() + numu);
      phi_user.assign(c.phi_user.begin(), c.phi_user.begin() + nphi);
      if (uses_phase_table(c)) {
        if ((int)c.mu_phase_function.size() < c.num_phase_func_angles)
          throw std::invalid_argument("mu_phase_function not allocated for num_phase_func_angles");
        mu_phase.assign(c.mu_phase_function.begin(), c.mu_phase_function.begin() + c.num_phase_func_angles);
      }
    }
    if (!mu_phase.empty()) {
      const size_t np = mu_phase.size();
      if (c.mu_phase_function.size() < np || c.phase_function.size() < (size_t)L)
        throw std::invalid_argument("mu_phase_function / phase_function not allocated for num_phase_func_angles");
      for (int l = 0; l < L; ++l) {
        if (c.phase_function[l].size() < np)
          throw std::invalid_argument("phase_function[" + std::to_string(l) + "] has fewer than num_phase_func_angles values");
        phase.insert(phase.end(), c.phase_function[l].begin(), c.phase_function[l].begin() + np);
      }
    }
    dtau.insert(dtau.end(), c.delta_tau.begin(), c.delta_tau.begin() + L);
    ssa.insert(ssa.end(), c.single_scat_albedo.begin(), c.single_scat_albedo.begin() + L);
    for (int l = 0; l < L; ++l)
      for (int k = 0; k < P; ++k) pmom.push_back(k == 0 ? 1.0 : c.phase_function_moments[l][k]);
    if (flags & DISORT_F_THERMAL) temper.insert(temper.end(), c.temperature.begin(), c.temperature.begin() + L + 1);
    if (flags & DISORT_F_SPHERICAL)
      alt.insert(alt.end(), c.level_altitudes.begin(), c.level_altitudes.begin() + L + 1);
    if (flags & DISORT_G_USER_TAU) tau_user.insert(tau_user.end(), c.tau_user.begin(), c.tau_user.begin() + ntau);
    double s[DISORT_NSCAL];
    fill_scal(s, c.bc.direct_beam_flux, c.bc.direct_beam_mu, c.bc.isotropic_flux_top, c.bc.isotropic_flux_bottom,
              c.bc.temperature_top, c.bc.temperature_bottom, c.bc.emissivity_top, c.bc.surface_albedo,
              c.wavenumber_low, c.wavenumber_high, c.bottom_radius);
    s[DISORT_S_BEAM_PHI] = c.bc.direct_beam_phi;
    scal.insert(scal.end(), s, s + DISORT_NSCAL);
    if (!c.flags.use_lambertian_surface) {
      // The BRDF's parameter block (include/disort_amd.h, DISORT_NRPV slots).
      double v[DISORT_NRPV] = {0, 0, 0, 0, 0, 0, 0};
      switch (c.flags.brdf_type) {
        case BrdfType::Hapke: {
          const HapkeBrdfSpec h = c.brdf.hapke ? *c.brdf.hapke : HapkeBrdfSpec{};
          v[0] = h.b0, v[1] = h.hh, v[2] = h.w;
          break;
        }
        case BrdfType::CoxMunk: {
          const CoxMunkBrdfSpec& m = *c.brdf.cox_munk;
          v[0] = m.u10, v[1] = m.refractive_index, v[2] = m.do_shadow ? 1.0 : 0.0;
          break;
        }
        case BrdfType::Ambrals: {
          const AmbralsBrdfSpec& r = *c.brdf.ambrals;
          v[0] = r.iso, v[1] = r.vol, v[2] = r.geo;
          break;
        }
        default: {
          const RpvBrdfSpec& r = *c.brdf.rpv;
          const double w[DISORT_NRPV] = {r.rho0, r.k, r.theta, r.sigma, r.t1, r.t2, r.scale};
          std::copy(w, w + DISORT_NRPV, v);
        }
      }
      rpv.insert(rpv.end(), v, v + DISORT_NRPV);
    }
    ++n;
  }

  void solve() {
    if (n == 0) return;
    flux.assign(static_cast<size_t>(DISORT_NOUT) * n * ntau, 0.0);
    inten.assign(static_cast<size_t>(n) * ntau * numu * std::max(nphi, 0), 0.0);
    avg.assign(static_cast<size_t>(n) * ntau * numu, 0.0);
    status.assign(static_cast<size_t>(n), 0);
    disort_general_batch b = view();
    b.flux = flux.data();
    b.inten = inten.empty() ? nullptr : inten.data();
    b.inten_avg = avg.data();
    b.status = status.data();
    run([&](disort_ctx* ctx) { return disort_general_solve(ctx, &b, DISORT_MEM_HOST); });
  }

  // Temperature Jacobians of every column into jac, [column][6][ntau][L+2].
  void jacobian() {
    jac.assign(static_cast<size_t>(n) * 6 * ntau * (L + 2), 0.0);
    if (n == 0) return;
    disort_general_batch b = view();
    run([&](disort_ctx* ctx) { return disort_general_jacobian(ctx, &b, jac.data()); });
  }

  // Column i's Jacobians into the six *_temperature_jac members of r, with
  // index_from_bottom's reversal of rows and level columns
  // (src/IndexConversion.hpp:57-68).
  void unpack_jacobian(int64_t i, const DisortConfig& c, DisortResult& r) const {
    const bool rev = c.flags.index_from_bottom;
    const int nd = L + 2;
    std::vector<std::vector<double>>* f[6] = {&r.flux_up_temperature_jac, &r.flux_down_temperature_jac,
                                              &r.mean_intensity_temperature_jac,
                                              &r.mean_intensity_down_temperature_jac,
                                              &r.mean_intensity_up_temperature_jac, &r.flux_divergence_temperature_jac};
    for (int q = 0; q < 6; ++q) {
      f[q]->assign(ntau, std::vector<double>(nd, 0.0));
      const double* src = jac.data() + (static_cast<size_t>(i) * 6 + q) * ntau * nd;
      for (int lu = 0; lu < ntau; ++lu)
        for (int k = 0; k < nd; ++k)
          (*f[q])[rev ? ntau - 1 - lu : lu][rev && k <= L ? L - k : k] = src[static_cast<size_t>(lu) * nd + k];
    }
  }

  std::vector<double> jac;

 private:
  disort_general_batch view() {
    disort_general_batch b{};
    b.n = n;
    b.flags = flags;
    b.nstr = N;
    b.nlyr = L;
    b.pstride = P;
    b.ntau = ntau;
    b.numu = numu;
    b.nphi = nphi;
    b.accuracy = accuracy;
    b.dtau = dtau.data();
    b.ssa = ssa.data();
    b.pmom = pmom.data();
    b.temper = temper.empty() ? nullptr : temper.data();
    b.alt = alt.empty() ? nullptr : alt.data();
    b.scal = scal.data();
    b.rpv = rpv.empty() ? nullptr : rpv.data();
    b.tau_user = tau_user.empty() ? nullptr : tau_user.data();
    b.mu_user = mu_user.data();
    b.phi_user = phi_user.empty() ? nullptr : phi_user.data();
    b.nphase = (int32_t)mu_phase.size();
    b.mu_phase = mu_phase.empty() ? nullptr : mu_phase.data();
    b.phase = phase.empty() ? nullptr : phase.data();
    return b;
  }

  template <class F>
  void run(F call) {
    Gpu& g = gpu();
    std::lock_guard<std::mutex> lock(g.mu);
    if (!g.ctx) {
      const int dev = g.device >= 0 ? g.device : default_device();
      if (disort_ctx_create(dev, &g.ctx) != DISORT_OK) {
        g.ctx = nullptr;
        throw std::runtime_error("disortpp: cannot create a GPU context on device " + std::to_string(dev));
      }
      g.device = dev;
    }
    const int rc = call(g.ctx);
    if (rc == DISORT_OK) return;
    std::string msg = disort_last_error(g.ctx);
    if (n == 1 && msg.rfind("column 0: ", 0) == 0) msg = msg.substr(10);
    if (rc == DISORT_E_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error("disortpp: " + msg);
  }

 public:

  // Column i as a DisortResult (levels reversed for index_from_bottom,
  // src/IndexConversion.hpp:70-97).
  DisortResult unpack(int64_t i, const DisortConfig& c) const {
    DisortResult r;
    const bool rev = c.flags.index_from_bottom;
    std::vector<double>* f[DISORT_NOUT] = {&r.flux_direct_beam, &r.flux_down, &r.flux_up,
                                           &r.flux_tau_divergence, &r.mean_intensity, &r.mean_intensity_down,
                                           &r.mean_intensity_up, &r.mean_intensity_direct_beam};
    for (int q = 0; q < DISORT_NOUT; ++q) {
      f[q]->resize(ntau);
      const double* src = flux.data() + (static_cast<size_t>(q) * n + i) * ntau;
      for (int l = 0; l < ntau; ++l) (*f[q])[rev ? ntau - 1 - l : l] = src[l];
    }
    r.mu_angles.assign(c.mu_user.begin(), c.mu_user.begin() + numu);
    if (!c.flags.comp_only_fluxes) {
      r.intensity.assign(ntau, std::vector<std::vector<double>>(numu, std::vector<double>(nphi, 0.0)));
      r.intensity_azimuth_avg.assign(ntau, std::vector<double>(numu, 0.0));
      for (int lu = 0; lu < ntau; ++lu) {
        const int lo = rev ? ntau - 1 - lu : lu;
        for (int iu = 0; iu < numu; ++iu) {
          const size_t base = (static_cast<size_t>(i) * ntau + lu) * numu + iu;
          r.intensity_azimuth_avg[lo][iu] = avg[base];
          for (int j = 0; j < nphi; ++j) r.intensity[lo][iu][j] = inten[base * nphi + j];
        }
      }
    }
    return r;
  }
};

// Temperature Jacobians (src/DisortSolver.cpp:2130-2377) of the configs
// that ask for them, in full-solver batches of equal shape.
void general_jacobians(const std::vector<const DisortConfig*>& cs, const std::vector<DisortResult*>& rs) {
  std::vector<bool> done(cs.size());
  for (size_t i = 0; i < cs.size(); ++i) done[i] = !wants_general_jacobian(*cs[i]);
  for (size_t i0 = 0; i0 < cs.size(); ++i0) {
    if (done[i0]) continue;
    GeneralPack g;
    std::vector<size_t> idx;
    for (size_t i = i0; i < cs.size(); ++i)
      if (!done[i] && g.fits(*cs[i])) {
        g.add(*cs[i]);
        idx.push_back(i);
        done[i] = true;
      }
    g.jacobian();
    for (size_t j = 0; j < idx.size(); ++j) g.unpack_jacobian(static_cast<int64_t>(j), *cs[idx[j]], *rs[idx[j]]);
  }
}

// SPECIAL_BC (src/DisortSolver.cpp:4025-4163): albedo and transmissivity
// of the medium at the positive user polar angles (the quadrature nodes
// without user angles) from its response to unit isotropic illumination at
// the top and, with several layers, at the bottom, over a black surface,
// without beam, thermal sources or absorption cut-off; with surface_albedo > 0
// the spherical-albedo correction (:4126-4153).  Each illumination is one
// GPU general solve of the same medium at the doubled angles -mu .. +mu: its
// azimuthally averaged intensity at the top (bottom) is the reference's
// source integral plus the transmitted unit boundary intensity exp(-tau*/mu)
// that the reference adds, and its fluxes are pi times the spherical albedo
// and transmissivity sums.  The config is left as the reference leaves it.
std::vector<DisortResult> special_bc_solve(std::vector<DisortConfig*>& cs) {
  const size_t n = cs.size();
  std::vector<DisortConfig> base(n);
  std::vector<int> half(n);
  std::vector<std::vector<double>> ref_mu(n);
  std::vector<std::vector<int>> rank(n);
  for (size_t i = 0; i < n; ++i) {
    DisortConfig b = *cs[i];
    if (b.flags.index_from_bottom) reverse_general_inputs(b);
    std::vector<double> pos;
    if (b.flags.use_user_mu) {
      pos.assign(b.mu_user.begin(), b.mu_user.begin() + b.num_user_mu);
    } else {
      std::vector<double> cmu(b.num_streams), cwt(b.num_streams);
      if (disort_quadrature(b.num_streams, cmu.data(), cwt.data()) != DISORT_OK)
        throw std::runtime_error("gaussianQuadrature: Failed to converge");
      pos.assign(cmu.begin(), cmu.begin() + b.num_streams / 2);
    }
    const int h = static_cast<int>(pos.size());
    half[i] = h;
    // The reference's angle order [-mu_h .. -mu_1, mu_1 .. mu_h] (the
    // quadrature nodes come in decreasing order); the base solves take the
    // same cosines sorted increasing, and rank[k] places mu_k among them.
    ref_mu[i].assign(2 * h, 0.0);
    for (int k = 0; k < h; ++k) {
      ref_mu[i][h + k] = pos[k];
      ref_mu[i][k] = -pos[h - 1 - k];
    }
    std::vector<double> sorted = pos;
    std::sort(sorted.begin(), sorted.end());
    rank[i].resize(h);
    for (int k = 0; k < h; ++k)
      rank[i][k] = static_cast<int>(std::lower_bound(sorted.begin(), sorted.end(), pos[k]) - sorted.begin());
    b.flags = DisortFlags{};
    b.flags.use_user_mu = true;
    b.flags.use_delta_m_plus = cs[i]->flags.use_delta_m_plus;
    b.num_user_mu = 2 * h;
    b.mu_user.assign(2 * h, 0.0);
    for (int k = 0; k < h; ++k) {
      b.mu_user[h + k] = sorted[k];
      b.mu_user[k] = -sorted[h - 1 - k];
    }
    b.num_user_tau = b.num_layers + 1;
    b.tau_user.clear();
    b.num_phi = 1;
    b.phi_user.assign(1, 0.0);
    b.accuracy_fourier_series = 0.0;
    b.num_phase_func_angles = 0;
    b.bc = BoundaryConditions{};
    b.bc.isotropic_flux_top = 1.0;
    b.bc.emissivity_top = cs[i]->bc.emissivity_top;
    std::vector<double> cmu;
    prepare_general(b, cmu);
    base[i] = std::move(b);
  }
  // One GPU general batch per illumination and shape.
  auto run = [&](std::vector<DisortConfig>& cfgs, const std::vector<size_t>& which) {
    std::vector<DisortResult> out(cfgs.size());
    std::vector<bool> done(cfgs.size(), true);
    for (size_t i : which) done[i] = false;
    for (size_t i0 = 0; i0 < cfgs.size(); ++i0) {
      if (done[i0]) continue;
      GeneralPack g;
      g.force_flags = DISORT_G_NO_CUTOFF;
      std::vector<size_t> idx;
      for (size_t i = i0; i < cfgs.size(); ++i)
        if (!done[i] && g.fits(cfgs[i])) {
          g.add(cfgs[i]);
          idx.push_back(i);
          done[i] = true;
        }
      g.solve();
      for (size_t j = 0; j < idx.size(); ++j) out[idx[j]] = g.unpack(static_cast<int64_t>(j), cfgs[idx[j]]);
    }
    return out;
  };
  std::vector<size_t> all(n), multi;
  for (size_t i = 0; i < n; ++i) {
    all[i] = i;
    if (base[i].num_layers > 1) multi.push_back(i);
  }
  const std::vector<DisortResult> top = run(base, all);
  std::vector<DisortConfig> bottom = base;
  for (size_t i : multi) {
    bottom[i].bc.isotropic_flux_top = 0.0;
    bottom[i].bc.isotropic_flux_bottom = 1.0;
  }
  const std::vector<DisortResult> bot = run(bottom, multi);

  std::vector<DisortResult> res(n);
  for (size_t i = 0; i < n; ++i) {
    DisortConfig& c = *cs[i];
    const int h = half[i], L = c.num_layers;
    DisortResult& r = res[i];
    const int ntau = c.flags.use_user_tau ? c.num_user_tau : L + 1;
    for (auto* f : {&r.flux_direct_beam, &r.flux_down, &r.flux_up, &r.flux_tau_divergence, &r.mean_intensity,
                    &r.mean_intensity_down, &r.mean_intensity_up, &r.mean_intensity_direct_beam})
      f->assign(ntau, 0.0);
    r.mu_angles = ref_mu[i];
    if (!c.flags.comp_only_fluxes) {
      r.intensity.assign(ntau, std::vector<std::vector<double>>(2 * h, std::vector<double>(c.num_phi, 0.0)));
      r.intensity_azimuth_avg.assign(ntau, std::vector<double>(2 * h, 0.0));
    }
    r.albedo_medium.assign(2 * h, 0.0);
    r.transmissivity_medium.assign(2 * h, 0.0);
    const DisortResult& a = top[i];
    const std::vector<int>& rk = rank[i];
    for (int k = 0; k < h; ++k) r.albedo_medium[k] = a.intensity_azimuth_avg[0][h + rk[k]];
    const DisortResult* sph = &a;
    if (L == 1) {
      for (int k = 0; k < h; ++k) r.transmissivity_medium[k] = a.intensity_azimuth_avg[L][h - 1 - rk[k]];
    } else {
      for (int k = 0; k < h; ++k) r.transmissivity_medium[k] = bot[i].intensity_azimuth_avg[0][h + rk[k]];
      sph = &bot[i];
    }
    const double alb = c.bc.surface_albedo;
    if (alb > 0.0) {
      const double sflup = sph->flux_up[0] / M_PI, sfldn = sph->flux_down[L] / M_PI;
      const double sphalb = (L == 1) ? sflup : sfldn, sphtrn = (L == 1) ? sfldn : sflup;
      const double corr = alb / (1.0 - alb * sphalb);
      for (int k = 0; k < h; ++k) {
        const double t = r.transmissivity_medium[k];
        r.albedo_medium[k] += corr * sphtrn * t;
        r.transmissivity_medium[k] += corr * sphalb * t;
      }
      r.spherical_albedo = sphalb;
      r.spherical_transmissivity = sphtrn;
    }
    // What the reference leaves in the config (:48-66, :842-865,
    // :4155-4162): user levels at the layer boundaries, num_user_mu = h,
    // and mu_user the doubled angles (2h + 1 entries with user angles, the
    // 2h quadrature angles without) whose entries [1..h] are then taken
    // from [1+h..2h] (a 1-based copy loop over a 0-based array).
    if (!c.flags.use_user_tau) {
      c.num_user_tau = L + 1;
      c.tau_user.resize(L + 1);
    }
    std::vector<double> m = ref_mu[i];
    if (c.flags.use_user_mu) m.resize(2 * h + 1, 0.0);
    for (int iu = 1; iu <= h && iu + h < static_cast<int>(m.size()); ++iu) m[iu] = m[iu + h];
    c.mu_user = m;
    c.num_user_mu = h;
  }
  return res;
}

}  // namespace

DisortResult DisortSolver::solve(DisortConfig& config) {
  std::vector<double> cmu;
  if (config.flags.ibcnd == BoundaryConditionType::Special) {
    if (config.delta_tau.empty()) config.allocate();
    config.validate();
    std::vector<DisortConfig*> one{&config};
    return special_bc_solve(one)[0];
  }
  prepare_general(config, cmu);
  DisortResult r;
  if (flux_route(config)) {
    Packed p;
    p.flags = general_flags(config);
    p.add(general_instance(config));
    p.solve();
    r = to_result(p, 0, config);
  } else {
    GeneralPack g;
    g.add(config);
    g.solve();
    r = g.unpack(0, config);
  }
  general_jacobians({&config}, {&r});
  return r;
}

std::vector<DisortResult> solveSpectralBands(DisortConfig config,
                                             const std::vector<std::pair<double, double>>& bands,
                                             const SpectralOverrides& ov) {
  std::vector<DisortConfig> cfgs(bands.size(), config);
  std::vector<double> cmu;
  for (size_t i = 0; i < bands.size(); ++i) {
    DisortConfig& c = cfgs[i];
    c.wavenumber_low = bands[i].first;
    c.wavenumber_high = bands[i].second;
    apply_overrides(c, ov, i, [&](DisortConfig& x) {
      if (!ov.surface_albedo.empty()) x.bc.surface_albedo = ov.surface_albedo.at(i);
      if (!ov.emissivity_top.empty()) x.bc.emissivity_top = ov.emissivity_top.at(i);
      if (!ov.direct_beam_flux.empty()) x.bc.direct_beam_flux = ov.direct_beam_flux.at(i);
      if (!ov.isotropic_flux_top.empty()) x.bc.isotropic_flux_top = ov.isotropic_flux_top.at(i);
      if (!ov.isotropic_flux_bottom.empty()) x.bc.isotropic_flux_bottom = ov.isotropic_flux_bottom.at(i);
    });
    if (c.flags.ibcnd == BoundaryConditionType::Special) {
      if (c.delta_tau.empty()) c.allocate();
      c.validate();
    } else {
      prepare_general(c, cmu);
    }
  }
  std::vector<DisortResult> results(bands.size());
  // SPECIAL_BC bands: albedo and transmissivity of the medium.
  std::vector<DisortConfig*> special;
  std::vector<size_t> special_idx;
  for (size_t i = 0; i < cfgs.size(); ++i)
    if (cfgs[i].flags.ibcnd == BoundaryConditionType::Special) {
      special.push_back(&cfgs[i]);
      special_idx.push_back(i);
    }
  if (!special.empty()) {
    std::vector<DisortResult> r = special_bc_solve(special);
    for (size_t j = 0; j < special_idx.size(); ++j) results[special_idx[j]] = std::move(r[j]);
  }
  // Flux-only bands: one flux batch per batch-uniform flag set.
  std::vector<int32_t> seen;
  for (const auto& c : cfgs)
    if (flux_route(c) && std::find(seen.begin(), seen.end(), general_flags(c)) == seen.end())
      seen.push_back(general_flags(c));
  for (int32_t fl : seen) {
    Packed p;
    p.flags = fl;
    std::vector<size_t> idx;
    for (size_t i = 0; i < cfgs.size(); ++i)
      if (flux_route(cfgs[i]) && general_flags(cfgs[i]) == fl) {
        p.add(general_instance(cfgs[i]));
        idx.push_back(i);
      }
    p.solve();
    for (size_t j = 0; j < idx.size(); ++j) results[idx[j]] = to_result(p, static_cast<int64_t>(j), cfgs[idx[j]]);
  }
  // The rest: full-solver batches of equal shape.
  std::vector<bool> done(cfgs.size(), false);
  for (size_t i = 0; i < cfgs.size(); ++i)
    done[i] = flux_route(cfgs[i]) || cfgs[i].flags.ibcnd == BoundaryConditionType::Special;
  for (size_t i0 = 0; i0 < cfgs.size(); ++i0) {
    if (done[i0]) continue;
    GeneralPack g;
    std::vector<size_t> idx;
    for (size_t i = i0; i < cfgs.size(); ++i)
      if (!done[i] && g.fits(cfgs[i])) {
        g.add(cfgs[i]);
        idx.push_back(i);
        done[i] = true;
      }
    g.solve();
    for (size_t j = 0; j < idx.size(); ++j) results[idx[j]] = g.unpack(static_cast<int64_t>(j), cfgs[idx[j]]);
  }
  std::vector<const DisortConfig*> cp;
  std::vector<DisortResult*> rp;
  for (size_t i = 0; i < cfgs.size(); ++i) {
    cp.push_back(&cfgs[i]);
    rp.push_back(&results[i]);
  }
  general_jacobians(cp, rp);
  return results;
}

std::vector<DisortResult> solveSpectral(DisortConfig config, const std::vector<double>& wavenumbers,
                                        const SpectralOverrides& ov) {
  std::vector<std::pair<double, double>> bands;
  bands.reserve(wavenumbers.size());
  for (double w : wavenumbers) bands.emplace_back(w, w);
  return solveSpectralBands(std::move(config), bands, ov);
}

// The general drivers build a configuration per band anyway (user grids,
// intensities, mutation of the config), so the arrays become the nested
// form in C++.
std::vector<DisortResult> solveSpectralBands(DisortConfig config, const std::vector<std::pair<double, double>>& bands,
                                             const SpectralArrays& ov) {
  return solveSpectralBands(std::move(config), bands, ov.to_overrides(bands.size()));
}

std::vector<DisortResult> solveSpectral(DisortConfig config, const std::vector<double>& wavenumbers,
                                        const SpectralArrays& ov) {
  return solveSpectral(std::move(config), wavenumbers, ov.to_overrides(wavenumbers.size()));
}

}  // namespace disortpp
