// bindings.cpp -- pybind11 module `disortpp` with the reference module's
// names and signatures (src/python_bindings.cpp:213-763), backed by the
// GPU facade in disortpp.cpp.  Solves release the GIL.
//
// Vector fields of the configs read back as tuples and are assigned as a
// whole (cfg.delta_tau = [...]), as in the reference module: element
// assignment on a read-back copy would silently do nothing, so the tuple
// makes it an error instead.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "disortpp.hpp"

namespace py = pybind11;
using namespace disortpp;

namespace {

template <class C, class V>
void vec_property(py::class_<C>& cls, const char* name, std::vector<V> C::*field) {
  cls.def_property(
      name, [field](const C& c) { return py::tuple(py::cast(c.*field)); },
      [field](C& c, const std::vector<V>& v) { c.*field = v; });
}

template <class C>
void mat_property(py::class_<C>& cls, const char* name, std::vector<std::vector<double>> C::*field) {
  cls.def_property(
      name,
      [field](const C& c) {
        py::list rows;
        for (const auto& r : c.*field) rows.append(py::tuple(py::cast(r)));
        return py::tuple(rows);
      },
      [field](C& c, const std::vector<std::vector<double>>& v) { c.*field = v; });
}

template <class C>
void phase_setters(py::class_<C>& cls) {
  cls.def("nmom_nstr", &C::nmomNstr)
      .def("allocate", &C::allocate)
      .def("validate", &C::validate)
      .def("set_henyey_greenstein", &C::setHenyeyGreenstein, py::arg("g"), py::arg("lc") = -1)
      .def("set_isotropic", &C::setIsotropic, py::arg("lc") = -1)
      .def("set_rayleigh", &C::setRayleigh, py::arg("lc") = -1)
      .def("set_haze_garcia_siewert", &C::setHazeGarciaSiewert, py::arg("lc") = -1)
      .def("set_cloud_garcia_siewert", &C::setCloudGarciaSiewert, py::arg("lc") = -1)
      .def("set_phase_function", &C::setPhaseFunction, py::arg("type"), py::arg("g") = 0.0, py::arg("lc") = -1);
}

template <class R>
void flux_fields(py::class_<R>& cls) {
  cls.def_readwrite("flux_direct_beam", &R::flux_direct_beam)
      .def_readwrite("flux_down", &R::flux_down)
      .def_readwrite("flux_up", &R::flux_up)
      .def_readwrite("flux_tau_divergence", &R::flux_tau_divergence)
      .def_readwrite("mean_intensity", &R::mean_intensity)
      .def_readwrite("mean_intensity_down", &R::mean_intensity_down)
      .def_readwrite("mean_intensity_up", &R::mean_intensity_up)
      .def_readwrite("mean_intensity_direct_beam", &R::mean_intensity_direct_beam)
      .def_readwrite("flux_up_temperature_jac", &R::flux_up_temperature_jac)
      .def_readwrite("flux_down_temperature_jac", &R::flux_down_temperature_jac)
      .def_readwrite("mean_intensity_temperature_jac", &R::mean_intensity_temperature_jac)
      .def_readwrite("mean_intensity_down_temperature_jac", &R::mean_intensity_down_temperature_jac)
      .def_readwrite("mean_intensity_up_temperature_jac", &R::mean_intensity_up_temperature_jac)
      .def_readwrite("flux_divergence_temperature_jac", &R::flux_divergence_temperature_jac);
}

template <int N>
void flux_solver_class(py::module_& m) {
  const std::string name = "DisortFluxSolver" + std::to_string(N);
  py::class_<DisortFluxSolver<N>>(m, name.c_str(), ("Flux-only solver with " + std::to_string(N) + " streams").c_str())
      .def(py::init<>())
      .def(
          "solve",
          [](DisortFluxSolver<N>& self, const DisortFluxConfig& config) {
            py::gil_scoped_release release;
            return self.solve(config);
          },
          py::arg("config"));
}

using Mat = std::vector<std::vector<double>>;
using Cube = std::vector<std::vector<std::vector<double>>>;
using Vec = std::vector<double>;

#define DISORTPP_OVERRIDE_ARGS                                                                                \
  py::arg("delta_tau") = py::none(), py::arg("single_scat_albedo") = py::none(),                                \
      py::arg("phase_function_moments") = py::none(), py::arg("surface_albedo") = py::none(),                   \
      py::arg("emissivity_top") = py::none(), py::arg("direct_beam_flux") = py::none(),                         \
      py::arg("isotropic_flux_top") = py::none(), py::arg("isotropic_flux_bottom") = py::none()

using F64 = py::array_t<double, py::array::c_style | py::array::forcecast>;

// The eight per-band override arguments, in either form.  None or an empty
// sequence means "not given" (the reference's default, an empty list).  When
// every given argument converts to a float64 C-contiguous array of its rank
// (2, 2, 3, then 1 for the scalars) the arrays are used in place (NumPy
// arrays of that dtype and layout are not copied); ragged nested lists take
// the nested-list form.
struct Overrides {
  std::vector<F64> keep;
  SpectralArrays arrays;
  bool as_arrays = true;
  SpectralOverrides lists;

  Overrides(py::object dtau, py::object ssa, py::object pmom, py::object alb, py::object emis, py::object beam,
            py::object itop, py::object ibot) {
    py::object src[8] = {dtau, ssa, pmom, alb, emis, beam, itop, ibot};
    const int rank[8] = {2, 2, 3, 1, 1, 1, 1, 1};
    ArrayView* dst[8] = {&arrays.delta_tau,      &arrays.single_scat_albedo, &arrays.phase_function_moments,
                         &arrays.surface_albedo, &arrays.emissivity_top,     &arrays.direct_beam_flux,
                         &arrays.isotropic_flux_top, &arrays.isotropic_flux_bottom};
    bool given[8];
    for (int k = 0; k < 8; ++k) given[k] = !src[k].is_none() && py::len(src[k]) > 0;
    for (int k = 0; k < 8; ++k) {
      if (!given[k]) continue;
      F64 a = F64::ensure(src[k]);
      if (!a || a.ndim() != rank[k]) {
        as_arrays = false;
        break;
      }
      keep.push_back(a);
      ArrayView& v = *dst[k];
      v.data = a.data();
      v.n = a.shape(0);
      if (rank[k] >= 2) v.cols = a.shape(rank[k] - 1);
      if (rank[k] == 3) v.rows = a.shape(1);
    }
    if (as_arrays) return;
    keep.clear();
    arrays = SpectralArrays{};
    if (given[0]) lists.delta_tau = src[0].cast<Mat>();
    if (given[1]) lists.single_scat_albedo = src[1].cast<Mat>();
    if (given[2]) lists.phase_function_moments = src[2].cast<Cube>();
    if (given[3]) lists.surface_albedo = src[3].cast<Vec>();
    if (given[4]) lists.emissivity_top = src[4].cast<Vec>();
    if (given[5]) lists.direct_beam_flux = src[5].cast<Vec>();
    if (given[6]) lists.isotropic_flux_top = src[6].cast<Vec>();
    if (given[7]) lists.isotropic_flux_bottom = src[7].cast<Vec>();
  }
};

}  // namespace

PYBIND11_MODULE(disortpp, m) {
  m.doc() = "DISORT discrete-ordinates radiative transfer on AMD Instinct MI355X (gfx950), DisORT++ API";
  m.attr("__version__") = "1.0";

  py::enum_<PhaseFunction>(m, "PhaseFunction")
      .value("Isotropic", PhaseFunction::Isotropic)
      .value("Rayleigh", PhaseFunction::Rayleigh)
      .value("HenyeyGreenstein", PhaseFunction::HenyeyGreenstein)
      .value("HazeGarciaSiewert", PhaseFunction::HazeGarciaSiewert)
      .value("CloudGarciaSiewert", PhaseFunction::CloudGarciaSiewert);
  py::enum_<BoundaryConditionType>(m, "BoundaryConditionType")
      .value("General", BoundaryConditionType::General)
      .value("Special", BoundaryConditionType::Special);
  py::enum_<BrdfType>(m, "BrdfType")
      .value("NONE", BrdfType::None)
      .value("RPV", BrdfType::RPV)
      .value("CoxMunk", BrdfType::CoxMunk)
      .value("Ambrals", BrdfType::Ambrals)
      .value("Hapke", BrdfType::Hapke);

  py::class_<DisortFlags>(m, "DisortFlags")
      .def(py::init<>())
      .def_readwrite("use_user_tau", &DisortFlags::use_user_tau)
      .def_readwrite("use_user_mu", &DisortFlags::use_user_mu)
      .def_readwrite("ibcnd", &DisortFlags::ibcnd)
      .def_readwrite("use_lambertian_surface", &DisortFlags::use_lambertian_surface)
      .def_readwrite("use_thermal_emission", &DisortFlags::use_thermal_emission)
      .def_readwrite("use_spherical_beam", &DisortFlags::use_spherical_beam)
      .def_readwrite("comp_only_fluxes", &DisortFlags::comp_only_fluxes)
      .def_readwrite("brdf_type", &DisortFlags::brdf_type)
      .def_readwrite("intensity_corr_buras", &DisortFlags::intensity_corr_buras)
      .def_readwrite("intensity_corr_nakajima", &DisortFlags::intensity_corr_nakajima)
      .def_readwrite("use_delta_m_plus", &DisortFlags::use_delta_m_plus)
      .def_readwrite("use_diffusion_lower_bc", &DisortFlags::use_diffusion_lower_bc)
      .def_readwrite("output_fourier_expansion", &DisortFlags::output_fourier_expansion)
      .def_readwrite("index_from_bottom", &DisortFlags::index_from_bottom)
      .def_readwrite("compute_temperature_jacobian", &DisortFlags::compute_temperature_jacobian);

  py::class_<BoundaryConditions>(m, "BoundaryConditions")
      .def(py::init<>())
      .def_readwrite("direct_beam_flux", &BoundaryConditions::direct_beam_flux)
      .def_readwrite("direct_beam_mu", &BoundaryConditions::direct_beam_mu)
      .def_readwrite("direct_beam_phi", &BoundaryConditions::direct_beam_phi)
      .def_readwrite("isotropic_flux_top", &BoundaryConditions::isotropic_flux_top)
      .def_readwrite("isotropic_flux_bottom", &BoundaryConditions::isotropic_flux_bottom)
      .def_readwrite("temperature_top", &BoundaryConditions::temperature_top)
      .def_readwrite("temperature_bottom", &BoundaryConditions::temperature_bottom)
      .def_readwrite("emissivity_top", &BoundaryConditions::emissivity_top)
      .def_readwrite("surface_albedo", &BoundaryConditions::surface_albedo);

  py::class_<RpvBrdfSpec>(m, "RpvBrdfSpec")
      .def(py::init<>())
      .def_readwrite("rho0", &RpvBrdfSpec::rho0)
      .def_readwrite("k", &RpvBrdfSpec::k)
      .def_readwrite("theta", &RpvBrdfSpec::theta)
      .def_readwrite("sigma", &RpvBrdfSpec::sigma)
      .def_readwrite("t1", &RpvBrdfSpec::t1)
      .def_readwrite("t2", &RpvBrdfSpec::t2)
      .def_readwrite("scale", &RpvBrdfSpec::scale);
  py::class_<AmbralsBrdfSpec>(m, "AmbralsBrdfSpec")
      .def(py::init<>())
      .def_readwrite("iso", &AmbralsBrdfSpec::iso)
      .def_readwrite("vol", &AmbralsBrdfSpec::vol)
      .def_readwrite("geo", &AmbralsBrdfSpec::geo);
  py::class_<CoxMunkBrdfSpec>(m, "CoxMunkBrdfSpec")
      .def(py::init<>())
      .def_readwrite("u10", &CoxMunkBrdfSpec::u10)
      .def_readwrite("pcl", &CoxMunkBrdfSpec::pcl)
      .def_readwrite("xsal", &CoxMunkBrdfSpec::xsal)
      .def_readwrite("refractive_index", &CoxMunkBrdfSpec::refractive_index)
      .def_readwrite("do_shadow", &CoxMunkBrdfSpec::do_shadow);
  py::class_<HapkeBrdfSpec>(m, "HapkeBrdfSpec")
      .def(py::init<>())
      .def_readwrite("b0", &HapkeBrdfSpec::b0)
      .def_readwrite("hh", &HapkeBrdfSpec::hh)
      .def_readwrite("w", &HapkeBrdfSpec::w);
  py::class_<BrdfSpecification>(m, "BrdfSpecification")
      .def(py::init<>())
      .def_readwrite("rpv", &BrdfSpecification::rpv)
      .def_readwrite("ambrals", &BrdfSpecification::ambrals)
      .def_readwrite("cox_munk", &BrdfSpecification::cox_munk)
      .def_readwrite("hapke", &BrdfSpecification::hapke)
      .def("set_rpv", &BrdfSpecification::setRpv, py::arg("spec"))
      .def("set_ambrals", &BrdfSpecification::setAmbrals, py::arg("spec"))
      .def("set_cox_munk", &BrdfSpecification::setCoxMunk, py::arg("spec"))
      .def("set_hapke", &BrdfSpecification::setHapke, py::arg("spec"));

  py::class_<DisortConfig> dc(m, "DisortConfig");
  dc.def(py::init<>())
      .def(py::init<int, int, int>(), py::arg("num_layers"), py::arg("num_streams"),
           py::arg("num_phase_func_moments") = 0)
      .def_readwrite("flags", &DisortConfig::flags)
      .def_readwrite("bc", &DisortConfig::bc)
      .def_readwrite("brdf", &DisortConfig::brdf)
      .def_readwrite("num_layers", &DisortConfig::num_layers)
      .def_readwrite("num_phase_func_moments", &DisortConfig::num_phase_func_moments)
      .def_readwrite("num_streams", &DisortConfig::num_streams)
      .def_readwrite("num_user_tau", &DisortConfig::num_user_tau)
      .def_readwrite("num_user_mu", &DisortConfig::num_user_mu)
      .def_readwrite("num_phi", &DisortConfig::num_phi)
      .def_readwrite("num_phase_func_angles", &DisortConfig::num_phase_func_angles)
      .def_readwrite("wavenumber_low", &DisortConfig::wavenumber_low)
      .def_readwrite("wavenumber_high", &DisortConfig::wavenumber_high)
      .def_readwrite("accuracy_fourier_series", &DisortConfig::accuracy_fourier_series)
      .def_readwrite("bottom_radius", &DisortConfig::bottom_radius);
  vec_property(dc, "delta_tau", &DisortConfig::delta_tau);
  vec_property(dc, "single_scat_albedo", &DisortConfig::single_scat_albedo);
  mat_property(dc, "phase_function_moments", &DisortConfig::phase_function_moments);
  vec_property(dc, "temperature", &DisortConfig::temperature);
  vec_property(dc, "level_altitudes", &DisortConfig::level_altitudes);
  vec_property(dc, "tau_user", &DisortConfig::tau_user);
  vec_property(dc, "mu_user", &DisortConfig::mu_user);
  vec_property(dc, "phi_user", &DisortConfig::phi_user);
  vec_property(dc, "mu_phase_function", &DisortConfig::mu_phase_function);
  mat_property(dc, "phase_function", &DisortConfig::phase_function);
  phase_setters(dc);

  py::class_<DisortFluxConfig> fc(m, "DisortFluxConfig");
  fc.def(py::init<>())
      .def(py::init<int, int, int>(), py::arg("num_layers"), py::arg("num_streams"),
           py::arg("num_phase_func_moments") = 0)
      .def_readwrite("num_layers", &DisortFluxConfig::num_layers)
      .def_readwrite("num_streams", &DisortFluxConfig::num_streams)
      .def_readwrite("num_phase_func_moments", &DisortFluxConfig::num_phase_func_moments)
      .def_readwrite("use_thermal_emission", &DisortFluxConfig::use_thermal_emission)
      .def_readwrite("use_spherical_beam", &DisortFluxConfig::use_spherical_beam)
      .def_readwrite("use_delta_m_plus", &DisortFluxConfig::use_delta_m_plus)
      .def_readwrite("use_diffusion_lower_bc", &DisortFluxConfig::use_diffusion_lower_bc)
      .def_readwrite("index_from_bottom", &DisortFluxConfig::index_from_bottom)
      .def_readwrite("compute_temperature_jacobian", &DisortFluxConfig::compute_temperature_jacobian)
      .def_readwrite("direct_beam_flux", &DisortFluxConfig::direct_beam_flux)
      .def_readwrite("direct_beam_mu", &DisortFluxConfig::direct_beam_mu)
      .def_readwrite("isotropic_flux_top", &DisortFluxConfig::isotropic_flux_top)
      .def_readwrite("isotropic_flux_bottom", &DisortFluxConfig::isotropic_flux_bottom)
      .def_readwrite("temperature_top", &DisortFluxConfig::temperature_top)
      .def_readwrite("temperature_bottom", &DisortFluxConfig::temperature_bottom)
      .def_readwrite("emissivity_top", &DisortFluxConfig::emissivity_top)
      .def_readwrite("surface_albedo", &DisortFluxConfig::surface_albedo)
      .def_readwrite("wavenumber_low", &DisortFluxConfig::wavenumber_low)
      .def_readwrite("wavenumber_high", &DisortFluxConfig::wavenumber_high)
      .def_readwrite("bottom_radius", &DisortFluxConfig::bottom_radius);
  vec_property(fc, "delta_tau", &DisortFluxConfig::delta_tau);
  vec_property(fc, "single_scat_albedo", &DisortFluxConfig::single_scat_albedo);
  mat_property(fc, "phase_function_moments", &DisortFluxConfig::phase_function_moments);
  vec_property(fc, "temperature", &DisortFluxConfig::temperature);
  vec_property(fc, "level_altitudes", &DisortFluxConfig::level_altitudes);
  phase_setters(fc);

  py::class_<DisortResult> dr(m, "DisortResult");
  dr.def(py::init<>());
  flux_fields(dr);
  dr.def_readwrite("albedo_medium", &DisortResult::albedo_medium)
      .def_readwrite("transmissivity_medium", &DisortResult::transmissivity_medium)
      .def_readwrite("spherical_albedo", &DisortResult::spherical_albedo)
      .def_readwrite("spherical_transmissivity", &DisortResult::spherical_transmissivity)
      .def_readwrite("mu_angles", &DisortResult::mu_angles)
      .def_readwrite("intensity", &DisortResult::intensity)
      .def_readwrite("intensity_azimuth_avg", &DisortResult::intensity_azimuth_avg)
      .def_readwrite("intensity_fourier_expansion", &DisortResult::intensity_fourier_expansion)
      .def("num_user_tau", &DisortResult::num_user_tau)
      .def("num_user_mu", &DisortResult::num_user_mu)
      .def("num_phi", &DisortResult::num_phi)
      .def("total_flux_down", &DisortResult::totalFluxDown, py::arg("lu"))
      .def("net_flux", &DisortResult::netFlux, py::arg("lu"))
      .def("intensities", &DisortResult::intensities, py::arg("iu"), py::arg("lu"), py::arg("j"))
      .def("intensities_azimuth_avg", &DisortResult::intensitiesAzimuthAvg, py::arg("iu"), py::arg("lu"));

  py::class_<FluxResult> fr(m, "FluxResult");
  fr.def(py::init<>());
  flux_fields(fr);
  fr.def("num_levels", &FluxResult::num_levels)
      .def("total_flux_down", &FluxResult::totalFluxDown, py::arg("lev"))
      .def("net_flux", &FluxResult::netFlux, py::arg("lev"));

  py::class_<DisortSolver>(m, "DisortSolver", "General discrete ordinates radiative transfer solver")
      .def(py::init<>())
      .def(
          "solve",
          [](DisortSolver& self, DisortConfig& config) {
            py::gil_scoped_release release;
            return self.solve(config);
          },
          py::arg("config"),
          "Solve the radiative transfer problem. Note: config may be modified by the solver to set automatic "
          "dimensions.");

  flux_solver_class<4>(m);
  flux_solver_class<6>(m);
  flux_solver_class<8>(m);
  flux_solver_class<10>(m);
  flux_solver_class<12>(m);
  flux_solver_class<14>(m);
  flux_solver_class<16>(m);
  flux_solver_class<32>(m);
  flux_solver_class<64>(m);

  m.def(
      "create_flux_solver",
      [](int nstr) -> py::object {
        switch (nstr) {
          case 4: return py::cast(DisortFluxSolver<4>());
          case 6: return py::cast(DisortFluxSolver<6>());
          case 8: return py::cast(DisortFluxSolver<8>());
          case 10: return py::cast(DisortFluxSolver<10>());
          case 12: return py::cast(DisortFluxSolver<12>());
          case 14: return py::cast(DisortFluxSolver<14>());
          case 16: return py::cast(DisortFluxSolver<16>());
          case 32: return py::cast(DisortFluxSolver<32>());
          case 64: return py::cast(DisortFluxSolver<64>());
        }
        throw std::invalid_argument("Unsupported nstr=" + std::to_string(nstr) +
                                    ". Supported values: 4, 6, 8, 10, 12, 14, 16, 32, 64");
      },
      py::arg("nstr"), "Create a flux-only solver for the given number of streams (4, 6, 8, 10, 12, 14, 16, 32, 64)");

  m.def(
      "solve_spectral",
      [](DisortConfig config, const Vec& wavenumbers, py::object dtau, py::object ssa, py::object pmom, py::object alb,
         py::object emis, py::object beam, py::object itop, py::object ibot) {
        Overrides ov(dtau, ssa, pmom, alb, emis, beam, itop, ibot);
        py::gil_scoped_release release;
        return ov.as_arrays ? solveSpectral(std::move(config), wavenumbers, ov.arrays)
                            : solveSpectral(std::move(config), wavenumbers, ov.lists);
      },
      py::arg("config"), py::arg("wavenumbers"), DISORTPP_OVERRIDE_ARGS,
      "Solve at single wavenumbers (wavenumber_low = wavenumber_high = wn), one GPU batch; returns DisortResult "
      "list. Per-band overrides are nested lists or NumPy arrays ([bands][layers], [bands][layers][moments], "
      "[bands]).");

  m.def(
      "solve_spectral_bands",
      [](DisortConfig config, const std::vector<std::pair<double, double>>& bands, py::object dtau, py::object ssa,
         py::object pmom, py::object alb, py::object emis, py::object beam, py::object itop, py::object ibot) {
        Overrides ov(dtau, ssa, pmom, alb, emis, beam, itop, ibot);
        py::gil_scoped_release release;
        return ov.as_arrays ? solveSpectralBands(std::move(config), bands, ov.arrays)
                            : solveSpectralBands(std::move(config), bands, ov.lists);
      },
      py::arg("config"), py::arg("wavenumber_bands"), DISORTPP_OVERRIDE_ARGS,
      "Solve for (low, high) wavenumber bands as one GPU batch; returns DisortResult list. Per-band overrides "
      "are nested lists or NumPy arrays.");

  m.def(
      "solve_flux_spectral",
      [](DisortFluxConfig& config, const Vec& wavenumbers, py::object dtau, py::object ssa, py::object pmom,
         py::object alb, py::object emis, py::object beam, py::object itop, py::object ibot) {
        Overrides ov(dtau, ssa, pmom, alb, emis, beam, itop, ibot);
        py::gil_scoped_release release;
        return ov.as_arrays ? solveFluxSpectral(config, wavenumbers, ov.arrays)
                            : solveFluxSpectral(config, wavenumbers, ov.lists);
      },
      py::arg("config"), py::arg("wavenumbers"), DISORTPP_OVERRIDE_ARGS,
      "Flux-only solve at single wavenumbers as one GPU batch; returns FluxResult list. NumPy override arrays "
      "are packed into the GPU batch directly.");

  m.def(
      "solve_flux_spectral_bands",
      [](DisortFluxConfig& config, const std::vector<std::pair<double, double>>& bands, py::object dtau,
         py::object ssa, py::object pmom, py::object alb, py::object emis, py::object beam, py::object itop,
         py::object ibot) {
        Overrides ov(dtau, ssa, pmom, alb, emis, beam, itop, ibot);
        py::gil_scoped_release release;
        return ov.as_arrays ? solveFluxSpectralBands(config, bands, ov.arrays)
                            : solveFluxSpectralBands(config, bands, ov.lists);
      },
      py::arg("config"), py::arg("wavenumber_bands"), DISORTPP_OVERRIDE_ARGS,
      "Flux-only solve for (low, high) wavenumber bands as one GPU batch; returns FluxResult list. NumPy "
      "override arrays are packed into the GPU batch directly.");

  m.def(
      "solve_flux_batch",
      [](const std::vector<DisortFluxConfig>& configs) {
        py::gil_scoped_release release;
        return solveFluxBatch(configs);
      },
      py::arg("configs"),
      "Solve independent flux configurations (any mix of num_layers and num_streams) as GPU batches.");

  m.def("set_device", &setDevice, py::arg("device"), "Select the GPU used by subsequent solves.");
}
