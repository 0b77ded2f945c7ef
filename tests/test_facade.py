"""Host-side behaviour of the DisORT++-compatible facade (`disortpp`):
construction, allocation, phase setters and validation with the reference's
messages and exception types.  Nothing here reaches the GPU: every case
fails validation (or is pure host logic) before a solve is attempted."""
import math

import pytest

import disort_amd.disortpp as d


def c1_flux_config():
    c = d.DisortFluxConfig(4, 16)
    c.allocate()
    c.delta_tau = [0.1, 0.5, 1.0, 2.0]
    c.single_scat_albedo = [0.9] * 4
    c.set_henyey_greenstein(0.8)
    c.direct_beam_flux = math.pi
    c.direct_beam_mu = 0.5
    c.surface_albedo = 0.1
    return c


def test_flux_config_allocate_and_phase_setters():
    c = c1_flux_config()
    assert c.nmom_nstr() == 16
    assert len(c.phase_function_moments) == 4 and len(c.phase_function_moments[0]) == 17
    assert c.phase_function_moments[2][3] == pytest.approx(0.8**3, rel=1e-15)
    c.set_rayleigh(lc=1)
    assert c.phase_function_moments[1][:4] == (1.0, 0.0, 0.1, 0.0)
    c.set_isotropic()
    assert all(row[0] == 1.0 and max(row[1:]) == 0.0 for row in c.phase_function_moments)
    c.set_phase_function(d.PhaseFunction.HenyeyGreenstein, 0.5, 0)
    assert c.phase_function_moments[0][2] == 0.25
    c.validate()
    with pytest.raises(TypeError):
        c.delta_tau[0] = 1.0  # read-back copies are immutable, as in the reference module


@pytest.mark.parametrize(
    "mutate,message",
    [
        (lambda c: setattr(c, "delta_tau", [0.1, -0.5, 1.0, 2.0]), "delta_tau[1] must be >= 0"),
        (lambda c: setattr(c, "single_scat_albedo", [0.9, 0.9, 1.5, 0.9]), "single_scat_albedo[2] must be in [0, 1]"),
        (lambda c: setattr(c, "delta_tau", [0.1]), "delta_tau size (1) < num_layers (4)"),
        (lambda c: setattr(c, "surface_albedo", 1.2), "surface_albedo must be in [0, 1]"),
        (lambda c: setattr(c, "direct_beam_mu", 0.0), "direct_beam_mu must be in (0, 1] when direct_beam_flux > 0"),
        (lambda c: setattr(c, "direct_beam_flux", -1.0), "direct_beam_flux must be >= 0"),
        (lambda c: setattr(c, "use_delta_m_plus", True),
         "use_delta_m_plus requires num_phase_func_moments >= num_streams + 1"),
        (lambda c: setattr(c, "use_diffusion_lower_bc", True),
         "use_diffusion_lower_bc requires use_thermal_emission = true"),
        (lambda c: setattr(c, "use_thermal_emission", True), "temperature size < num_layers + 1"),
    ],
)
def test_flux_config_validation_messages(mutate, message):
    c = c1_flux_config()
    mutate(c)
    with pytest.raises(ValueError, match=message.replace("(", r"\(").replace(")", r"\)").replace("[", r"\[")
                       .replace("]", r"\]").replace("+", r"\+")):
        c.validate()
    # The solver validates before touching the GPU.
    with pytest.raises(ValueError):
        d.DisortFluxSolver16().solve(c)


def test_constructor_rules():
    with pytest.raises(ValueError, match="num_layers must be > 0"):
        d.DisortFluxConfig(0, 16)
    with pytest.raises(ValueError, match="num_streams must be even and >= 4"):
        d.DisortFluxConfig(4, 7)
    with pytest.raises(ValueError, match="num_layers must be positive"):
        d.DisortConfig(0, 16, 0)


def test_stream_count_checks():
    c = c1_flux_config()
    with pytest.raises(ValueError, match=r"DisortFluxSolver<8>::solve: config.num_streams=16 does not match"):
        d.DisortFluxSolver8().solve(c)
    with pytest.raises(ValueError, match="Unsupported nstr=18"):
        d.create_flux_solver(18)
    assert type(d.create_flux_solver(16)).__name__ == "DisortFluxSolver16"
    c18 = d.DisortFluxConfig(4, 18)
    c18.allocate()
    with pytest.raises(ValueError, match="Unsupported num_streams=18"):
        d.solve_flux_spectral_bands(c18, [(100.0, 101.0)])


def test_index_from_bottom_validates_in_top_down_order():
    c = c1_flux_config()
    c.index_from_bottom = True
    c.delta_tau = [0.1, 0.5, 1.0, -2.0]  # bottom-up index 3 is top-down layer 0
    with pytest.raises(ValueError, match=r"delta_tau\[0\] must be >= 0"):
        d.DisortFluxSolver16().solve(c)


def test_general_solver_rejects_what_the_gpu_path_does_not_cover():
    """Checked before any GPU work: a BRDF selected without its parameters
    (the reference's bidirReflectivity messages)."""
    cfg = d.DisortConfig(4, 16, 0)
    cfg.allocate()
    cfg.delta_tau = [0.1, 0.5, 1.0, 2.0]
    cfg.single_scat_albedo = [0.9] * 4
    cfg.set_henyey_greenstein(0.8)
    cfg.num_phi = 1
    cfg.phi_user = [0.0]
    cfg.flags.use_lambertian_surface = False
    cfg.flags.brdf_type = d.BrdfType.CoxMunk
    with pytest.raises(RuntimeError, match="Cox-Munk BRDF selected but cox_munk spec is null"):
        d.DisortSolver().solve(cfg)
    cfg.flags.brdf_type = d.BrdfType.Ambrals
    with pytest.raises(RuntimeError, match="Ross-Li/Ambrals BRDF selected but ambrals spec is null"):
        d.DisortSolver().solve(cfg)
    cfg.flags.brdf_type = d.BrdfType.RPV
    with pytest.raises(RuntimeError, match="rpv spec is null"):
        d.DisortSolver().solve(cfg)


def test_general_config_validation_messages():
    cfg = d.DisortConfig(4, 16, 0)
    cfg.flags.comp_only_fluxes = True
    cfg.allocate()
    cfg.single_scat_albedo = [0.5, 0.5, 2.0, 0.5]
    with pytest.raises(ValueError, match=r"single_scat_albedo\[2\] must be in \[0, 1\]"):
        cfg.validate()
    cfg.single_scat_albedo = [0.5] * 4
    cfg.flags.use_thermal_emission = True
    with pytest.raises(ValueError, match="temperature array is too small"):
        cfg.validate()
    cfg.temperature = [250.0] * 5
    with pytest.raises(ValueError, match="Invalid wavenumber range"):
        cfg.wavenumber_low, cfg.wavenumber_high = 10.0, 5.0
        cfg.validate()
    cfg.wavenumber_low, cfg.wavenumber_high = 5.0, 5.0
    with pytest.raises(ValueError, match="requires 0 <= wavenumber_low < wavenumber_high"):
        cfg.validate()
    cfg.flags.use_delta_m_plus = True
    cfg.wavenumber_high = 6.0
    with pytest.raises(ValueError, match=r"use_delta_m_plus=true requires"):
        cfg.validate()


def test_flux_config_garcia_siewert_fill():
    c = c1_flux_config()
    c.set_haze_garcia_siewert()
    assert c.phase_function_moments[3][:3] == (1.0, 2.41260 / 3, 3.23047 / 5)
    c.set_cloud_garcia_siewert(lc=0)
    assert c.phase_function_moments[0][16] == 13.058 / 33



def test_spectral_override_rows_checked_before_the_gpu():
    """Fewer override rows than bands: IndexError in the nested-list form
    (.at(i)) and in the NumPy-array form alike; band 0's configuration is
    validated with the reference's messages before any GPU work."""
    import numpy as np

    c = c1_flux_config()
    bands = [(100.0, 101.0)] * 3
    with pytest.raises(IndexError):
        d.solve_flux_spectral_bands(c, bands, delta_tau=np.ones((2, 4)))
    with pytest.raises(IndexError):
        d.solve_flux_spectral_bands(c, bands, surface_albedo=[0.1, 0.2])
    with pytest.raises(ValueError, match=r"delta_tau\[2\] must be >= 0"):
        d.solve_flux_spectral_bands(c, bands, delta_tau=np.array([[0.1, 0.2, -1.0, 0.3]] * 3))
    with pytest.raises(ValueError, match=r"delta_tau size \(3\) < num_layers \(4\)"):
        d.solve_flux_spectral_bands(c, bands, delta_tau=np.ones((3, 3)))
