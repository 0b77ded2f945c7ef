// Oracle restatement of the general DISORT solver, DisortSolver::solve
// (src/DisortSolver.cpp:21-399): azimuthal Fourier modes, intensities at user
// angles and optical depths, Lambertian or RPV surface, plane-parallel or
// pseudo-spherical beam, thermal emission, Nakajima-Tanaka intensity
// correction and the BRDF beam-reflection correction.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// Not restated (the caller gets std::invalid_argument): SPECIAL_BC
// (computeAlbTrans), the Hapke / Cox-Munk / Ross-Li BRDFs, temperature
// Jacobians; the Buras-Emde correction needs tabulated phase functions,
// which this interface does not carry.
// The eigenproblem uses the same symmetric route as flux_oracle.cpp (see the
// deviations listed there).
#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>

#include "oracle.hpp"

namespace ORACLE_NS {

// src/DisortSolver.cpp:2734-2780 with badmu = 0 (the reference's dispatcher,
// :2912-2919, always passes 0).
Real rpv_reflectivity(const Real* p, Real mu1, Real mu2, Real phi) {
  return perturb1(PS_BRDF, rpv_reflectivity_exact(p, mu1, mu2, phi));
}

// Hapke (src/DisortSolver.cpp:2710-2731): p = {b0, hh, w}.
Real hapke_reflectivity(const Real* p, Real mu, Real mup, Real dphi) {
  const Real b0 = p[0], hh = p[1], w = p[2];
  const Real ct = mu * mup + std::sqrt((1.0 - mu * mu) * (1.0 - mup * mup)) * std::cos(dphi);
  const Real half = 0.5 * std::acos(ct);
  const Real b = b0 * hh / (hh + std::tan(half));
  const Real gam = std::sqrt(1.0 - w);
  const Real h0 = (1.0 + 2.0 * mup) / (1.0 + 2.0 * gam * mup);
  const Real h = (1.0 + 2.0 * mu) / (1.0 + 2.0 * gam * mu);
  return 0.25 * w * ((1.0 + b) * (1.0 + 0.5 * ct) + h0 * h - 1.0) / (mu + mup);
}

// Ross-thick volume + Li-sparse-reciprocal geometric kernels
// (src/DisortSolver.cpp:2785-2840): p = {iso, vol, geo}; crown
// height/base 2, branch/radius 1, hot-spot width 1.5 degrees.
Real rossli_reflectivity(const Real* p, Real mu, Real mup, Real dphi) {
  const Real hb = 2.0, br = 1.0, a0 = 1.5 * M_PI / 180.0;
  const Real si = std::sqrt(1.0 - mup * mup), sr = std::sqrt(1.0 - mu * mu);
  const Real ti = mup > 0.0 ? si / mup : 0.0, tr = mu > 0.0 ? sr / mu : 0.0;
  Real ca = mup * mu + si * sr * std::cos(dphi);
  ca = ca < -1.0 ? -1.0 : (ca > 1.0 ? 1.0 : ca);
  const Real al = std::acos(ca);
  const Real kvol = 4.0 / (3.0 * M_PI) * (1.0 / (mup + mu)) * ((M_PI / 2.0 - al) * ca + std::sin(al)) *
                        (1.0 + 1.0 / (1.0 + al / a0)) -
                    1.0 / 3.0;
  const Real ti1 = br * ti, tr1 = br * tr;
  const Real ci1 = 1.0 / std::sqrt(1.0 + ti1 * ti1), cr1 = 1.0 / std::sqrt(1.0 + tr1 * tr1);
  const Real si1 = ti1 * ci1, sr1 = tr1 * cr1;
  const Real ca1 = ci1 * cr1 + si1 * sr1 * std::cos(dphi);
  Real g2 = ti1 * ti1 + tr1 * tr1 - 2.0 * ti1 * tr1 * std::cos(dphi);
  if (g2 < 0.0) g2 = 0.0;
  const Real x = ti1 * tr1 * std::sin(dphi);
  const Real cost = hb * (ci1 * cr1) / (ci1 + cr1) * std::sqrt(g2 + x * x);
  const Real t = (cost >= -1.0 && cost <= 1.0) ? std::acos(cost) : 0.0;
  const Real kgeo = (ci1 + cr1) / (M_PI * ci1 * cr1) * (t - std::sin(t) * std::cos(t) - M_PI) +
                    (1.0 + ca1) / (2.0 * ci1 * cr1);
  const Real r = p[0] + p[1] * kvol + p[2] * kgeo;
  return r > 0.0 ? r : 0.0;
}

// Tsang shadowing term of the Cox-Munk surface (src/DisortSolver.cpp:2842-2856).
static Real shadow_eta(Real ct, Real s2) {
  const Real st = std::sqrt(1.0 - ct * ct);
  if (st < 1e-15) return 0.0;
  const Real cot = ct / st;
  return 0.5 * (std::sqrt(s2 / M_PI) / cot * std::exp(-cot * cot / s2) - std::erfc(cot / std::sqrt(s2)));
}

// Cox-Munk ocean surface (src/DisortSolver.cpp:2858-2900): p = {u10,
// refractive index, shadowing on (1) / off (0)}.
Real coxmunk_reflectivity(const Real* p, Real mu, Real mup, Real dphi) {
  const Real si = std::sqrt(1.0 - mup * mup), sr = std::sqrt(1.0 - mu * mu);
  const Real ct = -mup * mu + si * sr * std::cos(dphi);
  const Real omc = 1.0 - ct;
  if (omc < 1e-15) return 0.0;
  const Real mn2 = (mup + mu) * (mup + mu) / (2.0 * omc);
  if (mn2 < 1e-15) return 0.0;
  const Real s2 = 0.003 + 0.00512 * p[0];
  const Real pslope = 1.0 / (M_PI * s2) * std::exp(-(1.0 - mn2) / (s2 * mn2));
  const Real nt = p[1];
  const Real cli = std::sqrt(0.5 * (1.0 + ct));
  const Real q = 1.0 - cli * cli;
  const Real sli = std::sqrt(q > 0.0 ? q : 0.0);
  const Real slt = sli / nt;
  if (slt >= 1.0) return 0.0;
  const Real clt = std::sqrt(1.0 - slt * slt);
  const Real rs = (cli - nt * clt) / (cli + nt * clt);
  const Real rp = (nt * cli - clt) / (clt + nt * cli);
  Real r = (pslope * 0.5 * (rs * rs + rp * rp)) / (4.0 * mup * mu * mn2 * mn2);
  if (p[2] != 0.0) r *= 1.0 / (shadow_eta(mup, s2) + shadow_eta(mu, s2) + 1.0);
  return r > 0.0 ? r : 0.0;
}

Real brdf_reflectivity(int kind, const Real* p, Real mu, Real mup, Real dphi) {
  switch (kind) {
    case BRDF_HAPKE: return perturb1(PS_BRDF, hapke_reflectivity(p, mu, mup, dphi));
    case BRDF_COXMUNK: return perturb1(PS_BRDF, coxmunk_reflectivity(p, mu, mup, dphi));
    case BRDF_ROSSLI: return perturb1(PS_BRDF, rossli_reflectivity(p, mu, mup, dphi));
    default: return rpv_reflectivity(p, mu, mup, dphi);
  }
}

Real rpv_reflectivity_exact(const Real* p, Real mu1, Real mu2, Real phi) {
  const Real rho0 = p[0], k = p[1], theta = p[2], sigma = p[3], t1 = p[4], t2 = p[5], scale = p[6];
  phi = M_PI - phi;  // RPV measures azimuth from the backward direction
  const Real hspot = rho0 * (std::pow(2.0 * mu1 * mu1 * mu1, k - 1.0) * (1.0 - theta) /
                                 ((1.0 + theta) * (1.0 + theta)) * (2.0 - rho0) +
                             sigma / mu1) *
                     (t1 * std::exp(M_PI * t2) + 1.0);
  if (phi == 1.0e-4 && mu1 == mu2) return hspot * scale;
  const Real m = std::pow(mu1 * mu2 * (mu1 + mu2), k - 1.0);
  const Real cosphi = std::cos(phi);
  const Real sin1 = std::sqrt(1.0 - mu1 * mu1), sin2 = std::sqrt(1.0 - mu2 * mu2);
  const Real cosg = mu1 * mu2 + sin1 * sin2 * cosphi;
  const Real g = std::acos(cosg);
  const Real f = (1.0 - theta * theta) / std::pow(1.0 + 2.0 * theta * cosg + theta * theta, 1.5);
  const Real tan1 = sin1 / mu1, tan2 = sin2 / mu2;
  const Real capg = std::sqrt(tan1 * tan1 + tan2 * tan2 - 2.0 * tan1 * tan2 * cosphi);
  const Real h = 1.0 + (1.0 - rho0) / (1.0 + capg);
  const Real t = 1.0 + t1 * std::exp(t2 * (M_PI - g));
  Real ans = rho0 * (m * f * h + sigma / mu1) * t * scale;
  if (ans < 0.0) ans = 0.0;
  return ans;
}

void validate_general_config(const GeneralConfig& c) {
  // src/DisortConfig.cpp:72-249 (the checks that apply to what is restated)
  using std::to_string;
  auto bad = [](const std::string& m) { throw std::invalid_argument(m); };
  if (c.nlyr <= 0) bad("num_layers must be positive");
  if (c.nstr < 4) bad("num_streams must be >= 4");
  if (c.nstr % 2 != 0) bad("num_streams must be even");
  if (c.nmom < 0) bad("num_phase_func_moments must be non-negative");
  if ((int)c.dtau.size() < c.nlyr) bad("delta_tau array is too small");
  if ((int)c.ssa.size() < c.nlyr) bad("single_scat_albedo array is too small");
  if (c.pmom.size() < (size_t)c.nlyr * c.pstride())
    bad("phase_function_moments must have at least num_layers layers");
  for (int lc = 0; lc < c.nlyr; ++lc) {
    if (c.dtau[lc] < 0.0) bad("delta_tau[" + to_string(lc) + "] is negative");
    if (c.ssa[lc] < 0.0 || c.ssa[lc] > 1.0) bad("single_scat_albedo[" + to_string(lc) + "] must be in [0, 1]");
  }
  if (c.thermal) {
    if ((int)c.temper.size() < c.nlyr + 1)
      bad("temperature array is too small (need num_layers+1 when use_thermal_emission=true)");
    for (int l = 0; l <= c.nlyr; ++l)
      if (c.temper[l] < 0.0) bad("temperature[" + to_string(l) + "] is negative");
    if (c.wn_lo < 0.0 || c.wn_hi < c.wn_lo) bad("Invalid wavenumber range for Planck function");
  }
  if (c.diffusion_bc) {
    if (!c.thermal) bad("use_diffusion_lower_bc requires use_thermal_emission = true");
    if (!c.lambertian) bad("use_diffusion_lower_bc requires use_lambertian_surface = true (BRDF is incompatible)");
  }
  if (c.beam_flux < 0.0) bad("direct_beam_flux must be non-negative");
  if ((c.beam_mu <= 0.0 || c.beam_mu > 1.0) && c.beam_flux > 0.0)
    bad("direct_beam_mu must be in (0, 1] when direct_beam_flux > 0");
  if (c.lambertian && (c.albedo < 0.0 || c.albedo > 1.0))
    bad("surface_albedo must be in [0, 1] when use_lambertian_surface=true");
  if (c.use_user_mu) {
    if (c.mu_user.empty()) bad("num_user_mu must be positive when use_user_mu=true");
    for (size_t iu = 0; iu < c.mu_user.size(); ++iu) {
      if (std::abs(c.mu_user[iu]) > 1.0) bad("mu_user[" + to_string(iu) + "] out of range [-1, 1]");
      if (iu > 0 && c.mu_user[iu] <= c.mu_user[iu - 1]) bad("mu_user must be in increasing order");
    }
  }
  if (c.spherical) {
    if (c.radius <= 0.0) bad("bottom_radius must be positive when use_spherical_beam=true");
    if ((int)c.alt.size() < c.nlyr + 1)
      bad("level_altitudes array is too small (need num_layers+1 when use_spherical_beam=true)");
  }
  if (!c.only_fluxes && !c.special_bc) {
    if (c.phi_user.empty()) bad("num_phi must be positive when comp_only_fluxes=false");
    for (size_t j = 0; j < c.phi_user.size(); ++j)
      if (c.phi_user[j] < 0.0 || c.phi_user[j] > 360.0) bad("phi_user[" + to_string(j) + "] must be in [0, 360]");
  }
  if (c.deltam_plus && c.nmom < c.nstr + 1)
    bad("use_delta_m_plus=true requires num_phase_func_moments >= num_streams + 1 (need moments M and M+1)");
  if (c.accur < 0.0) bad("accuracy_fourier_series must be non-negative (0 = disabled)");
  if (c.thermal) {
    if (c.wn_lo < 0.0 || c.wn_hi <= c.wn_lo)
      bad("use_thermal_emission=true requires 0 <= wavenumber_low < wavenumber_high");
    if (c.temp_top < 0.0 || c.temp_bot < 0.0)
      bad("use_thermal_emission=true requires non-negative temperature_top and temperature_bottom");
    if (c.emis_top < 0.0 || c.emis_top > 1.0) bad("use_thermal_emission=true requires 0 <= emissivity_top <= 1");
  }
  if (!c.lambertian && (c.brdf < BRDF_RPV || c.brdf > BRDF_ROSSLI)) bad("bidirReflectivity: unknown BRDF type");
  if (c.use_user_tau && c.tau_user.empty()) bad("oracle: use_user_tau needs tau_user values");
}

namespace {

constexpr int NMUG = 200;  // BRDF azimuth quadrature (src/DisortSolver.cpp:484-496)

struct GeneralSolver {
  const GeneralConfig& c;
  int N, NN, L, ncut = 0, ntau = 0, numu = 0, nphi = 0;
  bool deltam = false, deltam_plus = false, lyrcut = false, scat_yes = false, do_user = false;

  std::vector<Real> cmu, cwt, sd, sdinv;
  std::vector<Real> ylmc, ylm0, ylmu;               // P_l^m at quadrature, beam, user angles
  std::vector<Real> tauc, dtaus, taus, ssas, fwd, gl;
  std::vector<Real> chap, chtau, btrans;
  std::vector<Real> kk, gc, cc, gu;                 // eigen-system, phase matrix, user-angle eigenvectors
  std::vector<Real> zb, zbs, xba, zraw, z0, z1, xr0, xr1;
  std::vector<Real> zbu, z0u, z1u;                  // sources at user angles [L][numu]
  std::vector<Real> rq, eq, ru, eu;                 // surface terms (quadrature / user)
  std::vector<Real> gmu, gwt;                       // BRDF azimuth quadrature
  std::vector<Real> rho_beam_user;                  // [numu][N] beam BRDF Fourier terms
  Real pl_top = 0, pl_bot = 0, pl_bot_deriv = 0;
  std::vector<Real> LL;
  std::vector<Real> utau, sutau, umu;               // user levels (unscaled, scaled), user angles
  std::vector<int> lyr_of;                          // layer holding each user level
  // Temperature-Jacobian state (see ThermalJacobianInputs).
  std::vector<Real> dz0, dz1, dslope, dicpt, band_lu;
  std::vector<int> band_piv;

  explicit GeneralSolver(const GeneralConfig& cfg) : c(cfg), N(cfg.nstr), NN(cfg.nstr / 2), L(cfg.nlyr) {}

  Real& G(int lc, int r, int col) { return gc[(size_t)lc * N * N + r + (size_t)col * N]; }
  Real& CC(int lc, int r, int col) { return cc[(size_t)lc * N * N + r + (size_t)col * N]; }
  Real& GL(int k, int lc) { return gl[(size_t)lc * N + k]; }
  Real& GU(int lc, int iu, int col) { return gu[((size_t)lc * numu + iu) * N + col]; }
  Real Y(int l, int i) const { return ylmc[(size_t)l * N + i]; }
  Real YU(int l, int iu) const { return ylmu[(size_t)l * numu + iu]; }
  Real& RQ(int iq, int col) { return rq[(size_t)iq * (NN + 1) + col]; }
  Real& RU(int iu, int col) { return ru[(size_t)iu * (NN + 1) + col]; }
  Real K(int lc, int q) const { return kk[(size_t)lc * N + q]; }
  Real rho(Real mu, Real mup, Real dphi) const { return brdf_reflectivity(c.brdf, c.rpv, mu, mup, dphi); }

  // src/DisortSolver.cpp:641-820 (disortSet); identical to the flux solver's
  // set-up plus the user-level mapping.
  void set_up() {
    const Real abscut = 10.0;
    tauc.assign(L + 1, 0.0);
    for (int lc = 0; lc < L; ++lc) tauc[lc + 1] = tauc[lc] + c.dtau[lc];
    dtaus.assign(L, 0.0);
    taus.assign(L + 1, 0.0);
    ssas.assign(L, 0.0);
    fwd.assign(L, 0.0);
    gl.assign((size_t)L * N, 0.0);
    Real abstau = 0.0;
    ncut = L;
    for (int lc = 0; lc < L; ++lc) {
      if (abstau < abscut) ncut = lc + 1;
      abstau += (1.0 - c.ssa[lc]) * c.dtau[lc];
      const Real w = c.ssa[lc];
      const Real pm0 = 1.0;  // P_0 is forced to 1 (src/DisortSolver.cpp:668)
      auto mom = [&](int k) { return k == 0 ? pm0 : c.pm(k, lc); };
      Real f;
      auto plain_delta_m = [&](Real fr) {
        ssas[lc] = w * (1.0 - fr) / (1.0 - fr * w);
        dtaus[lc] = (1.0 - fr * w) * c.dtau[lc];
        taus[lc + 1] = taus[lc] + dtaus[lc];
        for (int k = 0; k < N; ++k) GL(k, lc) = (2.0 * k + 1.0) * ssas[lc] * (mom(k) - fr) / (1.0 - fr);
      };
      if (!deltam && !deltam_plus) {
        ssas[lc] = w;
        dtaus[lc] = c.dtau[lc];
        taus[lc + 1] = tauc[lc + 1];
        for (int k = 0; k < N; ++k) GL(k, lc) = (2.0 * k + 1.0) * ssas[lc] * mom(k);
        f = 0.0;
      } else if (!deltam_plus || mom(N - 1) == mom(N)) {
        f = mom(N);
        plain_delta_m(f);
      } else {
        const Real pm = mom(N), pm1 = mom(N + 1);
        if (pm == pm1) {
          f = pm;
          plain_delta_m(f);
        } else {
          const Real M = static_cast<Real>(N);
          const Real sig2 = ((M + 1.0) * (M + 1.0) - M * M) / (std::log(pm * pm) - std::log(pm1 * pm1));
          const Real cf = std::exp(M * M / (2.0 * sig2));
          f = cf * pm;
          ssas[lc] = w * (1.0 - f) / (1.0 - f * w);
          dtaus[lc] = (1.0 - f * w) * c.dtau[lc];
          taus[lc + 1] = taus[lc] + dtaus[lc];
          for (int k = 0; k < N; ++k) {
            const Real kd = static_cast<Real>(k);
            GL(k, lc) = (2.0 * k + 1.0) * ssas[lc] * (mom(k) - f * std::exp(-kd * kd / (2.0 * sig2))) / (1.0 - f);
          }
        }
      }
      fwd[lc] = f;
    }
    chap.assign(L, 0.0);
    chtau.assign(L + 1, 0.0);
    btrans.assign(L + 1, 0.0);
    if (c.beam_flux > 0.0) {
      btrans[0] = 1.0;
      if (c.spherical) {
        for (int lc = 0; lc < ncut; ++lc) {
          const Real taup_mid = taus[lc] + dtaus[lc] / 2.0;
          const Real ch_mid = chapman(lc + 1, 0.5, c.alt.data(), dtaus.data(), L, c.radius, c.beam_mu);
          chap[lc] = taup_mid / ch_mid;
          chtau[lc + 1] = chapman(lc + 1, 0.0, c.alt.data(), dtaus.data(), L, c.radius, c.beam_mu);
          btrans[lc + 1] = std::exp(-chtau[lc + 1]);
        }
      } else {
        for (int lc = 0; lc < ncut; ++lc) {
          chap[lc] = c.beam_mu;
          btrans[lc + 1] = std::exp(-taus[lc + 1] / c.beam_mu);
        }
      }
    }
    lyrcut = (abstau >= abscut && !c.thermal && L > 1 && !c.no_cutoff);
    if (!lyrcut) ncut = L;
    // User levels (src/DisortSolver.cpp:803-820).
    lyr_of.assign(ntau, 0);
    sutau.assign(ntau, 0.0);
    for (int lu = 0; lu < ntau; ++lu) {
      int lc;
      for (lc = 0; lc < L - 1; ++lc)
        if (utau[lu] >= tauc[lc] && utau[lu] <= tauc[lc + 1]) break;
      sutau[lu] = utau[lu];
      if (deltam || deltam_plus) sutau[lu] = taus[lc] + (1.0 - c.ssa[lc] * fwd[lc]) * (utau[lu] - tauc[lc]);
      lyr_of[lu] = lc;
    }
    cmu.assign(N, 0.0);
    cwt.assign(N, 0.0);
    gauss_quadrature(NN, cmu.data(), cwt.data());
    sd.assign(NN, 0.0);
    sdinv.assign(NN, 0.0);
    for (int i = 0; i < NN; ++i) {
      sd[i] = std::sqrt(cwt[i] * cmu[i]);
      sdinv[i] = 1.0 / sd[i];
    }
    if (!c.use_user_mu || c.only_fluxes) {
      umu.assign(N, 0.0);
      for (int iu = 0; iu < NN; ++iu) umu[iu] = -cmu[NN - 1 - iu];
      for (int iu = NN; iu < N; ++iu) umu[iu] = cmu[iu - NN];
    }
  }

  // src/DisortSolver.cpp:872-1150 at azimuthal mode m (sums over l >= m).
  void solve_eigen(int lc, int m) {
    std::vector<Real> top((size_t)NN * N);
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < NN; ++i) {
        Real s = 0.0;
        for (int l = m; l < N; ++l) s += (GL(l, lc) * Y(l, i)) * Y(l, j);
        top[i + (size_t)j * NN] = s * 0.5 * cwt[j];
      }
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < NN; ++i) CC(lc, i, j) = top[i + (size_t)j * NN];
    for (int j = 0; j < NN; ++j)
      for (int i = 0; i < NN; ++i) {
        CC(lc, NN + i, j) = top[i + (size_t)(NN + j) * NN];
        CC(lc, NN + i, NN + j) = top[i + (size_t)j * NN];
      }
    std::vector<Real> amb(NN * NN), apb(NN * NN);
    for (int i = 0; i < NN; ++i) {
      const Real inv = 1.0 / cmu[i];
      for (int j = 0; j < NN; ++j) {
        amb[i + j * NN] = inv * (top[i + j * NN] - top[i + (NN + j) * NN]);
        apb[i + j * NN] = inv * (top[i + j * NN] + top[i + (NN + j) * NN]);
      }
      amb[i + i * NN] -= inv;
      apb[i + i * NN] -= inv;
    }
    std::vector<Real> nbs(NN * NN), nas(NN * NN);
    for (int i = 0; i < NN; ++i)
      for (int j = 0; j < NN; ++j) {
        nbs[i + j * NN] = -sd[i] * amb[i + j * NN] * sdinv[j];
        nas[i + j * NN] = -(sd[i] * apb[i + j * NN] * sdinv[j]);
      }
    std::vector<Real> lw(NN * NN), S(NN * NN), tmp(NN * NN), ev(NN), W(NN * NN), vec(NN * NN);
    auto congruence = [&](const std::vector<Real>& a) {  // S = lw^T a lw
      for (int j = 0; j < NN; ++j)
        for (int i = 0; i < NN; ++i) {
          Real s = 0.0;
          for (int k = 0; k < NN; ++k) s += a[i + k * NN] * lw[k + j * NN];
          tmp[i + j * NN] = s;
        }
      for (int j = 0; j < NN; ++j)
        for (int i = 0; i < NN; ++i) {
          Real s = 0.0;
          for (int k = 0; k < NN; ++k) s += lw[k + i * NN] * tmp[k + j * NN];
          S[i + j * NN] = s;
        }
    };
    if (cholesky_lower(NN, nbs.data(), lw.data())) {
      congruence(nas);
      sym_eigen_jacobi(NN, S.data(), ev.data(), W.data());
      for (int col = 0; col < NN; ++col) {
        std::vector<Real> v(NN);
        for (int i = NN - 1; i >= 0; --i) {
          Real s = W[i + col * NN];
          for (int k = i + 1; k < NN; ++k) s -= lw[k + i * NN] * v[k];
          v[i] = s / lw[i + i * NN];
        }
        for (int j = 0; j < NN; ++j) vec[j + col * NN] = sdinv[j] * v[j];
      }
    } else if (cholesky_lower(NN, nas.data(), lw.data())) {
      congruence(nbs);
      sym_eigen_jacobi(NN, S.data(), ev.data(), W.data());
      for (int col = 0; col < NN; ++col)
        for (int j = 0; j < NN; ++j) {
          Real s = 0.0;
          for (int k = 0; k <= j; ++k) s += lw[j + k * NN] * W[k + col * NN];
          vec[j + col * NN] = sdinv[j] * s;
        }
    } else {
      throw std::runtime_error("oracle: neither -B_s nor -A_s is positive definite (degenerate layer)");
    }
    std::vector<Real> rk(NN);
    for (int iq = 0; iq < NN; ++iq) {
      rk[iq] = std::sqrt(std::abs(ev[iq]));
      kk[(size_t)lc * N + iq + NN] = rk[iq];
      kk[(size_t)lc * N + NN - 1 - iq] = -rk[iq];
    }
    for (int jq = 0; jq < NN; ++jq)
      for (int iq = 0; iq < NN; ++iq) {
        Real s = 0.0;
        for (int k = 0; k < NN; ++k) s += amb[iq + k * NN] * vec[k + jq * NN];
        Real plus = s / rk[jq];
        const Real minus = vec[iq + jq * NN];
        G(lc, NN + iq, NN + jq) = 0.5 * (plus + minus);
        G(lc, NN - iq - 1, NN + jq) = 0.5 * (plus - minus);
        plus = -plus;
        G(lc, NN + iq, NN - jq - 1) = 0.5 * (plus + minus);
        G(lc, NN - iq - 1, NN - jq - 1) = 0.5 * (plus - minus);
      }
  }

  // Row of G for quadrature index jq (cmu order) -- the reference's
  // reduced_eigenvectors_ layout (src/DisortSolver.cpp:1124-1138).
  int grow(int jq) const { return jq < NN ? NN + jq : N - 1 - jq; }

  // src/DisortSolver.cpp:2468-2504
  void interp_eigvec(int lc, int m) {
    std::vector<Real> W((size_t)N * N, 0.0);  // W(l, gcol)
    for (int l = m; l < N; ++l)
      for (int col = 0; col < N; ++col) {
        Real s = 0.0;
        for (int jq = 0; jq < N; ++jq) s += (Y(l, jq) * cwt[jq]) * G(lc, grow(jq), col);
        W[(size_t)l * N + col] = s * (0.5 * GL(l, lc));
      }
    for (int iu = 0; iu < numu; ++iu)
      for (int col = 0; col < N; ++col) {
        Real s = 0.0;
        for (int l = m; l < N; ++l) s += YU(l, iu) * W[(size_t)l * N + col];
        GU(lc, iu, col) = s;
      }
  }

  void beam_legendre(int m) {
    const Real mu = -c.beam_mu;
    legendre_flat(1, m, 0, N - 1, &mu, ylm0.data());
  }

  // src/DisortSolver.cpp:1319-1395
  void beam_source(int m) {
    beam_legendre(m);
    const Real d0 = (m == 0) ? 1.0 : 0.0;
    const Real thresh = 1000.0 * std::numeric_limits<double>::epsilon();
    std::vector<Real> A((size_t)N * N), b(N);
    DenseLU lu;
    for (int lc = 0; lc < L; ++lc) {
      Real bmu = c.beam_mu;
      const Real mu0sq = c.beam_mu * c.beam_mu;
      for (int iq = 0; iq < NN; ++iq) {
        const Real k = K(lc, iq + NN);
        if (std::abs(1.0 - mu0sq * k * k) < thresh) {
          bmu = 0.999 * c.beam_mu;
          break;
        }
      }
      for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) A[i + (size_t)j * N] = -CC(lc, i, j);
      for (int i = 0; i < N; ++i) A[i + (size_t)i * N] += 1.0 + cmu[i] / bmu;
      for (int iq = 0; iq < N; ++iq) {
        Real s = 0.0;
        for (int k = m; k < N; ++k) s += GL(k, lc) * Y(k, iq) * ylm0[k];
        b[iq] = (2.0 - d0) * c.beam_flux * s / (4.0 * M_PI);
      }
      lu.factor(N, A.data());
      lu.solve(b.data());
      for (int iq = 0; iq < N; ++iq) zraw[(size_t)lc * N + iq] = b[iq];
      for (int iq = 0; iq < NN; ++iq) {
        zb[(size_t)lc * N + NN + iq] = b[iq];
        zb[(size_t)lc * N + NN - iq - 1] = b[iq + NN];
      }
    }
  }

  // src/DisortSolver.cpp:1397-1531; zraw keeps the right-hand side ZJ, as the
  // reference does (:1527).
  void beam_source_spherical(int m) {
    beam_legendre(m);
    const Real d0 = (m == 0) ? 1.0 : 0.0;
    const Real big = std::sqrt(std::numeric_limits<double>::max()) / 1.0e10;
    std::vector<Real> A((size_t)N * N), zj(N), xb0(N), xb1(N);
    DenseLU lu;
    for (int lc = 0; lc < L; ++lc) {
      for (int iq = 0; iq < N; ++iq) {
        Real s = 0.0;
        for (int k = m; k < N; ++k) s += GL(k, lc) * Y(k, iq) * ylm0[k];
        zj[iq] = (2.0 - d0) * c.beam_flux * s / (4.0 * M_PI);
      }
      const Real q0a = std::exp(-chtau[lc]);
      const Real q2a = std::exp(-chtau[lc + 1]);
      Real x = 1.0 / chap[lc];
      if (std::abs(x) > big && taus[lc + 1] > 1.0) x = 0.0;
      if (std::abs(x * taus[lc + 1]) > std::log(big)) x = 0.0;
      if (std::abs(x) > 1.0e-5)
        for (int iq = 0; iq < NN; ++iq)
          if (std::abs((std::abs(x) - 1.0 / cmu[iq]) / x) < 0.05) {
            x *= 1.001;
            break;
          }
      xba[lc] = x;
      const Real dt = taus[lc + 1] - taus[lc];
      for (int iq = 0; iq < N; ++iq) {
        const Real q0 = q0a * zj[iq], q2 = q2a * zj[iq];
        xb1[iq] = (q2 * std::exp(x * taus[lc + 1]) - q0 * std::exp(x * taus[lc])) / dt;
        xb0[iq] = q0 * std::exp(x * taus[lc]) - xb1[iq] * taus[lc];
      }
      auto build = [&](Real xv) {
        for (int j = 0; j < N; ++j)
          for (int i = 0; i < N; ++i) A[i + (size_t)j * N] = -CC(lc, i, j);
        for (int i = 0; i < N; ++i) A[i + (size_t)i * N] += 1.0 + xv * cmu[i];
      };
      build(x);
      lu.factor(N, A.data());
      if (1.0 / lu.frob_norm_of_input < 1.0e-4) {
        if (x == 0.0) x = 5.0e-9;
        x *= 1.00000005;
        xba[lc] = x;
        build(x);
        lu.factor(N, A.data());
      }
      std::vector<Real> s1 = xb1;
      lu.solve(s1.data());
      std::vector<Real> s0(N);
      for (int iq = 0; iq < N; ++iq) s0[iq] = xb0[iq] + cmu[iq] * s1[iq];
      lu.solve(s0.data());
      for (int iq = 0; iq < N; ++iq) zraw[(size_t)lc * N + iq] = zj[iq];
      for (int iq = 0; iq < NN; ++iq) {
        zb[(size_t)lc * N + NN + iq] = s0[iq];
        zb[(size_t)lc * N + NN - 1 - iq] = s0[iq + NN];
        zbs[(size_t)lc * N + NN + iq] = s1[iq];
        zbs[(size_t)lc * N + NN - 1 - iq] = s1[iq + NN];
      }
    }
  }

  // src/DisortSolver.cpp:1533-1590 (m = 0 only)
  void thermal_source() {
    std::vector<Real> pk(L + 1);
    for (int l = 0; l <= L; ++l) pk[l] = planck_band(c.wn_lo, c.wn_hi, c.temper[l]);
    std::vector<Real> A((size_t)N * N), b1(N), b0(N);
    DenseLU lu;
    for (int lc = 0; lc < L; ++lc) {
      for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) A[i + (size_t)j * N] = -CC(lc, i, j);
      for (int i = 0; i < N; ++i) A[i + (size_t)i * N] += 1.0;
      lu.factor(N, A.data());
      Real s1 = 0.0;
      if (dtaus[lc] > 1e-4) s1 = (pk[lc + 1] - pk[lc]) / dtaus[lc];
      const Real s0 = pk[lc] - s1 * taus[lc];
      xr0[lc] = s0;
      xr1[lc] = s1;
      for (int i = 0; i < N; ++i) b1[i] = (1.0 - ssas[lc]) * s1;
      lu.solve(b1.data());
      for (int i = 0; i < N; ++i) b0[i] = (1.0 - ssas[lc]) * s0 + cmu[i] * b1[i];
      lu.solve(b0.data());
      for (int iq = 0; iq < NN; ++iq) {
        z0[(size_t)lc * N + NN + iq] = b0[iq];
        z0[(size_t)lc * N + NN - iq - 1] = b0[iq + NN];
        z1[(size_t)lc * N + NN + iq] = b1[iq];
        z1[(size_t)lc * N + NN - iq - 1] = b1[iq + NN];
      }
      if (c.jacobian)
        thermal_source_derivs(lu, N, 1.0 - ssas[lc], dtaus[lc], taus[lc], cmu.data(), &dz0[(size_t)2 * lc * N],
                              &dz1[(size_t)2 * lc * N], &dslope[2 * lc], &dicpt[2 * lc]);
    }
  }

  // src/DisortSolver.cpp:2130-2377: temperature Jacobians of the m = 0
  // fluxes at the user levels, with the surface's m = 0 reflectivity and
  // emissivity at the quadrature angles.
  void jacobian(GeneralOutput& out) {
    std::vector<Real> refl((size_t)NN * NN);
    for (int iq = 0; iq < NN; ++iq)
      for (int jq = 0; jq < NN; ++jq) refl[(size_t)iq * NN + jq] = RQ(iq, jq + 1);
    ThermalJacobianInputs in;
    in.N = N;
    in.L = L;
    in.ncut = ncut;
    in.lyrcut = lyrcut;
    in.diffusion_bc = c.diffusion_bc;
    in.cmu = cmu.data();
    in.cwt = cwt.data();
    in.taus = taus.data();
    in.dtaus = dtaus.data();
    in.kk = kk.data();
    in.gc = gc.data();
    in.band = band_lu.data();
    in.ipvt = band_piv.data();
    in.dz0 = dz0.data();
    in.dz1 = dz1.data();
    in.dslope = dslope.data();
    in.dicpt = dicpt.data();
    in.ssa = c.ssa.data();
    in.refl = refl.data();
    in.emis = eq.data();
    in.nrows = ntau;
    in.row_layer = lyr_of.data();
    in.row_tau = sutau.data();
    in.wn_lo = c.wn_lo;
    in.wn_hi = c.wn_hi;
    in.temp_bot = c.temp_bot;
    in.temper = c.temper.data();
    in.perturb_stage = PS_JACOBIAN;
    thermal_jacobian(in, out.jac);
  }

  // src/DisortSolver.cpp:2506-2580
  void interp_source(int lc, int m) {
    const Real d0 = (m == 0) ? 1.0 : 0.0;
    if (c.beam_flux > 0.0) {
      std::vector<Real> psi0(N, 0.0);
      for (int l = m; l < N; ++l) {
        Real s = 0.0;
        for (int jq = 0; jq < N; ++jq) s += cwt[jq] * Y(l, jq) * zraw[(size_t)lc * N + jq];
        psi0[l] = 0.5 * GL(l, lc) * s;
      }
      const Real fact = (2.0 - d0) * c.beam_flux / (4.0 * M_PI);
      for (int iu = 0; iu < numu; ++iu) {
        Real s = 0.0;
        for (int l = m; l < N; ++l) s += YU(l, iu) * (psi0[l] + fact * GL(l, lc) * ylm0[l]);
        zbu[(size_t)lc * numu + iu] = s;
      }
    }
    if (c.thermal && m == 0) {
      std::vector<Real> p0(N, 0.0), p1(N, 0.0);
      for (int l = m; l < N; ++l) {
        Real s0 = 0.0, s1 = 0.0;
        for (int jq = 0; jq < N; ++jq) {
          const int r = grow(jq);
          s0 += cwt[jq] * Y(l, jq) * z0[(size_t)lc * N + r];
          s1 += cwt[jq] * Y(l, jq) * z1[(size_t)lc * N + r];
        }
        p0[l] = 0.5 * GL(l, lc) * s0;
        p1[l] = 0.5 * GL(l, lc) * s1;
      }
      for (int iu = 0; iu < numu; ++iu) {
        Real s0 = 0.0, s1 = 0.0;
        for (int l = m; l < N; ++l) {
          s0 += YU(l, iu) * p0[l];
          s1 += YU(l, iu) * p1[l];
        }
        z0u[(size_t)lc * numu + iu] = s0 + (1.0 - ssas[lc]) * xr0[lc];
        z1u[(size_t)lc * numu + iu] = s1 + (1.0 - ssas[lc]) * xr1[lc];
      }
    }
  }

  // src/DisortSolver.cpp:2582-2708
  void surface_bidir(int m) {
    const Real d0 = (m == 0) ? 1.0 : 0.0;
    std::fill(rq.begin(), rq.end(), 0.0);
    std::fill(eq.begin(), eq.end(), 0.0);
    const int K2 = NMUG / 2;
    auto cosmp = [&](int k) { return m == 0 ? Real(1.0) : std::cos(static_cast<Real>(m) * M_PI * gmu[k]); };
    auto emissivity = [&](Real mu) {
      Real dref = 0.0;
      for (int jg = 0; jg < NMUG; ++jg) {
        Real s = 0.0;
        for (int k = 0; k < K2; ++k) s += gwt[k] * gmu[k] * rho(mu, gmu[k], M_PI * gmu[jg]);
        dref += gwt[jg] * s;
      }
      return 1.0 - dref;
    };
    if (c.lambertian && m == 0) {
      for (int iq = 0; iq < NN; ++iq) {
        eq[iq] = 1.0 - c.albedo;
        for (int jq = 0; jq <= NN; ++jq) RQ(iq, jq) = c.albedo;
      }
    } else if (!c.lambertian) {
      for (int iq = 0; iq < NN; ++iq) {
        for (int jq = 0; jq < NN; ++jq) {
          Real s = 0.0;
          for (int k = 0; k < K2; ++k) s += gwt[k] * rho(cmu[iq], cmu[jq], M_PI * gmu[k]) * cosmp(k);
          RQ(iq, jq + 1) = (2.0 - d0) * s;
        }
        if (c.beam_flux > 0.0) {
          Real s = 0.0;
          for (int k = 0; k < K2; ++k) s += gwt[k] * rho(cmu[iq], c.beam_mu, M_PI * gmu[k]) * cosmp(k);
          RQ(iq, 0) = (2.0 - d0) * s;
        }
      }
      if (m == 0)
        for (int iq = 0; iq < NN; ++iq) eq[iq] = emissivity(cmu[iq]);
    }
    if (do_user) {
      std::fill(ru.begin(), ru.end(), 0.0);
      std::fill(eu.begin(), eu.end(), 0.0);
      for (int iu = 0; iu < numu; ++iu) {
        const Real u = umu[iu];
        if (u <= 0.0) continue;
        if (c.lambertian && m == 0) {
          for (int iq = 0; iq <= NN; ++iq) RU(iu, iq) = c.albedo;
          eu[iu] = 1.0 - c.albedo;
        } else if (!c.lambertian) {
          for (int jq = 0; jq < NN; ++jq) {
            Real s = 0.0;
            for (int k = 0; k < K2; ++k) s += gwt[k] * rho(u, cmu[jq], M_PI * gmu[k]) * cosmp(k);
            RU(iu, jq + 1) = (2.0 - d0) * s;
          }
          if (c.beam_flux > 0.0) {
            Real s = 0.0;
            for (int k = 0; k < K2; ++k) s += gwt[k] * rho(u, c.beam_mu, M_PI * gmu[k]) * cosmp(k);
            RU(iu, 0) = (2.0 - d0) * s;
          }
          if (m == 0) eu[iu] = emissivity(u);
        }
      }
    }
  }

  // Beam particular solution at scaled depth t of layer lc, row iq (G order).
  Real beam_at(int lc, int iq, Real t, Real plane_trans) {
    if (c.spherical) return std::exp(-xba[lc] * t) * (zb[(size_t)lc * N + iq] + zbs[(size_t)lc * N + iq] * t);
    return zb[(size_t)lc * N + iq] * plane_trans;
  }

  // src/DisortSolver.cpp:1153-1317 (setMatrix) and :1669-1974 (solve0)
  void boundary_solve(int m) {
    const Real d0 = (m == 0) ? 1.0 : 0.0;
    const int ncd = 3 * NN - 1;
    const int lda = 3 * ncd + 1;
    const int nshift = lda - 2 * N + 1;
    const int nn = N * L;
    std::vector<Real> band((size_t)lda * nn, 0.0);
    auto B = [&](int r, int col) -> Real& { return band[r + (size_t)col * lda]; };
    std::vector<Real> work(N, 0.0);
    int ncol = 0;
    for (int lc = 0; lc < ncut; ++lc) {
      for (int iq = 0; iq < NN; ++iq) work[iq] = std::exp(K(lc, iq) * dtaus[lc]);
      int jcol = 0;
      for (int iq = 0; iq < NN; ++iq) {
        int irow = nshift - jcol - 1;
        for (int jq = 0; jq < N; ++jq) {
          if (irow + N >= 0 && irow + N < lda && ncol < nn) B(irow + N, ncol) = G(lc, jq, iq);
          if (irow >= 0 && irow < lda && ncol < nn) B(irow, ncol) = -G(lc, jq, iq) * work[iq];
          ++irow;
        }
        ++ncol;
        ++jcol;
      }
      for (int iq = NN; iq < N; ++iq) {
        int irow = nshift - jcol - 1;
        for (int jq = 0; jq < N; ++jq) {
          B(irow + N, ncol) = G(lc, jq, iq) * work[N - iq - 1];
          B(irow, ncol) = -G(lc, jq, iq);
          ++irow;
        }
        ++ncol;
        ++jcol;
      }
    }
    int jcol = 0;
    for (int iq = 0; iq < NN; ++iq) {
      const Real expa = std::exp(K(0, iq) * taus[1]);
      int irow = nshift - jcol + NN - 1;
      for (int jq = NN - 1; jq >= 0; --jq) B(irow++, jcol) = G(0, jq, iq) * expa;
      ++jcol;
    }
    for (int iq = NN; iq < N; ++iq) {
      int irow = nshift - jcol + NN - 1;
      for (int jq = NN - 1; jq >= 0; --jq) B(irow++, jcol) = G(0, jq, iq);
      ++jcol;
    }
    const int lb = ncut - 1;
    const bool nocouple = lyrcut || c.diffusion_bc || (c.lambertian && m > 0);
    int nncol = ncol - N - 1;
    jcol = 0;
    auto bottom = [&](int jq, int iq) {
      if (nocouple) return G(lb, jq, iq);
      Real rs = 0.0;
      for (int k = 0; k < NN; ++k) rs += cwt[k] * cmu[k] * RQ(jq - NN, k + 1) * G(lb, NN - 1 - k, iq);
      return G(lb, jq, iq) - (1.0 + d0) * rs;
    };
    for (int iq = 0; iq < NN; ++iq) {
      ++nncol;
      int irow = nshift - jcol + N - 1;
      for (int jq = NN; jq < N; ++jq) B(irow++, nncol) = bottom(jq, iq);
      ++jcol;
    }
    for (int iq = NN; iq < N; ++iq) {
      ++nncol;
      int irow = nshift - jcol + N - 1;
      const Real expa = work[N - iq - 1];
      for (int jq = NN; jq < N; ++jq) B(irow++, nncol) = bottom(jq, iq) * expa;
      ++jcol;
    }

    const int nc = N * ncut;
    std::vector<Real> b(nc, 0.0);
    const bool beam = c.beam_flux > 0.0;
    const Real tb = taus[ncut];
    if (m > 0) {
      if (beam) {
        for (int iq = 0; iq < NN; ++iq) b[iq] = -zb[NN - 1 - iq];
        for (int iq = 0; iq < NN; ++iq) {
          Real v;
          if (lyrcut || c.lambertian) {
            v = -beam_at(lb, NN + iq, tb, btrans[ncut]);
          } else {
            // The reference uses the plane-parallel form here also for a
            // spherical beam (src/DisortSolver.cpp:1733-1744).
            Real s = 0.0;
            for (int jq = 0; jq < NN; ++jq)
              s += cwt[jq] * cmu[jq] * RQ(iq, jq + 1) * zb[(size_t)lb * N + NN - 1 - jq] * btrans[ncut];
            v = s + (RQ(iq, 0) * c.beam_mu * c.beam_flux / M_PI - zb[(size_t)lb * N + NN + iq]) * btrans[ncut];
          }
          b[nc - NN + iq] = v;
        }
        for (int lc = 1; lc < ncut; ++lc) {
          const int ip = lc * N - NN;
          for (int iq = 0; iq < N; ++iq)
            b[ip + iq] = c.spherical ? beam_at(lc, iq, taus[lc], 0.0) - beam_at(lc - 1, iq, taus[lc], 0.0)
                                     : (zb[(size_t)lc * N + iq] - zb[(size_t)(lc - 1) * N + iq]) * btrans[lc];
        }
      }
    } else {
      for (int iq = 0; iq < NN; ++iq) {
        Real s = 0.0;
        if (beam) s -= zb[NN - iq - 1];
        if (c.thermal) s -= z0[NN - iq - 1];
        s += c.iso_top + pl_top;
        b[iq] = s;
      }
      if (c.diffusion_bc) {
        for (int iq = 0; iq < NN; ++iq) {
          Real s = 0.0;
          if (beam) s -= beam_at(lb, NN + iq, tb, btrans[ncut]);
          if (c.thermal) {
            s -= z0[(size_t)lb * N + NN + iq];
            s -= z1[(size_t)lb * N + NN + iq] * tb;
          }
          s += pl_bot + cmu[iq] * pl_bot_deriv;
          b[nc - NN + iq] = s;
        }
      } else {
        for (int iq = 0; iq < NN; ++iq) {
          Real s = 0.0;
          if (!lyrcut) {
            Real rs = 0.0;
            for (int jq = 0; jq < NN; ++jq) {
              Real ps = 0.0;
              if (beam) ps += beam_at(lb, NN - 1 - jq, tb, btrans[ncut]);
              if (c.thermal) ps += z0[(size_t)lb * N + NN - 1 - jq] + z1[(size_t)lb * N + NN - 1 - jq] * tb;
              rs += cwt[jq] * cmu[jq] * RQ(iq, jq + 1) * ps;
            }
            s += 2.0 * rs;
          }
          if (beam) {
            s -= beam_at(lb, NN + iq, tb, btrans[ncut]);
            if (!lyrcut) s += RQ(iq, 0) * c.beam_mu * c.beam_flux * btrans[ncut] / M_PI;
          }
          if (c.thermal) {
            s -= z0[(size_t)lb * N + NN + iq];
            s -= z1[(size_t)lb * N + NN + iq] * tb;
          }
          s += c.iso_bot;
          if (c.thermal) s += eq[iq] * pl_bot;
          b[nc - NN + iq] = s;
        }
      }
      for (int lc = 1; lc < ncut; ++lc) {
        const int ip = lc * N - NN;
        for (int iq = 0; iq < N; ++iq) {
          Real s = 0.0;
          if (beam) {
            s += beam_at(lc, iq, taus[lc], btrans[lc]);
            s -= beam_at(lc - 1, iq, taus[lc], btrans[lc]);
          }
          if (c.thermal) {
            s += z0[(size_t)lc * N + iq] - z0[(size_t)(lc - 1) * N + iq];
            s += (z1[(size_t)lc * N + iq] - z1[(size_t)(lc - 1) * N + iq]) * taus[lc];
          }
          b[ip + iq] = s;
        }
      }
    }
    std::vector<int> ipvt(nc);
    band_factor(band.data(), lda, nc, ncd, ncd, ipvt.data());
    band_solve(band.data(), lda, nc, ncd, ncd, ipvt.data(), b.data());
    LL.assign(nn, 0.0);
    for (int i = 0; i < nc; ++i) LL[i] = b[i];
    perturb(PS_CONSTANTS, LL);
    if (m == 0 && c.jacobian) {
      band_lu.swap(band);
      band_piv.swap(ipvt);
    }
  }

  // src/DisortSolver.cpp:1980-2124 (m = 0), at the user levels.
  void fluxes(GeneralOutput& out) {
    for (auto& v : out.f) v.assign(ntau, 0.0);
    std::vector<Real> ec(N), inten(N);
    for (int lu = 0; lu < ntau; ++lu) {
      const int lyu = lyr_of[lu];
      if (lyrcut && lyu > ncut - 1) continue;
      const Real ut = sutau[lu];
      Real fact = 0.0, dirint = 0.0, fldir = 0.0, fdirect = 0.0;
      if (c.beam_flux > 0.0) {
        if (c.spherical) {
          fact = std::exp(-ut / chap[lyu]);
          fdirect = std::abs(c.beam_mu) * c.beam_flux * std::exp(-utau[lu] / chap[lyu]);
        } else {
          fact = std::exp(-ut / c.beam_mu);
          fdirect = c.beam_mu * c.beam_flux * std::exp(-utau[lu] / c.beam_mu);
        }
        dirint = c.beam_flux * fact;
        fldir = c.beam_mu * c.beam_flux * fact;
      }
      const int off = lyu * N;
      for (int jq = 0; jq < NN; ++jq) ec[jq] = LL[off + jq] * std::exp(-K(lyu, jq) * (ut - taus[lyu + 1]));
      for (int jq = NN; jq < N; ++jq) ec[jq] = LL[off + jq] * std::exp(-K(lyu, jq) * (ut - taus[lyu]));
      for (int i = 0; i < N; ++i) {
        Real s = 0.0;
        for (int j = 0; j < N; ++j) s += G(lyu, i, j) * ec[j];
        inten[i] = s;
      }
      if (c.beam_flux > 0.0)
        for (int i = 0; i < N; ++i) inten[i] += beam_at(lyu, i, ut, fact);
      if (c.thermal)
        for (int i = 0; i < N; ++i) inten[i] += z0[(size_t)lyu * N + i] + z1[(size_t)lyu * N + i] * ut;
      Real mi = 0.0, mid = 0.0, miu = 0.0, fldn = 0.0, flup = 0.0;
      for (int iq = 0; iq < NN; ++iq) {
        const int ci = NN - iq - 1;
        mi += cwt[ci] * inten[iq];
        mid += cwt[ci] * inten[iq];
        fldn += cwt[ci] * inten[iq] * cmu[ci];
      }
      for (int iq = NN; iq < N; ++iq) {
        const int ci = iq - NN;
        mi += cwt[ci] * inten[iq];
        miu += cwt[ci] * inten[iq];
        flup += cwt[ci] * inten[iq] * cmu[ci];
      }
      flup *= 2.0 * M_PI;
      fldn *= 2.0 * M_PI;
      const Real fdn = fldn + fldir - fdirect;
      mi = (2.0 * M_PI * mi + dirint) / (4.0 * M_PI);
      const Real midir = dirint / (4.0 * M_PI);
      mid = (2.0 * M_PI * mid) / (4.0 * M_PI);
      miu = (2.0 * M_PI * miu) / (4.0 * M_PI);
      Real pls = 0.0;
      if (c.thermal) pls = xr0[lyu] + xr1[lyu] * ut;
      out.f[0][lu] = fdirect;
      out.f[1][lu] = fdn;
      out.f[2][lu] = flup;
      out.f[3][lu] = (1.0 - c.ssa[lyu]) * 4.0 * M_PI * (mi - pls);
      out.f[4][lu] = mi;
      out.f[5][lu] = mid;
      out.f[6][lu] = miu;
      out.f[7][lu] = midir;
    }
  }

  // src/DisortSolver.cpp:2938-3268; returns the largest relative change of
  // the accumulated intensities (m > 0) for the convergence test.
  Real user_intensities(int m, GeneralOutput& out) {
    const Real d0 = (m == 0) ? 1.0 : 0.0;
    const bool beam = c.beam_flux > 0.0;
    for (int lc = 0; lc < ncut; ++lc)
      for (int iq = 0; iq < N; ++iq) {
        const Real ll = LL[(size_t)lc * N + iq];
        for (int iu = 0; iu < numu; ++iu) GU(lc, iu, iq) *= ll;
      }
    std::vector<Real> uum((size_t)ntau * numu, 0.0), wk(NN);
    for (int lu = 0; lu < ntau; ++lu) {
      const Real exp0 = beam ? std::exp(-sutau[lu] / c.beam_mu) : 0.0;
      const int lyu = lyr_of[lu];
      for (int iu = 0; iu < numu; ++iu) {
        const Real mu = umu[iu];
        Real& res = uum[(size_t)lu * numu + iu];
        if (lyrcut && lyu >= ncut) {
          res = 0.0;
          continue;
        }
        const bool neg = mu < 0.0;
        const int l0 = neg ? 0 : lyu + 1, l1 = neg ? lyu - 1 : ncut - 1;
        const Real sgn = neg ? -1.0 : 1.0;
        Real palint = 0.0, plkint = 0.0;
        for (int lc = l0; lc <= l1; ++lc) {
          const Real dt = dtaus[lc];
          const Real e1 = std::exp((sutau[lu] - taus[lc]) / mu);
          const Real e2 = std::exp((sutau[lu] - taus[lc + 1]) / mu);
          if (c.thermal && m == 0) {
            const Real f0n = sgn * (e1 - e2);
            const Real f1n = sgn * ((taus[lc] + mu) * e1 - (taus[lc + 1] + mu) * e2);
            plkint += z0u[(size_t)lc * numu + iu] * f0n + z1u[(size_t)lc * numu + iu] * f1n;
          }
          if (beam) {
            const Real den = 1.0 + mu / c.beam_mu;
            const Real expn = (std::abs(den) < 0.0001) ? (dt / c.beam_mu) * exp0
                                                       : (e1 * btrans[lc] - e2 * btrans[lc + 1]) * sgn / den;
            palint += zbu[(size_t)lc * numu + iu] * expn;
          }
          for (int iq = 0; iq < NN; ++iq) {
            wk[iq] = std::exp(K(lc, iq) * dt);
            const Real den = 1.0 + mu * K(lc, iq);
            const Real expn = (std::abs(den) < 0.0001) ? (dt / mu) * e2 : sgn * (e1 * wk[iq] - e2) / den;
            palint += GU(lc, iu, iq) * expn;
          }
          for (int iq = NN; iq < N; ++iq) {
            const Real den = 1.0 + mu * K(lc, iq);
            const Real expn = (std::abs(den) < 0.0001) ? -(dt / mu) * e1 : sgn * (e1 - e2 * wk[N - 1 - iq]) / den;
            palint += GU(lc, iu, iq) * expn;
          }
        }
        const Real dtau1 = sutau[lu] - taus[lyu], dtau2 = sutau[lu] - taus[lyu + 1];
        if ((std::abs(dtau1) >= 1e-6 || !neg) && (std::abs(dtau2) >= 1e-6 || neg)) {
          const Real e1 = neg ? std::exp(dtau1 / mu) : 0.0;
          const Real e2 = neg ? 0.0 : std::exp(dtau2 / mu);
          if (beam) {
            const Real den = 1.0 + mu / c.beam_mu;
            Real expn;
            if (std::abs(den) < 0.0001) expn = (dtau1 / c.beam_mu) * exp0;
            else if (neg) expn = (exp0 - btrans[lyu] * e1) / den;
            else expn = (exp0 - btrans[lyu + 1] * e2) / den;
            palint += zbu[(size_t)lyu * numu + iu] * expn;
          }
          const Real dl = dtaus[lyu];
          for (int iq = 0; iq < NN; ++iq) {
            const Real k = K(lyu, iq), den = 1.0 + mu * k;
            Real expn;
            if (std::abs(den) < 0.0001) expn = -dtau2 / mu * e2;
            else if (neg) expn = (std::exp(-k * dtau2) - std::exp(k * dl) * e1) / den;
            else expn = (std::exp(-k * dtau2) - e2) / den;
            palint += GU(lyu, iu, iq) * expn;
          }
          for (int iq = NN; iq < N; ++iq) {
            const Real k = K(lyu, iq), den = 1.0 + mu * k;
            Real expn;
            if (std::abs(den) < 0.0001) expn = -(dtau1 / mu) * e1;
            else if (neg) expn = (std::exp(-k * dtau1) - e1) / den;
            else expn = (std::exp(-k * dtau1) - std::exp(-k * dl) * e2) / den;
            palint += GU(lyu, iu, iq) * expn;
          }
          if (c.thermal && m == 0) {
            const Real expn = neg ? e1 : e2;
            const Real fact = neg ? taus[lyu] + mu : taus[lyu + 1] + mu;
            const Real f0n = 1.0 - expn;
            const Real f1n = sutau[lu] + mu - fact * expn;
            plkint += z0u[(size_t)lyu * numu + iu] * f0n + z1u[(size_t)lyu * numu + iu] * f1n;
          }
        }
        Real bndint = 0.0;
        if (neg && m == 0) {
          bndint = (c.iso_top + pl_top) * std::exp(sutau[lu] / mu);
        } else if (!neg) {
          if (c.diffusion_bc) {
            if (lyrcut || m > 0) {
              res = palint + plkint;
              continue;
            }
            bndint = (pl_bot + mu * pl_bot_deriv) * std::exp((sutau[lu] - taus[L]) / mu);
          } else if (lyrcut || (c.lambertian && m > 0)) {
            res = palint + plkint;
            continue;
          } else {
            const int lb = L - 1;
            std::vector<Real> wb(NN);
            for (int jq = NN; jq < N; ++jq) wb[jq - NN] = std::exp(-K(lb, jq) * dtaus[lb]);
            Real bnddfu = 0.0;
            for (int iq = NN - 1; iq >= 0; --iq) {
              Real dfu = 0.0;
              for (int jq = 0; jq < NN; ++jq) dfu += G(lb, iq, jq) * LL[(size_t)lb * N + jq];
              for (int jq = NN; jq < N; ++jq) dfu += G(lb, iq, jq) * LL[(size_t)lb * N + jq] * wb[jq - NN];
              if (beam) dfu += zb[(size_t)lb * N + iq] * btrans[L];
              dfu += d0 * (z0[(size_t)lb * N + iq] + z1[(size_t)lb * N + iq] * taus[L]);
              bnddfu += (1.0 + d0) * RU(iu, NN - iq) * cmu[NN - iq - 1] * cwt[NN - iq - 1] * dfu;
            }
            Real bnddir = 0.0;
            if (c.beam_flux > 0.0 || c.beam_mu > 0.0)
              bnddir = c.beam_mu * c.beam_flux / M_PI * RU(iu, 0) * btrans[L];
            bndint = (bnddfu + bnddir + d0 * eu[iu] * pl_bot + c.iso_bot) * std::exp((sutau[lu] - taus[L]) / mu);
          }
        }
        res = palint + plkint + bndint;
      }
    }
    if (m == 0) {
      for (int lu = 0; lu < ntau; ++lu)
        for (int iu = 0; iu < numu; ++iu) {
          const Real v = uum[(size_t)lu * numu + iu];
          out.inten_avg[(size_t)lu * numu + iu] = v;
          for (int j = 0; j < nphi; ++j) out.inten[((size_t)lu * numu + iu) * nphi + j] = v;
        }
      return 0.0;
    }
    Real maxr = 0.0;
    for (int j = 0; j < nphi; ++j) {
      const Real phirad = (c.phi_user[j] - c.beam_phi) * M_PI / 180.0;
      const Real cp = std::cos(static_cast<Real>(m) * phirad);
      for (int lu = 0; lu < ntau; ++lu)
        for (int iu = 0; iu < numu; ++iu) {
          const Real az = uum[(size_t)lu * numu + iu] * cp;
          Real& u = out.inten[((size_t)lu * numu + iu) * nphi + j];
          u += az;
          const Real ua = std::abs(u);
          Real r;
          if (ua == 0.0) r = 1.0 + std::abs(az);
          else if (std::abs(az) == 0.0) r = 0.0;
          else r = std::abs(az) / ua;
          maxr = std::max(maxr, r);
        }
    }
    return maxr;
  }

  // src/DisortSolver.cpp:3274-3285
  static Real xi_func(Real u1, Real u2, Real tau) {
    const Real x1 = (u2 - u1) / (u2 * u1);
    const Real e1 = std::exp(-tau / u1);
    if (x1 != 0.0) return ((tau * x1 - 1.0) * std::exp(-tau / u2) + e1) / (x1 * x1 * u1 * u2);
    return tau * tau * e1 / (2.0 * u1 * u2);
  }

  // src/DisortSolver.cpp:3287-3328 (1-based layer arrays, tau = scaled
  // cumulative depths with tau[0] = 0).
  Real single_scat(Real dither, int layru, int nl, const std::vector<Real>& ph, const std::vector<Real>& om,
                   Real mu, Real tu) const {
    const Real mu0 = c.beam_mu, f0 = c.beam_flux;
    Real ans = 0.0, e0 = std::exp(-tu / mu0);
    if (std::abs(mu + mu0) <= dither) {
      for (int l = 1; l <= layru - 1; ++l) ans += om[l] * ph[l] * (taus[l] - taus[l - 1]);
      return f0 / (4.0 * M_PI * mu0) * e0 * (ans + om[layru] * ph[layru] * (tu - taus[layru - 1]));
    }
    if (mu > 0.0) {
      for (int l = layru; l <= nl; ++l) {
        const Real e1 = std::exp(-((taus[l] - tu) / mu + taus[l] / mu0));
        ans += om[l] * ph[l] * (e0 - e1);
        e0 = e1;
      }
    } else {
      for (int l = layru; l >= 1; --l) {
        const Real e1 = std::exp(-((taus[l - 1] - tu) / mu + taus[l - 1] / mu0));
        ans += om[l] * ph[l] * (e0 - e1);
        e0 = e1;
      }
    }
    return ans * f0 / (4.0 * M_PI * (1.0 + mu / mu0));
  }

  // src/DisortSolver.cpp:3330-3401 (lu, iu 1-based as in the reference).
  Real secondary_scat(int iu, int lu, Real ctheta, int layru) const {
    const Real tiny = 1.0e-4;
    const Real ut = utau[lu - 1], um = umu[iu - 1];
    Real dt = ut - tauc[layru];
    Real wbar = c.ssa[layru] * dt, fbar = fwd[layru] * wbar, stau = dt;
    for (int l = 0; l < layru; ++l) {
      wbar += c.dtau[l] * c.ssa[l];
      fbar += c.dtau[l] * c.ssa[l] * fwd[l];
      stau += c.dtau[l];
    }
    if (wbar <= tiny || fbar <= tiny || stau <= tiny || c.beam_flux <= tiny) return 0.0;
    fbar /= wbar;
    wbar /= stau;
    Real pspike = 1.0, gbar = 1.0, pl, plm1 = 1.0, plm2 = 0.0;
    for (int k = 1; k <= N - 1; ++k) {
      pl = ((2 * k - 1) * ctheta * plm1 - (k - 1) * plm2) / k;
      plm2 = plm1;
      plm1 = pl;
      pspike += gbar * (2.0 - gbar) * (Real)(2 * k + 1) * pl;
    }
    for (int k = N; k <= c.nmom; ++k) {
      pl = ((2 * k - 1) * ctheta * plm1 - (k - 1) * plm2) / k;
      plm2 = plm1;
      plm1 = pl;
      dt = ut - tauc[layru];
      gbar = c.pm(k, layru) * c.ssa[layru] * dt;
      for (int l = 0; l < layru; ++l) gbar += c.pm(k, l) * c.ssa[l] * c.dtau[l];
      const Real t = fbar * wbar * stau;
      gbar = (t <= tiny) ? 0.0 : gbar / t;
      pspike += gbar * (2.0 - gbar) * (Real)(2 * k + 1) * pl;
    }
    const Real umu0p = c.beam_mu / (1.0 - fbar * wbar);
    return c.beam_flux / (4.0 * M_PI) * (fbar * wbar) * (fbar * wbar) / (1.0 - fbar * wbar) * pspike *
           xi_func(-um, umu0p, ut);
  }

  // src/DisortSolver.cpp:3403-3524
  void nt_correction(GeneralOutput& out) {
    const Real dither = 100.0 * std::numeric_limits<double>::epsilon();
    const Real dtheta = 10.0, deg = std::acos(-1.0) / 180.0;
    std::vector<Real> phasa(ncut + 1), phasm(ncut + 1), phast(ncut + 1), om_a(ncut + 1), om_m(ncut + 1);
    for (int lc = 1; lc <= ncut; ++lc) {
      om_a[lc] = c.ssa[lc - 1];
      om_m[lc] = ssas[lc - 1];
    }
    for (int iu = 1; iu <= numu; ++iu) {
      const Real um = umu[iu - 1];
      Real theta0 = 0.0, thetap = 0.0;
      if (um < 0.0) {
        theta0 = std::acos(-c.beam_mu) / deg;
        thetap = std::acos(um) / deg;
      }
      for (int jp = 1; jp <= nphi; ++jp) {
        const Real phirad = (c.phi_user[jp - 1] - c.beam_phi) * deg;
        const Real ctheta =
            -c.beam_mu * um + std::sqrt((1.0 - c.beam_mu * c.beam_mu) * (1.0 - um * um)) * std::cos(phirad);
        for (int lc = 1; lc <= ncut; ++lc) phasa[lc] = phasm[lc] = 1.0;
        Real pl, plm1 = 1.0, plm2 = 0.0;
        for (int k = 1; k <= c.nmom; ++k) {
          pl = ((2 * k - 1) * ctheta * plm1 - (k - 1) * plm2) / k;
          plm2 = plm1;
          plm1 = pl;
          for (int lc = 1; lc <= ncut; ++lc) phasa[lc] += (Real)(2 * k + 1) * pl * c.pm(k, lc - 1);
          if (k <= N - 1)
            for (int lc = 1; lc <= ncut; ++lc) {
              const Real fl = fwd[lc - 1];
              phasm[lc] += (Real)(2 * k + 1) * pl * (c.pm(k, lc - 1) - fl) / (1.0 - fl);
            }
        }
        for (int lc = 1; lc <= ncut; ++lc) phast[lc] = phasa[lc] / (1.0 - fwd[lc - 1] * c.ssa[lc - 1]);
        for (int lu = 1; lu <= ntau; ++lu) {
          if (lyrcut && lyr_of[lu - 1] >= ncut) continue;
          const Real a = single_scat(dither, lyr_of[lu - 1] + 1, ncut, phast, om_a, um, sutau[lu - 1]);
          const Real b = single_scat(dither, lyr_of[lu - 1] + 1, ncut, phasm, om_m, um, sutau[lu - 1]);
          out.inten[((size_t)(lu - 1) * numu + (iu - 1)) * nphi + (jp - 1)] += a - b;
        }
        if (um < 0.0 && std::abs(theta0 - thetap) <= dtheta) {
          const int ltau = (utau[0] <= dither) ? 2 : 1;
          for (int lu = ltau; lu <= ntau; ++lu) {
            if (lyrcut && lyr_of[lu - 1] >= ncut) continue;
            out.inten[((size_t)(lu - 1) * numu + (iu - 1)) * nphi + (jp - 1)] -=
                secondary_scat(iu, lu, ctheta, lyr_of[lu - 1]);
          }
        }
      }
    }
  }

  // ---- Buras-Emde correction (src/DisortSolver.cpp:3532-3745, :3812-4019) ----
  // Tabulated phase function of layer lc at table angle it (0-based).
  Real tab(int lc, int it) const { return c.phase[(size_t)lc * c.nphase + it]; }
  // Table interval of x, clamped to an interval of the table (the
  // reference's locate; at x beyond the last angle the reference reads one
  // past the table, which the clamp replaces by the last interval).
  int interval(Real x) const {
    int j = locate_sorted(c.mu_phase, x);
    if (j < 0) j = 0;
    if (j > c.nphase - 2) j = c.nphase - 2;
    return j;
  }

  // Residual phase function phas2 [ntau][nphase] normalised to integrate to
  // 2, and its equal-|area| integration grid (prepDoubleScatIntegr).
  struct Residual {
    std::vector<Real> phas2, mu_eq, norm;
    std::vector<int> neg;
  };
  static constexpr int kNF = 100;

  Residual residual_phase() const {
    const int np = c.nphase;
    const Real tiny = 1.0e-4;
    Residual r;
    r.phas2.assign((size_t)ntau * np, 0.0);
    std::vector<Real> trunc(L);
    for (int it = 0; it < np; ++it) {
      // Delta-M truncated phase function (not divided by 1 - f), N terms.
      const Real x = c.mu_phase[it];
      for (int lc = 0; lc < L; ++lc) trunc[lc] = 1.0 - fwd[lc];
      Real p1 = 1.0, p2 = 0.0;
      for (int k = 1; k <= N - 1; ++k) {
        const Real p = ((2 * k - 1) * x * p1 - (k - 1) * p2) / k;
        p2 = p1;
        p1 = p;
        for (int lc = 0; lc < L; ++lc) trunc[lc] += (2 * k + 1) * p * (c.pm(k, lc) - fwd[lc]);
      }
      for (int lu = 0; lu < ntau; ++lu) {
        const int ll = lyr_of[lu];
        Real acc = 0.0;
        for (int lc = 0; lc < ll; ++lc) acc += (tab(lc, it) - trunc[lc]) * c.ssa[lc] * c.dtau[lc];
        acc += (tab(ll, it) - trunc[ll]) * c.ssa[ll] * (utau[lu] - tauc[ll]);
        r.phas2[(size_t)lu * np + it] = acc;
      }
    }
    r.mu_eq.assign((size_t)ntau * kNF, 0.0);
    r.neg.assign((size_t)ntau * kNF, 0);
    r.norm.assign(ntau, 0.0);
    std::vector<Real> cum(np);
    for (int lu = 0; lu < ntau; ++lu) {
      Real* ph = &r.phas2[(size_t)lu * np];
      const int ll = lyr_of[lu];
      Real fbar = fwd[ll] * c.ssa[ll] * (utau[lu] - tauc[ll]);
      for (int lc = 0; lc < ll; ++lc) fbar += c.ssa[lc] * c.dtau[lc] * fwd[lc];
      const Real scale = (fbar <= tiny || c.beam_flux <= tiny) ? 0.0 : 1.0 / fbar;
      for (int it = 0; it < np; ++it) ph[it] *= scale;
      Real area = 0.0;
      for (int it = 1; it < np; ++it) area += (c.mu_phase[it] - c.mu_phase[it - 1]) * 0.5 * (ph[it] + ph[it - 1]);
      if (area != 0.0)
        for (int it = 0; it < np; ++it) ph[it] *= 2.0 / area;
      // Grid equidistant in the cumulative integral of |phas2|.
      cum[0] = 0.0;
      for (int it = 1; it < np; ++it)
        cum[it] = cum[it - 1] + (c.mu_phase[it] - c.mu_phase[it - 1]) * 0.5 * (std::abs(ph[it]) + std::abs(ph[it - 1]));
      Real* me = &r.mu_eq[(size_t)lu * kNF];
      int* ng = &r.neg[(size_t)lu * kNF];
      const Real df = cum[np - 1] / (kNF - 1);
      me[0] = -1.0;
      ng[0] = ph[0] > 0.0 ? 0 : 1;
      Real target = 0.0;
      int it = 0;
      for (int i = 1; i < kNF - 1; ++i) {
        target += df;
        while (it < np - 2 && cum[it + 1] < target) ++it;
        const Real w = (target - cum[it]) / (cum[it + 1] - cum[it]);
        me[i] = c.mu_phase[it] + w * (c.mu_phase[it + 1] - c.mu_phase[it]);
        if (ph[it] > 0.0 && ph[it + 1] > 0.0) ng[i] = 0;
        else if (ph[it] < 0.0 && ph[it + 1] < 0.0) ng[i] = 1;
        else ng[i] = (ph[it] + w * (ph[it + 1] - ph[it])) > 0.0 ? 0 : 1;
      }
      me[kNF - 1] = 1.0;
      ng[kNF - 1] = ph[np - 1] > 0.0 ? 0 : 1;
      r.norm[lu] = cum[np - 1] / ((kNF - 1) * M_PI);
    }
    return r;
  }

  // Azimuthal integral of phas2(mu_1) phas2(theta_12) over the equal-area
  // grid (calcPhaseSquared).
  Real phase_squared(const Residual& r, int lu, Real ct) const {
    const int np = c.nphase;
    const Real* ph = &r.phas2[(size_t)lu * np];
    const Real* me = &r.mu_eq[(size_t)lu * kNF];
    const int* ng = &r.neg[(size_t)lu * kNF];
    const Real st = std::sqrt(1.0 - ct * ct);
    const Real* mp = c.mu_phase.data();
    Real sum = 0.0;
    for (int j = 0; j < kNF; ++j) {
      const Real u = me[j];
      Real phint = 0.0;
      if (ct == 1.0 || u == 1.0) {
        const int k = interval(u * ct);
        phint = M_PI * (ph[k] + (u * ct - mp[k]) / (mp[k + 1] - mp[k]) * (ph[k + 1] - ph[k]));
        if (ct == 1.0) phint /= 2.0;
      } else {
        const Real su = std::sqrt(1.0 - u * u);
        const Real lo = ct * u - st * su;
        Real hi = ct * u + st * su;
        const bool cut = u < hi;
        if (cut) hi = u;
        if (lo < hi) {
          const int kmin = interval(lo), kmax = interval(hi);
          // Piecewise-linear phas2 = D mu + C on each table interval.
          Real D = (ph[kmin + 1] - ph[kmin]) / (mp[kmin + 1] - mp[kmin]);
          Real C = ph[kmin] - mp[kmin] * D;
          phint += (D * ct * u + C) * M_PI / 2.0;
          for (int k = kmin + 1; k <= kmax; ++k) {
            const Real D2 = (ph[k + 1] - ph[k]) / (mp[k + 1] - mp[k]);
            const Real C2 = ph[k] - mp[k] * D2;
            phint += (D2 - D) * std::sqrt(1.0 - ct * ct - u * u + 2.0 * ct * u * mp[k] - mp[k] * mp[k]) +
                     ((D2 - D) * ct * u + C2 - C) * std::asin((ct * u - mp[k]) / (su * st));
            D = D2;
            C = C2;
          }
          if (cut)
            phint += -D * std::sqrt(1.0 - ct * ct + 2.0 * u * u * (ct - 1.0)) -
                     (D * ct * u + C) * std::asin((ct - 1.0) * u / (su * st));
          else
            phint += (D * ct * u + C) * M_PI / 2.0;
        }
      }
      const Real w = (j == 0 || j == kNF - 1) ? 0.5 : 1.0;
      sum += (ng[j] ? -w : w) * phint;
    }
    return sum * r.norm[lu];
  }

  // IMS term of the Buras-Emde correction (newSecondaryScat), lu and iu 0-based.
  Real buras_secondary(const Residual& r, int iu, int lu, int k, Real ct) const {
    const Real tiny = 1.0e-4;
    const int ll = lyr_of[lu];
    const Real dt = utau[lu] - tauc[ll];
    Real wbar = c.ssa[ll] * dt, fbar = fwd[ll] * wbar, stau = dt;
    for (int lc = 0; lc < ll; ++lc) {
      wbar += c.dtau[lc] * c.ssa[lc];
      fbar += c.dtau[lc] * c.ssa[lc] * fwd[lc];
      stau += c.dtau[lc];
    }
    if (wbar <= tiny || fbar <= tiny || stau <= tiny || c.beam_flux <= tiny) return 0.0;
    fbar /= wbar;
    wbar /= stau;
    const Real* ph = &r.phas2[(size_t)lu * c.nphase];
    const Real* mp = c.mu_phase.data();
    const Real p1 = ph[k] + (ct - mp[k]) / (mp[k + 1] - mp[k]) * (ph[k + 1] - ph[k]);
    const Real pspike = 2.0 * p1 - phase_squared(r, lu, ct);
    const Real fw = fbar * wbar;
    const Real umu0p = c.beam_mu / (1.0 - fw);
    return c.beam_flux / (4.0 * M_PI) * fw * fw / (1.0 - fw) * pspike * xi_func(-umu[iu], umu0p, utau[lu]);
  }

  void buras_correction(GeneralOutput& out) {
    const Real dither = 100.0 * std::numeric_limits<double>::epsilon();
    const Real dtheta = 10.0, deg = M_PI / 180.0;
    const Real theta0 = std::acos(-c.beam_mu) / deg;
    bool aureole = false;
    for (int iu = 0; iu < numu; ++iu)
      if (umu[iu] < 0.0 && std::abs(theta0 - std::acos(umu[iu]) / deg) <= dtheta) aureole = true;
    Residual res;
    if (aureole) res = residual_phase();
    std::vector<Real> phasm(ncut + 1), phast(ncut + 1), om_a(ncut + 1), om_m(ncut + 1);
    for (int lc = 1; lc <= ncut; ++lc) {
      om_a[lc] = c.ssa[lc - 1];
      om_m[lc] = ssas[lc - 1];
    }
    for (int iu = 0; iu < numu; ++iu) {
      const Real um = umu[iu];
      const bool ims = um < 0.0 && std::abs(theta0 - std::acos(um) / deg) <= dtheta;
      for (int jp = 0; jp < nphi; ++jp) {
        const Real phirad = (c.phi_user[jp] - c.beam_phi) * deg;
        const Real ct = -c.beam_mu * um + std::sqrt((1.0 - c.beam_mu * c.beam_mu) * (1.0 - um * um)) * std::cos(phirad);
        // Exact phase function interpolated from the table, and the Delta-M
        // phase function from N moments (1-based layer arrays).
        const int k = interval(ct);
        const Real* mp = c.mu_phase.data();
        for (int lc = 1; lc <= ncut; ++lc) {
          const Real pa = tab(lc - 1, k) + (ct - mp[k]) / (mp[k + 1] - mp[k]) * (tab(lc - 1, k + 1) - tab(lc - 1, k));
          phast[lc] = pa / (1.0 - fwd[lc - 1] * c.ssa[lc - 1]);
          phasm[lc] = 1.0;
        }
        Real p1 = 1.0, p2 = 0.0;
        for (int l = 1; l <= N - 1; ++l) {
          const Real p = ((2 * l - 1) * ct * p1 - (l - 1) * p2) / l;
          p2 = p1;
          p1 = p;
          for (int lc = 1; lc <= ncut; ++lc)
            phasm[lc] += (2.0 * l + 1) * p * (c.pm(l, lc - 1) - fwd[lc - 1]) / (1.0 - fwd[lc - 1]);
        }
        for (int lu = 0; lu < ntau; ++lu) {
          if (lyrcut && lyr_of[lu] >= ncut) continue;
          const Real a = single_scat(dither, lyr_of[lu] + 1, ncut, phast, om_a, um, sutau[lu]);
          const Real b = single_scat(dither, lyr_of[lu] + 1, ncut, phasm, om_m, um, sutau[lu]);
          out.inten[((size_t)lu * numu + iu) * nphi + jp] += a - b;
        }
        if (ims) {
          const int l0 = (utau[0] <= dither) ? 1 : 0;
          for (int lu = l0; lu < ntau; ++lu) {
            if (lyrcut && lyr_of[lu] >= ncut) continue;
            out.inten[((size_t)lu * numu + iu) * nphi + jp] -= buras_secondary(res, iu, lu, k, ct);
          }
        }
      }
    }
  }

  // src/DisortSolver.cpp:3751-3808
  void brdf_correction(GeneralOutput& out) {
    const Real deg = M_PI / 180.0;
    for (int iu = 0; iu < numu; ++iu) {
      const Real mu = umu[iu];
      if (mu <= 0.0) continue;
      for (int j = 0; j < nphi; ++j) {
        const Real dphi = (c.phi_user[j] - c.beam_phi) * deg;
        const Real exact = rho(mu, c.beam_mu, dphi);
        if (exact < 0.0) continue;
        Real approx = 0.0;
        for (int m = 0; m < N; ++m) approx += rho_beam_user[(size_t)iu * N + m] * std::cos(static_cast<Real>(m) * dphi);
        const Real dr = exact - approx;
        if (std::abs(dr) < 1e-15) continue;
        const Real to_bottom = std::exp(-taus[ncut] / c.beam_mu);
        for (int lu = 0; lu < ntau; ++lu) {
          if (lyrcut && lyr_of[lu] >= ncut) continue;
          out.inten[((size_t)lu * numu + iu) * nphi + j] +=
              c.beam_mu * c.beam_flux * dr * to_bottom * std::exp((sutau[lu] - taus[ncut]) / mu);
        }
      }
    }
  }

  GeneralOutput run() {
    scat_yes = false;
    for (int lc = 0; lc < L; ++lc)
      if (c.ssa[lc] > 0.0) { scat_yes = true; break; }
    deltam = false;
    for (int lc = 0; lc < L; ++lc)
      if (c.pm(N, lc) != 0.0) { deltam = true; break; }
    const bool needdeltam = deltam;
    deltam_plus = c.deltam_plus;
    if (deltam_plus) {
      for (int lc = 0; lc < L; ++lc) {
        const Real pm = c.pm(N, lc), pm1 = c.pm(N + 1, lc);
        if (pm < 1e-4 || pm1 < 0.7 * pm) { deltam_plus = false; break; }
      }
      deltam = deltam_plus ? false : needdeltam;
    }
    ntau = c.use_user_tau ? (int)c.tau_user.size() : L + 1;
    if (c.use_user_tau) {
      utau = c.tau_user;
    } else {
      utau.assign(L + 1, 0.0);
      for (int lc = 0; lc < L; ++lc) utau[lc + 1] = utau[lc] + c.dtau[lc];
    }
    const bool user_mu = c.use_user_mu && !c.only_fluxes;
    if (user_mu) umu = c.mu_user;
    numu = user_mu ? (int)c.mu_user.size() : N;
    nphi = c.only_fluxes ? 0 : (int)c.phi_user.size();
    do_user = !c.only_fluxes && c.use_user_mu;
    set_up();
    perturb(PS_SCALED_COEF, gl);
    perturb(PS_SCALED_SSA, ssas);
    perturb(PS_QUAD, cwt);

    GeneralOutput out;
    out.ntau = ntau;
    out.numu = numu;
    out.nphi = nphi;
    out.mu_angles = umu;
    out.inten.assign((size_t)ntau * numu * nphi, 0.0);
    out.inten_avg.assign((size_t)ntau * numu, 0.0);

    pl_top = pl_bot = pl_bot_deriv = 0.0;
    if (c.thermal) {
      pl_top = planck_band(c.wn_lo, c.wn_hi, c.temp_top) * c.emis_top;
      pl_bot = planck_band(c.wn_lo, c.wn_hi, c.temp_bot);
      if (c.diffusion_bc) {
        const Real bl = planck_band(c.wn_lo, c.wn_hi, c.temper[L - 1]);
        if (dtaus[L - 1] > 0.0) pl_bot_deriv = (pl_bot - bl) / dtaus[L - 1];
      }
    }
    // src/DisortSolver.cpp:206-224.  Its two user-angle shortcuts read
    // mu_user[1] / mu_user[2] of a 0-based array (one past the entries they
    // mean), so for valid sorted input they never fire; the oracle omits them.
    // At mu = +-1 every m > 0 term carries P_l^m(+-1) = 0 anyway.
    int naz = N - 1;
    if (c.beam_flux == 0.0 || std::abs(1.0 - c.beam_mu) < 1.e-5 || c.only_fluxes) naz = 0;

    ylmc.assign((size_t)(N + 1) * N, 0.0);
    ylm0.assign((size_t)N + 1, 0.0);
    ylmu.assign((size_t)(N + 1) * std::max(numu, 1), 0.0);
    kk.assign((size_t)L * N, 0.0);
    gc.assign((size_t)L * N * N, 0.0);
    cc.assign((size_t)L * N * N, 0.0);
    gu.assign((size_t)L * numu * N, 0.0);
    zb.assign((size_t)L * N, 0.0);
    zbs.assign((size_t)L * N, 0.0);
    zraw.assign((size_t)L * N, 0.0);
    xba.assign(L, 0.0);
    z0.assign((size_t)L * N, 0.0);
    z1.assign((size_t)L * N, 0.0);
    xr0.assign(L, 0.0);
    xr1.assign(L, 0.0);
    zbu.assign((size_t)L * numu, 0.0);
    z0u.assign((size_t)L * numu, 0.0);
    z1u.assign((size_t)L * numu, 0.0);
    rq.assign((size_t)NN * (NN + 1), 0.0);
    eq.assign(NN, 0.0);
    ru.assign((size_t)numu * (NN + 1), 0.0);
    eu.assign(numu, 0.0);
    rho_beam_user.assign((size_t)numu * N, 0.0);
    gmu.assign(NMUG, 0.0);
    gwt.assign(NMUG, 0.0);
    gauss_quadrature(NMUG / 2, gmu.data(), gwt.data());
    if (c.jacobian) {
      dz0.assign((size_t)2 * L * N, 0.0);
      dz1.assign((size_t)2 * L * N, 0.0);
      dslope.assign((size_t)2 * L, 0.0);
      dicpt.assign((size_t)2 * L, 0.0);
    }

    int kconv = 0;
    for (int m = 0; m <= naz; ++m) {
      out.modes = m + 1;
      legendre_flat(N, m, N - 1, N - 1, cmu.data(), ylmc.data());
      if (do_user) legendre_flat(numu, m, numu - 1, N - 1, umu.data(), ylmu.data());
      perturb(PS_LEGENDRE, ylmc);
      perturb(PS_LEGENDRE, ylmu);
      for (int lc = 0; lc < L; ++lc) {
        solve_eigen(lc, m);
        if (do_user) interp_eigvec(lc, m);
      }
      perturb(PS_EIGVAL, kk);
      perturb(PS_EIGVEC, gc);
      perturb(PS_USER_EIGVEC, gu);
      if (c.beam_flux > 0.0) {
        if (c.spherical) beam_source_spherical(m);
        else beam_source(m);
      }
      perturb(PS_BEAM, zb);
      if (c.thermal && m == 0) thermal_source();
      perturb(PS_THERMAL, z0);
      perturb(PS_THERMAL, z1);
      if (do_user)
        for (int lc = 0; lc < L; ++lc) interp_source(lc, m);
      surface_bidir(m);
      if (!c.lambertian && do_user && c.beam_flux > 0.0)
        for (int iu = 0; iu < numu; ++iu) rho_beam_user[(size_t)iu * N + m] = RU(iu, 0);
      boundary_solve(m);
      if (m == 0) fluxes(out);
      if (m == 0 && c.jacobian && c.thermal) jacobian(out);
      if (do_user) {
        const Real r = user_intensities(m, out);
        if (m > 0 && c.accur > 0.0) {
          if (r <= c.accur) ++kconv;
          if (kconv >= 2) break;
        }
      }
    }
    const bool corint = c.corr_buras && deltam && c.beam_flux > 0.0 && scat_yes && !c.only_fluxes;
    if (corint) {
      // Nakajima-Tanaka unless a tabulated phase function and
      // intensity_corr_nakajima = false select Buras-Emde (:369-377).
      const bool buras = !c.corr_nakajima && c.nphase > 0;
      std::vector<Real> before = out.inten;
      if (buras) buras_correction(out);
      else nt_correction(out);
      // Mutation hook: scale the correction itself.
      std::vector<Real> d(out.inten.size());
      for (size_t q = 0; q < d.size(); ++q) d[q] = out.inten[q] - before[q];
      perturb(buras ? PS_BURAS : PS_NT, d);
      for (size_t q = 0; q < d.size(); ++q) out.inten[q] = before[q] + d[q];
    }
    if (c.corr_buras && !deltam_plus && !c.lambertian && !c.only_fluxes && c.use_user_mu && c.beam_flux > 0.0)
      brdf_correction(out);
    for (auto& f : out.f) perturb(PS_FLUX, f);
    perturb(PS_FLUX, out.inten);
    return out;
  }
};

}  // namespace

// SPECIAL_BC (src/DisortSolver.cpp:4025-4163): the albedo and the
// transmissivity of the medium from its response to unit isotropic
// illumination at the top (and, with several layers, at the bottom) over a
// black surface, without sources or absorption cut-off, at the user polar
// angles doubled to -mu .. +mu (:54-58, :854-865); with surface_albedo > 0
// the spherical-albedo correction (:4126-4153).  Each illumination is one
// general solve of the same medium: the azimuthally averaged intensity it
// gives at the top (bottom) is the reference's source integral plus the
// directly transmitted unit boundary intensity exp(-tau*/mu), and its
// fluxes are pi times the spherical albedo / transmissivity sums.
GeneralOutput special_bc_solve(const GeneralConfig& c) {
  GeneralConfig b = c;
  b.special_bc = false;
  b.no_cutoff = true;
  b.index_from_bottom = false;
  b.beam_flux = 0.0;
  b.iso_top = 1.0;
  b.iso_bot = 0.0;
  b.albedo = 0.0;
  b.lambertian = true;
  b.brdf = BRDF_NONE;
  b.thermal = false;
  b.spherical = false;
  b.diffusion_bc = false;
  b.only_fluxes = false;
  b.corr_buras = b.corr_nakajima = false;
  b.use_user_tau = false;
  b.tau_user.clear();
  b.accur = 0.0;
  b.phi_user.assign(1, 0.0);
  b.nphase = 0;
  // Positive user cosines mu_1 < ... < mu_h, or the positive quadrature
  // nodes without user angles; doubled to [-mu_h .. -mu_1, mu_1 .. mu_h].
  std::vector<Real> pos;
  if (c.use_user_mu) {
    pos = c.mu_user;
  } else {
    std::vector<Real> cm(c.nstr), cw(c.nstr);
    gauss_quadrature(c.nstr / 2, cm.data(), cw.data());
    pos.assign(cm.begin(), cm.begin() + c.nstr / 2);
  }
  const int h = (int)pos.size();
  // The base solves take the cosines sorted increasing (the quadrature
  // nodes come in decreasing order); rank[k] places mu_k among them.
  std::vector<Real> sorted = pos;
  std::sort(sorted.begin(), sorted.end());
  std::vector<int> rank(h);
  for (int k = 0; k < h; ++k) rank[k] = (int)(std::lower_bound(sorted.begin(), sorted.end(), pos[k]) - sorted.begin());
  b.use_user_mu = true;
  b.mu_user.assign(2 * h, 0.0);
  for (int k = 0; k < h; ++k) {
    b.mu_user[h + k] = sorted[k];
    b.mu_user[k] = -sorted[h - 1 - k];
  }
  const int L = c.nlyr;
  GeneralOutput top = general_solve(b);
  GeneralOutput out;
  out.ntau = L + 1;
  out.numu = 2 * h;
  out.nphi = (int)c.phi_user.size();
  for (auto& f : out.f) f.assign(out.ntau, 0.0);
  out.inten.assign((size_t)out.ntau * out.numu * out.nphi, 0.0);
  out.inten_avg.assign((size_t)out.ntau * out.numu, 0.0);
  out.mu_angles.assign(2 * h, 0.0);
  for (int k = 0; k < h; ++k) {
    out.mu_angles[h + k] = pos[k];
    out.mu_angles[k] = -pos[h - 1 - k];
  }
  out.modes = top.modes;
  out.albedo_medium.assign(2 * h, 0.0);
  out.transmissivity_medium.assign(2 * h, 0.0);
  auto avg = [&](const GeneralOutput& o, int lu, int iu) { return o.inten_avg[(size_t)lu * o.numu + iu]; };
  for (int k = 0; k < h; ++k) out.albedo_medium[k] = avg(top, 0, h + rank[k]);
  GeneralOutput bot;
  const GeneralOutput* sph = &top;
  if (L == 1) {
    for (int k = 0; k < h; ++k) out.transmissivity_medium[k] = avg(top, L, h - 1 - rank[k]);
  } else {
    b.iso_top = 0.0;
    b.iso_bot = 1.0;
    bot = general_solve(b);
    for (int k = 0; k < h; ++k) out.transmissivity_medium[k] = avg(bot, 0, h + rank[k]);
    sph = &bot;
  }
  if (c.albedo > 0.0) {
    const Real sflup = sph->f[2][0] / M_PI, sfldn = sph->f[1][L] / M_PI;
    const Real sphalb = (L == 1) ? sflup : sfldn, sphtrn = (L == 1) ? sfldn : sflup;
    const Real corr = c.albedo / (1.0 - c.albedo * sphalb);
    for (int k = 0; k < h; ++k) {
      const Real t = out.transmissivity_medium[k];
      out.albedo_medium[k] += corr * sphtrn * t;
      out.transmissivity_medium[k] += corr * sphalb * t;
    }
    out.spherical_albedo = sphalb;
    out.spherical_transmissivity = sphtrn;
  }
  return out;
}

GeneralOutput general_solve(const GeneralConfig& cin) {
  if (cin.special_bc) {
    validate_general_config(cin);
    return special_bc_solve(cin);
  }
  if (!cin.index_from_bottom) {
    validate_general_config(cin);
    GeneralSolver s(cin);
    return s.run();
  }
  // src/IndexConversion.hpp:16-38 and :70-97
  GeneralConfig c = cin;
  const int L = c.nlyr, ps = c.pstride();
  std::reverse(c.dtau.begin(), c.dtau.end());
  std::reverse(c.ssa.begin(), c.ssa.end());
  for (int lc = 0; lc < L / 2; ++lc)
    std::swap_ranges(c.pmom.begin() + (size_t)lc * ps, c.pmom.begin() + (size_t)(lc + 1) * ps,
                     c.pmom.begin() + (size_t)(L - 1 - lc) * ps);
  if (!c.temper.empty()) std::reverse(c.temper.begin(), c.temper.end());
  if (!c.alt.empty()) std::reverse(c.alt.begin(), c.alt.end());
  if (c.use_user_tau) {
    Real total = 0.0;
    for (int lc = 0; lc < L; ++lc) total += c.dtau[lc];
    for (auto& t : c.tau_user) t = total - t;
    std::reverse(c.tau_user.begin(), c.tau_user.end());
  }
  validate_general_config(c);
  GeneralSolver s(c);
  GeneralOutput o = s.run();
  for (auto& v : o.f) std::reverse(v.begin(), v.end());
  reverse_jacobian(o.jac, o.ntau, L);
  const size_t lvl = (size_t)o.numu * o.nphi, lva = (size_t)o.numu;
  for (int a = 0, b = o.ntau - 1; a < b; ++a, --b) {
    std::swap_ranges(o.inten.begin() + a * lvl, o.inten.begin() + (a + 1) * lvl, o.inten.begin() + b * lvl);
    std::swap_ranges(o.inten_avg.begin() + a * lva, o.inten_avg.begin() + (a + 1) * lva, o.inten_avg.begin() + b * lva);
  }
  return o;
}

}  // namespace ORACLE_NS
