// C entry points of the oracle for ctypes (tests and bench.py's cpu_baseline).
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <chrono>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/disort_amd.h"
#include "oracle.hpp"

namespace {
void put_err(char* err, int errlen, const char* msg) {
  if (err && errlen > 0) {
    std::strncpy(err, msg, errlen - 1);
    err[errlen - 1] = 0;
  }
}

oracle::FluxConfig from_batch(const disort_flux_batch* b, int64_t i) {
  oracle::FluxConfig c;
  c.nstr = b->nstr[i];
  c.nlyr = b->nlyr[i];
  c.nmom = b->pstride[i] - 1;  // rows hold nmom+1 = pstride entries (>= nstr+1)
  c.thermal = (b->flags & DISORT_F_THERMAL) != 0;
  c.spherical = (b->flags & DISORT_F_SPHERICAL) != 0;
  c.deltam_plus = (b->flags & DISORT_F_DELTA_M_PLUS) != 0;
  c.diffusion_bc = (b->flags & DISORT_F_DIFFUSION_BC) != 0;
  const double* s = b->scal + i * DISORT_NSCAL;
  c.beam_flux = s[DISORT_S_BEAM_FLUX];
  c.beam_mu = s[DISORT_S_BEAM_MU];
  c.iso_top = s[DISORT_S_ISO_TOP];
  c.iso_bot = s[DISORT_S_ISO_BOT];
  c.temp_top = s[DISORT_S_TEMP_TOP];
  c.temp_bot = s[DISORT_S_TEMP_BOT];
  c.emis_top = s[DISORT_S_EMIS_TOP];
  c.albedo = s[DISORT_S_ALBEDO];
  c.wn_lo = s[DISORT_S_WN_LO];
  c.wn_hi = s[DISORT_S_WN_HI];
  c.radius = s[DISORT_S_RADIUS];
  const int64_t lo = b->lay_off[i], vo = b->lay_off[i] + i;
  c.dtau.assign(b->dtau + lo, b->dtau + lo + c.nlyr);
  c.ssa.assign(b->ssa + lo, b->ssa + lo + c.nlyr);
  c.pmom.assign(b->pmom + b->mom_off[i], b->pmom + b->mom_off[i] + (size_t)c.nlyr * b->pstride[i]);
  if (c.thermal && b->temper) c.temper.assign(b->temper + vo, b->temper + vo + c.nlyr + 1);
  if (c.spherical && b->alt) c.alt.assign(b->alt + vo, b->alt + vo + c.nlyr + 1);
  return c;
}
}  // namespace

extern "C" {

int oracle_planck(double lo, double hi, double t, double* out, char* err, int errlen) {
  try {
    *out = oracle::planck_band(lo, hi, t);
    return 0;
  } catch (const std::exception& e) {
    put_err(err, errlen, e.what());
    return 1;
  }
}

int oracle_planck_deriv(double lo, double hi, double t, double* out, char* err, int errlen) {
  try {
    *out = oracle::planck_band_deriv(lo, hi, t);
    return 0;
  } catch (const std::exception& e) {
    put_err(err, errlen, e.what());
    return 1;
  }
}

void oracle_legendre_flat(int nmu, int m, int maxmu, int twonm1, const double* mu, double* ylm) {
  oracle::legendre_flat(nmu, m, maxmu, twonm1, mu, ylm);
}

int oracle_gauss(int nn, double* cmu, double* cwt) {
  try {
    oracle::gauss_quadrature(nn, cmu, cwt);
    return 0;
  } catch (...) {
    return 1;
  }
}

void oracle_band_factor(double* abd, int lda, int n, int ml, int mu, int* ipvt) {
  oracle::band_factor(abd, lda, n, ml, mu, ipvt);
}

void oracle_band_solve(const double* abd, int lda, int n, int ml, int mu, const int* ipvt, double* b) {
  oracle::band_solve(abd, lda, n, ml, mu, ipvt, b);
}

void oracle_sym_eigen(int n, const double* a, double* evals, double* evecs) {
  oracle::sym_eigen_jacobi(n, a, evals, evecs);
}

int oracle_cholesky(int n, const double* a, double* l) { return oracle::cholesky_lower(n, a, l) ? 0 : 1; }

double oracle_chapman(int lc, double tau_frac, const double* zd, const double* dtau, int nlyr, double r, double mu0) {
  return oracle::chapman(lc, tau_frac, zd, dtau, nlyr, r, mu0);
}

/* One instance of the batch; out is [8][nlyr+1].  Returns 0, 1 (invalid
 * argument) or 2 (runtime error), with the message in err. */
int oracle_flux_one(const disort_flux_batch* b, int64_t i, int index_from_bottom, double* out, char* err, int errlen) {
  try {
    oracle::FluxConfig c = from_batch(b, i);
    c.index_from_bottom = index_from_bottom != 0;
    oracle::FluxOutput o = oracle::flux_solve(c);
    const int nl = c.nlyr + 1;
    for (int f = 0; f < 8; ++f) std::memcpy(out + (size_t)f * nl, o.f[f].data(), sizeof(double) * nl);
    return 0;
  } catch (const std::invalid_argument& e) {
    put_err(err, errlen, e.what());
    return 1;
  } catch (const std::exception& e) {
    put_err(err, errlen, e.what());
    return 2;
  }
}

/* One instance with temperature Jacobians: out as oracle_flux_one, jac
 * [6][nlyr+1][nlyr+2] in the field order flux_up, flux_down,
 * mean_intensity, mean_intensity_down, mean_intensity_up,
 * flux_tau_divergence. */
int oracle_flux_jacobian_one(const disort_flux_batch* b, int64_t i, int index_from_bottom, double* out, double* jac,
                             char* err, int errlen) {
  try {
    oracle::FluxConfig c = from_batch(b, i);
    c.index_from_bottom = index_from_bottom != 0;
    c.jacobian = true;
    oracle::FluxOutput o = oracle::flux_solve(c);
    const int nl = c.nlyr + 1;
    for (int f = 0; f < 8; ++f) std::memcpy(out + (size_t)f * nl, o.f[f].data(), sizeof(double) * nl);
    const size_t per = (size_t)nl * (nl + 1);
    for (int f = 0; f < 6; ++f) {
      if (o.jac[f].size() != per) throw std::runtime_error("oracle: no Jacobian (thermal emission off)");
      for (size_t q = 0; q < per; ++q) jac[f * per + q] = (double)o.jac[f][q];
    }
    return 0;
  } catch (const std::invalid_argument& e) {
    put_err(err, errlen, e.what());
    return 1;
  } catch (const std::exception& e) {
    put_err(err, errlen, e.what());
    return 2;
  }
}

/* Whole batch into b->out ([8][total_levels]) with an OpenMP static loop over
 * instances -- the structure of solveFluxSpectralBandsImpl<N>
 * (src/python_bindings.cpp:170-190).  Returns the number of failed instances;
 * *seconds receives the wall time of the loop. */
int64_t oracle_flux_batch(const disort_flux_batch* b, int nthreads, double* seconds) {
  int64_t failed = 0;
  const auto t0 = std::chrono::steady_clock::now();
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(static) reduction(+ : failed)
  for (int64_t i = 0; i < b->n; ++i) {
    try {
      oracle::FluxConfig c = from_batch(b, i);
      oracle::FluxOutput o = oracle::flux_solve(c);
      const int64_t vo = b->lay_off[i] + i;
      for (int f = 0; f < 8; ++f)
        for (int l = 0; l <= c.nlyr; ++l) b->out[f * b->total_levels + vo + l] = o.f[f][l];
      if (b->status) b->status[i] = 0;
    } catch (...) {
      ++failed;
      if (b->status) b->status[i] = 1;
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return failed;
}

// Column i of a general batch as an oracle configuration.
static oracle::GeneralConfig general_config(const disort_general_batch* b, int64_t i, int index_from_bottom) {
  oracle::GeneralConfig c;
  const int L = b->nlyr, P = b->pstride;
  c.nlyr = L;
  c.nstr = b->nstr;
  c.nmom = P - 1;
  c.thermal = (b->flags & DISORT_F_THERMAL) != 0;
  c.spherical = (b->flags & DISORT_F_SPHERICAL) != 0;
  c.deltam_plus = (b->flags & DISORT_F_DELTA_M_PLUS) != 0;
  c.diffusion_bc = (b->flags & DISORT_F_DIFFUSION_BC) != 0;
  c.use_user_tau = (b->flags & DISORT_G_USER_TAU) != 0;
  c.use_user_mu = (b->flags & DISORT_G_USER_MU) != 0;
  c.only_fluxes = (b->flags & DISORT_G_ONLY_FLUXES) != 0;
  const int bf = b->flags;
  c.brdf = (bf & DISORT_G_BRDF_RPV)       ? oracle::BRDF_RPV
           : (bf & DISORT_G_BRDF_HAPKE)   ? oracle::BRDF_HAPKE
           : (bf & DISORT_G_BRDF_COXMUNK) ? oracle::BRDF_COXMUNK
           : (bf & DISORT_G_BRDF_ROSSLI)  ? oracle::BRDF_ROSSLI
                                          : oracle::BRDF_NONE;
  c.lambertian = c.brdf == oracle::BRDF_NONE;
  c.corr_buras = (b->flags & DISORT_G_CORR_BURAS) != 0;
  c.corr_nakajima = (b->flags & DISORT_G_CORR_NAKAJIMA) != 0;
  c.index_from_bottom = index_from_bottom != 0;
  c.no_cutoff = (b->flags & DISORT_G_NO_CUTOFF) != 0;
  c.accur = b->accuracy;
  const double* s = b->scal + i * DISORT_NSCAL;
  c.beam_flux = s[DISORT_S_BEAM_FLUX];
  c.beam_mu = s[DISORT_S_BEAM_MU];
  c.beam_phi = s[DISORT_S_BEAM_PHI];
  c.iso_top = s[DISORT_S_ISO_TOP];
  c.iso_bot = s[DISORT_S_ISO_BOT];
  c.temp_top = s[DISORT_S_TEMP_TOP];
  c.temp_bot = s[DISORT_S_TEMP_BOT];
  c.emis_top = s[DISORT_S_EMIS_TOP];
  c.albedo = s[DISORT_S_ALBEDO];
  c.wn_lo = s[DISORT_S_WN_LO];
  c.wn_hi = s[DISORT_S_WN_HI];
  c.radius = s[DISORT_S_RADIUS];
  if (!c.lambertian && b->rpv)
    for (int k = 0; k < DISORT_NRPV; ++k) c.rpv[k] = b->rpv[i * DISORT_NRPV + k];
  c.dtau.assign(b->dtau + i * L, b->dtau + (i + 1) * L);
  c.ssa.assign(b->ssa + i * L, b->ssa + (i + 1) * L);
  c.pmom.assign(b->pmom + i * L * P, b->pmom + (i + 1) * L * P);
  if (c.thermal && b->temper) c.temper.assign(b->temper + i * (L + 1), b->temper + (i + 1) * (L + 1));
  if (c.spherical && b->alt) c.alt.assign(b->alt + i * (L + 1), b->alt + (i + 1) * (L + 1));
  if (c.use_user_tau && b->tau_user) c.tau_user.assign(b->tau_user + i * b->ntau, b->tau_user + (i + 1) * b->ntau);
  if (c.use_user_mu && b->mu_user) c.mu_user.assign(b->mu_user, b->mu_user + b->numu);
  if (b->phi_user) c.phi_user.assign(b->phi_user, b->phi_user + b->nphi);
  if (b->nphase > 0 && b->mu_phase && b->phase) {
    c.nphase = b->nphase;
    c.mu_phase.assign(b->mu_phase, b->mu_phase + b->nphase);
    c.phase.assign(b->phase + (size_t)i * L * b->nphase, b->phase + (size_t)(i + 1) * L * b->nphase);
  }
  return c;
}

/* One column of a general batch (DisortSolver::solve).  flux is
 * [8][ntau], inten [ntau][numu][nphi], inten_avg [ntau][numu]; *modes gets
 * the number of Fourier modes computed.  Returns 0, 1 (invalid argument) or
 * 2 (runtime error). */
int oracle_general_one(const disort_general_batch* b, int64_t i, int index_from_bottom, double* flux,
                       double* inten, double* inten_avg, int* modes, char* err, int errlen) {
  try {
    const oracle::GeneralConfig c = general_config(b, i, index_from_bottom);
    oracle::GeneralOutput o = oracle::general_solve(c);
    if (o.ntau != b->ntau || o.numu != b->numu || o.nphi != (b->flags & DISORT_G_ONLY_FLUXES ? 0 : b->nphi))
      throw std::invalid_argument("oracle: output shape differs from the batch's ntau/numu/nphi");
    for (int f = 0; f < 8; ++f) std::memcpy(flux + (size_t)f * o.ntau, o.f[f].data(), sizeof(double) * o.ntau);
    if (inten) std::memcpy(inten, o.inten.data(), sizeof(double) * o.inten.size());
    if (inten_avg) std::memcpy(inten_avg, o.inten_avg.data(), sizeof(double) * o.inten_avg.size());
    if (modes) *modes = o.modes;
    return 0;
  } catch (const std::invalid_argument& e) {
    put_err(err, errlen, e.what());
    return 1;
  } catch (const std::exception& e) {
    put_err(err, errlen, e.what());
    return 2;
  }
}

/* Column i of a general batch with temperature Jacobians: flux as
 * oracle_general_one, jac [6][ntau][nlyr+2] (fields flux_up, flux_down,
 * mean_intensity, mean_intensity_down, mean_intensity_up,
 * flux_tau_divergence).  Returns 0, 1 or 2 as oracle_general_one. */
int oracle_general_jacobian_one(const disort_general_batch* b, int64_t i, int index_from_bottom, double* flux,
                                double* jac, char* err, int errlen) {
  try {
    oracle::GeneralConfig c = general_config(b, i, index_from_bottom);
    c.jacobian = true;
    const oracle::GeneralOutput o = oracle::general_solve(c);
    for (int f = 0; f < 8; ++f) std::memcpy(flux + (size_t)f * o.ntau, o.f[f].data(), sizeof(double) * o.ntau);
    const size_t per = (size_t)o.ntau * (c.nlyr + 2);
    for (int f = 0; f < 6; ++f) {
      if (o.jac[f].size() != per) throw std::runtime_error("oracle: no Jacobian (thermal emission off)");
      std::memcpy(jac + f * per, o.jac[f].data(), sizeof(double) * per);
    }
    return 0;
  } catch (const std::invalid_argument& e) {
    put_err(err, errlen, e.what());
    return 1;
  } catch (const std::exception& e) {
    put_err(err, errlen, e.what());
    return 2;
  }
}

/* SPECIAL_BC of column i of a general batch (DisortSolver::solve with
 * ibcnd = SPECIAL_BC): albedo and transmissivity of the medium [2 h] (h =
 * positive user polar angles, or nstr/2 without DISORT_G_USER_MU; the first
 * h filled), spherical albedo and transmissivity sph[2].  Returns 0, 1 or 2
 * as oracle_general_one. */
int oracle_general_albtrans(const disort_general_batch* b, int64_t i, double* albedo, double* trans, double* sph,
                            char* err, int errlen) {
  try {
    oracle::GeneralConfig c = general_config(b, i, 0);
    c.special_bc = true;
    const oracle::GeneralOutput o = oracle::general_solve(c);
    std::memcpy(albedo, o.albedo_medium.data(), sizeof(double) * o.albedo_medium.size());
    std::memcpy(trans, o.transmissivity_medium.data(), sizeof(double) * o.transmissivity_medium.size());
    sph[0] = o.spherical_albedo;
    sph[1] = o.spherical_transmissivity;
    return 0;
  } catch (const std::invalid_argument& e) {
    put_err(err, errlen, e.what());
    return 1;
  } catch (const std::exception& e) {
    put_err(err, errlen, e.what());
    return 2;
  }
}

/* Columns [c0, c0+cnt) of a general batch with an OpenMP static loop, the
 * structure of solveSpectralBands (src/python_bindings.cpp:96-128); outputs
 * into the batch's host arrays.  Returns the number of failed columns;
 * *seconds receives the wall time of the loop. */
int64_t oracle_general_batch(const disort_general_batch* b, int64_t c0, int64_t cnt, int nthreads, double* seconds) {
  int64_t failed = 0;
  const auto t0 = std::chrono::steady_clock::now();
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  const int64_t per = (int64_t)b->ntau * b->numu;
#pragma omp parallel for schedule(static) reduction(+ : failed)
  for (int64_t i = c0; i < c0 + cnt; ++i) {
    std::vector<double> flux((size_t)DISORT_NOUT * b->ntau);
    char err[256];
    int modes = 0;
    const int rc = oracle_general_one(b, i, 0, flux.data(), b->inten ? b->inten + i * per * b->nphi : nullptr,
                                      b->inten_avg ? b->inten_avg + i * per : nullptr, &modes, err, 256);
    if (rc) {
      ++failed;
      continue;
    }
    for (int f = 0; f < DISORT_NOUT; ++f)
      for (int l = 0; l < b->ntau; ++l) b->flux[((int64_t)f * b->n + i) * b->ntau + l] = flux[(size_t)f * b->ntau + l];
  }
  const auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return failed;
}

double oracle_rpv(const double* rpv, double mu1, double mu2, double phi) {
  return oracle::rpv_reflectivity(rpv, mu1, mu2, phi);
}

}  // extern "C"

namespace oracle_ld {
void set_perturbation(int stage, double eps);
}

extern "C" {

/* Mutation hook of the pin tests (see oracle::set_perturbation): selects
 * one stage of both the FP64 and the extended-precision oracle; stage 0
 * restores the unperturbed oracle. */
void oracle_set_perturbation(int stage, double eps) {
  oracle::set_perturbation(stage, eps);
  oracle_ld::set_perturbation(stage, eps);
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
