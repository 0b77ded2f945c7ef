// Extended-precision (x87 long double) twin of oracle_flux_one, used by tests
// as a higher-precision truth when FP64 implementations disagree.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <cstring>
#include <exception>
#include <stdexcept>

#include "../include/disort_amd.h"
#include "oracle.hpp"

namespace {
int solve_ld(const disort_flux_batch* b, int64_t i, double* out, double* jac, char* err, int errlen) {
  try {
    oracle_ld::FluxConfig c;
    c.jacobian = jac != nullptr;
    c.nstr = b->nstr[i];
    c.nlyr = b->nlyr[i];
    c.nmom = b->pstride[i] - 1;
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
    const int64_t lo = b->lay_off[i], vo = lo + i;
    c.dtau.assign(b->dtau + lo, b->dtau + lo + c.nlyr);
    c.ssa.assign(b->ssa + lo, b->ssa + lo + c.nlyr);
    c.pmom.assign(b->pmom + b->mom_off[i], b->pmom + b->mom_off[i] + (size_t)c.nlyr * b->pstride[i]);
    if (c.thermal && b->temper) c.temper.assign(b->temper + vo, b->temper + vo + c.nlyr + 1);
    if (c.spherical && b->alt) c.alt.assign(b->alt + vo, b->alt + vo + c.nlyr + 1);
    oracle_ld::FluxOutput o = oracle_ld::flux_solve(c);
    const int nl = c.nlyr + 1;
    for (int f = 0; f < 8; ++f)
      for (int l = 0; l < nl; ++l) out[(size_t)f * nl + l] = (double)o.f[f][l];
    if (jac) {
      const size_t per = (size_t)nl * (nl + 1);
      for (int f = 0; f < 6; ++f) {
        if (o.jac[f].size() != per) throw std::runtime_error("oracle: no Jacobian (thermal emission off)");
        for (size_t q = 0; q < per; ++q) jac[f * per + q] = (double)o.jac[f][q];
      }
    }
    return 0;
  } catch (const std::exception& e) {
    if (err && errlen > 0) {
      std::strncpy(err, e.what(), errlen - 1);
      err[errlen - 1] = 0;
    }
    return 1;
  }
}
}  // namespace

extern "C" int oracle_ld_flux_one(const disort_flux_batch* b, int64_t i, double* out, char* err, int errlen) {
  return solve_ld(b, i, out, nullptr, err, errlen);
}

extern "C" int oracle_ld_flux_jacobian_one(const disort_flux_batch* b, int64_t i, double* out, double* jac, char* err,
                                           int errlen) {
  return solve_ld(b, i, out, jac, err, errlen);
}
