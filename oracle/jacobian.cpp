// Oracle restatement of the analytic temperature Jacobians shared by the
// flux-only and the general solver (DisortFluxSolver::computeThermalJacobian,
// src/FluxSolver.cpp:786-831 and :1294-1497; DisortSolver::
// computeThermalJacobian, src/DisortSolver.cpp:2130-2377): the eigen-system
// and the boundary-value matrix do not depend on temperature, so the
// derivative of the integration constants with respect to a level Planck
// value is one more back-substitution on the LU of the forward solve.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <cmath>

#include "oracle.hpp"

namespace ORACLE_NS {

void thermal_source_derivs(const DenseLU& lu, int N, Real omm, Real dtau_s, Real tau_top, const Real* cmu, Real* dz0,
                           Real* dz1, Real* dslope, Real* dicpt) {
  // src/FluxSolver.cpp:786-830: the slope s1 = (B_bot - B_top) / dtau (0 for
  // dtau <= 1e-4) and intercept s0 = B_top - s1 tau_top of the layer's
  // linear-in-tau Planck source, differentiated by B_top (e = 0) and B_bot
  // (e = 1); [I - C] dz1 = (1 - w) ds1, [I - C] dz0 = (1 - w) ds0 + mu dz1.
  const int NN = N / 2;
  const Real ds1[2] = {dtau_s > 1e-4 ? -1.0 / dtau_s : 0.0, dtau_s > 1e-4 ? 1.0 / dtau_s : 0.0};
  const Real ds0[2] = {1.0 - ds1[0] * tau_top, -ds1[1] * tau_top};
  std::vector<Real> b0(N), b1(N);
  for (int e = 0; e < 2; ++e) {
    dslope[e] = ds1[e];
    dicpt[e] = ds0[e];
    for (int i = 0; i < N; ++i) b1[i] = omm * ds1[e];
    lu.solve(b1.data());
    for (int i = 0; i < N; ++i) b0[i] = omm * ds0[e] + cmu[i] * b1[i];
    lu.solve(b0.data());
    Real* d0 = dz0 + (size_t)e * N;
    Real* d1 = dz1 + (size_t)e * N;
    for (int iq = 0; iq < NN; ++iq) {
      d0[NN + iq] = b0[iq];
      d0[NN - iq - 1] = b0[iq + NN];
      d1[NN + iq] = b1[iq];
      d1[NN - iq - 1] = b1[iq + NN];
    }
  }
}

void thermal_jacobian(const ThermalJacobianInputs& in, std::vector<Real> jac[6]) {
  const int N = in.N, NN = N / 2, L = in.L, ncut = in.ncut;
  const int ncd = 3 * NN - 1, lda = 3 * ncd + 1, nc = N * ncut, lb = ncut - 1, nd = L + 2;
  for (int f = 0; f < 6; ++f) jac[f].assign((size_t)in.nrows * nd, 0.0);
  // d z(lc) / d B_k: nonzero for the layer's top (k = lc) and bottom (k = lc + 1) level.
  auto dz = [&](const Real* d, int lc, int iq, int k) -> Real {
    Real r = 0.0;
    if (k == lc) r += d[(size_t)(2 * lc) * N + iq];
    if (k == lc + 1) r += d[(size_t)(2 * lc + 1) * N + iq];
    return r;
  };
  auto dlin = [&](const Real* d, int lc, int k) -> Real {
    Real r = 0.0;
    if (k == lc) r += d[2 * lc];
    if (k == lc + 1) r += d[2 * lc + 1];
    return r;
  };
  auto G = [&](int lc, int r, int col) { return in.gc[(size_t)lc * N * N + r + (size_t)col * N]; };
  const Real tb = in.taus[ncut];
  const Real dtl = in.dtaus[L - 1];
  std::vector<Real> b(nc), ec(N), dv(N);
  for (int k = 0; k <= L + 1; ++k) {
    const bool surface = k == L + 1;
    std::fill(b.begin(), b.end(), 0.0);
    if (!surface) {
      // Top boundary, bottom boundary, layer interfaces: the forward
      // right-hand side's thermal terms differentiated by B_k.
      for (int iq = 0; iq < NN; ++iq) b[iq] = -dz(in.dz0, 0, NN - 1 - iq, k);
      for (int iq = 0; iq < NN; ++iq) {
        Real s = -dz(in.dz0, lb, NN + iq, k) - dz(in.dz1, lb, NN + iq, k) * tb;
        if (in.diffusion_bc) {
          // d/dB_k of mu (B_bottom - B_{L-1}) / dtau_{L-1}
          if (k == L - 1 && dtl > 0.0) s -= in.cmu[iq] / dtl;
        } else if (!in.lyrcut) {
          Real rs = 0.0;
          for (int jq = 0; jq < NN; ++jq)
            rs += in.cwt[jq] * in.cmu[jq] * in.refl[(size_t)iq * NN + jq] *
                  (dz(in.dz0, lb, NN - 1 - jq, k) + dz(in.dz1, lb, NN - 1 - jq, k) * tb);
          s += 2.0 * rs;
        }
        b[nc - NN + iq] = s;
      }
      for (int lc = 1; lc < ncut; ++lc)
        for (int iq = 0; iq < N; ++iq)
          b[lc * N - NN + iq] = dz(in.dz0, lc, iq, k) - dz(in.dz0, lc - 1, iq, k) +
                                (dz(in.dz1, lc, iq, k) - dz(in.dz1, lc - 1, iq, k)) * in.taus[lc];
    } else {
      // Surface emission: emissivity B_bottom, or with the diffusion lower
      // boundary B_bottom + mu (B_bottom - B_{L-1}) / dtau_{L-1}.
      for (int iq = 0; iq < NN; ++iq)
        b[nc - NN + iq] = in.diffusion_bc ? 1.0 + (dtl > 0.0 ? in.cmu[iq] / dtl : 0.0) : in.emis[iq];
    }
    band_solve(in.band, lda, nc, ncd, ncd, in.ipvt, b.data());
    const Real chain = planck_band_deriv(in.wn_lo, in.wn_hi, surface ? in.temp_bot : in.temper[k]);
    for (int lu = 0; lu < in.nrows; ++lu) {
      const int lyu = in.row_layer[lu];
      if (in.lyrcut && lyu > ncut - 1) continue;
      const Real ut = in.row_tau[lu];
      for (int jq = 0; jq < NN; ++jq)
        ec[jq] = b[lyu * N + jq] * std::exp(-in.kk[(size_t)lyu * N + jq] * (ut - in.taus[lyu + 1]));
      for (int jq = NN; jq < N; ++jq)
        ec[jq] = b[lyu * N + jq] * std::exp(-in.kk[(size_t)lyu * N + jq] * (ut - in.taus[lyu]));
      for (int i = 0; i < N; ++i) {
        Real s = 0.0;
        for (int j = 0; j < N; ++j) s += G(lyu, i, j) * ec[j];
        if (!surface) s += dz(in.dz0, lyu, i, k) + dz(in.dz1, lyu, i, k) * ut;
        dv[i] = s;
      }
      Real mi = 0.0, mid = 0.0, miu = 0.0, fdn = 0.0, fup = 0.0;
      for (int iq = 0; iq < NN; ++iq) {
        const int ci = NN - iq - 1;
        mi += in.cwt[ci] * dv[iq];
        mid += in.cwt[ci] * dv[iq];
        fdn += in.cwt[ci] * dv[iq] * in.cmu[ci];
      }
      for (int iq = NN; iq < N; ++iq) {
        const int ci = iq - NN;
        mi += in.cwt[ci] * dv[iq];
        miu += in.cwt[ci] * dv[iq];
        fup += in.cwt[ci] * dv[iq] * in.cmu[ci];
      }
      const Real dpls = surface ? 0.0 : dlin(in.dicpt, lyu, k) + dlin(in.dslope, lyu, k) * ut;
      const Real dmi = 2.0 * M_PI * mi / (4.0 * M_PI);
      const Real vals[6] = {2.0 * M_PI * fup,
                            2.0 * M_PI * fdn,
                            dmi,
                            2.0 * M_PI * mid / (4.0 * M_PI),
                            2.0 * M_PI * miu / (4.0 * M_PI),
                            (1.0 - in.ssa[lyu]) * 4.0 * M_PI * (dmi - dpls)};
      for (int f = 0; f < 6; ++f) jac[f][(size_t)lu * nd + k] = chain * vals[f];
    }
  }
  for (int f = 0; f < 6; ++f) perturb(in.perturb_stage, jac[f]);
}

void reverse_jacobian(std::vector<Real> jac[6], int nrows, int L) {
  // src/IndexConversion.hpp:57-68: rows reversed, level columns reversed,
  // the surface column kept last.
  const int nd = L + 2;
  for (int f = 0; f < 6; ++f) {
    if (jac[f].empty()) continue;
    std::vector<Real> r(jac[f].size());
    for (int lu = 0; lu < nrows; ++lu)
      for (int k = 0; k < nd; ++k) r[(size_t)(nrows - 1 - lu) * nd + (k <= L ? L - k : k)] = jac[f][(size_t)lu * nd + k];
    jac[f].swap(r);
  }
}

}  // namespace ORACLE_NS
