// gfx950 kernels of the batched flux-only DISORT solve for N <= 16 streams,
// with the whole per-lane working set in registers.
//
// One instance = one DisortFluxSolver<N>::solve (src/FluxSolver.cpp:23-171).
// The solve is split into four launches over a bucket of instances that share
// the stream count N = 2*NN:
//   k_column  (flux_common.hpp) one wavefront per instance: Delta-M /
//             Delta-M+ scaling, optical-depth prefix sums, absorption
//             cut-off, beam transmission, Planck function per level
//             (src/FluxSolver.cpp:309-452, :104-122; src/Planck.cpp:35-169).
//   k_layer   one lane per (instance, layer): the reduced NN x NN eigenproblem
//             (src/FluxSolver.cpp:459-588) solved in registers by Cholesky
//             symmetrisation and cyclic Jacobi, then the beam and thermal
//             particular solutions (:595-643, :745-831) expressed through
//             that eigen-decomposition and a Cholesky factor of -A_s instead
//             of N x N LU solves.
//   k_boundary one lane per panel row (16 or 32 lanes per instance): the
//             (N*L)-unknown boundary system of
//             setMatrix/solve0 (:838-1131) eliminated layer by layer by LU
//             with partial pivoting on 3NN-row panels (the NN equations
//             carried from above plus the N continuity equations of the next
//             interface), then block back substitution.
//   k_flux    (flux_common.hpp) a lane group per (instance, level), one lane
//             per quadrature angle pair: intensities at the quadrature
//             angles and the eight outputs (:1138-1287).
#include "flux_common.hpp"

namespace dsa {
namespace {

// ===========================================================================
// k_layer: eigen-system and particular solutions, one lane per layer.
// ===========================================================================
template <int NN>
struct LayerMath {
  static constexpr int N = 2 * NN;
  static constexpr int NT = NN * (NN + 1) / 2;
  double L[NT];       // Cholesky factor of -B_s (packed lower)
  double V[NN][NN];   // eigenvectors of S = L^T (-A_s) L (columns)
  double lam[NN];     // eigenvalues of S (= k^2)

  __device__ void lsolve(double* x) const {  // x <- L^{-1} x
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      double s = x[r];
#pragma unroll
      for (int m = 0; m < r; ++m) s -= L[tri(r, m)] * x[m];
      x[r] = s / L[tri(r, r)];
    }
  }
  __device__ void ltsolve(double* x) const {  // x <- L^{-T} x
#pragma unroll
    for (int r = NN - 1; r >= 0; --r) {
      double s = x[r];
#pragma unroll
      for (int m = r + 1; m < NN; ++m) s -= L[tri(m, r)] * x[m];
      x[r] = s / L[tri(r, r)];
    }
  }
  __device__ void lmul(double* x) const {  // x <- L x
#pragma unroll
    for (int r = NN - 1; r >= 0; --r) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m <= r; ++m) s += L[tri(r, m)] * x[m];
      x[r] = s;
    }
  }
  __device__ void ltmul(double* x) const {  // x <- L^T x
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      double s = 0.0;
#pragma unroll
      for (int m = r; m < NN; ++m) s += L[tri(m, r)] * x[m];
      x[r] = s;
    }
  }
  // x <- V diag(d) V^T x
  __device__ void vdv(double* x, const double* d) const {
    double y[NN];
#pragma unroll
    for (int j = 0; j < NN; ++j) {
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < NN; ++r) s += V[r][j] * x[r];
      y[j] = s * d[j];
    }
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < NN; ++j) s += V[r][j] * y[j];
      x[r] = s;
    }
  }
};

template <int NN>
__device__ void jacobi_eigen(double (&S)[NN * (NN + 1) / 2], double (&V)[NN][NN], double (&lam)[NN]) {
#pragma unroll
  for (int r = 0; r < NN; ++r)
#pragma unroll
    for (int c = 0; c < NN; ++c) V[r][c] = (r == c) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 16; ++sweep) {
    double off = 0.0, dia = 0.0;
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      dia += S[tri(r, r)] * S[tri(r, r)];
#pragma unroll
      for (int c = 0; c < r; ++c) off += S[tri(r, c)] * S[tri(r, c)];
    }
    const int done = (off == 0.0) || (off <= 1e-36 * dia);
    if (__all(done)) break;
#pragma unroll
    for (int p = 0; p < NN - 1; ++p) {
#pragma unroll
      for (int q = p + 1; q < NN; ++q) {
        const double apq = S[tri(q, p)];
        const double app = S[tri(p, p)], aqq = S[tri(q, q)];
        const double theta = (aqq - app) / (2.0 * apq);
        const double at = fabs(theta);
        // No overflow guard (for |theta| > 1e154, theta^2 = inf gives t = 0, a
        // rotation below rounding): a select between two divisions becomes a
        // branch inside the unrolled sweep and multiplies the register demand.
        double t = 1.0 / (at + sqrt(theta * theta + 1.0));
        t = (theta < 0.0) ? -t : t;
        t = (apq == 0.0) ? 0.0 : t;
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
        for (int k = 0; k < NN; ++k) {
          if (k == p || k == q) continue;
          const int ip = (k > p) ? tri(k, p) : tri(p, k);
          const int iq = (k > q) ? tri(k, q) : tri(q, k);
          const double akp = S[ip], akq = S[iq];
          S[ip] = c * akp - s * akq;
          S[iq] = s * akp + c * akq;
        }
        S[tri(p, p)] = app - t * apq;
        S[tri(q, q)] = aqq + t * apq;
        S[tri(q, p)] = 0.0;
#pragma unroll
        for (int k = 0; k < NN; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < NN; ++r) lam[r] = S[tri(r, r)];
}

template <int NN>
__global__ __launch_bounds__(64) void k_layer(FluxArgs a) {
  constexpr int N = 2 * NN;
  constexpr int NT = NN * (NN + 1) / 2;
  const int64_t t = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (t >= a.n_tasks) return;
  const int i = a.t_inst[t];
  const int lc = a.t_lc[t];
  const int64_t g = a.lay_off[i] + lc;
  const int64_t v = a.lay_off[i] + i + lc;  // level at the layer top
  const int P = a.pstride[i];
  const double* pm = a.pmom + a.mom_off[i] + (int64_t)lc * P;
  const double* sc = a.scal + (int64_t)i * DISORT_NSCAL;
  const int64_t TL = a.total_layers, TV = a.total_levels;
  const double* tb = a.tab;
  constexpr TabOff to(N);
  constexpr EigOff eo(N);

  const double ws = a.lay_d[LD_SSA_S * TL + g];
  const double f = a.lay_d[LD_F * TL + g];
  const double sig2 = a.lay_d[LD_SIG2 * TL + g];

  // Scaled phase coefficients chi'_l (src/FluxSolver.cpp:342-388).  They are
  // evaluated twice (phase matrix, beam source) rather than kept live across
  // the eigen-solve, to bound register pressure.
  auto scaled_coeffs = [&](double (&chi)[N]) {
#pragma unroll
    for (int l = 0; l < N; ++l) {
      const double pl = pm[l];
      double c;
      if (sig2 < 0.0) c = (2.0 * l + 1.0) * ws * pl;
      else if (sig2 == 0.0) c = (2.0 * l + 1.0) * ws * (pl - f) / (1.0 - f);
      else c = (2.0 * l + 1.0) * ws * (pl - f * exp(-(double)(l * l) / (2.0 * sig2))) / (1.0 - f);
      chi[l] = c;
    }
  };

  const double F0 = sc[DISORT_S_BEAM_FLUX];
  const double mu0 = sc[DISORT_S_BEAM_MU];
  const bool beam = F0 > 0.0;

  // Symmetric forms A_s = S_d (alpha+beta) S_d^-1 and -B_s (:494-512):
  // A_s(i,k) = sum_{l even} chi_l Q_l(i) Q_l(k) - delta_ik / mu_i, and the
  // odd-l sum for -B_s, with Q_l(i) = P_l(mu_i) sqrt(w_i / mu_i).
  double As[NT], nB[NT];
  {
  double chi[N];
  scaled_coeffs(chi);
#pragma unroll
  for (int r = 0; r < NN; ++r)
#pragma unroll
    for (int c = 0; c <= r; ++c) {
      double se = 0.0, so = 0.0;
#pragma unroll
      for (int l = 0; l < N; ++l) {
        const double q = chi[l] * tb[to.Q() + l * NN + r] * tb[to.Q() + l * NN + c];
        if (l % 2 == 0) se += q;
        else so += q;
      }
      const double dg = (r == c) ? 1.0 / tb[to.cmu() + r] : 0.0;
      As[tri(r, c)] = se - dg;
      nB[tri(r, c)] = dg - so;
    }
  }

  LayerMath<NN> m;
  // Cholesky of -B_s (positive definite for omega' <= 1).
  int ok = 1;
#pragma unroll
  for (int c = 0; c < NN; ++c) {
    double x = nB[tri(c, c)];
#pragma unroll
    for (int k = 0; k < c; ++k) x -= m.L[tri(c, k)] * m.L[tri(c, k)];
    ok &= (x > 0.0);
    const double lcc = sqrt(x > 0.0 ? x : 1.0);
    m.L[tri(c, c)] = lcc;
#pragma unroll
    for (int r = c + 1; r < NN; ++r) {
      double s = nB[tri(r, c)];
#pragma unroll
      for (int k = 0; k < c; ++k) s -= m.L[tri(r, k)] * m.L[tri(c, k)];
      m.L[tri(r, c)] = s / lcc;
    }
  }
  if (!ok) atomicOr(&a.inst_i[i * kInstI + II_STATUS], DISORT_ST_EIGEN);

  // S = L^T (-A_s) L, column by column.
  double S[NT];
#pragma unroll
  for (int c = 0; c < NN; ++c) {
    double tc[NN];
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      double s = 0.0;
#pragma unroll
      for (int k = c; k < NN; ++k) s -= ((r >= k) ? As[tri(r, k)] : As[tri(k, r)]) * m.L[tri(k, c)];
      tc[r] = s;
    }
#pragma unroll
    for (int r = 0; r <= c; ++r) {
      double s = 0.0;
#pragma unroll
      for (int k = r; k < NN; ++k) s += m.L[tri(k, r)] * tc[k];
      S[tri(c, r)] = s;
    }
  }
  // Park L in the (not yet written) particular-solution slots of this task
  // while the Jacobi sweep holds S and V in registers.
  static_assert(NT <= 3 * N, "L must fit in the ztop/zbot/zfl slots");
#pragma unroll
  for (int q = 0; q < NT; ++q) a.eig[eo.at(t, eo.ztop() + q)] = m.L[q];
  jacobi_eigen<NN>(S, m.V, m.lam);
#pragma unroll
  for (int q = 0; q < NT; ++q) m.L[q] = a.eig[eo.at(t, eo.ztop() + q)];

  // Eigenvalues k = sqrt|lambda| and the G+/- blocks (:560-587):
  // u = S_d^-1 L^-T w is the eigenvector of (alpha+beta)(alpha-beta);
  // (alpha-beta) u / k = -S_d^-1 L w / k.
  double kv[NN];
#pragma unroll
  for (int j = 0; j < NN; ++j) {
    kv[j] = sqrt(fabs(m.lam[j]));
    a.eig[eo.at(t, eo.k() + j)] = kv[j];
    double u[NN], y[NN];
#pragma unroll
    for (int r = 0; r < NN; ++r) u[r] = y[r] = m.V[r][j];
    m.ltsolve(u);
    m.lmul(y);
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      const double sdi = tb[to.sdi() + r];
      const double uu = sdi * u[r];
      const double tt = -sdi * y[r] / kv[j];
      a.eig[eo.at(t, eo.gp() + r + j * NN)] = 0.5 * (tt + uu);
      a.eig[eo.at(t, eo.gm() + r + j * NN)] = 0.5 * (tt - uu);
    }
  }

  // Particular solutions in the reference's angle order: index NN+r is the
  // upward node mu_r, index NN-1-r the downward node -mu_r.
  double zb[N], z0[N], z1[N];
#pragma unroll
  for (int r = 0; r < N; ++r) zb[r] = z0[r] = z1[r] = 0.0;

  // Beam source sums b+ + b- and b+ - b- over the positive nodes
  // (right-hand side of src/FluxSolver.cpp:628-632 split by parity of l).
  double bs[NN], bd[NN];
  {
    double chi[N];
    scaled_coeffs(chi);
    double p0[N];
    const double x = -mu0;
    p0[0] = 1.0;
    if (N > 1) p0[1] = x;
#pragma unroll
    for (int l = 2; l < N; ++l) p0[l] = ((2.0 * l - 1.0) * x * p0[l - 1] - (l - 1.0) * p0[l - 2]) / l;
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      double se = 0.0, so = 0.0;
#pragma unroll
      for (int l = 0; l < N; ++l) {
        const double term = chi[l] * tb[to.P() + l * NN + r] * p0[l];
        if (l % 2 == 0) se += term;
        else so += term;
      }
      bs[r] = F0 * se / (2.0 * kPi);
      bd[r] = F0 * so / (2.0 * kPi);
    }
  }

  if (beam) {
    // (I + diag(mu)/mu0 - C) x = b in sum/difference form: with s = x+ + x-,
    // d = x+ - x-, ((alpha+beta)(alpha-beta) - 1/mu0^2) d = (p + mu0 (alpha+beta) q)/mu0
    // and s = mu0((alpha-beta) d - q), p = -D^-1 (b+ + b-), q = -D^-1 (b+ - b-).
    double mub = mu0;
    const double mu0sq = mu0 * mu0;
#pragma unroll
    for (int j = 0; j < NN; ++j)
      if (fabs(1.0 - mu0sq * kv[j] * kv[j]) < 1000.0 * kEps) mub = 0.999 * mu0;  // dither (:604-618)
    double p[NN], q[NN], w[NN];
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      const double cm = tb[to.cmu() + r];
      p[r] = -bs[r] / cm;
      q[r] = -bd[r] / cm;
      w[r] = tb[to.sd() + r] * q[r];
    }
    // (alpha+beta) q = -S_d^-1 L^-T V Lambda V^T L^-1 S_d q
    m.lsolve(w);
    m.vdv(w, m.lam);
    m.ltsolve(w);
    double rr[NN];
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      const double apbq = -tb[to.sdi() + r] * w[r];
      rr[r] = (p[r] + mub * apbq) / mub;
      rr[r] *= tb[to.sd() + r];
    }
    // d = S_d^-1 L^-T V (Lambda - 1/mu0^2)^-1 V^T L^T S_d r
    double dinv[NN];
    const double im2 = 1.0 / (mub * mub);
#pragma unroll
    for (int j = 0; j < NN; ++j) dinv[j] = 1.0 / (m.lam[j] - im2);
    m.ltmul(rr);
    m.vdv(rr, dinv);
    m.ltsolve(rr);
    double d[NN], e[NN];
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      d[r] = tb[to.sdi() + r] * rr[r];
      e[r] = tb[to.sd() + r] * d[r];
    }
    // (alpha-beta) d = -S_d^-1 L L^T S_d d
    m.ltmul(e);
    m.lmul(e);
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      const double ambd = -tb[to.sdi() + r] * e[r];
      const double s = mub * (ambd - q[r]);
      zb[NN + r] = 0.5 * (s + d[r]);
      zb[NN - 1 - r] = 0.5 * (s - d[r]);
    }
  }

  const double taut = a.lev_d[VD_TAU_S * TV + v];
  const double taub = a.lev_d[VD_TAU_S * TV + v + 1];
  double xr0 = 0.0, xr1 = 0.0;
  if (a.flags & DISORT_F_THERMAL) {
    // (src/FluxSolver.cpp:755-776) B(tau) = xr0 + xr1 tau;
    // (I - C) z1 = (1-omega') xr1, (I - C) z0 = (1-omega') xr0 + mu z1.
    const double pt = a.lev_d[VD_PLANCK * TV + v], pb = a.lev_d[VD_PLANCK * TV + v + 1];
    const double dts = a.lay_d[LD_DTAU_S * TL + g];
    if (dts > 1e-4) xr1 = (pb - pt) / dts;
    xr0 = pt - xr1 * taut;
    // base = F^-1 (2*1) with F = I - X - Y = -D (alpha+beta)
    //      = S_d^-1 (-A_s)^-1 S_d D^-1 (2*1), through a Cholesky factor of
    // -A_s (conditioned like the reference's LU of I - C, about 1/(1-omega'),
    // unlike the eigen-decomposition of S whose small eigenvalues are only
    // known to eps * k_max^2).
    double base[NN];
    {
      double R[NT];
      {
        double chi[N];
        scaled_coeffs(chi);
#pragma unroll
        for (int r = 0; r < NN; ++r)
#pragma unroll
          for (int c = 0; c <= r; ++c) {
            double se = 0.0;
#pragma unroll
            for (int l = 0; l < N; l += 2) se += chi[l] * tb[to.Q() + l * NN + r] * tb[to.Q() + l * NN + c];
            R[tri(r, c)] = ((r == c) ? 1.0 / tb[to.cmu() + r] : 0.0) - se;
          }
      }
#pragma unroll
      for (int c = 0; c < NN; ++c) {
        double x = R[tri(c, c)];
#pragma unroll
        for (int k = 0; k < c; ++k) x -= R[tri(c, k)] * R[tri(c, k)];
        const double rcc = sqrt(x > 0.0 ? x : 1.0);
        R[tri(c, c)] = rcc;
#pragma unroll
        for (int r = c + 1; r < NN; ++r) {
          double s2 = R[tri(r, c)];
#pragma unroll
          for (int k = 0; k < c; ++k) s2 -= R[tri(r, k)] * R[tri(c, k)];
          R[tri(r, c)] = s2 / rcc;
        }
      }
#pragma unroll
      for (int r = 0; r < NN; ++r) {
        double s2 = tb[to.sd() + r] * (2.0 / tb[to.cmu() + r]);
#pragma unroll
        for (int k = 0; k < r; ++k) s2 -= R[tri(r, k)] * base[k];
        base[r] = s2 / R[tri(r, r)];
      }
#pragma unroll
      for (int r = NN - 1; r >= 0; --r) {
        double s2 = base[r];
#pragma unroll
        for (int k = r + 1; k < NN; ++k) s2 -= R[tri(k, r)] * base[k];
        base[r] = s2 / R[tri(r, r)];
      }
    }
#pragma unroll
    for (int r = 0; r < NN; ++r) base[r] *= tb[to.sdi() + r];
    const double c1 = (1.0 - ws) * (a.jbase ? 1.0 : xr1), c0 = (1.0 - ws) * xr0;
    double h[NN], dd[NN];
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      h[r] = 0.5 * c1 * base[r];  // z1 at +mu_r and -mu_r
      dd[r] = tb[to.sd() + r] * h[r];
    }
    // d0 = 2 S_d^-1 L^-T L^-1 S_d z1h  (solves H d0 = 2 D z1h, H = -D (alpha-beta))
    m.lsolve(dd);
    m.ltsolve(dd);
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      const double d0 = 2.0 * tb[to.sdi() + r] * dd[r];
      const double s0 = c0 * base[r];
      if (a.jbase) {
        double* jb = a.jbase + ((int64_t)i * a.jac_nlyr + lc) * 2 * NN;
        jb[r] = (1.0 - ws) * base[r];
        jb[NN + r] = d0;
      }
      z1[NN + r] = h[r];
      z1[NN - 1 - r] = h[r];
      z0[NN + r] = 0.5 * (s0 + d0);
      z0[NN - 1 - r] = 0.5 * (s0 - d0);
    }
  }

  const double btt = a.lay_d[LD_BT_TOP * TL + g], btb = a.lay_d[LD_BT_BOT * TL + g];
  const double fact = beam ? exp(-taub / mu0) : 0.0;
#pragma unroll
  for (int r = 0; r < N; ++r) {
    a.eig[eo.at(t, eo.ztop() + r)] = zb[r] * btt + z0[r] + z1[r] * taut;
    a.eig[eo.at(t, eo.zbot() + r)] = zb[r] * btb + z0[r] + z1[r] * taub;
    a.eig[eo.at(t, eo.zfl() + r)] = zb[r] * fact + z0[r] + z1[r] * taub;
  }
  a.eig[eo.at(t, eo.pls())] = xr0 + xr1 * taub;
}

// ===========================================================================
// k_layer2: the same eigen-system and particular solutions with two lanes
// per layer (a lane pair, lanes 2p and 2p+1 of the wavefront).  Both lanes
// hold the symmetric matrices (S, the Cholesky factors) and perform the same
// rotations on them, so they agree bit for bit; the eigenvector matrix V is
// split by rows (lane h owns rows h*RH .. h*RH+RH-1), which halves the
// register demand of the Jacobi sweep (S + half of V) so that two or three
// wavefronts fit per SIMD without spilling.  Products with V combine the
// two halves by a DPP exchange within the pair (quad_perm 1,0,3,2); IEEE
// addition is commutative, so both lanes form identical sums.  The
// Cholesky factor of -B_s is parked in LDS (one copy per pair) during the
// sweep.  Same formulas and operation order per element as k_layer except
// the V^T x products, which are summed in two halves.
// ===========================================================================
__device__ __forceinline__ double pair_swap(double x) { return dpp_f64<0xB1>(x); }  // quad_perm(1,0,3,2)

template <int NN>
__global__ __launch_bounds__(64, (NN <= 4 ? 3 : 2)) void k_layer2(FluxArgs a) {
  constexpr int N = 2 * NN;
  constexpr int NT = NN * (NN + 1) / 2;
  constexpr int LS = NT | 1;        // LDS pitch of a parked factor (odd: fewer bank conflicts)
  constexpr int RH = (NN + 1) / 2;  // V rows per lane (lane 1 pads one row when NN is odd)
  __shared__ double Lsh[32 * LS];   // Cholesky factor of -B_s, one per pair, for the whole kernel
  const int h = threadIdx.x & 1;
  const int64_t t = ((int64_t)blockIdx.x * 64 + threadIdx.x) >> 1;
  if (t >= a.n_tasks) return;  // pair-uniform: both lanes of a pair leave together
  double* Lp = Lsh + (threadIdx.x >> 1) * LS;
  const int i = a.t_inst[t];
  const int lc = a.t_lc[t];
  const int64_t g = a.lay_off[i] + lc;
  const int64_t v = a.lay_off[i] + i + lc;
  const int P = a.pstride[i];
  const double* pm = a.pmom + a.mom_off[i] + (int64_t)lc * P;
  const double* sc = a.scal + (int64_t)i * DISORT_NSCAL;
  const int64_t TL = a.total_layers, TV = a.total_levels;
  const double* tb = a.tab;
  constexpr TabOff to(N);
  constexpr EigOff eo(N);

  const double ws = a.lay_d[LD_SSA_S * TL + g];
  const double f = a.lay_d[LD_F * TL + g];
  const double sig2 = a.lay_d[LD_SIG2 * TL + g];
  auto scaled_coeffs = [&](double (&chi)[N]) {
#pragma unroll
    for (int l = 0; l < N; ++l) {
      const double pl_ = pm[l];
      double c;
      if (sig2 < 0.0) c = (2.0 * l + 1.0) * ws * pl_;
      else if (sig2 == 0.0) c = (2.0 * l + 1.0) * ws * (pl_ - f) / (1.0 - f);
      else c = (2.0 * l + 1.0) * ws * (pl_ - f * exp(-(double)(l * l) / (2.0 * sig2))) / (1.0 - f);
      chi[l] = c;
    }
  };
  // One scaled coefficient chi'_l (the same expression as scaled_coeffs).
  auto chi_of = [&](int l) -> double {
    const double pl_ = pm[l];
    if (sig2 < 0.0) return (2.0 * l + 1.0) * ws * pl_;
    if (sig2 == 0.0) return (2.0 * l + 1.0) * ws * (pl_ - f) / (1.0 - f);
    return (2.0 * l + 1.0) * ws * (pl_ - f * exp(-(double)(l * l) / (2.0 * sig2))) / (1.0 - f);
  };
  const double F0 = sc[DISORT_S_BEAM_FLUX];
  const double mu0 = sc[DISORT_S_BEAM_MU];
  const bool beam = F0 > 0.0;

  // Full vector x from the two halves (own rows in `mine`).
  auto join = [&](const double (&mine)[RH], double (&x)[NN]) {
#pragma unroll
    for (int q = 0; q < RH; ++q) {
      const double other = pair_swap(mine[q]);
      x[q] = h ? other : mine[q];
      if (RH + q < NN) x[RH + q] = h ? mine[q] : other;
    }
  };
  // x of the rows this lane owns.
  auto own = [&](const double (&x)[NN], int q) -> double {
    return (RH + q < NN) ? (h ? x[RH + q] : x[q]) : (h ? 0.0 : x[q]);
  };
  // Triangular products and solves with the parked factor L (LayerMath's
  // operations, L read from LDS).  Each starts with a compiler memory
  // barrier so that the factor is re-read from LDS rather than kept live in
  // registers between uses (the LDS reads are cheap, the registers are not).
  auto reload = []() { asm volatile("" ::: "memory"); };
  auto Lx = [&](int r, int m) -> double { return Lp[tri(r, m)]; };
  auto lsolve = [&](double (&x)[NN]) {
    reload();
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      double s2 = x[r];
#pragma unroll
      for (int m = 0; m < r; ++m) s2 -= Lx(r, m) * x[m];
      x[r] = s2 / Lx(r, r);
    }
  };
  auto ltsolve = [&](double (&x)[NN]) {
    reload();
#pragma unroll
    for (int r = NN - 1; r >= 0; --r) {
      double s2 = x[r];
#pragma unroll
      for (int m = r + 1; m < NN; ++m) s2 -= Lx(m, r) * x[m];
      x[r] = s2 / Lx(r, r);
    }
  };
  auto lmul = [&](double (&x)[NN]) {
    reload();
#pragma unroll
    for (int r = NN - 1; r >= 0; --r) {
      double s2 = 0.0;
#pragma unroll
      for (int m = 0; m <= r; ++m) s2 += Lx(r, m) * x[m];
      x[r] = s2;
    }
  };
  auto ltmul = [&](double (&x)[NN]) {
    reload();
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      double s2 = 0.0;
#pragma unroll
      for (int m = r; m < NN; ++m) s2 += Lx(m, r) * x[m];
      x[r] = s2;
    }
  };

  // -B_s, then its Cholesky factor in place; parked in LDS (one copy per
  // pair: lane h stores the entries q = h mod 2).
  {
    double nB[NT];
    {
      double chi[N];
      scaled_coeffs(chi);
      // Odd-l sums accumulated one l at a time (the per-element order of
      // the additions is unchanged), so only one table row is live.
#pragma unroll
      for (int q = 0; q < NT; ++q) nB[q] = 0.0;
#pragma unroll
      for (int l = 1; l < N; l += 2) {
        double ql[NN];
#pragma unroll
        for (int r = 0; r < NN; ++r) ql[r] = tb[to.Q() + l * NN + r];
#pragma unroll
        for (int r = 0; r < NN; ++r)
#pragma unroll
          for (int c = 0; c <= r; ++c) nB[tri(r, c)] += chi[l] * ql[r] * ql[c];
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int r = 0; r < NN; ++r)
#pragma unroll
        for (int c = 0; c <= r; ++c) {
          const double dg = (r == c) ? 1.0 / tb[to.cmu() + r] : 0.0;
          nB[tri(r, c)] = dg - nB[tri(r, c)];
        }
    }
    int ok = 1;
#pragma unroll
    for (int c = 0; c < NN; ++c) {
      double x = nB[tri(c, c)];
#pragma unroll
      for (int k = 0; k < c; ++k) x -= nB[tri(c, k)] * nB[tri(c, k)];
      ok &= (x > 0.0);
      const double lcc = sqrt(x > 0.0 ? x : 1.0);
      nB[tri(c, c)] = lcc;
#pragma unroll
      for (int r = c + 1; r < NN; ++r) {
        double s2 = nB[tri(r, c)];
#pragma unroll
        for (int k = 0; k < c; ++k) s2 -= nB[tri(r, k)] * nB[tri(c, k)];
        nB[tri(r, c)] = s2 / lcc;
      }
    }
    if (!ok && h == 0) atomicOr(&a.inst_i[i * kInstI + II_STATUS], DISORT_ST_EIGEN);
#pragma unroll
    for (int q = 0; q < NT; ++q)
      if ((q & 1) == h) Lp[q] = nB[q];
  }
  wave_lds_order();

  // S = L^T (-A_s) L, column by column.
  double S[NT];
  {
    double As[NT];
    {
      double chi[N];
      scaled_coeffs(chi);
#pragma unroll
      for (int q = 0; q < NT; ++q) As[q] = 0.0;
#pragma unroll
      for (int l = 0; l < N; l += 2) {
        double ql[NN];
#pragma unroll
        for (int r = 0; r < NN; ++r) ql[r] = tb[to.Q() + l * NN + r];
#pragma unroll
        for (int r = 0; r < NN; ++r)
#pragma unroll
          for (int c = 0; c <= r; ++c) As[tri(r, c)] += chi[l] * ql[r] * ql[c];
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int r = 0; r < NN; ++r)
#pragma unroll
        for (int c = 0; c <= r; ++c) {
          const double dg = (r == c) ? 1.0 / tb[to.cmu() + r] : 0.0;
          As[tri(r, c)] = As[tri(r, c)] - dg;
        }
    }
#pragma unroll
    for (int c = 0; c < NN; ++c) {
      double tc[NN];
#pragma unroll
      for (int r = 0; r < NN; ++r) {
        double s2 = 0.0;
#pragma unroll
        for (int k = c; k < NN; ++k) s2 -= ((r >= k) ? As[tri(r, k)] : As[tri(k, r)]) * Lx(k, c);
        tc[r] = s2;
      }
#pragma unroll
      for (int r = 0; r <= c; ++r) {
        double s2 = 0.0;
#pragma unroll
        for (int k = r; k < NN; ++k) s2 += Lx(k, r) * tc[k];
        S[tri(c, r)] = s2;
      }
    }
  }

  // Cyclic Jacobi: S in both lanes, V rows split.
  double Vh[RH][NN];
#pragma unroll
  for (int q = 0; q < RH; ++q)
#pragma unroll
    for (int c = 0; c < NN; ++c) Vh[q][c] = (h * RH + q == c) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 16; ++sweep) {
    double off = 0.0, dia = 0.0;
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      dia += S[tri(r, r)] * S[tri(r, r)];
#pragma unroll
      for (int c = 0; c < r; ++c) off += S[tri(r, c)] * S[tri(r, c)];
    }
    const int done = (off == 0.0) || (off <= 1e-36 * dia);
    if (__all(done)) break;
#pragma unroll
    for (int p = 0; p < NN - 1; ++p) {
#pragma unroll
      for (int q = p + 1; q < NN; ++q) {
        const double apq = S[tri(q, p)];
        const double app = S[tri(p, p)], aqq = S[tri(q, q)];
        const double theta = (aqq - app) / (2.0 * apq);
        const double at = fabs(theta);
        double tt = 1.0 / (at + sqrt(theta * theta + 1.0));  // as in jacobi_eigen
        tt = (theta < 0.0) ? -tt : tt;
        tt = (apq == 0.0) ? 0.0 : tt;
        const double c = 1.0 / sqrt(tt * tt + 1.0), sn = tt * c;
#pragma unroll
        for (int k = 0; k < NN; ++k) {
          if (k == p || k == q) continue;
          const int ip = (k > p) ? tri(k, p) : tri(p, k);
          const int iq = (k > q) ? tri(k, q) : tri(q, k);
          const double akp = S[ip], akq = S[iq];
          S[ip] = c * akp - sn * akq;
          S[iq] = sn * akp + c * akq;
        }
        S[tri(p, p)] = app - tt * apq;
        S[tri(q, q)] = aqq + tt * apq;
        S[tri(q, p)] = 0.0;
#pragma unroll
        for (int k = 0; k < RH; ++k) {
          const double vkp = Vh[k][p], vkq = Vh[k][q];
          Vh[k][p] = c * vkp - sn * vkq;
          Vh[k][q] = sn * vkp + c * vkq;
        }
      }
    }
  }
  double lam[NN];
#pragma unroll
  for (int r = 0; r < NN; ++r) lam[r] = S[tri(r, r)];

  // x <- V diag(d) V^T x with V split by rows.
  auto vdv = [&](double (&x)[NN], const double (&d)[NN]) {
    double y[NN];
#pragma unroll
    for (int j = 0; j < NN; ++j) {
      double s2 = 0.0;
#pragma unroll
      for (int q = 0; q < RH; ++q) s2 += Vh[q][j] * own(x, q);
      y[j] = (s2 + pair_swap(s2)) * d[j];
    }
    double mine[RH];
#pragma unroll
    for (int q = 0; q < RH; ++q) {
      double s2 = 0.0;
#pragma unroll
      for (int j = 0; j < NN; ++j) s2 += Vh[q][j] * y[j];
      mine[q] = s2;
    }
    join(mine, x);
  };

  // k and G+/- : lane h takes columns j = 2 jj + h.
  double kv[NN];
#pragma unroll
  for (int j = 0; j < NN; ++j) kv[j] = sqrt(fabs(lam[j]));
#pragma unroll
  for (int jj = 0; jj < (NN + 1) / 2; ++jj) {
    const int j0 = 2 * jj, j1 = (2 * jj + 1 < NN) ? 2 * jj + 1 : 2 * jj;
    // Full column j: our rows of it, and the partner's rows, received in
    // exchange for our rows of the partner's column.
    double col[NN];
#pragma unroll
    for (int q = 0; q < RH; ++q) {
      const double keep = h ? Vh[q][j1] : Vh[q][j0];
      const double other = pair_swap(h ? Vh[q][j0] : Vh[q][j1]);
      col[q] = h ? other : keep;
      if (RH + q < NN) col[RH + q] = h ? keep : other;
    }
    const int j = h ? j1 : j0;
    const bool write = (2 * jj + h) < NN;
    const double kj = h ? kv[j1] : kv[j0];
    double u[NN], y[NN];
#pragma unroll
    for (int r = 0; r < NN; ++r) u[r] = y[r] = col[r];
    ltsolve(u);
    lmul(y);
    if (write) {
      a.eig[eo.at(t, eo.k() + j)] = kj;
#pragma unroll
      for (int r = 0; r < NN; ++r) {
        const double sdi = tb[to.sdi() + r];
        const double uu = sdi * u[r];
        const double tt = -sdi * y[r] / kj;
        a.eig[eo.at(t, eo.gp() + r + j * NN)] = 0.5 * (tt + uu);
        a.eig[eo.at(t, eo.gm() + r + j * NN)] = 0.5 * (tt - uu);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  }

  double zb[N], z0[N], z1[N];
#pragma unroll
  for (int r = 0; r < N; ++r) zb[r] = z0[r] = z1[r] = 0.0;
  double bs[NN], bd[NN];
  {
    // Sums over l accumulated one l at a time (same per-element order),
    // with P_l(-mu0) by its recurrence.
    const double x = -mu0;
    double pm2 = 0.0, pm1 = 0.0;
#pragma unroll
    for (int r = 0; r < NN; ++r) bs[r] = bd[r] = 0.0;
#pragma unroll
    for (int l = 0; l < N; ++l) {
      const double p0l = (l == 0) ? 1.0 : (l == 1) ? x : ((2.0 * l - 1.0) * x * pm1 - (l - 1.0) * pm2) / l;
      pm2 = pm1;
      pm1 = p0l;
      const double cl = chi_of(l);
#pragma unroll
      for (int r = 0; r < NN; ++r) {
        const double term = cl * tb[to.P() + l * NN + r] * p0l;
        if (l % 2 == 0) bs[r] += term;
        else bd[r] += term;
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      bs[r] = F0 * bs[r] / (2.0 * kPi);
      bd[r] = F0 * bd[r] / (2.0 * kPi);
    }
  }
  if (beam) {
    double mub = mu0;
    const double mu0sq = mu0 * mu0;
#pragma unroll
    for (int j = 0; j < NN; ++j)
      if (fabs(1.0 - mu0sq * kv[j] * kv[j]) < 1000.0 * kEps) mub = 0.999 * mu0;
    double p[NN], q[NN], w[NN];
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      const double cm = tb[to.cmu() + r];
      p[r] = -bs[r] / cm;
      q[r] = -bd[r] / cm;
      w[r] = tb[to.sd() + r] * q[r];
    }
    lsolve(w);
    vdv(w, lam);
    ltsolve(w);
    double rr[NN];
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      const double apbq = -tb[to.sdi() + r] * w[r];
      rr[r] = (p[r] + mub * apbq) / mub;
      rr[r] *= tb[to.sd() + r];
    }
    double dinv[NN];
    const double im2 = 1.0 / (mub * mub);
#pragma unroll
    for (int j = 0; j < NN; ++j) dinv[j] = 1.0 / (lam[j] - im2);
    ltmul(rr);
    vdv(rr, dinv);
    ltsolve(rr);
    double d[NN], e[NN];
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      d[r] = tb[to.sdi() + r] * rr[r];
      e[r] = tb[to.sd() + r] * d[r];
    }
    ltmul(e);
    lmul(e);
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      const double ambd = -tb[to.sdi() + r] * e[r];
      const double s2 = mub * (ambd - q[r]);
      zb[NN + r] = 0.5 * (s2 + d[r]);
      zb[NN - 1 - r] = 0.5 * (s2 - d[r]);
    }
  }

  const double taut = a.lev_d[VD_TAU_S * TV + v];
  const double taub = a.lev_d[VD_TAU_S * TV + v + 1];
  double xr0 = 0.0, xr1 = 0.0;
  if (a.flags & DISORT_F_THERMAL) {
    const double pt = a.lev_d[VD_PLANCK * TV + v], pb = a.lev_d[VD_PLANCK * TV + v + 1];
    const double dts = a.lay_d[LD_DTAU_S * TL + g];
    if (dts > 1e-4) xr1 = (pb - pt) / dts;
    xr0 = pt - xr1 * taut;
    double base[NN];
    {
      double R[NT];
      {
#pragma unroll
        for (int q = 0; q < NT; ++q) R[q] = 0.0;
#pragma unroll
        for (int l = 0; l < N; l += 2) {
          const double cl = chi_of(l);
          double ql[NN];
#pragma unroll
          for (int r = 0; r < NN; ++r) ql[r] = tb[to.Q() + l * NN + r];
#pragma unroll
          for (int r = 0; r < NN; ++r)
#pragma unroll
            for (int c = 0; c <= r; ++c) R[tri(r, c)] += cl * ql[r] * ql[c];
          __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int r = 0; r < NN; ++r)
#pragma unroll
          for (int c = 0; c <= r; ++c) R[tri(r, c)] = ((r == c) ? 1.0 / tb[to.cmu() + r] : 0.0) - R[tri(r, c)];
      }
#pragma unroll
      for (int c = 0; c < NN; ++c) {
        double x = R[tri(c, c)];
#pragma unroll
        for (int k = 0; k < c; ++k) x -= R[tri(c, k)] * R[tri(c, k)];
        const double rcc = sqrt(x > 0.0 ? x : 1.0);
        R[tri(c, c)] = rcc;
#pragma unroll
        for (int r = c + 1; r < NN; ++r) {
          double s2 = R[tri(r, c)];
#pragma unroll
          for (int k = 0; k < c; ++k) s2 -= R[tri(r, k)] * R[tri(c, k)];
          R[tri(r, c)] = s2 / rcc;
        }
      }
#pragma unroll
      for (int r = 0; r < NN; ++r) {
        double s2 = tb[to.sd() + r] * (2.0 / tb[to.cmu() + r]);
#pragma unroll
        for (int k = 0; k < r; ++k) s2 -= R[tri(r, k)] * base[k];
        base[r] = s2 / R[tri(r, r)];
      }
#pragma unroll
      for (int r = NN - 1; r >= 0; --r) {
        double s2 = base[r];
#pragma unroll
        for (int k = r + 1; k < NN; ++k) s2 -= R[tri(k, r)] * base[k];
        base[r] = s2 / R[tri(r, r)];
      }
    }
#pragma unroll
    for (int r = 0; r < NN; ++r) base[r] *= tb[to.sdi() + r];
    const double c1 = (1.0 - ws) * (a.jbase ? 1.0 : xr1), c0 = (1.0 - ws) * xr0;
    double hh[NN], dd[NN];
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      hh[r] = 0.5 * c1 * base[r];
      dd[r] = tb[to.sd() + r] * hh[r];
    }
    lsolve(dd);
    ltsolve(dd);
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      const double d0 = 2.0 * tb[to.sdi() + r] * dd[r];
      const double s0 = c0 * base[r];
      if (a.jbase) {
        double* jb = a.jbase + ((int64_t)i * a.jac_nlyr + lc) * 2 * NN;
        jb[r] = (1.0 - ws) * base[r];
        jb[NN + r] = d0;
      }
      z1[NN + r] = hh[r];
      z1[NN - 1 - r] = hh[r];
      z0[NN + r] = 0.5 * (s0 + d0);
      z0[NN - 1 - r] = 0.5 * (s0 - d0);
    }
  }

  const double btt = a.lay_d[LD_BT_TOP * TL + g], btb = a.lay_d[LD_BT_BOT * TL + g];
  const double fact = beam ? exp(-taub / mu0) : 0.0;
#pragma unroll
  for (int r = 0; r < N; ++r) {
    if ((r & 1) != h) continue;
    a.eig[eo.at(t, eo.ztop() + r)] = zb[r] * btt + z0[r] + z1[r] * taut;
    a.eig[eo.at(t, eo.zbot() + r)] = zb[r] * btb + z0[r] + z1[r] * taub;
    a.eig[eo.at(t, eo.zfl() + r)] = zb[r] * fact + z0[r] + z1[r] * taub;
  }
  if (h == 0) a.eig[eo.at(t, eo.pls())] = xr0 + xr1 * taub;
}

// ===========================================================================
// k_boundary: elimination of the boundary-value problem by LU with partial
// pivoting, one layer panel at a time.
//
// The (N*L)-unknown band system of setMatrix/solve0 (src/FluxSolver.cpp:
// 838-1131) has, for the constants C_i of layer i, the equations carried
// from above (NN rows in C_i) and the N continuity equations of the
// interface below layer i (rows in C_i and C_{i+1}).  Eliminating C_i from
// these 3NN rows with partial pivoting -- the pivot sequence LINPACK's
// xGBFA would choose on the band (src/BandSolver.hpp:29-126) up to ties --
// leaves N pivot rows (the U rows of layer i) and NN rows in C_{i+1} that
// are carried to the next panel.
//
// One lane per panel row ("slot"): a group of LPI lanes (whole 16-lane DPP
// rows) per instance, 64/LPI instances per wavefront, one wavefront per
// workgroup.  A lane holds its row's left block (N coefficients of C_i),
// right block (N coefficients of C_{i+1}) and right-hand side.  Step c of a
// panel: a group max of |a_rc| over the candidate rows (DPP rotations),
// the first lane attaining it (ballot) is the pivot; it publishes its row in
// LDS and every candidate row subtracts its multiple.  Pivot rows stay in
// their lanes, are marked used, are written out at the end of the panel as
// the U rows of the layer, and the next panel refills the used slots with
// the new interface rows (row q = rank of the slot among the used ones).
// Slots are ordered as rows of the band, so "first largest" picks the row
// LINPACK picks.  The layer-task records are contiguous, so a row fill reads
// whole cache lines.
// ===========================================================================
template <int NN>
struct RowCfg {
  static constexpr int N = 2 * NN;
  static constexpr int R = 3 * NN;
  static constexpr int LPI = (R <= 16) ? 16 : 32;  // lanes per instance
  static constexpr int G = 64 / LPI;               // instances per wavefront
  static constexpr int W = 2 * N + 2;              // pivot-row record in LDS
  static_assert(R <= 32, "one slot per lane of a 32-lane group");
};

// Maximum of x over the LPI lanes of the caller's group (every lane gets it):
// the 16-lane row maximum, joined across the two rows of a 32-lane group by
// a permlane16 swap.
template <int LPI>
__device__ __forceinline__ double group_max(double x) {
  x = row_max(x);
  if constexpr (LPI == 32) x = row_pair_max(x);
  return x;
}

// Row q of the reference's N x N eigenvector matrix G of layer-task record
// rec (see EigOff): out[j] = G(q, j).
template <int NN>
__device__ __forceinline__ void g_row(const double* rec, int q, double* out) {
  constexpr EigOff eo(2 * NN);
  const bool lower = q >= NN;
  const int r = lower ? q - NN : NN - 1 - q;
  const double* up = rec + (lower ? eo.gp() : eo.gm()) + r;  // columns NN + j'
  const double* dn = rec + (lower ? eo.gm() : eo.gp()) + r;  // columns NN - 1 - j'
#pragma unroll
  for (int jp = 0; jp < NN; ++jp) {
    out[NN + jp] = up[jp * NN];
    out[NN - 1 - jp] = -dn[jp * NN];
  }
}

// Bits of the caller's group in a wavefront ballot.
template <int LPI>
__device__ __forceinline__ unsigned group_bits(unsigned long long bal, int g) {
  const unsigned long long m = (LPI == 32) ? 0xffffffffull : 0xffffull;
  return (unsigned)((bal >> (g * LPI)) & m);
}

template <int NN>
__global__ __launch_bounds__(64, 3) void k_boundary(FluxArgs a) {
  using C = RowCfg<NN>;
  constexpr int N = C::N, R = C::R, LPI = C::LPI, G = C::G, W = C::W;
  constexpr EigOff eo(N);
  constexpr TabOff to(N);
  constexpr int PS = panel_stride(N);
  constexpr int ES = eo.stride();
  __shared__ double sRow[G * W];     // pivot row of the current step, per group
  __shared__ double sE[G * 2 * NN];  // e^{-k dtau} of the two layers of a panel
  __shared__ double sRS[G * N];      // bottom reflection row

  const int g = threadIdx.x / LPI;
  const int s = threadIdx.x % LPI;  // slot
  const bool slot = s < R;
  const int bk0 = blockIdx.x * G + g;
  const bool live = bk0 < a.nb;
  const int bk = live ? bk0 : a.nb - 1;  // a spare group repeats the last instance, stores nothing
  const int i = a.b_inst[bk];
  const int64_t t0 = a.b_task0[bk];
  const int64_t lo = a.lay_off[i];
  const int64_t TL = a.total_layers;
  const double* sc = a.scal + (int64_t)i * DISORT_NSCAL;
  const double* tb = a.tab;
  const int M = a.inst_i[i * kInstI + II_NCUT];
  const bool lyrcut = a.inst_i[i * kInstI + II_LYRCUT] != 0;
  const bool diffusion = (a.flags & DISORT_F_DIFFUSION_BC) != 0;
  const bool thermal = (a.flags & DISORT_F_THERMAL) != 0;
  const double F0 = sc[DISORT_S_BEAM_FLUX], mu0 = sc[DISORT_S_BEAM_MU];
  const double alb = sc[DISORT_S_ALBEDO];
  double* const sr = sRow + g * W;
  double* const se = sE + g * 2 * NN;

  // e^{-k_j' dtau} of layer ta in se[j'] and of layer tb in se[NN + j'].
  auto exps = [&](int64_t ta, int64_t tb2) {
    wave_lds_order();
    if (s < 2 * NN) {
      const int64_t t = (s < NN) ? ta : tb2;
      const int jp = (s < NN) ? s : s - NN;
      se[s] = exp(-a.eig[eo.at(t, eo.k() + jp)] * a.lay_d[LD_DTAU_S * TL + lo + (t - t0)]);
    }
    wave_lds_order();
  };

  double A0[N], A1[N], rhs = 0.0;
#pragma unroll
  for (int j = 0; j < N; ++j) A0[j] = A1[j] = 0.0;

  // Top boundary in slots 0..NN-1: downward intensities at tau=0
  // (:890-907, :976-990), stored as if carried from a previous panel (right
  // block), which the first panel moves into the left block.  Slots NN..R-1
  // are free (used) for the first interface.
  exps(t0, t0);
  if (s < NN) {
    g_row<NN>(a.eig + t0 * ES, NN - 1 - s, A1);
#pragma unroll
    for (int j = 0; j < NN; ++j) A1[j] *= se[NN - 1 - j];  // e at the layer top for the decaying modes
    rhs = sc[DISORT_S_ISO_TOP] + a.inst_d[i * kInstD + ID_PL_TOP] - a.eig[eo.at(t0, eo.ztop() + NN - 1 - s)];
  }
  bool used = slot && s >= NN;

  // N elimination steps over the candidate rows; pivot rows are written to
  // pp as the U rows of the layer in pivot order: [c][j] left block
  // (diagonal as its reciprocal), N*N + [c][j] right block, 2N*N + c rhs.
  auto lu_panel = [&](bool excluded, double* pp) {
    int myc = -1;
    double myinv = 0.0;
    static_for<0, N>([&](auto cc) {
      constexpr int c = decltype(cc)::value;
      const bool cand = slot && !used && !excluded;
      const double key = cand ? fabs(A0[c]) : -1.0;
      const double m = group_max<LPI>(key);
      const int p = __ffs(group_bits<LPI>(__ballot(key == m), g)) - 1;
      wave_lds_order();
      if (s == p) {
#pragma unroll
        for (int k = c; k < N; ++k) sr[k] = A0[k];
#pragma unroll
        for (int k = 0; k < N; ++k) sr[N + k] = A1[k];
        sr[2 * N] = rhs;
      }
      wave_lds_order();
      const double piv = sr[c];
      const double inv = (piv != 0.0) ? 1.0 / (piv != 0.0 ? piv : 1.0) : 0.0;
      if (s == p) {
        used = true;
        myc = c;
        myinv = inv;
      } else if (cand) {
        const double l = A0[c] * inv;
#pragma unroll
        for (int k = c + 1; k < N; ++k) A0[k] -= l * sr[k];
#pragma unroll
        for (int k = 0; k < N; ++k) A1[k] -= l * sr[N + k];
        rhs -= l * sr[2 * N];
      }
    });
    if (live && myc >= 0) {
#pragma unroll
      for (int j = 0; j < N; ++j) {
        pp[myc * N + j] = (j == myc) ? myinv : A0[j];
        pp[N * N + myc * N + j] = A1[j];
      }
      pp[2 * N * N + myc] = rhs;
    }
  };

  // Interior panels: continuity at the interface below layer ip (:851-887, :1094-1118).
  for (int ip = 0; ip < M - 1; ++ip) {
    const int64_t tt = t0 + ip;
    exps(tt, tt + 1);
    const unsigned ub = group_bits<LPI>(__ballot(used), g);
    if (used) {
      const int q = __popc(ub & ((1u << s) - 1u));  // interface row of this slot
      g_row<NN>(a.eig + tt * ES, q, A0);
      g_row<NN>(a.eig + (tt + 1) * ES, q, A1);
#pragma unroll
      for (int j = 0; j < NN; ++j) {
        A0[NN + j] *= se[j];              // bottom of layer tt: growing modes decay
        A1[j] = -A1[j] * se[2 * NN - 1 - j];  // top of layer tt+1
        A1[NN + j] = -A1[NN + j];
      }
      rhs = a.eig[eo.at(tt + 1, eo.ztop() + q)] - a.eig[eo.at(tt, eo.zbot() + q)];
      used = false;
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) {
        A0[j] = A1[j];
        A1[j] = 0.0;
      }
    }
    lu_panel(false, a.panel + tt * PS);
  }

  // Last panel: carried equations plus the bottom boundary, Lambertian
  // reflection of the downward field (:909-959, :992-1091).  Only NN of the
  // free slots take bottom rows; the others stay out of the pivot search.
  {
    const int64_t tt = t0 + M - 1;
    const double* rec = a.eig + tt * ES;
    exps(tt, tt);
    const bool refl = !(lyrcut || diffusion);
    // Reflection row rs_j = sum_k w_k G(NN-1-k, j), one column per lane.
    if (s < N) {
      double rs = 0.0;
#pragma unroll
      for (int k = 0; k < NN; ++k) {
        const double wk = tb[to.cwt() + k] * tb[to.cmu() + k] * alb;
        rs += wk * g_entry<NN>(a, tt, NN - 1 - k, s);
      }
      sRS[g * N + s] = refl ? rs : 0.0;
    }
    wave_lds_order();
    double zrefl = 0.0;
#pragma unroll
    for (int k = 0; k < NN; ++k) zrefl += tb[to.cwt() + k] * tb[to.cmu() + k] * alb * rec[eo.zbot() + NN - 1 - k];
    const double btn = a.inst_d[i * kInstD + ID_BTRANS_NCUT];
    const double plb = a.inst_d[i * kInstD + ID_PL_BOT];
    const double plbd = a.inst_d[i * kInstD + ID_PL_BOT_DERIV];
    double extra = sc[DISORT_S_ISO_BOT];
    if (refl) extra += 2.0 * zrefl;
    if (F0 > 0.0 && !lyrcut) extra += alb * mu0 * F0 * btn / kPi;
    if (thermal) extra += (1.0 - alb) * plb;
    const unsigned ub = group_bits<LPI>(__ballot(used), g);
    const int q = __popc(ub & ((1u << s) - 1u));
    const bool bot = used && q < NN;
    const bool excluded = used && !bot;
    if (bot) {
      g_row<NN>(rec, NN + q, A0);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        A0[j] = (A0[j] - 2.0 * sRS[g * N + j]) * ((j < NN) ? 1.0 : se[j - NN]);
        A1[j] = 0.0;
      }
      rhs = -rec[eo.zbot() + NN + q] + (diffusion ? plb + tb[to.cmu() + q] * plbd : extra);
      used = false;
    } else if (!used) {
#pragma unroll
      for (int j = 0; j < N; ++j) {
        A0[j] = A1[j];
        A1[j] = 0.0;
      }
    }
    lu_panel(excluded, a.panel + tt * PS);
  }

  // Back substitution over the layers, lane = U row k of each panel (the
  // first 16-lane row of the group):
  // x_k = (r_k - sum_{c>k} U(k,c) x_c - sum_c V(k,c) xnext_c) / U(k,k).
  const int k = (s < N) ? s : 0;
  double xn = 0.0;  // this lane's entry of the next layer's constants
  for (int ip = M - 1; ip >= 0; --ip) {
    const int64_t tt = t0 + ip;
    const double* pp = a.panel + tt * PS;
    double v = pp[2 * N * N + k];
    if (ip < M - 1) {
      static_for<0, N>([&](auto cj) {
        constexpr int c = decltype(cj)::value;
        v -= pp[N * N + k * N + c] * row_bcast<c>(xn);
      });
    }
    const double rdg = pp[k * N + k];  // 1 / U(k,k)
    double xk = 0.0;
    static_for<0, N>([&](auto cj) {
      constexpr int c = N - 1 - decltype(cj)::value;
      const double xc = row_bcast<c>(v * rdg);
      xk = (k == c) ? xc : xk;
      v -= pp[k * N + c] * xc;
    });
    xn = xk;
    if (live && s < N) a.cvec[tt * N + k] = xk;
  }
}

// ===========================================================================
// Temperature Jacobians on the boundary system (src/FluxSolver.cpp:1294-1497
// computes them by back-substitution on its band LU).  With the optical
// properties fixed, a unit Planck value at level d changes only the thermal
// particular solution of the layers above and below level d (and, for the
// surface degree of freedom d = L+1, the bottom boundary row), so every
// degree of freedom is one more right-hand side of the same boundary
// system: k_jac_boundary eliminates K of them alongside one factorisation,
// k_jac_flux turns the mode amplitudes into the six flux and mean-intensity
// outputs and scales them by dB/dT.
// ===========================================================================

// Unit-source particular solution (BC form, row q) of layer lc of instance i
// at its top or bottom, for the degree of freedom d: the layer's Planck fit
// B(tau) = xr0 + xr1 tau through (B_top, B_bot) = ([d == lc], [d == lc+1])
// (src/FluxSolver.cpp:755-776), applied to the stored basis.
template <int NN>
__device__ __forceinline__ double jac_zunit(const FluxArgs& a, int i, int64_t lo, int64_t vo, int lc, int d, int q,
                                            bool bottom, double* pls = nullptr) {
  if (d != lc && d != lc + 1) {
    if (pls) *pls = 0.0;
    return 0.0;
  }
  const int64_t TL = a.total_layers, TV = a.total_levels;
  const double pt = (d == lc) ? 1.0 : 0.0, pb = 1.0 - pt;
  const double dts = a.lay_d[LD_DTAU_S * TL + lo + lc];
  const double taut = a.lev_d[VD_TAU_S * TV + vo + lc];
  const double taub = a.lev_d[VD_TAU_S * TV + vo + lc + 1];
  const double xr1 = (dts > 1e-4) ? (pb - pt) / dts : 0.0;
  const double xr0 = pt - xr1 * taut;
  if (pls) *pls = xr0 + xr1 * taub;
  const double* jb = a.jbase + ((int64_t)i * a.jac_nlyr + lc) * 2 * NN;
  const int r = (q >= NN) ? q - NN : NN - 1 - q;
  const double d0 = xr1 * jb[NN + r];
  const double z0 = 0.5 * (xr0 * jb[r] + ((q >= NN) ? d0 : -d0));
  const double z1 = 0.5 * xr1 * jb[r];
  return z0 + z1 * (bottom ? taub : taut);
}

// One LPI-lane group per (instance, group of K degrees of freedom): the
// panel elimination of k_boundary with K right-hand sides carried through
// every step, then K back substitutions.
template <int NN, int K>
__global__ __launch_bounds__(64, 2) void k_jac_boundary(FluxArgs a) {
  using C = RowCfg<NN>;
  constexpr int N = C::N, R = C::R, LPI = C::LPI, G = C::G, W = 2 * N + K;
  constexpr EigOff eo(N);
  constexpr TabOff to(N);
  constexpr int PS = 2 * N * N + K * N;
  constexpr int ES = eo.stride();
  __shared__ double sRow[G * W];
  __shared__ double sE[G * 2 * NN];
  __shared__ double sRS[G * N];

  const int L = a.jac_nlyr, D = L + 2, NG = (D + K - 1) / K;
  const int g = threadIdx.x / LPI;
  const int s = threadIdx.x % LPI;
  const bool slot = s < R;
  const int64_t nq = (int64_t)a.nb * NG;
  const int64_t q0 = (int64_t)blockIdx.x * G + g;
  const bool live = q0 < nq;
  const int64_t qq = live ? q0 : nq - 1;  // a spare group repeats the last task, stores nothing
  const int bk = (int)(qq / NG), dg = (int)(qq % NG);
  const int i = a.b_inst[bk];
  const int64_t t0 = a.b_task0[bk];
  const int64_t lo = a.lay_off[i], vo = lo + i;
  const int64_t TL = a.total_layers;
  const double* sc = a.scal + (int64_t)i * DISORT_NSCAL;
  const double* tb = a.tab;
  const int M = a.inst_i[i * kInstI + II_NCUT];
  const bool lyrcut = a.inst_i[i * kInstI + II_LYRCUT] != 0;
  const bool diffusion = (a.flags & DISORT_F_DIFFUSION_BC) != 0;
  const double alb = sc[DISORT_S_ALBEDO];
  double* const sr = sRow + g * W;
  double* const se = sE + g * 2 * NN;
  double* const pbase = a.jpanel + ((int64_t)i * NG + dg) * (int64_t)L * PS;

  if (live && s < K && dg * K + s < D) {
    const int d = dg * K + s;
    const double T = (d <= L) ? a.temper[vo + d] : sc[DISORT_S_TEMP_BOT];
    a.jdbdt[(int64_t)i * D + d] = planck_band_deriv(sc[DISORT_S_WN_LO], sc[DISORT_S_WN_HI], T);
  }
  if (live && dg == 0 && s == 0) a.jstatus[a.jac_c0 + i] = a.inst_i[i * kInstI + II_STATUS];

  auto exps = [&](int64_t ta, int64_t tb2) {
    wave_lds_order();
    if (s < 2 * NN) {
      const int64_t t = (s < NN) ? ta : tb2;
      const int jp = (s < NN) ? s : s - NN;
      se[s] = exp(-a.eig[eo.at(t, eo.k() + jp)] * a.lay_d[LD_DTAU_S * TL + lo + (t - t0)]);
    }
    wave_lds_order();
  };

  double A0[N], A1[N], rhs[K];
#pragma unroll
  for (int j = 0; j < N; ++j) A0[j] = A1[j] = 0.0;
#pragma unroll
  for (int kk = 0; kk < K; ++kk) rhs[kk] = 0.0;

  exps(t0, t0);
  if (s < NN) {
    g_row<NN>(a.eig + t0 * ES, NN - 1 - s, A1);
#pragma unroll
    for (int j = 0; j < NN; ++j) A1[j] *= se[NN - 1 - j];
#pragma unroll
    for (int kk = 0; kk < K; ++kk) rhs[kk] = -jac_zunit<NN>(a, i, lo, vo, 0, dg * K + kk, NN - 1 - s, false);
  }
  bool used = slot && s >= NN;

  auto lu_panel = [&](bool excluded, double* pp) {
    int myc = -1;
    double myinv = 0.0;
    static_for<0, N>([&](auto cc) {
      constexpr int c = decltype(cc)::value;
      const bool cand = slot && !used && !excluded;
      const double key = cand ? fabs(A0[c]) : -1.0;
      const double m = group_max<LPI>(key);
      const int p = __ffs(group_bits<LPI>(__ballot(key == m), g)) - 1;
      wave_lds_order();
      if (s == p) {
#pragma unroll
        for (int k = c; k < N; ++k) sr[k] = A0[k];
#pragma unroll
        for (int k = 0; k < N; ++k) sr[N + k] = A1[k];
#pragma unroll
        for (int kk = 0; kk < K; ++kk) sr[2 * N + kk] = rhs[kk];
      }
      wave_lds_order();
      const double piv = sr[c];
      const double inv = (piv != 0.0) ? 1.0 / (piv != 0.0 ? piv : 1.0) : 0.0;
      if (s == p) {
        used = true;
        myc = c;
        myinv = inv;
      } else if (cand) {
        const double l = A0[c] * inv;
#pragma unroll
        for (int k = c + 1; k < N; ++k) A0[k] -= l * sr[k];
#pragma unroll
        for (int k = 0; k < N; ++k) A1[k] -= l * sr[N + k];
#pragma unroll
        for (int kk = 0; kk < K; ++kk) rhs[kk] -= l * sr[2 * N + kk];
      }
    });
    if (live && myc >= 0) {
#pragma unroll
      for (int j = 0; j < N; ++j) {
        pp[myc * N + j] = (j == myc) ? myinv : A0[j];
        pp[N * N + myc * N + j] = A1[j];
      }
#pragma unroll
      for (int kk = 0; kk < K; ++kk) pp[2 * N * N + kk * N + myc] = rhs[kk];
    }
  };

  for (int ip = 0; ip < M - 1; ++ip) {
    const int64_t tt = t0 + ip;
    exps(tt, tt + 1);
    const unsigned ub = group_bits<LPI>(__ballot(used), g);
    if (used) {
      const int q = __popc(ub & ((1u << s) - 1u));
      g_row<NN>(a.eig + tt * ES, q, A0);
      g_row<NN>(a.eig + (tt + 1) * ES, q, A1);
#pragma unroll
      for (int j = 0; j < NN; ++j) {
        A0[NN + j] *= se[j];
        A1[j] = -A1[j] * se[2 * NN - 1 - j];
        A1[NN + j] = -A1[NN + j];
      }
#pragma unroll
      for (int kk = 0; kk < K; ++kk) {
        const int d = dg * K + kk;
        rhs[kk] = jac_zunit<NN>(a, i, lo, vo, ip + 1, d, q, false) - jac_zunit<NN>(a, i, lo, vo, ip, d, q, true);
      }
      used = false;
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) {
        A0[j] = A1[j];
        A1[j] = 0.0;
      }
    }
    lu_panel(false, pbase + (int64_t)ip * PS);
  }

  {
    const int64_t tt = t0 + M - 1;
    const double* rec = a.eig + tt * ES;
    exps(tt, tt);
    const bool refl = !(lyrcut || diffusion);
    if (s < N) {
      double rs = 0.0;
#pragma unroll
      for (int k = 0; k < NN; ++k) {
        const double wk = tb[to.cwt() + k] * tb[to.cmu() + k] * alb;
        rs += wk * g_entry<NN>(a, tt, NN - 1 - k, s);
      }
      sRS[g * N + s] = refl ? rs : 0.0;
    }
    wave_lds_order();
    const double dl = a.lay_d[LD_DTAU_S * TL + lo + L - 1];
    const unsigned ub = group_bits<LPI>(__ballot(used), g);
    const int q = __popc(ub & ((1u << s) - 1u));
    const bool bot = used && q < NN;
    const bool excluded = used && !bot;
    if (bot) {
      g_row<NN>(rec, NN + q, A0);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        A0[j] = (A0[j] - 2.0 * sRS[g * N + j]) * ((j < NN) ? 1.0 : se[j - NN]);
        A1[j] = 0.0;
      }
#pragma unroll
      for (int kk = 0; kk < K; ++kk) {
        const int d = dg * K + kk;
        double zrefl = 0.0;
        if (refl) {
#pragma unroll
          for (int k = 0; k < NN; ++k)
            zrefl += tb[to.cwt() + k] * tb[to.cmu() + k] * alb *
                     jac_zunit<NN>(a, i, lo, vo, M - 1, d, NN - 1 - k, true);
        }
        const double plb = (d == L + 1) ? 1.0 : 0.0;
        double extra = 0.0;
        if (refl) extra += 2.0 * zrefl;
        extra += (1.0 - alb) * plb;
        double plbd = 0.0;
        if (diffusion && dl > 0.0) plbd = (plb - ((d == L - 1) ? 1.0 : 0.0)) / dl;
        rhs[kk] = -jac_zunit<NN>(a, i, lo, vo, M - 1, d, NN + q, true) +
                  (diffusion ? plb + tb[to.cmu() + q] * plbd : extra);
      }
      used = false;
    } else if (!used) {
#pragma unroll
      for (int j = 0; j < N; ++j) {
        A0[j] = A1[j];
        A1[j] = 0.0;
      }
    }
    lu_panel(excluded, pbase + (int64_t)(M - 1) * PS);
  }

  // K back substitutions, lane = U row k of each panel.
  const int k = (s < N) ? s : 0;
  double xn[K];
#pragma unroll
  for (int kk = 0; kk < K; ++kk) xn[kk] = 0.0;
  for (int ip = M - 1; ip >= 0; --ip) {
    const double* pp = pbase + (int64_t)ip * PS;
    double v[K];
#pragma unroll
    for (int kk = 0; kk < K; ++kk) v[kk] = pp[2 * N * N + kk * N + k];
    if (ip < M - 1) {
      static_for<0, N>([&](auto cj) {
        constexpr int c = decltype(cj)::value;
        const double u = pp[N * N + k * N + c];
#pragma unroll
        for (int kk = 0; kk < K; ++kk) v[kk] -= u * row_bcast<c>(xn[kk]);
      });
    }
    const double rdg = pp[k * N + k];
    double xk[K];
#pragma unroll
    for (int kk = 0; kk < K; ++kk) xk[kk] = 0.0;
    static_for<0, N>([&](auto cj) {
      constexpr int c = N - 1 - decltype(cj)::value;
      const double u = pp[k * N + c];
#pragma unroll
      for (int kk = 0; kk < K; ++kk) {
        const double xc = row_bcast<c>(v[kk] * rdg);
        xk[kk] = (k == c) ? xc : xk[kk];
        v[kk] -= u * xc;
      }
    });
#pragma unroll
    for (int kk = 0; kk < K; ++kk) {
      xn[kk] = xk[kk];
      const int d = dg * K + kk;
      if (live && s < N && d < D) a.jcvec[(((int64_t)i * D + d) * L + ip) * N + k] = xk[kk];
    }
  }
}

// A group of LPL lanes per (level task, degree of freedom), consecutive
// groups on consecutive degrees of freedom of one level: the k_flux
// arithmetic with the unit-source amplitudes and particular solution, no
// beam, times dB/dT, written as [column][6][L+1][L+2].
template <int NN>
__global__ __launch_bounds__(64) void k_jac_flux(FluxArgs a) {
  constexpr int N = 2 * NN;
  constexpr EigOff eo(N);
  constexpr TabOff to(N);
  constexpr int LPL = FluxLanes<NN>::LPL, LPW = FluxLanes<NN>::LPW;
  __shared__ double sY[LPW * 2 * LPL];
  const int L = a.jac_nlyr, D = L + 2;
  const int lv = threadIdx.x / LPL;
  const int r = threadIdx.x % LPL;
  const int64_t q = (int64_t)blockIdx.x * LPW + lv;
  if (q >= a.n_lvtasks * D) return;  // whole groups leave together
  const int64_t vt = q / D;
  const int d = (int)(q - vt * D);
  const int bk = a.v_bk[vt];
  const int lu = a.v_lu[vt];
  const int i = a.b_inst[bk];
  const int64_t t0 = a.b_task0[bk];
  const int64_t lo = a.lay_off[i], vo = lo + i;
  const int64_t TL = a.total_layers;
  const double* tb = a.tab;
  const int ncut = a.inst_i[i * kInstI + II_NCUT];
  const bool lyrcut = a.inst_i[i * kInstI + II_LYRCUT] != 0;
  int lyu = (lu == 0) ? 0 : lu - 1;
  double o[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) o[f] = 0.0;
  if (!(lyrcut && lyu > ncut - 1)) {
    if (lyu >= ncut) lyu = ncut - 1;
    const int64_t tt = t0 + lyu;
    const double* rec = a.eig + eo.at(tt, 0);
    const double* cv = a.jcvec + (((int64_t)i * D + d) * L + lyu) * N;
    const double dts = a.lay_d[LD_DTAU_S * TL + lo + lyu];
    const bool act = r < NN;
    const int rr = act ? r : 0;
    double* y = sY + lv * 2 * LPL;
    {
      const double e = exp(-rec[eo.k() + rr] * dts);
      const double yd = cv[NN - 1 - rr] * ((lu == 0) ? e : 1.0);
      const double yu = cv[NN + rr] * ((lu == 0) ? 1.0 : e);
      wave_lds_order();
      if (act) {
        y[r] = yd;
        y[LPL + r] = yu;
      }
      wave_lds_order();
    }
    double pls = 0.0;
    double iu = jac_zunit<NN>(a, i, lo, vo, lyu, d, NN + rr, lu != 0, &pls);
    double id = jac_zunit<NN>(a, i, lo, vo, lyu, d, NN - 1 - rr, lu != 0);
    if (lu == 0) pls = (d == 0) ? 1.0 : 0.0;
    constexpr int kUnroll = NN <= 16 ? NN : 8;
#pragma unroll kUnroll
    for (int jp = 0; jp < NN; ++jp) {
      const double yd = y[jp], yu = y[LPL + jp];
      const double gp = rec[eo.gp() + rr + jp * NN];
      const double gm = rec[eo.gm() + rr + jp * NN];
      iu += gp * yu - gm * yd;
      id += gm * yu - gp * yd;
    }
    const double w = act ? tb[to.cwt() + rr] : 0.0, cm = tb[to.cmu() + rr];
    double mi = group_sum<LPL>(w * id + w * iu);
    double mid = group_sum<LPL>(w * id);
    double miu = group_sum<LPL>(w * iu);
    double fldn = group_sum<LPL>(w * id * cm);
    double flup = group_sum<LPL>(w * iu * cm);
    flup *= 2.0 * kPi;
    fldn *= 2.0 * kPi;
    mi = (2.0 * kPi * mi) / (4.0 * kPi);
    mid = (2.0 * kPi * mid) / (4.0 * kPi);
    miu = (2.0 * kPi * miu) / (4.0 * kPi);
    o[0] = flup;
    o[1] = fldn;
    o[2] = mi;
    o[3] = mid;
    o[4] = miu;
    o[5] = (1.0 - a.ssa[lo + lyu]) * 4.0 * kPi * (mi - pls);
  }
  if (r == 0) {
    const double sdb = a.jdbdt[(int64_t)i * D + d];
    const int64_t per = (int64_t)(L + 1) * D;
    double* jc = a.jac + (a.jac_c0 + i) * 6 * per + (int64_t)lu * D + d;
#pragma unroll
    for (int f = 0; f < 6; ++f) jc[f * per] = sdb * o[f];
  }
}

}  // namespace

// Host-callable launchers, one set per stream count.
template <int NN>
static void launch_all(const FluxArgs& a, hipStream_t s, hipEvent_t* ev, bool pair) {
  constexpr int N = 2 * NN;
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_column<N>, dim3(a.nb), dim3(64), 0, s, a);
  if (ev) (void)hipEventRecord(ev[1], s);
  if (pair)
    hipLaunchKernelGGL(k_layer2<NN>, dim3((unsigned)((a.n_tasks + 31) / 32)), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL(k_layer<NN>, dim3((unsigned)((a.n_tasks + 63) / 64)), dim3(64), 0, s, a);
  if (ev) (void)hipEventRecord(ev[2], s);
  if (a.jbase) {
    constexpr int K = kJacRhs;
    const int64_t NG = (a.jac_nlyr + 2 + K - 1) / K;
    const int64_t groups = (int64_t)a.nb * NG;
    hipLaunchKernelGGL((k_jac_boundary<NN, K>), dim3((unsigned)((groups + RowCfg<NN>::G - 1) / RowCfg<NN>::G)), dim3(64),
                       0, s, a);
    if (ev) (void)hipEventRecord(ev[3], s);
    constexpr int LPW = FluxLanes<NN>::LPW;
    const int64_t ft = a.n_lvtasks * (a.jac_nlyr + 2);
    hipLaunchKernelGGL(k_jac_flux<NN>, dim3((unsigned)((ft + LPW - 1) / LPW)), dim3(64), 0, s, a);
    if (ev) (void)hipEventRecord(ev[4], s);
    return;
  }
  hipLaunchKernelGGL(k_boundary<NN>, dim3((unsigned)((a.nb + RowCfg<NN>::G - 1) / RowCfg<NN>::G)), dim3(64), 0, s, a);
  if (ev) (void)hipEventRecord(ev[3], s);
  launch_flux_levels<NN>(a, s);
  if (ev) (void)hipEventRecord(ev[4], s);
}

bool launch_flux_bucket(int nstr, const FluxArgs& a, hipStream_t s, hipEvent_t* ev, bool pair) {
  switch (nstr) {
    case 4: launch_all<2>(a, s, ev, pair); return true;
    case 6: launch_all<3>(a, s, ev, pair); return true;
    case 8: launch_all<4>(a, s, ev, pair); return true;
    case 10: launch_all<5>(a, s, ev, pair); return true;
    case 12: launch_all<6>(a, s, ev, pair); return true;
    case 14: launch_all<7>(a, s, ev, pair); return true;
    case 16: launch_all<8>(a, s, ev, pair); return true;
    default: return false;
  }
}

}  // namespace dsa
