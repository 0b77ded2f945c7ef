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
//   k_boundary 16 lanes per instance: the (N*L)-unknown boundary system of
//             setMatrix/solve0 (:838-1131) eliminated layer by layer by LU
//             with partial pivoting on 3NN-row panels (the NN equations
//             carried from above plus the N continuity equations of the next
//             interface), then block back substitution.
//   k_flux    (flux_common.hpp) one lane per (instance, level): intensities
//             at the quadrature angles and the eight outputs (:1138-1287).
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
        double t = (at > 1e150) ? 0.5 / at : 1.0 / (at + sqrt(theta * theta + 1.0));
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
  const int64_t TL = a.total_layers, TV = a.total_levels, T = a.n_tasks;
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
  for (int q = 0; q < NT; ++q) a.eig[(eo.ztop() + q) * T + t] = m.L[q];
  jacobi_eigen<NN>(S, m.V, m.lam);
#pragma unroll
  for (int q = 0; q < NT; ++q) m.L[q] = a.eig[(eo.ztop() + q) * T + t];

  // Eigenvalues k = sqrt|lambda| and the G+/- blocks (:560-587):
  // u = S_d^-1 L^-T w is the eigenvector of (alpha+beta)(alpha-beta);
  // (alpha-beta) u / k = -S_d^-1 L w / k.
  double kv[NN];
#pragma unroll
  for (int j = 0; j < NN; ++j) {
    kv[j] = sqrt(fabs(m.lam[j]));
    a.eig[(eo.k() + j) * T + t] = kv[j];
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
      a.eig[(eo.gp() + r + j * NN) * T + t] = 0.5 * (tt + uu);
      a.eig[(eo.gm() + r + j * NN) * T + t] = 0.5 * (tt - uu);
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
    const double c1 = (1.0 - ws) * xr1, c0 = (1.0 - ws) * xr0;
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
    a.eig[(eo.ztop() + r) * T + t] = zb[r] * btt + z0[r] + z1[r] * taut;
    a.eig[(eo.zbot() + r) * T + t] = zb[r] * btb + z0[r] + z1[r] * taub;
    a.eig[(eo.zfl() + r) * T + t] = zb[r] * fact + z0[r] + z1[r] * taub;
  }
  a.eig[eo.pls() * T + t] = xr0 + xr1 * taub;
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
// 16 lanes per instance, 4 instances per 64-lane workgroup.  Lane j owns
// column j of the left block (C_i), column j of the right block (C_{i+1})
// and a copy of the right-hand side.  Rows live in 3NN fixed slots: pivot
// rows are marked used instead of being swapped, their values are picked by
// a select chain, and the next panel refills the used slots with the new
// interface rows (row = rank of the slot among the used ones), loaded from
// memory at computed addresses -- so no register is ever indexed dynamically.
// ===========================================================================
template <int NN>
struct Panel {
  static constexpr int N = 2 * NN;
  static constexpr int R = 3 * NN;
  double A0[R], A1[R], A2[R];
};

template <int SRC>
__device__ __forceinline__ int row_bcast_i(int x) {
  return __builtin_amdgcn_mov_dpp(x, 0x150 + SRC, 0xF, 0xF, false);
}


// One panel: N steps of partial-pivot elimination over the candidate slots
// (those not in `used`).  Pivot rows are written to pp as the U rows of this
// layer in pivot order: [c][j] left block, N*N + [c][j] right block, 2N*N + c
// right-hand side.  Returns the mask of pivot slots.
template <int NN>
__device__ __forceinline__ unsigned lu_panel(Panel<NN>& pn, unsigned used, int lane, bool live, double* pp) {
  constexpr int N = 2 * NN;
  constexpr int R = 3 * NN;
  static_for<0, N>([&](auto cc) {
    constexpr int c = decltype(cc)::value;
    // Pivot search in column c (meaningful in lane c): first largest |a|.
    int p = -1;
    double best = -1.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const bool cand = !((used >> r) & 1u);
      const double mag = fabs(pn.A0[r]);
      const bool take = cand && (mag > best);
      best = take ? mag : best;
      p = take ? r : p;
    }
    p = row_bcast_i<c>(p);
    // Pivot-row entries of this lane's columns (one-hot sum over the slots).
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double hit = (r == p) ? 1.0 : 0.0;
      v0 = fma(hit, pn.A0[r], v0);
      v1 = fma(hit, pn.A1[r], v1);
      v2 = fma(hit, pn.A2[r], v2);
    }
    // Multipliers of column c (lane c), broadcast to the row.  Every slot is
    // eliminated: the pivot rows of this panel are already stored and are
    // refilled by the next panel, so changing them is harmless.
    const double inv = (v0 != 0.0) ? 1.0 / (v0 != 0.0 ? v0 : 1.0) : 0.0;
    used |= 1u << p;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double l = row_bcast<c>(pn.A0[r] * inv);
      pn.A0[r] -= l * v0;
      pn.A1[r] -= l * v1;
      pn.A2[r] -= l * v2;
    }
    if (live) {
      pp[c * N + lane] = (lane == c) ? inv : v0;  // diagonal stored as its reciprocal
      pp[N * N + c * N + lane] = v1;
    }
    if (lane == 0) pp[2 * N * N + c] = v2;
  });
  return used;
}

template <int NN>
__global__ __launch_bounds__(64, 2) void k_boundary(FluxArgs a) {
  constexpr int N = 2 * NN;
  constexpr int R = 3 * NN;
  constexpr EigOff eo(N);
  constexpr TabOff to(N);
  constexpr int PS = panel_stride(N);
  static_assert(N <= 16, "one 16-lane row per instance");
  static_assert(R <= 32, "slot mask is 32 bits");
  const int grp = threadIdx.x >> 4;
  const int lane = threadIdx.x & 15;
  const int bk = blockIdx.x * 4 + grp;
  if (bk >= a.nb) return;
  const int i = a.b_inst[bk];
  const int64_t t0 = a.b_task0[bk];
  const int64_t lo = a.lay_off[i];
  const int64_t T = a.n_tasks, TL = a.total_layers;
  const double* sc = a.scal + (int64_t)i * DISORT_NSCAL;
  const double* tb = a.tab;
  const int M = a.inst_i[i * kInstI + II_NCUT];
  const bool lyrcut = a.inst_i[i * kInstI + II_LYRCUT] != 0;
  const bool diffusion = (a.flags & DISORT_F_DIFFUSION_BC) != 0;
  const bool thermal = (a.flags & DISORT_F_THERMAL) != 0;
  const double F0 = sc[DISORT_S_BEAM_FLUX], mu0 = sc[DISORT_S_BEAM_MU];
  const double alb = sc[DISORT_S_ALBEDO];
  const int j = (lane < N) ? lane : 0;  // idle lanes mirror column 0, results discarded
  const bool live = lane < N;

  // Mode column j over layer task tt: value at the layer top / bottom.
  auto efac = [&](int64_t tt, double& etop, double& ebot) {
    const int jp = (j >= NN) ? j - NN : NN - 1 - j;
    const double k = a.eig[(eo.k() + jp) * T + tt];
    const double e = exp(-k * a.lay_d[LD_DTAU_S * TL + lo + (tt - t0)]);
    etop = (j < NN) ? e : 1.0;
    ebot = (j < NN) ? 1.0 : e;
  };

  Panel<NN> pn;
  // Top boundary in slots 0..NN-1: downward intensities at tau=0
  // (:890-907, :976-990).  They are stored as if carried from a previous
  // panel (right block A1), which the first panel moves into the left block;
  // slots NN.. are free (marked used) for the first interface.
  {
    double etop, ebot;
    efac(t0, etop, ebot);
    const double top = sc[DISORT_S_ISO_TOP] + a.inst_d[i * kInstD + ID_PL_TOP];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const bool tbc = r < NN;
      pn.A0[r] = 0.0;
      pn.A1[r] = (tbc && live) ? g_entry<NN>(a, t0, tbc ? NN - 1 - r : 0, j) * etop : 0.0;
      pn.A2[r] = tbc ? top - a.eig[(eo.ztop() + (tbc ? NN - 1 - r : 0)) * T + t0] : 0.0;
    }
  }
  unsigned used = ((1u << R) - 1u) & ~((1u << NN) - 1u);

  // Interior panels: continuity at the interface below layer ip (:851-887, :1094-1118).
  for (int ip = 0; ip < M - 1; ++ip) {
    const int64_t tt = t0 + ip;
    double etop, ebot, etn, ebn;
    efac(tt, etop, ebot);
    efac(tt + 1, etn, ebn);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const bool fill = (used >> r) & 1u;
      const int q = fill ? __popc(used & ((1u << r) - 1u)) : 0;  // interface row of this slot
      const double gi = live ? g_entry<NN>(a, tt, q, j) : 0.0;
      const double gn = live ? g_entry<NN>(a, tt + 1, q, j) : 0.0;
      const double rhs = a.eig[(eo.ztop() + q) * T + tt + 1] - a.eig[(eo.zbot() + q) * T + tt];
      const double carried = pn.A1[r];
      pn.A0[r] = fill ? gi * ebot : carried;
      pn.A1[r] = fill ? -gn * etn : 0.0;
      pn.A2[r] = fill ? rhs : pn.A2[r];
    }
    used = lu_panel<NN>(pn, 0u, lane, live, a.panel + tt * PS);
  }

  // Last panel: carried equations plus the bottom boundary, Lambertian
  // reflection of the downward field (:909-959, :992-1091).  Only NN of the
  // free slots take bottom rows; the others stay out of the pivot search.
  {
    const int64_t tt = t0 + M - 1;
    double etop, ebot;
    efac(tt, etop, ebot);
    const bool refl = !(lyrcut || diffusion);
    double rs = 0.0, zrefl = 0.0;
#pragma unroll
    for (int k = 0; k < NN; ++k) {
      const double wk = tb[to.cwt() + k] * tb[to.cmu() + k] * alb;
      rs += wk * (live ? g_entry<NN>(a, tt, NN - 1 - k, j) : 0.0);
      zrefl += wk * a.eig[(eo.zbot() + NN - 1 - k) * T + tt];
    }
    rs = refl ? rs : 0.0;
    const double btn = a.inst_d[i * kInstD + ID_BTRANS_NCUT];
    const double plb = a.inst_d[i * kInstD + ID_PL_BOT];
    const double plbd = a.inst_d[i * kInstD + ID_PL_BOT_DERIV];
    double extra = sc[DISORT_S_ISO_BOT];
    if (refl) extra += 2.0 * zrefl;
    if (F0 > 0.0 && !lyrcut) extra += alb * mu0 * F0 * btn / kPi;
    if (thermal) extra += (1.0 - alb) * plb;
    unsigned excluded = 0u;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const bool fill = (used >> r) & 1u;
      const int q = fill ? __popc(used & ((1u << r) - 1u)) : 0;
      const bool bot = fill && q < NN;
      const int t = bot ? q : 0;
      const double gb = live ? g_entry<NN>(a, tt, NN + t, j) : 0.0;
      const double zb = -a.eig[(eo.zbot() + NN + t) * T + tt];
      const double carried = pn.A1[r];
      pn.A0[r] = bot ? (gb - 2.0 * rs) * ebot : (fill ? 0.0 : carried);
      pn.A1[r] = 0.0;
      pn.A2[r] = bot ? zb + (diffusion ? plb + tb[to.cmu() + t] * plbd : extra) : (fill ? 0.0 : pn.A2[r]);
      excluded |= (fill && !bot) ? (1u << r) : 0u;
    }
    lu_panel<NN>(pn, excluded, lane, live, a.panel + tt * PS);
  }

  // Back substitution over the layers, lane = U row k of each panel:
  // x_k = (r_k - sum_{c>k} U(k,c) x_c - sum_c V(k,c) xnext_c) / U(k,k).
  const int k = j;
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
    if (live) a.cvec[tt * N + k] = xk;
  }
}

}  // namespace

// Host-callable launchers, one set per stream count.
template <int NN>
static void launch_all(const FluxArgs& a, hipStream_t s, hipEvent_t* ev) {
  constexpr int N = 2 * NN;
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_column<N>, dim3(a.nb), dim3(64), 0, s, a);
  if (ev) (void)hipEventRecord(ev[1], s);
  hipLaunchKernelGGL(k_layer<NN>, dim3((unsigned)((a.n_tasks + 63) / 64)), dim3(64), 0, s, a);
  if (ev) (void)hipEventRecord(ev[2], s);
  hipLaunchKernelGGL(k_boundary<NN>, dim3((unsigned)((a.nb + 3) / 4)), dim3(64), 0, s, a);
  if (ev) (void)hipEventRecord(ev[3], s);
  hipLaunchKernelGGL(k_flux<NN>, dim3((unsigned)((a.n_lvtasks + 63) / 64)), dim3(64), 0, s, a);
  if (ev) (void)hipEventRecord(ev[4], s);
}

bool launch_flux_bucket(int nstr, const FluxArgs& a, hipStream_t s, hipEvent_t* ev) {
  switch (nstr) {
    case 4: launch_all<2>(a, s, ev); return true;
    case 6: launch_all<3>(a, s, ev); return true;
    case 8: launch_all<4>(a, s, ev); return true;
    case 10: launch_all<5>(a, s, ev); return true;
    case 12: launch_all<6>(a, s, ev); return true;
    case 14: launch_all<7>(a, s, ev); return true;
    case 16: launch_all<8>(a, s, ev); return true;
    default: return false;
  }
}

}  // namespace dsa
