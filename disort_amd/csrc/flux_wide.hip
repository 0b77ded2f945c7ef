// gfx950 workgroup kernels of the flux solve: the only kernels for the wide
// stream counts N = 32, 64 and for the pseudo-spherical beam, and an
// alternative schedule for N = 4 ... 16.
//
// The register kernels of flux_kernels.hip keep a whole NN x NN eigenproblem
// in one lane and a 3NN-row panel in 16 or 32 lanes; at N = 32 and 64 that
// no longer fits.  Here
//   k_layer_wg     one workgroup per (instance, layer): one wavefront up to
//                  N = 32, four at N = 64.  The matrices of the reduced
//                  eigenproblem (src/FluxSolver.cpp:459-588) live in three
//                  LDS slots; Cholesky factors are built column by column,
//                  the symmetric eigenproblem is solved by parallel cyclic
//                  Jacobi (round-robin pairing, NN/2 disjoint rotations per
//                  round), and the particular solutions (:595-643, :745-831)
//                  use the same eigen-decomposition / Cholesky formulation as
//                  the register kernel, with one thread per vector entry.
//   k_boundary_wg  one workgroup per instance: the panel LU with partial
//                  pivoting of the register kernel, with one thread per
//                  (panel row, block): rows stay in registers, the pivot row
//                  and column go through LDS, the elimination steps are a
//                  run-time loop.  k_boundary_wg1 is its one-wavefront form
//                  (lane = panel row holding both blocks), used up to N = 16.
//                  Pivot choices and operation order are those of the
//                  register kernel, so for N <= 16 both produce the same
//                  integration constants from the same eigen-system.
// k_column and k_flux (flux_common.hpp) are shared with the register path.
#include "wg_common.hpp"

namespace dsa {
namespace {

// ===========================================================================
// k_layer_wg: one workgroup per (instance, layer) task.
// ===========================================================================
template <int NN>
__global__ __launch_bounds__(WideCfg<NN>::LT, WideCfg<NN>::LW) void k_layer_wg(FluxArgs a) {
  using C = WideCfg<NN>;
  // Round-robin pairing over NNE = NN rounded up to even players; with odd
  // NN the extra player is a dummy and its pairs rotate nothing.
  constexpr int N = C::N, LD = C::LD, NNE = NN + (NN & 1), NP = NNE / 2;
  constexpr TabOff to(N);
  constexpr EigOff eo(N);
  constexpr bool OW = C::LT == 64;  // one wavefront per block
  // Three NN x NN slots: A_s, then S = L^T (-A_s) L (diagonalised in
  // place; formed once T = -A_s L is held in the third slot), then U, then
  // the Cholesky factor of -A_s (rebuilt when a thermal or spherical source
  // needs it) | -B_s, then its Cholesky factor L | T, then the eigenvectors.
  __shared__ double sbuf[3 * NN * LD];
  double* const sA = sbuf;
  double* const sS = sbuf;
  double* const sL = sbuf + NN * LD;
  double* const sV = sbuf + 2 * NN * LD;
  __shared__ double chi[N], p0[N];
  __shared__ double v0[NN], v1[NN], v2[NN], v3[NN], lam[NN], dv[NN];
  __shared__ double rc[NP], rs[NP], rt[NP], rapp[NP], raqq[NP], rapq[NP];
  __shared__ int rp[NP], rq[NP];
  __shared__ double red[C::LT / 64];
  __shared__ int flag;

  const int tid = threadIdx.x;
  const int64_t t = blockIdx.x;
  if (t >= a.n_tasks) return;
  const int i = a.t_inst[t];
  const int lc = a.t_lc[t];
  // No output depends on a layer below the absorption cut-off (ncut is nlyr
  // when the cut-off is not in effect).
  if (lc >= a.inst_i[(int64_t)i * kInstI + II_NCUT]) return;
  const int64_t g = a.lay_off[i] + lc;
  const int64_t v = a.lay_off[i] + i + lc;
  const int P = a.pstride[i];
  const double* pm = a.pmom + a.mom_off[i] + (int64_t)lc * P;
  const double* sc = a.scal + (int64_t)i * DISORT_NSCAL;
  const int64_t TL = a.total_layers, TV = a.total_levels;
  const double* tb = a.tab;
  const double ws = a.lay_d[LD_SSA_S * TL + g];
  const double f = a.lay_d[LD_F * TL + g];
  const double sig2 = a.lay_d[LD_SIG2 * TL + g];
  const double F0 = sc[DISORT_S_BEAM_FLUX], mu0 = sc[DISORT_S_BEAM_MU];
  const bool beam = F0 > 0.0;

  // Scaled phase coefficients chi'_l (src/FluxSolver.cpp:342-388) and
  // Legendre polynomials at -mu0.
  for (int l = tid; l < N; l += blockDim.x) {
    const double pl = pm[l];
    double c;
    if (sig2 < 0.0) c = (2.0 * l + 1.0) * ws * pl;
    else if (sig2 == 0.0) c = (2.0 * l + 1.0) * ws * (pl - f) / (1.0 - f);
    else c = (2.0 * l + 1.0) * ws * (pl - f * exp(-(double)(l * l) / (2.0 * sig2))) / (1.0 - f);
    chi[l] = c;
  }
  if (tid == 0) {
    const double x = -mu0;
    p0[0] = 1.0;
    p0[1] = x;
    for (int l = 2; l < N; ++l) p0[l] = ((2.0 * l - 1.0) * x * p0[l - 1] - (l - 1.0) * p0[l - 2]) / l;
  }
  __syncthreads();

  // A_s and -B_s from the even / odd Legendre sums (:494-512); the lower
  // triangle is computed and mirrored so both are exactly symmetric.
  for (int e = tid; e < NN * NN; e += blockDim.x) {
    const int r0 = e / NN, c0 = e % NN;
    const int r = r0 >= c0 ? r0 : c0, c = r0 >= c0 ? c0 : r0;
    double se = 0.0, so = 0.0;
    for (int l = 0; l < N; ++l) {
      const double q = chi[l] * tb[to.Q() + l * NN + r] * tb[to.Q() + l * NN + c];
      if (l % 2 == 0) se += q;
      else so += q;
    }
    const double dg = (r == c) ? 1.0 / tb[to.cmu() + r] : 0.0;
    sA[r0 * LD + c0] = se - dg;
    sL[r0 * LD + c0] = dg - so;
  }
  __syncthreads();
  if (!blk_cholesky<NN, OW>(sL, LD, &flag) && tid == 0) atomicOr(&a.inst_i[i * kInstI + II_STATUS], DISORT_ST_EIGEN);

  // S = L^T (-A_s) L: T = -A_s L into sV, then S (lower, mirrored).
  for (int e = tid; e < NN * NN; e += blockDim.x) {
    const int r = e / NN, c = e % NN;
    double s = 0.0;
    for (int k = c; k < NN; ++k) s -= sA[r * LD + k] * sL[k * LD + c];
    sV[r * LD + c] = s;
  }
  __syncthreads();
  for (int e = tid; e < NN * NN; e += blockDim.x) {
    const int c0 = e / NN, r0 = e % NN;
    if (r0 <= c0) {
      double s = 0.0;
      for (int k = r0; k < NN; ++k) s += sL[k * LD + r0] * sV[k * LD + c0];
      sS[c0 * LD + r0] = s;
      sS[r0 * LD + c0] = s;
    }
  }
  __syncthreads();
  for (int e = tid; e < NN * NN; e += blockDim.x) sV[e / NN * LD + e % NN] = (e / NN == e % NN) ? 1.0 : 0.0;
  __syncthreads();

  // Parallel cyclic Jacobi: each round rotates NN/2 disjoint index pairs
  // (round-robin schedule), each sweep visits every pair once.
  for (int sweep = 0; sweep < 40; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int e = tid; e < NN * NN; e += blockDim.x) {
      const int r = e / NN, c = e % NN;
      const double x = sS[r * LD + c];
      if (r > c) off += x * x;
      if (r == c) dia += x * x;
    }
    off = blk_sum<OW>(off, red);
    dia = blk_sum<OW>(dia, red);
    if (off == 0.0 || off <= 1e-36 * dia) break;
    for (int round = 0; round < NNE - 1; ++round) {
      if (tid < NP) {
        const int k = tid;
        const int pa = (k == 0) ? 0 : 1 + (k - 1 + round) % (NNE - 1);
        const int pb = 1 + (NNE - 2 - k + round) % (NNE - 1);
        const int p = pa < pb ? pa : pb, q = pa < pb ? pb : pa;
        rp[k] = p;
        rq[k] = q;
        if (q < NN) {
          const double apq = sS[q * LD + p], app = sS[p * LD + p], aqq = sS[q * LD + q];
          // t = tan of the rotation angle, the root of t^2 + 2 theta t = 1
          // (theta = d / h) of smaller magnitude, sign(theta) |h| /
          // (|d| + sqrt(d^2 + h^2)): one division instead of three.
          const double d = aqq - app, h = 2.0 * apq;
          const bool neg = (d != 0.0) && ((d < 0.0) != (h < 0.0));
          double tt = (neg ? -fabs(h) : fabs(h)) / (fabs(d) + sqrt(d * d + h * h));
          tt = (apq == 0.0) ? 0.0 : tt;
          const double cs = 1.0 / sqrt(tt * tt + 1.0);
          rc[k] = cs;
          rs[k] = tt * cs;
          rt[k] = tt;
          rapp[k] = app;
          raqq[k] = aqq;
          rapq[k] = apq;
        } else {  // pair with the dummy player: identity
          rc[k] = 1.0;
          rs[k] = 0.0;
          rt[k] = 0.0;
        }
      }
      blk_sync<OW>();
      // S <- J^T S J by 2x2 blocks (pair k rows, pair k2 columns), k >= k2;
      // V <- V J by column pairs.  Indices >= NN (the dummy) are absent.
      for (int e = tid; e < NP * NP; e += blockDim.x) {
        const int k = e / NP, k2 = e % NP;
        if (k2 > k) continue;
        const int p = rp[k], q = rq[k];
        const bool qd = q >= NN;
        if (k == k2) {
          if (!qd) {
            const double tt = rt[k], apq = rapq[k];
            sS[p * LD + p] = rapp[k] - tt * apq;
            sS[q * LD + q] = raqq[k] + tt * apq;
            sS[q * LD + p] = 0.0;
            sS[p * LD + q] = 0.0;
          }
        } else {
          const int p2 = rp[k2], q2 = rq[k2];
          const bool q2d = q2 >= NN;
          const double c1 = rc[k], s1 = rs[k], c2 = rc[k2], s2 = rs[k2];
          const double bpp = sS[p * LD + p2];
          const double bpq = q2d ? 0.0 : sS[p * LD + q2];
          const double bqp = qd ? 0.0 : sS[q * LD + p2];
          const double bqq = (qd || q2d) ? 0.0 : sS[q * LD + q2];
          const double upp = c1 * bpp - s1 * bqp, upq = c1 * bpq - s1 * bqq;  // rows
          const double uqp = s1 * bpp + c1 * bqp, uqq = s1 * bpq + c1 * bqq;
          const double npp = c2 * upp - s2 * upq, npq = s2 * upp + c2 * upq;  // columns
          const double nqp = c2 * uqp - s2 * uqq, nqq = s2 * uqp + c2 * uqq;
          sS[p * LD + p2] = npp;
          sS[p2 * LD + p] = npp;
          if (!q2d) {
            sS[p * LD + q2] = npq;
            sS[q2 * LD + p] = npq;
          }
          if (!qd) {
            sS[q * LD + p2] = nqp;
            sS[p2 * LD + q] = nqp;
          }
          if (!qd && !q2d) {
            sS[q * LD + q2] = nqq;
            sS[q2 * LD + q] = nqq;
          }
        }
      }
      for (int e = tid; e < NN * NP; e += blockDim.x) {
        const int m = e / NP, k = e % NP;
        const int p = rp[k], q = rq[k];
        if (q >= NN) continue;
        const double vp = sV[m * LD + p], vq = sV[m * LD + q];
        sV[m * LD + p] = rc[k] * vp - rs[k] * vq;
        sV[m * LD + q] = rs[k] * vp + rc[k] * vq;
      }
      blk_sync<OW>();
    }
  }
  if (tid < NN) {
    lam[tid] = sS[tid * LD + tid];
    a.eig[eo.at(t, eo.k() + tid)] = sqrt(fabs(lam[tid]));
  }
  __syncthreads();

  // G+/- (:560-587): u = S_d^-1 L^-T w, (alpha-beta) u / k = -S_d^-1 L w / k.
  // U = L^-T V column by column into sS (S is no longer needed), then every
  // entry (r, j) combines U with (L V)(r, j).
  blk_ltsolve_cols<NN>(sL, LD, sV, sS);
  __syncthreads();
  for (int e = tid; e < NN * NN; e += blockDim.x) {
    const int r = e % NN, j = e / NN;
    double y = 0.0;
    for (int m = 0; m <= r; ++m) y += sL[r * LD + m] * sV[m * LD + j];
    const double sdi = tb[to.sdi() + r];
    const double uu = sdi * sS[r * LD + j];
    const double tt = -sdi * y / sqrt(fabs(lam[j]));
    a.eig[eo.at(t, eo.gp() + r + j * NN)] = 0.5 * (tt + uu);
    a.eig[eo.at(t, eo.gm() + r + j * NN)] = 0.5 * (tt - uu);
  }

  // Particular solutions, reference angle order: NN+r is +mu_r, NN-1-r is -mu_r.
  const bool thermal = (a.flags & DISORT_F_THERMAL) != 0;
  const bool spherical = beam && (a.flags & DISORT_F_SPHERICAL) != 0;
  const double taut = a.lev_d[VD_TAU_S * TV + v];
  const double taub = a.lev_d[VD_TAU_S * TV + v + 1];

  // R = Cholesky factor of -A_s (in sA), for the thermal source and the
  // unattenuated spherical source.
  if (thermal || spherical) {
    // -A_s again (the slot held S and U meanwhile), entry by entry as built
    // above, then its Cholesky factor.
    __syncthreads();
    for (int e = tid; e < NN * NN; e += blockDim.x) {
      const int r0 = e / NN, c0 = e % NN;
      const int r = r0 >= c0 ? r0 : c0, c = r0 >= c0 ? c0 : r0;
      double se = 0.0;
      for (int l = 0; l < N; ++l) {
        const double q = chi[l] * tb[to.Q() + l * NN + r] * tb[to.Q() + l * NN + c];
        if (l % 2 == 0) se += q;
      }
      const double dg = (r == c) ? 1.0 / tb[to.cmu() + r] : 0.0;
      sA[r0 * LD + c0] = -(se - dg);
    }
    __syncthreads();
    blk_cholesky<NN, OW>(sA, LD, &flag);
  }

  // Solves (I + xba diag(mu) - C) z = b for z_up (+mu_r) and z_dn (-mu_r),
  // given this thread's bs = b_up + b_dn and bd = b_up - b_dn (thread r < NN).
  // With s = z_up + z_dn, d = z_up - z_dn:
  //   ((alpha+beta)(alpha-beta) - xba^2) d = xba p + (alpha+beta) q,
  //   s = ((alpha-beta) d - q) / xba,   p = -bs/mu, q = -bd/mu,
  // through the eigen-decomposition; for xba = 0 directly,
  //   d = (alpha-beta)^-1 q,  s = (alpha+beta)^-1 p  (Cholesky factors L, R).
  auto solve_ab = [&](double bs, double bd, bool xba_zero, double mub, double& zu, double& zd) {
    double pr = 0.0, qr = 0.0;
    if (tid < NN) {
      const double cm = tb[to.cmu() + tid];
      pr = -bs / cm;
      qr = -bd / cm;
    }
    double sr = 0.0, dr = 0.0;
    if (xba_zero) {
      if (tid < NN) {
        v0[tid] = tb[to.sd() + tid] * qr;
        v1[tid] = tb[to.sd() + tid] * pr;
      }
      __syncthreads();
      blk_lsolve<NN, OW>(sL, LD, v0);
      blk_ltsolve<NN, OW>(sL, LD, v0);
      blk_lsolve<NN, OW>(sA, LD, v1);
      blk_ltsolve<NN, OW>(sA, LD, v1);
      if (tid < NN) {
        dr = -tb[to.sdi() + tid] * v0[tid];
        sr = -tb[to.sdi() + tid] * v1[tid];
      }
    } else {
      if (tid < NN) v0[tid] = tb[to.sd() + tid] * qr;  // w
      __syncthreads();
      // (alpha+beta) q = -S_d^-1 L^-T V Lambda V^T L^-1 S_d q
      blk_lsolve<NN, OW>(sL, LD, v0);
      blk_vdv<NN, OW>(sV, LD, lam, v0, v1);
      blk_ltsolve<NN, OW>(sL, LD, v0);
      if (tid < NN) {
        const double apbq = -tb[to.sdi() + tid] * v0[tid];
        const double rr = (pr + mub * apbq) / mub;
        v2[tid] = rr * tb[to.sd() + tid];
        dv[tid] = 1.0 / (lam[tid] - 1.0 / (mub * mub));
      }
      __syncthreads();
      // d = S_d^-1 L^-T V (Lambda - xba^2)^-1 V^T L^T S_d r
      blk_lmul<NN, true, OW>(sL, LD, v2, v1);
      blk_vdv<NN, OW>(sV, LD, dv, v2, v1);
      blk_ltsolve<NN, OW>(sL, LD, v2);
      if (tid < NN) {
        dr = tb[to.sdi() + tid] * v2[tid];
        v3[tid] = tb[to.sd() + tid] * dr;
      }
      __syncthreads();
      // (alpha-beta) d = -S_d^-1 L L^T S_d d
      blk_lmul<NN, true, OW>(sL, LD, v3, v1);
      blk_lmul<NN, false, OW>(sL, LD, v3, v1);
      if (tid < NN) {
        const double ambd = -tb[to.sdi() + tid] * v3[tid];
        sr = mub * (ambd - qr);
      }
    }
    zu = 0.5 * (sr + dr);
    zd = 0.5 * (sr - dr);
    __syncthreads();
  };

  // Beam source sums over even / odd l (src/FluxSolver.cpp:628-632):
  // b_up + b_dn = bsr, b_up - b_dn = bdr.
  double bsr = 0.0, bdr = 0.0;
  if (beam && tid < NN) {
    double se = 0.0, so = 0.0;
    for (int l = 0; l < N; ++l) {
      const double term = chi[l] * tb[to.P() + l * NN + tid] * p0[l];
      if (l % 2 == 0) se += term;
      else so += term;
    }
    bsr = F0 * se / (2.0 * kPi);
    bdr = F0 * so / (2.0 * kPi);
  }

  // Beam particular solution in the form e^{-xba tau} (zu0 + zu1 tau) (up)
  // and likewise (down); plane-parallel: xba = 1/mu0, no slope, and the
  // exponential factors come from the beam transmission.
  double zbu = 0.0, zbd = 0.0, s1u = 0.0, s1d = 0.0, xba = 0.0;
  if (beam && !spherical) {
    // Dither mu0 when 1 - mu0^2 k^2 is tiny for any mode (:604-618).
    double mub = mu0;
    const double mu0sq = mu0 * mu0;
    for (int j = 0; j < NN; ++j) {
      const double kj = sqrt(fabs(lam[j]));
      if (fabs(1.0 - mu0sq * kj * kj) < 1000.0 * kEps) mub = 0.999 * mu0;
    }
    solve_ab(bsr, bdr, false, mub, zbu, zbd);
  } else if (spherical) {
    // Pseudo-spherical source (src/FluxSolver.cpp:650-738): xba = 1/CH with
    // its dithers, linear-in-tau fit of e^{xba tau} times the attenuated
    // source between the layer's top and bottom, two solves.
    const double big = sqrt(1.79769313486231571e308) / 1.0e10;
    xba = 1.0 / a.lay_d[LD_CH * TL + g];
    if (fabs(xba) > big && taub > 1.0) xba = 0.0;
    if (fabs(xba * taub) > log(big)) xba = 0.0;
    if (fabs(xba) > 1.0e-5) {
      for (int iq = 0; iq < NN; ++iq)
        if (fabs((fabs(xba) - 1.0 / tb[to.cmu() + iq]) / xba) < 0.05) {
          xba *= 1.001;
          break;
        }
    }
    // Frobenius-norm dither of I + xba diag(mu) - C (:719-724).
    double fro = 0.0;
    for (int e = tid; e < N * N; e += blockDim.x) {
      const int i0 = e / N, j0 = e % N;
      const int ri = i0 < NN ? i0 : i0 - NN, rj = j0 < NN ? j0 : j0 - NN;
      const double mi = i0 < NN ? tb[to.cmu() + ri] : -tb[to.cmu() + ri];
      double c = 0.0;
      for (int l = 0; l < N; ++l) {
        const double sgn = ((l & 1) && ((i0 < NN) != (j0 < NN))) ? -1.0 : 1.0;
        c += chi[l] * tb[to.P() + l * NN + ri] * tb[to.P() + l * NN + rj] * sgn;
      }
      c *= 0.5 * tb[to.cwt() + rj];
      const double aij = ((i0 == j0) ? 1.0 + xba * mi : 0.0) - c;
      fro += aij * aij;
    }
    fro = blk_sum(fro, red);
    if (1.0 / sqrt(fro) < 1.0e-4) {
      if (xba == 0.0) xba = 5.0e-9;
      xba *= 1.00000005;
    }
    const double q0a = exp(-((lc == 0) ? 0.0 : a.lay_d[LD_CHTAU * TL + g - 1]));
    const double q2a = exp(-a.lay_d[LD_CHTAU * TL + g]);
    const double et = exp(xba * taut), eb = exp(xba * taub), dtl = taub - taut;
    // Source at the up / down nodes: zj = (bs +/- bd) / 2.
    const double zju = 0.5 * (bsr + bdr), zjd = 0.5 * (bsr - bdr);
    const double x1u = (q2a * zju * eb - q0a * zju * et) / dtl, x1d = (q2a * zjd * eb - q0a * zjd * et) / dtl;
    const double x0u = q0a * zju * et - x1u * taut, x0d = q0a * zjd * et - x1d * taut;
    const bool xz = (xba == 0.0);
    const double mubs = xz ? 0.0 : 1.0 / xba;
    solve_ab(x1u + x1d, x1u - x1d, xz, mubs, s1u, s1d);
    double r0u = 0.0, r0d = 0.0;
    if (tid < NN) {
      const double cm = tb[to.cmu() + tid];
      r0u = x0u + cm * s1u;
      r0d = x0d - cm * s1d;
    }
    solve_ab(r0u + r0d, r0u - r0d, xz, mubs, zbu, zbd);
  }

  double xr0 = 0.0, xr1 = 0.0, z0u = 0.0, z0d = 0.0, z1h = 0.0;
  if (thermal) {
    // B(tau) = xr0 + xr1 tau (:755-776); z1 and z0 from the Cholesky factor
    // R of -A_s (see the register kernel).
    const double pt = a.lev_d[VD_PLANCK * TV + v], pb = a.lev_d[VD_PLANCK * TV + v + 1];
    const double dts = a.lay_d[LD_DTAU_S * TL + g];
    if (dts > 1e-4) xr1 = (pb - pt) / dts;
    xr0 = pt - xr1 * taut;
    if (tid < NN) v0[tid] = tb[to.sd() + tid] * (2.0 / tb[to.cmu() + tid]);
    __syncthreads();
    blk_lsolve<NN, OW>(sA, LD, v0);
    blk_ltsolve<NN, OW>(sA, LD, v0);
    const double c1 = (1.0 - ws) * xr1, c0 = (1.0 - ws) * xr0;
    double base = 0.0;
    if (tid < NN) {
      base = v0[tid] * tb[to.sdi() + tid];
      z1h = 0.5 * c1 * base;
      v1[tid] = tb[to.sd() + tid] * z1h;
    }
    __syncthreads();
    blk_lsolve<NN, OW>(sL, LD, v1);
    blk_ltsolve<NN, OW>(sL, LD, v1);
    if (tid < NN) {
      const double d0 = 2.0 * tb[to.sdi() + tid] * v1[tid];
      const double s0 = c0 * base;
      z0u = 0.5 * (s0 + d0);
      z0d = 0.5 * (s0 - d0);
    }
  }

  if (tid < NN) {
    const int ru = NN + tid, rd = NN - 1 - tid;
    double btu, btd, bbu, bbd, bfu, bfd;  // beam part at the top / bottom (BC form) / bottom (flux form)
    if (spherical) {
      const double at = exp(-xba * taut), ab = exp(-xba * taub);
      btu = at * (zbu + s1u * taut);
      btd = at * (zbd + s1d * taut);
      bbu = bfu = ab * (zbu + s1u * taub);
      bbd = bfd = ab * (zbd + s1d * taub);
    } else {
      const double btt = a.lay_d[LD_BT_TOP * TL + g], btb = a.lay_d[LD_BT_BOT * TL + g];
      const double fact = beam ? exp(-taub / mu0) : 0.0;
      btu = zbu * btt;
      btd = zbd * btt;
      bbu = zbu * btb;
      bbd = zbd * btb;
      bfu = zbu * fact;
      bfd = zbd * fact;
    }
    a.eig[eo.at(t, eo.ztop() + ru)] = btu + z0u + z1h * taut;
    a.eig[eo.at(t, eo.ztop() + rd)] = btd + z0d + z1h * taut;
    a.eig[eo.at(t, eo.zbot() + ru)] = bbu + z0u + z1h * taub;
    a.eig[eo.at(t, eo.zbot() + rd)] = bbd + z0d + z1h * taub;
    a.eig[eo.at(t, eo.zfl() + ru)] = bfu + z0u + z1h * taub;
    a.eig[eo.at(t, eo.zfl() + rd)] = bfd + z0d + z1h * taub;
  }
  if (tid == 0) a.eig[eo.at(t, eo.pls())] = xr0 + xr1 * taub;
}

// ===========================================================================
// k_boundary_wg: one workgroup per instance.  Thread (h, r): panel row r,
// block h of the row (N coefficients); h = 1 also owns the right-hand side.
// In panel ip the left block (columns of layer ip) is held by h = ip & 1, so
// the equations carried to the next panel stay in place and the other
// block takes the next layer's columns.
// ===========================================================================
template <int NN>
__global__ __launch_bounds__(WideCfg<NN>::BT, 2) void k_boundary_wg(FluxArgs a) {
  using C = WideCfg<NN>;
  constexpr int N = C::N, R = C::R, RP = C::RP, NWH = RP / 64;
  constexpr EigOff eo(N);
  constexpr TabOff to(N);
  constexpr int PS = panel_stride(N);
  __shared__ double prow[2 * N + 2];  // pivot row: left block (shifted), right block, rhs
  __shared__ double colc[2][RP];      // pivot-column entry of every row, by step parity
  __shared__ double red_v[NWH];
  __shared__ int red_i[NWH];
  __shared__ int uflag[RP];
  __shared__ double ecur[NN], enxt[NN], rsb[N];

  const int tid = threadIdx.x;
  const int h = tid / RP, r = tid % RP;
  const bool row = r < R;
  const int bk = blockIdx.x;
  if (bk >= a.nb) return;
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

  // exp(-k_jp dtau') of layer task tt into e[jp].
  auto layer_exp = [&](int64_t tt, double* e) {
    for (int jp = tid; jp < NN; jp += blockDim.x)
      e[jp] = exp(-a.eig[eo.at(tt, eo.k() + jp)] * a.lay_d[LD_DTAU_S * TL + lo + (tt - t0)]);
  };
  // Mode column j: factor at the layer top / bottom.
  auto etop_of = [](const double* e, int j) { return j < NN ? e[NN - 1 - j] : 1.0; };
  auto ebot_of = [](const double* e, int j) { return j < NN ? 1.0 : e[j - NN]; };

  double A[N];
  double rhs = 0.0;
  bool used;

  // Top boundary rows 0..NN-1 (:890-907, :976-990) in the left block of
  // panel 0 (h = 0); rows NN.. are free for the first interface.
  layer_exp(t0, ecur);
  __syncthreads();
  {
    const bool tbc = row && r < NN;
    const double top = sc[DISORT_S_ISO_TOP] + a.inst_d[i * kInstD + ID_PL_TOP];
#pragma unroll
    for (int j = 0; j < N; ++j) A[j] = 0.0;
    if (tbc && h == 0) {
#pragma unroll
      for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
        for (int j = j0; j < j0 + 8 && j < N; ++j) A[j] = g_entry<NN>(a, t0, NN - 1 - r, j) * etop_of(ecur, j);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    if (h == 1) rhs = tbc ? top - a.eig[eo.at(t0, eo.ztop() + (tbc ? NN - 1 - r : 0))] : 0.0;
    used = !(row && r < NN);
  }

  // N elimination steps of one panel, a run-time loop (the unrolled form
  // does not fit the instruction cache at N = 64).  The block holding the
  // pivot columns shifts left by one column per step, so the pivot column
  // is always A[0] and registers are indexed statically; `left` tells
  // whether this thread holds that block.  `used` rows are not candidates.
  // Barriers: partial maxima published, pivot row published.
  auto eliminate = [&](bool left, double* pp) __attribute__((always_inline)) {
#pragma unroll 1
    for (int c = 0; c < N; ++c) {
      const int b = c & 1;
      if (left) {
        colc[b][r] = A[0];
        const double key = (row && !used) ? fabs(A[0]) : -1.0;
        double m = row_max(key);
        m = row_pair_max(m);
        const double m0 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(m), 0),
                                           __builtin_amdgcn_readlane(__double2loint(m), 0));
        const double m1 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(m), 32),
                                           __builtin_amdgcn_readlane(__double2loint(m), 32));
        m = fmax(m0, m1);
        const unsigned long long bal = __ballot(key == m);
        if ((tid & 63) == 0) {
          red_v[r >> 6] = m;
          red_i[r >> 6] = (r & ~63) + __ffsll((long long)bal) - 1;
        }
      }
      __syncthreads();
      double best = red_v[0];
      int p = red_i[0];
#pragma unroll
      for (int w = 1; w < NWH; ++w) {
        if (red_v[w] > best) {  // strictly larger: ties keep the earlier row
          best = red_v[w];
          p = red_i[w];
        }
      }
      if (row && r == p) {
        const int o = left ? 0 : N;
#pragma unroll
        for (int j = 0; j < N; ++j) prow[o + j] = A[j];
        if (h == 1) prow[2 * N] = rhs;
      }
      __syncthreads();
      const double piv = prow[0];
      const double inv = (piv != 0.0) ? 1.0 / (piv != 0.0 ? piv : 1.0) : 0.0;
      const double l = (row && !used && r != p) ? colc[b][r] * inv : 0.0;
      // Chunks of 8 columns, kept apart by scheduling barriers, so that at
      // most one chunk of the pivot row is live in registers.
      if (left) {
#pragma unroll
        for (int j0 = 0; j0 < N - 1; j0 += 8) {
#pragma unroll
          for (int j = j0; j < j0 + 8 && j < N - 1; ++j) A[j] = A[j + 1] - l * prow[j + 1];
          __builtin_amdgcn_sched_barrier(0);
        }
        A[N - 1] = 0.0;
      } else {
#pragma unroll
        for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
          for (int j = j0; j < j0 + 8 && j < N; ++j) A[j] -= l * prow[N + j];
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      if (h == 1) rhs -= l * prow[2 * N];
      used = used || (r == p);
      // U row c (left block with the diagonal as its reciprocal, right
      // block, rhs), stored by the first wavefront from the LDS copy.
      if (tid < 64) {
        for (int t = tid; t < N; t += 64) {
          pp[c * N + t] = (t < c) ? 0.0 : ((t == c) ? inv : prow[t - c]);
          pp[N * N + c * N + t] = prow[N + t];
        }
        if (tid == 0) pp[2 * N * N + c] = prow[2 * N];
      }
    }
  };

  // Rank of this row among the used rows (the interface row it takes).
  auto used_rank = [&]() {
    if (h == 0) uflag[r] = (row && used) ? 1 : 0;
    __syncthreads();
    int q = 0;
    for (int k = 0; k < r; ++k) q += uflag[k];
    __syncthreads();
    return q;
  };

  // Interior panels: continuity at the interface below layer ip
  // (:851-887, :1094-1118).
  for (int ip = 0; ip < M - 1; ++ip) {
    const int64_t tt = t0 + ip;
    const bool left = h == (ip & 1);
    layer_exp(tt, ecur);
    layer_exp(tt + 1, enxt);
    const int q = used_rank();  // includes the barrier that publishes ecur/enxt
    const bool fill = row && used;
    // Fills in chunks of 8 columns (scheduling barriers bound the loads in
    // flight, and so the registers, next to the live row).
    if (left) {
      if (fill) {
#pragma unroll
        for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
          for (int j = j0; j < j0 + 8 && j < N; ++j) A[j] = g_entry<NN>(a, tt, q, j) * ebot_of(ecur, j);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    } else if (fill) {
#pragma unroll
      for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
        for (int j = j0; j < j0 + 8 && j < N; ++j) A[j] = -g_entry<NN>(a, tt + 1, q, j) * etop_of(enxt, j);
        __builtin_amdgcn_sched_barrier(0);
      }
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) A[j] = 0.0;
    }
    if (h == 1 && fill) rhs = a.eig[eo.at(tt + 1, eo.ztop() + q)] - a.eig[eo.at(tt, eo.zbot() + q)];
    used = !row;
    eliminate(left, a.panel + tt * PS);
  }

  // Last panel: bottom boundary, Lambertian reflection of the downward field
  // (:909-959, :992-1091).
  {
    const int64_t tt = t0 + M - 1;
    const bool left = h == ((M - 1) & 1);
    const bool refl = !(lyrcut || diffusion);
    layer_exp(tt, ecur);
    for (int j = tid; j < N; j += blockDim.x) {
      double s = 0.0;
      for (int k = 0; k < NN; ++k) s += tb[to.cwt() + k] * tb[to.cmu() + k] * alb * g_entry<NN>(a, tt, NN - 1 - k, j);
      rsb[j] = refl ? s : 0.0;
    }
    double zrefl = 0.0;
    for (int k = 0; k < NN; ++k) zrefl += tb[to.cwt() + k] * tb[to.cmu() + k] * alb * a.eig[eo.at(tt, eo.zbot() + NN - 1 - k)];
    const int q = used_rank();
    const bool fill = row && used;
    const bool bot = fill && q < NN;
    const int tq = bot ? q : 0;
    if (left) {
      if (bot) {
#pragma unroll
        for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
          for (int j = j0; j < j0 + 8 && j < N; ++j)
            A[j] = (g_entry<NN>(a, tt, NN + tq, j) - 2.0 * rsb[j]) * ebot_of(ecur, j);
          __builtin_amdgcn_sched_barrier(0);
        }
      } else if (fill) {
#pragma unroll
        for (int j = 0; j < N; ++j) A[j] = 0.0;
      }
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) A[j] = 0.0;
    }
    if (h == 1) {
      const double btn = a.inst_d[i * kInstD + ID_BTRANS_NCUT];
      const double plb = a.inst_d[i * kInstD + ID_PL_BOT];
      const double plbd = a.inst_d[i * kInstD + ID_PL_BOT_DERIV];
      double extra = sc[DISORT_S_ISO_BOT];
      if (refl) extra += 2.0 * zrefl;
      if (F0 > 0.0 && !lyrcut) extra += alb * mu0 * F0 * btn / kPi;
      if (thermal) extra += (1.0 - alb) * plb;
      const double zb = -a.eig[eo.at(tt, eo.zbot() + NN + tq)];
      rhs = bot ? zb + (diffusion ? plb + tb[to.cmu() + tq] * plbd : extra) : (fill ? 0.0 : rhs);
    }
    used = !row || (fill && !bot);
    eliminate(left, a.panel + tt * PS);
  }

  // Back substitution, lane k = U row k of each panel, first wavefront only.
  if (tid >= 64) return;
  panel_back_substitute<N>(a.panel, t0, M, a.cvec);
}

// One-wavefront form of k_boundary_wg (launched for N <= 16; it works up to
// 3NN <= 64 panel rows, but at N = 32 the two-block form is faster): lane r holds
// both blocks of panel row r (Ac, the block of the pivot columns, shifted
// left by one column per step; An, the next layer's columns) and the
// right-hand side, so every synchronisation is wave-local.  Same pivots,
// operations and stored factors as k_boundary_wg.
template <int NN>
__global__ __launch_bounds__(64) void k_boundary_wg1(FluxArgs a) {
  constexpr int N = 2 * NN, R = 3 * NN;
  static_assert(R <= 64, "one panel row per lane");
  constexpr EigOff eo(N);
  constexpr TabOff to(N);
  constexpr int PS = panel_stride(N);
  __shared__ double prow[2 * N + 2];  // pivot row: Ac (shifted), An, rhs
  __shared__ double ecur[NN], enxt[NN], rsb[N];

  const int r = threadIdx.x;
  const bool row = r < R;
  const int bk = blockIdx.x;
  if (bk >= a.nb) return;
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

  auto layer_exp = [&](int64_t tt, double* e) {
    if (r < NN) e[r] = exp(-a.eig[eo.at(tt, eo.k() + r)] * a.lay_d[LD_DTAU_S * TL + lo + (tt - t0)]);
  };
  auto etop_of = [](const double* e, int j) { return j < NN ? e[NN - 1 - j] : 1.0; };
  auto ebot_of = [](const double* e, int j) { return j < NN ? 1.0 : e[j - NN]; };
  auto used_rank = [&](bool used) {
    const unsigned long long b = __ballot(row && used);
    return __popcll(b & ((1ull << r) - 1ull));
  };

  double Ac[N], An[N];
  double rhs = 0.0;
  bool used;

  // Top boundary rows 0..NN-1 (:890-907, :976-990), carried into panel 0.
  layer_exp(t0, ecur);
  wave_lds_order();
  {
    const bool tbc = row && r < NN;
    const double top = sc[DISORT_S_ISO_TOP] + a.inst_d[i * kInstD + ID_PL_TOP];
#pragma unroll
    for (int j = 0; j < N; ++j) An[j] = 0.0;
    if (tbc) {
#pragma unroll
      for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
        for (int j = j0; j < j0 + 8 && j < N; ++j) An[j] = g_entry<NN>(a, t0, NN - 1 - r, j) * etop_of(ecur, j);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    rhs = tbc ? top - a.eig[eo.at(t0, eo.ztop() + (tbc ? NN - 1 - r : 0))] : 0.0;
    used = !tbc;
  }

  // N elimination steps of one panel (run-time loop; the pivot column is
  // always Ac[0]).
  auto eliminate = [&](double* pp) __attribute__((always_inline)) {
#pragma unroll 1
    for (int c = 0; c < N; ++c) {
      const double key = (row && !used) ? fabs(Ac[0]) : -1.0;
      double mx = row_pair_max(row_max(key));
      const double m0 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(mx), 0),
                                         __builtin_amdgcn_readlane(__double2loint(mx), 0));
      const double m1 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(mx), 32),
                                         __builtin_amdgcn_readlane(__double2loint(mx), 32));
      mx = fmax(m0, m1);
      const int p = __ffsll((long long)__ballot(key == mx)) - 1;
      wave_lds_order();
      if (row && r == p) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
          prow[j] = Ac[j];
          prow[N + j] = An[j];
        }
        prow[2 * N] = rhs;
      }
      wave_lds_order();
      const double piv = prow[0];
      const double inv = (piv != 0.0) ? 1.0 / (piv != 0.0 ? piv : 1.0) : 0.0;
      const double l = (row && !used && r != p) ? Ac[0] * inv : 0.0;
#pragma unroll
      for (int j0 = 0; j0 < N - 1; j0 += 8) {
#pragma unroll
        for (int j = j0; j < j0 + 8 && j < N - 1; ++j) Ac[j] = Ac[j + 1] - l * prow[j + 1];
        __builtin_amdgcn_sched_barrier(0);
      }
      Ac[N - 1] = 0.0;
#pragma unroll
      for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
        for (int j = j0; j < j0 + 8 && j < N; ++j) An[j] -= l * prow[N + j];
        __builtin_amdgcn_sched_barrier(0);
      }
      rhs -= l * prow[2 * N];
      used = used || (r == p);
      // U row c: left block with the diagonal as its reciprocal, right block, rhs.
      for (int t = r; t < N; t += 64) {
        pp[c * N + t] = (t < c) ? 0.0 : ((t == c) ? inv : prow[t - c]);
        pp[N * N + c * N + t] = prow[N + t];
      }
      if (r == 0) pp[2 * N * N + c] = prow[2 * N];
    }
  };

  // Interior panels: continuity at the interface below layer ip
  // (:851-887, :1094-1118).
  for (int ip = 0; ip < M - 1; ++ip) {
    const int64_t tt = t0 + ip;
    layer_exp(tt, ecur);
    layer_exp(tt + 1, enxt);
    wave_lds_order();
    const int q = used_rank(used);
    const bool fill = row && used;
    if (fill) {
#pragma unroll
      for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
        for (int j = j0; j < j0 + 8 && j < N; ++j) {
          Ac[j] = g_entry<NN>(a, tt, q, j) * ebot_of(ecur, j);
          An[j] = -g_entry<NN>(a, tt + 1, q, j) * etop_of(enxt, j);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      rhs = a.eig[eo.at(tt + 1, eo.ztop() + q)] - a.eig[eo.at(tt, eo.zbot() + q)];
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) {
        Ac[j] = An[j];
        An[j] = 0.0;
      }
    }
    used = !row;
    eliminate(a.panel + tt * PS);
  }

  // Last panel: bottom boundary, Lambertian reflection of the downward field
  // (:909-959, :992-1091).
  {
    const int64_t tt = t0 + M - 1;
    const bool refl = !(lyrcut || diffusion);
    layer_exp(tt, ecur);
    if (r < N) {
      double s = 0.0;
      for (int k = 0; k < NN; ++k) s += tb[to.cwt() + k] * tb[to.cmu() + k] * alb * g_entry<NN>(a, tt, NN - 1 - k, r);
      rsb[r] = refl ? s : 0.0;
    }
    wave_lds_order();
    double zrefl = 0.0;
    for (int k = 0; k < NN; ++k) zrefl += tb[to.cwt() + k] * tb[to.cmu() + k] * alb * a.eig[eo.at(tt, eo.zbot() + NN - 1 - k)];
    const int q = used_rank(used);
    const bool fill = row && used;
    const bool bot = fill && q < NN;
    const int tq = bot ? q : 0;
    if (bot) {
#pragma unroll
      for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
        for (int j = j0; j < j0 + 8 && j < N; ++j)
          Ac[j] = (g_entry<NN>(a, tt, NN + tq, j) - 2.0 * rsb[j]) * ebot_of(ecur, j);
        __builtin_amdgcn_sched_barrier(0);
      }
    } else if (fill) {
#pragma unroll
      for (int j = 0; j < N; ++j) Ac[j] = 0.0;
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) Ac[j] = An[j];
    }
#pragma unroll
    for (int j = 0; j < N; ++j) An[j] = 0.0;
    {
      const double btn = a.inst_d[i * kInstD + ID_BTRANS_NCUT];
      const double plb = a.inst_d[i * kInstD + ID_PL_BOT];
      const double plbd = a.inst_d[i * kInstD + ID_PL_BOT_DERIV];
      double extra = sc[DISORT_S_ISO_BOT];
      if (refl) extra += 2.0 * zrefl;
      if (F0 > 0.0 && !lyrcut) extra += alb * mu0 * F0 * btn / kPi;
      if (thermal) extra += (1.0 - alb) * plb;
      const double zb = -a.eig[eo.at(tt, eo.zbot() + NN + tq)];
      rhs = bot ? zb + (diffusion ? plb + tb[to.cmu() + tq] * plbd : extra) : (fill ? 0.0 : rhs);
    }
    used = !row || (fill && !bot);
    eliminate(a.panel + tt * PS);
  }

  // Back substitution, lane k = U row k of each panel.
  panel_back_substitute<N>(a.panel, t0, M, a.cvec);
}

template <int NN>
void launch_wide(const FluxArgs& a, hipStream_t s, hipEvent_t* ev) {
  using C = WideCfg<NN>;
  constexpr int N = 2 * NN;
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_column<N>, dim3(a.nb), dim3(64), 0, s, a);
  if (ev) (void)hipEventRecord(ev[1], s);
  hipLaunchKernelGGL(k_layer_wg<NN>, dim3((unsigned)a.n_tasks), dim3(C::LT), 0, s, a);
  if (ev) (void)hipEventRecord(ev[2], s);
  if constexpr (NN <= 8) hipLaunchKernelGGL(k_boundary_wg1<NN>, dim3((unsigned)a.nb), dim3(64), 0, s, a);  // faster there; not at N = 32
  else hipLaunchKernelGGL(k_boundary_wg<NN>, dim3((unsigned)a.nb), dim3(C::BT), 0, s, a);
  if (ev) (void)hipEventRecord(ev[3], s);
  launch_flux_levels<NN>(a, s);
  if (ev) (void)hipEventRecord(ev[4], s);
}

}  // namespace

bool launch_flux_bucket_wg(int nstr, const FluxArgs& a, hipStream_t s, hipEvent_t* ev) {
  switch (nstr) {
    case 4: launch_wide<2>(a, s, ev); return true;
    case 6: launch_wide<3>(a, s, ev); return true;
    case 8: launch_wide<4>(a, s, ev); return true;
    case 10: launch_wide<5>(a, s, ev); return true;
    case 12: launch_wide<6>(a, s, ev); return true;
    case 14: launch_wide<7>(a, s, ev); return true;
    case 16: launch_wide<8>(a, s, ev); return true;
    case 32: launch_wide<16>(a, s, ev); return true;
    case 64: launch_wide<32>(a, s, ev); return true;
    default: return false;
  }
}

}  // namespace dsa
