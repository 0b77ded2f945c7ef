// gfx950 kernels of the general solver (DisortSolver::solve,
// src/DisortSolver.cpp:21-399) for a batch of columns sharing one shape.
//
// All azimuthal modes of a chunk of columns are solved together (the modes
// are independent; the reference's sequential mode loop only adds the
// early-exit test, applied here when the modes are summed):
//   k_column        per column: Delta-M scaling, cut-off, beam transmission,
//                   Chapman factors, Planck (shared with the flux path)
//   k_glevels       per column: layer and scaled depth of each user level
//   k_gsurf_c/_e/_b BRDF Fourier coefficients of all modes per (RPV parameter
//                   set, entry), emissivities, the beam column per column
//   k_glayer        per (column, mode, layer) wavefront: phase matrix at
//                   mode m, reduced eigenproblem (parallel Jacobi in LDS),
//                   beam and thermal particular solutions, eigenvectors and
//                   sources interpolated to the user angles
//   k_gboundary1    per (column, mode) wavefront (3NN <= 64 panel rows) or
//   k_gboundary     workgroup: panel LU of the boundary problem with the
//                   mode's surface coupling
//   k_gflux         per (column, level): fluxes at the user levels (m = 0)
//   k_guser         per (column, mode) workgroup: intensity components at the
//                   user angles and levels
//   k_gaccum        per column: Fourier sum over the modes (with the
//                   accuracy_fourier_series cut), azimuthal average
//   k_gnt, k_gbe_prep/k_gbe, k_gbrdf
//                   Nakajima-Tanaka or Buras-Emde, and BRDF beam-reflection
//                   corrections
#include "general_layout.hpp"
#include "wg_common.hpp"

namespace dsa {
namespace {

constexpr int kNMUG = 200;

// Entry G(q, j) of a task record (reference ordering).
template <int NN>
__device__ __forceinline__ double ge_G(const double* e, int q, int j) {
  constexpr GE ge(2 * NN);
  const bool up = j >= NN;
  const int jp = up ? j - NN : NN - 1 - j;
  const bool lower = q >= NN;
  const int r = lower ? q - NN : NN - 1 - q;
  const int f = (lower == up) ? ge.gp() : ge.gm();
  const double v = e[f + r + jp * NN];
  return up ? v : -v;
}

// RPV bidirectional reflectivity (src/DisortSolver.cpp:2734-2780, badmu = 0).
__device__ double rpv_rho(const double* p, double mu1, double mu2, double phi) {
  const double rho0 = p[0], k = p[1], theta = p[2], sigma = p[3], t1 = p[4], t2 = p[5], scale = p[6];
  phi = kPi - phi;
  if (phi == 1.0e-4 && mu1 == mu2) {
    const double hs = rho0 * (pow(2.0 * mu1 * mu1 * mu1, k - 1.0) * (1.0 - theta) / ((1.0 + theta) * (1.0 + theta)) *
                                  (2.0 - rho0) +
                              sigma / mu1) *
                      (t1 * exp(kPi * t2) + 1.0);
    return hs * scale;
  }
  const double m = pow(mu1 * mu2 * (mu1 + mu2), k - 1.0);
  const double cosphi = cos(phi);
  const double sin1 = sqrt(1.0 - mu1 * mu1), sin2 = sqrt(1.0 - mu2 * mu2);
  const double cosg = mu1 * mu2 + sin1 * sin2 * cosphi;
  const double g = acos(cosg);
  const double f = (1.0 - theta * theta) / pow(1.0 + 2.0 * theta * cosg + theta * theta, 1.5);
  const double tan1 = sin1 / mu1, tan2 = sin2 / mu2;
  const double capg = sqrt(tan1 * tan1 + tan2 * tan2 - 2.0 * tan1 * tan2 * cosphi);
  const double h = 1.0 + (1.0 - rho0) / (1.0 + capg);
  const double t = 1.0 + t1 * exp(t2 * (kPi - g));
  const double ans = rho0 * (m * f * h + sigma / mu1) * t * scale;
  return ans < 0.0 ? 0.0 : ans;
}

// Hapke (src/DisortSolver.cpp:2710-2731): p = {b0, hh, w}.
__device__ double hapke_rho(const double* p, double mu, double mup, double dphi) {
  const double b0 = p[0], hh = p[1], w = p[2];
  const double ct = mu * mup + sqrt((1.0 - mu * mu) * (1.0 - mup * mup)) * cos(dphi);
  const double b = b0 * hh / (hh + tan(0.5 * acos(ct)));
  const double gam = sqrt(1.0 - w);
  const double h0 = (1.0 + 2.0 * mup) / (1.0 + 2.0 * gam * mup);
  const double h = (1.0 + 2.0 * mu) / (1.0 + 2.0 * gam * mu);
  return 0.25 * w * ((1.0 + b) * (1.0 + 0.5 * ct) + h0 * h - 1.0) / (mu + mup);
}

// Ross-thick + Li-sparse-reciprocal kernels (src/DisortSolver.cpp:2785-2840):
// p = {iso, vol, geo}; crown height/base 2, branch/radius 1, hot spot 1.5 deg.
__device__ double rossli_rho(const double* p, double mu, double mup, double dphi) {
  const double a0 = 1.5 * kPi / 180.0;
  const double si = sqrt(1.0 - mup * mup), sr = sqrt(1.0 - mu * mu);
  const double ti = mup > 0.0 ? si / mup : 0.0, tr = mu > 0.0 ? sr / mu : 0.0;
  const double cd = cos(dphi);
  double ca = mup * mu + si * sr * cd;
  ca = ca < -1.0 ? -1.0 : (ca > 1.0 ? 1.0 : ca);
  const double al = acos(ca);
  const double kvol =
      4.0 / (3.0 * kPi) * (1.0 / (mup + mu)) * ((kPi / 2.0 - al) * ca + sin(al)) * (1.0 + 1.0 / (1.0 + al / a0)) -
      1.0 / 3.0;
  // Branch-to-radius ratio 1: the scaled tangents are the tangents.
  const double ci1 = 1.0 / sqrt(1.0 + ti * ti), cr1 = 1.0 / sqrt(1.0 + tr * tr);
  const double si1 = ti * ci1, sr1 = tr * cr1;
  const double ca1 = ci1 * cr1 + si1 * sr1 * cd;
  double g2 = ti * ti + tr * tr - 2.0 * ti * tr * cd;
  if (g2 < 0.0) g2 = 0.0;
  const double x = ti * tr * sin(dphi);
  const double cost = 2.0 * (ci1 * cr1) / (ci1 + cr1) * sqrt(g2 + x * x);
  const double t = (cost >= -1.0 && cost <= 1.0) ? acos(cost) : 0.0;
  const double kgeo =
      (ci1 + cr1) / (kPi * ci1 * cr1) * (t - sin(t) * cos(t) - kPi) + (1.0 + ca1) / (2.0 * ci1 * cr1);
  const double r = p[0] + p[1] * kvol + p[2] * kgeo;
  return r > 0.0 ? r : 0.0;
}

// Tsang shadowing term (src/DisortSolver.cpp:2842-2856).
__device__ double shadow_eta(double ct, double s2) {
  const double st = sqrt(1.0 - ct * ct);
  if (st < 1e-15) return 0.0;
  const double cot = ct / st;
  return 0.5 * (sqrt(s2 / kPi) / cot * exp(-cot * cot / s2) - erfc(cot / sqrt(s2)));
}

// Cox-Munk ocean surface (src/DisortSolver.cpp:2858-2900): p = {u10,
// refractive index, shadowing 1 / 0}.
__device__ double coxmunk_rho(const double* p, double mu, double mup, double dphi) {
  const double si = sqrt(1.0 - mup * mup), sr = sqrt(1.0 - mu * mu);
  const double ct = -mup * mu + si * sr * cos(dphi);
  const double omc = 1.0 - ct;
  if (omc < 1e-15) return 0.0;
  const double mn2 = (mup + mu) * (mup + mu) / (2.0 * omc);
  if (mn2 < 1e-15) return 0.0;
  const double s2 = 0.003 + 0.00512 * p[0];
  const double pslope = 1.0 / (kPi * s2) * exp(-(1.0 - mn2) / (s2 * mn2));
  const double nt = p[1];
  const double cli = sqrt(0.5 * (1.0 + ct));
  const double q = 1.0 - cli * cli;
  const double sli = sqrt(q > 0.0 ? q : 0.0);
  const double slt = sli / nt;
  if (slt >= 1.0) return 0.0;
  const double clt = sqrt(1.0 - slt * slt);
  const double rs = (cli - nt * clt) / (cli + nt * clt);
  const double rp = (nt * cli - clt) / (clt + nt * cli);
  double r = (pslope * 0.5 * (rs * rs + rp * rp)) / (4.0 * mup * mu * mn2 * mn2);
  if (p[2] != 0.0) r *= 1.0 / (shadow_eta(mup, s2) + shadow_eta(mu, s2) + 1.0);
  return r > 0.0 ? r : 0.0;
}

// Bidirectional reflectivity of the batch's BRDF kind (bidirReflectivity,
// src/DisortSolver.cpp:2902-2936) with parameter block p.
__device__ double brdf_rho(int kind, const double* p, double mu, double mup, double dphi) {
  switch (kind) {
    case kBrdfHapke: return hapke_rho(p, mu, mup, dphi);
    case kBrdfCoxMunk: return coxmunk_rho(p, mu, mup, dphi);
    case kBrdfRossLi: return rossli_rho(p, mu, mup, dphi);
    default: return rpv_rho(p, mu, mup, dphi);
  }
}

// ===========================================================================
// k_glevels: user levels (src/DisortSolver.cpp:803-820), thread per column.
// ===========================================================================
__global__ __launch_bounds__(64) void k_glevels(GenArgs a) {
  const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (i >= a.n) return;
  const int L = a.nlyr;
  const int64_t lo = i * L, vo = i * (L + 1), TL = a.n * L, TV = a.n * (L + 1);
  const bool user_tau = (a.flags & DISORT_G_USER_TAU) != 0;
  const double* tauc = a.lev_d + VD_TAUC * TV + vo;
  const double* taus = a.lev_d + VD_TAU_S * TV + vo;
  for (int lu = 0; lu < a.ntau; ++lu) {
    const double ut = user_tau ? a.tau_user[i * a.ntau + lu] : tauc[lu];
    int lc;
    for (lc = 0; lc < L - 1; ++lc)
      if (ut >= tauc[lc] && ut <= tauc[lc + 1]) break;
    double s = ut;
    if (a.lay_d[LD_SIG2 * TL + lo + lc] >= 0.0)
      s = taus[lc] + (1.0 - a.ssa[lo + lc] * a.lay_d[LD_F * TL + lo + lc]) * (ut - tauc[lc]);
    a.u_lyr[i * a.ntau + lu] = lc;
    a.u_tau[i * a.ntau + lu] = s;
  }
}

// ===========================================================================
// Diffuse BRDF Fourier coefficients and directional emissivities of every
// RPV parameter set (src/DisortSolver.cpp:2604-2707).
//
// k_gsurf_c: one thread per (parameter set, coefficient entry); the entry's
// reflectivity rho(mu, mu', pi x_k) is evaluated once per azimuth node and
// enters the quadrature sums of all modes (each sum in the order of the
// reference's per-mode loop), so a set costs (NN + numu) NN x 100
// reflectivities instead of that times the mode count.
// k_gsurf_e: one block per (parameter set, angle with mu > 0), thread per
// outer azimuth node, the 200 x 100 emissivity double sum.  The emissivity
// enters only through the bottom thermal emission (multiplied by the
// bottom Planck function, zero without thermal emission), so it is computed
// only with DISORT_F_THERMAL; otherwise the slots hold 0.
// ===========================================================================
template <int NN>
__global__ __launch_bounds__(256) void k_gsurf_c(GenArgs a) {
  constexpr int N = 2 * NN, K = kNMUG / 2;
  __shared__ double cmk[N * K];  // cos(m pi x_k), m < N (1 at m = 0)
  const int nm = a.nm;
  const int numu = a.numu;
  const int ne = (NN + numu) * NN;  // coefficient entries per set and mode
  const int64_t sz = surf_d_size(NN, numu);
  const double* gmu = a.gaz;
  const double* gwt = a.gaz + kNMUG;
  for (int e = threadIdx.x; e < nm * K; e += blockDim.x) {
    const int m = e / K, k = e % K;
    cmk[e] = (m == 0) ? 1.0 : cos((double)m * kPi * gmu[k]);
  }
  __syncthreads();
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)a.nsets * ne) return;
  const int64_t s = t / ne;
  const int e = (int)(t % ne);
  const double* p = a.set_rpv + s * DISORT_NRPV;
  const double* cmu = a.tab;  // TabOff::cmu() == 0
  const bool do_user = (a.flags & DISORT_G_USER_MU) && !(a.flags & DISORT_G_ONLY_FLUXES);
  const double mup = cmu[e % NN];
  double mu;
  if (e < NN * NN) mu = cmu[e / NN];
  else mu = do_user ? a.mu_user[(e - NN * NN) / NN] : 0.0;
  const bool act = mu > 0.0;
  double sm[N];
#pragma unroll
  for (int m = 0; m < N; ++m) sm[m] = 0.0;
  if (act)
    for (int k = 0; k < K; ++k) {
      const double x = gwt[k] * brdf_rho(a.brdf_kind, p, mu, mup, kPi * gmu[k]);
#pragma unroll
      for (int m = 0; m < N; ++m)
        if (m < nm) sm[m] += x * cmk[m * K + k];
    }
  double* d = a.srf_d + s * nm * sz + e;
#pragma unroll
  for (int m = 0; m < N; ++m)
    if (m < nm) d[m * sz] = act ? (m == 0 ? 1.0 : 2.0) * sm[m] : 0.0;
  // Emissivity slots (written here for sets without thermal emission, and
  // for modes m > 0 and angles mu <= 0 in any case).
  const int ee = e;
  if (ee < NN + numu) {
    double mue = 0.0;
    if (ee < NN) mue = cmu[ee];
    else if (do_user) mue = a.mu_user[ee - NN];
    const bool computed = (a.flags & DISORT_F_THERMAL) && mue > 0.0;
    double* eq = a.srf_d + s * nm * sz + NN * NN + numu * NN + ee;
    for (int m = 0; m < nm; ++m)
      if (m > 0 || !computed) eq[m * sz] = 0.0;
  }
}

template <int NN>
__global__ __launch_bounds__(256) void k_gsurf_e(GenArgs a) {
  __shared__ double red[4];
  const int numu = a.numu;
  const int64_t s = blockIdx.x / (NN + numu);
  const int e = (int)(blockIdx.x % (NN + numu));
  const bool do_user = (a.flags & DISORT_G_USER_MU) && !(a.flags & DISORT_G_ONLY_FLUXES);
  double mu = 0.0;
  if (e < NN) mu = a.tab[e];
  else if (do_user) mu = a.mu_user[e - NN];
  if (!(mu > 0.0)) return;  // block-uniform
  const double* p = a.set_rpv + s * DISORT_NRPV;
  const double* gmu = a.gaz;
  const double* gwt = a.gaz + kNMUG;
  double part = 0.0;
  for (int jg = threadIdx.x; jg < kNMUG; jg += blockDim.x) {
    double sm = 0.0;
    for (int k = 0; k < kNMUG / 2; ++k) sm += gwt[k] * gmu[k] * brdf_rho(a.brdf_kind, p, mu, gmu[k], kPi * gmu[jg]);
    part += gwt[jg] * sm;
  }
  const double tot = blk_sum(part, red);
  if (threadIdx.x == 0) a.srf_d[s * a.nm * surf_d_size(NN, numu) + NN * NN + numu * NN + e] = 1.0 - tot;
}

// Beam column of the BRDF coefficients of every mode, block per chunk
// column: each reflectivity value rho(mu, mu0, phi_k) is evaluated once and
// enters all the modes' quadrature sums (in the order of the reference's per-mode loop, with
// the same cos(m pi x_k) table), rather than once per mode.
template <int NN>
__global__ __launch_bounds__(64) void k_gsurf_b(GenArgs a) {
  constexpr int N = 2 * NN, K = kNMUG / 2;
  __shared__ double cmk[N * K];  // cos(m pi x_k), m < N (1 at m = 0)
  const int64_t cl = blockIdx.x;
  if (cl >= a.nc) return;
  const int nm = a.nm;
  const int64_t i = a.c0 + cl;
  const double mu0 = a.scal[i * DISORT_NSCAL + DISORT_S_BEAM_MU];
  const bool beam = a.scal[i * DISORT_NSCAL + DISORT_S_BEAM_FLUX] > 0.0;
  const double* p = a.set_rpv + (int64_t)a.rpv_set[i] * DISORT_NRPV;
  const bool do_user = (a.flags & DISORT_G_USER_MU) && !(a.flags & DISORT_G_ONLY_FLUXES);
  const double* gmu = a.gaz;
  const double* gwt = a.gaz + kNMUG;
  for (int e = threadIdx.x; e < nm * K; e += blockDim.x) {
    const int m = e / K, k = e % K;
    cmk[e] = (m == 0) ? 1.0 : cos((double)m * kPi * gmu[k]);
  }
  __syncthreads();
  const int64_t bs = surf_b_size(NN, a.numu);
  double* b0 = a.srf_b + cl * nm * bs;
  for (int e = threadIdx.x; e < NN + a.numu; e += blockDim.x) {
    double mu = 0.0;
    if (e < NN) mu = a.tab[e];
    else if (do_user) mu = a.mu_user[e - NN];
    const bool act = beam && mu > 0.0;
    double sm[N];
#pragma unroll
    for (int m = 0; m < N; ++m) sm[m] = 0.0;
    if (act) {
      for (int k = 0; k < K; ++k) {
        const double x = gwt[k] * brdf_rho(a.brdf_kind, p, mu, mu0, kPi * gmu[k]);
#pragma unroll
        for (int m = 0; m < N; ++m)
          if (m < nm) sm[m] += x * cmk[m * K + k];
      }
    }
#pragma unroll
    for (int m = 0; m < N; ++m)
      if (m < nm) b0[m * bs + e] = act ? (m == 0 ? 1.0 : 2.0) * sm[m] : 0.0;
  }
}

// Row and column (r >= c) of entry e of a packed lower triangle.
__device__ __forceinline__ void tri_rc(int e, int& r, int& c) {
  r = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
  if ((r + 1) * (r + 2) / 2 <= e) ++r;
  if (r * (r + 1) / 2 > e) --r;
  c = e - r * (r + 1) / 2;
}

// ===========================================================================
// k_glayer: one workgroup per (column, mode, layer) task.
// ===========================================================================
// One wavefront per task: block-wide barriers are then wave-local and cost
// next to nothing, and the whole task stays within one SIMD.
constexpr int kGLT = 64;

template <int NN>
__global__ __launch_bounds__(kGLT) void k_glayer(GenArgs a) {
  using C = WideCfg<NN>;
  constexpr int N = C::N, LD = C::LD, NNE = NN + (NN & 1), NP = NNE / 2;
  constexpr TabOff to(N);
  constexpr GE ge(N);
  constexpr bool OW = kGLT == 64;  // one wavefront per block
  // LDS holds the dense NN x NN work matrices and short vectors; the Y_l^m
  // table (shared by every task of the mode) is read through the caches, to
  // keep the footprint small enough for more wavefronts per CU.
  // Three NN x NN slots: A_s, then S (formed once T = -A_s L is held in the
  // third slot), then U = L^-T V, then the Cholesky factor of -A_s (rebuilt
  // when a thermal or spherical source needs it) | -B_s, then L | T, then
  // V.  For the interpolation to the user angles: W, a block of its columns
  // at a time | G+ | G-.
  __shared__ double sbuf[3 * NN * LD];
  double* const sA = sbuf;
  double* const sS = sbuf;
  double* const sL = sbuf + NN * LD;
  double* const sV = sbuf + 2 * NN * LD;
  __shared__ double chi[N], p0[N];
  __shared__ double qn[NN];  // sqrt(w_r / mu_r)
  __shared__ double v0[NN], v1[NN], v2[NN], v3[NN], lam[NN], dv[NN];
  __shared__ double zr_u[NN], zr_d[NN], z0_u[NN], z0_d[NN], z1_h[NN];
  __shared__ double psi[3 * N];
  __shared__ double rc[NP], rs[NP], rt[NP], rapp[NP], raqq[NP], rapq[NP];
  __shared__ int rp[NP], rq[NP];
  __shared__ double red[kGLT / 64];
  __shared__ int flag;

  const int tid = threadIdx.x;
  const int64_t t = blockIdx.x;
  if (t >= a.n_tasks) return;
  const int L = a.nlyr;
  const int lc = (int)(t % L);
  const int64_t cm = t / L;
  const int m = (int)(cm % a.nm);
  const int64_t i = a.c0 + cm / a.nm;
  // A layer below the absorption cut-off enters no output: the boundary
  // problem spans the ncut layers above it, and fluxes and intensities at
  // levels below it are zero.  Its task ends here (ncut is nlyr when the
  // cut-off is not in effect).
  if (lc >= a.inst_i[i * kInstI + II_NCUT]) return;
  const int64_t g = i * L + lc, v = i * (L + 1) + lc;
  const int P = a.pstride;
  const double* pm = a.pmom + g * P;
  const double* sc = a.scal + i * DISORT_NSCAL;
  const int64_t TL = a.n * L, TV = a.n * (L + 1);
  const double* tb = a.tab;
  const double* Pm = a.ylm + (int64_t)m * N * NN;  // Y_l^m(mu_r), [l][r]
  const double ws = a.lay_d[LD_SSA_S * TL + g];
  const double f = a.lay_d[LD_F * TL + g];
  const double sig2 = a.lay_d[LD_SIG2 * TL + g];
  const double F0 = sc[DISORT_S_BEAM_FLUX], mu0 = sc[DISORT_S_BEAM_MU];
  const bool beam = F0 > 0.0;
  const double fm = (m == 0) ? 1.0 : 2.0;  // 2 - delta_m0
  double* E = a.eig + t * ge.size();

  // Scaled phase coefficients (src/DisortSolver.cpp:667-738); P_0 is 1.
  for (int l = tid; l < N; l += blockDim.x) {
    const double pl = (l == 0) ? 1.0 : pm[l];
    double c;
    if (sig2 < 0.0) c = (2.0 * l + 1.0) * ws * pl;
    else if (sig2 == 0.0) c = (2.0 * l + 1.0) * ws * (pl - f) / (1.0 - f);
    else c = (2.0 * l + 1.0) * ws * (pl - f * exp(-(double)(l * l) / (2.0 * sig2))) / (1.0 - f);
    chi[l] = c;
  }
  for (int r = tid; r < NN; r += blockDim.x) {
    qn[r] = tb[to.sd() + r] / tb[to.cmu() + r];
  }
  // Y_l^m(-mu0) by the recurrence of src/Legendre.cpp:67-120 (the diagonal
  // term as the product the reference builds over successive modes).
  if (tid == 0) {
    const double x = -mu0;
    for (int l = 0; l < N; ++l) p0[l] = 0.0;
    if (m == 0) {
      p0[0] = 1.0;
      if (N > 1) p0[1] = x;
      for (int l = 2; l < N; ++l) p0[l] = ((2.0 * l - 1.0) * x * p0[l - 1] - (l - 1.0) * p0[l - 2]) / l;
    } else {
      double y = 1.0;
      for (int k = 1; k <= m; ++k) y = -sqrt((1.0 - 1.0 / (2.0 * k)) * (1.0 - x * x)) * y;
      p0[m] = y;
      if (m + 1 < N) p0[m + 1] = sqrt(2.0 * m + 1.0) * x * y;
      for (int l = m + 2; l < N; ++l) {
        const double aa = sqrt((double)((l - m) * (l + m)));
        const double bb = sqrt((double)((l - m - 1) * (l + m - 1)));
        p0[l] = ((2.0 * l - 1.0) * x * p0[l - 1] - bb * p0[l - 2]) / aa;
      }
    }
  }
  __syncthreads();

  // Pseudo-spherical beam: xba = 1/CH with its dithers (src/DisortSolver.cpp:
  // 1442-1456); the Frobenius-norm dither (:1486-1497) follows below.
  const bool spherical = beam && (a.flags & DISORT_F_SPHERICAL) != 0;
  const double taut = a.lev_d[VD_TAU_S * TV + v];
  const double taub = a.lev_d[VD_TAU_S * TV + v + 1];
  double xba = 0.0;
  if (spherical) {
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
  }

  // A_s and -B_s at mode m from the sums over l >= m split by the parity of
  // l - m, e(r,c) = sum_even chi Y_r Y_c and o(r,c) = sum_odd: with
  // q_r = sqrt(w_r/mu_r), A_s = q_r q_c e - 1/mu_r delta, -B_s = 1/mu_r delta
  // - q_r q_c o.  The phase matrix C has the blocks w_c (e +/- o) / 2, from
  // which the Frobenius norm of I + xba diag(mu) - C is accumulated.
  // One lane per entry r >= c of the lower triangle, stored at (r, c) and
  // (c, r); the norm takes both entries (column weights w_c and w_r).
  double fro = 0.0;
  for (int e = tid; e < NN * (NN + 1) / 2; e += blockDim.x) {
    int r, c;
    tri_rc(e, r, c);
    double se = 0.0, so = 0.0;
    for (int l = m; l < N; ++l) {
      const double q = chi[l] * Pm[l * NN + r] * Pm[l * NN + c];
      if (((l - m) & 1) == 0) se += q;
      else so += q;
    }
    const double qq = qn[r] * qn[c];
    const double dg = (r == c) ? 1.0 / tb[to.cmu() + r] : 0.0;
    const double av = qq * se - dg, bv = dg - qq * so;
    sA[r * LD + c] = av;
    sA[c * LD + r] = av;
    sL[r * LD + c] = bv;
    sL[c * LD + r] = bv;
    if (spherical) {
      for (int side = 0; side < (r == c ? 1 : 2); ++side) {
        const int c0 = side ? r : c;
        const double hw = 0.5 * tb[to.cwt() + c0];
        const double tl = hw * (se + so), tr = hw * (se - so);
        const double du = (r == c) ? 1.0 + xba * tb[to.cmu() + r] : 0.0;
        const double dd = (r == c) ? 1.0 - xba * tb[to.cmu() + r] : 0.0;
        fro += (du - tl) * (du - tl) + (dd - tl) * (dd - tl) + 2.0 * tr * tr;
      }
    }
  }
  __syncthreads();
  if (spherical) {
    fro = blk_sum(fro, red);
    if (1.0 / sqrt(fro) < 1.0e-4) {
      if (xba == 0.0) xba = 5.0e-9;
      xba *= 1.00000005;
    }
  }
  if (!blk_cholesky<NN, OW>(sL, LD, &flag) && tid == 0) atomicOr(&a.status[i], DISORT_ST_EIGEN);

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

  // Parallel cyclic Jacobi (round-robin pairs), as the flux kernel.
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
        } else {
          rc[k] = 1.0;
          rs[k] = 0.0;
          rt[k] = 0.0;
        }
      }
      blk_sync<OW>();
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
          const double upp = c1 * bpp - s1 * bqp, upq = c1 * bpq - s1 * bqq;
          const double uqp = s1 * bpp + c1 * bqp, uqq = s1 * bpq + c1 * bqq;
          const double npp = c2 * upp - s2 * upq, npq = s2 * upp + c2 * upq;
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
        const int mm = e / NP, k = e % NP;
        const int p = rp[k], q = rq[k];
        if (q >= NN) continue;
        const double vp = sV[mm * LD + p], vq = sV[mm * LD + q];
        sV[mm * LD + p] = rc[k] * vp - rs[k] * vq;
        sV[mm * LD + q] = rs[k] * vp + rc[k] * vq;
      }
      blk_sync<OW>();
    }
  }
  if (tid < NN) {
    lam[tid] = sS[tid * LD + tid];
    E[ge.k() + tid] = sqrt(fabs(lam[tid]));
  }
  __syncthreads();

  // G+/- (src/DisortSolver.cpp:1110-1138).
  blk_ltsolve_cols<NN>(sL, LD, sV, sS);  // U = L^-T V
  __syncthreads();
#pragma unroll 1
  for (int e = tid; e < NN * NN; e += blockDim.x) {
    const int r = e % NN, j = e / NN;
    double y = 0.0;
#pragma unroll 1
    for (int mm = 0; mm <= r; ++mm) y += sL[r * LD + mm] * sV[mm * LD + j];
    const double sdi = tb[to.sdi() + r];
    const double uu = sdi * sS[r * LD + j];
    const double tt = -sdi * y / sqrt(fabs(lam[j]));
    const double gp = 0.5 * (tt + uu), gmv = 0.5 * (tt - uu);
    E[ge.gp() + r + j * NN] = gp;
    E[ge.gm() + r + j * NN] = gmv;
  }

  const bool thermal = (a.flags & DISORT_F_THERMAL) != 0 && m == 0;

  if (thermal || spherical) {
    // -A_s again (the slot held S and U meanwhile), entry by entry as built
    // above, then its Cholesky factor.
    __syncthreads();
    for (int e = tid; e < NN * (NN + 1) / 2; e += blockDim.x) {
      int r, c;
      tri_rc(e, r, c);
      double se = 0.0;
      for (int l = m; l < N; l += 2) se += chi[l] * Pm[l * NN + r] * Pm[l * NN + c];
      const double dg = (r == c) ? 1.0 / tb[to.cmu() + r] : 0.0;
      const double v = -(qn[r] * qn[c] * se - dg);
      sA[r * LD + c] = v;
      sA[c * LD + r] = v;
    }
    __syncthreads();
    blk_cholesky<NN, OW>(sA, LD, &flag);
  }

  // (I + xba diag(mu) - C) z = b through the eigen-decomposition, as in the
  // flux workgroup kernel (bs = b_up + b_dn, bd = b_up - b_dn per node r).
  auto solve_ab = [&](double bs, double bd, bool xba_zero, double mub, double& zu, double& zd) {
    double pr = 0.0, qr = 0.0;
    if (tid < NN) {
      const double cmr = tb[to.cmu() + tid];
      pr = -bs / cmr;
      qr = -bd / cmr;
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
      if (tid < NN) v0[tid] = tb[to.sd() + tid] * qr;
      __syncthreads();
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
      blk_lmul<NN, true, OW>(sL, LD, v2, v1);
      blk_vdv<NN, OW>(sV, LD, dv, v2, v1);
      blk_ltsolve<NN, OW>(sL, LD, v2);
      if (tid < NN) {
        dr = tb[to.sdi() + tid] * v2[tid];
        v3[tid] = tb[to.sd() + tid] * dr;
      }
      __syncthreads();
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

  // Beam right-hand side (src/DisortSolver.cpp:1366-1373).
  double bsr = 0.0, bdr = 0.0;
  if (beam && tid < NN) {
    double se = 0.0, so = 0.0;
    for (int l = m; l < N; ++l) {
      const double term = chi[l] * Pm[l * NN + tid] * p0[l];
      if (((l - m) & 1) == 0) se += term;
      else so += term;
    }
    bsr = fm * F0 * se / (2.0 * kPi);
    bdr = fm * F0 * so / (2.0 * kPi);
  }

  double zbu = 0.0, zbd = 0.0, s1u = 0.0, s1d = 0.0;
  if (beam && !spherical) {
    double mub = mu0;
    const double mu0sq = mu0 * mu0;
    for (int j = 0; j < NN; ++j) {
      const double kj = sqrt(fabs(lam[j]));
      if (fabs(1.0 - mu0sq * kj * kj) < 1000.0 * kEps) mub = 0.999 * mu0;
    }
    solve_ab(bsr, bdr, false, mub, zbu, zbd);
  } else if (spherical) {
    // src/DisortSolver.cpp:1397-1531 (xba and its dithers above).
    const double q0a = exp(-((lc == 0) ? 0.0 : a.lay_d[LD_CHTAU * TL + g - 1]));
    const double q2a = exp(-a.lay_d[LD_CHTAU * TL + g]);
    const double et = exp(xba * taut), eb = exp(xba * taub), dtl = taub - taut;
    const double zju = 0.5 * (bsr + bdr), zjd = 0.5 * (bsr - bdr);
    const double x1u = (q2a * zju * eb - q0a * zju * et) / dtl, x1d = (q2a * zjd * eb - q0a * zjd * et) / dtl;
    const double x0u = q0a * zju * et - x1u * taut, x0d = q0a * zjd * et - x1d * taut;
    const bool xz = (xba == 0.0);
    const double mubs = xz ? 0.0 : 1.0 / xba;
    solve_ab(x1u + x1d, x1u - x1d, xz, mubs, s1u, s1d);
    double r0u = 0.0, r0d = 0.0;
    if (tid < NN) {
      const double cmr = tb[to.cmu() + tid];
      r0u = x0u + cmr * s1u;
      r0d = x0d - cmr * s1d;
    }
    solve_ab(r0u + r0d, r0u - r0d, xz, mubs, zbu, zbd);
  }

  double xr0 = 0.0, xr1 = 0.0, z0u = 0.0, z0d = 0.0, z1h = 0.0;
  if (thermal) {
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
    const double btt = a.lay_d[LD_BT_TOP * TL + g], btb = a.lay_d[LD_BT_BOT * TL + g];
    double btu, btd, bbu, bbd;
    if (spherical) {
      const double at = exp(-xba * taut), ab = exp(-xba * taub);
      btu = at * (zbu + s1u * taut);
      btd = at * (zbd + s1d * taut);
      bbu = ab * (zbu + s1u * taub);
      bbd = ab * (zbd + s1d * taub);
    } else {
      btu = zbu * btt;
      btd = zbd * btt;
      bbu = zbu * btb;
      bbd = zbd * btb;
    }
    E[ge.ztop() + ru] = btu + z0u + z1h * taut;
    E[ge.ztop() + rd] = btd + z0d + z1h * taut;
    E[ge.zbot() + ru] = bbu + z0u + z1h * taub;
    E[ge.zbot() + rd] = bbd + z0d + z1h * taub;
    E[ge.zq() + ru] = zbu * btb;
    E[ge.zq() + rd] = zbd * btb;
    E[ge.zb() + ru] = zbu;
    E[ge.zb() + rd] = zbd;
    E[ge.zbs() + ru] = s1u;
    E[ge.zbs() + rd] = s1d;
    E[ge.z0() + ru] = z0u;
    E[ge.z0() + rd] = z0d;
    E[ge.z1() + ru] = z1h;
    E[ge.z1() + rd] = z1h;
    // Quadrature-order vectors for the source interpolation: the beam
    // solution, or with the spherical beam its right-hand side ZJ
    // (src/DisortSolver.cpp:1525-1528).
    zr_u[tid] = spherical ? 0.5 * (bsr + bdr) : zbu;
    zr_d[tid] = spherical ? 0.5 * (bsr - bdr) : zbd;
    z0_u[tid] = z0u;
    z0_d[tid] = z0d;
    z1_h[tid] = z1h;
  }
  if (tid == 0) {
    E[ge.xba()] = xba;
    E[ge.xr0()] = xr0;
    E[ge.xr1()] = xr1;
  }

  // Interpolation to the user angles (src/DisortSolver.cpp:2468-2580).
  const bool do_user = (a.flags & DISORT_G_USER_MU) && !(a.flags & DISORT_G_ONLY_FLUXES);
  if (!do_user) return;
  __syncthreads();
  const int numu = a.numu;
  const double* Yu = a.ylu + (int64_t)m * N * numu;  // [l][iu]
  const double* Pw = a.ylw + (int64_t)m * N * NN;    // [l][r]
  // psi for the beam and thermal sources.
  for (int e = tid; e < 3 * N; e += blockDim.x) {
    const int k = e / N, l = e % N;
    double s = 0.0;
    if (l >= m && (k == 0 ? beam : thermal)) {
      const double sgn = ((l - m) & 1) ? -1.0 : 1.0;
#pragma unroll 2
      for (int r = 0; r < NN; ++r) {
        const double w = Pw[l * NN + r];  // w_r Y_l^m(mu_r)
        const double xu = (k == 0) ? zr_u[r] : (k == 1 ? z0_u[r] : z1_h[r]);
        const double xd = (k == 0) ? zr_d[r] : (k == 1 ? z0_d[r] : z1_h[r]);
        s += w * (xu + sgn * xd);
      }
    }
    psi[e] = 0.5 * chi[l] * s;
  }
  // W(l, col) = chi_l / 2 sum_jq w_jq Y_l^m(mu_jq) G(row(jq), col), then
  // G_u = Y_u W, CW columns at a time (W[l][col - cb] in the first LDS
  // slot).  With j' the column's eigenvector and the parity of l - m, the
  // entry G(row, col) is +-(G+ + G-) or +-(G+ - G-) (exactly: a sign flip
  // and a commuted sum); both combinations are staged in the other two
  // slots (column j' contiguous), so the inner sums read LDS and the cached
  // weighted Y table only.
  constexpr int CW = (NN + 1) / 2;  // N * CW <= NN * LD
  double* const W = sbuf;
  double* const sGs = sbuf + NN * LD;
  double* const sGd = sbuf + 2 * NN * LD;
  for (int e = tid; e < NN * NN; e += blockDim.x) {
    const int j = e / NN, r = e % NN;
    const double gp = E[ge.gp() + e], gmv = E[ge.gm() + e];
    sGs[j * LD + r] = gp + gmv;
    sGd[j * LD + r] = gp - gmv;
  }
  __syncthreads();
  double* gu = a.gu + t * (int64_t)numu * N;
#pragma unroll 1
  for (int cb = 0; cb < N; cb += CW) {
#pragma unroll 1
    for (int e = tid; e < N * CW; e += blockDim.x) {
      const int l = e / CW, col = cb + e % CW;
      double s = 0.0;
      if (l >= m && col < N) {
        // up: G+ + sgn G-; down: -(G- + sgn G+), sgn = (-1)^(l-m).
        const bool even = ((l - m) & 1) == 0;
        const bool up = col >= NN;
        const int jp = up ? col - NN : NN - 1 - col;
        const double* gc = (even ? sGs : sGd) + jp * LD;
        const double sg = (!up && even) ? -1.0 : 1.0;
        const double* pw = Pw + l * NN;
#pragma unroll 4
        for (int r = 0; r < NN; ++r) s += pw[r] * (sg * gc[r]);
      }
      W[e] = s * (0.5 * chi[l]);
    }
    __syncthreads();
    for (int e = tid; e < numu * CW; e += blockDim.x) {
      const int iu = e / CW, cc = e % CW;
      if (cb + cc < N) {
        double s = 0.0;
#pragma unroll 8
        for (int l = m; l < N; ++l) s += Yu[l * numu + iu] * W[l * CW + cc];
        gu[iu * N + cb + cc] = s;
      }
    }
    __syncthreads();
  }
  double* us = a.usrc + t * (int64_t)kUSrc * numu;
  const double fact = fm * F0 / (4.0 * kPi);
  for (int iu = tid; iu < numu; iu += blockDim.x) {
    double sb = 0.0, s0 = 0.0, s1 = 0.0;
#pragma unroll 4
    for (int l = m; l < N; ++l) {
      const double y = Yu[l * numu + iu];
      sb += y * (psi[l] + fact * chi[l] * p0[l]);
      s0 += y * psi[N + l];
      s1 += y * psi[2 * N + l];
    }
    us[US_BEAM * numu + iu] = beam ? sb : 0.0;
    us[US_Z0 * numu + iu] = thermal ? s0 + (1.0 - ws) * xr0 : 0.0;
    us[US_Z1 * numu + iu] = thermal ? s1 + (1.0 - ws) * xr1 : 0.0;
  }
}


// ===========================================================================
// k_gboundary: one workgroup per (column, mode): the panel LU of the flux
// workgroup kernel with the mode's boundary conditions
// (src/DisortSolver.cpp:1153-1317 setMatrix, :1669-1974 solve0).
// ===========================================================================
template <int NN>
__global__ __launch_bounds__(WideCfg<NN>::BT) void k_gboundary(GenArgs a) {
  using C = WideCfg<NN>;
  constexpr int N = C::N, R = C::R, RP = C::RP, NWH = RP / 64;
  constexpr TabOff to(N);
  constexpr GE ge(N);
  constexpr int PS = panel_stride(N);
  __shared__ double prow[2 * N + 2];  // pivot row: left block (shifted), right block, rhs
  __shared__ double colc[2][RP];      // pivot-column entry of every row, by step parity
  __shared__ double red_v[NWH];
  __shared__ int red_i[NWH];
  __shared__ int uflag[RP];
  __shared__ double ecur[NN], enxt[NN];

  const int tid = threadIdx.x;
  const int h = tid / RP, r = tid % RP;
  const bool row = r < R;
  const int64_t bk = blockIdx.x;
  if (bk >= a.nc * a.nm) return;
  const int64_t cl = bk / a.nm;
  const int m = (int)(bk % a.nm);
  const int64_t i = a.c0 + cl;
  const int L = a.nlyr;
  const int64_t t0 = bk * L;
  const int64_t lo = i * L, TL = a.n * L;
  const double* sc = a.scal + i * DISORT_NSCAL;
  const double* tb = a.tab;
  const int M = a.inst_i[i * kInstI + II_NCUT];
  const bool lyrcut = a.inst_i[i * kInstI + II_LYRCUT] != 0;
  const bool diffusion = (a.flags & DISORT_F_DIFFUSION_BC) != 0;
  const bool thermal = (a.flags & DISORT_F_THERMAL) != 0;
  const bool lamb = (a.flags & kBrdfFlags) == 0;
  const double F0 = sc[DISORT_S_BEAM_FLUX], mu0 = sc[DISORT_S_BEAM_MU];
  const double alb = sc[DISORT_S_ALBEDO];
  const double d0 = (m == 0) ? 1.0 : 0.0;
  const double* sd = lamb ? nullptr : a.srf_d + ((int64_t)a.rpv_set[i] * a.nm + m) * surf_d_size(NN, a.numu);
  const double* sb = lamb ? nullptr : a.srf_b + (cl * a.nm + m) * surf_b_size(NN, a.numu);
  // Surface reflection refl_quad(iq, jq+1) and refl_quad(iq, 0) at mode m.
  auto rq = [&](int iq, int jq) { return lamb ? (m == 0 ? alb : 0.0) : sd[iq * NN + jq]; };
  auto rq0 = [&](int iq) { return lamb ? (m == 0 ? alb : 0.0) : sb[iq]; };
  auto eig = [&](int64_t tt) { return a.eig + tt * ge.size(); };

  auto layer_exp = [&](int64_t tt, double* e) {
    for (int jp = tid; jp < NN; jp += blockDim.x) e[jp] = exp(-eig(tt)[ge.k() + jp] * a.lay_d[LD_DTAU_S * TL + lo + (tt - t0)]);
  };
  auto etop_of = [](const double* e, int j) { return j < NN ? e[NN - 1 - j] : 1.0; };
  auto ebot_of = [](const double* e, int j) { return j < NN ? 1.0 : e[j - NN]; };

  double A[N];
  double rhs = 0.0;
  bool used;

  layer_exp(t0, ecur);
  __syncthreads();
  {
    const bool tbc = row && r < NN;
    const double top = (m == 0) ? sc[DISORT_S_ISO_TOP] + a.inst_d[i * kInstD + ID_PL_TOP] : 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) A[j] = 0.0;
    if (tbc && h == 0) {
#pragma unroll
      for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
        for (int j = j0; j < j0 + 8 && j < N; ++j) A[j] = ge_G<NN>(eig(t0), NN - 1 - r, j) * etop_of(ecur, j);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    if (h == 1) rhs = tbc ? top - eig(t0)[ge.ztop() + (tbc ? NN - 1 - r : 0)] : 0.0;
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

  auto used_rank = [&]() {
    if (h == 0) uflag[r] = (row && used) ? 1 : 0;
    __syncthreads();
    int q = 0;
    for (int k = 0; k < r; ++k) q += uflag[k];
    __syncthreads();
    return q;
  };

  for (int ip = 0; ip < M - 1; ++ip) {
    const int64_t tt = t0 + ip;
    const bool left = h == (ip & 1);
    layer_exp(tt, ecur);
    layer_exp(tt + 1, enxt);
    const int q = used_rank();
    const bool fill = row && used;
    // Fills in chunks of 8 columns (scheduling barriers bound the loads in
    // flight, and so the registers, next to the live row).
    if (left) {
      if (fill) {
#pragma unroll
        for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
          for (int j = j0; j < j0 + 8 && j < N; ++j) A[j] = ge_G<NN>(eig(tt), q, j) * ebot_of(ecur, j);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    } else if (fill) {
#pragma unroll
      for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
        for (int j = j0; j < j0 + 8 && j < N; ++j) A[j] = -ge_G<NN>(eig(tt + 1), q, j) * etop_of(enxt, j);
        __builtin_amdgcn_sched_barrier(0);
      }
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) A[j] = 0.0;
    }
    if (h == 1 && fill) rhs = eig(tt + 1)[ge.ztop() + q] - eig(tt)[ge.zbot() + q];
    used = !row;
    eliminate(left, a.panel + tt * PS);
  }

  // Bottom boundary: reflection of the downward field, coupled at m = 0 for
  // a Lambertian surface and at every mode for a BRDF.
  {
    const int64_t tt = t0 + M - 1;
    const double* Eb = eig(tt);
    const bool left = h == ((M - 1) & 1);
    const bool couple = !(lyrcut || diffusion || (lamb && m > 0));
    layer_exp(tt, ecur);
    const int q = used_rank();
    const bool fill = row && used;
    const bool bot = fill && q < NN;
    const int tq = bot ? q : 0;
    if (left) {
      if (bot) {
#pragma unroll
        for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
          for (int j = j0; j < j0 + 8 && j < N; ++j) {
            double rs = 0.0;
            if (couple)
              for (int k = 0; k < NN; ++k)
                rs += tb[to.cwt() + k] * tb[to.cmu() + k] * rq(tq, k) * ge_G<NN>(Eb, NN - 1 - k, j);
            A[j] = (ge_G<NN>(Eb, NN + tq, j) - (1.0 + d0) * rs) * ebot_of(ecur, j);
          }
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
      double v = 0.0;
      if (m == 0) {
        const double plb = a.inst_d[i * kInstD + ID_PL_BOT];
        if (diffusion) {
          v = -Eb[ge.zbot() + NN + tq] + plb + tb[to.cmu() + tq] * a.inst_d[i * kInstD + ID_PL_BOT_DERIV];
        } else {
          if (!lyrcut) {
            double rs = 0.0;
            for (int k = 0; k < NN; ++k) rs += tb[to.cwt() + k] * tb[to.cmu() + k] * rq(tq, k) * Eb[ge.zbot() + NN - 1 - k];
            v += 2.0 * rs;
          }
          v -= Eb[ge.zbot() + NN + tq];
          if (F0 > 0.0 && !lyrcut) v += rq0(tq) * mu0 * F0 * btn / kPi;
          v += sc[DISORT_S_ISO_BOT];
          if (thermal) {
            const double eq = lamb ? 1.0 - alb : sd[NN * NN + a.numu * NN + tq];
            v += eq * plb;
          }
        }
      } else if (F0 > 0.0) {
        if (lyrcut || lamb) {
          v = -Eb[ge.zbot() + NN + tq];
        } else {
          double s = 0.0;
          for (int k = 0; k < NN; ++k) s += tb[to.cwt() + k] * tb[to.cmu() + k] * rq(tq, k) * Eb[ge.zq() + NN - 1 - k];
          v = s + rq0(tq) * mu0 * F0 / kPi * btn - Eb[ge.zq() + NN + tq];
        }
      }
      rhs = bot ? v : (fill ? 0.0 : rhs);
    }
    used = !row || (fill && !bot);
    eliminate(left, a.panel + tt * PS);
  }

  // Layers below the cut-off carry no solution.
  for (int e = tid; e < (L - M) * N; e += blockDim.x) a.cvec[(t0 + M) * N + e] = 0.0;

  // Back substitution, lane k = U row k of each panel, first wavefront only.
  if (tid >= 64) return;
  panel_back_substitute<N>(a.panel, t0, M, a.cvec);
}

// One-wavefront form of k_gboundary for 3NN <= 64 panel rows: lane r holds
// both blocks of panel row r (Ac, the block of the pivot columns, shifted
// left by one column per step; An, the next layer's columns) and the
// right-hand side, so every synchronisation is wave-local.  Same pivots and
// operations as k_gboundary; the pivot row reaches the other lanes by
// readlane (no LDS traffic) and is stored from its lane with U in shifted
// form.  Three wavefronts per SIMD: the register budget (168) spills only
// outside the elimination loop.
template <int NN>
__global__ __launch_bounds__(64, 3) void k_gboundary1(GenArgs a) {
  constexpr int N = 2 * NN, R = 3 * NN;
  static_assert(R <= 64, "one panel row per lane");
  constexpr TabOff to(N);
  constexpr GE ge(N);
  constexpr int PS = panel_stride(N);
  __shared__ double ecur[NN], enxt[NN];

  const int r = threadIdx.x;
  const bool row = r < R;
  const int64_t bk = blockIdx.x;
  if (bk >= a.nc * a.nm) return;
  const int64_t cl = bk / a.nm;
  const int m = (int)(bk % a.nm);
  const int64_t i = a.c0 + cl;
  const int L = a.nlyr;
  const int64_t t0 = bk * L;
  const int64_t lo = i * L, TL = a.n * L;
  const double* sc = a.scal + i * DISORT_NSCAL;
  const double* tb = a.tab;
  const int M = a.inst_i[i * kInstI + II_NCUT];
  const bool lyrcut = a.inst_i[i * kInstI + II_LYRCUT] != 0;
  const bool diffusion = (a.flags & DISORT_F_DIFFUSION_BC) != 0;
  const bool thermal = (a.flags & DISORT_F_THERMAL) != 0;
  const bool lamb = (a.flags & kBrdfFlags) == 0;
  const double F0 = sc[DISORT_S_BEAM_FLUX], mu0 = sc[DISORT_S_BEAM_MU];
  const double alb = sc[DISORT_S_ALBEDO];
  const double d0 = (m == 0) ? 1.0 : 0.0;
  const double* sd = lamb ? nullptr : a.srf_d + ((int64_t)a.rpv_set[i] * a.nm + m) * surf_d_size(NN, a.numu);
  const double* sb = lamb ? nullptr : a.srf_b + (cl * a.nm + m) * surf_b_size(NN, a.numu);
  auto rq = [&](int iq, int jq) { return lamb ? (m == 0 ? alb : 0.0) : sd[iq * NN + jq]; };
  auto rq0 = [&](int iq) { return lamb ? (m == 0 ? alb : 0.0) : sb[iq]; };
  auto eig = [&](int64_t tt) { return a.eig + tt * ge.size(); };
  auto layer_exp = [&](int64_t tt, double* e) {
    if (r < NN) e[r] = exp(-eig(tt)[ge.k() + r] * a.lay_d[LD_DTAU_S * TL + lo + (tt - t0)]);
  };
  auto etop_of = [](const double* e, int j) { return j < NN ? e[NN - 1 - j] : 1.0; };
  auto ebot_of = [](const double* e, int j) { return j < NN ? 1.0 : e[j - NN]; };
  // Rank of this row among the used rows.
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
    const double top = (m == 0) ? sc[DISORT_S_ISO_TOP] + a.inst_d[i * kInstD + ID_PL_TOP] : 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) An[j] = 0.0;
    if (tbc) {
#pragma unroll
      for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
        for (int j = j0; j < j0 + 8 && j < N; ++j) An[j] = ge_G<NN>(eig(t0), NN - 1 - r, j) * etop_of(ecur, j);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    rhs = tbc ? top - eig(t0)[ge.ztop() + (tbc ? NN - 1 - r : 0)] : 0.0;
    used = !tbc;
  }

  // N elimination steps of one panel (run-time loop; the pivot column is
  // always Ac[0]).
  auto eliminate = [&](double* pp) __attribute__((always_inline)) {
#pragma unroll 1
    for (int c = 0; c < N; ++c) {
      const double key = (row && !used) ? fabs(Ac[0]) : -1.0;
      double mx = row_pair_max(row_max(key));
      mx = fmax(lane_bcast(mx, 0), lane_bcast(mx, 32));
      const int p = __ffsll((long long)__ballot(key == mx)) - 1;
      // The pivot row stays in lane p: the other lanes read its entries by
      // readlane (wave-uniform values, no LDS traffic), and lane p stores it
      // as U row c in shifted form, [1/U(c,c), U(c,c+1), ..., U(c,N-1), 0...]
      // | V(c,0..N-1) | r_c.
      const double piv = lane_bcast(Ac[0], p);
      const double inv = (piv != 0.0) ? 1.0 / (piv != 0.0 ? piv : 1.0) : 0.0;
      if (r == p) {
        double2* u = reinterpret_cast<double2*>(pp + c * N);
        double2* w = reinterpret_cast<double2*>(pp + N * N + c * N);
        u[0] = make_double2(inv, Ac[1]);
#pragma unroll
        for (int j = 2; j < N; j += 2) u[j / 2] = make_double2(Ac[j], Ac[j + 1]);
#pragma unroll
        for (int j = 0; j < N; j += 2) w[j / 2] = make_double2(An[j], An[j + 1]);
        pp[2 * N * N + c] = rhs;
      }
      const double l = (row && !used && r != p) ? Ac[0] * inv : 0.0;
#pragma unroll
      for (int j0 = 0; j0 < N - 1; j0 += 8) {
#pragma unroll
        for (int j = j0; j < j0 + 8 && j < N - 1; ++j) Ac[j] = Ac[j + 1] - l * lane_bcast(Ac[j + 1], p);
        __builtin_amdgcn_sched_barrier(0);
      }
      Ac[N - 1] = 0.0;
#pragma unroll
      for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
        for (int j = j0; j < j0 + 8 && j < N; ++j) An[j] -= l * lane_bcast(An[j], p);
        __builtin_amdgcn_sched_barrier(0);
      }
      rhs -= l * lane_bcast(rhs, p);
      used = used || (r == p);
    }
  };

  // Interior panels: continuity at the interface below layer ip.
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
          Ac[j] = ge_G<NN>(eig(tt), q, j) * ebot_of(ecur, j);
          An[j] = -ge_G<NN>(eig(tt + 1), q, j) * etop_of(enxt, j);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      rhs = eig(tt + 1)[ge.ztop() + q] - eig(tt)[ge.zbot() + q];
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

  // Bottom boundary: reflection of the downward field, coupled at m = 0 for
  // a Lambertian surface and at every mode for a BRDF.
  {
    const int64_t tt = t0 + M - 1;
    const double* Eb = eig(tt);
    const bool couple = !(lyrcut || diffusion || (lamb && m > 0));
    layer_exp(tt, ecur);
    wave_lds_order();
    const int q = used_rank(used);
    const bool fill = row && used;
    const bool bot = fill && q < NN;
    const int tq = bot ? q : 0;
    if (bot) {
#pragma unroll
      for (int j0 = 0; j0 < N; j0 += 8) {
#pragma unroll
        for (int j = j0; j < j0 + 8 && j < N; ++j) {
          double rs = 0.0;
          if (couple)
            for (int k = 0; k < NN; ++k)
              rs += tb[to.cwt() + k] * tb[to.cmu() + k] * rq(tq, k) * ge_G<NN>(Eb, NN - 1 - k, j);
          Ac[j] = (ge_G<NN>(Eb, NN + tq, j) - (1.0 + d0) * rs) * ebot_of(ecur, j);
        }
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
      double v = 0.0;
      if (m == 0) {
        const double plb = a.inst_d[i * kInstD + ID_PL_BOT];
        if (diffusion) {
          v = -Eb[ge.zbot() + NN + tq] + plb + tb[to.cmu() + tq] * a.inst_d[i * kInstD + ID_PL_BOT_DERIV];
        } else {
          if (!lyrcut) {
            double rs = 0.0;
            for (int k = 0; k < NN; ++k) rs += tb[to.cwt() + k] * tb[to.cmu() + k] * rq(tq, k) * Eb[ge.zbot() + NN - 1 - k];
            v += 2.0 * rs;
          }
          v -= Eb[ge.zbot() + NN + tq];
          if (F0 > 0.0 && !lyrcut) v += rq0(tq) * mu0 * F0 * btn / kPi;
          v += sc[DISORT_S_ISO_BOT];
          if (thermal) {
            const double eq = lamb ? 1.0 - alb : sd[NN * NN + a.numu * NN + tq];
            v += eq * plb;
          }
        }
      } else if (F0 > 0.0) {
        if (lyrcut || lamb) {
          v = -Eb[ge.zbot() + NN + tq];
        } else {
          double s = 0.0;
          for (int k = 0; k < NN; ++k) s += tb[to.cwt() + k] * tb[to.cmu() + k] * rq(tq, k) * Eb[ge.zq() + NN - 1 - k];
          v = s + rq0(tq) * mu0 * F0 / kPi * btn - Eb[ge.zq() + NN + tq];
        }
      }
      rhs = bot ? v : (fill ? 0.0 : rhs);
    }
    used = !row || (fill && !bot);
    eliminate(a.panel + tt * PS);
  }

  // Layers below the cut-off carry no solution.
  for (int e = r; e < (L - M) * N; e += 64) a.cvec[(t0 + M) * N + e] = 0.0;

  // Back substitution, lane k = U row k of each panel.
  panel_back_substitute<N, true>(a.panel, t0, M, a.cvec);
}

// ===========================================================================
// k_gflux: fluxes at the user levels (src/DisortSolver.cpp:1980-2124),
// thread per (chunk column, level), from the m = 0 tasks.
// ===========================================================================
template <int NN>
__global__ __launch_bounds__(64) void k_gflux(GenArgs a) {
  constexpr int N = 2 * NN;
  constexpr TabOff to(N);
  constexpr GE ge(N);
  const int64_t w = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (w >= a.nc * a.ntau) return;
  const int64_t cl = w / a.ntau;
  const int lu = (int)(w % a.ntau);
  const int64_t i = a.c0 + cl;
  const int L = a.nlyr;
  const int64_t lo = i * L, vo = i * (L + 1), TL = a.n * L, TV = a.n * (L + 1);
  const double* sc = a.scal + i * DISORT_NSCAL;
  const double* tb = a.tab;
  const int ncut = a.inst_i[i * kInstI + II_NCUT];
  const bool lyrcut = a.inst_i[i * kInstI + II_LYRCUT] != 0;
  const bool thermal = (a.flags & DISORT_F_THERMAL) != 0;
  const bool spherical = (a.flags & DISORT_F_SPHERICAL) != 0;
  const int lyu = a.u_lyr[i * a.ntau + lu];
  double o[DISORT_NOUT];
#pragma unroll
  for (int f = 0; f < DISORT_NOUT; ++f) o[f] = 0.0;
  if (!(lyrcut && lyu > ncut - 1)) {
    const int64_t tt = cl * a.nm * L + lyu;
    const double* E = a.eig + tt * ge.size();
    const double* LL = a.cvec + tt * N;
    const double F0 = sc[DISORT_S_BEAM_FLUX], mu0 = sc[DISORT_S_BEAM_MU];
    const double ut = a.u_tau[i * a.ntau + lu];
    const double utau = (a.flags & DISORT_G_USER_TAU) ? a.tau_user[i * a.ntau + lu] : a.lev_d[VD_TAUC * TV + vo + lu];
    const double ttop = a.lev_d[VD_TAU_S * TV + vo + lyu], tbot = a.lev_d[VD_TAU_S * TV + vo + lyu + 1];
    double fact = 0.0, dirint = 0.0, fldir = 0.0, fdir = 0.0;
    if (F0 > 0.0) {
      const double ch = a.lay_d[LD_CH * TL + lo + lyu];
      fact = exp(-ut / ch);
      fdir = (spherical ? fabs(mu0) : mu0) * F0 * exp(-utau / ch);
      dirint = F0 * fact;
      fldir = mu0 * F0 * fact;
    }
    double ec[N];
#pragma unroll
    for (int jq = 0; jq < NN; ++jq) ec[jq] = LL[jq] * exp(E[ge.k() + NN - 1 - jq] * (ut - tbot));
#pragma unroll
    for (int jq = NN; jq < N; ++jq) ec[jq] = LL[jq] * exp(-E[ge.k() + jq - NN] * (ut - ttop));
    const double xba = E[ge.xba()];
    double mi = 0.0, mid = 0.0, miu = 0.0, fldn = 0.0, flup = 0.0;
#pragma unroll 1
    for (int q = 0; q < N; ++q) {
      double s = 0.0;
      for (int j = 0; j < N; ++j) s += ge_G<NN>(E, q, j) * ec[j];
      if (F0 > 0.0) s += spherical ? exp(-xba * ut) * (E[ge.zb() + q] + E[ge.zbs() + q] * ut) : E[ge.zb() + q] * fact;
      if (thermal) s += E[ge.z0() + q] + E[ge.z1() + q] * ut;
      const int ci = q < NN ? NN - 1 - q : q - NN;
      const double wq = tb[to.cwt() + ci];
      mi += wq * s;
      if (q < NN) {
        mid += wq * s;
        fldn += wq * s * tb[to.cmu() + ci];
      } else {
        miu += wq * s;
        flup += wq * s * tb[to.cmu() + ci];
      }
    }
    flup *= 2.0 * kPi;
    fldn *= 2.0 * kPi;
    const double fdn = (fldn + fldir) - fdir;
    mi = (2.0 * kPi * mi + dirint) / (4.0 * kPi);
    const double midir = dirint / (4.0 * kPi);
    mid = (2.0 * kPi * mid) / (4.0 * kPi);
    miu = (2.0 * kPi * miu) / (4.0 * kPi);
    const double pls = thermal ? E[ge.xr0()] + E[ge.xr1()] * ut : 0.0;
    o[DISORT_O_FLUX_DIRECT_BEAM] = fdir;
    o[DISORT_O_FLUX_DOWN] = fdn;
    o[DISORT_O_FLUX_UP] = flup;
    o[DISORT_O_FLUX_TAU_DIVERGENCE] = (1.0 - a.ssa[lo + lyu]) * 4.0 * kPi * (mi - pls);
    o[DISORT_O_MEAN_INTENSITY] = mi;
    o[DISORT_O_MEAN_INTENSITY_DOWN] = mid;
    o[DISORT_O_MEAN_INTENSITY_UP] = miu;
    o[DISORT_O_MEAN_INTENSITY_DIRECT_BEAM] = midir;
  }
#pragma unroll
  for (int f = 0; f < DISORT_NOUT; ++f) a.flux[((int64_t)f * a.n + i) * a.ntau + lu] = o[f];
}

// ===========================================================================
// k_gucoef: per task, the user-angle eigenvectors times the integration
// constants (src/DisortSolver.cpp:2965-2974) and, per user angle, the
// coefficients of exp((tau_u - tau_top)/mu) and exp((tau_u - tau_bot)/mu)
// (and of exp(-tau_u/mu0)) with which a whole layer enters every level it
// does not contain (:3010-3066); for the last layer also the bottom-boundary
// intensity (:3156-3220).  Thread per user angle.
// ===========================================================================
enum { UC_E1 = 0, UC_E2 = 1, UC_E0 = 2, UC_BND = 3, kUCoef = 4 };

// Tasks per 64-thread block: lanes (task, user angle), so that short user
// angle lists still fill the wavefront.
constexpr int kUcMaxT = 8;
inline int ucoef_tasks_per_block(int numu) {
  const int t = 64 / (numu > 0 ? numu : 1);
  return t < 1 ? 1 : (t > kUcMaxT ? kUcMaxT : t);
}

template <int NN>
__global__ __launch_bounds__(64) void k_gucoef(GenArgs a, double* ucoef, int tpb) {
  constexpr int N = 2 * NN;
  constexpr TabOff to(N);
  constexpr GE ge(N);
  __shared__ double wk_s[kUcMaxT * NN];
  const int L = a.nlyr, numu = a.numu;
  // Lane (sub-task st, first user angle iu0); with numu > 64 one task per
  // block and each lane steps through the angles by 64.
  const int span = tpb > 1 ? numu : 64;
  const int st = (int)threadIdx.x / span, iu0 = (int)threadIdx.x % span;
  const int64_t t = (int64_t)blockIdx.x * tpb + (st < tpb ? st : 0);
  const bool live = st < tpb && t < a.n_tasks;
  const int lc = (int)(t % L);
  const int64_t cm = t / L;
  const int m = (int)(cm % a.nm);
  const int64_t cl = cm / a.nm;
  const int64_t i = a.c0 + cl;
  const int64_t lo = i * L, vo = i * (L + 1), TL = a.n * L, TV = a.n * (L + 1);
  const int ncut = live ? a.inst_i[i * kInstI + II_NCUT] : 0;
  const bool lyrcut = live && a.inst_i[i * kInstI + II_LYRCUT] != 0;
  // No record below the cut-off (k_glayer skipped the layer); k_guser reads
  // none of those tasks' coefficients.
  const bool run = live && lc < ncut;
  const double* sc = a.scal + i * DISORT_NSCAL;
  const double* E = a.eig + t * ge.size();
  double* const wk = wk_s + (st < kUcMaxT ? st : 0) * NN;
  const double dt = run ? a.lay_d[LD_DTAU_S * TL + lo + lc] : 0.0;
  for (int q = iu0; q < NN; q += span)
    if (run) wk[q] = exp(-E[ge.k() + NN - 1 - q] * dt);  // exp(kk_q dt), kk_q < 0
  __syncthreads();
  if (!run) return;
  const double F0 = sc[DISORT_S_BEAM_FLUX], mu0 = sc[DISORT_S_BEAM_MU];
  const double* LL = a.cvec + t * N;
  double* gu = a.gu + t * (int64_t)numu * N;
  const double* us = a.usrc + t * (int64_t)kUSrc * numu;
  double* uc = ucoef + t * (int64_t)kUCoef * numu;
  const bool thermal = (a.flags & DISORT_F_THERMAL) != 0 && m == 0;
  const double d0 = (m == 0) ? 1.0 : 0.0;
  const double ttop = a.lev_d[VD_TAU_S * TV + vo + lc], tbot = a.lev_d[VD_TAU_S * TV + vo + lc + 1];
  const double btt = a.lay_d[LD_BT_TOP * TL + lo + lc], btb = a.lay_d[LD_BT_BOT * TL + lo + lc];
  for (int iu = iu0; iu < numu; iu += span) {
    double* g = gu + (int64_t)iu * N;
    for (int q = 0; q < N; ++q) g[q] *= LL[q];
    const double mu = a.mu_user[iu];
    const double sgn = mu < 0.0 ? -1.0 : 1.0;
    double c1 = 0.0, c2 = 0.0, c0 = 0.0;
    if (thermal) {
      const double z0 = us[US_Z0 * numu + iu], z1 = us[US_Z1 * numu + iu];
      c1 += sgn * (z0 + z1 * (ttop + mu));
      c2 -= sgn * (z0 + z1 * (tbot + mu));
    }
    if (F0 > 0.0) {
      const double zb = us[US_BEAM * numu + iu];
      const double den = 1.0 + mu / mu0;
      if (fabs(den) < 0.0001) {
        c0 += zb * (dt / mu0);
      } else {
        c1 += zb * sgn * btt / den;
        c2 -= zb * sgn * btb / den;
      }
    }
    for (int q = 0; q < NN; ++q) {
      const double kq = -E[ge.k() + NN - 1 - q];
      const double den = 1.0 + mu * kq;
      if (fabs(den) < 0.0001) {
        c2 += g[q] * (dt / mu);
      } else {
        c1 += g[q] * sgn * wk[q] / den;
        c2 -= g[q] * sgn / den;
      }
    }
    for (int q = NN; q < N; ++q) {
      const double kq = E[ge.k() + q - NN];
      const double den = 1.0 + mu * kq;
      if (fabs(den) < 0.0001) {
        c1 -= g[q] * (dt / mu);
      } else {
        c1 += g[q] * sgn / den;
        c2 -= g[q] * sgn * wk[N - 1 - q] / den;
      }
    }
    uc[UC_E1 * numu + iu] = c1;
    uc[UC_E2 * numu + iu] = c2;
    uc[UC_E0 * numu + iu] = c0;
    // Bottom-boundary intensity leaving the surface upward at mu (> 0).
    double bnd = 0.0;
    if (lc == L - 1 && mu > 0.0) {
      const bool diffusion = (a.flags & DISORT_F_DIFFUSION_BC) != 0;
      const bool lamb = (a.flags & kBrdfFlags) == 0;
      if (diffusion) {
        if (!lyrcut && m == 0)
          bnd = a.inst_d[i * kInstD + ID_PL_BOT] + mu * a.inst_d[i * kInstD + ID_PL_BOT_DERIV];
      } else if (!(lyrcut || (lamb && m > 0))) {
        const double alb = sc[DISORT_S_ALBEDO];
        const double* sd = lamb ? nullptr : a.srf_d + ((int64_t)a.rpv_set[i] * a.nm + m) * surf_d_size(NN, numu);
        const double* sb = lamb ? nullptr : a.srf_b + (cl * a.nm + m) * surf_b_size(NN, numu);
        const double btL = a.lay_d[LD_BT_BOT * TL + lo + L - 1];
        double bnddfu = 0.0;
        for (int q = NN - 1; q >= 0; --q) {
          double dfu = 0.0;
          for (int j = 0; j < NN; ++j) dfu += ge_G<NN>(E, q, j) * LL[j];
          for (int j = NN; j < N; ++j) dfu += ge_G<NN>(E, q, j) * LL[j] * wk[N - 1 - j];
          if (F0 > 0.0) dfu += E[ge.zb() + q] * btL;
          dfu += d0 * (E[ge.z0() + q] + E[ge.z1() + q] * tbot);
          const double ru = lamb ? alb : sd[NN * NN + iu * NN + (NN - 1 - q)];
          bnddfu += (1.0 + d0) * ru * a.tab[to.cmu() + NN - 1 - q] * a.tab[to.cwt() + NN - 1 - q] * dfu;
        }
        const double ru0 = lamb ? alb : sb[NN + iu];
        const double bnddir = mu0 * F0 / kPi * ru0 * btL;
        const double eu = lamb ? 1.0 - alb : sd[NN * NN + numu * NN + NN + iu];
        bnd = bnddfu + bnddir + d0 * eu * a.inst_d[i * kInstD + ID_PL_BOT] + sc[DISORT_S_ISO_BOT];
      }
    }
    uc[UC_BND * numu + iu] = bnd;
  }
}

// ===========================================================================
// k_guser: intensity component of one mode at every user level and angle
// (src/DisortSolver.cpp:2976-3226); block per (chunk column, mode), thread
// per (level, angle).
// ===========================================================================
template <int NN>
__global__ __launch_bounds__(256) void k_guser(GenArgs a, const double* ucoef) {
  constexpr int N = 2 * NN;
  constexpr GE ge(N);
  const int64_t bk = blockIdx.x;
  if (bk >= a.nc * a.nm) return;
  const int64_t cl = bk / a.nm;
  const int m = (int)(bk % a.nm);
  const int64_t i = a.c0 + cl;
  const int L = a.nlyr, numu = a.numu;
  const int64_t t0 = bk * L;
  const int64_t vo = i * (L + 1), TL = a.n * L, TV = a.n * (L + 1);
  const int ncut = a.inst_i[i * kInstI + II_NCUT];
  const bool lyrcut = a.inst_i[i * kInstI + II_LYRCUT] != 0;
  const double* sc = a.scal + i * DISORT_NSCAL;
  const double F0 = sc[DISORT_S_BEAM_FLUX], mu0 = sc[DISORT_S_BEAM_MU];
  const bool thermal = (a.flags & DISORT_F_THERMAL) != 0 && m == 0;
  const double* taus = a.lev_d + VD_TAU_S * TV + vo;
  double* out = a.uum + bk * (int64_t)a.ntau * numu;
  const double pl_top = a.inst_d[i * kInstD + ID_PL_TOP];
  for (int e = threadIdx.x; e < a.ntau * numu; e += blockDim.x) {
    const int lu = e / numu, iu = e % numu;
    const int lyu = a.u_lyr[i * a.ntau + lu];
    double res = 0.0;
    if (!(lyrcut && lyu >= ncut)) {
      const double mu = a.mu_user[iu];
      const double ut = a.u_tau[i * a.ntau + lu];
      const bool neg = mu < 0.0;
      const double exp0 = (F0 > 0.0) ? exp(-ut / mu0) : 0.0;
      const int l0 = neg ? 0 : lyu + 1, l1 = neg ? lyu - 1 : ncut - 1;
      double pal = 0.0;
      for (int lc = l0; lc <= l1; ++lc) {
        const double* uc = ucoef + (t0 + lc) * (int64_t)kUCoef * numu;
        const double e1 = exp((ut - taus[lc]) / mu);
        const double e2 = exp((ut - taus[lc + 1]) / mu);
        pal += uc[UC_E1 * numu + iu] * e1 + uc[UC_E2 * numu + iu] * e2 + uc[UC_E0 * numu + iu] * exp0;
      }
      // Layer holding the level.
      const int64_t tt = t0 + lyu;
      const double* E = a.eig + tt * ge.size();
      const double dtau1 = ut - taus[lyu], dtau2 = ut - taus[lyu + 1];
      if ((fabs(dtau1) >= 1e-6 || !neg) && (fabs(dtau2) >= 1e-6 || neg)) {
        const double* g = a.gu + tt * (int64_t)numu * N + (int64_t)iu * N;
        const double* us = a.usrc + tt * (int64_t)kUSrc * numu;
        const double e1 = neg ? exp(dtau1 / mu) : 0.0;
        const double e2 = neg ? 0.0 : exp(dtau2 / mu);
        if (F0 > 0.0) {
          const double den = 1.0 + mu / mu0;
          const double btt = a.lay_d[LD_BT_TOP * TL + i * L + lyu], btb = a.lay_d[LD_BT_BOT * TL + i * L + lyu];
          double expn;
          if (fabs(den) < 0.0001) expn = (dtau1 / mu0) * exp0;
          else if (neg) expn = (exp0 - btt * e1) / den;
          else expn = (exp0 - btb * e2) / den;
          pal += us[US_BEAM * numu + iu] * expn;
        }
        const double dl = a.lay_d[LD_DTAU_S * TL + i * L + lyu];
        for (int q = 0; q < NN; ++q) {
          const double k = -E[ge.k() + NN - 1 - q], den = 1.0 + mu * k;
          double expn;
          if (fabs(den) < 0.0001) expn = -dtau2 / mu * e2;
          else if (neg) expn = (exp(-k * dtau2) - exp(k * dl) * e1) / den;
          else expn = (exp(-k * dtau2) - e2) / den;
          pal += g[q] * expn;
        }
        for (int q = NN; q < N; ++q) {
          const double k = E[ge.k() + q - NN], den = 1.0 + mu * k;
          double expn;
          if (fabs(den) < 0.0001) expn = -(dtau1 / mu) * e1;
          else if (neg) expn = (exp(-k * dtau1) - e1) / den;
          else expn = (exp(-k * dtau1) - exp(-k * dl) * e2) / den;
          pal += g[q] * expn;
        }
        if (thermal) {
          const double expn = neg ? e1 : e2;
          const double fact = neg ? taus[lyu] + mu : taus[lyu + 1] + mu;
          pal += us[US_Z0 * numu + iu] * (1.0 - expn) + us[US_Z1 * numu + iu] * (ut + mu - fact * expn);
        }
      }
      double bnd = 0.0;
      if (neg) {
        if (m == 0) bnd = (sc[DISORT_S_ISO_TOP] + pl_top) * exp(ut / mu);
      } else if (!lyrcut) {
        // With the cut-off in effect the surface term is zero (and the last
        // layer's task did not run).
        const double* ucL = ucoef + (t0 + L - 1) * (int64_t)kUCoef * numu;
        bnd = ucL[UC_BND * numu + iu] * exp((ut - taus[L]) / mu);
      }
      res = pal + bnd;
    }
    out[e] = res;
  }
}

// ===========================================================================
// k_gaccum: Fourier sum over the modes at the user azimuths, with the
// accuracy_fourier_series early exit (src/DisortSolver.cpp:336-346,
// :3228-3266), and the azimuthal average; block per chunk column.
// ===========================================================================
constexpr int kAccPhi = 64;

__global__ __launch_bounds__(256) void k_gaccum(GenArgs a, int32_t* modes_used) {
  __shared__ double red[4];
  __shared__ double cph[kAccPhi];
  const int64_t cl = blockIdx.x;
  if (cl >= a.nc) return;
  const int64_t i = a.c0 + cl;
  const int numu = a.numu, nphi = a.nphi, ntau = a.ntau;
  const double* sc = a.scal + i * DISORT_NSCAL;
  const double F0 = sc[DISORT_S_BEAM_FLUX], mu0 = sc[DISORT_S_BEAM_MU], phi0 = sc[DISORT_S_BEAM_PHI];
  const bool single = F0 == 0.0 || fabs(1.0 - mu0) < 1.e-5 || (a.flags & DISORT_G_ONLY_FLUXES);
  const int naz = single ? 0 : a.nm - 1;
  const int64_t per = (int64_t)ntau * numu;
  const double* u = a.uum + cl * a.nm * per;
  double* I = a.inten + i * per * nphi;
  double* avg = a.inten_avg + i * per;
  for (int64_t e = threadIdx.x; e < per; e += blockDim.x) {
    avg[e] = u[e];
    for (int j = 0; j < nphi; ++j) I[e * nphi + j] = u[e];
  }
  __syncthreads();
  int kconv = 0, used = 1;
  for (int m = 1; m <= naz; ++m) {
    used = m + 1;
    const double* um = u + m * per;
    // cos(m (phi_j - phi0)) once per mode and azimuth (the first kAccPhi in
    // LDS; beyond them evaluated in place).
    for (int j = threadIdx.x; j < nphi && j < kAccPhi; j += blockDim.x) {
      const double phirad = (a.phi_user[j] - phi0) * kPi / 180.0;
      cph[j] = cos((double)m * phirad);
    }
    __syncthreads();
    double mx = 0.0;
    for (int64_t q = threadIdx.x; q < per; q += blockDim.x) {
      const double uq = um[q];
      double* Iq = I + q * nphi;
      for (int j = 0; j < nphi; ++j) {
        double cj;
        if (j < kAccPhi) {
          cj = cph[j];
        } else {
          const double phirad = (a.phi_user[j] - phi0) * kPi / 180.0;
          cj = cos((double)m * phirad);
        }
        const double az = uq * cj;
        const double v = Iq[j] + az;
        Iq[j] = v;
        const double ua = fabs(v);
        double r;
        if (ua == 0.0) r = 1.0 + fabs(az);
        else if (fabs(az) == 0.0) r = 0.0;
        else r = fabs(az) / ua;
        mx = r > mx ? r : mx;
      }
    }
    if (a.accuracy > 0.0) {
      // Block maximum (values are >= 0).
      for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(mx, off);
        mx = o > mx ? o : mx;
      }
      __syncthreads();
      if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
      __syncthreads();
      double b = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) b = red[w] > b ? red[w] : b;
      if (b <= a.accuracy) ++kconv;
      if (kconv >= 2) break;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) modes_used[cl] = used;
}

// Single-scattered beam intensity at scaled depth tu in 1-based layer layru
// with phase function ph[1..ncut] and albedos om(l) (singleScat,
// src/DisortSolver.cpp:3287-3328); taus: scaled cumulative depths.
template <class Om>
__device__ double single_scat(double dither, int layru, int ncut, const double* ph, Om om, const double* taus,
                              double um, double mu0, double tu, double F0) {
  double ans = 0.0, e0 = exp(-tu / mu0);
  if (fabs(um + mu0) <= dither) {
    for (int l = 1; l <= layru - 1; ++l) ans += om(l) * ph[l] * (taus[l] - taus[l - 1]);
    return F0 / (4.0 * kPi * mu0) * e0 * (ans + om(layru) * ph[layru] * (tu - taus[layru - 1]));
  }
  if (um > 0.0) {
    for (int l = layru; l <= ncut; ++l) {
      const double e1 = exp(-((taus[l] - tu) / um + taus[l] / mu0));
      ans += om(l) * ph[l] * (e0 - e1);
      e0 = e1;
    }
  } else {
    for (int l = layru; l >= 1; --l) {
      const double e1 = exp(-((taus[l - 1] - tu) / um + taus[l - 1] / mu0));
      ans += om(l) * ph[l] * (e0 - e1);
      e0 = e1;
    }
  }
  return ans * F0 / (4.0 * kPi * (1.0 + um / mu0));
}

// xi function (src/DisortSolver.cpp:3274-3285).
__device__ double xi_func(double u1, double u2, double tau) {
  const double x1 = (u2 - u1) / (u2 * u1);
  const double ex1 = exp(-tau / u1);
  return (x1 != 0.0) ? ((tau * x1 - 1.0) * exp(-tau / u2) + ex1) / (x1 * x1 * u1 * u2)
                     : tau * tau * ex1 / (2.0 * u1 * u2);
}

// Gate of the intensity corrections (src/DisortSolver.cpp:364-368): Delta-M
// (not Delta-M+) applied, beam present, scattering present.
__device__ bool correction_gate(const GenArgs& a, int64_t i) {
  const int scal_kind = a.inst_i[i * kInstI + II_SCALING];
  if (!(scal_kind & 1) || !(a.scal[i * DISORT_NSCAL + DISORT_S_BEAM_FLUX] > 0.0)) return false;
  bool scat = false;
  for (int lc = 0; lc < a.nlyr; ++lc) scat = scat || a.ssa[i * a.nlyr + lc] > 0.0;
  return scat;
}

// ===========================================================================
// k_gnt: Nakajima-Tanaka correction (src/DisortSolver.cpp:3274-3524), block
// per (chunk column, user angle, azimuth); threads over layers, then levels.
// Dynamic LDS: phast[L+1], phasm[L+1], pl[nmom+1].
// ===========================================================================
__global__ __launch_bounds__(64) void k_gnt(GenArgs a) {
  extern __shared__ double dyn[];
  const int L = a.nlyr, numu = a.numu, nphi = a.nphi, P = a.pstride, N = a.nstr;
  const int nmom = P - 1;
  const int64_t bk = blockIdx.x;
  const int64_t cl = bk / ((int64_t)numu * nphi);
  const int iu = (int)(bk / nphi % numu), jp = (int)(bk % nphi);
  if (cl >= a.nc) return;
  const int64_t i = a.c0 + cl;
  const int64_t lo = i * L, vo = i * (L + 1), TL = a.n * L, TV = a.n * (L + 1);
  const double* sc = a.scal + i * DISORT_NSCAL;
  const double F0 = sc[DISORT_S_BEAM_FLUX], mu0 = sc[DISORT_S_BEAM_MU];
  const int ncut = a.inst_i[i * kInstI + II_NCUT];
  const bool lyrcut = a.inst_i[i * kInstI + II_LYRCUT] != 0;
  if (!correction_gate(a, i)) return;
  double* phast = dyn;
  double* phasm = dyn + (L + 1);
  double* pl = dyn + 2 * (L + 1);
  const double deg = acos(-1.0) / 180.0;
  const double um = a.mu_user[iu];
  const double phirad = (a.phi_user[jp] - sc[DISORT_S_BEAM_PHI]) * deg;
  const double ct = -mu0 * um + sqrt((1.0 - mu0 * mu0) * (1.0 - um * um)) * cos(phirad);
  if (threadIdx.x == 0) {
    double plm1 = 1.0, plm2 = 0.0;
    pl[0] = 1.0;
    for (int k = 1; k <= nmom; ++k) {
      const double p = ((2 * k - 1) * ct * plm1 - (k - 1) * plm2) / k;
      plm2 = plm1;
      plm1 = p;
      pl[k] = p;
    }
  }
  __syncthreads();
  for (int lc = 1 + threadIdx.x; lc <= ncut; lc += blockDim.x) {
    const double* pm = a.pmom + (lo + lc - 1) * P;
    const double fl = a.lay_d[LD_F * TL + lo + lc - 1];
    double pa = 1.0, pmm = 1.0;
    for (int k = 1; k <= nmom; ++k) {
      pa += (double)(2 * k + 1) * pl[k] * pm[k];
      if (k <= N - 1) pmm += (double)(2 * k + 1) * pl[k] * (pm[k] - fl) / (1.0 - fl);
    }
    phast[lc] = pa / (1.0 - fl * a.ssa[lo + lc - 1]);
    phasm[lc] = pmm;
  }
  __syncthreads();
  const double* taus = a.lev_d + VD_TAU_S * TV + vo;
  const double* tauc = a.lev_d + VD_TAUC * TV + vo;
  // omega arrays (1-based): unscaled and scaled albedo.
  auto om_unscaled = [&](int l) { return a.ssa[lo + l - 1]; };
  auto om_scaled = [&](int l) { return a.lay_d[LD_SSA_S * TL + lo + l - 1]; };
  const double theta0 = um < 0.0 ? acos(-mu0) / deg : 0.0;
  const double thetap = um < 0.0 ? acos(um) / deg : 0.0;
  const bool ims = um < 0.0 && fabs(theta0 - thetap) <= 10.0;
  const double u0 = (a.flags & DISORT_G_USER_TAU) ? a.tau_user[i * a.ntau] : tauc[0];
  const int ltau = (u0 <= 100.0 * kEps) ? 1 : 0;
  for (int lu = threadIdx.x; lu < a.ntau; lu += blockDim.x) {
    const int lyr = a.u_lyr[i * a.ntau + lu];
    if (lyrcut && lyr >= ncut) continue;
    const double tu = a.u_tau[i * a.ntau + lu];
    const int layru = lyr + 1;
    const double dither = 100.0 * kEps;
    double corr = single_scat(dither, layru, ncut, phast, om_unscaled, taus, um, mu0, tu, F0) -
                  single_scat(dither, layru, ncut, phasm, om_scaled, taus, um, mu0, tu, F0);
    if (ims && lu >= ltau) {
      // secondary_scat (src/DisortSolver.cpp:3330-3401) with unscaled depths.
      const double tiny = 1.0e-4;
      const double ut = (a.flags & DISORT_G_USER_TAU) ? a.tau_user[i * a.ntau + lu] : tauc[lu];
      double dtl = ut - tauc[lyr];
      double wbar = a.ssa[lo + lyr] * dtl, fbar = a.lay_d[LD_F * TL + lo + lyr] * wbar, stau = dtl;
      for (int l = 0; l < lyr; ++l) {
        wbar += a.dtau[lo + l] * a.ssa[lo + l];
        fbar += a.dtau[lo + l] * a.ssa[lo + l] * a.lay_d[LD_F * TL + lo + l];
        stau += a.dtau[lo + l];
      }
      double ims_v = 0.0;
      if (!(wbar <= tiny || fbar <= tiny || stau <= tiny || F0 <= tiny)) {
        fbar /= wbar;
        wbar /= stau;
        double pspike = 1.0, gbar = 1.0, plm1 = 1.0, plm2 = 0.0;
        for (int k = 1; k <= N - 1; ++k) {
          const double p = ((2 * k - 1) * ct * plm1 - (k - 1) * plm2) / k;
          plm2 = plm1;
          plm1 = p;
          pspike += gbar * (2.0 - gbar) * (double)(2 * k + 1) * p;
        }
        for (int k = N; k <= nmom; ++k) {
          const double p = ((2 * k - 1) * ct * plm1 - (k - 1) * plm2) / k;
          plm2 = plm1;
          plm1 = p;
          dtl = ut - tauc[lyr];
          gbar = a.pmom[(lo + lyr) * P + k] * a.ssa[lo + lyr] * dtl;
          for (int l = 0; l < lyr; ++l) gbar += a.pmom[(lo + l) * P + k] * a.ssa[lo + l] * a.dtau[lo + l];
          const double tt = fbar * wbar * stau;
          gbar = (tt <= tiny) ? 0.0 : gbar / tt;
          pspike += gbar * (2.0 - gbar) * (double)(2 * k + 1) * p;
        }
        const double umu0p = mu0 / (1.0 - fbar * wbar);
        ims_v = F0 / (4.0 * kPi) * (fbar * wbar) * (fbar * wbar) / (1.0 - fbar * wbar) * pspike *
                xi_func(-um, umu0p, ut);
      }
      corr -= ims_v;
    }
    a.inten[((i * a.ntau + lu) * numu + iu) * nphi + jp] += corr;
  }
}

// ===========================================================================
// Buras-Emde correction (src/DisortSolver.cpp:3532-3745, :3812-4019): exact
// single scattering from the tabulated phase function replaces the Delta-M
// one (TMS), and near the beam direction a secondary-scattering term from
// the residual phase function phas2 = (P - P_trunc) / (f omega) is
// subtracted (IMS).
// ===========================================================================

// Table interval j of x in the increasing table xx[0..n) -- the bisection of
// locate (src/Utils.hpp:27-64), clamped to an interval of the table (at x
// beyond the last entry the reference reads one past the table).
__device__ int phase_interval(const double* xx, int n, double x) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (x >= xx[mid]) lo = mid;
    else hi = mid;
  }
  if (x == xx[0]) lo = 0;
  else if (x == xx[n - 1]) lo = n - 2;
  return lo < 0 ? 0 : (lo > n - 2 ? n - 2 : lo);
}

// Some downward user angle within 10 degrees of the beam direction.
__device__ bool aureole_angles(const GenArgs& a, double mu0) {
  const double deg = kPi / 180.0, theta0 = acos(-mu0) / deg;
  for (int iu = 0; iu < a.numu; ++iu) {
    const double um = a.mu_user[iu];
    if (um < 0.0 && fabs(theta0 - acos(um) / deg) <= 10.0) return true;
  }
  return false;
}

// k_gbe_prep: block per (chunk column, user level).  Threads over the table
// angles form phas2 before normalisation (sum over the layers above the
// level of (P - P_trunc) omega dtau, P_trunc the N-term Delta-M truncation
// not divided by 1 - f); lane 0 then normalises it (1 / fbar, then to unit
// half-integral 2) and builds the grid equidistant in the cumulative
// integral of |phas2| (prepDoubleScatIntegr).  Dynamic LDS: nphase doubles.
__global__ __launch_bounds__(64) void k_gbe_prep(GenArgs a) {
  extern __shared__ double cum[];
  const int np = a.nphase, L = a.nlyr, N = a.nstr, P = a.pstride;
  const int64_t cl = blockIdx.x / a.ntau;
  const int lu = (int)(blockIdx.x % a.ntau);
  if (cl >= a.nc) return;
  const int64_t i = a.c0 + cl;
  const double F0 = a.scal[i * DISORT_NSCAL + DISORT_S_BEAM_FLUX], mu0 = a.scal[i * DISORT_NSCAL + DISORT_S_BEAM_MU];
  if (!correction_gate(a, i) || !aureole_angles(a, mu0)) return;  // block-uniform
  const int64_t lo = i * L, vo = i * (L + 1), TL = a.n * L, TV = a.n * (L + 1);
  const double* tauc = a.lev_d + VD_TAUC * TV + vo;
  const int ll = a.u_lyr[i * a.ntau + lu];
  const double ut = (a.flags & DISORT_G_USER_TAU) ? a.tau_user[i * a.ntau + lu] : tauc[lu];
  const double* mp = a.mu_phase;
  const double* tab = a.phase + lo * np;  // [L][np] of this column
  double* ph = a.be_ph2 + (cl * a.ntau + lu) * np;
  for (int it = threadIdx.x; it < np; it += blockDim.x) {
    const double x = mp[it];
    double acc = 0.0;
    for (int lc = 0; lc <= ll; ++lc) {
      const double fl = a.lay_d[LD_F * TL + lo + lc];
      const double* pm = a.pmom + (lo + lc) * P;
      double tr = 1.0 - fl, p1 = 1.0, p2 = 0.0;
      for (int k = 1; k <= N - 1; ++k) {
        const double p = ((2 * k - 1) * x * p1 - (k - 1) * p2) / k;
        p2 = p1;
        p1 = p;
        tr += (2 * k + 1) * p * (pm[k] - fl);
      }
      const double w = (lc < ll) ? a.dtau[lo + lc] : ut - tauc[ll];
      acc += (tab[(int64_t)lc * np + it] - tr) * a.ssa[lo + lc] * w;
    }
    ph[it] = acc;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double tiny = 1.0e-4;
  double fbar = a.lay_d[LD_F * TL + lo + ll] * a.ssa[lo + ll] * (ut - tauc[ll]);
  for (int lc = 0; lc < ll; ++lc) fbar += a.ssa[lo + lc] * a.dtau[lo + lc] * a.lay_d[LD_F * TL + lo + lc];
  const double sc = (fbar <= tiny || F0 <= tiny) ? 0.0 : 1.0 / fbar;
  for (int it = 0; it < np; ++it) ph[it] *= sc;
  double area = 0.0;
  for (int it = 1; it < np; ++it) area += (mp[it] - mp[it - 1]) * 0.5 * (ph[it] + ph[it - 1]);
  if (area != 0.0) {
    const double nrm = 2.0 / area;
    for (int it = 0; it < np; ++it) ph[it] *= nrm;
  }
  cum[0] = 0.0;
  for (int it = 1; it < np; ++it) cum[it] = cum[it - 1] + (mp[it] - mp[it - 1]) * 0.5 * (fabs(ph[it]) + fabs(ph[it - 1]));
  double* grid = a.be_grid + (cl * a.ntau + lu) * kBeNF;
  int32_t* neg = a.be_neg + (cl * a.ntau + lu) * kBeNF;
  const double df = cum[np - 1] / (kBeNF - 1);
  grid[0] = -1.0;
  neg[0] = ph[0] > 0.0 ? 0 : 1;
  double target = 0.0;
  int it = 0;
  for (int q = 1; q < kBeNF - 1; ++q) {
    target += df;
    while (it < np - 2 && cum[it + 1] < target) ++it;
    const double w = (target - cum[it]) / (cum[it + 1] - cum[it]);
    grid[q] = mp[it] + w * (mp[it + 1] - mp[it]);
    if (ph[it] > 0.0 && ph[it + 1] > 0.0) neg[q] = 0;
    else if (ph[it] < 0.0 && ph[it + 1] < 0.0) neg[q] = 1;
    else neg[q] = (ph[it] + w * (ph[it + 1] - ph[it])) > 0.0 ? 0 : 1;
  }
  grid[kBeNF - 1] = 1.0;
  neg[kBeNF - 1] = ph[np - 1] > 0.0 ? 0 : 1;
  a.be_norm[cl * a.ntau + lu] = cum[np - 1] / ((kBeNF - 1) * kPi);
}

// p" * p" at scattering-angle cosine ct: azimuthal integral of the
// piecewise-linear phas2 over the equal-area grid (calcPhaseSquared).
__device__ double phase_squared(const double* mp, int np, const double* ph, const double* grid, const int32_t* neg,
                                double norm, double ct) {
  const double st = sqrt(1.0 - ct * ct);
  double sum = 0.0;
  for (int j = 0; j < kBeNF; ++j) {
    const double u = grid[j];
    double phint = 0.0;
    if (ct == 1.0 || u == 1.0) {
      const int k = phase_interval(mp, np, u * ct);
      phint = kPi * (ph[k] + (u * ct - mp[k]) / (mp[k + 1] - mp[k]) * (ph[k + 1] - ph[k]));
      if (ct == 1.0) phint /= 2.0;
    } else {
      const double su = sqrt(1.0 - u * u);
      const double lo = ct * u - st * su;
      double hi = ct * u + st * su;
      const bool cut = u < hi;
      if (cut) hi = u;
      if (lo < hi) {
        const int kmin = phase_interval(mp, np, lo), kmax = phase_interval(mp, np, hi);
        double D = (ph[kmin + 1] - ph[kmin]) / (mp[kmin + 1] - mp[kmin]);
        double C = ph[kmin] - mp[kmin] * D;
        phint += (D * ct * u + C) * kPi / 2.0;
        for (int k = kmin + 1; k <= kmax; ++k) {
          const double D2 = (ph[k + 1] - ph[k]) / (mp[k + 1] - mp[k]);
          const double C2 = ph[k] - mp[k] * D2;
          phint += (D2 - D) * sqrt(1.0 - ct * ct - u * u + 2.0 * ct * u * mp[k] - mp[k] * mp[k]) +
                   ((D2 - D) * ct * u + C2 - C) * asin((ct * u - mp[k]) / (su * st));
          D = D2;
          C = C2;
        }
        if (cut)
          phint += -D * sqrt(1.0 - ct * ct + 2.0 * u * u * (ct - 1.0)) - (D * ct * u + C) * asin((ct - 1.0) * u / (su * st));
        else
          phint += (D * ct * u + C) * kPi / 2.0;
      }
    }
    const double w = (j == 0 || j == kBeNF - 1) ? 0.5 : 1.0;
    sum += (neg[j] ? -w : w) * phint;
  }
  return sum * norm;
}

// k_gbe: block per (chunk column, user angle, azimuth), threads over layers
// (the two phase functions at the scattering angle), then over levels.
// Dynamic LDS: phast[L+1], phasm[L+1], pl[N].
__global__ __launch_bounds__(64) void k_gbe(GenArgs a) {
  extern __shared__ double dyn[];
  const int L = a.nlyr, numu = a.numu, nphi = a.nphi, P = a.pstride, N = a.nstr, np = a.nphase;
  const int64_t bk = blockIdx.x;
  const int64_t cl = bk / ((int64_t)numu * nphi);
  const int iu = (int)(bk / nphi % numu), jp = (int)(bk % nphi);
  if (cl >= a.nc) return;
  const int64_t i = a.c0 + cl;
  if (!correction_gate(a, i)) return;
  const int64_t lo = i * L, vo = i * (L + 1), TL = a.n * L, TV = a.n * (L + 1);
  const double* sc = a.scal + i * DISORT_NSCAL;
  const double F0 = sc[DISORT_S_BEAM_FLUX], mu0 = sc[DISORT_S_BEAM_MU];
  const int ncut = a.inst_i[i * kInstI + II_NCUT];
  const bool lyrcut = a.inst_i[i * kInstI + II_LYRCUT] != 0;
  double* phast = dyn;
  double* phasm = dyn + (L + 1);
  double* pl = dyn + 2 * (L + 1);
  const double deg = kPi / 180.0;
  const double um = a.mu_user[iu];
  const double phirad = (a.phi_user[jp] - sc[DISORT_S_BEAM_PHI]) * deg;
  const double ct = -mu0 * um + sqrt((1.0 - mu0 * mu0) * (1.0 - um * um)) * cos(phirad);
  const double* mp = a.mu_phase;
  const int kt = phase_interval(mp, np, ct);
  if (threadIdx.x == 0) {
    double p1 = 1.0, p2 = 0.0;
    pl[0] = 1.0;
    for (int k = 1; k <= N - 1; ++k) {
      const double p = ((2 * k - 1) * ct * p1 - (k - 1) * p2) / k;
      p2 = p1;
      p1 = p;
      pl[k] = p;
    }
  }
  __syncthreads();
  for (int lc = 1 + threadIdx.x; lc <= ncut; lc += blockDim.x) {
    const double* tab = a.phase + (lo + lc - 1) * np;
    const double fl = a.lay_d[LD_F * TL + lo + lc - 1];
    const double pa = tab[kt] + (ct - mp[kt]) / (mp[kt + 1] - mp[kt]) * (tab[kt + 1] - tab[kt]);
    phast[lc] = pa / (1.0 - fl * a.ssa[lo + lc - 1]);
    const double* pm = a.pmom + (lo + lc - 1) * P;
    double pmm = 1.0;
    for (int k = 1; k <= N - 1; ++k) pmm += (2.0 * k + 1) * pl[k] * (pm[k] - fl) / (1.0 - fl);
    phasm[lc] = pmm;
  }
  __syncthreads();
  const double* taus = a.lev_d + VD_TAU_S * TV + vo;
  const double* tauc = a.lev_d + VD_TAUC * TV + vo;
  auto om_unscaled = [&](int l) { return a.ssa[lo + l - 1]; };
  auto om_scaled = [&](int l) { return a.lay_d[LD_SSA_S * TL + lo + l - 1]; };
  const bool ims = um < 0.0 && fabs(acos(-mu0) / deg - acos(um) / deg) <= 10.0;
  const double u0 = (a.flags & DISORT_G_USER_TAU) ? a.tau_user[i * a.ntau] : tauc[0];
  const int ltau = (u0 <= 100.0 * kEps) ? 1 : 0;
  const double dither = 100.0 * kEps;
  for (int lu = threadIdx.x; lu < a.ntau; lu += blockDim.x) {
    const int lyr = a.u_lyr[i * a.ntau + lu];
    if (lyrcut && lyr >= ncut) continue;
    const double tu = a.u_tau[i * a.ntau + lu];
    double corr = single_scat(dither, lyr + 1, ncut, phast, om_unscaled, taus, um, mu0, tu, F0) -
                  single_scat(dither, lyr + 1, ncut, phasm, om_scaled, taus, um, mu0, tu, F0);
    if (ims && lu >= ltau) {
      // newSecondaryScat (src/DisortSolver.cpp:3905-3945), unscaled depths.
      const double tiny = 1.0e-4;
      const double ut = (a.flags & DISORT_G_USER_TAU) ? a.tau_user[i * a.ntau + lu] : tauc[lu];
      const double dtl = ut - tauc[lyr];
      double wbar = a.ssa[lo + lyr] * dtl, fbar = a.lay_d[LD_F * TL + lo + lyr] * wbar, stau = dtl;
      for (int l = 0; l < lyr; ++l) {
        wbar += a.dtau[lo + l] * a.ssa[lo + l];
        fbar += a.dtau[lo + l] * a.ssa[lo + l] * a.lay_d[LD_F * TL + lo + l];
        stau += a.dtau[lo + l];
      }
      if (!(wbar <= tiny || fbar <= tiny || stau <= tiny || F0 <= tiny)) {
        fbar /= wbar;
        wbar /= stau;
        const int64_t r = cl * a.ntau + lu;
        const double* ph = a.be_ph2 + r * np;
        const double p1 = ph[kt] + (ct - mp[kt]) / (mp[kt + 1] - mp[kt]) * (ph[kt + 1] - ph[kt]);
        const double pspike =
            2.0 * p1 - phase_squared(mp, np, ph, a.be_grid + r * kBeNF, a.be_neg + r * kBeNF, a.be_norm[r], ct);
        const double fw = fbar * wbar;
        corr -= F0 / (4.0 * kPi) * fw * fw / (1.0 - fw) * pspike * xi_func(-um, mu0 / (1.0 - fw), ut);
      }
    }
    a.inten[((i * a.ntau + lu) * numu + iu) * nphi + jp] += corr;
  }
}

// ===========================================================================
// k_gbrdf: BRDF beam-reflection correction (src/DisortSolver.cpp:3751-3808),
// thread per (chunk column, user angle, azimuth).
// ===========================================================================
template <int NN>
__global__ __launch_bounds__(64) void k_gbrdf(GenArgs a, const int32_t* modes_used) {
  const int64_t w = (int64_t)blockIdx.x * 64 + threadIdx.x;
  const int numu = a.numu, nphi = a.nphi, L = a.nlyr;
  if (w >= a.nc * numu * nphi) return;
  const int64_t cl = w / ((int64_t)numu * nphi);
  const int iu = (int)(w / nphi % numu), j = (int)(w % nphi);
  const int64_t i = a.c0 + cl;
  const double* sc = a.scal + i * DISORT_NSCAL;
  const double F0 = sc[DISORT_S_BEAM_FLUX], mu0 = sc[DISORT_S_BEAM_MU];
  const double mu = a.mu_user[iu];
  if (!(F0 > 0.0) || mu <= 0.0) return;
  if (a.inst_i[i * kInstI + II_SCALING] & 2) return;  // Delta-M+ active
  const double* p = a.set_rpv + (int64_t)a.rpv_set[i] * DISORT_NRPV;
  const double dphi = (a.phi_user[j] - sc[DISORT_S_BEAM_PHI]) * (kPi / 180.0);
  const double exact = brdf_rho(a.brdf_kind, p, mu, mu0, dphi);
  if (exact < 0.0) return;
  double approx = 0.0;
  const int mu_ = modes_used[cl];
  for (int m = 0; m < mu_; ++m)
    approx += a.srf_b[(cl * a.nm + m) * surf_b_size(NN, numu) + NN + iu] * cos((double)m * dphi);
  const double dr = exact - approx;
  if (fabs(dr) < 1e-15) return;
  const int64_t vo = i * (L + 1), TV = a.n * (L + 1);
  const int ncut = a.inst_i[i * kInstI + II_NCUT];
  const bool lyrcut = a.inst_i[i * kInstI + II_LYRCUT] != 0;
  const double tn = a.lev_d[VD_TAU_S * TV + vo + ncut];
  const double to_bottom = exp(-tn / mu0);
  for (int lu = 0; lu < a.ntau; ++lu) {
    if (lyrcut && a.u_lyr[i * a.ntau + lu] >= ncut) continue;
    a.inten[((i * a.ntau + lu) * numu + iu) * nphi + j] +=
        mu0 * F0 * dr * to_bottom * exp((a.u_tau[i * a.ntau + lu] - tn) / mu);
  }
}


template <int NN>
void general_setup_n(const FluxArgs& fa, const GenArgs& a, hipStream_t s, bool brdf) {
  constexpr int N = 2 * NN;
  hipLaunchKernelGGL(k_column<N>, dim3((unsigned)fa.nb), dim3(64), 0, s, fa);
  hipLaunchKernelGGL(k_glevels, dim3((unsigned)((a.n + 63) / 64)), dim3(64), 0, s, a);
  if (brdf) {
    const int64_t ne = (int64_t)a.nsets * (NN + a.numu) * NN;
    hipLaunchKernelGGL(k_gsurf_c<NN>, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, a);
    if (a.flags & DISORT_F_THERMAL)
      hipLaunchKernelGGL(k_gsurf_e<NN>, dim3((unsigned)(a.nsets * (NN + a.numu))), dim3(256), 0, s, a);
  }
}

template <int NN>
void general_chunk_n(const GenArgs& a, const GenLaunch& g, hipStream_t s) {
  using C = WideCfg<NN>;
  const unsigned ncm = (unsigned)(a.nc * a.nm);
  auto rec = [&](int k) {
    if (g.ev) (void)hipEventRecord(g.ev[k], s);
  };
  rec(0);
  if (g.brdf) hipLaunchKernelGGL(k_gsurf_b<NN>, dim3((unsigned)a.nc), dim3(64), 0, s, a);
  rec(1);
  hipLaunchKernelGGL(k_glayer<NN>, dim3((unsigned)a.n_tasks), dim3(kGLT), 0, s, a);
  rec(2);
  if constexpr (3 * NN <= 64) hipLaunchKernelGGL(k_gboundary1<NN>, dim3(ncm), dim3(64), 0, s, a);
  else hipLaunchKernelGGL(k_gboundary<NN>, dim3(ncm), dim3(C::BT), 0, s, a);
  rec(3);
  hipLaunchKernelGGL(k_gflux<NN>, dim3((unsigned)((a.nc * a.ntau + 63) / 64)), dim3(64), 0, s, a);
  rec(4);
  if (g.do_user) {
    const int tpb = ucoef_tasks_per_block(a.numu);
    hipLaunchKernelGGL(k_gucoef<NN>, dim3((unsigned)((a.n_tasks + tpb - 1) / tpb)), dim3(64), 0, s, a, g.ucoef, tpb);
    rec(5);
    hipLaunchKernelGGL(k_guser<NN>, dim3(ncm), dim3(256), 0, s, a, (const double*)g.ucoef);
    rec(6);
    hipLaunchKernelGGL(k_gaccum, dim3((unsigned)a.nc), dim3(256), 0, s, a, g.modes_used);
    rec(7);
    if (g.nt && g.buras) {
      hipLaunchKernelGGL(k_gbe_prep, dim3((unsigned)(a.nc * a.ntau)), dim3(64), sizeof(double) * (size_t)a.nphase,
                         s, a);
      const size_t lds = sizeof(double) * (2 * (size_t)(a.nlyr + 1) + (size_t)a.nstr);
      hipLaunchKernelGGL(k_gbe, dim3((unsigned)(a.nc * a.numu * a.nphi)), dim3(64), lds, s, a);
    } else if (g.nt) {
      const size_t lds = sizeof(double) * (2 * (size_t)(a.nlyr + 1) + (size_t)a.pstride);
      hipLaunchKernelGGL(k_gnt, dim3((unsigned)(a.nc * a.numu * a.nphi)), dim3(64), lds, s, a);
    }
    if (g.brdf_corr)
      hipLaunchKernelGGL(k_gbrdf<NN>, dim3((unsigned)((a.nc * a.numu * a.nphi + 63) / 64)), dim3(64), 0, s, a,
                         (const int32_t*)g.modes_used);
    rec(8);
  }
}

}  // namespace

bool general_setup(int nstr, const FluxArgs& fa, const GenArgs& a, hipStream_t s, bool brdf) {
  switch (nstr) {
    case 4: general_setup_n<2>(fa, a, s, brdf); return true;
    case 6: general_setup_n<3>(fa, a, s, brdf); return true;
    case 8: general_setup_n<4>(fa, a, s, brdf); return true;
    case 10: general_setup_n<5>(fa, a, s, brdf); return true;
    case 12: general_setup_n<6>(fa, a, s, brdf); return true;
    case 14: general_setup_n<7>(fa, a, s, brdf); return true;
    case 16: general_setup_n<8>(fa, a, s, brdf); return true;
    case 32: general_setup_n<16>(fa, a, s, brdf); return true;
    case 64: general_setup_n<32>(fa, a, s, brdf); return true;
    default: return false;
  }
}

bool general_chunk(int nstr, const GenArgs& a, const GenLaunch& g, hipStream_t s) {
  switch (nstr) {
    case 4: general_chunk_n<2>(a, g, s); return true;
    case 6: general_chunk_n<3>(a, g, s); return true;
    case 8: general_chunk_n<4>(a, g, s); return true;
    case 10: general_chunk_n<5>(a, g, s); return true;
    case 12: general_chunk_n<6>(a, g, s); return true;
    case 14: general_chunk_n<7>(a, g, s); return true;
    case 16: general_chunk_n<8>(a, g, s); return true;
    case 32: general_chunk_n<16>(a, g, s); return true;
    case 64: general_chunk_n<32>(a, g, s); return true;
    default: return false;
  }
}

}  // namespace dsa
