// Oracle numerics: Planck band integral, Legendre recurrences, double-Gauss
// quadrature, Chapman path integral, LINPACK-semantics band LU and the small
// dense kernels (Cholesky, Jacobi symmetric eigensolver, partial-pivot LU).
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include "oracle.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>

namespace ORACLE_NS {

// ---------------------------------------------------------------------------
// Planck band integral.  Restates src/Planck.cpp:35-169 (planckFunction and
// planckFunction2): same physical constants (src/Planck.hpp:10-22, c1 at
// src/Planck.cpp:28), the same three branches (Simpson refinement when the
// band is narrow, power series below v=1.5, exponential series above) and the
// same convergence test, so that the returned values agree to the last bits.
// ---------------------------------------------------------------------------
namespace {
constexpr Real kC2 = 1.438786;        // hc/k [cm K]
constexpr Real kSigma = 5.67032e-8;   // Stefan-Boltzmann [W m^-2 K^-4]
constexpr Real kVcut = 1.5;
constexpr Real kA1 = 1.0 / 3.0, kA2 = -1.0 / 8.0, kA3 = 1.0 / 60.0;
constexpr Real kA4 = -1.0 / 5040.0, kA5 = 1.0 / 272160.0, kA6 = -1.0 / 13305600.0;
constexpr Real kC1Mono = 1.1911e-8;

inline Real shape(Real v) { return v * v * v / (std::exp(v) - 1.0); }

Real planck_band_strict(Real lo, Real hi, Real t) {
  if (t < 0.0 || hi <= lo || lo < 0.0)
    throw std::invalid_argument("planckFunction: Invalid arguments (temp=" + std::to_string(t) +
                                ", wnumlo=" + std::to_string(lo) + ", wnumhi=" + std::to_string(hi) + ")");
  if (t < 1.e-4) return 0.0;
  const Real sigdpi = kSigma / M_PI;
  const Real conc = 15.0 / std::pow(M_PI, 4.0);
  const Real vmax = std::log(std::numeric_limits<double>::max());
  const Real v0 = kC2 * lo / t, v1 = kC2 * hi / t;

  if (v0 > std::numeric_limits<double>::epsilon() && v1 < vmax && (hi - lo) / hi < 1.e-2) {
    // Composite Simpson with 2n panels, n = 1..10, stop at 1e-6 relative change.
    const Real span = v1 - v0;
    const Real ends = shape(v0) + shape(v1);
    Real prev = 0.0;
    for (int n = 1; n <= 10; ++n) {
      const Real h = span / (2.0 * n);
      Real acc = ends;
      for (int k = 1; k <= 2 * n - 1; ++k) acc += 2.0 * (1 + k % 2) * shape(v0 + k * h);
      acc *= h * kA1;
      if (std::abs((acc - prev) / acc) <= 1.e-6) return sigdpi * std::pow(t, 4.0) * conc * acc;
      prev = acc;
    }
    return sigdpi * std::pow(t, 4.0) * conc * prev;
  }

  static const Real vcp[7] = {10.25, 5.7, 3.9, 2.9, 2.3, 1.9, 0.0};
  Real p[2] = {0.0, 0.0}, d[2] = {0.0, 0.0};
  const Real v[2] = {v0, v1};
  int nsmall = 0;
  for (int i = 0; i < 2; ++i) {
    if (v[i] < kVcut) {
      ++nsmall;
      const Real vs = v[i] * v[i];
      p[i] = conc * vs * v[i] * (kA1 + v[i] * (kA2 + v[i] * (kA3 + vs * (kA4 + vs * (kA5 + vs * kA6)))));
    } else {
      int terms = 1;
      while (v[i] < vcp[terms - 1]) ++terms;
      const Real ex = std::exp(-v[i]);
      Real exm = 1.0, s = 0.0;
      for (int m = 1; m <= terms; ++m) {
        const Real mv = m * v[i];
        exm *= ex;
        s += exm * (6.0 + mv * (6.0 + mv * (3.0 + mv))) / (m * m * m * m);
      }
      d[i] = s * conc;
    }
  }
  Real ans;
  if (nsmall == 2) ans = p[1] - p[0];
  else if (nsmall == 1) ans = 1.0 - p[0] - d[1];
  else ans = d[0] - d[1];
  ans *= sigdpi * std::pow(t, 4.0);
  return ans == 0.0 ? 0.0 : ans;
}
}  // namespace

namespace {
int g_pstage = PS_NONE;
double g_peps = 0.0;
}  // namespace

void set_perturbation(int stage, double eps) {
  g_pstage = stage;
  g_peps = eps;
}

void perturb(int stage, std::vector<Real>& v) {
  if (stage != g_pstage || g_pstage == PS_NONE) return;
  static const double w[4] = {-1.0, -1.0 / 3.0, 1.0 / 3.0, 1.0};
  for (size_t i = 0; i < v.size(); ++i) v[i] *= 1.0 + g_peps * w[i % 4];
}

Real perturb1(int stage, Real x) {
  return (stage == g_pstage && g_pstage != PS_NONE) ? x * (1.0 + g_peps) : x;
}

Real planck_band(Real lo, Real hi, Real t) { return perturb1(PS_PLANCK, planck_band_exact(lo, hi, t)); }

Real planck_band_exact(Real lo, Real hi, Real t) {
  // src/Planck.cpp:146-169
  if (t < 0.0 || hi < lo || lo < 0.0)
    throw std::invalid_argument("planckFunction2: Invalid arguments (temp=" + std::to_string(t) +
                                ", wnumlo=" + std::to_string(lo) + ", wnumhi=" + std::to_string(hi) + ")");
  if (t < 1.e-4) return 0.0;
  if (hi == lo) {
    const Real arg = std::exp(-kC2 * hi / t);
    return kC1Mono * hi * hi * hi * arg / (1.0 - arg);
  }
  return planck_band_strict(lo, hi, t);
}

Real planck_band_deriv(Real lo, Real hi, Real t) {
  // src/Planck.cpp:171-214: dB/dT = 4B/T - (sigma/pi)(15/pi^4) T^3 (v1 f(v1) - v0 f(v0)).
  if (t < 0.0 || hi < lo || lo < 0.0)
    throw std::invalid_argument("planckFunctionDeriv2: Invalid arguments (temp=" + std::to_string(t) +
                                ", wnumlo=" + std::to_string(lo) + ", wnumhi=" + std::to_string(hi) + ")");
  if (t < 1.e-4) return 0.0;
  auto f = [](Real v) {
    if (v < 1.e-6) return v * v;
    const Real em = std::exp(-v);
    return v * v * v * em / (1.0 - em);
  };
  if (hi == lo) {
    const Real x = kC2 * hi / t;
    const Real em = std::exp(-x);
    const Real b = kC1Mono * hi * hi * hi * em / (1.0 - em);
    return b * (x / t) / (1.0 - em);
  }
  const Real sigdpi = kSigma / M_PI;
  const Real conc = 15.0 / std::pow(M_PI, 4.0);
  const Real v0 = kC2 * lo / t, v1 = kC2 * hi / t;
  const Real b = planck_band(lo, hi, t);
  return 4.0 * b / t - sigdpi * conc * t * t * t * (v1 * f(v1) - v0 * f(v0));
}

// ---------------------------------------------------------------------------
// Normalised associated Legendre functions by upward recurrence, flat output
// ylm[l*(maxmu+1)+i].  Restates src/Legendre.cpp:67-120 (Dave & Armstrong
// recurrences); for m>0 the caller must already hold row m-1.
// ---------------------------------------------------------------------------
void legendre_flat(int nmu, int m, int maxmu, int twonm1, const Real* mu, Real* ylm) {
  const int ld = maxmu + 1;
  if (m == 0) {
    for (int i = 0; i < nmu; ++i) {
      ylm[i] = 1.0;
      ylm[ld + i] = mu[i];
    }
    for (int l = 2; l <= twonm1; ++l)
      for (int i = 0; i < nmu; ++i)
        ylm[l * ld + i] = ((2.0 * l - 1.0) * mu[i] * ylm[(l - 1) * ld + i] - (l - 1.0) * ylm[(l - 2) * ld + i]) / l;
    return;
  }
  for (int i = 0; i < nmu; ++i) {
    ylm[m * ld + i] = -std::sqrt((1.0 - 1.0 / (2.0 * m)) * (1.0 - mu[i] * mu[i])) * ylm[(m - 1) * ld + i];
    ylm[(m + 1) * ld + i] = std::sqrt(2.0 * m + 1.0) * mu[i] * ylm[m * ld + i];
  }
  for (int l = m + 2; l <= twonm1; ++l) {
    const Real a = std::sqrt(static_cast<Real>((l - m) * (l + m)));
    const Real b = std::sqrt(static_cast<Real>((l - m - 1) * (l + m - 1)));
    for (int i = 0; i < nmu; ++i)
      ylm[l * ld + i] = ((2.0 * l - 1.0) * mu[i] * ylm[(l - 1) * ld + i] - b * ylm[(l - 2) * ld + i]) / a;
  }
}

// ---------------------------------------------------------------------------
// Double-Gauss quadrature: Gauss-Legendre roots on (-1,1) by Newton iteration
// (tolerance 3e-14, at most 10 steps), mapped to (0,1), mirrored to negative
// cosines.  Restates src/FluxSolver.cpp:257-290.
// ---------------------------------------------------------------------------
void gauss_quadrature(int n, Real* cmu, Real* cwt) {
  for (int i = 1; i <= n; ++i) {
    Real z = std::cos(M_PI * (i - 0.25) / (n + 0.5));
    Real z1, dp;
    int it = 0;
    do {
      Real p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        const Real p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      dp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / dp;
      if (++it > 10) throw std::runtime_error("gaussianQuadrature: Failed to converge");
    } while (std::abs(z - z1) > 3.0e-14);
    cmu[i - 1] = 0.5 * (1.0 + z);
    cwt[i - 1] = 1.0 / ((1.0 - z * z) * dp * dp);
  }
  for (int i = 0; i < n; ++i) {
    cmu[n + i] = -cmu[i];
    cwt[n + i] = cwt[i];
  }
}

// ---------------------------------------------------------------------------
// Chapman slant-path factor for pseudo-spherical geometry.
// Restates src/FluxSolver.cpp:1503-1566 (1-based layer index lc).
// ---------------------------------------------------------------------------
int locate_sorted(const std::vector<Real>& xx, Real x) {
  // src/Utils.hpp:27-64: bisection on an increasing or decreasing table.
  const int n = (int)xx.size();
  if (n == 0) return -1;
  if (n == 1) return 0;
  const bool up = xx[1] > xx[0];
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (up ? (x >= xx[mid]) : (x <= xx[mid])) lo = mid;
    else hi = mid;
  }
  if (x == xx[0]) return 0;
  if (x == xx[n - 1]) return n - 2;
  return lo;
}

Real chapman(int lc, Real tau_frac, const Real* zd, const Real* dtau, int nlyr, Real r, Real mu0) {
  return perturb1(PS_CHAPMAN, chapman_exact(lc, tau_frac, zd, dtau, nlyr, r, mu0));
}

Real chapman_exact(int lc, Real tau_frac, const Real* zd, const Real* dtau, int nlyr, Real r, Real mu0) {
  const Real deg = M_PI / 180.0;
  const Real zenang = std::acos(mu0) * 180.0 / M_PI;
  const Real zenrad = zenang * deg;
  const Real xp = r + zd[lc] + (zd[lc - 1] - zd[lc]) * tau_frac;
  const Real xpsinz = xp * std::sin(zenrad);
  if (zenang > 90.0 && xpsinz < r) return 1.0e+20;
  int id = lc;
  if (zenang > 90.0)
    for (int j = lc; j <= nlyr; ++j)
      if (xpsinz < (zd[j - 1] + r) && xpsinz >= (zd[j] + r)) id = j;
  Real sum = 0.0;
  for (int j = 1; j <= id; ++j) {
    Real fact = 1.0, fact2 = 1.0;
    if (j > lc) fact = 2.0;
    else if (j == lc && lc == id && zenang > 90.0) fact2 = -1.0;
    const Real rj = r + zd[j - 1];
    Real rjp1 = r + zd[j];
    if (j == lc && id == lc) rjp1 = xp;
    const Real dhj = zd[j - 1] - zd[j];
    Real dsj;
    if (id > lc && j == id) dsj = std::sqrt(rj * rj - xpsinz * xpsinz);
    else dsj = std::sqrt(rj * rj - xpsinz * xpsinz) - fact2 * std::sqrt(rjp1 * rjp1 - xpsinz * xpsinz);
    sum += dtau[j - 1] * fact * dsj / dhj;
  }
  if (id > lc) {
    const Real dhj = zd[lc - 1] - zd[lc];
    const Real dsj = std::sqrt(xp * xp - xpsinz * xpsinz) - std::sqrt((zd[lc] + r) * (zd[lc] + r) - xpsinz * xpsinz);
    sum += dtau[lc - 1] * dsj / dhj;
  }
  return sum;
}

// ---------------------------------------------------------------------------
// Banded LU with partial pivoting (LINPACK xGBFA / xGBSL, job=0), as used by
// src/BandSolver.hpp:29-172.  Storage: column-major, lda >= 2*ml+mu+1 rows;
// A(i,j) (0-based) lives at band row ml+mu+i-j of column j; the top ml rows
// take the fill-in.  ipvt[k] holds the 1-based row of the pivot of column k.
// Pivot choice: the first entry of largest magnitude, as in IDAMAX.
// ---------------------------------------------------------------------------
void band_factor(Real* abd, int lda, int n, int ml, int mu, int* ipvt) {
  const int d = ml + mu;  // 0-based band row of the diagonal
  auto at = [&](int r, int c) -> Real& { return abd[r + (size_t)c * lda]; };
  // Clear the fill-in rows of the columns that the first pivots can reach.
  for (int c = mu + 1; c <= std::min(n, d + 1) - 2; ++c)
    for (int r = d - c; r < ml; ++r) at(r, c) = 0.0;
  int jz = std::min(n, d + 1) - 1;  // 1-based index of last cleared column
  int ju = 0;                        // 1-based furthest column touched so far
  for (int k = 0; k < n - 1; ++k) {
    ++jz;
    if (jz <= n && ml > 0)
      for (int r = 0; r < ml; ++r) at(r, jz - 1) = 0.0;
    const int lm = std::min(ml, n - 1 - k);
    int off = 0;
    Real best = std::abs(at(d, k));
    for (int i = 1; i <= lm; ++i)
      if (std::abs(at(d + i, k)) > best) { best = std::abs(at(d + i, k)); off = i; }
    ipvt[k] = k + off + 1;
    if (at(d + off, k) == 0.0) continue;
    if (off != 0) std::swap(at(d + off, k), at(d, k));
    const Real t = -1.0 / at(d, k);
    for (int i = 1; i <= lm; ++i) at(d + i, k) *= t;
    ju = std::min(std::max(ju, mu + ipvt[k]), n);
    int lrow = d + off, mrow = d;
    for (int j = k + 1; j < ju; ++j) {
      --lrow;
      --mrow;
      const Real s = at(lrow, j);
      if (lrow != mrow) {
        at(lrow, j) = at(mrow, j);
        at(mrow, j) = s;
      }
      if (s != 0.0)
        for (int i = 1; i <= lm; ++i) at(mrow + i, j) += s * at(d + i, k);
    }
  }
  ipvt[n - 1] = n;
}

void band_solve(const Real* abd, int lda, int n, int ml, int mu, const int* ipvt, Real* b) {
  const int d = ml + mu;
  auto at = [&](int r, int c) { return abd[r + (size_t)c * lda]; };
  if (ml > 0) {
    for (int k = 0; k < n - 1; ++k) {
      const int lm = std::min(ml, n - 1 - k);
      const int l = ipvt[k] - 1;
      const Real t = b[l];
      if (l != k) {
        b[l] = b[k];
        b[k] = t;
      }
      for (int i = 1; i <= lm; ++i) b[k + i] += t * at(d + i, k);
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    b[k] /= at(d, k);
    const int lm = std::min(k + 1, d + 1) - 1;
    const Real t = -b[k];
    for (int i = 1; i <= lm; ++i) b[k - i] += t * at(d - i, k);
  }
}

// ---------------------------------------------------------------------------
// Dense kernels (column-major).  The reference delegates these to Eigen 3.4
// (LLT, SelfAdjointEigenSolver, PartialPivLU); they are restated here from
// the textbook algorithms.
// ---------------------------------------------------------------------------
bool cholesky_lower(int n, const Real* a, Real* l) {
  // Reads the lower triangle of a; fails on a non-positive pivot (as Eigen's LLT).
  for (int i = 0; i < n * n; ++i) l[i] = 0.0;
  for (int j = 0; j < n; ++j) {
    Real x = a[j + j * n];
    for (int k = 0; k < j; ++k) x -= l[j + k * n] * l[j + k * n];
    if (!(x > 0.0)) return false;
    const Real ljj = std::sqrt(x);
    l[j + j * n] = ljj;
    for (int i = j + 1; i < n; ++i) {
      Real s = a[i + j * n];
      for (int k = 0; k < j; ++k) s -= l[i + k * n] * l[j + k * n];
      l[i + j * n] = s / ljj;
    }
  }
  return true;
}

void sym_eigen_jacobi(int n, const Real* a_in, Real* evals, Real* evecs) {
  // Cyclic Jacobi rotations (Rutishauser's formulation) on the lower triangle
  // mirrored into a full symmetric matrix; runs to full convergence.
  std::vector<Real> a(n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) a[i + j * n] = (i >= j) ? a_in[i + j * n] : a_in[j + i * n];
  std::vector<Real> v(n * n, 0.0);
  for (int i = 0; i < n; ++i) v[i + i * n] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    Real off = 0.0, diag = 0.0;
    for (int j = 0; j < n; ++j) {
      diag += a[j + j * n] * a[j + j * n];
      for (int i = j + 1; i < n; ++i) off += a[i + j * n] * a[i + j * n];
    }
    if (off == 0.0 || off <= 1e-36 * diag) break;
    for (int p = 0; p < n - 1; ++p) {
      for (int q = p + 1; q < n; ++q) {
        const Real apq = a[p + q * n];
        if (apq == 0.0) continue;
        const Real app = a[p + p * n], aqq = a[q + q * n];
        const Real theta = (aqq - app) / (2.0 * apq);
        Real t = 1.0 / (std::abs(theta) + std::sqrt(theta * theta + 1.0));
        if (theta < 0.0) t = -t;
        const Real c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // columns p,q
          const Real akp = a[k + p * n], akq = a[k + q * n];
          a[k + p * n] = c * akp - s * akq;
          a[k + q * n] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p,q
          const Real apk = a[p + k * n], aqk = a[q + k * n];
          a[p + k * n] = c * apk - s * aqk;
          a[q + k * n] = s * apk + c * aqk;
        }
        a[p + q * n] = a[q + p * n] = 0.0;
        for (int k = 0; k < n; ++k) {
          const Real vkp = v[k + p * n], vkq = v[k + q * n];
          v[k + p * n] = c * vkp - s * vkq;
          v[k + q * n] = s * vkp + c * vkq;
        }
      }
    }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return a[x + x * n] < a[y + y * n]; });
  for (int j = 0; j < n; ++j) {
    evals[j] = a[idx[j] + idx[j] * n];
    for (int i = 0; i < n; ++i) evecs[i + j * n] = v[i + idx[j] * n];
  }
}

void DenseLU::factor(int n_, const Real* a) {
  n = n_;
  lu.assign(a, a + (size_t)n * n);
  piv.assign(n, 0);
  Real fro = 0.0;
  for (int i = 0; i < n * n; ++i) fro += a[i] * a[i];
  frob_norm_of_input = std::sqrt(fro);
  for (int k = 0; k < n; ++k) {
    int p = k;
    Real best = std::abs(lu[k + k * n]);
    for (int i = k + 1; i < n; ++i)
      if (std::abs(lu[i + k * n]) > best) { best = std::abs(lu[i + k * n]); p = i; }
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(lu[k + j * n], lu[p + j * n]);
    const Real dkk = lu[k + k * n];
    if (dkk != 0.0)
      for (int i = k + 1; i < n; ++i) lu[i + k * n] /= dkk;
    for (int j = k + 1; j < n; ++j) {
      const Real ukj = lu[k + j * n];
      if (ukj != 0.0)
        for (int i = k + 1; i < n; ++i) lu[i + j * n] -= lu[i + k * n] * ukj;
    }
  }
}

void DenseLU::solve(Real* b) const {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k) std::swap(b[k], b[piv[k]]);
  for (int k = 0; k < n; ++k)
    for (int i = k + 1; i < n; ++i) b[i] -= lu[i + k * n] * b[k];
  for (int k = n - 1; k >= 0; --k) {
    b[k] /= lu[k + k * n];
    for (int i = 0; i < k; ++i) b[i] -= lu[i + k * n] * b[k];
  }
}

}  // namespace ORACLE_NS
