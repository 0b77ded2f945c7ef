// Device-side layouts of the general solver (DisortSolver::solve): kernel
// arguments, per-task records and constant tables.  A task is one
// (column, azimuthal mode, layer); a chunk of columns is solved for all its
// modes at once, its tasks numbered t = ((c * NM) + m) * L + lc for the
// chunk-local column c, mode m < NM and layer lc.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/disort_amd.h"

namespace dsa {

// Per-task record (doubles), AoS: eig[t * GE(N).size() + field].  G is the
// N x N eigenvector matrix in the reference's ordering
// (src/DisortSolver.cpp:1124-1138) held as its two NN x NN blocks like the
// flux kernels (flux_layout.hpp EigOff); vectors of length N are in the
// reference's row order (row NN+r is +mu_r, row NN-1-r is -mu_r).
struct GE {
  int n, nn;
  __host__ __device__ constexpr GE(int n_) : n(n_), nn(n_ / 2) {}
  __host__ __device__ constexpr int k() const { return 0; }                  // [NN] eigenvalues k_j > 0
  __host__ __device__ constexpr int gp() const { return nn; }                // [NN*NN] col-major
  __host__ __device__ constexpr int gm() const { return nn + nn * nn; }      // [NN*NN]
  __host__ __device__ constexpr int ztop() const { return nn + 2 * nn * nn; } // [N] particular sol. at layer top (BC form)
  __host__ __device__ constexpr int zbot() const { return ztop() + n; }      // [N] at layer bottom (BC form)
  __host__ __device__ constexpr int zq() const { return zbot() + n; }        // [N] beam intercept x transmission at the bottom
  __host__ __device__ constexpr int zb() const { return zq() + n; }          // [N] beam solution (intercept)
  __host__ __device__ constexpr int zbs() const { return zb() + n; }         // [N] beam slope (spherical)
  __host__ __device__ constexpr int z0() const { return zbs() + n; }         // [N] thermal intercept (m = 0)
  __host__ __device__ constexpr int z1() const { return z0() + n; }          // [N] thermal slope (m = 0)
  __host__ __device__ constexpr int xba() const { return z1() + n; }         // spherical attenuation 1/CH
  __host__ __device__ constexpr int xr0() const { return xba() + 1; }        // Planck B(tau) = xr0 + xr1 tau
  __host__ __device__ constexpr int xr1() const { return xba() + 2; }
  __host__ __device__ constexpr int size() const { return xba() + 3; }
};

// Sources at user angles per task, [t][3][numu]: beam, thermal intercept,
// thermal slope (src/DisortSolver.cpp:2506-2580).
enum { US_BEAM = 0, US_Z0 = 1, US_Z1 = 2, kUSrc = 3 };

// Surface reflection, BRDF case.  Diffuse part per (parameter set, mode):
// [NN*NN refl_quad(iq, jq) | numu*NN refl_user(iu, jq) | NN emis_quad | numu emis_user];
// beam part per (column, mode): [NN refl_quad(iq, 0) | numu refl_user(iu, 0)].
__host__ __device__ constexpr int surf_d_size(int nn, int numu) { return nn * nn + numu * nn + nn + numu; }
__host__ __device__ constexpr int surf_b_size(int nn, int numu) { return nn + numu; }

struct GenArgs {
  int32_t flags, nstr, nlyr, pstride;
  int32_t ntau, numu, nphi, nm;  // nm: modes computed per column (N, or 1)
  int64_t n;                     // columns in the batch
  int64_t c0, nc;                // this chunk: columns c0 .. c0+nc-1
  int64_t n_tasks;               // nc * nm * nlyr
  double accuracy;
  const double* dtau;
  const double* ssa;
  const double* pmom;
  const double* temper;
  const double* alt;
  const double* scal;
  const double* rpv;
  const double* tau_user;
  const double* mu_user;
  const double* phi_user;
  double* flux;
  double* inten;
  double* inten_avg;
  // Per-column records of k_column (flux_layout.hpp), batch-wide.
  double* lay_d;   // [kLayD][n*nlyr]
  double* lev_d;   // [kLevD][n*(nlyr+1)]
  double* inst_d;  // [n][kInstD]
  int32_t* inst_i; // [n][kInstI]
  // User levels, batch-wide: layer index and scaled depth, [n][ntau].
  int32_t* u_lyr;
  double* u_tau;
  // Constant tables.
  const double* tab;  // TabOff(N)
  const double* ylm;  // [N][N][NN]: Y_l^m at the positive nodes, [m][l][r]
  const double* ylq;  // [N][N][NN]: Y_l^m(mu_r) sqrt(w_r / mu_r)
  const double* ylw;  // [N][N][NN]: w_r Y_l^m(mu_r)
  const double* ylu;  // [N][N][numu]: Y_l^m at the user angles
  const double* gaz;  // [2][NMUG]: BRDF azimuth nodes and weights
  // Surface (BRDF) tables.
  const int32_t* rpv_set;  // [n] parameter-set index of each column
  const double* set_rpv;   // [nsets][7] BRDF parameter blocks
  int32_t nsets, brdf_kind;  // brdf_kind: kBrdf* of the batch
  double* srf_d;  // [nsets][nm][surf_d_size]
  double* srf_b;  // [nc][nm][surf_b_size] (chunk-local)
  // Chunk scratch.
  double* eig;    // [n_tasks][GE::size]
  double* gu;     // [n_tasks][numu][N]
  double* usrc;   // [n_tasks][kUSrc][numu]
  double* panel;  // [n_tasks][panel_stride(N)]
  double* cvec;   // [n_tasks][N] integration constants
  double* uum;    // [nc][nm][ntau][numu] intensity Fourier components
  int32_t* status;
  // Buras-Emde correction: the tabulated phase function and, per (chunk
  // column, user level), the residual phase function and its equal-area
  // integration grid (kBeNF points).
  const double* mu_phase;  // [nphase] increasing
  const double* phase;     // [n][nlyr][nphase]
  int32_t nphase, pad1;
  double* be_ph2;          // [nc][ntau][nphase]
  double* be_grid;         // [nc][ntau][kBeNF]
  int32_t* be_neg;         // [nc][ntau][kBeNF]
  double* be_norm;         // [nc][ntau]
};
constexpr int kBeNF = 100;
enum { kBrdfRpv = 0, kBrdfHapke = 1, kBrdfCoxMunk = 2, kBrdfRossLi = 3 };
enum : int32_t {
  kBrdfFlags = DISORT_G_BRDF_RPV | DISORT_G_BRDF_HAPKE | DISORT_G_BRDF_COXMUNK | DISORT_G_BRDF_ROSSLI
};
constexpr int kMaxPhaseAngles = 2048;

// Per-chunk launch options of the general solver.
struct GenLaunch {
  double* ucoef;        // [n_tasks][4][numu] layer coefficients of the user intensities
  int32_t* modes_used;  // [nc] Fourier modes summed per column
  bool brdf, do_user, nt, buras, brdf_corr;  // nt: intensity correction on; buras: Buras-Emde form
  hipEvent_t* ev;       // 9 events for stage timing, or null
};

struct FluxArgs;
bool general_setup(int nstr, const FluxArgs& fa, const GenArgs& a, hipStream_t s, bool brdf);
bool general_chunk(int nstr, const GenArgs& a, const GenLaunch& g, hipStream_t s);

}  // namespace dsa
