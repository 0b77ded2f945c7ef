// Device-side layouts shared by the flux kernels (flux_kernels.hip) and the
// host driver (capi.cpp).  All scratch is FP64 structure-of-arrays in HBM.
#pragma once
#include <cstdint>

namespace dsa {

// Per-N constant table (doubles): double-Gauss nodes/weights, the similarity
// vector sqrt(w*mu), Legendre values at the positive nodes and the scaled
// Legendre values Q_l(i) = P_l(mu_i) sqrt(w_i/mu_i) used to build the
// symmetric forms of alpha +/- beta directly.
struct TabOff {
  int n, nn;
  __host__ __device__ constexpr TabOff(int n_) : n(n_), nn(n_ / 2) {}
  __host__ __device__ constexpr int cmu() const { return 0; }             // [N]
  __host__ __device__ constexpr int cwt() const { return n; }             // [N]
  __host__ __device__ constexpr int sd() const { return 2 * n; }          // [NN]
  __host__ __device__ constexpr int sdi() const { return 2 * n + nn; }    // [NN]
  __host__ __device__ constexpr int P() const { return 2 * n + 2 * nn; }  // [N][NN], l-major
  __host__ __device__ constexpr int Q() const { return P() + n * nn; }    // [N][NN]
  __host__ __device__ constexpr int size() const { return Q() + n * nn; }
};

// FluxArgs::flags bit (beside the DISORT_F_* bits) of batches solved
// without absorption cut-off (general solver, DISORT_G_NO_CUTOFF).
constexpr int32_t kFluxNoCutoff = 1 << 20;

// Per-instance record (doubles), index inst*kInstD + field.
enum InstD {
  ID_PL_TOP = 0,      // emissivity_top * B(T_top)
  ID_PL_BOT = 1,      // B(T_bottom)
  ID_PL_BOT_DERIV = 2,
  ID_BTRANS_NCUT = 3, // beam transmission at level ncut (0 below the absorption cut)
  kInstD = 4
};
// Per-instance record (ints), index inst*kInstI + field.
// II_SCALING: bit 0 Delta-M applied, bit 1 Delta-M+ applied.
enum InstI { II_NCUT = 0, II_LYRCUT = 1, II_STATUS = 2, II_SCALING = 3, kInstI = 4 };

// Per-layer record (doubles), SoA: field*total_layers + global_layer.
enum LayD {
  LD_DTAU_S = 0,   // delta-M scaled optical thickness
  LD_SSA_S = 1,    // scaled single-scattering albedo
  LD_F = 2,        // truncated forward fraction (0 without scaling)
  LD_SIG2 = 3,     // Delta-M+ Gaussian width (0 = plain Delta-M, -1 = no scaling)
  LD_TAU_TOP = 4,  // scaled cumulative depth at the layer top
  LD_BT_TOP = 5,   // beam transmission used by the boundary conditions
  LD_BT_BOT = 6,
  LD_CH = 7,       // beam path factor of the layer: mu0, or the Chapman factor
                   // tau'_mid / CH(tau'_mid) with the spherical beam
  LD_CHTAU = 8,    // spherical beam: slant optical depth at the layer bottom
  kLayD = 9
};
// Per-level record (doubles), SoA: field*total_levels + global_level.
enum LevD {
  VD_TAU_S = 0,  // scaled cumulative optical depth
  VD_TAUC = 1,   // unscaled cumulative optical depth
  VD_PLANCK = 2, // band-integrated Planck function of the level temperature
  kLevD = 3
};

// Per-layer output of the eigen kernel (doubles): one contiguous record per
// layer task, task*stride() + field, so that the boundary solve reads a
// layer's eigen-system as whole cache lines.  gp/gm are column-major NN x NN:
// entry (a, j') at a + j'*NN.  The N x N eigenvector matrix of the reference
// (src/FluxSolver.cpp:574-586) is
//   G(NN+a, NN+j') =  gp(a,j')   G(NN-1-a, NN+j')     =  gm(a,j')
//   G(NN+a, NN-1-j') = -gm(a,j') G(NN-1-a, NN-1-j')   = -gp(a,j')
// with eigenvalues kk(NN+j') = k(j'), kk(NN-1-j') = -k(j').
struct EigOff {
  int n, nn;
  __host__ __device__ constexpr EigOff(int n_) : n(n_), nn(n_ / 2) {}
  __host__ __device__ constexpr int k() const { return 0; }                       // [NN]
  __host__ __device__ constexpr int gp() const { return nn; }                     // [NN*NN]
  __host__ __device__ constexpr int gm() const { return nn + nn * nn; }           // [NN*NN]
  __host__ __device__ constexpr int ztop() const { return nn + 2 * nn * nn; }     // [N] particular sol. at top (BC form)
  __host__ __device__ constexpr int zbot() const { return ztop() + n; }           // [N] at bottom (BC form)
  __host__ __device__ constexpr int zfl() const { return zbot() + n; }            // [N] at bottom (flux form)
  __host__ __device__ constexpr int pls() const { return zfl() + n; }             // Planck source at bottom
  __host__ __device__ constexpr int size() const { return pls() + 1; }
  // Record pitch, even so that records start on 16-byte boundaries.
  __host__ __device__ constexpr int stride() const { return (size() + 1) & ~1; }
  __host__ __device__ constexpr int64_t at(int64_t task, int field) const { return task * stride() + field; }
};

// Boundary-solve factors per panel (doubles), AoS per task:
// task*stride + [R (N*N col-major) | U (N*N col-major) | r (N)].
__host__ __device__ constexpr int panel_stride(int n) { return 2 * n * n + n; }

// Temperature Jacobians: right-hand sides eliminated per factorisation of
// the boundary system, and the panel record of that solve, U (2 N*N) then
// the K eliminated right-hand sides (K*N).
constexpr int kJacRhs = 8;
__host__ __device__ constexpr int jac_panel_stride(int n) { return 2 * n * n + kJacRhs * n; }

}  // namespace dsa
