/*
 * disort_amd.h -- C ABI of the MI355X DISORT library (libdisort_amd.so).
 *
 * The reference (DisORT++, NewStrangeWorlds/DisORT) has no C ABI: its hot path
 * is reached through C++ classes and a pybind11 module.  Each entry point
 * below names the reference interface it replaces; the C++ facade in
 * disort_amd/host (DisortFluxSolver<N>, DisortSolver, solveFluxSpectralBands,
 * solveSpectralBands) and the pybind11 module `disortpp` sit on top of it.
 *
 * Conventions: plain pointers and sizes, no exceptions cross the boundary,
 * every function returns DISORT_OK (0) or a DISORT_E_* code and leaves a
 * message in disort_last_error(ctx).  A context owns one HIP stream on one
 * device; calls on one context are not re-entrant (as a reference solver
 * object is not: src/FluxSolver.hpp:80-194 keeps per-solve scratch).
 */
#ifndef DISORT_AMD_H
#define DISORT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes -------------------------------------------------------- */
enum {
  DISORT_OK = 0,
  DISORT_E_INVALID = 1,     /* input fails the reference's validate() rules     */
  DISORT_E_UNSUPPORTED = 2, /* configuration outside what the GPU path solves   */
  DISORT_E_NUMERIC = 3,     /* an instance hit a numerical failure (see status) */
  DISORT_E_HIP = 4,         /* HIP runtime error                                 */
  DISORT_E_NOMEM = 5
};

/* Per-instance status codes written to disort_flux_batch.status[i]. */
enum {
  DISORT_ST_OK = 0,
  DISORT_ST_EIGEN = 1       /* neither symmetrisation of the layer eigenproblem
                               is positive definite (degenerate layer)          */
};

/* ---- batch flags (uniform over a batch) ---------------------------------- */
enum {
  DISORT_F_THERMAL = 1,       /* DisortFluxConfig::use_thermal_emission      */
  DISORT_F_SPHERICAL = 2,     /* use_spherical_beam (workgroup kernels)      */
  DISORT_F_DELTA_M_PLUS = 4,  /* use_delta_m_plus                            */
  DISORT_F_DIFFUSION_BC = 8   /* use_diffusion_lower_bc                      */
};

/* ---- per-instance scalar block: scal[i*DISORT_NSCAL + slot] ------------- */
enum {
  DISORT_S_BEAM_FLUX = 0, /* direct_beam_flux       */
  DISORT_S_BEAM_MU = 1,   /* direct_beam_mu         */
  DISORT_S_ISO_TOP = 2,   /* isotropic_flux_top     */
  DISORT_S_ISO_BOT = 3,   /* isotropic_flux_bottom  */
  DISORT_S_TEMP_TOP = 4,  /* temperature_top        */
  DISORT_S_TEMP_BOT = 5,  /* temperature_bottom     */
  DISORT_S_EMIS_TOP = 6,  /* emissivity_top         */
  DISORT_S_ALBEDO = 7,    /* surface_albedo         */
  DISORT_S_WN_LO = 8,     /* wavenumber_low         */
  DISORT_S_WN_HI = 9,     /* wavenumber_high        */
  DISORT_S_RADIUS = 10,   /* bottom_radius          */
  DISORT_S_BEAM_PHI = 11, /* direct_beam_phi [deg]; general solver only */
  DISORT_NSCAL = 12
};

/* ---- output fields: out[field*total_levels + level] ---------------------- */
/* Same order as the members of FluxResult (src/FluxResult.hpp:14-22). */
enum {
  DISORT_O_FLUX_DIRECT_BEAM = 0,
  DISORT_O_FLUX_DOWN = 1,
  DISORT_O_FLUX_UP = 2,
  DISORT_O_FLUX_TAU_DIVERGENCE = 3,
  DISORT_O_MEAN_INTENSITY = 4,
  DISORT_O_MEAN_INTENSITY_DOWN = 5,
  DISORT_O_MEAN_INTENSITY_UP = 6,
  DISORT_O_MEAN_INTENSITY_DIRECT_BEAM = 7,
  DISORT_NOUT = 8
};

enum { DISORT_MEM_HOST = 0, DISORT_MEM_DEVICE = 1 };

/*
 * A batch of independent flux-only DISORT instances (one instance = one
 * (wavenumber band x column) solve of DisortFluxSolver<N>::solve,
 * src/FluxSolver.cpp:23-171).  Layouts are structure-of-arrays with
 * per-instance offsets, so stream count and layer count may differ between
 * instances ("ragged" batches).  Instance i owns
 *   layers  dtau[lay_off[i] .. lay_off[i]+nlyr[i])       (same for ssa)
 *   levels  temper[lay_off[i]+i .. +nlyr[i]+1)            (same for alt, out)
 *   moments pmom[mom_off[i] + lc*pstride[i] + k], k = 0..pstride[i]-1
 * Arrays are indexed from the top of the atmosphere (index_from_bottom is
 * handled by the host facade).  pstride[i] >= nstr[i]+1 (>= nstr[i]+2 with
 * DISORT_F_DELTA_M_PLUS), mirroring phase_function_moments[lc][0..nmomNstr()].
 */
typedef struct disort_flux_batch {
  int64_t n;
  int32_t flags;
  int32_t reserved;
  const int32_t* nstr;
  const int32_t* nlyr;
  const int32_t* pstride;
  const int64_t* lay_off;
  const int64_t* mom_off;
  const double* dtau;
  const double* ssa;
  const double* pmom;
  const double* temper; /* NULL unless DISORT_F_THERMAL */
  const double* alt;    /* NULL unless DISORT_F_SPHERICAL */
  const double* scal;
  double* out;
  int64_t total_levels; /* sum over i of nlyr[i]+1 */
  int32_t* status;      /* [n], may be NULL */
} disort_flux_batch;

typedef struct disort_ctx disort_ctx;

/* Context: one device, one stream, grow-only device scratch. */
int disort_ctx_create(int device, disort_ctx** ctx);
int disort_ctx_device(const disort_ctx* ctx);
int disort_ctx_destroy(disort_ctx* ctx);
const char* disort_last_error(const disort_ctx* ctx);
void* disort_ctx_stream(disort_ctx* ctx); /* the hipStream_t the kernels run on */
int disort_sync(disort_ctx* ctx);

/*
 * Batched flux solve on host arrays.  Replaces the OpenMP band loop of
 * solveFluxSpectralBandsImpl<N> (src/python_bindings.cpp:170-190) and the
 * per-instance DisortFluxSolver<N>::solve (src/FluxSolver.cpp:23-171).
 * memspace must be DISORT_MEM_HOST: inputs are validated with the rules of
 * DisortFluxConfig::validate (src/DisortFluxConfig.cpp:43-143), copied in,
 * solved, and the outputs and per-instance status copied back before return.
 * Supported on the GPU: nstr in {4,6,...,16,32,64} (the stream counts the
 * reference instantiates, src/FluxSolver.cpp:1572-1580), plane-parallel and
 * pseudo-spherical beam (level altitudes in batch->alt, radius in
 * scal[DISORT_S_RADIUS]), thermal emission, Lambertian surface, Delta-M /
 * Delta-M+, layer cut-off, diffusion lower boundary.
 */
int disort_flux_solve(disort_ctx* ctx, const disort_flux_batch* batch, int memspace);

/*
 * Temperature Jacobians of a thermal batch on host arrays (the
 * compute_temperature_jacobian option of DisortFluxSolver<N>::solve,
 * src/FluxSolver.cpp:1294-1497, FluxResult::*_temperature_jac).  For
 * instance i with L = nlyr[i], jac holds 6 (L+1) (L+2) values at offset
 * 6 sum_{j<i} (L_j+1)(L_j+2), as [field][level lu][dof k]: fields flux_up,
 * flux_down, mean_intensity, mean_intensity_down, mean_intensity_up,
 * flux_tau_divergence; dof k = 0..L the level temperature temper[k], k = L+1
 * temperature_bottom.  The derivatives are the GPU solves of the same media
 * with a unit Planck source at one degree of freedom and no other source,
 * times dB/dT.  Validated as disort_flux_solve; needs DISORT_F_THERMAL.
 */
int disort_flux_jacobian(disort_ctx* ctx, const disort_flux_batch* batch, double* jac);

/*
 * Device-resident path.  A plan holds the bucket schedule (instances grouped
 * by stream count, layer and level task lists) computed once from host-side
 * shapes; disort_flux_execute then only launches kernels, asynchronously on
 * the context stream, reading and writing device arrays.  The batch's
 * lay_off/mom_off are ignored (the plan's prefix offsets are used).  Values
 * of device arrays are not validated.  The same split as constructing a
 * DisortFluxSolver<N> once and calling solve() per band
 * (src/python_bindings.cpp:177-188).
 */
typedef struct disort_flux_plan disort_flux_plan;
int disort_flux_plan_create(disort_ctx* ctx, int64_t n, const int32_t* nstr, const int32_t* nlyr,
                            const int32_t* pstride, int32_t flags, disort_flux_plan** plan);
int disort_flux_plan_destroy(disort_ctx* ctx, disort_flux_plan* plan);
int disort_flux_execute(disort_ctx* ctx, disort_flux_plan* plan, const disort_flux_batch* device_batch);
/* Per-instance DISORT_ST_* codes of the last execute with this plan (host array [n]). */
int disort_flux_status(disort_ctx* ctx, disort_flux_plan* plan, int32_t* status_host);

/*
 * Device-resident temperature Jacobians of n uniform columns (nstr, nlyr,
 * pstride shared; lay_off[i] = i*nlyr): disort_flux_jacobian
 * (src/FluxSolver.cpp:1294-1497, FluxResult::*_temperature_jac) on device
 * arrays.  disort_jac_execute reads the batch's dtau, ssa, pmom, temper and
 * scal (device pointers; values are not validated) and writes the device
 * array jac [n][6][nlyr+1][nlyr+2] in the layout of disort_flux_jacobian,
 * asynchronously on the context stream.
 *
 * Factored form (nstr <= 16, the default): per column one eigen-system
 * pass, then the nlyr+2 unit-source right-hand sides of the boundary system
 * eliminated eight at a time alongside its factorisation, the reference's
 * back-substitution route.  Expansion form (DISORT_JAC_EXPAND, and nstr >
 * 16): one forward solve per column x degree of freedom.  Columns are
 * processed in chunks of disort_jac_plan_chunk (at most max_expanded / (nlyr
 * + 2) columns when max_expanded > 0; by default 131072 expanded instances,
 * or a 4 GiB buffer budget in the factored form) with an inner flux plan
 * (disort_jac_plan_flux: its schedule setters apply).  flags:
 * DISORT_F_THERMAL (required), DISORT_F_DELTA_M_PLUS, DISORT_F_DIFFUSION_BC,
 * DISORT_JAC_EXPAND.  disort_jac_status copies the per-column DISORT_ST_*
 * codes of the last execute (expansion: the first nonzero code among the
 * column's solves) to a host array [n].
 */
#define DISORT_JAC_EXPAND 0x100
typedef struct disort_jac_plan disort_jac_plan;
int disort_jac_plan_create(disort_ctx* ctx, int64_t n, int32_t nstr, int32_t nlyr, int32_t pstride, int32_t flags,
                           int64_t max_expanded, disort_jac_plan** plan);
int disort_jac_plan_destroy(disort_ctx* ctx, disort_jac_plan* plan);
int64_t disort_jac_plan_chunk(const disort_jac_plan* plan);
disort_flux_plan* disort_jac_plan_flux(disort_jac_plan* plan);
int disort_jac_execute(disort_ctx* ctx, disort_jac_plan* plan, const disort_flux_batch* device_batch, double* jac);
int disort_jac_status(disort_ctx* ctx, disort_jac_plan* plan, int32_t* status_host);

/*
 * Kernel variant of one stream-count bucket of a plan (no counterpart in the
 * reference: it selects between gfx950 implementations of the same
 * solve).  DISORT_V_LANE: register kernels, one lane per layer and 16 lanes
 * per instance, N in {4,...,16}.  DISORT_V_LANE_PAIR: the same with two
 * lanes per layer in the eigen-system kernel (half the registers, more
 * wavefronts per SIMD).  DISORT_V_WORKGROUP: LDS kernels, one workgroup per
 * layer and per instance, every N; the only variant for N = 32, 64 and for
 * the spherical beam.  All give results within the parity tolerance of each
 * other; the default is the register form where it exists.  The autotuner
 * (disort_flux_autotune) picks by timing.
 */
enum { DISORT_V_DEFAULT = 0, DISORT_V_LANE = 1, DISORT_V_WORKGROUP = 2, DISORT_V_LANE_PAIR = 3 };
int disort_flux_plan_set_variant(disort_ctx* ctx, disort_flux_plan* plan, int nstr, int variant);
int disort_flux_plan_get_variant(disort_ctx* ctx, disort_flux_plan* plan, int nstr, int* variant);

/*
 * Bucket schedule of a plan (no counterpart in the reference, whose band loop
 * is one OpenMP loop per stream count, src/python_bindings.cpp:170-190).
 * set_streams: 1 runs the buckets one after another on the context stream;
 * 2..4 forks them onto that many side streams (bucket-private scratch), each
 * bucket on the stream with the least work so far, largest first, and joins
 * them back, so that buckets overlap.  set_concurrent(on) is set_streams(3)
 * or set_streams(1).  set_chunks splits the bucket of one stream count into
 * 1..8 chunks of its instance list, each launched as its own bucket, so that
 * a large bucket spreads over the side streams and the tail of one chunk's
 * boundary sweep overlaps the next chunk's kernels.  Results do not depend
 * on the schedule (every instance runs the same kernels).
 */
int disort_flux_plan_set_streams(disort_ctx* ctx, disort_flux_plan* plan, int streams);
int disort_flux_plan_set_concurrent(disort_ctx* ctx, disort_flux_plan* plan, int on);
int disort_flux_plan_set_chunks(disort_ctx* ctx, disort_flux_plan* plan, int nstr, int chunks);

/*
 * disort_flux_execute captured once into a hipGraph (all buckets, all four
 * kernels each, with the fork/join of a concurrent schedule) and replayed by
 * one graph launch per call.  The graph is re-captured when the device
 * pointers of the batch, the schedule or a variant change.  No memset or
 * allocation is recorded in the graph.
 */
int disort_flux_execute_graph(disort_ctx* ctx, disort_flux_plan* plan, const disort_flux_batch* device_batch);

/*
 * Autotuner: times every kernel variant of every stream count, then the
 * bucket schedules (serial; 2 or 3 side streams with the large buckets in
 * 1, 2 or 3 chunks) on the given device batch (median of `reps` after a
 * warm-up; the batch's outputs are overwritten), and keeps the fastest in
 * the plan.  disort_flux_plan_report returns the measurements as JSON.
 */
int disort_flux_autotune(disort_ctx* ctx, disort_flux_plan* plan, const disort_flux_batch* device_batch, int reps);
const char* disort_flux_plan_report(disort_flux_plan* plan);

/*
 * Device-resident flux solve that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same solve and kernels as disort_flux_execute for a uniform batch --
 * every instance has nstr streams, nlyr layers and pstride moments per layer
 * row; the batch's nstr/nlyr/pstride/lay_off/mom_off arrays are ignored and
 * instance i owns layers [i*nlyr, (i+1)*nlyr) -- with every intermediate in
 * one static device array of the library (DISORT_STATIC_SCRATCH_BYTES per
 * device, held until the process ends).  Device pointers only; inputs are not
 * validated; outputs and, if batch->status is not NULL, the per-instance
 * status (device int32 array [n]) are written asynchronously on the context
 * stream.  A batch whose intermediates exceed the static array is rejected
 * with DISORT_E_INVALID (disort_flux_static_bytes gives the requirement; at
 * nstr=16, nlyr=60 about 350 KB per instance).  All calls on one device share
 * that array: calls that overlap in time on one device (other streams or
 * contexts) race.  The first call for a stream count on a device uploads its
 * constant table into the array (one synchronous host-to-device copy).
 */
#define DISORT_STATIC_SCRATCH_BYTES ((size_t)256 << 20)
int disort_flux_execute_static(disort_ctx* ctx, const disort_flux_batch* device_batch, int32_t nstr, int32_t nlyr,
                               int32_t pstride);
/* Bytes of static scratch a uniform batch needs, and the bytes available. */
size_t disort_flux_static_bytes(int64_t n, int32_t nstr, int32_t nlyr);
size_t disort_static_scratch_bytes(void);

/* Device memory helpers for callers without their own allocator (ctypes). */
int disort_device_alloc(disort_ctx* ctx, size_t bytes, void** ptr);
int disort_device_free(disort_ctx* ctx, void* ptr);
int disort_memcpy_h2d(disort_ctx* ctx, void* dst, const void* src, size_t bytes);
int disort_memcpy_d2h(disort_ctx* ctx, void* dst, const void* src, size_t bytes);

/*
 * Per-stage device timings of the most recent disort_flux_solve on this
 * context, in milliseconds, measured with hipEvents on the context stream
 * when timing is enabled.  Stages: 0 column set-up, 1 layer eigen-system,
 * 2 boundary solve, 3 flux assembly.  Returns the number of stages written.
 */
int disort_set_timing(disort_ctx* ctx, int enable);
int disort_stage_times(disort_ctx* ctx, double* ms, int max_stages);

/* Library version string. */
const char* disort_version(void);

/*
 * Double-Gauss computational angles and weights for nstr streams, in the
 * layout of gaussianQuadrature (src/FluxSolver.cpp:257-290): cmu[0..nstr/2)
 * the Gauss-Legendre nodes on (0,1), cmu[nstr/2+i] = -cmu[i], weights
 * mirrored.  Host-only; needs no context or GPU.
 */
int disort_quadrature(int nstr, double* cmu, double* cwt);
/* Band-integrated Planck function B(lo, hi, t[k]) into b[k] for n host
 * temperatures: src/Planck.cpp:35-169 (planckFunction2), the same arithmetic
 * the kernels use.  Returns DISORT_E_INVALID for lo < 0, hi < lo or t < 0. */
int disort_planck_band(double lo, double hi, int64_t n, const double* t, double* b);
/* Its temperature derivative dB/dT (planckFunctionDeriv2, src/Planck.cpp:171-214)
 * into d[k], the factor of the temperature Jacobians; same argument rules. */
int disort_planck_band_deriv(double lo, double hi, int64_t n, const double* t, double* d);

/* =========================================================================
 * General solver: DisortSolver::solve (src/DisortSolver.cpp:21-399) for a
 * batch of columns -- azimuthal Fourier modes, intensities at user polar
 * angles, azimuths and optical depths, Lambertian or RPV surface,
 * plane-parallel or pseudo-spherical beam, thermal emission, the
 * Nakajima-Tanaka intensity correction and the BRDF beam-reflection
 * correction.  Shapes are uniform over a batch (one DisortConfig with
 * per-column optical properties, as solveSpectralBands varies them,
 * src/python_bindings.cpp:96-128).
 * ========================================================================= */
enum {
  DISORT_G_USER_TAU = 16,       /* flags.use_user_tau                          */
  DISORT_G_USER_MU = 32,        /* flags.use_user_mu                           */
  DISORT_G_ONLY_FLUXES = 64,    /* flags.comp_only_fluxes                      */
  DISORT_G_BRDF_RPV = 128,      /* use_lambertian_surface = false, BrdfType::RPV */
  DISORT_G_CORR_BURAS = 256,    /* flags.intensity_corr_buras                  */
  DISORT_G_CORR_NAKAJIMA = 512, /* flags.intensity_corr_nakajima               */
  DISORT_G_BRDF_HAPKE = 1024,   /* use_lambertian_surface = false, BrdfType::Hapke   */
  DISORT_G_BRDF_COXMUNK = 2048, /* use_lambertian_surface = false, BrdfType::CoxMunk */
  DISORT_G_BRDF_ROSSLI = 4096,  /* use_lambertian_surface = false, BrdfType::Ambrals */
  DISORT_G_NO_CUTOFF = 8192     /* no absorption cut-off (ncut = nlyr, lyrcut = false), as the
                                   SPECIAL_BC base solves of src/DisortSolver.cpp:4037-4039 */
};
/* BRDF parameter block rpv[i*DISORT_NRPV + slot], one of the BRDF flags set
 * (src/DisortConfig.hpp:52-96): RPV (RpvBrdfSpec) in the slots below; Hapke
 * (HapkeBrdfSpec) {b0, hh, w}; Cox-Munk (CoxMunkBrdfSpec) {u10,
 * refractive_index, do_shadow as 1.0 / 0.0}; Ross-Li (AmbralsBrdfSpec)
 * {iso, vol, geo}; unused slots 0. */
enum {
  DISORT_RPV_RHO0 = 0, DISORT_RPV_K = 1, DISORT_RPV_THETA = 2, DISORT_RPV_SIGMA = 3,
  DISORT_RPV_T1 = 4, DISORT_RPV_T2 = 5, DISORT_RPV_SCALE = 6, DISORT_NRPV = 7
};

typedef struct disort_general_batch {
  int64_t n;            /* columns                                               */
  int32_t flags;        /* DISORT_F_* | DISORT_G_*                               */
  int32_t nstr;         /* num_streams                                           */
  int32_t nlyr;         /* num_layers                                            */
  int32_t pstride;      /* moments per layer row: max(nmom, nstr) + 1            */
  int32_t ntau;         /* output levels: num_user_tau, or nlyr+1 without USER_TAU */
  int32_t numu;         /* num_user_mu, or nstr without USER_MU / with ONLY_FLUXES */
  int32_t nphi;         /* num_phi (0 with ONLY_FLUXES)                          */
  int32_t reserved;
  double accuracy;      /* accuracy_fourier_series (0 = all nstr modes)          */
  const double* dtau;   /* [n][nlyr]                                             */
  const double* ssa;    /* [n][nlyr]                                             */
  const double* pmom;   /* [n][nlyr][pstride]                                    */
  const double* temper; /* [n][nlyr+1] with DISORT_F_THERMAL, else NULL          */
  const double* alt;    /* [n][nlyr+1] with DISORT_F_SPHERICAL, else NULL        */
  const double* scal;   /* [n][DISORT_NSCAL]                                     */
  const double* rpv;    /* [n][DISORT_NRPV] BRDF parameters with a BRDF flag, else NULL */
  const double* tau_user; /* [n][ntau] with DISORT_G_USER_TAU, else NULL         */
  const double* mu_user;  /* [numu], increasing, with DISORT_G_USER_MU           */
  const double* phi_user; /* [nphi] degrees                                      */
  double* flux;         /* out [DISORT_NOUT][n][ntau], FluxResult field order    */
  double* inten;        /* out [n][ntau][numu][nphi] (DisortResult::intensity)   */
  double* inten_avg;    /* out [n][ntau][numu] (intensity_azimuth_avg)           */
  int32_t* status;      /* out [n], may be NULL                                  */
  /* Tabulated phase function of the Buras-Emde correction
   * (DisortConfig::num_phase_func_angles, mu_phase_function, phase_function):
   * used when DISORT_G_CORR_BURAS is set, DISORT_G_CORR_NAKAJIMA is not and
   * nphase > 0 (src/DisortSolver.cpp:364-377); otherwise ignored. */
  int32_t nphase;       /* num_phase_func_angles (0: none)                       */
  int32_t reserved2;
  const double* mu_phase; /* [nphase] scattering-angle cosines, increasing, shared */
  const double* phase;    /* [n][nlyr][nphase] phase function values              */
} disort_general_batch;

/*
 * Solve a general batch.  Replaces DisortSolver::solve per column
 * (src/DisortSolver.cpp:21-399) and the band loop of solveSpectralBands
 * (src/python_bindings.cpp:96-128).  memspace DISORT_MEM_HOST: arrays are
 * host arrays, validated with the rules and messages of
 * DisortConfig::validate (src/DisortConfig.cpp:72-249), copied in, solved,
 * copied out.  DISORT_MEM_DEVICE: array pointers are device pointers (not
 * validated; mu_user, phi_user and rpv are read back once), outputs stay on
 * the device, the call returns after the launches are queued on the context
 * stream.  Supported on the GPU: nstr in {4,...,16,32,64}; user optical
 * depths, polar angles and azimuths; Lambertian or RPV surface;
 * plane-parallel or pseudo-spherical beam; thermal emission; Delta-M /
 * Delta-M+; layer cut-off; diffusion lower boundary; accuracy_fourier_series;
 * with DISORT_G_CORR_BURAS the Nakajima-Tanaka correction
 * (src/DisortSolver.cpp:3403-3524) or, with a tabulated phase function and
 * without DISORT_G_CORR_NAKAJIMA, the Buras-Emde correction (:3532-3745,
 * :3812-4019), and the BRDF beam-reflection correction (:3751-3808).  Not
 * here: SPECIAL_BC (the C++ facade builds it from general solves with
 * DISORT_G_NO_CUTOFF) and temperature Jacobians (disort_general_jacobian).
 * All modes of every column are solved in parallel; the
 * accuracy_fourier_series cut is applied when they are summed, which gives
 * the reference's result.  Stage times (disort_stage_times, 8 stages):
 * set-up, BRDF beam terms, layers, boundary, fluxes, user coefficients, user
 * intensities, mode sum and corrections.
 */
int disort_general_solve(disort_ctx* ctx, const disort_general_batch* batch, int memspace);

/*
 * Temperature Jacobians of a thermal general batch on host arrays (the
 * compute_temperature_jacobian option of DisortSolver::solve,
 * src/DisortSolver.cpp:2130-2377, DisortResult::*_temperature_jac).  jac
 * holds, per column, 6 ntau (nlyr+2) values as [field][output level
 * lu][dof k]: fields flux_up, flux_down, mean_intensity, mean_intensity_down,
 * mean_intensity_up, flux_tau_divergence; rows the output levels (user
 * optical depths with DISORT_G_USER_TAU); dof k = 0..nlyr the level
 * temperature, k = nlyr+1 temperature_bottom.  Lambertian or BRDF surface,
 * diffusion lower boundary.  Validated as disort_general_solve; needs
 * DISORT_F_THERMAL.  User angles, azimuths and intensity corrections do not
 * enter (the Jacobians are of fluxes and mean intensities).
 */
int disort_general_jacobian(disort_ctx* ctx, const disort_general_batch* batch, double* jac);


/* =========================================================================
 * Learned flux surrogate (no counterpart in the reference).  A multilayer
 * perceptron with BF16 weights and activations and FP32 accumulation on the
 * gfx950 matrix cores, trained by the library's own forward, backward and
 * Adam kernels on columns the batched solver generates; plain DISORT stays
 * the fallback and the measure of its error (disort_amd/surrogate.py holds
 * the training loop and the gate).  All array pointers are device pointers.
 * ========================================================================= */
typedef struct disort_mlp disort_mlp;

/* widths[0..n_layers]: layer widths, multiples of 64 (pad inputs/outputs
 * with zero columns); max_batch a multiple of 64.  Hidden layers use GELU
 * (tanh form), the last layer is linear.  Parameters are packed per layer
 * as W_l [widths[l] x widths[l+1]] row-major, then b_l [widths[l+1]], FP32. */
int disort_mlp_create(disort_ctx* ctx, int n_layers, const int* widths, int max_batch, disort_mlp** mlp);
int disort_mlp_destroy(disort_ctx* ctx, disort_mlp* mlp);
const char* disort_mlp_last_error(const disort_mlp* mlp);
int64_t disort_mlp_param_count(const disort_mlp* mlp);
int disort_mlp_set_params(disort_ctx* ctx, disort_mlp* mlp, const float* host_params); /* also resets Adam state */
int disort_mlp_get_params(disort_ctx* ctx, disort_mlp* mlp, float* host_params);
int disort_mlp_get_grads(disort_ctx* ctx, disort_mlp* mlp, float* host_grads);
/* y [B x widths[L]] FP32 = network(x [B x widths[0]] BF16 bits); B a multiple of 64.
 * disort_mlp_forward keeps what disort_mlp_backward needs; disort_mlp_infer
 * (serving) keeps nothing and writes the same y. */
int disort_mlp_forward(disort_ctx* ctx, disort_mlp* mlp, const uint16_t* x, int batch, float* y);
int disort_mlp_infer(disort_ctx* ctx, disort_mlp* mlp, const uint16_t* x, int batch, float* y);
/* Parameter gradients for output gradient dy [B x widths[L]] (BF16 bits) of the last forward. */
int disort_mlp_backward(disort_ctx* ctx, disort_mlp* mlp, const uint16_t* dy);
int disort_mlp_adam(disort_ctx* ctx, disort_mlp* mlp, float lr, float beta1, float beta2, float eps, int step);
/* forward, loss = sum mask (y - t)^2 / (B * n_valid) added to *loss_dev, backward, Adam. */
int disort_mlp_train_step(disort_ctx* ctx, disort_mlp* mlp, const uint16_t* x, const float* targets,
                          const float* mask, int batch, int n_valid, float lr, float beta1, float beta2, float eps,
                          int step, float* loss_dev);

/* Surrogate column kernels on a uniform device batch (n columns of nlyr
 * layers, lay_off[i] = i*nlyr).  Features: 4*nlyr+2 per column (log10 dtau,
 * albedo, g = pmom[1], layer-mean temperature per layer; mu0; surface
 * albedo), then with w0 >= 7*nlyr+5 the 3*(nlyr+1) level features, with
 * w0 >= 8*nlyr+7 the nlyr+2 thermal-source features, and with ts given and
 * w0 >= 10*nlyr+9 the two-stream baseline's diffuse net and sum flux in
 * beam units at every level, standardised as (f - shift) * scale into BF16
 * rows of width w0 (rows n..n_pad-1 zero); ood[i] = 1 when a raw feature
 * leaves [lo, hi].  ts: the baseline of disort_surrogate_twostream, or NULL. */
int disort_surrogate_features(disort_ctx* ctx, int64_t n, int64_t n_pad, int nlyr, int pstride,
                              const disort_flux_batch* batch, const float* shift, const float* scale,
                              const float* lo, const float* hi, int w0, uint16_t* x, int32_t* ood,
                              const double* ts);
/* Standardised targets in units of the incident beam mu0*F0 at nlev levels:
 * diffuse net (down - up), diffuse sum (down + up), direct beam; from the
 * solved batch's outputs (batch->out).  With ts the first two blocks are the
 * residuals over the two-stream baseline. */
int disort_surrogate_targets(disort_ctx* ctx, int64_t n, int64_t n_pad, int nlev, const disort_flux_batch* batch,
                             const float* shift, const float* scale, int wL, float* targets, const double* ts);
/* Network outputs of 1..8 ensemble members (ys: host array of n_members
 * device pointers) -> the members' mean fluxes [n][3][nlev] (up, down,
 * direct; the direct beam in closed form; with ts the outputs are residuals
 * added to the two-stream baseline) and their net-flux spread per column:
 * max over levels of twice the standard deviation of Fnet over the members
 * (|Fnet_1 - Fnet_2| for two; 0 for one) divided by
 * max(c_local |Fnet|, c_floor mu0 F0). */
int disort_surrogate_decode(disort_ctx* ctx, int64_t n, int nlyr, int wL, const float* const* ys, int n_members,
                            const float* shift, const float* scale, const disort_flux_batch* batch, double c_local,
                            double c_floor, double* flux, float* spread, const double* ts);
/* Route columns idx[0..m) of src into the compact batch dst, and the plain
 * solver's outputs for them back into the flux array. */
int disort_surrogate_gather(disort_ctx* ctx, int64_t m, const int32_t* idx, int nlyr, int pstride,
                            const disort_flux_batch* src, disort_flux_batch* dst);
int disort_surrogate_scatter(disort_ctx* ctx, int64_t m, const int32_t* idx, int nlev, const double* out,
                             int64_t total_levels, double* flux);
/* Gate routing on the device: column i goes to plain DISORT when it is out
 * of distribution (ood[i] != 0, ood may be null), its ensemble spread is not
 * <= tau (NaN included), or it is sampled (a counter-based draw from seed
 * and i below sample_frac); with route_all every column is routed.
 * flags[i]: bit 0 routed, bit 1 sampled, bit 2
 * passes the gate (in distribution, spread <= tau).  The
 * routed indices are written to idx (in no particular order) and their
 * number to *count (device). */
int disort_surrogate_route(disort_ctx* ctx, int64_t n, const float* spread, const int32_t* ood, float tau,
                           double sample_frac, uint64_t seed, int route_all, int32_t* flags, int32_t* idx,
                           int32_t* count);
/* idx[m .. mpad) = idx[0] (a routed list padded to a reusable sub-batch size). */
int disort_surrogate_pad_index(disort_ctx* ctx, int64_t m, int64_t mpad, int32_t* idx);
/* Online error monitor: among the m routed columns idx[k] (plain outputs in
 * out, sub-batch column k), the sampled ones that pass the gate (flags
 * bits 1 and 2) are compared
 * with the surrogate fluxes pred [n][3][nlev] of batch b:
 * counters[0] = sampled columns, counters[1] = those whose max over levels of
 * |dF_net| / max(c_local |F_net|, c_floor mu0 F0) is not <= tol. */
int disort_surrogate_monitor(disort_ctx* ctx, int64_t m, const int32_t* idx, const int32_t* flags, int nlev,
                             const double* out, int64_t total_levels, const double* pred, const disort_flux_batch* b,
                             double c_local, double c_floor, double tol, int32_t* counters);
/* Two-stream baseline of the flux surrogate (no counterpart in the
 * reference): delta-Eddington scaled, quadrature-coefficient two-stream
 * fluxes of n uniform columns, one thread per column, with the beam, the
 * thermal source linear in optical depth (DISORT_F_THERMAL) and a
 * Lambertian surface; no diffuse flux enters at the top.  Row i of ts
 * (4*nlyr+4 doubles): diffuse up at the levels, diffuse down without the
 * unscaled direct beam (DisortFluxSolver's output convention), the band
 * Planck values of the levels and of the surface temperature (with
 * DISORT_F_THERMAL), then the layers' g = pmom[1].  work: 12 * n * nlyr doubles of device scratch;
 * nlyr + 1 <= 64. */
int disort_surrogate_twostream(disort_ctx* ctx, int64_t n, int nlyr, int pstride, const disort_flux_batch* batch,
                               double* work, double* ts);
/* Jacobian surrogate (no counterpart in the reference): the temperature
 * Jacobians of the top-of-atmosphere flux_up and the surface flux_down
 * (rows (0, level 0) and (1, level nlyr) of disort_flux_jacobian's layout)
 * for uniform columns, D = nlyr + 2 degrees of freedom each.  The network
 * learns the responses R_k = dF/dB_k / pi, which depend on the optical
 * properties only; dF/dT_k = pi R_k dB/dT(T_k) with dB/dT evaluated
 * exactly.  Features: 3*nlyr + 1 + 6*(nlyr+1) per column (log10 dtau,
 * albedo, g per layer; surface albedo; per level log10 of the optical and
 * absorption depths, their transmissions exp(-x/0.6) from the top and to
 * the surface), standardised into BF16 rows of width w0 as
 * disort_surrogate_features.  Targets: [k] R_k of the top flux_up, [D+k]
 * R_k of the surface flux_down, from a device Jacobian jac
 * [n][6][nlyr+1][nlyr+2] (disort_jac_execute).  Decode writes jrows
 * [n][2][D] (dF/dT of the two fluxes) and the members' spread max_k
 * |R_1 - R_2|.  The monitor compares the sampled columns that pass the gate
 * (flags bits 1 and 2) among the m routed idx[k] with the sub-batch
 * Jacobian jac_sub [m][6][nlyr+1][nlyr+2]: counters[0] = sampled columns,
 * counters[1] = those whose max over both rows and all k of
 * |dF/dT_k - ref| / (pi dB/dT(T_k)) is not <= tol.  Scatter copies the two
 * rows of jac_sub column k into jrows row idx[k]. */
int disort_jac_surrogate_features(disort_ctx* ctx, int64_t n, int64_t n_pad, int nlyr, int pstride,
                                  const disort_flux_batch* batch, const float* shift, const float* scale,
                                  const float* lo, const float* hi, int w0, uint16_t* x, int32_t* ood,
                                  const double* resp);
int disort_jac_surrogate_targets(disort_ctx* ctx, int64_t n, int64_t n_pad, int nlyr, const double* jac,
                                 const disort_flux_batch* batch, const float* shift, const float* scale, int wL,
                                 float* targets, const double* resp);
int disort_jac_surrogate_decode(disort_ctx* ctx, int64_t n, int nlyr, int wL, const float* y1, const float* y2,
                                const float* shift, const float* scale, const disort_flux_batch* batch,
                                double* jrows, float* spread, const double* resp);
/* Two-stream responses for the Jacobian surrogate's residual form: with
 * the two-stream model of disort_surrogate_twostream, resp [n][2][D] gets
 * the top-level U and surface-level D of each column for a unit Planck
 * value at degree of freedom k (levels, then the surface) and no beam, one
 * thread per column (the eliminations once, then one right-hand side per
 * k).  Given resp (else NULL), the features append the
 * 2D responses / pi (w0 >= 3*nlyr + 1 + 6*(nlyr+1) + 2D), the targets are
 * R / pi minus them and decoding adds them back.  Columns are processed
 * work_cols at a time; work: 18 * nlyr * work_cols doubles; ts: n rows as
 * for disort_surrogate_twostream (filled here). */
int disort_jac_twostream(disort_ctx* ctx, int64_t n, int nlyr, int pstride, const disort_flux_batch* batch,
                         double* work, int64_t work_cols, double* ts, double* resp);
int disort_jac_surrogate_monitor(disort_ctx* ctx, int64_t m, const int32_t* idx, const int32_t* flags, int nlyr,
                                 const double* jac_sub, const double* jrows, const disort_flux_batch* batch,
                                 double tol, int32_t* counters);
int disort_jac_surrogate_scatter(disort_ctx* ctx, int64_t m, const int32_t* idx, int nlyr, const double* jac_sub,
                                 double* jrows);
/* dst row k <- src row idx[k] of a row-major matrix of 2-, 4- or 8-byte elements. */
int disort_rows_gather(disort_ctx* ctx, int64_t m, const int32_t* idx, int64_t row_elems, int elem_bytes,
                       const void* src, void* dst);

#ifdef __cplusplus
}
#endif

#endif /* DISORT_AMD_H */
