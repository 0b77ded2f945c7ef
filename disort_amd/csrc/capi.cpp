// Host implementation of the C ABI declared in include/disort_amd.h.
//
// Replaces the OpenMP band loop of solveFluxSpectralBandsImpl<N>
// (src/python_bindings.cpp:170-190): instances are grouped into buckets of
// equal stream count, each bucket is solved by four kernel launches on the
// context stream (flux_kernels.hip), scratch lives in grow-only device
// buffers owned by the context.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/disort_amd.h"
#include "flux_args.hpp"
#include "flux_layout.hpp"
#include "ctx.hpp"
#include "planck.hpp"

namespace dsa {
// Register kernels (flux_kernels.hip), N in {4,...,16}.
bool launch_flux_bucket(int nstr, const FluxArgs& a, hipStream_t s, hipEvent_t* ev, bool pair);
// Workgroup kernels (flux_wide.hip), N in {8, 16, 32, 64}.
bool launch_flux_bucket_wg(int nstr, const FluxArgs& a, hipStream_t s, hipEvent_t* ev);
// Jacobian expansion and scaling (jacobian_kernels.hip).
bool launch_jac_expand(int64_t n, int64_t c0, int64_t ne, int L, int P, const disort_flux_batch* b, double* xdtau,
                       double* xssa, double* xpmom, double* xtemper, double* xscal, double* ovr, double* dbdt,
                       hipStream_t st);
bool launch_jac_scale(int64_t cols, int64_t c0, int L, const double* out, int64_t TV, const double* dbdt,
                      double* jac, const int32_t* inst_i, int32_t* status, hipStream_t st);
}

using dsa::FluxArgs;

namespace dsa_host {

int fail(disort_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

// Grow-only named device buffer.
int get_buf(disort_ctx* ctx, const char* name, size_t bytes, void** out) {
  DevBuf& b = ctx->bufs[name];
  if (b.bytes < bytes) {
    if (b.p) {
      HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
      HIPCHK(ctx, hipFree(b.p));
      b.p = nullptr;
      b.bytes = 0;
    }
    const size_t want = std::max<size_t>(bytes, 256);
    HIPCHK(ctx, hipMalloc(&b.p, want));
    b.bytes = want;
  }
  *out = b.p;
  return DISORT_OK;
}

// Double-Gauss quadrature on (0,1): Newton iteration on P_n (tolerance
// 3e-14, at most 10 steps) as in src/FluxSolver.cpp:257-290, so that nodes
// and weights agree with the reference to the last bits.
bool gauss_nodes(int n, std::vector<double>& cmu, std::vector<double>& cwt) {
  cmu.assign(2 * n, 0.0);
  cwt.assign(2 * n, 0.0);
  for (int i = 1; i <= n; ++i) {
    double z = std::cos(M_PI * (i - 0.25) / (n + 0.5)), z1, dp;
    int it = 0;
    do {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      dp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / dp;
      if (++it > 10) return false;
    } while (std::abs(z - z1) > 3.0e-14);
    cmu[i - 1] = 0.5 * (1.0 + z);
    cwt[i - 1] = 1.0 / ((1.0 - z * z) * dp * dp);
  }
  for (int i = 0; i < n; ++i) {
    cmu[n + i] = -cmu[i];
    cwt[n + i] = cwt[i];
  }
  return true;
}

int table_for(disort_ctx* ctx, int nstr, double** out) {
  auto it = ctx->tables.find(nstr);
  if (it != ctx->tables.end()) {
    *out = it->second;
    return DISORT_OK;
  }
  std::vector<double> tab;
  if (!build_table(nstr, tab)) return fail(ctx, DISORT_E_NUMERIC, "gaussianQuadrature: Failed to converge");
  double* d = nullptr;
  HIPCHK(ctx, hipMalloc(&d, tab.size() * sizeof(double)));
  HIPCHK(ctx, hipMemcpy(d, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  ctx->tables[nstr] = d;
  *out = d;
  return DISORT_OK;
}

bool build_table(int nstr, std::vector<double>& tab) {
  const int nn = nstr / 2;
  std::vector<double> cmu, cwt;
  if (!gauss_nodes(nn, cmu, cwt)) return false;
  dsa::TabOff to(nstr);
  tab.assign(to.size(), 0.0);
  for (int k = 0; k < nstr; ++k) {
    tab[to.cmu() + k] = cmu[k];
    tab[to.cwt() + k] = cwt[k];
  }
  for (int r = 0; r < nn; ++r) {
    const double sd = std::sqrt(cwt[r] * cmu[r]);
    tab[to.sd() + r] = sd;
    tab[to.sdi() + r] = 1.0 / sd;
    // Legendre P_l(mu_r) by upward recurrence (src/Legendre.cpp:73-90).
    std::vector<double> p(nstr, 0.0);
    p[0] = 1.0;
    if (nstr > 1) p[1] = cmu[r];
    for (int l = 2; l < nstr; ++l) p[l] = ((2.0 * l - 1.0) * cmu[r] * p[l - 1] - (l - 1.0) * p[l - 2]) / l;
    const double sc = std::sqrt(cwt[r] / cmu[r]);
    for (int l = 0; l < nstr; ++l) {
      tab[to.P() + l * nn + r] = p[l];
      tab[to.Q() + l * nn + r] = p[l] * sc;
    }
  }
  return true;
}

bool gpu_supported_nstr(int n) { return (n >= 4 && n <= 16 && n % 2 == 0) || n == 32 || n == 64; }

}  // namespace dsa_host

using namespace dsa_host;

namespace {

// Side streams of the concurrent schedule: a process gets four hardware
// queues, and more concurrent streams than three measured slower.
constexpr int kMaxSideStreams = 4;
constexpr int kDefaultSideStreams = 3;
constexpr int kMaxChunks = 8;

bool lane_kernels(int n) { return n >= 4 && n <= 16 && n % 2 == 0; }
bool wg_kernels(int n) { return (n >= 4 && n <= 16 && n % 2 == 0) || n == 32 || n == 64; }
// Register kernels where they exist; the two-lanes-per-layer eigen kernel
// where it measured faster on the 10,000-band workload (6 to 12 streams:
// no spill at two or three wavefronts per SIMD).
int default_variant(int n) {
  if (!lane_kernels(n)) return DISORT_V_WORKGROUP;
  return (n >= 6 && n <= 12) ? DISORT_V_LANE_PAIR : DISORT_V_LANE;
}
bool variant_ok(int n, int v) {
  return ((v == DISORT_V_LANE || v == DISORT_V_LANE_PAIR) && lane_kernels(n)) || (v == DISORT_V_WORKGROUP && wg_kernels(n));
}
const char* variant_name(int v) {
  return v == DISORT_V_LANE ? "lane" : v == DISORT_V_LANE_PAIR ? "lane_pair" : "workgroup";
}

// Host-side validation of one instance, with the reference's messages
// (src/DisortFluxConfig.cpp:43-143).
std::string validate_instance(const disort_flux_batch* b, int64_t i) {
  const int L = b->nlyr[i], N = b->nstr[i], P = b->pstride[i];
  auto at = [&](const std::string& m) { return "instance " + std::to_string(i) + ": " + m; };
  if (L <= 0) return at("num_layers must be > 0");
  if (N < 4 || N % 2 != 0) return at("num_streams must be even and >= 4");
  if (P < N + 1) return at("phase_function_moments[0] size < nmomNstr()+1");
  if ((b->flags & DISORT_F_DELTA_M_PLUS) && P < N + 2)
    return at("use_delta_m_plus requires num_phase_func_moments >= num_streams + 1");
  const int64_t lo = b->lay_off[i], vo = lo + i;
  for (int lc = 0; lc < L; ++lc) {
    if (b->dtau[lo + lc] < 0.0) return at("delta_tau[" + std::to_string(lc) + "] must be >= 0");
    const double w = b->ssa[lo + lc];
    if (w < 0.0 || w > 1.0) return at("single_scat_albedo[" + std::to_string(lc) + "] must be in [0, 1]");
  }
  const double* s = b->scal + i * DISORT_NSCAL;
  if ((b->flags & DISORT_F_DIFFUSION_BC) && !(b->flags & DISORT_F_THERMAL))
    return at("use_diffusion_lower_bc requires use_thermal_emission = true");
  if (s[DISORT_S_BEAM_FLUX] < 0.0) return at("direct_beam_flux must be >= 0");
  if (s[DISORT_S_BEAM_FLUX] > 0.0 && (s[DISORT_S_BEAM_MU] <= 0.0 || s[DISORT_S_BEAM_MU] > 1.0))
    return at("direct_beam_mu must be in (0, 1] when direct_beam_flux > 0");
  if (s[DISORT_S_ALBEDO] < 0.0 || s[DISORT_S_ALBEDO] > 1.0) return at("surface_albedo must be in [0, 1]");
  if (b->flags & DISORT_F_THERMAL) {
    if (!b->temper) return at("temperature size < num_layers + 1");
    for (int l = 0; l <= L; ++l)
      if (b->temper[vo + l] < 0.0) return at("temperature[" + std::to_string(l) + "] must be >= 0");
    if (s[DISORT_S_WN_LO] < 0.0 || s[DISORT_S_WN_LO] > s[DISORT_S_WN_HI])
      return at("wavenumber: need 0 <= wavenumber_low <= wavenumber_high");
    if (s[DISORT_S_TEMP_TOP] < 0.0 || s[DISORT_S_TEMP_BOT] < 0.0)
      return at("temperature_top and temperature_bottom must be >= 0");
    if (s[DISORT_S_EMIS_TOP] < 0.0 || s[DISORT_S_EMIS_TOP] > 1.0) return at("emissivity_top must be in [0, 1]");
  }
  if (b->flags & DISORT_F_SPHERICAL) {
    if (s[DISORT_S_RADIUS] <= 0.0) return at("bottom_radius must be > 0 when use_spherical_beam is true");
    if (!b->alt) return at("level_altitudes size < num_layers + 1");
  }
  return std::string();
}

}  // namespace

// ---------------------------------------------------------------------------
// Plans: bucket schedule computed from host-side shapes.
// ---------------------------------------------------------------------------
// A bucket: the instances of one stream count, or one chunk of them (a
// contiguous share of the stream count's instance list).
struct Bucket {
  int nstr = 0;
  int variant = 0;
  int part = 0, parts = 1;  // chunk index and chunk count of this stream count
  double tuned_ms[4] = {-1.0, -1.0, -1.0, -1.0};  // per variant, from the autotuner
  int64_t nb = 0, n_tasks = 0, n_lv = 0;
  double cost = 0.0;  // relative cost used to place it on a stream (model, or the tuned time)
  int32_t* d_inst = nullptr;
  int64_t* d_task0 = nullptr;
  int32_t* d_t_inst = nullptr;
  int32_t* d_t_lc = nullptr;
  int32_t* d_v_bk = nullptr;
  int32_t* d_v_lu = nullptr;
  std::vector<void*> dev;  // device arrays above
};

struct disort_flux_plan {
  int64_t n = 0, total_layers = 0, total_levels = 0;
  int32_t flags = 0;
  std::vector<int64_t> lay_off, mom_off;
  int64_t* d_lay_off = nullptr;
  int64_t* d_mom_off = nullptr;
  int32_t* d_nstr = nullptr;
  int32_t* d_nlyr = nullptr;
  int32_t* d_pstride = nullptr;
  std::vector<Bucket> buckets;
  int64_t max_tasks = 0, max_eig = 0, max_panel = 0, max_cvec = 0;
  std::vector<void*> owned;
  // Shapes kept to rebuild the buckets when the chunking changes.
  std::map<int, std::vector<int32_t>> groups;  // stream count -> instances
  std::vector<int32_t> nlyr_h;
  std::map<int, int> chunks;      // stream count -> chunks (absent: 1)
  std::map<int, int> variant_of;  // stream count -> kernel variant
  std::map<int, double> tuned_of;  // stream count -> tuned time of the chosen variant (ms)
  // Schedule: buckets one after another on the context stream (streams = 1),
  // or concurrently on a pool of `streams` side streams with bucket-private
  // scratch, each bucket on the least loaded stream, largest first.
  int streams = 1;
  bool concurrent = false;  // streams > 1
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  // Captured execution (disort_flux_execute_graph) and the batch it was
  // captured for.
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  disort_flux_batch gkey{};
  std::vector<const void*> gbufs;  // scratch and tables the graph was captured with
  int gstreams = 1;
  std::vector<int> gvariant;
  std::string report;  // autotuner log (JSON)
};

// ---------------------------------------------------------------------------
// Execution of a plan.
// ---------------------------------------------------------------------------
namespace {

void release_graph(disort_flux_plan* plan) {
  if (plan->gexec) (void)hipGraphExecDestroy(plan->gexec);
  if (plan->graph) (void)hipGraphDestroy(plan->graph);
  plan->gexec = nullptr;
  plan->graph = nullptr;
}

// Every resource a plan holds (the caller has synchronised its stream).
void free_plan(disort_flux_plan* plan) {
  release_graph(plan);
  for (hipStream_t st : plan->side) (void)hipStreamDestroy(st);
  for (hipEvent_t e : plan->join) (void)hipEventDestroy(e);
  if (plan->fork) (void)hipEventDestroy(plan->fork);
  for (Bucket& bk : plan->buckets)
    for (void* q : bk.dev) (void)hipFree(q);
  for (void* p : plan->owned) (void)hipFree(p);
  delete plan;
}

// Batch-level arguments and every scratch buffer the plan needs, allocated
// before any launch (so that a capture never meets an allocation).
struct PlanArgs {
  FluxArgs base{};
  std::vector<double*> eig, panel, cvec;  // per bucket
};

int prepare(disort_ctx* ctx, disort_flux_plan* plan, const disort_flux_batch* b, PlanArgs& pa) {
  if (!ctx || !plan || !b) return fail(ctx, DISORT_E_INVALID, "null argument");
  if (b->n != plan->n) return fail(ctx, DISORT_E_INVALID, "batch size differs from the plan");
  if ((b->flags & DISORT_F_THERMAL) && !b->temper) return fail(ctx, DISORT_E_INVALID, "thermal batch without temperatures");
  if ((b->flags & DISORT_F_SPHERICAL) && !b->alt)
    return fail(ctx, DISORT_E_INVALID, "spherical batch without level altitudes");
  if ((b->flags & DISORT_F_SPHERICAL) != (plan->flags & DISORT_F_SPHERICAL))
    return fail(ctx, DISORT_E_INVALID, "the batch's spherical flag differs from the plan's");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  FluxArgs& a = pa.base;
  a = FluxArgs{};
  a.flags = b->flags;
  a.nstr = plan->d_nstr;
  a.nlyr = plan->d_nlyr;
  a.pstride = plan->d_pstride;
  a.lay_off = plan->d_lay_off;
  a.mom_off = plan->d_mom_off;
  a.dtau = b->dtau;
  a.ssa = b->ssa;
  a.pmom = b->pmom;
  a.temper = b->temper;
  a.alt = b->alt;
  a.scal = b->scal;
  a.out = b->out;
  a.total_levels = plan->total_levels;
  a.total_layers = plan->total_layers;
  int rc;
  void* p;
  if ((rc = get_buf(ctx, "inst_d", plan->n * dsa::kInstD * sizeof(double), &p))) return rc;
  a.inst_d = (double*)p;
  if ((rc = get_buf(ctx, "inst_i", plan->n * dsa::kInstI * sizeof(int32_t), &p))) return rc;
  a.inst_i = (int32_t*)p;
  if ((rc = get_buf(ctx, "lay_d", plan->total_layers * dsa::kLayD * sizeof(double), &p))) return rc;
  a.lay_d = (double*)p;
  if ((rc = get_buf(ctx, "lev_d", plan->total_levels * dsa::kLevD * sizeof(double), &p))) return rc;
  a.lev_d = (double*)p;
  const size_t nb = plan->buckets.size();
  pa.eig.assign(nb, nullptr);
  pa.panel.assign(nb, nullptr);
  pa.cvec.assign(nb, nullptr);
  if (plan->streams <= 1) {
    if ((rc = get_buf(ctx, "eig", plan->max_eig * sizeof(double), &p))) return rc;
    std::fill(pa.eig.begin(), pa.eig.end(), (double*)p);
    if ((rc = get_buf(ctx, "panel", plan->max_panel * sizeof(double), &p))) return rc;
    std::fill(pa.panel.begin(), pa.panel.end(), (double*)p);
    if ((rc = get_buf(ctx, "cvec", plan->max_cvec * sizeof(double), &p))) return rc;
    std::fill(pa.cvec.begin(), pa.cvec.end(), (double*)p);
  } else {
    // Bucket-private scratch, named by bucket index.
    for (size_t k = 0; k < nb; ++k) {
      const Bucket& bk = plan->buckets[k];
      const std::string tag = "#" + std::to_string(k);
      if ((rc = get_buf(ctx, ("eig" + tag).c_str(), (size_t)dsa::EigOff(bk.nstr).stride() * bk.n_tasks * 8, &p))) return rc;
      pa.eig[k] = (double*)p;
      if ((rc = get_buf(ctx, ("panel" + tag).c_str(), (size_t)dsa::panel_stride(bk.nstr) * bk.n_tasks * 8, &p))) return rc;
      pa.panel[k] = (double*)p;
      if ((rc = get_buf(ctx, ("cvec" + tag).c_str(), (size_t)bk.nstr * bk.n_tasks * 8, &p))) return rc;
      pa.cvec[k] = (double*)p;
    }
    while (plan->side.size() < (size_t)plan->streams) {
      hipStream_t st;
      HIPCHK(ctx, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      plan->side.push_back(st);
      hipEvent_t e;
      HIPCHK(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
      plan->join.push_back(e);
    }
    if (!plan->fork) HIPCHK(ctx, hipEventCreateWithFlags(&plan->fork, hipEventDisableTiming));
  }
  for (const Bucket& bk : plan->buckets) {
    double* tab = nullptr;
    if ((rc = table_for(ctx, bk.nstr, &tab))) return rc;
  }
  return DISORT_OK;
}

FluxArgs bucket_args(disort_ctx* ctx, const PlanArgs& pa, const Bucket& bk, size_t k) {
  FluxArgs a = pa.base;
  a.tab = ctx->tables[bk.nstr];
  a.nb = (int32_t)bk.nb;
  a.b_inst = bk.d_inst;
  a.b_task0 = bk.d_task0;
  a.n_tasks = bk.n_tasks;
  a.t_inst = bk.d_t_inst;
  a.t_lc = bk.d_t_lc;
  a.n_lvtasks = bk.n_lv;
  a.v_bk = bk.d_v_bk;
  a.v_lu = bk.d_v_lu;
  a.eig = pa.eig[k];
  a.panel = pa.panel[k];
  a.cvec = pa.cvec[k];
  return a;
}

int launch_bucket(disort_ctx* ctx, const FluxArgs& a, const Bucket& bk, hipStream_t st, hipEvent_t* ev) {
  // The temperature-Jacobian kernels (a.jbase) exist in the register form only.
  const bool ok = (bk.variant != DISORT_V_WORKGROUP || a.jbase)
                      ? dsa::launch_flux_bucket(bk.nstr, a, st, ev, bk.variant == DISORT_V_LANE_PAIR)
                                              : dsa::launch_flux_bucket_wg(bk.nstr, a, st, ev);
  if (!ok) return fail(ctx, DISORT_E_UNSUPPORTED, "no kernels for num_streams=" + std::to_string(bk.nstr));
  HIPCHK(ctx, hipGetLastError());
  return DISORT_OK;
}

// Enqueue the whole plan on the context stream (forking to side streams
// when concurrent).  With `timed`, buckets run one after another and the
// per-stage event times are accumulated into ctx->stage_ms.
int launch_plan(disort_ctx* ctx, disort_flux_plan* plan, const PlanArgs& pa, bool timed) {
  int rc;
  for (int k = 0; k < 4; ++k) ctx->stage_ms[k] = 0.0;
  ctx->n_stages = 4;
  if (plan->streams <= 1 || timed) {
    for (size_t k = 0; k < plan->buckets.size(); ++k) {
      const Bucket& bk = plan->buckets[k];
      if ((rc = launch_bucket(ctx, bucket_args(ctx, pa, bk, k), bk, ctx->stream, timed ? ctx->ev : nullptr))) return rc;
      if (timed) {
        HIPCHK(ctx, hipEventSynchronize(ctx->ev[4]));
        for (int q = 0; q < 4; ++q) {
          float ms = 0.f;
          HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[q], ctx->ev[q + 1]));
          ctx->stage_ms[q] += ms;
        }
      }
    }
    return DISORT_OK;
  }
  HIPCHK(ctx, hipEventRecord(plan->fork, ctx->stream));
  // Largest buckets first, each on the stream with the least work so far.
  std::vector<size_t> order(plan->buckets.size());
  for (size_t k = 0; k < order.size(); ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t x, size_t y) { return plan->buckets[x].cost > plan->buckets[y].cost; });
  const int S = std::max(1, std::min<int>(plan->streams, (int)plan->side.size()));
  std::vector<double> load(S, 0.0);
  std::vector<bool> used(S, false);
  for (size_t k : order) {
    const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    hipStream_t st = plan->side[q];
    if (!used[q]) HIPCHK(ctx, hipStreamWaitEvent(st, plan->fork, 0));
    used[q] = true;
    load[q] += plan->buckets[k].cost;
    if ((rc = launch_bucket(ctx, bucket_args(ctx, pa, plan->buckets[k], k), plan->buckets[k], st, nullptr))) return rc;
  }
  for (int q = 0; q < S; ++q) {
    if (!used[q]) continue;
    HIPCHK(ctx, hipEventRecord(plan->join[q], plan->side[q]));
    HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, plan->join[q], 0));
  }
  return DISORT_OK;
}

// Every device address a captured graph bakes in besides the batch arrays.
std::vector<const void*> scratch_key(disort_ctx* ctx, const disort_flux_plan* plan, const PlanArgs& pa) {
  std::vector<const void*> k = {pa.base.inst_d, pa.base.inst_i, pa.base.lay_d, pa.base.lev_d};
  for (size_t q = 0; q < pa.eig.size(); ++q) {
    k.push_back(pa.eig[q]);
    k.push_back(pa.panel[q]);
    k.push_back(pa.cvec[q]);
    k.push_back(ctx->tables[plan->buckets[q].nstr]);
  }
  return k;
}

// Key of a captured graph: every field of the batch descriptor (the graph
// holds the device addresses the kernels read and write).
bool same_batch(const disort_flux_batch& x, const disort_flux_batch& y) {
  return x.n == y.n && x.flags == y.flags && x.nstr == y.nstr && x.nlyr == y.nlyr && x.pstride == y.pstride &&
         x.lay_off == y.lay_off && x.mom_off == y.mom_off && x.dtau == y.dtau && x.ssa == y.ssa &&
         x.pmom == y.pmom && x.temper == y.temper && x.alt == y.alt && x.scal == y.scal && x.out == y.out &&
         x.total_levels == y.total_levels && x.status == y.status;
}

// (Re)build the buckets from the stream-count groups: one bucket per
// stream count, or `chunks` contiguous chunks of its instance list.  Device
// arrays of the previous buckets are freed; variants and tuned times are
// kept per stream count.
int build_buckets(disort_ctx* ctx, disort_flux_plan* plan) {
  if (!plan->buckets.empty()) HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  release_graph(plan);
  for (Bucket& bk : plan->buckets)
    for (void* q : bk.dev) (void)hipFree(q);
  plan->buckets.clear();
  plan->max_tasks = plan->max_eig = plan->max_panel = plan->max_cvec = 0;
  for (auto& kv : plan->groups) {
    const int nstr = kv.first;
    const std::vector<int32_t>& all = kv.second;
    const auto ch = plan->chunks.find(nstr);
    const int parts = (int)std::max<int64_t>(1, std::min<int64_t>(ch == plan->chunks.end() ? 1 : ch->second, (int64_t)all.size()));
    int64_t total_tasks = 0;
    for (int32_t i : all) total_tasks += plan->nlyr_h[i];
    for (int part = 0; part < parts; ++part) {
      Bucket bk;
      bk.nstr = nstr;
      bk.part = part;
      bk.parts = parts;
      bk.variant = plan->variant_of[nstr];
      const size_t lo = all.size() * part / parts, hi = all.size() * (part + 1) / parts;
      const std::vector<int32_t> inst(all.begin() + lo, all.begin() + hi);
      bk.nb = (int64_t)inst.size();
      std::vector<int64_t> task0(bk.nb);
      std::vector<int32_t> t_inst, t_lc, v_bk, v_lu;
      for (int64_t k = 0; k < bk.nb; ++k) {
        const int32_t i = inst[k];
        task0[k] = (int64_t)t_inst.size();
        for (int lc = 0; lc < plan->nlyr_h[i]; ++lc) {
          t_inst.push_back(i);
          t_lc.push_back(lc);
        }
        for (int lu = 0; lu <= plan->nlyr_h[i]; ++lu) {
          v_bk.push_back((int32_t)k);
          v_lu.push_back(lu);
        }
      }
      bk.n_tasks = (int64_t)t_inst.size();
      bk.n_lv = (int64_t)v_bk.size();
      const auto tu = plan->tuned_of.find(nstr);
      bk.cost = (tu != plan->tuned_of.end() ? tu->second : 1e-9 * total_tasks * nstr * nstr * nstr) *
                ((double)bk.n_tasks / (double)std::max<int64_t>(total_tasks, 1));
      auto upload = [&](const void* src, size_t bytes, void** dst) -> int {
        HIPCHK(ctx, hipMalloc(dst, std::max<size_t>(bytes, 8)));
        bk.dev.push_back(*dst);
        HIPCHK(ctx, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return DISORT_OK;
      };
      int rc;
      if ((rc = upload(inst.data(), inst.size() * 4, (void**)&bk.d_inst)) ||
          (rc = upload(task0.data(), task0.size() * 8, (void**)&bk.d_task0)) ||
          (rc = upload(t_inst.data(), t_inst.size() * 4, (void**)&bk.d_t_inst)) ||
          (rc = upload(t_lc.data(), t_lc.size() * 4, (void**)&bk.d_t_lc)) ||
          (rc = upload(v_bk.data(), v_bk.size() * 4, (void**)&bk.d_v_bk)) ||
          (rc = upload(v_lu.data(), v_lu.size() * 4, (void**)&bk.d_v_lu))) {
        for (void* q : bk.dev) (void)hipFree(q);
        return rc;
      }
      plan->max_tasks = std::max(plan->max_tasks, bk.n_tasks);
      plan->max_eig = std::max<int64_t>(plan->max_eig, (int64_t)dsa::EigOff(bk.nstr).stride() * bk.n_tasks);
      plan->max_panel = std::max<int64_t>(plan->max_panel, (int64_t)dsa::panel_stride(bk.nstr) * bk.n_tasks);
      plan->max_cvec = std::max<int64_t>(plan->max_cvec, (int64_t)bk.nstr * bk.n_tasks);
      plan->buckets.push_back(std::move(bk));
    }
  }
  return DISORT_OK;
}

}  // namespace

extern "C" {

const char* disort_version(void) { return "disort_amd 0.1 (gfx950)"; }

int disort_quadrature(int nstr, double* cmu, double* cwt) {
  if (nstr < 4 || nstr % 2 != 0 || !cmu || !cwt) return DISORT_E_INVALID;
  std::vector<double> m, w;
  if (!gauss_nodes(nstr / 2, m, w)) return DISORT_E_NUMERIC;
  std::copy(m.begin(), m.end(), cmu);
  std::copy(w.begin(), w.end(), cwt);
  return DISORT_OK;
}

int disort_planck_band(double lo, double hi, int64_t n, const double* t, double* b) {
  if (!(lo >= 0.0) || !(hi >= lo) || n < 0 || (n > 0 && (!t || !b))) return DISORT_E_INVALID;
  for (int64_t k = 0; k < n; ++k) {
    if (!(t[k] >= 0.0)) return DISORT_E_INVALID;
    b[k] = dsa::planck_band(lo, hi, t[k]);
  }
  return DISORT_OK;
}

int disort_planck_band_deriv(double lo, double hi, int64_t n, const double* t, double* d) {
  if (!(lo >= 0.0) || !(hi >= lo) || n < 0 || (n > 0 && (!t || !d))) return DISORT_E_INVALID;
  for (int64_t k = 0; k < n; ++k) {
    if (!(t[k] >= 0.0)) return DISORT_E_INVALID;
    d[k] = dsa::planck_band_deriv(lo, hi, t[k]);
  }
  return DISORT_OK;
}

int disort_ctx_create(int device, disort_ctx** out) {
  if (!out) return DISORT_E_INVALID;
  auto ctx = std::make_unique<disort_ctx>();
  ctx->device = device;
  HIPCHK(ctx, hipSetDevice(device));
  HIPCHK(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  for (auto& e : ctx->ev) HIPCHK(ctx, hipEventCreate(&e));
  for (auto& e : ctx->gev) HIPCHK(ctx, hipEventCreate(&e));
  *out = ctx.release();
  return DISORT_OK;
}

int disort_ctx_destroy(disort_ctx* ctx) {
  if (!ctx) return DISORT_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->bufs)
    if (kv.second.p) (void)hipFree(kv.second.p);
  for (auto& kv : ctx->tables) (void)hipFree(kv.second);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->gev)
    if (e) (void)hipEventDestroy(e);
  for (auto& kv : ctx->mode_tables) (void)hipFree(kv.second);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return DISORT_OK;
}

int disort_ctx_device(const disort_ctx* ctx) { return ctx ? ctx->device : -1; }

const char* disort_last_error(const disort_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void* disort_ctx_stream(disort_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int disort_sync(disort_ctx* ctx) {
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return DISORT_OK;
}

int disort_set_timing(disort_ctx* ctx, int enable) {
  ctx->timing = enable != 0;
  return DISORT_OK;
}

int disort_stage_times(disort_ctx* ctx, double* ms, int max_stages) {
  const int n = std::min(max_stages, ctx->n_stages);
  for (int k = 0; k < n; ++k) ms[k] = (ctx->n_stages == 8) ? ctx->gstage_ms[k] : ctx->stage_ms[k];
  return n;
}

int disort_device_alloc(disort_ctx* ctx, size_t bytes, void** ptr) {
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipMalloc(ptr, std::max<size_t>(bytes, 1)));
  return DISORT_OK;
}

int disort_device_free(disort_ctx* ctx, void* ptr) {
  HIPCHK(ctx, hipFree(ptr));
  return DISORT_OK;
}

int disort_memcpy_h2d(disort_ctx* ctx, void* dst, const void* src, size_t bytes) {
  HIPCHK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return DISORT_OK;
}

int disort_memcpy_d2h(disort_ctx* ctx, void* dst, const void* src, size_t bytes) {
  HIPCHK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return DISORT_OK;
}

int disort_flux_plan_destroy(disort_ctx* ctx, disort_flux_plan* plan) {
  if (!plan) return DISORT_OK;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  free_plan(plan);
  return DISORT_OK;
}

int disort_flux_plan_set_variant(disort_ctx* ctx, disort_flux_plan* plan, int nstr, int variant) {
  if (!plan) return fail(ctx, DISORT_E_INVALID, "null plan");
  for (Bucket& bk : plan->buckets)
    if (bk.nstr == nstr) {
      if (variant == DISORT_V_DEFAULT) variant = default_variant(nstr);
      if ((plan->flags & DISORT_F_SPHERICAL) && variant != DISORT_V_WORKGROUP)
        return fail(ctx, DISORT_E_UNSUPPORTED, "the spherical beam needs the workgroup kernels");
      if (!variant_ok(nstr, variant))
        return fail(ctx, DISORT_E_UNSUPPORTED,
                    "kernel variant " + std::to_string(variant) + " has no kernels for num_streams=" + std::to_string(nstr));
      for (Bucket& b2 : plan->buckets)
        if (b2.nstr == nstr) b2.variant = variant;
      plan->variant_of[nstr] = variant;
      return DISORT_OK;
    }
  return fail(ctx, DISORT_E_INVALID, "the plan has no bucket with num_streams=" + std::to_string(nstr));
}

int disort_flux_plan_get_variant(disort_ctx* ctx, disort_flux_plan* plan, int nstr, int* variant) {
  if (!plan || !variant) return fail(ctx, DISORT_E_INVALID, "null argument");
  for (Bucket& bk : plan->buckets)
    if (bk.nstr == nstr) {
      *variant = bk.variant;
      return DISORT_OK;
    }
  return fail(ctx, DISORT_E_INVALID, "the plan has no bucket with num_streams=" + std::to_string(nstr));
}

int disort_flux_plan_create(disort_ctx* ctx, int64_t n, const int32_t* nstr, const int32_t* nlyr,
                            const int32_t* pstride, int32_t flags, disort_flux_plan** out) {
  if (!ctx || !out || n <= 0 || !nstr || !nlyr || !pstride) return fail(ctx, DISORT_E_INVALID, "invalid plan arguments");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  auto plan = std::make_unique<disort_flux_plan>();
  plan->n = n;
  plan->flags = flags;
  plan->lay_off.resize(n);
  plan->mom_off.resize(n);
  int64_t lay = 0, mom = 0;
  std::map<int, std::vector<int32_t>> groups;
  for (int64_t i = 0; i < n; ++i) {
    if (nlyr[i] <= 0) return fail(ctx, DISORT_E_INVALID, "instance " + std::to_string(i) + ": num_layers must be > 0");
    if (nstr[i] < 4 || nstr[i] % 2)
      return fail(ctx, DISORT_E_INVALID, "instance " + std::to_string(i) + ": num_streams must be even and >= 4");
    if (!gpu_supported_nstr(nstr[i]))
      return fail(ctx, DISORT_E_UNSUPPORTED,
                  "instance " + std::to_string(i) + ": num_streams=" + std::to_string(nstr[i]) +
                      " is not supported by the GPU flux path (4, 6, ..., 16, 32, 64)");
    if (pstride[i] < nstr[i] + 1 + ((flags & DISORT_F_DELTA_M_PLUS) ? 1 : 0))
      return fail(ctx, DISORT_E_INVALID, "instance " + std::to_string(i) + ": phase moment rows too short");
    plan->lay_off[i] = lay;
    plan->mom_off[i] = mom;
    lay += nlyr[i];
    mom += (int64_t)nlyr[i] * pstride[i];
    groups[nstr[i]].push_back((int32_t)i);
  }
  plan->total_layers = lay;
  plan->total_levels = lay + n;
  auto upload = [&](const void* src, size_t bytes, void** dst) -> int {
    HIPCHK(ctx, hipMalloc(dst, std::max<size_t>(bytes, 8)));
    plan->owned.push_back(*dst);
    HIPCHK(ctx, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return DISORT_OK;
  };
  int rc;
  if ((rc = upload(plan->lay_off.data(), n * sizeof(int64_t), (void**)&plan->d_lay_off))) return rc;
  if ((rc = upload(plan->mom_off.data(), n * sizeof(int64_t), (void**)&plan->d_mom_off))) return rc;
  if ((rc = upload(nstr, n * sizeof(int32_t), (void**)&plan->d_nstr))) return rc;
  if ((rc = upload(nlyr, n * sizeof(int32_t), (void**)&plan->d_nlyr))) return rc;
  if ((rc = upload(pstride, n * sizeof(int32_t), (void**)&plan->d_pstride))) return rc;
  plan->nlyr_h.assign(nlyr, nlyr + n);
  for (auto& kv : groups)
    // The pseudo-spherical beam source is built by the workgroup kernels only.
    plan->variant_of[kv.first] = (flags & DISORT_F_SPHERICAL) ? DISORT_V_WORKGROUP : default_variant(kv.first);
  plan->groups = std::move(groups);
  if ((rc = build_buckets(ctx, plan.get()))) {
    free_plan(plan.release());
    return rc;
  }
  *out = plan.release();
  return DISORT_OK;
}

// Launch a planned solve on device-resident batch arrays (asynchronous).
// The batch's lay_off/mom_off are ignored: the plan's offsets are used.
int disort_flux_execute(disort_ctx* ctx, disort_flux_plan* plan, const disort_flux_batch* b) {
  PlanArgs pa;
  int rc;
  if ((rc = prepare(ctx, plan, b, pa))) return rc;
  return launch_plan(ctx, plan, pa, ctx->timing);
}

int disort_flux_plan_set_streams(disort_ctx* ctx, disort_flux_plan* plan, int streams) {
  if (!plan) return fail(ctx, DISORT_E_INVALID, "null plan");
  if (streams < 1 || streams > kMaxSideStreams)
    return fail(ctx, DISORT_E_INVALID, "streams must be in 1.." + std::to_string(kMaxSideStreams));
  plan->streams = streams;
  plan->concurrent = streams > 1;
  return DISORT_OK;
}

int disort_flux_plan_set_concurrent(disort_ctx* ctx, disort_flux_plan* plan, int on) {
  return disort_flux_plan_set_streams(ctx, plan, on ? kDefaultSideStreams : 1);
}

int disort_flux_plan_set_chunks(disort_ctx* ctx, disort_flux_plan* plan, int nstr, int chunks) {
  if (!plan) return fail(ctx, DISORT_E_INVALID, "null plan");
  if (chunks < 1 || chunks > kMaxChunks) return fail(ctx, DISORT_E_INVALID, "chunks must be in 1.." + std::to_string(kMaxChunks));
  if (!plan->groups.count(nstr))
    return fail(ctx, DISORT_E_INVALID, "the plan has no bucket with num_streams=" + std::to_string(nstr));
  plan->chunks[nstr] = chunks;
  return build_buckets(ctx, plan);
}

int disort_flux_execute_graph(disort_ctx* ctx, disort_flux_plan* plan, const disort_flux_batch* b) {
  PlanArgs pa;
  int rc;
  if ((rc = prepare(ctx, plan, b, pa))) return rc;
  std::vector<int> variants;
  for (const Bucket& bk : plan->buckets) variants.push_back(bk.variant);
  const std::vector<const void*> bufs = scratch_key(ctx, plan, pa);
  if (!plan->gexec || !same_batch(plan->gkey, *b) || plan->gstreams != plan->streams || plan->gvariant != variants ||
      plan->gbufs != bufs) {
    release_graph(plan);
    HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    rc = launch_plan(ctx, plan, pa, false);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(ctx, DISORT_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    plan->graph = g;
    HIPCHK(ctx, hipGraphInstantiate(&plan->gexec, g, nullptr, nullptr, 0));
    plan->gkey = *b;
    plan->gbufs = bufs;
    plan->gstreams = plan->streams;
    plan->gvariant = variants;
  }
  HIPCHK(ctx, hipGraphLaunch(plan->gexec, ctx->stream));
  return DISORT_OK;
}

// Autotuner.  (1) For every stream count, time each kernel variant that
// exists for it on the whole bucket (median of `reps` runs after one
// warm-up, hipEvents on the context stream) and keep the fastest.  (2) Time
// the bucket schedule: serial on the context stream, or concurrent on a pool
// of 2 or 3 side streams with the buckets larger than half a stream's share
// split into 2 or 3 chunks (each chunk's own kernels; a chunk's tail overlaps
// the other streams' work), and keep the fastest.  Runs the real solve on the
// given device batch (its outputs are overwritten).
int disort_flux_autotune(disort_ctx* ctx, disort_flux_plan* plan, const disort_flux_batch* b, int reps) {
  if (!ctx || !plan || !b) return fail(ctx, DISORT_E_INVALID, "null argument");
  if (reps < 1) reps = 3;
  int rc;
  plan->chunks.clear();
  plan->streams = 1;
  plan->concurrent = false;
  if ((rc = build_buckets(ctx, plan))) return rc;
  PlanArgs pa;
  if ((rc = prepare(ctx, plan, b, pa))) return rc;
  auto time_it = [&](auto&& enqueue, double* ms_out) -> int {
    std::vector<float> t;
    for (int r = 0; r <= reps; ++r) {
      HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
      int rc2 = enqueue();
      if (rc2) return rc2;
      HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
      HIPCHK(ctx, hipEventSynchronize(ctx->ev[1]));
      float ms = 0.f;
      HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
      if (r > 0) t.push_back(ms);
    }
    std::sort(t.begin(), t.end());
    *ms_out = t[t.size() / 2];
    return DISORT_OK;
  };
  std::string js = "{\"buckets\": [";
  double total_ms = 0.0;
  for (size_t k = 0; k < plan->buckets.size(); ++k) {
    Bucket& bk = plan->buckets[k];
    const FluxArgs a = bucket_args(ctx, pa, bk, k);
    int best = bk.variant;
    double best_ms = 1e300;
    for (int v : {DISORT_V_LANE, DISORT_V_LANE_PAIR, DISORT_V_WORKGROUP}) {
      if (!variant_ok(bk.nstr, v)) continue;
      if ((plan->flags & DISORT_F_SPHERICAL) && v != DISORT_V_WORKGROUP) continue;
      bk.variant = v;
      double ms = 0.0;
      if ((rc = time_it([&]() { return launch_bucket(ctx, a, bk, ctx->stream, nullptr); }, &ms))) return rc;
      bk.tuned_ms[v] = ms;
      if (ms < best_ms) {
        best_ms = ms;
        best = v;
      }
    }
    bk.variant = best;
    plan->variant_of[bk.nstr] = best;
    plan->tuned_of[bk.nstr] = best_ms;
    total_ms += best_ms;
    char buf[320];
    std::snprintf(buf, sizeof buf,
                  "%s{\"nstr\": %d, \"instances\": %lld, \"layers\": %lld, \"lane_ms\": %.6g, \"lane_pair_ms\": %.6g, "
                  "\"workgroup_ms\": %.6g, \"chosen\": \"%s\"}",
                  k ? ", " : "", bk.nstr, (long long)bk.nb, (long long)bk.n_tasks, bk.tuned_ms[DISORT_V_LANE],
                  bk.tuned_ms[DISORT_V_LANE_PAIR], bk.tuned_ms[DISORT_V_WORKGROUP], variant_name(best));
    js += buf;
  }
  js += "], \"schedules\": [";
  // The variant of the bucket with the most work interacts with the
  // schedule (a variant with fewer registers overlaps differently with the
  // other chunks' kernels), so the schedule search times every valid
  // variant of that stream count under every schedule; the other stream
  // counts keep their isolated choice.
  int dom = -1;
  double dom_ms = -1.0;
  for (auto& kv : plan->tuned_of)
    if (kv.second > dom_ms) dom = kv.first, dom_ms = kv.second;
  std::vector<int> dom_variants;
  for (int v : {DISORT_V_LANE, DISORT_V_LANE_PAIR, DISORT_V_WORKGROUP})
    if (dom >= 0 && variant_ok(dom, v) && (!(plan->flags & DISORT_F_SPHERICAL) || v == DISORT_V_WORKGROUP))
      dom_variants.push_back(v);
  if (dom_variants.empty()) dom_variants.push_back(dom >= 0 ? plan->variant_of[dom] : DISORT_V_DEFAULT);
  struct Cand {
    int streams, chunks, variant;
    double ms;
  };
  std::vector<Cand> cands;
  for (int v : dom_variants)
    for (auto sc : {std::pair<int, int>{1, 1}, {2, 1}, {2, 2}, {3, 1}, {3, 2}, {3, 3}})
      cands.push_back({sc.first, sc.second, v, 0.0});
  auto apply = [&](int S, int c, int v) -> int {
    plan->streams = S;
    plan->concurrent = S > 1;
    plan->chunks.clear();
    if (dom >= 0) plan->variant_of[dom] = v;
    for (auto& kv : plan->tuned_of)
      if (S > 1 && kv.second >= 0.5 * total_ms / S) plan->chunks[kv.first] = c;
    return build_buckets(ctx, plan);
  };
  size_t best = 0;
  double serial_ms = -1.0;  // the isolated choice run serially
  const int iso = dom >= 0 ? plan->variant_of[dom] : DISORT_V_DEFAULT;
  for (size_t q = 0; q < cands.size(); ++q) {
    Cand& cd = cands[q];
    if ((rc = apply(cd.streams, cd.chunks, cd.variant))) return rc;
    PlanArgs pc;
    if ((rc = prepare(ctx, plan, b, pc))) return rc;
    if ((rc = time_it([&]() { return launch_plan(ctx, plan, pc, false); }, &cd.ms))) return rc;
    if (cd.ms < cands[best].ms) best = q;
    if (cd.streams == 1 && cd.variant == iso) serial_ms = cd.ms;
    char buf[200];
    std::snprintf(buf, sizeof buf,
                  "%s{\"streams\": %d, \"chunks\": %d, \"variant\": \"%s\", \"buckets\": %zu, \"ms\": %.6g}",
                  q ? ", " : "", cd.streams, cd.chunks, variant_name(cd.variant), plan->buckets.size(), cd.ms);
    js += buf;
  }
  if ((rc = apply(cands[best].streams, cands[best].chunks, cands[best].variant))) return rc;
  char buf[360];
  std::snprintf(buf, sizeof buf,
                "], \"serial_ms\": %.6g, \"best_ms\": %.6g, \"streams\": %d, \"chunks\": %d, \"dominant_nstr\": %d, "
                "\"dominant_variant\": \"%s\", \"schedule\": \"%s\"}",
                serial_ms, cands[best].ms, cands[best].streams, cands[best].chunks, dom,
                variant_name(cands[best].variant), plan->streams > 1 ? "concurrent" : "serial");
  js += buf;
  plan->report = js;
  return DISORT_OK;
}

const char* disort_flux_plan_report(disort_flux_plan* plan) { return plan ? plan->report.c_str() : ""; }

// Per-instance status codes of the last execute (copied to host).
int disort_flux_status(disort_ctx* ctx, disort_flux_plan* plan, int32_t* status_host) {
  void* p = nullptr;
  int rc;
  if ((rc = get_buf(ctx, "inst_i", plan->n * dsa::kInstI * sizeof(int32_t), &p))) return rc;
  std::vector<int32_t> ii(plan->n * dsa::kInstI);
  HIPCHK(ctx, hipMemcpyAsync(ii.data(), p, ii.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int64_t i = 0; i < plan->n; ++i) status_host[i] = ii[i * dsa::kInstI + dsa::II_STATUS];
  return DISORT_OK;
}

int disort_flux_solve(disort_ctx* ctx, const disort_flux_batch* b, int memspace) {
  if (!ctx || !b) return fail(ctx, DISORT_E_INVALID, "null argument");
  if (memspace == DISORT_MEM_DEVICE)
    return fail(ctx, DISORT_E_INVALID, "device-resident batches go through disort_flux_plan_create/disort_flux_execute");
  if (b->n <= 0) return DISORT_OK;
  for (int64_t i = 0; i < b->n; ++i) {
    const std::string m = validate_instance(b, i);
    if (!m.empty()) return fail(ctx, DISORT_E_INVALID, m);
  }
  disort_flux_plan* plan = nullptr;
  int rc = disort_flux_plan_create(ctx, b->n, b->nstr, b->nlyr, b->pstride, b->flags, &plan);
  if (rc) return rc;
  std::unique_ptr<disort_flux_plan, void (*)(disort_flux_plan*)> guard(plan, free_plan);
  const int64_t TL = plan->total_layers, TV = plan->total_levels;
  int64_t mom_total = 0;
  for (int64_t i = 0; i < b->n; ++i) mom_total += (int64_t)b->nlyr[i] * b->pstride[i];
  struct Up {
    const char* name;
    const void* src;
    size_t bytes;
    void** dst;
  };
  disort_flux_batch d = *b;
  void *p_dtau, *p_ssa, *p_pmom, *p_temp = nullptr, *p_scal, *p_out;
  std::vector<Up> ups = {{"in_dtau", b->dtau, (size_t)TL * 8, &p_dtau},
                         {"in_ssa", b->ssa, (size_t)TL * 8, &p_ssa},
                         {"in_pmom", b->pmom, (size_t)mom_total * 8, &p_pmom},
                         {"in_scal", b->scal, (size_t)b->n * DISORT_NSCAL * 8, &p_scal}};
  void* p_alt = nullptr;
  if (b->flags & DISORT_F_THERMAL) ups.push_back({"in_temp", b->temper, (size_t)TV * 8, &p_temp});
  if (b->flags & DISORT_F_SPHERICAL) ups.push_back({"in_alt", b->alt, (size_t)TV * 8, &p_alt});
  for (auto& u : ups) {
    if ((rc = get_buf(ctx, u.name, u.bytes, u.dst))) return rc;
    HIPCHK(ctx, hipMemcpyAsync(*u.dst, u.src, u.bytes, hipMemcpyHostToDevice, ctx->stream));
  }
  if ((rc = get_buf(ctx, "out", (size_t)DISORT_NOUT * TV * 8, &p_out))) return rc;
  d.dtau = (const double*)p_dtau;
  d.ssa = (const double*)p_ssa;
  d.pmom = (const double*)p_pmom;
  d.temper = (const double*)p_temp;
  d.alt = (const double*)p_alt;
  d.scal = (const double*)p_scal;
  d.out = (double*)p_out;
  d.total_levels = TV;
  if ((rc = disort_flux_execute(ctx, plan, &d))) return rc;
  HIPCHK(ctx, hipMemcpyAsync(b->out, p_out, (size_t)DISORT_NOUT * TV * 8, hipMemcpyDeviceToHost, ctx->stream));
  std::vector<int32_t> st(b->n);
  if ((rc = disort_flux_status(ctx, plan, st.data()))) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  int64_t bad = 0;
  for (int64_t i = 0; i < b->n; ++i) {
    if (b->status) b->status[i] = st[i];
    bad += st[i] != 0;
  }
  if (bad) return fail(ctx, DISORT_E_NUMERIC, std::to_string(bad) + " instance(s) failed (see status)");
  return DISORT_OK;
}

// ---------------------------------------------------------------------------
// Temperature Jacobians (DisortFluxSolver::computeThermalJacobian,
// src/FluxSolver.cpp:1294-1497).  With the optical properties fixed, every
// output is affine in the level Planck values B_k and in B(T_bottom): its
// derivative with respect to B_k is the solve of the same medium with the
// unit source B = e_k and no other source (no beam, no isotropic or top
// emission), and with respect to B(T_bottom) the solve with only a unit
// surface emission.  The reference obtains the same columns from the same
// linear system by extra right-hand sides on its band LU.  Here each
// (instance, degree of freedom) is one instance of an expanded GPU batch
// whose Planck values are given (FluxArgs::planck_ovr); the columns are then
// scaled by dB/dT (planckFunctionDeriv2, src/Planck.cpp:171-214).
// ---------------------------------------------------------------------------
int disort_flux_jacobian(disort_ctx* ctx, const disort_flux_batch* b, double* jac) {
  if (!ctx || !b || !jac) return fail(ctx, DISORT_E_INVALID, "null argument");
  if (!(b->flags & DISORT_F_THERMAL))
    return fail(ctx, DISORT_E_INVALID, "temperature Jacobians need use_thermal_emission = true");
  if (b->n <= 0) return DISORT_OK;
  for (int64_t i = 0; i < b->n; ++i) {
    const std::string m = validate_instance(b, i);
    if (!m.empty()) return fail(ctx, DISORT_E_INVALID, m);
  }
  const int32_t flags = b->flags & (DISORT_F_THERMAL | DISORT_F_DELTA_M_PLUS | DISORT_F_DIFFUSION_BC);
  // Chunks of instances whose degrees of freedom (nlyr + 2 each) fit an
  // expanded batch of about 16k instances.
  const int64_t kMaxExpanded = 16384;
  int64_t jac_off = 0;
  for (int64_t c0 = 0; c0 < b->n;) {
    int64_t c1 = c0, ne = 0;
    while (c1 < b->n && (c1 == c0 || ne + b->nlyr[c1] + 2 <= kMaxExpanded)) ne += b->nlyr[c1++] + 2;
    std::vector<int32_t> vn, vl, vp;
    std::vector<int64_t> src;  // expanded instance -> batch instance
    std::vector<int> dof;      // expanded instance -> degree of freedom
    vn.reserve(ne), vl.reserve(ne), vp.reserve(ne), src.reserve(ne), dof.reserve(ne);
    for (int64_t i = c0; i < c1; ++i)
      for (int k = 0; k <= b->nlyr[i] + 1; ++k) {
        vn.push_back(b->nstr[i]);
        vl.push_back(b->nlyr[i]);
        vp.push_back(b->pstride[i]);
        src.push_back(i);
        dof.push_back(k);
      }
    std::vector<double> dtau, ssa, pmom, temper, scal, ovr;
    int64_t TV = 0;
    for (int64_t e = 0; e < ne; ++e) TV += vl[e] + 1;
    ovr.assign(TV + 2 * ne, 0.0);
    int64_t v = 0;
    for (int64_t e = 0; e < ne; ++e) {
      const int64_t i = src[e];
      const int L = b->nlyr[i];
      const int64_t lo = b->lay_off[i];
      dtau.insert(dtau.end(), b->dtau + lo, b->dtau + lo + L);
      ssa.insert(ssa.end(), b->ssa + lo, b->ssa + lo + L);
      pmom.insert(pmom.end(), b->pmom + b->mom_off[i], b->pmom + b->mom_off[i] + (int64_t)L * b->pstride[i]);
      temper.insert(temper.end(), b->temper + lo + i, b->temper + lo + i + L + 1);
      double sc[DISORT_NSCAL];
      std::memcpy(sc, b->scal + i * DISORT_NSCAL, sizeof sc);
      sc[DISORT_S_BEAM_FLUX] = 0.0;
      sc[DISORT_S_ISO_TOP] = 0.0;
      sc[DISORT_S_ISO_BOT] = 0.0;
      scal.insert(scal.end(), sc, sc + DISORT_NSCAL);
      if (dof[e] <= L) ovr[v + dof[e]] = 1.0;
      else ovr[TV + 2 * e + 1] = 1.0;
      v += L + 1;
    }
    disort_flux_plan* plan = nullptr;
    int rc = disort_flux_plan_create(ctx, ne, vn.data(), vl.data(), vp.data(), flags, &plan);
    if (rc) return rc;
    std::unique_ptr<disort_flux_plan, void (*)(disort_flux_plan*)> guard(plan, free_plan);
    struct Up {
      const char* name;
      const void* src;
      size_t bytes;
      void** dst;
    };
    void *p_dtau, *p_ssa, *p_pmom, *p_temp, *p_scal, *p_ovr, *p_out;
    const std::vector<Up> ups = {{"jac_dtau", dtau.data(), dtau.size() * 8, &p_dtau},
                                 {"jac_ssa", ssa.data(), ssa.size() * 8, &p_ssa},
                                 {"jac_pmom", pmom.data(), pmom.size() * 8, &p_pmom},
                                 {"jac_temp", temper.data(), temper.size() * 8, &p_temp},
                                 {"jac_scal", scal.data(), scal.size() * 8, &p_scal},
                                 {"jac_ovr", ovr.data(), ovr.size() * 8, &p_ovr}};
    for (const Up& u : ups) {
      if ((rc = get_buf(ctx, u.name, u.bytes, u.dst))) return rc;
      HIPCHK(ctx, hipMemcpyAsync(*u.dst, u.src, u.bytes, hipMemcpyHostToDevice, ctx->stream));
    }
    if ((rc = get_buf(ctx, "jac_out", (size_t)DISORT_NOUT * TV * 8, &p_out))) return rc;
    disort_flux_batch d{};
    d.n = ne;
    d.flags = flags;
    d.dtau = (const double*)p_dtau;
    d.ssa = (const double*)p_ssa;
    d.pmom = (const double*)p_pmom;
    d.temper = (const double*)p_temp;
    d.scal = (const double*)p_scal;
    d.out = (double*)p_out;
    d.total_levels = TV;
    PlanArgs pa;
    if ((rc = prepare(ctx, plan, &d, pa))) return rc;
    pa.base.planck_ovr = (const double*)p_ovr;
    if ((rc = launch_plan(ctx, plan, pa, false))) return rc;
    std::vector<double> out((size_t)DISORT_NOUT * TV);
    HIPCHK(ctx, hipMemcpyAsync(out.data(), p_out, out.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<int32_t> st(ne);
    if ((rc = disort_flux_status(ctx, plan, st.data()))) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t e = 0; e < ne; ++e)
      if (st[e]) return fail(ctx, DISORT_E_NUMERIC, "instance " + std::to_string(src[e]) + ": Jacobian solve failed");
    // jac[i][f][lu][k] = d out_f(lu) / d T_k; f in the order flux_up,
    // flux_down, mean_intensity, mean_intensity_down, mean_intensity_up,
    // flux_tau_divergence (the *_temperature_jac members of FluxResult).
    static const int field[6] = {DISORT_O_FLUX_UP, DISORT_O_FLUX_DOWN, DISORT_O_MEAN_INTENSITY,
                                 DISORT_O_MEAN_INTENSITY_DOWN, DISORT_O_MEAN_INTENSITY_UP,
                                 DISORT_O_FLUX_TAU_DIVERGENCE};
    v = 0;
    int64_t base = jac_off;
    for (int64_t e = 0; e < ne; ++e) {
      const int64_t i = src[e];
      const int L = b->nlyr[i], k = dof[e];
      const double* s2 = b->scal + i * DISORT_NSCAL;
      const double t = (k <= L) ? b->temper[b->lay_off[i] + i + k] : s2[DISORT_S_TEMP_BOT];
      const double dbdt = dsa::planck_band_deriv(s2[DISORT_S_WN_LO], s2[DISORT_S_WN_HI], t);
      const int64_t per = (int64_t)(L + 1) * (L + 2);
      for (int f = 0; f < 6; ++f)
        for (int lu = 0; lu <= L; ++lu)
          jac[base + f * per + (int64_t)lu * (L + 2) + k] = dbdt * out[(size_t)field[f] * TV + v + lu];
      v += L + 1;
      if (k == L + 1) base += 6 * per;
    }
    jac_off = base;
    c0 = c1;
  }
  return DISORT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Device-resident temperature Jacobians of uniform columns: the expansion of
// disort_flux_jacobian done on the device, chunk by chunk, into buffers the
// plan owns (jacobian_kernels.hip), solved by an inner flux plan of
// chunk * (nlyr + 2) instances.
// ---------------------------------------------------------------------------
struct disort_jac_plan {
  int64_t n = 0, chunk = 0, ne = 0;
  int32_t nstr = 0, nlyr = 0, pstride = 0, flags = 0;
  disort_flux_plan* flux = nullptr;
  double *dtau = nullptr, *ssa = nullptr, *pmom = nullptr, *temper = nullptr, *scal = nullptr;
  double *ovr = nullptr, *out = nullptr, *dbdt = nullptr;
  int32_t* status = nullptr;  // per column
  // Factored form: inner plans over chunk columns and over the remainder.
  bool factored = false;
  disort_flux_plan* flux_rem = nullptr;
  double *jbase = nullptr, *jpanel = nullptr, *jcvec = nullptr;
  std::vector<void*> owned;
};

namespace {
void free_jac_plan(disort_jac_plan* jp) {
  if (jp->flux) free_plan(jp->flux);
  if (jp->flux_rem) free_plan(jp->flux_rem);
  for (void* p : jp->owned) (void)hipFree(p);
  delete jp;
}
}  // namespace

extern "C" {

int disort_jac_plan_create(disort_ctx* ctx, int64_t n, int32_t nstr, int32_t nlyr, int32_t pstride, int32_t flags,
                           int64_t max_expanded, disort_jac_plan** out) {
  if (!ctx || !out) return fail(ctx, DISORT_E_INVALID, "null argument");
  *out = nullptr;
  if (n <= 0 || nlyr <= 0) return fail(ctx, DISORT_E_INVALID, "need n > 0 columns of nlyr > 0 layers");
  if (!(flags & DISORT_F_THERMAL))
    return fail(ctx, DISORT_E_INVALID, "temperature Jacobians need use_thermal_emission = true");
  if (flags & ~(DISORT_F_THERMAL | DISORT_F_DELTA_M_PLUS | DISORT_F_DIFFUSION_BC | DISORT_JAC_EXPAND))
    return fail(ctx, DISORT_E_INVALID, "device Jacobians support the thermal, Delta-M+ and diffusion-BC flags only");
  const bool factored = !(flags & DISORT_JAC_EXPAND) && nstr <= 16;
  flags &= ~DISORT_JAC_EXPAND;
  if (!gpu_supported_nstr(nstr)) return fail(ctx, DISORT_E_UNSUPPORTED, "no kernels for num_streams=" + std::to_string(nstr));
  if (pstride < nstr + 1 + ((flags & DISORT_F_DELTA_M_PLUS) ? 1 : 0))
    return fail(ctx, DISORT_E_INVALID, "pstride below the moments the solver reads");
  const int D = nlyr + 2;
  auto* jp = new disort_jac_plan;
  std::unique_ptr<disort_jac_plan, void (*)(disort_jac_plan*)> guard(jp, free_jac_plan);
  jp->n = n;
  jp->nstr = nstr, jp->nlyr = nlyr, jp->pstride = pstride, jp->flags = flags;
  jp->factored = factored;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  auto alloc = [&](size_t bytes, void** p) -> int {
    HIPCHK(ctx, hipMalloc(p, std::max<size_t>(bytes, 256)));
    jp->owned.push_back(*p);
    return DISORT_OK;
  };
  int rc;
  if (factored) {
    // Per column: the layer basis, the panels of each group of kJacRhs
    // right-hand sides, the mode amplitudes per degree of freedom, dB/dT.
    const int64_t L = nlyr, N = nstr, NG = (D + dsa::kJacRhs - 1) / dsa::kJacRhs;
    const int64_t per_col = 8 * (L * N + NG * L * dsa::jac_panel_stride(nstr) + D * L * N + D);
    int64_t chunk = std::max<int64_t>(1, (int64_t(4) << 30) / per_col);
    if (max_expanded > 0) chunk = std::min<int64_t>(chunk, std::max<int64_t>(1, max_expanded / D));
    jp->chunk = std::min<int64_t>(n, chunk);
    const int64_t c = jp->chunk;
    if ((rc = alloc(c * L * N * 8, (void**)&jp->jbase)) ||
        (rc = alloc(c * NG * L * dsa::jac_panel_stride(nstr) * 8, (void**)&jp->jpanel)) ||
        (rc = alloc(c * D * L * N * 8, (void**)&jp->jcvec)) || (rc = alloc(c * D * 8, (void**)&jp->dbdt)) ||
        (rc = alloc((size_t)DISORT_NOUT * c * (L + 1) * 8, (void**)&jp->out)) || (rc = alloc(n * 4, (void**)&jp->status)))
      return rc;
    std::vector<int32_t> vn(c, nstr), vl(c, nlyr), vp(c, pstride);
    if ((rc = disort_flux_plan_create(ctx, c, vn.data(), vl.data(), vp.data(), flags, &jp->flux))) return rc;
    const int64_t rem = n % c;
    if (rem && (rc = disort_flux_plan_create(ctx, rem, vn.data(), vl.data(), vp.data(), flags, &jp->flux_rem))) return rc;
    *out = guard.release();
    return DISORT_OK;
  }
  if (max_expanded <= 0) max_expanded = 131072;
  jp->chunk = std::max<int64_t>(1, std::min<int64_t>(n, max_expanded / D));
  jp->ne = jp->chunk * D;
  const int64_t ne = jp->ne, L = nlyr;
  if ((rc = alloc(ne * L * 8, (void**)&jp->dtau)) || (rc = alloc(ne * L * 8, (void**)&jp->ssa)) ||
      (rc = alloc(ne * L * pstride * 8, (void**)&jp->pmom)) || (rc = alloc(ne * (L + 1) * 8, (void**)&jp->temper)) ||
      (rc = alloc(ne * DISORT_NSCAL * 8, (void**)&jp->scal)) ||
      (rc = alloc((ne * (L + 1) + 2 * ne) * 8, (void**)&jp->ovr)) ||
      (rc = alloc((size_t)DISORT_NOUT * ne * (L + 1) * 8, (void**)&jp->out)) || (rc = alloc(ne * 8, (void**)&jp->dbdt)) ||
      (rc = alloc(n * 4, (void**)&jp->status)))
    return rc;
  std::vector<int32_t> vn(ne, nstr), vl(ne, nlyr), vp(ne, pstride);
  if ((rc = disort_flux_plan_create(ctx, ne, vn.data(), vl.data(), vp.data(), flags, &jp->flux))) return rc;
  *out = guard.release();
  return DISORT_OK;
}

int disort_jac_plan_destroy(disort_ctx* ctx, disort_jac_plan* jp) {
  if (!jp) return DISORT_OK;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  free_jac_plan(jp);
  return DISORT_OK;
}

int64_t disort_jac_plan_chunk(const disort_jac_plan* jp) { return jp ? jp->chunk : 0; }

disort_flux_plan* disort_jac_plan_flux(disort_jac_plan* jp) { return jp ? jp->flux : nullptr; }

int disort_jac_execute(disort_ctx* ctx, disort_jac_plan* jp, const disort_flux_batch* b, double* jac) {
  if (!ctx || !jp || !b || !jac) return fail(ctx, DISORT_E_INVALID, "null argument");
  if (b->n != jp->n) return fail(ctx, DISORT_E_INVALID, "batch size differs from the plan");
  if (!b->dtau || !b->ssa || !b->pmom || !b->temper || !b->scal)
    return fail(ctx, DISORT_E_INVALID, "the batch needs dtau, ssa, pmom, temper and scal");
  const int L = jp->nlyr;
  if (jp->factored) {
    for (int64_t c0 = 0; c0 < jp->n; c0 += jp->chunk) {
      const int64_t cols = std::min(jp->chunk, jp->n - c0);
      disort_flux_plan* fp = (cols == jp->chunk) ? jp->flux : jp->flux_rem;
      disort_flux_batch d{};
      d.n = cols;
      d.flags = jp->flags;
      d.dtau = b->dtau + c0 * L;
      d.ssa = b->ssa + c0 * L;
      d.pmom = b->pmom + c0 * L * jp->pstride;
      d.temper = b->temper + c0 * (L + 1);
      d.scal = b->scal + c0 * DISORT_NSCAL;
      d.out = jp->out;
      d.total_levels = cols * (L + 1);
      PlanArgs pa;
      int rc;
      if ((rc = prepare(ctx, fp, &d, pa))) return rc;
      pa.base.jbase = jp->jbase;
      pa.base.jpanel = jp->jpanel;
      pa.base.jcvec = jp->jcvec;
      pa.base.jdbdt = jp->dbdt;
      pa.base.jac = jac;
      pa.base.jstatus = jp->status;
      pa.base.jac_c0 = c0;
      pa.base.jac_nlyr = L;
      if ((rc = launch_plan(ctx, fp, pa, false))) return rc;
    }
    return DISORT_OK;
  }
  const int64_t ne = jp->ne, TV = ne * (L + 1);
  for (int64_t c0 = 0; c0 < jp->n; c0 += jp->chunk) {
    const int64_t cols = std::min(jp->chunk, jp->n - c0);
    if (!dsa::launch_jac_expand(jp->n, c0, ne, L, jp->pstride, b, jp->dtau, jp->ssa, jp->pmom, jp->temper, jp->scal,
                                jp->ovr, jp->dbdt, ctx->stream))
      return fail(ctx, DISORT_E_HIP, "Jacobian expansion launch failed");
    disort_flux_batch d{};
    d.n = ne;
    d.flags = jp->flags;
    d.dtau = jp->dtau;
    d.ssa = jp->ssa;
    d.pmom = jp->pmom;
    d.temper = jp->temper;
    d.scal = jp->scal;
    d.out = jp->out;
    d.total_levels = TV;
    PlanArgs pa;
    int rc;
    if ((rc = prepare(ctx, jp->flux, &d, pa))) return rc;
    pa.base.planck_ovr = jp->ovr;
    if ((rc = launch_plan(ctx, jp->flux, pa, false))) return rc;
    if (!dsa::launch_jac_scale(cols, c0, L, jp->out, TV, jp->dbdt, jac, pa.base.inst_i, jp->status, ctx->stream))
      return fail(ctx, DISORT_E_HIP, "Jacobian scaling launch failed");
  }
  return DISORT_OK;
}

int disort_jac_status(disort_ctx* ctx, disort_jac_plan* jp, int32_t* status_host) {
  if (!ctx || !jp || !status_host) return fail(ctx, DISORT_E_INVALID, "null argument");
  HIPCHK(ctx, hipMemcpyAsync(status_host, jp->status, jp->n * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return DISORT_OK;
}

}  // extern "C"
