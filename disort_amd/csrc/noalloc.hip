// disort_flux_execute_static: the device-resident flux solve of a uniform
// batch (every instance has the same stream count, layer count and moment
// stride) with every intermediate in one static device array, so that the
// call allocates no device memory.  It launches the same bucket kernels as
// disort_flux_execute; only the origin of the scratch and of the task lists
// differs (the task lists of a uniform batch are filled by a device kernel).
//
// Static scratch layout (`scratch`, one copy per device, kScratchBytes):
//   [0, kTableBytes)        constant tables of the nine stream counts,
//                           uploaded once per device and stream count
//   [kTableBytes, end)      per-call work region: index arrays, per-instance,
//                           per-layer and per-level records, the layer-task
//                           records, panel factors and integration constants
// The work region is rewritten by every call; calls on one device that
// overlap in time (other streams, other contexts) race on it.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "ctx.hpp"
#include "flux_args.hpp"
#include "flux_layout.hpp"

using dsa_host::fail;

namespace dsa {
bool launch_flux_bucket(int nstr, const FluxArgs& a, hipStream_t s, hipEvent_t* ev, bool pair);
bool launch_flux_bucket_wg(int nstr, const FluxArgs& a, hipStream_t s, hipEvent_t* ev);
}  // namespace dsa

namespace {

constexpr size_t kScratchBytes = DISORT_STATIC_SCRATCH_BYTES;
constexpr size_t kTableBytes = size_t(128) << 10;
constexpr int kStreamCounts[9] = {4, 6, 8, 10, 12, 14, 16, 32, 64};
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceScratch {
  unsigned char* ptr = nullptr;  // address of this device's copy of `scratch`
  unsigned tables = 0;           // bit k: table of kStreamCounts[k] uploaded
};
std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

int count_index(int nstr) {
  for (int k = 0; k < 9; ++k)
    if (kStreamCounts[k] == nstr) return k;
  return -1;
}

// Byte offset of the table of stream count kStreamCounts[k] in the table region.
constexpr size_t table_offset(int k) {
  size_t off = 0;
  for (int q = 0; q < k; ++q) off += ((size_t)dsa::TabOff(kStreamCounts[q]).size() * 8 + 255) & ~size_t(255);
  return off;
}
static_assert(table_offset(9) <= kTableBytes, "constant tables exceed their region");

// Sub-buffers of the work region for n instances of (N, L).
struct WorkLayout {
  size_t nstr, nlyr, pstride, lay_off, mom_off, b_inst, b_task0, inst_i, inst_d, t_inst, t_lc, v_bk, v_lu, lay_d,
      lev_d, eig, panel, cvec, end;
  WorkLayout(int64_t n, int N, int L) {
    const int64_t T = n * L, V = n * (L + 1);
    size_t o = kTableBytes;
    auto take = [&](size_t bytes) {
      const size_t at = o;
      o += (bytes + 255) & ~size_t(255);
      return at;
    };
    nstr = take(n * 4);
    nlyr = take(n * 4);
    pstride = take(n * 4);
    lay_off = take(n * 8);
    mom_off = take(n * 8);
    b_inst = take(n * 4);
    b_task0 = take(n * 8);
    inst_i = take(n * dsa::kInstI * 4);
    inst_d = take(n * dsa::kInstD * 8);
    t_inst = take(T * 4);
    t_lc = take(T * 4);
    v_bk = take(V * 4);
    v_lu = take(V * 4);
    lay_d = take(T * dsa::kLayD * 8);
    lev_d = take(V * dsa::kLevD * 8);
    eig = take(T * (size_t)dsa::EigOff(N).stride() * 8);
    panel = take(T * (size_t)dsa::panel_stride(N) * 8);
    cvec = take(T * (size_t)N * 8);
    end = o;
  }
};

// Index arrays of a uniform batch (what disort_flux_plan_create builds on
// the host for one bucket): instance i owns layer tasks [i L, (i+1) L) and
// level tasks [i (L+1), (i+1)(L+1)).
__global__ __launch_bounds__(256) void k_static_index(int64_t n, int N, int L, int P, int32_t* nstr, int32_t* nlyr,
                                                      int32_t* pstride, int64_t* lay_off, int64_t* mom_off,
                                                      int32_t* b_inst, int64_t* b_task0, int32_t* t_inst,
                                                      int32_t* t_lc, int32_t* v_bk, int32_t* v_lu) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t V = n * (L + 1);
  if (x >= V) return;
  if (x < n) {
    nstr[x] = N;
    nlyr[x] = L;
    pstride[x] = P;
    lay_off[x] = x * L;
    mom_off[x] = x * (int64_t)L * P;
    b_inst[x] = (int32_t)x;
    b_task0[x] = x * L;
  }
  if (x < n * L) {
    t_inst[x] = (int32_t)(x / L);
    t_lc[x] = (int32_t)(x % L);
  }
  v_bk[x] = (int32_t)(x / (L + 1));
  v_lu[x] = (int32_t)(x % (L + 1));
}

__global__ __launch_bounds__(256) void k_static_status(int64_t n, const int32_t* inst_i, int32_t* status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) status[i] = inst_i[i * dsa::kInstI + dsa::II_STATUS];
}

// This device's scratch address, resolved once per device (with the device
// current) by hipGetSymbolAddress; and the table of `nstr` in it.
int device_scratch(disort_ctx* ctx, int nstr, unsigned char** base, const double** tab) {
  const int dev = ctx->device;
  if (dev < 0 || dev >= kMaxDevices) return fail(ctx, DISORT_E_UNSUPPORTED, "device index out of range for the static scratch");
  const int k = count_index(nstr);
  std::lock_guard<std::mutex> lock(g_mu);
  DeviceScratch& d = g_dev[dev];
  if (!d.ptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ctx, DISORT_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    d.ptr = (unsigned char*)p;
  }
  if (!(d.tables & (1u << k))) {
    std::vector<double> t;
    if (!dsa_host::build_table(nstr, t)) return fail(ctx, DISORT_E_NUMERIC, "gaussianQuadrature: Failed to converge");
    HIPCHK(ctx, hipMemcpy(d.ptr + table_offset(k), t.data(), t.size() * 8, hipMemcpyHostToDevice));
    d.tables |= 1u << k;
  }
  *base = d.ptr;
  *tab = (const double*)(d.ptr + table_offset(k));
  return DISORT_OK;
}

}  // namespace

extern "C" {

size_t disort_static_scratch_bytes(void) { return kScratchBytes - kTableBytes; }

size_t disort_flux_static_bytes(int64_t n, int32_t nstr, int32_t nlyr) {
  if (n <= 0 || nlyr <= 0 || nstr <= 0) return 0;
  return WorkLayout(n, nstr, nlyr).end - kTableBytes;
}

int disort_flux_execute_static(disort_ctx* ctx, const disort_flux_batch* b, int32_t nstr, int32_t nlyr,
                               int32_t pstride) {
  if (!ctx || !b) return fail(ctx, DISORT_E_INVALID, "null argument");
  if (b->n <= 0) return fail(ctx, DISORT_E_INVALID, "batch is empty");
  if (nlyr <= 0) return fail(ctx, DISORT_E_INVALID, "num_layers must be > 0");
  if (nstr < 4 || nstr % 2) return fail(ctx, DISORT_E_INVALID, "num_streams must be even and >= 4");
  if (count_index(nstr) < 0)
    return fail(ctx, DISORT_E_UNSUPPORTED,
                "num_streams=" + std::to_string(nstr) + " is not supported by the GPU flux path (4, 6, ..., 16, 32, 64)");
  if (pstride < nstr + 1 + ((b->flags & DISORT_F_DELTA_M_PLUS) ? 1 : 0))
    return fail(ctx, DISORT_E_INVALID, "phase moment rows too short");
  if ((b->flags & DISORT_F_THERMAL) && !b->temper) return fail(ctx, DISORT_E_INVALID, "thermal batch without temperatures");
  if ((b->flags & DISORT_F_SPHERICAL) && !b->alt)
    return fail(ctx, DISORT_E_INVALID, "spherical batch without level altitudes");
  if (!b->dtau || !b->ssa || !b->pmom || !b->scal || !b->out) return fail(ctx, DISORT_E_INVALID, "null batch array");
  const int64_t n = b->n;
  if (n > (int64_t)1 << 30) return fail(ctx, DISORT_E_INVALID, "batch too large for the static scratch");
  const WorkLayout w(n, nstr, nlyr);
  if (w.end > kScratchBytes) {
    const size_t per = WorkLayout(1, nstr, nlyr).end - kTableBytes;
    return fail(ctx, DISORT_E_INVALID,
                "batch needs " + std::to_string(w.end - kTableBytes) + " bytes of static scratch; the maximum is " +
                    std::to_string(kScratchBytes - kTableBytes) + " (about " +
                    std::to_string((kScratchBytes - kTableBytes) / per) + " instances of num_streams=" +
                    std::to_string(nstr) + ", num_layers=" + std::to_string(nlyr) +
                    "); split the batch or use disort_flux_execute");
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  unsigned char* base = nullptr;
  const double* tab = nullptr;
  int rc;
  if ((rc = device_scratch(ctx, nstr, &base, &tab))) return rc;

  dsa::FluxArgs a{};
  a.flags = b->flags;
  a.nstr = (const int32_t*)(base + w.nstr);
  a.nlyr = (const int32_t*)(base + w.nlyr);
  a.pstride = (const int32_t*)(base + w.pstride);
  a.lay_off = (const int64_t*)(base + w.lay_off);
  a.mom_off = (const int64_t*)(base + w.mom_off);
  a.dtau = b->dtau;
  a.ssa = b->ssa;
  a.pmom = b->pmom;
  a.temper = b->temper;
  a.alt = b->alt;
  a.scal = b->scal;
  a.out = b->out;
  a.total_layers = n * nlyr;
  a.total_levels = n * (nlyr + 1);
  a.inst_d = (double*)(base + w.inst_d);
  a.inst_i = (int32_t*)(base + w.inst_i);
  a.lay_d = (double*)(base + w.lay_d);
  a.lev_d = (double*)(base + w.lev_d);
  a.nb = (int32_t)n;
  a.b_inst = (const int32_t*)(base + w.b_inst);
  a.b_task0 = (const int64_t*)(base + w.b_task0);
  a.n_tasks = n * nlyr;
  a.t_inst = (const int32_t*)(base + w.t_inst);
  a.t_lc = (const int32_t*)(base + w.t_lc);
  a.n_lvtasks = n * (nlyr + 1);
  a.v_bk = (const int32_t*)(base + w.v_bk);
  a.v_lu = (const int32_t*)(base + w.v_lu);
  a.eig = (double*)(base + w.eig);
  a.panel = (double*)(base + w.panel);
  a.cvec = (double*)(base + w.cvec);
  a.tab = tab;

  const int64_t V = n * (nlyr + 1);
  hipLaunchKernelGGL(k_static_index, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, ctx->stream, n, (int)nstr,
                     (int)nlyr, (int)pstride, (int32_t*)a.nstr, (int32_t*)a.nlyr, (int32_t*)a.pstride,
                     (int64_t*)a.lay_off, (int64_t*)a.mom_off, (int32_t*)a.b_inst, (int64_t*)a.b_task0,
                     (int32_t*)a.t_inst, (int32_t*)a.t_lc, (int32_t*)a.v_bk, (int32_t*)a.v_lu);
  HIPCHK(ctx, hipGetLastError());
  // Kernel variant as a plan's default: register kernels up to 16 streams,
  // workgroup kernels at 32 and 64 streams and for the spherical beam.
  const bool lane = nstr <= 16 && !(b->flags & DISORT_F_SPHERICAL);
  // The plan's default kernels for this stream count (two lanes per layer
  // in the eigen kernel at 6 to 12 streams), so that results equal
  // disort_flux_execute's bit for bit.
  const bool pair = nstr >= 6 && nstr <= 12;
  const bool ok = lane ? dsa::launch_flux_bucket(nstr, a, ctx->stream, nullptr, pair)
                       : dsa::launch_flux_bucket_wg(nstr, a, ctx->stream, nullptr);
  if (!ok) return fail(ctx, DISORT_E_UNSUPPORTED, "no kernels for num_streams=" + std::to_string(nstr));
  HIPCHK(ctx, hipGetLastError());
  if (b->status) {
    hipLaunchKernelGGL(k_static_status, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, n,
                       (const int32_t*)a.inst_i, b->status);
    HIPCHK(ctx, hipGetLastError());
  }
  return DISORT_OK;
}

}  // extern "C"
