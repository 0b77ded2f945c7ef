// Host-side context of the C ABI (one device, one stream, grow-only device
// buffers and constant tables) and the helpers the entry points share
// (capi.cpp: flux path, general_capi.cpp: general solver).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/disort_amd.h"

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct disort_flux_plan;

struct disort_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool timing = false;
  hipEvent_t ev[5] = {};
  double stage_ms[4] = {0, 0, 0, 0};
  std::map<std::string, DevBuf> bufs;
  std::map<int, double*> tables;  // per stream count
  std::map<int, double*> mode_tables;  // general solver: Y_l^m at the nodes, per stream count
  hipEvent_t gev[10] = {};
  double gstage_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int n_stages = 4;  // stages of the last timed solve (4 flux, 8 general)
};

#define HIPCHK(ctx, call)                                                              \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                  \
      return DISORT_E_HIP;                                                             \
    }                                                                                  \
  } while (0)


namespace dsa_host {
int fail(disort_ctx* ctx, int code, const std::string& msg);
int get_buf(disort_ctx* ctx, const char* name, size_t bytes, void** out);
bool gauss_nodes(int n, std::vector<double>& cmu, std::vector<double>& cwt);
int table_for(disort_ctx* ctx, int nstr, double** out);
// Constant table of stream count nstr (TabOff layout: nodes, weights,
// symmetrisation factors, Legendre values at the nodes).
bool build_table(int nstr, std::vector<double>& tab);
bool gpu_supported_nstr(int n);
}  // namespace dsa_host
