// Kernel argument block for the flux kernels (device pointers only).
#pragma once
#include <cstdint>

namespace dsa {

struct FluxArgs {
  // Batch (device copies of disort_flux_batch arrays).
  int32_t flags;
  int32_t pad0;
  const int32_t* nstr;
  const int32_t* nlyr;
  const int32_t* pstride;
  const int64_t* lay_off;
  const int64_t* mom_off;
  const double* dtau;
  const double* ssa;
  const double* pmom;
  const double* temper;
  const double* alt;  // level altitudes (spherical beam)
  const double* scal;
  double* out;
  int64_t total_levels;
  int64_t total_layers;
  // Scratch shared by all buckets.
  double* inst_d;   // [n][kInstD]
  int32_t* inst_i;  // [n][kInstI]
  double* lay_d;    // [kLayD][total_layers]
  double* lev_d;    // [kLevD][total_levels]
  // Bucket (instances with the same stream count).
  int32_t nb;                // instances in the bucket
  int32_t pad1;
  const int32_t* b_inst;     // [nb] batch instance index
  const int64_t* b_task0;    // [nb] first layer task of the instance
  int64_t n_tasks;           // layer tasks (sum of nlyr over the bucket)
  const int32_t* t_inst;     // [n_tasks] batch instance index
  const int32_t* t_lc;       // [n_tasks] layer index
  int64_t n_lvtasks;         // level tasks (sum of nlyr+1)
  const int32_t* v_bk;       // [n_lvtasks] bucket-local instance index
  const int32_t* v_lu;       // [n_lvtasks] level index
  double* eig;               // [n_tasks][EigOff::stride]
  double* panel;             // [n_tasks][panel_stride]
  double* cvec;              // [n_tasks][N]
  const double* tab;         // TabOff layout for this N
  // Planck values given instead of computed from the temperatures (the
  // temperature-Jacobian solves): [total_levels] at the levels, then [n][2]
  // the top emission (emissivity_top B(T_top)) and B(T_bottom); null = off.
  const double* planck_ovr;
  // Temperature Jacobians on the factored boundary system (null = off).
  // When jbase is set the layer kernels also store the layer's thermal
  // particular-solution basis ([instance][layer][2 NN]: (1-omega') base,
  // then d0 for a unit Planck slope), and the bucket launch solves the
  // (L+2) unit-source right-hand sides instead of the forward problem.
  double* jbase;
  double* jpanel;   // [instance][dof group][layer][2 N N + K N]
  double* jcvec;    // [instance][dof][layer][N] mode amplitudes
  double* jdbdt;    // [instance][dof] dB/dT
  double* jac;      // [column][6][L+1][L+2]
  int32_t* jstatus; // [column]
  int64_t jac_c0;   // column of batch instance 0
  int32_t jac_nlyr; // layers per column (uniform)
  int32_t pad2;
};

}  // namespace dsa
