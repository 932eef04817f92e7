// Host-visible interface of the resident Lindblad RK4 kernel (lindblad_resident.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace lime {

struct ResidentArgs {
  int batch, n, group, n_dense, n_eops, eop_group, flags;
  const float2* heff;
  const float2* dmat;
  const float2* lops;
  const float2* rops;
  const double2* eops_t;
  double2* rho;
  double2* obs;
  double2* states;
  double dt;
  int nsteps, nout, step_offset, rec_capacity;
};

int resident_nb_max(int n_dense);
// n = 16 m + 1 without dense jump operators runs the bordered kernels
bool resident_bordered(int n, int n_dense);
int resident_variants(int n, int n_dense);
size_t resident_lds_bytes(int n, int variant, int n_dense, int n_eops);
// does this launch use the double-buffered stage operand (no dense operators, n <= 80,
// and the layout fits the LDS)?
bool resident_double_buffered(int n, int variant, int n_dense, int n_eops);
hipError_t resident_launch(const ResidentArgs& a, int variant, hipStream_t stream);

// 3xBF16 split-product resident kernel (lindblad_resident_bf3.hip): bordered n = 33, 49, 65
// without dense operators; one launch shape
bool resident_bf3_eligible(int n, int n_dense);
size_t resident_bf3_lds_bytes(int n, int n_eops);
hipError_t resident_bf3_launch(const ResidentArgs& a, hipStream_t stream);

}  // namespace lime
