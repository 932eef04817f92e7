// Host-visible interface of the tiled (per-stage launch) Lindblad RK4 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace lime {

struct TiledArgs {
  int batch, n, group, n_dense, n_eops, eop_group, flags;
  int stage, rec, rec_capacity;
  double dt;
  const float2* heff;
  const float2* dmat;
  const float2* lops;
  const float2* rops;      // right factors of the sandwich terms (NULL: R = L)
  const double2* eops_t;
  double2* rho;
  double2* obs;
  double2* states;
  // workspace
  const float2* s_cur;
  float2* s_next;
  float2* inc;
  float2* tbuf;
};

size_t tiled_workspace_bytes(int batch, int n, int n_dense);
int tiled_pick(int batch, int n);
hipError_t tiled_run(TiledArgs a, int nsteps, int step_offset, int nout, void* work, hipStream_t stream,
                     int tile);

}  // namespace lime
