// extern "C" entry points of liblime_hip.so (declared in include/lime_hip.h).
#include <cstdio>
#include <string>

#include "common.hpp"
#include "lime_hip.h"

#include "lindblad_resident_decl.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  return fail(LIME_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

}  // namespace

extern "C" {

const char* lime_last_error(void) { return g_last_error.c_str(); }

const char* lime_version(void) { return "lime_hip 0.1 (gfx950)"; }

int lime_lindblad_max_n(int32_t n_dense, int32_t flags) {
  if ((flags & LIME_F_PREC_MASK) != LIME_F_PREC_FP32) return 0;
  return 16 * lime::resident_nb_max(n_dense);
}

int lime_lindblad_rk4(const lime_lindblad_batch* d, const lime_step_params* p, void* stream) {
  if (d == nullptr || p == nullptr) return fail(LIME_ERR_ARG, "lime_lindblad_rk4: null descriptor");
  if (d->batch < 0 || d->n < 1) return fail(LIME_ERR_ARG, "lime_lindblad_rk4: batch must be >= 0 and n >= 1");
  if (d->batch == 0 || p->nsteps == 0) return LIME_OK;
  if (p->nsteps < 0) return fail(LIME_ERR_ARG, "lime_lindblad_rk4: nsteps < 0");
  if (d->group < 1 || d->eop_group < 1) return fail(LIME_ERR_ARG, "lime_lindblad_rk4: group and eop_group must be >= 1");
  if (d->heff == nullptr || d->rho == nullptr) return fail(LIME_ERR_ARG, "lime_lindblad_rk4: heff and rho are required");
  if (d->n_dense < 0 || (d->n_dense > 0 && d->lops == nullptr))
    return fail(LIME_ERR_ARG, "lime_lindblad_rk4: n_dense > 0 needs lops");
  if (d->n_eops < 0) return fail(LIME_ERR_ARG, "lime_lindblad_rk4: n_eops < 0");
  if (d->obs != nullptr && (d->n_eops == 0 || d->eops_t == nullptr))
    return fail(LIME_ERR_ARG, "lime_lindblad_rk4: obs needs n_eops > 0 and eops_t");
  if ((d->obs != nullptr || d->states != nullptr) && (p->nout < 1 || p->rec_capacity < 0))
    return fail(LIME_ERR_ARG, "lime_lindblad_rk4: recording needs nout >= 1 and rec_capacity >= 0");
  if (p->step_offset < 0) return fail(LIME_ERR_ARG, "lime_lindblad_rk4: step_offset < 0");
  if ((d->flags & LIME_F_PREC_MASK) != LIME_F_PREC_FP32)
    return fail(LIME_ERR_UNSUPPORTED, "lime_lindblad_rk4: only the FP32-MFMA precision is built");

  const int nb = lime::cdiv(d->n, 16);
  if (nb > lime::resident_nb_max(d->n_dense))
    return fail(LIME_ERR_UNSUPPORTED, "lime_lindblad_rk4: n = " + std::to_string(d->n) +
                                          " exceeds the resident kernel's limit " +
                                          std::to_string(16 * lime::resident_nb_max(d->n_dense)));
  const size_t lds = lime::resident_lds_bytes(nb, d->n_dense, d->obs ? d->n_eops : 0);
  if (lds > 160 * 1024)
    return fail(LIME_ERR_UNSUPPORTED, "lime_lindblad_rk4: n = " + std::to_string(d->n) + " with " +
                                          std::to_string(d->n_dense) + " dense operators and " +
                                          std::to_string(d->n_eops) + " observables needs " +
                                          std::to_string(lds) + " B of LDS (> 160 KiB)");
  lime::ResidentArgs a;
  a.batch = d->batch;
  a.n = d->n;
  a.group = d->group;
  a.n_dense = d->n_dense;
  a.n_eops = d->obs ? d->n_eops : 0;
  a.eop_group = d->eop_group;
  a.flags = d->flags;
  a.heff = static_cast<const float2*>(d->heff);
  a.dmat = static_cast<const float2*>(d->dmat);
  a.lops = static_cast<const float2*>(d->lops);
  a.eops_t = static_cast<const double2*>(d->eops_t);
  a.rho = static_cast<double2*>(d->rho);
  a.obs = static_cast<double2*>(d->obs);
  a.states = static_cast<double2*>(d->states);
  a.dt = p->dt;
  a.nsteps = p->nsteps;
  a.nout = (d->obs || d->states) ? p->nout : 0;
  a.step_offset = p->step_offset;
  a.rec_capacity = p->rec_capacity;
  hipError_t e = lime::resident_launch(a, nb, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "lime_lindblad_rk4 launch");
  return LIME_OK;
}

}  // extern "C"
