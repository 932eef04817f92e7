// extern "C" entry points of liblime_hip.so (declared in include/lime_hip.h).
#include <cstddef>
#include <cstdio>
#include <mutex>
#include <string>

#include "common.hpp"
#include "lime_hip.h"
#include "lindblad_resident_decl.hpp"
#include "lindblad_tiled_decl.hpp"

static_assert(LIME_F_ONE_SIDED == lime::kFlagOneSided, "one flag value");
static_assert(sizeof(lime_ragged_instance) == sizeof(lime::RaggedInstance) &&
                  offsetof(lime_ragged_instance, work) == offsetof(lime::RaggedInstance, work),
              "the ABI instance record is the kernel's");

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  return fail(LIME_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

constexpr size_t kLdsLimit = 160 * 1024;

// Does this descriptor run on the resident (single-workgroup-per-operator) kernel?
bool is_fp64(const lime_lindblad_batch* d) { return (d->flags & LIME_F_PREC_MASK) == LIME_F_PREC_FP64; }

// Resident launch-shape variant requested by d->variant (0 = default).
int resident_variant(const lime_lindblad_batch* d) {
  return (d->variant > 0 && d->variant < lime::resident_variants(d->n, d->n_dense)) ? d->variant : 0;
}

bool is_split_bf16(const lime_lindblad_batch* d) {
  return (d->flags & LIME_F_PREC_MASK) == LIME_F_PREC_3XBF16;
}

// 3xBF16 on the resident kernel: the bordered sizes n = 33, 49, 65 without dense operators
bool use_resident_bf3(const lime_lindblad_batch* d) {
  if (d->flags & (LIME_F_FORCE_TILED | LIME_F_ONE_SIDED)) return false;
  if (!is_split_bf16(d) || d->rops != nullptr || !lime::resident_bf3_eligible(d->n, d->n_dense)) return false;
  return lime::resident_bf3_lds_bytes(d->n, d->obs ? d->n_eops : 0) <= kLdsLimit;
}

bool use_resident(const lime_lindblad_batch* d) {
  if (d->flags & (LIME_F_FORCE_TILED | LIME_F_ONE_SIDED)) return false;
  if (use_resident_bf3(d)) return true;
  // FP64 (and 3xBF16 at other sizes) run on the tiled kernels
  if ((d->flags & LIME_F_PREC_MASK) != LIME_F_PREC_FP32) return false;
  const int nb = lime::cdiv(d->n, 16);
  if (nb > lime::resident_nb_max(d->n_dense)) return false;
  return lime::resident_lds_bytes(d->n, resident_variant(d), d->n_dense, d->obs ? d->n_eops : 0) <= kLdsLimit;
}

// Tiled launch shape requested by d->variant: 1 -> 32 x 32 tiles, 2 -> 64 x 64 tiles, 3 / 4
// -> 32 x 32 tiles with the stage reduction split in 2 / 8 parts (more workgroups for small
// batches; dense512 at batch 1: 3xBF16 fastest with 2 parts, FP32 with 8); 0 -> the
// heuristic tile, no split.
int tiled_tile(const lime_lindblad_batch* d) {
  if (d->variant == 1 || d->variant == 3 || d->variant == 4) return 32;
  if (d->variant == 2) return 64;
  return lime::tiled_pick(d->batch, d->n);
}
int tiled_ksplit(const lime_lindblad_batch* d) { return d->variant == 3 ? 2 : d->variant == 4 ? 8 : 1; }

// ---- static scratch of lime_lindblad_rk4_noalloc: one array per GPU, held until the
// process ends.  Host code never uses the array's name as an address (on the host it is a
// stand-in); the device address comes from hipGetSymbolAddress, once per device.
constexpr size_t kScratchBytes = (size_t)LIME_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array (looked up on the first call on
// that device, with that device current).
int scratch_for_current_device(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "lime_lindblad_rk4_noalloc: hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return fail(LIME_ERR_UNSUPPORTED, "lime_lindblad_rk4_noalloc: device index " + std::to_string(dev) + " >= 64");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "lime_lindblad_rk4_noalloc: hipGetSymbolAddress(scratch)");
    if (ptr == nullptr) return fail(LIME_ERR_HIP, "lime_lindblad_rk4_noalloc: scratch has no device address");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return LIME_OK;
}

// lime_lindblad_rk4 with the per-stage workspace given explicitly (the caller's, or the
// library's static scratch array for lime_lindblad_rk4_noalloc); `name` prefixes errors.
int rk4_impl(const lime_lindblad_batch* d, const lime_step_params* p, void* stream, void* work,
             int64_t work_bytes, const char* name, bool noalloc) {
  if (d == nullptr || p == nullptr) return fail(LIME_ERR_ARG, std::string(name) + ": null descriptor");
  if (d->batch < 0 || d->n < 1) return fail(LIME_ERR_ARG, std::string(name) + ": batch must be >= 0 and n >= 1");
  if (d->batch == 0 || p->nsteps == 0) return LIME_OK;
  if (p->nsteps < 0) return fail(LIME_ERR_ARG, std::string(name) + ": nsteps < 0");
  if (d->group < 1 || d->eop_group < 1) return fail(LIME_ERR_ARG, std::string(name) + ": group and eop_group must be >= 1");
  if (d->heff == nullptr || d->rho == nullptr) return fail(LIME_ERR_ARG, std::string(name) + ": heff and rho are required");
  if (d->n_dense < 0 || (d->n_dense > 0 && d->lops == nullptr))
    return fail(LIME_ERR_ARG, std::string(name) + ": n_dense > 0 needs lops");
  if (d->n_eops < 0) return fail(LIME_ERR_ARG, std::string(name) + ": n_eops < 0");
  if (d->obs != nullptr && (d->n_eops == 0 || d->eops_t == nullptr))
    return fail(LIME_ERR_ARG, std::string(name) + ": obs needs n_eops > 0 and eops_t");
  if ((d->obs != nullptr || d->states != nullptr || d->stage_states != nullptr) && (p->nout < 1 || p->rec_capacity < 0))
    return fail(LIME_ERR_ARG, std::string(name) + ": recording needs nout >= 1 and rec_capacity >= 0");
  if (p->step_offset < 0) return fail(LIME_ERR_ARG, std::string(name) + ": step_offset < 0");
  if ((d->flags & LIME_F_PREC_MASK) == 0x30)
    return fail(LIME_ERR_UNSUPPORTED, std::string(name) + ": precision flag not built");
  const hipStream_t s = static_cast<hipStream_t>(stream);

  if (use_resident(d)) {
    if (d->stage_states != nullptr)
      return fail(LIME_ERR_UNSUPPORTED, std::string(name) + ": stage operands are recorded by the per-stage "
                                        "kernels only (set LIME_F_FORCE_TILED)");
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
    a.rops = static_cast<const float2*>(d->rops);
    a.eops_t = static_cast<const double2*>(d->eops_t);
    a.rho = static_cast<double2*>(d->rho);
    a.obs = static_cast<double2*>(d->obs);
    a.states = static_cast<double2*>(d->states);
    a.dt = p->dt;
    a.nsteps = p->nsteps;
    a.nout = (d->obs || d->states) ? p->nout : 0;
    a.step_offset = p->step_offset;
    a.rec_capacity = p->rec_capacity;
    hipError_t e = use_resident_bf3(d) ? lime::resident_bf3_launch(a, s) : lime::resident_launch(a, resident_variant(d), s);
    if (e != hipSuccess) return hip_fail(e, (std::string(name) + " (resident) launch").c_str());
    return LIME_OK;
  }

  // ---- tiled path: one launch per RK4 stage on the workspace
  const size_t need = lime::tiled_workspace_bytes(d->batch, d->n, d->group, d->n_dense, d->flags, is_fp64(d),
                                                 d->obs ? d->n_eops : 0, d->rops != nullptr, tiled_tile(d),
                                                 tiled_ksplit(d));
  if (noalloc && need > kScratchBytes)
    return fail(LIME_ERR_ARG, std::string(name) + ": this batch needs " + std::to_string(need) +
                                  " bytes of scratch, above the static " + std::to_string(kScratchBytes) +
                                  " (LIME_NOALLOC_SCRATCH_BYTES); use lime_lindblad_rk4 with a workspace");
  if (noalloc) {
    const int r = scratch_for_current_device(&work);
    if (r != LIME_OK) return r;
    work_bytes = (int64_t)kScratchBytes;
  }
  if (work == nullptr || work_bytes < (int64_t)need)
    return fail(LIME_ERR_ARG, std::string(name) + ": n = " + std::to_string(d->n) + " needs a workspace of " +
                                  std::to_string(need) + " bytes (lime_lindblad_workspace_bytes)");
  lime::TiledArgs a = {};
  a.batch = d->batch;
  a.n = d->n;
  a.group = d->group;
  a.n_dense = d->n_dense;
  a.n_eops = d->obs ? d->n_eops : 0;
  a.eop_group = d->eop_group;
  a.flags = d->flags;
  a.rec_capacity = p->rec_capacity;
  a.dt = p->dt;
  a.heff = d->heff;
  a.dmat = d->dmat;
  a.lops = d->lops;
  a.rops = d->rops;
  a.eops_t = static_cast<const double2*>(d->eops_t);
  a.rho = static_cast<double2*>(d->rho);
  a.obs = static_cast<double2*>(d->obs);
  a.states = static_cast<double2*>(d->states);
  a.stage_rec = static_cast<double2*>(d->stage_states);
  a.ks = tiled_ksplit(d);
  const int nout = (d->obs || d->states || d->stage_states) ? p->nout : 0;
  const lime::TiledPrec prec = is_fp64(d) ? lime::TiledPrec::FP64
                              : is_split_bf16(d) ? lime::TiledPrec::SPLIT_BF16 : lime::TiledPrec::FP32;
  hipError_t e = lime::tiled_run(a, p->nsteps, p->step_offset, nout, work, s, tiled_tile(d), prec);
  if (e != hipSuccess) return hip_fail(e, (std::string(name) + " (tiled) launch").c_str());
  return LIME_OK;
}

}  // namespace

extern "C" {

const char* lime_last_error(void) { return g_last_error.c_str(); }

const char* lime_version(void) { return "lime_hip 0.2 (gfx950)"; }

int lime_lindblad_max_n(int32_t n_dense, int32_t flags) {
  if ((flags & LIME_F_PREC_MASK) != LIME_F_PREC_FP32) return 0;
  return 16 * lime::resident_nb_max(n_dense);
}

int lime_lindblad_variants(const lime_lindblad_batch* d) {
  if (d == nullptr || d->n < 1) return 0;
  lime_lindblad_batch probe = *d;
  probe.variant = 0;
  if (use_resident_bf3(&probe)) return 1;
  if (use_resident(&probe)) return lime::resident_variants(d->n, d->n_dense);
  return 5;   // tiled: 0 = heuristic, 1 = 32x32, 2 = 64x64 tiles, 3 / 4 = 32x32 K-split in 2 / 8
}

int64_t lime_lindblad_workspace_bytes(const lime_lindblad_batch* d) {
  if (d == nullptr || d->n < 1 || d->batch < 1) return 0;
  if (use_resident(d)) return 0;
  return (int64_t)lime::tiled_workspace_bytes(d->batch, d->n, d->group, d->n_dense, d->flags, is_fp64(d),
                                             d->obs ? d->n_eops : 0, d->rops != nullptr, tiled_tile(d),
                                             tiled_ksplit(d));
}

int lime_lindblad_rk4(const lime_lindblad_batch* d, const lime_step_params* p, void* stream) {
  return rk4_impl(d, p, stream, d ? d->work : nullptr, d ? d->work_bytes : 0, "lime_lindblad_rk4", false);
}

int lime_lindblad_rk4_noalloc(const lime_lindblad_batch* d, const lime_step_params* p, void* stream) {
  return rk4_impl(d, p, stream, nullptr, 0, "lime_lindblad_rk4_noalloc", true);
}

int lime_lindblad_apply(const lime_lindblad_batch* d, const void* x64, const void* z, void* y,
                        void* y64, void* acc, const double* c_lx, const double* c_z,
                        const double* c_acc, void* stream) {
  if (d == nullptr || x64 == nullptr) return fail(LIME_ERR_ARG, "lime_lindblad_apply: null descriptor or x");
  if (d->batch < 0 || d->n < 1 || d->group < 1) return fail(LIME_ERR_ARG, "lime_lindblad_apply: bad batch/n/group");
  if (d->batch == 0) return LIME_OK;
  if (d->heff == nullptr) return fail(LIME_ERR_ARG, "lime_lindblad_apply: heff is required");
  if (d->n_dense < 0 || (d->n_dense > 0 && d->lops == nullptr)) return fail(LIME_ERR_ARG, "lime_lindblad_apply: n_dense > 0 needs lops");
  if ((y || y64) && c_lx == nullptr) return fail(LIME_ERR_ARG, "lime_lindblad_apply: y needs c_lx");
  if (z && c_z == nullptr) return fail(LIME_ERR_ARG, "lime_lindblad_apply: z needs c_z");
  if (acc && c_acc == nullptr) return fail(LIME_ERR_ARG, "lime_lindblad_apply: acc needs c_acc");
  if ((d->flags & LIME_F_PREC_MASK) == LIME_F_PREC_3XBF16 || (d->flags & LIME_F_PREC_MASK) == 0x30)
    return fail(LIME_ERR_UNSUPPORTED, "lime_lindblad_apply: built for the FP32 and FP64 variants only");
  const size_t need = (size_t)d->batch * d->n_dense * d->n * d->n * (is_fp64(d) ? sizeof(double2) : sizeof(float2));
  if (d->n_dense > 0 && (d->work == nullptr || d->work_bytes < (int64_t)need))
    return fail(LIME_ERR_ARG, "lime_lindblad_apply: dense operators need a workspace of " + std::to_string(need) + " bytes");
  lime::TiledArgs a = {};
  a.batch = d->batch;
  a.n = d->n;
  a.group = d->group;
  a.n_dense = d->n_dense;
  a.flags = d->flags;
  a.heff = d->heff;
  a.dmat = d->dmat;
  a.lops = d->lops;
  a.rops = d->rops;
  a.tbuf = d->work;
  a.s_cur = x64;
  a.z = static_cast<const double2*>(z);
  a.y = static_cast<double2*>(y);
  a.y64 = y64;
  a.acc = static_cast<double2*>(acc);
  if (c_lx) a.c_lx = make_double2(c_lx[0], c_lx[1]);
  if (c_z) a.c_z = make_double2(c_z[0], c_z[1]);
  if (c_acc) a.c_acc = make_double2(c_acc[0], c_acc[1]);
  hipError_t e = lime::tiled_apply(a, static_cast<hipStream_t>(stream), tiled_tile(d), is_fp64(d));
  if (e != hipSuccess) return hip_fail(e, "lime_lindblad_apply launch");
  return LIME_OK;
}

int lime_ragged_variants(void) { return lime::ragged_variants(); }

int lime_lindblad_rk4_ragged(const lime_ragged_batch* r, const lime_step_params* p, void* stream) {
  if (r == nullptr || p == nullptr) return fail(LIME_ERR_ARG, "lime_lindblad_rk4_ragged: null descriptor");
  if (r->count < 0 || p->nsteps < 0) return fail(LIME_ERR_ARG, "lime_lindblad_rk4_ragged: count and nsteps must be >= 0");
  if (r->count == 0 || p->nsteps == 0) return LIME_OK;
  if (!r->inst || !r->order || !r->ops || !r->rho || !r->work || !r->queue)
    return fail(LIME_ERR_ARG, "lime_lindblad_rk4_ragged: inst, order, ops, rho, work and queue are required");
  if (r->variant < 0 || r->variant >= lime::ragged_variants())
    return fail(LIME_ERR_ARG, "lime_lindblad_rk4_ragged: variant out of range");
  if (r->max_groups < 0) return fail(LIME_ERR_ARG, "lime_lindblad_rk4_ragged: max_groups < 0");
  lime::RaggedArgs a;
  a.inst = reinterpret_cast<const lime::RaggedInstance*>(r->inst);
  a.order = r->order;
  a.count = r->count;
  a.nsteps = p->nsteps;
  a.dt = p->dt;
  a.ops = static_cast<const float2*>(r->ops);
  a.rho = static_cast<double2*>(r->rho);
  a.work = static_cast<float2*>(r->work);
  a.queue = static_cast<unsigned*>(r->queue);
  hipError_t e = lime::ragged_run(a, r->variant, r->max_groups, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "lime_lindblad_rk4_ragged launch");
  return LIME_OK;
}

int lime_adjoint_contract_pairs(int32_t batch, int32_t n, int32_t pairs, const void* const* kappa,
                                const void* const* s, double w, void* out_a, void* out_b, void* out_u,
                                void* stream) {
  if (batch < 0 || n < 1) return fail(LIME_ERR_ARG, "lime_adjoint_contract: bad batch/n");
  if (pairs < 1 || pairs > lime::kMaxContractPairs)
    return fail(LIME_ERR_ARG, "lime_adjoint_contract_pairs: pairs must be in 1..4");
  if (lime::contract_lds_bytes(n) == 0)
    return fail(LIME_ERR_UNSUPPORTED, "lime_adjoint_contract: n = " + std::to_string(n) +
                                          " exceeds the contraction's LDS (n <= 1996)");
  if (batch == 0) return LIME_OK;
  if (!kappa || !s || !out_a || !out_b || (n > 1 && !out_u))
    return fail(LIME_ERR_ARG, "lime_adjoint_contract: null pointer");
  lime::ContractArgs c = {};
  c.batch = batch;
  c.n = n;
  c.pairs = pairs;
  for (int p = 0; p < pairs; ++p) {
    if (!kappa[p] || !s[p]) return fail(LIME_ERR_ARG, "lime_adjoint_contract: null pointer");
    c.kappa[p] = static_cast<const double2*>(kappa[p]);
    c.s[p] = static_cast<const double2*>(s[p]);
  }
  c.weight = w;
  c.out_a = static_cast<double2*>(out_a);
  c.out_b = static_cast<double2*>(out_b);
  c.out_u = static_cast<double2*>(out_u);
  hipError_t e = lime::contract_launch(c, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "lime_adjoint_contract launch");
  return LIME_OK;
}

int lime_adjoint_contract(int32_t batch, int32_t n, const void* kappa, const void* s, double w,
                          void* out_a, void* out_b, void* out_u, void* stream) {
  return lime_adjoint_contract_pairs(batch, n, 1, &kappa, &s, w, out_a, out_b, out_u, stream);
}

int lime_adjoint_seed(int32_t batch, int32_t n, int32_t flags, void* lam, const void* g, const void* m,
                      double c, void* k4, void* k4_op, void* acc, void* stream) {
  if (batch < 0 || n < 1) return fail(LIME_ERR_ARG, "lime_adjoint_seed: bad batch/n");
  if (batch == 0) return LIME_OK;
  if (!lam || !g || !m || !k4 || !k4_op || !acc) return fail(LIME_ERR_ARG, "lime_adjoint_seed: null pointer");
  lime::SeedArgs a;
  a.batch = batch;
  a.n = n;
  a.op_fp64 = (flags & LIME_F_PREC_MASK) == LIME_F_PREC_FP64;
  a.lam = static_cast<double2*>(lam);
  a.g = static_cast<const double2*>(g);
  a.m = static_cast<const double2*>(m);
  a.c = c;
  a.k4 = static_cast<double2*>(k4);
  a.k4_op = k4_op;
  a.acc = static_cast<double2*>(acc);
  hipError_t e = lime::seed_launch(a, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "lime_adjoint_seed launch");
  return LIME_OK;
}

}  // extern "C"
