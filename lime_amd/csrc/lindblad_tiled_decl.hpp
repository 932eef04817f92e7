// Host-visible interface of the tiled (per-stage launch) Lindblad RK4 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace lime {

struct TiledArgs {
  int batch, n, group, n_dense, n_eops, eop_group, flags;
  int nc, border_wgs;      // tile-covered core size and border workgroups per operator
  int stage, rec, rec_capacity;
  double dt;
  // operator sets and workspace are complex64 (FP32 variant) or complex128 (FP64 variant)
  const void* heff;
  const void* dmat;
  const void* lops;
  const void* rops;        // right factors of the sandwich terms (NULL: R = L)
  // 3xBF16: H_eff, L_k and R_k pre-split into BF16 planes (Re hi, Re lo, Im hi, Im lo) of n
  // rows x ldp (n rounded up to 8, zero beyond n), one set of planes per matrix
  const unsigned short* hsp;
  const unsigned short* lsp;
  const unsigned short* rsp;  // NULL: R = L
  int ldp;
  const double2* eops_t;
  double2* rho;
  double2* obs;
  double2* states;
  // workspace
  const void* s_cur;
  void* s_next;
  void* inc;
  void* tbuf;
  double* tscale;          // per-operator 1 / Re Tr rho of the trace renormalisation
  double2* obs_part;       // [batch][slots][n_eops] per-workgroup partial sums of Tr[E rho]
  int slots;               // workgroups per operator of the stage kernel (= obs_part slots)
  // K-split stage launches (ks > 1): part p reduces k in [p klen, (p + 1) klen) into
  // part[p][batch][n][n] (operand precision); lindblad_stage_sum finishes the stage
  int ks, klen;
  void* part;
  double2* stage_rec;      // [rec_capacity][3][batch][n][n] FP64 stage operands, or NULL
  // generator application (lindblad_apply_tiled): V = L(s_cur);
  //   y = c_lx V + c_z z (complex128), y64 = copy of y in the operand precision,
  //   acc += c_acc V
  const double2* z;
  double2* y;
  void* y64;
  double2* acc;
  double2 c_lx, c_z, c_acc;
};

// GEMM arithmetic of the tiled kernels: complex64 on FP32 MFMA, complex64 as 3-term BF16
// split products on BF16 MFMA, complex128 on FP64 MFMA.
enum class TiledPrec { FP32 = 0, SPLIT_BF16 = 1, FP64 = 2 };

// Workspace of lime_lindblad_rk4's per-stage path: S0, S1, the increment, T_k (operand
// precision; 3xBF16: BF16 planes), the trace scales (renormalisation), the per-workgroup
// observable partials (n_eops > 0), for 3xBF16 the pre-split planes of H_eff, L_k and R_k
// (one set per operator set, ceil(batch / group) sets), and with a K-split (ks > 1 for this
// tile edge) the ks partial sums of the stage operators, each region 256-byte aligned.
size_t tiled_workspace_bytes(int batch, int n, int group, int n_dense, int flags, bool fp64, int n_eops,
                             bool rops, int tile, int ks);

// lime_lindblad_batch.flags bit of one-sided generators (LIME_F_ONE_SIDED)
constexpr int kFlagOneSided = 0x200;
int tiled_pick(int batch, int n);
hipError_t tiled_run(TiledArgs a, int nsteps, int step_offset, int nout, void* work, hipStream_t stream,
                     int tile, TiledPrec prec);
hipError_t tiled_apply(TiledArgs a, hipStream_t stream, int tile, bool fp64);

constexpr int kMaxContractPairs = 4;
struct ContractArgs {
  int batch, n, pairs;    // pairs (<= kMaxContractPairs) of operators summed into one result
  const double2* kappa[kMaxContractPairs];   // adjoint stage operators  [batch][n][n]
  const double2* s[kMaxContractPairs];       // forward stage operators  [batch][n][n]
  double weight;
  double2* out_a;         // [batch][n]   += w (rowsum - colsum) of conj(kappa) o s
  double2* out_b;         // [batch][n]   += w (diag - (rowsum + colsum)/2)
  double2* out_u;         // [batch][n-1] += w (nearest-neighbour coupling contraction)
};
size_t contract_lds_bytes(int n);

// Start of one backward RK4 step of the adjoint (elementwise over batch n^2):
//   lam += g[b] m;  k4 = c lam;  k4_op = k4 in the operand precision;  acc = lam
struct SeedArgs {
  int batch, n, op_fp64;
  double2* lam;           // [batch][n][n]
  const double2* g;       // [batch] source weights
  const double2* m;       // [n][n] source operator (shared)
  double c;
  double2* k4;            // [batch][n][n]
  void* k4_op;            // [batch][n][n] complex64 (complex128 when op_fp64)
  double2* acc;           // [batch][n][n]
};
hipError_t seed_launch(const SeedArgs& s, hipStream_t stream);

// Uneven batch in one persistent launch (complex64 operands on FP32 MFMA, FP64 state).
struct RaggedInstance {
  int n, n_dense;
  long long heff, dmat, lops;   // element offsets into ops (complex64); dmat < 0: none
  long long rho;                // element offset into rho (complex128)
  long long work;               // element offset into work (complex64): (3 + n_dense) n^2
};
struct RaggedArgs {
  const RaggedInstance* inst;   // [count]
  const int* order;             // [count] processing order (largest cost first)
  int count, nsteps;
  double dt;
  const float2* ops;
  double2* rho;
  float2* work;
  unsigned* queue;              // [2]: next item, finished workgroups (zero between launches)
};
int ragged_variants();
hipError_t ragged_run(const RaggedArgs& r, int variant, int max_groups, hipStream_t stream);
hipError_t contract_launch(const ContractArgs& c, hipStream_t stream);

}  // namespace lime
