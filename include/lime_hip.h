/*
 * lime_hip.h -- C ABI of liblime_hip.so, the MI355X (gfx950) propagator library for
 * lime's open-quantum-system RK4 path.
 *
 * Conventions shared by every entry point
 *   - All pointers are DEVICE pointers owned by the caller (e.g. torch tensors); the
 *     library never frees them.  Matrices are dense, row-major, contiguous.
 *   - complex64 = interleaved (re, im) float pairs (torch.complex64 layout);
 *     complex128 = interleaved double pairs (torch.complex128 layout).
 *   - Every call is asynchronous on `stream` (a hipStream_t; NULL = default stream).
 *   - Return value: LIME_OK (0) or a negative LIME_ERR_* code; lime_last_error() then
 *     returns a thread-local, human-readable message.  No C++ exception crosses the ABI.
 *
 * The reference (binggu56/lime) is pure Python: it has no FFI.  Each entry point below
 * replaces the NumPy loop named in its comment; the Python facade in lime_amd/oqs.py keeps
 * the reference's call surface and calls these through ctypes.
 */
#ifndef LIME_HIP_H
#define LIME_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LIME_OK 0
#define LIME_ERR_ARG (-1)          /* invalid argument (shape, null pointer, size)      */
#define LIME_ERR_UNSUPPORTED (-2)  /* valid request this build has no kernel for        */
#define LIME_ERR_HIP (-3)          /* a HIP runtime call failed                         */

/* lime_lindblad_batch.flags */
#define LIME_F_RENORM_TRACE 1      /* renormalise Tr rho = 1 after every step (density
                                      matrices only; never set for QRT operators)       */
#define LIME_F_PREC_MASK 0x30      /* arithmetic of the right-hand-side GEMMs:          */
#define LIME_F_PREC_FP32 0x00      /*   complex64 operands on FP32 MFMA (default)       */
#define LIME_F_PREC_3XBF16 0x10    /*   complex64 operands split x = hi + lo (BF16) and  */
                                   /*   multiplied as hi*hi + hi*lo + lo*hi on BF16 MFMA */
                                   /*   (tiled kernels; lime_lindblad_rk4 only)          */
#define LIME_F_PREC_FP64 0x20      /*   complex128 operands on FP64 MFMA (tiled kernels;
                                        operator sets and x/y64 of lime_lindblad_apply are
                                        then complex128)                                  */
#define LIME_F_FORCE_TILED 0x100   /* use the per-stage tiled kernels even when n fits the
                                      resident kernel (testing / tuning)                  */
#define LIME_F_ONE_SIDED 0x200     /* drop the X H_eff^dag product: f(X) = -i H_eff X + D o X
                                      + sum_k L_k X R_k^dag, i.e. dX/dt = G X with H_eff = i G
                                      (Liouville-space generators acting on columns;
                                      per-stage kernels)                                   */

/*
 * A uniform batch of `batch` propagated operators, all of Hilbert dimension n.
 * Operator sets (H_eff, D, dense L) are shared by `group` consecutive operators:
 * operator b uses set b / group.  Observable sets are shared by `eop_group` operators.
 *
 * Right-hand side (H_eff folding of every jump operator):
 *   f(X) = -i (H_eff X - X H_eff^dag) + D o X + sum_k L_k X R_k^dag
 *   H_eff = H - (i/2) sum_(all jump ops) L^dag L;  D = sum_(diagonal L=diag(l)) l l^dag;
 *   L_k = the dense (non-diagonal) jump operators; R_k = L_k unless `rops` is given
 *   (general sandwich terms, e.g. the Hilbert-space form of the Redfield tensor).
 *
 * FP32 with n <= 80 (n <= 64 with dense operators) runs the resident kernel: one
 * workgroup per operator for all steps, no workspace; n = 16m + 1 without dense operators
 * uses its bordered layout.  Larger n, PREC_FP64, PREC_3XBF16 and FORCE_TILED run one
 * launch per RK4 stage and need `work` of lime_lindblad_workspace_bytes() bytes
 * (caller-owned device memory; the library never allocates).
 */
typedef struct lime_lindblad_batch {
  int32_t batch;       /* number of propagated operators                               */
  int32_t n;           /* Hilbert dimension                                            */
  int32_t group;       /* operators per operator set (>= 1)                            */
  int32_t n_dense;     /* dense jump operators per set                                 */
  int32_t n_eops;      /* observables per operator                                     */
  int32_t eop_group;   /* operators per observable set (>= 1)                          */
  int32_t flags;       /* LIME_F_*                                                     */
  int32_t variant;     /* launch-shape variant (0 = default; see lime_lindblad_variants) */
  const void* heff;    /* complex64  [sets][n][n]  (complex128 if PREC_FP64)           */
  const void* dmat;    /* complex64  [sets][n][n]  or NULL (complex128 if PREC_FP64)   */
  const void* lops;    /* complex64  [sets][n_dense][n][n] or NULL (c128 if PREC_FP64) */
  const void* eops_t;  /* complex128 [eop_sets][n_eops][n][n], each TRANSPOSED (E^T)  */
  void* rho;           /* complex128 [batch][n][n], state, updated in place            */
  void* obs;           /* complex128 [batch][rec_capacity][n_eops] or NULL             */
  void* states;        /* complex128 [batch][rec_capacity][n][n] or NULL               */
  const void* rops;    /* complex64  [sets][n_dense][n][n] right factors, or NULL (= lops) */
  void* work;          /* device workspace of the per-stage kernels (lime_lindblad_workspace_bytes) */
  int64_t work_bytes;
  void* stage_states;  /* complex128 [rec_capacity][3][batch][n][n] (record-major) or NULL:
                          the FP64 RK4 stage operands X + h/2 k1, X + h/2 k2, X + h k3 of
                          every recorded step (per-stage kernels only: set
                          LIME_F_FORCE_TILED for n <= 80)                                  */
} lime_lindblad_batch;

typedef struct lime_step_params {
  double dt;
  int32_t nsteps;       /* RK4 steps taken by this call                                */
  int32_t nout;         /* record after every nout-th step (global step count)         */
  int32_t step_offset;  /* global index of this call's first step (chunked runs)       */
  int32_t rec_capacity; /* records allocated per operator in obs / states              */
} lime_step_params;

/*
 * Fused batched Lindblad RK4: all four stages of `nsteps` steps in one launch, the
 * operator resident on chip, GEMMs on MFMA, state update / trace / observables in FP64.
 * Records Tr[E rho] = sum_ij E^T_ij rho_ij after global step g when (g+1) % nout == 0,
 * at record index (g+1)/nout - 1.  Records are STORED (obs and states need no
 * initialisation; entries of records outside this call are left untouched) and are
 * bitwise reproducible from run to run on every path.
 *
 * Replaces the time loop of lime/oqs.py:1590-1688 (_lindblad: rk4 at phys.py:636-649
 * over liouvillian/lindbladian at oqs.py:706-723, obs_dm at phys.py:837-844) and of the
 * QRT drivers oqs.py:726-800 (_correlation_2p_1t), oqs.py:1227-1331 and
 * lime/correlation.py:17-69, which all step the same right-hand side.
 */
int lime_lindblad_rk4(const lime_lindblad_batch* d, const lime_step_params* p, void* stream);

/*
 * An uneven batch: `count` independent Lindblad systems of their own dimensions, packed.
 * Each instance's operators are complex64 regions of `ops` (H_eff, optional Hadamard
 * dissipator D, n_dense dense jump operators L_k, all n x n row-major), its state a
 * complex128 region of `rho` (n x n, updated in place), and its per-stage operands a
 * complex64 region of `work` of (3 + n_dense) n^2 elements.  Offsets count elements.
 * The same right-hand side as lime_lindblad_batch with R_k = L_k.
 */
typedef struct lime_ragged_instance {
  int32_t n;
  int32_t n_dense;
  int64_t heff;        /* offset of H_eff in ops                                        */
  int64_t dmat;        /* offset of D in ops, or -1 (no diagonal jump operators)        */
  int64_t lops;        /* offset of L_0 (L_k follow contiguously) in ops when n_dense > 0 */
  int64_t rho;         /* offset of the state in rho                                    */
  int64_t work;        /* offset of the instance's workspace in work                    */
} lime_ragged_instance;

typedef struct lime_ragged_batch {
  int32_t count;                      /* instances                                       */
  int32_t variant;                    /* launch shape, [0, lime_ragged_variants())       */
  int32_t max_groups;                 /* cap on persistent workgroups (0 = fill the GPU) */
  int32_t reserved;
  const lime_ragged_instance* inst;   /* DEVICE array [count]                            */
  const int32_t* order;               /* DEVICE array [count]: processing order, largest
                                         n^3 (1 + n_dense) first                         */
  const void* ops;                    /* complex64 operator pool                         */
  void* rho;                          /* complex128 state pool                           */
  void* work;                         /* complex64 workspace pool                        */
  void* queue;                        /* DEVICE uint32 [2], zero before the first call;
                                         the kernel leaves it zero when it ends          */
} lime_ragged_batch;

/*
 * Advance every instance of an uneven batch by p->nsteps RK4 steps (p->dt) in ONE kernel
 * launch: a persistent grid drains the instances in `order` through an atomic queue, each
 * instance on one workgroup for all its steps.  Nothing is recorded (p->nout and
 * rec_capacity are ignored); read the states from `rho`.  The caller guarantees that the
 * offsets address its pools (the library does not read the device descriptor on the host).
 * Capturable in a HIP graph (no host synchronisation, no allocation).
 * Replaces lime's one-system-at-a-time loop: _lindblad (lime/oqs.py:1590-1688) called once
 * per instance, each running rk4 (phys.py:636-649) over liouvillian (oqs.py:706-723).
 */
int lime_lindblad_rk4_ragged(const lime_ragged_batch* r, const lime_step_params* p, void* stream);

/* Number of launch shapes of lime_lindblad_rk4_ragged (the autotuner times each one). */
int lime_ragged_variants(void);

/* Size of the static device scratch array of lime_lindblad_rk4_noalloc (per GPU). */
#define LIME_NOALLOC_SCRATCH_BYTES (64ll << 20)

/*
 * lime_lindblad_rk4 without any allocation and without a caller workspace: the per-stage
 * kernels run on a static device array of LIME_NOALLOC_SCRATCH_BYTES per GPU (held by
 * each GPU the library runs on until the process ends); d->work / d->work_bytes are
 * ignored.  Calls no hipMalloc / hipFree.  Returns LIME_ERR_ARG when the batch needs more
 * scratch than that (lime_lindblad_workspace_bytes tells).  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (concurrent streams or threads) race.
 * Nothing in the array needs to start at zero; it keeps the last call's contents.
 * No counterpart in lime: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 */
int lime_lindblad_rk4_noalloc(const lime_lindblad_batch* d, const lime_step_params* p, void* stream);

/*
 * One application of the generator to every operator of the batch (adjoint and stage
 * recomputation building block):
 *   V = f(X)   (the right-hand side above; X in the operand precision, [batch][n][n])
 *   y   = c_lx V + c_z z   (complex128, optional; z optional)
 *   y64 = copy of y in the operand precision (complex64, or complex128 with PREC_FP64)
 *   acc += c_acc V         (complex128, optional)
 * Complex scalars are passed as (re, im) pairs.  Uses d->heff/dmat/lops/rops/group and,
 * when n_dense > 0, d->work (>= batch * n_dense * n * n elements of the operand
 * precision: 8 bytes, 16 with PREC_FP64) for L_k X.  Built for PREC_FP32 and PREC_FP64.
 * Replaces one call of lime's liouvillian (lime/oqs.py:706-713); with the adjoint operator
 * set (H_eff -> -H_eff^dag, D -> conj(D), (L, R) -> (L^dag, R^dag)) it applies f^dag.
 */
int lime_lindblad_apply(const lime_lindblad_batch* d, const void* x64, const void* z, void* y,
                        void* y64, void* acc, const double* c_lx, const double* c_z,
                        const double* c_acc, void* stream);

/*
 * Parameter-gradient contractions of the discrete RK4 adjoint: with P = conj(kappa) o s
 * (both complex128 [batch][n][n]) accumulate, per operator,
 *   out_a[j] += w (sum_m P_jm - sum_m P_mj)          (Im(out_a) = d/dH_jj contribution)
 *   out_b[j] += w (P_jj - (sum_m P_jm + sum_m P_mj)/2)  (Re(out_b) = d/d gamma_j)
 *   out_u[j] += w (coupling contraction of H_j,j+1 = H_j+1,j) (Im(out_u))
 * out_a, out_b: complex128 [batch][n]; out_u: complex128 [batch][n-1].
 * Deterministic (fixed summation order, no atomics).
 * No counterpart in lime (it has no gradients).
 */
int lime_adjoint_contract(int32_t batch, int32_t n, const void* kappa, const void* s, double w,
                          void* out_a, void* out_b, void* out_u, void* stream);

/* The same with P = sum_p conj(kappa[p]) o s[p] over `pairs` (1..4) operator pairs in one
 * launch (`kappa`, `s`: HOST arrays of device pointers): the four stage contractions of an
 * RK4 step.  No counterpart in lime. */
int lime_adjoint_contract_pairs(int32_t batch, int32_t n, int32_t pairs, const void* const* kappa,
                                const void* const* s, double w, void* out_a, void* out_b, void* out_u,
                                void* stream);

/*
 * Start of one backward RK4 step of the adjoint, elementwise over complex128 [batch][n][n]:
 *   lam += g[b] m,  k4 = c lam,  k4_op = k4 in the operand precision of `flags`
 *   (complex64, or complex128 with LIME_F_PREC_FP64),  acc = lam.
 * g: complex128 [batch]; m: complex128 [n][n] shared by the batch.  No counterpart in lime.
 */
int lime_adjoint_seed(int32_t batch, int32_t n, int32_t flags, void* lam, const void* g, const void* m,
                      double c, void* k4, void* k4_op, void* acc, void* stream);

/* Largest Hilbert dimension the resident (workspace-free) kernel accepts. */
int lime_lindblad_max_n(int32_t n_dense, int32_t flags);

/* Number of launch-shape variants for this descriptor (the autotuner times each one;
 * any value in [0, count) is valid in lime_lindblad_batch.variant). */
int lime_lindblad_variants(const lime_lindblad_batch* d);

/* Workspace bytes lime_lindblad_rk4 needs for this descriptor (0 when none). */
int64_t lime_lindblad_workspace_bytes(const lime_lindblad_batch* d);

/* Thread-local message describing the last error returned on this thread. */
const char* lime_last_error(void);

/* Library version string. */
const char* lime_version(void);

#ifdef __cplusplus
}
#endif

#endif /* LIME_HIP_H */
