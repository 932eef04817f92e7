// Lindblad RK4 for operators too large to stay resident in one CU (n > 96).
//
// Same right-hand side as the resident kernel,
//   f(X) = -i (H_eff X - X H_eff^dag) + D o X + sum_k L_k X R_k^dag   (R_k = L_k for Lindblad),
// but one launch per RK4 stage: the kernel boundary is the grid-wide barrier between
// stages.  Each workgroup owns a TILE x TILE output tile of one operator:
//   * K-panels of 16 complex columns are staged global -> registers -> LDS (the next panel
//     is fetched into registers while the current one feeds the MFMAs);
//   * 4 waves in a 2 x 2 arrangement, each owning (TILE/32)^2 16x16 MFMA blocks;
//   * the RK4 stage update runs in the epilogue: k = -i C + D o S, the FP32 increment,
//     the next stage operand (complex64, double buffered so no tile reads a half-written
//     operand), and at the last stage rho += increment in FP64, records (observables by
//     FP64 atomics, states) and the next S = rho.
// The dense products T_k = L_k S of a stage come from lindblad_left_tiled, launched first.
#include "common.hpp"
#include "lindblad_tiled_decl.hpp"

namespace lime {

namespace {

constexpr int KP = 16;          // K-panel width (complex elements)
constexpr int KPP = KP + 1;     // LDS pitch of [row][k] panels (odd: conflict-free fragments)

// Complex accumulation modes of one product term:
//   PLAIN : C += A B                     (B[k][j] given row-major)
//   NEG_AH: C -= A M^dag                 (M given row-major; B[k][j] = conj(M[j][k]))
//   I_AH  : C += i A M^dag
enum Mode { PLAIN = 0, NEG_AH = 1, I_AH = 2 };

template <int TILE>
struct Frag {
  static constexpr int BW = TILE / 32;   // MFMA blocks per wave per dimension
};

// Load a TILE x KP panel of a row-major n x n complex64 matrix: rows r0.., cols k0..
template <int TILE>
__device__ __forceinline__ void load_rows(const float2* __restrict__ M, int n, int r0, int k0,
                                          float2 (&reg)[TILE * KP / 256], int tid) {
#pragma unroll
  for (int q = 0; q < TILE * KP / 256; ++q) {
    const int e = tid + q * 256;
    const int r = e / KP, c = e % KP;
    const int gr = r0 + r, gc = k0 + c;
    reg[q] = (gr < n && gc < n) ? M[(size_t)gr * n + gc] : make_float2(0.f, 0.f);
  }
}

template <int TILE>
__device__ __forceinline__ void store_rows(float2* sM, const float2 (&reg)[TILE * KP / 256], int tid) {
#pragma unroll
  for (int q = 0; q < TILE * KP / 256; ++q) {
    const int e = tid + q * 256;
    sM[(e / KP) * KPP + e % KP] = reg[q];
  }
}

// Load a KP x TILE panel: rows k0.. (the k index), cols c0..
template <int TILE>
__device__ __forceinline__ void load_cols(const float2* __restrict__ M, int n, int k0, int c0,
                                          float2 (&reg)[TILE * KP / 256], int tid) {
#pragma unroll
  for (int q = 0; q < TILE * KP / 256; ++q) {
    const int e = tid + q * 256;
    const int r = e / TILE, c = e % TILE;
    const int gr = k0 + r, gc = c0 + c;
    reg[q] = (gr < n && gc < n) ? M[(size_t)gr * n + gc] : make_float2(0.f, 0.f);
  }
}

template <int TILE>
__device__ __forceinline__ void store_cols(float2* sM, const float2 (&reg)[TILE * KP / 256], int tid) {
#pragma unroll
  for (int q = 0; q < TILE * KP / 256; ++q) {
    const int e = tid + q * 256;
    sM[(e / TILE) * (TILE + 1) + e % TILE] = reg[q];
  }
}

// Accumulate one product term into the wave's MFMA blocks.
//   A: row-major n x n (rows of the tile); Bsrc: row-major n x n;
//   MODE PLAIN uses B[k][j] = Bsrc[k][j]; the other modes B[k][j] = conj(Bsrc[j][k]).
template <int TILE, int MODE>
__device__ __forceinline__ void gemm_term(const float2* __restrict__ A, const float2* __restrict__ Bsrc,
                                          int n, int row0, int col0, float2* sA, float2* sB,
                                          floatx4 (&cr)[TILE / 32][TILE / 32],
                                          floatx4 (&ci)[TILE / 32][TILE / 32],
                                          int tid, int wr, int wc, int lr, int lg) {
  constexpr int BW = TILE / 32;
  constexpr int NR = TILE * KP / 256;
  float2 ra[NR], rb[NR];
  load_rows<TILE>(A, n, row0, 0, ra, tid);
  if (MODE == PLAIN) load_cols<TILE>(Bsrc, n, 0, col0, rb, tid);
  else load_rows<TILE>(Bsrc, n, col0, 0, rb, tid);
  for (int k0 = 0; k0 < n; k0 += KP) {
    __syncthreads();                          // previous panel fully consumed
    store_rows<TILE>(sA, ra, tid);
    if (MODE == PLAIN) store_cols<TILE>(sB, rb, tid);
    else store_rows<TILE>(sB, rb, tid);
    __syncthreads();
    if (k0 + KP < n) {                        // prefetch the next panel into registers
      load_rows<TILE>(A, n, row0, k0 + KP, ra, tid);
      if (MODE == PLAIN) load_cols<TILE>(Bsrc, n, k0 + KP, col0, rb, tid);
      else load_rows<TILE>(Bsrc, n, col0, k0 + KP, rb, tid);
    }
#pragma unroll
    for (int kk = 0; kk < KP; kk += 4) {
      float2 x[BW], y[BW];
#pragma unroll
      for (int p = 0; p < BW; ++p) {
        x[p] = sA[(wr * (TILE / 2) + p * 16 + lr) * KPP + kk + lg];
        if (MODE == PLAIN) y[p] = sB[(kk + lg) * (TILE + 1) + wc * (TILE / 2) + p * 16 + lr];
        else y[p] = sB[(wc * (TILE / 2) + p * 16 + lr) * KPP + kk + lg];
      }
#pragma unroll
      for (int p = 0; p < BW; ++p)
#pragma unroll
        for (int q = 0; q < BW; ++q) {
          const float2 a = x[p], h = y[q];
          if (MODE == PLAIN) {
            cr[p][q] = mfma_f32(a.x, h.x, cr[p][q]);
            ci[p][q] = mfma_f32(a.x, h.y, ci[p][q]);
            cr[p][q] = mfma_f32(-a.y, h.y, cr[p][q]);
            ci[p][q] = mfma_f32(a.y, h.x, ci[p][q]);
          } else if (MODE == NEG_AH) {   // C -= a conj(h)
            cr[p][q] = mfma_f32(-a.x, h.x, cr[p][q]);
            ci[p][q] = mfma_f32(a.x, h.y, ci[p][q]);
            cr[p][q] = mfma_f32(-a.y, h.y, cr[p][q]);
            ci[p][q] = mfma_f32(-a.y, h.x, ci[p][q]);
          } else {                       // C += i a conj(h)
            cr[p][q] = mfma_f32(a.x, h.y, cr[p][q]);
            ci[p][q] = mfma_f32(a.x, h.x, ci[p][q]);
            cr[p][q] = mfma_f32(-a.y, h.x, cr[p][q]);
            ci[p][q] = mfma_f32(a.y, h.y, ci[p][q]);
          }
        }
    }
  }
}

template <int TILE>
__device__ __forceinline__ int tiles_per_dim(int n) { return (n + TILE - 1) / TILE; }

}  // namespace

// T[b][k] = L[set][k] @ S[b]  for every dense operator k.
template <int TILE>
__global__ void __launch_bounds__(256) lindblad_left_tiled(TiledArgs a) {
  __shared__ float2 sA[TILE * KPP];
  __shared__ float2 sB[KP * (TILE + 1) > TILE * KPP ? KP * (TILE + 1) : TILE * KPP];
  const int tpd = tiles_per_dim<TILE>(a.n);
  const int per = tpd * tpd;
  const int blk = blockIdx.x;
  const int tile = blk % per;
  const int bk = blk / per;                   // = b * n_dense + k
  const int b = bk / a.n_dense, k = bk % a.n_dense;
  const int set = b / a.group;
  const int row0 = (tile / tpd) * TILE, col0 = (tile % tpd) * TILE;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = w >> 1, wc = w & 1, lr = lane & 15, lg = lane >> 4;
  const size_t nn = (size_t)a.n * a.n;
  constexpr int BW = TILE / 32;
  floatx4 cr[BW][BW], ci[BW][BW];
#pragma unroll
  for (int p = 0; p < BW; ++p)
#pragma unroll
    for (int q = 0; q < BW; ++q) cr[p][q] = ci[p][q] = floatx4{0.f, 0.f, 0.f, 0.f};
  gemm_term<TILE, PLAIN>(a.lops + ((size_t)set * a.n_dense + k) * nn, a.s_cur + (size_t)b * nn, a.n,
                         row0, col0, sA, sB, cr, ci, tid, wr, wc, lr, lg);
  float2* T = a.tbuf + ((size_t)b * a.n_dense + k) * nn;
#pragma unroll
  for (int p = 0; p < BW; ++p)
#pragma unroll
    for (int q = 0; q < BW; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = row0 + wr * (TILE / 2) + p * 16 + 4 * lg + v;
        const int j = col0 + wc * (TILE / 2) + q * 16 + lr;
        if (i < a.n && j < a.n) T[(size_t)i * a.n + j] = make_float2(cr[p][q][v], ci[p][q][v]);
      }
}

// One RK4 stage for every operator of the batch.
template <int TILE>
__global__ void __launch_bounds__(256) lindblad_stage_tiled(TiledArgs a) {
  __shared__ float2 sA[TILE * KPP];
  __shared__ float2 sB[KP * (TILE + 1) > TILE * KPP ? KP * (TILE + 1) : TILE * KPP];
  __shared__ double2 sRed[4];
  const int n = a.n;
  const int tpd = tiles_per_dim<TILE>(n);
  const int per = tpd * tpd;
  const int b = blockIdx.x / per, tile = blockIdx.x % per;
  const int set = b / a.group;
  const int row0 = (tile / tpd) * TILE, col0 = (tile % tpd) * TILE;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = w >> 1, wc = w & 1, lr = lane & 15, lg = lane >> 4;
  const size_t nn = (size_t)n * n;
  constexpr int BW = TILE / 32;

  floatx4 cr[BW][BW], ci[BW][BW];
#pragma unroll
  for (int p = 0; p < BW; ++p)
#pragma unroll
    for (int q = 0; q < BW; ++q) cr[p][q] = ci[p][q] = floatx4{0.f, 0.f, 0.f, 0.f};

  const float2* H = a.heff + (size_t)set * nn;
  const float2* S = a.s_cur + (size_t)b * nn;
  gemm_term<TILE, PLAIN>(H, S, n, row0, col0, sA, sB, cr, ci, tid, wr, wc, lr, lg);
  gemm_term<TILE, NEG_AH>(S, H, n, row0, col0, sA, sB, cr, ci, tid, wr, wc, lr, lg);
  for (int k = 0; k < a.n_dense; ++k) {
    const float2* T = a.tbuf + ((size_t)b * a.n_dense + k) * nn;
    const float2* R = (a.rops ? a.rops : a.lops) + ((size_t)set * a.n_dense + k) * nn;
    gemm_term<TILE, I_AH>(T, R, n, row0, col0, sA, sB, cr, ci, tid, wr, wc, lr, lg);
  }

  // ---- RK4 stage epilogue
  const int st = a.stage;
  const double dt = a.dt;
  const float cacc = (float)((st == 0 || st == 3) ? dt / 6.0 : dt / 3.0);
  const double cnext = (st == 2) ? dt : 0.5 * dt;
  double2* rho = a.rho + (size_t)b * nn;
  float2* inc = a.inc + (size_t)b * nn;
  float2* Sn = a.s_next + (size_t)b * nn;
  const float2* Dm = a.dmat ? a.dmat + (size_t)set * nn : nullptr;
  const bool rec = (st == 3) && a.rec >= 0;
  double2 rfin[BW][BW][4];
#pragma unroll
  for (int p = 0; p < BW; ++p)
#pragma unroll
    for (int q = 0; q < BW; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        rfin[p][q][v] = make_double2(0.0, 0.0);
        const int i = row0 + wr * (TILE / 2) + p * 16 + 4 * lg + v;
        const int j = col0 + wc * (TILE / 2) + q * 16 + lr;
        if (i >= n || j >= n) continue;
        const size_t o = (size_t)i * n + j;
        float kr = ci[p][q][v], ki = -cr[p][q][v];
        if (Dm) {
          const float2 d = Dm[o], x = S[o];
          kr += d.x * x.x - d.y * x.y;
          ki += d.x * x.y + d.y * x.x;
        }
        float2 ic = (st == 0) ? make_float2(0.f, 0.f) : inc[o];
        ic.x = fmaf(cacc, kr, ic.x);
        ic.y = fmaf(cacc, ki, ic.y);
        double2 r = rho[o];
        if (st < 3) {
          inc[o] = ic;
          Sn[o] = make_float2((float)(r.x + cnext * (double)kr), (float)(r.y + cnext * (double)ki));
        } else {
          r.x += (double)ic.x;
          r.y += (double)ic.y;
          rho[o] = r;
          rfin[p][q][v] = r;
          Sn[o] = make_float2((float)r.x, (float)r.y);
          if (rec && a.states) a.states[((size_t)b * a.rec_capacity + a.rec) * nn + o] = r;
        }
      }
  if (rec && a.obs) {
    const double2* E0 = a.eops_t + (size_t)(b / a.eop_group) * a.n_eops * nn;
    for (int e = 0; e < a.n_eops; ++e) {
      double pr = 0.0, pi = 0.0;
#pragma unroll
      for (int p = 0; p < BW; ++p)
#pragma unroll
        for (int q = 0; q < BW; ++q)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = row0 + wr * (TILE / 2) + p * 16 + 4 * lg + v;
            const int j = col0 + wc * (TILE / 2) + q * 16 + lr;
            if (i >= n || j >= n) continue;
            const double2 ev = E0[(size_t)e * nn + (size_t)i * n + j];
            pr += ev.x * rfin[p][q][v].x - ev.y * rfin[p][q][v].y;
            pi += ev.x * rfin[p][q][v].y + ev.y * rfin[p][q][v].x;
          }
      const double sr = wave_sum(pr), si = wave_sum(pi);
      __syncthreads();
      if (lane == 0) sRed[w] = make_double2(sr, si);
      __syncthreads();
      if (tid == 0) {
        double2 t = make_double2(0.0, 0.0);
        for (int x = 0; x < 4; ++x) {
          t.x += sRed[x].x;
          t.y += sRed[x].y;
        }
        double* dst = reinterpret_cast<double*>(a.obs + ((size_t)b * a.rec_capacity + a.rec) * a.n_eops + e);
        atomicAdd(dst, t.x);
        atomicAdd(dst + 1, t.y);
      }
    }
  }
}

template <int TILE>
static hipError_t run_step(TiledArgs a, hipStream_t stream, float2* s0, float2* s1) {
  const int tpd = (a.n + TILE - 1) / TILE;
  const int per = tpd * tpd;
  for (int st = 0; st < 4; ++st) {
    a.stage = st;
    a.s_cur = (st % 2 == 0) ? s0 : s1;
    a.s_next = (st % 2 == 0) ? s1 : s0;
    if (a.n_dense > 0) {
      hipLaunchKernelGGL(lindblad_left_tiled<TILE>, dim3(per * a.batch * a.n_dense), dim3(256), 0,
                         stream, a);
    }
    hipLaunchKernelGGL(lindblad_stage_tiled<TILE>, dim3(per * a.batch), dim3(256), 0, stream, a);
  }
  return hipGetLastError();
}

__global__ void init_stage_operand(const double2* __restrict__ rho, float2* __restrict__ s, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) s[i] = make_float2((float)rho[i].x, (float)rho[i].y);
}

size_t tiled_workspace_bytes(int batch, int n, int n_dense) {
  const size_t nn = (size_t)n * n;
  return (size_t)batch * nn * sizeof(float2) * (3 + (size_t)n_dense);   // S0, S1, inc, T_k
}

int tiled_pick(int batch, int n) {
  const int t64 = (n + 63) / 64;
  return ((size_t)batch * t64 * t64 >= 512) ? 64 : 32;
}

hipError_t tiled_run(TiledArgs a, int nsteps, int step_offset, int nout, void* work, hipStream_t stream,
                     int tile) {
  const size_t nn = (size_t)a.n * a.n;
  float2* s0 = static_cast<float2*>(work);
  float2* s1 = s0 + (size_t)a.batch * nn;
  a.inc = s1 + (size_t)a.batch * nn;
  a.tbuf = a.inc + (size_t)a.batch * nn;
  const size_t cnt = (size_t)a.batch * nn;
  hipLaunchKernelGGL(init_stage_operand, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, stream,
                     a.rho, s0, cnt);
  for (int s = 0; s < nsteps; ++s) {
    const int g = step_offset + s + 1;
    a.rec = -1;
    if (nout > 0 && (a.obs || a.states) && g % nout == 0 && g / nout - 1 < a.rec_capacity) a.rec = g / nout - 1;
    // stage 3 writes S = rho into s0 (stages alternate s0 -> s1 -> s0 -> s1 -> s0)
    hipError_t e = (tile == 64) ? run_step<64>(a, stream, s0, s1) : run_step<32>(a, stream, s0, s1);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

}  // namespace lime
