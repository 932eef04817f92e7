// Fused batched Lindblad RK4, one workgroup per propagated operator, state resident on chip.
//
// Replaces the NumPy time loop of lime's _lindblad (lime/oqs.py:1590-1688):
//   rho = rk4(rho, liouvillian, dt, H, c_ops)                   lime/phys.py:636-649
//   liouvillian = -1j*comm(H, rho) + sum lindbladian(c, rho)    lime/oqs.py:706-723
//   observables[k] = [obs_dm(rho, e) for e in e_ops]            lime/phys.py:837-844
// with the right-hand side folded to
//   f(X) = -i (H_eff X - X H_eff^dag) + D o X + sum_k L_k X L_k^dag.
//
// Work decomposition (wave64, 16x16x4 FP32 MFMA tiles):
//   * the operator is zero-padded to NP = 16*NB (bordered layouts: below); its NB*NB output tiles of 16x16 are
//     dealt to the W waves of the workgroup, tile t = w + u*W for u < BPW;
//   * lane l owns, in each of its tiles, the 4 elements (4*(l/16)+v, l%16), v = 0..3 --
//     the MFMA accumulator layout -- so the FP64 state (rho and the RK4 accumulator)
//     stays in VGPRs for the whole launch and never round-trips through memory;
//   * LDS holds H_eff and the current stage operand S as complex64 with a row pitch of
//     NP+1 elements (odd pitch: column-wise A-fragment reads hit distinct banks);
//   * per stage: C = H_eff S - S H_eff^dag (+ i sum_k T_k L_k^dag, T_k = L_k S staged
//     through LDS) on FP32 MFMA; k = -i C + D o S; the RK4 combination in FP64; the next
//     stage operand is rounded to complex64 into LDS;
//   * records (observables Tr[E rho], optional states) are reduced in FP64.
// Padded rows/columns hold exact zeros in H_eff, D, L and S, so they stay zero.
#include "common.hpp"
#include "lindblad_resident_decl.hpp"

namespace lime {

namespace {

// D o S at one element (complex product d s)
__device__ __forceinline__ float2 cmulf(float2 d, float2 s) {
  return make_float2(d.x * s.x - d.y * s.y, d.x * s.y + d.y * s.x);
}

// Copy a row-major n x n complex64 operator into an NP x P LDS buffer (zero padded).
template <int NP, int P, int T>
__device__ __forceinline__ void stage_operator(float2* dst, const float2* __restrict__ src, int n, int tid) {
  for (int idx = tid; idx < NP * NP; idx += T) {
    const int i = idx / NP, j = idx - (idx / NP) * NP;
    float2 v = make_float2(0.f, 0.f);
    if (i < n && j < n) v = src[(size_t)i * n + j];
    dst[i * P + j] = v;
  }
}

// Software-pipelined complex product over LDS operands for the wave's tiles:
//   MODE 0: C += A B          (A rows from sA, B[k][j] = sB[k][j])
//   MODE 2: C += i A M^dag    (A rows from sA, B[k][j] = conj(sM[j][k]), sM = sB)
// The reduction runs over k < ke (operands are zero beyond n).  sA / sB may be read up to
// 4 columns / rows past ke (padding, or the spare LDS after each buffer).
template <int NP, int P, int BPW, int MODE>
__device__ __forceinline__ void lds_gemm(const float2* sA, const float2* sB, const int (&trow)[BPW],
                                         const int (&tcol)[BPW], int lr, int lg, int ke, floatx4 (&cr)[BPW],
                                         floatx4 (&ci)[BPW]) {
  int oA[BPW], oB[BPW];
#pragma unroll
  for (int u = 0; u < BPW; ++u) {
    oA[u] = (trow[u] + lr) * P + lg;
    oB[u] = (MODE == 0) ? lg * P + tcol[u] + lr : (tcol[u] + lr) * P + lg;
  }
  float2 x[BPW], y[BPW];
#pragma unroll
  for (int u = 0; u < BPW; ++u) {
    x[u] = sA[oA[u]];
    y[u] = sB[oB[u]];
  }
#pragma unroll 2
  for (int k0 = 0; k0 < ke; k0 += 4) {
    float2 xn[BPW], yn[BPW];
#pragma unroll
    for (int u = 0; u < BPW; ++u) {
      xn[u] = sA[oA[u] + k0 + 4];
      yn[u] = sB[oB[u] + (MODE == 0 ? (k0 + 4) * P : k0 + 4)];
    }
#pragma unroll
    for (int u = 0; u < BPW; ++u) {
      const float2 a = x[u], h = y[u];
      if (MODE == 0) {
        cr[u] = mfma_f32(a.x, h.x, cr[u]);
        ci[u] = mfma_f32(a.x, h.y, ci[u]);
        cr[u] = mfma_f32(-a.y, h.y, cr[u]);
        ci[u] = mfma_f32(a.y, h.x, ci[u]);
      } else {
        cr[u] = mfma_f32(a.x, h.y, cr[u]);
        ci[u] = mfma_f32(a.x, h.x, ci[u]);
        cr[u] = mfma_f32(-a.y, h.x, cr[u]);
        ci[u] = mfma_f32(a.y, h.y, ci[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < BPW; ++u) {
      x[u] = xn[u];
      y[u] = yn[u];
    }
  }
}

}  // namespace

// Bordered sizes (n = 16 m + 1, no dense jump operators): the last index is the border.
// The core block (rows/columns 0..16m-1) runs on MFMA over the reduction range 0..16m+3,
// which already contains the border terms H[i][16m] S[16m][j] of core outputs; the 2*16m+1
// border outputs (row 16m and column 16m) are 65-term-or-shorter complex dot products on
// VALU, each split over BG lanes.  The padded size of the plain kernel would spend
// (m+1)^2 tiles on these outputs instead of m^2.
template <int NB, int W, int BPW, int MINW, bool DENSE, bool REC, bool BRD, bool DB>
__global__ void __launch_bounds__(W * 64) __attribute__((amdgpu_waves_per_eu(MINW)))
lindblad_rk4_resident(ResidentArgs a) {
  // LDS rows / columns held: bordered layouts hold the core, the border index and 7 zero
  // ones (the reductions and the fragment prefetch read up to index 16 (NB - 1) + 7)
  constexpr int NP = BRD ? 16 * (NB - 1) + 8 : NB * 16;
  constexpr int P = NP + 1;
  constexpr int T = W * 64;
  constexpr int TG = BRD ? NB - 1 : NB;     // MFMA tile grid edge
  constexpr int NT = TG * TG;
  static_assert(W * BPW >= NT, "every tile needs an owner");
  constexpr int NC = 16 * (NB - 1);          // bordered: core size = border index
  // bordered: BG lanes per border output, 2 NC groups of them (the corner is summed
  // apart, below); the groups fill whole waves (BG NC is a multiple of 32)
  constexpr int BG = (T >= 8 * NC) ? 4 : 2;
  static_assert(!BRD || T >= 4 * NC, "every border output needs an owner");
  constexpr int BWAVES = BRD ? (2 * NC * BG) / 64 : 0;   // waves holding border groups
  static_assert(!(BRD && DENSE), "bordered kernels are built without dense jump operators");
  // per-lane cache of D (needs the register budget of one tile per lane at <= 4 waves per
  // SIMD, or of at most two waves per SIMD); of a single observable only at <= 2 waves per
  // SIMD (at 4 the registers go to the split-sign accumulators, which pay more)
  constexpr bool CACHE = (BPW == 1 && MINW <= 4) || MINW <= 2;
  constexpr bool ECACHE = CACHE && MINW <= 2;
  // split-sign accumulation: the positive and the negative real products of C go to
  // separate accumulators (subtracted once per stage), so the MFMA stream carries no
  // sign-flip VALU instructions (+8 % at n = 65).  Built where the register budget
  // (512 / MINW VGPRs) covers ~56 VGPRs per tile + 72 of fixed state.
  constexpr bool SPLIT = 512 / MINW >= 56 * BPW + 72;
  // G3 (with SPLIT, no dense operators): the two complex products take 3 real products each
  // (Gauss), 6 MFMAs per k-step instead of 8:
  //   R1 = Hx Sx - Sx Hx',  R2 = Hy Sy + Sy Hy',  M1 = (Hx + Hy)(Sx + Sy) - (Sx + Sy)(Hx' - Hy')
  // (H' = H[j][k] of the S H^dag product), Re C = R1 - R2, Im C = M1 - R1 - R2.  R1 is
  // held as cr - di (Hx Sx in cr, Sx Hx' in di: no negated operand in the MFMA stream), R2
  // in dr, M1 in ci.  A value (cr0, ci0) of C enters them as cr = cr0, ci = ci0 + cr0.
  constexpr bool G3 = SPLIT && !DENSE;
  // PF: the H_eff / S fragments of k-step k0 + 4 are read while step k0's MFMAs issue (one
  // tile per wave; with more, the tiles' MFMAs interleave and the prefetch registers spill)
  constexpr bool PF = BPW == 1;
  // DB: double-buffered stage operand (when it fits the LDS; never with dense operators):
  // the epilogue writes the next S into the other buffer, so a stage needs one barrier
  static_assert(!(DB && DENSE), "the double buffer is built without dense operators");

  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  float2* sH = reinterpret_cast<float2*>(lds_raw);
  float2* sS0 = sH + NP * P;
  // 4 spare rows after each S buffer (and after sT, sL) absorb the unused fragment
  // prefetch past the last k step
  float2* sS1 = DB ? sS0 + NP * P + 4 * P + 4 : sS0;
  float2* sT = sS1 + NP * P + 4 * P + 4;
  float2* sL = sT + (DENSE ? NP * P + 4 * P + 4 : 0);
  double2* sRed = reinterpret_cast<double2*>(sL + (DENSE ? NP * P + 4 * P + 4 : 0));

  const int b = xcd_block();   // the operators of one generator set run on one XCD (shared L2)
  if (b >= a.batch) return;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  const int n = a.n;
  const int set = b / a.group;
  const int nn = n * n;   // n <= 80: element offsets fit in 32 bits
  const int lr = lane & 15;
  const int lg = lane >> 4;
  // reduction extent: operands are zero beyond n, so round n up to the MFMA k-step
  const int ke = BRD ? NC + 4 : ((n + 3) & ~3);

  // ---- stage H_eff into LDS (zero padded to NP); bordered: S is zeroed too, since its
  // padding is not covered by any owner
  {
    const float2* H = a.heff + (size_t)set * nn;
    for (int idx = tid; idx < NP * NP; idx += T) {
      const int i = idx / NP, j = idx - (idx / NP) * NP;
      float2 v = make_float2(0.f, 0.f);
      if (i < n && j < n) v = H[(size_t)i * n + j];
      sH[i * P + j] = v;
      if (BRD) {
        sS0[i * P + j] = make_float2(0.f, 0.f);
        sS1[i * P + j] = make_float2(0.f, 0.f);
      }
    }
  }
  if (BRD) __syncthreads();

  // ---- tile ownership: first row / column of each owned tile (wave-uniform)
  int trow[BPW], tcol[BPW];
  bool tval[BPW];
#pragma unroll
  for (int u = 0; u < BPW; ++u) {
    const int t = w + u * W;
    tval[u] = (W * BPW == NT) || t < NT;   // every slot owns a tile when W * BPW == NT
    const int tt = tval[u] ? t : NT - 1;   // idle slots shadow a real tile, never store
    trow[u] = (tt / TG) * 16;
    tcol[u] = (tt % TG) * 16;
  }
  // ---- bordered: lane group g = tid / BG (g < 2 NC) sums border output (gi, gj), its lane
  // bq taking k = bq, bq + BG, ...; groups 0..NC-1 are the row outputs (NC, g), groups
  // NC..2NC-1 the column outputs (g - NC, NC).  The group's lane 0 owns its output's FP64
  // state (bi, bj) = (gi, gj).  The corner (NC, NC) is summed after the k-loop, one k per
  // lane and a wave sum, by the first wave without border groups (CW = BWAVES, which has
  // slack at the stage barrier; its lane 0 owns it), or, when every wave holds groups, by
  // wave 0 (lane 1 owns it).
  int gi = 0, gj = 0, bq = 0;
  bool bown = false;
  if (BRD) {
    const int grp = tid / BG;
    bq = tid % BG;
    bown = grp < 2 * NC;
    const int gg = bown ? grp : 0;
    if (gg < NC) {
      gi = NC;
      gj = gg;
    } else {
      gi = gg - NC;
      gj = NC;
    }
  }
  constexpr int CW = (W > BWAVES) ? BWAVES : 0;   // wave summing the corner
  const bool bcorner = BRD && tid == (CW > 0 ? 64 * CW : 1);
  const bool bwrite = (bown && bq == 0) || bcorner;
  const int bi = bcorner ? NC : gi, bj = bcorner ? NC : gj;

  double2 rho[BPW][4];   // FP64 state
  float2 inc[BPW][4];    // RK4 increment dt*(k1 + 2k2 + 2k3 + k4)/6, summed in FP32
  double2 rho_b = make_double2(0.0, 0.0);   // bordered: state of the border output
  float2 inc_b = make_float2(0.f, 0.f);
  {
    const double2* R = a.rho + (size_t)b * nn;
#pragma unroll
    for (int u = 0; u < BPW; ++u) {
      const int j = tcol[u] + lr;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = trow[u] + 4 * lg + v;
        double2 x = make_double2(0.0, 0.0);
        if (tval[u] && i < n && j < n) x = R[i * n + j];
        rho[u][v] = x;
        inc[u][v] = make_float2(0.f, 0.f);
        if (tval[u]) sS0[i * P + j] = make_float2((float)x.x, (float)x.y);
      }
    }
    if (bown || bcorner) {
      rho_b = R[bi * n + bj];
      if (bwrite) sS0[bi * P + bj] = make_float2((float)rho_b.x, (float)rho_b.y);
    }
  }
  // ---- register caches: D at the owned elements, the single observable's E^T
  const bool has_d = a.dmat != nullptr;
  float2 dc[CACHE ? BPW : 1][4];
  float2 dc_b = make_float2(0.f, 0.f);
  const bool ecached = ECACHE && REC && a.obs != nullptr && a.n_eops == 1;
  double2 ec[ECACHE && REC ? BPW : 1][4];
  double2 ec_b = make_double2(0.0, 0.0);
  if constexpr (CACHE) {
    const float2* Dm = a.dmat + (size_t)set * nn;
    const double2* E = a.eops_t + (size_t)(b / a.eop_group) * a.n_eops * nn;
#pragma unroll
    for (int u = 0; u < BPW; ++u) {
      const int j = tcol[u] + lr;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = trow[u] + 4 * lg + v;
        const bool in = tval[u] && i < n && j < n;
        dc[u][v] = (has_d && in) ? Dm[i * n + j] : make_float2(0.f, 0.f);
        if constexpr (ECACHE && REC) ec[u][v] = (ecached && in) ? E[i * n + j] : make_double2(0.0, 0.0);
      }
    }
    if (bown || bcorner) {
      if (has_d) dc_b = Dm[bi * n + bj];
      if (ecached) ec_b = E[bi * n + bj];
    }
  } else if (BRD && (bown || bcorner)) {
    if (has_d) dc_b = a.dmat[(size_t)set * nn + bi * n + bj];
    if (REC && ecached) ec_b = a.eops_t[(size_t)(b / a.eop_group) * a.n_eops * nn + bi * n + bj];
  }
  __syncthreads();

  // C accumulators of the stage.  DFOLD (D cached in registers): D o S is their initial
  // value, formed by the epilogue that wrote S -- every lane writes the S elements it owns,
  // so no stage reads S back for the Hadamard term.  br, bm: C at the border output (the
  // lane's partial sum until the group sum); dbr, dbm: the D o S part of the owned border
  // output, added after the group sum
  // (D o S only on the group's lane 0, since the group's partial sums are added up).
  constexpr bool DFOLD = CACHE;
  floatx4 cr[BPW], ci[BPW];
  float br = 0.f, bm = 0.f, dbr = 0.f, dbm = 0.f;
  // the accumulators hold C and the stage derivative is k = -i C, so D o S enters C as
  // (Re C, Im C) = (-Im(D o S), Re(D o S))
  if constexpr (DFOLD) {
#pragma unroll
    for (int u = 0; u < BPW; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float2 p = cmulf(dc[u][v], make_float2((float)rho[u][v].x, (float)rho[u][v].y));
        cr[u][v] = -p.y;
        ci[u][v] = G3 ? p.x - p.y : p.x;
      }
    if (BRD && bwrite) {
      const float2 p = cmulf(dc_b, make_float2((float)rho_b.x, (float)rho_b.y));
      dbr = -p.y;
      dbm = p.x;
    }
  }

  const double dt = a.dt;
  const bool renorm = (a.flags & 1) != 0;
  const bool record_any = REC && a.nout > 0 && (a.obs != nullptr || a.states != nullptr);

  int ph = 0;   // buffer holding the current stage operand
  for (int s = 0; s < a.nsteps; ++s) {
#pragma unroll 1
    for (int st = 0; st < 4; ++st) {
      const float2* sS = ph ? sS1 : sS0;      // read: operand of this stage
      float2* sN = ph ? sS0 : sS1;            // written: operand of the next stage (== sS without DB)
      // Opaque copies of the lane indices: global addresses derived from them are
      // rebuilt inside the loop instead of being hoisted (and spilled) as 64-bit values.
      int olr = lr, olg = lg;
      asm volatile("" : "+v"(olr), "+v"(olg));
      if constexpr (!DFOLD) {
#pragma unroll
        for (int u = 0; u < BPW; ++u) {
          cr[u] = floatx4{0.f, 0.f, 0.f, 0.f};
          ci[u] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
      }
      br = 0.f;
      bm = 0.f;
      float qr = 0.f, qm = 0.f;   // bordered: the corner's sum (wave CW)
      // C += H_eff S  and  C -= S H_eff^dag  ((S H^dag)_ij = sum_k S_ik conj(H_jk)).
      // Software pipelined by hand: fragments of step k0+4 are read from LDS while the
      // MFMAs of step k0 issue, so at most two fragment sets are live.  Bordered: the
      // border dot products take their terms k0..k0+3 in the same step, on VALU while the
      // MFMAs run (lane bq of an output's group takes k = k0 + bq + BG j).
      {
        int oR[BPW], oC[BPW];
#pragma unroll
        for (int u = 0; u < BPW; ++u) {
          oR[u] = (trow[u] + lr) * P + lg;   // row-fragment offset (A of H S, A of S H^dag)
          oC[u] = lg * P + tcol[u] + lr;     // column fragment of S (B of H S)
        }
        int oHc[BPW];
#pragma unroll
        for (int u = 0; u < BPW; ++u) oHc[u] = (tcol[u] + lr) * P + lg;  // B of S H^dag
        float2 x[BPW], y[BPW], z[BPW], h[BPW];
#pragma unroll
        for (int u = 0; u < BPW; ++u) {
          x[u] = sH[oR[u]];
          y[u] = sS[oC[u]];
          z[u] = sS[oR[u]];
          h[u] = sH[oHc[u]];
        }
        // SPLIT: cr, ci collect the positive products of Re C, Im C; dr, di the negative
        // (G3: the four accumulators of the three-product form, see above)
        floatx4 dr[SPLIT ? BPW : 1], di[SPLIT ? BPW : 1];
#pragma unroll
        for (int u = 0; u < (SPLIT ? BPW : 1); ++u) {
          dr[u] = floatx4{0.f, 0.f, 0.f, 0.f};
          di[u] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
        // bordered: the core's k = NC term, H[i][NC] S[NC][j] - S[i][NC] conj(H[j][NC]), on
        // VALU into the accumulators, so the MFMA k-steps stop at NC instead of issuing a
        // fifth-of-a-step of border index and zero padding (k = NC..NC+3)
        if constexpr (BRD) {
#pragma unroll
          for (int u = 0; u < BPW; ++u) {
            const float2 sj = sS[NC * P + tcol[u] + lr], gj = sH[(tcol[u] + lr) * P + NC];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int i = trow[u] + 4 * lg + v;
              const float2 hi = sH[i * P + NC], zi = sS[i * P + NC];
              const float tr = hi.x * sj.x - hi.y * sj.y - (zi.x * gj.x + zi.y * gj.y);
              const float ti = hi.x * sj.y + hi.y * sj.x - (zi.y * gj.x - zi.x * gj.y);
              cr[u][v] += tr;
              ci[u][v] += G3 ? ti + tr : ti;
            }
          }
        }
#pragma unroll 2
        for (int k0 = 0; k0 < ke; k0 += 4) {
          // bordered: MFMA steps k0 < NC only; the step k0 = NC carries the border terms
          const bool mstep = !BRD || k0 < NC;
          float2 xn[BPW], yn[BPW], zn[BPW], hn[BPW];
          if constexpr (!PF) {
            // several tiles per wave interleave their MFMAs: the fragments of this step are
            // read here instead of one step ahead (their registers would spill)
#pragma unroll
            for (int u = 0; u < BPW; ++u) {
              x[u] = sH[oR[u] + k0];
              y[u] = sS[oC[u] + k0 * P];
              z[u] = sS[oR[u] + k0];
              h[u] = sH[oHc[u] + k0];
            }
          }
          // prefetch step k0+4; at k0+4 == ke this reads padding or spare rows, discarded
#pragma unroll
          for (int u = 0; u < BPW && PF; ++u) {
            if (!BRD || k0 + 4 < NC) {
              xn[u] = sH[oR[u] + k0 + 4];
              yn[u] = sS[oC[u] + (k0 + 4) * P];
              zn[u] = sS[oR[u] + k0 + 4];
              hn[u] = sH[oHc[u] + k0 + 4];
            } else {
              xn[u] = x[u];
              yn[u] = y[u];
              zn[u] = z[u];
              hn[u] = h[u];
            }
          }
#pragma unroll
          for (int u = 0; u < BPW && mstep; ++u) {
            if constexpr (G3) {
              const float sx = x[u].x + x[u].y, sy = y[u].x + y[u].y;
              const float sz = z[u].x + z[u].y, dh = h[u].y - h[u].x;
              cr[u] = mfma_f32(x[u].x, y[u].x, cr[u]);
              dr[u] = mfma_f32(x[u].y, y[u].y, dr[u]);
              ci[u] = mfma_f32(sx, sy, ci[u]);
              di[u] = mfma_f32(z[u].x, h[u].x, di[u]);
              dr[u] = mfma_f32(z[u].y, h[u].y, dr[u]);
              ci[u] = mfma_f32(sz, dh, ci[u]);
            } else if constexpr (SPLIT) {
              // Re C = Hx Sx - (Hy Sy + Sx Hx' + Sy Hy'),  Im C = Hx Sy + Hy Sx + Sx Hy' - Sy Hx'
              // (H' = H[j][k] of the S H^dag product)
              cr[u] = mfma_f32(x[u].x, y[u].x, cr[u]);
              ci[u] = mfma_f32(x[u].x, y[u].y, ci[u]);
              dr[u] = mfma_f32(x[u].y, y[u].y, dr[u]);
              di[u] = mfma_f32(z[u].y, h[u].x, di[u]);
              ci[u] = mfma_f32(x[u].y, y[u].x, ci[u]);
              dr[u] = mfma_f32(z[u].x, h[u].x, dr[u]);
              ci[u] = mfma_f32(z[u].x, h[u].y, ci[u]);
              dr[u] = mfma_f32(z[u].y, h[u].y, dr[u]);
            } else {
              cr[u] = mfma_f32(x[u].x, y[u].x, cr[u]);
              ci[u] = mfma_f32(x[u].x, y[u].y, ci[u]);
              cr[u] = mfma_f32(-x[u].y, y[u].y, cr[u]);
              ci[u] = mfma_f32(x[u].y, y[u].x, ci[u]);
              cr[u] = mfma_f32(-z[u].x, h[u].x, cr[u]);
              ci[u] = mfma_f32(z[u].x, h[u].y, ci[u]);
              cr[u] = mfma_f32(-z[u].y, h[u].y, cr[u]);
              ci[u] = mfma_f32(-z[u].y, h[u].x, ci[u]);
            }
          }
          if constexpr (BRD) {
            // border terms k = k0 + bq + BG j (up to NC + 3: zero padding beyond NC); the
            // border waves are whole waves, so these branches are wave-uniform
            if (w < BWAVES) {
#pragma unroll
              for (int j = 0; j < 4 / BG; ++j) {
                const int k = k0 + bq + BG * j;
                const float2 hh = sH[gi * P + k], ss = sS[k * P + gj];
                const float2 zz = sS[gi * P + k], gg = sH[gj * P + k];
                br += hh.x * ss.x - hh.y * ss.y - (zz.x * gg.x + zz.y * gg.y);
                bm += hh.x * ss.y + hh.y * ss.x - (zz.y * gg.x - zz.x * gg.y);
              }
            }
          }
#pragma unroll
          for (int u = 0; u < BPW && PF; ++u) {
            x[u] = xn[u];
            y[u] = yn[u];
            z[u] = zn[u];
            h[u] = hn[u];
          }
        }
        if constexpr (G3) {
#pragma unroll
          for (int u = 0; u < BPW; ++u) {
            cr[u] -= di[u];                   // R1
            ci[u] -= cr[u] + dr[u];
            cr[u] -= dr[u];
          }
        } else if constexpr (SPLIT) {
#pragma unroll
          for (int u = 0; u < BPW; ++u) {
            cr[u] -= dr[u];
            ci[u] -= di[u];
          }
        }
      }
      // bordered: C at the border output, (H S - S H^dag)_ij over k = 0..NC, summed over
      // the output's lane group
      if (BRD) {
        br = group_sum<BG>(br);
        bm = group_sum<BG>(bm);
        if (w == CW) {
          // the corner, after the k-loop: lane l takes k = l (and l + 64), then a wave sum
          // (inside the loop every lane group of the wave would repeat the same 4 terms)
          for (int k = lane; k <= NC; k += 64) {
            const float2 hh = sH[NC * P + k], ss = sS[k * P + NC], zz = sS[NC * P + k];
            qr += hh.x * ss.x - hh.y * ss.y - (zz.x * hh.x + zz.y * hh.y);
            qm += hh.x * ss.y + hh.y * ss.x - (zz.y * hh.x - zz.x * hh.y);
          }
#pragma unroll
          for (int off = 32; off >= 1; off >>= 1) {
            qr += __shfl_xor(qr, off, 64);
            qm += __shfl_xor(qm, off, 64);
          }
          if (bcorner) {
            br = qr;
            bm = qm;
          }
        }
        br += dbr;
        bm += dbm;
      }
      // C += i sum_k (L_k S) R_k^dag, every factor staged through LDS (sL) so the fragment
      // loops are the same pipelined LDS reads as the H_eff products.
      for (int kk = 0; DENSE && kk < a.n_dense; ++kk) {
        const float2* L = a.lops + ((size_t)set * a.n_dense + kk) * nn;
        __syncthreads();                        // previous users of sL / sT are done
        stage_operator<NP, P, T>(sL, L, n, tid);
        __syncthreads();
        floatx4 tr[BPW], ti[BPW];
#pragma unroll
        for (int u = 0; u < BPW; ++u) {
          tr[u] = floatx4{0.f, 0.f, 0.f, 0.f};
          ti[u] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
        lds_gemm<NP, P, BPW, 0>(sL, sS, trow, tcol, lr, lg, ke, tr, ti);    // T = L S
#pragma unroll
        for (int u = 0; u < BPW; ++u) {
          if (!tval[u]) continue;
#pragma unroll
          for (int v = 0; v < 4; ++v)
            sT[(trow[u] + 4 * lg + v) * P + tcol[u] + lr] = make_float2(tr[u][v], ti[u][v]);
        }
        if (a.rops) {                           // general sandwich: R_k replaces L_k in sL
          __syncthreads();
          stage_operator<NP, P, T>(sL, a.rops + ((size_t)set * a.n_dense + kk) * nn, n, tid);
        }
        __syncthreads();
        lds_gemm<NP, P, BPW, 2>(sT, sL, trow, tcol, lr, lg, ke, cr, ci);    // C += i T R^dag
      }
      // k = -i C + D o S, formed in place: afterwards kr = ci, ki = -cr.
      if (!DFOLD && has_d) {
        const float2* Dm = a.dmat + (size_t)set * nn;
#pragma unroll
        for (int u = 0; u < BPW; ++u) {
          if (!tval[u]) continue;
          const int j = tcol[u] + lr;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = trow[u] + 4 * lg + v;
            float2 d;
            if constexpr (CACHE) {
              d = dc[u][v];
            } else {
              d = make_float2(0.f, 0.f);
              if (i < n && j < n) d = Dm[(trow[u] + v) * n + tcol[u] + 4 * olg * n + olr];
            }
            const float2 x = sS[i * P + j];
            ci[u][v] += d.x * x.x - d.y * x.y;
            cr[u][v] -= d.x * x.y + d.y * x.x;
            if (!CACHE) __builtin_amdgcn_sched_barrier(0);  // bound the loads in flight
          }
        }
        if (BRD) {
          const float2 x = sS[bi * P + bj];
          bm += dc_b.x * x.x - dc_b.y * x.y;
          br -= dc_b.x * x.y + dc_b.y * x.x;
        }
      }
      if (!DB) __syncthreads();  // every wave has finished reading S

      const float cacc = (float)((st == 0 || st == 3) ? dt / 6.0 : dt / 3.0);
      const float keep = (st == 0) ? 0.f : 1.f;   // the increment restarts at stage 0
      if (st < 3) {
        // next stage operand S = rho + c k (FP64, rounded to complex64 into LDS)
        const double cnext = (st == 2) ? dt : 0.5 * dt;
#pragma unroll
        for (int u = 0; u < BPW; ++u) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float kr = ci[u][v], ki = -cr[u][v];
            inc[u][v].x = fmaf(cacc, kr, keep * inc[u][v].x);
            inc[u][v].y = fmaf(cacc, ki, keep * inc[u][v].y);
            const float2 sv = make_float2((float)(rho[u][v].x + cnext * (double)kr),
                                          (float)(rho[u][v].y + cnext * (double)ki));
            if (tval[u]) sN[(trow[u] + 4 * lg + v) * P + tcol[u] + lr] = sv;
            if constexpr (DFOLD) {
              const float2 p = tval[u] ? cmulf(dc[u][v], sv) : make_float2(0.f, 0.f);
              cr[u][v] = -p.y;
              ci[u][v] = G3 ? p.x - p.y : p.x;
            }
          }
        }
        if (BRD) {
          const float kr = bm, ki = -br;
          inc_b.x = fmaf(cacc, kr, keep * inc_b.x);
          inc_b.y = fmaf(cacc, ki, keep * inc_b.y);
          const float2 sv = make_float2((float)(rho_b.x + cnext * (double)kr), (float)(rho_b.y + cnext * (double)ki));
          if (bwrite) sN[bi * P + bj] = sv;
          if constexpr (DFOLD) {
            const float2 p = bwrite ? cmulf(dc_b, sv) : make_float2(0.f, 0.f);
            dbr = -p.y;
            dbm = p.x;
          }
        }
        __syncthreads();  // S of the next stage is complete
        ph ^= (int)DB;
        continue;
      }
      // last stage: the state update itself is FP64
#pragma unroll
      for (int u = 0; u < BPW; ++u) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          rho[u][v].x += (double)fmaf(cacc, ci[u][v], inc[u][v].x);
          rho[u][v].y += (double)fmaf(cacc, -cr[u][v], inc[u][v].y);
        }
      }
      if (BRD) {
        rho_b.x += (double)fmaf(cacc, bm, inc_b.x);
        rho_b.y += (double)fmaf(cacc, -br, inc_b.y);
      }

      // ---------------- end of step: optional trace renormalisation, records
      if (renorm) {
        double tr = 0.0;
#pragma unroll
        for (int u = 0; u < BPW; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if (tval[u] && trow[u] == tcol[u] && 4 * lg + v == lr) tr += rho[u][v].x;
        if (BRD && bwrite && bi == bj) tr += rho_b.x;
        tr = wave_sum(tr);
        if (lane == 0) sRed[w].x = tr;
        __syncthreads();
        double tot = 0.0;
        for (int x = 0; x < W; ++x) tot += sRed[x].x;
        const double inv = 1.0 / tot;
#pragma unroll
        for (int u = 0; u < BPW; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            rho[u][v].x *= inv;
            rho[u][v].y *= inv;
          }
        rho_b.x *= inv;
        rho_b.y *= inv;
      }
#pragma unroll
      for (int u = 0; u < BPW; ++u) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float2 sv = make_float2((float)rho[u][v].x, (float)rho[u][v].y);
          if (tval[u]) sN[(trow[u] + 4 * lg + v) * P + tcol[u] + lr] = sv;
          if constexpr (DFOLD) {
            const float2 p = tval[u] ? cmulf(dc[u][v], sv) : make_float2(0.f, 0.f);
            cr[u][v] = -p.y;
            ci[u][v] = G3 ? p.x - p.y : p.x;
          }
        }
      }
      if (BRD) {
        const float2 sv = make_float2((float)rho_b.x, (float)rho_b.y);
        if (bwrite) sN[bi * P + bj] = sv;
        if constexpr (DFOLD) {
          const float2 p = bwrite ? cmulf(dc_b, sv) : make_float2(0.f, 0.f);
          dbr = -p.y;
          dbm = p.x;
        }
      }

      const int g = a.step_offset + s + 1;
      const bool rec_now = record_any && (g % a.nout) == 0 && (g / a.nout - 1) < a.rec_capacity;
      if (rec_now) {
        const int rec = g / a.nout - 1;
        if (a.obs != nullptr) {
          const double2* E0 = a.eops_t + (size_t)(b / a.eop_group) * a.n_eops * nn;
          for (int e = 0; e < a.n_eops; ++e) {
            const double2* E = E0 + (size_t)e * nn;
            double pr = 0.0, pi = 0.0;
            if constexpr (CACHE && BPW == 1) {
              // the owned elements of E^T (unless cached), loaded at clamped, always valid
              // addresses so that the loads leave together; zero outside the operator
              double2 ev[4];
              const int j = tcol[0] + lr;
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                const int i = trow[0] + 4 * lg + v;
                const bool in = tval[0] && i < n && j < n;
                const double2 x = E[in ? (trow[0] + v) * n + tcol[0] + 4 * olg * n + olr : 0];
                ev[v] = in ? x : make_double2(0.0, 0.0);
                if constexpr (ECACHE && REC) {
                  if (ecached) ev[v] = ec[0][v];   // zero outside the operator
                }
              }
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                pr += ev[v].x * rho[0][v].x - ev[v].y * rho[0][v].y;
                pi += ev[v].x * rho[0][v].y + ev[v].y * rho[0][v].x;
              }
            } else {
              // no register room: one load in flight
#pragma unroll
              for (int u = 0; u < BPW; ++u) {
                const int j = tcol[u] + lr;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                  const int i = trow[u] + 4 * lg + v;
                  if (tval[u] && i < n && j < n) {
                    double2 x;
                    if constexpr (ECACHE && REC) {
                      x = ecached ? ec[u][v] : E[(trow[u] + v) * n + tcol[u] + 4 * olg * n + olr];
                    } else {
                      x = E[(trow[u] + v) * n + tcol[u] + 4 * olg * n + olr];
                    }
                    pr += x.x * rho[u][v].x - x.y * rho[u][v].y;
                    pi += x.x * rho[u][v].y + x.y * rho[u][v].x;
                  }
                  if (!CACHE) __builtin_amdgcn_sched_barrier(0);
                }
              }
            }
            if (BRD && bwrite) {
              const double2 ev = ecached ? ec_b : E[bi * n + bj];
              pr += ev.x * rho_b.x - ev.y * rho_b.y;
              pi += ev.x * rho_b.y + ev.y * rho_b.x;
            }
            pr = wave_sum(pr);
            pi = wave_sum(pi);
            // slots W.. are free: slots 0..W-1 may still be read by the trace step
            if (lane == 0) sRed[W + e * W + w] = make_double2(pr, pi);
          }
        }
        if (a.states != nullptr) {
          double2* Sg = a.states + ((size_t)b * a.rec_capacity + rec) * nn;
#pragma unroll
          for (int u = 0; u < BPW; ++u) {
            const int j = tcol[u] + lr;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int i = trow[u] + 4 * lg + v;
              if (tval[u] && i < n && j < n) Sg[(trow[u] + v) * n + tcol[u] + 4 * olg * n + olr] = rho[u][v];
            }
          }
          if (BRD && bwrite) Sg[bi * n + bj] = rho_b;
        }
      }
      __syncthreads();  // S of the next step (and the record partials) are complete
      ph ^= (int)DB;
      // the record's final sums go to the last wave: in the bordered kernels it holds no
      // border group, so it reaches the stage barrier first and has the slack
      if (rec_now && a.obs != nullptr && w == W - 1) {
        const int rec = g / a.nout - 1;
        // the lane's record address is formed here, not hoisted into a register (or a
        // scratch spill) held across the whole step
        int e0 = lane;
        asm volatile("" : "+v"(e0));
        for (int e = e0; e < a.n_eops; e += 64) {
          double2 t = make_double2(0.0, 0.0);
          for (int x = 0; x < W; ++x) {
            const double2 y = sRed[W + e * W + x];
            t.x += y.x;
            t.y += y.y;
          }
          a.obs[((size_t)b * a.rec_capacity + rec) * a.n_eops + e] = t;
        }
      }
    }
  }

  // ---- write the state back
  {
    double2* R = a.rho + (size_t)b * nn;
#pragma unroll
    for (int u = 0; u < BPW; ++u) {
      const int j = tcol[u] + lr;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = trow[u] + 4 * lg + v;
        if (tval[u] && i < n && j < n) R[i * n + j] = rho[u][v];
      }
    }
    if (BRD && bwrite) R[bi * n + bj] = rho_b;
  }
}

// ------------------------------------------------------------------------ host launch
namespace {

// Launch shapes per padded size NB: W waves, BPW tiles per wave, MINW the waves per SIMD
// the register budget is compiled for (512 / MINW VGPRs).  The FP64 state costs 16
// doubles per tile per lane, so BPW sets the budget.  Variant 0 is the default; the
// autotuner times the others.  Bordered shapes (n = 16 (NB-1) + 1) own (NB-1)^2 tiles.
struct Shape {
  int w, bpw;
};
constexpr int kMaxVariants = 4;
constexpr Shape kShapes[7][kMaxVariants] = {
    {},
    {{1, 1}, {0, 0}, {0, 0}, {0, 0}},
    {{4, 1}, {2, 2}, {0, 0}, {0, 0}},
    {{9, 1}, {5, 2}, {0, 0}, {0, 0}},
    {{16, 1}, {8, 2}, {0, 0}, {0, 0}},
    {{13, 2}, {0, 0}, {0, 0}, {0, 0}},
    {},
};
constexpr Shape kBorderShapes[7][kMaxVariants] = {
    {},
    {},
    {{4, 1}, {2, 1}, {0, 0}, {0, 0}},
    {{4, 1}, {8, 1}, {0, 0}, {0, 0}},
    {{9, 1}, {0, 0}, {0, 0}, {0, 0}},
    {{16, 1}, {0, 0}, {0, 0}, {0, 0}},
    {},
};

template <int NB, int W, int BPW, int MINW, bool DENSE, bool REC, bool BRD, bool DB>
hipError_t launch_one(const ResidentArgs& a, size_t lds, hipStream_t stream) {
  auto k = lindblad_rk4_resident<NB, W, BPW, MINW, DENSE, REC, BRD, DB>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(a.batch), dim3(W * 64), lds, stream, a);
  return hipGetLastError();
}

template <int NB, int W, int BPW, int MINW, bool DENSE, bool BRD, bool DB>
hipError_t launch_rec2(const ResidentArgs& a, size_t lds, hipStream_t stream) {
  const bool rec = a.nout > 0 && (a.obs != nullptr || a.states != nullptr);
  return rec ? launch_one<NB, W, BPW, MINW, DENSE, true, BRD, DB>(a, lds, stream)
             : launch_one<NB, W, BPW, MINW, DENSE, false, BRD, DB>(a, lds, stream);
}

template <int NB, int W, int BPW, int MINW, bool DENSE, bool BRD>
hipError_t launch_rec(const ResidentArgs& a, size_t lds, hipStream_t stream, bool db) {
  if constexpr (!DENSE && NB <= 5) {
    if (db) return launch_rec2<NB, W, BPW, MINW, DENSE, BRD, true>(a, lds, stream);
  }
  return launch_rec2<NB, W, BPW, MINW, DENSE, BRD, false>(a, lds, stream);
}

template <int NB, int W, int BPW>
hipError_t launch_shape(const ResidentArgs& a, size_t lds, hipStream_t stream, bool db) {
  static_assert(W * 64 <= 1024, "a workgroup holds at most 16 waves");
  // two tiles per wave: 4 waves per SIMD only where the workgroup needs them (W > 12)
  constexpr int MINW = BPW == 1 ? 6 : BPW == 2 ? (W > 12 ? 4 : 3) : BPW == 3 ? 3 : 2;
  // the dense products stage T = L S and their own accumulators: two waves per SIMD fewer
  // for one tile per wave
  constexpr int MINW_D = (BPW == 1 ? MINW - 2 : MINW) > (W + 3) / 4 ? (BPW == 1 ? MINW - 2 : MINW) : (W + 3) / 4;
  if constexpr (NB <= 4) {
    if (a.n_dense > 0) return launch_rec<NB, W, BPW, MINW_D, true, false>(a, lds, stream, false);
  } else {
    if (a.n_dense > 0) return hipErrorInvalidValue;
  }
  return launch_rec<NB, W, BPW, MINW, false, false>(a, lds, stream, db);
}

// Bordered shapes: one workgroup per CU at the larger sizes, so the register budget is
// that of ceil(W / 4) waves per SIMD (at least 2 workgroups per CU below NB = 4).
template <int NB, int W, int BPW>
hipError_t launch_border(const ResidentArgs& a, size_t lds, hipStream_t stream, bool db) {
  static_assert(W * 64 <= 1024, "a workgroup holds at most 16 waves");
  constexpr int PER_SIMD = (W + 3) / 4;
  constexpr int MINW = NB <= 3 ? 2 * PER_SIMD : PER_SIMD;
  if (a.n_dense > 0 || a.n != 16 * (NB - 1) + 1) return hipErrorInvalidValue;
  return launch_rec<NB, W, BPW, MINW, false, true>(a, lds, stream, db);
}

const Shape& shape_of(int nb, int variant, bool brd) {
  return brd ? kBorderShapes[nb][variant] : kShapes[nb][variant];
}

constexpr size_t kLdsMax = 160 * 1024;

// LDS bytes of the layout: H_eff + nbuf operator buffers (NP x (NP+1) complex64 plus the
// prefetch spare rows) + the FP64 reduction slots
size_t lds_layout(int n, int variant, int n_dense, int n_eops, int nbuf) {
  const int nb = cdiv(n, 16);
  const int np = resident_bordered(n, n_dense) ? 16 * (nb - 1) + 8 : nb * 16, p = np + 1;
  const size_t buf = (size_t)np * p + 4 * p + 4;
  size_t bytes = (np * p + nbuf * buf) * sizeof(float2);
  bytes += (size_t)(1 + n_eops) * shape_of(nb, variant, resident_bordered(n, n_dense)).w * sizeof(double2);
  return bytes;
}

}  // namespace

// n <= 80 (64 with dense operators).  At 81..96 (NB = 6) the FP64 state of 3 tiles per wave
// (or 2 with the border outputs) does not fit the register budget of the 12-13 waves those
// sizes need: the builds spilled 20-92 VGPRs, so those sizes run the per-stage kernels.
int resident_nb_max(int n_dense) { return n_dense > 0 ? 4 : 5; }

bool resident_bordered(int n, int n_dense) { return n_dense == 0 && n >= 17 && n % 16 == 1 && n <= 65; }

int resident_variants(int n, int n_dense) {
  const int nb = cdiv(n, 16);
  if (nb < 1 || nb > 6) return 0;
  const bool brd = resident_bordered(n, n_dense);
  int c = 0;
  while (c < kMaxVariants && shape_of(nb, c, brd).w > 0) ++c;
  return c;
}

// H_eff + S + T + L with dense operators; without, H_eff + S double-buffered (nb <= 5)
// when that fits, else a single S
size_t resident_lds_bytes(int n, int variant, int n_dense, int n_eops) {
  if (n_dense > 0) return lds_layout(n, variant, n_dense, n_eops, 3);
  return lds_layout(n, variant, n_dense, n_eops, resident_double_buffered(n, variant, n_dense, n_eops) ? 2 : 1);
}

bool resident_double_buffered(int n, int variant, int n_dense, int n_eops) {
  return n_dense == 0 && cdiv(n, 16) <= 5 && lds_layout(n, variant, n_dense, n_eops, 2) <= kLdsMax;
}

hipError_t resident_launch(const ResidentArgs& a, int variant, hipStream_t stream) {
  if (variant < 0 || variant >= resident_variants(a.n, a.n_dense)) return hipErrorInvalidValue;
  const size_t lds = resident_lds_bytes(a.n, variant, a.n_dense, a.n_eops);
  const bool db = resident_double_buffered(a.n, variant, a.n_dense, a.n_eops);
  const int nb = cdiv(a.n, 16);
  if (resident_bordered(a.n, a.n_dense)) {
    switch (nb * 10 + variant) {
      case 20: return launch_border<2, 4, 1>(a, lds, stream, db);
      case 21: return launch_border<2, 2, 1>(a, lds, stream, db);
      case 30: return launch_border<3, 4, 1>(a, lds, stream, db);
      case 31: return launch_border<3, 8, 1>(a, lds, stream, db);
      case 40: return launch_border<4, 9, 1>(a, lds, stream, db);
      case 50: return launch_border<5, 16, 1>(a, lds, stream, db);
      default: return hipErrorInvalidValue;
    }
  }
  switch (nb * 10 + variant) {
    case 10: return launch_shape<1, 1, 1>(a, lds, stream, db);
    case 20: return launch_shape<2, 4, 1>(a, lds, stream, db);
    case 21: return launch_shape<2, 2, 2>(a, lds, stream, db);
    case 30: return launch_shape<3, 9, 1>(a, lds, stream, db);
    case 31: return launch_shape<3, 5, 2>(a, lds, stream, db);
    case 40: return launch_shape<4, 16, 1>(a, lds, stream, db);
    case 41: return launch_shape<4, 8, 2>(a, lds, stream, db);
    case 50: return launch_shape<5, 13, 2>(a, lds, stream, db);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace lime
