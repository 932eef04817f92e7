// Lindblad RK4 for operators too large to stay resident in one CU (n > 96), and the
// single-application kernel used by the adjoint.
//
// Same right-hand side as the resident kernel,
//   f(X) = -i (H_eff X - X H_eff^dag) + D o X + sum_k L_k X R_k^dag   (R_k = L_k for Lindblad),
// but one launch per RK4 stage: the kernel boundary is the grid-wide barrier between
// stages.  Each workgroup owns a TILE x TILE output tile of one operator:
//   * K-panels of 16 complex columns are staged global -> registers -> LDS (the next panel
//     is fetched into registers while the current one feeds the MFMAs);
//   * 4 waves in a 2 x 2 arrangement, each owning (TILE/32)^2 16x16 MFMA blocks;
//   * the RK4 stage update runs in the epilogue: k = -i C + D o S, the increment, the next
//     stage operand (double buffered so no tile reads a half-written operand), and at the
//     last stage rho += increment in FP64, records (each workgroup's share of Tr[E rho] to
//     its own workspace slot, summed over the slots in a fixed order by obs_reduce, so the
//     records are stored, never accumulated, and bitwise reproducible; states) and the
//     next S = rho.
// Three arithmetic variants share the code (their MFMA accumulator row layouts differ, see
// CT::row):
//   R = float,  P = 0: complex64 operands on v_mfma_f32_16x16x4_f32, FP32 increment;
//   R = float,  P = 1: complex64 operands split into BF16 pairs x = hi + lo as they are
//                      staged into LDS; every real product uses the three terms
//                      hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16 (FP32 accumulate);
//   R = double, P = 0: complex128 operands on v_mfma_f64_16x16x4_f64, FP64 everywhere
//                      (the FP64 GPU reference).
// The dense products T_k = L_k S of a stage come from lindblad_left_tiled, launched first.
#include "common.hpp"
#include "lindblad_tiled_decl.hpp"

namespace lime {

namespace {

constexpr int KW = 16;          // K-panel width of the FP32 / FP64 staging
constexpr int KWP = KW + 1;         // LDS pitch of [row][k] panels (odd: conflict-free fragments)

template <typename R> struct CT;
template <> struct CT<float> {
  using C = float2;
  using V4 = floatx4;
  __device__ static V4 mfma(float a, float b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static C make(double x, double y) { return make_float2((float)x, (float)y); }
  // accumulator register v of lane group lg holds row 4*lg + v of the 16x16 block
  __device__ static int row(int lg, int v) { return 4 * lg + v; }
};
template <> struct CT<double> {
  using C = double2;
  using V4 = doublex4;
  __device__ static V4 mfma(double a, double b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static C make(double x, double y) { return make_double2(x, y); }
  // FP64 16x16x4: register v of lane group lg holds row lg + 4*v (measured on gfx950)
  __device__ static int row(int lg, int v) { return lg + 4 * v; }
};

// Complex accumulation modes of one product term:
//   PLAIN : C += A B                     (B[k][j] given row-major)
//   NEG_AH: C -= A M^dag                 (M given row-major; B[k][j] = conj(M[j][k]))
//   I_AH  : C += i A M^dag
enum Mode { PLAIN = 0, NEG_AH = 1, I_AH = 2 };

// Load a TILE x PW panel of a row-major n x n matrix: rows r0.., cols k0..
template <int TILE, int PW, typename C, bool FAST = false>
__device__ __forceinline__ void load_rows(const C* __restrict__ M, int n, int r0, int k0,
                                          C (&reg)[TILE * PW / 256], int tid) {
  if (FAST && r0 + TILE <= n && k0 + PW <= n) {   // interior panel (uniform): no bounds tests
#pragma unroll
    for (int q = 0; q < TILE * PW / 256; ++q) {
      const int e = tid + q * 256;
      reg[q] = M[(size_t)(r0 + e / PW) * n + k0 + e % PW];
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < TILE * PW / 256; ++q) {
    const int e = tid + q * 256;
    const int r = e / PW, c = e % PW;
    const int gr = r0 + r, gc = k0 + c;
    C v;
    v.x = 0;
    v.y = 0;
    if (gr < n && gc < n) v = M[(size_t)gr * n + gc];
    reg[q] = v;
  }
}

template <int TILE, int PW, typename C>
__device__ __forceinline__ void store_rows(C* sM, const C (&reg)[TILE * PW / 256], int tid) {
#pragma unroll
  for (int q = 0; q < TILE * PW / 256; ++q) {
    const int e = tid + q * 256;
    sM[(e / PW) * (PW + 1) + e % PW] = reg[q];
  }
}

// Load a PW x TILE panel: rows k0.. (the k index), cols c0..
template <int TILE, int PW, typename C, bool FAST = false>
__device__ __forceinline__ void load_cols(const C* __restrict__ M, int n, int k0, int c0,
                                          C (&reg)[TILE * PW / 256], int tid) {
  if (FAST && k0 + PW <= n && c0 + TILE <= n) {   // interior panel (uniform): no bounds tests
#pragma unroll
    for (int q = 0; q < TILE * PW / 256; ++q) {
      const int e = tid + q * 256;
      reg[q] = M[(size_t)(k0 + e / TILE) * n + c0 + e % TILE];
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < TILE * PW / 256; ++q) {
    const int e = tid + q * 256;
    const int r = e / TILE, c = e % TILE;
    const int gr = k0 + r, gc = c0 + c;
    C v;
    v.x = 0;
    v.y = 0;
    if (gr < n && gc < n) v = M[(size_t)gr * n + gc];
    reg[q] = v;
  }
}

template <int TILE, int PW, typename C>
__device__ __forceinline__ void store_cols(C* sM, const C (&reg)[TILE * PW / 256], int tid) {
#pragma unroll
  for (int q = 0; q < TILE * PW / 256; ++q) {
    const int e = tid + q * 256;
    sM[(e / TILE) * (TILE + 1) + e % TILE] = reg[q];
  }
}

template <int TILE>
constexpr int smem_b() { return KW * (TILE + 1) > TILE * KWP ? KW * (TILE + 1) : TILE * KWP; }

// Accumulate one product term into the wave's MFMA blocks, in the three-multiplication
// (Gauss) form: three real products per complex product into
//   cr = R1, dr = R2, ci = M1, with Re C = R1 - R2 and Im C = M1 - R1 - R2
// (the conversion runs once, after the last term); per mode, with the stored B element h:
//   PLAIN  (C += a h)         R1 += ax hx,   R2 += ay hy,  M1 += (ax + ay)(hx + hy)
//   NEG_AH (C -= a conj(h))   R1 += -ax hx,  R2 += ay hy,  M1 += (ax + ay)(hy - hx)
//   I_AH   (C += i a conj(h)) R1 += ax hy,   R2 += ay hx,  M1 += (ax + ay)(hx + hy)
template <int TILE, int MODE, typename R, bool FAST = false>
__device__ __forceinline__ void gemm_term(const typename CT<R>::C* __restrict__ A,
                                          const typename CT<R>::C* __restrict__ Bsrc, int n, int row0,
                                          int col0, typename CT<R>::C* sA, typename CT<R>::C* sB,
                                          typename CT<R>::V4 (&cr)[TILE / 32][TILE / 32],
                                          typename CT<R>::V4 (&ci)[TILE / 32][TILE / 32],
                                          typename CT<R>::V4 (&dr)[TILE / 32][TILE / 32], int tid, int wr,
                                          int wc, int lr, int lg, int kb, int ke) {
  using C = typename CT<R>::C;
  constexpr int BW = TILE / 32;
  constexpr int NR = TILE * KW / 256;
  // 16x16 blocks entirely outside the operator hold zeros: their MFMAs are skipped
  // (wave-uniform), which frees the matrix pipe for the other workgroups of the CU
  bool rlive[BW], clive[BW];
#pragma unroll
  for (int p = 0; p < BW; ++p) {
    rlive[p] = row0 + wr * (TILE / 2) + p * 16 < n;
    clive[p] = col0 + wc * (TILE / 2) + p * 16 < n;
  }
  // the reduction runs over k in [kb, ke): kb a multiple of the panel width, ke = n or a
  // multiple of it (a K-split part), so no panel straddles two parts
  C ra[NR], rb[NR];
  load_rows<TILE, KW, C, FAST>(A, n, row0, kb, ra, tid);
  if (MODE == PLAIN) load_cols<TILE, KW, C, FAST>(Bsrc, n, kb, col0, rb, tid);
  else load_rows<TILE, KW, C, FAST>(Bsrc, n, col0, kb, rb, tid);
  for (int k0 = kb; k0 < ke; k0 += KW) {
    __syncthreads();                          // previous panel fully consumed
    store_rows<TILE, KW>(sA, ra, tid);
    if (MODE == PLAIN) store_cols<TILE, KW>(sB, rb, tid);
    else store_rows<TILE, KW>(sB, rb, tid);
    __syncthreads();
    if (k0 + KW < ke) {                       // prefetch the next panel into registers
      load_rows<TILE, KW, C, FAST>(A, n, row0, k0 + KW, ra, tid);
      if (MODE == PLAIN) load_cols<TILE, KW, C, FAST>(Bsrc, n, k0 + KW, col0, rb, tid);
      else load_rows<TILE, KW, C, FAST>(Bsrc, n, col0, k0 + KW, rb, tid);
    }
#pragma unroll
    for (int kk = 0; kk < KW; kk += 4) {
      if (k0 + kk >= n) break;                // the last panel: k-steps past n are zeros
      C x[BW], y[BW];
#pragma unroll
      for (int p = 0; p < BW; ++p) {
        x[p] = sA[(wr * (TILE / 2) + p * 16 + lr) * KWP + kk + lg];
        if (MODE == PLAIN) y[p] = sB[(kk + lg) * (TILE + 1) + wc * (TILE / 2) + p * 16 + lr];
        else y[p] = sB[(wc * (TILE / 2) + p * 16 + lr) * KWP + kk + lg];
      }
#pragma unroll
      for (int p = 0; p < BW; ++p)
#pragma unroll
        for (int q = 0; q < BW; ++q) {
          if (!(rlive[p] && clive[q])) continue;
          const C a = x[p], h = y[q];
          const R as = a.x + a.y;
          if (MODE == PLAIN) {
            cr[p][q] = CT<R>::mfma(a.x, h.x, cr[p][q]);
            dr[p][q] = CT<R>::mfma(a.y, h.y, dr[p][q]);
            ci[p][q] = CT<R>::mfma(as, h.x + h.y, ci[p][q]);
          } else if (MODE == NEG_AH) {
            cr[p][q] = CT<R>::mfma(-a.x, h.x, cr[p][q]);
            dr[p][q] = CT<R>::mfma(a.y, h.y, dr[p][q]);
            ci[p][q] = CT<R>::mfma(as, h.y - h.x, ci[p][q]);
          } else {
            cr[p][q] = CT<R>::mfma(a.x, h.y, cr[p][q]);
            dr[p][q] = CT<R>::mfma(a.y, h.x, dr[p][q]);
            ci[p][q] = CT<R>::mfma(as, h.x + h.y, ci[p][q]);
          }
        }
    }
  }
}

// ---- 3xBF16 split products
// Every real operand is split x = hi + lo (v_cvt_pk_bf16_f32, round to nearest even; x - hi
// is exact in FP32), each part in its own plane (Re hi, Re lo, Im hi, Im lo): the constant
// H_eff, L_k, R_k once per call into workspace planes (split_planes), T_k by the 32-tile
// left kernel that produces it, the stage operand (and the 64-tiles' T_k) while it is staged
// into LDS.  K-panels of KB = 32 complex elements (one 16x16x32 MFMA k-step).  A rows,
// and the B operand of the conjugate-transpose modes (rows j of M), are planes [row][k]:
// one 16-byte read per fragment.  The PLAIN B operand (rows k of M) is stored as planes
// [k][column] and its fragments (8 consecutive k of one column) come through
// ds_read_b64_tr_b16, whose result lane i, element e is element i % 4 of what lane
// 4 e + i / 4 of the 16-lane group loaded.  Each real product X Y is Xlo Yhi + Xhi Ylo +
// Xhi Yhi in FP32 (the lo*lo term, <= 2^-16 |X Y|, is dropped); 4 real products per complex
// product into Re+ (cr), Re- (dr) and Im (ci) accumulators, with the conjugate-transpose
// mode's sign staged into its B planes:
//   PLAIN  (C += a b)          B = (bx, by):   Re += ax bx - ay by,  Im += ax by + ay bx
//   NEG_AH (C -= a conj(h))    B = (-hx, hy):  Re += ax (-hx) - ay hy,  Im += ax hy + ay (-hx)
//   I_AH   (C += i a conj(h))  B = (hx, hy):   Re += ax hy - ay hx,  Im += ax hx + ay hy
constexpr int KB = 32;          // K-panel width (complex elements) of the 3xBF16 staging
constexpr int PP = KB + 16;     // BF16 pitch of [row][k] planes: 96-byte rows, so the 16-byte
                                // fragment reads of 8 consecutive rows hit distinct banks and
                                // the staging stores of 4 rows overlap on fewer banks than at
                                // 80 bytes (dense512: +3.5 %; 64-byte rows with an XOR
                                // swizzle: +0.6 %); the 64-tile's 48 KB still fits 3 per CU

typedef unsigned short u16;
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));
typedef short short4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) short4v lds_short4;

// BF16 pitch of the PLAIN B planes [k][column]
template <int TILE>
constexpr int pb_pitch() { return TILE + 8; }

// h = hi(a) | hi(b) << 16, l = lo(a) | lo(b) << 16 with x = hi + lo
__device__ __forceinline__ void split2(float a, float b, unsigned& h, unsigned& l) {
  const float2v x = {a, b};
  const bf16x2 hb = __builtin_convertvector(x, bf16x2);
  const float2v hf = __builtin_convertvector(hb, float2v);
  const bf16x2 lb = __builtin_convertvector(x - hf, bf16x2);
  h = __builtin_bit_cast(unsigned, hb);
  l = __builtin_bit_cast(unsigned, lb);
}

__device__ __forceinline__ floatx4 mfma_bf(u16x8 a, u16x8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                 0, 0, 0);
}

// c += X Y for split operands (small terms first)
__device__ __forceinline__ floatx4 mfma3(u16x8 xh, u16x8 xl, u16x8 yh, u16x8 yl, floatx4 c) {
  c = mfma_bf(xl, yh, c);
  c = mfma_bf(xh, yl, c);
  return mfma_bf(xh, yh, c);
}

// Stage element pair (e0, e1) (consecutive along the plane row) at plane offset o; NEG
// negates the real parts.
template <int PLANE, bool NEG>
__device__ __forceinline__ void put_pair(u16* sM, int o, float2v e0, float2v e1) {
  unsigned xh, xl, yh, yl;
  split2(e0.x, e1.x, xh, xl);
  split2(e0.y, e1.y, yh, yl);
  if (NEG) {
    xh ^= 0x80008000u;
    xl ^= 0x80008000u;
  }
  *reinterpret_cast<unsigned*>(sM + o) = xh;
  *reinterpret_cast<unsigned*>(sM + PLANE + o) = xl;
  *reinterpret_cast<unsigned*>(sM + 2 * PLANE + o) = yh;
  *reinterpret_cast<unsigned*>(sM + 3 * PLANE + o) = yl;
}

// Element pairs of a ROWS x COLS panel of a row-major n x n matrix (rows r0.., cols c0..),
// NP pairs per thread, pair q = tid + 256 i along the columns.
template <int ROWS, int COLS>
struct PanelPairs {
  static constexpr int NP = ROWS * COLS / 2 / 256;
};
template <int ROWS, int COLS>
__device__ __forceinline__ void load_pairs(const float2* __restrict__ M, int n, int r0, int c0,
                                           float2v (&reg)[2 * PanelPairs<ROWS, COLS>::NP], int tid) {
  const float2v* __restrict__ Mv = reinterpret_cast<const float2v*>(M);
  if (r0 + ROWS <= n && c0 + COLS <= n) {     // interior panel (uniform): no bounds tests
#pragma unroll
    for (int i = 0; i < PanelPairs<ROWS, COLS>::NP; ++i) {
      const int q = tid + 256 * i;
      const int r = q / (COLS / 2), c = 2 * (q % (COLS / 2));
      const float2v* p = Mv + (size_t)(r0 + r) * n + c0 + c;
      reg[2 * i] = p[0];
      reg[2 * i + 1] = p[1];
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < PanelPairs<ROWS, COLS>::NP; ++i) {
    const int q = tid + 256 * i;
    const int r = q / (COLS / 2), c = 2 * (q % (COLS / 2));
    const int gr = r0 + r, gc = c0 + c;
    float2v v0 = {0.f, 0.f}, v1 = {0.f, 0.f};
    if (gr < n) {
      if (gc < n) v0 = Mv[(size_t)gr * n + gc];
      if (gc + 1 < n) v1 = Mv[(size_t)gr * n + gc + 1];
    }
    reg[2 * i] = v0;
    reg[2 * i + 1] = v1;
  }
}
template <int ROWS, int COLS, int PITCH, bool NEG>
__device__ __forceinline__ void store_pairs(u16* sM, const float2v (&reg)[2 * PanelPairs<ROWS, COLS>::NP], int tid) {
#pragma unroll
  for (int i = 0; i < PanelPairs<ROWS, COLS>::NP; ++i) {
    const int q = tid + 256 * i;
    const int r = q / (COLS / 2), c = 2 * (q % (COLS / 2));
    put_pair<ROWS * PITCH, NEG>(sM, r * PITCH + c, reg[2 * i], reg[2 * i + 1]);
  }
}

// Pre-split operands (the constant H_eff, L_k, R_k): four BF16 planes of n rows x ldp
// (n rounded up to 8, zero beyond n), so a ROWS x KB panel of [row][k] planes is staged as
// 16-byte chunks (8 consecutive k of one row of one plane), NC chunks per thread.
template <int ROWS>
struct PanelChunks {
  static constexpr int PER_PLANE = ROWS * KB / 8;
  static constexpr int NC = 4 * PER_PLANE / 256;
};
template <int ROWS>
__device__ __forceinline__ void load_chunks(const u16* __restrict__ M, int n, int ldp, int r0, int k0,
                                            u16x8 (&reg)[PanelChunks<ROWS>::NC], int tid) {
  const size_t plane = (size_t)n * ldp;
#pragma unroll
  for (int i = 0; i < PanelChunks<ROWS>::NC; ++i) {
    const int q = tid + 256 * i;
    const int pl = q / PanelChunks<ROWS>::PER_PLANE, e = q % PanelChunks<ROWS>::PER_PLANE;
    const int gr = r0 + e / (KB / 8), gk = k0 + 8 * (e % (KB / 8));
    u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
    if (gr < n && gk < ldp) v = *reinterpret_cast<const u16x8*>(M + pl * plane + (size_t)gr * ldp + gk);
    reg[i] = v;
  }
}
// NEG negates the real parts (planes 0 and 1)
template <int ROWS, bool NEG>
__device__ __forceinline__ void store_chunks(u16* sM, const u16x8 (&reg)[PanelChunks<ROWS>::NC], int tid) {
#pragma unroll
  for (int i = 0; i < PanelChunks<ROWS>::NC; ++i) {
    const int q = tid + 256 * i;
    const int pl = q / PanelChunks<ROWS>::PER_PLANE, e = q % PanelChunks<ROWS>::PER_PLANE;
    u16x8 v = reg[i];
    if (NEG && pl < 2) v ^= (u16)0x8000;
    *reinterpret_cast<u16x8*>(sM + pl * ROWS * PP + (e / (KB / 8)) * PP + 8 * (e % (KB / 8))) = v;
  }
}

// One ROWS x KB [row][k] panel held in registers between its load and its LDS store: from a
// complex64 operand (split while staged) or from pre-split planes (PRE).
template <int ROWS, bool PRE>
struct RowPanel {
  float2v r[2 * PanelPairs<ROWS, KB>::NP];
  __device__ __forceinline__ void load(const void* M, int n, int, int r0, int k0, int tid) {
    load_pairs<ROWS, KB>(static_cast<const float2*>(M), n, r0, k0, r, tid);
  }
  template <bool NEG>
  __device__ __forceinline__ void store(u16* sM, int tid) const {
    store_pairs<ROWS, KB, PP, NEG>(sM, r, tid);
  }
};
template <int ROWS>
struct RowPanel<ROWS, true> {
  u16x8 r[PanelChunks<ROWS>::NC];
  __device__ __forceinline__ void load(const void* M, int n, int ldp, int r0, int k0, int tid) {
    load_chunks<ROWS>(static_cast<const u16*>(M), n, ldp, r0, k0, r, tid);
  }
  template <bool NEG>
  __device__ __forceinline__ void store(u16* sM, int tid) const {
    store_chunks<ROWS, NEG>(sM, r, tid);
  }
};

// APS / BPS: the A operand / the B operand of a conjugate-transpose mode comes pre-split
// (planes with row pitch ldp); otherwise complex64, split while staged.
template <int TILE, int MODE, bool APS, bool BPS>
__device__ __forceinline__ void gemm_term_bf3(const void* __restrict__ A, const void* __restrict__ Bsrc, int n, int ldp,
                                              int row0, int col0, unsigned char* smem,
                                              floatx4 (&cr)[TILE / 32][TILE / 32],
                                              floatx4 (&ci)[TILE / 32][TILE / 32],
                                              floatx4 (&dr)[TILE / 32][TILE / 32], int tid, int wr, int wc,
                                              int lr, int lg, int kb, int ke) {
  static_assert(!(BPS && MODE == PLAIN), "the PLAIN B operand is staged as [k][column] planes");
  constexpr int BW = TILE / 32;
  constexpr int PB = pb_pitch<TILE>();
  constexpr int APLANE = TILE * PP;
  bool rlive[BW], clive[BW];
#pragma unroll
  for (int p = 0; p < BW; ++p) {
    rlive[p] = row0 + wr * (TILE / 2) + p * 16 < n;
    clive[p] = col0 + wc * (TILE / 2) + p * 16 < n;
  }
  u16* sA = reinterpret_cast<u16*>(smem);
  u16* sB = sA + 4 * APLANE;
  constexpr int BPLANE = MODE == PLAIN ? KB * PB : TILE * PP;
  // PF: the next panel is loaded into registers while the current one feeds the MFMAs.
  // Measured at n = 512: +11 % for 32 x 32 tiles; for 64 x 64 tiles the 32 prefetch VGPRs
  // cost a wave per SIMD (3 -> 2) and loading each panel when it is staged is 8 % faster.
  constexpr bool PF = TILE == 32;
  RowPanel<TILE, APS> ra;
  RowPanel<TILE, BPS> rb;                     // B rows (conjugate-transpose modes)
  float2v rp[2 * PanelPairs<KB, TILE>::NP];   // B = KB rows of TILE columns (PLAIN)
  auto load_panel = [&](int k) {
    ra.load(A, n, ldp, row0, k, tid);
    if constexpr (MODE == PLAIN) load_pairs<KB, TILE>(static_cast<const float2*>(Bsrc), n, k, col0, rp, tid);
    else rb.load(Bsrc, n, ldp, col0, k, tid);
  };
  if constexpr (PF) load_panel(kb);
  for (int k0 = kb; k0 < ke; k0 += KB) {
    if constexpr (!PF) load_panel(k0);
    __syncthreads();                          // previous panel fully consumed
    ra.template store<false>(sA, tid);
    if constexpr (MODE == PLAIN) store_pairs<KB, TILE, PB, false>(sB, rp, tid);
    else rb.template store<MODE == NEG_AH>(sB, tid);
    __syncthreads();
    if constexpr (PF) {
      if (k0 + KB < ke) load_panel(k0 + KB);  // prefetch the next panel into registers
    }
    // fragment reads and MFMAs at raised wave priority: a wave in its products wins issue
    // over waves staging the next panel (dense512: +2.5 %; the FP32 terms measured 0.3 %
    // slower with it)
    __builtin_amdgcn_s_setprio(1);
    u16x8 a0h[BW], a0l[BW], a1h[BW], a1l[BW], b0h[BW], b0l[BW], b1h[BW], b1l[BW];
#pragma unroll
    for (int p = 0; p < BW; ++p) {
      const int o = (wr * (TILE / 2) + p * 16 + lr) * PP + 8 * lg;
      a0h[p] = *reinterpret_cast<const u16x8*>(sA + o);
      a0l[p] = *reinterpret_cast<const u16x8*>(sA + APLANE + o);
      a1h[p] = *reinterpret_cast<const u16x8*>(sA + 2 * APLANE + o);
      a1l[p] = *reinterpret_cast<const u16x8*>(sA + 3 * APLANE + o);
    }
#pragma unroll
    for (int q = 0; q < BW; ++q) {
      if constexpr (MODE == PLAIN) {
        // transposed reads: lane (lg, lr) loads plane row 8 lg + lr / 4 (+ 4), columns
        // c + 4 (lr % 4) .. + 3 of the block's 16 columns
        const int o = (8 * lg + (lr >> 2)) * PB + wc * (TILE / 2) + q * 16 + 4 * (lr & 3);
        u16x8 f[4];
#pragma unroll
        for (int pl = 0; pl < 4; ++pl) {
          const u16* base = sB + pl * BPLANE + o;
          const short4v x = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_short4*)(base));
          const short4v y = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_short4*)(base + 4 * PB));
          f[pl] = u16x8{(u16)x[0], (u16)x[1], (u16)x[2], (u16)x[3], (u16)y[0], (u16)y[1], (u16)y[2], (u16)y[3]};
        }
        b0h[q] = f[0];
        b0l[q] = f[1];
        b1h[q] = f[2];
        b1l[q] = f[3];
      } else {
        const int o = (wc * (TILE / 2) + q * 16 + lr) * PP + 8 * lg;
        b0h[q] = *reinterpret_cast<const u16x8*>(sB + o);
        b0l[q] = *reinterpret_cast<const u16x8*>(sB + BPLANE + o);
        b1h[q] = *reinterpret_cast<const u16x8*>(sB + 2 * BPLANE + o);
        b1l[q] = *reinterpret_cast<const u16x8*>(sB + 3 * BPLANE + o);
      }
    }
#pragma unroll
    for (int p = 0; p < BW; ++p)
#pragma unroll
      for (int q = 0; q < BW; ++q) {
        if (!(rlive[p] && clive[q])) continue;
        if constexpr (MODE == I_AH) {
          cr[p][q] = mfma3(a0h[p], a0l[p], b1h[q], b1l[q], cr[p][q]);
          dr[p][q] = mfma3(a1h[p], a1l[p], b0h[q], b0l[q], dr[p][q]);
          ci[p][q] = mfma3(a0h[p], a0l[p], b0h[q], b0l[q], ci[p][q]);
          ci[p][q] = mfma3(a1h[p], a1l[p], b1h[q], b1l[q], ci[p][q]);
        } else {
          cr[p][q] = mfma3(a0h[p], a0l[p], b0h[q], b0l[q], cr[p][q]);
          dr[p][q] = mfma3(a1h[p], a1l[p], b1h[q], b1l[q], dr[p][q]);
          ci[p][q] = mfma3(a0h[p], a0l[p], b1h[q], b1l[q], ci[p][q]);
          ci[p][q] = mfma3(a1h[p], a1l[p], b0h[q], b0l[q], ci[p][q]);
        }
      }
    __builtin_amdgcn_s_setprio(0);
  }
}

// Waves per SIMD the compiler must leave room for: the 64 x 64 3xBF16 kernels would take
// more than 168 VGPRs with the interior-panel load path and drop to 2 waves per SIMD.
template <int TILE, int P>
constexpr int kMinWaves = (TILE == 64 && P == 1) ? 3 : 1;

// LDS bytes of a GEMM-staging area.
template <int TILE, typename R, int P>
constexpr int smem_bytes() {
  return P == 1 ? (4 * TILE * PP + 4 * (KB * pb_pitch<TILE>() > TILE * PP ? KB * pb_pitch<TILE>() : TILE * PP)) * 2
                : (TILE * KWP + smem_b<TILE>()) * (int)sizeof(typename CT<R>::C);
}

// One product term in the selected arithmetic.
// T_k = L_k S between the two kernels of a stage: 3xBF16 32-tiles write it as split planes
// (dense512 batch 1: +16 %); the 64-tiles keep complex64 (their 4x more 2-byte plane stores
// cost 2.6 % at batch 32).
template <int TILE, int P>
constexpr bool tk_split() { return P == 1 && TILE == 32; }

// (3xBF16: APS / BPS mark pre-split operands, with row pitch ldp.)
template <int TILE, int MODE, typename R, int P, bool APS = false, bool BPS = false, bool RAG = false>
__device__ __forceinline__ void term(const void* __restrict__ A, const void* __restrict__ B, int n, int ldp, int row0,
                                     int col0, unsigned char* smem,
                                     typename CT<R>::V4 (&cr)[TILE / 32][TILE / 32],
                                     typename CT<R>::V4 (&ci)[TILE / 32][TILE / 32],
                                     typename CT<R>::V4 (&dr)[TILE / 32][TILE / 32], int tid, int wr, int wc,
                                     int lr, int lg, int kb, int ke) {
  if constexpr (P == 1) {
    gemm_term_bf3<TILE, MODE, APS, BPS>(A, B, n, ldp, row0, col0, smem, cr, ci, dr, tid, wr, wc, lr, lg, kb, ke);
  } else {
    using C = typename CT<R>::C;
    C* sA = reinterpret_cast<C*>(smem);
    // interior panels without bounds tests for the per-stage 64-tiles (dense512 FP32: +3.5 %);
    // the 32-tiles (n = 129 gradient forward: 3.6 % slower) and the persistent ragged kernel
    // (9.5 % slower) keep the bounds-tested loads
    gemm_term<TILE, MODE, R, !RAG && TILE == 64>(static_cast<const C*>(A), static_cast<const C*>(B), n, row0, col0, sA,
                             sA + TILE * KWP, cr, ci, dr, tid, wr, wc, lr, lg, kb, ke);
  }
}

// (R1, R2, M1) of the FP32 / FP64 panels -> (Re C, Im C) in cr, ci; the 3xBF16 panels
// leave (Re+, Im, Re-) -> Re C = Re+ - Re-.
template <int TILE, typename R, int P>
__device__ __forceinline__ void finish_terms(typename CT<R>::V4 (&cr)[TILE / 32][TILE / 32],
                                             typename CT<R>::V4 (&ci)[TILE / 32][TILE / 32],
                                             const typename CT<R>::V4 (&dr)[TILE / 32][TILE / 32]) {
  if constexpr (P == 0) {
#pragma unroll
    for (int p = 0; p < TILE / 32; ++p)
#pragma unroll
      for (int q = 0; q < TILE / 32; ++q) {
        ci[p][q] -= cr[p][q] + dr[p][q];
        cr[p][q] -= dr[p][q];
      }
  } else {
#pragma unroll
    for (int p = 0; p < TILE / 32; ++p)
#pragma unroll
      for (int q = 0; q < TILE / 32; ++q) cr[p][q] -= dr[p][q];
  }
}

__device__ __forceinline__ int tiles_per_dim(int n, int tile) { return (n + tile - 1) / tile; }

// Common per-workgroup geometry.
struct Geo {
  int row0, col0, tid, lane, w, wr, wc, lr, lg;
};

// tid: the thread's index within its group of 256 threads (the workgroup, or one tile
// group of the persistent ragged kernel).
template <int TILE>
__device__ __forceinline__ Geo geometry(int n, int tile_index, int tid) {
  Geo g;
  const int tpd = tiles_per_dim(n, TILE);
  g.row0 = (tile_index / tpd) * TILE;
  g.col0 = (tile_index % tpd) * TILE;
  g.tid = tid;
  g.lane = g.tid & 63;
  g.w = __builtin_amdgcn_readfirstlane(g.tid >> 6);
  g.wr = g.w >> 1;
  g.wc = g.w & 1;
  g.lr = g.lane & 15;
  g.lg = g.lane >> 4;
  return g;
}

// C = H S - S H^dag + i sum_k T_k R_k^dag for operator b (set = b / group), the reductions
// over k in [kb, ke) (a K-split part; all of them: kb = 0, ke = n).
template <int TILE, typename R, int P, bool RAG = false>
__device__ __forceinline__ void generator_gemms(const TiledArgs& a, int b, const Geo& g, unsigned char* smem,
                                                typename CT<R>::V4 (&cr)[TILE / 32][TILE / 32],
                                                typename CT<R>::V4 (&ci)[TILE / 32][TILE / 32], int kb, int ke) {
  using C = typename CT<R>::C;
  const int n = a.n, set = b / a.group;
  const size_t nn = (size_t)n * n;
  const C* H = static_cast<const C*>(a.heff) + (size_t)set * nn;
  const C* S = static_cast<const C*>(a.s_cur) + (size_t)b * nn;
  using V4 = typename CT<R>::V4;
  V4 dr[TILE / 32][TILE / 32];
#pragma unroll
  for (int p = 0; p < TILE / 32; ++p)
#pragma unroll
    for (int q = 0; q < TILE / 32; ++q) dr[p][q] = V4{0, 0, 0, 0};
  // 3xBF16: H_eff, R_k (pre-split once per call) and, for 32-tiles, T_k (written split)
  // from planes
  constexpr bool PS = P == 1;
  const size_t np = (size_t)4 * n * a.ldp;   // u16 elements of one pre-split matrix
  const void* Hop = PS ? static_cast<const void*>(a.hsp + (size_t)set * np) : static_cast<const void*>(H);
  term<TILE, PLAIN, R, P, PS, false, RAG>(Hop, S, n, a.ldp, g.row0, g.col0, smem, cr, ci, dr, g.tid, g.wr, g.wc, g.lr,
                                          g.lg, kb, ke);
  // one-sided generators (dX/dt = G X, H_eff = i G) have no X H_eff^dag product
  if (!(a.flags & kFlagOneSided))
    term<TILE, NEG_AH, R, P, false, PS, RAG>(S, Hop, n, a.ldp, g.row0, g.col0, smem, cr, ci, dr, g.tid, g.wr, g.wc,
                                             g.lr, g.lg, kb, ke);
  for (int k = 0; k < a.n_dense; ++k) {
    constexpr bool TS = tk_split<TILE, P>();
    const size_t tk = (size_t)b * a.n_dense + k;
    const void* T = TS ? static_cast<const void*>(static_cast<const u16*>(a.tbuf) + tk * np)
                       : static_cast<const void*>(static_cast<const C*>(a.tbuf) + tk * nn);
    const size_t rk = (size_t)set * a.n_dense + k;
    const void* Rk = PS ? static_cast<const void*>((a.rsp ? a.rsp : a.lsp) + rk * np)
                        : static_cast<const void*>(static_cast<const C*>(a.rops ? a.rops : a.lops) + rk * nn);
    term<TILE, I_AH, R, P, TS, PS, RAG>(T, Rk, n, a.ldp, g.row0, g.col0, smem, cr, ci, dr, g.tid, g.wr, g.wc, g.lr,
                                        g.lg, kb, ke);
  }
  finish_terms<TILE, R, P>(cr, ci, dr);
}

}  // namespace

// One TILE x TILE tile of T[b][k] = L[set][k] @ S[b] (set = b / group).
template <int TILE, typename R, int P, bool RAG = false>
__device__ __forceinline__ void left_tile(const TiledArgs& a, int b, int k, const Geo& g, unsigned char* smem) {
  using C = typename CT<R>::C;
  using V4 = typename CT<R>::V4;
  const int set = b / a.group;
  const size_t nn = (size_t)a.n * a.n;
  constexpr int BW = TILE / 32;
  V4 cr[BW][BW], ci[BW][BW], dr[BW][BW];
#pragma unroll
  for (int p = 0; p < BW; ++p)
#pragma unroll
    for (int q = 0; q < BW; ++q) cr[p][q] = ci[p][q] = dr[p][q] = V4{0, 0, 0, 0};
  // 3xBF16: L_k from its pre-split planes
  constexpr bool PS = P == 1;
  const size_t lk = (size_t)set * a.n_dense + k;
  const void* Lk = PS ? static_cast<const void*>(a.lsp + lk * 4 * a.n * a.ldp)
                      : static_cast<const void*>(static_cast<const C*>(a.lops) + lk * nn);
  term<TILE, PLAIN, R, P, PS, false, RAG>(Lk, static_cast<const C*>(a.s_cur) + (size_t)b * nn, a.n, a.ldp, g.row0,
                                     g.col0, smem, cr, ci, dr, g.tid, g.wr, g.wc, g.lr, g.lg, 0, a.n);
  finish_terms<TILE, R, P>(cr, ci, dr);
  if constexpr (tk_split<TILE, P>()) {
    // 3xBF16, 32-tiles: T_k goes out as split planes (row pitch ldp; the columns n .. ldp - 1
    // hold the products with the zero columns beyond n, i.e. zeros)
    const size_t plane = (size_t)a.n * a.ldp;
    u16* T = static_cast<u16*>(a.tbuf) + ((size_t)b * a.n_dense + k) * 4 * plane;
#pragma unroll
    for (int p = 0; p < BW; ++p)
#pragma unroll
      for (int q = 0; q < BW; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = g.row0 + g.wr * (TILE / 2) + p * 16 + CT<R>::row(g.lg, v);
          const int j = g.col0 + g.wc * (TILE / 2) + q * 16 + g.lr;
          if (i < a.n && j < a.ldp) {
            unsigned h, l;
            split2(cr[p][q][v], ci[p][q][v], h, l);
            u16* t = T + (size_t)i * a.ldp + j;
            t[0] = (u16)h;
            t[plane] = (u16)l;
            t[2 * plane] = (u16)(h >> 16);
            t[3 * plane] = (u16)(l >> 16);
          }
        }
  } else {
    C* T = static_cast<C*>(a.tbuf) + ((size_t)b * a.n_dense + k) * nn;
#pragma unroll
    for (int p = 0; p < BW; ++p)
#pragma unroll
      for (int q = 0; q < BW; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = g.row0 + g.wr * (TILE / 2) + p * 16 + CT<R>::row(g.lg, v);
          const int j = g.col0 + g.wc * (TILE / 2) + q * 16 + g.lr;
          if (i < a.n && j < a.n) T[(size_t)i * a.n + j] = CT<R>::make(cr[p][q][v], ci[p][q][v]);
        }
  }
}

// T[b][k] = L[set][k] @ S[b]  for every dense operator k.
template <int TILE, typename R, int P>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kMinWaves<TILE, P>))) lindblad_left_tiled(TiledArgs a) {
  __shared__ __align__(16) unsigned char smem[smem_bytes<TILE, R, P>()];
  const int per = tiles_per_dim(a.n, TILE) * tiles_per_dim(a.n, TILE);
  const int blk = xcd_block();
  const int bk = blk / per;                   // = b * n_dense + k
  const Geo g = geometry<TILE>(a.n, blk % per, threadIdx.x);
  left_tile<TILE, R, P>(a, bk / a.n_dense, bk % a.n_dense, g, smem);
}

// ---- per-element epilogues shared by the MFMA tiles and the border workgroups

// RK4 stage update of element o of operator b from C = (cre, cim) (k = -i C + D o S):
// the FP32/FP64 increment, the next stage operand, and at the last stage rho += increment
// and the records.  Returns the final rho of the element at the last stage (else 0).
template <typename R>
__device__ __forceinline__ double2 stage_elem(const TiledArgs& a, int b, int set, size_t o, R cre, R cim) {
  using C = typename CT<R>::C;
  const size_t nn = (size_t)a.n * a.n;
  const int st = a.stage;
  const double dt = a.dt;
  const R cacc = (R)((st == 0 || st == 3) ? dt / 6.0 : dt / 3.0);
  const double cnext = (st == 2) ? dt : 0.5 * dt;
  const size_t ob = (size_t)b * nn + o;
  R kr = cim, ki = -cre;
  if (a.dmat) {
    const C d = static_cast<const C*>(a.dmat)[(size_t)set * nn + o];
    const C x = static_cast<const C*>(a.s_cur)[ob];
    kr += d.x * x.x - d.y * x.y;
    ki += d.x * x.y + d.y * x.x;
  }
  C* inc = static_cast<C*>(a.inc);
  C* Sn = static_cast<C*>(a.s_next);
  C ic;
  if (st == 0) {
    ic.x = 0;
    ic.y = 0;
  } else {
    ic = inc[ob];
  }
  ic.x += cacc * kr;
  ic.y += cacc * ki;
  double2 r = a.rho[ob];
  if (st < 3) {
    inc[ob] = ic;
    const double sx = r.x + cnext * (double)kr, sy = r.y + cnext * (double)ki;
    Sn[ob] = CT<R>::make(sx, sy);
    if (a.stage_rec && a.rec >= 0)        // FP64 stage operand of this recorded step
      a.stage_rec[(((size_t)a.rec * 3 + st) * a.batch + b) * nn + o] = make_double2(sx, sy);
    return make_double2(0.0, 0.0);
  }
  r.x += (double)ic.x;
  r.y += (double)ic.y;
  a.rho[ob] = r;
  Sn[ob] = CT<R>::make(r.x, r.y);
  if (a.rec >= 0 && a.states && !(a.flags & 1)) a.states[((size_t)b * a.rec_capacity + a.rec) * nn + o] = r;
  return r;
}

// Generator application epilogue of element o: V = -i C + D o X; y = c_lx V + c_z z,
// y_op = y in the operand precision, acc += c_acc V.
template <typename R>
__device__ __forceinline__ void apply_elem(const TiledArgs& a, int b, int set, size_t o, R cre, R cim) {
  using C = typename CT<R>::C;
  const size_t nn = (size_t)a.n * a.n;
  const size_t ob = (size_t)b * nn + o;
  double vr = (double)cim, vi = -(double)cre;
  if (a.dmat) {
    const C d = static_cast<const C*>(a.dmat)[(size_t)set * nn + o];
    const C x = static_cast<const C*>(a.s_cur)[ob];
    vr += (double)(d.x * x.x - d.y * x.y);
    vi += (double)(d.x * x.y + d.y * x.x);
  }
  C* yop = static_cast<C*>(a.y64);
  if (a.y || yop) {
    double yr = a.c_lx.x * vr - a.c_lx.y * vi;
    double yi = a.c_lx.x * vi + a.c_lx.y * vr;
    if (a.z) {
      const double2 zz = a.z[ob];
      yr += a.c_z.x * zz.x - a.c_z.y * zz.y;
      yi += a.c_z.x * zz.y + a.c_z.y * zz.x;
    }
    if (a.y) a.y[ob] = make_double2(yr, yi);
    if (yop) yop[ob] = CT<R>::make(yr, yi);
  }
  if (a.acc) {
    double2 t = a.acc[ob];
    t.x += a.c_acc.x * vr - a.c_acc.y * vi;
    t.y += a.c_acc.x * vi + a.c_acc.y * vr;
    a.acc[ob] = t;
  }
}

// Stores the workgroup's share (pr, pi: this lane's part) of Tr[E_e rho] at this step in
// the workgroup's slot of obs_part; obs_reduce sums the slots.  Called by every thread.
__device__ __forceinline__ void obs_partial(const TiledArgs& a, int b, int slot, double pr, double pi, int e,
                                            double2* sRed) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const double sr = wave_sum(pr), si = wave_sum(pi);
  __syncthreads();
  if (lane == 0) sRed[w] = make_double2(sr, si);
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 t = make_double2(0.0, 0.0);
    for (int x = 0; x < 4; ++x) {
      t.x += sRed[x].x;
      t.y += sRed[x].y;
    }
    a.obs_part[((size_t)b * a.slots + slot) * a.n_eops + e] = t;
  }
}

// ---- bordered operators (n = nc + 1, nc a multiple of TILE, no dense operators): the
// MFMA tiles cover the nc x nc core -- their reduction runs over all n indices, so the
// border terms of core outputs are included -- and BORDER_WGS extra workgroups per
// operator form the 2 nc + 1 outputs of row nc and column nc as complex dot products,
// BG lanes per output (shuffle-reduced), BG_OUT outputs per workgroup.
constexpr int BG = 8;
constexpr int BG_OUT = 256 / BG;

// C = (H S - S H^dag)_(bi, bj) over the lanes q, q + BG, ... of this output's group
template <typename R>
__device__ __forceinline__ void border_dot(const TiledArgs& a, int b, int set, int bi, int bj, int q, bool own,
                                           R& cre, R& cim) {
  using C = typename CT<R>::C;
  const int n = a.n;
  const size_t nn = (size_t)n * n;
  const C* H = static_cast<const C*>(a.heff) + (size_t)set * nn;
  const C* S = static_cast<const C*>(a.s_cur) + (size_t)b * nn;
  R xr = 0, xi = 0;
  if (own) {
    for (int k = q; k < n; k += BG) {
      const C h = H[(size_t)bi * n + k], sv = S[(size_t)k * n + bj];
      const C z = S[(size_t)bi * n + k], g = H[(size_t)bj * n + k];
      xr += h.x * sv.x - h.y * sv.y - (z.x * g.x + z.y * g.y);
      xi += h.x * sv.y + h.y * sv.x - (z.y * g.x - z.x * g.y);
    }
  }
#pragma unroll
  for (int off = 1; off < BG; off <<= 1) {
    xr += __shfl_xor(xr, off, 64);
    xi += __shfl_xor(xi, off, 64);
  }
  cre = xr;
  cim = xi;
}

__device__ __forceinline__ void border_index(int nc, int oidx, int& bi, int& bj) {
  if (oidx <= nc) {
    bi = nc;
    bj = oidx;
  } else {
    bi = oidx - nc - 1;
    bj = nc;
  }
}

// A border workgroup of the stage (bordered layout): the last row and column of the
// operator as complex dot products, with the same element update and records.
template <typename R>
__device__ __forceinline__ void stage_border(const TiledArgs& a, int b, int set, int idx, int tiles, double2* sRed) {
  const int n = a.n, nc = a.nc;
  const size_t nn = (size_t)n * n;
  const bool rec = (a.stage == 3) && a.rec >= 0;
  const int q = threadIdx.x % BG;
  const int oidx = (idx - tiles) * BG_OUT + threadIdx.x / BG;
  const bool own = oidx < 2 * nc + 1;
  int bi = 0, bj = 0;
  if (own) border_index(nc, oidx, bi, bj);
  R cre, cim;
  border_dot<R>(a, b, set, bi, bj, q, own, cre, cim);
  double2 rf = make_double2(0.0, 0.0);
  const bool lead = own && q == 0;
  if (lead) rf = stage_elem<R>(a, b, set, (size_t)bi * n + bj, cre, cim);
  if (rec && a.obs) {
    const double2* E0 = a.eops_t + (size_t)(b / a.eop_group) * a.n_eops * nn;
    for (int e = 0; e < a.n_eops; ++e) {
      double pr = 0.0, pi = 0.0;
      if (lead) {
        const double2 ev = E0[(size_t)e * nn + (size_t)bi * n + bj];
        pr = ev.x * rf.x - ev.y * rf.y;
        pi = ev.x * rf.y + ev.y * rf.x;
      }
      obs_partial(a, b, idx, pr, pi, e, sRed);
    }
  }
}

// RK4 stage epilogue of one tile: stage_elem on every element of the tile, then the
// tile's share of the records.  Called by every thread of the workgroup.
template <int TILE, typename R, int P>
__device__ __forceinline__ void stage_tile_epilogue(const TiledArgs& a, int b, int set, int idx, const Geo& g,
                                                    typename CT<R>::V4 (&cr)[TILE / 32][TILE / 32],
                                                    typename CT<R>::V4 (&ci)[TILE / 32][TILE / 32],
                                                    double2* sRed) {
  constexpr int BW = TILE / 32;
  const int n = a.n, nc = a.nc;
  const size_t nn = (size_t)n * n;
  const bool rec = (a.stage == 3) && a.rec >= 0;
  double2 rfin[BW][BW][4];
#pragma unroll
  for (int p = 0; p < BW; ++p)
#pragma unroll
    for (int q = 0; q < BW; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        rfin[p][q][v] = make_double2(0.0, 0.0);
        const int i = g.row0 + g.wr * (TILE / 2) + p * 16 + CT<R>::row(g.lg, v);
        const int j = g.col0 + g.wc * (TILE / 2) + q * 16 + g.lr;
        if (i >= nc || j >= nc) continue;
        rfin[p][q][v] = stage_elem<R>(a, b, set, (size_t)i * n + j, cr[p][q][v], ci[p][q][v]);
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
            const int i = g.row0 + g.wr * (TILE / 2) + p * 16 + CT<R>::row(g.lg, v);
            const int j = g.col0 + g.wc * (TILE / 2) + q * 16 + g.lr;
            if (i >= nc || j >= nc) continue;
            const double2 ev = E0[(size_t)e * nn + (size_t)i * n + j];
            pr += ev.x * rfin[p][q][v].x - ev.y * rfin[p][q][v].y;
            pi += ev.x * rfin[p][q][v].y + ev.y * rfin[p][q][v].x;
          }
      obs_partial(a, b, idx, pr, pi, e, sRed);
    }
  }
}

// One RK4 stage for every operator of the batch.
template <int TILE, typename R, int P>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kMinWaves<TILE, P>))) lindblad_stage_tiled(TiledArgs a) {
  using C = typename CT<R>::C;
  using V4 = typename CT<R>::V4;
  __shared__ __align__(16) unsigned char smem[smem_bytes<TILE, R, P>()];
  __shared__ double2 sRed[4];
  const int n = a.n, nc = a.nc;
  const int tiles = tiles_per_dim(nc, TILE) * tiles_per_dim(nc, TILE);
  const int per = tiles + a.border_wgs;
  const int blk = xcd_block();
  const int b = blk / per, idx = blk % per;
  const int set = b / a.group;
  const size_t nn = (size_t)n * n;
  const bool rec = (a.stage == 3) && a.rec >= 0;

  if (idx >= tiles) {                         // border workgroup (uniform branch)
    stage_border<R>(a, b, set, idx, tiles, sRed);
    return;
  }

  const Geo g = geometry<TILE>(nc, idx, threadIdx.x);
  constexpr int BW = TILE / 32;
  V4 cr[BW][BW], ci[BW][BW];
#pragma unroll
  for (int p = 0; p < BW; ++p)
#pragma unroll
    for (int q = 0; q < BW; ++q) cr[p][q] = ci[p][q] = V4{0, 0, 0, 0};
  generator_gemms<TILE, R, P>(a, b, g, smem, cr, ci, 0, n);
  stage_tile_epilogue<TILE, R, P>(a, b, set, idx, g, cr, ci, sRed);
}

// K-split stage products (a.ks > 1): ks consecutive workgroups per core tile, part p
// reducing k in [p klen, (p + 1) klen) into its partial sums; lindblad_stage_sum then
// finishes the stage (and forms the border of the bordered layout).  More workgroups for
// small batches.
template <int TILE, typename R, int P>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kMinWaves<TILE, P>))) lindblad_stage_part(TiledArgs a) {
  using C = typename CT<R>::C;
  using V4 = typename CT<R>::V4;
  __shared__ __align__(16) unsigned char smem[smem_bytes<TILE, R, P>()];
  const int n = a.n, nc = a.nc;
  const int tiles = tiles_per_dim(nc, TILE) * tiles_per_dim(nc, TILE);
  const int blk0 = xcd_block();
  const int part = blk0 % a.ks, blk = blk0 / a.ks;
  const int b = blk / tiles, idx = blk % tiles;
  const size_t nn = (size_t)n * n;
  const Geo g = geometry<TILE>(nc, idx, threadIdx.x);
  constexpr int BW = TILE / 32;
  V4 cr[BW][BW], ci[BW][BW];
#pragma unroll
  for (int p = 0; p < BW; ++p)
#pragma unroll
    for (int q = 0; q < BW; ++q) cr[p][q] = ci[p][q] = V4{0, 0, 0, 0};
  const int kb = part * a.klen, ke = kb + a.klen < n ? kb + a.klen : n;
  generator_gemms<TILE, R, P>(a, b, g, smem, cr, ci, kb, ke);
  C* Pp = static_cast<C*>(a.part) + ((size_t)part * a.batch + b) * nn;
#pragma unroll
  for (int p = 0; p < BW; ++p)
#pragma unroll
    for (int q = 0; q < BW; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = g.row0 + g.wr * (TILE / 2) + p * 16 + CT<R>::row(g.lg, v);
        const int j = g.col0 + g.wc * (TILE / 2) + q * 16 + g.lr;
        if (i < nc && j < nc) Pp[(size_t)i * n + j] = CT<R>::make(cr[p][q][v], ci[p][q][v]);
      }
}

// The stage of a K-split launch: the partial sums of every part, in part order, then the
// same epilogue as lindblad_stage_tiled (one workgroup per tile; the border workgroups of
// the bordered layout as in lindblad_stage_tiled).
template <int TILE, typename R, int P>
__global__ void __launch_bounds__(256) lindblad_stage_sum(TiledArgs a) {
  using C = typename CT<R>::C;
  using V4 = typename CT<R>::V4;
  __shared__ double2 sRed[4];
  const int n = a.n, nc = a.nc;
  const int tiles = tiles_per_dim(nc, TILE) * tiles_per_dim(nc, TILE);
  const int per = tiles + a.border_wgs;
  const int blk = xcd_block();
  const int b = blk / per, idx = blk % per;
  const size_t nn = (size_t)n * n;
  if (idx >= tiles) {                         // border workgroup (uniform branch)
    stage_border<R>(a, b, b / a.group, idx, tiles, sRed);
    return;
  }
  const Geo g = geometry<TILE>(nc, idx, threadIdx.x);
  constexpr int BW = TILE / 32;
  V4 cr[BW][BW], ci[BW][BW];
#pragma unroll
  for (int p = 0; p < BW; ++p)
#pragma unroll
    for (int q = 0; q < BW; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = g.row0 + g.wr * (TILE / 2) + p * 16 + CT<R>::row(g.lg, v);
        const int j = g.col0 + g.wc * (TILE / 2) + q * 16 + g.lr;
        R xr = 0, xi = 0;
        if (i < nc && j < nc) {
          const C* Pp = static_cast<const C*>(a.part) + (size_t)b * nn + (size_t)i * n + j;
          for (int part = 0; part < a.ks; ++part) {
            const C x = Pp[(size_t)part * a.batch * nn];
            xr += x.x;
            xi += x.y;
          }
        }
        cr[p][q][v] = xr;
        ci[p][q][v] = xi;
      }
  stage_tile_epilogue<TILE, R, P>(a, b, b / a.group, idx, g, cr, ci, sRed);
}

// One application of the generator to every operator of the batch, with a flexible
// epilogue: V = f(s_cur); y = c_lx V + c_z z; y_op = y in the operand precision;
// acc += c_acc V.
template <int TILE, typename R, int P>
__global__ void __launch_bounds__(256) lindblad_apply_tiled(TiledArgs a) {
  using V4 = typename CT<R>::V4;
  __shared__ __align__(16) unsigned char smem[smem_bytes<TILE, R, P>()];
  const int n = a.n, nc = a.nc;
  const int tiles = tiles_per_dim(nc, TILE) * tiles_per_dim(nc, TILE);
  const int per = tiles + a.border_wgs;
  const int blk = xcd_block();
  const int b = blk / per, idx = blk % per;
  const int set = b / a.group;

  if (idx >= tiles) {                         // border workgroup (uniform branch)
    const int q = threadIdx.x % BG;
    const int oidx = (idx - tiles) * BG_OUT + threadIdx.x / BG;
    const bool own = oidx < 2 * nc + 1;
    int bi = 0, bj = 0;
    if (own) border_index(nc, oidx, bi, bj);
    R cre, cim;
    border_dot<R>(a, b, set, bi, bj, q, own, cre, cim);
    if (own && q == 0) apply_elem<R>(a, b, set, (size_t)bi * n + bj, cre, cim);
    return;
  }

  const Geo g = geometry<TILE>(nc, idx, threadIdx.x);
  constexpr int BW = TILE / 32;
  V4 cr[BW][BW], ci[BW][BW];
#pragma unroll
  for (int p = 0; p < BW; ++p)
#pragma unroll
    for (int q = 0; q < BW; ++q) cr[p][q] = ci[p][q] = V4{0, 0, 0, 0};
  generator_gemms<TILE, R, P>(a, b, g, smem, cr, ci, 0, n);
#pragma unroll
  for (int p = 0; p < BW; ++p)
#pragma unroll
    for (int q = 0; q < BW; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = g.row0 + g.wr * (TILE / 2) + p * 16 + CT<R>::row(g.lg, v);
        const int j = g.col0 + g.wc * (TILE / 2) + q * 16 + g.lr;
        if (i >= nc || j >= nc) continue;
        apply_elem<R>(a, b, set, (size_t)i * n + j, cr[p][q][v], ci[p][q][v]);
      }
}

// ---- uneven batches in one persistent launch
// A fixed grid of workgroups drains a queue of instances ordered by cost (largest n^3 (1 +
// K_d) first): a workgroup takes the next instance with one atomic and advances it through
// every RK4 step of the call before taking another, so the largest systems start first and
// the small ones fill the tail.  Each workgroup holds SG groups of 256 threads; every group
// runs the TILE x TILE tile machinery of the per-stage kernels (K-panels staged through its
// own LDS area) on tiles t, t + SG, ... of the instance, and a workgroup barrier separates
// the dense products T_k = L_k S of a stage, the stage's tiles, and the next stage.  An
// instance never leaves its workgroup, so every operand it writes is read back by the same
// compute unit: no synchronisation between workgroups and no cross-XCD coherence traffic.
// The last workgroup to finish resets the queue, so the next launch (or graph replay)
// starts from an empty queue with no host work.
__device__ __forceinline__ long long uniform64(long long x) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(unsigned long long)x);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)((unsigned long long)x >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ RaggedInstance uniform_instance(const RaggedInstance& v) {
  RaggedInstance u;
  u.n = __builtin_amdgcn_readfirstlane(v.n);
  u.n_dense = __builtin_amdgcn_readfirstlane(v.n_dense);
  u.heff = uniform64(v.heff);
  u.dmat = uniform64(v.dmat);
  u.lops = uniform64(v.lops);
  u.rho = uniform64(v.rho);
  u.work = uniform64(v.work);
  return u;
}

template <int TILE, int SG, int MINW>
__global__ void __launch_bounds__(256 * SG) __attribute__((amdgpu_waves_per_eu(MINW)))
lindblad_ragged_persistent(RaggedArgs r) {
  using C = float2;
  __shared__ __align__(16) unsigned char smem[SG][smem_bytes<TILE, float, 0>()];
  __shared__ int s_item;
  const int grp = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), ltid = threadIdx.x & 255;   // grp: wave-uniform
  for (;;) {
    __syncthreads();                          // s_item of the previous instance is consumed
    if (threadIdx.x == 0) s_item = (int)atomicAdd(r.queue, 1u);
    __syncthreads();
    // the item and the instance record are wave-uniform: kept in scalar registers, so the
    // instance's pointers cost no vector registers in the tile code below
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    if (item >= r.count) break;
    const RaggedInstance in = uniform_instance(r.inst[__builtin_amdgcn_readfirstlane(r.order[item])]);
    const int n = in.n;
    const size_t nn = (size_t)n * n;
    TiledArgs a = {};
    a.batch = 1;
    a.n = n;
    a.nc = n;
    a.group = 1;
    a.n_dense = in.n_dense;
    a.dt = r.dt;
    a.rec = -1;
    a.heff = r.ops + in.heff;
    a.dmat = in.dmat >= 0 ? r.ops + in.dmat : nullptr;
    a.lops = in.n_dense > 0 ? r.ops + in.lops : nullptr;
    a.rho = r.rho + in.rho;
    C* s0 = r.work + in.work;
    C* s1 = s0 + nn;
    a.inc = s1 + nn;
    a.tbuf = s1 + 2 * nn;
    for (size_t i = threadIdx.x; i < nn; i += blockDim.x) s0[i] = make_float2((float)a.rho[i].x, (float)a.rho[i].y);
    __syncthreads();
    const int tiles = tiles_per_dim(n, TILE) * tiles_per_dim(n, TILE);
    for (int step = 0; step < r.nsteps; ++step) {
      for (int st = 0; st < 4; ++st) {
        a.stage = st;
        a.s_cur = (st % 2 == 0) ? s0 : s1;
        a.s_next = (st % 2 == 0) ? s1 : s0;
        // every group runs the same number of tiles (a group past the last tile runs a
        // dummy one: all its MFMA blocks lie outside the operator and are skipped), so the
        // panel barriers inside the GEMMs match across the workgroup
        for (int k = 0; k < a.n_dense; ++k) {
          for (int t0 = 0; t0 < tiles; t0 += SG) {
            const int t = t0 + grp;
            const Geo g = geometry<TILE>(n, t < tiles ? t : tiles, ltid);   // tiles: rows >= n
            left_tile<TILE, float, 0, true>(a, 0, k, g, smem[grp]);
          }
        }
        if (a.n_dense > 0) __syncthreads();   // T_k complete
        constexpr int BW = TILE / 32;
        for (int t0 = 0; t0 < tiles; t0 += SG) {
          const int t = t0 + grp;
          const Geo g = geometry<TILE>(n, t < tiles ? t : tiles, ltid);   // tiles: rows >= n
          floatx4 cr[BW][BW], ci[BW][BW];
#pragma unroll
          for (int p = 0; p < BW; ++p)
#pragma unroll
            for (int q = 0; q < BW; ++q) cr[p][q] = ci[p][q] = floatx4{0.f, 0.f, 0.f, 0.f};
          generator_gemms<TILE, float, 0, true>(a, 0, g, smem[grp], cr, ci, 0, n);
#pragma unroll
          for (int p = 0; p < BW; ++p)
#pragma unroll
            for (int q = 0; q < BW; ++q)
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                const int i = g.row0 + g.wr * (TILE / 2) + p * 16 + CT<float>::row(g.lg, v);
                const int j = g.col0 + g.wc * (TILE / 2) + q * 16 + g.lr;
                if (i < n && j < n) stage_elem<float>(a, 0, 0, (size_t)i * n + j, cr[p][q][v], ci[p][q][v]);
              }
        }
        __syncthreads();                        // the next stage operand is complete
      }
    }
  }
  if (threadIdx.x == 0 && atomicAdd(r.queue + 1, 1u) == gridDim.x - 1) {
    // every other workgroup has left the loop: reset the queue for the next launch
    atomicExch(r.queue, 0u);
    atomicExch(r.queue + 1, 0u);
  }
}

// Gradient contractions of an adjoint stage operator kappa with a forward stage operator
// s (P = conj(kappa) o s, per operator b):
//   out_a[j] += w (sum_m P_jm - sum_m P_mj)                -> d/dH_jj   (times -i)
//   out_b[j] += w (P_jj - (sum_m P_jm + sum_m P_mj) / 2)    -> d/d(dephasing rate of j)
//   out_u[j] += w (sum_m conj(k_jm) s_j+1,m + conj(k_j+1,m) s_jm
//                  - sum_m conj(k_m,j+1) s_mj + conj(k_mj) s_m,j+1)   -> d/dH_j,j+1 (times -i)
// One 1024-thread workgroup per operator and ONE pass over kappa and s: wave w sweeps the
// rows of its band (ceil(n / 16) consecutive rows), 64 columns at a time, lanes along the
// row (coalesced).  Row sums and row-pair terms are FP64 wave sums, added by the row's own
// wave into LDS; the previous row of the row-pair terms stays in registers, the next column
// of the column-pair terms is the neighbouring lane's cache line.  Column sums and column-pair terms stay in the lane's registers
// for the band and meet in LDS in a fixed order (waves in NS slots, slot-sharing waves in
// rounds).  Deterministic: no atomics.
constexpr int CT_THREADS = 1024;
constexpr int CT_WAVES = CT_THREADS / 64;
constexpr size_t CT_LDS = 156 * 1024;
// column-partial slots that fit the LDS next to the row arrays
__host__ __device__ inline int contract_slots(int n) {
  const long long free_bytes = (long long)CT_LDS - 3ll * n * 16;
  long long s = free_bytes / (2ll * n * 16);
  if (s > CT_WAVES) s = CT_WAVES;
  return s < 1 ? 0 : (int)s;
}

__global__ void __launch_bounds__(CT_THREADS) adjoint_contract(ContractArgs c) {
  extern __shared__ __attribute__((aligned(16))) unsigned char craw[];
  const int b = blockIdx.x, n = c.n;
  const size_t nn = (size_t)n * n;
  const int NS = contract_slots(n);
  double2* colC = reinterpret_cast<double2*>(craw);   // [NS][n] column-sum partials
  double2* colP = colC + (size_t)NS * n;               // [NS][n] column-pair partials
  double2* rowR = colP + (size_t)NS * n;               // [n] row sums
  double2* rowP = rowR + n;                            // [n] row-pair terms
  double2* dia = rowP + n;                             // [n] diagonal
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int band = (n + CT_WAVES - 1) / CT_WAVES;
  const int r0 = min(n, wave * band), r1 = min(n, r0 + band);
  for (int i = threadIdx.x; i < 3 * n; i += CT_THREADS) rowR[i] = make_double2(0.0, 0.0);
  __syncthreads();
  const int rounds = (CT_WAVES + NS - 1) / NS;
  for (int c0 = 0; c0 < n; c0 += 64) {
    const int m = c0 + lane;
    const bool in = m < n, has_next = m + 1 < n;
    double cr = 0, ci = 0, qr = 0, qi = 0;      // column sum and column-pair term of column m
    // kappa and s of the previous row at column m, per pair (registers: each element is read
    // from memory once, plus one row above the band)
    double2 kp[kMaxContractPairs], xp[kMaxContractPairs];
#pragma unroll
    for (int pp = 0; pp < kMaxContractPairs; ++pp) {
      kp[pp] = make_double2(0.0, 0.0);
      xp[pp] = make_double2(0.0, 0.0);
      if (pp < c.pairs && in && r0 > 0 && r0 < r1) {
        kp[pp] = c.kappa[pp][(size_t)b * nn + (size_t)(r0 - 1) * n + m];
        xp[pp] = c.s[pp][(size_t)b * nn + (size_t)(r0 - 1) * n + m];
      }
    }
    for (int r = r0; r < r1; ++r) {
      double pr = 0, pi = 0, tr = 0, ti = 0;    // P_rm; the pair term of rows r - 1, r at m
      double ur = 0, ui = 0;                     // the pair term of columns m, m + 1 at row r
#pragma unroll
      for (int pp = 0; pp < kMaxContractPairs; ++pp) {
        if (pp < c.pairs && in) {
          const double2* K = c.kappa[pp] + (size_t)b * nn;
          const double2* S = c.s[pp] + (size_t)b * nn;
          const double2 k = K[(size_t)r * n + m], x = S[(size_t)r * n + m];
          pr += k.x * x.x + k.y * x.y;
          pi += k.x * x.y - k.y * x.x;
          if (has_next) {   // the next column of the same row: the neighbouring lane's line
            const double2 k1 = K[(size_t)r * n + m + 1], x1 = S[(size_t)r * n + m + 1];
            ur += (k1.x * x.x + k1.y * x.y) + (k.x * x1.x + k.y * x1.y);
            ui += (k1.x * x.y - k1.y * x.x) + (k.x * x1.y - k.y * x1.x);
          }
          if (r > 0) {
            tr += (kp[pp].x * x.x + kp[pp].y * x.y) + (k.x * xp[pp].x + k.y * xp[pp].y);
            ti += (kp[pp].x * x.y - kp[pp].y * x.x) + (k.x * xp[pp].y - k.y * xp[pp].x);
          }
          kp[pp] = k;
          xp[pp] = x;
        }
      }
      cr += pr;
      ci += pi;
      qr += ur;
      qi += ui;
      if (in && m == r) dia[r] = make_double2(pr, pi);
      // row quantities: this wave owns row r (and the pair term of rows r - 1, r)
      const double sr = wave_sum(pr), si = wave_sum(pi);
      const double wr = wave_sum(tr), wi = wave_sum(ti);
      if (lane == 0) {
        rowR[r].x += sr;
        rowR[r].y += si;
        if (r > 0) {
          rowP[r - 1].x += wr;
          rowP[r - 1].y += wi;
        }
      }
    }
    // column partials of this 64-column block: slot wave % NS, waves sharing a slot in
    // rounds (round 0 stores, later rounds add), so the sum order is fixed
    for (int k = 0; k < rounds; ++k) {
      if (wave / NS == k && in) {
        double2* C = colC + (size_t)(wave % NS) * n + m;
        double2* Q = colP + (size_t)(wave % NS) * n + m;
        if (k == 0) {
          *C = make_double2(cr, ci);
          *Q = make_double2(qr, qi);
        } else {
          C->x += cr;
          C->y += ci;
          Q->x += qr;
          Q->y += qi;
        }
      }
      if (rounds > 1) __syncthreads();
    }
  }
  __syncthreads();
  const int used = min(NS, CT_WAVES);
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double2 Cc = make_double2(0.0, 0.0), Cp = make_double2(0.0, 0.0);
    for (int x = 0; x < used; ++x) {
      Cc.x += colC[(size_t)x * n + j].x;
      Cc.y += colC[(size_t)x * n + j].y;
      Cp.x += colP[(size_t)x * n + j].x;
      Cp.y += colP[(size_t)x * n + j].y;
    }
    const double2 R = rowR[j], D = dia[j];
    double2* A = c.out_a + (size_t)b * n + j;
    A->x += c.weight * (R.x - Cc.x);
    A->y += c.weight * (R.y - Cc.y);
    double2* B = c.out_b + (size_t)b * n + j;
    B->x += c.weight * (D.x - 0.5 * (R.x + Cc.x));
    B->y += c.weight * (D.y - 0.5 * (R.y + Cc.y));
    if (j + 1 < n) {
      double2* U = c.out_u + (size_t)b * (n - 1) + j;
      U->x += c.weight * (rowP[j].x - Cp.x);
      U->y += c.weight * (rowP[j].y - Cp.y);
    }
  }
}

// Start of a backward RK4 step: lam += g[b] m;  k4 = c lam;  k4_op = (operand precision) k4;
// acc = lam.  One pass over the operators instead of four.
template <typename C>
__global__ void __launch_bounds__(256) adjoint_seed(SeedArgs s) {
  const size_t nn = (size_t)s.n * s.n, total = (size_t)s.batch * nn;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const size_t b = i / nn, o = i - b * nn;
    const double2 g = s.g[b], m = s.m[o];
    double2 l = s.lam[i];
    l.x += g.x * m.x - g.y * m.y;
    l.y += g.x * m.y + g.y * m.x;
    s.lam[i] = l;
    s.acc[i] = l;
    const double2 k = make_double2(s.c * l.x, s.c * l.y);
    s.k4[i] = k;
    C kc;
    kc.x = k.x;
    kc.y = k.y;
    static_cast<C*>(s.k4_op)[i] = kc;
  }
}

hipError_t seed_launch(const SeedArgs& s, hipStream_t stream) {
  const size_t total = (size_t)s.batch * s.n * s.n;
  size_t blocks = (total + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (s.op_fp64)
    hipLaunchKernelGGL(adjoint_seed<double2>, dim3((unsigned)blocks), dim3(256), 0, stream, s);
  else
    hipLaunchKernelGGL(adjoint_seed<float2>, dim3((unsigned)blocks), dim3(256), 0, stream, s);
  return hipGetLastError();
}

// 0 when no slot fits (n too large for the contraction's LDS)
size_t contract_lds_bytes(int n) {
  const int ns = contract_slots(n);
  return ns < 1 ? 0 : (size_t)(2 * ns + 3) * n * sizeof(double2);
}

hipError_t contract_launch(const ContractArgs& c, hipStream_t stream) {
  const size_t lds = contract_lds_bytes(c.n);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(adjoint_contract),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(adjoint_contract, dim3(c.batch), dim3(CT_THREADS), lds, stream, c);
  return hipGetLastError();
}

// Bordered layout when n = nc + 1 with nc a multiple of the tile and no dense operators
// (their products T_k = L_k S would need a border of their own); sets a.nc, a.border_wgs.
// Returns the workgroups per operator of the stage / apply kernels.
// (The border dot products carry both H_eff products, so one-sided generators never use it.)
static int layout(TiledArgs& a, int tile) {
  const bool brd = a.n_dense == 0 && !(a.flags & kFlagOneSided) && a.n > tile && (a.n - 1) % tile == 0;
  a.nc = brd ? a.n - 1 : a.n;
  a.border_wgs = brd ? (2 * a.nc + 1 + BG_OUT - 1) / BG_OUT : 0;
  const int tpd = (a.nc + tile - 1) / tile;
  return tpd * tpd + a.border_wgs;
}

template <int TILE, typename R>
static hipError_t apply_impl(TiledArgs a, hipStream_t stream) {
  const int per = layout(a, TILE);
  if (a.n_dense > 0)
    hipLaunchKernelGGL((lindblad_left_tiled<TILE, R, 0>), dim3(per * a.batch * a.n_dense), dim3(256), 0, stream, a);
  hipLaunchKernelGGL((lindblad_apply_tiled<TILE, R, 0>), dim3(per * a.batch), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t tiled_apply(TiledArgs a, hipStream_t stream, int tile, bool fp64) {
  if (fp64) return tile == 64 ? apply_impl<64, double>(a, stream) : apply_impl<32, double>(a, stream);
  return tile == 64 ? apply_impl<64, float>(a, stream) : apply_impl<32, float>(a, stream);
}

template <int TILE, typename R, int P>
static hipError_t run_step(TiledArgs a, hipStream_t stream, void* s0, void* s1) {
  const int per = layout(a, TILE);
  // K-split: 32-tiles only
  if (a.ks < 1 || TILE != 32) a.ks = 1;
  for (int st = 0; st < 4; ++st) {
    a.stage = st;
    a.s_cur = (st % 2 == 0) ? s0 : s1;
    a.s_next = (st % 2 == 0) ? s1 : s0;
    if (a.n_dense > 0) {
      hipLaunchKernelGGL((lindblad_left_tiled<TILE, R, P>), dim3(per * a.batch * a.n_dense), dim3(256), 0,
                         stream, a);
    }
    bool split = false;
    if constexpr (TILE == 32) {
      if (a.ks > 1) {
        hipLaunchKernelGGL((lindblad_stage_part<TILE, R, P>), dim3((per - a.border_wgs) * a.batch * a.ks), dim3(256),
                           0, stream, a);
        hipLaunchKernelGGL((lindblad_stage_sum<TILE, R, P>), dim3(per * a.batch), dim3(256), 0, stream, a);
        split = true;
      }
    }
    if (!split) hipLaunchKernelGGL((lindblad_stage_tiled<TILE, R, P>), dim3(per * a.batch), dim3(256), 0, stream, a);
  }
  return hipGetLastError();
}

// Record of the step's observables: the sum of the per-workgroup partials of every
// operator, in slot order (deterministic), stored at record index a.rec.
__global__ void __launch_bounds__(64) obs_reduce(TiledArgs a) {
  const int b = blockIdx.x;
  for (int e = threadIdx.x; e < a.n_eops; e += 64) {
    double2 t = make_double2(0.0, 0.0);
    const double2* p = a.obs_part + (size_t)b * a.slots * a.n_eops + e;
    for (int x = 0; x < a.slots; ++x) {
      const double2 y = p[(size_t)x * a.n_eops];
      t.x += y.x;
      t.y += y.y;
    }
    a.obs[((size_t)b * a.rec_capacity + a.rec) * a.n_eops + e] = t;
  }
}

// Trace renormalisation after a step (flag bit 0; density matrices only): the stage-3
// kernels have written rho and obs_reduce the step's observables.  renorm_trace computes
// s_b = 1 / Re Tr rho_b in FP64 (one workgroup per operator) and rescales the observables
// recorded at this step; renorm_scale applies s_b to rho, the next stage operand and the
// recorded state.  Observables and states are therefore recorded after renormalisation.
__global__ void __launch_bounds__(256) renorm_trace(TiledArgs a) {
  __shared__ double part[4];
  const int b = blockIdx.x, n = a.n;
  const double2* rho = a.rho + (size_t)b * n * n;
  double t = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) t += rho[(size_t)i * n + i].x;
  t = wave_sum(t);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = t;
  __syncthreads();
  const double s = 1.0 / (part[0] + part[1] + part[2] + part[3]);
  if (threadIdx.x == 0) a.tscale[b] = s;
  if (a.rec >= 0 && a.obs)
    for (int e = threadIdx.x; e < a.n_eops; e += 256) {
      double2* o = a.obs + ((size_t)b * a.rec_capacity + a.rec) * a.n_eops + e;
      o->x *= s;
      o->y *= s;
    }
}

template <typename C>
__global__ void renorm_scale(TiledArgs a, C* __restrict__ s0) {
  const size_t nn = (size_t)a.n * a.n;
  const size_t total = (size_t)a.batch * nn;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const size_t b = i / nn;
    const double s = a.tscale[b];
    double2 r = a.rho[i];
    r.x *= s;
    r.y *= s;
    a.rho[i] = r;
    C v;
    v.x = r.x;
    v.y = r.y;
    s0[i] = v;
    if (a.rec >= 0 && a.states) a.states[(b * a.rec_capacity + a.rec) * nn + (i - b * nn)] = r;
  }
}

template <typename C>
__global__ void init_stage_operand(const double2* __restrict__ rho, C* __restrict__ s, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    C v;
    v.x = rho[i].x;
    v.y = rho[i].y;
    s[i] = v;
  }
}

// Byte offsets of the workspace regions (see tiled_workspace_bytes).
struct WorkLayout {
  size_t s0, s1, inc, tbuf, tscale, obs_part, split, part, total;
  int ldp, nsets, ks, klen;
};

// Row pitch of the pre-split planes: n rounded up to whole 16-byte chunks of BF16.
static int split_pitch(int n) { return (n + 7) & ~7; }

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Workgroups per operator of the stage kernel for a tile edge (the bordered layout adds
// the border workgroups).
static int stage_wgs(int n, int n_dense, int flags, int tile) {
  TiledArgs t = {};
  t.n = n;
  t.n_dense = n_dense;
  t.flags = flags;
  return layout(t, tile);
}

// K-split parts actually used (32-tiles only).
static int effective_ks(int tile, int ks) { return (ks <= 1 || tile != 32) ? 1 : ks; }

static WorkLayout work_layout(int batch, int n, int group, int n_dense, int flags, bool fp64, int n_eops,
                              bool rops, int tile, int ks) {
  const bool renorm = (flags & 1) != 0;
  const size_t op = (size_t)batch * n * n * (fp64 ? sizeof(double2) : sizeof(float2));
  WorkLayout w;
  w.s0 = 0;
  w.s1 = align256(op);
  w.inc = w.s1 + align256(op);
  w.tbuf = w.inc + align256(op);
  // T_k: complex64 / complex128, or (3xBF16) split planes of row pitch split_pitch(n)
  const bool split = !fp64 && (flags & 0x30) == 0x10;
  const size_t tsz = split ? (size_t)batch * n_dense * 4 * n * split_pitch(n) * sizeof(unsigned short)
                           : op * (size_t)n_dense;
  w.tscale = w.tbuf + align256(tsz);
  w.obs_part = w.tscale + (renorm ? align256((size_t)batch * sizeof(double)) : 0);
  // slots for the larger of the two tile choices (32-tiles: more workgroups)
  const int w32 = stage_wgs(n, n_dense, flags, 32), w64 = stage_wgs(n, n_dense, flags, 64);
  const int slots = w32 > w64 ? w32 : w64;
  w.split = w.obs_part + (n_eops > 0 ? align256((size_t)batch * slots * n_eops * sizeof(double2)) : 0);
  // 3xBF16: the pre-split planes of H_eff, L_k (and R_k) of every operator set
  w.ldp = split_pitch(n);
  w.nsets = (batch + (group > 1 ? group : 1) - 1) / (group > 1 ? group : 1);
  const size_t mats = (size_t)w.nsets * (1 + (size_t)n_dense * (rops ? 2 : 1));
  w.part = w.split + (split ? align256(mats * 4 * n * w.ldp * sizeof(unsigned short)) : 0);
  // K-split partial sums: ks x batch operators in the operand precision; parts of klen
  // (a multiple of both panel widths) reduction indices
  w.ks = effective_ks(tile, ks);
  w.klen = ((n + w.ks - 1) / w.ks + 31) / 32 * 32;
  w.total = w.part + (w.ks > 1 ? align256((size_t)w.ks * op) : 0);
  return w;
}

size_t tiled_workspace_bytes(int batch, int n, int group, int n_dense, int flags, bool fp64, int n_eops,
                             bool rops, int tile, int ks) {
  return work_layout(batch, n, group, n_dense, flags, fp64, n_eops, rops, tile, ks).total;
}

int tiled_pick(int batch, int n) {
  const int t64 = (n + 63) / 64;
  return ((size_t)batch * t64 * t64 >= 512) ? 64 : 32;
}

// Pre-split planes of count n x n complex64 matrices: thread t forms one 16-byte chunk (8
// consecutive columns of one row) of each of the four planes, zero beyond column n.
__global__ void __launch_bounds__(256) split_planes(const float2* __restrict__ src, unsigned short* __restrict__ dst,
                                                    int n, int ldp, size_t chunks) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= chunks) return;
  const int cpr = ldp / 8;
  const size_t row = t / cpr;                 // matrix * n + row within the matrix
  const int k0 = 8 * (int)(t % cpr);
  const size_t m = row / n, r = row % n;
  const float2* s = src + row * n;
  unsigned xh[4], xl[4], yh[4], yl[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int k = k0 + 2 * e;
    const float2 v0 = k < n ? s[k] : make_float2(0.f, 0.f);
    const float2 v1 = k + 1 < n ? s[k + 1] : make_float2(0.f, 0.f);
    split2(v0.x, v1.x, xh[e], xl[e]);
    split2(v0.y, v1.y, yh[e], yl[e]);
  }
  const size_t plane = (size_t)n * ldp;
  unsigned short* d = dst + m * 4 * plane + r * ldp + k0;
  *reinterpret_cast<uint4*>(d) = make_uint4(xh[0], xh[1], xh[2], xh[3]);
  *reinterpret_cast<uint4*>(d + plane) = make_uint4(xl[0], xl[1], xl[2], xl[3]);
  *reinterpret_cast<uint4*>(d + 2 * plane) = make_uint4(yh[0], yh[1], yh[2], yh[3]);
  *reinterpret_cast<uint4*>(d + 3 * plane) = make_uint4(yl[0], yl[1], yl[2], yl[3]);
}

template <typename R, int P>
static hipError_t run_all(TiledArgs a, int nsteps, int step_offset, int nout, void* work, hipStream_t stream,
                          int tile) {
  using C = typename CT<R>::C;
  const size_t nn = (size_t)a.n * a.n;
  const WorkLayout wl = work_layout(a.batch, a.n, a.group, a.n_dense, a.flags, sizeof(R) == 8, a.obs ? a.n_eops : 0,
                                    a.rops != nullptr, tile, a.ks);
  a.ks = wl.ks;
  a.klen = wl.klen;
  a.part = a.ks > 1 ? static_cast<void*>(static_cast<unsigned char*>(work) + wl.part) : nullptr;
  unsigned char* base = static_cast<unsigned char*>(work);
  if constexpr (P == 1) {
    // split the constant operators once per call: H_eff, then L_k, then R_k
    unsigned short* sp = reinterpret_cast<unsigned short*>(base + wl.split);
    const size_t np = (size_t)4 * a.n * wl.ldp;
    a.ldp = wl.ldp;
    a.hsp = sp;
    a.lsp = sp + (size_t)wl.nsets * np;
    a.rsp = a.rops ? a.lsp + (size_t)wl.nsets * a.n_dense * np : nullptr;
    unsigned short* dst[3] = {sp, const_cast<unsigned short*>(a.lsp), const_cast<unsigned short*>(a.rsp)};
    const void* src[3] = {a.heff, a.lops, a.rops};
    const int count[3] = {wl.nsets, wl.nsets * a.n_dense, a.rops ? wl.nsets * a.n_dense : 0};
    for (int m = 0; m < 3; ++m) {
      const size_t chunks = (size_t)count[m] * a.n * (wl.ldp / 8);
      if (chunks == 0) continue;
      hipLaunchKernelGGL(split_planes, dim3((unsigned)((chunks + 255) / 256)), dim3(256), 0, stream,
                         static_cast<const float2*>(src[m]), dst[m], a.n, wl.ldp, chunks);
    }
  }
  C* s0 = reinterpret_cast<C*>(base + wl.s0);
  C* s1 = reinterpret_cast<C*>(base + wl.s1);
  a.inc = base + wl.inc;
  a.tbuf = base + wl.tbuf;
  a.tscale = reinterpret_cast<double*>(base + wl.tscale);
  a.obs_part = reinterpret_cast<double2*>(base + wl.obs_part);
  a.slots = stage_wgs(a.n, a.n_dense, a.flags, tile);
  const size_t cnt = (size_t)a.batch * nn;
  hipLaunchKernelGGL(init_stage_operand<C>, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, stream, a.rho,
                     s0, cnt);
  for (int s = 0; s < nsteps; ++s) {
    const int gstep = step_offset + s + 1;
    a.rec = -1;
    if (nout > 0 && (a.obs || a.states || a.stage_rec) && gstep % nout == 0 && gstep / nout - 1 < a.rec_capacity)
      a.rec = gstep / nout - 1;
    // stage 3 writes S = rho into s0 (stages alternate s0 -> s1 -> s0 -> s1 -> s0)
    hipError_t e = (tile == 64) ? run_step<64, R, P>(a, stream, s0, s1) : run_step<32, R, P>(a, stream, s0, s1);
    if (e != hipSuccess) return e;
    if (a.rec >= 0 && a.obs) hipLaunchKernelGGL(obs_reduce, dim3(a.batch), dim3(64), 0, stream, a);
    if (a.flags & 1) {
      hipLaunchKernelGGL(renorm_trace, dim3(a.batch), dim3(256), 0, stream, a);
      const size_t blocks = (cnt + 255) / 256;
      hipLaunchKernelGGL(renorm_scale<C>, dim3((unsigned)(blocks < 8192 ? blocks : 8192)), dim3(256), 0, stream,
                         a, s0);
    }
  }
  return hipGetLastError();
}


// ---- persistent ragged launch: shapes (TILE, SG, MINW) by variant; MINW is the waves per
// SIMD the register budget is compiled for (512 / MINW VGPRs)
namespace {
constexpr int kRaggedVariants = 5;

template <int TILE, int SG, int MINW>
hipError_t ragged_launch_shape(const RaggedArgs& r, int max_groups, hipStream_t stream) {
  auto k = lindblad_ragged_persistent<TILE, SG, MINW>;
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k), 256 * SG, 0);
  if (e != hipSuccess) return e;
  int groups = cus * (per_cu > 0 ? per_cu : 1);
  if (max_groups > 0 && groups > max_groups) groups = max_groups;
  if (groups > r.count) groups = r.count;
  hipLaunchKernelGGL(k, dim3(groups), dim3(256 * SG), 0, stream, r);
  return hipGetLastError();
}
}  // namespace

int ragged_variants() { return kRaggedVariants; }

hipError_t ragged_run(const RaggedArgs& r, int variant, int max_groups, hipStream_t stream) {
  if (r.count <= 0 || r.nsteps <= 0) return hipSuccess;
  switch (variant) {
    case 0: return ragged_launch_shape<64, 2, 2>(r, max_groups, stream);   // 8 waves, 1 per CU
    case 1: return ragged_launch_shape<64, 1, 2>(r, max_groups, stream);   // 4 waves, 2 per CU
    case 2: return ragged_launch_shape<32, 4, 4>(r, max_groups, stream);   // 16 waves
    case 3: return ragged_launch_shape<32, 2, 4>(r, max_groups, stream);   // 8 waves, 2 per CU
    case 4: return ragged_launch_shape<32, 1, 4>(r, max_groups, stream);   // 4 waves, 4 per CU
    default: return hipErrorInvalidValue;
  }
}

hipError_t tiled_run(TiledArgs a, int nsteps, int step_offset, int nout, void* work, hipStream_t stream,
                     int tile, TiledPrec prec) {
  switch (prec) {
    case TiledPrec::FP64: return run_all<double, 0>(a, nsteps, step_offset, nout, work, stream, tile);
    case TiledPrec::SPLIT_BF16: return run_all<float, 1>(a, nsteps, step_offset, nout, work, stream, tile);
    default: return run_all<float, 0>(a, nsteps, step_offset, nout, work, stream, tile);
  }
}

}  // namespace lime
