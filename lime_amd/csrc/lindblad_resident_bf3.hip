// Fused batched Lindblad RK4 on 3xBF16 split products, one workgroup per propagated
// operator, state resident on chip: the split-product arithmetic of lindblad_rk4_resident
// for the bordered size n = 65 (an 8 x 8 lattice or a 64-site chain plus the ground state).
//
// Same right-hand side and time stepping as lindblad_rk4_resident (lime/oqs.py:706-723,
// lime/phys.py:636-649): f(X) = -i (H_eff X - X H_eff^dag) + D o X, RK4 with the FP64 state
// in VGPRs.  The two complex products run on v_mfma_f32_16x16x32_bf16: every real operand is
// split x = hi + lo (two BF16 values, hi rounded to nearest even, lo = x - hi rounded again,
// so |x - hi - lo| <= 2^-17 |x|) and every real product X Y is accumulated in FP32 as
// Xhi Yhi + Xhi Ylo + Xlo Yhi.
//
// Work decomposition (n = NC + 1, NC = 64; 8 waves, two 16 x 16 core tiles each):
//   * LDS holds the split operands as BF16 planes, one per part (Re hi, Re lo, Im hi, Im lo):
//     H_eff by rows (A operand of H S; B operand of S H^dag, whose B[k][j] = conj(H[j][k]))
//     and the stage operand S by columns (plane row j holds S[k][j]: B operand of H S, one
//     16-byte read of 8 consecutive k).  The A operand of S H^dag (S rows) comes from the
//     same column planes through ds_read_b64_tr_b16, whose result lane i, element e is
//     element i % 4 of what lane 4 e + i / 4 of its 16-lane group loaded: lane 4 e + q
//     loads S[4 q .. 4 q + 3][k + e] (8 bytes of plane row k + e), so lane i receives
//     S[i][k .. k + 3].  The lanes that own S elements write 4 consecutive rows of one
//     column per plane in the epilogue (one 8-byte store each), into the other of two
//     buffers: one barrier per stage.
//   * C = H S - S H^dag over k < NC on MFMA, 4 real products per complex product into split-
//     sign FP32 accumulators (no negated operand in the MFMA stream); the border index
//     k = NC is an FP32 VALU term from side arrays (row and column NC of H and S).
//   * Border outputs ride on the core fragments.  Row NC, columns of tile t: a 16-row A
//     operand whose rows 0 and 1 hold the (Re, Im) coefficients built from H[NC][:] and
//     S[NC][:] meets the B fragments of tile column t (4 products) and leaves Re C[NC][j]
//     in row 0, Im C[NC][j] in row 1 of one accumulator.  Column NC, rows of tile t: a
//     16-column B operand built from S[:][NC] and H[NC][:] meets the A fragments of tile
//     row t the same way.  The corner is a VALU wave sum.
//   * D o S is folded into the accumulators' start values by the epilogue (as in the FP32
//     kernel); the RK4 combination and the state update are FP64, the increment FP32.
#include "common.hpp"
#include "lindblad_resident_decl.hpp"

namespace lime {

namespace {

// a wave role as a type: the generic lambda below is instantiated once per role
template <int V> struct RoleK { static constexpr int value = V; };

typedef unsigned short u16;
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ floatx4 mfma_bf(u16x8 a, u16x8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                 0, 0, 0);
}

// c += X Y for split operands X = xh + xl, Y = yh + yl (the lo*lo term, <= 2^-16 |X Y|
// elementwise, is dropped); small terms first
__device__ __forceinline__ floatx4 mfma3(u16x8 xh, u16x8 xl, u16x8 yh, u16x8 yl, floatx4 c) {
  c = mfma_bf(xl, yh, c);
  c = mfma_bf(xh, yl, c);
  return mfma_bf(xh, yh, c);
}

// Split a pair of floats: h = hi(a) | hi(b) << 16, l = lo(a) | lo(b) << 16 with x = hi + lo
// (v_cvt_pk_bf16_f32 rounds to nearest even; x - hi is exact in FP32).
__device__ __forceinline__ void split2(float a, float b, unsigned& h, unsigned& l) {
  const float2v x = {a, b};
  const bf16x2 hb = __builtin_convertvector(x, bf16x2);
  const float2v hf = __builtin_convertvector(hb, float2v);
  const bf16x2 lb = __builtin_convertvector(x - hf, bf16x2);
  h = __builtin_bit_cast(unsigned, hb);
  l = __builtin_bit_cast(unsigned, lb);
}

__device__ __forceinline__ u16 neg(unsigned x) { return (u16)(x ^ 0x8000u); }

__device__ __forceinline__ float2 cmul(float2 d, float2 s) {
  return make_float2(d.x * s.x - d.y * s.y, d.x * s.y + d.y * s.x);
}

// a b - c conj(d)
__device__ __forceinline__ float2 border_term(float2 a, float2 b, float2 c, float2 d) {
  return make_float2(a.x * b.x - a.y * b.y - (c.x * d.x + c.y * d.y),
                     a.x * b.y + a.y * b.x - (c.y * d.x - c.x * d.y));
}

// Border coefficient operands (8 of them, each (hi, lo) x 2 rows of PITCH; operand m, part
// p, row r at row index 4 m + 2 p + r).  With h = g = H[NC][k], s = S[NC][k], c = S[k][NC]:
//   row outputs  C[NC][j]: A rows (Re coefficient; Im coefficient)
//     m = 0  [ hx;  hy] with B = Re S       m = 1  [-hy;  hx] with B = Im S
//     m = 2  [-sx; -sy] with B = Re H'      m = 3  [-sy;  sx] with B = Im H'   (H'[k][j] = H[j][k])
//   column outputs C[i][NC]: B columns (Re coefficient, Im coefficient)
//     m = 4  [ cx,  cy] with A = Re H       m = 5  [-cy,  cx] with A = Im H
//     m = 6  [-gx,  gy] with A = Re S       m = 7  [-gy, -gx] with A = Im S
template <int PITCH>
__device__ __forceinline__ void put_coef(u16* sK, int m, int k, u16 r0h, u16 r0l, u16 r1h, u16 r1l) {
  u16* p = sK + 4 * m * PITCH + k;
  p[0] = r0h;
  p[PITCH] = r1h;
  p[2 * PITCH] = r0l;
  p[3 * PITCH] = r1l;
}

// the constant operands (m = 0, 1, 6, 7) from H[NC][k]
template <int PITCH>
__device__ __forceinline__ void put_coef_h(u16* sK, int k, float2 v) {
  unsigned h, l;
  split2(v.x, v.y, h, l);
  const unsigned xh = h & 0xffffu, yh = h >> 16, xl = l & 0xffffu, yl = l >> 16;
  put_coef<PITCH>(sK, 0, k, xh, xl, yh, yl);
  put_coef<PITCH>(sK, 1, k, neg(yh), neg(yl), xh, xl);
  put_coef<PITCH>(sK, 6, k, neg(xh), neg(xl), yh, yl);
  put_coef<PITCH>(sK, 7, k, neg(yh), neg(yl), neg(xh), neg(xl));
}

// m = 2, 3 from S[NC][k]
template <int PITCH>
__device__ __forceinline__ void put_coef_srow(u16* sK, int k, float2 v) {
  unsigned h, l;
  split2(v.x, v.y, h, l);
  const unsigned xh = h & 0xffffu, yh = h >> 16, xl = l & 0xffffu, yl = l >> 16;
  put_coef<PITCH>(sK, 2, k, neg(xh), neg(xl), neg(yh), neg(yl));
  put_coef<PITCH>(sK, 3, k, neg(yh), neg(yl), xh, xl);
}

// m = 4, 5 from S[k][NC]
template <int PITCH>
__device__ __forceinline__ void put_coef_scol(u16* sK, int k, float2 v) {
  unsigned h, l;
  split2(v.x, v.y, h, l);
  const unsigned xh = h & 0xffffu, yh = h >> 16, xl = l & 0xffffu, yl = l >> 16;
  put_coef<PITCH>(sK, 4, k, xh, xl, yh, yl);
  put_coef<PITCH>(sK, 5, k, neg(yh), neg(yl), xh, xl);
}

// Wave roles.  n = 65: the 4 x 4 core tiles go two per wave to 8 waves, wave w taking
// tiles (ti, 2c) and (ti, 2c + 1) with ti = w / 2, c = w % 2, so both share the A
// fragments of tile row ti.  Each wave also forms one set of border outputs: the row
// outputs of tile column ti when c = ti / 2 (that column is one of its own, tile u = ti % 2,
// so the B fragments are shared), else the column outputs of tile row ti (A fragments
// shared).  The corner goes to lane 2 of wave 1 (a column-output wave, whose lanes
// lr >= 2 own nothing).
struct Bf3Layout {
  static constexpr int TG = 4;                       // core tile grid edge
  static constexpr int W = 8;                        // waves, two core tiles each
  static constexpr int NC = 64;                      // core size = border index (n = 65)
  static constexpr int KP = 64;                      // MFMA reduction extent
  // BF16 per plane row (multiples of 8: 16-byte rows for the fragment reads), chosen by
  // measured LDS bank conflicts: H_eff planes, S column planes, coefficient rows
  static constexpr int HP = KP + 16, SP = KP + 16, KPT = KP + 16;
  static constexpr int HPLANE = NC * HP, SPLANE = NC * SP;
  static constexpr int NCP = NC + 2;                 // side-array length (keeps 16-byte alignment)
  static constexpr int NBO = 2 * NC + 2;             // border-output slots (2 NC + 1 used)
  // one stage buffer: S column planes, the coefficient operands (32 rows), S[NC][:], S[:][NC]
  static constexpr size_t kStageBytes = (size_t)(4 * SPLANE + 32 * KPT) * 2 + (size_t)2 * NCP * 8;
  static constexpr size_t kHBytes = (size_t)4 * HPLANE * 2;   // H_eff row planes
  // H_eff planes, two stage buffers, H[:][NC] and H[NC][:], border-output state (FP64 rho,
  // FP32 increment, D)
  static constexpr size_t kFixedBytes = kHBytes + 2 * kStageBytes + (size_t)2 * NCP * 8 + (size_t)NBO * 32;
  static constexpr int CORNER_WAVE = 1, CORNER_LANE = 2;
};

struct Bf3Stage {   // pointers into one stage buffer
  u16* sC;
  u16* sK;
  float2* srow;
  float2* scol;
};

__device__ __forceinline__ Bf3Stage bf3_stage(unsigned char* base) {
  using L = Bf3Layout;
  Bf3Stage b;
  b.sC = reinterpret_cast<u16*>(base);
  b.sK = b.sC + 4 * L::SPLANE;
  b.srow = reinterpret_cast<float2*>(b.sK + 32 * L::KPT);
  b.scol = b.srow + L::NCP;
  return b;
}

typedef short short4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) short4v lds_short4;

// A-operand fragment of S rows (rows I0 + lr, k = k0 + 8 lg .. + 7) from a column plane:
// two transposed 8-byte reads, lane (lg, lr) loading plane rows k0 + 8 lg + lr / 4 (+ 4),
// positions I0 + 4 (lr % 4) .. + 3
__device__ __forceinline__ u16x8 frag_rows(const u16* plane, int off) {
  const short4v a = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_short4*)(plane + off));
  const short4v b = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_short4*)(plane + off + 4 * Bf3Layout::SP));
  return u16x8{(u16)a[0], (u16)a[1], (u16)a[2], (u16)a[3], (u16)b[0], (u16)b[1], (u16)b[2], (u16)b[3]};
}

}  // namespace

template <bool REC>
__global__ void __launch_bounds__(Bf3Layout::W * 64) __attribute__((amdgpu_waves_per_eu(2)))
lindblad_rk4_resident_bf3(ResidentArgs a) {
  using L = Bf3Layout;
  constexpr int W = L::W, NC = L::NC, KP = L::KP, HP = L::HP, SP = L::SP, KPT = L::KPT;
  constexpr int HPLANE = L::HPLANE, SPLANE = L::SPLANE;
  constexpr int T = W * 64;

  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  u16* sH = reinterpret_cast<u16*>(lds_raw);   // H_eff rows: planes Re hi, Re lo, Im hi, Im lo
  auto stage_buf = [&](int x) { return bf3_stage(lds_raw + L::kHBytes + x * L::kStageBytes); };
  float2* hcol = reinterpret_cast<float2*>(lds_raw + L::kHBytes + 2 * L::kStageBytes);   // H[r][NC]
  float2* hrow = hcol + L::NCP;                                                          // H[NC][k]
  // border outputs' state, one slot each (row output j: slot j; column output i: NC + i;
  // the corner: 2 NC), read and written only by the owning lane
  double2* bRho = reinterpret_cast<double2*>(hrow + L::NCP);
  float2* bInc = reinterpret_cast<float2*>(bRho + L::NBO);
  float2* bD = bInc + L::NBO;
  double2* sRed = reinterpret_cast<double2*>(lds_raw + L::kFixedBytes);

  const int b = xcd_block();   // the operators of one generator set run on one XCD (shared L2)
  if (b >= a.batch) return;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int n = a.n;   // NC + 1
  const int set = b / a.group;
  const int nn = n * n;
  const int lr = lane & 15;
  const int lg = lane >> 4;
  const int ti = w >> 1, c2 = w & 1;           // tile row, column pair (wave-uniform)
  const int I0 = 16 * ti, J0 = 32 * c2;        // the wave's tiles: columns J0 .. J0 + 31
  const bool rowsk = c2 == (ti >> 1);          // row outputs of tile column ti (tile u = ti & 1)
  const int usk = ti & 1;

  // ---- LDS: zero the planes and coefficient rows (k >= NC stays zero), then fill them
  {
    uint4* z = reinterpret_cast<uint4*>(lds_raw);
    for (int i = tid; i < (int)((L::kHBytes + 2 * L::kStageBytes) / 16); i += T) z[i] = make_uint4(0u, 0u, 0u, 0u);
  }
  __syncthreads();
  {
    const float2* H = a.heff + (size_t)set * nn;
    const double2* R = a.rho + (size_t)b * nn;
    const Bf3Stage s0 = stage_buf(0), s1 = stage_buf(1);
    for (int idx = tid; idx < NC * (NC / 2); idx += T) {   // element pairs (i, k), (i, k + 1)
      const int i = idx / (NC / 2), k = 2 * (idx - i * (NC / 2));
      unsigned xh, xl, yh, yl;
      const float2 h0 = H[i * n + k], h1 = H[i * n + k + 1];
      split2(h0.x, h1.x, xh, xl);
      split2(h0.y, h1.y, yh, yl);
      u16* hr = sH + i * HP + k;
      *reinterpret_cast<unsigned*>(hr) = xh;
      *reinterpret_cast<unsigned*>(hr + HPLANE) = xl;
      *reinterpret_cast<unsigned*>(hr + 2 * HPLANE) = yh;
      *reinterpret_cast<unsigned*>(hr + 3 * HPLANE) = yl;
      // S[k][i], S[k + 1][i]: column plane row i, positions k, k + 1
      const double2 r0 = R[k * n + i], r1 = R[(k + 1) * n + i];
      split2((float)r0.x, (float)r1.x, xh, xl);
      split2((float)r0.y, (float)r1.y, yh, yl);
      u16* cr = s0.sC + i * SP + k;
      *reinterpret_cast<unsigned*>(cr) = xh;
      *reinterpret_cast<unsigned*>(cr + SPLANE) = xl;
      *reinterpret_cast<unsigned*>(cr + 2 * SPLANE) = yh;
      *reinterpret_cast<unsigned*>(cr + 3 * SPLANE) = yl;
    }
    for (int k = tid; k <= NC; k += T) {
      const float2 hc = H[k * n + NC], hr = H[NC * n + k];
      const double2 rr = R[NC * n + k], rc = R[k * n + NC];
      const float2 sr = make_float2((float)rr.x, (float)rr.y), sc = make_float2((float)rc.x, (float)rc.y);
      hcol[k] = hc;
      hrow[k] = hr;
      s0.srow[k] = sr;
      s0.scol[k] = sc;
      if (k < NC) {
        put_coef_h<KPT>(s0.sK, k, hr);   // the constant operands, in both buffers
        put_coef_h<KPT>(s1.sK, k, hr);
        put_coef_srow<KPT>(s0.sK, k, sr);
        put_coef_scol<KPT>(s0.sK, k, sc);
      }
    }
  }

  // ---- border outputs owned by this lane (row outputs: lanes 0..15; column outputs: lanes
  // with lr < 2, rows 4 lg + 2 lr + u of tile row ti; the corner)
  // (of lane ln: output u's row bi, column bj and state slot bslot; nbo outputs)
  auto owns_row = [&](int ln) { return rowsk && ln < 16; };
  auto owns_col = [&](int ln) { return !rowsk && (ln & 15) < 2; };
  auto owns_corner = [&](int ln) { return w == L::CORNER_WAVE && ln == L::CORNER_LANE; };
  auto nbo_of = [&](int ln) { return owns_col(ln) ? 2 : (owns_row(ln) || owns_corner(ln)) ? 1 : 0; };
  auto bi_of = [&](int ln, int u) { return owns_col(ln) ? I0 + 4 * (ln >> 4) + 2 * (ln & 15) + u : NC; };
  auto bj_of = [&](int ln) { return owns_row(ln) ? I0 + ln : NC; };   // row outputs: columns of tile column ti
  auto bslot_of = [&](int ln, int u) {
    return owns_row(ln) ? I0 + ln : owns_col(ln) ? NC + I0 + 4 * (ln >> 4) + 2 * (ln & 15) + u : 2 * NC;
  };
  const bool own_row = owns_row(lane);
  const bool own_col = owns_col(lane);
  const bool own_corner = owns_corner(lane);
  const int nbo = nbo_of(lane);
  auto bi = [&](int u) { return bi_of(lane, u); };
  auto bj = [&](int u) { return bj_of(lane); };
  auto bslot = [&](int u) { return bslot_of(lane, u); };

  // FP64 state of the owned core elements (I0 + 4 lg + v, J0 + 16 u + lr), the FP32 RK4
  // increment and D there
  double2 rho[2][4];
  float2 inc[2][4];
  float2 dc[2][4];
  const bool has_d = a.dmat != nullptr;
  {
    const double2* R = a.rho + (size_t)b * nn;
    const float2* Dm = a.dmat + (size_t)set * nn;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int j = J0 + 16 * u + lr;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = I0 + 4 * lg + v;
        rho[u][v] = R[i * n + j];
        inc[u][v] = make_float2(0.f, 0.f);
        dc[u][v] = has_d ? Dm[i * n + j] : make_float2(0.f, 0.f);
      }
    }
    for (int u = 0; u < nbo; ++u) {
      const int x = bslot(u);
      bRho[x] = R[bi(u) * n + bj(u)];
      bInc[x] = make_float2(0.f, 0.f);
      bD[x] = has_d ? Dm[bi(u) * n + bj(u)] : make_float2(0.f, 0.f);
    }
  }
  __syncthreads();

  // accumulators of C: Re C = rp - rn, Im C = ip - in.  The stage derivative is k = -i C
  // (+ D o S folded in: C starts at i (D o S), i.e. Re C = -Im(D o S), Im C = Re(D o S))
  floatx4 rp[2], rn[2], ip[2], in[2], q;
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float2 p = cmul(dc[u][v], make_float2((float)rho[u][v].x, (float)rho[u][v].y));
      rp[u][v] = -p.y;
      ip[u][v] = p.x;
    }

  const double dt = a.dt;
  const bool renorm = (a.flags & 1) != 0;
  const bool record_any = REC && a.nout > 0 && (a.obs != nullptr || a.states != nullptr);
  const int aoff = (I0 + lr) * HP + 8 * lg;                    // H rows of tile row ti (A)
  const int bHoff = (J0 + lr) * HP + 8 * lg;                   // H rows of the wave's first tile column (B)
  const int bSoff = (J0 + lr) * SP + 8 * lg;                   // S columns of the wave's first tile column (B)
  const int troff = (8 * lg + (lr >> 2)) * SP + I0 + 4 * (lr & 3);   // S rows of tile row ti (A, transposed reads)
  const int koff = (lr & 1) * KPT + 8 * lg;    // coefficient rows (rows 2..15 repeat 0, 1:
                                                 // their outputs are not used)
  auto frag = [&](const u16* base, int off) { return *reinterpret_cast<const u16x8*>(base + off); };

  int ph = 0;   // buffer holding the current stage operand
  for (int s = 0; s < a.nsteps; ++s) {
#pragma unroll 1
    for (int st = 0; st < 4; ++st) {
      const Bf3Stage cur = stage_buf(ph), nxt = stage_buf(ph ^ 1);
      const u16* sC = cur.sC;
      const u16* sK = cur.sK;
      const float2* srow = cur.srow;
      const float2* scol = cur.scol;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        rn[u] = floatx4{0.f, 0.f, 0.f, 0.f};
        in[u] = floatx4{0.f, 0.f, 0.f, 0.f};
      }
      q = floatx4{0.f, 0.f, 0.f, 0.f};
      // the border index k = NC of the core outputs: H[i][NC] S[NC][j] - S[i][NC] conj(H[j][NC])
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const float2 sj = srow[J0 + 16 * u + lr], hj = hcol[J0 + 16 * u + lr];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = I0 + 4 * lg + v;
          const float2 t = border_term(hcol[i], sj, scol[i], hj);
          rp[u][v] += t.x;
          ip[u][v] += t.y;
        }
      }
      // The products of one stage, instantiated once per wave role (QK 0: column outputs of
      // tile row ti through coefficient columns; 1, 2: row outputs through coefficient rows,
      // sharing the B fragments of tile 0 or 1): the role is a compile-time constant inside,
      // so the 120 MFMAs and their fragment reads form one basic block that the scheduler
      // can interleave (wave-uniform branches per k-chunk cost 1 % of the launch).  The
      // products run at raised wave priority: of the two waves on a SIMD, one still issuing
      // MFMAs wins issue over one already in its epilogue (0.6 % of the launch).
      auto products = [&](auto qk) {
        constexpr int QK = decltype(qk)::value;
        __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int k0 = 0; k0 < KP; k0 += 32) {
          // H S: A = H rows of tile row ti (shared), B = S columns of the wave's two tiles
          {
            const u16x8 hxh = frag(sH, aoff + k0), hxl = frag(sH + HPLANE, aoff + k0);
            const u16x8 hyh = frag(sH + 2 * HPLANE, aoff + k0), hyl = frag(sH + 3 * HPLANE, aoff + k0);
            u16x8 sxh[2], sxl[2], syh[2], syl[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              const int o = bSoff + 16 * u * SP + k0;
              sxh[u] = frag(sC, o);
              sxl[u] = frag(sC + SPLANE, o);
              syh[u] = frag(sC + 2 * SPLANE, o);
              syl[u] = frag(sC + 3 * SPLANE, o);
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              rp[u] = mfma3(hxh, hxl, sxh[u], sxl[u], rp[u]);   // Hx Sx
              ip[u] = mfma3(hxh, hxl, syh[u], syl[u], ip[u]);   // Hx Sy
              rn[u] = mfma3(hyh, hyl, syh[u], syl[u], rn[u]);   // Hy Sy
              ip[u] = mfma3(hyh, hyl, sxh[u], sxl[u], ip[u]);   // Hy Sx
            }
            if constexpr (QK != 0) {
              const u16x8 k0h = frag(sK, koff + k0), k0l = frag(sK, koff + 2 * KPT + k0);
              const u16x8 k1h = frag(sK, koff + 4 * KPT + k0), k1l = frag(sK, koff + 6 * KPT + k0);
              if constexpr (QK == 1) {
                q = mfma3(k0h, k0l, sxh[0], sxl[0], q);
                q = mfma3(k1h, k1l, syh[0], syl[0], q);
              } else {
                q = mfma3(k0h, k0l, sxh[1], sxl[1], q);
                q = mfma3(k1h, k1l, syh[1], syl[1], q);
              }
            } else {
              q = mfma3(hxh, hxl, frag(sK, koff + 16 * KPT + k0), frag(sK, koff + 18 * KPT + k0), q);
              q = mfma3(hyh, hyl, frag(sK, koff + 20 * KPT + k0), frag(sK, koff + 22 * KPT + k0), q);
            }
          }
          // S H^dag: A = S rows of tile row ti (shared; transposed reads of the column planes),
          // B[k][j] = conj(H[j][k]) from H rows of the wave's tile columns
          {
            const int to = troff + k0 * SP;
            const u16x8 zxh = frag_rows(sC, to), zxl = frag_rows(sC + SPLANE, to);
            const u16x8 zyh = frag_rows(sC + 2 * SPLANE, to), zyl = frag_rows(sC + 3 * SPLANE, to);
            u16x8 gxh[2], gxl[2], gyh[2], gyl[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              const int o = bHoff + 16 * u * HP + k0;
              gxh[u] = frag(sH, o);
              gxl[u] = frag(sH + HPLANE, o);
              gyh[u] = frag(sH + 2 * HPLANE, o);
              gyl[u] = frag(sH + 3 * HPLANE, o);
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              rn[u] = mfma3(zxh, zxl, gxh[u], gxl[u], rn[u]);   // Sx Hx'
              ip[u] = mfma3(zxh, zxl, gyh[u], gyl[u], ip[u]);   // Sx Hy'
              rn[u] = mfma3(zyh, zyl, gyh[u], gyl[u], rn[u]);   // Sy Hy'
              in[u] = mfma3(zyh, zyl, gxh[u], gxl[u], in[u]);   // Sy Hx'
            }
            if constexpr (QK != 0) {
              const u16x8 k2h = frag(sK, koff + 8 * KPT + k0), k2l = frag(sK, koff + 10 * KPT + k0);
              const u16x8 k3h = frag(sK, koff + 12 * KPT + k0), k3l = frag(sK, koff + 14 * KPT + k0);
              if constexpr (QK == 1) {
                q = mfma3(k2h, k2l, gxh[0], gxl[0], q);
                q = mfma3(k3h, k3l, gyh[0], gyl[0], q);
              } else {
                q = mfma3(k2h, k2l, gxh[1], gxl[1], q);
                q = mfma3(k3h, k3l, gyh[1], gyl[1], q);
              }
            } else {
              q = mfma3(zxh, zxl, frag(sK, koff + 24 * KPT + k0), frag(sK, koff + 26 * KPT + k0), q);
              q = mfma3(zyh, zyl, frag(sK, koff + 28 * KPT + k0), frag(sK, koff + 30 * KPT + k0), q);
            }
          }
        }
        __builtin_amdgcn_s_setprio(0);
      };
      if (!rowsk) products(RoleK<0>{});
      else if (usk == 0) products(RoleK<1>{});
      else products(RoleK<2>{});
      // C at the owned border outputs (before the border index term and D o S)
      float bre[2] = {0.f, 0.f}, bim[2] = {0.f, 0.f};
      if (rowsk) {   // lanes 0..15: Re C[NC][j] in row 0, Im in row 1
        bre[0] = q[0];
        bim[0] = q[1];
      } else {
        // lane lr = 0 holds Re C[I0 + 4 lg + v][NC] in q[v], lane lr = 1 the Im part; the
        // even lane owns rows v = 0, 1, the odd lane rows v = 2, 3
        const float o0 = dpp_f32<0xB1>(q[0]), o1 = dpp_f32<0xB1>(q[1]);   // lane l ^ 1
        const float o2 = dpp_f32<0xB1>(q[2]), o3 = dpp_f32<0xB1>(q[3]);
        const bool odd = (lr & 1) != 0;
        bre[0] = odd ? o2 : q[0];
        bim[0] = odd ? q[2] : o0;
        bre[1] = odd ? o3 : q[1];
        bim[1] = odd ? q[3] : o1;
      }
      if (w == L::CORNER_WAVE) {   // the corner over k < NC: one k per lane, then a wave sum
        float qr = 0.f, qm = 0.f;
        for (int k = lane; k < NC; k += 64) {
          const float2 t = border_term(hrow[k], scol[k], srow[k], hrow[k]);
          qr += t.x;
          qm += t.y;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
          qr += __shfl_xor(qr, off, 64);
          qm += __shfl_xor(qm, off, 64);
        }
        if (own_corner) {
          bre[0] = qr;
          bim[0] = qm;
        }
      }
      // the border index term and D o S of the border outputs: k_b = -i C_b + D o S
      float2 kb[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        kb[u] = make_float2(0.f, 0.f);
        if (u < nbo) {
          const int r = bi(u), c = bj(u);
          const float2 t = border_term(hcol[r], srow[c], scol[r], hcol[c]);
          const float2 sv = r == NC ? srow[c] : scol[r];
          const float2 p = cmul(bD[bslot(u)], sv);
          kb[u] = make_float2(bim[u] + t.y + p.x, -(bre[u] + t.x) + p.y);
        }
      }
      // (no barrier: the next stage operand goes to the other buffer)

      const float cacc = (float)((st == 0 || st == 3) ? dt / 6.0 : dt / 3.0);
      float2 sv[2][4], svb[2];
      if (st < 3) {
        // next stage operand S = rho + c k, formed in FP32 (it is split into BF16 pairs, 2^-17
        // relative, right after; the state itself stays FP64)
        const double cnext = (st == 2) ? dt : 0.5 * dt;
        const float cnf = (float)cnext;
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float kr = ip[u][v] - in[u][v], ki = -(rp[u][v] - rn[u][v]);
            inc[u][v].x = fmaf(cacc, kr, inc[u][v].x);
            inc[u][v].y = fmaf(cacc, ki, inc[u][v].y);
            sv[u][v] = make_float2(fmaf(cnf, kr, (float)rho[u][v].x), fmaf(cnf, ki, (float)rho[u][v].y));
          }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          svb[u] = make_float2(0.f, 0.f);
          if (u < nbo) {
            const int x = bslot(u);
            const double2 r = bRho[x];
            float2 ib = bInc[x];
            ib.x = fmaf(cacc, kb[u].x, ib.x);
            ib.y = fmaf(cacc, kb[u].y, ib.y);
            bInc[x] = ib;
            svb[u] = make_float2((float)(r.x + cnext * (double)kb[u].x), (float)(r.y + cnext * (double)kb[u].y));
          }
        }
      } else {
        // last stage: the state update itself is FP64
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float kr = ip[u][v] - in[u][v], ki = -(rp[u][v] - rn[u][v]);
            rho[u][v].x += (double)fmaf(cacc, kr, inc[u][v].x);
            rho[u][v].y += (double)fmaf(cacc, ki, inc[u][v].y);
            inc[u][v] = make_float2(0.f, 0.f);   // the next step's increment starts from zero
          }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          if (u < nbo) {
            const int x = bslot(u);
            double2 r = bRho[x];
            const float2 ib = bInc[x];
            r.x += (double)fmaf(cacc, kb[u].x, ib.x);
            r.y += (double)fmaf(cacc, kb[u].y, ib.y);
            bRho[x] = r;
            bInc[x] = make_float2(0.f, 0.f);
          }
        }
        if (renorm) {
          double tr = 0.0;
#pragma unroll
          for (int u = 0; u < 2; ++u)
            if (J0 + 16 * u == I0) {
#pragma unroll
              for (int v = 0; v < 4; ++v)
                if (4 * lg + v == lr) tr += rho[u][v].x;
            }
          if (own_corner) tr += bRho[2 * NC].x;
          tr = wave_sum(tr);
          if (lane == 0) sRed[w].x = tr;
          __syncthreads();
          double tot = 0.0;
          for (int x = 0; x < W; ++x) tot += sRed[x].x;
          const double inv = 1.0 / tot;
#pragma unroll
          for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              rho[u][v].x *= inv;
              rho[u][v].y *= inv;
            }
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            if (u < nbo) {
              double2 r = bRho[bslot(u)];
              r.x *= inv;
              r.y *= inv;
              bRho[bslot(u)] = r;
            }
          }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) sv[u][v] = make_float2((float)rho[u][v].x, (float)rho[u][v].y);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const double2 r = u < nbo ? bRho[bslot(u)] : make_double2(0.0, 0.0);
          svb[u] = make_float2((float)r.x, (float)r.y);
        }
      }

      // ---- write the next stage operand into the other buffer: core elements (4
      // consecutive rows of one column per plane), border outputs into the side arrays and
      // the coefficient operands; D o S into the accumulators
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        unsigned xh0, xl0, yh0, yl0, xh1, xl1, yh1, yl1;
        split2(sv[u][0].x, sv[u][1].x, xh0, xl0);
        split2(sv[u][0].y, sv[u][1].y, yh0, yl0);
        split2(sv[u][2].x, sv[u][3].x, xh1, xl1);
        split2(sv[u][2].y, sv[u][3].y, yh1, yl1);
        u16* c0 = nxt.sC + (J0 + 16 * u + lr) * SP + I0 + 4 * lg;   // column plane row j, positions i..i+3
        *reinterpret_cast<uint2*>(c0) = make_uint2(xh0, xh1);
        *reinterpret_cast<uint2*>(c0 + SPLANE) = make_uint2(xl0, xl1);
        *reinterpret_cast<uint2*>(c0 + 2 * SPLANE) = make_uint2(yh0, yh1);
        *reinterpret_cast<uint2*>(c0 + 3 * SPLANE) = make_uint2(yl0, yl1);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float2 p = cmul(dc[u][v], sv[u][v]);
          rp[u][v] = -p.y;
          ip[u][v] = p.x;
        }
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        if (u < nbo) {
          const int r = bi(u), c = bj(u);
          if (r == NC) {
            nxt.srow[c] = svb[u];
            if (c < NC) put_coef_srow<KPT>(nxt.sK, c, svb[u]);
          }
          if (c == NC) {
            nxt.scol[r] = svb[u];
            if (r < NC) put_coef_scol<KPT>(nxt.sK, r, svb[u]);
          }
        }
      }
      if (st < 3) {
        __syncthreads();   // S of the next stage is complete
        ph ^= 1;
        continue;
      }

      // ---------------- end of step: records
      const int g = a.step_offset + s + 1;
      const bool rec_now = record_any && (g % a.nout) == 0 && (g / a.nout - 1) < a.rec_capacity;
      if (rec_now) {
        const int rec = g / a.nout - 1;
        int olr = lr, olg = lg, ob0 = bi(0) * n + bj(0), ob1 = bi(1) * n + bj(1);
        asm volatile("" : "+v"(olr), "+v"(olg), "+v"(ob0), "+v"(ob1));   // formed here, not held
        const int ob[2] = {ob0, ob1};
        if (a.obs != nullptr) {
          const double2* E0 = a.eops_t + (size_t)(b / a.eop_group) * a.n_eops * nn;
          for (int e = 0; e < a.n_eops; ++e) {
            const double2* E = E0 + (size_t)e * nn;
            double pr = 0.0, pi = 0.0;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              double2 ev[4];
#pragma unroll
              for (int v = 0; v < 4; ++v) ev[v] = E[(I0 + v) * n + J0 + 16 * u + 4 * olg * n + olr];
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                pr += ev[v].x * rho[u][v].x - ev[v].y * rho[u][v].y;
                pi += ev[v].x * rho[u][v].y + ev[v].y * rho[u][v].x;
              }
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              if (u < nbo) {
                const double2 x = E[ob[u]];
                const double2 r = bRho[bslot(u)];
                pr += x.x * r.x - x.y * r.y;
                pi += x.x * r.y + x.y * r.x;
              }
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
          for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) Sg[(I0 + v) * n + J0 + 16 * u + 4 * olg * n + olr] = rho[u][v];
#pragma unroll
          for (int u = 0; u < 2; ++u)
            if (u < nbo) Sg[ob[u]] = bRho[bslot(u)];
        }
      }
      __syncthreads();   // S of the next step (and the record partials) are complete
      ph ^= 1;
      if (rec_now && a.obs != nullptr && w == W - 1) {
        const int rec = g / a.nout - 1;
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
    int wlr = lr, wlg = lg;   // opaque: the addresses are formed here, not held from the start
    asm volatile("" : "+v"(wlr), "+v"(wlg));
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) R[(I0 + 4 * wlg + v) * n + J0 + 16 * u + wlr] = rho[u][v];
    int wl = lane;   // the same for the border outputs
    asm volatile("" : "+v"(wl));
#pragma unroll
    for (int u = 0; u < 2; ++u)
      if (u < nbo_of(wl)) R[bi_of(wl, u) * n + bj_of(wl)] = bRho[bslot_of(wl, u)];
  }
}

namespace {

template <bool REC>
hipError_t launch_bf3_one(const ResidentArgs& a, size_t lds, hipStream_t stream) {
  auto k = lindblad_rk4_resident_bf3<REC>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(a.batch), dim3(Bf3Layout::W * 64), lds, stream, a);
  return hipGetLastError();
}

}  // namespace

bool resident_bf3_eligible(int n, int n_dense) { return n_dense == 0 && n == 65; }

size_t resident_bf3_lds_bytes(int n, int n_eops) {
  if (n != 65) return 0;
  return Bf3Layout::kFixedBytes + (size_t)(1 + n_eops) * Bf3Layout::W * sizeof(double2);
}

hipError_t resident_bf3_launch(const ResidentArgs& a, hipStream_t stream) {
  if (!resident_bf3_eligible(a.n, a.n_dense) || a.rops != nullptr) return hipErrorInvalidValue;
  const size_t lds = resident_bf3_lds_bytes(a.n, a.n_eops);
  const bool rec = a.nout > 0 && (a.obs != nullptr || a.states != nullptr);
  return rec ? launch_bf3_one<true>(a, lds, stream) : launch_bf3_one<false>(a, lds, stream);
}

}  // namespace lime
