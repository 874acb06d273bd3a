// Register-resident Stockham FFT for gfx950 (lengths 2^k and 3 * 2^k).
//
// Twiddles come from the FP64-built bank exp(-2 pi i t / L) (twiddle_bank,
// offset tw_offset(L)), so a length-L transform has the accuracy of its radix
// arithmetic only.
#pragma once

#include <type_traits>

#include "mm_common.h"

namespace mm {

constexpr int kBatchElems = 16384;  // complex points per full-size FFT workgroup (128 KiB of LDS)

// LDS index of element i of a transform: one pad slot per 16 complex values,
// so strided Stockham writes hit distinct banks.
__host__ __device__ __forceinline__ constexpr int lpad(int i) { return i + (i >> 4); }

// Row-pass workgroups hold half as many points (64 KiB + pad of LDS), so two of
// them share a CU and one's global loads overlap the other's LDS stages.
constexpr int kRowBatch = 8192;

template <bool INV>
__host__ __device__ __forceinline__ void dft2(float2& a, float2& b) {
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <bool INV>
__host__ __device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_mi<INV>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <bool INV>
__host__ __device__ __forceinline__ void dft8(float2* a) {
  float2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  float2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4<INV>(e0, e1, e2, e3);
  dft4<INV>(o0, o1, o2, o3);
  const float c = 0.70710678118654752440f;
  // W8^k o_k (forward W8 = exp(-i pi/4), inverse its conjugate)
  float2 w1 = INV ? make_float2(c * (o1.x - o1.y), c * (o1.x + o1.y)) : make_float2(c * (o1.x + o1.y), c * (o1.y - o1.x));
  float2 w2 = mul_mi<INV>(o2);
  float2 w3 = INV ? make_float2(-c * (o3.x + o3.y), c * (o3.x - o3.y)) : make_float2(c * (o3.y - o3.x), -c * (o3.x + o3.y));
  a[0] = cadd(e0, o0);
  a[4] = csub(e0, o0);
  a[1] = cadd(e1, w1);
  a[5] = csub(e1, w1);
  a[2] = cadd(e2, w2);
  a[6] = csub(e2, w2);
  a[3] = cadd(e3, w3);
  a[7] = csub(e3, w3);
}

// 3-point DFT: y1,2 = a - (b + c)/2 -/+ i (sqrt 3 / 2)(b - c) (forward; INV conjugates).
template <bool INV>
__host__ __device__ __forceinline__ void dft3(float2& a, float2& b, float2& c) {
  constexpr float H3 = 0.86602540378443864676f;
  const float2 s = cadd(b, c), d = csub(b, c);
  const float2 m = make_float2(a.x - 0.5f * s.x, a.y - 0.5f * s.y);
  const float2 r = cscale(mul_mi<INV>(d), H3);
  a = cadd(a, s);
  b = cadd(m, r);
  c = csub(m, r);
}

// W12^e times a (forward kernel exp(-2 pi i e / 12); INV conjugates).
template <int E, bool INV>
__host__ __device__ __forceinline__ float2 w12(float2 a) {
  constexpr float H3 = 0.86602540378443864676f;
  constexpr int e = E % 12;
  float wr, wi;
  if constexpr (e == 0) return a;
  else if constexpr (e == 3) return mul_mi<INV>(a);
  else if constexpr (e == 6) return make_float2(-a.x, -a.y);
  else {
    if constexpr (e == 1) { wr = H3; wi = -0.5f; }
    else if constexpr (e == 2) { wr = 0.5f; wi = -H3; }
    else { static_assert(e == 4, "exponents of the 3 x 4 split"); wr = -0.5f; wi = -H3; }
    if (INV) wi = -wi;
    return make_float2(a.x * wr - a.y * wi, a.x * wi + a.y * wr);
  }
}

// 12-point DFT as 3 x 4: n = n2 + 4 n1, k = k1 + 3 k2;
// X[k1 + 3 k2] = sum_n2 W4^(n2 k2) W12^(n2 k1) sum_n1 W3^(n1 k1) x[n2 + 4 n1].
template <bool INV>
__host__ __device__ __forceinline__ void dft12(float2* a) {
  float2 b[4][3];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    float2 t0 = a[n2], t1 = a[n2 + 4], t2 = a[n2 + 8];
    dft3<INV>(t0, t1, t2);
    b[n2][0] = t0;
    b[n2][1] = t1;
    b[n2][2] = t2;
  }
  b[1][1] = w12<1, INV>(b[1][1]);
  b[1][2] = w12<2, INV>(b[1][2]);
  b[2][1] = w12<2, INV>(b[2][1]);
  b[2][2] = w12<4, INV>(b[2][2]);
  b[3][1] = w12<3, INV>(b[3][1]);
  b[3][2] = w12<6, INV>(b[3][2]);
#pragma unroll
  for (int k1 = 0; k1 < 3; ++k1) {
    float2 u0 = b[0][k1], u1 = b[1][k1], u2 = b[2][k1], u3 = b[3][k1];
    dft4<INV>(u0, u1, u2, u3);
    a[k1] = u0;
    a[k1 + 3] = u1;
    a[k1 + 6] = u2;
    a[k1 + 9] = u3;
  }
}

template <int R, bool INV>
__host__ __device__ __forceinline__ void dft(float2* v) {
  if constexpr (R == 2) {
    dft2<INV>(v[0], v[1]);
  } else if constexpr (R == 4) {
    dft4<INV>(v[0], v[1], v[2], v[3]);
  } else {
    static_assert(R == 8, "radix 2, 4 or 8");
    dft8<INV>(v);
  }
}

// Lengths: 2^k (16 points per thread, radix-16 passes) and 3 * 2^k (24 points
// per thread from 768 up: one radix-12 pass of two butterflies per thread,
// then radix-8 passes and a radix-4 or radix-2 for the leftover, so a
// 6144-point transform makes three LDS exchanges; 12 points per thread below
// 768: radix-12, then radix-4 / radix-2).
__host__ __device__ constexpr bool pow2c(int L) { return L > 0 && (L & (L - 1)) == 0; }
__host__ __device__ constexpr bool fft_len_ok(int L) { return pow2c(L) || (L % 3 == 0 && pow2c(L / 3)); }
// 3 * 2^k: 24 points per thread from 768 points up (radix-8 passes after the
// radix-12 one), 12 below (short transforms keep 16+ threads per transform).
__host__ __device__ constexpr int pt_of(int L) { return pow2c(L) ? 16 : (L >= 768 ? 24 : 12); }
__host__ __device__ constexpr int next_radix(int L, int NS) {
  return pow2c(L) ? (L / NS >= 16 ? 16 : L / NS)
                  : (NS == 1 ? 12
                             : ((L / NS) % 8 == 0 && pt_of(L) % 8 == 0 ? 8 : ((L / NS) % 4 == 0 ? 4 : 2)));
}

// ---- register-resident Stockham FFT -----------------------------------------
// Thread t (0 <= t < TPC = L / PT) of a transform holds v[m] = x[t + m TPC],
// m < PT (PT = pt_of(L): 16, or 12 / 24 for 3 * 2^k), before and after the
// transform.  A radix-R pass (R divides PT) runs the thread's PT / R
// butterflies j = t + b TPC on its own registers (butterfly j reads
// x[j + s L / R] = v[b + s PT / R]); its outputs go to LDS at their Stockham
// positions and come back in natural order for the next pass.  The last
// pass's Stockham positions are the natural ones, so its outputs stay in
// registers: a length 16^a r transform (r in {1, 2, 4, 8}) costs
// a + [r > 1] - 1 LDS exchanges.  Global loads and stores of v are direct
// (element t + m TPC: consecutive t, consecutive addresses).

// Workgroup shape: C transforms of length L (at most 1024 threads), stride S
// between them in LDS (room for XTRA more values per transform, for kernels
// that stage more than L inputs there).  S = 16 (mod 32) complex values, with
// the interleaved thread mapping of map(), keeps the Stockham writes and
// natural reads of power-of-two lengths free of bank conflicts (two
// transforms per 32-lane half-wave at a 16-bank-pair offset).  The plain
// mapping (64 consecutive t of one transform per wave) reads rows in longer
// contiguous runs; it leaves some 2-way conflicts on 64-bank LDS.
template <int L, int BATCH, int XTRA = 0, bool INTERLEAVE = true>
struct RegShape {
  static_assert(L >= 16 && L <= kMaxFFT && fft_len_ok(L), "length 2^k or 3 * 2^k in [16, 16384]");
  static constexpr int PT = pt_of(L);
  static constexpr int TPC = L / PT;
  // interleaved shapes take whole groups of 4 transforms (the map below; column kernels: whole blocks)
  static constexpr int CB = BATCH > L ? BATCH / L : 1;
  static constexpr int C0 = INTERLEAVE && L / PT >= 16 && CB >= 4 ? CB / 4 * 4 : CB;
  static constexpr int C = C0 * TPC > 1024 ? 1024 / TPC : C0;
  static constexpr int NT = C * TPC;
  static constexpr int S0 = L + L / 16 + XTRA;  // > lpad(L - 1) + XTRA
  static constexpr int S = TPC < 16 ? S0 : S0 + ((16 - S0 % 32) + 32) % 32;  // S = 16 (mod 32)
  static constexpr int LDS = C * S;
  // Thread -> (transform c, index t).  Interleaved (TPC >= 16): a wave holds
  // 16 consecutive t of min(C, 4) transforms, so loads of interleaved columns
  // (4 per 32-byte row segment) are contiguous across the wave.
  __device__ static __forceinline__ void map(int tid, int& c, int& t) {
    if constexpr (TPC < 16 || !INTERLEAVE) {
      c = tid / TPC;
      t = tid % TPC;
    } else {
      constexpr int G = C < 4 ? C : 4, PER = TPC / 16;
      const int g = tid >> 4, tg = g / G;
      c = g % G + G * (tg / PER);
      t = (tid & 15) + 16 * (tg % PER);
    }
  }
};

// Register budget of kernels that run two transforms or two stages (Fresnel,
// Paganin, FBP): at least 4 waves per SIMD (<= 128 VGPRs), so that two
// 512-lane workgroups share a CU; left free, the compiler keeps both
// transforms' twiddle loads in flight and drops to one workgroup per CU.
#define MM_TWO_FFT_BUDGET __attribute__((amdgpu_waves_per_eu(4)))

// The twiddle table pointer and thread index for a kernel's second transform, opaque to the compiler:
// otherwise it reuses the first transform's twiddle loads and LDS addresses and keeps them live across
// both (past the register budget).
__device__ __forceinline__ const float2* opaque_twiddles(const float2* tw) {
  asm volatile("" : "+s"(tw));
  return tw;
}
__device__ __forceinline__ int opaque_index(int t) {
  asm volatile("" : "+v"(t));
  return t;
}

// Column-fastest lane mapping for column transforms (columns of a row-major array, row stride = the array
// width): consecutive lanes on consecutive columns of one row, so a wave's load or store is one contiguous
// run of its workgroup's C columns (8-byte pieces 16 lanes apart, on 16 different rows, in the interleaved
// mapping of RegShape).  With the lanes of one wave on different transforms at stride S in LDS, an odd S
// keeps every 32-lane 8-byte LDS access on distinct bank pairs.
template <int L, int BATCH>
struct ColFastShape : RegShape<L, BATCH> {
  using Base = RegShape<L, BATCH>;
  static constexpr int S = Base::S0 | 1;
  static constexpr int LDS = Base::C * S;
  __device__ static __forceinline__ void map(int tid, int& c, int& t) {
    c = tid % Base::C;
    t = tid / Base::C;
  }
};
// Column transforms of fewer than 16 threads or of 2048 points and more run column-fastest; the others keep
// the interleaved mapping.
template <int L, int BATCH>
using ColumnShape = std::conditional_t<(RegShape<L, BATCH>::TPC < 16 || L >= 2048), ColFastShape<L, BATCH>,
                                       RegShape<L, BATCH>>;

// Row-pass shape of the register FFT: 8192 points per workgroup, plain mapping.
template <int L, int XTRA = 0>
using RowReg = RegShape<L, kRowBatch, XTRA, false>;

// W16^e times a (forward kernel exp(-2 pi i e / 16); INV conjugates).
template <int E, bool INV>
__host__ __device__ __forceinline__ float2 w16(float2 a) {
  constexpr float C1 = 0.92387953251128675613f, S1 = 0.38268343236508977173f, R2 = 0.70710678118654752440f;
  constexpr int e = E % 16;
  float wr, wi;
  if constexpr (e == 0) return a;
  else if constexpr (e == 4) return mul_mi<INV>(a);
  else {
    if constexpr (e == 1) { wr = C1; wi = -S1; }
    else if constexpr (e == 2) { wr = R2; wi = -R2; }
    else if constexpr (e == 3) { wr = S1; wi = -C1; }
    else if constexpr (e == 6) { wr = -R2; wi = -R2; }
    else { static_assert(e == 9, "exponents of the 4 x 4 split"); wr = -C1; wi = S1; }
    if (INV) wi = -wi;
    return make_float2(a.x * wr - a.y * wi, a.x * wi + a.y * wr);
  }
}

// 16-point DFT as 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2.
template <bool INV>
__host__ __device__ __forceinline__ void dft16(float2* a) {
  float2 b[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    float2 t0 = a[n2], t1 = a[4 + n2], t2 = a[8 + n2], t3 = a[12 + n2];
    dft4<INV>(t0, t1, t2, t3);
    b[n2][0] = t0;
    b[n2][1] = t1;
    b[n2][2] = t2;
    b[n2][3] = t3;
  }
  b[1][1] = w16<1, INV>(b[1][1]);
  b[1][2] = w16<2, INV>(b[1][2]);
  b[1][3] = w16<3, INV>(b[1][3]);
  b[2][1] = w16<2, INV>(b[2][1]);
  b[2][2] = w16<4, INV>(b[2][2]);
  b[2][3] = w16<6, INV>(b[2][3]);
  b[3][1] = w16<3, INV>(b[3][1]);
  b[3][2] = w16<6, INV>(b[3][2]);
  b[3][3] = w16<9, INV>(b[3][3]);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 u0 = b[0][k1], u1 = b[1][k1], u2 = b[2][k1], u3 = b[3][k1];
    dft4<INV>(u0, u1, u2, u3);
    a[k1] = u0;
    a[k1 + 4] = u1;
    a[k1 + 8] = u2;
    a[k1 + 12] = u3;
  }
}

// u[s] *= W_{NS R}^{s k}, s < R.  The radix-16 pass after a radix-16 pass
// (NS R = 256) reads all 15 from the W_256 grid (one 128-byte line per e for a
// wave's 16 values of k); other passes read the powers 1, 2, 4, 8 from the
// length's table and form the rest as products of at most three of them.
template <int L, int R, int NS, bool INV, bool GRID>
__device__ __forceinline__ void twiddle(float2* u, int k, const float2* __restrict__ tw) {
  if constexpr (GRID && R == 16 && NS == 16 && pow2c(L)) {
    // W_256^(e k) straight from the grid (tw points at tw_offset(L) of the bank)
    const float2* g = tw + (kTwGrid256 - tw_offset(L)) + k;
#pragma unroll
    for (int e = 1; e < 16; ++e) {
      float2 w = g[16 * e];
      if (INV) w.y = -w.y;
      u[e] = cmul(u[e], w);
    }
    return;
  }
  constexpr int STEP = L / (NS * R);
  auto ld = [&](int e) {
    float2 w = tw[e * k * STEP];
    if (INV) w.y = -w.y;
    return w;
  };
  float2 w[16];
  w[1] = ld(1);
  if constexpr (R > 2) {
    w[2] = ld(2);
    w[3] = cmul(w[1], w[2]);
  }
  if constexpr (R > 4) {
    w[4] = ld(4);
#pragma unroll
    for (int e = 5; e < 8; ++e) w[e] = cmul(w[e - 4], w[4]);
  }
  if constexpr (R > 8) {
    w[8] = ld(8);
#pragma unroll
    for (int e = 9; e < 16; ++e) w[e] = cmul(w[e - 8], w[8]);
  }
#pragma unroll
  for (int e = 1; e < R; ++e) u[e] = cmul(u[e], w[e]);
}

// One radix-R pass of fft_reg (NS = product of the radices already applied).
template <int L, int R, int NS, int TPC, bool INV, int PT, bool GRID>
__device__ __forceinline__ void reg_pass(float2 (&v)[PT], float2* __restrict__ x, int t,
                                         const float2* __restrict__ tw) {
  static_assert(PT % R == 0, "the radix divides the points per thread");
  constexpr int BPT = PT / R;
  constexpr bool LAST = NS * R == L;
  float2 u[BPT][R];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = t + b * TPC;
#pragma unroll
    for (int s = 0; s < R; ++s) u[b][s] = v[b + s * BPT];
    if constexpr (NS > 1) twiddle<L, R, NS, INV, GRID>(u[b], j % NS, tw);
    if constexpr (R == 16)
      dft16<INV>(u[b]);
    else if constexpr (R == 12)
      dft12<INV>(u[b]);
    else
      dft<R, INV>(u[b]);
  }
  if constexpr (LAST) {
#pragma unroll
    for (int b = 0; b < BPT; ++b)
#pragma unroll
      for (int q = 0; q < R; ++q) v[b + q * BPT] = u[b][q];
  } else {
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = t + b * TPC;
      const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
      for (int q = 0; q < R; ++q) x[lpad(base + q * NS)] = u[b][q];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < PT; ++m) v[m] = x[lpad(t + m * TPC)];
    __syncthreads();
  }
}

// Transform of the PT points each of the TPC threads of one transform holds
// in v (natural order in and out), x = that transform's LDS (RegShape stride).
// Contains barriers when L > PT: every thread of the workgroup must call it.
// Unnormalised in both directions (INV only conjugates the kernel).  GRID = false
// keeps the product twiddles in every pass (kernels that run two transforms:
// the grid's 15 loads per pass raise their registers past two workgroups per CU).
template <int L, int TPC, bool INV, int NS = 1, bool GRID = true, int PT>
__device__ __forceinline__ void fft_reg(float2 (&v)[PT], float2* __restrict__ x, int t,
                                        const float2* __restrict__ tw) {
  static_assert(PT == pt_of(L) && TPC * PT == L, "register FFT shape");
  if constexpr (NS < L) {
    constexpr int R = next_radix(L, NS);
    reg_pass<L, R, NS, TPC, INV, PT, GRID>(v, x, t, tw);
    fft_reg<L, TPC, INV, NS * R, GRID>(v, x, t, tw);
  }
}

}  // namespace mm

// Runs CALL with constexpr int LL = L_RUNTIME for every FFT length the kernels
// are built for (2^k in [16, 16384], 3 * 2^k in [48, 12288]); any other length
// returns MM_ERR_UNSUPPORTED from the enclosing function.
#define MM_DISPATCH_FFT(L_RUNTIME, CALL)                                                             \
  switch (L_RUNTIME) {                                                                               \
    case 16: { constexpr int LL = 16; CALL; } break;                                                 \
    case 32: { constexpr int LL = 32; CALL; } break;                                                 \
    case 64: { constexpr int LL = 64; CALL; } break;                                                 \
    case 128: { constexpr int LL = 128; CALL; } break;                                               \
    case 256: { constexpr int LL = 256; CALL; } break;                                               \
    case 512: { constexpr int LL = 512; CALL; } break;                                               \
    case 1024: { constexpr int LL = 1024; CALL; } break;                                             \
    case 2048: { constexpr int LL = 2048; CALL; } break;                                             \
    case 4096: { constexpr int LL = 4096; CALL; } break;                                             \
    case 8192: { constexpr int LL = 8192; CALL; } break;                                             \
    case 16384: { constexpr int LL = 16384; CALL; } break;                                           \
    case 48: { constexpr int LL = 48; CALL; } break;                                                 \
    case 96: { constexpr int LL = 96; CALL; } break;                                                 \
    case 192: { constexpr int LL = 192; CALL; } break;                                               \
    case 384: { constexpr int LL = 384; CALL; } break;                                               \
    case 768: { constexpr int LL = 768; CALL; } break;                                               \
    case 1536: { constexpr int LL = 1536; CALL; } break;                                             \
    case 3072: { constexpr int LL = 3072; CALL; } break;                                             \
    case 6144: { constexpr int LL = 6144; CALL; } break;                                             \
    case 12288: { constexpr int LL = 12288; CALL; } break;                                           \
    default: return ::mm::fail(MM_ERR_UNSUPPORTED, "FFT length " + std::to_string(L_RUNTIME) + " not built"); \
  }
