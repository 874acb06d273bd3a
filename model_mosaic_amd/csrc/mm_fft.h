// LDS-resident Stockham FFT for gfx950.
//
// A workgroup holds C = BATCH / L transforms of length L in LDS (transform c
// occupies x[c*S .. c*S + L), the stride S padded so that the C transforms start
// in different banks).  TPC = L / 16 threads work on one transform; every
// thread owns 16 points per stage.  Stages are radix-8 (radix-4/2 for the leftover factor), each one a
// read-all / barrier / twiddle + DFT_R in registers / write / barrier sweep
// (Stockham autosort: natural order in and out, no bit reversal).  Twiddles come
// from the FP64-built bank exp(-2 pi i t / L), so a length-L transform has the
// accuracy of its radix arithmetic only.
#pragma once

#include "mm_common.h"

namespace mm {

constexpr int kThreads = 1024;      // threads of a full-size FFT workgroup
constexpr int kBatchElems = 16384;  // complex points per full-size FFT workgroup (128 KiB of LDS)
constexpr int kPerThread = 16;      // points per thread per stage

// Shape of a workgroup holding BATCH / L transforms of length L.
template <int L, int BATCH = kBatchElems>
struct FftShape {
  static_assert(L >= 16 && L <= kMaxFFT && (L & (L - 1)) == 0, "power-of-two length in [16, 16384]");
  static_assert(BATCH >= L && BATCH % L == 0, "whole transforms per workgroup");
  static constexpr int C = BATCH / L;                    // transforms per workgroup
  static constexpr int TPC = L / kPerThread;             // threads per transform
  static constexpr int NT = C * TPC;                     // threads per workgroup
  // LDS stride between transforms: room for lpad(L + 1) plus a bank offset
  static constexpr int S = L + L / 16 + (L >= 256 ? 16 : 2);
  static constexpr int LDS = C * S;                      // float2 elements of LDS
};

// LDS index of element i of a transform: one pad slot per 16 complex values,
// so the stride-R writes of the first Stockham stages hit distinct banks.
__host__ __device__ __forceinline__ constexpr int lpad(int i) { return i + (i >> 4); }

// Row-pass workgroups hold half as many points (64 KiB + pad of LDS), so two of
// them share a CU and one's global loads overlap the other's LDS stages.
constexpr int kRowBatch = 8192;
template <int L>
using RowShape = FftShape<L, (L > kRowBatch ? L : kRowBatch)>;

constexpr int pick_radix(int rem) { return (rem % 8 == 0 && rem != 16) ? 8 : ((rem % 4 == 0) ? 4 : 2); }

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

template <int R, bool INV>
__host__ __device__ __forceinline__ void dft(float2* v) {
  if constexpr (R == 2) {
    dft2<INV>(v[0], v[1]);
  } else if constexpr (R == 4) {
    dft4<INV>(v[0], v[1], v[2], v[3]);
  } else {
    dft8<INV>(v);
  }
}

// One Stockham radix-R stage over a length-L transform at x (LDS), NS = product
// of the radices already applied.  Thread t in [0, TPC) handles butterflies
// j = t + b*TPC.
template <int L, int R, int NS, int TPC, bool INV>
__device__ __forceinline__ void fft_stage(float2* __restrict__ x, int t, const float2* __restrict__ tw) {
  constexpr int NB = L / R;
  constexpr int BPT = NB / TPC;
  static_assert(BPT >= 1 && NB % TPC == 0, "butterflies per thread");
  float2 v[BPT][R];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = t + b * TPC;
#pragma unroll
    for (int r = 0; r < R; ++r) v[b][r] = x[lpad(j + r * NB)];
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = t + b * TPC;
    const int k = j % NS;
    if constexpr (NS > 1) {
      // W^(k r): three table reads, the other powers by products
      constexpr int STEP = L / (NS * R);
      float2 w1 = tw[k * STEP];
      if (INV) w1.y = -w1.y;
      if constexpr (R == 2) {
        v[b][1] = cmul(v[b][1], w1);
      } else {
        float2 w2 = tw[2 * k * STEP];
        if (INV) w2.y = -w2.y;
        const float2 w3 = cmul(w1, w2);
        v[b][1] = cmul(v[b][1], w1);
        v[b][2] = cmul(v[b][2], w2);
        v[b][3] = cmul(v[b][3], w3);
        if constexpr (R == 8) {
          float2 w4 = tw[4 * k * STEP];
          if (INV) w4.y = -w4.y;
          v[b][4] = cmul(v[b][4], w4);
          v[b][5] = cmul(v[b][5], cmul(w1, w4));
          v[b][6] = cmul(v[b][6], cmul(w2, w4));
          v[b][7] = cmul(v[b][7], cmul(w3, w4));
        }
      }
    }
    dft<R, INV>(v[b]);
    const int base = (j / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) x[lpad(base + r * NS)] = v[b][r];
  }
  __syncthreads();
}

// Full in-place transform of x[lpad(0..L)) by the TPC threads with index t.
// Must be reached by every thread of the workgroup (it contains barriers).
// Caller synchronises before the first stage reads x.  Unnormalised in both
// directions (INV only conjugates the kernel).
template <int L, int TPC, bool INV, int NS = 1>
__device__ __forceinline__ void fft_lds(float2* __restrict__ x, int t, const float2* __restrict__ tw) {
  if constexpr (NS < L) {
    constexpr int R = pick_radix(L / NS);
    fft_stage<L, R, NS, TPC, INV>(x, t, tw);
    fft_lds<L, TPC, INV, NS * R>(x, t, tw);
  }
}

// ---- register-resident Stockham FFT -----------------------------------------
// Thread t (0 <= t < TPC = L / 16) of a transform holds v[m] = x[t + m TPC],
// m < 16, before and after the transform.  A radix-R pass runs the thread's
// 16 / R butterflies j = t + b TPC on its own registers (butterfly j reads
// x[j + s L / R] = v[b + s 16 / R]); its outputs go to LDS at their Stockham
// positions and come back in natural order for the next pass.  The last
// pass's Stockham positions are the natural ones, so its outputs stay in
// registers: a length 16^a r transform (r in {1, 2, 4, 8}) costs a + [r > 1] - 1
// LDS exchanges, against one LDS round trip per radix-8 stage plus the
// global <-> LDS staging of fft_lds.  Global loads and stores of v are
// therefore direct (element t + m TPC: consecutive t, consecutive addresses).

// Workgroup shape: C transforms of length L, stride S between them in LDS
// (room for XTRA more values per transform, for kernels that stage more than L
// inputs there).  S = 16 (mod 32) complex values, with the interleaved thread
// mapping of map(), keeps the Stockham writes and natural reads free of bank
// conflicts (two transforms per 32-lane half-wave at a 16-bank-pair offset).
// The plain mapping (64 consecutive t of one transform per wave) reads rows
// in longer contiguous runs; it leaves some 2-way conflicts on 64-bank LDS.
template <int L, int BATCH, int XTRA = 0, bool INTERLEAVE = true>
struct RegShape {
  static_assert(L >= 16 && L <= kMaxFFT && (L & (L - 1)) == 0, "power-of-two length in [16, 16384]");
  static constexpr int C = BATCH > L ? BATCH / L : 1;
  static constexpr int TPC = L / kPerThread;
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

// u[s] *= W_{NS R}^{s k}, s < R: table reads for the powers 1, 2, 4, 8, the
// rest as products of at most three of them.
template <int L, int R, int NS, bool INV>
__device__ __forceinline__ void twiddle(float2* u, int k, const float2* __restrict__ tw) {
  constexpr int STEP = L / (NS * R);
  auto ld = [&](int e) {
    float2 w = tw[e * k * STEP];
    if (INV) w.y = -w.y;
    return w;
  };
  const float2 w1 = ld(1);
  if constexpr (R == 2) {
    u[1] = cmul(u[1], w1);
  } else {
    const float2 w2 = ld(2), w3 = cmul(w1, w2);
    u[1] = cmul(u[1], w1);
    u[2] = cmul(u[2], w2);
    u[3] = cmul(u[3], w3);
    if constexpr (R >= 8) {
      const float2 w4 = ld(4), w5 = cmul(w1, w4), w6 = cmul(w2, w4), w7 = cmul(w3, w4);
      u[4] = cmul(u[4], w4);
      u[5] = cmul(u[5], w5);
      u[6] = cmul(u[6], w6);
      u[7] = cmul(u[7], w7);
      if constexpr (R == 16) {
        const float2 w8 = ld(8);
        u[8] = cmul(u[8], w8);
        u[9] = cmul(u[9], cmul(w1, w8));
        u[10] = cmul(u[10], cmul(w2, w8));
        u[11] = cmul(u[11], cmul(w3, w8));
        u[12] = cmul(u[12], cmul(w4, w8));
        u[13] = cmul(u[13], cmul(w5, w8));
        u[14] = cmul(u[14], cmul(w6, w8));
        u[15] = cmul(u[15], cmul(w7, w8));
      }
    }
  }
}

// One radix-R pass of fft_reg (NS = product of the radices already applied).
template <int L, int R, int NS, int TPC, bool INV>
__device__ __forceinline__ void reg_pass(float2 (&v)[kPerThread], float2* __restrict__ x, int t,
                                         const float2* __restrict__ tw) {
  constexpr int BPT = kPerThread / R;
  constexpr bool LAST = NS * R == L;
  float2 u[BPT][R];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int j = t + b * TPC;
#pragma unroll
    for (int s = 0; s < R; ++s) u[b][s] = v[b + s * BPT];
    if constexpr (NS > 1) twiddle<L, R, NS, INV>(u[b], j & (NS - 1), tw);
    if constexpr (R == 16)
      dft16<INV>(u[b]);
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
      const int base = (j / NS) * NS * R + (j & (NS - 1));
#pragma unroll
      for (int q = 0; q < R; ++q) x[lpad(base + q * NS)] = u[b][q];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kPerThread; ++m) v[m] = x[lpad(t + m * TPC)];
    __syncthreads();
  }
}

// Transform of the 16 points each of the TPC threads of one transform holds in
// v (natural order in and out), x = that transform's LDS (RegShape stride).
// Contains barriers when L > 16: every thread of the workgroup must call it.
// Unnormalised in both directions (INV only conjugates the kernel).
template <int L, int TPC, bool INV, int NS = 1>
__device__ __forceinline__ void fft_reg(float2 (&v)[kPerThread], float2* __restrict__ x, int t,
                                        const float2* __restrict__ tw) {
  if constexpr (NS < L) {
    constexpr int R = L / NS >= 16 ? 16 : L / NS;
    reg_pass<L, R, NS, TPC, INV>(v, x, t, tw);
    fft_reg<L, TPC, INV, NS * R>(v, x, t, tw);
  }
}

}  // namespace mm
