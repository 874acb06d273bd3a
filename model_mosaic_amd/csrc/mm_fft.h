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

}  // namespace mm
