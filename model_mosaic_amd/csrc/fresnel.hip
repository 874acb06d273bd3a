// Fresnel propagation D and its adjoint DT for gfx950.
//
// Operator (processing/propagation.py:47-64): pad (symmetric mirror,
// processing/kernels.py:530-550, or zero) the (nz, n) frame to (2nz, 2n),
// FFT2, multiply by exp(-i pi lambda z (fx^2 + fy^2)) / (4 n nz), unnormalised
// IFFT2, centre crop.  The kernel and the padding are separable
// (processing/propagation_large.py:7-13), so D = Dx o Dy with
//   Dy = crop_y IFFT_y K_y FFT_y pad_y   (one column pass, both FFTs in LDS)
//   Dx = crop_x IFFT_x K_x FFT_x pad_x   (one row pass)
// i.e. two HBM round trips of the complex64 frame instead of the reference's
// pad / fft2 / multiply / ifft2 / crop passes over a 4x larger buffer.
// DT = pad^T IFFT conj(K) FFT crop^T per axis (crop^T = zero embedding,
// pad^T = fold-add of the mirrored borders).
// The per-axis chirps K_x (2n) and K_y (2nz) are built on the GPU: the phase
// pi lambda z f^2 is formed and reduced mod 2 pi in FP64, then the FP32 sincos
// of the reduced phase is taken and 1/(2n) resp. 1/(2nz) is folded in.
#include <cmath>
#include <complex>
#include <type_traits>
#include <vector>

#include "mm_fft.h"

namespace mm {
namespace fresnel {

// table[j * L + k] = exp(-i (pi lambda z_j f_k^2 mod 2 pi)) / L, f_k = fftfreq(L, vox)[k]
__global__ void chirp_table(float2* __restrict__ table, const double* __restrict__ dist, int ndist, int L,
                            double wavelength, double voxelsize) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ndist * L) return;
  const int j = idx / L, k = idx - j * L;
  const double f = (double)(k < L / 2 ? k : k - L) / ((double)L * voxelsize);
  const double twopi = 6.283185307179586476925286766559;
  double ph = M_PI * wavelength * dist[j] * f * f;
  ph = fmod(ph, twopi);
  float s, c;
  sincosf((float)ph, &s, &c);
  const float inv = 1.0f / (float)L;
  table[idx] = make_float2(c * inv, -s * inv);
}

// Source row of padded point i = t + m TPC of a length-L transform over a
// frame of len = L/2 rows (len/2 padding on each side; -1: zero).  The frame
// is m in [PT/4, 3PT/4) (L/4 is a multiple of TPC), so after unrolling over m
// the frame/padding split is known at compile time.
template <class Sh>
__device__ __forceinline__ int pad_src(int m, int t, bool mirror) {
  static_assert(Sh::PT % 4 == 0, "L/4 is a multiple of TPC");
  constexpr int M0 = Sh::PT / 4, M1 = 3 * Sh::PT / 4, h = M0 * Sh::TPC, len = 2 * h;
  const int i = t + m * Sh::TPC;
  if (m < M0) return mirror ? h - i - 1 : -1;
  if (m >= M1) return mirror ? 2 * len - i + h - 1 : -1;
  return i - h;
}
// Frame row of point t + m TPC (-1: padding).
template <class Sh>
__device__ __forceinline__ int frame_row(int m, int t) {
  return (m >= Sh::PT / 4 && m < 3 * Sh::PT / 4) ? t + (m - Sh::PT / 4) * Sh::TPC : -1;
}

// Column-pass shape: long transforms take 8192 points per workgroup (two
// columns of a 4096-point transform, 70 KB of LDS), so two workgroups share a
// CU and one's loads and stores overlap the other's two transforms (4 %
// faster D than four columns per workgroup, which holds a CU alone; the
// column-fastest mapping of the DCT passes with four columns was 2 % slower
// here); short ones keep 16384 points (many columns per workgroup).
template <int L>
using ColShape = RegShape<L, (L >= 2048 ? kRowBatch : kBatchElems)>;

// Column pass: frames (b, nz, n); blockIdx.y = frame, C columns per block.
// ADJ = false: out = crop IFFT K FFT pad (in);  ADJ = true: out = pad^T IFFT conj(K) FFT crop^T (in).
// Both transforms and the chirp run on the register FFT; only the fold-add of
// the adjoint mirror padding reads other threads' points (through LDS).
template <int L, bool ADJ>
__global__ MM_TWO_FFT_BUDGET __launch_bounds__(ColShape<L>::NT) void col_pass(const float2* __restrict__ in, float2* __restrict__ out,
                                                            int n, int nz, bool mirror, const float2* __restrict__ K,
                                                            const float2* __restrict__ tw) {
  using Sh = ColShape<L>;
  __shared__ float2 lds[Sh::LDS];
  int c, t;
  Sh::map(threadIdx.x, c, t);
  // the column groups that share each 128-byte line of a row run on one XCD
  const unsigned lw = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int bx = (int)(lw % gridDim.x), frame = (int)(lw / gridDim.x);
  const int col = bx * Sh::C + c;
  const bool live = col < n;
  const long long fo = (long long)frame * nz * n + col;
  const float2* src = in + fo;
  float2* dst = out + fo;
  float2* x = lds + c * Sh::S;
  const int h = nz / 2;
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int sr = ADJ ? frame_row<Sh>(m, t) : pad_src<Sh>(m, t, mirror);
    v[m] = (live && sr >= 0) ? src[(long long)sr * n] : make_float2(0.f, 0.f);
  }
  fft_reg<L, Sh::TPC, false, 1, false>(v, x, t, tw);
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    float2 kk = K[t + m * Sh::TPC];
    if (ADJ) kk.y = -kk.y;
    v[m] = cmul(v[m], kk);
  }
  fft_reg<L, Sh::TPC, true, 1, false>(v, x, t, tw);
  if (ADJ && mirror) {
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) x[lpad(t + m * Sh::TPC)] = v[m];
    __syncthreads();
    if (!live) return;  // no barrier below
    for (int r = t; r < nz; r += Sh::TPC) {
      const int pr = (r < h) ? (h - 1 - r) : (2 * nz + h - 1 - r);
      dst[(long long)r * n] = cadd(x[lpad(r + h)], x[lpad(pr)]);
    }
  } else {
    if (!live) return;
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) {
      const int r = frame_row<Sh>(m, t);
      if (r >= 0) dst[(long long)r * n] = v[m];
    }
  }
}

// Row pass: rows of length n, C rows per block, nrows = b * nz.
template <int L, bool ADJ>
__global__ MM_TWO_FFT_BUDGET __launch_bounds__(RowReg<L>::NT) void row_pass(const float2* __restrict__ in, float2* __restrict__ out,
                                                          int n, long long nrows, bool mirror,
                                                          const float2* __restrict__ K, const float2* __restrict__ tw) {
  using Sh = RowReg<L>;
  __shared__ float2 lds[Sh::LDS];
  int c, t;
  Sh::map(threadIdx.x, c, t);
  const long long row = (long long)blockIdx.x * Sh::C + c;
  const bool valid = row < nrows;
  const float2* src = in + (valid ? row : 0) * n;
  float2* x = lds + c * Sh::S;
  const int h = n / 2;
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int sx = ADJ ? frame_row<Sh>(m, t) : pad_src<Sh>(m, t, mirror);
    v[m] = (valid && sx >= 0) ? src[sx] : make_float2(0.f, 0.f);
  }
  fft_reg<L, Sh::TPC, false, 1, false>(v, x, t, tw);
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    float2 kk = K[t + m * Sh::TPC];
    if (ADJ) kk.y = -kk.y;
    v[m] = cmul(v[m], kk);
  }
  fft_reg<L, Sh::TPC, true, 1, false>(v, x, t, tw);
  float2* dst = out + (valid ? row : 0) * n;
  if (ADJ && mirror) {
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) x[lpad(t + m * Sh::TPC)] = v[m];
    __syncthreads();
    if (!valid) return;  // no barrier below
    for (int r = t; r < n; r += Sh::TPC) {
      const int pr = (r < h) ? (h - 1 - r) : (2 * n + h - 1 - r);
      dst[r] = cadd(x[lpad(r + h)], x[lpad(pr)]);
    }
  } else {
    if (!valid) return;
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) {
      const int r = frame_row<Sh>(m, t);
      if (r >= 0) dst[r] = v[m];
    }
  }
}

// ---- mirror padding as a DCT: transforms of the frame length ---------------
// Workgroup sizes of the DCT-form passes (complex points per workgroup; build-time shape knobs).
#ifndef MM_FRESNEL_DCT_ROW_BATCH
#define MM_FRESNEL_DCT_ROW_BATCH 4096
#endif
#ifndef MM_FRESNEL_DCT_COL_BATCH
#define MM_FRESNEL_DCT_COL_BATCH kBatchElems
#endif
// Column transforms of fewer than 16 threads, and of 2048 points and more, take the column-fastest lane
// mapping (ColFastShape): 8 columns per workgroup for 2048-point columns, 6.13 -> 5.62 ms per D of 128
// frames of 2048^2 (4 per workgroup in the interleaved mapping was the best before).  The short columns
// of the mosaic's 96-row stacks and the model's 128-row frames take 64 and 32 columns per workgroup
// (several workgroups per CU; a wave still reads one 256- or 512-byte row run): mosaic batch 0.584 ->
// 0.572 s, model 13.87 -> 13.72 ms, same bits.
constexpr int dct_col_batch(int L) {
  return L >= 2048 ? MM_FRESNEL_DCT_COL_BATCH : (L == 96 ? 6144 : (L == 128 ? 4096 : kBatchElems));
}
template <int L, int BATCH = dct_col_batch(L)>
using DctColShape = ColumnShape<L, BATCH>;
template <int L>
using DctRowShape = RegShape<L, MM_FRESNEL_DCT_ROW_BATCH, 0, false>;

// The symmetric pad of a line x (length N) to 2N is the half-sample-symmetric
// extension of x shifted by N/2, and the chirp K is even in the frequency
// index, so  crop IFFT_2N K FFT_2N pad  =  DCT-III . K . DCT-II  on N points:
//   C_k = sum_j x_j 2 cos(pi k (2j+1) / 2N),   z_j = sum_k w_k K_k C_k cos(pi k (2j+1) / 2N)
// (w_0 = 1, else 2; K carries 1/(2N)).  The operator is complex-symmetric, so
// its adjoint is the same pipeline with conj(K): no fold-add.  Both DCTs run on
// one complex FFT of length N (Makhoul's reordering v_j = x_{2j}, v_{N-1-j} =
// x_{2j+1}):  with V = FFT(v) and c_k = exp(-i pi k / 2N),
//   C_k = c_k V_k + conj(c_k) V_{N-k},   C_{N-k} = i (c_k V_k - conj(c_k) V_{N-k}),
//   U_k = conj(c_k) (K_k C_k - i K_{N-k} C_{N-k}),   x_{perm(j)} = (IFFT U)_j.
// Multiplied out (|c_k| = 1):  U_k = A_k V_k + B_k V_{N-k}  with
//   A_k = K_k + K_{N-k},   B_k = conj(c_k)^2 (K_k - K_{N-k})
// (k = 0: V_{N-0} = V_0, K_{N-0} = K_N, U_0 = 2 K_0 V_0); with conj(K) for DT:
// A -> conj(A), B -> conj(c_k)^2 conj(K_k - K_{N-k}).  The plan holds A, B
// and the DT B per distance, built in FP64: two complex products per point
// between the transforms.  The partner V_{N-k} of thread t's points sits in
// thread (TPC - t) mod TPC: one LDS exchange between the two transforms.  The
// j < N/2 half of the reordering is the registers m < PT/2, so it is fixed at
// compile time.
template <class Sh>
__device__ __forceinline__ int dct_src(int m, int t) {
  constexpr int N = Sh::PT * Sh::TPC;
  const int j = t + m * Sh::TPC;
  return m < Sh::PT / 2 ? 2 * j : 2 * N - 1 - 2 * j;
}

// Between the forward and the inverse transform: v holds V_k (k = t + m TPC) on entry and U_k on exit.
// M = [A | B | B of DT] of one distance (3 N values).
template <class Sh, bool CONJ_K>
__device__ __forceinline__ void dct_mid(float2 (&v)[Sh::PT], float2* __restrict__ x, int t,
                                        const float2* __restrict__ M) {
  constexpr int N = Sh::PT * Sh::TPC;
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) x[lpad(t + m * Sh::TPC)] = v[m];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int k = t + m * Sh::TPC;
    const float2 vr = x[lpad(k == 0 ? 0 : N - k)];
    float2 a = M[k];
    if (CONJ_K) a.y = -a.y;
    const float2 b = M[(CONJ_K ? 2 * N : N) + k];
    const float2 p = cmul(a, v[m]), q = cmul(b, vr);
    v[m] = make_float2(p.x + q.x, p.y + q.y);
  }
  __syncthreads();  // the inverse transform's first pass rewrites x
}

// Load / store policies of the DCT-form passes.  An index is the flat element
// index of the (frames, nz, n) array.  The plain operator moves complex64; the
// forward model builds psi = exp(-k b) e^{i k p} from the two float32
// sinograms while loading, and its backward forms 4 (|u|^2 - d) u while
// loading and the chain rule while storing, so none of these element-wise
// steps is a pass of its own.  Each policy splits into the raw global reads
// (fetch / pre) and the arithmetic on them (make / the store): the passes
// issue every point's reads first and compute afterwards, so the reads stay in
// flight together (reads interleaved with the libm calls of psi were issued
// one or two at a time).
struct NoAux {};
struct LoadC64 {
  const float2* __restrict__ in;
  static constexpr bool kSplit = false;  // plain reads: kept predicated (the split form costs registers)
  using Raw = float2;
  __device__ __forceinline__ Raw fetch(long long i) const { return in[i]; }
  __device__ __forceinline__ float2 make(Raw r) const { return r; }
};
struct LoadPsi {  // psi of step3_propagation.py:94-103 with separate phase / attenuation sinograms
  const float* __restrict__ p;
  const float* __restrict__ b;
  float k;
  static constexpr bool kSplit = true;
  using Raw = float2;  // (p, b)
  __device__ __forceinline__ Raw fetch(long long i) const { return make_float2(p[i], b[i]); }
  __device__ __forceinline__ float2 make(Raw r) const {
    float sn, cs;
    sincosf(k * r.x, &sn, &cs);
    const float at = expf(-k * r.y);
    return make_float2(at * cs, at * sn);
  }
};
struct LoadPsiRatio {  // psi of one projection sinogram (step3_propagation.py:94-103)
  const float* __restrict__ p;
  float inv_norm, inv_beta_ratio;
  static constexpr bool kSplit = true;
  using Raw = float;
  __device__ __forceinline__ Raw fetch(long long i) const { return p[i]; }
  __device__ __forceinline__ float2 make(Raw r) const { return psi_of(r, inv_norm, inv_beta_ratio); }
};
struct LoadResidual {  // 4 (|u|^2 - d) u
  const float2* __restrict__ u;
  const float* __restrict__ d;
  static constexpr bool kSplit = true;
  using Raw = float3;  // (u, d)
  __device__ __forceinline__ Raw fetch(long long i) const {
    const float2 v = u[i];
    return make_float3(v.x, v.y, d[i]);
  }
  __device__ __forceinline__ float2 make(Raw r) const {
    const float q = 4.0f * (r.x * r.x + r.y * r.y - r.z);
    return make_float2(q * r.x, q * r.y);
  }
};
struct StoreC64 {
  float2* __restrict__ out;
  static constexpr bool kLoss = false;
  using Aux = NoAux;
  __device__ __forceinline__ Aux pre(long long) const { return {}; }
  __device__ __forceinline__ void operator()(long long i, float2 v, double&, Aux) const { out[i] = v; }
};
struct StoreLoss {  // u kept for the backward; (|u|^2 - d)^2 summed in FP64
  float2* __restrict__ out;
  const float* __restrict__ d;
  static constexpr bool kLoss = true;
  using Aux = float;  // d
  __device__ __forceinline__ Aux pre(long long i) const { return d[i]; }
  __device__ __forceinline__ void operator()(long long i, float2 v, double& acc, Aux dd) const {
    out[i] = v;
    const double r = (double)(v.x * v.x + v.y * v.y) - (double)dd;
    acc += r * r;
  }
};
struct StoreAbs2 {  // intensity |u|^2 (step3_propagation.py:105-110)
  float* __restrict__ out;
  static constexpr bool kLoss = false;
  using Aux = NoAux;
  __device__ __forceinline__ Aux pre(long long) const { return {}; }
  __device__ __forceinline__ void operator()(long long i, float2 v, double&, Aux) const { out[i] = abs2_of(v); }
};
struct StoreChain {  // dL/dp = -k Im(conj(G) psi), dL/db = -k Re(conj(G) psi), times gscale, over p and b
  float* __restrict__ p;
  float* __restrict__ b;
  float k, gscale;
  static constexpr bool kLoss = false;
  using Aux = float2;  // (p, b) before they are overwritten
  __device__ __forceinline__ Aux pre(long long i) const { return make_float2(p[i], b[i]); }
  __device__ __forceinline__ void operator()(long long i, float2 g, double&, Aux pb) const {
    float sn, cs;
    sincosf(k * pb.x, &sn, &cs);
    const float at = expf(-k * pb.y);
    const float2 psi = make_float2(at * cs, at * sn);
    const float re = g.x * psi.x + g.y * psi.y;  // Re(conj(g) psi)
    const float im = g.x * psi.y - g.y * psi.x;  // Im(conj(g) psi)
    p[i] = -k * im * gscale;
    b[i] = -k * re * gscale;
  }
};

// Sum of one double per thread into partial[blockIdx.x] (x: the workgroup's LDS, free after the transforms).
template <int NT>
__device__ __forceinline__ void block_sum(double acc, float2* __restrict__ x, double* __restrict__ partial,
                                          unsigned slot) {
  double* red = reinterpret_cast<double*>(x);
  __syncthreads();
  red[threadIdx.x] = acc;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[slot] = red[0];
}

// Column pass on the DCT form: frames (b, nz, n), transforms of length L = nz down the columns.
template <int L, bool CONJ_K, class Load, class Store>
__global__ MM_TWO_FFT_BUDGET __launch_bounds__(DctColShape<L>::NT) void dct_col_pass(
    Load ld, Store st, int n, const float2* __restrict__ M, const float2* __restrict__ tw,
    double* __restrict__ partial) {
  using Sh = DctColShape<L>;
  __shared__ float2 lds[Sh::LDS];
  int c, t;
  Sh::map(threadIdx.x, c, t);
  const unsigned lw = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int bx = (int)(lw % gridDim.x), frame = (int)(lw / gridDim.x);
  const int col = bx * Sh::C + c;
  const bool live = col < n;
  // split reads of a column past the frame width (last block) come from column n - 1 and are dropped
  const long long fo = (long long)frame * L * n + (Load::kSplit && !live ? n - 1 : col);
  float2* x = lds + c * Sh::S;
  float2 v[Sh::PT];
  if constexpr (Load::kSplit) {
    typename Load::Raw r[Sh::PT];
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) r[m] = ld.fetch(fo + (long long)dct_src<Sh>(m, t) * n);
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) v[m] = live ? ld.make(r[m]) : make_float2(0.f, 0.f);
  } else {
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m)
      v[m] = live ? ld.make(ld.fetch(fo + (long long)dct_src<Sh>(m, t) * n)) : make_float2(0.f, 0.f);
  }
  fft_reg<L, Sh::TPC, false, 1, false>(v, x, t, tw);
  dct_mid<Sh, CONJ_K>(v, x, t, M);
  fft_reg<L, Sh::TPC, true, 1, false>(v, x, t, tw);
  double acc = 0.0;
  if (live) {
    typename Store::Aux a[Sh::PT];
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) a[m] = st.pre(fo + (long long)dct_src<Sh>(m, t) * n);
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) st(fo + (long long)dct_src<Sh>(m, t) * n, v[m], acc, a[m]);
  }
  if constexpr (Store::kLoss) block_sum<Sh::NT>(acc, lds, partial, blockIdx.x + gridDim.x * blockIdx.y);
}

// Row pass on the DCT form: rows of length L = n, C rows per block, nrows = b * nz.
template <int L, bool CONJ_K, class Load, class Store>
__global__ MM_TWO_FFT_BUDGET __launch_bounds__(DctRowShape<L>::NT) void dct_row_pass(
    Load ld, Store st, long long nrows, const float2* __restrict__ M, const float2* __restrict__ tw,
    double* __restrict__ partial) {
  using Sh = DctRowShape<L>;
  __shared__ float2 lds[Sh::LDS];
  int c, t;
  Sh::map(threadIdx.x, c, t);
  const long long row = (long long)blockIdx.x * Sh::C + c;
  const bool valid = row < nrows;
  const long long ro = (valid ? row : 0) * L;
  float2* x = lds + c * Sh::S;
  float2 v[Sh::PT];
  if constexpr (Load::kSplit) {
    typename Load::Raw r[Sh::PT];
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) r[m] = ld.fetch(ro + dct_src<Sh>(m, t));
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) v[m] = valid ? ld.make(r[m]) : make_float2(0.f, 0.f);
  } else {
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) v[m] = valid ? ld.make(ld.fetch(ro + dct_src<Sh>(m, t))) : make_float2(0.f, 0.f);
  }
  fft_reg<L, Sh::TPC, false, 1, false>(v, x, t, tw);
  dct_mid<Sh, CONJ_K>(v, x, t, M);
  fft_reg<L, Sh::TPC, true, 1, false>(v, x, t, tw);
  double acc = 0.0;
  if (valid) {
    typename Store::Aux a[Sh::PT];
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) a[m] = st.pre(ro + dct_src<Sh>(m, t));
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) st(ro + dct_src<Sh>(m, t), v[m], acc, a[m]);
  }
  if constexpr (Store::kLoss) block_sum<Sh::NT>(acc, lds, partial, blockIdx.x);
}

// Rows longer than one workgroup's transform (n = 2Q, Q <= 16384) on the DCT
// form: the n-point FFT of the reordered row v (v_s = x_(2s), v_(s+Q) =
// x_(n-1-2s), s < Q) splits by frequency parity, V[2m + e] =
// FFT_Q(w^(2es) (v_s + (-1)^e v_(s+Q)))[m] with w = exp(-i pi / n) (table h);
// the partner V[n - k] of V[k] has k's parity, so each half takes its own
// step U = A V + B V(n - .) through LDS; the inverse n-point transform splits by
// input parity, E = IFFT_Q(U[2m]), O = IFFT_Q(U[2m + 1]), y_p = E_p + conj(w^(2p)) O_p,
// y_(p+Q) = E_p - conj(w^(2p)) O_p, and y_j goes back to x position dct_src(j).
// One row per workgroup, in place: both halves' inputs are in registers before
// the first barrier.
template <int Q>
using RowSplitShape = RegShape<Q, Q>;  // one Q-point transform per workgroup
template <int Q, bool CONJ_K>
__global__ __launch_bounds__(RowSplitShape<Q>::NT) void dct_row_split(
    float2* __restrict__ buf, long long nrows, const float2* __restrict__ M, const float2* __restrict__ h,
    const float2* __restrict__ tw) {
  using Sh = RowSplitShape<Q>;
  static_assert(Sh::C == 1 && Sh::NT == Sh::TPC, "one transform per workgroup");
  __shared__ float2 lds[Sh::LDS];
  constexpr int N = 2 * Q;
  const long long row = blockIdx.x;
  if (row >= nrows) return;  // the whole workgroup, before any barrier
  float2* x = buf + row * N;
  const int t = threadIdx.x;
  // Thread t reads and writes only the row positions 2s and N - 1 - 2s of its own s = t + m TPC, so it
  // parks the odd half's input and then the even half's result there: one half is live in registers.
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int s = t + m * Sh::TPC;
    const float2 lo = x[2 * s], hi = x[N - 1 - 2 * s];
    v[m] = make_float2(lo.x + hi.x, lo.y + hi.y);
    x[N - 1 - 2 * s] = cmul(make_float2(lo.x - hi.x, lo.y - hi.y), h[2 * s]);
  }
  // one half: forward transform, the step between the transforms, inverse transform (in place in v)
  auto half = [&](int e, int tt, const float2* twp) {
    fft_reg<Q, Sh::TPC, false, 1, false>(v, lds, tt, twp);
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) lds[lpad(tt + m * Sh::TPC)] = v[m];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) {
      const int mm = tt + m * Sh::TPC, k = 2 * mm + e;
      const int mr = e ? Q - 1 - mm : (mm == 0 ? 0 : Q - mm);  // V[N - k] (k = 0: V[0])
      const float2 vr = lds[lpad(mr)];
      float2 a = M[k];
      if (CONJ_K) a.y = -a.y;
      const float2 b = M[(CONJ_K ? 2 * N : N) + k];
      const float2 p = cmul(a, v[m]), q = cmul(b, vr);
      v[m] = make_float2(p.x + q.x, p.y + q.y);
    }
    __syncthreads();  // the inverse transform's first pass rewrites lds
    fft_reg<Q, Sh::TPC, true, 1, false>(v, lds, tt, twp);
  };
  half(0, t, tw);
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {  // E parked at 2p, the odd input taken back
    const int s = t + m * Sh::TPC;
    const float2 odd = x[N - 1 - 2 * s];
    x[2 * s] = v[m];
    v[m] = odd;
  }
  half(1, opaque_index(t), opaque_twiddles(tw));
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int p = t + m * Sh::TPC;
    const float2 e = x[2 * p], o = cmul(v[m], cconj(h[2 * p]));
    x[2 * p] = make_float2(e.x + o.x, e.y + o.y);          // y_p at dct_src(p) = 2p
    x[N - 1 - 2 * p] = make_float2(e.x - o.x, e.y - o.y);  // y_(p+Q) at dct_src(p+Q)
  }
}

}  // namespace fresnel
}  // namespace mm

using namespace mm;
using namespace mm::fresnel;

struct mm_fresnel {
  int dev = 0;
  int n = 0, nz = 0, ndist = 0, pad_mode = 0;
  bool dct = false;      // mirror padding on the DCT form (both frame sides are DCT lengths)
  float2* Kx = nullptr;  // ndist x 2n
  float2* Ky = nullptr;  // ndist x 2nz
  float2* Mx = nullptr;  // DCT form: ndist x [A | B | B of DT] x n (see dct_mid)
  float2* My = nullptr;  // ndist x 3 x nz
  float2* hx = nullptr;  // rows above 16384 (n = 2Q): exp(-i pi j / n), j < n (dct_row_split)
  const float2* tw = nullptr;
  size_t bytes = 0;
};

static void fresnel_free(mm_fresnel* p) {
  if (!p) return;
  if (p->Kx) (void)hipFree(p->Kx);
  if (p->Ky) (void)hipFree(p->Ky);
  if (p->Mx) (void)hipFree(p->Mx);
  if (p->My) (void)hipFree(p->My);
  if (p->hx) (void)hipFree(p->hx);
  delete p;
}

template <int L, bool ADJ>
static void launch_col_pass(const mm_fresnel* p, const float2* in, float2* out, int nframes, int j, hipStream_t st) {
  using Sh = ColShape<L>;
  hipLaunchKernelGGL((col_pass<L, ADJ>), dim3((unsigned)((p->n + Sh::C - 1) / Sh::C), (unsigned)nframes),
                     dim3(Sh::NT), 0, st, in, out, p->n, p->nz, p->pad_mode == 0, p->Ky + (size_t)j * L, p->tw + tw_offset(L));
}

template <int L, bool ADJ>
static void launch_row_pass(const mm_fresnel* p, const float2* in, float2* out, int nframes, int j, hipStream_t st) {
  using Sh = RowReg<L>;
  const long long rows = (long long)nframes * p->nz;
  hipLaunchKernelGGL((row_pass<L, ADJ>), dim3((unsigned)((rows + Sh::C - 1) / Sh::C)), dim3(Sh::NT), 0, st, in, out,
                     p->n, rows, p->pad_mode == 0, p->Kx + (size_t)j * L, p->tw + tw_offset(L));
}


// The DCT form needs the length-N register FFT (N >= 16, 3 * 2^k >= 48).
static bool dct_len_ok(int N) { return ((is_pow2(N) && N >= 16) || (is_fft_len(N) && N >= 48)) && N <= kMaxFFT; }
// Rows may also be twice that (dct_row_split: two N/2-point halves).
static bool dct_row_len_ok(int N) { return dct_len_ok(N) || (N > kMaxFFT && N <= 2 * kMaxFFT && dct_len_ok(N / 2)); }

template <int Q, bool CONJ_K>
static void launch_dct_row_split(const mm_fresnel* p, float2* buf, int nframes, int j, hipStream_t st) {
  const long long rows = (long long)nframes * p->nz;
  hipLaunchKernelGGL((dct_row_split<Q, CONJ_K>), dim3((unsigned)rows), dim3(RowSplitShape<Q>::NT), 0, st, buf, rows,
                     p->Mx + (size_t)j * 3 * 2 * Q, p->hx, p->tw + tw_offset(Q));
}

// [A | B | B of DT] (dct_mid) of a DCT side of N points for every distance, in FP64 from the chirp of the
// 2N-point padded transform (the same phase as chirp_table, K carrying 1/(2N)).
static std::vector<float2> dct_mid_tables(int N, const double* dist, int ndist, double wavelength, double voxelsize) {
  const int L = 2 * N;
  std::vector<float2> out((size_t)ndist * 3 * N);
  std::vector<std::complex<double>> K(N + 1);
  for (int j = 0; j < ndist; ++j) {
    for (int k = 0; k <= N; ++k) {
      const double f = (double)(k < L / 2 ? k : k - L) / ((double)L * voxelsize);
      const double ph = std::fmod(M_PI * wavelength * dist[j] * f * f, 2.0 * M_PI);
      K[k] = std::polar(1.0 / L, -ph);
    }
    float2* o = out.data() + (size_t)j * 3 * N;
    for (int k = 0; k < N; ++k) {
      const std::complex<double> kk = K[k], kr = K[N - k];  // k = 0: K_N
      const std::complex<double> c2 = std::polar(1.0, M_PI * k / N);  // conj(c_k)^2
      const std::complex<double> a = kk + kr, b = c2 * (kk - kr), bt = c2 * std::conj(kk - kr);
      o[k] = make_float2((float)a.real(), (float)a.imag());
      o[N + k] = make_float2((float)b.real(), (float)b.imag());
      o[2 * N + k] = make_float2((float)bt.real(), (float)bt.imag());
    }
  }
  return out;
}

template <int L, bool CONJ_K, class Load, class Store>
static void launch_dct_col(const mm_fresnel* p, Load ld, Store sto, int nframes, int j, double* partial,
                           hipStream_t st) {
  using Sh = DctColShape<L>;
  hipLaunchKernelGGL((dct_col_pass<L, CONJ_K, Load, Store>), dim3((unsigned)((p->n + Sh::C - 1) / Sh::C), (unsigned)nframes),
                     dim3(Sh::NT), 0, st, ld, sto, p->n, p->My + (size_t)j * 3 * L, p->tw + tw_offset(L), partial);
}

template <int L, bool CONJ_K, class Load, class Store>
static void launch_dct_row(const mm_fresnel* p, Load ld, Store sto, int nframes, int j, double* partial,
                           hipStream_t st) {
  using Sh = DctRowShape<L>;
  const long long rows = (long long)nframes * p->nz;
  hipLaunchKernelGGL((dct_row_pass<L, CONJ_K, Load, Store>), dim3((unsigned)((rows + Sh::C - 1) / Sh::C)), dim3(Sh::NT),
                     0, st, ld, sto, rows, p->Mx + (size_t)j * 3 * L, p->tw + tw_offset(L), partial);
}

template <int L>
static long long dct_row_parts(const mm_fresnel* p, int nframes) {
  return ((long long)nframes * p->nz + DctRowShape<L>::C - 1) / DctRowShape<L>::C;
}

static int fresnel_apply(mm_fresnel* p, const void* in, void* out, int nframes, int j, void* stream, bool adj) {
  MM_ARG_CHECK(p != nullptr, "mm_fresnel: plan is NULL");
  MM_ARG_CHECK(nframes >= 0, "mm_fresnel: nframes < 0");
  MM_ARG_CHECK(j >= 0 && j < p->ndist,
               "mm_fresnel: distance index " + std::to_string(j) + " outside [0, " + std::to_string(p->ndist) + ")");
  if (nframes == 0) return MM_OK;
  MM_ARG_CHECK(in && out, "mm_fresnel: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const float2* src = (const float2*)in;
  float2* dst = (float2*)out;
  if (p->pad_mode == 0 && p->dct) {  // mirror padding: DCT form, DT = the same with conj(K)
    const LoadC64 ld_in{src}, ld_out{dst};
    const StoreC64 sto{dst};
    if (adj) {
      MM_DISPATCH_FFT(p->nz, (launch_dct_col<LL, true>(p, ld_in, sto, nframes, j, nullptr, st)));
    } else {
      MM_DISPATCH_FFT(p->nz, (launch_dct_col<LL, false>(p, ld_in, sto, nframes, j, nullptr, st)));
    }
    int rc = check_launch("fresnel dct_col_pass");
    if (rc) return rc;
    if (p->n > kMaxFFT) {  // n = 24576 or 32768 (dct_row_len_ok)
      if (p->n == 24576) {
        if (adj) launch_dct_row_split<12288, true>(p, dst, nframes, j, st);
        else launch_dct_row_split<12288, false>(p, dst, nframes, j, st);
      } else if (p->n == 32768) {
        if (adj) launch_dct_row_split<16384, true>(p, dst, nframes, j, st);
        else launch_dct_row_split<16384, false>(p, dst, nframes, j, st);
      } else {
        return fail(MM_ERR_UNSUPPORTED, "fresnel: split rows not built for n=" + std::to_string(p->n));
      }
      return check_launch("fresnel dct_row_split");
    }
    if (adj) {
      MM_DISPATCH_FFT(p->n, (launch_dct_row<LL, true>(p, ld_out, sto, nframes, j, nullptr, st)));
    } else {
      MM_DISPATCH_FFT(p->n, (launch_dct_row<LL, false>(p, ld_out, sto, nframes, j, nullptr, st)));
    }
    return check_launch("fresnel dct_row_pass");
  }
  const int Ly = 2 * p->nz, Lx = 2 * p->n;
  if (adj) {
    MM_DISPATCH_FFT(Ly, (launch_col_pass<LL, true>(p, src, dst, nframes, j, st)));
  } else {
    MM_DISPATCH_FFT(Ly, (launch_col_pass<LL, false>(p, src, dst, nframes, j, st)));
  }
  int rc = check_launch("fresnel col_pass");
  if (rc) return rc;
  if (adj) {
    MM_DISPATCH_FFT(Lx, (launch_row_pass<LL, true>(p, dst, dst, nframes, j, st)));
  } else {
    MM_DISPATCH_FFT(Lx, (launch_row_pass<LL, false>(p, dst, dst, nframes, j, st)));
  }
  return check_launch("fresnel row_pass");
}

namespace mm {
// Forward of the model on the DCT form: psi from (P, B) -> u = D_j psi (kept), loss partials of
// (|u|^2 - d)^2 into partial[0 .. *nparts).  Column pass first (psi built while loading), row pass
// second (coalesced loads of d next to the stores of u).
int fresnel_model_fwd(const mm_fresnel* p, const float* P, const float* B, float k, float2* u, const float* d,
                      double* partial, int* nparts, int nframes, int j, hipStream_t st) {
  if (!(p->pad_mode == 0 && p->dct)) return fail(MM_ERR_UNSUPPORTED, "fused model passes need the DCT form");
  const LoadPsi lpsi{P, B, k};
  const LoadC64 lu{u};
  const StoreC64 su{u};
  const StoreLoss sl{u, d};
  MM_DISPATCH_FFT(p->nz, (launch_dct_col<LL, false>(p, lpsi, su, nframes, j, nullptr, st)));
  int rc = check_launch("model psi + D columns");
  if (rc) return rc;
  MM_DISPATCH_FFT(p->n, (launch_dct_row<LL, false>(p, lu, sl, nframes, j, partial, st)));
  MM_DISPATCH_FFT(p->n, (*nparts = (int)dct_row_parts<LL>(p, nframes)));
  return check_launch("model D rows + loss");
}

// Backward of the model on the DCT form: G = DT_j(4 (|u|^2 - d) u) over u, then P, B <- gscale times
// dL/dp, dL/db (psi recomputed from P, B while storing).
int fresnel_model_bwd(const mm_fresnel* p, float2* u, const float* d, float* P, float* B, float k, float gscale,
                      int nframes, int j, hipStream_t st) {
  if (!(p->pad_mode == 0 && p->dct)) return fail(MM_ERR_UNSUPPORTED, "fused model passes need the DCT form");
  const LoadResidual lr{u, d};
  const LoadC64 lu{u};
  const StoreC64 su{u};
  const StoreChain sc{P, B, k, gscale};
  MM_DISPATCH_FFT(p->nz, (launch_dct_col<LL, true>(p, lr, su, nframes, j, nullptr, st)));
  int rc = check_launch("model residual + DT columns");
  if (rc) return rc;
  MM_DISPATCH_FFT(p->n, (launch_dct_row<LL, true>(p, lu, sc, nframes, j, nullptr, st)));
  return check_launch("model DT rows + chain");
}

bool fresnel_fused_ok(const mm_fresnel* p) { return p->pad_mode == 0 && p->dct; }

// out = |D_j psi(p)|^2 with psi built while the column pass loads p and |.|^2 taken while the row pass stores;
// work (nframes, nz, n) complex64 holds the wavefield between the passes.
int fresnel_intensity(const mm_fresnel* p, const float* P, float inv_norm, float inv_beta_ratio, float2* work,
                      float* out, int nframes, int j, hipStream_t st) {
  if (!(p->pad_mode == 0 && p->dct)) return fail(MM_ERR_UNSUPPORTED, "fused intensity passes need the DCT form");
  const LoadPsiRatio lpsi{P, inv_norm, inv_beta_ratio};
  const LoadC64 lw{work};
  const StoreC64 sw{work};
  const StoreAbs2 sa{out};
  MM_DISPATCH_FFT(p->nz, (launch_dct_col<LL, false>(p, lpsi, sw, nframes, j, nullptr, st)));
  int rc = check_launch("psi + D columns");
  if (rc) return rc;
  MM_DISPATCH_FFT(p->n, (launch_dct_row<LL, false>(p, lw, sa, nframes, j, nullptr, st)));
  return check_launch("D rows + |.|^2");
}

// Loss partials the fused forward writes for nframes frames (one per row-pass workgroup).
int fresnel_model_parts(const mm_fresnel* p, int nframes, long long* parts) {
  MM_DISPATCH_FFT(p->n, (*parts = dct_row_parts<LL>(p, nframes)));
  return MM_OK;
}

void fresnel_dims(const mm_fresnel* p, int* n, int* nz, int* ndist) {
  *n = p->n;
  *nz = p->nz;
  *ndist = p->ndist;
}
}  // namespace mm

extern "C" {

int mm_fresnel_create(mm_fresnel** plan, int n, int nz, int ndist, double wavelength, double voxelsize,
                      const double* dist_host, int pad_mode) {
  MM_ARG_CHECK(plan != nullptr, "mm_fresnel_create: plan is NULL");
  *plan = nullptr;
  auto size_ok = [](int v) { return ((is_pow2(v) && v >= 8) || (is_fft_len(v) && v >= 24)) && v <= kMaxFFT; };
  MM_ARG_CHECK(size_ok(n) || (pad_mode == 0 && n > kMaxFFT && dct_row_len_ok(n)),
               "mm_fresnel_create: n must be 2^k in [8, 32768] or 3 * 2^k in [24, 24576] (above 16384 with symmetric "
               "padding), got " + std::to_string(n));
  MM_ARG_CHECK(size_ok(nz), "mm_fresnel_create: nz must be 2^k in [8, 16384] or 3 * 2^k in [24, 12288], got " +
                                std::to_string(nz));
  MM_ARG_CHECK(ndist >= 1 && dist_host != nullptr, "mm_fresnel_create: need ndist >= 1 distances");
  MM_ARG_CHECK(pad_mode == 0 || pad_mode == 1, "mm_fresnel_create: pad_mode must be 0 (symmetric) or 1 (zero)");
  // a side past 8192 runs only on the frame-length DCT form (the padded form's transforms are 2N long)
  MM_ARG_CHECK((2 * n <= kMaxFFT && 2 * nz <= kMaxFFT) || (pad_mode == 0 && dct_row_len_ok(n) && dct_len_ok(nz)),
               "mm_fresnel_create: sides above 8192 need symmetric padding and both sides 2^k >= 16 or 3 * 2^k >= 48");
  MM_ARG_CHECK(voxelsize > 0 && wavelength >= 0, "mm_fresnel_create: bad wavelength / voxelsize");
  mm_fresnel* p = new mm_fresnel();
  auto bail = [&](int code) {
    fresnel_free(p);
    return code;
  };
  int rc = twiddle_bank(&p->tw);
  if (rc) return bail(rc);
  if (hipGetDevice(&p->dev) != hipSuccess) return bail(fail(MM_ERR_HIP, "hipGetDevice"));
  p->n = n;
  p->nz = nz;
  p->ndist = ndist;
  p->pad_mode = pad_mode;
  p->dct = pad_mode == 0 && dct_row_len_ok(n) && dct_len_ok(nz);
  double* ddist = nullptr;
  if (hipMalloc(&ddist, sizeof(double) * ndist) != hipSuccess) return bail(fail(MM_ERR_OOM, "fresnel distances"));
  if (n > kMaxFFT) {  // the split rows' half-angle table, FP64-built
    std::vector<float2> hh(n);
    for (int j = 0; j < n; ++j) {
      const double a = -M_PI * (double)j / n;
      hh[j] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    if (hipMalloc((void**)&p->hx, sizeof(float2) * n) != hipSuccess ||
        hipMemcpy(p->hx, hh.data(), sizeof(float2) * n, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(ddist);
      return bail(fail(MM_ERR_OOM, "fresnel split-row table"));
    }
    p->bytes += sizeof(float2) * n;
  }
  if (hipMalloc(&p->Kx, sizeof(float2) * (size_t)ndist * 2 * n) != hipSuccess ||
      hipMalloc(&p->Ky, sizeof(float2) * (size_t)ndist * 2 * nz) != hipSuccess) {
    (void)hipFree(ddist);
    return bail(fail(MM_ERR_OOM, "fresnel chirp tables"));
  }
  p->bytes = sizeof(float2) * (size_t)ndist * 2 * (n + nz);
  for (int axis = 0; axis < 2 && p->dct; ++axis) {
    const int N = axis ? nz : n;
    const std::vector<float2> h = dct_mid_tables(N, dist_host, ndist, wavelength, voxelsize);
    float2** dst = axis ? &p->My : &p->Mx;
    if (hipMalloc((void**)dst, sizeof(float2) * h.size()) != hipSuccess ||
        hipMemcpy(*dst, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(ddist);
      return bail(fail(MM_ERR_OOM, "fresnel DCT tables"));
    }
    p->bytes += sizeof(float2) * h.size();
  }
  hipError_t e = hipMemcpy(ddist, dist_host, sizeof(double) * ndist, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(chirp_table, dim3((ndist * 2 * n + 255) / 256), dim3(256), 0, 0, p->Kx, ddist, ndist, 2 * n,
                       wavelength, voxelsize);
    hipLaunchKernelGGL(chirp_table, dim3((ndist * 2 * nz + 255) / 256), dim3(256), 0, 0, p->Ky, ddist, ndist, 2 * nz,
                       wavelength, voxelsize);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
  }
  (void)hipFree(ddist);
  if (e != hipSuccess) return bail(fail(MM_ERR_HIP, std::string("fresnel chirp build: ") + hipGetErrorString(e)));
  *plan = p;
  return MM_OK;
}

int mm_fresnel_fwd(mm_fresnel* p, const void* psi, void* out, int nframes, int j, void* stream) {
  return fresnel_apply(p, psi, out, nframes, j, stream, false);
}

int mm_fresnel_adj(mm_fresnel* p, const void* in, void* out, int nframes, int j, void* stream) {
  return fresnel_apply(p, in, out, nframes, j, stream, true);
}

size_t mm_fresnel_workspace_bytes(const mm_fresnel* p) { return p ? p->bytes : 0; }

void mm_fresnel_destroy(mm_fresnel* p) { fresnel_free(p); }

}  // extern "C"
