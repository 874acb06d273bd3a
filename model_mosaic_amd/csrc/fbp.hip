// FBP filter of the sinogram rows (processing/fbp_filter.py:128-210): edge
// replication to ne = 4 n, FFT, multiplication by real weights w (ne / 2 + 1
// values), inverse FFT / ne, crop back to n -- in place, one kernel.
//
// Two real rows ride in one complex transform of length ne (row A in the real
// part, row B in the imaginary part): the weights are real and even
// (w[k] = w[ne - k] extends the half-spectrum weights), so the inverse
// transform returns filtered A and filtered B as its real and imaginary
// parts -- the same values irfft(rfft(row) w, ne) gives for each row.  A
// workgroup reads its rows once into the register FFT and writes the cropped
// results once: the sinogram makes one HBM round trip.
#include <cmath>

#include "mm_fft.h"

namespace mm {
namespace fbp {

template <int NE>
__global__ MM_TWO_FFT_BUDGET __launch_bounds__(RowReg<NE>::NT) void filter_rows(float* __restrict__ sino, const float* __restrict__ w,
                                                             long long rows, int n, const float2* __restrict__ tw) {
  using Sh = RowReg<NE>;
  __shared__ float2 lds[Sh::LDS];
  int c, t;
  Sh::map(threadIdx.x, c, t);
  const long long ra = 2 * ((long long)blockIdx.x * Sh::C + c), rb = ra + 1;
  const bool has_a = ra < rows, has_b = rb < rows;
  float* pa = sino + (has_a ? ra : 0) * n;
  float* pb = sino + (has_b ? rb : 0) * n;
  const int pad = (NE - n) / 2;
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int i = t + m * Sh::TPC;
    const int j = min(max(i - pad, 0), n - 1);  // edge replication
    v[m] = make_float2(has_a ? pa[j] : 0.f, has_b ? pb[j] : 0.f);
  }
  fft_reg<NE, Sh::TPC, false, 1, false>(v, lds + c * Sh::S, t, tw);
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int i = t + m * Sh::TPC;
    v[m] = cscale(v[m], w[i <= NE / 2 ? i : NE - i]);
  }
  fft_reg<NE, Sh::TPC, true, 1, false>(v, lds + c * Sh::S, t, tw);
  const float s = 1.0f / (float)NE;
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int i = t + m * Sh::TPC - pad;
    if (i < 0 || i >= n) continue;
    if (has_a) pa[i] = v[m].x * s;
    if (has_b) pb[i] = v[m].y * s;
  }
}

// The same filter on transforms of length L = 2 n.  With h = irfft(w, 4 n)
// (real and even) the padded convolution splits into
//   y[i] = sum_j x[j] h[i - j] + x[0] A[i] + x[n - 1] B[i],   i, j in [0, n),
// A[i] = sum of h over the lags of the left padding, B[i] over the right
// padding.  The first term only sees lags |i - j| < n, so it is a circular
// convolution of length 2 n with g = h restricted to those lags, whose
// spectrum G is real (g is even): the rows occupy points [0, n) (m < PT/2,
// the zero half folds out of the first radix pass and the unused half of the
// outputs out of the last), two rows ride in one complex transform as above.
// G (with the 1/L), A and B come from the host in FP64.
template <int L>
__global__ MM_TWO_FFT_BUDGET __launch_bounds__(RowReg<L>::NT) void filter_rows_2n(float* __restrict__ sino,
                                                                const float* __restrict__ G,
                                                                const float* __restrict__ A,
                                                                const float* __restrict__ B, long long rows,
                                                                const float2* __restrict__ tw) {
  using Sh = RowReg<L>;
  constexpr int n = L / 2, MH = Sh::PT / 2;  // points [0, n) = m < PT/2
  static_assert(Sh::PT % 2 == 0 && Sh::PT * Sh::TPC == L, "n is a multiple of TPC");
  __shared__ float2 lds[Sh::LDS];
  int c, t;
  Sh::map(threadIdx.x, c, t);
  const long long ra = 2 * ((long long)blockIdx.x * Sh::C + c), rb = ra + 1;
  const bool has_a = ra < rows, has_b = rb < rows;
  float* pa = sino + (has_a ? ra : 0) * n;
  float* pb = sino + (has_b ? rb : 0) * n;
  // edge values, read before any lane of the workgroup writes
  const float a0 = has_a ? pa[0] : 0.f, a1 = has_a ? pa[n - 1] : 0.f;
  const float b0 = has_b ? pb[0] : 0.f, b1 = has_b ? pb[n - 1] : 0.f;
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int i = t + m * Sh::TPC;
    v[m] = m < MH ? make_float2(has_a ? pa[i] : 0.f, has_b ? pb[i] : 0.f) : make_float2(0.f, 0.f);
  }
  fft_reg<L, Sh::TPC, false, 1, false>(v, lds + c * Sh::S, t, tw);
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int i = t + m * Sh::TPC;
    v[m] = cscale(v[m], G[i <= L / 2 ? i : L - i]);
  }
  fft_reg<L, Sh::TPC, true, 1, false>(v, lds + c * Sh::S, t, tw);
  // the edge tables for every point first (read pairwise next to the stores they were waited on in pairs)
  float ai[MH], bi[MH];
#pragma unroll
  for (int m = 0; m < MH; ++m) {
    ai[m] = A[t + m * Sh::TPC];
    bi[m] = B[t + m * Sh::TPC];
  }
#pragma unroll
  for (int m = 0; m < MH; ++m) {
    const int i = t + m * Sh::TPC;
    if (has_a) pa[i] = fmaf(a1, bi[m], fmaf(a0, ai[m], v[m].x));
    if (has_b) pb[i] = fmaf(b1, bi[m], fmaf(b0, ai[m], v[m].y));
  }
}

// The length-2n form for n above 8192 (2 n past one workgroup's transform), on transforms of length N = n:
// with the rows at points [0, n) and h_j = exp(-i pi j / n),
//   Z[2m] = FFT_n(x)[m],  Z[2m+1] = FFT_n(x_j h_j)[m]   (frequency parity),
// times G, then by output parity  y[2r] = IFFT_n(Z[k] + Z[k+n])[r],  y[2r+1] = IFFT_n((Z[k] - Z[k+n]) conj h_k)[r],
// r < n/2.  Z[k] and Z[k+n] have the parity of k: Z[2m] + Z[2m+n] = V[m] + V[m + n/2] sits in one thread
// (registers q and q + PT/2), and one LDS pass interleaves the even and odd k for each inverse transform.
// h_j = tw_n[j / 2] (j even) or tw_n[(j - 1) / 2] exp(-i pi / n) (j odd), from the length-n bank table.
template <int N>
using SplitRowShape = RegShape<N, N>;

template <int N>
__global__ __launch_bounds__(SplitRowShape<N>::NT) void filter_rows_2n_split(
    float* __restrict__ sino, const float* __restrict__ G, const float* __restrict__ A, const float* __restrict__ B,
    long long rows, const float2* __restrict__ tw, float2 hodd) {
  using Sh = SplitRowShape<N>;
  constexpr int n = N, PT = Sh::PT, TPC = Sh::TPC, H = PT / 2;
  static_assert(Sh::C == 1 && Sh::NT == TPC && PT % 2 == 0, "one transform per workgroup");
  __shared__ float2 lds[Sh::LDS];
  const int t = threadIdx.x;
  const long long ra = 2 * (long long)blockIdx.x, rb = ra + 1;
  const bool has_a = ra < rows, has_b = rb < rows;
  float* pa = sino + (has_a ? ra : 0) * n;
  float* pb = sino + (has_b ? rb : 0) * n;
  const float a0 = has_a ? pa[0] : 0.f, a1 = has_a ? pa[n - 1] : 0.f;
  const float b0 = has_b ? pb[0] : 0.f, b1 = has_b ? pb[n - 1] : 0.f;
  auto hsplit = [&](int j) {
    const float2 w = tw[j >> 1];
    return (j & 1) ? cmul(w, hodd) : w;
  };
  float2 x[PT], v[PT], g[PT];
#pragma unroll
  for (int q = 0; q < PT; ++q) {
    const int j = t + q * TPC;
    x[q] = make_float2(has_a ? pa[j] : 0.f, has_b ? pb[j] : 0.f);
    v[q] = x[q];
  }
  fft_reg<N, TPC, false, 1, false>(v, lds, t, tw);
  // every later index goes through opaque_index: the table loads of one stage stay after the transform
  // before it (hoisted, they would hold registers across it)
  const int t1 = opaque_index(t);
#pragma unroll
  for (int q = 0; q < PT; ++q) {
    const int k = 2 * (t1 + q * TPC);  // Z[k] G[k], G[2n - k] past n (G holds n + 1 values)
    g[q] = cscale(v[q], G[k <= n ? k : 2 * n - k]);
    v[q] = cmul(x[q], hsplit(t1 + q * TPC));
  }
  fft_reg<N, TPC, false, 1, false>(v, lds, t1, opaque_twiddles(tw));
  const int t2 = opaque_index(t1);
#pragma unroll
  for (int q = 0; q < PT; ++q) {
    const int m = t2 + q * TPC, k = 2 * m + 1;
    v[q] = cscale(v[q], G[k <= n ? k : 2 * n - k]);
  }
  // even / odd k of the two inverse inputs: a (sum) into g, b (difference) into x
#pragma unroll
  for (int q = 0; q < H; ++q) {
    const int m = t2 + q * TPC;  // m < n/2: k = 2m and 2m + 1
    const float2 se = cadd(g[q], g[q + H]), de = csub(g[q], g[q + H]);
    const float2 so = cadd(v[q], v[q + H]), dodd = csub(v[q], v[q + H]);
    g[q] = se;
    g[q + H] = so;
    x[q] = cmul(de, cconj(hsplit(2 * m)));
    x[q + H] = cmul(dodd, cconj(hsplit(2 * m + 1)));
  }
  // natural order of a (k = 2m from g[q], 2m + 1 from g[q + H]) through LDS, then IFFT: y[2r]
  __syncthreads();  // the forward transforms' last LDS reads are done
#pragma unroll
  for (int q = 0; q < H; ++q) {
    const int m = t2 + q * TPC;
    lds[lpad(2 * m)] = g[q];
    lds[lpad(2 * m + 1)] = g[q + H];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PT; ++q) v[q] = lds[lpad(t2 + q * TPC)];
  __syncthreads();
  fft_reg<N, TPC, true, 1, false>(v, lds, t2, opaque_twiddles(tw));
  const int t3 = opaque_index(t2);
#pragma unroll
  for (int q = 0; q < H; ++q) {
    const int i = 2 * (t3 + q * TPC);
    const float ai = A[i], bi = B[i];
    if (has_a) pa[i] = fmaf(a1, bi, fmaf(a0, ai, v[q].x));
    if (has_b) pb[i] = fmaf(b1, bi, fmaf(b0, ai, v[q].y));
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < H; ++q) {
    const int m = t3 + q * TPC;
    lds[lpad(2 * m)] = x[q];
    lds[lpad(2 * m + 1)] = x[q + H];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PT; ++q) v[q] = lds[lpad(t3 + q * TPC)];
  __syncthreads();
  fft_reg<N, TPC, true, 1, false>(v, lds, t3, opaque_twiddles(tw));
  const int t4 = opaque_index(t3);
#pragma unroll
  for (int q = 0; q < H; ++q) {
    const int i = 2 * (t4 + q * TPC) + 1;
    const float ai = A[i], bi = B[i];
    if (has_a) pa[i] = fmaf(a1, bi, fmaf(a0, ai, v[q].x));
    if (has_b) pb[i] = fmaf(b1, bi, fmaf(b0, ai, v[q].y));
  }
}

template <int N>
static void launch_2n_split(float* sino, const float* G, const float* A, const float* B, long long rows,
                            const float2* tw, hipStream_t st) {
  const double ang = -M_PI / N;
  const float2 hodd = make_float2((float)std::cos(ang), (float)std::sin(ang));
  hipLaunchKernelGGL(filter_rows_2n_split<N>, dim3((unsigned)((rows + 1) / 2)), dim3(SplitRowShape<N>::NT), 0, st,
                     sino, G, A, B, rows, tw + tw_offset(N), hodd);
}

template <int L>
static void launch_2n(float* sino, const float* G, const float* A, const float* B, long long rows, const float2* tw,
                      hipStream_t st) {
  using Sh = RowReg<L>;
  const long long pairs = (rows + 1) / 2;
  hipLaunchKernelGGL(filter_rows_2n<L>, dim3((unsigned)((pairs + Sh::C - 1) / Sh::C)), dim3(Sh::NT), 0, st, sino, G,
                     A, B, rows, tw + tw_offset(L));
}

template <int NE>
static void launch(float* sino, const float* w, long long rows, int n, const float2* tw, hipStream_t st) {
  using Sh = RowReg<NE>;
  const long long pairs = (rows + 1) / 2;
  hipLaunchKernelGGL(filter_rows<NE>, dim3((unsigned)((pairs + Sh::C - 1) / Sh::C)), dim3(Sh::NT), 0, st, sino, w,
                     rows, n, tw + tw_offset(NE));
}

}  // namespace fbp
}  // namespace mm

extern "C" {

int mm_fbp_filter(float* sino_dev, const float* w_dev, long long rows, int n, void* stream) {
  MM_ARG_CHECK(rows >= 0, "mm_fbp_filter: rows < 0");
  MM_ARG_CHECK(n >= 4 && mm::is_fft_len(4L * n) && 4 * n >= 64 && 4 * n <= mm::kMaxFFT,
               "mm_fbp_filter: 4n must be an FFT length 2^k or 3 * 2^k in [64, 16384], got n=" + std::to_string(n));
  if (rows == 0) return MM_OK;
  MM_ARG_CHECK(sino_dev && w_dev, "mm_fbp_filter: null pointer");
  MM_ARG_CHECK((rows + 1) / 2 < (1LL << 31), "mm_fbp_filter: too many rows");
  const float2* tw = nullptr;
  int rc = mm::twiddle_bank(&tw);
  if (rc) return rc;
  MM_DISPATCH_FFT(4 * n, (mm::fbp::launch<LL>(sino_dev, w_dev, rows, n, tw, (hipStream_t)stream)));
  return mm::check_launch("fbp filter_rows");
}

int mm_fbp_filter_2n(float* sino_dev, const float* G_dev, const float* A_dev, const float* B_dev, long long rows,
                     int n, void* stream) {
  MM_ARG_CHECK(rows >= 0, "mm_fbp_filter_2n: rows < 0");
  MM_ARG_CHECK(n >= 8 && mm::is_fft_len(2L * n) && 2 * n >= 16 && n <= mm::kMaxFFT,
               "mm_fbp_filter_2n: 2n must be an FFT length 2^k or 3 * 2^k in [16, 32768], got n=" + std::to_string(n));
  if (rows == 0) return MM_OK;
  MM_ARG_CHECK(sino_dev && G_dev && A_dev && B_dev, "mm_fbp_filter_2n: null pointer");
  MM_ARG_CHECK((rows + 1) / 2 < (1LL << 31), "mm_fbp_filter_2n: too many rows");
  const float2* tw = nullptr;
  int rc = mm::twiddle_bank(&tw);
  if (rc) return rc;
  if (2 * n > mm::kMaxFFT) {  // two n-point halves per row pair
    switch (n) {
      case 12288: mm::fbp::launch_2n_split<12288>(sino_dev, G_dev, A_dev, B_dev, rows, tw, (hipStream_t)stream); break;
      case 16384: mm::fbp::launch_2n_split<16384>(sino_dev, G_dev, A_dev, B_dev, rows, tw, (hipStream_t)stream); break;
      default: return mm::fail(MM_ERR_UNSUPPORTED, "mm_fbp_filter_2n: n=" + std::to_string(n) + " not built");
    }
    return mm::check_launch("fbp filter_rows_2n_split");
  }
  MM_DISPATCH_FFT(2 * n, (mm::fbp::launch_2n<LL>(sino_dev, G_dev, A_dev, B_dev, rows, tw, (hipStream_t)stream)));
  return mm::check_launch("fbp filter_rows_2n");
}

}  // extern "C"
