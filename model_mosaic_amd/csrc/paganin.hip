// Single-distance Paganin phase retrieval (processing/paganin.py:29-92):
//   out = -log(max(irfft2(H . rfft2(I)), 1e-6)),  H = alpha / (lambda z |k|^2 / (4 pi) + alpha)
// over each (nz, n) frame.  H is real and even in (ky, kx), so two real frames
// ride in one complex frame (frame a in the real part, frame b in the
// imaginary part): the full complex 2-D transform, the product with H and the
// inverse return filtered a and filtered b as the real and imaginary parts --
// the same values irfft2(H rfft2(.)) gives for each frame.  H is not
// separable, so the 2-D transform is three HBM passes over a (pairs, nz, n)
// complex64 workspace:
//   rows_fwd: FFT along x of the packed rows (reads the two real frames);
//   cols:     FFT along y, times H / (n nz), IFFT along y (in place);
//   rows_inv: IFFT along x, clip, -log, unpack (writes the two real frames).
// H is formed on the fly from the per-axis tables fy^2 and fx^2 (float32, as
// the reference builds them) with one reciprocal: H = 1 / (1 + c |k|^2),
// c = lambda z / (4 pi alpha).
#include <cmath>
#include <string>
#include <vector>

#include "mm_fft.h"

namespace mm {
namespace paganin {

// Row pass, forward: unit u = (pair p, row y); transform of (a[y], b[y]).
template <int L>
__global__ MM_TWO_FFT_BUDGET __launch_bounds__(RowReg<L>::NT) void rows_fwd(const float* __restrict__ in, float2* __restrict__ ws,
                                                         int nz, int nframes, long long units,
                                                         const float2* __restrict__ tw) {
  using Sh = RowReg<L>;
  __shared__ float2 lds[Sh::LDS];
  int c, t;
  Sh::map(threadIdx.x, c, t);
  const long long u = (long long)blockIdx.x * Sh::C + c;
  const bool valid = u < units;
  const long long uu = valid ? u : 0;
  const int p = (int)(uu / nz), y = (int)(uu - (long long)p * nz);
  const bool has_b = valid && 2 * p + 1 < nframes;
  const float* ra = in + ((long long)(2 * p) * nz + y) * L;
  const float* rb = ra + (long long)nz * L;
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int x = t + m * Sh::TPC;
    v[m] = make_float2(valid ? ra[x] : 0.f, has_b ? rb[x] : 0.f);
  }
  fft_reg<L, Sh::TPC, false>(v, lds + c * Sh::S, t, tw);
  if (!valid) return;
  float2* dst = ws + uu * L;
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) dst[t + m * Sh::TPC] = v[m];
}

// Column pass: C columns of one pair (blockIdx.y); FFT along y, times H / (n nz), IFFT.
// Column-fastest lanes (8 columns per workgroup at 2048 points: 3.44 -> 3.10 ms per 128 frames of 2048^2
// against 4 columns in the interleaved mapping).
template <int L>
using PagColShape = ColumnShape<L, kBatchElems>;
template <int L>
__global__ MM_TWO_FFT_BUDGET __launch_bounds__(PagColShape<L>::NT) void cols(float2* __restrict__ ws, int n,
                                                                       const float* __restrict__ fy2,
                                                                       const float* __restrict__ fx2, float ca,
                                                                       float scale, const float2* __restrict__ tw) {
  using Sh = PagColShape<L>;
  __shared__ float2 lds[Sh::LDS];
  int c, t;
  Sh::map(threadIdx.x, c, t);
  // the column groups sharing each 128-byte row line run on one XCD
  const unsigned lw = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int bx = (int)(lw % gridDim.x), pair = (int)(lw / gridDim.x);
  const int col = bx * Sh::C + c;
  const bool live = col < n;
  float2* b = ws + (long long)pair * L * n + (live ? col : 0);
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) v[m] = live ? b[(long long)(t + m * Sh::TPC) * n] : make_float2(0.f, 0.f);
  fft_reg<L, Sh::TPC, false, 1, false>(v, lds + c * Sh::S, t, tw);
  const float kx2 = live ? fx2[col] : 0.f;
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const float h = scale * __builtin_amdgcn_rcpf(fmaf(ca, fy2[t + m * Sh::TPC] + kx2, 1.0f));
    v[m] = cscale(v[m], h);
  }
  fft_reg<L, Sh::TPC, true, 1, false>(v, lds + c * Sh::S, t, tw);
  if (!live) return;
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) b[(long long)(t + m * Sh::TPC) * n] = v[m];
}

// Row pass, inverse: IFFT along x, clip at 1e-6, -log, unpack the two frames.
template <int L>
__global__ MM_TWO_FFT_BUDGET __launch_bounds__(RowReg<L>::NT) void rows_inv(const float2* __restrict__ ws, float* __restrict__ out,
                                                         int nz, int nframes, long long units,
                                                         const float2* __restrict__ tw) {
  using Sh = RowReg<L>;
  __shared__ float2 lds[Sh::LDS];
  int c, t;
  Sh::map(threadIdx.x, c, t);
  const long long u = (long long)blockIdx.x * Sh::C + c;
  const bool valid = u < units;
  const long long uu = valid ? u : 0;
  const int p = (int)(uu / nz), y = (int)(uu - (long long)p * nz);
  const bool has_b = valid && 2 * p + 1 < nframes;
  const float2* src = ws + uu * L;
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) v[m] = valid ? src[t + m * Sh::TPC] : make_float2(0.f, 0.f);
  fft_reg<L, Sh::TPC, true>(v, lds + c * Sh::S, t, tw);
  if (!valid) return;
  float* oa = out + ((long long)(2 * p) * nz + y) * L;
  float* ob = oa + (long long)nz * L;
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int x = t + m * Sh::TPC;
    oa[x] = -logf(fmaxf(v[m].x, 1e-6f));
    if (has_b) ob[x] = -logf(fmaxf(v[m].y, 1e-6f));
  }
}

}  // namespace paganin
}  // namespace mm

using namespace mm;
using namespace mm::paganin;

struct mm_paganin {
  int dev = 0;
  int n = 0, nz = 0, nframes_max = 0;
  float ca = 0.f;
  float* fy2 = nullptr;  // nz
  float* fx2 = nullptr;  // n
  float2* ws = nullptr;  // ceil(nframes_max / 2) x nz x n
  const float2* tw = nullptr;
  size_t bytes = 0;
};

static void paganin_free(mm_paganin* p) {
  if (!p) return;
  if (p->fy2) (void)hipFree(p->fy2);
  if (p->fx2) (void)hipFree(p->fx2);
  if (p->ws) (void)hipFree(p->ws);
  delete p;
}

template <int L>
static void launch_rows(const mm_paganin* p, const float* in, float* out, int nframes, bool inv, hipStream_t st) {
  using Sh = RowReg<L>;
  const long long units = (long long)((nframes + 1) / 2) * p->nz;
  const dim3 grid((unsigned)((units + Sh::C - 1) / Sh::C));
  if (inv)
    hipLaunchKernelGGL(rows_inv<L>, grid, dim3(Sh::NT), 0, st, p->ws, out, p->nz, nframes, units, p->tw + tw_offset(L));
  else
    hipLaunchKernelGGL(rows_fwd<L>, grid, dim3(Sh::NT), 0, st, in, p->ws, p->nz, nframes, units, p->tw + tw_offset(L));
}

template <int L>
static void launch_cols(const mm_paganin* p, int nframes, hipStream_t st) {
  using Sh = PagColShape<L>;
  const float scale = (float)(1.0 / ((double)p->n * p->nz));
  hipLaunchKernelGGL(cols<L>, dim3((unsigned)((p->n + Sh::C - 1) / Sh::C), (unsigned)((nframes + 1) / 2)),
                     dim3(Sh::NT), 0, st, p->ws, p->n, p->fy2, p->fx2, p->ca, scale, p->tw + tw_offset(L));
}

// squared fftfreq(len, d) in float32, as the reference builds fy / fx (float32 grid, squared in float32)
static std::vector<float> freq2(int len, double d) {
  std::vector<float> f(len);
  for (int k = 0; k < len; ++k) {
    const float v = (float)((double)(k < (len + 1) / 2 ? k : k - len) / (len * d));
    f[k] = v * v;
  }
  return f;
}

extern "C" {

int mm_paganin_create(mm_paganin** plan, int n, int nz, int nframes_max, double wavelength, double voxelsize,
                      double distance, double alpha) {
  MM_ARG_CHECK(plan != nullptr, "mm_paganin_create: plan is NULL");
  *plan = nullptr;
  auto size_ok = [](int v) { return is_fft_len(v) && v >= 16 && v <= kMaxFFT; };
  MM_ARG_CHECK(size_ok(n), "mm_paganin_create: n must be 2^k or 3 * 2^k in [16, 16384], got " + std::to_string(n));
  MM_ARG_CHECK(size_ok(nz), "mm_paganin_create: nz must be 2^k or 3 * 2^k in [16, 16384], got " + std::to_string(nz));
  MM_ARG_CHECK(nframes_max >= 1, "mm_paganin_create: nframes_max must be >= 1");
  MM_ARG_CHECK(voxelsize > 0 && wavelength >= 0 && distance >= 0 && alpha > 0,
               "mm_paganin_create: need voxelsize > 0, alpha > 0, wavelength and distance >= 0");
  mm_paganin* p = new mm_paganin();
  auto bail = [&](int code) {
    paganin_free(p);
    return code;
  };
  int rc = twiddle_bank(&p->tw);
  if (rc) return bail(rc);
  if (hipGetDevice(&p->dev) != hipSuccess) return bail(fail(MM_ERR_HIP, "hipGetDevice"));
  p->n = n;
  p->nz = nz;
  p->nframes_max = nframes_max;
  p->ca = (float)(wavelength * distance / (4.0 * M_PI * alpha));
  const std::vector<float> fy2 = freq2(nz, voxelsize), fx2 = freq2(n, voxelsize);
  const size_t ws = sizeof(float2) * (size_t)((nframes_max + 1) / 2) * nz * n;
  if (hipMalloc(&p->fy2, sizeof(float) * nz) != hipSuccess || hipMalloc(&p->fx2, sizeof(float) * n) != hipSuccess)
    return bail(fail(MM_ERR_OOM, "paganin frequency tables"));
  if (hipMalloc(&p->ws, ws) != hipSuccess)
    return bail(fail(MM_ERR_OOM, "paganin workspace (" + std::to_string(ws) + " bytes)"));
  p->bytes = ws + sizeof(float) * (size_t)(n + nz);
  if (hipMemcpy(p->fy2, fy2.data(), sizeof(float) * nz, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->fx2, fx2.data(), sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(MM_ERR_HIP, "paganin table upload"));
  *plan = p;
  return MM_OK;
}

int mm_paganin_retrieve(mm_paganin* p, const float* in, float* out, int nframes, void* stream) {
  MM_ARG_CHECK(p != nullptr, "mm_paganin_retrieve: plan is NULL");
  MM_ARG_CHECK(nframes >= 0 && nframes <= p->nframes_max,
               "mm_paganin_retrieve: nframes=" + std::to_string(nframes) + " outside [0, " +
                   std::to_string(p->nframes_max) + "]");
  if (nframes == 0) return MM_OK;
  MM_ARG_CHECK(in && out, "mm_paganin_retrieve: null pointer");
  hipStream_t st = (hipStream_t)stream;
  MM_DISPATCH_FFT(p->n, (launch_rows<LL>(p, in, out, nframes, false, st)));
  int rc = check_launch("paganin rows_fwd");
  if (rc) return rc;
  MM_DISPATCH_FFT(p->nz, (launch_cols<LL>(p, nframes, st)));
  if ((rc = check_launch("paganin cols"))) return rc;
  MM_DISPATCH_FFT(p->n, (launch_rows<LL>(p, in, out, nframes, true, st)));
  return check_launch("paganin rows_inv");
}

size_t mm_paganin_workspace_bytes(const mm_paganin* p) { return p ? p->bytes : 0; }

void mm_paganin_destroy(mm_paganin* p) { paganin_free(p); }

}  // extern "C"
