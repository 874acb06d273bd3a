// Shared helpers for the gfx950 USFFT Radon / Fresnel library: status codes,
// the thread-local error message, complex arithmetic and the twiddle bank.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/mosaic_hip.h"

namespace mm {

// ---- error state -----------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define MM_HIP_TRY(expr)                                                             \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess)                                                            \
      return ::mm::fail(MM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define MM_ARG_CHECK(cond, msg) \
  do {                          \
    if (!(cond)) return ::mm::fail(MM_ERR_ARG, msg); \
  } while (0)

// Checks the launch that was just issued.
int check_launch(const char* what);

// ---- complex helpers -------------------------------------------------------
__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__host__ __device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__host__ __device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ---- element-wise steps of step3_propagation.py:94-110 ----------------------
// One definition for the stand-alone kernels and for the Fresnel passes that
// fuse them, so both give the same bits.
// psi = exp(-phase * inv_beta_ratio) * (cos phase + i sin phase), phase = p * inv_norm (HIP's precise expf / sincosf)
__device__ __forceinline__ float2 psi_of(float p, float inv_norm, float inv_beta_ratio) {
  const float phase = p * inv_norm;
  const float atten = expf(-phase * inv_beta_ratio);
  float s, c;
  sincosf(phase, &s, &c);
  return make_float2(atten * c, atten * s);
}
__device__ __forceinline__ float abs2_of(float2 v) { return v.x * v.x + v.y * v.y; }

// ---- twiddle bank ----------------------------------------------------------
// For every FFT length L = 2^k in [2, kMaxFFT] and L = 3 * 2^k in [12, kMaxFFT]
// the bank holds exp(-2 pi i t / L), t in [0, L), at element offset
// tw_offset(L) (computed in FP64 on the host).
// After them, at kTwGrid256, the grid exp(-2 pi i e k / 256) at 16 e + k (e, k < 16):
// the twiddles of a radix-16 pass after one radix-16 pass, with the 16 values of
// k for one e in one 128-byte line.
constexpr int kMaxFFT = 16384;
constexpr long kTwGrid256 = 2L * kMaxFFT + 2L * (3 * kMaxFFT / 4);
constexpr long kTwiddleBank = kTwGrid256 + 256;  // elements
__host__ __device__ inline constexpr long tw_offset(int L) {
  return (L & (L - 1)) == 0 ? (long)L : 2L * kMaxFFT + L;  // 3 * 2^k: [2 kMax + L, 2 kMax + 2 L)
}
// Returns the device pointer of the bank on the current device (built on first use).
int twiddle_bank(const float2** out);

// Plan dimensions for the modules that compose plans (model.hip, mosaic.hip).
void radon_dims(const mm_radon* p, int* n, int* ntheta, int* nz_max);
// Radon plan creation with optional caller-owned half-spectrum (fde) and
// representative (gh) buffers shared by plans that never run concurrently,
// and optionally without the adjoint tables (radon.hip).
int radon_create_impl(mm_radon** plan, int n, int nz_max, int ntheta, const double* theta_host, unsigned flags,
                      float2* fde_ext, float2* gh_ext, bool fwd_only);
// R on caller-given half-spectrum and representative buffers (at least radon_fde_elems / radon_gh_elems of
// the call's nz) instead of the plan's: plans shared by concurrent graph branches, one workspace each.
int radon_fwd_ws(mm_radon* p, const float* obj, float* sino, int nz, hipStream_t st, float2* fde, float2* gh);
size_t radon_fde_elems(int n, int nz_max);
size_t radon_gh_elems(int n, int nz_max, int ntheta);
void fresnel_dims(const mm_fresnel* p, int* n, int* nz, int* ndist);
// Fused model passes on the DCT form of a mirror-padded Fresnel plan (fresnel.hip); MM_ERR_UNSUPPORTED
// when the plan has no DCT form (fresnel_fused_ok false).
bool fresnel_fused_ok(const mm_fresnel* p);
int fresnel_model_parts(const mm_fresnel* p, int nframes, long long* parts);
int fresnel_model_fwd(const mm_fresnel* p, const float* P, const float* B, float k, float2* u, const float* d,
                      double* partial, int* nparts, int nframes, int j, hipStream_t st);
int fresnel_model_bwd(const mm_fresnel* p, float2* u, const float* d, float* P, float* B, float k, float gscale,
                      int nframes, int j, hipStream_t st);
int fresnel_intensity(const mm_fresnel* p, const float* P, float inv_norm, float inv_beta_ratio, float2* work,
                      float* out, int nframes, int j, hipStream_t st);

inline bool is_pow2(long v) { return v > 0 && (v & (v - 1)) == 0; }
// FFT lengths the kernels are built for: 2^k and 3 * 2^k.
inline bool is_fft_len(long v) { return is_pow2(v) || (v % 3 == 0 && is_pow2(v / 3)); }

// gfx950 dispatches workgroup w (linear id, x fastest) to XCD w mod 8, and each
// XCD has its own L2.  xcd_remap turns w into a logical id such that every XCD
// runs one contiguous range of logical ids: workgroups whose data share cache
// lines (neighbouring columns, tiles with overlapping halos) then meet in one
// L2 instead of each XCD fetching the lines again.  A bijection on [0, total).
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_remap(unsigned w, unsigned total) {
  const unsigned per = total / kXcds, rem = total % kXcds;
  const unsigned k = w % kXcds, pos = w / kXcds;
  return k * per + (k < rem ? k : rem) + pos;
}
// Grouped XCD order: for every k, logical ids [8 G k + G x, 8 G k + G x + G)
// run on XCD x, so G neighbouring work items (tiles that share halo and
// input lines) meet in one L2 while dense regions still spread over all
// XCDs.  Identity for G <= 1 and past the last whole span of 8 G.  A
// bijection on [0, total).
__device__ __forceinline__ unsigned xcd_group(unsigned w, unsigned total, unsigned G) {
  if (G <= 1) return w;
  const unsigned span = kXcds * G, full = total - total % span;
  if (w >= full) return w;
  const unsigned x = w % kXcds, pos = w / kXcds;
  return (pos / G) * span + x * G + pos % G;
}
inline int ilog2(long v) { int r = 0; while ((1L << r) < v) ++r; return r; }

}  // namespace mm
