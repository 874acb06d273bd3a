// USFFT Radon transform R and its adjoint RT for gfx950.
//
// Operator (processing/tomo.py:121-188, processing/kernels.py:16-93,274-349):
//   R x  = Re( c1d . IFFT_n( c1d . G( FFT2_2n( pad(phi . x) ) ) ) ) / 4
//   RT y = Re or complex( phi . crop( IFFT2_2n( G^T( c1d . FFT_n(c1d . y) ) ) ) )
// G samples the 2n x 2n Fourier grid at the polar points
//   X = 2n x0, Y = 2n y0,  x0 = (r - n/2)/n cos(theta),  y0 = -(r - n/2)/n sin(theta)
// with the (2m+1)^2-tap Gaussian window (pi/mu) exp(-pi^2/mu |xi - l/2n|^2), m = 4.
//
// Design (all FP32 arithmetic except the FP64 polar coordinates):
//  * Real input => Hermitian spectrum.  Only the half plane l0 in [0, n] of the
//    rfft2 is ever stored (fde: nz planes of 2n raw fftfreq rows x (n+1) columns,
//    in blocks of 4 columns, see hs_idx).
//  * The samples r and n - r of one angle sit at mirrored frequencies, so for
//    real data g(n - r) = conj(g(r)).  Each angle is therefore gathered at the
//    H = n/2 + 1 "representative" samples of its half line in the right half
//    plane (X >= 0); the other half of the line is their conjugate.  The
//    unpaired r = 0 sample is represented by its mirror.
//  * The adjoint scatter folds the Hermitian part of the full grid into the
//    half plane: a contribution c at tap k lands at k when k0 in [0, n] and as
//    conj(c) at -k when k0 in {0} U [n, 2n); the self-mirrored (r = n/2) and
//    unpaired (r = 0) samples carry weight 1/2.  A C2R row transform then
//    returns Re(IFFT2(full grid)), which is what the reference keeps.
//    A complex sinogram is split into the two Hermitian channels of its real
//    and imaginary parts (one complex row FFT gives both).
//  * The gather is tiled over the Fourier plane: samples are binned (host,
//    FP64, at plan time) into quadtree tiles of at most 1024 samples.  A
//    gather workgroup stages its tile + 4-point halo of the half spectrum in
//    LDS (applying the conj reflection while loading) and each lane sums its
//    81 taps from LDS; per-sample weights stay in registers across the z loop.
//  * The scatter is owner-computes, without atomics: a workgroup owns a 16 x 16
//    tile of the output half plane and reads the values of the entries touching
//    it from LDS.  By default (scatter_mfma) each 4 x 4 patch of the tile is a
//    dense product of the entries' values and Gaussian weights on the matrix
//    cores (FP32); with MM_RADON_LIST_TILES (FP32), MM_RADON_FP64_TILES (FP64)
//    or MM_RADON_KAHAN_TILES (compensated FP32) each lane owns one point and
//    walks its plan-time list of (entry, tap) slots (scatter_lists).  The
//    summation order is fixed either way, so RT is deterministic.  Tiles with
//    more entries than one work item holds (near the origin, where every angle
//    passes) are split into partial tiles that scatter_reduce sums.
//  * FFTs are the register-resident Stockham transforms of mm_fft.h (up to
//    16384 points per workgroup); the 2n-point passes of n > 8192 run as two
//    n-point halves (r1_rows_split, col_fft_split, rt4_rows_c2r_split).
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#include "mm_fft.h"

namespace mm {
namespace radon {

constexpr int M = 4;             // window half width (m of processing/tomo.py:41 for every n)
constexpr int TAPS = 2 * M + 1;  // 9
constexpr int ROOT = 64;         // quadtree root tile
constexpr int FLOOR = 8;         // quadtree floor tile

constexpr int SMAX = 1024;  // samples per gather work item

// Gather patch pitch (complex values per LDS row) for tile widths 64/32/16/8:
// the smallest pitch >= w + 2M that is 4 (mod 8) with pitch / 4 odd, so the
// staging writes (4 columns x 8 rows per half-wave) hit 32 distinct bank
// pairs.  Taps are read as 16-byte pieces (a 9-tap row = five ds_read_b128
// from the window base rounded down to even; ds_read2_b64 moves half the bytes
// per cycle), which the LDS serves 16 lanes per cycle: the plan gives the 16
// samples of each quarter-wave distinct (base / 2) mod 16, so the pieces of
// one read never share a bank (64 banks = 16 groups of 16 bytes).
__host__ __device__ __forceinline__ constexpr int patch_pitch(int w) {
  return w >= 64 ? 76 : (w >= 32 ? 44 : (w >= 16 ? 28 : 20));
}

// Gather kernel shapes; each holds two workgroups per CU, so that one's
// staging loads overlap the other's taps.  Root class (tiles wider than 32):
// 512 lanes with two samples each (<= 128 VGPRs, 8 waves) and a float2 patch
// of one z slice (72 x 76, 43 KB: a float4 patch of two slices would take
// 87.5 KB and hold the CU alone).  Leaf class (tiles of width <= 32): 1024
// lanes with one sample each (<= 64 VGPRs, 16 waves: twice the waves per CU
// of the root shape) and a float4 patch of two slices (40 x 44, 28 KB) unless
// the plan asks for one slice.
//   PE: patch capacity (the sink slot for idle staging lanes is PE itself);
//   STG: staging slots per lane (4-column blocks x rows over NT lanes).
template <int WMAX, int NT_>
struct GatherShape {
  static constexpr int NT = NT_, W = WMAX;
  static constexpr int SPT = SMAX / NT;
  static constexpr int PE = (WMAX + 2 * M) * patch_pitch(WMAX);  // even: the 16-byte reads stay inside
  static constexpr int STG = (((WMAX + 2 * M + 3) / 4 + 1) * 4 * (WMAX + 2 * M) + NT - 1) / NT;
  static_assert(SMAX % NT == 0, "samples per lane");
  static constexpr int WAVES_PER_SIMD = NT == SMAX ? 8 : 4;  // the register budget: 64 or 128 VGPRs
};
using GatherRoot = GatherShape<ROOT, SMAX / 2>;
using GatherLeaf = GatherShape<ROOT / 2, SMAX>;
// z slices per patch element of a gather item of width w (plan flag zp for the
// leaf class); the lane colouring of the plan follows it
__host__ __device__ __forceinline__ constexpr int gather_zp(int w, int zp_flag) {
  return w > GatherLeaf::W ? 1 : zp_flag;
}

struct Item {
  int x0, y0;        // tile origin (l0, l1), centred coordinates
  int w, h;          // tile extent
  int start, count;  // samples[start, start + count)
  int slot;          // scatter: -1 = own the tile, >= 0 = partial slot
  int pad;
};

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Half-spectrum layout: column blocks of 4; inside a block the 2n rows are
// contiguous, 4 complex values (one per column) each.  A column FFT workgroup
// then reads one contiguous block, and the row passes read/write 32-byte
// pieces that are contiguous over consecutive rows.
__host__ __device__ __forceinline__ long long hs_idx(int row, int col, int twon) {
  return ((long long)(col >> 2) * twon + row) * 4 + (col & 3);
}

// Workgroup shapes of the row and line passes.  The forward row pass takes
// 16384 points per workgroup for the power-of-two lengths 4096 and 16384 (four
// transforms = 8 rows at L = 4096: 256-byte runs per half-spectrum block, one
// workgroup per CU; 4 % faster than two transforms); L = 8192 (n = 4096),
// 3 * 2^k lengths, the C2R rows and the line passes take 8192 (two workgroups
// per CU at 70 KB of LDS; 16384 points made them slower: 12 % at L = 8192).
template <int L>
using R1Shape = RegShape<L, (pow2c(L) && L >= 4096 && L != 8192 ? kBatchElems : kRowBatch), 0, false>;
template <int L>
using LineShape = RowReg<L>;

// Row pass of R: two object rows per complex FFT.  Row y of slice z, times
// phi_y[y] phi_x[x], zero-padded to 2n (object centred), is transformed and
// the half spectrum kx in [0, n] is written to half-spectrum row n/2 + y.  The
// inputs go straight into the register FFT; the spectrum passes once through
// LDS to pair Z_k with Z_{L-k} and to write the half-spectrum blocks in order.
template <int L>
__global__ __launch_bounds__(R1Shape<L>::NT) void r1_rows_rfft(const float* __restrict__ obj, float2* __restrict__ fde,
                                                           const float* __restrict__ phx, const float* __restrict__ phy,
                                                           int n, int nz, long long fzs, const float2* __restrict__ tw) {
  using Sh = R1Shape<L>;
  constexpr int R2 = 2 * Sh::C;
  __shared__ float2 lds[Sh::LDS];
  // offset of destination row r of this workgroup in the half spectrum (-1: none); offsets from the
  // kernel argument keep the stores global (a pointer read back from LDS would make them flat
  // stores, which the LDS reads of the next element then wait for)
  __shared__ long long rowoff[R2];
  // with one transform (two rows) per workgroup, two workgroups share each 128-byte line of a block
  // (4 rows x 32 bytes): keep them adjacent in the XCD-grouped order, so the line is merged in one L2
  const unsigned wg = Sh::C == 1 ? xcd_group(blockIdx.x, gridDim.x, 2) : blockIdx.x;
  int c, t;
  Sh::map(threadIdx.x, c, t);
  const int half = n / 2;
  const long long npairs = (long long)nz * half;
  if (threadIdx.x < R2) {
    const long long pr = (long long)wg * Sh::C + (threadIdx.x >> 1);
    long long ro = -1;
    if (pr < npairs) {
      const int zz = (int)(pr / half), yy = 2 * (int)(pr - (long long)zz * half) + (threadIdx.x & 1);
      ro = (long long)zz * fzs + 4LL * (half + yy);  // hs_idx(half + yy, col) = this + col part
    }
    rowoff[threadIdx.x] = ro;
  }
  const long long pair = (long long)wg * Sh::C + c;
  const bool valid = pair < npairs;
  int z = 0, y0 = 0;
  if (valid) {
    z = (int)(pair / half);
    y0 = 2 * (int)(pair - (long long)z * half);
  }
  float2* x = lds + c * Sh::S;
  const float* ra = obj + ((long long)z * n + y0) * n;
  const float* rb = ra + n;
  const float pya = valid ? phy[y0] : 0.f, pyb = valid ? phy[y0 + 1] : 0.f;
  // the row (n = L/2 values) occupies points [L/4, 3L/4): m in [PT/4, 3PT/4)
  static_assert(Sh::PT % 4 == 0 && Sh::PT * Sh::TPC == L, "L/4 is a multiple of TPC");
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int xi = t + (m - Sh::PT / 4) * Sh::TPC;
    v[m] = make_float2(0.f, 0.f);
    // no per-lane predicate on the reads (a pair past the last reads rows 0 and 1 and is never stored):
    // predicated reads were issued a few at a time
    if (m >= Sh::PT / 4 && m < 3 * Sh::PT / 4) {
      const float p = phx[xi];
      v[m] = make_float2(ra[xi] * pya * p, rb[xi] * pyb * p);
    }
  }
  fft_reg<L, Sh::TPC, false>(v, x, t, tw);
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) x[lpad(t + m * Sh::TPC)] = v[m];
  __syncthreads();
  // unpack A = (Z_k + conj Z_{L-k})/2 (row y0), B = (Z_k - conj Z_{L-k})/(2i) (row y0+1);
  // thread order: column within block, then the 2C rows of this workgroup, then block.
  // Unrolled: every element's LDS reads are independent of the global stores.
  constexpr int NBLK = L / 8 + 1;  // n / 4 + 1 column blocks
  constexpr int TOT = NBLK * R2 * 4, IT = (TOT + Sh::NT - 1) / Sh::NT;
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int idx = threadIdx.x + i * Sh::NT;
    if (idx >= TOT) continue;
    const int q = idx & 3, rest = idx >> 2;
    const int r = rest % R2, blk = rest / R2;
    const int col = 4 * blk + q;
    const long long ro = rowoff[r];
    if (col > n || ro < 0) continue;
    const float2* xc = lds + (r >> 1) * Sh::S;
    const float2 zk = xc[lpad(col)], zm = xc[lpad(col == 0 ? 0 : L - col)];
    const float2 w = (r & 1) ? make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x))
                             : make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
    fde[ro + (long long)blk * 4 * L + q] = w;  // hs_idx(row, col, L) = 4 row + 4 L (col >> 2) + (col & 3)
  }
}

// Column transform over C adjacent columns of one z plane of the half
// spectrum (blockIdx.y = z), in place.  Forward: the object occupies rows
// [L/4, 3L/4) (the others are zero padding and are not read); inverse: only
// rows [L/4, 3L/4) are written back (the row pass crops to them; it also
// applies this pass's 1/L).  Thread t holds rows t + m TPC, and L/4 is a multiple of TPC
// (PT = 16 or 12), so both ranges are m in [PT/4, 3PT/4), fixed at compile
// time: the zero inputs fold out of the first radix pass and the unused
// outputs out of the last.  supp holds per 4-column block the centred row
// range [lo, hi) (row i is l = i for i < L/2, i - L otherwise) that carries
// spectrum data: the forward pass writes only those rows (the gather reads
// nothing else), the inverse pass reads only those (the scatter writes nothing
// else; the disc of radius n leaves the corners of the plane empty).  The
template <int L>
using ColShape = RegShape<L, kBatchElems>;

//
// BATCH = kBatchElems: 16384 points per workgroup (C = 4 columns, one 32-byte
// block row piece, at L = 4096: one workgroup per CU).  BATCH = kRowBatch (the
// column split knob): half as many columns, two workgroups per CU where the
// LDS allows, the workgroups of a block kept on one XCD so that its L2 fetches
// each 128-byte line once.  (The column-fastest lane
// mapping of the Fresnel and Paganin column passes is slower here: 1.27 ->
// 1.49 ms for the forward pass at n = 2048; a wave already reads one
// contiguous 512-byte run of the block layout.)
template <int L, int BATCH = kBatchElems>
using ColShapeB = RegShape<L, BATCH>;

template <int L, bool INV, int BATCH = kBatchElems>
__global__ __launch_bounds__((ColShapeB<L, BATCH>::NT)) void col_fft(float2* __restrict__ buf, int ncols,
                                                                   long long zs, const int2* __restrict__ supp,
                                                                   const float2* __restrict__ tw, int xg) {
  using Sh = ColShapeB<L, BATCH>;
  __shared__ float2 lds[Sh::LDS];
  int c, t;
  Sh::map(threadIdx.x, c, t);
  const unsigned lw = xcd_group(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y, xg);
  const int col = (int)(lw % gridDim.x) * Sh::C + c;
  const bool live = col < ncols;
  float2* b = buf + (long long)(lw / gridDim.x) * zs;
  const int2 sp = live ? supp[col >> 2] : make_int2(0, 0);
  auto carries = [&](int i) {
    const int l = i < L / 2 ? i : i - L;
    return l >= sp.x && l < sp.y;
  };
  constexpr int M0 = Sh::PT / 4, M1 = 3 * Sh::PT / 4;  // rows [L/4, 3L/4)
  static_assert(Sh::PT % 4 == 0, "L/4 is a multiple of TPC");
  float2 v[Sh::PT];
  // Inverse: every row loaded from an in-bounds column (the grid may hold columns past the plane), the rows
  // off the support zeroed afterwards.  Loads under the per-lane support predicate were issued one at a
  // time, each waited on, for every length but the power-of-two ones from 4096 on one-workgroup-per-CU
  // shapes, which kept all 16 in flight and measured 0.03 ms faster that way at 4096.
  constexpr bool kUncond = INV && !(pow2c(L) && L >= 4096 && BATCH == kBatchElems);
  if constexpr (kUncond) {
    const int colc = min(col, ncols - 1);
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) v[m] = b[hs_idx(t + m * Sh::TPC, colc, L)];
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m)
      if (!(live && carries(t + m * Sh::TPC))) v[m] = make_float2(0.f, 0.f);
  } else {
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) {
      const int i = t + m * Sh::TPC;
      const bool rd = INV ? (live && carries(i)) : (live && m >= M0 && m < M1);
      v[m] = rd ? b[hs_idx(i, col, L)] : make_float2(0.f, 0.f);
    }
  }
  fft_reg<L, Sh::TPC, INV>(v, lds + c * Sh::S, t, tw);
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int i = t + m * Sh::TPC;
    if (INV) {
      if (live && m >= M0 && m < M1) b[hs_idx(i, col, L)] = v[m];
    } else if (live && carries(i)) {
      b[hs_idx(i, col, L)] = v[m];
    }
  }
}

// Polar coordinates of representative sample id = theta * H + k: X = 2k|cos|,
// Y = -2k sigma sin (FP64), split into the nearest grid point and the FP32
// fractional offset.
struct SampleGeom {
  int th, k, rx, ry;
  float fx, fy;
};

// (direction d = dirs[id / H] already read)
__device__ __forceinline__ SampleGeom sample_geom_d(int id, int H, double2 d) {
  SampleGeom g;
  g.th = id / H;
  g.k = id - g.th * H;
  const double X = (2.0 * g.k) * d.x, Y = (2.0 * g.k) * d.y;
  const double rx = rint(X), ry = rint(Y);
  g.rx = (int)rx;
  g.ry = (int)ry;
  g.fx = (float)(X - rx);
  g.fy = (float)(Y - ry);
  return g;
}

__device__ __forceinline__ SampleGeom sample_geom(int id, int H, const double2* __restrict__ dirs) {
  SampleGeom g;
  g.th = id / H;
  g.k = id - g.th * H;
  const double2 d = dirs[g.th];
  const double X = (2.0 * g.k) * d.x, Y = (2.0 * g.k) * d.y;
  const double rx = rint(X), ry = rint(Y);
  g.rx = (int)rx;
  g.ry = (int)ry;
  g.fx = (float)(X - rx);
  g.fy = (float)(Y - ry);
  return g;
}

// Tiled polar gather (forward R): gh[theta, z, k] = C0 sum ey sum ex F_c.
// Thread s computes sample samples[start + s]; its results go through LDS so
// that thread t writes sorted sample t (ascending theta, k: coalesced).
// Each thread's patch staging addresses and reflections are worked out once
// per item; the centring sign (-1)^(k0 + k1) is separable and rides in the
// weights (reflection keeps the parity of k0 and k1), so the z loop only moves
// data.
//
// ZP = z slices per patch element:
//  * ZP = 1: float2 patch; a 9-tap row is read as five 16-byte pieces from the
//    window base rounded down to even (a tenth weight of zero absorbs the
//    parity); lanes of a quarter-wave have distinct (base / 2) mod 16.
//  * ZP = 2: float4 patch holding slices z and z + 1 of each grid point; a tap
//    is one 16-byte read for both slices (9 per row, half the barriers per
//    slice); lanes of a quarter-wave have distinct base mod 16.
// Either way the 16 pieces of one read hit 16 distinct bank groups.
// Staging descriptor idx of a gather item: the half-spectrum offset (within a
// z plane) of patch element idx and its LDS slot << 1 | conj flag; slots
// outside the patch read offset 0 into the sink slot pe (the patch capacity).  Patch
// columns and rows are raw fftfreq indices taken mod 2n (tiles near the top
// and right edges reach past 2n for small n); columns k0 > n are the
// conjugate of the Hermitian mirror (2n - k1, 2n - k0).
__host__ __device__ __forceinline__ int wrap_mod(int v, int m) {
  const int r = v % m;
  return r < 0 ? r + m : r;
}
// soff: element offset inside one z plane of the half spectrum (unsigned: a plane of n = 32768 holds more
// than 2^31 elements).
__host__ __device__ __forceinline__ void gather_stage_desc(const Item& it, int idx, int n, int pe, unsigned& soff,
                                                           int& sdst) {
  const int pw = it.w + 2 * M, ph = it.h + 2 * M;
  const int pp = patch_pitch(it.w);
  const int px0 = it.x0 - M, py0 = it.y0 - M;
  const int bx0 = px0 & ~3;  // first 4-column block touched by the patch
  const int nstage = (((px0 + pw - 1) - bx0) / 4 + 1) * ph * 4;
  const int twon = 2 * n;
  soff = 0;
  sdst = pe << 1;
  if (idx >= nstage) return;
  const int q = idx & 3, rest = idx >> 2;
  const int pr = rest % ph, blk = rest / ph;
  const int l0 = bx0 + 4 * blk + q;
  const int pc = l0 - px0;
  if (pc < 0 || pc >= pw) return;
  const int k0 = wrap_mod(l0, twon), k1 = wrap_mod(py0 + pr, twon);
  int fl = 0;
  if (k0 <= n) {
    soff = (unsigned)hs_idx(k1, k0, twon);
  } else {
    soff = (unsigned)hs_idx(k1 == 0 ? 0 : twon - k1, twon - k0, twon);
    fl = 1;
  }
  sdst = ((pr * pp + pc) << 1) | fl;
}

template <int ZP, class S>
__global__ __launch_bounds__(S::NT) __attribute__((amdgpu_waves_per_eu(S::WAVES_PER_SIMD, S::WAVES_PER_SIMD)))
void gather_tiles(const float2* __restrict__ fde, float2* __restrict__ gh, const Item* __restrict__ items,
                  const int* __restrict__ samples, const unsigned short* __restrict__ order,
                  const double2* __restrict__ dirs, int n, int H, int nz, long long fzs, int zchunk, float a,
                  float C0, int xg) {
  using PT = typename std::conditional<ZP == 1, float2, float4>::type;
  constexpr int NT = S::NT, SPT = S::SPT, STG = S::STG, PE = S::PE;
  __shared__ __align__(16) PT P[PE + 2];  // + a sink for idle staging slots
  __shared__ PT O[SMAX];
  const unsigned lw = xcd_group(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y, xg);
  const Item it = items[lw % gridDim.x];
  const int z0 = (int)(lw / gridDim.x) * zchunk, z1 = min(nz, z0 + zchunk);
  const int pp = patch_pitch(it.w);  // LDS row pitch >= w + 2M
  // staging descriptors: source offset in the z plane; LDS slot << 1 | conj.
  // Idle slots read offset 0 and write the sink, so the loads issue together.
  unsigned soff[STG];
  int sdst[STG];
#pragma unroll
  for (int e = 0; e < STG; ++e) gather_stage_desc(it, threadIdx.x + e * NT, n, PE, soff[e], sdst[e]);
  // Lane threadIdx.x computes samples s_j = threadIdx.x + j NT (j < SPT).
  // ZP = 1: ex[j] weighs patch column (base & ~1) + j, the 9 taps shifted by
  // the parity of base (ex[9] = 0 or ex[0] = 0).  ZP = 2: ex[j] weighs base + j.
  constexpr int NX = ZP == 1 ? TAPS + 1 : TAPS;
  float ex[SPT][NX], ey[SPT][TAPS];
  int base[SPT];
  // output of sorted sample t_j = threadIdx.x + j NT
  int src_lane[SPT];
  long long out0[SPT];
  // The sample reads of all SPT samples go out together, in two dependent rounds (ids and lane order; then
  // directions and the ids in lane order), from clamped indices: one sample at a time behind the
  // s < count predicate they were issued one by one, each waited on.
  int idj[SPT], idl[SPT];
  double2 dj[SPT];
  if (it.count > 0) {
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
      const int sc = it.start + min((int)threadIdx.x + j * NT, it.count - 1);
      idj[j] = samples[sc];
      src_lane[j] = order[sc];
    }
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
      dj[j] = dirs[idj[j] / H];
      idl[j] = samples[it.start + src_lane[j]];
    }
  }
#pragma unroll
  for (int j = 0; j < SPT; ++j) {
    const int s = threadIdx.x + j * NT;
    base[j] = 0;
    out0[j] = 0;
    if (s >= it.count) {
      src_lane[j] = 0;
      continue;
    }
    const SampleGeom g = sample_geom_d(idj[j], H, dj[j]);
    const int b = (g.ry - it.y0) * pp + (g.rx - it.x0);
    const int par = ZP == 1 ? (b & 1) : 0;
    base[j] = b - par;
    // centring sign of tap (i0, i1): (-1)^(rx - M + i0) (-1)^(ry - M + i1)
    const float sx = ((g.rx - M) & 1) ? -1.f : 1.f, sy = ((g.ry - M) & 1) ? -C0 : C0;
#pragma unroll
    for (int q = 0; q < NX; ++q) {
      const int i = q - par;
      const float dx = (float)(i - M) - g.fx;
      const float sg = (i & 1) ? -sx : sx;
      ex[j][q] = (i >= 0 && i < TAPS) ? sg * __expf(-a * dx * dx) : 0.f;
    }
#pragma unroll
    for (int i = 0; i < TAPS; ++i) {
      const float dy = (float)(i - M) - g.fy;
      ey[j][i] = ((i & 1) ? -sy : sy) * __expf(-a * dy * dy);
    }
    const int id = idl[j];
    const int th = id / H;
    out0[j] = (long long)th * nz * H + (id - th * H);
  }
  float2 sa[STG], sb[STG];
  auto fetch = [&](int z) {
    const float2* f0 = fde + (long long)z * fzs;
#pragma unroll
    for (int e = 0; e < STG; ++e) sa[e] = f0[soff[e]];
    if constexpr (ZP == 2) {
      const bool two = z + 1 < z1;
      const float2* f1 = two ? f0 + fzs : f0;
#pragma unroll
      for (int e = 0; e < STG; ++e) sb[e] = f1[soff[e]];
      if (!two)
#pragma unroll
        for (int e = 0; e < STG; ++e) sb[e] = make_float2(0.f, 0.f);
    }
  };
  // ZP = 2 prefetches the next round during the taps; ZP = 1 (11 staging slots
  // per lane in the root class) has no registers to spare and fetches at the
  // top.
  if (ZP == 2 && z0 < z1) fetch(z0);
  for (int z = z0; z < z1; z += ZP) {
    if constexpr (ZP == 1) fetch(z);
#pragma unroll
    for (int e = 0; e < STG; ++e) {
      float2 u = sa[e];
      if (sdst[e] & 1) u.y = -u.y;
      if constexpr (ZP == 1) {
        P[sdst[e] >> 1] = u;
      } else {
        float2 v = sb[e];
        if (sdst[e] & 1) v.y = -v.y;
        P[sdst[e] >> 1] = make_float4(u.x, u.y, v.x, v.y);
      }
    }
    __syncthreads();
    if (ZP == 2 && z + ZP < z1) fetch(z + ZP);
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
      if (threadIdx.x + j * NT >= it.count) continue;
      if constexpr (ZP == 1) {
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int i1 = 0; i1 < TAPS; ++i1) {
          const float4* prow = reinterpret_cast<const float4*>(P + base[j] + i1 * pp);  // base, pp even
          float2 row = make_float2(0.f, 0.f);
#pragma unroll
          for (int c = 0; c < (TAPS + 1) / 2; ++c) {
            const float4 v = prow[c];
            row.x = fmaf(ex[j][2 * c], v.x, row.x);
            row.y = fmaf(ex[j][2 * c], v.y, row.y);
            row.x = fmaf(ex[j][2 * c + 1], v.z, row.x);
            row.y = fmaf(ex[j][2 * c + 1], v.w, row.y);
          }
          acc.x = fmaf(ey[j][i1], row.x, acc.x);
          acc.y = fmaf(ey[j][i1], row.y, acc.y);
        }
        O[threadIdx.x + j * NT] = acc;
      } else {
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int i1 = 0; i1 < TAPS; ++i1) {
          const float4* prow = P + base[j] + i1 * pp;
          float4 row = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
          for (int i0 = 0; i0 < TAPS; ++i0) {
            const float4 v = prow[i0];
            row.x = fmaf(ex[j][i0], v.x, row.x);
            row.y = fmaf(ex[j][i0], v.y, row.y);
            row.z = fmaf(ex[j][i0], v.z, row.z);
            row.w = fmaf(ex[j][i0], v.w, row.w);
          }
          acc.x = fmaf(ey[j][i1], row.x, acc.x);
          acc.y = fmaf(ey[j][i1], row.y, acc.y);
          acc.z = fmaf(ey[j][i1], row.z, acc.z);
          acc.w = fmaf(ey[j][i1], row.w, acc.w);
        }
        O[threadIdx.x + j * NT] = acc;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
      if (threadIdx.x + j * NT >= it.count) continue;
      const PT o = O[src_lane[j]];
      if constexpr (ZP == 1) {
        gh[out0[j] + (long long)z * H] = o;
      } else {
        gh[out0[j] + (long long)z * H] = make_float2(o.x, o.y);
        if (z + 1 < z1) gh[out0[j] + (long long)(z + 1) * H] = make_float2(o.z, o.w);
      }
    }
  }
}

// Line pass of R: rebuild the full line g[r] from its representatives,
// sino = c1d . IFFT(c1d . g) . real / (4n)   (c1d[r] = (-1)^(r+1)).  Inputs and
// outputs go straight between HBM and the register FFT.
//
// Two lines (angles theta, theta + 1 at one z) ride in one complex transform:
// the real part of IFFT(a) is IFFT(Herm a), Herm a[r] = (a[r] + conj a[L - r]) / 2,
// and a rebuilt line is Hermitian already except at r = 0 and r = L/2 (the
// k = -n/2 and k = 0 samples), where Herm takes the real part.  So
// IFFT(Herm a + i Herm b) has line a's result as its real part and line b's as
// its imaginary part.
template <int L>
__global__ __launch_bounds__(LineShape<L>::NT) void r4_line_ifft(const float2* __restrict__ gh, float* __restrict__ sino,
                                                           const int* __restrict__ sig, int H, long long nrows, int nz,
                                                           float scale, const float2* __restrict__ tw) {
  using Sh = LineShape<L>;
  __shared__ float2 lds[Sh::LDS];
  int c, t;
  Sh::map(threadIdx.x, c, t);
  // rows (theta, z) and (theta + 1, z), theta even: the pairing does not depend on how z is batched
  const long long q = (long long)blockIdx.x * Sh::C + c, th2 = q / nz;
  const long long ra = th2 * 2 * nz + (q - th2 * nz), rb = ra + nz;
  const bool va = ra < nrows, vb = rb < nrows;
  // reads from in-bounds rows (row 0 for a missing line), predicates applied afterwards: every read of
  // the workgroup is in flight at once (the sign read and per-lane predicated reads were waited on first)
  const int sga0 = sig[(va ? ra : 0) / nz], sgb0 = sig[(vb ? rb : 0) / nz];
  const int sga = va ? sga0 : 1, sgb = vb ? sgb0 : 1;
  const float2* srca = gh + (va ? ra : 0) * H;
  const float2* srcb = gh + (vb ? rb : 0) * H;
  constexpr int half = L / 2;
  // line value at r from the representatives (k = r - L/2: |k|, conjugated by the sign of the angle)
  auto rebuild = [&](const float2* src, int sg, bool valid, int r) {
    const int d = r - half;
    const int kk = d >= 0 ? d : -d;
    const bool cj = sg > 0 ? d < 0 : d > 0;
    float2 w = src[kk];
    if (!valid) w = make_float2(0.f, 0.f);
    if (cj) w.y = -w.y;
    if (r == 0 || r == half) w.y = 0.f;  // Hermitian part
    return w;
  };
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int r = t + m * Sh::TPC;
    const float2 a = rebuild(srca, sga, va, r), b = rebuild(srcb, sgb, vb, r);
    float2 w = make_float2(a.x - b.y, a.y + b.x);  // a + i b
    if (!(r & 1)) w = make_float2(-w.x, -w.y);
    v[m] = w;
  }
  fft_reg<L, Sh::TPC, true>(v, lds + c * Sh::S, t, tw);
  float* dsta = sino + (va ? ra : 0) * L;
  float* dstb = sino + (vb ? rb : 0) * L;
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int s = t + m * Sh::TPC;
    const float sc = (s & 1) ? scale : -scale;
    if (va) dsta[s] = v[m].x * sc;
    if (vb) dstb[s] = v[m].y * sc;
  }
}

// Line pass of RT: yhat = c1d . FFT(c1d . y), written at the representatives
// with the fold weights (1/2 at k = 0 and k = n/2) times `scale`.  A complex
// sinogram yields the two channels of its real and imaginary parts; a real one
// puts two rows (angles theta, theta + 1 at one z) into one transform as its
// real and imaginary parts.  Either way the split pairs yhat(r) with
// yhat(L - r), so the output passes through LDS.
// Representatives of one real line from its own transform (used behind the FBP
// filter; thread t holds r = t + m TPC):
// yhat(r) = c1d(r) v, k = r - L/2 (sigma > 0; k = L/2 from r = 0, conjugated) or
// k = L/2 - r (sigma < 0), fold weights 1/2 at k = 0 and k = L/2.
template <class Sh>
__device__ __forceinline__ void rt1_emit_real(const float2 (&v)[Sh::PT], int t, int sg, float2* __restrict__ dst,
                                              float scale) {
  constexpr int half = Sh::PT * Sh::TPC / 2;
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int r = t + m * Sh::TPC;
    const float2 y = (r & 1) ? v[m] : make_float2(-v[m].x, -v[m].y);  // yhat(r)
    if (sg > 0) {
      if (r >= half) {
        const int k = r - half;
        dst[k] = cscale(y, (k == 0 ? 0.5f : 1.0f) * scale);
      } else if (r == 0) {
        dst[half] = cscale(cconj(y), 0.5f * scale);
      }
    } else if (r <= half) {
      const int k = half - r;
      dst[k] = cscale(y, ((k == 0 || k == half) ? 0.5f : 1.0f) * scale);
    }
  }
}

template <int L, bool CPLX>
__global__ __launch_bounds__(LineShape<L>::NT) void rt1_line_fft(const void* __restrict__ sino_v, float2* __restrict__ yre,
                                                           float2* __restrict__ yim, const int* __restrict__ sig, int H,
                                                           long long nrows, int nz, float scale,
                                                           const float2* __restrict__ tw) {
  using Sh = LineShape<L>;
  __shared__ float2 lds[Sh::LDS];
  int c, t;
  Sh::map(threadIdx.x, c, t);
  float2* x = lds + c * Sh::S;
  constexpr int half = L / 2;
  if constexpr (!CPLX) {
    // two real rows per transform, (theta, z) and (theta + 1, z) with theta even (the pairing does
    // not depend on how z is batched), as its real and imaginary parts; each row's spectrum is
    // (V[r] +- conj V[L - r]) / 2 (times -i for the imaginary one), V[0] splits into its real and
    // imaginary part
    const long long q = (long long)blockIdx.x * Sh::C + c, th2 = q / nz;
    const long long ra = th2 * 2 * nz + (q - th2 * nz), rb = ra + nz;
    const bool va = ra < nrows, vb = rb < nrows;
    const int sga0 = sig[(va ? ra : 0) / nz], sgb0 = sig[(vb ? rb : 0) / nz];
    const int sga = va ? sga0 : 1, sgb = vb ? sgb0 : 1;
    const float* pa = (const float*)sino_v + (va ? ra : 0) * L;
    const float* pb = (const float*)sino_v + (vb ? rb : 0) * L;
    float2 v[Sh::PT];
    // reads from in-bounds rows (row 0 for a missing line), zeroed afterwards: all in flight at once
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) {
      const int r = t + m * Sh::TPC;
      float2 w = make_float2(pa[r], pb[r]);
      if (!va) w.x = 0.f;
      if (!vb) w.y = 0.f;
      if (!(r & 1)) w = make_float2(-w.x, -w.y);
      v[m] = w;
    }
    fft_reg<L, Sh::TPC, false>(v, x, t, tw);
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) x[lpad(t + m * Sh::TPC)] = v[m];
    __syncthreads();
    auto yhat = [&](int r) {
      float2 w = x[lpad(r)];
      return (r & 1) ? w : make_float2(-w.x, -w.y);
    };
    float2* da = yre + (va ? ra : 0) * H;
    float2* db = yre + (vb ? rb : 0) * H;
    for (int k = t; k < H; k += Sh::TPC) {
      const float f = ((k == 0 || k == half) ? 0.5f : 1.0f) * scale;
      if (k == half) {  // from r = 0 for either sign
        const float2 v0 = yhat(0);
        if (va) da[k] = make_float2(v0.x * f, 0.f);
        if (vb) db[k] = make_float2(v0.y * f, 0.f);
        continue;
      }
      if (va) {
        const int r = sga > 0 ? half + k : half - k;
        const float2 p = yhat(r), pm = yhat(L - r);
        da[k] = make_float2(0.5f * (p.x + pm.x) * f, 0.5f * (p.y - pm.y) * f);
      }
      if (vb) {
        const int r = sgb > 0 ? half + k : half - k;
        const float2 p = yhat(r), pm = yhat(L - r);
        db[k] = make_float2(0.5f * (p.y + pm.y) * f, -0.5f * (p.x - pm.x) * f);
      }
    }
    return;
  }
  const long long row = (long long)blockIdx.x * Sh::C + c;
  const bool valid = row < nrows;
  const long long rr = valid ? row : 0;
  const int sg = valid ? sig[rr / nz] : 1;
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int r = t + m * Sh::TPC;
    float2 w = valid ? ((const float2*)sino_v)[rr * L + r] : make_float2(0.f, 0.f);
    if (!(r & 1)) w = make_float2(-w.x, -w.y);
    v[m] = w;
  }
  fft_reg<L, Sh::TPC, false>(v, x, t, tw);
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) x[lpad(t + m * Sh::TPC)] = v[m];
  __syncthreads();
  if (!valid) return;  // no barrier below
  auto yhat = [&](int r) {
    float2 w = x[lpad(r)];
    return (r & 1) ? w : make_float2(-w.x, -w.y);
  };
  for (int k = t; k < H; k += Sh::TPC) {
    const float f = ((k == 0 || k == half) ? 0.5f : 1.0f) * scale;
    float2 vr, vi;
    if (k == half) {
      const float2 v0 = yhat(0);
      if (sg > 0) {
        vr = cconj(v0);
        vi = make_float2(v0.y, v0.x);  // conj(-i v0)
      } else {
        vr = v0;
        vi = make_float2(v0.y, -v0.x);  // -i v0
      }
    } else {
      const int r = sg > 0 ? half + k : half - k;
      const float2 p = yhat(r), q = yhat(r == 0 ? 0 : L - r);
      vr = make_float2(0.5f * (p.x + q.x), 0.5f * (p.y - q.y));
      vi = make_float2(0.5f * (p.y + q.y), -0.5f * (p.x - q.x));
    }
    yre[rr * H + k] = cscale(vr, f);
    yim[rr * H + k] = cscale(vi, f);
  }
}

// Line pass of RT behind the FBP filter (processing/fbp_filter.py:128-210, then
// tomo.py:160-170), for real sinograms: each row is filtered in the length-2n
// form of the edge-padded filter (G, A, B: see mm_fbp_filter_2n; two rows per
// complex transform of length 2L), handed through LDS to the line transform of
// length L, and its representatives are written as by rt1_line_fft.  The
// filtered sinogram never goes to HBM and the input is not modified.  Both
// stages use the same threads: C2 filter transforms = C1 = 2 C2 lines.
template <int L>
struct FbpLineShape {  // C2 filter transforms of length 2L and C1 = 2 C2 lines of length L per workgroup
  static constexpr int C2 = kRowBatch / (2 * L) > 0 ? kRowBatch / (2 * L) : 1;
  using S1 = RegShape<L, C2 * 2 * L, 0, false>;
  using S2 = RegShape<2 * L, C2 * 2 * L, 0, false>;
};

template <int L>
__global__ MM_TWO_FFT_BUDGET __launch_bounds__(FbpLineShape<L>::S1::NT) void rt1_fbp_line_fft(const float* __restrict__ sino,
                                                                  float2* __restrict__ yre,
                                                                  const int* __restrict__ sig, int H,
                                                                  long long nrows, int nz, float scale,
                                                                  const float* __restrict__ G,
                                                                  const float* __restrict__ A,
                                                                  const float* __restrict__ B,
                                                                  const float2* __restrict__ tw1,
                                                                  const float2* __restrict__ tw2) {
  using S1 = typename FbpLineShape<L>::S1;
  using S2 = typename FbpLineShape<L>::S2;
  static_assert(S1::NT == S2::NT && S1::C == 2 * S2::C && S1::PT == S2::PT, "one thread layout for both stages");
  static_assert(S2::PT % 2 == 0, "the row is points m < PT/2 of the filter transform");
  constexpr int LDSN = S1::LDS > S2::LDS ? S1::LDS : S2::LDS;
  constexpr int MH = S2::PT / 2;
  __shared__ float2 lds[LDSN];
  const long long row0 = (long long)blockIdx.x * S1::C;
  // stage 1: filter rows 2 c2 and 2 c2 + 1 of the workgroup (real / imaginary part)
  int c2, t2;
  S2::map(threadIdx.x, c2, t2);
  const long long ra = row0 + 2 * c2, rb = ra + 1;
  const bool has_a = ra < nrows, has_b = rb < nrows;
  const float* pa = sino + (has_a ? ra : 0) * L;
  const float* pb = sino + (has_b ? rb : 0) * L;
  const float a0 = has_a ? pa[0] : 0.f, a1 = has_a ? pa[L - 1] : 0.f;
  const float b0 = has_b ? pb[0] : 0.f, b1 = has_b ? pb[L - 1] : 0.f;
  float2 v[S2::PT];
#pragma unroll
  for (int m = 0; m < S2::PT; ++m) {
    const int i = t2 + m * S2::TPC;
    v[m] = m < MH ? make_float2(has_a ? pa[i] : 0.f, has_b ? pb[i] : 0.f) : make_float2(0.f, 0.f);
  }
  fft_reg<2 * L, S2::TPC, false, 1, false>(v, lds + c2 * S2::S, t2, tw2);
#pragma unroll
  for (int m = 0; m < S2::PT; ++m) {
    const int i = t2 + m * S2::TPC;
    v[m] = cscale(v[m], G[i <= L ? i : 2 * L - i]);
  }
  fft_reg<2 * L, S2::TPC, true, 1, false>(v, lds + c2 * S2::S, t2, tw2);
  __syncthreads();  // the filter's last exchange is read out
#pragma unroll
  for (int m = 0; m < MH; ++m) {
    const int i = t2 + m * S2::TPC;
    const float ai = A[i], bi = B[i];
    lds[(2 * c2) * S1::S + lpad(i)] = make_float2(fmaf(a1, bi, fmaf(a0, ai, v[m].x)), 0.f);
    lds[(2 * c2 + 1) * S1::S + lpad(i)] = make_float2(fmaf(b1, bi, fmaf(b0, ai, v[m].y)), 0.f);
  }
  __syncthreads();
  // stage 2: line c1 = row row0 + c1, as rt1_line_fft
  int c1, t1;
  S1::map(threadIdx.x, c1, t1);
  const long long row = row0 + c1;
  const bool valid = row < nrows;
  const int sg = valid ? sig[row / nz] : 1;
  float2 w[S1::PT];
#pragma unroll
  for (int m = 0; m < S1::PT; ++m) {
    const int r = t1 + m * S1::TPC;
    const float2 u = lds[c1 * S1::S + lpad(r)];
    w[m] = (r & 1) ? u : make_float2(-u.x, -u.y);
  }
  __syncthreads();  // the transform's exchanges reuse the staging area
  fft_reg<L, S1::TPC, false, 1, false>(w, lds + c1 * S1::S, t1, tw1);
  if (valid) rt1_emit_real<S1>(w, t1, sg, yre + row * H, scale);
}

// ---- adjoint scatter: per-output-point contribution lists (owner computes) --
// The folded scatter of a representative sample s is rewritten as the direct
// scatter of two "entries": s itself and its mirror (position -X, -Y, value
// conj v).  A tap contributes to the half plane iff its raw column
// (l0 mod 2n) lies in [0, n]; for the mirror entry this selects exactly the
// folded taps of k0 in {0} U [n, 2n).  At plan time every 16 x 16 output tile
// gets the list of entries touching it and, for each of its points, the
// (entry, i0, i1) taps landing there.  A workgroup owns one tile: thread p
// keeps the list of point p for the whole z loop, the tile's entry
// values are staged SZC slices at a time, and each thread accumulates its own
// point in registers -- no atomics, fixed summation order (bitwise
// deterministic), one write per output point.  Tiles whose lists exceed the
// per-item budget (around the origin) are split over entry chunks into
// partial tiles that scatter_reduce sums.
//
// A point's list is private to its thread, so it lives in registers: two
// 16-bit slots per VGPR, in a 32 + 16 element vector indexed by the (uniform)
// loop counter, which the compiler lowers to relative register addressing.
// Each entry has one 80-byte LDS row: its SZC staged values, its offset F and
// its (theta, k | mirror) key.  A tap reads the row as five 16-byte pieces
// (ds_read_b128; ds_read2_b64 moves half the bytes per cycle), which the LDS
// serves 16 lanes at a time; the plan orders every 16-lane group's slots so
// that the rows read in one step differ mod 16 or coincide (schedule_group),
// keeping the reads free of bank conflicts.  Each wave runs to the longest
// schedule among its four groups (wnp); idle slots read a row another lane
// reads with tap code SL_IDLE_TAP, whose weight is zero.  Four work items
// share a CU (4 x 39.7 KB of LDS).
constexpr int STS = 16;          // output tile edge
constexpr int SNP = STS * STS;   // points = threads per work item
constexpr int SWAVES = SNP / 64; // waves per work item
constexpr int SGROUP = 16;       // lanes per LDS cycle of a 16-byte read
#ifndef MM_SCATTER_SE_MAX
#define MM_SCATTER_SE_MAX 496
#endif
#ifndef MM_SCATTER_SZC
#define MM_SCATTER_SZC 8
#endif
#ifndef MM_SCATTER_WAVES
#define MM_SCATTER_WAVES 4
#endif
#ifndef MM_SCATTER_RELPOS
#define MM_SCATTER_RELPOS 0
#endif
constexpr int SE_MAX = MM_SCATTER_SE_MAX;  // entry rows in LDS per work item
constexpr int SL_MAX = 96;       // list slots per point per work item (48 VGPRs; fewer split tiles than 80)
constexpr int SL_CHUNK = SL_MAX - 4;  // taps per point per work item (room for idle steps)
constexpr int SL_PAIRS_A = 32, SL_PAIRS_B = SL_MAX / 2 - SL_PAIRS_A;
constexpr int SZC = MM_SCATTER_SZC;  // z slices per staging round
constexpr unsigned SL_IDLE_TAP = 127;  // slot = row | (9 * i1 + i0) << 9; tap code 127: weight 0
static_assert(SE_MAX <= 512 && SL_MAX % 2 == 0 && SL_PAIRS_B == 16 && SWAVES <= 4 && SL_MAX / 2 < 256,
              "slot encoding");

struct __align__(16) SRow {
  float2 v[SZC];  // staged values, z = zc + 0 .. SZC - 1
  float2 f;       // fractional offset (mirror: negated)
  int2 tk;        // theta, k | mirror << 30
};
static_assert(sizeof(SRow) == 16 * (SZC / 2 + 1), "SZC / 2 + 1 16-byte pieces per row");

struct SItem {
  long long loff;  // pairs of slots: lists[loff + (j / 2) * 2 SNP + 2 p + (j & 1)], j < nl
  int eoff, ne;    // entries[eoff, eoff + ne): (sample id << 1) | mirror, ne <= SE_MAX
  int x0, y0;      // tile origin (l0, l1), centred coordinates
  int nl, slot;    // list slots (even); -1: the item owns the tile, >= 0: partial slot
  unsigned wnp;    // slot pairs wave w runs (<= nl / 2): bits [8 w, 8 w + 8)
  int pad_;
};

struct SplitTile {
  int x0, y0, slot0, nslots;
};

// Tile point (row-major, 16 x 16) of index p in 4 x 4 patch order.
__host__ __device__ __forceinline__ int patch_major_point(int p) {
  const int P = p >> 4, q = p & 15;
  return (4 * (P >> 2) + (q >> 2)) * STS + 4 * (P & 3) + (q & 3);
}

typedef unsigned u32x32 __attribute__((ext_vector_type(SL_PAIRS_A)));
typedef unsigned u32x16 __attribute__((ext_vector_type(SL_PAIRS_B)));  // power-of-two sizes stay in VGPRs

// Compensated (Kahan) FP32 accumulator for the scatter tiles (MM_RADON_KAHAN_TILES): the running sum s and
// the negated low-order part c lost by its last addition; the value is s - c.
struct KahanAcc {
  float s = 0.f, c = 0.f;
  __device__ KahanAcc() = default;
  __device__ explicit KahanAcc(float v) : s(v) {}
  __device__ __forceinline__ KahanAcc& operator+=(KahanAcc x) {
    const float y = x.s - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
    return *this;
  }
  __device__ explicit operator float() const { return s - c; }
};

template <typename Acc>
__global__ __launch_bounds__(SNP) __attribute__((amdgpu_waves_per_eu(MM_SCATTER_WAVES, MM_SCATTER_WAVES))) void scatter_lists(
    const float2* __restrict__ yh, float2* __restrict__ Hb, float2* __restrict__ partial,
    const SItem* __restrict__ items, const unsigned char* __restrict__ perm, const int* __restrict__ entries,
    const unsigned short* __restrict__ lists, const double2* __restrict__ dirs, int n, int H, int nz, long long hzs,
    int zchunk, float a2, int xg) {
  __shared__ SRow R[SE_MAX];
#if !MM_SCATTER_RELPOS
  __shared__ float2 TT[128];  // tap code t = 9 i1 + i0 -> (i0 - M, i1 - M); idle code: far away (weight 0)
#endif
  const unsigned lw = xcd_group(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y, xg);
  const unsigned item = lw % gridDim.x;
  const SItem it = items[item];
#if !MM_SCATTER_RELPOS
  if (threadIdx.x < 128) {
    const int t = threadIdx.x;
    TT[t] = t < TAPS * TAPS ? make_float2((float)(t % TAPS - M), (float)(t / TAPS - M)) : make_float2(1e3f, 1e3f);
  }
#endif
  const int z0 = (int)(lw / gridDim.x) * zchunk, z1 = min(nz, z0 + zchunk);
  const int p = threadIdx.x;                                 // lane: list lp[.. p ..]
  const int pt = perm[(long long)item * SNP + p];           // the tile point it owns
  const int l0 = it.x0 + (pt % STS), l1 = it.y0 + (pt / STS);
#if MM_SCATTER_RELPOS
  // entry position relative to the tile centre (x0 + 7.5, y0 + 7.5), wrapped onto this tile: the
  // tap distance is then (point - centre) - position, with no tap code involved
  const float pcx = (float)(pt % STS) - 7.5f, pcy = (float)(pt / STS) - 7.5f;
#endif
  for (int e = p; e < it.ne; e += SNP) {
    const int code = entries[it.eoff + e];
    const SampleGeom g = sample_geom(code >> 1, H, dirs);
    const bool mir = (code & 1) != 0;
#if MM_SCATTER_RELPOS
    const int cx = mir ? -g.rx : g.rx, cy = mir ? -g.ry : g.ry;
    const int twon = 2 * n;
    const int ix = ((cx - it.x0 + M) % twon + twon) % twon - M, iy = ((cy - it.y0 + M) % twon + twon) % twon - M;
    const double2 d = dirs[g.th];
    const double X = (2.0 * g.k) * d.x, Y = (2.0 * g.k) * d.y;
    const double fx = X - (double)g.rx, fy = Y - (double)g.ry;
    R[e].f = make_float2((float)((double)ix - 7.5 + (mir ? -fx : fx)), (float)((double)iy - 7.5 + (mir ? -fy : fy)));
#else
    R[e].f = mir ? make_float2(-g.fx, -g.fy) : make_float2(g.fx, g.fy);
#endif
    R[e].tk = make_int2(g.th, g.k | (mir ? (1 << 30) : 0));
  }
  const int npairs = __builtin_amdgcn_readfirstlane((int)((it.wnp >> (8 * (p >> 6))) & 0xFFu));
  const int na = min(npairs, SL_PAIRS_A);
  const unsigned* lp = reinterpret_cast<const unsigned*>(lists + it.loff) + p;
  u32x32 la;
  u32x16 lb;
#pragma unroll
  for (int jp = 0; jp < SL_PAIRS_A; ++jp) la[jp] = jp < npairs ? lp[jp * SNP] : 0u;
#pragma unroll
  for (int jp = 0; jp < SL_PAIRS_B; ++jp) lb[jp] = jp + SL_PAIRS_A < npairs ? lp[(jp + SL_PAIRS_A) * SNP] : 0u;
  __syncthreads();
  const float sgn = ((l0 + l1) & 1) ? -1.f : 1.f;
  for (int zc = z0; zc < z1; zc += SZC) {
    const int nzz = min(SZC, z1 - zc);
    for (int e = p; e < it.ne; e += SNP) {
      const int2 tk = R[e].tk;
      const float2* src = yh + ((long long)tk.x * nz + zc) * H + (tk.y & 0x3FFFFFFF);
      const bool cj = (tk.y >> 30) != 0;
      float2 v[SZC];
#pragma unroll
      for (int zz = 0; zz < SZC; ++zz) {
        v[zz] = zz < nzz ? src[(long long)zz * H] : make_float2(0.f, 0.f);
        if (cj) v[zz].y = -v[zz].y;
      }
      float4* dst = reinterpret_cast<float4*>(R[e].v);
#pragma unroll
      for (int q = 0; q < SZC / 2; ++q) dst[q] = make_float4(v[2 * q].x, v[2 * q].y, v[2 * q + 1].x, v[2 * q + 1].y);
    }
    __syncthreads();
    Acc ax[SZC], ay[SZC];
#pragma unroll
    for (int zz = 0; zz < SZC; ++zz) {
      ax[zz] = (Acc)0;
      ay[zz] = (Acc)0;
    }
    auto tap = [&](unsigned u) {
      const int e = (int)(u & 0x1FFu), t = (int)(u >> 9);
      const float4* row = reinterpret_cast<const float4*>(&R[e]);
      // f is read with the value pieces; the tap's integer offsets come from TT (no
      // division by 9, no conversions), and an idle slot's offsets are so far off that
      // its weight underflows to exactly 0 (no select); a2 = -a log2(e)
#if MM_SCATTER_RELPOS
      const float2 ft = *reinterpret_cast<const float2*>(&row[SZC / 2]);  // position relative to the tile centre
      const float dx = pcx - ft.x, dy = pcy - ft.y;
      float w = __builtin_amdgcn_exp2f(a2 * (dx * dx + dy * dy));
      if (t == (int)SL_IDLE_TAP) w = 0.f;  // an idle step reads another lane's row
#else
      const float4 ft = row[SZC / 2];  // f, tk
      const float2 tb = TT[t];
      const float dx = tb.x - ft.x, dy = tb.y - ft.y;
      const float w = __builtin_amdgcn_exp2f(a2 * (dx * dx + dy * dy));
#endif
#pragma unroll
      for (int q = 0; q < SZC / 2; ++q) {
        const float4 v = row[q];
        ax[2 * q] += (Acc)(w * v.x);
        ay[2 * q] += (Acc)(w * v.y);
        ax[2 * q + 1] += (Acc)(w * v.z);
        ay[2 * q + 1] += (Acc)(w * v.w);
      }
    };
    for (int jp = 0; jp < na; ++jp) {
      const unsigned u = la[jp];
      tap(u & 0xFFFFu);
      tap(u >> 16);
    }
    for (int jp = SL_PAIRS_A; jp < npairs; ++jp) {
      const unsigned u = lb[jp - SL_PAIRS_A];
      tap(u & 0xFFFFu);
      tap(u >> 16);
    }
    // Results leave through LDS: lane p owns point pt (lanes are ordered by list length, so a
    // wave's points are scattered over the tile); staged in the value slots of the entry rows
    // (index zz * SNP + point -> row / SZC, slice % SZC), then written in tile order -- a wave
    // covers 16 rows x 4 columns of one 4-column block, i.e. one contiguous 512-byte run of the
    // half plane per slice (partial tiles: 64 consecutive points).
    __syncthreads();  // every lane is done reading the entry values
    static_assert(SZC * SNP <= SE_MAX * SZC, "result staging fits the value slots");
#pragma unroll
    for (int zz = 0; zz < SZC; ++zz) {
      const int i = zz * SNP + pt;
      R[i / SZC].v[i % SZC] = make_float2(sgn * (float)ax[zz], sgn * (float)ay[zz]);
    }
    __syncthreads();
    if (it.slot < 0) {
      const int blk = p >> 6, r = (p >> 2) & 15, c4 = p & 3;
      const int q = r * STS + 4 * blk + c4;  // tile point this lane writes
      const int col = it.x0 + 4 * blk + c4, row = it.y0 + r;
      if (col <= n) {
        const long long o = hs_idx(row < 0 ? row + 2 * n : row, col, 2 * n);
        for (int zz = 0; zz < nzz; ++zz) {
          const int i = zz * SNP + q;
          Hb[(long long)(zc + zz) * hzs + o] = R[i / SZC].v[i % SZC];
        }
      }
    } else {
      const int q = patch_major_point(p);  // partial tiles are stored in 4 x 4 patch order
      for (int zz = 0; zz < nzz; ++zz) {
        const int i = zz * SNP + q;
        partial[((long long)it.slot * nz + zc + zz) * SNP + p] = R[i / SZC].v[i % SZC];
      }
    }
    __syncthreads();
  }
}

// Sum of the partial tiles of split tiles (stored in 4 x 4 patch order: index
// p = 16 patch + point, patch = 4 patch row + patch column, point = 4 row + column).
__global__ __launch_bounds__(SNP) void scatter_reduce(const float2* __restrict__ partial, float2* __restrict__ Hb,
                                                      const SplitTile* __restrict__ splits, int n, int nz,
                                                      long long hzs) {
  const SplitTile st = splits[blockIdx.x];
  const int z = blockIdx.y;
  const int p = threadIdx.x;
  const int q = patch_major_point(p);
  const int l0 = st.x0 + (q % STS), l1 = st.y0 + (q / STS);
  if (l0 > n) return;
  float2 acc = make_float2(0.f, 0.f);
#pragma unroll 8  // eight loads in flight; the summation order stays slot order
  for (int sl = st.slot0; sl < st.slot0 + st.nslots; ++sl) {
    const float2 v = partial[((long long)sl * nz + z) * SNP + p];
    acc.x += v.x;
    acc.y += v.y;
  }
  Hb[(long long)z * hzs + hs_idx(l1 < 0 ? l1 + 2 * n : l1, l0, 2 * n)] = acc;
}

// ---- adjoint scatter on the matrix cores (the default FP32 path) ------------
// The 16 x 16 output tile is sixteen 4 x 4 patches, aligned to the 4-column
// blocks of the half plane (a patch of one slice is 16 consecutive complex
// values).  An entry's 9 x 9 window covers exactly 3 x 3 patches (9 = 2 * 4 + 1)
// wherever it sits, so the scatter into one patch is the dense product
//   C[zc][q] = sum_e V[zc][e] W[e][q],   zc = 2 slice + (re | im), q = patch point,
// over the entries e touching the patch, with W[e][q] the Gaussian weight of
// the tap q - (window base of e) and 0 off the window: 144 (entry, point)
// pairs per entry for its 81 taps (1.78x zero fill).  One
// v_mfma_f32_16x16x4_f32 takes 4 entries x 16 zc x 16 points; a staging round
// holds MZ = 16 slices (two 16-row zc blocks of 8 slices, real and imaginary
// parts interleaved), so each weight is evaluated once per 16 slices and each
// staged value is read once per patch.  A workgroup of
// 8 waves owns the tile, two patches per wave.
// At plan time the item's entries are sorted by window anchor class
// (floor(bx / 4) + 2, floor(by / 4) + 2) in [0, 6)^2, bx, by = window base
// relative to the tile: the entries touching patch (px, py) are then three
// contiguous runs (anchor rows py .. py + 2, anchor columns px .. px + 2 each),
// so no per-point list exists.  The fold needs no test: a tap counts iff its
// raw column is <= n, i.e. iff the output point exists; points past column n
// are computed and not written.
// Lane l of a wave: k = l / 16 picks the chunk's entry, q = l % 16 the patch
// point (B operand W[k][q]) and the zc row (A operand V[zc = q][k]); the result
// C[zc][q] lands in register r of lane l with zc = 4 (l / 16) + r, q = l % 16.
// Accumulation order is fixed by the plan: RT stays bitwise deterministic.
constexpr int MZ = 16;            // z slices per staging round (two 16-row zc blocks)
constexpr int MZB = MZ / 8;       // zc blocks of 16
constexpr int MROW = 2 * MZ + 4;  // floats per LDS entry row: values, then bx + M, by + M, fx, fy
// entries per work item: 52 KB of LDS and at most 80 registers hold three workgroups per CU (480 entries,
// 78 KB and two per CU, measured 2 - 3 % slower: fewer tiles split, but less overlap of one workgroup's
// staging and stores with another's products)
constexpr int ME_MAX = 320;
constexpr int MNT = 512;          // threads: 8 waves, two patches each
constexpr int MCLS = 36;          // anchor classes
constexpr int MNULL = ME_MAX;     // LDS row of zeros outside every window: pads the patch lists
// patch lists: <= 9 patches per entry, <= 3 padding slots per patch, and room for the read-ahead of a
// patch's last chunk (4 slots past the last list)
constexpr int MLIST = 9 * ME_MAX + 16 * 3 + 4;
static_assert(ME_MAX <= MNT, "one staging thread per entry");
static_assert(MROW % 4 == 0 && (MROW / 4) % 2 == 1, "16-byte rows at an odd 16-byte stride: conflict-free staging");
static_assert(MROW * (ME_MAX + 1) < 65536, "row offsets in 16 bits");
static_assert(4 * (ME_MAX + 1) * MROW + 2 * (MLIST + 40) <= 160 * 1024 / 3, "three workgroups per CU");

struct MItem {
  int eoff, ne;            // entries[eoff, eoff + ne), sorted by anchor class
  int x0, y0;              // tile origin (l0, l1), centred coordinates
  int slot;                // -1: the item owns the tile; >= 0: partial slot
  int pad_[3];
  unsigned short cls[40];  // class c: entries [cls[c], cls[c + 1]), c < MCLS; cls[MCLS] = ne
};

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// Entries of patch (px, py) in run order (three runs of anchor rows py .. py + 2)
// padded with MNULL to a multiple of 4; returns the padded length.
__device__ __forceinline__ int patch_list_len(const unsigned short* cls, int px, int py) {
  int k = 0;
#pragma unroll
  for (int r = 0; r < 3; ++r) k += cls[(py + r) * 6 + px + 3] - cls[(py + r) * 6 + px];
  return (k + 3) & ~3;
}

// OFF64: byte offsets of the representatives in 64 bits (their buffer spans 4 GiB or more); otherwise
// 32-bit offsets from a wave-uniform base.
template <bool OFF64>
__global__ __launch_bounds__(MNT) __attribute__((amdgpu_waves_per_eu(6, 8))) void scatter_mfma(const float2* __restrict__ yh, float2* __restrict__ Hb,
                                                    float2* __restrict__ partial, const MItem* __restrict__ items,
                                                    const int* __restrict__ entries,
                                                    const unsigned short* __restrict__ rows,
                                                    const double2* __restrict__ dirs, int n, int H, int nz,
                                                    long long hzs, int zchunk, float a2, int xg) {
  __shared__ __align__(16) float R[(ME_MAX + 1) * MROW];
  __shared__ unsigned short LST[MLIST];
  __shared__ unsigned short CLS[40];
  const unsigned lw = xcd_group(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y, xg);
  const unsigned item = lw % gridDim.x;
  const int z0 = (int)(lw / gridDim.x) * zchunk, z1 = min(nz, z0 + zchunk);
  const MItem* it = items + item;
  const int ne = it->ne, x0 = it->x0, y0 = it->y0, slot = it->slot;
  const int t = threadIdx.x;
  if (t < 40) CLS[t] = it->cls[t];
  for (int j = t; j < MLIST; j += MNT) LST[j] = (unsigned short)(MNULL * MROW);  // read-ahead slots: valid rows
  // Staging: thread j takes entry j (item entries are in sample order, so neighbouring lanes stage
  // neighbouring samples) and writes its window base relative to the tile and its fractional
  // offset (mirror: negated) into the entry's LDS row (its place in anchor-class order).
  using Off = typename std::conditional<OFF64, long long, unsigned>::type;
  Off src = 0;  // byte offset of the entry's (theta, k) at z = 0
  bool cj = false;
  int myrow = MNULL;
  if (t < ne) {
    const int code = entries[it->eoff + t];
    myrow = rows[it->eoff + t];
    const SampleGeom g = sample_geom(code >> 1, H, dirs);
    cj = (code & 1) != 0;
    const int twon = 2 * n;
    const int cx = cj ? -g.rx : g.rx, cy = cj ? -g.ry : g.ry;
    const int bx = ((cx - M - x0 + 8) % twon + twon) % twon - 8;
    const int by = ((cy - M - y0 + 8) % twon + twon) % twon - 8;
    *reinterpret_cast<float4*>(&R[myrow * MROW + 2 * MZ]) =
        make_float4((float)(bx + M), (float)(by + M), cj ? -g.fx : g.fx, cj ? -g.fy : g.fy);
    src = (Off)(((long long)g.th * nz * H + g.k) * (long long)sizeof(float2));
  }
  if (t == MNULL) {
    float4* r = reinterpret_cast<float4*>(&R[MNULL * MROW]);
#pragma unroll
    for (int j = 0; j < MROW / 4 - 1; ++j) r[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    r[MROW / 4 - 1] = make_float4(1e4f, 1e4f, 0.f, 0.f);
  }
  __syncthreads();
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int l = t & 63, kk = l >> 4, q = l & 15;
  const int py = w >> 1;  // this wave's patch row; patches px = 2 (w & 1) + {0, 1}
  const float fqy = (float)(4 * py + (q >> 2));
  // patch lists: patch P = 4 py + px at LST[off(P)], off = the padded lengths of the patches before
  // it; each wave writes its own two
  int off[2], len[2];
  {
    int o = 0;
    for (int P = 0; P < 4 * py + 2 * (w & 1); ++P) o += patch_list_len(CLS, P & 3, P >> 2);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int px = 2 * (w & 1) + i;
      off[i] = __builtin_amdgcn_readfirstlane(o);
      len[i] = __builtin_amdgcn_readfirstlane(patch_list_len(CLS, px, py));
      o += len[i];
      const int s0 = CLS[py * 6 + px], L0 = CLS[py * 6 + px + 3] - s0;
      const int s1 = CLS[(py + 1) * 6 + px], L1 = CLS[(py + 1) * 6 + px + 3] - s1;
      const int s2 = CLS[(py + 2) * 6 + px];
      const int L01 = L0 + L1, K = L01 + CLS[(py + 2) * 6 + px + 3] - s2;
      for (int j = l; j < len[i]; j += 64) {
        const int e = j < L0 ? s0 + j : (j < L01 ? s1 + j - L0 : (j < K ? s2 + j - L01 : MNULL));
        LST[off[i] + j] = (unsigned short)(e * MROW);  // the row's first float
      }
    }
  }
  // Staging at the start of each round.  Every lane issues all its loads from valid addresses
  // (the round's last slice and sample 0 stand in for missing ones) and zeroes what it does not
  // use, so the loads stay in flight together.  (Loading round r + 1 into registers during round
  // r measured slower: the second workgroup on the CU already overlaps one's staging with the
  // other's products.)
  float2 v[MZ];
  // (the representatives span less than 4 GiB unless OFF64: byte offsets fit the buffer form, whose
  // range check would return zeros rather than fault)
  const __amdgpu_buffer_rsrc_t ybuf = __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(yh), (short)0, -1, 0x00020000);
  auto load_round = [&](int zr) {
    const int nzr = min(MZ, z1 - zr);
#pragma unroll
    for (int zz = 0; zz < MZ; ++zz) {
      const long long zoff = (long long)(zr + min(zz, nzr - 1)) * H * (long long)sizeof(float2);
      if constexpr (OFF64) {
        v[zz] = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(yh) + zoff + src);
      } else {  // buffer load: per-lane 32-bit offset plus the slice's (uniform) offset
        const u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(ybuf, (int)src, (int)zoff, 0);
        v[zz] = make_float2(__uint_as_float(u.x), __uint_as_float(u.y));
      }
    }
#pragma unroll
    for (int zz = 0; zz < MZ; ++zz)
      if (t >= ne || zz >= nzr) v[zz] = make_float2(0.f, 0.f);
  };
  for (int zc = z0; zc < z1; zc += MZ) {
    const int nzz = min(MZ, z1 - zc);
    load_round(zc);
    if (t < ne) {
      float4* dst = reinterpret_cast<float4*>(&R[myrow * MROW]);
      const float sy = cj ? -1.f : 1.f;
#pragma unroll
      for (int j = 0; j < MZ / 2; ++j)
        dst[j] = make_float4(v[2 * j].x, sy * v[2 * j].y, v[2 * j + 1].x, sy * v[2 * j + 1].y);
    }
    __syncthreads();
#pragma unroll 1
    for (int i = 0; i < 2; ++i) {
      const int px = 2 * (w & 1) + i;
      const float fqx = (float)(4 * px + (q & 3));
      const unsigned short* lp = LST + off[i] + kk;
      // one chunk of 4 entries per step (MZB independent accumulators); the next chunk's row
      // offset is read at the top of the step, off the dependent path
      f32x4 a[MZB];
#pragma unroll
      for (int j = 0; j < MZB; ++j) a[j] = f32x4{0.f, 0.f, 0.f, 0.f};
      int ea = lp[0];
      for (int c = 0; c < len[i]; c += 4) {
        const int na = lp[c + 4];
        __builtin_amdgcn_sched_barrier(0);  // issue that read first: it is waited for at the step's end
        const float* ra = R + ea;
        const float4 ma = *reinterpret_cast<const float4*>(ra + 2 * MZ);
        float va[MZB];
#pragma unroll
        for (int j = 0; j < MZB; ++j) va[j] = ra[16 * j + q];
        // tap offset from the window centre (exact), then the offset from the sample
        const float tx = fqx - ma.x, ty = fqy - ma.y;
        const float dx = tx - ma.z, dy = ty - ma.w;
        float wgt = __builtin_amdgcn_exp2f(a2 * (dx * dx + dy * dy));
        wgt = fmaxf(fabsf(tx), fabsf(ty)) <= (float)M ? wgt : 0.f;
#pragma unroll
        for (int j = 0; j < MZB; ++j) a[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(va[j], wgt, a[j], 0, 0, 0);
        ea = na;
        asm volatile("" : "+v"(ea));  // keeps the offset read a step ahead (not re-read here)
      }
      // lane (kk, q) holds slices 8 j + 2 kk and 8 j + 2 kk + 1 of point q in block j
      const int col = x0 + 4 * px + (q & 3), lrow = y0 + 4 * py + (q >> 2);
      const float sgn = ((col + lrow) & 1) ? -1.f : 1.f;
      float2* dst;
      long long zs;
      bool live = true;
      if (slot < 0) {
        dst = Hb + (long long)zc * hzs + hs_idx(lrow < 0 ? lrow + 2 * n : lrow, col, 2 * n);
        zs = hzs;
        live = col <= n;
      } else {
        dst = partial + ((long long)slot * nz + zc) * SNP + (4 * py + px) * 16 + q;
        zs = SNP;
      }
      if (live) {
#pragma unroll
        for (int j = 0; j < MZB; ++j) {
          const int s = 8 * j + 2 * kk;
          if (s < nzz) dst[(long long)s * zs] = make_float2(sgn * a[j][0], sgn * a[j][1]);
          if (s + 1 < nzz) dst[(long long)(s + 1) * zs] = make_float2(sgn * a[j][2], sgn * a[j][3]);
        }
      }
    }
    __syncthreads();
  }
}

// Row pass of RT: C2R inverse transform of the half-spectrum rows n/2 + y
// (after the column IFFT), centre crop, times phi.  Real output packs two rows
// per complex FFT; complex output packs the real/imaginary channels of one row.
// The 2C input rows of a workgroup are first staged block-ordered (coalesced):
// row A of transform c at x_c[lpad(col)], row B at x_c[OFF + lpad(col)], with
// OFF = 8 (mod 32) bank pairs so that the staging writes of a half-wave (8
// columns x rows A, B of 2 transforms at a 16-pair stride) hit 32 distinct
// bank pairs.  Row addresses are worked out once per workgroup.
template <int L>
struct C2rOff {
  static constexpr int A = lpad(L / 2) + 1;  // > every lpad(col), col <= n
  // = 8 (mod 32); short transforms (many per workgroup) skip the alignment to fit LDS
  static constexpr int OFF = L >= 512 ? A + ((8 - A % 32) + 32) % 32 : A;
};
template <int L>
using C2rShape = RowReg<L, C2rOff<L>::OFF + C2rOff<L>::A - (L + L / 16)>;

template <int L, bool CPLX>
__global__ __launch_bounds__(C2rShape<L>::NT) void rt4_rows_c2r(const float2* __restrict__ Hre,
                                                              const float2* __restrict__ Him, void* __restrict__ out_v,
                                                              const float* __restrict__ phx, const float* __restrict__ phy,
                                                              int n, int nz, long long hzs, float scale,
                                                              const float2* __restrict__ tw) {
  using Sh = C2rShape<L>;
  constexpr int OFF = C2rOff<L>::OFF;
  constexpr int R2 = 2 * Sh::C;
  static_assert(OFF + C2rOff<L>::A <= Sh::S, "staging fits the transform's LDS");
  __shared__ float2 lds[Sh::LDS];
  // offset of source row r (transform r / 2, A/B = r & 1; complex output: B from Him) in its half
  // plane (-1: none).  Offsets from the kernel arguments keep the loads global; a pointer read back
  // from LDS made them flat loads that each waited before the next.
  __shared__ long long rowoff[R2];
  // with one transform (two rows) per workgroup, two workgroups share each 128-byte line of a block
  // (4 rows x 32 bytes): keep them adjacent in the XCD-grouped order, so the line is merged in one L2
  const unsigned wg = Sh::C == 1 ? xcd_group(blockIdx.x, gridDim.x, 2) : blockIdx.x;
  const int half = n / 2;
  const long long nunits = CPLX ? (long long)nz * n : (long long)nz * half;
  if (threadIdx.x < R2) {
    const long long u = (long long)wg * Sh::C + (threadIdx.x >> 1);
    const int which = threadIdx.x & 1;
    long long ro = -1;
    if (u < nunits) {
      int z, row;
      if (CPLX) {
        z = (int)(u / n);
        row = half + (int)(u - (long long)z * n);
      } else {
        z = (int)(u / half);
        row = half + 2 * (int)(u - (long long)z * half) + which;
      }
      ro = (long long)z * hzs + 4LL * row;  // hs_idx(row, col) = this + col part
    }
    rowoff[threadIdx.x] = ro;
  }
  __syncthreads();
  // all of this thread's staging loads are issued before its LDS stores
  constexpr int NBLK = L / 8 + 1;  // n / 4 + 1 column blocks
  constexpr int TOT = NBLK * R2 * 4, IT = (TOT + Sh::NT - 1) / Sh::NT;
  float2 sv[IT];
  int sd[IT];
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int idx = threadIdx.x + i * Sh::NT;
    sd[i] = -1;
    sv[i] = make_float2(0.f, 0.f);
    if (idx >= TOT) continue;
    const int q = idx & 3, rest = idx >> 2;
    const int r = rest % R2, blk = rest / R2;
    const int col = 4 * blk + q;
    const long long ro = rowoff[r];
    if (col > n || ro < 0) continue;
    const float2* src = (CPLX && (r & 1)) ? Him : Hre;
    sv[i] = src[ro + (long long)blk * 4 * L + q];
    sd[i] = (r >> 1) * Sh::S + (r & 1) * OFF + lpad(col);
  }
#pragma unroll
  for (int i = 0; i < IT; ++i)
    if (sd[i] >= 0) lds[sd[i]] = sv[i];
  __syncthreads();
  int c, t;
  Sh::map(threadIdx.x, c, t);
  float2* x = lds + c * Sh::S;
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int kx = t + m * Sh::TPC;
    float2 a, b;
    if (kx <= L / 2) {  // n = L/2
      a = x[lpad(kx)];
      b = x[OFF + lpad(kx)];
      if (kx == 0 || kx == L / 2) {
        a.y = 0.f;
        b.y = 0.f;
      }
    } else {
      a = cconj(x[lpad(L - kx)]);
      b = cconj(x[OFF + lpad(L - kx)]);
    }
    v[m] = make_float2(a.x - b.y, a.y + b.x);
  }
  __syncthreads();  // the transform's exchanges reuse the staging area
  fft_reg<L, Sh::TPC, true>(v, x, t, tw);
  const long long unit = (long long)wg * Sh::C + c;
  if (unit >= nunits) return;
  int z, y;
  if (CPLX) {
    z = (int)(unit / n);
    y = (int)(unit - (long long)z * n);
  } else {
    z = (int)(unit / half);
    y = 2 * (int)(unit - (long long)z * half);
  }
  // centre crop: points [L/4, 3L/4) = m in [PT/4, 3PT/4); `scale` (both
  // inverse passes' 1/L) rides in the row weights
  static_assert(Sh::PT % 4 == 0 && Sh::PT * Sh::TPC == L, "L/4 is a multiple of TPC");
  const float pya = phy[y] * scale, pyb = CPLX ? 0.f : phy[y + 1] * scale;
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    if (m < Sh::PT / 4 || m >= 3 * Sh::PT / 4) continue;
    const int xx = t + (m - Sh::PT / 4) * Sh::TPC;
    const float px = phx[xx];
    if (CPLX) {
      const float p = pya * px;
      ((float2*)out_v)[((long long)z * n + y) * n + xx] = make_float2(v[m].x * p, v[m].y * p);
    } else {
      float* o = (float*)out_v + ((long long)z * n + y) * n;
      o[xx] = v[m].x * (pya * px);
      o[n + xx] = v[m].y * (pyb * px);
    }
  }
}

// ---------------------------------------------------------------------------
// 2n-point transforms longer than one workgroup holds (n > kMaxFFT / 2)
// ---------------------------------------------------------------------------
// The 2n-point rows and columns of R and RT have n nonzero inputs (forward) or
// n kept outputs (inverse), centred at n/2; each becomes two n-point transforms
// of one workgroup, with h_j = exp(-i pi j / n) from a per-plan FP64-built
// table (hsplit):
//   forward, split by the parity of the frequency (x centred at n/2):
//     Z[2m] = (-1)^m V[m],  Z[2m+1] = -i (-1)^m W[m],  V = FFT_n(x),  W = FFT_n(x_j h_j);
//   inverse, split by the parity of the output point:
//     y[2r] = IFFT_n(Z[k] + Z[k+n])[r],  y[2r+1] = IFFT_n((Z[k] - Z[k+n]) conj(h_k))[r],
//     and the kept points p in [n/2, 3n/2) are r in [n/4, 3n/4) of both halves.
// Both halves are formed by one workgroup, one after the other: every input is
// in registers before the first transform's first barrier, so the in-place
// column passes may then write, and the first half's results leave before the
// second transform (one register copy of the inputs is live, not two).
template <int L>  // L = n: one transform per workgroup
using SplitShape = RegShape<L, L>;


// R's row pass: workgroup 2 u + par forms the even (par = 0) or odd (par = 1)
// half-spectrum columns of row pair u (rows y0, y0 + 1 of slice z), unpacking
// the two real rows from the complex transform as r1_rows_rfft does:
// A_k = (Z_k + conj Z_{2n-k}) / 2, B_k = (Z_k - conj Z_{2n-k}) / (2i), where
// Z_{2n-k} is V[n-m] (k = 2m) or W[n-1-m] (k = 2m+1) with the same sign.
template <int L>
__global__ __launch_bounds__(SplitShape<L>::NT) void r1_rows_split(const float* __restrict__ obj,
                                                                 float2* __restrict__ fde,
                                                                 const float* __restrict__ phx,
                                                                 const float* __restrict__ phy,
                                                                 const float2* __restrict__ hsplit, int nz,
                                                                 long long fzs, const float2* __restrict__ tw) {
  using Sh = SplitShape<L>;
  static_assert(Sh::C == 1 && Sh::NT == Sh::TPC, "one transform per workgroup");
  __shared__ float2 lds[Sh::LDS];
  constexpr int half = L / 2;
  // the four workgroups of two row pairs (one 128-byte line of every block) share an L2
  const unsigned wg = xcd_group(blockIdx.x, gridDim.x, 4);
  const long long pair = wg >> 1;
  const int par = (int)(wg & 1u);
  if (pair >= (long long)nz * half) return;  // the whole workgroup, before any barrier
  const int t = threadIdx.x;
  const int z = (int)(pair / half), y0 = 2 * (int)(pair - (long long)z * half);
  const float* ra = obj + ((long long)z * L + y0) * L;
  const float* rb = ra + L;
  const float pya = phy[y0], pyb = phy[y0 + 1];
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int j = t + m * Sh::TPC;
    const float p = phx[j];
    const float2 w = make_float2(ra[j] * pya * p, rb[j] * pyb * p);
    v[m] = par ? cmul(w, hsplit[j]) : w;
  }
  fft_reg<L, Sh::TPC, false>(v, lds, t, tw);
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) lds[lpad(t + m * Sh::TPC)] = v[m];
  __syncthreads();
  const long long rowa = (long long)z * fzs + 4LL * (half + y0);  // hs_idx(half + y0, col) minus the column part
  const int cnt = par ? half : half + 1;
  for (int mm = t; mm < cnt; mm += Sh::NT) {
    const int col = 2 * mm + par;
    const float2 zk = lds[lpad(mm)];
    const float2 zm = lds[lpad(par ? L - 1 - mm : (mm == 0 ? 0 : L - mm))];
    float2 a = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));  // (zk + conj zm) / 2
    float2 d = make_float2(0.5f * (zk.x - zm.x), 0.5f * (zk.y + zm.y));  // (zk - conj zm) / 2
    float2 b;
    if (par) {
      a = make_float2(a.y, -a.x);  // -i a
      b = make_float2(-d.x, -d.y);
    } else {
      b = make_float2(d.y, -d.x);  // d / i
    }
    if (mm & 1) {
      a = make_float2(-a.x, -a.y);
      b = make_float2(-b.x, -b.y);
    }
    const long long o = rowa + (long long)(col >> 2) * 4 * (2 * L) + (col & 3);
    fde[o] = a;      // row half + y0
    fde[o + 4] = b;  // row half + y0 + 1
  }
}

// Column pass (in place) of one half-spectrum column of a z plane: forward, the
// object rows [n/2, 3n/2) to the rows of the disc (supp); inverse, the disc's
// rows to rows [n/2, 3n/2), unnormalised (the C2R row pass applies both 1/(2n)).
template <int L, bool INV>
__global__ __launch_bounds__(SplitShape<L>::NT) void col_fft_split(float2* __restrict__ buf, int ncols, long long zs,
                                                                 const int2* __restrict__ supp,
                                                                 const float2* __restrict__ hsplit,
                                                                 const float2* __restrict__ tw) {
  using Sh = SplitShape<L>;
  static_assert(Sh::C == 1 && Sh::NT == Sh::TPC, "one transform per workgroup");
  __shared__ float2 lds[Sh::LDS];
  constexpr int L2 = 2 * L;
  // the four columns of a block (one 32-byte row piece) share an L2
  const unsigned lw = xcd_group(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y, 4);
  const int col = (int)(lw % gridDim.x);
  if (col >= ncols) return;  // the whole workgroup, before any barrier
  // column col of the z plane: row i at cb[4 i] (hs_idx with 32-bit offsets from one base)
  float2* cb = buf + (long long)(lw / gridDim.x) * zs + (long long)(col >> 2) * (4 * L2) + (col & 3);
  const int2 sp = supp[col >> 2];
  auto carries = [&](int i) {
    const int l = i < L ? i : i - L2;
    return l >= sp.x && l < sp.y;
  };
  const int t = threadIdx.x;
  float2 v[Sh::PT];
  if constexpr (!INV) {
    float2 x[Sh::PT];
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) {
      x[m] = cb[4 * (L / 2 + t + m * Sh::TPC)];
      v[m] = x[m];
    }
    fft_reg<L, Sh::TPC, false>(v, lds, t, tw);
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) {
      const int mm = t + m * Sh::TPC;
      if (carries(2 * mm)) cb[4 * (2 * mm)] = (mm & 1) ? make_float2(-v[m].x, -v[m].y) : v[m];
      v[m] = cmul(x[m], hsplit[mm]);
    }
    fft_reg<L, Sh::TPC, false>(v, lds, opaque_index(t), opaque_twiddles(tw));
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) {
      const int mm = t + m * Sh::TPC;
      const float2 w = make_float2(v[m].y, -v[m].x);  // -i W
      if (carries(2 * mm + 1)) cb[4 * (2 * mm + 1)] = (mm & 1) ? make_float2(-w.x, -w.y) : w;
    }
  } else {
    float2 d[Sh::PT];  // (Z[k] - Z[k+n]) conj(h_k): the odd output points' half
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) {
      const int k = t + m * Sh::TPC;
      const float2 lo = carries(k) ? cb[4 * k] : make_float2(0.f, 0.f);
      const float2 hi = carries(k + L) ? cb[4 * (k + L)] : make_float2(0.f, 0.f);
      v[m] = make_float2(lo.x + hi.x, lo.y + hi.y);
      d[m] = cmul(make_float2(lo.x - hi.x, lo.y - hi.y), cconj(hsplit[k]));
    }
    constexpr int M0 = Sh::PT / 4, M1 = 3 * Sh::PT / 4;  // r in [n/4, 3n/4)
    static_assert(Sh::PT % 4 == 0, "n/4 is a multiple of TPC");
    fft_reg<L, Sh::TPC, true>(v, lds, t, tw);
#pragma unroll
    for (int m = M0; m < M1; ++m) cb[8 * (t + m * Sh::TPC)] = v[m];
    fft_reg<L, Sh::TPC, true>(d, lds, opaque_index(t), opaque_twiddles(tw));
#pragma unroll
    for (int m = M0; m < M1; ++m) cb[8 * (t + m * Sh::TPC) + 4] = d[m];
  }
}

// RT's row pass (C2R, crop, phi) for 2n-point rows: Z = A + i B over the full
// 2n spectrum from the half-spectrum rows (Hermitian extension, imaginary parts
// dropped at k = 0 and k = n as rt4_rows_c2r does), the output-parity halves
// above, and output point p = 2r + parity to column p - n/2.  One unit (a row
// pair, or one complex row) per workgroup; Z[k] and Z[k+n] are read straight
// from HBM (columns k and n - k: consecutive lanes, consecutive columns).
template <int L, bool CPLX>
__global__ __launch_bounds__(SplitShape<L>::NT) void rt4_rows_c2r_split(const float2* __restrict__ Hre,
                                                                      const float2* __restrict__ Him,
                                                                      void* __restrict__ out_v,
                                                                      const float* __restrict__ phx,
                                                                      const float* __restrict__ phy,
                                                                      const float2* __restrict__ hsplit, int nz,
                                                                      long long hzs, float scale,
                                                                      const float2* __restrict__ tw) {
  using Sh = SplitShape<L>;
  static_assert(Sh::C == 1 && Sh::NT == Sh::TPC, "one transform per workgroup");
  __shared__ float2 lds[Sh::LDS];
  constexpr int half = L / 2, L2 = 2 * L;
  const long long nunits = CPLX ? (long long)nz * L : (long long)nz * half;
  // the units of one 128-byte line of every block (4 rows) share an L2
  const unsigned wg = xcd_group(blockIdx.x, gridDim.x, CPLX ? 4 : 2);
  const long long u = wg;
  if (u >= nunits) return;  // the whole workgroup, before any barrier
  int z, y;
  if (CPLX) {
    z = (int)(u / L);
    y = (int)(u - (long long)z * L);
  } else {
    z = (int)(u / half);
    y = 2 * (int)(u - (long long)z * half);
  }
  const float2* pa = Hre + (long long)z * hzs + 4LL * (half + y);
  const float2* pb = CPLX ? Him + (long long)z * hzs + 4LL * (half + y) : pa + 4;
  auto at = [&](const float2* row, int col) { return row[(col >> 2) * (4 * L2) + (col & 3)]; };  // < 2^31
  auto comb = [](float2 a, float2 bb) { return make_float2(a.x - bb.y, a.y + bb.x); };  // a + i b
  const int t = threadIdx.x;
  float2 v[Sh::PT], d[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int k = t + m * Sh::TPC;
    float2 a0 = at(pa, k), b0 = at(pb, k);        // Z[k], k < n
    float2 a1 = at(pa, L - k), b1 = at(pb, L - k);  // Z[k + n] = conj of column n - k (k > 0), column n (k = 0)
    if (k == 0) {
      a0.y = b0.y = 0.f;
      a1.y = b1.y = 0.f;
    } else {
      a1.y = -a1.y;
      b1.y = -b1.y;
    }
    const float2 lo = comb(a0, b0), hi = comb(a1, b1);
    v[m] = make_float2(lo.x + hi.x, lo.y + hi.y);
    d[m] = cmul(make_float2(lo.x - hi.x, lo.y - hi.y), cconj(hsplit[k]));
  }
  constexpr int M0 = Sh::PT / 4, M1 = 3 * Sh::PT / 4;  // r in [n/4, 3n/4): output columns 2r + parity - n/2
  static_assert(Sh::PT % 4 == 0, "n/4 is a multiple of TPC");
  const float pya = phy[y] * scale, pyb = CPLX ? 0.f : phy[y + 1] * scale;
  auto emit = [&](float2 yv, int xx) {
    const float px = phx[xx];
    if (CPLX) {
      const float pw = pya * px;
      ((float2*)out_v)[((long long)z * L + y) * L + xx] = make_float2(yv.x * pw, yv.y * pw);
    } else {
      float* o2 = (float*)out_v + ((long long)z * L + y) * L;
      o2[xx] = yv.x * (pya * px);
      o2[L + xx] = yv.y * (pyb * px);
    }
  };
  fft_reg<L, Sh::TPC, true>(v, lds, t, tw);
#pragma unroll
  for (int m = M0; m < M1; ++m) emit(v[m], 2 * (t + m * Sh::TPC) - half);
  fft_reg<L, Sh::TPC, true>(d, lds, opaque_index(t), opaque_twiddles(tw));
#pragma unroll
  for (int m = M0; m < M1; ++m) emit(d[m], 2 * (t + m * Sh::TPC) + 1 - half);
}

// ---------------------------------------------------------------------------
// Detector widths above 16384 (n = 2 Q, Q <= 16384): every transform on Q points
// ---------------------------------------------------------------------------
// The 2n = 4Q-point rows and columns split four ways by decimation in frequency
// (forward) or by output residue (inverse); the n = 2Q-point lines split two
// ways, as the halves above.  With w = exp(-i pi / n) (the hsplit table,
// w^j for j < n):
//   forward, x centred in the 2n-point line (x_s at s + Q, s < n):
//     Z[4m + r] = FFT_Q(y)[m],  y_s = w^(r s) ((-i)^r x_s + (-1)^r x_(s+Q)),  s < Q;
//   inverse, kept points p = 4t + r in [Q, 3Q), i.e. t in [Q/4, 3Q/4):
//     z[4t + r] = IFFT_Q(u)[t],  u_k = conj(w^(r k)) sum_q i^(r q) Z[k + q Q],  k < Q;
//   lines, split by frequency parity e (forward) or output parity e (inverse):
//     V[2m + e] = FFT_Q(w^(2 e s) (x_s + (-1)^e x_(s+Q)))[m],
//     y[2t + e] = IFFT_Q(conj(w^(2 e k)) (Z[k] + (-1)^e Z[k+Q]))[t].
// One workgroup per (line, residue): each reads its whole line, so the
// residues of one column cannot share it in place -- R's row pass writes, and
// RT's inverse column pass writes, a second half-spectrum buffer B holding only
// the n object rows (B row s = padded row s + Q, block layout of n rows).  Rows
// are transformed one real row at a time (the residues r and 4 - r of a packed
// pair would need each other).  These passes read each input four times
// (columns, rows) or twice (lines): they serve the size range, not speed.
__device__ __forceinline__ float2 big_w(const float2* __restrict__ h, int j, int n) {  // w^j, 0 <= j < 2n
  return j < n ? h[j] : make_float2(-h[j - n].x, -h[j - n].y);
}
__device__ __forceinline__ float2 mul_negi_pow(float2 x, int r) {  // (-i)^r x
  switch (r & 3) {
    case 0: return x;
    case 1: return make_float2(x.y, -x.x);
    case 2: return make_float2(-x.x, -x.y);
    default: return make_float2(-x.y, x.x);
  }
}
__device__ __forceinline__ float2 mul_i_pow(float2 x, int r) { return mul_negi_pow(x, (4 - (r & 3)) & 3); }  // i^r x
__host__ __device__ __forceinline__ long long big_idx(int row, int col, int n) {  // B: blocks of 4 columns x n rows
  return ((long long)(col >> 2) * n + row) * 4 + (col & 3);
}

// R's row pass: workgroup ((z, y), r) -> B row y, columns k = 4m + r <= n.
template <int Q>
__global__ __launch_bounds__(SplitShape<Q>::NT) void big_r1_rows(const float* __restrict__ obj, float2* __restrict__ B,
                                                               const float* __restrict__ phx,
                                                               const float* __restrict__ phy,
                                                               const float2* __restrict__ h, long long bzs,
                                                               const float2* __restrict__ tw) {
  using Sh = SplitShape<Q>;
  __shared__ float2 lds[Sh::LDS];
  constexpr int n = 2 * Q;
  const long long u = blockIdx.x;  // the grid is exactly nz * n * 4 workgroups
  const int r = (int)(u & 3);
  const long long zy = u >> 2;
  const int z = (int)(zy / n), y = (int)(zy - (long long)z * n);
  const float* row = obj + ((long long)z * n + y) * n;
  const float py = phy[y];
  const int t = threadIdx.x;
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int j = t + m * Sh::TPC;
    const float x0 = row[j] * py * phx[j], x1 = row[j + Q] * py * phx[j + Q];
    const float2 a = mul_negi_pow(make_float2(x0, 0.f), r);
    const float2 yv = make_float2(a.x + ((r & 1) ? -x1 : x1), a.y);
    v[m] = r ? cmul(yv, big_w(h, r * j, n)) : yv;
  }
  fft_reg<Q, Sh::TPC, false>(v, lds, t, tw);
  float2* b = B + (long long)z * bzs;
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int k = 4 * (t + m * Sh::TPC) + r;
    if (k <= n) b[big_idx(y, k, n)] = v[m];
  }
}

// Column passes.  Forward: B column col (the n object rows) -> the rows of the
// disc (supp) of fde column col, raw frequency order.  Inverse: the disc's rows
// of fde column col -> B rows s = p - Q, unnormalised (the row pass applies
// both 1/(2n)).  Workgroup (col, r), blockIdx.y = z.
template <int Q, bool INV>
__global__ __launch_bounds__(SplitShape<Q>::NT) void big_col(float2* __restrict__ fde, float2* __restrict__ B,
                                                           int ncols, long long zs, long long bzs,
                                                           const int2* __restrict__ supp,
                                                           const float2* __restrict__ h,
                                                           const float2* __restrict__ tw) {
  using Sh = SplitShape<Q>;
  __shared__ float2 lds[Sh::LDS];
  constexpr int n = 2 * Q, L2 = 4 * Q;
  const int col = (int)(blockIdx.x >> 2), r = (int)(blockIdx.x & 3);
  if (col >= ncols) return;  // the whole workgroup, before any barrier
  float2* fb = fde + (long long)blockIdx.y * zs + (long long)(col >> 2) * (4LL * L2) + (col & 3);  // row k: fb[4 k]
  float2* bb = B + (long long)blockIdx.y * bzs + (long long)(col >> 2) * (4LL * n) + (col & 3);    // row s: bb[4 s]
  const int2 sp = supp[col >> 2];
  auto carries = [&](int k) {
    const int l = k < n ? k : k - L2;
    return l >= sp.x && l < sp.y;
  };
  const int t = threadIdx.x;
  float2 v[Sh::PT];
  if constexpr (!INV) {
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) {
      const int j = t + m * Sh::TPC;
      const float2 x0 = bb[4LL * j], x1 = bb[4LL * (j + Q)];
      const float2 a = mul_negi_pow(x0, r);
      const float2 yv = (r & 1) ? make_float2(a.x - x1.x, a.y - x1.y) : make_float2(a.x + x1.x, a.y + x1.y);
      v[m] = r ? cmul(yv, big_w(h, r * j, n)) : yv;
    }
    fft_reg<Q, Sh::TPC, false>(v, lds, t, tw);
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) {
      const int k = 4 * (t + m * Sh::TPC) + r;
      if (carries(k)) fb[4LL * k] = v[m];
    }
  } else {
#pragma unroll
    for (int m = 0; m < Sh::PT; ++m) {
      const int k = t + m * Sh::TPC;
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) {
        const int kk = k + qq * Q;
        if (carries(kk)) {
          const float2 zq = mul_i_pow(fb[4LL * kk], r * qq);
          acc.x += zq.x;
          acc.y += zq.y;
        }
      }
      v[m] = r ? cmul(acc, cconj(big_w(h, r * k, n))) : acc;
    }
    constexpr int M0 = Sh::PT / 4, M1 = 3 * Sh::PT / 4;  // t in [Q/4, 3Q/4): p = 4t + r in [Q, 3Q)
    static_assert(Sh::PT % 4 == 0, "Q/4 is a multiple of TPC");
    fft_reg<Q, Sh::TPC, true>(v, lds, t, tw);
#pragma unroll
    for (int m = M0; m < M1; ++m) bb[4LL * (4 * (t + m * Sh::TPC) + r - Q)] = v[m];
  }
}

// RT's row pass: C2R of the 2n-point rows from B (Hermitian half rows k <= n,
// imaginary parts dropped at k = 0 and k = n), kept points p = 4t + r in
// [Q, 3Q) to columns p - Q, times phi.  A real output packs rows y, y + 1 as
// Z = A + i B; a complex one takes Z = Hre + i Him of row y.
template <int Q, bool CPLX>
__global__ __launch_bounds__(SplitShape<Q>::NT) void big_rows_c2r(const float2* __restrict__ Bre,
                                                                const float2* __restrict__ Bim,
                                                                void* __restrict__ out_v,
                                                                const float* __restrict__ phx,
                                                                const float* __restrict__ phy,
                                                                const float2* __restrict__ h, long long bzs,
                                                                float scale, const float2* __restrict__ tw) {
  using Sh = SplitShape<Q>;
  __shared__ float2 lds[Sh::LDS];
  constexpr int n = 2 * Q, L2 = 4 * Q;
  const long long u = blockIdx.x >> 2;  // the grid is exactly units * 4 workgroups
  const int r = (int)(blockIdx.x & 3);
  int z, y;
  if (CPLX) {
    z = (int)(u / n);
    y = (int)(u - (long long)z * n);
  } else {
    z = (int)(u / Q);
    y = 2 * (int)(u - (long long)z * Q);
  }
  const float2* pa = Bre + (long long)z * bzs + 4LL * y;
  const float2* pb = CPLX ? Bim + (long long)z * bzs + 4LL * y : pa + 4;
  auto at = [&](const float2* row, int col) { return row[(long long)(col >> 2) * (4LL * n) + (col & 3)]; };
  auto spec = [&](const float2* row, int k) {  // Hermitian extension of the half row to k in [0, 2n)
    if (k == 0 || k == n) return make_float2(at(row, k).x, 0.f);
    return k < n ? at(row, k) : cconj(at(row, L2 - k));
  };
  const int t = threadIdx.x;
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int k = t + m * Sh::TPC;
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      const float2 a = spec(pa, k + qq * Q), b = spec(pb, k + qq * Q);
      const float2 zq = mul_i_pow(make_float2(a.x - b.y, a.y + b.x), r * qq);  // i^(r q) (a + i b)
      acc.x += zq.x;
      acc.y += zq.y;
    }
    v[m] = r ? cmul(acc, cconj(big_w(h, r * k, n))) : acc;
  }
  fft_reg<Q, Sh::TPC, true>(v, lds, t, tw);
  constexpr int M0 = Sh::PT / 4, M1 = 3 * Sh::PT / 4;
  const float pya = phy[y] * scale, pyb = CPLX ? 0.f : phy[y + 1] * scale;
#pragma unroll
  for (int m = M0; m < M1; ++m) {
    const int xx = 4 * (t + m * Sh::TPC) + r - Q;
    const float px = phx[xx];
    if (CPLX) {
      const float pw = pya * px;
      ((float2*)out_v)[((long long)z * n + y) * n + xx] = make_float2(v[m].x * pw, v[m].y * pw);
    } else {
      float* o = (float*)out_v + ((long long)z * n + y) * n;
      o[xx] = v[m].x * (pya * px);
      o[n + xx] = v[m].y * (pyb * px);
    }
  }
}

// R's line pass on n = 2Q points: rows (theta, z), (theta + 1, z) as one complex
// transform (r4_line_ifft), output parity e = blockIdx.x & 1.
template <int Q>
__global__ __launch_bounds__(SplitShape<Q>::NT) void big_line_ifft(const float2* __restrict__ gh, float* __restrict__ sino,
                                                                 const int* __restrict__ sig, int H, long long nrows,
                                                                 int nz, float scale, const float2* __restrict__ h,
                                                                 const float2* __restrict__ tw) {
  using Sh = SplitShape<Q>;
  __shared__ float2 lds[Sh::LDS];
  constexpr int L = 2 * Q, half = Q;
  const int e = (int)(blockIdx.x & 1);
  const long long q = blockIdx.x >> 1, th2 = q / nz;
  const long long ra = th2 * 2 * nz + (q - th2 * nz), rb = ra + nz;
  const bool va = ra < nrows, vb = rb < nrows;
  const int sga = va ? sig[ra / nz] : 1, sgb = vb ? sig[rb / nz] : 1;
  const float2* srca = gh + (va ? ra : 0) * H;
  const float2* srcb = gh + (vb ? rb : 0) * H;
  auto rebuild = [&](const float2* src, int sg, bool valid, int r) {
    const int d = r - half;
    int kk;
    bool cj;
    if (sg > 0) {
      kk = d >= 0 ? d : -d;
      cj = d < 0;
    } else {
      kk = d <= 0 ? -d : d;
      cj = d > 0;
    }
    float2 w = valid ? src[kk] : make_float2(0.f, 0.f);
    if (cj) w.y = -w.y;
    if (r == 0 || r == half) w.y = 0.f;  // Hermitian part
    return w;
  };
  auto line = [&](int r) {  // c1d . (a + i b) at r
    const float2 a = rebuild(srca, sga, va, r), b = rebuild(srcb, sgb, vb, r);
    const float2 w = make_float2(a.x - b.y, a.y + b.x);
    return (r & 1) ? w : make_float2(-w.x, -w.y);
  };
  const int t = threadIdx.x;
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int k = t + m * Sh::TPC;
    const float2 lo = line(k), hi = line(k + Q);
    const float2 sum = e ? make_float2(lo.x - hi.x, lo.y - hi.y) : make_float2(lo.x + hi.x, lo.y + hi.y);
    v[m] = e ? cmul(sum, cconj(h[2 * k])) : sum;
  }
  fft_reg<Q, Sh::TPC, true>(v, lds, t, tw);
  float* dsta = sino + (va ? ra : 0) * L;
  float* dstb = sino + (vb ? rb : 0) * L;
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int s = 2 * (t + m * Sh::TPC) + e;
    const float sc = (s & 1) ? scale : -scale;
    if (va) dsta[s] = v[m].x * sc;
    if (vb) dstb[s] = v[m].y * sc;
  }
}

// RT's line pass on n = 2Q points (rt1_line_fft), frequency parity e =
// blockIdx.x & 1: the representatives k of parity e use V[r], V[L - r] with
// r = L/2 +- k of the same parity, all in this workgroup's half.
template <int Q, bool CPLX>
__global__ __launch_bounds__(SplitShape<Q>::NT) void big_line_fft(const void* __restrict__ sino_v,
                                                                float2* __restrict__ yre, float2* __restrict__ yim,
                                                                const int* __restrict__ sig, int H, long long nrows,
                                                                int nz, float scale, const float2* __restrict__ h,
                                                                const float2* __restrict__ tw) {
  using Sh = SplitShape<Q>;
  __shared__ float2 lds[Sh::LDS];
  constexpr int L = 2 * Q, half = Q;
  const int e = (int)(blockIdx.x & 1);
  const long long q = blockIdx.x >> 1;
  long long ra, rb;
  bool va, vb;
  if (CPLX) {
    ra = q;
    rb = 0;
    va = ra < nrows;
    vb = false;
  } else {
    const long long th2 = q / nz;
    ra = th2 * 2 * nz + (q - th2 * nz);
    rb = ra + nz;
    va = ra < nrows;
    vb = rb < nrows;
  }
  const int sga = va ? sig[ra / nz] : 1, sgb = vb ? sig[rb / nz] : 1;
  auto line = [&](int j) {  // c1d . y at j (real rows: a + i b)
    float2 w;
    if (CPLX) {
      w = va ? ((const float2*)sino_v)[ra * L + j] : make_float2(0.f, 0.f);
    } else {
      const float* sv = (const float*)sino_v;
      w = make_float2(va ? sv[ra * L + j] : 0.f, vb ? sv[rb * L + j] : 0.f);
    }
    return (j & 1) ? w : make_float2(-w.x, -w.y);
  };
  const int t = threadIdx.x;
  float2 v[Sh::PT];
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) {
    const int j = t + m * Sh::TPC;
    const float2 lo = line(j), hi = line(j + Q);
    const float2 sum = e ? make_float2(lo.x - hi.x, lo.y - hi.y) : make_float2(lo.x + hi.x, lo.y + hi.y);
    v[m] = e ? cmul(sum, h[2 * j]) : sum;
  }
  fft_reg<Q, Sh::TPC, false>(v, lds, t, tw);
#pragma unroll
  for (int m = 0; m < Sh::PT; ++m) lds[lpad(t + m * Sh::TPC)] = v[m];  // V[2 m' + e] at m'
  __syncthreads();
  if (!va) return;  // no barrier below
  auto yhat = [&](int r) {  // c1d . V at r (r of parity e)
    const float2 w = lds[lpad(r >> 1)];
    return (r & 1) ? w : make_float2(-w.x, -w.y);
  };
  for (int k = 2 * t + e; k < H; k += 2 * Sh::TPC) {
    const float f = ((k == 0 || k == half) ? 0.5f : 1.0f) * scale;
    if (k == half) {  // from r = 0 for either sign (half is even: e = 0)
      const float2 v0 = yhat(0);
      if (CPLX) {
        yre[ra * H + k] = cscale(sga > 0 ? cconj(v0) : v0, f);
        yim[ra * H + k] = cscale(sga > 0 ? make_float2(v0.y, v0.x) : make_float2(v0.y, -v0.x), f);
      } else {
        yre[ra * H + k] = make_float2(v0.x * f, 0.f);
        if (vb) yre[rb * H + k] = make_float2(v0.y * f, 0.f);
      }
      continue;
    }
    {
      const int r = sga > 0 ? half + k : half - k;
      const float2 p = yhat(r), pm = yhat(L - r);
      if (CPLX) {
        yre[ra * H + k] = cscale(make_float2(0.5f * (p.x + pm.x), 0.5f * (p.y - pm.y)), f);
        yim[ra * H + k] = cscale(make_float2(0.5f * (p.y + pm.y), -0.5f * (p.x - pm.x)), f);
      } else {
        yre[ra * H + k] = make_float2(0.5f * (p.x + pm.x) * f, 0.5f * (p.y - pm.y) * f);
      }
    }
    if (!CPLX && vb) {
      const int r = sgb > 0 ? half + k : half - k;
      const float2 p = yhat(r), pm = yhat(L - r);
      yre[rb * H + k] = make_float2(0.5f * (p.y + pm.y) * f, -0.5f * (p.x - pm.x) * f);
    }
  }
}

// ---------------------------------------------------------------------------
// Host side: geometry, binning, launches
// ---------------------------------------------------------------------------

struct Rect {
  int c0, c1, r0, r1;  // inclusive, centred coordinates inside the half plane domain
};

static inline int wrapc(int v, int n) {  // into [-n, n)
  const int r = (v + n) % (2 * n);
  return (r < 0 ? r + 2 * n : r) - n;
}

// Splits rows [r0, r1] (centred, may leave [-n, n)) into wrapped pieces.
static void push_wrapped(std::vector<Rect>& out, int c0, int c1, int r0, int r1, int n) {
  if (c0 > c1) return;
  if (r0 < -n) {
    if (r1 < -n) {
      out.push_back({c0, c1, r0 + 2 * n, r1 + 2 * n});
    } else {
      out.push_back({c0, c1, r0 + 2 * n, n - 1});
      out.push_back({c0, c1, -n, r1});
    }
  } else if (r1 >= n) {
    if (r0 >= n) {
      out.push_back({c0, c1, r0 - 2 * n, r1 - 2 * n});
    } else {
      out.push_back({c0, c1, r0, n - 1});
      out.push_back({c0, c1, -n, r1 - 2 * n});
    }
  } else {
    out.push_back({c0, c1, r0, r1});
  }
}

struct HostGeom {
  int n, H, ntheta;
  std::vector<double> dx, dy;  // per theta: |cos|, -sigma sin
  void centre(int id, int& rx, int& ry) const {
    const int th = id / H, k = id - th * H;
    const double X = (2.0 * k) * dx[th], Y = (2.0 * k) * dy[th];
    rx = (int)std::rint(X);
    ry = (int)std::rint(Y);
  }
  // Output-plane rectangles touched by the taps of sample id (direct + folded).
  void rects(int id, std::vector<Rect>& out) const {
    out.clear();
    int rx, ry;
    centre(id, rx, ry);
    const int bx = rx - M, by = ry - M;
    push_wrapped(out, std::max(bx, 0), std::min(bx + 2 * M, n), by, by + 2 * M, n);
    if (bx <= 0) push_wrapped(out, 0, -bx, -(by + 2 * M), -by, n);
    if (bx + 2 * M >= n) push_wrapped(out, 2 * n - (bx + 2 * M), n, -(by + 2 * M), -by, n);
  }
};

struct Binning {
  int zp = 2;                          // z slices per patch element of the gather kernel
  std::vector<Item> items;             // root class (w > GatherLeaf::W) first, then the leaf class
  int n_root = 0;
  std::vector<int> samples;            // per item, in lane order
  std::vector<unsigned short> order;   // per item: lane of the t-th sample in ascending id order
};

// Lane order of one gather item (ids ascending): every quarter-wave of 16
// lanes takes samples from 16 different colours (the LDS bank group of each
// 16-byte tap read: (window base / 2) mod 16 for float2 patch elements, base
// mod 16 for float4 ones), drawing from the fullest
// colours first, so the tap reads are conflict-free wherever the colours allow.
static void gather_lanes(const HostGeom& G, int x0, int y0, int w, const int* ids, int count, Binning& B) {
  const int pp = patch_pitch(w);
  // 16-byte pieces: two float2 (zp = 1) or one float4 (zp = 2) patch elements
  const int shift = gather_zp(w, B.zp) == 1 ? 1 : 0;
  std::vector<int> bucket[16];
  for (int t = 0; t < count; ++t) {
    int rx, ry;
    G.centre(ids[t], rx, ry);
    bucket[(((ry - y0) * pp + (rx - x0)) >> shift) & 15].push_back(t);
  }
  size_t head[16] = {};
  std::vector<int> lane_of(count);
  const size_t start = B.samples.size();
  B.samples.resize(start + count);
  int lane = 0;
  while (lane < count) {
    const int size = std::min(16, count - lane);
    int col[16];
    for (int c = 0; c < 16; ++c) col[c] = c;
    std::sort(col, col + 16, [&](int u, int v) { return bucket[u].size() - head[u] > bucket[v].size() - head[v]; });
    int taken = 0;
    for (int i = 0; i < 16 && taken < size; ++i) {
      const int c = col[i];
      if (head[c] == bucket[c].size()) break;
      const int t = bucket[c][head[c]++];
      lane_of[t] = lane + taken++;
    }
    for (int i = 0; taken < size; i = (i + 1) % 16) {  // fewer colours than lanes left
      const int c = col[i];
      if (head[c] == bucket[c].size()) continue;
      const int t = bucket[c][head[c]++];
      lane_of[t] = lane + taken++;
    }
    lane += size;
  }
  for (int t = 0; t < count; ++t) {
    B.samples[start + lane_of[t]] = ids[t];
    B.order.push_back((unsigned short)lane_of[t]);
  }
}

static bool rect_hits(const Rect& r, int x0, int y0, int w, int h) {
  return r.c0 <= x0 + w - 1 && r.c1 >= x0 && r.r0 <= y0 + h - 1 && r.r1 >= y0;
}

// Gather: samples binned by their centre.
static void gather_node(const HostGeom& G, int x0, int y0, int sz, std::vector<int>& ids, Binning& B) {
  if (ids.empty()) return;
  if ((int)ids.size() <= SMAX || sz <= FLOOR) {
    for (size_t s = 0; s < ids.size(); s += SMAX) {
      Item it{};
      it.x0 = x0;
      it.y0 = y0;
      it.w = sz;
      it.h = sz;
      it.start = (int)B.samples.size();
      it.count = (int)std::min<size_t>(SMAX, ids.size() - s);
      it.slot = -1;
      gather_lanes(G, x0, y0, sz, ids.data() + s, it.count, B);
      B.items.push_back(it);
    }
    return;
  }
  const int h = sz / 2;
  std::vector<int> q[4];
  for (int id : ids) {
    int rx, ry;
    G.centre(id, rx, ry);
    const int qx = (rx - x0) >= h, qy = (ry - y0) >= h;
    q[qy * 2 + qx].push_back(id);
  }
  std::vector<int>().swap(ids);
  for (int c = 0; c < 4; ++c) gather_node(G, x0 + (c & 1) * h, y0 + (c >> 1) * h, h, q[c], B);
}

static void build_gather(const HostGeom& G, int root, Binning& B) {
  const int n = G.n;
  const int nrx = n / root + 1, nry = 2 * n / root + 1;
  std::vector<std::vector<int>> roots((size_t)nrx * nry);
  const int total = G.ntheta * G.H;
  for (int id = 0; id < total; ++id) {
    int rx, ry;
    G.centre(id, rx, ry);
    roots[(size_t)((ry + n) / root) * nrx + rx / root].push_back(id);
  }
  for (int ty = 0; ty < nry; ++ty)
    for (int tx = 0; tx < nrx; ++tx) gather_node(G, tx * root, ty * root - n, root, roots[(size_t)ty * nrx + tx], B);
  // neighbouring items stay neighbours inside each class (XCD grouping)
  const auto mid = std::stable_partition(B.items.begin(), B.items.end(),
                                         [](const Item& it) { return it.w > GatherLeaf::W; });
  B.n_root = (int)(mid - B.items.begin());
}

// ---- scatter lists (host, at plan time) -------------------------------------

struct SBinning {
  std::vector<SItem> items;
  std::vector<unsigned char> perm;  // per item, SNP bytes: the tile point lane p owns
  std::vector<int> entries;
  std::vector<unsigned short> lists;
  std::vector<SplitTile> splits;
  int nslots = 0;
};

static inline int pmod(int v, int m) {
  const int r = v % m;
  return r < 0 ? r + m : r;
}

// Step order of the taps (row | tap << 9) of one 16-lane group (one tile row
// of points).  The kernel reads each tap's 80-byte entry row as 16-byte
// pieces, which the LDS serves 16 lanes per cycle; two rows collide iff they
// are equal mod 16 (their colour) and differ.  Greedy per step: lanes with the
// most taps left choose first, a tap whose row another lane already reads
// (broadcast) before a tap of a free colour, free colours by descending
// demand.  Lanes without a usable tap idle on a row read in the same step,
// with tap code SL_IDLE_TAP (weight 0).  If the steps would exceed SL_MAX the
// lists are used in their own order (bank conflicts, never more steps than
// the longest list, which chunking keeps <= SL_CHUNK).  Returns the steps.
static int schedule_group(const std::vector<unsigned short>* L, std::vector<unsigned short>* S) {
  constexpr int G = SGROUP;
  unsigned short tap[G][SL_MAX];
  int beg[G][G], cur[G][G], rem[G] = {}, demand[G] = {};
  int longest = 0, total = 0;
  for (int l = 0; l < G; ++l) {
    const int nt = (int)L[l].size();
    longest = std::max(longest, nt);
    total += nt;
    rem[l] = nt;
    int k = 0;
    for (int c = 0; c < G; ++c) {  // counting sort by colour
      beg[l][c] = cur[l][c] = k;
      for (unsigned short u : L[l])
        if ((u & 15) == c) tap[l][k++] = u;
      demand[c] += k - beg[l][c];
    }
  }
  auto end_of = [&](int l, int c) { return c + 1 < G ? beg[l][c + 1] : (int)L[l].size(); };
  for (int l = 0; l < G; ++l) S[l].clear();
  int steps = 0;
  bool ok = true;
  while (total > 0) {
    if (steps == SL_MAX) {
      ok = false;
      break;
    }
    int row_of[G];
    for (int c = 0; c < G; ++c) row_of[c] = -1;
    int order[G];
    for (int l = 0; l < G; ++l) order[l] = l;
    std::sort(order, order + G, [&](int u, int v) { return rem[u] > rem[v]; });
    int pick[G];
    int any_row = -1;
    for (int oi = 0; oi < G; ++oi) {
      const int l = order[oi];
      pick[l] = -1;
      if (rem[l] == 0) continue;
      int best = -1, best_c = -1;
      for (int c = 0; c < G && best < 0; ++c) {  // broadcast: a row already read this step
        if (row_of[c] < 0) continue;
        for (int k = cur[l][c]; k < end_of(l, c); ++k)
          if ((tap[l][k] & 0x1FF) == row_of[c]) {
            best = k;
            best_c = c;
            break;
          }
      }
      if (best < 0) {  // the free colour in most demand
        for (int c = 0; c < G; ++c)
          if (row_of[c] < 0 && cur[l][c] < end_of(l, c) && (best_c < 0 || demand[c] > demand[best_c])) best_c = c;
        if (best_c >= 0) best = cur[l][best_c];
      }
      if (best < 0) continue;
      const unsigned short u = tap[l][best];
      std::swap(tap[l][best], tap[l][cur[l][best_c]]);  // consume from the colour range
      ++cur[l][best_c];
      --rem[l];
      --total;
      --demand[best_c];
      row_of[best_c] = u & 0x1FF;
      any_row = u & 0x1FF;
      pick[l] = u;
    }
    for (int l = 0; l < G; ++l)
      S[l].push_back(pick[l] >= 0 ? (unsigned short)pick[l] : (unsigned short)(any_row | (SL_IDLE_TAP << 9)));
    ++steps;
  }
  if (ok) return steps;
  for (int l = 0; l < G; ++l) S[l] = L[l];  // conflicting but no longer than the longest list
  return longest;
}

// Items of one 16 x 16 output tile (tx, ty) from the entry codes touching it.
static void scatter_tile_items(const HostGeom& G, int tx, int ty, const std::vector<int>& codes, SBinning& B) {
  const int n = G.n, twon = 2 * n;
  const int x0 = tx * STS, y0 = ty * STS - n;
  const int ne = (int)codes.size();
  if (ne == 0) {  // no sample touches the tile: the item writes zeros
    for (int pt = 0; pt < SNP; ++pt) B.perm.push_back((unsigned char)pt);
    SItem it{};
    it.loff = (long long)B.lists.size();
    it.eoff = (int)B.entries.size();
    it.x0 = x0;
    it.y0 = y0;
    it.slot = -1;
    B.items.push_back(it);
    return;
  }
  // taps of every entry that land in this tile: (point, i0, i1)
  std::vector<int> tap_pt;
  std::vector<unsigned short> tap_code;
  std::vector<int> tap_off(ne + 1, 0);
  for (int e = 0; e < ne; ++e) {
    int rx, ry;
    G.centre(codes[e] >> 1, rx, ry);
    const int cx = (codes[e] & 1) ? -rx : rx, cy = (codes[e] & 1) ? -ry : ry;
    for (int i1 = 0; i1 < TAPS; ++i1) {
      const int lr = wrapc(cy - M + i1, n) - y0;
      if (lr < 0 || lr >= STS) continue;
      for (int i0 = 0; i0 < TAPS; ++i0) {
        const int raw = pmod(cx - M + i0, twon);
        const int lc = raw - x0;
        if (raw > n || lc < 0 || lc >= STS) continue;
        tap_pt.push_back(lr * STS + lc);
        tap_code.push_back((unsigned short)((TAPS * i1 + i0) << 9));
      }
    }
    tap_off[e + 1] = (int)tap_pt.size();
  }
  // greedy chunks of entries: <= SE_MAX entries, <= SL_CHUNK taps per point
  std::vector<int> cnt(SNP, 0);
  std::vector<std::pair<int, int>> chunks;
  int e0 = 0;
  for (int e = 0; e < ne; ++e) {
    bool fits = (e - e0) < SE_MAX;
    for (int t = tap_off[e]; fits && t < tap_off[e + 1]; ++t)
      if (cnt[tap_pt[t]] + 1 > SL_CHUNK) fits = false;
    if (!fits) {
      chunks.push_back({e0, e});
      std::fill(cnt.begin(), cnt.end(), 0);
      e0 = e;
    }
    for (int t = tap_off[e]; t < tap_off[e + 1]; ++t) cnt[tap_pt[t]]++;
  }
  chunks.push_back({e0, ne});
  const bool split = chunks.size() > 1;
  int slot0 = -1;
  if (split) {
    slot0 = B.nslots;
    B.nslots += (int)chunks.size();
    B.splits.push_back({x0, y0, slot0, (int)chunks.size()});
  }
  std::vector<std::vector<unsigned short>> L(SNP), LP(SNP), S(SNP);
  for (size_t c = 0; c < chunks.size(); ++c) {
    const int a = chunks[c].first, b = chunks[c].second;
    for (auto& l : L) l.clear();
    for (auto& l : LP) l.clear();
    for (int e = a; e < b; ++e)
      for (int t = tap_off[e]; t < tap_off[e + 1]; ++t) L[tap_pt[t]].push_back((unsigned short)((e - a) | tap_code[t]));
    // lanes take the points in order of decreasing list length, so each wave (and each 16-lane
    // group) holds lists of similar length; then the step order of each group
    int perm[SNP];
    for (int pt = 0; pt < SNP; ++pt) perm[pt] = pt;
    std::stable_sort(perm, perm + SNP, [&](int u, int v) { return L[u].size() > L[v].size(); });
    for (int l = 0; l < SNP; ++l) {
      LP[l].swap(L[perm[l]]);
      B.perm.push_back((unsigned char)perm[l]);
    }
    int glen[SNP / SGROUP];
    for (int g = 0; g < SNP / SGROUP; ++g) glen[g] = schedule_group(&LP[g * SGROUP], &S[g * SGROUP]);
    SItem it{};
    int nl = 0;
    for (int w = 0; w < SWAVES; ++w) {
      int m = 0;
      for (int g = w * 64 / SGROUP; g < (w + 1) * 64 / SGROUP; ++g) m = std::max(m, glen[g]);
      m = (m + 1) & ~1;  // slots are read in pairs
      nl = std::max(nl, m);
      it.wnp |= (unsigned)(m / 2) << (8 * w);
    }
    it.x0 = x0;
    it.y0 = y0;
    it.ne = b - a;
    it.eoff = (int)B.entries.size();
    it.nl = nl;
    it.loff = (long long)B.lists.size();
    it.slot = split ? slot0 + (int)c : -1;
    B.entries.insert(B.entries.end(), codes.begin() + a, codes.begin() + b);
    const unsigned short idle = (unsigned short)(SL_IDLE_TAP << 9);  // row 0, weight 0: a broadcast read
    for (int j = 0; j < nl; j += 2)
      for (int pt = 0; pt < SNP; ++pt) {
        B.lists.push_back(j < (int)S[pt].size() ? S[pt][j] : idle);
        B.lists.push_back(j + 1 < (int)S[pt].size() ? S[pt][j + 1] : idle);
      }
    B.items.push_back(it);
  }
}

// Entry codes ((sample id << 1) | mirror) of every 16 x 16 output tile that
// one of the entry's counted taps (raw column <= n) lands in.
static std::vector<std::vector<int>> scatter_tile_codes(const HostGeom& G) {
  const int n = G.n;
  const int ntx = n / STS + 1, nty = 2 * n / STS;
  std::vector<std::vector<int>> tiles((size_t)ntx * nty);
  const int total = G.ntheta * G.H;
  std::vector<Rect> rects;
  for (int id = 0; id < total; ++id) {
    int rx, ry;
    G.centre(id, rx, ry);
    for (int mir = 0; mir < 2; ++mir) {
      const int cx = mir ? -rx : rx, cy = mir ? -ry : ry;
      int c0 = n + 1, c1 = -1;
      for (int i = 0; i < TAPS; ++i) {
        const int raw = pmod(cx - M + i, 2 * n);
        if (raw <= n) {
          c0 = std::min(c0, raw);
          c1 = std::max(c1, raw);
        }
      }
      if (c1 < 0) continue;
      rects.clear();
      push_wrapped(rects, c0, c1, cy - M, cy + M, n);
      for (const Rect& r : rects)
        for (int ty = (r.r0 + n) / STS; ty <= (r.r1 + n) / STS; ++ty)
          for (int tx = r.c0 / STS; tx <= r.c1 / STS; ++tx) tiles[(size_t)ty * ntx + tx].push_back((id << 1) | mir);
    }
  }
  return tiles;
}

static void build_scatter_lists(const HostGeom& G, SBinning& B) {
  const int ntx = G.n / STS + 1;
  const std::vector<std::vector<int>> tiles = scatter_tile_codes(G);
  const size_t ntiles = tiles.size();
  unsigned nw = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<SBinning> part(nw);
  std::vector<std::thread> pool;
  for (unsigned w = 0; w < nw; ++w)
    pool.emplace_back([&, w]() {
      const size_t t0 = ntiles * w / nw, t1 = ntiles * (w + 1) / nw;
      for (size_t t = t0; t < t1; ++t)
        scatter_tile_items(G, (int)(t % ntx), (int)(t / ntx), tiles[t], part[w]);
    });
  for (auto& th : pool) th.join();
  for (auto& P : part) {
    const int eoff = (int)B.entries.size(), soff = B.nslots;
    const long long loff = (long long)B.lists.size();
    for (SItem it : P.items) {
      it.eoff += eoff;
      it.loff += loff;
      if (it.slot >= 0) it.slot += soff;
      B.items.push_back(it);
    }
    for (SplitTile s : P.splits) {
      s.slot0 += soff;
      B.splits.push_back(s);
    }
    B.entries.insert(B.entries.end(), P.entries.begin(), P.entries.end());
    B.perm.insert(B.perm.end(), P.perm.begin(), P.perm.end());
    B.lists.insert(B.lists.end(), P.lists.begin(), P.lists.end());
    B.nslots += P.nslots;
  }
}

// ---- scatter patch runs (host, at plan time) --------------------------------

struct MBinning {
  std::vector<MItem> items;
  std::vector<int> entries;            // per item, in sample order
  std::vector<unsigned short> rows;    // per entry: its LDS row (place in anchor-class order)
  std::vector<SplitTile> splits;
  int nslots = 0;
};

// Window base (bx, by) of an entry relative to the tile origin (x0, y0): the
// tap (i0, i1) lands on tile point (bx + i0, by + i1) when that lies in the
// tile (columns and centred rows wrap mod 2n; the representative in
// [-8, 2n - 8) is the one that can reach the tile).
static inline void window_base(const HostGeom& G, int code, int x0, int y0, int& bx, int& by) {
  const int twon = 2 * G.n;
  int rx, ry;
  G.centre(code >> 1, rx, ry);
  const int cx = (code & 1) ? -rx : rx, cy = (code & 1) ? -ry : ry;
  bx = pmod(cx - M - x0 + 8, twon) - 8;
  by = pmod(cy - M - y0 + 8, twon) - 8;
}

// Items of one 16 x 16 output tile for scatter_mfma: the LDS rows of the
// entries in anchor-class order ((by + 8) / 4) * 6 + (bx + 8) / 4; more than
// ME_MAX entries are split into near-equal runs of consecutive rows whose
// partial tiles scatter_reduce sums.  Each item lists its entries in sample
// order with their rows, so the staging loads of neighbouring lanes hit
// neighbouring samples.
static void scatter_tile_items_mfma(const HostGeom& G, int tx, int ty, const std::vector<int>& codes, MBinning& B) {
  const int x0 = tx * STS, y0 = ty * STS - G.n;
  MItem base{};
  base.x0 = x0;
  base.y0 = y0;
  base.slot = -1;
  base.eoff = (int)B.entries.size();
  std::vector<std::pair<int, int>> ce;  // (class, code)
  ce.reserve(codes.size());
  for (int code : codes) {
    int bx, by;
    window_base(G, code, x0, y0, bx, by);
    if (bx < -8 || bx >= STS || by < -8 || by >= STS) continue;  // no tap in the tile
    ce.push_back({((by + 8) >> 2) * 6 + ((bx + 8) >> 2), code});
  }
  const int ne = (int)ce.size();
  if (ne == 0) {  // no sample touches the tile: the item writes zeros
    B.items.push_back(base);
    return;
  }
  std::stable_sort(ce.begin(), ce.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
    return a.first < b.first;
  });
  const int nch = (ne + ME_MAX - 1) / ME_MAX;
  int slot0 = -1;
  if (nch > 1) {
    slot0 = B.nslots;
    B.nslots += nch;
    B.splits.push_back({x0, y0, slot0, nch});
  }
  for (int c = 0; c < nch; ++c) {
    const int a = (int)((long long)ne * c / nch), b = (int)((long long)ne * (c + 1) / nch);
    MItem it = base;
    it.eoff = (int)B.entries.size();
    it.ne = b - a;
    it.slot = nch > 1 ? slot0 + c : -1;
    int cnt[MCLS] = {};
    std::vector<std::pair<int, int>> order;  // (code, row): staged in sample order
    for (int e = a; e < b; ++e) {
      ++cnt[ce[e].first];
      order.push_back({ce[e].second, e - a});
    }
    std::sort(order.begin(), order.end());
    for (const auto& o : order) {
      B.entries.push_back(o.first);
      B.rows.push_back((unsigned short)o.second);
    }
    int acc = 0;
    for (int k = 0; k < MCLS; ++k) {
      it.cls[k] = (unsigned short)acc;
      acc += cnt[k];
    }
    for (int k = MCLS; k < 40; ++k) it.cls[k] = (unsigned short)acc;
    B.items.push_back(it);
  }
}

static void build_scatter_patches(const HostGeom& G, MBinning& B) {
  const int ntx = G.n / STS + 1;
  const std::vector<std::vector<int>> tiles = scatter_tile_codes(G);
  const size_t ntiles = tiles.size();
  unsigned nw = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<MBinning> part(nw);
  std::vector<std::thread> pool;
  for (unsigned w = 0; w < nw; ++w)
    pool.emplace_back([&, w]() {
      const size_t t0 = ntiles * w / nw, t1 = ntiles * (w + 1) / nw;
      for (size_t t = t0; t < t1; ++t)
        scatter_tile_items_mfma(G, (int)(t % ntx), (int)(t / ntx), tiles[t], part[w]);
    });
  for (auto& th : pool) th.join();
  for (auto& P : part) {
    const int eoff = (int)B.entries.size(), soff = B.nslots;
    for (MItem it : P.items) {
      it.eoff += eoff;
      if (it.slot >= 0) it.slot += soff;
      B.items.push_back(it);
    }
    for (SplitTile s : P.splits) {
      s.slot0 += soff;
      B.splits.push_back(s);
    }
    B.entries.insert(B.entries.end(), P.entries.begin(), P.entries.end());
    B.rows.insert(B.rows.end(), P.rows.begin(), P.rows.end());
    B.nslots += P.nslots;
  }
}

// ---- spectrum support (host, at plan time) ----------------------------------
// Per 4-column block of the half spectrum, the centred row range [lo, hi) that
// carries data between the column pass and the polar pass (hull of the rows
// touched in the block's 4 columns; empty: lo = hi).

static void support_add(std::vector<int2>& s, int col, int lo, int hi) {
  int2& r = s[col >> 2];
  if (r.x == r.y) {
    r = make_int2(lo, hi);
  } else {
    r.x = std::min(r.x, lo);
    r.y = std::max(r.y, hi);
  }
}

// Rows the gather patches read: patch rows [y0 - M, y0 + h + M) in columns
// [x0 - M, x0 + w + M); a column k0 > n is read as row -l of column 2n - k0.
// A patch reaching past row -n or n (small n: rows wrap mod 2n) takes every row.
static std::vector<int2> gather_support(const Binning& B, int n) {
  const int twon = 2 * n;
  std::vector<int2> s((size_t)(n / 4 + 1), make_int2(0, 0));
  for (const Item& it : B.items) {
    int lo = it.y0 - M, hi = it.y0 + it.h + M;
    const bool all = lo <= -n || hi > n;
    if (all) {
      lo = -n;
      hi = n;
    }
    for (int c = it.x0 - M; c < it.x0 + it.w + M; ++c) {
      const int k0 = wrap_mod(c, twon);
      if (k0 <= n)
        support_add(s, k0, lo, hi);
      else if (all)
        support_add(s, twon - k0, -n, n);
      else
        support_add(s, twon - k0, 1 - hi, 1 - lo);
    }
  }
  return s;
}

// Rows the scatter writes.  Per tile column, the tiles between the lowest and
// the highest tile that has entries stay work items (empty ones write zeros);
// empty tiles outside that range are dropped, and the column pass reads the
// range only.  Returns the support; compacts the items (and the lane
// permutations of list items).
template <class ItemT>
static std::vector<int2> trim_scatter_items(std::vector<ItemT>& all, std::vector<unsigned char>* perm, int n) {
  const int ntx = n / STS + 1;
  std::vector<int> tmin(ntx, INT_MAX), tmax(ntx, INT_MIN);
  for (const ItemT& it : all)
    if (it.ne > 0) {
      const int tx = it.x0 / STS, ty = (it.y0 + n) / STS;
      tmin[tx] = std::min(tmin[tx], ty);
      tmax[tx] = std::max(tmax[tx], ty);
    }
  std::vector<ItemT> items;
  std::vector<unsigned char> kept;
  for (size_t i = 0; i < all.size(); ++i) {
    const ItemT& it = all[i];
    const int tx = it.x0 / STS, ty = (it.y0 + n) / STS;
    if (it.ne == 0 && (ty < tmin[tx] || ty > tmax[tx])) continue;
    items.push_back(it);
    if (perm) kept.insert(kept.end(), perm->begin() + (long long)i * SNP, perm->begin() + (long long)(i + 1) * SNP);
  }
  all.swap(items);
  if (perm) perm->swap(kept);
  std::vector<int2> s((size_t)(n / 4 + 1), make_int2(0, 0));
  for (int tx = 0; tx < ntx; ++tx)
    if (tmin[tx] <= tmax[tx])
      for (int c = tx * STS; c < std::min(tx * STS + STS, n + 1); c += 4)
        support_add(s, c, tmin[tx] * STS - n, (tmax[tx] + 1) * STS - n);
  return s;
}
static std::vector<int2> trim_scatter(SBinning& B, int n) { return trim_scatter_items(B.items, &B.perm, n); }
static std::vector<int2> trim_scatter(MBinning& B, int n) { return trim_scatter_items(B.items, nullptr, n); }

// Detector widths: 2^k in [16, 32768] or 3 * 2^k in [96, 24576].  Up to 8192 the 2n-point passes take one
// workgroup per transform, up to 16384 two n-point halves, above that four n/2-point quarters (big_*).
static bool radon_width_ok(int n) {
  if (!((is_pow2(n) && n >= 16) || (is_fft_len(n) && n >= 96))) return false;
  return n <= kMaxFFT || (n <= 2 * kMaxFFT && fft_len_ok(n / 2) && pt_of(n / 2) % 4 == 0);
}

template <typename T>
static int upload(T** dst, const std::vector<T>& src, size_t& bytes) {
  *dst = nullptr;
  if (src.empty()) return MM_OK;
  const size_t nb = src.size() * sizeof(T);
  hipError_t e = hipMalloc(dst, nb);
  if (e != hipSuccess) return fail(MM_ERR_OOM, std::string("radon plan upload: ") + hipGetErrorString(e));
  bytes += nb;
  MM_HIP_TRY(hipMemcpy(*dst, src.data(), nb, hipMemcpyHostToDevice));
  return MM_OK;
}

}  // namespace radon
}  // namespace mm

using namespace mm;
using namespace mm::radon;

struct mm_radon {
  int dev = 0;
  int n = 0, nz_max = 0, ntheta = 0, H = 0, NCB = 0;
  long long zs = 0;  // elements per z plane of the half spectrum (NCB blocks x 2n rows x 4)
  unsigned flags = 0;
  double mu = 0;
  float a = 0, C0g = 0, C0s = 0;
  float2* fde = nullptr;   // nz_max x zs, block layout (hs_idx): R half spectrum / RT channel "re" half plane
  float2* fde2 = nullptr;  // RT channel "im" (complex sinograms, allocated on first use)
  float2* gh = nullptr;    // ntheta x nz_max x H: representatives
  float2* gh2 = nullptr;   // RT channel "im" representatives
  float2* partial = nullptr;
  float* phx = nullptr;
  float* phy = nullptr;
  double2* dirs = nullptr;
  int* sig = nullptr;
  Item* gitems = nullptr;
  int* gsamples = nullptr;
  unsigned short* gorder = nullptr;
  int n_gitems = 0;
  int n_groot = 0;  // gather items of the root class (the first n_groot of gitems)
  int2* gsupp = nullptr;  // per 4-column block: centred rows the gather reads
  int2* ssupp = nullptr;  // per 4-column block: centred rows the scatter writes
  float2* hsplit = nullptr;  // exp(-i pi j / n), j < n: the 2n-point passes run as two n-point halves (split)
  bool split = false;        // 2n > kMaxFFT
  bool big = false;          // n > kMaxFFT: four-way 2n passes through bbuf (big_*)
  float2* bbuf = nullptr;    // nz_max x bzs: the n object rows of the half spectrum (big)
  float2* bbuf2 = nullptr;   // second channel (complex sinograms, allocated on first use)
  long long bzs = 0;         // elements per z plane of bbuf
  int gzp = 2;  // z slices per gather patch element (the lane colouring depends on it)
  int groot = ROOT;                  // largest gather quadtree tile (64, or 32 with MM_RADON_GATHER_ROOT32)
  std::vector<double> theta_h;       // the angles, for rebuilding the gather tables (radon_regather)
  size_t gbytes = 0;                 // device bytes of the gather tables (part of bytes)
  int tune[MM_RADON_TUNE_NKNOBS] = {0, 8, 32, 0, ROOT, 2};  // knobs (mm_radon_tune): XCD groups of 8 gather, 32 scatter items
  // adjoint scatter: patch runs on the matrix cores (FP32, default) or per-point lists (FP64 / Kahan tiles)
  bool mfma = true;
  MItem* mitems = nullptr;
  int* mentries = nullptr;
  unsigned short* mrows = nullptr;
  int n_mitems = 0;
  SItem* sitems = nullptr;
  unsigned char* sperm = nullptr;
  int* sentries = nullptr;
  unsigned short* slists = nullptr;
  int n_sitems = 0;
  SplitTile* splits = nullptr;
  int n_splits = 0, n_slots = 0;
  const float2* tw = nullptr;
  size_t bytes = 0;
  bool own_ws = true;     // fde / gh allocated by the plan (false: a batch's shared workspace)
  bool fwd_only = false;  // no adjoint scatter tables (mm_radon_adj unsupported)
  bool profile = false;
  bool ran[MM_RADON_NSTAGES] = {};
  hipEvent_t ev[MM_RADON_NSTAGES + 1][2] = {};
};

// Records the start (which = 0) or end (which = 1) of `stage` when profiling.
static int mark(mm_radon* p, int stage, int which, hipStream_t st) {
  if (!p->profile) return MM_OK;
  MM_HIP_TRY(hipEventRecord(p->ev[stage][which], st));
  if (which == 1) p->ran[stage] = true;
  return MM_OK;
}

static void radon_free(mm_radon* p) {
  if (!p) return;
  for (auto& e : p->ev)
    for (auto& x : e)
      if (x) (void)hipEventDestroy(x);
  if (!p->own_ws) p->fde = p->gh = nullptr;
  void* ptrs[] = {p->fde, p->fde2, p->gh, p->gh2, p->partial, p->bbuf, p->bbuf2, p->phx, p->phy, p->dirs, p->sig,
                  p->gitems, p->gsamples, p->gorder, p->sitems, p->sperm, p->sentries, p->slists,
                  p->mitems, p->mentries, p->mrows, p->splits, p->gsupp, p->ssupp, p->hsplit};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete p;
}

static int dmalloc(void** p, size_t nb, size_t& bytes, const char* what) {
  hipError_t e = hipMalloc(p, nb);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(MM_ERR_OOM, std::string("radon ") + what + " (" + std::to_string(nb) + " bytes): " + hipGetErrorString(e));
  }
  bytes += nb;
  return MM_OK;
}

// ---- launch helpers (templated on the FFT length) --------------------------

template <int L>
static void launch_r1(const mm_radon* p, const float* obj, int nz, hipStream_t st) {
  using Sh = R1Shape<L>;
  const long long units = (long long)nz * (p->n / 2);
  hipLaunchKernelGGL(r1_rows_rfft<L>, dim3((unsigned)((units + Sh::C - 1) / Sh::C)), dim3(Sh::NT), 0, st, obj,
                     p->fde, p->phx, p->phy, p->n, nz, p->zs, p->tw + tw_offset(L));
}

// Workgroups of fewer than four columns (long transforms, or the split knob)
// read part of every 32-byte row piece of a block: the 4 / C workgroups of a
// block are kept adjacent in the XCD-grouped order (xg = 4 / C, the number of
// workgroups per z plane rounded up to a multiple of it, the spare ones idle)
// so that they meet in one L2 and each 128-byte line is fetched once.
template <int L, bool INV, int BATCH>
static void launch_col_shape(const mm_radon* p, float2* buf, int nz, hipStream_t st) {
  using Sh = ColShapeB<L, BATCH>;
  constexpr int G = Sh::C < 4 ? 4 / Sh::C : 1;
  const int ncols = p->n + 1;
  const int2* supp = INV ? p->ssupp : p->gsupp;
  const int gx = ((ncols + Sh::C - 1) / Sh::C + G - 1) / G * G;
  hipLaunchKernelGGL((col_fft<L, INV, BATCH>), dim3((unsigned)gx, (unsigned)nz), dim3(Sh::NT), 0, st, buf, ncols,
                     p->zs, supp, p->tw + tw_offset(L), G);
}

template <int L, bool INV>
static void launch_col(const mm_radon* p, float2* buf, int nz, hipStream_t st) {
  if constexpr (ColShapeB<L, kRowBatch>::C < ColShapeB<L, kBatchElems>::C) {
    if (p->tune[MM_RADON_TUNE_COL_SPLIT]) {
      launch_col_shape<L, INV, kRowBatch>(p, buf, nz, st);
      return;
    }
  }
  launch_col_shape<L, INV, kBatchElems>(p, buf, nz, st);
}

template <int L>
static void launch_r4(const mm_radon* p, float* sino, int nz, hipStream_t st) {
  using Sh = LineShape<L>;
  const long long rows = (long long)p->ntheta * nz, pairs = (long long)((p->ntheta + 1) / 2) * nz;  // two rows per transform
  hipLaunchKernelGGL(r4_line_ifft<L>, dim3((unsigned)((pairs + Sh::C - 1) / Sh::C)), dim3(Sh::NT), 0, st, p->gh,
                     sino, p->sig, p->H, rows, nz, 1.0f / (4.0f * p->n), p->tw + tw_offset(L));
}

template <int L>
static void launch_rt1(const mm_radon* p, const void* sino, bool cplx, int nz, hipStream_t st) {
  using Sh = LineShape<L>;
  const long long rows = (long long)p->ntheta * nz;
  if (cplx)  // one complex row per transform
    hipLaunchKernelGGL((rt1_line_fft<L, true>), dim3((unsigned)((rows + Sh::C - 1) / Sh::C)), dim3(Sh::NT), 0, st,
                       sino, p->gh, p->gh2, p->sig, p->H, rows, nz, p->C0s, p->tw + tw_offset(L));
  else  // two real rows per transform
    hipLaunchKernelGGL((rt1_line_fft<L, false>), dim3((unsigned)((((p->ntheta + 1) / 2) * (long long)nz + Sh::C - 1) / Sh::C)), dim3(Sh::NT),
                       0, st, sino, p->gh, p->gh2, p->sig, p->H, rows, nz, p->C0s, p->tw + tw_offset(L));
}

// the fused kernel needs one thread layout for the filter (2L) and the line (L) transforms
__host__ __device__ constexpr bool fbp_fusable(int L) { return 2 * L <= kMaxFFT && pt_of(L) == pt_of(2 * L); }

template <int L>
static void launch_rt1_fbp(const mm_radon* p, const float* sino, const float* G, const float* A, const float* B,
                           int nz, hipStream_t st) {
  if constexpr (fbp_fusable(L)) {
    using Sh = typename FbpLineShape<L>::S1;
    const long long rows = (long long)p->ntheta * nz;
    hipLaunchKernelGGL(rt1_fbp_line_fft<L>, dim3((unsigned)((rows + Sh::C - 1) / Sh::C)), dim3(Sh::NT), 0, st, sino,
                       p->gh, p->sig, p->H, rows, nz, p->C0s, G, A, B, p->tw + tw_offset(L), p->tw + tw_offset(2 * L));
  }
}

template <int L>
static void launch_rt4(const mm_radon* p, void* out, bool cplx, int nz, hipStream_t st) {
  using Sh = C2rShape<L>;
  const long long units = cplx ? (long long)nz * p->n : (long long)nz * (p->n / 2);
  const dim3 grid((unsigned)((units + Sh::C - 1) / Sh::C));
  const float scale = (float)(1.0 / (4.0 * (double)p->n * p->n));  // column and row inverse passes, 1/(2n) each
  if (cplx)
    hipLaunchKernelGGL((rt4_rows_c2r<L, true>), grid, dim3(Sh::NT), 0, st, p->fde, p->fde2, out, p->phx, p->phy,
                       p->n, nz, p->zs, scale, p->tw + tw_offset(L));
  else
    hipLaunchKernelGGL((rt4_rows_c2r<L, false>), grid, dim3(Sh::NT), 0, st, p->fde, p->fde2, out, p->phx, p->phy,
                       p->n, nz, p->zs, scale, p->tw + tw_offset(L));
}

// 2n-point passes as two n-point halves (n > kMaxFFT / 2): one workgroup per row pair and parity, per
// column, per C2R unit
template <int L>
static void launch_r1_split(const mm_radon* p, const float* obj, int nz, hipStream_t st) {
  const long long wgs = 2LL * nz * (L / 2);
  hipLaunchKernelGGL(r1_rows_split<L>, dim3((unsigned)wgs), dim3(SplitShape<L>::NT), 0, st, obj, p->fde, p->phx,
                     p->phy, p->hsplit, nz, p->zs, p->tw + tw_offset(L));
}

template <int L, bool INV>
static void launch_col_split(const mm_radon* p, float2* buf, int nz, hipStream_t st) {
  const int ncols = L + 1;
  hipLaunchKernelGGL((col_fft_split<L, INV>), dim3((unsigned)ncols, (unsigned)nz), dim3(SplitShape<L>::NT), 0, st,
                     buf, ncols, p->zs, INV ? p->ssupp : p->gsupp, p->hsplit, p->tw + tw_offset(L));
}

template <int L>
static void launch_rt4_split(const mm_radon* p, void* out, bool cplx, int nz, hipStream_t st) {
  const long long units = cplx ? (long long)nz * L : (long long)nz * (L / 2);
  const float scale = (float)(1.0 / (4.0 * (double)L * L));  // column and row inverse passes, 1/(2n) each
  if (cplx)
    hipLaunchKernelGGL((rt4_rows_c2r_split<L, true>), dim3((unsigned)units), dim3(SplitShape<L>::NT), 0, st, p->fde,
                       p->fde2, out, p->phx, p->phy, p->hsplit, nz, p->zs, scale, p->tw + tw_offset(L));
  else
    hipLaunchKernelGGL((rt4_rows_c2r_split<L, false>), dim3((unsigned)units), dim3(SplitShape<L>::NT), 0, st, p->fde,
                       p->fde2, out, p->phx, p->phy, p->hsplit, nz, p->zs, scale, p->tw + tw_offset(L));
}

// Runs CALL with constexpr int LL = n for the detector widths whose 2n-point passes are split.
#define MM_DISPATCH_SPLIT(N_RUNTIME, CALL)                                                            \
  switch (N_RUNTIME) {                                                                                \
    case 12288: { constexpr int LL = 12288; CALL; } break;                                            \
    case 16384: { constexpr int LL = 16384; CALL; } break;                                            \
    default: return ::mm::fail(MM_ERR_UNSUPPORTED, "split 2n transforms not built for n=" + std::to_string(N_RUNTIME)); \
  }

// Runs CALL with constexpr int QQ = n / 2 for the detector widths above 16384 (big_*).
#define MM_DISPATCH_BIG(N_RUNTIME, CALL)                                                              \
  switch ((N_RUNTIME) / 2) {                                                                          \
    case 12288: { constexpr int QQ = 12288; CALL; } break;                                            \
    case 16384: { constexpr int QQ = 16384; CALL; } break;                                            \
    default: return ::mm::fail(MM_ERR_UNSUPPORTED, "four-way transforms not built for n=" + std::to_string(N_RUNTIME)); \
  }

template <int Q>
static void launch_big_r1(const mm_radon* p, const float* obj, int nz, hipStream_t st) {
  hipLaunchKernelGGL(big_r1_rows<Q>, dim3((unsigned)(4LL * nz * 2 * Q)), dim3(SplitShape<Q>::NT), 0, st, obj, p->bbuf,
                     p->phx, p->phy, p->hsplit, p->bzs, p->tw + tw_offset(Q));
}
template <int Q, bool INV>
static void launch_big_col(const mm_radon* p, float2* fde, float2* B, int nz, hipStream_t st) {
  const int ncols = 2 * Q + 1;
  hipLaunchKernelGGL((big_col<Q, INV>), dim3((unsigned)(4 * ncols), (unsigned)nz), dim3(SplitShape<Q>::NT), 0, st, fde,
                     B, ncols, p->zs, p->bzs, INV ? p->ssupp : p->gsupp, p->hsplit, p->tw + tw_offset(Q));
}
template <int Q>
static void launch_big_c2r(const mm_radon* p, void* out, bool cplx, int nz, hipStream_t st) {
  const long long units = cplx ? (long long)nz * 2 * Q : (long long)nz * Q;
  const double n = 2.0 * Q;
  const float scale = (float)(1.0 / (4.0 * n * n));  // column and row inverse passes, 1/(2n) each
  if (cplx)
    hipLaunchKernelGGL((big_rows_c2r<Q, true>), dim3((unsigned)(4 * units)), dim3(SplitShape<Q>::NT), 0, st, p->bbuf,
                       p->bbuf2, out, p->phx, p->phy, p->hsplit, p->bzs, scale, p->tw + tw_offset(Q));
  else
    hipLaunchKernelGGL((big_rows_c2r<Q, false>), dim3((unsigned)(4 * units)), dim3(SplitShape<Q>::NT), 0, st, p->bbuf,
                       p->bbuf2, out, p->phx, p->phy, p->hsplit, p->bzs, scale, p->tw + tw_offset(Q));
}
template <int Q>
static void launch_big_r4(const mm_radon* p, float* sino, int nz, hipStream_t st) {
  const long long rows = (long long)p->ntheta * nz, pairs = (long long)((p->ntheta + 1) / 2) * nz;
  hipLaunchKernelGGL(big_line_ifft<Q>, dim3((unsigned)(2 * pairs)), dim3(SplitShape<Q>::NT), 0, st, p->gh, sino, p->sig,
                     p->H, rows, nz, 1.0f / (4.0f * p->n), p->hsplit, p->tw + tw_offset(Q));
}
template <int Q>
static void launch_big_rt1(const mm_radon* p, const void* sino, bool cplx, int nz, hipStream_t st) {
  const long long rows = (long long)p->ntheta * nz, pairs = (long long)((p->ntheta + 1) / 2) * nz;
  if (cplx)
    hipLaunchKernelGGL((big_line_fft<Q, true>), dim3((unsigned)(2 * rows)), dim3(SplitShape<Q>::NT), 0, st, sino, p->gh,
                       p->gh2, p->sig, p->H, rows, nz, p->C0s, p->hsplit, p->tw + tw_offset(Q));
  else
    hipLaunchKernelGGL((big_line_fft<Q, false>), dim3((unsigned)(2 * pairs)), dim3(SplitShape<Q>::NT), 0, st, sino,
                       p->gh, p->gh2, p->sig, p->H, rows, nz, p->C0s, p->hsplit, p->tw + tw_offset(Q));
}

static int zsplit_for(int items, int nz) {
  if (items <= 0) return 1;
  int zs = (2048 + items - 1) / items;
  return std::max(1, std::min(zs, nz));
}

template <class S>
static void launch_gather_class(const mm_radon* p, const Item* items, int count, int nz, hipStream_t st) {
  const int zsplit = zsplit_for(count, nz);
  const int zchunk = (nz + zsplit - 1) / zsplit;
  const int zblocks = (nz + zchunk - 1) / zchunk;
  const dim3 grid((unsigned)count, (unsigned)zblocks);
  const int xg = p->tune[MM_RADON_TUNE_GATHER_XCD_GROUP];
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, grid, dim3(S::NT), 0, st, p->fde, p->gh, items, p->gsamples, p->gorder, p->dirs, p->n,
                       p->H, nz, p->zs, zchunk, p->a, p->C0g, xg);
  };
  if constexpr (S::W > GatherLeaf::W)
    go(gather_tiles<1, S>);
  else if (p->gzp == 1)
    go(gather_tiles<1, S>);
  else
    go(gather_tiles<2, S>);
}

// Root-class items (tiles wider than 32) come first in the item table, then
// the leaf class; each class is one launch.
static int launch_gather(const mm_radon* p, int nz, hipStream_t st) {
  if (p->n_groot > 0) launch_gather_class<GatherRoot>(p, p->gitems, p->n_groot, nz, st);
  if (p->n_gitems > p->n_groot)
    launch_gather_class<GatherLeaf>(p, p->gitems + p->n_groot, p->n_gitems - p->n_groot, nz, st);
  return check_launch("gather_tiles");
}

static int launch_scatter(const mm_radon* p, const float2* yh, float2* Hb, int nz, hipStream_t st) {
  const int zsplit = zsplit_for(p->mfma ? p->n_mitems : p->n_sitems, nz);
  const int zchunk = (nz + zsplit - 1) / zsplit;
  const int zblocks = (nz + zchunk - 1) / zchunk;
  const long long hzs = p->zs;
  const int xg = p->tune[MM_RADON_TUNE_SCATTER_XCD_GROUP];
  if (p->mfma) {
    const bool off64 = (unsigned long long)p->ntheta * nz * p->H * sizeof(float2) >= (1ULL << 32);
    hipLaunchKernelGGL(off64 ? scatter_mfma<true> : scatter_mfma<false>,
                       dim3((unsigned)p->n_mitems, (unsigned)zblocks), dim3(MNT), 0, st, yh, Hb, p->partial,
                       p->mitems, p->mentries, p->mrows, p->dirs, p->n, p->H, nz, hzs, zchunk,
                       (float)(-(double)p->a * M_LOG2E), xg);
  }
  else if (p->flags & MM_RADON_FP64_TILES)
    hipLaunchKernelGGL(scatter_lists<double>, dim3((unsigned)p->n_sitems, (unsigned)zblocks), dim3(SNP), 0, st, yh, Hb,
                       p->partial, p->sitems, p->sperm, p->sentries, p->slists, p->dirs, p->n, p->H, nz, hzs, zchunk,
                       (float)(-(double)p->a * M_LOG2E), xg);
  else if (p->flags & MM_RADON_KAHAN_TILES)
    hipLaunchKernelGGL(scatter_lists<KahanAcc>, dim3((unsigned)p->n_sitems, (unsigned)zblocks), dim3(SNP), 0, st, yh, Hb,
                       p->partial, p->sitems, p->sperm, p->sentries, p->slists, p->dirs, p->n, p->H, nz, hzs, zchunk,
                       (float)(-(double)p->a * M_LOG2E), xg);
  else
    hipLaunchKernelGGL(scatter_lists<float>, dim3((unsigned)p->n_sitems, (unsigned)zblocks), dim3(SNP), 0, st, yh, Hb,
                       p->partial, p->sitems, p->sperm, p->sentries, p->slists, p->dirs, p->n, p->H, nz, hzs, zchunk,
                       (float)(-(double)p->a * M_LOG2E), xg);
  int rc = check_launch(p->mfma ? "scatter_mfma" : "scatter_lists");
  if (rc) return rc;
  if (p->n_splits > 0) {
    hipLaunchKernelGGL(scatter_reduce, dim3((unsigned)p->n_splits, (unsigned)nz), dim3(SNP), 0, st, p->partial, Hb,
                       p->splits, p->n, nz, hzs);
    rc = check_launch("scatter_reduce");
  }
  return rc;
}

// Rebuilds the gather tables (items, sample lists, lane order, column-pass support) for another
// quadtree root tile and z slices per patch element: same result, other launch shape.  Host work and
// device allocations; the device is synchronised before the old tables are freed.
static int radon_regather(mm_radon* p, int root, int zp) {
  if (root == p->groot && zp == p->gzp) return MM_OK;
  HostGeom G;
  G.n = p->n;
  G.H = p->H;
  G.ntheta = p->ntheta;
  G.dx.resize(p->ntheta);
  G.dy.resize(p->ntheta);
  for (int t = 0; t < p->ntheta; ++t) {
    const double c = std::cos(p->theta_h[t]), sn = std::sin(p->theta_h[t]);
    G.dx[t] = std::fabs(c);
    G.dy[t] = -(c >= 0 ? 1 : -1) * sn;
  }
  Binning BG;
  BG.zp = zp;
  build_gather(G, std::min(root, p->n), BG);
  const std::vector<int2> gsupp = gather_support(BG, p->n);
  Item* gi = nullptr;
  int* gs = nullptr;
  unsigned short* go = nullptr;
  int2* gp = nullptr;
  size_t nb = 0;
  int rc;
  if ((rc = upload(&gi, BG.items, nb)) || (rc = upload(&gs, BG.samples, nb)) || (rc = upload(&go, BG.order, nb)) ||
      (rc = upload(&gp, gsupp, nb))) {
    for (void* q : {(void*)gi, (void*)gs, (void*)go, (void*)gp})
      if (q) (void)hipFree(q);
    return rc;
  }
  if (hipDeviceSynchronize() != hipSuccess) {  // no launch may still read the old tables
    for (void* q : {(void*)gi, (void*)gs, (void*)go, (void*)gp})
      if (q) (void)hipFree(q);
    return fail(MM_ERR_HIP, "mm_radon_tune: device synchronisation before the gather tables are rebuilt");
  }
  for (void* q : {(void*)p->gitems, (void*)p->gsamples, (void*)p->gorder, (void*)p->gsupp})
    if (q) (void)hipFree(q);
  p->gitems = gi;
  p->gsamples = gs;
  p->gorder = go;
  p->gsupp = gp;
  p->bytes = p->bytes - p->gbytes + nb;
  p->gbytes = nb;
  p->n_gitems = (int)BG.items.size();
  p->n_groot = BG.n_root;
  p->gzp = zp;
  p->groot = root;
  p->tune[MM_RADON_TUNE_GATHER_ROOT] = root;
  p->tune[MM_RADON_TUNE_GATHER_ZP] = zp;
  return MM_OK;
}

namespace mm {
void radon_dims(const mm_radon* p, int* n, int* ntheta, int* nz_max) {
  *n = p->n;
  *ntheta = p->ntheta;
  *nz_max = p->nz_max;
}
}  // namespace mm

extern "C" {

int mm_radon_create(mm_radon** plan, int n, int nz_max, int ntheta, const double* theta_host, unsigned flags) {
  return radon_create_impl(plan, n, nz_max, ntheta, theta_host, flags, nullptr, nullptr, false);
}

}  // extern "C"

namespace mm {
size_t radon_fde_elems(int n, int nz_max) { return (size_t)nz_max * (size_t)(n / 4 + 1) * 2 * n * 4; }
size_t radon_gh_elems(int n, int nz_max, int ntheta) { return (size_t)ntheta * nz_max * (n / 2 + 1); }

// Plan creation.  fde_ext / gh_ext (both or neither): caller-owned half-spectrum
// and representative buffers of at least radon_fde_elems / radon_gh_elems
// complex values, shared by plans that never run concurrently (a mosaic
// batch).  fwd_only skips the adjoint scatter tables.
int radon_create_impl(mm_radon** plan, int n, int nz_max, int ntheta, const double* theta_host, unsigned flags,
                      float2* fde_ext, float2* gh_ext, bool fwd_only) {
  MM_ARG_CHECK(plan != nullptr, "mm_radon_create: plan is NULL");
  *plan = nullptr;
  MM_ARG_CHECK(radon_width_ok(n),
               "mm_radon_create: n must be 2^k in [16, 32768] or 3 * 2^k in [96, 24576], got " + std::to_string(n));
  MM_ARG_CHECK(nz_max >= 1, "mm_radon_create: nz_max must be >= 1");
  MM_ARG_CHECK(ntheta >= 1 && theta_host != nullptr, "mm_radon_create: need ntheta >= 1 angles");
  MM_ARG_CHECK((long long)ntheta * (n / 2 + 1) < (1LL << 30),
               "mm_radon_create: ntheta * (n/2+1) must stay below 2^30 (entry codes carry it shifted by one bit)");
  MM_ARG_CHECK((flags & ~(MM_RADON_FP64_TILES | MM_RADON_GATHER_ZP1 | MM_RADON_KAHAN_TILES | MM_RADON_LIST_TILES |
                           MM_RADON_GATHER_ROOT32)) == 0,
               "mm_radon_create: unknown flags");
  MM_ARG_CHECK(__builtin_popcount(flags & (MM_RADON_FP64_TILES | MM_RADON_KAHAN_TILES | MM_RADON_LIST_TILES)) <= 1,
               "mm_radon_create: at most one of MM_RADON_FP64_TILES, MM_RADON_KAHAN_TILES, MM_RADON_LIST_TILES");
  const double eps = 1e-3;
  const double mu = -std::log(eps) / (2.0 * n * n);
  const int m = (int)std::ceil(2.0 * n / M_PI * std::sqrt(-mu * std::log(eps) + (mu * n) * (mu * n) / 4.0));
  if (m != M) return fail(MM_ERR_UNSUPPORTED, "mm_radon_create: window half width m=" + std::to_string(m) + " (kernels built for 4)");

  mm_radon* p = new mm_radon();
  int rc = MM_OK;
  auto bail = [&](int code) {
    radon_free(p);
    return code;
  };
  if ((rc = twiddle_bank(&p->tw))) return bail(rc);
  if (hipGetDevice(&p->dev) != hipSuccess) return bail(fail(MM_ERR_HIP, "hipGetDevice"));
  p->n = n;
  p->nz_max = nz_max;
  p->ntheta = ntheta;
  p->H = n / 2 + 1;
  p->NCB = n / 4 + 1;
  p->zs = (long long)p->NCB * 2 * n * 4;
  p->flags = flags;
  p->mu = mu;
  p->a = (float)(M_PI * M_PI / (mu * 4.0 * n * n));
  p->C0g = (float)(M_PI / mu / n);
  p->C0s = (float)(M_PI / mu);
  p->split = 2 * n > kMaxFFT;
  p->big = n > kMaxFFT;
  p->bzs = (long long)p->NCB * n * 4;
  if (p->big && (fde_ext || gh_ext))
    return bail(fail(MM_ERR_UNSUPPORTED, "mm_radon_create: shared workspaces are not built for n > 16384"));

  // separable phi (processing/tomo.py:46-49,59): phi(y,x) = phy[y] * phx[x]
  std::vector<float> phx(n), phy(n);
  const double sgn_n = (double)(1 - n % 4);
  const double cst = 1.0 / (n * std::sqrt((double)n * ntheta));
  for (int i = 0; i < n; ++i) {
    const double t = (double)i / n - 0.5;
    const double e = std::exp(mu * (double)n * n * t * t);
    phy[i] = (float)e;
    phx[i] = (float)(e * sgn_n * cst);
  }
  HostGeom G;
  G.n = n;
  G.H = p->H;
  G.ntheta = ntheta;
  G.dx.resize(ntheta);
  G.dy.resize(ntheta);
  std::vector<double2> dirs(ntheta);
  std::vector<int> sig(ntheta);
  for (int t = 0; t < ntheta; ++t) {
    const double c = std::cos(theta_host[t]), s = std::sin(theta_host[t]);
    const int sg = c >= 0 ? 1 : -1;
    G.dx[t] = std::fabs(c);
    G.dy[t] = -sg * s;
    dirs[t] = make_double2(G.dx[t], G.dy[t]);
    sig[t] = sg;
  }
  Binning BG;
  BG.zp = (flags & MM_RADON_GATHER_ZP1) ? 1 : 2;
  p->gzp = BG.zp;
  SBinning BS;
  MBinning BM;
  const int root = std::min((flags & MM_RADON_GATHER_ROOT32) ? ROOT / 2 : ROOT, n);
  build_gather(G, root, BG);
  p->groot = (flags & MM_RADON_GATHER_ROOT32) ? ROOT / 2 : ROOT;
  p->tune[MM_RADON_TUNE_GATHER_ROOT] = p->groot;
  // half as many columns per column-FFT workgroup measured faster (same bits) at these widths: 15 - 17 % at
  // n = 1024 and 1536, 9 % at 3072; neutral at 2048, slower at 4096
  p->tune[MM_RADON_TUNE_COL_SPLIT] = (n == 1024 || n == 1536 || n == 3072) ? 1 : 0;
  // scatter work items in XCD groups of 64 from n = 2048 on, 32 below (three scatter workgroups per CU: larger
  // groups keep the entries neighbouring tiles share in one L2; 2 - 3 % faster than groups of 4)
  p->tune[MM_RADON_TUNE_SCATTER_XCD_GROUP] = n >= 2048 ? 64 : 32;
  p->tune[MM_RADON_TUNE_GATHER_ZP] = BG.zp;
  p->theta_h.assign(theta_host, theta_host + ntheta);
  p->fwd_only = fwd_only;
  // the matrix-core products pad every call to MZ slices: a plan that never sees MZ slices takes the lists
  p->mfma = (flags & (MM_RADON_FP64_TILES | MM_RADON_KAHAN_TILES | MM_RADON_LIST_TILES)) == 0 && nz_max >= MZ;
  const std::vector<int2> gsupp = gather_support(BG, n);
  std::vector<int2> ssupp;
  if (!fwd_only) {
    if (p->mfma) {
      build_scatter_patches(G, BM);
      ssupp = trim_scatter(BM, n);
    } else {
      build_scatter_lists(G, BS);
      ssupp = trim_scatter(BS, n);
    }
  }

  size_t& bytes = p->bytes;
  if ((rc = upload(&p->phx, phx, bytes)) || (rc = upload(&p->phy, phy, bytes)) || (rc = upload(&p->dirs, dirs, bytes)) ||
      (rc = upload(&p->sig, sig, bytes)) || (rc = upload(&p->gitems, BG.items, bytes)) ||
      (rc = upload(&p->gsamples, BG.samples, bytes)) || (rc = upload(&p->gorder, BG.order, bytes)) ||
      (rc = upload(&p->sitems, BS.items, bytes)) || (rc = upload(&p->sperm, BS.perm, bytes)) ||
      (rc = upload(&p->sentries, BS.entries, bytes)) || (rc = upload(&p->slists, BS.lists, bytes)) ||
      (rc = upload(&p->mitems, BM.items, bytes)) || (rc = upload(&p->mentries, BM.entries, bytes)) ||
      (rc = upload(&p->mrows, BM.rows, bytes)) ||
      (rc = upload(&p->splits, p->mfma ? BM.splits : BS.splits, bytes)) || (rc = upload(&p->gsupp, gsupp, bytes)) ||
      (rc = upload(&p->ssupp, ssupp, bytes)))
    return bail(rc);
  if (p->split) {
    std::vector<float2> h(n);
    for (int j = 0; j < n; ++j) {
      const double a = -M_PI * (double)j / n;
      h[j] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    if ((rc = upload(&p->hsplit, h, bytes))) return bail(rc);
  }
  p->n_gitems = (int)BG.items.size();
  p->n_groot = BG.n_root;
  p->gbytes = BG.items.size() * sizeof(Item) + BG.samples.size() * sizeof(int) +
              BG.order.size() * sizeof(unsigned short) + gsupp.size() * sizeof(int2);
  p->n_sitems = (int)BS.items.size();
  p->n_mitems = (int)BM.items.size();
  p->n_splits = (int)(p->mfma ? BM.splits.size() : BS.splits.size());
  p->n_slots = p->mfma ? BM.nslots : BS.nslots;
  if (fde_ext && gh_ext) {
    p->own_ws = false;
    p->fde = fde_ext;
    p->gh = gh_ext;
  } else if ((rc = dmalloc((void**)&p->fde, sizeof(float2) * (size_t)nz_max * p->zs, bytes, "fde")) ||
             (rc = dmalloc((void**)&p->gh, sizeof(float2) * (size_t)ntheta * nz_max * p->H, bytes, "representatives"))) {
    return bail(rc);
  }
  if (p->big && (rc = dmalloc((void**)&p->bbuf, sizeof(float2) * (size_t)nz_max * p->bzs, bytes, "object rows")))
    return bail(rc);
  if (p->n_slots > 0 &&
      (rc = dmalloc((void**)&p->partial, sizeof(float2) * (size_t)p->n_slots * nz_max * SNP, bytes, "partial tiles")))
    return bail(rc);
  *plan = p;
  return MM_OK;
}
}  // namespace mm

// R on the half-spectrum buffer p->fde and the representatives p->gh (the plan's own, a batch's shared ones,
// or the static scratch below).  Every element a stage reads was written by the stage before it in the same
// call, so the buffers need no clearing.
static int radon_fwd_chain(mm_radon* p, const float* obj, float* sino, int nz, hipStream_t st) {
  const int L2 = 2 * p->n, L1 = p->n;
  int rc;
  if ((rc = mark(p, 0, 0, st))) return rc;
  if (p->big) {
    MM_DISPATCH_BIG(L1, (launch_big_r1<QQ>(p, obj, nz, st)));
  } else if (p->split) {
    MM_DISPATCH_SPLIT(L1, (launch_r1_split<LL>(p, obj, nz, st)));
  } else {
    MM_DISPATCH_FFT(L2, (launch_r1<LL>(p, obj, nz, st)));
  }
  if ((rc = check_launch("r1_rows_rfft")) || (rc = mark(p, 0, 1, st)) || (rc = mark(p, 1, 0, st))) return rc;
  if (p->big) {
    MM_DISPATCH_BIG(L1, (launch_big_col<QQ, false>(p, p->fde, p->bbuf, nz, st)));
  } else if (p->split) {
    MM_DISPATCH_SPLIT(L1, (launch_col_split<LL, false>(p, p->fde, nz, st)));
  } else {
    MM_DISPATCH_FFT(L2, (launch_col<LL, false>(p, p->fde, nz, st)));
  }
  if ((rc = check_launch("col_fft (R)")) || (rc = mark(p, 1, 1, st)) || (rc = mark(p, 2, 0, st))) return rc;
  if ((rc = launch_gather(p, nz, st)) || (rc = mark(p, 2, 1, st)) || (rc = mark(p, 3, 0, st))) return rc;
  if (p->big) {
    MM_DISPATCH_BIG(L1, (launch_big_r4<QQ>(p, sino, nz, st)));
  } else {
    MM_DISPATCH_FFT(L1, (launch_r4<LL>(p, sino, nz, st)));
  }
  if ((rc = check_launch("r4_line_ifft"))) return rc;
  return mark(p, 3, 1, st);
}

// ---- R without allocation ---------------------------------------------------
// The half spectrum and the representatives of mm_radon_fwd_noalloc live in one file-scope device array.  Its
// name is only a host-side handle: host code uses the address hipGetSymbolAddress returns for the current
// device (each device holds its own copy), resolved once per device and kept in scratch_dev.
constexpr size_t kScratchBytes = size_t(128) << 20;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static unsigned char* scratch_dev[kScratchDevices];
static std::mutex scratch_mu;

// Bytes of scratch R needs for nz slices: the half spectrum (rounded up to 256 bytes), then the representatives.
static size_t noalloc_bytes(const mm_radon* p, int nz) {
  const size_t fde_b = (sizeof(float2) * (size_t)nz * (size_t)p->zs + 255) / 256 * 256;
  return fde_b + sizeof(float2) * (size_t)p->ntheta * (size_t)nz * (size_t)p->H;
}

static int scratch_base(int dev, unsigned char** out) {
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* ptr = nullptr;
    MM_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    MM_ARG_CHECK(ptr != nullptr, "mm_radon_fwd_noalloc: no device address for the scratch array");
    scratch_dev[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = scratch_dev[dev];
  return MM_OK;
}

namespace mm {
int radon_fwd_ws(mm_radon* p, const float* obj, float* sino, int nz, hipStream_t st, float2* fde, float2* gh) {
  MM_ARG_CHECK(p != nullptr && obj && sino && fde && gh, "radon_fwd_ws: null pointer");
  MM_ARG_CHECK(nz >= 1 && nz <= p->nz_max, "radon_fwd_ws: nz outside [1, nz_max]");
  float2* const own_fde = p->fde;
  float2* const own_gh = p->gh;
  p->fde = fde;
  p->gh = gh;
  const int rc = radon_fwd_chain(p, obj, sino, nz, st);
  p->fde = own_fde;
  p->gh = own_gh;
  return rc;
}
}  // namespace mm

extern "C" {

int mm_radon_fwd(mm_radon* p, const float* obj, float* sino, int nz, void* stream) {
  MM_ARG_CHECK(p != nullptr, "mm_radon_fwd: plan is NULL");
  MM_ARG_CHECK(nz >= 1 && nz <= p->nz_max,
               "mm_radon_fwd: nz=" + std::to_string(nz) + " outside [1, " + std::to_string(p->nz_max) + "]");
  MM_ARG_CHECK(obj && sino, "mm_radon_fwd: null pointer");
  return radon_fwd_chain(p, obj, sino, nz, (hipStream_t)stream);
}

int mm_radon_fwd_noalloc(mm_radon* p, const float* obj, float* sino, int nz, void* stream) {
  MM_ARG_CHECK(p != nullptr, "mm_radon_fwd_noalloc: plan is NULL");
  MM_ARG_CHECK(nz >= 1 && nz <= p->nz_max,
               "mm_radon_fwd_noalloc: nz=" + std::to_string(nz) + " outside [1, " + std::to_string(p->nz_max) + "]");
  MM_ARG_CHECK(obj && sino, "mm_radon_fwd_noalloc: null pointer");
  const size_t need = noalloc_bytes(p, nz);
  MM_ARG_CHECK(need <= kScratchBytes, "mm_radon_fwd_noalloc: n=" + std::to_string(p->n) + ", ntheta=" +
                                          std::to_string(p->ntheta) + ", nz=" + std::to_string(nz) + " needs " +
                                          std::to_string(need) + " bytes of scratch, the static buffer holds " +
                                          std::to_string(kScratchBytes) + " (use mm_radon_fwd or fewer slices)");
  int dev = -1;
  MM_HIP_TRY(hipGetDevice(&dev));
  MM_ARG_CHECK(dev == p->dev, "mm_radon_fwd_noalloc: the plan belongs to device " + std::to_string(p->dev) +
                                  ", device " + std::to_string(dev) + " is current");
  MM_ARG_CHECK(dev >= 0 && dev < kScratchDevices, "mm_radon_fwd_noalloc: device index " + std::to_string(dev) +
                                                      " beyond the " + std::to_string(kScratchDevices) + " supported");
  unsigned char* base = nullptr;
  int rc = scratch_base(dev, &base);
  if (rc) return rc;
  float2* const own_fde = p->fde;
  float2* const own_gh = p->gh;
  p->fde = reinterpret_cast<float2*>(base);
  p->gh = reinterpret_cast<float2*>(base + (need - sizeof(float2) * (size_t)p->ntheta * nz * p->H));
  rc = radon_fwd_chain(p, obj, sino, nz, (hipStream_t)stream);
  p->fde = own_fde;
  p->gh = own_gh;
  return rc;
}

static int radon_adj_tail(mm_radon* p, bool cplx, void* obj, int nz, hipStream_t st);

int mm_radon_adj(mm_radon* p, const void* sino, int sino_is_c64, void* obj, int nz, void* stream) {
  MM_ARG_CHECK(p != nullptr, "mm_radon_adj: plan is NULL");
  MM_ARG_CHECK(nz >= 1 && nz <= p->nz_max,
               "mm_radon_adj: nz=" + std::to_string(nz) + " outside [1, " + std::to_string(p->nz_max) + "]");
  MM_ARG_CHECK(sino && obj, "mm_radon_adj: null pointer");
  MM_ARG_CHECK(sino_is_c64 == 0 || sino_is_c64 == 1, "mm_radon_adj: sino_is_c64 must be 0 or 1");
  if (p->fwd_only) return fail(MM_ERR_UNSUPPORTED, "mm_radon_adj: forward-only plan (no scatter tables)");
  const bool cplx = sino_is_c64 != 0;
  if (cplx) {
    int rc;
    if (!p->fde2 &&
        (rc = dmalloc((void**)&p->fde2, sizeof(float2) * (size_t)p->nz_max * p->zs, p->bytes, "fde (im)")))
      return rc;
    if (!p->gh2 &&
        (rc = dmalloc((void**)&p->gh2, sizeof(float2) * (size_t)p->ntheta * p->nz_max * p->H, p->bytes, "representatives (im)")))
      return rc;
  }
  if (cplx && p->big && !p->bbuf2) {
    int rc;
    if ((rc = dmalloc((void**)&p->bbuf2, sizeof(float2) * (size_t)p->nz_max * p->bzs, p->bytes, "object rows (im)")))
      return rc;
  }
  hipStream_t st = (hipStream_t)stream;
  const int L1 = p->n;
  int rc;
  if ((rc = mark(p, 4, 0, st))) return rc;
  if (p->big) {
    MM_DISPATCH_BIG(L1, (launch_big_rt1<QQ>(p, sino, cplx, nz, st)));
  } else {
    MM_DISPATCH_FFT(L1, (launch_rt1<LL>(p, sino, cplx, nz, st)));
  }
  if ((rc = check_launch("rt1_line_fft")) || (rc = mark(p, 4, 1, st))) return rc;
  return radon_adj_tail(p, cplx, obj, nz, st);
}

// RT after the line pass: scatter, column IFFT, C2R rows.
static int radon_adj_tail(mm_radon* p, bool cplx, void* obj, int nz, hipStream_t st) {
  const int L2 = 2 * p->n, L1 = p->n;
  int rc;
  if ((rc = mark(p, 5, 0, st))) return rc;
  if ((rc = launch_scatter(p, p->gh, p->fde, nz, st))) return rc;
  if (cplx && (rc = launch_scatter(p, p->gh2, p->fde2, nz, st))) return rc;
  if ((rc = mark(p, 5, 1, st)) || (rc = mark(p, 6, 0, st))) return rc;
  for (int ch = 0; ch < (cplx ? 2 : 1); ++ch) {
    float2* buf = ch ? p->fde2 : p->fde;
    if (p->big) {
      MM_DISPATCH_BIG(L1, (launch_big_col<QQ, true>(p, buf, ch ? p->bbuf2 : p->bbuf, nz, st)));
    } else if (p->split) {
      MM_DISPATCH_SPLIT(L1, (launch_col_split<LL, true>(p, buf, nz, st)));
    } else {
      MM_DISPATCH_FFT(L2, (launch_col<LL, true>(p, buf, nz, st)));
    }
    if ((rc = check_launch(ch ? "col_fft (RT im)" : "col_fft (RT)"))) return rc;
  }
  if ((rc = mark(p, 6, 1, st)) || (rc = mark(p, 7, 0, st))) return rc;
  if (p->big) {
    MM_DISPATCH_BIG(L1, (launch_big_c2r<QQ>(p, obj, cplx, nz, st)));
  } else if (p->split) {
    MM_DISPATCH_SPLIT(L1, (launch_rt4_split<LL>(p, obj, cplx, nz, st)));
  } else {
    MM_DISPATCH_FFT(L2, (launch_rt4<LL>(p, obj, cplx, nz, st)));
  }
  if ((rc = check_launch("rt4_rows_c2r"))) return rc;
  return mark(p, 7, 1, st);
}

int mm_radon_adj_fbp(mm_radon* p, const float* sino, const float* G, const float* A, const float* B, float* obj,
                     int nz, void* stream) {
  MM_ARG_CHECK(p != nullptr, "mm_radon_adj_fbp: plan is NULL");
  MM_ARG_CHECK(nz >= 1 && nz <= p->nz_max,
               "mm_radon_adj_fbp: nz=" + std::to_string(nz) + " outside [1, " + std::to_string(p->nz_max) + "]");
  MM_ARG_CHECK(sino && G && A && B && obj, "mm_radon_adj_fbp: null pointer");
  if (p->fwd_only) return fail(MM_ERR_UNSUPPORTED, "mm_radon_adj_fbp: forward-only plan (no scatter tables)");
  if (p->big) return fail(MM_ERR_UNSUPPORTED, "mm_radon_adj_fbp: the fused filter is not built for n > 16384");
  hipStream_t st = (hipStream_t)stream;
  int rc;
  if ((rc = mark(p, 4, 0, st))) return rc;
  if (fbp_fusable(p->n)) {
    MM_DISPATCH_FFT(p->n, (launch_rt1_fbp<LL>(p, sino, G, A, B, nz, st)));
    if ((rc = check_launch("rt1_fbp_line_fft"))) return rc;
  } else {
    // a length whose filter transform holds a different number of points per thread (n = 384): filter a
    // copy in the second representative buffer (unused by real sinograms, large enough), then the line pass
    if (!p->gh2 && (rc = dmalloc((void**)&p->gh2, sizeof(float2) * radon_gh_elems(p->n, p->nz_max, p->ntheta),
                                 p->bytes, "filtered sinogram")))
      return rc;
    float* tmp = reinterpret_cast<float*>(p->gh2);
    const size_t nb = sizeof(float) * (size_t)p->ntheta * nz * p->n;
    if (hipMemcpyAsync(tmp, sino, nb, hipMemcpyDeviceToDevice, st) != hipSuccess)
      return fail(MM_ERR_HIP, "mm_radon_adj_fbp: sinogram copy");
    if ((rc = mm_fbp_filter_2n(tmp, G, A, B, (long long)p->ntheta * nz, p->n, stream))) return rc;
    MM_DISPATCH_FFT(p->n, (launch_rt1<LL>(p, tmp, false, nz, st)));
    if ((rc = check_launch("rt1_line_fft"))) return rc;
  }
  if ((rc = mark(p, 4, 1, st))) return rc;
  return radon_adj_tail(p, false, obj, nz, st);
}

int mm_radon_tune(mm_radon* p, int knob, int value) {
  MM_ARG_CHECK(p != nullptr, "mm_radon_tune: plan is NULL");
  MM_ARG_CHECK(knob >= 1 && knob < MM_RADON_TUNE_NKNOBS, "mm_radon_tune: unknown knob " + std::to_string(knob));
  if (knob == MM_RADON_TUNE_GATHER_ROOT) {
    MM_ARG_CHECK(value == 32 || value == 64, "mm_radon_tune: gather root tile must be 32 or 64");
    return radon_regather(p, value, p->gzp);
  }
  if (knob == MM_RADON_TUNE_GATHER_ZP) {
    MM_ARG_CHECK(value == 1 || value == 2, "mm_radon_tune: gather slices per patch element must be 1 or 2");
    return radon_regather(p, p->groot, value);
  }
  if (knob == MM_RADON_TUNE_COL_SPLIT)
    MM_ARG_CHECK(value == 0 || value == 1, "mm_radon_tune: column split must be 0 or 1");
  else
    MM_ARG_CHECK(value >= 1 && value <= 64, "mm_radon_tune: XCD group size must be in [1, 64]");
  p->tune[knob] = value;
  return MM_OK;
}

int mm_radon_autotune(mm_radon* p, const float* obj, float* sino, float* back, int nz, void* stream) {
  MM_ARG_CHECK(p != nullptr, "mm_radon_autotune: plan is NULL");
  MM_ARG_CHECK(obj && sino, "mm_radon_autotune: null pointer");
  MM_ARG_CHECK(nz >= 1 && nz <= p->nz_max, "mm_radon_autotune: nz outside [1, nz_max]");
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  MM_HIP_TRY(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(MM_ERR_HIP, "mm_radon_autotune: hipEventCreate");
  }
  int rc = MM_OK;
  // best of two timed calls for every candidate value; the knob keeps the fastest
  auto sweep = [&](int knob, std::initializer_list<int> cand, auto&& call) {
    float best_ms = 1e30f;
    int best = p->tune[knob];
    if ((rc = call())) return;  // warm-up
    for (int c : cand) {
      p->tune[knob] = c;
      for (int r = 0; r < 2; ++r) {
        if (hipEventRecord(e0, st) != hipSuccess || (rc = call()) || hipEventRecord(e1, st) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess) {
          if (!rc) rc = fail(MM_ERR_HIP, "mm_radon_autotune: event timing");
          return;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best_ms) {
          best_ms = ms;
          best = c;
        }
      }
    }
    p->tune[knob] = best;
  };
  // gather: every launch shape (root tile 64 / 32, one / two slices per patch element; bitwise the same
  // result) with every XCD group, the fastest pair kept
  {
    float best_ms = 1e30f;
    int best_root = p->groot, best_zp = p->gzp, best_xg = p->tune[MM_RADON_TUNE_GATHER_XCD_GROUP];
    const int roots[2] = {ROOT, ROOT / 2};
    for (int ri = 0; ri < (p->n > ROOT / 2 ? 2 : 1) && !rc; ++ri)
      for (int zp = 2; zp >= 1 && !rc; --zp) {
        if ((rc = radon_regather(p, roots[ri], zp))) break;
        if ((rc = mm_radon_fwd(p, obj, sino, nz, stream))) break;  // warm-up
        for (int xg : {1, 2, 4, 8, 16}) {
          p->tune[MM_RADON_TUNE_GATHER_XCD_GROUP] = xg;
          for (int r = 0; r < 2 && !rc; ++r) {
            if (hipEventRecord(e0, st) != hipSuccess || (rc = mm_radon_fwd(p, obj, sino, nz, stream)) ||
                hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
              if (!rc) rc = fail(MM_ERR_HIP, "mm_radon_autotune: event timing");
              break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best_ms) {
              best_ms = ms;
              best_root = roots[ri];
              best_zp = zp;
              best_xg = xg;
            }
          }
        }
      }
    if (!rc) rc = radon_regather(p, best_root, best_zp);
    p->tune[MM_RADON_TUNE_GATHER_XCD_GROUP] = best_xg;
  }
  // the column split halves the columns per workgroup of the 2n-point column transforms (same bits)
  if (!rc && !p->big && !p->split && p->n >= 512) {
    auto both = [&]() {
      int r = mm_radon_fwd(p, obj, sino, nz, stream);
      if (!r && back && !p->fwd_only) r = mm_radon_adj(p, sino, 0, back, nz, stream);
      return r;
    };
    sweep(MM_RADON_TUNE_COL_SPLIT, {0, 1}, both);
  }
  if (!rc && back && !p->fwd_only)
    sweep(MM_RADON_TUNE_SCATTER_XCD_GROUP, {4, 8, 16, 32, 64}, [&]() { return mm_radon_adj(p, sino, 0, back, nz, stream); });
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int mm_radon_get_tune(const mm_radon* p, int knob) {
  if (!p || knob < 1 || knob >= MM_RADON_TUNE_NKNOBS) return MM_ERR_ARG;
  return p->tune[knob];
}

int mm_radon_profile(mm_radon* p, int enable) {
  MM_ARG_CHECK(p != nullptr, "mm_radon_profile: plan is NULL");
  if (enable && !p->ev[0][0]) {
    for (auto& e : p->ev)
      for (auto& x : e) MM_HIP_TRY(hipEventCreate(&x));
  }
  p->profile = enable != 0;
  for (bool& r : p->ran) r = false;
  return MM_OK;
}

int mm_radon_stage_ms(mm_radon* p, float* ms, int max_stages) {
  MM_ARG_CHECK(p != nullptr && ms != nullptr, "mm_radon_stage_ms: null pointer");
  for (int s = 0; s < max_stages && s < MM_RADON_NSTAGES; ++s) {
    ms[s] = -1.f;
    if (!p->ran[s]) continue;
    MM_HIP_TRY(hipEventSynchronize(p->ev[s][1]));
    MM_HIP_TRY(hipEventElapsedTime(&ms[s], p->ev[s][0], p->ev[s][1]));
  }
  return MM_OK;
}

// Host-only check of the plan tables against the index arithmetic of the
// gather and scatter kernels (every staging offset inside a half-spectrum
// plane, every patch element staged once, every tap read inside the patch,
// every list slot naming a staged entry row).  Builds the tables like
// mm_radon_create, touches no device.
int mm_radon_plan_check(int n, int ntheta, const double* theta_host) {
  MM_ARG_CHECK(radon_width_ok(n), "mm_radon_plan_check: unsupported n " + std::to_string(n));
  MM_ARG_CHECK(ntheta >= 1 && theta_host != nullptr, "mm_radon_plan_check: need ntheta >= 1 angles");
  HostGeom G;
  G.n = n;
  G.H = n / 2 + 1;
  G.ntheta = ntheta;
  G.dx.resize(ntheta);
  G.dy.resize(ntheta);
  for (int t = 0; t < ntheta; ++t) {
    const double c = std::cos(theta_host[t]), sn = std::sin(theta_host[t]);
    G.dx[t] = std::fabs(c);
    G.dy[t] = -(c >= 0 ? 1 : -1) * sn;
  }
  const long long zs = (long long)(n / 4 + 1) * 2 * n * 4;
  auto bad = [&](const std::string& what) { return fail(MM_ERR_ARG, "mm_radon_plan_check n=" + std::to_string(n) + ": " + what); };
  for (int zp = 1; zp <= 2; ++zp) {
    Binning B;
    B.zp = zp;
    build_gather(G, std::min(ROOT, n), B);
    const std::vector<int2> gs = gather_support(B, n);
    for (size_t i = 0; i < B.items.size(); ++i) {
      const Item& it = B.items[i];
      const bool root_class = (int)i < B.n_root;
      if (root_class != (it.w > GatherLeaf::W)) return bad("item class order");
      const int PE = root_class ? GatherRoot::PE : GatherLeaf::PE;
      const int nslots = root_class ? GatherRoot::STG * GatherRoot::NT : GatherLeaf::STG * GatherLeaf::NT;
      if (it.count > SMAX) return bad("samples per item");
      const int pw = it.w + 2 * M, ph = it.h + 2 * M, pp = patch_pitch(it.w);
      if (pp < pw || ph * pp > PE) return bad("patch pitch");
      std::vector<int> cover(PE + 1, 0);
      for (int idx = 0; idx < nslots; ++idx) {
        unsigned so;
        int sd;
        gather_stage_desc(it, idx, n, PE, so, sd);
        if ((long long)so >= zs || (sd >> 1) > PE)
          return bad("staging offset " + std::to_string(so) + " / slot " + std::to_string(sd >> 1) + " out of range");
        if ((sd >> 1) < PE) {  // a patch element: its row is one the forward column pass writes
          const int blk = (int)(so / (8LL * n)), row = (int)((so / 4) % (2 * n));
          const int l = row < n ? row : row - 2 * n;
          if (l < gs[blk].x || l >= gs[blk].y) return bad("patch row outside the column-pass support");
        }
        ++cover[sd >> 1];
      }
      for (int r = 0; r < ph; ++r)
        for (int c = 0; c < pw; ++c)
          if (cover[r * pp + c] != 1) return bad("patch element staged " + std::to_string(cover[r * pp + c]) + " times");
      for (int k = 0; k < it.count; ++k) {
        int rx, ry;
        G.centre(B.samples[it.start + k], rx, ry);
        if (rx < it.x0 || rx >= it.x0 + it.w || ry < it.y0 || ry >= it.y0 + it.h) return bad("sample outside its tile");
        const int b = (ry - it.y0) * pp + (rx - it.x0);
        const int kzp = gather_zp(it.w, zp);
        const int last = kzp == 1 ? (b & ~1) + 8 * pp + 9 : b + 8 * pp + 8;
        if (last >= PE + (kzp == 1 ? 2 : 0)) return bad("tap read past the patch");
        if (B.order[it.start + k] >= it.count) return bad("output order");
      }
    }
  }
  SBinning S;
  build_scatter_lists(G, S);
  const std::vector<int2> ss = trim_scatter(S, n);
  {  // every row the inverse column pass reads is written by one owning item or one split tile
    const int ntx = n / STS + 1, nty = 2 * n / STS;
    std::vector<int> wr((size_t)ntx * nty, 0);
    if (S.perm.size() != S.items.size() * SNP) return bad("scatter permutation table");
    for (const SItem& it : S.items)
      if (it.slot < 0) ++wr[(size_t)((it.y0 + n) / STS) * ntx + it.x0 / STS];
    for (const SplitTile& t : S.splits) ++wr[(size_t)((t.y0 + n) / STS) * ntx + t.x0 / STS];
    for (size_t i = 0; i < wr.size(); ++i)
      if (wr[i] > 1) return bad("scatter tile written twice");
    for (int c = 0; c <= n; ++c)
      for (int l = ss[c >> 2].x; l < ss[c >> 2].y; ++l)
        if (wr[(size_t)((l + n) / STS) * ntx + c / STS] != 1) return bad("support row not written by the scatter");
    for (const SItem& it : S.items)
      if (it.ne > 0 && (it.y0 < ss[it.x0 >> 2].x || it.y0 + STS > ss[it.x0 >> 2].y))
        return bad("scatter tile outside the support");
  }
  for (const SItem& it : S.items) {
    if (it.ne > SE_MAX) return bad("scatter entries per item");
    if (it.loff + (long long)it.nl / 2 * 2 * SNP > (long long)S.lists.size()) return bad("list offset");
    for (int w = 0; w < SWAVES; ++w) {
      const int np = (int)((it.wnp >> (8 * w)) & 0xFFu);
      if (2 * np > it.nl || 2 * np > SL_MAX) return bad("wave schedule length");
      for (int j = 0; j < 2 * np; ++j)
        for (int l = 64 * w; l < 64 * w + 64; ++l) {
          const unsigned u = S.lists[it.loff + (long long)(j / 2) * 2 * SNP + 2 * l + (j & 1)];
          if ((int)(u & 0x1FFu) >= std::max(it.ne, 1) || ((u >> 9) >= (unsigned)(TAPS * TAPS) && (u >> 9) != SL_IDLE_TAP))
            return bad("list slot");
        }
    }
  }
  {  // the slots the kernel reads give every tile point in the half plane exactly the taps of the
     // item's entries that land on it: (entry, i0, i1) from the lists against a direct enumeration
    const int twon = 2 * n;
    std::vector<std::vector<int>> want(SNP), got(SNP);
    for (size_t ii = 0; ii < S.items.size(); ++ii) {
      const SItem& it = S.items[ii];
      for (auto& v : want) v.clear();
      for (auto& v : got) v.clear();
      for (int e = 0; e < it.ne; ++e) {
        const int code = S.entries[it.eoff + e];
        int rx, ry;
        G.centre(code >> 1, rx, ry);
        const int cx = (code & 1) ? -rx : rx, cy = (code & 1) ? -ry : ry;
        for (int i1 = 0; i1 < TAPS; ++i1) {
          const int lr = wrapc(cy - M + i1, n) - it.y0;
          if (lr < 0 || lr >= STS) continue;
          for (int i0 = 0; i0 < TAPS; ++i0) {
            const int raw = pmod(cx - M + i0, twon), lc = raw - it.x0;
            if (raw <= n && lc >= 0 && lc < STS) want[lr * STS + lc].push_back((e * TAPS + i1) * TAPS + i0);
          }
        }
      }
      std::vector<int> seen(SNP, 0);
      for (int l = 0; l < SNP; ++l) {
        const int pt = S.perm[ii * SNP + l];
        if (seen[pt]++) return bad("scatter lane permutation");
        const int np = (int)((it.wnp >> (8 * (l / 64))) & 0xFFu);
        for (int j = 0; j < 2 * np; ++j) {
          const unsigned u = S.lists[it.loff + (long long)(j / 2) * 2 * SNP + 2 * l + (j & 1)];
          const int e = (int)(u & 0x1FFu), t = (int)(u >> 9);
          if (t != (int)SL_IDLE_TAP) got[pt].push_back((e * TAPS + t / TAPS) * TAPS + t % TAPS);
        }
      }
      for (int pt = 0; pt < SNP; ++pt) {
        if (it.x0 + pt % STS > n) continue;  // never written
        std::sort(want[pt].begin(), want[pt].end());
        std::sort(got[pt].begin(), got[pt].end());
        if (want[pt] != got[pt]) return bad("scatter taps of a point differ from its entries' taps");
      }
    }
  }
  {  // patch runs of scatter_mfma: every tile point past column n aside, the (entry, tap) pairs the
     // kernel's three runs and window test give a point equal a direct enumeration of the taps
    MBinning B;
    build_scatter_patches(G, B);
    const std::vector<int2> ms = trim_scatter(B, n);
    if (ms.size() != ss.size()) return bad("patch-run support size");
    for (size_t c = 0; c < ms.size(); ++c)
      if (ms[c].x != ss[c].x || ms[c].y != ss[c].y) return bad("patch-run support differs from the list support");
    const int ntx = n / STS + 1, nty = 2 * n / STS, twon = 2 * n;
    std::vector<int> wr((size_t)ntx * nty, 0);
    for (const MItem& it : B.items)
      if (it.slot < 0) ++wr[(size_t)((it.y0 + n) / STS) * ntx + it.x0 / STS];
    for (const SplitTile& t : B.splits) ++wr[(size_t)((t.y0 + n) / STS) * ntx + t.x0 / STS];
    for (int c = 0; c <= n; ++c)
      for (int l = ms[c >> 2].x; l < ms[c >> 2].y; ++l)
        if (wr[(size_t)((l + n) / STS) * ntx + c / STS] != 1) return bad("support row not written by the patch runs");
    std::vector<std::vector<int>> want(SNP), got(SNP);
    for (const MItem& it : B.items) {
      if (it.ne > ME_MAX) return bad("patch-run entries per item");
      if (it.eoff < 0 || it.eoff + it.ne > (int)B.entries.size() || B.rows.size() != B.entries.size())
        return bad("patch-run entry offset");
      if (it.cls[0] != 0 || it.cls[MCLS] != it.ne) return bad("anchor class offsets");
      for (int k = 0; k < MCLS; ++k)
        if (it.cls[k] > it.cls[k + 1]) return bad("anchor class offsets not ascending");
      for (auto& v : want) v.clear();
      for (auto& v : got) v.clear();
      std::vector<int> code_of(it.ne, -1);  // LDS row -> entry code
      for (int j = 0; j < it.ne; ++j) {
        const int r = B.rows[it.eoff + j];
        if (r >= it.ne || code_of[r] >= 0) return bad("patch-run row map");
        if (j > 0 && B.entries[it.eoff + j] <= B.entries[it.eoff + j - 1])
          return bad("patch-run staging order");
        code_of[r] = B.entries[it.eoff + j];
      }
      for (int e = 0; e < it.ne; ++e) {
        const int code = code_of[e];
        int rx, ry;
        G.centre(code >> 1, rx, ry);
        const int cx = (code & 1) ? -rx : rx, cy = (code & 1) ? -ry : ry;
        for (int i1 = 0; i1 < TAPS; ++i1) {
          const int lr = wrapc(cy - M + i1, n) - it.y0;
          if (lr < 0 || lr >= STS) continue;
          for (int i0 = 0; i0 < TAPS; ++i0) {
            const int raw = pmod(cx - M + i0, twon), lc = raw - it.x0;
            if (raw <= n && lc >= 0 && lc < STS) want[lr * STS + lc].push_back((e * TAPS + i1) * TAPS + i0);
          }
        }
        int bx, by;
        window_base(G, code, it.x0, it.y0, bx, by);
        const int cls = ((by + 8) >> 2) * 6 + ((bx + 8) >> 2);
        if (bx < -8 || bx >= STS || by < -8 || by >= STS || e < it.cls[cls] || e >= it.cls[cls + 1])
          return bad("entry outside its anchor class");
      }
      // the kernel: patch (px, py), runs [cls[ay 6 + px], cls[ay 6 + px + 3]) for ay = py .. py + 2,
      // tap (tx, ty) = point - (bx + M, by + M) counted iff max(|tx|, |ty|) <= M
      for (int P = 0; P < 16; ++P) {
        const int px = P & 3, py = P >> 2;
        for (int ay = py; ay < py + 3; ++ay)
          for (int e = it.cls[ay * 6 + px]; e < it.cls[ay * 6 + px + 3]; ++e) {
            int bx, by;
            window_base(G, code_of[e], it.x0, it.y0, bx, by);
            for (int q = 0; q < 16; ++q) {
              const int qx = 4 * px + (q & 3), qy = 4 * py + (q >> 2);
              const int tx = qx - (bx + M), ty = qy - (by + M);
              if (std::max(std::abs(tx), std::abs(ty)) <= M)
                got[qy * STS + qx].push_back((e * TAPS + (ty + M)) * TAPS + (tx + M));
            }
          }
      }
      for (int pt = 0; pt < SNP; ++pt) {
        if (it.x0 + pt % STS > n) continue;  // never written
        std::sort(want[pt].begin(), want[pt].end());
        std::sort(got[pt].begin(), got[pt].end());
        if (want[pt] != got[pt]) return bad("patch-run taps of a point differ from its entries' taps");
      }
    }
  }
  return MM_OK;
}

size_t mm_radon_workspace_bytes(const mm_radon* p) { return p ? p->bytes : 0; }

void mm_radon_destroy(mm_radon* p) { radon_free(p); }

}  // extern "C"
