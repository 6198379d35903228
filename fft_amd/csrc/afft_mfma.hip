// afft_mfma.hip -- split-precision MFMA path: 256-point DFT tiles whose two
// radix-16 stages run on the matrix cores.
//
// No counterpart in afft (its butterflies are FP32/FP64 SIMD code,
// onedim.hpp:55-133); the transform computed is the same DFT.
//
// A radix-16 stage over 16 butterflies is one real matrix product
//     [Yr; Yi] (32 x 16) = A (32 x 32) * [Xr; Xi] (32 x 16)
// with A = [[Re C, -Im C], [Im C, Re C]] and C the 16 x 16 complex stage
// matrix.  Operands are FP16 split in two: x*s = hi + lo, A = A_hi + A_lo; the
// product keeps A_hi*hi + A_hi*lo + A_lo*hi (the dropped A_lo*lo term is
// ~2^-22 relative), accumulated in FP32 by v_mfma_f32_16x16x32_f16.  s is a
// power of two per transform, taken from the transform's max |Re|, |Im| in the
// load phase so that max*s lies in [2^10, 2^11): stage-1 outputs then stay
// below 16 * 2^11 = 2^15, inside FP16's range, and feed stage 2 unscaled; the
// store multiplies by 1/s exactly.  Stage 1 uses the
// plain DFT-16 matrix; stage 2 (Ns = 16) folds the Stockham twiddles
// W_256^(r*k) into 16 matrices C_k = F16 * diag(W_256^(r*k)), so no twiddle
// multiply runs on the VALU.
//
// BF16 variant (afft_gpu_mode_mfma_bf16, v_mfma_f32_16x16x32_bf16): BF16 keeps
// 8 significand bits, so operands are split three ways, x = hi + mid + lo and
// A = A_hi + A_mid + A_lo, and the six products whose order is at least 2^-16
// are kept (hi*hi, hi*mid, mid*hi, hi*lo, mid*mid, lo*hi; the dropped ones are
// ~2^-24 relative).  BF16 has FP32's exponent range; the same per-transform
// scaling is applied anyway, so both variants share the data path.
//
// Operand layouts (v_mfma_f32_16x16x32_f16 / _bf16, wave64): lane l holds row l%16 of
// A and column l%16 of B, eight consecutive K values starting at 8*(l/16);
// the result holds column l%16, rows 4*(l/16)..+3.  Only the result layout
// matters for correctness (A and B share one K order); it is checked on the
// device by afft_gpu_mfma_selftest.
#include <hip/hip_runtime.h>

#include "../../include/afft_gpu.h"
#include "afft_device.h"
#include "afft_mfma.h"
#include "afft_pass.h"

namespace afgpu {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kML = 256;   // points per transform
constexpr int kMCT = 16;   // transforms per tile
constexpr int kMLP = 264;  // padded LDS row (complex entries)

// A-operand table: [matrix (0 = stage 1, 1 + k = stage 2, k < 16)][mtile][hi/lo][lane] -> h8
constexpr int kMats = 17;
__host__ __device__ constexpr int mat_index(int mat, int mtile, int part, int lane) {
  return ((mat * 2 + mtile) * 2 + part) * 64 + lane;
}

// One A operand value: row m of M-tile `mtile`, K column kk, for matrix `mat`.
__device__ double a_value(int mat, int mtile, int m, int kk) {
  const int part = kk >> 4;  // 0: multiplies Re x_r, 1: multiplies Im x_r
  const int r = kk & 15;
  // C[m][r] = W16^(m r) * W256^(r k) = exp(-2 pi i (16 m r + r k) / 256)
  const int k = mat == 0 ? 0 : mat - 1;
  const int e = (16 * m * r + r * k) & 255;
  double s, c;
  sincospi(-2.0 * (double)e / 256.0, &s, &c);  // c + i s = C[m][r]
  if (mtile == 0) return part == 0 ? c : -s;     // Re(C x) = Re C Re x - Im C Im x
  return part == 0 ? s : c;                      // Im(C x) = Im C Re x + Re C Im x
}

__global__ void gen_mfma_mats_kernel(h8* out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // (mat, mtile, lane)
  if (idx >= kMats * 2 * 64) return;
  const int lane = idx & 63;
  const int mtile = (idx >> 6) & 1;
  const int mat = idx >> 7;
  const int m = lane & 15;
  const int k0 = 8 * (lane >> 4);
  h8 hi, lo;
  for (int i = 0; i < 8; i++) {
    double a = a_value(mat, mtile, m, k0 + i);
    _Float16 h = (_Float16)a;
    hi[i] = h;
    lo[i] = (_Float16)(a - (double)h);
  }
  out[mat_index(mat, mtile, 0, lane)] = hi;
  out[mat_index(mat, mtile, 1, lane)] = lo;
}

// BF16 table: [matrix][mtile][hi/mid/lo][lane] -> b8
__host__ __device__ constexpr int mat_index3(int mat, int mtile, int part, int lane) {
  return ((mat * 2 + mtile) * 3 + part) * 64 + lane;
}

__global__ void gen_mfma_mats_bf16_kernel(b8* out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // (mat, mtile, lane)
  if (idx >= kMats * 2 * 64) return;
  const int lane = idx & 63;
  const int mtile = (idx >> 6) & 1;
  const int mat = idx >> 7;
  const int m = lane & 15;
  const int k0 = 8 * (lane >> 4);
  b8 hi, mid, lo;
  for (int i = 0; i < 8; i++) {
    const double a = a_value(mat, mtile, m, k0 + i);
    const __bf16 h = (__bf16)(float)a;
    const double r1 = a - (double)(float)h;
    const __bf16 md = (__bf16)(float)r1;
    const double r2 = r1 - (double)(float)md;
    hi[i] = h;
    mid[i] = md;
    lo[i] = (__bf16)(float)r2;
  }
  out[mat_index3(mat, mtile, 0, lane)] = hi;
  out[mat_index3(mat, mtile, 1, lane)] = mid;
  out[mat_index3(mat, mtile, 2, lane)] = lo;
}



// FP16 hi/lo split of 8 values (already scaled into FP16 range): hi = x
// rounded to FP16, lo = the FP32 remainder x - hi (exact) rounded to FP16.
__device__ __forceinline__ void split8(const float (&x)[8], h8& hi, h8& lo) {
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const _Float16 h = (_Float16)x[i];
    hi[i] = h;
    lo[i] = (_Float16)(x[i] - (float)h);
  }
}

// Power-of-two exponent that puts max |value| of a transform into [2^10, 2^11).
__device__ __forceinline__ int transform_shift(float mx) {
  if (!(mx > 0.f) || !isfinite(mx)) return 0;
  int e;
  (void)frexpf(mx, &e);  // mx in [2^(e-1), 2^e)
  return 11 - e;
}

struct AOps {
  h8 a0h, a0l, a1h, a1l;  // M-tile 0 (Re rows) and 1 (Im rows), hi and lo
};

__device__ __forceinline__ AOps load_aops(const h8* __restrict__ mats, int mat, int lane) {
  return {mats[mat_index(mat, 0, 0, lane)], mats[mat_index(mat, 0, 1, lane)], mats[mat_index(mat, 1, 0, lane)],
          mats[mat_index(mat, 1, 1, lane)]};
}

// BF16: three-way split of 8 values, x = hi + mid + lo (each remainder exact
// in FP32, each part rounded to BF16).
__device__ __forceinline__ void split8_bf16(const float (&x)[8], b8& hi, b8& mid, b8& lo) {
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const __bf16 h = (__bf16)x[i];
    const float r1 = x[i] - (float)h;
    const __bf16 m = (__bf16)r1;
    hi[i] = h;
    mid[i] = m;
    lo[i] = (__bf16)(r1 - (float)m);
  }
}

struct AOps3 {
  b8 h[2], m[2], l[2];  // per M-tile (Re rows, Im rows): hi, mid, lo
};

__device__ __forceinline__ AOps3 load_aops3(const b8* __restrict__ mats, int mat, int lane) {
  AOps3 A;
#pragma unroll
  for (int t = 0; t < 2; t++) {
    A.h[t] = mats[mat_index3(mat, t, 0, lane)];
    A.m[t] = mats[mat_index3(mat, t, 1, lane)];
    A.l[t] = mats[mat_index3(mat, t, 2, lane)];
  }
  return A;
}

__device__ __forceinline__ void mfma_group3(const AOps3& A, const float (&x)[8], f4& re, f4& im) {
  b8 bh, bm, bl;
  split8_bf16(x, bh, bm, bl);
  f4 acc[2];
#pragma unroll
  for (int t = 0; t < 2; t++) {
    f4 c = {0.f, 0.f, 0.f, 0.f};
    // the 2^-16 terms, the two 2^-8 terms, then hi*hi
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.h[t], bl, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.m[t], bm, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.l[t], bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.h[t], bm, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.m[t], bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.h[t], bh, c, 0, 0, 0);
    acc[t] = c;
  }
  re = acc[0];
  im = acc[1];
}

// One group: 16 butterflies (B columns), operands already in registers and
// already scaled into FP16 range.
__device__ __forceinline__ void mfma_group(const AOps& A, const float (&x)[8], f4& re, f4& im) {
  h8 bh, bl;
  split8(x, bh, bl);
  // small cross terms first, the large hi*hi product last
  f4 z = {0.f, 0.f, 0.f, 0.f};
  re = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.a0h, bl, z, 0, 0, 0);
  re = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.a0l, bh, re, 0, 0, 0);
  re = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.a0h, bh, re, 0, 0, 0);
  im = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.a1h, bl, z, 0, 0, 0);
  im = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.a1l, bh, im, 0, 0, 0);
  im = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.a1h, bh, im, 0, 0, 0);
}

// 16 transforms of 256 points per work-group of 256 threads (4 waves).  The
// tile lives in LDS as separate re / im planes so each lane reads exactly the
// components its B-operand quarter needs.
template <bool BF>
__global__ void __launch_bounds__(256) mfma_dft256_kernel(MfmaParams p) {
  __shared__ float lre[kMCT * kMLP];
  __shared__ float lim[kMCT * kMLP];
  __shared__ float smax[4][kMCT];  // per-wave partial max |x| of each transform
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int tile = blockIdx.x;
  const int bidx = tile / p.col_tiles;
  const int col0 = (tile - bidx * p.col_tiles) * kMCT;
  const int bhi = bidx / p.n_blo;
  const int blo = bidx - bhi * p.n_blo;
  const idx_t ib = (idx_t)bhi * p.in_bs + (idx_t)blo * p.in_bs2;
  const idx_t ob = (idx_t)bhi * p.out_bs + (idx_t)blo * p.out_bs2;

  // stage-1 A operands issued before the tile load so their latency overlaps
  // it; stage 2's are issued once the tile is in LDS (overlapping stage 1);
  // the BF16 operands (three parts) are loaded per group instead
  const b8* mats3 = reinterpret_cast<const b8*>(p.mats);
  AOps A1{};
  if constexpr (!BF) A1 = load_aops(p.mats, 0, lane);

  // load: 16 x 256 complex, one transform per thread and 16 of its points;
  // columns contiguous in memory: 16 lanes read one 128-byte row segment;
  // transforms contiguous: 16 lanes read 16 consecutive points of one
  const int lc = p.in_cfast ? (tid & 15) : (tid >> 4);
  const int lj0 = p.in_cfast ? (tid >> 4) : (tid & 15);
  cx<float> v[kMCT];
#pragma unroll
  for (int it = 0; it < kMCT; it++) {
    const int j = lj0 + 16 * it;
    v[it] = {0.f, 0.f};
    if (col0 + lc < p.n_col)
      v[it] = ld_elem<float>(p.in, nullptr, 1, ib + (idx_t)(col0 + lc) * p.in_cs + (idx_t)j * p.in_js, p.in_nt);
  }
  float mx = 0.f;
#pragma unroll
  for (int it = 0; it < kMCT; it++) mx = fmaxf(mx, fmaxf(fabsf(v[it].x), fabsf(v[it].y)));
  if (p.in_cfast) {
    // the 4 lanes of this wave that hold transform lc: lanes lc, lc+16, lc+32, lc+48
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    if (lane < 16) smax[wave][lc] = mx;
  } else {
    // the 16 lanes 16k..16k+15 hold transform lc
    mx = fmaxf(mx, __shfl_xor(mx, 1));
    mx = fmaxf(mx, __shfl_xor(mx, 2));
    mx = fmaxf(mx, __shfl_xor(mx, 4));
    mx = fmaxf(mx, __shfl_xor(mx, 8));
    if ((tid & 15) < 4) smax[tid & 15][lc] = mx;
  }
#pragma unroll
  for (int it = 0; it < kMCT; it++) {
    const int j = lj0 + 16 * it;
    const cx<float> z = p.swap_in ? swapri(v[it]) : v[it];
    lre[lc * kMLP + j] = z.x;
    lim[lc * kMLP + j] = z.y;
  }
  __syncthreads();
  AOps A2[4];
  if constexpr (!BF) {
#pragma unroll
    for (int g = 0; g < 4; g++) A2[g] = load_aops(p.mats, 1 + wave * 4 + g, lane);
  }
  auto shift_of = [&](int c) -> int {
    return transform_shift(fmaxf(fmaxf(smax[0][c], smax[1][c]), fmaxf(smax[2][c], smax[3][c])));
  };

  const int n = lane & 15;   // B column within the group
  const int q = lane >> 4;   // quarter of K
  const int rb = 8 * (q & 1);
  const float* plane = q >= 2 ? lim : lre;

  // stage 1: wave w handles transforms 4w..4w+3; group = transform (scaled
  // into FP16 range on the way in), B column = butterfly j
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const int c = wave * 4 + g;
    const int j = n;
    const float sc = ldexpf(1.f, shift_of(c));
    float x[8];
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = plane[c * kMLP + j + 16 * (rb + i)] * sc;
    f4 re, im;
    if constexpr (BF)
      mfma_group3(load_aops3(mats3, 0, lane), x, re, im);
    else
      mfma_group(A1, x, re, im);
    // the MFMA consumed every lane's loads, so the column can be overwritten
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
      const int m = 4 * q + jj;
      lre[c * kMLP + 16 * j + m] = re[jj];
      lim[c * kMLP + 16 * j + m] = im[jj];
    }
  }
  __syncthreads();

  // stage 2: wave w handles butterflies j = 4w..4w+3 of all 16 transforms;
  // group = j (one twiddled matrix), B column = transform c (already scaled)
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const int j = wave * 4 + g;
    const int c = n;
    float x[8];
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = plane[c * kMLP + j + 16 * (rb + i)];
    f4 re, im;
    if constexpr (BF)
      mfma_group3(load_aops3(mats3, 1 + wave * 4 + g, lane), x, re, im);
    else
      mfma_group(A2[g], x, re, im);
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
      const int m = 4 * q + jj;
      lre[c * kMLP + j + 16 * m] = re[jj];
      lim[c * kMLP + j + 16 * m] = im[jj];
    }
  }
  __syncthreads();

  // store, undoing the transform's scale (optional four-step twiddle
  // W_N^(n2 * k1), n2 = column, k1 = bin)
  const int sc_ = p.out_cfast ? (tid & 15) : (tid >> 4);
  const int sk0 = p.out_cfast ? (tid >> 4) : (tid & 15);
  const float inv = ldexpf(1.f, -shift_of(sc_));
  const unsigned tmask = (1u << p.tw4_bits) - 1u;
#pragma unroll
  for (int it = 0; it < kMCT; it++) {
    const int k = sk0 + 16 * it;
    v[it] = {lre[sc_ * kMLP + k] * inv, lim[sc_ * kMLP + k] * inv};
  }
  if (col0 + sc_ >= p.n_col) return;
#pragma unroll
  for (int it = 0; it < kMCT; it++) {
    const int k = sk0 + 16 * it;
    cx<float> z = v[it];
    if (p.tw4_hi) {
      const unsigned mi = (unsigned)((col0 + sc_) >> p.tw4_shift) * (unsigned)k;  // < N <= 2^24
      z = cmul(z, p.tw4_log2n > 0 ? tw_pow2f(mi, p.tw4_log2n) : tw4_lookup32(p.tw4_hi, p.tw4_lo, p.tw4_bits, tmask, mi));
    }
    if (p.swap_out) z = swapri(z);
    st_elem<float>(p.out, nullptr, 1, ob + (idx_t)(col0 + sc_) * p.out_cs + (idx_t)k * p.out_ks, z, p.out_nt);
  }
}

// Layout self-test: D = A * B for A 16 x 32 (rows 16), B 32 x 16 with small
// integers packed the way mfma_group packs them, through both the FP16 and the
// BF16 instruction; D read back with the assumed result layout.  out[0] receives the number of mismatching entries.
__global__ void mfma_selftest_kernel(int* out) {
  const int lane = threadIdx.x;
  h8 a, b;
  for (int i = 0; i < 8; i++) {
    const int k = 8 * (lane >> 4) + i;
    const int m = lane & 15;
    a[i] = (_Float16)(float)(((m * 7 + k * 3) % 11) - 5);
    b[i] = (_Float16)(float)(((k * 5 + m * 2) % 13) - 6);  // here m is the B column
  }
  b8 ab, bb;
  for (int i = 0; i < 8; i++) {
    ab[i] = (__bf16)(float)a[i];
    bb[i] = (__bf16)(float)b[i];
  }
  f4 z = {0.f, 0.f, 0.f, 0.f};
  const f4 d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, z, 0, 0, 0);
  const f4 db = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ab, bb, z, 0, 0, 0);
  int bad = 0;
  for (int jj = 0; jj < 4; jj++) {
    const int row = 4 * (lane >> 4) + jj;
    const int col = lane & 15;
    float ref = 0.f;
    for (int k = 0; k < 32; k++)
      ref += (float)(((row * 7 + k * 3) % 11) - 5) * (float)(((k * 5 + col * 2) % 13) - 6);
    if (d[jj] != ref) bad++;
    if (db[jj] != ref) bad++;
  }
  atomicAdd(out, bad);
}

hipError_t gen_mfma_mats(void* out, int bf16, hipStream_t st) {
  if (bf16)
    hipLaunchKernelGGL(gen_mfma_mats_bf16_kernel, dim3((kMats * 2 * 64 + 255) / 256), dim3(256), 0, st, (b8*)out);
  else
    hipLaunchKernelGGL(gen_mfma_mats_kernel, dim3((kMats * 2 * 64 + 255) / 256), dim3(256), 0, st, (h8*)out);
  return hipGetLastError();
}
size_t mfma_mats_bytes(int bf16) { return sizeof(h8) * (size_t)kMats * 2 * (bf16 ? 3 : 2) * 64; }

hipError_t launch_mfma_dft256(const MfmaParams& p, int n_tiles, hipStream_t st) {
  if (n_tiles <= 0) return hipSuccess;
  if (p.bf16)
    hipLaunchKernelGGL(mfma_dft256_kernel<true>, dim3(n_tiles), dim3(256), 0, st, p);
  else
    hipLaunchKernelGGL(mfma_dft256_kernel<false>, dim3(n_tiles), dim3(256), 0, st, p);
  return hipGetLastError();
}

}  // namespace afgpu

extern "C" int afft_gpu_mfma_selftest(void) {
  int* d = nullptr;
  if (hipMalloc(&d, sizeof(int)) != hipSuccess) return -1;
  int h = -1;
  hipError_t e = hipMemset(d, 0, sizeof(int));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(afgpu::mfma_selftest_kernel, dim3(1), dim3(64), 0, nullptr, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof(int), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? h : -1;
}
