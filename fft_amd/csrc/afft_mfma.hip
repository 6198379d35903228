// afft_mfma.hip -- split-precision MFMA path: 256-point DFT tiles whose two
// radix-16 stages run on the matrix cores.
//
// No counterpart in afft (its butterflies are FP32/FP64 SIMD code,
// onedim.hpp:55-133); the transform computed is the same DFT.
//
// A radix-16 stage over 16 butterflies is one real matrix product
//     [Yr; Yi] (32 x 16) = A (32 x 32) * [Xr; Xi] (32 x 16)
// with A = [[Re C, -Im C], [Im C, Re C]] and C the 16 x 16 complex stage
// matrix.  Operands are FP16 split in two: x*s = hi + lo (s a power of two per
// butterfly, chosen from the butterfly's max |x| so hi uses the full FP16
// range), A = A_hi + A_lo; the product keeps A_hi*hi + A_hi*lo + A_lo*hi (the
// dropped A_lo*lo term is ~2^-22 relative), accumulated in FP32 by
// v_mfma_f32_16x16x32_f16, then scaled back by 1/s exactly.  Stage 1 uses the
// plain DFT-16 matrix; stage 2 (Ns = 16) folds the Stockham twiddles
// W_256^(r*k) into 16 matrices C_k = F16 * diag(W_256^(r*k)), so no twiddle
// multiply runs on the VALU.
//
// Operand layouts (v_mfma_f32_16x16x32_f16, wave64): lane l holds row l%16 of
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



// Splits the 8 values of a lane (one quarter of a butterfly's 32 reals) into
// FP16 hi/lo after scaling by 2^sh, sh shared by the 4 lanes of the butterfly.
__device__ __forceinline__ int block_shift(const float (&x)[8]) {
  float mx = 0.f;
#pragma unroll
  for (int i = 0; i < 8; i++) mx = fmaxf(mx, fabsf(x[i]));
  mx = fmaxf(mx, __shfl_xor(mx, 16));
  mx = fmaxf(mx, __shfl_xor(mx, 32));
  if (!(mx > 0.f) || !isfinite(mx)) return 0;
  int e;
  (void)frexpf(mx, &e);  // mx in [2^(e-1), 2^e)
  return 14 - e;         // scaled max in [2^13, 2^14)
}

__device__ __forceinline__ void split8(const float (&x)[8], int sh, h8& hi, h8& lo) {
#pragma unroll
  for (int i = 0; i < 8; i++) {
    float v = ldexpf(x[i], sh);
    _Float16 h = (_Float16)v;
    hi[i] = h;
    lo[i] = (_Float16)(v - (float)h);
  }
}

struct AOps {
  h8 a0h, a0l, a1h, a1l;  // M-tile 0 (Re rows) and 1 (Im rows), hi and lo
};

__device__ __forceinline__ AOps load_aops(const h8* __restrict__ mats, int mat, int lane) {
  return {mats[mat_index(mat, 0, 0, lane)], mats[mat_index(mat, 0, 1, lane)], mats[mat_index(mat, 1, 0, lane)],
          mats[mat_index(mat, 1, 1, lane)]};
}

// One group: 16 butterflies (B columns), operands already in registers.
__device__ __forceinline__ void mfma_group(const AOps& A, const float (&x)[8], f4& re, f4& im) {
  const int sh = block_shift(x);
  h8 bh, bl;
  split8(x, sh, bh, bl);
  // small cross terms first, the large hi*hi product last
  f4 z = {0.f, 0.f, 0.f, 0.f};
  re = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.a0h, bl, z, 0, 0, 0);
  re = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.a0l, bh, re, 0, 0, 0);
  re = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.a0h, bh, re, 0, 0, 0);
  im = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.a1h, bl, z, 0, 0, 0);
  im = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.a1l, bh, im, 0, 0, 0);
  im = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.a1h, bh, im, 0, 0, 0);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    re[j] = ldexpf(re[j], -sh);
    im[j] = ldexpf(im[j], -sh);
  }
}

// 16 transforms of 256 points per work-group of 256 threads (4 waves).  The
// tile lives in LDS as separate re / im planes so each lane reads exactly the
// components its B-operand quarter needs.
__global__ void __launch_bounds__(256) mfma_dft256_kernel(MfmaParams p) {
  __shared__ float lre[kMCT * kMLP];
  __shared__ float lim[kMCT * kMLP];
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

  // the A operands of both stages, issued before the tile load so their
  // latency overlaps it
  const AOps A1 = load_aops(p.mats, 0, lane);
  AOps A2[4];
#pragma unroll
  for (int g = 0; g < 4; g++) A2[g] = load_aops(p.mats, 1 + wave * 4 + g, lane);

  // load: 16 x 256 complex, every load of the thread issued before the first
  // LDS write so the whole tile is in flight at once
  cx<float> v[kMCT];
#pragma unroll
  for (int it = 0; it < kMCT; it++) {
    const int c = p.in_cfast ? (tid & 15) : it;
    const int j = p.in_cfast ? (tid >> 4) + 16 * it : tid;
    v[it] = {0.f, 0.f};
    if (col0 + c < p.n_col)
      v[it] = ld_elem<float>(p.in, nullptr, 1, ib + (idx_t)(col0 + c) * p.in_cs + (idx_t)j * p.in_js, p.in_nt);
  }
#pragma unroll
  for (int it = 0; it < kMCT; it++) {
    const int c = p.in_cfast ? (tid & 15) : it;
    const int j = p.in_cfast ? (tid >> 4) + 16 * it : tid;
    const cx<float> z = p.swap_in ? swapri(v[it]) : v[it];
    lre[c * kMLP + j] = z.x;
    lim[c * kMLP + j] = z.y;
  }
  __syncthreads();

  const int n = lane & 15;   // B column within the group
  const int q = lane >> 4;   // quarter of K
  const int rb = 8 * (q & 1);
  const float* plane = q >= 2 ? lim : lre;

  // stage 1: wave w handles columns 4w..4w+3; group = column, B column = butterfly j
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const int c = wave * 4 + g;
    const int j = n;
    float x[8];
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = plane[c * kMLP + j + 16 * (rb + i)];
    f4 re, im;
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

  // stage 2: wave w handles butterflies j = 4w..4w+3 of all 16 columns;
  // group = j (one twiddled matrix), B column = transform c
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const int j = wave * 4 + g;
    const int c = n;
    float x[8];
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = plane[c * kMLP + j + 16 * (rb + i)];
    f4 re, im;
    mfma_group(A2[g], x, re, im);
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
      const int m = 4 * q + jj;
      lre[c * kMLP + j + 16 * m] = re[jj];
      lim[c * kMLP + j + 16 * m] = im[jj];
    }
  }
  __syncthreads();

  // store (optional four-step twiddle W_N^(n2 * k1), n2 = column, k1 = bin)
  const unsigned tmask = (1u << p.tw4_bits) - 1u;
#pragma unroll
  for (int it = 0; it < kMCT; it++) {
    const int c = p.out_cfast ? (tid & 15) : it;
    const int k = p.out_cfast ? (tid >> 4) + 16 * it : tid;
    v[it] = {lre[c * kMLP + k], lim[c * kMLP + k]};
  }
#pragma unroll
  for (int it = 0; it < kMCT; it++) {
    const int c = p.out_cfast ? (tid & 15) : it;
    const int k = p.out_cfast ? (tid >> 4) + 16 * it : tid;
    if (col0 + c >= p.n_col) continue;
    cx<float> z = v[it];
    if (p.tw4_hi) {
      const unsigned mi = (unsigned)((col0 + c) >> p.tw4_shift) * (unsigned)k;  // < N <= 2^24
      z = cmul(z, tw4_lookup32(p.tw4_hi, p.tw4_lo, p.tw4_bits, tmask, mi));
    }
    if (p.swap_out) z = swapri(z);
    st_elem<float>(p.out, nullptr, 1, ob + (idx_t)(col0 + c) * p.out_cs + (idx_t)k * p.out_ks, z, p.out_nt);
  }
}

// Layout self-test: D = A * B for A 16 x 32 (rows 16), B 32 x 16 with small
// integers packed the way mfma_group packs them; D read back with the assumed
// result layout.  out[0] receives the number of mismatching entries.
__global__ void mfma_selftest_kernel(int* out) {
  const int lane = threadIdx.x;
  h8 a, b;
  for (int i = 0; i < 8; i++) {
    const int k = 8 * (lane >> 4) + i;
    const int m = lane & 15;
    a[i] = (_Float16)(float)(((m * 7 + k * 3) % 11) - 5);
    b[i] = (_Float16)(float)(((k * 5 + m * 2) % 13) - 6);  // here m is the B column
  }
  f4 z = {0.f, 0.f, 0.f, 0.f};
  f4 d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, z, 0, 0, 0);
  int bad = 0;
  for (int jj = 0; jj < 4; jj++) {
    const int row = 4 * (lane >> 4) + jj;
    const int col = lane & 15;
    float ref = 0.f;
    for (int k = 0; k < 32; k++)
      ref += (float)(((row * 7 + k * 3) % 11) - 5) * (float)(((k * 5 + col * 2) % 13) - 6);
    if (d[jj] != ref) bad++;
  }
  atomicAdd(out, bad);
}

hipError_t gen_mfma_mats(void* out, hipStream_t st) {
  hipLaunchKernelGGL(gen_mfma_mats_kernel, dim3((kMats * 2 * 64 + 255) / 256), dim3(256), 0, st, (h8*)out);
  return hipGetLastError();
}
size_t mfma_mats_bytes() { return sizeof(h8) * (size_t)kMats * 2 * 2 * 64; }

hipError_t launch_mfma_dft256(const MfmaParams& p, int n_tiles, hipStream_t st) {
  if (n_tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(mfma_dft256_kernel, dim3(n_tiles), dim3(256), 0, st, p);
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
