// afft_aux.hip -- twiddle generation (FP64 on the device, rounded once), the
// stand-alone real pass for transforms too long to fuse it, the FP64 error
// reduction and a plain streaming copy used as the measured bandwidth ceiling.
#include <hip/hip_runtime.h>

#include "afft_aux.h"
#include "afft_device.h"

namespace afgpu {

// exp(-2 pi i e / M) evaluated as cospi/sinpi of the exact rational 2e/M (so
// multiples of M/8 come out exact), then rounded once to T.
__device__ __forceinline__ void wexp(idx_t e, idx_t M, double& c, double& s) {
  idx_t r = e % M;
  // reduce to [-M/2, M/2) so the argument of sincospi stays in [-1, 1)
  if (2 * r >= M) r -= M;
  double a = -2.0 * (double)r / (double)M;
  sincospi(a, &s, &c);
}

template <typename T>
__global__ void gen_stage_tw_kernel(cx<T>* out, StageList sl) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sl.total) return;
  // stage 0 (Ns = 1) multiplies by nothing and has no entries
  for (int s = 1; s < sl.nstage; s++) {
    int beg = sl.off[s], cnt = sl.ns[s] * (sl.radix[s] - 1);
    if (i >= beg && i < beg + cnt) {
      int rel = i - beg;
      int k = rel / (sl.radix[s] - 1);
      int r = rel % (sl.radix[s] - 1) + 1;
      double c, sn;
      wexp((idx_t)r * k, (idx_t)sl.ns[s] * sl.radix[s], c, sn);
      out[i] = cx<T>{(T)c, (T)sn};
      return;
    }
  }
}

template <typename T>
__global__ void gen_tw4_kernel(cx<T>* hi, cx<T>* lo, idx_t N, int bits) {
  idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
  idx_t nhi = N >> bits, nlo = idx_t(1) << bits;
  if (i < nhi) {
    double c, s;
    wexp(i << bits, N, c, s);
    hi[i] = cx<T>{(T)c, (T)s};
  }
  if (i < nlo) {
    // (cos - 1, sin) with cos - 1 = -2 sin^2(theta/2) to keep its relative accuracy
    idx_t r = i % N;
    double h = (double)r / (double)N;  // theta / (2 pi)
    double sh = sinpi(h);
    double s = sinpi(-2.0 * h);
    lo[i] = cx<T>{(T)(-2.0 * sh * sh), (T)s};
  }
}

template <typename T>
__global__ void gen_real_tw_kernel(cx<T>* out, idx_t L) {
  idx_t k = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= L) return;
  double c, s;
  wexp(k, 2 * L, c, s);
  out[k] = cx<T>{(T)c, (T)s};
}

template <typename T>
hipError_t gen_stage_tw(cx<T>* out, const StageList& sl, hipStream_t st) {
  if (sl.total == 0) return hipSuccess;
  hipLaunchKernelGGL(gen_stage_tw_kernel<T>, dim3((sl.total + 255) / 256), dim3(256), 0, st, out, sl);
  return hipGetLastError();
}
template <typename T>
hipError_t gen_tw4(cx<T>* hi, cx<T>* lo, idx_t N, int bits, hipStream_t st) {
  idx_t n = (N >> bits) > (idx_t(1) << bits) ? (N >> bits) : (idx_t(1) << bits);
  hipLaunchKernelGGL(gen_tw4_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, hi, lo, N, bits);
  return hipGetLastError();
}
template <typename T>
hipError_t gen_real_tw(cx<T>* out, idx_t L, hipStream_t st) {
  if (L <= 0) return hipSuccess;
  hipLaunchKernelGGL(gen_real_tw_kernel<T>, dim3((unsigned)((L + 255) / 256)), dim3(256), 0, st, out, L);
  return hipGetLastError();
}

// Real pass over rows of m complex entries (h + 1 rows per batch entry): row k
// is paired with row h - k.  Forward (r2c post-pass) and inverse (c2r
// pre-pass) follow onedim.hpp:1031-1108 (m = 1) and multi.hpp:343-414; src and
// dst may alias.  One thread per (batch, pair k, column c), c fastest.
template <typename T, bool INV>
__global__ void real_rows_kernel(RealRowsArgs<T> a) {
  const idx_t npair = a.h / 2 + 1;
  const idx_t total = a.nbatch * npair * a.m;
  const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const idx_t c = i % a.m;
  const idx_t k = (i / a.m) % npair;
  const idx_t b = i / (a.m * npair);
  const idx_t ib = b * a.in_bs + c * a.in_cs;
  const idx_t ob = b * a.out_bs + c * a.out_cs;
  auto ld = [&](idx_t row) -> cx<T> {
    idx_t e = ib + row * a.in_rs;
    if (a.in_il) return reinterpret_cast<const cx<T>*>(a.in_re)[e];
    return cx<T>{a.in_re[e], a.in_im[e]};
  };
  auto st = [&](idx_t row, cx<T> v) {
    idx_t e = ob + row * a.out_rs;
    if (a.out_il) {
      reinterpret_cast<cx<T>*>(a.out_re)[e] = v;
    } else {
      a.out_re[e] = v.x;
      a.out_im[e] = v.y;
    }
  };
  const idx_t h = a.h;
  if (k == 0) {
    if (INV) {
      cx<T> s0 = ld(0), sh = ld(h);
      st(0, cx<T>{s0.x + sh.x, s0.x - sh.x});
    } else {
      cx<T> z = ld(0);
      st(0, cx<T>{z.x + z.y, T(0)});
      st(h, cx<T>{z.x - z.y, T(0)});
    }
    return;
  }
  cx<T> s0 = ld(k), s1 = ld(h - k);
  cx<T> w = real_tw_pow2(a.real_tw, k, h);
  cx<T> s1c = conjg(s1);
  cx<T> d0, d1;
  if (INV) {
    cx<T> aa = s0 + s1c;
    cx<T> bb = cmulc(s1c - s0, w);
    d0 = aa + mul_neg_i(bb);
    d1 = conjg(aa) + mul_neg_i(conjg(bb));
  } else {
    const T hf = T(0.5);
    cx<T> aa = {(s0.x + s1c.x) * hf, (s0.y + s1c.y) * hf};
    cx<T> bw = cmul(s0 - s1c, w);
    cx<T> bb = {bw.x * hf, bw.y * hf};
    d0 = aa + mul_neg_i(bb);
    d1 = conjg(aa) + mul_neg_i(conjg(bb));
  }
  st(k, d0);
  if (h - k != k) st(h - k, d1);
}

template <typename T>
hipError_t real_rows(const RealRowsArgs<T>& a, bool inverse, hipStream_t st) {
  if (a.h <= 0 || a.m <= 0 || a.nbatch <= 0) return hipSuccess;
  const idx_t total = a.nbatch * (a.h / 2 + 1) * a.m;
  const idx_t blocks = (total + 255) / 256;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  if (inverse)
    hipLaunchKernelGGL((real_rows_kernel<T, true>), dim3((unsigned)blocks), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((real_rows_kernel<T, false>), dim3((unsigned)blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

// Transposing forward real pass.  A work-group takes a block of 32 x 128
// (k1, k2) entries (32 x 64 in FP64): it reads Y[k1][k2] and the partner
// Z[h - k] (row n1 - k1, column n2 - 1 - k2; row 0: column (n2 - k2) mod n2),
// both as 1 KiB row segments, forms X[k] = (Z[k] + conj Z[h-k])/2 - i (Z[k] - conj Z[h-k]) W_2h^k
// / 2 (onedim.hpp:1066-1106 arithmetic, W_2h^k = W_2h^k1 * W_2h^(n1 k2)), and
// writes the bins k1 + n1*k2 through LDS as 256-byte runs of consecutive k1.
// Bin 0 and bin h come from Z[0] (its block writes both).  Each thread holds
// 16 (FP64: 8) entries, so a work-group keeps 32 KiB of loads in flight.
//
// Block order.  The partner entries of tile t = i1 + nb1 * i2 (nb1 x nb2 tiles
// per transform) lie in its mirror tile T - 1 - t (rows n1 - k1 sit one row
// off the tile grid, so 31 of its 32 rows).  Work-groups are dispatched
// round-robin over the 8 XCDs, block B on XCD B % 8; blocks 16 j + x and
// 16 j + 8 + x run tile pair 8 j + x on the same XCD one slot apart, so the
// second of the two reads of every entry hits that XCD's L2.  The grid is
// rounded up to whole groups of 16; blocks past the last pair exit.
constexpr int kTrRows = 32;
template <typename T>
constexpr int tr_cols() {
  return sizeof(T) == 4 ? 128 : 64;  // 33 KiB of LDS either way
}

template <typename T>
__global__ void __launch_bounds__(256) real_tr_kernel(RealTrArgs<T> a) {
  constexpr int TB = kTrRows, TC = tr_cols<T>(), PER = TB * TC / 256;
  __shared__ cx<T> sx[TC][TB + 1];
  const int nb1 = a.n1 / TB, nb2 = a.n2 / TC;
  const int ntile = nb1 * nb2;
  idx_t b;
  int tile;
  if (ntile >= 2) {
    const idx_t B = blockIdx.x;
    const idx_t pair = (B >> 4) * 8 + (B & 7);
    const int half = ntile / 2;
    b = pair / half;
    if (b >= a.nbatch) return;
    const int pp = (int)(pair - b * half);
    tile = (B & 8) ? ntile - 1 - pp : pp;
  } else {
    b = blockIdx.x;
    tile = 0;
  }
  const int a0 = (tile % nb1) * TB, c0 = (tile / nb1) * TC;
  const cx<T>* Y = a.in + b * a.in_bs;
  const idx_t h = a.h;
  // this thread's column is fixed (256 is a multiple of TC): one twiddle
  // factor W_2h^(n1 k2) and one partner column per thread
  const int q = threadIdx.x % TC;
  const int k2 = c0 + q;
  const cx<T> w2 = a.real_tw[(idx_t)a.n1 * k2];
  const int p2n = a.n2 - 1 - k2;              // partner column, rows k1 > 0
  const int p20 = (a.n2 - k2) & (a.n2 - 1);   // partner column, row 0
#pragma unroll
  for (int i = 0; i < PER; i++) {
    const int r = threadIdx.x / TC + (256 / TC) * i;
    const int k1 = a0 + r;
    const cx<T> z = Y[(idx_t)k1 * a.n2 + k2];
    const int p1 = k1 == 0 ? 0 : a.n1 - k1;
    const cx<T> zp = Y[(idx_t)p1 * a.n2 + (k1 == 0 ? p20 : p2n)];
    cx<T> x;
    if (k1 == 0 && k2 == 0) {
      x = cx<T>{z.x + z.y, T(0)};
      const cx<T> nyq{z.x - z.y, T(0)};
      const idx_t o = b * a.out_bs + h;
      if (a.out_il) {
        reinterpret_cast<cx<T>*>(a.out_re)[o] = nyq;
      } else {
        a.out_re[o] = nyq.x;
        a.out_im[o] = nyq.y;
      }
    } else {
      const cx<T> w = cmul(a.real_tw[k1], w2);
      const cx<T> s1c = conjg(zp);
      const T hf = T(0.5);
      const cx<T> aa = {(z.x + s1c.x) * hf, (z.y + s1c.y) * hf};
      const cx<T> bw = cmul(z - s1c, w);
      const cx<T> bb = {bw.x * hf, bw.y * hf};
      x = aa + mul_neg_i(bb);
    }
    sx[q][r] = x;
  }
  __syncthreads();
  const idx_t ob = b * a.out_bs + a0 + (idx_t)a.n1 * c0;
#pragma unroll
  for (int i = 0; i < PER; i++) {
    const int e = threadIdx.x + 256 * i;
    const int r = e % TB, qq = e / TB;
    const idx_t o = ob + r + (idx_t)a.n1 * qq;
    const cx<T> x = sx[qq][r];
    if (a.out_il) {
      reinterpret_cast<cx<T>*>(a.out_re)[o] = x;
    } else {
      a.out_re[o] = x.x;
      a.out_im[o] = x.y;
    }
  }
}

template <typename T>
hipError_t real_tr(const RealTrArgs<T>& a, hipStream_t st) {
  constexpr int TC = tr_cols<T>();
  if (a.n1 % kTrRows || a.n2 % TC || a.n1 < kTrRows || a.n2 < TC) return hipErrorInvalidValue;
  const idx_t ntile = (idx_t)(a.n1 / kTrRows) * (a.n2 / TC);
  if (ntile > 1 && (ntile & 1)) return hipErrorInvalidValue;
  // tile pairs in groups of 8 (16 blocks), one block per tile when ntile = 1
  const idx_t blocks = ntile >= 2 ? (a.nbatch * (ntile / 2) + 7) / 8 * 16 : a.nbatch;
  if (blocks <= 0) return hipSuccess;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(real_tr_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

// Single-value real transform (total size 1): X[0] = (x, 0) / x = Re X[0].
template <typename T>
__global__ void real_one_kernel(const T* src, T* dst_re, T* dst_im, int inverse, int il, idx_t nb, idx_t in_s,
                                idx_t out_s) {
  idx_t b = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  if (!inverse) {
    T v = src[b * in_s];
    if (il) {
      dst_re[2 * b * out_s] = v;
      dst_re[2 * b * out_s + 1] = T(0);
    } else {
      dst_re[b * out_s] = v;
      dst_im[b * out_s] = T(0);
    }
  } else {
    dst_re[b * out_s] = il ? src[2 * b * in_s] : src[b * in_s];
  }
}

template <typename T>
hipError_t real_one(const T* src, T* dst_re, T* dst_im, int inverse, int il, idx_t nb, idx_t in_s, idx_t out_s,
                    hipStream_t st) {
  hipLaunchKernelGGL(real_one_kernel<T>, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, src, dst_re, dst_im,
                     inverse, il, nb, in_s, out_s);
  return hipGetLastError();
}

// sum (y - r)^2 and sum r^2 in FP64; acc[0..1] must be zero on entry.
template <typename TY>
__global__ void err_kernel(const TY* __restrict__ y, const double* __restrict__ r, idx_t n, double* acc) {
  double num = 0.0, den = 0.0;
  for (idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (idx_t)gridDim.x * blockDim.x) {
    double d = (double)y[i] - r[i];
    num += d * d;
    den += r[i] * r[i];
  }
  __shared__ double sn[256], sd[256];
  sn[threadIdx.x] = num;
  sd[threadIdx.x] = den;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sn[threadIdx.x] += sn[threadIdx.x + s];
      sd[threadIdx.x] += sd[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&acc[0], sn[0]);
    atomicAdd(&acc[1], sd[0]);
  }
}

__global__ void zero2_kernel(double* acc) {
  if (threadIdx.x < 2) acc[threadIdx.x] = 0.0;
}

template <typename TY>
hipError_t err_sums(const TY* y, const double* r, idx_t n, double* acc, hipStream_t st) {
  hipLaunchKernelGGL(zero2_kernel, dim3(1), dim3(64), 0, st, acc);
  idx_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(err_kernel<TY>, dim3((unsigned)blocks), dim3(256), 0, st, y, r, n, acc);
  return hipGetLastError();
}

// Streaming copy with 16-byte accesses: the practical HBM ceiling a
// read-once / write-once kernel can reach on this part.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ void copy16_kernel(const u32x4* __restrict__ src, u32x4* __restrict__ dst, idx_t n16) {
  for (idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (idx_t)gridDim.x * blockDim.x)
    __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
}

hipError_t copy16(const void* src, void* dst, idx_t bytes, hipStream_t st) {
  idx_t n16 = bytes / 16;
  idx_t blocks = (n16 + 255) / 256;
  if (blocks > 256 * 64) blocks = 256 * 64;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (const u32x4*)src, (u32x4*)dst, n16);
  return hipGetLastError();
}

template hipError_t gen_stage_tw<float>(cx<float>*, const StageList&, hipStream_t);
template hipError_t gen_stage_tw<double>(cx<double>*, const StageList&, hipStream_t);
template hipError_t gen_tw4<float>(cx<float>*, cx<float>*, idx_t, int, hipStream_t);
template hipError_t gen_tw4<double>(cx<double>*, cx<double>*, idx_t, int, hipStream_t);
template hipError_t gen_real_tw<float>(cx<float>*, idx_t, hipStream_t);
template hipError_t gen_real_tw<double>(cx<double>*, idx_t, hipStream_t);
template hipError_t real_rows<float>(const RealRowsArgs<float>&, bool, hipStream_t);
template hipError_t real_rows<double>(const RealRowsArgs<double>&, bool, hipStream_t);
template hipError_t real_tr<float>(const RealTrArgs<float>&, hipStream_t);
template hipError_t real_tr<double>(const RealTrArgs<double>&, hipStream_t);
template hipError_t real_one<float>(const float*, float*, float*, int, int, idx_t, idx_t, idx_t, hipStream_t);
template hipError_t real_one<double>(const double*, double*, double*, int, int, idx_t, idx_t, idx_t, hipStream_t);
template hipError_t err_sums<float>(const float*, const double*, idx_t, double*, hipStream_t);
template hipError_t err_sums<double>(const double*, const double*, idx_t, double*, hipStream_t);

}  // namespace afgpu
