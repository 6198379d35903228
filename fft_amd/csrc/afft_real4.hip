// afft_real4.hip -- 1-D FP32 r2c and c2r of a long real sequence as a
// real-input four-step (N = n1 * n2 reals, two passes over half-size data;
// see afft_real4.h).
//
// afft computes its r2c as an N/2-point complex FFT of the packed sequence
// x[2j] + i x[2j+1] followed by a real post-pass pairing bins k and N/2 - k
// (onedim.hpp:1031-1108, 1175-1189).  On the GPU that is three round trips
// through memory for long N (the two four-step passes and the post-pass, which
// needs the bins of two different rows).  Here the real input enters the
// four-step directly:
//   pass A  reads two adjacent real columns 2c, 2c+1 of the n1 x n2 matrix as
//           one complex column (they are adjacent floats), runs its n1-point
//           DFT Z, separates the two real columns' spectra through LDS,
//             A[k1] = (Z[k1] + conj Z[n1 - k1]) / 2,
//             B[k1] = (Z[k1] - conj Z[n1 - k1]) / (2i),
//           multiplies by W_N^(2c*k1) and W_N^((2c+1)*k1) and stores rows
//           k1 = 0..n1/2 only (the others are their conjugates);
//   pass B  runs the n2-point DFTs of those rows and stores, through LDS,
//           X[k1 + n1*k2] for the bins up to N/2 and, for 0 < k1 < n1/2, the
//           bins of row n1 - k1 as conj X[k1 + n1*(n2 - 1 - k2)].
// The c2r runs the transposed arrangement backwards (passes B' and A' below).
// Memory traffic: N reals in, N/2 complex through scratch, N/2 + 1 complex
// out -- two round trips instead of three.
#include "afft_real4.h"

namespace afgpu {

// Work-group ids in XCD-contiguous order: the hardware deals consecutive
// work-groups to the 8 XCDs in turn; this mapping gives XCD x a contiguous
// range of tiles, so neighbouring row tiles -- which share the cache lines at
// their edges when a transform's output does not start on a line boundary --
// run on one XCD and meet in its L2.
__device__ __forceinline__ unsigned xcd_contiguous(unsigned id, unsigned n) {
  const unsigned per = n / 8, rem = n % 8, x = id % 8;
  return x * per + (x < rem ? x : rem) + id / 8;
}

// The four-step twiddle W_N^e: computed (FP32, tw4_log2n > 0) or from the
// two-level table.
template <typename T>
__device__ __forceinline__ cx<T> real4_tw(const Real4Args<T>& a, unsigned mask, unsigned e) {
  if constexpr (std::is_same_v<T, float>) {
    if (a.tw4_log2n > 0) return tw_pow2f(e, a.tw4_log2n);
  }
  return tw4_lookup32(a.tw4_hi, a.tw4_lo, a.tw4_bits, mask, e);
}

template <typename T, int L, int CT, int SV, bool HX>
struct R4AShape {
  using PS = PassShape<L, CT, SV>;
  static constexpr int E = PS::E;
  static constexpr int TPT = PS::TPT;
  static constexpr int NT = PS::NT;
  static constexpr size_t lds = PS::template lds_bytes<T, HX>();
  static constexpr int groups = 2 * lds <= 160 * 1024 ? 2 : 1;
  static constexpr int min_waves = groups * NT / 256 > 1 ? groups * NT / 256 : 1;
};

template <typename T, int L, int CT, int SV, bool HX>
struct R4BShape {
  using PS = PassShape<L, CT, SV, HX ? (int)sizeof(T) : 2 * (int)sizeof(T)>;
  static constexpr int E = PS::E;
  static constexpr int TPT = PS::TPT;
  static constexpr int NT = PS::NT;
  static constexpr int LL = PS::LL;
  static constexpr size_t lds = PS::template lds_bytes<T, HX>();
  static constexpr int groups = 2 * lds <= 160 * 1024 ? 2 : 1;
  static constexpr int min_waves = groups * NT / 256 > 1 ? groups * NT / 256 : 1;
};

// Pass A: CT packed columns (2 CT real columns) of L = n1 points per group.
template <typename T, int L, int CT, int SV, bool HX>
__global__ void __launch_bounds__((R4AShape<T, L, CT, SV, HX>::NT), (R4AShape<T, L, CT, SV, HX>::min_waves))
    real4a_kernel(Real4Args<T> a) {
  using S = R4AShape<T, L, CT, SV, HX>;
  constexpr int E = S::E;
  constexpr int TPT = S::TPT;
  constexpr int MH = E / 2;  // registers m <= MH hold bins k1 <= L/2 (m == MH: thread 0 only)
  extern __shared__ __align__(16) unsigned char afgpu_smem[];
  cx<T>* __restrict__ lds = reinterpret_cast<cx<T>*>(afgpu_smem);
  T* __restrict__ ldsf = reinterpret_cast<T*>(afgpu_smem);
  const int tid = threadIdx.x;
  const int c = tid % CT, t = tid / CT;
  const idx_t npc = a.n2 / 2;  // packed columns per transform
  const int col_tiles = (int)(npc / CT);
  const idx_t b = blockIdx.x / col_tiles;
  const idx_t col = (idx_t)(blockIdx.x - b * col_tiles) * CT + c;
  auto lat = [&](int i) -> cx<T>& { return lds[lds_pad(i) * CT + c]; };
  auto latf = [&](int i) -> T& { return ldsf[lds_pad(i) * CT + c]; };

  // element j1 of packed column col: the reals x[j1*n2 + 2col], x[j1*n2 + 2col + 1]
  const T* src = a.in + b * a.in_bs + 2 * col;
  cx<T> v[E];
  static_for<E>([&](auto mc) AFGPU_INLINE {
    constexpr int m = decltype(mc)::value;
    v[m] = ld_elem<T>(src, nullptr, 1, (idx_t)(t + m * TPT) * npc, a.in_nt);
  });
  if constexpr (HX)
    run_stages<T, L, SV, 0, true, CT>(v, t, latf, a.stage_tw);
  else
    run_stages<T, L, SV, 0, false, CT>(v, t, lat, a.stage_tw);

  // partner Z[n1 - k1] of every bin k1 = t + m*TPT <= L/2
  cx<T> zp[MH + 1];
  auto partner = [&](int m) AFGPU_INLINE { return (L - (t + m * TPT)) & (L - 1); };
  if constexpr (!HX) {
    static_for<E>([&](auto mc) AFGPU_INLINE { lat(t + decltype(mc)::value * TPT) = v[decltype(mc)::value]; });
    __syncthreads();
    static_for<MH + 1>([&](auto mc) AFGPU_INLINE {
      constexpr int m = decltype(mc)::value;
      zp[m] = lat(partner(m));
    });
  } else {
    static_for<E>([&](auto mc) AFGPU_INLINE { latf(t + decltype(mc)::value * TPT) = v[decltype(mc)::value].x; });
    __syncthreads();
    static_for<MH + 1>([&](auto mc) AFGPU_INLINE {
      constexpr int m = decltype(mc)::value;
      zp[m].x = latf(partner(m));
    });
    __syncthreads();
    static_for<E>([&](auto mc) AFGPU_INLINE { latf(t + decltype(mc)::value * TPT) = v[decltype(mc)::value].y; });
    __syncthreads();
    static_for<MH + 1>([&](auto mc) AFGPU_INLINE {
      constexpr int m = decltype(mc)::value;
      zp[m].y = latf(partner(m));
    });
  }

  // unpack, twiddle, store rows k1 <= n1/2 of Y (columns 2col, 2col + 1)
  const unsigned mask = (1u << a.tw4_bits) - 1u;
  cx<T>* dst = a.scr + b * a.scr_bs + 2 * col;
  const T h = T(0.5);
  static_for<MH + 1>([&](auto mc) AFGPU_INLINE {
    constexpr int m = decltype(mc)::value;
    const int k1 = t + m * TPT;
    if (k1 <= L / 2) {
      const cx<T> zc = conjg(zp[m]);
      const cx<T> s = v[m] + zc, d = v[m] - zc;
      cx<T> A = {s.x * h, s.y * h};
      cx<T> B = mul_neg_i(cx<T>{d.x * h, d.y * h});
      const unsigned e0 = (unsigned)(2 * col) * (unsigned)k1;
      A = cmul(A, real4_tw(a, mask, e0));
      B = cmul(B, real4_tw(a, mask, e0 + (unsigned)k1));
      cx<T>* row = dst + (idx_t)k1 * a.n2;
      row[0] = A;
      row[1] = B;
    }
  });
}

template <typename T>
__device__ __forceinline__ void r4_store(const Real4Args<T>& a, idx_t e, cx<T> z) {
  if (a.out_il)
    st_elem<T>(a.out_re, nullptr, 1, e, z, a.out_nt);
  else
    st_elem<T>(a.out_re, a.out_im, 0, e, z, a.out_nt);
}

// Pass B: CT rows of L = n2 points per group (rows 0..n1/2), transposed store
// of the own and the mirrored bins.
template <typename T, int L, int CT, int SV, bool HX>
__global__ void __launch_bounds__((R4BShape<T, L, CT, SV, HX>::NT), (R4BShape<T, L, CT, SV, HX>::min_waves))
    real4b_kernel(Real4Args<T> a) {
  using S = R4BShape<T, L, CT, SV, HX>;
  constexpr int E = S::E;
  constexpr int TPT = S::TPT;
  constexpr int LL = S::LL;
  constexpr int MH = E / 2;  // registers m <= MH of the store phase hold bins k2 <= L/2
  extern __shared__ __align__(16) unsigned char afgpu_smem[];
  cx<T>* __restrict__ lds = reinterpret_cast<cx<T>*>(afgpu_smem);
  T* __restrict__ ldsf = reinterpret_cast<T*>(afgpu_smem);
  const int tid = threadIdx.x;
  const int c = tid / TPT, t = tid % TPT;
  const idx_t half = a.n1 / 2;
  const idx_t nrows = half + 1;
  const int row_tiles = (int)((nrows + CT - 1) / CT);
  const unsigned bid = xcd_contiguous(blockIdx.x, gridDim.x);
  const idx_t b = bid / row_tiles;
  const idx_t row0 = (idx_t)(bid - b * row_tiles) * CT;
  const idx_t row = row0 + c;
  auto lat = [&](int i) -> cx<T>& { return lds[c * LL + lds_pad(i)]; };
  auto latf = [&](int i) -> T& { return ldsf[c * LL + lds_pad(i)]; };

  cx<T> v[E];
  const cx<T>* src = a.scr + b * a.scr_bs + row * a.n2;
  static_for<E>([&](auto mc) AFGPU_INLINE {
    constexpr int m = decltype(mc)::value;
    v[m] = row < nrows ? src[t + m * TPT] : cx<T>{T(0), T(0)};
  });
  if constexpr (HX)
    run_stages<T, L, SV, 0, true, 1>(v, t, latf, a.stage_tw);
  else
    run_stages<T, L, SV, 0, false, 1>(v, t, lat, a.stage_tw);

  // store phase: lane c2 = tid % CT takes row row0 + c2, bins k2 = t2 + m*TPT
  const int c2 = tid % CT, t2 = tid / CT;
  const idx_t r2 = row0 + c2;
  const idx_t n1 = a.n1;
  const idx_t ob = b * a.out_bs;
  auto own = [&](int k2) AFGPU_INLINE { return r2 <= half && (k2 < L / 2 || (r2 == 0 && k2 == L / 2)); };
  auto mirror = [&](int k2) AFGPU_INLINE { return r2 > 0 && r2 < half && k2 < L / 2; };
  if constexpr (!HX) {
    static_for<E>([&](auto mc) AFGPU_INLINE { lat(t + decltype(mc)::value * TPT) = v[decltype(mc)::value]; });
    __syncthreads();
    static_for<MH + 1>([&](auto mc) AFGPU_INLINE {
      constexpr int m = decltype(mc)::value;
      const int k2 = t2 + m * TPT;
      if (own(k2)) r4_store(a, ob + r2 + n1 * k2, lds[c2 * LL + lds_pad(k2)]);
      if (mirror(k2)) r4_store(a, ob + (n1 - r2) + n1 * k2, conjg(lds[c2 * LL + lds_pad(L - 1 - k2)]));
    });
  } else {
    T xo[MH + 1], xm[MH + 1];
    static_for<E>([&](auto mc) AFGPU_INLINE { latf(t + decltype(mc)::value * TPT) = v[decltype(mc)::value].x; });
    __syncthreads();
    static_for<MH + 1>([&](auto mc) AFGPU_INLINE {
      constexpr int m = decltype(mc)::value;
      const int k2 = t2 + m * TPT;
      xo[m] = ldsf[c2 * LL + lds_pad(k2)];
      xm[m] = ldsf[c2 * LL + lds_pad(L - 1 - k2)];
    });
    __syncthreads();
    static_for<E>([&](auto mc) AFGPU_INLINE { latf(t + decltype(mc)::value * TPT) = v[decltype(mc)::value].y; });
    __syncthreads();
    static_for<MH + 1>([&](auto mc) AFGPU_INLINE {
      constexpr int m = decltype(mc)::value;
      const int k2 = t2 + m * TPT;
      if (own(k2)) r4_store(a, ob + r2 + n1 * k2, cx<T>{xo[m], ldsf[c2 * LL + lds_pad(k2)]});
      if (mirror(k2)) r4_store(a, ob + (n1 - r2) + n1 * k2, cx<T>{xm[m], -ldsf[c2 * LL + lds_pad(L - 1 - k2)]});
    });
  }
}

template <typename T>
__device__ __forceinline__ cx<T> r4_load_in(const Real4Args<T>& a, idx_t e) {
  return a.in_il ? ld_elem<T>(a.in, nullptr, 1, e, a.in_nt) : ld_elem<T>(a.in, a.in_im, 0, e, a.in_nt);
}

// c2r pass B': CT consecutive columns k1 of the spectrum viewed as n2 rows of
// n1 bins (bin k1 + n1*k2), columns k1 = 0..n1/2, L = n2 points each.  Bins
// past N/2 come from the mirrored column: for k2 >= n2/2 (k1 > 0) conj X at
// (n1 - k1) + n1*(n2 - 1 - k2), for column 0 conj X at n1*(n2 - k2); bins 0
// and N/2 keep their real parts.  Own and mirrored loads are runs of CT
// consecutive bins straight into registers.  Inverse DFT, times
// W_N^-(j2*k1), stored as Dt[j2][k1] (rows of rs entries).
template <typename T, int L, int CT, int SV, bool HX>
__global__ void __launch_bounds__((R4AShape<T, L, CT, SV, HX>::NT), (R4AShape<T, L, CT, SV, HX>::min_waves))
    real4c_kernel(Real4Args<T> a) {
  using S = R4AShape<T, L, CT, SV, HX>;
  constexpr int E = S::E;
  constexpr int TPT = S::TPT;
  extern __shared__ __align__(16) unsigned char afgpu_smem[];
  cx<T>* __restrict__ lds = reinterpret_cast<cx<T>*>(afgpu_smem);
  T* __restrict__ ldsf = reinterpret_cast<T*>(afgpu_smem);
  const int tid = threadIdx.x;
  const int c = tid % CT, t = tid / CT;
  const idx_t n1 = a.n1, half = n1 / 2;
  const int col_tiles = (int)((half + 1 + CT - 1) / CT);
  const unsigned bid = xcd_contiguous(blockIdx.x, gridDim.x);
  const idx_t b = bid / col_tiles;
  const idx_t k1 = (idx_t)(bid - b * col_tiles) * CT + c;
  const bool live = k1 <= half;
  const idx_t ib = b * a.in_bs;
  auto lat = [&](int i) -> cx<T>& { return lds[lds_pad(i) * CT + c]; };
  auto latf = [&](int i) -> T& { return ldsf[lds_pad(i) * CT + c]; };

  cx<T> v[E];
  static_for<E>([&](auto mc) AFGPU_INLINE {
    constexpr int m = decltype(mc)::value;
    const int k2 = t + m * TPT;
    cx<T> z = {T(0), T(0)};
    if (live) {
      if (m < E / 2 || (k1 == 0 && k2 == L / 2)) {
        z = r4_load_in(a, ib + k1 + n1 * k2);
        if (k1 == 0 && (k2 == 0 || k2 == L / 2)) z.y = T(0);  // bins 0 and N/2: real parts only
      } else if (k1 == 0) {
        z = conjg(r4_load_in(a, ib + n1 * (L - k2)));
      } else {
        z = conjg(r4_load_in(a, ib + (n1 - k1) + n1 * (L - 1 - k2)));
      }
    }
    v[m] = swapri(z);  // inverse DFT as swap(DFT(swap(.)))
  });
  if constexpr (HX)
    run_stages<T, L, SV, 0, true, CT>(v, t, latf, a.stage_tw);
  else
    run_stages<T, L, SV, 0, false, CT>(v, t, lat, a.stage_tw);
  if (live) {
    const unsigned mask = (1u << a.tw4_bits) - 1u;
    cx<T>* dst = a.scr + b * a.scr_bs + k1;
    static_for<E>([&](auto mc) AFGPU_INLINE {
      constexpr int m = decltype(mc)::value;
      const int j2 = t + m * TPT;
      const cx<T> wt = real4_tw(a, mask, (unsigned)j2 * (unsigned)k1);
      dst[(idx_t)j2 * a.rs] = cmulc(swapri(v[m]), wt);
    });
  }
}

// c2r pass A': CT packed columns c of L = n1 points per group, each the
// column spectrum of the real columns 2c and 2c+1 at once:
// Z[k1] = Dt[2c][k1] + i Dt[2c+1][k1] (k1 > n1/2 from the conjugates at
// n1 - k1), read along rows of Dt; inverse DFT; z[j1] = x[j1*n2 + 2c] +
// i x[j1*n2 + 2c + 1] stored through LDS so that a group writes runs of CT
// consecutive pairs of each output row.
template <typename T, int L, int CT, int SV, bool HX>
__global__ void __launch_bounds__((R4BShape<T, L, CT, SV, HX>::NT), (R4BShape<T, L, CT, SV, HX>::min_waves))
    real4d_kernel(Real4Args<T> a) {
  using S = R4BShape<T, L, CT, SV, HX>;
  constexpr int E = S::E;
  constexpr int TPT = S::TPT;
  constexpr int LL = S::LL;
  extern __shared__ __align__(16) unsigned char afgpu_smem[];
  cx<T>* __restrict__ lds = reinterpret_cast<cx<T>*>(afgpu_smem);
  T* __restrict__ ldsf = reinterpret_cast<T*>(afgpu_smem);
  const int tid = threadIdx.x;
  const int c = tid / TPT, t = tid % TPT;
  const idx_t npc = a.n2 / 2;
  const int col_tiles = (int)(npc / CT);
  const idx_t b = blockIdx.x / col_tiles;
  const idx_t c0 = (idx_t)(blockIdx.x - b * col_tiles) * CT;
  auto lat = [&](int i) -> cx<T>& { return lds[c * LL + lds_pad(i)]; };
  auto latf = [&](int i) -> T& { return ldsf[c * LL + lds_pad(i)]; };

  const cx<T>* ra = a.scr + b * a.scr_bs + 2 * (c0 + c) * a.rs;
  const cx<T>* rb = ra + a.rs;
  cx<T> v[E];
  static_for<E>([&](auto mc) AFGPU_INLINE {
    constexpr int m = decltype(mc)::value;
    const int k1 = t + m * TPT;
    const bool lo = k1 <= L / 2;
    const int kk = lo ? k1 : L - k1;
    cx<T> da = ra[kk], db = rb[kk];
    if (!lo) {
      da = conjg(da);
      db = conjg(db);
    }
    v[m] = cx<T>{da.y + db.x, da.x - db.y};  // swap(da + i db)
  });
  if constexpr (HX)
    run_stages<T, L, SV, 0, true, 1>(v, t, latf, a.stage_tw);
  else
    run_stages<T, L, SV, 0, false, 1>(v, t, lat, a.stage_tw);

  // v[m] = swap(z[j1]), j1 = t + m*TPT.  Store phase: lane c2 = tid % CT
  // takes packed column c0 + c2, j1 = t2 + m*TPT.
  const int c2 = tid % CT, t2 = tid / CT;
  T* dst = a.out_re + b * a.out_bs + 2 * (c0 + c2);
  if constexpr (!HX) {
    static_for<E>([&](auto mc) AFGPU_INLINE { lat(t + decltype(mc)::value * TPT) = swapri(v[decltype(mc)::value]); });
    __syncthreads();
    static_for<E>([&](auto mc) AFGPU_INLINE {
      constexpr int m = decltype(mc)::value;
      const int j1 = t2 + m * TPT;
      st_elem<T>(dst, nullptr, 1, (idx_t)j1 * npc, lds[c2 * LL + lds_pad(j1)], a.out_nt);
    });
  } else {
    T xr[E];
    static_for<E>([&](auto mc) AFGPU_INLINE { latf(t + decltype(mc)::value * TPT) = v[decltype(mc)::value].y; });
    __syncthreads();
    static_for<E>([&](auto mc) AFGPU_INLINE {
      constexpr int m = decltype(mc)::value;
      xr[m] = ldsf[c2 * LL + lds_pad(t2 + m * TPT)];
    });
    __syncthreads();
    static_for<E>([&](auto mc) AFGPU_INLINE { latf(t + decltype(mc)::value * TPT) = v[decltype(mc)::value].x; });
    __syncthreads();
    static_for<E>([&](auto mc) AFGPU_INLINE {
      constexpr int m = decltype(mc)::value;
      const int j1 = t2 + m * TPT;
      st_elem<T>(dst, nullptr, 1, (idx_t)j1 * npc, cx<T>{xr[m], ldsf[c2 * LL + lds_pad(j1)]}, a.out_nt);
    });
  }
}

// ---- launchers --------------------------------------------------------------
template <typename T, int L, int CT, int SV, bool HX>
static hipError_t launch_r4a(const Real4Args<T>& a, hipStream_t st) {
  using S = R4AShape<T, L, CT, SV, HX>;
  auto k = real4a_kernel<T, L, CT, SV, HX>;
  if (S::lds > 65536) {
    static bool attr = false;
    if (!attr) {
      hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)S::lds);
      if (e != hipSuccess) return e;
      attr = true;
    }
  }
  const idx_t npc = a.n2 / 2;
  if (npc % CT != 0 || a.n1 != L) return hipErrorInvalidValue;
  const idx_t blocks = a.nbatch * (npc / CT);
  if (blocks <= 0) return hipSuccess;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(S::NT), S::lds, st, a);
  return hipGetLastError();
}

template <typename T, int L, int CT, int SV, bool HX>
static hipError_t launch_r4b(const Real4Args<T>& a, hipStream_t st) {
  using S = R4BShape<T, L, CT, SV, HX>;
  auto k = real4b_kernel<T, L, CT, SV, HX>;
  if (S::lds > 65536) {
    static bool attr = false;
    if (!attr) {
      hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)S::lds);
      if (e != hipSuccess) return e;
      attr = true;
    }
  }
  if (a.n2 != L) return hipErrorInvalidValue;
  const idx_t blocks = a.nbatch * ((a.n1 / 2 + 1 + CT - 1) / CT);
  if (blocks <= 0) return hipSuccess;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(S::NT), S::lds, st, a);
  return hipGetLastError();
}

template <typename T, int L, int CT, int SV, bool HX>
static hipError_t launch_r4c(const Real4Args<T>& a, hipStream_t st) {
  using S = R4AShape<T, L, CT, SV, HX>;
  auto k = real4c_kernel<T, L, CT, SV, HX>;
  if (S::lds > 65536) {
    static bool attr = false;
    if (!attr) {
      hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)S::lds);
      if (e != hipSuccess) return e;
      attr = true;
    }
  }
  if (a.n2 != L || a.rs < a.n1 / 2 + 1) return hipErrorInvalidValue;
  const idx_t blocks = a.nbatch * ((a.n1 / 2 + 1 + CT - 1) / CT);
  if (blocks <= 0) return hipSuccess;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(S::NT), S::lds, st, a);
  return hipGetLastError();
}

template <typename T, int L, int CT, int SV, bool HX>
static hipError_t launch_r4d(const Real4Args<T>& a, hipStream_t st) {
  using S = R4BShape<T, L, CT, SV, HX>;
  auto k = real4d_kernel<T, L, CT, SV, HX>;
  if (S::lds > 65536) {
    static bool attr = false;
    if (!attr) {
      hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)S::lds);
      if (e != hipSuccess) return e;
      attr = true;
    }
  }
  const idx_t npc = a.n2 / 2;
  if (npc % CT != 0 || a.n1 != L || a.rs < a.n1 / 2 + 1) return hipErrorInvalidValue;
  const idx_t blocks = a.nbatch * (npc / CT);
  if (blocks <= 0) return hipSuccess;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(S::NT), S::lds, st, a);
  return hipGetLastError();
}

// Compiled shapes (length, tile, stage schedule), the default of each length
// first.  Column-mapped passes (A over n1-point columns of the reals, C over
// n2-point columns of the spectrum): tiles of CT columns, 16 giving 128-byte
// runs.  Row-mapped passes with a transposed store (B over n2-point rows, D
// over n1-point rows of the c2r scratch): tiles of CT rows.  All use the half
// exchange, so two groups of <= 80 KiB share a CU.
struct R4Shape {
  int L, CT, SV, HX;
};
// 2048-point columns: 16 per group for the r2c pass A (one group per CU,
// 128-byte segments: 16 x r2c 2^22 0.469 -> 0.427 ms)
#define AFGPU_R4A_SHAPES(X) X(128, 16, 0, 1) X(256, 16, 0, 1) X(512, 16, 1, 1) X(1024, 16, 1, 1) X(2048, 16, 1, 1) X(2048, 8, 1, 1)
// r2c row pass: 2048-point rows with full complex entries in LDS (one group
// per CU, no spills): 16 x 2^22 0.385 -> 0.350 ms, 32 x 2^21 0.391 -> 0.359 ms
#define AFGPU_R4B_SHAPES(X)                                                                                 \
  X(256, 16, 0, 1) X(512, 16, 0, 1) X(1024, 8, 0, 1) X(1024, 16, 1, 1) X(1024, 16, 0, 0) X(2048, 8, 0, 0)  \
  X(2048, 8, 0, 1) X(2048, 8, 1, 1) X(2048, 16, 1, 1) X(4096, 4, 0, 1)
// c2r spectrum pass: 8 columns per group at 1024 and 2048 points (64 x 2^20
// 0.237 -> 0.207 ms, 16 x 2^22 0.254 -> 0.228 ms against 16 columns)
#define AFGPU_R4C_SHAPES(X)                                                                                 \
  X(256, 16, 0, 1) X(512, 16, 1, 1) X(1024, 8, 1, 1) X(1024, 16, 1, 1) X(1024, 16, 0, 0) X(2048, 8, 1, 1)   \
  X(2048, 16, 1, 1) X(2048, 8, 0, 0) X(4096, 4, 0, 1)
#define AFGPU_R4D_SHAPES(X)                                                                                 \
  X(128, 16, 0, 1) X(256, 16, 0, 1) X(512, 16, 0, 1) X(1024, 8, 0, 1) X(1024, 16, 0, 0) X(1024, 16, 1, 1)  \
  X(2048, 8, 0, 0) X(2048, 8, 1, 1) X(2048, 16, 1, 1)
#define AFGPU_R4_ENTRY(LL, CTT, SVV, HXX) {LL, CTT, SVV, HXX},
constexpr R4Shape kR4A[] = {AFGPU_R4A_SHAPES(AFGPU_R4_ENTRY)};
constexpr R4Shape kR4B[] = {AFGPU_R4B_SHAPES(AFGPU_R4_ENTRY)};
constexpr R4Shape kR4C[] = {AFGPU_R4C_SHAPES(AFGPU_R4_ENTRY)};
constexpr R4Shape kR4D[] = {AFGPU_R4D_SHAPES(AFGPU_R4_ENTRY)};
#undef AFGPU_R4_ENTRY

template <size_t N>
static bool r4_pick(const R4Shape (&tab)[N], int L, int want_ct, int want_sv, int want_hx, int& ct, int& sv,
                    int& hx) {
  const R4Shape* first = nullptr;
  for (const R4Shape& s : tab) {
    if (s.L != L) continue;
    if (!first) first = &s;
    if ((want_ct <= 0 || s.CT == want_ct) && (want_sv < 0 || s.SV == want_sv) && (want_hx < 0 || s.HX == want_hx)) {
      ct = s.CT;
      sv = s.SV;
      hx = s.HX;
      return true;
    }
  }
  if (!first) return false;
  ct = first->CT;
  sv = first->SV;
  hx = first->HX;
  return true;
}

bool real4_shape(int pass, int elem_bytes, int len, int want_ct, int want_sv, int want_hx, int& ct, int& sv,
                 int& hx) {
  if (elem_bytes != 4) return false;
  switch (pass) {
    case 0: return r4_pick(kR4A, len, want_ct, want_sv, want_hx, ct, sv, hx);
    case 1: return r4_pick(kR4B, len, want_ct, want_sv, want_hx, ct, sv, hx);
    case 2: return r4_pick(kR4C, len, want_ct, want_sv, want_hx, ct, sv, hx);
    case 3: return r4_pick(kR4D, len, want_ct, want_sv, want_hx, ct, sv, hx);
    default: return false;
  }
}

#define AFGPU_R4_N1(LL, CTT, SVV, HXX) \
  if (a.n1 == LL && a.ct == CTT && a.sv == SVV && a.hx == HXX) return LAUNCH<float, LL, CTT, SVV, (HXX != 0)>(a, st);
#define AFGPU_R4_N2(LL, CTT, SVV, HXX) \
  if (a.n2 == LL && a.ct == CTT && a.sv == SVV && a.hx == HXX) return LAUNCH<float, LL, CTT, SVV, (HXX != 0)>(a, st);

template <>
hipError_t real4a<float>(const Real4Args<float>& a, hipStream_t st) {
#define LAUNCH launch_r4a
  AFGPU_R4A_SHAPES(AFGPU_R4_N1)
#undef LAUNCH
  return hipErrorInvalidValue;
}
template <>
hipError_t real4b<float>(const Real4Args<float>& a, hipStream_t st) {
#define LAUNCH launch_r4b
  AFGPU_R4B_SHAPES(AFGPU_R4_N2)
#undef LAUNCH
  return hipErrorInvalidValue;
}
template <>
hipError_t real4c<float>(const Real4Args<float>& a, hipStream_t st) {
#define LAUNCH launch_r4c
  AFGPU_R4C_SHAPES(AFGPU_R4_N2)
#undef LAUNCH
  return hipErrorInvalidValue;
}
template <>
hipError_t real4d<float>(const Real4Args<float>& a, hipStream_t st) {
#define LAUNCH launch_r4d
  AFGPU_R4D_SHAPES(AFGPU_R4_N1)
#undef LAUNCH
  return hipErrorInvalidValue;
}
#undef AFGPU_R4_N1
#undef AFGPU_R4_N2
#undef AFGPU_R4A_SHAPES
#undef AFGPU_R4B_SHAPES
#undef AFGPU_R4C_SHAPES
#undef AFGPU_R4D_SHAPES

template <>
hipError_t real4c<double>(const Real4Args<double>&, hipStream_t) {
  return hipErrorInvalidValue;
}
template <>
hipError_t real4d<double>(const Real4Args<double>&, hipStream_t) {
  return hipErrorInvalidValue;
}

template <>
hipError_t real4a<double>(const Real4Args<double>&, hipStream_t) {
  return hipErrorInvalidValue;
}
template <>
hipError_t real4b<double>(const Real4Args<double>&, hipStream_t) {
  return hipErrorInvalidValue;
}

}  // namespace afgpu
