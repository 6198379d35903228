// afft_pass.h -- the batched LDS Stockham pass kernel.
//
// One work-group computes CT independent L-point FFTs ("columns").  Element j of
// column c of batch entry b is read from
//     in + b*in_bs + c*in_cs + j*in_js            (complex-element strides)
// and output bin k is written to
//     out + b*out_bs + c*out_cs + k*out_ks.
// This one addressing form covers every pass the plans need:
//   - contiguous 1-D rows (the whole transform for L <= 16384, the last axis
//     of an N-D transform, pass B of a four-step transform),
//   - strided columns (outer axes of N-D transforms -- the job afft's
//     multi.hpp row passes do -- and pass A of a four-step transform),
//   - the real-transform packings (rows 2j / 2j+1 of dim 0 read as re / im,
//     fft_internal.hpp:326-330; N reals read as N/2 interleaved complex,
//     onedim.hpp:1175-1189).
//
// Thread mapping.  Each of the TPT = L/E threads of a column holds E elements
// j = t + m*TPT (m < E).  COLMAP=false puts t fastest across lanes (coalesced
// when in_js == 1); COLMAP=true puts c fastest (coalesced when in_cs == 1).
//
// Stages.  Stockham autosort: stage s of radix R with Ns = R0*...*R(s-1) reads
// entries j + r*L/R, multiplies by W_{Ns*R}^{r*(j mod Ns)}, runs the radix-R
// butterfly and writes entry (j/Ns)*Ns*R + (j mod Ns) + r*Ns.  The first stage
// reads straight from global memory and the last leaves bin t + m*TPT in
// register m, so only the L-1 interior exchanges go through LDS and the output
// is in natural order with no bit-reversal pass (afft needs one,
// onedim.hpp:397-447).
//
// Epilogues (EPI):
//   EPI_PLAIN  store.
//   EPI_TW4    multiply bin k1 of column n2 by W_N^(n2*k1) before the store
//              (pass A of the four-step N = N1*N2 split).
//   EPI_R2C    forward real post-pass fused after the last stage (pairs bin k
//              with L-k through LDS; writes L+1 bins): onedim.hpp:1031-1108 and
//              multi.hpp:343-414, inverse=false.
//   EPI_C2R    inverse real pre-pass fused into the load (reads bins k and L-k,
//              L+1 input bins): the inverse=true branch of the same functions.
//   EPI_C2R4   pass A of the four-step inverse c2c of a long 1-D c2r: the
//              inverse real pre-pass on the load, pairing input bin k (the
//              element's offset in its batch entry, col*in_cs + j*in_js) with
//              bin real_h - k, which lies in another column; then the EPI_TW4
//              twiddle on the store (onedim.hpp:1031-1108, inverse=true).
// STORE_T routes the final store through LDS so a row-mapped pass can write
// with columns fastest (pass B of a four-step transform writes bins k2*N1+k1).
#pragma once

#include "afft_device.h"

namespace afgpu {

enum { EPI_PLAIN = 0, EPI_TW4 = 1, EPI_R2C = 2, EPI_C2R = 3, EPI_C2R4 = 4 };

template <typename T>
struct PassParams {
  const T* in_re;
  const T* in_im;
  T* out_re;
  T* out_im;
  idx_t in_js, in_cs, in_bs;
  idx_t out_ks, out_cs, out_bs;
  idx_t in_bs2, out_bs2;  // stride of the low batch index (two-level batch)
  int n_blo;              // batch entries of the low index (1 = one-level)
  int n_col;              // columns per batch entry
  int col_tiles;          // ceil(n_col / CT)
  int n_tiles;            // col_tiles * n_batch
  int in_il, out_il;
  int swap_in, swap_out;
  int in_nt, out_nt;  // streaming (non-temporal) loads / stores
  int xcd_remap;      // give each XCD a contiguous run of tiles
  int mirror_pairs;   // EPI_C2R4: n_tiles counts whole groups of 16 blocks; blocks
                      // 16q + x and 16q + 8 + x run column tile c and its mirror
                      // col_tiles - 1 - c of one batch entry on the same XCD
  int n_batch;        // mirror_pairs: batch entries (tiles = n_batch * col_tiles)
  const cx<T>* stage_tw;  // per-length stage twiddles
  const cx<T>* tw4_hi;    // W_N^(i << lo_bits)
  const cx<T>* tw4_lo;    // W_N^j - 1, j < 2^lo_bits
  int tw4_lo_bits;
  int tw4_col_shift;      // n2 = column >> shift
  int tw4_blo;            // 1: n2 = low batch index instead (columns of one n2 are the c index)
  int tw4_log2n;          // > 0 (FP32): compute W_N^m, N = 2^tw4_log2n, instead of the table lookup
  const cx<T>* real_tw;   // W_{2L}^k, k < L (EPI_C2R4: W_{2h}^k, k < h)
  idx_t real_h;           // EPI_C2R4: h, the number of complex points
  // Byte extents of the in / out regions of one batch entry (fast path only).
  unsigned in_bytes, out_bytes;
  // Optional per-column buffers (uneven batches): column c reads from
  // col_ptrs[2c] and writes to col_ptrs[2c+1], both interleaved, with the
  // batch and column offsets taken as zero.
  const unsigned long long* col_ptrs;
};

template <typename T>
__device__ __forceinline__ cx<T> ld_elem(const T* __restrict__ re, const T* __restrict__ im, int il, idx_t e,
                                         int nt = 0) {
  if (nt) {
    if (il) {
      typename Vec2<T>::type w = __builtin_nontemporal_load(reinterpret_cast<const typename Vec2<T>::type*>(re) + e);
      return {w.x, w.y};
    }
    return {__builtin_nontemporal_load(re + e), __builtin_nontemporal_load(im + e)};
  }
  if (il) return reinterpret_cast<const cx<T>*>(re)[e];
  return {re[e], im[e]};
}
template <typename T>
__device__ __forceinline__ void st_elem(T* __restrict__ re, T* __restrict__ im, int il, idx_t e, cx<T> v,
                                        int nt = 0) {
  if (nt) {
    if (il) {
      typename Vec2<T>::type w = {v.x, v.y};
      __builtin_nontemporal_store(w, reinterpret_cast<typename Vec2<T>::type*>(re) + e);
    } else {
      __builtin_nontemporal_store(v.x, re + e);
      __builtin_nontemporal_store(v.y, im + e);
    }
    return;
  }
  if (il) {
    reinterpret_cast<cx<T>*>(re)[e] = v;
  } else {
    re[e] = v.x;
    im[e] = v.y;
  }
}

// Work-groups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  Remapping block b to tile (b % 8) * (n / 8) + b / 8 hands each XCD a
// contiguous run of tiles, so tiles that share cache lines meet in one L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
  if ((n & 7) != 0) return b;
  return (b & 7) * (n >> 3) + (b >> 3);
}

// W_N^m from the two-level table: hi[m >> b] * (1 + lo[m & mask]).
template <typename T>
__device__ __forceinline__ cx<T> tw4_lookup(const cx<T>* __restrict__ hi, const cx<T>* __restrict__ lo, int bits,
                                            idx_t m) {
  cx<T> h = hi[m >> bits];
  cx<T> l = lo[m & ((idx_t(1) << bits) - 1)];
  return {h.x + fma(h.x, l.x, -(h.y * l.y)), h.y + fma(h.x, l.y, h.y * l.x)};
}

// Same with a 32-bit index (m < 2^32) and a precomputed mask.
template <typename T>
__device__ __forceinline__ cx<T> tw4_lookup32(const cx<T>* __restrict__ hi, const cx<T>* __restrict__ lo, int bits,
                                              unsigned mask, unsigned m) {
  cx<T> h = hi[m >> bits];
  cx<T> l = lo[m & mask];
  return {h.x + fma(h.x, l.x, -(h.y * l.y)), h.y + fma(h.x, l.y, h.y * l.x)};
}

// The E registers of one thread.
template <typename T, int L, int SV>
using Regs = cx<T>[Sched<L, SV>::E];

// Step, in padded LDS entries, between index i and i + k*d when d is a
// multiple of 16 (lds_pad(i + 16q) = lds_pad(i) + 17q): 0 when d is not.
template <int D>
constexpr int pad_step() {
  return D % 16 == 0 ? D + D / 16 : 0;
}

// Stage S and the ones after it.  HX (half exchange): `lat` addresses one real
// per entry and each exchange between stages moves the real parts, then the
// imaginary parts, through the same LDS words -- half the LDS of a tile for
// two more barriers per exchange.
//
// ST > 0 is the distance between consecutive LDS indices of `lat` (1 for
// row-mapped tiles, CT for column-mapped ones).  Where an exchange touches
// index i0 + k*d with d a multiple of 16 (or i0 + r, r < R, with R and i0
// multiples of 16), the accesses are then written as one address plus
// constant offsets, which the LDS instructions take as immediates instead of
// one address register per element.
template <typename T, int L, int SV, int S, bool HX, int ST = 0, typename LdsAt>
__device__ __forceinline__ void run_stages(Regs<T, L, SV>& v, int t, LdsAt& lat, const cx<T>* __restrict__ tw) {
  constexpr int E = Sched<L, SV>::E;
  constexpr int TPT = L / E;
  constexpr int R = Sched<L, SV>::R[S];
  constexpr int NS = StageNs<L, SV, S>::value;
  constexpr int NB = E / R;
  constexpr bool LAST = (S == Sched<L, SV>::NST - 1);
  // entry offset of write r (from base) and of read m (from t), or -1 where
  // the padded index is not affine
  auto woff = [](int r) constexpr -> int {
    if constexpr (ST > 0 && pad_step<NS>() > 0) return r * pad_step<NS>() * ST;
    if constexpr (ST > 0 && NS == 1 && R % 16 == 0) return (r + r / 16) * ST;
    return -1;
  };
  constexpr bool WAFF = ST > 0 && (pad_step<NS>() > 0 || (NS == 1 && R % 16 == 0));
  constexpr bool RAFF = ST > 0 && pad_step<TPT>() > 0;
  static_for<NB>([&](auto bc) AFGPU_INLINE {
    constexpr int b = decltype(bc)::value;
    cx<T> a[R];
    static_for<R>([&](auto rc) AFGPU_INLINE { a[decltype(rc)::value] = v[b + decltype(rc)::value * NB]; });
    if constexpr (NS > 1) {
      const int j = t + b * TPT;
      const int k = j & (NS - 1);
      const cx<T>* tp = tw + StageTwOff<L, SV, S>::value + k * (R - 1);
      static_for<R - 1>([&](auto rc) AFGPU_INLINE {
        constexpr int r = decltype(rc)::value + 1;
        a[r] = cmul(a[r], tp[r - 1]);
      });
    }
    Dft<T, R>::run(a);
    static_for<R>([&](auto rc) AFGPU_INLINE { v[b + decltype(rc)::value * NB] = a[decltype(rc)::value]; });
  });
  if constexpr (!LAST) {
    // entry r of butterfly b, and entry t + m*TPT
    auto wref = [&](auto bc, auto rc) AFGPU_INLINE -> decltype(lat(0))& {
      constexpr int b = decltype(bc)::value;
      constexpr int r = decltype(rc)::value;
      const int j = t + b * TPT;
      const int base = (j / NS) * (NS * R) + (j & (NS - 1));
      if constexpr (WAFF)
        return (&lat(base))[woff(r)];
      else
        return lat(base + r * NS);
    };
    auto* const rbase = &lat(t);
    auto rref = [&](auto mc) AFGPU_INLINE -> decltype(lat(0))& {
      constexpr int m = decltype(mc)::value;
      if constexpr (RAFF)
        return rbase[m * pad_step<TPT>() * ST];
      else
        return lat(t + m * TPT);
    };
    if constexpr (!HX) {
      static_for<NB>([&](auto bc) AFGPU_INLINE {
        constexpr int b = decltype(bc)::value;
        static_for<R>([&](auto rc) AFGPU_INLINE { wref(bc, rc) = v[b + decltype(rc)::value * NB]; });
      });
      __syncthreads();
      static_for<E>([&](auto mc) AFGPU_INLINE { v[decltype(mc)::value] = rref(mc); });
      __syncthreads();
    } else {
      static_for<2>([&](auto hc) AFGPU_INLINE {
        constexpr int h = decltype(hc)::value;
        static_for<NB>([&](auto bc) AFGPU_INLINE {
          constexpr int b = decltype(bc)::value;
          static_for<R>([&](auto rc) AFGPU_INLINE {
            constexpr int r = decltype(rc)::value;
            wref(bc, rc) = h ? v[b + r * NB].y : v[b + r * NB].x;
          });
        });
        __syncthreads();
        static_for<E>([&](auto mc) AFGPU_INLINE {
          constexpr int m = decltype(mc)::value;
          if constexpr (h)
            v[m].y = rref(mc);
          else
            v[m].x = rref(mc);
        });
        __syncthreads();
      });
    }
    run_stages<T, L, SV, S + 1, HX, ST>(v, t, lat, tw);
  }
}

// Row length (in LDS entries of EB bytes) of a transposed-store tile: the
// store phase reads entry c2 * LL + k with c2 = lane % CT and k running over
// lane / CT.  The LDS serves mod = 128 / EB lanes per cycle (32 four-byte
// banks), so LL = mod / CT (mod mod) puts the CT rows of those lanes mod / CT
// banks apart and the mod / CT values of k in between: no bank is hit twice.
template <int L, int CT, int EB>
constexpr int transposed_row_len() {
  constexpr int base = LdsLen<L>::value;
  constexpr int mod = EB > 0 ? 128 / EB : 1;
  constexpr int want = CT <= mod ? (mod / CT) % mod : 1;
  return base + ((want - base % mod) % mod + mod) % mod;
}

// RT > 0: the tile is row-mapped with a transposed store and LDS entries of
// RT bytes (its rows use transposed_row_len).
template <int L, int CT, int SV = 0, int RT = 0>
struct PassShape {
  static constexpr int E = Sched<L, SV>::E;
  static constexpr int TPT = L / E;
  static constexpr int NT = CT * TPT;
  static constexpr int LL = RT ? transposed_row_len<L, CT, RT>() : LdsLen<L>::value;
  static constexpr size_t lds_elems = size_t(CT) * LL;
  // LDS bytes of a tile of element type T (HX: one real per entry)
  template <typename T, bool HX>
  static constexpr size_t lds_bytes() {
    return lds_elems * (HX ? sizeof(T) : 2 * sizeof(T));
  }
};

// One tile (CT columns) of a pass, executed by the NT = CT * L/E threads
// tid = 0..NT-1 of a work-group, with `lds` holding CT * LdsLen<L> entries.
// ---- element access ---------------------------------------------------------
// Generic access follows PassParams in full (split / interleaved, one-array
// split, per-column buffers).  FAST access is for interleaved data on both
// sides: one buffer resource per tile side (base and byte extent uniform over
// the work-group, so reads past the extent return 0 and writes past it are
// dropped) and 32-bit byte offsets per element, which saves the 64-bit address
// arithmetic of every load and store.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <typename T, int AUX>
__device__ __forceinline__ cx<T> buf_ld(__amdgpu_buffer_rsrc_t r, unsigned off);
template <>
__device__ __forceinline__ cx<float> buf_ld<float, 0>(__amdgpu_buffer_rsrc_t r, unsigned off) {
  u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
  return {__uint_as_float(w.x), __uint_as_float(w.y)};
}
template <>
__device__ __forceinline__ cx<float> buf_ld<float, 2>(__amdgpu_buffer_rsrc_t r, unsigned off) {
  u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 2);
  return {__uint_as_float(w.x), __uint_as_float(w.y)};
}
template <>
__device__ __forceinline__ cx<double> buf_ld<double, 0>(__amdgpu_buffer_rsrc_t r, unsigned off) {
  u32x4 w = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  return {__hiloint2double(w.y, w.x), __hiloint2double(w.w, w.z)};
}
template <>
__device__ __forceinline__ cx<double> buf_ld<double, 2>(__amdgpu_buffer_rsrc_t r, unsigned off) {
  u32x4 w = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 2);
  return {__hiloint2double(w.y, w.x), __hiloint2double(w.w, w.z)};
}
template <int AUX>
__device__ __forceinline__ void buf_st(cx<float> v, __amdgpu_buffer_rsrc_t r, unsigned off) {
  u32x2 w = {__float_as_uint(v.x), __float_as_uint(v.y)};
  __builtin_amdgcn_raw_buffer_store_b64(w, r, off, 0, AUX);
}
template <int AUX>
__device__ __forceinline__ void buf_st(cx<double> v, __amdgpu_buffer_rsrc_t r, unsigned off) {
  unsigned long long a = __double_as_longlong(v.x), b = __double_as_longlong(v.y);
  u32x4 w = {(unsigned)a, (unsigned)(a >> 32), (unsigned)b, (unsigned)(b >> 32)};
  __builtin_amdgcn_raw_buffer_store_b128(w, r, off, 0, AUX);
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}

// W_2L^k (k < L) of the real pass beside an L-point row DFT: computed in FP32
// (one load per bin less), read from the table in FP64.
template <typename T, int L>
__device__ __forceinline__ cx<T> real_twiddle(const cx<T>* __restrict__ tab, int k) {
  if constexpr (std::is_same_v<T, float> && L >= 4)
    return tw_pow2f((unsigned)k, __builtin_ctz(2 * L));
  else
    return tab[k];
}

// A cache policy known only at run time (one branch per access).
struct RtPol {
  int nt;
};

// One tile (CT columns) of a pass, executed by the NT = CT * L/E threads
// tid = 0..NT-1 of a work-group, with `lds` holding CT * LdsLen<L> entries.

template <typename T, int L, int CT, bool COLMAP, int EPI, bool STORE_T, bool FAST = false, int SV = 0, bool HX = false>
__device__ __forceinline__ void fft_tile(const PassParams<T>& p, int tile, int tid, cx<T>* __restrict__ lds) {
  static_assert(EPI != EPI_C2R4 || (COLMAP && !STORE_T), "the paired load is for four-step column passes");
  using PS = PassShape<L, CT, SV, (!COLMAP && STORE_T) ? (HX ? (int)sizeof(T) : 2 * (int)sizeof(T)) : 0>;
  constexpr int E = PS::E;
  constexpr int TPT = PS::TPT;
  constexpr int LL = PS::LL;
  constexpr unsigned ES = sizeof(cx<T>);

  const int c = COLMAP ? (tid % CT) : (tid / TPT);
  const int t = COLMAP ? (tid / CT) : (tid % TPT);
  const int bidx = tile / p.col_tiles;
  const int col0 = (tile - bidx * p.col_tiles) * CT;
  const int col = col0 + c;
  const bool active = col < p.n_col;
  const int bhi = bidx / p.n_blo;
  const int blo = bidx - bhi * p.n_blo;

  auto lat = [&](int i) -> cx<T>& {
    if constexpr (COLMAP)
      return lds[lds_pad(i) * CT + c];
    else
      return lds[c * LL + lds_pad(i)];
  };
  T* __restrict__ ldsf = reinterpret_cast<T*>(lds);
  auto latf = [&](int i) -> T& {
    if constexpr (COLMAP)
      return ldsf[lds_pad(i) * CT + c];
    else
      return ldsf[c * LL + lds_pad(i)];
  };

  // generic addressing
  const T* in_re = p.in_re;
  const T* in_im = p.in_im;
  T* out_re = p.out_re;
  T* out_im = p.out_im;
  idx_t ib = 0, ob = 0;
  // fast addressing: per-side resource + 32-bit byte offsets
  __amdgpu_buffer_rsrc_t rin = make_rsrc(p.in_re, 0u), rout = make_rsrc(p.out_re, 0u);
  unsigned in0 = 0, out0 = 0, js8 = 0, ks8 = 0;
  if constexpr (FAST) {
    rin = make_rsrc(reinterpret_cast<const cx<T>*>(p.in_re) + ((idx_t)bhi * p.in_bs + (idx_t)blo * p.in_bs2),
                    p.in_bytes);
    rout = make_rsrc(reinterpret_cast<const cx<T>*>(p.out_re) + ((idx_t)bhi * p.out_bs + (idx_t)blo * p.out_bs2),
                     p.out_bytes);
    in0 = (unsigned)col * (unsigned)p.in_cs * ES;
    out0 = (unsigned)col * (unsigned)p.out_cs * ES;
    js8 = (unsigned)p.in_js * ES;
    ks8 = (unsigned)p.out_ks * ES;
  } else {
    ib = (idx_t)bhi * p.in_bs + (idx_t)blo * p.in_bs2 + (idx_t)col * p.in_cs;
    ob = (idx_t)bhi * p.out_bs + (idx_t)blo * p.out_bs2 + (idx_t)col * p.out_cs;
    if (p.col_ptrs) {
      ib = 0;
      ob = 0;
      if (active) {
        in_re = reinterpret_cast<const T*>(p.col_ptrs[2 * col]);
        out_re = reinterpret_cast<T*>(p.col_ptrs[2 * col + 1]);
      }
    }
  }
  // element j of this column / bin k of this column.  POL is the cache policy
  // of the fast path's buffer accesses (2 = non-temporal), chosen once per
  // phase below so that no access carries a branch of its own.
  auto load = [&](auto polc, int j) AFGPU_INLINE -> cx<T> {
    if constexpr (FAST) {
      const unsigned off = in0 + (unsigned)j * js8;
      return buf_ld<T, decltype(polc)::value>(rin, off);
    } else {
      return ld_elem(in_re, in_im, p.in_il, ib + (idx_t)j * p.in_js, p.in_nt);
    }
  };
  // element at offset e (elements) from the batch entry's base (EPI_C2R4)
  auto load_at = [&](auto polc, idx_t e) AFGPU_INLINE -> cx<T> {
    if constexpr (FAST) {
      return buf_ld<T, decltype(polc)::value>(rin, (unsigned)e * ES);
    } else {
      return ld_elem(p.in_re, p.in_im, p.in_il, (idx_t)bhi * p.in_bs + (idx_t)blo * p.in_bs2 + e, p.in_nt);
    }
  };
  auto store = [&](auto polc, int k, cx<T> z) AFGPU_INLINE {
    if (p.swap_out) z = swapri(z);
    if constexpr (FAST) {
      const unsigned off = out0 + (unsigned)k * ks8;
      if constexpr (std::is_same_v<decltype(polc), RtPol>) {
        if (polc.nt)
          buf_st<2>(z, rout, off);
        else
          buf_st<0>(z, rout, off);
      } else {
        buf_st<decltype(polc)::value>(z, rout, off);
      }
    } else {
      st_elem(out_re, out_im, p.out_il, ob + (idx_t)k * p.out_ks, z, p.out_nt);
    }
  };

  // runs f(pol) with pol = the compile-time cache policy for a side whose
  // streaming flag is nt
  auto with_pol = [&](bool nt, auto&& f) AFGPU_INLINE {
    if constexpr (FAST) {
      if (nt)
        f(std::integral_constant<int, 2>{});
      else
        f(std::integral_constant<int, 0>{});
    } else {
      f(std::integral_constant<int, 0>{});
    }
  };

  cx<T> v[E];
  if (active) {
    with_pol(p.in_nt, [&](auto pol) AFGPU_INLINE {
      if constexpr (EPI == EPI_C2R4) {
        // Z[k] = (X[k] + conj X[h-k]) - i conj(W_2h^k) (conj X[h-k] - X[k]),
        // Z[0] = (Re X[0] + Re X[h]) + i (Re X[0] - Re X[h]); eight elements at
        // a time (two loads and a twiddle each)
        const idx_t h = p.real_h;
        const idx_t kc = (idx_t)col * p.in_cs;
        // W_2h^k = W_2h^(col*in_cs) * W_2h^(j*in_js): one factor per thread and
        // one per row of the tile, both from small, cached parts of the table
        // (reading W_2h^k itself would stream a table as large as the data)
        const cx<T> wc = p.real_tw[kc];
        constexpr int G = E < 8 ? E : 8;
        static_for<E / G>([&](auto gc) AFGPU_INLINE {
          static_for<G>([&](auto ic) AFGPU_INLINE {
            constexpr int m = G * decltype(gc)::value + decltype(ic)::value;
            const int j = t + m * TPT;
            const idx_t k = kc + (idx_t)j * p.in_js;
            const cx<T> s0 = load(pol, j);
            const cx<T> s1 = load_at(pol, h - k);
            cx<T> z;
            if (k == 0) {
              z = {s0.x + s1.x, s0.x - s1.x};
            } else {
              const cx<T> s1c = conjg(s1);
              const cx<T> bb = cmulc(s1c - s0, cmul(wc, p.real_tw[(idx_t)j * p.in_js]));
              z = s0 + s1c + mul_neg_i(bb);
            }
            v[m] = p.swap_in ? swapri(z) : z;
          });
          __builtin_amdgcn_sched_barrier(0);
        });
      } else if constexpr (EPI == EPI_C2R) {
        static_for<E>([&](auto mc) AFGPU_INLINE {
          constexpr int m = decltype(mc)::value;
          const int k = t + m * TPT;
          cx<T> s0 = load(pol, k);
          cx<T> s1 = load(pol, L - k);
          cx<T> z;
          if (k == 0) {
            z = {s0.x + s1.x, s0.x - s1.x};
          } else {
            cx<T> w = real_twiddle<T, L>(p.real_tw, k);
            cx<T> s1c = conjg(s1);
            cx<T> a = s0 + s1c;
            cx<T> bb = cmulc(s1c - s0, w);
            z = a + mul_neg_i(bb);
          }
          v[m] = p.swap_in ? swapri(z) : z;
        });
      } else {
        static_for<E>([&](auto mc) AFGPU_INLINE {
          constexpr int m = decltype(mc)::value;
          cx<T> z = load(pol, t + m * TPT);
          v[m] = p.swap_in ? swapri(z) : z;
        });
      }
    });
  } else {
    static_for<E>([&](auto mc) AFGPU_INLINE { v[decltype(mc)::value] = cx<T>{T(0), T(0)}; });
  }

  if constexpr (HX)
    run_stages<T, L, SV, 0, true, COLMAP ? CT : 1>(v, t, latf, p.stage_tw);
  else
    run_stages<T, L, SV, 0, false, COLMAP ? CT : 1>(v, t, lat, p.stage_tw);

  // The output policy is hoisted over the whole epilogue for column tiles
  // (fewer live registers in their twiddled store); row tiles keep one branch
  // per store, which measured fewer spills in the transposed store.
  auto with_out_pol = [&](auto&& f) AFGPU_INLINE {
    if constexpr (COLMAP)
      with_pol(p.out_nt, f);
    else
      f(RtPol{p.out_nt});
  };
  with_out_pol([&](auto pol) AFGPU_INLINE {
    if constexpr (EPI == EPI_R2C) {
      // X[k] from Z[k] and its partner Z[L-k]; the first bin also gives X[L]
      auto finish = [&](auto mc, cx<T> s1) AFGPU_INLINE {
        constexpr int m = decltype(mc)::value;
        const int k = t + m * TPT;
        const cx<T> s0 = v[m];
        cx<T> x;
        if (k == 0) {
          x = {s0.x + s0.y, T(0)};
          store(pol, L, cx<T>{s0.x - s0.y, T(0)});
        } else {
          const cx<T> s1c = conjg(s1);
          const cx<T> w = real_twiddle<T, L>(p.real_tw, k);
          const T h = T(0.5);
          const cx<T> a = {(s0.x + s1c.x) * h, (s0.y + s1c.y) * h};
          const cx<T> bw = cmul(s0 - s1c, w);
          x = a + mul_neg_i(cx<T>{bw.x * h, bw.y * h});
        }
        store(pol, k, x);
      };
      if constexpr (HX) {
        // half-LDS tiles (one row): two rounds.  The partners of the lower
        // bins k < L/2 (registers m < E/2) are upper bins and the other way
        // round, so each round stages one half of the row as whole complex
        // entries -- L/2 of them fill the half-LDS tile -- and no partner is
        // held in registers
        static_assert(CT == 1 && !COLMAP, "the two-round r2c epilogue is for single-row half-LDS tiles");
        static_assert(E % 2 == 0, "the two rounds split the registers in halves");
        cx<T>* __restrict__ half_row = reinterpret_cast<cx<T>*>(ldsf);
        constexpr int MH = E / 2;
        static_for<E - MH>([&](auto ic) AFGPU_INLINE {
          constexpr int m = MH + decltype(ic)::value;
          half_row[lds_pad(t + m * TPT - L / 2)] = v[m];
        });
        __syncthreads();
        if (active) {
          static_for<MH>([&](auto mc) AFGPU_INLINE {
            const int k = t + decltype(mc)::value * TPT;
            finish(mc, k == 0 ? v[0] : half_row[lds_pad(L - k - L / 2)]);
          });
        }
        __syncthreads();
        static_for<MH>([&](auto mc) AFGPU_INLINE {
          constexpr int m = decltype(mc)::value;
          half_row[lds_pad(t + m * TPT)] = v[m];
        });
        __syncthreads();
        if (active) {
          static_for<E - MH>([&](auto ic) AFGPU_INLINE {
            constexpr int m = MH + decltype(ic)::value;
            const int k = t + m * TPT;
            // bin L/2 is its own partner
            finish(std::integral_constant<int, m>{}, k == L / 2 ? v[m] : half_row[lds_pad(L - k)]);
          });
        }
      } else {
        static_for<E>([&](auto mc) AFGPU_INLINE {
          constexpr int m = decltype(mc)::value;
          lat(t + m * TPT) = v[m];
        });
        __syncthreads();
        if (active) {
          static_for<E>([&](auto mc) AFGPU_INLINE {
            const int k = t + decltype(mc)::value * TPT;
            finish(mc, k == 0 ? v[0] : lat(L - k));
          });
        }
      }
    } else {
      if constexpr ((EPI == EPI_TW4 || EPI == EPI_C2R4) && !STORE_T) {
        // twiddle and store eight bins at a time: the scheduling barrier keeps
        // the table loads of later groups from being hoisted, which would hold
        // every twiddle in registers at once.  Group g takes registers
        // m = g (mod E/8), which the last radix-32 stage's inner radix-8
        // butterflies produce together, so their registers free early.
        const unsigned n2 = p.tw4_blo ? (unsigned)blo : (unsigned)(col >> p.tw4_col_shift);
        const unsigned mask = (1u << p.tw4_lo_bits) - 1u;
        constexpr int GS = E < 8 ? E : 8;
        static_for<E / GS>([&](auto gc) AFGPU_INLINE {
          constexpr int g = decltype(gc)::value;
          static_for<GS>([&](auto ic) AFGPU_INLINE {
            constexpr int m = g + (E / GS) * decltype(ic)::value;
            const unsigned mi = n2 * (unsigned)(t + m * TPT);  // < N <= 2^28
            cx<T> w;
            if constexpr (std::is_same_v<T, float>) {
              if (p.tw4_log2n > 0)
                w = tw_pow2f(mi, p.tw4_log2n);
              else
                w = tw4_lookup32(p.tw4_hi, p.tw4_lo, p.tw4_lo_bits, mask, mi);
            } else {
              w = tw4_lookup32(p.tw4_hi, p.tw4_lo, p.tw4_lo_bits, mask, mi);
            }
            v[m] = cmul(v[m], w);
            if (active) store(pol, t + m * TPT, v[m]);
          });
          __builtin_amdgcn_sched_barrier(0);
        });
        return;
      }
      if constexpr (EPI == EPI_TW4) {
        const idx_t n2 = p.tw4_blo ? (idx_t)blo : (idx_t)(col >> p.tw4_col_shift);
        static_for<E>([&](auto mc) AFGPU_INLINE {
          constexpr int m = decltype(mc)::value;
          const idx_t k1 = t + m * TPT;
          v[m] = cmul(v[m], tw4_lookup(p.tw4_hi, p.tw4_lo, p.tw4_lo_bits, n2 * k1));
        });
      }
      if constexpr (STORE_T) {
        static_assert(!COLMAP, "transposed store is for row-mapped passes");
        const int c2 = tid % CT;
        const int t2 = tid / CT;
        const int colb = col0 + c2;
        // re-aim the output at column colb (the thread's column changes here)
        if constexpr (FAST) {
          out0 = (unsigned)colb * (unsigned)p.out_cs * ES;
        } else {
          ob = (idx_t)bhi * p.out_bs + (idx_t)blo * p.out_bs2 + (idx_t)colb * p.out_cs;
          out_re = p.out_re;
          out_im = p.out_im;
        }
        if constexpr (HX) {
          T xs[E];
          static_for<E>([&](auto mc) AFGPU_INLINE { latf(t + decltype(mc)::value * TPT) = v[decltype(mc)::value].x; });
          __syncthreads();
          static_for<E>([&](auto mc) AFGPU_INLINE {
            constexpr int m = decltype(mc)::value;
            xs[m] = ldsf[c2 * LL + lds_pad(t2 + m * TPT)];
          });
          __syncthreads();
          static_for<E>([&](auto mc) AFGPU_INLINE { latf(t + decltype(mc)::value * TPT) = v[decltype(mc)::value].y; });
          __syncthreads();
          if (colb < p.n_col) {
            static_for<E>([&](auto mc) AFGPU_INLINE {
              constexpr int m = decltype(mc)::value;
              const int k = t2 + m * TPT;
              store(pol, k, cx<T>{xs[m], ldsf[c2 * LL + lds_pad(k)]});
            });
          }
        } else {
          static_for<E>([&](auto mc) AFGPU_INLINE {
            constexpr int m = decltype(mc)::value;
            lat(t + m * TPT) = v[m];
          });
          __syncthreads();
          if (colb < p.n_col) {
            static_for<E>([&](auto mc) AFGPU_INLINE {
              constexpr int m = decltype(mc)::value;
              const int k = t2 + m * TPT;
              store(pol, k, lds[c2 * LL + lds_pad(k)]);
            });
          }
        }
      } else if (active) {
        static_for<E>([&](auto mc) AFGPU_INLINE {
          constexpr int m = decltype(mc)::value;
          store(pol, t + m * TPT, v[m]);
        });
      }
    }
  });
}

// LDS bytes of a pass tile (the shape fft_tile uses).
template <typename T, int L, int CT, bool COLMAP, bool STORE_T, int SV, bool HX>
constexpr size_t tile_lds_bytes() {
  return PassShape<L, CT, SV, (!COLMAP && STORE_T) ? (HX ? (int)sizeof(T) : 2 * (int)sizeof(T)) : 0>::template lds_bytes<T, HX>();
}

// Minimum waves per SIMD the compiler must allow for: half-exchange tiles
// that fit twice in a CU's 160 KiB of LDS are meant to run two groups per CU,
// so their registers are bounded accordingly; everything else runs one.

template <typename T, int L, int CT, bool COLMAP, bool STORE_T, int SV, bool HX>
constexpr int pass_min_waves() {
  constexpr int nt = PassShape<L, CT, SV>::NT;
  constexpr int groups = (HX && 2 * tile_lds_bytes<T, L, CT, COLMAP, STORE_T, SV, HX>() <= 160 * 1024) ? 2 : 1;
  return groups * nt / 256 > 1 ? groups * nt / 256 : 1;
}

template <typename T, int L, int CT, bool COLMAP, int EPI, bool STORE_T, bool FAST, int SV, bool HX>
__global__ void __launch_bounds__((PassShape<L, CT, SV>::NT), (pass_min_waves<T, L, CT, COLMAP, STORE_T, SV, HX>()))
    fft_pass_kernel(PassParams<T> p) {
  extern __shared__ __align__(16) unsigned char afgpu_smem[];
  int tile = p.xcd_remap ? xcd_tile(blockIdx.x, p.n_tiles) : (int)blockIdx.x;
  if constexpr (EPI == EPI_C2R4) {
    // the partner bins of column tile c lie in its mirror tile: run both on
    // one XCD (blocks go round-robin over the 8 XCDs) so that the second read
    // of every bin hits that XCD's L2
    if (p.mirror_pairs) {
      const int B = blockIdx.x;
      const int pair = (B >> 4) * 8 + (B & 7);
      const int half = p.col_tiles / 2;
      const int b = pair / half;
      if (b >= p.n_batch) return;
      const int pp = pair - b * half;
      tile = b * p.col_tiles + ((B & 8) ? p.col_tiles - 1 - pp : pp);
    }
  }
  fft_tile<T, L, CT, COLMAP, EPI, STORE_T, FAST, SV, HX>(p, tile, threadIdx.x,
                                                         reinterpret_cast<cx<T>*>(afgpu_smem));
}

}  // namespace afgpu
