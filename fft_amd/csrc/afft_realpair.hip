// afft_realpair.hip -- real passes fused into FFT passes: pass B of the
// four-step 1-D r2c with the real post-pass and the output transposition; the
// N-D r2c / c2r real pass over row pairs with the inner-axis row FFTs; and the
// same pass inside the dim-0 column four-step (r2c: pass B, c2r: pass A).
//
// A long r2c (N reals, h = N/2 = n1 * n2 complex points) runs its h-point c2c
// as a four-step: pass A (n1-point columns with the W_h twiddle) leaves
// Y[k1][n2'] in scratch, and the n2-point FFT of row k1 gives
// Z[k1 + n1*k2] at bin k2.  The real post-pass (onedim.hpp:1066-1106,
// inverse=false) forms
//     X[k] = (Z[k] + conj Z[h-k]) / 2 - i W_2h^k (Z[k] - conj Z[h-k]) / 2,
// and h - k = (n1 - k1) + n1*(n2 - 1 - k2) for k1 > 0, n1*((n2 - k2) mod n2)
// for k1 = 0.  So the partner of every bin of row k1 is a bin of row n1 - k1:
// a work-group that transforms rows k1 and n1 - k1 together has both members
// of every pair.  It holds 2R rows:
//   slots c <  R: rows j*R + c,
//   slots c >= R: rows n1 - j*R - (c - R)   (tile 0: slot R is row n1/2),
// so slot c pairs with slot c +- R, and rows 0 and n1/2 (tile 0, slots 0 and
// R) pair with themselves.  After the row FFTs the partner values cross the
// LDS (real parts, then imaginary parts), each thread forms X for its bins,
// and a second exchange turns the rows around so that 2R lanes store R
// consecutive bins of each row set (bins k1 + n1*k2 are contiguous in k1).
// Bin 0 of row 0 gives X[0] = Re Z[0] + Im Z[0] through the same formula
// (W^0 = 1, partner itself); its thread also writes X[h] = Re Z[0] - Im Z[0].
//
// This replaces a row pass into a second scratch plus real_tr_kernel: the
// intermediate Z never goes back to memory, and the plan needs half the
// scratch.  It is a variant the caller asks for (afft_gpu_variant.real_pair):
// a work-group of 16 rows holds 64 registers of data per lane plus the
// partner values, is bounded to 128 VGPRs (two groups per CU for 1024-point
// rows, one for 2048) and spills; measured, it runs 5-10 % behind the two
// kernels it replaces for 1024-point rows and 2.5x behind for 2048.
#include <climits>

#include "afft_aux.h"
#include "afft_pass.h"

namespace afgpu {

// Row held by slot c of tile j (see above).
__device__ __forceinline__ int rp_row(int j, int c, int R, int n1) {
  if (c < R) return j * R + c;
  const int c2 = c - R;
  if (j == 0 && c2 == 0) return n1 / 2;
  return n1 - j * R - c2;
}

template <typename T, int L, int R, int SV>
constexpr size_t rp_lds_bytes() {
  // one real per entry (half exchange), rows padded for the transposed reads
  return PassShape<L, 2 * R, SV, (int)sizeof(T)>::template lds_bytes<T, true>();
}

template <typename T, int L, int R, int SV>
constexpr int rp_min_waves() {
  constexpr int nt = PassShape<L, 2 * R, SV>::NT;
  constexpr int groups = 2 * rp_lds_bytes<T, L, R, SV>() <= 160 * 1024 ? 2 : 1;
  return groups * nt / 256 > 1 ? groups * nt / 256 : 1;
}

template <typename T, int L, int R, int SV>
__global__ void __launch_bounds__((PassShape<L, 2 * R, SV>::NT), (rp_min_waves<T, L, R, SV>()))
    real_pair_kernel(RealPairArgs<T> a) {
  static_assert(sizeof(T) == 4, "FP32 rows (the split store writes 4-byte reals)");
  using PS = PassShape<L, 2 * R, SV, (int)sizeof(T)>;
  constexpr int E = PS::E;
  constexpr int TPT = PS::TPT;
  constexpr int LL = PS::LL;
  constexpr int CT = 2 * R;
  extern __shared__ __align__(16) unsigned char afgpu_smem[];
  T* __restrict__ ldsf = reinterpret_cast<T*>(afgpu_smem);

  const int tiles = a.n1 / CT;
  const idx_t b = blockIdx.x / tiles;
  const int j = (int)(blockIdx.x - b * tiles);
  const int tid = threadIdx.x;
  const int c = tid / TPT, t = tid % TPT;
  const int row = rp_row(j, c, R, a.n1);
  const bool self = j == 0 && (c == 0 || c == R);
  const bool row0 = j == 0 && c == 0;
  const int pc = self ? c : (c < R ? c + R : c - R);

  auto latf = [&](int i) -> T& { return ldsf[c * LL + lds_pad(i)]; };
  // entries t + m*TPT of this row: one base, constant steps (TPT is a multiple
  // of 16, so lds_pad(t + m*TPT) = lds_pad(t) + m*MS); written this way the
  // LDS accesses take immediate offsets instead of one address register each
  constexpr int MS = TPT + TPT / 16;
  T* __restrict__ own = ldsf + c * LL + lds_pad(t);
  // LDS entry of the partner of bin k = t + m*TPT: bin L-1-k of the partner
  // row, or (L-k) mod L for row 0.  Both are base - m*TPT with one base per
  // thread (bin 0 of row 0, its own partner, aside); TPT is a multiple of 16,
  // so the padded index steps by a constant and the LDS reads take immediate
  // offsets.
  static_assert(TPT % 16 == 0, "padded partner index must be affine in m");
  const int pbase = row0 ? L - t : L - 1 - t;
  const int pent = pc * LL + lds_pad(pbase);
  const int pent0 = (row0 && t == 0) ? pc * LL : pent;
  auto partner = [&](auto mc) -> T {
    constexpr int m = decltype(mc)::value;
    if constexpr (m == 0)
      return ldsf[pent0];
    else
      return ldsf[pent - m * MS];
  };

  // rows in (buffer addressing: one descriptor per transform, 32-bit byte
  // offsets), FFT along each row
  constexpr unsigned ES = sizeof(cx<T>);
  const unsigned h = (unsigned)a.n1 * (unsigned)L;
  const __amdgpu_buffer_rsrc_t rin = make_rsrc(a.in + b * a.in_bs, h * ES);
  const unsigned rowoff = (unsigned)row * (unsigned)L * ES + (unsigned)t * ES;
  cx<T> v[E];
  static_for<E>([&](auto mc) AFGPU_INLINE {
    constexpr int m = decltype(mc)::value;
    v[m] = buf_ld<T, 0>(rin, rowoff + (unsigned)(m * TPT) * ES);
  });
  run_stages<T, L, SV, 0, true, 1>(v, t, latf, a.stage_tw);

  // real post-pass: partner real parts, then imaginary parts, through the LDS
  const cx<T> wr = a.real_tw[row];
  T pre[E];
  static_for<E>([&](auto mc) AFGPU_INLINE { own[decltype(mc)::value * MS] = v[decltype(mc)::value].x; });
  __syncthreads();
  static_for<E>([&](auto mc) AFGPU_INLINE { pre[decltype(mc)::value] = partner(mc); });
  __syncthreads();
  static_for<E>([&](auto mc) AFGPU_INLINE { own[decltype(mc)::value * MS] = v[decltype(mc)::value].y; });
  __syncthreads();
  const T nyq = v[0].x - v[0].y;  // used by row 0, bin 0 only
  // eight bins at a time: the scheduling barrier keeps the twiddle loads of
  // later groups from being hoisted (all of them would not fit in registers)
  static_for<E / 8>([&](auto gc) AFGPU_INLINE {
    static_for<8>([&](auto ic) AFGPU_INLINE {
      constexpr int m = 8 * decltype(gc)::value + decltype(ic)::value;
      const int k = t + m * TPT;
      const cx<T> z = v[m];
      const cx<T> zc{pre[m], -partner(std::integral_constant<int, m>{})};  // conj Z[h - k]
      const cx<T> w = cmul(wr, a.row_tw[k]);
      const T hf = T(0.5);
      const cx<T> s = {(z.x + zc.x) * hf, (z.y + zc.y) * hf};
      const cx<T> dw = cmul(z - zc, w);
      v[m] = s + mul_neg_i(cx<T>{dw.x * hf, dw.y * hf});
    });
    __builtin_amdgcn_sched_barrier(0);
  });
  // pin the results here: without it the compiler sinks the arithmetic above
  // past the barrier, where its inputs then compete with the store phase for
  // registers
#pragma unroll
  for (int m = 0; m < E; m++) asm volatile("" : "+v"(v[m].x), "+v"(v[m].y));
  __syncthreads();

  // transpose through the LDS: lane c2 of each group of 2R lanes stores row
  // slot c2, so R lanes write R consecutive bins of one row set
  const int c2 = tid % CT, t2 = tid / CT;
  const int orow = rp_row(j, c2, R, a.n1);
  T xs[E];
  static_for<E>([&](auto mc) AFGPU_INLINE { own[decltype(mc)::value * MS] = v[decltype(mc)::value].x; });
  __syncthreads();
  const T* __restrict__ col = ldsf + c2 * LL + lds_pad(t2);
  static_for<E>([&](auto mc) AFGPU_INLINE { xs[decltype(mc)::value] = col[decltype(mc)::value * MS]; });
  __syncthreads();
  static_for<E>([&](auto mc) AFGPU_INLINE { own[decltype(mc)::value * MS] = v[decltype(mc)::value].y; });
  __syncthreads();
  // bins 0..h of this transform: interleaved (one 8-byte store per bin) or
  // split (a 4-byte store into each of the re and im arrays)
  const idx_t ob = b * a.out_bs;
  const bool il = a.out_il != 0;
  const unsigned obytes = (h + 1) * (il ? ES : (unsigned)sizeof(T));
  const __amdgpu_buffer_rsrc_t rre =
      make_rsrc(il ? (void*)(reinterpret_cast<cx<T>*>(a.out_re) + ob) : (void*)(a.out_re + ob), obytes);
  const __amdgpu_buffer_rsrc_t rim = make_rsrc(il ? (void*)a.out_re : (void*)(a.out_im + ob), il ? 0u : obytes);
  // bin orow + n1*(t2 + m*TPT): the per-lane part in the vector offset, the
  // per-m part (uniform) in the scalar offset
  const unsigned esz = il ? ES : (unsigned)sizeof(T);
  const unsigned vbin = (unsigned)orow + (unsigned)a.n1 * (unsigned)t2;
  const unsigned mstep = (unsigned)a.n1 * (unsigned)TPT * esz;
  auto put = [&](auto polc, unsigned voff, unsigned soff, cx<T> x) AFGPU_INLINE {
    constexpr int POL = decltype(polc)::value;
    if (il) {
      const u32x2 w = {__float_as_uint(x.x), __float_as_uint(x.y)};
      __builtin_amdgcn_raw_buffer_store_b64(w, rre, voff, soff, POL);
    } else {
      __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(x.x), rre, voff, soff, POL);
      __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(x.y), rim, voff, soff, POL);
    }
  };
  auto store_all = [&](auto polc) AFGPU_INLINE {
    static_for<E>([&](auto mc) AFGPU_INLINE {
      constexpr int m = decltype(mc)::value;
      const cx<T> x{xs[m], col[m * MS]};
      put(polc, vbin * esz, (unsigned)m * mstep, x);
    });
    if (row0 && t == 0) put(polc, h * esz, 0u, cx<T>{nyq, T(0)});
  };
  if (a.out_nt)
    store_all(std::integral_constant<int, 2>{});
  else
    store_all(std::integral_constant<int, 0>{});
}

// ---- N-D r2c: real pass over row pairs fused into the first inner-axis pass --
//
// After the dim-0 c2c of an N-D r2c, dst rows 0..h-1 hold Z[k][j] (j = index
// along the untransformed last axis).  The real pass (multi.hpp:343-414,
// inverse=false) is elementwise per column j between rows k and h-k:
//     X[k]   = (Z[k] + conj Z[h-k]) / 2 - i W_2h^k   (Z[k] - conj Z[h-k]) / 2,
//     X[h-k] = the same with k and h-k swapped (W_2h^(h-k) = -conj W_2h^k),
// X[0] = Re Z[0] + Im Z[0], X[h] = Re Z[0] - Im Z[0], X[h/2] = conj Z[h/2].
// Work-group p (0 <= p <= h/2) reads rows p and h-p, forms X for both (slot 0:
// row p, slot 1: row h-p; p = 0: rows 0 and h, both from Z[0]; p = h/2: slot 0
// only), runs the L-point FFT along each row and writes the rows in place:
// the real pass costs no pass of its own.  With the own and the other row's
// value at the same j in hand, the uniform formula
//     X = (own + conj oth) / 2 - i w (own - conj oth) / 2,
//     w = W_2h^p (slot 0), -conj W_2h^p (slot 1)
// covers every case (p = 0 gives Re +- Im, p = h/2 gives conj Z).
// FP32: two work-groups per CU where their LDS fits (registers bounded to
// match); FP64 rows need the registers of one.
template <typename T, int L, int SV>
constexpr int rrf_min_waves() {
  constexpr int nt = PassShape<L, 2, SV>::NT;
  constexpr int groups =
      sizeof(T) == 4 && 2 * PassShape<L, 2, SV>::template lds_bytes<T, true>() <= 160 * 1024 ? 2 : 1;
  return groups * nt / 256 > 1 ? groups * nt / 256 : 1;
}

template <typename T, int L, int SV>
__global__ void __launch_bounds__((PassShape<L, 2, SV>::NT), (rrf_min_waves<T, L, SV>()))
    real_rowfft_kernel(RealRowFftArgs<T> a) {
  using PS = PassShape<L, 2, SV>;
  constexpr int E = PS::E;
  constexpr int TPT = PS::TPT;
  constexpr int LL = PS::LL;
  extern __shared__ __align__(16) unsigned char afgpu_smem[];
  T* __restrict__ ldsf = reinterpret_cast<T*>(afgpu_smem);
  const idx_t np = a.h / 2 + 1;
  const idx_t b = blockIdx.x / np;
  const idx_t p = blockIdx.x - b * np;
  const int tid = threadIdx.x;
  const int c = tid / TPT, t = tid % TPT;
  const idx_t h = a.h;
  const idx_t own_row = c == 0 ? p : (p == 0 ? 0 : h - p);
  const idx_t oth_row = c == 0 ? (p == 0 ? 0 : h - p) : p;
  const idx_t out_row = c == 0 ? p : (p == 0 ? h : h - p);
  const bool store = !(c == 1 && 2 * p == h);
  const cx<T> wp = a.real_tw[p];
  const cx<T> w = c == 0 ? wp : cx<T>{-wp.x, wp.y};
  const idx_t base = b * a.bs;
  auto ld = [&](idx_t row, int j) -> cx<T> {
    const idx_t e = base + row * a.rs + j;
    if (a.il) return reinterpret_cast<const cx<T>*>(a.re)[e];
    return cx<T>{a.re[e], a.im[e]};
  };
  // eight elements (sixteen loads) at a time: hoisting all of them would
  // double the registers the row needs
  cx<T> v[E];
  constexpr int G = E < 8 ? E : 8;
  static_for<E / G>([&](auto gc) AFGPU_INLINE {
    static_for<G>([&](auto ic) AFGPU_INLINE {
      constexpr int m = G * decltype(gc)::value + decltype(ic)::value;
      const int j = t + m * TPT;
      const cx<T> own = ld(own_row, j), oth = ld(oth_row, j);
      const T hf = T(0.5);
      const cx<T> oc = conjg(oth);
      const cx<T> s = {(own.x + oc.x) * hf, (own.y + oc.y) * hf};
      const cx<T> dw = cmul(own - oc, w);
      v[m] = s + mul_neg_i(cx<T>{dw.x * hf, dw.y * hf});
    });
    __builtin_amdgcn_sched_barrier(0);
  });
  auto latf = [&](int i) -> T& { return ldsf[c * LL + lds_pad(i)]; };
  run_stages<T, L, SV, 0, true, 1>(v, t, latf, a.stage_tw);
  if (store) {
    static_for<E>([&](auto mc) AFGPU_INLINE {
      constexpr int m = decltype(mc)::value;
      st_elem(a.re, a.im, a.il, base + out_row * a.rs + t + m * TPT, v[m], 0);
    });
  }
}

// ---- N-D c2r: inverse real pass over row pairs fused into the row FFTs -------
//
// A c2r first runs the inverse FFT along the last axis of the h+1 input rows
// (Y = IDFT X), then the inverse real pass (multi.hpp:343-414, inverse=true)
// pairs rows k and h-k elementwise:
//     Z[k]   = (Y[k] + conj Y[h-k]) - i conj(W_2h^k) (conj Y[h-k] - Y[k]),
//     Z[h-k] = its conjugate-symmetric partner,  Z[0] = (Re Y[0] + Re Y[h])
//              + i (Re Y[0] - Re Y[h]).
// Both steps are linear, and conj Y[r] is the IDFT of conj X[r][(-m) mod L],
// so the pairing can run before the row transform on the input rows with the
// partner row read at the mirrored column:
//     pre[m] = (own[m] + conj oth[-m]) + i u (own[m] - conj oth[-m]),
//     u = conj W_2h^p (slot 0, row p), -W_2h^p (slot 1, row h-p),
// and for row 0 the Hermitian parts H_r[m] = (X[r][m] + conj X[r][-m]) / 2 of
// rows 0 and h: pre = (H_0 + H_h) + i (H_0 - H_h) (this is what taking the
// real parts after the row transform amounts to).  Work-group p reads input
// rows p and h-p (p = 0: rows 0 and h), writes rows p and h-p of Z (p = 0,
// p = h/2: one row) into the plan's scratch; the real pass costs no pass.
template <typename T, int L, int SV>
__global__ void __launch_bounds__((PassShape<L, 2, SV>::NT), (rrf_min_waves<T, L, SV>()))
    real_irowfft_kernel(RealIRowFftArgs<T> a) {
  using PS = PassShape<L, 2, SV>;
  constexpr int E = PS::E;
  constexpr int TPT = PS::TPT;
  constexpr int LL = PS::LL;
  extern __shared__ __align__(16) unsigned char afgpu_smem[];
  T* __restrict__ ldsf = reinterpret_cast<T*>(afgpu_smem);
  const idx_t np = a.h / 2 + 1;
  const idx_t b = blockIdx.x / np;
  const idx_t p = blockIdx.x - b * np;
  const int tid = threadIdx.x;
  const int c = tid / TPT, t = tid % TPT;
  const idx_t h = a.h;
  const bool row0 = p == 0 && c == 0;
  const bool act = c == 0 || (p != 0 && 2 * p != h);
  const idx_t own_row = c == 0 ? p : h - p;
  const idx_t oth_row = c == 0 ? h - p : p;
  const cx<T> wp = a.real_tw[p];
  const cx<T> u = c == 0 ? conjg(wp) : cx<T>{-wp.x, -wp.y};
  const idx_t ib = b * a.in_bs;
  auto ld = [&](idx_t row, int j) -> cx<T> {
    const idx_t e = ib + row * a.in_rs + j;
    if (a.in_il) return reinterpret_cast<const cx<T>*>(a.in_re)[e];
    return cx<T>{a.in_re[e], a.in_im[e]};
  };
  cx<T> v[E];
  constexpr int G = E < 8 ? E : 8;
  static_for<E / G>([&](auto gc) AFGPU_INLINE {
    static_for<G>([&](auto ic) AFGPU_INLINE {
      constexpr int m = G * decltype(gc)::value + decltype(ic)::value;
      const int j = t + m * TPT;
      const int jm = (L - j) & (L - 1);
      cx<T> z;
      if (row0) {
        const T hf = T(0.5);
        const cx<T> x0 = ld(0, j), x0m = conjg(ld(0, jm)), xh = ld(h, j), xhm = conjg(ld(h, jm));
        const cx<T> h0 = {(x0.x + x0m.x) * hf, (x0.y + x0m.y) * hf};
        const cx<T> hh = {(xh.x + xhm.x) * hf, (xh.y + xhm.y) * hf};
        const cx<T> d = h0 - hh;
        z = h0 + hh + cx<T>{-d.y, d.x};
      } else {
        const cx<T> own = ld(own_row, j), oc = conjg(ld(oth_row, jm));
        const cx<T> du = cmul(own - oc, u);
        z = own + oc + cx<T>{-du.y, du.x};
      }
      v[m] = swapri(z);  // inverse transform: swap, forward FFT, swap
    });
    __builtin_amdgcn_sched_barrier(0);
  });
  auto latf = [&](int i) -> T& { return ldsf[c * LL + lds_pad(i)]; };
  run_stages<T, L, SV, 0, true, 1>(v, t, latf, a.stage_tw);
  if (act) {
    cx<T>* __restrict__ out = a.out + b * a.out_bs + own_row * a.out_rs;
    static_for<E>([&](auto mc) AFGPU_INLINE {
      constexpr int m = decltype(mc)::value;
      out[t + m * TPT] = swapri(v[m]);
    });
  }
}

namespace {
constexpr int kRpR = 8;  // row pairs per work-group: 8 consecutive bins per store run

template <typename T, int L>
hipError_t launch_rp(const RealPairArgs<T>& a, idx_t blocks, hipStream_t st) {
  constexpr int SV = 1;
  constexpr size_t lds = rp_lds_bytes<T, L, kRpR, SV>();
  auto k = real_pair_kernel<T, L, kRpR, SV>;
  if (lds > 65536) {
    static bool attr = false;
    if (!attr) {
      hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
      attr = true;
    }
  }
  hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(PassShape<L, 2 * kRpR, SV>::NT), lds, st, a);
  return hipGetLastError();
}
}  // namespace

bool real_pair_supported(int elem_bytes, int n1, int n2) {
  return elem_bytes == 4 && (n2 == 1024 || n2 == 2048) && n1 >= 2 * kRpR && n1 % (2 * kRpR) == 0;
}

template <typename T>
hipError_t real_pair(const RealPairArgs<T>& a, hipStream_t st) {
  if (!real_pair_supported((int)sizeof(T), a.n1, a.n2)) return hipErrorInvalidValue;
  const idx_t blocks = a.nbatch * (idx_t)(a.n1 / (2 * kRpR));
  if (blocks <= 0) return hipSuccess;
  if (blocks > INT_MAX) return hipErrorInvalidValue;
  if constexpr (sizeof(T) == 4) {
    if (a.n2 == 1024) return launch_rp<T, 1024>(a, blocks, st);
    return launch_rp<T, 2048>(a, blocks, st);
  }
  return hipErrorInvalidValue;
}

template <typename T, int L>
static hipError_t launch_rrf(const RealRowFftArgs<T>& a, idx_t blocks, hipStream_t st) {
  constexpr int SV = 0;
  constexpr size_t lds = PassShape<L, 2, SV>::template lds_bytes<T, true>();
  auto k = real_rowfft_kernel<T, L, SV>;
  if (lds > 65536) {
    static bool attr = false;
    if (!attr) {
      hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
      attr = true;
    }
  }
  hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(PassShape<L, 2, SV>::NT), lds, st, a);
  return hipGetLastError();
}

bool real_rowfft_supported(int elem_bytes, idx_t h, int L, bool all) {
  (void)elem_bytes;
  // 1024..4096-point rows measured 0-11 % faster than the real pass plus the
  // row pass; 8192-point rows (in a full 8192^2 r2c) 11 % slower: on request
  return h >= 2 && (L == 1024 || L == 2048 || L == 4096 || (all && L == 8192));
}

template <typename T>
hipError_t real_rowfft(const RealRowFftArgs<T>& a, hipStream_t st) {
  if (!real_rowfft_supported((int)sizeof(T), a.h, a.L, true)) return hipErrorInvalidValue;
  const idx_t blocks = a.nbatch * (a.h / 2 + 1);
  if (blocks <= 0) return hipSuccess;
  if (blocks > INT_MAX) return hipErrorInvalidValue;
  switch (a.L) {
    case 1024: return launch_rrf<T, 1024>(a, blocks, st);
    case 2048: return launch_rrf<T, 2048>(a, blocks, st);
    case 4096: return launch_rrf<T, 4096>(a, blocks, st);
    default: return launch_rrf<T, 8192>(a, blocks, st);
  }
}

template <typename T, int L>
static hipError_t launch_irrf(const RealIRowFftArgs<T>& a, idx_t blocks, hipStream_t st) {
  constexpr int SV = 0;
  constexpr size_t lds = PassShape<L, 2, SV>::template lds_bytes<T, true>();
  auto k = real_irowfft_kernel<T, L, SV>;
  if (lds > 65536) {
    static bool attr = false;
    if (!attr) {
      hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
      attr = true;
    }
  }
  hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(PassShape<L, 2, SV>::NT), lds, st, a);
  return hipGetLastError();
}

template <typename T>
hipError_t real_irowfft(const RealIRowFftArgs<T>& a, hipStream_t st) {
  if (!real_rowfft_supported((int)sizeof(T), a.h, a.L, true)) return hipErrorInvalidValue;
  const idx_t blocks = a.nbatch * (a.h / 2 + 1);
  if (blocks <= 0) return hipSuccess;
  if (blocks > INT_MAX) return hipErrorInvalidValue;
  switch (a.L) {
    case 1024: return launch_irrf<T, 1024>(a, blocks, st);
    case 2048: return launch_irrf<T, 2048>(a, blocks, st);
    case 4096: return launch_irrf<T, 4096>(a, blocks, st);
    default: return launch_irrf<T, 8192>(a, blocks, st);
  }
}

// ---- column pass B of a dim-0 four-step with the real pass ------------------
//
// N-D r2c over a dim 0 of n0 = 2h rows (h = h1 * L): the complex rows Z[r] =
// x[2r] + i x[2r+1] are transformed along r as a column four-step.  Pass A
// leaves Y[k1][j][c]; the L-point FFT over j gives Z[k1 + h1*k2][c] at bin
// k2.  The real pass over row pairs (multi.hpp:343-414, forward) forms
//     X[k] = (Z[k] + conj Z[h-k]) / 2 - i W_2h^k (Z[k] - conj Z[h-k]) / 2,
//     X[h-k] = conj of the two halves' terms in the same way,
// per column, and h - k = (h1 - k1) + h1*(L - 1 - k2) for k1 > 0, h1*((L -
// k2) mod L) for k1 = 0.  A work-group takes a block of CT columns and the
// tiles k1 = q and h1 - q (q = 0 .. h1/2; q = 0 and q = h1/2 pair with
// themselves): both column FFTs, tile h1 - q's bins through the LDS, then
// X for each bin k2 of tile q and its partner bin, written straight from the
// registers (rows of CT consecutive columns).  This replaces pass B plus the
// separate real pass, which read and wrote the output rows once more.
template <typename T, int L>
constexpr int rcp_ct() {
  // 34 KiB of LDS per group (one tile of L-point columns)
  return (sizeof(T) == 4 ? 64 : 32) * 64 / L;
}

template <typename T, int L>
__global__ void __launch_bounds__((PassShape<L, rcp_ct<T, L>(), 0>::NT)) real_colpair_kernel(RealColPairArgs<T> a) {
  constexpr int CT = rcp_ct<T, L>();
  using PS = PassShape<L, CT, 0>;
  constexpr int E = PS::E;
  constexpr int TPT = PS::TPT;
  __shared__ cx<T> lds[CT * LdsLen<L>::value];

  const int tid = threadIdx.x;
  const int c = tid % CT, t = tid / CT;
  const idx_t ncb = a.m / CT;
  const int nq = a.h1 / 2 + 1;
  idx_t g = blockIdx.x;
  const idx_t cb = g % ncb;
  g /= ncb;
  const int q = (int)(g % nq);
  const idx_t b = g / nq;
  const int ka = q;
  const int kb = (a.h1 - q) & (a.h1 - 1);
  const idx_t col = cb * CT + c;
  const cx<T>* __restrict__ src = a.in + b * a.in_bs + col;

  auto lat = [&](int i) -> cx<T>& { return lds[lds_pad(i) * CT + c]; };
  cx<T> va[E], vb[E];
  static_for<E>([&](auto mc) AFGPU_INLINE {
    constexpr int mm = decltype(mc)::value;
    va[mm] = src[((idx_t)ka * L + t + mm * TPT) * a.m];
  });
  if (kb != ka) {
    static_for<E>([&](auto mc) AFGPU_INLINE {
      constexpr int mm = decltype(mc)::value;
      vb[mm] = src[((idx_t)kb * L + t + mm * TPT) * a.m];
    });
  }
  run_stages<T, L, 0, 0, false, CT>(va, t, lat, a.stage_tw);
  // (kb != ka is uniform over the work-group: the barriers inside match)
  if (kb != ka) {
    run_stages<T, L, 0, 0, false, CT>(vb, t, lat, a.stage_tw);
  } else {
    static_for<E>([&](auto mc) AFGPU_INLINE { vb[decltype(mc)::value] = va[decltype(mc)::value]; });
  }
  // every exchange of run_stages ends with a barrier after its reads, so the
  // LDS is free: tile kb's bins k2 = t + mm*TPT at entry k2*CT + c
  static_for<E>([&](auto mc) AFGPU_INLINE {
    constexpr int mm = decltype(mc)::value;
    lds[(t + mm * TPT) * CT + c] = vb[mm];
  });
  __syncthreads();

  const idx_t h = (idx_t)a.h1 * L;
  const idx_t ob = b * a.out_bs + col;
  auto st = [&](idx_t row, cx<T> v) AFGPU_INLINE {
    const idx_t e = ob + row * a.m;
    if (a.out_il) {
      reinterpret_cast<cx<T>*>(a.out_re)[e] = v;
    } else {
      a.out_re[e] = v.x;
      a.out_im[e] = v.y;
    }
  };
  static_for<E>([&](auto mc) AFGPU_INLINE {
    constexpr int mm = decltype(mc)::value;
    const int k2 = t + mm * TPT;
    // a tile paired with itself writes each pair once
    const bool mine = ka != kb || (ka == 0 ? k2 <= L / 2 : k2 < L / 2);
    const int pk2 = ka == 0 ? ((L - k2) & (L - 1)) : (L - 1 - k2);
    const cx<T> s1 = lds[pk2 * CT + c];
    if (mine) {
      const idx_t k = (idx_t)ka + (idx_t)a.h1 * k2;
      const cx<T> s0 = va[mm];
      const cx<T> w = real_tw_pow2(a.real_tw, k, h);
      const cx<T> s1c = conjg(s1);
      const T hf = T(0.5);
      const cx<T> aa = {(s0.x + s1c.x) * hf, (s0.y + s1c.y) * hf};
      const cx<T> bw = cmul(s0 - s1c, w);
      const cx<T> bb = {bw.x * hf, bw.y * hf};
      st(k, aa + mul_neg_i(bb));
      if (h - k != k) st(h - k, conjg(aa) + mul_neg_i(conjg(bb)));
    }
  });
}

int real_colpair_ct(int elem_bytes, int L, idx_t m, int h1) {
  int ct = 0;
  if (L == 64) ct = elem_bytes == 4 ? rcp_ct<float, 64>() : rcp_ct<double, 64>();
  if (L == 128) ct = elem_bytes == 4 ? rcp_ct<float, 128>() : rcp_ct<double, 128>();
  if (L == 256) ct = elem_bytes == 4 ? rcp_ct<float, 256>() : rcp_ct<double, 256>();
  if (ct == 0 || m < ct || m % ct != 0 || h1 < 2 || (h1 & (h1 - 1)) != 0) return 0;
  return ct;
}

template <typename T, int L>
static hipError_t launch_rcp(const RealColPairArgs<T>& a, hipStream_t st) {
  constexpr int CT = rcp_ct<T, L>();
  const idx_t blocks = a.nbatch * (idx_t)(a.h1 / 2 + 1) * (a.m / CT);
  if (blocks <= 0) return hipSuccess;
  if (blocks > INT_MAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL((real_colpair_kernel<T, L>), dim3((unsigned)blocks), dim3(PassShape<L, CT, 0>::NT), 0, st, a);
  return hipGetLastError();
}

template <typename T>
hipError_t real_colpair(const RealColPairArgs<T>& a, hipStream_t st) {
  if (real_colpair_ct((int)sizeof(T), a.L, a.m, a.h1) == 0) return hipErrorInvalidValue;
  switch (a.L) {
    case 64: return launch_rcp<T, 64>(a, st);
    case 128: return launch_rcp<T, 128>(a, st);
    default: return launch_rcp<T, 256>(a, st);
  }
}

// ---- column pass A of a dim-0 four-step with the inverse real pass --------
//
// N-D c2r over a dim 0 of n0 = 2h rows (h = L * h2): after the inner axes the
// spectrum rows S[0..h] are paired by the inverse real pass (multi.hpp:343-414,
// inverse)
//     Z[k] = (S[k] + conj S[h-k]) - i conj(W_2h^k) (conj S[h-k] - S[k]),
//     Z[0] = (Re S[0] + Re S[h]) + i (Re S[0] - Re S[h]),
// and the inverse h-point c2c of Z runs as a column four-step whose pass A
// transforms, for each n2, the L-point column over j of rows k = j*h2 + n2.
// h - k = (L - 1 - j)*h2 + (h2 - n2) for n2 > 0 and (L - j)*h2 for n2 = 0, so
// the partners of block n2 lie in block h2 - n2.  A work-group takes a block
// of CT columns of blocks n2 = q and h2 - q (q = 0 .. h2/2; q = 0 and h2/2
// pair with themselves, block 0 also with row h), forms Z for both from the
// raw rows (each tile's partner values cross the LDS), and runs pass A on
// both: re / im swapped in (the inverse as a forward transform), the column
// FFTs, the W_h^(n2*k1) twiddle, the store into the scratch.  This replaces
// the separate inverse real pass, which read and wrote the spectrum once more.
template <typename T, int L>
__global__ void __launch_bounds__((PassShape<L, rcp_ct<T, L>(), 0>::NT)) real_icolpair_kernel(RealIColPairArgs<T> a) {
  constexpr int CT = rcp_ct<T, L>();
  using PS = PassShape<L, CT, 0>;
  constexpr int E = PS::E;
  constexpr int TPT = PS::TPT;
  __shared__ cx<T> lds[CT * LdsLen<L>::value];

  const int tid = threadIdx.x;
  const int c = tid % CT, t = tid / CT;
  const idx_t ncb = a.m / CT;
  const int nq = a.h2 / 2 + 1;
  idx_t g = blockIdx.x;
  const idx_t cb = g % ncb;
  g /= ncb;
  const int q = (int)(g % nq);
  const idx_t b = g / nq;
  const int na = q;
  const int nb = (a.h2 - q) & (a.h2 - 1);
  const bool self = na == nb;
  const idx_t col = cb * CT + c;
  const idx_t h = (idx_t)L * a.h2;
  const cx<T>* __restrict__ src = a.in + b * a.in_bs + col;

  // raw rows j*h2 + n2 of both blocks, j = t + mm*TPT
  cx<T> va[E], vb[E], pa[E], pb[E];
  static_for<E>([&](auto mc) AFGPU_INLINE {
    constexpr int mm = decltype(mc)::value;
    va[mm] = src[((idx_t)(t + mm * TPT) * a.h2 + na) * a.m];
  });
  if (!self) {
    static_for<E>([&](auto mc) AFGPU_INLINE {
      constexpr int mm = decltype(mc)::value;
      vb[mm] = src[((idx_t)(t + mm * TPT) * a.h2 + nb) * a.m];
    });
  } else {
    static_for<E>([&](auto mc) AFGPU_INLINE { vb[decltype(mc)::value] = va[decltype(mc)::value]; });
  }
  // row h: the partner of row 0 (block 0, j = 0: thread t = 0, register 0)
  const cx<T> sh = (na == 0 && t == 0) ? src[h * a.m] : cx<T>{T(0), T(0)};
  // partner entry j' of entry j of a block n2: L - 1 - j, or (L - j) mod L for n2 = 0
  auto pj = [&](int n2, int j) AFGPU_INLINE -> int { return n2 == 0 ? ((L - j) & (L - 1)) : (L - 1 - j); };
  // partners of block na's entries from block nb's raw values, then the reverse
  static_for<E>([&](auto mc) AFGPU_INLINE {
    constexpr int mm = decltype(mc)::value;
    lds[(t + mm * TPT) * CT + c] = vb[mm];
  });
  __syncthreads();
  static_for<E>([&](auto mc) AFGPU_INLINE {
    constexpr int mm = decltype(mc)::value;
    pa[mm] = lds[pj(na, t + mm * TPT) * CT + c];
  });
  __syncthreads();
  if (!self) {
    static_for<E>([&](auto mc) AFGPU_INLINE {
      constexpr int mm = decltype(mc)::value;
      lds[(t + mm * TPT) * CT + c] = va[mm];
    });
    __syncthreads();
    static_for<E>([&](auto mc) AFGPU_INLINE {
      constexpr int mm = decltype(mc)::value;
      pb[mm] = lds[pj(nb, t + mm * TPT) * CT + c];
    });
    __syncthreads();
  }
  // Z of every entry, re / im swapped for the forward-transform inverse
  auto pre = [&](cx<T> s0, cx<T> s1, idx_t k) AFGPU_INLINE -> cx<T> {
    const cx<T> s1c = conjg(s1);
    const cx<T> aa = s0 + s1c;
    const cx<T> bb = cmulc(s1c - s0, a.real_tw[k]);
    return swapri(aa + mul_neg_i(bb));
  };
  static_for<E>([&](auto mc) AFGPU_INLINE {
    constexpr int mm = decltype(mc)::value;
    const int j = t + mm * TPT;
    const idx_t ka = (idx_t)j * a.h2 + na;
    if (mm == 0 && na == 0 && t == 0)
      va[mm] = swapri(cx<T>{va[mm].x + sh.x, va[mm].x - sh.x});
    else
      va[mm] = pre(va[mm], pa[mm], ka);
    if (!self) vb[mm] = pre(vb[mm], pb[mm], (idx_t)j * a.h2 + nb);
  });

  auto lat = [&](int i) -> cx<T>& { return lds[lds_pad(i) * CT + c]; };
  const unsigned mask = (1u << a.tw4_bits) - 1u;
  cx<T>* __restrict__ dst = a.out + b * a.out_bs + col;
  // column FFTs, twiddle W_h^(n2*k1), bins k1 = t + mm*TPT to rows k1*h2 + n2
  auto finish = [&](cx<T>(&v)[E], int n2) AFGPU_INLINE {
    run_stages<T, L, 0, 0, false, CT>(v, t, lat, a.stage_tw);
    static_for<E>([&](auto mc) AFGPU_INLINE {
      constexpr int mm = decltype(mc)::value;
      const int k1 = t + mm * TPT;
      const unsigned e = (unsigned)n2 * (unsigned)k1;
      cx<T> w;
      if constexpr (std::is_same_v<T, float>)
        w = a.tw4_log2n > 0 ? tw_pow2f(e, a.tw4_log2n) : tw4_lookup32(a.tw4_hi, a.tw4_lo, a.tw4_bits, mask, e);
      else
        w = tw4_lookup32(a.tw4_hi, a.tw4_lo, a.tw4_bits, mask, e);
      dst[((idx_t)k1 * a.h2 + n2) * a.m] = cmul(v[mm], w);
    });
  };
  finish(va, na);
  // (self is uniform over the work-group: the barriers inside match)
  if (!self) finish(vb, nb);
}

template <typename T, int L>
static hipError_t launch_ircp(const RealIColPairArgs<T>& a, hipStream_t st) {
  constexpr int CT = rcp_ct<T, L>();
  const idx_t blocks = a.nbatch * (idx_t)(a.h2 / 2 + 1) * (a.m / CT);
  if (blocks <= 0) return hipSuccess;
  if (blocks > INT_MAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL((real_icolpair_kernel<T, L>), dim3((unsigned)blocks), dim3(PassShape<L, CT, 0>::NT), 0, st, a);
  return hipGetLastError();
}

template <typename T>
hipError_t real_icolpair(const RealIColPairArgs<T>& a, hipStream_t st) {
  if (real_colpair_ct((int)sizeof(T), a.L, a.m, a.h2) == 0) return hipErrorInvalidValue;
  switch (a.L) {
    case 64: return launch_ircp<T, 64>(a, st);
    case 128: return launch_ircp<T, 128>(a, st);
    default: return launch_ircp<T, 256>(a, st);
  }
}

template hipError_t real_icolpair<float>(const RealIColPairArgs<float>&, hipStream_t);
template hipError_t real_icolpair<double>(const RealIColPairArgs<double>&, hipStream_t);
template hipError_t real_colpair<float>(const RealColPairArgs<float>&, hipStream_t);
template hipError_t real_colpair<double>(const RealColPairArgs<double>&, hipStream_t);
template hipError_t real_irowfft<float>(const RealIRowFftArgs<float>&, hipStream_t);
template hipError_t real_irowfft<double>(const RealIRowFftArgs<double>&, hipStream_t);
template hipError_t real_rowfft<float>(const RealRowFftArgs<float>&, hipStream_t);
template hipError_t real_rowfft<double>(const RealRowFftArgs<double>&, hipStream_t);
template hipError_t real_pair<float>(const RealPairArgs<float>&, hipStream_t);
template hipError_t real_pair<double>(const RealPairArgs<double>&, hipStream_t);

}  // namespace afgpu
