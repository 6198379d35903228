// afft_tiny.hip -- rows of L <= 32 complex points, one row per thread.
//
// The general row kernel gives each thread of a short row pass a whole row
// read in place, so a wave's load of element j touches 64 rows, 8 bytes each,
// L*8 bytes apart (16-point rows: 0.37 TB/s).  Here a work-group stages its R
// rows through LDS instead: the loads and stores run over the R rows' elements
// in order (consecutive lanes, consecutive addresses; 16-point rows: 4.6
// TB/s), and each thread reads its row from LDS, runs the L-point DFT in
// registers (compile-time butterflies and twiddles, the real pass's W_2L^k
// included) with the real pass of r2c / c2r beside it, and writes its row back
// to LDS for the store.
//   c2c: X = DFT(x) (inverse: re / im exchanged around it, on each side as
//        the pass asks);
//   r2c: the reals as L pairs z[j] = x[2j] + i x[2j+1], Z = DFT(z), then afft's
//        real post-pass (onedim.hpp:1031-1108): X[0] = Re Z0 + Im Z0,
//        X[L] = Re Z0 - Im Z0, and for 0 < k < L
//        X[k] = (Z[k] + conj Z[L-k]) / 2 - i W_2L^k (Z[k] - conj Z[L-k]) / 2;
//   c2r: the inverse pre-pass z[k] = (X[k] + conj X[L-k]) - i conj(W_2L^k)
//        (conj X[L-k] - X[k]) (z[0] from the real parts of X[0] and X[L]),
//        then the inverse DFT, whose L outputs are the reals as pairs.
#include "afft_tiny.h"

namespace afgpu {

// LDS row pitch (complex entries): odd, so that the threads' reads of one
// element of their rows fall in different banks
template <int N>
constexpr int tiny_pitch() {
  return N | 1;
}

// R rows (threads) per work-group
template <typename T, int L, int EPI, int R>
__global__ void __launch_bounds__(R) tiny_rows_kernel(TinyArgs<T> a) {
  constexpr int LIN = EPI == EPI_C2R ? L + 1 : L;   // entries per input row
  constexpr int LOUT = EPI == EPI_R2C ? L + 1 : L;  // entries per output row
  constexpr int PIN = tiny_pitch<LIN>(), POUT = tiny_pitch<LOUT>();
  extern __shared__ __align__(16) unsigned char afgpu_smem[];
  cx<T>* __restrict__ lds = reinterpret_cast<cx<T>*>(afgpu_smem);
  const int tid = threadIdx.x;
  const idx_t r0 = (idx_t)blockIdx.x * R;
  const int nr = (int)(a.rows - r0 < R ? a.rows - r0 : R);

  // coalesced load: entry e of the group's rows (row e / LIN, element e % LIN)
  for (int e = tid; e < nr * LIN; e += R) {
    const int r = e / LIN, j = e - r * LIN;
    lds[r * PIN + j] = ld_elem(a.in_re, a.in_im, a.in_il, (r0 + r) * a.in_rs + j, a.in_nt);
  }
  __syncthreads();
  cx<T> v[L];
  const bool live = tid < nr;
  if (live) {
    const cx<T>* row = lds + tid * PIN;
    if constexpr (EPI == EPI_C2R) {
      static_for<L>([&](auto kc) AFGPU_INLINE {
        constexpr int k = decltype(kc)::value;
        const cx<T> s0 = row[k], s1 = row[L - k];
        cx<T> z;
        if constexpr (k == 0) {
          z = {s0.x + s1.x, s0.x - s1.x};
        } else {
          const cx<T> s1c = conjg(s1);
          z = s0 + s1c + mul_neg_i(twc<T, 2 * L, -k>(s1c - s0));
        }
        v[k] = swapri(z);
      });
    } else {
      static_for<L>([&](auto kc) AFGPU_INLINE {
        constexpr int k = decltype(kc)::value;
        v[k] = a.swap_in ? swapri(row[k]) : row[k];
      });
    }
  }
  __syncthreads();  // rows are rewritten below
  if (live) {
    Dft<T, L>::run(v);
    cx<T>* row = lds + tid * POUT;
    if constexpr (EPI == EPI_R2C) {
      const T h = T(0.5);
      static_for<L>([&](auto kc) AFGPU_INLINE {
        constexpr int k = decltype(kc)::value;
        const cx<T> s0 = v[k];
        if constexpr (k == 0) {
          row[0] = {s0.x + s0.y, T(0)};
          row[L] = {s0.x - s0.y, T(0)};
        } else {
          const cx<T> s1c = conjg(v[L - k]);
          const cx<T> av = {(s0.x + s1c.x) * h, (s0.y + s1c.y) * h};
          const cx<T> bw = twc<T, 2 * L, k>(s0 - s1c);
          row[k] = av + mul_neg_i(cx<T>{bw.x * h, bw.y * h});
        }
      });
    } else if constexpr (EPI == EPI_C2R) {
      static_for<L>([&](auto kc) AFGPU_INLINE { row[decltype(kc)::value] = swapri(v[decltype(kc)::value]); });
    } else {
      static_for<L>([&](auto kc) AFGPU_INLINE {
        constexpr int k = decltype(kc)::value;
        row[k] = a.swap_out ? swapri(v[k]) : v[k];
      });
    }
  }
  __syncthreads();
  for (int e = tid; e < nr * LOUT; e += R) {
    const int r = e / LOUT, j = e - r * LOUT;
    st_elem(a.out_re, a.out_im, a.out_il, (r0 + r) * a.out_rs + j, lds[r * POUT + j], a.out_nt);
  }
}

template <typename T, int L, int EPI, int R>
static hipError_t launch_tiny(const TinyArgs<T>& a, hipStream_t st) {
  constexpr int LMAX = EPI == EPI_PLAIN ? L : L + 1;
  constexpr size_t lds = (size_t)R * tiny_pitch<LMAX>() * sizeof(cx<T>);
  static_assert(lds <= 160 * 1024, "LDS of one work-group");
  auto k = tiny_rows_kernel<T, L, EPI, R>;
  if (lds > 65536) {
    static bool attr = false;
    if (!attr) {
      hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
      attr = true;
    }
  }
  if (a.rows <= 0) return hipSuccess;
  const idx_t blocks = (a.rows + R - 1) / R;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(R), lds, st, a);
  return hipGetLastError();
}

int tiny_default_rows(int elem_bytes, int L, int epi) {
  // 32-point rows: the general row kernel (64 rows per group, four threads
  // per row, cached accesses) measured faster in every kind and precision
  // (FP32 c2c 4.3 TB/s against 2.8 here, r2c 3.4 / 2.7, c2r 3.6 / 2.5)
  (void)epi;
  if (L >= 32) return 0;
  // 64, 128 and 256 rows per group measured within 3% of each other on every
  // length (256 MiB per batch); smaller groups for longer rows
  const int row_bytes = L * 2 * elem_bytes;
  return row_bytes <= 128 ? 256 : row_bytes <= 256 ? 128 : 64;
}

bool tiny_supported(int L, int epi, int rows) {
  if (epi != EPI_PLAIN && epi != EPI_R2C && epi != EPI_C2R) return false;
  if (rows != 64 && rows != 128 && rows != 256) return false;
  return L == 2 || L == 4 || L == 8 || L == 16 || L == 32;
}

template <typename T>
hipError_t tiny_rows(const TinyArgs<T>& a, int L, int epi, int rows, hipStream_t st) {
#define AFGPU_TINY_R(LL, EE)                                              \
  if (rows == 64) return launch_tiny<T, LL, EE, 64>(a, st);               \
  if (rows == 128) return launch_tiny<T, LL, EE, 128>(a, st);             \
  if (rows == 256) return launch_tiny<T, LL, EE, 256>(a, st);
#define AFGPU_TINY_L(LL)                                                  \
  if (L == LL) {                                                          \
    if (epi == EPI_PLAIN) { AFGPU_TINY_R(LL, EPI_PLAIN) }                 \
    if (epi == EPI_R2C) { AFGPU_TINY_R(LL, EPI_R2C) }                     \
    if (epi == EPI_C2R) { AFGPU_TINY_R(LL, EPI_C2R) }                     \
  }
  AFGPU_TINY_L(2)
  AFGPU_TINY_L(4)
  AFGPU_TINY_L(8)
  AFGPU_TINY_L(16)
  AFGPU_TINY_L(32)
#undef AFGPU_TINY_L
#undef AFGPU_TINY_R
  return hipErrorInvalidValue;
}

template hipError_t tiny_rows<float>(const TinyArgs<float>&, int, int, int, hipStream_t);
template hipError_t tiny_rows<double>(const TinyArgs<double>&, int, int, int, hipStream_t);

}  // namespace afgpu
