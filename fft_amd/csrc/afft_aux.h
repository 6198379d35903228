// afft_aux.h -- declarations of the auxiliary kernels in afft_aux.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "afft_device.h"

namespace afgpu {

struct StageList {
  int nstage;
  int total;
  int radix[8];
  int ns[8];
  int off[8];
};

template <typename T>
struct RealRowsArgs {
  const T* in_re;
  const T* in_im;
  T* out_re;
  T* out_im;
  int in_il, out_il;
  idx_t in_rs, in_cs, in_bs;    // row, column, batch strides (elements)
  idx_t out_rs, out_cs, out_bs;
  idx_t h, m, nbatch;
  const cx<T>* real_tw;  // W_{2h}^k, k < h
};

// Forward real pass that also undoes the four-step transposition: the input
// is Y[b][k1][k2] (n1 x n2, k2 contiguous, interleaved) holding the h-point
// c2c result Z[k1 + n1*k2]; the output is the h+1 bins of the r2c in natural
// order.
template <typename T>
struct RealTrArgs {
  const cx<T>* in;
  T* out_re;
  T* out_im;
  int out_il;
  idx_t in_bs, out_bs;  // batch strides (complex / scalar elements)
  int n1, n2;
  idx_t h, nbatch;
  const cx<T>* real_tw;  // W_{2h}^k, k < h
};

// Pass B of the four-step 1-D r2c fused with the real post-pass and the
// four-step transposition (afft_realpair.hip): the input is the pass-A
// output Y[b][k1][n2'] (n1 rows of n2, interleaved); each work-group runs the
// n2-point FFTs of 2R rows k1 and n1 - k1, pairs bin k2 of row k1 with bin
// n2 - 1 - k2 of row n1 - k1 (Z[k] with Z[h - k], k = k1 + n1*k2) and writes
// the h+1 bins of the r2c in natural order.
template <typename T>
struct RealPairArgs {
  const cx<T>* in;
  T* out_re;
  T* out_im;
  int out_il;
  int out_nt;            // non-temporal stores
  idx_t in_bs, out_bs;   // batch strides (complex / output elements)
  int n1, n2;
  idx_t nbatch;
  const cx<T>* stage_tw; // stage twiddles of the n2-point rows (radix-32 schedule)
  const cx<T>* real_tw;  // W_{2h}^k, k < h
  const cx<T>* row_tw;   // W_{2 n2}^k, k < n2  (= W_{2h}^(n1 k))
};

// N-D r2c: the real pass over dst row pairs (k, h-k) fused into the FFT along
// the last axis (L points, rows of dst in place; afft_realpair.hip).  Row r of
// batch entry b starts at element b*bs + r*rs (complex units when il, scalar
// otherwise); rows 0..h-1 hold Z on entry, rows 0..h hold the r2c on exit.
template <typename T>
struct RealRowFftArgs {
  T* re;
  T* im;
  int il;
  int L;
  idx_t h, rs, bs, nbatch;
  const cx<T>* stage_tw;  // stage twiddles of the L-point rows (radix-16 schedule)
  const cx<T>* real_tw;   // W_{2h}^k, k < h
};

// N-D c2r: the inverse real pass over row pairs fused into the inverse FFT
// along the last axis: input rows 0..h of the user's spectrum (in_rs / in_bs,
// complex units when in_il, scalar otherwise), output rows 0..h-1 of Z into
// the plan's scratch (interleaved; afft_realpair.hip).
template <typename T>
struct RealIRowFftArgs {
  const T* in_re;
  const T* in_im;
  int in_il;
  cx<T>* out;
  int L;
  idx_t h, in_rs, in_bs, out_rs, out_bs, nbatch;
  const cx<T>* stage_tw;  // stage twiddles of the L-point rows (radix-16 schedule)
  const cx<T>* real_tw;   // W_{2h}^k, k < h
};

// N-D r2c with the halved dim 0 as a column four-step h = h1 * L: pass B
// (the L-point FFTs over j of Y[k1][j][c], bins k1 + h1*k2) with the real
// pass over row pairs (k, h-k) fused in (afft_realpair.hip).  Y is the plan's
// scratch (interleaved, entry (k1, j, c) at k1*L*m + j*m + c of each batch
// entry); the output holds rows 0..h of m entries (row r at r*m; complex
// units when out_il, scalar otherwise).
template <typename T>
struct RealColPairArgs {
  const cx<T>* in;
  T* out_re;
  T* out_im;
  int out_il;
  int L;                  // points of the column FFTs
  int h1;                 // rows k1 of the four-step
  idx_t m;                // columns
  idx_t in_bs, out_bs;    // batch strides (complex / output elements)
  idx_t nbatch;
  const cx<T>* stage_tw;  // stage twiddles of the L-point columns (radix-16 schedule)
  const cx<T>* real_tw;   // W_{2h}^k, k < h
};

// Columns per work-group of real_colpair / real_icolpair (0: no kernel for
// these shapes; L = the column FFT length, h1 = the other factor of h).
int real_colpair_ct(int elem_bytes, int L, idx_t m, int h1);

// N-D c2r with the halved dim 0 as a column four-step h = L * h2: the inverse
// real pass over row pairs (k, h-k) fused into pass A (the L-point FFTs over
// j of the rows j*h2 + n2 with the W_h^(n2*k1) twiddle; afft_realpair.hip).
// Input: spectrum rows 0..h of m entries (row r at r*m, interleaved); output:
// the plan's scratch, entry (k1, n2, c) at k1*h2*m + n2*m + c.
template <typename T>
struct RealIColPairArgs {
  const cx<T>* in;
  cx<T>* out;
  int L;                  // points of the column FFTs (h1)
  int h2;                 // column blocks n2
  idx_t m;                // columns
  idx_t in_bs, out_bs;    // batch strides (complex entries)
  idx_t nbatch;
  const cx<T>* stage_tw;  // stage twiddles of the L-point columns (radix-16 schedule)
  const cx<T>* real_tw;   // W_{2h}^k, k < h
  const cx<T>* tw4_hi;    // W_h^e = tw4_hi[e >> tw4_bits] * tw4_lo[e & mask]
  const cx<T>* tw4_lo;
  int tw4_bits;
  int tw4_log2n;          // > 0 (FP32): compute W_h^e, h = 2^tw4_log2n, instead of the table lookup
};

// all = false: only the row lengths where the fused pass is the default
bool real_rowfft_supported(int elem_bytes, idx_t h, int L, bool all);

// Whether real_pair has a kernel for rows of n2 points and n1 rows.
bool real_pair_supported(int elem_bytes, int n1, int n2);

template <typename T>
hipError_t gen_stage_tw(cx<T>* out, const StageList& sl, hipStream_t st);
template <typename T>
hipError_t gen_tw4(cx<T>* hi, cx<T>* lo, idx_t N, int bits, hipStream_t st);
template <typename T>
hipError_t gen_real_tw(cx<T>* out, idx_t L, hipStream_t st);
template <typename T>
hipError_t real_rows(const RealRowsArgs<T>& a, bool inverse, hipStream_t st);
template <typename T>
hipError_t real_tr(const RealTrArgs<T>& a, hipStream_t st);
template <typename T>
hipError_t real_pair(const RealPairArgs<T>& a, hipStream_t st);
template <typename T>
hipError_t real_rowfft(const RealRowFftArgs<T>& a, hipStream_t st);
template <typename T>
hipError_t real_irowfft(const RealIRowFftArgs<T>& a, hipStream_t st);
template <typename T>
hipError_t real_colpair(const RealColPairArgs<T>& a, hipStream_t st);
template <typename T>
hipError_t real_icolpair(const RealIColPairArgs<T>& a, hipStream_t st);
template <typename T>
hipError_t real_one(const T* src, T* dst_re, T* dst_im, int inverse, int il, idx_t nb, idx_t in_s, idx_t out_s,
                    hipStream_t st);
template <typename TY>
hipError_t err_sums(const TY* y, const double* r, idx_t n, double* acc, hipStream_t st);
hipError_t copy16(const void* src, void* dst, idx_t bytes, hipStream_t st);

}  // namespace afgpu
