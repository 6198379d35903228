// afft_real4.h -- the 1-D r2c of a long real sequence as a real-input
// four-step: two passes over half-size data instead of a complex four-step of
// the half-length sequence followed by a real post-pass.
#pragma once

#include <hip/hip_runtime.h>

#include "afft_pass.h"

namespace afgpu {

// N = n1 * n2 reals per transform, x[j1 * n2 + j2] (n1 rows of n2).
//   Pass A: the n1-point DFTs down the columns j2 of each transform (two real
//   columns 2c, 2c+1 read as one complex column), unpacked, multiplied by
//   W_N^(j2*k1) and stored for k1 = 0..n1/2 as Y[k1][j2] (row k1 at k1*n2,
//   interleaved scratch).  Rows n1/2 < k1 < n1 are not needed: the columns
//   are real, so C[n1 - k1] = conj C[k1].
//   Pass B: the n2-point DFTs of rows k1 = 0..n1/2 give X[k1 + n1*k2]; for
//   0 < k1 < n1/2 the same output also gives the bins of row n1 - k1:
//   X[(n1 - k1) + n1*k2] = conj X[k1 + n1*(n2 - 1 - k2)].  Every bin 0..N/2
//   is stored once, in natural order.
// The c2r runs the same two passes backwards (real4c, real4d):
//   pass B' takes, for k1 = 0..n1/2, the bins X[k1 + n1*k2] up to N/2 and
//   the rest as conj X[(n1 - k1) + n1*(n2 - 1 - k2)] (k1 = 0: conj
//   X[n1*(n2 - k2)]), imaginary parts of bins 0 and N/2 taken as zero -- CT
//   consecutive k1 per group, so both loads are runs of CT bins -- runs the
//   inverse n2-point DFTs, multiplies by W_N^-(j2*k1) and stores
//   Dt[j2][k1] (rows of rs >= n1/2 + 1 entries);
//   pass A' forms, for two real columns 2c, 2c+1 at once, the column
//   spectrum Dt[2c][k1] + i Dt[2c+1][k1] for all k1 (k1 > n1/2 from the
//   conjugates at n1 - k1), runs the inverse n1-point DFT and stores the two
//   columns of reals through LDS, CT column pairs per output row.
template <typename T>
struct Real4Args {
  const T* in;           // r2c pass A: the reals (an interleaved complex view of pairs);
                         // c2r pass B': the spectrum (re array)
  const T* in_im;        // c2r pass B', split format: imaginary parts
  int in_il;
  cx<T>* scr;            // the intermediate: r2c (n1/2 + 1) rows of n2; c2r n2 rows of rs
  T* out_re;             // r2c pass B: bins 0..N/2; c2r pass A': the reals
  T* out_im;             // (split format: imaginary parts; unused when interleaved)
  int out_il;
  idx_t n1, n2;
  idx_t in_bs;           // input elements (reals, or complex when interleaved) between transforms
  idx_t scr_bs;          // complex entries between transforms in the scratch
  idx_t out_bs;          // output elements (complex when interleaved, reals for c2r) between transforms
  idx_t nbatch;
  int in_nt, out_nt;     // streaming loads of the input / stores of the output
  const cx<T>* stage_tw; // stage twiddles of the pass's DFT length
  const cx<T>* tw4_hi;   // W_N^e = hi[e >> bits] * (1 + lo[e & mask]) (pass A)
  const cx<T>* tw4_lo;
  int tw4_bits;
  int tw4_log2n;         // > 0 (FP32): compute W_N^e, N = 2^tw4_log2n, instead of the table lookup
  int ct, sv, hx;        // the compiled shape to launch (tile, stage schedule, half exchange)
  idx_t rs;              // c2r: entries per row of the scratch Dt[j2][k1] (>= n1/2 + 1)
};

// Compiled shapes of pass 0 (A: n1-point columns), 1 (B: n2-point rows),
// 2 (c2r B': n2-point columns of the spectrum) or 3 (c2r A': n1-point rows
// of the scratch).  Picks the shape with tile want_ct (> 0) and schedule
// want_sv (>= 0) and half exchange want_hx (>= 0) when compiled, else the
// pass's default for the length; false when none exists.
bool real4_shape(int pass, int elem_bytes, int len, int want_ct, int want_sv, int want_hx, int& ct, int& sv,
                 int& hx);

template <typename T>
hipError_t real4a(const Real4Args<T>& a, hipStream_t st);
template <typename T>
hipError_t real4b(const Real4Args<T>& a, hipStream_t st);
template <typename T>
hipError_t real4c(const Real4Args<T>& a, hipStream_t st);
template <typename T>
hipError_t real4d(const Real4Args<T>& a, hipStream_t st);

}  // namespace afgpu
