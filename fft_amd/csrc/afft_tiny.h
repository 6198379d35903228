// afft_tiny.h -- batches of tiny rows (afft's tiny path, onedim.hpp:548-753):
// rows of L <= 32 complex points (c2c n <= 32, r2c / c2r n <= 64 reals), one
// row per thread, staged through LDS so that every global access of a block
// is a contiguous run over its rows.
#pragma once

#include <hip/hip_runtime.h>

#include "afft_pass.h"

namespace afgpu {

template <typename T>
struct TinyArgs {
  const T* in_re;   // rows of L complex (c2c, r2c: the reals as pairs) or L + 1 bins (c2r)
  const T* in_im;   // split input: imaginary parts
  int in_il;
  int in_nt;        // non-temporal loads
  T* out_re;        // rows of L complex (c2c, c2r: the reals as pairs) or L + 1 bins (r2c)
  T* out_im;        // split output: imaginary parts
  int out_il;
  int out_nt;       // non-temporal stores
  idx_t rows;       // rows in this launch
  idx_t in_rs;      // elements between rows (complex entries, or reals of each split array)
  idx_t out_rs;
  int swap_in;      // exchange re / im of the loaded rows (inverse c2c: around a forward DFT)
  int swap_out;     // exchange re / im of the stored rows
};

// Default rows per work-group (one per thread: 64, 128 or 256; 0 where the
// general row kernel is faster) and whether a kernel exists for (L,
// epilogue, rows per work-group).
int tiny_default_rows(int elem_bytes, int L, int epi);
bool tiny_supported(int L, int epi, int rows);

// epi: EPI_PLAIN (c2c), EPI_R2C, EPI_C2R; rows: rows per work-group.
template <typename T>
hipError_t tiny_rows(const TinyArgs<T>& a, int L, int epi, int rows, hipStream_t st);

}  // namespace afgpu
