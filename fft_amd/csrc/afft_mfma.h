// afft_mfma.h -- host interface of the split-precision MFMA DFT-256 tile.
#pragma once

#include <hip/hip_runtime.h>

#include "afft_device.h"

namespace afgpu {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));

struct MfmaParams {
  const float* in;  // interleaved complex
  float* out;
  idx_t in_js, in_cs, in_bs, in_bs2;
  idx_t out_ks, out_cs, out_bs, out_bs2;
  int n_blo;
  int n_col, col_tiles;
  int swap_in, swap_out;
  int in_cfast, out_cfast;  // 1: columns contiguous in memory (c fastest when moving data)
  int in_nt, out_nt;        // streaming (non-temporal) loads / stores
  int bf16;                 // 1: BF16 three-way split operands (table from gen_mfma_mats(.., 1, ..))
  const h8* mats;
  const cx<float>* tw4_hi;  // optional four-step twiddle on the output
  const cx<float>* tw4_lo;
  int tw4_bits;
  int tw4_shift;
  int tw4_log2n;  // > 0: compute W_N^m (N = 2^tw4_log2n) instead of the table lookup
};

hipError_t gen_mfma_mats(void* out, int bf16, hipStream_t st);
size_t mfma_mats_bytes(int bf16);
hipError_t launch_mfma_dft256(const MfmaParams& p, int n_tiles, hipStream_t st);

}  // namespace afgpu
