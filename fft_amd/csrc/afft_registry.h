// afft_registry.h -- host-side table of the compiled pass-kernel variants.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "afft_pass.h"

namespace afgpu {

template <typename T>
using PassLaunchFn = hipError_t (*)(const PassParams<T>&, hipStream_t, bool fast);

template <typename T>
struct PassVariant {
  int L, CT, colmap, epi, store_t, sv;
  int hx;  // half exchange: real and imaginary parts cross the LDS in turn
  int threads;
  size_t lds_bytes;
  PassLaunchFn<T> launch;
};

// Looks up the compiled variant; returns nullptr when none exists.
template <typename T>
const PassVariant<T>* find_pass(int L, int CT, int colmap, int epi, int store_t, int sv = 0, int hx = 0);

// Largest L with a compiled row / column kernel, and the column tile (CT) used
// for a column pass of length L.
template <typename T>
int max_row_len();  // longest row pass (the longest ones have half-LDS kernels only)
template <typename T>
int max_row_len_full();  // longest row pass with a full-LDS kernel
template <typename T>
int max_col_len();
template <typename T>
int col_tile(int L);
template <typename T>
int row_tile(int L);
template <typename T>
int rowt_tile(int L);  // transposed-store row pass (four-step pass B)

// Fills host-side descriptors of the stages of schedule (L, sv): radix and Ns
// per stage; returns the number of stages.
int stage_list(int L, int sv, int* radix, int* ns);

}  // namespace afgpu
