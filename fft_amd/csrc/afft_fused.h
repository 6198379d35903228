// afft_fused.h -- both passes of a 1-D four-step c2c in one persistent launch.
#pragma once

#include <hip/hip_runtime.h>

#include "afft_pass.h"

namespace afgpu {

// Counter words (ints) of one fused launch, placed after its scratch ring:
// 8 pop counters, a status word and two done counters per transform, each on
// its own 64-byte line.
constexpr int kFusedXcd = 8;     // XCDs of an MI355X (one task list each)
constexpr int kFusedLine = 16;   // ints per 64-byte line
inline size_t fused_counter_ints(idx_t batch) { return (size_t)kFusedLine * (kFusedXcd + 1 + 2 * (size_t)batch); }

// One launch over `batch` transforms of N = n1 * n2 points (n1 = the pass-A
// column length, n2 = the pass-B row length, both compile-time in the kernel),
// interleaved FP32, transform t at src + t * src_stride / dst + t * dst_stride.
struct FusedArgs {
  const float* src;
  float* dst;
  float* ring;           // kFusedXcd * slots slots of N complex
  idx_t src_stride;      // floats between consecutive transforms of the input
  idx_t dst_stride;      // floats between consecutive transforms of the output
  const cx<float>* stage_a;  // stage twiddles of the n1-point columns
  const cx<float>* stage_b;  // stage twiddles of the n2-point rows
  const cx<float>* tw4_hi;   // W_N^m = hi[m >> bits] * (1 + lo[m & mask])
  const cx<float>* tw4_lo;
  int tw4_bits;
  int swap;              // inverse transform: re / im swapped on load and store
  int in_nt, out_nt;     // streaming loads of the input / stores of the output
  int batch;             // transforms in this launch
  int slots;             // ring slots per XCD
  int lag;               // rounds between a transform's pass A and its pass B
  int* ctr;              // fused_counter_ints(batch) ints, zeroed by the launch's reset kernel
  long long spin_limit;  // dependency-wait iterations before the status word is set
};

// Shapes with a fused kernel: pass A of la-point columns (cta per tile,
// schedule sva, half exchange hxa) and pass B of lb-point rows (ctb per tile,
// svb, hxb, transposed store).
bool fused_supported(int la, int cta, int sva, int hxa, int lb, int ctb, int svb, int hxb);
// Launches the counter reset and the persistent kernel on `st` (kernels only).
hipError_t fused_launch(const FusedArgs& a, int la, int cta, int sva, int hxa, int lb, int ctb, int svb, int hxb,
                        hipStream_t st);

}  // namespace afgpu
