// afft_fused.hip -- a batch of 1-D four-step c2c transforms (N = n1 * n2) with
// both passes in ONE persistent launch, so that HBM sees only the input reads
// and the output writes.
//
// The two-pass plan (afft_plan.hip) runs pass A (n1-point columns with the
// W_N twiddle, input -> scratch) and pass B (n2-point rows stored columns-
// fastest, scratch -> output) as separate launches over groups of transforms:
// the scratch round trip then crosses the memory side of the chip twice per
// transform.  Here every transform goes to one XCD (transform t to XCD t % 8)
// and lives in one of `slots` ring slots of that XCD while its two passes run:
//   list of XCD x, round r:  pass-A tiles of its transform r,
//                            then pass-B tiles of its transform r - lag;
// work-groups read their XCD from the hardware (HW_REG_XCC_ID) and pop that
// XCD's list through an atomic counter.  A pass-B tile waits until every
// pass-A tile of its transform is done; a pass-A tile of transform k >= slots
// waits until every pass-B tile of transform k - slots (the previous user of
// its slot) is done.  Both waits are on tasks listed earlier in the same
// list, which were popped by running work-groups, so every wait ends; a spin
// limit bounds them anyway and sets the status word.
//
// Coherence follows from the placement: producer and consumer tiles of a
// transform share one L2 (the vector L1 writes through), so a tile publishes
// after all its stores completed (s_waitcnt) with one relaxed counter add, and
// a pass-B tile invalidates its CU's L1 after its wait before it reads the
// slot.  The ring (8 XCDs x slots x N complex) is small enough to stay in the
// 256 MB MALL; the input is read and the output written with streaming
// (non-temporal) accesses.
//
// The tile bodies are the pass kernels' own (fft_tile), so the result is bit
// for bit that of the two-pass plan.
#include "afft_fused.h"

namespace afgpu {

__device__ __forceinline__ int fused_xcc_id() {
  unsigned v;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
  return (int)(v & (kFusedXcd - 1));
}

// One tile of each pass, as functions of their own: inlined into the task
// loop, the compiler keeps values of one task live across the next and
// spills; called, each body gets the register allocation of its standalone
// pass kernel.  Pointers arrive in vector registers and are made uniform
// first (buffer resources need scalar registers).
template <typename P>
__device__ __forceinline__ P uniform(P p) {
  // readfirstlane moves 32 bits: a 64-bit address takes two
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (P)(((unsigned long long)hi << 32) | lo);
}

template <int LA, int CTA, int SVA, bool HXA, int LB>
__device__ __noinline__ void fused_tile_a(const float* src, float* slot, const cx<float>* stage,
                                          const cx<float>* tw4_hi, const cx<float>* tw4_lo, int tw4_bits, int swap,
                                          int in_nt, int tile, int tid, cx<float>* lds) {
  // LA-point columns (stride LB) of one transform, W_N twiddle, into its slot
  constexpr idx_t N = (idx_t)LA * LB;
  PassParams<float> p{};
  p.in_il = p.out_il = 1;
  p.n_blo = 1;
  p.in_bytes = p.out_bytes = (unsigned)(N * sizeof(cx<float>));
  p.in_re = uniform(src);
  p.out_re = uniform(slot);
  p.in_js = p.out_ks = LB;
  p.in_cs = p.out_cs = 1;
  p.in_bs = p.out_bs = N;
  p.n_col = LB;
  p.col_tiles = p.n_tiles = LB / CTA;
  p.swap_in = __builtin_amdgcn_readfirstlane(swap);
  p.in_nt = __builtin_amdgcn_readfirstlane(in_nt);
  p.stage_tw = uniform(stage);
  p.tw4_hi = uniform(tw4_hi);
  p.tw4_lo = uniform(tw4_lo);
  p.tw4_lo_bits = __builtin_amdgcn_readfirstlane(tw4_bits);
  fft_tile<float, LA, CTA, true, EPI_TW4, false, true, SVA, HXA>(p, __builtin_amdgcn_readfirstlane(tile), tid, lds);
}

template <int LA, int LB, int CTB, int SVB, bool HXB>
__device__ __noinline__ void fused_tile_b(const float* slot, float* dst, const cx<float>* stage, int swap, int out_nt,
                                          int tile, int tid, cx<float>* lds) {
  // LB-point rows of the slot, bins k1 + LA*k2 into the transform's output
  constexpr idx_t N = (idx_t)LA * LB;
  PassParams<float> p{};
  p.in_il = p.out_il = 1;
  p.n_blo = 1;
  p.in_bytes = p.out_bytes = (unsigned)(N * sizeof(cx<float>));
  p.in_re = uniform(slot);
  p.out_re = uniform(dst);
  p.in_js = 1;
  p.in_cs = LB;
  p.out_ks = LA;
  p.out_cs = 1;
  p.in_bs = p.out_bs = N;
  p.n_col = LA;
  p.col_tiles = p.n_tiles = LA / CTB;
  p.swap_out = __builtin_amdgcn_readfirstlane(swap);
  p.out_nt = __builtin_amdgcn_readfirstlane(out_nt);
  p.stage_tw = uniform(stage);
  fft_tile<float, LB, CTB, false, EPI_PLAIN, true, true, SVB, HXB>(p, __builtin_amdgcn_readfirstlane(tile), tid, lds);
}

template <int LA, int CTA, int SVA, bool HXA, int LB, int CTB, int SVB, bool HXB>
struct FusedShape {
  static constexpr int NT = PassShape<LA, CTA, SVA>::NT;
  static_assert(NT == PassShape<LB, CTB, SVB>::NT, "both passes use one group size");
  static constexpr size_t lds_a = tile_lds_bytes<float, LA, CTA, true, false, SVA, HXA>();
  static constexpr size_t lds_b = tile_lds_bytes<float, LB, CTB, false, true, SVB, HXB>();
  static constexpr size_t lds = lds_a > lds_b ? lds_a : lds_b;
  // work-groups per CU the registers are bounded for
  static constexpr int groups = 2 * (lds + 64) <= 160 * 1024 ? 2 : 1;
  static constexpr int min_waves = groups * NT / 256 > 1 ? groups * NT / 256 : 1;
};

template <int LA, int CTA, int SVA, bool HXA, int LB, int CTB, int SVB, bool HXB>
__global__ void __launch_bounds__((FusedShape<LA, CTA, SVA, HXA, LB, CTB, SVB, HXB>::NT),
                                  (FusedShape<LA, CTA, SVA, HXA, LB, CTB, SVB, HXB>::min_waves))
    fused4_kernel(FusedArgs a) {
  extern __shared__ __align__(16) unsigned char afgpu_smem[];
  cx<float>* lds = reinterpret_cast<cx<float>*>(afgpu_smem);
  __shared__ int s_task;
  const int tid = threadIdx.x;
  const int x = fused_xcc_id();
  const int nxcd = kFusedXcd;
  const int T = x < a.batch ? (a.batch - x + nxcd - 1) / nxcd : 0;  // transforms of this XCD
  constexpr int nA = LB / CTA, nB = LA / CTB;  // tiles per transform of pass A / pass B
  const int total = T * (nA + nB);
  int* const pop = a.ctr + x * kFusedLine;
  int* const status = a.ctr + nxcd * kFusedLine;
  int* const done = a.ctr + (nxcd + 1) * kFusedLine;  // transform t: A at 2t, B at 2t+1 (lines)
  for (;;) {
    if (tid == 0) s_task = __hip_atomic_fetch_add(pop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    // uniform over the group: keeps the task decode and every address derived
    // from it in scalar registers
    const int i = __builtin_amdgcn_readfirstlane(s_task);
    __syncthreads();
    if (i >= total) break;
    // round r holds the A tiles of transform r (r < T) and the B tiles of
    // transform r - lag (lag <= r < T + lag)
    int base = 0, k = 0, tile = 0;
    bool is_a = false;
    for (int r = 0;; r++) {
      const int na = r < T ? nA : 0;
      const int nb = (r >= a.lag && r - a.lag < T) ? nB : 0;
      if (i < base + na) {
        is_a = true;
        k = r;
        tile = i - base;
        break;
      }
      if (i < base + na + nb) {
        k = r - a.lag;
        tile = i - base - na;
        break;
      }
      base += na + nb;
    }
    const int t = x + nxcd * k;
    const idx_t slot = (idx_t)x * a.slots + k % a.slots;
    int* wait_on = nullptr;
    int need = 0;
    if (is_a) {
      if (k >= a.slots) {
        wait_on = done + (2 * (t - nxcd * a.slots) + 1) * kFusedLine;
        need = nB;
      }
    } else {
      wait_on = done + 2 * t * kFusedLine;
      need = nA;
    }
    if (wait_on) {
      if (tid == 0) {
        long long it = 0;
        // polled with an atomic, performed in the L2 (never a stale L1 line)
        while (__hip_atomic_fetch_add(wait_on, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
          if (++it > a.spin_limit) {
            __hip_atomic_fetch_or(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
      }
      __syncthreads();
      // pass B reads the slot: drop lines of it this CU's L1 may still hold
      if (!is_a) asm volatile("buffer_inv sc0" ::: "memory");
    }
    if (is_a)
      fused_tile_a<LA, CTA, SVA, HXA, LB>(a.src + (idx_t)t * a.src_stride, a.ring + slot * (2 * (idx_t)LA * LB),
                                         a.stage_a, a.tw4_hi, a.tw4_lo, a.tw4_bits, a.swap, a.in_nt, tile, tid, lds);
    else
      fused_tile_b<LA, LB, CTB, SVB, HXB>(a.ring + slot * (2 * (idx_t)LA * LB), a.dst + (idx_t)t * a.dst_stride,
                                         a.stage_b, a.swap, a.out_nt, tile, tid, lds);
    // publish: every wave's stores reached the L2, then one add
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0)
      __hip_atomic_fetch_add(done + (2 * t + (is_a ? 0 : 1)) * kFusedLine, 1, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Clears the pop, status and done counters (a kernel, not a memset, so that a
// captured graph replays it reliably).
__global__ void fused_reset_kernel(int* ctr, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    ctr[i] = 0;
}

template <int LA, int CTA, int SVA, bool HXA, int LB, int CTB, int SVB, bool HXB>
static hipError_t fused_launch_t(const FusedArgs& a, hipStream_t st) {
  using FS = FusedShape<LA, CTA, SVA, HXA, LB, CTB, SVB, HXB>;
  auto k = fused4_kernel<LA, CTA, SVA, HXA, LB, CTB, SVB, HXB>;
  static int ncu = 0;
  if (!ncu) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    if (FS::lds > 65536) {
      e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)FS::lds);
      if (e != hipSuccess) return e;
    }
  }
  const long long nctr = (long long)fused_counter_ints(a.batch);
  hipLaunchKernelGGL(fused_reset_kernel, dim3((unsigned)std::min<long long>(64, (nctr + 255) / 256)), dim3(256), 0,
                     st, a.ctr, nctr);
  // the persistent grid: `groups` work-groups per CU, dispatched round-robin
  // over the XCDs (every XCD gets its share, and with it its list)
  hipLaunchKernelGGL(k, dim3((unsigned)(ncu * FS::groups)), dim3(FS::NT), FS::lds, st, a);
  return hipGetLastError();
}

bool fused_supported(int la, int cta, int sva, int hxa, int lb, int ctb, int svb, int hxb) {
  return la == 1024 && cta == 16 && sva == 1 && hxa == 1 && lb == 1024 && ctb == 8 && svb == 0 && hxb == 1;
}

hipError_t fused_launch(const FusedArgs& a, int la, int cta, int sva, int hxa, int lb, int ctb, int svb, int hxb,
                        hipStream_t st) {
  if (la == 1024 && cta == 16 && sva == 1 && hxa == 1 && lb == 1024 && ctb == 8 && svb == 0 && hxb == 1)
    return fused_launch_t<1024, 16, 1, true, 1024, 8, 0, true>(a, st);
  return hipErrorInvalidValue;
}

}  // namespace afgpu
