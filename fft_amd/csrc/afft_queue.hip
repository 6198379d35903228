// afft_queue.hip -- thousands of 1-D FP32 transforms of uneven power-of-two
// sizes, kinds and directions in ONE persistent launch.
//
// There is no counterpart in afft (one plan per shape, one call per
// transform); the individual transforms computed are exactly afft's c / ci /
// r / ri (fft.cpp:139-217 semantics, unnormalised, r2c giving n/2+1 bins).
//
// Host side (afft_gpu_queue_prepare): jobs are grouped into tasks:
//   - a GROUP task holds up to CT = 256/TPT(L) jobs of one (kind, length) with
//     complex length L <= 4096 (c2c n <= 4096, real n <= 8192); the tile runs
//     them side by side exactly like a row pass, each column with its own
//     src / dst (PassParams::col_ptrs);
//   - a longer job (complex length m = 8192 .. 65536) becomes a chain of tile
//     tasks over its own scratch slot: [real pre-pass tiles (c2r)] -> four-step
//     pass-A tiles (m/256-point columns, 16..64 columns each, W_m twiddle) ->
//     pass-B tiles (16 rows of 256 points, columns-fastest store) -> [real
//     post-pass tiles (r2c)].  Every tile of one step can run on a different
//     CU; a step waits for the previous step's tiles of the same job through a
//     per-job counter.
// Placement: each long job is assigned to one of the 8 XCDs (largest first,
// to the least loaded), and its tiles go to that XCD's task list; a
// work-group reads its XCD id from the hardware (HW_REG_XCC_ID) and pops its
// own list, so every tile of a job runs under one L2.  Within a list the
// chains are interleaved with a lag of kLag jobs between a step and the next
// (step s of job j is listed beside step 0 of job j + s*kLag), so a dependent
// tile is normally popped long after the tiles it waits for (each chain job
// has a scratch slot of its own, so the lag changes only the order).  The GROUP tasks form one shared list,
// longest first, that every work-group drains once its own list is empty.
// Device side: a persistent kernel; lists are popped through atomic counters.
// A task only ever waits for tasks listed (hence popped) before it in its own
// list, and a popped task is being run by a resident work-group, so by
// induction every wait ends; the wait is also bounded by a spin limit that,
// when hit, flags the queue (afft_gpu_queue_status) and lets the kernel finish.
// Coherence: producer and consumer tiles of a job share the XCD's L2 and the
// L1 is write-through, so a step publishes its tiles with every wave's stores
// completed (s_waitcnt) before one relaxed counter add, and a waiting tile
// polls the counter with atomics (performed in the L2) and then invalidates
// its CU's L1 (buffer_inv sc0); no L2-wide write-back or invalidation is
// needed.  afft_gpu_queue_launch issues only kernels (a
// reset kernel for the counters and the persistent kernel), so it can be
// captured in a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/afft_gpu.h"
#include "afft_aux.h"
#include "afft_pass.h"

namespace afgpu {

const void* queue_table(int type, idx_t n, int bits);  // afft_plan.hip (twiddle cache)
void queue_set_err(const char* msg);

enum { QT_GROUP = 0, QT_RPRE = 1, QT_BIGA = 2, QT_BIGB = 3, QT_RPOST = 4 };
constexpr int kQThreads = 256;
constexpr int kQMaxLog2 = 16;
constexpr int kBigN2 = 256;
constexpr int kRealTile = 2048;  // bin pairs per real-pass tile
constexpr int kLag = 128;        // default jobs between consecutive steps of one chain (per XCD list)
constexpr int kXcd = 8;          // XCDs of an MI355X: one chain list each
constexpr int kCtr = 16;         // ints between counters (one 64-byte line each)

struct QTask {
  int type;
  int lg;     // log2 of the complex length (group: L, chain: m)
  int kind;   // afft_gpu_c / ci / r / ri
  int count;  // group: jobs; chain: index of this step within the job
  int first;  // first (src, dst) pair in the pointer table; a chain's scratch follows
  int job;    // chain: job index (dependency counters)
  int tile;   // chain: tile index within the step
  int need;   // chain: tiles of the previous step to wait for (0: none)
};

struct QTables {
  const cx<float>* stage[kQMaxLog2 + 1];  // stage twiddles of an L-point tile, by log2 L
  const cx<float>* tw4[kQMaxLog2 + 1];    // four-step hi table (lo follows), by log2 m
  const cx<float>* real[kQMaxLog2 + 1];   // W_{2L}^k, by log2 L
};

// Written by prepare, read by every launch: a launch (or a replay of a captured
// one) runs the task list of the latest prepare.
struct QListHdr {
  int list_off[kXcd + 2];  // kXcd chain lists, then the group list, then the end
  int ndeps;               // step counters to clear (4 per chain job)
};

struct QArgs {
  const QTask* tasks;  // kXcd chain lists, then the shared group list
  const unsigned long long* ptrs;
  int* counter;        // list l pops through counter[l * kCtr]
  const QListHdr* hdr;
  int* deps;       // 4 step counters per chain job
  int* status;     // set to 1 when a dependency wait hits the spin limit
  long long spin_limit;
  const QTables* tab;  // device copy: indexed by log2 length at run time
};

template <int L>
struct QGroupCT {
  static constexpr int value = kQThreads / (L / Sched<L, 0>::E);
};

// Each task body is its own function: inlined into one loop they blow the
// register budget, as separate functions each gets its own allocation.
template <int L>
__device__ __noinline__ void q_group(const QArgs& a, const QTask& tk, int tid, cx<float>* lds) {
  constexpr int CT = QGroupCT<L>::value;
  PassParams<float> p{};
  p.in_js = 1;
  p.out_ks = 1;
  p.n_col = tk.count;
  p.col_tiles = 1;
  p.n_tiles = 1;
  p.n_blo = 1;
  p.in_il = 1;
  p.out_il = 1;
  p.col_ptrs = a.ptrs + 2 * (idx_t)tk.first;
  p.stage_tw = a.tab->stage[tk.lg];
  p.real_tw = a.tab->real[tk.lg];
  if (tk.kind == afft_gpu_c || tk.kind == afft_gpu_ci) {
    p.swap_in = p.swap_out = tk.kind == afft_gpu_ci;
    fft_tile<float, L, CT, false, EPI_PLAIN, false>(p, 0, tid, lds);
  } else if (tk.kind == afft_gpu_r) {
    fft_tile<float, L, CT, false, EPI_R2C, false>(p, 0, tid, lds);
  } else {
    p.swap_in = p.swap_out = 1;
    fft_tile<float, L, CT, false, EPI_C2R, false>(p, 0, tid, lds);
  }
}

// Real pass over one transform of h complex bins held in `z` (in place),
// forward (bins 0..h written) or inverse (reads bins 0..h of `s`, writes
// z[0..h)).  Same arithmetic as real_rows_kernel with m = 1.
__device__ void q_real_pass(const cx<float>* s, cx<float>* z, idx_t h, bool inv, const cx<float>* w_tab, int tid,
                            idx_t k0, idx_t k1) {
  for (idx_t k = k0 + tid; k < k1 && k <= h / 2; k += kQThreads) {
    if (k == 0) {
      if (inv) {
        cx<float> s0 = s[0], sh = s[h];
        z[0] = cx<float>{s0.x + sh.x, s0.x - sh.x};
      } else {
        cx<float> v = s[0];
        z[0] = cx<float>{v.x + v.y, 0.f};
        z[h] = cx<float>{v.x - v.y, 0.f};
      }
      continue;
    }
    cx<float> s0 = s[k], s1 = s[h - k];
    cx<float> w = w_tab[k];
    cx<float> s1c = conjg(s1);
    cx<float> d0, d1;
    if (inv) {
      cx<float> aa = s0 + s1c;
      cx<float> bb = cmulc(s1c - s0, w);
      d0 = aa + mul_neg_i(bb);
      d1 = conjg(aa) + mul_neg_i(conjg(bb));
    } else {
      cx<float> aa = {(s0.x + s1c.x) * 0.5f, (s0.y + s1c.y) * 0.5f};
      cx<float> bw = cmul(s0 - s1c, w);
      cx<float> bb = {bw.x * 0.5f, bw.y * 0.5f};
      d0 = aa + mul_neg_i(bb);
      d1 = conjg(aa) + mul_neg_i(conjg(bb));
    }
    z[k] = d0;
    if (h - k != k) z[h - k] = d1;
  }
}

template <int N1>
__device__ __forceinline__ void q_pass_a_tile(PassParams<float>& pa, int tile, int tid, cx<float>* lds) {
  constexpr int CT = QGroupCT<N1>::value;
  pa.col_tiles = kBigN2 / CT;
  fft_tile<float, N1, CT, true, EPI_TW4, false>(pa, tile, tid, lds);
}

// Pass-A tiles of a chain job (m/256-point columns, CT of them per tile).
int q_chain_a_tiles(int lg) {
  switch (lg - 8) {
    case 5: return kBigN2 / QGroupCT<32>::value;
    case 6: return kBigN2 / QGroupCT<64>::value;
    case 7: return kBigN2 / QGroupCT<128>::value;
    default: return kBigN2 / QGroupCT<256>::value;
  }
}

// One step tile of a chain job.  Buffers: src / dst of the job, its scratch
// (m complex); c2r runs the real pre-pass from src into dst and the c2c from
// dst; r2c runs the real post-pass in place on dst.
__device__ __noinline__ void q_chain(const QArgs& a, const QTask& tk, int tid, cx<float>* lds) {
  int* dep = a.deps + 4 * (idx_t)tk.job;
  if (tk.need > 0) {
    if (tid == 0) {
      long long it = 0;
      // polled with an atomic (performed in the L2, never served by a stale L1 line)
      while (__hip_atomic_fetch_add(&dep[tk.count - 1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < tk.need) {
        if (++it > a.spin_limit) {
          __hip_atomic_fetch_or(a.status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __syncthreads();
    // drop this CU's L1 copies: a line shared with another job's buffer may
    // have been read here before this job's producer tiles wrote it
    asm volatile("buffer_inv sc0" ::: "memory");
  }
  const idx_t m = idx_t(1) << tk.lg;
  const int n1 = (int)(m / kBigN2);
  const unsigned long long* pp = a.ptrs + 2 * (idx_t)tk.first;
  const cx<float>* src = reinterpret_cast<const cx<float>*>(pp[0]);
  cx<float>* dst = reinterpret_cast<cx<float>*>(pp[1]);
  cx<float>* scr = reinterpret_cast<cx<float>*>(pp[2]);
  const bool swap = tk.kind == afft_gpu_ci || tk.kind == afft_gpu_ri;
  if (tk.type == QT_RPRE || tk.type == QT_RPOST) {
    const bool inv = tk.type == QT_RPRE;
    q_real_pass(inv ? src : dst, dst, m, inv, a.tab->real[tk.lg], tid, (idx_t)tk.tile * kRealTile,
                (idx_t)(tk.tile + 1) * kRealTile);
  } else if (tk.type == QT_BIGA) {
    // pass A: n1-point columns (stride 256) with the W_m twiddle, into scratch
    PassParams<float> pa{};
    pa.in_re = reinterpret_cast<const float*>(tk.kind == afft_gpu_ri ? dst : src);
    pa.out_re = reinterpret_cast<float*>(scr);
    pa.in_js = kBigN2;
    pa.out_ks = kBigN2;
    pa.in_cs = 1;
    pa.out_cs = 1;
    pa.n_col = kBigN2;
    pa.n_blo = 1;
    pa.in_il = pa.out_il = 1;
    pa.swap_in = swap;
    pa.tw4_hi = a.tab->tw4[tk.lg];
    pa.tw4_lo = a.tab->tw4[tk.lg] + (m >> ((tk.lg + 1) / 2));
    pa.tw4_lo_bits = (tk.lg + 1) / 2;
    pa.tw4_log2n = tk.lg;  // the twiddle is computed, not read from the tables
    pa.tw4_col_shift = 0;
    pa.stage_tw = a.tab->stage[tk.lg - 8];
    switch (n1) {
      case 32: q_pass_a_tile<32>(pa, tk.tile, tid, lds); break;
      case 64: q_pass_a_tile<64>(pa, tk.tile, tid, lds); break;
      case 128: q_pass_a_tile<128>(pa, tk.tile, tid, lds); break;
      default: q_pass_a_tile<256>(pa, tk.tile, tid, lds); break;
    }
  } else {
    // pass B: 16 rows of 256 points of the scratch, bins k1 + n1*k2 into dst
    PassParams<float> pb{};
    pb.in_re = reinterpret_cast<const float*>(scr);
    pb.out_re = reinterpret_cast<float*>(dst);
    pb.in_js = 1;
    pb.in_cs = kBigN2;
    pb.out_ks = n1;
    pb.out_cs = 1;
    pb.n_col = n1;
    pb.n_blo = 1;
    pb.in_il = pb.out_il = 1;
    pb.swap_out = swap;
    pb.stage_tw = a.tab->stage[8];
    pb.col_tiles = n1 / 16;
    fft_tile<float, kBigN2, 16, false, EPI_PLAIN, true>(pb, tk.tile, tid, lds);
  }
  // publish: every wave's stores have reached the L2 (s_waitcnt 0), then one
  // add announces the tile
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0) __hip_atomic_fetch_add(&dep[tk.count], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int xcc_id() {
  unsigned v;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
  return (int)(v & (kXcd - 1));
}

__device__ __noinline__ void q_group_any(const QArgs& a, const QTask& tk, int tid, cx<float>* lds) {
  switch (tk.lg) {
    case 0: q_group<1>(a, tk, tid, lds); break;
    case 1: q_group<2>(a, tk, tid, lds); break;
    case 2: q_group<4>(a, tk, tid, lds); break;
    case 3: q_group<8>(a, tk, tid, lds); break;
    case 4: q_group<16>(a, tk, tid, lds); break;
    case 5: q_group<32>(a, tk, tid, lds); break;
    case 6: q_group<64>(a, tk, tid, lds); break;
    case 7: q_group<128>(a, tk, tid, lds); break;
    case 8: q_group<256>(a, tk, tid, lds); break;
    case 9: q_group<512>(a, tk, tid, lds); break;
    case 10: q_group<1024>(a, tk, tid, lds); break;
    case 11: q_group<2048>(a, tk, tid, lds); break;
    default: q_group<4096>(a, tk, tid, lds); break;
  }
}

// GROUPS = work-groups per CU the registers are bounded for (2: 256 VGPRs per
// lane, 4: 128; the queue picks one, see afft_gpu_queue_set_groups_per_cu).
template <int GROUPS>
__global__ void __launch_bounds__(kQThreads, GROUPS) queue_kernel(QArgs a) {
  extern __shared__ __align__(16) unsigned char q_smem[];
  cx<float>* lds = reinterpret_cast<cx<float>*>(q_smem);
  __shared__ int s_task;
  const int tid = threadIdx.x;
  // this XCD's chain list first, then the shared group list
  const int lists[2] = {xcc_id(), kXcd};
  for (int li = 0; li < 2; li++) {
    const int l = lists[li];
    const int off = a.hdr->list_off[l];
    const int n = a.hdr->list_off[l + 1] - off;
    for (;;) {
      if (tid == 0) s_task = atomicAdd(a.counter + l * kCtr, 1);
      __syncthreads();
      const int ti = s_task;
      __syncthreads();
      if (ti >= n) break;
      const QTask tk = a.tasks[off + ti];
      if (tk.type != QT_GROUP)
        q_chain(a, tk, tid, lds);
      else
        q_group_any(a, tk, tid, lds);
      __syncthreads();
    }
  }
}

// Clears the task counter, the status flag and the chain step counters (a
// kernel, not a memset, so a captured graph replays it reliably).
__global__ void queue_reset_kernel(int* counter, int* status, int* deps, const QListHdr* hdr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= kXcd) counter[i * kCtr] = 0;
  if (i == 0) *status = 0;
  const int ndeps = hdr->ndeps;
  for (int k = i; k < ndeps; k += gridDim.x * blockDim.x) deps[k] = 0;
}

template <int L>
constexpr size_t q_lds(int ct) {
  return (size_t)ct * LdsLen<L>::value * sizeof(cx<float>);
}
constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }
constexpr size_t kQLds = cmax(
    cmax(cmax(cmax(q_lds<16>(QGroupCT<16>::value), q_lds<4096>(QGroupCT<4096>::value)),
              cmax(q_lds<128>(QGroupCT<128>::value), q_lds<256>(QGroupCT<256>::value))),
         cmax(cmax(q_lds<2>(QGroupCT<2>::value), q_lds<32>(QGroupCT<32>::value)),
              cmax(q_lds<64>(QGroupCT<64>::value), q_lds<2048>(QGroupCT<2048>::value)))),
    // pass-B tiles of the chains: transposed store, padded rows
    tile_lds_bytes<float, kBigN2, 16, false, true, 0, false>());

int q_group_ct(int lg) {
  switch (lg) {
    case 0: return QGroupCT<1>::value;
    case 1: return QGroupCT<2>::value;
    case 2: return QGroupCT<4>::value;
    case 3: return QGroupCT<8>::value;
    case 4: return QGroupCT<16>::value;
    case 5: return QGroupCT<32>::value;
    case 6: return QGroupCT<64>::value;
    case 7: return QGroupCT<128>::value;
    case 8: return QGroupCT<256>::value;
    case 9: return QGroupCT<512>::value;
    case 10: return QGroupCT<1024>::value;
    case 11: return QGroupCT<2048>::value;
    default: return QGroupCT<4096>::value;
  }
}

}  // namespace afgpu

using namespace afgpu;

namespace {

constexpr uint32_t kQMagic = 0xAFF7A0EEu;

struct Queue {
  size_t max_jobs = 0;
  int max_lg = 0;
  int grid = 0;
  int ntasks = 0;
  int nchain = 0;                   // jobs run as tile chains
  int groups_per_cu = 4;            // persistent work-groups per CU (2 or 4; 4 measured faster)
  int lag = kLag;                   // jobs between consecutive steps of one chain in a list
  int ncu = 256;
  long long spin_limit = 1LL << 26;  // dependency-wait iterations before the queue is flagged
  QTask* d_tasks = nullptr;
  size_t cap_tasks = 0;
  unsigned long long* d_ptrs = nullptr;
  size_t cap_ptrs = 0;
  int* d_counter = nullptr;  // list counters at [l * kCtr], l = 0..kXcd; status at [(kXcd + 1) * kCtr]
  QListHdr hdr{};
  QListHdr* d_hdr = nullptr;
  int* d_deps = nullptr;
  size_t cap_deps = 0;
  hipEvent_t done = nullptr;  // recorded after each eager launch; prepare waits for it
  bool launched = false;
  cx<float>* d_scratch = nullptr;
  size_t cap_scratch = 0;  // complex elements
  QTables tab{};
  QTables* d_tab = nullptr;
  std::vector<QTask> h_tasks;
  std::vector<unsigned long long> h_ptrs;
};

// Grows a device array to at least `need` elements (contents not kept).
template <typename X>
hipError_t q_reserve(X*& p, size_t& cap, size_t need) {
  if (need <= cap) return hipSuccess;
  (void)hipFree(p);
  p = nullptr;
  cap = 0;
  hipError_t e = hipMalloc(&p, sizeof(X) * need);
  if (e == hipSuccess) cap = need;
  return e;
}

struct QHeader {
  uint32_t magic;
  uint32_t pad;
  Queue* q;
};

Queue* get_queue(const void* mem) {
  if (!mem) return nullptr;
  QHeader h;
  std::memcpy(&h, mem, sizeof(h));
  return h.magic == kQMagic ? h.q : nullptr;
}

void free_queue(Queue* q) {
  if (!q) return;
  if (q->done) {
    (void)hipEventSynchronize(q->done);
    (void)hipEventDestroy(q->done);
  }
  (void)hipFree(q->d_hdr);
  (void)hipFree(q->d_tasks);
  (void)hipFree(q->d_ptrs);
  (void)hipFree(q->d_counter);
  (void)hipFree(q->d_deps);
  (void)hipFree(q->d_scratch);
  (void)hipFree(q->d_tab);
  delete q;
}

}  // namespace

extern "C" {

size_t afft_gpu_queue_memsize(void) { return sizeof(QHeader); }

afft_gpu_queue* afft_gpu_queue_create(size_t max_jobs, int max_log2n, void* mem) {
  if (!mem || max_jobs == 0 || max_log2n < 0 || max_log2n > kQMaxLog2) {
    queue_set_err("queue: invalid arguments (max_log2n must be 0..16)");
    return nullptr;
  }
  Queue* q = new (std::nothrow) Queue();
  if (!q) return nullptr;
  q->max_jobs = max_jobs;
  q->max_lg = max_log2n;
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ncu = prop.multiProcessorCount;
  }
  // persistent grid: groups_per_cu 256-thread groups per CU
  q->ncu = ncu;
  q->grid = ncu * q->groups_per_cu;
  // The task, pointer and dependency arrays are sized for max_jobs here, so
  // their addresses never change (a captured launch keeps reading them); only
  // the chain jobs' scratch grows in prepare.
  int max_chain_tasks = 1;
  for (int lgm = 13; lgm <= max_log2n; lgm++) {
    const int rt = (int)((((size_t)1 << lgm) / 2 + 1 + kRealTile - 1) / kRealTile);
    max_chain_tasks = std::max(max_chain_tasks, rt + q_chain_a_tiles(lgm) + (1 << (lgm - 8)) / 16);
  }
  hipError_t e = hipMalloc(&q->d_counter, sizeof(int) * kCtr * (kXcd + 2));
  if (e == hipSuccess) e = hipMalloc(&q->d_hdr, sizeof(QListHdr));
  if (e == hipSuccess) e = hipMemset(q->d_hdr, 0, sizeof(QListHdr));
  if (e == hipSuccess) e = q_reserve(q->d_tasks, q->cap_tasks, max_jobs * (size_t)max_chain_tasks);
  if (e == hipSuccess) e = q_reserve(q->d_ptrs, q->cap_ptrs, 4 * max_jobs);
  if (e == hipSuccess) e = q_reserve(q->d_deps, q->cap_deps, 4 * max_jobs);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&q->done, hipEventDisableTiming);
  if (e != hipSuccess) {
    queue_set_err(hipGetErrorString(e));
    free_queue(q);
    return nullptr;
  }
  // twiddle tables for every length a job may need
  for (int lg = 0; lg <= max_log2n && lg <= 12; lg++) {
    q->tab.stage[lg] = (const cx<float>*)queue_table(0, idx_t(1) << lg, 0);
    q->tab.real[lg] = (const cx<float>*)queue_table(2, idx_t(1) << lg, 0);
    if (!q->tab.stage[lg] || !q->tab.real[lg]) {
      free_queue(q);
      return nullptr;
    }
  }
  for (int lg = 13; lg <= max_log2n; lg++) {
    int bits = (lg + 1) / 2;
    q->tab.tw4[lg] = (const cx<float>*)queue_table(1, idx_t(1) << lg, bits);
    q->tab.real[lg] = (const cx<float>*)queue_table(2, idx_t(1) << lg, 0);
    if (!q->tab.tw4[lg] || !q->tab.real[lg]) {
      free_queue(q);
      return nullptr;
    }
  }
  e = hipMalloc(&q->d_tab, sizeof(QTables));
  if (e == hipSuccess) e = hipMemcpy(q->d_tab, &q->tab, sizeof(QTables), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    queue_set_err(hipGetErrorString(e));
    free_queue(q);
    return nullptr;
  }
  QHeader h{kQMagic, 0, q};
  std::memcpy(mem, &h, sizeof(h));
  return (afft_gpu_queue*)mem;
}

int afft_gpu_queue_prepare(afft_gpu_queue* qh, const afft_gpu_job* jobs, size_t njobs) {
  Queue* q = get_queue(qh);
  if (!q) {
    queue_set_err("queue: not a queue");
    return -1;
  }
  if (njobs > q->max_jobs || (njobs && !jobs)) {
    queue_set_err("queue: too many jobs for this queue");
    return -1;
  }
  // the previous eager launch may still read the arrays rewritten below
  if (q->launched && hipEventSynchronize(q->done) != hipSuccess) {
    queue_set_err("queue: waiting for the previous launch failed");
    return -1;
  }
  // bucket by (kind, complex length)
  std::vector<std::vector<size_t>> small((size_t)4 * 13);
  std::vector<size_t> big;
  for (size_t i = 0; i < njobs; i++) {
    const afft_gpu_job& j = jobs[i];
    const bool real = j.kind == afft_gpu_r || j.kind == afft_gpu_ri;
    if (j.kind < 0 || j.kind > 3 || j.log2n < (real ? 1 : 0) || j.log2n > q->max_lg || !j.src || !j.dst) {
      queue_set_err("queue: invalid job (kind, log2n or null buffer)");
      return -1;
    }
    const int lgm = real ? j.log2n - 1 : j.log2n;
    if (lgm <= 12)
      small[(size_t)j.kind * 13 + lgm].push_back(i);
    else
      big.push_back(i);
  }
  q->h_tasks.clear();
  q->h_ptrs.clear();
  std::stable_sort(big.begin(), big.end(), [&](size_t x, size_t y) {
    auto lgm = [&](size_t i) { return jobs[i].log2n - (jobs[i].kind >= afft_gpu_r ? 1 : 0); };
    return lgm(x) > lgm(y);
  });
  // chains: pointer pairs (src, dst) and (scratch, 0) per job, steps per job
  struct Step {
    int type, tiles;
  };
  std::vector<std::vector<Step>> steps(big.size());
  size_t scratch_elems = 0;
  std::vector<size_t> scratch_off(big.size());
  for (size_t c = 0; c < big.size(); c++) {
    const afft_gpu_job& j = jobs[big[c]];
    const int lgm = j.log2n - (j.kind >= afft_gpu_r ? 1 : 0);
    const int rt = (int)((((size_t)1 << lgm) / 2 + 1 + kRealTile - 1) / kRealTile);
    if (j.kind == afft_gpu_ri) steps[c].push_back({QT_RPRE, rt});
    steps[c].push_back({QT_BIGA, q_chain_a_tiles(lgm)});
    steps[c].push_back({QT_BIGB, (1 << (lgm - 8)) / 16});
    if (j.kind == afft_gpu_r) steps[c].push_back({QT_RPOST, rt});
    scratch_off[c] = scratch_elems;
    scratch_elems += (size_t)1 << lgm;
  }
  hipError_t e0 = q_reserve(q->d_scratch, q->cap_scratch, std::max<size_t>(scratch_elems, 1));
  if (e0 != hipSuccess) {
    queue_set_err(hipGetErrorString(e0));
    return -1;
  }
  std::vector<int> first(big.size());
  for (size_t c = 0; c < big.size(); c++) {
    const afft_gpu_job& j = jobs[big[c]];
    first[c] = (int)(q->h_ptrs.size() / 2);
    q->h_ptrs.push_back(j.src);
    q->h_ptrs.push_back(j.dst);
    q->h_ptrs.push_back((unsigned long long)(q->d_scratch + scratch_off[c]));
    q->h_ptrs.push_back(0);
  }
  // jobs to XCDs: largest first (big is sorted), each to the least loaded
  const size_t nb = big.size();
  std::vector<std::vector<size_t>> on_xcd(kXcd);
  {
    std::vector<double> load(kXcd, 0.0);
    for (size_t c = 0; c < nb; c++) {
      const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      on_xcd[x].push_back(c);
      load[x] += (double)((size_t)1 << (jobs[big[c]].log2n));
    }
  }
  // per XCD list: step s of its k-th job is listed at round k + s*lag (any
  // lag >= 0 lists a step after the one it waits for)
  const size_t lag = (size_t)q->lag;
  for (int x = 0; x < kXcd; x++) {
    q->hdr.list_off[x] = (int)q->h_tasks.size();
    const auto& js = on_xcd[x];
    for (size_t round = 0; !js.empty() && round < js.size() + 3 * lag; round++)
      for (int st = 0; st < 4; st++) {
        if (round < (size_t)st * lag) continue;
        const size_t k = round - (size_t)st * lag;
        if (k >= js.size()) continue;
        const size_t c = js[k];
        if (st >= (int)steps[c].size()) continue;
        const afft_gpu_job& j = jobs[big[c]];
        for (int t = 0; t < steps[c][st].tiles; t++) {
          QTask tk{};
          tk.type = steps[c][st].type;
          tk.kind = j.kind;
          tk.lg = j.log2n - (j.kind >= afft_gpu_r ? 1 : 0);
          tk.count = st;
          tk.first = first[c];
          tk.job = (int)c;
          tk.tile = t;
          tk.need = st > 0 ? steps[c][st - 1].tiles : 0;
          q->h_tasks.push_back(tk);
        }
      }
  }
  q->hdr.list_off[kXcd] = (int)q->h_tasks.size();
  q->nchain = (int)nb;
  for (int lgm = 12; lgm >= 0; lgm--)
    for (int kind = 0; kind < 4; kind++) {
      const auto& v = small[(size_t)kind * 13 + lgm];
      const int ct = q_group_ct(lgm);
      for (size_t s = 0; s < v.size(); s += (size_t)ct) {
        QTask t{};
        t.type = QT_GROUP;
        t.kind = kind;
        t.lg = lgm;
        t.count = (int)std::min((size_t)ct, v.size() - s);
        t.first = (int)(q->h_ptrs.size() / 2);
        for (int c = 0; c < t.count; c++) {
          q->h_ptrs.push_back(jobs[v[s + c]].src);
          q->h_ptrs.push_back(jobs[v[s + c]].dst);
        }
        q->h_tasks.push_back(t);
      }
    }
  q->ntasks = (int)q->h_tasks.size();
  q->hdr.list_off[kXcd + 1] = q->ntasks;
  q->hdr.ndeps = 4 * q->nchain;
  if (q->h_tasks.size() > q->cap_tasks || q->h_ptrs.size() > q->cap_ptrs || (size_t)q->hdr.ndeps > q->cap_deps) {
    queue_set_err("queue: internal: task list exceeds the arrays sized at create");
    return -1;
  }
  hipError_t e = hipMemcpy(q->d_hdr, &q->hdr, sizeof(QListHdr), hipMemcpyHostToDevice);
  if (e == hipSuccess && q->ntasks) {
    e = hipMemcpy(q->d_tasks, q->h_tasks.data(), sizeof(QTask) * q->h_tasks.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(q->d_ptrs, q->h_ptrs.data(), sizeof(unsigned long long) * q->h_ptrs.size(),
                    hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    queue_set_err(hipGetErrorString(e));
    return -1;
  }
  return q->ntasks;
}

int afft_gpu_queue_launch(afft_gpu_queue* qh, void* stream) {
  Queue* q = get_queue(qh);
  if (!q) {
    queue_set_err("queue: not a queue");
    return -1;
  }
  if (q->ntasks == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  static_assert(kQLds * 4 <= 160 * 1024, "four groups per CU must fit the LDS");
  const size_t ndeps_max = q->cap_deps;
  hipLaunchKernelGGL(queue_reset_kernel, dim3((unsigned)std::max<size_t>(1, std::min<size_t>(64, (ndeps_max + 255) / 256))),
                     dim3(256), 0, st, q->d_counter, q->d_counter + (kXcd + 1) * kCtr, q->d_deps, q->d_hdr);
  QArgs a{};
  a.tasks = q->d_tasks;
  a.ptrs = q->d_ptrs;
  a.counter = q->d_counter;
  a.hdr = q->d_hdr;
  a.deps = q->d_deps;
  a.status = q->d_counter + (kXcd + 1) * kCtr;
  a.spin_limit = q->spin_limit;
  a.tab = q->d_tab;
  // the full grid even for few tasks: every XCD needs work-groups for its list
  const int grid = q->grid;
  if (q->groups_per_cu == 4)
    hipLaunchKernelGGL(queue_kernel<4>, dim3(grid), dim3(kQThreads), kQLds, st, a);
  else
    hipLaunchKernelGGL(queue_kernel<2>, dim3(grid), dim3(kQThreads), kQLds, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    queue_set_err(hipGetErrorString(e));
    return -1;
  }
  // an eager launch is remembered so that the next prepare waits for it; a
  // captured one is not (the caller finishes its replays before a prepare)
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
    if (hipEventRecord(q->done, st) != hipSuccess) {
      queue_set_err("queue: event record failed");
      return -1;
    }
    q->launched = true;
  }
  return 0;
}

int afft_gpu_queue_run(afft_gpu_queue* qh, const afft_gpu_job* jobs, size_t njobs, void* stream) {
  if (afft_gpu_queue_prepare(qh, jobs, njobs) < 0) return -1;
  return afft_gpu_queue_launch(qh, stream);
}

int afft_gpu_queue_status(const afft_gpu_queue* qh) {
  Queue* q = get_queue(qh);
  if (!q) {
    queue_set_err("queue: not a queue");
    return -1;
  }
  int v = 0;
  if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(&v, q->d_counter + (kXcd + 1) * kCtr, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
    queue_set_err("queue: status read failed");
    return -1;
  }
  return v;
}

int afft_gpu_queue_set_spin_limit(afft_gpu_queue* qh, long long limit) {
  Queue* q = get_queue(qh);
  if (!q || limit < 0) {
    queue_set_err("queue: not a queue or negative limit");
    return -1;
  }
  q->spin_limit = limit;
  return 0;
}

int afft_gpu_queue_set_groups_per_cu(afft_gpu_queue* qh, int groups) {
  Queue* q = get_queue(qh);
  if (!q || (groups != 2 && groups != 4)) {
    queue_set_err("queue: not a queue or groups per CU not 2 or 4");
    return -1;
  }
  q->groups_per_cu = groups;
  q->grid = q->ncu * groups;
  return 0;
}

int afft_gpu_queue_set_lag(afft_gpu_queue* qh, int lag) {
  Queue* q = get_queue(qh);
  if (!q || lag < 0 || lag > 1024) {
    queue_set_err("queue: not a queue or lag outside 0..1024");
    return -1;
  }
  q->lag = lag;
  return 0;
}

int afft_gpu_queue_get_config(const afft_gpu_queue* qh, int* groups_per_cu, int* lag) {
  Queue* q = get_queue(qh);
  if (!q) {
    queue_set_err("queue: not a queue");
    return -1;
  }
  if (groups_per_cu) *groups_per_cu = q->groups_per_cu;
  if (lag) *lag = q->lag;
  return 0;
}

// Times every (groups per CU, lag) candidate on this batch -- prepare, one
// warm launch, `iters` timed launches -- and keeps the fastest whose launches
// met every dependency; the queue is left prepared with it.
int afft_gpu_queue_tune(afft_gpu_queue* qh, const afft_gpu_job* jobs, size_t njobs, void* stream, int iters,
                        float* best_ms) {
  Queue* q = get_queue(qh);
  if (!q || iters < 1) {
    queue_set_err("queue: not a queue or iters < 1");
    return -1;
  }
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    queue_set_err("queue: event create failed");
    return -1;
  }
  const int keep_groups = q->groups_per_cu, keep_lag = q->lag;
  float best = 1e30f;
  int bg = keep_groups, bl = keep_lag;
  bool failed = false;
  for (int groups : {4, 2}) {
    for (int lag : {kLag, 32, 8}) {
      if (afft_gpu_queue_set_groups_per_cu(qh, groups) != 0 || afft_gpu_queue_set_lag(qh, lag) != 0 ||
          afft_gpu_queue_prepare(qh, jobs, njobs) < 0 || afft_gpu_queue_launch(qh, stream) != 0) {
        failed = true;
        break;
      }
      float ms = 0.f;
      if (hipEventRecord(e0, st) != hipSuccess) failed = true;
      for (int i = 0; i < iters && !failed; i++)
        if (afft_gpu_queue_launch(qh, stream) != 0) failed = true;
      if (failed || hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        failed = true;
        break;
      }
      if (afft_gpu_queue_status(qh) != 0) continue;  // a dependency wait ran out: not a candidate
      ms /= (float)iters;
      if (ms < best) {
        best = ms;
        bg = groups;
        bl = lag;
      }
    }
    if (failed) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  q->groups_per_cu = bg;
  q->grid = q->ncu * bg;
  q->lag = bl;
  if (failed) {
    afft_gpu_queue_set_groups_per_cu(qh, keep_groups);
    q->lag = keep_lag;
    return -1;
  }
  if (afft_gpu_queue_prepare(qh, jobs, njobs) < 0) return -1;
  if (best_ms) *best_ms = best;
  return 0;
}

int afft_gpu_queue_num_tasks(const afft_gpu_queue* qh) {
  Queue* q = get_queue(qh);
  return q ? q->ntasks : -1;
}

void afft_gpu_queue_destroy(afft_gpu_queue* qh) {
  Queue* q = get_queue(qh);
  if (!q) return;
  free_queue(q);
  std::memset((void*)qh, 0, sizeof(QHeader));
}

}  // extern "C"
