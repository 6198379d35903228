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
//   - a BIG task is one longer job (complex length 8192 .. 65536) that a single
//     work-group runs as a four-step n1 x 256 transform through a private
//     scratch slot, with the real pre/post pass as extra sweeps.
// Tasks are ordered longest first.  The device side is a persistent kernel:
// every work-group pops the next task index from one atomic counter and runs
// it until the list is empty.  No task waits on another, so the launch needs
// no co-residency and cannot deadlock.  afft_gpu_queue_launch issues only
// kernels (a counter-reset kernel and the persistent kernel), so it can be
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

enum { QT_GROUP = 0, QT_BIG = 1 };
constexpr int kQThreads = 256;
constexpr int kQMaxLog2 = 16;
constexpr int kBigN2 = 256;

struct QTask {
  int type;
  int lg;     // log2 of the complex length (group: L, big: m)
  int kind;   // afft_gpu_c / ci / r / ri
  int count;  // jobs in a group
  int first;  // first (src, dst) pair in the pointer table
  int pad[3];
};

struct QTables {
  const cx<float>* stage[kQMaxLog2 + 1];  // stage twiddles of an L-point tile, by log2 L
  const cx<float>* tw4[kQMaxLog2 + 1];    // four-step hi table (lo follows), by log2 m
  const cx<float>* real[kQMaxLog2 + 1];   // W_{2L}^k, by log2 L
};

struct QArgs {
  const QTask* tasks;
  const unsigned long long* ptrs;
  int* counter;
  int ntasks;
  cx<float>* scratch;
  idx_t scratch_per_wg;
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
__device__ void q_real_pass(const cx<float>* s, cx<float>* z, idx_t h, bool inv, const cx<float>* w_tab, int tid) {
  for (idx_t k = tid; k <= h / 2; k += kQThreads) {
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
__device__ __forceinline__ void q_big_pass_a(PassParams<float>& pa, int tid, cx<float>* lds) {
  constexpr int CT = QGroupCT<N1>::value;
  pa.col_tiles = kBigN2 / CT;
  for (int tile = 0; tile < pa.col_tiles; tile++) {
    fft_tile<float, N1, CT, true, EPI_TW4, false>(pa, tile, tid, lds);
    __syncthreads();
  }
}

__device__ __noinline__ void q_big(const QArgs& a, const QTask& tk, int tid, cx<float>* lds) {
  const idx_t m = idx_t(1) << tk.lg;  // complex length of the c2c part
  const int n1 = (int)(m / kBigN2);
  const unsigned long long* pp = a.ptrs + 2 * (idx_t)tk.first;
  const cx<float>* src = reinterpret_cast<const cx<float>*>(pp[0]);
  cx<float>* dst = reinterpret_cast<cx<float>*>(pp[1]);
  cx<float>* scr = a.scratch + (idx_t)blockIdx.x * a.scratch_per_wg;
  const bool swap = tk.kind == afft_gpu_ci || tk.kind == afft_gpu_ri;
  const cx<float>* a_in = src;
  if (tk.kind == afft_gpu_ri) {
    q_real_pass(src, dst, m, true, a.tab->real[tk.lg], tid);
    __syncthreads();
    a_in = dst;
  }
  // pass A: n1-point columns (stride 256) with the W_m twiddle, into scratch
  PassParams<float> pa{};
  pa.in_re = reinterpret_cast<const float*>(a_in);
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
  pa.tw4_col_shift = 0;
  int ln1 = tk.lg - 8;
  pa.stage_tw = a.tab->stage[ln1];
  switch (n1) {
    case 32: q_big_pass_a<32>(pa, tid, lds); break;
    case 64: q_big_pass_a<64>(pa, tid, lds); break;
    case 128: q_big_pass_a<128>(pa, tid, lds); break;
    default: q_big_pass_a<256>(pa, tid, lds); break;
  }
  // pass B: 256-point rows of the scratch, bins k1 + n1*k2 into dst
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
  constexpr int CTB = 16;
  pb.col_tiles = n1 / CTB;
  for (int tile = 0; tile < pb.col_tiles; tile++) {
    fft_tile<float, kBigN2, CTB, false, EPI_PLAIN, true>(pb, tile, tid, lds);
    __syncthreads();
  }
  if (tk.kind == afft_gpu_r) {
    q_real_pass(dst, dst, m, false, a.tab->real[tk.lg], tid);
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kQThreads, 2) queue_kernel(QArgs a) {
  extern __shared__ __align__(16) unsigned char q_smem[];
  cx<float>* lds = reinterpret_cast<cx<float>*>(q_smem);
  __shared__ int s_task;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) s_task = atomicAdd(a.counter, 1);
    __syncthreads();
    const int ti = s_task;
    __syncthreads();
    if (ti >= a.ntasks) return;
    const QTask tk = a.tasks[ti];
    if (tk.type == QT_BIG) {
      q_big(a, tk, tid, lds);
    } else {
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
    __syncthreads();
  }
}

__global__ void queue_reset_kernel(int* counter) {
  if (threadIdx.x == 0) *counter = 0;
}

template <int L>
constexpr size_t q_lds(int ct) {
  return (size_t)ct * LdsLen<L>::value * sizeof(cx<float>);
}
constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }
constexpr size_t kQLds = cmax(
    cmax(cmax(q_lds<16>(QGroupCT<16>::value), q_lds<4096>(QGroupCT<4096>::value)),
         cmax(q_lds<128>(QGroupCT<128>::value), q_lds<256>(QGroupCT<256>::value))),
    cmax(cmax(q_lds<2>(QGroupCT<2>::value), q_lds<32>(QGroupCT<32>::value)),
         cmax(q_lds<64>(QGroupCT<64>::value), q_lds<2048>(QGroupCT<2048>::value))));

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
  QTask* d_tasks = nullptr;
  unsigned long long* d_ptrs = nullptr;
  int* d_counter = nullptr;
  cx<float>* d_scratch = nullptr;
  idx_t scratch_per_wg = 0;
  QTables tab{};
  QTables* d_tab = nullptr;
  std::vector<QTask> h_tasks;
  std::vector<unsigned long long> h_ptrs;
};

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
  (void)hipFree(q->d_tasks);
  (void)hipFree(q->d_ptrs);
  (void)hipFree(q->d_counter);
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
  // persistent grid: two 256-thread groups per CU (the kernel is compiled for
  // two waves per SIMD: 256 VGPRs, no spills)
  int per_cu = (int)(160 * 1024 / kQLds);
  if (per_cu > 2) per_cu = 2;
  if (per_cu < 1) per_cu = 1;
  q->grid = ncu * per_cu;
  // scratch: one slot per group, sized for the largest complex length in use
  idx_t maxm = max_log2n >= 1 ? (idx_t(1) << max_log2n) : 1;
  q->scratch_per_wg = maxm >= 8192 ? maxm : 0;
  hipError_t e = hipMalloc(&q->d_tasks, sizeof(QTask) * max_jobs);
  if (e == hipSuccess) e = hipMalloc(&q->d_ptrs, sizeof(unsigned long long) * 2 * max_jobs);
  if (e == hipSuccess) e = hipMalloc(&q->d_counter, 64);
  if (e == hipSuccess && q->scratch_per_wg)
    e = hipMalloc(&q->d_scratch, sizeof(cx<float>) * (size_t)q->scratch_per_wg * (size_t)q->grid);
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
  for (size_t i : big) {
    const afft_gpu_job& j = jobs[i];
    QTask t{};
    t.type = QT_BIG;
    t.kind = j.kind;
    t.lg = j.log2n - (j.kind >= afft_gpu_r ? 1 : 0);
    t.count = 1;
    t.first = (int)(q->h_ptrs.size() / 2);
    q->h_ptrs.push_back(j.src);
    q->h_ptrs.push_back(j.dst);
    q->h_tasks.push_back(t);
  }
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
  hipError_t e = hipSuccess;
  if (q->ntasks) {
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
  static bool attr = false;
  if (!attr && kQLds > 65536) {
    (void)hipFuncSetAttribute((const void*)queue_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kQLds);
    attr = true;
  }
  hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, st, q->d_counter);
  QArgs a{};
  a.tasks = q->d_tasks;
  a.ptrs = q->d_ptrs;
  a.counter = q->d_counter;
  a.ntasks = q->ntasks;
  a.scratch = q->d_scratch;
  a.scratch_per_wg = q->scratch_per_wg;
  a.tab = q->d_tab;
  int grid = std::min(q->grid, q->ntasks);
  hipLaunchKernelGGL(queue_kernel, dim3(grid), dim3(kQThreads), kQLds, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    queue_set_err(hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int afft_gpu_queue_run(afft_gpu_queue* qh, const afft_gpu_job* jobs, size_t njobs, void* stream) {
  if (afft_gpu_queue_prepare(qh, jobs, njobs) < 0) return -1;
  return afft_gpu_queue_launch(qh, stream);
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
