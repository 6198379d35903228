// afft_plan.hip -- plan construction, execution and the C ABI of fft.h /
// afft_gpu.h.
//
// A plan is a short list of pass launches over three kinds of buffer: the
// caller's src, the caller's dst and device scratch owned by the plan.  How
// each afft operation decomposes (reference structure in brackets):
//   c2c, 1-D, N <= 16384       one row pass          [onedim small_transform]
//   c2c, 1-D, N  > 16384       four-step N = N1*N2: column pass of N1 with the
//                              W_N twiddle fused (src -> scratch), then a row
//                              pass of N2 storing columns-fastest (scratch ->
//                              dst)                  [onedim large_transform]
//   c2c, N-D                   last axis as above, every outer axis as one
//                              strided column pass in place on dst
//                              [fft_internal.hpp:121-152 + multi.hpp]
//   r2c / c2r, 1-D             N/2-point c2c with the real pass fused into
//                              the last / first stage (separate real-rows
//                              kernel when N/2 > 16384) [onedim.hpp:1110-1254]
//   r2c / c2r, N-D             dim 0 halved: fused column r2c/c2r along dim 0
//                              plus the inner c2c on each of the N0/2+1 rows
//                              [fft_internal.hpp:200-493]
// Inverse transforms swap re/im on load and store of the forward passes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/afft_gpu.h"
#include "../../include/fft.h"
#include "afft_aux.h"
#include "afft_fused.h"
#include "afft_mfma.h"
#include "afft_real4.h"
#include "afft_registry.h"
#include "afft_tiny.h"

using namespace afgpu;

namespace {

thread_local std::string g_err;

void set_err(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

#define HIP_OK(expr)                                                     \
  do {                                                                   \
    hipError_t e_ = (expr);                                              \
    if (e_ != hipSuccess) {                                              \
      set_err("%s failed: %s", #expr, hipGetErrorString(e_));            \
      return false;                                                      \
    }                                                                    \
  } while (0)

constexpr uint32_t kMagic = 0xAFF7C0DEu;
constexpr int kMaxDim = 64;
constexpr int kMaxPass = 160;

enum Role { R_SRC = 0, R_DST = 1, R_SCR = 2, R_SCR2 = 3 };
enum PassType {
  PT_FFT = 0,
  PT_REALROWS = 1,
  PT_REALONE = 2,
  PT_REALTR = 3,
  PT_REALPAIR = 4,
  PT_REALROWFFT = 5,
  PT_REALIROWFFT = 6,
  PT_REALCOLPAIR = 7,
  PT_REALICOLPAIR = 8,
  PT_REAL4A = 9,
  PT_REAL4B = 10,
  PT_REAL4C = 11,
  PT_REAL4D = 12,
  PT_TINY = 13
};

struct Side {
  int role = R_SRC;
  int il = 1;         // interleaved (complex-element units) vs split (scalar units)
  int one_array = 0;  // split addressing with re and im both in the re array
  idx_t re_off = 0, im_off = 0;
  idx_t s_j = 0, s_c = 0, s_b = 0, s_b2 = 0;
  idx_t s_u = 0;      // offset between consecutive user transforms (src/dst only)
};

struct PassRec {
  int type = PT_FFT;
  int L = 1, CT = 1, colmap = 0, epi = EPI_PLAIN, store_t = 0;
  const void* variant = nullptr;
  Side in, out;
  idx_t n_col = 1;
  int fold_user = 0;     // user transforms folded into the column index
  idx_t b_per_user = 1;  // batch entries per user transform
  int n_blo = 1;
  int swap_in = 0, swap_out = 0;
  const void* stage_tw = nullptr;
  const void* tw4_hi = nullptr;
  const void* tw4_lo = nullptr;
  int tw4_bits = 0, tw4_shift = 0;
  idx_t tw4_n = 0;  // N of the four-step twiddle W_N (tables: tw4_hi / tw4_lo)
  int tw4_blo = 0;  // four-step twiddle index from the low batch index
  const void* real_tw = nullptr;
  idx_t rr_h = 0, rr_m = 0, rr_rs = 0;
  int rr_inv = 0;
  int in_nt = 0, out_nt = 0, xcd = 0;
  int sv = 0;                    // stage schedule variant of the kernel
  int hx = 0;                    // half-LDS kernel (real / imaginary parts exchanged in turn)
  int mfma = 0;                  // run on the split-precision MFMA DFT-256 tile
  const void* mfma_mats = nullptr;
};

// The fused four-step launch (afft_fused.hip) stays off: its first GPU run
// after a pointer fix did not finish within its time limit, and no cause has
// been shown yet.  A variant asking for it is rejected.
constexpr bool kFusedEnabled = false;

// Strided (column) axes up to this length run as one column pass; longer
// ones as a two-pass column four-step with balanced factors, whose tiles span
// more contiguous columns than one long column tile can.
constexpr idx_t kColDirectMax = 2048;
// 1-D FP32 real transforms up to these lengths run as the real-input
// four-step by default (variant real_pair = 2 asks for it at any length it
// has kernels for); longer ones as the complex four-step with a real pass.
// Measured over 2^16..2^23 with 256 MiB of reals per batch: r2c 0.31 vs 0.38
// ms (2^16), 0.34 vs 0.43 (2^20), 0.43 vs 0.45 (2^22), 0.47 vs 0.49 (2^23);
// c2r 0.29 vs 0.35 (2^16), 0.32 vs 0.33 (2^20), 0.36 vs 0.40 (2^22), but
// 0.52 vs 0.43 (2^23, whose 4096-point spectrum columns fit 4 per group).
constexpr idx_t kReal4R2cMax = (idx_t)1 << 23;
constexpr idx_t kReal4C2rMax = (idx_t)1 << 22;

struct Variant {
  int col_tile = 0;   // columns per group of column passes with L >= 512 (0 = default)
  int rowt_tile = 0;  // rows per group of transposed-store passes (0 = default)
  int stream_io = 2;  // non-temporal loads of src / stores of dst (keeps the MALL for scratch):
                      // 1 always, 0 never, 2 (default) unless an N-D plan's arrays fit the MALL
  int xcd_order = 0;  // XCD-contiguous tile order
  idx_t chunk = 0;    // transforms per pass group (0 = default)
  int sched = 0;      // 1 / 0 (default): radix-32 schedules where compiled, -1: never
  int half_lds = 0;   // 1 / 0 (default): half-LDS kernels where compiled, -1: never
  int pipeline = 0;   // removed schedule (pass A of group k beside pass B of group k-1): must be <= 0
  int real_pair = 0;  // fused real passes: 1 = all (also 1-D r2c pass B), 0 (default) = the N-D row-pair one, -1 = none
  int a_split = 0;    // two-pass plans: launches of pass A per group (0 / 1: one)
  int row_log2 = 0;   // 1-D four-step: row length n2 = 2^row_log2 (0 = default split)
  int fused = 0;      // 1: both four-step passes in one persistent launch where supported
  int fused_slots = 0;  // ring slots per XCD of the fused launch (0 = 3)
  int fused_lag = 0;    // rounds between pass A and pass B of a transform (0 = 1)
  int tiny_rows = 0;    // rows of 2..32 points: LDS-staged kernel with 64 / 128 / 256 rows per group
                        // (0 = default rows), -1: the general row kernel
};

// A plan lives in the caller's `mem`, as afft's does (fft_internal.hpp:24-106):
// create constructs it there, nothing else is allocated per plan, and the
// caller frees `mem` when done.  Its device resources are shared by all plans
// of the process (twiddle tables in TwCache, scratch in ScratchPool), so a plan
// whose memory is dropped without a destroy call leaks nothing.
struct Plan {
  uint32_t magic = 0;
  Variant var;
  int mode = 0;  // afft_gpu_mode_*
  int prec = 32;
  int kind = 0;
  int format = 1;
  int ndim = 0;
  idx_t dims[kMaxDim];
  idx_t batch = 1;
  idx_t chunk = 1;
  idx_t scr_per_user[2] = {0, 0};  // complex elements of scratch per user transform
  size_t scr_bytes[2] = {0, 0};    // bytes of scratch 0 / 1 one execute uses (from the pool)
  int npass = 0;
  PassRec pass[kMaxPass];
  hipStream_t stream = nullptr;
  int device = 0;
  // both passes of a 1-D FP32 c2c four-step in one persistent launch
  // (afft_fused.hip): ring slots per XCD and the pass-B lag, 0 = off
  int fused_slots = 0, fused_lag = 0;
};

// Plans sit at the first kPlanAlign boundary of `mem` (any alignment accepted,
// like afft's aligned_increment, common.hpp:173-183).
constexpr size_t kPlanAlign = 64;

Plan* plan_at(void* mem) {
  uintptr_t a = ((uintptr_t)mem + kPlanAlign - 1) & ~(uintptr_t)(kPlanAlign - 1);
  return reinterpret_cast<Plan*>(a);
}

Plan* get_plan(const void* state) {
  if (!state) {
    set_err("null plan");
    return nullptr;
  }
  Plan* P = reinterpret_cast<Plan*>(const_cast<void*>(state));
  if (P->magic != kMagic) {
    set_err("not a plan created by this library");
    return nullptr;
  }
  return P;
}

// ---- device twiddle cache ---------------------------------------------------
struct TwCache {
  std::mutex mu;
  std::map<std::tuple<int, int, int, long long, int>, void*> tabs;  // (device, prec, type, n, bits)
};
TwCache& twcache() {
  static TwCache* c = new TwCache;  // never destroyed: device memory lives for the process
  return *c;
}

// ---- device scratch pool ------------------------------------------------------
// Four-step and real plans need device scratch while they execute.  It comes
// from one pool per device instead of per-plan allocations, so plans hold no
// device memory of their own.  Slabs only grow: a request larger than the
// largest slab adds a slab of at least twice that size (all slabs together stay
// below twice the largest), and older slabs are kept because a HIP graph
// captured from an execute may still address them.  A slab is used by one
// execute at a time in stream order: an execute on another stream than the
// slab's previous user first waits for that user's last pass (an event), and an
// execute under graph capture skips that wait and the event (a captured plan's
// replays must not overlap executes of other plans on other streams).
struct Slab {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipEvent_t ev = nullptr;
  hipStream_t last = nullptr;
  bool used = false;
};
struct ScratchPool {
  std::mutex mu;
  std::map<int, std::vector<Slab>> slabs;  // by device; sizes increasing
};
ScratchPool& scratch_pool() {
  static ScratchPool* p = new ScratchPool;  // never destroyed: device memory lives for the process
  return *p;
}

// Exclusive use of `bytes` of the device's pool on stream `st` for the
// lifetime of the lease (the pool lock is held; launches go in between).
class ScratchLease {
 public:
  ScratchLease(int dev, size_t bytes, hipStream_t st) : lk_(scratch_pool().mu), st_(st) {
    if (bytes == 0) {
      ok_ = true;
      return;
    }
    std::vector<Slab>& v = scratch_pool().slabs[dev];
    if (v.empty() || v.back().bytes < bytes) {
      const size_t prev = v.empty() ? 0 : v.back().bytes;
      size_t want = std::max(bytes, 2 * prev);
      want = (want + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);
      Slab s;
      hipError_t e = hipMalloc(&s.ptr, want);
      if (e != hipSuccess && want > bytes) {
        want = bytes;
        e = hipMalloc(&s.ptr, want);
      }
      if (e == hipSuccess) e = hipEventCreateWithFlags(&s.ev, hipEventDisableTiming);
      if (e != hipSuccess) {
        set_err("scratch allocation of %zu bytes failed: %s", want, hipGetErrorString(e));
        if (s.ptr) (void)hipFree(s.ptr);
        return;
      }
      s.bytes = want;
      v.push_back(s);
    }
    slab_ = &v.back();
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    capturing_ = hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
    if (!capturing_ && slab_->used && slab_->last != st && hipStreamWaitEvent(st, slab_->ev, 0) != hipSuccess) {
      set_err("scratch: stream wait failed");
      return;
    }
    ok_ = true;
  }
  ~ScratchLease() {
    if (!ok_ || !slab_ || capturing_) return;
    if (hipEventRecord(slab_->ev, st_) == hipSuccess) {
      slab_->last = st_;
      slab_->used = true;
    }
  }
  bool ok() const { return ok_; }
  void* ptr() const { return slab_ ? slab_->ptr : nullptr; }

 private:
  std::unique_lock<std::mutex> lk_;
  hipStream_t st_;
  Slab* slab_ = nullptr;
  bool ok_ = false;
  bool capturing_ = false;
};

thread_local bool g_dry = false;  // host-only plan inspection: no device calls

template <typename T>
const void* get_table(int type, idx_t n, int bits) {
  if (g_dry) return (const void*)(uintptr_t)(0x1000 + type);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  auto key = std::make_tuple(dev, (int)sizeof(T), type, (long long)n, bits);
  TwCache& c = twcache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.tabs.find(key);
  if (it != c.tabs.end()) return it->second;
  void* p = nullptr;
  hipError_t e = hipSuccess;
  if (type == 0) {  // stage twiddles of an n-point pass
    int radix[8], ns[8];
    int nst = stage_list((int)n, bits, radix, ns);  // `bits` carries the schedule variant
    StageList sl{};
    sl.nstage = nst;
    int off = 0;
    for (int s = 0; s < nst; s++) {
      sl.radix[s] = radix[s];
      sl.ns[s] = ns[s];
      sl.off[s] = off;
      if (s >= 1) off += ns[s] * (radix[s] - 1);
    }
    sl.total = off;
    size_t bytes = sizeof(cx<T>) * (size_t)(off > 0 ? off : 1);
    e = hipMalloc(&p, bytes);
    if (e == hipSuccess) e = gen_stage_tw<T>((cx<T>*)p, sl, nullptr);
  } else if (type == 3) {  // A operands of the MFMA DFT-256 tile (bits: 1 = BF16 three-way)
    e = hipMalloc(&p, mfma_mats_bytes(bits));
    if (e == hipSuccess) e = gen_mfma_mats(p, bits, nullptr);
  } else if (type == 1) {  // four-step twiddles, hi then lo
    idx_t nhi = n >> bits, nlo = idx_t(1) << bits;
    e = hipMalloc(&p, sizeof(cx<T>) * (size_t)(nhi + nlo));
    if (e == hipSuccess) e = gen_tw4<T>((cx<T>*)p, (cx<T>*)p + nhi, n, bits, nullptr);
  } else {  // real-pass twiddles W_{2n}^k, k < n
    e = hipMalloc(&p, sizeof(cx<T>) * (size_t)(n > 0 ? n : 1));
    if (e == hipSuccess) e = gen_real_tw<T>((cx<T>*)p, n, nullptr);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    set_err("twiddle table: %s", hipGetErrorString(e));
    if (p) (void)hipFree(p);
    return nullptr;
  }
  c.tabs[key] = p;
  return p;
}

int ilog2(idx_t v) {
  int r = 0;
  while ((idx_t(1) << r) < v) r++;
  return r;
}
bool is_pow2(idx_t v) { return v >= 1 && (v & (v - 1)) == 0; }

// ---- plan construction --------------------------------------------------------
template <typename T>
struct Builder {
  Plan& P;
  bool ok = true;
  explicit Builder(Plan& p) : P(p) {}

  PassRec* add() {
    if (P.npass >= kMaxPass) {
      set_err("too many passes");
      ok = false;
      return nullptr;
    }
    PassRec* r = &P.pass[P.npass++];
    *r = PassRec();
    return r;
  }

  void need_scratch(int role, idx_t per_user) {
    if (role == R_SCR || role == R_SCR2) {
      idx_t& s = P.scr_per_user[role - R_SCR];
      if (per_user > s) s = per_user;
    }
  }

  // Row length n2 of a 1-D four-step n = n1 * n2: 1024-point rows up to 2^20;
  // beyond, 2048-point rows keep the column pass at <= 2048 points (2^21 as
  // 1024 x 2048 measured 3-5 % faster than 2048 x 1024); a variant may set it.
  // Longest row pass: the half-LDS-only lengths unless the variant turns the
  // half-LDS kernels off.
  idx_t row_max() const { return P.var.half_lds < 0 ? max_row_len_full<T>() : max_row_len<T>(); }

  // Row length n2 of the 1-D four-step n = n1 * n2: 1024-point rows up to
  // 2^20, but columns of at least 64 points (a pass A over 16- or 32-point
  // columns measured 1.29 TB/s for FP64 2^14, 2.84 as 64 x 256; FP64 2^15
  // 2.39 -> 2.75 as 64 x 512); longer rows beyond 2^20.
  idx_t row_len_1d(idx_t n) const {
    if (P.var.row_log2 > 0) return P.var.row_log2 < 40 ? (idx_t)1 << P.var.row_log2 : 0;
    if (n <= ((idx_t)1 << 20)) return std::min<idx_t>(1024, std::max<idx_t>(2, n / 64));
    return std::max<idx_t>(2048, n / max_col_len<T>());
  }

  // pref_ct > 0: a tile tried before the defaults (split-addressed column
  // passes ask for wide tiles, whose 4-byte re / im accesses then still cover
  // whole 256-byte row segments)
  bool fft_pass(PassRec* r, int L, int colmap, int epi, int store_t, int pref_ct = 0) {
    // tile: the variant's choice for column / transposed-store passes, else
    // the compiled default; schedule: radix-32 where compiled unless the
    // variant opts out; then the half-LDS kernel of that shape if one exists
    const int deflt = colmap ? col_tile<T>(L) : (store_t ? rowt_tile<T>(L) : row_tile<T>(L));
    const int want = colmap ? P.var.col_tile : (store_t ? P.var.rowt_tile : 0);
    const PassVariant<T>* v = nullptr;
    // column passes without an explicit tile first try a half-LDS kernel with
    // twice the default tile (wider row segments at the same LDS footprint;
    // compiled for the 256- and 512-point four-step pass A, where it measured
    // 5-8 % faster; a plain 256-point 32-column half-LDS tile measured 3 %
    // slower on 256^3 and is not compiled); short columns (<= 128 points) a
    // 4x-wide tile, whose LDS is still small (not for 1024-point columns:
    // 32-column tiles, one group per CU, measured 0.285 against 0.255 ms per
    // 64 x 2^20 direction for two 16-column groups; they stay a variant,
    // col_tile = 32)
    const bool wide = want <= 0 && colmap && P.var.half_lds >= 0 && L != 1024;
    // (the dim-0 real column passes too: r2c 256^3 0.135 -> 0.080 ms, 128^3 x 8
    // 0.124 -> 0.076, 256 x 512 x 256 0.320 -> 0.160; c2r 256^3 0.0885 -> 0.085,
    // but c2r with 64-point columns 0.087 -> 0.092, so those keep the default
    // tile; col_tile = 64 still selects the wide one)
    if (pref_ct <= 0 && want <= 0 && colmap && L <= 128 && !(epi == EPI_C2R && L < 128)) pref_ct = 4 * deflt;
    // and twice the default tile for the 256-point dim-0 r2c columns
    if (pref_ct <= 0 && want <= 0 && colmap && epi == EPI_R2C && L == 256) pref_ct = 2 * deflt;
    struct Cand {
      int ct, hx_only;
    };
    const Cand cands[4] = {
        {want > 0 ? 0 : pref_ct, 0}, {wide ? 2 * deflt : 0, 1}, {want > 0 ? want : deflt, 0}, {deflt, 0}};
    for (const Cand& c : cands)
      for (int sv : {1, 0})
        for (int hx : {1, 0}) {
          if (v || c.ct <= 0 || (c.hx_only && !hx) || (sv && P.var.sched < 0) || (hx && P.var.half_lds < 0))
            continue;
          v = find_pass<T>(L, c.ct, colmap, epi, store_t, sv, hx);
        }
    if (!v) {
      set_err("no kernel for L=%d colmap=%d epi=%d store_t=%d", L, colmap, epi, store_t);
      ok = false;
      return false;
    }
    const int CT = v->CT;
    r->type = PT_FFT;
    r->L = L;
    r->CT = CT;
    r->colmap = colmap;
    r->epi = epi;
    r->store_t = store_t;
    r->variant = v;
    r->sv = v->sv;
    r->hx = v->hx;
    r->stage_tw = get_table<T>(0, L, v->sv);
    if (!r->stage_tw) ok = false;
    if (epi == EPI_R2C || epi == EPI_C2R) {
      r->real_tw = L >= 2 ? get_table<T>(2, L, 0) : get_table<T>(2, 1, 0);
      if (!r->real_tw) ok = false;
    }
    return ok;
  }

  // Fold user transforms into the column index when every side is a plain
  // stack of the per-user tensor.
  static bool can_fold(const Side& a, const Side& b, idx_t outer_u) {
    auto f = [&](const Side& s) { return s.role >= R_SCR || s.s_u == outer_u * s.s_c; };
    return f(a) && f(b);
  }

  // A row pass of 2..32-point rows whose rows form one stack (every user
  // transform folded into the row index, entries of a row adjacent) runs on
  // the LDS-staged one-row-per-thread kernel: its loads and stores are runs
  // over the work-group's rows instead of one 8-byte entry per row per lane.
  void tiny(PassRec* r) {
    if (!ok || P.var.tiny_rows < 0 || r->type != PT_FFT || r->colmap || r->store_t || !r->fold_user) return;
    const int rows = P.var.tiny_rows > 0 ? P.var.tiny_rows : tiny_default_rows((int)sizeof(T), r->L, r->epi);
    if (r->in.s_j != 1 || r->out.s_j != 1 || !tiny_supported(r->L, r->epi, rows)) return;
    r->type = PT_TINY;
    r->variant = nullptr;
    r->CT = rows;
    r->sv = r->hx = 0;
    r->stage_tw = nullptr;
  }

  void scratch_side(Side& s, int role) {
    s = Side();
    s.role = role;
    s.il = 1;
  }

  // c2c along one axis of the per-user tensor (outer_u, n, inner); in and out
  // carry role/format/offsets and s_u, strides are filled here.
  void axis(Side in, Side out, idx_t outer_u, idx_t n, idx_t inner, int swi, int swo) {
    if (!ok) return;
    const idx_t tens = outer_u * n * inner;
    need_scratch(in.role, tens);
    need_scratch(out.role, tens);
    if (in.role < R_SCR && in.s_u == 0) in.s_u = tens;
    if (out.role < R_SCR && out.s_u == 0) out.s_u = tens;
    const int tmp_role = (in.role == R_SCR || out.role == R_SCR) ? R_SCR2 : R_SCR;

    if (inner == 1 && n <= row_max()) {
      PassRec* r = add();
      if (!r || !fft_pass(r, (int)n, 0, EPI_PLAIN, 0)) return;
      r->in = in;
      r->out = out;
      r->in.s_j = 1;
      r->out.s_j = 1;
      r->in.s_c = n;
      r->out.s_c = n;
      r->n_col = outer_u;
      if (can_fold(r->in, r->out, outer_u)) {
        r->fold_user = 1;
      } else {
        r->in.s_b = in.role >= R_SCR ? tens : in.s_u;
        r->out.s_b = out.role >= R_SCR ? tens : out.s_u;
        r->b_per_user = 1;
      }
      r->swap_in = swi;
      r->swap_out = swo;
      tiny(r);
      return;
    }
    // batch stride of a side whose batch index runs over (user, outer)
    auto bstride = [&](const Side& s, idx_t per_outer) -> idx_t {
      if (s.role < R_SCR && outer_u == 1) return s.s_u;
      return per_outer;
    };
    if (outer_u > 1 && ((in.role < R_SCR && in.s_u != tens) || (out.role < R_SCR && out.s_u != tens))) {
      set_err("internal: non-contiguous batch with outer axes");
      ok = false;
      return;
    }
    if (inner > 1 && n <= kColDirectMax) {
      PassRec* r = add();
      if (!r || !fft_pass(r, (int)n, 1, EPI_PLAIN, 0)) return;
      r->in = in;
      r->out = out;
      r->in.s_j = inner;
      r->out.s_j = inner;
      r->in.s_c = 1;
      r->out.s_c = 1;
      r->in.s_b = bstride(in, n * inner);
      r->out.s_b = bstride(out, n * inner);
      r->n_col = inner;
      r->b_per_user = outer_u;
      r->swap_in = swi;
      r->swap_out = swo;
      return;
    }
    // 1-D lengths beyond 2048 x 2048: three passes.  Two passes would need
    // 4096-point columns, four per tile (32-byte row segments): 8 x c2c 2^23
    // took 1.17 ms and 4 x 2^24 1.73 ms as two passes
    if (inner == 1 && P.var.row_log2 <= 0 && n > (idx_t)2048 * 2048) {
      fourstep3(in, out, outer_u, n, swi, swo, tmp_role);
      return;
    }
    // four-step along the axis: n = n1 * n2
    idx_t n2, n1;
    if (inner == 1) {
      n2 = row_len_1d(n);
      if (P.mode != afft_gpu_mode_fma && n == 65536 && P.var.row_log2 <= 0) n2 = 256;  // 256 x 256 on the matrix cores
      if (n2 < 2 || n2 > 4096 || n2 >= n || !find_pass<T>((int)n2, 0, 0, EPI_PLAIN, 1)) {
        set_err("transform length %lld is too long", (long long)n);
        ok = false;
        return;
      }
    } else {
      // strided axis: balanced split, both column passes <= 4096 points
      n2 = n >> (ilog2(n) / 2);
      if (n2 > max_col_len<T>() || n / n2 > max_col_len<T>()) {
        set_err("axis length %lld is too long", (long long)n);
        ok = false;
        return;
      }
    }
    n1 = n / n2;
    if (n1 > max_col_len<T>() || n1 < 1) {
      set_err("transform length %lld is too long", (long long)n);
      ok = false;
      return;
    }
    need_scratch(tmp_role, tens);
    Side tA, tB;
    scratch_side(tA, tmp_role);
    scratch_side(tB, tmp_role);
    const int bits = (ilog2(n) + 1) / 2;
    const void* tw4 = get_table<T>(1, n, bits);
    if (!tw4) {
      ok = false;
      return;
    }
    // pass A: length-n1 columns (stride n2*inner), twiddle W_n^(n2idx*k1)
    PassRec* a = add();
    if (!a || !fft_pass(a, (int)n1, 1, EPI_TW4, 0)) return;
    a->in = in;
    a->out = tA;
    a->in.s_j = n2 * inner;
    a->out.s_j = n2 * inner;
    a->in.s_c = 1;
    a->out.s_c = 1;
    a->in.s_b = bstride(in, n * inner);
    a->out.s_b = n * inner;
    a->n_col = n2 * inner;
    a->b_per_user = outer_u;
    a->swap_in = swi;
    a->tw4_hi = tw4;
    a->tw4_lo = (const cx<T>*)tw4 + (n >> bits);
    a->tw4_n = n;
    a->tw4_bits = bits;
    a->tw4_shift = ilog2(inner);
    // pass B: length-n2 transforms of every (k1, c); output bin k1 + n1*k2
    PassRec* b = add();
    if (inner == 1) {
      if (!b || !fft_pass(b, (int)n2, 0, EPI_PLAIN, 1)) return;
      b->in = tB;
      b->out = out;
      b->in.s_j = 1;
      b->in.s_c = n2;
      b->in.s_b = n;
      b->out.s_j = n1;
      b->out.s_c = 1;
      b->out.s_b = bstride(out, n);
      b->n_col = n1;
      b->b_per_user = outer_u;
      if (outer_u > 1) {
        // outer index within a user transform, then the user index
        b->n_blo = (int)outer_u;
        b->in.s_b2 = n;
        b->out.s_b2 = n;
        b->in.s_b = outer_u * n;
        b->out.s_b = (out.role >= R_SCR) ? outer_u * n : out.s_u;
      }
    } else {
      if (!b || !fft_pass(b, (int)n2, 1, EPI_PLAIN, 0)) return;
      b->in = tB;
      b->out = out;
      b->in.s_j = inner;
      b->in.s_c = 1;
      b->in.s_b = n * inner;
      b->in.s_b2 = n2 * inner;
      b->out.s_j = n1 * inner;
      b->out.s_c = 1;
      b->out.s_b = bstride(out, n * inner);
      b->out.s_b2 = inner;
      b->n_col = inner;
      b->n_blo = (int)n1;
      b->b_per_user = outer_u * n1;
    }
    b->swap_out = swo;
  }

  // 1-D c2c of n = p * q * n2 points in three passes: the four-step
  // n = p * M (M = q * n2) whose second step, the M-point DFTs of the p rows,
  // is itself a four-step.
  //   pass 1: p-point columns (stride M) of the input, times W_n^(jM * kp),
  //           into scratch rows kp of M;
  //   pass 2: in place, q-point columns (stride n2) of each row kp, times
  //           W_M^(j2 * kq);
  //   pass 3: n2-point rows (kp, kq) with the transposed store of bin
  //           kp + p * (kq + q * k2): for a block kq, the rows of consecutive
  //           kp land next to each other (s_c = 1), the blocks p apart.
  void fourstep3(Side in, Side out, idx_t outer_u, idx_t n, int swi, int swo, int tmp_role) {
    const idx_t n2 = 2048;
    const idx_t rem = n / n2;
    const idx_t p = (idx_t)1 << ((ilog2(rem) + 1) / 2), q = rem / p, M = q * n2;
    if (p > max_col_len<T>() || q > max_col_len<T>() || q < 2) {
      set_err("transform length %lld is too long", (long long)n);
      ok = false;
      return;
    }
    const idx_t tens = outer_u * n;
    need_scratch(tmp_role, tens);
    auto bstride = [&](const Side& s, idx_t per_outer) -> idx_t {
      if (s.role < R_SCR && outer_u == 1) return s.s_u;
      return per_outer;
    };
    const int bits = (ilog2(n) + 1) / 2, bits_m = (ilog2(M) + 1) / 2;
    const void* tw_n = get_table<T>(1, n, bits);
    const void* tw_m = get_table<T>(1, M, bits_m);
    if (!tw_n || !tw_m) {
      ok = false;
      return;
    }
    Side t;
    scratch_side(t, tmp_role);
    PassRec* a = add();
    if (!a || !fft_pass(a, (int)p, 1, EPI_TW4, 0)) return;
    a->in = in;
    a->out = t;
    a->in.s_j = a->out.s_j = M;
    a->in.s_c = a->out.s_c = 1;
    a->in.s_b = bstride(in, n);
    a->out.s_b = n;
    a->n_col = M;
    a->b_per_user = outer_u;
    a->swap_in = swi;
    a->tw4_hi = tw_n;
    a->tw4_lo = (const cx<T>*)tw_n + (n >> bits);
    a->tw4_n = n;
    a->tw4_bits = bits;
    PassRec* c = add();
    if (!c || !fft_pass(c, (int)q, 1, EPI_TW4, 0)) return;
    c->in = c->out = t;
    c->in.s_j = c->out.s_j = n2;
    c->in.s_c = c->out.s_c = 1;
    c->in.s_b = c->out.s_b = M;
    c->n_col = n2;
    c->b_per_user = outer_u * p;
    c->tw4_hi = tw_m;
    c->tw4_lo = (const cx<T>*)tw_m + (M >> bits_m);
    c->tw4_n = M;
    c->tw4_bits = bits_m;
    PassRec* b = add();
    if (!b || !fft_pass(b, (int)n2, 0, EPI_PLAIN, 1)) return;
    b->in = t;
    b->out = out;
    b->in.s_j = 1;
    b->in.s_c = M;
    b->in.s_b = n;
    b->in.s_b2 = n2;
    b->out.s_j = p * q;
    b->out.s_c = 1;
    b->out.s_b = bstride(out, n);
    b->out.s_b2 = p;
    b->n_col = p;
    b->n_blo = (int)q;
    b->b_per_user = outer_u * q;
    b->swap_out = swo;
  }

  // c2c over every axis of the per-user tensor dims[0..nd): last axis first
  // (reads `in`), then the outer axes in place on `out`.
  void all_axes(Side in, Side out, const idx_t* d, int nd, int swi, int swo) {
    idx_t total = 1;
    for (int i = 0; i < nd; i++) total *= d[i];
    Side cur = in;
    for (int a = nd - 1; a >= 0; a--) {
      idx_t outer = 1, inner = 1;
      for (int i = 0; i < a; i++) outer *= d[i];
      for (int i = a + 1; i < nd; i++) inner *= d[i];
      if (d[a] == 1 && !(a == nd - 1 && nd == 1)) continue;
      Side o = out;
      axis(cur, o, outer, d[a], inner, swi, swo);
      cur = out;
      (void)total;
    }
  }

  // 1-D r2c with h = n/2 > the longest row pass: four-step h = n1 * n2 whose
  // pass B stores rows (k1-major, contiguous) into a second scratch, then the
  // transposing real pass reads Z[k1 + n1*k2] = Y[k1][k2] and writes the h+1
  // bins in natural order (one kernel for the transposition and the real
  // post-pass instead of a transposed store and a separate real pass).
  // Row length of r2c_fourstep_tr: its pass B stores rows (not transposed),
  // so long rows cost nothing there, while 4096-point columns fit only four
  // per pass-A tile: 1024-point columns where the rows fit one row pass, else
  // the 1-D default (FP32 4 x r2c 2^24: 0.742 ms as 4096 x 2048, 0.428 as
  // 2048 x 4096, 0.410 as 1024 x 8192).
  idx_t r2c_tr_row_len(idx_t h) const {
    if (P.var.row_log2 > 0 || h < ((idx_t)1 << 21)) return row_len_1d(h);
    if (h / 1024 <= row_max()) return h / 1024;
    if (h / 2048 <= row_max()) return h / 2048;
    return row_len_1d(h);
  }

  void r2c_fourstep_tr(const Side& in, const Side& out, idx_t h) {
    const idx_t n2 = r2c_tr_row_len(h);
    const idx_t n1 = n2 >= 2 ? h / n2 : 0;
    if (n1 > max_col_len<T>() || n1 < 32) {
      set_err("transform length %lld is too long", (long long)(2 * h));
      ok = false;
      return;
    }
    need_scratch(R_SCR, h);
    const int bits = (ilog2(h) + 1) / 2;
    const void* tw4 = get_table<T>(1, h, bits);
    if (!tw4) {
      ok = false;
      return;
    }
    PassRec* a = add();
    if (!a || !fft_pass(a, (int)n1, 1, EPI_TW4, 0)) return;
    a->in = in;
    a->in.s_j = n2;
    a->in.s_c = 1;
    a->in.s_b = in.s_u;
    scratch_side(a->out, R_SCR);
    a->out.s_j = n2;
    a->out.s_c = 1;
    a->out.s_b = h;
    a->n_col = n2;
    a->b_per_user = 1;
    a->tw4_hi = tw4;
    a->tw4_lo = (const cx<T>*)tw4 + (h >> bits);
    a->tw4_n = h;
    a->tw4_bits = bits;
    if (P.var.real_pair > 0 && real_pair_supported((int)sizeof(T), (int)n1, (int)n2)) {
      // rows k1 and n1 - k1 transformed together, the real post-pass applied
      // between them and the bins stored in natural order (afft_realpair.hip)
      PassRec* r = add();
      if (!r) return;
      r->type = PT_REALPAIR;
      r->L = (int)n2;
      r->CT = 16;
      r->sv = 1;
      r->hx = 1;
      scratch_side(r->in, R_SCR);
      r->in.s_j = 1;
      r->in.s_c = n2;
      r->in.s_b = h;
      r->out = out;
      r->out.s_j = 1;
      r->out.s_b = out.s_u;
      r->rr_h = h;
      r->rr_m = n1;
      r->stage_tw = get_table<T>(0, (int)n2, 1);
      r->real_tw = get_table<T>(2, h, 0);
      r->tw4_hi = get_table<T>(2, n2, 0);
      if (!r->stage_tw || !r->real_tw || !r->tw4_hi) ok = false;
      return;
    }
    need_scratch(R_SCR2, h);
    PassRec* b = add();
    if (!b || !fft_pass(b, (int)n2, 0, EPI_PLAIN, 0)) return;
    scratch_side(b->in, R_SCR);
    b->in.s_j = 1;
    b->in.s_c = n2;
    b->in.s_b = h;
    scratch_side(b->out, R_SCR2);
    b->out.s_j = 1;
    b->out.s_c = n2;
    b->out.s_b = h;
    b->n_col = n1;
    b->b_per_user = 1;
    PassRec* r = add();
    if (!r) return;
    r->type = PT_REALTR;
    scratch_side(r->in, R_SCR2);
    r->in.s_b = h;
    r->out = out;
    r->out.s_j = 1;
    r->out.s_b = out.s_u;
    r->rr_h = h;
    r->rr_m = n1;
    r->real_tw = get_table<T>(2, h, 0);
    if (!r->real_tw) ok = false;
  }

  // 1-D FP32 r2c of N = 2h reals as a real-input four-step N = n1 * n2
  // (afft_real4.hip): pass A over two real columns per complex column, rows
  // k1 <= n1/2 of the twiddled column spectra into scratch; pass B the row
  // DFTs, storing each bin once.  False when no kernels exist for the split.
  bool r2c_real4(const Side& in, const Side& out, idx_t h) {
    const idx_t N = 2 * h;
    if (sizeof(T) != 4 || !(P.var.real_pair == 2 || (P.var.real_pair == 0 && N <= kReal4R2cMax))) return false;
    const int lg = ilog2(N);
    idx_t n2 = P.var.row_log2 > 0 && P.var.row_log2 < 40 ? (idx_t)1 << P.var.row_log2 : (idx_t)1 << (lg - lg / 2);
    if (n2 >= N) return false;
    const idx_t n1 = N / n2;
    int cta = 0, sva = 0, hxa = 0, ctb = 0, svb = 0, hxb = 0;
    const int want_sv = P.var.sched < 0 ? 0 : (P.var.sched > 0 ? 1 : -1);
    const int want_hx = P.var.half_lds < 0 ? 0 : (P.var.half_lds > 0 ? 1 : -1);
    if (!real4_shape(0, (int)sizeof(T), (int)std::min<idx_t>(n1, 1 << 30), P.var.col_tile, want_sv, want_hx, cta, sva,
                     hxa) ||
        !real4_shape(1, (int)sizeof(T), (int)std::min<idx_t>(n2, 1 << 30), P.var.rowt_tile, want_sv, want_hx, ctb,
                     svb, hxb) ||
        (n2 / 2) % cta != 0)
      return false;
    const int bits = (lg + 1) / 2;
    const void* tw4 = get_table<T>(1, N, bits);
    const void* sa = get_table<T>(0, n1, sva);
    const void* sb = get_table<T>(0, n2, svb);
    if (!tw4 || !sa || !sb) {
      ok = false;
      return true;
    }
    const idx_t rows = n1 / 2 + 1;
    need_scratch(R_SCR, rows * n2);
    PassRec* a = add();
    if (!a) return true;
    a->type = PT_REAL4A;
    a->L = (int)n1;
    a->CT = cta;
    a->sv = sva;
    a->colmap = 1;
    a->in = in;  // N reals as h interleaved complex pairs
    a->in.s_j = n2 / 2;
    a->in.s_c = 1;
    a->in.s_b = in.s_u;
    scratch_side(a->out, R_SCR);
    a->out.s_j = n2;
    a->out.s_c = 2;
    a->out.s_b = rows * n2;
    a->n_col = n2 / 2;
    a->rr_h = h;
    a->rr_m = n2;
    a->stage_tw = sa;
    a->hx = hxa;
    a->tw4_hi = tw4;
    a->tw4_lo = (const cx<T>*)tw4 + (N >> bits);
    a->tw4_n = N;
    a->tw4_bits = bits;
    PassRec* b = add();
    if (!b) return true;
    b->type = PT_REAL4B;
    b->L = (int)n2;
    b->CT = ctb;
    b->sv = svb;
    b->hx = hxb;
    scratch_side(b->in, R_SCR);
    b->in.s_j = 1;
    b->in.s_c = n2;
    b->in.s_b = rows * n2;
    b->out = out;
    b->out.s_j = n1;
    b->out.s_c = 1;
    b->out.s_b = out.s_u;
    b->n_col = rows;
    b->rr_h = h;
    b->rr_m = n1;
    b->stage_tw = sb;
    return true;
  }

  // 1-D FP32 c2r of h+1 bins to N = 2h reals, the same split run backwards:
  // pass B' the inverse DFTs of the spectrum's columns k1 <= n1/2 (bins
  // k1 + n1*k2, past N/2 the conjugates of the mirrored column) into the
  // scratch Dt[j2][k1], pass A' the inverse DFTs along the rows of Dt, two
  // real columns per complex column, stored through LDS.
  bool c2r_real4(const Side& in, const Side& out, idx_t h) {
    const idx_t N = 2 * h;
    if (sizeof(T) != 4 || !(P.var.real_pair == 2 || (P.var.real_pair == 0 && N <= kReal4C2rMax))) return false;
    const int lg = ilog2(N);
    idx_t n2 = P.var.row_log2 > 0 && P.var.row_log2 < 40 ? (idx_t)1 << P.var.row_log2 : (idx_t)1 << (lg - lg / 2);
    if (n2 >= N) return false;
    const idx_t n1 = N / n2;
    int ctc = 0, svc = 0, hxc = 0, ctd = 0, svd = 0, hxd = 0;
    const int want_sv = P.var.sched < 0 ? 0 : (P.var.sched > 0 ? 1 : -1);
    const int want_hx = P.var.half_lds < 0 ? 0 : (P.var.half_lds > 0 ? 1 : -1);
    if (!real4_shape(2, (int)sizeof(T), (int)std::min<idx_t>(n2, 1 << 30), P.var.col_tile, want_sv, want_hx, ctc, svc,
                     hxc) ||
        !real4_shape(3, (int)sizeof(T), (int)std::min<idx_t>(n1, 1 << 30), P.var.rowt_tile, want_sv, want_hx, ctd,
                     svd, hxd) ||
        (n2 / 2) % ctd != 0)
      return false;
    const int bits = (lg + 1) / 2;
    const void* tw4 = get_table<T>(1, N, bits);
    const void* sc = get_table<T>(0, n2, svc);
    const void* sd = get_table<T>(0, n1, svd);
    if (!tw4 || !sc || !sd) {
      ok = false;
      return true;
    }
    // rows of Dt padded to whole 128-byte lines
    const idx_t rs = (n1 / 2 + 1 + 15) & ~(idx_t)15;
    need_scratch(R_SCR, n2 * rs);
    PassRec* c = add();
    if (!c) return true;
    c->type = PT_REAL4C;
    c->L = (int)n2;
    c->CT = ctc;
    c->sv = svc;
    c->colmap = 1;
    c->hx = hxc;
    c->in = in;  // bins 0..h: column k1 at k1, bin k2 of it at n1*k2
    c->in.s_j = n1;
    c->in.s_c = 1;
    c->in.s_b = in.s_u;
    scratch_side(c->out, R_SCR);
    c->out.s_j = rs;
    c->out.s_c = 1;
    c->out.s_b = n2 * rs;
    c->n_col = n1 / 2 + 1;
    c->rr_h = h;
    c->rr_m = n1;
    c->stage_tw = sc;
    c->tw4_hi = tw4;
    c->tw4_lo = (const cx<T>*)tw4 + (N >> bits);
    c->tw4_n = N;
    c->tw4_bits = bits;
    PassRec* d = add();
    if (!d) return true;
    d->type = PT_REAL4D;
    d->L = (int)n1;
    d->CT = ctd;
    d->sv = svd;
    d->store_t = 1;
    d->hx = hxd;
    scratch_side(d->in, R_SCR);  // packed column c: rows 2c and 2c+1 of Dt
    d->in.s_j = 1;
    d->in.s_c = 2 * rs;
    d->in.s_b = n2 * rs;
    d->out = out;  // N reals as h interleaved complex pairs
    d->out.s_j = n2 / 2;
    d->out.s_c = 1;
    d->out.s_b = out.s_u;
    d->n_col = n2 / 2;
    d->rr_h = h;
    d->rr_m = n2;
    d->stage_tw = sd;
    return true;
  }

  // The inner c2c axes of an N-D r2c, on each of the h+1 rows of dst.
  void inner_axes_on_dst(const Side& usr_dst, const idx_t* d, int nd, idx_t h) {
    idx_t inner = 1;
    for (int i = 1; i < nd; i++) inner *= d[i];
    idx_t dd[kMaxDim];
    dd[0] = h + 1;
    for (int i = 1; i < nd; i++) dd[i] = d[i];
    Side o = usr_dst;
    o.s_u = (h + 1) * inner;
    for (int a = nd - 1; a >= 1; a--) {
      idx_t outer = 1, inn = 1;
      for (int i = 0; i < a; i++) outer *= dd[i];
      for (int i = a + 1; i < nd; i++) inn *= dd[i];
      axis(o, o, outer, dd[a], inn, 0, 0);
    }
  }

  // Dim 0 of an N-D r2c when h = n0/2 is too long for one column pass: the
  // h-point c2c over the row pairs (x[2j] + i x[2j+1]) as a column four-step
  // h = h1 * h2 (pass A: h1-point columns of rows n1*h2 + n2, batch index
  // n2, twiddle W_h^(n2*k1); pass B: h2-point columns, output row k1 + h1*k2
  // of dst), then the real pass over row pairs (k, h-k) in place on dst.
  void real_dim0_fourstep_r2c(const Side& usr_src, const Side& usr_dst, idx_t n0, idx_t inner, bool real_pass = true) {
    const idx_t h = n0 / 2;
    const idx_t h1 = (idx_t)1 << (ilog2(h) / 2), h2 = h / h1;
    need_scratch(R_SCR, h * inner);
    const int bits = (ilog2(h) + 1) / 2;
    const void* tw4 = get_table<T>(1, h, bits);
    if (!tw4) {
      ok = false;
      return;
    }
    PassRec* a = add();
    if (!a || !fft_pass(a, (int)h1, 1, EPI_TW4, 0, 4 * col_tile<T>((int)h1))) return;
    a->in = usr_src;
    a->in.il = 0;
    a->in.one_array = 1;
    a->in.re_off = 0;
    a->in.im_off = inner;
    a->in.s_j = 2 * h2 * inner;
    a->in.s_c = 1;
    a->in.s_b2 = 2 * inner;
    a->in.s_b = a->in.s_u = n0 * inner;
    scratch_side(a->out, R_SCR);
    a->out.s_j = h2 * inner;
    a->out.s_c = 1;
    a->out.s_b2 = inner;
    a->out.s_b = h * inner;
    a->n_col = inner;
    a->n_blo = (int)h2;
    a->b_per_user = h2;
    a->tw4_hi = tw4;
    a->tw4_lo = (const cx<T>*)tw4 + (h >> bits);
    a->tw4_n = h;
    a->tw4_bits = bits;
    a->tw4_blo = 1;
    const int rct = real_pass && P.var.real_pair >= 0 ? real_colpair_ct((int)sizeof(T), (int)h2, inner, (int)h1) : 0;
    if (rct > 0) {
      // pass B with the real pass over row pairs (k, h - k): tiles k1 and
      // h1 - k1 per work-group (afft_realpair.hip)
      PassRec* r = add();
      if (!r) return;
      r->type = PT_REALCOLPAIR;
      r->L = (int)h2;
      r->CT = rct;
      r->colmap = 1;
      scratch_side(r->in, R_SCR);
      r->in.s_j = inner;
      r->in.s_c = 1;
      r->in.s_b2 = h2 * inner;
      r->in.s_b = h * inner;
      r->out = usr_dst;
      r->out.s_j = inner;
      r->out.s_c = 1;
      r->out.s_b = r->out.s_u = (h + 1) * inner;
      r->n_col = inner;
      r->n_blo = (int)h1;
      r->rr_h = h;
      r->rr_m = inner;
      r->stage_tw = get_table<T>(0, (int)h2, 0);
      r->real_tw = get_table<T>(2, h, 0);
      if (!r->stage_tw || !r->real_tw) ok = false;
      return;
    }
    PassRec* b = add();
    if (!b || !fft_pass(b, (int)h2, 1, EPI_PLAIN, 0)) return;
    scratch_side(b->in, R_SCR);
    b->in.s_j = inner;
    b->in.s_c = 1;
    b->in.s_b2 = h2 * inner;
    b->in.s_b = h * inner;
    b->out = usr_dst;
    b->out.s_j = h1 * inner;
    b->out.s_c = 1;
    b->out.s_b2 = inner;
    b->out.s_b = b->out.s_u = (h + 1) * inner;
    b->n_col = inner;
    b->n_blo = (int)h1;
    b->b_per_user = h1;
    if (!real_pass) return;
    Side rr = usr_dst;
    rr.s_j = inner;
    rr.s_c = 1;
    rr.s_b = rr.s_u = (h + 1) * inner;
    real_rows(rr, rr, h, inner, 0);
  }

  // Dim 0 of an N-D c2r with a long dim 0: `cur` holds the h+1 spectrum rows
  // after the inner axes (plan scratch); the inverse real pass runs in place
  // on it, then the inverse h-point c2c as a column four-step whose pass B
  // writes row k1 + h1*k2 as the real rows 2(k1 + h1*k2), 2(k1 + h1*k2) + 1.
  void real_dim0_fourstep_c2r(const Side& cur, const Side& usr_dst, idx_t n0, idx_t inner, bool real_pass = true) {
    const idx_t h = n0 / 2;
    const idx_t h1 = (idx_t)1 << (ilog2(h) / 2), h2 = h / h1;
    const int rct = real_pass && P.var.real_pair >= 0 && cur.il && cur.role >= R_SCR
                        ? real_colpair_ct((int)sizeof(T), (int)h1, inner, (int)h2)
                        : 0;
    if (rct > 0) {
      // the inverse real pass over row pairs inside pass A: blocks n2 and
      // h2 - n2 per work-group (afft_realpair.hip)
      need_scratch(R_SCR2, h * inner);
      const int bits = (ilog2(h) + 1) / 2;
      const void* tw4 = get_table<T>(1, h, bits);
      PassRec* r = add();
      if (!r || !tw4) {
        ok = false;
        return;
      }
      r->type = PT_REALICOLPAIR;
      r->L = (int)h1;
      r->CT = rct;
      r->colmap = 1;
      r->in = cur;
      r->in.s_j = inner;
      r->in.s_c = 1;
      r->in.s_b = (h + 1) * inner;
      scratch_side(r->out, R_SCR2);
      r->out.s_j = h2 * inner;
      r->out.s_c = 1;
      r->out.s_b = h * inner;
      r->n_col = inner;
      r->n_blo = (int)h2;
      r->rr_h = h;
      r->rr_m = inner;
      r->rr_inv = 1;
      r->stage_tw = get_table<T>(0, (int)h1, 0);
      r->real_tw = get_table<T>(2, h, 0);
      r->tw4_hi = tw4;
      r->tw4_lo = (const cx<T>*)tw4 + (h >> bits);
      r->tw4_n = h;
      r->tw4_bits = bits;
      if (!r->stage_tw || !r->real_tw) ok = false;
    }
    if (real_pass && rct == 0) {
      Side rr = cur;
      rr.s_j = inner;
      rr.s_c = 1;
      rr.s_b = (h + 1) * inner;
      real_rows(rr, rr, h, inner, 1);
    }
    need_scratch(R_SCR2, h * inner);
    const int bits = (ilog2(h) + 1) / 2;
    const void* tw4 = get_table<T>(1, h, bits);
    if (!tw4) {
      ok = false;
      return;
    }
    // pass A: h1-point columns of rows n1*h2 + n2 (column index n2*inner + c)
    PassRec* a = rct > 0 ? nullptr : add();
    if (rct == 0 && (!a || !fft_pass(a, (int)h1, 1, EPI_TW4, 0))) return;
    if (a) {
      a->in = cur;
      a->in.s_j = h2 * inner;
      a->in.s_c = 1;
      a->in.s_b = (h + 1) * inner;
      scratch_side(a->out, R_SCR2);
      a->out.s_j = h2 * inner;
      a->out.s_c = 1;
      a->out.s_b = h * inner;
      a->n_col = h2 * inner;
      a->b_per_user = 1;
      a->swap_in = 1;
      a->tw4_hi = tw4;
      a->tw4_lo = (const cx<T>*)tw4 + (h >> bits);
      a->tw4_n = h;
      a->tw4_bits = bits;
      a->tw4_shift = ilog2(inner);
    }
    // pass B: h2-point columns (batch index k1), real rows out
    PassRec* b = add();
    if (!b || !fft_pass(b, (int)h2, 1, EPI_PLAIN, 0, 4 * col_tile<T>((int)h2))) return;
    scratch_side(b->in, R_SCR2);
    b->in.s_j = inner;
    b->in.s_c = 1;
    b->in.s_b2 = h2 * inner;
    b->in.s_b = h * inner;
    b->out = usr_dst;
    b->out.il = 0;
    b->out.one_array = 1;
    b->out.re_off = 0;
    b->out.im_off = inner;
    b->out.s_j = 2 * h1 * inner;
    b->out.s_c = 1;
    b->out.s_b2 = 2 * inner;
    b->out.s_b = b->out.s_u = n0 * inner;
    b->n_col = inner;
    b->n_blo = (int)h1;
    b->b_per_user = h1;
    b->swap_out = 1;
  }

  void real_rows(Side in, Side out, idx_t h, idx_t m, int inv) {
    PassRec* r = add();
    if (!r) return;
    r->type = PT_REALROWS;
    r->in = in;
    r->out = out;
    r->rr_h = h;
    r->rr_m = m;
    r->rr_inv = inv;
    r->real_tw = get_table<T>(2, h, 0);
    if (!r->real_tw) ok = false;
  }

  bool build(int kind, int format, const idx_t* d, int nd) {
    P.npass = 0;
    P.scr_per_user[0] = P.scr_per_user[1] = 0;
    bool r = build_passes(kind, format, d, nd);
    // stream_io 2: an N-D plan (its outer axes re-read dst) whose arrays hold at
    // most 128 MiB per side keeps cached I/O, so input and output share the
    // 256 MB MALL with the passes between them (measured: r2c 256^3 0.165 ->
    // 0.137 ms, c2c 1024^2 x 16 0.098 -> 0.092, c2c 256^3 0.131 -> 0.127);
    // 256 MiB arrays stream (c2c 4096^2 x 2 0.277 streaming, 0.293 cached)
    bool nd_reread = false;
    for (int i = 0; i < P.npass && !nd_reread; i++)
      for (int j = i + 1; j < P.npass; j++) nd_reread = nd_reread || (P.pass[j].in.role == P.pass[i].out.role &&
                                                                     P.pass[i].out.role == R_DST);
    double side_bytes = (double)P.batch * (double)sizeof(T) * (kind < 2 ? 2.0 : 1.0);
    for (int k = 0; k < nd; k++) side_bytes *= (double)d[k];
    const int sio = P.var.stream_io == 2 ? !(nd_reread && side_bytes <= 128.0 * 1048576.0) : P.var.stream_io;
    for (int i = 0; i < P.npass; i++) {
      PassRec& q = P.pass[i];
      // Streaming (non-temporal) accesses for the caller's input and for the
      // final write of the output; a pass whose output a later pass reads
      // again (the outer axes of N-D plans run in place on dst) stores
      // through the caches, so the array can stay in the MALL in between.
      bool reread = false;
      for (int j = i + 1; j < P.npass; j++) reread = reread || P.pass[j].in.role == q.out.role;
      q.in_nt = sio && q.in.role == R_SRC;
      q.out_nt = sio && q.out.role == R_DST && !reread;
      // the real-input four-step's row passes store (r2c) or load (c2r) the
      // spectrum in runs of a tile's rows; a transform's N/2 + 1 bins rarely
      // start on a line boundary, so neighbouring tiles share lines: cached
      // accesses let those halves meet in L2 (measured 0.27 -> 0.16 ms for
      // 64 x 2^20 with the XCD-contiguous tile order)
      if (q.type == PT_REAL4B) q.out_nt = 0;
      if (q.type == PT_REAL4C) q.in_nt = 0;
      // The same holds for row passes whose waves cover only part of a line of
      // each row per access (c2c rows of up to 128 points: a wave's access
      // spans 64 B of each of eight rows) and for real row passes, whose rows
      // of L+1 bins start off line boundaries: cached, the partial lines of
      // neighbouring accesses merge in L2 before they reach HBM (256 MiB per
      // batch: r2c 2^7 1.62 -> 4.21 TB/s, c2r 2^7..2^13 2.6-3.3 -> 4.1-4.5,
      // c2c 2^6 4.48 -> 5.03; c2c rows of 256 points and more keep streaming,
      // 2-7 % faster)
      if (q.type == PT_FFT && !q.colmap && !q.store_t && (q.L <= 128 || q.epi == EPI_R2C || q.epi == EPI_C2R))
        q.in_nt = q.out_nt = 0;
      q.xcd = P.var.xcd_order;
      q.mfma = 0;
      if (P.mode != afft_gpu_mode_fma && sizeof(T) == 4 && q.type == PT_FFT && q.L == 256 &&
          (q.epi == EPI_PLAIN || q.epi == EPI_TW4) && q.in.il && q.out.il) {
        q.mfma = P.mode;
        q.mfma_mats = get_table<T>(3, 256, P.mode == afft_gpu_mode_mfma_bf16 ? 1 : 0);
        if (!q.mfma_mats) ok = false;
      }
    }
    if (r && ok && P.var.tiny_rows > 0 && P.var.tiny_rows != 64 && P.var.tiny_rows != 128 &&
        P.var.tiny_rows != 256) {
      set_err("variant tiny_rows must be 64, 128 or 256 (rows per work-group), 0 (default) or -1");
      ok = false;
    }
    if (r && ok && P.var.pipeline > 0) {
      set_err("variant pipeline is no longer provided (pass A of one group beside pass B of the previous one "
              "measured slower on every shape); use pipeline = 0");
      ok = false;
    }
    if (r && ok) pick_fused();
    return r && ok;
  }

  // Batched 1-D FP32 c2c four-step plans whose two pass shapes have a fused
  // kernel run both passes in one persistent launch (afft_fused.hip) when the
  // variant asks for it.
  void pick_fused() {
    P.fused_slots = P.fused_lag = 0;
    if (P.var.fused > 0 && !kFusedEnabled) {
      set_err("the fused four-step launch is switched off: its first run did not finish (see the fused kernel's "
              "notes); use fused = 0");
      ok = false;
      return;
    }
    if (P.var.fused <= 0 || sizeof(T) != 4 || P.npass != 2 || P.mode != afft_gpu_mode_fma ||
        P.kind > 1 || P.ndim != 1 || P.format != afft_interleaved)
      return;
    const PassRec& a = P.pass[0];
    const PassRec& b = P.pass[1];
    if (a.type != PT_FFT || b.type != PT_FFT || !a.colmap || a.epi != EPI_TW4 || a.store_t || b.colmap ||
        !b.store_t || b.epi != EPI_PLAIN || a.in.role != R_SRC || a.out.role != R_SCR || b.in.role != R_SCR ||
        b.out.role != R_DST || a.mfma || b.mfma || a.n_blo != 1 || b.n_blo != 1 || a.fold_user || b.fold_user ||
        (idx_t)a.L * b.L != P.dims[0] || a.n_col != b.L || b.n_col != a.L)
      return;
    if (!fused_supported(a.L, a.CT, a.sv, a.hx, b.L, b.CT, b.sv, b.hx)) return;
    P.fused_slots = P.var.fused_slots > 0 ? P.var.fused_slots : 3;
    P.fused_lag = P.var.fused_lag > 0 ? P.var.fused_lag : 1;
    if (P.fused_lag >= P.fused_slots) P.fused_lag = P.fused_slots - 1;
  }

  bool build_passes(int kind, int format, const idx_t* d, int nd) {
    const int swap = (kind == 1) ? 1 : 0;
    idx_t total = 1;
    for (int i = 0; i < nd; i++) total *= d[i];
    Side usr_src, usr_dst;
    usr_src.role = R_SRC;
    usr_dst.role = R_DST;
    usr_src.il = usr_dst.il = format == afft_interleaved;

    if (kind == 0 || kind == 1) {
      all_axes(usr_src, usr_dst, d, nd, swap, swap);
      return ok;
    }
    // real transforms
    if (total == 1) {
      PassRec* r = add();
      if (!r) return false;
      r->type = PT_REALONE;
      r->in = usr_src;
      r->out = usr_dst;
      r->rr_inv = kind == 3;
      return ok;
    }
    const idx_t n0 = d[0], h = n0 / 2;
    idx_t inner = 1;
    for (int i = 1; i < nd; i++) inner *= d[i];
    if (kind == 2) {
      if (nd == 1) {
        Side in = usr_src;
        in.il = 1;  // N reals read as N/2 interleaved complex
        in.s_u = h;
        Side out = usr_dst;
        out.s_u = h + 1;
        // beyond the full-LDS rows the real-input four-step comes first (FP32
        // 2^16 x 512: 0.117-0.128 ms against 0.118 for one half-LDS row pass)
        if (h > max_row_len_full<T>() && r2c_real4(in, out, h)) return ok;
        if (h <= row_max()) {
          PassRec* r = add();
          if (!r || !fft_pass(r, (int)h, 0, EPI_R2C, 0)) return false;
          r->in = in;
          r->out = out;
          r->in.s_j = r->out.s_j = 1;
          r->in.s_c = h;
          r->out.s_c = h + 1;
          r->n_col = 1;
          r->fold_user = 1;
          tiny(r);
          return ok;
        }
        if (r2c_real4(in, out, h)) return ok;
        const idx_t n2 = r2c_tr_row_len(h);
        if (n2 >= 2 && h / n2 >= 32 && h <= (idx_t)max_col_len<T>() * 4096) {
          r2c_fourstep_tr(in, out, h);
          return ok;
        }
        // short column pass (FP64 h = 16384), or h beyond two passes (the
        // three-pass four-step): the c2c with a transposed store, then the
        // real pass in place
        axis(in, out, 1, h, 1, 0, 0);
        Side rin = out, rout = out;
        rin.s_j = rout.s_j = 1;
        rin.s_b = rout.s_b = h + 1;
        real_rows(rin, rout, h, 1, 0);
        return ok;
      }
      if (h > kColDirectMax) {
        // the real pass over row pairs runs inside pass B of the dim-0 column
        // four-step (real_dim0_fourstep_r2c; 8192 x 1024: 0.048 ms, 16384 x
        // 4096: 0.337 ms) or, on request (real_pair = 1, 2-D with a compiled
        // row length), inside the row FFTs (0.056 and 0.375 ms)
        const idx_t f1 = (idx_t)1 << (ilog2(h) / 2);
        const bool col_ok = real_colpair_ct((int)sizeof(T), (int)(h / f1), inner, (int)f1) > 0;
        const bool fuse = nd == 2 && P.var.real_pair >= 0 && (P.var.real_pair > 0 || !col_ok) &&
                          real_rowfft_supported((int)sizeof(T), h, (int)inner, P.var.real_pair > 0);
        real_dim0_fourstep_r2c(usr_src, usr_dst, n0, inner, !fuse);
        if (fuse) {
          PassRec* r = add();
          if (!r) return false;
          r->type = PT_REALROWFFT;
          r->L = (int)inner;
          r->CT = 2;
          r->hx = 1;
          r->in = usr_dst;
          r->in.s_j = 1;
          r->in.s_c = inner;
          r->in.s_b = r->in.s_u = (h + 1) * inner;
          r->out = r->in;
          r->rr_h = h;
          r->rr_m = inner;
          r->stage_tw = get_table<T>(0, inner, 0);
          r->real_tw = get_table<T>(2, h, 0);
          if (!r->stage_tw || !r->real_tw) ok = false;
          return ok;
        }
        inner_axes_on_dst(usr_dst, d, nd, h);
        return ok;
      }
      PassRec* r = add();
      if (!r || !fft_pass(r, (int)h, 1, EPI_R2C, 0)) return false;
      r->in = usr_src;
      r->in.il = 0;
      r->in.one_array = 1;
      r->in.re_off = 0;
      r->in.im_off = inner;
      r->in.s_j = 2 * inner;
      r->in.s_c = 1;
      r->in.s_b = r->in.s_u = n0 * inner;
      r->out = usr_dst;
      r->out.s_j = inner;
      r->out.s_c = 1;
      r->out.s_b = r->out.s_u = (h + 1) * inner;
      r->n_col = inner;
      r->b_per_user = 1;
      // inner c2c on each of the h+1 rows, in place on dst
      idx_t dd[kMaxDim];
      dd[0] = h + 1;
      for (int i = 1; i < nd; i++) dd[i] = d[i];
      Side o = usr_dst;
      o.s_u = (h + 1) * inner;
      for (int a = nd - 1; a >= 1; a--) {
        idx_t outer = 1, inn = 1;
        for (int i = 0; i < a; i++) outer *= dd[i];
        for (int i = a + 1; i < nd; i++) inn *= dd[i];
        axis(o, o, outer, dd[a], inn, 0, 0);
      }
      return ok;
    }
    // kind == 3: c2r
    if (nd == 1) {
      Side in = usr_src;
      in.s_u = h + 1;
      Side out = usr_dst;
      out.il = 1;  // N reals written as N/2 interleaved complex
      out.s_u = h;
      // variant real_pair = 2 asks for the real-input four-step wherever it
      // has kernels; by default the half-LDS row pass holds 2^16 (FP32 2^16 x
      // 512: 0.113 ms against 0.155 for the real-input four-step)
      if (P.var.real_pair == 2 && h > max_row_len_full<T>() && c2r_real4(in, out, h)) return ok;
      if (h <= row_max()) {
        PassRec* r = add();
        if (!r || !fft_pass(r, (int)h, 0, EPI_C2R, 0)) return false;
        r->in = in;
        r->out = out;
        r->in.s_j = r->out.s_j = 1;
        r->in.s_c = h + 1;
        r->out.s_c = h;
        r->n_col = 1;
        r->fold_user = 1;
        r->swap_in = r->swap_out = 1;
        tiny(r);
        return ok;
      }
      if (c2r_real4(in, out, h)) return ok;
      // long FP32 c2r whose four-step pass A has a kernel with the inverse
      // real pre-pass fused into its load: pass A reads the spectrum, pairs
      // bins k and h-k itself and no real pass runs on its own
      if (P.var.real_pair > 0 && sizeof(T) == 4) {
        const int np0 = P.npass;
        axis(in, out, 1, h, 1, 1, 1);
        PassRec* a = ok && P.npass == np0 + 2 ? &P.pass[np0] : nullptr;
        if (a && a->type == PT_FFT && a->colmap && a->epi == EPI_TW4 && a->in.role == R_SRC) {
          // the same kernel selection with the fused epilogue; a shape or
          // variant without such a kernel keeps the unfused form
          PassRec f = *a;
          if (fft_pass(&f, a->L, 1, EPI_C2R4, 0)) {
            a->epi = EPI_C2R4;
            a->CT = f.CT;
            a->variant = f.variant;
            a->sv = f.sv;
            a->hx = f.hx;
            a->stage_tw = f.stage_tw;
            a->rr_h = h;
            a->real_tw = get_table<T>(2, h, 0);
            if (!a->real_tw) ok = false;
            return ok;
          }
        }
        P.npass = np0;
        ok = true;
        g_err.clear();
      }
      Side rin = in, rout = out;
      rin.s_j = rout.s_j = 1;
      rin.s_b = h + 1;
      rout.s_b = h;
      real_rows(rin, rout, h, 1, 1);
      axis(out, out, 1, h, 1, 1, 1);
      return ok;
    }
    idx_t dd[kMaxDim];
    dd[0] = h + 1;
    for (int i = 1; i < nd; i++) dd[i] = d[i];
    Side in = usr_src;
    in.s_u = (h + 1) * inner;
    Side scr;
    scratch_side(scr, R_SCR);
    need_scratch(R_SCR, (h + 1) * inner);
    Side cur = in;
    // long dim 0: the inverse real pass over row pairs runs inside pass A of
    // the dim-0 column four-step (real_dim0_fourstep_c2r; 16384 x 2048: 0.167
    // ms, 16384 x 4096: 0.371 ms) or, on request (real_pair = 1, 2-D with a
    // compiled row length) or where that has no kernel, inside the inverse row
    // FFTs (0.213 and 0.418 ms; 8192 x 1024: 0.053 against 0.055 ms)
    const idx_t f1 = h > 1 ? (idx_t)1 << (ilog2(h) / 2) : 1;
    const bool col_ok = real_colpair_ct((int)sizeof(T), (int)f1, inner, (int)(h / f1)) > 0;
    if (h > kColDirectMax && nd == 2 && P.var.real_pair >= 0 && (P.var.real_pair > 0 || !col_ok) &&
        real_rowfft_supported((int)sizeof(T), h, (int)inner, P.var.real_pair > 0)) {
      PassRec* r = add();
      if (!r) return false;
      r->type = PT_REALIROWFFT;
      r->L = (int)inner;
      r->CT = 2;
      r->hx = 1;
      r->in = in;
      r->in.s_j = 1;
      r->in.s_c = inner;
      r->in.s_b = in.s_u;
      r->out = scr;
      r->out.s_j = 1;
      r->out.s_c = inner;
      r->out.s_b = (h + 1) * inner;
      r->rr_h = h;
      r->rr_m = inner;
      r->rr_inv = 1;
      r->stage_tw = get_table<T>(0, inner, 0);
      r->real_tw = get_table<T>(2, h, 0);
      if (!r->stage_tw || !r->real_tw) ok = false;
      real_dim0_fourstep_c2r(scr, usr_dst, n0, inner, false);
      return ok;
    }
    for (int a = nd - 1; a >= 1; a--) {
      idx_t outer = 1, inn = 1;
      for (int i = 0; i < a; i++) outer *= dd[i];
      for (int i = a + 1; i < nd; i++) inn *= dd[i];
      axis(cur, scr, outer, dd[a], inn, 1, 1);
      cur = scr;
    }
    if (h > kColDirectMax) {
      real_dim0_fourstep_c2r(cur, usr_dst, n0, inner);
      return ok;
    }
    PassRec* r = add();
    if (!r || !fft_pass(r, (int)h, 1, EPI_C2R, 0)) return false;
    r->in = scr;
    r->in.s_j = inner;
    r->in.s_c = 1;
    r->in.s_b = (h + 1) * inner;
    r->out = usr_dst;
    r->out.il = 0;
    r->out.one_array = 1;
    r->out.re_off = 0;
    r->out.im_off = inner;
    r->out.s_j = 2 * inner;
    r->out.s_c = 1;
    r->out.s_b = r->out.s_u = n0 * inner;
    r->n_col = inner;
    r->b_per_user = 1;
    r->swap_in = r->swap_out = 1;
    return ok;
  }
};

// ---- execution --------------------------------------------------------------
template <typename T>
struct Bufs {
  const T* src_re;
  const T* src_im;
  T* dst_re;
  T* dst_im;
  T* scr[2];
};

template <typename T>
void resolve(const Side& s, const Bufs<T>& B, idx_t start, const T*& re, const T*& im) {
  const T* bre = nullptr;
  const T* bim = nullptr;
  idx_t uoff = 0;
  switch (s.role) {
    case R_SRC:
      bre = B.src_re;
      bim = B.src_im;
      uoff = start * s.s_u;
      break;
    case R_DST:
      bre = B.dst_re;
      bim = B.dst_im;
      uoff = start * s.s_u;
      break;
    default:
      bre = B.scr[s.role - R_SCR];
      bim = nullptr;
      break;
  }
  if (s.il) {
    re = bre + 2 * (s.re_off + uoff);
    im = nullptr;
  } else if (s.one_array) {
    re = bre + s.re_off + uoff;
    im = bre + s.im_off + uoff;
  } else {
    re = bre + s.re_off + uoff;
    im = bim + s.im_off + uoff;
  }
}

// Kernel arguments of an FFT pass over user transforms [start, start+cnt);
// log2 N where the kernels compute the pass's four-step twiddle W_N instead of
// reading its tables (FP32, power-of-two N <= 2^27, whose folded index N/8
// is exact in a float), else 0.  64 x 2^20 pass A: 0.285 -> 0.22 ms per
// direction with both table loads per element gone.
template <typename T>
int tw4_log2n(const PassRec& r) {
  const idx_t n = r.tw4_n;
  return (sizeof(T) == 4 && n >= 8 && n <= ((idx_t)1 << 27) && (n & (n - 1)) == 0) ? ilog2(n) : 0;
}

// `fast` tells whether the buffer-resource path applies.
template <typename T>
bool fft_params(const PassRec& r, const Bufs<T>& B, idx_t start, idx_t cnt, PassParams<T>& p, bool& fast) {
  const T *ire, *iim, *ore_c, *oim_c;
  resolve(r.in, B, start, ire, iim);
  resolve(r.out, B, start, ore_c, oim_c);
  T* ore = const_cast<T*>(ore_c);
  T* oim = const_cast<T*>(oim_c);
  p = PassParams<T>{};
  p.in_re = ire;
  p.in_im = iim;
  p.out_re = ore;
  p.out_im = oim;
  p.in_js = r.in.s_j;
  p.in_cs = r.in.s_c;
  p.in_bs = r.in.s_b;
  p.in_bs2 = r.in.s_b2;
  p.out_ks = r.out.s_j;
  p.out_cs = r.out.s_c;
  p.out_bs = r.out.s_b;
  p.out_bs2 = r.out.s_b2;
  p.n_blo = r.n_blo;
  idx_t ncol = r.fold_user ? r.n_col * cnt : r.n_col;
  idx_t nb = r.fold_user ? 1 : r.b_per_user * cnt;
  idx_t tiles = (ncol + r.CT - 1) / r.CT;
  if (ncol > 0x7fffffff || tiles * nb > 0x7fffffff) {
    set_err("launch too large");
    return false;
  }
  p.n_col = (int)ncol;
  p.col_tiles = (int)tiles;
  p.n_tiles = (int)(tiles * nb);
  p.in_il = r.in.il;
  p.out_il = r.out.il;
  p.swap_in = r.swap_in;
  p.swap_out = r.swap_out;
  p.in_nt = r.in_nt;
  p.out_nt = r.out_nt;
  p.xcd_remap = r.xcd;
  if (r.epi == EPI_C2R4 && tiles >= 2 && tiles % 2 == 0 && !r.xcd) {
    // mirror tiles paired on one XCD (see PassParams::mirror_pairs)
    p.mirror_pairs = 1;
    p.n_batch = (int)nb;
    p.n_tiles = (int)((nb * (tiles / 2) + 7) / 8 * 16);
  }
  p.stage_tw = (const cx<T>*)r.stage_tw;
  p.tw4_hi = (const cx<T>*)r.tw4_hi;
  p.tw4_lo = (const cx<T>*)r.tw4_lo;
  p.tw4_lo_bits = r.tw4_bits;
  p.tw4_col_shift = r.tw4_shift;
  p.tw4_blo = r.tw4_blo;
  p.tw4_log2n = (r.epi == EPI_TW4 || r.epi == EPI_C2R4) ? tw4_log2n<T>(r) : 0;
  p.real_tw = (const cx<T>*)r.real_tw;
  // byte extent touched inside one batch entry (both sides); the buffer
  // path needs it below 4 GiB
  const idx_t lin = (idx_t)r.L + (r.epi == EPI_C2R ? 1 : 0);
  const idx_t lout = (idx_t)r.L + (r.epi == EPI_R2C ? 1 : 0);
  idx_t ext_in = ((ncol - 1) * r.in.s_c + (lin - 1) * r.in.s_j + 1) * (idx_t)sizeof(cx<T>);
  if (r.epi == EPI_C2R4) {
    // the paired load reaches bin h of each batch entry
    p.real_h = r.rr_h;
    ext_in = std::max(ext_in, (r.rr_h + 1) * (idx_t)sizeof(cx<T>));
  }
  const idx_t ext_out = ((ncol - 1) * r.out.s_c + (lout - 1) * r.out.s_j + 1) * (idx_t)sizeof(cx<T>);
  const bool fast_ok = r.in.il && r.out.il && ext_in > 0 && ext_out > 0 && ext_in < ((idx_t)1 << 32) &&
                       ext_out < ((idx_t)1 << 32);
  p.in_bytes = fast_ok ? (unsigned)ext_in : 0u;
  p.out_bytes = fast_ok ? (unsigned)ext_out : 0u;
  fast = fast_ok;
  return true;
}

template <typename T>
bool run_pass(const PassRec& r, const Bufs<T>& B, idx_t start, idx_t cnt, hipStream_t st) {
  const T *ire, *iim, *ore_c, *oim_c;
  resolve(r.in, B, start, ire, iim);
  resolve(r.out, B, start, ore_c, oim_c);
  T* ore = const_cast<T*>(ore_c);
  T* oim = const_cast<T*>(oim_c);
  if (r.type == PT_FFT && r.mfma) {
    MfmaParams m{};
    m.in = (const float*)ire;
    m.out = (float*)ore;
    m.in_js = r.in.s_j;
    m.in_cs = r.in.s_c;
    m.in_bs = r.in.s_b;
    m.in_bs2 = r.in.s_b2;
    m.out_ks = r.out.s_j;
    m.out_cs = r.out.s_c;
    m.out_bs = r.out.s_b;
    m.out_bs2 = r.out.s_b2;
    m.n_blo = r.n_blo;
    idx_t ncol = r.fold_user ? r.n_col * cnt : r.n_col;
    idx_t nb = r.fold_user ? 1 : r.b_per_user * cnt;
    idx_t tiles = (ncol + 15) / 16;
    if (ncol > 0x7fffffff || tiles * nb > 0x7fffffff) {
      set_err("launch too large");
      return false;
    }
    m.n_col = (int)ncol;
    m.col_tiles = (int)tiles;
    m.swap_in = r.swap_in;
    m.swap_out = r.swap_out;
    m.in_cfast = r.in.s_c == 1;
    m.out_cfast = r.out.s_c == 1;
    m.in_nt = r.in_nt;
    m.out_nt = r.out_nt;
    m.mats = (const h8*)r.mfma_mats;
    m.bf16 = r.mfma == afft_gpu_mode_mfma_bf16;
    if (r.epi == EPI_TW4) {
      m.tw4_hi = (const cx<float>*)r.tw4_hi;
      m.tw4_lo = (const cx<float>*)r.tw4_lo;
      m.tw4_bits = r.tw4_bits;
      m.tw4_shift = r.tw4_shift;
      m.tw4_log2n = tw4_log2n<float>(r);
    }
    HIP_OK(launch_mfma_dft256(m, (int)(tiles * nb), st));
    return true;
  }
  if (r.type == PT_FFT) {
    const PassVariant<T>* v = (const PassVariant<T>*)r.variant;
    PassParams<T> p;
    bool fast = false;
    if (!fft_params<T>(r, B, start, cnt, p, fast)) return false;
    HIP_OK(v->launch(p, st, fast));
    return true;
  }
  if (r.type == PT_REALROWS) {
    RealRowsArgs<T> a{};
    a.in_re = ire;
    a.in_im = iim;
    a.out_re = ore;
    a.out_im = oim;
    a.in_il = r.in.il;
    a.out_il = r.out.il;
    a.in_rs = r.in.s_j;
    a.in_cs = r.in.s_c ? r.in.s_c : 1;
    a.in_bs = r.in.s_b;
    a.out_rs = r.out.s_j;
    a.out_cs = r.out.s_c ? r.out.s_c : 1;
    a.out_bs = r.out.s_b;
    a.h = r.rr_h;
    a.m = r.rr_m;
    a.nbatch = cnt;
    a.real_tw = (const cx<T>*)r.real_tw;
    HIP_OK(real_rows<T>(a, r.rr_inv != 0, st));
    return true;
  }
  if (r.type == PT_REALIROWFFT) {
    RealIRowFftArgs<T> a{};
    a.in_re = ire;
    a.in_im = iim;
    a.in_il = r.in.il;
    a.out = reinterpret_cast<cx<T>*>(ore);
    a.L = r.L;
    a.h = r.rr_h;
    a.in_rs = r.in.s_c;
    a.in_bs = r.in.s_b;
    a.out_rs = r.out.s_c;
    a.out_bs = r.out.s_b;
    a.nbatch = cnt;
    a.stage_tw = (const cx<T>*)r.stage_tw;
    a.real_tw = (const cx<T>*)r.real_tw;
    HIP_OK(real_irowfft<T>(a, st));
    return true;
  }
  if (r.type == PT_REALROWFFT) {
    RealRowFftArgs<T> a{};
    a.re = ore;
    a.im = oim;
    a.il = r.out.il;
    a.L = r.L;
    a.h = r.rr_h;
    a.rs = r.out.s_c;
    a.bs = r.out.s_b;
    a.nbatch = cnt;
    a.stage_tw = (const cx<T>*)r.stage_tw;
    a.real_tw = (const cx<T>*)r.real_tw;
    HIP_OK(real_rowfft<T>(a, st));
    return true;
  }
  if (r.type == PT_REALPAIR) {
    RealPairArgs<T> a{};
    a.in = reinterpret_cast<const cx<T>*>(ire);
    a.out_re = ore;
    a.out_im = oim;
    a.out_il = r.out.il;
    a.out_nt = r.out_nt;
    a.in_bs = r.in.s_b;
    a.out_bs = r.out.s_b;
    a.n1 = (int)r.rr_m;
    a.n2 = r.L;
    a.nbatch = cnt;
    a.stage_tw = (const cx<T>*)r.stage_tw;
    a.real_tw = (const cx<T>*)r.real_tw;
    a.row_tw = (const cx<T>*)r.tw4_hi;
    HIP_OK(real_pair<T>(a, st));
    return true;
  }
  if (r.type == PT_REALICOLPAIR) {
    RealIColPairArgs<T> a{};
    a.in = reinterpret_cast<const cx<T>*>(ire);
    a.out = reinterpret_cast<cx<T>*>(ore);
    a.L = r.L;
    a.h2 = r.n_blo;
    a.m = r.rr_m;
    a.in_bs = r.in.s_b;
    a.out_bs = r.out.s_b;
    a.nbatch = cnt;
    a.stage_tw = (const cx<T>*)r.stage_tw;
    a.real_tw = (const cx<T>*)r.real_tw;
    a.tw4_hi = (const cx<T>*)r.tw4_hi;
    a.tw4_lo = (const cx<T>*)r.tw4_lo;
    a.tw4_bits = r.tw4_bits;
    a.tw4_log2n = tw4_log2n<T>(r);
    HIP_OK(real_icolpair<T>(a, st));
    return true;
  }
  if (r.type == PT_REALCOLPAIR) {
    RealColPairArgs<T> a{};
    a.in = reinterpret_cast<const cx<T>*>(ire);
    a.out_re = ore;
    a.out_im = oim;
    a.out_il = r.out.il;
    a.L = r.L;
    a.h1 = r.n_blo;
    a.m = r.rr_m;
    a.in_bs = r.in.s_b;
    a.out_bs = r.out.s_b;
    a.nbatch = cnt;
    a.stage_tw = (const cx<T>*)r.stage_tw;
    a.real_tw = (const cx<T>*)r.real_tw;
    HIP_OK(real_colpair<T>(a, st));
    return true;
  }
  if (r.type == PT_TINY) {
    TinyArgs<T> a{};
    a.in_re = ire;
    a.in_im = iim;
    a.in_il = r.in.il;
    a.in_nt = r.in_nt;
    a.out_re = ore;
    a.out_im = oim;
    a.out_il = r.out.il;
    a.out_nt = r.out_nt;
    a.rows = r.n_col * cnt;
    a.in_rs = r.in.s_c;
    a.out_rs = r.out.s_c;
    a.swap_in = r.swap_in;
    a.swap_out = r.swap_out;
    HIP_OK(tiny_rows<T>(a, r.L, r.epi, r.CT, st));
    return true;
  }
  if (r.type == PT_REAL4C || r.type == PT_REAL4D) {
    Real4Args<T> a{};
    const bool c = r.type == PT_REAL4C;
    a.ct = r.CT;
    a.sv = r.sv;
    a.hx = r.hx;
    a.n1 = c ? r.rr_m : r.L;
    a.n2 = c ? r.L : r.rr_m;
    a.nbatch = cnt;
    a.stage_tw = (const cx<T>*)r.stage_tw;
    a.rs = c ? r.out.s_j : r.in.s_c / 2;
    a.scr_bs = a.n2 * a.rs;
    if (c) {
      a.in = ire;
      a.in_im = iim;
      a.in_il = r.in.il;
      a.in_bs = r.in.s_b;
      a.in_nt = r.in_nt;
      a.scr = reinterpret_cast<cx<T>*>(ore);
      a.tw4_hi = (const cx<T>*)r.tw4_hi;
      a.tw4_lo = (const cx<T>*)r.tw4_lo;
      a.tw4_bits = r.tw4_bits;
      a.tw4_log2n = tw4_log2n<T>(r);
      HIP_OK(real4c<T>(a, st));
    } else {
      a.scr = const_cast<cx<T>*>(reinterpret_cast<const cx<T>*>(ire));
      a.out_re = ore;
      a.out_il = 1;
      a.out_bs = 2 * r.out.s_b;  // complex pairs -> reals
      a.out_nt = r.out_nt;
      HIP_OK(real4d<T>(a, st));
    }
    return true;
  }
  if (r.type == PT_REAL4A || r.type == PT_REAL4B) {
    Real4Args<T> a{};
    const idx_t N = 2 * r.rr_h;
    a.ct = r.CT;
    a.sv = r.sv;
    a.hx = r.hx;
    a.n1 = r.type == PT_REAL4A ? r.L : r.rr_m;
    a.n2 = r.type == PT_REAL4A ? r.rr_m : r.L;
    a.nbatch = cnt;
    a.stage_tw = (const cx<T>*)r.stage_tw;
    a.scr_bs = (a.n1 / 2 + 1) * a.n2;
    if (r.type == PT_REAL4A) {
      a.in = ire;
      a.in_bs = 2 * r.in.s_b;  // complex pairs -> reals
      a.scr = reinterpret_cast<cx<T>*>(ore);
      a.in_nt = r.in_nt;
      a.tw4_hi = (const cx<T>*)r.tw4_hi;
      a.tw4_lo = (const cx<T>*)r.tw4_lo;
      a.tw4_bits = r.tw4_bits;
      a.tw4_log2n = tw4_log2n<T>(r);
      (void)N;
      HIP_OK(real4a<T>(a, st));
    } else {
      a.scr = const_cast<cx<T>*>(reinterpret_cast<const cx<T>*>(ire));
      a.out_re = ore;
      a.out_im = oim;
      a.out_il = r.out.il;
      a.out_bs = r.out.s_b;
      a.out_nt = r.out_nt;
      HIP_OK(real4b<T>(a, st));
    }
    return true;
  }
  if (r.type == PT_REALTR) {
    RealTrArgs<T> a{};
    a.in = reinterpret_cast<const cx<T>*>(ire);
    a.out_re = ore;
    a.out_im = oim;
    a.out_il = r.out.il;
    a.in_bs = r.in.s_b;
    a.out_bs = r.out.s_b;
    a.n1 = (int)r.rr_m;
    a.n2 = (int)(r.rr_h / r.rr_m);
    a.h = r.rr_h;
    a.nbatch = cnt;
    a.real_tw = (const cx<T>*)r.real_tw;
    HIP_OK(real_tr<T>(a, st));
    return true;
  }
  // PT_REALONE
  if (!r.rr_inv) {
    HIP_OK(real_one<T>(ire, ore, oim, 0, r.out.il, cnt, 1, 1, st));
  } else {
    HIP_OK(real_one<T>(ire, ore, nullptr, 1, r.in.il, cnt, 1, 1, st));
  }
  return true;
}

// Launches per group for pass A of a four-step plan (variant a_split): pass A
// reads the caller's input into a scratch that holds the group's transforms
// back to back, and the later passes (pass B; for 1-D r2c also the
// transposing real pass) start from that scratch, so pass A can run the group
// in parts (each writing its part of the scratch) while the others run it
// whole.
bool split_shape(const Plan& P) {
  if (P.npass < 2) return false;
  const PassRec& a = P.pass[0];
  const PassRec& b = P.pass[1];
  if (a.type != PT_FFT || a.mfma || a.fold_user || a.in.role >= R_SCR || a.out.role < R_SCR) return false;
  if (a.b_per_user != 1 || a.n_blo > 1 || a.out.s_b != P.scr_per_user[a.out.role - R_SCR]) return false;
  return b.in.role == a.out.role && P.pass[P.npass - 1].out.role < R_SCR;
}

// Default (a_split = 0) for four-step plans whose pass A has long (>= 1024-
// point) columns: pass A in groups of 64 MiB of scratch (which stay in the
// MALL between its launches), pass B over four of them at once (fewer
// launches, several waves per launch).  64 x 2^20 c2c: 0.44 vs 0.48-0.51 ms
// per direction; 32 x 2^21: 0.52 vs 0.57.  With the transposing real pass
// after pass B (1-D r2c) only where a group holds at most 4 transforms:
// 16 x r2c 2^22 0.42 -> 0.40 ms, while 32 x r2c 2^21 measured 0.408 against
// 0.402 ms and keeps its groups of 8.
constexpr int kAutoSplit = 4;
bool auto_split(const Plan& P) {
  if (P.var.a_split != 0 || P.var.chunk != 0 || !split_shape(P) || !P.pass[0].colmap ||
      P.pass[0].L < 1024)
    return false;
  if (P.npass == 2) return true;
  const idx_t eb = P.prec == 32 ? 8 : 16;
  return P.npass == 3 && P.pass[2].type == PT_REALTR && P.scr_per_user[0] * eb >= ((idx_t)16 << 20);
}

int a_split(const Plan& P) {
  if (!split_shape(P)) return 1;
  if (P.var.a_split > 1) return P.var.a_split;
  if (!auto_split(P)) return 1;
  // pass A launches of one 64 MiB group each
  const idx_t per = P.scr_per_user[0] > P.scr_per_user[1] ? P.scr_per_user[0] : P.scr_per_user[1];
  const idx_t eb = P.prec == 32 ? 8 : 16;
  const idx_t g = per > 0 ? ((idx_t)64 << 20) / (per * eb) : 0;
  if (g < 1 || P.chunk % g != 0) return 1;
  return (int)(P.chunk / g);
}

// Bytes of the fused launch's scratch ring (counters follow it).
size_t fused_ring_bytes(const Plan& P) {
  return ((size_t)kFusedXcd * P.fused_slots * (size_t)P.dims[0] * sizeof(cx<float>) + 255) & ~(size_t)255;
}

// Scratch 1 starts at this offset into the plan's scratch (scratch 0 first).
size_t scratch1_offset(const Plan& P) { return (P.scr_bytes[0] + 255) & ~(size_t)255; }
size_t scratch_total(const Plan& P) { return P.scr_bytes[1] ? scratch1_offset(P) + P.scr_bytes[1] : P.scr_bytes[0]; }

// Launches passes [first, last) of the plan on `st` with its scratch at `scr`
// (scratch_total(P) bytes).
template <typename T>
bool execute_on(Plan& P, const void* sre, const void* sim, void* dre, void* dim, hipStream_t st, int first, int last,
                void* scr, idx_t chunk = 0, bool plain = false) {
  // chunk > 0: groups of that many transforms (scratch sized for it) and,
  // with plain, one launch per pass only (no fused schedule)
  if (last < 0) last = P.npass;
  if (chunk <= 0) chunk = P.chunk;
  // scratch 1 follows scratch 0 of the group (the pool layout also makes
  // room for the fused ring, which plain runs skip)
  const size_t scr1_off = !plain ? scratch1_offset(P)
                                 : ((size_t)P.scr_per_user[0] * (size_t)chunk * sizeof(cx<T>) + 255) & ~(size_t)255;
  Bufs<T> B{(const T*)sre, (const T*)sim, (T*)dre, (T*)dim,
            {(T*)scr, P.scr_bytes[1] ? (T*)((char*)scr + scr1_off) : nullptr}};
  if (P.format == afft_interleaved) {
    B.src_im = nullptr;
    B.dst_im = nullptr;
  }
  if (P.fused_slots && !plain && first == 0 && last == P.npass) {
    // both passes in one persistent launch over the whole batch
    const PassRec& a = P.pass[0];
    const PassRec& b = P.pass[1];
    FusedArgs f{};
    f.src = (const float*)B.src_re;
    f.dst = (float*)B.dst_re;
    f.ring = (float*)B.scr[0];
    f.src_stride = 2 * a.in.s_u;
    f.dst_stride = 2 * b.out.s_u;
    f.stage_a = (const cx<float>*)a.stage_tw;
    f.stage_b = (const cx<float>*)b.stage_tw;
    f.tw4_hi = (const cx<float>*)a.tw4_hi;
    f.tw4_lo = (const cx<float>*)a.tw4_lo;
    f.tw4_bits = a.tw4_bits;
    f.swap = a.swap_in;
    f.in_nt = a.in_nt;
    f.out_nt = b.out_nt;
    f.batch = (int)P.batch;
    f.slots = P.fused_slots;
    f.lag = P.fused_lag;
    f.ctr = (int*)((char*)B.scr[0] + fused_ring_bytes(P));
    f.spin_limit = 1LL << 26;
    HIP_OK(fused_launch(f, a.L, a.CT, a.sv, a.hx, b.L, b.CT, b.sv, b.hx, st));
    return true;
  }
  const int split = a_split(P);
  for (idx_t start = 0; start < P.batch; start += chunk) {
    idx_t cnt = P.batch - start < chunk ? P.batch - start : chunk;
    for (int i = first; i < last; i++) {
      if (i == 0 && split > 1 && cnt % split == 0) {
        // pass A of the group in `split` launches, each into its part of the
        // scratch (the scratch holds the group's transforms back to back)
        const idx_t sub = cnt / split;
        for (int q = 0; q < split; q++) {
          Bufs<T> Bq = B;
          const int role = P.pass[0].out.role - R_SCR;
          Bq.scr[role] = B.scr[role] + 2 * (size_t)(q * sub) * (size_t)P.scr_per_user[role];
          if (!run_pass<T>(P.pass[0], Bq, start + q * sub, sub, st)) return false;
        }
        continue;
      }
      if (!run_pass<T>(P.pass[i], B, start, cnt, st)) return false;
    }
  }
  return true;
}

// Bytes of each input array and each output array one execute touches
// (split format: per re / im array).
void io_bytes(const Plan& P, size_t elem, size_t& in_b, size_t& out_b) {
  idx_t n = 1, inner = 1;
  for (int i = 0; i < P.ndim; i++) n *= P.dims[i];
  for (int i = 1; i < P.ndim; i++) inner *= P.dims[i];
  const idx_t m = (P.dims[0] / 2 + 1) * inner;  // complex side of a real transform
  const idx_t c = P.format == afft_interleaved ? 2 : 1;
  idx_t in_s = c * n, out_s = c * n;
  if (P.kind == 2) {
    in_s = n;
    out_s = c * m;
  } else if (P.kind == 3) {
    in_s = c * m;
    out_s = n;
  }
  in_b = (size_t)(in_s * P.batch) * elem;
  out_b = (size_t)(out_s * P.batch) * elem;
}

bool overlaps(const void* a, size_t na, const void* b, size_t nb) {
  if (!a || !b) return false;
  const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
  return x < y + nb && y < x + na;
}

// Buffers may overlap: complex transforms run exactly in place (src == dst:
// every pass reads a tile completely before it writes the same tile, and the
// four-step passes go through scratch); any other overlap of input and output
// (real transforms in place, shifted buffers) first copies the input into
// scratch and runs from there.  Staging says how many input arrays to copy
// (0 = none) and their size.
struct Staging {
  int n = 0;          // arrays to copy (re; im too for split complex input)
  size_t bytes = 0;   // bytes of one array
  size_t stride = 0;  // bytes between the copies in scratch
};

template <typename T>
Staging staging(const Plan& P, const void* sre, const void* sim, const void* dre, const void* dim) {
  const bool il = P.format == afft_interleaved;
  size_t in_b = 0, out_b = 0;
  io_bytes(P, sizeof(T), in_b, out_b);
  const void* ins[2] = {sre, (il || P.kind == 2) ? nullptr : sim};
  const void* outs[2] = {dre, (il || P.kind == 3) ? nullptr : dim};
  bool ov = false;
  for (const void* a : ins)
    for (const void* b : outs) ov = ov || overlaps(a, in_b, b, out_b);
  const bool exact = P.kind <= 1 && sre == dre && (il || sim == dim);
  Staging sg;
  if (ov && !exact) {
    sg.n = ins[1] ? 2 : 1;
    sg.bytes = in_b;
    sg.stride = (in_b + 255) & ~(size_t)255;
  }
  return sg;
}

// Copies the input into scratch at `at` (sg.n arrays) and points sre / sim there.
bool stage_input(const Staging& sg, char* at, const void*& sre, const void*& sim, hipStream_t st) {
  if (sg.n == 0) return true;
  HIP_OK(hipMemcpyAsync(at, sre, sg.bytes, hipMemcpyDeviceToDevice, st));
  sre = at;
  if (sg.n == 2) {
    HIP_OK(hipMemcpyAsync(at + sg.stride, sim, sg.bytes, hipMemcpyDeviceToDevice, st));
    sim = at + sg.stride;
  }
  return true;
}

// Launches passes [first, last) of the plan on `st`, its scratch (and room
// for a staged input) leased from the device's pool.
template <typename T>
bool execute(Plan& P, const void* sre, const void* sim, void* dre, void* dim, hipStream_t st, int first = 0,
             int last = -1) {
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess && dev != P.device) HIP_OK(hipSetDevice(P.device));
  const Staging sg = staging<T>(P, sre, sim, dre, dim);
  const size_t base = scratch_total(P);
  const size_t stage_off = (base + 255) & ~(size_t)255;
  ScratchLease lease(P.device, sg.n ? stage_off + sg.n * sg.stride : base, st);
  if (!lease.ok()) return false;
  if (!stage_input(sg, (char*)lease.ptr() + stage_off, sre, sim, st)) return false;
  return execute_on<T>(P, sre, sim, dre, dim, st, first, last, lease.ptr());
}

// Scratch bytes one execute needs (taken from the pool when it runs).
void size_scratch(Plan& P, size_t elem_bytes) {
  for (int i = 0; i < 2; i++) P.scr_bytes[i] = (size_t)P.scr_per_user[i] * (size_t)P.chunk * elem_bytes;
  if (P.fused_slots)  // the ring and its counters in scratch 0 (pass ranges still run unfused)
    P.scr_bytes[0] = std::max(P.scr_bytes[0], fused_ring_bytes(P) + sizeof(int) * fused_counter_ints(P.batch));
}

idx_t default_chunk(const Plan& P, size_t elem_bytes) {
  idx_t per = P.scr_per_user[0] > P.scr_per_user[1] ? P.scr_per_user[0] : P.scr_per_user[1];
  if (per == 0) return P.batch;
  if (P.var.chunk > 0) return P.var.chunk < P.batch ? P.var.chunk : P.batch;
  // Four-step groups: 64 MiB of scratch when pass A has long (>= 1024-point)
  // columns; 256 MiB when they are short, and for the real-input four-step,
  // where fewer, larger launches measured faster than keeping the scratch
  // within the MALL
  const bool short_cols = P.npass >= 1 && ((P.pass[0].type == PT_FFT && P.pass[0].colmap && P.pass[0].L <= 512) ||
                                           P.pass[0].type == PT_REAL4A || P.pass[0].type == PT_REAL4C);
  idx_t target = (idx_t)((short_cols ? 256u : 64u) << 20) / (idx_t)(per * elem_bytes);
  // pass A split into launches of the 64 MiB groups, pass B over kAutoSplit of them
  const idx_t unit = !short_cols && auto_split(P) && target >= 1 ? target : 1;
  if (unit > 1) target *= kAutoSplit;
  if (target < 1) target = 1;
  if (target >= P.batch) return P.batch;
  // as many groups as the budget needs, of equal size (16 x r2c 2^22: 8 + 8
  // transforms instead of 15 + 1, whose single-transform launches fill a
  // fraction of the GPU), in whole pass-A launches where pass A is split
  const idx_t groups = (P.batch + target - 1) / target;
  const idx_t even = (P.batch + groups - 1) / groups;
  return std::min(target, (even + unit - 1) / unit * unit);
}

// Validates arguments; returns squeezed dims in d / nd.  status: 0 ok,
// 1 bad format, 2 other.
int validate(size_t ndim, const size_t* dim, int format, int impl, idx_t* d, int& nd) {
  if (format != afft_split && format != afft_interleaved) {
    set_err("invalid format %d", format);
    return 1;
  }
  if (impl != afft_auto && impl != afft_gpu) {
    set_err("implementation %d is not provided by this library", impl);
    return 2;
  }
  if (ndim > (size_t)kMaxDim) {
    set_err("ndim %zu exceeds %d", ndim, kMaxDim);
    return 2;
  }
  if (ndim > 0 && !dim) {
    set_err("null dim array");
    return 2;
  }
  nd = 0;
  for (size_t i = 0; i < ndim; i++) {
    if (!is_pow2((idx_t)dim[i])) {
      set_err("dimension %zu (%zu) is not a power of two", i, dim[i]);
      return 2;
    }
    if (dim[i] != 1) d[nd++] = (idx_t)dim[i];
  }
  if (nd == 0) d[nd++] = 1;
  return 0;
}

bool tcache_get(const Plan& P, Variant& v);  // defined with the autotuner below
void* noalloc_scratch(int dev);              // defined with afft_gpu_execute_noalloc below

// Bytes of `mem` a plan needs: the Plan itself at the first 64-byte boundary.
constexpr size_t kPlanMem = sizeof(Plan) + kPlanAlign - 1;

template <typename T>
void* create_impl(int kind, size_t ndim, const size_t* dim, int format, size_t batch, void* mem, int impl) {
  g_err.clear();
  idx_t d[kMaxDim];
  int nd = 0;
  if (validate(ndim, dim, format, impl, d, nd) != 0) return nullptr;
  if (!mem) {
    set_err("null mem");
    return nullptr;
  }
  if (batch < 1) {
    set_err("batch must be >= 1");
    return nullptr;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_err("no GPU visible");
    return nullptr;
  }
  Plan* P = new (plan_at(mem)) Plan();
  P->prec = sizeof(T) == 4 ? 32 : 64;
  P->kind = kind;
  P->format = format;
  P->ndim = nd;
  for (int i = 0; i < nd; i++) P->dims[i] = d[i];
  P->batch = (idx_t)batch;
  (void)hipGetDevice(&P->device);
  Variant cached;
  if (tcache_get(*P, cached)) P->var = cached;
  Builder<T> b(*P);
  if (!b.build(kind, format, d, nd)) {
    P->magic = 0;
    return nullptr;
  }
  P->chunk = default_chunk(*P, sizeof(cx<T>));
  size_scratch(*P, sizeof(cx<T>));
  // set-up of the no-allocation entry point: resolve this device's fixed
  // buffer now, so that its first call allocates nothing either (a failure
  // here only matters to that entry point, which reports it)
  if (!noalloc_scratch(P->device)) g_err.clear();
  P->magic = kMagic;
  return P;
}

// The plan's device resources are shared (twiddle cache, scratch pool), so
// destroying a plan only invalidates it; dropping its memory is as good.
void destroy_impl(void* state) {
  if (!state) return;
  Plan* P = reinterpret_cast<Plan*>(state);
  if (P->magic == kMagic) P->magic = 0;
}

// The plan size, or 0 when no plan of this shape can be built (a length
// beyond the compiled passes); checked by building the pass list on the host.
template <typename T>
size_t memsize_impl(int kind, size_t ndim, const size_t* dim, int format, int impl) {
  idx_t d[kMaxDim];
  int nd = 0;
  int st = validate(ndim, dim, format, impl, d, nd);
  if (st == 1) return (size_t)-1;
  if (st != 0) return 0;
  Plan* P = new (std::nothrow) Plan();
  if (!P) return 0;
  P->prec = sizeof(T) == 4 ? 32 : 64;
  P->kind = kind;
  P->format = format;
  P->ndim = nd;
  for (int i = 0; i < nd; i++) P->dims[i] = d[i];
  g_dry = true;
  const bool ok = Builder<T>(*P).build(kind, format, d, nd);
  g_dry = false;
  delete P;
  return ok ? kPlanMem : 0;
}

template <typename T>
void transform_impl(void* state, const void* sre, const void* sim, void* dre, void* dim, int kind) {
  g_err.clear();
  Plan* P = get_plan(state);
  if (!P) return;
  if (P->prec != (int)(sizeof(T) * 8) || P->kind != kind) {
    set_err("plan kind / precision mismatch");
    return;
  }
  if (!execute<T>(*P, sre, sim, dre, dim, P->stream)) return;
  hipError_t e = hipStreamSynchronize(P->stream);
  if (e != hipSuccess) set_err("transform: %s", hipGetErrorString(e));
}


// ---- variants and the autotuner ---------------------------------------------------
std::string plan_key(const Plan& P) {
  std::string k = std::to_string(P.prec) + "/" + std::to_string(P.kind) + "/" + std::to_string(P.format) + "/" +
                  std::to_string((long long)P.batch) + "/";
  for (int i = 0; i < P.ndim; i++) k += (i ? "x" : "") + std::to_string((long long)P.dims[i]);
  return k;
}

struct TuneCache {
  std::mutex mu;
  bool loaded = false;
  std::map<std::string, Variant> best;
};
TuneCache& tcache() {
  static TuneCache* c = new TuneCache;
  return *c;
}

// The cache file (AFFT_GPU_TUNE_CACHE) holds one line per tuned shape:
//   <key> <col_tile> <rowt_tile> <stream_io> <xcd_order> <chunk> <sched> [<half_lds> [<pipeline> [<real_pair>
//   [<a_split> [<row_log2> [<fused> [<fused_slots> [<fused_lag> [<tiny_rows>]]]]]]]]]
void tcache_load_locked(TuneCache& c) {
  if (c.loaded) return;
  c.loaded = true;
  const char* path = getenv("AFFT_GPU_TUNE_CACHE");
  if (!path) return;
  FILE* f = fopen(path, "r");
  if (!f) return;
  char line[1024], key[512];
  while (fgets(line, sizeof(line), f)) {
    Variant v;
    long long ch = 0;
    const int got = sscanf(line, "%511s %d %d %d %d %lld %d %d %d %d %d %d %d %d %d %d", key, &v.col_tile,
                           &v.rowt_tile, &v.stream_io, &v.xcd_order, &ch, &v.sched, &v.half_lds, &v.pipeline,
                           &v.real_pair, &v.a_split, &v.row_log2, &v.fused, &v.fused_slots, &v.fused_lag,
                           &v.tiny_rows);
    if (got < 7) continue;
    v.chunk = ch;
    c.best[key] = v;
  }
  fclose(f);
}

bool tcache_get(const Plan& P, Variant& v) {
  TuneCache& c = tcache();
  std::lock_guard<std::mutex> lk(c.mu);
  tcache_load_locked(c);
  auto it = c.best.find(plan_key(P));
  if (it == c.best.end()) return false;
  v = it->second;
  return true;
}

void tcache_put(const Plan& P, const Variant& v) {
  TuneCache& c = tcache();
  std::lock_guard<std::mutex> lk(c.mu);
  tcache_load_locked(c);
  c.best[plan_key(P)] = v;
  const char* path = getenv("AFFT_GPU_TUNE_CACHE");
  if (!path) return;
  FILE* f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "%s %d %d %d %d %lld %d %d %d %d %d %d %d %d %d %d\n", plan_key(P).c_str(), v.col_tile,
          v.rowt_tile, v.stream_io, v.xcd_order, (long long)v.chunk, v.sched, v.half_lds, v.pipeline, v.real_pair,
          v.a_split, v.row_log2, v.fused, v.fused_slots, v.fused_lag, v.tiny_rows);
  fclose(f);
}

template <typename T>
bool rebuild(Plan& P) {
  Builder<T> b(P);
  if (!b.build(P.kind, P.format, P.dims, P.ndim)) return false;
  P.chunk = default_chunk(P, sizeof(cx<T>));
  size_scratch(P, sizeof(cx<T>));
  return true;
}

bool rebuild_any(Plan& P) { return P.prec == 32 ? rebuild<float>(P) : rebuild<double>(P); }

template <typename T>
bool tune_impl(Plan& P, const void* sre, const void* sim, void* dre, void* dim, hipStream_t st, int iters,
               float* best_ms) {
  const Variant keep = P.var;
  std::vector<Variant> cands;
  idx_t base_chunk = 0;
  {
    P.var = Variant();
    if (!rebuild<T>(P)) return false;
    base_chunk = P.chunk;
  }
  // streaming I/O forced on and off, and the XCD-contiguous tile order, on the
  // default shape (each moved N-D transforms by a few per cent either way)
  for (int sx : {0, 1, 2}) {
    Variant v;
    if (sx < 2) v.stream_io = sx;
    else v.xcd_order = 1;
    cands.push_back(v);
  }
  for (int col : {0, 8, 32})
    for (int rowt : {0, 4, 16})
      for (int hl : {0, -1})
        for (int sch : {0, -1})
          for (int cm : {1, 2, 4}) {
            Variant v;
            v.sched = sch;
            v.col_tile = col;
            v.rowt_tile = rowt;
            v.half_lds = hl;
            v.chunk = cm == 1 ? 0 : (cm == 2 ? base_chunk / 2 : base_chunk * 2);
            if (cm != 1 && (v.chunk < 1 || v.chunk > P.batch || base_chunk >= P.batch)) continue;
            cands.push_back(v);
          }
  // plans with a real pass that has a fused form also try the other form
  const bool long_1d_real =
      sizeof(T) == 4 && P.kind >= 2 && P.ndim == 1 && P.dims[0] / 2 > (idx_t)max_row_len<T>();
  if (long_1d_real) {
    // 1-D FP32 real: the real-input four-step (with its wider column tiles)
    // and the complex four-step with the real pass on its own or fused
    for (int rp : {2, -1, 1}) {
      Variant v;
      v.real_pair = rp;
      cands.push_back(v);
      if (rp == 2)
        for (int ct : {8, 16}) {
          v.col_tile = ct;
          cands.push_back(v);
        }
    }
  }
  for (int i = 0; i < P.npass && !long_1d_real; i++) {
    if (P.pass[i].type == PT_REALPAIR || P.pass[i].type == PT_REALTR || P.pass[i].type == PT_REALROWFFT ||
        P.pass[i].type == PT_REALIROWFFT || P.pass[i].type == PT_REALCOLPAIR ||
        P.pass[i].type == PT_REALICOLPAIR) {
      Variant v;
      v.real_pair = P.pass[i].type == PT_REALTR ? 1 : -1;
      cands.push_back(v);
      if (P.pass[i].type == PT_REALCOLPAIR || P.pass[i].type == PT_REALICOLPAIR) {
        // and the row-FFT form (2-D with 1024..8192-point rows)
        Variant w;
        w.real_pair = 1;
        cands.push_back(w);
      }
      break;
    }
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return false;
  float best = 1e30f;
  Variant bestv = keep;
  for (const Variant& v : cands) {
    P.var = v;
    if (!rebuild<T>(P)) continue;
    if (!execute<T>(P, sre, sim, dre, dim, st)) continue;
    (void)hipEventRecord(e0, st);
    for (int i = 0; i < iters; i++) execute<T>(P, sre, sim, dre, dim, st);
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ms /= iters;
    if (ms < best) {
      best = ms;
      bestv = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  P.var = bestv;
  if (!rebuild<T>(P)) return false;
  if (best_ms) *best_ms = best;
  tcache_put(P, bestv);
  return best < 1e29f;
}

}  // namespace

// ---- stage schedule shared with the twiddle generator --------------------------
namespace afgpu {
// entry points for afft_queue.hip
const void* queue_table(int type, idx_t n, int bits) { return get_table<float>(type, n, bits); }
void queue_set_err(const char* msg) { set_err("%s", msg); }

template <int L, int SV>
static int fill_stages(int* radix, int* ns) {
  int acc = 1;
  for (int s = 0; s < Sched<L, SV>::NST; s++) {
    radix[s] = Sched<L, SV>::R[s];
    ns[s] = acc;
    acc *= Sched<L, SV>::R[s];
  }
  return Sched<L, SV>::NST;
}
int stage_list(int L, int sv, int* radix, int* ns) {
  if (sv == 1) {
    switch (L) {
      case 512: return fill_stages<512, 1>(radix, ns);
      case 1024: return fill_stages<1024, 1>(radix, ns);
      case 2048: return fill_stages<2048, 1>(radix, ns);
      default: return 0;
    }
  }
  switch (L) {
    case 1: return fill_stages<1, 0>(radix, ns);
    case 2: return fill_stages<2, 0>(radix, ns);
    case 4: return fill_stages<4, 0>(radix, ns);
    case 8: return fill_stages<8, 0>(radix, ns);
    case 16: return fill_stages<16, 0>(radix, ns);
    case 32: return fill_stages<32, 0>(radix, ns);
    case 64: return fill_stages<64, 0>(radix, ns);
    case 128: return fill_stages<128, 0>(radix, ns);
    case 256: return fill_stages<256, 0>(radix, ns);
    case 512: return fill_stages<512, 0>(radix, ns);
    case 1024: return fill_stages<1024, 0>(radix, ns);
    case 2048: return fill_stages<2048, 0>(radix, ns);
    case 4096: return fill_stages<4096, 0>(radix, ns);
    case 8192: return fill_stages<8192, 0>(radix, ns);
    case 16384: return fill_stages<16384, 0>(radix, ns);
    case 32768: return fill_stages<32768, 0>(radix, ns);
    default: return 0;
  }
}
}  // namespace afgpu

namespace {
// One line describing pass i of a plan (kernel shape and buffer roles).
int describe_pass(const Plan& P, int i, char* buf, size_t len) {
  if (i < 0 || i >= P.npass || !buf) return -1;
  const PassRec& r = P.pass[i];
  static const char* ep[] = {"plain", "tw4", "r2c", "c2r", "c2r+tw4"};
  static_assert(sizeof(ep) / sizeof(ep[0]) == EPI_C2R4 + 1, "one name per epilogue");
  if (r.type == PT_FFT)
    snprintf(buf, len,
             "fft%s L=%d CT=%d%s%s %s epi=%s%s ncol=%lld bpu=%lld blo=%d in(role=%d il=%d) out(role=%d il=%d)",
             r.mfma ? (r.mfma == afft_gpu_mode_mfma_bf16 ? "[mfma-bf16]" : "[mfma]")
                    : (P.fused_slots ? "[fused4]" : ""),
             r.L, r.CT, r.sv ? " sv1" : "",
             r.hx ? " hx" : "", r.colmap ? "col" : "row", ep[r.epi],
             r.store_t ? " store_t" : "", (long long)r.n_col,
             (long long)r.b_per_user, r.n_blo, r.in.role, r.in.il, r.out.role, r.out.il);
  else if (r.type == PT_REALTR)
    snprintf(buf, len, "real_tr h=%lld n1=%lld", (long long)r.rr_h, (long long)r.rr_m);
  else if (r.type == PT_REALIROWFFT)
    snprintf(buf, len, "real_irowfft L=%d h=%lld in(role=%d il=%d)", r.L, (long long)r.rr_h, r.in.role, r.in.il);
  else if (r.type == PT_REALROWFFT)
    snprintf(buf, len, "real_rowfft L=%d h=%lld out(role=%d il=%d)", r.L, (long long)r.rr_h, r.out.role, r.out.il);
  else if (r.type == PT_REALPAIR)
    snprintf(buf, len, "real_pair L=%d rows=%d h=%lld n1=%lld out(role=%d il=%d)", r.L, r.CT, (long long)r.rr_h,
             (long long)r.rr_m, r.out.role, r.out.il);
  else if (r.type == PT_REALROWS)
    snprintf(buf, len, "real_rows h=%lld m=%lld inv=%d", (long long)r.rr_h, (long long)r.rr_m, r.rr_inv);
  else if (r.type == PT_REALICOLPAIR)
    snprintf(buf, len, "real_icolpair L=%d CT=%d h2=%d h=%lld m=%lld in(role=%d il=%d)", r.L, r.CT, r.n_blo,
             (long long)r.rr_h, (long long)r.rr_m, r.in.role, r.in.il);
  else if (r.type == PT_REAL4A)
    snprintf(buf, len, "real4a n1=%d CT=%d sv%d %s n2=%lld in(role=%d) out(role=%d)", r.L, r.CT, r.sv, r.hx ? "hx" : "lds", (long long)r.rr_m,
             r.in.role, r.out.role);
  else if (r.type == PT_REAL4B)
    snprintf(buf, len, "real4b n2=%d CT=%d sv%d %s n1=%lld in(role=%d) out(role=%d il=%d)", r.L, r.CT, r.sv, r.hx ? "hx" : "lds", (long long)r.rr_m,
             r.in.role, r.out.role, r.out.il);
  else if (r.type == PT_REAL4C)
    snprintf(buf, len, "real4c n2=%d CT=%d sv%d %s n1=%lld in(role=%d il=%d) out(role=%d)", r.L, r.CT, r.sv, r.hx ? "hx" : "lds", (long long)r.rr_m,
             r.in.role, r.in.il, r.out.role);
  else if (r.type == PT_REAL4D)
    snprintf(buf, len, "real4d n1=%d CT=%d sv%d %s n2=%lld in(role=%d) out(role=%d)", r.L, r.CT, r.sv, r.hx ? "hx" : "lds", (long long)r.rr_m,
             r.in.role, r.out.role);
  else if (r.type == PT_TINY)
    snprintf(buf, len, "tiny L=%d rows=%d epi=%s ncol=%lld in(role=%d il=%d) out(role=%d il=%d)", r.L, r.CT,
             ep[r.epi], (long long)r.n_col, r.in.role, r.in.il, r.out.role, r.out.il);
  else if (r.type == PT_REALCOLPAIR)
    snprintf(buf, len, "real_colpair L=%d CT=%d h1=%d h=%lld m=%lld out(role=%d il=%d)", r.L, r.CT, r.n_blo,
             (long long)r.rr_h, (long long)r.rr_m, r.out.role, r.out.il);
  else
    snprintf(buf, len, "real_one inv=%d", r.rr_inv);
  return 0;
}

// The plan a shape and variant build, on the host only (no device calls;
// table addresses are placeholders).  nullptr on error.
Plan* dry_plan(int precision, int kind, size_t ndim, const size_t* dim, int format, size_t batch,
               const afft_gpu_variant* v) {
  idx_t d[kMaxDim];
  int nd = 0;
  if (validate(ndim, dim, format, afft_gpu, d, nd) != 0) return nullptr;
  if (batch < 1 || kind < 0 || kind > 3 || (precision != 32 && precision != 64)) {
    set_err("invalid arguments");
    return nullptr;
  }
  Plan* P = new (std::nothrow) Plan();
  if (!P) return nullptr;
  P->prec = precision;
  P->kind = kind;
  P->format = format;
  P->ndim = nd;
  for (int i = 0; i < nd; i++) P->dims[i] = d[i];
  P->batch = (idx_t)batch;
  if (v) {
    P->var.col_tile = v->col_tile;
    P->var.rowt_tile = v->rowt_tile;
    P->var.stream_io = v->stream_io;
    P->var.xcd_order = v->xcd_order;
    P->var.chunk = v->chunk;
    P->var.sched = v->sched;
    P->var.half_lds = v->half_lds;
    P->var.pipeline = v->pipeline;
    P->var.real_pair = v->real_pair;
    P->var.a_split = v->a_split;
    P->var.row_log2 = v->row_log2;
    P->var.fused = v->fused;
    P->var.fused_slots = v->fused_slots;
    P->var.fused_lag = v->fused_lag;
    P->var.tiny_rows = v->tiny_rows;
  }
  g_dry = true;
  bool ok = precision == 32 ? Builder<float>(*P).build(kind, format, d, nd)
                            : Builder<double>(*P).build(kind, format, d, nd);
  g_dry = false;
  if (!ok) {
    delete P;
    return nullptr;
  }
  const size_t eb = precision == 32 ? sizeof(cx<float>) : sizeof(cx<double>);
  P->chunk = default_chunk(*P, eb);
  size_scratch(*P, eb);
  return P;
}

// ---- execution with no allocation ---------------------------------------------
// afft_gpu_execute_noalloc runs a plan on a fixed scratch buffer compiled into
// the library (kNoAllocBytes of device memory per GPU, held from the first
// plan created on that GPU to the end of the process) instead of the pool: no
// allocation, so callers that must not allocate (code under graph capture
// that may not grow the pool, tight serving loops) can use it.  The array's
// name is a host-side stand-in; its device address comes from
// hipGetSymbolAddress, once per device, with that device current, during plan
// creation (the set-up every call needs anyway).
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};

void* noalloc_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) {
    set_err("device %d out of range", dev);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_err("no-allocation scratch: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
      return nullptr;
    }
    g_noalloc_ptr[dev] = ptr;
  }
  return g_noalloc_ptr[dev];
}

// Runs the plan with its scratch (and a staged input) in the fixed buffer:
// transforms are grouped as the plan groups them, in smaller groups if that
// is what fits, and always one launch per pass.
template <typename T>
bool execute_noalloc(Plan& P, const void* sre, const void* sim, void* dre, void* dim, hipStream_t st) {
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess && dev != P.device) HIP_OK(hipSetDevice(P.device));
  void* base = noalloc_scratch(P.device);
  if (!base) return false;
  const Staging sg = staging<T>(P, sre, sim, dre, dim);
  const size_t stage_need = sg.n * sg.stride;
  auto need = [&](idx_t c) -> size_t {  // scratch of a group of c transforms, then the staged input
    const size_t s0 = (size_t)P.scr_per_user[0] * (size_t)c * sizeof(cx<T>);
    const size_t s1 = (size_t)P.scr_per_user[1] * (size_t)c * sizeof(cx<T>);
    const size_t tot = s1 ? ((s0 + 255) & ~(size_t)255) + s1 : s0;
    return ((tot + 255) & ~(size_t)255) + stage_need;
  };
  idx_t c = P.chunk;
  while (c > 1 && need(c) > kNoAllocBytes) c = (c + 1) / 2;
  if (need(c) > kNoAllocBytes) {
    set_err("the plan needs %zu bytes of scratch for one transform, above the %zu bytes of the no-allocation "
            "buffer",
            need(1), kNoAllocBytes);
    return false;
  }
  const size_t stage_off = need(c) - stage_need;
  if (!stage_input(sg, (char*)base + stage_off, sre, sim, st)) return false;
  return execute_on<T>(P, sre, sim, dre, dim, st, 0, P.npass, base, c, true);
}

}  // namespace

// ---- C ABI ----------------------------------------------------------------------
extern "C" {

const size_t afft_alignment = 64;

size_t afft_32_align_size(size_t sz) { return (sz + 15) & ~(size_t)15; }
size_t afft_64_align_size(size_t sz) { return (sz + 7) & ~(size_t)7; }

int afft_supported(int impl) {
  if (impl != afft_auto && impl != afft_gpu) return 0;
  int n = 0;
  return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

#define AFFT_KIND_API(P, T, K, KID)                                                                    \
  size_t afft_##P##_##K##_memsize(size_t ndim, const size_t* dim, int format, int impl) {              \
    return memsize_impl<T>(KID, ndim, dim, format, impl);                                             \
  }                                                                                                   \
  afft_##P##_##K##_type* afft_##P##_##K##_create(size_t ndim, const size_t* dim, int format, void* mem, \
                                                 int impl) {                                          \
    return (afft_##P##_##K##_type*)create_impl<T>(KID, ndim, dim, format, 1, mem, impl);              \
  }                                                                                                   \
  void afft_##P##_##K##_destroy(afft_##P##_##K##_type* state) { destroy_impl(state); }

AFFT_KIND_API(32, float, c, 0)
AFFT_KIND_API(32, float, ci, 1)
AFFT_KIND_API(32, float, r, 2)
AFFT_KIND_API(32, float, ri, 3)
AFFT_KIND_API(64, double, c, 0)
AFFT_KIND_API(64, double, ci, 1)
AFFT_KIND_API(64, double, r, 2)
AFFT_KIND_API(64, double, ri, 3)

void afft_32_c_transform(afft_32_c_type* s, const float* sr, const float* si, float* dr, float* di) {
  transform_impl<float>(s, sr, si, dr, di, 0);
}
void afft_32_ci_transform(afft_32_ci_type* s, const float* sr, const float* si, float* dr, float* di) {
  transform_impl<float>(s, sr, si, dr, di, 1);
}
void afft_32_r_transform(afft_32_r_type* s, const float* src, float* dr, float* di) {
  transform_impl<float>(s, src, nullptr, dr, di, 2);
}
void afft_32_ri_transform(afft_32_ri_type* s, const float* sr, const float* si, float* dst) {
  transform_impl<float>(s, sr, si, dst, nullptr, 3);
}
void afft_64_c_transform(afft_64_c_type* s, const double* sr, const double* si, double* dr, double* di) {
  transform_impl<double>(s, sr, si, dr, di, 0);
}
void afft_64_ci_transform(afft_64_ci_type* s, const double* sr, const double* si, double* dr, double* di) {
  transform_impl<double>(s, sr, si, dr, di, 1);
}
void afft_64_r_transform(afft_64_r_type* s, const double* src, double* dr, double* di) {
  transform_impl<double>(s, src, nullptr, dr, di, 2);
}
void afft_64_ri_transform(afft_64_ri_type* s, const double* sr, const double* si, double* dst) {
  transform_impl<double>(s, sr, si, dst, nullptr, 3);
}

size_t afft_gpu_plan_memsize(void) { return kPlanMem; }

afft_gpu_plan* afft_gpu_plan_create(int precision, int kind, size_t ndim, const size_t* dim, int format,
                                    size_t batch, void* mem) {
  if (kind < 0 || kind > 3) {
    set_err("invalid kind %d", kind);
    return nullptr;
  }
  if (precision == 32) return (afft_gpu_plan*)create_impl<float>(kind, ndim, dim, format, batch, mem, afft_gpu);
  if (precision == 64) return (afft_gpu_plan*)create_impl<double>(kind, ndim, dim, format, batch, mem, afft_gpu);
  set_err("precision must be 32 or 64");
  return nullptr;
}

int afft_gpu_execute(afft_gpu_plan* plan, const void* sre, const void* sim, void* dre, void* dim, void* stream) {
  return afft_gpu_execute_range(plan, 0, -1, sre, sim, dre, dim, stream);
}

int afft_gpu_execute_noalloc(afft_gpu_plan* plan, const void* sre, const void* sim, void* dre, void* dim,
                             void* stream) {
  g_err.clear();
  Plan* P = get_plan(plan);
  if (!P) return -1;
  bool ok = P->prec == 32 ? execute_noalloc<float>(*P, sre, sim, dre, dim, (hipStream_t)stream)
                          : execute_noalloc<double>(*P, sre, sim, dre, dim, (hipStream_t)stream);
  return ok ? 0 : -1;
}

int afft_gpu_execute_range(afft_gpu_plan* plan, int first, int last, const void* sre, const void* sim, void* dre,
                           void* dim, void* stream) {
  g_err.clear();
  Plan* Pp = get_plan(plan);
  if (!Pp) return -1;
  Plan& P = *Pp;
  if (last < 0 || last > P.npass) last = P.npass;
  if (first < 0) first = 0;
  bool ok = P.prec == 32 ? execute<float>(P, sre, sim, dre, dim, (hipStream_t)stream, first, last)
                         : execute<double>(P, sre, sim, dre, dim, (hipStream_t)stream, first, last);
  return ok ? 0 : -1;
}

int afft_gpu_plan_launches(const afft_gpu_plan* plan, int pass) {
  Plan* Pp = get_plan(plan);
  if (!Pp) return -1;
  const Plan& P = *Pp;
  if (pass < 0 || pass >= P.npass) return -1;
  const int split = pass == 0 ? a_split(P) : 1;
  int n = 0;
  for (idx_t start = 0; start < P.batch; start += P.chunk) {
    const idx_t cnt = P.batch - start < P.chunk ? P.batch - start : P.chunk;
    n += (split > 1 && cnt % split == 0) ? split : 1;
  }
  return n;
}

void afft_gpu_plan_destroy(afft_gpu_plan* plan) { destroy_impl(plan); }

int afft_gpu_set_stream(void* state, void* stream) {
  Plan* P = get_plan(state);
  if (!P) return -1;
  P->stream = (hipStream_t)stream;
  return 0;
}

int afft_gpu_plan_set_chunk(afft_gpu_plan* plan, size_t chunk) {
  Plan* Pp = get_plan(plan);
  if (!Pp) return -1;
  Plan& P = *Pp;
  size_t eb = P.prec == 32 ? sizeof(cx<float>) : sizeof(cx<double>);
  idx_t c = chunk == 0 ? default_chunk(P, eb) : (idx_t)chunk;
  if (c > P.batch) c = P.batch;
  if (c < 1) c = 1;
  P.chunk = c;
  size_scratch(P, eb);
  return 0;
}

int afft_gpu_plan_set_variant(afft_gpu_plan* plan, const afft_gpu_variant* v) {
  g_err.clear();
  Plan* Pp = get_plan(plan);
  if (!Pp || !v) return -1;
  Plan& P = *Pp;
  Variant old = P.var;
  P.var.col_tile = v->col_tile;
  P.var.rowt_tile = v->rowt_tile;
  P.var.stream_io = v->stream_io;
  P.var.xcd_order = v->xcd_order;
  P.var.chunk = v->chunk;
  P.var.sched = v->sched;
  P.var.half_lds = v->half_lds;
  P.var.pipeline = v->pipeline;
  P.var.real_pair = v->real_pair;
  P.var.a_split = v->a_split;
  P.var.row_log2 = v->row_log2;
  P.var.fused = v->fused;
  P.var.fused_slots = v->fused_slots;
  P.var.fused_lag = v->fused_lag;
  P.var.tiny_rows = v->tiny_rows;
  if (!rebuild_any(P)) {
    P.var = old;
    rebuild_any(P);
    return -1;
  }
  return 0;
}

int afft_gpu_plan_set_mode(afft_gpu_plan* plan, int mode) {
  g_err.clear();
  Plan* Pp = get_plan(plan);
  if (!Pp) return -1;
  Plan& P = *Pp;
  if (mode != afft_gpu_mode_fma && mode != afft_gpu_mode_mfma_f16 && mode != afft_gpu_mode_mfma_bf16) {
    set_err("unknown mode %d", mode);
    return -1;
  }
  const int old = P.mode;
  P.mode = mode;
  if (!rebuild_any(P)) {
    P.mode = old;
    rebuild_any(P);
    return -1;
  }
  if (mode != afft_gpu_mode_fma) {
    int used = 0;
    for (int i = 0; i < P.npass; i++) used += P.pass[i].mfma != 0;
    if (!used) {
      set_err("no pass of this plan can run on the MFMA tile (needs FP32 c2c, interleaved, 256-point passes)");
      P.mode = old;
      rebuild_any(P);
      return -1;
    }
  }
  return 0;
}

int afft_gpu_plan_get_variant(const afft_gpu_plan* plan, afft_gpu_variant* v) {
  Plan* Pp = get_plan(plan);
  if (!Pp || !v) return -1;
  const Plan& P = *Pp;
  v->col_tile = P.var.col_tile;
  v->rowt_tile = P.var.rowt_tile;
  v->stream_io = P.var.stream_io;
  v->xcd_order = P.var.xcd_order;
  v->chunk = (int64_t)P.chunk;
  v->sched = P.var.sched;
  v->half_lds = P.var.half_lds;
  v->pipeline = P.var.pipeline;
  v->real_pair = P.var.real_pair;
  v->a_split = P.var.a_split;
  v->row_log2 = P.var.row_log2;
  v->fused = P.var.fused;
  v->fused_slots = P.var.fused_slots;
  v->fused_lag = P.var.fused_lag;
  v->tiny_rows = P.var.tiny_rows;
  return 0;
}

int afft_gpu_plan_tune(afft_gpu_plan* plan, const void* sre, const void* sim, void* dre, void* dim, void* stream,
                       int iters, float* best_ms) {
  g_err.clear();
  Plan* Pp = get_plan(plan);
  if (!Pp) return -1;
  Plan& P = *Pp;
  if (iters < 1) iters = 3;
  bool ok = P.prec == 32 ? tune_impl<float>(P, sre, sim, dre, dim, (hipStream_t)stream, iters, best_ms)
                         : tune_impl<double>(P, sre, sim, dre, dim, (hipStream_t)stream, iters, best_ms);
  return ok ? 0 : -1;
}

int afft_gpu_plan_num_passes(const afft_gpu_plan* plan) {
  Plan* P = get_plan(plan);
  return P ? P->npass : -1;
}

size_t afft_gpu_plan_scratch_bytes(const afft_gpu_plan* plan) {
  Plan* P = get_plan(plan);
  return P ? scratch_total(*P) : 0;
}

int afft_gpu_plan_describe(const afft_gpu_plan* plan, int i, char* buf, size_t len) {
  Plan* P = get_plan(plan);
  if (!P) return -1;
  return describe_pass(*P, i, buf, len);
}

const char* afft_gpu_last_error(void) { return g_err.c_str(); }

static void fill_side(const Side& s, afft_gpu_side_info& o) {
  o.role = s.role;
  o.il = s.il;
  o.one_array = s.one_array;
  o.re_off = s.re_off;
  o.im_off = s.im_off;
  o.s_j = s.s_j;
  o.s_c = s.s_c;
  o.s_b = s.s_b;
  o.s_b2 = s.s_b2;
  o.s_u = s.s_u;
}

int afft_gpu_plan_inspect(int precision, int kind, size_t ndim, const size_t* dim, int format, size_t batch,
                          afft_gpu_plan_info* info, afft_gpu_pass_info* passes, int max_passes) {
  return afft_gpu_plan_inspect_variant(precision, kind, ndim, dim, format, batch, nullptr, info, passes, max_passes);
}

int afft_gpu_plan_inspect_variant(int precision, int kind, size_t ndim, const size_t* dim, int format, size_t batch,
                                  const afft_gpu_variant* v, afft_gpu_plan_info* info, afft_gpu_pass_info* passes,
                                  int max_passes) {
  Plan* P = dry_plan(precision, kind, ndim, dim, format, batch, v);
  if (!P) return -1;
  if (info) {
    info->npass = P->npass;
    info->fused_slots = P->fused_slots;
    info->chunk = P->chunk;
    info->scr_per_user[0] = P->scr_per_user[0];
    info->scr_per_user[1] = P->scr_per_user[1];
  }
  int n = P->npass < max_passes ? P->npass : max_passes;
  for (int i = 0; i < n && passes; i++) {
    const PassRec& r = P->pass[i];
    afft_gpu_pass_info& o = passes[i];
    std::memset(&o, 0, sizeof(o));
    o.type = r.type;
    o.L = r.L;
    o.CT = r.CT;
    o.colmap = r.colmap;
    o.epi = r.epi;
    o.store_t = r.store_t;
    fill_side(r.in, o.in);
    fill_side(r.out, o.out);
    o.n_col = r.n_col;
    o.fold_user = r.fold_user;
    o.b_per_user = r.b_per_user;
    o.n_blo = r.n_blo;
    o.swap_in = r.swap_in;
    o.swap_out = r.swap_out;
    o.tw4_bits = r.tw4_bits;
    o.tw4_shift = r.tw4_shift;
    o.tw4_blo = r.tw4_blo;
    o.sv = r.sv;
    o.nt_io = (r.in_nt ? 1 : 0) | (r.out_nt ? 2 : 0);
    o.rr_h = r.rr_h;
    o.rr_m = r.rr_m;
    o.rr_inv = r.rr_inv;
  }
  int np = P->npass;
  delete P;
  return np;
}

int afft_gpu_plan_describe_variant(int precision, int kind, size_t ndim, const size_t* dim, int format, size_t batch,
                                   const afft_gpu_variant* v, int i, char* buf, size_t len) {
  Plan* P = dry_plan(precision, kind, ndim, dim, format, batch, v);
  if (!P) return -1;
  const int r = describe_pass(*P, i, buf, len);
  delete P;
  return r;
}

int afft_gpu_error_sums_f32(const float* y, const double* ref, size_t n, double* acc, void* stream) {
  return err_sums<float>(y, ref, (idx_t)n, acc, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}
int afft_gpu_error_sums_f64(const double* y, const double* ref, size_t n, double* acc, void* stream) {
  return err_sums<double>(y, ref, (idx_t)n, acc, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}
int afft_gpu_copy(const void* src, void* dst, size_t bytes, void* stream) {
  return copy16(src, dst, (idx_t)bytes, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}

}  // extern "C"
