/*
 * afft_gpu.h -- extensions of the MI355X afft backend beyond fft.h.
 *
 * None of these has a counterpart in the reference (afft has no batch, stream,
 * destroy, precision mode or work-queue API).  They exist for callers that
 * need asynchronous, batched or mixed-size execution on the GPU.  Every
 * pointer argument named src/dst is a device pointer; `stream` is a
 * hipStream_t passed as void* (NULL = the null stream).
 */
#ifndef AFFT_MI355X_GPU_H
#define AFFT_MI355X_GPU_H

#include <stddef.h>
#include <stdint.h>

#include "fft.h"

#ifdef __cplusplus
extern "C"
{
#endif

/* impl id of this backend; afft_auto selects it too */
enum { afft_gpu = 16 };

/* transform kinds */
enum { afft_gpu_c = 0, afft_gpu_ci = 1, afft_gpu_r = 2, afft_gpu_ri = 3 };

/* arithmetic modes of FP32 complex plans */
enum
{
  afft_gpu_mode_fma = 0,       /* FP32 FMA Stockham passes (default) */
  afft_gpu_mode_mfma_f16 = 1,  /* 256-point passes as FP16 hi+lo MFMA products, FP32 accumulation */
  afft_gpu_mode_mfma_bf16 = 2  /* the same with BF16 hi+mid+lo operands (six products) */
};

typedef struct afft_gpu_plan_struct afft_gpu_plan;

/* Bytes of host memory a plan needs (any alignment).  The plan lives in that
 * memory, as afft's does; the caller frees it when done. */
size_t afft_gpu_plan_memsize(void);

/* A batch of `batch` transforms of shape dim[0..ndim) laid out back to back:
 *   c / ci : N complex in, N complex out per transform;
 *   r      : N reals in, (dim0/2+1)*inner complex out per transform;
 *   ri     : (dim0/2+1)*inner complex in, N reals out per transform;
 * (split format: each of the re / im arrays advances by the per-transform
 * element count).  precision is 32 or 64.  Returns NULL on error
 * (afft_gpu_last_error tells why). */
afft_gpu_plan* afft_gpu_plan_create(int precision, int kind, size_t ndim, const size_t* dim, int format,
                                    size_t batch, void* mem);

/* Enqueues the transform on `stream` and returns without waiting.  Returns 0
 * on success, -1 on error. */
int afft_gpu_execute(afft_gpu_plan* plan, const void* src_re, const void* src_im, void* dst_re, void* dst_im,
                     void* stream);

/* The same transform as afft_gpu_execute without any allocation: no
 * counterpart in the reference (afft allocates nothing at transform time
 * because its plan memory holds its work arrays).  For callers that cannot
 * allocate device memory: code under HIP graph capture that must not grow the
 * scratch pool, serving loops and harnesses that forbid hipMalloc.  The plan's
 * scratch (and the copy of an input that overlaps the output) comes from a
 * fixed 64 MiB device buffer compiled into the library, one per GPU, held from
 * the first plan created on that GPU until the process ends (plan creation
 * resolves it, so this call itself allocates nothing); transforms run in groups that
 * fit it, one launch per pass (the fused and pipelined schedules are not
 * used).  A plan whose single transform needs more than 64 MiB of scratch is
 * rejected with -1 and a message (afft_gpu_last_error), and nothing runs.
 * Every call on a GPU shares that GPU's buffer and it keeps the last call's
 * contents (no kernel here reads scratch it did not write first): calls that
 * overlap on one GPU -- on two streams, or from two threads -- race.
 * Returns 0 on success, -1 on error. */
int afft_gpu_execute_noalloc(afft_gpu_plan* plan, const void* src_re, const void* src_im, void* dst_re, void* dst_im,
                             void* stream);

/* Enqueues only passes [first, last) of the plan (last = -1: to the end),
 * for every chunk of the batch.  For measuring one kernel of a multi-pass
 * plan; running the ranges in order equals afft_gpu_execute. */
int afft_gpu_execute_range(afft_gpu_plan* plan, int first, int last, const void* src_re, const void* src_im,
                           void* dst_re, void* dst_im, void* stream);

/* Kernel launches pass `pass` makes per execute (one per chunk of the
 * batch), or -1. */
int afft_gpu_plan_launches(const afft_gpu_plan* plan, int pass);

/* Invalidates the plan.  Plans hold no device memory of their own (twiddle
 * tables and scratch are shared by the process, see afft_gpu_plan_create), so
 * freeing the plan memory without this call leaks nothing. */
void afft_gpu_plan_destroy(afft_gpu_plan* plan);

/* Works for a plan made by any afft_XX_Y_create of fft.h as well: the stream
 * its transform call runs on (and then waits for). */
int afft_gpu_set_stream(void* state, void* stream);

/* Number of transforms processed per pass group in large four-step plans
 * (0 = automatic).  Tuning knob; the result does not depend on it. */
int afft_gpu_plan_set_chunk(afft_gpu_plan* plan, size_t chunk);

/* Selects the arithmetic mode.  afft_gpu_mode_mfma_f16 / _bf16 run every
 * 256-point pass of an FP32 interleaved c2c plan (1-D N = 256 or 65536 =
 * 256 x 256, and 256-long axes of N-D shapes) on the matrix cores: each
 * radix-16 stage is a product of split operands (FP16 hi+lo, or BF16
 * hi+mid+lo) with FP32 accumulation, after a per-transform power-of-two
 * scaling.  Returns -1 if no pass of the plan qualifies or the mode is unknown. */
int afft_gpu_plan_set_mode(afft_gpu_plan* plan, int mode);

/* Runs a device check of the MFMA operand layout the mode relies on; returns
 * the number of mismatching entries (0 = correct) or -1. */
int afft_gpu_mfma_selftest(void);

/* Kernel-shape variant of a plan.  Every variant computes the same result;
 * they differ in speed.  0 means "default" for the integer fields. */
typedef struct
{
  int32_t col_tile;  /* columns per work-group of column passes (8, 16; in FP32 plans 32
                        for the 256-, 512- and 1024-point and 64 for the 256-point
                        four-step columns) */
  int32_t rowt_tile; /* rows per work-group of transposed-store row passes (4, 8 or 16) */
  int32_t stream_io; /* 1: non-temporal loads of the caller's input, stores of the output;
                        0: cached; 2 (the default of a plan created without a variant):
                        cached for N-D plans whose arrays hold at most 128 MiB per side
                        (input, output and the passes between them share the MALL),
                        non-temporal otherwise */
  int32_t xcd_order; /* 1: hand each XCD a contiguous run of tiles */
  int64_t chunk;     /* transforms per pass group of four-step plans */
  int32_t sched;     /* 1 or 0 (default): radix-32 stage schedules for the long passes
                        where compiled, -1: never */
  int32_t half_lds;  /* 1: half-LDS kernels (real and imaginary parts exchanged in turn)
                        where compiled, -1: never, 0: default (on) */
  int32_t pipeline;  /* must be 0 or -1: the schedule that ran pass A of group k beside
                        pass B of group k-1 in one launch measured slower on every shape
                        and is no longer provided (1 is refused with a message) */
  int32_t real_pair; /* fused real passes.  N-D r2c / c2r with dim0 > 4096: the
                        (inverse) real pass over row pairs (k, h-k) runs inside
                        pass B (r2c) or pass A (c2r) of the dim-0 column
                        four-step (0, default; 64..256-point columns); with 1
                        (or, where that has no kernel, rows of up to 4096
                        points) and a 1024..8192-point last axis of a 2-D
                        transform, inside the row FFTs; with -1 as a pass of
                        its own.  FP32 1-D real transforms with N > 2^15:
                        2 runs the real-input four-step (two passes over N/2
                        complex: pass A reads two real columns as one complex
                        column, pass B stores each bin once; 2^16 <= N <= 2^23),
                        the default for r2c up to 2^23 and c2r 2^17..2^22 (c2r
                        2^16 runs as one half-LDS row pass).
                        Longer ones (or -1) run the complex four-step of the
                        N/2 packed points with the real pass on its own: r2c
                        as a row pass into a second scratch and the transposing
                        real pass, c2r with the inverse real pre-pass first.
                        1: r2c with pass B, the real post-pass and the
                        transposition fused (rows k1 and n1-k1 per work-group),
                        c2r (2^20 <= N <= 2^22) with the pre-pass in the load of
                        pass A -- both measured slower */
  int32_t a_split;   /* two-pass plans (pass A: input -> scratch, pass B: scratch ->
                        output): pass A of each group runs in this many launches,
                        pass B in one (0 or 1: one) */
  int32_t row_log2;  /* 1-D four-step plans: row length n2 = 2^row_log2 of the split
                        n = n1 * n2 (pass A: n2 columns of n1 points, pass B: n1 rows
                        of n2 points); 0 (default): 1024 up to 2^20 (n/64 where
                        that is shorter, so columns have at least 64 points),
                        longer rows beyond */
  int32_t fused;     /* batched 1-D FP32 c2c four-step plans: 1 would run both passes in
                        one persistent launch (pass-A and pass-B tiles of each transform
                        on one XCD, the intermediate in a ring of scratch slots that
                        stays in the MALL); that launch is switched off in this build
                        and a variant with fused = 1 is rejected; -1 or 0: one launch
                        per pass */
  int32_t fused_slots; /* ring slots per XCD of the fused launch (0: 3) */
  int32_t fused_lag;   /* rounds between a transform's pass A and its pass B in the
                          fused launch (0: 1; < fused_slots) */
  int32_t tiny_rows;   /* rows of 2..32 complex points (c2c N <= 32, 1-D r2c / c2r
                          N <= 64, last axes of N-D c2c of that length) can run one
                          row per thread, each work-group's rows staged through LDS
                          so its loads and stores are contiguous: 64, 128 or 256
                          rows per work-group; 0 (default): so for rows of up to 16
                          points (by row size), 32-point rows on the general row
                          kernel; -1: the general row kernel for all */
} afft_gpu_variant;

int afft_gpu_plan_set_variant(afft_gpu_plan* plan, const afft_gpu_variant* v);
int afft_gpu_plan_get_variant(const afft_gpu_plan* plan, afft_gpu_variant* v);

/* Autotuner: runs the plan `iters` times per candidate variant on the given
 * buffers (dst is overwritten; src is read only) on `stream`, keeps the
 * fastest and records it in the process-wide cache keyed by (precision,
 * kind, format, batch, dims).  Later plans of the same key start from the
 * tuned variant; with AFFT_GPU_TUNE_CACHE=<file> the cache persists across
 * processes.  best_ms (optional) receives the winner's time per execute. */
int afft_gpu_plan_tune(afft_gpu_plan* plan, const void* src_re, const void* src_im, void* dst_re, void* dst_im,
                       void* stream, int iters, float* best_ms);

/* Plan introspection: number of kernel launches per execute, and the bytes
 * of the process-wide scratch pool one execute uses.  Scratch comes from one
 * pool per device that only grows (to at most twice the largest request) and
 * is shared by all plans: executes on one stream run in order; an execute on
 * another stream waits for the pool's previous user; executes under HIP graph
 * capture skip that wait, so replays of a captured plan must not overlap
 * executes of other plans on other streams. */
int afft_gpu_plan_num_passes(const afft_gpu_plan* plan);
size_t afft_gpu_plan_scratch_bytes(const afft_gpu_plan* plan);
/* Writes a one-line description of pass i (kernel shape) into buf. */
int afft_gpu_plan_describe(const afft_gpu_plan* plan, int i, char* buf, size_t len);

/* The last error message of the calling thread ("" if none). */
const char* afft_gpu_last_error(void);

/* ---- uneven batches in one persistent launch ------------------------------ */
/* One job of a mixed batch: a 1-D FP32 transform of n = 2^log2n points,
 * interleaved complex, with afft's semantics for the kind:
 *   afft_gpu_c / afft_gpu_ci : n complex in, n complex out;
 *   afft_gpu_r               : n reals in, n/2+1 complex out;
 *   afft_gpu_ri              : n/2+1 complex in, n reals out. */
typedef struct
{
  int32_t kind;
  int32_t log2n;
  uint64_t src; /* device address of the input */
  uint64_t dst; /* device address of the output (distinct from src) */
} afft_gpu_job;

typedef struct afft_gpu_queue_struct afft_gpu_queue;

/* Bytes of host memory a queue header needs. */
size_t afft_gpu_queue_memsize(void);
/* A queue for up to max_jobs jobs of sizes up to 2^max_log2n (<= 2^16). */
afft_gpu_queue* afft_gpu_queue_create(size_t max_jobs, int max_log2n, void* mem);
/* Groups the jobs into tasks and uploads the task list (synchronous host
 * work).  It first waits for the previous eager launch of this queue; the
 * task, pointer and counter arrays were sized for max_jobs at create, and only
 * the scratch of the long jobs grows here.  Jobs of complex length <= 4096 run
 * in groups of one tile; longer ones as chains of tiles (real pre-pass,
 * four-step pass A, pass B, real post-pass) spread over the whole GPU.
 * Returns the number of tasks, or -1. */
int afft_gpu_queue_prepare(afft_gpu_queue* q, const afft_gpu_job* jobs, size_t njobs);
/* Runs the prepared task list in one persistent launch on `stream`.  Issues
 * kernels only, so it can be captured in a HIP graph and replayed.  A launch
 * and every replay of a captured one run the task list of the LATEST prepare
 * (the list offsets and counts are read from device memory at run time): a
 * prepare between replays changes what the graph computes, and the caller
 * must let replays finish before calling prepare (prepare only waits for
 * eager launches). */
int afft_gpu_queue_launch(afft_gpu_queue* q, void* stream);
/* prepare + launch. */
int afft_gpu_queue_run(afft_gpu_queue* q, const afft_gpu_job* jobs, size_t njobs, void* stream);
int afft_gpu_queue_num_tasks(const afft_gpu_queue* q);
/* Waits for the device, then returns 0 if the last launch ran every task with
 * its dependencies met, 1 if a dependency wait of a long job's tile hit the
 * spin limit (the launch still completes; that job's output is then not
 * defined), -1 on error. */
int afft_gpu_queue_status(const afft_gpu_queue* q);
/* Iterations a tile waits for the tiles it depends on before it flags the
 * queue and proceeds (default 2^26, about a second). */
int afft_gpu_queue_set_spin_limit(afft_gpu_queue* q, long long limit);
/* Persistent work-groups per CU: 4 (default; registers bounded to 128 per
 * lane, some spills) or 2 (registers for the largest tiles, half the tiles in
 * flight). */
int afft_gpu_queue_set_groups_per_cu(afft_gpu_queue* q, int groups);
/* Jobs between consecutive steps of one long job in a task list (default 128;
 * 0..1024): larger lags let more of a chain's tiles finish before a dependent
 * tile is popped, smaller ones keep a job's data in cache.  Takes effect at
 * the next prepare. */
int afft_gpu_queue_set_lag(afft_gpu_queue* q, int lag);
int afft_gpu_queue_get_config(const afft_gpu_queue* q, int* groups_per_cu, int* lag);
/* Autotuning for a batch (no counterpart in afft, which has no queue): times
 * each (groups per CU, lag) candidate on these jobs -- prepare, one warm
 * launch, `iters` timed launches on `stream` -- keeps the fastest whose
 * launches met every dependency and leaves the queue prepared with it.
 * best_ms (optional) receives its time per launch.  The jobs' outputs are
 * written by every candidate (the same values).  Returns 0, or -1. */
int afft_gpu_queue_tune(afft_gpu_queue* q, const afft_gpu_job* jobs, size_t njobs, void* stream, int iters,
                        float* best_ms);
void afft_gpu_queue_destroy(afft_gpu_queue* q);

/* ---- plan inspection (host only, no GPU needed) ---------------------------- */
/* Buffer side of a pass: role 0 = src, 1 = dst, 2/3 = plan scratch.  il = 1:
 * interleaved complex addressing (offsets/strides in complex elements); il = 0:
 * split addressing in scalar elements, with im taken from the re array when
 * one_array = 1.  Element (b, c, j) of a pass is at
 * re_off + s_u*user + (b / n_blo)*s_b + (b % n_blo)*s_b2 + c*s_c + j*s_j. */
typedef struct
{
  int32_t role, il, one_array, pad_;
  int64_t re_off, im_off, s_j, s_c, s_b, s_b2, s_u;
} afft_gpu_side_info;

typedef struct
{
  int32_t type; /* 0 = FFT pass, 1 = real-rows pass, 2 = single-value real,
                   3 = transposing real pass (rr_h = h, rr_m = n1),
                   4 = pass B of a 1-D r2c four-step fused with the real pass
                       (L = n2-point rows of the input, rr_h = h, rr_m = n1),
                   5 = N-D r2c real pass over row pairs fused into the L-point
                       row FFTs, in place (rr_h = h, row stride in.s_c),
                   6 = N-D c2r inverse real pass over row pairs fused into the
                       L-point inverse row FFTs (input rows 0..h, output rows
                       0..h-1; rr_h = h, row strides in.s_c / out.s_c),
                   7 = N-D r2c: pass B of the dim-0 column four-step h = h1 * L
                       with the real pass over row pairs fused in (input
                       entry (k1, j, c) at k1*in.s_b2 + j*in.s_j + c, output
                       rows 0..h at out.s_j; n_blo = h1, rr_h = h, rr_m = the
                       number of columns),
                   8 = N-D c2r: the inverse real pass over row pairs fused
                       into pass A of the dim-0 column four-step h = L * h2
                       (input rows 0..h at in.s_j, output entry (k1, n2, c) at
                       k1*out.s_j + n2*rr_m + c; n_blo = h2, rr_h = h),
                   9 / 10 = 1-D r2c real-input four-step pass A (L = n1-point
                       columns, rr_m = n2) / pass B (L = n2-point rows,
                       rr_m = n1), rr_h = N/2,
                   11 / 12 = 1-D c2r real-input four-step pass B' (L = n2,
                       rr_m = n1) / pass A' (L = n1, rr_m = n2),
                   13 = rows of L = 2..32 points, one per thread, CT rows per
                       work-group staged through LDS (row r at r*s_c, entries
                       adjacent; epi plain, r2c or c2r as for type 0) */
  int32_t L, CT, colmap, epi, store_t;
  afft_gpu_side_info in, out;
  int64_t n_col;
  int32_t fold_user, n_blo;
  int64_t b_per_user;
  int32_t swap_in, swap_out, tw4_bits, tw4_shift;
  int64_t rr_h, rr_m;
  int32_t rr_inv;
  int32_t tw4_blo; /* 1: the four-step twiddle's column index n2 is the low batch index
                      (b % n_blo, N = L * n_blo) instead of c >> tw4_shift */
  int32_t sv;      /* stage schedule of the pass kernel (0: default radices, 1: radix-32) */
  int32_t nt_io;   /* bit 0: non-temporal loads of the pass input, bit 1: non-temporal stores */
} afft_gpu_pass_info;

typedef struct
{
  int32_t npass;
  int32_t fused_slots; /* > 0: both passes run in one persistent launch with this many
                          ring slots per XCD (afft_gpu_variant.fused) */
  int64_t chunk;
  int64_t scr_per_user[2];
} afft_gpu_plan_info;

/* Builds the pass list of a plan on the host only (no GPU calls, nothing
 * allocated on the device) and copies up to max_passes descriptors out.
 * Returns the number of passes, or -1 on error.  For tests and tooling that
 * check a plan's addressing without a GPU. */
int afft_gpu_plan_inspect(int precision, int kind, size_t ndim, const size_t* dim, int format, size_t batch,
                          afft_gpu_plan_info* info, afft_gpu_pass_info* passes, int max_passes);
/* The same for the plan a variant (afft_gpu_plan_set_variant) would build;
   v = NULL: the default variant. */
int afft_gpu_plan_inspect_variant(int precision, int kind, size_t ndim, const size_t* dim, int format, size_t batch,
                                  const afft_gpu_variant* v, afft_gpu_plan_info* info, afft_gpu_pass_info* passes,
                                  int max_passes);

/* The one-line description afft_gpu_plan_describe gives for pass i of the
 * plan this shape and variant (NULL: default) build, built on the host only.
 * Returns 0, or -1 on error or when i is out of range. */
int afft_gpu_plan_describe_variant(int precision, int kind, size_t ndim, const size_t* dim, int format, size_t batch,
                                   const afft_gpu_variant* v, int i, char* buf, size_t len);

/* ---- helpers -------------------------------------------------------------- */
/* acc[0] = sum (y[i] - ref[i])^2, acc[1] = sum ref[i]^2 over n values, in
 * FP64 on the device (acc is a device array of 2 doubles). */
int afft_gpu_error_sums_f32(const float* y, const double* ref, size_t n, double* acc, void* stream);
int afft_gpu_error_sums_f64(const double* y, const double* ref, size_t n, double* acc, void* stream);
/* Streaming device copy (16-byte accesses); the measured bandwidth ceiling. */
int afft_gpu_copy(const void* src, void* dst, size_t bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif
