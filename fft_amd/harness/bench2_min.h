/*
 * bench2_min.h -- the part of FFTW's libbench2 user interface (bench-user.h)
 * that the afft adapter in bench2_afft.cpp uses, declared here so that the
 * adapter also builds without an FFTW source tree.  When built against FFTW
 * (-DAFFT_WITH_LIBBENCH2 -I <fftw>/libbench2), the real bench-user.h is used
 * instead and libbench2 supplies main(), timing and verification.
 *
 * The field names and meanings follow libbench2's public header: a problem
 * has a kind, a size tensor (rank and per-dimension n / input stride / output
 * stride), a vector tensor (batch), a sign (-1 forward, +1 backward), an
 * in-place flag and host input / output buffers.
 */
#ifndef AFFT_BENCH2_MIN_H
#define AFFT_BENCH2_MIN_H

typedef float bench_real;
typedef bench_real bench_complex[2];

typedef struct
{
  int n;
  int is; /* input stride */
  int os; /* output stride */
} bench_iodim;

typedef struct
{
  int rnk;
  bench_iodim* dims;
} bench_tensor;

typedef enum { PROBLEM_COMPLEX, PROBLEM_REAL, PROBLEM_R2R } problem_kind_t;

typedef struct
{
  problem_kind_t kind;
  bench_tensor* sz;
  bench_tensor* vecsz;
  int sign;
  int in_place;
  int destroy_input;
  int split;
  void* in;
  void* out;
  char* pstring;
  void* userinfo;
} bench_problem;

struct bench_doc
{
  const char* key;
  const char* val;
  const char* (*f)(void);
};

#define BEGIN_BENCH_DOC struct bench_doc bench_doc[] = {
#define BENCH_DOC(key, val) {key, val, 0},
#define BENCH_DOCF(key, f) {key, 0, f},
#define END_BENCH_DOC {0, 0, 0}};

#ifdef __cplusplus
extern "C"
{
#endif
int can_do(bench_problem* p);
void setup(bench_problem* p);
void doit(int iter, bench_problem* p);
void done(bench_problem* p);
void cleanup(void);
void main_init(int* argc, char*** argv);
#ifdef __cplusplus
}
#endif

#endif
