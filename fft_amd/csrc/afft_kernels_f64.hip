// FP64 instantiations of the pass kernels (16-byte elements halve what fits
// in LDS, so tiles are half the FP32 ones and rows stop at 8192, 16384 with
// half-LDS tiles).
#define AFGPU_T double
#define AFGPU_ROW_MAX 8192
#define AFGPU_ROW_HX_MAX 16384
#define AFGPU_ROWT_CT64 8
#define AFGPU_ROWT_CT128 8
#define AFGPU_ROWT_CT256 8
#define AFGPU_ROWT_CT512 8
#define AFGPU_ROWT_CT1024 8
#define AFGPU_ROWT_SV1_CT1024 8
#define AFGPU_ROWT_CT2048 4
#define AFGPU_ROWT_CT4096 2
#define AFGPU_COL_CT_SMALL 8
#define AFGPU_COL_CT256 8
#define AFGPU_COL_CT512 8
#define AFGPU_COL_CT1024 8
#define AFGPU_COL_CT2048 4
#define AFGPU_COL_CT4096 2
#include "afft_variants.inc"
