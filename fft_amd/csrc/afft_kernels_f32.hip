// FP32 instantiations of the pass kernels.
// Column tiles: 16 columns (128-byte segments) up to L = 1024; beyond that the
// 160 KiB LDS of a CU bounds CT * L.
#define AFGPU_T float
#define AFGPU_ROW_MAX 16384
#define AFGPU_ROW_HX_MAX 32768
#define AFGPU_ROWT_CT64 16
#define AFGPU_ROWT_CT128 16
#define AFGPU_ROWT_CT256 16
#define AFGPU_ROWT_CT512 16
#define AFGPU_ROWT_CT1024 8
#define AFGPU_ROWT_SV1_CT1024 16
#define AFGPU_ROWT_CT2048 8
#define AFGPU_ROWT_CT4096 4
#define AFGPU_COL_CT_SMALL 16
#define AFGPU_COL_CT256 16
#define AFGPU_COL_CT512 16
#define AFGPU_COL_CT1024 16
#define AFGPU_COL_CT2048 8
#define AFGPU_COL_CT4096 4
#define AFGPU_WIDE_COL1024 1
#define AFGPU_C2R4 1
#define AFGPU_WIDE_REAL_COL 1
#include "afft_variants.inc"
