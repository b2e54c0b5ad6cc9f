// Decode kernels of kind TSK_KIND_F32_CAL (see decode_core.h).
#include "decode_core.h"

TSK_DEFINE_DECODE_TABLE(tsk_decode_table_f32_cal, TSK_KIND_F32_CAL)
