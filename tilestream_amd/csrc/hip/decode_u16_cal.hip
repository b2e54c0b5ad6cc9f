// Decode kernels of kind TSK_KIND_U16_CAL (see decode_core.h).
#include "decode_core.h"

TSK_DEFINE_DECODE_TABLE(tsk_decode_table_u16_cal, TSK_KIND_U16_CAL)
