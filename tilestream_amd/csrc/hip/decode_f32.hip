// Decode kernels of kind TSK_KIND_F32 (see decode_core.h).
// Compiled once per lane-group width (-DTSK_DECODE_GROUP=1, 2, 4, 8, 16).
#include "decode_core.h"

#ifndef TSK_DECODE_GROUP
#error "decode_f32.hip is compiled with -DTSK_DECODE_GROUP set to 1, 2, 4, 8 or 16"
#endif

TSK_DEFINE_DECODE_TABLE(f32, TSK_KIND_F32)
