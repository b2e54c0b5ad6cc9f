// Decode kernels of kind TSK_KIND_U16 (see decode_core.h).
#include "decode_core.h"

TSK_DEFINE_DECODE_TABLE(tsk_decode_table_u16, TSK_KIND_U16)
