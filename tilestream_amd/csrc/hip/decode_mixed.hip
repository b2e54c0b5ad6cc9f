// Decode kernels of kind TSK_KIND_MIXED (see decode_core.h).
#include "decode_core.h"

TSK_DEFINE_DECODE_TABLE(tsk_decode_table_mixed, TSK_KIND_MIXED)
