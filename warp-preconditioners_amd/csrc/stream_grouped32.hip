// Batched sliced products over several systems (runtime vector count), float32
// vectors; the entry point is in stream_grouped.hip.
#define WPC_STREAM_NV 0
#define WPC_GROUPED_F32 1
#include "stream_impl.h"
