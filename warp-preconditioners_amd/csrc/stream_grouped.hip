// Batched sliced products over several systems, each with its own number of
// interleaved vectors (one runtime-count instantiation of stream_impl.h): float64
// vectors and the entry point; float32 vectors in stream_grouped32.hip.
#define WPC_STREAM_NV 0
#include "stream_impl.h"
