// Streamed CSR product kernels instantiated for 1 interleaved vector(s).
#define WPC_STREAM_NV 1
#include "stream_impl.h"
