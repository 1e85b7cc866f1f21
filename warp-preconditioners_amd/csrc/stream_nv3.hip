// Streamed CSR product kernels instantiated for 3 interleaved vector(s).
#define WPC_STREAM_NV 3
#include "stream_impl.h"
