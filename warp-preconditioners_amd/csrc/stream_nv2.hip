// Streamed CSR product kernels instantiated for 2 interleaved vector(s).
#define WPC_STREAM_NV 2
#include "stream_impl.h"
