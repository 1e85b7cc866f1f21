// Streamed CSR product kernels instantiated for 8 interleaved vector(s).
#define WPC_STREAM_NV 8
#include "stream_impl.h"
