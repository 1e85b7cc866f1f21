// Streamed CSR product kernels instantiated for 4 interleaved vector(s).
#define WPC_STREAM_NV 4
#include "stream_impl.h"
