// Streamed CSR product kernels instantiated for 5 interleaved vector(s).
#define WPC_STREAM_NV 5
#include "stream_impl.h"
