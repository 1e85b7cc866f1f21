// Streamed CSR product kernels instantiated for 7 interleaved vector(s).
#define WPC_STREAM_NV 7
#include "stream_impl.h"
