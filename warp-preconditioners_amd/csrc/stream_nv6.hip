// Streamed CSR product kernels instantiated for 6 interleaved vector(s).
#define WPC_STREAM_NV 6
#include "stream_impl.h"
