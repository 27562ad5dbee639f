// Broadband (BPM) kernels for gfx950 -- filled in with the broadband path.
#include <hip/hip_runtime.h>
