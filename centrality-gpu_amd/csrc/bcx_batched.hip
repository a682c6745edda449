// Batched multi-source kernels (placeholder translation unit).
#include "bcx_device.h"
