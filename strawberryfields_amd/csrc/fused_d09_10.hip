// Stage kernels for cutoffs 9..10 (see fused_impl.h).
#include "fused_impl.h"

SFAMD_STAGE_DEFINE_D(9)
SFAMD_STAGE_DEFINE_D(10)
