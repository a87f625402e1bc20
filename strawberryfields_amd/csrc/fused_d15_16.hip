// Stage kernels for cutoffs 15..16 (see fused_impl.h).
#include "fused_impl.h"

SFAMD_STAGE_DEFINE_D(15)
SFAMD_STAGE_DEFINE_D(16)
