// Stage kernels for cutoffs 13..14 (see fused_impl.h).
#include "fused_impl.h"

SFAMD_STAGE_DEFINE_D(13)
SFAMD_STAGE_DEFINE_D(14)
