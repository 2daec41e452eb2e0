// Refinement and wall markers (reference include/markerFunctions.h) run inside
// lbm3_build_grids, calling the case's getRefinementModifier through the hook
// callbacks of gridBuilderFunctions.h.
#pragma once
#include "types.h"
