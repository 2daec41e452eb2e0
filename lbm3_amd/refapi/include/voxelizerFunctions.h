// The STL voxelizer (reference include/voxelizerFunctions.h) runs inside
// lbm3_build_grids on the GPU; a case file includes this header for the
// reference's include order only.
#pragma once
#include "types.h"
