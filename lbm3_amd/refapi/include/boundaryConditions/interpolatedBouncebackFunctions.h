// Wall data word of a wall-adjacent cell (reference
// include/boundaryConditions/interpolatedBouncebackFunctions.h:662-702): bit 0
// marks an interface overlap, bits 1-26 the links that exist, bits 27-31 the
// wall id.  The bounceback itself runs in the library's wall kernel.
#pragma once
#include "../types.h"

__host__ __device__ inline void unpackWallID(const uint32_t &wallData, int &wallID, bool &interfaceOverlapMarker)
{
    interfaceOverlapMarker = (wallData & 1u) != 0u;
    wallID = (int)((wallData >> 27) & 31u);
}
