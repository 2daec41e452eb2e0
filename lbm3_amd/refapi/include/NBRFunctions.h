// The four "plus" neighbours that follow from the three stored ones
// (reference include/NBRFunctions.h:5-11): x-neighbours are the next cell in
// memory, wrapping to cell 0 past the last cell.
#pragma once
#include "types.h"

__host__ __device__ inline void finishNBRPlus(NBRStruct &NBR, const InfoStruct &Info)
{
    auto next = [&](int c) { return c + 1 >= Info.cellCount ? 0 : c + 1; };
    NBR.iPlus = next(NBR.self);
    NBR.ijPlus = next(NBR.jPlus);
    NBR.ikPlus = next(NBR.kPlus);
    NBR.ijkPlus = next(NBR.jkPlus);
}
