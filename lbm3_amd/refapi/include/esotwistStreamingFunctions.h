// Esoteric-twist addressing (reference include/esotwistStreamingFunctions.h:
// 6-255) from the library's neighbour maps (lattice.hpp pre_nb / post_nb,
// pre_plane / post_plane): direction q of a cell is read before collision from
// slot pre_plane(q) of the neighbour pre_nb(q), and written after collision to
// slot post_plane(q) of the neighbour post_nb(q).
#pragma once
#include "types.h"

__host__ __device__ inline int lbm3_nb_cell(const NBRStruct &NBR, int slot)
{
    switch (slot) {
    case 0: return NBR.self;
    case 1: return NBR.iPlus;
    case 2: return NBR.jPlus;
    case 3: return NBR.kPlus;
    case 4: return NBR.ijPlus;
    case 5: return NBR.ikPlus;
    case 6: return NBR.jkPlus;
    default: return NBR.ijkPlus;
    }
}

__host__ __device__ inline void getPreCollisionIndex(int (&cellIndex)[27], int (&fIndex)[27], const NBRStruct &NBR,
                                                     const bool &esotwistFlipper)
{
    for (int q = 0; q < 27; ++q) {
        cellIndex[q] = lbm3_nb_cell(NBR, lbm3::pre_nb(q));
        fIndex[q] = lbm3::pre_plane(q, esotwistFlipper);
    }
}

__host__ __device__ inline void getPostCollisionIndex(int (&cellIndex)[27], int (&fIndex)[27], const NBRStruct &NBR,
                                                      const bool &esotwistFlipper)
{
    for (int q = 0; q < 27; ++q) {
        cellIndex[q] = lbm3_nb_cell(NBR, lbm3::post_nb(q));
        fIndex[q] = lbm3::post_plane(q, esotwistFlipper);
    }
}
