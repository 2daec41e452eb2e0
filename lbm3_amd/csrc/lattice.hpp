// D3Q27 lattice: direction ids, velocity components, inverse directions and
// weights (reference include/D3Q27Directions.h:3-56, same numbering), plus the
// esotwist neighbour maps (include/esotwistStreamingFunctions.h:17-255).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace lbm3 {

enum Dir : int {
    OOO = 0, POO, MOO, OOM, OOP, OMO, OPO, POM, MOP, POP, MOM, MMO, PPO, OPM, OMP,
    MPO, PMO, OPP, OMM, MPM, PMP, MMP, PPM, PMM, MPP, MMM, PPP
};

__host__ __device__ constexpr int cx(int q)
{
    constexpr int t[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1};
    return t[q];
}
__host__ __device__ constexpr int cy(int q)
{
    constexpr int t[27] = {0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, -1, 1};
    return t[q];
}
__host__ __device__ constexpr int cz(int q)
{
    constexpr int t[27] = {0, 0, 0, -1, 1, 0, 0, -1, 1, 1, -1, 0, 0, -1, 1, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, -1, 1};
    return t[q];
}
// opposite directions come in pairs (2k-1, 2k)
__host__ __device__ constexpr int inv(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }

template <class T> __host__ __device__ constexpr T weight(int q)
{
    return q == 0 ? T(8) / T(27) : (q < 7 ? T(2) / T(27) : (q < 19 ? T(1) / T(54) : T(1) / T(216)));
}

// Neighbour slots: 0 self, 1 i+, 2 j+, 3 k+, 4 ij+, 5 ik+, 6 jk+, 7 ijk+.
// A cell reads population q before collision from neighbour PRE_NB(q) and
// writes it after collision to neighbour POST_NB(q); the neighbour of q is the
// one that lies in +x/+y/+z along the components where c_q is negative (pre)
// or positive (post).
__host__ __device__ constexpr int nb_slot(int dx, int dy, int dz)
{
    return (dx ? 1 : 0) + (dy ? 2 : 0) + (dz ? 4 : 0) == 0 ? 0
         : ((dx && !dy && !dz) ? 1 : (!dx && dy && !dz) ? 2 : (!dx && !dy && dz) ? 3
         : (dx && dy && !dz) ? 4 : (dx && !dy && dz) ? 5 : (!dx && dy && dz) ? 6 : 7);
}
__host__ __device__ constexpr int pre_nb(int q) { return nb_slot(cx(q) < 0, cy(q) < 0, cz(q) < 0); }
__host__ __device__ constexpr int post_nb(int q) { return nb_slot(cx(q) > 0, cy(q) > 0, cz(q) > 0); }
// memory slot (population plane) used by direction q before / after collision
__host__ __device__ constexpr int pre_plane(int q, bool flip) { return flip ? inv(q) : q; }
__host__ __device__ constexpr int post_plane(int q, bool flip) { return flip ? q : inv(q); }

} // namespace lbm3
