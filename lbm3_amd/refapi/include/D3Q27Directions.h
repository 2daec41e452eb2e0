// D3Q27 direction ids, velocity components, inverse pairs and weights under the
// reference's names (include/D3Q27Directions.h:8-56), taken from the library's
// lattice tables so case code and kernels agree on the numbering.
#pragma once
#include "lattice.hpp"

constexpr int OOO = lbm3::OOO, POO = lbm3::POO, MOO = lbm3::MOO, OOM = lbm3::OOM, OOP = lbm3::OOP, OMO = lbm3::OMO,
              OPO = lbm3::OPO, POM = lbm3::POM, MOP = lbm3::MOP, POP = lbm3::POP, MOM = lbm3::MOM, MMO = lbm3::MMO,
              PPO = lbm3::PPO, OPM = lbm3::OPM, OMP = lbm3::OMP, MPO = lbm3::MPO, PMO = lbm3::PMO, OPP = lbm3::OPP,
              OMM = lbm3::OMM, MPM = lbm3::MPM, PMP = lbm3::PMP, MMP = lbm3::MMP, PPM = lbm3::PPM, PMM = lbm3::PMM,
              MPP = lbm3::MPP, MMM = lbm3::MMM, PPP = lbm3::PPP;

// indexable tables with the reference's names, usable in host and device code
struct Lbm3CX { __host__ __device__ constexpr int operator[](int q) const { return lbm3::cx(q); } };
struct Lbm3CY { __host__ __device__ constexpr int operator[](int q) const { return lbm3::cy(q); } };
struct Lbm3CZ { __host__ __device__ constexpr int operator[](int q) const { return lbm3::cz(q); } };
struct Lbm3Inv { __host__ __device__ constexpr int operator[](int q) const { return lbm3::inv(q); } };
struct Lbm3W { __host__ __device__ constexpr float operator[](int q) const { return lbm3::weight<float>(q); } };

constexpr Lbm3CX CX_DIRECTIONS{};
constexpr Lbm3CY CY_DIRECTIONS{};
constexpr Lbm3CZ CZ_DIRECTIONS{};
constexpr Lbm3Inv INVERSE_DIRECTIONS{};
constexpr Lbm3W DIRECTION_WEIGHTS{};
constexpr float INVSQRT3 = 0.57735026918962576f;
