// readSTL (reference include/STLFunctions.h:317-462): a binary STL file into
// an STLStruct on the host.  Dropping triangles with a repeated vertex, the
// bins and the closed-mesh check (checkSTLEdges, :130-315) happen when
// buildGrids hands the bodies to the library (lbm3_build_grids), which refuses
// an open mesh as the reference does.
#pragma once
#include "types.h"

inline void readSTL(STLStruct &STL, const std::string &filename)
{
    std::cout << "Reading STL: " << filename << std::endl;
    std::ifstream file(filename, std::ios::binary);
    if (!file.is_open()) throw std::runtime_error("Failed to open STL file");
    char header[80];
    uint32_t count = 0;
    file.read(header, 80);
    file.read(reinterpret_cast<char *>(&count), sizeof(count));
    if (!file) throw std::runtime_error("Failed to read the STL header of " + filename);
    STL.triangles.assign((size_t)count * 9, 0.f);
    const float inf = std::numeric_limits<float>::infinity();
    BoundsStruct B;
    B.xMin = B.yMin = B.zMin = inf;
    B.xMax = B.yMax = B.zMax = -inf;
    for (uint32_t t = 0; t < count; ++t) {
        float rec[12];
        uint16_t attribute = 0;
        file.read(reinterpret_cast<char *>(rec), sizeof(rec));
        file.read(reinterpret_cast<char *>(&attribute), sizeof(attribute));
        if (!file) throw std::runtime_error("STL file " + filename + " ends inside triangle " + std::to_string(t));
        for (int v = 0; v < 9; ++v) STL.triangles[(size_t)t * 9 + v] = rec[3 + v];
        for (int p = 0; p < 3; ++p) {
            const float x = rec[3 + 3 * p], y = rec[4 + 3 * p], z = rec[5 + 3 * p];
            B.xMin = std::min(B.xMin, x); B.xMax = std::max(B.xMax, x);
            B.yMin = std::min(B.yMin, y); B.yMax = std::max(B.yMax, y);
            B.zMin = std::min(B.zMin, z); B.zMax = std::max(B.zMax, z);
        }
    }
    STL.triangleCount = (int)count;
    STL.Bounds = B;
    std::cout << "\tTriangle count: " << count << "\n\txMin xMax: " << B.xMin << " " << B.xMax << "\n\tyMin yMax: " << B.yMin
              << " " << B.yMax << "\n\tzMin zMax: " << B.zMin << " " << B.zMax << std::endl;
}
