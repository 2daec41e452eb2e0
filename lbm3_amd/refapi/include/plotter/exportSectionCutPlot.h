// exportSectionCutPlot{XY,ZY,ZX} (reference include/plotter/
// exportSectionCutPlot.h:11-461): the section through every level computed by
// the library (lbm3_section_cut) and written to /dev/shm/sim_data.bin in the
// reference's format: int32 {plot number, vertical pixels, horizontal pixels,
// 6}, then per pixel six float32 {pressure [Pa], horizontal, vertical and
// normal velocity [m/s], solid marker, cell size [mm]}.  The library draws the
// plane in the grid frame: the rotor-frame overloads crop as the others do.
#pragma once
#include "../types.h"
#include "../gridBuilderFunctions.h"

namespace lbm3facade {

inline void section_cut(std::vector<GridStruct> &grids, const BoundsStruct *Bounds, float cut, int plotNumber, int plane)
{
    lbm3_sim *sim = grids[0].sim;
    lbm3_bounds b{};
    if (Bounds) b = bounds_to(*Bounds);
    int dims[2] = {0, 0};
    check(lbm3_section_cut(sim, plane, cut, &b, nullptr, 0, dims), "section cut");
    std::vector<float> px((size_t)dims[0] * (size_t)dims[1] * 6);
    check(lbm3_section_cut(sim, plane, cut, &b, px.data(), (long long)px.size(), dims), "section cut");
    FILE *fp = std::fopen("/dev/shm/sim_data.bin", "wb");
    if (!fp) {
        std::perror("Section cut: cannot open /dev/shm/sim_data.bin");
        return;
    }
    const int header[4] = {plotNumber, dims[0], dims[1], 6};
    std::fwrite(header, sizeof(int), 4, fp);
    std::fwrite(px.data(), sizeof(float), px.size(), fp);
    std::fclose(fp);
}

} // namespace lbm3facade

inline void exportSectionCutPlotXY(std::vector<GridStruct> &grids, BoundsStruct &Bounds, RotorInfoStruct &, const float &zCut,
                                   const int &plotNumber)
{
    std::cout << "Exporting XY section cut plot " << plotNumber << " ... " << std::flush;
    lbm3facade::section_cut(grids, &Bounds, zCut, plotNumber, LBM3_PLANE_XY);
}
inline void exportSectionCutPlotZY(std::vector<GridStruct> &grids, BoundsStruct &Bounds, RotorInfoStruct &, const float &xCut,
                                   const int &plotNumber)
{
    std::cout << "Exporting ZY section cut plot " << plotNumber << " ... " << std::flush;
    lbm3facade::section_cut(grids, &Bounds, xCut, plotNumber, LBM3_PLANE_ZY);
}
inline void exportSectionCutPlotZX(std::vector<GridStruct> &grids, BoundsStruct &Bounds, RotorInfoStruct &, const float &yCut,
                                   const int &plotNumber)
{
    std::cout << "Exporting ZX section cut plot " << plotNumber << " ... " << std::flush;
    lbm3facade::section_cut(grids, &Bounds, yCut, plotNumber, LBM3_PLANE_ZX);
}
inline void exportSectionCutPlotXY(std::vector<GridStruct> &grids, const float &zCut, const int &plotNumber)
{
    BoundsStruct B; RotorInfoStruct R;
    exportSectionCutPlotXY(grids, B, R, zCut, plotNumber);
}
inline void exportSectionCutPlotZY(std::vector<GridStruct> &grids, const float &xCut, const int &plotNumber)
{
    BoundsStruct B; RotorInfoStruct R;
    exportSectionCutPlotZY(grids, B, R, xCut, plotNumber);
}
inline void exportSectionCutPlotZX(std::vector<GridStruct> &grids, const float &yCut, const int &plotNumber)
{
    BoundsStruct B; RotorInfoStruct R;
    exportSectionCutPlotZX(grids, B, R, yCut, plotNumber);
}
inline void exportSectionCutPlotXY(std::vector<GridStruct> &grids, BoundsStruct &Bounds, const float &zCut, const int &plotNumber)
{
    RotorInfoStruct R;
    exportSectionCutPlotXY(grids, Bounds, R, zCut, plotNumber);
}
inline void exportSectionCutPlotZY(std::vector<GridStruct> &grids, BoundsStruct &Bounds, const float &xCut, const int &plotNumber)
{
    RotorInfoStruct R;
    exportSectionCutPlotZY(grids, Bounds, R, xCut, plotNumber);
}
inline void exportSectionCutPlotZX(std::vector<GridStruct> &grids, BoundsStruct &Bounds, const float &yCut, const int &plotNumber)
{
    RotorInfoStruct R;
    exportSectionCutPlotZX(grids, Bounds, R, yCut, plotNumber);
}
inline void exportSectionCutPlotXY(std::vector<GridStruct> &grids, RotorInfoStruct &RotorInfo, const float &zCut, const int &plotNumber)
{
    BoundsStruct B;
    exportSectionCutPlotXY(grids, B, RotorInfo, zCut, plotNumber);
}
inline void exportSectionCutPlotZY(std::vector<GridStruct> &grids, RotorInfoStruct &RotorInfo, const float &xCut, const int &plotNumber)
{
    BoundsStruct B;
    exportSectionCutPlotZY(grids, B, RotorInfo, xCut, plotNumber);
}
inline void exportSectionCutPlotZX(std::vector<GridStruct> &grids, RotorInfoStruct &RotorInfo, const float &yCut, const int &plotNumber)
{
    BoundsStruct B;
    exportSectionCutPlotZX(grids, B, RotorInfo, yCut, plotNumber);
}
