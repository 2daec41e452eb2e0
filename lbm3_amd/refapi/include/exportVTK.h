// exportVTK (reference include/exportVTK.h:17-158): every stored cell of every
// level as a vertex of a legacy ASCII VTK file, written by the library
// (lbm3_export_vtk).
#pragma once
#include "types.h"
#include "gridBuilderFunctions.h"

inline void exportVTK(const std::vector<GridStruct> &grids, const std::string &filename)
{
    lbm3facade::check(lbm3_export_vtk(grids[0].sim, filename.c_str()), "exportVTK");
}
