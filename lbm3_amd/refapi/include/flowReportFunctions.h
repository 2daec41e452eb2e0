// getFlowReport{XY,ZY,ZX} (reference include/flowReportFunctions.h:11-372):
// the flow through a section, from the library's FP64 reduction over the
// section image (lbm3_flow_report).
#pragma once
#include "types.h"
#include "gridBuilderFunctions.h"

namespace lbm3facade {
inline void flow_report(FlowReportStruct &R, std::vector<GridStruct> &grids, BoundsStruct &Bounds, float cut, int plane)
{
    const lbm3_bounds b = bounds_to(Bounds);
    double o[7];
    check(lbm3_flow_report(grids[0].sim, plane, cut, &b, o), "flow report");
    R.normalVelocity = (float)o[0]; R.massFlow = (float)o[1]; R.momentumThrust = (float)o[2];
    R.pressure = (float)o[3]; R.pressurePower = (float)o[4]; R.normalKineticPower = (float)o[5]; R.areamm2 = (float)o[6];
}
} // namespace lbm3facade

inline void getFlowReportXY(FlowReportStruct &FlowReport, std::vector<GridStruct> &grids, BoundsStruct &Bounds, const float &zCut)
{
    lbm3facade::flow_report(FlowReport, grids, Bounds, zCut, LBM3_PLANE_XY);
}
inline void getFlowReportZY(FlowReportStruct &FlowReport, std::vector<GridStruct> &grids, BoundsStruct &Bounds, const float &xCut)
{
    lbm3facade::flow_report(FlowReport, grids, Bounds, xCut, LBM3_PLANE_ZY);
}
inline void getFlowReportZX(FlowReportStruct &FlowReport, std::vector<GridStruct> &grids, BoundsStruct &Bounds, const float &yCut)
{
    lbm3facade::flow_report(FlowReport, grids, Bounds, yCut, LBM3_PLANE_ZX);
}
