// updateAllGrids / updateSingleGrid (reference include/updateGrid.h:16-396):
// one (sub-cycled) time step of every level from `level` down, run by the
// library (lbm3_update_all_grids: the fused collide-stream, bounceback and
// open-boundary kernels, the 2:1 sub-cycling and the interface transfers).
// Before it, the Info controls and rotor motions the case changed on the host
// are handed over (hooks evaluated again where they read them); after it, the
// levels' step state (iterations, esotwist parity, rotor angles) is read back.
#pragma once
#include "types.h"
#include "cellFunctions.h"
#include "boundaryConditions/interpolatedBouncebackFunctions.h"
#include "gridBuilderFunctions.h"

namespace lbm3facade {

inline void push_controls(std::vector<GridStruct> &grids)
{
    bool changed = false;
    for (GridStruct &G : grids) {
        if (!G.sim) continue;
        if (G.Info.iRegulatorInlet != G.sentIRegulatorInlet || G.Info.iRegulatorOutlet != G.sentIRegulatorOutlet ||
            G.Info.useRotors != G.sentUseRotors) {
            check(lbm3_set_info_controls(G.sim, G.level, G.Info.iRegulatorInlet, G.Info.iRegulatorOutlet, G.Info.useRotors),
                  "Info controls");
            G.sentIRegulatorInlet = G.Info.iRegulatorInlet;
            G.sentIRegulatorOutlet = G.Info.iRegulatorOutlet;
            G.sentUseRotors = G.Info.useRotors;
            changed = true;
        }
        for (size_t r = 0; r < G.rotors.size(); ++r) {
            RotorStruct &R = G.rotors[r];
            const int axis = R.Info.rotateAlongX ? LBM3_AXIS_X : (R.Info.rotateAlongY ? LBM3_AXIS_Y : (R.Info.rotateAlongZ ? LBM3_AXIS_Z : LBM3_AXIS_NONE));
            if (R.Info.radiansPerSecond != R.sentRadiansPerSecond || axis != R.sentAxis) {
                check(lbm3_set_rotor(G.sim, G.level, (int)r, R.Info.radiansPerSecond, axis, R.Info.ox, R.Info.oy, R.Info.oz),
                      "rotor motion");
                R.sentRadiansPerSecond = R.Info.radiansPerSecond;
                R.sentAxis = axis;
            }
        }
    }
    if (changed && !grids.empty() && grids[0].sim) check(lbm3_refresh_hooks(grids[0].sim), "hooks");
}

} // namespace lbm3facade

inline void updateAllGrids(std::vector<GridStruct> &grids, const int level)
{
    lbm3facade::push_controls(grids);
    lbm3facade::check(lbm3_update_all_grids(grids[level].sim, level), "updateAllGrids");
    lbm3facade::sync_step_state(grids);
}

inline void updateSingleGrid(GridStruct &Grid)
{
    lbm3facade::check(lbm3_update_single_grid(Grid.sim, Grid.level), "updateSingleGrid");
    lbm3_info I;
    lbm3facade::check(lbm3_get_info(Grid.sim, Grid.level, &I), "lbm3_get_info");
    Grid.Info.iterationsFinished = I.iterationsFinished;
    Grid.Info.iterationsSinceTrackerReset = I.iterationsSinceTrackerReset;
    Grid.Info.esotwistFlipper = I.esotwistFlipper != 0;
}
