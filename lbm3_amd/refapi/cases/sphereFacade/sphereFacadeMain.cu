// Flow around a small sphere written as an LBM3 case file: constants, the four
// device hooks, the reference's include order and a main() that builds the
// grid, steps it with updateAllGrids, tracks the wall force every step and
// reduces the kinetic energy with TNL::Algorithms::reduce, as
// cases/flowAroundSphere/flowAroundSphereMain.cu and
// cases/taylorGreenVortex/taylorGreenVortexMain.cu do.
//
//   sphereFacadeMain <sphere.stl> <iterations> <output prefix>
//
// writes <prefix>.macros (dRho, ux, uy, uz per cell of level 0, float32, cell
// order, read through fArray.getConstView()), <prefix>.txt (iterations, kinetic
// energy, and per iteration Fx Fy Fz of wall 0 from the tracker, the drag
// coefficient) and <prefix>.cut (the XY section-cut file).

constexpr float reynoldsNumber = 100.f;
constexpr float uxInlet = 0.05f;
constexpr int ITERATION_COUNT = 100000;
constexpr int TRACKER_PERIOD = 1;
constexpr bool TRACK_WALL_FORCE = true;
constexpr bool TRACK_ROTOR_FORCE = false;
constexpr bool TRACK_OPEN_BOUNDARIES = true;

constexpr float sphereDiameterPhys = 6.f;                     // mm
constexpr float uxInletPhys = 1.f;                            // m/s
constexpr int GRID_LEVEL_COUNT = 1;
constexpr int WALL_REFINEMENT_COUNT = 0;
constexpr float RES_GLOBAL = 1.f;                             // mm
constexpr float NU_PHYS = uxInletPhys * (sphereDiameterPhys / 1000.f) / reynoldsNumber;
constexpr float RHO_PHYS = 1.f;
constexpr float DT_PHYS_GLOBAL = (uxInlet / uxInletPhys) * (RES_GLOBAL / 1000.f);

#include "../../include/types.h"
#include "../../include/STLFunctions.h"
#include "../../include/voxelizerFunctions.h"
#include "../../include/cellFunctions.h"

__cuda_callable__ void getRefinementModifier(const int &iCell, const int &jCell, const int &kCell, bool &refinementMarker,
                                             const InfoStruct &Info)
{
    refinementMarker = false;
}

__cuda_callable__ void getInitialCondition(BCStruct &BC, const int &iCell, const int &jCell, const int &kCell, const InfoStruct &Info)
{
    BC.ux = uxInlet;
}

__cuda_callable__ void getOpenBC(BCStruct &BC, const int &iCell, const int &jCell, const int &kCell, const InfoStruct &Info)
{
    const bool interior = jCell != 0 && jCell != Info.cellCountY - 1 && kCell != 0 && kCell != Info.cellCountZ - 1;
    if (iCell == Info.cellCountX - 1 && interior) {          // outlet
        BC.dirichletRho = true;
        BC.dRho = 0.f;
        BC.openBCID = 1;
        BC.nonReflective = true;
    } else if (iCell == 0 && interior) {                     // inlet
        BC.dirichletU = true;
        BC.ux = uxInlet;
        BC.nonReflective = false;
        BC.openBCID = 0;
    } else {                                                 // far field
        BC.dirichletU = true;
        BC.ux = uxInlet;
        BC.nonReflective = false;
        BC.openBCID = 2;
    }
}

__cuda_callable__ void getLocalBC(BCStruct &BC, const int &iCell, const int &jCell, const int &kCell, const InfoStruct &Info)
{
    BC.collisionLimiter = 0.1f;
    if (BC.wallID == 0) {
        BC.ux = 0.f;
        BC.uy = 0.f;
        BC.uz = 0.f;
    }
    if (Info.gridID == 0) BC.collisionLimiter = 0.f;        // K15 where the open boundaries are
}

#include "../../include/gridBuilderFunctions.h"
#include "../../include/updateGrid.h"
#include "../../include/trackerFunctions.h"
#include "../../include/plotter/exportSectionCutPlot.h"
#include "../../include/plotter/plotTracker.h"

double getKineticEnergy(const std::vector<GridStruct> &grids)
{
    double energy = 0.0;
    for (const GridStruct &Grid : grids) {
        const InfoStruct Info = Grid.Info;
        auto fView = Grid.fArray.getConstView();
        auto wallMapView = Grid.Wall.wallMapArray.getConstView();
        auto wallDataView = Grid.Wall.wallDataArray.getConstView();
        auto shifterView = Grid.IJKNBR.shifterArray.getConstView();
        auto iView = Grid.IJKNBR.iArray.getConstView();
        auto jView = Grid.IJKNBR.jArray.getConstView();
        auto kView = Grid.IJKNBR.kArray.getConstView();
        auto jPlusView = Grid.IJKNBR.jPlusArray.getConstView();
        auto kPlusView = Grid.IJKNBR.kPlusArray.getConstView();
        auto jkPlusView = Grid.IJKNBR.jkPlusArray.getConstView();
        auto energyAtCell = [=] __cuda_callable__(const int cell) mutable -> double {
            const int wallMap = wallMapView(cell);
            if (wallMap == -3 || wallMap == -2) return 0.0;
            if (wallMap >= 0 && (wallDataView(wallMap) & 1u)) return 0.0;
            int iCell, jCell, kCell;
            NBRStruct NBR;
            getCompressedIJKNBR(cell, iCell, jCell, kCell, NBR, shifterView, iView, jView, kView, jPlusView, kPlusView,
                                jkPlusView, Info);
            float f[27];
            int cellReadIndex[27], fReadIndex[27];
            getPreCollisionIndex(cellReadIndex, fReadIndex, NBR, Info.esotwistFlipper);
            for (int q = 0; q < 27; ++q) f[q] = fView(fReadIndex[q], cellReadIndex[q]);
            float dRho, ux, uy, uz;
            getDRhoUxUyUz(dRho, ux, uy, uz, f);
            return 0.5 * (1.0 + double(dRho)) * (double(ux) * ux + double(uy) * uy + double(uz) * uz);
        };
        energy += TNL::Algorithms::reduce<TNL::Devices::Cuda>(0, Info.cellCount, energyAtCell, TNL::Plus{}, 0.0);
    }
    return energy;
}

// (dRho, ux, uy, uz) of every cell of a level, read through the case-facing views
void writeMacros(const GridStruct &Grid, const std::string &path)
{
    const InfoStruct Info = Grid.Info;
    auto fView = Grid.fArray.getConstView();
    auto shifterView = Grid.IJKNBR.shifterArray.getConstView();
    auto iView = Grid.IJKNBR.iArray.getConstView();
    auto jView = Grid.IJKNBR.jArray.getConstView();
    auto kView = Grid.IJKNBR.kArray.getConstView();
    auto jPlusView = Grid.IJKNBR.jPlusArray.getConstView();
    auto kPlusView = Grid.IJKNBR.kPlusArray.getConstView();
    auto jkPlusView = Grid.IJKNBR.jkPlusArray.getConstView();
    FloatArrayType macros(4 * (size_t)Info.cellCount);
    auto out = macros.getView();
    auto cellLambda = [=] __cuda_callable__(const int cell) mutable {
        int iCell, jCell, kCell;
        NBRStruct NBR;
        getCompressedIJKNBR(cell, iCell, jCell, kCell, NBR, shifterView, iView, jView, kView, jPlusView, kPlusView, jkPlusView,
                            Info);
        float f[27];
        int cellReadIndex[27], fReadIndex[27];
        getPreCollisionIndex(cellReadIndex, fReadIndex, NBR, Info.esotwistFlipper);
        for (int q = 0; q < 27; ++q) f[q] = fView(fReadIndex[q], cellReadIndex[q]);
        float dRho, ux, uy, uz;
        getDRhoUxUyUz(dRho, ux, uy, uz, f);
        out(4 * (long long)cell) = dRho;
        out(4 * (long long)cell + 1) = ux;
        out(4 * (long long)cell + 2) = uy;
        out(4 * (long long)cell + 3) = uz;
    };
    TNL::Algorithms::parallelFor<TNL::Devices::Cuda>(0, Info.cellCount, cellLambda);
    FloatArrayTypeCPU host = macros;
    std::ofstream file(path, std::ios::binary);
    file.write(reinterpret_cast<const char *>(host.getData()), sizeof(float) * 4 * (size_t)Info.cellCount);
}

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::cerr << "usage: " << argv[0] << " <sphere.stl> <iterations> <output prefix>" << std::endl;
        return EXIT_FAILURE;
    }
    const std::string prefix = argv[3];
    const int iterations = std::atoi(argv[2]);
    std::vector<STLStruct> gridStaticSTLs(1);
    readSTL(gridStaticSTLs[0], argv[1]);
    std::vector<STLStruct> rotorSTLs(0);
    std::vector<GridStruct> grids(GRID_LEVEL_COUNT);
    BoundsStruct DomainBounds;
    DomainBounds.xMin = 0.f;
    DomainBounds.xMax = 48.f;
    DomainBounds.yMin = 0.f;
    DomainBounds.yMax = 24.f;
    DomainBounds.zMin = 0.f;
    DomainBounds.zMax = 24.f;
    const long long fluidUpdatesPerIteration = buildGrids(grids, gridStaticSTLs, rotorSTLs, DomainBounds);

    TrackerStruct Tracker;
    Tracker.TRACK_CUSTOM_VARIABLES = true;
    Tracker.customNames = {"Sphere Drag Coefficient"};
    Tracker.customUnits = {"[1]"};
    initializeTracker(Tracker, grids);

    std::ofstream log(prefix + ".txt");
    log.precision(17);
    log << "fluid_updates " << fluidUpdatesPerIteration << "\n";
    TNL::Timer lapTimer;
    lapTimer.reset();
    lapTimer.start();
    for (int iterationsFinished = 1; iterationsFinished <= iterations; iterationsFinished++) {
        updateAllGrids(grids, 0);
        if (iterationsFinished % TRACKER_PERIOD == 0) {
            updateTracker(Tracker, grids);
            const float drag = -Tracker.wallFx[0];
            const float d = sphereDiameterPhys / 1000.f;
            const float dragCoefficient = (8.f * drag) / (RHO_PHYS * uxInletPhys * uxInletPhys * 3.14159f * d * d);
            trackCustomVariables(Tracker, {dragCoefficient});
            log << "force " << iterationsFinished << " " << Tracker.wallFx[0] << " " << Tracker.wallFy[0] << " "
                << Tracker.wallFz[0] << " " << dragCoefficient << "\n";
        }
    }
    lapTimer.stop();
    log << "iterations " << grids[0].Info.iterationsFinished << "\n";
    log << "kinetic_energy " << getKineticEnergy(grids) << "\n";
    log << "glups " << double(fluidUpdatesPerIteration) * iterations / lapTimer.getRealTime() / 1e9 << "\n";
    writeMacros(grids[0], prefix + ".macros");
    exportSectionCutPlotXY(grids, 12.f, 1);
    std::ifstream cut("/dev/shm/sim_data.bin", std::ios::binary);
    std::ofstream cutCopy(prefix + ".cut", std::ios::binary);
    cutCopy << cut.rdbuf();
    std::cout << "done" << std::endl;
    return EXIT_SUCCESS;
}
