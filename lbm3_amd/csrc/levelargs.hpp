// Kernel argument blocks of one grid level, built from the level's buffers.
#pragma once
#include "sim.hpp"

namespace lbm3 {

inline Topology topo_of(const Level &l)
{
    return Topology{l.N, l.shifter.p, l.ri.p, l.rj.p, l.rk.p, l.rjp.p, l.rkp.p, l.rjkp.p};
}

template <class S> inline StepArgs<S> step_args(Sim &s, Level &l)
{
    StepArgs<S> a;
    a.f = reinterpret_cast<S *>(l.f.p);
    a.topo = topo_of(l);
    a.fast = FastTopo{l.N, l.cellRow.p, l.rowTab.p, l.bulkSkip.p, l.tileRow.p, l.nRows,
                      l.openInBulk ? l.openBits.p : nullptr, l.rotorInBulk ? l.rotorBits.p : nullptr};
    a.wallMap = l.wallMap.p;
    a.wallIndex = l.wallIndex.p;
    a.wallData = l.wallData.p;
    a.linkLength = l.linkLength.p;
    a.W = l.W;
    a.gWall = s.C.trackWallForce ? l.gWall.p : nullptr;
    a.nu = l.info.nu;
    a.bulk = l.bulk;
    a.wallBC = l.wallBC.p;
    a.wallVel = l.wallVel.p;
    a.param = nullptr;
    a.xcdRemap = s.tune.xcdRemap;
    a.bulkLds = s.tune.bulkLds;
    a.laneRot = l.openInBulk && l.laneRot ? 1 : 0;
    a.wallBlocks = 0;
    a.geom = l.geom;
    a.cd = s.codec;
    a.open = FusedOpen<typename StoreTraits<S>::compute>{};
    a.rotors = nullptr;
    a.nRotors = 0;
    a.trackRotor = 0;
    a.rotorCells = nullptr;
    a.rotorCellCount = 0;
    for (double &v : a.rotorCS) v = 0.0;
    return a;
}

template <class S> inline OpenArgs<S> open_args(Sim &s, Level &l, int iter)
{
    using T = typename StoreTraits<S>::compute;
    OpenArgs<S> a;
    a.f = reinterpret_cast<S *>(l.f.p);
    a.topo = topo_of(l);
    a.count = l.nOpen;
    a.cell = l.openCell.p;
    a.nb = l.openNb.p;
    a.meta = l.openMeta.p;
    a.records = l.openRecords.p;
    a.scale = l.openScale.p;
    T *refs = reinterpret_cast<T *>(l.openRefs.p);
    a.dRhoRef = refs;
    a.uNormalRef = refs + (l.nOpen > 0 ? l.nOpen : 1);
    a.dRhoCum = s.C.trackOpenBoundaries ? l.openCum.p : nullptr;
    a.uNormalCum = s.C.trackOpenBoundaries ? l.openCum.p + (l.nOpen > 0 ? l.nOpen : 1) : nullptr;
    a.iterationsFinished = iter;
    a.cd = s.codec;
    a.param = nullptr;
    a.list = nullptr;
    return a;
}

} // namespace lbm3
