// toiletPaperPlot{X,Y,Z} (reference include/plotter/exportToiletPaperPlot.h):
// unrolled cylindrical section plots.  Not part of this build's observables;
// the overloads exist so case files that include the header compile, and a
// call reports that it is not available.
#pragma once
#include "../types.h"

namespace lbm3facade {
inline void toilet_paper_unavailable(const char *axis)
{
    std::cerr << "toiletPaperPlot" << axis << ": cylindrical section plots are not available in this build; skipped.\n";
}
} // namespace lbm3facade

#define LBM3_TOILET_PAPER(AXIS)                                                                                            \
    inline void toiletPaperPlot##AXIS(std::vector<GridStruct> &, const BoundsStruct &, const RotorInfoStruct &,          \
                                      const float &, const int &) { lbm3facade::toilet_paper_unavailable(#AXIS); }       \
    inline void toiletPaperPlot##AXIS(std::vector<GridStruct> &, const float &, const int &)                             \
    { lbm3facade::toilet_paper_unavailable(#AXIS); }                                                                     \
    inline void toiletPaperPlot##AXIS(std::vector<GridStruct> &, const BoundsStruct &, const float &, const int &)       \
    { lbm3facade::toilet_paper_unavailable(#AXIS); }                                                                     \
    inline void toiletPaperPlot##AXIS(std::vector<GridStruct> &, const RotorInfoStruct &, const float &, const int &)    \
    { lbm3facade::toilet_paper_unavailable(#AXIS); }
LBM3_TOILET_PAPER(X)
LBM3_TOILET_PAPER(Y)
LBM3_TOILET_PAPER(Z)
#undef LBM3_TOILET_PAPER
