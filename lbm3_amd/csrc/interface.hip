// Multilevel coarse<->fine interface transfer (reference include/updateInterface.h).
#include "sim.hpp"

namespace lbm3 {

void update_interface(Sim &s, int coarse)
{
    (void)s;
    (void)coarse;
}

} // namespace lbm3
