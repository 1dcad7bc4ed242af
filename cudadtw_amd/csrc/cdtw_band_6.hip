// cdtw_band_6.hip -- exact-W instantiations of the banded cDTW kernels
// (cdtw_band.hpp, cdtw_pair.hpp), L2 metric, W in [177, 208] (split over two
// waves); split across units so they compile in parallel.
#include "cdtw_band.hpp"

namespace mdtw {
namespace {
template <int LO, int HI>
BandLaunchFn band_pick_l2(int w) {
    if constexpr (LO > HI) {
        return nullptr;
    } else {
        if (w == LO) return &band_launch<LO, true>;
        return band_pick_l2<LO + 1, HI>(w);
    }
}
}  // namespace

BandLaunchFn band_table_6(int w) { return band_pick_l2<177, 208>(w); }

}  // namespace mdtw
