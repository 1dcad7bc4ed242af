// cdtw_band_10.hip -- exact-W instantiations of the banded cDTW kernels
// (cdtw_band.hpp, cdtw_pair.hpp), L1 metric, W in [177, 208] (split over two
// waves; the L2 instantiations of the same range are in cdtw_band_6.hip); split
// across units so they compile in parallel.
#include "cdtw_band.hpp"

namespace mdtw {
namespace {
template <int LO, int HI>
BandLaunchFn band_pick_l1(int w) {
    if constexpr (LO > HI) {
        return nullptr;
    } else {
        if (w == LO) return &band_launch<LO, false>;
        return band_pick_l1<LO + 1, HI>(w);
    }
}
}  // namespace

BandLaunchFn band_table_10(int w) { return band_pick_l1<177, 208>(w); }

}  // namespace mdtw
