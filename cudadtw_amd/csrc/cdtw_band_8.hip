// cdtw_band_8.hip -- exact-W instantiations of the banded cDTW kernels
// (cdtw_band.hpp, cdtw_pair.hpp), L1 metric, W in [113, 144] (one column at a
// time to W = 121, split over two waves beyond; the L2 instantiations of the
// same range are in cdtw_band_4.hip); split across units so they compile in
// parallel.
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

BandLaunchFn band_table_8(int w) { return band_pick_l1<113, 144>(w); }

}  // namespace mdtw
