// cdtw_band_9.hip -- exact-W instantiations of the banded cDTW kernels
// (cdtw_band.hpp, cdtw_pair.hpp), L1 metric, W in [145, 176] (split over two
// waves; the L2 instantiations of the same range are in cdtw_band_5.hip); split
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

BandLaunchFn band_table_9(int w) { return band_pick_l1<145, 176>(w); }

}  // namespace mdtw
