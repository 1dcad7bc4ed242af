// cdtw_band_14.hip -- exact-W instantiations of the register-band cDTW kernel
// (cdtw_band.hpp) for W in [70, 79]; split across units so they compile in
// parallel.
#include "cdtw_band.hpp"

namespace mdtw {
namespace {
template <int LO, int HI>
BandLaunchFn band_pick(int w, bool l2) {
    if constexpr (LO > HI) {
        return nullptr;
    } else {
        if (w == LO) return l2 ? &band_launch<LO, true> : &band_launch<LO, false>;
        return band_pick<LO + 1, HI>(w, l2);
    }
}
}  // namespace

BandLaunchFn band_table_14(int w, bool l2) { return band_pick<70, 79>(w, l2); }

}  // namespace mdtw
