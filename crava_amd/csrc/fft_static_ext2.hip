// Specialised transforms of the extended lengths, group 2 (see
// fft_static_impl.hpp): direct-I/O launch shapes, float, double and float
// storage with double arithmetic.
#include "fft_static_impl.hpp"

namespace crava {

CRAVA_EXT_LAUNCH_DEF(2)

}  // namespace crava
