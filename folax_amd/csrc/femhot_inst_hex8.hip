// Kernel instantiations for element type HEX8 (one translation unit per element type so the
// build compiles them in parallel).
#include "femhot_launch.h"

namespace femhot {
template <> int launch_elem<HEX8>(const ElemLaunch &L) { return launch_elem_impl<HEX8>(L); }
template <> int launch_tangent<HEX8>(const TangentLaunch &L) { return launch_tangent_impl<HEX8>(L); }
}  // namespace femhot
