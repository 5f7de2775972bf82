// Kernel instantiations for element type QUAD4 (one translation unit per element type so the
// build compiles them in parallel).
#include "femhot_launch.h"

namespace femhot {
template <> int launch_elem<QUAD4>(const ElemLaunch &L) { return launch_elem_impl<QUAD4>(L); }
template <> int launch_tangent<QUAD4>(const TangentLaunch &L) { return launch_tangent_impl<QUAD4>(L); }
}  // namespace femhot
