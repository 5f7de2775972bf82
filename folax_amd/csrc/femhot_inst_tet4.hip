// Kernel instantiations for element type TET4 (one translation unit per element type so the
// build compiles them in parallel).
#include "femhot_launch.h"

namespace femhot {
template <> int launch_elem<TET4>(const ElemLaunch &L) { return launch_elem_impl<TET4>(L); }
template <> int launch_tangent<TET4>(const TangentLaunch &L) { return launch_tangent_impl<TET4>(L); }
}  // namespace femhot
