// Launch requests shared by the C-ABI layer and the per-element-type instantiation units.
#pragma once
#include <type_traits>

#include "femhot_kernels.h"
#include "femhot_rows.h"

namespace femhot {

struct ElemLaunch {
  int dtype, ph, ngp, mode;
  bool plain;
  dim3 grid;
  hipStream_t stream;
  KArgs args;
};

struct TangentLaunch {
  int dtype, ph, ngp;
  bool plain;
  bool rows;  // row-owner assembly (k_tangent_rows, one 64-lane workgroup per node); grid = nodes
  dim3 grid;
  hipStream_t stream;
  TArgs args;
};

// Chunk (LDS-staged, persistent) launch: see femhot_chunk.h
struct ChunkArgs;
struct ChunkLaunch;

template <int ET> int launch_elem(const ElemLaunch &L);
template <int ET> int launch_tangent(const TangentLaunch &L);

// Calls f(R*, integral_constant<PH>, integral_constant<NGP>) for the runtime (dtype, ph, ngp).
template <class R, class F> int dispatch_ph_ngp(int ph, int ngp, F &&f) {
  using std::integral_constant;
  switch (ph * 4 + ngp) {
    case THERMAL * 4 + 1: return f((R *)nullptr, integral_constant<int, THERMAL>{}, integral_constant<int, 1>{});
    case THERMAL * 4 + 2: return f((R *)nullptr, integral_constant<int, THERMAL>{}, integral_constant<int, 2>{});
    case THERMAL * 4 + 3: return f((R *)nullptr, integral_constant<int, THERMAL>{}, integral_constant<int, 3>{});
    case MECH * 4 + 1: return f((R *)nullptr, integral_constant<int, MECH>{}, integral_constant<int, 1>{});
    case MECH * 4 + 2: return f((R *)nullptr, integral_constant<int, MECH>{}, integral_constant<int, 2>{});
    case MECH * 4 + 3: return f((R *)nullptr, integral_constant<int, MECH>{}, integral_constant<int, 3>{});
    case NEOHOOKE * 4 + 1: return f((R *)nullptr, integral_constant<int, NEOHOOKE>{}, integral_constant<int, 1>{});
    case NEOHOOKE * 4 + 2: return f((R *)nullptr, integral_constant<int, NEOHOOKE>{}, integral_constant<int, 2>{});
    case NEOHOOKE * 4 + 3: return f((R *)nullptr, integral_constant<int, NEOHOOKE>{}, integral_constant<int, 3>{});
    default: set_error("unsupported physics / Gauss rule"); return FEMHOT_ERR_UNSUPPORTED;
  }
}

template <class F> int dispatch_r_ph_ngp(int dtype, int ph, int ngp, F &&f) {
  if (dtype == FEMHOT_F32) return dispatch_ph_ngp<float>(ph, ngp, f);
  return dispatch_ph_ngp<double>(ph, ngp, f);
}

template <int ET> int launch_elem_impl(const ElemLaunch &L) {
  return dispatch_r_ph_ngp(L.dtype, L.ph, L.ngp, [&](auto rp, auto phc, auto ngpc) -> int {
    using R = std::remove_pointer_t<decltype(rp)>;
    constexpr int PH = decltype(phc)::value, NGP = decltype(ngpc)::value;
    void (*k)(KArgs) = nullptr;
    if (L.mode == MODE_RES)
      k = L.plain ? k_elem<R, ET, PH, NGP, MODE_RES, true> : k_elem<R, ET, PH, NGP, MODE_RES, false>;
    else if (L.mode == MODE_VJP)
      k = L.plain ? k_elem<R, ET, PH, NGP, MODE_VJP, true> : k_elem<R, ET, PH, NGP, MODE_VJP, false>;
    else
      k = L.plain ? k_elem<R, ET, PH, NGP, MODE_EGRAD, true> : k_elem<R, ET, PH, NGP, MODE_EGRAD, false>;
    hipLaunchKernelGGL(k, L.grid, dim3(256), 0, L.stream, L.args);
    FH_HIP(hipGetLastError());
    return FEMHOT_OK;
  });
}

template <int ET> int launch_tangent_impl(const TangentLaunch &L) {
  return dispatch_r_ph_ngp(L.dtype, L.ph, L.ngp, [&](auto rp, auto phc, auto ngpc) -> int {
    using R = std::remove_pointer_t<decltype(rp)>;
    constexpr int PH = decltype(phc)::value, NGP = decltype(ngpc)::value;
    if (L.rows) {
      if constexpr (PH == NEOHOOKE) {
        set_error("femhot_tangent_csr: the row-owner assembly covers thermal and linear-elastic physics");
        return FEMHOT_ERR_UNSUPPORTED;
      } else {
        hipLaunchKernelGGL((k_tangent_rows<R, ET, PH, NGP>), L.grid, dim3(64), 0, L.stream, L.args);
        FH_HIP(hipGetLastError());
        return FEMHOT_OK;
      }
    }
    void (*k)(TArgs) = L.plain ? k_tangent<R, ET, PH, NGP, true> : k_tangent<R, ET, PH, NGP, false>;
    hipLaunchKernelGGL(k, L.grid, dim3(256), 0, L.stream, L.args);
    FH_HIP(hipGetLastError());
    return FEMHOT_OK;
  });
}

}  // namespace femhot
