// Kokkos::sort, Kokkos::BinSort, BinOp1D / BinOp3D on the MI355X execution
// space (reference algorithms/src/Kokkos_Sort.hpp).
#ifndef KKMX_KOKKOS_SORT_HPP
#define KKMX_KOKKOS_SORT_HPP

#include <Kokkos_Core.hpp>
#include <kkmx/Sort.hpp>

#endif
