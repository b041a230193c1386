// Kokkos::sort, Kokkos::Experimental::sort_by_key, Kokkos::BinSort,
// BinOp1D / BinOp3D and the team/thread-level sorts on the MI355X execution
// space (reference algorithms/src/Kokkos_Sort.hpp).
#ifndef KKMX_KOKKOS_SORT_HPP
#define KKMX_KOKKOS_SORT_HPP

#include <Kokkos_Core.hpp>
#include <kkmx/Sort.hpp>
#include <kkmx/NestedSort.hpp>

#endif
