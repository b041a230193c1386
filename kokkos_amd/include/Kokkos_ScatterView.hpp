// Kokkos::Experimental::ScatterView on the MI355X execution space
// (reference containers/src/Kokkos_ScatterView.hpp).
#ifndef KKMX_KOKKOS_SCATTER_VIEW_HPP
#define KKMX_KOKKOS_SCATTER_VIEW_HPP

#include <Kokkos_Core.hpp>
#include <kkmx/ScatterView.hpp>

#endif
