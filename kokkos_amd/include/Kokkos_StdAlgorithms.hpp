// Kokkos::Experimental std-algorithm scans and reductions on the MI355X
// execution space (reference algorithms/src/Kokkos_StdAlgorithms.hpp).
#ifndef KKMX_KOKKOS_STD_ALGORITHMS_HPP
#define KKMX_KOKKOS_STD_ALGORITHMS_HPP

#include <Kokkos_Core.hpp>
#include <kkmx/StdAlgorithms.hpp>

#endif
