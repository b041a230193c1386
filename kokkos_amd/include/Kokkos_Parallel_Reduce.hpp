// Kokkos_Parallel_Reduce.hpp under the reference's name (core/src/Kokkos_Parallel_Reduce.hpp): its contents are
// part of this execution space's Kokkos_Core.hpp, which this header includes.
#ifndef KKMX_FWD_KOKKOS_PARALLEL_REDUCE_HPP
#define KKMX_FWD_KOKKOS_PARALLEL_REDUCE_HPP

#include <Kokkos_Core.hpp>

#endif
