// Kokkos_Tuners.hpp under the reference's name (core/src/Kokkos_Tuners.hpp):
// its contents are part of this execution space's Kokkos_Core.hpp, which this
// header includes.
#ifndef KKMX_FWD_KOKKOS_TUNERS_HPP
#define KKMX_FWD_KOKKOS_TUNERS_HPP

#include <Kokkos_Core.hpp>

#endif
