// Kokkos_Timer.hpp under the reference's name (core/src/Kokkos_Timer.hpp): its contents are
// part of this execution space's Kokkos_Core.hpp, which this header includes.
#ifndef KKMX_FWD_KOKKOS_TIMER_HPP
#define KKMX_FWD_KOKKOS_TIMER_HPP

#include <Kokkos_Core.hpp>

#endif
