// Kokkos_Core_fwd.hpp under the reference's name (core/src/Kokkos_Core_fwd.hpp): its contents are
// part of this execution space's Kokkos_Core.hpp, which this header includes.
#ifndef KKMX_FWD_KOKKOS_CORE_FWD_HPP
#define KKMX_FWD_KOKKOS_CORE_FWD_HPP

#include <Kokkos_Core.hpp>

#endif
