// impl/Kokkos_Tools_Generic.hpp under the reference's name
// (core/src/impl/Kokkos_Tools_Generic.hpp): the tuning calculators and the
// built-in tuners are part of this execution space's Kokkos_Core.hpp, which
// this header includes.
#ifndef KKMX_FWD_KOKKOS_TOOLS_GENERIC_HPP
#define KKMX_FWD_KOKKOS_TOOLS_GENERIC_HPP

#include <Kokkos_Core.hpp>

#endif
