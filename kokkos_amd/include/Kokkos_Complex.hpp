// Kokkos::complex (reference core/src/Kokkos_Complex.hpp), part of this
// execution space's Kokkos_Core.hpp, which this header includes.
#ifndef KKMX_FWD_KOKKOS_COMPLEX_HPP
#define KKMX_FWD_KOKKOS_COMPLEX_HPP

#include <Kokkos_Core.hpp>

#endif
