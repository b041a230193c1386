// Kokkos::Experimental::BatchDispatch, the batched uneven dispatch of the
// MI355X execution space.  It has no counterpart in the reference, so it is
// not part of Kokkos_Core.hpp: code that uses it includes this header.
#ifndef KKMX_KOKKOS_BATCH_HPP
#define KKMX_KOKKOS_BATCH_HPP

#include <Kokkos_Core.hpp>
#include <kkmx/Batch.hpp>

#endif
