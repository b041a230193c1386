// Kokkos programming model on the MI355X execution space.
//
// Include this header instead of the reference's core/src/Kokkos_Core.hpp:
// user functors, lambdas and the public API (parallel_for / parallel_reduce /
// parallel_scan, RangePolicy / MDRangePolicy / TeamPolicy, View, atomic_*,
// Experimental::Graph, Tools) keep their names and meaning, and
// Kokkos::DefaultExecutionSpace is Kokkos::MI355X.  Compile with
//   hipcc --offload-arch=gfx950 -std=c++17 -I<repo>/kokkos_amd/include -I<repo>/include
// and link against libkkmx_rt.so (the C-ABI runtime).
#ifndef KOKKOS_CORE_HPP
#define KOKKOS_CORE_HPP

#include <Kokkos_Macros.hpp>
#include <kkmx/Core_fwd.hpp>
#include <kkmx/TypeInfo.hpp>
#include <kkmx/Math.hpp>
#include <kkmx/Complex.hpp>
#include <kkmx/Spaces.hpp>
#include <kkmx/PartitionSpace.hpp>
#include <kkmx/Atomic.hpp>
#include <kkmx/Desul.hpp>
#include <kkmx/View.hpp>
#include <kkmx/Subview.hpp>
#include <kkmx/Reducers.hpp>
#include <kkmx/Policy.hpp>
#include <kkmx/Tools.hpp>
#include <kkmx/FunctorAnalysis.hpp>
#include <kkmx/Wave.hpp>
#include <kkmx/KernelRange.hpp>
#include <kkmx/HostExec.hpp>
#include <kkmx/KernelMDRange.hpp>
#include <kkmx/KernelTeam.hpp>
#include <kkmx/TeamMDRange.hpp>
#include <kkmx/Parallel.hpp>
#include <kkmx/CopyViews.hpp>
#include <kkmx/LocalDeepCopy.hpp>
#include <kkmx/KokkosMalloc.hpp>
#include <kkmx/Graph.hpp>
#include <kkmx/Tuners.hpp>
#include <kkmx/SpacesImpl.hpp>

namespace Kokkos {
// Wall-clock timer with the reference's interface (core/src/Kokkos_Timer.hpp).
class Timer {
 public:
  Timer() { reset(); }
  void reset() { m_start = now(); }
  double seconds() const { return now() - m_start; }

 private:
  static double now() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return double(ts.tv_sec) + 1e-9 * double(ts.tv_nsec);
  }
  double m_start;
};
}  // namespace Kokkos

#endif
