// Host atomics on a 16-byte type from two threads: one adds, the other
// exchanges and compare-exchanges the same Kokkos::complex<double>.  Every
// operation on one address must take the same lock, so the final value is
// exact: the adder's total plus what the other thread put back.
#include <Kokkos_Core.hpp>

#include <cstdio>
#include <thread>

int main() {
  Kokkos::complex<double> z(0.0, 0.0);
  constexpr int n = 200000;
  double swapped_out = 0.0;  // what thread B took out of z by exchange
  std::thread a([&] {
    for (int i = 0; i < n; ++i) Kokkos::atomic_fetch_add(&z, Kokkos::complex<double>(1.0, -1.0));
  });
  std::thread b([&] {
    for (int i = 0; i < n; ++i) {
      const Kokkos::complex<double> old = Kokkos::atomic_exchange(&z, Kokkos::complex<double>(0.0, 0.0));
      swapped_out += old.real();
      if (old.real() != -old.imag()) {
        std::printf("torn value %g %g\n", old.real(), old.imag());
        std::exit(1);
      }
      // put it back through a compare-exchange that must succeed unless the
      // adder got in between, then by exchange-and-add
      const Kokkos::complex<double> zero(0.0, 0.0);
      const Kokkos::complex<double> seen = Kokkos::atomic_compare_exchange(&z, zero, old);
      if (seen.real() == 0.0 && seen.imag() == 0.0) {
        swapped_out -= old.real();
      }
    }
  });
  a.join();
  b.join();
  const double total = z.real() + swapped_out;
  const Kokkos::complex<double> last = Kokkos::atomic_load(&z);
  if (total != double(n) || z.imag() != -z.real() || last.real() != z.real()) {
    std::printf("lost updates: z = (%g, %g), taken out %g, want total %d\n", z.real(), z.imag(), swapped_out, n);
    return 1;
  }
  std::printf("ok %g\n", total);
  return 0;
}
