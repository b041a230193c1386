// Result layouts of Kokkos::subview, checked at compile time against the
// reference's rules (core/src/View/Kokkos_ViewMapping.hpp:85-170,
// 3215-3290): LayoutLeft survives when the kept dimensions start at
// dimension 0, LayoutRight when they end at the last one, and any rank-1 or
// rank-2 result whose unit-stride dimension is kept keeps the layout.
#include <Kokkos_Core.hpp>
using E = Kokkos::MI355X;
template <class A, class B> constexpr bool same = std::is_same<A, B>::value;
int main() {
  using L2 = Kokkos::View<double**, Kokkos::LayoutLeft, E>;
  using R2 = Kokkos::View<double**, Kokkos::LayoutRight, E>;
  using R3 = Kokkos::View<double***, Kokkos::LayoutRight, E>;
  using L3 = Kokkos::View<double***, Kokkos::LayoutLeft, E>;
  using P = Kokkos::pair<int, int>;
  using ALL = Kokkos::ALL_t;
  static_assert(same<Kokkos::Subview<R2, int, ALL>::array_layout, Kokkos::LayoutRight>);
  static_assert(same<Kokkos::Subview<R2, ALL, int>::array_layout, Kokkos::LayoutStride>);
  static_assert(same<Kokkos::Subview<L2, ALL, int>::array_layout, Kokkos::LayoutLeft>);
  static_assert(same<Kokkos::Subview<L2, int, ALL>::array_layout, Kokkos::LayoutStride>);
  static_assert(same<Kokkos::Subview<L2, P, ALL>::array_layout, Kokkos::LayoutLeft>);
  static_assert(same<Kokkos::Subview<R2, ALL, P>::array_layout, Kokkos::LayoutRight>);
  static_assert(same<Kokkos::Subview<R3, int, P, ALL>::array_layout, Kokkos::LayoutRight>);
  static_assert(same<Kokkos::Subview<R3, P, int, ALL>::array_layout, Kokkos::LayoutRight>);  // rank 2, last kept
  static_assert(same<Kokkos::Subview<R3, P, ALL, ALL>::array_layout, Kokkos::LayoutRight>);
  static_assert(same<Kokkos::Subview<R3, ALL, P, ALL>::array_layout, Kokkos::LayoutStride>);
  static_assert(same<Kokkos::Subview<L3, P, P, int>::array_layout, Kokkos::LayoutLeft>);
  static_assert(same<Kokkos::Subview<L3, ALL, int, ALL>::array_layout, Kokkos::LayoutLeft>);  // rank 2, first kept
  static_assert(same<Kokkos::Subview<L3, int, ALL, ALL>::array_layout, Kokkos::LayoutStride>);
  static_assert(same<Kokkos::Subview<L3, int, int, int>::array_layout, Kokkos::LayoutLeft>);
  static_assert(Kokkos::Subview<L3, int, int, int>::rank == 0);
  static_assert(Kokkos::Subview<R3, int, P, ALL>::rank == 2);
  // rank and rank_dynamic are usable as values and as calls
  static_assert(R3::rank() == 3 && R3::rank == 3 && R3::rank_dynamic() == 3);
  static_assert(Kokkos::View<double* [4], E>::rank() == 2 && Kokkos::View<double* [4], E>::rank_dynamic == 1);
  // type names, and MemoryTraits<> as the default traits
  static_assert(Kokkos::Impl::TypeInfo<int>::name() == "int");
  static_assert(Kokkos::Impl::TypeInfo<double[3]>::name() == "double[3]");
  static_assert(Kokkos::Impl::TypeInfo<Kokkos::LayoutLeft>::name() == "Kokkos::LayoutLeft");
  static_assert(Kokkos::MemoryTraits<>::impl_value == 0 && Kokkos::MemoryTraits<>::is_unmanaged == false);
  return 0;
}
