// The built-in tuners of the Kokkos Tools tuning interface: a sparse
// multidimensional tuning problem over nested std::map / std::vector
// spaces, and the team-size, MDRange-tile, RangePolicy-occupancy,
// single-dimension and categorical tuners built on it.
//
// Reference: core/src/Kokkos_Tuners.hpp:277-746 (public classes and their
// behaviour), core/src/impl/Kokkos_Tools_Generic.hpp:83-161 (the
// calculators that tell a tuner the legal team sizes and tile products).
//
// A tuning problem of dimension D asks the tool for D doubles in [0, 1)
// (output variables declared as ranges with step 0.999 / max size) and maps
// them to a point of the space: at each level the value at index
// floor(size * fraction) is taken, and for a map its sub-space at the same
// index is descended into.  The point is a std::tuple of one value per level.
#ifndef KKMX_TUNERS_HPP
#define KKMX_TUNERS_HPP

#include <array>
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

namespace Kokkos {
namespace Tools {
namespace Experimental {
namespace Impl {

// A search space as a tree of value lists: a vector is a leaf level, a map
// contributes its keys as one level and its values as the sub-spaces.
template <class Container>
struct SparseSpace;

template <class T, class A>
struct SparseSpace<std::vector<T, A>> {
  static constexpr int dims = 1;
  using point_type          = std::tuple<T>;
  std::vector<T> values;
  SparseSpace() = default;
  explicit SparseSpace(const std::vector<T, A>& v) : values(v.begin(), v.end()) {}
  point_type point(const double* f) const { return point_type(values[pick(values.size(), f[0])]); }
  static size_t pick(size_t n, double f) { return static_cast<size_t>(static_cast<double>(n) * f); }
};

template <class K, class V, class Cmp, class A>
struct SparseSpace<std::map<K, V, Cmp, A>> {
  using child               = SparseSpace<V>;
  static constexpr int dims = 1 + child::dims;
  using point_type = decltype(std::tuple_cat(std::declval<std::tuple<K>>(), std::declval<typename child::point_type>()));
  std::vector<K> keys;
  std::vector<child> children;
  SparseSpace() = default;
  explicit SparseSpace(const std::map<K, V, Cmp, A>& m) {
    for (const auto& kv : m) {
      keys.push_back(kv.first);
      children.emplace_back(kv.second);
    }
  }
  point_type point(const double* f) const {
    const size_t i = static_cast<size_t>(static_cast<double>(keys.size()) * f[0]);
    return std::tuple_cat(std::tuple<K>(keys[i]), children[i].point(f + 1));
  }
};

template <class Container>
struct get_space_dimensionality {
  static constexpr int value = SparseSpace<Container>::dims;
};

// The space of all tile shapes of an MDRange of rank N: nested maps of
// powers of two whose product stays below a bound.
template <class T, int N>
struct n_dimensional_sparse_structure {
  using type = std::map<T, typename n_dimensional_sparse_structure<T, N - 1>::type>;
};
template <class T>
struct n_dimensional_sparse_structure<T, 1> {
  using type = std::vector<T>;
};
template <class T>
void fill_tile(std::vector<T>& level, int bound) {
  for (int x = 1; x < bound; x *= 2) level.push_back(x);
}
template <class T, class Sub>
void fill_tile(std::map<T, Sub>& level, int bound) {
  for (int x = 1; x < bound; x *= 2) fill_tile(level[x], bound / x);
}

// What a tuner asks of the backend about the legal configurations
// (core/src/impl/Kokkos_Tools_Generic.hpp:83-161; the reference's tests name
// them through this nested namespace).
namespace Impl {
struct SimpleTeamSizeCalculator {
  template <class Policy, class Functor, class Tag>
  int get_max_team_size(const Policy& policy, const Functor& functor, const Tag tag) {
    return policy.team_size_max(functor, tag);
  }
  template <class Policy, class Functor, class Tag>
  int get_recommended_team_size(const Policy& policy, const Functor& functor, const Tag tag) {
    return policy.team_size_recommended(functor, tag);
  }
  template <class Policy, class Functor, class Tag>
  int get_mdrange_max_tile_size_product(const Policy& policy, const Functor& functor, const Tag&) {
    using exec_space = typename Policy::execution_space;
    return Kokkos::Impl::ParallelFor<Functor, Policy, exec_space>::max_tile_size_product(policy, functor);
  }
};
struct ComplexReducerSizeCalculator {
  template <class Policy, class Functor, class Tag>
  int get_max_team_size(const Policy& policy, const Functor& functor, const Tag tag) {
    return policy.team_size_max(functor, tag);
  }
  template <class Policy, class Functor, class Tag>
  int get_recommended_team_size(const Policy& policy, const Functor& functor, const Tag tag) {
    return policy.team_size_recommended(functor, tag);
  }
  template <class Policy, class Functor, class Tag>
  int get_mdrange_max_tile_size_product(const Policy& policy, const Functor& functor, const Tag&) {
    using exec_space = typename Policy::execution_space;
    return Kokkos::Impl::ParallelFor<Functor, Policy, exec_space>::max_tile_size_product(policy, functor);
  }
};
}  // namespace Impl
}  // namespace Impl

template <template <class...> class Container, size_t MaxDimensionSize = 100, class... TemplateArguments>
class MultidimensionalSparseTuningProblem {
 public:
  using ProblemSpaceInput                            = Container<TemplateArguments...>;
  using space_type                                   = Impl::SparseSpace<ProblemSpaceInput>;
  static constexpr int space_dimensionality          = space_type::dims;
  static constexpr size_t max_space_dimension_size   = MaxDimensionSize;
  static constexpr double tuning_min                 = 0.0;
  static constexpr double tuning_max                 = 0.999;
  double tuning_step                                 = tuning_max / max_space_dimension_size;
  using ValueArray = std::array<VariableValue, size_t(space_dimensionality)>;

  template <class Key>
  using extended_problem = MultidimensionalSparseTuningProblem<std::map, MaxDimensionSize, Key, ProblemSpaceInput>;

  MultidimensionalSparseTuningProblem() = default;
  MultidimensionalSparseTuningProblem(const ProblemSpaceInput& space, const std::vector<std::string>& names)
      : MultidimensionalSparseTuningProblem(space_type(space), names) {}
  MultidimensionalSparseTuningProblem(space_type space, const std::vector<std::string>& names)
      : m_space(std::move(space)), m_names(names) {
    for (size_t x = 0; x < names.size() && x < size_t(space_dimensionality); ++x) {
      VariableInfo info;
      info.type          = ValueType::kokkos_value_double;
      info.category      = StatisticalCategory::kokkos_value_interval;
      info.valueQuantity = CandidateValueType::kokkos_value_range;
      info.candidates    = make_candidate_range(tuning_min, tuning_max, tuning_step, true, true);
      m_ids[x]           = declare_output_type(names[x], info);
    }
  }

  // A new problem with one more level in front: every key of the new axis
  // leads to a copy of this space.
  template <class Key>
  extended_problem<Key> extend(const std::string& axis_name, const std::vector<Key>& axis) const {
    typename extended_problem<Key>::space_type s;
    for (const auto& k : axis) {
      s.keys.push_back(k);
      s.children.push_back(m_space);
    }
    std::vector<std::string> names{axis_name};
    names.insert(names.end(), m_names.begin(), m_names.end());
    return extended_problem<Key>(std::move(s), names);
  }

  template <class... Coordinates>
  auto get_point(Coordinates... coordinates) const {
    static_assert(sizeof...(Coordinates) == size_t(space_dimensionality),
                  "get_point takes one coordinate per dimension");
    const double f[] = {static_cast<double>(coordinates)...};
    return m_space.point(f);
  }

  // Opens a context, asks the tool for one value per dimension (0.0 when
  // no tool answers) and returns the point they select.
  auto begin() {
    m_context = get_new_context_id();
    ValueArray values;
    for (int x = 0; x < space_dimensionality; ++x) values[x] = make_variable_value(m_ids[x], 0.0);
    begin_context(m_context);
    request_output_values(m_context, space_dimensionality, values.data());
    double f[space_dimensionality];
    for (int x = 0; x < space_dimensionality; ++x) f[x] = values[x].value.double_value;
    return m_space.point(f);
  }
  void end() { end_context(m_context); }

 private:
  template <template <class...> class, size_t, class...>
  friend class MultidimensionalSparseTuningProblem;
  space_type m_space{};
  std::array<size_t, size_t(space_dimensionality)> m_ids{};
  std::vector<std::string> m_names;
  size_t m_context = 0;
};

template <size_t MaxDimensionSize = 100, template <class...> class Container, class... TemplateArguments>
auto make_multidimensional_sparse_tuning_problem(const Container<TemplateArguments...>& in,
                                                 std::vector<std::string> names) {
  return MultidimensionalSparseTuningProblem<Container, MaxDimensionSize, TemplateArguments...>(in, names);
}

// combine(): a tuner's problem with one more axis in front.
template <class Tuner>
struct ExtendableTunerMixin {
  template <class Key>
  auto combine(const std::string& axis_name, const std::vector<Key>& new_axis) const {
    return static_cast<const Tuner*>(this)->get_tuner().extend(axis_name, new_axis);
  }
  template <class... Coordinates>
  auto get_point(Coordinates... coordinates) const {
    return static_cast<const Tuner*>(this)->get_tuner().get_point(coordinates...);
  }

 private:
  ExtendableTunerMixin() = default;
  friend Tuner;
};

// Team size and vector length of a TeamPolicy whose fields are AUTO: the
// space maps each legal vector length (powers of two up to the maximum) to
// the legal team sizes (powers of two down from the maximum at that length).
class TeamSizeTuner : public ExtendableTunerMixin<TeamSizeTuner> {
  using SpaceDescription = std::map<int64_t, std::vector<int64_t>>;
  using TunerType        = decltype(make_multidimensional_sparse_tuning_problem<20>(
      std::declval<SpaceDescription>(), std::declval<std::vector<std::string>>()));
  TunerType tuner;

 public:
  TeamSizeTuner() = default;
  template <class Calculator, class Functor, class TagType, class... Properties>
  TeamSizeTuner(const std::string& name, const Kokkos::TeamPolicy<Properties...>& policy_in, const Functor& functor,
                const TagType& tag, Calculator calc) {
    Kokkos::TeamPolicy<Properties...> policy(policy_in);
    const int initial_vl = policy.impl_vector_length();
    if (initial_vl < 1) policy.impl_set_vector_length(1);
    std::vector<int64_t> lengths;
    if (policy.impl_auto_vector_length()) {
      for (int vl = Kokkos::TeamPolicy<Properties...>::vector_length_max(); vl >= 1; vl /= 2) {
        policy.impl_set_vector_length(vl);
        if (policy.impl_auto_team_size() || policy.team_size() <= calc.get_max_team_size(policy, functor, tag))
          lengths.push_back(vl);
      }
    } else {
      lengths.push_back(policy.impl_vector_length());
    }
    SpaceDescription space;
    for (const int64_t vl : lengths) {
      policy.impl_set_vector_length(int(vl));
      std::vector<int64_t> sizes;
      if (policy.impl_auto_team_size()) {
        for (int ts = calc.get_max_team_size(policy, functor, tag); ts >= 1; ts /= 2) sizes.push_back(ts);
      } else {
        sizes.push_back(policy.team_size());
      }
      space[vl] = sizes;
    }
    tuner = make_multidimensional_sparse_tuning_problem<20>(space, {name + "_vector_length", name + "_team_size"});
  }
  template <class... Properties>
  auto tune(const Kokkos::TeamPolicy<Properties...>& policy_in) {
    Kokkos::TeamPolicy<Properties...> policy(policy_in);
    if (have_tuning_tool()) {
      const auto c = tuner.begin();
      if (std::get<0>(c) > 0) {
        policy.impl_set_team_size(int(std::get<1>(c)));
        policy.impl_set_vector_length(int(std::get<0>(c)));
      }
    }
    return policy;
  }
  void end() {
    if (have_tuning_tool()) tuner.end();
  }
  TunerType get_tuner() const { return tuner; }
};

// One tuned value of an interval [lower, upper] with the given step.
template <class Bound>
class SingleDimensionalRangeTuner {
  size_t m_id      = 0;
  size_t m_context = 0;
  Bound m_default{};

  static Bound read(const VariableValue& v) {
    if constexpr (std::is_floating_point<Bound>::value) return static_cast<Bound>(v.value.double_value);
    else return static_cast<Bound>(v.value.int_value);
  }

 public:
  SingleDimensionalRangeTuner() = default;
  SingleDimensionalRangeTuner(const std::string& name, StatisticalCategory category, Bound default_val, Bound lower,
                              Bound upper, Bound step = Bound(0))
      : m_default(default_val) {
    VariableInfo info;
    info.category      = category;
    info.candidates    = make_candidate_range(lower, upper, step, false, false);
    info.valueQuantity = CandidateValueType::kokkos_value_range;
    info.type = std::is_floating_point<Bound>::value ? ValueType::kokkos_value_double : ValueType::kokkos_value_int64;
    m_id      = declare_output_type(name, info);
  }
  Bound begin() {
    m_context = get_new_context_id();
    begin_context(m_context);
    auto v = make_variable_value(m_id, m_default);
    request_output_values(m_context, 1, &v);
    return read(v);
  }
  void end() { end_context(m_context); }
  template <class Functor>
  void with_tuned_value(Functor& func) {
    func(begin());
    end();
  }
};

// Desired occupancy (percent, 5 to 100 in steps of 5) of a RangePolicy.
class RangePolicyOccupancyTuner {
  using TunerType = SingleDimensionalRangeTuner<int64_t>;
  TunerType tuner;

 public:
  RangePolicyOccupancyTuner() = default;
  template <class Calculator, class Functor, class TagType, class... Properties>
  RangePolicyOccupancyTuner(const std::string& name, const Kokkos::RangePolicy<Properties...>&, const Functor&,
                            const TagType&, Calculator)
      : tuner(name, StatisticalCategory::kokkos_value_ratio, int64_t(100), int64_t(5), int64_t(100), int64_t(5)) {}
  template <class... Properties>
  auto tune(const Kokkos::RangePolicy<Properties...>& policy_in) {
    Kokkos::RangePolicy<Properties...> policy(policy_in);
    if (have_tuning_tool())
      policy.impl_set_desired_occupancy(Kokkos::Experimental::DesiredOccupancy{static_cast<int>(tuner.begin())});
    return policy;
  }
  void end() {
    if (have_tuning_tool()) tuner.end();
  }
  TunerType get_tuner() const { return tuner; }
};

// Tile shape of an MDRangePolicy of rank R: powers of two per dimension
// whose product stays below the backend's largest tile.
template <int MDRangeRank>
struct MDRangeTuner : public ExtendableTunerMixin<MDRangeTuner<MDRangeRank>> {
 private:
  static constexpr int rank       = MDRangeRank;
  static constexpr int max_slices = 15;
  using SpaceDescription          = typename Impl::n_dimensional_sparse_structure<int, rank>::type;
  using TunerType                 = decltype(make_multidimensional_sparse_tuning_problem<max_slices>(
      std::declval<SpaceDescription>(), std::declval<std::vector<std::string>>()));
  TunerType tuner;

  template <class Policy, class Tuple, size_t... Is>
  static void set_tile(Policy& p, const Tuple& t, std::index_sequence<Is...>) {
    p.impl_change_tile_size({std::get<Is>(t)...});
  }

 public:
  MDRangeTuner() = default;
  template <class Functor, class TagType, class Calculator, class... Properties>
  MDRangeTuner(const std::string& name, const Kokkos::MDRangePolicy<Properties...>& policy, const Functor& functor,
               const TagType& tag, Calculator calc) {
    SpaceDescription desc;
    Impl::fill_tile(desc, calc.get_mdrange_max_tile_size_product(policy, functor, tag));
    std::vector<std::string> names;
    for (int x = 0; x < rank; ++x) names.push_back(name + "_tile_size_" + std::to_string(x));
    tuner = make_multidimensional_sparse_tuning_problem<max_slices>(desc, names);
  }
  template <class... Properties>
  auto tune(const Kokkos::MDRangePolicy<Properties...>& policy_in) {
    Kokkos::MDRangePolicy<Properties...> policy(policy_in);
    if (have_tuning_tool()) set_tile(policy, tuner.begin(), std::make_index_sequence<rank>{});
    return policy;
  }
  void end() {
    if (have_tuning_tool()) tuner.end();
  }
  TunerType get_tuner() const { return tuner; }
};

// A choice among a list: the tool picks an index (default 0).
template <class Choice>
struct CategoricalTuner {
  using choice_list = std::vector<Choice>;
  choice_list choices;
  size_t context            = 0;
  size_t tuning_variable_id = 0;
  CategoricalTuner(std::string name, choice_list list) : choices(std::move(list)) {
    std::vector<int64_t> indices(choices.size());
    for (size_t x = 0; x < choices.size(); ++x) indices[x] = int64_t(x);
    VariableInfo info;
    info.category      = StatisticalCategory::kokkos_value_categorical;
    info.valueQuantity = CandidateValueType::kokkos_value_set;
    info.type          = ValueType::kokkos_value_int64;
    info.candidates    = make_candidate_set(indices.size(), indices.data());
    tuning_variable_id = declare_output_type(name, info);
  }
  const Choice& begin() {
    context = get_new_context_id();
    begin_context(context);
    VariableValue v = make_variable_value(tuning_variable_id, int64_t(0));
    request_output_values(context, 1, &v);
    return choices[size_t(v.value.int_value)];
  }
  void end() { end_context(context); }
};

template <class Choice>
auto make_categorical_tuner(std::string name, std::vector<Choice> choices) -> CategoricalTuner<Choice> {
  return CategoricalTuner<Choice>(name, choices);
}

}  // namespace Experimental
}  // namespace Tools
}  // namespace Kokkos

#endif
