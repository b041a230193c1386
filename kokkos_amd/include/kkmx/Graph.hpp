// Kokkos::Experimental::Graph on MI355X: a DAG of dispatches recorded once
// and replayed as a hipGraph.
//
// API (reference core/src/Kokkos_Graph.hpp:43-185, Kokkos_GraphNode.hpp:134-531):
// create_graph(exec, closure) / Graph{exec} + root_node(), node.then_parallel_
// {for,reduce,scan}(...), node.then(label, functor), when_all(nodes...),
// graph.instantiate() (once), graph.submit([exec]).
//
// MI355X design (the reference adds one hipGraph kernel node per Kokkos node,
// core/src/HIP/Kokkos_HIP_Graph_Impl.hpp:82-221):
//  * every node owns a private execution-space instance (own control block,
//    partials, scan state), so branches that run concurrently never share
//    counters;
//  * instantiate() first runs each node in size-only mode (scratch is sized
//    and zeroed on an ordinary stream, nothing is launched), then stream-
//    captures the DAG: each node is launched on a stream of a small pool,
//    joined to its predecessors with events, and all streams are joined back
//    to the origin before the capture ends.  No allocation or memset is ever
//    recorded; the kernels' self-cleaning counters and scan epochs make every
//    replay start from a clean state;
//  * submit() is one hipGraphLaunch on the submitting instance's stream,
//    preceded once every 2^29 scan replays by a kernel that zeroes the scan
//    nodes' status words (their epochs keep 30 bits).
// Reductions inside a graph must write to device-accessible Views.
#ifndef KKMX_GRAPH_HPP
#define KKMX_GRAPH_HPP

#include <kkmx/CopyViews.hpp>

#include <functional>

namespace Kokkos {
namespace Impl {

template <class Exec>
struct GraphNodeImpl {
  std::vector<GraphNodeImpl*> preds;
  std::function<void(const Exec&)> launch;  // empty for root / aggregate nodes
  std::shared_ptr<void> instance;           // node-private instance state
  std::string label;
  int id = 0;
};

template <class Exec>
class GraphImpl;
template <class Exec>
struct GraphNodeAggregateImpl;
template <class Exec>
struct GraphNodeRootImpl;

template <>
class GraphImpl<MI355X> {
 public:
  using node_t           = GraphNodeImpl<MI355X>;
  using aggregate_impl_t = GraphNodeAggregateImpl<MI355X>;
  using root_node_impl_t = GraphNodeRootImpl<MI355X>;
  explicit GraphImpl(const MI355X& ex) : m_exec(ex) {}
  ~GraphImpl() {
    if (m_exec_graph) {
      (void)mx_stream_sync(m_exec.stream());
      (void)mx_graph_exec_destroy(m_exec_graph);
    }
    if (m_graph) (void)mx_graph_destroy(m_graph);
    for (auto e : m_events) (void)mx_event_destroy(e);
    for (auto s : m_streams) (void)mx_stream_destroy(s);
  }
  node_t* create_node() {
    m_nodes.emplace_back(new node_t());
    m_nodes.back()->id = static_cast<int>(m_nodes.size()) - 1;
    return m_nodes.back().get();
  }
  node_t* root() {
    if (m_nodes.empty()) create_node();
    return m_nodes.front().get();
  }
  const MI355X& exec() const { return m_exec; }
  bool instantiated() const { return m_exec_graph != nullptr; }

  void instantiate() {
    if (m_exec_graph)
      Kokkos::abort("Kokkos::Graph::instantiate: a graph can only be instantiated once");
    const int n = static_cast<int>(m_nodes.size());
    // Node-private instances; size-only pass on the graph's stream.
    std::vector<MI355X> spaces;
    spaces.reserve(n);
    for (auto& nd : m_nodes) {
      auto inst = std::make_shared<MI355XInstance>(m_exec.stream(), false, mi355x_globals().next_id++);
      nd->instance = inst;
      spaces.emplace_back(inst);
    }
    {
      struct SizeOnly {
        SizeOnly() { g_size_only_dispatch = true; }
        ~SizeOnly() { g_size_only_dispatch = false; }
      } guard;
      for (int i = 0; i < n; ++i)
        if (m_nodes[i]->launch) m_nodes[i]->launch(spaces[i]);
    }
    KKMX_SAFE_CALL(mx_stream_sync(m_exec.stream()));
    // Capture, nodes in creation order (a topological order: a node can only
    // be created after its predecessors).
    mx_stream_t origin = m_exec.stream();
    std::vector<mx_stream_t> node_stream(n, nullptr);
    std::vector<mx_event_t> node_done(n, nullptr);
    std::vector<int> children_started(n, 0);
    // Branches go to a small pool of capture streams (the hardware has 4
    // queues per process); reusing a stream only adds ordering the DAG
    // already allows, never removes a dependency.
    constexpr size_t kMaxBranchStreams = 4;
    size_t pool_next = 0;
    auto take_stream = [&]() -> mx_stream_t {
      if (m_streams.size() < kMaxBranchStreams) {
        mx_stream_t s;
        KKMX_SAFE_CALL(mx_stream_create(&s));
        m_streams.push_back(s);
        return s;
      }
      return m_streams[pool_next++ % m_streams.size()];
    };
    auto new_event = [&]() -> mx_event_t {
      mx_event_t e;
      KKMX_SAFE_CALL(mx_event_create(&e));
      m_events.push_back(e);
      return e;
    };
    KKMX_SAFE_CALL(mx_graph_capture_begin(origin));
    for (int i = 0; i < n; ++i) {
      node_t* nd = m_nodes[i].get();
      mx_stream_t s;
      if (nd->preds.empty()) {
        s = (i == 0) ? origin : take_stream();
        if (s != origin) {
          mx_event_t e = new_event();
          KKMX_SAFE_CALL(mx_event_record(e, origin));
          KKMX_SAFE_CALL(mx_stream_wait_event(s, e));
        }
      } else {
        // Continue on the first predecessor's stream when this is its first
        // successor, otherwise fork onto a fresh stream.
        const int p0 = nd->preds.front()->id;
        if (children_started[p0]++ == 0) s = node_stream[p0];
        else s = take_stream();
        for (auto* p : nd->preds)
          if (node_stream[p->id] != s) KKMX_SAFE_CALL(mx_stream_wait_event(s, node_done[p->id]));
      }
      node_stream[i] = s;
      if (nd->launch) {
        auto inst    = std::static_pointer_cast<MI355XInstance>(nd->instance);
        inst->stream = s;
        nd->launch(spaces[i]);
        inst->stream = m_exec.stream();
      }
      node_done[i] = new_event();
      KKMX_SAFE_CALL(mx_event_record(node_done[i], s));
    }
    // Join every stream back to the origin.
    for (int i = 0; i < n; ++i)
      if (node_stream[i] != origin) KKMX_SAFE_CALL(mx_stream_wait_event(origin, node_done[i]));
    KKMX_SAFE_CALL(mx_graph_capture_end(origin, &m_graph));
    KKMX_SAFE_CALL(mx_graph_instantiate(&m_exec_graph, m_graph));
    // Scan launches each replay makes per node instance; they are counted
    // at submit, where the status arrays are zeroed when the period is due.
    for (auto& nd : m_nodes) {
      if (!nd->instance) continue;
      auto* inst = static_cast<MI355XInstance*>(nd->instance.get());
      if (inst->scan_since_clear) m_scan_nodes.emplace_back(inst, inst->scan_since_clear);
      inst->scan_since_clear = 0;
    }
  }

  void submit(const MI355X& ex) {
    if (!m_exec_graph) instantiate();
    for (auto& sn : m_scan_nodes) note_scan_launches(sn.first, sn.second, ex.stream());
    KKMX_SAFE_CALL(mx_graph_launch(m_exec_graph, ex.stream()));
  }
  mx_graph_t native_graph() const { return m_graph; }
  mx_graph_exec_t native_graph_exec() const { return m_exec_graph; }
  size_t node_count() const { return m_nodes.size(); }

 private:
  MI355X m_exec;
  std::vector<std::unique_ptr<node_t>> m_nodes;
  std::vector<mx_stream_t> m_streams;
  std::vector<mx_event_t> m_events;
  std::vector<std::pair<MI355XInstance*, uint64_t>> m_scan_nodes;
  mx_graph_t m_graph           = nullptr;
  mx_graph_exec_t m_exec_graph = nullptr;
};

// Host execution: nodes run in creation (topological) order at submit.
template <>
class GraphImpl<Serial> {
 public:
  using node_t           = GraphNodeImpl<Serial>;
  using aggregate_impl_t = GraphNodeAggregateImpl<Serial>;
  using root_node_impl_t = GraphNodeRootImpl<Serial>;
  explicit GraphImpl(const Serial& ex) : m_exec(ex) {}
  node_t* create_node() {
    m_nodes.emplace_back(new node_t());
    m_nodes.back()->id = static_cast<int>(m_nodes.size()) - 1;
    return m_nodes.back().get();
  }
  node_t* root() {
    if (m_nodes.empty()) create_node();
    return m_nodes.front().get();
  }
  const Serial& exec() const { return m_exec; }
  bool instantiated() const { return m_instantiated; }
  void instantiate() {
    if (m_instantiated) Kokkos::abort("Kokkos::Graph::instantiate: a graph can only be instantiated once");
    m_instantiated = true;
  }
  void submit(const Serial& ex) {
    if (!m_instantiated) instantiate();
    for (auto& n : m_nodes)
      if (n->launch) n->launch(ex);
  }
  size_t node_count() const { return m_nodes.size(); }

 private:
  Serial m_exec;
  std::vector<std::unique_ptr<node_t>> m_nodes;
  bool m_instantiated = false;
};

template <class F>
struct ThenWrapper {
  F f;
  KOKKOS_INLINE_FUNCTION void operator()(const int64_t) const { f(); }
};

template <class Policy, class Exec>
Policy rebind(const Policy& p, const Exec& e) {
  Policy q = p;
  q.impl_set_space(e);
  return q;
}

// Node kinds at the type level.  A GraphNodeRef's type names the kind of node
// it refers to and the type of its predecessor's reference, as the
// reference's node references do (core/src/Kokkos_GraphNode.hpp:41-60); the
// nodes themselves are type-erased GraphNodeImpl records, so these types only
// describe: a kernel recorded from parallel_for / reduce / scan, a single
// call (`then`), a user capture, and the join of `when_all`.
template <class Exec, class Policy, class Functor, class PatternTag, class... Args>
struct GraphNodeKernelImpl {
  using execution_space = Exec;
  using policy_type     = Policy;
  using functor_type    = Functor;
  using pattern_tag     = PatternTag;
};
template <class Exec, class Functor>
struct GraphNodeThenImpl {
  using execution_space = Exec;
  using functor_type    = Functor;
};
template <class Exec, class Functor>
struct GraphNodeCaptureImpl {
  using execution_space = Exec;
  using functor_type    = Functor;
};
template <class Exec>
struct GraphNodeAggregateImpl {
  using execution_space = Exec;
};
template <class Exec>
struct GraphNodeRootImpl {
  using execution_space = Exec;
};

template <class T>
struct is_graph_kernel : std::false_type {};
template <class Exec, class Policy, class Functor, class PatternTag, class... Args>
struct is_graph_kernel<GraphNodeKernelImpl<Exec, Policy, Functor, PatternTag, Args...>> : std::true_type {};
template <class Exec, class Functor>
struct is_graph_kernel<GraphNodeThenImpl<Exec, Functor>> : std::true_type {};
template <class T>
inline constexpr bool is_graph_kernel_v = is_graph_kernel<T>::value;

template <class T>
struct is_graph_capture : std::false_type {};
template <class Exec, class Functor>
struct is_graph_capture<GraphNodeCaptureImpl<Exec, Functor>> : std::true_type {};
template <class T>
inline constexpr bool is_graph_capture_v = is_graph_capture<T>::value;

// The policy type of a kernel node: the user's policy with IsGraphKernelTag
// appended to its traits.
template <class Policy>
struct graph_kernel_policy;
template <template <class...> class P, class... Traits>
struct graph_kernel_policy<P<Traits...>> {
  using type = P<Traits..., IsGraphKernelTag>;
};
template <class Policy>
using graph_kernel_policy_t = typename graph_kernel_policy<Policy>::type;

}  // namespace Impl

namespace Experimental {

// The kernel / predecessor parameter of a reference whose node kind is not
// named in its type (core/src/Kokkos_Graph_fwd.hpp:29).
struct TypeErasedTag {};

template <class Exec = DefaultExecutionSpace>
class Graph;

template <class Exec, class Kernel = TypeErasedTag, class Predecessor = TypeErasedTag>
class GraphNodeRef {
  template <class, class, class>
  friend class GraphNodeRef;

 public:
  using execution_space   = Exec;
  using graph_kernel      = Kernel;
  using graph_predecessor = Predecessor;

  GraphNodeRef() = default;
  GraphNodeRef(std::shared_ptr<Impl::GraphImpl<Exec>> g, Impl::GraphNodeImpl<Exec>* n) : m_graph(std::move(g)), m_node(n) {}
  // A reference converts to one that names less of its type (the kernel,
  // the predecessor or both replaced by TypeErasedTag).
  template <class K2, class P2,
            class = std::enable_if_t<!std::is_same<GraphNodeRef, GraphNodeRef<Exec, K2, P2>>::value &&
                                     (std::is_same<Kernel, TypeErasedTag>::value || std::is_same<Kernel, K2>::value) &&
                                     (std::is_same<Predecessor, TypeErasedTag>::value ||
                                      std::is_same<Predecessor, P2>::value)>>
  GraphNodeRef(const GraphNodeRef<Exec, K2, P2>& other) : m_graph(other.m_graph), m_node(other.m_node) {}

  template <class Policy, class F, class = std::enable_if_t<is_execution_policy<Impl::remove_cvref_t<Policy>>::value>>
  auto then_parallel_for(std::string label, const Policy& policy_in, const F& f) const {
    using kernel_t = Impl::GraphNodeKernelImpl<Exec, Impl::graph_kernel_policy_t<Impl::remove_cvref_t<Policy>>,
                                               std::decay_t<F>, ParallelForTag>;
    const auto policy = policy_in;
    return add<kernel_t>([label, policy, f](const Exec& e) { Kokkos::parallel_for(label, Impl::rebind(policy, e), f); });
  }
  template <class Policy, class F, class = std::enable_if_t<is_execution_policy<Impl::remove_cvref_t<Policy>>::value>>
  auto then_parallel_for(const Policy& policy, const F& f) const {
    return then_parallel_for(std::string(), policy, f);
  }
  template <class F>
  auto then_parallel_for(std::string label, std::size_t n, const F& f) const {
    return then_parallel_for(std::move(label), RangePolicy<Exec>(0, n), f);
  }
  template <class F>
  auto then_parallel_for(std::size_t n, const F& f) const {
    return then_parallel_for(std::string(), RangePolicy<Exec>(0, n), f);
  }

  template <class Policy, class F, class Ret,
            class = std::enable_if_t<is_execution_policy<Impl::remove_cvref_t<Policy>>::value>>
  auto then_parallel_reduce(std::string label, const Policy& policy_in, const F& f, const Ret& ret) const {
    check_result<Ret>();
    using kernel_t = Impl::GraphNodeKernelImpl<Exec, Impl::graph_kernel_policy_t<Impl::remove_cvref_t<Policy>>,
                                               std::decay_t<F>, ParallelReduceTag, Ret>;
    const auto policy = policy_in;
    return add<kernel_t>([label, policy, f, ret](const Exec& e) {
      Ret r = ret;
      Kokkos::parallel_reduce(label, Impl::rebind(policy, e), f, r);
    });
  }
  template <class Policy, class F, class Ret,
            class = std::enable_if_t<is_execution_policy<Impl::remove_cvref_t<Policy>>::value>>
  auto then_parallel_reduce(const Policy& policy, const F& f, const Ret& ret) const {
    return then_parallel_reduce(std::string(), policy, f, ret);
  }
  template <class F, class Ret>
  auto then_parallel_reduce(std::string label, typename Exec::size_type n, const F& f, const Ret& ret) const {
    return then_parallel_reduce(std::move(label), RangePolicy<Exec>(0, n), f, ret);
  }
  template <class F, class Ret>
  auto then_parallel_reduce(typename Exec::size_type n, const F& f, const Ret& ret) const {
    return then_parallel_reduce(std::string(), RangePolicy<Exec>(0, n), f, ret);
  }

  template <class Policy, class F, class = std::enable_if_t<is_execution_policy<Impl::remove_cvref_t<Policy>>::value>>
  auto then_parallel_scan(std::string label, const Policy& policy_in, const F& f) const {
    using kernel_t = Impl::GraphNodeKernelImpl<Exec, Impl::graph_kernel_policy_t<Impl::remove_cvref_t<Policy>>,
                                               std::decay_t<F>, ParallelScanTag>;
    const auto policy = policy_in;
    return add<kernel_t>([label, policy, f](const Exec& e) { Kokkos::parallel_scan(label, Impl::rebind(policy, e), f); });
  }
  template <class Policy, class F, class = std::enable_if_t<is_execution_policy<Impl::remove_cvref_t<Policy>>::value>>
  auto then_parallel_scan(const Policy& policy, const F& f) const {
    return then_parallel_scan(std::string(), policy, f);
  }
  template <class F>
  auto then_parallel_scan(std::string label, typename Exec::size_type n, const F& f) const {
    return then_parallel_scan(std::move(label), RangePolicy<Exec>(0, n), f);
  }
  template <class F>
  auto then_parallel_scan(typename Exec::size_type n, const F& f) const {
    return then_parallel_scan(std::string(), RangePolicy<Exec>(0, n), f);
  }

  // Run functor() once.  The instance argument names the space the call is
  // meant for; inside the graph it runs on the graph's node instance.
  template <class Label, class F,
            class = std::enable_if_t<std::is_invocable_r<void, const Impl::remove_cvref_t<F>>::value>>
  auto then(Label&& label, const Exec&, const F& f) const {
    using kernel_t = Impl::GraphNodeThenImpl<Exec, Impl::remove_cvref_t<F>>;
    std::string name(std::forward<Label>(label));
    return add<kernel_t>([name, f](const Exec& e) {
      Kokkos::parallel_for(name, RangePolicy<Exec, IndexType<int64_t>>(e, 0, 1), Impl::ThenWrapper<F>{f});
    });
  }
  template <class Label, class F,
            class = std::enable_if_t<std::is_invocable_r<void, const Impl::remove_cvref_t<F>>::value>>
  auto then(Label&& label, const F& f) const {
    return then(std::forward<Label>(label), Exec{}, f);
  }

  // Work the caller launches itself on the given instance's stream (a
  // library call, a hand-written kernel), recorded into the graph by stream
  // capture like every other node.
  template <class F, class = std::enable_if_t<std::is_invocable_r<void, const Impl::remove_cvref_t<F>, const Exec&>::value>>
  auto mi355x_capture(const Exec&, const F& f) const {
    using kernel_t = Impl::GraphNodeCaptureImpl<Exec, Impl::remove_cvref_t<F>>;
    return add<kernel_t>([f](const Exec& e) {
      if (!g_size_only_dispatch_for(e)) f(e);
    });
  }

  std::shared_ptr<Impl::GraphImpl<Exec>> impl_graph() const { return m_graph; }
  Impl::GraphNodeImpl<Exec>* impl_node() const { return m_node; }

 private:
  std::shared_ptr<Impl::GraphImpl<Exec>> m_graph;
  Impl::GraphNodeImpl<Exec>* m_node = nullptr;

  static bool g_size_only_dispatch_for(const Exec&) {
    if constexpr (std::is_same<Exec, MI355X>::value) return Impl::g_size_only_dispatch;
    else return false;
  }

  template <class Ret>
  static void check_result() {
    if constexpr (is_view<Ret>::value) {
      static_assert(Impl::exec_can_access<Exec, typename Ret::memory_space>::value,
                    "Kokkos::Graph: reduction results must be Views the execution space can write");
    } else if constexpr (is_reducer<Ret>::value) {
      static_assert(Impl::exec_can_access<Exec, typename Ret::result_view_type::memory_space>::value,
                    "Kokkos::Graph: reducer results must live in memory the execution space can write");
    } else {
      static_assert(is_view<Ret>::value || is_reducer<Ret>::value,
                    "Kokkos::Graph: reductions need a View or a reducer on a View as result");
    }
  }
  template <class KernelT>
  GraphNodeRef<Exec, KernelT, GraphNodeRef> add(std::function<void(const Exec&)> launch) const {
    if (m_graph->instantiated())
      Kokkos::Impl::throw_runtime_exception("Kokkos::Graph: nodes cannot be added after instantiate()");
    auto* n   = m_graph->create_node();
    n->launch = std::move(launch);
    n->preds.push_back(m_node);
    return GraphNodeRef<Exec, KernelT, GraphNodeRef>(m_graph, n);
  }
};

template <class Exec>
class Graph {
 public:
  using execution_space = Exec;
  using graph           = Graph;
  using root_t          = GraphNodeRef<Exec>;
  explicit Graph(Exec ex = Exec{}) : m_impl(std::make_shared<Impl::GraphImpl<Exec>>(ex)) { m_impl->root(); }
  const Exec& get_execution_space() const { return m_impl->exec(); }
  root_t root_node() const { return root_t(m_impl, m_impl->root()); }
  void instantiate() { m_impl->instantiate(); }
  void submit(const Exec& ex) const { m_impl->submit(ex); }
  void submit() const { m_impl->submit(m_impl->exec()); }
  auto native_graph() const { return m_impl->native_graph(); }
  auto native_graph_exec() const { return m_impl->native_graph_exec(); }
  size_t impl_node_count() const { return m_impl->node_count(); }

 private:
  std::shared_ptr<Impl::GraphImpl<Exec>> m_impl;
};

template <class Exec, class Closure>
std::enable_if_t<is_execution_space<Exec>::value, Graph<Exec>> create_graph(Exec ex, Closure&& closure) {
  Graph<Exec> g{ex};
  closure(g.root_node());
  return g;
}
template <class Exec = DefaultExecutionSpace, class Closure>
std::enable_if_t<!is_execution_space<Impl::remove_cvref_t<Closure>>::value, Graph<Exec>> create_graph(
    Closure&& closure) {
  return create_graph(Exec{}, static_cast<Closure&&>(closure));
}

template <class Exec, class K, class P, class... More>
GraphNodeRef<Exec, Impl::GraphNodeAggregateImpl<Exec>, TypeErasedTag> when_all(const GraphNodeRef<Exec, K, P>& first,
                                                                                const More&... more) {
  auto g = first.impl_graph();
  if (g->instantiated())
    Kokkos::Impl::throw_runtime_exception("Kokkos::Graph: nodes cannot be added after instantiate()");
  auto* n = g->create_node();
  n->preds.push_back(first.impl_node());
  (n->preds.push_back(more.impl_node()), ...);
  return GraphNodeRef<Exec, Impl::GraphNodeAggregateImpl<Exec>, TypeErasedTag>(g, n);
}

}  // namespace Experimental
}  // namespace Kokkos

#endif
