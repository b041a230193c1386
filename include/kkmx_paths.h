/*
 * kkmx_paths.h -- C-ABI entry points of the hot path, for callers that cannot
 * instantiate C++ templates (ctypes, cgo, JNI, other languages) and for the
 * parity tests and the benchmark.
 *
 * Each entry point is ordinary Kokkos user code compiled against the MI355X
 * execution space (kokkos_amd/csrc/paths.hip): it wraps the caller's device
 * pointers in unmanaged Views and calls parallel_for / parallel_reduce /
 * parallel_scan.  Pointers named *_dev are device memory (mx_malloc); results
 * land in device memory asynchronously on the default instance stream
 * (mx_default_stream), the behaviour of a View result in the reference
 * (core/src/Kokkos_Parallel_Reduce.hpp:1592-1635: no fence for View results).
 * Every function returns 0 on success, a negative value for a rejected
 * argument and a positive HIP status for a runtime failure; kkmx_last_error()
 * gives the message.
 *
 * Reference interfaces:
 *   kkmx_stream_*           benchmarks/stream/stream-kokkos.cpp:36-77 (set/copy/scale/add/triad functors)
 *   kkmx_reduce_*           core/src/Kokkos_Parallel_Reduce.hpp:1679-1833 (parallel_reduce with
 *                           Sum / Min / Max / MinMax reducers, :33-212,536-587)
 *   kkmx_scan_*             core/src/Kokkos_Parallel.hpp:348-454 (parallel_scan with and without total)
 *   kkmx_fill_*             algorithms/src/Kokkos_Random.hpp:943-1010 (fill_random) -- seeded inputs,
 *                           counter-based so host and device produce identical values
 */
#ifndef KKMX_PATHS_H
#define KKMX_PATHS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Kokkos::initialize / finalize for C callers (idempotent). */
int kkmx_init(int device);
int kkmx_finalize(void);
const char* kkmx_last_error(void);
/* The stream every entry point below is ordered on. */
void* kkmx_stream(void);
int kkmx_fence(void);

/* ---- seeded inputs (splitmix64 counter hash of (seed, i)) ---- */
int kkmx_fill_uniform_f64(double* x_dev, int64_t n, uint64_t seed, double lo, double hi);
int kkmx_fill_uniform_i64(int64_t* x_dev, int64_t n, uint64_t seed, int64_t lo, int64_t hi);
int kkmx_fill_uniform_i32(int32_t* x_dev, int64_t n, uint64_t seed, int32_t lo, int32_t hi);

/* ---- STREAM: RangePolicy<IndexType<int>> parallel_for ---- */
int kkmx_stream_set(double* a_dev, int64_t n, double scalar);
int kkmx_stream_copy(const double* a_dev, double* b_dev, int64_t n);
int kkmx_stream_scale(double* b_dev, const double* c_dev, int64_t n, double scalar);
int kkmx_stream_add(const double* a_dev, const double* b_dev, double* c_dev, int64_t n);
int kkmx_stream_triad(double* a_dev, const double* b_dev, const double* c_dev, int64_t n, double scalar);

/* ---- deep_copy between device Views (ViewFill / ViewCopy,
 * core/src/Kokkos_CopyViews.hpp:58-420; HIP core/src/HIP/Kokkos_HIP_DeepCopy.cpp:23-51).
 * kkmx_deep_copy_f64: dst(i,j,k) = src(i,j,k) for rank-3 Views of
 * n0 x n1 x n2 doubles; layout 0 = LayoutLeft, 1 = LayoutRight (same
 * layouts: one contiguous copy; different: an index-space copy).
 * kkmx_deep_fill_f64: deep_copy(View<double*>(dst, n), value). ---- */
int kkmx_deep_copy_f64(double* dst_dev, const double* src_dev, int64_t n0, int64_t n1, int64_t n2, int dst_layout,
                       int src_layout);
int kkmx_deep_fill_f64(double* dst_dev, int64_t n, double value);

/* ---- Kokkos::Experimental std-algorithm scans and reductions over device
 * Views (algorithms/src/std_algorithms/Kokkos_InclusiveScan.hpp,
 * Kokkos_ExclusiveScan.hpp, Kokkos_TransformExclusiveScan.hpp,
 * Kokkos_Reduce.hpp, Kokkos_TransformReduce.hpp).  in and out may be the
 * same array.  Reductions return their value through a host pointer.
 * kkmx_std_inclusive_scan_max_i32: running maximum (a user operator with no
 *   neutral element).
 * kkmx_std_transform_exclusive_scan_sq_f64: out[i] = init + sum_{j<i} in[j]^2. */
int kkmx_std_inclusive_scan_i32(const int32_t* in_dev, int32_t* out_dev, int64_t n);
int kkmx_std_exclusive_scan_i64(const int64_t* in_dev, int64_t* out_dev, int64_t n, int64_t init);
int kkmx_std_inclusive_scan_max_i32(const int32_t* in_dev, int32_t* out_dev, int64_t n);
int kkmx_std_transform_exclusive_scan_sq_f64(const double* in_dev, double* out_dev, int64_t n, double init);
int kkmx_std_reduce_i64(const int64_t* in_dev, int64_t n, int64_t init, int64_t* out_host);
int kkmx_std_reduce_max_i64(const int64_t* in_dev, int64_t n, int64_t init, int64_t* out_host);
int kkmx_std_transform_reduce_dot_f64(const double* a_dev, const double* b_dev, int64_t n, double init,
                                      double* out_host);

/* ---- sorting (algorithms/src/sorting/Kokkos_SortPublicAPI.hpp,
 * Kokkos_BinSortPublicAPI.hpp, Kokkos_BinOpsPublicAPI.hpp).
 * kkmx_sort: Kokkos::sort of n keys in place, ascending; type 0 = int32,
 *   1 = int64, 2 = float, 3 = double, 4 = uint32, 5 = uint64.
 * kkmx_binsort_f64: BinSort(keys, BinOp1D(nbins, kmin, kmax),
 *   sort_within_bins).create_permute_vector(), then sort(values); perm_out
 *   (n ints, may be null) receives the permutation vector. ---- */
int kkmx_sort(void* keys_dev, int64_t n, int type);
/* kkmx_sort_by_key: Kokkos::Experimental::sort_by_key(exec, keys, values)
 *   (Kokkos_SortByKeyPublicAPI.hpp:27-64); keys as kkmx_sort's type codes,
 *   values are n words of value_bytes (4 or 8) moved with their keys.
 *   descending = 1 passes a greater-than comparator (the comparator
 *   overload, :67-108). Equal keys keep their input order. */
int kkmx_sort_by_key(void* keys_dev, void* values_dev, int64_t n, int key_type, int value_bytes, int descending);
int kkmx_binsort_f64(const double* keys_dev, int64_t* values_dev, int64_t n, int nbins, double kmin, double kmax,
                     int sort_within_bins, int32_t* perm_out_dev);

/* ---- parallel_reduce into a device result ---- */
int kkmx_reduce_sum_f64(const double* x_dev, int64_t n, double* out_dev);
/* op: 0 = Sum, 1 = Min, 2 = Max */
int kkmx_reduce_i64(const int64_t* x_dev, int64_t n, int op, int64_t* out_dev);
int kkmx_reduce_i32(const int32_t* x_dev, int64_t n, int op, int64_t* out_dev);
/* One pass: out_dev[0] = min, [1] = max, [2] = sum (MinMax + Sum in one struct value). */
int kkmx_reduce_minmaxsum_i64(const int64_t* x_dev, int64_t n, int64_t* out_dev);

/* ---- parallel_scan ---- */
/* Exclusive prefix sum of int32 input accumulated in int64; total_dev may be NULL. */
int kkmx_scan_excl_i32_i64(const int32_t* in_dev, int64_t* out_dev, int64_t n, int64_t* total_dev);
/* Inclusive prefix sum of int32 input in int32.  The input is read through a
 * const RandomAccess View when it shares no memory with the output, else
 * through a plain const View (in-place calls are allowed). */
int kkmx_scan_incl_i32(const int32_t* in_dev, int32_t* out_dev, int64_t n);
/* The same scan with the input View's memory traits chosen by the caller
 * (core/src/Kokkos_MemoryTraits.hpp): 0 = Unmanaged, 1 = Unmanaged |
 * RandomAccess, 2 = Unmanaged | Restrict.  1 and 2 promise the kernel never
 * writes the input and are rejected (status -1) when the buffers overlap. */
int kkmx_scan_incl_i32_traits(const int32_t* in_dev, int32_t* out_dev, int64_t n, int in_traits);
/* Which register-resident tile body every Range scan runs: -1 = chosen per
 * functor type from the runtime's own timings (the default), 0 = keep the
 * per-round wave scans, 1 = keep the functor's contributions.  Both give
 * bitwise-identical results; no counterpart in the reference (its scan has
 * one kernel, core/src/HIP/Kokkos_HIP_ParallelScan_Range.hpp:29-388), used by
 * tests and the benchmark to pin the body. */
int kkmx_set_scan_body(int body);
/* The tile body the automatic choice kept for kkmx_scan_incl_i32_traits'
 * functor with these input traits: 0 or 1, or -2 while its trial launches
 * are still running (the choice is made at the first large launch after the
 * trials have completed). */
int kkmx_scan_incl_i32_body(int in_traits);
/* kkmx_scan_incl_i32 without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Its scratch (a 256-byte control block and
 * one 8-byte status word per tile of 32768 elements) is a fixed device array
 * of 512 KiB + 256 B per GPU, held until the process ends; n above 2^31 is
 * rejected (status -1).  Call kkmx_init first.  Each call clears the part of
 * the array it uses, so the previous call's contents do not matter; calls
 * that overlap on one GPU share the array and race. */
int kkmx_scan_incl_i32_noalloc(const int32_t* in_dev, int32_t* out_dev, int64_t n);
int kkmx_scan_excl_i64(const int64_t* in_dev, int64_t* out_dev, int64_t n, int64_t* total_dev);
int kkmx_scan_excl_f64(const double* in_dev, double* out_dev, int64_t n, double* total_dev);

/* ---- MDRangePolicy<Rank<3>> 7-point stencil (config 3) ----
 * u_dev, v_dev: (n+2)^3 doubles, LayoutLeft (first index fastest), 1-cell
 * halo; v(i,j,k) = c0*u(i,j,k) + c1*(sum of the 6 neighbours) for the
 * interior.  t0..t2 = tile shape (0 = the space's default).  The sweeps
 * variant ping-pongs u -> v -> u ...  (reference: core/src/KokkosExp_MDRangePolicy.hpp:158-426) */
int kkmx_stencil7(const double* u_dev, double* v_dev, int64_t n, double c0, double c1, int t0, int t1, int t2);
int kkmx_stencil7_sweeps(double* u_dev, double* v_dev, int64_t n, int sweeps, double c0, double c1, int t0, int t1,
                         int t2);

/* ---- TeamPolicy CSR SpMV (config 3) ----
 * Seeded CSR: nnz of row r uniform in [nnz_lo, nnz_hi], columns uniform in
 * [0, ncols), values uniform in [-1, 1).  kkmx_csr_generate fills rowptr
 * (nrows+1 entries) and, when col/val are given, the nonzeros; *nnz_out gets
 * the count.  kkmx_spmv_csr: variant 0 = one row per index (RangePolicy),
 * 1 = TeamPolicy with the x-gather staged in team scratch (LDS).
 * (reference: core/src/Kokkos_ExecPolicy.hpp:365-980, Kokkos_ScratchSpace.hpp:38-170) */
int kkmx_csr_generate(int64_t* rowptr_dev, int32_t* col_dev, double* val_dev, int64_t nrows, int64_t ncols,
                      int64_t nnz_lo, int64_t nnz_hi, uint64_t seed, int64_t* nnz_out);
int kkmx_csr_fill(int32_t* col_dev, double* val_dev, int64_t nnz, int64_t ncols, uint64_t seed);
int kkmx_spmv_csr(const int64_t* rowptr_dev, const int32_t* col_dev, const double* val_dev, const double* x_dev,
                  double* y_dev, int64_t nrows, int64_t nnz, int variant, int team_size, int vector_length);

/* ---- atomics (config 4) ----
 * kkmx_xorshift64_stream: n outputs of Random_XorShift64::urand64 from seed
 *   (algorithms/src/Kokkos_Random.hpp:788-808), generated in parallel by
 *   GF(2) jump-ahead.
 * kkmx_gups: table[stream[2u] & (table_size-1)] += stream[2u+1] with
 *   atomic_add (core/src/Kokkos_Atomics_Desul_Wrapper.hpp:120).
 * kkmx_zipf_keys: Zipf(s) keys over [0, nbins) by inverse CDF.
 * kkmx_histogram: mode 0 = global atomics, 1 = bins [0, lds_bins) privatised
 *   in team scratch (LDS atomics), the rest global; 2 = as 1 with 16-byte key
 *   loads and the lowest 256 bins replicated 16 times in LDS (keys 16-byte
 *   aligned; otherwise mode 1 runs); 3 = every bin counted in LDS when they
 *   fit (up to about 70 K bins; lds_bins is then not used): the lowest 256
 *   bins replicated 16 times as 32-bit counters, the others as 16-bit halves
 *   of LDS words with exact carry-out, each team's counts stored into its row
 *   of a workspace and the rows summed into bins by a second kernel; with
 *   more bins (or more than 2^24 keys per CU) as 2 for the LDS part, with
 *   bins at or above lds_bins counted in the team's row.  No device-scope
 *   atomics in mode 3; the workspace, teams x nbins int32, is kept between
 *   calls and released by kkmx_finalize.  Modes 2 and 3 lower lds_bins to
 *   what fits in LDS beside the replicas.  Every mode adds into bins.
 * kkmx_scatter_add_f64: bins[idx[i]] += vals[i] with FP64 atomic_add. */
int kkmx_xorshift64_stream(uint64_t* out_dev, int64_t n, uint64_t seed);
int kkmx_gups(int64_t* table_dev, int64_t table_size, const uint64_t* stream_dev, int64_t nupdates);
int kkmx_zipf_keys(int32_t* keys_dev, int64_t n, double s, int32_t nbins, uint64_t seed);
int kkmx_histogram(const int32_t* keys_dev, int64_t n, int32_t* bins_dev, int32_t nbins, int mode, int lds_bins);
/* The same scatter and histogram through Kokkos::Experimental::ScatterView
 * (containers/src/Kokkos_ScatterView.hpp): create_scatter_view(bins),
 * access()(idx) += value in the loop, contribute(bins, sv).  op for the
 * float64 form: 0 = ScatterSum, 1 = ScatterMin, 2 = ScatterMax. */
int kkmx_scatter_view_f64(const double* vals_dev, const int32_t* idx_dev, int64_t n, double* bins_dev,
                          int32_t nbins, int op);
/* duplicated = 0: ScatterNonDuplicated (one atomic copy); 1: ScatterDuplicated
 * (replicas folded by contribute). */
int kkmx_scatter_view_histogram(const int32_t* keys_dev, int64_t n, int32_t* bins_dev, int32_t nbins,
                                int duplicated);
int kkmx_scatter_add_f64(const double* vals_dev, const int32_t* idx_dev, int64_t n, double* bins_dev, int32_t nbins);

/* ---- batched uneven dispatch (config 5) ----
 * A fixed set of independent ops over device arrays, planned once and run
 * many times.  op: 0 = Sum, 1 = Min, 2 = Max (into *result_dev),
 * 3 = exclusive scan, 4 = inclusive scan (into out_dev, total into
 * result_dev when not NULL), 5 = out = alpha*in + beta.  type of the input:
 * 0 = int32, 1 = int64, 2 = float64; the accumulator and outputs are int64
 * for both integer types and double for float64.
 * kkmx_batch_run modes: 0 = one parallel_reduce / parallel_scan /
 * parallel_for per op (plain per-call launches), 1 = one persistent fused
 * kernel per step, 2 = that kernel captured once into a hipGraph and
 * replayed, 3 = the per-op dispatches as an Experimental::Graph, replayed.
 * Runs are asynchronous on kkmx_stream(); kkmx_fence() waits.
 * (reference: the same ops through core/src/Kokkos_Parallel_Reduce.hpp:1679-1833
 *  and core/src/Kokkos_Parallel.hpp:130-173,348-454; graph replay through
 *  core/src/Kokkos_Graph.hpp:43-185 and HIP/Kokkos_HIP_Graph_Impl.hpp:82-221) */
typedef struct kkmx_batch kkmx_batch;
int kkmx_batch_create(kkmx_batch** out);
int kkmx_batch_destroy(kkmx_batch* b);
int kkmx_batch_add(kkmx_batch* b, int op, int type, const void* in_dev, int64_t n, void* out_dev, void* result_dev,
                   double alpha, double beta);
/* Tile shape of the fused kernel, before plan (lanes x consecutive elements
 * per lane and round x rounds): 0 = 256x1x16, 1 = 256x4x4, 2 = 256x4x8,
 * 3 = 512x4x4, 4 = 1024x1x16, 5 = 512x4x8 (the default), 6 = 1024x4x4,
 * 7 = 512x1x32.  Without it, the plan asks a loaded tuning tool for the
 * shape (output variable kokkos.mi355x.batch_shape). */
int kkmx_batch_set_shape(kkmx_batch* b, int shape);
/* The tile shape in use: the one set, else after the plan the learned
 * tuner's choice when a tuning tool is loaded, else the default. */
int kkmx_batch_get_shape(kkmx_batch* b, int* shape);
int kkmx_batch_plan(kkmx_batch* b, int64_t* tiles_out);
int kkmx_batch_run(kkmx_batch* b, int mode, int steps);

/* ---- Kokkos Tools ----
 * Load tool libraries (';'-separated) into the running process, as
 * KOKKOS_TOOLS_LIBS does at initialisation (core/src/impl/Kokkos_Profiling.cpp:
 * 520-700); a tuning tool then answers the dispatches' tuning requests. */
int kkmx_tools_load(const char* libs);
int kkmx_tools_unload(void);

/* ---- launch-shape overrides (what the tuning hooks set), -1 = heuristic ---- */
int kkmx_set_range_config(int pattern, int block, int items, int max_blocks);
/* The launch shape of tuning candidate `candidate` (0-15) of the Range
 * launch-shape variable for a dispatch of n indices (pattern 0 for, 1
 * reduce, 2 scan): what a tuning tool's answer selects, and what
 * kkmx_set_range_config then fixes for an exhaustive timing. */
int kkmx_range_candidate(int pattern, int candidate, int64_t n, int* block, int* items, int* max_blocks);
/* MDRange parallel_for coarsening: tiles per workgroup along the slowest tile
 * dimension (0 = heuristic). */
int kkmx_set_mdrange_depth(int depth);

#ifdef __cplusplus
}
#endif

#endif /* KKMX_PATHS_H */
