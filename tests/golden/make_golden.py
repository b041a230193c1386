"""Writes tests/golden/known_answers.json.

Every value here comes from a published source or a closed form, not from the
code under test:
  * SplitMix64 outputs for seed 1234567 (the test vector published with
    Vigna's splitmix64.c reference implementation);
  * the closed forms of the reference's own unit tests:
      core/unit_test/TestParallelScanRangePolicy.hpp:42-88 (prefix i(i-1)/2,
      postfix i(i+1)/2, total N(N-1)/2 for the scan over 0..N-1),
      core/unit_test/TestReduce.hpp:27-140 (count N, sum(i+1) = N(N+1)/2,
      sum(N-i) = N(N+1)/2; the final() variants negate),
      core/unit_test/hip/TestHIP_ScanUnit.cpp:26-70 (intra-block inclusive
      scan of 1..n = (i+1)(i+2)/2);
  * the STREAM validation recurrence of benchmarks/stream/stream-kokkos.cpp:79-100.
"""
import json
from pathlib import Path

out = {
    "splitmix64_seed_1234567": [
        6457827717110365317,
        3203168211198807973,
        9817491932198370423,
        4593380528125082431,
        16408922859458223821,
    ],
    "scan_sizes": [0, 1, 2, 63, 64, 65, 1000, 4095, 4096, 4097, 16384, 100003],
}
# Reference unit-test closed forms, materialised for a few sizes.
scan = {}
for n in [1, 2, 7, 64, 1000, 4097]:
    scan[str(n)] = {
        "prefix_last": (n - 1) * (n - 2) // 2,
        "postfix_last": (n - 1) * n // 2,
        "total": n * (n - 1) // 2,
    }
out["scan_closed_forms"] = scan
red = {}
for n in [1, 10, 1000, 1 << 20]:
    red[str(n)] = {"count": n, "sum_i_plus_1": n * (n + 1) // 2, "sum_n_minus_i": n * (n + 1) // 2}
out["reduce_closed_forms"] = red
# STREAM validation values after STREAM_NTIMES iterations (ntimes=20, scalar 3).
ai, bi, ci = 1.0, 2.0, 0.0
for _ in range(20):
    ci = ai
    bi = 3.0 * ci
    ci = ai + bi
    ai = bi + 3.0 * ci
out["stream_recurrence_20"] = {"a": ai, "b": bi, "c": ci}
Path(__file__).with_name("known_answers.json").write_text(json.dumps(out, indent=1) + "\n")
