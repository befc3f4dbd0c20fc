"""Writes tests/golden/known_answers.json: the known-answer vectors of the
reference's own tests and example programs for the hot path, transcribed as
data (inputs and expected outputs), each with the reference file:line it
comes from.  Entries marked "derived" have no literal expected value in the
reference (its end-to-end tests compare interpreter and compiled C, or the
example prints its result); their expected values are worked out by hand from
the program at the cited lines and noted as such.

    python tests/golden/make_golden.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def cases():
    c = []

    def add(name, kind, source, inputs, expected, derived=False):
        c.append(dict(name=name, kind=kind, source=source, inputs=inputs, expected=expected,
                      derived=derived))

    # ---- scan family, segmented_reduce, sort, COO/CSR (literal answers) ----
    add("scan_exclusive", "scan", "test/Language/Hydrangea/InterpreterSpec.hs:250-255",
        dict(a=[1, 2, 3, 4], init=0, inclusive=False, reverse=False), [0, 1, 3, 6])
    add("scan_inclusive", "scan", "test/Language/Hydrangea/InterpreterSpec.hs:257-262",
        dict(a=[1, 2, 3, 4], init=0, inclusive=True, reverse=False), [1, 3, 6, 10])
    add("scanr_exclusive", "scan", "test/Language/Hydrangea/InterpreterSpec.hs:264-269",
        dict(a=[1, 2, 3, 4], init=0, inclusive=False, reverse=True), [9, 7, 4, 0])
    add("scanr_inclusive", "scan", "test/Language/Hydrangea/InterpreterSpec.hs:271-276",
        dict(a=[1, 2, 3, 4], init=0, inclusive=True, reverse=True), [10, 9, 7, 4])
    add("scan_prefix_sum_example", "scan", "examples/scan_prefix_sum.hyd:1-3",
        dict(a=[1, 2, 3, 4, 5, 6], init=0, inclusive=False, reverse=False), [0, 1, 3, 6, 10, 15],
        derived=True)
    add("segmented_reduce", "segmented_reduce", "test/Language/Hydrangea/InterpreterSpec.hs:278-287",
        dict(offsets=[0, 2, 5, 5], vals=[1, 2, 10, 20, 30], init=0), [3, 60, 0])
    add("sort_indices_stable", "sort_indices", "test/Language/Hydrangea/InterpreterSpec.hs:289-301",
        dict(keys=[3, 1, 3, 2, 1]), [1, 4, 3, 0, 2])
    add("sort_indices_gather_example", "sort_gather", "examples/sort_indices_gather.hyd:1-13",
        dict(keys=[3, 1, 3, 2, 1], vals=[10, 11, 12, 13, 14]), [11, 14, 13, 10, 12], derived=True)
    add("coo_sum_duplicates", "coo_sum_duplicates", "test/Language/Hydrangea/InterpreterSpec.hs:303-314",
        dict(rows=[0, 0, 1, 1, 2, 2], cols=[1, 1, 0, 2, 0, 0], vals=[1, 2, 10, 12, 20, 21], nnz=6),
        dict(vals=[3, 10, 12, 41, 0, 0], nnz=4, rows=[0, 1, 1, 2, 0, 0], cols=[1, 0, 2, 0, 0, 0]))
    add("csr_from_sorted_coo", "csr_from_sorted_coo", "test/Language/Hydrangea/InterpreterSpec.hs:316-327",
        dict(nrows=3, nnz=4, rows=[0, 1, 1, 2], cols=[1, 0, 2, 0], vals=[3, 10, 12, 41]),
        dict(row_ptr=[0, 1, 3, 4], col_idx=[1, 0, 2, 0], vals=[3, 10, 12, 41]))
    add("coo_to_csr_pipeline_example", "coo_to_csr", "examples/coo_to_csr_pipeline.hyd:1-78",
        dict(nrows=3, ncols=4, rows=[2, 0, 1, 0, 2, 1, 1, 0], cols=[1, 3, 0, 1, 0, 2, 2, 1],
             vals=[21, 3, 10, 1, 20, 12, 5, 4]),
        dict(row_ptr=[0, 2, 4, 6], col_idx=[1, 3, 0, 2, 0, 1, 0, 0], vals=[5, 3, 10, 17, 20, 21, 0, 0],
             nnz=6), derived=True)

    # ---- scatter (+) ----
    add("scatter_non_overlapping", "scatter", "test/Language/Hydrangea/EndToEndSpec.hs:581-586",
        dict(idx=[0, 1, 2], vals=[10, 11, 12], ndst=3), [10, 11, 12], derived=True)
    add("scatter_overlapping", "scatter", "test/Language/Hydrangea/EndToEndSpec.hs:588-593",
        dict(idx=[0, 0, 1, 1], vals=[1, 1, 1, 1], ndst=2), [2, 2], derived=True)
    add("scatter_all_to_one_bin", "scatter", "test/Language/Hydrangea/EndToEndSpec.hs:595-600",
        dict(idx=[0, 0, 0, 0, 0], vals=[1, 1, 1, 1, 1], ndst=1), [5], derived=True)
    add("scatter_fused_values_map", "scatter", "test/Language/Hydrangea/EndToEndSpec.hs:602-608",
        dict(idx=[0, 0, 1, 1], vals=[0, 3, 6, 9], ndst=2), [3, 15], derived=True)
    add("weighted_histogram_scalar_routes", "scatter", "test/Language/Hydrangea/EndToEndSpec.hs:610-616",
        dict(idx=[0, 0, 1, 1, 2, 2, 3, 3], vals=[1, 4, 7, 10, 13, 16, 19, 22], ndst=4), [5, 17, 29, 41],
        derived=True)
    add("scatter_zipwith_routes", "scatter", "test/Language/Hydrangea/EndToEndSpec.hs:618-625",
        dict(idx=[0, 0, 1, 1, 2, 2, 3, 3], vals=[100, 102, 104, 106, 108, 110, 112, 114], ndst=4),
        [202, 210, 218, 226], derived=True)
    add("scatter_guarded_even_only", "scatter", "test/Language/Hydrangea/EndToEndSpec.hs:636-642",
        dict(idx=[0, 1, 2, 3], vals=[1, 7, 13, 19], ndst=4), [1, 7, 13, 19], derived=True)
    # the guarded form itself: src = iota 8, route x/2, weight 3x+1, keep x/2 == (x+1)/2
    # (true for even x), defaults fill [4] 0
    add("scatter_guarded_fused_example", "scatter_guarded",
        "examples/scatter_guarded_fused.hyd:1-5; test/Language/Hydrangea/EndToEndSpec.hs:636-642",
        dict(idx=[x // 2 for x in range(8)], vals=[3 * x + 1 for x in range(8)],
             guard=[1 if x // 2 == (x + 1) // 2 else 0 for x in range(8)], defaults=[0, 0, 0, 0]),
        [1, 7, 13, 19], derived=True)
    add("scatter_parallel_privatized_example", "scatter", "examples/scatter_parallel_privatized.hyd:1-9",
        dict(idx=[i // 8 for i in range(64)], vals=[1] * 64, ndst=8), [8] * 8, derived=True)
    add("scatter_parallel_direct_example", "scatter", "examples/scatter_parallel_direct.hyd:1-8",
        dict(idx=list(range(8)), vals=[(i + 1) * 10 for i in range(8)], ndst=8),
        [(i + 1) * 10 for i in range(8)], derived=True)
    # weighted histogram example: bin k = sum_{j<16} (k + 64 j)^2 (closed form in the comment)
    add("weighted_histogram_example", "weighted_histogram", "examples/weighted_histogram.hyd:1-24",
        dict(data=list(range(1024)), bins=64), [sum((k + 64 * j) ** 2 for j in range(16)) for k in range(64)])

    # ---- reductions ----
    add("dot_int_example", "dot", "examples/dot.hyd:1-7",
        dict(xs=[1.0, 2.0, 3.0], ys=[2.0, 3.0, 4.0]), 20.0, derived=True)
    add("dot_float", "dot", "test/Language/Hydrangea/EndToEndSpec.hs:991-997",
        dict(xs=[1.0, 2.0, 3.0, 4.0], ys=[1.0, 2.0, 3.0, 4.0]), 30.0, derived=True)

    # ---- stencils ----
    add("stencil1d_box3", "stencil1d", "test/Language/Hydrangea/StencilSpec.hs:24-35",
        dict(a=[1, 2, 3, 4, 5], offsets=[-1, 0, 1]), [4, 6, 9, 12, 14])
    add("stencil1d_identity", "stencil1d", "test/Language/Hydrangea/StencilSpec.hs:37-41",
        dict(a=[0, 2, 4, 6], offsets=[0]), [0, 2, 4, 6])
    add("stencil1d_neg5", "stencil1d", "test/Language/Hydrangea/StencilSpec.hs:43-48",
        dict(a=[10, 11, 12], offsets=[-5]), [10, 10, 10])
    add("stencil1d_pos5", "stencil1d", "test/Language/Hydrangea/StencilSpec.hs:50-55",
        dict(a=[10, 11, 12], offsets=[5]), [12, 12, 12])
    add("stencil1d_single", "stencil1d", "test/Language/Hydrangea/StencilSpec.hs:57-62",
        dict(a=[42], offsets=[-1, 0, 1]), [126])
    add("stencil2d_corner_4sum", "stencil2d_jacobi4_times4",
        "test/Language/Hydrangea/StencilSpec.hs:176-185",
        dict(grid=[[1, 2, 3], [4, 5, 6], [7, 8, 9]], at=[0, 0]), 8.0)
    add("laplacian_centre_example", "stencil2d_laplacian", "examples/stencil_2d.hyd:1-19",
        dict(grid=[[1, 2, 3], [4, 5, 6], [7, 8, 9]], at=[1, 1]), 0.0)
    add("jacobi_ones_fixed_point", "jacobi_iterate", "bench/stencil/jacobi_2d.hyd:1-7",
        dict(h=16, w=24, iters=5, value=1.0), 1.0)
    return c


def main():
    out = os.path.join(HERE, "known_answers.json")
    with open(out, "w") as f:
        json.dump(cases(), f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main()
