"""The algorithmic-byte accounting of bench.py (CPU only).

The per-iteration figure B_alg = 136 T + 140 V is split over the kernels
that do the work: the projection (72 T + 28 V, plus the 60 V Chebyshev blend
of the previous iteration in every launch but a substep's first), the Jacobi
average (40 V), the residual (64 T + 12 V) and, once per substep, the final
blend (60 V).  Over a substep the parts must add up to iters x B_alg exactly,
so that no byte is counted twice or dropped.
"""
import pytest

import bench

# (T, V, iterations per substep): the 64^3 jittered box, the 16^3 cube, the
# 256-scene batch, and a substep of one iteration
CASES = [(1_572_864, 274_625, 100), (24_576, 4_913, 30), (19_999_224, 3_777_906, 50),
         (6, 8, 1)]


@pytest.mark.parametrize("T,V,iters", CASES)
def test_kernel_bytes_partition_the_iteration(T, V, iters):
    project_total = iters * (72 * T + 28 * V) + (iters - 1) * 60 * V
    average_total = iters * 40 * V
    residual_total = iters * (64 * T + 12 * V)
    final_blend = 60 * V
    assert (project_total + average_total + residual_total + final_blend
            == iters * bench.b_alg_iteration(T, V))
    # the per-launch figure is the projection's total spread over its launches
    # (rounded down to whole bytes)
    assert bench.b_alg_project(T, V, iters) == project_total // iters


def test_jittered_box_figures():
    T, V = 1_572_864, 274_625
    assert bench.b_alg_iteration(T, V) == 252_357_004
    assert bench.b_alg_project(T, V, 100) == 137_248_433
