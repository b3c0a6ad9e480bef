"""The autotuner's selection rule (sq_tune_select, the rule sq_create_from_deck
applies to its probe times), on the CPU: candidates 0/1 = one-node kernel with
512/1024 threads, 2/3 = 2 nodes a workgroup with gather trip a/b, 4/5 = 4 nodes
with trip a/b (FP64 trips 1/2 planes, FP32 trips 4 planes on the matrix cores / 2
on the vector ALU); the environment overrides
SQ_ZMARCH_WG, SQ_BATCH and SQ_BATCH_TRIP narrow the set."""
import pytest

import sqminiapp_amd as sq

ALL = [1] * 6
FP64_AV = [1, 1, 1, 1, 0, 0]     # the batched FP64 kernel is built for B = 2 only


def test_fastest_available_candidate():
    assert sq.tune_select(sq.SQ_FP32, ALL, [5.0, 4.0, 6.0, 3.0, 2.0, 2.5]) == 4
    assert sq.tune_select(sq.SQ_FP32, ALL, [1.0, 4.0, 6.0, 3.0, 2.0, 2.5]) == 0
    # the C128 probe times of the bench run: FP64 keeps the one-node 1024-thread kernel, FP32 takes B = 4
    assert sq.tune_select(sq.SQ_FP64, FP64_AV, [79.665, 63.802, 86.345, 76.232, 0.0, 0.0]) == 1
    assert sq.tune_select(sq.SQ_FP32, ALL, [47.684, 51.946, 37.336, 37.476, 30.7, 32.661]) == 4


def test_unavailable_candidates_are_never_chosen():
    # a zero probe time of a candidate that cannot launch does not win
    assert sq.tune_select(sq.SQ_FP64, FP64_AV, [3.9, 4.1, 6.0, 5.1, 0.0, 0.0]) == 0
    assert sq.tune_select(sq.SQ_FP32, [1, 0, 0, 0, 0, 0], [9.0, 1.0, 1.0, 1.0, 1.0, 1.0]) == 0
    assert sq.tune_select(sq.SQ_FP32, [0] * 6, [1.0] * 6) == -1


def test_ties_take_the_lower_index():
    assert sq.tune_select(sq.SQ_FP32, ALL, [2.0, 2.0, 2.0, 2.0, 2.0, 2.0]) == 0
    assert sq.tune_select(sq.SQ_FP32, [0, 0, 1, 1, 1, 1], [0.0, 0.0, 3.0, 2.0, 2.0, 4.0]) == 3


@pytest.mark.parametrize("wg,expect", [("512", 0), ("1024", 1)])
def test_workgroup_override_keeps_the_one_node_kernel(wg, expect):
    assert sq.tune_select(sq.SQ_FP32, ALL, [9.0, 8.0, 1.0, 1.0, 1.0, 1.0], workgroup=wg) == expect


def test_batch_and_trip_overrides():
    ms = [5.0, 4.0, 6.0, 3.0, 2.0, 2.5]
    assert sq.tune_select(sq.SQ_FP32, ALL, ms, batch=1) == 1
    assert sq.tune_select(sq.SQ_FP32, ALL, ms, batch=2) == 3
    assert sq.tune_select(sq.SQ_FP32, ALL, ms, batch=4) == 4
    assert sq.tune_select(sq.SQ_FP32, ALL, ms, batch=4, trip=2) == 5
    assert sq.tune_select(sq.SQ_FP32, ALL, ms, batch=2, trip=4) == 2
    assert sq.tune_select(sq.SQ_FP64, FP64_AV, [5.0, 4.0, 3.0, 6.0, 0.0, 0.0], batch=2, trip=2) == 3
    # a trip that the precision is not built with leaves the one-node kernels
    assert sq.tune_select(sq.SQ_FP64, FP64_AV, [5.0, 4.0, 3.0, 6.0, 0.0, 0.0], trip=8) == 1


def test_narrowed_to_nothing_falls_back_to_candidate_zero():
    # FP64 has no B = 4 kernel: SQ_BATCH=4 leaves nothing, and the 512-thread one-node kernel runs
    assert sq.tune_select(sq.SQ_FP64, FP64_AV, [5.0, 4.0, 3.0, 6.0, 0.0, 0.0], batch=4) == 0
    assert sq.tune_select(sq.SQ_FP32, [0, 1, 1, 1, 1, 1], [1.0] * 6, batch=3) == -1


def test_bad_precision_rejected():
    with pytest.raises(sq.SQError):
        sq.tune_select(7, ALL, [1.0] * 6)
