"""The frame-level parity tests again with two waves per search (the
analysis kernel's other launch shape): results must not depend on it."""
import pytest

import test_gpu_parity as base

pytestmark = pytest.mark.gpu


@pytest.fixture
def ctx2(ctx):
    ctx.set_tuning(waves_per_search=2)
    assert ctx.tuning()[0]["waves_per_search"] == 2
    yield ctx
    ctx.set_tuning(waves_per_search=1)


@pytest.mark.parametrize("quality,noise", [(8, True), (13, False)])
def test_frame_p_single_ref(ctx2, oracle, quality, noise):
    base.test_frame_p_single_ref(ctx2, oracle, quality, noise)


@pytest.mark.parametrize("n_refs,quality", [(2, 8), (4, 8), (3, 31), (4, 0)])
def test_frame_p_multi_ref(ctx2, oracle, n_refs, quality):
    base.test_frame_p_multi_ref(ctx2, oracle, n_refs, quality)


@pytest.mark.parametrize("span", base.RANGE_CASES[1:])
def test_frame_p_sample_ranges(ctx2, oracle, span):
    base.test_frame_p_sample_ranges(ctx2, oracle, span)


def test_frame_p_temporal_chain(ctx2, oracle):
    base.test_frame_p_temporal_chain(ctx2, oracle)


@pytest.mark.parametrize("quality", [9, 13, 63])
def test_frame_b(ctx2, oracle, quality):
    base.test_frame_b(ctx2, oracle, quality)


def test_gop8_hierarchical_sequence(ctx2, oracle):
    base.test_gop8_hierarchical_sequence(ctx2, oracle)


def test_batch_uneven_streams_match_alone(ctx2, oracle):
    base.test_batch_uneven_streams_match_alone(ctx2, oracle)


def test_graph_replays_match_eager(ctx2, oracle):
    base.test_graph_replays_match_eager(ctx2, oracle)
