"""Host-side mapping of the joins' d_result words to their stats records
(vsjoin.h documents the word order): each field reads its own word."""

from __future__ import annotations

from vibespatial_amd import join as J


def test_learned_stats_fields_follow_the_result_words():
    h = [1000 + i for i in range(16)]
    st = J._learned_stats(h, cap=77, retries=2)
    assert (st.pairs, st.staged, st.candidates, st.capacity, st.retries) == (1000, 1001, 1002, 77, 2)
    assert st.candidate_vertices == 1003
    assert (st.orient_tests, st.escalated_edges, st.escalated_pairs) == (1004, 1005, 1006)
    assert (st.lookup_probes, st.mlp_fallbacks, st.window_fallbacks) == (1007, 1008, 1009)
    assert (st.scanned, st.mlp_leaf_error) == (1010, 1011)


def test_batch_stats_extend_join_stats_without_shifting_them():
    st = J.BatchStats(1, 2, 3, 4, 5, 6, 7)
    assert (st.pairs, st.staged, st.candidates, st.capacity, st.retries, st.candidate_vertices, st.refined) == (
        1, 2, 3, 4, 5, 6, 7)
    assert isinstance(st, J.JoinStats)
    # the base record keeps its positional layout (the learned stats extend it)
    assert [f.name for f in J.JoinStats.__dataclass_fields__.values()] == [
        "pairs", "staged", "candidates", "capacity", "retries", "candidate_vertices"]
