"""Sub-island regroups and DE configuration, host side (no GPU):
speo_island_regroup against the oracle's restatement of Migrate::RegroupIslands
(src/migrate.cc:147-186), Random::Permutate (src/random.cc:29-53) and
DE_CUDA::ConfigureEA (CUDA_EA_library/DE_CUDA.cc:89-119)."""
import itertools

import numpy as np
import pytest


def test_permutate_hand_worked(oracle):
    # n = 5, injected words 2, 0, 1, 0: take g[2] = 2 (g -> 0 1 4 3), g[0] = 0
    # (g -> 3 1 4), g[1] = 1 (g -> 3 4), g[0] = 3 (g -> 4), then the last one
    # without a draw (RandIntUnif(0, 0) returns 0 and draws nothing).
    d = oracle.HostDraws(0, 0, 0, words=[2, 0, 1, 0])
    assert oracle.permutate(d, 5) == [2, 0, 1, 3, 4]
    assert d.k == 4


def test_rand_int_matches_the_modulo_form(oracle):
    d = oracle.HostDraws(0, 0, 0, words=[7, 10, 3])
    assert d.rand_int(0, 2) == 7 % 3
    assert d.rand_int(4, 4) == 4 and d.k == 1    # a == b: no draw
    assert d.rand_int(5, 9) == 5 + 10 % 5


@pytest.mark.parametrize("option,configure", list(itertools.product(range(4), range(2))))
def test_library_regroup_equals_oracle(speo, oracle, option, configure):
    base = (0.9, 0.5, 4)
    for np_, seed, island in ((16, 1, 0), (64, 2026, 3), (256, 99, 17)):
        subs, perm = np_ // 4, list(range(np_))
        s_o, p_o, i_o = oracle.island_regroup(seed, island, 0, option, configure, np_, subs, perm,
                                              base, None)
        s_l, p_l, i_l = speo.island_regroup(seed, island, 0, option, configure, np_, subs, perm,
                                            base)
        assert (s_o, list(p_o)) == (s_l, list(p_l))
        assert np.allclose(i_o, i_l)
        subs, perm, info = s_l, list(p_l), i_l
        for event in range(1, 6):
            s_o, p_o, i_o = oracle.island_regroup(seed, island, event, option, configure, np_,
                                                  subs, perm, base, info)
            s_l, p_l, i_l = speo.island_regroup(seed, island, event, option, configure, np_,
                                                subs, perm, base, info)
            assert s_o == s_l and list(p_o) == list(p_l), (np_, event)
            assert [tuple(map(float, t)) for t in i_o] == [tuple(map(float, t)) for t in i_l]
            subs, perm, info = s_l, list(p_l), i_l
            # invariants
            assert sorted(perm) == list(range(np_))
            assert np_ % subs == 0 and np_ // subs >= 4
            if option in (2, 3):
                size = np_ // subs
                assert size & (size - 1) == 0
            else:
                assert subs == np_ // 4
            if option in (0, 3):
                assert perm == list(p_o)
            for cr, f, st, ps in info:
                assert ps == np_ // subs
                if configure == 1:
                    assert st in (4, 6) and round(cr, 6) in (0.1, 0.9) and f == 0.5


def test_static_ordered_keeps_partition_and_configuration(speo):
    info = [(0.1, 0.5, 6, 16), (0.9, 0.5, 4, 16)]
    s, p, i = speo.island_regroup(5, 0, 3, 0, 1, 32, 2, list(range(32))[::-1], (0.9, 0.5, 4),
                                  info)
    assert s == 2 and list(p) == list(range(32))[::-1]
    assert [(round(a, 6), b, c, d) for a, b, c, d in i] == [(0.1, 0.5, 6, 16), (0.9, 0.5, 4, 16)]


def test_dynamic_regroup_draw_count(oracle):
    # dynamic_and_random with dynamic configuration: one size draw, n - 1
    # permutation draws, two draws per sub-island
    d_words = list(range(1000))
    np_ = 64
    d = oracle.HostDraws(0, 0, 1, words=d_words)
    s, p, i = oracle.island_regroup(0, 0, 1, 2, 1, np_, 16, list(range(np_)), (0.9, 0.5, 4),
                                    [(0.9, 0.5, 4, 4)] * 16, words=d_words)
    size = 4 << (0 % 5)                        # first word 0 -> size 4
    assert s == np_ // size
    # permutation: words 1 .. 63; configuration: words 64 .. 64 + 2 s - 1
    d2 = oracle.HostDraws(0, 0, 1, words=d_words[1:])
    assert p == oracle.permutate(d2, np_)


def test_regroup_rejects_bad_arguments(speo):
    with pytest.raises(speo.SpeoError):
        speo.island_regroup(1, 0, 1, 7, 0, 64, 16, list(range(64)), (0.9, 0.5, 4))
    with pytest.raises(speo.SpeoError, match="4 \\* 2\\^k"):
        # 48 rows: a drawn size of 32 does not divide it
        for seed in range(64):
            speo.island_regroup(seed, 0, 1, 2, 0, 48, 12, list(range(48)), (0.9, 0.5, 4))
