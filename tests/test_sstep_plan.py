"""The s-step shape rule of the bench and the profiles (mek.sstep_plan): the
first 8 columns one by one, then the fewest equal blocks of at most 16 that
cover the rest (no GPU needed)."""
import pytest

from conftest import mek


@pytest.mark.parametrize("m,want", [(40, (16, 8)), (50, (14, 8)), (30, (11, 8)), (60, (13, 8)), (20, (12, 8)),
                                    (10, (2, 8)), (9, (0, -1)), (8, (0, -1)), (1, (0, -1))])
def test_sstep_plan(m, want):
    s, first = mek.sstep_plan(m)
    assert (s, first) == want
    if s:
        rest = m - first
        nb = -(-rest // s)
        assert 2 <= s <= 16 and nb == -(-rest // 16) and s * (nb - 1) < rest <= s * nb


def test_sstep_plan_keeps_explicit_shapes():
    assert mek.sstep_plan(40, 12, 6) == (12, 6)
    assert mek.sstep_plan(40, 0, -1) == (0, -1)
    assert mek.sstep_plan(40, -1, 4) == (12, 4)
