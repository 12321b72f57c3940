"""Host logic of the fused slot decoder (no GPU): the kernel variant of a
TB's code blocks by block rows, and how row classes are merged into
launches (aerial_amd.slot.launch_classes)."""
import pytest

pytest.importorskip("torch")


def test_row_class_boundaries():
    from aerial_amd.slot import row_class
    assert [row_class(mb) for mb in (4, 8, 9, 16, 17, 42, 46)] == [0, 0, 1, 1, 2, 2, 2]


@pytest.mark.parametrize("counts, split_min, expect", [
    # the 64-cell fused slot: BG1 mb = 4 and BG2 mb = 12 code blocks, each
    # class fills the GPU: two launches
    ((8960, 1792, 0), 256, {0: 0, 1: 1}),
    # too few high-rate CBs to fill the GPU: they join the full variant
    ((100, 0, 3000), 256, {0: 2, 2: 2}),
    ((100, 50, 3000), 256, {0: 2, 1: 2, 2: 2}),
    # small + middle together reach split_min: one launch of the 16-row variant
    ((200, 100, 3000), 256, {0: 1, 1: 1, 2: 2}),
    # the largest class present takes what is left, however few
    ((300, 100, 0), 256, {0: 0, 1: 1}),
    ((100, 0, 0), 256, {0: 0}),
    # split_min = 1: every class present has its own launch
    ((5, 3, 2), 1, {0: 0, 1: 1, 2: 2}),
])
def test_launch_classes(counts, split_min, expect):
    from aerial_amd.slot import launch_classes
    got = launch_classes(list(counts), split_min)
    # only classes with code blocks matter
    assert {c: got[c] for c in range(3) if counts[c] > 0} == expect
    # a class never goes to a kernel variant with fewer rows
    for c in range(3):
        assert got[c] >= c
