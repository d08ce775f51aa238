"""Cylinder layout (layout_cylinder.h:152-195) pinned by brute force.

Both transcriptions (the oracle's or_cylinder and the library's host-side
sct_cylinder_layout) are checked against disc membership computed directly:
voxel (row, col) is in the mask iff |row - yc| < sqrt(r^2 - (col - xc)^2),
rows clipped to [0, height - 1].
"""
import math

import numpy as np
import pytest

import oracle as O
import segmentctscans_amd as S

CASES = [
    (64, 64, 64, 31.5, 31.5, 28.8),      # config 1
    (10, 12, 16, 7.5, 5.5, 5.0),         # the 16 x 12 x 10 golden cylinder
    (5, 5, 5, 2.0, 2.0, 2.0),            # radius hits integer rows exactly
    (7, 9, 11, 5.0, 4.0, 4.0),           # odd sizes, integer centre
    (4, 6, 6, 2.5, 2.5, 0.7),            # r < 1
    (3, 8, 8, 3.5, 3.5, 100.0),          # full box (r > half diagonal)
    (6, 10, 10, 4.5, 4.5, 3.2),
    (2, 1, 1, 0.0, 0.0, 5.0),            # a single line
]


def brute(height, width, xc, yc, r):
    member = np.zeros((width, height), bool)
    for col in range(width):
        d2 = r * r - (col - xc) ** 2
        if d2 <= 0:
            continue
        d = math.sqrt(d2)
        for row in range(height):
            member[col, row] = abs(row - yc) < d
    return member


@pytest.mark.parametrize("n,h,w,xc,yc,r", CASES)
def test_cylinder_membership(n, h, w, xc, yc, r):
    m = S.Mapping.cylinder(n, h, w, xc, yc, r)
    lay = O.Layout(n, h, w, xc, yc, r)
    assert np.array_equal(m.base, lay.base)
    assert np.array_equal(m.start_row, lay.start_row)
    assert (m.col_begin, m.col_end, m.data_size) == (lay.col_begin, lay.col_end, lay.data_size)
    member = brute(h, w, xc, yc, r)
    got = np.zeros((w, h), bool)
    for col in range(m.col_begin, m.col_end):
        got[col, m.row_begin(col):m.row_end(col)] = True
    assert np.array_equal(got, member)
    assert m.data_size == member.sum() * n
    # columns outside [col_begin, col_end) hold no rows
    for col in range(w):
        if not (m.col_begin <= col < m.col_end):
            assert m.base[col + 1] == m.base[col]
            assert m.start_row[col] == 0xFFFFFFFF


@pytest.mark.parametrize("n,h,w,xc,yc,r", CASES)
def test_offsets_are_a_bijection(n, h, w, xc, yc, r):
    m = S.Mapping.cylinder(n, h, w, xc, yc, r)
    assert m.nstride == S.get_stride(n) and m.nstride % 32 == 0
    seen = np.zeros(m.span, np.int32)
    for col in range(m.col_begin, m.col_end):
        for row in range(m.row_begin(col), m.row_end(col)):
            for k in range(n):
                off = int(m.base[col]) + (row - m.row_begin(col)) * m.nstride + k
                seen[off] += 1
    assert seen.max(initial=0) <= 1
    assert seen.sum() == m.data_size
    assert np.array_equal(seen.astype(bool), m.valid_mask())


def test_stride_helpers():
    lib = S._native.lib()
    for n in (0, 1, 31, 32, 33, 1426, 1440):
        assert lib.sct_get_stride(n) == S.get_stride(n) == ((n + 31) // 32) * 32
        assert lib.sct_prev_stride(n) == S.prev_stride(n) == (n // 32) * 32


def test_config3_counts():
    """Config 3 in-mask voxel count quoted in the measurement plan."""
    m = S.Mapping.cylinder(1426, 1024, 1024, 511.5, 511.5, 460.0)
    assert m.data_size == 947_999_096
    assert m.lines == 664_796 and m.nstride == 1440 and m.span == 957_306_240
