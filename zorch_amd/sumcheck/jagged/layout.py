"""Segment layout recurrences of the jagged sumcheck (drop-in for reference
zorch/sumcheck/jagged/layout.py:14-53): the gather that re-lays a jagged
layout from one set of segment counts to another, and the even-prepad /
folded-count recurrence of one transition."""

from __future__ import annotations

import numpy as np


def _segment_gather_np(src_counts, dst_counts):
    """Indices re-laying `src_counts` segments into `dst_counts` ones: the
    first min(src, dst) rows of each segment copy over, later positions get
    the sentinel sum(src_counts); None when the layouts agree."""
    src_counts, dst_counts = tuple(src_counts), tuple(dst_counts)
    if src_counts == dst_counts:
        return None
    if len(src_counts) != len(dst_counts):
        raise ValueError("segment count mismatch")
    sentinel = sum(src_counts)
    gather = np.full(sum(dst_counts), sentinel, dtype=np.int32)
    src_off = np.concatenate([[0], np.cumsum(src_counts)[:-1]]).astype(np.int64)
    dst_off = np.concatenate([[0], np.cumsum(dst_counts)[:-1]]).astype(np.int64)
    for s, (src, dst) in enumerate(zip(src_counts, dst_counts)):
        c = min(src, dst)
        gather[dst_off[s]:dst_off[s] + c] = np.arange(src_off[s], src_off[s] + c)
    return gather


def _prepad_folded(row_counts):
    """(even-prepadded counts, folded counts) of one transition."""
    prepad = tuple(rc + rc % 2 for rc in row_counts)
    return prepad, tuple(pc // 2 for pc in prepad)
