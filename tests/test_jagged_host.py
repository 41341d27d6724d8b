"""Host-side jagged LogUp-GKR (no GPU): the jagged circuit against the
oracle's, the layout helpers, and the prover's structural rejections, which
fire before any device work as the reference's do (zorch/logup_gkr/
jagged_prover.py:159-170, 320-345; circuit.py:221-245, 446-470, 518-522)."""

from __future__ import annotations

import random

import numpy as np
import pytest

from oracle import fields as OF
from oracle import jagged as OJ
from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import FieldArray, as_field, from_storage
from zorch_amd.logup_gkr import (JaggedGkrLayer, RoundWidthCaps, build_jagged_pyramid,
                                 extract_jagged_outputs, jagged_layer_transition,
                                 prove_jagged_layer)
from zorch_amd.sumcheck.jagged import _prepad_folded, _segment_gather_np

F = OF.BABYBEAR4
FIELD = "babybearx4_mont"


def _fa(vals):
    f = as_field(FIELD)
    return from_storage(np.array([F.to_words(v) for v in vals], dtype=np.uint32)
                        .reshape(len(vals), 4), f)


def _words(fa):
    return fa.normalized().words.reshape(-1).tolist()


def _ow(vals):
    return [int(w) for v in vals for w in F.to_words(v)]


def _pair(counts, seed=3):
    rng = random.Random(seed)
    w = sum(counts)
    vals = [[F.from_int(rng.randrange(1 << 62)) for _ in range(w)] for _ in range(4)]
    return OJ.JaggedLayer(*vals, tuple(counts)), JaggedGkrLayer(*(_fa(v) for v in vals),
                                                              tuple(counts))


def test_pyramid_and_outputs_match_oracle():
    counts = (3, 1, 5, 2, 6, 1, 8, 4)
    ol, dl = _pair(counts)
    sched = OJ.jagged_fold_schedules(counts)
    olayers = OJ.build_jagged_pyramid(F, ol, sched)
    dlayers = build_jagged_pyramid(dl, sched)
    for o, d in zip(olayers, dlayers):
        assert d.row_counts == o.row_counts
        for a, b in zip((d.numerator_0, d.numerator_1, d.denominator_0, d.denominator_1),
                        (o.n0, o.n1, o.d0, o.d1)):
            assert _words(a) == _ow(b)
    num, den = OJ.extract_jagged_outputs(F, olayers[-1])
    out = extract_jagged_outputs(dlayers[-1])
    assert _words(out.numerator) == _ow(num) and _words(out.denominator) == _ow(den)
    with pytest.raises(ValueError, match="expects the batch floor"):
        extract_jagged_outputs(dlayers[0])


def test_transition_capacity_and_schedule_checks():
    _, dl = _pair((3, 1, 5, 2))
    wide = jagged_layer_transition(dl, (2, 1, 3, 1), out_width=10)
    assert wide.width == 10 and _words(wide.numerator_0[7:]) == [0] * 12
    assert _words(wide.denominator_0[7:]) == [0] * 12  # dead slots are zero, not neutral
    with pytest.raises(ValueError, match="cannot hold"):
        jagged_layer_transition(dl, (2, 1, 3, 1), out_width=6)
    with pytest.raises(ValueError, match="truncates"):
        jagged_layer_transition(dl, (1, 1, 3, 1))
    with pytest.raises(ValueError, match="cover all 4 batches"):
        jagged_layer_transition(dl, (2, 1, 3))
    with pytest.raises(ValueError, match="share one capacity width"):
        JaggedGkrLayer(dl.numerator_0, dl.numerator_1[:3], dl.denominator_0,
                       dl.denominator_1, dl.row_counts)
    with pytest.raises(ValueError):
        JaggedGkrLayer(dl.numerator_0, dl.numerator_1, dl.denominator_0, dl.denominator_1,
                       (3, 1, 7))


def test_layout_helpers():
    assert _segment_gather_np((3, 1), (3, 1)) is None
    assert _segment_gather_np((3, 1, 5, 2), (4, 2, 6, 2)).tolist() == \
        [0, 1, 2, 11, 3, 11, 4, 5, 6, 7, 8, 11, 9, 10]
    assert _prepad_folded((3, 1, 5, 2)) == ((4, 2, 6, 2), (2, 1, 3, 1))


def test_prover_rejects_before_device_work():
    _, dl = _pair((3, 1, 5, 2))
    one = _fa([F.one()])[0]
    ch = ChallengePolicy(FIELD)
    z5 = _fa([F.from_int(i + 3) for i in range(5)])
    caps = RoundWidthCaps(elements=12, eq_row=8, interaction=4)
    with pytest.raises(ValueError, match="proves under caps"):
        prove_jagged_layer(dl, one, one, z5, None, challenges=ch)
    with pytest.raises(ValueError, match="at least one row variable"):
        prove_jagged_layer(dl, one, one, z5[:2], None, challenges=ch, caps=caps)
    with pytest.raises(ValueError, match="exceeds the virtual row space"):
        prove_jagged_layer(dl, one, one, z5[:4], None, challenges=ch, caps=caps)
    for bad, msg in [(RoundWidthCaps(10, 8, 4), "multiple of 4"),
                     (RoundWidthCaps(12, 4, 4), "row-eq table"),
                     (RoundWidthCaps(12, 8, 2), "multiple of 4"),
                     (RoundWidthCaps(8, 8, 4), "plane width")]:
        with pytest.raises(ValueError, match=msg):
            prove_jagged_layer(dl, one, one, z5, None, challenges=ch, caps=bad)


def test_mixed_field_first_layer_promotes():
    """Base-field numerators under extension denominators: the transition
    promotes and equals the all-extension layer's (reference
    jagged_circuit_test.py:222-259)."""
    import numpy as np
    from zorch_amd.fields import babybear_mont
    rng = np.random.default_rng(8)
    counts = (3, 1, 5, 2)
    w = sum(counts)
    nb = [FieldArray(rng.integers(0, OF.BABYBEAR.p, size=(w, 1), dtype=np.uint64)
                     .astype(np.uint32), babybear_mont) for _ in range(2)]
    db = [FieldArray(rng.integers(0, OF.BABYBEAR.p, size=(w, 4), dtype=np.uint64)
                     .astype(np.uint32), as_field(FIELD)) for _ in range(2)]
    mixed = JaggedGkrLayer(nb[0], nb[1], db[0], db[1], counts)
    ext = JaggedGkrLayer(nb[0].astype(FIELD), nb[1].astype(FIELD), db[0], db[1], counts)
    a = jagged_layer_transition(mixed, (2, 1, 3, 1))
    b = jagged_layer_transition(ext, (2, 1, 3, 1))
    for x, y in zip((a.numerator_0, a.numerator_1, a.denominator_0, a.denominator_1),
                    (b.numerator_0, b.numerator_1, b.denominator_0, b.denominator_1)):
        assert x.field is y.field and _words(x) == _words(y)
