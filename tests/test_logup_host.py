"""Structural checks of ZS_MODE_LOGUP requests (no GPU): the C ABI rejects a
malformed LogUp stage before any device work, as the reference's summand
rejects a wrong factor count (zorch/logup_gkr/prover.py:95-101), and the
host-side circuit and combine agree with the oracle."""

from __future__ import annotations

import random

import numpy as np
import pytest

from oracle import fields as OF
from oracle import logup as OL
from zorch_amd import _lib
from zorch_amd.fields import FieldArray, as_field, from_storage, stack
from zorch_amd.challenge import ChallengePolicy
from zorch_amd.logup_gkr import (GkrLayer, GkrProof, LogUpGkrProver, LogUpGkrVerifier,
                                 LogUpOutputClaim, LogupSummand, build_pyramid, extract_outputs,
                                 logup_combine)
from zorch_amd.logup_gkr.prover import expand_eq_to_hypercube, eval_mle
from zorch_amd.runtime import message_width


def _args(k=5, deg=3, tf=_lib.ZS_F_BABYBEAR4, cf=_lib.ZS_F_BABYBEAR4, lam=True, eq=False,
          wire=None, refold=0, n=4):
    keep = np.zeros(64, np.uint32)
    return _lib.ProveArgsC(table_field=tf, chal_field=cf, mode=_lib.ZS_MODE_LOGUP, degree=deg,
                           num_tables=k, rounds=n, tables=keep.ctypes.data, tables_on_device=0,
                           refold_rounds=refold, eq_point=keep.ctypes.data if eq else None,
                           claim=keep.ctypes.data, wire=wire,
                           lambda_=keep.ctypes.data if lam else None), keep


def test_logup_request_shape():
    a, keep = _args()
    assert message_width(a) == 4
    for kw, msg in [(dict(k=4), "5 factors"), (dict(deg=2), "5 factors"),
                    (dict(tf=_lib.ZS_F_BABYBEAR), "challenge field"),
                    (dict(lam=False), "batching challenge"), (dict(eq=True), "eq table"),
                    (dict(refold=1), "refold_rounds")]:
        a, keep = _args(**kw)
        with pytest.raises(ValueError, match=msg):
            message_width(a)
    w = _lib.WireC(kind=_lib.ZS_WIRE_COEFFS)
    import ctypes
    a, keep = _args(wire=ctypes.pointer(w))
    with pytest.raises(ValueError, match="natural values"):
        message_width(a)


def _fa(F, field, vals):
    f = as_field(field)
    words = np.array([F.to_words(v) for v in vals], dtype=f.word_dtype).reshape(len(vals), f.words)
    return from_storage(words, f)


def _words(fa: FieldArray) -> list:
    return fa.normalized().words.reshape(-1).tolist()


def _owords(F, vals) -> list:
    return [int(w) for v in vals for w in F.to_words(v)]


@pytest.mark.parametrize("name", ["babybearx4_mont", "goldilocks"])
def test_host_circuit_and_combine_match_oracle(name):
    F = {"babybearx4_mont": OF.BABYBEAR4, "goldilocks": OF.GOLDILOCKS}[name]
    rng = random.Random(4)
    vals = [[F.from_int(rng.randrange(1 << 62)) for _ in range(16)] for _ in range(4)]
    layers = build_pyramid(GkrLayer(*(_fa(F, name, v) for v in vals), 1))
    olayers = OL.build_pyramid(F, OL.Layer(*vals, 1))
    for l, o in zip(layers, olayers):
        assert _words(stack([l.numerator_0, l.numerator_1, l.denominator_0, l.denominator_1])) \
            == _owords(F, o.n0 + o.n1 + o.d0 + o.d1)
    out = extract_outputs(layers[-1])
    onum, oden = OL.extract_outputs(F, olayers[-1])
    assert _words(out.numerator) == _owords(F, onum) and _words(out.denominator) == _owords(F, oden)
    lam = _fa(F, name, [F.from_int(6)])[0]
    f5 = [_fa(F, name, [F.from_int(v)])[0] for v in (2, 3, 4, 5, 7)]
    got = LogupSummand(lam).combine((lam,), *f5)
    assert _words(got) == _owords(F, [OL.logup_combine(F, F.from_int(6), *[F.from_int(v) for v in
                                                                          (2, 3, 4, 5, 7)])])
    assert _words(logup_combine(lam, *f5)) == _words(got)
    with pytest.raises(ValueError, match="5 factors"):
        LogupSummand(lam).combine((lam,), *f5[:4])
    point = [F.from_int(rng.randrange(1 << 62)) for _ in range(3)]
    eq = expand_eq_to_hypercube(_fa(F, name, point), _fa(F, name, [F.one()])[0])
    from oracle.sumcheck import eq_table, eval_mle as o_eval_mle
    assert _words(eq) == _owords(F, eq_table(F, point))
    assert _words(eval_mle(_fa(F, name, vals[0][:8]), _fa(F, name, point))) == \
        _owords(F, [o_eval_mle(F, vals[0][:8], point)])


def test_stage_rejects_a_layer_count_mismatch():
    """The verifier owns the layer count: a witness or proof of another
    height is rejected before any work (reference zorch/logup_gkr/stage.py:75-79,114-118)."""
    F = OF.BABYBEAR4
    rng = random.Random(2)
    vals = [[F.from_int(rng.randrange(1 << 62)) for _ in range(16)] for _ in range(4)]
    layer = GkrLayer(*(_fa(F, "babybearx4_mont", v) for v in vals), 1)
    out = extract_outputs(build_pyramid(layer)[-1])
    ch = ChallengePolicy("babybearx4_mont")
    with pytest.raises(ValueError, match="claim expects 2 GKR layers, witness needs 3"):
        LogUpGkrProver(ch).prove(LogUpOutputClaim(out, 2), layer, None)
    with pytest.raises(ValueError, match="expected 3 GKR layers, got 2"):
        LogUpGkrVerifier(ch).verify(LogUpOutputClaim(out, 3), GkrProof((None, None)), None)
