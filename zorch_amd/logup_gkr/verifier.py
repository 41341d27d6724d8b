"""Dense LogUp-GKR verifier (drop-in for reference zorch/logup_gkr/verifier.py).

``GkrLayerRound`` replays one layer's sumcheck (``verify`` over
``SumcheckRound(3)``, the library's host replay), checks the LogUp oracle at
the bound point, combine(lam, eq(point, bound point), openings) == the final
claim, then binds the openings at the child selector (verifier.py:37-72).
It stops at the reduced point claim: the check against the input's MLE is
the consumer's.
"""

from __future__ import annotations

from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import FieldArray, as_field, ones, stack
from zorch_amd.logup_gkr.prover import (DEGREE, LayerProof, bind_output, fold_carry,
                                        logup_combine)
from zorch_amd.round import verify_rounds
from zorch_amd.sumcheck.verifier import SumcheckRound
from zorch_amd.verify import verify


def eval_eq(a: FieldArray, b: FieldArray) -> FieldArray:
    """prod_i (a_i b_i + (1 - a_i)(1 - b_i)) (reference zorch/poly/eq.py eval_eq)."""
    one = ones((), a.field)
    acc = one
    for i in range(a.shape[0]):
        acc = acc * (a[i] * b[i] + (one - a[i]) * (one - b[i]))
    return acc


class GkrLayerRound:
    """Verify one GKR layer; the chain of these is the GKR verifier."""

    def __init__(self, challenges: ChallengePolicy) -> None:
        self.challenges = challenges

    def __call__(self, claim, transcript, layer_proof: LayerProof):
        num_eval, den_eval, eval_point = claim
        n0, n1 = layer_proof.numerator_0, layer_proof.numerator_1
        d0, d1 = layer_proof.denominator_0, layer_proof.denominator_1
        transcript, lam = self.challenges.sample(transcript)
        lam = lam.normalized()
        claim_value = lam * num_eval + den_eval
        point, final_claim, transcript, ok = verify(SumcheckRound(DEGREE, self.challenges),
                                                    claim_value, layer_proof.round_polys,
                                                    transcript)
        point = point.normalized()
        if eval_point.shape[0] != point.shape[0]:
            raise ValueError(f"eq point mismatch: claim has {eval_point.shape[0]} coords, "
                             f"layer ran {point.shape[0]} rounds")
        combined = logup_combine(lam, eval_eq(eval_point, point), n0, d1, n1, d0)
        ok = ok and bool((combined == final_claim).all())
        transcript, r = self.challenges.observe_and_sample(transcript, stack([n0, n1, d0, d1]))
        return fold_carry(n0, n1, d0, d1, point, r.normalized()), transcript, ok


def verify_gkr(output, proofs, transcript, challenges: ChallengePolicy):
    """The verifier chain: bind the output, then every layer from the floor
    out. Returns (final carry, transcript, ok)."""
    carry, transcript = bind_output(output, transcript, challenges)
    return verify_rounds([GkrLayerRound(challenges) for _ in proofs], carry, proofs, transcript)


__all__ = ["GkrLayerRound", "verify_gkr", "eval_eq"]
