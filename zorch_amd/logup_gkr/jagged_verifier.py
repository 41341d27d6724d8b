"""Jagged LogUp-GKR verifier (drop-in for reference zorch/logup_gkr/jagged_verifier.py).

The prover's closed-form virtual mass makes its round polynomials those of
the virtual dense hypercube, so the verifier never sees the jagged layout:
it replays the coefficient-form sumcheck (``verify`` over
``CoeffsSumcheckRound(3)``), reverses the sampled challenges into the
most-significant-first bound point (the prover binds LSB first), checks
combine(lam, eq(eval_point, point), openings) against the final claim, and
reduces the carry across the child selector (jagged_verifier.py:40-89). The
final check against the input layer's MLE is the consumer's.
"""

from __future__ import annotations

from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import stack
from zorch_amd.logup_gkr.jagged_prover import DEGREE, JaggedLayerProof
from zorch_amd.logup_gkr.prover import fold_carry, logup_combine
from zorch_amd.logup_gkr.verifier import eval_eq
from zorch_amd.sumcheck.verifier import CoeffsSumcheckRound
from zorch_amd.verify import verify


class JaggedGkrLayerRound:
    """Verify one jagged GKR layer; the chain of these is the jagged verifier."""

    def __init__(self, challenges: ChallengePolicy) -> None:
        self.challenges = challenges

    def __call__(self, claim, transcript, layer_proof: JaggedLayerProof):
        num_eval, den_eval, eval_point = claim
        n0, n1 = layer_proof.numerator_0, layer_proof.numerator_1
        d0, d1 = layer_proof.denominator_0, layer_proof.denominator_1
        transcript, lam = self.challenges.sample(transcript)
        lam = lam.normalized()
        batched = lam * num_eval + den_eval
        point, final_claim, transcript, ok = verify(CoeffsSumcheckRound(DEGREE, self.challenges),
                                                    batched, layer_proof.round_polys, transcript)
        point = point.normalized()[::-1]
        if eval_point.shape[0] != point.shape[0]:
            raise ValueError(f"eq point mismatch: claim has {eval_point.shape[0]} coords, "
                             f"layer ran {point.shape[0]} rounds")
        combined = logup_combine(lam, eval_eq(eval_point, point), n0, d1, n1, d0)
        ok = ok and bool((combined == final_claim).all())
        transcript, r = self.challenges.observe_and_sample(transcript, stack([n0, n1, d0, d1]))
        return fold_carry(n0, n1, d0, d1, point, r.normalized()), transcript, ok


__all__ = ["JaggedGkrLayerRound"]
