"""Univariate skip (drop-in for reference zorch/sumcheck/univariate_skip.py).

The first ``skip_rounds`` variables collapse into one univariate round over
the order-2^skip_rounds two-adic subgroup D: round 0 sends the ascending
coefficients of s0(Z) = sum_x prod_k f_k(Z, x) (base field, degree
degree * (|D| - 1)), the verifier checks claim == sum_{z in D} s0(z), samples
r0 in the challenge field, and the prism is bound at r0; the remaining
rounds are ordinary product rounds over the bound challenge-field state.

On the GPU (``zs_prove_univariate_skip``) round 0 is one streaming pass over
the input: each column's |D| values are mapped to their values on the
order-M subgroup that the message needs (a fixed matrix, the iNTT then the
padded NTT), multiplied across factors and summed; the host interpolates the M
sums to the coefficients and runs the hop; a second pass binds the prism at
r0 (Lagrange weights of D), and the dense prover runs the tail. skip_rounds 0
is the plain StandardRound run.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any

import numpy as np

from zorch_amd import _lib
from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import FieldArray, as_field, concatenate, stack, zeros
from zorch_amd.prove import fold_rounds
from zorch_amd.round import RunningClaim
from zorch_amd.runtime import Context, DeviceTables, default_context, log2_strict
from zorch_amd.stage import ProveResult, VerifyResult
from zorch_amd.sumcheck._device import make_job, prove_stage
from zorch_amd.sumcheck.domain import subgroup_to_coeffs
from zorch_amd.sumcheck.prover import FoldingClaim, ProductSummand, StandardRound
from zorch_amd.sumcheck.stage import SumcheckWitness, SumClaim
from zorch_amd.sumcheck.verifier import SumcheckRound, UnivariateSkipRound


def _check(p_initial, skip_rounds: int, summand) -> int:
    total = log2_strict(p_initial.shape[1])
    if not 0 <= skip_rounds <= total:
        raise ValueError(f"skip_rounds must be in [0, {total}], got {skip_rounds}")
    if type(summand) is not ProductSummand:
        raise NotImplementedError("the device univariate skip takes the product summand")
    return total


def _skip_call(p_initial, claim: FieldArray, skip_rounds: int, transcript, summand,
               challenges: ChallengePolicy, whole: bool, ctx: Context | None = None):
    """One library call: the whole skip proof (whole) or round 0 alone."""
    ctx = ctx or default_context()
    total = log2_strict(p_initial.shape[1])
    job = make_job(p_initial, claim, challenges.dtype, _lib.ZS_MODE_PRODUCT, summand.degree)
    tf = as_field(p_initial.dtype)
    d0 = summand.degree * ((1 << skip_rounds) - 1)
    msg0 = np.zeros((d0 + 1, tf.words), dtype=tf.word_dtype)
    st = transcript.state.copy()
    c = job.chal
    if whole:
        _lib.check(_lib.lib().zs_prove_univariate_skip(
            ctx.handle, ctypes.byref(job.args), skip_rounds, ctypes.byref(st),
            msg0.ctypes.data, ctypes.byref(job.proof)))
        return job, FieldArray(msg0, tf), type(transcript)(st, transcript.dtype)
    k = p_initial.shape[0]
    bound = DeviceTables.from_host(zeros((k, 1 << (total - skip_rounds)), c), ctx)
    chal = np.zeros((c.words,), dtype=c.word_dtype)
    red = np.zeros((c.words,), dtype=c.word_dtype)
    _lib.check(_lib.lib().zs_univariate_skip_round0(
        ctx.handle, ctypes.byref(job.args), skip_rounds, ctypes.byref(st), msg0.ctypes.data,
        chal.ctypes.data, red.ctypes.data, ctypes.c_void_p(bound.data_ptr())))
    return (bound, FieldArray(chal, c), FieldArray(red, c)), FieldArray(msg0, tf), \
        type(transcript)(st, transcript.dtype)


def round0_message(p_initial, skip_rounds: int, summand) -> tuple[FieldArray, FieldArray]:
    """The round-0 message s0 (ascending coefficients, base field, degree
    degree * (|D| - 1)) and the factors' D-coefficients (m, H_n, |D|)
    (reference univariate_skip.py:65-101). The message comes from the device
    round 0 (a scratch transcript absorbs its hop); the coefficients are the
    host transform of the columns, for small inputs."""
    from zorch_amd.sha256_field_transcript import Sha256FieldTranscript
    m = p_initial.shape[0]
    size = 1 << skip_rounds
    hn = p_initial.shape[1] // size
    tables = p_initial.to_host() if isinstance(p_initial, DeviceTables) else p_initial
    _, msg0, _ = _skip_call(p_initial, zeros((), tables.field), skip_rounds,
                            Sha256FieldTranscript.new(b"", tables.field), summand,
                            ChallengePolicy(tables.field), whole=False)
    cols = FieldArray(np.ascontiguousarray(
        np.swapaxes(tables.words.reshape(m, size, hn, tables.field.words), 1, 2)), tables.field)
    return msg0, subgroup_to_coeffs(cols)


def skip_round0(p_initial, claim: RunningClaim, skip_rounds: int, transcript, summand,
                challenges: ChallengePolicy):
    """Round 0 of the skip (reference univariate_skip.py:104-133): emit s0,
    observe it, sample r0, bind the prism at r0. The bound state (m, H_n) is
    device-resident challenge-field tables any product tail continues from."""
    total = _check(p_initial, skip_rounds, summand)
    if skip_rounds < 1:
        raise ValueError("skip_round0 needs skip_rounds >= 1")
    (bound, r0, reduced), msg0, transcript = _skip_call(
        p_initial, claim.value, skip_rounds, transcript, summand, challenges, whole=False)
    del total
    return FoldingClaim(bound, claim.bind(reduced, r0)), transcript, msg0


def prove_univariate_skip(p_initial, claim: FieldArray, skip_rounds: int, transcript,
                          summand: ProductSummand | None = None, *,
                          challenges: ChallengePolicy):
    """Prove with the first `skip_rounds` rounds collapsed into one univariate
    round (reference univariate_skip.py:136-190). Returns (carry, transcript,
    messages): carry.state the final folded factors (m, 1), carry.claim the
    reduced claim over [r0, r1, ..., r_n], messages [msg0] + the n tail
    messages. skip_rounds 0 is the plain StandardRound run."""
    summand = summand or ProductSummand(degree=p_initial.shape[0])
    total = _check(p_initial, skip_rounds, summand)
    n = total - skip_rounds
    if skip_rounds == 0:
        out = prove_stage(p_initial, claim, challenges.dtype, _lib.ZS_MODE_PRODUCT,
                          summand.degree, transcript)
        carry = FoldingClaim(out.final.reshape(p_initial.shape[0], 1),
                             RunningClaim(out.value, out.point, total))
        return carry, out.transcript, [out.msgs[i] for i in range(total)]
    job, msg0, transcript = _skip_call(p_initial, claim, skip_rounds, transcript, summand,
                                       challenges, whole=True)
    c = job.chal
    msgs = FieldArray(job.msgs[:n], c)
    point = FieldArray(job.point[:n + 1], c)
    final = FieldArray(job.final[:job.k], c).reshape(job.k, 1)
    carry = FoldingClaim(final, RunningClaim(FieldArray(job.value, c), point, n + 1))
    return carry, transcript, [msg0] + [msgs[i] for i in range(n)]


def verify_univariate_skip(claim: FieldArray, msgs: list, skip_rounds: int, total: int,
                           transcript, degree: int, *, challenges: ChallengePolicy):
    """Replay the skip prover (reference univariate_skip.py:193-245): the
    subgroup round then the natural-domain tail; returns (value, transcript,
    point [r0, ..., r_n], ok)."""
    if not 0 <= skip_rounds <= total:
        raise ValueError(f"skip_rounds must be in [0, {total}], got {skip_rounds}")
    cf = as_field(challenges.dtype)
    if skip_rounds == 0:
        rnd = SumcheckRound(degree=degree, challenges=challenges)
        state = RunningClaim(claim.astype(cf), zeros((len(msgs),), cf), 0)
        ok = True
        for msg in msgs:
            state, transcript, ok_r = rnd(state, transcript, msg)
            ok = ok and ok_r
        return state.value, transcript, state.point, ok
    n = total - skip_rounds
    if len(msgs) != 1 + n:
        raise ValueError(f"need {1 + n} messages, got {len(msgs)}")
    state = RunningClaim(claim, zeros((1 + n,), cf), 0)
    state, transcript, ok = UnivariateSkipRound(skip_rounds=skip_rounds, degree=degree,
                                                challenges=challenges)(state, transcript, msgs[0])
    tail = SumcheckRound(degree=degree, challenges=challenges)
    for msg in msgs[1:]:
        state, transcript, ok_r = tail(state, transcript, msg)
        ok = ok and ok_r
    return state.value, transcript, state.point, ok


@dataclass(frozen=True)
class UnivariateSkipProof:
    """The subgroup message followed by ordinary sumcheck messages."""

    head: FieldArray
    tail: tuple


@dataclass(frozen=True)
class PrismEvaluationClaim:
    """The claim reduced at one prism-prefix coordinate and the remaining MLE
    point."""

    prism_point: FieldArray
    value: FieldArray


class UnivariateSkipProver:
    """Univariate-skip prover stage (reference univariate_skip.py:263-299)."""

    def __init__(self, skip_rounds: int, summand, *, challenges: ChallengePolicy) -> None:
        if skip_rounds < 1:
            raise ValueError("univariate skip requires skip_rounds >= 1")
        self.skip_rounds = skip_rounds
        self.summand = summand
        self.challenges = challenges

    def prove(self, claim: SumClaim, witness: SumcheckWitness, transcript) -> ProveResult:
        carry, transcript, messages = prove_univariate_skip(
            witness.state, claim.value, self.skip_rounds, transcript, self.summand,
            challenges=self.challenges)
        reduced = carry.claim
        return ProveResult(PrismEvaluationClaim(reduced.point, reduced.value),
                           UnivariateSkipProof(messages[0], tuple(messages[1:])), transcript)


class UnivariateSkipVerifier:
    """Univariate-skip verifier stage (reference univariate_skip.py:302-332)."""

    def __init__(self, skip_rounds: int, summand, *, challenges: ChallengePolicy) -> None:
        if skip_rounds < 1:
            raise ValueError("univariate skip requires skip_rounds >= 1")
        self.skip_rounds = skip_rounds
        self.summand = summand
        self.challenges = challenges

    def verify(self, claim: SumClaim, reduction_proof: UnivariateSkipProof,
               transcript) -> VerifyResult:
        messages = [reduction_proof.head, *reduction_proof.tail]
        value, transcript, point, ok = verify_univariate_skip(
            claim.value, messages, self.skip_rounds, claim.rounds, transcript,
            self.summand.degree, challenges=self.challenges)
        return VerifyResult(PrismEvaluationClaim(point, value), transcript, ok)


__all__ = ["round0_message", "skip_round0", "prove_univariate_skip", "verify_univariate_skip",
           "UnivariateSkipProof", "PrismEvaluationClaim", "UnivariateSkipProver",
           "UnivariateSkipVerifier"]
