"""Oracle dense LogUp-GKR (TEST INFRASTRUCTURE ONLY).

Plain-Python restatement, on canonical field integers (``oracle.fields``) and
the hashlib transcript (``oracle.transcript``), of the reference's dense
LogUp-GKR block; only ``tests/`` and the bench's CPU baseline use it, as the
checker. Small sizes only.

* ``logup_combine``  the summand eq (lam (n0 d1 + n1 d0) + d0 d1)
  (zorch/logup_gkr/prover.py:56-68).
* ``fold_pairs`` / ``layer_transition`` / ``build_pyramid`` /
  ``extract_outputs``  the fractional-sum circuit (zorch/logup_gkr/circuit.py:
  _fold_pairs, layer_transition, build_pyramid, _interleave, extract_outputs):
  stride-2 pairs n0/d0 + n1/d1 = (n0 d1 + n1 d0) / (d0 d1).
* ``logup_round_poly``  LogupSumcheckRound._round_poly (prover.py:282-290):
  the round polynomial at t = 0..3 of the combine over the stacked state
  [eq, n0, d1, n1, d0] (summand_evals on the natural domain, MSB halves).
* ``prove_layer``  GkrLayerRound.__call__ (prover.py:355-378): lam, the eq
  table of the carry's point, fold_rounds over LogupSumcheckRound, the final
  openings observed for the child selector r, fold_carry (prover.py:159-172).
* ``bind_output``  (prover.py:307-322, _bind_output_body :325-341): one
  observe_and_sample_many of the concatenated output MLEs.
* ``verify_layer``  the verifier's GkrLayerRound (zorch/logup_gkr/verifier.py:
  37-72): lam, claim = lam num + den, the sumcheck replay at degree 3, the
  oracle check combine(lam, eq(point, bound point), openings) == final claim.
* ``prove_gkr`` / ``verify_gkr``  the chains of the reference's test kit
  (zorch/logup_gkr/testing/__init__.py: prove_gkr_with_transcript).
"""

from __future__ import annotations

from dataclasses import dataclass

from oracle.fields import limbs_over
from oracle.sumcheck import (_fold, eq_factor, eq_table, eval_mle, fs_hop, lagrange_eval,
                             log2_strict)
from oracle.sumcheck import verify as sumcheck_verify
from oracle.transcript import OracleTranscript

DEGREE = 3  # eq (degree 1) times the degree-2 bracket


def logup_combine(F, lam, eq, n0, d1, n1, d0):
    return F.mul(eq, F.add(F.mul(lam, F.add(F.mul(n0, d1), F.mul(n1, d0))), F.mul(d0, d1)))


@dataclass
class Layer:
    """A dense layer: four equal-length MLEs and the batch-variable floor."""

    n0: list
    n1: list
    d0: list
    d1: list
    num_batch_variables: int

    @property
    def num_variables(self) -> int:
        return log2_strict(len(self.n0))

    @property
    def num_row_variables(self) -> int:
        return self.num_variables - self.num_batch_variables


def fold_pairs(F, n0, n1, d0, d1):
    """Even nodes become the next layer's 0-child, odd nodes its 1-child."""
    e, o = slice(0, None, 2), slice(1, None, 2)

    def frac_num(a0, a1, b0, b1):  # numerator of a0/b0 + a1/b1 ... per index
        return [F.add(F.mul(x, y), F.mul(z, w)) for x, y, z, w in zip(a0, b1, a1, b0)]

    return (frac_num(n0[e], n1[e], d0[e], d1[e]), frac_num(n0[o], n1[o], d0[o], d1[o]),
            [F.mul(x, y) for x, y in zip(d0[e], d1[e])], [F.mul(x, y) for x, y in zip(d0[o], d1[o])])


def layer_transition(F, layer: Layer) -> Layer:
    if layer.num_row_variables < 1:
        raise ValueError("no row variable left to fold")
    rn0, rn1, rd0, rd1 = fold_pairs(F, layer.n0, layer.n1, layer.d0, layer.d1)
    return Layer(rn0, rn1, rd0, rd1, layer.num_batch_variables)


def build_pyramid(F, first: Layer) -> list:
    layers = [first]
    while layers[-1].num_row_variables > 0:
        layers.append(layer_transition(F, layers[-1]))
    return layers


def extract_outputs(F, floor: Layer):
    """The floor's two children interleaved into the output MLEs."""
    if floor.num_row_variables != 0:
        raise ValueError("extract_outputs expects the batch floor")
    num, den = [], []
    for a, b, c, d in zip(floor.n0, floor.n1, floor.d0, floor.d1):
        num += [a, b]
        den += [c, d]
    return num, den


def logup_round_poly(F, lam, state):
    """s(t), t = 0..3, of the combine over the stacked [eq, n0, d1, n1, d0]."""
    h = len(state[0]) // 2
    msg = []
    for t in range(DEGREE + 1):
        tt = F.from_int(t)
        acc = F.zero()
        for j in range(h):
            v = [F.add(tab[j], F.mul(tt, F.sub(tab[j + h], tab[j]))) for tab in state]
            acc = F.add(acc, logup_combine(F, lam, *v))
        msg.append(acc)
    return msg


def _sample(F, T, transcript: OracleTranscript, count: int):
    """ChallengePolicy.sample_many (zorch/challenge.py:57-78): count * limbs
    raw words, regrouped into elements of F."""
    limbs = limbs_over(F, T)
    transcript, raw = transcript.sample(count * limbs)
    words = F.words_from_bytes(raw) if F.degree > 1 else T.words_from_bytes(raw)
    per = len(words) // count
    return transcript, [F.from_words(words[i * per:(i + 1) * per]) for i in range(count)]


def _observe_and_sample_many(F, T, transcript: OracleTranscript, values, count: int):
    payload = b"".join(F.to_bytes(v) for v in values)
    return _sample(F, T, transcript.observe(payload), count)


def fold_carry(F, n0, n1, d0, d1, point, r):
    num = F.add(n0, F.mul(F.sub(n1, n0), r))
    den = F.add(d0, F.mul(F.sub(d1, d0), r))
    return num, den, list(point) + [r]


@dataclass
class LayerProof:
    round_polys: list
    point: list
    n0: object
    n1: object
    d0: object
    d1: object


def bind_output(F, T, num, den, transcript: OracleTranscript):
    nv = log2_strict(len(num))
    transcript, point = _observe_and_sample_many(F, T, transcript, list(num) + list(den), nv)
    return (eval_mle(F, num, point), eval_mle(F, den, point), point), transcript


def prove_layer(F, T, layer: Layer, carry, transcript: OracleTranscript):
    """One GKR layer (prover side). Returns (carry, transcript, proof, lam)."""
    num_eval, den_eval, eval_point = carry
    transcript, (lam,) = _sample(F, T, transcript, 1)
    state = [eq_table(F, eval_point), [F.embed(x) for x in layer.n0],
             [F.embed(x) for x in layer.d1], [F.embed(x) for x in layer.n1],
             [F.embed(x) for x in layer.d0]]
    rounds = log2_strict(len(state[0]))
    polys, point = [], []
    for _ in range(rounds):
        msg = logup_round_poly(F, lam, state)
        transcript, _, r = fs_hop(transcript, F, T, msg)
        state = [_fold(F, tab, r) for tab in state]
        polys.append(msg)
        point.append(r)
    _, n0, d1, n1, d0 = (tab[0] for tab in state)
    transcript, (r,) = _observe_and_sample_many(F, T, transcript, [n0, n1, d0, d1], 1)
    carry = fold_carry(F, n0, n1, d0, d1, point, r)
    return carry, transcript, LayerProof(polys, point, n0, n1, d0, d1), lam


def eval_eq(F, a, b):
    """eq(a, b) = prod_i (a_i b_i + (1 - a_i)(1 - b_i)) (zorch/poly/eq.py:16-40)."""
    acc = F.one()
    for x, y in zip(a, b):
        acc = F.mul(acc, eq_factor(F, x, y))
    return acc


def verify_layer(F, T, proof: LayerProof, carry, transcript: OracleTranscript):
    """One GKR layer (verifier side). Returns (carry, transcript, ok)."""
    num_eval, den_eval, eval_point = carry
    transcript, (lam,) = _sample(F, T, transcript, 1)
    claim = F.add(F.mul(lam, num_eval), den_eval)
    point, final_claim, transcript, ok = sumcheck_verify(F, T, claim, proof.round_polys,
                                                         transcript, DEGREE)
    if len(eval_point) != len(point):
        raise ValueError("eq point mismatch")
    combined = logup_combine(F, lam, eval_eq(F, eval_point, point), proof.n0, proof.d1,
                             proof.n1, proof.d0)
    ok = ok and combined == final_claim
    transcript, (r,) = _observe_and_sample_many(F, T, transcript,
                                                [proof.n0, proof.n1, proof.d0, proof.d1], 1)
    return fold_carry(F, proof.n0, proof.n1, proof.d0, proof.d1, point, r), transcript, ok


def prove_gkr(F, T, first: Layer, transcript: OracleTranscript):
    """The prover chain: pyramid, output, bind, layers from the floor out.
    Returns (layers, (num, den), proofs, final carry, transcript)."""
    layers = build_pyramid(F, first)
    num, den = extract_outputs(F, layers[-1])
    carry, transcript = bind_output(F, T, num, den, transcript)
    proofs = []
    for layer in reversed(layers[:-1]):
        carry, transcript, proof, _ = prove_layer(F, T, layer, carry, transcript)
        proofs.append(proof)
    return layers, (num, den), proofs, carry, transcript


def verify_gkr(F, T, num, den, proofs, transcript: OracleTranscript):
    carry, transcript = bind_output(F, T, num, den, transcript)
    ok = True
    for proof in proofs:
        carry, transcript, ok_i = verify_layer(F, T, proof, carry, transcript)
        ok = ok and ok_i
    return carry, transcript, ok


__all__ = ["logup_combine", "Layer", "fold_pairs", "layer_transition", "build_pyramid",
           "extract_outputs", "logup_round_poly", "bind_output", "prove_layer", "verify_layer",
           "eval_eq", "prove_gkr", "verify_gkr", "fold_carry", "lagrange_eval", "DEGREE"]
