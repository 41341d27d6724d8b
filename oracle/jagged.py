"""Oracle jagged LogUp-GKR (TEST INFRASTRUCTURE ONLY).

Plain-Python restatement, on canonical field integers, of the reference's
jagged LogUp-GKR: a layer materialises only ``sum(row_counts)`` of its
virtual 2^(niv+nrv) positions, batch-major (segment s holds row_counts[s]
rows); every other position holds the fold-neutral fraction (n = 0, d = 1).
Only ``tests/`` and the bench's CPU baseline use it, as the checker.

* ``prepad_folded``  the even-prepad / folded-count recurrence
  (zorch/sumcheck/jagged/layout.py:41-53).
* ``jagged_layer_transition``  one row variable folded per segment into the
  consumer's out counts (zorch/logup_gkr/circuit.py:296-470): output row j of
  segment s folds the segment's rows 2j, 2j+1 (the odd one neutral past the
  count) while j < ceil(rc/2); later rows of the out count are neutral.
* ``jagged_fold_schedules`` / ``build_jagged_pyramid`` /
  ``extract_jagged_outputs``  the test kit's halving policy (over-pad to
  even unless saturated at one row; zorch/logup_gkr/testing/__init__.py:
  124-142) and the floor's interleaved outputs (circuit.py:503-523).
* ``virtual_planes``  each segment extended to 2^nrv rows with the neutral
  fraction (testing/__init__.py:92-121): the brute-force dense view.
* ``prove_jagged_layer``  the layer's sumcheck (zorch/logup_gkr/
  jagged_prover.py:272-448 with _jagged_rounds.py and
  zorch/sumcheck/jagged/rounds.py): variables bind LSB first; row rounds
  re-pad odd segments to even and pair (2j, 2j+1) inside a segment with eq
  weights eq_row[2j], eq_row[2j+1] times eq_int[s]; after the row variables
  the state is one element per segment and the batch variables fold
  densely under eq_int. A round sends the ascending coefficients of the
  cubic through (0, s(0)), (1, claim - s(0)), (1/2, s(1/2)), (b, 0) with
  b = (1 - z)/(1 - 2z) the root of the current variable's eq factor, from
  the materialised sums at 0 and (doubled) 1/2 plus the virtual mass in
  closed form: pad_adj - eq_sum, at u = 0 times (1 - z), at the doubled
  u = 1/2 scale 4 per unit of mass (rounds.py:36-85). Then claim = the
  polynomial at r and pad_adj *= eq(r, z) (zorch/sumcheck/gruen.py:145-157);
  after the row phase eq_adj takes pad_adj and pad_adj restarts at 1.
* ``JaggedLayerRound`` / ``verify_jagged_layer``  the layer carry (lam, the
  batched claim, the openings observed for the child selector,
  jagged_prover.py:453-548) and the verifier (zorch/logup_gkr/
  jagged_verifier.py:40-89: the coefficient-form replay, the bound point
  reversed, the combine at eq(eval_point, point)).
"""

from __future__ import annotations

from dataclasses import dataclass

from oracle.logup import (DEGREE, _observe_and_sample_many, _sample, bind_output, eval_eq,
                          fold_carry, logup_combine)
from oracle.sumcheck import (eq_factor, eq_table, eval_coeffs, eval_mle, fs_hop, interp,
                             log2_strict)


@dataclass
class JaggedLayer:
    """Four flat MLEs over the live prefix, batch-major, and the row counts."""

    n0: list
    n1: list
    d0: list
    d1: list
    row_counts: tuple

    @property
    def num_batch_variables(self) -> int:
        return log2_strict(len(self.row_counts))


def prepad_folded(counts):
    prepad = tuple(rc + rc % 2 for rc in counts)
    return prepad, tuple(pc // 2 for pc in prepad)


def _segments(layer: JaggedLayer):
    out, pos = [], 0
    for rc in layer.row_counts:
        out.append(tuple(p[pos:pos + rc] for p in (layer.n0, layer.n1, layer.d0, layer.d1)))
        pos += rc
    return out


def jagged_layer_transition(F, layer: JaggedLayer, out_row_counts) -> JaggedLayer:
    out_row_counts = tuple(int(c) for c in out_row_counts)
    if len(out_row_counts) != len(layer.row_counts):
        raise ValueError("schedule must cover every batch")
    zero, one = F.zero(), F.one()
    n0s, n1s, d0s, d1s = [], [], [], []
    for (a0, a1, b0, b1), rc, orc in zip(_segments(layer), layer.row_counts, out_row_counts):
        folded = (rc + 1) // 2
        if orc < folded:
            raise ValueError("out count truncates a segment")
        for j in range(orc):
            if j < folded:
                e = 2 * j
                ne0, ne1, de0, de1 = a0[e], a1[e], b0[e], b1[e]
                if e + 1 < rc:
                    no0, no1, do0, do1 = a0[e + 1], a1[e + 1], b0[e + 1], b1[e + 1]
                else:
                    no0, no1, do0, do1 = zero, zero, one, one
            else:
                ne0 = ne1 = no0 = no1 = zero
                de0 = de1 = do0 = do1 = one
            n0s.append(F.add(F.mul(ne0, de1), F.mul(ne1, de0)))
            n1s.append(F.add(F.mul(no0, do1), F.mul(no1, do0)))
            d0s.append(F.mul(de0, de1))
            d1s.append(F.mul(do0, do1))
    return JaggedLayer(n0s, n1s, d0s, d1s, out_row_counts)


def jagged_fold_schedules(row_counts):
    out, counts = [], tuple(row_counts)
    while max(counts) > 1:
        folded = tuple((rc + 1) // 2 for rc in counts)
        counts = tuple(fc if fc == 1 else fc + fc % 2 for fc in folded)
        out.append(counts)
    return out


def build_jagged_pyramid(F, first: JaggedLayer, schedules=None):
    if schedules is None:
        schedules = jagged_fold_schedules(first.row_counts)
    layers = [first]
    for sched in schedules:
        layers.append(jagged_layer_transition(F, layers[-1], sched))
    return layers


def extract_jagged_outputs(F, floor: JaggedLayer):
    if len(floor.n0) != len(floor.row_counts):
        raise ValueError("extract_jagged_outputs expects the batch floor")
    num, den = [], []
    for a, b, c, d in zip(floor.n0, floor.n1, floor.d0, floor.d1):
        num += [a, b]
        den += [c, d]
    return num, den


def virtual_planes(F, layer: JaggedLayer, nrv: int):
    rows = 1 << nrv
    planes = [[], [], [], []]
    for seg, rc in zip(_segments(layer), layer.row_counts):
        for q, (vals, neutral) in enumerate(zip(seg, (F.zero(), F.zero(), F.one(), F.one()))):
            planes[q] += list(vals) + [neutral] * (rows - rc)
    return planes


def _fold_lsb(F, tab, r):
    return [F.add(tab[2 * j], F.mul(r, F.sub(tab[2 * j + 1], tab[2 * j])))
            for j in range(len(tab) // 2)]


def round_coeffs(F, eval_zero, eval_half, eq_sum, eq_adj, pad_adj, z, claim):
    """The round polynomial's ascending coefficients (rounds.py:36-85)."""
    one = F.one()
    two = F.from_int(2)
    correction = F.sub(pad_adj, eq_sum)
    s_zero = F.mul(F.add(eval_zero, F.mul(correction, F.sub(one, z))), eq_adj)
    s_half = F.mul(F.mul(F.add(eval_half, F.mul(correction, F.from_int(4))),
                         F.inv(F.from_int(8))), eq_adj)
    s_one = F.sub(claim, s_zero)
    b_root = F.mul(F.sub(one, z), F.inv(F.sub(one, F.mul(two, z))))
    xs = [F.zero(), one, F.inv(two), b_root]
    return interp(F, xs, [s_zero, s_one, s_half, F.zero()])


def _paired_sums(F, lam, pairs):
    """(s(0), 8 s(1/2), eq mass) over [(eq0, eq1, even, odd)], each of even /
    odd the (n0, n1, d0, d1) of a pair's elements (_jagged_rounds.py:22-60)."""
    ez, eh, es = F.zero(), F.zero(), F.zero()
    for eq0, eq1, (n0e, n1e, d0e, d1e), (n0o, n1o, d0o, d1o) in pairs:
        ez = F.add(ez, logup_combine(F, lam, eq0, n0e, d1e, n1e, d0e))
        eqh = F.add(eq0, eq1)
        eh = F.add(eh, logup_combine(F, lam, eqh, F.add(n0e, n0o), F.add(d1e, d1o),
                                     F.add(n1e, n1o), F.add(d0e, d0o)))
        es = F.add(es, eqh)
    return ez, eh, es


@dataclass
class JaggedLayerProof:
    lam: object
    claim: object       # the batched claim the layer entered with
    round_polys: list   # rounds x 4 ascending coefficients
    point: list         # MSB-first: the sampled challenges reversed
    n0: object
    n1: object
    d0: object
    d1: object
    point_words: list = None  # raw squeezed words, sampling order


def prove_jagged_layer(F, T, layer: JaggedLayer, lam, claim, eval_point, transcript):
    """Returns (bound point, transcript, proof)."""
    niv = layer.num_batch_variables
    nrv = len(eval_point) - niv
    if nrv < 1:
        raise ValueError("eval_point must carry at least one row variable")
    if max(layer.row_counts) > (1 << nrv) or min(layer.row_counts) < 1:
        raise ValueError("row counts must lie in [1, 2^nrv]")
    zero, one = F.zero(), F.one()
    neutral = (zero, zero, one, one)
    entry_claim = claim
    eq_row = eq_table(F, eval_point[niv:])
    eq_int = eq_table(F, eval_point[:niv])
    segs = [[tuple(p[i] for p in seg) for i in range(rc)]
            for seg, rc in zip(_segments(layer), layer.row_counts)]
    dense = None
    eq_adj, pad_adj = one, one
    polys, chals, words = [], [], []
    r_prev = None
    for rnd in range(nrv + niv):
        z = eval_point[-(rnd + 1)]
        if rnd < nrv:
            if rnd > 0:
                # bind the previous round's challenge (pairs of the padded layout)
                segs = [[tuple(F.add(a, F.mul(r_prev, F.sub(b, a))) for a, b in
                               zip(seg[2 * j], seg[2 * j + 1]))
                         for j in range(len(seg) // 2)] for seg in segs]
                eq_row = _fold_lsb(F, eq_row, r_prev)
            # re-pad odd segments to even with the neutral fraction
            segs = [seg + [neutral] * (len(seg) % 2) for seg in segs]
            pairs = []
            for s, seg in enumerate(segs):
                for j in range(len(seg) // 2):
                    pairs.append((F.mul(eq_row[2 * j], eq_int[s]),
                                  F.mul(eq_row[2 * j + 1], eq_int[s]), seg[2 * j], seg[2 * j + 1]))
        else:
            if rnd == nrv:
                dense = [tuple(F.add(a, F.mul(r_prev, F.sub(b, a))) for a, b in
                               zip(seg[0], seg[1])) for seg in segs]
            else:
                dense = _fold_lsb_tuples(F, dense, r_prev)
                eq_int = _fold_lsb(F, eq_int, r_prev)
            pairs = [(eq_int[2 * b], eq_int[2 * b + 1], dense[2 * b], dense[2 * b + 1])
                     for b in range(len(dense) // 2)]
        ez, eh, es = _paired_sums(F, lam, pairs)
        poly = round_coeffs(F, ez, eh, es, eq_adj, pad_adj, z, claim)
        transcript, rw, r = fs_hop(transcript, F, T, poly)
        claim = eval_coeffs(F, poly, r)
        pad_adj = F.mul(pad_adj, eq_factor(F, r, z))
        polys.append(poly)
        chals.append(r)
        words.append(rw)
        if rnd == nrv - 1:
            eq_adj, pad_adj = pad_adj, one
        r_prev = r
    if niv == 0:
        last = [tuple(F.add(a, F.mul(r_prev, F.sub(b, a))) for a, b in zip(seg[0], seg[1]))
                for seg in segs]
    else:
        last = _fold_lsb_tuples(F, dense, r_prev)
    fn0, fn1, fd0, fd1 = last[0]
    point = chals[::-1]
    return point, transcript, JaggedLayerProof(lam, entry_claim, polys, point, fn0, fn1, fd0,
                                               fd1, words)


def _fold_lsb_tuples(F, dense, r):
    return [tuple(F.add(a, F.mul(r, F.sub(b, a))) for a, b in zip(dense[2 * j], dense[2 * j + 1]))
            for j in range(len(dense) // 2)]


def prove_jagged_layer_round(F, T, layer: JaggedLayer, carry, transcript):
    """JaggedGkrLayerRound (prover): returns (carry, transcript, proof)."""
    num_eval, den_eval, eval_point = carry
    transcript, (lam,) = _sample(F, T, transcript, 1)
    claim = F.add(F.mul(lam, num_eval), den_eval)
    point, transcript, proof = prove_jagged_layer(F, T, layer, lam, claim, eval_point,
                                                  transcript)
    transcript, (r,) = _observe_and_sample_many(F, T, transcript,
                                                [proof.n0, proof.n1, proof.d0, proof.d1], 1)
    return fold_carry(F, proof.n0, proof.n1, proof.d0, proof.d1, point, r), transcript, proof


def verify_coeffs_sumcheck(F, T, claim, polys, transcript, degree: int = DEGREE):
    """verify over CoeffsSumcheckRound(degree): s(0) + s(1) == claim, then
    claim = s(r) (zorch/sumcheck/verifier.py:76-103)."""
    ok, point = True, []
    for poly in polys:
        if len(poly) != degree + 1:
            raise ValueError(f"round polynomial must have {degree + 1} coefficients")
        ok = ok and F.add(eval_coeffs(F, poly, F.zero()), eval_coeffs(F, poly, F.one())) == claim
        transcript, _, r = fs_hop(transcript, F, T, poly)
        claim = eval_coeffs(F, poly, r)
        point.append(r)
    return point, claim, transcript, ok


def verify_jagged_layer(F, T, proof: JaggedLayerProof, carry, transcript):
    num_eval, den_eval, eval_point = carry
    transcript, (lam,) = _sample(F, T, transcript, 1)
    claim = F.add(F.mul(lam, num_eval), den_eval)
    point, final_claim, transcript, ok = verify_coeffs_sumcheck(F, T, claim, proof.round_polys,
                                                                transcript)
    point = point[::-1]
    if len(point) != len(eval_point):
        raise ValueError("eq point mismatch")
    combined = logup_combine(F, lam, eval_eq(F, eval_point, point), proof.n0, proof.d1,
                             proof.n1, proof.d0)
    ok = ok and combined == final_claim
    transcript, (r,) = _observe_and_sample_many(F, T, transcript,
                                                [proof.n0, proof.n1, proof.d0, proof.d1], 1)
    return fold_carry(F, proof.n0, proof.n1, proof.d0, proof.d1, point, r), transcript, ok


def prove_jagged_gkr(F, T, first: JaggedLayer, transcript, schedules=None):
    """The chain: pyramid, output, bind, layers from the floor out. Returns
    (layers, (num, den), proofs, final carry, transcript)."""
    layers = build_jagged_pyramid(F, first, schedules)
    num, den = extract_jagged_outputs(F, layers[-1])
    carry, transcript = bind_output(F, T, num, den, transcript)
    proofs = []
    for layer in reversed(layers[:-1]):
        carry, transcript, proof = prove_jagged_layer_round(F, T, layer, carry, transcript)
        proofs.append(proof)
    return layers, (num, den), proofs, carry, transcript


def verify_jagged_gkr(F, T, num, den, proofs, transcript):
    carry, transcript = bind_output(F, T, num, den, transcript)
    ok = True
    for proof in proofs:
        carry, transcript, ok_i = verify_jagged_layer(F, T, proof, carry, transcript)
        ok = ok and ok_i
    return carry, transcript, ok


def interleave(a, b):
    out = []
    for x, y in zip(a, b):
        out += [x, y]
    return out


__all__ = ["JaggedLayer", "prepad_folded", "jagged_layer_transition", "jagged_fold_schedules",
           "build_jagged_pyramid", "extract_jagged_outputs", "virtual_planes", "round_coeffs",
           "JaggedLayerProof", "prove_jagged_layer", "prove_jagged_layer_round",
           "verify_coeffs_sumcheck", "verify_jagged_layer", "prove_jagged_gkr",
           "verify_jagged_gkr", "interleave", "eval_mle", "interp"]
