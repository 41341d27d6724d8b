"""LogUp-GKR fractional-sum circuit (drop-in for reference
zorch/logup_gkr/circuit.py: GkrLayer :44-92, _fold_pairs :100-117,
layer_transition :120-140, build_pyramid :143-154, extract_outputs :163-181;
the jagged layout: JaggedGkrLayer :192-257, jagged_layer_transition
:409-470, build_jagged_pyramid :472-500, extract_jagged_outputs :503-523).

A layer is four equal-length MLEs (n0, n1, d0, d1): index i holds the two
children n0[i]/d0[i] and n1[i]/d1[i]. A transition adds each stride-2 pair,
n0/d0 + n1/d1 = (n0 d1 + n1 d0) / (d0 d1), halving every MLE, until only the
batch variables remain. Host field arrays: the circuit is built once per
proof; the layers' sumchecks run on the device (prover.py).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from zorch_amd.fields import FieldArray, concatenate, ones, stack, zeros
from zorch_amd.runtime import log2_strict


@dataclass(frozen=True)
class GkrLayer:
    numerator_0: FieldArray
    numerator_1: FieldArray
    denominator_0: FieldArray
    denominator_1: FieldArray
    num_batch_variables: int

    def __post_init__(self) -> None:
        shape = self.numerator_0.shape
        for name in ("numerator_1", "denominator_0", "denominator_1"):
            if getattr(self, name).shape != shape:
                raise ValueError(f"all MLEs must share a shape; {name} is "
                                 f"{getattr(self, name).shape}, numerator_0 is {shape}")
        if not 0 <= self.num_batch_variables <= self.num_variables:
            raise ValueError(f"num_batch_variables must be in [0, {self.num_variables}], "
                             f"got {self.num_batch_variables}")

    @property
    def num_variables(self) -> int:
        return log2_strict(self.numerator_0.shape[0])

    @property
    def num_row_variables(self) -> int:
        return self.num_variables - self.num_batch_variables


@dataclass(frozen=True)
class LogUpGkrOutput:
    numerator: FieldArray
    denominator: FieldArray


def _common(*arrays: FieldArray):
    """The arrays in one field: base-field ones lifted to the extension when
    any is an extension (the reference promotes a first layer of base-field
    numerators under extension denominators, circuit.py:360-372)."""
    f = max((a.field for a in arrays), key=lambda x: x.degree)
    return tuple(a if a.field is f else a.astype(f) for a in arrays)


def _fold_pairs(n0, n1, d0, d1):
    """Even nodes become the next layer's 0-child, odd nodes its 1-child."""
    n0, n1, d0, d1 = _common(n0, n1, d0, d1)
    n0e, n0o = n0[0::2], n0[1::2]
    n1e, n1o = n1[0::2], n1[1::2]
    d0e, d0o = d0[0::2], d0[1::2]
    d1e, d1o = d1[0::2], d1[1::2]
    return (n0e * d1e + n1e * d0e, n0o * d1o + n1o * d0o, d0e * d1e, d0o * d1o)


def layer_transition(layer: GkrLayer) -> GkrLayer:
    if layer.num_row_variables < 1:
        raise ValueError("no row variable left to fold")
    rn0, rn1, rd0, rd1 = _fold_pairs(layer.numerator_0, layer.numerator_1,
                                     layer.denominator_0, layer.denominator_1)
    return GkrLayer(rn0, rn1, rd0, rd1, layer.num_batch_variables)


def build_pyramid(first: GkrLayer) -> list[GkrLayer]:
    layers = [first]
    while layers[-1].num_row_variables > 0:
        layers.append(layer_transition(layers[-1]))
    return layers


def _interleave(child_0: FieldArray, child_1: FieldArray) -> FieldArray:
    """[c0[0], c1[0], c0[1], c1[1], ...]."""
    child_0, child_1 = _common(child_0, child_1)
    both = stack([child_0, child_1], axis=1)
    return both.reshape(2 * child_0.shape[0])


def extract_outputs(layer: GkrLayer) -> LogUpGkrOutput:
    if layer.num_row_variables != 0:
        raise ValueError("extract_outputs expects the batch floor (num_row_variables == 0), "
                         f"got {layer.num_row_variables}")
    return LogUpGkrOutput(_interleave(layer.numerator_0, layer.numerator_1),
                          _interleave(layer.denominator_0, layer.denominator_1))


@dataclass(frozen=True)
class JaggedGkrLayer:
    """One jagged layer, batch-major: the four MLEs hold row_counts[0] rows of
    batch element 0, then batch element 1's, ... over a capacity width >=
    sum(row_counts); slots past the live rows are dead and zero. The number
    of batch elements is a power of two."""

    numerator_0: FieldArray
    numerator_1: FieldArray
    denominator_0: FieldArray
    denominator_1: FieldArray
    row_counts: tuple

    def __post_init__(self) -> None:
        counts = tuple(int(c) for c in np.asarray(self.row_counts).reshape(-1))
        if np.asarray(self.row_counts).ndim != 1:
            raise ValueError(f"row_counts must be a flat vector, got "
                             f"{np.asarray(self.row_counts).shape}")
        object.__setattr__(self, "row_counts", counts)
        log2_strict(len(counts))
        width = self.width
        for name in ("numerator_0", "numerator_1", "denominator_0", "denominator_1"):
            shape = getattr(self, name).shape
            if shape != (width,):
                raise ValueError(f"the four MLEs must share one capacity width; {name} is "
                                 f"{shape}, expected ({width},)")

    @property
    def num_batches(self) -> int:
        return len(self.row_counts)

    @property
    def num_batch_variables(self) -> int:
        return log2_strict(self.num_batches)

    @property
    def width(self) -> int:
        return self.numerator_0.shape[0]

    @property
    def live(self) -> int:
        return sum(self.row_counts)


def _transition_gathers(row_counts, out_row_counts, out_width):
    """Per output slot, the (even, odd) source rows of its fold: row 2j and
    2j + 1 of its segment while j < ceil(rc / 2) (the odd one past the count
    and every later slot of the out count resolve to the neutral fraction,
    index `width`); slots past the live out rows resolve to zero (index
    `width + 1`)."""
    src_base = np.concatenate([[0], np.cumsum(row_counts)[:-1]]).astype(np.int64)
    neutral = int(sum(row_counts))
    even = np.full(out_width, neutral + 1, dtype=np.int64)
    odd = np.full(out_width, neutral + 1, dtype=np.int64)
    pos = 0
    for s, (rc, orc) in enumerate(zip(row_counts, out_row_counts)):
        j = np.arange(orc)
        folds = j < (rc + 1) // 2
        even[pos:pos + orc] = np.where(folds, src_base[s] + 2 * j, neutral)
        odd[pos:pos + orc] = np.where(folds & (2 * j + 1 < rc), src_base[s] + 2 * j + 1,
                                      neutral)
        pos += orc
    return even, odd


def jagged_layer_transition(layer: JaggedGkrLayer, out_row_counts, out_width=None) -> JaggedGkrLayer:
    """Fold one row variable per segment into `out_row_counts` (the consumer's
    halving policy; each must be at least ceil(rc / 2)): odd segments pre-pad
    with the neutral fraction (n = 0, d = 1) so a pair never crosses a
    segment, out rows past a segment's folded count are neutral, and slots
    past sum(out_row_counts) are dead zeros."""
    out = tuple(int(c) for c in out_row_counts)
    if len(out) != layer.num_batches:
        raise ValueError(f"schedule must cover all {layer.num_batches} batches, got "
                         f"{len(out)} entries")
    if out_width is None:
        out_width = sum(out)
    elif out_width < sum(out):
        raise ValueError(f"out_width {out_width} cannot hold the schedule's live size "
                         f"({sum(out)}); widen the capacity for this transition")
    for rc, orc in zip(layer.row_counts, out):
        if orc < (rc + 1) // 2:
            raise ValueError(f"out row count {orc} truncates a segment of {rc} rows "
                             f"(needs at least {(rc + 1) // 2})")
    even, odd = _transition_gathers(layer.row_counts, out, out_width)
    live = layer.live
    n0, n1, d0, d1 = _common(layer.numerator_0, layer.numerator_1, layer.denominator_0,
                             layer.denominator_1)
    f = n0.field

    def gather(a: FieldArray, neutral: FieldArray):
        ext = concatenate([a[:live], neutral, zeros((1,), f)])
        return ext[even], ext[odd]

    z1, o1 = zeros((1,), f), ones((1,), f)
    n0e, n0o = gather(n0, z1)
    n1e, n1o = gather(n1, z1)
    d0e, d0o = gather(d0, o1)
    d1e, d1o = gather(d1, o1)
    return JaggedGkrLayer(n0e * d1e + n1e * d0e, n0o * d1o + n1o * d0o, d0e * d1e, d0o * d1o,
                          out)


def build_jagged_pyramid(first: JaggedGkrLayer, schedules) -> list[JaggedGkrLayer]:
    """[first, ..., floor]: one transition per schedule entry (a sequence of
    out counts, or (out counts, out width))."""
    layers = [first]
    for sched in schedules:
        if isinstance(sched, tuple) and len(sched) == 2 and not np.isscalar(sched[0]):
            layers.append(jagged_layer_transition(layers[-1], sched[0], sched[1]))
        else:
            layers.append(jagged_layer_transition(layers[-1], sched))
    return layers


def extract_jagged_outputs(layer: JaggedGkrLayer) -> LogUpGkrOutput:
    """The floor (one live row per batch element, width == num_batches)
    interleaved into the output MLEs."""
    if layer.width != layer.num_batches:
        raise ValueError(f"extract_jagged_outputs expects the batch floor (width == "
                         f"num_batches == {layer.num_batches}), got width {layer.width}")
    return LogUpGkrOutput(_interleave(layer.numerator_0, layer.numerator_1),
                          _interleave(layer.denominator_0, layer.denominator_1))
