"""Seeded synthetic inputs for the benchmark and test workloads.

Generator: ``np.random.default_rng(seed).integers(0, p, 2**n, dtype=uint64)``
then the storage codec (Montgomery for BabyBear); labels are
``default_rng(label_seed).bytes(32)``; the zerocheck point z has four seeded
limbs per coordinate in BabyBear^4. Claims are computed from the tables
(sum_x prod_k T_k(x), or sum_x eq(z, x) prod_k T_k(x)) with vectorised host
arithmetic.

Workloads (sizes as named):
* ``product_bb_n16``      BabyBear, 2 tables, degree 2, n = 16, BabyBear challenges.
* ``zerocheck_bb4_n24``   eq(z,x) A(x) B(x), BabyBear tables, BabyBear^4 challenges
                          and z, n = 24, messages of degree 3.
* ``product_gl_n27``      Goldilocks, 2 tables, degree 2, n = 27.
* ``batch_512``           512 instances, n ~ U{12..22}, half product / half zerocheck.
* ``serve_poisson``       Poisson arrivals, zerocheck requests with n ~ U{14..24}.
"""

from __future__ import annotations

from dataclasses import dataclass, field as dc_field

import numpy as np

from zorch_amd.fields import (
    BB_P,
    GL_P,
    Field,
    FieldArray,
    babybear_mont,
    babybearx4_mont,
    c_add,
    c_mul,
    c_sum,
    decode,
    encode,
    goldilocks,
)


def label(seed: int = 0, nbytes: int = 32) -> bytes:
    return np.random.default_rng(seed).bytes(nbytes)


def table(seed: int, n: int, field: Field) -> FieldArray:
    """2^n uniform elements of a prime field from `seed`."""
    vals = np.random.default_rng(seed).integers(0, field.p, size=1 << n, dtype=np.uint64)
    return FieldArray(encode(field, vals)[:, None], field)


def ext_point(seed: int, n: int) -> FieldArray:
    """n seeded BabyBear^4 coordinates."""
    vals = np.random.default_rng(seed).integers(0, BB_P, size=(n, 4), dtype=np.uint64)
    return FieldArray(encode(babybearx4_mont, vals), babybearx4_mont)


def stack_tables(tables: list[FieldArray]) -> FieldArray:
    return FieldArray(np.stack([t.words for t in tables]), tables[0].field)


def product_claim(tables: FieldArray, chal: Field) -> FieldArray:
    """sum_x prod_k T_k(x), lifted into the challenge field."""
    f = tables.field
    can = decode(f, tables.words)  # (k, N, words)
    prod = can[0]
    for q in range(1, can.shape[0]):
        prod = c_mul(f, prod, can[q])
    s = c_sum(f, prod, axis=0)
    return FieldArray(encode(f, s), f).astype(chal)


def _eq_table_can(field: Field, z_can: np.ndarray) -> np.ndarray:
    """eq(z, x) over {0,1}^len(z), z[0] on the index MSB; canonical (2^m, words)."""
    state = np.zeros((1, field.words), dtype=np.uint64)
    state[0, 0] = 1
    for j in range(z_can.shape[0] - 1, -1, -1):
        hi = c_mul(field, state, z_can[j][None, :])
        lo = _sub_vec(field, state, hi)
        state = np.concatenate([lo, hi], axis=0)
    return state


def _sub_vec(field: Field, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    p = np.uint64(field.p)
    return np.where(a >= b, a - b, a + (p - b))


def eq_claim(tables: FieldArray, z: FieldArray) -> FieldArray:
    """sum_x eq(z, x) prod_k T_k(x) for base-field tables and BabyBear^4 z."""
    f = tables.field
    ef = z.field
    n = z.shape[0]
    can = decode(f, tables.words)[..., 0]  # (k, N)
    p = np.uint64(f.p)
    prod = can[0]
    for q in range(1, can.shape[0]):
        prod = (prod * can[q]) % p
    h = n // 2
    zc = decode(ef, z.words)
    eql = _eq_table_can(ef, zc[:h])        # (2^h, 4)
    eqr = _eq_table_can(ef, zc[h:])        # (2^(n-h), 4)
    mat = prod.reshape(1 << h, 1 << (n - h))
    inner = np.zeros((1 << h, 4), dtype=np.uint64)
    rows = max(1, (1 << 22) >> (n - h))
    for r0 in range(0, 1 << h, rows):
        blk = mat[r0:r0 + rows]
        for c in range(4):
            t = (blk * eqr[None, :, c]) % p
            inner[r0:r0 + rows, c] = np.sum(t, axis=1, dtype=np.uint64) % p
    tot = c_sum(ef, c_mul(ef, eql, inner), axis=0)
    return FieldArray(encode(ef, tot), ef)


@dataclass
class Instance:
    """One sumcheck instance of a workload."""

    name: str
    mode: str                 # "product" or "zerocheck"
    n: int
    tables: FieldArray        # (k, 2^n) table field
    claim: FieldArray         # challenge field
    chal: Field
    transcript_field: Field
    label: bytes
    degree: int = 2           # ProductSummand degree (= number of tables)
    z: FieldArray | None = None
    transcript_kind: str = "sha256"  # or "blake3"

    @property
    def table_elems(self) -> int:
        return self.tables.shape[0] * self.tables.shape[1]


def new_transcript(inst: Instance):
    """The instance's initial field transcript (its hash, label and word field)."""
    if inst.transcript_kind == "blake3":
        from zorch_amd.blake3_field_transcript import Blake3FieldTranscript
        return Blake3FieldTranscript.new(inst.label, inst.transcript_field)
    from zorch_amd.sha256_field_transcript import Sha256FieldTranscript
    return Sha256FieldTranscript.new(inst.label, inst.transcript_field)


def product_bb(n: int = 16, seeds=(1, 2), label_seed: int = 0) -> Instance:
    t = stack_tables([table(s, n, babybear_mont) for s in seeds])
    return Instance(f"product_bb_n{n}", "product", n, t, product_claim(t, babybear_mont),
                    babybear_mont, babybear_mont, label(label_seed), degree=len(seeds))


def zerocheck_bb4(n: int = 24, seeds=(1, 2), z_seed: int = 3, label_seed: int = 0,
                  transcript_field: Field = babybear_mont) -> Instance:
    t = stack_tables([table(s, n, babybear_mont) for s in seeds])
    z = ext_point(z_seed, n)
    return Instance(f"zerocheck_bb4_n{n}", "zerocheck", n, t, eq_claim(t, z), babybearx4_mont,
                    transcript_field, label(label_seed), degree=len(seeds), z=z)


def product_gl(n: int = 27, seeds=(1, 2), label_seed: int = 0) -> Instance:
    t = stack_tables([table(s, n, goldilocks) for s in seeds])
    return Instance(f"product_gl_n{n}", "product", n, t, product_claim(t, goldilocks),
                    goldilocks, goldilocks, label(label_seed), degree=len(seeds))


def product_bb_ext(n: int, seeds, label_seed: int) -> Instance:
    """Product of two BabyBear tables with BabyBear^4 challenges (tables lifted)."""
    t = stack_tables([table(s, n, babybear_mont) for s in seeds])
    return Instance(f"product_bb_bb4_n{n}", "product", n, t,
                    product_claim(t, babybearx4_mont), babybearx4_mont, babybear_mont,
                    label(label_seed), degree=len(seeds))


@dataclass
class BatchSpec:
    sizes: list[int]
    kinds: list[str]
    seed: int


def batch_spec(count: int = 512, seed: int = 4, lo: int = 12, hi: int = 22) -> BatchSpec:
    """Log-sizes uniform in {lo..hi}, product (degree 2) or zerocheck (degree 3) with
    probability 1/2 each, from `seed`."""
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.integers(lo, hi + 1, size=count)]
    kinds = ["product" if b else "zerocheck" for b in rng.integers(0, 2, size=count)]
    return BatchSpec(sizes, kinds, seed)


def batch_instance(spec: BatchSpec, i: int) -> Instance:
    n, kind = spec.sizes[i], spec.kinds[i]
    base = 10_000 + 4 * i
    if kind == "product":
        return product_bb_ext(n, (base, base + 1), base + 3)
    return zerocheck_bb4(n, (base, base + 1), base + 2, base + 3)


def _workers(workers: int | None) -> int:
    import os
    return workers or max(1, min(16, os.cpu_count() or 1))


def batch_instances(spec: BatchSpec, workers: int | None = None) -> list[Instance]:
    """Every instance of a batch spec, generated on host threads (numpy
    releases the interpreter lock in its array kernels)."""
    from concurrent.futures import ThreadPoolExecutor
    from functools import partial
    with ThreadPoolExecutor(_workers(workers)) as ex:
        return list(ex.map(partial(batch_instance, spec), range(len(spec.sizes))))


@dataclass
class ServeSpec:
    arrivals: np.ndarray        # seconds
    sizes: list[int]
    rate: float
    duration: float
    seed: int = 5


def serve_spec(rate: float = 2000.0, duration: float = 60.0, seed: int = 5, lo: int = 14,
               hi: int = 24) -> ServeSpec:
    """Poisson arrivals at `rate`/s over `duration` s; n uniform in {lo..hi}."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=int(rate * duration * 1.2) + 16)
    t = np.cumsum(gaps)
    t = t[t < duration]
    sizes = [int(x) for x in rng.integers(lo, hi + 1, size=t.size)]
    return ServeSpec(t, sizes, rate, duration, seed)


def serve_inputs(spec: ServeSpec, per_size: int = 4,
                 workers: int | None = None) -> dict[int, list[Instance]]:
    """A pool of `per_size` distinct zerocheck inputs for every size in a serve
    spec (request i uses input i mod per_size of its size, with a transcript
    label of its own, ``serve_label(i)``)."""
    from concurrent.futures import ThreadPoolExecutor
    keys = [(n, j) for n in sorted(set(spec.sizes)) for j in range(per_size)]

    def make(key):
        n, j = key
        base = 20_000 + 100 * n + 4 * j
        return zerocheck_bb4(n, (base, base + 1), base + 2, base + 3)

    with ThreadPoolExecutor(_workers(workers)) as ex:
        made = list(ex.map(make, keys))
    pool: dict[int, list[Instance]] = {}
    for (n, _), inst in zip(keys, made):
        pool.setdefault(n, []).append(inst)
    return pool


def serve_label(i: int) -> bytes:
    """Transcript label of request i of a serve trace."""
    return label(1_000_000 + i)


# Algorithmic HBM bytes of one proof under the fused-pass model: pass 0 reads
# the input tables, pass r >= 1 reads the round-(r-1) tables and writes the
# round-r tables; eq weights are factored and cost nothing.
def algorithmic_bytes(k: int, n: int, w0: int, wc: int) -> int:
    total = k * (1 << n) * w0
    for r in range(1, n):
        w_prev = w0 if r == 1 else wc
        total += k * (1 << (n - r + 1)) * w_prev + k * (1 << (n - r)) * wc
    return total
