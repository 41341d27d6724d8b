"""Oracle field arithmetic on plain Python integers (TEST INFRASTRUCTURE ONLY).

Restates the third-party ``zk_dtypes`` fields the reference computes in
(pinned ``zk-dtypes==0.0.17``, reference requirements.in:31), as used at
zorch/sumcheck/domain.py:101-102,232-233 and zorch/poly/univariate.py:55-65:

* ``babybear_mont``   p = 2^31 - 2^27 + 1, stored word = a * 2^32 mod p
                      (reference zorch/testkit/random_field.py:4-5).
* ``babybearx4_mont`` F_p[X]/(X^4 - 11), four Montgomery words c0..c3
                      (limb order: zorch/utils/field.py:44-60).
* ``goldilocks``      p = 2^64 - 2^32 + 1, canonical u64 storage
                      (reference docs/reference/conventions.md:266-268).
* ``koalabear_mont``  p = 2^31 - 2^24 + 1, Montgomery words as BabyBear's;
                      only the LogUp-GKR circuit's golden vectors use it
                      (zorch/logup_gkr/testing/__init__.py:33,161-173).

Elements are canonical Python ints (prime fields) or 4-tuples of canonical
ints (the extension). ``to_words``/``from_words`` are the storage codecs; a
stored word that is not reduced (a squeezed word >= p) decodes to its value
``w * R^-1 mod p`` (Montgomery) or ``w mod p`` (canonical), which is what the
reference's arithmetic sees for such a word.
"""

from __future__ import annotations

from dataclasses import dataclass

BB_P = (1 << 31) - (1 << 27) + 1
KB_P = (1 << 31) - (1 << 24) + 1
GL_P = (1 << 64) - (1 << 32) + 1
BB4_W = 11
_R = 1 << 32
_RINV = {BB_P: pow(_R, -1, BB_P), KB_P: pow(_R, -1, KB_P)}


@dataclass(frozen=True)
class PrimeField:
    name: str
    p: int
    word_bytes: int  # storage bytes per element
    montgomery: bool
    degree: int = 1

    @property
    def item_bytes(self) -> int:
        return self.word_bytes

    @property
    def base(self) -> "PrimeField":
        return self

    # --- arithmetic on canonical ints ---
    def zero(self) -> int:
        return 0

    def one(self) -> int:
        return 1

    def from_int(self, v: int) -> int:
        return v % self.p

    def add(self, a: int, b: int) -> int:
        return (a + b) % self.p

    def sub(self, a: int, b: int) -> int:
        return (a - b) % self.p

    def neg(self, a: int) -> int:
        return (-a) % self.p

    def mul(self, a: int, b: int) -> int:
        return (a * b) % self.p

    def inv(self, a: int) -> int:
        if a % self.p == 0:
            raise ZeroDivisionError("inverse of zero")
        return pow(a, self.p - 2, self.p)

    def embed(self, a: int) -> int:
        return a

    # --- storage codec ---
    def to_words(self, a: int) -> list[int]:
        """Storage word(s) of a canonical value."""
        if self.montgomery:
            return [(a * _R) % self.p]
        return [a % self.p]

    def from_words(self, words) -> int:
        """Value of a storage word (reduced or not)."""
        (w,) = words
        if self.montgomery:
            return (w * _RINV[self.p]) % self.p
        return w % self.p

    def to_bytes(self, a: int) -> bytes:
        return b"".join(w.to_bytes(self.word_bytes, "little") for w in self.to_words(a))

    def words_from_bytes(self, data: bytes) -> list[int]:
        return [int.from_bytes(data[i : i + self.word_bytes], "little")
                for i in range(0, len(data), self.word_bytes)]


@dataclass(frozen=True)
class QuarticField:
    """BabyBear^4 = F_p[X]/(X^4 - W), elements (c0, c1, c2, c3)."""

    name: str
    base: PrimeField
    w: int
    degree: int = 4

    @property
    def p(self) -> int:
        return self.base.p

    @property
    def word_bytes(self) -> int:
        return self.base.word_bytes

    @property
    def item_bytes(self) -> int:
        return 4 * self.base.word_bytes

    def zero(self):
        return (0, 0, 0, 0)

    def one(self):
        return (1, 0, 0, 0)

    def from_int(self, v: int):
        return (v % self.p, 0, 0, 0)

    def embed(self, a):
        if isinstance(a, tuple):
            return a
        return (a % self.p, 0, 0, 0)

    def add(self, a, b):
        a, b = self.embed(a), self.embed(b)
        return tuple((x + y) % self.p for x, y in zip(a, b))

    def sub(self, a, b):
        a, b = self.embed(a), self.embed(b)
        return tuple((x - y) % self.p for x, y in zip(a, b))

    def neg(self, a):
        return tuple((-x) % self.p for x in self.embed(a))

    def mul(self, a, b):
        a, b = self.embed(a), self.embed(b)
        p = self.p
        prod = [0] * 7
        for i in range(4):
            for j in range(4):
                prod[i + j] += a[i] * b[j]
        # X^4 = W folds degrees 4..6 onto 0..2.
        return tuple((prod[k] + (self.w * prod[k + 4] if k + 4 < 7 else 0)) % p
                     for k in range(4))

    def pow(self, a, e: int):
        out = self.one()
        base = self.embed(a)
        while e:
            if e & 1:
                out = self.mul(out, base)
            base = self.mul(base, base)
            e >>= 1
        return out

    def inv(self, a):
        a = self.embed(a)
        if a == (0, 0, 0, 0):
            raise ZeroDivisionError("inverse of zero")
        return self.pow(a, self.p ** 4 - 2)

    def to_words(self, a) -> list[int]:
        return [w for c in self.embed(a) for w in self.base.to_words(c)]

    def from_words(self, words):
        words = list(words)
        assert len(words) == 4
        return tuple(self.base.from_words([w]) for w in words)

    def to_bytes(self, a) -> bytes:
        return b"".join(w.to_bytes(4, "little") for w in self.to_words(a))

    def words_from_bytes(self, data: bytes) -> list[int]:
        return [int.from_bytes(data[i : i + 4], "little") for i in range(0, len(data), 4)]


BABYBEAR = PrimeField("babybear_mont", BB_P, 4, True)
GOLDILOCKS = PrimeField("goldilocks", GL_P, 8, False)
BABYBEAR4 = QuarticField("babybearx4_mont", BABYBEAR, BB4_W)
KOALABEAR = PrimeField("koalabear_mont", KB_P, 4, True)

FIELDS = {f.name: f for f in (BABYBEAR, BABYBEAR4, GOLDILOCKS)}


def challenge_limbs(field) -> int:
    """Base-field squeezes per element (reference zorch/challenge.py:16-21)."""
    return field.degree


def limbs_over(chal_field, transcript_field) -> int:
    """Transcript words per challenge (reference zorch/challenge.py:44-55)."""
    words = challenge_limbs(chal_field)
    per_word = challenge_limbs(transcript_field)
    if words % per_word:
        raise ValueError(f"{chal_field.name} does not tile {transcript_field.name}")
    return words // per_word
