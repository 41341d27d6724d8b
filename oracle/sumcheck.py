"""Oracle dense sumcheck prover and verifier (TEST INFRASTRUCTURE ONLY).

Plain-Python restatement of the reference hot path on canonical field
integers (``oracle.fields``) and the hashlib transcript
(``oracle.transcript``). Small n only; ``sumcheck_ref.c`` is the same
algorithm for large n.

Reference call stack restated here:

* ``prove_product``  SumcheckProver.prove (zorch/sumcheck/stage.py:52-69)
  over StandardRound(ProductSummand(d)) (zorch/sumcheck/prover.py:97-174):
  per round msg[t] = sum_j prod_k (lo_k[j] + t*(hi_k[j]-lo_k[j])), t=0..d
  (domain.py:131-137,158-163,176-205), Fiat-Shamir hop
  (challenge.py:97-106), reduce_evals (reduce.py:27-32,
  poly/univariate.py:36-65), MSB fold (domain.py:226-233), bind (round.py:41-49).
* ``prove_eq``  EqPolyProver.prove (zorch/sumcheck/eq/stage.py:62-92) over
  EqPolyRound (zorch/sumcheck/eq/eq_poly.py:114-221): suffix eq tables
  (poly/eq.py:130-214), _eq_tables (eq_poly.py:144-187), _weighted_summand
  (63-97), sumcheck_poly_from_t (100-111), reduce_domain (reduce.py:43-52),
  fold and eq_prev update (eq_poly.py:208-212, poly/eq.py:16-26).
* ``verify``  SumcheckVerifier.verify (zorch/sumcheck/stage.py:72-92),
  verify (zorch/verify.py:26-48), SumcheckRound (zorch/sumcheck/verifier.py:44-71).

Decision pinned here (no reference test fixes it): messages, claim and
point live in the challenge field; base-field tables are lifted into it
before round 0 (the reference's tested pattern for extension challenges,
zorch/sumcheck/testing/verifier_test.py:97-114).
"""

from __future__ import annotations

from dataclasses import dataclass, field as dc_field

from oracle.fields import limbs_over
from oracle.transcript import OracleTranscript


def log2_strict(n: int) -> int:
    if n <= 0 or n & (n - 1):
        raise ValueError(f"{n} is not a power of two")
    return n.bit_length() - 1


# --- univariate helpers (zorch/poly/univariate.py) -------------------------

def lagrange_eval(F, evals, r):
    """eval_univariate: value at r of the polynomial with values `evals` on
    {0..len-1} (univariate.py:36-65, direct Lagrange form)."""
    n = len(evals)
    out = F.zero()
    for k in range(n):
        num, den = F.one(), F.one()
        for j in range(n):
            if j == k:
                continue
            num = F.mul(num, F.sub(r, F.from_int(j)))
            den = F.mul(den, F.from_int(k - j))
        out = F.add(out, F.mul(evals[k], F.mul(num, F.inv(den))))
    return out


def natural_to_coeffs(F, evals):
    """EvalDomain.to_coeffs for the naturals: inverse Vandermonde times values
    (domain.py:67-70, univariate.py:68-91)."""
    n = len(evals)
    coeffs = [F.zero()] * n
    for j in range(n):
        poly = [F.one()]
        den = F.one()
        for k in range(n):
            if k == j:
                continue
            nxt = [F.zero()] * (len(poly) + 1)
            for i, c in enumerate(poly):
                nxt[i] = F.add(nxt[i], F.mul(c, F.from_int(-k)))
                nxt[i + 1] = F.add(nxt[i + 1], c)
            poly = nxt
            den = F.mul(den, F.from_int(j - k))
        scale = F.mul(evals[j], F.inv(den))
        for i in range(n):
            coeffs[i] = F.add(coeffs[i], F.mul(poly[i], scale))
    return coeffs


def eval_coeffs(F, coeffs, r):
    """Horner form of eval_coeffs (univariate.py:113-...)."""
    out = coeffs[-1]
    for c in reversed(coeffs[:-1]):
        out = F.add(F.mul(out, r), c)
    return out


# --- Fiat-Shamir hop (zorch/challenge.py:80-106) ---------------------------

def fs_hop(transcript: OracleTranscript, F, T, msg):
    """observe(msg) then sample(L); challenge = the L raw words regrouped as one
    element of F (reinterpret_challenge, zorch/transcript.py:97-119)."""
    limbs = limbs_over(F, T)
    payload = b"".join(F.to_bytes(m) for m in msg)
    transcript, raw = transcript.observe_and_sample(payload, limbs)
    words = F.words_from_bytes(raw) if F.degree > 1 else T.words_from_bytes(raw)
    return transcript, words, F.from_words(words)


@dataclass
class OracleProof:
    msgs: list          # rounds x (d+1) elements of F
    point_words: list   # rounds x raw squeezed words (storage)
    point: list         # rounds x values
    value: object
    transcript: OracleTranscript
    final_tables: list = dc_field(default_factory=list)
    eq_prev: object = None


def _fold(F, tab, r):
    h = len(tab) // 2
    return [F.add(tab[j], F.mul(r, F.sub(tab[j + h], tab[j]))) for j in range(h)]


def prove_product(F, T, tables, claim, transcript: OracleTranscript, degree: int,
                  rounds: int | None = None) -> OracleProof:
    """StandardRound(ProductSummand(degree)) under fold_rounds."""
    if degree < 1:
        raise ValueError("degree must be >= 1")
    state = [[F.embed(x) for x in t] for t in tables]
    n = log2_strict(len(state[0])) if rounds is None else rounds
    msgs, words, point = [], [], []
    for _ in range(n):
        h = len(state[0]) // 2
        msg = []
        for t in range(degree + 1):
            tt = F.from_int(t)
            acc = F.zero()
            for j in range(h):
                prod = F.one()
                for tab in state:
                    lo, hi = tab[j], tab[j + h]
                    prod = F.mul(prod, F.add(lo, F.mul(tt, F.sub(hi, lo))))
                acc = F.add(acc, prod)
            msg.append(acc)
        transcript, rw, r = fs_hop(transcript, F, T, msg)
        claim = lagrange_eval(F, msg, r)
        state = [_fold(F, tab, r) for tab in state]
        msgs.append(msg)
        words.append(rw)
        point.append(r)
    return OracleProof(msgs, words, point, claim, transcript, [t for t in state])


# --- eq tables (zorch/poly/eq.py) -------------------------------------------

def expand_step_msb(F, state, coord):
    """expand_hypercube_step(msb=True): [state*(1-c), state*c] (eq.py:55-62)."""
    high = [F.mul(s, coord) for s in state]
    low = [F.sub(s, hh) for s, hh in zip(state, high)]
    return low + high


def suffix_eq_tables(F, w):
    """compute_eq_evaluations: [eq(w[-1:]), ..., eq(w)], each new coordinate
    prepended as the MSB (eq_poly.py:46-51, eq.py:130-214)."""
    state = [F.one()]
    tables = []
    for j in range(len(w) - 1, -1, -1):
        state = expand_step_msb(F, state, w[j])
        tables.append(state)
    return tables


def eq_table(F, w):
    """eq(w, x) over {0,1}^n with w[0] on the index MSB (eq.py:92-115)."""
    return suffix_eq_tables(F, w)[-1] if len(w) else [F.one()]


def eq_factor(F, t, z):
    """t*z + (1-t)*(1-z) (eq.py:16-26)."""
    one = F.one()
    return F.add(F.mul(t, z), F.mul(F.sub(one, t), F.sub(one, z)))


def prove_eq(F, T, factors, w, claim, transcript: OracleTranscript,
             summand_degree: int) -> OracleProof:
    """EqPolyProver(ProductSummand(summand_degree)).prove over the natural
    domain {0..summand_degree+1}."""
    if summand_degree < 1:
        raise ValueError("degree must be >= 1")
    state = [[F.embed(x) for x in t] for t in factors]
    w = [F.embed(x) for x in w]
    l = len(w)
    if log2_strict(len(state[0])) != l:
        raise ValueError("equality point needs one coordinate per variable")
    degree = summand_degree + 1
    l_half = l // 2
    eq_l_list = suffix_eq_tables(F, w[:l_half])
    eq_r_list = suffix_eq_tables(F, w[l_half:])
    eq_prev = F.one()
    msgs, words, point = [], [], []
    for _ in range(l):
        nlen = len(state[0])
        i = l - log2_strict(nlen) + 1
        rest = l - i
        if rest == 0:
            eq_wl, eq_wr = None, [F.one()]
        elif i < l_half:
            eq_wl, eq_wr = eq_l_list[(l_half - i) - 1], eq_r_list[-1]
        else:
            eq_wl, eq_wr = None, eq_r_list[rest - 1]
        h = nlen // 2
        t_evals = []
        for node in range(degree + 1):
            nd = F.from_int(node)
            acc = F.zero()
            for x in range(h):
                prod = F.one()
                for tab in state:
                    lo, hi = tab[x], tab[x + h]
                    prod = F.mul(prod, F.add(lo, F.mul(nd, F.sub(hi, lo))))
                if eq_wl is not None:
                    nxr = len(eq_wr)
                    wt = F.mul(eq_wl[x // nxr], eq_wr[x % nxr])
                else:
                    wt = eq_wr[x]
                acc = F.add(acc, F.mul(prod, wt))
            t_evals.append(acc)
        w_i = w[i - 1]
        l1 = F.mul(eq_prev, w_i)
        l0 = F.sub(eq_prev, l1)
        ldiff = F.sub(l1, l0)
        msg = [F.mul(F.add(l0, F.mul(F.from_int(nd), ldiff)), t_evals[nd])
               for nd in range(degree + 1)]
        transcript, rw, r = fs_hop(transcript, F, T, msg)
        claim = eval_coeffs(F, natural_to_coeffs(F, msg), r)
        state = [_fold(F, tab, r) for tab in state]
        eq_prev = F.mul(eq_prev, eq_factor(F, r, w_i))
        msgs.append(msg)
        words.append(rw)
        point.append(r)
    return OracleProof(msgs, words, point, claim, transcript, state, eq_prev)


def verify(F, T, claim, msgs, transcript: OracleTranscript, degree: int):
    """SumcheckVerifier.verify with SumcheckRound(degree): returns
    (point, value, transcript, ok)."""
    if len(msgs) == 0:
        raise ValueError("proof must be a non-empty 2-D array (one row per round)")
    ok = True
    point = []
    for msg in msgs:
        if len(msg) != degree + 1:
            raise ValueError(f"round message must have {degree + 1} evals, got {len(msg)}")
        transcript, _, r = fs_hop(transcript, F, T, msg)
        ok = ok and (claim == F.add(msg[0], msg[1]))
        claim = lagrange_eval(F, msg, r)
        point.append(r)
    return point, claim, transcript, ok


def eval_mle(F, evals, point):
    """Bind-one-variable-at-a-time MLE evaluation, MSB first (reference test
    oracle zorch/sumcheck/testing/__init__.py:19-29)."""
    cur = [F.embed(x) for x in evals]
    for r in point:
        cur = _fold(F, cur, r)
    return cur[0]


# --- message wire forms (TEST INFRASTRUCTURE ONLY) ---------------------------
# A wire is ("values",), ("domain", nodes, inf_index), ("coeffs",) or
# ("compressed",): the samples or coefficients a round sends, restated
# directly from the reference's sampling (EvalDomain.sample, domain.py:95-106;
# summand_evals, domain.py:176-205; sumcheck_poly_from_t, eq_poly.py:100-111),
# coefficient maps (EvalDomain.to_coeffs, domain.py:72-93;
# product_round_coeffs, domain.py:216-223; EqPolyRound._round_coeffs,
# eq_poly.py:189-206; CompressedProductRound._round_poly, prover.py:191-204)
# and reductions (reduce.py:27-66).

def interp(F, xs, ys):
    """Ascending coefficients of the interpolant through (xs[i], ys[i]):
    the inverse Vandermonde applied to the values, solved by Gauss-Jordan
    elimination on the Vandermonde system (an independent route from the
    Lagrange expansion of natural_to_coeffs)."""
    m = len(xs)
    rows = []
    for i in range(m):
        p, row = F.one(), []
        for _ in range(m):
            row.append(p)
            p = F.mul(p, xs[i])
        rows.append(row + [ys[i]])
    for col in range(m):
        piv = next(r for r in range(col, m) if rows[r][col] != F.zero())
        rows[col], rows[piv] = rows[piv], rows[col]
        s = F.inv(rows[col][col])
        rows[col] = [F.mul(v, s) for v in rows[col]]
        for r in range(m):
            if r != col and rows[r][col] != F.zero():
                f = rows[r][col]
                rows[r] = [F.sub(a, F.mul(f, b)) for a, b in zip(rows[r], rows[col])]
    return [rows[i][m] for i in range(m)]


def to_coeffs(F, nodes, inf_index, values):
    """EvalDomain(nodes, inf_index).to_coeffs(values) (domain.py:72-93); nodes
    None stands for the naturals."""
    if inf_index is None:
        xs = nodes if nodes is not None else [F.from_int(i) for i in range(len(values))]
        return interp(F, xs, values)
    pos = inf_index % len(values)
    v_inf = values[pos]
    finite = values[:pos] + values[pos + 1:]
    d = len(finite)
    xs = nodes if nodes is not None else [F.from_int(i) for i in range(d)]
    shifted = []
    for x, y in zip(xs, finite):
        xd = F.one()
        for _ in range(d):
            xd = F.mul(xd, x)
        shifted.append(F.sub(y, F.mul(v_inf, xd)))
    low = interp(F, xs, shifted) if d else []
    return low + [v_inf]


def _points(F, nodes, inf_index):
    """Sample points of a domain in order; None marks infinity."""
    pts = list(nodes)
    if inf_index is not None:
        pts.insert(inf_index % (len(nodes) + 1), None)
    return pts


def _sample(F, lo, hi, x):
    d = F.sub(hi, lo)
    return d if x is None else F.add(lo, F.mul(x, d))


def _product_samples(F, state, pts):
    h = len(state[0]) // 2
    out = []
    for x in pts:
        acc = F.zero()
        for j in range(h):
            prod = F.one()
            for tab in state:
                prod = F.mul(prod, _sample(F, tab[j], tab[j + h], x))
            acc = F.add(acc, prod)
        out.append(acc)
    return out


def _wire_message(F, wire, samples_of, degree):
    """The message of one round: samples_of(points) samples the round
    polynomial (of the given degree) at a point list."""
    kind = wire[0]
    if kind == "values":
        return samples_of([F.from_int(t) for t in range(degree + 1)])
    if kind == "domain":
        return samples_of(_points(F, wire[1], wire[2]))
    if kind == "coeffs":
        full = [None] + [F.from_int(t) for t in range(degree)]
        return to_coeffs(F, None, 0, samples_of(full))
    if kind == "compressed":
        s = samples_of([F.zero(), None])
        return [s[0], s[1]]
    raise ValueError(kind)


def wire_reduce(F, wire, claim, msg, r):
    """The wire form's claim reduction (reduce.py:27-66)."""
    kind = wire[0]
    if kind == "values":
        return lagrange_eval(F, msg, r)
    if kind == "domain":
        return eval_coeffs(F, to_coeffs(F, wire[1], wire[2], msg), r)
    if kind == "coeffs":
        return eval_coeffs(F, msg, r)
    c0, c2 = msg
    c1 = F.sub(F.sub(F.sub(claim, c0), c0), c2)
    return eval_coeffs(F, [c0, c1, c2], r)


def prove_product_wire(F, T, tables, claim, transcript: OracleTranscript, degree: int,
                       wire) -> OracleProof:
    """A product-summand stage whose rounds send `wire` messages:
    StandardRound on a domain, CompressedProductRound, or coefficients."""
    state = [[F.embed(x) for x in t] for t in tables]
    n = log2_strict(len(state[0]))
    msgs, words, point = [], [], []
    for _ in range(n):
        msg = _wire_message(F, wire, lambda pts: _product_samples(F, state, pts), degree)
        transcript, rw, r = fs_hop(transcript, F, T, msg)
        claim = wire_reduce(F, wire, claim, msg, r)
        state = [_fold(F, tab, r) for tab in state]
        msgs.append(msg)
        words.append(rw)
        point.append(r)
    return OracleProof(msgs, words, point, claim, transcript, state)


def prove_eq_wire(F, T, factors, w, claim, transcript: OracleTranscript, summand_degree: int,
                  wire, eq_mass=None) -> OracleProof:
    """EqPolyRound under fold_rounds with messages on `wire` (the reference's
    default is ("domain", Û nodes, 0)); `eq_mass` starts the bound eq mass."""
    state = [[F.embed(x) for x in t] for t in factors]
    w = [F.embed(x) for x in w]
    l = len(w)
    l_half = l // 2
    eq_l_list = suffix_eq_tables(F, w[:l_half])
    eq_r_list = suffix_eq_tables(F, w[l_half:])
    eq_prev = F.one() if eq_mass is None else eq_mass
    msgs, words, point = [], [], []
    for _ in range(l):
        nlen = len(state[0])
        i = l - log2_strict(nlen) + 1
        rest = l - i
        if rest == 0:
            eq_wl, eq_wr = None, [F.one()]
        elif i < l_half:
            eq_wl, eq_wr = eq_l_list[(l_half - i) - 1], eq_r_list[-1]
        else:
            eq_wl, eq_wr = None, eq_r_list[rest - 1]
        h = nlen // 2
        w_i = w[i - 1]
        l1 = F.mul(eq_prev, w_i)
        l0 = F.sub(eq_prev, l1)
        ldiff = F.sub(l1, l0)

        def samples_of(pts, state=state, eq_wl=eq_wl, eq_wr=eq_wr, h=h, l0=l0, ldiff=ldiff):
            out = []
            for x in pts:
                acc = F.zero()
                for y in range(h):
                    prod = F.one()
                    for tab in state:
                        prod = F.mul(prod, _sample(F, tab[y], tab[y + h], x))
                    if eq_wl is not None:
                        nxr = len(eq_wr)
                        wt = F.mul(eq_wl[y // nxr], eq_wr[y % nxr])
                    else:
                        wt = eq_wr[y]
                    acc = F.add(acc, F.mul(prod, wt))
                lx = ldiff if x is None else F.add(l0, F.mul(x, ldiff))
                out.append(F.mul(lx, acc))
            return out

        msg = _wire_message(F, wire, samples_of, summand_degree + 1)
        transcript, rw, r = fs_hop(transcript, F, T, msg)
        claim = wire_reduce(F, wire, claim, msg, r)
        state = [_fold(F, tab, r) for tab in state]
        eq_prev = F.mul(eq_prev, eq_factor(F, r, w_i))
        msgs.append(msg)
        words.append(rw)
        point.append(r)
    return OracleProof(msgs, words, point, claim, transcript, state, eq_prev)
