"""Oracle round-cost engines (TEST INFRASTRUCTURE ONLY).

Plain-Python restatements of the reference's three alternative sumcheck
engines, on canonical field integers (``oracle.fields``) and the oracle
transcripts. Small n only. Each one is pinned in ``tests/test_oracle_engines.py``
by the reference's own assertion for it (its messages equal a linear-time
prover's, or its round trip verifies), by closed forms, and by brute force.

* ``prove_sqrt_space``     zorch/sumcheck/sqrt_space.py:40-132: the first
  l//2 rounds refold the untouched factors over the bound prefix with the
  running eq(r[<i], .) table (compute_folded_evaluations, :40-48); the
  boundary materialises the fold; a plain product tail finishes.
* ``prove_univariate_skip`` zorch/sumcheck/univariate_skip.py:65-190: the
  first `skip` rounds collapse into one univariate round over the order-2^skip
  two-adic subgroup (round0_message :65-101, skip_round0 :104-133); the
  round-0 message is base-field ascending coefficients; the prism is bound at
  r0 and a product tail runs in the challenge field. The verifier
  (verify_univariate_skip :193-245) checks the subgroup sum.
* ``prove_eq_small_value``  zorch/sumcheck/eq/small_value.py:43-192 with
  zorch/sumcheck/eq/accumulators.py:41-128: accumulator rounds 1..l0, the
  transition round over [P_1..P_d, eq(w, .)], then the eq-poly tail on the
  compressed domain.

Parity unpinned here: the generator of the two-adic subgroup (FRX's
``lax.ntt`` canonical root, from a third-party library absent here). The
restatement uses the order-2^k power of the field's multiplicative generator
raised to (p-1)/2^k (BabyBear 31, Goldilocks 7); every property the reference
tests of the skip (subgroup sum = claim, prover/verifier agreement, skip 0 =
plain run) holds for any primitive root.
"""

from __future__ import annotations

from oracle.fields import limbs_over
from oracle.sumcheck import (OracleProof, _fold, _product_samples, _wire_message, eq_factor,
                             eval_coeffs, fs_hop, log2_strict, to_coeffs, wire_reduce)

# --- eq helpers (zorch/poly/eq.py:55-115) -----------------------------------


def expand_step_lsb(F, state, coord):
    """expand_hypercube_step(msb=False): the new variable is the LSB of the
    index, out[2b] = s_b (1 - c), out[2b + 1] = s_b c (eq.py:55-62)."""
    out = []
    for s in state:
        hi = F.mul(s, coord)
        out.append(F.sub(s, hi))
        out.append(hi)
    return out


def eq_hypercube(F, coords, scalar=None):
    """expand_eq_to_hypercube(coords, scalar): scalar * eq(coords, x), coords[0]
    on the index MSB (eq.py:92-115)."""
    state = [F.one() if scalar is None else scalar]
    for c in coords:
        state = expand_step_lsb(F, state, c)
    return state


# --- sqrt-space (zorch/sumcheck/sqrt_space.py) ------------------------------


def compute_folded_evaluations(F, tables, eq_evals):
    """p_k(r[<i], x') = sum_b eq(r[<i], b) p_k(b, x'): each table reshaped to
    (2^i, 2^(l-i)) with the bound prefix b on the high index bits
    (sqrt_space.py:40-48)."""
    i = log2_strict(len(eq_evals))
    out = []
    for tab in tables:
        L = len(tab) >> i
        out.append([_dot(F, eq_evals, [tab[b * L + x] for b in range(len(eq_evals))])
                    for x in range(L)])
    return out


def _dot(F, ws, xs):
    acc = F.zero()
    for w, x in zip(ws, xs):
        acc = F.add(acc, F.mul(w, x))
    return acc


def uhat_wire(F, degree):
    """The compressed domain U-hat_degree = {inf, 0, 2, ..., degree-1}
    (domain.py:140-146) as an oracle wire."""
    nodes = [F.from_int(0)] + [F.from_int(u) for u in range(2, degree)]
    return ("domain", nodes, 0)


def prove_sqrt_space(F, T, tables, claim, transcript, degree, wire=None) -> OracleProof:
    """prove_sqrt_space(p, claim, transcript, ProductSummand(degree), domain):
    l//2 refold rounds, the boundary fold, then l - l//2 plain rounds, all
    sending `wire` messages (U-hat_degree by default, sqrt_space.py:103-104)."""
    wire = wire or uhat_wire(F, degree)
    state = [[F.embed(x) for x in t] for t in tables]
    l = log2_strict(len(state[0]))
    l_half = l // 2
    eq = [F.one()]
    msgs, words, point = [], [], []
    for _ in range(l_half):
        folded = compute_folded_evaluations(F, state, eq)
        msg = _wire_message(F, wire, lambda pts, f=folded: _product_samples(F, f, pts), degree)
        transcript, rw, r = fs_hop(transcript, F, T, msg)
        claim = wire_reduce(F, wire, claim, msg, r)
        eq = expand_step_lsb(F, eq, r)
        msgs.append(msg)
        words.append(rw)
        point.append(r)
    folded = compute_folded_evaluations(F, state, eq)
    for _ in range(l - l_half):
        msg = _wire_message(F, wire, lambda pts, f=folded: _product_samples(F, f, pts), degree)
        transcript, rw, r = fs_hop(transcript, F, T, msg)
        claim = wire_reduce(F, wire, claim, msg, r)
        folded = [_fold(F, tab, r) for tab in folded]
        msgs.append(msg)
        words.append(rw)
        point.append(r)
    return OracleProof(msgs, words, point, claim, transcript, folded)


# --- two-adic subgroup (zorch/sumcheck/domain.py:237-311) --------------------

_GENERATOR = {"babybear_mont": 31, "goldilocks": 7}


def subgroup_root(F, k: int) -> int:
    """A primitive 2^k-th root of unity of the prime field F: the
    multiplicative generator raised to (p - 1) / 2^k."""
    p = F.p
    if (p - 1) % (1 << k):
        raise ValueError(f"{F.name} has no subgroup of order 2^{k}")
    return pow(_GENERATOR[F.name], (p - 1) >> k, p)


def ntt(F, values, inverse=False):
    """Natural-order transform over the subgroup of order len(values):
    out[j] = sum_i values[i] w^(ij) (NTT(e_1)_j = w^j), the inverse scaled by
    1/len (lax.ntt NTT / INTT as the reference uses them, domain.py:245-266)."""
    n = len(values)
    k = log2_strict(n)
    w = subgroup_root(F, k) if n > 1 else 1
    if inverse:
        w = pow(w, F.p - 2, F.p)
    out = []
    for j in range(n):
        acc, wj, x = 0, pow(w, j, F.p), 1
        for v in values:
            acc = (acc + v * x) % F.p
            x = x * wj % F.p
        out.append(acc)
    if inverse:
        ninv = pow(n, F.p - 2, F.p)
        out = [v * ninv % F.p for v in out]
    return out


def subgroup_to_coeffs(F, values):
    return ntt(F, values, inverse=True)


def subgroup_evals(F, coeffs, size):
    if size < len(coeffs):
        raise ValueError(f"subgroup_evals size {size} < coeff count {len(coeffs)}")
    return ntt(F, list(coeffs) + [0] * (size - len(coeffs)))


def subgroup_sum(F, coeffs, skip: int):
    """sum over the order-2^skip subgroup = |D| * sum of the coefficients at
    multiples of |D| (domain.py:269-278)."""
    size = 1 << skip
    acc = 0
    for c in coeffs[::size]:
        acc = (acc + c) % F.p
    return acc * size % F.p


def _next_pow2(x: int) -> int:
    n = 1
    while n < x:
        n <<= 1
    return n


def round0_message(F, tables, skip: int, degree: int):
    """(msg0 ascending coefficients of degree degree*(|D|-1), the factors'
    D-coefficients per column) in the base field F (univariate_skip.py:65-101):
    the skip MSB variables index the subgroup points; each column's |D|
    values are interpolated (iNTT), extended to the order-M subgroup, the
    factors multiplied, the columns summed, and the sums interpolated."""
    size = 1 << skip
    hn = len(tables[0]) // size
    coeffs_z = [[subgroup_to_coeffs(F, [tab[b * hn + x] for b in range(size)])
                 for x in range(hn)] for tab in tables]
    d0 = degree * (size - 1)
    big = _next_pow2(d0 + 1)
    sums = [0] * big
    for x in range(hn):
        prod = [1] * big
        for k in range(len(tables)):
            ev = subgroup_evals(F, coeffs_z[k][x], big)
            prod = [a * b % F.p for a, b in zip(prod, ev)]
        sums = [(a + b) % F.p for a, b in zip(sums, prod)]
    return subgroup_to_coeffs(F, sums)[: d0 + 1], coeffs_z


def _hop_base(transcript, C, T, Fb, msg):
    """observe(msg in the base field Fb) then sample a challenge of C: the
    observed bytes are the base-field storage bytes (the round-0 message keeps
    the tables' dtype, univariate_skip.py:121-123)."""
    limbs = limbs_over(C, T)
    payload = b"".join(Fb.to_bytes(m) for m in msg)
    transcript, raw = transcript.observe_and_sample(payload, limbs)
    words = C.words_from_bytes(raw) if C.degree > 1 else T.words_from_bytes(raw)
    return transcript, words, C.from_words(words)


def prove_univariate_skip(Fb, C, T, tables, claim, skip: int, transcript, degree: int):
    """prove_univariate_skip(p, claim, skip, transcript, ProductSummand(degree),
    challenges=ChallengePolicy(C)): msg0 in the base field, then a natural-domain
    product tail over the bound prism in C (univariate_skip.py:136-190).
    Returns (msg0 or None, OracleProof of the tail with r0 first in the point)."""
    from oracle.sumcheck import prove_product

    total = log2_strict(len(tables[0]))
    if not 0 <= skip <= total:
        raise ValueError(f"skip_rounds must be in [0, {total}], got {skip}")
    if skip == 0:
        return None, prove_product(C, T, tables, C.embed(claim), transcript, degree)
    msg0, coeffs_z = round0_message(Fb, tables, skip, degree)
    transcript, rw, r0 = _hop_base(transcript, C, T, Fb, msg0)
    reduced = eval_coeffs(C, [C.embed(c) for c in msg0], r0)
    state = [[eval_coeffs(C, [C.embed(c) for c in col], r0) for col in tab] for tab in coeffs_z]
    tail = prove_product(C, T, state, reduced, transcript, degree, rounds=total - skip)
    tail.msgs = tail.msgs
    tail.point_words = [rw] + tail.point_words
    tail.point = [r0] + tail.point
    return msg0, tail


def verify_univariate_skip(Fb, C, T, claim, msg0, msgs, skip: int, total: int, transcript,
                           degree: int):
    """verify_univariate_skip (univariate_skip.py:193-245): the subgroup-sum
    check and s0(r0) reduction, then the natural-domain tail. Returns
    (value, transcript, point, ok)."""
    from oracle.sumcheck import verify

    if not 0 <= skip <= total:
        raise ValueError(f"skip_rounds must be in [0, {total}], got {skip}")
    if skip == 0:
        point, value, transcript, ok = verify(C, T, C.embed(claim), msgs, transcript, degree)
        return value, transcript, point, ok
    n = total - skip
    if len(msgs) != n:
        raise ValueError(f"need {1 + n} messages, got {1 + len(msgs)}")
    if len(msg0) != degree * ((1 << skip) - 1) + 1:
        raise ValueError("round message must have "
                         f"{degree * ((1 << skip) - 1) + 1} coefficients, got {len(msg0)}")
    ok = Fb.embed(claim) == subgroup_sum(Fb, msg0, skip)
    transcript, _, r0 = _hop_base(transcript, C, T, Fb, msg0)
    value = eval_coeffs(C, [C.embed(c) for c in msg0], r0)
    point = [r0]
    if msgs:
        tail_point, value, transcript, ok_t = verify(C, T, value, msgs, transcript, degree)
        ok = ok and ok_t
        point += tail_point
    return value, transcript, point, ok


# --- small-value eq-poly (zorch/sumcheck/eq/small_value.py) -----------------


def extend_to_round_domain(F, p0, p1, d):
    """[inf, 0, 1, 2, ..., d-1] samples of p0 + x (p1 - p0) (domain.py:109-127)."""
    diff = F.sub(p1, p0)
    if d == 1:
        return [diff, p0]
    return [diff, p0, p1] + [F.add(p0, F.mul(F.from_int(u), diff)) for u in range(2, d)]


def extend_prefix_to_domain(F, evals, d, k):
    """The first k variables extended from {0,1} to U_d, the rest left on the
    cube, in lexicographic order (accumulators.py:22-38)."""
    l = log2_strict(len(evals))
    size = d + 1
    for i in range(k):
        prefix, suffix = size ** i, 1 << (l - i - 1)
        out = [None] * (prefix * size * suffix)
        for p in range(prefix):
            for s in range(suffix):
                lo = evals[p * 2 * suffix + s]
                hi = evals[p * 2 * suffix + suffix + s]
                for n, v in enumerate(extend_to_round_domain(F, lo, hi, d)):
                    out[p * size * suffix + n * suffix + s] = v
        evals = out
    return evals


def _decode_beta(beta_idx, l0, d):
    digits = []
    for _ in range(l0):
        digits.append(beta_idx % (d + 1))
        beta_idx //= d + 1
    return tuple(reversed(digits))


def idx4(beta_idx, l0, d, literal=False):
    """Routes of beta to (round i, v, u, y) accumulator slots
    (accumulators.py:101-128): u in U-hat_d (digit 2, the node 1, dropped), y on
    the cube. v and y are read most-significant digit first, the order of the
    R tensor (small_value.py:96-98) and of the e_out rows (small_value.py:
    138-145) they are contracted with. `literal` reads them least-significant
    first, as the reference's _idx4 writes them; the two agree while v and y
    have at most one digit (l0 <= 2) and differ from l0 = 3 on, where only the
    most-significant-first order reproduces the eq-poly messages the engine
    is specified to send."""
    beta = _decode_beta(beta_idx, l0, d)
    out = []
    for i in range(1, l0 + 1):
        u_digit = beta[i - 1]
        if u_digit == 2:
            continue
        y_digits = beta[i:]
        if not all(y in (1, 2) for y in y_digits):
            continue
        u = u_digit - 1 if u_digit > 2 else u_digit
        if literal:
            y = sum((yv - 1) << j for j, yv in enumerate(y_digits))
            v = sum(vv * ((d + 1) ** j) for j, vv in enumerate(beta[: i - 1]))
        else:
            y = 0
            for yv in y_digits:
                y = 2 * y + (yv - 1)
            v = 0
            for vv in beta[: i - 1]:
                v = (d + 1) * v + vv
        out.append((i, v, u, y))
    return out


def precompute_accumulators(F, tables, e_in, e_out, literal=False):
    """[A_1, ..., A_l0], A_i of (d+1)^(i-1) x d over (v, u)
    (accumulators.py:41-91)."""
    d = len(tables)
    l = log2_strict(len(tables[0]))
    l_half = l // 2
    l0 = len(e_out)
    x_in_size = 1 << l_half
    x_out_size = 1 << (l - l_half - l0)
    beta_size = (d + 1) ** l0
    ext = [extend_prefix_to_domain(F, t, d, l0) for t in tables]
    t_a = [[F.zero()] * x_out_size for _ in range(beta_size)]
    for beta in range(beta_size):
        for xi in range(x_in_size):
            for xo in range(x_out_size):
                idx = (beta * x_in_size + xi) * x_out_size + xo
                prod = F.one()
                for e in ext:
                    prod = F.mul(prod, e[idx])
                t_a[beta][xo] = F.add(t_a[beta][xo], F.mul(e_in[xi], prod))
    acc = [[F.zero()] * ((d + 1) ** i * d) for i in range(l0)]
    for beta in range(beta_size):
        for i, v, u, y in idx4(beta, l0, d, literal):
            row = e_out[i - 1][y]
            contrib = _dot(F, row, t_a[beta])
            acc[i - 1][v * d + u] = F.add(acc[i - 1][v * d + u], contrib)
    return acc


def _lagrange_basis(F, r, nodes):
    """compute_lagrange_basis(r, nodes): L_j(r) over the finite nodes."""
    out = []
    for j, xj in enumerate(nodes):
        num, den = F.one(), F.one()
        for k, xk in enumerate(nodes):
            if k != j:
                num = F.mul(num, F.sub(r, xk))
                den = F.mul(den, F.sub(xj, xk))
        out.append(F.mul(num, F.inv(den)))
    return out


def _lagrange_over_round_domain(F, r, d):
    """[L_inf, L_0, ..., L_{d-1}] at r over U_d: L_inf the vanishing polynomial
    of the finite nodes (small_value.py:47-55)."""
    finite = [F.from_int(u) for u in range(d)]
    van = F.one()
    for x in finite:
        van = F.mul(van, F.sub(r, x))
    return [van] + _lagrange_basis(F, r, finite)


def _poly_from_t(F, t_evals, l0, l1, wire):
    """sumcheck_poly_from_t on a domain with infinity first (eq_poly.py:100-111)."""
    ldiff = F.sub(l1, l0)
    nodes = wire[1]
    finite = [F.mul(F.add(l0, F.mul(x, ldiff)), t) for x, t in zip(nodes, t_evals[1:])]
    return [F.mul(ldiff, t_evals[0])] + finite


def _hop_one(transcript, F, T, msg):
    """transcript.observe_and_sample(msg, 1): one transcript word, the
    challenge in the transcript field (small_value.py:95, 121)."""
    payload = b"".join(F.to_bytes(m) for m in msg)
    transcript, raw = transcript.observe_and_sample(payload, 1)
    words = T.words_from_bytes(raw)
    return transcript, words, T.from_words(words)


def prove_eq_small_value(F, T, tables, w, l0: int, claim, transcript, literal=False):
    """prove_eq_poly_small_value(p, w, l0, claim, transcript, challenges) with the
    challenge field F equal to the transcript field T (one word a challenge)
    (small_value.py:148-192). Returns an OracleProof (final tables and eq mass).
    `literal`: the reference's own accumulator routing (see idx4)."""
    from oracle.sumcheck import prove_eq_wire

    d = len(tables)
    state = [[F.embed(x) for x in t] for t in tables]
    w = [F.embed(x) for x in w]
    l = log2_strict(len(state[0]))
    if d < 2:
        raise ValueError("the small-value engine needs degree >= 2: U_1 = {inf, 0} routes no "
                         "cube point through node 1")
    if not 1 <= l0 <= l - l // 2:
        raise ValueError(f"l_0 must be in [1, {l - l // 2}] (the small-value rounds fit before "
                         f"the out-half); got l_0={l0} for l={l}")
    l_half = l // 2
    e_in = eq_hypercube(F, w[l0:l0 + l_half])
    w_x_out = w[l_half + l0:]
    x_out_size = 1 << (l - l_half - l0)
    e_out = []
    for i in range(1, l0 + 1):
        flat = eq_hypercube(F, w[i:l0] + w_x_out)
        e_out.append([flat[y * x_out_size:(y + 1) * x_out_size] for y in range(1 << (l0 - i))])
    accs = precompute_accumulators(F, state, e_in, e_out, literal)
    wire = uhat_wire(F, d)
    r_tensor, eq_prev, eq_evals = [F.one()], F.one(), [F.one()]
    msgs, words, point = [], [], []
    for i in range(l0):
        A = accs[i]
        t_evals = [_dot(F, r_tensor, [A[v * d + u] for v in range(len(r_tensor))])
                   for u in range(d)]
        l1 = F.mul(eq_prev, w[i])
        lz = F.sub(eq_prev, l1)
        msg = _poly_from_t(F, t_evals, lz, l1, wire)
        transcript, rw, r = _hop_one(transcript, F, T, msg)
        claim = eval_coeffs(F, to_coeffs(F, wire[1], wire[2], msg), r)
        lag = _lagrange_over_round_domain(F, r, d)
        r_tensor = [F.mul(a, b) for a in r_tensor for b in lag]
        eq_prev = F.mul(eq_prev, eq_factor(F, r, w[i]))
        eq_evals = expand_step_lsb(F, eq_evals, r)
        msgs.append(msg)
        words.append(rw)
        point.append(r)
    # Transition: refold [P_1..P_d, eq(w, .)] over the bound prefix, one
    # product round over the d + 1 factors on U-hat_d.
    weights = eq_hypercube(F, w)
    folded = compute_folded_evaluations(F, state + [weights], eq_evals)
    msg = _wire_message(F, wire, lambda pts: _product_samples(F, folded, pts), d + 1)
    transcript, rw, r = _hop_one(transcript, F, T, msg)
    claim = eval_coeffs(F, to_coeffs(F, wire[1], wire[2], msg), r)
    folded = [_fold(F, tab, r) for tab in folded]
    eq_prev = F.mul(eq_prev, eq_factor(F, r, w[l0]))
    msgs.append(msg)
    words.append(rw)
    point.append(r)
    # Tail: the eq-poly rounds over the remaining variables of w, starting
    # from the eq mass bound so far.
    tail = prove_eq_wire(F, T, folded[:d], w[l0 + 1:], claim, transcript, d, wire,
                         eq_mass=eq_prev) if l - l0 - 1 > 0 else None
    if tail is None:
        return OracleProof(msgs, words, point, claim, transcript, folded[:d], eq_prev)
    return OracleProof(msgs + tail.msgs, words + tail.point_words, point + tail.point,
                       tail.value, tail.transcript, tail.final_tables, tail.eq_prev)
