// Host construction of a round-message wire form's maps (Ctl::wmat, cmat,
// cclaim) from the C ABI's zs_wire.
//
// The kernels always compute the round polynomial s on the natural nodes
// {0..D} (D = npts - 1: the degree, plus one with the eq factor); every other
// wire form is a fixed linear map of those D + 1 values, and its claim
// reduction a fixed linear map of the message (and, for the compressed form,
// the running claim) to ascending coefficients, evaluated at the challenge.
// Because field arithmetic is exact, these maps give the reference's values
// bit for bit whatever formula computes them:
//  * ZS_WIRE_DOMAIN: s sampled at an evaluation domain's points, the finite
//    nodes in order with s(inf) inserted at position inf_pos
//    (EvalDomain.sample, reference zorch/sumcheck/domain.py:95-106). s(x) for
//    a node x is the Lagrange form over {0..D}; s(inf) is the leading
//    coefficient, sum_i (-1)^(D-i) C(D, i) s(i) / D! (what the reference
//    samples as the product of the factor slopes). Reduction: reduce_domain
//    (reduce.py:43-52) = to_coeffs (domain.py:72-93) then eval_coeffs.
//  * ZS_WIRE_COEFFS: the ascending coefficients (inverse Vandermonde over the
//    naturals; product_round_coeffs, domain.py:216-223, and EqPolyRound's
//    coefficient form, eq_poly.py:189-206); reduce_coeffs (reduce.py:35-40).
//  * ZS_WIRE_COMPRESSED: [c0, c2] of a degree-2 product round
//    (CompressedProductRound, prover.py:177-212); reduce_compressed rebuilds
//    c1 = claim - 2 c0 - c2 (reduce.py:55-66).
#pragma once
#include <stdint.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/zorch_amd.h"
#include "kernels.h"
#include "runtime.h"

namespace zs {

template <class FC>
using FMat = std::vector<std::vector<typename FC::T>>;

// Inverse of a square matrix over the field; false if it is singular.
template <class FC>
bool field_inverse(FMat<FC> a, FMat<FC>& inv) {
  using T = typename FC::T;
  const size_t n = a.size();
  inv.assign(n, std::vector<T>(n, FC::zero()));
  for (size_t i = 0; i < n; ++i) inv[i][i] = FC::one();
  for (size_t col = 0; col < n; ++col) {
    size_t piv = col;
    while (piv < n && FC::eq(a[piv][col], FC::zero())) ++piv;
    if (piv == n) return false;
    std::swap(a[piv], a[col]);
    std::swap(inv[piv], inv[col]);
    const T s = FC::inv(a[col][col]);
    for (size_t j = 0; j < n; ++j) {
      a[col][j] = FC::mul(a[col][j], s);
      inv[col][j] = FC::mul(inv[col][j], s);
    }
    for (size_t row = 0; row < n; ++row) {
      if (row == col || FC::eq(a[row][col], FC::zero())) continue;
      const T f = a[row][col];
      for (size_t j = 0; j < n; ++j) {
        a[row][j] = FC::sub(a[row][j], FC::mul(f, a[col][j]));
        inv[row][j] = FC::sub(inv[row][j], FC::mul(f, inv[col][j]));
      }
    }
  }
  return true;
}

template <class FC>
typename FC::T field_pow(typename FC::T x, uint32_t e) {
  typename FC::T r = FC::one();
  for (uint32_t i = 0; i < e; ++i) r = FC::mul(r, x);
  return r;
}

// Inverse Vandermonde of the nodes x: row k maps values at x to coefficient k.
template <class FC>
bool inv_vandermonde(const std::vector<typename FC::T>& x, FMat<FC>& out) {
  const size_t m = x.size();
  FMat<FC> v(m, std::vector<typename FC::T>(m));
  for (size_t i = 0; i < m; ++i)
    for (size_t k = 0; k < m; ++k) v[i][k] = field_pow<FC>(x[i], (uint32_t)k);
  return field_inverse<FC>(v, out);
}

// Lagrange basis over the naturals {0..d} at x: s(x) = sum_i L[i] s(i).
template <class FC>
std::vector<typename FC::T> natural_basis_at(uint32_t d, typename FC::T x) {
  using T = typename FC::T;
  std::vector<T> out(d + 1);
  for (uint32_t i = 0; i <= d; ++i) {
    T num = FC::one();
    int64_t den = 1;
    for (uint32_t j = 0; j <= d; ++j) {
      if (j == i) continue;
      num = FC::mul(num, FC::sub(x, FC::from_int(j)));
      den *= (int64_t)i - (int64_t)j;
    }
    out[i] = FC::mul(num, FC::inv(FC::from_int(den)));
  }
  return out;
}

// Leading coefficient of the degree-d interpolant over {0..d}.
template <class FC>
std::vector<typename FC::T> natural_leading(uint32_t d) {
  std::vector<typename FC::T> out(d + 1);
  int64_t fact = 1;
  for (uint32_t i = 2; i <= d; ++i) fact *= i;
  const typename FC::T inv_fact = FC::inv(FC::from_int(fact));
  int64_t binom = 1;  // C(d, i)
  for (uint32_t i = 0; i <= d; ++i) {
    out[i] = FC::mul(FC::from_int(((d - i) & 1) ? -binom : binom), inv_fact);
    binom = binom * (int64_t)(d - i) / (int64_t)(i + 1);
  }
  return out;
}

// Fill the control block's wire fields; ZS_OK or ZS_EINVAL (message set).
template <class FC>
int set_ctl_wire(Ctl<FC>& c, uint32_t npts, const zs_wire* w) {
  using T = typename FC::T;
  for (int i = 0; i < MAXPTS; ++i) {
    c.cclaim[i] = FC::zero();
    for (int j = 0; j < MAXPTS; ++j) c.wmat[i][j] = c.cmat[i][j] = FC::zero();
  }
  c.wire = WIRE_VALUES;
  c.mlen = npts;
  c.ncoef = npts;
  c.claim_every = 0;
  if (!w || w->kind == ZS_WIRE_VALUES) return ZS_OK;
  const uint32_t d = npts - 1;
  const uint32_t m = wire_width(w, npts, true);
  if (!m) return ZS_EINVAL;
  c.wire = w->kind;
  c.mlen = m;
  FMat<FC> W(m, std::vector<T>(npts, FC::zero()));
  FMat<FC> C;
  std::vector<T> gam;
  if (w->kind == ZS_WIRE_DOMAIN) {
    std::vector<T> x(w->num_nodes);
    for (uint32_t i = 0; i < w->num_nodes; ++i) {
      T v;
      memcpy(&v, (const uint8_t*)w->nodes + i * sizeof(T), sizeof(T));
      x[i] = FC::canon(v);
    }
    const uint32_t pos = w->has_inf ? w->inf_pos : m;
    for (uint32_t j = 0, f = 0; j < m; ++j) {
      if (j == pos) {
        W[j] = natural_leading<FC>(d);
      } else {
        W[j] = natural_basis_at<FC>(d, x[f++]);
      }
    }
    FMat<FC> vi;
    if (!inv_vandermonde<FC>(x, vi))
      return fail(ZS_EINVAL, "evaluation-domain nodes must be distinct");
    C.assign(m, std::vector<T>(m, FC::zero()));
    if (!w->has_inf) {
      C = vi;
    } else {
      // The finite samples interpolate s - s(inf) x^d' (d' finite nodes);
      // s(inf) is the top coefficient.
      const uint32_t dp = w->num_nodes;
      for (uint32_t k = 0; k < dp; ++k) {
        T lead = FC::zero();
        for (uint32_t j = 0, f = 0; j < m; ++j) {
          if (j == pos) continue;
          C[k][j] = vi[k][f];
          lead = FC::add(lead, FC::mul(vi[k][f], field_pow<FC>(x[f], dp)));
          ++f;
        }
        C[k][pos] = FC::sub(FC::zero(), lead);
      }
      C[dp][pos] = FC::one();
    }
    gam.assign(m, FC::zero());
  } else if (w->kind == ZS_WIRE_COEFFS) {
    std::vector<T> nat(npts);
    for (uint32_t i = 0; i < npts; ++i) nat[i] = FC::from_int(i);
    FMat<FC> vi;
    if (!inv_vandermonde<FC>(nat, vi)) return fail(ZS_EINVAL, "singular natural Vandermonde");
    W = vi;
    C.assign(m, std::vector<T>(m, FC::zero()));
    for (uint32_t k = 0; k < m; ++k) C[k][k] = FC::one();
    gam.assign(m, FC::zero());
  } else {  // ZS_WIRE_COMPRESSED: [c0, c2]; c1 = claim - 2 c0 - c2
    W[0][0] = FC::one();
    W[1] = natural_leading<FC>(d);
    C.assign(3, std::vector<T>(2, FC::zero()));
    C[0][0] = FC::one();
    C[1][0] = FC::from_int(-2);
    C[1][1] = FC::from_int(-1);
    C[2][1] = FC::one();
    gam = {FC::zero(), FC::one(), FC::zero()};
    c.claim_every = 1;
  }
  c.ncoef = (uint32_t)C.size();
  for (uint32_t j = 0; j < m; ++j)
    for (uint32_t i = 0; i < npts; ++i) c.wmat[j][i] = W[j][i];
  for (uint32_t k = 0; k < c.ncoef; ++k) {
    c.cclaim[k] = gam[k];
    for (uint32_t j = 0; j < m; ++j) c.cmat[k][j] = C[k][j];
  }
  return ZS_OK;
}

}  // namespace zs
