// G1 group law of the MSM bucket stages on Fq29 (lazy reduction).
//
// XYZZ coordinates as in curve.hpp (x = X / ZZ, y = Y / ZZZ, ZZ^3 = ZZZ^2,
// identity: ZZ = 0 with all-zero limbs).  Every operation keeps
//   X < 10 p,  Y < 4 p,  ZZ, ZZZ < 2 p   (normalised limbs),
// which fixes the multiple of p each subtraction adds:
//   madd-2008-s:  P = U2 + 16p - X1 (< 18p), R = S2 + 4p - Y1 (< 6p),
//                 X3 = R^2 + 8p - (PPP + 2Q) (< 10p),
//                 Y3 = R (Q + 16p - X3) + (4p - Y1) PPP;
//   add-2008-s:   P = U2 + 2p - U1, R = S2 + 2p - S1 (< 4p), X3 as above,
//                 Y3 = R (Q + 16p - X3) + (2p - S1) PPP;
//   dbl-2008-s-1: M = 3 X^2 (< 6p), X3 = M^2 + 4p - 2S (< 6p),
//                 Y3 = M (S + 16p - X3) + (4p - Y) W.
// Products stay below p (1 + 2^-15) for inputs below 32 p; each Y3 is one
// sum of two products with a single reduction (mul_sum), below p (1 + 2^-18).
// Exceptional cases (identity operands, P == Q, P == -Q) are exact, tested
// with is_zero_mod, so the group element does not depend on the order of
// accumulation.
#pragma once
#include "curve.hpp"
#include "field29.hpp"

namespace hp {
namespace f29 {

struct Xyzz29 {
  Fq29 X, Y, ZZ, ZZZ;
  HP_HD bool is_identity() const { return all_zero(ZZ); }
  HP_HD static Xyzz29 identity() { return {one(), one(), zero(), zero()}; }
};

struct Aff29 {
  Fq29 x, y;
};

HP_HD Xyzz29 from_affine(const Aff29& a) { return {a.x, a.y, one(), one()}; }

// 2 (x, y), y != 0 (mdbl-2008-s-1)
HP_HD Xyzz29 dbl_affine(const Aff29& a) {
  const Fq29 U = add(a.y, a.y);
  const Fq29 V = sqr(U);
  const Fq29 W = mul(U, V);
  const Fq29 S = mul(a.x, V);
  const Fq29 xx = sqr(a.x);
  const Fq29 M = add_twice(xx, xx);
  const Fq29 X3 = sub4p(sqr(M), add(S, S));
  const Fq29 Y3 = mul_sum(M, sub16p(S, X3), sub2p(zero(), a.y), W);
  return {X3, Y3, V, W};
}

HP_HD Xyzz29 dbl(const Xyzz29& p) {
  if (p.is_identity() || is_zero_mod(p.Y)) return Xyzz29::identity();
  const Fq29 U = add(p.Y, p.Y);
  const Fq29 V = sqr(U);
  const Fq29 W = mul(U, V);
  const Fq29 S = mul(p.X, V);
  const Fq29 xx = sqr(p.X);
  const Fq29 M = add_twice(xx, xx);
  const Fq29 X3 = sub4p(sqr(M), add(S, S));
  const Fq29 Y3 = mul_sum(M, sub16p(S, X3), sub4p(zero(), p.Y), W);
  return {X3, Y3, mul(V, p.ZZ), mul(W, p.ZZZ)};
}

// P + Q, Q affine and finite with canonical coordinates (madd-2008-s)
HP_HD Xyzz29 add_mixed(const Xyzz29& p, const Aff29& q) {
  if (p.is_identity()) return from_affine(q);
  const Fq29 U2 = mul(q.x, p.ZZ);
  const Fq29 S2 = mul(q.y, p.ZZZ);
  const Fq29 P = sub16p(U2, p.X);
  const Fq29 R = sub4p(S2, p.Y);
  if (is_zero_mod(P)) {
    if (is_zero_mod(R)) return dbl_affine(q);
    return Xyzz29::identity();
  }
  const Fq29 PP = sqr(P);
  const Fq29 PPP = mul(P, PP);
  const Fq29 Qv = mul(p.X, PP);
  const Fq29 X3 = sub8p(sqr(R), add_twice(PPP, Qv));
  const Fq29 Y3 = mul_sum(R, sub16p(Qv, X3), sub4p(zero(), p.Y), PPP);
  return {X3, Y3, mul(p.ZZ, PP), mul(p.ZZZ, PPP)};
}

// P + Q, both XYZZ (add-2008-s)
HP_HD Xyzz29 add(const Xyzz29& p, const Xyzz29& q) {
  if (q.is_identity()) return p;
  if (p.is_identity()) return q;
  const Fq29 U1 = mul(p.X, q.ZZ);
  const Fq29 U2 = mul(q.X, p.ZZ);
  const Fq29 S1 = mul(p.Y, q.ZZZ);
  const Fq29 S2 = mul(q.Y, p.ZZZ);
  const Fq29 P = sub2p(U2, U1);
  const Fq29 R = sub2p(S2, S1);
  if (is_zero_mod(P)) {
    if (is_zero_mod(R)) return dbl(p);
    return Xyzz29::identity();
  }
  const Fq29 PP = sqr(P);
  const Fq29 PPP = mul(P, PP);
  const Fq29 Qv = mul(U1, PP);
  const Fq29 X3 = sub8p(sqr(R), add_twice(PPP, Qv));
  const Fq29 Y3 = mul_sum(R, sub16p(Qv, X3), sub2p(zero(), S1), PPP);
  return {X3, Y3, mul(mul(p.ZZ, q.ZZ), PP), mul(mul(p.ZZZ, q.ZZZ), PPP)};
}

// P + Q by two lanes (half = 0, 1) that both hold P and Q: each lane does
// half of the independent products and `ex(v)` returns the partner lane's v
// (every exchange is reached by both lanes of the pair together).  Per lane:
// 5 products, 1 squaring and the Y3 sum against 12 products and 2 squarings
// for add(); both lanes return the same point, equal to add(P, Q) as a group
// element.  For latency-bound reductions whose wavefronts leave SIMDs idle.
//   stage 1: lane 0 U1 = X1 ZZ2, S1 = Y1 ZZZ2, ZZ12 = ZZ1 ZZ2;
//            lane 1 U2 = X2 ZZ1, S2 = Y2 ZZZ1, ZZZ12 = ZZZ1 ZZZ2
//   stage 2: lane 0 PP = P^2, lane 1 RR = R^2
//   stage 3: lane 0 Q = U1 PP, ZZ3 = ZZ12 PP; lane 1 PPP = P PP, ZZZ3 = ZZZ12 PPP
//   stage 4: both X3 and Y3 (the bounds of add()).
template <class Ex>
HP_HD Xyzz29 add_pair(const Xyzz29& p, const Xyzz29& q, uint32_t half, Ex ex) {
  if (q.is_identity()) return p;
  if (p.is_identity()) return q;
  const bool h = half != 0;
  const Fq29 m1 = mul(h ? q.X : p.X, h ? p.ZZ : q.ZZ);     // U1 | U2
  const Fq29 m2 = mul(h ? q.Y : p.Y, h ? p.ZZZ : q.ZZZ);   // S1 | S2
  const Fq29 m3 = mul(h ? p.ZZZ : p.ZZ, h ? q.ZZZ : q.ZZ); // ZZ12 | ZZZ12
  const Fq29 o1 = ex(m1), o2 = ex(m2);
  const Fq29 U1 = h ? o1 : m1, U2 = h ? m1 : o1;
  const Fq29 S1 = h ? o2 : m2, S2 = h ? m2 : o2;
  const Fq29 P = sub2p(U2, U1);
  const Fq29 R = sub2p(S2, S1);
  if (is_zero_mod(P)) {
    if (is_zero_mod(R)) return dbl(p);
    return Xyzz29::identity();
  }
  const Fq29 m4 = sqr(h ? R : P);  // PP | RR
  const Fq29 o4 = ex(m4);
  const Fq29 PP = h ? o4 : m4, RR = h ? m4 : o4;
  const Fq29 m5 = mul(h ? P : U1, PP);  // Q | PPP
  const Fq29 o5 = ex(m5);
  const Fq29 Qv = h ? o5 : m5, PPP = h ? m5 : o5;
  const Fq29 m6 = mul(m3, h ? PPP : PP);  // ZZ3 | ZZZ3
  const Fq29 o6 = ex(m6);
  const Fq29 X3 = sub8p(RR, add_twice(PPP, Qv));
  const Fq29 Y3 = mul_sum(R, sub16p(Qv, X3), sub2p(zero(), S1), PPP);
  return {X3, Y3, h ? o6 : m6, h ? m6 : o6};
}

// Per-lane choice among four values by the lane's index q in its quad.
HP_HD Fq29 pick4(uint32_t q, const Fq29& a0, const Fq29& a1, const Fq29& a2, const Fq29& a3) {
  Fq29 r;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const uint32_t lo = (q & 1u) ? a1.l[i] : a0.l[i];
    const uint32_t hi = (q & 1u) ? a3.l[i] : a2.l[i];
    r.l[i] = (q & 2u) ? hi : lo;
  }
  return r;
}

// P + Q by the four lanes q = 0..3 of a quad that all hold P and Q, one
// product per lane and stage; `bc.template get<K>(v)` returns lane K's v
// (every broadcast is reached by the four lanes together: P and Q, hence
// every branch, are the same in all of them).  Per lane: four products
// against add()'s 12 products and 2 squarings, so for latency-bound
// reductions a dependent addition costs about a third of add()'s chain.
//   stage 1: U1 = X1 ZZ2 | U2 = X2 ZZ1 | S1 = Y1 ZZZ2 | S2 = Y2 ZZZ1
//   stage 2: PP = P P    | RR = R R    | ZZ12 = ZZ1 ZZ2 | ZZZ12 = ZZZ1 ZZZ2
//   stage 3: PPP = P PP  | Q = U1 PP   | ZZ3 = ZZ12 PP  | (PPP again)
//   stage 4: R (Q + 16p - X3) | ZZZ3 = ZZZ12 PPP | (2p - S1) PPP | (lane 0's)
// X3, ZZ3, ZZZ3 equal add()'s limb for limb (the squarings as products give
// the same integers); Y3 is the sum of two reduced products instead of
// add()'s single reduction of their sum: the same field element, below
// 2 p (1 + 2^-15) < 4 p, so the bounds of every formula hold.
template <class Bc>
HP_HD Xyzz29 add_quad(const Xyzz29& p, const Xyzz29& q, uint32_t lane, Bc& bc) {
  if (q.is_identity()) return p;
  if (p.is_identity()) return q;
  const Fq29 m1 = mul(pick4(lane, p.X, q.X, p.Y, q.Y), pick4(lane, q.ZZ, p.ZZ, q.ZZZ, p.ZZZ));
  const Fq29 U1 = bc.template get<0>(m1), U2 = bc.template get<1>(m1);
  const Fq29 S1 = bc.template get<2>(m1), S2 = bc.template get<3>(m1);
  const Fq29 P = sub2p(U2, U1);
  const Fq29 R = sub2p(S2, S1);
  if (is_zero_mod(P)) {
    if (is_zero_mod(R)) return dbl(p);
    return Xyzz29::identity();
  }
  const Fq29 m2 = mul(pick4(lane, P, R, p.ZZ, p.ZZZ), pick4(lane, P, R, q.ZZ, q.ZZZ));
  const Fq29 PP = bc.template get<0>(m2), RR = bc.template get<1>(m2);
  const Fq29 ZZ12 = bc.template get<2>(m2), ZZZ12 = bc.template get<3>(m2);
  const Fq29 m3 = mul(pick4(lane, P, U1, ZZ12, P), PP);
  const Fq29 PPP = bc.template get<0>(m3), Qv = bc.template get<1>(m3);
  const Fq29 ZZ3 = bc.template get<2>(m3);
  const Fq29 X3 = sub8p(RR, add_twice(PPP, Qv));
  const Fq29 m4 = mul(pick4(lane, R, ZZZ12, sub2p(zero(), S1), R),
                      pick4(lane, sub16p(Qv, X3), PPP, PPP, sub16p(Qv, X3)));
  const Fq29 Y3 = add(bc.template get<0>(m4), bc.template get<2>(m4));
  return {X3, Y3, ZZ3, bc.template get<1>(m4)};
}

// a^(p - 2) (zero maps to zero); the one inversion per thread of the batch
// normalisations
HP_HD Fq29 inverse(const Fq29& a) {
  uint32_t exp[12];  // p - 2
  uint32_t borrow = 2;
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    const uint64_t d = (uint64_t)FqParams::MOD[i] - borrow;
    exp[i] = (uint32_t)d;
    borrow = (uint32_t)(d >> 63);
  }
  Fq29 acc = one();
  for (int i = 11; i >= 0; --i)
    for (int b = 31; b >= 0; --b) {
      acc = sqr(acc);
      if ((exp[i] >> b) & 1u) acc = mul(acc, a);
    }
  return acc;
}

// Conversions to the 32-bit-limb XYZZ form (identity: ZZ = ZZZ = 0).
HP_HD G1Xyzz to_xyzz(const Xyzz29& p) {
  if (p.is_identity()) return G1Xyzz::identity();
  return {from29(p.X), from29(p.Y), from29(p.ZZ), from29(p.ZZZ)};
}

}  // namespace f29
}  // namespace hp
