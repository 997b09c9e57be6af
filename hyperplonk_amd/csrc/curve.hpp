// BLS12-381 G1 (y^2 = x^3 + 4) arithmetic for the MSM kernels.
//
// Accumulators use XYZZ coordinates (x = X / ZZ, y = Y / ZZZ, ZZ^3 = ZZZ^2),
// whose mixed addition with an affine point costs 8M + 2S and needs no
// doubling-specific state; ZZ = 0 encodes the identity.  Affine bases keep
// (0, 0) as the identity marker, which is not on the curve.  Every
// exceptional case (identity operands, P == Q, P == -Q) is handled exactly, so
// any summation order produces the same group element and, after
// normalisation, the same affine bytes.
#pragma once
#include "field.hpp"

namespace hp {

struct G1Affine {
  Fq x, y;
  HP_HD bool is_identity() const { return x.is_zero() && y.is_zero(); }
  HP_HD static G1Affine identity() { return {Fq::zero(), Fq::zero()}; }
};

struct G1Xyzz {
  Fq X, Y, ZZ, ZZZ;
  HP_HD bool is_identity() const { return ZZ.is_zero(); }
  HP_HD static G1Xyzz identity() {
    return {Fq::one(), Fq::one(), Fq::zero(), Fq::zero()};
  }
  HP_HD static G1Xyzz from_affine(const G1Affine& a) {
    if (a.is_identity()) return identity();
    return {a.x, a.y, Fq::one(), Fq::one()};
  }
};

HP_HD G1Affine neg(const G1Affine& a) {
  if (a.is_identity()) return a;
  return {a.x, neg(a.y)};
}

// 2 * (x, y) for an affine point with y != 0 (mdbl-2008-s-1).
HP_HD G1Xyzz dbl_affine(const G1Affine& a) {
  Fq U = dbl(a.y);
  Fq V = sqr(U);
  Fq W = U * V;
  Fq S = a.x * V;
  Fq xx = sqr(a.x);
  Fq M = xx + xx + xx;
  Fq X3 = sqr(M) - dbl(S);
  Fq Y3 = M * (S - X3) - W * a.y;
  return {X3, Y3, V, W};
}

// 2 * P in XYZZ (dbl-2008-s-1 with a = 0).
HP_HD G1Xyzz dbl(const G1Xyzz& p) {
  if (p.is_identity() || p.Y.is_zero()) return G1Xyzz::identity();
  Fq U = dbl(p.Y);
  Fq V = sqr(U);
  Fq W = U * V;
  Fq S = p.X * V;
  Fq xx = sqr(p.X);
  Fq M = xx + xx + xx;
  Fq X3 = sqr(M) - dbl(S);
  Fq Y3 = M * (S - X3) - W * p.Y;
  return {X3, Y3, V * p.ZZ, W * p.ZZZ};
}

// P + Q with Q affine (madd-2008-s), exact in every exceptional case.
HP_HD G1Xyzz add_mixed(const G1Xyzz& p, const G1Affine& q) {
  if (q.is_identity()) return p;
  if (p.is_identity()) return G1Xyzz::from_affine(q);
  Fq U2 = q.x * p.ZZ;
  Fq S2 = q.y * p.ZZZ;
  Fq P = U2 - p.X;
  Fq R = S2 - p.Y;
  if (P.is_zero()) {
    if (R.is_zero()) return dbl_affine(q);
    return G1Xyzz::identity();
  }
  Fq PP = sqr(P);
  Fq PPP = P * PP;
  Fq Qv = p.X * PP;
  Fq X3 = sqr(R) - PPP - dbl(Qv);
  Fq Y3 = R * (Qv - X3) - p.Y * PPP;
  return {X3, Y3, p.ZZ * PP, p.ZZZ * PPP};
}

// P + Q, both XYZZ (add-2008-s).
HP_HD G1Xyzz add(const G1Xyzz& p, const G1Xyzz& q) {
  if (q.is_identity()) return p;
  if (p.is_identity()) return q;
  Fq U1 = p.X * q.ZZ;
  Fq U2 = q.X * p.ZZ;
  Fq S1 = p.Y * q.ZZZ;
  Fq S2 = q.Y * p.ZZZ;
  Fq P = U2 - U1;
  Fq R = S2 - S1;
  if (P.is_zero()) {
    if (R.is_zero()) return dbl(p);
    return G1Xyzz::identity();
  }
  Fq PP = sqr(P);
  Fq PPP = P * PP;
  Fq Qv = U1 * PP;
  Fq X3 = sqr(R) - PPP - dbl(Qv);
  Fq Y3 = R * (Qv - X3) - S1 * PPP;
  return {X3, Y3, p.ZZ * q.ZZ * PP, p.ZZZ * q.ZZZ * PPP};
}

HP_HD G1Xyzz neg(const G1Xyzz& p) { return {p.X, neg(p.Y), p.ZZ, p.ZZZ}; }

// Unique affine representative (identity -> (0, 0)).
HP_HD G1Affine to_affine(const G1Xyzz& p) {
  if (p.is_identity()) return G1Affine::identity();
  Fq inv = inverse(p.ZZ * p.ZZZ);  // 1 / (ZZ ZZZ)
  Fq zz_inv = inv * p.ZZZ;         // 1 / ZZ
  Fq zzz_inv = inv * p.ZZ;         // 1 / ZZZ
  return {p.X * zz_inv, p.Y * zzz_inv};
}

}  // namespace hp
