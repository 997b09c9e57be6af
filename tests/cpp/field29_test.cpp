// Host-side checks of the radix-2^29 Fq arithmetic and its group law
// (hyperplonk_amd/csrc/field29.hpp, curve29.hpp; the same source the MSM
// kernels compile), against the 32-bit-limb arithmetic of field.hpp /
// curve.hpp and against exact integer identities:
//   * pack / unpack round trips; conversions Fq <-> Fq29 round trip;
//   * products of lazy operands (up to 31 p) against the Fq product;
//   * add / add_twice / sub_kp / neg against Fq;
//   * is_zero_mod on every k p (k < 32) and on k p +- 1, canonical;
//   * mixed and general XYZZ additions, doublings and every exceptional case
//     (identity operands, P + P, P + (-P)) against curve.hpp, compared as
//     affine points;
//   * Fr29 on stored Fr values: prescaled products and product chains of lazy
//     operands against the Fr product, reduce_est, canonical, is_zero_mod,
//     inverse_fr against the 32-bit-limb inverse;
//   * the host finish's Fq inverse (host_g1.hpp, binary extended Euclid)
//     against the 32-bit-limb inverse, including 0, 1 and -1;
//   * the two-lane addition add_pair, its lanes played by two host threads,
//     limb for limb against add() (general, P + P, P + (-P), identities);
//   * the four-lane addition add_quad, its lanes played by four host threads,
//     against add() (X, ZZ, ZZZ limb for limb, Y modulo p and below 4 p),
//     also fed its own outputs.
// Prints "FIELD29 PASS" or the failing checks.
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <condition_variable>
#include <mutex>
#include <random>
#include <thread>

#include "../../hyperplonk_amd/csrc/curve29.hpp"
#include "../../hyperplonk_amd/csrc/host_g1.hpp"

using namespace hp;
using namespace hp::f29;

static int failures = 0;
#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);     \
      ++failures;                                                 \
    }                                                             \
  } while (0)

static std::mt19937_64 rng(2024);

static Fq rand_fq() {  // canonical, Montgomery form of a random value
  Fq x;
  for (;;) {
    for (int i = 0; i < 12; ++i) x.v[i] = (uint32_t)rng();
    x.v[11] &= 0x1fffffffu;
    if (!geq_modulus<FqParams>(x.v)) break;
  }
  return to_mont(x);
}

static Fq29 to29(const Fq& x) { return unpack(to29_packed(x).v); }
static Fq back(const Fq29& a) { return from29(a); }

// a + k p as limbs (a canonical, k small): a lazy representative of a
static Fq29 lazy(const Fq29& a, uint32_t k) {
  Fq29 kp = zero();
  uint64_t c = 0;
  for (int i = 0; i < L; ++i) {
    c += (uint64_t)k * FqP::P[i] + a.l[i];
    kp.l[i] = i < L - 1 ? ((uint32_t)c & kMask) : (uint32_t)c;
    c >>= kBits;
  }
  return kp;
}

static G1Affine rand_point(const G1Affine& g) {
  G1Xyzz acc = G1Xyzz::from_affine(g);
  const int steps = 5 + (int)(rng() % 40);
  for (int i = 0; i < steps; ++i) acc = (rng() & 1) ? dbl(acc) : add_mixed(acc, g);
  return to_affine(acc);
}

static bool same_point(const G1Xyzz& a, const G1Xyzz& b) {
  const G1Affine x = to_affine(a), y = to_affine(b);
  return x.x == y.x && x.y == y.y;
}

static Aff29 aff29(const G1Affine& a) { return {to29(a.x), to29(a.y)}; }

// Two host threads stand for the two lanes of add_pair: an exchange publishes
// the lane's value, waits for the partner's, and waits again before the slot
// is reused.
struct PairExchange {
  std::mutex mu;
  std::condition_variable cv;
  Fq29 slot[2];
  unsigned arrived = 0, generation = 0;
  void barrier() {
    std::unique_lock<std::mutex> lk(mu);
    const unsigned gen = generation;
    if (++arrived == 2) {
      arrived = 0;
      ++generation;
      cv.notify_all();
    } else {
      cv.wait(lk, [&] { return generation != gen; });
    }
  }
  Fq29 swap(unsigned half, const Fq29& v) {
    {
      std::lock_guard<std::mutex> g(mu);
      slot[half] = v;
    }
    barrier();
    Fq29 other;
    {
      std::lock_guard<std::mutex> g(mu);
      other = slot[half ^ 1u];
    }
    barrier();
    return other;
  }
};

static bool same_limbs(const Xyzz29& a, const Xyzz29& b) {
  for (int i = 0; i < L; ++i)
    if (a.X.l[i] != b.X.l[i] || a.Y.l[i] != b.Y.l[i] || a.ZZ.l[i] != b.ZZ.l[i] ||
        a.ZZZ.l[i] != b.ZZZ.l[i])
      return false;
  return true;
}

// Four host threads stand for the four lanes of a quad in add_quad: get<K>
// publishes every lane's value, waits, reads lane K's and waits again.
struct QuadExchange {
  std::mutex mu;
  std::condition_variable cv;
  Fq29 slot[4];
  unsigned arrived = 0, generation = 0;
  void barrier() {
    std::unique_lock<std::mutex> lk(mu);
    const unsigned gen = generation;
    if (++arrived == 4) {
      arrived = 0;
      ++generation;
      cv.notify_all();
    } else {
      cv.wait(lk, [&] { return generation != gen; });
    }
  }
};
struct QuadLane {
  QuadExchange* ex;
  unsigned lane;
  template <int K>
  Fq29 get(const Fq29& v) {
    {
      std::lock_guard<std::mutex> g(ex->mu);
      ex->slot[lane] = v;
    }
    ex->barrier();
    Fq29 r;
    {
      std::lock_guard<std::mutex> g(ex->mu);
      r = ex->slot[K];
    }
    ex->barrier();
    return r;
  }
};

// a < the normalised constant k (limbs compared from the top)
static bool below(const Fq29& a, const uint32_t* k) {
  for (int i = L - 1; i >= 0; --i)
    if (a.l[i] != k[i]) return a.l[i] < k[i];
  return false;
}

static bool same_fq(const Fq29& a, const Fq29& b) {
  for (int i = 0; i < L; ++i)
    if (a.l[i] != b.l[i]) return false;
  return true;
}

// add_quad on four threads: every lane returns add(p, q)'s X, ZZ and ZZZ limb
// for limb and the same Y modulo p (the sum of two reduced products below 4 p).
static bool quad_add_matches(const Xyzz29& p, const Xyzz29& q) {
  QuadExchange ex;
  Xyzz29 out[4];
  std::thread th[3];
  for (unsigned l = 1; l < 4; ++l)
    th[l - 1] = std::thread([&, l] {
      QuadLane bc{&ex, l};
      out[l] = add_quad(p, q, l, bc);
    });
  QuadLane bc0{&ex, 0};
  out[0] = add_quad(p, q, 0u, bc0);
  for (auto& t : th) t.join();
  const Xyzz29 want = f29::add(p, q);
  bool ok = true;
  for (int l = 0; l < 4; ++l) {
    ok = ok && same_fq(out[l].X, want.X) && same_fq(out[l].ZZ, want.ZZ) && same_fq(out[l].ZZZ, want.ZZZ);
    ok = ok && same_fq(canonical(out[l].Y), canonical(want.Y));
    ok = ok && below(out[l].Y, FqP::P4);  // the bound every formula assumes of Y
  }
  return ok;
}

// add_pair on two threads: both lanes must return add(p, q) limb for limb.
static bool pair_add_matches(const Xyzz29& p, const Xyzz29& q) {
  PairExchange ex;
  Xyzz29 out[2];
  std::thread t1([&] {
    out[1] = add_pair(p, q, 1u, [&](const Fq29& v) { return ex.swap(1u, v); });
  });
  out[0] = add_pair(p, q, 0u, [&](const Fq29& v) { return ex.swap(0u, v); });
  t1.join();
  const Xyzz29 want = f29::add(p, q);
  return same_limbs(out[0], want) && same_limbs(out[1], want);
}

int main() {
  // field
  for (int it = 0; it < 2000; ++it) {
    const Fq a = rand_fq(), b = rand_fq();
    const Fq29 a29 = to29(a), b29 = to29(b);
    CHECK(back(a29) == a);
    uint32_t w[12];
    pack(a29, w);
    Fq29 u = unpack(w);
    bool same = true;
    for (int i = 0; i < L; ++i) same = same && u.l[i] == a29.l[i];
    CHECK(same);
    const uint32_t ka = (uint32_t)(rng() % 16), kb = (uint32_t)(rng() % 16);
    CHECK(back(mul(lazy(a29, ka), lazy(b29, kb))) == a * b);
    CHECK(back(sqr(lazy(a29, ka))) == a * a);
    CHECK(back(add(lazy(a29, ka), lazy(b29, kb))) == a + b);
    CHECK(back(add_twice(lazy(a29, ka), b29)) == a + b + b);
    CHECK(back(sub2p(a29, lazy(b29, 1))) == a - b);
    CHECK(back(sub4p(lazy(a29, 3), lazy(b29, 3))) == a - b);
    CHECK(back(sub8p(a29, lazy(b29, 7))) == a - b);
    CHECK(back(sub16p(lazy(a29, 15), lazy(b29, 15))) == a - b);
    CHECK(back(neg_canonical(a29)) == neg(a));
    // product bound: below p (1 + 2^-15) for inputs below 32 p
    const Fq29 m = mul(lazy(a29, 31), lazy(b29, 31));
    CHECK(m.l[L - 1] <= FqP::P[L - 1]);
    // the squaring takes each cross product once: same integer as mul(x, x)
    const Fq29 x = lazy(a29, (uint32_t)(rng() % 32));
    const Fq29 s = sqr(x), sm = mul(x, x);
    bool same_sq = true;
    for (int i = 0; i < L; ++i) same_sq = same_sq && s.l[i] == sm.l[i];
    CHECK(same_sq);
    // one reduction for a sum of two products, and the same bound
    const Fq c = rand_fq(), d = rand_fq();
    const uint32_t kc = (uint32_t)(rng() % 32), kd = (uint32_t)(rng() % 32);
    const Fq29 ms = mul_sum(lazy(a29, ka + 16), lazy(b29, kb + 15), lazy(to29(c), kc),
                            lazy(to29(d), kd));
    CHECK(back(ms) == a * b + c * d);
    CHECK(ms.l[L - 1] <= FqP::P[L - 1]);
  }
  {
    // largest limbs: 2^29 - 1 below the top limb, top limb of 31 p
    Fq29 big = lazy(zero(), 31);
    for (int i = 0; i < L - 1; ++i) big.l[i] = kMask;
    const Fq29 s = sqr(big), sm = mul(big, big);
    bool same_sq = true;
    for (int i = 0; i < L; ++i) same_sq = same_sq && s.l[i] == sm.l[i];
    CHECK(same_sq);
    const Fq29 ms = mul_sum(big, big, big, big), twice = add(sm, sm);
    CHECK(back(ms) == back(twice));
  }
  for (int it = 0; it < 100; ++it) {
    const Fq a = rand_fq();
    CHECK(back(f29::inverse(lazy(to29(a), (uint32_t)(rng() % 8)))) == inverse(a));
  }
  {
    // host inverse on the same Montgomery bytes (R = 2^384 on 6 x 64 bits)
    auto host_inv_matches = [](const Fq& a) {
      host::Fq64 h;
      std::memcpy(h.v, a.v, 48);
      const host::Fq64 hi = host::inverse(h);
      const Fq want = inverse(a);
      return std::memcmp(hi.v, want.v, 48) == 0;
    };
    for (int it = 0; it < 300; ++it) CHECK(host_inv_matches(rand_fq()));
    Fq z;
    for (int i = 0; i < 12; ++i) z.v[i] = 0;
    CHECK(host_inv_matches(z));  // zero maps to zero
    const Fq one = Fq::one();  // Montgomery one
    CHECK(host_inv_matches(one));
    CHECK(host_inv_matches(neg(one)));
  }
  for (uint32_t k = 0; k < 32; ++k) {
    const Fq29 z = lazy(zero(), k);
    CHECK(is_zero_mod(z));
    Fq29 plus = z;
    plus.l[0] += 1;  // k p + 1 (no carry: low limb of k p is never 2^29 - 1 here)
    CHECK(plus.l[0] <= kMask && !is_zero_mod(plus));
    // same low limb as k p, another limb off: the full comparison decides
    for (int limb = 1; limb < L; ++limb) {
      Fq29 t = z;
      t.l[limb] ^= 1u;
      CHECK(!is_zero_mod(t));
    }
    if (k > 0) {
      const Fq29 one_raw = lazy(zero(), 0);
      Fq29 o = one_raw;
      o.l[0] = 1;
      CHECK(!is_zero_mod(sub_kp(z, o, FqP::P2)) || k == 0);  // k p - 1 + 2p
    }
    const Fq29 c = canonical(lazy(to29(rand_fq()), k));
    CHECK(!is_zero_mod(c) || all_zero(c));
  }
  for (int it = 0; it < 200; ++it) {
    const Fq a = rand_fq();
    const uint32_t k = (uint32_t)(rng() % 31);
    const Fq29 c = canonical(lazy(to29(a), k));
    bool same = true;
    const Fq29 ref = to29(a);
    for (int i = 0; i < L; ++i) same = same && c.l[i] == ref.l[i];
    CHECK(same);
    CHECK(!is_zero_mod(lazy(to29(a), k)));
  }
  // group law
  G1Affine g;
  {
    static const uint32_t gx[12] = {0xdb22c6bbu, 0xfb3af00au, 0xf97a1aefu, 0x6c55e83fu,
                                    0x171bac58u, 0xa14e3a3fu, 0x9774b905u, 0xc3688c4fu,
                                    0x4fa9ac0fu, 0x2695638cu, 0x3197d794u, 0x17f1d3a7u};
    static const uint32_t gy[12] = {0x46c5e7e1u, 0x0caa2329u, 0xa2888ae4u, 0xd03cc744u,
                                    0x2c04b3edu, 0x00db18cbu, 0xd5d00af6u, 0xfcf5e095u,
                                    0x741d8ae4u, 0xa09e30edu, 0xe3aaa0f1u, 0x08b3f481u};
    for (int i = 0; i < 12; ++i) {
      g.x.v[i] = gx[i];
      g.y.v[i] = gy[i];
    }
    g.x = to_mont(g.x);
    g.y = to_mont(g.y);
  }
  for (int it = 0; it < 200; ++it) {
    // a chain of mixed additions with repeated and negated points
    std::vector<G1Affine> pts;
    for (int i = 0; i < 6; ++i) pts.push_back(rand_point(g));
    G1Xyzz ref = G1Xyzz::identity();
    Xyzz29 acc = Xyzz29::identity();
    for (int s = 0; s < 24; ++s) {
      G1Affine q = pts[rng() % pts.size()];
      const int kind = (int)(rng() % 6);
      if (kind == 0 && !ref.is_identity()) q = to_affine(ref);         // P + P
      if (kind == 1 && !ref.is_identity()) q = neg(to_affine(ref));    // P + (-P)
      ref = add_mixed(ref, q);
      Aff29 q29 = aff29(q);
      if (rng() & 1) {  // the negated form used for negative digits
        q29.y = neg_canonical(to29(neg(q.y)));
      }
      acc = f29::add_mixed(acc, q29);
      CHECK(same_point(to_xyzz(acc), ref));
    }
    // general additions and doublings of accumulated points
    Xyzz29 b29 = Xyzz29::identity();
    G1Xyzz bref = G1Xyzz::identity();
    for (int s = 0; s < 6; ++s) {
      const G1Affine q = pts[rng() % pts.size()];
      bref = add_mixed(bref, q);
      b29 = f29::add_mixed(b29, aff29(q));
    }
    CHECK(same_point(to_xyzz(f29::add(acc, b29)), add(ref, bref)));
    CHECK(same_point(to_xyzz(f29::add(acc, acc)), add(ref, ref)));
    CHECK(same_point(to_xyzz(f29::dbl(acc)), dbl(ref)));
    CHECK(same_point(to_xyzz(f29::add(acc, Xyzz29::identity())), ref));
    CHECK(same_point(to_xyzz(f29::add(Xyzz29::identity(), b29)), bref));
    // P + (-P) in XYZZ form
    Xyzz29 nacc = acc;
    nacc.Y = sub4p(zero(), acc.Y);
    CHECK(f29::add(acc, nacc).is_identity() || ref.is_identity());
    // the two-lane addition: general, doubling, cancelling and identity cases
    if (it < 60) {
      CHECK(pair_add_matches(acc, b29));
      CHECK(pair_add_matches(acc, acc));
      CHECK(pair_add_matches(acc, nacc));
      CHECK(pair_add_matches(acc, Xyzz29::identity()));
      CHECK(pair_add_matches(Xyzz29::identity(), b29));
      // the four-lane addition, also on its own outputs (Y as a sum of two
      // reduced products) as the reductions chain it
      CHECK(quad_add_matches(acc, b29));
      CHECK(quad_add_matches(acc, acc));
      CHECK(quad_add_matches(acc, nacc));
      CHECK(quad_add_matches(acc, Xyzz29::identity()));
      CHECK(quad_add_matches(Xyzz29::identity(), b29));
      QuadExchange qx;
      Xyzz29 r[4];
      std::thread qt[3];
      for (unsigned l = 1; l < 4; ++l)
        qt[l - 1] = std::thread([&, l] {
          QuadLane bc{&qx, l};
          r[l] = add_quad(acc, b29, l, bc);
        });
      QuadLane bq{&qx, 0};
      r[0] = add_quad(acc, b29, 0u, bq);
      for (auto& t : qt) t.join();
      CHECK(quad_add_matches(r[0], acc));
      CHECK(quad_add_matches(r[0], r[0]));
      CHECK(quad_add_matches(b29, r[0]));
    }
  }
  // ---- Fr29: prescaled operands keep products in the stored form
  auto rand_fr = [] {
    Fr x;
    for (;;) {
      for (int i = 0; i < 8; ++i) x.v[i] = (uint32_t)rng();
      x.v[7] &= 0x7fffffffu;
      if (!geq_modulus<FrParams>(x.v)) break;
    }
    return x;  // any canonical value is a stored (Montgomery) form
  };
  auto lazy_r = [](const Fr29& a, uint32_t k) {
    Fr29 v;
    uint64_t c = 0;
    for (int i = 0; i < FrP::L; ++i) {
      c += (uint64_t)k * FrP::P[i] + a.l[i];
      v.l[i] = i < FrP::L - 1 ? ((uint32_t)c & kMask) : (uint32_t)c;
      c >>= kBits;
    }
    return v;
  };
  auto same_fr = [](const Fr& a, const Fr& b) { return a == b; };
  Fr k5;
  for (int i = 0; i < 8; ++i) k5.v[i] = FrConsts::K5[i];
  for (int it = 0; it < 2000; ++it) {
    const Fr c = rand_fr(), a = rand_fr(), b = rand_fr(), d = rand_fr();
    const uint32_t ka = (uint32_t)(rng() % 11), kb = (uint32_t)(rng() % 3);
    // one product: c 2^261 times a lazy stored value
    CHECK(same_fr(pack_fr(mul(prescale_fr(c), lazy_r(unpack_fr(a), ka))), c * a));
    // a chain of three factors: coefficient prescaled by 2^(5 * 3)
    const Fr c3 = ((c * k5) * k5) * k5;
    Fr29 t = mul(unpack_fr(c3), lazy_r(unpack_fr(a), ka));
    t = mul(lazy_r(t, kb), lazy_r(unpack_fr(b), ka));
    t = mul(t, lazy_r(unpack_fr(d), kb));
    CHECK(same_fr(pack_fr(t), ((c * a) * b) * d));
    // the fold f = lo + c (hi - lo)
    const Fr29 lo = unpack_fr(a), hi = unpack_fr(b);
    const Fr29 f = add(lo, mul(prescale_fr(c), sub1p(hi, lo)));
    CHECK(same_fr(pack_fr(f), a + c * (b - a)));
    // reduce_est: a representative below 3 r of values below 2^261 (up to
    // 70 r, past the 41 r the sumcheck accumulators reach)
    const uint32_t kk = (uint32_t)(rng() % 70);
    const Fr29 big = lazy_r(unpack_fr(a), kk);
    const Fr29 red = reduce_est(big);
    CHECK(same_fr(pack_fr(red), a));
    CHECK(cond_sub(cond_sub(cond_sub(red, FrP::P2), FrP::P), FrP::P).l[FrP::L - 1] ==
          canonical(red).l[FrP::L - 1]);
    CHECK(!is_zero_mod(big) || all_zero(unpack_fr(a)));
  }
  for (int it = 0; it < 200; ++it) {
    const Fr a = rand_fr();
    const Fr inv = inverse_fr(a);
    CHECK(inv == inverse(a));
    CHECK(a.is_zero() || a * inv == Fr::one());
  }
  CHECK(inverse_fr(Fr::zero()).is_zero());
  for (uint32_t k = 0; k < 32; ++k) {
    const Fr29 z = lazy_r(zero_t<FrP>(), k);
    CHECK(is_zero_mod(z));
    for (int limb = 1; limb < FrP::L; ++limb) {
      Fr29 t = z;
      t.l[limb] ^= 1u;
      CHECK(!is_zero_mod(t));
    }
    CHECK(all_zero(canonical(z)));
  }
  std::printf("FIELD29 %s\n", failures ? "FAIL" : "PASS");
  return failures ? 1 : 0;
}
