/* Oracle (test infrastructure only): one prime field in Montgomery form over
 * 64-bit limbs.  Instantiated twice by field.c, once for BLS12-381 Fr (4 limbs)
 * and once for Fq (6 limbs), by defining FN (limb count) and FP (name prefix)
 * before inclusion.  Plain CIOS Montgomery multiplication with unsigned
 * __int128 products; inversion by Fermat exponentiation.  This restates the
 * arithmetic that ark-ff 0.6.0 provides to the CPU backend (see
 * backend/src/cpu/ (all .rs files), which only call field +,-,*,inverse) and is pinned by
 * Python-integer fixtures, not by reuse of any device code. */
#define FCAT2(a, b) a##_##b
#define FCAT(a, b) FCAT2(a, b)
#define F(name) FCAT(FP, name)

typedef struct { uint64_t v[FN]; } F(t);

typedef struct {
  uint64_t mod[FN];
  uint64_t inv;        /* -mod^{-1} mod 2^64 */
  F(t) r;              /* 2^(64 FN) mod p  (Montgomery one) */
  F(t) r2;             /* 2^(128 FN) mod p */
  uint64_t pm2[FN];    /* p - 2, exponent for inversion */
} F(params_t);

extern F(params_t) F(P);

static inline int F(geq_mod)(const uint64_t a[FN]) {
  for (int i = FN - 1; i >= 0; --i) {
    if (a[i] > F(P).mod[i]) return 1;
    if (a[i] < F(P).mod[i]) return 0;
  }
  return 1;
}

static inline void F(sub_mod_raw)(uint64_t a[FN]) {
  uint64_t borrow = 0;
  for (int i = 0; i < FN; ++i) {
    unsigned __int128 d = (unsigned __int128)a[i] - F(P).mod[i] - borrow;
    a[i] = (uint64_t)d;
    borrow = (uint64_t)(d >> 64) & 1;
  }
}

static inline void F(add)(F(t) *r, const F(t) *a, const F(t) *b) {
  uint64_t t[FN];
  uint64_t carry = 0;
  for (int i = 0; i < FN; ++i) {
    unsigned __int128 s = (unsigned __int128)a->v[i] + b->v[i] + carry;
    t[i] = (uint64_t)s;
    carry = (uint64_t)(s >> 64);
  }
  if (carry || F(geq_mod)(t)) F(sub_mod_raw)(t);
  memcpy(r->v, t, sizeof t);
}

static inline void F(sub)(F(t) *r, const F(t) *a, const F(t) *b) {
  uint64_t t[FN];
  uint64_t borrow = 0;
  for (int i = 0; i < FN; ++i) {
    unsigned __int128 d = (unsigned __int128)a->v[i] - b->v[i] - borrow;
    t[i] = (uint64_t)d;
    borrow = (uint64_t)(d >> 64) & 1;
  }
  if (borrow) {
    uint64_t carry = 0;
    for (int i = 0; i < FN; ++i) {
      unsigned __int128 s = (unsigned __int128)t[i] + F(P).mod[i] + carry;
      t[i] = (uint64_t)s;
      carry = (uint64_t)(s >> 64);
    }
  }
  memcpy(r->v, t, sizeof t);
}

static inline void F(neg)(F(t) *r, const F(t) *a) {
  F(t) z;
  memset(&z, 0, sizeof z);
  F(sub)(r, &z, a);
}

static inline int F(is_zero)(const F(t) *a) {
  uint64_t acc = 0;
  for (int i = 0; i < FN; ++i) acc |= a->v[i];
  return acc == 0;
}

static inline int F(eq)(const F(t) *a, const F(t) *b) {
  return memcmp(a->v, b->v, sizeof a->v) == 0;
}

/* CIOS Montgomery product a*b*2^(-64 FN) mod p. */
static inline void F(mul)(F(t) *r, const F(t) *a, const F(t) *b) {
  uint64_t t[FN + 2];
  memset(t, 0, sizeof t);
  for (int i = 0; i < FN; ++i) {
    uint64_t c = 0;
    for (int j = 0; j < FN; ++j) {
      unsigned __int128 s = (unsigned __int128)a->v[j] * b->v[i] + t[j] + c;
      t[j] = (uint64_t)s;
      c = (uint64_t)(s >> 64);
    }
    unsigned __int128 s = (unsigned __int128)t[FN] + c;
    t[FN] = (uint64_t)s;
    t[FN + 1] = (uint64_t)(s >> 64);
    uint64_t m = t[0] * F(P).inv;
    s = (unsigned __int128)m * F(P).mod[0] + t[0];
    c = (uint64_t)(s >> 64);
    for (int j = 1; j < FN; ++j) {
      s = (unsigned __int128)m * F(P).mod[j] + t[j] + c;
      t[j - 1] = (uint64_t)s;
      c = (uint64_t)(s >> 64);
    }
    s = (unsigned __int128)t[FN] + c;
    t[FN - 1] = (uint64_t)s;
    t[FN] = t[FN + 1] + (uint64_t)(s >> 64);
  }
  if (t[FN] || F(geq_mod)(t)) F(sub_mod_raw)(t);
  memcpy(r->v, t, sizeof r->v);
}

static inline void F(sqr)(F(t) *r, const F(t) *a) { F(mul)(r, a, a); }

static inline void F(one)(F(t) *r) { *r = F(P).r; }
static inline void F(zero)(F(t) *r) { memset(r, 0, sizeof *r); }

/* Canonical integer (limbs) -> Montgomery.  Returns 0 if non-canonical. */
static inline int F(from_canonical_limbs)(F(t) *r, const uint64_t in[FN]) {
  if (F(geq_mod)(in)) {
    memset(r, 0, sizeof *r);
    return 0;
  }
  F(t) x;
  memcpy(x.v, in, sizeof x.v);
  F(mul)(r, &x, &F(P).r2);
  return 1;
}

static inline void F(to_canonical_limbs)(uint64_t out[FN], const F(t) *a) {
  F(t) one;
  memset(&one, 0, sizeof one);
  one.v[0] = 1;
  F(t) x;
  F(mul)(&x, a, &one);
  memcpy(out, x.v, sizeof x.v);
}

static inline int F(from_bytes)(F(t) *r, const uint8_t *bytes) {
  uint64_t limbs[FN];
  for (int i = 0; i < FN; ++i) {
    uint64_t w = 0;
    for (int k = 7; k >= 0; --k) w = (w << 8) | bytes[8 * i + k];
    limbs[i] = w;
  }
  return F(from_canonical_limbs)(r, limbs);
}

static inline void F(to_bytes)(uint8_t *bytes, const F(t) *a) {
  uint64_t limbs[FN];
  F(to_canonical_limbs)(limbs, a);
  for (int i = 0; i < FN; ++i)
    for (int k = 0; k < 8; ++k) bytes[8 * i + k] = (uint8_t)(limbs[i] >> (8 * k));
}

static inline void F(from_u64)(F(t) *r, uint64_t x) {
  uint64_t limbs[FN];
  memset(limbs, 0, sizeof limbs);
  limbs[0] = x;
  F(from_canonical_limbs)(r, limbs);
}

static inline void F(pow_limbs)(F(t) *r, const F(t) *a, const uint64_t *e, int nlimbs) {
  F(t) acc = F(P).r;
  for (int i = nlimbs - 1; i >= 0; --i) {
    for (int b = 63; b >= 0; --b) {
      F(sqr)(&acc, &acc);
      if ((e[i] >> b) & 1) F(mul)(&acc, &acc, a);
    }
  }
  *r = acc;
}

/* Inverse by Fermat; the inverse of zero is reported as zero. */
static inline void F(inv)(F(t) *r, const F(t) *a) {
  F(pow_limbs)(r, a, F(P).pm2, FN);
}

/* Derive the Montgomery constants from the modulus alone. */
static inline void F(init_params)(const uint64_t mod[FN]) {
  memcpy(F(P).mod, mod, sizeof F(P).mod);
  /* inv = -mod^{-1} mod 2^64 by Newton iteration */
  uint64_t x = 1;
  for (int i = 0; i < 7; ++i) x *= 2 - mod[0] * x;
  F(P).inv = (uint64_t)0 - x;
  /* r = 2^(64FN) mod p by doubling 1 (64 FN) times; r2 by doubling further */
  F(t) acc;
  memset(&acc, 0, sizeof acc);
  acc.v[0] = 1;
  for (int i = 0; i < 64 * FN; ++i) F(add)(&acc, &acc, &acc);
  F(P).r = acc;
  for (int i = 0; i < 64 * FN; ++i) F(add)(&acc, &acc, &acc);
  F(P).r2 = acc;
  uint64_t borrow = 2;
  for (int i = 0; i < FN; ++i) {
    unsigned __int128 d = (unsigned __int128)mod[i] - borrow;
    F(P).pm2[i] = (uint64_t)d;
    borrow = (uint64_t)(d >> 64) & 1;
  }
}

#undef F
#undef FCAT
#undef FCAT2
