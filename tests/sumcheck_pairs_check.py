"""Whole sumchecks compared round by round with the oracle, run in a process of
its own (the round decomposition thresholds are read once per process):
under HP_MI355X_SC_ITEMS_BELOW=0 every round, down to the last single pair,
runs on a pair-per-thread kernel: k_round_classes<D> (degrees up to 6) or,
under HP_MI355X_SC_CLASSES=0 and above degree 6, k_round_pairs<D>.  Cases:
degrees 1..8 at num_vars 8, 10 and 12 with two product lists:
  * (c0, [f0..f_{D-1}]) + (c1, [f_D, f0]) + (c2, [f1]) as in the 64-sumcheck
    batch workload (factor repeats, a product of one factor, one product per
    class);
  * products sharing coefficients so that classes hold several products:
    (1, [f0, f1]), (1, [f2, f0]), (r - 1, [f1, f2]), (1, [f0..f_{D-1}]),
    (c, [f1]), (c, [f2]), (c, [f0]) (arity-1 and arity-2 classes of three,
    a negated coefficient in a class of its own), (c', f1^D) and (1, f2^D)
    (powers of one table: alone in a class and in the class of
    (1, [f0..f_{D-1}])).
Prints PAIRS PASS and exits 0 when every message matches
(cf. backend/src/cpu/sum_check.rs:74-180)."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main() -> int:
    from hyperplonk_amd import VirtualPolynomial
    from hyperplonk_amd.backend import Mi355xBackend
    from hyperplonk_amd.synth import SplitMix64Stream
    from oracle import oracle
    from tests import pyfield as pf

    be = Mi355xBackend(0)
    failures = 0
    cases = [(degree, nv, shared) for degree in range(1, 9) for nv in (8, 10, 12)
             for shared in (False, True)]
    for degree, nv, shared in cases:
        if True:
            rng = SplitMix64Stream(0x5041 + 97 * degree + nv)
            k = max(degree + 1, 3)
            tabs = [rng.fr_bytes(1 << nv) for _ in range(k)]
            crng = pf.SplitMix64(degree * 1000 + nv)
            prods = [(crng.nonboolean(), list(range(degree))), (crng.nonboolean(), [degree, 0]),
                     (crng.nonboolean(), [1 % k])]
            if degree == 1:
                prods = [(crng.nonboolean(), [0]), (crng.nonboolean(), [1])]
            if shared:
                c = crng.nonboolean()
                prods = [(1, list(range(degree))), (c, [1]), (c, [2]), (c, [0])]
                if degree >= 2:
                    prods += [(1, [0, 1]), (1, [2, 0]), (pf.R - 1, [1, 2])]
                    # powers of one table: alone in a class, and beside a
                    # product of distinct tables in the degree-D class
                    prods += [(crng.nonboolean(), [1] * degree), (1, [2] * degree)]
            mles = [be.mle_from_evaluations(nv, t) for t in tabs]
            vp = VirtualPolynomial(nv)
            for c, idx in prods:
                vp.add_mle_list([mles[i] for i in idx], c)
            assert vp.aux_info.max_degree == degree
            st = be.prover_init(vp)
            ost = oracle.SumCheckState(nv, degree, tabs, prods)
            ch = None
            for rnd in range(nv):
                got = be.prove_round_and_update_state(st, ch)
                want = ost.round(ch)
                if got != want:
                    print(f"MISMATCH degree {degree} nv {nv} round {rnd}", flush=True)
                    failures += 1
                    break
                ch = crng.nonboolean()
            print(f"degree {degree} nv {nv} shared {shared}: {'ok' if failures == 0 else 'FAIL'}",
                  flush=True)
    print("PAIRS PASS" if failures == 0 else "PAIRS FAIL", flush=True)
    return 1 if failures else 0


if __name__ == "__main__":
    sys.exit(main())
