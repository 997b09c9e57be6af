"""Whole sumchecks compared round by round with the oracle, run in a process of
its own (the round decomposition thresholds are read once per process):
under HP_MI355X_SC_ITEMS_BELOW=0 every round, down to the last single pair,
runs on the pair-per-thread kernel, k_round_pairs<D>.  Cases: degrees 1..8
at num_vars 8, 10 and 12, products (c0, [f0..f_{D-1}]) + (c1, [f_D, f0]) +
(c2, [f1]) as in the 64-sumcheck batch workload, so the factor repeats and a
product of one factor are covered.  Prints PAIRS PASS and exits 0 when every
message matches (cf. backend/src/cpu/sum_check.rs:74-180)."""
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
    for degree in range(1, 9):
        for nv in (8, 10, 12):
            rng = SplitMix64Stream(0x5041 + 97 * degree + nv)
            k = degree + 1
            tabs = [rng.fr_bytes(1 << nv) for _ in range(k)]
            crng = pf.SplitMix64(degree * 1000 + nv)
            prods = [(crng.nonboolean(), list(range(degree))), (crng.nonboolean(), [degree, 0]),
                     (crng.nonboolean(), [1 % k])]
            if degree == 1:
                prods = [(crng.nonboolean(), [0]), (crng.nonboolean(), [1])]
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
            print(f"degree {degree} nv {nv}: {'ok' if failures == 0 else 'FAIL'}", flush=True)
    print("PAIRS PASS" if failures == 0 else "PAIRS FAIL", flush=True)
    return 1 if failures else 0


if __name__ == "__main__":
    sys.exit(main())
