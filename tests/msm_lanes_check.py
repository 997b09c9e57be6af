"""MSMs compared with the oracle's Pippenger, run in a process of its own (the
threads-per-lane switch of the latency-bound bucket stages is read once per
process): under HP_MI355X_MSM_QUAD=1 (the default) the bucket sums and the
level-1 reductions of batches that leave SIMDs idle run four threads per
addition (f29::add_quad), under HP_MI355X_MSM_QUAD=0 two (f29::add_pair).
Cases, each byte for byte against oracle.msm_pippenger_packets
(backend/src/cpu/multilinear_kzg.rs:40-59):
  * one multi_commit of mixed sizes (num_vars 0..14, a repeat): one batch with
    64-bucket reduction groups;
  * single commits at num_vars 10, 13 and 16 (the small-batch bucket sums and
    reductions) and 18 (the 256-bucket groups with two buckets per lane);
  * a commit whose scalars hit few buckets hard (every scalar one of eight
    values), so the bucket sums' heavy path and the identity cases of the
    additions run.
Prints LANES PASS and exits 0 when every commitment matches."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main() -> int:
    import numpy as np

    from hyperplonk_amd.backend import Mi355xBackend
    from hyperplonk_amd.fields import fr_from_bytes
    from hyperplonk_amd.synth import SplitMix64Stream
    from oracle import oracle

    be = Mi355xBackend(0)
    rng = SplitMix64Stream(0x1A4E5)
    N = 18
    t = fr_from_bytes(rng.fr_bytes(N).tobytes())
    pp = be.generate_prover_param(N, t)
    layers = {}

    def layer(nv):
        if nv not in layers:
            layers[nv] = be.download_layer(pp, N - nv)
        return layers[nv]

    failures = 0

    def check(name, nv, tab, got):
        nonlocal failures
        want = oracle.msm_pippenger_packets(layer(nv), tab.tobytes(), 1 << nv, 16)
        ok = got == want
        print(f"{name} nv={nv}: {'ok' if ok else 'MISMATCH'}", flush=True)
        failures += 0 if ok else 1

    nvs = [12, 0, 9, 3, 14, 6, 12, 1, 11]
    tabs = [rng.fr_bytes(1 << nv) for nv in nvs]
    mles = [be.mle_from_evaluations(nv, tb) for nv, tb in zip(nvs, tabs)]
    mles.append(mles[4])
    tabs.append(tabs[4])
    nvs.append(nvs[4])
    for k, (nv, tb, g) in enumerate(zip(nvs, tabs, be.multi_commit_packets(pp, mles))):
        check(f"multi_commit[{k}]", nv, tb, g)
    for nv in (10, 13, 16, 18):
        tb = rng.fr_bytes(1 << nv)
        check("commit", nv, tb, be.commit_packet(pp, be.mle_from_evaluations(nv, tb)))
    # eight distinct scalars over 2^14 entries: a few very full buckets
    nv = 14
    few = rng.fr_bytes(8)
    idx = np.arange(1 << nv) % 8
    tb = np.ascontiguousarray(few[idx])
    check("hot buckets", nv, tb, be.commit_packet(pp, be.mle_from_evaluations(nv, tb)))
    if failures:
        print(f"LANES FAIL ({failures})")
        return 1
    print("LANES PASS")
    return 0


if __name__ == "__main__":
    sys.exit(main())
