"""Every decomposition of the sumcheck round kernels and every shape of the
batched MSM against the oracle, byte for byte.

* the pair-per-thread kernels for D = 1..8 (processes with
  HP_MI355X_SC_ITEMS_BELOW=0 and the round loops off, so every round runs on
  them; tests/sumcheck_pairs_check.py): k_round_classes<D> by default and
  k_round_pairs<D> with HP_MI355X_SC_CLASSES=0, and the class kernels replayed
  from round graphs (HP_MI355X_ROUND_GRAPHS=1); the same cases through the
  persistent round loops (wide loops of 256, 16 and 2 blocks, then fewer
  blocks, then one); and the 14-table degree-5 jellyfish-style gate of the
  n = 20 workload at num_vars 16 on the default thresholds (round 0 on the
  pairs kernel, then the loops), each round against oracle.SumCheckState
  (backend/src/cpu/sum_check.rs:74-180);
* multilinear_kzg_batch and multi_commit of mixed sizes (num_vars 0..16, with
  repeats) entry by entry against the oracle's Pippenger and KZG open
  (backend/src/cpu/multilinear_kzg.rs:40-137);
* the bucket sums and level-1 reductions on four and on two threads per
  addition (processes with HP_MI355X_MSM_QUAD=1 / 0; tests/msm_lanes_check.py)
  against the oracle's Pippenger;
* the 2^18 MSM against the oracle's Pippenger and the 2^22 MSM of the
  standalone MSM workload against the closed form g * f(t) (layer 0 of an SRS
  with trapdoor t holds g * eq(t, x), so the commitment of f is g * f(t);
  f(t) by the oracle's evaluation of the same bytes)."""
import os
import subprocess
import sys

import pytest

from hyperplonk_amd import VirtualPolynomial
from hyperplonk_amd.fields import fr_from_bytes
from tests import pyfield as pf

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("classes,graphs", [("1", "0"), ("0", "0"), ("1", "1")])
def test_pairs_kernel_every_degree_matches_oracle(classes, graphs):
    # round loops off: every round is a launched pair-per-thread round (with
    # graphs "1", replayed from hipGraphs that read their packs from device memory)
    env = dict(os.environ, HP_MI355X_SC_ITEMS_BELOW="0", HP_MI355X_SC_CLASSES=classes,
               HP_MI355X_SC_WIDE_ITEMS="0", HP_MI355X_SC_TAIL_SINGLE_ITEMS="0",
               HP_MI355X_ROUND_GRAPHS=graphs)
    res = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "tests", "sumcheck_pairs_check.py")],
                         capture_output=True, text=True, timeout=300, env=env, cwd=ROOT)
    print(res.stdout[-3000:])
    assert res.returncode == 0 and "PAIRS PASS" in res.stdout, res.stdout[-3000:] + res.stderr[-3000:]


@pytest.mark.parametrize("blocks", ["256", "16", "2"])
def test_round_loops_every_degree_match_oracle(blocks):
    """The same cases through the persistent round loops (the defaults: a wide
    loop of up to `blocks` blocks from round 0, then fewer blocks, then one):
    with 2 blocks every block takes its pairs in many LDS chunks."""
    env = dict(os.environ, HP_MI355X_SC_WIDE_BLOCKS=blocks)
    res = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "tests", "sumcheck_pairs_check.py")],
                         capture_output=True, text=True, timeout=300, env=env, cwd=ROOT)
    print(res.stdout[-3000:])
    assert res.returncode == 0 and "PAIRS PASS" in res.stdout, res.stdout[-3000:] + res.stderr[-3000:]


@pytest.mark.parametrize("quad", ["1", "0"])
def test_msm_bucket_lanes_match_oracle(quad):
    env = dict(os.environ, HP_MI355X_MSM_QUAD=quad)
    res = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "tests", "msm_lanes_check.py")],
                         capture_output=True, text=True, timeout=300, env=env, cwd=ROOT)
    print(res.stdout[-3000:])
    assert res.returncode == 0 and "LANES PASS" in res.stdout, res.stdout[-3000:] + res.stderr[-3000:]


def jellyfish_products(rng, ids):
    """Degree-5 gate over 5 wires and 9 selectors plus a scalar q_H w1^5 term."""
    R = pf.R
    q1, q2, q3, q4, qm1, qm2, qe, qo, qc, w1, w2, w3, w4, w5 = ids
    return [
        (1, [q1, w1]), (1, [q2, w2]), (1, [q3, w3]), (1, [q4, w4]), (1, [qm1, w1, w2]),
        (1, [qm2, w3, w4]), (1, [qe, w1, w2, w3, w4]), (R - 1, [qo, w5]), (1, [qc]),
        (rng.nonboolean(), [w1, w1, w1, w1, w1]),
    ]


def test_jellyfish_degree5_nv16_matches_oracle(backend, oracle, rng_factory):
    nv = 16
    rng = rng_factory(1616)
    tabs = [rng.fr_bytes(1 << nv) for _ in range(14)]
    prods = jellyfish_products(pf.SplitMix64(1617), list(range(14)))
    mles = [backend.mle_from_evaluations(nv, t) for t in tabs]
    vp = VirtualPolynomial(nv)
    for c, idx in prods:
        vp.add_mle_list([mles[i] for i in idx], c)
    assert vp.aux_info.max_degree == 5
    st = backend.prover_init(vp)
    ost = oracle.SumCheckState(nv, 5, tabs, prods)
    crng = pf.SplitMix64(1618)
    ch = None
    for rnd in range(nv):
        assert backend.prove_round_and_update_state(st, ch) == ost.round(ch), rnd
        ch = crng.nonboolean()


@pytest.fixture(scope="module")
def srs16(backend, rng_factory):
    rng = rng_factory(1600)
    N = 16
    t = fr_from_bytes(rng.fr_bytes(N).tobytes())
    pp = backend.generate_prover_param(N, t)
    layers = [backend.download_layer(pp, i) for i in range(N + 1)]
    return N, pp, layers


def test_kzg_batch_mixed_sizes_match_oracle(backend, oracle, rng_factory, srs16):
    N, pp, layers = srs16
    rng = rng_factory(1601)
    nvs = [0, 16, 1, 3, 12, 5, 7, 16, 9, 2, 14, 11, 6, 12]
    tabs = [rng.fr_bytes(1 << nv) for nv in nvs]
    mles = [backend.mle_from_evaluations(nv, t) for nv, t in zip(nvs, tabs)]
    ops, want = [], []
    for i, (nv, m, tab) in enumerate(zip(nvs, mles, tabs)):
        if i % 3 == 1:
            point = fr_from_bytes(rng.fr_bytes(nv).tobytes()) if nv else []
            ops.append(("open", m, point))
            opened = b"".join(layers[N - nv + 1 + j] for j in range(nv))
            want.append(oracle.kzg_open_packets(opened, tab, point, threads=16))
        else:
            ops.append(("commit", m))
            want.append(oracle.msm_pippenger_packets(layers[N - nv], tab.tobytes(), 1 << nv, 16))
    # repeats of a commit and of an open keep their places
    ops.append(ops[1])
    want.append(want[1])
    ops.append(ops[4])
    want.append(want[4])
    got = backend.kzg_batch(pp, ops)
    assert len(got) == len(ops)
    for k, (op, g, w) in enumerate(zip(ops, got, want)):
        if op[0] == "commit":
            assert g[0] == w, k
        else:
            assert (g[0], g[1]) == w, k


def test_multi_commit_mixed_sizes_match_oracle(backend, oracle, rng_factory, srs16):
    N, pp, layers = srs16
    rng = rng_factory(1602)
    nvs = [16, 0, 14, 3, 16, 10, 1, 12, 8]
    tabs = [rng.fr_bytes(1 << nv) for nv in nvs]
    mles = [backend.mle_from_evaluations(nv, t) for nv, t in zip(nvs, tabs)]
    mles.append(mles[2])
    tabs.append(tabs[2])
    nvs.append(nvs[2])
    got = backend.multi_commit_packets(pp, mles)
    for k, (nv, tab, g) in enumerate(zip(nvs, tabs, got)):
        assert g == oracle.msm_pippenger_packets(layers[N - nv], tab.tobytes(), 1 << nv, 16), k


def test_commit_2_18_matches_oracle(backend, oracle, rng_factory):
    rng = rng_factory(1800)
    N = 18
    t = fr_from_bytes(rng.fr_bytes(N).tobytes())
    pp = backend.generate_prover_param(N, t)
    tab = rng.fr_bytes(1 << N)
    m = backend.mle_from_evaluations(N, tab)
    layer = backend.download_layer(pp, 0)
    assert backend.commit_packet(pp, m) == oracle.msm_pippenger_packets(layer, tab.tobytes(), 1 << N, 16)


def test_commit_2_22_closed_form(backend, oracle, rng_factory):
    rng = rng_factory(2200)
    N = 22
    t = fr_from_bytes(rng.fr_bytes(N).tobytes())
    pp = backend.generate_prover_param(N, t)
    tab = rng.fr_bytes(1 << N)
    m = backend.mle_from_evaluations(N, tab)
    value = oracle.mle_evaluate(tab, t)
    assert backend.commit(pp, m) == pf.aff_mul(pf.G, value)
    # a smaller layer of the same SRS: layer k commits against eq(t[k..], .)
    tab20 = tab[: 1 << 20]
    m20 = backend.mle_from_evaluations(20, tab20)
    assert backend.commit(pp, m20) == pf.aff_mul(pf.G, oracle.mle_evaluate(tab20, t[2:]))
