"""SumCheck prover state on the GPU, round by round, byte-exact against the
oracle; cases follow the reference's own tests (backend/src/cpu/sum_check.rs:224-253,
backend/src/ascend910b3/native/tests/sumcheck_tensor_edge.cpp:164-187) plus the
degree-5 jellyfish-style gate of the n = 20 workload at a smaller size."""
import ctypes

import pytest

from hyperplonk_amd import InvalidParameters, InvalidState, VirtualPolynomial, abi
from hyperplonk_amd.fields import R_MODULUS as R
from hyperplonk_amd.fields import fr_from_bytes
from tests import pyfield as pf

pytestmark = pytest.mark.gpu


def test_round_lifecycle_known_answer(backend):
    # backend/src/cpu/sum_check.rs:225-253
    f = backend.mle_from_evaluations(2, [2, 3, 5, 7])
    vp = VirtualPolynomial.new_from_mle(f, 2)
    vp.add_mle_list([f, f], 3)
    st = backend.prover_init(vp)
    with pytest.raises(InvalidState):
        backend.prove_round_and_update_state(st, 2)
    assert backend.prove_round_and_update_state(st, None) == [101, 194, 317]
    with pytest.raises(InvalidState):
        backend.prove_round_and_update_state(st, None)
    assert backend.prove_round_and_update_state(st, 2) == [56, 261, 616]
    with pytest.raises(InvalidState):
        backend.prove_round_and_update_state(st, 2)
    assert backend.mle_values(f) == [2, 3, 5, 7]


EDGE_CASES = [
    ("n1_d2", 1, 2, 0), ("n2_d2", 2, 2, 0), ("n3_multi", 3, 2, 1), ("n9_tail", 9, 2, 0),
    ("n12_d2", 12, 2, 0), ("n12_multi", 12, 2, 1), ("n9_mixed", 9, 2, 2), ("n9_d3", 9, 3, 3),
    ("n2_empty", 2, 2, 4),
]


def edge_inputs(n, kind, seed):
    rng = pf.SplitMix64(seed)
    tables = [[rng.field() for _ in range(1 << n)] for _ in range(2)]
    if n > 1:
        tables[0][0], tables[0][1] = 0, R - 1
        tables[1][0], tables[1][1] = R - 1, 1
    products = {
        0: [(rng.nonboolean(), [2, 1])],
        1: [(rng.nonboolean(), [2, 1]), (R - 1, [0, 2]), (rng.nonboolean(), [1, 1])],
        2: [(3, [0]), (7, [2, 1])],
        3: [(rng.nonboolean(), [0, 1, 2]), (R - 1, [1])],
        4: [],
    }[kind]
    return tables, products, rng


def raw_state(backend, n, degree, handles, products):
    """Create a state through the C ABI with handle list `handles` (repeats allowed)."""
    coeffs = b"".join(c.to_bytes(32, "little") for c, _ in products)
    offsets, idx = [0], []
    for _, ids in products:
        idx.extend(ids)
        offsets.append(len(idx))
    out = ctypes.c_void_p()
    st = backend.lib.hp_mi355x_v1_sumcheck_prover_state_create(
        backend.context, n, degree, abi.ptr_array(handles), len(handles), coeffs or None,
        abi.size_array(offsets), len(products), abi.size_array(idx), len(idx), ctypes.byref(out),
    )
    return st, out.value


def round_raw(backend, state, challenge, count):
    buf = ctypes.create_string_buffer(b"\xa7" * (32 * count), 32 * count)
    ch = challenge.to_bytes(32, "little") if isinstance(challenge, int) else challenge
    st = backend.lib.hp_mi355x_v1_sumcheck_prover_state_round(backend.context, state, ch, buf, count)
    return st, buf.raw


@pytest.mark.parametrize("name,n,degree,kind", EDGE_CASES)
def test_edge_cases_round_by_round(backend, oracle, name, n, degree, kind):
    tables, products, rng = edge_inputs(n, kind, 0x534354454E534F52 + 17 * n + kind)
    a = backend.mle_from_evaluations(n, tables[0])
    b = backend.mle_from_evaluations(n, tables[1])
    st, state = raw_state(backend, n, degree, [a.address, b.address, a.address], products)
    assert st.kind == 0, name
    # the state owns its inputs: retire the originals before round zero
    del a, b
    ost = oracle.SumCheckState(n, degree, tables + [tables[0]], products)
    ch = rng.nonboolean()
    # premature challenge / wrong count / null output leave the state untouched
    s, out = round_raw(backend, state, ch, degree + 1)
    assert s.kind == abi.INVALID_STATE and out == b"\xa7" * len(out)
    s, _ = round_raw(backend, state, None, degree)
    assert s.kind == abi.INVALID_PARAMETERS
    s = backend.lib.hp_mi355x_v1_sumcheck_prover_state_round(backend.context, state, None, None, degree + 1)
    assert s.kind == abi.INVALID_PARAMETERS
    for rnd in range(n):
        c = None
        if rnd:
            s, _ = round_raw(backend, state, None, degree + 1)
            assert s.kind == abi.INVALID_STATE
            s, _ = round_raw(backend, state, b"\xff" * 32, degree + 1)
            assert s.kind == abi.INVALID_PARAMETERS
            c = rng.nonboolean()
        s, out = round_raw(backend, state, c, degree + 1)
        assert s.kind == 0, (name, rnd)
        assert fr_from_bytes(out) == ost.round(c), (name, rnd)
    s, _ = round_raw(backend, state, 5, degree + 1)
    assert s.kind == abi.INVALID_STATE
    assert backend.lib.hp_mi355x_v1_sumcheck_prover_state_destroy(backend.context, state).kind == 0


def test_create_rejects_bad_metadata(backend):
    a = backend.mle_from_evaluations(2, [1, 2, 3, 4])
    h = [a.address]
    for n, d, prods in ((0, 2, [(1, [0])]), (2, 0, [(1, [0])]), (2, 1, [(1, [0, 0])]), (2, 2, [(1, [1])])):
        st, _ = raw_state(backend, n, d, h, prods)
        assert st.kind == abi.INVALID_PARAMETERS
    st, _ = raw_state(backend, 3, 2, h, [(1, [0])])  # dimension mismatch
    assert st.kind == abi.INVALID_PARAMETERS
    st, _ = raw_state(backend, 2, 2, h, [(R, [0])])  # non-canonical coefficient
    assert st.kind == abi.INVALID_PARAMETERS


def run_vp(backend, oracle, nv, degree, tables, products, seed, check_every=True):
    mles = [backend.mle_from_evaluations(nv, t) for t in tables]
    vp = VirtualPolynomial(nv)
    for c, idx in products:
        vp.add_mle_list([mles[i] for i in idx], c)
    assert vp.aux_info.max_degree == degree
    st = backend.prover_init(vp)
    ints = [fr_from_bytes(t.tobytes()) for t in tables] if check_every else None
    ost = oracle.SumCheckState(nv, degree, ints, products) if check_every else None
    rng = pf.SplitMix64(seed)
    c = None
    msgs = []
    for _ in range(nv):
        got = backend.prove_round_and_update_state(st, c)
        if check_every:
            assert got == ost.round(c)
        msgs.append((c, got))
        c = rng.nonboolean()
    return msgs, c, mles


def jellyfish_products(rng, ids):
    """Degree-5 gate over 5 wires and 9 selectors plus a scalar q_H w1^5 term."""
    q1, q2, q3, q4, qm1, qm2, qe, qo, qc, w1, w2, w3, w4, w5 = ids
    return [
        (1, [q1, w1]), (1, [q2, w2]), (1, [q3, w3]), (1, [q4, w4]), (1, [qm1, w1, w2]),
        (1, [qm2, w3, w4]), (1, [qe, w1, w2, w3, w4]), (R - 1, [qo, w5]), (1, [qc]),
        (rng.nonboolean(), [w1, w1, w1, w1, w1]),
    ]


@pytest.mark.parametrize("nv", [6, 13])
def test_degree2_and_degree5_match_oracle(backend, oracle, rng_factory, nv):
    rng = rng_factory(300 + nv)
    tabs = [rng.fr_bytes(1 << nv) for _ in range(6)]
    prods = [(fr_from_bytes(rng.fr_bytes(1).tobytes())[0], [2 * i, 2 * i + 1]) for i in range(3)]
    run_vp(backend, oracle, nv, 2, tabs, prods, 11)
    tabs = [rng.fr_bytes(1 << nv) for _ in range(14)]
    run_vp(backend, oracle, nv, 5, tabs, jellyfish_products(pf.SplitMix64(5), list(range(14))), 12)


@pytest.mark.parametrize("degree,nprod,nmle", [(5, 50, 20), (5, 64, 12), (2, 64, 30), (8, 40, 9),
                                               (1, 3, 3)])
def test_many_products_match_oracle(backend, oracle, rng_factory, degree, nprod, nmle):
    """Block shapes of the smallest-round kernel: 50 and 64 degree-5 products
    span two blocks per pair (42 slots per point), 64 degree-2 products leave
    one pair per block, degree 8 leaves 28 slots, degree 1 with 3 products
    packs 42 pairs per block.  Random arities 1..degree, repeated factors,
    shared and negated coefficients."""
    nv = 7
    rng = rng_factory(500 + degree * 100 + nprod)
    tabs = [rng.fr_bytes(1 << nv) for _ in range(nmle)]
    crng = pf.SplitMix64(900 + nprod)
    prods = []
    for i in range(nprod):
        m = degree if i == 0 else 1 + crng.next() % degree
        ids = [crng.next() % nmle for _ in range(m)]
        c = (1, R - 1, crng.nonboolean())[crng.next() % 3]
        prods.append((c, ids))
    run_vp(backend, oracle, nv, degree, tabs, prods, 17)


def test_unreferenced_factors_match_oracle(backend, oracle, rng_factory):
    """40 distinct factor handles of which two products use three: the
    smallest-round kernel would need 40 x 85 (factor, pair) lines per block in
    LDS, so those rounds take the per-(product, pair) kernel; messages equal
    the oracle's over the same tables."""
    nv = 8
    rng = rng_factory(777)
    tabs = [rng.fr_bytes(1 << nv) for _ in range(40)]
    mles = [backend.mle_from_evaluations(nv, t) for t in tabs]
    products = [(3, [5, 17]), (R - 1, [39])]
    st, state = raw_state(backend, nv, 2, [m.address for m in mles], products)
    assert st.kind == 0
    ost = oracle.SumCheckState(nv, 2, [fr_from_bytes(t.tobytes()) for t in tabs], products)
    crng = pf.SplitMix64(778)
    c = None
    try:
        for _ in range(nv):
            st, raw = round_raw(backend, state, c, 3)
            assert st.kind == 0
            assert fr_from_bytes(raw) == ost.round(c)
            c = crng.nonboolean()
    finally:
        backend.lib.hp_mi355x_v1_sumcheck_prover_state_destroy(backend.context, state)


@pytest.mark.parametrize("degree", [7, 9, 34])
def test_high_degree_products(backend, oracle, rng_factory, degree):
    # degrees beyond the register-resident kernel (cf. sum_check/mod.rs:663-914 degrees {2,3,7,34})
    nv = 5
    rng = rng_factory(400 + degree)
    tabs = [rng.fr_bytes(1 << nv) for _ in range(3)]
    prods = [(3, [i % 3 for i in range(degree)]), (5, [1, 2]), (R - 2, [0])]
    run_vp(backend, oracle, nv, degree, tabs, prods, 13)


@pytest.mark.parametrize("degree", [2, 5])
def test_n20_invariants(backend, rng_factory, degree):
    """Full-size (n = 20) check through size-independent properties: each
    round's g(0) + g(1) equals the previous round polynomial at the challenge,
    and the final polynomial at the last challenge equals the virtual
    polynomial evaluated at the challenge point (per-MLE device evaluations)."""
    nv = 20
    rng = rng_factory(500 + degree)
    k = 6 if degree == 2 else 14
    tabs = [rng.fr_bytes(1 << nv) for _ in range(k)]
    if degree == 2:
        prods = [(fr_from_bytes(rng.fr_bytes(1).tobytes())[0], [2 * i, 2 * i + 1]) for i in range(3)]
    else:
        prods = jellyfish_products(pf.SplitMix64(9), list(range(14)))
    mles = [backend.mle_from_evaluations(nv, t) for t in tabs]
    vp = VirtualPolynomial(nv)
    for c, idx in prods:
        vp.add_mle_list([mles[i] for i in idx], c)
    st = backend.prover_init(vp)
    srng = pf.SplitMix64(77)
    prev, point = None, []
    c = None
    for rnd in range(nv):
        got = backend.prove_round_and_update_state(st, c)
        assert len(got) == degree + 1
        if prev is not None:
            assert (got[0] + got[1]) % R == interpolate(prev, c)
        prev = got
        c = srng.nonboolean()
        point.append(c)
    evals = [backend.evaluate_mle(m, point) for m in mles]
    expect = 0
    for coeff, idx in prods:
        term = coeff
        for i in idx:
            term = term * evals[i] % R
        expect += term
    assert interpolate(prev, point[-1]) == expect % R


def interpolate(values, x):
    n = len(values)
    total = 0
    for j in range(n):
        num, den = 1, 1
        for m in range(n):
            if m != j:
                num = num * (x - m) % R
                den = den * (j - m) % R
        total += values[j] * num * pow(den, R - 2, R)
    return total % R


def test_invalid_challenge_rejected_by_host_mirror(backend):
    f = backend.mle_from_evaluations(2, [2, 3, 5, 7])
    st = backend.prover_init(VirtualPolynomial.new_from_mle(f, 1))
    backend.prove_round_and_update_state(st, None)
    with pytest.raises(InvalidParameters):
        backend.prove_round_and_update_state(st, R)


def test_round_batch_matches_oracle(backend, oracle, rng_factory):
    """Rounds of several states of different sizes and degrees (including the
    generic path above degree 8) advanced together through round_batch,
    every message byte-exact against the oracle."""
    rng = rng_factory(900)
    specs = [(3, [(1, [0, 1])]), (9, [(5, [0, 1, 2]), (R - 1, [2])]), (6, [(3, [0, 0, 1, 1, 2])]),
             (11, [(7, [0, 1, 2, 0, 1, 2, 0])]), (5, [(2, [0, 1, 2] * 3)]), (1, [(9, [0, 1])]),
             (17, [(1, [0, 1]), (R - 3, [2, 0])])]
    states, osts = [], []
    for nv, prods in specs:
        tabs = [rng.fr_bytes(1 << nv) for _ in range(3)]
        mles = [backend.mle_from_evaluations(nv, t) for t in tabs]
        vp = VirtualPolynomial(nv)
        for c, idx in prods:
            vp.add_mle_list([mles[i] for i in idx], c)
        states.append(backend.prover_init(vp))
        osts.append(oracle.SumCheckState(nv, vp.aux_info.max_degree,
                                         [fr_from_bytes(t.tobytes()) for t in tabs], prods))
    srng = pf.SplitMix64(31)
    chal = [None] * len(specs)
    rounds = [0] * len(specs)
    while True:
        live = [i for i, (nv, _) in enumerate(specs) if rounds[i] < nv]
        if not live:
            break
        got = backend.prove_rounds_batch([states[i] for i in live], [chal[i] for i in live])
        for i, msg in zip(live, got):
            assert msg == osts[i].round(chal[i]), (i, rounds[i])
            rounds[i] += 1
            chal[i] = srng.nonboolean()


def test_round_batch_checks_before_work(backend):
    f = backend.mle_from_evaluations(2, [2, 3, 5, 7])
    a = backend.prover_init(VirtualPolynomial.new_from_mle(f, 2))
    b = backend.prover_init(VirtualPolynomial.new_from_mle(f, 3))
    with pytest.raises(InvalidParameters):  # the same state twice
        backend.prove_rounds_batch([a, a], [None, None])
    with pytest.raises(InvalidState):  # b has a challenge at round zero
        backend.prove_rounds_batch([a, b], [None, 5])
    # nothing advanced: both still answer round zero
    assert backend.prove_rounds_batch([a, b], [None, None]) == [[14, 20], [21, 30]]


@pytest.mark.parametrize("degree", ["2", "5"])
def test_round_graph_replay_gives_same_messages(degree):
    """Rounds replayed from hipGraphs (HP_MI355X_ROUND_GRAPHS=1, the round
    loops off so that every round is a launched one) produce the same messages
    as the defaults: whole n = 12 sumchecks, 1 + 2 runs each (the later runs
    replay the graphs built by the first, with the packs copied in by the
    graph's first node), on the default decompositions, on the class kernels
    and their finish for every round, and on the pair kernels and the split
    finish for every round."""
    import json
    import os
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    prog = os.path.join(root, "_build", "prove_bench")
    loops_off = {"HP_MI355X_SC_WIDE_ITEMS": "0", "HP_MI355X_SC_TAIL_SINGLE_ITEMS": "0"}
    digests = []
    for env_extra in ({}, dict(loops_off, HP_MI355X_ROUND_GRAPHS="1"),
                      dict(loops_off, HP_MI355X_ROUND_GRAPHS="1", HP_MI355X_SC_ITEMS_BELOW="0"),
                      dict(loops_off, HP_MI355X_ROUND_GRAPHS="1", HP_MI355X_SC_ITEMS_BELOW="0",
                           HP_MI355X_SC_CLASSES="0")):
        env = dict(os.environ, **env_extra)
        res = subprocess.run([prog, "sumcheck", "12", degree, "2", "3"], capture_output=True, text=True,
                             timeout=120, env=env)
        assert res.returncode == 0, (env_extra, res.stderr)
        digests.append(json.loads(res.stdout.strip().splitlines()[-1])["messages_sha3_8"])
    assert all(d == digests[0] for d in digests), digests


@pytest.mark.parametrize("degree", ["2", "5"])
def test_round_loops_give_the_same_messages(degree):
    """The persistent round loops change no message: whole n = 12 sumchecks
    of one state with the defaults (a wide loop from round 0, then fewer
    blocks, then the one-block loop for rounds of at most 16384 items), every
    loop off, the wide loop off, the one-block loop off (a one-block stage
    that takes its pairs in chunks), wide loops of 2 blocks (many chunks per
    block), the one-block loop for rounds of at most 4096 items or for every
    round whose lines fit the block (a large item limit), and with a zero
    device wait limit (every block stops right after its message, so each
    round relaunches the loop from the state's round and a wide launch finds
    its ticket dirty)."""
    import json
    import os
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    prog = os.path.join(root, "_build", "prove_bench")
    digests = []
    for env_extra in ({}, {"HP_MI355X_SC_TAIL_SINGLE_ITEMS": "0", "HP_MI355X_SC_WIDE_ITEMS": "0"},
                      {"HP_MI355X_SC_WIDE_ITEMS": "0"}, {"HP_MI355X_SC_TAIL_SINGLE_ITEMS": "0"},
                      {"HP_MI355X_SC_WIDE_BLOCKS": "2"},
                      {"HP_MI355X_SC_WIDE_ITEMS": "0", "HP_MI355X_SC_TAIL_SINGLE_ITEMS": "4096"},
                      {"HP_MI355X_SC_WIDE_ITEMS": "0", "HP_MI355X_SC_TAIL_SINGLE_ITEMS": "100000000"},
                      {"HP_MI355X_SC_TAIL_WAIT_US": "0"},
                      {"HP_MI355X_SC_TAIL_SINGLE_ITEMS": "4096", "HP_MI355X_SC_TAIL_WAIT_US": "0"}):
        env = dict(os.environ, **env_extra)
        res = subprocess.run([prog, "sumcheck", "12", degree, "2", "3"], capture_output=True, text=True,
                             timeout=120, env=env)
        assert res.returncode == 0, (env_extra, res.stderr[-2000:])
        digests.append(json.loads(res.stdout.strip().splitlines()[-1])["messages_sha3_8"])
    assert len(set(digests)) == 1, digests


def test_round_loops_interleaved_states_and_other_work(backend, oracle, rng_factory):
    """Three states advanced two at a time in alternating pairs (each batch
    stops the running loop and starts one for its pair), one state alone
    between them (the launch path), and other work on the context between
    rounds (which stops the running loop), each round byte-exact against the
    oracle; a state destroyed while its loop waits, then the context keeps
    working."""
    nv = 9
    rng = rng_factory(0x7A11)
    tabs = [rng.fr_bytes(1 << nv) for _ in range(4)]
    mles = [backend.mle_from_evaluations(nv, t) for t in tabs]
    specs = [[(3, [0, 1]), (R - 5, [2])], [(7, [1, 2, 3]), (2, [0, 0])], [(5, [3, 3, 1])]]
    states, osts = [], []
    for prods in specs:
        vp = VirtualPolynomial(nv)
        for c, idx in prods:
            vp.add_mle_list([mles[i] for i in idx], c)
        states.append(backend.prover_init(vp))
        osts.append(oracle.SumCheckState(nv, vp.aux_info.max_degree, tabs, prods))
    srng = pf.SplitMix64(0x7A12)
    ch = [None, None, None]
    for rnd in range(nv):
        for group in ((0, 1), (2,)) if rnd % 2 == 0 else ((1, 2), (0,)):
            got = backend.prove_rounds_batch([states[i] for i in group], [ch[i] for i in group])
            for i, msg in zip(group, got):
                assert msg == osts[i].round(ch[i]), (i, rnd)
                ch[i] = srng.nonboolean()
        if rnd % 3 == 1:
            pt = [srng.field() for _ in range(nv)]
            assert backend.evaluate_mle(mles[0], pt) == oracle.mle_evaluate(tabs[0], pt)
    # states stopped in the middle of their loop
    sts, osts2 = [], []
    for prods in ([(11, [0, 3])], [(4, [1, 2])]):
        vp = VirtualPolynomial(nv)
        for c, idx in prods:
            vp.add_mle_list([mles[i] for i in idx], c)
        sts.append(backend.prover_init(vp))
        osts2.append(oracle.SumCheckState(nv, 2, tabs, prods))
    c = [None, None]
    for _ in range(nv - 2):
        got = backend.prove_rounds_batch(sts, c)
        assert got == [osts2[0].round(c[0]), osts2[1].round(c[1])]
        c = [srng.nonboolean(), srng.nonboolean()]
    del sts
    assert backend.evaluate_mle(mles[1], [1] * nv) == oracle.mle_evaluate(tabs[1], [1] * nv)


def test_round_loops_batches_gain_members(backend, oracle, rng_factory):
    """round_batch through the loops: states join the batch's loop in later
    rounds (the running loop is stopped and relaunched with every state), all
    messages byte-exact against the oracle."""
    rng = rng_factory(0x7A13)
    specs = [(5, [(1, [0, 1])]), (8, [(5, [0, 1, 2]), (R - 1, [2])]), (6, [(3, [0, 0, 1, 1, 2])]),
             (7, [(2, [0, 1])]), (9, [(4, [2, 1])])]
    states, osts = [], []
    for nv, prods in specs:
        tabs = [rng.fr_bytes(1 << nv) for _ in range(3)]
        mles = [backend.mle_from_evaluations(nv, t) for t in tabs]
        vp = VirtualPolynomial(nv)
        for c, idx in prods:
            vp.add_mle_list([mles[i] for i in idx], c)
        states.append(backend.prover_init(vp))
        osts.append(oracle.SumCheckState(nv, vp.aux_info.max_degree, tabs, prods))
    srng = pf.SplitMix64(0x7A14)
    chal = [None] * len(specs)
    rounds = [0] * len(specs)
    step = 0
    while True:
        # states join one by one over the first steps
        live = [i for i, (nv, _) in enumerate(specs) if rounds[i] < nv and i <= step]
        if not live:
            break
        got = backend.prove_rounds_batch([states[i] for i in live], [chal[i] for i in live])
        for i, msg in zip(live, got):
            assert msg == osts[i].round(chal[i]), (i, rounds[i])
            rounds[i] += 1
            chal[i] = srng.nonboolean()
        step += 1
