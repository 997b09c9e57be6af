"""Host protocol layer on the CPU (tests/cpp/protocol_test.cpp over the oracle
backend): transcript, serialization and pairing known answers, and HyperPlonk
prove + verify with tampering checks.

Pins: Keccak-f[1600] through SHA3-256 against hashlib at lengths around the
rate; the Merlin construction through merlin's published simple-transcript
challenge; from_le_bytes_mod_order through Python integers; the compressed G1
encoding through the standard compressed BLS12-381 generator; the pairing
through bilinearity, non-degeneracy and order r (checked inside the program).
SRS files: the compressed G2 generator against its standard encoding, square
roots in Fq and Fq2 against Euler's criterion, decompression round trips of
random points and their negations, a small SRS written, read and decompressed
point by point, and malformed files refused."""
import hashlib
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "_build", "protocol_test")
R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
Q = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
MERLIN_SIMPLE = "d5a21972d0d5fe320c0d263fac7fffb8145aa640af6e9bca177c03c7efcf0615"
G2_GEN_COMPRESSED = (
    "93e02b6052719f607dacd3a088274f65596bd0d09920b61ab5da61bbdc7f5049334cf11213945d57e5ac7d055d042b7e"
    "024aa2b2f08f0a91260805272dc51051c6e47ad4fa403b02b4510b647ae3d1770bac0326a805bbefd48056c8c121bdb8"
)
G1_GEN_COMPRESSED = (
    "97f1d3a73197d7942695638c4fa9ac0fc3688c4f9774b905a14e3a3f171bac586c55e83ff97a1aeffb3af00adb22c6bb"
)


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), os.path.join(ROOT, "_build", "protocol_test")],
                   check=True)


@pytest.fixture(scope="module")
def kat_lines():
    build()
    res = subprocess.run([BIN, "kat"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "KAT PASS" in res.stdout
    return [line.split() for line in res.stdout.splitlines()]


def test_keccak_against_hashlib(kat_lines):
    rows = [l for l in kat_lines if l[0] == "sha3"]
    assert len(rows) == 8
    for _, n, digest in rows:
        msg = bytes((i * 7 + 3) & 0xFF for i in range(int(n)))
        assert hashlib.sha3_256(msg).hexdigest() == digest


def test_merlin_known_answer(kat_lines):
    assert [l[1] for l in kat_lines if l[0] == "merlin"] == [MERLIN_SIMPLE]


def test_mod_order_reduction(kat_lines):
    got = [int(l[1], 16) for l in kat_lines if l[0] == "modorder"]
    assert got == [2**256 % R, (2**512 - 1) % R]


def test_mod_order_reduction_lengths(kat_lines):
    rows = [l for l in kat_lines if l[0] == "modorderx"]
    assert len(rows) == 28
    for _, field, hx, got in rows:
        data = b"" if hx == "-" else bytes.fromhex(hx)
        assert int(got, 16) == int.from_bytes(data, "little") % (R if field == "r" else Q), (field, len(data))


def test_compressed_g1(kat_lines):
    d = {l[0]: l[1] for l in kat_lines}
    assert d["g1gen"] == G1_GEN_COMPRESSED
    # -g flips only the sort flag (0x20); infinity is 0xC0 then zeros
    assert d["g1neg"][2:] == G1_GEN_COMPRESSED[2:] and int(d["g1neg"][:2], 16) == 0x97 | 0x20
    assert d["g1inf"] == "c0" + "00" * 47


@pytest.mark.parametrize("nv,gate,seed", [(4, "vanilla", 1), (5, "jellyfish", 2), (6, "vanilla", 3)])
def test_prove_verify_on_oracle_backend(nv, gate, seed):
    build()
    res = subprocess.run([BIN, "prove", str(nv), gate, str(seed)], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "PROVE PASS" in res.stdout


def test_srs_file_encoding(tmp_path):
    build()
    res = subprocess.run([BIN, "srs", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "SRS PASS" in res.stdout
    rows = {l.split()[0]: l.split()[1:] for l in res.stdout.splitlines() if l.strip()}
    assert rows["g2gen"] == [G2_GEN_COMPRESSED]
    # num_vars 3: 16 header bytes, 4 layers of 8 + 48 n bytes, g, h, masks
    assert int(rows["srs_nv3"][0]) == 16 + 4 * 8 + 48 * 15 + 48 + 96 + 8 + 96 * 3
