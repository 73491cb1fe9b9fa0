"""Groth16 on the device: the proving key and every proof equal the
oracle's (same trapdoor, r, s) point for point; proofs verify with the
pairing; batched proofs equal proofs made alone; bad witnesses are rejected."""
import numpy as np
import pytest

from oracle import curve as C1
from oracle import curve2 as C2
from oracle import groth16 as OG
from pyzkp_amd.curve import normalize
from pyzkp_amd.field import words_to_ints
from pyzkp_amd.g2 import normalize as normalize2
from pyzkp_amd.groth16 import prove, prove_many, setup, synthetic_circuit, synthetic_witness, verify

pytestmark = pytest.mark.gpu


def _g2w(words):
    from pyzkp_amd.g2 import point_from_words
    return normalize2(point_from_words(words))


@pytest.mark.parametrize("log_n", [5, 7])
def test_groth16_matches_oracle(zctx, log_n):
    circ = synthetic_circuit(log_n, 0x5000 + log_n)
    w_words = synthetic_witness(circ, 0x5100 + log_n)
    w = words_to_ints(w_words)
    pk = setup(circ, zctx, seed=0x5200 + log_n, window_bits=6)
    td = (pk.trapdoor_tau, pk.trapdoor_alpha, pk.trapdoor_beta, pk.trapdoor_gamma, pk.trapdoor_delta)
    A, B, C = circ.A.row_dicts(), circ.B.row_dicts(), circ.C.row_dicts()
    opk = OG.setup(A, B, C, circ.n, circ.n_vars, circ.public_ids, list(circ.aux_ids), td)
    # proving key, point for point
    for mine, theirs in (("a", "a_query"), ("b_g1", "b_query_g1"), ("h", "h_query"), ("l", "l_query")):
        got = [normalize(p) for p in pk.g1_query(mine)]
        assert got == opk[theirs], mine
    assert [_g2w(pk.b_query_g2.words[i]) for i in range(circ.n_vars)] == opk["b_query_g2"]
    assert [normalize(p) for p in pk.vk.ic] == opk["ic"]
    assert normalize2(pk.vk.delta_g2) == opk["delta_g2"]
    # proofs
    r, s = 987654321, 123456789
    pf = prove(pk, w_words, r, s)
    opf = OG.prove(opk, A, B, C, w, r, s)
    assert normalize(pf.a) == opf["a"]
    assert normalize2(pf.b) == opf["b"]
    assert normalize(pf.c) == opf["c"]
    assert OG.check_in_exponent(opk, w, r, s, {"a": normalize(pf.a), "b": normalize2(pf.b),
                                                "c": normalize(pf.c)}, circ.n)
    pub = [w[i] for i in circ.public_ids]
    assert verify(pk.vk, pub, pf)
    assert not verify(pk.vk, [pub[0], (pub[1] + 1)], pf)
    # a witness that does not satisfy the R1CS is rejected
    bad = w_words.copy()
    bad[0, 0] ^= 1
    with pytest.raises(ValueError):
        prove(pk, bad, r, s)
    pk.free()


def test_groth16_batch_equals_alone(zctx):
    circ = synthetic_circuit(12, 0x5300)
    pk = setup(circ, zctx, seed=0x5301)
    ws = [synthetic_witness(circ, 0x5400 + i) for i in range(3)]
    rs = [(1000 + i, 2000 + i) for i in range(3)]
    many = prove_many(pk, ws, rs)
    for i in range(3):
        alone = prove(pk, ws[i], *rs[i])
        assert normalize(alone.a) == normalize(many[i].a)
        assert normalize2(alone.b) == normalize2(many[i].b)
        assert normalize(alone.c) == normalize(many[i].c)
        wi = words_to_ints(ws[i][[circ.public_ids[1]]])
        assert verify(pk.vk, [1, wi[0]], many[i])
    pk.free()


def test_groth16_device_g2_equals_host(zctx):
    # the G2 MSM on the device (b_query_g2 . w over the witness segment the
    # G1 MSMs read) gives the host G2 MSM's proof byte for byte
    circ = synthetic_circuit(11, 0x5320)
    pk = setup(circ, zctx, seed=0x5321)
    ws = [synthetic_witness(circ, 0x5322 + i) for i in range(3)]
    rs = [(3000 + i, 4000 + i) for i in range(3)]
    host = prove_many(pk, ws, rs)
    dev = prove_many(pk, ws, rs, g2_device=True)
    for i in range(3):
        assert dev[i].to_bytes() == host[i].to_bytes()
        pub = words_to_ints(ws[i][circ.public_ids])
        assert verify(pk.vk, pub, dev[i])
    assert prove(pk, ws[1], *rs[1], g2_device=True).to_bytes() == host[1].to_bytes()
    bad = ws[0].copy()
    bad[0, 0] ^= 1
    with pytest.raises(ValueError):
        prove(pk, bad, 1, 2, g2_device=True)
    assert pk.b_query_g2_dev is not None
    pk.free()
    assert pk.b_query_g2_dev is None


def test_key_cache_roundtrip(zctx, tmp_path):
    from pyzkp_amd.groth16.keycache import CacheMismatchError, load_key, save_key
    circ = synthetic_circuit(9, 0x5310)
    pk = setup(circ, zctx, seed=0x5311)
    path = str(tmp_path / "key.npz")
    save_key(path, pk, circ)
    pk2 = load_key(path, circ, zctx)
    w = synthetic_witness(circ, 0x5312)
    a = prove(pk, w, 11, 22)
    b = prove(pk2, w, 11, 22)
    assert a.to_bytes() == b.to_bytes()
    pub = words_to_ints(w[circ.public_ids])
    assert verify(pk2.vk, pub, b)
    with pytest.raises(CacheMismatchError):
        load_key(path, synthetic_circuit(9, 0x5313), zctx)
    pk.free()
    pk2.free()
