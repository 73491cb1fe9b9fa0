"""The hip backend behind the runtime API, on the reference's own Metal
test inputs (tests/correctness/metal/test_runtime_metal_fr_ntt.py,
test_runtime_metal_fr_coset.py, test_runtime_metal_g1_msm.py) and the
golden PLONK-key commitments, with expectations from the oracle."""
import json
import os

import pytest

from oracle import curve as C
from oracle import msm as M
from oracle import ntt as N
from oracle.field import R_MOD
from pyzkp_amd.curve import normalize
from pyzkp_amd.runtime import (Backend, Device, DType, Executor, Graph, KernelRegistry, MemoryPool, OpType,
                               RuntimeConfig, register_hip_kernels)

pytestmark = pytest.mark.gpu


def _exe():
    r = KernelRegistry()
    register_hip_kernels(r)
    return Executor(registry=r)


def test_roots_ntt_matches_cpu():
    n = 16
    omega = N.omega_for_size(n)
    coeff = [i * i % R_MOD for i in range(n)]
    g = Graph()
    g.add_buffer(id="a", device=Device.CPU, dtype=DType.FR, data=list(coeff))
    g.add_node(op=OpType.TO_DEVICE, inputs=["a"], outputs=["a_m"])
    g.add_node(op=OpType.ROOTS_EVALS_FROM_COEFFS, inputs=["a_m"], outputs=["ev_m"], attrs={"n": n, "omega": omega})
    g.add_node(op=OpType.FROM_DEVICE, inputs=["ev_m"], outputs=["ev"])
    _exe().run(g, runtime_config=RuntimeConfig(backend=Backend.HIP), keep=["ev"])
    assert g.buffers["ev"].data == N.evals_from_coeffs_on_roots(coeff, n, omega)


def test_roots_intt_roundtrip_auto_transfer():
    n = 16
    omega = N.omega_for_size(n)
    ev = [(i + 3) * 7 % R_MOD for i in range(n)]
    g = Graph()
    g.add_buffer(id="ev", device=Device.CPU, dtype=DType.FR, data=list(ev))
    # no explicit TO_DEVICE / FROM_DEVICE: the executor inserts them
    g.add_node(op=OpType.ROOTS_COEFFS_FROM_EVALS, inputs=["ev"], outputs=["coeff_m"], attrs={"omega": omega})
    g.add_node(op=OpType.FROM_DEVICE, inputs=["coeff_m"], outputs=["coeff"])
    _exe().run(g, runtime_config=RuntimeConfig(backend=Backend.HIP), keep=["coeff"])
    assert g.buffers["coeff"].data == N.coeffs_from_evals_on_roots(ev, omega)


@pytest.mark.parametrize("n,m", [(16, 16), (16, 64), (5, 8)])
def test_coset_matches_cpu(n, m):
    omega_m = N.omega_for_size(m)
    coeff = [(3 * i + 1) % R_MOD for i in range(n)]
    g = Graph()
    g.add_buffer(id="c", device=Device.CPU, dtype=DType.FR, data=list(coeff))
    g.add_node(op=OpType.COSET_EVALS_FROM_COEFFS, inputs=["c"], outputs=["e"],
               attrs={"n": m, "omega": omega_m, "shift": 5})
    g.add_node(op=OpType.COSET_COEFFS_FROM_EVALS, inputs=["e"], outputs=["c2"], attrs={"omega": omega_m, "shift": 5})
    g.add_node(op=OpType.FROM_DEVICE, inputs=["e"], outputs=["e_h"])
    g.add_node(op=OpType.FROM_DEVICE, inputs=["c2"], outputs=["c2_h"])
    _exe().run(g, runtime_config=RuntimeConfig(backend=Backend.HIP), keep=["e_h", "c2_h"])
    assert g.buffers["e_h"].data == N.evals_from_coeffs_on_coset(coeff, n=m, omega=omega_m, shift=5)
    padded = (coeff + [0] * m)[:m]
    assert g.buffers["c2_h"].data == padded


def test_pointwise_and_sub():
    a = [i * 7 + 1 for i in range(32)]
    b = [R_MOD - i * 3 for i in range(32)]
    g = Graph()
    g.add_buffer(id="a", device=Device.CPU, dtype=DType.FR, data=a)
    g.add_buffer(id="b", device=Device.CPU, dtype=DType.FR, data=b)
    g.add_node(op=OpType.POINTWISE_MUL, inputs=["a", "b"], outputs=["m"])
    g.add_node(op=OpType.POLY_SUB, inputs=["m", "a"], outputs=["s"])
    g.add_node(op=OpType.FROM_DEVICE, inputs=["s"], outputs=["s_h"])
    pool = MemoryPool()
    _exe().run(g, runtime_config=RuntimeConfig(backend=Backend.HIP), pool=pool, keep=["s_h"])
    assert g.buffers["s_h"].data == [(x * (y % R_MOD) - x) % R_MOD for x, y in zip(a, b)]
    # mismatched lengths -> ValueError
    g2 = Graph()
    g2.add_buffer(id="a", device=Device.CPU, dtype=DType.FR, data=a)
    g2.add_buffer(id="b", device=Device.CPU, dtype=DType.FR, data=b[:5])
    g2.add_node(op=OpType.POLY_SUB, inputs=["a", "b"], outputs=["s"])
    with pytest.raises(ValueError):
        _exe().run(g2, runtime_config=RuntimeConfig(backend=Backend.HIP))


def test_msm_g1_matches_cpu():
    n = 100
    scalars = [(i * 12345 + 6789) % R_MOD for i in range(n)]
    pts_aff = [C.g1_mul_affine(C.G1_GENERATOR, i + 1) for i in range(n)]
    points = [(p[0], p[1], 1) for p in pts_aff]
    g = Graph()
    g.add_buffer(id="points", device=Device.CPU, dtype=DType.OBJ, data=points)
    g.add_buffer(id="scalars", device=Device.CPU, dtype=DType.FR, data=scalars)
    g.add_node(op=OpType.MSM_G1, inputs=["points", "scalars"], outputs=["msm_res"])
    _exe().run(g, runtime_config=RuntimeConfig(backend=Backend.HIP))
    assert normalize(g.buffers["msm_res"].data) == M.msm_pippenger(pts_aff, scalars, M.default_window_bits(n))


def test_msm_g2_matches_host():
    # MSM_G2 through the registry (the reference registers it on the CPU
    # only): the device sum over bases [k_i] Q equals [sum s_i k_i] Q, with
    # the scalars handed over on the host as the reference's graph does
    from pyzkp_amd import g2 as G2
    n = 200
    ks = [(i * 7919 + 3) % R_MOD for i in range(n)]
    scalars = [(i * 12345 + 6789) % R_MOD for i in range(n)]
    bases = G2.from_words(G2.g2_mul_many(G2.G2_GENERATOR, ks))
    expect = G2.g2_mul(G2.G2_GENERATOR, sum(k * s for k, s in zip(ks, scalars)) % R_MOD)
    for window in (0, 5):
        g = Graph()
        g.add_buffer(id="b_query_g2", device=Device.CPU, dtype=DType.OBJ, data=bases)
        g.add_buffer(id="w", device=Device.CPU, dtype=DType.FR, data=scalars)
        g.add_node(op=OpType.MSM_G2, inputs=["b_query_g2", "w"], outputs=["b_lin_g2"],
                   attrs={"hip_window_bits": window})
        _exe().run(g, runtime_config=RuntimeConfig(backend=Backend.HIP))
        assert G2.g2_eq(g.buffers["b_lin_g2"].data, expect)


def test_kzg_batch_commit_golden():
    with open(os.path.join(os.path.dirname(__file__), "golden", "plonk_pk_n16.json")) as f:
        pk = json.load(f)

    class SRS:
        g1_powers = tuple((p[0], p[1], 1) for p in pk["srs_g1"])

    names = list(pk["coeffs"])
    g = Graph()
    g.add_buffer(id="srs", device=Device.CPU, dtype=DType.OBJ, data=SRS())
    g.add_buffer(id="polys", device=Device.CPU, dtype=DType.OBJ, data=[pk["coeffs"][k] for k in names] + [[]])
    g.add_node(op=OpType.KZG_BATCH_COMMIT, inputs=["srs", "polys"], outputs=["cms"])
    _exe().run(g, runtime_config=RuntimeConfig(backend=Backend.HIP))
    got = g.buffers["cms"].data
    for k, cm in zip(names, got):
        assert normalize(cm) == tuple(pk["commits"][k]), k
    assert normalize(got[-1]) is None
    # the same commitments as one fixed-base batch launch (attrs fixed_base)
    for wbits in (0, 5):
        g3 = Graph()
        g3.add_buffer(id="srs", device=Device.CPU, dtype=DType.OBJ, data=SRS.g1_powers)
        g3.add_buffer(id="polys", device=Device.CPU, dtype=DType.OBJ, data=[pk["coeffs"][k] for k in names] + [[]])
        g3.add_node(op=OpType.KZG_BATCH_COMMIT, inputs=["srs", "polys"], outputs=["cms"],
                    attrs={"fixed_base": True, "fixed_base_window_bits": wbits})
        _exe().run(g3, runtime_config=RuntimeConfig(backend=Backend.HIP))
        got3 = g3.buffers["cms"].data
        for k, cm in zip(names, got3):
            assert normalize(cm) == tuple(pk["commits"][k]), (k, wbits)
        assert normalize(got3[-1]) is None
    # polynomial longer than the SRS -> ValueError
    g2 = Graph()
    g2.add_buffer(id="srs", device=Device.CPU, dtype=DType.OBJ, data=SRS())
    g2.add_buffer(id="polys", device=Device.CPU, dtype=DType.OBJ, data=[[1] * 17])
    g2.add_node(op=OpType.KZG_BATCH_COMMIT, inputs=["srs", "polys"], outputs=["cms"])
    with pytest.raises(ValueError):
        _exe().run(g2, runtime_config=RuntimeConfig(backend=Backend.HIP))
