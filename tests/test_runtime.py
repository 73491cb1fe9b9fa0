"""Runtime mirror: graph analysis, registry errors, executor dispatch and
transfer insertion, pool reuse, trace — CPU only, with stand-in kernels
registered under the HIP backend the way the reference tests register CPU
kernels under the Metal backend (tests/correctness/cpu/test_runtime_backend_switch.py:11-40)."""
import json

import pytest

from pyzkp_amd.runtime import (Backend, Buffer, DeviceContext, Device, DType, Executor, Graph, KernelRegistry,
                               MemoryPool, OpType, Trace, export_chrome_tracing)


def test_topo_order_and_errors():
    g = Graph()
    g.add_buffer(id="a", device=Device.CPU, dtype=DType.FR, data=[1])
    g.add_node(op=OpType.POLY_SUB, inputs=["b", "a"], outputs=["c"])
    g.add_node(op=OpType.POINTWISE_MUL, inputs=["a", "a"], outputs=["b"])
    an = g.analyze()
    assert an.topo_order.index(1) < an.topo_order.index(0)
    assert an.initial == ("a",)
    assert g.analyze_cached() is g.analyze_cached()
    g2 = Graph()
    g2.add_buffer(id="a", device=Device.CPU, dtype=DType.FR, data=[1])
    g2.add_node(op=OpType.POLY_SUB, inputs=["x"], outputs=["y"])
    with pytest.raises(ValueError):
        g2.analyze()
    g3 = Graph()
    g3.add_buffer(id="a", device=Device.CPU, dtype=DType.FR, data=[1])
    g3.add_node(op=OpType.POLY_SUB, inputs=["c"], outputs=["b"])
    g3.add_node(op=OpType.POLY_SUB, inputs=["b"], outputs=["c"])
    with pytest.raises(ValueError):
        g3.analyze()
    g4 = Graph()
    g4.add_buffer(id="a", device=Device.CPU, dtype=DType.FR, data=[1])
    g4.add_node(op=OpType.POLY_SUB, inputs=["a"], outputs=["a"])
    with pytest.raises(ValueError):
        g4.analyze()
    with pytest.raises(ValueError):
        g4.add_buffer(id="a", device=Device.CPU, dtype=DType.FR, data=[])


def test_registry_errors():
    r = KernelRegistry()
    f = lambda ctx: {"outputs": {}}
    r.register(OpType.MSM_G1, Device.HIP, f, backend=Backend.HIP)
    with pytest.raises(ValueError):
        r.register(OpType.MSM_G1, Device.HIP, f, backend=Backend.HIP)
    with pytest.raises(KeyError):
        r.get(OpType.MSM_G1, Device.CPU, backend=Backend.HIP)
    assert r.has(OpType.MSM_G1, Device.HIP, backend=Backend.HIP)


class _Dev:
    """Stand-in for device-resident data."""
    def __init__(self, vals):
        self.vals = list(vals)
        self.n = len(vals)


def _fake_registry(log):
    r = KernelRegistry()
    B = Backend.HIP

    def to_dev(ctx):
        inp = ctx["inputs"][0]
        oid = ctx["node"].outputs[0]
        log.append(("to_device", inp.id))
        return {"outputs": {oid: Buffer(id=oid, device=Device.HIP, dtype=inp.dtype, data=_Dev(inp.data))}}

    def from_dev(ctx):
        inp = ctx["inputs"][0]
        oid = ctx["node"].outputs[0]
        log.append(("from_device", inp.id))
        data = inp.data.vals if isinstance(inp.data, _Dev) else inp.data
        return {"outputs": {oid: Buffer(id=oid, device=Device.CPU, dtype=inp.dtype, data=list(data))}}

    def dev_add1(ctx):
        inp = ctx["inputs"][0]
        oid = ctx["node"].outputs[0]
        assert inp.device == Device.HIP
        log.append(("hip_add1", inp.id))
        return {"outputs": {oid: Buffer(id=oid, device=Device.HIP, dtype=DType.FR,
                                        data=_Dev([x + 1 for x in inp.data.vals]))}}

    def cpu_double(ctx):
        inp = ctx["inputs"][0]
        oid = ctx["node"].outputs[0]
        assert inp.device == Device.CPU and isinstance(inp.data, list)
        log.append(("cpu_double", inp.id))
        pool = ctx["pool"]
        out = pool.alloc_fr(len(inp.data)) if pool else [0] * len(inp.data)
        for i, x in enumerate(inp.data):
            out[i] = 2 * x
        return {"outputs": {oid: Buffer(id=oid, device=Device.CPU, dtype=DType.FR, data=out)}}

    r.register(OpType.TO_DEVICE, Device.CPU, to_dev, backend=B)
    r.register(OpType.FROM_DEVICE, Device.CPU, from_dev, backend=B)
    r.register(OpType.ROOTS_EVALS_FROM_COEFFS, Device.HIP, dev_add1, backend=B)
    r.register(OpType.BATCH_INV, Device.CPU, cpu_double, backend=B)  # no HIP kernel: runs on CPU
    return r


def test_executor_inserts_transfers_and_falls_back():
    log = []
    r = _fake_registry(log)
    g = Graph()
    g.add_buffer(id="x", device=Device.CPU, dtype=DType.FR, data=[1, 2, 3])
    g.add_node(op=OpType.ROOTS_EVALS_FROM_COEFFS, inputs=["x"], outputs=["y"])
    g.add_node(op=OpType.BATCH_INV, inputs=["y"], outputs=["z"])
    ctx = DeviceContext(backend=Backend.HIP, device=Device.HIP)
    tr = Trace()
    pool = MemoryPool()
    Executor(registry=r).run(g, context=ctx, trace=tr, pool=pool, keep=["z"])
    assert log == [("to_device", "x"), ("hip_add1", "x_hip"), ("from_device", "y"), ("cpu_double", "y_cpu")]
    assert g.buffers["z"].data == [4, 6, 8]
    # dead intermediates were released, kept output survives
    assert set(g.buffers) == {"z"}
    assert [e.device for e in tr.events] == [Device.CPU, Device.HIP, Device.CPU, Device.CPU]
    assert [e.op for e in tr.events][1] == OpType.ROOTS_EVALS_FROM_COEFFS
    assert tr.by_op()["roots_evals_from_coeffs"]["count"] == 1


def test_run_repeated_reuses_pool(tmp_path):
    log = []
    r = _fake_registry(log)
    g = Graph()
    g.add_buffer(id="x", device=Device.CPU, dtype=DType.FR, data=[5, 6])
    g.add_node(op=OpType.BATCH_INV, inputs=["x"], outputs=["y"])
    g.add_node(op=OpType.BATCH_INV, inputs=["y"], outputs=["z"])
    pool = MemoryPool()
    ctx = DeviceContext(backend=Backend.HIP, device=Device.HIP, pool=pool)
    seen = []

    def before(i, graph):
        graph.reset_to_initial(graph.analyze_cached())
        seen.append(i)

    tr = Trace()
    Executor(registry=r).run_repeated(g, 3, context=ctx, keep=["z"], before_each=before, trace=tr)
    assert seen == [0, 1, 2]
    assert g.buffers["z"].data == [20, 24]
    assert pool.cpu_stats.alloc_calls == 6 and pool.cpu_stats.reuse_calls >= 2
    p = tmp_path / "trace.json"
    export_chrome_tracing(tr, str(p))
    ev = json.load(open(p))["traceEvents"]
    assert sum(1 for e in ev if e.get("ph") == "X") == 6


def test_cpu_backend_has_no_hot_path_kernels():
    from pyzkp_amd.runtime import register_hip_kernels
    r = KernelRegistry()
    register_hip_kernels(r)
    for op in (OpType.MSM_G1, OpType.ROOTS_EVALS_FROM_COEFFS, OpType.ROOTS_COEFFS_FROM_EVALS,
               OpType.COSET_EVALS_FROM_COEFFS, OpType.COSET_COEFFS_FROM_EVALS, OpType.KZG_BATCH_COMMIT,
               OpType.POINTWISE_MUL, OpType.POLY_SUB):
        assert r.has(op, Device.HIP, backend=Backend.HIP)
        assert not r.has(op, Device.CPU, backend=Backend.HIP)
        assert not r.has(op, Device.CPU, backend=Backend.CPU)


def test_hip_context_requires_device():
    from pyzkp_amd import native_available
    from pyzkp_amd.runtime import HipContext
    if native_available():
        pytest.skip("a HIP device is present")
    with pytest.raises(RuntimeError):
        HipContext.create_default()


def test_graph_unchanged_by_runs_and_repeatable_with_transfers():
    log = []
    r = _fake_registry(log)
    g = Graph()
    g.add_buffer(id="x", device=Device.CPU, dtype=DType.FR, data=[1, 2, 3])
    g.add_node(op=OpType.ROOTS_EVALS_FROM_COEFFS, inputs=["x"], outputs=["y"])
    g.add_node(op=OpType.BATCH_INV, inputs=["y"], outputs=["z"])
    before = g.analyze()
    nodes = list(g.nodes)
    ctx = DeviceContext(backend=Backend.HIP, device=Device.HIP)

    def reset(i, graph):
        graph.reset_to_initial(graph.analyze_cached())

    Executor(registry=r).run_repeated(g, 2, context=ctx, keep=["z"], before_each=reset)
    assert g.nodes == nodes                          # inputs never rewritten to the copies
    assert g.buffers["z"].data == [4, 6, 8] and set(g.buffers) == {"x", "z"}
    assert log.count(("to_device", "x")) == 2
    after = g.analyze()                              # outputs of a run are not taken for sources
    assert after.initial == before.initial == ("x",) and after.topo_order == before.topo_order


def test_insertion_order_kept_when_legal():
    g = Graph()
    g.add_buffer(id="a", device=Device.CPU, dtype=DType.FR, data=[1])
    g.add_node(op=OpType.POLY_SUB, inputs=["a"], outputs=["p"])
    g.add_node(op=OpType.POLY_SUB, inputs=["a"], outputs=["q"])
    g.add_node(op=OpType.POLY_SUB, inputs=["q", "p"], outputs=["r"])
    g.add_node(op=OpType.POLY_SUB, inputs=["a"], outputs=["s"])
    assert g.analyze().topo_order == (0, 1, 2, 3)
    g.add_node(op=OpType.POLY_SUB, inputs=["r"], outputs=["p"])
    with pytest.raises(ValueError):                  # p defined twice
        g.analyze()


def test_no_kept_list_releases_nothing():
    log = []
    r = _fake_registry(log)
    g = Graph()
    g.add_buffer(id="x", device=Device.CPU, dtype=DType.FR, data=[1])
    g.add_node(op=OpType.BATCH_INV, inputs=["x"], outputs=["y"])
    g.add_node(op=OpType.BATCH_INV, inputs=["y"], outputs=["z"])
    Executor(registry=r).run(g, context=DeviceContext(backend=Backend.HIP, device=Device.HIP))
    assert set(g.buffers) == {"x", "y", "z"}


def test_kernel_without_outputs_is_an_error():
    r = KernelRegistry()
    r.register(OpType.BATCH_INV, Device.CPU, lambda ctx: {}, backend=Backend.CPU)
    g = Graph()
    g.add_buffer(id="x", device=Device.CPU, dtype=DType.FR, data=[1])
    g.add_node(op=OpType.BATCH_INV, inputs=["x"], outputs=["y"])
    with pytest.raises(ValueError):
        Executor(registry=r).run(g)


def test_graph_runs_twice_without_reanalysis_and_cache_invalidates():
    """The reference's runtime graph tests (tests/correctness/cpu/test_runtime_graph.py,
    test_runtime_graph_batch.py): out-of-order nodes run in dependency order,
    a graph runs twice in a row without a reset, and adding a node replaces
    the cached analysis."""
    log = []
    r = _fake_registry(log)
    g = Graph()
    g.add_buffer(id="a", device=Device.CPU, dtype=DType.FR, data=[2, 3, 4, 5])
    g.add_node(op=OpType.BATCH_INV, inputs=["c"], outputs=["d"])
    g.add_node(op=OpType.BATCH_INV, inputs=["a"], outputs=["c"])
    exe = Executor(registry=r)
    ctx = DeviceContext(backend=Backend.HIP, device=Device.HIP)
    a1 = g.analyze_cached()
    exe.run(g, context=ctx)
    assert g.buffers["d"].data == [8, 12, 16, 20]
    exe.run(g, context=ctx)
    assert g.buffers["d"].data == [8, 12, 16, 20]
    assert g.analyze_cached() is a1
    g.add_node(op=OpType.BATCH_INV, inputs=["d"], outputs=["e"])
    a2 = g.analyze_cached()
    assert a2 is not a1 and a2.topo_order != a1.topo_order
    exe.run(g, context=ctx)
    assert g.buffers["e"].data == [16, 24, 32, 40]
