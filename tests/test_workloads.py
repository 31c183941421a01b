"""CPU: the synthetic workloads of the batched and served paths (seeded
generators of the quadrupole-trap batch and the serving stream) and the
serving loop's memory model."""
import numpy as np

from depsolver_amd import mesh


def test_serving_stream_follows_its_description():
    st = mesh.config5_stream(24, seed=3)
    t = np.array([a for a, _ in st])
    assert np.all(np.diff(t) > 0)
    for _, p in st:
        assert 5000 < p.n_nodes <= 30000
        assert 1e4 <= p.freq <= 5e7
        assert p.ifaces.tolist() == [[1, 2], [2, 3]]
        # two concentric surfaces per particle: bc 6 outside (interface 1), bc 0 inside (interface 2)
        n = p.n_nodes
        assert set(np.unique(p.bctype[:n])) == {0, 1, 6}
        out = p.bctype[:n] == 6
        assert np.all(p.bc[:n][out, 1] == 1) and np.all(p.bc[:n][p.bctype[:n] == 0, 1] == 2)
        assert p.xyz[~(p.bctype[:n] == 1), 2].min() > 0.2   # particles clear of the electrodes


def test_serving_stream_is_seeded():
    a = mesh.config5_stream(4, seed=5)
    b = mesh.config5_stream(4, seed=5)
    for (ta, pa), (tb, pb) in zip(a, b):
        assert ta == tb and pa.freq == pb.freq and np.array_equal(pa.xyz, pb.xyz)


def test_position_batch_sizes():
    ps = mesh.config4_problems(16)
    assert {p.n_nodes for p in ps} <= {4998, 6918, 14598}
    for p in ps:
        sph = p.bctype[:p.n_nodes] == 6
        c = p.xyz[sph].mean(axis=0)
        r = np.linalg.norm(p.xyz[sph] - c, axis=1)
        assert 0.5 <= r.max() <= 1.5 + 1e-9 and abs(c[0]) <= 2 and abs(c[1]) <= 2 and 1 <= c[2] <= 4


def test_request_footprint_orders_the_solvers():
    import depsolver_amd as dep
    p = mesh.config3_geometry()
    f = {s: dep.Server.footprint(p, s) for s in dep.SOLVER}
    N = p.n_nodes
    assert f["gmres-ir-fp32"] > 12 * N * N           # P and its FP32 copy
    assert f["lu-fp64"] > f["lu-fp16"] > 8 * N * N   # complex FP64 planes beside P
    assert dep.Server.footprint(p, want_mst=False) < f["gmres-ir-fp32"]
