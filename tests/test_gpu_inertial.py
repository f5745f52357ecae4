"""The device inertial error term (mlba_inertial_evaluate) against the CPU
oracle: residuals and all 8 ambient Jacobian blocks of many edges with
uneven sample counts."""
import numpy as np
import pytest

from maplab_amd import inertial
from test_oracle_inertial import G, NOISE

pytestmark = pytest.mark.gpu


def random_chain(seed, n_vertices=12):
    rng = np.random.default_rng(seed)
    P = n_vertices
    q = rng.normal(size=(P, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q *= np.sign(q[:, 3:4])
    verts = np.concatenate([q, rng.normal(size=(P, 3))], axis=1)
    counts = rng.integers(2, 60, size=P - 1)
    begin = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    S = int(begin[-1])
    t = np.zeros(S, dtype=np.int64)
    imu = np.zeros((S, 6))
    for e in range(P - 1):
        n = counts[e]
        ts = np.arange(n) * 0.005
        t[begin[e]:begin[e + 1]] = (ts * 1e9).astype(np.int64) + int(e * 1e9)
        phase = rng.uniform(0, 6, size=6)
        imu[begin[e]:begin[e + 1]] = np.stack(
            [0.5 * np.sin(3 * ts + phase[0]), 0.3 * np.cos(2 * ts + phase[1]), G + 0.2 * np.sin(5 * ts + phase[2]),
             0.2 * np.sin(4 * ts + phase[3]), -0.3 * np.cos(3 * ts + phase[4]), 0.4 * np.sin(ts + phase[5])], axis=1)
    return inertial.InertialEdges(
        vertices=verts, gyro_bias=rng.normal(size=(P, 3)) * 0.01, velocity=rng.normal(size=(P, 3)),
        accel_bias=rng.normal(size=(P, 3)) * 0.02, edge_from=np.arange(P - 1), edge_to=np.arange(1, P),
        edge_sample_begin=begin, imu=imu, timestamps_ns=t, noise=tuple(NOISE[:4]), gravity=G)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_inertial_matches_oracle(oracle, seed):
    ed = random_chain(seed)
    print(f"inertial parity: seed {seed} edges {len(ed.edge_from)} samples {len(ed.timestamps_ns)}", flush=True)
    res, jac = inertial.evaluate(ed)
    for e in range(len(ed.edge_from)):
        a, b = ed.edge_from[e], ed.edge_to[e]
        params = [ed.vertices[a], ed.gyro_bias[a], ed.velocity[a], ed.accel_bias[a],
                  ed.vertices[b], ed.gyro_bias[b], ed.velocity[b], ed.accel_bias[b]]
        s0, s1 = ed.edge_sample_begin[e], ed.edge_sample_begin[e + 1]
        r, J, _, _, _ = oracle.inertial_evaluate(ed.imu[s0:s1], ed.timestamps_ns[s0:s1], NOISE, params)
        scale = max(1.0, np.abs(r).max())
        np.testing.assert_allclose(res[e], r, atol=1e-9 * scale, rtol=1e-9)
        for k in range(8):
            np.testing.assert_allclose(jac[k][e], J[k], atol=1e-9 * max(1.0, np.abs(J[k]).max()), rtol=1e-9)
