"""Learned correction of the block-Jacobi preconditioner of the reduced camera
system (include/maplab_ba_learned.h): training records from the device BA,
the network's training on the device, and the model file.

The preconditioner is M^-1 = G'G, G = (I - K) L^-1 with per-row coefficients
theta_i of K predicted by a 3-layer message-passing network over the camera
covisibility graph; the network is trained on the device (its own forward,
backward and Adam kernels) to minimize the Kaporin condition number of
G S G' row by row on reduced camera systems the BA produces. No counterpart
in the reference (maplab's Ceres ITERATIVE_SCHUR uses the fixed SCHUR_JACOBI
preconditioner, solver-options.h:57-68).

python -m maplab_amd.learned train [--maps N] [--walk-maps W] [--epochs E] [--out FILE]
    generates records on maps of the circle and random-walk workload
    generators with varied seeds and sizes, trains, and writes the model file.
"""
from __future__ import annotations

import argparse
import base64
import ctypes
import json
import os
import time

import numpy as np

from . import abi, problem
from .solver import DeviceProblem

DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models", "covis_precond.json")
NODE_F, EDGE_F = abi.GNN_NODE_FEATURES, abi.GNN_EDGE_FEATURES


def num_params() -> int:
    return abi.load_library().mlba_gnn_num_params()


def _p(a):
    return None if a is None else a.ctypes.data


class Records:
    """One reduced camera system: graph (CSR by row, diagonal slots kept),
    raw features and per-row traces (a, b, e, d, has)."""

    def __init__(self, row_ptr, col, rev, node_feat, edge_feat, qc):
        self.row_ptr, self.col, self.rev = row_ptr, col, rev
        self.node_feat, self.edge_feat, self.qc = node_feat, edge_feat, qc

    @property
    def nodes(self):
        return len(self.row_ptr) - 1


def rcs_records(dev: DeviceProblem, options: abi.Options, radius: float) -> Records:
    lib = dev.lib
    rows, slots = ctypes.c_int32(), ctypes.c_int32()
    abi.check(lib.mlba_rcs_size(dev.handle, ctypes.byref(rows), ctypes.byref(slots)))
    n, z = rows.value, slots.value
    row_ptr = np.zeros(n + 1, np.int32)
    col, rev = np.zeros(z, np.int32), np.zeros(z, np.int32)
    nf = np.zeros((n, NODE_F), np.float32)
    ef = np.zeros((z, EDGE_F), np.float32)
    qc = np.zeros((n, 5), np.float32)
    dev._check(lib.mlba_rcs_records(dev.handle, ctypes.byref(options), radius, _p(row_ptr), _p(col), _p(rev),
                                    _p(nf), _p(ef), _p(qc)))
    return Records(row_ptr, col, rev, nf, ef, qc)


def rcs_system(dev: DeviceProblem, options: abi.Options, radius: float):
    """The RCS at the current state and radius as a dense matrix (6 rows per
    variable vertex, Jacobi-scaled variables) and its right-hand side."""
    lib = dev.lib
    rows, slots = ctypes.c_int32(), ctypes.c_int32()
    abi.check(lib.mlba_rcs_size(dev.handle, ctypes.byref(rows), ctypes.byref(slots)))
    n, z = rows.value, slots.value
    row_ptr = np.zeros(n + 1, np.int32)
    col = np.zeros(z, np.int32)
    blocks = np.zeros((z, 6, 6))
    rhs = np.zeros(6 * n)
    dev._check(lib.mlba_rcs_system(dev.handle, ctypes.byref(options), radius, _p(row_ptr), _p(col), _p(blocks),
                                   _p(rhs)))
    S = np.zeros((6 * n, 6 * n))
    for i in range(n):
        for k in range(row_ptr[i], row_ptr[i + 1]):
            j = col[k]
            S[6 * i:6 * i + 6, 6 * j:6 * j + 6] = blocks[k]
    return S, rhs


def rcs_pcg_iterations(dev: DeviceProblem, options: abi.Options, radius: float, preconditioner: int,
                       tolerance: float = 1e-6, max_iterations: int = 5000) -> int:
    it = ctypes.c_int32()
    dev._check(dev.lib.mlba_rcs_pcg_iterations(dev.handle, ctypes.byref(options), radius, preconditioner, tolerance,
                                               max_iterations, ctypes.byref(it)))
    return it.value


def normalization(records) -> np.ndarray:
    """(node mean, node std, edge mean, edge std) over rows and edge slots."""
    nf = np.concatenate([r.node_feat for r in records]).astype(np.float64)
    ef = np.concatenate([r.edge_feat[r.col != np.repeat(np.arange(r.nodes), np.diff(r.row_ptr))]
                         for r in records]).astype(np.float64)
    out = np.concatenate([nf.mean(0), nf.std(0) + 1e-3, ef.mean(0), ef.std(0) + 1e-3])
    return out.astype(np.float32)


class Dataset:
    """Disjoint union of records, features normalized, as mlba_gnn_dataset."""

    def __init__(self, records, norm: np.ndarray):
        rp, col, rev, nf, ef, qc = [np.zeros(1, np.int32)], [], [], [], [], []
        node_off = slot_off = 0
        for r in records:
            rp.append(r.row_ptr[1:] + slot_off)
            col.append(r.col + node_off)
            rev.append(r.rev + slot_off)
            nf.append(r.node_feat)
            dst = np.repeat(np.arange(r.nodes), np.diff(r.row_ptr))
            e = (r.edge_feat - norm[2 * NODE_F:2 * NODE_F + EDGE_F]) / norm[2 * NODE_F + EDGE_F:]
            e[r.col == dst] = 0.0
            ef.append(e)
            qc.append(r.qc)
            node_off += r.nodes
            slot_off += len(r.col)
        self.row_ptr = np.ascontiguousarray(np.concatenate(rp), np.int32)
        self.col = np.ascontiguousarray(np.concatenate(col), np.int32)
        self.rev = np.ascontiguousarray(np.concatenate(rev), np.int32)
        self.node_feat = np.ascontiguousarray((np.concatenate(nf) - norm[:NODE_F]) / norm[NODE_F:2 * NODE_F],
                                              np.float32)
        self.edge_feat = np.ascontiguousarray(np.concatenate(ef), np.float32)
        self.qc = np.ascontiguousarray(np.concatenate(qc), np.float32)
        self.nodes, self.slots = len(self.row_ptr) - 1, len(self.col)

    def struct(self) -> abi.GnnDataset:
        return abi.GnnDataset(self.nodes, self.slots, _p(self.row_ptr), _p(self.col), _p(self.rev),
                              _p(self.node_feat), _p(self.edge_feat), _p(self.qc))


def _gnn_check(rc):
    if rc != abi.MLBA_OK:
        msg = abi.load_library().mlba_gnn_last_error()
        raise RuntimeError(f"maplab_ba error {rc}: {msg.decode() if msg else ''}")


def evaluate(data: Dataset, params: np.ndarray, want_grad: bool = True):
    """(theta, loss, grad) of the network on the device."""
    lib = abi.load_library()
    params = np.ascontiguousarray(params, np.float32)
    theta = np.zeros((data.nodes, 2), np.float32)
    loss = np.zeros(1, np.float32)
    grad = np.zeros(num_params(), np.float32) if want_grad else None
    ds = data.struct()
    _gnn_check(lib.mlba_gnn_evaluate(ctypes.byref(ds), _p(params), _p(theta), _p(loss), _p(grad)))
    return theta, float(loss[0]), grad


def train(data: Dataset, params: np.ndarray, epochs: int, learning_rate: float = 3e-3):
    """Adam on the device; returns (params, loss before each step)."""
    lib = abi.load_library()
    p = np.ascontiguousarray(params, np.float32).copy()
    log = np.zeros(max(epochs, 1), np.float32)
    ds = data.struct()
    _gnn_check(lib.mlba_gnn_train(ctypes.byref(ds), epochs, learning_rate, _p(p), _p(log)))
    return p, log[:epochs]


def init_params(seed: int = 0, theta0=(0.5, -4.0)) -> np.ndarray:
    """He-initialized weights, output layer near zero with bias theta0."""
    rng = np.random.default_rng(seed)
    H, F, E, T = 64, NODE_F, EDGE_F, 2
    parts = [rng.normal(size=F * H) * np.sqrt(2.0 / F), np.zeros(H)]
    for _ in range(3):
        parts += [rng.normal(size=H * H) * np.sqrt(2.0 / H), np.zeros(H), np.zeros(E), np.zeros(1),
                  rng.normal(size=2 * H * H) * np.sqrt(2.0 / (2 * H)), np.zeros(H)]
    parts += [rng.normal(size=H * T) * 0.01 * np.sqrt(2.0 / H), np.asarray(theta0, dtype=np.float64)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == num_params()
    return p


class Model:
    """Network parameters and the feature normalization."""

    def __init__(self, params: np.ndarray, norm: np.ndarray, meta: dict | None = None):
        self.params = np.ascontiguousarray(params, np.float32)
        self.norm = np.ascontiguousarray(norm, np.float32)
        self.meta = meta or {}

    def save(self, path: str = DEFAULT_MODEL) -> None:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        doc = {"format": "maplab_amd covisibility preconditioner v1", "num_params": int(self.params.size),
               "norm": [float(x) for x in self.norm],
               "params_f32_base64": base64.b64encode(self.params.tobytes()).decode(), "meta": self.meta}
        with open(path, "w") as fh:
            json.dump(doc, fh, indent=1)
            fh.write("\n")

    @staticmethod
    def load(path: str = DEFAULT_MODEL) -> "Model":
        with open(path) as fh:
            doc = json.load(fh)
        params = np.frombuffer(base64.b64decode(doc["params_f32_base64"]), dtype=np.float32).copy()
        if params.size != num_params():
            raise ValueError(f"{path}: {params.size} parameters, the library expects {num_params()}")
        return Model(params, np.asarray(doc["norm"], np.float32), doc.get("meta", {}))

    def install(self, dev: DeviceProblem) -> None:
        dev._check(dev.lib.mlba_set_preconditioner_model(dev.handle, _p(self.params), _p(self.norm)))


def stats(dev: DeviceProblem) -> abi.PreconditionerStats:
    st = abi.PreconditionerStats()
    dev._check(dev.lib.mlba_get_preconditioner_stats(dev.handle, ctypes.byref(st)))
    return st


def training_records(n_maps: int, records_per_map: int, seed: int = 1000, walk_maps: int = 0,
                     walk_records: int = 8, log=print):
    """Records of reduced camera systems the BA produces: n_maps maps of the
    circle workload generator (10-120 keyframes, 500-8000 landmarks) and
    walk_maps maps of the random-walk generator (30-300 keyframes, 150-400
    landmarks per keyframe), varied seeds; at each map's start and along its
    LM iterations (two iterations between every third record), radius
    log-uniform in [1e3, 1e9]."""
    rng = np.random.default_rng(seed)
    out = []
    t0 = time.time()
    total = n_maps + walk_maps
    for k in range(total):
        if k < n_maps:
            kf = int(rng.integers(10, 121))
            m = problem.workload_map(1, seed=seed + k, num_keyframes=kf, num_landmarks=int(rng.integers(500, 8001)))
            per = records_per_map
        else:
            kf = int(rng.integers(30, 301))
            m = problem.workload_map(2, seed=seed + k, num_keyframes=kf, num_landmarks=kf * int(rng.integers(150, 401)))
            per = walk_records
        o = abi.default_options()
        with DeviceProblem(m) as dev:
            for r in range(per):
                if r and r % 3 == 0:
                    step = abi.default_options()
                    step.max_num_iterations = 2
                    dev.solve(step)
                out.append(rcs_records(dev, o, float(10 ** rng.uniform(3, 9))))
        if log and (k + 1) % max(1, total // 10) == 0:
            log(f"records: {k + 1}/{total} maps, {len(out)} systems, {time.time() - t0:.1f}s")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the covisibility preconditioner network on the device")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--maps", type=int, default=300, help="circle-generator maps")
    t.add_argument("--records", type=int, default=10, help="systems per circle map")
    t.add_argument("--walk-maps", type=int, default=250, help="random-walk-generator maps")
    t.add_argument("--walk-records", type=int, default=8, help="systems per random-walk map")
    t.add_argument("--epochs", type=int, default=400)
    t.add_argument("--lr", type=float, default=3e-3)
    t.add_argument("--seed", type=int, default=1000)
    t.add_argument("--out", default=DEFAULT_MODEL)
    a = ap.parse_args(argv)
    recs = training_records(a.maps, a.records, a.seed, a.walk_maps, a.walk_records,
                            log=lambda s: print(s, flush=True))
    norm = normalization(recs)
    data = Dataset(recs, norm)
    print(f"dataset: {len(recs)} systems, {data.nodes} rows, {data.slots} slots", flush=True)
    p0 = init_params(a.seed)
    t0 = time.time()
    params, log = train(data, p0, a.epochs, a.lr)
    print(f"trained {a.epochs} epochs in {time.time() - t0:.1f}s: loss {log[0]:.6f} -> {log[-1]:.6f}", flush=True)
    Model(params, norm, {"systems": len(recs), "rows": int(data.nodes), "slots": int(data.slots),
                         "epochs": a.epochs, "learning_rate": a.lr, "seed": a.seed,
                         "loss_first": float(log[0]), "loss_last": float(log[-1]),
                         "generator": (f"{a.maps} circle maps (10-120 keyframes, 500-8000 landmarks) x {a.records} "
                                       f"systems, {a.walk_maps} random-walk maps (30-300 keyframes) x "
                                       f"{a.walk_records} systems")}).save(a.out)
    print(f"wrote {a.out}", flush=True)


if __name__ == "__main__":
    main()
