"""The library's C++ map folder reader and writer (include/maplab_ba_vimap.h),
the entry a C / C++ host uses, against the Python reader
(maplab_amd/vimap.py, itself pinned by tests/test_vimap_io.py and the
reference .proto field table): the same problem from the same folder
(compressed or not, one or two missions, pinhole and unified-projection
cameras, sensors.yaml in block or flow style), and states written by the
C++ writer read back exactly by the Python reader with every other field
unchanged."""
import os
import shutil

import numpy as np
import pytest

from cases import tiny_map
from maplab_amd import abi, problem, vimap

FIELDS = ["vertices", "baseframes", "landmarks", "camera_intrinsics", "camera_distortion",
          "camera_distortion_type", "camera_resolution", "vertex_mission", "landmark_store_vertex", "obs_vertex",
          "obs_landmark", "obs_camera", "obs_keypoint", "obs_sigma", "vertex_constant", "landmark_constant"]


def same_problem(a, b):
    for f in FIELDS:
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    np.testing.assert_allclose(a.camera_q_CI__C_p_CI, b.camera_q_CI__C_p_CI, rtol=0, atol=1e-15)
    assert (a.camera_model is None) == (b.camera_model is None)
    if a.camera_model is not None:
        np.testing.assert_array_equal(a.camera_model, b.camera_model)
        np.testing.assert_array_equal(a.camera_xi, b.camera_xi)


@pytest.mark.parametrize("compress", [True, False])
@pytest.mark.parametrize("kind", ["walk", "two_missions", "unified"])
def test_native_reader_equals_python_reader(tmp_path, compress, kind):
    if kind == "walk":
        m = problem.workload_map(2, seed=5, num_keyframes=25, num_landmarks=2500)
    elif kind == "two_missions":
        m = tiny_map(np.random.default_rng(3), dist_type=2, n_vertices=8, n_landmarks=50, global_missions=True,
                     n_cameras=2)
    else:
        m = tiny_map(np.random.default_rng(4), dist_type=1, n_vertices=6, n_landmarks=40, xi=0.9)
    folder = str(tmp_path / "map")
    vimap.write_vimap(m, folder, compress=compress)
    py, rec = vimap.load_vimap(folder)
    nat = vimap.NativeMap(folder)
    same_problem(nat.problem(), py)
    np.testing.assert_array_equal(nat.landmark_quality(), [lm.quality for lm in rec.landmark_protos])


def test_flow_style_sensors_yaml(tmp_path):
    # maplab's own sensors.yaml files write matrices as flow sequences.
    m = problem.workload_map(2, seed=6, num_keyframes=12, num_landmarks=1200)
    folder = str(tmp_path / "map")
    vimap.write_vimap(m, folder)
    path = os.path.join(folder, vimap.FOLDER, vimap.SENSORS)
    import yaml
    doc = yaml.safe_load(open(path))
    text = yaml.safe_dump(doc, sort_keys=False, default_flow_style=None)
    text = "# written by a maplab tool\n---\n" + text
    open(path, "w").write(text)
    assert "data: [" in text
    py, _ = vimap.load_vimap(folder)
    same_problem(vimap.NativeMap(folder).problem(), py)


def test_native_writer_round_trip(tmp_path):
    m = tiny_map(np.random.default_rng(7), dist_type=1, n_vertices=7, n_landmarks=45, global_missions=True)
    src = str(tmp_path / "src")
    rec0 = vimap.write_vimap(m, src)
    dst = str(tmp_path / "dst")
    shutil.copytree(src, dst)
    nat = vimap.NativeMap(src)
    p = nat.problem()
    rng = np.random.default_rng(1)
    v = p.vertices.copy()
    v[:, 4:] += rng.normal(size=(p.num_vertices, 3))
    l = p.landmarks + rng.normal(size=p.landmarks.shape)
    q = np.where(np.arange(p.num_landmarks) % 3 == 0, 1, 0).astype(np.uint8)   # mark every third bad
    nat.save(dst, v, l, q)
    _, rec_src = vimap.load_vimap(src)
    back, rec = vimap.load_vimap(dst)
    np.testing.assert_array_equal(back.vertices, v)
    np.testing.assert_array_equal(back.landmarks, l)
    qual = np.array([lm.quality for lm in rec.landmark_protos])
    qual0 = np.array([lm.quality for lm in rec_src.landmark_protos])
    np.testing.assert_array_equal(qual[q == 1], 1)
    np.testing.assert_array_equal(qual[q == 0], qual0[q == 0])   # 0 leaves the quality as it was
    # Everything else is unchanged: the observations, and the vertex messages
    # apart from T_M_I (and their landmark stores apart from position / quality).
    for f in ["obs_vertex", "obs_landmark", "obs_camera", "obs_keypoint", "obs_sigma", "baseframes",
              "camera_intrinsics"]:
        np.testing.assert_array_equal(getattr(back, f), getattr(p, f), err_msg=f)
    for a, b in zip(rec.vertex_protos, rec0.vertex_protos):
        a2, b2 = type(a)(), type(b)()
        a2.CopyFrom(a)
        b2.CopyFrom(b)
        a2.ClearField("T_M_I"), b2.ClearField("T_M_I")
        a2.ClearField("landmark_store"), b2.ClearField("landmark_store")
        assert a2.SerializeToString() == b2.SerializeToString()


def test_errors_are_reported(tmp_path):
    with pytest.raises(RuntimeError, match="cannot read"):
        vimap.NativeMap(str(tmp_path / "missing"))
