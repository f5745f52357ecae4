"""The C ABI library loads and exports every declared symbol, the option
defaults are maplab's flags, and invalid problems are rejected on the host
before any device work (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from cases import tiny_map
from maplab_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return set(re.findall(r"^\S[^\n(]*?\b(mlba_\w+)\s*\(", text, re.M))


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    syms = declared("maplab_ba.h")
    assert syms == set(abi.LIB_SYMBOLS), syms ^ set(abi.LIB_SYMBOLS)
    for s in syms:
        assert hasattr(lib, s), s
    assert lib.mlba_abi_version() == 1


def test_library_exports_every_optimizer_symbol():
    lib = abi.load_library()
    syms = declared("maplab_ba_optimizer.h")
    assert syms == set(abi.OPTIMIZER_SYMBOLS), syms ^ set(abi.OPTIMIZER_SYMBOLS)
    for s in syms:
        assert hasattr(lib, s), s


def test_library_exports_every_inertial_symbol():
    lib = abi.load_library()
    syms = declared("maplab_ba_inertial.h")
    assert syms == set(abi.INERTIAL_SYMBOLS), syms ^ set(abi.INERTIAL_SYMBOLS)
    for s in syms:
        assert hasattr(lib, s), s


def test_library_exports_every_learned_preconditioner_symbol():
    lib = abi.load_library()
    syms = declared("maplab_ba_learned.h")
    assert syms == set(abi.LEARNED_SYMBOLS), syms ^ set(abi.LEARNED_SYMBOLS)
    for s in syms:
        assert hasattr(lib, s), s
    assert lib.mlba_gnn_num_params() == 9 * 64 + 64 + 3 * (64 * 64 + 64 + 3 + 1 + 128 * 64 + 64) + 64 * 2 + 2


def test_library_exports_every_vimap_symbol():
    lib = abi.load_library()
    syms = declared("maplab_ba_vimap.h")
    assert syms == set(abi.VIMAP_SYMBOLS), syms ^ set(abi.VIMAP_SYMBOLS)
    for s in syms:
        assert hasattr(lib, s), s


def test_synth_library_exports_every_declared_symbol():
    lib = abi.load_synth()
    syms = declared("maplab_ba_synth.h")
    assert syms == set(abi.SYNTH_SYMBOLS)
    for s in syms:
        assert hasattr(lib, s), s


def test_struct_layouts_match_header():
    # Field offsets are what the C compiler produces for the declarations.
    assert ctypes.sizeof(abi.ProblemDesc) == 24 + 19 * 8
    assert ctypes.sizeof(abi.IterationSummary) == 16 + 8 * 8
    assert abi.Summary.kernel_launches.offset == ctypes.sizeof(abi.Summary) - 4 * abi.NUM_KERNELS
    assert abi.Summary.kernel_ms.offset == abi.Summary.iterations.offset + 8


@pytest.mark.parametrize("struct, cname, header", [
    (abi.ProblemDesc, "mlba_problem_desc", "maplab_ba.h"),
    (abi.Options, "mlba_options", "maplab_ba.h"),
    (abi.IterationSummary, "mlba_iteration_summary", "maplab_ba.h"),
    (abi.Summary, "mlba_summary", "maplab_ba.h"),
    (abi.KernelTimes, "mlba_kernel_times", "maplab_ba.h"),
    (abi.Schedule, "mlba_schedule", "maplab_ba.h"),
    (abi.CalibrationOptions, "mlba_calibration_options", "maplab_ba.h"),
    (abi.SynthConfig, "mlba_synth_config", "maplab_ba_synth.h"),
    (abi.QualityOptions, "mlba_quality_options", "maplab_ba.h"),
    (abi.ViProblemOptions, "mlba_vi_problem_options", "maplab_ba_optimizer.h"),
    (abi.ProblemSelection, "mlba_problem_selection", "maplab_ba_optimizer.h"),
    (abi.InertialProblem, "mlba_inertial_problem", "maplab_ba_inertial.h"),
])
def test_struct_layout_matches_compiled_header(tmp_path, struct, cname, header):
    # The C compiler's sizeof / offsetof of every field equals the ctypes mirror.
    fields = [f[0] for f in struct._fields_]
    src = tmp_path / "layout.c"
    src.write_text(
        "#include <stdio.h>\n#include <stddef.h>\n#include \"%s\"\nint main(void) {\n"
        "  printf(\"%%zu\\n\", sizeof(%s));\n" % (header, cname)
        + "".join("  printf(\"%%zu\\n\", offsetof(%s, %s));\n" % (cname, f) for f in fields)
        + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert out[0] == ctypes.sizeof(struct)
    assert out[1:] == [getattr(struct, f).offset for f in fields]


def test_quality_and_problem_defaults_are_maplab_flags():
    # landmark-quality-metrics.cc:11-33; vi-optimization-builder.cc:10-60.
    q = abi.default_quality_options()
    assert (q.min_observers, q.min_observation_angle_deg) == (4, 5.0)
    assert (q.max_distance_from_closest_observer, q.min_distance_from_closest_observer) == (40.0, 0.05)
    assert q.max_reprojection_error_px == -1.0
    from maplab_amd.optimizer import default_vi_problem_options
    o = default_vi_problem_options()
    assert (o.add_visual_constraints, o.fix_landmark_positions, o.fix_vertices) == (1, 0, 0)
    assert (o.min_landmarks_per_frame, o.enable_visual_outlier_rejection) == (0, 1)
    assert o.solver.max_num_iterations == 30 and o.landmark_quality.min_observers == 4


def test_option_defaults_are_maplab_flags():
    # algorithms/map-optimization/src/solver-options.cc:3-52 and
    # outlier-rejection-solver.cc:9-18.
    o = abi.default_options()
    assert o.max_num_iterations == 30
    assert o.use_jacobi_scaling == 1
    assert o.initial_trust_region_radius == 1e5
    assert o.max_trust_region_radius == 1e20
    assert o.min_trust_region_radius == 1e-32
    assert o.function_tolerance == 1e-12
    assert o.gradient_tolerance == 1e-10
    assert o.parameter_tolerance == 1e-8
    assert o.max_solver_time_in_seconds == 1e6
    assert o.reject_outliers_every_n_iters == 3
    assert o.use_reprojection_error == 0
    assert o.max_reprojection_error_px == 50
    assert o.loss_type == abi.LOSS_MAPLAB_HUBER


def _create(m):
    lib = abi.load_library()
    d = m.desc()
    h = ctypes.c_void_p()
    rc = lib.mlba_create(ctypes.byref(d), ctypes.byref(h))
    return rc, lib.mlba_last_error(None).decode()


@pytest.mark.parametrize("mutate, code, text", [
    (lambda m: m.obs_sigma.__setitem__(0, 0.0), abi.MLBA_EINVAL, "sigma"),
    (lambda m: m.obs_landmark.__setitem__(0, 10**6), abi.MLBA_EINVAL, "obs_landmark"),
    (lambda m: m.obs_vertex.__setitem__(0, -1), abi.MLBA_EINVAL, "obs_vertex"),
    (lambda m: m.vertices.__setitem__((1, 3), -0.5), abi.MLBA_ESTATE, "quaternion"),
    (lambda m: m.landmark_store_vertex.__setitem__(0, 99), abi.MLBA_EINVAL, "store_vertex"),
    (lambda m: m.camera_distortion_type.__setitem__(0, 7), abi.MLBA_EINVAL, "distortion"),
    (lambda m: (m.camera_distortion_type.__setitem__(0, 3), m.camera_distortion.__setitem__((0, 0), 0.3)),
     abi.MLBA_EINVAL, "fisheye"),
    (lambda m: setattr(m, "camera_model", np.array([5], np.int32)), abi.MLBA_EINVAL, "camera model"),
    (lambda m: (setattr(m, "camera_model", np.array([1], np.int32)), setattr(m, "camera_xi", np.array([-0.1]))),
     abi.MLBA_EINVAL, "unified-projection"),
])
def test_invalid_problem_is_rejected(mutate, code, text):
    # The reference CHECK-fails on these (visual-error-term.h:38,
    # quaternion-param-jpl.cc:44); the library returns a status instead.
    m = tiny_map(np.random.default_rng(0), n_vertices=4, n_landmarks=10)
    mutate(m)
    rc, msg = _create(m)
    assert rc == code and text in msg, (rc, msg)


def _edges(**kw):
    from maplab_amd.inertial import InertialEdges
    base = dict(vertices=np.tile([0, 0, 0, 1.0, 0, 0, 0], (2, 1)), gyro_bias=np.zeros((2, 3)),
                velocity=np.zeros((2, 3)), accel_bias=np.zeros((2, 3)), edge_from=[0], edge_to=[1],
                edge_sample_begin=[0, 3], imu=np.tile([0, 0, 9.81, 0, 0, 0], (3, 1)),
                timestamps_ns=np.array([0, 5000000, 10000000], dtype=np.int64))
    base.update(kw)
    return InertialEdges(**base)


@pytest.mark.parametrize("kw, text", [
    (dict(noise=(0.0, 1e-4, 1e-2, 1e-3)), "sigmas"),
    (dict(edge_sample_begin=[0, 0]), "without IMU data"),
    (dict(timestamps_ns=np.array([0, 5000000, 4000000], dtype=np.int64)), "not properly ordered"),
    (dict(edge_to=[5]), "out of range"),
])
def test_inertial_rejects_invalid_edges(kw, text):
    # The reference CHECK-fails on these (inertial-error-term.h:84-89,
    # inertial-error-term.cc:36-37); the library returns EINVAL before any
    # device work.
    from maplab_amd import inertial
    with pytest.raises(RuntimeError, match=text):
        inertial.evaluate(_edges(**kw))


def test_solve_many_rejects_bad_arguments_without_touching_the_device():
    lib = abi.load_library()
    assert lib.mlba_solve_many(None, None, None, None, 0, 0, 0) == abi.MLBA_OK
    assert lib.mlba_solve_many(None, None, None, None, -1, 0, 0) == abi.MLBA_EINVAL
    assert lib.mlba_solve_many(None, None, None, None, 2, 0, 0) == abi.MLBA_EINVAL
    handles = (ctypes.c_void_p * 2)(None, None)
    o = abi.default_options()
    opts = (ctypes.POINTER(abi.Options) * 2)(ctypes.pointer(o), ctypes.pointer(o))
    assert lib.mlba_solve_many(handles, opts, None, None, 2, 0, 0) == abi.MLBA_EINVAL


def test_schedule_calls_reject_null_arguments_without_touching_the_device():
    lib = abi.load_library()
    o = abi.default_options()
    sc = abi.Schedule(item_band_rows=4, item_waves=4, xcd_mapping=1)
    assert lib.mlba_autotune(None, ctypes.byref(o), 5, ctypes.byref(sc)) == abi.MLBA_EINVAL
    assert lib.mlba_set_schedule(None, ctypes.byref(sc)) == abi.MLBA_EINVAL
    assert lib.mlba_get_schedule(None, ctypes.byref(sc)) == abi.MLBA_EINVAL
