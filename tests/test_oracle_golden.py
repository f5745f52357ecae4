"""Pins the FP64 CPU oracle against the known answers held by the reference's
own tests and against mathematics that fixes the result independently.

Reference tests restated here:
  aslam_cv2/aslam_cv_cameras/test/test-cameras.cc:387-425 (manual projection),
  :491-520 (projection statuses), :88-93 / :136-146 / :190-199 (Jacobians vs
  numdiff); test-distortions.cc:145-150, :191-197 (distortion Jacobians);
  algorithms/ceres-error-terms/test/test_quaternion_parameterization_test.cc:13-31
  (JPL local-parameterization Jacobian vs numdiff of Plus); test-cameras.cc:427-487
  (unified-projection manual projection and statuses).
Test fixtures of the reference: test camera (400, 300, 320, 240, 640x480)
(camera-pinhole.h:251); radtan (-0.28, 0.08, -0.00026, -0.00024)
(distortion-radtan.h:97); equidistant (0.2, 0.01, 0.3, 0.05)
(distortion-equidistant.h:109); fisheye keypoint (399.139364932153,
200.430317533923) at w = 1 (test_visual_term_test.cc:300-301).
"""
import numpy as np
import pytest

import independent as ind

TEST_INTR = np.array([400.0, 300.0, 320.0, 240.0])
RADTAN = np.array([-0.28, 0.08, -0.00026, -0.00024])
EQUI = np.array([0.2, 0.01, 0.3, 0.05])
NONE, RAD, EQD, FOV = 0, 1, 2, 3
FISHEYE = np.array([0.95, 0.0, 0.0, 0.0])


def test_manual_projection_without_distortion(oracle):
    # test-cameras.cc:387-425: exact to 1e-15.
    st, kp, jp, _, _ = oracle.project(NONE, TEST_INTR, None, 640, 480, [1.0, 1.2, 10.0])
    assert st == 0
    np.testing.assert_allclose(kp, [400 * 0.1 + 320, 300 * 0.12 + 240], atol=1e-15, rtol=0)
    # Jacobian of (fu x/z + cu, fv y/z + cv): closed form.
    np.testing.assert_allclose(jp, [[40.0, 0, -4.0], [0, 30.0, -3.6]], rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("p, status", [((1, 1, 10), 0), ((1, 1, -10), 2), ((1000, 1000, 10), 1), ((1, 1, 0.0), 3)])
def test_projection_status(oracle, p, status):
    # test-cameras.cc:491-520
    st, *_ = oracle.project(NONE, TEST_INTR, None, 640, 480, p)
    assert st == status


@pytest.mark.parametrize("dist_type, dist", [(RAD, RADTAN), (EQD, EQUI), (FOV, FISHEYE)])
def test_distortion_matches_textbook_model(oracle, dist_type, dist):
    rng = np.random.default_rng(3)
    f = {RAD: ind.distort_radtan, EQD: ind.distort_equidistant, FOV: ind.distort_fov}[dist_type]
    for _ in range(50):
        x, y = rng.uniform(-0.8, 0.8, 2)
        pt, _ = oracle.distort(dist_type, dist, [x, y])
        np.testing.assert_allclose(pt, f(dist, x, y), rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("dist_type, dist", [(RAD, RADTAN), (EQD, EQUI), (FOV, FISHEYE)])
def test_distortion_keypoint_jacobian_numdiff(oracle, dist_type, dist):
    # test-distortions.cc:145-150 at (0.3, -0.2), tolerance 1e-4 there; the
    # 5-point stencil allows a much tighter bound.
    for point in ([0.3, -0.2], [0.05, 0.4], [-0.6, 0.1]):
        _, jd = oracle.distort(dist_type, dist, point)
        f = lambda d: oracle.distort(dist_type, dist, np.asarray(point) + d)[0]
        np.testing.assert_allclose(jd, ind.numdiff(f, 2, 1e-5), rtol=1e-8, atol=1e-10)


def test_equidistant_optical_centre_special_case(oracle):
    # distortion-equidistant.cc:38-43: point unchanged, Jacobian zero.
    pt, jd = oracle.distort(EQD, EQUI, [1e-12, -1e-12])
    np.testing.assert_array_equal(pt, [1e-12, -1e-12])
    np.testing.assert_array_equal(jd, np.zeros((2, 2)))


@pytest.mark.parametrize("dist_type, dist", [(NONE, np.zeros(4)), (RAD, RADTAN), (EQD, EQUI), (FOV, FISHEYE)])
def test_projection_jacobians_numdiff(oracle, dist_type, dist):
    # test-cameras.cc:88-93 (point, 1e-5), :136-146 (intrinsics 1e-6),
    # :190-199 (distortion 1e-4): random points at depth ~10.
    rng = np.random.default_rng(7)
    for _ in range(10):
        p = np.array([rng.uniform(-3, 3), rng.uniform(-2, 2), 10.0])
        st, kp, jp, ji, jd = oracle.project(dist_type, TEST_INTR, dist, 640, 480, p)
        f_p = lambda d: oracle.project(dist_type, TEST_INTR, dist, 640, 480, p + d, False)[1]
        f_i = lambda d: oracle.project(dist_type, TEST_INTR + d, dist, 640, 480, p, False)[1]
        np.testing.assert_allclose(jp, ind.numdiff(f_p, 3, 1e-4), rtol=1e-7, atol=1e-8)
        np.testing.assert_allclose(ji, ind.numdiff(f_i, 4, 1e-4), rtol=1e-7, atol=1e-8)
        if dist_type != NONE:
            f_d = lambda d: oracle.project(dist_type, TEST_INTR, dist + d, 640, 480, p, False)[1]
            np.testing.assert_allclose(jd, ind.numdiff(f_d, 4, 1e-5), rtol=1e-6, atol=1e-7)


def test_fisheye_keypoint_of_the_visual_term_test(oracle):
    # test_visual_term_test.cc:300-301 (VisualErrorTermNonZeroDistortionOptimization):
    # keypoint (399.139364932153, 200.430317533923) generated in Matlab for
    # w = 1 with intrinsics (100, 100, 320, 240), 640 x 480, landmark at the
    # origin of an identity base pose, observer at M_p_MI = (-1, 0.5, -1)
    # with identity orientation and extrinsics.
    intr = np.array([100.0, 100.0, 320.0, 240.0])
    st, kp, *_ = oracle.project(FOV, intr, [1.0, 0, 0, 0], 640, 480, [1.0, -0.5, 1.0])
    assert st == 0
    np.testing.assert_allclose(kp, [399.139364932153, 200.430317533923], rtol=0, atol=1e-9)
    ident = np.array([0.0, 0, 0, 1, 0, 0, 0])
    observer = np.array([0.0, 0, 0, 1, -1.0, 0.5, -1.0])
    params = [np.zeros(3), ident, ident, ident, observer, ident[:4], np.zeros(3), intr, np.array([1.0, 0, 0, 0])]
    failed, r, _ = oracle.evaluate_visual([399.139364932153, 200.430317533923], 0.7, 1, FOV, 640, 480, params)
    assert not failed
    np.testing.assert_allclose(r, [0.0, 0.0], atol=2e-9)


def test_fisheye_limits_follow_the_reference(oracle):
    # distortion-fisheye.cc:31-42 / :50-56 / :99-104: w^2 < 1e-5 leaves the
    # point and an identity Jacobian; r^2 < 1e-5 scales by 2 tan(w/2) / w and
    # the parameter Jacobian's two rows are that factor's w-derivative.
    pt, jd = oracle.distort(FOV, [1e-3, 0, 0, 0], [0.3, -0.2])
    np.testing.assert_array_equal(pt, [0.3, -0.2])
    np.testing.assert_array_equal(jd, np.eye(2))
    w = 0.95
    f = 2 * np.tan(w / 2) / w
    pt, jd = oracle.distort(FOV, [w, 0, 0, 0], [1e-3, 2e-3])
    np.testing.assert_allclose(pt, [1e-3 * f, 2e-3 * f], rtol=1e-15)
    np.testing.assert_allclose(jd, f * np.eye(2), rtol=1e-15)
    # Parameter Jacobian outside the limit: numdiff in w; inside: the
    # closed-form derivative of 2 tan(w/2)/w, the same in both rows.
    # (At r = 1/2 the factor is 1 for every w, so the point has r != 1/2.)
    p = np.array([0.6, -0.2, 1.0])
    _, _, _, _, jdist = oracle.project(FOV, TEST_INTR, [w, 0, 0, 0], 640, 480, p)
    f_w = lambda d: oracle.project(FOV, TEST_INTR, [w + d[0], 0, 0, 0], 640, 480, p, False)[1]
    np.testing.assert_allclose(jdist[:, :1], ind.numdiff(f_w, 1, 1e-5), rtol=1e-7)
    np.testing.assert_array_equal(jdist[:, 1:], 0.0)
    _, _, _, _, jdist = oracle.project(FOV, TEST_INTR, [w, 0, 0, 0], 640, 480, [1e-3, 1e-3, 1.0])
    df = ((1 + np.tan(w / 2) ** 2) * w - 2 * np.tan(w / 2)) / (w * w)
    np.testing.assert_allclose(jdist[:, 0], [TEST_INTR[0] * df, TEST_INTR[1] * df], rtol=1e-12)


UPM = 1
UPM_TEST = np.array([400.0, 300.0, 320.0, 240.0])  # createTestCamera: xi 0.9, (400, 300, 320, 240), 640x480
UPM_XI = 0.9


def test_unified_projection_manual_projection_and_statuses(oracle):
    # test-cameras.cc:427-466: keypoint of (1, 1.2, 10) to 1e-15.
    p = np.array([1.0, 1.2, 10.0])
    rz = 1.0 / (p[2] + UPM_XI * np.linalg.norm(p))
    st, kp, *_ = oracle.project_camera(UPM, UPM_XI, NONE, UPM_TEST, None, 640, 480, p)
    assert st == 0
    np.testing.assert_allclose(kp, [400 * p[0] * rz + 320, 300 * p[1] * rz + 240], atol=1e-15, rtol=0)
    # :468-487: visible, outside the image box, invalid (z <= -fov(xi) |p|).
    for q, status in [((1, 1, 10), 0), ((0.0, 10, -10), 1), ((0, 0, -10), 3)]:
        st, kp, *_ = oracle.project_camera(UPM, UPM_XI, NONE, UPM_TEST, None, 640, 480, q)
        assert st == status, q
    st, kp, *_ = oracle.project_camera(UPM, UPM_XI, NONE, UPM_TEST, None, 640, 480, [0, 0, -10])
    np.testing.assert_array_equal(kp, [0.0, 0.0])


def test_unified_projection_closed_forms(oracle):
    # xi = 0 is the pinhole; the optical axis maps to (cu, cv) for every xi;
    # with xi = 1 a ray at 90 degrees lands at unit normalized radius; the
    # fov limit uses 1 / xi above xi = 1.
    rng = np.random.default_rng(5)
    for _ in range(20):
        p = np.array([rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(3, 10)])
        for dt, dist in [(NONE, np.zeros(4)), (RAD, RADTAN), (FOV, FISHEYE)]:
            a = oracle.project_camera(UPM, 0.0, dt, UPM_TEST, dist, 640, 480, p)
            b = oracle.project(dt, UPM_TEST, dist, 640, 480, p)
            np.testing.assert_allclose(a[1], b[1], rtol=1e-14)
            np.testing.assert_allclose(a[2], b[2], rtol=1e-12, atol=1e-12)
            np.testing.assert_allclose(a[3], b[3], rtol=1e-14, atol=1e-15)
    for xi in (0.3, 0.9, 1.7):
        _, kp, *_ = oracle.project_camera(UPM, xi, NONE, UPM_TEST, None, 640, 480, [0, 0, 4.0])
        np.testing.assert_allclose(kp, [320, 240], atol=1e-13)
    _, kp, *_ = oracle.project_camera(UPM, 1.0, NONE, UPM_TEST, None, 640, 480, [2.5, 0, 0])
    np.testing.assert_allclose(kp, [400 + 320, 240], rtol=1e-15)
    # xi = 1.6: fov parameter 1/1.6; z = -0.6 |p| is valid, z = -0.65 |p| is not.
    for zf, status in [(-0.6, 1), (-0.65, 3)]:
        q = np.array([np.sqrt(1 - zf * zf), 0.0, zf])
        st, *_ = oracle.project_camera(UPM, 1.6, NONE, UPM_TEST, None, 640, 480, q)
        assert st == status, zf
    # The independent sphere model.
    for _ in range(20):
        p = rng.normal(size=3) + [0, 0, 3]
        for dt, dist in [(NONE, np.zeros(4)), (RAD, RADTAN), (EQD, EQUI), (FOV, FISHEYE)]:
            _, kp, *_ = oracle.project_camera(UPM, 1.3, dt, UPM_TEST, dist, 640, 480, p, False)
            np.testing.assert_allclose(kp, ind.project(dt, UPM_TEST, dist, p, xi=1.3), rtol=1e-13)


@pytest.mark.parametrize("dist_type, dist", [(NONE, np.zeros(4)), (RAD, RADTAN), (EQD, EQUI), (FOV, FISHEYE)])
@pytest.mark.parametrize("xi", [0.3, 0.9, 1.4])
def test_unified_projection_jacobians_numdiff(oracle, dist_type, dist, xi):
    rng = np.random.default_rng(11)
    for _ in range(8):
        p = np.array([rng.uniform(-3, 3), rng.uniform(-2, 2), rng.uniform(2, 10)])
        st, kp, jp, ji, jd, jx = oracle.project_camera(UPM, xi, dist_type, UPM_TEST, dist, 640, 480, p)
        f_p = lambda d: oracle.project_camera(UPM, xi, dist_type, UPM_TEST, dist, 640, 480, p + d, False)[1]
        f_i = lambda d: oracle.project_camera(UPM, xi, dist_type, UPM_TEST + d, dist, 640, 480, p, False)[1]
        f_x = lambda d: oracle.project_camera(UPM, xi + d[0], dist_type, UPM_TEST, dist, 640, 480, p, False)[1]
        np.testing.assert_allclose(jp, ind.numdiff(f_p, 3, 1e-4), rtol=1e-7, atol=1e-8)
        np.testing.assert_allclose(ji, ind.numdiff(f_i, 4, 1e-4), rtol=1e-7, atol=1e-8)
        np.testing.assert_allclose(jx, ind.numdiff(f_x, 1, 1e-6)[:, 0], rtol=1e-7, atol=1e-7)
        if dist_type != NONE:
            f_d = lambda d: oracle.project_camera(UPM, xi, dist_type, UPM_TEST, dist + d, 640, 480, p, False)[1]
            np.testing.assert_allclose(jd, ind.numdiff(f_d, 4, 1e-5), rtol=1e-6, atol=1e-7)


def test_rotation_is_transposed_hamilton(oracle):
    # quaternion-math-inl.h:7-27 is the transpose of the Hamilton matrix.
    rng = np.random.default_rng(1)
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        np.testing.assert_allclose(oracle.rotation_jpl(q), ind.R_jpl(q), atol=1e-15)


def test_quaternion_parameterization_jacobian(oracle):
    # test_quaternion_parameterization_test.cc:13-31: q = (1, 1, 0, 1) / |.|,
    # Jacobian of Plus at delta = 0 equal to numdiff (1e-10 in the reference).
    q = np.array([1.0, 1.0, 0.0, 1.0]) / np.sqrt(3.0)
    J = oracle.quat_jacobian(q)
    f = lambda d: oracle.quat_plus(q, d)
    np.testing.assert_allclose(J, ind.numdiff(f, 3, 1e-5), atol=1e-10)


def test_quaternion_plus_properties(oracle):
    rng = np.random.default_rng(2)
    for _ in range(200):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        q *= np.sign(q[3])
        d = rng.normal(size=3) * rng.choice([1e-3, 0.1, 0.6, 2.0])
        out = oracle.quat_plus(q, d)
        assert abs(np.linalg.norm(out) - 1) < 1e-15 and out[3] >= 0
        np.testing.assert_allclose(out, ind.jpl_plus(q, d), atol=1e-14)
    # Large-angle branch is the exact rotation by |d|: composing with -d undoes it.
    q = np.array([0.0, 0.0, 0.0, 1.0])
    d = np.array([0.3, -0.5, 0.4])
    back = oracle.quat_plus(oracle.quat_plus(q, d), -d)
    np.testing.assert_allclose(back, q, atol=1e-14)


def test_huber_and_cauchy_closed_forms(oracle):
    a = 1.5
    for s in (0.0, 0.5, 2.25, 3.0, 100.0):
        rho = oracle.loss(0, a, s)
        if s <= a * a:
            np.testing.assert_allclose(rho, [s, 1.0, 0.0])
        else:
            np.testing.assert_allclose(rho, [2 * a * np.sqrt(s) - a * a, a / np.sqrt(s), -a / np.sqrt(s) / (2 * s)])
        rc = oracle.loss(1, a, s)
        b = a * a
        np.testing.assert_allclose(rc, [b * np.log1p(s / b), 1 / (1 + s / b), -1 / b / (1 + s / b) ** 2], rtol=1e-14)
        # rho' and rho'' are the derivatives of rho
        for lt in (0, 1):
            if lt == 0 and abs(s - b) < 1e-3:
                continue
            h = 1e-6 * max(1.0, s)
            d1 = (oracle.loss(lt, a, s + h)[0] - oracle.loss(lt, a, max(s - h, 0))[0]) / (s + h - max(s - h, 0))
            assert abs(d1 - oracle.loss(lt, a, s)[1]) < 1e-6
