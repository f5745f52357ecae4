"""maplab VI-map folders <-> the library's structure-of-arrays problem.

A maplab map is a folder holding `vi_map/` with:
  - `vi_map_metadata`: a VIMapMetadata protobuf listing the files of the map
    (vi-map-serialization.cc:228-251; names at vi-map-serialization.h:25-32);
  - those files: partial VIMap protobufs (missions, vertices<i>, edges<i>,
    landmark_index), gzip-compressed by default
    (proto-serialization-helper.cc:59-77, flag proto_use_compression);
  - `sensors.yaml`: the sensor manager (sensor-manager.cc:186-219): NCAMERA
    sensors with their cameras (ncamera.cc:177-199, camera.cc:253-297) and
    the extrinsics between sensors.

`load_vimap` reads such a folder into a VisualMap, doing what
OptimizationStateBuffer::importStatesOfMissions does
(optimization-state-buffer.cc:153-263): vertex blocks are the Hamilton
q_M_I coefficients with w >= 0 (passive JPL q_IM), baseframe and camera
blocks the inverted rotations (JPL q_GM, q_CI), T_C_I = inverse(T_B_C) of
the NCamera. Observations are the keypoints of every visual frame whose
landmark id is valid and known (vertex.cc:441-470 and visual-frame.proto).
`save_vimap` writes optimized vertex poses and landmark positions back into
the same protobufs and files (Vertex::serialize, vertex.cc:365-407;
Landmark position p_B), everything else unchanged. `write_vimap` writes a
new map folder from a VisualMap.

The protobuf schema is declared here (field numbers and types of
vi-map/proto/vi-map/vi_map.proto, aslam-serialization/visual-frame.proto and
aslam/common/id.proto); fields that are not declared are kept as unknown
fields and written back unchanged.
"""
from __future__ import annotations

import ctypes
import gzip
import os
import shutil
from dataclasses import dataclass, field

import numpy as np
import yaml

from . import abi
from .problem import VisualMap

_D, _I64, _U64, _I32, _BOOL, _STR, _MSG, _BYTES, _U32 = 1, 3, 4, 5, 8, 9, 11, 12, 13
_OPT, _REP = 1, 3

# (message, [(field, number, label, type, message type name or None)]).
# Enums are declared as int32 (same wire format).
_ASLAM = ("aslam/common/id+visual-frame.proto", "aslam.proto", [
    ("Id", [("uint", 1, _REP, _U64, None)]),
    ("VisualFrame", [
        ("id", 1, _OPT, _MSG, ".aslam.proto.Id"), ("timestamp", 2, _OPT, _I64, None),
        ("keypoint_measurements", 3, _REP, _D, None), ("keypoint_measurement_sigmas", 4, _REP, _D, None),
        ("keypoint_descriptors", 5, _OPT, _BYTES, None), ("landmark_ids", 7, _REP, _MSG, ".aslam.proto.Id"),
        ("keypoint_scales", 8, _REP, _D, None), ("is_valid", 9, _OPT, _BOOL, None),
        ("track_ids", 10, _REP, _I32, None), ("descriptor_types", 11, _REP, _I32, None),
        ("keypoint_3d_positions", 12, _REP, _D, None), ("keypoint_time_offsets", 13, _REP, _I32, None)]),
    ("VisualNFrame", [("id", 1, _OPT, _MSG, ".aslam.proto.Id"),
                      ("frames", 2, _REP, _MSG, ".aslam.proto.VisualFrame")]),
])
_VIMAP = ("vi-map/vi_map.proto", "vi_map.proto", [
    ("FileTypeWithPath", [("file_type", 1, _OPT, _I32, None), ("path", 2, _OPT, _STR, None)]),
    ("VIMapMetadata", [("files", 1, _REP, _MSG, ".vi_map.proto.FileTypeWithPath")]),
    ("ViwlsVertex", [
        ("incoming", 1, _REP, _MSG, ".aslam.proto.Id"), ("outgoing", 2, _REP, _MSG, ".aslam.proto.Id"),
        ("T_M_I", 3, _REP, _D, None), ("v_M", 4, _REP, _D, None), ("accel_bias", 5, _REP, _D, None),
        ("gyro_bias", 6, _REP, _D, None), ("n_visual_frame", 7, _OPT, _MSG, ".aslam.proto.VisualNFrame"),
        ("landmark_store", 8, _OPT, _MSG, ".vi_map.proto.LandmarkStore"),
        ("mission_id", 14, _OPT, _MSG, ".aslam.proto.Id")]),
    ("ViwlsEdge", [("from", 1, _OPT, _MSG, ".aslam.proto.Id"), ("to", 2, _OPT, _MSG, ".aslam.proto.Id"),
                   ("imu_timestamps", 3, _REP, _I64, None), ("imu_data", 4, _REP, _D, None),
                   ("mission_id", 5, _OPT, _MSG, ".aslam.proto.Id")]),
    ("Edge", [("viwls", 1, _OPT, _MSG, ".vi_map.proto.ViwlsEdge")]),
    ("Landmark", [("id", 1, _OPT, _MSG, ".aslam.proto.Id"), ("position", 2, _REP, _D, None),
                  ("covariance", 3, _REP, _D, None), ("vertex_ids", 4, _REP, _MSG, ".aslam.proto.Id"),
                  ("keypoint_indices", 5, _REP, _U32, None), ("frame_indices", 6, _REP, _U32, None),
                  ("quality", 7, _OPT, _I32, None), ("feature_type", 9, _OPT, _I32, None)]),
    ("LandmarkStore", [("landmarks", 1, _REP, _MSG, ".vi_map.proto.Landmark")]),
    ("Mission", [("baseframe_id", 1, _OPT, _MSG, ".aslam.proto.Id"),
                 ("root_vertex_id", 2, _OPT, _MSG, ".aslam.proto.Id"), ("backbone", 3, _OPT, _I32, None),
                 ("ncamera_id", 8, _OPT, _MSG, ".aslam.proto.Id"), ("imu_id", 9, _OPT, _MSG, ".aslam.proto.Id")]),
    ("MissionBaseframe", [("T_G_M", 1, _REP, _D, None), ("T_G_M_covariance", 2, _REP, _D, None),
                          ("is_T_G_M_known", 3, _OPT, _BOOL, None)]),
    ("LandmarkToVertexReference", [("vertex_id", 1, _OPT, _MSG, ".aslam.proto.Id"),
                                   ("landmark_id", 2, _OPT, _MSG, ".aslam.proto.Id")]),
    ("VIMap", [
        ("vertex_ids", 1, _REP, _MSG, ".aslam.proto.Id"), ("vertices", 2, _REP, _MSG, ".vi_map.proto.ViwlsVertex"),
        ("edge_ids", 3, _REP, _MSG, ".aslam.proto.Id"), ("edges", 4, _REP, _MSG, ".vi_map.proto.Edge"),
        ("mission_ids", 5, _REP, _MSG, ".aslam.proto.Id"), ("missions", 6, _REP, _MSG, ".vi_map.proto.Mission"),
        ("mission_base_frame_ids", 7, _REP, _MSG, ".aslam.proto.Id"),
        ("mission_base_frames", 8, _REP, _MSG, ".vi_map.proto.MissionBaseframe"),
        ("landmark_index_ids", 9, _REP, _MSG, ".aslam.proto.Id"),
        ("landmark_index", 10, _REP, _MSG, ".vi_map.proto.LandmarkToVertexReference")]),
])

# VIMapFileType (vi-map-metadata.h:18-23) and the file names
# (vi-map-serialization.h:25-32).
FILE_MISSIONS, FILE_VERTICES, FILE_EDGES, FILE_LANDMARK_INDEX = 0, 1, 2, 3
FOLDER, METADATA, SENSORS = "vi_map", "vi_map_metadata", "sensors.yaml"
LANDMARK_BAD, LANDMARK_GOOD = 1, 2      # vi_map.proto Landmark.Quality

_classes = None


def messages():
    """The message classes, built once from the declared schema."""
    global _classes
    if _classes is None:
        from google.protobuf import descriptor_pb2, descriptor_pool, message_factory
        pool = descriptor_pool.DescriptorPool()
        for name, package, msgs in (_ASLAM, _VIMAP):
            fd = descriptor_pb2.FileDescriptorProto(name=name, package=package, syntax="proto2")
            if package == "vi_map.proto":
                fd.dependency.append(_ASLAM[0])
            for mname, fields in msgs:
                m = fd.message_type.add(name=mname)
                for fname, num, label, ftype, tname in fields:
                    f = m.field.add(name=fname, number=num, label=label, type=ftype)
                    if tname:
                        f.type_name = tname
            pool.Add(fd)
        _classes = {}
        for name, package, msgs in (_ASLAM, _VIMAP):
            for mname, _ in msgs:
                _classes[mname] = message_factory.GetMessageClass(pool.FindMessageTypeByName(f"{package}.{mname}"))
    return _classes


# ------------------------------------------------------------------ ids

def id_key(pid) -> tuple:
    """The 128-bit id of an aslam.proto.Id as a tuple (u64[0], u64[1])."""
    u = list(pid.uint)
    return (u[0], u[1]) if len(u) == 2 else (0, 0)


def id_hex(key: tuple) -> str:
    """HashId::hexString (hash-id.h:61-69): the 16 bytes in memory order."""
    return (int(key[0]).to_bytes(8, "little") + int(key[1]).to_bytes(8, "little")).hex()


def hex_key(text: str) -> tuple:
    b = bytes.fromhex(text)
    return (int.from_bytes(b[:8], "little"), int.from_bytes(b[8:], "little"))


def _set_id(pid, key: tuple) -> None:
    del pid.uint[:]
    pid.uint.extend([int(key[0]), int(key[1])])


# ------------------------------------------------------------------ rotations

def _quat_to_matrix(q) -> np.ndarray:
    """Hamilton quaternion (x, y, z, w) -> rotation matrix."""
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _matrix_to_quat(R) -> np.ndarray:
    """Rotation matrix -> Hamilton quaternion (x, y, z, w), w >= 0."""
    t = np.trace(R)
    if t > 0:
        s = 2.0 * np.sqrt(t + 1.0)
        q = np.array([(R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s, 0.25 * s])
    else:
        i = int(np.argmax(np.diag(R)))
        j, k = (i + 1) % 3, (i + 2) % 3
        s = 2.0 * np.sqrt(1.0 + R[i, i] - R[j, j] - R[k, k])
        q = np.zeros(4)
        q[i] = 0.25 * s
        q[j] = (R[j, i] + R[i, j]) / s
        q[k] = (R[k, i] + R[i, k]) / s
        q[3] = (R[k, j] - R[j, k]) / s
    q /= np.linalg.norm(q)
    return -q if q[3] < 0 else q


def _positive(q) -> np.ndarray:
    q = np.asarray(q, dtype=np.float64)
    return -q if q[3] < 0 else q


def _yaml_matrix(node) -> np.ndarray:
    """YAML::convert<Eigen::Matrix> (eigen-yaml-serialization.h): rows, cols, row-major data."""
    return np.asarray(node["data"], dtype=np.float64).reshape(int(node["rows"]), int(node["cols"]))


def _matrix_yaml(M) -> dict:
    M = np.asarray(M, dtype=np.float64)
    return {"rows": int(M.shape[0]), "cols": int(M.shape[1]), "data": [float(v) for v in M.reshape(-1)]}


_DIST_NAMES = {"none": abi.DIST_NONE, "radial-tangential": abi.DIST_RADTAN, "equidistant": abi.DIST_EQUIDISTANT,
               "fisheye": abi.DIST_FISHEYE}  # camera.cc:128-160, 281-287


# ------------------------------------------------------------------ loading

@dataclass
class MapRecord:
    """What `save_vimap` needs: the parsed files and where each state lives."""
    folder: str
    files: list = field(default_factory=list)           # [(file_type, path, VIMap proto, gzip?)]
    vertex_protos: list = field(default_factory=list)   # per vertex index: its ViwlsVertex
    landmark_protos: list = field(default_factory=list) # per landmark index: its Landmark
    vertex_ids: list = field(default_factory=list)
    landmark_ids: list = field(default_factory=list)
    mission_ids: list = field(default_factory=list)


def _read_proto(path: str, cls):
    with open(path, "rb") as f:
        data = f.read()
    zipped = data[:2] == b"\x1f\x8b"
    msg = cls()
    msg.ParseFromString(gzip.decompress(data) if zipped else data)
    return msg, zipped


def _cameras_of(sensors: dict) -> dict:
    """NCAMERA sensors of sensors.yaml: ncamera id -> [(intrinsics (fu fv cu
    cv), distortion type, distortion (4), resolution, T_B_C, camera model,
    xi)]. Camera types pinhole and unified-projection (camera.cc:201-215;
    the latter's intrinsics are xi fu fv cu cv)."""
    out = {}
    for s in sensors.get("sensors") or []:
        if str(s.get("sensor_type", "")).upper() != "NCAMERA":
            continue
        cams = []
        for c in s.get("cameras") or []:
            cam = c["camera"]
            ctype = str(cam.get("type", "pinhole"))
            if ctype not in ("pinhole", "unified-projection"):
                raise ValueError(f"camera type {ctype!r} is not supported (pinhole, unified-projection)")
            intr = _yaml_matrix(cam["intrinsics"]).reshape(-1) if isinstance(cam["intrinsics"], dict) \
                else np.asarray(cam["intrinsics"], dtype=np.float64)
            model, xi = abi.CAMERA_PINHOLE, 0.0
            if ctype == "unified-projection":
                model, xi, intr = abi.CAMERA_UNIFIED_PROJECTION, float(intr[0]), intr[1:]
            dist = cam.get("distortion") or {"type": "none"}
            dname = str(dist.get("type", "none"))
            if dname not in _DIST_NAMES:
                raise ValueError(f"distortion {dname!r} is not supported")
            dpar = np.zeros(4)
            if dname != "none":
                p = dist["parameters"]
                p = _yaml_matrix(p).reshape(-1) if isinstance(p, dict) else np.asarray(p, dtype=np.float64)
                dpar[:len(p)] = p
            if "T_B_C" in c:
                T_B_C = _yaml_matrix(c["T_B_C"])
            else:
                T_B_C = np.linalg.inv(_yaml_matrix(c["T_C_B"]))
            cams.append((intr[:4], _DIST_NAMES[dname], dpar, (int(cam["image_width"]), int(cam["image_height"])),
                         T_B_C, model, xi))
        out[hex_key(str(s["id"]))] = cams
    return out


def load_vimap(folder: str):
    """Reads the maplab map in `folder` (the folder containing `vi_map/`).
    Returns (VisualMap, MapRecord). Vertices come mission by mission along
    the VIWLS edges from each mission's root vertex; the first vertex is
    constant (the gauge fix of the visual problem, as the synthetic maps)."""
    M = messages()
    base = os.path.join(folder, FOLDER)
    meta, _ = _read_proto(os.path.join(base, METADATA), M["VIMapMetadata"])
    rec = MapRecord(folder=folder)
    missions, baseframes, vertices, edges, stores = {}, {}, {}, {}, {}
    for ft in meta.files:
        proto, zipped = _read_proto(os.path.join(base, ft.path), M["VIMap"])
        rec.files.append((ft.file_type, ft.path, proto, zipped))
        for i, m in zip(proto.mission_ids, proto.missions):
            missions[id_key(i)] = m
        for i, b in zip(proto.mission_base_frame_ids, proto.mission_base_frames):
            baseframes[id_key(i)] = b
        for i, v in zip(proto.vertex_ids, proto.vertices):
            vertices[id_key(i)] = v
        for i, e in zip(proto.edge_ids, proto.edges):
            edges[id_key(i)] = e
    with open(os.path.join(base, SENSORS)) as f:
        sensors = yaml.safe_load(f) or {}
    ncams = _cameras_of(sensors)
    # Missions in file order; vertices along the VIWLS chain from the root.
    order, mission_of = [], []
    rec.mission_ids = list(missions.keys())
    for mi, mk in enumerate(rec.mission_ids):
        chain, seen = [], set()
        cur = id_key(missions[mk].root_vertex_id)
        while cur in vertices and cur not in seen:
            seen.add(cur)
            chain.append(cur)
            nxt = None
            for eid in vertices[cur].outgoing:
                e = edges.get(id_key(eid))
                if e is not None and e.HasField("viwls"):
                    nxt = id_key(e.viwls.to)
                    break
            cur = nxt
        rest = [k for k, v in vertices.items() if id_key(v.mission_id) == mk and k not in seen]
        order += chain + rest
        mission_of += [mi] * (len(chain) + len(rest))
    vindex = {k: i for i, k in enumerate(order)}
    rec.vertex_ids = order
    rec.vertex_protos = [vertices[k] for k in order]
    # Cameras: the NCamera of every mission, in mission order.
    cam_base, cams = {}, []
    for mk in rec.mission_ids:
        nk = id_key(missions[mk].ncamera_id)
        if nk not in cam_base:
            if nk not in ncams:
                raise ValueError(f"mission {id_hex(mk)} has no NCAMERA sensor in {SENSORS}")
            cam_base[nk] = len(cams)
            cams += ncams[nk]
    # Landmarks from the landmark stores of the vertices.
    lindex, lm_pos, lm_store = {}, [], []
    for vi, vk in enumerate(order):
        for lm in vertices[vk].landmark_store.landmarks:
            lindex[id_key(lm.id)] = len(lm_pos)
            rec.landmark_ids.append(id_key(lm.id))
            rec.landmark_protos.append(lm)
            lm_pos.append(list(lm.position))
            lm_store.append(vi)
    # Observations.
    ov, ol, oc, kp, sg = [], [], [], [], []
    for vi, vk in enumerate(order):
        v = vertices[vk]
        cb = cam_base[id_key(missions[rec.mission_ids[mission_of[vi]]].ncamera_id)]
        for fi, fr in enumerate(v.n_visual_frame.frames):
            meas = np.asarray(fr.keypoint_measurements, dtype=np.float64).reshape(-1, 2)
            sig = np.asarray(fr.keypoint_measurement_sigmas, dtype=np.float64)
            for k, lid in enumerate(fr.landmark_ids):
                li = lindex.get(id_key(lid))
                if li is None:
                    continue
                ov.append(vi)
                ol.append(li)
                oc.append(cb + fi)
                kp.append(meas[k])
                sg.append(sig[k])
    # State blocks (optimization-state-buffer.cc:161-263).
    verts = np.zeros((len(order), 7))
    for i, v in enumerate(rec.vertex_protos):
        t = np.asarray(v.T_M_I, dtype=np.float64)
        verts[i, :4] = _positive(t[:4])
        verts[i, 4:] = t[4:]
    bfs = np.zeros((len(rec.mission_ids), 7))
    for i, mk in enumerate(rec.mission_ids):
        t = np.asarray(baseframes[id_key(missions[mk].baseframe_id)].T_G_M, dtype=np.float64)
        bfs[i, :4] = _positive([-t[0], -t[1], -t[2], t[3]])  # JPL q_GM = Hamilton q_G_M inverse
        bfs[i, 4:] = t[4:]
    cq = np.zeros((len(cams), 7))
    for i, (_, _, _, _, T_B_C, _, _) in enumerate(cams):
        R_B_C, p_B_C = T_B_C[:3, :3], T_B_C[:3, 3]
        cq[i, :4] = _matrix_to_quat(R_B_C)       # JPL q_CI = Hamilton q_C_I inverse = q(R_B_C)
        cq[i, 4:] = -R_B_C.T @ p_B_C            # C_p_CI
    L = len(lm_pos)
    vmap = VisualMap(
        vertices=verts, baseframes=bfs, landmarks=np.asarray(lm_pos, dtype=np.float64).reshape(L, 3),
        camera_q_CI__C_p_CI=cq,
        camera_intrinsics=np.asarray([c[0] for c in cams], dtype=np.float64).reshape(-1, 4),
        camera_distortion=np.asarray([c[2] for c in cams], dtype=np.float64).reshape(-1, 4),
        camera_distortion_type=np.asarray([c[1] for c in cams], dtype=np.int32),
        camera_resolution=np.asarray([c[3] for c in cams], dtype=np.int32).reshape(-1, 2),
        vertex_mission=np.asarray(mission_of, dtype=np.int32),
        landmark_store_vertex=np.asarray(lm_store, dtype=np.int32),
        obs_vertex=np.asarray(ov, dtype=np.int32), obs_landmark=np.asarray(ol, dtype=np.int32),
        obs_camera=np.asarray(oc, dtype=np.int32), obs_keypoint=np.asarray(kp, dtype=np.float64).reshape(-1, 2),
        obs_sigma=np.asarray(sg, dtype=np.float64),
        vertex_constant=np.eye(1, len(order), dtype=np.uint8).reshape(-1),
        landmark_constant=np.zeros(L, dtype=np.uint8))
    if any(c[5] != abi.CAMERA_PINHOLE for c in cams):
        vmap.camera_model = np.asarray([c[5] for c in cams], dtype=np.int32)
        vmap.camera_xi = np.asarray([c[6] for c in cams], dtype=np.float64)
    return vmap.normalized(), rec


# ------------------------------------------------------------------ saving

def _write_proto(path: str, msg, zipped: bool) -> None:
    data = msg.SerializeToString()
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(gzip.compress(data, mtime=0) if zipped else data)
    os.replace(tmp, path)


def save_vimap(rec: MapRecord, vertices, landmarks, out_folder: str | None = None,
               landmark_quality=None) -> str:
    """Writes optimized vertex poses (P x 7, [q_IM xyzw | p_MI]) and landmark
    positions (L x 3, p_B) into the map read by `load_vimap`, in place or
    into a copy at `out_folder`. landmark_quality (L, MLBA_QUALITY_*;
    optional) updates Landmark.quality for entries that are not UNKNOWN."""
    dst = rec.folder if out_folder is None else out_folder
    if out_folder is not None and os.path.abspath(out_folder) != os.path.abspath(rec.folder):
        if os.path.exists(out_folder):
            shutil.rmtree(out_folder)
        shutil.copytree(rec.folder, out_folder)
    vertices = np.asarray(vertices, dtype=np.float64).reshape(-1, 7)
    landmarks = np.asarray(landmarks, dtype=np.float64).reshape(-1, 3)
    for v, row in zip(rec.vertex_protos, vertices):
        del v.T_M_I[:]
        v.T_M_I.extend([float(x) for x in row])
    for i, (lm, row) in enumerate(zip(rec.landmark_protos, landmarks)):
        del lm.position[:]
        lm.position.extend([float(x) for x in row])
        if landmark_quality is not None and landmark_quality[i] in (LANDMARK_BAD, LANDMARK_GOOD):
            lm.quality = int(landmark_quality[i])
    base = os.path.join(dst, FOLDER)
    for _, path, proto, zipped in rec.files:
        _write_proto(os.path.join(base, path), proto, zipped)
    return dst


def write_vimap(vmap: VisualMap, folder: str, compress: bool = True, seed: int = 7) -> MapRecord:
    """Writes `vmap` as a new maplab map folder (one NCamera per mission
    layout: every camera of the map in one NCamera shared by all missions;
    VIWLS edges without IMU data between consecutive vertices of a mission;
    each landmark stored in its store vertex). Ids are drawn from `seed`.
    Returns the record of the written map (as `load_vimap` would)."""
    M = messages()
    vmap = vmap.normalized()
    rng = np.random.default_rng(seed)
    new_id = lambda: (int(rng.integers(1, 2**63)), int(rng.integers(1, 2**63)))
    P, L, C, nm = vmap.num_vertices, vmap.num_landmarks, vmap.camera_intrinsics.shape[0], vmap.baseframes.shape[0]
    vid = [new_id() for _ in range(P)]
    lid = [new_id() for _ in range(L)]
    mid = [new_id() for _ in range(nm)]
    bid = [new_id() for _ in range(nm)]
    ncam_id, sensor_ids = new_id(), [new_id() for _ in range(C)]
    base = os.path.join(folder, FOLDER)
    os.makedirs(base, exist_ok=True)
    # Missions and baseframes.
    mp = M["VIMap"]()
    for m in range(nm):
        verts_m = np.where(vmap.vertex_mission == m)[0]
        _set_id(mp.mission_ids.add(), mid[m])
        ms = mp.missions.add()
        _set_id(ms.baseframe_id, bid[m])
        _set_id(ms.root_vertex_id, vid[int(verts_m[0])] if len(verts_m) else (0, 0))
        ms.backbone = 0
        _set_id(ms.ncamera_id, ncam_id)
        _set_id(mp.mission_base_frame_ids.add(), bid[m])
        q = vmap.baseframes[m, :4]
        mp.mission_base_frames.add().T_G_M.extend([-q[0], -q[1], -q[2], q[3], *vmap.baseframes[m, 4:]])
    # Edges between consecutive vertices of a mission.
    ep = M["VIMap"]()
    out_edges, in_edges = {}, {}
    for m in range(nm):
        verts_m = np.where(vmap.vertex_mission == m)[0]
        for a, b in zip(verts_m, verts_m[1:]):
            eid = new_id()
            _set_id(ep.edge_ids.add(), eid)
            e = ep.edges.add().viwls
            _set_id(getattr(e, "from"), vid[a])
            _set_id(e.to, vid[b])
            _set_id(e.mission_id, mid[m])
            out_edges[int(a)] = eid
            in_edges[int(b)] = eid
    # Vertices with their frames and landmark stores.
    vp = M["VIMap"]()
    order = np.lexsort((vmap.obs_camera, vmap.obs_vertex))
    by_vertex = {}
    for o in order:
        by_vertex.setdefault(int(vmap.obs_vertex[o]), []).append(int(o))
    lmi = M["VIMap"]()
    for v in range(P):
        _set_id(vp.vertex_ids.add(), vid[v])
        vx = vp.vertices.add()
        if v in in_edges:
            _set_id(vx.incoming.add(), in_edges[v])
        if v in out_edges:
            _set_id(vx.outgoing.add(), out_edges[v])
        vx.T_M_I.extend([float(x) for x in vmap.vertices[v]])
        vx.v_M.extend([0.0] * 3)
        vx.accel_bias.extend([0.0] * 3)
        vx.gyro_bias.extend([0.0] * 3)
        _set_id(vx.mission_id, mid[int(vmap.vertex_mission[v])])
        nf = vx.n_visual_frame
        _set_id(nf.id, new_id())
        frames = [nf.frames.add() for _ in range(C)]
        for f in frames:
            f.is_valid = True
        for o in by_vertex.get(v, []):
            f = frames[int(vmap.obs_camera[o])]
            f.keypoint_measurements.extend([float(x) for x in vmap.obs_keypoint[o]])
            f.keypoint_measurement_sigmas.append(float(vmap.obs_sigma[o]))
            _set_id(f.landmark_ids.add(), lid[int(vmap.obs_landmark[o])])
        store = vx.landmark_store
        for l in np.where(vmap.landmark_store_vertex == v)[0]:
            lm = store.landmarks.add()
            _set_id(lm.id, lid[l])
            lm.position.extend([float(x) for x in vmap.landmarks[l]])
            lm.quality = LANDMARK_GOOD
            ref = lmi.landmark_index.add()
            _set_id(ref.vertex_id, vid[v])
            _set_id(ref.landmark_id, lid[l])
            _set_id(lmi.landmark_index_ids.add(), lid[l])
    files = [(FILE_MISSIONS, "missions", mp), (FILE_VERTICES, "vertices0", vp), (FILE_EDGES, "edges0", ep),
             (FILE_LANDMARK_INDEX, "landmark_index", lmi)]
    meta = M["VIMapMetadata"]()
    for ft, path, proto in files:
        _write_proto(os.path.join(base, path), proto, compress)
        fp = meta.files.add()
        fp.file_type, fp.path = ft, path
    _write_proto(os.path.join(base, METADATA), meta, compress)
    # sensors.yaml: one NCamera with every camera; T_B_C = inverse(T_C_I).
    names = {v: k for k, v in _DIST_NAMES.items()}
    cams = []
    for c in range(C):
        R_C_I = _quat_to_matrix(vmap.camera_q_CI__C_p_CI[c, :4]).T  # JPL q_CI -> R_C_I
        T = np.eye(4)
        T[:3, :3] = R_C_I.T
        T[:3, 3] = -R_C_I.T @ vmap.camera_q_CI__C_p_CI[c, 4:]
        upm = vmap.camera_model is not None and int(vmap.camera_model[c]) == abi.CAMERA_UNIFIED_PROJECTION
        intr = vmap.camera_intrinsics[c]
        if upm:
            intr = np.concatenate([[vmap.camera_xi[c] if vmap.camera_xi is not None else 0.0], intr])
        cam = {"id": id_hex(sensor_ids[c]), "sensor_type": "CAMERA",
               "type": "unified-projection" if upm else "pinhole",
               "image_width": int(vmap.camera_resolution[c, 0]), "image_height": int(vmap.camera_resolution[c, 1]),
               "intrinsics": _matrix_yaml(intr.reshape(-1, 1))}
        dt = int(vmap.camera_distortion_type[c])
        cam["distortion"] = {"type": names[dt]}
        if dt != abi.DIST_NONE:
            n = 1 if dt == abi.DIST_FISHEYE else 4
            cam["distortion"]["parameters"] = _matrix_yaml(vmap.camera_distortion[c, :n].reshape(n, 1))
        cams.append({"camera": cam, "T_B_C": _matrix_yaml(T)})
    sensors = {"sensors": [{"id": id_hex(ncam_id), "sensor_type": "NCAMERA", "cameras": cams}], "extrinsics": []}
    with open(os.path.join(base, SENSORS), "w") as f:
        yaml.safe_dump(sensors, f, sort_keys=False)
    return load_vimap(folder)[1]


# ------------------------------------------------------------------ native reader

class NativeMap:
    """A map folder read by the library's C++ reader (include/maplab_ba_vimap.h),
    the entry a C / C++ host uses; same problem as `load_vimap`."""

    def __init__(self, folder: str):
        self.lib = abi.load_library()
        h = ctypes.c_void_p()
        rc = self.lib.mlba_vimap_load(folder.encode(), ctypes.byref(h))
        if rc != abi.MLBA_OK:
            raise RuntimeError(f"mlba_vimap_load: {self.lib.mlba_vimap_last_error().decode()}")
        self.handle = h

    def close(self) -> None:
        if self.handle:
            self.lib.mlba_vimap_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def problem(self) -> VisualMap:
        d = abi.ProblemDesc()
        abi.check(self.lib.mlba_vimap_describe(self.handle, ctypes.byref(d)))
        P, L, N, C, M = d.num_vertices, d.num_landmarks, d.num_observations, d.num_cameras, d.num_missions

        def arr(ptr, n, dtype, shape=None):
            ct = {np.float64: ctypes.c_double, np.int32: ctypes.c_int32, np.uint8: ctypes.c_uint8}[dtype]
            out = np.zeros(n, dtype=dtype) if n == 0 or not ptr else \
                np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), (n,)).copy()
            return out.reshape(shape) if shape else out

        m = VisualMap(
            vertices=arr(d.vertex_q_IM__M_p_MI, 7 * P, np.float64, (P, 7)),
            baseframes=arr(d.baseframe_q_GM__G_p_GM, 7 * M, np.float64, (M, 7)),
            landmarks=arr(d.landmark_p_B, 3 * L, np.float64, (L, 3)),
            camera_q_CI__C_p_CI=arr(d.camera_q_CI__C_p_CI, 7 * C, np.float64, (C, 7)),
            camera_intrinsics=arr(d.camera_intrinsics, 4 * C, np.float64, (C, 4)),
            camera_distortion=arr(d.camera_distortion, 4 * C, np.float64, (C, 4)),
            camera_distortion_type=arr(d.camera_distortion_type, C, np.int32),
            camera_resolution=arr(d.camera_resolution, 2 * C, np.int32, (C, 2)),
            vertex_mission=arr(d.vertex_mission, P, np.int32),
            landmark_store_vertex=arr(d.landmark_store_vertex, L, np.int32),
            obs_vertex=arr(d.obs_vertex, N, np.int32), obs_landmark=arr(d.obs_landmark, N, np.int32),
            obs_camera=arr(d.obs_camera, N, np.int32), obs_keypoint=arr(d.obs_keypoint, 2 * N, np.float64, (N, 2)),
            obs_sigma=arr(d.obs_sigma, N, np.float64), vertex_constant=arr(d.vertex_constant, P, np.uint8),
            landmark_constant=arr(d.landmark_constant, L, np.uint8))
        if d.camera_model:
            m.camera_model = arr(d.camera_model, C, np.int32)
            m.camera_xi = arr(d.camera_xi, C, np.float64)
        return m.normalized()

    def landmark_quality(self) -> np.ndarray:
        d = abi.ProblemDesc()
        abi.check(self.lib.mlba_vimap_describe(self.handle, ctypes.byref(d)))
        q = np.zeros(d.num_landmarks, dtype=np.uint8)
        abi.check(self.lib.mlba_vimap_landmark_quality(self.handle, q.ctypes.data))
        return q

    def save(self, folder: str, vertices, landmarks, landmark_quality=None) -> None:
        """Writes the states into `folder` (which must hold vi_map/, e.g. a copy)."""
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        l = np.ascontiguousarray(landmarks, dtype=np.float64)
        q = None if landmark_quality is None else np.ascontiguousarray(landmark_quality, dtype=np.uint8)
        rc = self.lib.mlba_vimap_save(self.handle, folder.encode(), v.ctypes.data, l.ctypes.data,
                                      None if q is None else q.ctypes.data)
        if rc != abi.MLBA_OK:
            raise RuntimeError(f"mlba_vimap_save: {self.lib.mlba_vimap_last_error().decode()}")


def optimize_map_folder(folder: str, out_folder: str | None = None, options=None, missions=None):
    """`optv` on a maplab map folder: load, VIMapOptimizer::optimize on the
    device (landmark admission, problem selection and gauge, solve), write
    the optimized vertices, landmarks and landmark qualities back (in place
    or into a copy at out_folder). Returns the OptimizationResult."""
    from .optimizer import VIMapOptimizer
    vmap, rec = load_vimap(folder)
    quality = np.asarray([lm.quality for lm in rec.landmark_protos], dtype=np.uint8)
    res = VIMapOptimizer(options).optimize(vmap, missions=missions, landmark_quality=quality)
    save_vimap(rec, res.vertices, res.landmarks, out_folder=out_folder, landmark_quality=res.landmark_quality)
    return res


def main(argv=None) -> int:
    import argparse
    from .optimizer import default_vi_problem_options
    ap = argparse.ArgumentParser(description="Visual BA (optv) on a maplab map folder, on the GPU.")
    ap.add_argument("map_folder")
    ap.add_argument("--out", default=None, help="write the optimized map here (default: in place)")
    ap.add_argument("--ba_num_iterations", type=int, default=None)
    ap.add_argument("--ba_use_visual_outlier_rejection_solver", type=int, default=None)
    args = ap.parse_args(argv)
    o = default_vi_problem_options()
    if args.ba_num_iterations is not None:
        o.solver.max_num_iterations = args.ba_num_iterations
    if args.ba_use_visual_outlier_rejection_solver is not None:
        o.enable_visual_outlier_rejection = args.ba_use_visual_outlier_rejection_solver
    res = optimize_map_folder(args.map_folder, args.out, o)
    sm = res.summary
    print(f"optv: {sm.num_iterations - 1} iterations, cost {sm.initial_cost:.6e} -> {sm.final_cost:.6e}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
