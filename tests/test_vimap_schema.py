"""The VI-map reader declares maplab's protobuf schema in code (no generated
module ships with the reference). This pins every declared field -- number,
label and type -- against tests/golden/vimap_schema.json, the field table
parsed from the reference's own id.proto, visual-frame.proto and vi_map.proto
by tests/golden/make_vimap_schema.py."""
import json
import os

import pytest

from maplab_amd import vimap

SCHEMA = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "vimap_schema.json")))
SCALAR = {"double": vimap._D, "int64": vimap._I64, "uint64": vimap._U64, "int32": vimap._I32,
          "bool": vimap._BOOL, "string": vimap._STR, "bytes": vimap._BYTES, "uint32": vimap._U32,
          "enum": vimap._I32}
LABEL = {"optional": vimap._OPT, "repeated": vimap._REP, "required": 2}


def _declared():
    for _, package, msgs in (vimap._ASLAM, vimap._VIMAP):
        for msg, fields in msgs:
            yield package, msg, fields


def _resolve(package, typ):
    if typ in SCHEMA:
        return typ
    return f"{package}.{typ}"


@pytest.mark.parametrize("package,msg,fields", list(_declared()), ids=lambda x: x if isinstance(x, str) else "")
def test_declared_fields_match_reference_proto(package, msg, fields):
    ref = SCHEMA[f"{package}.{msg}"]["fields"]
    assert len(fields) == len({f[0] for f in fields})
    for name, number, label, typ, type_name in fields:
        assert name in ref, f"{msg}.{name} is not a field of the reference message"
        ref_num, ref_label, ref_type = ref[name]
        assert number == ref_num, (msg, name)
        assert label == LABEL[ref_label], (msg, name)
        if ref_type in SCALAR:
            assert typ == SCALAR[ref_type] and type_name is None, (msg, name)
        else:
            assert typ == vimap._MSG, (msg, name)
            assert type_name == "." + _resolve(package, ref_type), (msg, name)


def test_fields_the_reader_uses_are_declared():
    # The fields the importer reads or writes (optimization-state-buffer.cc
    # and vi-map-serialization.cc use) are all part of the declared schema.
    declared = {f"{p}.{m}": {f[0] for f in fs} for p, m, fs in _declared()}
    need = {
        "vi_map.proto.ViwlsVertex": {"T_M_I", "n_visual_frame", "landmark_store", "mission_id", "outgoing"},
        "vi_map.proto.Landmark": {"id", "position", "vertex_ids", "keypoint_indices", "frame_indices", "quality"},
        "aslam.proto.VisualFrame": {"keypoint_measurements", "keypoint_measurement_sigmas", "landmark_ids"},
        "vi_map.proto.Mission": {"baseframe_id", "root_vertex_id", "ncamera_id"},
        "vi_map.proto.MissionBaseframe": {"T_G_M"},
        "vi_map.proto.VIMap": {"vertex_ids", "vertices", "missions", "mission_base_frames", "landmark_index"},
    }
    for msg, fields in need.items():
        assert fields <= declared[msg], msg


def test_enum_values_of_landmark_quality():
    values = SCHEMA["vi_map.proto.Landmark.Quality"]["values"]
    assert (vimap.LANDMARK_BAD, vimap.LANDMARK_GOOD) == (values["kBad"], values["kGood"])
