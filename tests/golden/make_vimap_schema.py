"""Writes tests/golden/vimap_schema.json: the field table (name, number,
label, type) of every message declared in the reference's VI-map protobuf
files, read from their text:

  aslam_cv2/aslam_cv_common/proto/aslam/common/id.proto
  common/aslam-serialization/proto/aslam-serialization/visual-frame.proto
  map-structure/vi-map/proto/vi-map/vi_map.proto

Run in a checkout that has the reference at /root/reference:
  python tests/golden/make_vimap_schema.py
The output is data (field numbers and types), the form in which the map
reader's declared schema is checked.
"""
from __future__ import annotations

import json
import os
import re

REF = "/root/reference"
FILES = [
    "aslam_cv2/aslam_cv_common/proto/aslam/common/id.proto",
    "common/aslam-serialization/proto/aslam-serialization/visual-frame.proto",
    "map-structure/vi-map/proto/vi-map/vi_map.proto",
]
FIELD = re.compile(r"^(optional|repeated|required)\s+([\w.]+)\s+(\w+)\s*=\s*(\d+)")


def parse(text):
    """Returns (package, {message: {field: [number, label, type]}}, enum names).
    Nested messages and enums are named Outer.Inner."""
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    package = re.search(r"package\s+([\w.]+)\s*;", text).group(1)
    tokens = re.findall(r"[{};]|[^{};]+", text)
    stack, messages, enums = [], {}, {}
    for tok in tokens:
        t = " ".join(tok.split())
        if not t:
            continue
        if t == "{":
            continue
        if t == "}":
            stack.pop()
            continue
        if t == ";":
            continue
        m = re.match(r"^(message|enum|oneof)\s+(\w+)$", t)
        if m:
            kind, name = m.groups()
            stack.append((kind, name))
            full = ".".join(n for k, n in stack if k != "oneof")
            if kind == "message":
                messages[full] = {}
            elif kind == "enum":
                enums[full] = {}
            continue
        e = re.match(r"^(\w+)\s*=\s*(-?\d+)$", t)
        if e and stack and stack[-1][0] == "enum":
            enums[".".join(n for _, n in stack)][e.group(1)] = int(e.group(2))
            continue
        f = FIELD.match(t)
        if f and stack and stack[-1][0] in ("message", "oneof"):
            owner = ".".join(n for k, n in stack if k != "oneof")
            label, typ, name, num = f.groups()
            messages[owner][name] = [int(num), label, typ]
    return package, messages, enums


def main():
    out = {}
    for rel in FILES:
        with open(os.path.join(REF, rel)) as fh:
            package, messages, enums = parse(fh.read())
        for msg, fields in messages.items():
            scope = msg.split(".")
            for name, (num, label, typ) in fields.items():
                # Resolve enum types declared in this file to "enum".
                cands = [".".join(scope[:i] + [typ]) for i in range(len(scope), -1, -1)]
                if any(c in enums for c in cands):
                    fields[name] = [num, label, "enum"]
            out[f"{package}.{msg}"] = {"file": rel, "fields": fields}
        for enum, values in enums.items():
            out[f"{package}.{enum}"] = {"file": rel, "values": values}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vimap_schema.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
        fh.write("\n")
    print(f"wrote {path}: {len(out)} messages")


if __name__ == "__main__":
    main()
