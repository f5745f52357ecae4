"""The map-optimization console plugin (include/maplab_ba_plugin.h): the
dlsym entry points, plugin id and commands of optimizer-plugin.cc:27-55, flag
parsing and the CommandStatus of user errors (no device work)."""
import ctypes
import os
import re

import numpy as np
import pytest

from cases import tiny_map
from maplab_amd import plugin

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_plugin_entry_points():
    lib = ctypes.CDLL(plugin.PLUGIN_PATH)
    text = open(os.path.join(ROOT, "include", "maplab_ba_plugin.h")).read()
    declared = set(re.findall(r"^(?![/ *#])[^\n(]*?\b(\w+)\s*\(", text, re.M))
    assert declared == set(plugin.PLUGIN_SYMBOLS), declared ^ set(plugin.PLUGIN_SYMBOLS)
    for name in ("mlba_plugin_create", "mlba_plugin_destroy"):
        assert hasattr(lib, name)


def test_does_not_export_the_cpp_console_names():
    # maplab's console dlsyms createConsolePlugin / destroyConsolePlugin and
    # calls them with C++ types; the C library must not answer to those names
    # (a shim against maplab's headers exports them and forwards).
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", plugin.PLUGIN_PATH], capture_output=True, text=True,
                         check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert "createConsolePlugin" not in names and "destroyConsolePlugin" not in names
    assert {"mlba_plugin_create", "mlba_plugin_destroy"} <= names


def test_command_names_stay_valid():
    # A returned name stays valid after later calls (same library, raw pointers).
    c = plugin.MapConsole()
    f = ctypes.CDLL(plugin.PLUGIN_PATH).mlba_plugin_command_name
    f.restype, f.argtypes = ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_int32]
    first = f(c.plugin, 0)
    later = [f(c.plugin, i) for i in range(4)]
    assert ctypes.cast(first, ctypes.c_char_p).value == b"optimize_visual"
    assert first == later[0]
    c.close()


def test_plugin_id_and_commands():
    c = plugin.MapConsole()
    assert c.plugin_id() == "map-optimization"
    assert c.commands() == ["optimize_visual", "optv", "optimize_visual_inertial", "optvi"]
    c.close()


@pytest.mark.parametrize("line, code, text", [
    ("", plugin.CMD_STUPID_USER_ERROR, "empty"),
    ("relaxx", plugin.CMD_STUPID_USER_ERROR, "unknown command"),
    ("optv --ba_num_iterationz=3", plugin.CMD_STUPID_USER_ERROR, "unknown flag"),
    ("optv --ba_num_iterations=three", plugin.CMD_STUPID_USER_ERROR, "bad value"),
    ("optv --ba_use_jacobi_scaling=maybe", plugin.CMD_STUPID_USER_ERROR, "bad value"),
    ("optv", plugin.CMD_STUPID_USER_ERROR, "No map selected"),
    ("optvi", plugin.CMD_UNKNOWN_ERROR, "inertial"),
])
def test_user_errors_without_a_map(line, code, text):
    c = plugin.MapConsole()
    rc, msg = c.run(line)
    assert rc == code and text in msg
    c.close()


@pytest.mark.parametrize("line, text", [
    ("optv --map_mission=0 --map_mission_list=0", "only one of"),
    ("optv --map_mission=7", "is not valid"),
    ("optv --map_mission_list=0,x", "is not valid"),
])
def test_mission_selection_errors(line, text):
    m = tiny_map(np.random.default_rng(0), n_vertices=6, n_landmarks=30)
    c = plugin.MapConsole(m)
    rc, msg = c.run(line)
    assert rc == plugin.CMD_STUPID_USER_ERROR and text in msg
    c.close()
