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


def test_exports_console_entry_points():
    lib = ctypes.CDLL(plugin.PLUGIN_PATH)
    text = open(os.path.join(ROOT, "include", "maplab_ba_plugin.h")).read()
    declared = set(re.findall(r"^(?![/ *#])[^\n(]*?\b(\w+)\s*\(", text, re.M))
    assert declared == set(plugin.PLUGIN_SYMBOLS), declared ^ set(plugin.PLUGIN_SYMBOLS)
    for name in ("createConsolePlugin", "destroyConsolePlugin"):
        assert hasattr(lib, name)


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
