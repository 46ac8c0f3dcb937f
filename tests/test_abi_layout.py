"""The ctypes mirror of the C ABI against the header itself: every struct
of include/strata_hip.h that the host binds has the same size, and every
field the same offset, as a C compiler lays them out. Runs on the CPU
(gcc compiles a small program over the header; nothing is loaded)."""
import ctypes
import os
import shutil
import subprocess

import pytest

from strata_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "strata_raster": _lib.Raster,
    "strata_kernel": _lib.Kernel,
    "strata_options": _lib.Options,
    "strata_stats": _lib.Stats,
    "strata_tuning": _lib.Tuning,
    "strata_raw": _lib.Raw,
    "strata_source": _lib.Source,
    "strata_sink": _lib.Sink,
    "strata_fold": _lib.Fold,
    "strata_task": _lib.Task,
    "strata_batch_options": _lib.BatchOptions,
    "strata_batch_info": _lib.BatchInfo,
    "strata_serve_config": _lib.ServeConfig,
    "strata_window": _lib.Window,
    "strata_serve_stats": _lib.ServeStats,
}

# ctypes names that differ from the C ones (a Python keyword).
C_NAME = {"async_": "async"}


def c_field(name):
    return C_NAME.get(name, name)


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    d = tmp_path_factory.mktemp("abi")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "strata_hip.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'    printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {c_field(f[0])}));')
    lines += ["    return 0;", "}"]
    src = d / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    layout = {}
    for line in out.splitlines():
        struct, field, value = line.split()
        layout[(struct, field)] = int(value)
    return layout


@pytest.mark.parametrize("cname", list(STRUCTS))
def test_struct_layout(c_layout, cname):
    cls = STRUCTS[cname]
    assert ctypes.sizeof(cls) == c_layout[(cname, "size")], cname
    for f in cls._fields_:
        assert getattr(cls, f[0]).offset == c_layout[(cname, f[0])], (cname, f[0])
