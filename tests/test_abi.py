"""The C ABI without a GPU: the library loads, exports every entry point the header declares,
and the Python mirrors of the header's structs have the C layout.

No compute entry point is called here (there is no device); the GPU tests do that.
"""

from __future__ import annotations

import ctypes
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "waam_hip.h"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w ]+?\**\s*\b(waam_\w+)\s*\(", text, re.M)))


def test_header_declares_the_step_and_per_kernel_entries():
    names = declared_functions()
    for must in ("waam_abi_version", "waam_last_error", "waam_scratch_bytes", "waam_collide_mrt",
                 "waam_stream", "waam_advect_diffuse_temperature", "waam_step_fused",
                 "waam_deposit_droplet", "waam_telemetry_pool_reduce"):
        assert must in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    from waam_solver_amd._lib import SIGNATURES, exported_symbols, lib

    L = lib()
    missing = [n for n in declared_functions() if not hasattr(L.dll, n)]
    assert not missing, f"declared in {HEADER.name} but not exported: {missing}"
    # every ctypes signature names a declared entry and vice versa
    assert set(exported_symbols()) == set(declared_functions())
    assert set(SIGNATURES) <= set(declared_functions())


def test_abi_version_and_scratch_size():
    from waam_solver_amd._lib import ABI_VERSION, lib

    L = lib()
    assert L.dll.waam_abi_version() == ABI_VERSION
    assert L.scratch_bytes() >= 8 * 1024 * 2
    assert L.dll.waam_last_error() is not None


def test_argument_validation_without_device():
    """Entry points reject bad arguments before touching the device."""
    from waam_solver_amd._lib import WaamError, lib

    L = lib()
    with pytest.raises(WaamError, match="NULL|null"):
        L.call("waam_step_fused", None, None, None)
    with pytest.raises(WaamError):
        L.call("waam_stream", None, None, None, 4, 4, 4, 0, None)
    with pytest.raises(WaamError):
        L.call("waam_stream", 1, 2, 3, 0, 4, 4, 0, None)
    with pytest.raises(WaamError, match="layout"):
        L.call("waam_stream", 1, 2, 3, 4, 4, 4, 7, None)
    with pytest.raises(WaamError, match="range"):
        L.call("waam_timing_begin", 99, 10)


def test_graph_capture_rejects_unsteady_steps_without_device():
    """waam_graph_capture takes only steady steps and a bounded run; the checks come before
    any HIP call."""
    import ctypes

    from waam_solver_amd._lib import WaamError, lib
    from waam_solver_amd.solvers.fused import StepFields, StepParams

    L = lib()
    p, f = StepParams(), StepFields()
    h = ctypes.c_void_p()
    args = (ctypes.addressof(p), ctypes.addressof(f), ctypes.addressof(f))
    with pytest.raises(WaamError, match="NULL"):
        L.call("waam_graph_capture", *args, 8, None)
    with pytest.raises(WaamError, match="steps"):
        L.call("waam_graph_capture", *args, 0, ctypes.byref(h))
    with pytest.raises(WaamError, match="steady"):
        L.call("waam_graph_capture", *args, 8, ctypes.byref(h))  # no pending stream
    p.pending_stream, p.thermal_list_ready, p.force_reset_all = 1, 1, 1
    with pytest.raises(WaamError, match="steady"):
        L.call("waam_graph_capture", *args, 8, ctypes.byref(h))
    assert h.value is None
    with pytest.raises(WaamError, match="NULL"):
        L.call("waam_graph_launch", None, None, None, None, None, None)
    assert L.fns["waam_graph_steps"](None) == 0
    assert L.fns["waam_graph_destroy"](None) == 0


_CODES = {"float": "f", "double": "d", "int": "i", "int32_t": "i", "int64_t": "l",
          "uint32_t": "u"}


def header_signatures() -> dict[str, str]:
    """ctypes argument codes of every declaration in the header, derived from its text."""
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    out = {}
    for name, args in re.findall(r"^\s*(?:const\s+)?[\w ]+?\**\s*\b(waam_\w+)\s*\(([^)]*)\)\s*;",
                                 text, re.M | re.S):
        args = args.strip()
        if args in ("", "void"):
            continue
        codes = []
        for a in args.split(","):
            a = " ".join(a.split())
            codes.append("p" if "*" in a else _CODES[a.rsplit(" ", 1)[0]])
        out[name] = "".join(codes)
    return out


def test_ctypes_signatures_follow_the_header():
    from waam_solver_amd._lib import SIGNATURES

    hdr = header_signatures()
    for name, sig in SIGNATURES.items():
        assert hdr[name] == sig.replace(" ", ""), name


_OFFSETS_SRC = r"""
#include <stddef.h>
#include <stdio.h>
#include "waam_hip.h"
#define F(s, m) printf("%s %s %zu\n", #s, #m, offsetof(s, m));
int main(void) {
%s
printf("size waam_step_fields %zu\n", sizeof(waam_step_fields));
printf("size waam_step_params %zu\n", sizeof(waam_step_params));
printf("size waam_adj_consts %zu\n", sizeof(waam_adj_consts));
printf("size waam_droplet %zu\n", sizeof(waam_droplet));
return 0; }
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_struct_layouts_match_header(tmp_path):
    from waam_solver_amd.adjoint import AdjConsts
    from waam_solver_amd.solvers.fused import DropletC, StepFields, StepParams

    structs = (("waam_step_fields", StepFields), ("waam_step_params", StepParams),
               ("waam_adj_consts", AdjConsts), ("waam_droplet", DropletC))
    lines = []
    for cname, py in structs:
        for field, _ in py._fields_:
            lines.append(f"F({cname}, {field})")
    src = tmp_path / "offsets.c"
    src.write_text(_OFFSETS_SRC.replace("%s\n", "\n".join(lines) + "\n", 1))
    exe = tmp_path / "offsets"
    subprocess.run(["gcc", "-std=c11", "-I", str(HEADER.parent), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    c_off = {}
    for line in out.splitlines():
        a, b, c = line.split()
        c_off[(a, b)] = int(c)
    for cname, py in structs:
        for field, _ in py._fields_:
            assert getattr(py, field).offset == c_off[(cname, field)], (cname, field)
        # same size: the C struct has no member the ctypes one lacks (all members are 4 or 8 B,
        # so a missing one would change the size)
        assert ctypes.sizeof(py) == c_off[("size", cname)], cname
