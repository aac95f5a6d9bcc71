"""CPU tests of the C ABI library: it loads, exports every entry point that
include/openph.h declares, and its host-side argument checks answer without
touching a device."""
import ctypes
import os
import re

import numpy as np
import pytest

import openph_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "openph.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(openph[a-z_0-9]*)\s*\(", text, re.M)
    return sorted(set(names))


def test_library_loads_and_exports_header_symbols():
    L = openph_amd.lib()
    names = declared_functions()
    assert "openph" in names and "openph_v2" in names and "openph_pms_device" in names
    for n in names:
        assert hasattr(L, n), n
    nm = os.popen(f"nm -D --defined-only {openph_amd._SO}").read()
    for n in names:
        assert re.search(rf"\bT {n}\b", nm), n


def test_version_and_strerror():
    L = openph_amd.lib()
    assert L.openph_version().startswith(b"openph-mi355x")
    for code in range(0, -10, -1):
        assert L.openph_strerror(code)


def test_unknown_algorithm_rejected_on_host():
    with pytest.raises(openph_amd.OpenPHError) as e:
        openph_amd.ph("not_an_algorithm", [1], [2], 2, 1, [0, 0])
    assert e.value.status == -3


def test_negative_m_rejected_on_host():
    L = openph_amd.lib()
    low = np.zeros(1, np.int32)
    ess = np.zeros(1, np.int32)
    rc = L.openph_v2(b"pms", None, None, -1, 0, None, None,
                     low.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                     ess.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None, None)
    assert rc == -1


def test_mismatched_coo_lengths_rejected():
    with pytest.raises(openph_amd.OpenPHError):
        openph_amd.ph("pms", [1, 2], [2], 3, 1, None)


def test_library_has_gfx950_code_object():
    blob = open(openph_amd._SO, "rb").read()
    assert b"gfx950" in blob


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of the public structs have the C layout (sizes and
    field offsets from a C program compiled against include/openph.h)."""
    from openph_amd.server import Request, RequestInfo, ServerOptions
    structs = {
        "openph_request": Request,
        "openph_options": openph_amd.Options,
        "openph_stats": openph_amd.Stats,
        "openph_tune_row": openph_amd.TuneRow,
        "openph_server_options": ServerOptions,
        "openph_request_info": RequestInfo,
    }
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "openph.h"', "int main(void) {"]
    for cname, py in structs.items():
        src.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            src.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    src.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    rc = os.system(f"gcc -I{os.path.join(ROOT, 'include')} {c} -o {exe}")
    assert rc == 0
    got = dict(line.rsplit(" ", 1) for line in os.popen(str(exe)).read().splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


def test_submit_request_argument_checks_on_host():
    """A request with one output buffer but not the other, unknown flags or
    an unknown kind is refused before the server is looked at."""
    from openph_amd.server import Request, _bind
    L = _bind()
    t = ctypes.c_int64()
    buf = np.zeros(4, np.int32)
    for q in (Request(kind=0, low1=buf.ctypes.data), Request(kind=0, flags=6),
              Request(kind=9)):
        assert L.openph_server_submit_request(None, ctypes.byref(q), ctypes.byref(t)) == -1
    assert L.openph_server_submit_request(None, None, ctypes.byref(t)) == -1


def test_autotune_argument_checks_on_host():
    L = openph_amd.lib()
    best = openph_amd.Options()
    ms = np.zeros(1, np.int32)
    ip = ms.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    # reps < 1, a missing best, a negative report size: refused before any device work
    assert L.openph_autotune_batch_device(1, ip, ip, None, None, None, None, None, None, 0,
                                          ctypes.byref(best), None, 0, None) == -1
    assert L.openph_autotune_batch_device(1, ip, ip, None, None, None, None, None, None, 1,
                                          None, None, 0, None) == -1
    assert L.openph_autotune_batch_device(1, ip, ip, None, None, None, None, None, None, 1,
                                          ctypes.byref(best), None, -1, None) == -1


def test_filtration_sizes_on_host():
    import math
    L = openph_amd.lib()
    m = ctypes.c_int64()
    nnz = ctypes.c_int64()
    for n in (0, 1, 2, 3, 40, 200):
        assert L.openph_vr2_size(n, ctypes.byref(m), ctypes.byref(nnz)) == 0
        assert m.value == n + math.comb(n, 2) + math.comb(n, 3)
        assert nnz.value == 2 * math.comb(n, 2) + 3 * math.comb(n, 3)
    from openph_amd import complexes as C
    for shape in ((1, 1, 1), (2, 3, 4), (8, 8, 8)):
        assert L.openph_cubical_size(*shape, ctypes.byref(m), ctypes.byref(nnz)) == 0
        b = C.cubical_lower_star(np.zeros(shape, np.float32))
        assert m.value == b.m and nnz.value == b.nnz
    assert L.openph_cubical_size(0, 2, 2, ctypes.byref(m), ctypes.byref(nnz)) == -1
    assert L.openph_vr2_size(-1, ctypes.byref(m), ctypes.byref(nnz)) == -1
