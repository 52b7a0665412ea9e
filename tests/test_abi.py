"""The C ABI library loads, exports every declared entry point, and its
host-only logic (budget parsing, init validation) matches the reference.
No GPU compute happens here."""
import ctypes as C
import os
import re

import pytest

from mettle_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mettle_hip.h")


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(mtl_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = abi.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def parse(text):
    out = C.c_uint64(0)
    st = abi.lib().mtl_parse_bytes(text.encode(), C.byref(out))
    return st, out.value


def test_parse_human_sizes_matches_config_rs():
    # src/config.rs:129-135
    assert parse("64") == (0, 64)
    assert parse("64 KB") == (0, 64 * 1024)
    assert parse("1.5 MB") == (0, 1_572_864)
    assert parse("2gb") == (0, 2 * 1024 ** 3)
    assert parse("64 GB") == (0, 64 * 1024 ** 3)
    assert parse("1 TiB") == (0, 1024 ** 4)


@pytest.mark.parametrize("bad", ["", "   ", "GB", "12 XB", "1.2.3 MB", "1 G 2"])
def test_parse_rejects_malformed_sizes(bad):
    st, _ = parse(bad)
    assert st == abi.MTL_ERR_INVALID_BUFFER_CONFIG
    assert abi.last_error().startswith("invalid buffer config:")


def test_init_rejects_zero_budgets_before_touching_the_device():
    L = abi.lib()
    st = L.mtl_buffer_init(b"0", b"1 GB", None, None, 0)
    assert st == abi.MTL_ERR_INVALID_BUFFER_CONFIG
    assert "non-zero cache region" in abi.last_error()
    st = L.mtl_buffer_init(b"1 GB", b"0 KB", None, None, 0)
    assert st == abi.MTL_ERR_INVALID_BUFFER_CONFIG
    assert "non-zero processing region" in abi.last_error()
    st = L.mtl_buffer_init(b"1 GB", b"2 GB", b"0", None, 0)
    assert st == abi.MTL_ERR_INVALID_BUFFER_CONFIG
    assert "non-zero pinned region" in abi.last_error()


def test_budgets_before_init_report_not_initialized():
    L = abi.lib()
    L.mtl_shutdown()
    st = L.mtl_budgets(None, None, None)
    assert st == abi.MTL_ERR_NOT_INITIALIZED
    assert "call gpu_buffer_init before gpu_processing" in abi.last_error()


def test_star_status_decoding():
    L = abi.lib()
    assert L.mtl_star_status(0) == abi.MTL_OK
    assert L.mtl_star_status(2) == abi.MTL_UNSUPPORTED
    assert L.mtl_star_status(16) == abi.MTL_ERR_OVERFLOW
    assert L.mtl_star_status(32) == abi.MTL_UNSUPPORTED
    assert L.mtl_star_status(8) == abi.MTL_ERR_INVALID_ARGUMENT


def test_struct_layouts_match_header():
    # sizes computed from the C declarations (x86-64 SysV)
    assert C.sizeof(abi.Pred) == 24
    assert C.sizeof(abi.StarRow) == 24
    assert C.sizeof(abi.Expr) == 4 + 4 + 64 + 128 + 4 + 4 + 32 + 16


def test_struct_layouts_match_c_compiler(tmp_path):
    """Every ctypes mirror has the size and field offsets the C compiler
    gives the header's struct."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    structs = {"mtl_pred": abi.Pred, "mtl_expr": abi.Expr, "mtl_star_desc": abi.StarDesc,
               "mtl_star_row": abi.StarRow, "mtl_agg_spec": abi.AggSpec,
               "mtl_group_desc": abi.GroupDesc, "mtl_query": abi.Query,
               "mtl_query_result": abi.QueryResult}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void) {']
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        cname, field, value = line.split()
        cls = structs[cname]
        got = C.sizeof(cls) if field == "size" else getattr(cls, field).offset
        assert got == int(value), (cname, field, got, value)
