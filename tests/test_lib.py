"""The C-ABI library loads and exports every entry point of include/swsph.h
(no compute: this runs without a GPU), and the host-side checks that do not
need a device behave."""
from __future__ import annotations

import ctypes as C

import pytest


def test_library_exports_declared_symbols():
    from swift_amd._lib import declared_symbols, lib
    L = lib()
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_null_arguments_are_errors():
    from swift_amd._lib import lib
    L = lib()
    assert L.swsph_engine_create(None, None) != 0
    assert b"null" in L.swsph_last_error()
    assert L.swsph_pass_batched(None, None) != 0
    assert L.swsph_get_field(None, 0, None) != 0


def test_bad_grid_rejected_before_device():
    """Fewer than 3 top-level cells per axis is an error (the 27-cell stencil
    would visit a cell twice); checked before any device call."""
    from swift_amd import Engine, HydroProps, SwsphError
    with pytest.raises(SwsphError, match="3 top-level cells"):
        Engine(HydroProps(ncell=(2, 4, 4)))


def test_no_cpu_fallback(monkeypatch, tmp_path):
    """A missing library fails loudly instead of falling back to the CPU."""
    import swift_amd._lib as L
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "libswsph.so"))
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(L.SwsphError, match="no CPU fallback"):
        L.lib()


def test_structs_match_header(tmp_path):
    """The ctypes mirrors have the C layout of include/swsph.h: a C program
    compiled against the header prints sizeof/offsetof of every field."""
    import shutil
    import subprocess
    from swift_amd._lib import HEADER_PATH
    from swift_amd.engine import (_FieldDst, _HostParts, _Props, _TuneResult,
                                  PassStats, SchedStats)
    from swift_amd.server import _Info, _Opts
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {"swsph_props": _Props, "swsph_host_parts": _HostParts,
               "swsph_pass_stats": PassStats, "swsph_tune_result": _TuneResult,
               "swsph_server_opts": _Opts, "swsph_request_info": _Info,
               "swsph_field_dst": _FieldDst, "swsph_sched_stats": SchedStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>',
             f'#include "{HEADER_PATH}"', 'int main(void) {']
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ['return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = dict(l.rsplit(" ", 1) for l in
               subprocess.run([str(exe)], check=True, capture_output=True,
                              text=True).stdout.split("\n") if l)
    for cname, py in structs.items():
        assert int(out[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


def test_c_abi_program_builds(tmp_path):
    """The C-ABI test program (tests/c/abi_27cells.c) compiles with gcc
    against include/swsph.h and links with the library and the oracle."""
    import os
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "abi_27cells"
    subprocess.run(["gcc", "-O2", "-std=c11", "-Wall", "-Wextra", "-Werror",
                    "-o", str(exe), os.path.join(root, "tests", "c",
                                                 "abi_27cells.c"),
                    "-L" + os.path.join(root, "swift_amd"),
                    "-L" + os.path.join(root, "oracle"), "-lswsph", "-lorc",
                    "-lm"], check=True)
    assert exe.exists()


def _header_int(path, name):
    import re
    m = re.search(r"constexpr int %s = (\d+);" % name, open(path).read())
    assert m, name
    return int(m.group(1))


def test_tile_tables_hold_every_shape():
    """The tile tables are sized n / kMinTile + ncells + 1: a cell of k
    particles cuts into ceil(k / C) tiles, at most k / C + 1, so every
    accepted tile capacity C >= kMinTile fits (the sparse tiles' list of
    flagged tiles reuses an n-entry buffer: tiles are never empty).  Checked
    against the shapes swsph_set_tile_shape accepts and, by brute force, on
    skewed random cell occupancies."""
    import os
    import re
    import numpy as np
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    kh = os.path.join(root, "swift_amd", "csrc", "sph_kernels.h")
    kmin = _header_int(kh, "kMinTile")
    src = open(os.path.join(root, "swift_amd", "csrc", "sph_engine.cpp")).read()
    body = src[src.index("int swsph_set_tile_shape("):]
    body = body[:body.index("if (!ok)")]
    caps = {int(c) for c in re.findall(r"CI == (\d+)", body)}
    assert caps and min(caps) >= kmin, caps
    rng = np.random.default_rng(0)
    for ncells in (1, 27, 1000):
        for _ in range(20):
            w = rng.pareto(0.7, ncells) + (rng.random(ncells) < 0.3) * 50
            n = int(rng.integers(1, 20000))
            k = np.floor(w / w.sum() * n).astype(np.int64)
            k[0] += n - k.sum()
            cap = n // kmin + ncells + 1
            for C in sorted(caps):
                tiles = int(((k + C - 1) // C).sum())
                assert tiles + 1 <= cap, (ncells, n, C, tiles, cap)
                assert tiles <= max(n, 1)
