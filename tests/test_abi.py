"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(",
                             text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "sizeof", "defined"):
                names.add(name)
    return names


def test_header_declares_drop_ins():
    names = _declared()
    assert {"letkf_core_", "letkf_core_batch_", "dida_letkf_sweep"} <= names


def test_library_exports_every_declared_symbol(dl):
    lib = ctypes.CDLL(dl.LIB_PATH)
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_declared()) == sorted(dl.EXPORTED)


def test_synth_library_loads(dl):
    s = dl.synth_lib()
    cfg = dl.synth_cfg(8, 16, 2, 4, 10, 1)
    assert s.dida_synth_scratch_bytes(ctypes.byref(cfg)) > 0


def test_params_default_and_errors(dl):
    p = dl.default_params()
    assert p.max_obs == 10000 and p.gc_filter == 1 and p.jacobi_max_sweeps > 0
    assert p.form == dl.FORM_PLAIN and p.gram == dl.GRAM_FP64
    # argument checking happens before any device call
    bad = dl.default_params(nanals=1, nrow=1, nlat=2, nlon=2)
    h = ctypes.c_void_p()
    rc = dl.lib().dida_letkf_create(ctypes.byref(h), ctypes.byref(bad), 10, 0)
    assert rc < 0 and b"nanals" in dl.lib().dida_letkf_last_error()
    # k up to 128 in the plain form, up to 64 in the batched form
    bad = dl.default_params(nanals=129, nrow=1, nlat=2, nlon=2)
    rc = dl.lib().dida_letkf_create(ctypes.byref(h), ctypes.byref(bad), 10, 0)
    assert rc == -5 and b"2..128" in dl.lib().dida_letkf_last_error()
    # the chained form (64 < k <= 128) takes no vertical localisation
    bad = dl.default_params(nanals=96, nrow=1, nlat=2, nlon=2, form=dl.FORM_BATCHED,
                            gram=dl.GRAM_BF16X3, vert_dist=0.5)
    rc = dl.lib().dida_letkf_create(ctypes.byref(h), ctypes.byref(bad), 10, 0)
    assert rc == -5 and b"vertical" in dl.lib().dida_letkf_last_error()
    # the selection addresses the binned obs with 32-bit byte offsets
    ok = dl.default_params(nanals=8, nrow=3, nlat=2, nlon=2)
    rc = dl.lib().dida_letkf_create(ctypes.byref(h), ctypes.byref(ok), 1 << 29, 0)
    assert rc == -5 and b"at most 536870911 obs" in dl.lib().dida_letkf_last_error()


def test_struct_layouts_match_header(dl, tmp_path):
    """The ctypes mirrors have the size and field offsets of the C structs."""
    import shutil
    import subprocess
    import pytest
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    pairs = {"dida_letkf_params": dl.Params, "dida_letkf_obs": dl.ObsC,
             "dida_letkf_stats": dl.StatsC, "dida_letkf_tuning": dl.TuningC,
             "dida_vert_coord": dl.VertCoordC, "dida_obs_records": dl.ObsRecordsC}
    lines = ['#include <stdio.h>', '#include "dida_letkf.h"', "int main(void) {"]
    for cname, py in pairs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        cname, field, val = line.split()
        py = pairs[cname]
        want = ctypes.sizeof(py) if field == "size" else getattr(py, field).offset
        assert int(val) == want, (cname, field, int(val), want)


def test_vertical_params(dl):
    p = dl.default_params()
    assert p.vert_dist == 0.0 and p.cube_nz == 1 and p.gc_filter_vertical == 1
    h = ctypes.c_void_p()
    # vertical localisation: batched form only, cube_nz >= 1
    bad = dl.default_params(nanals=8, nrow=3, nlat=2, nlon=2, vert_dist=1.0)
    rc = dl.lib().dida_letkf_create(ctypes.byref(h), ctypes.byref(bad), 10, 0)
    assert rc == -5 and b"batched" in dl.lib().dida_letkf_last_error()
    bad = dl.default_params(nanals=8, nrow=3, nlat=2, nlon=2, vert_dist=1.0, cube_nz=0,
                            form=dl.FORM_BATCHED)
    rc = dl.lib().dida_letkf_create(ctypes.byref(h), ctypes.byref(bad), 10, 0)
    assert rc < 0 and b"cube_nz" in dl.lib().dida_letkf_last_error()


def test_cube_params(dl):
    p = dl.default_params()
    assert p.cube_n == 1 and p.overlap == 0
    h = ctypes.c_void_p()
    for cube_n, overlap in ((0, 0), (2, 2), (3, -1)):
        bad = dl.default_params(nanals=8, nrow=3, nlat=4, nlon=4, cube_n=cube_n, overlap=overlap)
        rc = dl.lib().dida_letkf_create(ctypes.byref(h), ctypes.byref(bad), 10, 0)
        assert rc < 0 and b"cube_n" in dl.lib().dida_letkf_last_error()
    bad = dl.default_params(nanals=8, nrow=3, nlat=4, nlon=4, cube_n=3, overlap=1,
                            vert_dist=1.0, form=dl.FORM_BATCHED)
    rc = dl.lib().dida_letkf_create(ctypes.byref(h), ctypes.byref(bad), 10, 0)
    assert rc == -5 and b"overlap = 0" in dl.lib().dida_letkf_last_error()
