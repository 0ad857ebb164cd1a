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
    bad = dl.default_params(nanals=96, nrow=1, nlat=2, nlon=2, form=dl.FORM_BATCHED,
                            gram=dl.GRAM_BF16X3)
    rc = dl.lib().dida_letkf_create(ctypes.byref(h), ctypes.byref(bad), 10, 0)
    assert rc == -5 and b"batched" in dl.lib().dida_letkf_last_error()
