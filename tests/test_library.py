"""CPU checks of the C-ABI library: it loads without a GPU and exports every
entry point include/lam.h declares; the ABI structs match the header."""
import ctypes
import re

import lam


def test_library_loads_and_exports_all_symbols():
    L = lam.load_library()
    names = lam.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.lam_abi_version() == 1


def test_struct_layouts_match_header():
    text = open(lam.HEADER).read()
    shape_fields = re.search(r"typedef struct lam_shape \{(.*?)\} lam_shape;", text, re.S).group(1)
    names = re.findall(r"int32_t\s+(\w+);", shape_fields)
    assert names == [f[0] for f in lam.LamShape._fields_]
    io_fields = re.search(r"typedef struct lam_block_io \{(.*?)\} lam_block_io;", text, re.S).group(1)
    io_names = re.findall(r"(\w+);", re.sub(r"/\*.*?\*/", "", io_fields, flags=re.S))
    assert io_names == [f[0] for f in lam.LamBlockIO._fields_]
    assert ctypes.sizeof(lam.LamShape) == 40


def test_create_without_gpu_fails_cleanly():
    """No GPU here: lam_create must return an error code, not crash or exit."""
    import lam_synth as syn
    L = lam.load_library()
    sh = lam.LamShape(1, 1, 8, 4, 2, 4, 1, 1, 32, 1)
    h = ctypes.c_void_p()
    rc = L.lam_create(ctypes.byref(sh), 0, 1, ctypes.byref(h))
    assert rc in (lam.LAM_ERR_NO_DEVICE, lam.LAM_ERR_HIP, lam.LAM_OK)
    if rc == lam.LAM_OK:
        L.lam_destroy(h)
    bad = lam.LamShape(1, 1, 8, 4, 3, 4, 1, 1, 32, 1)  # 3 pols
    assert L.lam_create(ctypes.byref(bad), 0, 1, ctypes.byref(h)) == lam.LAM_ERR_UNSUPPORTED
    bad = lam.LamShape(1, 2, 8, 4, 2, 4, 1, 1, 32, 1)  # R != F*RPP
    assert L.lam_create(ctypes.byref(bad), 0, 1, ctypes.byref(h)) == lam.LAM_ERR_ARG
