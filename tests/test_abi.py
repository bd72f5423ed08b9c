"""The C ABI: libligesis_hip.so loads and exports every function include/*.h declares, and the
host-side entry points fail cleanly (no crash) without a GPU or with bad arguments. No GPU
compute is called here."""
import ctypes
import glob
import os
import re

import pytest

import ligesis_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\**\s+(lgs_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_reference_boundary():
    names = _declared()
    for must in ("lgs_init", "lgs_setup", "lgs_commit", "lgs_commit_batch", "lgs_commitment_release",
                 "lgs_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(ligesis_amd.LIB_PATH)
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_null_arguments_are_rejected():
    lib = ligesis_amd.lib()
    assert lib.lgs_init(0, None) == ligesis_amd.LGS_ERR_ARG
    assert lib.lgs_set_option(None, b"hash", 1) == ligesis_amd.LGS_ERR_ARG
    assert lib.lgs_setup(None, 3, 10, None, 0, None) == ligesis_amd.LGS_ERR_ARG
    assert lib.lgs_commit(None, None, None, 0, 0, None, None) == ligesis_amd.LGS_ERR_ARG
    lib.lgs_shutdown(None)
    lib.lgs_key_release(None)
    lib.lgs_commitment_release(None)


def test_init_without_device_fails_cleanly():
    if os.environ.get("HIP_VISIBLE_DEVICES", None) is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be visible here")
    h = ctypes.c_void_p()
    assert ligesis_amd.lib().lgs_init(0, ctypes.byref(h)) == ligesis_amd.LGS_ERR_DEVICE
