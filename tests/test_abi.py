"""The C-ABI library loads and exports every function include/xpcs_mt.h declares
(no compute calls: no device is needed)."""
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared():
    text = (ROOT / "include" / "xpcs_mt.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xpcs_[a-z0-9_]+)\s*\(", text)))


def test_exports(xmt):
    names = declared()
    assert len(names) >= 20
    lib = xmt.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(xmt.EXPORTS)


def test_no_device_is_an_error_not_a_crash(xmt):
    import numpy as np
    try:
        import torch
        if torch.cuda.is_available():
            return
    except Exception:
        pass
    dq = np.ones((4, 4), np.int32)
    try:
        xmt.Plan(4, 4, 10, 2, dq, dq)
    except xmt.XpcsError as e:
        assert e.status == 8
