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


def test_stream_bytes_and_unsupported_streams(xmt):
    """Admission sizes are host-only arithmetic; streams refuse what they do not do."""
    import numpy as np
    import pytest
    import synth
    dq, sq = synth.annular_qmaps(64, 64, 6, 3)
    P = int(np.count_nonzero((dq > 0) & (sq > 0)))
    b4 = xmt.stream_bytes(64, 64, 5000, 4, dq, sq)
    b8 = xmt.stream_bytes(64, 64, 5000, 8, dq, sq)
    b8_long = xmt.stream_bytes(64, 64, 50000, 8, dq, sq)
    # level state alone is P * levels * 2dpl * 16 bytes
    assert b8 > b4 > P * 10 * 8 * 16
    assert b8_long > b8
    with pytest.raises(xmt.XpcsError, match="UNSUPPORTED"):
        xmt.stream_bytes(64, 64, 5000, 2, dq, sq)
