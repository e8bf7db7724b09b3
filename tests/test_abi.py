"""The C-ABI library loads and exports every symbol include/*.h declares.
No compute calls: these run without a GPU."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names += re.findall(r"GPUD_API\s+[\w\s\*]+?\b(gpud_\w+)\s*\(", txt)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = _declared()
    assert "gpud_launch_streaming_polyphase_stateful_hip" in names
    assert "gpud_runner_streaming_extract" in names
    assert len(names) > 30


def test_library_exports_every_declared_symbol():
    import argus
    lib = argus.lib()  # raises ErrUnavailable if the .so is missing
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_declares_every_export():
    import argus
    declared = set(_declared())
    bound = set(argus.EXPORTS)
    assert declared <= bound | _late_bound(), sorted(declared - bound)


def _late_bound():
    # modules that register their own entry points on import
    import importlib
    names = set()
    for mod in ("argus.surveillance", "argus.serve", "argus.wfm", "argus.ingest", "argus.jobs"):
        try:
            importlib.import_module(mod)
        except ModuleNotFoundError:
            continue
    import argus
    names |= set(argus._lib.EXPORTS)
    return names


def test_version_string():
    import argus
    assert b"gfx950" in argus.lib().gpud_version()


def test_pure_host_helpers_match_oracle():
    import numpy as np
    import oracle as O
    from argus import gpudemod as G
    assert G.StreamingConfigHash(7, 1.0, 2.0, 200000, 65, 4000000) == O.StreamingConfigHash(7, 0, 0, 200000, 65,
                                                                                            4000000)
    assert G.ExactIntegerDecimation(20_000_000, 250_000) == 80
    with pytest.raises(ValueError):
        G.ExactIntegerDecimation(20_000_000, 256_000)
    np.testing.assert_array_equal(G.LowpassFIR(7500, 250000, 101), O.LowpassFIR(7500, 250000, 101))
    base = np.arange(1, 8, dtype=np.float32)
    np.testing.assert_array_equal(G.BuildPolyphaseTapsPhaseMajor(base, 3), [1, 4, 7, 2, 5, 0, 3, 6, 0])


def test_library_resolves_its_own_symbols():
    """Every argus_/gpud_ symbol the library uses is defined in it (a C-linkage
    declaration of a C++ helper would leave one undefined until load)."""
    import subprocess
    lib = ROOT / "argus-sdr_amd" / "libargus_hip.so"
    out = subprocess.run(["nm", "-D", "--undefined-only", str(lib)], capture_output=True, text=True, check=True).stdout
    own = [l for l in out.splitlines() if " argus_" in l or " gpud_" in l]
    assert not own, own


def test_entry_points_refuse_missing_handles():
    """Entry points check their handle before any device work: a null
    service or engine is refused with -1 (no GPU needed)."""
    import ctypes as C
    import argus
    _late_bound()
    lib = argus.lib()
    assert lib.gpud_serve_set_download_split(None, 4, 0) == -1
    assert lib.gpud_serve_set_download_parts(None, 1) == -1
    assert lib.gpud_serve_set_demod_fusion(None, 1) == -1
    from argus import surveillance as SV
    cfar = SV.cfar_params("CA", guard=4, train=16)
    assert lib.gpud_surv_cfar_runs_device_ex(None, None, 1, C.byref(cfar), None, None, 16, None) == -1
