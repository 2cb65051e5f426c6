"""The C ABI library loads and exports every entry point include/racer.h
declares; no GPU compute is called here."""
import ctypes as C
import subprocess

import pytest

import warporacer3d_amd as R


def test_library_exports_every_declared_symbol():
    lib = R.lib()
    declared = R.exported_symbols()
    assert len(declared) >= 60
    missing = [name for name in declared if not hasattr(lib, name)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(R.LIB_PATH)], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(declared) <= exported


def test_version_and_defaults():
    assert b"gfx950" in R.lib().racer_version()
    cfg = R.Config()
    R.lib().racer_config_default(C.byref(cfg))
    assert (cfg.beams, cfg.hidden, cfg.rollout, cfg.epochs, cfg.minibatches, cfg.precision) == (108, 64, 32, 4, 4, 0)
    assert cfg.params() == 64 * 113 + 64 * 65 + 3 * 65 + 2


def test_cpu_device_is_refused():
    with pytest.raises(R.RacerError, match="GPU only"):
        R.Device(False)


def test_no_gpu_here_fails_loudly():
    import os

    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU is present")
    with pytest.raises(R.RacerError):
        R.Device(True)
