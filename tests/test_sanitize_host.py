"""Host code under sanitizers (no GPU): the HDF5 readers (the raw multi-threaded path and the H5Dread fallbacks),
the FITS writer and the oracle's C API built with AddressSanitizer and UndefinedBehaviorSanitizer (undefined
behaviour is fatal), the readers again under ThreadSanitizer, and the command line under AddressSanitizer through
its argument checks, file reads and the device start-up thread (on a machine without a GPU the run ends at the
device setup, as it should).  The driver is tests/sanitize/driver.cpp."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDF5 = "/opt/conda"
PKG = os.path.join(ROOT, "snapshot-raytracer_amd")
DRIVER = os.path.join(ROOT, "tests", "sanitize", "driver.cpp")
IO_SRC = os.path.join(PKG, "csrc", "srt_io.cpp")
ORACLE_SRC = [os.path.join(ROOT, "oracle", "oracle_capi.cpp"), os.path.join(ROOT, "oracle", "opcount.cpp")]
ASAN_UBSAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(HDF5, "include", "hdf5.h")), reason="no HDF5")


def build(out, sources, flags, libs=()):
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-pthread", "-fno-omit-frame-pointer", *flags, f"-I{HDF5}/include",
           *sources, *libs, f"{HDF5}/lib/libhdf5.so", f"-Wl,-rpath,/usr/lib/x86_64-linux-gnu:{HDF5}/lib",
           "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    return str(out)


def run(cmd, cwd=None, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run(cmd, capture_output=True, text=True, timeout=600, cwd=cwd, env=e)


@pytest.fixture(scope="module")
def inputs(tmp_path_factory):
    from snapshot_raytracer_amd import snapshot_io as sio
    from test_io_host import write_h5
    d = tmp_path_factory.mktemp("sanitize")
    n = 160  # 16 MB potential, 49 MB shift: several read threads
    seq = np.arange(n ** 3, dtype=np.float32).reshape(n, n, n)
    sio.write_potential(str(d / "seq.h5"), seq)
    sio.write_potential(str(d / "seqd.h5"), seq.astype(np.float64))
    sio.write_shift(str(d / "seq3.h5"), np.arange(3 * n ** 3, dtype=np.float32).reshape(n, n, n, 3))
    native = sio._global("H5T_NATIVE_FLOAT_g")
    write_h5(sio, str(d / "seqc.h5"), seq, native, native, chunk=(40, 80, 160))
    write_h5(sio, str(d / "sequb.h5"), seq, native, native, userblock=512)
    return d


def test_io_and_oracle_under_asan_ubsan(inputs, tmp_path):
    exe = build(tmp_path / "driver_asan", [DRIVER, IO_SRC, *ORACLE_SRC], ASAN_UBSAN + ["-fopenmp"])
    for what in ("io", "oracle"):
        r = run([exe, str(inputs), what], env={"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0"})
        assert r.returncode == 0, (what, r.stdout[-2000:], r.stderr[-4000:])
        assert f"{what} checks: 0 failures" in r.stdout


def test_io_under_tsan(inputs, tmp_path):
    exe = build(tmp_path / "driver_tsan", [DRIVER, IO_SRC], ["-fsanitize=thread", "-DSANITIZE_IO_ONLY"])
    r = run([exe, str(inputs), "io"])
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "ThreadSanitizer" not in r.stderr
    assert "io checks: 0 failures" in r.stdout


def test_cli_under_asan_ubsan(inputs, tmp_path):
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("the command line under AddressSanitizer is checked on machines without a GPU")
    if not os.path.exists(os.path.join(PKG, "libsrt.so")):
        pytest.skip("libsrt.so not built")
    from snapshot_raytracer_amd import snapshot_io as sio
    exe = build(tmp_path / "cli_asan", [os.path.join(PKG, "csrc", "srt_cli.cpp"), IO_SRC], ASAN_UBSAN,
                libs=[f"-L{PKG}", "-lsrt", f"-Wl,-rpath,{PKG}"])
    sio.write_potential(str(tmp_path / "p.h5"), np.zeros((8, 8, 8), np.float32))
    sio.write_shift(str(tmp_path / "b.h5"), np.zeros((8, 8, 8, 3), np.float32))
    sio.write_shift(str(tmp_path / "b6.h5"), np.zeros((8, 8, 6, 3), np.float32))
    env = {"ASAN_OPTIONS": "detect_leaks=0"}  # the HIP runtime keeps allocations to the end of the process
    cases = [([], "Usage:"), (["p.h5", "b.h5", "0", "4"], "Invalid Nside value"),
             (["missing.h5", "b.h5", "4", "4"], "Failed to open HDF5 file"),
             (["p.h5", "p.h5", "4", "4"], "not an array type"), (["p.h5", "b6.h5", "4", "4"], "dimensions differ"),
             (["p.h5", "b.h5", "4", "4", "12"], "Batch exceeds total number of pixels"),
             (["p.h5", "b.h5", "4", "4"], "Device setup failed")]
    for args, msg in cases:
        r = run([exe, *args], cwd=tmp_path, env=env)
        assert r.returncode == 255, (args, r.stderr[-3000:])
        assert msg in r.stderr, (args, r.stderr[-3000:])
        assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, (args, r.stderr[-3000:])
