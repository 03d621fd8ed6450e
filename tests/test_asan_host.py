"""The C-ABI layer's host code under AddressSanitizer (CPU only): every entry
point's argument checks, with NULL pointers and zero extents, reject the call
without a memory error; host-side tile arithmetic on valid extents; the
error-message path.  The library is built from the same sources with the
host part sanitized (hfm._build.build_asan_host)."""
import glob
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _asan_runtime():
    libs = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return libs[-1] if libs else None


def test_abi_argument_checks_under_address_sanitizer(tmp_path):
    from hfm import _build
    rt = _asan_runtime()
    if rt is None or not os.path.exists(_build.HIPCC):
        pytest.skip("hipcc or the ASan runtime is not installed")
    lib = _build.build_asan_host(str(tmp_path / "libhfm_asan.so"))
    env = dict(os.environ)
    # the sanitizer runtime goes first; whatever the environment preloads stays
    env["LD_PRELOAD"] = " ".join(x for x in (rt, os.environ.get("LD_PRELOAD", "")) if x)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1:verify_asan_link_order=0"
    env["HFM_NO_TORCH"] = "1"
    res = subprocess.run([sys.executable, os.path.join(HERE, "asan_abi_sweep.py"), lib], env=env,
                         capture_output=True, text=True, timeout=600)
    out = res.stdout + res.stderr
    assert res.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out
    assert "ok:" in res.stdout
