"""Argument checks of every libhfm entry point, run against a library whose
host code is built with AddressSanitizer (tests/test_asan_host.py starts this
script with the sanitizer runtime preloaded).  No GPU is needed: every call
here must be rejected, or answered, by host code before any device work.

Usage: python asan_abi_sweep.py <path to the sanitized libhfm.so>
Prints one line per entry point and exits non-zero if a call that should be
rejected is accepted; a memory error aborts the process (the sanitizer's
report names it)."""
import ctypes as C
import os
import sys

os.environ["HFM_NO_TORCH"] = "1"
sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..",
                                "ai-hfm-solver_amd"))
from hfm import _lib as L  # noqa: E402

# entry points that take no arguments or only report sizes
QUERIES = {"hfm_last_error", "hfm_version"}
# entry points that answer from host state when given valid out-pointers
SIZE_QUERIES = {"hfm_struct_sizes", "hfm_mg_persist_struct_sizes"}


def zero_args(argtypes):
    out = []
    for a in argtypes:
        if a in (C.c_float, C.c_double):
            out.append(0.0)
        elif a in (C.c_int, C.c_int64, C.c_int32, C.c_uint32):
            out.append(0)
        else:
            out.append(None)
    return out


def main(path):
    lib = C.CDLL(path)
    lib.hfm_last_error.restype = C.c_char_p
    bad = []
    # every function the header declares has a binding here
    import re
    hdr = open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "include",
                            "hfm.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char\*)\s+(hfm_\w+)\s*\(", hdr, re.M))
    missing = sorted(declared - set(L.SIGNATURES))
    if missing:
        bad += [f"{m} (no binding)" for m in missing]
    for name, (res, args) in sorted(L.SIGNATURES.items()):
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
        if name in QUERIES:
            fn()
            continue
        if name in SIZE_QUERIES:
            a, b = C.c_int64(0), C.c_int64(0)
            rc = fn(C.byref(a), C.byref(b))
            ok = rc == 0 and a.value > 0 and b.value > 0
            print(f"{name}: rc={rc} sizes=({a.value}, {b.value})", flush=True)
            if not ok:
                bad.append(name)
            rc = fn(None, None)
            if rc >= 0:
                bad.append(name + "(NULL)")
            continue
        # every pointer NULL, every extent and count zero: must be rejected
        rc = fn(*zero_args(args))
        msg = lib.hfm_last_error()
        print(f"{name}: rc={rc} {msg.decode() if msg else ''}", flush=True)
        if rc >= 0 or not msg:
            bad.append(name)
    # host-side arithmetic on valid extents (no device work)
    tile = lib.hfm_mg_tile
    tile.argtypes = L.SIGNATURES["hfm_mg_tile"][1]
    for shape in [(8, 8, 8), (128, 128, 128), (33, 17, 64), (4, 1024, 8), (512, 512, 512)]:
        for mp in (0, 8, 64):
            r, z, n = C.c_int32(0), C.c_int32(0), C.c_int64(0)
            rc = tile(*shape, mp, C.byref(r), C.byref(z), C.byref(n))
            if rc != 0 or r.value <= 0 or z.value <= 0 or n.value <= 0:
                bad.append(f"hfm_mg_tile{shape}")
    # a single plane of 2^30 cells: within the cell limit (2^31 - 1), beyond the
    # plane limit (a cell is its plane's base plus a 32-bit byte offset)
    conv = lib.hfm_conv27_f32
    conv.argtypes = L.SIGNATURES["hfm_conv27_f32"][1]
    K = (C.c_float * 27)()
    rc = conv(C.c_void_p(8), C.c_void_p(16), None, K, 1.0, 1, 32768, 32768, 0, None)
    if rc >= 0:
        bad.append("hfm_conv27_f32(1 x 32768 x 32768)")
    # a message longer than any buffer the library keeps
    lib.hfm_bc_velocity_inplace_noalloc_f32.argtypes = \
        L.SIGNATURES["hfm_bc_velocity_inplace_noalloc_f32"][1]
    rc = lib.hfm_bc_velocity_inplace_noalloc_f32(C.c_void_p(8), C.c_void_p(8), C.c_void_p(8),
                                                 1024, 1024, 1024, 1.0, None, None)
    if rc >= 0:
        bad.append("hfm_bc_velocity_inplace_noalloc_f32(1024^3)")
    print("rejected-size message:", lib.hfm_last_error().decode(), flush=True)
    if bad:
        print("accepted bad arguments:", ", ".join(bad), flush=True)
        return 1
    print(f"ok: {len(L.SIGNATURES)} entry points", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1]))
