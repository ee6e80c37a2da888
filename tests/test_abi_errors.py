"""The C ABI's failure contract, on the CPU: entry points given bad
arguments return a negative status and leave the reason in
pgld_last_error(), without touching the GPU.  Every case below fails an
argument check that precedes the entry point's first HIP call, so the suite
runs with or without a device.  The same cases run again in a subprocess
against the AddressSanitizer build of the host code when it is built
(python -m pg_gpu_amd.build --asan)."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

from pg_gpu_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_vp, _i32, _i64, _f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double


def _arr(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data_as(_vp)


def cases(lib):
    """(name, status, message substring, thunk) for every failure case, and
    the arrays the thunks point into."""
    pos, p_pos = _arr([0.0, 10.0, 5.0, 20.0], np.float64)            # not sorted
    ok_pos, p_ok = _arr([0.0, 10.0, 20.0, 30.0], np.float64)
    edges_bad, p_eb = _arr([0.0, 100.0, 100.0], np.float64)           # not increasing
    edges, p_e = _arr([0.0, 100.0, 1000.0], np.float64)
    data, p_d = _arr(np.zeros((4, 8)), np.int8)
    fake = _vp(64)  # a non-null pointer that is never dereferenced: its check fails first
    out = _vp()
    kpad1 = (C.c_int64 * 1)(256)
    ns1 = (C.c_int32 * 1)(100)
    pp1 = (C.c_void_p * 1)(64)
    nv, p_nv = _arr([3, 0], np.int64)
    nv4, p_nv4 = _arr([4], np.int64)
    return [
        ("plan null output", -1, "null output",
         lambda: lib.pgld_band_plan_create(p_ok, 4, 50.0, L.KEY_BIN, None, p_e, 2, 0.0, 1.0, None)),
        ("plan unsorted positions", -1, "not sorted",
         lambda: lib.pgld_band_plan_create(p_pos, 4, 50.0, L.KEY_BIN, None, p_e, 2, 0.0, 1.0, C.byref(out))),
        ("plan negative max_dist", -1, "negative max_dist",
         lambda: lib.pgld_band_plan_create(p_ok, 4, -1.0, L.KEY_BIN, None, p_e, 2, 0.0, 1.0, C.byref(out))),
        ("plan unknown key mode", -1, "unknown key mode",
         lambda: lib.pgld_band_plan_create(p_ok, 4, 50.0, 7, None, p_e, 2, 0.0, 1.0, C.byref(out))),
        ("plan edges not increasing", -1, "edges must increase",
         lambda: lib.pgld_band_plan_create(p_ok, 4, 500.0, L.KEY_BIN, None, p_eb, 2, 0.0, 1.0, C.byref(out))),
        ("plan window without width", -1, "width > 0",
         lambda: lib.pgld_band_plan_create(p_ok, 4, 50.0, L.KEY_WINDOW, None, None, 2, 0.0, 0.0, C.byref(out))),
        ("plan null positions", -1, "null pos",
         lambda: lib.pgld_band_plan_create(None, 4, 50.0, L.KEY_BIN, None, p_e, 2, 0.0, 1.0, C.byref(out))),
        ("batch no regions", -1, "n_regions >= 1",
         lambda: lib.pgld_batch_plan_create(0, p_nv, p_ok, 50.0, L.KEY_BIN, None, p_e, 2, 0.0, 1.0, C.byref(out))),
        ("batch unsorted region", -1, "not sorted",
         lambda: lib.pgld_batch_plan_create(1, p_nv4, p_pos, 50.0, L.KEY_BIN, None, p_e, 2, 0.0, 1.0,
                                            C.byref(out))),
        ("pack bad family", -1, "bad family",
         lambda: lib.pgld_pack_planes(5, p_d, 4, 8, None, 4, 128, fake, fake, fake, None)),
        ("pack bad k_pad", -1, "multiple of 128",
         lambda: lib.pgld_pack_planes(L.FAMILY_HAP, p_d, 4, 8, None, 4, 100, fake, fake, fake, None)),
        ("pack fp4 k_pad", -1, "multiple of 256",
         lambda: lib.pgld_pack_planes_fp4(L.FAMILY_HAP, p_d, 4, 8, None, 4, 128, fake, fake, fake, None)),
        ("pack over 32767 samples", -1, "at most 32767",
         lambda: lib.pgld_pack_planes(L.FAMILY_HAP, p_d, 40000, 8, None, 40000, 40064, fake, fake, fake, None)),
        ("pack null pointer", -1, "null pointer",
         lambda: lib.pgld_pack_planes(L.FAMILY_GENO, None, 4, 8, None, 4, 128, fake, fake, fake, None)),
        ("band sums null plan", -1, "null pointer",
         lambda: lib.pgld_band_sums(None, 0, 1, C.cast(pp1, L._PP), C.cast(pp1, L._PP), kpad1, ns1, 1, 0, fake,
                                    None, None)),
        ("band sums fp4 null plan", -1, "null pointer",
         lambda: lib.pgld_band_sums_fp4(None, 0, 1, C.cast(pp1, L._PP), C.cast(pp1, L._PP), kpad1, ns1, 1, 0, fake,
                                        None, None)),
        ("noalloc null plan", -1, "null pointer",
         lambda: lib.pgld_band_sums_fp4_noalloc(None, 0, 1, C.cast(pp1, L._PP), C.cast(pp1, L._PP), kpad1, ns1, 1,
                                                0, fake, None, None)),
        ("counts null data", -1, "null pointer",
         lambda: lib.pgld_counts_pairs_hap(None, 4, 8, None, 4, fake, fake, 5, fake, None, None)),
        ("counts n_var beyond int32", -1, "fit int32",
         lambda: lib.pgld_counts_pairs_geno(fake, 4, 1 << 31, None, 4, fake, fake, 5, fake, None, None)),
        ("stats bad family", -1, "bad family",
         lambda: lib.pgld_stats_pairs(9, 1, C.cast(pp1, L._PP), None, 5, fake, fake, 3, fake, None)),
        ("stats five populations", -1, "1..4",
         lambda: lib.pgld_stats_pairs(0, 5, C.cast(pp1, L._PP), None, 5, fake, fake, 3, fake, None)),
        ("r family bad which", -1, "which must be",
         lambda: lib.pgld_r_family(fake, None, 5, 3, fake, None)),
        ("pair keys bad mode", -1, "pgld_pair_keys",
         lambda: lib.pgld_pair_keys(fake, fake, fake, 5, 9, fake, 4, 0.0, 1.0, fake, None)),
        ("keyed sums too wide", -1, "above 8192",
         lambda: lib.pgld_keyed_sums(fake, fake, 5, 100, 100, fake, None)),
        ("batch sums null plan", -1, "pgld_batch_sums",
         lambda: lib.pgld_batch_sums(None, 1, fake, None, None)),
        ("graph launch null", -1, "null graph",
         lambda: lib.pgld_graph_launch(None, None)),
    ], (pos, ok_pos, edges_bad, edges, data, nv, nv4)


def run_cases(lib):
    """Runs every case; returns a list of failure descriptions (empty = all
    statuses and messages as documented)."""
    bad = []
    todo, _keep = cases(lib)  # _keep holds the arrays the calls point into
    for name, status, msg, thunk in todo:
        rc = thunk()
        err = lib.pgld_last_error().decode(errors="replace")
        if rc != status or msg not in err:
            bad.append(f"{name}: status {rc}, message {err!r}")
    return bad


def test_abi_failure_contract():
    bad = run_cases(L.lib())
    assert not bad, "\n".join(bad)


def test_last_error_is_per_thread():
    """pgld_last_error() is thread-local: a failure on another thread leaves
    this thread's message alone."""
    import threading
    lib = L.lib()
    lib.pgld_r_family(_vp(64), None, 5, 3, _vp(64), None)
    mine = lib.pgld_last_error().decode()
    t = threading.Thread(target=lambda: lib.pgld_keyed_sums(_vp(64), _vp(64), 5, 100, 100, _vp(64), None))
    t.start()
    t.join()
    assert lib.pgld_last_error().decode() == mine
    assert lib.pgld_version() > 0


def _asan_runtime():
    import glob
    hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


def test_abi_failure_contract_under_asan():
    """The same cases and a host-side band-schedule build (plan creation up to
    its first device allocation) in a process running the AddressSanitizer
    build of the host code."""
    from pg_gpu_amd import build as B
    so = B.ASAN_OUT
    rt = _asan_runtime()
    if rt is None:
        pytest.skip("no AddressSanitizer runtime in this toolchain")
    if not os.path.exists(so):
        B.build_asan()
    code = (
        "import sys, numpy as np, ctypes as C\n"
        f"sys.path.insert(0, {ROOT!r}); sys.path.insert(0, {os.path.join(ROOT, 'tests')!r})\n"
        "from pg_gpu_amd import _lib as L\n"
        f"L.LIB_PATH = {so!r}\n"
        "import test_abi_errors as T\n"
        "bad = T.run_cases(L.lib())\n"
        "assert not bad, bad\n"
        "rng = np.random.default_rng(3)\n"
        "pos = np.sort(rng.uniform(0, 1e6, 3000))\n"
        "out = C.c_void_p()\n"
        "edges = np.logspace(0, 6, 17)\n"
        "rc = L.lib().pgld_band_plan_create(pos.ctypes.data_as(C.c_void_p), len(pos), 1e6, L.KEY_BIN, None,\n"
        "                                   edges.ctypes.data_as(C.c_void_p), 16, 0.0, 1.0, C.byref(out))\n"
        "print('plan_create status', rc)\n"
        "if rc == 0: L.lib().pgld_band_plan_destroy(out)\n"
        "print('asan run ok')\n")
    env = dict(os.environ, LD_PRELOAD=rt, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=23")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "asan run ok" in r.stdout, (r.stdout[-2000:], r.stderr[-4000:])
    assert "ERROR: AddressSanitizer" not in r.stderr
