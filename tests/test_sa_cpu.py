"""SA-1TDVP host rules of the engine (holstein/sa_rule.h) against the oracle on the CPU:
the seeded draws, the inverse width model and the enrichment count through the native
harness, the SVS record, and the width fit through the C ABI (host only, no device)."""
import ctypes
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from oracle import model as M
from oracle import sa
from oracle import tdvp as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "kylin-v_amd", "libkylinv.so")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("native") / "plan_host_exec")
    subprocess.run([gxx, "-std=c++17", "-O2", "-o", out,
                    os.path.join(ROOT, "tests", "native", "plan_host_exec.cpp")], check=True)
    return out


def run(harness, tmp_path, mode, payload):
    fi, fo = tmp_path / "in.bin", tmp_path / "out.bin"
    fi.write_bytes(payload)
    subprocess.run([harness, mode, str(fi), str(fo)], check=True)
    return fo.read_bytes()


SPECTRA = [
    [1.0, 0.5, 0.25, 0.1, 1e-2, 1e-3],
    list(np.sort(10.0 ** np.random.default_rng(1).uniform(-9, 0, 40))[::-1]),
    list(np.sort(10.0 ** np.random.default_rng(2).uniform(-4, 0, 12))[::-1]),
]


@pytest.mark.parametrize("spec", range(len(SPECTRA)))
@pytest.mark.parametrize("beta,seed,call", [(0.5, 1, 0), (3.0, 12345, 7), (20.0, 2 ** 63 + 5, 41)])
def test_extension_count_and_draws_match_oracle(harness, tmp_path, spec, beta, seed, call):
    sv = SPECTRA[spec]
    maxdim, thres = 64, 1e-10
    payload = struct.pack("i", len(sv)) + np.array(sv).tobytes() + struct.pack("i", maxdim)
    payload += struct.pack("dd", thres, beta) + struct.pack("QQ", seed, call)
    out = run(harness, tmp_path, "sacount", payload)
    cnt = struct.unpack_from("i", out, 0)[0]
    d = np.frombuffer(out, dtype=np.float64, offset=4).reshape(4, 3)
    want = sa.extension_count(sv, maxdim, thres, beta,
                              lambda j: sa.inv_quasi_gauss(sa.draw_u01(seed, call, 0, j), beta))
    assert cnt == want
    for j in range(4):
        assert d[j, 0] == sa.draw_u01(seed, call, 0, j)
        assert d[j, 1] == sa.draw_u01(seed, call, 3, j)
        assert abs(d[j, 2] - sa.inv_quasi_gauss(d[j, 0], beta)) <= 1e-15


def test_svs_record_matches_oracle(harness, tmp_path):
    sv = SPECTRA[1]
    out = run(harness, tmp_path, "svsrec", struct.pack("i", len(sv)) + np.array(sv).tobytes())
    got = np.frombuffer(out, dtype=np.float64)
    want = np.array(sa.svs_record(sv))
    assert got.shape == want.shape and np.abs(got - want).max() <= 1e-14


def _lib():
    if not os.path.exists(LIB):
        pytest.skip("libkylinv.so not built")
    lib = ctypes.CDLL(LIB)
    fn = lib.kv_sa_fit_widths
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                   ctypes.POINTER(ctypes.c_int)]
    return fn


def _fit_cpp(fn, path, nmax=64):
    w = np.zeros(nmax)
    n = ctypes.c_int()
    rc = fn(str(path).encode(), w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nmax, ctypes.byref(n))
    return rc, list(w[:n.value])


def _objective(vals, b):
    ni, edges = np.histogram(vals, 50)
    y = ni / len(vals)
    return float(np.sum((sa.fit_model(edges[:-1], b) - y) ** 2))


def _check_widths(txt, got):
    want = sa.fit_widths(txt)
    lists = sa.parse_svs(txt)
    assert len(got) == len(want) == len(lists)
    for vals, g, w in zip(lists, got, want):
        if len(vals) <= 2:
            assert g == w == 1.0
            continue
        # at least as good a minimum as scipy's, and the same one
        assert _objective(vals, g) <= _objective(vals, w) * (1 + 1e-8)
        assert abs(np.log(g) - np.log(w)) < 2e-3


def test_width_fit_matches_oracle_on_synthetic_logs(tmp_path):
    fn = _lib()
    rng = np.random.default_rng(8)
    recs = []
    for k in range(10):  # N = 6: 10 records per sweep
        n = int(rng.integers(3, 120))
        recs.append((n, list(rng.laplace(0.0, 0.05 + 0.05 * k, size=n - 2))))
    recs += [(2, []), (4, [0.3, -0.1])]  # a second sweep: short records
    txt = sa.svs_text(6, recs)
    p = tmp_path / "SVS"
    p.write_text(txt)
    rc, got = _fit_cpp(fn, p)
    assert rc == 0
    _check_widths(txt, got)


def test_width_fit_matches_oracle_on_a_two_site_run(tmp_path):
    """The SVS log the oracle's two-site sweeps write (holstein/svd.h:390-400) on a small
    chain, fitted by the engine and by the oracle."""
    fn = _lib()
    N, Ls, om, J, dt = 4, [3], [1.0], 1.0, 0.05
    g = np.array([0.9, 1.1, 0.8, 1.0]).reshape(4, 1)
    lat = M.holstein(N, om, g, Ls, J, "merged")
    eng = T.TDVP(M.product_state(lat), M.build_mpo(lat), cutoff=1e-10, maxdim=30, dt=dt)
    for _ in range(8):
        eng.two_site()
    txt = sa.svs_text(N, eng.svs)
    p = tmp_path / "SVS"
    p.write_text(txt)
    rc, got = _fit_cpp(fn, p)
    assert rc == 0
    _check_widths(txt, got)


def test_width_fit_rejects_bad_logs(tmp_path):
    fn = _lib()
    p = tmp_path / "SVS"
    p.write_text("not an SVS file\n")
    assert _fit_cpp(fn, p)[0] == -5
    assert _fit_cpp(fn, tmp_path / "missing")[0] == -1
