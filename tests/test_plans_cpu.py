"""CPU checks of the engine's host-side planners against the oracle.

tests/native/plan_host_exec.cpp runs the exact work lists the HIP kernels consume
(grouped-GEMM segments, MPO-fold terms, permutations; holstein/plan.h), plus the Krylov
exponential coefficients (herm_eig.h) and the truncation rule (svd_rule.h), with naive
host loops.  Agreement with the oracle here pins the planning logic; the GPU tests then
pin the kernels.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from oracle import model as M
from oracle import tdvp as T
from oracle.blocksparse import Tensor, product

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("native") / "plan_host_exec")
    subprocess.run([gxx, "-std=c++17", "-O2", "-o", out,
                    os.path.join(ROOT, "tests", "native", "plan_host_exec.cpp")], check=True)
    return out


def write_tensor(f, t: Tensor):
    f.write(struct.pack("i", t.rank))
    f.write(struct.pack(f"{t.rank}i", *t.dirs))
    items = t.sorted_items()
    f.write(struct.pack("i", len(items)))
    for q, b in items:
        f.write(struct.pack(f"{t.rank}i", *q))
        f.write(struct.pack(f"{t.rank}i", *b.shape))
        f.write(np.ascontiguousarray(b, dtype=np.complex128).tobytes())


def read_tensor(buf, pos=0):
    rank = struct.unpack_from("i", buf, pos)[0]
    pos += 4
    dirs = struct.unpack_from(f"{rank}i", buf, pos)
    pos += 4 * rank
    nb = struct.unpack_from("i", buf, pos)[0]
    pos += 4
    t = Tensor(dirs)
    for _ in range(nb):
        q = struct.unpack_from(f"{rank}i", buf, pos)
        pos += 4 * rank
        sh = struct.unpack_from(f"{rank}i", buf, pos)
        pos += 4 * rank
        n = int(np.prod(sh))
        a = np.frombuffer(buf, dtype=np.complex128, count=n, offset=pos).reshape(sh)
        pos += 16 * n
        t.blocks[tuple(q)] = a.copy()
    return t, pos


def run(harness, tmp_path, mode, tensors=None, raw=None):
    fin = tmp_path / "in.bin"
    fout = tmp_path / "out.bin"
    with open(fin, "wb") as f:
        for t in tensors or []:
            write_tensor(f, t)
        if raw is not None:
            f.write(raw)
    subprocess.run([harness, mode, str(fin), str(fout)], check=True)
    return open(fout, "rb").read()


def assert_same(native: Tensor, ref: Tensor, tol=1e-12):
    """Blocks of ref equal; blocks only in native (closure padding) must be zero."""
    assert native.dirs == ref.dirs
    for q, b in ref.blocks.items():
        assert q in native.blocks, f"missing block {q}"
        np.testing.assert_allclose(native.blocks[q], b, atol=tol, rtol=0)
    for q, b in native.blocks.items():
        if q not in ref.blocks:
            assert np.abs(b).max() <= tol


def evolved_state(layout, N=3, L=3, steps=2, D=12, seed=0):
    rng = np.random.default_rng(seed)
    g = rng.uniform(0.6, 1.4, size=(N, 1))
    lat = M.holstein(N, [1.0], g, [L], 0.8, layout)
    mpo = M.build_mpo(lat)
    eng = T.TDVP(M.product_state(lat), mpo, cutoff=1e-12, maxdim=D, dt=0.3)
    for _ in range(steps):
        eng.two_site()
    return lat, mpo, eng


@pytest.mark.parametrize("layout", ["merged", "pp", "split"])
def test_heff_plans_match_oracle(harness, tmp_path, layout):
    lat, mpo, eng = evolved_state(layout)
    envl, envr = eng.envL[-1], eng.envR[-1]  # environments of site 0
    # one-site
    y = T.apply_single(envl, eng.mps[0], mpo[0], envr)
    got, _ = read_tensor(run(harness, tmp_path, "heff1", [envl, mpo[0], envr, eng.mps[0]]))
    assert_same(got, y)
    # two-site on (0, 1): right env of site 1
    envr2 = eng.envR[-2] if len(eng.envR) >= 2 else None
    x2 = product(eng.mps[0], eng.mps[1], 1)
    y2 = T.apply_double(envl, x2, mpo[0], mpo[1], envr2)
    got2, _ = read_tensor(run(harness, tmp_path, "heff2", [envl, mpo[0], mpo[1], envr2, x2]))
    assert_same(got2, y2)
    # zero-site on bond (0|1)
    cur, nex = T.svd_impl(eng.mps[0], 2, -1.0, 10 ** 6, True)
    envl1 = T.env_sweep(envl, cur, mpo[0], True)
    y0 = T.apply_zero(envl1, nex, envr)
    got0, _ = read_tensor(run(harness, tmp_path, "heff0", [envl1, envr, nex]))
    assert_same(got0, y0)


@pytest.mark.parametrize("layout", ["merged", "pp"])
def test_env_plans_match_oracle(harness, tmp_path, layout):
    lat, mpo, eng = evolved_state(layout)
    envl = eng.envL[-1]
    ref = T.env_sweep(envl, eng.mps[0], mpo[0], True)
    got, _ = read_tensor(run(harness, tmp_path, "env_l2r", [envl, mpo[0], eng.mps[0]]))
    assert_same(got, ref)
    envr = T.boundary_envs(eng.mps)[1]
    ref_r = T.env_sweep(envr, eng.mps[-1], mpo[-1], False)
    got_r, _ = read_tensor(run(harness, tmp_path, "env_r2l", [envr, mpo[-1], eng.mps[-1]]))
    assert_same(got_r, ref_r)


@pytest.mark.parametrize("spec", ["1NN", "2CN", "2NT", "1TN", "2NC"])
def test_generic_product_plans(harness, tmp_path, spec):
    lat, mpo, eng = evolved_state("merged", N=4, L=2)
    a, b = eng.mps[1], eng.mps[2]
    nc, o1, o2 = int(spec[0]), spec[1], spec[2]
    if spec == "1NN":
        A, B = a, b
    elif spec == "2CN":
        A, B = a, a
    elif spec == "2NT":
        A, B = a, a.copy()
        B.dirs = tuple(-d for d in a.dirs)  # (a,s,b) against (a',s,b) with flipped legs
    elif spec == "1TN":
        A = a
        B = Tensor((-1, 1))
        for q, blk in a.blocks.items():
            B.push_back_ch((q[0], q[0]), np.eye(blk.shape[0]) * (1 + 0.5j))
    else:
        A, B = a, a
    ref = product(A, B, nc, o1, o2)
    got, _ = read_tensor(run(harness, tmp_path, f"product:{nc}{o1}{o2}", [A, B]))
    assert_same(got, ref)


def test_krylov_coefficients_match_eigh(harness, tmp_path):
    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 14):
        al = rng.normal(size=n) + 1e-14j * rng.normal(size=n)
        be = np.abs(rng.normal(size=n - 1)) + 1e-14j * rng.normal(size=n - 1)
        for rev in (0, 1):
            raw = struct.pack("i", n) + al.astype(np.complex128).tobytes() + be.astype(np.complex128).tobytes()
            out = np.frombuffer(run(harness, tmp_path, f"krycoef:{rev}", raw=raw), dtype=np.complex128)
            h = np.zeros((n, n), dtype=complex)
            h[np.arange(n), np.arange(n)] = al
            h[np.arange(n - 1), np.arange(1, n)] = be
            w, v = T.herm_eig_small(h)
            ref = v @ (np.exp((1j if rev else -1j) * w) * np.conj(v[0, :]))
            np.testing.assert_allclose(out, ref, atol=1e-13)


def test_truncation_rule_matches_oracle(harness, tmp_path):
    cases = [([[0.9, 0.3, 1e-11], [0.5, 0.3, 0.1, 2e-10]], 1e-10, 100, [2, 4]),
             ([[0.9, 0.3, 1e-11], [0.5, 0.3, 0.1, 2e-10]], 2e-10, 100, [2, 3]),
             ([[0.9, 0.3, 1e-11], [0.5, 0.3, 0.1, 2e-10]], 0.0, 3, [2, 1]),
             ([[0.9, 0.3, 1e-11], [0.5, 0.3, 0.1, 2e-10]], 10.0, 100, [1, 0]),
             ([[0.0, 0.0], [0.0]], -1.0, 2, [2, 0])]
    for sv, cut, md, expect in cases:
        raw = struct.pack("i", len(sv))
        for s in sv:
            raw += struct.pack("i", len(s)) + np.asarray(s, dtype=np.float64).tobytes()
        raw += struct.pack("d", cut) + struct.pack("i", md)
        out = list(np.frombuffer(run(harness, tmp_path, "trunc", raw=raw), dtype=np.int32))
        assert out == expect, (sv, cut, md, out)


@pytest.mark.parametrize("layout", ["merged", "pp"])
def test_two_site_heff_plans_match_oracle_on_every_bond(harness, tmp_path, layout):
    """The two-site H_eff, with the two MPO folds one after the other and as the one-pass
    composite, on every bond of an evolved chain, with environments built from both ends:
    interior bonds have sectors the boundary bond lacks."""
    lat, mpo, eng = evolved_state(layout, N=5, L=3, steps=3, D=16, seed=4)
    n = len(eng.mps)
    envl0, envr0 = eng.envL[0], eng.envR[0]
    left = [envl0]
    for i in range(n - 1):
        left.append(T.env_sweep(left[-1], eng.mps[i], mpo[i], True))
    right = [envr0]
    for i in range(n - 1, 0, -1):
        right.append(T.env_sweep(right[-1], eng.mps[i], mpo[i], False))
    for i in range(n - 1):
        envl, envr = left[i], right[n - 2 - i]
        x2 = product(eng.mps[i], eng.mps[i + 1], 1)
        y2 = T.apply_double(envl, x2, mpo[i], mpo[i + 1], envr)
        scale = max(np.abs(b).max() for b in y2.blocks.values())
        for mode in ("heff2", "heff2c"):
            got2, _ = read_tensor(run(harness, tmp_path, mode, [envl, mpo[i], mpo[i + 1], envr, x2]))
            assert_same(got2, y2, tol=1e-12 * max(scale, 1.0))
