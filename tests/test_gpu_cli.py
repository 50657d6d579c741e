"""kln-tdvp command-line driver (src/kln-tdvp.cpp contract) on the GPU: reference-format
files in, State/T-<k>.mps out, C(t) against the oracle."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from oracle import model as M
from oracle import tdvp as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import __graft_entry__ as GE  # noqa: E402

pytestmark = pytest.mark.gpu


def write_train(path, tensors, magic, rank):
    with open(path, "wb") as f:
        f.write(struct.pack("ii", magic, len(tensors)))
        for t in tensors:
            items = t.sorted_items()
            f.write(struct.pack("i", len(items)))
            for q, b in items:
                f.write(struct.pack(f"{rank}i", *b.shape))
                f.write(struct.pack(f"{rank}i", *q))
                f.write(np.ascontiguousarray(b).tobytes())
        f.write(b"\n")


@pytest.mark.parametrize("layout", ["merged", "pp"])
def test_kln_tdvp_cli_matches_oracle(tmp_path, layout):
    print(f"case kln-tdvp {layout}", flush=True)
    kv = GE.load_package()
    g = np.array([1.0, 0.8, 1.2])
    lat = M.holstein(3, [1.0], g, [3], 1.0, layout)
    mpo = M.build_mpo(lat)
    psi0 = M.product_state(lat)
    write_train(tmp_path / "Ele-1.mps", psi0, 111, 3)
    write_train(tmp_path / "H_tot", mpo, 112, 4)
    exe = os.path.join(ROOT, "kylin-v_amd", "kln-tdvp")
    r = subprocess.run([exe, "-w", "Ele-1.mps", "-o", "H_tot", "-s", "[0,3,0,1]", "-v", "0.05",
                        "-D", "12", "-t", "1e-10"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-800:], r.stderr[-800:], flush=True)
    assert r.returncode == 0
    assert "Max bond dimension" in r.stdout
    ref = T.TDVP(psi0, mpo, cutoff=1e-10, maxdim=12, dt=0.05)
    for k in range(4):
        if k < 3:
            ref.two_site()
        else:
            ref.one_site()
        sites = kv.load_mps(str(tmp_path / "State" / f"T-{k}.mps"))
        o_sites = []
        from oracle.blocksparse import Tensor as OT
        for s in sites:
            dirs, blocks = s.to_blocks()
            o_sites.append(OT(dirs, blocks))
        c_cli = T.overlap_states(psi0, o_sites)
        c_ref = T.overlap_states(psi0, ref.mps)
        assert abs(c_cli - c_ref) < 1e-9
    # an SA-1TDVP stage without two-site sweeps before it has no widths: a message, no crash
    r2 = subprocess.run([exe, "-w", "Ele-1.mps", "-o", "H_tot", "-s", "[0,0,1,0]", "-v", "0.05"],
                        cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r2.returncode == 1 and "needs a width" in r2.stderr


def test_kln_tdvp_cli_sa_stage(tmp_path):
    """-s [0,6,2,1]: two-site sweeps write ./SVS, the SA stage fits ./beta and runs two
    enrichment sweeps (seed 1 + sweep), then one plain one-site sweep; every written state
    against the oracle run with the CLI's own beta file."""
    print("case kln-tdvp sa", flush=True)
    kv = GE.load_package()
    from oracle import sa
    from oracle.blocksparse import Tensor as OT
    N = 4
    lat = M.holstein(N, [1.0], np.array([0.9, 1.1, 0.8, 1.0]), [3], 1.0, "merged")
    mpo = M.build_mpo(lat)
    psi0 = M.product_state(lat)
    write_train(tmp_path / "Ele-1.mps", psi0, 111, 3)
    write_train(tmp_path / "H_tot", mpo, 112, 4)
    exe = os.path.join(ROOT, "kylin-v_amd", "kln-tdvp")
    r = subprocess.run([exe, "-w", "Ele-1.mps", "-o", "H_tot", "-s", "[0,6,2,1]", "-v", "0.05",
                        "-D", "30", "-t", "1e-10"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-600:], r.stderr[-600:], flush=True)
    assert r.returncode == 0 and "Fitting OK!" in r.stdout
    svs = (tmp_path / "SVS").read_text()
    beta_txt = (tmp_path / "beta").read_text()
    assert len(beta_txt.splitlines()) == N - 1
    assert all(len(l) == 21 for l in beta_txt.splitlines())
    w_cli = [float(l) for l in beta_txt.splitlines()]
    w_ora = sa.fit_widths(svs)
    assert np.allclose(w_cli, w_ora, rtol=2e-3, atol=0)
    betas = sa.betas_from_widths(beta_txt)
    eng = T.TDVP(psi0, mpo, cutoff=1e-10, maxdim=30, dt=0.05)
    for k in range(9):
        if k < 6:
            eng.two_site()
        elif k < 8:
            eng.one_site(betas=betas, seed=1 + k)
        else:
            eng.one_site()
        got = [OT(*s.to_blocks()) for s in kv.load_mps(str(tmp_path / "State" / f"T-{k}.mps"))]
        assert [sum(t.legs()[2].values()) for t in got[:-1]] == eng.bond_dims(), k
        assert abs(T.overlap_states(psi0, got) - T.overlap_states(psi0, eng.mps)) < 1e-9, k
    # the oracle's SVS records equal the CLI's
    a, b = sa.parse_svs(svs), sa.parse_svs(sa.svs_text(N, eng.svs))
    assert [len(x) for x in a] == [len(x) for x in b]


def test_kln_tdvp_cli_global_krylov_stage(tmp_path):
    """-s [2,3,0,0] in the projected-purification layout: two global Krylov steps open the
    exciton-hop sectors (holstein/tdvp.h:318-327), then 2TDVP moves the exciton; C(t) and
    the populations of every written state against the oracle."""
    print("case kln-tdvp nK pp", flush=True)
    kv = GE.load_package()
    from oracle import krylov as K
    from oracle.blocksparse import Tensor as OT
    lat = M.holstein(3, [1.0], np.array([1.0, 0.8, 1.2]), [3], 1.0, "pp")
    mpo = M.build_mpo(lat)
    psi0 = M.product_state(lat)
    write_train(tmp_path / "Ele-1.mps", psi0, 111, 3)
    write_train(tmp_path / "H_tot", mpo, 112, 4)
    exe = os.path.join(ROOT, "kylin-v_amd", "kln-tdvp")
    r = subprocess.run([exe, "-w", "Ele-1.mps", "-o", "H_tot", "-s", "[2,3,0,0]", "-v", "0.05",
                        "-D", "24", "-t", "1e-10"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-600:], r.stderr[-600:], flush=True)
    assert r.returncode == 0
    Hc = K.scale_mpo(mpo, 0.05)
    g1, _ = K.global_main(Hc, psi0, 1e-10, 24)
    g2, _ = K.global_main(Hc, g1, 1e-10, 24)
    expected = [g1, g2]
    eng = T.TDVP(g2, mpo, cutoff=1e-10, maxdim=24, dt=0.05)
    for _ in range(3):
        eng.two_site()
        expected.append([t.copy() for t in eng.mps])
    for k, want in enumerate(expected):
        sites = kv.load_mps(str(tmp_path / "State" / f"T-{k}.mps"))
        got = [OT(*s.to_blocks()) for s in sites]
        assert abs(T.overlap_states(psi0, got) - T.overlap_states(psi0, want)) < 1e-9
    n1 = M.density_mpo(lat, 1)
    assert T.join(n1, got, got).real > 1e-5  # the exciton left its first site
