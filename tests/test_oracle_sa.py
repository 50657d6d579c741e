"""Pins for the oracle's SA-1TDVP pieces (CPU): the width model against its normalisation
and an independent root finder, the count rule against hand-derived spectra, the
enrichment SVD against the invariants it must keep, the SVS bond mapping against a
hand-traced file, the width fit against its own objective, and the SA sweep against
exact dynamics."""
import numpy as np
import pytest
import scipy.integrate as si
import scipy.optimize as so

from oracle import model as M
from oracle import sa
from oracle import tdvp as T
from oracle.blocksparse import Tensor, product


@pytest.mark.parametrize("b", [0.3, 1.0, 4.0, 25.0])
def test_fit_model_is_a_normalised_density(b):
    val, err = si.quad(lambda x: float(sa.fit_model(np.array([x]), b)[0]), -np.inf, np.inf)
    assert abs(val - 1.0) < 1e-8


@pytest.mark.parametrize("b", [0.3, 1.0, 4.0, 25.0])
def test_quasi_gauss_is_the_fitted_density_scaled_to_peak_one(b):
    assert abs(sa.quasi_gauss(0.0, b) - 1.0) < 1e-14
    scale = 4.0 * np.sqrt(b / np.pi) / b  # 1 / fit_model(0, b)
    for x in (0.05, 0.3, 0.9, 1.7, -0.6):
        ref = float(sa.fit_model(np.array([x]), b)[0]) * scale
        # the C++ form uses pi = 3.1415926 inside the bracket; with the cancellation far in
        # the tail the shapes then differ by up to ~2e-7 relative
        assert abs(sa.quasi_gauss(x, b) - ref) <= 2e-6 * ref


@pytest.mark.parametrize("b", [0.5, 2.0, 8.0])
def test_inv_quasi_gauss_matches_an_independent_root_finder(b):
    q2 = sa.quasi_gauss(2.0, b)
    for y in np.linspace(q2 + 1e-3, 0.999, 9):
        root = so.brentq(lambda x: sa.quasi_gauss(x, b) - y, 0.0, 2.0, xtol=1e-14)
        assert abs(sa.inv_quasi_gauss(y, b) - root) < 2e-8
    # no root in [0, 2]: the bracket closes on its upper end
    assert abs(sa.inv_quasi_gauss(1.5, b) - 2.0) < 2e-8
    assert abs(sa.inv_quasi_gauss(q2 * 0.5, b) - 2.0) < 2e-8


def test_extension_count_hand_derived_spectra():
    spec = [1.0, 0.5, 0.25, 0.1, 1e-2, 1e-3]
    zero = lambda j: 0.0
    # r = 0: the geometric continuation 1e-4 ... 1e-8 stays above 3e-9, 1e-9 does not
    assert sa.extension_count(spec, 100, 3e-9, 1.0, zero) == 5
    # maxdim caps the spectrum at 9 values
    assert sa.extension_count(spec, 9, 3e-9, 1.0, zero) == 3
    # nothing is added at 5 values or at maxdim
    assert sa.extension_count(spec[:5], 100, 3e-9, 1.0, zero) == 0
    assert sa.extension_count(spec, 6, 3e-9, 1.0, zero) == 0
    # r = 0.5, sign + at even i + n: log10 values -3.5 -4.5 -5 -6 -6.5 -7.5 -8, then -9
    assert sa.extension_count(spec, 100, 3e-9, 1.0, lambda j: 0.5) == 7
    # r = 1 from a flat tail: -2 is clamped to -3, then -4 -4 -5 -5 ... -8 -8, then -9
    flat = [1.0, 0.5, 0.25, 0.1, 1e-3, 1e-3]
    assert sa.extension_count(flat, 100, 3e-9, 1.0, lambda j: 1.0) == 11


def _site_tensor(rng):
    # (left, phys, right) dirs (+1, +1, -1); right leg narrower than left x phys
    dl = {0: 3, 1: 4, 2: 2}
    ds = {0: 2, 1: 2}
    dr = {0: 2, 1: 3, 2: 3, 3: 1}
    t = Tensor((1, 1, -1))
    for a, na in dl.items():
        for s, ns in ds.items():
            b = a + s
            if b in dr:
                t.push_back_ch((a, s, b), rng.standard_normal((na, ns, dr[b])) + 1j * rng.standard_normal((na, ns, dr[b])))
    return t


def _dense_sectors(t, r1):
    rows, cols = T._sector_layout(t, r1)
    out = {}
    for qr in rows:
        if -qr not in cols:
            continue
        rlist, nr = rows[qr]
        clist, nc = cols[-qr]
        m = np.zeros((nr, nc), dtype=np.complex128)
        for (rq, _, rp, rs) in rlist:
            for (cq, _, cp, cs) in clist:
                b = t.blocks.get(tuple(rq) + tuple(cq))
                if b is not None:
                    m[rp:rp + rs, cp:cp + cs] = b.reshape(rs, cs)
        out[qr] = m
    return out


@pytest.mark.parametrize("l2r", [True, False])
def test_sa_svd_keeps_the_tensor_and_extends_an_orthonormal_side(l2r):
    rng = np.random.default_rng(11)
    A = _site_tensor(rng)
    if not l2r:  # the mirrored tensor (left, phys, right) -> columns wider than rows
        A = Tensor((1, -1, -1), {(q[2], q[1], q[0]): b.transpose(2, 1, 0) for q, b in A.blocks.items()})
        r1 = 1
    else:
        r1 = 2
    n = sum(min(m.shape) for m in _dense_sectors(A, r1).values())
    lef, rig, added = sa.sa_svd_impl(A, r1, 1e-12, 40, l2r, 2.0, seed=5, call=3)
    assert n > 5 and added > 0
    rec = product(lef, rig, 1)
    for q, b in A.blocks.items():
        assert np.abs(rec.blocks[q] - b).max() < 1e-12
    # the isometric side, new states included, is orthonormal sector by sector
    iso = lef if l2r else rig
    legs = iso.legs()
    bond = 2 if l2r else 0
    dims = sum(legs[bond].values())
    assert dims == n + added
    for qb, d in legs[bond].items():
        blocks = [(q, b) for q, b in iso.blocks.items() if q[bond] == qb]
        if l2r:
            m = np.concatenate([b.reshape(-1, d) for _, b in sorted(blocks)], axis=0)
        else:
            m = np.concatenate([b.reshape(d, -1) for _, b in sorted(blocks)], axis=1).T
        assert np.abs(m.conj().T @ m - np.eye(d)).max() < 1e-13
    # without room for new states it is the untruncated SVD
    l0, r0 = T.svd_impl(A, r1, -1.0, 10 ** 6, l2r)
    l1, r1_, add1 = sa.sa_svd_impl(A, r1, 1e-12, n, l2r, 2.0, seed=5, call=3)
    assert add1 == 0
    for x, y in ((l0, l1), (r0, r1_)):
        assert set(x.blocks) == set(y.blocks)
        for q in x.blocks:
            assert np.abs(x.blocks[q] - y.blocks[q]).max() == 0.0


def test_parse_svs_follows_the_reference_bond_mapping():
    # N = 3: records 1-2 open bonds 0, 1; 3 -> bond 1, 4 -> bond 0 (right-to-left);
    # second sweep: 5 -> 1 (5 mod 4), 6 -> list 2 missing -> 1, 7 -> 4 - 3 = 1, 8 -> 0
    recs = [(3 + k, [float(10 * k + j) for j in range(k + 1)]) for k in range(8)]
    lists = sa.parse_svs(sa.svs_text(3, recs))
    v = [r[1] for r in recs]
    assert lists[0] == v[0] + v[3] + v[7]
    assert lists[1] == v[1] + v[2] + v[4] + v[5] + v[6]
    assert len(lists) == 2


def test_svs_text_layout():
    txt = sa.svs_text(4, [(5, [-1.25e-3, 3.5, 0.0])])
    lines = txt.splitlines(keepends=True)
    assert lines[0] == "Total Site = 4\n" and lines[1] == "5\n"
    assert all(len(l) == 21 for l in lines[2:])
    assert lines[2].startswith("-1.25000000000e-03")


def test_fit_widths_minimises_the_bin_residual():
    rng = np.random.default_rng(3)
    vals = rng.laplace(0.0, 0.2, size=400)
    txt = sa.svs_text(2, [(len(vals) + 2, list(vals)), (3, [0.1])])
    w = sa.fit_widths(txt)
    assert len(w) == 1
    lst = sa.parse_svs(txt)[0]  # the second record (bond 0, right to left) joins the first
    assert len(lst) == 401
    ni, edges = np.histogram(lst, 50)
    x, y = edges[:-1], ni / len(lst)

    def f(b):
        return float(np.sum((sa.fit_model(x, b) - y) ** 2))

    b = w[0]
    # a minimum to the fit's own tolerance (relative cost change 1e-8)
    assert b > 0 and f(b) <= f(b * 1.001) * (1 + 1e-8) and f(b) <= f(b / 1.001) * (1 + 1e-8)
    grid = np.exp(np.linspace(np.log(1e-6), np.log(1e3), 400))
    assert f(b) <= min(f(g) for g in grid) * (1 + 1e-8)
    # bonds with at most two values get width 1
    txt2 = sa.svs_text(3, [(4, [0.1, 0.2]), (5, [0.1, 0.2, 0.3])])
    assert sa.fit_widths(txt2)[0] == 1.0
    assert sa.betas_from_widths(sa.widths_text([0.25, 2.0])) == [4.0, 0.5]


def _ed_ct(N, om, g, Ls, J, dt, nsteps):
    H = M.ed_hamiltonian(N, om, g, Ls, J)
    v0 = M.ed_initial_state(N, Ls)
    w, V = np.linalg.eigh(H)
    c0 = V.conj().T @ v0
    return np.array([np.vdot(v0, V @ (np.exp(-1j * w * dt * k) * c0)) for k in range(1, nsteps + 1)])


def test_sa_one_site_tracks_exact_dynamics_closer_than_one_site():
    """From a truncated 2TDVP state, plain 1TDVP is stuck at its bond dimensions; the SA
    sweep adds directions and follows the exact dynamics more closely."""
    N, Ls, om, J, dt = 4, [3], [1.0], 1.0, 0.05
    g = np.array([0.9, 1.1, 0.8, 1.0]).reshape(4, 1)
    ed = _ed_ct(N, om, g, Ls, J, dt, 22)
    lat = M.holstein(N, om, g, Ls, J, "merged")
    mpo = M.build_mpo(lat)
    psi0 = M.product_state(lat)
    eng = T.TDVP(psi0, mpo, cutoff=1e-10, maxdim=6, dt=dt)
    for _ in range(10):
        eng.two_site()
    assert max(eng.bond_dims()) == 6
    start = eng.mps
    one = T.TDVP(start, mpo, cutoff=1e-10, maxdim=6, dt=dt)
    sae = T.TDVP(start, mpo, cutoff=1e-10, maxdim=40, dt=dt)
    betas = [1.0 / 0.3] * (N - 1)
    e1, e2 = [], []
    for k in range(12):
        one.one_site()
        sae.one_site(betas=betas, seed=100 + k)
        e1.append(abs(T.overlap_states(psi0, one.mps) - ed[10 + k]))
        e2.append(abs(T.overlap_states(psi0, sae.mps) - ed[10 + k]))
    assert max(sae.bond_dims()) > 6
    assert max(e2) < 0.5 * max(e1)


def _perturbed(sites, rng, eps):
    out = []
    for t in sites:
        c = t.copy()
        for q in c.blocks:
            b = c.blocks[q]
            c.blocks[q] = b + eps * np.abs(b).max() * (rng.standard_normal(b.shape) + 1j * rng.standard_normal(b.shape))
        out.append(c)
    return out


@pytest.mark.parametrize("layout", ["merged", "pp"])
def test_sa_sweep_sensitivity_to_start_noise(layout):
    """How far 1e-13 noise on the start moves the oracle's own SA sweeps (the workload of
    the GPU comparison): merged keeps every bond dimension; pp does not (its counts hang on
    singular values near 1e-8), but its state stays within 1e-8."""
    from oracle import krylov as K
    N, Ls, om, J, dt = 4, [4], [1.0], 1.0, 0.05
    g = np.array([0.9, 1.1, 0.8, 1.0]).reshape(4, 1)
    lat = M.holstein(N, om, g, Ls, J, layout)
    mpo = M.build_mpo(lat)
    p0 = M.product_state(lat)
    psi0 = K.global_main(K.scale_mpo(mpo, dt), p0, 1e-10, 24)[0] if layout == "pp" else p0
    ref = T.TDVP(psi0, mpo, cutoff=1e-10, maxdim=6, dt=dt)
    for _ in range(6):
        ref.two_site()
    betas = [1.0 / w for w in sa.fit_widths(sa.svs_text(len(psi0), ref.svs))]
    runs = []
    for start in (ref.mps, _perturbed(ref.mps, np.random.default_rng(0), 1e-13)):
        e = T.TDVP(start, mpo, cutoff=1e-10, maxdim=40, dt=dt)
        dims, cs = [], []
        for k in range(4):
            e.one_site(betas=betas, seed=77 + k)
            dims.append(e.bond_dims())
            cs.append(T.overlap_states(p0, e.mps))
        runs.append((dims, np.array(cs), e.mps))
    (d0, c0, m0), (d1, c1, m1) = runs
    assert np.abs(c0 - c1).max() < 1e-8
    assert 1 - abs(T.overlap_states(m0, m1)) < 1e-8
    if layout == "merged":
        assert d0 == d1
    else:
        assert d0 != d1
