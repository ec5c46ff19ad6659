"""Wavepacket files in the reference's save-file formats (evr_psi_save /
evr_psi_read, host code: no GPU needed).  Pinned by the reference's own
initial-wavepacket line of the pyrazine tests (the version-0 line format),
by the oracle's packed-basis enumeration (the index tuples) and by exact
round trips of every form."""
import os

import numpy as np
import pytest

from pyrazine import P12, P24

FORMS = [(0, True), (1, True), (1, False)]


@pytest.fixture(scope="module")
def basis(evr):
    return evr.Basis.sg4(4, 3, 3, 1, 2, 2, 2)


def _psi(n, nwp, cplx, seed=0):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((n, nwp))
    if cplx:
        p = p + 1j * rng.standard_normal((n, nwp))
    return p


@pytest.mark.parametrize("version,formatted", FORMS)
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("nb0", [1, 2])
def test_round_trip_exact(evr, basis, tmp_path, version, formatted, cplx, nb0):
    psi = _psi(nb0 * basis.nb_packed, 3, cplx)
    path = tmp_path / "wp"
    basis.save_psi(path, psi, nb0=nb0, version=version, formatted=formatted)
    got = basis.read_psi(path, nb0=nb0, cplx=cplx, max_wp=5)
    assert got.shape == psi.shape
    assert np.array_equal(got, psi)          # 17 significant digits: exact


def test_index_tuples_follow_packed_enumeration(evr, oracle_mod, basis, tmp_path):
    """The index list of a version-1 file is the oracle's packed-basis order
    (type-5 nDindex, last mode fastest), 1-based, with i_h = 1 and i_e."""
    s = oracle_mod.SG4(4, 3, 3, 1, 2, 2, 2)
    path = tmp_path / "wp1"
    basis.save_psi(path, np.zeros((2 * basis.nb_packed, 1)), nb0=2, version=1)
    lines = open(path).read().splitlines()
    k = [i for i, l in enumerate(lines) if l.strip() == "/"][0]
    idx = np.array([[int(v) for v in l.split()] for l in lines[k + 1:k + 1 + 2 * basis.nb_packed]])
    tup = s.packed() + 1
    n = s.nb_packed
    assert np.array_equal(idx[:n, :4], tup) and np.array_equal(idx[n:, :4], tup)
    assert (idx[:, 4] == 1).all()
    assert (idx[:n, 5] == 1).all() and (idx[n:, 5] == 2).all()


@pytest.mark.parametrize("P", [P12, P24], ids=["12d", "24d"])
def test_reference_wp0_line(evr, tmp_path, P):
    """The pyrazine inputs' initial wavepacket line (functions 1 of every mode,
    i_h = 1, state 2, 1.0 + 0i) read as a version-0 file lands on the ground
    function of state 2."""
    D, LG, LB, nbA, nbB, nqA, nqB = P.sg4_args()
    b = evr.Basis.sg4_modes(D, LG, nbA, nbB, nqA, nqB, LB=LB)
    path = tmp_path / "wp0"
    path.write_text("1 %d\n%s\nend wp0\n" % (2 * b.nb_packed, P.fix["wp0_line"]))
    got = b.read_psi(path, nb0=2, cplx=True, max_wp=1)[:, 0]
    want = np.zeros(2 * b.nb_packed, complex)
    want[b.nb_packed] = 1.0
    assert np.array_equal(got, want)


def test_projection_between_bases(evr, tmp_path):
    """A file written on the L=2 basis read on L=3 keeps every common function
    and zeros the new ones; read back on L=2 the extra functions are dropped."""
    b2 = evr.Basis.sg4(4, 2, 2, 1, 2, 2, 2)
    b3 = evr.Basis.sg4(4, 3, 3, 1, 2, 2, 2)
    p2, p3 = tmp_path / "l2", tmp_path / "l3"
    psi2 = _psi(b2.nb_packed, 2, True, 4)
    b2.save_psi(p2, psi2, version=1)
    up = b3.read_psi(p2, cplx=True, max_wp=2)
    # L=2 functions are the L=3 ones whose level sum is <= 2
    t2, t3 = _tuples(b2, tmp_path), _tuples(b3, tmp_path)
    pos = {tuple(t): i for i, t in enumerate(t3)}
    rows = np.array([pos[tuple(t)] for t in t2])
    assert np.array_equal(up[rows], psi2)
    others = np.setdiff1d(np.arange(b3.nb_packed), rows)
    assert not up[others].any()
    b3.save_psi(p3, up, version=0)
    assert np.array_equal(b2.read_psi(p3, cplx=True, max_wp=2), psi2)


def _tuples(b, tmp_path):
    path = tmp_path / ("t%d" % b.nb_packed)
    b.save_psi(path, np.zeros((b.nb_packed, 1)), version=1)
    lines = open(path).read().splitlines()
    k = [i for i, l in enumerate(lines) if l.strip() == "/"][0]
    return [[int(v) for v in l.split()[:4]] for l in lines[k + 1:k + 1 + b.nb_packed]]


def test_fortran_list_directed_forms(evr, basis, tmp_path):
    """Values as gfortran and other compilers write them: padded complex
    constants, D exponents, repeat counts, several values per line."""
    n = basis.nb_packed
    path = tmp_path / "wp1"
    basis.save_psi(path, np.zeros((n, 1)), version=1)
    text = open(path).read()
    head = text[:text.index("/") + 2]
    idx = text[len(head):].splitlines()[:n]
    vals = ["(  1.0000000000000000     , -2.5000000000000000D-01 )", "%d*(0.0,0.0)" % (n - 2),
            "( 3.0E+00 , 0.5 )"]
    path.write_text(head + "\n".join(idx) + "\n" + "  ".join(vals) + "\n")
    got = basis.read_psi(path, cplx=True, max_wp=1)[:, 0]
    want = np.zeros(n, complex)
    want[0], want[-1] = 1.0 - 0.25j, 3.0 + 0.5j
    assert np.array_equal(got, want)


def test_errors(evr, basis, tmp_path):
    with pytest.raises(RuntimeError, match="cannot open"):
        basis.read_psi(tmp_path / "missing")
    bad = tmp_path / "bad"
    bad.write_text("hello\n")
    with pytest.raises(RuntimeError, match="header"):
        basis.read_psi(bad)
    with pytest.raises(RuntimeError, match="unformatted"):
        basis.save_psi(tmp_path / "x", np.zeros(basis.nb_packed), version=0, formatted=False)
