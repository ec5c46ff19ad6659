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


@pytest.mark.parametrize("value", [-5, 2 ** 31 - 1])
def test_corrupted_unformatted_header_is_an_error(evr, basis, tmp_path, value):
    """nb_tot read from a file is checked against the file before anything is
    sized from it: a negative or huge count returns EVR_ERR_IO (the host
    process survives)."""
    path = tmp_path / "wpu"
    basis.save_psi(path, _psi(basis.nb_packed, 1, True), version=1, formatted=False)
    raw = bytearray(open(path, "rb").read())
    # records: [4]'X'[4] [4]nb_psi[4] [4]nb_tot[4] ...: nb_tot at byte 25
    assert int.from_bytes(raw[25:29], "little", signed=True) == basis.nb_packed
    raw[25:29] = int(value).to_bytes(4, "little", signed=True)
    open(path, "wb").write(bytes(raw))
    with pytest.raises(RuntimeError, match="do not fit the file"):
        basis.read_psi(path, cplx=True)


@pytest.mark.parametrize("field,value", [("nb_tot", 2000000000), ("nb_tot", -3), ("nb_psi", -1)])
def test_corrupted_namelist_header_is_an_error(evr, basis, tmp_path, field, value):
    path = tmp_path / "wpn"
    basis.save_psi(path, _psi(basis.nb_packed, 1, True), version=1, formatted=True)
    text = open(path).read()
    import re
    text, n = re.subn(r"(?i)(%s\s*=\s*)(-?\d+)" % field, lambda m: m.group(1) + str(value), text, count=1)
    assert n == 1
    open(path, "w").write(text)
    with pytest.raises(RuntimeError, match="do not fit the file"):
        basis.read_psi(path, cplx=True)


def test_restart_round_trip(evr, tmp_path):
    """SaveWP_restart / ReadWP_restart: "T nb_WP size_WP" then the complex
    coefficients, read back exactly."""
    rng = np.random.default_rng(2)
    psi = rng.standard_normal((37, 3)) + 1j * rng.standard_normal((37, 3))
    path = tmp_path / "restart"
    evr.save_restart(path, 123.456, psi)
    first = open(path).readline().split()
    assert float(first[0]) == 123.456 and first[1:] == ["3", "37"]
    T, got, found = evr.read_restart(path, 3, 37)
    assert found and T == 123.456
    assert np.array_equal(got, psi)


def test_restart_gfortran_list_directed(evr, tmp_path):
    """A restart file in gfortran's list-directed layout (several complex
    constants per record, blanks inside the parentheses) is read."""
    psi = np.array([1.5 - 2j, 0.25 + 0j, -3e-17 + 4.5j, 2 + 1j])
    items = ["(%25.17E,%25.17E)" % (z.real, z.imag) for z in psi]
    text = "   2.0000000000000000                1           4\n" + \
        " " + "  ".join(items[:3]) + "\n " + items[3] + "\n"
    path = tmp_path / "restart_gf"
    open(path, "w").write(text.replace("E", "E"))
    T, got, found = evr.read_restart(path, 1, 4)
    assert found and T == 2.0
    assert np.allclose(got[:, 0], psi, rtol=0, atol=0)


def test_restart_missing_header_and_size_mismatch(evr, tmp_path):
    """No readable "T nb_WP size_WP" record: no restart (T0 = 0, found =
    False), as the reference's WARNING; wrong sizes: an error (the reference
    STOPs)."""
    path = tmp_path / "r0"
    open(path, "w").write("not a restart file\n")
    T, _, found = evr.read_restart(path, 1, 4)
    assert not found and T == 0.0
    evr.save_restart(path, 1.0, np.ones(5, complex))
    with pytest.raises(RuntimeError, match="restart file holds 1 wavepackets of 5"):
        evr.read_restart(path, 1, 4)
