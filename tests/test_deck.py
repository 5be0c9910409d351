"""The deck reader (fvc_deck_read, host/fvc_io.c; ompVC/parameters.c:20-292)
against known answers: the values the reference program itself echoed after
reading each deck shipped with ompVC (tests/golden/decks, written by
make_deck_goldens.py).  The entries are laid out again between comment and
blank lines of our own, as the shipped decks interleave theirs."""
import ctypes
import json
import os

import pytest

import cases
from fvcpy import deck, host

DECKS = ["cartesian"]
# names the reference's echo uses for the entries (parameters.c:142-230)
ECHO_NAMES = {"Particle_ICType": "Particle_ICflag"}


def _echo_name(key):
    key = ECHO_NAMES.get(key, key)
    for a in ("FTLE", "Trace", "VelOut"):
        key = key.replace(a + "_MeshBounds.", a + "_CartMesh.")
    return key.replace("Gravity_Vector[", "Gravity[")


def _parse(path, tmp):
    L = host.lib()
    L.fvc_deck_read.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
    L.fvc_deck_write.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    buf = ctypes.create_string_buffer(1 << 20)
    err = ctypes.create_string_buffer(512)
    rc = L.fvc_deck_read(path.encode(), buf, err, 512)
    assert rc == 0, err.value
    out = os.path.join(tmp, "canonical.in")
    assert L.fvc_deck_write(out.encode(), buf) == 0
    return deck.read_deck(out)


@pytest.mark.parametrize("name", DECKS)
def test_shipped_decks_read_as_the_reference_reads_them(name, tmp_path, host_built):
    fx = os.path.join(cases.HERE, "decks")
    entries = open(os.path.join(fx, name + ".entries")).read().splitlines()
    echo = json.load(open(os.path.join(fx, name + ".echo.json")))["echo"]
    text = "########\n# known-answer deck\n#\n\n" + "".join("# entry %d\n\n%s\n" % (k, e) for k, e in enumerate(entries))
    path = os.path.join(str(tmp_path), name + ".in")
    with open(path, "w") as f:
        f.write(text)
    got = _parse(path, str(tmp_path))
    keys = [(k, t) for k, t, _ in deck.KEYS if k != "Num_threads"]  # the one entry the reference does not echo
    assert len(echo) == len(keys) == 96
    for (ename, evalue), (key, t) in zip(echo, keys):
        assert ename == _echo_name(key)
        if t == "s":
            assert got[key] == evalue, key
        elif t == "d":
            assert got[key] == int(evalue), key
        else:  # echoed with %g: six significant digits
            assert got[key] == pytest.approx(float(evalue), rel=1e-5, abs=0.0), key


def test_decks_without_num_threads(tmp_path, host_built):
    """Decks written before Num_threads was added (96 entries, FTLE_Compute
    where Num_threads is expected) are read with Num_threads = 0 and every
    later entry in its place; the reference reads them shifted by one."""
    fx = os.path.join(cases.HERE, "decks")
    entries = open(os.path.join(fx, "cartesian.entries")).read().splitlines()
    assert entries[38].startswith("Num_threads")
    path = os.path.join(str(tmp_path), "old.in")
    with open(path, "w") as f:
        f.write("\n".join(entries[:38] + entries[39:]) + "\n")
    got = _parse(path, str(tmp_path))
    ref = _parse(os.path.join(fx, "cartesian.entries"), str(tmp_path))
    assert got["Num_threads"] == 0
    for k, _, _ in deck.KEYS:
        if k != "Num_threads":
            assert got[k] == ref[k], k
