"""The reference's triples.restart file (src/triples-ijk/program.F90:44-46,195-218) through the
library's C reader and writer.  The framing is checked by an independent parser here and by
the file sizes GAMESS logged for the reference's water runs.  No GPU needed."""

import os

import numpy as np
import pytest

from libcc_amd import CcxError, synth
from libcc_amd.triples import read_restart, restart_dims, write_restart


def parse(path):
    """Fortran unformatted sequential records with 4-byte markers -> list of raw bytes."""
    raw = open(path, "rb").read()
    recs, o = [], 0
    while o < len(raw):
        n = int(np.frombuffer(raw, np.int32, 1, o)[0])
        recs.append(raw[o + 4:o + 4 + n])
        assert int(np.frombuffer(raw, np.int32, 1, o + 4 + n)[0]) == n
        o += 8 + n
    return recs


def expected_size(no, nu):
    nou, nutr = no * nu, nu * (nu + 1) // 2
    return 16 + 8 * (no + nu + nou + 2 * no * no * nu * nu + no ** 3 * nu + nou * nutr) + 8 * (nou + 8)


@pytest.mark.parametrize("no,nu,size", [(4, 36, 1120144), (4, 87, 12645736), (4, 167, 82202856)])
def test_size_matches_gamess_logged_files(no, nu, size):
    """The formula the writer follows gives the byte counts GAMESS printed for the reference's
    water aug-cc-pVDZ / TZ / QZ restart files."""
    assert expected_size(no, nu) == size


def test_round_trip_and_framing(tmp_path):
    no, nu, seed = 3, 5, 7
    inp = synth.make_inputs(no, nu, seed)
    path = str(tmp_path / "triples.restart")
    write_restart(path, inp)
    assert os.path.getsize(path) == expected_size(no, nu)
    recs = parse(path)
    assert len(recs) == 8 + no * nu
    assert np.frombuffer(recs[0], np.int64)[0] == no and np.frombuffer(recs[1], np.int64)[0] == nu
    for r, key in zip(recs[2:8], ("eh", "ep", "t1", "t2", "vm", "voe")):
        np.testing.assert_array_equal(np.frombuffer(r, np.float64), np.ravel(inp[key], order="F"))
    # VVVO record j = column c + nu*x, lower-packed rows p(p+1)/2 + q (q <= p)
    vv = inp["vvvo"]
    for x in range(no):
        for c in range(nu):
            col = np.frombuffer(recs[8 + c + nu * x], np.float64)
            want = [vv[p, q, c, x] for p in range(nu) for q in range(p + 1)]
            np.testing.assert_array_equal(col, want)
    back = read_restart(path)
    assert restart_dims(path) == (no, nu)
    for key in ("eh", "ep", "t1", "t2", "vm", "voe"):
        np.testing.assert_array_equal(back[key], inp[key])
    np.testing.assert_array_equal(back["vvvo"], synth.pack_vvvo(vv))
    # packed input writes the same bytes
    path2 = str(tmp_path / "packed.restart")
    write_restart(path2, dict(inp, vvvo=synth.pack_vvvo(vv)), packed_vvvo=True)
    assert open(path, "rb").read() == open(path2, "rb").read()


def test_bad_files_rejected(tmp_path):
    inp = synth.make_inputs(2, 4, 3)
    path = str(tmp_path / "t.restart")
    write_restart(path, inp)
    raw = open(path, "rb").read()
    short = str(tmp_path / "short.restart")
    open(short, "wb").write(raw[:-5])
    with pytest.raises(CcxError) as e:
        read_restart(short)
    assert e.value.status == 7
    bad = bytearray(raw)
    bad[0] = 9                                    # first marker no longer 8
    badp = str(tmp_path / "bad.restart")
    open(badp, "wb").write(bytes(bad))
    with pytest.raises(CcxError):
        restart_dims(badp)
    with pytest.raises(CcxError):
        restart_dims(str(tmp_path / "missing.restart"))
