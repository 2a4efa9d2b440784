"""A compiled Fortran program calls the library the way the reference's callers call libcc:
common_cc_init, ccx_register_vvvo and cc_triples_acc through the Fortran-77 ABI with 8-byte
INTEGER (src/triples-ijk/triples.F90:1-12, src/common/common_cc.f90:342), and the context API
through an iso_c_binding interface block (tests/fortran/ccx_caller.f90, built by
tests/fortran/Makefile).  Checked at no=4, nu=36 (seed 1001) against the compiled reference's
golden energies: |dE| <= 1e-9 Eh on both routes, the in-place t1 transpose of cc_triples_acc
(triples.F90:311), and the two routes' v1 equal."""

import os
import subprocess
import tempfile

import numpy as np
import pytest

from libcc_amd import synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CALLER = os.path.join(HERE, "fortran", "_build", "ccx_caller")
KEYS = ("eh", "ep", "t1", "t2", "vm", "voe", "vvvo")


def test_fortran_caller_matches_reference_golden(golden):
    if not os.path.exists(CALLER):
        pytest.fail(f"{CALLER} is not built (make -C tests/fortran)")
    no, nu, seed = 4, 36, 1001
    inp = synth.make_inputs(no, nu, seed)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([no, nu], dtype=np.int64).tofile(f)
            for k in KEYS:
                np.ravel(inp[k], order="F").astype(np.float64).tofile(f)
        print(f"running {CALLER} no={no} nu={nu} seed={seed}", flush=True)
        subprocess.run([CALLER, fin, fout], check=True, timeout=120)
        o = np.fromfile(fout, dtype=np.float64)
    st_legacy, gms_ets, gms_etd, st_ctx, ets, etd, tdiff, dv1 = o
    ets_want, etd_want = golden[f"e_{no}_{nu}_{seed}"]
    e_want = ets_want + etd_want
    print(f"legacy |dE| {abs(gms_ets + gms_etd - e_want):.2e}, context |dE| {abs(ets + etd - e_want):.2e} Eh")
    assert st_legacy == 0 and st_ctx == 0
    assert abs(gms_etd - etd_want) <= 1e-9 and abs(gms_ets - ets_want) <= 1e-9
    assert abs(etd - etd_want) <= 1e-9 and abs(ets - ets_want) <= 1e-9
    assert tdiff == 0.0          # t1 came back as its exact transpose (nu,no)
    assert dv1 == 0.0            # both routes run the same kernels in the same order
