"""ccx_energy_static, the entry point that allocates nothing (every work buffer in one static
device array per GPU), against the context entry points it mirrors (ccx_load_device +
ccx_energy): bitwise equal E[T], E_ST and v1 on two calls in a row, the second on a smaller
problem (the array keeps the first call's contents), plus the size and input rejections."""

import ctypes

import numpy as np
import pytest

from libcc_amd import lib, synth
from libcc_amd.triples import TriplesContext

pytestmark = pytest.mark.gpu

CCX_ERR_INPUT, CCX_ERR_SCRATCH = 5, 8
KEYS = ("eh", "ep", "t1", "t2", "vm", "voe", "vvvo")


def _device_buffers(torch, inp):
    # Fortran-ordered float64 buffers on the GPU, allocated before the call (by the caller)
    return [torch.from_numpy(np.asfortranarray(inp[k]).ravel(order="F").copy()).to("cuda:0") for k in KEYS]


def _static(torch, no, nu, bufs):
    e = torch.full((2,), np.nan, dtype=torch.float64, device="cuda:0")
    v1 = torch.full((nu * no,), np.nan, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    rc = lib().ccx_energy_static(no, nu, *[ctypes.c_void_p(b.data_ptr()) for b in bufs],
                                 ctypes.c_void_p(e.data_ptr()), ctypes.c_void_p(v1.data_ptr()), None)
    torch.cuda.synchronize()
    return rc, e.cpu().numpy(), v1.cpu().numpy().reshape(nu, no, order="F")


@pytest.mark.parametrize("order", ["large_then_small"])
def test_static_entry_matches_context_entry(order):
    import torch

    for no, nu, seed in ((4, 87, 1002), (3, 40, 7)):
        print(f"ccx_energy_static no={no} nu={nu} seed={seed}", flush=True)
        inp = synth.make_inputs(no, nu, seed)
        bufs = _device_buffers(torch, inp)
        rc, e, v1 = _static(torch, no, nu, bufs)
        assert rc == 0, lib().ccx_error_string(rc).decode()
        ctx = TriplesContext(no, nu)
        ctx.load({k: b for k, b in zip(KEYS, bufs)})
        want = ctx.energy("fp64")
        ctx.close()
        print(f"  etd {e[0]:.15e} / {want['etd']:.15e}, ets {e[1]:.15e} / {want['ets']:.15e}", flush=True)
        assert e[0] == want["etd"] and e[1] == want["ets"]
        assert np.array_equal(v1, want["v1"])


def test_static_entry_rejections():
    import torch

    # no=8, nu=120 needs about 270 MB of slabs: refused before any device work
    rc = lib().ccx_energy_static(8, 120, *[ctypes.c_void_p(16)] * 9, ctypes.c_void_p(16), ctypes.c_void_p(16), None)
    assert rc == CCX_ERR_SCRATCH
    no, nu = 3, 20
    inp = synth.make_inputs(no, nu, 5)
    inp["t2"] = np.array(inp["t2"], order="F")
    inp["t2"][1, 2, 0, 1] = np.nan
    rc, _, _ = _static(torch, no, nu, _device_buffers(torch, inp))
    assert rc == CCX_ERR_INPUT
    # one occupied orbital: no tuple, E(T) = 0 and v1 = 0
    inp = synth.make_inputs(1, 9, 3)
    rc, e, v1 = _static(torch, 1, 9, _device_buffers(torch, inp))
    assert rc == 0 and np.all(e == 0.0) and np.all(v1 == 0.0)
