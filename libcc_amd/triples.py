"""Python mirror of the reference's triples driver interface over libcc_mi355x.so.

``cc_triples(eh, ep, t1, t2, vm, voe, vvvo)`` mirrors ``cc_triples_acc``
(src/triples-ijk/triples.F90:1-12,319-320): same arrays, same layouts, returns
(ets, etd, v1) with v1(nu,no).  Unlike the Fortran routine it does not transpose
``t1`` in place (the legacy C entry ``cc_triples_acc_`` does, for drop-in
fidelity).  ``TriplesContext`` exposes the context API: host or device-resident
inputs (numpy arrays or torch tensors on the GPU), FP64 energies, per-tuple W.
``read_restart`` / ``write_restart`` handle the reference's ``triples.restart`` file.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import CcxError, lib

FORMS = {"fp64": 0, "fp32": 1, "bf16x3": 2, "fp8x4": 3, "bf16x4": 4}
_dp = ctypes.POINTER(ctypes.c_double)


def _host(a) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    return np.asfortranarray(a)


def _fortran_buffer_np(a, shape: tuple, name: str) -> np.ndarray:
    """The Fortran-ordered float64 buffer of a host array given in its logical shape or as a
    1-D buffer of the right length."""
    a = np.asarray(a)
    n = int(np.prod(shape))
    if a.ndim == 1 and len(shape) > 1:
        if a.size != n:
            raise ValueError(f"{name}: 1-D buffer of {a.size} elements, expected {n} for {shape}")
        return np.ascontiguousarray(a, dtype=np.float64)
    if tuple(a.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(a.shape)}, expected {tuple(shape)} (or a 1-D buffer of {n})")
    return np.asfortranarray(a, dtype=np.float64)


def _fortran_buffer_t(t, shape: tuple, name: str):
    """The Fortran-ordered buffer (1-D, contiguous) of a device tensor in its logical shape, or
    the tensor itself when it already is a contiguous 1-D buffer of the right length."""
    n = int(np.prod(shape))
    if t.dim() == 1 and len(shape) > 1:
        if t.numel() != n:
            raise ValueError(f"{name}: 1-D buffer of {t.numel()} elements, expected {n} for {shape}")
        if not t.is_contiguous():
            raise ValueError(f"{name}: a 1-D buffer must be contiguous")
        return t
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)} (or a 1-D buffer of {n})")
    return t.permute(*reversed(range(t.dim()))).contiguous().reshape(-1)


def _check(status: int, what: str) -> None:
    if status != 0:
        raise CcxError(status, what)


class TriplesContext:
    """A device-resident (T) problem of no occupied / nu virtual orbitals."""

    def __init__(self, no: int, nu: int):
        self.no, self.nu = int(no), int(nu)
        self._ctx = ctypes.c_void_p()
        _check(lib().ccx_create(self.no, self.nu, ctypes.byref(self._ctx)), "ccx_create")
        self._keep = None

    def close(self) -> None:
        if self._ctx:
            lib().ccx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shapes(self, packed_vvvo: bool = False) -> dict:
        """Logical shapes of the inputs in the reference layouts."""
        no, nu = self.no, self.nu
        return {"eh": (no,), "ep": (nu,), "t1": (no, nu), "t2": (nu, nu, no, no), "vm": (no, nu, no, no),
                "voe": (nu, nu, no, no),
                "vvvo": (nu * (nu + 1) // 2, nu * no) if packed_vvvo else (nu, nu, nu, no)}

    def load(self, inp: dict, packed_vvvo: bool = False) -> None:
        """Upload inputs in the reference layouts (numpy, host) or use torch GPU tensors in place.

        Every array is either in its logical shape (indexed as the reference indexes it, e.g.
        t1[i, a] = T1(i,a), any strides) or a 1-D buffer holding the Fortran-ordered array; a
        1-D tensor on the device must be contiguous.  Sizes, dtype (float64) and, for tensors,
        the device (the current one) are checked before anything is handed to the library."""
        keys = ("eh", "ep", "t1", "t2", "vm", "voe", "vvvo")
        shapes = self.shapes(packed_vvvo)
        layout = 1 if packed_vvvo else 0
        if any(hasattr(inp[k], "is_cuda") and inp[k].is_cuda for k in keys):
            import torch

            dev = torch.cuda.current_device()
            bufs = []
            for k in keys:
                t = inp[k]
                if not (hasattr(t, "is_cuda") and t.is_cuda):
                    raise TypeError(f"{k}: every input must be a tensor on the GPU when one is")
                if t.dtype != torch.float64:
                    raise TypeError(f"{k} must be float64, got {t.dtype}")
                if t.device.index != dev:
                    raise ValueError(f"{k} is on cuda:{t.device.index}, the library uses cuda:{dev}")
                bufs.append(_fortran_buffer_t(t, shapes[k], k))
            self._keep = bufs
            _check(lib().ccx_load_device(self._ctx, *[ctypes.c_void_p(b.data_ptr()) for b in bufs], layout),
                   "ccx_load_device")
        else:
            arrs = [_fortran_buffer_np(inp[k], shapes[k], k) for k in keys]
            _check(lib().ccx_load_host(self._ctx, *[a.ctypes.data_as(_dp) for a in arrs], layout),
                   "ccx_load_host")

    def energy(self, form: str = "fp64", v1_device=None) -> dict:
        ets, etd = ctypes.c_double(), ctypes.c_double()
        if form not in FORMS:
            raise ValueError(f"unknown form {form!r}; one of {sorted(FORMS)}")
        if v1_device is not None:
            import torch

            if (not v1_device.is_cuda or v1_device.dtype != torch.float64 or not v1_device.is_contiguous()
                    or v1_device.numel() != self.no * self.nu or v1_device.device.index != torch.cuda.current_device()):
                raise ValueError("v1_device must be a contiguous float64 tensor of no*nu elements on the current GPU")
            _check(lib().ccx_energy(self._ctx, FORMS[form], ctypes.byref(ets), ctypes.byref(etd),
                                    ctypes.c_void_p(v1_device.data_ptr()), 1), "ccx_energy")
            v1 = v1_device
        else:
            v1 = np.zeros((self.nu, self.no), order="F")
            _check(lib().ccx_energy(self._ctx, FORMS[form], ctypes.byref(ets), ctypes.byref(etd),
                                    ctypes.c_void_p(v1.ctypes.data), 0), "ccx_energy")
        return {"ets": ets.value, "etd": etd.value, "e": ets.value + etd.value, "v1": v1}

    def energy_tuples(self, tuples, form: str = "fp64") -> dict:
        """ccx_energy over the given 1-based (i, j, k) tuples (i >= j >= k, not all equal): the
        partial E[T], E_ST and v1 the reference's driver accumulates over exactly those tuples."""
        arr = np.ascontiguousarray(np.asarray(tuples, dtype=np.int64).reshape(-1, 3))
        ets, etd = ctypes.c_double(), ctypes.c_double()
        v1 = np.zeros((self.nu, self.no), order="F")
        _check(lib().ccx_energy_tuples(self._ctx, FORMS[form], arr.shape[0],
                                       arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(ets),
                                       ctypes.byref(etd), ctypes.c_void_p(v1.ctypes.data), 0), "ccx_energy_tuples")
        return {"ets": ets.value, "etd": etd.value, "e": ets.value + etd.value, "v1": v1}

    def gradient(self, device: bool = False, form: str = "fp64") -> dict:
        """E(T) and dE/d(t1, t2, vm, voe, vvvo) in the reference layouts (ccx_gradient_form; form
        "fp64", "fp32" or "bf16x3" for the W recomputation, everything after it FP64).

        With ``device=True`` the gradients are torch float64 tensors on the current GPU,
        written in place by the library; otherwise Fortran-ordered numpy arrays."""
        no, nu = self.no, self.nu
        shapes = {"t1": (no, nu), "t2": (nu, nu, no, no), "vm": (no, nu, no, no),
                  "voe": (nu, nu, no, no), "vvvo": (nu, nu, nu, no)}
        if device:
            import torch
            # Fortran order = reversed-shape contiguous tensor, transposed back
            out = {k: torch.empty(tuple(reversed(v)), dtype=torch.float64, device="cuda").permute(
                *reversed(range(len(v)))) for k, v in shapes.items()}
            ptrs = {k: ctypes.c_void_p(t.data_ptr()) for k, t in out.items()}
        else:
            out = {k: np.zeros(v, order="F") for k, v in shapes.items()}
            ptrs = {k: ctypes.c_void_p(a.ctypes.data) for k, a in out.items()}
        ets, etd = ctypes.c_double(), ctypes.c_double()
        _check(lib().ccx_gradient_form(self._ctx, FORMS[form], ctypes.byref(ets), ctypes.byref(etd), ptrs["t1"],
                                       ptrs["t2"], ptrs["vm"], ptrs["voe"], ptrs["vvvo"], 1 if device else 0),
               "ccx_gradient_form")
        out.update({"ets": ets.value, "etd": etd.value, "e": ets.value + etd.value})
        return out

    def load_restart(self, path: str) -> None:
        """Load a reference ``triples.restart`` file whose no, nu match (ccx_load_restart)."""
        _check(lib().ccx_load_restart(self._ctx, str(path).encode()), "ccx_load_restart")

    @classmethod
    def from_restart(cls, path: str) -> "TriplesContext":
        no, nu = restart_dims(path)
        ctx = cls(no, nu)
        ctx.load_restart(path)
        return ctx

    def tuple_w(self, i: int, j: int, k: int) -> np.ndarray:
        """W_ijk cube for a 1-based tuple (no W(a,a,a)=0 step)."""
        w = np.zeros((self.nu, self.nu, self.nu), order="F")
        _check(lib().ccx_tuple_w(self._ctx, i, j, k, w.ctypes.data_as(_dp)), "ccx_tuple_w")
        return w

    def w_abc(self, a: int, b: int, c: int) -> np.ndarray:
        """Occupied cube v3(i,j,k) = W_ijk(a,b,c) of a 1-based virtual triple, v3(i,i,i) = 0
        (the cube of the triples-abc prototype, src/triples-abc/abc_tuple.f90:1-81)."""
        v3 = np.zeros((self.no, self.no, self.no), order="F")
        _check(lib().ccx_w_abc(self._ctx, a, b, c, v3.ctypes.data_as(_dp)), "ccx_w_abc")
        return v3

    def energy_abc(self) -> float:
        """E[T] re-summed over virtual triples (ccx_energy_abc); a cross-check route."""
        etd = ctypes.c_double()
        _check(lib().ccx_energy_abc(self._ctx, ctypes.byref(etd)), "ccx_energy_abc")
        return etd.value

    def timing(self) -> tuple:
        k, t = ctypes.c_double(), ctypes.c_double()
        _check(lib().ccx_last_timing(self._ctx, ctypes.byref(k), ctypes.byref(t)), "ccx_last_timing")
        return k.value, t.value

    def work_stats(self) -> tuple:
        n, f = ctypes.c_int64(), ctypes.c_double()
        _check(lib().ccx_work_stats(self._ctx, ctypes.byref(n), ctypes.byref(f)), "ccx_work_stats")
        return n.value, f.value


def cc_triples(eh, ep, t1, t2, vm, voe, vvvo, form: str = "fp64", packed_vvvo: bool = False) -> tuple:
    """(ets, etd, v1) for the given reference-layout arrays; see module docstring."""
    no, nu = len(eh), len(ep)
    ctx = TriplesContext(no, nu)
    try:
        ctx.load({"eh": eh, "ep": ep, "t1": t1, "t2": t2, "vm": vm, "voe": voe, "vvvo": vvvo},
                 packed_vvvo=packed_vvvo)
        r = ctx.energy(form)
    finally:
        ctx.close()
    return r["ets"], r["etd"], r["v1"]


def restart_dims(path: str) -> tuple:
    """(no, nu) from the header records of a ``triples.restart`` file (ccx_restart_dims)."""
    no, nu = ctypes.c_int64(), ctypes.c_int64()
    _check(lib().ccx_restart_dims(str(path).encode(), ctypes.byref(no), ctypes.byref(nu)), "ccx_restart_dims")
    return no.value, nu.value


def read_restart(path: str) -> dict:
    """Every array of a ``triples.restart`` file in the reference layouts (Fortran-ordered
    numpy arrays); ``vvvo`` is the packed (nutr, no*nu) array the file holds."""
    no, nu = restart_dims(path)
    nutr = nu * (nu + 1) // 2
    shapes = {"eh": (no,), "ep": (nu,), "t1": (no, nu), "t2": (nu, nu, no, no), "vm": (no, nu, no, no),
              "voe": (nu, nu, no, no), "vvvo": (nutr, no * nu)}
    out = {k: np.zeros(v, order="F") for k, v in shapes.items()}
    _check(lib().ccx_read_restart(str(path).encode(), no, nu,
                                  *[out[k].ctypes.data_as(_dp) for k in ("eh", "ep", "t1", "t2", "vm", "voe", "vvvo")]),
           "ccx_read_restart")
    out["no"], out["nu"] = no, nu
    return out


def write_restart(path: str, inp: dict, packed_vvvo: bool = False) -> None:
    """Write the arrays (reference layouts) as a ``triples.restart`` file (ccx_write_restart)."""
    eh = _host(inp["eh"])
    no, nu = eh.shape[0], _host(inp["ep"]).shape[0]
    arrs = [_host(inp[k]) for k in ("eh", "ep", "t1", "t2", "vm", "voe", "vvvo")]
    _check(lib().ccx_write_restart(str(path).encode(), no, nu, *[a.ctypes.data_as(_dp) for a in arrs],
                                   1 if packed_vvvo else 0), "ccx_write_restart")


def algorithmic_flops(no: int, nu: int) -> float:
    """GEMM FLOPs of the reference decomposition: 12 nu^3 (no+nu) per distinct tuple and
    6 nu^3 (no+nu) per degenerate one (src/triples-ijk/ijk_tuple.f90:19-28) = 2 (no^3-no) nu^3 (no+nu)."""
    return 2.0 * (no ** 3 - no) * nu ** 3 * (no + nu)
