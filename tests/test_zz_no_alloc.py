"""ns_totals_noalloc (the allocation-free form of ns_totals) through the C-ABI: on fields held in device
memory by the caller it gives ns_totals' sums bit for bit, on two calls in a row (the second on a smaller
grid, so the shared scratch holds the first call's partials when it starts), in FP64 and FP32."""
import ctypes as C

import numpy as np
import pytest
import torch

import cases
import nsfsi

pytestmark = pytest.mark.gpu


def _entry():
    L = nsfsi.lib()
    f = L.ns_totals_noalloc
    f.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double),
                  C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    L.ns_noalloc_last_error.restype = C.c_char_p
    return f, L


def _case_on_device(N, precision, steps=6):
    cfg = cases.sphere48(N=N, steps=steps, precision=precision)
    print(f"no-alloc totals: sphere {N}^3 {precision}", flush=True)
    s = nsfsi.Solver(cfg)
    s.initialize()
    s.marchTimeStep(steps)
    ref = s.mesh.totals()
    U = s.mesh.conserved()
    types = s.mesh.nodeType
    sp = np.array([s.mesh.dx, s.mesh.dy, s.mesh.dz])
    s.close()
    dt = np.float64 if precision == "fp64" else np.float32
    fields = torch.from_numpy(np.ascontiguousarray(U.astype(dt))).cuda()
    dtypes = torch.from_numpy(np.ascontiguousarray(types)).cuda()
    return cfg, ref, fields, dtypes, sp


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_totals_equal_handle_totals(precision):
    f, L = _entry()
    prec = 0 if precision == "fp64" else 1
    stream = torch.cuda.current_stream()
    for N in (32, 20):  # the second call is smaller
        cfg, ref, fields, types, sp = _case_on_device(N, precision)
        out = torch.full((5,), float("nan"), dtype=torch.float64, device="cuda")
        st = f(fields.data_ptr(), prec, types.data_ptr(), cfg.NI, cfg.NJ, cfg.NK,
               sp.ctypes.data_as(C.POINTER(C.c_double)), out.data_ptr(), stream.cuda_stream)
        assert st == 0, L.ns_noalloc_last_error()
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        print(f"  ns_totals {ref} no-alloc {got}")
        assert np.array_equal(got, ref)


def test_noalloc_refuses_oversized_grid():
    f, L = _entry()
    buf = torch.zeros(8, dtype=torch.float64, device="cuda")
    types = torch.zeros(8, dtype=torch.uint8, device="cuda")
    sp = np.ones(3)
    st = f(buf.data_ptr(), 0, types.data_ptr(), 1 << 13, 1 << 12, 1 << 12, sp.ctypes.data_as(C.POINTER(C.c_double)),
           buf.data_ptr(), None)
    assert st == 5
    assert b"exceeds the maximum" in L.ns_noalloc_last_error()
    assert st == 5 and f(0, 0, types.data_ptr(), 2, 2, 2, sp.ctypes.data_as(C.POINTER(C.c_double)), buf.data_ptr(),
                         None) == 5
