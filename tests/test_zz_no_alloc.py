"""The allocation-free filter entry point (arecon_filter_dev_noalloc) against
arecon_filter_dev, through the C ABI.

Both run the same kernels; the no-alloc one builds its twiddle, spectrum and
preweight tables in a fixed per-device array instead of the allocating caches.
Bar: the filter stage's (1e-5 relative L2 of the FP64 oracle for FP32 output),
and agreement with arecon_filter_dev to FP32 rounding of the tables (1e-6
relative L2; in practice bit-identical).  Two calls in a row, the second with a
smaller detector, so the second must not read the first call's tables.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from analyticalreconbigpatient_amd import _lib

pytestmark = pytest.mark.gpu

SO = 541.0


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _out_shape(ViewN, ZL, YL, layout):
    if layout == _lib.ARECON_FP_ROWS:
        return (ViewN, ZL, YL)
    if layout == _lib.ARECON_FP_QUADS:
        return (ViewN, ZL - 1, YL - 1, 4)
    return (ViewN, ZL - 1, YL, 2)


def _run(fn_name, P, dtype, layout):
    ViewN, ZL, YL = P.shape
    tdt = {_lib.ARECON_F32: torch.float32, _lib.ARECON_F16: torch.float16,
           _lib.ARECON_BF16: torch.bfloat16}[dtype]
    din = torch.from_numpy(P).cuda()
    dout = torch.empty(_out_shape(ViewN, ZL, YL, layout), dtype=tdt, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    fn = getattr(_lib.lib(), fn_name)
    _lib.check(fn(C.c_void_p(dout.data_ptr()), dtype, layout, C.c_void_p(din.data_ptr()), YL,
                  ZL, ViewN, np.float32((YL - 1) / 2.0), np.float32((ZL - 1) / 2.0),
                  np.float32(0.9493), np.float32(1.9231), np.float32(SO), stream), fn_name)
    torch.cuda.synchronize()
    return dout.float().cpu().numpy()


# first call at the clinical detector, the second smaller, then a 2048-channel
# detector (4096-point transforms: the in-place rows path) and back to small
CALLS = [((9, 64, 888), _lib.ARECON_F32, _lib.ARECON_FP_ROWDIFF),
         ((5, 16, 368), _lib.ARECON_F32, _lib.ARECON_FP_ROWDIFF),
         ((4, 3, 2048), _lib.ARECON_F32, _lib.ARECON_FP_ROWS),
         ((6, 8, 200), _lib.ARECON_F16, _lib.ARECON_FP_ROWDIFF),
         ((5, 6, 130), _lib.ARECON_F32, _lib.ARECON_FP_QUADS)]


def test_noalloc_filter_matches_allocating_entry_point(gpu):
    for i, ((ViewN, ZL, YL), dtype, layout) in enumerate(CALLS):
        print(f"no-alloc filter call {i}: ViewN={ViewN} ZL={ZL} YL={YL} dtype={dtype} "
              f"layout={layout}", flush=True)
        P = np.random.default_rng(100 + i).normal(size=(ViewN, ZL, YL)).astype(np.float32)
        want = _run("arecon_filter_dev", P, dtype, layout)
        got = _run("arecon_filter_dev_noalloc", P, dtype, layout)
        assert rel_l2(got, want) <= 1e-6, (i, rel_l2(got, want))
        if layout == _lib.ARECON_FP_ROWS and dtype == _lib.ARECON_F32:
            ref = oracle.filter_f64(P.astype(np.float64), float(np.float32((YL - 1) / 2.0)),
                                    float(np.float32((ZL - 1) / 2.0)),
                                    float(np.float32(0.9493)), float(np.float32(1.9231)), SO)
            assert rel_l2(got, ref) < 1e-5
        if layout == _lib.ARECON_FP_ROWDIFF and dtype == _lib.ARECON_F32:
            rows = oracle.filter_f64(P.astype(np.float64), float(np.float32((YL - 1) / 2.0)),
                                     float(np.float32((ZL - 1) / 2.0)),
                                     float(np.float32(0.9493)), float(np.float32(1.9231)), SO)
            assert rel_l2(got[..., 0], rows[:, :-1, :]) < 1e-5


def test_noalloc_filter_rejects_shapes_beyond_its_scratch(gpu):
    """ZL * YL above 249,856 is refused on the host before any launch."""
    lib = _lib.lib()
    null = C.c_void_p(0)
    rc = lib.arecon_filter_dev_noalloc(null, _lib.ARECON_F32, _lib.ARECON_FP_ROWDIFF, null,
                                       1024, 256, 4, np.float32(511.5), np.float32(127.5),
                                       np.float32(0.9), np.float32(1.9), np.float32(SO),
                                       C.c_void_p(torch.cuda.current_stream().cuda_stream))
    code, msg = _lib.last_error()
    assert rc == 1 and code == 1
    assert "249,856" in msg
