"""dts_solve_noalloc: the solve on device buffers with every intermediate plane
in the library's static scratch array, called through the C ABI and checked
bit for bit against the allocating path (dts_init_frame_device + dts_solve +
download) on the same inputs, on calls in a row with a smaller frame second;
device memory stays unchanged across the calls."""
import ctypes

import numpy as np
import pytest

from domain_transform_solver_amd import _lib, scenes

pytestmark = pytest.mark.gpu


def _opts(p, loss=None, overwrite=-1.0, conf=(8.0, 8.0, 0.2, 16.0)):
    fp = _lib.FilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"])
    op = _lib.OptimizeParams(p["num_iter"], p["loss"] if loss is None else loss, p["lam"],
                             p["sigma_z"], p["step"], 0.02, 0.2)
    cp = _lib.FilterParams(conf[0], conf[1], conf[2])
    return _lib.SolveOptions(fp, op, cp, conf[3], overwrite)


def _case(torch, kind, w, h, seed):
    if kind == "depth":
        sc = scenes.depth_scene(w, h, 8, seed=seed)
        p = scenes.depth_sr_params(8)
        rgba, t, c = sc.rgba, sc.target, sc.confidence
        o = _opts(p)
        cs = p["color_space"]
    else:
        sc = scenes.colorization_scene(w, h, seed=seed, regions=8, scribble_frac=0.05)
        p = scenes.depth_sr_params(8)
        p.update(sigma_x=64.0, sigma_y=64.0, sigma_r=0.25, sigma_z=30.0, num_iter=20)
        rgba, t, c = sc.rgba, sc.target_u, sc.confidence
        o = _opts(p, loss=2, overwrite=0.85)
        cs = 0
    d = dict(rgba=torch.from_numpy(np.ascontiguousarray(rgba)).cuda(),
             t=torch.from_numpy(np.ascontiguousarray(t, np.float32)).cuda(),
             c=torch.from_numpy(np.ascontiguousarray(c, np.float32)).cuda(),
             out=torch.empty((h, w), dtype=torch.float32, device="cuda"))
    return d, o, cs


def _reference(lib, d, o, cs, w, h, with_conf, compute_conf):
    """The allocating path on the same device inputs."""
    hd = ctypes.c_void_p()
    _lib.check(lib.dts_create(w, h, 0, ctypes.byref(hd)))
    try:
        _lib.check(lib.dts_init_frame_device(hd, cs, ctypes.c_void_p(d["rgba"].data_ptr()),
                                             ctypes.c_void_p(d["t"].data_ptr()),
                                             ctypes.c_void_p(d["c"].data_ptr()) if with_conf
                                             else None))
        _lib.check(lib.dts_solve(hd, ctypes.byref(o), compute_conf))
        out = np.empty((h, w), np.float32)
        _lib.check(lib.dts_download(hd, 5, out.ctypes.data_as(ctypes.c_void_p)))
    finally:
        lib.dts_destroy(hd)
    return out


def _noalloc(lib, torch, d, o, cs, w, h, with_conf, compute_conf):
    print(f"dts_solve_noalloc {w}x{h} conf={with_conf} compute_conf={compute_conf}", flush=True)
    rc = lib.dts_solve_noalloc(w, h, cs, ctypes.c_void_p(d["rgba"].data_ptr()),
                               ctypes.c_void_p(d["t"].data_ptr()),
                               ctypes.c_void_p(d["c"].data_ptr()) if with_conf else None,
                               ctypes.byref(o), compute_conf,
                               ctypes.c_void_p(d["out"].data_ptr()),
                               ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, lib.dts_last_error()


def test_noalloc_matches_dts_solve_on_consecutive_calls(gpu_lib):
    import torch
    lib = gpu_lib
    torch.cuda.init()
    # (kind, w, h, seed, confidence given, compute confidence): the second
    # call is smaller than the first; the last one computes the confidence.
    plan = [("depth", 640, 480, 1001, True, 0), ("depth", 320, 240, 1002, True, 0),
            ("color", 512, 384, 7, True, 0), ("depth", 200, 150, 1003, False, 1),
            ("depth", 97, 61, 1004, True, 0)]
    cases = []
    for kind, w, h, seed, wc, cc in plan:
        d, o, cs = _case(torch, kind, w, h, seed)
        cases.append((d, o, cs, w, h, wc, cc))
    torch.cuda.synchronize()
    refs = [_reference(lib, d, o, cs, w, h, wc, cc) for d, o, cs, w, h, wc, cc in cases]
    torch.cuda.synchronize()
    # First call on this device: looks up the scratch array's address.
    d, o, cs, w, h, wc, cc = cases[0]
    _noalloc(lib, torch, d, o, cs, w, h, wc, cc)
    for d, o, cs, w, h, wc, cc in cases:
        d["out"].fill_(float("nan"))
    assert np.isnan(cases[0][0]["out"].cpu().numpy()).all()
    torch.cuda.synchronize()
    # Device memory in use is read only around the calls themselves (torch's
    # own kernels and copies load code objects and staging on first use).
    free0 = torch.cuda.mem_get_info()[0]
    for d, o, cs, w, h, wc, cc in cases:
        _noalloc(lib, torch, d, o, cs, w, h, wc, cc)
        torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    for (d, o, cs, w, h, wc, cc), ref in zip(cases, refs):
        got = d["out"].cpu().numpy()
        assert np.array_equal(got, ref), (w, h, float(np.abs(got - ref).max()))
    assert free1 == free0


def test_noalloc_rejects_what_it_cannot_hold(gpu_lib):
    import torch
    lib = gpu_lib
    p = scenes.depth_sr_params(8)
    o = _opts(p)
    dummy = torch.zeros(16, device="cuda")
    ptr = ctypes.c_void_p(dummy.data_ptr())
    # Over 2^20 pixels, and a side over 4096.
    assert lib.dts_solve_noalloc(1200, 1000, 1, ptr, ptr, ptr, ctypes.byref(o), 0, ptr,
                                 None) == 2
    assert b"exceeds" in lib.dts_last_error()
    assert lib.dts_solve_noalloc(5000, 10, 1, ptr, ptr, ptr, ctypes.byref(o), 0, ptr, None) == 2
    # Windows beyond the LDS segments take the allocating wide path.
    ow = _opts(dict(p, sigma_z=400.0))
    assert lib.dts_solve_noalloc(1024, 1000, 1, ptr, ptr, ptr, ctypes.byref(ow), 0, ptr,
                                 None) == 5
    assert lib.dts_solve_noalloc(64, 64, 1, None, ptr, ptr, ctypes.byref(o), 0, ptr, None) == 1
