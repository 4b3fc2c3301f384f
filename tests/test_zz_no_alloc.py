"""me_ncp_sensitivity_noalloc (the entry point that allocates nothing) against me_ncp_sensitivity,
through the C ABI on the GPU: two calls in a row, the second with a smaller input, both with systems
beyond the LDS limit so that the fixed scratch array carries the matrices. The outputs must be the
same bytes (same kernel, same arithmetic); a call beyond the array's size is refused before any launch."""
import ctypes as C
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _run(fn_name, G, b, mu, lam, with_workspace):
    import torch
    L = importlib.import_module("3fold-motion-engine_amd._lib")
    lib = L.lib()
    dev = torch.device("cuda:0")
    E, n3 = b.shape
    n = n3 // 3
    t = lambda x: torch.tensor(np.asarray(x), dtype=torch.float64, device=dev).contiguous()
    Gd, bd, md, ld = t(G), t(b), t(mu), t(lam)
    dmu = torch.full((E, n3, n), -1.0, dtype=torch.float64, device=dev)
    dlb = torch.full((E, n3, n3), -1.0, dtype=torch.float64, device=dev)
    lab = torch.full((E, n), -1, dtype=torch.int32, device=dev)
    st = torch.full((E,), -1, dtype=torch.int32, device=dev)
    stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    p = lambda x: C.c_void_p(x.data_ptr())
    head = (E, n, n3, p(Gd), p(bd), p(md), p(ld), None, None, 1e-10, 1e-6, 1e-12, p(dmu), p(dlb), p(lab), p(st))
    if with_workspace:
        nb = C.c_size_t()
        L.check(lib.me_ncp_workspace_size(E, n, n3, C.byref(nb)), "me_ncp_workspace_size")
        assert nb.value > 0, "the case must need scratch beyond LDS"
        ws = torch.empty(int(nb.value), dtype=torch.uint8, device=dev)
        L.check(lib.me_ncp_sensitivity(*head, p(ws), int(nb.value), stream), "me_ncp_sensitivity")
    else:
        L.check(lib.me_ncp_sensitivity_noalloc(*head, stream), "me_ncp_sensitivity_noalloc")
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in (dmu, dlb, lab, st)]


def test_noalloc_matches_workspace_entry_point_twice():
    ncp = importlib.import_module("3fold-motion-engine_amd.ncp")
    cases = [(6, 40, 160, 9), (3, 30, 120, 11)]  # second call smaller
    for E, n_c, nv, seed in cases:
        _say(f"ncp noalloc vs workspace: {E} envs x {n_c} contacts")
        G, b, mu, lam, _ = ncp.synthetic_problems(E, n_c=n_c, nv=nv, seed=seed)
        ref = _run("ws", G, b, mu, lam, True)
        got = _run("noalloc", G, b, mu, lam, False)
        for r, g in zip(ref, got):
            assert r.tobytes() == g.tobytes()
        assert np.all(got[3] >= 0) and np.isfinite(got[0]).any()


def test_noalloc_refuses_beyond_its_array():
    import torch
    L = importlib.import_module("3fold-motion-engine_amd._lib")
    lib = L.lib()
    _say("ncp noalloc capacity refusal (host check, no launch)")
    E, n = 200, 40  # 200 x 268 800 B of matrices > 16 MiB
    nb = C.c_size_t()
    L.check(lib.me_ncp_workspace_size(E, n, 3 * n, C.byref(nb)), "me_ncp_workspace_size")
    assert nb.value > (16 << 20)
    x = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    p = C.c_void_p(x.data_ptr())
    rc = lib.me_ncp_sensitivity_noalloc(E, n, 3 * n, p, p, p, p, None, None, 1e-10, 1e-6, 1e-12, p, p, p, p, None)
    assert rc == L.ME_ECAPACITY
    assert b"fixed scratch" in lib.me_last_error()
