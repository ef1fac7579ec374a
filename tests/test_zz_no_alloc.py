"""GPU: jump_energies_noalloc, the allocation-free variant of jump_energies,
called through the C ABI on particles held in device memory (torch tensors),
equals jump_energies of a context holding the same particles bit for bit,
on calls in a row with shrinking inputs; sizes past its fixed partial-sum
array are refused before any launch."""
import ctypes

import numpy as np
import pytest

from jump_amd import _ffi, scenes
from jump_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _device_state(eng, nsolids):
    import torch
    rows = [eng.get_solid(s) for s in range(nsolids)]
    cat = {k: np.ascontiguousarray(np.concatenate([r[k] for r in rows])) for k in
           ("mass", "vel", "volume", "strain", "stress")}
    return {k: torch.from_numpy(v).to("cuda") for k, v in cat.items()}, cat["mass"].shape[0]


def _noalloc(lib, dim, n, t, out, stream=None):
    return lib.jump_energies_noalloc(dim, n, t["mass"].data_ptr(), t["vel"].data_ptr(), t["volume"].data_ptr(),
                                     t["strain"].data_ptr(), t["stress"].data_ptr(), out.data_ptr(), stream)


def test_noalloc_energies_equal_the_context_entry_point(hip_lib):
    import torch
    cases = [scenes.block3d(nodes=(21, 21, 21), material="j2", steps=5),
             scenes.block3d(nodes=(11, 11, 11), material="elastic", steps=3),  # smaller
             scenes.rings2d(cells=40, basis=_ffi.BASIS_LINEAR, material="j2", steps=4),
             scenes.block2d(cells=8, steps=2)]
    sizes = []
    for i, sc in enumerate(cases):
        print(f"noalloc case {i}: {sc.name} {sc.n_particles} particles", flush=True)
        e = Engine(sc.params)
        sc.load(e)
        e.step(sc.steps, 0.0)
        want = e.energies()
        t, n = _device_state(e, len(sc.solids))
        sizes.append(n)
        out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
        assert _noalloc(hip_lib, sc.params.dim, n, t, out) == 0
        torch.cuda.synchronize()
        got = tuple(float(x) for x in out.cpu())
        assert got == want, (i, got, want)
        # a second call at once, on a stream of the caller's, with a prefix
        out2 = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
        k = max(1, n // 3)
        s = torch.cuda.Stream()
        sub = {key: v[: k * (v.numel() // n)] for key, v in t.items()}
        torch.cuda.synchronize()
        assert _noalloc(hip_lib, sc.params.dim, k, sub, out2, ctypes.c_void_p(s.cuda_stream)) == 0
        s.synchronize()
        sub_eng = Engine(sc.params)
        rows = [e.get_solid(q) for q in range(len(sc.solids))]
        cat = {key: np.concatenate([r[key] for r in rows])[:k] for key in rows[0]}
        sub_eng.add_solid(sc.solids[0]["material"], cat["pos"], cat["vel"], cat["mass"], cat["volume"],
                          F=cat["F"], strain=cat["strain"], stress=cat["stress"])
        assert tuple(float(x) for x in out2.cpu()) == sub_eng.energies(), i
        sub_eng.close()
        e.close()
    assert sizes[1] < sizes[0] and sizes[3] < sizes[2]
    # n = 0 writes zeros
    out = torch.full((2,), 7.0, dtype=torch.float64, device="cuda")
    assert _noalloc(hip_lib, 3, 0, t, out) == 0
    torch.cuda.synchronize()
    assert out.cpu().tolist() == [0.0, 0.0]


def test_noalloc_refuses_sizes_past_its_partial_sum_array(hip_lib):
    import torch
    t = {k: torch.zeros(9, dtype=torch.float64, device="cuda") for k in ("mass", "vel", "volume", "strain", "stress")}
    out = torch.full((2,), 5.0, dtype=torch.float64, device="cuda")
    assert _noalloc(hip_lib, 3, (1 << 24) + 1, t, out) == _ffi.JUMP_ERR_UNSUPPORTED
    assert _noalloc(hip_lib, 4, 1, t, out) == _ffi.JUMP_ERR_ARG
    assert hip_lib.jump_energies_noalloc(3, 1, None, None, None, None, None, out.data_ptr(), None) == \
        _ffi.JUMP_ERR_ARG
    torch.cuda.synchronize()
    assert out.cpu().tolist() == [5.0, 5.0]  # refused calls launch nothing
