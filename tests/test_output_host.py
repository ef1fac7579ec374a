"""Host build of the library's output code (jump_amd/csrc/mpm_output.hpp):
Julia's Float64 text (known answers of its print rule, agreement with the
independent Python formatter, exact round trip) and the LAMMPS dump against
the format restated from Output.jl (tests/lammps_ref.py)."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from jump_amd.fix import julia_float
from lammps_ref import dump_text

HERE = os.path.dirname(os.path.abspath(__file__))
pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")

# print(x::Float64) in Julia (its Ryu shortest form; plain for -4 < pt <= 6)
KNOWN = [(1.0, "1.0"), (0.1, "0.1"), (1e6, "1.0e6"), (1e5, "100000.0"), (123456.0, "123456.0"),
         (1234567.0, "1.234567e6"), (0.0001, "0.0001"), (1e-5, "1.0e-5"), (1.5e-5, "1.5e-5"),
         (12345.678, "12345.678"), (-0.0, "-0.0"), (0.0, "0.0"), (-2.5, "-2.5"), (5e-324, "5.0e-324"),
         (1.7976931348623157e308, "1.7976931348623157e308"), (1e16, "1.0e16"), (0.001, "0.001"),
         (float("nan"), "NaN"), (float("inf"), "Inf"), (float("-inf"), "-Inf"), (1010e-9, "1.01e-6"),
         (2.0e-3, "0.002"), (999999.9, "999999.9")]


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("out") / "output_check")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(HERE, "host", "output_check.cpp"), "-o", path],
                   check=True)
    return path


def _cpp_format(exe, xs):
    inp = "".join(f"{struct.unpack('<Q', struct.pack('<d', float(x)))[0]:016x}\n" for x in xs)
    out = subprocess.run([exe, "num"], input=inp, capture_output=True, text=True, check=True)
    return out.stdout.splitlines()


def test_julia_float_known_answers(exe):
    xs = [x for x, _ in KNOWN]
    want = [t for _, t in KNOWN]
    assert _cpp_format(exe, xs) == want
    assert [julia_float(x) for x in xs] == want


def test_julia_float_cpp_matches_python_and_round_trips(exe):
    rng = np.random.default_rng(20261202)
    bits = rng.integers(0, 2 ** 64, size=20000, dtype=np.uint64)
    xs = list(bits.view(np.float64))
    # decade boundaries and their neighbours, where the notation switches
    for k in range(-12, 20):
        b = 10.0 ** k
        xs += [b, np.nextafter(b, 0.0), np.nextafter(b, np.inf), -b]
    got = _cpp_format(exe, xs)
    for x, t in zip(xs, got):
        assert t == julia_float(x), (x, t, julia_float(x))
        if np.isfinite(x):
            y = float(t.replace("Inf", "inf"))
            assert y == x and np.signbit(y) == np.signbit(x), (x, t)


def _write_state(path, dim, lims, solids, outs):
    S = 6 if dim == 3 else 3
    with open(path, "wb") as f:
        n = sum(len(s["pos"]) for s in solids)
        f.write(struct.pack("<iqi", dim, n, len(solids)))
        end = 0
        for s in solids:
            end += len(s["pos"])
            f.write(struct.pack("<qB", end, 1 if s["elastic"] else 0))
        f.write(struct.pack("<3d", *lims))
        ob = " ".join(outs).encode()
        f.write(struct.pack("<i", len(ob)) + ob)
        cat = lambda k: np.concatenate([s[k] for s in solids]).astype(np.float64)
        f.write(cat("pos").tobytes())
        f.write(cat("vel").tobytes())
        sig = np.concatenate([s["stress"] for s in solids])
        idx = [(0, 0), (1, 1), (0, 1)] if dim == 2 else [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
        f.write(np.stack([sig[:, a, b] for a, b in idx], axis=1).astype(np.float64).tobytes())
        f.write(cat("alpha").tobytes())
        f.write(cat("vm").tobytes())
    assert S == len(idx)


def _solids(rng, dim, counts, elastic):
    out = []
    for n, el in zip(counts, elastic):
        sig = rng.normal(size=(n, dim, dim)) * 10.0 ** rng.integers(-8, 9, size=(n, 1, 1))
        sig = 0.5 * (sig + sig.transpose(0, 2, 1))
        out.append(dict(pos=rng.uniform(0, 2, size=(n, dim)), vel=rng.normal(size=(n, dim)) * 1e-3,
                        stress=sig, alpha=np.abs(rng.normal(size=n)) * 1e-4, vm=rng.uniform(0, 1e8, size=n),
                        elastic=el))
    return out


@pytest.mark.parametrize("dim,outs", [(2, ["pstrain", "vonMises", "sigmaxx", "sigmayy", "pressure", "vx", "vy"]),
                                      (2, []), (3, ["pstrain", "vonMises", "vx", "vy", "vz", "pressure"])])
def test_lammps_dump_matches_reference_format(exe, tmp_path, dim, outs):
    rng = np.random.default_rng(7 + dim + len(outs))
    solids = _solids(rng, dim, [5, 7, 3], [True, False, True])
    lims = (200.0, 100.0, 50.0 if dim == 3 else 0.0)
    state = str(tmp_path / "state.bin")
    _write_state(state, dim, lims, solids, outs)
    dump = str(tmp_path / "dump_p.0.LAMMPS")
    subprocess.run([exe, "dump", state, dump], check=True)
    subprocess.run([exe, "dump", state, dump], check=True)  # the reference appends
    want = dump_text(dim, lims, solids, outs)
    assert open(dump).read() == want + want
