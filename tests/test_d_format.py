"""The Fortran host's D23.15 record formatter (ixchel2d_amd/fortran/
d_format.c, used by postproceso_dat for the out_*.dat files of upstream's
mod_postproceso.f90) against flang's own formatted write of the same
records: byte for byte, over values of every binary exponent, subnormals,
signed zeros, powers of ten and two with their neighbours, and exact
rounding ties at the 16th significant digit."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")


def _values() -> np.ndarray:
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2**64 - 1, size=200_000, dtype=np.uint64, endpoint=True)
    anyexp = bits.view(np.float64)
    anyexp = anyexp[np.isfinite(anyexp)]
    typical = rng.uniform(-1.0, 1.0, 200_000) * 10.0 ** rng.uniform(-9.0, 9.0, 200_000)
    edges = [0.0, -0.0, 5e-324, -5e-324, 2.2250738585072014e-308, 2.2250738585072009e-308,
             1.7976931348623157e308, -1.7976931348623157e308, 1.0, -1.0, 0.5, 0.1, 1e-100, 9.9999999999999995e99]
    for k in range(-323, 309):
        p = float(f"1e{k}")
        edges += [p, np.nextafter(p, 0.0), np.nextafter(p, np.inf), -p]
    for k in range(-1074, 1024):
        p = float(np.ldexp(1.0, k))
        edges += [p, np.nextafter(p, 0.0), np.nextafter(p, np.inf)]
    # exact ties: k / 2^(15+t) with 16 significant digits, the last a 5
    ties = []
    for t in range(-10, 30):
        m = 15 + t
        lo, hi = 10.0 ** (-t), 10.0 ** (1 - t)
        k0, k1 = int(np.ceil(lo * 2.0**m)), int(np.floor(hi * 2.0**m))
        if k1 <= k0:
            continue
        ks = rng.integers(k0, k1, size=400) | 1
        ties += list(np.ldexp(ks.astype(np.float64), -m))
    ties = np.array(ties)
    ties = np.concatenate([ties, -ties])
    v = np.concatenate([anyexp, typical, np.array(edges), ties])
    return v[np.isfinite(v)]


@pytest.mark.skipif(not (os.path.exists(FLANG) and shutil.which("gcc")), reason="needs flang and gcc")
def test_d23_records_equal_flang_formatted_write():
    v = _values()
    v = v[: (len(v) // 12) * 12]
    with tempfile.TemporaryDirectory(prefix="ixh_dfmt_") as d:
        src = os.path.join(ROOT, "ixchel2d_amd", "fortran", "d_format.c")
        subprocess.run(["gcc", "-O2", "-Wall", "-pthread", "-c", src, "-o", "d_format.o"], cwd=d, check=True)
        subprocess.run([FLANG, "-O2", os.path.join(ROOT, "tests", "dfmt_driver.f90"), "d_format.o", "-lm", "-lpthread",
                        "-o", "dfmt_driver"], cwd=d, check=True)
        with open(os.path.join(d, "valores.raw"), "wb") as fh:
            fh.write(np.int32(len(v)).tobytes())
            fh.write(v.astype("<f8").tobytes())
        subprocess.run([os.path.join(d, "dfmt_driver")], cwd=d, check=True, timeout=300)
        for ncol in (3, 4):
            f = open(os.path.join(d, f"f{ncol}.txt"), "rb").read()
            c = open(os.path.join(d, f"c{ncol}.txt"), "rb").read()
            if f != c:
                fl, cl = f.splitlines(), c.splitlines()
                bad = [(i, a, b) for i, (a, b) in enumerate(zip(fl, cl)) if a != b][:5]
                pytest.fail(f"{ncol} columns: {len(fl)} vs {len(cl)} lines; first differences {bad}")
