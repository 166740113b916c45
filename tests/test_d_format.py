"""The Fortran host's D23.15 record formatter and reader (ixchel2d_amd/
fortran/d_format.c: postproceso_dat writes the out_*.dat files of upstream's
mod_postproceso.f90 with it, lectura_mallas_escalonadas reads the mesh and
restart files of mod_malla.f90 with it) against flang's own formatted write
and read of the same records: byte for byte and bit for bit, over values of
every binary exponent, subnormals, signed zeros, powers of ten and two with
their neighbours, exact rounding ties at the 16th significant digit, and
the other spellings of a D23.15 input field."""
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


needs = pytest.mark.skipif(not (os.path.exists(FLANG) and shutil.which("gcc")), reason="needs flang and gcc")


def _build(d):
    src = os.path.join(ROOT, "ixchel2d_amd", "fortran", "d_format.c")
    subprocess.run(["gcc", "-O2", "-Wall", "-pthread", "-c", src, "-o", "d_format.o"], cwd=d, check=True)
    subprocess.run([FLANG, "-O2", os.path.join(ROOT, "tests", "dfmt_driver.f90"), "d_format.o", "-lm", "-lpthread",
                    "-o", "dfmt_driver"], cwd=d, check=True)


def _read_both(d, name, nrec, ncol):
    r = subprocess.run([os.path.join(d, "dfmt_driver"), "leer", name, str(nrec), str(ncol)], cwd=d, check=True,
                       stdout=subprocess.PIPE, text=True, timeout=300)
    status = int(r.stdout.split()[-1])
    f = np.fromfile(os.path.join(d, "fleer.raw"), "<f8")
    c = np.fromfile(os.path.join(d, "cleer.raw"), "<f8")
    return status, f, c


@needs
def test_d23_records_equal_flang_formatted_write():
    v = _values()
    v = v[: (len(v) // 12) * 12]
    with tempfile.TemporaryDirectory(prefix="ixh_dfmt_") as d:
        _build(d)
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


@needs
def test_d23_read_equals_flang_formatted_read():
    """The records the formatter wrote, and fields in the other spellings
    Fortran's D23.15 input accepts (E and lower-case exponents, an exponent
    given by its sign alone, no exponent, no leading digit, more than 15
    digits, subnormal and signed values), read back bit for bit as flang
    reads them; fields the parser declines (no decimal point, where Fortran
    scales by 10^-15; blanks inside) make it return 1."""
    v = _values()
    v = v[np.abs(v) < 1.79769313486231e308]  # 15 digits of the largest doubles round past the range
    v = v[: (len(v) // 12) * 12]
    spellings = ["1.5", "-.25E+03", "3.0e-5", "1.0d0", "+2.5D-310", "7.", "0.12345678901234567891",
                 "1.0+5", "-4.2-3", "9.99999999999999999D+99", "2.2250738585072011d-308", "-0.0", "1.E2",
                 ".5", "123456789.123456789", "6.02214076D+23"]
    with tempfile.TemporaryDirectory(prefix="ixh_dfmt_") as d:
        _build(d)
        with open(os.path.join(d, "valores.raw"), "wb") as fh:
            fh.write(np.int32(len(v)).tobytes())
            fh.write(v.astype("<f8").tobytes())
        subprocess.run([os.path.join(d, "dfmt_driver")], cwd=d, check=True, timeout=300)
        for ncol in (3, 4):
            with open(os.path.join(d, f"c{ncol}.txt"), "rb") as fh:
                body = fh.read()
            with open(os.path.join(d, f"r{ncol}.txt"), "wb") as fh:
                fh.write(b" 2 16 20 1.\n" + body)
            status, f, c = _read_both(d, f"r{ncol}.txt", len(v) // ncol, ncol)
            assert status == 0
            assert np.array_equal(f.view(np.uint64), c.view(np.uint64))
        rng = np.random.default_rng(5)
        fields = [rng.choice(spellings) for _ in range(3 * 400)]
        lines = ["".join(f"{x:>23s}" for x in fields[3 * r:3 * r + 3]) for r in range(400)]
        with open(os.path.join(d, "s.txt"), "w") as fh:
            fh.write(" 1 2 3 4.\n" + "\n".join(lines) + "\n")
        status, f, c = _read_both(d, "s.txt", 400, 3)
        assert status == 0
        assert np.array_equal(f.view(np.uint64), c.view(np.uint64))
        for bad in ("15", "1 .5"):
            lines2 = list(lines)
            lines2[7] = f"{bad:>23s}" + lines2[7][23:]
            with open(os.path.join(d, "b.txt"), "w") as fh:
                fh.write(" 1 2 3 4.\n" + "\n".join(lines2) + "\n")
            status, f, c = _read_both(d, "b.txt", 400, 3)
            assert status == 1, bad
