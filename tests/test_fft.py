"""The register FFT of the fast spectral transforms (kitbase.jl_amd/csrc/kbg_fft.h: mixed-radix
decimation in time with radices 4, 3, 2, unrolled at compile time) built for the host and compared
with numpy's FFT (FFTW's forward sign) on every length the library instantiates and their radix
factors, in FP32 and FP64.  The GPU kernels k_fsm_pair / k_fsm_slab64 run the same code."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def fft_bin(tmp_path_factory):
    if not Path(HIPCC).exists():
        pytest.skip("hipcc is not available")
    out = tmp_path_factory.mktemp("fft") / "fft_check"
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", str(HERE / "native" / "fft_check.cpp"), "-o", str(out)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return out


def transform(binary, x, prec):
    inp = f"{x.size} {prec}\n" + "\n".join(f"{float(v.real)!r} {float(v.imag)!r}" for v in x) + "\n"
    r = subprocess.run([str(binary)], input=inp, capture_output=True, text=True, check=True)
    v = np.array([[float(t) for t in line.split()] for line in r.stdout.strip().splitlines()])
    return v[:, 0] + 1j * v[:, 1]


@pytest.mark.parametrize("n", [2, 3, 4, 6, 8, 12, 16, 18, 24])
@pytest.mark.parametrize("prec,tol", [("f32", 4e-7), ("f64", 2e-15)])
def test_register_fft_matches_numpy(fft_bin, n, prec, tol):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    if prec == "f32":
        x = x.astype(np.complex64).astype(np.complex128)  # the inputs the FP32 transform sees
    ref = np.fft.fft(x)
    got = transform(fft_bin, x, prec)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err <= tol * np.log2(max(n, 2)) * 2, (n, prec, err)


def test_register_fft_of_a_delta_and_a_mode(fft_bin):
    """Closed forms: the transform of a unit impulse at j0 is exp(-2 pi i j0 k / n); of the mode
    exp(2 pi i m j / n) it is n at k = m and 0 elsewhere."""
    n = 24
    for j0 in (0, 1, 7):
        x = np.zeros(n, complex)
        x[j0] = 1.0
        got = transform(fft_bin, x, "f64")
        np.testing.assert_allclose(got, np.exp(-2j * np.pi * j0 * np.arange(n) / n), rtol=0, atol=1e-14)
    m = 5
    x = np.exp(2j * np.pi * m * np.arange(n) / n)
    got = transform(fft_bin, x, "f64")
    ref = np.zeros(n, complex)
    ref[m] = n
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)
