"""The azimuthal FFT of the cardinal kernels, checked on the host: the C
ABI's axs3d_fft_check runs the library's own plan (pass radices, twiddle
table, digit reversal) pass by pass and lane by lane as the device waves
do, in FP64.  Every FFT lucky number Nr up to 2048 (prime factors <= 13,
at most one 11 or 13, NrField.cpp:108-133) with the pass radix limited to
16, 8 and 4, both directions, against numpy's FFT in the sign conventions
of SolverFFTW (SolverFFTW.hpp:101-168: exp(+2 pi i jk / Nr) from Fourier
orders to slices, exp(-2 pi i jk / Nr) back, unscaled here).  No GPU."""
import ctypes as C

import numpy as np
import pytest

from axisem3d_amd import capi, spectral


def _lucky(nmax):
    out = []
    for n in range(1, nmax + 1):
        m, big = n, 0
        for p in (2, 3, 5, 7, 11, 13):
            while m % p == 0:
                m //= p
                big += p > 7
        if m == 1 and big <= 1:
            out.append(n)
    return out


def _run(nr, max_radix, inverse, x):
    L = capi.load()
    d = np.ascontiguousarray(x, dtype=np.complex128).copy()
    rad = (C.c_int * 12)()
    npass = C.c_int()
    rc = L.axs3d_fft_check(nr, max_radix, int(inverse), d.ctypes.data, rad,
                           C.byref(npass))
    assert rc == 0, (nr, rc)
    return d, [rad[i] for i in range(npass.value)]


def test_lucky_numbers_match_the_library_table():
    luck = _lucky(2048)
    assert 2048 in luck and 2016 in luck and 2002 not in luck
    assert spectral.next_lucky_number(2001) == 2016
    for n in luck:
        assert spectral.next_lucky_number(n) == n


@pytest.mark.parametrize("max_radix", [16, 8, 4])
def test_every_lucky_nr_matches_numpy(max_radix):
    worst = 0.0
    for nr in _lucky(2048):
        rng = np.random.default_rng(nr)
        x = rng.standard_normal(nr) + 1j * rng.standard_normal(nr)
        for inverse in (True, False):
            got, rad = _run(nr, max_radix, inverse, x)
            ref = np.fft.ifft(x) * nr if inverse else np.fft.fft(x)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            worst = max(worst, err)
            assert err < 1e-13, (nr, inverse, rad, err)
            assert int(np.prod(rad)) == nr
            primes = {2, 3, 5, 7, 11, 13}
            assert all(r <= max_radix or r in primes for r in rad)
    print(f"max radix {max_radix}: worst relative error {worst:.2e}")


def test_round_trip_and_pass_counts():
    """Inverse then forward is Nr times the identity; with radix-16 passes
    the long series (several butterflies per lane) need at most four
    passes (the radix-2..8 design needed up to eight)."""
    for nr in (22, 96, 384, 512, 810, 2016, 2048):
        x = np.random.default_rng(nr).standard_normal(nr) + 0j
        y, rad = _run(nr, 16, True, x)
        z, _ = _run(nr, 16, False, y)
        assert np.abs(z / nr - x).max() < 1e-13
        if nr >= 384:
            assert len(rad) <= 4, (nr, rad)
