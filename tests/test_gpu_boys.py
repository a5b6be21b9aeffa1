"""The device Boys function (Taylor tables + division-free recursions) against
the integral definition F_m(T) = int_0^1 t^(2m) exp(-T t^2) dt at 40 digits:
the full table (dT = 0.05, 8 terms: cooperative and one-electron kernels) and
the compact per-L table (dT = 0.2, 12 coefficients F_{L+k}/k! with the
derivative giving F_{L+1}: the straight-line quartet kernels, staged in LDS)."""
import mpmath as mp
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
mp.mp.dps = 40

TS = np.array([0.0, 1e-12, 1e-6, 0.013, 0.024999, 0.025001, 0.5, 1.0, 3.7, 9.99, 17.3, 29.0, 35.99, 36.01,
               44.4, 59.97, 59.999, 60.0, 60.5, 75.0, 200.0, 1e4])


def ref(m, T):
    return float(mp.quad(lambda t: t ** (2 * m) * mp.e ** (-T * t * t), [0, 1]))


@pytest.mark.parametrize("compact", [False, True])
@pytest.mark.parametrize("L", [0, 1, 4, 8])
def test_device_boys(L, compact):
    from valence_amd import lib
    F = lib.device_boys(L, TS, compact=compact)
    worst = 0.0
    for i, T in enumerate(TS):
        for m in range(L + 1):
            r = ref(m, float(T))
            worst = max(worst, abs(F[i, m] - r) / r)
    print(f"L={L} worst relative error {worst:.2e}")
    assert worst < 5e-15
