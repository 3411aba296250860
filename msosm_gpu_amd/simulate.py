"""Seeded synthetic complex baseband with a dispersed pulsar, packed as the
reference's offset-binary uint16 pairs (utils/include/globals.h:21-25, code 0
reserved for "no data").

Pulse model: phase phi_n = frac(n / (P fs)); envelope
g(phi) = exp(-(phi - 1/2)^2 / (2 sigma^2)) with FWHM = duty (sigma =
duty / 2.3548); pulsar s_n = sqrt(A g(phi_n)) z_n with z_n ~ CN(0, 1) and A set
so the period-averaged pulsar power is `snr` times the noise power.  The
pulsar is dispersed by multiplying its spectrum with exp(+i phi(f)),
phi(f) = 2 pi D DM f^2 / (f0^2 (f0 + f)), D = 4.15e15, f measured from the band
bottom f0 (FP64, one circular FFT over the whole signal), then CN(0, 1) noise
is added.  The reference's own simulator (simulated_pulsar/) injects two unit
impulses per period instead; this model follows the benchmark configurations.
"""
import numpy as np

KDM = 4.15e15


def _fft(x, inverse=False):
    try:
        import scipy.fft as sf
        return sf.ifft(x, workers=-1) if inverse else sf.fft(x, workers=-1)
    except ImportError:  # pragma: no cover
        return np.fft.ifft(x) if inverse else np.fft.fft(x)


def dispersion_phase(L, fs, f0, dm):
    """phi(f) for the bins of an L-point FFT of complex baseband at fs, band
    bottom f0 (FFT bin b has baseband frequency b*fs/L, wrapped)."""
    b = np.arange(L, dtype=np.float64)
    nu = np.where(b < L / 2, b, b - L) * (fs / L)
    f = nu + fs / 2
    return 2 * np.pi * KDM * dm * f * f / (f0 * f0 * (f0 + f))


def pulsar_baseband(L, fs, f0, dm, period, duty=0.05, snr=0.1, seed=0x5EED0000, noise=True):
    """Complex64 baseband (noise power 1) with a dispersed pulsar."""
    rng = np.random.Generator(np.random.Philox(seed))
    n = np.arange(L, dtype=np.float64)
    phase = np.mod(n / (period * fs), 1.0)
    sigma = duty / 2.3548
    g = np.exp(-(phase - 0.5) ** 2 / (2 * sigma * sigma))
    A = snr / g.mean() if g.mean() > 0 else 0.0
    z = (rng.standard_normal(L, dtype=np.float32) + 1j * rng.standard_normal(L, dtype=np.float32)) * np.float32(
        np.sqrt(0.5))
    s = (np.sqrt(A * g).astype(np.float32) * z).astype(np.complex64)
    if dm != 0:
        S = _fft(s)
        S *= np.exp(1j * dispersion_phase(L, fs, f0, dm)).astype(np.complex64)
        s = _fft(S, inverse=True).astype(np.complex64)
    if noise:
        s += ((rng.standard_normal(L, dtype=np.float32) + 1j * rng.standard_normal(L, dtype=np.float32))
              * np.float32(np.sqrt(0.5)))
    return s


def to_u16(x, scale=1000.0):
    """Offset-binary uint16 pairs (2L values), clipped to [1, 65535]."""
    v = np.empty(2 * x.size, np.float32)
    v[0::2] = x.real
    v[1::2] = x.imag
    u = np.rint(v * np.float32(scale)) + np.float32(32768)
    return np.clip(u, 1, 65535).astype(np.uint16)


def pulsar_u16(L, fs, f0, dm, period, duty=0.05, snr=0.1, seed=0x5EED0000, scale=1000.0):
    return to_u16(pulsar_baseband(L, fs, f0, dm, period, duty, snr, seed), scale)
