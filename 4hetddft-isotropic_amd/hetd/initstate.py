"""Synthetic initial states and the mode-2 checkpoint format.

The reference has no density builder: it reads the starting state from a
text file.  Mode 2 (readenc.f90:104-113) reads, after the '#' comment lines
skipped by titols (titols.f90:12-42), one list-directed record
``xmax ymax zmax hx hy hz nx ny nz``, then ``rimp(3)``, ``vimp(3)`` (internal
units, Angstrom*K) and the complex wave function psi(nx,ny,nz) with x fastest.
That is exactly what printoutc writes (printoutc.f90:52-76), so the same file
serves as input and as checkpoint.

``woods_saxon_state`` builds the synthetic droplet used for every workload:
Woods-Saxon density rho0/(1+exp((r-R)/a)), rho0=0.0218 A^-3, R=2.22 N^(1/3) A,
a=0.7 A, renormalised to N atoms; phase = seeded white noise (splitmix64) on
[-1,1), smoothed with a Gaussian of width sigma in k-space and scaled to a
maximum of ``phase_amp`` rad; psi = sqrt(rho) exp(i phi).
"""
from __future__ import annotations

import numpy as np

KTOPS = 7.63822291  # 1 ps in K^-1 (DFT4He3d.f90:238)

_MASK64 = (1 << 64) - 1


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """n doubles on [-1, 1) from the splitmix64 sequence started at ``seed``.

    splitmix64 (Steele, Lea & Flood 2014, constants as published by Vigna):
    state += 0x9E3779B97F4A7C15; z = state; z = (z ^ z>>30) * 0xBF58476D1CE4E5B9;
    z = (z ^ z>>27) * 0x94D049BB133111EB; z ^= z>>31.  The top 53 bits give a
    double in [0,1), mapped to [-1,1).
    """
    with np.errstate(over="ignore"):
        st = np.uint64(seed & _MASK64)
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = st + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def grid_axes(nx, ny, nz, xmax, ymax, zmax):
    """Real-space axes x(i) = -xmax + h*(i-1), h = 2|xmax|/n (DFT4He3d.f90:274-276,309-317)."""
    hx, hy, hz = 2.0 * abs(xmax) / nx, 2.0 * abs(ymax) / ny, 2.0 * abs(zmax) / nz
    x = -xmax + hx * np.arange(nx)
    y = -ymax + hy * np.arange(ny)
    z = -zmax + hz * np.arange(nz)
    return x, y, z, (hx, hy, hz)


def woods_saxon_state(nx, ny, nz, xmax, ymax, zmax, n4, seed,
                      rho0=0.0218, diffuseness=0.7, sigma_phase=2.0,
                      phase_amp=1e-3, center=(0.0, 0.0, 0.0)):
    """Return psi with shape (nz, ny, nx) complex128 (x fastest in memory)."""
    x, y, z, (hx, hy, hz) = grid_axes(nx, ny, nz, xmax, ymax, zmax)
    radius = 2.22 * n4 ** (1.0 / 3.0)
    zz, yy, xx = np.meshgrid(z - center[2], y - center[1], x - center[0], indexing="ij")
    r = np.sqrt(xx * xx + yy * yy + zz * zz)
    arg = np.minimum((r - radius) / diffuseness, 700.0)
    rho = rho0 / (1.0 + np.exp(arg))
    dv = hx * hy * hz
    rho *= n4 / (rho.sum() * dv)
    noise = splitmix64_uniform(seed, nx * ny * nz).reshape(nz, ny, nx)
    kx = np.fft.fftfreq(nx, d=hx)
    ky = np.fft.fftfreq(ny, d=hy)
    kz = np.fft.fftfreq(nz, d=hz)
    kzz, kyy, kxx = np.meshgrid(kz, ky, kx, indexing="ij")
    k2 = kxx ** 2 + kyy ** 2 + kzz ** 2
    smooth = np.fft.ifftn(np.fft.fftn(noise) * np.exp(-2.0 * (np.pi * sigma_phase) ** 2 * k2)).real
    peak = np.abs(smooth).max()
    phi = smooth * (phase_amp / peak) if peak > 0 else smooth
    return np.sqrt(rho) * np.exp(1j * phi)


def _fmt(v: float) -> str:
    return "%.17E" % v


def write_mode2(path, psi, xmax, ymax, zmax, rimp, vimp_internal, header=None):
    """Write a mode-2 file (the printoutc layout) readable by readenc mode 2.

    psi has shape (nz, ny, nx); vimp_internal is in Angstrom*K (the reference
    multiplies the namelist velocity in A/ps by 7.63822291, DFT4He3d.f90:302).
    """
    nz, ny, nx = psi.shape
    hx, hy, hz = 2.0 * abs(xmax) / nx, 2.0 * abs(ymax) / ny, 2.0 * abs(zmax) / nz
    flat = np.ascontiguousarray(psi).reshape(-1)
    with open(path, "w") as f:
        f.write("#  \n")
        f.write("#  %s\n" % (header or "synthetic Woods-Saxon state"))
        f.write("#  \n")
        f.write(" ".join(_fmt(v) for v in (xmax, ymax, zmax, hx, hy, hz))
                + " %d %d %d\n" % (nx, ny, nz))
        f.write(" ".join(_fmt(v) for v in rimp) + "\n")
        f.write(" ".join(_fmt(v) for v in vimp_internal) + "\n")
        re = flat.real
        im = flat.imag
        chunk = 4
        lines = []
        for i in range(0, flat.size, chunk):
            lines.append(" ".join("(%s,%s)" % (_fmt(a), _fmt(b))
                                  for a, b in zip(re[i:i + chunk], im[i:i + chunk])))
        f.write("\n".join(lines))
        f.write("\n")


def read_mode2(path):
    """Read a mode-2 file; returns (psi (nz,ny,nx), grid dict, rimp, vimp)."""
    with open(path) as f:
        lines = f.readlines()
    i = 0
    while i < len(lines) and lines[i].startswith("#"):
        i += 1
    tok = lines[i].split()
    xmax, ymax, zmax, hx, hy, hz = (float(t.replace("D", "E")) for t in tok[:6])
    nx, ny, nz = (int(t) for t in tok[6:9])
    rest = " ".join(lines[i + 1:]).replace("D", "E")
    rest = rest.replace("(", " ").replace(")", " ").replace(",", " ")
    vals = np.array(rest.split(), dtype=np.float64)
    rimp = vals[0:3].copy()
    vimp = vals[3:6].copy()
    c = vals[6:6 + 2 * nx * ny * nz]
    psi = (c[0::2] + 1j * c[1::2]).reshape(nz, ny, nx)
    grid = dict(nx=nx, ny=ny, nz=nz, xmax=xmax, ymax=ymax, zmax=zmax, hx=hx, hy=hy, hz=hz)
    return psi, grid, rimp, vimp


def smooth_sizes(lo, hi):
    """Even grid edges in [lo, hi] with no prime factor above 5 (what the FFT
    passes accept), ascending."""
    out = []
    for n in range(lo + (lo & 1), hi + 1, 2):
        m = n
        for q in (2, 3, 5):
            while m % q == 0:
                m //= q
        if m == 1:
            out.append(n)
    return out


def ensemble_members(nruns=32, seed0=100, spacing=0.4, nmin=64, nmax=192):
    """Independent droplet + impurity runs of uneven size: per seed, He number
    N uniform in 200..3000, a cubic grid of the smallest FFT-friendly edge
    >= (2R + 16 A)/spacing (R = 2.22 N^(1/3)) clamped to [nmin, nmax], and an
    impurity at the centre moving at 0.5..4.0 A/ps in a uniform random
    direction.  Returns dicts n4, n, xmax, seed, v (A/ps)."""
    sizes = smooth_sizes(nmin, nmax)
    out = []
    for k in range(nruns):
        seed = seed0 + k
        u = splitmix64_uniform(seed, 4)  # uniform on [-1, 1)
        w = 0.5 * (u + 1.0)               # uniform on [0, 1)
        n4 = 200 + int(w[0] * 2801)
        radius = 2.22 * n4 ** (1.0 / 3.0)
        need = (2.0 * radius + 16.0) / spacing
        n = next((s for s in sizes if s >= need), sizes[-1])
        speed = 0.5 + 3.5 * w[1]
        cz = 2.0 * w[2] - 1.0
        ph = 2.0 * np.pi * w[3]
        sz = np.sqrt(max(0.0, 1.0 - cz * cz))
        v = (speed * sz * np.cos(ph), speed * sz * np.sin(ph), speed * cz)
        out.append({"n4": n4, "n": n, "xmax": 0.5 * n * spacing, "seed": seed, "v": v})
    return out
