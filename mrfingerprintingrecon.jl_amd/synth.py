"""Host-side input generators: trajectories, temporal basis, coil maps and phantoms.

Array convention used throughout this package: a NumPy/torch array is the
row-major view of the reference's column-major Julia array, so every shape is
reversed and the memory is identical.  Julia ``trj (d, Nsamp, Nt)`` is here
``(Nt, Nsamp, d)``; ``data (Nsamp, Nt, C)`` is ``(C, Nt, Nsamp)``; ``U (Nt, K)``
is ``(K, Nt)``; an image ``(Nx, Ny, K)`` is ``(K, Ny, Nx)``.

The trajectory functions restate the reference's exported generators
(`src/Trajectories.jl`).  They produce inputs; they are not on the accelerated
path.
"""
from __future__ import annotations

import math

import numpy as np


def calculate_golden_means():
    """3D golden means (src/Trajectories.jl:168-174): eigenvector of the largest
    eigenvalue of [[0,1,0],[0,0,1],[1,0,1]], normalised by its third entry."""
    m = np.array([[0, 1, 0], [0, 0, 1], [1, 0, 1]], dtype=np.float64)
    w, v = np.linalg.eig(m)
    i = int(np.argmax(w.real))
    vec = v[:, i]
    return float((vec[0] / vec[2]).real), float((vec[1] / vec[2]).real)


def traj_kooshball(nr, theta, phi, delay=(0.0, 0.0, 0.0)):
    """3D radial readouts for given angles (src/Trajectories.jl:108-154, the
    thetaRot = phiRot = 0 branch).

    theta, phi: Julia-shaped angle tables indexed [ic, it] (Ncyc, Nt).  Returns
    k of shape (Nt, nr*Ncyc, 3) (Julia (3, nr*Ncyc, Nt)) in the dtype of theta,
    clamped to [-0.5, 0.5] (src/Trajectories.jl:131)."""
    theta = np.asarray(theta)
    phi = np.asarray(phi)
    dt = theta.dtype
    ncyc, nt = theta.shape
    kr = (np.arange(nr, dtype=np.float64) - (nr - 1) / 2.0) / nr  # :113
    st = np.sin(theta).astype(np.float64)
    ct = np.cos(theta).astype(np.float64)
    sp = np.sin(phi).astype(np.float64)
    cp = np.cos(phi).astype(np.float64)
    # k[c, ir, ic, it]
    k = np.empty((3, nr, ncyc, nt), dtype=np.float64)
    k[0] = -(st * cp)[None] * (kr[:, None, None] + delay[0])
    k[1] = (st * sp)[None] * (kr[:, None, None] + delay[1])
    k[2] = ct[None] * (kr[:, None, None] + delay[2])
    k = k.astype(dt)
    k = np.clip(k, dt.type(-0.5), dt.type(0.5))
    # k[c, ir, ic, it] -> out[it, ir + nr*ic, c]
    out = np.transpose(k, (3, 2, 1, 0)).reshape(nt, ncyc * nr, 3)
    return np.ascontiguousarray(out)


def traj_2d_radial_goldenratio(nr, ncyc, nt, n_tiny=1, dtype=np.float32):
    """2D golden-angle radial trajectory (src/Trajectories.jl:81-93).

    Spoke s = it + Nt*ic (0-based) has angle s*pi/(tau + n_tiny - 1) computed in
    `dtype`.  Returns (nt, nr*ncyc, 2) (Julia (2, nr*ncyc, nt))."""
    tau = (math.sqrt(5.0) + 1.0) / 2.0
    ang = dtype(math.pi / (tau + n_tiny - 1))
    s = np.arange(ncyc * nt)
    phi = (s.astype(dtype) * ang).astype(dtype)
    phi = phi.reshape(nt, ncyc, order="F")  # phi[it, ic]
    theta = np.full_like(phi, dtype(math.pi / 2))
    k = traj_kooshball(nr, theta.T, phi.T)
    return np.ascontiguousarray(k[:, :, :2])


def traj_kooshball_goldenratio(nr, ncyc, nt):
    """3D golden-means koosh-ball (src/Trajectories.jl:51-60); float64 like the
    reference, callers cast to float32."""
    gm1, gm2 = calculate_golden_means()
    s = np.arange(ncyc * nt, dtype=np.float64)
    theta = np.arccos(np.mod(s * gm1, 1.0))
    phi = s * 2.0 * math.pi * gm2
    theta = theta.reshape(nt, ncyc, order="F")
    phi = phi.reshape(nt, ncyc, order="F")
    return traj_kooshball(nr, theta.T, phi.T)


def t2_basis(nt, k, n_atoms=200, seed=0, t2_range=(20.0, 300.0), tr_total_ms=1000.0,
             complex_out=True):
    """Temporal basis U, Julia (Nt, K), returned as (K, Nt): first K left
    singular vectors of a dictionary of exp(-t/T2) curves, T2 uniform in
    t2_range (ms), t_n = n * tr_total/Nt."""
    rng = np.random.default_rng(seed)
    t2 = rng.uniform(t2_range[0], t2_range[1], size=n_atoms)
    t = (np.arange(1, nt + 1) * (tr_total_ms / nt))[:, None]
    d = np.exp(-t / t2[None, :])
    u, _, _ = np.linalg.svd(d, full_matrices=False)
    u = np.ascontiguousarray(u[:, :k].T)
    if complex_out:
        return u.astype(np.complex64)
    return u.astype(np.float32)


def gaussian_coil_maps(img_shape, ncoils, seed=0, radius_frac=0.6, sigma_frac=1.0 / 3.0):
    """Smooth complex coil sensitivities: Gaussian magnitude profiles centred on
    a circle (2D) or on a golden-spiral sphere (3D), a seeded constant phase per
    coil, normalised to root-sum-of-squares 1 per voxel.  Returns
    (ncoils, *img_shape[::-1]) complex64 in row-major order, i.e. each coil is a
    Julia column-major array of size img_shape."""
    rng = np.random.default_rng(seed)
    d = len(img_shape)
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64) - n / 2 for n in img_shape],
                        indexing="ij")  # grids[i] has shape img_shape (x first)
    fov = float(max(img_shape))
    r0 = radius_frac * fov / 2
    sig = sigma_frac * fov
    maps = np.empty((ncoils,) + tuple(img_shape), dtype=np.complex128)
    for c in range(ncoils):
        if d == 2:
            a = 2 * math.pi * c / ncoils
            centre = (r0 * math.cos(a), r0 * math.sin(a))
        else:
            zc = 1 - 2 * (c + 0.5) / ncoils
            rc = math.sqrt(max(0.0, 1 - zc * zc))
            a = c * math.pi * (3 - math.sqrt(5))
            centre = (r0 * rc * math.cos(a), r0 * rc * math.sin(a), r0 * zc)
        dist2 = sum((g - cc) ** 2 for g, cc in zip(grids, centre))
        ph = rng.uniform(0, 2 * math.pi)
        maps[c] = np.exp(-dist2 / (2 * sig * sig)) * np.exp(1j * ph)
    rss = np.sqrt(np.sum(np.abs(maps) ** 2, axis=0))
    maps /= rss[None]
    # to row-major storage of a column-major array: transpose spatial axes
    out = np.transpose(maps, (0,) + tuple(range(d, 0, -1)))
    return np.ascontiguousarray(out.astype(np.complex64))


def ellipse_phantom(img_shape, k, seed=0, n_ell=6):
    """K piecewise-constant random complex ellipse (ellipsoid) maps.
    Returns (k, *img_shape[::-1]) complex64 (column-major images)."""
    rng = np.random.default_rng(seed)
    d = len(img_shape)
    grids = np.meshgrid(*[(np.arange(n) - n / 2) / (n / 2) for n in img_shape],
                        indexing="ij")
    out = np.zeros((k,) + tuple(img_shape), dtype=np.complex128)
    for ik in range(k):
        for _ in range(n_ell):
            c = rng.uniform(-0.5, 0.5, size=d)
            ax = rng.uniform(0.1, 0.45, size=d)
            val = rng.normal() + 1j * rng.normal()
            if d == 2:
                ang = rng.uniform(0, math.pi)
                x = grids[0] - c[0]
                y = grids[1] - c[1]
                xr = math.cos(ang) * x + math.sin(ang) * y
                yr = -math.sin(ang) * x + math.cos(ang) * y
                inside = (xr / ax[0]) ** 2 + (yr / ax[1]) ** 2 <= 1
            else:
                inside = sum(((g - cc) / a) ** 2 for g, cc, a in zip(grids, c, ax)) <= 1
            out[ik][inside] += val
    out = np.transpose(out, (0,) + tuple(range(d, 0, -1)))
    return np.ascontiguousarray(out.astype(np.complex64))


def gaussian_blob_kspace(trj, nsamp_frames, img_shape, U, ncoils, seed=0, n_blobs=4,
                         coil_radius_frac=0.6, coil_sigma_frac=1.0 / 3.0, snr_db=30.0):
    """Analytic k-space data for coefficient maps made of Gaussian blobs seen
    through Gaussian coil profiles (no NUFFT needed: the product of Gaussians is a
    Gaussian whose continuous Fourier transform is closed-form).

    trj: (Nt, Nsamp, d) float32; U: (K, Nt).  Returns data (C, Nt, Nsamp)
    complex64 (Julia (Nsamp, Nt, C)) plus complex Gaussian noise at `snr_db`."""
    rng = np.random.default_rng(seed)
    nt, ns, d = trj.shape
    k = np.asarray(U).shape[0]
    fov = float(max(img_shape))
    r0 = coil_radius_frac * fov / 2
    sc = coil_sigma_frac * fov
    kx = np.transpose(trj.astype(np.float64), (2, 1, 0))  # (d, ns, nt)
    data = np.zeros((ncoils, nt, ns), dtype=np.complex128)
    blobs = []
    for ik in range(k):
        for _ in range(n_blobs):
            c = rng.uniform(-0.25, 0.25, size=d) * fov
            s = rng.uniform(0.03, 0.12) * fov
            a = rng.normal() + 1j * rng.normal()
            blobs.append((ik, c, s, a))
    for ic in range(ncoils):
        if d == 2:
            ang = 2 * math.pi * ic / ncoils
            cc = np.array([r0 * math.cos(ang), r0 * math.sin(ang)])
        else:
            zc = 1 - 2 * (ic + 0.5) / ncoils
            rc = math.sqrt(max(0.0, 1 - zc * zc))
            ang = ic * math.pi * (3 - math.sqrt(5))
            cc = np.array([r0 * rc * math.cos(ang), r0 * rc * math.sin(ang), r0 * zc])
        cph = np.exp(1j * rng.uniform(0, 2 * math.pi))
        for ik, c, s, a in blobs:
            # product of exp(-|r-c|^2/2s^2) and exp(-|r-cc|^2/2sc^2)
            var = 1.0 / (1.0 / s ** 2 + 1.0 / sc ** 2)
            mu = var * (c / s ** 2 + cc / sc ** 2)
            amp = a * cph * math.exp(-0.5 * float(np.sum((c - cc) ** 2)) / (s ** 2 + sc ** 2))
            # FT convention of the forward model: sum_r x(r) exp(-2 pi i k.r)
            k2 = np.sum(kx ** 2, axis=0)  # (ns, nt)
            phase = np.tensordot(mu, kx, axes=(0, 0))
            ft = amp * (2 * math.pi * var) ** (d / 2) * np.exp(-2 * math.pi ** 2 * var * k2) \
                * np.exp(-2j * math.pi * phase)
            data[ic] += (ft * np.asarray(U)[ik, :][None, :]).T
    rms = math.sqrt(float(np.mean(np.abs(data) ** 2)))
    sigma = rms * 10 ** (-snr_db / 20) / math.sqrt(2)
    data += sigma * (rng.standard_normal(data.shape) + 1j * rng.standard_normal(data.shape))
    return np.ascontiguousarray(data.astype(np.complex64))
