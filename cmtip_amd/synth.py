"""Seeded synthetic inputs for the hot path (tests and benchmark).

The reference simulates with skopi (cmtip/simulate.py:119-162), which needs a GPU
and a PDB file; here the same kinds of arrays come from closed-form generators:

* ``gaussian_ac``: autocorrelation of a density of isotropic Gaussian atoms (sigma
  in voxels, centres uniform in a ball), computed on a zero-padded (2M)^3 grid,
  fftshifted and cropped to the central M^3 (real and centrosymmetric).
* ``detector``: flat square panel with Ewald curvature, s = (p/|p| - z)/lambda in
  m^-1, laid out as cmtip/prep_data.py:55,86-89 leaves it: (3, 1, P) float32 and
  reciprocal_extent = max |s| (float32).
* ``random_quaternions``: unit quaternions uniform on SO(3) (normalised 4-D
  Gaussians), scalar first, q0 >= 0.  skopi's get_uniform_quat / get_random_quat are
  not available, so reference sets drawn here are not skopi's.
"""
import numpy as np

BEAM_WAVELENGTH_M = 2.695e-10  # 4600 eV, the beam of examples/input/amo86615.beam


def gaussian_density(M, n_atoms, seed=0, sigma=1.0, radius_frac=0.3):
    rng = np.random.default_rng(seed)
    r = radius_frac * M
    # centres uniform in a ball of radius r (voxels), relative to the grid centre
    d = rng.standard_normal((n_atoms, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rad = r * rng.random(n_atoms) ** (1.0 / 3.0)
    centres = d * rad[:, None]
    ax = np.arange(M) - M // 2
    g = [np.exp(-((ax[None, :] - centres[:, k:k + 1]) ** 2) / (2 * sigma ** 2)) for k in range(3)]
    return np.einsum("ai,aj,ak->ijk", g[0], g[1], g[2], optimize=True)


def gaussian_ac(M, n_atoms, seed=0, sigma=1.0, radius_frac=0.3, dtype=np.float32):
    """Autocorrelation (M, M, M) of a Gaussian-atom density; its FT is |rho_hat|^2 >= 0."""
    rho = gaussian_density(M, n_atoms, seed, sigma, radius_frac)
    L = 2 * M
    pad = np.zeros((L, L, L))
    pad[:M, :M, :M] = rho
    F = np.fft.rfftn(pad)
    ac = np.fft.irfftn(np.abs(F) ** 2, s=(L, L, L), axes=(0, 1, 2))
    ac = np.fft.fftshift(ac)
    c = L // 2
    h = M // 2
    out = ac[c - h:c - h + M, c - h:c - h + M, c - h:c - h + M]
    out = out / np.abs(out).max()
    return np.ascontiguousarray(out.astype(dtype))


def detector(n_side, det_width=0.08, det_dist=0.2, wavelength=BEAM_WAVELENGTH_M):
    """(pixel_position_reciprocal (3, 1, P) float32, reciprocal_extent float32)."""
    px = det_width / n_side
    c = (np.arange(n_side) - (n_side - 1) / 2.0) * px
    X, Y = np.meshgrid(c, c, indexing="ij")
    Z = np.full_like(X, det_dist)
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()])
    u = P / np.linalg.norm(P, axis=0, keepdims=True)
    s = (u - np.array([[0.0], [0.0], [1.0]])) / wavelength
    s = s.astype(np.float32).reshape(3, 1, -1)
    ext = np.linalg.norm(s, axis=0).max()
    return s, np.float32(ext)


def random_quaternions(n, seed=0):
    rng = np.random.default_rng(seed)
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[q[:, 0] < 0] *= -1
    return q.astype(np.float32)


def poisson_images(slices, mean_photons, seed=2):
    """Poisson counts with the given mean photons per pixel (slices clipped at 0)."""
    rng = np.random.default_rng(seed)
    sl = np.clip(np.asarray(slices, dtype=np.float64), 0, None)
    scale = mean_photons / max(sl.mean(), 1e-30)
    return rng.poisson(sl * scale).astype(np.float32)


def panel_detector(px=64, py=64, grid=(2, 2), gap=4, det_width=0.08, det_dist=0.2, wavelength=BEAM_WAVELENGTH_M):
    """A flat detector of grid[0] x grid[1] panels of px x py pixels with gaps of ``gap``
    pixels between panels, in the layout load_h5 leaves (prep_data.py:54-93):
    (pixel_index_map (n_panels, px, py, 2) int64, pixel_position_reciprocal
    (3, 1, n_panels * px * py) float32, det_shape (n_panels, px, py), reciprocal_extent)."""
    n_panels = grid[0] * grid[1]
    X = grid[0] * px + (grid[0] - 1) * gap
    Y = grid[1] * py + (grid[1] - 1) * gap
    pix = det_width / max(X, Y)
    imap = np.zeros((n_panels, px, py, 2), dtype=np.int64)
    for a in range(grid[0]):
        for b in range(grid[1]):
            l = a * grid[1] + b
            imap[l, :, :, 0] = a * (px + gap) + np.arange(px)[:, None]
            imap[l, :, :, 1] = b * (py + gap) + np.arange(py)[None, :]
    cx = (imap[..., 0] - (X - 1) / 2.0) * pix
    cy = (imap[..., 1] - (Y - 1) / 2.0) * pix
    P = np.stack([cx.ravel(), cy.ravel(), np.full(cx.size, det_dist)])
    u = P / np.linalg.norm(P, axis=0, keepdims=True)
    s = ((u - np.array([[0.0], [0.0], [1.0]])) / wavelength).astype(np.float32).reshape(3, 1, -1)
    return imap, s, (n_panels, px, py), np.float32(np.linalg.norm(s, axis=0).max())
