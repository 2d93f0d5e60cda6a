"""Seeded synthetic microstructures for parity tests, benchmarks and training
data.  Images are uint8 arrays shaped (z, y, x) in file order, phase colours 0
and 255.  Every generator draws from numpy's PCG64 with the given seed."""
import numpy as np


def grf(shape, sigma=2.0, vf=0.4, seed=1):
    """Thresholded periodic Gaussian random field: i.i.d. N(0,1), periodic
    Gaussian blur of standard deviation sigma voxels, rank threshold so that
    exactly round(vf*N) voxels get colour 255 (the rest 0)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    noise = rng.standard_normal(shape)
    k = [np.fft.fftfreq(n) for n in shape]
    kz, ky, kx = np.meshgrid(*k, indexing="ij")
    filt = np.exp(-2.0 * (np.pi * sigma) ** 2 * (kx ** 2 + ky ** 2 + kz ** 2))
    field = np.real(np.fft.ifftn(np.fft.fftn(noise) * filt))
    n_on = int(round(vf * field.size))
    order = np.argsort(field, axis=None, kind="stable")
    img = np.zeros(field.size, dtype=np.uint8)
    img[order[field.size - n_on:]] = 255
    return img.reshape(shape)


def spheres(shape, rmin=6, rmax=12, fraction=0.3, seed=2, color=255, background=0):
    """Overlapping spheres with centres uniform in the periodic box and radii
    uniform in [rmin, rmax], added until the sphere phase covers at least
    `fraction` of the voxels."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nz, ny, nx = shape
    img = np.full(shape, background, dtype=np.uint8)
    target = fraction * img.size
    count = 0
    while count < target:
        c = rng.uniform(0, 1, 3) * np.array([nz, ny, nx])
        r = rng.uniform(rmin, rmax)
        ri = int(np.ceil(r))
        zz = np.arange(int(np.floor(c[0])) - ri, int(np.floor(c[0])) + ri + 2)
        yy = np.arange(int(np.floor(c[1])) - ri, int(np.floor(c[1])) + ri + 2)
        xx = np.arange(int(np.floor(c[2])) - ri, int(np.floor(c[2])) + ri + 2)
        d2 = ((zz[:, None, None] + 0.5 - c[0]) ** 2 + (yy[None, :, None] + 0.5 - c[1]) ** 2 +
              (xx[None, None, :] + 0.5 - c[2]) ** 2)
        inside = d2 <= r * r
        iz, iy, ix = np.nonzero(inside)
        sub = img[np.ix_(zz % nz, yy % ny, xx % nx)]
        new = (sub[iz, iy, ix] != color).sum()
        sub[iz, iy, ix] = color
        img[np.ix_(zz % nz, yy % ny, xx % nx)] = sub
        count += new
    return img


def percolates(pore_mask):
    """True if the pore voxels connect opposite faces (6-connectivity) along
    each of x, y and z."""
    from scipy import ndimage
    lab, _ = ndimage.label(pore_mask)
    for axis in range(3):
        a = np.take(lab, 0, axis=axis)
        b = np.take(lab, -1, axis=axis)
        common = np.intersect1d(a[a > 0], b[b > 0])
        if common.size == 0:
            return False
    return True


def porous_spheres(shape, rmin=6, rmax=12, porosity=0.35, seed=4):
    """Solid overlapping spheres (colour 255) in a pore matrix (colour 0) until
    the porosity is at most `porosity`; the seed is bumped until the pore space
    percolates in x, y and z.  Returns (image, seed used)."""
    s = seed
    while True:
        img = spheres(shape, rmin, rmax, 1.0 - porosity, seed=s, color=255, background=0)
        if percolates(img == 0):
            return img, s
        s += 1


def elem_map(img, colors):
    """Material index per element in element order (row + col*ny + layer*nx*ny)
    for an image in file order; colors lists the colours in .nf order."""
    lut = np.zeros(256, dtype=np.uint16)
    for i, c in enumerate(colors):
        lut[c] = i
    return np.ascontiguousarray(lut[img].transpose(0, 2, 1)).reshape(-1)
