"""Quaternion helpers on float64 numpy arrays (component order (1, i, j, k)).

Replaces python/proj/quat.py, which works on spt3g's quat / G3VectorQuat.
Here a single quaternion is a ``Quat`` of shape (4,) and a vector of them a
``Quat`` of shape (n, 4); ``*`` between two Quat objects is the Hamilton
product (broadcast over leading axes), ``*`` with a plain number or array
is elementwise, and ``~q`` is the conjugate.  Anything that is not a
quaternion (slices of single components, results of numpy functions) comes
back as a plain ndarray, so component arithmetic is elementwise.

The Hamilton product uses the term order of boost::math::quaternion's
operator*= (the product the reference's C++ engine uses,
src/Projection.cxx:47); spt3g's own order is not visible from the
reference sources.
"""
import numpy as np

DEG = np.pi / 180

__all__ = ['DEG', 'Quat', 'quat', 'G3VectorQuat', 'qmul', 'euler', 'rotation_iso',
           'rotation_lonlat', 'rotation_xieta', 'decompose_iso', 'decompose_lonlat',
           'decompose_xieta']


def qmul(l, r):
    """Hamilton product l*r of (..., 4) arrays, broadcasting leading axes."""
    l = np.asarray(l, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    a, b, c, d = l[..., 0], l[..., 1], l[..., 2], l[..., 3]
    ar, br, cr, dr = r[..., 0], r[..., 1], r[..., 2], r[..., 3]
    out = np.empty(np.broadcast(a, ar).shape + (4,))
    out[..., 0] = a * ar - b * br - c * cr - d * dr
    out[..., 1] = a * br + b * ar + c * dr - d * cr
    out[..., 2] = a * cr - b * dr + c * ar + d * br
    out[..., 3] = a * dr + b * cr - c * br + d * ar
    return out


class Quat(np.ndarray):
    """float64 array whose last axis (length 4) holds quaternion components."""

    def __new__(cls, data):
        a = np.array(data, dtype=np.float64)
        if a.shape[-1:] != (4,):
            raise ValueError(f'quaternion arrays need a last axis of length 4, got {a.shape}')
        return a.view(cls)

    # Results of numpy functions and ufuncs are plain arrays.
    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        inputs = tuple(np.asarray(x) if isinstance(x, Quat) else x for x in inputs)
        if 'out' in kwargs:
            kwargs['out'] = tuple(np.asarray(x) if isinstance(x, Quat) else x
                                  for x in kwargs['out'])
        return getattr(ufunc, method)(*inputs, **kwargs)

    def __array_function__(self, func, types, args, kwargs):
        def plain(x):
            if isinstance(x, Quat):
                return np.asarray(x)
            if isinstance(x, (list, tuple)):
                return type(x)(plain(y) for y in x)
            return x
        return func(*plain(args), **{k: plain(v) for k, v in kwargs.items()})

    def __getitem__(self, key):
        out = np.asarray(self)[key]
        if not isinstance(out, np.ndarray):
            return out
        # Indexing that leaves the component axis intact keeps the type.
        keeps_last = (self.ndim >= 2 and (not isinstance(key, tuple) or
                                          (len(key) < self.ndim and Ellipsis not in key)))
        if keeps_last and out.shape[-1:] == (4,):
            return out.view(Quat)
        return out

    def __mul__(self, other):
        if isinstance(other, Quat):
            return Quat(qmul(self, other))
        return np.multiply(np.asarray(self), other)

    def __rmul__(self, other):
        if isinstance(other, Quat):
            return Quat(qmul(other, self))
        return np.multiply(other, np.asarray(self))

    def __imul__(self, other):
        res = self * other
        self[...] = res
        return self

    def __invert__(self):
        out = np.array(self)
        out[..., 1:] *= -1
        return Quat(out)

    @property
    def a(self):
        return np.asarray(self)[..., 0]

    @property
    def b(self):
        return np.asarray(self)[..., 1]

    @property
    def c(self):
        return np.asarray(self)[..., 2]

    @property
    def d(self):
        return np.asarray(self)[..., 3]


def quat(a, b=0., c=0., d=0.):
    """A single quaternion a + b i + c j + d k."""
    return Quat([a, b, c, d])


def G3VectorQuat(data=()):
    """A vector of quaternions, shape (n, 4)."""
    a = np.array(data, dtype=np.float64)
    if a.size == 0:
        a = np.zeros((0, 4))
    return Quat(a.reshape(-1, 4))


def euler(axis, angle):
    """Quaternion(s) of rotation by angle (radians) about cartesian axis
    0, 1 or 2: (cos(angle/2), ..., sin(angle/2) in slot axis+1, ...).
    Same construction as python/proj/quat.py:15-44."""
    angle = np.asarray(angle)
    shape = np.broadcast(axis, angle).shape + (4,)
    c, s = np.cos(angle / 2), np.sin(angle / 2)
    q = np.zeros(shape)
    q[..., 0] = c
    q[..., axis + 1] = s
    return Quat(q)


def rotation_iso(theta, phi, psi=None):
    """Qz(phi) Qy(theta) Qz(psi)."""
    out = euler(2, phi) * euler(1, theta)
    if psi is None:
        return out
    return out * euler(2, psi)


def rotation_lonlat(lon, lat, psi=0., azel=False):
    """Qz(lon) Qy(pi/2 - lat) Qz(psi); azel flips the sign of lon."""
    if azel:
        return rotation_iso(np.pi / 2 - lat, -lon, psi)
    return rotation_iso(np.pi / 2 - lat, lon, psi)


def rotation_xieta(xi, eta, gamma=0):
    """Rotation taking the focal-plane centre to (xi, eta, gamma):
    xi = -sin(theta) sin(phi), eta = -sin(theta) cos(phi), gamma = psi + phi."""
    phi = np.arctan2(-xi, -eta)
    theta = np.arcsin(np.sqrt(np.asarray(xi) ** 2 + np.asarray(eta) ** 2))
    psi = gamma - phi
    return rotation_iso(theta, phi, psi)


def _components(q):
    q = np.asarray(q, dtype=np.float64)
    return q[..., 0], q[..., 1], q[..., 2], q[..., 3]


def decompose_iso(q):
    """(theta, phi, psi) with q = Qz(phi) Qy(theta) Qz(psi)."""
    a, b, c, d = _components(q)
    psi = np.arctan2(a * b + c * d, a * c - b * d)
    phi = np.arctan2(c * d - a * b, a * c + b * d)
    theta = 2 * np.arctan2((b ** 2 + c ** 2) ** .5, (a ** 2 + d ** 2) ** .5)
    return (theta, phi, psi)


def decompose_lonlat(q, azel=False):
    theta, phi, psi = decompose_iso(q)
    if azel:
        return (-phi, np.pi / 2 - theta, psi)
    return (phi, np.pi / 2 - theta, psi)


def decompose_xieta(q):
    a, b, c, d = _components(q)
    return (2 * (a * b - c * d), -2 * (a * c + b * d),
            np.arctan2(2 * a * d, a * a - d * d))
