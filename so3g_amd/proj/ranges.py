"""Ranges / RangesMatrix: sample-interval sets, the carrier type of the
``threads=`` (thread_intervals) argument of the projection routines.

Mirrors the Python-visible part of so3g's RangesInt32 (src/Ranges.cxx:
712-784) and RangesMatrix (python/proj/ranges.py:11-297) that the
projection path uses; the full interval algebra of so3g is out of scope.
Intervals are half-open [lo, hi) within [0, count).
"""
import numpy as np

__all__ = ['Ranges', 'RangesInt32', 'RangesMatrix']


class Ranges:
    """Sorted, disjoint, non-touching half-open intervals within [0, count)."""

    def __init__(self, count=0, reference=0):
        self._count = int(count)
        self._reference = int(reference)
        self._segs = []          # list of [lo, hi]

    # -- properties -------------------------------------------------------
    @property
    def count(self):
        return self._count

    @count.setter
    def count(self, v):
        self._count = int(v)
        self._clip()

    @property
    def reference(self):
        return self._reference

    @property
    def shape(self):
        return (self._count,)

    def __len__(self):
        return self._count

    def __repr__(self):
        return f'Ranges(n={self._count}:rs={self._segs})'

    __str__ = __repr__

    # -- construction ---------------------------------------------------------
    def _clip(self):
        out = []
        for lo, hi in self._segs:
            lo, hi = max(lo, 0), min(hi, self._count)
            if hi > lo:
                out.append([lo, hi])
        self._segs = out

    def _normalise(self):
        if not self._segs:
            return
        self._segs.sort()
        out = [list(self._segs[0])]
        for lo, hi in self._segs[1:]:
            if lo <= out[-1][1]:
                out[-1][1] = max(out[-1][1], hi)
            else:
                out.append([lo, hi])
        self._segs = out
        self._clip()

    def add_interval(self, lo, hi):
        self._segs.append([int(lo), int(hi)])
        self._normalise()
        return self

    def append_interval_no_check(self, lo, hi):
        """Append [lo, hi) that lies after every existing interval
        (Ranges.h append_interval_no_check)."""
        lo, hi = int(lo), int(hi)
        if self._segs and self._segs[-1][1] == lo:
            self._segs[-1][1] = hi
        else:
            self._segs.append([lo, hi])
        return self

    def merge(self, other):
        self._segs += [list(s) for s in other._segs]
        self._normalise()
        return self

    def ranges(self):
        """(n, 2) int32 array of the intervals."""
        if not self._segs:
            return np.zeros((0, 2), np.int32)
        return np.array(self._segs, np.int32)

    @classmethod
    def from_array(cls, data, count):
        r = cls(count)
        for lo, hi in np.asarray(data).reshape(-1, 2):
            r._segs.append([int(lo), int(hi)])
        r._normalise()
        return r

    @classmethod
    def from_mask(cls, mask):
        mask = np.asarray(mask, bool)
        r = cls(len(mask))
        if len(mask):
            d = np.diff(np.concatenate(([0], mask.astype(np.int8), [0])))
            starts = np.flatnonzero(d == 1)
            ends = np.flatnonzero(d == -1)
            r._segs = [[int(a), int(b)] for a, b in zip(starts, ends)]
        return r

    from_bitmask = from_mask

    def mask(self):
        m = np.zeros(self._count, bool)
        for lo, hi in self._segs:
            m[lo:hi] = True
        return m

    def copy(self):
        r = Ranges(self._count, self._reference)
        r._segs = [list(s) for s in self._segs]
        return r

    def zeros_like(self):
        return Ranges(self._count, self._reference)

    def ones_like(self):
        r = Ranges(self._count, self._reference)
        if self._count > 0:
            r._segs = [[0, self._count]]
        return r

    def buffer(self, n):
        self._segs = [[lo - n, hi + n] for lo, hi in self._segs]
        self._normalise()
        return self

    def buffered(self, n):
        return self.copy().buffer(n)

    def close_gaps(self, gap_size=0):
        out = []
        for lo, hi in self._segs:
            if out and lo - out[-1][1] <= gap_size:
                out[-1][1] = hi
            else:
                out.append([lo, hi])
        self._segs = out
        return self

    def complement(self):
        r = Ranges(self._count, self._reference)
        r._segs = [list(s) for s in self._segs]
        return ~r

    def intersect(self, other):
        return self * other

    def __invert__(self):
        r = Ranges(self._count, self._reference)
        pos = 0
        for lo, hi in self._segs:
            if lo > pos:
                r._segs.append([pos, lo])
            pos = hi
        if pos < self._count:
            r._segs.append([pos, self._count])
        return r

    def __add__(self, other):
        return self.copy().merge(other)

    def __iadd__(self, other):
        return self.merge(other)

    def __mul__(self, other):
        return Ranges.from_mask(self.mask() & other.mask())

    def __imul__(self, other):
        self._segs = Ranges.from_mask(self.mask() & other.mask())._segs
        return self

    def __getitem__(self, sl):
        if not isinstance(sl, slice):
            raise IndexError('Ranges supports only slice indexing')
        return Ranges.from_mask(self.mask()[sl])


RangesInt32 = Ranges


class RangesMatrix:
    """A grid of Ranges; the right-most axis is the sample axis."""

    def __init__(self, items=(), child_shape=None, skip_shape_check=False):
        self.ranges = list(items)
        if len(self.ranges):
            child_shape = self.ranges[0].shape
            if not skip_shape_check:
                assert all(it.shape == child_shape for it in self.ranges)
        elif child_shape is None:
            child_shape = ()
        self._child_shape = tuple(child_shape)

    def __repr__(self):
        return 'RangesMatrix(' + ','.join(map(str, self.shape)) + ')'

    def __len__(self):
        return self.shape[0]

    @property
    def shape(self):
        if not self.ranges:
            return (0,) + self._child_shape
        return (len(self.ranges),) + tuple(self.ranges[0].shape)

    def __getitem__(self, index):
        if isinstance(index, tuple):
            if len(index) == 0:
                return self
            head, rest = index[0], index[1:]
        else:
            head, rest = index, ()
        if isinstance(head, (int, np.integer)):
            sub = self.ranges[head]
            return sub[rest] if rest else sub
        if isinstance(head, slice):
            items = self.ranges[head]
            if rest:
                items = [x[rest] for x in items]
            return RangesMatrix(items, child_shape=self._child_shape)
        idx = np.arange(len(self.ranges))[head]
        items = [self.ranges[i] for i in np.atleast_1d(idx)]
        if rest:
            items = [x[rest] for x in items]
        return RangesMatrix(items, child_shape=self._child_shape)

    def copy(self):
        return RangesMatrix([x.copy() for x in self.ranges], child_shape=self.shape[1:])

    def zeros_like(self):
        return RangesMatrix([x.zeros_like() for x in self.ranges], child_shape=self.shape[1:])

    def ones_like(self):
        return RangesMatrix([x.ones_like() for x in self.ranges], child_shape=self.shape[1:])

    def buffer(self, n):
        for x in self.ranges:
            x.buffer(n)
        return self

    def buffered(self, n):
        return self.copy().buffer(n)

    def __invert__(self):
        return RangesMatrix([~x for x in self.ranges], child_shape=self.shape[1:])

    def __mul__(self, other):
        if isinstance(other, RangesMatrix):
            return RangesMatrix([a * b for a, b in zip(self.ranges, other.ranges)],
                                child_shape=self.shape[1:])
        return RangesMatrix([a * other for a in self.ranges], child_shape=self.shape[1:])

    def __add__(self, other):
        if isinstance(other, RangesMatrix):
            return RangesMatrix([a + b for a, b in zip(self.ranges, other.ranges)],
                                child_shape=self.shape[1:])
        return RangesMatrix([a + other for a in self.ranges], child_shape=self.shape[1:])

    def mask(self, dest=None):
        """Boolean array of self.shape."""
        out = np.zeros(self.shape, bool) if dest is None else dest
        for i, r in enumerate(self.ranges):
            out[i] = r.mask()
        return out

    @staticmethod
    def full(shape, fill_value):
        shape = tuple(shape)
        if len(shape) == 1:
            r = Ranges(shape[0])
            return r.ones_like() if fill_value else r
        return RangesMatrix([RangesMatrix.full(shape[1:], fill_value)
                             for _ in range(shape[0])], child_shape=shape[1:])

    @classmethod
    def zeros(cls, shape):
        return cls.full(shape, False)

    @classmethod
    def ones(cls, shape):
        return cls.full(shape, True)

    @classmethod
    def from_mask(cls, mask):
        mask = np.asarray(mask, bool)
        if mask.ndim == 1:
            return Ranges.from_mask(mask)
        return cls([cls.from_mask(m) for m in mask], child_shape=mask.shape[1:])

    @staticmethod
    def concatenate(items, axis=0):
        if axis != 0:
            raise NotImplementedError('only axis=0 concatenation is supported')
        out = []
        for it in items:
            out += list(it.ranges)
        return RangesMatrix(out)
