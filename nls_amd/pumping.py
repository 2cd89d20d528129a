"""Pumping profiles P(x[, y]) -- the functors of the reference's
``nls/pumping.py`` with the same class names, constructor arguments and
composition rules, plus the single-Gaussian ring used by the standard
workloads.

Behaviour kept from the reference on purpose (a user switching over gets the
same arrays):
  * ``AbstractPumping.__init__`` ignores its ``power`` argument and sets 1.0
    (pumping.py:15-16); classes that want a power set it themselves, so
    ``GaussianElipticPumping2D`` always has power 1.0 (pumping.py:184);
  * ``OpSumPumping`` / ``OpSubPumping`` scale the sum / difference by their
    own power (pumping.py:51-52, :67-68); ``OpMulPumping`` refuses to run;
  * the rings are sums of two Gaussians (or rectangles) at +radius and
    -radius of the radial coordinate (pumping.py:144-178, :224-228).
"""
import numpy as np


class AbstractPumping(object):
    """Base of the pumping tree (pumping.py:11-40)."""

    def __init__(self, power=1.0):
        self.power = 1.0

    def setPower(self, power):
        self.power = power

    def __call__(self, *args, **kwargs):
        raise Exception("Nothing to call: abstract class could not represent pumping.")

    def __add__(self, other):
        return OpSumPumping(self, other)

    def __sub__(self, other):
        return OpSubPumping(self, other)

    def __str__(self):
        return str(repr(self))

    def __repr__(self):
        return u"AbstractPumping"


class OpSumPumping(AbstractPumping):
    """power * (lhs + rhs) (pumping.py:43-58)."""

    def __init__(self, lhs, rhs):
        super(OpSumPumping, self).__init__()
        self.lhs, self.rhs = lhs, rhs

    def __call__(self, *args, **kwargs):
        return self.power * (self.lhs(*args, **kwargs) + self.rhs(*args, **kwargs))

    def __repr__(self):
        return repr(self.lhs) + u" + " + repr(self.rhs)


class OpSubPumping(AbstractPumping):
    """power * (lhs - rhs) (pumping.py:61-74)."""

    def __init__(self, lhs, rhs):
        super(OpSubPumping, self).__init__()
        self.lhs, self.rhs = lhs, rhs

    def __call__(self, *args, **kwargs):
        return self.power * (self.lhs(*args, **kwargs) - self.rhs(*args, **kwargs))

    def __repr__(self):
        return repr(self.lhs) + u" - " + repr(self.rhs)


class OpMulPumping(AbstractPumping):
    """Cartesian product of two profiles; not evaluable (pumping.py:77-92)."""

    def __init__(self, lhs, rhs):
        super(OpMulPumping, self).__init__()
        self.lhs, self.rhs = lhs, rhs

    def __call__(self, *args, **kwargs):
        raise Exception("Not supported yet!")

    def __repr__(self):
        return repr(self.lhs) + u" x " + repr(self.rhs)


class GridPumping(AbstractPumping):
    """A profile given as an array, returned as is (pumping.py:95-110)."""

    def __init__(self, pumping, desciption=None):
        super(GridPumping, self).__init__()
        self.pumping = pumping
        self.desciption = desciption if desciption else "<class GridPumping>"

    def __call__(self, *args, **kwargs):
        return self.pumping

    def __repr__(self):
        return self.desciption


class GaussianPumping(AbstractPumping):
    """power * exp(-((x-x0)^2 + (y-y0)^2) / (2 variation^2)) (pumping.py:113-130)."""

    def __init__(self, power=1.0, x0=0.0, y0=0.0, variation=5.0):
        super(GaussianPumping, self).__init__()
        self.power, self.x0, self.y0, self.variation = power, x0, y0, variation

    def __call__(self, x, y, t=None):
        d2 = (np.asarray(x) - self.x0) ** 2 + (np.asarray(y) - self.y0) ** 2
        return self.power * np.exp(-d2 / (2.0 * self.variation ** 2))

    def __repr__(self):
        return u"{0} exp(-{1} ((x - {2})^2 - (y - {3})^2))".format(
            self.power, 1.0 / (2.0 * self.variation ** 2), self.x0, self.y0)


class GaussianPumping1D(GaussianPumping):
    """The Gaussian on a line, y = 0 (pumping.py:133-141)."""

    def __call__(self, x, t=None):
        return GaussianPumping.__call__(self, x, 0.0, t)


class GaussianPumping2D(GaussianPumping):
    """The Gaussian on the plane (pumping.py:144-149)."""


class GaussianRingPumping1D(OpSumPumping):
    """Gaussians at +radius and -radius of the line (pumping.py:152-159)."""

    def __init__(self, power=1.0, radius=0.0, variation=5.0):
        super(GaussianRingPumping1D, self).__init__(GaussianPumping1D(power, +radius, 0.0, variation),
                                                    GaussianPumping1D(power, -radius, 0.0, variation))


class GaussianRingPumping2D(AbstractPumping):
    """The 1-D ring evaluated on r = |(x, y) - (x0, y0)| (pumping.py:162-178)."""

    def __init__(self, power=1.0, x0=0.0, y0=0.0, variation=5.0, radius=1.0):
        super(GaussianRingPumping2D, self).__init__()
        self.x0, self.y0 = x0, y0
        self.pumping = GaussianRingPumping1D(power, radius, variation)

    def __call__(self, x, y, t=None):
        r = np.sqrt((np.asarray(x) - self.x0) ** 2 + (np.asarray(y) - self.y0) ** 2)
        return self.pumping(r, t)

    def __repr__(self):
        return "<class GaussianRingPumping2D(AbstractPumping)>"


class GaussianElipticPumping2D(AbstractPumping):
    """Gaussian rim of the ellipse with half-axes a, b (pumping.py:181-202):
    power * (exp(-(re - rp)^2 / (2 variation)) + exp(-(re + rp)^2 / (2 variation)))
    with t = atan2(x - x0, y - y0), re = |(a cos t, b sin t)|, rp = |(x, y)|."""

    def __init__(self, power=1.0, x0=0.0, y0=0.0, variation=5.0, a=1.0, b=1.0):
        super(GaussianElipticPumping2D, self).__init__(power)
        self.x0, self.y0, self.var, self.a, self.b = x0, y0, variation, a, b

    def __call__(self, x, y, t=None):
        x, y = np.asarray(x), np.asarray(y)
        th = np.arctan2(x - self.x0, y - self.y0)
        re = np.sqrt((self.a * np.cos(th)) ** 2 + (self.b * np.sin(th)) ** 2)
        rp = np.sqrt(x ** 2 + y ** 2)
        return self.power * (np.exp(-(re - rp) ** 2 / (2 * self.var)) + np.exp(-(re + rp) ** 2 / (2 * self.var)))

    def __repr__(self):
        return u"<class GaussianElipticPumping2D(AbstractPumping)>"


class RectangularPumping1D(AbstractPumping):
    """power on [x0 - width/2, x0 + width/2], 0 elsewhere (pumping.py:205-221)."""

    def __init__(self, power=10.0, x0=0.0, width=1.0):
        super(RectangularPumping1D, self).__init__(power)
        self.x0, self.width, self.power = x0, width, power

    def __call__(self, x, t=None):
        x = np.asarray(x)
        out = np.zeros(x.shape)
        out[(x >= self.x0 - self.width / 2.0) & (x <= self.x0 + self.width / 2.0)] = self.power
        return out

    def __repr__(self):
        return u"{0} * (\\theta(x - {1}) - \\theta({2} - x))".format(
            self.power, self.x0 + self.width / 2.0, self.x0 - self.width / 2.0)


class RectangularRingPumping1D(OpSumPumping):
    """Rectangles at +radius and -radius (pumping.py:224-228)."""

    def __init__(self, power=10.0, radius=10.0, width=2.0):
        super(RectangularRingPumping1D, self).__init__(RectangularPumping1D(power, +radius, width),
                                                       RectangularPumping1D(power, -radius, width))


class RingPumping2D(AbstractPumping):
    """Single Gaussian ring power * exp(-(r - radius)^2 / (2 width^2)),
    r = |(x, y) - (x0, y0)| -- the profile of the standard workloads (no
    mirror term at -radius).  Not in the reference."""

    def __init__(self, power=1.0, radius=1.0, width=1.0, x0=0.0, y0=0.0):
        super(RingPumping2D, self).__init__()
        self.power, self.radius, self.width, self.x0, self.y0 = power, radius, width, x0, y0

    def __call__(self, x, y, t=None):
        r = np.sqrt((np.asarray(x) - self.x0) ** 2 + (np.asarray(y) - self.y0) ** 2)
        return self.power * np.exp(-(r - self.radius) ** 2 / (2.0 * self.width ** 2))

    def __repr__(self):
        return u"<class RingPumping2D power=%g radius=%g width=%g>" % (self.power, self.radius, self.width)
