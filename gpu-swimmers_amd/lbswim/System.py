"""The coupled step driver (reference: lbswim/System.py:1-17)."""
from . import _core


class System:
    """A lattice plus the particle arrays ("things") coupled to it.

    Step() runs ZeroForce, AddForces of every thing, Step, CalcHydro, then
    Move of every thing -- the reference's order (System.py:9-16).  When
    every thing is a native swimmer or tracer array, the whole step is one
    native call with no Python between kernel launches; any other object is
    driven through its AddForces/Move methods in the same order.
    """

    def __init__(self, lat):
        self.lat = lat
        self.things = []

    def AddThing(self, th):
        self.things.append(th)

    def _native_split(self):
        sw, tr = [], []
        for th in self.things:
            if isinstance(th, _core.SwimmerArray):
                sw.append(th)
            elif isinstance(th, _core.TracerArray):
                tr.append(th)
            else:
                return None
        # The fused call moves all swimmer arrays before the tracer arrays;
        # that equals the user's order only if no tracer precedes a swimmer.
        kinds = [isinstance(th, _core.SwimmerArray) for th in self.things]
        if kinds != sorted(kinds, reverse=True):
            return None
        return sw, tr

    def Autotune(self, steps_per_trial=5):
        """Times the native coupled step's launch choices (two launches, fused
        tile shapes) on real steps -- the system advances by them -- and
        keeps the fastest; returns (rows, strips), rows 0 for two launches.
        Systems with non-native things are left as they are (returns None)."""
        native = self._native_split() if isinstance(self.lat, _core.Lattice) else None
        if native is None:
            return None
        return _core.system_autotune(self.lat, native[0], native[1], steps_per_trial)

    def Step(self, nsteps=1):
        native = self._native_split() if isinstance(self.lat, _core.Lattice) else None
        if native is not None:
            _core.system_step(self.lat, native[0], native[1], nsteps)
            return
        for _ in range(nsteps):
            self.lat.ZeroForce()
            for th in self.things:
                th.AddForces(self.lat)
            self.lat.Step()
            self.lat.CalcHydro()
            for th in self.things:
                th.Move(self.lat)
