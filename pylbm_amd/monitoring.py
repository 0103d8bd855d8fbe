"""MLUPS monitoring with pylbm's definition (pylbm/monitoring.py:48-131):
MLUPS = ncalls x interior cells / total seconds / 1e6 (formula at line 111,
cell count from pylbm/simulation.py:95-96).  Enabled with ``Monitor.enable()``
or the ``--monitoring`` command-line flag (pylbm/options.py:26-29); the
timing of ``one_time_step`` then includes a stream synchronisation."""
import atexit
import sys

__all__ = ["Monitor", "mlups"]


def mlups(ncalls, cells, seconds):
    return ncalls * cells / seconds / 1e6


class Monitor:
    enabled = "--monitoring" in sys.argv
    size = 0
    stats = {}

    @classmethod
    def enable(cls, flag=True):
        cls.enabled = flag

    @classmethod
    def set_size(cls, size):
        cls.size = size

    @classmethod
    def record(cls, name, seconds):
        n, t = cls.stats.get(name, (0, 0.0))
        cls.stats[name] = (n + 1, t + seconds)

    @classmethod
    def report(cls):
        lines = []
        for name, (n, t) in cls.stats.items():
            rate = mlups(n, cls.size, t) if t > 0 else float("nan")
            lines.append("%-20s ncalls=%-8d time=%10.4fs  MLUPS=%10.2f" % (name, n, t, rate))
        return "\n".join(lines)

    @classmethod
    def print(cls):
        if cls.stats:
            print(cls.report())


atexit.register(lambda: Monitor.enabled and Monitor.print())
