"""Timer report of VelocityVerlet.getTimers (src/tools/timers.py:25-80)."""
import sys

KEYS = ["run", "pair", "FENE", "angle", "comm1", "comm2", "int1", "int2", "resort", "other"]


def to_dict(alltimers):
    t = list(alltimers[0])
    return dict(zip(KEYS, t))


def show(alltimers, precision=1):
    d = to_dict(alltimers)
    run = d["run"] or 1.0
    for k in KEYS:
        fmt = '%-6s time (%%) = %.' + str(precision) + 'f (%.1f)\n'
        sys.stdout.write(fmt % (k, d[k], 100.0 * d[k] / run))
