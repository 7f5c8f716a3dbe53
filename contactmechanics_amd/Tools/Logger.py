"""Column logger with the interface the solver calls (``pr``, ``st``).

Mirrors ``ContactMechanics/Tools/Logger.py:102-230`` in interface.  ``quiet``
drops everything (it does not buffer lines without bound).
"""

import sys


class Logger:
    """``fobj``: a file object, a file name (opened for writing) or None
    (drop everything).  The column header is printed before the first row
    and again every ``sepevery`` rows."""

    def __init__(self, fobj=None, sepevery=10):
        if isinstance(fobj, str):
            fobj = open(fobj, "w")
        self.fobj = fobj
        self.sepevery = sepevery
        self._rows = 0

    def pr(self, msg):
        if self.fobj is not None:
            print(f"# {msg}", file=self.fobj)
            self.fobj.flush()

    def warn(self, msg):
        self.pr("Warning: " + msg)
        if self.fobj is not sys.stdout:
            print(f"# Warning: {msg}", file=sys.stdout)

    def flush(self):
        if self.fobj is not None:
            self.fobj.flush()

    def st(self, headers, values, force_print=False):
        if self.fobj is None:
            return
        if self._rows % max(1, self.sepevery) == 0:
            print("#" + " ".join(f"{i + 1:>5}:{h:>16}" for i, h in enumerate(headers)), file=self.fobj)
        self._rows += 1
        cols = []
        for v in values:
            if isinstance(v, str):
                cols.append(f"{v:>22}")
            elif isinstance(v, int):
                cols.append(f"{v:>22d}")
            else:
                cols.append(f"{float(v):>22.12e}")
        print(" ".join(cols), file=self.fobj)
        self.fobj.flush()

    @property
    def active(self):
        return self.fobj is not None


quiet = Logger(None)
screen = Logger(sys.stdout)
