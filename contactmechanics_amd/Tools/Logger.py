"""Column logger with the interface the solver calls (``pr``, ``st``).

Mirrors ``ContactMechanics/Tools/Logger.py:102-230`` in interface.  ``quiet``
drops everything (it does not buffer lines without bound).
"""

import sys


class Logger:
    def __init__(self, fobj=None):
        self.fobj = fobj
        self._header_done = False

    def pr(self, msg):
        if self.fobj is not None:
            print(f"# {msg}", file=self.fobj)

    def st(self, headers, values, force_print=False):
        if self.fobj is None:
            return
        if not self._header_done:
            print("#" + " ".join(f"{i + 1:>5}:{h:>16}" for i, h in enumerate(headers)), file=self.fobj)
            self._header_done = True
        cols = []
        for v in values:
            if isinstance(v, str):
                cols.append(f"{v:>22}")
            elif isinstance(v, int):
                cols.append(f"{v:>22d}")
            else:
                cols.append(f"{float(v):>22.12e}")
        print(" ".join(cols), file=self.fobj)

    @property
    def active(self):
        return self.fobj is not None


quiet = Logger(None)
screen = Logger(sys.stdout)
