"""Minimal logger with the reference's interface (src/utils/logger.py:7-98).

``InfoLog`` / ``WarnLog`` print a timestamped line, ``ErrorLog`` prints and
raises ``RuntimeError`` as the reference's does.  No TensorBoard writer.
"""

import time


class Logger:
    def __init__(self, root_path: str = "", log_to_disk: bool = False, quiet: bool = False):
        self.root_path = root_path
        self.log_to_disk = log_to_disk
        self.quiet = quiet
        self.initialized = False

    def Initialization(self):
        self.initialized = True

    def _log(self, kind: str, sentences: str):
        if not self.quiet:
            print(time.strftime("[%Y-%m-%d %H:%M:%S] (") + kind + ") " + str(sentences), flush=True)

    def InfoLog(self, sentences=""):
        self._log("EInfo", sentences)

    def WarnLog(self, sentences=""):
        self._log("EWarn", sentences)

    def ErrorLog(self, sentences=""):
        self._log("EError", sentences)
        raise RuntimeError(sentences)
