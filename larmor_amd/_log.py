"""The package logger (the reference logs through larmor._log.LarmorLogger)."""

import logging

LarmorLogger = logging.getLogger("larmor_amd")
