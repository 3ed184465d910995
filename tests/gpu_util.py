"""Shared helpers of the GPU tests."""
import sys

import numpy as np


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


RGB_CAM = np.array([[1.9, -0.7, -0.2, 0.0], [-0.3, 1.5, -0.2, 0.0], [0.0, -0.4, 1.6, 0.0]], np.float32)
CAM_MUL = np.array([1.8, 1.0, 1.5, 1.0], np.float32)
FILTERS = {"RGGB": 0x94949494, "BGGR": 0x16161616, "GRBG": 0x61616161, "GBRG": 0x49494949}
