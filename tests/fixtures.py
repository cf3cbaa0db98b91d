"""Reference input / calibration vectors (reference data/, data/data_generation.py).

The small files are committed under tests/golden/nexus_data/ as copied data.
The 32768-value GELU vectors are regenerated from data_generation.py's
formulas; the SHA-256 of the regenerated text equals that of the reference's
files (data/input/gelu_input_32768.txt, data/calibration/gelu_calibration_32768.txt).
"""
import hashlib
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nexus_data")
GELU_INPUT_SHA256 = "fd15fa8f5c0bb4c2478e971037281dcc2fe6c65d63e9039d938f995da7167151"
GELU_CALIB_SHA256 = "6f964e9d7f99ef150b4246df3a23ac84a522f22102ee3eb063d63fa5e72b490f"


def _as_file_text(v):
    return "".join("%f\n" % x for x in v)


def gelu_vectors():
    x = np.linspace(-5, 5, 32768)
    y = x * 0.5 * (1 + np.tanh(np.sqrt(2 / np.pi) * (x + 0.044715 * x ** 3)))
    tx, ty = _as_file_text(x), _as_file_text(y)
    assert hashlib.sha256(tx.encode()).hexdigest() == GELU_INPUT_SHA256
    assert hashlib.sha256(ty.encode()).hexdigest() == GELU_CALIB_SHA256
    return np.array([float(s) for s in tx.split()]), np.array([float(s) for s in ty.split()])


def load(name):
    return np.loadtxt(os.path.join(HERE, name))


def layernorm_vectors():
    return load("layernorm_input_16_768.txt"), load("layernorm_calibration_16_768.txt")


def softmax_vectors():
    return load("softmax_input_128_128.txt"), load("softmax_calibration_128_128.txt")


def argmax_vectors():
    return load("argmax_input_8.txt"), load("argmax_calibration_8.txt")
