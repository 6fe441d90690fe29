"""Writes to_chars_vectors.txt: 3,000 doubles and their std::to_chars text (g++ -std=c++17).

The doubles mix magnitudes from 1e-30 to 1e30, random bit patterns, short
decimals, integers and powers of two. Run from this directory:
``python make_to_chars_vectors.py``.
"""

import subprocess
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def main():
    rng = np.random.default_rng(20261203)
    dec = rng.integers(0, 6, 600)
    vals = np.concatenate([
        rng.standard_normal(1000) * 10.0 ** rng.integers(-30, 30, 1000),
        np.frombuffer(rng.bytes(8 * 800), dtype=np.float64),
        np.array([round(float(a), int(d)) for a, d in zip(rng.standard_normal(600) * 1000, dec)]),
        rng.integers(-10**6, 10**6, 300).astype(np.float64),
        2.0 ** rng.integers(-1074, 1023, 280),
        np.array([1.0, 0.5, 1e-5, 1e-4, 0.001, 1e16, 1e15, 123456789012345678.0, 1e22, 0.1 + 0.2, 5e-324,
                  1.7976931348623157e308, 100.0, -9999.0, 2.0 ** 60, 1234567.0, 0.000123, 1e21, 1e-7,
                  12345678901234567890.0]),
    ])
    vals = vals[np.isfinite(vals)]
    with tempfile.TemporaryDirectory() as d:
        exe = Path(d) / "to_chars"
        subprocess.run(["g++", "-std=c++17", "-O2", str(HERE / "to_chars.cpp"), "-o", str(exe)], check=True)
        out = subprocess.run([str(exe)], input=vals.astype("<f8").tobytes(), capture_output=True, check=True).stdout
    (HERE / "to_chars_vectors.txt").write_bytes(out)


if __name__ == "__main__":
    main()
