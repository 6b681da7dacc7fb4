"""Regenerates the committed fixtures in this folder.

mt19937_64.npz: raw std::mt19937_64 streams (libstdc++, compiled here with
g++) for the seeds the tests use: the lanes of Vmt19937_64(42) (42..45), of
engines 0 and 1 of a range set (0..7), and the default seed 5489 (whose
10000th draw is the C++ standard's known answer 9981545732273789042,
[rand.predef]/3).
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SEEDS = [42, 43, 44, 45, 0, 1, 2, 3, 4, 5, 6, 7, 5489, 16380, 16381, 16382, 16383]
COUNT = 1000


def main():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "mt_golden")
        subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "mt_golden.cpp")],
                       check=True)
        raw = subprocess.run([exe, str(COUNT), *map(str, SEEDS)], check=True,
                             capture_output=True).stdout
        streams = np.frombuffer(raw, dtype="<u8").reshape(len(SEEDS), COUNT)
        kat = subprocess.run([exe, "10000", "5489"], check=True, capture_output=True).stdout
        d10000 = np.frombuffer(kat, dtype="<u8")[-1]
    assert int(d10000) == 9981545732273789042, d10000
    np.savez(os.path.join(HERE, "mt19937_64.npz"), seeds=np.array(SEEDS, dtype=np.uint64),
             streams=streams, default_seed_draw_10000=np.array([d10000], dtype=np.uint64))
    print("wrote mt19937_64.npz", streams.shape)


if __name__ == "__main__":
    main()
