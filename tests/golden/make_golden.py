"""Write tests/golden/known_answers.json: known-answer data from the reference's tests.

Every entry is data taken from the reference's own test files (inputs and the
expected outputs those tests assert), cited by file:line. Nothing is imported
from the reference. Run: python tests/golden/make_golden.py
"""

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def cistem_zyz_matrix(phi, theta, psi):
    # Expected matrix of tests/test_agree_with_cistem.py:100-114 (cisTEM's ZYZ
    # convention); the reference asserts EulerAnglePose(-phi, -theta, -psi)
    # .rotation.as_matrix() equals it to atol 1e-12.
    p, t, s = (np.deg2rad(a) for a in (phi, theta, psi))
    cp, sp, ct, st, cs, ss = np.cos(p), np.sin(p), np.cos(t), np.sin(t), np.cos(s), np.sin(s)
    m = np.zeros((3, 3))
    m[0, 0] = cp * ct * cs - sp * ss
    m[0, 1] = sp * ct * cs + cp * ss
    m[0, 2] = -st * cs
    m[1, 0] = -cp * ct * ss - sp * cs
    m[1, 1] = -sp * ct * ss + cp * cs
    m[1, 2] = st * ss
    m[2, 0] = st * cp
    m[2, 1] = st * sp
    m[2, 2] = ct
    return m


def main():
    zyz = [(10.0, 90.0, 170.0), (10.0, 80.0, -20.0), (-1.2, 90.5, 67.0), (-50.0, 62.0, -21.0)]
    data = {
        "source": "reference tests (data only): see each entry's 'cite'",
        "zyz_cistem": {
            "cite": "tests/test_agree_with_cistem.py:88-116",
            "atol": 1e-12,
            "cases": [
                {"phi": a, "theta": b, "psi": c, "matrix": cistem_zyz_matrix(a, b, c).tolist()}
                for a, b, c in zyz
            ],
        },
        "euler_roundtrip": {
            "cite": "tests/test_pose.py:74-99",
            "cases": [[2.0, 15.0, -40.0], [10.0, 90.0, 170.0], [-10.0, 120.0, 140.0], [-120.0, 40.0, -80.0]],
        },
        "pose_conversion_wxyz": {"cite": "tests/test_pose.py:33-40", "wxyz": [1.0, 2.0, 3.0, 0.5]},
        "axis_angle_degrees": {"cite": "tests/test_pose.py:51-62", "angle": 2.0},
        "ctf_parameter_sets": {
            "cite": "tests/test_agree_with_cistem.py:28-41 (cisTEM outputs unavailable; inputs only)",
            "columns": ["defocus1", "defocus2", "asti_angle", "kV", "cs", "ac", "pixel_size"],
            "cases": [
                [12000.0, 12000.0, 0.0, 300.0, 2.7, 0.07, 1.0],
                [12000.0, 12000.0, 0.0, 200.0, 0.01, 0.12, 1.3],
                [1200.0, 1200.0, 0.0, 300.0, 2.7, 0.07, 1.5],
                [24000.0, 12000.0, 30.0, 300.0, 2.7, 0.07, 0.9],
                [24000.0, 24000.0, 0.0, 300.0, 2.7, 0.07, 2.0],
                [9000.0, 7000.0, 180.0, 300.0, 2.7, 0.07, 1.0],
                [12000.0, 9000.0, 0.0, 200.0, 2.7, 0.07, 0.9],
                [12000.0, 12000.0, 60.0, 200.0, 2.7, 0.02, 0.75],
                [12000.0, 3895.0, 45.0, 200.0, 2.7, 0.07, 2.2],
            ],
        },
        "grid_search_landscape": {
            "cite": "tests/test_jax_util.py:124-167",
            "columns": ["batch_size", "dim", "offset_x", "offset_y", "variance"],
            "cases": [[None, 200, -1.0, 2.0, 10.0], [1, 200, -1.0, 2.0, 10.0], [10, 200, -1.0, 2.0, 10.0],
                      [33, 200, 99.0, 99.0, 10.0]],
        },
        "mrc_fixture": {
            "cite": "tests/data/3j9g_potential_ps4_4.mrc (copied data file; header nx=ny=nz=64, mode 2, cell 281.6 A)",
            "file": "3j9g_potential_ps4_4.mrc",
            "voxel_size": 4.4,
        },
    }
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
