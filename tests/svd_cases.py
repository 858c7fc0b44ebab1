"""Special 3x3 matrices for SVD checks (identity, diagonal, rotations,
reflections, rank-deficient, repeated and clustered singular values)."""
import numpy as np


def rot(axis, ang):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


def special_matrices():
    R1 = rot([1, 2, 3], 0.7)
    R2 = rot([-1, 0.5, 2], 2.1)
    mats = [
        np.eye(3),
        np.zeros((3, 3)),
        np.diag([3.0, 2.0, 1.0]),
        np.diag([1.0, 2.0, 3.0]),
        np.diag([1.0, -1.0, 1.0]),
        np.diag([-1.0, -1.0, -1.0]),
        np.diag([2.0, 2.0, 2.0]),
        np.diag([1.0, 1.0, 1e-7]),
        np.diag([1.0, 0.0, 0.0]),
        R1,
        R1 @ np.diag([1.0, 1.0, -1.0]),
        R1 @ np.diag([1.2, 1.0, 0.9]) @ R2.T,
        R1 @ np.diag([1.0 + 1e-6, 1.0, 1.0 - 1e-6]) @ R2.T,
        R1 @ np.diag([1.0075, 1.0, 0.975]) @ R2.T,
        R1 @ np.diag([5.0, 5.0, 0.1]) @ R2.T,
        R1 @ np.diag([1.0, 1e-4, 1e-4]) @ R2.T,
        np.outer([1, 2, 3], [4, 5, 6]).astype(np.float64),
        np.array([[1, 2, 3], [4, 5, 6], [7, 8, 9]], np.float64),
        np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]], np.float64),
        np.array([[1e3, 0, 0], [0, 1e-3, 0], [0, 0, 1.0]]),
    ]
    return np.stack(mats)
